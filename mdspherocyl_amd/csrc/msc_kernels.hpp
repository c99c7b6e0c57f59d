// Device state handed to every kernel, and the host-side launch wrappers.
#pragma once

#include <hip/hip_runtime.h>

#include "msc_common.hpp"

namespace msc {

struct State {
  int Ntot, nd, nrep;
  int ntask, nrtask, ncells;  // ntask: tasks in the selected range
  int ntask_all;             // every wave task of the context: the value stride
                             // of per-task rows (value k of task t at k * ntask_all + t)
  int task_begin;            // first selected wave task
  int g_begin, g_end;        // particle range of per-particle kernels
  int rep_begin, rep_end;    // replica range of per-replica reductions
  double lever_max;          // largest |dpc| of the context
  const RepDev *reps;        // [nrep]
  const Task *tasks;         // [ntask] wave tasks, grouped by replica
  const Task *rtasks;        // [nrtask] reciprocal particle chunks, grouped by replica
  const int *rep_of;         // [Ntot]
  // particle state (component-major)
  double *x, *p, *l, *ori, *F, *T;   // 3N, 3N, 3N, 9N, 3N, 3N
  const double *len, *dia, *mass, *inert;  // N, N, N, 3N
  const double *dpc, *doc, *dmu;     // nd N, 3 nd N, nd N
  double *dpos, *dori;               // 3 nd N lab dipole positions / orientations
  double *snap;                      // 3N neighbour-list reference positions
  double *snap_e;                    // 6N centre and axis when the contact list was built
  double *snap_u;                    // 3N axis when the reference list was built
  // cell grid
  int *cell_of, *cell_cnt, *cell_start, *cell_fill, *cell_tmp, *cell_list;
  double *cell_pack;         // [Ntot][8] centre, axis, length, index in cell order
  const int *cell_rep;       // [ncells] owning replica
  // neighbour rows (layout in RepDev): reference list, contact list,
  // real-space dipole pairs; counts per particle
  int *nl_cnt, *nl_idx;
  int *sl_cnt, *sl_idx;
  int *dl_cnt, *dl_idx;
  // reciprocal weights per slot (0 outside the half space)
  const double *kval;
  // Partner records, rewritten with the dipoles whenever the particles move
  // (null: not kept): rec_rod [N][8] centre, axis, length, diameter;
  // rec_dip [nd][N][8] dipole position (3 doubles), then as 8 floats the
  // orientation, the offset dpc u and the moment (one 64-byte line each).
  double *rec_rod, *rec_dip;
};

// List maintenance of the step loop (v25/main.cpp:679-701).  k_drift
// reduces the displacements into dmax; extra blocks of k_force_sc decide
// (rebuild, srebuild, obs[.][11]) and, with those of the step's later
// launches, rebuild the cell grid.
constexpr int kDmaxBuckets = 64;  // power of two, at most 64
struct ListStep {
  int *rebuild;              // [nrep] rebuild the reference rows this step
  int *srebuild;             // [nrep] refresh the contact rows (null: no contact rows)
  int *ctr;                  // [nrep] arrival counters, zero between launches
  unsigned long long *dmax;  // [2][nrep][kDmaxBuckets][2] by step parity: largest
                             // centre-component and segment-point displacements
                             // (double bits); zero before the step that fills them
  double *obs;               // observables (12 per replica and step) or null
  const int *step;           // the step counter
};

// First failed kernel launch since the record was last taken.
struct LaunchRecord {
  hipError_t err = hipSuccess;
  const char *kernel = "";
};
// Makes r the record that launches from the calling host thread report into
// (null: the shared record); returns the previous one.
LaunchRecord *set_launch_record(LaunchRecord *r);
// first kernel launch that failed since the last call (hipSuccess if none),
// with the kernel's name, from the calling thread's current record; clears it
hipError_t take_launch_error(const char **kernel);
// MSC_SYNC_CHECK is set: every launch is followed by a synchronisation
bool launches_checked();
// records a failed stream operation (event record or wait) like a failed
// launch, under the name what
void note_stream_op(const char *what, hipError_t e);

void launch_dipoles(const State &s, hipStream_t st);
// zero: clear every cell counter first (they are zero outside a rebuild:
// the scan clears those it consumes)
void launch_cells(const State &s, const int *rebuild, bool zero, hipStream_t st);
void launch_lists(const State &s, const int *rebuild, int *flags, hipStream_t st);
void launch_short(const State &s, const int *srebuild, int *flags, hipStream_t st);
void launch_check(const State &s, double *task_out, int *rebuild, int *srebuild, double *obs_rebuild,
                  int obs_stride, bool standalone, hipStream_t st);
// ls (step loop only): the cell assignment and scan of rebuilt replicas run
// in extra blocks of the same launch
void launch_force_sc(const State &s, bool strict, bool mixed, double *task_out, int *flags, hipStream_t st,
                     const ListStep *ls = nullptr);
void launch_dipR(const State &s, bool mixed, bool all_pairs, const void *tab, double *task_out, hipStream_t st);
// mixed contexts, every replica with a real-space cut-off and contact rows:
// k_force_sc<false, true> and k_dipR_list<float> in one launch (step loop; ls as
// launch_force_sc's)
// waves: per 64-particle task, 2, 4 or 8 (the row entries q = w mod waves go
// to wave w; the choice changes the rounding of the pair sums only)
void launch_pairs_mixed(const State &s, int waves, const void *tab, double *sc_out, double *r_out, int *flags,
                        const ListStep &ls, hipStream_t st);
void launch_ext_field(const State &s, double *task_out, hipStream_t st);
void launch_self_ener(const State &s, double *task_out, hipStream_t st);
// out[rep*ostride + ooff + k] = replica sum of value k < nv of its items:
// which 0, wave tasks, value-major rows in[k*ntask_all + task] (istride is
// not used); which 1, k slots, in[slot*istride + k].  With nv2 > 0 the
// following nv2 values go to out2[rep*o2stride + o2off + k]
void launch_reduce(const State &s, const double *in, int istride, int nv, int which, double *out, int ostride,
                   int ooff, hipStream_t st, int nv2 = 0, double *out2 = nullptr, int o2stride = 0, int o2off = 0);
void launch_recip_tables(const State &s, bool mixed, void *ph, void *amp, double *task_out, hipStream_t st);
int launch_recip_struct(const State &s, bool mixed, int KZ, int rt0, int nrt, int maxcol, const void *ph,
                        const void *amp, void *part, hipStream_t st);
// M2: per-replica sums of c c^T, or null to sum them from the per-task rows
// m2rows; kblk (may be null): per-block sums of the 10 slot values over
// kKblkSlots slots, nblk(maxslot, kKblkSlots) rows per replica
constexpr int kKblkSlots = 64;
// bop (mixed contexts): the operands of the matrix-core back projection,
// kBopPerTile groups of 4 fp16 values per column tile of 16 (RepDev::bop0,
// a multiple of 2: the kernels read them as groups of 8), scaled by
// 2^bexp[3 col + l] (RepDev::bexp0)
// The last kick fused into the back projection (mixed contexts of the step
// loop without an external field, where the back projection is the step's
// last force launch): k_recip_reduce takes moveTher mode 1 of each replica
// (from the first kick's rows, as k_kick2) into ts[2 rep + i]; the back
// projection then kicks its task's particles with the final F and T (the
// last moveP, as k_kick2), writes the kick2 rows, and its extra blocks fill
// the cells of rebuilt replicas (k_kick2's extra blocks).
struct KickFuse {
  const Thermo *th;
  const double *gs;   // the run's deviates
  const int *step;
  const double *rows1;  // kick1 rows (k_drift)
  double *ts;           // [nrep][2]
  double *rows2;        // kick2 rows
  ListStep ls;
};
void launch_recip_reduce(const State &s, bool mixed, const void *part, const double *M2, const double *m2rows,
                         void *A, double *slot_out, double *kblk, int maxslot, void *bop, const int *bexp,
                         hipStream_t st, const KickFuse *kf = nullptr);
constexpr int kBopPerTile = 24 * 64;  // 3 weights x 4 operands x 64 lanes x 8 fp16
int launch_recip_force(const State &s, bool mixed, int KZ, const void *ph, const void *amp, const void *A,
                       const void *bop, const int *bexp, hipStream_t st, const KickFuse *kf = nullptr);
// Reciprocal sum of contexts with several dipole slots per particle, FP64:
// slot phase tables ph (nd * 3 * kMaxKZ * Ntot complex doubles), chunk
// partials of the chunk tasks rt0 .. rt0 + nrt - 1 (10 doubles per slot and
// chunk, at RepDev::part0 / 6 * 10), slot energies and A (complex doubles per
// slot), kblk as in launch_recip_reduce, and the forces and torques added to
// F, T of the selected tasks.
void launch_recip_nd(const State &s, int rt0, int nrt, int maxslot, void *ph, double *part, void *A,
                     double *slot_out, double *kblk, hipStream_t st);
void launch_moveP(const State &s, bool last, const Thermo *th, double *task_out, hipStream_t st);
void launch_moveX(const State &s, hipStream_t st);
void launch_renorm(const State &s, hipStream_t st);
void launch_thermo0(const State &s, Thermo *th, const double *gs, hipStream_t st);
void launch_step_scalars(const State &s, int mode, const double *red, double *Kin, Thermo *th, const RepEnergy *E,
                         double *obs, hipStream_t st);
void launch_zero_FT(const State &s, hipStream_t st);
void launch_fill_int(int *p, int n, int v, hipStream_t st);
// one-lane kernel writing v to p[0] (completion stamp)
void launch_stamp(double *p, double v, hipStream_t st);
// the step loop's fused kernels; rows + k*region is per-task row set k:
// 0 kick1, 1 drift, 2 sc, 3 R, 4 ext, 5 kick2 (16 doubles per task)
constexpr int kRowKick1 = 0, kRowDrift = 1, kRowSC = 2, kRowR = 3, kRowExt = 4, kRowKick2 = 5, kRowSets = 6;
// step: device {k, first absolute step of the run}; gs and obs are the run's
// deviate and observable buffers, indexed by k in the kernels
void launch_drift(const State &s, bool mixed, const Thermo *th, const double *gs, const int *step, double *rows1,
                  void *ph, void *amp, double *rows, const ListStep &ls, hipStream_t st);
void launch_kick2(const State &s, const Thermo *th, const double *gs, const int *step, const double *rows1,
                  double *rows2, const ListStep &ls, hipStream_t st);
// End-of-step sums per replica (sums[kEndSums rp + v]): v < kEndVals the
// replica sums of the per-task rows and k-space block rows the step's
// thermostat, energies and observables need (k_end, several blocks per
// replica so that the row reads spread over several CUs), v = kEndVals the
// step index k, then the step's four thermostat deviates of the replica
constexpr int kEndVals = 44, kEndSums = 56, kEndBlocks = 4;  // kEndBlocks * 4 waves >= kEndVals / 4
void launch_end(const State &s, const int *step, const double *gs, const double *rows, long region,
                bool has_field, const double *kblk, int maxslot, double *sums, const ListStep &ls,
                hipStream_t st);
// Outputs of the end of a step, written from the sums by one extra block
// per replica of k_lists_step: moveTher mode 2 into th, the energies E in
// the layout of RepEnergy, Kin, the pair counts and the observables of
// print_out (obs + 12 (k nrep + rp)) but the rebuild flag
struct EndArgs {
  const double *sums;  // launch_end's sums
  Thermo *th;
  RepEnergy *E;
  double *Kin, *counts, *obs;
};
// the step's last launch: rows of rebuilt replicas, contact rows of
// refreshed ones, and the end-of-step outputs (EndArgs); advances step[0]
void launch_lists_step(const State &s, const ListStep &ls, int *flags, int *step, const EndArgs &ea,
                       hipStream_t st);
// contact-graph components of the selected replica: par[g] = smallest
// particle of g's component (global index); flags[0] counts particles that
// moved too far for the reference rows to hold every contact, flags[1] pairs
// whose contact test depends on the order of the two particles
// u_cut: contacts also need the bare dipolar pair energy below u_cut (+inf: off)
void launch_clusters(const State &s, double min_s, double reach_max, double u_cut, int *par, int *flags,
                     hipStream_t st);

// order-parameter sums of replica rep (oo_params): out[0..5] = sum u_a u_b
// (xx xy xz yy yz zz), out[6..8] = sum u, out[9..11] = total dipole moment
constexpr int kOrderSums = 12;
constexpr int kOrderThreads = 1024;  // one block; its LDS tree holds 12 x 1024 doubles
void launch_order_sums(const State &s, int rep, double *out, hipStream_t st);
// oo_params as the reference computes it (single-precision Q and Jacobi,
// ordered sums): out = G1, G2, M[3], rotations; at most kOrderRefSlots
// dipole slots per particle
constexpr int kOrderRefSlots = 4;
void launch_order_ref(const State &s, int rep, double *out, hipStream_t st);

// pair correlations of one replica (acc_correl): histograms of the sets
// [sets][4][n_bars] added into hist, pairs beyond the last bar counted in
// *dropped.  Bars go through LDS in ranges of nbr (kCorrelLds doubles per
// block hold sets*4*nbr of them).
constexpr int kCorrelTile = 512;          // particles per tile (threads per block)
constexpr int kCorrelLds = 19456;         // LDS histogram, doubles (152 KiB)
constexpr int kCorrelMaxSpec = 8;
struct CorrelArgs {
  int n0, N, nspec, sets;
  int spec_end[kCorrelMaxSpec];  // cumulative species counts; spec_end[nspec-1] = N
  double L[3], invL[3];  // box and 1 / box
  double bar_w, inv_bw;
  int n_bars, nbr;
};
void launch_pair_correl(const double *x, const double *ori, int Ntot, const CorrelArgs &a, double *hist,
                        unsigned long long *dropped, hipStream_t st);

constexpr int kChunk = 64;  // particles per reciprocal structure-factor chunk
constexpr int kRedStride = 16;  // doubles per replica in the reduction scratch d_red
constexpr int kRfGroups = 8;    // waves sharing one task in the reciprocal back projection
constexpr int kMaxPairWaves = 8;  // most waves per task of the fused mixed pair kernel
constexpr int kPairSplit = 4;   // waves sharing one task in the WCA and real-space pair kernels
constexpr int kListStepBlocks = 2048;  // grid of the step loop's row launch (4 particles per wave at a time)

}  // namespace msc
