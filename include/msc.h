/*
 * msc.h -- C ABI of libmsc_hip.so, the MI355X implementation of mdSpherocyl's
 * per-step force path (neighbour list, short-range spherocylinder force,
 * Ewald dipolar real-space and reciprocal sums) and the step loop around it.
 *
 * Plain C: pointers, sizes and status codes.  No exceptions cross this
 * boundary.  A context owns one HIP stream on one device; every call is
 * ordered on that stream and returns after the result it reports is ready.
 * Host arrays passed in are read during the call only.
 *
 * Layout of host arrays (one replica, particles in the reference's global-id
 * order, i.e. species-major as rev_II at v25/functions.cpp:1166-1179):
 *   r, p, l, F, T : [N][3]     ori : [N][3][3] row-major, row 2 = axis u
 *   len, dia, mass: [N]        dpc, dmu : [N][nd]     doc : [N][nd][3]
 * A context may hold several independent replicas ("systems"); most entry
 * points take a replica index, MSC_ALL meaning every replica.
 *
 * Conventions follow the reference routines each entry point replaces:
 * force routines ADD into F and T (the caller zeroes them, v25/main.cpp:582-588,
 * here msc_zero_forces) and OVERWRITE the energy / virial / pressure-tensor
 * outputs; energies are totals, not per particle.
 */
#ifndef MSC_H
#define MSC_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define MSC_OK          0
#define MSC_E_OVERLAP   1   /* force_sc: segment distance^2 < 0.49 sigma^2  */
#define MSC_E_DISTANCE  2   /* force_sc: segment distance^2 > 1.00001 r^2   */
#define MSC_E_ARG       3   /* bad argument                                  */
#define MSC_E_HIP       4   /* HIP runtime error, or kernels that did not
                               execute (each call that enqueues kernels reads
                               back a completion stamp written after them)  */
#define MSC_E_CAPACITY  5   /* a device buffer bound was exceeded            */
#define MSC_E_STATE     6   /* call order (e.g. force before neighbour list) */

/* precision of pair and reciprocal arithmetic */
#define MSC_FP64  0         /* plain FP64 throughout (drop-in path)           */
#define MSC_MIXED 1         /* FP32 pair/reciprocal arithmetic, FP64 state,
                               reductions, thermostat and rotation            */

#define MSC_ALL  (-1)

/* observables per replica and step, as printed by print_out
 * (v25/functions.cpp:2830-2853) plus the list-rebuild flag */
#define MSC_OBS_K      0
#define MSC_OBS_U      1
#define MSC_OBS_P      2
#define MSC_OBS_USS    3
#define MSC_OBS_PSS    4
#define MSC_OBS_UDR    5
#define MSC_OBS_PDR    6
#define MSC_OBS_UDK    7
#define MSC_OBS_PDK    8
#define MSC_OBS_MOM2   9
#define MSC_OBS_ANG2   10
#define MSC_OBS_REBUILD 11
#define MSC_OBS_N      12

typedef struct msc_ctx msc_ctx;

/* Parameters of one replica.  Fields map onto input.txt (r_input,
 * v25/functions.cpp:10-87) as noted. */
typedef struct msc_replica_desc {
  int N;              /* particles ("No. of Particles", summed)            */
  int nd;             /* dipole slots per particle (1..8)                  */
  double L[3];        /* box                                               */
  double alpha;       /* "Convergence param(a)"                            */
  int K;              /* "No. of Kvectors": |n_a| <= K-1                   */
  double sr_cut;      /* "Cutoff dist(SRange)"                             */
  double rc;          /* real-space Ewald cutoff; <= 0: every minimum-image
                         pair, as force_dipR (v25/functions.cpp:1570-1600) */
  double B[3];        /* "Ext Mag.Field"                                   */
  double temp;        /* "Temperature"                                     */
  double dt;          /* "Time step size"                                  */
  double Q;           /* "Thermostat\"mass\""                               */
  double gamma;       /* "Thermostat\"shearRate\""                          */
  double s0;          /* initial thermostat coordinate (0 in v25, 1 in v251) */
  double dens;        /* "Density" (number density, enters P only)         */
  double dmin, dmax;  /* smallest diameter, largest L+D over species
                         (v25/main.cpp:153-158); <= 0: taken from the state */
  unsigned long thermo_seed[2]; /* MT19937 seeds (1 and 2 in v25/main.cpp:183,186) */
} msc_replica_desc;

/* ---- context ---------------------------------------------------------- */
int  msc_create(int device, int precision, int nrep, const msc_replica_desc *reps,
                msc_ctx **out);
void msc_destroy(msc_ctx *ctx);
const char *msc_last_error(const msc_ctx *ctx);
int  msc_version(void);

/* ---- state ------------------------------------------------------------ */
/* Upload the state of replica `rep` (host arrays, layout above).  Dipole lab
 * positions and orientations are derived as Spherocyl::set_dip does
 * (v25/classes.cpp:148-173).  Per-particle inertia follows set_inertia
 * (v25/classes.cpp:175-185). */
int msc_set_state(msc_ctx *ctx, int rep, const double *r, const double *p,
                  const double *l, const double *ori, const double *len,
                  const double *dia, const double *mass, const double *dpc,
                  const double *doc, const double *dmu);
int msc_get_state(msc_ctx *ctx, int rep, double *r, double *p, double *l, double *ori);
int msc_get_forces(msc_ctx *ctx, int rep, double *F, double *T);
int msc_set_forces(msc_ctx *ctx, int rep, const double *F, const double *T);
/* thermostat coordinate s and stored normal deviate of the two thermostats */
int msc_get_thermostat(msc_ctx *ctx, int rep, double s[2], double strn[2]);
int msc_set_thermostat(msc_ctx *ctx, int rep, const double s[2], const double strn[2]);

/* ---- routines of v25/functions.hpp:29-54 ------------------------------ */
/* neigh_list (v25/functions.cpp:1298-1334): build the neighbour pairs with
 * segment distance < r_list on the device; savec_list snapshot included. */
int msc_neigh_list(msc_ctx *ctx, int rep, double r_list);
/* The caller's list as the replica's neighbour rows: force_sc's nblist
 * argument (v25/functions.cpp:1397-1398, built by neigh_list) in CSR form,
 * off[N+1] and nbr[off[N]], row f holding ascending partners g > f.  The
 * list snapshot of savec_list (v25/functions.cpp:1336-1353) is taken from the
 * present state and r_list becomes the replica's list radius.  A row that is
 * not ascending or names a partner outside (f, N) is MSC_E_ARG. */
int msc_set_neigh_list(msc_ctx *ctx, int rep, double r_list, const int *off, const int *nbr);
/* The list in the reference's form: for each f, ascending g > f.  off has
 * N+1 entries; nbr holds cap entries; *total receives the pair count. */
int msc_get_neigh_list(msc_ctx *ctx, int rep, int *off, int *nbr, long cap, long *total);
/* check_list (v25/functions.cpp:1356-1394) against the savec_list snapshot */
int msc_check_list(msc_ctx *ctx, int rep, double r_list_min, double r_list_max,
                   int *rebuild);
/* force_sc (v25/functions.cpp:1397-1553); MSC_E_OVERLAP / MSC_E_DISTANCE
 * where the reference returns false */
int msc_force_sc(msc_ctx *ctx, int rep, double r_cut, double *U, double *W, double PT[9]);
/* force_dipR (v25/functions.cpp:1555-1678), table of tabul (1680-1704) */
int msc_force_dipR(msc_ctx *ctx, int rep, double *U, double PT[9]);
/* force_dipK (v25/functions.cpp:1706-1961), weights of wave_v (1963-1990).
 * With several dipole slots per particle the interactions between the slots
 * of one particle are excluded (v25/functions.cpp:1780-1827) and the sum is
 * FP64 in every context. */
int msc_force_dipK(msc_ctx *ctx, int rep, double *U, double PT[9]);
/* ext_field (v25/functions.cpp:1992-2025) */
int msc_ext_field(msc_ctx *ctx, int rep, double *U);
/* self_ener (v25/functions.cpp:2027-2069) */
int msc_self_ener(msc_ctx *ctx, int rep, double *U);
int msc_zero_forces(msc_ctx *ctx, int rep);
/* moveX (v25/functions.cpp:2071-2092) */
int msc_moveX(msc_ctx *ctx, int rep, double dt);
/* moveP (v25/functions.cpp:2094-2184) with the replica's thermostat
 * coordinates; K[2] kinetic energies, PT the kinetic pressure tensor */
int msc_moveP(msc_ctx *ctx, int rep, double dt, int last, double K[2], double PT[9]);
/* renorm every particle (v25/classes.cpp:307-325) */
int msc_renorm(msc_ctx *ctx, int rep);

/* ---- the step loop (v25/main.cpp:520-708) ------------------------------ */
/* Initial list, forces and self energy as v25/main.cpp:460-497. */
int msc_init_run(msc_ctx *ctx);
/* Advance every replica by nsteps steps.  obs (may be NULL) receives
 * [nsteps][nrep][MSC_OBS_N] values.  step0 is the global step number of the
 * first step (renormalisation every 100 steps counts from it). */
int msc_run(msc_ctx *ctx, int step0, int nsteps, double *obs);
/* 1: msc_run replays each step as one captured HIP graph; 0 (default):
 * direct kernel launches (about 15 per step).  Runs with timing on, or with
 * MSC_SYNC_CHECK set, launch directly.  Both give identical results; on
 * MI355X at N=32768 the direct launches are about 1% faster (the step is
 * bound by the GPU, not by the host's launches). */
int msc_set_graph(msc_ctx *ctx, int on);
/* Launch choices of a context; no counterpart in the reference.  graph and
 * fused_pairs give the same results bit for bit; pair_waves changes how the
 * row entries of a 64-particle task are split over waves, so the rounding of
 * the real-space and WCA sums (each choice is deterministic).  Defaults:
 * graph 0, fused_pairs 1, pair_waves 4, overlap 1, fuse_kick 1. */
typedef struct msc_tuning {
  int graph;        /* msc_run replays each step as one captured HIP graph */
  int fused_pairs;  /* MSC_MIXED contexts with contact rows whose replicas all
                       have a real-space cut-off run the WCA over the contact
                       rows and the real-space dipole pairs in one launch,
                       partners read as packed records */
  int pair_waves;   /* waves per task of that launch: 2, 4 or 8 */
  double list_skin; /* MSC_MIXED: > 0 gives the context its own rows, every
                       pair within 2^(1/6) D_max + list_skin (WCA and
                       real-space pairs within r_c + list_skin), rebuilt
                       together when twice the largest segment-point
                       displacement exceeds list_skin; 0 keeps the
                       reference's rows and rebuild test (v25/main.cpp:
                       461-463, 679-701) with contact rows refreshed from
                       them.  MSC_FP64: always 0.  A change takes effect at
                       the next msc_init_run. */
  int chunk;        /* particles per structure-factor chunk of the reciprocal
                       sum (64, 128, 256 or 512; default 128): fewer, larger
                       chunks write and re-read fewer partial sums.  Changes
                       the summation order of S(k) only. */
  int overlap;      /* 1: the fused pair launch runs on a second stream of the
                       device, concurrently with the structure factors and
                       their reduction, and is joined before the back
                       projection (each writes its own outputs, so results
                       are the same bit for bit); 0: one stream.  Runs with
                       timing on or MSC_SYNC_CHECK set use one stream. */
  int fuse_kick;    /* 1: in MSC_MIXED contexts without an external field the
                       back projection, the step's last force launch, also
                       does the last moveP of its particles (one launch
                       fewer per step; the same results bit for bit); 0: a
                       launch of its own. */
} msc_tuning;
int msc_get_tuning(msc_ctx *ctx, msc_tuning *t);
int msc_set_tuning(msc_ctx *ctx, const msc_tuning *t);
/* The autotuner: times `steps` steps of msc_run under each candidate tuning
 * (graph off/on x, for contexts that can fuse, unfused or fused with 2, 4 or
 * 8 waves per task; then, for MSC_MIXED contexts, list skins 0.5, 0.75, 1.0
 * and 1.5 with the fastest of those; then chunks of 64, 128, 256 and 512
 * particles with the fastest so far), each from the state at the call, and
 * keeps the fastest.  The state, thermostats and thermostat generators are restored
 * after every trial and the lists rebuilt from the restored state (as after a
 * chunk rollback of msc_run), so a run continues from the same state; only
 * the summation order of the rebuilt lists can differ from an uninterrupted
 * run.  *best (may be NULL) receives the choice; tried and ms (each may be
 * NULL, room for ntry values) the candidates and their milliseconds per
 * step, in the order tried; *ntried (may be NULL) their number.  No
 * counterpart in the reference. */
int msc_autotune(msc_ctx *ctx, int steps, msc_tuning *best, msc_tuning *tried, double *ms, int ntry, int *ntried);
/* msc_run advances in chunks of 100 steps.  A chunk during which a list
 * rebuild outgrew its device region is rolled back, the regions are enlarged
 * and the lists rebuilt from the chunk's first state, and the chunk is run
 * again; *n receives how many chunks were repeated since msc_create.  No
 * reference counterpart (the reference's lists are host deques). */
int msc_chunk_repeats(msc_ctx *ctx, long *n);

/* ---- cluster labels (clust_conf, v25/functions.cpp:3270-3324) ----------
 * labels[i] (replica-local i) = the cluster clust_conf assigns particle i:
 * clusters are the components of the contact graph
 * d_seg^2 < ((D_i+D_j)/2 + min_s)^2 (dist_test), numbered in the order the
 * reference creates them.  mt_state[624]/mt_index are the MT19937 state of
 * the reference's `ran` generator (gsl_rng_mt19937: words mt[] and position
 * mti), advanced in place by the one gsl_rng_uniform draw rnd_part makes per
 * cluster.  *nclust receives the number of clusters.  n_asym (may be NULL)
 * receives the number of pairs whose contact test gives different answers
 * for (i, j) and (j, i); such a pair is taken as a contact, where the
 * reference's answer depends on which of the two its search meets first
 * (0 unless a distance lies within rounding of the threshold).
 * Needs the neighbour list of the current state (msc_neigh_list, or the
 * list msc_run keeps) and min_s + largest diameter < r_list_max; fails with
 * MSC_E_STATE if particles moved too far since the list was built for its
 * rows to hold every contact. */
int msc_cluster_labels(msc_ctx *ctx, int rep, double min_s, unsigned int *mt_state,
                       int *mt_index, int *labels, int *nclust, long *n_asym);
/* msc_cluster_labels with a dipolar contact criterion (no reference
 * counterpart; the reference's test is purely geometric): a pair is in
 * contact when dist_test holds AND its bare dipole-dipole energy, summed
 * over the dipole slots of both particles at the minimum image of each site
 * separation, mu_a mu_b (u_a.u_b r^2 - 3 (u_a.r)(u_b.r)) / r^5, is below
 * u_cut (e.g. -1.0 times the temperature for "below -1 kT").  u_cut = +inf
 * is msc_cluster_labels; NaN or -inf is MSC_E_ARG. */
int msc_cluster_labels_dd(msc_ctx *ctx, int rep, double min_s, double u_cut, unsigned int *mt_state,
                          int *mt_index, int *labels, int *nclust, long *n_asym);

/* Which neighbour rows force_sc (and the WCA launch of msc_run) walks.
 * on = 0: every pair of the reference rows, as force_sc does
 * (v25/functions.cpp:1397-1553); the default of MSC_FP64 contexts.  on = 1:
 * the contact rows (pairs within 2^(1/6) D_max plus a skin of D_min / 2,
 * refreshed from the reference rows when the skin is used up); the default
 * of MSC_MIXED contexts.  Every pair that interacts or overlaps lies in the
 * contact rows, which keep the reference rows' order, so forces, energies
 * and MSC_E_OVERLAP reports are the same; only the work differs.  One
 * difference: the reference's second abort test, d^2 > 1.00001 r^2 on every
 * list pair (v25/functions.cpp:1466, MSC_E_DISTANCE), runs on the contact
 * pairs only, so an inconsistency of a distant list pair is not reported.
 * No counterpart in the reference. */
int msc_set_contact_rows(msc_ctx *ctx, int on);

/* The reference's Signal=0 initial state (v25/main.cpp:215-260), host only:
 * ini_pos's stacked triangular lattice with every axis along z
 * (v25/functions.cpp:491-676, including its overlap test), nmix GCAmix
 * pivot-reflection cluster moves (v25/functions.cpp:3328-3409; the
 * reference makes 1000) and ini_mom's Maxwell momenta with zeroed totals
 * (v25/functions.cpp:677-1146), all drawn from one MT19937 generator given
 * as its 624 words and position (the reference's `ran`, gsl_rng_mt19937
 * seeded 1) and advanced in place.  Inputs: N particles in nspec species of
 * counts[] (species-major), box L, per-particle len, dia, mass, temperature.
 * Outputs r[3N], p[3N], l[3N], ori[9N] (row-major, third row the axis), the
 * arrays msc_set_state takes.  One species only (ini_mom indexes its
 * per-species scratch by the global particle counter, defined only then):
 * nspec > 1 is MSC_E_ARG.  MSC_E_OVERLAP: the lattice does not fit the box
 * or its sites overlap, where the reference stops.  O(N^2) segment distances
 * per move: meant for the reference's system sizes. */
int msc_ini_state(int N, int nspec, const int *counts, const double L[3], const double *len,
                  const double *dia, const double *mass, double temp, int nmix, unsigned int *mt_state,
                  int *mt_index, double *r, double *p, double *l, double *ori);

/* oo_params (v25/functions.cpp:2702-2760) with Qmat_3d (3426-3453) and
 * jacobi (v25/jacobi.c): of replica `rep`, G2 = the largest |eigenvalue| of
 * Q = sum_i (3 u_i u_i - I) / (2N), G1 = |sum_i Ve . u_i| / N along its unit
 * eigenvector Ve, and M = the total dipole moment (sum of moment times the
 * lab-frame dipole axis; overwritten, as the reference's caller zeroes it
 * before each call).  Sums in FP64 on the device, Q and its eigenvectors in
 * FP64 on the host: G1 and G2 differ from the reference's single-precision Q
 * by its rounding (about 1e-6 at N = 500). */
int msc_order_params(msc_ctx *ctx, int rep, double *G1, double *G2, double M[3]);
/* oo_params (v25/functions.cpp:2702-2760) computed as the reference computes
 * it, for output-file parity (counters.dat, magmom.dat): the Q tensor of
 * Qmat_3d (3426-3453) in single precision summed in particle order, the
 * single-precision jacobi rotations (v25/jacobi.c:12-80, as compiled by the
 * reference's C++ build), and G1 and M as double sums in particle order.
 * Given the same state it returns the reference's numbers bit for bit.  One
 * wave adds the terms in order: meant for output files, not for speed.  At
 * most 4 dipole slots per particle; MSC_E_ARG if the rotations do not
 * converge in 50 sweeps (as jacobi gives up). */
int msc_order_params_ref(msc_ctx *ctx, int rep, double *G1, double *G2, double M[3]);

/* acc_correl (v25/functions.cpp:2399-2466): orientational pair histograms
 * 000, 110, 112 and 220 of replica `rep` over every pair f < g, binned by the
 * minimum-image centre distance into bars of width bar_w, ADDED into
 * hist[(set*4 + channel)*n_bars + bar] (host memory) as the reference adds
 * into Hista/Histe.  Particles form species of counts[0..nspec-1] in order
 * (sum = N, nspec <= 8); sets: one per species (pairs within it; weights 2,
 * 2 dot, 2(3 d_i d_j - dot), 3 dot^2 - 1), then one per species pair m < n in
 * row-major order (weights 1, dot, 3 d_i d_j - dot, (3 dot^2 - 1)/2).  The
 * reference sizes n_bars = floor(sqrt(3) Lmin / (2 bar_w)) (v25/main.cpp:251).
 * *dropped (may be NULL) = pairs at bar n_bars or beyond, not binned (the
 * reference writes past its array there).  FP64 in every precision mode: the
 * bars and the counts channel equal a serial evaluation exactly.  nor_correl's
 * normalisation is host work (mdspherocyl_amd.refio.correlation_text). */
int msc_pair_correl(msc_ctx *ctx, int rep, int nspec, const int *counts, double bar_w, int n_bars,
                    double *hist, long *dropped);

/* msc_pair_correl without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, checks and results as msc_pair_correl, but the histogram
 * scratch is a fixed device array of MSC_NOALLOC_CORREL_MAX doubles (2 MiB)
 * that the library holds on each device it runs on until the process ends;
 * sets * 4 * n_bars above that limit is MSC_E_CAPACITY.  The array is shared
 * by every call on one device and keeps the last call's contents (each call
 * clears what it uses first): calls that overlap in time on one device race. */
#define MSC_NOALLOC_CORREL_MAX 262144
int msc_pair_correl_noalloc(msc_ctx *ctx, int rep, int nspec, const int *counts, double bar_w, int n_bars,
                            double *hist, long *dropped);

/* ---- conf.dat (print_conf / chrg_conf, v25/functions.cpp:1181-1296) ----
 * Host-only reader and writer of the reference's configuration file, so
 * that states pass between the reference and this library.  Arrays are
 * particle-major: r, p, l [3N], ori [9N] (rows = body axes), mass [N],
 * dipole lab position and orientation dpos, dori [3 N nd], moment dmu
 * [N nd], cluster label clust [N] (NULL writes 0), thermo = {s0, strn0,
 * s1, strn1} of the two thermostats.  The writer formats as the reference's
 * stream does (first line default 6 digits, then fixed with 12 decimals), so
 * a file read and written back is byte-identical.  Readers may pass NULL for
 * arrays they do not need; MSC_E_STATE reports a file that does not hold N
 * particles of nd dipoles. */
int msc_write_conf(const char *path, int N, int nd, const double L[3], const double *r,
                   const double *p, const double *ori, const double *l, const double *mass,
                   const double *dpos, const double *dori, const double *dmu, const int *clust,
                   const double thermo[4]);
int msc_read_conf(const char *path, int N, int nd, double L[3], double *r, double *p,
                  double *ori, double *l, double *mass, double *dpos, double *dori, double *dmu,
                  int *clust, double thermo[4]);

/* ---- synthetic initial states (no reference counterpart) --------------
 * Random sequential addition of N rods of lengths len[] into a periodic
 * cubic box of edge `box` centred on the origin: r[3N] centres, u[3N] unit
 * axes, every pair at segment distance >= dmin.  Deterministic for a given
 * seed (xoshiro256** via splitmix64).  Host only.  MSC_E_CAPACITY: the box
 * could not be filled. */
int msc_place_rods(int N, double box, const double *len, double dmin, unsigned long long seed, double *r,
                   double *u);

/* ---- measurement -------------------------------------------------------
 * With timing on, the library brackets each launch of the named kernels with
 * HIP events on its stream; msc_kernel_time returns the summed milliseconds
 * and launch count since the last reset.  Names: "recip_struct",
 * "recip_force", "real_pairs", "wca", "nlist", "integrate". */
int msc_set_timing(msc_ctx *ctx, int on);
int msc_kernel_time(msc_ctx *ctx, const char *name, double *ms, long *launches);
int msc_reset_timing(msc_ctx *ctx);
/* counts of the algorithmic work units of the last force evaluation of a
 * replica: [0] WCA pairs inside 2^(1/6) sigma, [1] real-space dipole pairs
 * inside the cutoff, [2] reciprocal k-vectors (half space), [3] list pairs */
int msc_work_counts(msc_ctx *ctx, int rep, long counts[4]);

#ifdef __cplusplus
}
#endif
#endif
