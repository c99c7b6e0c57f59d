// Shared definitions of the MI355X force-path library (device and host).
//
// Device state is structure-of-arrays over the concatenation of all replicas
// ("global particle index" g); vector fields are stored component-major,
// v[c * Ntot + g], so that a wavefront of 64 consecutive particles reads
// contiguous 512-byte lines.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace msc {

constexpr int kWave = 64;
constexpr int kTabN = 32768;               // table length of tabul
constexpr double kTabExt = 200 * 1.001;    // table span of tabul/force_dipR
constexpr int kMaxKZ = 16;                 // largest reference K input supported
constexpr int kMaxNd = 8;                  // most dipole slots per particle

// Per-replica parameters as the kernels see them.
struct RepDev {
  int n0, N;            // first global particle, count
  int nd;               // dipole slots per particle
  int K, kmax, KZ;      // reference K input, kmax = K-1, KZ = K
  int ncol;             // (2 kmax + 1)^2 columns (nx, ny) of the k half space
  int slot0;            // first slot (ncol * KZ per replica) in the k arrays
  int tab0;             // first row of this replica's real-space table
  int cell0, ncell;     // first cell and cell count of this replica's grid
  int nc[3];            // cells per dimension (1 if the box is too small)
  int reach[3];         // stencil half-width per dimension: cells within +-reach are searched
  // Neighbour rows: entry q of particle g sits at idx[base + q*N + (g - n0)]
  // (column-major, so the lanes of a wave read consecutive words); w is the
  // row capacity.  nl: reference list (segment distance < r_list_max); sl:
  // contact list (segment distance < r_short); dl: real-space dipole pairs.
  int nl_w, sl_w, dl_w;
  // 1: the rows are the context's own (MSC_MIXED with a list skin): every
  // pair within r_short = r_wca + skin, rebuilt when twice the largest
  // segment-point displacement exceeds r_short - r_wca; 0: the reference's
  // rows at r_list_max and its rebuild test, contact rows refreshed from them
  int own_lists;
  long nl_base, sl_base, dl_base;
  long part0;           // first element of this replica's structure-factor partials
  long bop0;            // first element (4 fp16 values) of this replica's back-projection operands
  int bexp0;            // first of this replica's 3 operand scale exponents per column tile
  int task0, ntask;     // wave tasks (64 particles) of this replica
  int rtask0, nrtask;   // reciprocal chunk tasks of this replica
  double L[3], V, P2[3];
  double iL[3];         // 1 / L (single-precision pair sums' minimum image)
  double cw[3];         // cell widths
  double alpha, rc, sr_cut;
  double r_list_min, r_list_max;
  double rl_dip;        // real-space pair-list radius (rc + skin)
  double r_short;       // contact-list radius: 2^(1/6) max diameter + skin_short
  double r_wca;         // WCA range of the replica: 2^(1/6) max diameter
  double skin_short;
  double lmax;          // largest cylinder length
  double diamax;        // largest diameter
  double B[3];
  double temp, dt, Q, gamma, dens;
  // moveTher's factors (v25/functions.cpp:2196-2197), taken on the host
  // from the four above whenever the replicas are uploaded:
  // th_ex1 = exp(-gamma dt / 2), th_sq = sqrt(2 temp gamma dt / Q)
  double th_ex1, th_sq;
  double G[9];          // sum over half-space k of val_k k k^T
  double Uself;
};

// A wave task: up to 64 consecutive particles of one replica.
struct Task {
  int rep, g0, n, pad;
};

// Per-replica energies of one force evaluation (device resident).
struct RepEnergy {
  double Usc, Wsc, PTsc[9];
  double UR, PTR[9];
  double UK, PTK[9];
  double Uext;
  double Kin[2], PTkin[9];
  double pad[3];
};

// Thermostat state of a replica (two thermostats: translation, rotation).
struct Thermo {
  double s[2], strn[2];
};

#define MSC_HD __host__ __device__ __forceinline__

MSC_HD double wrapd(double d, double L) { return d - L * floor(d / L + 0.5); }

}  // namespace msc
