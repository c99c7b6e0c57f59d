// Drop-in definitions of the reference's force-path routines
// (v25/functions.hpp:29-54, 59-61, 70) over libmsc_hip.so.
//
// Build, next to the reference's own sources (v25/):
//
//   g++ -O2 -fopenmp -c functions.cpp -o functions.o \
//       -Dtabul=ref_tabul -Dwave_v=ref_wave_v -Dneigh_list=ref_neigh_list \
//       -Dforce_sc=ref_force_sc -Dforce_dipR=ref_force_dipR -Dforce_dipK=ref_force_dipK \
//       -Dext_field=ref_ext_field -Dself_ener=ref_self_ener -Dacc_correl=ref_acc_correl \
//       -Doo_params=ref_oo_params
//   g++ -O2 -std=c++17 -c msc_v25_adapter.cpp -I. -I<repo>/include
//   g++ -fopenmp main.cpp classes.cpp functions.o msc_v25_adapter.o jacobi.o \
//       -L<repo>/mdspherocyl_amd -lmsc_hip -lgsl -lgslcblas -o DSCmd
//
// main.cpp links unmodified: the routines below carry the declarations of
// functions.hpp, and the reference's own definitions stay available under
// the ref_ names (tabul and wave_v still fill the caller's tables; force_dipR
// uses them for a table the adapter has not seen built).
//
// One FP64 context on device 0 holds the system; moveX, moveP, moveTher,
// check_list, savec_list, clust_conf and the file routines stay the
// reference's.  For a whole run on the device, msc_init_run + msc_run replace
// the step loop (v25/main.cpp:559-708) instead.
#include <cstdio>

#include "functions.hpp"
#include "msc_adapter.hpp"

// the reference's own definitions, renamed when functions.cpp is compiled
void ref_tabul(double Tab[32768][3], double alpha);
void ref_wave_v(Dynvec<Matrix<double> > &Kvec, const double L[3], double alpha, int K_max);
void ref_force_dipR(Dynvec<Dynvec<Spherocyl> > &part, const part_num &NP, const double L[3],
                    const double Tab[32768][3], double &U, double PT[3][3]);

namespace {
using V25Adapter = msc_adapter::Adapter<Dynvec<Dynvec<Spherocyl> >, part_num, Matrix<double> >;
V25Adapter &adapter() {
  static V25Adapter a;
  return a;
}
}  // namespace

// tabul (v25/functions.cpp:1680-1704): the caller's table, and alpha for the
// library's own table
void tabul(double Tab[32768][3], double alpha) {
  ref_tabul(Tab, alpha);
  adapter().note_alpha(Tab, alpha);
}

// wave_v (v25/functions.cpp:1963-1990): the caller's k weights, and alpha, K
// for the library's own
void wave_v(Dynvec<Matrix<double> > &Kvec, const double L[3], double alpha, int K_max) {
  ref_wave_v(Kvec, L, alpha, K_max);
  adapter().note_kspace(alpha, K_max);
}

void neigh_list(const Dynvec<Dynvec<Spherocyl> > &part, const part_num &NP, const double L[3], deque<int> nblist[],
                double r_list) {
  if (!adapter().neigh_list(part, NP, L, nblist, r_list)) std::fprintf(stderr, "msc adapter: neigh_list failed\n");
}

// false where the reference returns false (an overlap, or a list pair whose
// segment distance contradicts its centre distance) and on a library error
bool force_sc(Dynvec<Dynvec<Spherocyl> > &part, const part_num &NP, const double L[3], double r_cut, double &U,
              double &W, double PT[3][3], const deque<int> nblist[]) {
  return adapter().force_sc(part, NP, L, r_cut, U, W, PT, nblist) == MSC_OK;
}

void force_dipR(Dynvec<Dynvec<Spherocyl> > &part, const part_num &NP, const double L[3], const double Tab[32768][3],
                double &U, double PT[3][3]) {
  if (!adapter().knows_table(Tab)) {
    // a table not built through tabul above: its alpha is unknown here
    ref_force_dipR(part, NP, L, Tab, U, PT);
    return;
  }
  adapter().force_dipR(part, NP, L, U, PT);
}

void force_dipK(Dynvec<Dynvec<Spherocyl> > &part, const part_num &NP, const double L[3], int K_max, double alpha,
                const Dynvec<Matrix<double> > & /*Kvec: the library's weights come from K_max, alpha*/, double &U,
                double PT[3][3]) {
  adapter().force_dipK(part, NP, L, K_max, alpha, U, PT);
}

void ext_field(Dynvec<Dynvec<Spherocyl> > &part, const part_num &NP, const double B[3], double &U) {
  adapter().ext_field(part, NP, B, U);
}

double self_ener(const Dynvec<Dynvec<Spherocyl> > &part, const part_num &NP) {
  double U = 0.0;
  adapter().self_ener(part, NP, U);
  return U;
}

void acc_correl(const Dynvec<Dynvec<Spherocyl> > &part, Dynvec<Matrix<double> > &Hista,
                Dynvec<Matrix<double> > &Histe, const double L[3], double bar_w, const part_num &NP,
                const cross Vc[]) {
  adapter().acc_correl(part, Hista, Histe, L, bar_w, NP, Vc);
}

void oo_params(const Dynvec<Dynvec<Spherocyl> > &part, const part_num &NP, double &G1, double &G2, double M[3]) {
  adapter().oo_params(part, NP, G1, G2, M);
}
