// The reference's force routines over libmsc_hip.so: the part of the
// drop-in adapter that does not depend on the reference's headers.
//
// `Adapter<Parts, NPart, Mat>` works on the reference's containers through
// the accessors of v25/classes.hpp: Parts is Dynvec<Dynvec<Spherocyl> >
// (part[species][i]), NPart is part_num (get_Ncomp, get_Npart, get_Tpart),
// Mat is Matrix<double> (Spherocyl::get_omat).  Particles are numbered
// species-major, as rev_II (v25/functions.cpp:1166-1179) numbers them.
//
// Every routine gathers the state (skipping the upload when it equals the
// last one sent), runs the library routine on one FP64 replica, and adds the
// forces and torques back with add_F / add_T, as the reference routines add
// into the particles.  Energies, virials and pressure tensors are
// overwritten totals, as in the reference.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <vector>

#include "msc.h"

namespace msc_adapter {

// Run parameters the library's context is created with.  alpha and K come
// from the reference's own start-up calls tabul(Tab, alpha) and
// wave_v(Kvec, L, alpha, K) (v25/main.cpp:485-496) or from force_dipK's
// arguments; a change of N, slots, box, alpha, K or B makes a new context.
struct RunParams {
  int N = 0, nd = 0;
  double L[3] = {0, 0, 0};
  double alpha = 0.0;
  int K = 0;
  double B[3] = {0, 0, 0};
  bool operator==(const RunParams &o) const {
    return N == o.N && nd == o.nd && alpha == o.alpha && K == o.K && std::memcmp(L, o.L, sizeof L) == 0 &&
           std::memcmp(B, o.B, sizeof B) == 0;
  }
};

template <class Parts, class NPart, class Mat>
class Adapter {
 public:
  ~Adapter() {
    if (ctx_) msc_destroy(ctx_);
  }

  // tabul / wave_v record what the context needs; the tables themselves are
  // the reference's (the library builds its own from the same parameters)
  void note_alpha(const void *tab, double alpha) {
    want_.alpha = alpha;
    tab_ = tab;
  }
  void note_kspace(double alpha, int K) {
    want_.alpha = alpha;
    want_.K = K;
  }
  bool knows_table(const void *tab) const { return tab_ != nullptr && tab == tab_; }

  // neigh_list (v25/functions.cpp:1298-1334): the device builds the list;
  // the caller's nblist receives it in the reference's form
  bool neigh_list(const Parts &part, const NPart &NP, const double L[3], std::deque<int> nblist[], double r_list) {
    if (!upload(part, NP, L)) return false;
    if (!ok(msc_neigh_list(ctx_, 0, r_list), "msc_neigh_list")) return false;
    off_.assign(N_ + 1, 0);
    long tot = 0;
    if (!ok(msc_get_neigh_list(ctx_, 0, off_.data(), nullptr, 0, &tot), "msc_get_neigh_list")) return false;
    nbr_.assign(tot > 0 ? tot : 1, 0);
    if (!ok(msc_get_neigh_list(ctx_, 0, off_.data(), nbr_.data(), tot, &tot), "msc_get_neigh_list")) return false;
    for (int f = 0; f < N_; ++f) {
      nblist[f].clear();
      for (int q = off_[f]; q < off_[f + 1]; ++q) nblist[f].push_back(nbr_[q]);
    }
    list_src_ = nblist;
    list_sig_ = signature(nblist);
    r_list_ = r_list;
    list_on_device_ = true;
    return true;
  }

  // force_sc (v25/functions.cpp:1397-1553) over the caller's nblist: a list
  // other than the one the device holds is uploaded first
  int force_sc(Parts &part, const NPart &NP, const double L[3], double r_cut, double &U, double &W, double PT[3][3],
               const std::deque<int> nblist[]) {
    if (!upload(part, NP, L)) return MSC_E_HIP;
    const unsigned long long sig = signature(nblist);
    if (!list_on_device_ || nblist != list_src_ || sig != list_sig_) {
      off_.assign(N_ + 1, 0);
      nbr_.clear();
      for (int f = 0; f < N_; ++f) {
        for (int g : nblist[f]) nbr_.push_back(g);
        off_[f + 1] = (int)nbr_.size();
      }
      if (nbr_.empty()) nbr_.push_back(0);
      // the list radius is the last neigh_list's (force_sc walks the rows
      // whatever the radius; without one, r_cut stands in)
      const double r_list = r_list_ > 0 ? r_list_ : r_cut;
      const int rc = msc_set_neigh_list(ctx_, 0, r_list, off_.data(), nbr_.data());
      if (rc) return report(rc, "msc_set_neigh_list");
      list_src_ = nblist;
      list_sig_ = sig;
      list_on_device_ = true;
    }
    msc_zero_forces(ctx_, 0);
    double pt[9];
    const int rc = msc_force_sc(ctx_, 0, r_cut, &U, &W, pt);
    for (int k = 0; k < 9; ++k) PT[k / 3][k % 3] = pt[k];
    if (rc == MSC_E_OVERLAP || rc == MSC_E_DISTANCE) return rc;
    if (rc) return report(rc, "msc_force_sc");
    return add_back(part, NP) ? MSC_OK : MSC_E_HIP;
  }

  // force_dipR (v25/functions.cpp:1555-1704)
  int force_dipR(Parts &part, const NPart &NP, const double L[3], double &U, double PT[3][3]) {
    if (!upload(part, NP, L)) return MSC_E_HIP;
    msc_zero_forces(ctx_, 0);
    double pt[9];
    const int rc = msc_force_dipR(ctx_, 0, &U, pt);
    if (rc) return report(rc, "msc_force_dipR");
    for (int k = 0; k < 9; ++k) PT[k / 3][k % 3] = pt[k];
    return add_back(part, NP) ? MSC_OK : MSC_E_HIP;
  }

  // force_dipK (v25/functions.cpp:1706-1990) with the caller's K and alpha
  int force_dipK(Parts &part, const NPart &NP, const double L[3], int K, double alpha, double &U, double PT[3][3]) {
    note_kspace(alpha, K);
    if (!upload(part, NP, L)) return MSC_E_HIP;
    msc_zero_forces(ctx_, 0);
    double pt[9];
    const int rc = msc_force_dipK(ctx_, 0, &U, pt);
    if (rc) return report(rc, "msc_force_dipK");
    for (int k = 0; k < 9; ++k) PT[k / 3][k % 3] = pt[k];
    return add_back(part, NP) ? MSC_OK : MSC_E_HIP;
  }

  // ext_field (v25/functions.cpp:1992-2025): B is the context's
  int ext_field(Parts &part, const NPart &NP, const double B[3], double &U) {
    for (int a = 0; a < 3; ++a) want_.B[a] = B[a];
    if (!upload(part, NP, L_)) return MSC_E_HIP;
    msc_zero_forces(ctx_, 0);
    const int rc = msc_ext_field(ctx_, 0, &U);
    if (rc) return report(rc, "msc_ext_field");
    return add_back(part, NP) ? MSC_OK : MSC_E_HIP;
  }

  // self_ener (v25/functions.cpp:2027-2069)
  int self_ener(const Parts &part, const NPart &NP, double &U) {
    if (!upload(part, NP, L_)) return MSC_E_HIP;
    const int rc = msc_self_ener(ctx_, 0, &U);
    return rc ? report(rc, "msc_self_ener") : MSC_OK;
  }

  // acc_correl (v25/functions.cpp:2399-2468): histograms added into Hista
  // (per species) and Histe (per species pair, at the index the cross table
  // gives the pair); the library returns the species sets first, then the
  // pairs m < n in order
  template <class HistVec, class Cross>
  int acc_correl(const Parts &part, HistVec &Hista, HistVec &Histe, const double L[3], double bar_w, const NPart &NP,
                 const Cross Vc[]) {
    if (!upload(part, NP, L)) return MSC_E_HIP;
    const int nsp = NP.get_Ncomp();
    int cst = 0, nb = 0;
    Hista[0].get_size(cst, nb);
    std::vector<int> counts(nsp);
    for (int n = 0; n < nsp; ++n) counts[n] = NP.get_Npart(n);
    const int sets = nsp + nsp * (nsp - 1) / 2;
    hist_.assign((size_t)sets * 4 * nb, 0.0);
    long dropped = 0;
    const int rc = msc_pair_correl(ctx_, 0, nsp, counts.data(), bar_w, nb, hist_.data(), &dropped);
    if (rc) return report(rc, "msc_pair_correl");
    int set = nsp;
    for (int n = 0; n < nsp; ++n)
      for (int c = 0; c < 4; ++c)
        for (int b = 0; b < nb; ++b) Hista[n][c][b] += hist_[((size_t)n * 4 + c) * nb + b];
    for (int m = 0; m < nsp; ++m)
      for (int n = m + 1; n < nsp; ++n, ++set) {
        int s = 0;
        while (s < nsp - 1 && Vc[m].P[s] != n) ++s;
        const int cc = Vc[m].C[s];
        for (int c = 0; c < 4; ++c)
          for (int b = 0; b < nb; ++b) Histe[cc][c][b] += hist_[((size_t)set * 4 + c) * nb + b];
      }
    return MSC_OK;
  }

  // oo_params (v25/functions.cpp:2702-2760): G1, G2 and the total moment M
  // (overwritten); Q and its eigenvectors in FP64
  int oo_params(const Parts &part, const NPart &NP, double &G1, double &G2, double M[3]) {
    if (!upload(part, NP, L_)) return MSC_E_HIP;
    const int rc = msc_order_params(ctx_, 0, &G1, &G2, M);
    return rc ? report(rc, "msc_order_params") : MSC_OK;
  }

  msc_ctx *context() const { return ctx_; }

 private:
  msc_ctx *ctx_ = nullptr;
  RunParams have_, want_;
  const void *tab_ = nullptr;
  int N_ = 0, nd_ = 0;
  double L_[3] = {0, 0, 0};
  bool sent_ = false;
  std::vector<double> r_, p_, l_, ori_, len_, dia_, mass_, dpc_, doc_, dmu_, F_, T_, last_;
  std::vector<int> off_, nbr_;
  std::vector<double> hist_;
  const std::deque<int> *list_src_ = nullptr;
  unsigned long long list_sig_ = 0;
  double r_list_ = 0.0;
  bool list_on_device_ = false;

  bool ok(int rc, const char *what) { return report(rc, what) == MSC_OK; }
  int report(int rc, const char *what) {
    if (rc) std::fprintf(stderr, "msc adapter: %s failed (%d): %s\n", what, rc, msc_last_error(ctx_));
    return rc;
  }

  // order-sensitive fingerprint of a list (sizes and entries)
  unsigned long long signature(const std::deque<int> nblist[]) const {
    unsigned long long h = 1469598103934665603ull;
    for (int f = 0; f < N_; ++f) {
      h = (h ^ (unsigned long long)nblist[f].size()) * 1099511628211ull;
      for (int g : nblist[f]) h = (h ^ (unsigned long long)(unsigned)g) * 1099511628211ull;
    }
    return h;
  }

  // gather the particles (species-major) into the SoA host layout of msc.h
  void gather(const Parts &part, const NPart &NP) {
    const int N = N_, nd = nd_;
    r_.resize(3 * N); p_.resize(3 * N); l_.resize(3 * N); ori_.resize(9 * N);
    len_.resize(N); dia_.resize(N); mass_.resize(N);
    dpc_.resize((size_t)N * nd); dmu_.resize((size_t)N * nd); doc_.resize((size_t)3 * N * nd);
    Mat O(3, 3);  // get_omat assigns, which needs the sizes to agree
    int g = 0;
    for (int n = 0; n < NP.get_Ncomp(); ++n)
      for (int i = 0; i < NP.get_Npart(n); ++i, ++g) {
        const auto &s = part[n][i];
        s.get_pos(&r_[3 * g]);
        s.get_mom(&p_[3 * g]);
        s.get_angm(&l_[3 * g]);
        s.get_omat(O);
        for (int k = 0; k < 9; ++k) ori_[9 * g + k] = O[k / 3][k % 3];
        double sh[2];
        s.get_shape(sh);
        len_[g] = sh[0];
        dia_[g] = sh[1];
        mass_[g] = s.get_mass();
        for (int w = 0; w < nd; ++w) {
          // body-frame constraints from the lab-frame dipole: the lever
          // dpc u (get_lev) along the axis u = row 2, the orientation
          // O u_lab (rows of O are the body axes)
          double d[3], o[3];
          s.get_lev(w, d);
          const auto dip = s.get_dip(w);
          dip.get_ori(o);
          dpc_[(size_t)g * nd + w] = d[0] * O[2][0] + d[1] * O[2][1] + d[2] * O[2][2];
          dmu_[(size_t)g * nd + w] = dip.get_dip();
          for (int k = 0; k < 3; ++k)
            doc_[((size_t)g * nd + w) * 3 + k] = O[k][0] * o[0] + O[k][1] * o[1] + O[k][2] * o[2];
        }
      }
  }

  // (re)create the context when the run parameters changed, then send the
  // state unless it equals the last one sent
  bool upload(const Parts &part, const NPart &NP, const double L[3]) {
    const int N = NP.get_Tpart();
    const int nd = N > 0 ? part[0][0].get_ndips() : 0;
    RunParams want = want_;
    want.N = N;
    want.nd = nd;
    // self_ener, ext_field and oo_params get no box: the last one seen, or
    // a unit box before any (none of them depends on it)
    for (int a = 0; a < 3; ++a) want.L[a] = L[a] > 0 ? L[a] : 1.0;
    // the library needs alpha > 0 and K >= 2 even before the run's values
    // are known (force_sc and the list need neither); force_dipK brings them
    if (!(want.alpha > 0)) want.alpha = 1.0;
    if (want.K < 2) want.K = 2;
    if (!ctx_ || !(want == have_)) {
      if (ctx_) msc_destroy(ctx_);
      ctx_ = nullptr;
      msc_replica_desc d;
      std::memset(&d, 0, sizeof d);
      d.N = N;
      d.nd = nd;
      for (int a = 0; a < 3; ++a) {
        d.L[a] = want.L[a];
        d.B[a] = want.B[a];
      }
      d.alpha = want.alpha;
      d.K = want.K;
      d.sr_cut = 1.0;
      d.rc = 0.0;  // real space over every minimum-image pair, as force_dipR
      d.temp = 1.0;
      d.dt = 0.001;
      d.Q = 1.0;
      const int rc = msc_create(0, MSC_FP64, 1, &d, &ctx_);
      if (rc) {
        std::fprintf(stderr, "msc adapter: msc_create failed (%d): %s\n", rc, msc_last_error(nullptr));
        ctx_ = nullptr;
        return false;
      }
      have_ = want;
      N_ = N;
      nd_ = nd;
      for (int a = 0; a < 3; ++a) L_[a] = want.L[a];
      sent_ = false;
      list_on_device_ = false;
    }
    gather(part, NP);
    // the state as one vector, to skip an upload of an unchanged state (the
    // force routines of one step share it)
    std::vector<double> cur;
    cur.reserve(r_.size() + p_.size() + l_.size() + ori_.size() + dpc_.size() + doc_.size() + dmu_.size() +
                3 * (size_t)N);
    for (const auto *v : {&r_, &p_, &l_, &ori_, &len_, &dia_, &mass_, &dpc_, &doc_, &dmu_})
      cur.insert(cur.end(), v->begin(), v->end());
    if (sent_ && cur == last_) return true;
    const int rc = msc_set_state(ctx_, 0, r_.data(), p_.data(), l_.data(), ori_.data(), len_.data(), dia_.data(),
                                 mass_.data(), dpc_.data(), doc_.data(), dmu_.data());
    if (rc) return ok(rc, "msc_set_state");
    last_.swap(cur);
    sent_ = true;
    return true;
  }

  // the library's F and T added to the particles (add_F / add_T)
  bool add_back(Parts &part, const NPart &NP) {
    F_.resize(3 * (size_t)N_);
    T_.resize(3 * (size_t)N_);
    if (!ok(msc_get_forces(ctx_, 0, F_.data(), T_.data()), "msc_get_forces")) return false;
    int g = 0;
    for (int n = 0; n < NP.get_Ncomp(); ++n)
      for (int i = 0; i < NP.get_Npart(n); ++i, ++g) {
        part[n][i].add_F(&F_[3 * g]);
        part[n][i].add_T(&T_[3 * g]);
      }
    return true;
  }
};

}  // namespace msc_adapter
