/*
 * msc_oracle.c -- CPU FP64 restatement of mdSpherocyl's per-step force path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.
 *
 * Each function names the reference routine it follows (paths relative to the
 * reference tree).  Expressions keep the reference's operand order and
 * C left-to-right association so that the results agree to the last bit with a
 * serial (non-OpenMP) reference build; compile with -ffp-contract=off.
 *
 * Built a second time with OpenMP (libmsc_oracle_omp.so) only to time the
 * host baseline on every core: the neighbour list, force_sc, force_dipR and
 * force_dipK loops then run on all threads, each with its own force arrays,
 * so that build's sums are taken in another order and it is never the
 * checker.
 */
#include "msc_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* MT19937, the generator behind gsl_rng_mt19937.  Published algorithm:     */
/* M. Matsumoto, T. Nishimura, ACM TOMACS 8 (1998) 3, with the 2002          */
/* init_genrand seeding; GSL maps seed 0 to 4357 and returns x/2^32 from     */
/* gsl_rng_uniform.  Used by Thermostat (v25/classes.hpp:153-160), GCAmix,   */
/* ini_mom and rnd_part.                                                     */
/* ------------------------------------------------------------------------ */
void omt_seed(omt *g, unsigned long seed) {
  if (seed == 0) seed = 4357;
  g->mt[0] = (uint32_t)(seed & 0xffffffffUL);
  for (int i = 1; i < 624; ++i) {
    uint32_t prev = g->mt[i - 1];
    g->mt[i] = (uint32_t)(1812433253UL * (prev ^ (prev >> 30)) + (uint32_t)i);
  }
  g->mti = 624;
}

uint32_t omt_u32(omt *g) {
  static const uint32_t mag01[2] = {0x0u, 0x9908b0dfu};
  const uint32_t UPPER = 0x80000000u, LOWER = 0x7fffffffu;
  if (g->mti >= 624) {
    int k;
    for (k = 0; k < 624 - 397; ++k) {
      uint32_t y = (g->mt[k] & UPPER) | (g->mt[k + 1] & LOWER);
      g->mt[k] = g->mt[k + 397] ^ (y >> 1) ^ mag01[y & 1u];
    }
    for (; k < 623; ++k) {
      uint32_t y = (g->mt[k] & UPPER) | (g->mt[k + 1] & LOWER);
      g->mt[k] = g->mt[k + (397 - 624)] ^ (y >> 1) ^ mag01[y & 1u];
    }
    uint32_t y = (g->mt[623] & UPPER) | (g->mt[0] & LOWER);
    g->mt[623] = g->mt[396] ^ (y >> 1) ^ mag01[y & 1u];
    g->mti = 0;
  }
  uint32_t y = g->mt[g->mti++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

double omt_uniform(omt *g) { return omt_u32(g) / 4294967296.0; }

/* Box-Muller pair, v25/functions.cpp:1149-1156 */
void o_gauss(double variance, double *rnd1, double *rnd2) {
  double a = sqrt(-2 * log(*rnd1)) * cos(2 * M_PI * *rnd2);
  double b = sqrt(-2 * log(*rnd1)) * sin(2 * M_PI * *rnd2);
  double sv = sqrt(variance);
  *rnd1 = a * sv;
  *rnd2 = b * sv;
}

/* ------------------------------------------------------------------------ */
/* small vector helpers: dot/cros of v25/functions.cpp:3411-3419            */
/* ------------------------------------------------------------------------ */
static inline double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void cross3(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
/* per-component minimum image, as written at v25/functions.cpp:1322-1325 */
static inline double wrapd(double d, double L) { return d - L * floor(d / L + 0.5); }

void o_minimg(const double ri[3], const double rj[3], const double L[3], double rij[3]) {
  for (int a = 0; a < 3; ++a) {
    double d = ri[a] - rj[a];
    rij[a] = d - L[a] * floor(d / L[a] + 0.5);
  }
}

/* C = A*B for 3x3 row-major, accumulating from +0 in index order like
 * Matrix<T>::operator* (v25/Matrix.hpp:383-392). */
static void mm3(const double *A, const double *B, double *C) {
  double R[9];
  for (int m = 0; m < 3; ++m)
    for (int n = 0; n < 3; ++n) {
      double acc = 0.0;
      for (int i = 0; i < 3; ++i) acc += A[m * 3 + i] * B[i * 3 + n];
      R[m * 3 + n] = acc;
    }
  memcpy(C, R, sizeof R);
}
static void mv3(const double *A, const double *x, double *y) {
  double r[3];
  for (int m = 0; m < 3; ++m) {
    double acc = 0.0;
    for (int i = 0; i < 3; ++i) acc += A[m * 3 + i] * x[i];
    r[m] = acc;
  }
  memcpy(y, r, sizeof r);
}
static void tr3(const double *A, double *B) {
  double R[9];
  for (int m = 0; m < 3; ++m)
    for (int n = 0; n < 3; ++n) R[m * 3 + n] = A[n * 3 + m];
  memcpy(B, R, sizeof R);
}

/* ------------------------------------------------------------------------ */
/* spherocyl_dist, v25/functions.cpp:89-489                                  */
/* ------------------------------------------------------------------------ */
static inline void seg_vec(const double rAB[3], double s, double t, double LA,
                           const double uA[3], double LB, const double uB[3],
                           double d[3]) {
  for (int i = 0; i < 3; ++i) d[i] = rAB[i] + (s - 0.5) * LA * uA[i] - (t - 0.5) * LB * uB[i];
}

double o_segdist(const double rA[3], const double uA[3], double LA,
                 const double rB[3], const double uB[3], double LB,
                 const double Lb[3], const double rAB[3], double dAB[3],
                 double st[2]) {
  double doto = dot3(uA, uB), dotrA = dot3(rAB, uA), dotrB = dot3(rAB, uB);
  double a = LA * LA;
  double b = doto * LA * LB;
  double c = LB * LB;
  double d = dotrA * LA - a / 2 + b / 2;
  double e = dotrB * LB + c / 2 - b / 2;
  double det = a * c - b * b;
  double s = b * e - c * d;
  double t = a * e - b * d;
  double dis[3], dsq = 0.0;

  if (det > 1.e-12) {
    double s0 = s / det, t0 = t / det, dsq0;
    seg_vec(rAB, s0, t0, LA, uA, LB, uB, dis);
    dsq0 = dot3(dis, dis);
    if (dsq0 < 1.e-12) {
      /* infinite lines intersect: endpoint search on absolute positions,
       * v25/functions.cpp:127-276 */
      double pp[4][3], dd[4], dv[4][3];
      for (int k = 0; k < 3; ++k) {
        pp[0][k] = rA[k] + LA * uA[k] / 2;
        pp[1][k] = rA[k] - LA * uA[k] / 2;
        pp[2][k] = rB[k] + LB * uB[k] / 2;
        pp[3][k] = rB[k] - LB * uB[k] / 2;
      }
      for (int k = 0; k < 3; ++k) {
        dv[0][k] = wrapd(pp[0][k] - pp[2][k], Lb[k]);
        dv[1][k] = wrapd(pp[0][k] - pp[3][k], Lb[k]);
        dv[2][k] = wrapd(pp[1][k] - pp[2][k], Lb[k]);
        dv[3][k] = wrapd(pp[1][k] - pp[3][k], Lb[k]);
      }
      for (int q = 0; q < 4; ++q) dd[q] = dot3(dv[q], dv[q]);
      int im = 0;
      double mind = 1.e6;
      for (int q = 0; q < 4; ++q)
        if (dd[q] < mind) { im = q; mind = dd[q]; }
      /* per case: t of the first try, the endpoint pair projected on A,
       * the s of the second try, the pair projected on B, the final t */
      static const int tfirst[4] = {1, 0, 1, 0};
      static const int pa_to[4] = {2, 3, 2, 3};  /* vec = pp[1] - pp[pa_to] */
      static const int ssec[4] = {1, 1, 0, 0};
      static const int pb_to[4] = {0, 0, 1, 1};  /* vec = pp[3] - pp[pb_to] */
      double vec[3];
      t = tfirst[im];
      for (int k = 0; k < 3; ++k) vec[k] = wrapd(pp[1][k] - pp[pa_to[im]][k], Lb[k]);
      s = dot3(vec, uA);
      s /= -LA;
      if (s >= 0 && s <= 1) {
        seg_vec(rAB, s, t, LA, uA, LB, uB, dAB);
        dsq = dot3(dAB, dAB);
      } else {
        s = ssec[im];
        for (int k = 0; k < 3; ++k) vec[k] = wrapd(pp[3][k] - pp[pb_to[im]][k], Lb[k]);
        t = dot3(vec, uB);
        t /= -LB;
        if (t >= 0 && t <= 1) {
          seg_vec(rAB, s, t, LA, uA, LB, uB, dAB);
          dsq = dot3(dAB, dAB);
        } else {
          t = tfirst[im];
          for (int k = 0; k < 3; ++k) dAB[k] = dv[im][k];
          dsq = dot3(dAB, dAB);
        }
      }
    } else {
      /* 3x3 region clamp of the (s,t) square, v25/functions.cpp:277-445 */
      int region;
      if (s >= 0) {
        if (s <= det) region = (t >= 0) ? ((t <= det) ? 5 : 6) : 4;
        else region = (t >= 0) ? ((t <= det) ? 2 : 3) : 1;
      } else {
        region = (t >= 0) ? ((t <= det) ? 8 : 9) : 7;
      }
#define CLAMP01(x) do { if ((x) > 1) (x) = 1; else if ((x) < 0) (x) = 0; } while (0)
      switch (region) {
        case 2: s = 1; t = (e + b) / c; CLAMP01(t); break;
        case 8: s = 0; t = e / c; CLAMP01(t); break;
        case 6: t = 1; s = -(d - b) / a; CLAMP01(s); break;
        case 4: t = 0; s = -d / a; CLAMP01(s); break;
        case 5: s /= det; t /= det; break;
        default: break;
      }
      if (region == 2 || region == 8 || region == 6 || region == 4 || region == 5) {
        seg_vec(rAB, s, t, LA, uA, LB, uB, dAB);
        dsq = dot3(dAB, dAB);
      } else {
        /* corner regions try two edges and keep the shorter */
        double s1, t1;
        if (region == 1 || region == 3) { s = 1; t = (e + b) / c; }
        else if (region == 7) { t = 0; s = -d / a; }
        else { t = 1; s = -(d - b) / a; }            /* region 9 */
        CLAMP01(s); CLAMP01(t);
        seg_vec(rAB, s, t, LA, uA, LB, uB, dAB);
        dsq = dot3(dAB, dAB);
        if (region == 1) { t1 = 0; s1 = -d / a; }
        else if (region == 3) { t1 = 1; s1 = -(d - b) / a; }
        else { s1 = 0; t1 = e / c; }                 /* regions 7, 9 */
        CLAMP01(s1); CLAMP01(t1);
        seg_vec(rAB, s1, t1, LA, uA, LB, uB, dis);
        dsq0 = dot3(dis, dis);
        if (dsq0 < dsq) {
          for (int k = 0; k < 3; ++k) dAB[k] = dis[k];
          s = s1; t = t1; dsq = dsq0;
        }
      }
#undef CLAMP01
    }
  } else {
    /* parallel axes, v25/functions.cpp:447-486: sentinel-coded (s,t) */
    double le = (LA + LB) / 2;
    if (fabs(dotrA) < le) {
      for (int k = 0; k < 3; ++k) dAB[k] = rAB[k] - dotrA * uA[k];
      dsq = dot3(dAB, dAB);
      s = t = -1.e+6;
    } else {
      double d2[3], dsq0;
      if (doto > 0) {
        for (int k = 0; k < 3; ++k) {
          dAB[k] = rAB[k] + LA * uA[k] / 2 + LB * uB[k] / 2;
          d2[k] = rAB[k] - LA * uA[k] / 2 - LB * uB[k] / 2;
        }
        dsq = dot3(dAB, dAB);
        dsq0 = dot3(d2, d2);
        s = t = -1.e+10;
        if (dsq0 < dsq) { for (int k = 0; k < 3; ++k) dAB[k] = d2[k]; dsq = dsq0; s = t = -2.e+10; }
      } else {
        for (int k = 0; k < 3; ++k) {
          dAB[k] = rAB[k] + LA * uA[k] / 2 - LB * uB[k] / 2;
          d2[k] = rAB[k] - LA * uA[k] / 2 + LB * uB[k] / 2;
        }
        dsq = dot3(dAB, dAB);
        dsq0 = dot3(d2, d2);
        s = t = -3.e+10;
        if (dsq0 < dsq) { for (int k = 0; k < 3; ++k) dAB[k] = d2[k]; dsq = dsq0; s = t = -4.e+10; }
      }
    }
  }
  st[0] = s;
  st[1] = t;
  return dsq;
}

double o_spherocyl_dist(const osys *S, int A, int B, const double rAB[3],
                        double dAB[3], double st[2]) {
  return o_segdist(S->r + 3 * A, S->ori + 9 * A + 6, S->len[A], S->r + 3 * B,
                   S->ori + 9 * B + 6, S->len[B], S->L, rAB, dAB, st);
}

/* ------------------------------------------------------------------------ */
/* neighbour list, v25/functions.cpp:1298-1394                               */
/* ------------------------------------------------------------------------ */
long o_neigh_list(const osys *S, double r_list, int *off, int *nbr, long cap) {
  double rl_sq = r_list * r_list;
  long cnt = 0;
#ifdef _OPENMP
  /* timing build: rows found in parallel, then concatenated in row order
   * (the same list as the serial loop) */
  int **rows = (int **)calloc((size_t)S->N, sizeof(int *));
  int *rn = (int *)calloc((size_t)S->N, sizeof(int));
#pragma omp parallel for schedule(dynamic, 16)
  for (int f = 0; f < S->N - 1; ++f) {
    int capr = 64, n = 0;
    int *buf = (int *)malloc(sizeof(int) * (size_t)capr);
    for (int g = f + 1; g < S->N; ++g) {
      double rij[3], st[2], dij[3];
      o_minimg(S->r + 3 * f, S->r + 3 * g, S->L, rij);
      if (o_spherocyl_dist(S, f, g, rij, dij, st) < rl_sq) {
        if (n == capr) buf = (int *)realloc(buf, sizeof(int) * (size_t)(capr *= 2));
        buf[n++] = g;
      }
    }
    rows[f] = buf;
    rn[f] = n;
  }
  for (int f = 0; f < S->N; ++f) {
    if (off) off[f] = (int)(cnt < 0x7fffffff ? cnt : 0x7fffffff);
    for (int q = 0; q < rn[f]; ++q) {
      if (nbr && cnt < cap) nbr[cnt] = rows[f][q];
      ++cnt;
    }
    free(rows[f]);
  }
  free(rows);
  free(rn);
  if (off) off[S->N] = (int)(cnt < 0x7fffffff ? cnt : 0x7fffffff);
  return cnt;
#endif
  for (int f = 0; f < S->N; ++f) {
    if (off) off[f] = (int)(cnt < 0x7fffffff ? cnt : 0x7fffffff);
    if (f == S->N - 1) continue;
    for (int g = f + 1; g < S->N; ++g) {
      double rij[3], st[2], dij[3];
      o_minimg(S->r + 3 * f, S->r + 3 * g, S->L, rij);
      double dsq = o_spherocyl_dist(S, f, g, rij, dij, st);
      if (dsq < rl_sq) {
        if (nbr && cnt < cap) nbr[cnt] = g;
        ++cnt;
      }
    }
  }
  if (off) off[S->N] = (int)(cnt < 0x7fffffff ? cnt : 0x7fffffff);
  return cnt;
}

/* savec_list writes positions as fixed with 12 decimals and check_list reads
 * them back (v25/functions.cpp:1344-1349, 1372): the reference positions are
 * the doubles nearest to the 12-decimal rounding. */
double o_round12(double x) {
  char buf[64];
  snprintf(buf, sizeof buf, "%.12f", x);
  return strtod(buf, NULL);
}

void o_savec_list(const osys *S, double *snap) {
  for (int i = 0; i < 3 * S->N; ++i) snap[i] = o_round12(S->r[i]);
}

int o_check_list(const osys *S, const double *snap, double r_list_min,
                 double r_list_max) {
  double dismax = 0.;
  for (int i = 0; i < S->N; ++i)
    for (int a = 0; a < 3; ++a) {
      double D = S->r[3 * i + a] - snap[3 * i + a];
      D -= S->L[a] * floor(D / S->L[a] + 0.5);
      double ad = fabs(D);
      dismax = (ad < dismax) ? dismax : ad; /* std::max(fabs(D), dismax) */
    }
  dismax = 2. * sqrt(3. * dismax * dismax);
  return dismax > r_list_max - r_list_min;
}

/* ------------------------------------------------------------------------ */
/* force_sc, v25/functions.cpp:1397-1553                                     */
/* returns 0, or 1 on overlap (d^2 < 0.49 sigma^2), 2 on d^2 > 1.00001 r^2.   */
/* ------------------------------------------------------------------------ */
/* Row id of the list (the pairs (id, nbr[q])): forces and torques into F
 * and T, energy and virial into *Uc, *Wc, the pressure tensor into P; the
 * overlap tests set *abort_code as the reference's loop does. */
static void sc_row(const osys *S, int id, const int *off, const int *nbr, double r_cut_sq, double srtsq,
                   double *F, double *T, double *Uc, double *Wc, double P[9], int *abort_code) {
  const double *ri = S->r + 3 * id, *oi = S->ori + 9 * id + 6;
  for (int q = off[id]; q < off[id + 1]; ++q) {
    int jd = nbr[q];
    const double *rj = S->r + 3 * jd, *oj = S->ori + 9 * jd + 6;
    double sigma = (S->dia[id] + S->dia[jd]) / 2;
    double sig_sq = sigma * sigma;
    double rsh = srtsq * sig_sq;
    double rij[3], rsq = 0;
    for (int a = 0; a < 3; ++a) {
      rij[a] = ri[a] - rj[a];
      rij[a] -= S->L[a] * floor(rij[a] / S->L[a] + 0.5);
      rsq += rij[a] * rij[a];
    }
    double st[2], dij[3];
    double dsq = o_spherocyl_dist(S, id, jd, rij, dij, st);
    if (dsq < 0.49 * sig_sq) *abort_code = 1;
    if (dsq > rsq * 1.00001 && !*abort_code) *abort_code = 2;
    if (dsq < r_cut_sq) {
      double SR2s = sig_sq / dsq;
      double SR6s = SR2s * SR2s * SR2s;
      double pot = 4 * SR6s * (SR6s - 1.);
      double UIJ = 0., WIJ = 0.;
      double ddr = dot3(rij, dij);
      if (dsq < rsh) {
        UIJ = pot + 1.;
        WIJ = 6 * (pot + 4 * SR6s * SR6s) * ddr / dsq;
      }
      *Uc += UIJ;
      *Wc += WIJ;
      double FIJ = WIJ / ddr, f[3], t[3];
      for (int a = 0; a < 3; ++a) f[a] = FIJ * dij[a];
      for (int a = 0; a < 3; ++a) F[3 * id + a] += f[a];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) P[3 * a + b] += rij[a] * f[b];
      for (int a = 0; a < 3; ++a) f[a] *= -1;
      for (int a = 0; a < 3; ++a) F[3 * jd + a] += f[a];
      double icroj[3], icror[3], jcror[3];
      cross3(oi, oj, icroj);
      cross3(oi, rij, icror);
      cross3(oj, rij, jcror);
      double fact1 = (st[0] - 0.5) * S->len[id];
      double fact2 = -(st[1] - 0.5) * S->len[jd];
      if (st[0] == -1.e+6) fact1 = fact2 = 0.;
      /* sentinel overrides; the reference tests -3e10 twice, so the (-/+)
       * code -4e10 keeps the factors computed above (v25/functions.cpp:1526-1532) */
      if (st[0] == -1.e+10) { fact1 = 0.5 * S->len[id]; fact2 = 0.5 * S->len[jd]; }
      else if (st[0] == -2.e+10) { fact1 = -0.5 * S->len[id]; fact2 = -0.5 * S->len[jd]; }
      else if (st[0] == -3.e+10) { fact1 = 0.5 * S->len[id]; fact2 = -0.5 * S->len[jd]; }
      double fmul = fact1 * fact2;
      for (int a = 0; a < 3; ++a) t[a] = FIJ * (fact1 * icror[a] + fmul * icroj[a]);
      for (int a = 0; a < 3; ++a) T[3 * id + a] += t[a];
      for (int a = 0; a < 3; ++a) t[a] = FIJ * (fact2 * jcror[a] - fmul * icroj[a]);
      for (int a = 0; a < 3; ++a) T[3 * jd + a] += t[a];
    }
  }
}

int o_force_sc(osys *S, double r_cut, const int *off, const int *nbr,
               double *U, double *W, double PT[9]) {
  double r_cut_sq = r_cut * r_cut;
  double Uc = 0., Wc = 0., P[9] = {0};
  double srt = pow(2., 1. / 6);
  double srtsq = srt * srt;
  int abort_code = 0;
#ifdef _OPENMP
  /* timing build: rows split over threads, each with its own force arrays */
#pragma omp parallel
  {
    double *Ft = (double *)calloc(6 * (size_t)S->N, sizeof(double)), *Tt = Ft + 3 * (size_t)S->N;
    double Ut = 0., Wt = 0., Pt[9] = {0};
    int at = 0;
#pragma omp for schedule(dynamic, 64)
    for (int id = 0; id < S->N; ++id) sc_row(S, id, off, nbr, r_cut_sq, srtsq, Ft, Tt, &Ut, &Wt, Pt, &at);
#pragma omp critical
    {
      for (size_t q = 0; q < 3 * (size_t)S->N; ++q) { S->F[q] += Ft[q]; S->T[q] += Tt[q]; }
      Uc += Ut; Wc += Wt;
      for (int k = 0; k < 9; ++k) P[k] += Pt[k];
      if (at && !abort_code) abort_code = at;
    }
    free(Ft);
  }
#else
  for (int id = 0; id < S->N && !abort_code; ++id)
    sc_row(S, id, off, nbr, r_cut_sq, srtsq, S->F, S->T, &Uc, &Wc, P, &abort_code);
#endif
  if (abort_code) return abort_code;
  *U = Uc;
  *W = Wc;
  for (int k = 0; k < 9; ++k) PT[k] = P[k];
  return 0;
}

/* ------------------------------------------------------------------------ */
/* tabul + force_dipR, v25/functions.cpp:1555-1704                           */
/* ------------------------------------------------------------------------ */
void o_tabul(double *Tab, double alpha) {
  double extL = 200 * 1.001;
  double alsq = alpha * alpha;
  double alft = alsq * alsq;
  double spi = sqrt(M_PI);
  for (int i = 1; i < 32769; i++) {
    double r = 0.5 * extL * i / 32768;
    double rsq = r * r;
    double sr1 = 1.0 / r;
    double sr2 = sr1 * sr1;
    double sr3 = sr2 * sr1;
    double ERFC = erfc(alpha * r);
    double aer = 2.0 * alpha * exp(-alsq * rsq) / spi;
    double aa = 2.0 * alsq + 3.0 * sr2;
    Tab[3 * (i - 1) + 0] = ERFC * sr3 + aer * sr2;
    Tab[3 * (i - 1) + 1] = (3.0 * ERFC * sr3 + aer * aa) * sr2;
    Tab[3 * (i - 1) + 2] = 4. * alft * aer;
  }
}

/* rc <= 0: every minimum-image pair of dipoles on distinct particles, as the
 * reference; rc > 0: only dipole pairs closer than rc (no reference
 * counterpart; used for the cut-off configurations). */
/* Row k of the pair loop (the pairs (k, l > k)), accumulating into F, T,
 * *Uc and P. */
static void dipR_row(const osys *S, const double *Tab, double rc, double Lh, int k, double *F, double *T,
                     double *Uc, double P[9]) {
  int nd = S->nd;
  for (int l = k + 1; l < S->N; ++l) {
    for (int v = 0; v < nd; ++v) {
      int dv = k * nd + v;
      double di[3];
      for (int a = 0; a < 3; ++a) di[a] = S->dpc[dv] * S->ori[9 * k + 6 + a];
      const double *ri = S->dr + 3 * dv, *oi = S->du + 3 * dv;
      double mui = S->dmu[dv];
      if (mui == 0.) continue;
      for (int w = 0; w < nd; ++w) {
        int dw = l * nd + w;
        double dj[3];
        for (int a = 0; a < 3; ++a) dj[a] = S->dpc[dw] * S->ori[9 * l + 6 + a];
        const double *rj = S->dr + 3 * dw, *oj = S->du + 3 * dw;
        double muj = S->dmu[dw];
        if (muj == 0.) continue;
        double rij[3];
        for (int a = 0; a < 3; ++a) {
          rij[a] = ri[a] - rj[a];
          rij[a] -= S->L[a] * floor(rij[a] / S->L[a] + 0.5);
        }
        double RIJSQ = dot3(rij, rij);
        double uu = mui * muj;
        double SR2 = 1. / RIJSQ;
        double RIJ = sqrt(RIJSQ);
        if (rc > 0 && !(RIJ < rc)) continue;
        int ix = (int)floor(RIJ / Lh);
        /* the reference reads Tab[-1] below Lh and past the end above
         * 100.1; both are undefined there, clamped here */
        if (ix < 1) ix = 1;
        if (ix > 32767) ix = 32767;
        double R_L0 = RIJ - Lh * ix;
        double R_L1 = RIJ - Lh * (ix + 1);
        const double *T1 = Tab + 3 * ix, *T0 = Tab + 3 * (ix - 1);
        double A = (T1[0] * R_L0 - T0[0] * R_L1) / Lh;
        double B = (T1[1] * R_L0 - T0[1] * R_L1) / Lh;
        double C = (T1[2] * R_L0 - T0[2] * R_L1) / Lh;
        double UIR = dot3(oi, rij), UJR = dot3(oj, rij);
        double UIJR = UIR * UJR;
        double UIUJ = dot3(oi, oj);
        double UIJ = uu * (UIUJ * A - UIJR * B);
        *Uc += UIJ;
        double fact = (UIUJ - 5. * UIJR * SR2) * B - UIJR * SR2 * C;
        double f[3], t[3];
        for (int a = 0; a < 3; ++a) f[a] = uu * (fact * rij[a] + (UIR * oj[a] + UJR * oi[a]) * B);
        for (int a = 0; a < 3; ++a) F[3 * k + a] += f[a];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) P[3 * a + b] += rij[a] * f[b];
        for (int a = 0; a < 3; ++a) f[a] *= -1;
        for (int a = 0; a < 3; ++a) F[3 * l + a] += f[a];
        double icror[3], dicror[3], dicroi[3], jcror[3], djcror[3], djcroj[3];
        double dicroj[3], djcroi[3], icroj[3];
        cross3(oi, rij, icror);
        cross3(di, rij, dicror);
        cross3(di, oi, dicroi);
        cross3(oj, rij, jcror);
        cross3(dj, rij, djcror);
        cross3(dj, oj, djcroj);
        cross3(di, oj, dicroj);
        cross3(dj, oi, djcroi);
        cross3(oi, oj, icroj);
        for (int a = 0; a < 3; ++a) icroj[a] *= A;
        fact = -UIR * UJR * (5 * B + C) * SR2 + UIUJ * B;
        double fact1;
        for (int a = 0; a < 3; ++a) {
          fact1 = (UJR * (icror[a] + dicroi[a]) + UIR * dicroj[a]) * B;
          t[a] = uu * (fact * dicror[a] + fact1 - icroj[a]);
        }
        for (int a = 0; a < 3; ++a) T[3 * k + a] += t[a];
        for (int a = 0; a < 3; ++a) {
          fact1 = (UIR * (jcror[a] - djcroj[a]) - UJR * djcroi[a]) * B;
          t[a] = uu * (-fact * djcror[a] + fact1 + icroj[a]);
        }
        for (int a = 0; a < 3; ++a) T[3 * l + a] += t[a];
      }
    }
  }
}

void o_force_dipR(osys *S, const double *Tab, double rc, double *U, double PT[9]) {
  double Uc = 0., P[9] = {0};
  double Lh = 0.5 * (200 * 1.001) / 32768;
#ifdef _OPENMP
  /* timing build: rows split over threads, each with its own force arrays */
#pragma omp parallel
  {
    double *Ft = (double *)calloc(6 * (size_t)S->N, sizeof(double)), *Tt = Ft + 3 * (size_t)S->N;
    double Ut = 0., Pt[9] = {0};
#pragma omp for schedule(dynamic, 16)
    for (int k = 0; k < S->N - 1; ++k) dipR_row(S, Tab, rc, Lh, k, Ft, Tt, &Ut, Pt);
#pragma omp critical
    {
      for (size_t q = 0; q < 3 * (size_t)S->N; ++q) { S->F[q] += Ft[q]; S->T[q] += Tt[q]; }
      Uc += Ut;
      for (int q = 0; q < 9; ++q) P[q] += Pt[q];
    }
    free(Ft);
  }
#else
  for (int k = 0; k < S->N - 1; ++k) dipR_row(S, Tab, rc, Lh, k, S->F, S->T, &Uc, P);
#endif
  *U = Uc;
  for (int k = 0; k < 9; ++k) PT[k] = P[k];
}

/* ------------------------------------------------------------------------ */
/* wave_v + force_dipK, v25/functions.cpp:1706-1990                          */
/* ------------------------------------------------------------------------ */
void o_wave_v(double *Kvec, const double L[3], double alpha, int K) {
  double c = 4. * alpha * alpha;
  double P2[3] = {2. * M_PI / L[0], 2. * M_PI / L[1], 2. * M_PI / L[2]};
  int K1 = K + 1;
  for (int i = 0; i < K1 * K1 * K1; ++i) Kvec[i] = 0.;
  for (int nz = 0; nz <= K; ++nz) {
    double Kz = P2[2] * nz;
    for (int ny = 0; ny <= K; ++ny) {
      double Ky = P2[1] * ny;
      for (int nx = 0; nx <= K; ++nx) {
        double Kx = P2[0] * nx;
        double Ksq = Kx * Kx + Ky * Ky + Kz * Kz;
        if (Ksq != 0.) Kvec[(nx * K1 + ny) * K1 + nz] = 4. * M_PI * exp(-Ksq / c) / Ksq;
      }
    }
  }
}

typedef struct { double re, im; } cplx;
static inline cplx cmul(cplx a, cplx b) {  /* GCC _Complex product, finite operands */
  cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
static inline cplx cscale(cplx a, double x) { cplx r = {a.re * x, a.im * x}; return r; }
static inline cplx cadd(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx cconj(cplx a) { cplx r = {a.re, -a.im}; return r; }
static inline cplx csub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static inline double cnorm(cplx a) { return a.re * a.re + a.im * a.im; }

/* k-vector kn of force_dipK (its four sign variants): forces and torques
 * into F and T, energy into *Uc, pressure tensor into PP; M is scratch of
 * 4 N nd values. */
static void dipK_kvec(const osys *S, int kn, int K, const double *Kvec, double a4, cplx *M, double *F,
                      double *T, double *Uc, double PP[9]) {
  int NC = S->nspec, nd = S->nd, N = S->N;
  double V = S->L[0] * S->L[1] * S->L[2];
  double P2[3] = {2. * M_PI / S->L[0], 2. * M_PI / S->L[1], 2. * M_PI / S->L[2]};
  int K2 = K * K, K1 = K + 1;
  const cplx cc = {0., 0.};
  int nz = (int)floorf((float)kn / K2);
  float lf = (float)(kn - K2 * nz);
  int ny = (int)floorf(lf / K);
  int nx = (int)lf - K * ny;
  double Kz = P2[2] * nz, Ky = P2[1] * ny, Kx = P2[0] * nx;
  double Ksq = Kx * Kx + Ky * Ky + Kz * Kz;
  if (Ksq == 0) return;
  double KK = Kvec[(nx * K1 + ny) * K1 + nz];
  double val = 2. * KK / V;
  double pval = 2 * (a4 + 1. / Ksq);
  double Kb[4][3] = {{Kx, Ky, Kz}, {-Kx, Ky, Kz}, {Kx, -Ky, Kz}, {-Kx, -Ky, Kz}};
  cplx sum[4][OSYS_MAXSPEC], uve[4][3];
  double self[4][OSYS_MAXSPEC], uvslf[4][3];
  for (int b = 0; b < 4; ++b) {
    for (int k = 0; k < 3; ++k) { uve[b][k] = cc; uvslf[b][k] = 0.; }
    for (int n = 0; n < NC; ++n) { sum[b][n] = cc; self[b][n] = 0.; }
  }
  for (size_t q = 0; q < 4 * (size_t)N * nd; ++q) M[q] = cc;
  /* generate values, v25/functions.cpp:1780-1827 */
  int base = 0;
  for (int n = 0; n < NC; ++n) {
    for (int ii = 0; ii < S->counts[n]; ++ii) {
      int i = base + ii;
      cplx selfpart[4] = {cc, cc, cc, cc}, uvslfpart[4][3];
      for (int b = 0; b < 4; ++b)
        for (int k = 0; k < 3; ++k) uvslfpart[b][k] = cc;
      for (int w = 0; w < nd; ++w) {
        int dw = i * nd + w;
        double ri[3], oi[3];
        for (int k = 0; k < 3; ++k) {
          ri[k] = S->dr[3 * dw + k];
          ri[k] -= S->L[k] * floor(ri[k] / S->L[k] + 0.5);
          oi[k] = S->du[3 * dw + k];
        }
        double mu = S->dmu[dw];
        if (mu == 0.) continue;
        double RIK[4], UIK[4];
        cplx z[4];
        for (int b = 0; b < 4; ++b) {
          RIK[b] = dot3(ri, Kb[b]);
          UIK[b] = dot3(oi, Kb[b]);
          UIK[b] *= mu;
          z[b].re = cos(RIK[b]);
          z[b].im = sin(RIK[b]);
        }
        for (int b = 0; b < 4; ++b) {
          cplx m = cconj(z[b]);
          M[((size_t)b * N + i) * nd + w] = m;
          cplx ukexp = cscale(m, UIK[b]);
          sum[b][n] = cadd(sum[b][n], ukexp);
          selfpart[b] = cadd(selfpart[b], ukexp);
          for (int k = 0; k < 3; ++k) {
            ukexp = cscale(m, mu * oi[k]);
            uve[b][k] = cadd(uve[b][k], ukexp);
            uvslfpart[b][k] = cadd(uvslfpart[b][k], ukexp);
          }
        }
      }
      for (int b = 0; b < 4; ++b) {
        self[b][n] += cnorm(selfpart[b]);
        for (int k = 0; k < 3; ++k)
          uvslf[b][k] += cmul(cconj(selfpart[b]), uvslfpart[b][k]).re;
      }
    }
    base += S->counts[n];
  }
  cplx adds[4] = {cc, cc, cc, cc};
  double addslf[4] = {0., 0., 0., 0.};
  for (int b = 0; b < 4; ++b)
    for (int n = 0; n < NC; ++n) { adds[b] = cadd(adds[b], sum[b][n]); addslf[b] += self[b][n]; }
  double Upar = 0;
  for (int b = 0; b < 4; ++b) Upar += val * 0.5 * (cnorm(adds[b]) - addslf[b]);
  int quarter = (nx == 0 && ny == 0) || (nx == 0 && nz == 0) || (ny == 0 && nz == 0);
  int half = !quarter && ((nz == 0 && nx != 0 && ny != 0) || (ny == 0 && nz != 0 && nx != 0) ||
                          (nx == 0 && ny != 0 && nz != 0));
  if (quarter) Upar /= 4.;
  else if (half) Upar /= 2.;
  *Uc += Upar;
  /* force and torque, v25/functions.cpp:1853-1925 */
  for (int i = 0; i < N; ++i) {
    cplx local[4];
    double UIKw[4][8];
    for (int b = 0; b < 4; ++b) {
      local[b] = cc;
      for (int w = 0; w < nd; ++w) {
        int dw = i * nd + w;
        double mu = S->dmu[dw];
        if (mu != 0) {
          UIKw[b][w] = dot3(S->du + 3 * dw, Kb[b]);
          UIKw[b][w] *= mu;
          local[b] = cadd(local[b], cscale(M[((size_t)b * N + i) * nd + w], UIKw[b][w]));
        }
      }
    }
    for (int w = 0; w < nd; ++w) {
      int dw = i * nd + w;
      double mu = S->dmu[dw];
      if (mu == 0) continue;
      double di[3], oi[3];
      for (int a = 0; a < 3; ++a) {
        di[a] = S->dpc[dw] * S->ori[9 * i + 6 + a];
        oi[a] = S->du[3 * dw + a];
      }
      double icrok[4][3], dicrok[4][3];
      for (int b = 0; b < 4; ++b) {
        cross3(di, Kb[b], dicrok[b]);
        cross3(oi, Kb[b], icrok[b]);
        for (int j = 0; j < 3; ++j) icrok[b][j] *= mu;
      }
      double f[3] = {0., 0., 0.}, t[3] = {0., 0., 0.};
      for (int b = 0; b < 4; ++b) {
        cplx m = M[((size_t)b * N + i) * nd + w];
        cplx X = csub(cconj(adds[b]), cconj(local[b]));
        double factF = -val * cmul(cscale(m, UIKw[b][w]), X).im;
        for (int j = 0; j < 3; ++j) f[j] += factF * Kb[b][j];
        double factT = -val * cmul(m, X).re;
        for (int j = 0; j < 3; ++j) t[j] += factT * icrok[b][j] + factF * dicrok[b][j];
      }
      if (quarter) { for (int j = 0; j < 3; ++j) { f[j] /= 4.; t[j] /= 4.; } }
      else if (half) { for (int j = 0; j < 3; ++j) { f[j] /= 2.; t[j] /= 2.; } }
      for (int j = 0; j < 3; ++j) { F[3 * i + j] += f[j]; }
      for (int j = 0; j < 3; ++j) { T[3 * i + j] += t[j]; }
    }
  }
  /* pressure tensor, v25/functions.cpp:1926-1952 (a4 = 1/(4 alpha) as written) */
  double Pk[9] = {0};
  for (int b = 0; b < 4; ++b)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        Pk[3 * j + k] += val * 0.5 *
            (((double)(j == k) - pval * Kb[b][j] * Kb[b][k]) * (cnorm(adds[b]) - addslf[b]) +
             2 * (cmul(cconj(adds[b]), uve[b][j]).re - uvslf[b][j]) * Kb[b][k]);
  if (quarter) { for (int q = 0; q < 9; ++q) Pk[q] /= 4.; }
  else if (half) { for (int q = 0; q < 9; ++q) Pk[q] /= 2.; }
  for (int q = 0; q < 9; ++q) PP[q] += Pk[q];
}

void o_force_dipK(osys *S, int K, double alpha, const double *Kvec, double *U,
                  double PT[9]) {
  int nd = S->nd, N = S->N;
  int K3 = K * K * K;
  double a4 = 1. / (4. * alpha);
  double Uc = 0., PP[9] = {0};
#ifdef _OPENMP
  /* timing build: k-vectors split over threads, each with its own arrays */
#pragma omp parallel
  {
    cplx *Mt = (cplx *)malloc(sizeof(cplx) * 4 * (size_t)N * nd);
    double *Ft = (double *)calloc(6 * (size_t)N, sizeof(double)), *Tt = Ft + 3 * (size_t)N;
    double Ut = 0., Pt[9] = {0};
#pragma omp for schedule(dynamic)
    for (int kn = 0; kn < K3; ++kn) dipK_kvec(S, kn, K, Kvec, a4, Mt, Ft, Tt, &Ut, Pt);
#pragma omp critical
    {
      for (size_t q = 0; q < 3 * (size_t)N; ++q) { S->F[q] += Ft[q]; S->T[q] += Tt[q]; }
      Uc += Ut;
      for (int q = 0; q < 9; ++q) PP[q] += Pt[q];
    }
    free(Mt);
    free(Ft);
  }
#else
  cplx *M = (cplx *)malloc(sizeof(cplx) * 4 * (size_t)N * nd);
  for (int kn = 0; kn < K3; ++kn) dipK_kvec(S, kn, K, Kvec, a4, M, S->F, S->T, &Uc, PP);
  free(M);
#endif
  *U = Uc;
  for (int q = 0; q < 9; ++q) PT[q] = PP[q];
}

/* ext_field, v25/functions.cpp:1992-2025 */
void o_ext_field(osys *S, const double B[3], double *U) {
  double Uext = 0.;
  for (int i = 0; i < S->N; ++i)
    for (int w = 0; w < S->nd; ++w) {
      int dw = i * S->nd + w;
      const double *oi = S->du + 3 * dw;
      double mu = S->dmu[dw], doti = dot3(oi, B), t[3];
      Uext -= mu * doti;
      cross3(oi, B, t);
      t[0] *= mu; t[1] *= mu; t[2] *= mu;
      for (int a = 0; a < 3; ++a) S->T[3 * i + a] += t[a];
    }
  *U = Uext;
}

/* self_ener, v25/functions.cpp:2027-2069: intra-particle dipole pairs of the
 * first particle of each species, times the species count. */
double o_self_ener(const osys *S) {
  double U = 0.;
  int base = 0;
  for (int n = 0; n < S->nspec; ++n) {
    double Un = 0.;
    int i = base;
    for (int v = 0; v < S->nd - 1; ++v) {
      int dv = i * S->nd + v;
      for (int w = v + 1; w < S->nd; ++w) {
        int dw = i * S->nd + w;
        double uu = S->dmu[dv] * S->dmu[dw];
        double rij[3];
        for (int a = 0; a < 3; ++a) rij[a] = S->dr[3 * dv + a] - S->dr[3 * dw + a];
        double RIJSQ = dot3(rij, rij);
        double RIJ = sqrt(RIJSQ);
        double SR2 = 1. / RIJSQ;
        double SR3 = SR2 / RIJ;
        double UIUJ = dot3(S->du + 3 * dv, S->du + 3 * dw);
        double UIR = dot3(S->du + 3 * dv, rij);
        double UJR = dot3(S->du + 3 * dw, rij);
        Un += uu * (UIUJ * SR3 - 3. * UIR * UJR * SR2 * SR3);
      }
    }
    U += S->counts[n] * Un;
    base += S->counts[n];
  }
  return U;
}

/* ------------------------------------------------------------------------ */
/* rigid body, v25/classes.cpp:148-325                                        */
/* ------------------------------------------------------------------------ */
void o_set_inertia(osys *S, int i) {
  double R = S->dia[i] / 2, Rsq = R * R, Rft = Rsq * Rsq, Lq = S->len[i];
  double Lsq = Lq * Lq;
  double vol = M_PI * Rsq * (Lq + 3. * R / 4);   /* reference formula, as written */
  double mdens = S->mass[i] / vol;
  S->inert[3 * i + 0] = mdens * M_PI * (Lsq * Lq * Rsq / 12 + Lsq * Rsq * R / 3 + Lq * Rft / 4 + 8. * Rft * R / 15);
  S->inert[3 * i + 1] = S->inert[3 * i + 0];
  S->inert[3 * i + 2] = mdens * M_PI * (Lq * Rft / 2 + 8. * Rft * R / 15);
}

/* Spherocyl::set_dip: lab dipole position and orientation from the body
 * constraints (v25/classes.cpp:148-173). */
static void dip_update(osys *S, int i, int do_ori) {
  const double *o = S->ori + 9 * i;
  for (int w = 0; w < S->nd; ++w) {
    int dw = i * S->nd + w;
    for (int a = 0; a < 3; ++a) S->dr[3 * dw + a] = S->r[3 * i + a] + S->dpc[dw] * o[6 + a];
    if (do_ori) {
      double ot[9], v[3];
      tr3(o, ot);
      mv3(ot, S->doc + 3 * dw, v);
      for (int a = 0; a < 3; ++a) S->du[3 * dw + a] = v[a];
    }
  }
}
void o_set_dip(osys *S, int i) { dip_update(S, i, 1); }

void o_translate(osys *S, int i, double dt) {
  for (int j = 0; j < 3; j++) S->r[3 * i + j] += dt * S->p[3 * i + j] / S->mass[i];
  dip_update(S, i, 0);
}

/* 5-factor symplectic splitting Rx Ry Rz Ry Rx on the body-frame angular
 * momentum (v25/classes.cpp:241-305, J. Chem. Phys. 107, 5840). */
void o_rotate(osys *S, int i, double dt) {
  double *O = S->ori + 9 * i;
  const double *I = S->inert + 3 * i;
  double R[9], Q[9], l0[3], l1[3], ang, ca, sa, dt2 = dt / 2;
  for (int k = 0; k < 3; ++k) l0[k] = S->l[3 * i + k];
  mv3(O, l0, l1);
#define RX(c, s) { R[0]=1.; R[1]=0.; R[2]=0.; R[3]=0.; R[4]=c; R[5]=-s; R[6]=0.; R[7]=s; R[8]=c; }
#define RY(c, s) { R[0]=c; R[1]=0.; R[2]=s; R[3]=0.; R[4]=1.; R[5]=0.; R[6]=-s; R[7]=0.; R[8]=c; }
#define RZ(c, s) { R[0]=c; R[1]=-s; R[2]=0.; R[3]=s; R[4]=c; R[5]=0.; R[6]=0.; R[7]=0.; R[8]=1.; }
  ang = -l1[0] * dt2 / I[0]; ca = cos(ang); sa = sin(ang); RX(ca, sa);
  mv3(R, l1, l0); mm3(R, O, Q);
  ang = -l0[1] * dt2 / I[1]; ca = cos(ang); sa = sin(ang); RY(ca, sa);
  mv3(R, l0, l1); mm3(R, Q, O);
  ang = -l1[2] * dt / I[2]; ca = cos(ang); sa = sin(ang); RZ(ca, sa);
  mv3(R, l1, l0); mm3(R, O, Q);
  ang = -l0[1] * dt2 / I[1]; ca = cos(ang); sa = sin(ang); RY(ca, sa);
  mv3(R, l0, l1); mm3(R, Q, O);
  ang = -l1[0] * dt2 / I[0]; ca = cos(ang); sa = sin(ang); RX(ca, sa);
  mv3(R, l1, l0); mm3(R, O, Q);
  memcpy(O, Q, sizeof Q);
#undef RX
#undef RY
#undef RZ
  dip_update(S, i, 1);
}

/* re-orthonormalise the body basis, v25/classes.cpp:307-325 */
void o_renorm(osys *S, int i) {
  double *o = S->ori + 9 * i;
  double norm = sqrt(o[6] * o[6] + o[7] * o[7] + o[8] * o[8]);
  o[6] /= norm; o[7] /= norm; o[8] /= norm;
  double Up[3] = {o[6] + o[3], o[7] + o[4], o[8] + o[5]};
  double d = Up[0] * o[6] + Up[1] * o[7] + Up[2] * o[8];
  o[3] = Up[0] - d * o[6];
  o[4] = Up[1] - d * o[7];
  o[5] = Up[2] - d * o[8];
  norm = sqrt(o[3] * o[3] + o[4] * o[4] + o[5] * o[5]);
  o[3] /= norm; o[4] /= norm; o[5] /= norm;
  o[0] = o[4] * o[8] - o[7] * o[5];
  o[1] = o[5] * o[6] - o[8] * o[3];
  o[2] = o[3] * o[7] - o[6] * o[4];
}

/* moveX, v25/functions.cpp:2071-2092 */
void o_moveX(osys *S, double dt) {
  for (int i = 0; i < S->N; ++i) { o_translate(S, i, dt); o_rotate(S, i, dt); }
}

/* moveP, v25/functions.cpp:2094-2184 */
void o_moveP(osys *S, const otherm th[2], double dt, double K[2], double PT[9], int last) {
  double Dt2 = dt / 2, Kc0 = 0., Kc1 = 0.;
  double P00 = 0., P01 = 0., P02 = 0., P11 = 0., P12 = 0., P22 = 0.;
  for (int i = 0; i < S->N; ++i) {
    double m = S->mass[i];
    const double *I = S->inert + 3 * i;
    double *p = S->p + 3 * i, *l = S->l + 3 * i;
    const double *f = S->F + 3 * i, *t = S->T + 3 * i;
    double I0[9] = {1. / I[0], 0., 0., 0., 1. / I[1], 0., 0., 0., 1. / I[2]};
    double ot[9], A[9], Ib[9];
    tr3(S->ori + 9 * i, ot);
    mm3(ot, I0, A);
    mm3(A, S->ori + 9 * i, Ib);
    double psq = 0., lsq = 0., sfac;
    if (!last) {
      sfac = 1. + Dt2 * th[0].s;
      for (int j = 0; j < 3; j++) { p[j] = (p[j] + Dt2 * f[j]) / sfac; psq += p[j] * p[j]; }
      sfac = 1. + Dt2 * th[1].s;
      for (int j = 0; j < 3; j++) l[j] = (l[j] + Dt2 * t[j]) / sfac;
    } else {
      sfac = 1. - Dt2 * th[0].s;
      for (int j = 0; j < 3; j++) { p[j] = p[j] * sfac + Dt2 * f[j]; psq += p[j] * p[j]; }
      P00 += p[0] * p[0] / m; P01 += p[0] * p[1] / m; P02 += p[0] * p[2] / m;
      P11 += p[1] * p[1] / m; P12 += p[1] * p[2] / m;
      P22 += p[2] * p[2] / m;
      sfac = 1. - Dt2 * th[1].s;
      for (int j = 0; j < 3; j++) l[j] = l[j] * sfac + Dt2 * t[j];
    }
    /* lres = transpose(ll) * Ib * ll */
    double v[3];
    for (int n = 0; n < 3; ++n) {
      double acc = 0.0;
      for (int k = 0; k < 3; ++k) acc += l[k] * Ib[k * 3 + n];
      v[n] = acc;
    }
    double lres = 0.0;
    for (int k = 0; k < 3; ++k) lres += v[k] * l[k];
    lsq += lres;
    Kc0 += psq / m;
    Kc1 += lsq;
  }
  K[0] = Kc0 / 2;
  K[1] = Kc1 / 2;
  double P[9] = {P00, P01, P02, P01, P11, P12, P02, P12, P22};
  for (int q = 0; q < 9; ++q) PT[q] = P[q];
}

/* Nose-Hoover-Langevin thermostat variable, v25/functions.cpp:2186-2218 */
void o_moveTher(otherm th[2], double temp, double dt, const double K[2], double Tpart, int mode) {
  for (int i = 0; i < 2; ++i) {
    double Q = th[i].Q, s, g, ex1, r0, r1;
    switch (mode) {
      case 0:
        r0 = omt_uniform(&th[i].ran);
        r1 = omt_uniform(&th[i].ran);
        o_gauss(1, &r0, &r1);
        th[i].strn = r1;
        g = th[i].g;
        ex1 = exp(-g * dt / 2);
        s = ex1 * th[i].s + sqrt(2. * temp * g * dt / Q) * r0;
        th[i].s = s;
        break;
      case 1:
        th[i].s = th[i].s + dt * (2. * K[i] - 3. * Tpart * temp) / Q;
        break;
      case 2:
        g = th[i].g;
        ex1 = exp(-g * dt / 2);
        th[i].s = ex1 * th[i].s + sqrt(2. * temp * g * dt / Q) * th[i].strn;
        break;
      default: break;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* initial state: ini_pos, GCAmix, ini_mom                                    */
/* ------------------------------------------------------------------------ */
static int spec_start(const osys *S, int n) {
  int b = 0;
  for (int m = 0; m < n; ++m) b += S->counts[m];
  return b;
}

/* stacked hexagonal lattice with all axes along z, v25/functions.cpp:491-675 */
int o_ini_pos(osys *S) {
  double sq = sqrt(3.) / 2., sqt = 1. / (2 * sqrt(3.)), sqtt = sqrt(6.) / 3;
  const double *L = S->L;
  static const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int n = 0; n < S->nspec; ++n) {
    int b0 = spec_start(S, n), NN = S->counts[n];
    double dia = S->dia[b0], len = S->len[b0];
    double m1 = floor(L[0] / dia), m2 = floor(L[1] / (dia * sq)), m3 = floor(L[2] / (dia * sqtt + len));
    int num1 = (int)(m1 < 1. ? 1. : m1), num2 = (int)(m2 < 1. ? 1. : m2), num3 = (int)(m3 < 1. ? 1. : m3);
    int cont = 0, passes = 0;
    while (cont < NN) {
      if (++passes > 2) return 1; /* lattice too small: the reference re-fills the same sites */
      for (int nz = 0; nz < num3; ++nz)
        for (int ny = 0; ny < num2; ++ny)
          for (int nx = 0; nx < num1; ++nx) {
            if (cont >= NN) continue;
            double r[3];
            if (fmod((double)nz, 2.) == 0) {
              if (fmod((double)ny, 2.) == 0) r[0] = dia * (double)nx * 1.0001;
              else r[0] = dia * ((double)nx + 0.5) * 1.0001;
              r[1] = dia * (double)ny * sq * 1.0001;
            } else {
              if (fmod((double)ny, 2.) != 0) r[0] = dia * (double)nx * 1.0001;
              else r[0] = dia * ((double)nx + 0.5) * 1.0001;
              r[1] = dia * ((double)ny * sq + sqt) * 1.0001;
            }
            r[2] = (dia * sqtt + len) * (double)nz * 1.0001;
            for (int a = 0; a < 3; ++a) r[a] -= L[a] * floor(r[a] / L[a] + 0.5);
            int i = b0 + cont;
            memcpy(S->r + 3 * i, r, sizeof r);
            memcpy(S->ori + 9 * i, I3, sizeof I3);
            if (n > 0) {
              int clash = 0;
              for (int m = 0; m < n && !clash; ++m) {
                int c0 = spec_start(S, m);
                double dd = (dia + S->dia[c0]) / 2, rsq = dd * dd;
                for (int j = c0; j < c0 + S->counts[m]; ++j) {
                  double rij[3], st[2], dij[3];
                  for (int a = 0; a < 3; ++a) rij[a] = wrapd(r[a] - S->r[3 * j + a], L[a]);
                  if (o_spherocyl_dist(S, i, j, rij, dij, st) < rsq) { clash = 1; break; }
                }
              }
              if (!clash) cont++;
            } else {
              cont++;
            }
          }
    }
  }
  /* overlap test, v25/functions.cpp:611-673 */
  for (int n = 0; n < S->nspec; ++n) {
    int b0 = spec_start(S, n);
    double rsq = S->dia[b0] * S->dia[b0];
    for (int i = b0; i < b0 + S->counts[n] - 1; ++i)
      for (int j = i + 1; j < b0 + S->counts[n]; ++j) {
        double rij[3], st[2], dij[3];
        o_minimg(S->r + 3 * i, S->r + 3 * j, L, rij);
        if (o_spherocyl_dist(S, i, j, rij, dij, st) < rsq) return 1;
      }
  }
  for (int m = 0; m < S->nspec - 1; ++m)
    for (int n = m + 1; n < S->nspec; ++n) {
      int a0 = spec_start(S, m), b0 = spec_start(S, n);
      double dd = (S->dia[a0] + S->dia[b0]) / 2, rsq = dd * dd;
      for (int i = a0; i < a0 + S->counts[m]; ++i)
        for (int j = b0; j < b0 + S->counts[n]; ++j) {
          double rij[3], st[2], dij[3];
          o_minimg(S->r + 3 * i, S->r + 3 * j, L, rij);
          if (o_spherocyl_dist(S, i, j, rij, dij, st) < rsq) return 1;
        }
    }
  return 0;
}

/* pivot-reflection cluster move, v25/functions.cpp:3328-3409 */
void o_gcamix(osys *S, omt *ran) {
  int N = S->N;
  const double *L = S->L;
  char *added = (char *)calloc((size_t)N, 1);
  int *queue = (int *)malloc(sizeof(int) * (size_t)N);
  double pivot[3], rnd;
  for (int a = 0; a < 3; ++a) {
    rnd = omt_uniform(ran);
    pivot[a] = (2 * rnd - 1) * L[a] / 2;
  }
  rnd = omt_uniform(ran);
  int id = (int)floor(rnd * N);
  int seed = 0, Nto = 0;
  for (int n = 0; n < S->nspec; ++n) {
    Nto += S->counts[n];
    if (id < Nto) { int loc = Nto - id; seed = (Nto - S->counts[n]) + (S->counts[n] - loc); break; }
  }
  int head = 0, tail = 0;
  added[seed] = 1;
  queue[tail++] = seed;
  while (head < tail) {
    int q = queue[head];
    double ri[3];
    for (int a = 0; a < 3; ++a) {
      double rij = pivot[a] - S->r[3 * q + a];
      rij -= L[a] * floor(rij / L[a] + 0.5);
      ri[a] = S->r[3 * q + a] + 2 * rij;
      ri[a] -= L[a] * floor(ri[a] / L[a] + 0.5);
    }
    memcpy(S->r + 3 * q, ri, sizeof ri);
    int qs = 0;
    for (int n = 0, b = 0; n < S->nspec; b += S->counts[n], ++n) if (q >= b) qs = b;
    for (int n = 0, b = 0; n < S->nspec; b += S->counts[n], ++n) {
      double dd = (S->dia[qs] + S->dia[b]) / 2, rsq = dd * dd;
      for (int i = b; i < b + S->counts[n]; ++i) {
        if (added[i]) continue;
        double rij[3], st[2], dij[3];
        for (int a = 0; a < 3; ++a) rij[a] = wrapd(ri[a] - S->r[3 * i + a], L[a]);
        if (o_spherocyl_dist(S, q, i, rij, dij, st) < rsq) { queue[tail++] = i; added[i] = 1; }
      }
    }
    ++head;
  }
  free(added);
  free(queue);
}

/* Maxwell momenta with sign-flip zeroing, v25/functions.cpp:677-1146.  The
 * reference indexes the per-species scratch with the global counter, which is
 * only defined for a single species; more species are rejected here. */
static void zero_sum(double *pp, int Ntot, double pr) {
  double P = 0.;
  for (int i = 0; i < Ntot; ++i) P += pp[i];
  int II = 0;
  long scont = 0;
  while (fabs(P) > pr) {
    while (P > pr) {
      if (pp[II] > 0) { pp[II] = -pp[II]; P += 2 * pp[II]; }
      II++;
      if (II == Ntot - 1) II = 0;
    }
    while (P < -pr) {
      if (pp[II] < 0) { pp[II] = -pp[II]; P += 2 * pp[II]; }
      II++;
      if (II == Ntot - 1) II = 0;
    }
    if (scont > 1e8) break;
    scont++;
  }
  P -= pp[Ntot - 1];
  pp[Ntot - 1] = -P;
}

/* ini_mom's per-species scratch (v25/functions.cpp:683-689, 915-920): for
 * each species n and component j an array of N_n + 1 doubles, allocated
 * species-major (n outer, j inner).  The copy into the global list reads
 * rnd[j][n][cnt] with the GLOBAL counter cnt (v25/functions.cpp:717-726,
 * 945-954), i.e. past the end of species n's array for every n > 0.  In the
 * reference's build those arrays are consecutive heap chunks of
 * 8 (N_n + 1) + 8 bytes rounded up to 16, each chunk's size word (size | 1,
 * "previous chunk in use") sitting in the 8 bytes after the array, and the
 * memory behind the last array is untouched (zero).  The reads therefore
 * land on the size word and on the following arrays' values; this image
 * lays the arrays out that way so that the same reads return the same
 * values.  The word after the last array (the heap's free-space size) is
 * taken as zero: it is read as a denormal of magnitude below 1e-300, which
 * no printed value can see.  Pinned by v25/output.dat lines 1220-1880
 * (two species of 50, every value equal at 12 decimals). */
typedef struct heap_image {
  double *img;           /* arrays and size words, then Ntot zeros */
  long start[OSYS_MAXSPEC][3];
} heap_image;

static void heap_image_make(heap_image *H, const osys *S) {
  long tot = 0;
  for (int n = 0; n < S->nspec; ++n) tot += 3L * (S->counts[n] + 2);
  H->img = (double *)calloc((size_t)(tot + S->N + 1), sizeof(double));
  long at = 0;
  for (int n = 0; n < S->nspec; ++n)
    for (int j = 0; j < 3; ++j) {
      H->start[n][j] = at;
      at += S->counts[n] + 1;
      const int last = (n == S->nspec - 1 && j == 2);
      const uint64_t chunk = ((uint64_t)(8 * (S->counts[n] + 1) + 8 + 15) & ~(uint64_t)15) | 1u;
      double word = 0.0;
      if (!last) memcpy(&word, &chunk, sizeof word);
      H->img[at++] = word;
    }
}

/* Maxwell draws of one kind (momenta or angular momenta) into the image,
 * species by species as v25/functions.cpp:690-710 and 921-941 */
static void ini_mom_draw(heap_image *H, const osys *S, omt *ran, const double *var /* [nspec][3] */) {
  for (int n = 0; n < S->nspec; ++n)
    for (int i = 0; i < S->counts[n]; ++i)
      for (int j = 0; j < 3; ++j) H->img[H->start[n][j] + i] = omt_uniform(ran);
  for (int n = 0; n < S->nspec; ++n) {
    const int NN = S->counts[n];
    for (int i = 0; i < NN; i += 2) {
      if ((fmodf((float)NN, 2) != 0.) && (i == NN - 1)) {
        for (int j = 0; j < 3; j++) {
          double rr = omt_uniform(ran);
          o_gauss(var[3 * n + j], &rr, &H->img[H->start[n][j] + i]);
        }
      } else {
        for (int j = 0; j < 3; j++) o_gauss(var[3 * n + j], &H->img[H->start[n][j] + i], &H->img[H->start[n][j] + i + 1]);
      }
    }
  }
}

int o_ini_mom(osys *S, double temp, omt *ran) {
  const int Ntot = S->N;
  double var[3 * OSYS_MAXSPEC];
  double *pp[3], *tt[3];
  for (int j = 0; j < 3; ++j) {
    pp[j] = (double *)calloc((size_t)Ntot + 1, sizeof(double));
    tt[j] = (double *)calloc((size_t)Ntot + 1, sizeof(double));
  }
  /* translation: variance temp * mass of the species' first particle */
  heap_image H;
  heap_image_make(&H, S);
  for (int n = 0, b = 0; n < S->nspec; b += S->counts[n], ++n)
    for (int j = 0; j < 3; ++j) var[3 * n + j] = temp * S->mass[b];
  ini_mom_draw(&H, S, ran, var);
  for (int n = 0, cnt = 0; n < S->nspec; ++n)
    for (int i = 0; i < S->counts[n]; ++i, ++cnt)
      for (int j = 0; j < 3; ++j) pp[j][cnt] = H.img[H.start[n][j] + cnt];
  free(H.img);
  for (int j = 0; j < 3; ++j) zero_sum(pp[j], Ntot, 0.1);
  for (int i = 0; i < Ntot; ++i)
    for (int j = 0; j < 3; ++j) S->p[3 * i + j] = pp[j][i];
  /* rotation: variance temp * I_j of the species' first particle */
  heap_image_make(&H, S);
  for (int n = 0, b = 0; n < S->nspec; b += S->counts[n], ++n)
    for (int j = 0; j < 3; ++j) var[3 * n + j] = temp * S->inert[3 * b + j];
  ini_mom_draw(&H, S, ran, var);
  for (int n = 0, cnt = 0; n < S->nspec; ++n)
    for (int i = 0; i < S->counts[n]; ++i, ++cnt)
      for (int j = 0; j < 3; ++j) tt[j][cnt] = H.img[H.start[n][j] + cnt];
  free(H.img);
  for (int j = 0; j < 3; ++j) zero_sum(tt[j], Ntot, 0.1);
  for (int i = 0; i < Ntot; ++i)
    for (int j = 0; j < 3; ++j) S->l[3 * i + j] = tt[j][i];
  for (int j = 0; j < 3; ++j) { free(pp[j]); free(tt[j]); }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* clusters: rnd_part / dist_test / clust_gen / clust_conf,                  */
/* v25/functions.cpp:2950-3148, 3282-3324                                     */
/* ------------------------------------------------------------------------ */
/* acc_correl (v25/functions.cpp:2399-2466): orientational pair histograms
 * 000/110/112/220 over every pair f < g, by minimum-image centre distance.
 * hist[(set*4 + channel)*n_bars + bin]; sets: species n intra (n < nspec),
 * then the species pairs m < n in row-major order (the reference's Vc[m].C,
 * v25/main.cpp:120-142).  Added into hist.  A pair whose bin is n_bars or
 * beyond is counted in the return value and not binned (the reference
 * writes past its array there).  Serial, so the sums are in pair order. */
long o_acc_correl(const osys *S, double bar_w, int n_bars, double *hist) {
  const int N = S->N, NC = S->nspec;
  int spec[OSYS_MAXSPEC + 1];
  spec[0] = 0;
  for (int n = 0; n < NC; ++n) spec[n + 1] = spec[n] + S->counts[n];
  long dropped = 0;
  for (int f = 0; f < N - 1; ++f) {
    int m = 0;
    while (f >= spec[m + 1]) ++m;
    const double *ri = S->r + 3 * f, *oi = S->ori + 9 * f + 6;
    for (int g = f + 1; g < N; ++g) {
      int n = m;
      while (g >= spec[n + 1]) ++n;
      const double *rj = S->r + 3 * g, *oj = S->ori + 9 * g + 6;
      double XIJ = ri[0] - rj[0];
      XIJ -= S->L[0] * floor(XIJ / S->L[0] + 0.5);
      double YIJ = ri[1] - rj[1];
      YIJ -= S->L[1] * floor(YIJ / S->L[1] + 0.5);
      double ZIJ = ri[2] - rj[2];
      ZIJ -= S->L[2] * floor(ZIJ / S->L[2] + 0.5);
      double RIJ = sqrt(XIJ * XIJ + YIJ * YIJ + ZIJ * ZIJ);
      double dot = oi[0] * oj[0] + oi[1] * oj[1] + oi[2] * oj[2];
      double dotsq = dot * dot;
      double doti = (oi[0] * XIJ + oi[1] * YIJ + oi[2] * ZIJ) / RIJ;
      double dotj = (oj[0] * XIJ + oj[1] * YIJ + oj[2] * ZIJ) / RIJ;
      int bin = (int)floor(RIJ / bar_w);
      if (bin >= n_bars) {
        ++dropped;
        continue;
      }
      double *h;
      if (m == n) {
        h = hist + (size_t)m * 4 * n_bars;
        h[bin] += 2.;
        h[n_bars + bin] += 2. * dot;
        h[2 * n_bars + bin] += 2. * (3. * doti * dotj - dot);
        h[3 * n_bars + bin] += 3. * dotsq - 1.;
      } else {
        int cc = 0;
        for (int a = 0; a < m; ++a) cc += NC - 1 - a;
        cc += n - m - 1;
        h = hist + (size_t)(NC + cc) * 4 * n_bars;
        h[bin] += 1.;
        h[n_bars + bin] += dot;
        h[2 * n_bars + bin] += 3. * doti * dotj - dot;
        h[3 * n_bars + bin] += (3. * dotsq - 1.) / 2;
      }
    }
  }
  return dropped;
}

/* Jacobi eigenvalue iteration of a symmetric 3x3 matrix: the cyclic method
 * with the threshold and skip rules of Numerical Recipes in C (2nd ed.,
 * section 11.1, routine jacobi), as v25/jacobi.c:12-80 runs it in single
 * precision when compiled as C++ (v25/Makefile:10): fabs and sqrt of a float
 * stay float, the double constants 0.2, 100.0, 0.5 and 1.0 promote their
 * expression to double before the float store.  a is overwritten above the
 * diagonal; d gets the eigenvalues, the columns of v the eigenvectors.
 * Returns the rotations made, -1 after 50 sweeps. */
int o_jacobi3(float a[3][3], float d[3], float v[3][3]) {
  float b[3], z[3];
  int nrot = 0;
  for (int p = 0; p < 3; ++p) {
    for (int q = 0; q < 3; ++q) v[p][q] = 0.0f;
    v[p][p] = 1.0f;
    b[p] = d[p] = a[p][p];
    z[p] = 0.0f;
  }
  for (int sweep = 1; sweep <= 50; ++sweep) {
    float sm = 0.0f;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) sm += fabsf(a[p][q]);
    if (sm == 0.0f) return nrot;
    const float tresh = sweep < 4 ? (float)(0.2 * sm / (3 * 3)) : 0.0f;
    for (int p = 0; p < 2; ++p) {
      for (int q = p + 1; q < 3; ++q) {
        const float g = (float)(100.0 * fabsf(a[p][q]));
        if (sweep > 4 && (float)(fabsf(d[p]) + g) == fabsf(d[p]) && (float)(fabsf(d[q]) + g) == fabsf(d[q])) {
          a[p][q] = 0.0f;
        } else if (fabsf(a[p][q]) > tresh) {
          float h = d[q] - d[p], t;
          if ((float)(fabsf(h) + g) == fabsf(h)) {
            t = a[p][q] / h;
          } else {
            const float theta = (float)(0.5 * h / a[p][q]);
            t = (float)(1.0 / (fabsf(theta) + sqrt(1.0 + theta * theta)));
            if (theta < 0.0) t = -t;
          }
          const float c = (float)(1.0 / sqrtf(1 + t * t));
          const float s = t * c;
          const float tau = (float)(s / (1.0 + c));
          h = t * a[p][q];
          z[p] -= h;
          z[q] += h;
          d[p] -= h;
          d[q] += h;
          a[p][q] = 0.0f;
          /* plane rotation of the pair (x, y): x - s(y + x tau), y + s(x - y tau) */
#define O_ROT(X, Y)                      \
  do {                                   \
    const float gx = (X), hy = (Y);      \
    (X) = gx - s * (hy + gx * tau);      \
    (Y) = hy + s * (gx - hy * tau);      \
  } while (0)
          for (int j = 0; j < p; ++j) O_ROT(a[j][p], a[j][q]);
          for (int j = p + 1; j < q; ++j) O_ROT(a[p][j], a[j][q]);
          for (int j = q + 1; j < 3; ++j) O_ROT(a[p][j], a[q][j]);
          for (int j = 0; j < 3; ++j) O_ROT(v[j][p], v[j][q]);
#undef O_ROT
          ++nrot;
        }
      }
    }
    for (int p = 0; p < 3; ++p) {
      b[p] += z[p];
      d[p] = b[p];
      z[p] = 0.0f;
    }
  }
  return -1;
}

/* oo_params + Qmat_3d (v25/functions.cpp:2702-2760, 3426-3453): total
 * dipole moment M (sum of moment times the stored lab-frame dipole axis),
 * the nematic order G2 = the largest |eigenvalue| of the single-precision
 * Q tensor sum_i (3 u_i u_i - I) / (2N), and the polar order G1 = |sum_i
 * Ve . u_i| / N along its unit eigenvector Ve.  M is overwritten (the
 * reference's caller zeroes it before every call, v25/main.cpp:723). */
int o_oo_params(const osys *S, double *G1, double *G2, double M[3]) {
  const int N = S->N;
  M[0] = M[1] = M[2] = 0.;
  for (int i = 0; i < N; ++i)
    for (int w = 0; w < S->nd; ++w) {
      const size_t q = (size_t)i * S->nd + w;
      for (int l = 0; l < 3; ++l) M[l] += S->dmu[q] * S->du[3 * q + l];
    }
  float Q[3][3] = {{0}}, d[3], V[3][3];
  for (int k = 0; k < N; ++k) {
    const double *o = S->ori + 9 * k + 6;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        Q[i][j] += ((float)(3 * o[i] * o[j]) - (float)(i == j ? 1 : 0)) / (2 * N);
  }
  const int nrot = o_jacobi3(Q, d, V);
  const float e1 = fabsf(d[0]), e2 = fabsf(d[1]), e3 = fabsf(d[2]);
  int col;
  if (e1 >= e2 && e1 >= e3) {
    *G2 = (double)e1;
    col = 0;
  } else if (e2 > e1 && e2 >= e3) {
    *G2 = (double)e2;
    col = 1;
  } else {
    *G2 = (double)e3;
    col = 2;
  }
  float Ve[3] = {V[0][col], V[1][col], V[2][col]};
  const float Vesq = sqrtf(Ve[0] * Ve[0] + Ve[1] * Ve[1] + Ve[2] * Ve[2]);
  Ve[0] /= Vesq;
  Ve[1] /= Vesq;
  Ve[2] /= Vesq;
  double g1 = 0.;
  for (int k = 0; k < N; ++k) {
    const double *o = S->ori + 9 * k + 6;
    g1 += Ve[0] * o[0] + Ve[1] * o[1] + Ve[2] * o[2];
  }
  *G1 = fabs(g1) / N;
  return nrot;
}

int o_dist_test(const osys *S, int A, int B, double min_s) {
  double sigma = (S->dia[A] + S->dia[B]) / 2 + min_s;
  double sig_sq = sigma * sigma, rij[3], st[2], dij[3];
  o_minimg(S->r + 3 * A, S->r + 3 * B, S->L, rij);
  return o_spherocyl_dist(S, A, B, rij, dij, st) < sig_sq;
}

/* Bare dipole-dipole energy of particles A and B summed over their dipole
 * slots (minimum image of the site separation):
 *   mu_a mu_b (u_a.u_b r^2 - 3 (u_a.r)(u_b.r)) / (r^4 r).
 * No reference counterpart: the energy clause of the workloads' dipolar
 * contact criterion.  Callers pass the lower index as A, so the operation
 * order (slots of A outer, of B inner) does not depend on which particle a
 * search meets first. */
double o_udd(const osys *S, int A, int B) {
  double U = 0.;
  for (int v = 0; v < S->nd; ++v) {
    const int da = A * S->nd + v;
    const double ma = S->dmu[da];
    if (ma == 0.) continue;
    for (int w = 0; w < S->nd; ++w) {
      const int db = B * S->nd + w;
      const double mb = S->dmu[db];
      if (mb == 0.) continue;
      double r[3];
      for (int a = 0; a < 3; ++a) {
        r[a] = S->dr[3 * da + a] - S->dr[3 * db + a];
        r[a] -= S->L[a] * floor(r[a] / S->L[a] + 0.5);
      }
      const double *ua = S->du + 3 * da, *ub = S->du + 3 * db;
      const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      const double uar = ua[0] * r[0] + ua[1] * r[1] + ua[2] * r[2];
      const double ubr = ub[0] * r[0] + ub[1] * r[1] + ub[2] * r[2];
      const double uab = ua[0] * ub[0] + ua[1] * ub[1] + ua[2] * ub[2];
      U += ma * mb * (uab * r2 - 3.0 * uar * ubr) / (r2 * r2 * sqrt(r2));
    }
  }
  return U;
}

/* The contact test of the cluster search: dist_test(A, B) and, when u_cut
 * is finite, the dipolar pair energy below u_cut. */
static int contact_test(const osys *S, int A, int B, double min_s, double u_cut) {
  if (!o_dist_test(S, A, B, min_s)) return 0;
  if (isinf(u_cut)) return 1;
  return (A < B ? o_udd(S, A, B) : o_udd(S, B, A)) < u_cut;
}

long o_contact_pairs(const osys *S, long n, const int *A, const int *B, double min_s, double u_cut,
                     unsigned char *out) {
  long hits = 0;
  for (long k = 0; k < n; ++k) {
    const int a = A[k], b = B[k];
    out[k] = (unsigned char)(contact_test(S, a, b, min_s, u_cut) || contact_test(S, b, a, min_s, u_cut));
    hits += out[k];
  }
  return hits;
}

int o_clust_conf(osys *S, double min_s, omt *ran) { return o_clust_conf_dd(S, min_s, INFINITY, ran); }

/* clust_conf (v25/functions.cpp:3282-3324) with rnd_part (2950-2970) and
 * clust_gen's breadth-first search (3018-3148); u_cut = +inf is the
 * reference's purely geometric criterion. */
int o_clust_conf_dd(osys *S, double min_s, double u_cut, omt *ran) {
  int N = S->N, NC = S->nspec;
  int *NI[OSYS_MAXSPEC], nI[OSYS_MAXSPEC], base[OSYS_MAXSPEC];
  for (int n = 0, b = 0; n < NC; b += S->counts[n], ++n) {
    base[n] = b;
    nI[n] = S->counts[n];
    NI[n] = (int *)malloc(sizeof(int) * (size_t)(nI[n] + 1));
    for (int j = 0; j < nI[n]; ++j) NI[n][j] = j;
  }
  char *added = (char *)calloc((size_t)N, 1), *locadd = (char *)malloc((size_t)N);
  int *queue = (int *)malloc(sizeof(int) * (size_t)N);
  int n_clust = 0, tsize = 1;
  while (tsize > 0) {
    int Tot = 0;
    for (int n = 0; n < NC; ++n) Tot += nI[n];
    double rnd = omt_uniform(ran);
    int ip = (int)floor(rnd * Tot), comp = 0, loc = 0, Ntn = 0, Nto = 0;
    for (int n = 0; n < NC; ++n) {
      Nto = Ntn;
      Ntn += nI[n];
      if (ip < Ntn) { comp = n; loc = ip - Nto; break; }
    }
    int seed = base[comp] + NI[comp][loc];
    memset(locadd, 0, (size_t)N);
    added[seed] = locadd[seed] = 1;
    S->clust[seed] = n_clust;
    memmove(NI[comp] + loc, NI[comp] + loc + 1, sizeof(int) * (size_t)(nI[comp] - loc - 1));
    nI[comp]--;
    int head = 0, tail = 0;
    queue[tail++] = seed;
    while (head < tail) {
      int q = queue[head];
      for (int o = 0; o < NC; ++o)
        for (int jj = 0; jj < S->counts[o]; ++jj) {
          int II = base[o] + jj;
          if (locadd[II]) continue;
          if (!contact_test(S, q, II, min_s, u_cut)) continue;
          if (added[II]) continue; /* the reference logs "classifies for 2 clusters" */
          queue[tail++] = II;
          int l2 = 0;
          for (int k = 0; k < nI[o]; ++k) if (NI[o][k] == jj) { l2 = k; break; }
          memmove(NI[o] + l2, NI[o] + l2 + 1, sizeof(int) * (size_t)(nI[o] - l2 - 1));
          nI[o]--;
          locadd[II] = added[II] = 1;
          S->clust[II] = n_clust;
        }
      ++head;
    }
    tsize = 0;
    for (int n = 0; n < NC; ++n) tsize += nI[n];
    n_clust++;
  }
  for (int n = 0; n < NC; ++n) free(NI[n]);
  free(added); free(locadd); free(queue);
  return n_clust;
}

/* ------------------------------------------------------------------------ */
/* run driver, restating v25/main.cpp:153-177, 401-497, 520-708               */
/* ------------------------------------------------------------------------ */
static void zero_FT(osys *S) {
  memset(S->F, 0, sizeof(double) * 3 * (size_t)S->N);
  memset(S->T, 0, sizeof(double) * 3 * (size_t)S->N);
}

static int list_build(osys *S, orun *R) {
  long need = o_neigh_list(S, R->r_list_max, R->off, NULL, 0);
  if (need > R->cap) {
    free(R->nbr);
    R->cap = need + need / 4 + 16;
    R->nbr = (int *)malloc(sizeof(int) * (size_t)R->cap);
  }
  o_neigh_list(S, R->r_list_max, R->off, R->nbr, R->cap);
  o_savec_list(S, R->snap);
  return 0;
}

int o_forces(osys *S, const orun_params *P, orun *R, double *e) {
  double Uss = 0, Wss = 0, UR = 0, UK = 0, Ue = 0;
  double PTss[9] = {0}, PTR[9] = {0}, PTK[9] = {0};
  zero_FT(S);
  int rc = o_force_sc(S, P->sr_cut, R->off, R->nbr, &Uss, &Wss, PTss);
  if (rc) return rc;
  if (P->dipR_on) o_force_dipR(S, R->Tab, P->rc_real, &UR, PTR);
  if (P->dipK_on) o_force_dipK(S, P->K, P->alpha, R->Kvec, &UK, PTK);
  o_ext_field(S, P->B, &Ue);
  if (e) {
    e[0] = Uss; e[1] = Wss; e[2] = UR; e[3] = UK; e[4] = Ue;
    memcpy(e + 5, PTss, sizeof PTss);
    memcpy(e + 14, PTR, sizeof PTR);
    memcpy(e + 23, PTK, sizeof PTK);
  }
  return 0;
}

int o_run_init(osys *S, const orun_params *P, orun *R, int seed_init) {
  memset(R, 0, sizeof *R);
  int N = S->N;
  R->dmax = 0; R->dmin = 1.e6;
  for (int n = 0, b = 0; n < S->nspec; b += S->counts[n], ++n) {
    double dm = S->len[b] + S->dia[b];
    R->dmax = (R->dmax < dm) ? dm : R->dmax;
    R->dmin = (S->dia[b] < R->dmin) ? S->dia[b] : R->dmin;
  }
  for (int i = 0; i < 2; ++i) {
    R->th[i].Q = P->Q; R->th[i].g = P->gamma; R->th[i].s = P->s0; R->th[i].strn = 0.;
    omt_seed(&R->th[i].ran, (unsigned long)(i + 1));
  }
  omt_seed(&R->ran, 1);
  if (seed_init) {
    if (o_ini_pos(S)) return 10;
    for (int i = 0; i < 1000; ++i) o_gcamix(S, &R->ran);
    for (int i = 0; i < N; ++i) o_set_dip(S, i);
    if (o_ini_mom(S, P->temp, &R->ran)) return 11;
  }
  R->Tab = (double *)calloc(32768 * 3, sizeof(double));
  o_tabul(R->Tab, P->alpha);
  int K1 = P->K + 1;
  R->Kvec = (double *)calloc((size_t)K1 * K1 * K1, sizeof(double));
  o_wave_v(R->Kvec, S->L, P->alpha, P->K);
  R->r_list_min = P->sr_cut + 0.1 * R->dmin;
  if (R->r_list_min < 1.5 * R->dmax) R->r_list_min = 1.5 * R->dmax;
  R->r_list_max = R->r_list_min + 0.9 * R->dmin;
  R->off = (int *)malloc(sizeof(int) * ((size_t)N + 1));
  R->snap = (double *)malloc(sizeof(double) * 3 * (size_t)N);
  list_build(S, R);
  int rc = o_forces(S, P, R, NULL);
  if (rc) return rc;
  R->Uself = o_self_ener(S) / N;
  return 0;
}

int o_run_steps(osys *S, const orun_params *P, orun *R, int step0, int nsteps, double *obs) {
  int N = S->N;
  double V = S->L[0] * S->L[1] * S->L[2];
  for (int step = step0; step < step0 + nsteps; ++step) {
    if (fmodf((float)step, (float)100) == 0)
      for (int i = 0; i < N; ++i) o_renorm(S, i);
    double Ktr[2] = {0., 0.}, PTkin[9], e[32];
    o_moveTher(R->th, P->temp, P->dt, Ktr, N, 0);
    o_moveP(S, R->th, P->dt, Ktr, PTkin, 0);
    o_moveTher(R->th, P->temp, P->dt, Ktr, N, 1);
    o_moveX(S, P->dt);
    int rc = o_forces(S, P, R, e);
    if (rc) return rc;
    o_moveP(S, R->th, P->dt, Ktr, PTkin, 1);
    o_moveTher(R->th, P->temp, P->dt, Ktr, N, 2);
    double Uss = e[0], Wss = e[1], UdipR = e[2], UdipK = e[3], Uext = e[4];
    double WdipR = 3 * UdipR, WdipK = 3 * UdipK;
    double K = (Ktr[0] + Ktr[1]) / N;
    Uss /= N; UdipR /= N; UdipK /= N; Uext /= N;
    double U = R->Uself + Uss + UdipR + UdipK + Uext;
    double Pss = Wss / (3 * V), PdipR = WdipR / (3 * V), PdipK = WdipK / (3 * V);
    double ctemp = K / 3;
    double Pr = ctemp * P->dens + Pss + PdipR + PdipK;
    /* each part divided by V, then summed in this order (v25/v251/main.cpp:635-642) */
    for (int q = 0; q < 9; ++q) {
      double kin = PTkin[q] / V, ss = e[5 + q] / V, dR = e[14 + q] / V, dK = e[23 + q] / V;
      R->PT[q] = kin + ss + dR + dK;
    }
    int reb = o_check_list(S, R->snap, R->r_list_min, R->r_list_max);
    if (reb) list_build(S, R);
    if (P->cl_fr > 0 && fmod((double)step, P->cl_fr) == 0) o_clust_conf(S, P->min_s, &R->ran);
    if (obs) {
      double Mom[6] = {0, 0, 0, 0, 0, 0};
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < 3; ++j) { Mom[j] += S->p[3 * i + j]; Mom[j + 3] += S->l[3 * i + j]; }
      double *o = obs + (size_t)(step - step0) * OBS_N;
      o[OBS_K] = K; o[OBS_U] = U; o[OBS_P] = Pr; o[OBS_USS] = Uss; o[OBS_PSS] = Pss;
      o[OBS_UDR] = UdipR; o[OBS_PDR] = PdipR; o[OBS_UDK] = UdipK; o[OBS_PDK] = PdipK;
      o[OBS_MOM2] = Mom[0] * Mom[0] + Mom[1] * Mom[1] + Mom[2] * Mom[2];
      o[OBS_ANG2] = Mom[3] * Mom[3] + Mom[4] * Mom[4] + Mom[5] * Mom[5];
      o[OBS_REBUILD] = reb;
    }
  }
  return 0;
}

/* threads the timing build's parallel loops use (1 in the serial build) */
int o_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void o_run_free(orun *R) {
  free(R->Tab); free(R->Kvec); free(R->off); free(R->nbr); free(R->snap);
  memset(R, 0, sizeof *R);
}
