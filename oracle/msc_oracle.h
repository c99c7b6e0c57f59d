/*
 * msc_oracle.h -- CPU FP64 restatement of mdSpherocyl's per-step force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU library is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path (mdspherocyl_amd) never does.
 *
 * Every routine restates one reference routine (file:line relative to the
 * reference tree) in the same floating-point operation order, so that a serial
 * build of the reference and this file produce the same doubles.  Build with
 * -ffp-contract=off (see oracle/Makefile).
 *
 * Data layout: structure of arrays, particles in the reference's species-major
 * global-id order (v25/functions.cpp:1158-1179).  Every particle has `nd`
 * dipole slots; a slot with moment 0 is skipped exactly as the reference skips
 * zero dipoles (v25/functions.cpp:1589, 1797).
 */
#ifndef MSC_ORACLE_H
#define MSC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OSYS_MAXSPEC 8

typedef struct osys {
  int N;          /* particles */
  int nd;         /* dipole slots per particle */
  int nspec;      /* species (contiguous blocks of the global id) */
  int counts[OSYS_MAXSPEC];
  double L[3];    /* box */
  double *r;      /* [N*3] centre position (never wrapped, v25/classes.cpp:228) */
  double *p;      /* [N*3] momentum */
  double *l;      /* [N*3] angular momentum (lab frame) */
  double *ori;    /* [N*9] orientation matrix, row-major; row 2 = axis u */
  double *F;      /* [N*3] force */
  double *T;      /* [N*3] torque */
  double *len;    /* [N] cylinder length L */
  double *dia;    /* [N] diameter D */
  double *mass;   /* [N] */
  double *inert;  /* [N*3] principal moments */
  double *dpc;    /* [N*nd] signed dipole offset along the axis (m_pos_constr) */
  double *doc;    /* [N*nd*3] dipole orientation in body axes (m_ori_constr) */
  double *dmu;    /* [N*nd] dipole moment */
  double *dr;     /* [N*nd*3] dipole lab position (stored state, as Dipole::m_r) */
  double *du;     /* [N*nd*3] dipole lab orientation (stored state, as Dipole::m_o) */
  int *clust;     /* [N] cluster label */
} osys;

/* ---- MT19937 (gsl_rng_mt19937 semantics) ---- */
typedef struct omt { uint32_t mt[624]; int mti; } omt;
void   omt_seed(omt *g, unsigned long seed);
uint32_t omt_u32(omt *g);
double omt_uniform(omt *g);
void   o_gauss(double variance, double *rnd1, double *rnd2);

/* ---- geometry ---- */
double o_segdist(const double rA[3], const double uA[3], double LA,
                 const double rB[3], const double uB[3], double LB,
                 const double Lbox[3], const double rAB[3], double dAB[3],
                 double st[2]);
double o_spherocyl_dist(const osys *s, int A, int B, const double rAB[3],
                        double dAB[3], double st[2]);
void   o_minimg(const double ri[3], const double rj[3], const double L[3],
                double rij[3]);

/* ---- neighbour list ---- */
long   o_neigh_list(const osys *s, double r_list, int *off, int *nbr, long cap);
void   o_savec_list(const osys *s, double *snap);
int    o_check_list(const osys *s, const double *snap, double r_list_min,
                    double r_list_max);
double o_round12(double x);

/* ---- forces ---- */
int    o_force_sc(osys *s, double r_cut, const int *off, const int *nbr,
                  double *U, double *W, double PT[9]);
void   o_tabul(double *Tab, double alpha);
void   o_force_dipR(osys *s, const double *Tab, double rc, double *U, double PT[9]);
void   o_wave_v(double *Kvec, const double L[3], double alpha, int K);
void   o_force_dipK(osys *s, int K, double alpha, const double *Kvec,
                    double *U, double PT[9]);
void   o_ext_field(osys *s, const double B[3], double *U);
double o_self_ener(const osys *s);

/* ---- integrator ---- */
typedef struct otherm { double Q, s, g, strn; omt ran; } otherm;
void   o_set_inertia(osys *s, int i);
void   o_set_dip(osys *s, int i);
void   o_translate(osys *s, int i, double dt);
void   o_rotate(osys *s, int i, double dt);
void   o_renorm(osys *s, int i);
void   o_moveX(osys *s, double dt);
void   o_moveP(osys *s, const otherm th[2], double dt, double K[2], double PT[9],
               int last);
void   o_moveTher(otherm th[2], double temp, double dt, const double K[2],
                  double Tpart, int mode);

/* ---- initial state (v25/functions.cpp:491-1146, 3328-3409) ---- */
int    o_ini_pos(osys *s);
void   o_gcamix(osys *s, omt *ran);
int    o_ini_mom(osys *s, double temp, omt *ran);

/* ---- clusters (v25/functions.cpp:2950-3324) ---- */
int    o_dist_test(const osys *s, int A, int B, double min_s);
int    o_clust_conf(osys *s, double min_s, omt *ran);
/* with the opt-in dipolar energy clause: contacts also need o_udd < u_cut */
int    o_clust_conf_dd(osys *s, double min_s, double u_cut, omt *ran);
double o_udd(const osys *s, int A, int B);
/* out[k] = contact of pair (A[k], B[k]) in either order; returns the count */
long   o_contact_pairs(const osys *s, long n, const int *A, const int *B, double min_s,
                       double u_cut, unsigned char *out);

/* ---- order parameters (v25/functions.cpp:2702-2760, v25/jacobi.c) ---- */
int    o_jacobi3(float a[3][3], float d[3], float v[3][3]);
int    o_oo_params(const osys *s, double *G1, double *G2, double M[3]);

/* ---- pair correlations (v25/functions.cpp:2399-2466) ---- */
long   o_acc_correl(const osys *s, double bar_w, int n_bars, double *hist);

/* ---- whole run, restating v25/main.cpp:401-818 ---- */
typedef struct orun_params {
  double dens, temp, dt, sr_cut, alpha, min_s;
  double Q, gamma, s0;      /* thermostat mass, "shear rate", initial coordinate */
  double B[3];
  int K;                    /* reference "No. of Kvectors" input */
  int cl_fr;                /* clustering period (0: never) */
  double rc_real;           /* <=0: all min-image pairs (reference); >0 cutoff */
  int dipK_on, dipR_on;
} orun_params;

/* Per-step observables written to obs[step*OBS_N + k], OBS_* below. */
enum { OBS_K = 0, OBS_U, OBS_P, OBS_USS, OBS_PSS, OBS_UDR, OBS_PDR, OBS_UDK,
       OBS_PDK, OBS_MOM2, OBS_ANG2, OBS_REBUILD, OBS_N };

typedef struct orun {
  double Tab_alpha;
  double *Tab;              /* 32768*3 */
  double *Kvec;             /* (K+1)^3 */
  int *off, *nbr;           /* neighbour list CSR */
  long cap;
  double *snap;             /* N*3 list reference positions */
  double r_list_min, r_list_max, Uself;
  double dmax, dmin;
  otherm th[2];
  omt ran;
  double PT[9];             /* last step's total pressure tensor
                               (PTkin + PTss + PTdipR + PTdipK) / V,
                               v25/v251/main.cpp:635-642 */
} orun;

int  o_run_init(osys *s, const orun_params *P, orun *R, int seed_init);
int  o_run_steps(osys *s, const orun_params *P, orun *R, int step0, int nsteps,
                 double *obs);
void o_run_free(orun *R);
/* threads of the OpenMP timing build's loops (1 in the serial build) */
int  o_threads(void);

/* One force evaluation in the reference's step order (v25/main.cpp:582-614).
 * e[0..]: Uss Wss UdipR UdipK Uext, then PTss[9] PTdipR[9] PTdipK[9]. */
int  o_forces(osys *s, const orun_params *P, orun *R, double *e);

#ifdef __cplusplus
}
#endif
#endif
