// Host side of libmsc_hip.so: context, replica layout, buffers and the C ABI
// declared in include/msc.h.  The step driver restates the order of the
// reference step loop (v25/main.cpp:559-708) over device kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/msc.h"
#include "msc_geom.hpp"
#include "msc_kernels.hpp"

using namespace msc;

namespace {

// MT19937 (published algorithm, GSL seeding) for the thermostat streams and
// the cluster seed draws.
struct MT {
  uint32_t mt[624];
  int mti = 625;
  void seed(unsigned long s) {
    if (s == 0) s = 4357;
    mt[0] = (uint32_t)(s & 0xffffffffUL);
    for (int i = 1; i < 624; ++i) mt[i] = (uint32_t)(1812433253UL * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i);
    mti = 624;
  }
  uint32_t next() {
    const uint32_t UP = 0x80000000u, LO = 0x7fffffffu, mag[2] = {0u, 0x9908b0dfu};
    if (mti >= 624) {
      int k;
      for (k = 0; k < 624 - 397; ++k) {
        uint32_t y = (mt[k] & UP) | (mt[k + 1] & LO);
        mt[k] = mt[k + 397] ^ (y >> 1) ^ mag[y & 1u];
      }
      for (; k < 623; ++k) {
        uint32_t y = (mt[k] & UP) | (mt[k + 1] & LO);
        mt[k] = mt[k + (397 - 624)] ^ (y >> 1) ^ mag[y & 1u];
      }
      uint32_t y = (mt[623] & UP) | (mt[0] & LO);
      mt[623] = mt[396] ^ (y >> 1) ^ mag[y & 1u];
      mti = 0;
    }
    uint32_t y = mt[mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
  }
  double uniform() { return next() / 4294967296.0; }
};

// Box-Muller pair as gauss(1, r1, r2) (v25/functions.cpp:1149-1156)
void gauss1(double &r1, double &r2) {
  double a = std::sqrt(-2 * std::log(r1)) * std::cos(2 * M_PI * r2);
  double b = std::sqrt(-2 * std::log(r1)) * std::sin(2 * M_PI * r2);
  double sv = std::sqrt(1.0);
  r1 = a * sv;
  r2 = b * sv;
}

struct TimerSet {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double ms = 0;
  long launches = 0;
};

}  // namespace

// Default skin of a MSC_MIXED context's own rows (length units of the
// state; the autotuner tries others).
static constexpr double kDefaultListSkin = 1.0;

struct msc_ctx {
  int device = 0, precision = MSC_FP64;
  bool mixed = false;
  bool contact = false;  // force_sc walks the contact rows (else every reference-row pair)
  hipStream_t stream = nullptr;
  int nrep = 0, Ntot = 0, nd = 1;
  int rchunk = 128;           // particles per structure-factor chunk task
  std::vector<msc_replica_desc> desc;
  std::vector<RepDev> reps;
  std::vector<Task> tasks, rtasks;
  std::vector<MT> thermo_rng;  // 2 per replica
  std::vector<int> kz_groups;  // distinct KZ values
  std::string err;
  State S{};
  bool have_state = false, have_lists = false, graph_on = false, timing = false;
  // launch choices (msc_tuning): the fused mixed pair launch and its waves per task
  bool fused_pairs = true;
  int pair_waves = 4;
  // msc_tuning.overlap: the fused pair launch on stream2, forked after the
  // drift (ev_fork) and joined before the back projection (ev_join)
  bool overlap = true;
  // msc_tuning.fuse_kick: the last kick in the mixed back projection's epilogue
  bool fuse_kick = true;
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // MSC_MIXED: the skin of the context's own rows (0: the reference's rows
  // and rebuild test, contact rows refreshed from them); per replica the
  // smallest diameter and largest length + diameter the radii derive from
  double list_skin = 0.0;
  std::vector<double> sh_dmin, sh_dmax;
  long nl_alloc = 0, sl_alloc = 0, dl_alloc = 0;  // entries allocated for the row arrays
  size_t part_elems = 0;      // structure-factor partials the buffers hold (at the present chunk)
  size_t part_cap = 0;        // and at most (allocated)
  size_t rtask_cap = 0;       // chunk tasks d_rtasks holds
  int nslot_tot = 0, maxslot = 0, maxcol = 0;
  // device buffers
  RepDev *d_reps = nullptr;
  Task *d_tasks = nullptr, *d_rtasks = nullptr;
  int *d_rep_of = nullptr, *d_cell_rep = nullptr;
  double *d_x = nullptr, *d_p = nullptr, *d_l = nullptr, *d_ori = nullptr, *d_F = nullptr, *d_T = nullptr;
  double *d_len = nullptr, *d_dia = nullptr, *d_mass = nullptr, *d_inert = nullptr;
  double *d_dpc = nullptr, *d_doc = nullptr, *d_dmu = nullptr, *d_dpos = nullptr, *d_dori = nullptr, *d_snap = nullptr,
         *d_snap_e = nullptr, *d_snap_u = nullptr;
  int *d_clu = nullptr, *d_clu_flag = nullptr;
  double *d_corr = nullptr;  // pair-correlation histograms (grown on demand)
  size_t corr_cap = 0;
  unsigned long long *d_corr_drop = nullptr;
  double *d_order = nullptr;  // [kOrderSums] order-parameter sums
  double *d_rows = nullptr, *d_kblk = nullptr;  // per-task row sets and k-space block sums of the step loop
  double *d_endsum = nullptr;  // end-of-step sums per replica (kEndSums each)
  double *d_kick_ts = nullptr;  // thermostat coordinates of the fused last kick, 2 per replica
  int *d_step = nullptr;  // {step within the run, first absolute step of the run}
  int *d_cell_of = nullptr, *d_cell_cnt = nullptr, *d_cell_start = nullptr, *d_cell_fill = nullptr,
      *d_cell_list = nullptr;
  int *d_cell_tmp = nullptr;
  double *d_cell_pack = nullptr;
  double *d_rec_rod = nullptr, *d_rec_dip = nullptr;  // partner records (mixed contexts)
  int *d_nl_cnt = nullptr, *d_nl_idx = nullptr;
  int *d_sl_cnt = nullptr, *d_sl_idx = nullptr;
  int *d_dl_cnt = nullptr, *d_dl_idx = nullptr;
  int ncells = 0;
  double *d_kval = nullptr;
  void *d_tab = nullptr, *d_ph = nullptr, *d_amp = nullptr, *d_part = nullptr, *d_A = nullptr;
  void *d_bop = nullptr;  // matrix-core back-projection operands (mixed contexts)
  // several dipole slots per particle (nd > 1): FP64 slot phase tables,
  // chunk partials and amplitudes of the reciprocal sum (launch_recip_nd)
  void *d_ph_nd = nullptr, *d_A_nd = nullptr;
  double *d_part_nd = nullptr;
  long bop_elems = 0;     // their count, in groups of 4 fp16 values
  int *d_bexp = nullptr;  // their per-column scale exponents
  std::vector<double> kval_host;  // reciprocal weights of every slot (build_kspace)
  double *d_slot_out = nullptr, *d_task_out = nullptr, *d_red = nullptr, *d_M2 = nullptr, *d_Kin = nullptr,
         *d_gs = nullptr, *d_obs = nullptr, *d_counts = nullptr;
  RepEnergy *d_E = nullptr;
  Thermo *d_th = nullptr;
  int *d_flags = nullptr, *d_rebuild = nullptr, *d_srebuild = nullptr, *d_one = nullptr;
  int *d_lctr = nullptr;  // [nrep] arrival counters of the step loop's cell scan (zero between launches)
  unsigned long long *d_dmax = nullptr;  // [2][nrep][kDmaxBuckets][2] displacement maxima of the step loop
  std::map<std::string, TimerSet> timers;
  std::vector<hipEvent_t> event_pool;
  double *d_run_gs = nullptr, *d_run_obs = nullptr;
  char *d_save = nullptr;  // chunk rollback copy of the step state
  char *d_tune_save = nullptr;  // msc_autotune's copy of the step state
  size_t tune_save_cap = 0;
  long chunk_repeats = 0;  // chunks of msc_run repeated after a list overflow
  size_t save_cap = 0;
  void *pin = nullptr;  // pinned staging buffer of kPinBytes
  size_t run_gs_cap = 0, run_obs_cap = 0;
  // the step captured as a graph; dropped when anything it holds by value
  // (row regions, run buffers, replica parameters) changes
  hipGraphExec_t graph_exec = nullptr;
  LaunchRecord launch;      // failed launches of this context's entry points
  double *d_stamp = nullptr;  // completion stamp (finish)
  double seq = 0;             // last stamp value enqueued
};

// Makes the context's launch record current on the calling thread for the
// duration of an entry point (launch failures are kept per context).
struct LaunchScope {
  LaunchRecord *prev;
  explicit LaunchScope(msc_ctx *c) : prev(set_launch_record(c ? &c->launch : nullptr)) {}
  ~LaunchScope() { set_launch_record(prev); }
};
#define SCOPE(ctx) LaunchScope launch_scope_(ctx)

static std::string g_err;  // failures that have no context to hold them
static constexpr size_t kPinBytes = (size_t)4 << 20;

// One stream per device, created on first use and kept until the process
// ends: contexts on a device share it, so creating and destroying contexts
// never creates or destroys streams.
static hipError_t device_stream(int device, hipStream_t *out) {
  static std::mutex mu;
  static std::map<int, hipStream_t> streams;
  std::lock_guard<std::mutex> lock(mu);
  auto it = streams.find(device);
  if (it != streams.end()) {
    *out = it->second;
    return hipSuccess;
  }
  hipStream_t st = nullptr;
  hipError_t e = hipStreamCreate(&st);
  if (e == hipSuccess) streams[device] = st;
  *out = st;
  return e;
}

// The second stream of a device (the pair launch of an overlapped step; see
// msc_tuning.overlap), shared by the contexts on the device, default priority.
static hipError_t device_stream2(int device, hipStream_t *out) {
  static std::mutex mu;
  static std::map<int, hipStream_t> streams;
  std::lock_guard<std::mutex> lock(mu);
  auto it = streams.find(device);
  if (it != streams.end()) {
    *out = it->second;
    return hipSuccess;
  }
  hipStream_t st = nullptr;
  hipError_t e = hipStreamCreate(&st);
  if (e == hipSuccess) streams[device] = st;
  *out = st;
  return e;
}

#define HIPC(expr)                                                                   \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                  \
      g_err = ctx->err;                                                              \
      return MSC_E_HIP;                                                              \
    }                                                                                \
  } while (0)

// A failed launch is recorded by the launch wrapper (take_launch_error);
// every result read back from the device (d2h) and the entry points below
// check that record first.
static int launch_status(msc_ctx *ctx, const char *what) {
  const char *kernel = "";
  hipError_t e = take_launch_error(&kernel);
  if (e == hipSuccess) {
    e = hipGetLastError();
    kernel = "(runtime)";
  }
  if (e == hipSuccess) return MSC_OK;
  ctx->err = std::string("kernel launch ") + kernel + " in " + what + ": " + hipGetErrorString(e);
  g_err = ctx->err;
  return MSC_E_HIP;
}
#define LAUNCHED(what)                                              \
  do {                                                              \
    if (int rc_l_ = launch_status(ctx, (what))) return rc_l_;       \
  } while (0)

// Host <-> device transfers.  The runtime is never handed pageable host
// memory: every transfer goes through the context's pinned staging buffer
// (allocated with hipHostMalloc), chunk by chunk, ordered on the context
// stream and complete when the call returns.  Pageable transfers of a
// megabyte or more would otherwise make the runtime lock the caller's pages
// in place for the DMA engine.
static int d2h(msc_ctx *ctx, void *dst, const void *src, size_t bytes) {
  LAUNCHED("a device read-back");
  char *d = (char *)dst;
  const char *s_ = (const char *)src;
  for (size_t off = 0; off < bytes; off += kPinBytes) {
    size_t n = std::min(kPinBytes, bytes - off);
    HIPC(hipMemcpyAsync(ctx->pin, s_ + off, n, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    memcpy(d + off, ctx->pin, n);
  }
  return MSC_OK;
}
static int h2d(msc_ctx *ctx, void *dst, const void *src, size_t bytes) {
  char *d = (char *)dst;
  const char *s_ = (const char *)src;
  for (size_t off = 0; off < bytes; off += kPinBytes) {
    size_t n = std::min(kPinBytes, bytes - off);
    HIPC(hipStreamSynchronize(ctx->stream));  // the staging buffer is free again
    memcpy(ctx->pin, s_ + off, n);
    HIPC(hipMemcpyAsync(d + off, ctx->pin, n, hipMemcpyHostToDevice, ctx->stream));
  }
  HIPC(hipStreamSynchronize(ctx->stream));
  return MSC_OK;
}

// Proof that the kernels an entry point enqueued have executed: a one-lane
// kernel enqueued after them writes a fresh sequence number, which is read
// back and compared.  A launch the runtime accepted but never ran (results
// would stay at their previous or zero values) makes the call fail.
static int finish(msc_ctx *ctx, const char *what) {
  LAUNCHED(what);
  const double want = ++ctx->seq;
  launch_stamp(ctx->d_stamp, want, ctx->stream);
  LAUNCHED(what);
  double got = -1;
  if (int rc_ = d2h(ctx, &got, ctx->d_stamp, sizeof got)) return rc_;
  if (got != want) {
    ctx->err = std::string(what) + ": the device did not execute the library's kernels (completion stamp " +
               std::to_string(got) + ", expected " + std::to_string(want) + ")";
    g_err = ctx->err;
    return MSC_E_HIP;
  }
  return MSC_OK;
}
#define FINISH(what)                                        \
  do {                                                      \
    if (int rc_f_ = finish(ctx, (what))) return rc_f_;      \
  } while (0)

namespace {

constexpr int kEoff_sc = 0, kEoff_R = 11, kEoff_K = 21, kEoff_ext = 31;
constexpr int kEstride = (int)(sizeof(RepEnergy) / sizeof(double));

int fail(msc_ctx *ctx, int code, const std::string &m) {
  ctx->err = m;
  return code;
}

template <typename T>
int dalloc(msc_ctx *ctx, T **p, size_t n) {
  if (*p) hipFree(*p);
  *p = nullptr;
  if (n == 0) n = 1;
  HIPC(hipMalloc((void **)p, n * sizeof(T)));
  HIPC(hipMemsetAsync(*p, 0, n * sizeof(T), ctx->stream));
  return MSC_OK;
}

// event bracket around a launch when timing is on
struct Bracket {
  msc_ctx *c;
  const char *name;
  hipEvent_t a = nullptr, b = nullptr;
  Bracket(msc_ctx *c_, const char *n) : c(c_), name(n) {
    if (!c->timing) return;
    a = take();
    b = take();
    hipEventRecord(a, c->stream);
  }
  hipEvent_t take() {
    hipEvent_t e = nullptr;
    if (!c->event_pool.empty()) {
      e = c->event_pool.back();
      c->event_pool.pop_back();
    } else {
      hipEventCreate(&e);
    }
    return e;
  }
  ~Bracket() {
    if (!c->timing) return;
    hipEventRecord(b, c->stream);
    c->timers[name].pending.push_back({a, b});
  }
};

void settle_timers(msc_ctx *ctx) {
  hipStreamSynchronize(ctx->stream);
  for (auto &kv : ctx->timers) {
    for (auto &ev : kv.second.pending) {
      float ms = 0;
      hipEventElapsedTime(&ms, ev.first, ev.second);
      kv.second.ms += ms;
      kv.second.launches += 1;
      ctx->event_pool.push_back(ev.first);
      ctx->event_pool.push_back(ev.second);
    }
    kv.second.pending.clear();
  }
}

State sel(const msc_ctx *ctx, int rep) {
  State s = ctx->S;
  if (rep == MSC_ALL) {
    s.task_begin = 0;
    s.ntask = (int)ctx->tasks.size();
    s.g_begin = 0;
    s.g_end = ctx->Ntot;
    s.rep_begin = 0;
    s.rep_end = ctx->nrep;
  } else {
    const RepDev &R = ctx->reps[rep];
    s.task_begin = R.task0;
    s.ntask = R.ntask;
    s.g_begin = R.n0;
    s.g_end = R.n0 + R.N;
    s.rep_begin = rep;
    s.rep_end = rep + 1;
  }
  return s;
}

bool rep_ok(msc_ctx *ctx, int rep) { return rep == MSC_ALL || (rep >= 0 && rep < ctx->nrep); }

// Contact-row radius never beyond the rows it is taken from.
void clamp_short(RepDev &R) { R.r_short = std::min(R.r_wca + R.skin_short, R.r_list_max); }

// List radii of a replica.  The reference's (v25/main.cpp:461-463):
// r_list_min = max(sr_cut + 0.1 D_min, 1.5 (L + D)_max), r_list_max =
// r_list_min + 0.9 D_min, rebuilt when 2 sqrt(3) max|dx| exceeds their
// difference; contact rows within the WCA range 2^(1/6) D_max plus half the
// smallest diameter; real-space rows within r_c plus the list skin.  With a
// list skin (MSC_MIXED) the rows are the context's own: every pair within
// r_wca + skin, the real-space rows within r_c + skin, all rebuilt together
// when twice the largest segment-point displacement exceeds the skin.
void set_radii(msc_ctx *ctx, int rp) {
  RepDev &R = ctx->reps[rp];
  const msc_replica_desc &D = ctx->desc[rp];
  const double dmin = ctx->sh_dmin[rp], dmax = ctx->sh_dmax[rp];
  R.r_wca = std::pow(2.0, 1.0 / 6.0) * R.diamax;
  if (ctx->list_skin > 0) {
    const double skin = ctx->list_skin;
    R.own_lists = 1;
    R.r_list_min = R.r_wca;
    R.r_list_max = R.r_wca + skin;
    R.skin_short = skin;
    R.rl_dip = D.rc > 0 ? D.rc + skin : 0;
  } else {
    R.own_lists = 0;
    R.r_list_min = D.sr_cut + 0.1 * dmin;
    if (R.r_list_min < 1.5 * dmax) R.r_list_min = 1.5 * dmax;
    R.r_list_max = R.r_list_min + 0.9 * dmin;
    R.rl_dip = D.rc > 0 ? D.rc + (R.r_list_max - R.r_list_min) : 0;
    R.skin_short = 0.5 * dmin;
  }
  clamp_short(R);
}

// derived per-replica quantities that depend on the state (shapes, dipoles)
void derive(msc_ctx *ctx, int rp, const double *len, const double *dia, const double *dpc) {
  RepDev &R = ctx->reps[rp];
  const msc_replica_desc &D = ctx->desc[rp];
  double dmax = 0, dmin = 1.e6, lmax = 0, lev = 0, diamax = 0;
  for (int i = 0; i < R.N; ++i) {
    dmax = std::max(dmax, len[i] + dia[i]);
    dmin = std::min(dmin, dia[i]);
    diamax = std::max(diamax, dia[i]);
    lmax = std::max(lmax, len[i]);
    for (int w = 0; w < R.nd; ++w) lev = std::max(lev, std::fabs(dpc[i * R.nd + w]));
  }
  if (D.dmax > 0) dmax = D.dmax;
  if (D.dmin > 0) dmin = D.dmin;
  R.lmax = lmax;
  R.diamax = diamax;
  ctx->sh_dmin[rp] = dmin;
  ctx->sh_dmax[rp] = dmax;
  ctx->S.lever_max = std::max(ctx->S.lever_max, lev);
  set_radii(ctx, rp);
}

// cell grid of a replica for the current list radii
void grid(msc_ctx *ctx, int rp) {
  RepDev &R = ctx->reps[rp];
  double need = R.r_list_max + R.lmax;
  if (R.rc > 0) need = std::max(need, R.rl_dip + 2 * ctx->S.lever_max);
  need *= 1.0 + 1e-9;
  // cells of half the search radius with a 5-cell stencil where 5 fit (a
  // reach of at most 2 cells, which the list kernel relies on; a
  // search volume (5/2)^3 need^3 instead of 3^3 need^3), else cells of the
  // whole radius with a 3-cell stencil, else the whole box
  for (int a = 0; a < 3; ++a) {
    int n = (int)std::floor(R.L[a] / (0.5 * need)), reach = 2;
    if (n < 5) {
      n = (int)std::floor(R.L[a] / need);
      reach = 1;
    }
    if (n < 3) {
      n = 1;
      reach = 0;
    }
    n = std::min(n, 256);
    R.nc[a] = n;
    R.reach[a] = reach;
    R.cw[a] = R.L[a] / n;
  }
  R.ncell = R.nc[0] * R.nc[1] * R.nc[2];
}

// the captured step holds row regions, run buffers and launch choices
static void drop_graph(msc_ctx *ctx) {
  if (ctx->graph_exec) hipGraphExecDestroy(ctx->graph_exec);
  ctx->graph_exec = nullptr;
}

int upload_reps(msc_ctx *ctx) {
  // replica parameters decide which launches a step makes (field, modes)
  drop_graph(ctx);
  // the thermostat's constant factors, with the host's libm as the
  // reference evaluates them once per call (v25/functions.cpp:2196-2197)
  for (RepDev &R : ctx->reps) {
    R.th_ex1 = std::exp(-R.gamma * R.dt / 2);
    R.th_sq = std::sqrt(2. * R.temp * R.gamma * R.dt / R.Q);
  }
  return h2d(ctx, ctx->d_reps, ctx->reps.data(), sizeof(RepDev) * ctx->nrep);
}

void refresh_state_ptrs(msc_ctx *ctx) {
  State &s = ctx->S;
  s.Ntot = ctx->Ntot;
  s.nd = ctx->nd;
  s.nrep = ctx->nrep;
  s.ntask = (int)ctx->tasks.size();
  s.ntask_all = (int)ctx->tasks.size();
  s.nrtask = (int)ctx->rtasks.size();
  s.ncells = ctx->ncells;
  s.reps = ctx->d_reps;
  s.tasks = ctx->d_tasks;
  s.rtasks = ctx->d_rtasks;
  s.rep_of = ctx->d_rep_of;
  s.x = ctx->d_x; s.p = ctx->d_p; s.l = ctx->d_l; s.ori = ctx->d_ori; s.F = ctx->d_F; s.T = ctx->d_T;
  s.len = ctx->d_len; s.dia = ctx->d_dia; s.mass = ctx->d_mass; s.inert = ctx->d_inert;
  s.dpc = ctx->d_dpc; s.doc = ctx->d_doc; s.dmu = ctx->d_dmu; s.dpos = ctx->d_dpos; s.dori = ctx->d_dori;
  s.snap = ctx->d_snap;
  s.snap_e = ctx->d_snap_e;
  s.snap_u = ctx->d_snap_u;
  s.cell_of = ctx->d_cell_of; s.cell_cnt = ctx->d_cell_cnt; s.cell_start = ctx->d_cell_start;
  s.cell_fill = ctx->d_cell_fill; s.cell_tmp = ctx->d_cell_tmp; s.cell_pack = ctx->d_cell_pack; s.cell_list = ctx->d_cell_list;
  s.cell_rep = ctx->d_cell_rep;
  s.nl_cnt = ctx->d_nl_cnt; s.nl_idx = ctx->d_nl_idx;
  s.sl_cnt = ctx->d_sl_cnt; s.sl_idx = ctx->d_sl_idx;
  bool anyrc = false;
  for (auto &R : ctx->reps) anyrc |= R.rc > 0;
  (void)anyrc;
  s.dl_cnt = ctx->d_dl_cnt;
  s.dl_idx = ctx->d_dl_idx;
  s.kval = ctx->d_kval;
  s.rec_rod = ctx->d_rec_rod;
  s.rec_dip = ctx->d_rec_dip;
}

// reciprocal weights val = 2 KK / V of wave_v/force_dipK (v25/functions.cpp:1733-1737,
// 1963-1990) for the half-space slots; G = sum val k k^T
void build_kspace(msc_ctx *ctx, std::vector<double> &kval) {
  kval.assign(ctx->nslot_tot, 0.0);
  for (int rp = 0; rp < ctx->nrep; ++rp) {
    RepDev &R = ctx->reps[rp];
    const int kmax = R.kmax, side = 2 * kmax + 1;
    double c = 4. * R.alpha * R.alpha;
    for (int q = 0; q < 9; ++q) R.G[q] = 0;
    for (int col = 0; col < R.ncol; ++col) {
      int nx = col % side - kmax, ny = col / side - kmax;
      for (int nz = 0; nz < R.KZ; ++nz) {
        bool half = nz > 0 || ny > 0 || (ny == 0 && nx > 0);
        if (!half) continue;
        int ax = std::abs(nx), ay = std::abs(ny);
        double Kx = R.P2[0] * ax, Ky = R.P2[1] * ay, Kz = R.P2[2] * nz;
        double Ksq = Kx * Kx + Ky * Ky + Kz * Kz;
        double KK = 4. * M_PI * std::exp(-Ksq / c) / Ksq;
        double val = 2. * KK / R.V;
        kval[R.slot0 + col * R.KZ + nz] = val;
        double k[3] = {R.P2[0] * nx, R.P2[1] * ny, R.P2[2] * nz};
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) R.G[3 * a + b] += val * k[a] * k[b];
      }
    }
  }
}

// tabul (v25/functions.cpp:1680-1704), host libm, one table per replica
void build_table(const RepDev &R, std::vector<double> &T) {
  double extL = 200 * 1.001, alpha = R.alpha;
  double alsq = alpha * alpha, alft = alsq * alsq, spi = std::sqrt(M_PI);
  for (int i = 1; i < kTabN + 1; i++) {
    double r = 0.5 * extL * i / kTabN;
    double rsq = r * r, sr1 = 1.0 / r, sr2 = sr1 * sr1, sr3 = sr2 * sr1;
    double ERFC = std::erfc(alpha * r);
    double aer = 2.0 * alpha * std::exp(-alsq * rsq) / spi;
    double aa = 2.0 * alsq + 3.0 * sr2;
    T[(size_t)R.tab0 * 3 + 3 * (i - 1) + 0] = ERFC * sr3 + aer * sr2;
    T[(size_t)R.tab0 * 3 + 3 * (i - 1) + 1] = (3.0 * ERFC * sr3 + aer * aa) * sr2;
    T[(size_t)R.tab0 * 3 + 3 * (i - 1) + 2] = 4. * alft * aer;
  }
}

// Scale exponents of the matrix-core back-projection operands of replica
// rp: the operands of column tile ct (16 columns (nx, ny)) and weight l
// (nz^l A) are multiplied by 2^e before their fp16 split, with e chosen so
// that the hard bound
//   |nz^l A(k)| <= nz^l val(k) |k| sum_i |mu_i|  (|S(k)| <= |k| sum |c_i|)
// maps to at most 2^15 over the tile: nothing can overflow fp16, and the
// entries that carry the sums stay in fp16's normal range, where the split
// keeps 22 bits.  The kernel multiplies its column sums by 2^-e.
void build_bop_exp(const msc_ctx *ctx, int rp, double musum, std::vector<int> &out) {
  const RepDev &R = ctx->reps[rp];
  const int kmax = R.kmax, side = 2 * kmax + 1, nct = (R.ncol + 15) / 16;
  out.assign((size_t)3 * nct, 0);
  for (int ct = 0; ct < nct; ++ct) {
    double bmax[3] = {0, 0, 0};
    for (int col = 16 * ct; col < std::min(R.ncol, 16 * ct + 16); ++col) {
      const int nx = col % side - kmax, ny = col / side - kmax;
      for (int nz = 0; nz < R.KZ; ++nz) {
        const double val = ctx->kval_host[R.slot0 + col * R.KZ + nz];
        if (val == 0.0) continue;
        const double k[3] = {R.P2[0] * nx, R.P2[1] * ny, R.P2[2] * nz};
        const double b = val * std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]) * musum;
        double w = 1.0;
        for (int l = 0; l < 3; ++l, w *= nz) bmax[l] = std::max(bmax[l], w * b);
      }
    }
    for (int l = 0; l < 3; ++l) {
      int e = 0;
      if (bmax[l] > 0) e = (int)std::floor(std::log2(32768.0 / bmax[l]));
      out[(size_t)3 * ct + l] = std::min(100, std::max(-100, e));
    }
  }
}

}  // namespace

// ==========================================================================
// C ABI
// ==========================================================================
// MT19937 of Matsumoto and Nishimura (ACM TOMACS 8, 1998) with GSL's state
// convention: words mt[] and the position mti of the next word (624: the
// block is used up); gsl_rng_uniform is the tempered word / 2^32.
// principal moments of inertia of a spherocylinder, Spherocyl::set_inertia
// (v25/classes.cpp:175-185), the reference's formula as written
static void inertia(double len, double dia, double mass, double I[3]) {
  const double Rr = dia / 2, Rsq = Rr * Rr, Rft = Rsq * Rsq, Lq = len, Lsq = Lq * Lq;
  const double vol = M_PI * Rsq * (Lq + 3. * Rr / 4);
  const double mdens = mass / vol;
  I[0] = mdens * M_PI * (Lsq * Lq * Rsq / 12 + Lsq * Rsq * Rr / 3 + Lq * Rft / 4 + 8. * Rft * Rr / 15);
  I[1] = I[0];
  I[2] = mdens * M_PI * (Lq * Rft / 2 + 8. * Rft * Rr / 15);
}

struct Mt19937 {
  static constexpr int kN = 624, kM = 397;
  uint32_t mt[kN];
  int mti = kN;
  uint32_t next() {
    if (mti >= kN) {
      for (int k = 0; k < kN; ++k) {
        const uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % kN] & 0x7fffffffu);
        mt[k] = mt[(k + kM) % kN] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
      }
      mti = 0;
    }
    uint32_t y = mt[mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
  double uniform() { return next() / 4294967296.0; }
};

// --------------------------------------------------------------------------
// Signal=0 initial state (v25/main.cpp:215-260): lattice, mixing moves,
// momenta.  Host code: every step depends on the generator's previous draw.
// --------------------------------------------------------------------------
namespace init {

struct Parts {
  int N, nspec;
  const int *counts;
  const double *L, *len, *dia;
  double *r, *ori;
  int start(int n) const {
    int b = 0;
    for (int m = 0; m < n; ++m) b += counts[m];
    return b;
  }
  double dist2(int A, const double rA[3], int B) const {
    double rAB[3], dAB[3], st[2];
    for (int a = 0; a < 3; ++a) rAB[a] = wrapd(rA[a] - r[3 * B + a], L[a]);
    return segdist(rA, ori + 9 * A + 6, len[A], r + 3 * B, ori + 9 * B + 6, len[B], L, rAB, dAB, st);
  }
};

// ini_pos (v25/functions.cpp:491-676): species by species, a stacked
// triangular lattice of layers one diameter-and-length apart, every axis
// along z, sites of later species skipped where they touch an earlier one;
// then the reference's overlap test of the whole configuration.
static int ini_pos(Parts &P) {
  const double sq = std::sqrt(3.) / 2., sqt = 1. / (2 * std::sqrt(3.)), sqtt = std::sqrt(6.) / 3;
  const double *L = P.L;
  for (int n = 0; n < P.nspec; ++n) {
    const int b0 = P.start(n), NN = P.counts[n];
    const double dia = P.dia[b0], len = P.len[b0];
    const double m1 = std::floor(L[0] / dia), m2 = std::floor(L[1] / (dia * sq)),
                 m3 = std::floor(L[2] / (dia * sqtt + len));
    const int num1 = (int)(m1 < 1. ? 1. : m1), num2 = (int)(m2 < 1. ? 1. : m2), num3 = (int)(m3 < 1. ? 1. : m3);
    int cont = 0;
    for (int pass = 0; cont < NN; ++pass) {
      if (pass == 2) return 1;  // the lattice cannot hold the species
      for (int nz = 0; nz < num3; ++nz)
        for (int ny = 0; ny < num2; ++ny)
          for (int nx = 0; nx < num1 && cont < NN; ++nx) {
            double q[3];
            const bool even_z = std::fmod((double)nz, 2.) == 0, even_y = std::fmod((double)ny, 2.) == 0;
            const bool shift = even_z ? !even_y : even_y;
            q[0] = shift ? dia * ((double)nx + 0.5) * 1.0001 : dia * (double)nx * 1.0001;
            q[1] = even_z ? dia * (double)ny * sq * 1.0001 : dia * ((double)ny * sq + sqt) * 1.0001;
            q[2] = (dia * sqtt + len) * (double)nz * 1.0001;
            for (int a = 0; a < 3; ++a) q[a] -= L[a] * std::floor(q[a] / L[a] + 0.5);
            const int i = b0 + cont;
            for (int a = 0; a < 3; ++a) P.r[3 * i + a] = q[a];
            for (int k = 0; k < 9; ++k) P.ori[9 * i + k] = (k % 4 == 0) ? 1.0 : 0.0;
            bool clash = false;
            for (int m = 0; m < n && !clash; ++m) {
              const int c0 = P.start(m);
              const double dd = (dia + P.dia[c0]) / 2, rsq = dd * dd;
              for (int j = c0; j < c0 + P.counts[m] && !clash; ++j) clash = P.dist2(i, q, j) < rsq;
            }
            if (!clash) ++cont;
          }
    }
  }
  // overlap test (v25/functions.cpp:611-673): within and across species
  for (int m = 0; m < P.nspec; ++m)
    for (int n = m; n < P.nspec; ++n) {
      const int a0 = P.start(m), b0 = P.start(n);
      const double dd = (P.dia[a0] + P.dia[b0]) / 2, rsq = dd * dd;
      for (int i = a0; i < a0 + P.counts[m]; ++i)
        for (int j = (m == n ? i + 1 : b0); j < b0 + P.counts[n]; ++j)
          if (P.dist2(i, P.r + 3 * i, j) < rsq) return 1;
    }
  return 0;
}

// GCAmix (v25/functions.cpp:3328-3409): a random pivot; a random seed
// particle and, breadth first, every particle that touches one already moved
// are reflected through the pivot (minimum image, wrapped into the box).
static void gca_move(Parts &P, Mt19937 &mt, std::vector<char> &added, std::vector<int> &queue) {
  const double *L = P.L;
  double pivot[3];
  for (int a = 0; a < 3; ++a) pivot[a] = (2 * mt.uniform() - 1) * L[a] / 2;
  const int seed = (int)std::floor(mt.uniform() * P.N);  // species-major order is global order
  std::fill(added.begin(), added.end(), 0);
  int head = 0, tail = 0;
  added[seed] = 1;
  queue[tail++] = seed;
  while (head < tail) {
    const int q = queue[head++];
    double ri[3];
    for (int a = 0; a < 3; ++a) {
      double d = pivot[a] - P.r[3 * q + a];
      d -= L[a] * std::floor(d / L[a] + 0.5);
      ri[a] = P.r[3 * q + a] + 2 * d;
      ri[a] -= L[a] * std::floor(ri[a] / L[a] + 0.5);
      P.r[3 * q + a] = ri[a];
    }
    int qs = 0;
    for (int n = 0, b = 0; n < P.nspec; b += P.counts[n], ++n)
      if (q >= b) qs = b;
    for (int n = 0, b = 0; n < P.nspec; b += P.counts[n], ++n) {
      const double dd = (P.dia[qs] + P.dia[b]) / 2, rsq = dd * dd;
      for (int i = b; i < b + P.counts[n]; ++i)
        if (!added[i] && P.dist2(q, ri, i) < rsq) {
          added[i] = 1;
          queue[tail++] = i;
        }
    }
  }
}

// Box-Muller pair (gauss, v25/functions.cpp:1149-1157) with the given variance
static void gauss_pair(double variance, double &u1, double &u2) {
  const double a = std::sqrt(-2 * std::log(u1)) * std::cos(2 * M_PI * u2);
  const double b = std::sqrt(-2 * std::log(u1)) * std::sin(2 * M_PI * u2);
  const double sv = std::sqrt(variance);
  u1 = a * sv;
  u2 = b * sv;
}

// The reference's zeroing of a momentum component (v25/functions.cpp:
// 728-770): flip signs, walking cyclically over all but the last particle,
// until |sum| <= pr, then let the last particle cancel the rest.
static void zero_sum(std::vector<double> &v, int n, double pr) {
  double S = 0.;
  for (int i = 0; i < n; ++i) S += v[i];
  int k = 0;
  for (long guard = 0; std::fabs(S) > pr && guard <= 100000000L; ++guard) {
    while (S > pr) {
      if (v[k] > 0) {
        v[k] = -v[k];
        S += 2 * v[k];
      }
      if (++k == n - 1) k = 0;
    }
    while (S < -pr) {
      if (v[k] < 0) {
        v[k] = -v[k];
        S += 2 * v[k];
      }
      if (++k == n - 1) k = 0;
    }
  }
  S -= v[n - 1];
  v[n - 1] = -S;
}

// ini_mom (v25/functions.cpp:677-1146), one species: uniforms for every
// component, Gaussian pairs of variance kT m (kT I_a for the angular
// momenta; an odd last particle draws one more uniform), zero totals.
static void ini_mom(int N, double temp, double mass, const double I[3], Mt19937 &mt, double *p, double *l) {
  std::vector<double> c[3];
  for (int pass = 0; pass < 2; ++pass) {
    for (auto &v : c) v.assign((size_t)N + 1, 0.0);
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < 3; ++j) c[j][i] = mt.uniform();
    for (int i = 0; i < N; i += 2)
      for (int j = 0; j < 3; ++j) {
        const double var = pass == 0 ? temp * mass : temp * I[j];
        if ((std::fmod((float)N, 2.f) != 0.) && i == N - 1) {
          double u = mt.uniform();
          gauss_pair(var, u, c[j][i]);
        } else {
          gauss_pair(var, c[j][i], c[j][i + 1]);
        }
      }
    for (int j = 0; j < 3; ++j) zero_sum(c[j], N, 0.1);
    double *out = pass == 0 ? p : l;
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < 3; ++j) out[3 * i + j] = c[j][i];
  }
}

}  // namespace init

// Cluster numbering of clust_conf (v25/functions.cpp:3270-3301): while
// particles are left, rnd_part (v25/functions.cpp:2950-2970) draws
// ip = floor(u * left) and takes the ip-th particle left in global order
// (species-major, as the per-species index lists NI concatenate); its whole
// component gets the next number.  The particles left are a bit set with a
// Fenwick tree over the popcounts of its words, so each draw finds its
// particle in O(log N).
static int order_clusters(const std::vector<int> &root, Mt19937 &mt, int *labels) {
  const int N = (int)root.size();
  std::vector<int> start(N + 1, 0), members(N);
  for (int r : root) ++start[r + 1];
  for (int i = 0; i < N; ++i) start[i + 1] += start[i];
  {
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int i = 0; i < N; ++i) members[fill[root[i]]++] = i;
  }
  const int W = (N + 63) / 64;
  std::vector<uint64_t> bits(W, ~0ull);
  if (N % 64) bits[W - 1] = (1ull << (N % 64)) - 1;
  std::vector<int> tree(W + 1, 0);  // Fenwick tree, 1-based
  for (int w = 0; w < W; ++w) {
    tree[w + 1] += __builtin_popcountll(bits[w]);
    const int up = (w + 1) + ((w + 1) & -(w + 1));
    if (up <= W) tree[up] += tree[w + 1];
  }
  int top = 1;
  while (top * 2 <= W) top *= 2;
  int left = N, nc = 0;
  while (left > 0) {
    int ip = (int)std::floor(mt.uniform() * left);
    int pos = 0;
    for (int step = top; step; step >>= 1)
      if (pos + step <= W && tree[pos + step] <= ip) {
        pos += step;
        ip -= tree[pos];
      }
    uint64_t word = bits[pos];
    for (int k = 0; k < ip; ++k) word &= word - 1;
    const int seed = pos * 64 + __builtin_ctzll(word);
    const int r = root[seed];
    for (int k = start[r]; k < start[r + 1]; ++k) {
      const int m = members[k];
      labels[m] = nc;
      bits[m >> 6] &= ~(1ull << (m & 63));
      for (int i = (m >> 6) + 1; i <= W; i += i & -i) --tree[i];
    }
    left -= start[r + 1] - start[r];
    ++nc;
  }
  return nc;
}

extern "C" {

int msc_version(void) { return 1; }

const char *msc_last_error(const msc_ctx *ctx) {
  if (ctx) return ctx->err.c_str();
  return g_err.empty() ? "no context" : g_err.c_str();
}

int msc_create(int device, int precision, int nrep, const msc_replica_desc *reps, msc_ctx **out) {
  if (!out || nrep <= 0 || !reps || (precision != MSC_FP64 && precision != MSC_MIXED)) return MSC_E_ARG;
  *out = nullptr;
  msc_ctx *ctx = new msc_ctx();
  ctx->device = device;
  ctx->precision = precision;
  ctx->mixed = precision == MSC_MIXED;
  ctx->contact = ctx->mixed;
  ctx->list_skin = ctx->mixed ? kDefaultListSkin : 0.0;
  ctx->nrep = nrep;
  ctx->desc.assign(reps, reps + nrep);
  int nd = reps[0].nd;
  for (int r = 0; r < nrep; ++r) {
    const msc_replica_desc &D = reps[r];
    if (D.N <= 0 || D.nd != nd || nd < 1 || nd > kMaxNd || D.K < 2 || D.K > kMaxKZ || D.L[0] <= 0 || D.L[1] <= 0 ||
        D.L[2] <= 0 || D.alpha <= 0) {
      delete ctx;
      return MSC_E_ARG;
    }
  }
  ctx->nd = nd;
  {
    hipError_t e1 = hipSetDevice(device);
    hipError_t e2 = e1 == hipSuccess ? device_stream(device, &ctx->stream) : e1;
    if (e2 == hipSuccess) e2 = device_stream2(device, &ctx->stream2);
    if (e2 == hipSuccess) e2 = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming);
    if (e2 == hipSuccess) e2 = hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming);
    if (e2 != hipSuccess) {
      g_err = std::string(e1 != hipSuccess ? "hipSetDevice: " : "stream or event creation: ") + hipGetErrorString(e2);
      if (ctx->ev_fork) hipEventDestroy(ctx->ev_fork);
      if (ctx->ev_join) hipEventDestroy(ctx->ev_join);
      delete ctx;
      return MSC_E_HIP;
    }
    hipError_t e3 = hipHostMalloc(&ctx->pin, kPinBytes, hipHostMallocDefault);
    if (e3 != hipSuccess) {
      g_err = std::string("hipHostMalloc: ") + hipGetErrorString(e3);
      ctx->pin = nullptr;
      delete ctx;
      return MSC_E_HIP;
    }
  }
  // replica layout: particles, tasks, chunk tasks, k slots, tables, partials
  int n0 = 0, slot0 = 0, bexp_cols = 0;
  long part0 = 0;
  ctx->reps.resize(nrep);
  ctx->sh_dmin.assign(nrep, 0.0);
  ctx->sh_dmax.assign(nrep, 0.0);
  for (int r = 0; r < nrep; ++r) {
    const msc_replica_desc &D = reps[r];
    RepDev &R = ctx->reps[r];
    memset(&R, 0, sizeof R);
    R.n0 = n0;
    R.N = D.N;
    R.nd = nd;
    R.K = D.K;
    R.kmax = D.K - 1;
    R.KZ = D.K;
    R.ncol = (2 * R.kmax + 1) * (2 * R.kmax + 1);
    R.slot0 = slot0;
    R.tab0 = r * kTabN;
    for (int a = 0; a < 3; ++a) {
      R.L[a] = D.L[a];
      R.iL[a] = 1.0 / D.L[a];
      R.P2[a] = 2. * M_PI / D.L[a];
      R.B[a] = D.B[a];
    }
    R.V = D.L[0] * D.L[1] * D.L[2];
    R.alpha = D.alpha;
    R.rc = D.rc;
    R.sr_cut = D.sr_cut;
    R.temp = D.temp;
    R.dt = D.dt;
    R.Q = D.Q;
    R.gamma = D.gamma;
    R.dens = D.dens;
    R.task0 = (int)ctx->tasks.size();
    for (int g = 0; g < D.N; g += kWave) ctx->tasks.push_back({r, n0 + g, std::min(kWave, D.N - g), 0});
    R.ntask = (int)ctx->tasks.size() - R.task0;
    R.rtask0 = (int)ctx->rtasks.size();
    for (int g = 0; g < D.N; g += ctx->rchunk)
      ctx->rtasks.push_back({r, n0 + g, std::min(ctx->rchunk, D.N - g), 0});
    R.nrtask = (int)ctx->rtasks.size() - R.rtask0;
    R.part0 = part0;
    part0 += (long)R.nrtask * R.ncol * R.KZ * 6;
    R.bop0 = ctx->bop_elems;
    ctx->bop_elems += (long)((R.ncol + 15) / 16) * kBopPerTile;
    R.bexp0 = 3 * bexp_cols;
    bexp_cols += (R.ncol + 15) / 16;
    slot0 += R.ncol * R.KZ;
    ctx->maxslot = std::max(ctx->maxslot, R.ncol * R.KZ);
    ctx->maxcol = std::max(ctx->maxcol, R.ncol);
    if (std::find(ctx->kz_groups.begin(), ctx->kz_groups.end(), R.KZ) == ctx->kz_groups.end())
      ctx->kz_groups.push_back(R.KZ);
    n0 += D.N;
  }
  ctx->Ntot = n0;
  ctx->nslot_tot = slot0;
  ctx->part_elems = (size_t)part0;
  ctx->part_cap = ctx->part_elems;
  ctx->rtask_cap = ctx->rtasks.size();
  ctx->thermo_rng.resize(2 * nrep);
  for (int r = 0; r < nrep; ++r)
    for (int i = 0; i < 2; ++i) ctx->thermo_rng[2 * r + i].seed(reps[r].thermo_seed[i]);
  const int N = ctx->Ntot;
  const size_t rs = ctx->mixed ? sizeof(float) : sizeof(double);
  int rc = MSC_OK;
  auto A = [&](auto **p, size_t n) {
    if (rc == MSC_OK) rc = dalloc(ctx, p, n);
  };
  A(&ctx->d_reps, nrep);
  A(&ctx->d_tasks, ctx->tasks.size());
  A(&ctx->d_rtasks, ctx->rtasks.size());
  A(&ctx->d_rep_of, N);
  A(&ctx->d_x, 3 * N); A(&ctx->d_p, 3 * N); A(&ctx->d_l, 3 * N); A(&ctx->d_ori, 9 * N);
  A(&ctx->d_F, 3 * N); A(&ctx->d_T, 3 * N);
  A(&ctx->d_len, N); A(&ctx->d_dia, N); A(&ctx->d_mass, N); A(&ctx->d_inert, 3 * N);
  A(&ctx->d_dpc, nd * N); A(&ctx->d_doc, 3 * nd * N); A(&ctx->d_dmu, nd * N);
  A(&ctx->d_dpos, 3 * nd * N); A(&ctx->d_dori, 3 * nd * N); A(&ctx->d_snap, 3 * N);
  A(&ctx->d_snap_e, 6 * N);
  A(&ctx->d_snap_u, 3 * N);
  A(&ctx->d_clu, N);
  A(&ctx->d_clu_flag, 2);
  A(&ctx->d_rows, ctx->tasks.size() * 16 * kRowSets);
  A(&ctx->d_kblk, (size_t)nrep * ((ctx->maxslot + kKblkSlots - 1) / kKblkSlots) * 10);
  A(&ctx->d_endsum, (size_t)nrep * kEndSums);
  A(&ctx->d_kick_ts, (size_t)nrep * 2);
  A(&ctx->d_step, 2);
  A(&ctx->d_cell_of, N); A(&ctx->d_cell_list, N);
  A(&ctx->d_cell_tmp, N);
  A(&ctx->d_cell_pack, (size_t)8 * N);
  A(&ctx->d_nl_cnt, N); A(&ctx->d_sl_cnt, N); A(&ctx->d_dl_cnt, N);
  A(&ctx->d_kval, ctx->nslot_tot);
  A(&ctx->d_slot_out, (size_t)ctx->nslot_tot * 10);
  A(&ctx->d_task_out, ctx->tasks.size() * 16);
  A(&ctx->d_red, (size_t)nrep * kRedStride);
  A(&ctx->d_M2, (size_t)nrep * 6);
  A(&ctx->d_Kin, (size_t)nrep * 2);
  A(&ctx->d_gs, (size_t)nrep * 4);
  A(&ctx->d_obs, (size_t)nrep * MSC_OBS_N);
  A(&ctx->d_counts, (size_t)nrep * 4);
  A(&ctx->d_E, nrep);
  A(&ctx->d_th, nrep);
  A(&ctx->d_flags, (size_t)nrep * 4);
  A(&ctx->d_rebuild, nrep);
  A(&ctx->d_srebuild, nrep);
  A(&ctx->d_lctr, nrep);
  A(&ctx->d_dmax, (size_t)4 * kDmaxBuckets * nrep);
  A(&ctx->d_one, nrep);
  A(&ctx->d_stamp, 1);
  // zeroed here once: entries for nz >= K and columns past the last stay 0
  if (ctx->mixed && rc == MSC_OK) rc = dalloc(ctx, (unsigned long long **)&ctx->d_bop, ctx->bop_elems);
  if (ctx->mixed && rc == MSC_OK) rc = dalloc(ctx, &ctx->d_bexp, (size_t)3 * bexp_cols);
  if (ctx->mixed && rc == MSC_OK) rc = dalloc(ctx, &ctx->d_rec_rod, (size_t)8 * N);
  if (ctx->mixed && rc == MSC_OK) rc = dalloc(ctx, &ctx->d_rec_dip, (size_t)8 * nd * N);
  if (nd > 1 && rc == MSC_OK) rc = dalloc(ctx, (unsigned char **)&ctx->d_ph_nd, (size_t)nd * 3 * kMaxKZ * N * 16);
  if (nd > 1 && rc == MSC_OK) rc = dalloc(ctx, &ctx->d_part_nd, ctx->part_elems / 6 * 10);
  if (nd > 1 && rc == MSC_OK) rc = dalloc(ctx, (unsigned char **)&ctx->d_A_nd, (size_t)ctx->nslot_tot * 16);
  if (rc == MSC_OK) {
    unsigned char **raw[] = {(unsigned char **)&ctx->d_tab, (unsigned char **)&ctx->d_ph,
                             (unsigned char **)&ctx->d_amp, (unsigned char **)&ctx->d_part,
                             (unsigned char **)&ctx->d_A};
    size_t sz[] = {(size_t)nrep * kTabN * 3 * rs, (size_t)3 * kMaxKZ * N * 2 * rs, (size_t)3 * N * rs,
                   ctx->part_elems * rs, (size_t)ctx->nslot_tot * 2 * rs};
    for (int q = 0; q < 5 && rc == MSC_OK; ++q) rc = dalloc(ctx, raw[q], sz[q]);
  }
  if (rc != MSC_OK) {
    g_err = "msc_create: " + ctx->err;
    msc_destroy(ctx);
    return MSC_E_HIP;
  }
  std::vector<int> rep_of(N);
  for (int r = 0; r < nrep; ++r)
    for (int i = 0; i < ctx->reps[r].N; ++i) rep_of[ctx->reps[r].n0 + i] = r;
  std::vector<int> ones(nrep, 1);
  std::vector<Thermo> th(nrep);
  for (int r = 0; r < nrep; ++r) {
    th[r].s[0] = th[r].s[1] = reps[r].s0;
    th[r].strn[0] = th[r].strn[1] = 0;
  }
  std::vector<double> &kval = ctx->kval_host;
  build_kspace(ctx, kval);
  std::vector<double> tab((size_t)nrep * kTabN * 3);
  for (int r = 0; r < nrep; ++r) build_table(ctx->reps[r], tab);
  std::vector<float> tf;
  if (ctx->mixed) tf.assign(tab.begin(), tab.end());
  auto U = [&](void *dst, const void *src, size_t bytes) {
    if (rc == MSC_OK) rc = h2d(ctx, dst, src, bytes);
  };
  U(ctx->d_rep_of, rep_of.data(), sizeof(int) * N);
  U(ctx->d_tasks, ctx->tasks.data(), sizeof(Task) * ctx->tasks.size());
  U(ctx->d_rtasks, ctx->rtasks.data(), sizeof(Task) * ctx->rtasks.size());
  U(ctx->d_one, ones.data(), sizeof(int) * nrep);
  U(ctx->d_th, th.data(), sizeof(Thermo) * nrep);
  U(ctx->d_kval, kval.data(), sizeof(double) * kval.size());
  if (ctx->mixed)
    U(ctx->d_tab, tf.data(), tf.size() * sizeof(float));
  else
    U(ctx->d_tab, tab.data(), tab.size() * sizeof(double));
  if (rc == MSC_OK) rc = upload_reps(ctx);
  if (rc != MSC_OK) {
    g_err = "msc_create: " + ctx->err;
    msc_destroy(ctx);
    return MSC_E_HIP;
  }
  refresh_state_ptrs(ctx);
  *out = ctx;
  return MSC_OK;
}

void msc_destroy(msc_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  if (ctx->stream2) hipStreamSynchronize(ctx->stream2);
  hipDeviceSynchronize();
  if (ctx->graph_exec) hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->ev_fork) hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) hipEventDestroy(ctx->ev_join);
  void *bufs[] = {ctx->d_reps, ctx->d_tasks, ctx->d_rtasks, ctx->d_rep_of, ctx->d_cell_rep, ctx->d_x, ctx->d_p,
                  ctx->d_l, ctx->d_ori, ctx->d_F, ctx->d_T, ctx->d_len, ctx->d_dia, ctx->d_mass, ctx->d_inert,
                  ctx->d_dpc, ctx->d_doc, ctx->d_dmu, ctx->d_dpos, ctx->d_dori, ctx->d_snap, ctx->d_cell_of,
                  ctx->d_cell_cnt, ctx->d_cell_start, ctx->d_cell_fill, ctx->d_cell_list, ctx->d_nl_cnt,
                  ctx->d_nl_idx, ctx->d_dl_cnt, ctx->d_dl_idx, ctx->d_kval, ctx->d_cell_tmp, ctx->d_sl_cnt,
                  ctx->d_sl_idx, ctx->d_snap_e, ctx->d_srebuild, ctx->d_cell_pack, 
                  ctx->d_snap_u, ctx->d_clu, ctx->d_clu_flag, ctx->d_corr, ctx->d_corr_drop, ctx->d_order, ctx->d_rows, ctx->d_kblk, ctx->d_endsum, ctx->d_kick_ts, ctx->d_step,
                  ctx->d_tab, ctx->d_ph, ctx->d_amp, ctx->d_part, ctx->d_A, ctx->d_slot_out, ctx->d_task_out,
                  ctx->d_red, ctx->d_M2, ctx->d_Kin, ctx->d_gs, ctx->d_obs, ctx->d_counts, ctx->d_E,
                  ctx->d_th, ctx->d_flags, ctx->d_rebuild, ctx->d_one, ctx->d_run_gs, ctx->d_run_obs, ctx->d_save, ctx->d_tune_save,
                  ctx->d_stamp, ctx->d_bop, ctx->d_bexp, ctx->d_lctr, ctx->d_dmax, ctx->d_rec_rod,
                  ctx->d_rec_dip, ctx->d_ph_nd, ctx->d_part_nd, ctx->d_A_nd};
  for (void *b : bufs)
    if (b) hipFree(b);
  for (auto &kv : ctx->timers)
    for (auto &ev : kv.second.pending) {
      hipEventDestroy(ev.first);
      hipEventDestroy(ev.second);
    }
  for (hipEvent_t e : ctx->event_pool) hipEventDestroy(e);
  if (ctx->pin) hipHostFree(ctx->pin);
  delete ctx;
}

int msc_set_state(msc_ctx *ctx, int rep, const double *r, const double *p, const double *l, const double *ori,
                  const double *len, const double *dia, const double *mass, const double *dpc, const double *doc,
                  const double *dmu) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep || !r || !p || !l || !ori || !len || !dia || !mass || !dpc || !doc || !dmu)
    return MSC_E_ARG;
  RepDev &R = ctx->reps[rep];
  const int n = R.N, N = ctx->Ntot, nd = ctx->nd;
  std::vector<double> buf;
  auto put = [&](double *dst, const double *src, int ncomp) -> int {
    buf.resize((size_t)n);
    for (int c = 0; c < ncomp; ++c) {
      for (int i = 0; i < n; ++i) buf[i] = src[(size_t)i * ncomp + c];
      if (int rc_ = h2d(ctx, dst + (size_t)c * N + R.n0, buf.data(), sizeof(double) * n)) return rc_;
    }
    return MSC_OK;
  };
  int rc;
  if ((rc = put(ctx->d_x, r, 3)) || (rc = put(ctx->d_p, p, 3)) || (rc = put(ctx->d_l, l, 3)) ||
      (rc = put(ctx->d_ori, ori, 9)) || (rc = put(ctx->d_len, len, 1)) || (rc = put(ctx->d_dia, dia, 1)) ||
      (rc = put(ctx->d_mass, mass, 1)))
    return rc;
  // dipole constraints: host layout [N][nd], [N][nd][3] -> [nd][N], [3nd][N]
  if ((rc = put(ctx->d_dpc, dpc, nd)) || (rc = put(ctx->d_dmu, dmu, nd)) || (rc = put(ctx->d_doc, doc, 3 * nd)))
    return rc;
  std::vector<double> I((size_t)3 * n);
  for (int i = 0; i < n; ++i) inertia(len[i], dia[i], mass[i], &I[3 * (size_t)i]);
  if ((rc = put(ctx->d_inert, I.data(), 3))) return rc;
  derive(ctx, rep, len, dia, dpc);
  upload_reps(ctx);
  if (ctx->mixed) {
    double musum = 0;
    for (int i = 0; i < n * nd; ++i) musum += std::fabs(dmu[i]);
    std::vector<int> ex;
    build_bop_exp(ctx, rep, musum, ex);
    if ((rc = h2d(ctx, ctx->d_bexp + R.bexp0, ex.data(), sizeof(int) * ex.size()))) return rc;
  }
  launch_dipoles(sel(ctx, rep), ctx->stream);
  FINISH("msc_set_state");
  ctx->have_state = true;
  return MSC_OK;
}

int msc_get_state(msc_ctx *ctx, int rep, double *r, double *p, double *l, double *ori) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep) return MSC_E_ARG;
  const RepDev &R = ctx->reps[rep];
  const int n = R.N, N = ctx->Ntot;
  std::vector<double> buf(n);
  auto get = [&](double *dst, const double *src, int ncomp) -> int {
    if (!dst) return MSC_OK;
    for (int c = 0; c < ncomp; ++c) {
      if (int rc_ = d2h(ctx, buf.data(), src + (size_t)c * N + R.n0, sizeof(double) * n)) return rc_;
      for (int i = 0; i < n; ++i) dst[(size_t)i * ncomp + c] = buf[i];
    }
    return MSC_OK;
  };
  HIPC(hipStreamSynchronize(ctx->stream));
  int rc;
  if ((rc = get(r, ctx->d_x, 3)) || (rc = get(p, ctx->d_p, 3)) || (rc = get(l, ctx->d_l, 3)) ||
      (rc = get(ori, ctx->d_ori, 9)))
    return rc;
  return MSC_OK;
}

int msc_get_forces(msc_ctx *ctx, int rep, double *F, double *T) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep) return MSC_E_ARG;
  const RepDev &R = ctx->reps[rep];
  const int n = R.N, N = ctx->Ntot;
  std::vector<double> buf(n);
  HIPC(hipStreamSynchronize(ctx->stream));
  for (int c = 0; c < 3; ++c) {
    if (F) {
      if (int rc_ = d2h(ctx, buf.data(), ctx->d_F + (size_t)c * N + R.n0, sizeof(double) * n)) return rc_;
      for (int i = 0; i < n; ++i) F[3 * i + c] = buf[i];
    }
    if (T) {
      if (int rc_ = d2h(ctx, buf.data(), ctx->d_T + (size_t)c * N + R.n0, sizeof(double) * n)) return rc_;
      for (int i = 0; i < n; ++i) T[3 * i + c] = buf[i];
    }
  }
  return MSC_OK;
}

int msc_set_forces(msc_ctx *ctx, int rep, const double *F, const double *T) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep || !F || !T) return MSC_E_ARG;
  const RepDev &R = ctx->reps[rep];
  const int n = R.N, N = ctx->Ntot;
  std::vector<double> buf(n);
  for (int c = 0; c < 3; ++c) {
    for (int i = 0; i < n; ++i) buf[i] = F[3 * i + c];
    if (int rc_ = h2d(ctx, ctx->d_F + (size_t)c * N + R.n0, buf.data(), sizeof(double) * n)) return rc_;
    for (int i = 0; i < n; ++i) buf[i] = T[3 * i + c];
    if (int rc_ = h2d(ctx, ctx->d_T + (size_t)c * N + R.n0, buf.data(), sizeof(double) * n)) return rc_;
  }
  return MSC_OK;
}

int msc_get_thermostat(msc_ctx *ctx, int rep, double s[2], double strn[2]) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep) return MSC_E_ARG;
  Thermo t;
  HIPC(hipStreamSynchronize(ctx->stream));
  if (int rc_ = d2h(ctx, &t, ctx->d_th + rep, sizeof t)) return rc_;
  if (s) { s[0] = t.s[0]; s[1] = t.s[1]; }
  if (strn) { strn[0] = t.strn[0]; strn[1] = t.strn[1]; }
  return MSC_OK;
}

int msc_set_thermostat(msc_ctx *ctx, int rep, const double s[2], const double strn[2]) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep || !s || !strn) return MSC_E_ARG;
  Thermo t;
  t.s[0] = s[0]; t.s[1] = s[1]; t.strn[0] = strn[0]; t.strn[1] = strn[1];
  return h2d(ctx, ctx->d_th + rep, &t, sizeof t);
}

// --------------------------------------------------------------------------
// neighbour lists
// --------------------------------------------------------------------------
static int lists_build(msc_ctx *ctx, const int *d_flagmask, bool zero_cells = true) {
  State s = sel(ctx, MSC_ALL);
  {
    Bracket b(ctx, "nlist");
    launch_cells(s, d_flagmask, zero_cells, ctx->stream);
    launch_lists(s, d_flagmask, ctx->d_flags, ctx->stream);
  }
  return MSC_OK;
}

// Expected full-list entries per particle of a fluid at the replica's number
// density: rho times the volume within segment distance r of a rod, averaged
// over random relative orientations, 4/3 pi r^3 + pi r^2 (L_i + L_j) +
// (pi/2) r L_i L_j, with L_i = L_j = the largest length.
static double fluid_neighbours(const RepDev &R, double r) {
  double rho = R.N / R.V, L = R.lmax;
  return rho * (4.0 / 3.0 * M_PI * r * r * r + M_PI * r * r * 2 * L + 0.5 * M_PI * r * L * L);
}

// MSC_TIGHT_LISTS=1 in the environment sizes rows to the present longest row
// only, so that any growth takes the overflow path of msc_run.
static bool tight_lists() {
  const char *v = std::getenv("MSC_TIGHT_LISTS");
  return v && std::atoi(v) != 0;
}

// Cells and neighbour-row capacities of every replica.  A measuring pass
// builds the rows with capacity 0, which records the longest row of each
// list per replica; the capacity is that length with headroom (and at least
// 1.5 times the mean row of a fluid at the replica's density), or what a
// rebuild of a running system found missing (need_*: longest rows seen by
// an overflowing rebuild, 0 where nothing overflowed).
static int lists_layout(msc_ctx *ctx, const std::vector<long> &need_nl = {},
                        const std::vector<long> &need_sl = {}, const std::vector<long> &need_dl = {}) {
  drop_graph(ctx);
  const int nrep = ctx->nrep;
  int ncells = 0;
  std::vector<int> cell_rep;
  for (int r = 0; r < nrep; ++r) {
    grid(ctx, r);
    ctx->reps[r].cell0 = ncells;
    ncells += ctx->reps[r].ncell;
    cell_rep.insert(cell_rep.end(), ctx->reps[r].ncell, r);
  }
  if (ncells != ctx->ncells || !ctx->d_cell_cnt) {
    int rc;
    if ((rc = dalloc(ctx, &ctx->d_cell_cnt, ncells)) || (rc = dalloc(ctx, &ctx->d_cell_start, ncells)) ||
        (rc = dalloc(ctx, &ctx->d_cell_fill, ncells)) || (rc = dalloc(ctx, &ctx->d_cell_rep, ncells)))
      return rc;
    ctx->ncells = ncells;
  }
  if (int rc_ = h2d(ctx, ctx->d_cell_rep, cell_rep.data(), sizeof(int) * ncells)) return rc_;
  // measuring pass
  for (int r = 0; r < nrep; ++r) {
    RepDev &R = ctx->reps[r];
    R.nl_w = R.sl_w = R.dl_w = 0;
    R.nl_base = R.sl_base = R.dl_base = 0;
  }
  int rc;
  if ((rc = upload_reps(ctx))) return rc;
  refresh_state_ptrs(ctx);
  State s = sel(ctx, MSC_ALL);
  launch_fill_int(ctx->d_flags, 4 * nrep, 0, ctx->stream);
  launch_cells(s, ctx->d_one, true, ctx->stream);
  launch_lists(s, ctx->d_one, ctx->d_flags, ctx->stream);
  std::vector<int> fl(4 * nrep);
  if (int rc_ = d2h(ctx, fl.data(), ctx->d_flags, sizeof(int) * fl.size())) return rc_;
  const bool tight = tight_lists();
  long nl_tot = 0, sl_tot = 0, dl_tot = 0;
  for (int r = 0; r < nrep; ++r) {
    RepDev &R = ctx->reps[r];
    long mw = fl[4 * r + 2], ms = fl[4 * r + 1], md = fl[4 * r + 3];
    auto width = [&](long longest, double fluid, long need) {
      long w = tight ? longest : std::max(longest + longest / 4 + 8, (long)std::ceil(1.5 * fluid));
      if (need > 0) w = std::max(w, need + need / 4 + 8);
      return std::max(w, 1L);
    };
    R.nl_w = (int)width(mw, fluid_neighbours(R, R.r_list_max), need_nl.empty() ? 0 : need_nl[r]);
    R.sl_w = (int)width(ms, fluid_neighbours(R, R.r_short), need_sl.empty() ? 0 : need_sl[r]);
    R.dl_w = 0;
    if (R.rc > 0) {
      double rd = R.rl_dip + 2 * ctx->S.lever_max;
      R.dl_w = (int)width(md, R.N / R.V * 4.0 / 3.0 * M_PI * rd * rd * rd, need_dl.empty() ? 0 : need_dl[r]);
    }
    R.nl_base = nl_tot;
    R.sl_base = sl_tot;
    R.dl_base = dl_tot;
    nl_tot += (long)R.nl_w * R.N;
    sl_tot += (long)R.sl_w * R.N;
    dl_tot += (long)R.dl_w * R.N;
  }
  if (nl_tot > ctx->nl_alloc) {
    if ((rc = dalloc(ctx, &ctx->d_nl_idx, nl_tot))) return rc;
    ctx->nl_alloc = nl_tot;
  }
  if (sl_tot > ctx->sl_alloc) {
    if ((rc = dalloc(ctx, &ctx->d_sl_idx, sl_tot))) return rc;
    ctx->sl_alloc = sl_tot;
  }
  if (dl_tot > ctx->dl_alloc) {
    if ((rc = dalloc(ctx, &ctx->d_dl_idx, std::max(dl_tot, 1L)))) return rc;
    ctx->dl_alloc = dl_tot;
  }
  if ((rc = upload_reps(ctx))) return rc;
  refresh_state_ptrs(ctx);
  return MSC_OK;
}

int msc_neigh_list(msc_ctx *ctx, int rep, double r_list) {
  SCOPE(ctx);
  if (!ctx || !rep_ok(ctx, rep) || !(r_list > 0)) return MSC_E_ARG;
  if (!ctx->have_state) return fail(ctx, MSC_E_STATE, "no state");
  // the rows are valid again only once this build has fitted its regions
  ctx->have_lists = false;
  for (int r = 0; r < ctx->nrep; ++r)
    if (rep == MSC_ALL || r == rep) {
      ctx->reps[r].r_list_max = r_list;
      clamp_short(ctx->reps[r]);
    }
  int rc = lists_layout(ctx);
  if (rc) return rc;
  launch_fill_int(ctx->d_flags, 4 * ctx->nrep, 0, ctx->stream);
  if ((rc = lists_build(ctx, ctx->d_one))) return rc;
  FINISH("msc_neigh_list");
  std::vector<int> fl(4 * ctx->nrep);
  if (int rc_ = d2h(ctx, fl.data(), ctx->d_flags, sizeof(int) * fl.size())) return rc_;
  for (int r = 0; r < ctx->nrep; ++r)
    if (fl[4 * r + 1] || fl[4 * r + 2] || fl[4 * r + 3])
      return fail(ctx, MSC_E_CAPACITY, "neighbour rows longer than measured");
  ctx->have_lists = true;
  return MSC_OK;
}

// The caller's list in the reference's form (rows f of ascending g > f) as
// replica rep's reference rows.  Every replica's rows and list snapshots are
// built first (regions sized to hold the caller's longest row), then rep's
// rows are replaced by the caller's pairs, both directions, in ascending
// partner order, and its contact rows are refreshed from them.
int msc_set_neigh_list(msc_ctx *ctx, int rep, double r_list, const int *off, const int *nbr) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep || !(r_list > 0) || !off) return MSC_E_ARG;
  if (!ctx->have_state) return fail(ctx, MSC_E_STATE, "no state");
  RepDev &R0 = ctx->reps[rep];
  const int n = R0.N;
  if (off[0] != 0) return fail(ctx, MSC_E_ARG, "off[0] must be 0");
  for (int f = 0; f < n; ++f)
    if (off[f + 1] < off[f]) return fail(ctx, MSC_E_ARG, "off is not ascending");
  if (off[n] > 0 && !nbr) return fail(ctx, MSC_E_ARG, "no pairs given");
  std::vector<int> cnt(n, 0);
  for (int f = 0; f < n; ++f)
    for (int q = off[f]; q < off[f + 1]; ++q) {
      const int g = nbr[q];
      if (g <= f || g >= n || (q > off[f] && g <= nbr[q - 1]))
        return fail(ctx, MSC_E_ARG, "row f must hold ascending partners g > f below N");
      ++cnt[f];
      ++cnt[g];
    }
  long longest = 0;
  for (int c : cnt) longest = std::max<long>(longest, c);
  ctx->have_lists = false;
  R0.r_list_max = r_list;
  clamp_short(R0);
  std::vector<long> need(ctx->nrep, 0);
  need[rep] = longest;
  int rc = lists_layout(ctx, need);
  if (rc) return rc;
  launch_fill_int(ctx->d_flags, 4 * ctx->nrep, 0, ctx->stream);
  if ((rc = lists_build(ctx, ctx->d_one))) return rc;
  const RepDev &R = ctx->reps[rep];
  if (longest > R.nl_w) return fail(ctx, MSC_E_CAPACITY, "row region smaller than the caller's longest row");
  // rows in the device layout: entry q of particle i at q * N + i (global
  // partner indices); ascending partners, as f < g visits them
  std::vector<int> idx((size_t)std::max<long>((long)R.nl_w * n, 1), 0), fill(n, 0);
  for (int f = 0; f < n; ++f)
    for (int q = off[f]; q < off[f + 1]; ++q) {
      const int g = nbr[q];
      idx[(size_t)fill[g]++ * n + g] = R.n0 + f;
    }
  // every row g so far holds its partners f < g in ascending order; the
  // partners above follow, ascending
  for (int f = 0; f < n; ++f)
    for (int q = off[f]; q < off[f + 1]; ++q) idx[(size_t)fill[f]++ * n + f] = R.n0 + nbr[q];
  if ((rc = h2d(ctx, ctx->d_nl_cnt + R.n0, cnt.data(), sizeof(int) * n))) return rc;
  if ((rc = h2d(ctx, ctx->d_nl_idx + R.nl_base, idx.data(), sizeof(int) * (size_t)R.nl_w * n))) return rc;
  if (ctx->contact) launch_short(sel(ctx, rep), ctx->d_one, ctx->d_flags, ctx->stream);
  FINISH("msc_set_neigh_list");
  std::vector<int> fl(4 * ctx->nrep);
  if (int rc_ = d2h(ctx, fl.data(), ctx->d_flags, sizeof(int) * fl.size())) return rc_;
  for (int r = 0; r < ctx->nrep; ++r)
    if (fl[4 * r + 1] || fl[4 * r + 2] || fl[4 * r + 3])
      return fail(ctx, MSC_E_CAPACITY, "neighbour rows longer than measured");
  ctx->have_lists = true;
  return MSC_OK;
}

int msc_get_neigh_list(msc_ctx *ctx, int rep, int *off, int *nbr, long cap, long *total) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep || !off || !total) return MSC_E_ARG;
  if (!ctx->have_lists) return fail(ctx, MSC_E_STATE, "no neighbour list");
  const RepDev &R = ctx->reps[rep];
  std::vector<int> cnt(R.N), idx(std::max<size_t>((size_t)R.nl_w * R.N, 1));
  if (int rc_ = d2h(ctx, cnt.data(), ctx->d_nl_cnt + R.n0, sizeof(int) * R.N)) return rc_;
  if (int rc_ = d2h(ctx, idx.data(), ctx->d_nl_idx + R.nl_base, sizeof(int) * (size_t)R.nl_w * R.N)) return rc_;
  long t = 0;
  std::vector<int> row;
  for (int i = 0; i < R.N; ++i) {
    off[i] = (int)t;
    row.clear();
    for (int q = 0; q < cnt[i]; ++q) {
      int j = idx[(size_t)q * R.N + i] - R.n0;
      if (j > i) row.push_back(j);
    }
    std::sort(row.begin(), row.end());
    for (int j : row) {
      if (nbr && t < cap) nbr[t] = j;
      ++t;
    }
  }
  off[R.N] = (int)t;
  *total = t;
  return MSC_OK;
}

int msc_check_list(msc_ctx *ctx, int rep, double r_list_min, double r_list_max, int *rebuild) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep || !rebuild) return MSC_E_ARG;
  if (!ctx->have_lists) return fail(ctx, MSC_E_STATE, "no neighbour list");
  ctx->reps[rep].r_list_min = r_list_min;
  ctx->reps[rep].r_list_max = r_list_max;
  clamp_short(ctx->reps[rep]);
  int rc;
  if ((rc = upload_reps(ctx))) return rc;
  State s = sel(ctx, rep);
  launch_check(s, ctx->d_task_out, ctx->d_rebuild, nullptr, nullptr, 0, false, ctx->stream);
  FINISH("msc_check_list");
  if (int rc_ = d2h(ctx, rebuild, ctx->d_rebuild + rep, sizeof(int))) return rc_;
  return MSC_OK;
}

// --------------------------------------------------------------------------
// forces
// --------------------------------------------------------------------------
// energies of the last force evaluation (after finish())
static int read_energy(msc_ctx *ctx, int rep, RepEnergy *e) {
  return d2h(ctx, e, ctx->d_E + rep, sizeof(RepEnergy));
}

static void do_force_sc(msc_ctx *ctx, int rep) {
  State s = sel(ctx, rep);
  if (ctx->contact && rep != MSC_ALL) {
    // a force call outside the step loop: refresh contact rows whose skin
    // the particles have used up since they were built
    launch_check(s, ctx->d_task_out, ctx->d_rebuild, ctx->d_srebuild, nullptr, 0, true, ctx->stream);
    launch_short(s, ctx->d_srebuild, ctx->d_flags, ctx->stream);
  }
  {
    Bracket b(ctx, "wca");
    launch_force_sc(s, !ctx->contact, ctx->mixed, ctx->d_task_out, ctx->d_flags, ctx->stream);
  }
  launch_reduce(s, ctx->d_task_out, 16, 11, 0, (double *)ctx->d_E, kEstride, kEoff_sc, ctx->stream, 1, ctx->d_counts,
                4, 0);
}

int msc_force_sc(msc_ctx *ctx, int rep, double r_cut, double *U, double *W, double PT[9]) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep) return MSC_E_ARG;
  if (!ctx->have_lists) return fail(ctx, MSC_E_STATE, "no neighbour list");
  if (r_cut != ctx->reps[rep].sr_cut) {
    ctx->reps[rep].sr_cut = r_cut;
    int rc = upload_reps(ctx);
    if (rc) return rc;
  }
  HIPC(hipMemsetAsync(ctx->d_flags + 4 * rep, 0, sizeof(int) * 4, ctx->stream));
  do_force_sc(ctx, rep);
  FINISH("msc_force_sc");
  int fl[4];
  if (int rc_ = d2h(ctx, fl, ctx->d_flags + 4 * rep, sizeof fl)) return rc_;
  if (fl[0]) return fail(ctx, fl[0] == 1 ? MSC_E_OVERLAP : MSC_E_DISTANCE, "force_sc abort");
  RepEnergy e;
  int rc = read_energy(ctx, rep, &e);
  if (rc) return rc;
  if (U) *U = e.Usc;
  if (W) *W = e.Wsc;
  if (PT) memcpy(PT, e.PTsc, sizeof e.PTsc);
  return MSC_OK;
}

// real-space launches; per-task sums into task_out
static void dipR_launches(msc_ctx *ctx, int rep, double *task_out) {
  State s = sel(ctx, rep);
  bool all = true;
  for (int r = (rep == MSC_ALL ? 0 : rep); r < (rep == MSC_ALL ? ctx->nrep : rep + 1); ++r) all &= ctx->reps[r].rc <= 0;
  bool none = true;
  for (int r = (rep == MSC_ALL ? 0 : rep); r < (rep == MSC_ALL ? ctx->nrep : rep + 1); ++r) none &= !(ctx->reps[r].rc <= 0);
  {
    Bracket b(ctx, "real_pairs");
    if (all) launch_dipR(s, ctx->mixed, true, ctx->d_tab, task_out, ctx->stream);
    else if (none) launch_dipR(s, ctx->mixed, false, ctx->d_tab, task_out, ctx->stream);
    else {
      for (int r = 0; r < ctx->nrep; ++r) {
        State s1 = sel(ctx, r);
        launch_dipR(s1, ctx->mixed, ctx->reps[r].rc <= 0, ctx->d_tab, task_out, ctx->stream);
      }
    }
  }
}

static void do_dipR(msc_ctx *ctx, int rep) {
  State s = sel(ctx, rep);
  dipR_launches(ctx, rep, ctx->d_task_out);
  launch_reduce(s, ctx->d_task_out, 16, 10, 0, (double *)ctx->d_E, kEstride, kEoff_R, ctx->stream, 1, ctx->d_counts,
                4, 1);
}

int msc_force_dipR(msc_ctx *ctx, int rep, double *U, double PT[9]) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep) return MSC_E_ARG;
  if (ctx->reps[rep].rc > 0 && !ctx->have_lists) return fail(ctx, MSC_E_STATE, "no pair list");
  do_dipR(ctx, rep);
  FINISH("msc_force_dipR");
  RepEnergy e;
  int rc = read_energy(ctx, rep, &e);
  if (rc) return rc;
  if (U) *U = e.UR;
  if (PT) memcpy(PT, e.PTR, sizeof e.PTR);
  return MSC_OK;
}

// Reciprocal launches.  Standalone (m2rows null): phase tables, M2 and the
// per-replica energy sums here.  In the step loop the tables and the c c^T
// rows come from k_drift (m2rows) and the energies go to kblk for k_end.
static int dipK_launches(msc_ctx *ctx, int rep, const double *m2rows, double *kblk, bool join = false,
                         const KickFuse *kf = nullptr) {
  State s = sel(ctx, rep);
  if (ctx->nd > 1) {
    // several slots per particle: intra-particle exclusion by the particles'
    // own sums (FP64 in every context)
    const int rt0 = rep == MSC_ALL ? 0 : ctx->reps[rep].rtask0;
    const int nrt = rep == MSC_ALL ? (int)ctx->rtasks.size() : ctx->reps[rep].nrtask;
    {
      Bracket b(ctx, "recip_force");
      launch_recip_nd(s, rt0, nrt, ctx->maxslot, ctx->d_ph_nd, ctx->d_part_nd, ctx->d_A_nd, ctx->d_slot_out, kblk,
                      ctx->stream);
    }
    if (!kblk) launch_reduce(s, ctx->d_slot_out, 10, 10, 1, (double *)ctx->d_E, kEstride, kEoff_K, ctx->stream);
    return MSC_OK;
  }
  if (!m2rows) {
    launch_recip_tables(s, ctx->mixed, ctx->d_ph, ctx->d_amp, ctx->d_task_out, ctx->stream);
    launch_reduce(s, ctx->d_task_out, 16, 6, 0, ctx->d_M2, 6, 0, ctx->stream);
  }
  // structure factors per group of replicas sharing KZ
  for (int kz : ctx->kz_groups) {
    int rt0 = -1, rt1 = -1;
    for (int r = 0; r < ctx->nrep; ++r) {
      if (ctx->reps[r].KZ != kz || (rep != MSC_ALL && r != rep)) continue;
      if (rt0 < 0) rt0 = ctx->reps[r].rtask0;
      rt1 = ctx->reps[r].rtask0 + ctx->reps[r].nrtask;
    }
    if (rt0 < 0) continue;
    Bracket b(ctx, "recip_struct");
    // groups are contiguous when replicas sharing KZ are adjacent; otherwise
    // launch replica by replica
    bool contiguous = true;
    for (int r = 0; r < ctx->nrep; ++r)
      if (ctx->reps[r].rtask0 >= rt0 && ctx->reps[r].rtask0 < rt1 && ctx->reps[r].KZ != kz) contiguous = false;
    if (contiguous) {
      launch_recip_struct(s, ctx->mixed, kz, rt0, rt1 - rt0, ctx->maxcol, ctx->d_ph, ctx->d_amp, ctx->d_part,
                          ctx->stream);
    } else {
      for (int r = 0; r < ctx->nrep; ++r)
        if (ctx->reps[r].KZ == kz && (rep == MSC_ALL || r == rep))
          launch_recip_struct(s, ctx->mixed, kz, ctx->reps[r].rtask0, ctx->reps[r].nrtask, ctx->maxcol, ctx->d_ph,
                              ctx->d_amp, ctx->d_part, ctx->stream);
    }
  }
  launch_recip_reduce(s, ctx->mixed, ctx->d_part, m2rows ? nullptr : ctx->d_M2, m2rows, ctx->d_A, ctx->d_slot_out,
                      kblk, ctx->maxslot, ctx->d_bop, ctx->d_bexp, ctx->stream, kf);
  if (!kblk) launch_reduce(s, ctx->d_slot_out, 10, 10, 1, (double *)ctx->d_E, kEstride, kEoff_K, ctx->stream);
  // an overlapped step's pair launch adds into F and T as the back
  // projection does: it must have finished first
  if (join) note_stream_op("join of the pair stream", hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
  for (int kz : ctx->kz_groups) {
    for (int r = 0; r < ctx->nrep; ++r) {
      if (ctx->reps[r].KZ != kz || (rep != MSC_ALL && r != rep)) continue;
      // merge consecutive replicas of this KZ into one launch
      int r1 = r;
      while (r1 + 1 < ctx->nrep && ctx->reps[r1 + 1].KZ == kz && rep == MSC_ALL) ++r1;
      State s1 = sel(ctx, r);
      s1.ntask = ctx->reps[r1].task0 + ctx->reps[r1].ntask - ctx->reps[r].task0;
      s1.g_end = ctx->reps[r1].n0 + ctx->reps[r1].N;
      s1.rep_end = r1 + 1;
      {
        Bracket b(ctx, "recip_force");
        launch_recip_force(s1, ctx->mixed, kz, ctx->d_ph, ctx->d_amp, ctx->d_A, ctx->d_bop, ctx->d_bexp,
                           ctx->stream, kf);
      }
      r = r1;
    }
  }
  return MSC_OK;
}

static int do_dipK(msc_ctx *ctx, int rep) { return dipK_launches(ctx, rep, nullptr, nullptr); }

int msc_force_dipK(msc_ctx *ctx, int rep, double *U, double PT[9]) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep) return MSC_E_ARG;
  int rc = do_dipK(ctx, rep);
  if (rc) return rc;
  FINISH("msc_force_dipK");
  RepEnergy e;
  if ((rc = read_energy(ctx, rep, &e))) return rc;
  if (U) *U = e.UK;
  if (PT) memcpy(PT, e.PTK, sizeof e.PTK);
  return MSC_OK;
}

int msc_ext_field(msc_ctx *ctx, int rep, double *U) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep) return MSC_E_ARG;
  State s = sel(ctx, rep);
  launch_ext_field(s, ctx->d_task_out, ctx->stream);
  launch_reduce(s, ctx->d_task_out, 16, 1, 0, (double *)ctx->d_E, kEstride, kEoff_ext, ctx->stream);
  FINISH("msc_ext_field");
  RepEnergy e;
  int rc = read_energy(ctx, rep, &e);
  if (rc) return rc;
  if (U) *U = e.Uext;
  return MSC_OK;
}

int msc_self_ener(msc_ctx *ctx, int rep, double *U) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep || !U) return MSC_E_ARG;
  State s = sel(ctx, rep);
  launch_self_ener(s, ctx->d_task_out, ctx->stream);
  launch_reduce(s, ctx->d_task_out, 16, 1, 0, ctx->d_red, kRedStride, 15, ctx->stream);
  FINISH("msc_self_ener");
  double v[16];
  if (int rc_ = d2h(ctx, v, ctx->d_red + kRedStride * rep, sizeof v)) return rc_;
  *U = v[15];
  return MSC_OK;
}

int msc_zero_forces(msc_ctx *ctx, int rep) {
  SCOPE(ctx);
  if (!ctx || !rep_ok(ctx, rep)) return MSC_E_ARG;
  launch_zero_FT(sel(ctx, rep), ctx->stream);
  FINISH("msc_zero_forces");
  return MSC_OK;
}

// --------------------------------------------------------------------------
// integrator
// --------------------------------------------------------------------------
int msc_moveX(msc_ctx *ctx, int rep, double dt) {
  SCOPE(ctx);
  if (!ctx || !rep_ok(ctx, rep)) return MSC_E_ARG;
  for (int r = 0; r < ctx->nrep; ++r)
    if (rep == MSC_ALL || r == rep) ctx->reps[r].dt = dt;
  int rc = upload_reps(ctx);
  if (rc) return rc;
  State s = sel(ctx, rep);
  launch_moveX(s, ctx->stream);
  launch_dipoles(s, ctx->stream);
  FINISH("msc_moveX");
  return MSC_OK;
}

int msc_moveP(msc_ctx *ctx, int rep, double dt, int last, double K[2], double PT[9]) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep) return MSC_E_ARG;
  ctx->reps[rep].dt = dt;
  int rc = upload_reps(ctx);
  if (rc) return rc;
  State s = sel(ctx, rep);
  launch_moveP(s, last != 0, ctx->d_th, ctx->d_task_out, ctx->stream);
  launch_reduce(s, ctx->d_task_out, 16, 8, 0, ctx->d_red, kRedStride, 0, ctx->stream);
  FINISH("msc_moveP");
  double v[16];
  if (int rc_ = d2h(ctx, v, ctx->d_red + kRedStride * rep, sizeof v)) return rc_;
  if (K) { K[0] = v[0] / 2; K[1] = v[1] / 2; }
  if (PT) {
    double P[9] = {v[2], v[3], v[4], v[3], v[5], v[6], v[4], v[6], v[7]};
    if (!last) for (double &x : P) x = 0;
    memcpy(PT, P, sizeof P);
  }
  return MSC_OK;
}

int msc_renorm(msc_ctx *ctx, int rep) {
  SCOPE(ctx);
  if (!ctx || !rep_ok(ctx, rep)) return MSC_E_ARG;
  launch_renorm(sel(ctx, rep), ctx->stream);
  FINISH("msc_renorm");
  return MSC_OK;
}

// --------------------------------------------------------------------------
// the step loop
// --------------------------------------------------------------------------
static bool any_field(const msc_ctx *ctx) {
  bool field = false;
  for (auto &R : ctx->reps) field |= R.B[0] != 0 || R.B[1] != 0 || R.B[2] != 0;
  return field;
}

static void forces_all(msc_ctx *ctx) {
  State s = sel(ctx, MSC_ALL);
  launch_zero_FT(s, ctx->stream);
  do_force_sc(ctx, MSC_ALL);
  do_dipR(ctx, MSC_ALL);
  do_dipK(ctx, MSC_ALL);
  // ext_field adds exactly nothing without a field: skipped then (U_ext
  // stays 0)
  if (any_field(ctx)) {
    launch_ext_field(s, ctx->d_task_out, ctx->stream);
    launch_reduce(s, ctx->d_task_out, 16, 1, 0, (double *)ctx->d_E, kEstride, kEoff_ext, ctx->stream);
  }
}

int msc_init_run(msc_ctx *ctx) {
  SCOPE(ctx);
  if (!ctx || !ctx->have_state) return MSC_E_ARG;
  ctx->have_lists = false;
  int rc = lists_layout(ctx);
  if (rc) return rc;
  launch_fill_int(ctx->d_flags, 4 * ctx->nrep, 0, ctx->stream);
  if ((rc = lists_build(ctx, ctx->d_one))) return rc;
  forces_all(ctx);
  State s = sel(ctx, MSC_ALL);
  launch_self_ener(s, ctx->d_task_out, ctx->stream);
  launch_reduce(s, ctx->d_task_out, 16, 1, 0, ctx->d_red, kRedStride, 15, ctx->stream);
  FINISH("msc_init_run");
  std::vector<double> red(kRedStride * ctx->nrep);
  std::vector<int> fl(4 * ctx->nrep);
  if (int rc_ = d2h(ctx, red.data(), ctx->d_red, sizeof(double) * red.size())) return rc_;
  if (int rc_ = d2h(ctx, fl.data(), ctx->d_flags, sizeof(int) * fl.size())) return rc_;
  for (int r = 0; r < ctx->nrep; ++r) {
    ctx->reps[r].Uself = red[kRedStride * r + 15] / ctx->reps[r].N;
    if (fl[4 * r + 1] || fl[4 * r + 2] || fl[4 * r + 3])
      return fail(ctx, MSC_E_CAPACITY, "neighbour rows longer than measured");
    if (fl[4 * r + 0]) return fail(ctx, fl[4 * r] == 1 ? MSC_E_OVERLAP : MSC_E_DISTANCE, "initial overlap");
  }
  // rows fitted their regions and no pair overlaps: the lists hold every pair
  ctx->have_lists = true;
  return upload_reps(ctx);
}

int msc_set_graph(msc_ctx *ctx, int on) {
  SCOPE(ctx);
  if (!ctx) return MSC_E_ARG;
  ctx->graph_on = on != 0;
  if (!ctx->graph_on) drop_graph(ctx);
  return MSC_OK;
}

// Particles per structure-factor chunk task: the chunk tasks and each
// replica's partial offsets laid out again, the buffers grown when needed.
static int set_chunk(msc_ctx *ctx, int chunk) {
  if (chunk == ctx->rchunk) return MSC_OK;
  drop_graph(ctx);
  std::vector<Task> rt;
  long part0 = 0;
  for (int r = 0; r < ctx->nrep; ++r) {
    RepDev &R = ctx->reps[r];
    R.rtask0 = (int)rt.size();
    for (int g = 0; g < R.N; g += chunk) rt.push_back({r, R.n0 + g, std::min(chunk, R.N - g), 0});
    R.nrtask = (int)rt.size() - R.rtask0;
    R.part0 = part0;
    part0 += (long)R.nrtask * R.ncol * R.KZ * 6;
  }
  const size_t rs = ctx->mixed ? sizeof(float) : sizeof(double);
  if ((size_t)part0 > ctx->part_cap) {
    if (int rc_ = dalloc(ctx, (unsigned char **)&ctx->d_part, (size_t)part0 * rs)) return rc_;
    if (ctx->nd > 1)
      if (int rc_ = dalloc(ctx, &ctx->d_part_nd, (size_t)part0 / 6 * 10)) return rc_;
    ctx->part_cap = (size_t)part0;
  }
  if (rt.size() > ctx->rtask_cap) {
    if (int rc_ = dalloc(ctx, &ctx->d_rtasks, rt.size())) return rc_;
    ctx->rtask_cap = rt.size();
  }
  ctx->rtasks.swap(rt);
  ctx->part_elems = (size_t)part0;
  ctx->rchunk = chunk;
  if (int rc_ = h2d(ctx, ctx->d_rtasks, ctx->rtasks.data(), sizeof(Task) * ctx->rtasks.size())) return rc_;
  refresh_state_ptrs(ctx);
  return upload_reps(ctx);
}

int msc_get_tuning(msc_ctx *ctx, msc_tuning *t) {
  SCOPE(ctx);
  if (!ctx || !t) return MSC_E_ARG;
  t->graph = ctx->graph_on ? 1 : 0;
  t->fused_pairs = ctx->fused_pairs ? 1 : 0;
  t->pair_waves = ctx->pair_waves;
  t->list_skin = ctx->list_skin;
  t->chunk = ctx->rchunk;
  t->overlap = ctx->overlap ? 1 : 0;
  t->fuse_kick = ctx->fuse_kick ? 1 : 0;
  return MSC_OK;
}

int msc_set_tuning(msc_ctx *ctx, const msc_tuning *t) {
  SCOPE(ctx);
  if (!ctx || !t || (t->pair_waves != 2 && t->pair_waves != 4 && t->pair_waves != 8)) return MSC_E_ARG;
  if (!(t->list_skin >= 0) || !std::isfinite(t->list_skin) || (!ctx->mixed && t->list_skin != 0))
    return fail(ctx, MSC_E_ARG, "list_skin must be >= 0, and 0 for MSC_FP64 contexts");
  if (t->chunk != 64 && t->chunk != 128 && t->chunk != 256 && t->chunk != 512)
    return fail(ctx, MSC_E_ARG, "chunk must be 64, 128, 256 or 512");
  if (int rc_ = set_chunk(ctx, t->chunk)) return rc_;
  drop_graph(ctx);  // the captured step holds the previous launch shapes
  ctx->graph_on = t->graph != 0;
  ctx->fused_pairs = t->fused_pairs != 0;
  ctx->pair_waves = t->pair_waves;
  ctx->overlap = t->overlap != 0;
  ctx->fuse_kick = t->fuse_kick != 0;
  if (t->list_skin != ctx->list_skin) {
    // other rows: msc_init_run (or the autotuner) builds them again
    ctx->list_skin = t->list_skin;
    if (ctx->have_state) {
      for (int r = 0; r < ctx->nrep; ++r) set_radii(ctx, r);
      ctx->have_lists = false;
      return upload_reps(ctx);
    }
  }
  return MSC_OK;
}

// One step of every replica, enqueued on the context stream, in the order of
// v25/main.cpp:559-708 (the fused kernels are described in msc_kernels.hip).
// The step reads its index from d_step, so the same launches serve every
// step of a run and can be captured once as a graph.
static void enqueue_step(msc_ctx *ctx) {
  State s = sel(ctx, MSC_ALL);
  hipStream_t st = ctx->stream;
  const long region = (long)ctx->tasks.size() * 16;
  double *rows = ctx->d_rows;
  const double *gs = ctx->d_run_gs;
  // list maintenance (v25/main.cpp:679-701): displacements reduced by
  // k_drift, decision and cell grid by extra blocks of k_force_sc (or the
  // fused pair launch), k_kick2 (or the back projection that does the last
  // kick) and k_end, the rows by k_lists_step; for replicas that keep their
  // rows these return at once
  const ListStep ls = {ctx->d_rebuild, ctx->contact ? ctx->d_srebuild : nullptr, ctx->d_lctr, ctx->d_dmax,
                       ctx->d_run_obs, ctx->d_step};
  {
    Bracket b(ctx, "integrate");
    launch_drift(s, ctx->mixed, ctx->d_th, gs, ctx->d_step, rows + kRowKick1 * region, ctx->d_ph, ctx->d_amp,
                 rows + kRowDrift * region, ls, st);
  }
  bool cut_all = true;
  for (const RepDev &R : ctx->reps) cut_all &= R.rc > 0;
  const bool fused = ctx->mixed && ctx->contact && cut_all && ctx->fused_pairs;
  // the pair launch on the second stream, concurrent with the structure
  // factors and their reduction (they read the drift's outputs only and
  // write buffers of their own); joined before the back projection
  const bool overlap = fused && ctx->overlap && ctx->nd == 1 && !ctx->timing && !launches_checked();
  if (fused) {
    // WCA over the contact rows and real-space pairs in one launch
    Bracket b(ctx, "real_pairs");
    hipStream_t sp = st;
    if (overlap) {
      note_stream_op("fork of the pair stream", hipEventRecord(ctx->ev_fork, st));
      note_stream_op("fork of the pair stream", hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0));
      sp = ctx->stream2;
    }
    launch_pairs_mixed(s, ctx->pair_waves, ctx->d_tab, rows + kRowSC * region, rows + kRowR * region,
                       ctx->d_flags, ls, sp);
    if (overlap) note_stream_op("join of the pair stream", hipEventRecord(ctx->ev_join, ctx->stream2));
  } else {
    {
      Bracket b(ctx, "wca");
      launch_force_sc(s, !ctx->contact, ctx->mixed, rows + kRowSC * region, ctx->d_flags, st, &ls);
    }
    dipR_launches(ctx, MSC_ALL, rows + kRowR * region);
  }
  // ext_field adds exactly nothing without a field: skipped then
  const bool field = any_field(ctx);
  // mixed contexts without a field: the back projection is the step's last
  // force launch and does the last kick (k_kick2's work) in its epilogue
  const bool fuse_kick = ctx->fuse_kick && ctx->mixed && ctx->nd == 1 && !field;
  const KickFuse kf = {ctx->d_th, gs, ctx->d_step, rows + kRowKick1 * region, ctx->d_kick_ts,
                       rows + kRowKick2 * region, ls};
  // the drift's rows hold m2 as values 2..7 (value k of task t at
  // k * ntask_all + t)
  dipK_launches(ctx, MSC_ALL, rows + kRowDrift * region + 2 * (long)ctx->tasks.size(), ctx->d_kblk, overlap,
                fuse_kick ? &kf : nullptr);
  if (field) launch_ext_field(s, rows + kRowExt * region, st);
  {
    Bracket b(ctx, "integrate");
    if (!fuse_kick)
      launch_kick2(s, ctx->d_th, gs, ctx->d_step, rows + kRowKick1 * region, rows + kRowKick2 * region, ls, st);
    launch_end(s, ctx->d_step, gs, rows, region, field, ctx->d_kblk, ctx->maxslot, ctx->d_endsum, ls, st);
  }
  {
    // the rows, and the thermostat, energies and observables from k_end's sums
    Bracket b(ctx, "nlist");
    const EndArgs ea = {ctx->d_endsum, ctx->d_th, ctx->d_E, ctx->d_Kin, ctx->d_counts, ctx->d_run_obs};
    launch_lists_step(s, ls, ctx->d_flags, ctx->d_step, ea, st);
  }
}

// Device copy of everything a step carries into the next one (positions,
// momenta, orientations, forces, list snapshot, thermostats, energies), for
// rolling back a chunk of steps.
static int state_copy(msc_ctx *ctx, bool restore, char *&save, size_t &cap) {
  const size_t N = ctx->Ntot, nrep = ctx->nrep;
  struct Item { void *p; size_t bytes; } items[] = {
      {ctx->d_x, 3 * N * 8}, {ctx->d_p, 3 * N * 8}, {ctx->d_l, 3 * N * 8}, {ctx->d_ori, 9 * N * 8},
      {ctx->d_F, 3 * N * 8}, {ctx->d_T, 3 * N * 8}, {ctx->d_snap, 3 * N * 8}, {ctx->d_snap_e, 6 * N * 8},
      {ctx->d_snap_u, 3 * N * 8},
      {ctx->d_th, nrep * sizeof(Thermo)}, {ctx->d_E, nrep * sizeof(RepEnergy)}, {ctx->d_Kin, 2 * nrep * 8}};
  size_t total = 0;
  for (auto &it : items) total += (it.bytes + 255) / 256 * 256;
  if (total > cap) {
    if (int rc_ = dalloc(ctx, &save, total)) return rc_;
    cap = total;
  }
  size_t off = 0;
  for (auto &it : items) {
    char *b = save + off;
    if (restore)
      HIPC(hipMemcpyAsync(it.p, b, it.bytes, hipMemcpyDeviceToDevice, ctx->stream));
    else
      HIPC(hipMemcpyAsync(b, it.p, it.bytes, hipMemcpyDeviceToDevice, ctx->stream));
    off += (it.bytes + 255) / 256 * 256;
  }
  return MSC_OK;
}
static int state_copy(msc_ctx *ctx, bool restore) { return state_copy(ctx, restore, ctx->d_save, ctx->save_cap); }

// Steps run in chunks of kRunChunk.  A chunk in which a list rebuild found
// its region too small is rolled back, the regions are enlarged to what the
// rebuild needed, every list is rebuilt from the chunk's first state, and the
// chunk is run again (the redone steps agree with an uninterrupted run up to
// the summation order of the rebuilt lists).
static constexpr int kRunChunk = 100;

// Capture one step (enqueue_step) as a graph on the context stream.
static int capture_step(msc_ctx *ctx) {
  drop_graph(ctx);
  hipGraph_t g = nullptr;
  HIPC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  enqueue_step(ctx);
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (e != hipSuccess) {
    if (g) hipGraphDestroy(g);
    ctx->err = std::string("capturing the step: ") + hipGetErrorString(e);
    g_err = ctx->err;
    return MSC_E_HIP;
  }
  const hipError_t e2 = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e2 != hipSuccess) {
    ctx->graph_exec = nullptr;
    ctx->err = std::string("instantiating the step graph: ") + hipGetErrorString(e2);
    g_err = ctx->err;
    return MSC_E_HIP;
  }
  return MSC_OK;
}

int msc_run(msc_ctx *ctx, int step0, int nsteps, double *obs) {
  SCOPE(ctx);
  if (!ctx || nsteps < 0) return MSC_E_ARG;
  if (!ctx->have_lists) return fail(ctx, MSC_E_STATE, "call msc_init_run first");
  const int nrep = ctx->nrep;
  // thermostat deviates: two uniforms per thermostat per step
  // (moveTher mode 0, v25/functions.cpp:2193-2201)
  std::vector<double> gs((size_t)nsteps * nrep * 4);
  for (int k = 0; k < nsteps; ++k)
    for (int r = 0; r < nrep; ++r)
      for (int i = 0; i < 2; ++i) {
        double a = ctx->thermo_rng[2 * r + i].uniform();
        double b = ctx->thermo_rng[2 * r + i].uniform();
        gauss1(a, b);
        gs[((size_t)k * nrep + r) * 4 + 2 * i] = a;
        gs[((size_t)k * nrep + r) * 4 + 2 * i + 1] = b;
      }
  // run buffers grow with the longest run and are kept by the context
  const size_t nobs = (size_t)nsteps * nrep * MSC_OBS_N;
  if (gs.size() > ctx->run_gs_cap) {
    drop_graph(ctx);
    if (int rc_ = dalloc(ctx, &ctx->d_run_gs, gs.size())) return rc_;
    ctx->run_gs_cap = gs.size();
  }
  if (nobs > ctx->run_obs_cap) {
    drop_graph(ctx);
    if (int rc_ = dalloc(ctx, &ctx->d_run_obs, nobs)) return rc_;
    ctx->run_obs_cap = nobs;
  }
  double *d_gs_all = ctx->d_run_gs, *d_obs_all = ctx->d_run_obs;
  if (!gs.empty())
    if (int rc_ = h2d(ctx, d_gs_all, gs.data(), sizeof(double) * gs.size())) return rc_;
  std::vector<int> fl(4 * nrep);
  int retries = 0;
  for (int k0 = 0; k0 < nsteps;) {
    const int nk = std::min(kRunChunk, nsteps - k0);
    if (int rc_ = state_copy(ctx, false)) return rc_;
    launch_fill_int(ctx->d_flags, 4 * nrep, 0, ctx->stream);
    // both parities of the displacement maxima start at zero (k_drift then
    // clears the next step's before it is used)
    launch_fill_int(reinterpret_cast<int *>(ctx->d_dmax), 8 * kDmaxBuckets * nrep, 0, ctx->stream);
    const int step_pos[2] = {k0, step0};
    if (int rc_ = h2d(ctx, ctx->d_step, step_pos, sizeof step_pos)) return rc_;
    // event brackets and per-launch synchronisation cannot sit in a replayed
    // graph: timed or checked runs launch directly
    const bool graph = ctx->graph_on && !ctx->timing && !launches_checked();
    if (graph && !ctx->graph_exec)
      if (int rc_ = capture_step(ctx)) return rc_;
    for (int k = k0; k < k0 + nk; ++k) {
      if (graph) HIPC(hipGraphLaunch(ctx->graph_exec, ctx->stream));
      else enqueue_step(ctx);
    }
    LAUNCHED("msc_run");
    if (int rc_ = d2h(ctx, fl.data(), ctx->d_flags, sizeof(int) * fl.size())) return rc_;
    // every step advances the device step counter (k_lists_step): a chunk
    // whose kernels did not all execute leaves it behind
    int pos[2] = {-1, -1};
    if (int rc_ = d2h(ctx, pos, ctx->d_step, sizeof pos)) return rc_;
    if (pos[0] != k0 + nk) {
      ctx->have_lists = false;
      return fail(ctx, MSC_E_HIP, "msc_run: the device did not execute the step kernels (step counter " +
                                      std::to_string(pos[0]) + ", expected " + std::to_string(k0 + nk) + ")");
    }
    // an overflowing rebuild empties the replica's rows, so an overlap later in
    // the same chunk may be its consequence: overflow is handled first
    bool overflow = false;
    for (int r = 0; r < nrep; ++r) overflow |= fl[4 * r + 1] != 0 || fl[4 * r + 2] != 0 || fl[4 * r + 3] != 0;
    if (!overflow) {
      std::string m;
      int code = 0;
      for (int r = 0; r < nrep; ++r)
        if (fl[4 * r + 0]) {
          if (!code) code = fl[4 * r] == 1 ? MSC_E_OVERLAP : MSC_E_DISTANCE;
          m += " replica " + std::to_string(r) + (fl[4 * r] == 1 ? " (overlap)" : " (distance)");
        }
      if (code) {
        ctx->have_lists = false;  // the caller restarts from msc_init_run
        return fail(ctx, code, "force_sc abort during steps " + std::to_string(step0 + k0) + ".." +
                                   std::to_string(step0 + k0 + nk - 1) + ":" + m);
      }
    }
    if (overflow) {
      if (++retries > 4) {
        std::string m = "neighbour list regions keep overflowing:";
        for (int r = 0; r < nrep; ++r)
          if (fl[4 * r + 1] || fl[4 * r + 2] || fl[4 * r + 3])
            m += " replica " + std::to_string(r) + " rows need " + std::to_string(fl[4 * r + 2]) + "/" +
                 std::to_string(fl[4 * r + 1]) + "/" + std::to_string(fl[4 * r + 3]) + " of " +
                 std::to_string(ctx->reps[r].nl_w) + "/" + std::to_string(ctx->reps[r].sl_w) + "/" +
                 std::to_string(ctx->reps[r].dl_w) + ";";
        ctx->have_lists = false;  // the rows of this chunk are truncated
        return fail(ctx, MSC_E_CAPACITY, m);
      }
      ++ctx->chunk_repeats;
      // overflowing rows grow to at least 1.5 times their old capacity; the
      // others keep theirs
      std::vector<long> need_nl(nrep), need_sl(nrep), need_dl(nrep);
      for (int r = 0; r < nrep; ++r) {
        const RepDev &R = ctx->reps[r];
        need_nl[r] = fl[4 * r + 2] ? std::max<long>(fl[4 * r + 2], R.nl_w + R.nl_w / 2) : R.nl_w;
        need_sl[r] = fl[4 * r + 1] ? std::max<long>(fl[4 * r + 1], R.sl_w + R.sl_w / 2) : R.sl_w;
        need_dl[r] = fl[4 * r + 3] ? std::max<long>(fl[4 * r + 3], R.dl_w + R.dl_w / 2) : R.dl_w;
      }
      if (int rc_ = state_copy(ctx, true)) return rc_;
      if (int rc_ = lists_layout(ctx, need_nl, need_sl, need_dl)) return rc_;
      launch_fill_int(ctx->d_flags, 4 * nrep, 0, ctx->stream);
      if (int rc_ = lists_build(ctx, ctx->d_one)) return rc_;
      continue;
    }
    k0 += nk;
  }
  FINISH("msc_run");
  if (obs && nobs)
    if (int rc_ = d2h(ctx, obs, d_obs_all, sizeof(double) * nobs)) return rc_;
  if (ctx->timing) settle_timers(ctx);
  return MSC_OK;
}

// The autotuner (no reference counterpart): each candidate launch choice
// runs `steps` steps from the state at the call, after an untimed run of a
// few steps from the same state (graph capture, first-touch costs); the state
// is restored and the lists rebuilt after every run.
int msc_autotune(msc_ctx *ctx, int steps, msc_tuning *best, msc_tuning *tried, double *ms, int ntry,
                 int *ntried) {
  SCOPE(ctx);
  if (!ctx || steps < 1 || ((tried || ms) && ntry < 0)) return MSC_E_ARG;
  if (!ctx->have_lists) return fail(ctx, MSC_E_STATE, "call msc_init_run first");
  bool cut_all = true;
  for (const RepDev &R : ctx->reps) cut_all &= R.rc > 0;
  const bool fusable = ctx->mixed && ctx->contact && cut_all;
  msc_tuning start;
  msc_get_tuning(ctx, &start);
  const std::vector<MT> rng = ctx->thermo_rng;
  if (int rc_ = state_copy(ctx, false, ctx->d_tune_save, ctx->tune_save_cap)) return rc_;
  // the state at the call, with rows built for the present radii
  auto restore = [&]() -> int {
    if (int rc_ = state_copy(ctx, true, ctx->d_tune_save, ctx->tune_save_cap)) return rc_;
    ctx->thermo_rng = rng;
    if (int rc_ = lists_layout(ctx)) return rc_;
    launch_fill_int(ctx->d_flags, 4 * ctx->nrep, 0, ctx->stream);
    if (int rc_ = lists_build(ctx, ctx->d_one)) return rc_;
    ctx->have_lists = true;
    return MSC_OK;
  };
  int n = 0;
  msc_tuning win = start;
  double win_ms = 0;
  auto trial = [&](const msc_tuning &c) -> int {
    if (int rc_ = msc_set_tuning(ctx, &c)) return rc_;
    if (int rc_ = restore()) return rc_;
    if (int rc_ = msc_run(ctx, 0, std::min(steps, 10), nullptr)) return rc_;
    if (int rc_ = restore()) return rc_;
    HIPC(hipStreamSynchronize(ctx->stream));
    const auto t0 = std::chrono::steady_clock::now();
    if (int rc_ = msc_run(ctx, 0, steps, nullptr)) return rc_;  // returns after its stream drained
    const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
    if (tried && n < ntry) tried[n] = c;
    if (ms && n < ntry) ms[n] = t;
    if (n == 0 || t < win_ms) {
      win = c;
      win_ms = t;
    }
    ++n;
    return MSC_OK;
  };
  // launch choices at the present list skin, then (MSC_MIXED) list skins
  // with the fastest launch choice
  for (int g = 0; g < 2; ++g) {
    msc_tuning c = start;
    c.graph = g;
    c.fused_pairs = 0;
    c.pair_waves = 4;
    if (int rc_ = trial(c)) return rc_;
    if (fusable)
      for (int w : {2, 4, 8}) {
        c.fused_pairs = 1;
        c.pair_waves = w;
        if (int rc_ = trial(c)) return rc_;
      }
  }
  if (fusable) {
    // the pair launch on its own stream or not, with the fastest so far
    msc_tuning c = win;
    c.overlap = start.overlap ? 0 : 1;
    if (int rc_ = trial(c)) return rc_;
  }
  if (ctx->mixed) {
    const msc_tuning launch = win;
    for (double skin : {0.5, 0.75, 1.0, 1.5}) {
      if (skin == start.list_skin) continue;
      msc_tuning c = launch;
      c.list_skin = skin;
      if (int rc_ = trial(c)) return rc_;
    }
  }
  {
    const msc_tuning sofar = win;
    for (int chunk : {64, 128, 256, 512}) {
      if (chunk == start.chunk) continue;
      msc_tuning c = sofar;
      c.chunk = chunk;
      if (int rc_ = trial(c)) return rc_;
    }
  }
  if (int rc_ = msc_set_tuning(ctx, &win)) return rc_;
  if (int rc_ = restore()) return rc_;
  FINISH("msc_autotune");
  if (best) *best = win;
  if (ntried) *ntried = n;
  return MSC_OK;
}

// --------------------------------------------------------------------------
// measurement
// --------------------------------------------------------------------------
int msc_set_timing(msc_ctx *ctx, int on) {
  SCOPE(ctx);
  if (!ctx) return MSC_E_ARG;
  ctx->timing = on != 0;
  return MSC_OK;
}
int msc_reset_timing(msc_ctx *ctx) {
  SCOPE(ctx);
  if (!ctx) return MSC_E_ARG;
  settle_timers(ctx);
  ctx->timers.clear();
  return MSC_OK;
}
int msc_kernel_time(msc_ctx *ctx, const char *name, double *ms, long *launches) {
  SCOPE(ctx);
  if (!ctx || !name) return MSC_E_ARG;
  settle_timers(ctx);
  auto it = ctx->timers.find(name);
  if (ms) *ms = it == ctx->timers.end() ? 0 : it->second.ms;
  if (launches) *launches = it == ctx->timers.end() ? 0 : it->second.launches;
  return MSC_OK;
}
int msc_work_counts(msc_ctx *ctx, int rep, long counts[4]) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep || !counts) return MSC_E_ARG;
  double c[4];
  HIPC(hipStreamSynchronize(ctx->stream));
  if (int rc_ = d2h(ctx, c, ctx->d_counts + 4 * rep, sizeof c)) return rc_;
  const RepDev &R = ctx->reps[rep];
  long nk = ((long)(2 * R.kmax + 1) * (2 * R.kmax + 1) * (2 * R.kmax + 1) - 1) / 2;
  counts[0] = (long)c[0];
  counts[1] = (long)c[1];
  counts[2] = nk;
  std::vector<int> cnt(R.N);
  if (int rc_ = d2h(ctx, cnt.data(), ctx->d_nl_cnt + R.n0, sizeof(int) * R.N)) return rc_;
  long t = 0;
  for (int v : cnt) t += v;
  counts[3] = t / 2;
  return MSC_OK;
}

int msc_chunk_repeats(msc_ctx *ctx, long *n) {
  SCOPE(ctx);
  if (!ctx || !n) return MSC_E_ARG;
  *n = ctx->chunk_repeats;
  return MSC_OK;
}

// Eigen-decomposition of a symmetric 3x3 matrix by cyclic Jacobi rotations
// in double precision (the method of the reference's jacobi routine,
// v25/jacobi.c:12-80, without its single-precision storage): a is
// destroyed above the diagonal, d gets the eigenvalues, the columns of v the
// eigenvectors.
static void jacobi3_sym(double a[3][3], double d[3], double v[3][3]) {
  double b[3], z[3];
  for (int p = 0; p < 3; ++p) {
    for (int q = 0; q < 3; ++q) v[p][q] = p == q ? 1.0 : 0.0;
    b[p] = d[p] = a[p][p];
    z[p] = 0.0;
  }
  for (int sweep = 1; sweep <= 50; ++sweep) {
    const double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
    if (off == 0.0) return;
    const double tresh = sweep < 4 ? 0.2 * off / 9 : 0.0;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        const double g = 100.0 * std::fabs(a[p][q]);
        if (sweep > 4 && std::fabs(d[p]) + g == std::fabs(d[p]) && std::fabs(d[q]) + g == std::fabs(d[q])) {
          a[p][q] = 0.0;
          continue;
        }
        if (!(std::fabs(a[p][q]) > tresh)) continue;
        const double h = d[q] - d[p];
        double t;
        if (std::fabs(h) + g == std::fabs(h)) {
          t = a[p][q] / h;
        } else {
          const double theta = 0.5 * h / a[p][q];
          t = 1.0 / (std::fabs(theta) + std::sqrt(1.0 + theta * theta));
          if (theta < 0.0) t = -t;
        }
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c, tau = sn / (1.0 + c), ht = t * a[p][q];
        z[p] -= ht;
        z[q] += ht;
        d[p] -= ht;
        d[q] += ht;
        a[p][q] = 0.0;
        auto rot = [&](double &x, double &y) {
          const double gx = x, hy = y;
          x = gx - sn * (hy + gx * tau);
          y = hy + sn * (gx - hy * tau);
        };
        for (int j = 0; j < p; ++j) rot(a[j][p], a[j][q]);
        for (int j = p + 1; j < q; ++j) rot(a[p][j], a[j][q]);
        for (int j = q + 1; j < 3; ++j) rot(a[p][j], a[q][j]);
        for (int j = 0; j < 3; ++j) rot(v[j][p], v[j][q]);
      }
    for (int p = 0; p < 3; ++p) {
      b[p] += z[p];
      d[p] = b[p];
      z[p] = 0.0;
    }
  }
}

int msc_order_params(msc_ctx *ctx, int rep, double *G1, double *G2, double M[3]) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep || !G1 || !G2 || !M) return MSC_E_ARG;
  if (!ctx->have_state) return fail(ctx, MSC_E_STATE, "no state");
  if (!ctx->d_order)
    if (int rc_ = dalloc(ctx, &ctx->d_order, kOrderSums)) return rc_;
  {
    Bracket b(ctx, "order_params");
    launch_order_sums(sel(ctx, rep), rep, ctx->d_order, ctx->stream);
  }
  FINISH("msc_order_params");
  double o[kOrderSums];
  if (int rc_ = d2h(ctx, o, ctx->d_order, sizeof o)) return rc_;
  const double N = ctx->reps[rep].N;
  // Q = sum_i (3 u_i u_i - I) / (2N) (Qmat_3d)
  const double uu[3][3] = {{o[0], o[1], o[2]}, {o[1], o[3], o[4]}, {o[2], o[4], o[5]}};
  double Q[3][3], d[3], V[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Q[i][j] = (3.0 * uu[i][j] - (i == j ? N : 0.0)) / (2.0 * N);
  jacobi3_sym(Q, d, V);
  // the reference's choice among equal magnitudes (v25/functions.cpp:2738-2748)
  const double e1 = std::fabs(d[0]), e2 = std::fabs(d[1]), e3 = std::fabs(d[2]);
  const int col = (e1 >= e2 && e1 >= e3) ? 0 : (e2 > e1 && e2 >= e3) ? 1 : 2;
  *G2 = col == 0 ? e1 : col == 1 ? e2 : e3;
  const double ve = std::sqrt(V[0][col] * V[0][col] + V[1][col] * V[1][col] + V[2][col] * V[2][col]);
  // sum_i Ve . u_i = Ve . sum_i u_i
  *G1 = std::fabs((V[0][col] * o[6] + V[1][col] * o[7] + V[2][col] * o[8]) / ve) / N;
  for (int a = 0; a < 3; ++a) M[a] = o[9 + a];
  return MSC_OK;
}

int msc_order_params_ref(msc_ctx *ctx, int rep, double *G1, double *G2, double M[3]) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep || !G1 || !G2 || !M) return MSC_E_ARG;
  if (!ctx->have_state) return fail(ctx, MSC_E_STATE, "no state");
  if (ctx->nd > kOrderRefSlots) return fail(ctx, MSC_E_ARG, "at most 4 dipole slots per particle");
  if (!ctx->d_order)
    if (int rc_ = dalloc(ctx, &ctx->d_order, kOrderSums)) return rc_;
  {
    Bracket b(ctx, "order_params");
    launch_order_ref(sel(ctx, rep), rep, ctx->d_order, ctx->stream);
  }
  FINISH("msc_order_params_ref");
  double o[6];
  if (int rc_ = d2h(ctx, o, ctx->d_order, sizeof o)) return rc_;
  if (o[5] < 0) return fail(ctx, MSC_E_ARG, "jacobi: no convergence in 50 sweeps");
  *G1 = o[0];
  *G2 = o[1];
  for (int a = 0; a < 3; ++a) M[a] = o[2 + a];
  return MSC_OK;
}

int msc_ini_state(int N, int nspec, const int *counts, const double L[3], const double *len, const double *dia,
                  const double *mass, double temp, int nmix, unsigned int *mt_state, int *mt_index, double *r,
                  double *p, double *l, double *ori) {
  if (N < 1 || nspec < 1 || nspec > 8 || !counts || !L || !len || !dia || !mass || nmix < 0 || !mt_state ||
      !mt_index || !r || !p || !l || !ori)
    return MSC_E_ARG;
  if (*mt_index < 0 || *mt_index > Mt19937::kN) return MSC_E_ARG;
  long sum = 0;
  for (int n = 0; n < nspec; ++n) {
    if (counts[n] < 1) return MSC_E_ARG;
    sum += counts[n];
  }
  if (sum != N || !(L[0] > 0 && L[1] > 0 && L[2] > 0) || !(temp >= 0)) return MSC_E_ARG;
  // ini_mom's scratch is per species but indexed by the global particle
  // counter (v25/functions.cpp:721, 954): only one species is defined
  if (nspec != 1) return MSC_E_ARG;
  Mt19937 mt;
  std::copy(mt_state, mt_state + Mt19937::kN, mt.mt);
  mt.mti = *mt_index;
  init::Parts P{N, nspec, counts, L, len, dia, r, ori};
  if (init::ini_pos(P)) return MSC_E_OVERLAP;
  std::vector<char> added(N);
  std::vector<int> queue(N);
  for (int k = 0; k < nmix; ++k) init::gca_move(P, mt, added, queue);
  double I[3];
  inertia(len[0], dia[0], mass[0], I);  // the first particle's, as ini_mom reads it
  init::ini_mom(N, temp, mass[0], I, mt, p, l);
  std::copy(mt.mt, mt.mt + Mt19937::kN, mt_state);
  *mt_index = mt.mti;
  return MSC_OK;
}

int msc_set_contact_rows(msc_ctx *ctx, int on) {
  SCOPE(ctx);
  if (!ctx || (on != 0 && on != 1)) return MSC_E_ARG;
  if (ctx->contact != (on == 1)) {
    ctx->contact = on == 1;
    drop_graph(ctx);  // the captured step launches the other WCA kernel
    if (ctx->contact && ctx->have_lists) {
      // the contact rows were last refreshed at a list build: refresh those
      // whose skin the particles have used up since, before any force
      for (int r = 0; r < ctx->nrep; ++r) {
        const State s = sel(ctx, r);
        launch_check(s, ctx->d_task_out, ctx->d_rebuild, ctx->d_srebuild, nullptr, 0, true, ctx->stream);
        launch_short(s, ctx->d_srebuild, ctx->d_flags, ctx->stream);
      }
      FINISH("msc_set_contact_rows");
    }
  }
  return MSC_OK;
}

}  // extern "C"

// Arguments of k_pair_correl for replica rep (checks of msc_pair_correl).
static int correl_args(msc_ctx *ctx, int rep, int nspec, const int *counts, double bar_w, int n_bars,
                       CorrelArgs *out) {
  if (nspec < 1 || nspec > kCorrelMaxSpec) return fail(ctx, MSC_E_ARG, "nspec outside [1, 8]");
  if (!(bar_w > 0) || n_bars < 1) return fail(ctx, MSC_E_ARG, "bar_w and n_bars must be positive");
  if (!ctx->have_state) return fail(ctx, MSC_E_STATE, "no state");
  const RepDev &R = ctx->reps[rep];
  CorrelArgs a{};
  a.n0 = R.n0;
  a.N = R.N;
  a.nspec = nspec;
  a.sets = nspec + nspec * (nspec - 1) / 2;
  long sum = 0;
  for (int m = 0; m < nspec; ++m) {
    if (counts[m] < 0) return fail(ctx, MSC_E_ARG, "negative species count");
    sum += counts[m];
    a.spec_end[m] = (int)std::min(sum, (long)R.N);
  }
  if (sum != R.N) return fail(ctx, MSC_E_ARG, "species counts do not add up to the replica's particles");
  for (int c = 0; c < 3; ++c) {
    a.L[c] = R.L[c];
    a.invL[c] = 1.0 / R.L[c];
  }
  a.bar_w = bar_w;
  a.inv_bw = 1.0 / bar_w;
  a.n_bars = n_bars;
  a.nbr = kCorrelLds / (4 * a.sets);
  *out = a;
  return MSC_OK;
}

// dst[i] += device src[i], staged through the pinned buffer (no allocation)
static int d2h_add(msc_ctx *ctx, double *dst, const double *src, size_t n) {
  LAUNCHED("a device read-back");
  const size_t per = kPinBytes / sizeof(double);
  const double *pin = (const double *)ctx->pin;
  for (size_t off = 0; off < n; off += per) {
    const size_t m = std::min(per, n - off);
    HIPC(hipMemcpyAsync(ctx->pin, src + off, m * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < m; ++i) dst[off + i] += pin[i];
  }
  return MSC_OK;
}

// Scratch of msc_pair_correl_noalloc: element 0 holds the dropped-pair
// counter (a 64-bit integer), the histograms start at element
// kNoAllocHistOff (256 bytes in).  One copy per device, shared by every call
// on that device; its address is resolved once per device (noalloc_scratch).
static constexpr size_t kNoAllocHistOff = 32;
static __device__ __align__(256) double scratch[kNoAllocHistOff + MSC_NOALLOC_CORREL_MAX];

// Device address of `scratch` on the current device, resolved by
// hipGetSymbolAddress at the first call on each device and kept.  Host code
// uses only this address: the array's own name is a host-side stand-in.
static int noalloc_scratch(msc_ctx *ctx, double **out) {
  static std::mutex mu;
  static std::map<int, double *> addr;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(ctx->device);
  if (it != addr.end()) {
    *out = it->second;
    return MSC_OK;
  }
  HIPC(hipSetDevice(ctx->device));
  void *p = nullptr;
  HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(ctx, MSC_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
  addr[ctx->device] = (double *)p;
  *out = (double *)p;
  return MSC_OK;
}

extern "C" {

int msc_pair_correl(msc_ctx *ctx, int rep, int nspec, const int *counts, double bar_w, int n_bars, double *hist,
                    long *dropped) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep || !counts || !hist) return MSC_E_ARG;
  CorrelArgs a;
  if (int rc_ = correl_args(ctx, rep, nspec, counts, bar_w, n_bars, &a)) return rc_;
  const size_t nh = (size_t)a.sets * 4 * n_bars;
  if (nh > ctx->corr_cap) {
    if (int rc_ = dalloc(ctx, &ctx->d_corr, nh)) return rc_;
    ctx->corr_cap = nh;
  }
  if (!ctx->d_corr_drop)
    if (int rc_ = dalloc(ctx, &ctx->d_corr_drop, 1)) return rc_;
  HIPC(hipMemsetAsync(ctx->d_corr, 0, nh * sizeof(double), ctx->stream));
  HIPC(hipMemsetAsync(ctx->d_corr_drop, 0, sizeof(unsigned long long), ctx->stream));
  if (a.N >= 2) {
    Bracket b(ctx, "pair_correl");
    launch_pair_correl(ctx->d_x, ctx->d_ori, ctx->Ntot, a, ctx->d_corr, ctx->d_corr_drop, ctx->stream);
  }
  FINISH("msc_pair_correl");
  unsigned long long dr = 0;
  if (int rc_ = d2h_add(ctx, hist, ctx->d_corr, nh)) return rc_;
  if (int rc_ = d2h(ctx, &dr, ctx->d_corr_drop, sizeof dr)) return rc_;
  if (dropped) *dropped = (long)dr;
  return MSC_OK;
}

int msc_pair_correl_noalloc(msc_ctx *ctx, int rep, int nspec, const int *counts, double bar_w, int n_bars,
                            double *hist, long *dropped) {
  SCOPE(ctx);
  if (!ctx || rep < 0 || rep >= ctx->nrep || !counts || !hist) return MSC_E_ARG;
  CorrelArgs a;
  if (int rc_ = correl_args(ctx, rep, nspec, counts, bar_w, n_bars, &a)) return rc_;
  const size_t nh = (size_t)a.sets * 4 * n_bars;
  if (nh > (size_t)MSC_NOALLOC_CORREL_MAX)
    return fail(ctx, MSC_E_CAPACITY, "msc_pair_correl_noalloc: sets * 4 * n_bars = " + std::to_string(nh) +
                                         " exceeds MSC_NOALLOC_CORREL_MAX = " +
                                         std::to_string(MSC_NOALLOC_CORREL_MAX));
  double *base = nullptr;
  if (int rc_ = noalloc_scratch(ctx, &base)) return rc_;
  // the previous call's contents stay in the array: clear the counter and the
  // histograms this call uses with a fill kernel
  launch_fill_int((int *)base, (int)(2 * (kNoAllocHistOff + nh)), 0, ctx->stream);
  unsigned long long *d_drop = (unsigned long long *)base;
  double *d_hist = base + kNoAllocHistOff;
  if (a.N >= 2) launch_pair_correl(ctx->d_x, ctx->d_ori, ctx->Ntot, a, d_hist, d_drop, ctx->stream);
  FINISH("msc_pair_correl_noalloc");
  unsigned long long dr = 0;
  if (int rc_ = d2h_add(ctx, hist, d_hist, nh)) return rc_;
  if (int rc_ = d2h(ctx, &dr, d_drop, sizeof dr)) return rc_;
  if (dropped) *dropped = (long)dr;
  return MSC_OK;
}

int msc_cluster_labels(msc_ctx *ctx, int rep, double min_s, unsigned int *mt_state, int *mt_index, int *labels,
                       int *nclust, long *n_asym) {
  return msc_cluster_labels_dd(ctx, rep, min_s, INFINITY, mt_state, mt_index, labels, nclust, n_asym);
}

int msc_cluster_labels_dd(msc_ctx *ctx, int rep, double min_s, double u_cut, unsigned int *mt_state, int *mt_index,
                          int *labels, int *nclust, long *n_asym) {
  SCOPE(ctx);
  if (std::isnan(u_cut) || u_cut == -INFINITY) return ctx ? fail(ctx, MSC_E_ARG, "u_cut must be a number or +inf") : MSC_E_ARG;
  if (!ctx || rep < 0 || rep >= ctx->nrep || !(min_s >= 0) || !mt_state || !mt_index || !labels || !nclust)
    return MSC_E_ARG;
  if (*mt_index < 0 || *mt_index > Mt19937::kN) return fail(ctx, MSC_E_ARG, "mt_index outside [0, 624]");
  if (!ctx->have_state) return fail(ctx, MSC_E_STATE, "no state");
  if (!ctx->have_lists) return fail(ctx, MSC_E_STATE, "no neighbour list (msc_neigh_list or msc_init_run)");
  const RepDev &R = ctx->reps[rep];
  const double reach_max = R.diamax + min_s;
  if (!(reach_max < R.r_list_max))
    return fail(ctx, MSC_E_ARG, "min_s + largest diameter must stay below the list radius r_list_max");
  std::vector<int> root(R.N);
  int fl[2];
  for (int attempt = 0;; ++attempt) {
    const State s = sel(ctx, rep);
    {
      Bracket b(ctx, "clusters");
      launch_clusters(s, min_s, reach_max, u_cut, ctx->d_clu, ctx->d_clu_flag, ctx->stream);
    }
    FINISH("msc_cluster_labels");
    if (int rc_ = d2h(ctx, root.data(), ctx->d_clu + R.n0, sizeof(int) * R.N)) return rc_;
    if (int rc_ = d2h(ctx, fl, ctx->d_clu_flag, sizeof fl)) return rc_;
    if (!fl[0]) break;
    if (!R.own_lists || attempt > 0)
      return fail(ctx, MSC_E_STATE, "particles moved too far since the neighbour list was built to find every contact");
    // the context's own rows are short: built again from the present state
    // (a run continues from them as after any rebuild)
    if (int rc_ = lists_layout(ctx)) return rc_;
    launch_fill_int(ctx->d_flags, 4 * ctx->nrep, 0, ctx->stream);
    if (int rc_ = lists_build(ctx, ctx->d_one)) return rc_;
  }
  for (int &r : root) r -= R.n0;
  Mt19937 mt;
  std::copy(mt_state, mt_state + Mt19937::kN, mt.mt);
  mt.mti = *mt_index;
  *nclust = order_clusters(root, mt, labels);
  std::copy(mt.mt, mt.mt + Mt19937::kN, mt_state);
  *mt_index = mt.mti;
  if (n_asym) *n_asym = fl[1];
  return MSC_OK;
}

}  // extern "C"
