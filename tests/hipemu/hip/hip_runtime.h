// CPU emulation of the subset of the HIP runtime the library uses -- TEST
// INFRASTRUCTURE for bounds checking under AddressSanitizer on a machine
// without a GPU.  Blocks run one after another; the threads of a block run
// as OS threads, with real barriers for __syncthreads and a per-wavefront
// exchange for __shfl_xor.  Device memory is host heap memory.
#pragma once
#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))
#define __builtin_amdgcn_rsqf(x) (1.0f / std::sqrt((float)(x)))
#define __builtin_amdgcn_readfirstlane(x) (x)

struct double2 { double x, y; };
struct float2 { float x, y; };
struct float4 { float x, y, z, w; };
struct int2 { int x, y; };
struct int4 { int x, y, z, w; };
inline double __longlong_as_double(long long v) {
  double d;
  std::memcpy(&d, &v, sizeof d);
  return d;
}
inline long long __double_as_longlong(double d) {
  long long v;
  std::memcpy(&v, &d, sizeof v);
  return v;
}
inline float __int_as_float(int v) {
  float f;
  std::memcpy(&f, &v, sizeof f);
  return f;
}
inline int __float_as_int(float f) {
  int v;
  std::memcpy(&v, &f, sizeof v);
  return v;
}

using std::floor;
using std::sqrt;
using std::fabs;
using std::isinf;
using std::isnan;
using std::isfinite;
using std::fmax;
using std::fma;
using std::exp;
using std::min;
using std::max;

struct dim3 {
  unsigned x = 1, y = 1, z = 1;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace hipemu {
struct Block {
  std::unique_ptr<std::barrier<>> bar;
  std::vector<std::unique_ptr<std::barrier<>>> wbar;
  std::vector<double> xch;  // [thread] exchange slots for shuffles
  std::vector<float> mxa, mxb;  // [thread][8] operand exchange of matrix instructions
};
inline Block *&cur_block() { static Block *b = nullptr; return b; }
inline thread_local dim3 t_threadIdx, t_blockIdx, t_blockDim, t_gridDim;
inline thread_local bool t_done = false;
}  // namespace hipemu

#define threadIdx (hipemu::t_threadIdx)
#define blockIdx (hipemu::t_blockIdx)
#define blockDim (hipemu::t_blockDim)
#define gridDim (hipemu::t_gridDim)

inline void __syncthreads() { hipemu::cur_block()->bar->arrive_and_wait(); }

template <typename T>
inline T __shfl_xor(T v, int o) {
  hipemu::Block *b = hipemu::cur_block();
  unsigned t = hipemu::t_threadIdx.x, w = t / 64, lane = t % 64;
  double *x = b->xch.data() + w * 64;
  x[lane] = (double)v;
  b->wbar[w]->arrive_and_wait();
  double r = x[lane ^ (unsigned)o];
  b->wbar[w]->arrive_and_wait();
  return (T)r;
}

// value of lane (lane - d), or the caller's own where that is below 0
template <typename T>
inline T __shfl_up(T v, unsigned d) {
  hipemu::Block *b = hipemu::cur_block();
  unsigned t = hipemu::t_threadIdx.x, w = t / 64, lane = t % 64;
  double *x = b->xch.data() + w * 64;
  x[lane] = (double)v;
  b->wbar[w]->arrive_and_wait();
  double r = lane >= d ? x[lane - d] : (double)v;
  b->wbar[w]->arrive_and_wait();
  return (T)r;
}
// value of lane src
template <typename T>
inline T __shfl(T v, int src) {
  hipemu::Block *b = hipemu::cur_block();
  unsigned t = hipemu::t_threadIdx.x, w = t / 64, lane = t % 64;
  double *x = b->xch.data() + w * 64;
  x[lane] = (double)v;
  b->wbar[w]->arrive_and_wait();
  double r = x[src & 63];
  b->wbar[w]->arrive_and_wait();
  return (T)r;
}

// DPP moves (quad_perm, row_mirror, row_half_mirror) and v_readlane
namespace hipemu {
inline int update_dpp(int src, int ctrl) {
  Block *b = cur_block();
  unsigned t = t_threadIdx.x, w = t / 64, lane = t % 64;
  double *x = b->xch.data() + w * 64;
  x[lane] = (double)src;
  b->wbar[w]->arrive_and_wait();
  unsigned from = lane;
  if (ctrl < 0x100) from = (lane & ~3u) | ((unsigned)(ctrl >> (2 * (lane & 3))) & 3u);
  else if (ctrl == 0x140) from = (lane & ~15u) | (15u - (lane & 15u));
  else if (ctrl == 0x141) from = (lane & ~7u) | (7u - (lane & 7u));
  else std::abort();
  int r = (int)x[from];
  b->wbar[w]->arrive_and_wait();
  return r;
}
inline int readlane(int v, int l) {
  Block *b = cur_block();
  unsigned t = t_threadIdx.x, w = t / 64, lane = t % 64;
  double *x = b->xch.data() + w * 64;
  x[lane] = (double)v;
  b->wbar[w]->arrive_and_wait();
  int r = (int)x[l & 63];
  b->wbar[w]->arrive_and_wait();
  return r;
}
}  // namespace hipemu
#define __builtin_amdgcn_update_dpp(old, src, ctrl, rm, bm, bc) hipemu::update_dpp((src), (ctrl))
#define __builtin_amdgcn_readlane(v, l) hipemu::readlane((v), (l))

// all 64 lanes of a wave take part (the kernels only call these from
// wave-uniform control flow)
inline unsigned long long __ballot(int pred) {
  hipemu::Block *b = hipemu::cur_block();
  unsigned t = hipemu::t_threadIdx.x, w = t / 64, lane = t % 64;
  double *x = b->xch.data() + w * 64;
  x[lane] = pred ? 1.0 : 0.0;
  b->wbar[w]->arrive_and_wait();
  unsigned long long m = 0;
  for (unsigned i = 0; i < 64; ++i)
    if (x[i] != 0.0) m |= 1ull << i;
  b->wbar[w]->arrive_and_wait();
  return m;
}
inline int __popcll(unsigned long long m) { return __builtin_popcountll(m); }
inline void hipemu_wave_barrier() {
  hipemu::Block *b = hipemu::cur_block();
  b->wbar[hipemu::t_threadIdx.x / 64]->arrive_and_wait();
}
#define __builtin_amdgcn_wave_barrier() hipemu_wave_barrier()

#ifdef __clang__
// v_mfma_f32_16x16x16f16 on the lanes of one wave: A[i][k] in lane
// i + 16 (k/4), element k % 4; B[k][n] in lane n + 16 (k/4); D[i][n] in lane
// n + 16 (i/4), element i % 4 (the products of two halves are exact in float)
typedef _Float16 hipemu_half4 __attribute__((ext_vector_type(4)));
typedef float hipemu_float4 __attribute__((ext_vector_type(4)));
inline hipemu_float4 hipemu_mfma16(hipemu_half4 a, hipemu_half4 b, hipemu_float4 c) {
  hipemu::Block *bk = hipemu::cur_block();
  const unsigned t = hipemu::t_threadIdx.x, w = t / 64, lane = t % 64;
  float *xa = bk->mxa.data() + w * 512, *xb = bk->mxb.data() + w * 512;
  for (int j = 0; j < 4; ++j) {
    xa[lane * 4 + j] = (float)a[j];
    xb[lane * 4 + j] = (float)b[j];
  }
  bk->wbar[w]->arrive_and_wait();
  hipemu_float4 d;
  for (int j = 0; j < 4; ++j) {
    const int i = 4 * (lane / 16) + j, n = lane % 16;
    float acc = c[j];
    for (int k = 0; k < 16; ++k) acc += xa[(i + 16 * (k / 4)) * 4 + k % 4] * xb[(n + 16 * (k / 4)) * 4 + k % 4];
    d[j] = acc;
  }
  bk->wbar[w]->arrive_and_wait();
  return d;
}
#define __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, cbsz, abid, blgp) hipemu_mfma16(a, b, c)
// v_mfma_f32_16x16x32_f16: A[i][k] in lane i + 16 (k/8), element k % 8; B[k][n]
// in lane n + 16 (k/8); D as above.  (The hardware's inner-index order may
// differ; a product does not depend on it when A and B share it.)
typedef _Float16 hipemu_half8 __attribute__((ext_vector_type(8)));
inline hipemu_float4 hipemu_mfma32(hipemu_half8 a, hipemu_half8 b, hipemu_float4 c) {
  hipemu::Block *bk = hipemu::cur_block();
  const unsigned t = hipemu::t_threadIdx.x, w = t / 64, lane = t % 64;
  float *xa = bk->mxa.data() + w * 512, *xb = bk->mxb.data() + w * 512;
  for (int j = 0; j < 8; ++j) {
    xa[lane * 8 + j] = (float)a[j];
    xb[lane * 8 + j] = (float)b[j];
  }
  bk->wbar[w]->arrive_and_wait();
  hipemu_float4 d;
  for (int j = 0; j < 4; ++j) {
    const int i = 4 * (lane / 16) + j, n = lane % 16;
    float acc = c[j];
    for (int k = 0; k < 32; ++k) acc += xa[(i + 16 * (k / 8)) * 8 + k % 8] * xb[(n + 16 * (k / 8)) * 8 + k % 8];
    d[j] = acc;
  }
  bk->wbar[w]->arrive_and_wait();
  return d;
}
#define __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, cbsz, abid, blgp) hipemu_mfma32(a, b, c)
#endif
#define __builtin_amdgcn_fence(...) ((void)0)
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }

inline int atomicAdd(int *p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned atomicAdd(unsigned *p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicAdd(unsigned long long *p, unsigned long long v) {
  return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline double atomicAdd(double *p, double v) {
  double o, n;
  __atomic_load(p, &o, __ATOMIC_SEQ_CST);
  do {
    n = o + v;
  } while (!__atomic_compare_exchange(p, &o, &n, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST));
  return o;
}
#ifdef __clang__
// clang provides the scoped atomic builtins on the host as well
#ifndef __HIP_MEMORY_SCOPE_AGENT
#define __HIP_MEMORY_SCOPE_AGENT 3
#endif
#else
#define __HIP_MEMORY_SCOPE_AGENT 3
template <class T> inline T __hip_atomic_load(T *p, int, int) { return __atomic_load_n(p, __ATOMIC_SEQ_CST); }
template <class T> inline void __hip_atomic_store(T *p, T v, int, int) { __atomic_store_n(p, v, __ATOMIC_SEQ_CST); }
template <class T> inline bool __hip_atomic_compare_exchange_strong(T *p, T *e, T d, int, int, int) {
  return __atomic_compare_exchange_n(p, e, d, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
}
#endif
inline int atomicMax(int *p, int v) {
  int old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return old;
}
inline unsigned long long atomicMax(unsigned long long *p, unsigned long long v) {
  unsigned long long old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return old;
}
inline int atomicCAS(int *p, int cmp, int val) {
  __atomic_compare_exchange_n(p, &cmp, val, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
  return cmp;
}

typedef int hipError_t;
typedef void *hipStream_t;
typedef void *hipEvent_t;
// graphs: launches made during a capture are recorded and replayed in order
struct hipemu_graph {
  std::vector<std::function<void()>> ops;
};
typedef hipemu_graph *hipGraph_t;
typedef hipemu_graph *hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
namespace hipemu {
inline hipemu_graph *&capturing() { static hipemu_graph *g = nullptr; return g; }
}  // namespace hipemu
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost };

inline const char *hipGetErrorString(hipError_t e) { return e ? "emulated error" : "no error"; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipStreamCreate(hipStream_t *s) { *s = (void *)0x1; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  hipemu::capturing() = new hipemu_graph();
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
  *g = hipemu::capturing();
  hipemu::capturing() = nullptr;
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, char *, size_t) {
  *e = new hipemu_graph(*g);
  return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto &op : e->ops) op();
  return hipSuccess;
}
inline hipError_t hipMalloc(void **p, size_t n) {
  *p = std::malloc(n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
enum { hipHostMallocDefault = 0 };
inline hipError_t hipHostMalloc(void **p, size_t n, unsigned) {
  *p = std::malloc(n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipHostFree(void *p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t) { std::memset(p, v, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = nullptr; return hipSuccess; }
// launches run in order on the host: a second stream and its events are no-ops
#define hipEventDisableTiming 0x2
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { *e = nullptr; return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0; return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
// a device symbol is host memory here: its address is the symbol's own
#define HIP_SYMBOL(s) (&(s))
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) {
  *p = const_cast<void *>(sym);
  return hipSuccess;
}
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }

namespace hipemu {
template <typename F, typename... A>
void launch(F f, dim3 grid, dim3 block, size_t shm, hipStream_t st, A... args) {
  if (capturing()) {
    hipemu_graph *g = capturing();
    capturing() = nullptr;  // the recorded launch runs when the graph is replayed
    g->ops.push_back([=]() { launch(f, grid, block, shm, st, args...); });
    capturing() = g;
    return;
  }
  unsigned nt = block.x * block.y * block.z;
  if (nt == 0 || grid.x == 0 || grid.y == 0 || grid.z == 0) {
    std::fprintf(stderr, "hipemu: empty launch\n");
    std::abort();
  }
  if (nt % 64 != 0 && nt > 64) {
    std::fprintf(stderr, "hipemu: block size %u not a multiple of 64\n", nt);
    std::abort();
  }
  for (unsigned bz = 0; bz < grid.z; ++bz)
    for (unsigned by = 0; by < grid.y; ++by)
      for (unsigned bx = 0; bx < grid.x; ++bx) {
        Block B;
        B.bar = std::make_unique<std::barrier<>>(nt);
        unsigned nw = (nt + 63) / 64;
        for (unsigned w = 0; w < nw; ++w) B.wbar.push_back(std::make_unique<std::barrier<>>(std::min(64u, nt - 64 * w)));
        B.xch.assign(nw * 64, 0.0);
        B.mxa.assign(nw * 64 * 8, 0.f);
        B.mxb.assign(nw * 64 * 8, 0.f);
        cur_block() = &B;
        std::vector<std::thread> th;
        th.reserve(nt);
        for (unsigned t = 0; t < nt; ++t)
          th.emplace_back([&, t]() {
            t_threadIdx = dim3(t, 0, 0);
            t_blockIdx = dim3(bx, by, bz);
            t_blockDim = block;
            t_gridDim = grid;
            f(args...);
            // a thread that leaves drops out of every barrier it could meet
            B.bar->arrive_and_drop();
            B.wbar[t / 64]->arrive_and_drop();
          });
        for (auto &x : th) x.join();
        cur_block() = nullptr;
      }
}
}  // namespace hipemu

#define hipLaunchKernelGGL(K, G, B, S, ST, ...) hipemu::launch(K, dim3(G), dim3(B), S, ST, __VA_ARGS__)
