// cps_common.hpp -- shared device/host helpers for the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <cstdio>
#include <string>

#include "../../include/coldplate_hip.h"

namespace cps {

// XCD-aware tile order for tiled kernels whose blocks read halos: the hardware hands block b to
// XCD b mod 8, so the logical tile L = (b mod 8) (nb / 8) + b / 8 gives each
// XCD a contiguous band of tiles, whose shared halo lines then meet in that
// XCD's L2.  Block partials are stored at L, so every reduction keeps its
// order (and its result).
__device__ inline int xcd_tile_of(int b, int nb) { return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3); }
__device__ inline int xcd_tile() {
  return xcd_tile_of((int)(blockIdx.x + blockIdx.y * gridDim.x), (int)(gridDim.x * gridDim.y));
}

// The same XCD-aware order for any block count of a 1-D grid: XCD x (which
// receives blocks x, x + 8, ...) takes the contiguous band that starts after
// the bands of XCDs 0 .. x-1 (a bijection of the block indices).
__device__ inline int xcd_band() {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

// Streaming store for results the next kernel reads (possibly on another
// XCD): marked non-temporal, so the line is not kept dirty in this XCD's L2
// for the end-of-kernel write-back.
template <class T>
__device__ __forceinline__ void st_nt(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}

// Mask decisions in the reference's float arithmetic on its float gamma
// (a MatrixXf, SIMPLE.h:479-520), so the FP64 path takes the reference's
// discrete choices bit for bit:
//   face_fluid   gammaAtU/V >= 0.999f, the average formed in float
//                (SIMPLE.h:740-778, boundaries.cpp:148-206);
//   inlet_fluid  gamma(cellRow, 0) > 0.5f (boundaries.cpp:82-100);
//   pure_fluid   !(gamma < 1.0f - tol) (postprocessing.cpp:155-160, 268).
__host__ __device__ inline bool face_fluid(double a, double b) { return 0.5f * ((float)a + (float)b) >= 0.999f; }
__host__ __device__ inline bool inlet_fluid(double g) { return (float)g > 0.5f; }
__host__ __device__ inline bool pure_fluid(double g, double tol) { return !((float)g < 1.0f - (float)tol); }

// ---------------------------------------------------------------------------
// error plumbing: a failed call records its message (thread-local) and the
// C entry point returns the status.
void set_error(const std::string& msg);
void clear_error();

struct Error {
  int code;
  std::string msg;
};

#define CPS_HIP(expr)                                                                           \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      throw ::cps::Error{CPS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)};        \
  } while (0)

#define CPS_CHECK(cond, code, msg)                  \
  do {                                              \
    if (!(cond)) throw ::cps::Error{(code), (msg)}; \
  } while (0)

// Runs fn() and converts exceptions into a status code.
template <class F>
int guard(F&& fn) {
  try {
    clear_error();
    fn();
    return CPS_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::exception& e) {
    set_error(e.what());
    return CPS_ERR_STATE;
  }
}

// ---------------------------------------------------------------------------
// Class-batched launches: the fields of one size class keep their buffers
// in one allocation, field k's copy of every buffer sitting k * zs bytes
// after field 0's.  A kernel launched with gridDim.z = fields gets field 0's
// pointers and zs, and block (x, y, z) shifts them to field z at entry
// (nullptr, a "not used" marker, stays nullptr).  gridDim.z = 1 shifts by 0.
__device__ __forceinline__ size_t zoff(size_t zs) { return (size_t)blockIdx.z * zs; }
// The shift is a byte offset on the pointer itself (no round trip through an
// integer), so the compiler keeps the global address space it inferred for
// the kernel argument and emits global, not flat, memory instructions.
template <class T>
__device__ __forceinline__ T* zsh(T* p, size_t off) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return p ? reinterpret_cast<T*>(reinterpret_cast<B*>(p) + off) : p;
}

// Bump allocator over one region: while a thread has an arena installed
// (ArenaScope), DevBuf::alloc takes its memory from it instead of hipMalloc,
// so two solvers built the same way under arenas at base and base + zs hold
// every buffer zs bytes apart.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, off = 0;
};
inline thread_local Arena* t_arena = nullptr;
struct ArenaScope {
  Arena* prev;
  explicit ArenaScope(Arena* a) : prev(t_arena) { t_arena = a; }
  ~ArenaScope() { t_arena = prev; }
  ArenaScope(const ArenaScope&) = delete;
  ArenaScope& operator=(const ArenaScope&) = delete;
};

// device buffer (RAII)
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  bool owned = false;  // from hipMalloc (else from an arena, freed with it)
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    if (Arena* a = t_arena) {
      const size_t off = (a->off + 255) & ~(size_t)255;
      if (off + count * sizeof(T) > a->cap)
        throw Error{CPS_ERR_STATE, "batch arena exhausted (" + std::to_string(a->cap) + " bytes)"};
      p = reinterpret_cast<T*>(a->base + off);
      a->off = off + count * sizeof(T);
      owned = false;
    } else {
      CPS_HIP(hipMalloc(&p, count * sizeof(T)));
      owned = true;
    }
    n = count;
  }
  void release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    n = 0;
    owned = false;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// ---------------------------------------------------------------------------
// 64-wide wavefront + block reductions (deterministic order)
constexpr int kWave = 64;

template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_down(v, o, kWave);
  return v;
}
template <class T>
__device__ inline T wave_max(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    T w = __shfl_down(v, o, kWave);
    v = v > w ? v : w;
  }
  return v;
}

// Sum over the block (NT threads, flattened); result valid in thread 0.
template <int NT, class T>
__device__ inline T block_sum(T v, T* sh) {
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  v = wave_sum(v);
  if ((t & (kWave - 1)) == 0) sh[t / kWave] = v;
  __syncthreads();
  T r = T(0);
  if (t == 0) {
#pragma unroll
    for (int k = 0; k < NT / kWave; ++k) r += sh[k];
  }
  __syncthreads();
  return r;
}
template <int NT, class T>
__device__ inline T block_max(T v, T* sh) {
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  v = wave_max(v);
  if ((t & (kWave - 1)) == 0) sh[t / kWave] = v;
  __syncthreads();
  T r = T(0);
  if (t == 0) {
#pragma unroll
    for (int k = 0; k < NT / kWave; ++k) r = r > sh[k] ? r : sh[k];
  }
  __syncthreads();
  return r;
}

// Sum over the block, result in thread 0, without the trailing barrier of
// block_sum: for the last use of sh before a barrier the caller places.
template <int NT, class T>
__device__ inline T block_sum_t0(T v, T* sh) {
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  v = wave_sum(v);
  if ((t & (kWave - 1)) == 0) sh[t / kWave] = v;
  __syncthreads();
  T r = T(0);
  if (t == 0) {
#pragma unroll
    for (int k = 0; k < NT / kWave; ++k) r += sh[k];
  }
  return r;
}
// Two sums over the block in block_sum's order (wave sums, then the
// wavefronts in order), delivered to every thread after one barrier.  sh2
// holds 2 * NT / kWave entries and must not be written again before a later
// barrier.  Equal, bit for bit, to parts_total / block_sum of each value.
template <int NT>
__device__ inline void block_sum2_all(double& a, double& b, double* sh2) {
  constexpr int NW = NT / kWave;
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  a = wave_sum(a);
  b = wave_sum(b);
  if ((t & (kWave - 1)) == 0) {
    sh2[t / kWave] = a;
    sh2[NW + t / kWave] = b;
  }
  __syncthreads();
  double ra = 0.0, rb = 0.0;
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    ra += sh2[k];
    rb += sh2[NW + k];
  }
  a = ra;
  b = rb;
}

// M sums over the block in block_sum's order (wave sums, then the
// wavefronts in order), delivered to every thread after one barrier.  sh
// holds M * NT / kWave entries and must not be written again before a later
// barrier.  Equal, bit for bit, to parts_total / block_sum of each value.
template <int NT, int M>
__device__ inline void block_sumn_all(double (&v)[M], double* sh) {
  constexpr int NW = NT / kWave;
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
#pragma unroll
  for (int m = 0; m < M; ++m) v[m] = wave_sum(v[m]);
  if ((t & (kWave - 1)) == 0) {
#pragma unroll
    for (int m = 0; m < M; ++m) sh[m * NW + t / kWave] = v[m];
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < M; ++m) {
    double r = 0.0;
#pragma unroll
    for (int k = 0; k < NW; ++k) r += sh[m * NW + k];
    v[m] = r;
  }
}

// block_sum2_all's two sums (same order, same bits) in thread 0 only, with
// no trailing barrier.
template <int NT>
__device__ inline void block_sum2_t0(double& a, double& b, double* sh2) {
  constexpr int NW = NT / kWave;
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  a = wave_sum(a);
  b = wave_sum(b);
  if ((t & (kWave - 1)) == 0) {
    sh2[t / kWave] = a;
    sh2[NW + t / kWave] = b;
  }
  __syncthreads();
  if (t == 0) {
    double ra = 0.0, rb = 0.0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      ra += sh2[k];
      rb += sh2[NW + k];
    }
    a = ra;
    b = rb;
  }
}

// Fixed-order sum / max of n block partials written by an earlier kernel on
// the same stream (the kernel boundary makes them visible; no in-kernel
// fences).  Every thread of the block gets the result.  sh needs
// NT / kWave + 1 entries.
template <int NT>
__device__ inline double parts_sum(const double* p, int n, double* sh) {
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  double a = 0.0;
  for (int q = t; q < n; q += NT) a += p[q];
  a = block_sum<NT>(a, sh);
  if (t == 0) sh[NT / kWave] = a;
  __syncthreads();
  a = sh[NT / kWave];
  __syncthreads();
  return a;
}
template <int NT>
__device__ inline double parts_max(const double* p, int n, double* sh) {
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  double a = 0.0;
  for (int q = t; q < n; q += NT) a = fmax(a, p[q]);
  a = block_max<NT>(a, sh);
  if (t == 0) sh[NT / kWave] = a;
  __syncthreads();
  a = sh[NT / kWave];
  __syncthreads();
  return a;
}
// Two-phase form of parts_sum for latency hiding: parts_load issues the
// loads of this thread's partials (n <= 2 * NT) early; parts_total (after
// any other loads have been issued) reduces them.  Same fixed order as
// parts_sum: thread t holds p[t] + p[t + NT].
// parts_load for n <= K NT partials (n >= 1): thread t holds p[t] +
// p[t + NT] + ... + p[t + (K-1) NT] in that order (K = 2: parts_load).
// Every load is issued unconditionally at an index clamped into [0, n) and
// the out-of-range ones are dropped by a select, so the K loads are in
// flight together (a load under a branch is waited for inside the branch).
template <int NT, int K>
__device__ inline double parts_load_k(const double* p, int n) {
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  if (n <= NT) {  // block-uniform: one partial per thread at most
    const double v0 = p[min(t, n - 1)];
    return t < n ? v0 : 0.0;
  }
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = p[min(t + k * NT, n - 1)];
  double a = t < n ? v[0] : 0.0;
#pragma unroll
  for (int k = 1; k < K; ++k) a = t + k * NT < n ? a + v[k] : a;
  return a;
}
template <int NT>
__device__ inline double parts_load(const double* p, int n) {
  return parts_load_k<NT, 2>(p, n);
}
template <int NT>
__device__ inline double parts_total(double a, double* sh) {
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  a = block_sum<NT>(a, sh);
  if (t == 0) sh[NT / kWave] = a;
  __syncthreads();
  a = sh[NT / kWave];
  __syncthreads();
  return a;
}

// Reads one int of device state once per block (uniform within the block).
__device__ inline int block_read(const volatile int* p, int* sh) {
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  if (t == 0) *sh = *p;
  __syncthreads();
  const int v = *sh;
  __syncthreads();
  return v;
}

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

}  // namespace cps
