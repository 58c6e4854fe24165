// cps_mg.hip -- geometric-multigrid preconditioned CG for the SIMPLE
// pressure-correction system, FP64 or FP32-inner / FP64-refined.
//
// The fine operator is the reference's 5-point pressure matrix
// (pressure_solver.cpp:531-559): face coefficients aE = rho dE hy,
// aN = rho dN hx, Neumann at walls and inlet, the east face of column N-2
// Dirichlet (p'(i,N-1) = 0), diagonal = sum of the faces (1 if < 1e-20).
//
// Hierarchy: 2x2 cell aggregation; the coarse face coefficient is half the
// sum of the fine faces it covers (half the Galerkin P0 operator, which is
// the rediscretised operator for smooth coefficients; it keeps the Brinkman
// 1e8 coefficient contrast because faces are summed, not averaged).
// Smoother: damped Jacobi (omega 0.8), 2 pre + 2 post sweeps, so the V-cycle
// is symmetric and PCG-safe.  Levels down to 1024 (FP64) / 2048 (FP32)
// cells run as grid kernels; all smaller levels, down to a single cell
// solved exactly in FP64, run inside one workgroup in LDS (k_mg_coarse).
//
// Mixed precision: r64 = b - A x64 (FP64), FP32 MG-PCG on A e = r64, x64 += e
// (FP64), repeated until ||r64|| meets the tolerance.
#include <algorithm>
#include <cmath>
#include <vector>

#include "cps_mg.hpp"

namespace cps {

namespace {

constexpr int MBX = 64, MBY = 4, MNT = MBX * MBY;
constexpr double kOmega = 0.8;
constexpr double kCoarseScale = 0.5;
constexpr int kCoarseThreads = 1024;
constexpr double kTiny20 = (double)1e-20f;

template <class T>
struct Lv {  // device view of one level; cell (r,c) at off + r*pitch + c
  int nx, ny, pitch, off;
  T *x, *y, *b, *ex, *ey;
  __device__ size_t at(int r, int c) const { return (size_t)off + (size_t)r * pitch + c; }
};

// Face coefficients around cell k: ex(k) is the east face (the Dirichlet
// outlet face in the last column), ey(k) the north face (0 in the last row).
// The diagonal is their sum in the reference order aE + aW + aN + aS, so it
// is never stored (pressure_solver.cpp:589-594; floor of :594).
template <class T>
struct Faces {
  T e, w, n, s, d;
};
template <class T>
__device__ inline Faces<T> faces(const Lv<T>& L, size_t k, int r, int c) {
  Faces<T> f;
  f.e = L.ex[k];
  f.w = c > 0 ? L.ex[k - 1] : (T)0;
  f.n = L.ey[k];
  f.s = r > 0 ? L.ey[k - L.pitch] : (T)0;
  T d = f.e + f.w + f.n + f.s;
  f.d = d < (T)kTiny20 ? (T)1 : d;
  return f;
}

template <class T>
__device__ inline T apply_at(const Lv<T>& L, const T* x, int r, int c) {
  const size_t k = L.at(r, c);
  const Faces<T> f = faces(L, k, r, c);
  T ax = f.d * x[k];
  if (c + 1 < L.nx) ax -= f.e * x[k + 1];
  if (c > 0) ax -= f.w * x[k - 1];
  if (r + 1 < L.ny) ax -= f.n * x[k + L.pitch];
  if (r > 0) ax -= f.s * x[k - L.pitch];
  return ax;
}

// level-0 faces from dE, dN (pressure_solver.cpp:589-594) in type T; the
// diagonal keeps the reference floor.
template <class T>
__global__ __launch_bounds__(MNT) void k_mg_build0(Geo g, const double* dE, const double* dN, Lv<T> L) {
  const int c = blockIdx.x * MBX + threadIdx.x, r = blockIdx.y * MBY + threadIdx.y;
  if (r >= L.ny || c >= L.nx) return;
  const int i = r + 1, j = c + 1, M = g.M;
  const size_t f = (size_t)i * g.P + j;
  const double aE = g.rho * dE[f] * g.hy;
  const double aN = (i < M - 1) ? g.rho * dN[f] * g.hx : 0.0;
  const size_t k = L.at(r, c);
  L.ex[k] = (T)aE;
  L.ey[k] = (T)aN;
}

// coarse faces: half the sum of the covered fine faces
template <class T>
__device__ inline T coarse_ex(const Lv<T>& F, int R, int C) {
  const int col = min(2 * C + 1, F.nx - 1);
  T s = F.ex[F.at(2 * R, col)];
  if (2 * R + 1 < F.ny) s += F.ex[F.at(2 * R + 1, col)];
  return (T)kCoarseScale * s;
}
template <class T>
__device__ inline T coarse_ey(const Lv<T>& F, int R, int C) {
  const int row = 2 * R + 1;
  if (row > F.ny - 1) return (T)0;
  T s = F.ey[F.at(row, 2 * C)];
  if (2 * C + 1 < F.nx) s += F.ey[F.at(row, 2 * C + 1)];
  return (T)kCoarseScale * s;
}

template <class T>
__global__ __launch_bounds__(MNT) void k_mg_coarsen(Lv<T> F, Lv<T> Cl) {
  const int C = blockIdx.x * MBX + threadIdx.x, R = blockIdx.y * MBY + threadIdx.y;
  if (R >= Cl.ny || C >= Cl.nx) return;
  const size_t k = Cl.at(R, C);
  Cl.ex[k] = coarse_ex(F, R, C);
  Cl.ey[k] = coarse_ey(F, R, C);
}

// first pre-smoothing sweep from a zero guess: y = omega b / diag
template <class T>
__global__ __launch_bounds__(MNT) void k_mg_jac0(Lv<T> L, const int* stop) {
  if (*stop) return;
  const int c = blockIdx.x * MBX + threadIdx.x, r = blockIdx.y * MBY + threadIdx.y;
  if (r >= L.ny || c >= L.nx) return;
  const size_t k = L.at(r, c);
  L.x[k] = (T)kOmega * L.b[k] / faces(L, k, r, c).d;
}

// dst = xs + omega (b - A xs) / diag, xs = src (+ coarse correction of the
// parent aggregate when Cl.y != nullptr: fused prolongation)
template <class T>
__global__ __launch_bounds__(MNT) void k_mg_jac(Lv<T> L, const T* src, T* dst, Lv<T> Cl, int add_coarse,
                                               const int* stop) {
  if (*stop) return;
  const int c = blockIdx.x * MBX + threadIdx.x, r = blockIdx.y * MBY + threadIdx.y;
  if (r >= L.ny || c >= L.nx) return;
  auto xs = [&](int rr, int cc) -> T {
    T v = src[L.at(rr, cc)];
    if (add_coarse) v += Cl.y[Cl.at(rr >> 1, cc >> 1)];
    return v;
  };
  const size_t k = L.at(r, c);
  const Faces<T> f = faces(L, k, r, c);
  const T xc = xs(r, c);
  T ax = f.d * xc;
  if (c + 1 < L.nx) ax -= f.e * xs(r, c + 1);
  if (c > 0) ax -= f.w * xs(r, c - 1);
  if (r + 1 < L.ny) ax -= f.n * xs(r + 1, c);
  if (r > 0) ax -= f.s * xs(r - 1, c);
  dst[k] = xc + (T)kOmega * (L.b[k] - ax) / f.d;
}

// coarse rhs = sum of the fine residuals b - A x over the 2x2 aggregate
template <class T>
__global__ __launch_bounds__(MNT) void k_mg_restrict(Lv<T> L, const T* x, Lv<T> Cl, const int* stop) {
  if (*stop) return;
  const int C = blockIdx.x * MBX + threadIdx.x, R = blockIdx.y * MBY + threadIdx.y;
  if (R >= Cl.ny || C >= Cl.nx) return;
  T s = (T)0;
  for (int dr = 0; dr < 2; ++dr)
    for (int dc = 0; dc < 2; ++dc) {
      const int r = 2 * R + dr, c = 2 * C + dc;
      if (r < L.ny && c < L.nx) s += L.b[L.at(r, c)] - apply_at(L, x, r, c);
    }
  Cl.b[Cl.at(R, C)] = s;
}

// ---------------------------------------------------------------------------
// All levels from `top` down inside one workgroup: the top level's rhs and
// faces come from global memory, every smaller level is built and solved in
// LDS; the top level's correction is written to top.y.
struct CoarseDesc {
  int nlev;
  int nx[16], ny[16], off[16];  // compact LDS layout, pitch = nx
  int total;
};

template <class T>
__global__ __launch_bounds__(kCoarseThreads) void k_mg_coarse(Lv<T> top, CoarseDesc d, const int* stop) {
  if (*stop) return;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  T* sx = reinterpret_cast<T*>(smem_raw);
  T* sb = sx + d.total;
  T* se = sb + d.total;
  T* sn = se + d.total;
  T* sd = sn + d.total;
  const int t = threadIdx.x;
  // load top level
  {
    const int n = d.nx[0] * d.ny[0];
    for (int q = t; q < n; q += kCoarseThreads) {
      const int r = q / d.nx[0], c = q % d.nx[0];
      const size_t k = top.at(r, c);
      sb[q] = top.b[k];
      se[q] = top.ex[k];
      sn[q] = top.ey[k];
      sd[q] = faces(top, k, r, c).d;
    }
  }
  __syncthreads();
  // build coarser faces in LDS
  for (int l = 1; l < d.nlev; ++l) {
    const int fx = d.nx[l - 1], fy = d.ny[l - 1], fo = d.off[l - 1];
    const int cx = d.nx[l], cy = d.ny[l], co = d.off[l];
    auto cex = [&](int R, int C) -> T {
      const int col = min(2 * C + 1, fx - 1);
      T s = se[fo + 2 * R * fx + col];
      if (2 * R + 1 < fy) s += se[fo + (2 * R + 1) * fx + col];
      return (T)kCoarseScale * s;
    };
    auto cey = [&](int R, int C) -> T {
      const int row = 2 * R + 1;
      if (row > fy - 1) return (T)0;
      T s = sn[fo + row * fx + 2 * C];
      if (2 * C + 1 < fx) s += sn[fo + row * fx + 2 * C + 1];
      return (T)kCoarseScale * s;
    };
    for (int q = t; q < cx * cy; q += kCoarseThreads) {
      const int R = q / cx, C = q % cx;
      const T e = cex(R, C), nn = cey(R, C);
      T dd = e + nn;
      if (C > 0) dd += cex(R, C - 1);
      if (R > 0) dd += cey(R - 1, C);
      se[co + q] = e;
      sn[co + q] = nn;
      sd[co + q] = dd > (T)0 ? dd : (T)1;
    }
    __syncthreads();
  }
  auto Ax = [&](int l, int r, int c) -> T {
    const int nx = d.nx[l], ny = d.ny[l], o = d.off[l];
    const int k = o + r * nx + c;
    T a = sd[k] * sx[k];
    if (c + 1 < nx) a -= se[k] * sx[k + 1];
    if (c > 0) a -= se[k - 1] * sx[k - 1];
    if (r + 1 < ny) a -= sn[k] * sx[k + nx];
    if (r > 0) a -= sn[k - nx] * sx[k - nx];
    return a;
  };
  constexpr int KMAX = 4;  // cells per thread held in registers during a sweep
  auto jacobi = [&](int l, bool from_zero) {
    const int nx = d.nx[l], n = nx * d.ny[l], o = d.off[l];
    T nv[KMAX];
    int cnt = 0;
    for (int q = t; q < n && cnt < KMAX; q += kCoarseThreads, ++cnt) {
      const int k = o + q;
      nv[cnt] = from_zero ? (T)kOmega * sb[k] / sd[k]
                          : sx[k] + (T)kOmega * (sb[k] - Ax(l, q / nx, q % nx)) / sd[k];
    }
    __syncthreads();
    cnt = 0;
    for (int q = t; q < n && cnt < KMAX; q += kCoarseThreads, ++cnt) sx[o + q] = nv[cnt];
    __syncthreads();
  };
  const int last = d.nlev - 1;
  // down
  for (int l = 0; l < last; ++l) {
    jacobi(l, true);
    jacobi(l, false);
    const int fx = d.nx[l], fy = d.ny[l], cx = d.nx[l + 1], cy = d.ny[l + 1], co = d.off[l + 1];
    for (int q = t; q < cx * cy; q += kCoarseThreads) {
      const int R = q / cx, C = q % cx;
      T s = (T)0;
      for (int dr = 0; dr < 2; ++dr)
        for (int dc = 0; dc < 2; ++dc) {
          const int r = 2 * R + dr, c = 2 * C + dc;
          if (r < fy && c < fx) s += sb[d.off[l] + r * fx + c] - Ax(l, r, c);
        }
      sb[co + q] = s;
    }
    __syncthreads();
  }
  // coarsest level is a single cell: exact solve in FP64
  if (t == 0) {
    const int o = d.off[last];
    sx[o] = (T)((double)sb[o] / (double)sd[o]);
  }
  __syncthreads();
  // up
  for (int l = last - 1; l >= 0; --l) {
    const int fx = d.nx[l], n = fx * d.ny[l], o = d.off[l];
    const int cx = d.nx[l + 1], co = d.off[l + 1];
    for (int q = t; q < n; q += kCoarseThreads) {
      const int r = q / fx, c = q % fx;
      sx[o + q] += sx[co + (r >> 1) * cx + (c >> 1)];
    }
    __syncthreads();
    jacobi(l, false);
    jacobi(l, false);
  }
  {
    const int n = d.nx[0] * d.ny[0];
    for (int q = t; q < n; q += kCoarseThreads) top.y[top.at(q / d.nx[0], q % d.nx[0])] = sx[q];
  }
}

// ---------------------------------------------------------------------------
// PCG pieces on level 0.  Reductions: each block stores its partial; the next
// kernel on the stream sums them in fixed order (no in-kernel fences).
struct PcgScal {
  double rzs[2], rr0, rr, tol2;
  int stop, iters, pad_[2];
};

enum MgSlot { S_RR0 = 0, S_RZ = 1, S_PAP = 2, S_ERR = 3, S_IR = 4, S_COUNT = 5 };

struct Grid0 {  // capped grid of the level-0 vector kernels
  int gx, gy, NB;
};

#define L0_LOOP(G, L, r, c)                                  \
  const int c = blockIdx.x * MBX + threadIdx.x;              \
  for (int r = blockIdx.y * MBY + threadIdx.y; r < (L).ny; r += (G).gy * MBY)

__device__ inline int bid() { return blockIdx.x + blockIdx.y * gridDim.x; }
__device__ inline int tid() { return threadIdx.x + threadIdx.y * MBX; }

// r = src (cast), x = 0, partial ||r||^2.  An outer refinement loop that has
// already converged (outer_stop) stops the inner solve at once.
template <class T, class S>
__global__ __launch_bounds__(MNT) void k_pcg_start(Grid0 G, Lv<T> L, const S* src, T* xv, T* rv, PcgScal* ps,
                                                  const int* outer_stop, double* part) {
  __shared__ double sh[MNT / kWave + 1];
  if (outer_stop && *outer_stop) {
    if (bid() == 0 && tid() == 0) { ps->stop = 1; ps->iters = 0; }
    return;
  }
  if (bid() == 0 && tid() == 0) { ps->stop = 0; ps->iters = 0; }
  double e = 0.0;
  L0_LOOP(G, L, r, c) {
    if (c >= L.nx) continue;
    const size_t k = L.at(r, c);
    const T v = (T)src[k];
    rv[k] = v;
    xv[k] = (T)0;
    e += (double)v * (double)v;
  }
  const double s = block_sum<MNT>(e, sh);
  if (tid() == 0) part[(size_t)S_RR0 * G.NB + bid()] = s;
}

// last post-smoothing sweep of level 0, fused with the partial r.z
template <class T>
__global__ __launch_bounds__(MNT) void k_mg_jac_dot(Grid0 G, Lv<T> L, const T* src, T* dst, const int* stop,
                                                   double* part) {
  __shared__ double sh[MNT / kWave + 1];
  if (*stop) return;
  double e = 0.0;
  L0_LOOP(G, L, r, c) {
    if (c >= L.nx) continue;
    const size_t k = L.at(r, c);
    const Faces<T> f = faces(L, k, r, c);
    T ax = f.d * src[k];
    if (c + 1 < L.nx) ax -= f.e * src[k + 1];
    if (c > 0) ax -= f.w * src[k - 1];
    if (r + 1 < L.ny) ax -= f.n * src[k + L.pitch];
    if (r > 0) ax -= f.s * src[k - L.pitch];
    const T z = src[k] + (T)kOmega * (L.b[k] - ax) / f.d;
    dst[k] = z;
    e += (double)L.b[k] * (double)z;
  }
  const double s = block_sum<MNT>(e, sh);
  if (tid() == 0) part[(size_t)S_RZ * G.NB + bid()] = s;
}

// stop test (start: ||r0|| against max(abs, rtol ||r0||); later: ||r|| of
// the previous update), beta = rz / rz_old, p = z + beta p_old fused into
// Ap, partial pAp
template <class T>
__global__ __launch_bounds__(MNT) void k_pcg_ap(Grid0 G, Lv<T> L, const T* z, const T* po, T* pn, T* Ap, int it,
                                               double abs_tol, double rtol, PcgScal* ps, double* part) {
  __shared__ double sh[MNT / kWave + 1];
  __shared__ int shi;
  if (block_read(&ps->stop, &shi)) return;
  const bool b0 = bid() == 0 && tid() == 0;
  if (it == 0) {
    const double rr0 = parts_sum<MNT>(part + (size_t)S_RR0 * G.NB, G.NB, sh);
    const double tgt = fmax(abs_tol, rtol * sqrt(rr0));
    if (b0) { ps->rr0 = rr0; ps->rr = rr0; ps->tol2 = tgt * tgt; }
    if (rr0 <= tgt * tgt) {
      if (b0) { ps->stop = 1; ps->iters = 0; }
      return;
    }
  } else {
    const double err = parts_sum<MNT>(part + (size_t)S_ERR * G.NB, G.NB, sh);
    if (b0) { ps->rr = err; ps->iters = it; }
    if (err <= ps->tol2) {
      if (b0) ps->stop = 1;
      return;
    }
  }
  const double rz = parts_sum<MNT>(part + (size_t)S_RZ * G.NB, G.NB, sh);
  const T beta = it == 0 ? (T)0 : (T)(rz / ps->rzs[(it - 1) & 1]);
  if (b0) ps->rzs[it & 1] = rz;
  double e = 0.0;
  L0_LOOP(G, L, r, c) {
    if (c >= L.nx) continue;
    const size_t k = L.at(r, c);
    auto pv = [&](size_t q) -> T { return z[q] + beta * po[q]; };
    const Faces<T> f = faces(L, k, r, c);
    const T pc = pv(k);
    T ax = f.d * pc;
    if (c + 1 < L.nx) ax -= f.e * pv(k + 1);
    if (c > 0) ax -= f.w * pv(k - 1);
    if (r + 1 < L.ny) ax -= f.n * pv(k + L.pitch);
    if (r > 0) ax -= f.s * pv(k - L.pitch);
    pn[k] = pc;
    Ap[k] = ax;
    e += (double)pc * (double)ax;
  }
  const double s = block_sum<MNT>(e, sh);
  if (tid() == 0) part[(size_t)S_PAP * G.NB + bid()] = s;
}

// alpha = rz / pAp, x += alpha p, r -= alpha Ap, partial ||r||^2
template <class T>
__global__ __launch_bounds__(MNT) void k_pcg_xr(Grid0 G, Lv<T> L, T* xv, T* rv, const T* pn, const T* Ap, int it,
                                               PcgScal* ps, double* part) {
  __shared__ double sh[MNT / kWave + 1];
  __shared__ int shi;
  if (block_read(&ps->stop, &shi)) return;
  const double pap = parts_sum<MNT>(part + (size_t)S_PAP * G.NB, G.NB, sh);
  if (fabs(pap) < kTiny20) {
    if (bid() == 0 && tid() == 0) { ps->stop = 1; ps->iters = it; }
    return;
  }
  const T alpha = (T)(ps->rzs[it & 1] / pap);
  double e = 0.0;
  L0_LOOP(G, L, r, c) {
    if (c >= L.nx) continue;
    const size_t k = L.at(r, c);
    xv[k] += alpha * pn[k];
    const T rn = rv[k] - alpha * Ap[k];
    rv[k] = rn;
    e += (double)rn * (double)rn;
  }
  const double s = block_sum<MNT>(e, sh);
  if (tid() == 0) part[(size_t)S_ERR * G.NB + bid()] = s;
}

// after K launched iterations: stop test of iteration K-1 (one block)
__global__ __launch_bounds__(MNT) void k_pcg_end(Grid0 G, int K, PcgScal* ps, const double* part) {
  __shared__ double sh[MNT / kWave + 1];
  __shared__ int shi;
  if (block_read(&ps->stop, &shi)) return;
  const double err = parts_sum<MNT>(part + (size_t)S_ERR * G.NB, G.NB, sh);
  if (tid() == 0) {
    ps->rr = err;
    ps->iters = K;
    if (err <= ps->tol2) ps->stop = 1;
  }
}

// FP64 refinement: stage 0: x64 += e; stage 1: r64 = b64 - A x64 with the
// FP64 fine operator from dE, dN and the partial ||r64||^2.
__global__ __launch_bounds__(MNT) void k_ir_add(Lv<float> L, const float* e, double* x64, int P, const int* stop) {
  if (*stop) return;
  const int c = blockIdx.x * MBX + threadIdx.x, r = blockIdx.y * MBY + threadIdx.y;
  if (r >= L.ny || c >= L.nx) return;
  x64[(size_t)(r + 1) * P + c + 1] += (double)e[L.at(r, c)];
}

__global__ __launch_bounds__(MNT) void k_ir_resid(Grid0 G, Geo g, Lv<float> L, const double* dE, const double* dN,
                                                 const double* x64, const double* b64, double* r64, const int* stop,
                                                 double* part) {
  __shared__ double sh[MNT / kWave + 1];
  if (*stop) return;
  const int M = g.M, N = g.N, P = g.P;
  double e2 = 0.0;
  L0_LOOP(G, L, r, c) {
    if (c >= L.nx) continue;
    const int i = r + 1, j = c + 1;
    const size_t f = (size_t)i * P + j;
    const double aE = g.rho * dE[f] * g.hy;
    const double aW = (j > 1) ? g.rho * dE[f - 1] * g.hy : 0.0;
    const double aN = (i < M - 1) ? g.rho * dN[f] * g.hx : 0.0;
    const double aS = (i > 1) ? g.rho * dN[f - P] * g.hx : 0.0;
    double aP = aE + aW + aN + aS;
    if (aP < kTiny20) aP = 1.0;
    double ax = aP * x64[f];
    if (i > 1) ax -= aS * x64[f - P];
    if (j > 1) ax -= aW * x64[f - 1];
    if (j + 1 <= N - 2) ax -= aE * x64[f + 1];
    if (i < M - 1) ax -= aN * x64[f + P];
    const double rn = b64[f] - ax;
    r64[f] = rn;
    e2 += rn * rn;
  }
  const double s = block_sum<MNT>(e2, sh);
  if (tid() == 0) part[(size_t)S_IR * G.NB + bid()] = s;
}

// outer test: ||r64||^2 <= max(abs, rtol ||b||)^2 (||b||^2 = rr0 of the first
// inner solve); one block
__global__ __launch_bounds__(MNT) void k_ir_check(Grid0 G, PcgScal* outer, const PcgScal* first, int o,
                                                 double abs_tol, double rtol, const double* part) {
  __shared__ double sh[MNT / kWave + 1];
  __shared__ int shi;
  if (block_read(&outer->stop, &shi)) return;
  const double rr = parts_sum<MNT>(part + (size_t)S_IR * G.NB, G.NB, sh);
  if (tid() == 0) {
    if (o == 0) {
      const double t = fmax(abs_tol, rtol * sqrt(first->rr0));
      outer->tol2 = t * t;
      outer->rr0 = first->rr0;
    }
    outer->rr = rr;
    outer->iters = o + 1;
    if (rr <= outer->tol2) outer->stop = 1;
  }
}

__global__ void k_outer_reset(PcgScal* outer) {
  if (threadIdx.x == 0) { outer->stop = 0; outer->iters = 0; }
}

__global__ void k_copy_state(IterState* st, const PcgScal* inner, const PcgScal* outer, int mixed) {
  if (threadIdx.x == 0) {
    st->pcg_iters = inner->iters;
    st->err = mixed ? outer->rr : inner->rr;
    st->ir_outer = mixed ? outer->iters : 0;
  }
}

template <class T>
__global__ __launch_bounds__(MNT) void k_copy_out(Lv<T> L, const T* src, double* x64, int P) {
  const int c = blockIdx.x * MBX + threadIdx.x, r = blockIdx.y * MBY + threadIdx.y;
  if (r >= L.ny || c >= L.nx) return;
  x64[(size_t)(r + 1) * P + c + 1] = (double)src[L.at(r, c)];
}

}  // namespace

// ===========================================================================
template <class T>
struct MGImpl {
  Ctx* ctx;
  Geo g;
  struct Level {
    int nx, ny, pitch, off;
    DevBuf<T> x, y, b, ex, ey;
    Lv<T> view() { return Lv<T>{nx, ny, pitch, off, x.p, y.p, b.p, ex.p, ey.p}; }
    dim3 grid() const { return dim3(ceil_div(nx, MBX), ceil_div(ny, MBY)); }
  };
  std::vector<std::unique_ptr<Level>> lv;  // grid-kernel levels 0 .. nGrid-1, lv[nGrid] = coarse top
  int nGrid = 0;
  CoarseDesc cd{};
  size_t coarseSmem = 0;
  DevBuf<T> r0, xs, pa, pb, Ap;  // PCG vectors, level-0 layout (xs = solution)
  DevBuf<double> part;
  DevBuf<PcgScal> scal;          // [0] inner, [1] outer
  Grid0 G0{};
  int coarseCells;
  KernelTimer* timer = nullptr;

  MGImpl(Ctx* c, const Geo& geo) : ctx(c), g(geo) {
    coarseCells = sizeof(T) == 8 ? 1024 : 2048;
    // level 0: the field layout, cell (r,c) = (i-1, j-1)
    auto L0 = std::make_unique<Level>();
    L0->nx = g.N - 2;
    L0->ny = g.M - 1;
    L0->pitch = g.P;
    L0->off = g.P + 1;
    const size_t S0 = (size_t)(g.M + 1) * g.P;
    for (auto* b : {&L0->x, &L0->y, &L0->ex, &L0->ey}) b->alloc(S0);
    for (auto* b : {&r0, &xs, &pa, &pb, &Ap}) b->alloc(S0);
    lv.push_back(std::move(L0));
    // at least one grid level: its last post-smoothing sweep carries the
    // fused r.z partials of the PCG
    while (lv.size() == 1 || (size_t)lv.back()->nx * lv.back()->ny > (size_t)coarseCells) {
      auto L = std::make_unique<Level>();
      L->nx = (lv.back()->nx + 1) / 2;
      L->ny = (lv.back()->ny + 1) / 2;
      L->pitch = ((L->nx + 63) / 64) * 64;
      L->off = 0;
      const size_t S = (size_t)L->ny * L->pitch;
      for (auto* b : {&L->x, &L->y, &L->b, &L->ex, &L->ey}) b->alloc(S);
      lv.push_back(std::move(L));
    }
    nGrid = (int)lv.size() - 1;
    int nx = lv.back()->nx, ny = lv.back()->ny, off = 0, n = 0;
    while (true) {
      cd.nx[n] = nx;
      cd.ny[n] = ny;
      cd.off[n] = off;
      off += nx * ny;
      ++n;
      if ((nx == 1 && ny == 1) || n == 16) break;
      nx = (nx + 1) / 2;
      ny = (ny + 1) / 2;
    }
    cd.nlev = n;
    cd.total = off;
    coarseSmem = (size_t)5 * off * sizeof(T);
    CPS_CHECK(coarseSmem <= 64 * 1024, CPS_ERR_STATE, "coarse multigrid levels exceed LDS");
    for (auto* b : {&lv[0]->x, &lv[0]->y, &lv[0]->ex, &lv[0]->ey, &r0, &xs, &pa, &pb, &Ap})
      CPS_HIP(hipMemsetAsync(b->p, 0, b->bytes(), ctx->stream));
    G0.gx = ceil_div(lv[0]->nx, MBX);
    G0.gy = std::max(1, std::min(ceil_div(lv[0]->ny, MBY), std::max(1, 1024 / G0.gx)));
    G0.NB = G0.gx * G0.gy;
    part.alloc((size_t)S_COUNT * G0.NB);
    CPS_HIP(hipMemsetAsync(part.p, 0, part.bytes(), ctx->stream));
    scal.alloc(2);
    CPS_HIP(hipMemsetAsync(scal.p, 0, 2 * sizeof(PcgScal), ctx->stream));
  }

  dim3 blk() const { return dim3(MBX, MBY); }
  dim3 grid0() const { return dim3(G0.gx, G0.gy); }
  // level views; the level-0 rhs of the V-cycle is the PCG residual r0
  Lv<T> view(int l) {
    Lv<T> v = lv[l]->view();
    if (l == 0) v.b = r0.p;
    return v;
  }

  void setup(const double* dE, const double* dN) {
    hipStream_t s = ctx->stream;
    k_mg_build0<T><<<lv[0]->grid(), blk(), 0, s>>>(g, dE, dN, lv[0]->view());
    for (size_t l = 1; l < lv.size(); ++l)
      k_mg_coarsen<T><<<lv[l]->grid(), blk(), 0, s>>>(lv[l - 1]->view(), lv[l]->view());
    CPS_HIP(hipGetLastError());
  }

  // z = V(r0) into level-0 y, with the partial r.z of the last sweep
  void vcycle(const int* stop) {
    hipStream_t s = ctx->stream;
    Lv<T> none{};
    for (int l = 0; l < nGrid; ++l) {
      Lv<T> L = view(l);
      k_mg_jac0<T><<<lv[l]->grid(), blk(), 0, s>>>(L, stop);
      if (l == 0) timer->begin(KT_MG_SMOOTH0);
      k_mg_jac<T><<<lv[l]->grid(), blk(), 0, s>>>(L, L.x, L.y, none, 0, stop);
      if (l == 0) timer->end(KT_MG_SMOOTH0);
      k_mg_restrict<T><<<lv[l + 1]->grid(), blk(), 0, s>>>(L, L.y, view(l + 1), stop);
    }
    k_mg_coarse<T><<<1, kCoarseThreads, coarseSmem, s>>>(view(nGrid), cd, stop);
    for (int l = nGrid - 1; l >= 0; --l) {
      Lv<T> L = view(l);
      k_mg_jac<T><<<lv[l]->grid(), blk(), 0, s>>>(L, L.y, L.x, view(l + 1), 1, stop);
      if (l > 0) k_mg_jac<T><<<lv[l]->grid(), blk(), 0, s>>>(L, L.x, L.y, none, 0, stop);
      else k_mg_jac_dot<T><<<grid0(), blk(), 0, s>>>(G0, L, L.x, L.y, stop, part.p);
    }
    CPS_HIP(hipGetLastError());
  }

  // MG-PCG: A xs = rhs (rhs FP64 in the field layout).  Iterations run with
  // device-side termination; with sync_every > 0 the host polls the stop
  // flag every sync_every iterations.
  void pcg(const double* rhs, double abs_tol, double rtol, int maxit, int sync_every, const int* outer_stop) {
    hipStream_t s = ctx->stream;
    const double nP = (double)lv[0]->nx * (double)lv[0]->ny;
    // smoother: x, b, east and north faces read, x' written; PCG stencil:
    // z, p_old, faces read, p and Ap written
    if (timer->which == KT_MG_SMOOTH0) timer->bytes_per_launch = 5.0 * sizeof(T) * nP;
    if (timer->which == KT_PCG_STENCIL) timer->bytes_per_launch = 6.0 * sizeof(T) * nP;
    const Lv<T> L = view(0);
    PcgScal* ps = scal.p;
    k_pcg_start<T, double><<<grid0(), blk(), 0, s>>>(G0, L, rhs, xs.p, r0.p, ps, outer_stop, part.p);
    // p_old of iteration 0 is pa: zero since creation, finite afterwards, and
    // multiplied by beta = 0
    for (int it = 0; it < maxit; ++it) {
      vcycle(&ps->stop);
      T* po = (it & 1) ? pb.p : pa.p;
      T* pn = (it & 1) ? pa.p : pb.p;
      timer->begin(KT_PCG_STENCIL);
      k_pcg_ap<T><<<grid0(), blk(), 0, s>>>(G0, L, lv[0]->y.p, po, pn, Ap.p, it, abs_tol, rtol, ps, part.p);
      timer->end(KT_PCG_STENCIL);
      k_pcg_xr<T><<<grid0(), blk(), 0, s>>>(G0, L, xs.p, r0.p, pn, Ap.p, it, ps, part.p);
      CPS_HIP(hipGetLastError());
      if (sync_every > 0 && (it + 1) % sync_every == 0 && it + 1 < maxit) {
        k_pcg_end<<<1, blk(), 0, s>>>(G0, it + 1, ps, part.p);
        int stop = 0;
        CPS_HIP(hipMemcpyAsync(&stop, &ps->stop, sizeof(int), hipMemcpyDeviceToHost, s));
        CPS_HIP(hipStreamSynchronize(s));
        if (stop) return;
      }
    }
    k_pcg_end<<<1, blk(), 0, s>>>(G0, maxit, ps, part.p);
  }
};

PressureMG::PressureMG(Ctx* ctx, const Geo& g) : ctx_(ctx), g_(g) {}

PressureMG::~PressureMG() {
  delete static_cast<MGImpl<double>*>(impl64_);
  delete static_cast<MGImpl<float>*>(impl32_);
  delete static_cast<DevBuf<double>*>(b64_);
}

// Host polling happens only when the iteration caps are large (reference-
// style runs); small caps launch a fixed kernel chain with device-side
// termination (no host synchronisation, capturable in a graph).
static constexpr int kNoSyncCap = 32;

void PressureMG::solve(const double* dE, const double* dN, double* r, double* x, IterState* st,
                       const cps_simple_config& cfg, bool fp64, KernelTimer* timer) {
  hipStream_t s = ctx_->stream;
  const int maxit = std::max(0, cfg.max_pressure_iter);
  const int sync_every = maxit > kNoSyncCap ? 4 : 0;
  if (fp64) {
    if (!impl64_) impl64_ = new MGImpl<double>(ctx_, g_);
    auto* m = static_cast<MGImpl<double>*>(impl64_);
    m->timer = timer;
    m->setup(dE, dN);
    m->pcg(r, cfg.p_tol, cfg.p_rtol, maxit, sync_every, nullptr);
    k_copy_out<double><<<m->lv[0]->grid(), dim3(MBX, MBY), 0, s>>>(m->view(0), m->xs.p, x, g_.P);
    k_copy_state<<<1, 64, 0, s>>>(st, m->scal.p, m->scal.p + 1, 0);
  } else {
    if (!impl32_) impl32_ = new MGImpl<float>(ctx_, g_);
    auto* m = static_cast<MGImpl<float>*>(impl32_);
    const size_t S = (size_t)(g_.M + 1) * g_.P;
    if (!b64_) {
      auto* b = new DevBuf<double>;
      b->alloc(S);
      b64_ = b;
    }
    double* b64 = static_cast<DevBuf<double>*>(b64_)->p;
    m->timer = timer;
    m->setup(dE, dN);
    CPS_HIP(hipMemcpyAsync(b64, r, S * sizeof(double), hipMemcpyDeviceToDevice, s));
    PcgScal* inner = m->scal.p;
    PcgScal* outer = m->scal.p + 1;
    k_outer_reset<<<1, 64, 0, s>>>(outer);
    const Lv<float> L = m->view(0);
    const double inner_rtol = std::max(cfg.p_rtol, 1e-5);
    const int max_outer = std::max(1, cfg.max_refine);
    const bool poll = sync_every > 0 || max_outer > 4;
    for (int o = 0; o < max_outer; ++o) {
      m->pcg(r, 0.0, inner_rtol, maxit, sync_every, &outer->stop);
      k_ir_add<<<m->lv[0]->grid(), dim3(MBX, MBY), 0, s>>>(L, m->xs.p, x, g_.P, &outer->stop);
      k_ir_resid<<<m->grid0(), dim3(MBX, MBY), 0, s>>>(m->G0, g_, L, dE, dN, x, b64, r, &outer->stop, m->part.p);
      k_ir_check<<<1, dim3(MBX, MBY), 0, s>>>(m->G0, outer, inner, o, cfg.p_tol, cfg.p_rtol, m->part.p);
      CPS_HIP(hipGetLastError());
      if (poll && o + 1 < max_outer) {
        int stop = 0;
        CPS_HIP(hipMemcpyAsync(&stop, &outer->stop, sizeof(int), hipMemcpyDeviceToHost, s));
        CPS_HIP(hipStreamSynchronize(s));
        if (stop) break;
      }
    }
    k_copy_state<<<1, 64, 0, s>>>(st, inner, outer, 1);
  }
  CPS_HIP(hipGetLastError());
}

}  // namespace cps
