// cps_mg_coarse.hpp -- device pieces of the pressure multigrid that the
// SIMPLE step kernels share with the multigrid kernels: the level view, the
// coarse solver's descriptor and its once-per-solve set-up.  The set-up is
// one workgroup's work; it runs either as its own launch
// (k_mg_coarse_setup) or as one extra block of the step's momentum
// finalize kernel, which reads nothing the set-up writes and so carries it
// beside its own tiles.
#pragma once

#include "cps_simple.hpp"

namespace cps {
namespace mgc {

#ifndef CPS_MG_COARSE_SCALE
#define CPS_MG_COARSE_SCALE 0.5
#endif
#ifndef CPS_MG_DIRICHLET_SCALE
#define CPS_MG_DIRICHLET_SCALE 1.25
#endif
// Coarse faces are this share of the covered fine faces' sum (0.5: twice
// the Galerkin coarse correction).
constexpr double kCoarseScale = CPS_MG_COARSE_SCALE;
// The east face of a level's last column couples it to the pinned outlet
// column p'(., N-1) = 0.  On a coarse level that column's centre lies half a
// coarse cell or less from the pinned one, while its face sum keeps the fine
// face's size, so kCoarseScale alone under-weights the outlet and the
// smoothest error mode (the only one the outlet pins) is over-corrected.
// The last coarse column's east face carries this extra factor.
constexpr double kDirichletScale = CPS_MG_DIRICHLET_SCALE;
// the factor of a coarse east face whose covered fine column is fcol of a
// finer level fnx wide (the last coarse column covers the last fine one)
__host__ __device__ constexpr double coarse_e_scale(int fcol, int fnx) {
  return fcol >= fnx - 1 ? kCoarseScale * kDirichletScale : kCoarseScale;
}
#ifndef CPS_COARSE_THREADS
#define CPS_COARSE_THREADS 1024
#endif
constexpr int kCoarseThreads = CPS_COARSE_THREADS;
// the coarse hierarchy ends at its first level of at most kDirectCells cells,
// which is solved with its inverse (one column per lane of a wavefront)
constexpr int kDirectCells = 32;
static_assert(kDirectCells * kDirectCells <= kCoarseThreads && kDirectCells <= 64, "direct level staging");
constexpr double kTiny20 = (double)1e-20f;

template <class T>
struct Lv {  // device view of one level; cell (r,c) at off + r*pitch + c
  int nx, ny, pitch, off;
  T *x, *y, *b, *ex, *ey;
  size_t zs;  // byte stride between the fields of a class-batched launch
  __device__ size_t at(int r, int c) const { return (size_t)off + (size_t)r * pitch + c; }
  __device__ void zshift(size_t o) {
    x = zsh(x, o);
    y = zsh(y, o);
    b = zsh(b, o);
    ex = zsh(ex, o);
    ey = zsh(ey, o);
  }
};

// All levels from `top` down inside one workgroup: the top level's rhs and
// faces come from global memory, every smaller level is built and solved in
// LDS; the top level's correction is written to top.y.
struct CoarseDesc {
  int nlev;
  int nx[16], ny[16], off[16];  // compact LDS layout, pitch = nx
  int total;
  int wl;  // the last level (at most kDirectCells cells): solved with its inverse
  int dn;  // its cell count; the inverse sits at cf[4 total ..), column-major
  unsigned long long mag[16];  // ceil(2^32 / nx): row = (q mag) >> 32 for a cell index q < 2^20
};

// the cells of level l, one per thread in row-major order (every lane of a
// wave busy on narrow levels); r = q / nx by the reciprocal multiply
#define MGC_CELLS(l)                                                             \
  for (int q_ = t, n_ = d.nx[l] * d.ny[l]; q_ < n_; q_ += kCoarseThreads)       \
    if (const int r = (int)(((unsigned long long)q_ * d.mag[l]) >> 32), c = q_ - r * d.nx[l]; true)

// Once per pressure solve: faces and inverse diagonals of every level of the
// coarse solver, built from the top level's faces (half the sum of the
// covered fine faces), stored as cf[0 .. 4 total) = east, north, 1/diag, diag.
// With par.nx > 0 the top level's faces are not in memory: they are built
// here from its parent grid level par, with k_mg_coarsen's arithmetic (no
// other kernel reads the top level's faces).  Run by kCoarseThreads threads
// (t = linear thread index) of one block, pointers already shifted to the
// block's field; smem_raw holds coarse_setup_smem() bytes.
template <class T>
__device__ __forceinline__ void coarse_setup_body(const Lv<T>& top, const Lv<T>& par, const CoarseDesc& d, T* cf,
                                                  unsigned char* smem_raw, int t) {
  T* se = reinterpret_cast<T*>(smem_raw);
  T* sn = se + d.total;
  if (par.nx > 0) {
    T* pe = sn + d.total;  // the parent's faces, pitch par.nx
    T* pn = pe + par.nx * par.ny;
    const int fx = par.nx, fy = par.ny;
    for (int q = t; q < fx * fy; q += kCoarseThreads) {
      const int r = q / fx, c = q - r * fx;
      const size_t k = par.at(r, c);
      pe[q] = par.ex[k];
      pn[q] = par.ey[k];
    }
    __syncthreads();
    MGC_CELLS(0) {
      const int col = min(2 * c + 1, fx - 1), row = 2 * r + 1;
      T e = pe[2 * r * fx + col];
      if (2 * r + 1 < fy) e += pe[(2 * r + 1) * fx + col];
      T n = (T)0;
      if (row <= fy - 1) {
        n = pn[row * fx + 2 * c];
        if (2 * c + 1 < fx) n += pn[row * fx + 2 * c + 1];
      }
      se[r * d.nx[0] + c] = (T)coarse_e_scale(col, fx) * e;
      sn[r * d.nx[0] + c] = (T)kCoarseScale * n;
    }
  } else {
    MGC_CELLS(0) {
      const size_t k = top.at(r, c);
      se[r * d.nx[0] + c] = top.ex[k];
      sn[r * d.nx[0] + c] = top.ey[k];
    }
  }
  __syncthreads();
  for (int l = 1; l < d.nlev; ++l) {
    const int fx = d.nx[l - 1], fy = d.ny[l - 1], fo = d.off[l - 1], co = d.off[l], cx = d.nx[l];
    MGC_CELLS(l) {
      const int col = min(2 * c + 1, fx - 1), row = 2 * r + 1;
      T e = se[fo + 2 * r * fx + col];
      if (2 * r + 1 < fy) e += se[fo + (2 * r + 1) * fx + col];
      T n = (T)0;
      if (row <= fy - 1) {
        n = sn[fo + row * fx + 2 * c];
        if (2 * c + 1 < fx) n += sn[fo + row * fx + 2 * c + 1];
      }
      se[co + r * cx + c] = (T)coarse_e_scale(col, fx) * e;
      sn[co + r * cx + c] = (T)kCoarseScale * n;
    }
    __syncthreads();
  }
  for (int l = 0; l < d.nlev; ++l) {
    const int o = d.off[l], nx = d.nx[l];
    MGC_CELLS(l) {
      const int k = o + r * nx + c;
      T dd = se[k] + (c > 0 ? se[k - 1] : (T)0) + sn[k] + (r > 0 ? sn[k - nx] : (T)0);
      dd = dd < (T)kTiny20 ? (T)1 : dd;
      cf[k] = se[k];
      cf[d.total + k] = sn[k];
      cf[2 * d.total + k] = (T)1 / dd;
      cf[3 * d.total + k] = dd;
    }
  }
  // Inverse of the last level's operator (the faces and floored diagonal the
  // V-cycle applies) by Gauss-Jordan in FP64 over the whole block: thread t
  // < n^2 holds entry (i, j) = (t / n, t mod n) in a register; each pivot's
  // scaled row and its column travel through LDS (double-buffered, one
  // barrier per pivot).  The operator is a symmetric M-matrix, so no
  // pivoting is needed.  Entry (i, j) is built from cell j's faces and
  // updated as m_j = (j == p ? 1 : a_pj) / a_pp (formed by row p's threads,
  // which share a wavefront and take a_pp by a shuffle), row p <- m_j,
  // a_ij <- (j == p ? 0 : a_ij) - a_ip m_j.
  __shared__ double gj_row[2][kDirectCells], gj_col[2][kDirectCells];
  __syncthreads();
  {
    const int l = d.wl, o = d.off[l], nx = d.nx[l], ny = d.ny[l], n = d.dn;
    const bool mine = t < n * n;
    const int gi = mine ? t / n : 0, gj = mine ? t - gi * n : 0;
    double a = 0.0;
    if (mine) {
      const int r = gj / max(nx, 1), c = gj - r * nx, k = o + gj;
      const T dT = se[k] + (c > 0 ? se[k - 1] : (T)0) + sn[k] + (r > 0 ? sn[k - nx] : (T)0);
      const double dg = dT < (T)kTiny20 ? 1.0 : (double)dT;
      const double ee = c + 1 < nx ? -(double)se[k] : 0.0;
      const double ww = c > 0 ? -(double)se[k - 1] : 0.0;
      const double nn = r + 1 < ny ? -(double)sn[k] : 0.0;
      const double ss = r > 0 ? -(double)sn[k - nx] : 0.0;
      if (gi == gj) a = dg;
      else if (gi == gj + 1) a = ee;
      else if (gi == gj - 1) a = ww;
      else if (gi == gj + nx) a = nn;
      else if (gi == gj - nx) a = ss;
    }
    if (kWave % n == 0) {  // block-uniform: every row of the matrix lies in one wavefront
      for (int p = 0; p < n; ++p) {
        const int b = p & 1;
        // a_pp from thread (p, p) of the same wavefront (every lane of the
        // wavefront takes part in the shuffle)
        const double app = __shfl(a, (p * n + p) & (kWave - 1));
        if (mine && gi == p) gj_row[b][gj] = (gj == p ? 1.0 : a) * (1.0 / app);
        if (mine && gj == p) gj_col[b][gi] = a;
        __syncthreads();
        if (mine) a = gi == p ? gj_row[b][gj] : (gj == p ? 0.0 : a) - gj_col[b][gi] * gj_row[b][gj];
      }
    } else {  // the same values, a_pp through LDS
      for (int p = 0; p < n; ++p) {
        const int b = p & 1;
        if (mine && gi == p) gj_row[b][gj] = a;
        if (mine && gj == p) gj_col[b][gi] = a;
        __syncthreads();
        if (mine) {
          const double mp = (gj == p ? 1.0 : gj_row[b][gj]) * (1.0 / gj_row[b][p]);
          a = gi == p ? mp : (gj == p ? 0.0 : a) - gj_col[b][gi] * mp;
        }
      }
    }
    if (mine) cf[4 * d.total + gj * n + gi] = (T)a;
  }
}

// Level 0 and up to kHierLevels coarser levels of the grid hierarchy from dE,
// dN: coarse faces are half the sums of the faces they cover, a purely local
// 2x2 aggregation, so a block that owns a TW x TH level-0 tile (TW, TH
// multiples of 2^kHierLevels) owns the nested tiles of every coarser level,
// and the results do not depend on the tile shape.  Level 0: aE = rho dE hy,
// aN = rho dN hx (0 on the last row), the pressure operator's faces
// (pressure_solver.cpp:589-594); coarser levels: the arithmetic of
// k_mg_coarsen.  NTH threads (t = linear index) of block (bx, by); dE, dN
// already shifted to the block's field, the levels' face pointers shifted by
// zo here; LDS: sEa, sNa hold TW x TH, sEb, sNb a quarter of that (even levels
// in the first pair, odd levels in the second).
constexpr int kHierLevels = 4;
template <class T>
struct HierLevels {
  Lv<T> l[kHierLevels + 1];
  int n;  // coarser levels built here (<= kHierLevels)
};
// The coarser levels from the tile's level-0 faces in sEa, sNa (cells in the
// level only, row pitch TW), the part of hier_body after level 0.
template <class T, int NTH, int TW, int TH>
__device__ __forceinline__ void hier_coarsen(const HierLevels<T>& H, int bx, int by, int t, size_t zo, T* sEa, T* sNa,
                                             T* sEb, T* sNb) {
  static_assert(TW % (1 << kHierLevels) == 0 && TH % (1 << kHierLevels) == 0, "nested tiles");
  int cx0 = bx * TW, cy0 = by * TH;  // tile origin at the current level
  int tw = TW, th = TH;              // tile size at the current level
  for (int l = 1; l <= H.n; ++l) {
    __syncthreads();
    const Lv<T>& F = H.l[l - 1];
    const Lv<T>& Cl = H.l[l];
    const T* fe = (l & 1) ? sEa : sEb;  // the finer level's faces
    const T* fn = (l & 1) ? sNa : sNb;
    T* ce = (l & 1) ? sEb : sEa;
    T* cn = (l & 1) ? sNb : sNa;
    const int fw = tw;  // row length of the finer tile in LDS
    cx0 >>= 1;
    cy0 >>= 1;
    tw >>= 1;
    th >>= 1;
    for (int q = t; q < tw * th; q += NTH) {
      const int lr = q / tw, lc = q - lr * tw, R = cy0 + lr, C = cx0 + lc;
      if (R >= Cl.ny || C >= Cl.nx) continue;
      // fine local (2 lr + dr, 2 lc + dc); the east column clamps to the
      // last fine column, the second row exists when 2R + 1 < F.ny
      const int lcol = 2 * C + 1 <= F.nx - 1 ? 2 * lc + 1 : 2 * lc;
      T e = fe[(2 * lr) * fw + lcol];
      if (2 * R + 1 < F.ny) e += fe[(2 * lr + 1) * fw + lcol];
      T n = (T)0;
      if (2 * R + 1 <= F.ny - 1) {
        n = fn[(2 * lr + 1) * fw + 2 * lc];
        if (2 * C + 1 < F.nx) n += fn[(2 * lr + 1) * fw + 2 * lc + 1];
      }
      e = (T)coarse_e_scale(min(2 * C + 1, F.nx - 1), F.nx) * e;
      n = (T)kCoarseScale * n;
      const size_t k = Cl.at(R, C);
      zsh(Cl.ex, zo)[k] = e;
      zsh(Cl.ey, zo)[k] = n;
      ce[lr * tw + lc] = e;
      cn[lr * tw + lc] = n;
    }
  }
}
template <class T, int NTH, int TW, int TH>
__device__ __forceinline__ void hier_body(const Geo& g, const double* dE, const double* dN, const HierLevels<T>& H,
                                          int bx, int by, int t, size_t zo, T* sEa, T* sNa, T* sEb, T* sNb) {
  const int M = g.M;
  const int cx0 = bx * TW, cy0 = by * TH;
  {
    const Lv<T>& L = H.l[0];
    for (int q = t; q < TW * TH; q += NTH) {
      const int lr = q / TW, lc = q - lr * TW, r = cy0 + lr, c = cx0 + lc;
      if (r >= L.ny || c >= L.nx) continue;
      const int i = r + 1, j = c + 1;
      const size_t f = (size_t)i * g.P + j;
      const double aE = g.rho * dE[f] * g.hy;
      const double aN = (i < M - 1) ? g.rho * dN[f] * g.hx : 0.0;
      const size_t k = L.at(r, c);
      zsh(L.ex, zo)[k] = (T)aE;
      zsh(L.ey, zo)[k] = (T)aN;
      sEa[q] = (T)aE;
      sNa[q] = (T)aN;
    }
  }
  hier_coarsen<T, NTH, TW, TH>(H, bx, by, t, zo, sEa, sNa, sEb, sNb);
}

// The hierarchy build handed to the kernel that forms dE, dN (on = 0: none).
template <class T>
struct HierJob {
  HierLevels<T> H;
  int on;
};

// State of one multigrid PCG solve (inner) or of its refinement loop (outer).
struct PcgScal {
  double rzs[2], rr0, rr, tol2;
  int stop, iters, pad_[2];
  double alphas[2];  // step lengths of the single-reduction form, by iteration parity
};

// The start of the coming multigrid PCG solve (r = b in the inner type, x =
// 0, the FP64 right side kept, the stop flags cleared), handed to the kernel
// that forms b (k_prhs, which then leaves its FP64 b nowhere else); the
// first V-cycle's fine down kernel adds the ||r0||^2 partials (on = 0: the
// solve starts itself).
template <class T>
struct StartJob {
  T* r;            // level-0 residual (field layout)
  T* x;            // level-0 solution
  double* keep;    // FP64 right side kept (nullptr: none)
  PcgScal* ps;     // inner solve state
  PcgScal* reset;  // outer refinement state (nullptr: none)
  int on;
};

// The coarse set-up of the coming pressure solve, handed to a kernel that
// carries it as one extra block (on = 0: none).
template <class T>
struct CoarseJob {
  Lv<T> top, par;  // par.nx == 0: the top's faces are in memory
  CoarseDesc d;
  T* cf;
  int on;
};

}  // namespace mgc
}  // namespace cps
