// cps_simple.hip -- the SIMPLE iteration on gfx950, device resident.
//
// One SIMPLE step (SIMPLE::calculateStep, iterations.cpp:82-659) is a fixed
// chain of kernels on the context stream:
//
//   k_bc            velocity + pressure BCs          boundaries.cpp:50-146
//   k_assemble      U and V momentum coefficients    momentum_solver.cpp:195-314
//                   (+ SOU, half-cell flags, Brinkman / 2.5-D sinks, dE, dN)
//   k_jacobi  x K   fused U+V point-Jacobi sweep      jacobi_solver.cpp:44-67
//                   with per-component sum r^2 test on the device
//   k_finalize      clamp + scatter to uStar/vStar,  momentum_solver.cpp:335-351
//                   BCs on uStar/vStar, RMS change   iterations.cpp:499-533
//   k_prhs          mass residual, pressure RHS,     pressure_solver.cpp:470-496,
//                   PCG start                        :613-623, :121-150
//   pressure solve  Jacobi-PCG (k_pcg_a/k_pcg_b) or geometric-MG PCG
//   k_correct       p += pAlpha p', u/v correction   iterations.cpp:556-600
//                   with clamps, BCs
//
// Outside the step: k_sample reduces the core-plane and boundary-patch
// pressure drops (postprocessing.cpp:102-327) that the outer loop's dP
// window (SIMPLE.cpp:310-491) reads back with the residuals.
//
// Fields live in HBM as row-major (i = y row, j = x column) arrays with one
// common pitch P >= N+1 (a multiple of 64 elements), so every kernel
// reads whole 64-lane rows: one wavefront = 64 consecutive j.  Reductions are
// wave-shuffle + block partials + a last-block finalisation in fixed order,
// so results are bit-reproducible run to run.
//
// Coefficient storage C is double (CPS_PREC_FP64) or float
// (CPS_PREC_MIXED).  In mixed mode the FP64 assembly writes the FP32
// correction system A d = b - A u; FP32 Jacobi sweeps solve for d and the
// FP64 state gets u* = u + d, so the fixed point is the FP64 one.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <type_traits>
#include <vector>

#include <chrono>

#include "cps_mg.hpp"
#include "cps_simple.hpp"

namespace cps {

// float literals of the reference (SIMPLE.h / *.cpp), widened exactly
constexpr double kTenthF = (double)0.1f;
constexpr double kVelTol = (double)1e-8f;
constexpr double kAP0Floor = (double)1e-10f;
constexpr double kAlphaEps = (double)1e-10f;
constexpr double kTiny20 = (double)1e-20f;
constexpr double kKminFloor = (double)1e-20f;

constexpr int BX = 64, BY = 16, NT = BX * BY;  // 1024-thread blocks, one row per thread

struct Fields {
  double *u, *v, *p, *us, *vs, *ps, *dE, *dN, *gam, *alp, *inlet;
  __device__ void zshift(size_t o) {
    u = zsh(u, o);
    v = zsh(v, o);
    p = zsh(p, o);
    us = zsh(us, o);
    vs = zsh(vs, o);
    ps = zsh(ps, o);
    dE = zsh(dE, o);
    dN = zsh(dN, o);
    gam = zsh(gam, o);
    alp = zsh(alp, o);
    inlet = zsh(inlet, o);
  }
};

template <class C>
struct MomSys {  // coefficients of one component, indexed like the field
  C *cS, *cW, *aP, *cE, *cN, *b;
  C *x0, *x1;  // Jacobi double buffer (x-form: velocity, delta-form: correction)
  C* x2;       // third buffer of the paired sweeps (delta form), x_j in buffer j mod 3
  __device__ void zshift(size_t o) {
    cS = zsh(cS, o);
    cW = zsh(cW, o);
    aP = zsh(aP, o);
    cE = zsh(cE, o);
    cN = zsh(cN, o);
    b = zsh(b, o);
    x0 = zsh(x0, o);
    x1 = zsh(x1, o);
    x2 = zsh(x2, o);
  }
};

struct PcgBufs {
  double *r, *z, *pa, *pb, *Ap, *diag;
  __device__ void zshift(size_t o) {
    r = zsh(r, o);
    z = zsh(z, o);
    pa = zsh(pa, o);
    pb = zsh(pb, o);
    Ap = zsh(Ap, o);
    diag = zsh(diag, o);
  }
};

__device__ __forceinline__ double clampv(double x, double m) { return fmin(fmax(x, -m), m); }

__device__ __forceinline__ size_t at(const Geo& g, int i, int j) { return (size_t)i * g.P + j; }

// ---------------------------------------------------------------------------
// Boundary conditions (boundaries.cpp:50-146).  Segment 0 walks j = 0..N
// (walls, p top/bottom ghosts); segment 1 walks i = 0..M (inlet, outlet,
// p left/right ghosts).  Every element has exactly one writer.
__global__ void k_bc(Geo g, double* u, double* v, double* p, const double* inlet, int do_p) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int M = g.M, N = g.N;
  if (blockIdx.y == 0) {
    const int j = t;
    if (j > N) return;
    if (j >= 1 && j <= N - 2) {
      u[at(g, 0, j)] = 0.0;
      u[at(g, M, j)] = 0.0;
    }
    v[at(g, 0, j)] = 0.0;
    v[at(g, M - 1, j)] = 0.0;
    if (do_p && j >= 1 && j <= N - 1) {
      p[at(g, 0, j)] = p[at(g, 1, j)];
      p[at(g, M, j)] = p[at(g, M - 1, j)];
    }
  } else {
    const int i = t;
    if (i > M) return;
    u[at(g, i, 0)] = inlet[i];
    u[at(g, i, N - 1)] = (i == 0 || i == M) ? 0.0 : u[at(g, i, N - 2)];
    if (i >= 1 && i <= M - 2) {
      v[at(g, i, 0)] = 0.0;
      v[at(g, i, N)] = v[at(g, i, N - 1)];
    }
    if (do_p && i >= 1 && i <= M - 1) {
      p[at(g, i, 0)] = p[at(g, i, 1)];
      p[at(g, i, N)] = 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// Momentum assembly (momentum_solver.cpp:195-314) for one U or V DOF.
template <bool IsU>
struct MomCoef {
  double aS, aW, aP, aE, aN;  // positive link coefficients, aP = aP0/uvAlpha
  double b;                   // source with non-DOF neighbours folded in
  double aP0;
  bool dS, dW, dE, dN;        // neighbour is a DOF of the same system
};

template <bool IsU>
__device__ inline MomCoef<IsU> mom_coef(const Geo& g, const Fields& f, int i, int j) {
  const int M = g.M, N = g.N;
  auto U = [&](int a, int b) { return f.u[at(g, a, b)]; };
  auto V = [&](int a, int b) { return f.v[at(g, a, b)]; };
  auto A = [&](int a, int b) { return f.alp[at(g, a, b)]; };
  auto G = [&](int a, int b) { return f.gam[at(g, a, b)]; };
  const double mv = g.maxVel;
  double alphaLocal, ue, uw, vn, vs;
  if (IsU) {
    alphaLocal = 0.5 * (A(i - 1, j - 1) + A(i - 1, j));
    ue = clampv(0.5 * (U(i, j) + U(i, j + 1)), mv);
    uw = clampv(0.5 * (U(i, j - 1) + U(i, j)), mv);
    vn = clampv(0.5 * (V(i, j) + V(i, j + 1)), mv);
    vs = clampv(0.5 * (V(i - 1, j) + V(i - 1, j + 1)), mv);
  } else {
    alphaLocal = 0.5 * (A(i - 1, j - 1) + A(i, j - 1));
    ue = clampv(0.5 * (U(i, j) + U(i + 1, j)), mv);
    uw = clampv(0.5 * (U(i, j - 1) + U(i + 1, j - 1)), mv);
    vn = clampv(0.5 * (V(i, j) + V(i + 1, j)), mv);
    vs = clampv(0.5 * (V(i - 1, j) + V(i, j)), mv);
  }
  const double Fe = g.convScale * g.rho * g.hy * ue, Fw = g.convScale * g.rho * g.hy * uw;
  const double Fn = g.convScale * g.rho * g.hx * vn, Fs = g.convScale * g.rho * g.hx * vs;
  const double De = g.De, Dn = g.Dn;
  double aE = De + fmax(0.0, -Fe);
  double aW = De + fmax(0.0, Fw);
  double aN = Dn + fmax(0.0, -Fn);
  double aS = Dn + fmax(0.0, Fs);
  // half-cell wall diffusion (boundaries.cpp:148-206)
  if (IsU) {
    if (i == 1 && face_fluid(G(0, j - 1), G(0, j)) && fabs(U(0, j)) <= kVelTol &&
        fabs(V(0, j)) <= kVelTol && fabs(V(0, j + 1)) <= kVelTol)
      aS += Dn;
    if (i == M - 1 && face_fluid(G(M - 2, j - 1), G(M - 2, j)) && fabs(U(M, j)) <= kVelTol &&
        fabs(V(M - 1, j)) <= kVelTol && fabs(V(M - 1, j + 1)) <= kVelTol)
      aN += Dn;
  } else {
    if (j == 1 && face_fluid(G(i - 1, 0), G(i, 0)) && !(fabs(g.targetVel) > kVelTol) &&
        fabs(V(i, 0)) <= kVelTol && fabs(U(i, 0)) <= kVelTol && fabs(U(i + 1, 0)) <= kVelTol)
      aW += De;
    if (j == N - 1 && face_fluid(G(i - 1, N - 2), G(i, N - 2)) &&
        !(fabs(f.p[at(g, i, N)]) <= kVelTol) && fabs(V(i, N)) <= kVelTol && fabs(U(i, N - 1)) <= kVelTol &&
        fabs(U(i + 1, N - 1)) <= kVelTol)
      aE += De;
  }
  const double sumA = aE + aW + aN + aS;
  double aP0 = sumA + (Fe - Fw + Fn - Fs) + 0.0 + g.sinkDiag + alphaLocal * g.vol;
  aP0 = fmax(aP0, kAP0Floor);
  MomCoef<IsU> c;
  c.aP0 = aP0;
  c.aP = aP0 / g.uvAlpha;
  double Sp, velCur;
  if (IsU) {
    Sp = (f.p[at(g, i, j)] - f.p[at(g, i, j + 1)]) * g.hy;
    velCur = U(i, j);
  } else {
    Sp = (f.p[at(g, i, j)] - f.p[at(g, i + 1, j)]) * g.hx;
    velCur = V(i, j);
  }
  Sp += (1.0 - g.uvAlpha) / g.uvAlpha * aP0 * velCur;
  if (g.sou) {
    // convection.cpp:53-186
    double Sdc = 0.0;
    if (IsU) {
      if (Fe >= 0.0) { if (j >= 2) Sdc -= Fe * ((1.5 * U(i, j) - 0.5 * U(i, j - 1)) - U(i, j)); }
      else { if (j + 2 < N) Sdc -= Fe * ((1.5 * U(i, j + 1) - 0.5 * U(i, j + 2)) - U(i, j + 1)); }
      if (Fw >= 0.0) { if (j >= 2) Sdc += Fw * ((1.5 * U(i, j - 1) - 0.5 * U(i, j - 2)) - U(i, j - 1)); }
      else { if (j + 1 < N) Sdc += Fw * ((1.5 * U(i, j) - 0.5 * U(i, j + 1)) - U(i, j)); }
      if (Fn >= 0.0) { if (i >= 1) Sdc -= Fn * ((1.5 * U(i, j) - 0.5 * U(i - 1, j)) - U(i, j)); }
      else { if (i + 2 <= M) Sdc -= Fn * ((1.5 * U(i + 1, j) - 0.5 * U(i + 2, j)) - U(i + 1, j)); }
      if (Fs >= 0.0) { if (i >= 2) Sdc += Fs * ((1.5 * U(i - 1, j) - 0.5 * U(i - 2, j)) - U(i - 1, j)); }
      else { if (i + 1 <= M) Sdc += Fs * ((1.5 * U(i, j) - 0.5 * U(i + 1, j)) - U(i, j)); }
    } else {
      if (Fe >= 0.0) { if (j >= 2) Sdc -= Fe * ((1.5 * V(i, j) - 0.5 * V(i, j - 1)) - V(i, j)); }
      else { if (j + 2 <= N) Sdc -= Fe * ((1.5 * V(i, j + 1) - 0.5 * V(i, j + 2)) - V(i, j + 1)); }
      if (Fw >= 0.0) { if (j >= 2) Sdc += Fw * ((1.5 * V(i, j - 1) - 0.5 * V(i, j - 2)) - V(i, j - 1)); }
      else { if (j + 1 <= N) Sdc += Fw * ((1.5 * V(i, j) - 0.5 * V(i, j + 1)) - V(i, j)); }
      if (Fn >= 0.0) { if (i >= 1) Sdc -= Fn * ((1.5 * V(i, j) - 0.5 * V(i - 1, j)) - V(i, j)); }
      else { if (i + 2 < M) Sdc -= Fn * ((1.5 * V(i + 1, j) - 0.5 * V(i + 2, j)) - V(i + 1, j)); }
      if (Fs >= 0.0) { if (i >= 2) Sdc += Fs * ((1.5 * V(i - 1, j) - 0.5 * V(i - 2, j)) - V(i - 1, j)); }
      else { if (i + 1 < M) Sdc += Fs * ((1.5 * V(i, j) - 0.5 * V(i + 1, j)) - V(i, j)); }
    }
    Sp += Sdc;
  }
  // DOF neighbours (momentum_solver.cpp:98-111) and folding of the others
  // into the source, in the reference order E, W, N, S (:299-303)
  if (IsU) {
    c.dE = j + 1 <= N - 2; c.dW = j - 1 >= 1; c.dN = i + 1 <= M - 1; c.dS = i - 1 >= 1;
    if (!c.dE) Sp += aE * U(i, j + 1);
    if (!c.dW) Sp += aW * U(i, j - 1);
    if (!c.dN) Sp += aN * U(i + 1, j);
    if (!c.dS) Sp += aS * U(i - 1, j);
  } else {
    c.dE = j + 1 <= N - 1; c.dW = j - 1 >= 1; c.dN = i + 1 <= M - 2; c.dS = i - 1 >= 1;
    if (!c.dE) Sp += aE * V(i, j + 1);
    if (!c.dW) Sp += aW * V(i, j - 1);
    if (!c.dN) Sp += aN * V(i + 1, j);
    if (!c.dS) Sp += aS * V(i - 1, j);
  }
  c.aE = aE; c.aW = aW; c.aN = aN; c.aS = aS;
  c.b = Sp;
  return c;
}

// Writes the CSR-equivalent row: c_nb = -a_nb for DOF neighbours, 0 for
// folded ones (their value already sits in b).  Delta form: b - A x0 with
// x0 = current velocity, evaluated in FP64 in the reference sum order.
template <bool IsU, class C, bool Delta>
__device__ inline void store_mom(const Geo& g, const MomCoef<IsU>& c, const double* x, MomSys<C> s, int i, int j) {
  const size_t k = at(g, i, j);
  const double cS = c.dS ? -c.aS : 0.0, cW = c.dW ? -c.aW : 0.0;
  const double cE = c.dE ? -c.aE : 0.0, cN = c.dN ? -c.aN : 0.0;
  s.cS[k] = (C)cS; s.cW[k] = (C)cW; s.aP[k] = (C)c.aP; s.cE[k] = (C)cE; s.cN[k] = (C)cN;
  if (Delta) {
    double Ax = 0.0;
    if (c.dS) Ax += cS * x[k - g.P];
    if (c.dW) Ax += cW * x[k - 1];
    Ax += c.aP * x[k];
    if (c.dE) Ax += cE * x[k + 1];
    if (c.dN) Ax += cN * x[k + g.P];
    s.b[k] = (C)(c.b - Ax);
    s.x0[k] = (C)0;
  } else {
    s.b[k] = (C)c.b;
    s.x0[k] = (C)x[k];
  }
}

// Launched on the active-cell grid (tiles from cell (1, 1), one cell per
// thread, see active_grid): every U and V cell lies in it.
//
// hj.on: the block's tile is also a 64 x 16 tile of the pressure multigrid's
// level 0 (DOF (i, j) = level-0 cell (i-1, j-1)), whose faces come from the
// dE, dN just formed here, so the block also builds its tile of level 0 and
// the nested tiles of the next hj.H.n levels (hier_body's arithmetic; the
// 2x2 aggregation is local to 16-aligned tiles).
template <class C, bool Delta>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void k_assemble(Geo g, Fields f, MomSys<C> su, MomSys<C> sv, IterState* st,
                                                                                         mgc::HierJob<C> hj) {
  const size_t zo_ = zoff(g.zs);
  f.zshift(zo_); su.zshift(zo_); sv.zshift(zo_); st = zsh(st, zo_);
  const int lt = xcd_tile();  // tiles in XCD-aware order (the SOU stencil reads two-cell halos)
  const int j = 1 + lt % (int)gridDim.x * BX + (int)threadIdx.x;
  const int M = g.M, N = g.N;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    // the Jacobi loop's state for this step
    st->jac_active[0] = g.nU > 0;
    st->jac_active[1] = g.nV > 0;
    st->jac_iters[0] = 0;
    st->jac_iters[1] = 0;
  }
  const int i = 1 + lt / (int)gridDim.x * BY + (int)threadIdx.y;
  double dEv = 0.0, dNv = 0.0;
  if (i >= 1 && i <= M - 1 && j >= 1 && j <= N - 2) {
    MomCoef<true> c = mom_coef<true>(g, f, i, j);
    dEv = g.hy / c.aP0;
    f.dE[at(g, i, j)] = dEv;
    store_mom<true, C, Delta>(g, c, f.u, su, i, j);
  }
  if (i >= 1 && i <= M - 2 && j >= 1 && j <= N - 1) {
    MomCoef<false> c = mom_coef<false>(g, f, i, j);
    dNv = g.hx / c.aP0;
    f.dN[at(g, i, j)] = dNv;
    store_mom<false, C, Delta>(g, c, f.v, sv, i, j);
  }
  if (hj.on) {  // kernel-uniform
    constexpr int QN = (BX / 2) * (BY / 2);
    __shared__ C sEa[NT], sNa[NT], sEb[QN], sNb[QN];
    const int t = (int)threadIdx.x + (int)threadIdx.y * BX;
    const mgc::Lv<C>& L = hj.H.l[0];
    const int r = i - 1, c = j - 1;
    if (r < L.ny && c < L.nx) {  // level 0: aE = rho dE hy, aN = rho dN hx (0 on the last row)
      const double aE = g.rho * dEv * g.hy;
      const double aN = (i < M - 1) ? g.rho * dNv * g.hx : 0.0;
      const size_t k = L.at(r, c);
      zsh(L.ex, zo_)[k] = (C)aE;
      zsh(L.ey, zo_)[k] = (C)aN;
      sEa[t] = (C)aE;
      sNa[t] = (C)aN;
    }
    mgc::hier_coarsen<C, NT, BX, BY>(hj.H, lt % (int)gridDim.x, lt / (int)gridDim.x, t, zo_, sEa, sNa, sEb, sNb);
  }
}

// Row loop of the capped field grid: thread (tx, ty) of block (bx, by)
// visits rows by*BY + ty + k*gy*BY.
#define FIELD_LOOP(g, i, j)                                   \
  const int j = blockIdx.x * BX + threadIdx.x;                \
  for (int i = blockIdx.y * BY + threadIdx.y; i <= (g).M; i += (g).gy * BY)

__device__ inline int block_id() { return blockIdx.x + blockIdx.y * gridDim.x; }
__device__ inline int thread_id() { return threadIdx.x + threadIdx.y * BX; }

// Jacobi stop rule of jacobi_solver.cpp:63 applied to the partials of sweep
// k-1 (slot parity (k-1)&1): returns the updated active flags and records
// the sweep count of a component that just converged.
__device__ inline void jac_detect(const Geo& g, int k, double tol2, IterState* st, const double* part, int& au,
                                  int& av, double* sh) {
  if (k <= 0) return;
  const double* pin = part + (size_t)(SLOT_JAC + 2 * ((k - 1) & 1)) * g.NB;
  if (au) {
    const double a = parts_sum<NT>(pin, g.NB, sh);
    if (a < tol2 * (double)g.nU) {
      au = 0;
      if (block_id() == 0 && thread_id() == 0) { st->jac_active[0] = 0; st->jac_iters[0] = k; }
    }
  }
  if (av) {
    const double b = parts_sum<NT>(pin + g.NB, g.NB, sh);
    if (b < tol2 * (double)g.nV) {
      av = 0;
      if (block_id() == 0 && thread_id() == 0) { st->jac_active[1] = 0; st->jac_iters[1] = k; }
    }
  }
}

// ---------------------------------------------------------------------------
// Fused U+V Jacobi sweep k (jacobi_solver.cpp:44-67): r = b - A x (sorted
// column order S, W, P, E, N), x' = x + r / aP, sum r^2 per component into
// the parity-k slots.  The stop test of sweep k-1 runs first.
template <class C>
__global__ __launch_bounds__(NT) void k_jacobi(Geo g, MomSys<C> su, MomSys<C> sv, int sweep, double tol2,
                                               IterState* st, double* part) {
  const size_t zo_ = zoff(g.zs);
  su.zshift(zo_); sv.zshift(zo_); st = zsh(st, zo_); part = zsh(part, zo_);
  __shared__ double sh[NT / kWave + 1];
  __shared__ int sfl[2];
  // both components stopped in an earlier launch: nothing to load (uniform
  // scalar test; a flag cleared by block 0 in this launch means the same
  // decision in every block)
  if (!st->jac_active[0] && !st->jac_active[1]) return;
  const int t = thread_id();
  // flags (thread 0) and the partials of sweep-1 are loaded together
  const int f0 = t == 0 ? st->jac_active[0] : 0, f1 = t == 0 ? st->jac_active[1] : 0;
  const double* pin = part + (size_t)(SLOT_JAC + 2 * ((sweep + 1) & 1)) * g.NB;
  const double qu = sweep > 0 ? parts_load<NT>(pin, g.NB) : 0.0;
  const double qv = sweep > 0 ? parts_load<NT>(pin + g.NB, g.NB) : 0.0;
  if (t == 0) { sfl[0] = f0; sfl[1] = f1; }
  __syncthreads();
  int au = sfl[0], av = sfl[1];
  if (sweep > 0) {  // stop rule of jacobi_solver.cpp:63 on sweep-1
    const bool b0 = block_id() == 0 && t == 0;
    if (au) {
      const double a = parts_total<NT>(qu, sh);
      if (a < tol2 * (double)g.nU) { au = 0; if (b0) { st->jac_active[0] = 0; st->jac_iters[0] = sweep; } }
    }
    if (av) {
      const double b = parts_total<NT>(qv, sh);
      if (b < tol2 * (double)g.nV) { av = 0; if (b0) { st->jac_active[1] = 0; st->jac_iters[1] = sweep; } }
    }
  }
  if (!au && !av) return;  // uniform over the grid
  const int M = g.M, N = g.N, P = g.P;
  double ru = 0.0, rv = 0.0;
  auto sweep1 = [&](const MomSys<C>& s, size_t k) -> double {
    const C* x = (sweep & 1) ? s.x1 : s.x0;
    C* y = (sweep & 1) ? s.x0 : s.x1;
    C Ax = (C)0;
    Ax += s.cS[k] * x[k - P];
    Ax += s.cW[k] * x[k - 1];
    Ax += s.aP[k] * x[k];
    Ax += s.cE[k] * x[k + 1];
    Ax += s.cN[k] * x[k + P];
    const C r = s.b[k] - Ax;
    y[k] = x[k] + r / s.aP[k];
    return (double)r * (double)r;
  };
  FIELD_LOOP(g, i, j) {
    const size_t k = at(g, i, j);
    if (au && i >= 1 && i <= M - 1 && j >= 1 && j <= N - 2) ru += sweep1(su, k);
    if (av && i >= 1 && i <= M - 2 && j >= 1 && j <= N - 1) rv += sweep1(sv, k);
  }
  const double tu = block_sum<NT>(ru, sh);
  const double tv = block_sum<NT>(rv, sh);
  if (thread_id() == 0) {
    double* pout = part + (size_t)(SLOT_JAC + 2 * (sweep & 1)) * g.NB;
    pout[block_id()] = tu;
    pout[g.NB + block_id()] = tv;
  }
}

// After K launched sweeps: stop test of sweep K-1 and sweep counts of the
// components still iterating (one block).
__global__ __launch_bounds__(NT) void k_jac_end(Geo g, int K, double tol2, IterState* st, const double* part) {
  const size_t zo_ = zoff(g.zs);
  st = zsh(st, zo_); part = zsh(part, zo_);
  __shared__ double sh[NT / kWave + 1];
  __shared__ int shi;
  int au = block_read(&st->jac_active[0], &shi), av = block_read(&st->jac_active[1], &shi);
  jac_detect(g, K, tol2, st, const_cast<double*>(part), au, av, sh);
  if (thread_id() == 0) {
    if (au) st->jac_iters[0] = K;
    if (av) st->jac_iters[1] = K;
  }
}

// Two fused U+V Jacobi sweeps k and k+1 (jacobi_solver.cpp:44-67) in one
// launch: the delta form without host polling.  Blocks own 64 x 16 tiles
// from cell (1, 1), the first active one, so 1024 x 512 active cells are
// exactly 512 blocks, one resident wave of two 1024-thread blocks per CU.
// Sweep k runs on the tile and a one-cell ring, from x_k on a two-cell ring
// in LDS, and its result stays in LDS for sweep k+1 on the tile.  Every cell
// value is the same expression as in k_jacobi; the r^2 sums are the same
// terms, summed per block of this grid.  x_j lives in buffer j mod 3: the launch reads x_k and
// writes x_(k+1) and x_(k+2) to the two other buffers, so neighbouring
// blocks' halos, and a result the stop rule may still pick, stay intact.
// Stop rule first: the test of sweep k-2, then of sweep k-1 (the previous
// launch's partials; slots by sweep index mod 4, so this launch's own
// partials never overwrite ones its blocks are still reading).  two = 0
// runs sweep k only (an odd sweep cap).
constexpr int JR1W = BX + 2, JR1H = BY + 2, JR1N = JR1W * JR1H;  // sweep-k region
constexpr int JR2W = BX + 4, JR2H = BY + 4, JR2N = JR2W * JR2H;  // x_k region
constexpr int JRING = JR1N - BX * BY;                             // ring cells of sweep k
static_assert(JRING == 2 * JR1W + 2 * BY && JRING <= NT, "one ring cell per thread");
#ifndef CPS_JAC_ROWS
#define CPS_JAC_ROWS 2
#endif
// tile rows per thread of the paired sweeps (blocks of BX x BY / kJacRows
// threads: more blocks per CU, one block's loads overlapping another's sweeps)
constexpr int kJacRows = CPS_JAC_ROWS;

template <class C>
__device__ __forceinline__ C* jac_buf(const MomSys<C>& s, int j) {
  const int m = j % 3;
  C* const b0 = s.x0;  // a select of values (not of member addresses), so the
  C* const b1 = s.x1;  // struct can stay in registers
  C* const b2 = s.x2;
  return m == 0 ? b0 : m == 1 ? b1 : b2;
}

// r = b - A x (column order S, W, P, E, N) and x + r / aP at LDS cell q of a
// region of width W, as in k_jacobi
template <class C>
__device__ __forceinline__ C jac_cell(const C* X, int W, int q, const C* c, C& r) {
  C Ax = (C)0;
  Ax += c[0] * X[q - W];
  Ax += c[1] * X[q - 1];
  Ax += c[2] * X[q];
  Ax += c[3] * X[q + 1];
  Ax += c[4] * X[q + W];
  r = c[5] - Ax;
  return X[q] + r / c[2];
}

template <class C, int RT>
__global__ __launch_bounds__(NT / RT)
void k_jacobi2(Geo g, MomSys<C> su, MomSys<C> sv, int k, int two, double tol2, IterState* st, double* part) {
  constexpr int NTH = NT / RT, BYR = BY / RT, KP = 2 * NT / NTH;  // KP: partials per thread (<= 2048 tiles)
  static_assert(JRING <= NTH, "one ring cell per thread");
  const size_t zo_ = zoff(g.zs);
  su.zshift(zo_); sv.zshift(zo_); st = zsh(st, zo_); part = zsh(part, zo_);
  __shared__ C xa[2][JR2N];  // x_k, U and V
  __shared__ C xb[2][JR1N];  // x_(k+1)
  __shared__ double sh[NTH / kWave + 1];
  __shared__ int sfl[2];
  if (!st->jac_active[0] && !st->jac_active[1]) return;  // as in k_jacobi
  const int t = thread_id();
  const int f0 = t == 0 ? st->jac_active[0] : 0, f1 = t == 0 ? st->jac_active[1] : 0;
  const bool tst = k >= 2;
  const double* p2 = part + (size_t)(SLOT_JAC4 + 2 * ((k + 2) & 3)) * g.NB;  // sweep k-2
  const double* p1 = part + (size_t)(SLOT_JAC4 + 2 * ((k + 3) & 3)) * g.NB;  // sweep k-1
  const int nb = (int)(gridDim.x * gridDim.y);  // partials per slot (slot stride g.NB >= nb)
  const double qu2 = tst ? parts_load_k<NTH, KP>(p2, nb) : 0.0, qv2 = tst ? parts_load_k<NTH, KP>(p2 + g.NB, nb) : 0.0;
  const double qu1 = tst ? parts_load_k<NTH, KP>(p1, nb) : 0.0, qv1 = tst ? parts_load_k<NTH, KP>(p1 + g.NB, nb) : 0.0;
  // the flags as thread 0 found them, uniform over the block (a block that
  // reads a flag block 0 has already cleared takes the same decision: every
  // block runs the same test on the same partials)
  if (t == 0) { sfl[0] = f0; sfl[1] = f1; }
  __syncthreads();
  int au0 = sfl[0], av0 = sfl[1];
  if (!au0 && !av0) return;  // block-uniform
  // jacobi_solver.cpp:63 on sweep k-2, and on sweep k-1 if that sweep counts,
  // before any tile load: once the solve has converged (in most SIMPLE steps
  // after the first pair) the launch loads nothing
  if (tst) {
    const bool b0 = block_id() == 0 && t == 0;
    // the four sums in one reduction (one barrier), each equal to parts_total's
    __shared__ double sh4[4 * (NTH / kWave)];
    double q4[4] = {qu2, qu1, qv2, qv1};
    block_sumn_all<NTH, 4>(q4, sh4);
    if (au0) {
      if (q4[0] < tol2 * (double)g.nU) {
        au0 = 0;
        if (b0) { st->jac_active[0] = 0; st->jac_iters[0] = k - 1; }
      } else if (q4[1] < tol2 * (double)g.nU) {
        au0 = 0;
        if (b0) { st->jac_active[0] = 0; st->jac_iters[0] = k; }
      }
    }
    if (av0) {
      if (q4[2] < tol2 * (double)g.nV) {
        av0 = 0;
        if (b0) { st->jac_active[1] = 0; st->jac_iters[1] = k - 1; }
      } else if (q4[3] < tol2 * (double)g.nV) {
        av0 = 0;
        if (b0) { st->jac_active[1] = 0; st->jac_iters[1] = k; }
      }
    }
    if (!au0 && !av0) return;  // uniform over the grid
  }
  const int M = g.M, N = g.N, P = g.P;
  const int lt = xcd_tile();
  const int i0 = 1 + lt / (int)gridDim.x * BY, j0 = 1 + lt % (int)gridDim.x * BX;
  auto actU = [&](int i, int j) { return i >= 1 && i <= M - 1 && j >= 1 && j <= N - 2; };
  auto actV = [&](int i, int j) { return i >= 1 && i <= M - 2 && j >= 1 && j <= N - 1; };
  const C* xu = jac_buf(su, k);
  const C* xv = jac_buf(sv, k);
  // the RT tile cells and (t < JRING) one ring cell of sweep k, in region-1 coordinates
  const int tc = (int)threadIdx.x + 1;
  int rr = 0, rc = 0;
  if (t < JR1W) {
    rc = t;
  } else if (t < 2 * JR1W) {
    rr = JR1H - 1;
    rc = t - JR1W;
  } else {
    rr = 1 + ((t - 2 * JR1W) >> 1);
    rc = ((t - 2 * JR1W) & 1) ? JR1W - 1 : 0;
  }
  const bool ring = t < JRING;
  const int tj = j0 - 1 + tc, ri = i0 - 1 + rr, rj = j0 - 1 + rc;
  // every global load first, for the components still active after the tests
  int ti[RT];
  bool tu[RT], tv[RT];
#pragma unroll
  for (int o = 0; o < RT; ++o) {
    ti[o] = i0 - 1 + (int)threadIdx.y + 1 + o * BYR;
    tu[o] = au0 && actU(ti[o], tj);
    tv[o] = av0 && actV(ti[o], tj);
  }
  bool gu = ring && au0 && actU(ri, rj), gv = ring && av0 && actV(ri, rj);
  constexpr int NX = (JR2N + NTH - 1) / NTH;
  C vxu[NX], vxv[NX], cu[RT][6], cv[RT][6], du[6], dv[6];
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    const int q = t + i * NTH;
    const int lr = q / JR2W, lc = q - (q / JR2W) * JR2W;
    const int gi = i0 - 2 + lr, gj = j0 - 2 + lc;
    const bool in = q < JR2N && gi >= 0 && gi <= M && gj >= 0 && gj < P;
    const size_t kk = in ? (size_t)gi * P + gj : 0;
    vxu[i] = (in && au0) ? xu[kk] : (C)0;
    vxv[i] = (in && av0) ? xv[kk] : (C)0;
  }
  auto coefs = [&](const MomSys<C>& s, bool on, int i, int j, C* c) {
    const size_t kk = on ? (size_t)i * P + j : 0;
    c[0] = on ? s.cS[kk] : (C)0;
    c[1] = on ? s.cW[kk] : (C)0;
    c[2] = on ? s.aP[kk] : (C)1;
    c[3] = on ? s.cE[kk] : (C)0;
    c[4] = on ? s.cN[kk] : (C)0;
    c[5] = on ? s.b[kk] : (C)0;
  };
#pragma unroll
  for (int o = 0; o < RT; ++o) {
    coefs(su, tu[o], ti[o], tj, cu[o]);
    coefs(sv, tv[o], ti[o], tj, cv[o]);
  }
  if (RT == 1) {  // with more tile rows per thread the ring's loads wait for the tests (registers)
    coefs(su, gu, ri, rj, du);
    coefs(sv, gv, ri, rj, dv);
  }
  const int au = au0, av = av0;
  if (!au && !av) return;  // uniform over the grid
#pragma unroll
  for (int o = 0; o < RT; ++o) {
    tu[o] = tu[o] && au;
    tv[o] = tv[o] && av;
  }
  gu = gu && au;
  gv = gv && av;
  if (RT > 1) {
    coefs(su, gu, ri, rj, du);
    coefs(sv, gv, ri, rj, dv);
  }
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    const int q = t + i * NTH;
    if (q < JR2N) {
      xa[0][q] = au ? vxu[i] : (C)0;
      xa[1][q] = av ? vxv[i] : (C)0;
    }
  }
  __syncthreads();
  // sweep k: x_(k+1) on the tile (to memory) and the ring (LDS only)
  double ru = 0.0, rv = 0.0;
  const int qr2 = (rr + 1) * JR2W + rc + 1, qr1 = rr * JR1W + rc;
#pragma unroll
  for (int o = 0; o < RT; ++o) {
    const int tr = (int)threadIdx.y + 1 + o * BYR;
    const int qt2 = (tr + 1) * JR2W + tc + 1, qt1 = tr * JR1W + tc;
    const size_t kt = (size_t)ti[o] * P + tj;
    C r;
    if (tu[o]) {
      const C x = jac_cell(xa[0], JR2W, qt2, cu[o], r);
      xb[0][qt1] = x;
      jac_buf(su, k + 1)[kt] = x;
      ru += (double)r * (double)r;
    } else {
      xb[0][qt1] = xa[0][qt2];
    }
    if (tv[o]) {
      const C x = jac_cell(xa[1], JR2W, qt2, cv[o], r);
      xb[1][qt1] = x;
      jac_buf(sv, k + 1)[kt] = x;
      rv += (double)r * (double)r;
    } else {
      xb[1][qt1] = xa[1][qt2];
    }
  }
  if (ring) {
    C r;
    xb[0][qr1] = gu ? jac_cell(xa[0], JR2W, qr2, du, r) : xa[0][qr2];
    xb[1][qr1] = gv ? jac_cell(xa[1], JR2W, qr2, dv, r) : xa[1][qr2];
  }
  {
    const double su2 = block_sum<NTH>(ru, sh);
    const double sv2 = block_sum<NTH>(rv, sh);
    if (t == 0) {
      double* pout = part + (size_t)(SLOT_JAC4 + 2 * (k & 3)) * g.NB;
      pout[lt] = su2;
      pout[g.NB + lt] = sv2;
    }
  }
  if (!two) return;
  __syncthreads();
  // sweep k+1 on the tile
  ru = 0.0;
  rv = 0.0;
#pragma unroll
  for (int o = 0; o < RT; ++o) {
    const int tr = (int)threadIdx.y + 1 + o * BYR;
    const int qt1 = tr * JR1W + tc;
    const size_t kt = (size_t)ti[o] * P + tj;
    C r;
    if (tu[o]) {
      jac_buf(su, k + 2)[kt] = jac_cell(xb[0], JR1W, qt1, cu[o], r);
      ru += (double)r * (double)r;
    }
    if (tv[o]) {
      jac_buf(sv, k + 2)[kt] = jac_cell(xb[1], JR1W, qt1, cv[o], r);
      rv += (double)r * (double)r;
    }
  }
  const double su2 = block_sum<NTH>(ru, sh);
  const double sv2 = block_sum<NTH>(rv, sh);
  if (t == 0) {
    double* pout = part + (size_t)(SLOT_JAC4 + 2 * ((k + 1) & 3)) * g.NB;
    pout[lt] = su2;
    pout[g.NB + lt] = sv2;
  }
}

// ---------------------------------------------------------------------------
// uStar / vStar = clamp(solution) (momentum_solver.cpp:335-351), their BCs
// (iterations.cpp:169) and the change sums of iterations.cpp:499-533.
// K >= 0: the sweep counts are settled here (K sweeps were launched and no
// k_jac_end ran): a component still active after sweep K-1 has done K
// sweeps whatever that sweep's stop test gives.  pairs: the sweeps ran as
// k_jacobi2 pairs (x_j in buffer j mod 3) on a grid of `pairs` blocks; if
// the last launch was a pair, its second sweep counts only if the first
// one's stop test failed.
//
// cj.on: the grid has one extra column of blocks (blockIdx.x == g.gx), and
// block (g.gx, 0) runs the coarse set-up of the coming multigrid pressure
// solve (faces, inverse diagonals and last-level inverse, from the grid
// levels' faces that dE, dN gave) beside the finalize tiles, which touch
// none of its data; the column's other blocks return at once.  The tiles'
// block index stays bx + by gx, so their partials are those of the plain
// grid.
template <class C, bool Delta>
__global__ __launch_bounds__(NT) void k_finalize(Geo g, Fields f, MomSys<C> su, MomSys<C> sv, IterState* st,
                                                 double* part, int K, int pairs, double tol2, mgc::CoarseJob<C> cj) {
  const size_t zo_ = zoff(g.zs);
  if ((int)blockIdx.x >= g.gx) {  // block-uniform
    if (cj.on && blockIdx.y == 0) {
      // shifted copies of the two level views only: the descriptor stays in
      // the argument segment (its arrays are indexed by level)
      mgc::Lv<C> top = cj.top, par = cj.par;
      top.zshift(zo_);
      par.zshift(zo_);
      extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
      mgc::coarse_setup_body<C>(top, par, cj.d, zsh(cj.cf, zo_), smem_raw, thread_id());
    }
    return;
  }
  f.zshift(zo_); su.zshift(zo_); sv.zshift(zo_); st = zsh(st, zo_); part = zsh(part, zo_);
  __shared__ double sh[NT / kWave + 1];
  const int bid = (int)blockIdx.x + (int)blockIdx.y * g.gx;
  const int M = g.M, N = g.N;
  const double mv = g.maxVel;
  double su2 = 0.0, sv2 = 0.0, umx = 0.0, vmx = 0.0;
  int itU = st->jac_iters[0], itV = st->jac_iters[1];
  if (K >= 0) {
    const int au = st->jac_active[0], av = st->jac_active[1];
    if (pairs && K >= 2 && !(K & 1)) {  // jacobi_solver.cpp:63 on sweep K-2
      const double* p2 = part + (size_t)(SLOT_JAC4 + 2 * ((K + 2) & 3)) * g.NB;
      double q2[2] = {au ? parts_load<NT>(p2, pairs) : 0.0, av ? parts_load<NT>(p2 + g.NB, pairs) : 0.0};
      __shared__ double sh2[2 * (NT / kWave)];  // not sh: block_sum below writes it with no barrier between
      block_sumn_all<NT, 2>(q2, sh2);
      if (au) itU = q2[0] < tol2 * (double)g.nU ? K - 1 : K;
      if (av) itV = q2[1] < tol2 * (double)g.nV ? K - 1 : K;
    } else {
      if (au) itU = K;
      if (av) itV = K;
    }
    if (bid == 0 && thread_id() == 0) { st->jac_iters[0] = itU; st->jac_iters[1] = itV; }
  }
  FIELD_LOOP(g, i, j) {
    if (j > N) continue;
    const size_t k = at(g, i, j);
    if (i >= 1 && i <= M - 1 && j >= 1 && j <= N - 2) {
      const C* x = pairs ? jac_buf(su, itU) : (itU & 1) ? su.x1 : su.x0;
      const double cur = f.u[k];
      const double nv = clampv(Delta ? cur + (double)x[k] : (double)x[k], mv);
      f.us[k] = nv;
      const double d = nv - cur;
      su2 += d * d;
      umx = fmax(umx, fabs(d));
      if (j == N - 2) f.us[k + 1] = nv;  // outlet copy u(i,N-1) = u(i,N-2)
    }
    if (i >= 1 && i <= M - 2 && j >= 1 && j <= N - 1) {
      const C* x = pairs ? jac_buf(sv, itV) : (itV & 1) ? sv.x1 : sv.x0;
      const double cur = f.v[k];
      const double nv = clampv(Delta ? cur + (double)x[k] : (double)x[k], mv);
      f.vs[k] = nv;
      const double d = nv - cur;
      sv2 += d * d;
      vmx = fmax(vmx, fabs(d));
      if (j == N - 1) f.vs[k + 1] = nv;  // outlet copy v(i,N) = v(i,N-1)
    }
    // walls / inlet of uStar, vStar (boundaries.cpp:59-106)
    if ((i == 0 || i == M) && j >= 1 && j <= N - 1) f.us[k] = 0.0;
    if (j == 0) f.us[k] = f.inlet[i];
    if (i <= M - 1 && (i == 0 || i == M - 1)) f.vs[k] = 0.0;
    if (i >= 1 && i <= M - 2 && j == 0) f.vs[k] = 0.0;
  }
  const double a = block_sum<NT>(su2, sh), b = block_sum<NT>(sv2, sh);
  const double c = block_max<NT>(umx, sh), d = block_max<NT>(vmx, sh);
  if (thread_id() == 0) {
    double* p = part + (size_t)SLOT_FIN * g.NB + bid;
    p[0] = a;
    p[g.NB] = b;
    p[2 * g.NB] = c;
    p[3 * g.NB] = d;
  }
}

// ---------------------------------------------------------------------------
// Pressure operator pieces from dE, dN (pressure_solver.cpp:589-594).
struct PCoef {
  double aS, aW, aP, aE, aN;
};
__device__ inline PCoef p_coef(const Geo& g, const double* dE, const double* dN, int i, int j) {
  const int M = g.M;
  PCoef c;
  c.aE = g.rho * dE[at(g, i, j)] * g.hy;
  c.aW = (j > 1) ? g.rho * dE[at(g, i, j - 1)] * g.hy : 0.0;
  c.aN = (i < M - 1) ? g.rho * dN[at(g, i, j)] * g.hx : 0.0;
  c.aS = (i > 1) ? g.rho * dN[at(g, i - 1, j)] * g.hx : 0.0;
  double aP = c.aE + c.aW + c.aN + c.aS;
  if (aP < kTiny20) aP = 1.0;
  c.aP = aP;
  return c;
}
__device__ inline double mass_b(const Geo& g, const Fields& f, int i, int j) {
  const double mE = g.rho * f.us[at(g, i, j)] * g.hy;
  const double mW = g.rho * f.us[at(g, i, j - 1)] * g.hy;
  const double mN = g.rho * f.vs[at(g, i, j)] * g.hx;
  const double mS = g.rho * f.vs[at(g, i - 1, j)] * g.hx;
  return mW - mE + mS - mN;
}

// Mass residual (j = 1..N-1) + RHS and PCG start (x0 = 0, r = b,
// z = r/diag, p_old = 0) for the DOFs j = 1..N-2.  pStar is cleared.
// PcgStart: also the Jacobi-PCG start (diagonal, z = r / diag, cleared
// search directions, r.z).  The SOR and multigrid solves read only r, so
// their steps skip those stores and the pressure coefficients.
//
// sj.on (multigrid solves without host polling): the right side goes
// straight into the solve's start state instead of q.r: sj.r = (C) b,
// sj.x = 0, sj.keep = b (FP64), and block 0 clears the solve's stop flags,
// the work of the solve's own start kernel.  The level-0 layout of the
// multigrid is the field layout, so DOF (i, j) is element k of both.
template <bool PcgStart, class C>
__global__ __launch_bounds__(NT) void k_prhs(Geo g, Fields f, PcgBufs q, double* part, mgc::StartJob<C> sj) {
  const size_t zo_ = zoff(g.zs);
  f.zshift(zo_); q.zshift(zo_); part = zsh(part, zo_);
  __shared__ double sh[NT / kWave + 1];
  const int M = g.M, N = g.N;
  C* const sr = zsh(sj.r, zo_);
  C* const sx = zsh(sj.x, zo_);
  double* const skeep = zsh(sj.keep, zo_);
  if (sj.on && block_id() == 0 && thread_id() == 0) {
    mgc::PcgScal* ps = zsh(sj.ps, zo_);
    ps->stop = 0;
    ps->iters = 0;
    if (sj.reset) {
      mgc::PcgScal* rs = zsh(sj.reset, zo_);
      rs->stop = 0;
      rs->iters = 0;
    }
  }
  double mx = 0, ss = 0, rz = 0, bb = 0;
  FIELD_LOOP(g, i, j) {
    if (j > N) continue;
    const size_t k = at(g, i, j);
    f.ps[k] = 0.0;
    if (i >= 1 && i <= M - 1 && j >= 1 && j <= N - 1) {
      const double b = mass_b(g, f, i, j);
      const double ab = fabs(b);
      mx = fmax(mx, ab);
      ss += ab * ab;
      if (j <= N - 2) {
        if (sj.on) {
          if (skeep) skeep[k] = b;
          sr[k] = (C)b;
          sx[k] = (C)0;
        } else {
          q.r[k] = b;
        }
        bb += b * b;
        if (PcgStart) {
          PCoef c = p_coef(g, f.dE, f.dN, i, j);
          q.diag[k] = c.aP;
          const double z = (fabs(c.aP) > kTiny20) ? b / c.aP : b;
          q.z[k] = z;
          q.pa[k] = 0.0;
          q.pb[k] = 0.0;
          rz += b * z;
        }
      }
    }
  }
  const double A = block_max<NT>(mx, sh), B = block_sum<NT>(ss, sh);
  const double Cc = block_sum<NT>(rz, sh), D = block_sum<NT>(bb, sh);
  if (thread_id() == 0) {
    double* p = part + (size_t)SLOT_RHS * g.NB + block_id();
    p[0] = A;
    p[g.NB] = B;
    p[2 * g.NB] = Cc;
    p[3 * g.NB] = D;
  }
}

// One block: residuals of the step (iterations.cpp:499-533, 628-658,
// pressure_solver.cpp:489-495) and the PCG start scalars.  All eight
// reductions share one pass over the block partials and one barrier; per
// reduction the order is that of parts_sum / parts_max (thread t adds
// p[t], p[t + NT], ..., then the wavefronts in order), so the results are
// the same.
__global__ __launch_bounds__(NT) void k_stats(Geo g, IterState* st, const double* part, int reset_pcg) {
  const size_t zo_ = zoff(g.zs);
  st = zsh(st, zo_); part = zsh(part, zo_);
  constexpr int NW = NT / kWave;
  __shared__ double sh[8][NW];
  const int NB = g.NB, t = thread_id();
  const double* pf = part + (size_t)SLOT_FIN * NB;
  const double* pr = part + (size_t)SLOT_RHS * NB;
  // su, sv, umax, vmax, mass max, mass sum of squares, r.z, ||b||^2
  double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int q = t; q < NB; q += NT) {
    v[0] += pf[q];
    v[1] += pf[NB + q];
    v[2] = fmax(v[2], pf[2 * NB + q]);
    v[3] = fmax(v[3], pf[3 * NB + q]);
    v[4] = fmax(v[4], pr[q]);
    v[5] += pr[NB + q];
    v[6] += pr[2 * NB + q];
    v[7] += pr[3 * NB + q];
  }
  constexpr bool kMax[8] = {false, false, true, true, true, false, false, false};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const double w = kMax[k] ? wave_max(v[k]) : wave_sum(v[k]);
    if ((t & (kWave - 1)) == 0) sh[k][t / kWave] = w;
  }
  __syncthreads();
  if (t != 0) return;
  double r[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    double a = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) a = kMax[k] ? (a > sh[k][w] ? a : sh[k][w]) : a + sh[k][w];
    r[k] = a;
  }
  const double su = r[0], sv = r[1], um = r[2], vm = r[3], mm = r[4], ms = r[5], rz = r[6], bb = r[7];
  {
    st->su = su;
    st->sv = sv;
    st->umax = um;
    st->vmax = vm;
    st->resid_u = g.nU > 0 ? sqrt(su / (double)g.nU) : 0.0;
    st->resid_v = g.nV > 0 ? sqrt(sv / (double)g.nV) : 0.0;
    const double cnt = (double)(g.M - 1) * (double)(g.N - 1);
    st->mass_max = mm;
    st->mass_ss = ms;
    st->resid_mass = cnt > 0 ? sqrt(ms / cnt) : 0.0;
    st->error = fmax(st->resid_u, fmax(st->resid_v, st->resid_mass));
    st->rz0 = rz;
    st->bnorm2 = bb;
    if (reset_pcg) {
      st->pcg_stop = 0;
      st->pcg_iters = 0;
      st->err = bb;
    }
  }
}

// PCG part A (pressure_solver.cpp:153-159, 168-181 of the previous loop
// body): stop test on ||r|| of iteration it-1, beta = rz_new / rz_old,
// p = z + beta p_old fused into Ap = A p (CSR sorted order S, W, P, E, N).
__global__ __launch_bounds__(NT) void k_pcg_a(Geo g, const double* dE, const double* dN, PcgBufs q, int it,
                                             double tol, IterState* st, double* part) {
  const size_t zo_ = zoff(g.zs);
  dE = zsh(dE, zo_); dN = zsh(dN, zo_); q.zshift(zo_); st = zsh(st, zo_); part = zsh(part, zo_);
  __shared__ double sh[NT / kWave + 1];
  __shared__ int shi;
  if (block_read(&st->pcg_stop, &shi)) return;
  double* pp = part + (size_t)SLOT_PCG * g.NB;
  double rzc, beta = 0.0;
  if (it == 0) {
    rzc = st->rz0;
  } else {
    const double err = parts_sum<NT>(pp + g.NB, g.NB, sh);
    if (sqrt(err) < tol) {  // pressure_solver.cpp:171
      if (block_id() == 0 && thread_id() == 0) { st->pcg_stop = 1; st->pcg_iters = it - 1; st->err = err; }
      return;
    }
    rzc = parts_sum<NT>(pp + 2 * g.NB, g.NB, sh);
    beta = rzc / st->rzs[(it - 1) & 1];
  }
  if (block_id() == 0 && thread_id() == 0) { st->rzs[it & 1] = rzc; st->pcg_iters = it; }
  const int M = g.M, N = g.N, P = g.P;
  const double* po = (it & 1) ? q.pb : q.pa;
  double* pn = (it & 1) ? q.pa : q.pb;
  double pap = 0.0;
  FIELD_LOOP(g, i, j) {
    if (!(i >= 1 && i <= M - 1 && j >= 1 && j <= N - 2)) continue;
    const size_t k = at(g, i, j);
    PCoef c = p_coef(g, dE, dN, i, j);
    const double pc = q.z[k] + beta * po[k];
    double Ap = 0.0;
    if (i > 1) Ap += -c.aS * (q.z[k - P] + beta * po[k - P]);
    if (j > 1) Ap += -c.aW * (q.z[k - 1] + beta * po[k - 1]);
    Ap += c.aP * pc;
    if (j + 1 <= N - 2) Ap += -c.aE * (q.z[k + 1] + beta * po[k + 1]);
    if (i < M - 1) Ap += -c.aN * (q.z[k + P] + beta * po[k + P]);
    pn[k] = pc;
    q.Ap[k] = Ap;
    pap += pc * Ap;
  }
  const double s = block_sum<NT>(pap, sh);
  if (thread_id() == 0) pp[block_id()] = s;
}

// PCG part B (pressure_solver.cpp:158-170): |pAp| test, alpha = rz / pAp,
// x += alpha p, r -= alpha Ap, z = r / diag, partial ||r||^2 and r.z.
__global__ __launch_bounds__(NT) void k_pcg_b(Geo g, double* x, PcgBufs q, int it, IterState* st, double* part) {
  const size_t zo_ = zoff(g.zs);
  x = zsh(x, zo_); q.zshift(zo_); st = zsh(st, zo_); part = zsh(part, zo_);
  __shared__ double sh[NT / kWave + 1];
  __shared__ int shi;
  if (block_read(&st->pcg_stop, &shi)) return;
  double* pp = part + (size_t)SLOT_PCG * g.NB;
  const double pap = parts_sum<NT>(pp, g.NB, sh);
  if (fabs(pap) < kTiny20) {  // pressure_solver.cpp:158
    if (block_id() == 0 && thread_id() == 0) { st->pcg_stop = 1; st->pcg_iters = it; }
    return;
  }
  const double alpha = st->rzs[it & 1] / pap;
  const int M = g.M, N = g.N;
  const double* pn = (it & 1) ? q.pa : q.pb;
  double e = 0.0, rz = 0.0;
  FIELD_LOOP(g, i, j) {
    if (!(i >= 1 && i <= M - 1 && j >= 1 && j <= N - 2)) continue;
    const size_t k = at(g, i, j);
    x[k] += alpha * pn[k];
    const double r = q.r[k] - alpha * q.Ap[k];
    q.r[k] = r;
    const double d = q.diag[k];
    const double z = (fabs(d) > kTiny20) ? r / d : r;
    q.z[k] = z;
    e += r * r;
    rz += r * z;
  }
  const double A = block_sum<NT>(e, sh), B = block_sum<NT>(rz, sh);
  if (thread_id() == 0) {
    pp[g.NB + block_id()] = A;
    pp[2 * g.NB + block_id()] = B;
  }
}

// After K launched PCG iterations: stop test of iteration K-1 (one block).
__global__ __launch_bounds__(NT) void k_pcg_end(Geo g, int K, double tol, IterState* st, const double* part) {
  const size_t zo_ = zoff(g.zs);
  st = zsh(st, zo_); part = zsh(part, zo_);
  __shared__ double sh[NT / kWave + 1];
  __shared__ int shi;
  if (block_read(&st->pcg_stop, &shi)) return;
  const double err = parts_sum<NT>(part + (size_t)(SLOT_PCG + 1) * g.NB, g.NB, sh);
  if (thread_id() == 0) {
    st->err = err;
    if (sqrt(err) < tol) { st->pcg_stop = 1; st->pcg_iters = K - 1; }
    else st->pcg_iters = K;
  }
}

// ---------------------------------------------------------------------------
// Red-black SOR pressure solve (pressureSolverType 0, pressure_solver.cpp:
// 819-919): cells of one colour ((i + j) % 2) over i = 1..M-1, j = 1..N-1
// read only the other colour, so a colour is one launch.  Column N-1 and
// cells with aP < 1e-20 are pinned to 0; p' is clamped to +-1e4; the stop is
// max |change| < pTol over both colours of an iteration.  Each launch writes
// its per-block max |change| into the slots of parity it & 1; colour 0 of
// iteration it first tests the slots of it - 1 (the other parity, so no
// block reads a slot this launch writes).  The right side b is the mass
// residual k_prhs left in q.r; pStar starts at 0 (cleared by k_prhs).
__global__ __launch_bounds__(NT) void k_sor(Geo g, Fields f, const double* b, int it, int colour, double omega,
                                           double tol, IterState* st, double* part) {
  const size_t zo_ = zoff(g.zs);
  f.zshift(zo_); b = zsh(b, zo_); st = zsh(st, zo_); part = zsh(part, zo_);
  __shared__ double sh[NT / kWave + 1];
  __shared__ int shi;
  if (block_read(&st->pcg_stop, &shi)) return;
  if (colour == 0 && it > 0) {
    const double* pv = part + (size_t)(SLOT_SOR + 2 * ((it - 1) & 1)) * g.NB;
    const double mc = fmax(parts_max<NT>(pv, g.NB, sh), parts_max<NT>(pv + g.NB, g.NB, sh));
    if (mc < tol) {
      if (block_id() == 0 && thread_id() == 0) {
        st->pcg_stop = 1;
        st->pcg_iters = it;
      }
      return;
    }
  }
  if (colour == 0 && block_id() == 0 && thread_id() == 0) st->pcg_iters = it + 1;
  const int M = g.M, N = g.N, P = g.P;
  double* ps = f.ps;
  double mx = 0.0;
  FIELD_LOOP(g, i, j) {
    if (!(i >= 1 && i <= M - 1 && j >= 1 && j <= N - 1) || ((i + j) & 1) != colour) continue;
    const size_t k = at(g, i, j);
    if (j == N - 1) {
      ps[k] = 0.0;
      continue;
    }
    const PCoef c = p_coef(g, f.dE, f.dN, i, j);
    const double aP = c.aE + c.aW + c.aN + c.aS;
    if (aP < kTiny20) {
      ps[k] = 0.0;
      continue;
    }
    const double pE = ps[k + 1];
    const double pW = (j > 1) ? ps[k - 1] : 0.0;
    const double pN = (i < M - 1) ? ps[k + P] : 0.0;
    const double pS = (i > 1) ? ps[k - P] : 0.0;
    const double pGS = (c.aE * pE + c.aW * pW + c.aN * pN + c.aS * pS + b[k]) / aP;
    const double pOld = ps[k];
    const double pNew = fmax(-1e4, fmin(1e4, pOld + omega * (pGS - pOld)));
    mx = fmax(mx, fabs(pNew - pOld));
    ps[k] = pNew;
  }
  mx = block_max<NT>(mx, sh);
  if (thread_id() == 0) part[(size_t)(SLOT_SOR + 2 * (it & 1) + colour) * g.NB + block_id()] = mx;
}

// After K launched SOR iterations: stop test of iteration K-1 (one block).
__global__ __launch_bounds__(NT) void k_sor_end(Geo g, int K, double tol, IterState* st, const double* part) {
  const size_t zo_ = zoff(g.zs);
  st = zsh(st, zo_); part = zsh(part, zo_);
  __shared__ double sh[NT / kWave + 1];
  __shared__ int shi;
  if (block_read(&st->pcg_stop, &shi)) return;
  const double* pv = part + (size_t)(SLOT_SOR + 2 * ((K - 1) & 1)) * g.NB;
  const double mc = fmax(parts_max<NT>(pv, g.NB, sh), parts_max<NT>(pv + g.NB, g.NB, sh));
  if (thread_id() == 0) {
    if (mc < tol) st->pcg_stop = 1;
    st->pcg_iters = K;
  }
}

// ---------------------------------------------------------------------------
// Corrections (iterations.cpp:556-600) with the BCs of u, v, p fused: the
// thread that owns an interior value also writes the ghost copied from it.
__global__ __launch_bounds__(NT) void k_correct(Geo g, Fields f) {
  const size_t zo_ = zoff(g.zs);
  f.zshift(zo_);
  const int j = blockIdx.x * BX + threadIdx.x;
  const int M = g.M, N = g.N, P = g.P;
  if (j > N) return;
  for (int i = blockIdx.y * BY + threadIdx.y; i <= M; i += g.gy * BY) {
  const size_t k = at(g, i, j);
  const double mv = g.maxVel;
  if (i >= 1 && i <= M - 1 && j >= 1 && j <= N - 1) {
    const double pn = f.p[k] + g.pAlpha * f.ps[k];
    f.p[k] = pn;
    if (i == 1) f.p[k - P] = pn;        // p(0,j) = p(1,j)
    if (i == M - 1) f.p[k + P] = pn;    // p(M,j) = p(M-1,j)
    if (j == 1) f.p[k - 1] = pn;        // p(i,0) = p(i,1)
    if (j == N - 1) f.p[k + 1] = 0.0;   // p(i,N) = 0
  }
  if (i >= 1 && i <= M - 1 && j >= 1 && j <= N - 2) {
    double du = f.dE[k] * (f.ps[k] - f.ps[k + 1]);
    du = clampv(du, mv);
    const double nu = clampv(f.us[k] + du, mv);
    f.u[k] = nu;
    if (j == N - 2) f.u[k + 1] = nu;
  }
  if (i >= 1 && i <= M - 2 && j >= 1 && j <= N - 1) {
    double dv = f.dN[k] * (f.ps[k] - f.ps[k + P]);
    dv = clampv(dv, mv);
    const double nv = clampv(f.vs[k] + dv, mv);
    f.v[k] = nv;
    if (j == N - 1) f.v[k + 1] = nv;
  }
  if ((i == 0 || i == M) && j >= 1 && j <= N - 1) f.u[k] = 0.0;
  if (j == 0) f.u[k] = f.inlet[i];
  if (i <= M - 1 && (i == 0 || i == M - 1)) f.v[k] = 0.0;
  if (i >= 1 && i <= M - 2 && j == 0) f.v[k] = 0.0;
  }
}

// Columns of one pressure-drop sample.  A boundary patch
// (postprocessing.cpp:245-327) reads one column of each array (left = right,
// t = 0); a plane (samplePlaneAtX, postprocessing.cpp:102-243) lerps between
// the two cell columns that bracket x with weight t on the right one.
struct SampleCols {
  int gl, gr;  // gamma columns of the pure-fluid row test
  int pl, pr;  // p columns
  int ul, ur;  // u columns
  int vl, vr;  // v columns (rows i - 1 and i)
  double t;
};
struct SampleSet {
  SampleCols c[4];
};

// Sums (area, |mass flux|, static * area, dynamic * area, total * |flux|) of
// one sample per block, over rows i = 1..M-1 whose cells are pure fluid.
__global__ void k_sample(Geo g, Fields f, double tol, SampleSet set, double* out) {
  __shared__ double sh[256 / kWave];
  const SampleCols c = set.c[blockIdx.x];
  const int M = g.M;
  const double w = 1.0 - c.t;
  double sA = 0, dA = 0, tw = 0, area = 0, flux = 0;
  for (int i = 1 + threadIdx.x; i < M; i += blockDim.x) {
    const int row = i - 1;
    if (!pure_fluid(f.gam[at(g, row, c.gl)], tol) || !pure_fluid(f.gam[at(g, row, c.gr)], tol)) continue;
    const double fa = g.hy;
    area += fa;
    const int vTop = i - 1, vBot = min(i, M - 1);
    const double sp = w * f.p[at(g, i, c.pl)] + c.t * f.p[at(g, i, c.pr)];
    const double un = w * f.u[at(g, i, c.ul)] + c.t * f.u[at(g, i, c.ur)];
    const double vt = w * f.v[at(g, vTop, c.vl)] + c.t * f.v[at(g, vTop, c.vr)];
    const double vb = w * f.v[at(g, vBot, c.vl)] + c.t * f.v[at(g, vBot, c.vr)];
    const double vtan = 0.5 * (vt + vb);
    const double vm = sqrt(un * un + vtan * vtan);
    const double dyn = 0.5 * g.rho * vm * vm;
    sA += sp * fa;
    dA += dyn * fa;
    const double fl = fabs(g.rho * un * fa);
    if (fl > 0.0) { flux += fl; tw += (sp + dyn) * fl; }
  }
  sA = block_sum<256>(sA, sh); dA = block_sum<256>(dA, sh); tw = block_sum<256>(tw, sh);
  area = block_sum<256>(area, sh); flux = block_sum<256>(flux, sh);
  double* o = out + 5 * blockIdx.x;
  if (threadIdx.x == 0) { o[0] = area; o[1] = flux; o[2] = sA; o[3] = dA; o[4] = tw; }
}

// inlet = s * inlet0 (the ramped inlet of SIMPLE.cpp:381-386)
__global__ void k_inlet_scale(int n, double s, const double* __restrict__ inlet0, double* __restrict__ inlet) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) inlet[i] = s * inlet0[i];
}

// Averages of one sample from its k_sample sums (postprocessing.cpp:226-240).
__host__ __device__ inline void finish_sample(const double* h, cps_patch* out) {
  out->area = h[0];
  out->mass_flux = h[1];
  out->avg_static = h[0] > 0 ? h[2] / h[0] : 0.0;
  out->avg_dynamic = h[0] > 0 ? h[3] / h[0] : 0.0;
  out->valid = h[0] > 0;
  out->avg_total = h[1] > 0 ? h[4] / h[1] : (h[0] > 0 ? out->avg_static + out->avg_dynamic : 0.0);
}

// The statistics the host reads after an iteration (iterations.cpp:628-658,
// SIMPLE.cpp:421-439), from the step's IterState.
__host__ __device__ inline void fill_stats(const IterState& h, int iteration, cps_step_stats* out) {
  out->error = h.error;
  out->resid_u = h.resid_u;
  out->resid_v = h.resid_v;
  out->resid_mass = h.resid_mass;
  out->resid_u_max = h.umax;
  out->resid_v_max = h.vmax;
  out->resid_mass_max = h.mass_max;
  out->u_iters = h.jac_iters[0];
  out->v_iters = h.jac_iters[1];
  out->p_iters = h.pcg_iters;
  out->iteration = iteration;
  cps_patch m[4];
  for (int k = 0; k < 4; ++k) finish_sample(h.mon[k], &m[k]);
  out->dp_core_total = m[0].avg_total - m[1].avg_total;
  out->dp_core_static = m[0].avg_static - m[1].avg_static;
  out->dp_full_total = m[2].avg_total - m[3].avg_total;
  out->dp_full_static = m[2].avg_static - m[3].avg_static;
  out->dp_converged = 0;
}

// Outer-loop state of runIterations kept on the device (SIMPLE.cpp:378-491).
struct RunState {
  int stopped;     // the loop condition failed after iteration stop_it - 1
  int stop_it;     // iterations the loop runs
  int conv;        // stopped on the settled dP window
  int nonfinite;   // the residual was not finite
  int idx, filled; // window ring position, window full
  int sn, spos;    // slope-gate entries held, next position
  cps_step_stats last;  // statistics of iteration stop_it - 1
};
struct RunParams {
  int max_iterations, ramp_steps, use_dp, win, use_slope, slope_win, log_cap, pad_;
  double epsilon, dp_percent, slope_max_delta;
};

__global__ void k_run_reset(RunState* rs) {
  if (threadIdx.x == 0) {
    RunState z{};
    *rs = z;
  }
}

// One thread, after iteration `it` (its k_sample and k_stats are done): the
// log row, the slope-gate ring, the dP window and the loop condition of
// SIMPLE.cpp:378, in the host loop's operation order (no contraction, sums in
// the same order), so the stop iteration is the host loop's.  Nothing
// changes once stopped.
__global__ void k_run_monitor(int it, int iteration, RunParams rp, const IterState* st, RunState* rs, double* hist,
                              double* slx, double* sly, cps_iter_log* log) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || rs->stopped) return;
  cps_step_stats sst;
  fill_stats(*st, iteration, &sst);
  const double core = sst.dp_core_total;
  bool stop = false, conv = false;
  if (!isfinite(sst.error)) {
    rs->nonfinite = 1;
    stop = true;
  } else {
    if (log && it < rp.log_cap) {
      cps_iter_log& r = log[it];
      r.resid_mass = sst.resid_mass;
      r.resid_u = sst.resid_u;
      r.resid_v = sst.resid_v;
      r.dp_core_total = sst.dp_core_total;
      r.dp_full_total = sst.dp_full_total;
      r.dp_core_static = sst.dp_core_static;
      r.dp_full_static = sst.dp_full_static;
      r.p_iters = sst.p_iters;
      r.pad_ = 0;
    }
    if (rp.use_slope && rp.slope_win > 0) {
      slx[rs->spos] = (double)it;
      sly[rs->spos] = core;
      rs->spos = (rs->spos + 1) % rp.slope_win;
      if (rs->sn < rp.slope_win) rs->sn++;
    }
    if (rp.use_dp && it >= rp.ramp_steps) {
      const int win = rp.win;
      hist[rs->idx] = core;
      rs->idx = (rs->idx + 1) % win;
      if (rs->idx == 0) rs->filled = 1;
      if (rs->filled) {
        // SIMPLE.cpp:456-490
        double sum = 0.0;
        for (int k = 0; k < win; ++k) sum += hist[k];
        const double mean = sum / win;
        double var = 0.0;
        for (int k = 0; k < win; ++k) var += (hist[k] - mean) * (hist[k] - mean);
        const double den = fmax(fabs(mean), 1.0);
        const double sdPct = sqrt(var / win) / den * 100.0;
        const double lastPct = fabs(hist[(rs->idx - 1 + win) % win] - mean) / den * 100.0;
        bool slopeOk = true;
        if (rp.use_slope && rp.slope_win > 0 && rs->sn >= rp.slope_win) {
          // least-squares slope over the ring in time order (SIMPLE.cpp:347-366)
          const int n0 = rs->sn, first = (rs->spos - n0 + rp.slope_win) % rp.slope_win;
          const double x0 = slx[first], n = (double)n0;
          double sx = 0, sy = 0, sxx = 0, sxy = 0;
          for (int k = 0; k < n0; ++k) {
            const int q = (first + k) % rp.slope_win;
            const double x = slx[q] - x0;
            sx += x;
            sy += sly[q];
            sxx += x * x;
            sxy += x * sly[q];
          }
          const double d = n * sxx - sx * sx;
          const double sl = fabs(d) < (double)1e-12f ? 0.0 : (n * sxy - sx * sy) / d;
          slopeOk = fabs(sl) <= rp.slope_max_delta / rp.slope_win;
        }
        conv = sdPct < rp.dp_percent && lastPct < rp.dp_percent && slopeOk;
      }
    }
    stop = it + 1 >= rp.max_iterations || !(sst.error > rp.epsilon) || conv;
  }
  if (stop) {
    sst.dp_converged = conv ? 1 : 0;
    rs->conv = conv ? 1 : 0;
    rs->last = sst;
    rs->stop_it = it + 1;
    rs->stopped = 1;
  }
}

// ===========================================================================
// host side
// ===========================================================================
class SimpleSolver {
 public:
  Ctx* ctx;
  cps_simple_config cfg;
  Geo g;
  int mode_fp64 = 1;
  size_t S = 0;  // elements per padded field
  DevBuf<double> u, v, p, us, vs, ps, dE, dN, gam, alp, inlet, inlet0;
  DevBuf<double> momD[16];  // FP64 coefficient arrays (x-form)
  DevBuf<float> momF[16];   // FP32 coefficient arrays (delta form)
  DevBuf<float> momX2[2];   // third x buffer of the paired Jacobi sweeps (delta form)
  DevBuf<double> pcg[6];
  DevBuf<double> part;  // SLOT_COUNT x NB block partials
  DevBuf<IterState> st;
  std::unique_ptr<PressureMG> mg;
  FlowAdjoint* adjoint = nullptr;
  KernelTimer timer;
  int iteration = 0;
  dim3 grid, block;
  int* hflag = nullptr;

  // pressureSolverType 2 / 4 (AMGCL CG with AMG, CPU / CUDA) and 3 (sparse
  // LDLT) have no build of their own: like the reference when it is built
  // without those libraries (SIMPLE.cpp:147-186), they fall back to another
  // backend.  Here that backend is the device multigrid PCG with the stop
  // rule of the solver asked for: AMGCL's ||r|| <= pTol ||b|| within
  // maxPressureIter iterations (pressure_solver.cpp:754-755,
  // pressure_solver.cu:115-116), and for the direct solve a residual at
  // FP64 round-off level (1e-12 ||b||, at least 100 iterations allowed).
  static void resolve_pressure_solver(cps_simple_config& c) {
    if (c.pressure_solver == CPS_PRESSURE_AMGCL || c.pressure_solver == CPS_PRESSURE_AMGCL_CUDA) {
      c.p_rtol = c.p_tol;
      c.p_tol = 0.0;
      c.pressure_solver = CPS_PRESSURE_GMG;
    } else if (c.pressure_solver == CPS_PRESSURE_LDLT) {
      c.p_rtol = 1e-12;
      c.p_tol = 0.0;
      c.max_pressure_iter = std::max(c.max_pressure_iter, 100);
      c.max_refine = std::max(c.max_refine, 8);
      c.pressure_solver = CPS_PRESSURE_GMG;
    }
  }

  SimpleSolver(Ctx* c, const cps_simple_config& conf, const double* gamma, const double* inlet_u)
      : ctx(c), cfg(conf) {
    CPS_CHECK(cfg.nx >= 2 && cfg.ny >= 2, CPS_ERR_ARG, "nx, ny must be >= 2");
    CPS_CHECK(cfg.hx > 0 && cfg.hy > 0, CPS_ERR_ARG, "hx, hy must be > 0");
    CPS_CHECK(cfg.uv_alpha > 0 && cfg.p_alpha > 0, CPS_ERR_ARG, "relaxation factors must be > 0");
    CPS_CHECK(cfg.precision == CPS_PREC_FP64 || cfg.precision == CPS_PREC_MIXED, CPS_ERR_ARG,
              "precision must be 64 or 32");
    CPS_CHECK(cfg.pressure_solver >= CPS_PRESSURE_SOR && cfg.pressure_solver <= CPS_PRESSURE_GMG, CPS_ERR_ARG,
              "pressure_solver must be 0 (SOR), 1 (PCG), 2 (AMGCL), 3 (LDLT), 4 (AMGCL-CUDA) or 5 (GMG)");
    resolve_pressure_solver(cfg);
    CPS_CHECK(!(cfg.precision == CPS_PREC_MIXED && cfg.pressure_solver != CPS_PRESSURE_GMG), CPS_ERR_ARG,
              "mixed precision needs the GMG pressure solver (5)");
    CPS_CHECK(gamma != nullptr, CPS_ERR_ARG, "gamma is required");
    mode_fp64 = cfg.precision == CPS_PREC_FP64;
    const int M = cfg.ny + 1, N = cfg.nx + 1;
    g.M = M;
    g.N = N;
    g.P = ((N + 1 + 63) / 64) * 64;
    g.nU = (M - 1) * (N - 2);
    g.nV = (M - 2) * (N - 1);
    g.nP = (M - 1) * (N - 2);
    g.hx = cfg.hx;
    g.hy = cfg.hy;
    g.rho = cfg.rho;
    g.eta = cfg.eta;
    g.vol = g.hx * g.hy;
    g.targetVel = cfg.u_in;
    g.maxVel = 3.0 * std::max(std::fabs(cfg.u_in), kTenthF);
    const bool on25 = cfg.enable_25d != 0 && cfg.ht > 0.0;  // SIMPLE.cpp:67-73
    g.convScale = on25 ? cfg.conv_factor_25d : 1.0;
    const double sinkCoeff = on25 ? cfg.sink_mult_25d * (12.0 * cfg.eta / (cfg.ht * cfg.ht)) : 0.0;
    g.sinkDiag = sinkCoeff * g.vol;
    g.De = cfg.eta * (g.hy / g.hx);
    g.Dn = cfg.eta * (g.hx / g.hy);
    g.uvAlpha = cfg.uv_alpha;
    g.pAlpha = cfg.p_alpha;
    g.sou = cfg.convection_scheme == 1;
    g.zs = ctx->zs;
    S = (size_t)(M + 1) * g.P;

    for (auto* b : {&u, &v, &p, &us, &vs, &ps, &dE, &dN, &gam, &alp}) b->alloc(S);
    inlet.alloc(M + 1);
    inlet0.alloc(M + 1);
    timer.stream = ctx->stream;
    for (auto* b : {&u, &v, &p, &us, &vs, &ps, &dE, &dN, &gam, &alp})
      CPS_HIP(hipMemsetAsync(b->p, 0, b->bytes(), ctx->stream));
    if (mode_fp64) {
      for (auto& b : momD) { b.alloc(S); CPS_HIP(hipMemsetAsync(b.p, 0, b.bytes(), ctx->stream)); }
    } else {
      for (auto& b : momF) { b.alloc(S); CPS_HIP(hipMemsetAsync(b.p, 0, b.bytes(), ctx->stream)); }
      for (auto& b : momX2) { b.alloc(S); CPS_HIP(hipMemsetAsync(b.p, 0, b.bytes(), ctx->stream)); }
    }
    for (auto& b : pcg) { b.alloc(S); CPS_HIP(hipMemsetAsync(b.p, 0, b.bytes(), ctx->stream)); }
    // 64 x 16 tiles; rows are strided only on grids large enough to need
    // more than 2048 blocks (bounds the partial sums every block re-reads)
    block = dim3(BX, BY);
    g.gx = ceil_div(N + 1, BX);
    g.gy = std::max(1, std::min(ceil_div(M + 1, BY), std::max(1, 2048 / g.gx)));
    g.NB = g.gx * g.gy;
    grid = dim3(g.gx, g.gy);
    part.alloc((size_t)SLOT_COUNT * g.NB);
    CPS_HIP(hipMemsetAsync(part.p, 0, part.bytes(), ctx->stream));
    st.alloc(1);
    CPS_HIP(hipMemsetAsync(st.p, 0, sizeof(IterState), ctx->stream));

    // gamma (clamped, padded: SIMPLE.cpp:707-729) and alpha (SIMPLE.cpp:766-771)
    std::vector<double> hg((size_t)M * g.P, 1.0), ha((size_t)M * g.P, 0.0);
    const int pr = M - 1, pc = N - 1;
    for (int i = 0; i < pr; ++i)
      for (int j = 0; j < pc; ++j) hg[(size_t)i * g.P + j] = std::max(0.0, std::min(1.0, gamma[(size_t)i * pc + j]));
    for (int i = 0; i < pr; ++i) hg[(size_t)i * g.P + pc] = hg[(size_t)i * g.P + pc - 1];
    for (int j = 0; j < N; ++j) hg[(size_t)pr * g.P + j] = hg[(size_t)(pr - 1) * g.P + std::min(j, pc - 1)];
    const double alphaMax = cfg.eta / std::max(cfg.brinkman_kmin, kKminFloor);
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < N; ++j) {
        const double gg = hg[(size_t)i * g.P + j];
        ha[(size_t)i * g.P + j] = alphaMax * (1.0 - gg) / (1.0 + cfg.brinkman_q * gg + kAlphaEps);
      }
    CPS_HIP(hipMemcpyAsync(gam.p, hg.data(), hg.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    CPS_HIP(hipMemcpyAsync(alp.p, ha.data(), ha.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    // inlet values with the gamma > 0.5 mask (boundaries.cpp:82-100)
    hinlet.assign(M + 1, 0.0);
    for (int i = 0; i <= M; ++i) {
      int cr = i <= 0 ? 0 : (i >= M ? std::max(0, M - 2) : i - 1);
      cr = std::max(0, std::min(cr, M - 1));
      const double speed = inlet_u ? inlet_u[i] : cfg.u_in;
      hinlet[i] = inlet_fluid(hg[(size_t)cr * g.P]) ? speed : 0.0;
    }
    CPS_HIP(hipMemcpyAsync(inlet.p, hinlet.data(), (M + 1) * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    CPS_HIP(hipMemcpyAsync(inlet0.p, hinlet.data(), (M + 1) * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    CPS_CHECK(!cfg.use_dp_convergence || cfg.dp_window >= 1, CPS_ERR_ARG, "dp_window must be >= 1");
    CPS_CHECK(!cfg.use_dp_slope_gate || cfg.dp_slope_window >= 0, CPS_ERR_ARG, "dp_slope_window must be >= 0");
    monitor.c[0] = plane_cols(cfg.x_core_in);
    monitor.c[1] = plane_cols(cfg.x_core_out);
    monitor.c[2] = patch_cols(1);
    monitor.c[3] = patch_cols(0);
    if (cfg.pressure_solver == CPS_PRESSURE_GMG) mg_create();
    CPS_HIP(hipStreamSynchronize(ctx->stream));
  }

  std::vector<double> hinlet;
  SampleSet monitor{};       // core in / out planes, inlet / outlet patches
  double inlet_scale = 1.0;  // current inlet ramp factor

  ~SimpleSolver() { flow_adjoint_free(adjoint); }

  // sampleBoundaryPatch columns (postprocessing.cpp:258-262)
  SampleCols patch_cols(int inl) const {
    const int N = g.N;
    SampleCols c{};
    c.gl = c.gr = inl ? 0 : N - 2;
    c.pl = c.pr = inl ? 0 : N;
    c.ul = c.ur = inl ? 0 : N - 1;
    c.vl = c.vr = inl ? 0 : N;
    c.t = 0.0;
    return c;
  }
  // samplePlaneAtX columns: cell columns cl, cl + 1 whose centres bracket x,
  // clamped to the first / last centre (postprocessing.cpp:62-96); p sits
  // one column right of its cell, u and v at the cell's own column index.
  SampleCols plane_cols(double x) const {
    const int NCells = g.N - 1;
    int cl = 0;
    double t = 0.0;
    if (NCells > 1) {
      const double xc = std::max(0.5 * g.hx, std::min(x, (NCells - 0.5) * g.hx));
      cl = (int)std::floor(xc / g.hx - 0.5);
      cl = std::max(0, std::min(cl, NCells - 2));
      const double xl = (cl + 0.5) * g.hx, xr = (cl + 1 + 0.5) * g.hx;
      t = std::fabs(xr - xl) < (double)1e-12f ? 0.0 : (xc - xl) / (xr - xl);
      t = std::max(0.0, std::min(1.0, t));
    }
    const int cr = NCells > 1 ? cl + 1 : cl;
    SampleCols c{};
    c.gl = cl;
    c.gr = cr;
    c.pl = cl + 1;
    c.pr = cr + 1;
    c.ul = cl;
    c.ur = cr;
    c.vl = cl;
    c.vr = cr;
    c.t = t;
    return c;
  }
  // the four samples into st->mon
  void sample() {
    k_sample<<<4, 256, 0, ctx->stream>>>(g, fields(), cfg.dp_fluid_tol, monitor, &st.p->mon[0][0]);
    CPS_HIP(hipGetLastError());
  }
  // ramped inlet (SIMPLE.cpp:381-386) and the velocity clamp that follows it
  // (iterations.cpp:104, momentum_solver.cpp:186)
  void set_inlet_scale(double sc) {
    if (sc == inlet_scale) return;
    inlet_scale = sc;
    g.maxVel = 3.0 * std::max(std::fabs(sc * cfg.u_in), kTenthF);
    const int n = g.M + 1;
    k_inlet_scale<<<ceil_div(n, 256), 256, 0, ctx->stream>>>(n, sc, inlet0.p, inlet.p);
    CPS_HIP(hipGetLastError());
    apply_bc();
  }

  SimpleView view() const {
    return SimpleView{g, S, cfg, u.p, v.p, p.p, gam.p, alp.p, inlet.p, dE.p, dN.p};
  }
  FlowAdjoint* adj() {
    if (!adjoint) adjoint = flow_adjoint_new(ctx, g, S);
    return adjoint;
  }

  Fields fields() {
    return Fields{u.p, v.p, p.p, us.p, vs.p, ps.p, dE.p, dN.p, gam.p, alp.p, inlet.p};
  }
  // 64 x 16 tiles over the active cells i = 1 .. M-1, j = 1 .. N-1 (one
  // thread each): 512 blocks for 1024 x 512, one resident wave of two
  // 1024-thread blocks per CU, where the field grid's 561 leave a tail
  dim3 active_grid() const { return dim3(ceil_div(g.N - 1, BX), ceil_div(g.M - 1, BY)); }
  template <class C>
  MomSys<C> sys(DevBuf<C>* b, int comp) {
    DevBuf<C>* q = b + 8 * comp;
    C* x2 = nullptr;
    if constexpr (std::is_same<C, float>::value) x2 = momX2[comp].p;
    return MomSys<C>{q[0].p, q[1].p, q[2].p, q[3].p, q[4].p, q[5].p, q[6].p, q[7].p, x2};
  }
  PcgBufs pbufs() { return PcgBufs{pcg[0].p, pcg[1].p, pcg[2].p, pcg[3].p, pcg[4].p, pcg[5].p}; }

  void apply_bc() {
    dim3 bg(ceil_div(std::max(g.M, g.N) + 1, 256), 2);
    k_bc<<<bg, 256, 0, ctx->stream>>>(g, u.p, v.p, p.p, inlet.p, 1);
    CPS_HIP(hipGetLastError());
  }

  // Reads one int of device state (blocking).
  int read_int(const int* dptr) {
    int h = 0;
    CPS_HIP(hipMemcpyAsync(&h, dptr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    CPS_HIP(hipStreamSynchronize(ctx->stream));
    return h;
  }

  // Launches up to `maxit` device-side-terminating iterations; when the cap is
  // large the host checks the device stop flag between chunks.
  static constexpr int kNoSyncCap = 32;

  void momentum() {
    hipStream_t s = ctx->stream;
    const Fields f = fields();
    const double tol2 = cfg.momentum_tol * cfg.momentum_tol;
    const int maxit = cfg.max_momentum_iter;
    const int chunk = maxit > kNoSyncCap ? 8 : std::max(1, maxit);
    // the multigrid pressure solve's hierarchy is built as soon as dE, dN
    // are final: its grid levels by the assembly's own blocks (else one
    // launch after it), the coarse set-up by k_finalize
    const bool mgpre = mg && cfg.pressure_solver == CPS_PRESSURE_GMG;
    if (mode_fp64) {
      mgc::HierJob<double> hj{};
      if (mgpre && !mg->prebuild(&hj)) mg->prebuild_launch(dE.p, dN.p, true);
      k_assemble<double, false><<<ctx->zg(active_grid()), block, 0, s>>>(g, f, sys(momD, 0), sys(momD, 1), st.p, hj);
    } else {
      mgc::HierJob<float> hj{};
      if (mgpre && !mg->prebuild(&hj)) mg->prebuild_launch(dE.p, dN.p, false);
      k_assemble<float, true><<<ctx->zg(active_grid()), block, 0, s>>>(g, f, sys(momF, 0), sys(momF, 1), st.p, hj);
    }
    CPS_HIP(hipGetLastError());
    int k = 0;
    // two sweeps per launch on the unpolled delta path (the timed single
    // sweeps keep their own kernel); the pair grid tiles the active cells
    // from (1, 1) and its partials fit the slots of the field grid
    const dim3 jgrid = active_grid();
    const bool pairs = !mode_fp64 && maxit > 0 && chunk == maxit && timer.which != KT_MOM_JACOBI &&
                       jgrid.x > 0 && jgrid.y > 0 && (int)(jgrid.x * jgrid.y) <= g.NB;
    if (pairs) {
      for (; k < maxit; k += 2)
        k_jacobi2<float, kJacRows><<<ctx->zg(jgrid), dim3(BX, BY / kJacRows), 0, s>>>(
            g, sys(momF, 0), sys(momF, 1), k, k + 1 < maxit ? 1 : 0, tol2, st.p, part.p);
      CPS_HIP(hipGetLastError());
    }
    while (!pairs && k < maxit) {
      const int k1 = std::min(maxit, k + chunk);
      for (; k < k1; ++k) {
        const bool timed = timer.active(KT_MOM_JACOBI);
        if (mode_fp64 && timed)
          timer.ext_launch(k_jacobi<double>, ctx->zg(grid), block, 0, s, g, sys(momD, 0), sys(momD, 1), k, tol2, st.p, part.p);
        else if (mode_fp64)
          k_jacobi<double><<<ctx->zg(grid), block, 0, s>>>(g, sys(momD, 0), sys(momD, 1), k, tol2, st.p, part.p);
        else if (timed)
          timer.ext_launch(k_jacobi<float>, ctx->zg(grid), block, 0, s, g, sys(momF, 0), sys(momF, 1), k, tol2, st.p, part.p);
        else
          k_jacobi<float><<<ctx->zg(grid), block, 0, s>>>(g, sys(momF, 0), sys(momF, 1), k, tol2, st.p, part.p);
      }
      // x, cS, cW, aP, cE, cN, b read + x' written, both components
      if (timer.which == KT_MOM_JACOBI)
        timer.bytes_per_launch = 8.0 * (mode_fp64 ? 8 : 4) * (double)(g.nU + g.nV);
      if (k >= maxit && chunk == maxit) break;  // the counts are settled by k_finalize
      k_jac_end<<<ctx->zg(1), block, 0, s>>>(g, k, tol2, st.p, part.p);
      CPS_HIP(hipGetLastError());
      if (k >= maxit) break;
      int act[2];
      CPS_HIP(hipMemcpyAsync(act, st.p->jac_active, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
      CPS_HIP(hipStreamSynchronize(s));
      if (!act[0] && !act[1]) break;
    }
    const int K = maxit == 0 ? 0 : chunk == maxit ? maxit : -1;
    // the multigrid's coarse set-up rides on the finalize launch as one
    // extra column of blocks (the grid levels' faces were built after the
    // assembly)
    size_t smem = 0;
    if (mode_fp64) {
      mgc::CoarseJob<double> cj{};
      const bool job = mg && mg->coarse_job(&cj, &smem);
      const dim3 fg(grid.x + (job ? 1 : 0), grid.y);
      k_finalize<double, false><<<ctx->zg(fg), block, smem, s>>>(g, f, sys(momD, 0), sys(momD, 1), st.p, part.p, K, 0,
                                                                 tol2, cj);
    } else {
      mgc::CoarseJob<float> cj{};
      const bool job = mg && mg->coarse_job(&cj, &smem);
      const dim3 fg(grid.x + (job ? 1 : 0), grid.y);
      k_finalize<float, true><<<ctx->zg(fg), block, smem, s>>>(g, f, sys(momF, 0), sys(momF, 1), st.p, part.p, K,
                                                               pairs ? (int)(jgrid.x * jgrid.y) : 0, tol2, cj);
    }
    CPS_HIP(hipGetLastError());
  }

  void pressure_pcg() {
    hipStream_t s = ctx->stream;
    // z, p_old, dE, dN read; p, Ap written
    if (timer.which == KT_PCG_STENCIL) timer.bytes_per_launch = 6.0 * 8.0 * (double)g.nP;
    const int maxit = cfg.max_pressure_iter;
    const int chunk = maxit > kNoSyncCap ? 16 : std::max(1, maxit);
    const PcgBufs q = pbufs();
    int k = 0;
    while (k < maxit) {
      const int k1 = std::min(maxit, k + chunk);
      for (; k < k1; ++k) {
        if (timer.active(KT_PCG_STENCIL))
          timer.ext_launch(k_pcg_a, ctx->zg(grid), block, 0, s, g, (const double*)dE.p, (const double*)dN.p, q, k, cfg.p_tol,
                           st.p, part.p);
        else
          k_pcg_a<<<ctx->zg(grid), block, 0, s>>>(g, dE.p, dN.p, q, k, cfg.p_tol, st.p, part.p);
        k_pcg_b<<<ctx->zg(grid), block, 0, s>>>(g, ps.p, q, k, st.p, part.p);
      }
      k_pcg_end<<<ctx->zg(1), block, 0, s>>>(g, k, cfg.p_tol, st.p, part.p);
      CPS_HIP(hipGetLastError());
      if (k >= maxit) break;
      if (read_int(&st.p->pcg_stop)) break;
    }
  }

  // red-black SOR with the reference's automatic omega (pressure_solver.cpp:
  // 820-826): 2 / (1 + sqrt(1 - cos(pi / max(M, N))^2)), clamped to [1, 1.95]
  void pressure_sor() {
    hipStream_t s = ctx->stream;
    const double sr = std::cos(M_PI / (double)std::max(g.M, g.N));
    const double omega = std::max(1.0, std::min(1.95, 2.0 / (1.0 + std::sqrt(1.0 - sr * sr))));
    const int maxit = cfg.max_pressure_iter;
    const int chunk = maxit > kNoSyncCap ? 16 : std::max(1, maxit);
    const Fields f = fields();
    const PcgBufs q = pbufs();
    int k = 0;
    while (k < maxit) {
      const int k1 = std::min(maxit, k + chunk);
      for (; k < k1; ++k)
        for (int colour = 0; colour < 2; ++colour)
          k_sor<<<ctx->zg(grid), block, 0, s>>>(g, f, q.r, k, colour, omega, cfg.p_tol, st.p, part.p);
      k_sor_end<<<ctx->zg(1), block, 0, s>>>(g, k, cfg.p_tol, st.p, part.p);
      CPS_HIP(hipGetLastError());
      if (k >= maxit) break;
      if (read_int(&st.p->pcg_stop)) break;
    }
  }

  void mg_create();
  void pressure_mg();

  void step_once() {
    hipStream_t s = ctx->stream;
    const Fields f = fields();
    momentum();
    // the multigrid solve's start rides on the right side's kernel
    if (cfg.pressure_solver == CPS_PRESSURE_PCG) {
      k_prhs<true, double><<<ctx->zg(grid), block, 0, s>>>(g, f, pbufs(), part.p, mgc::StartJob<double>{});
    } else if (mode_fp64) {
      mgc::StartJob<double> sj{};
      if (mg) mg->start_job(cfg, &sj);
      k_prhs<false, double><<<ctx->zg(grid), block, 0, s>>>(g, f, pbufs(), part.p, sj);
    } else {
      mgc::StartJob<float> sj{};
      if (mg) mg->start_job(cfg, &sj);
      k_prhs<false, float><<<ctx->zg(grid), block, 0, s>>>(g, f, pbufs(), part.p, sj);
    }
    // the Jacobi-PCG and SOR solves start from the statistics' r.z and
    // ||b||^2; the multigrid solve reads none of them, so there the
    // reductions wait until the host asks (stats())
    if (cfg.pressure_solver != CPS_PRESSURE_GMG) k_stats<<<ctx->zg(1), block, 0, s>>>(g, st.p, part.p, 1);
    CPS_HIP(hipGetLastError());
    if (cfg.pressure_solver == CPS_PRESSURE_PCG) pressure_pcg();
    else if (cfg.pressure_solver == CPS_PRESSURE_SOR) pressure_sor();
    else pressure_mg();
    k_correct<<<ctx->zg(grid), block, 0, s>>>(g, f);
    CPS_HIP(hipGetLastError());
    iteration++;
  }

  void stats(cps_step_stats* out) {
    // multigrid steps leave their statistics' partials in place (SLOT_FIN,
    // SLOT_RHS are written once per step and read by nothing else)
    if (cfg.pressure_solver == CPS_PRESSURE_GMG) {
      k_stats<<<1, block, 0, ctx->stream>>>(g, st.p, part.p, 0);
      CPS_HIP(hipGetLastError());
      if (mg) mg->finish(st.p);
    }
    IterState h{};
    CPS_HIP(hipMemcpyAsync(&h, st.p, sizeof(IterState), hipMemcpyDeviceToHost, ctx->stream));
    CPS_HIP(hipStreamSynchronize(ctx->stream));
    if (!out) return;
    fill_stats(h, iteration, out);
  }
  // the statistics kernels of stats() without the read-back
  void stats_launch() {
    if (cfg.pressure_solver == CPS_PRESSURE_GMG) {
      k_stats<<<1, block, 0, ctx->stream>>>(g, st.p, part.p, 0);
      CPS_HIP(hipGetLastError());
      if (mg) mg->finish(st.p);
    }
  }

  void step(int n, cps_step_stats* out) {
    apply_bc();
    for (int k = 0; k < n; ++k) step_once();
    sample();
    cps_step_stats tmp;
    stats(out ? out : &tmp);
    const double e = (out ? out : &tmp)->error;
    CPS_CHECK(std::isfinite(e), CPS_ERR_DIVERGED, "SIMPLE residual is not finite");
  }

  // SIMPLE.cpp:310-491 without console output, checkpoints and the
  // pseudo-time controller, with the loop's services on the device: after
  // every iteration k_run_monitor writes the log row, updates the dP window
  // and the slope gate and evaluates the loop condition.  The host enqueues
  // `poll` iterations, then reads the monitor (one synchronisation per
  // chunk).  A chunk that ran past the stop iteration is undone: u, v, p,
  // the inlet ramp and the iteration count are restored from the chunk's
  // start and the iterations up to the stop replayed, which repeats them bit
  // for bit (every kernel is deterministic).  Returns the iterations run.
  DevBuf<double> snap[3];
  DevBuf<RunState> runst;
  DevBuf<double> hist, slx, sly;
  DevBuf<cps_iter_log> dlog;
  int host_syncs = 0;

  static double ramp_of(const cps_simple_config& c, int it) {
    return c.inlet_ramp && c.ramp_steps > 0 && it < c.ramp_steps ? double(it + 1) / double(c.ramp_steps) : 1.0;
  }

  int run(cps_step_stats* out, cps_iter_log* log = nullptr, int log_cap = 0, int poll = kDefaultPoll) {
    CPS_CHECK(poll >= 1, CPS_ERR_ARG, "poll interval must be >= 1");
    host_syncs = 0;
    apply_bc();
    cps_step_stats sst{};
    sst.error = 1.0;
    if (!(cfg.max_iterations > 0 && sst.error > cfg.epsilon)) {  // the loop body never runs
      if (out) *out = sst;
      return 0;
    }
    const size_t nf[3] = {u.n, v.n, p.n};
    for (int k = 0; k < 3; ++k)
      if (snap[k].n != nf[k]) snap[k].alloc(nf[k]);
    if (!runst.p) runst.alloc(1);
    const int win = std::max(1, cfg.dp_window);
    const int swin = cfg.use_dp_slope_gate ? std::max(0, cfg.dp_slope_window) : 0;
    if ((int)hist.n < win) hist.alloc(win);
    if ((int)slx.n < std::max(1, swin)) { slx.alloc(std::max(1, swin)); sly.alloc(std::max(1, swin)); }
    const int lcap = log ? std::min(log_cap, cfg.max_iterations) : 0;
    if ((int)dlog.n < lcap) dlog.alloc(lcap);
    RunParams rp{};
    rp.max_iterations = cfg.max_iterations;
    rp.ramp_steps = cfg.ramp_steps;
    rp.use_dp = cfg.use_dp_convergence;
    rp.win = win;
    rp.use_slope = cfg.use_dp_slope_gate;
    rp.slope_win = swin;
    rp.log_cap = lcap;
    rp.epsilon = cfg.epsilon;
    rp.dp_percent = cfg.dp_percent;
    rp.slope_max_delta = cfg.dp_slope_max_delta;
    hipStream_t s = ctx->stream;
    k_run_reset<<<1, 64, 0, s>>>(runst.p);
    CPS_HIP(hipGetLastError());
    DevBuf<double>* fld[3] = {&u, &v, &p};
    RunState h{};
    int k0 = 0;
    for (;;) {
      // chunk start: what the next iterations read from earlier ones
      for (int k = 0; k < 3; ++k)
        CPS_HIP(hipMemcpyAsync(snap[k].p, fld[k]->p, fld[k]->bytes(), hipMemcpyDeviceToDevice, s));
      const double scale0 = inlet_scale;
      const int iter0 = iteration;
      const int k1 = std::min(cfg.max_iterations, k0 + poll);
      for (int it = k0; it < k1; ++it) {
        set_inlet_scale(ramp_of(cfg, it));
        step_once();
        sample();
        stats_launch();
        k_run_monitor<<<1, 64, 0, s>>>(it, iteration, rp, st.p, runst.p, hist.p, slx.p, sly.p,
                                       lcap > 0 ? dlog.p : nullptr);
        CPS_HIP(hipGetLastError());
      }
      CPS_HIP(hipMemcpyAsync(&h, runst.p, sizeof(RunState), hipMemcpyDeviceToHost, s));
      CPS_HIP(hipStreamSynchronize(s));
      host_syncs++;
      if (h.stopped) {
        if (h.stop_it < k1) {  // ran past the stop: restore the chunk start, replay to the stop
          for (int k = 0; k < 3; ++k)
            CPS_HIP(hipMemcpyAsync(fld[k]->p, snap[k].p, fld[k]->bytes(), hipMemcpyDeviceToDevice, s));
          restore_inlet_scale(scale0);
          iteration = iter0;
          for (int it = k0; it < h.stop_it; ++it) {  // the same launches as the first pass, minus the monitor
            set_inlet_scale(ramp_of(cfg, it));
            step_once();
            sample();
            stats_launch();
          }
        }
        break;
      }
      CPS_CHECK(k1 < cfg.max_iterations, CPS_ERR_STATE, "outer loop monitor missed the iteration cap");
      k0 = k1;
    }
    CPS_CHECK(!h.nonfinite, CPS_ERR_DIVERGED, "SIMPLE residual is not finite");
    if (lcap > 0) {
      const int rows = std::min(lcap, h.stop_it);
      CPS_HIP(hipMemcpyAsync(log, dlog.p, (size_t)rows * sizeof(cps_iter_log), hipMemcpyDeviceToHost, s));
      CPS_HIP(hipStreamSynchronize(s));
      host_syncs++;
    }
    sst = h.last;
    sst.iteration = iteration;
    if (out) *out = sst;
    return h.stop_it;
  }
  static constexpr int kDefaultPoll = 32;

  // sets the inlet ramp factor to sc without touching u (the fields already
  // hold the boundary values of that moment)
  void restore_inlet_scale(double sc) {
    inlet_scale = sc;
    g.maxVel = 3.0 * std::max(std::fabs(sc * cfg.u_in), kTenthF);
    const int n = g.M + 1;
    k_inlet_scale<<<ceil_div(n, 256), 256, 0, ctx->stream>>>(n, sc, inlet0.p, inlet.p);
    CPS_HIP(hipGetLastError());
  }

  DevBuf<double>* field(int which) {
    switch (which) {
      case CPS_FIELD_U: return &u;
      case CPS_FIELD_V: return &v;
      case CPS_FIELD_P: return &p;
      case CPS_FIELD_USTAR: return &us;
      case CPS_FIELD_VSTAR: return &vs;
      case CPS_FIELD_PSTAR: return &ps;
      case CPS_FIELD_DE: return &dE;
      case CPS_FIELD_DN: return &dN;
      case CPS_FIELD_GAMMA: return &gam;
      case CPS_FIELD_ALPHA: return &alp;
    }
    return nullptr;
  }
  void shape(int which, int* r, int* c) {
    const int M = g.M, N = g.N;
    switch (which) {
      case CPS_FIELD_U: case CPS_FIELD_USTAR: case CPS_FIELD_DE: *r = M + 1; *c = N; break;
      case CPS_FIELD_V: case CPS_FIELD_VSTAR: case CPS_FIELD_DN: *r = M; *c = N + 1; break;
      case CPS_FIELD_P: case CPS_FIELD_PSTAR: *r = M + 1; *c = N + 1; break;
      default: *r = M; *c = N;
    }
  }
  void get(int which, double* out) {
    DevBuf<double>* b = field(which);
    CPS_CHECK(b && out, CPS_ERR_ARG, "bad field id or null output");
    int r, c;
    shape(which, &r, &c);
    CPS_HIP(hipMemcpy2DAsync(out, c * sizeof(double), b->p, g.P * sizeof(double), c * sizeof(double), r,
                             hipMemcpyDeviceToHost, ctx->stream));
    CPS_HIP(hipStreamSynchronize(ctx->stream));
  }
  void set(int which, const double* in) {
    CPS_CHECK(which >= CPS_FIELD_U && which <= CPS_FIELD_PSTAR, CPS_ERR_ARG, "only u, v, p, uStar, vStar, pStar are settable");
    DevBuf<double>* b = field(which);
    CPS_CHECK(in, CPS_ERR_ARG, "null input");
    int r, c;
    shape(which, &r, &c);
    CPS_HIP(hipMemcpy2DAsync(b->p, g.P * sizeof(double), in, c * sizeof(double), c * sizeof(double), r,
                             hipMemcpyHostToDevice, ctx->stream));
    CPS_HIP(hipStreamSynchronize(ctx->stream));
  }
  void sample_one(const SampleCols& c, cps_patch* out) {
    DevBuf<double> o;
    o.alloc(5);
    SampleSet set{};
    set.c[0] = c;
    k_sample<<<1, 256, 0, ctx->stream>>>(g, fields(), cfg.dp_fluid_tol, set, o.p);
    CPS_HIP(hipGetLastError());
    double h[5];
    CPS_HIP(hipMemcpyAsync(h, o.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    CPS_HIP(hipStreamSynchronize(ctx->stream));
    finish_sample(h, out);
  }
  void patch(int inl, cps_patch* out) { sample_one(patch_cols(inl), out); }
  void plane(double x, cps_patch* out) { sample_one(plane_cols(x), out); }
};

// ---------------------------------------------------------------------------
// GMG pressure path: defined with the multigrid kernels.
void SimpleSolver::mg_create() { mg.reset(new PressureMG(ctx, g)); }
void SimpleSolver::pressure_mg() {
  const PcgBufs q = pbufs();
  mg->solve(dE.p, dN.p, q.r, ps.p, st.p, cfg, mode_fp64 != 0, &timer);
}

// ===========================================================================
// Batch of fields.  Fields whose grids and step settings agree form a size
// class.  A class keeps its fields' buffers in one allocation, field k's at
// k * stride: each solver of the class is built (and takes its first step,
// which creates the multigrid state) under an Arena at its own slot, so all
// of them lay their buffers out alike.  One SIMPLE iteration of the whole
// class is then its first field's step launched with gridDim.z = fields in
// the class: every kernel of the chain steps all of the class's fields at
// once, so the small multigrid levels and the one-workgroup kernels carry
// the whole class instead of one field.  A class's iteration (or `unroll`
// of them) is captured into one hipGraph, and the classes' graphs run on
// streams of their own (this runtime executes the parallel branches of a
// single graph one after another, so concurrency comes from streams).  The
// per-field plan, each field's own chain with the fields spread over up to
// 32 streams, stays as the tuner's alternative.  Either way a field's
// results are those of stepping it alone, bit for bit: the same kernels do
// the same arithmetic on the same data.
class BatchSolver {
 public:
  enum Plan { kClass = 0, kPerField = 1 };
  struct Group {
    std::vector<int> members;  // field indices; members[0] launches the class
    size_t stride = 0;         // bytes between consecutive members' buffers
  };
  struct Graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    hipStream_t s = nullptr;
  };
  struct Streams {
    std::vector<hipStream_t> v;
    ~Streams() {
      for (auto s : v) (void)hipStreamDestroy(s);
    }
  };
  static constexpr int kMaxStreams = 32;
  // arena bytes per field: 56 padded FP64 fields bound what a solver and its
  // multigrid allocate (at most ~47 in the FP64 form), plus the small
  // buffers (partials, coarse levels, scalars)
  static constexpr size_t kArenaFields = 56, kArenaExtra = size_t(8) << 20, kArenaAlign = size_t(2) << 20;

  Ctx* base;
  Streams streams;  // declared first: destroyed after the solvers
  std::vector<std::unique_ptr<DevBuf<char>>> arenas;
  std::vector<std::unique_ptr<Ctx>> fctx;
  std::vector<std::unique_ptr<SimpleSolver>> f;
  std::vector<Group> groups;
  std::vector<int> leader;       // field -> the field whose captured step covers it
  std::vector<Graph> one, many;  // graphs of one iteration / of `unroll` iterations
  int plan = kClass, plan_ns = 1, nstreams = 0, unroll = 1;  // plan_ns: chains per class / streams
  double us_per_step = 0.0;

  // the settings one step reads (the outer loop's and the samples' are per field)
  static bool same_step(const cps_simple_config& a, const cps_simple_config& b) {
    return a.nx == b.nx && a.ny == b.ny && a.hx == b.hx && a.hy == b.hy && a.rho == b.rho && a.eta == b.eta &&
           a.u_in == b.u_in && a.convection_scheme == b.convection_scheme && a.enable_25d == b.enable_25d &&
           a.ht == b.ht && a.conv_factor_25d == b.conv_factor_25d && a.sink_mult_25d == b.sink_mult_25d &&
           a.brinkman_kmin == b.brinkman_kmin && a.brinkman_q == b.brinkman_q && a.uv_alpha == b.uv_alpha &&
           a.p_alpha == b.p_alpha && a.momentum_tol == b.momentum_tol && a.max_momentum_iter == b.max_momentum_iter &&
           a.p_tol == b.p_tol && a.max_pressure_iter == b.max_pressure_iter && a.pressure_solver == b.pressure_solver &&
           a.precision == b.precision && a.p_rtol == b.p_rtol && a.max_refine == b.max_refine;
  }

  BatchSolver(Ctx* c, int n, const cps_simple_config* cfgs, const double* const* gammas,
              const double* const* inlets)
      : base(c) {
    CPS_CHECK(n >= 1 && cfgs && gammas, CPS_ERR_ARG, "batch needs n >= 1 fields with configs and gammas");
    std::vector<cps_simple_config> rc(cfgs, cfgs + n);
    for (auto& cf : rc) {
      SimpleSolver::resolve_pressure_solver(cf);
      // capture-safe caps: no host synchronisation inside one iteration
      CPS_CHECK(cf.max_momentum_iter <= SimpleSolver::kNoSyncCap && cf.max_pressure_iter <= SimpleSolver::kNoSyncCap &&
                    (cf.precision != CPS_PREC_MIXED || cf.max_refine <= 4),
                CPS_ERR_ARG, "batched fields need max_momentum_iter, max_pressure_iter <= 32 and max_refine <= 4");
    }
    for (int k = 0; k < kMaxStreams; ++k) {
      hipStream_t st;
      CPS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));  // default priority
      streams.v.push_back(st);
    }
    for (int i = 0; i < n; ++i) {
      size_t gi = 0;
      while (gi < groups.size() && !same_step(rc[groups[gi].members[0]], rc[i])) ++gi;
      if (gi == groups.size()) groups.emplace_back();
      groups[gi].members.push_back(i);
    }
    fctx.resize(n);
    f.resize(n);
    leader.assign(n, 0);
    for (size_t gi = 0; gi < groups.size(); ++gi) build_group(gi, cfgs, gammas, inlets);
    sync_all();
  }

  // The class's solvers, each under its own arena slot, each taking one
  // eager iteration (which creates the lazily built multigrid state:
  // allocations and clears stay out of the captured graphs).
  void build_group(size_t gi, const cps_simple_config* cfgs, const double* const* gammas,
                   const double* const* inlets) {
    Group& G = groups[gi];
    const cps_simple_config& c0 = cfgs[G.members[0]];
    const size_t P = ((size_t)c0.nx + 2 + 63) / 64 * 64, S = ((size_t)c0.ny + 2) * P;
    G.stride = (kArenaFields * S * sizeof(double) + kArenaExtra + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
    arenas.emplace_back(new DevBuf<char>);
    arenas.back()->alloc(G.stride * G.members.size());
    char* base_p = arenas.back()->p;
    size_t used = 0;
    for (size_t k = 0; k < G.members.size(); ++k) {
      const int i = G.members[k];
      Arena a{base_p + k * G.stride, G.stride, 0};
      ArenaScope scope(&a);
      fctx[i].reset(new Ctx(*base));
      fctx[i]->stream = streams.v[i % kMaxStreams];
      fctx[i]->zs = G.stride;
      fctx[i]->nz = 1;
      f[i].reset(new SimpleSolver(fctx[i].get(), cfgs[i], gammas[i], inlets ? inlets[i] : nullptr));
      f[i]->apply_bc();
      f[i]->step_once();
      if (k == 0) used = a.off;
      CPS_CHECK(a.off == used, CPS_ERR_STATE, "fields of one size class laid out their buffers differently");
      leader[i] = G.members[0];
    }
  }

  ~BatchSolver() {
    (void)sync_all_nothrow();
    release_graphs();
    f.clear();
    arenas.clear();
  }

  bool sync_all_nothrow() {
    bool ok = true;
    for (auto st : streams.v) ok = hipStreamSynchronize(st) == hipSuccess && ok;
    return ok;
  }
  void sync_all() {
    for (auto st : streams.v) CPS_HIP(hipStreamSynchronize(st));
  }

  void release_graphs() {
    for (auto* v : {&one, &many})
      for (auto& G : *v) {
        if (G.e) (void)hipGraphExecDestroy(G.e);
        if (G.g) (void)hipGraphDestroy(G.g);
      }
    one.clear();
    many.clear();
  }

  // `steps` iterations of the listed fields on stream s as one graph
  Graph capture_steps(const std::vector<int>& fields, hipStream_t s, int steps) {
    Graph G;
    G.s = s;
    CPS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      for (int r = 0; r < steps; ++r)
        for (int i : fields) {
          f[i]->step_once();
          f[i]->iteration--;  // capture does not run an iteration
        }
    } catch (...) {
      hipGraph_t dead = nullptr;
      (void)hipStreamEndCapture(s, &dead);
      if (dead) (void)hipGraphDestroy(dead);
      throw;
    }
    CPS_HIP(hipStreamEndCapture(s, &G.g));
    CPS_HIP(hipGraphInstantiate(&G.e, G.g, nullptr, nullptr, 0));
    return G;
  }

  // greedy longest-processing-time assignment (cost ~ cells)
  std::vector<int> assign(int ns) {
    std::vector<int> order(f.size()), owner(f.size(), 0);
    for (size_t i = 0; i < f.size(); ++i) order[i] = (int)i;
    auto cells = [&](int i) { return (double)f[i]->cfg.nx * f[i]->cfg.ny; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells(a) > cells(b); });
    std::vector<double> load(ns, 0.0);
    for (int i : order) {
      const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      owner[i] = k;
      load[k] += cells(i);
    }
    return owner;
  }

  // Captures plan p: kClass, each size class split into ns (>= 1) chains of
  // consecutive fields, one graph per chain (of `un` iterations, plus one of
  // a single iteration for remainders), chain k on stream k mod 32;
  // kPerField, the fields spread over ns streams, one graph per stream.
  void capture(int p, int ns, int un) {
    CPS_CHECK(p == kClass || p == kPerField, CPS_ERR_ARG, "plan must be 0 (class) or 1 (per field)");
    CPS_CHECK(un >= 1 && un <= 64, CPS_ERR_ARG, "unroll must be in 1..64");
    sync_all();
    release_graphs();
    const int n = (int)f.size();
    if (p == kClass) {
      // each class in `ns` chains of consecutive members (slots at the
      // class stride, so a chain's first member launches for the chain)
      const int split = std::max(1, ns);
      plan_ns = split;
      int chain = 0;
      for (size_t gi = 0; gi < groups.size(); ++gi) {
        const Group& G = groups[gi];
        const int m = (int)G.members.size(), nc = std::min(split, m);
        for (int c = 0; c < nc; ++c, ++chain) {
          const int k0 = c * m / nc, k1 = (c + 1) * m / nc;
          hipStream_t s = streams.v[chain % kMaxStreams];
          const int L = G.members[k0];
          fctx[L]->stream = s;  // the chain's data are written by its first member's launches
          for (int k = k0; k < k1; ++k) leader[G.members[k]] = L;
          fctx[L]->nz = k1 - k0;
          try {
            one.push_back(capture_steps({L}, s, 1));
            if (un > 1) many.push_back(capture_steps({L}, s, un));
          } catch (...) {
            fctx[L]->nz = 1;
            throw;
          }
          fctx[L]->nz = 1;
        }
      }
      nstreams = std::min(chain, kMaxStreams);
    } else {
      ns = std::max(1, std::min(ns, std::min(n, kMaxStreams)));
      plan_ns = ns;
      const std::vector<int> owner = assign(ns);
      for (int i = 0; i < n; ++i) {
        fctx[i]->stream = streams.v[owner[i]];
        leader[i] = i;
      }
      for (int k = 0; k < ns; ++k) {
        std::vector<int> mine;
        for (int i = 0; i < n; ++i)
          if (owner[i] == k) mine.push_back(i);
        one.push_back(capture_steps(mine, streams.v[k], 1));
        if (un > 1) many.push_back(capture_steps(mine, streams.v[k], un));
      }
      nstreams = ns;
    }
    plan = p;
    unroll = un;
  }

  void launch(int n) {
    CPS_CHECK(!one.empty(), CPS_ERR_STATE, "batch graphs not built: call cps_batch_tune or cps_batch_plan first");
    int it = 0;
    if (!many.empty())
      for (; it + unroll <= n; it += unroll)
        for (auto& G : many) CPS_HIP(hipGraphLaunch(G.e, G.s));
    for (; it < n; ++it)
      for (auto& G : one) CPS_HIP(hipGraphLaunch(G.e, G.s));
    for (auto& s : f) s->iteration += n;
    if (n > 0)
      for (size_t i = 0; i < f.size(); ++i)
        if (f[i]->mg) f[i]->mg->rearm(*f[leader[i]]->mg);
  }

  // Times `reps` batch iterations per candidate plan and keeps the fastest.
  // fixed > 0: the per-field plan on that many streams only (unroll 1);
  // fixed == 0: the class plan with 1 .. 4 chains per class and unroll 1,
  // 4, 8, and the per-field plan on 1 .. 32 streams.  Every candidate gives
  // the same results.
  static constexpr int kTuneWarmup = 128;
  void tune(int fixed, int reps) {
    struct Cand {
      int p, ns, un;
    };
    std::vector<Cand> cand;
    if (fixed > 0) cand.push_back({kPerField, fixed, 1});
    else {
      for (int split : {1, 2, 3, 4})
        for (int un : {1, 4, 8}) cand.push_back({kClass, split, un});
      for (int ns : {1, 2, 4, 8, 16, 24, 32})
        if (ns <= (int)f.size()) cand.push_back({kPerField, ns, 1});
    }
    // The fields keep iterating while the candidates are timed, and their
    // steps get cheaper as they converge (fewer live multigrid cycles), so
    // a single pass favours the candidates timed last.  Two passes, the
    // second in reverse order, cancel that drift to first order: each
    // candidate's score is the sum of its two timings.
    std::vector<double> us(cand.size(), 0.0);
    const size_t nc = cand.size();
    // The drift is fastest in a field's first iterations from rest (the
    // multigrid needs 7-8 cycles per step there, 2-3 later), where the
    // two passes no longer cancel it: the candidates at both ends of the
    // order were penalised.  Fields that start there first take
    // kTuneWarmup iterations under the first candidate.
    if (nc > 1) {
      int young = 0;
      for (auto& s : f) young = std::max(young, kTuneWarmup - s->iteration);
      if (young > 0) {
        capture(cand[0].p, cand[0].ns, cand[0].un);
        launch(young);
        sync_all();
      }
    }
    for (size_t pass = 0; pass < (nc > 1 ? 2u : 1u); ++pass)
      for (size_t kk = 0; kk < nc; ++kk) {
        const size_t k = pass == 0 ? kk : nc - 1 - kk;
        const Cand& c = cand[k];
        capture(c.p, c.ns, c.un);
        // reps graph replays, at least 16 iterations: the steps of a
        // converging field vary in cost (1-5 live multigrid cycles), so
        // shorter samples picked plans 10 % slower than the best
        const int r = std::max(reps * c.un, 16);
        launch(c.un);  // warm-up
        sync_all();
        const auto t0 = std::chrono::steady_clock::now();
        launch(r);
        sync_all();
        us[k] += std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() / r;
      }
    size_t best = 0;
    for (size_t k = 1; k < nc; ++k)
      if (us[k] < us[best]) best = k;
    capture(cand[best].p, cand[best].ns, cand[best].un);
    us_per_step = us[best] / (nc > 1 ? 2.0 : 1.0);
  }
};

}  // namespace cps

struct cps_simple {
  cps::SimpleSolver* s;
};

using cps::guard;

extern "C" {

void cps_simple_config_default(cps_simple_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->rho = 997.0;
  c->eta = 0.00089;
  c->conv_factor_25d = 1.2;
  c->sink_mult_25d = 1.0;
  c->brinkman_kmin = 1e-14;
  c->brinkman_q = 1.0;
  c->uv_alpha = 0.7;
  c->p_alpha = 0.3;
  c->momentum_tol = 1e-6;
  c->max_momentum_iter = 400;
  c->p_tol = 1e-6;
  c->max_pressure_iter = 30;
  c->pressure_solver = CPS_PRESSURE_PCG;
  c->precision = CPS_PREC_FP64;
  c->max_refine = 4;
  c->epsilon = 1e-7;
  c->max_iterations = 1000;
  c->dp_fluid_tol = 1e-6;
  c->x_core_in = 0.01;
  c->x_core_out = 0.05;
  c->use_dp_convergence = 1;
  c->dp_window = 500;
  c->dp_percent = 1.0;
  c->use_dp_slope_gate = 0;
  c->dp_slope_window = 1000;
  c->dp_slope_max_delta = 10.0;
  c->inlet_ramp = 0;
  c->ramp_steps = 1000;
}

int cps_simple_create(cps_ctx* ctx, const cps_simple_config* cfg, const double* gamma, const double* inlet_u,
                      cps_simple** out) {
  return guard([&] {
    CPS_CHECK(ctx && cfg && out, CPS_ERR_ARG, "null argument");
    CPS_HIP(hipSetDevice(ctx->c.device));
    auto* h = new cps_simple;
    try {
      h->s = new cps::SimpleSolver(&ctx->c, *cfg, gamma, inlet_u);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

void cps_simple_destroy(cps_simple* s) {
  if (!s) return;
  delete s->s;
  delete s;
}

int cps_simple_step(cps_simple* s, int n_steps, cps_step_stats* last) {
  return guard([&] {
    CPS_CHECK(s && n_steps >= 0, CPS_ERR_ARG, "bad arguments");
    s->s->step(n_steps, last);
  });
}

int cps_simple_run(cps_simple* s, cps_step_stats* last) {
  return guard([&] {
    CPS_CHECK(s, CPS_ERR_ARG, "null solver");
    s->s->run(last);
  });
}

int cps_simple_run_logged(cps_simple* s, cps_step_stats* last, cps_iter_log* log, int log_cap, int* n_iters) {
  return guard([&] {
    CPS_CHECK(s, CPS_ERR_ARG, "null solver");
    CPS_CHECK(log_cap >= 0 && (log || log_cap == 0), CPS_ERR_ARG, "log_cap needs a log of that many rows");
    const int n = s->s->run(last, log, log_cap);
    if (n_iters) *n_iters = n;
  });
}

int cps_simple_run_polled(cps_simple* s, int poll_every, cps_step_stats* last, cps_iter_log* log, int log_cap,
                          int* n_iters, int* host_syncs) {
  return guard([&] {
    CPS_CHECK(s, CPS_ERR_ARG, "null solver");
    CPS_CHECK(log_cap >= 0 && (log || log_cap == 0), CPS_ERR_ARG, "log_cap needs a log of that many rows");
    CPS_CHECK(poll_every >= 1, CPS_ERR_ARG, "poll_every must be >= 1");
    const int n = s->s->run(last, log, log_cap, poll_every);
    if (n_iters) *n_iters = n;
    if (host_syncs) *host_syncs = s->s->host_syncs;
  });
}

int cps_simple_get_field(cps_simple* s, int field, double* out) {
  return guard([&] {
    CPS_CHECK(s, CPS_ERR_ARG, "null solver");
    s->s->get(field, out);
  });
}

int cps_simple_set_field(cps_simple* s, int field, const double* in) {
  return guard([&] {
    CPS_CHECK(s, CPS_ERR_ARG, "null solver");
    s->s->set(field, in);
  });
}

int cps_simple_plane(cps_simple* s, double x, cps_patch* out) {
  return guard([&] {
    CPS_CHECK(s && out, CPS_ERR_ARG, "null argument");
    CPS_CHECK(std::isfinite(x), CPS_ERR_ARG, "x must be finite");
    s->s->plane(x, out);
  });
}

int cps_simple_time_kernel(cps_simple* s, int kernel) {
  return guard([&] {
    CPS_CHECK(s, CPS_ERR_ARG, "null solver");
    CPS_CHECK(kernel >= -1 && kernel <= 2, CPS_ERR_ARG, "unknown kernel id");
    s->s->timer.reset(kernel);
  });
}

int cps_simple_kernel_timing(cps_simple* s, int* launches, double* avg_us, double* bytes_per_launch) {
  return guard([&] {
    CPS_CHECK(s, CPS_ERR_ARG, "null solver");
    cps::KernelTimer& t = s->s->timer;
    t.drain();
    if (launches) *launches = (int)t.launches;
    if (avg_us) *avg_us = t.launches ? 1e3 * t.total_ms / (double)t.launches : 0.0;
    if (bytes_per_launch) *bytes_per_launch = t.bytes_per_launch;
  });
}

int cps_simple_gradient_dp(cps_simple* s, double rtol, int max_iter, double* dp, double* dJ_dgamma,
                           cps_solve_stats* stats) {
  return guard([&] {
    CPS_CHECK(s, CPS_ERR_ARG, "null solver");
    CPS_CHECK(rtol > 0 && max_iter >= 1, CPS_ERR_ARG, "rtol must be > 0 and max_iter >= 1");
    s->s->apply_bc();
    cps::flow_adjoint_run(s->s->adj(), s->s->view(), rtol, max_iter, dp, dJ_dgamma, stats);
  });
}

int cps_simple_gradient_vjp(cps_simple* s, const double* dJ_du, const double* dJ_dv, const double* dJ_dp,
                            double rtol, int max_iter, double* dJ_dgamma, cps_solve_stats* stats) {
  return guard([&] {
    CPS_CHECK(s && dJ_dgamma, CPS_ERR_ARG, "null argument");
    CPS_CHECK(rtol > 0 && max_iter >= 1, CPS_ERR_ARG, "rtol must be > 0 and max_iter >= 1");
    cps::flow_adjoint_vjp(s->s->adj(), s->s->view(), dJ_du, dJ_dv, dJ_dp, rtol, max_iter, dJ_dgamma, stats);
  });
}

int cps_simple_state_residual(cps_simple* s, double* out3) {
  return guard([&] {
    CPS_CHECK(s && out3, CPS_ERR_ARG, "null argument");
    s->s->apply_bc();
    cps::flow_adjoint_residual(s->s->adj(), s->s->view(), out3);
  });
}

int cps_simple_jacobian_apply(cps_simple* s, int transpose, const double* x, double* y) {
  return guard([&] {
    CPS_CHECK(s && x && y, CPS_ERR_ARG, "null argument");
    s->s->apply_bc();
    cps::flow_adjoint_apply(s->s->adj(), s->s->view(), transpose, x, y);
  });
}

int cps_simple_adjoint_options(cps_simple* s, int preconditioner, int line_sweeps, double omega,
                               int schur_iters) {
  return guard([&] {
    CPS_CHECK(s, CPS_ERR_ARG, "null argument");
    CPS_CHECK(preconditioner == CPS_ADJ_PRECOND_SIMPLE || preconditioner == CPS_ADJ_PRECOND_LSC_LINE ||
                  preconditioner == CPS_ADJ_PRECOND_LSC_ADI,
              CPS_ERR_ARG, "unknown adjoint preconditioner");
    CPS_CHECK(line_sweeps >= 0 && line_sweeps <= 64, CPS_ERR_ARG, "line_sweeps must be in [0, 64]");
    CPS_CHECK(omega > 0.0 && omega <= 1.0, CPS_ERR_ARG, "omega must be in (0, 1]");
    CPS_CHECK(schur_iters >= 1 && schur_iters <= 1000, CPS_ERR_ARG, "schur_iters must be in [1, 1000]");
    cps::flow_adjoint_options(s->s->adj(), preconditioner, line_sweeps, omega, schur_iters);
  });
}

int cps_simple_boundary_patch(cps_simple* s, int inlet, cps_patch* out) {
  return guard([&] {
    CPS_CHECK(s && out, CPS_ERR_ARG, "null argument");
    s->s->patch(inlet, out);
  });
}

}  // extern "C"

struct cps_batch {
  cps::BatchSolver* b;
};

extern "C" {

int cps_batch_create(cps_ctx* ctx, int n_fields, const cps_simple_config* cfgs, const double* const* gammas,
                     const double* const* inlets, cps_batch** out) {
  return guard([&] {
    CPS_CHECK(ctx && out, CPS_ERR_ARG, "null argument");
    CPS_HIP(hipSetDevice(ctx->c.device));
    auto* h = new cps_batch{nullptr};
    try {
      h->b = new cps::BatchSolver(&ctx->c, n_fields, cfgs, gammas, inlets);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

void cps_batch_destroy(cps_batch* b) {
  if (!b) return;
  delete b->b;
  delete b;
}

int cps_batch_tune(cps_batch* b, int streams, int reps, int* chosen, double* us_per_step) {
  return guard([&] {
    CPS_CHECK(b && streams >= 0 && reps >= 1, CPS_ERR_ARG, "bad arguments");
    b->b->tune(streams, reps);
    if (chosen) *chosen = b->b->nstreams;
    if (us_per_step) *us_per_step = b->b->us_per_step;
  });
}

int cps_batch_plan(cps_batch* b, int plan, int streams, int unroll) {
  return guard([&] {
    CPS_CHECK(b && streams >= 0, CPS_ERR_ARG, "bad arguments");
    b->b->capture(plan, streams, unroll);
    b->b->us_per_step = 0.0;
  });
}

int cps_batch_plan_info(cps_batch* b, int* plan, int* param, int* streams, int* unroll, int* classes,
                        double* us_per_step) {
  return guard([&] {
    CPS_CHECK(b, CPS_ERR_ARG, "null batch");
    if (plan) *plan = b->b->plan;
    if (param) *param = b->b->plan_ns;
    if (streams) *streams = b->b->nstreams;
    if (unroll) *unroll = b->b->unroll;
    if (classes) *classes = (int)b->b->groups.size();
    if (us_per_step) *us_per_step = b->b->us_per_step;
  });
}

int cps_batch_step(cps_batch* b, int n_steps) {
  return guard([&] {
    CPS_CHECK(b && n_steps >= 0, CPS_ERR_ARG, "bad arguments");
    b->b->launch(n_steps);
  });
}

int cps_batch_synchronize(cps_batch* b) {
  return guard([&] {
    CPS_CHECK(b, CPS_ERR_ARG, "null batch");
    b->b->sync_all();
  });
}

int cps_batch_field_stats(cps_batch* b, int field, cps_step_stats* out) {
  return guard([&] {
    CPS_CHECK(b && out && field >= 0 && field < (int)b->b->f.size(), CPS_ERR_ARG, "bad field index");
    b->b->sync_all();
    b->b->f[field]->stats(out);
  });
}

int cps_batch_get_field(cps_batch* b, int field, int which, double* out) {
  return guard([&] {
    CPS_CHECK(b && field >= 0 && field < (int)b->b->f.size(), CPS_ERR_ARG, "bad field index");
    b->b->sync_all();
    b->b->f[field]->get(which, out);
  });
}

int cps_batch_gradient_dp(cps_batch* b, int field, double rtol, int max_iter, double* dp, double* dJ_dgamma,
                          cps_solve_stats* stats) {
  return guard([&] {
    CPS_CHECK(b && field >= 0 && field < (int)b->b->f.size(), CPS_ERR_ARG, "bad field index");
    CPS_CHECK(rtol > 0 && max_iter >= 1, CPS_ERR_ARG, "rtol must be > 0 and max_iter >= 1");
    cps::SimpleSolver& s = *b->b->f[field];
    // the field's data are written on its class leader's stream (class plan)
    CPS_HIP(hipStreamSynchronize(b->b->f[b->b->leader[field]]->ctx->stream));
    CPS_HIP(hipStreamSynchronize(s.ctx->stream));
    s.apply_bc();
    cps::flow_adjoint_run(s.adj(), s.view(), rtol, max_iter, dp, dJ_dgamma, stats);
  });
}

int cps_batch_boundary_patch(cps_batch* b, int field, int inlet, cps_patch* out) {
  return guard([&] {
    CPS_CHECK(b && out && field >= 0 && field < (int)b->b->f.size(), CPS_ERR_ARG, "bad field index");
    b->b->sync_all();
    b->b->f[field]->patch(inlet, out);
  });
}

}  // extern "C"
