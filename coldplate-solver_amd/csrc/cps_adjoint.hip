// cps_adjoint.hip -- discrete adjoint of the converged SIMPLE state:
// d(pressure drop)/d(gamma) on gfx950.
//
// The reference has no flow adjoint (its sensitivities ignore the flow).  At
// a converged SIMPLE iteration the relaxation cancels and the state w = (u, v,
// p) satisfies the discrete equations the solver iterates on:
//   U / V momentum  aP0 x_P - sum_nb a_nb x_nb - (p_P - p_E) h - Sdc = 0
//                   (momentum_solver.cpp:195-314, convection.cpp:53-186),
//   continuity      rho hy (u_W - u_E) + rho hx (v_S - v_N) = 0 on the
//                   pressure DOFs (pressure_solver.cpp:470-496),
// with the boundary rules of boundaries.cpp:50-146 (ghost values resolve to
// the DOF they copy).  The objective is the full static pressure drop of
// sampleBoundaryPatch (postprocessing.cpp:245-327): the mean of p(i, 0) =
// p(i, 1) over the pure-fluid inlet rows minus the outlet mean (p = 0).
//
//   J^T lambda = dJ/dw,   dJ/dgamma = - lambda^T dR/dgamma,
//
// where R depends on gamma through the Brinkman term alpha(gamma) vol x_P of
// aP0 (the inlet mask, wall flags and row selection are piecewise constant).
//
// J is built exactly, one thread per equation, by evaluating the equation
// with 16-slot forward dual numbers seeded on the DOFs it reads; it is stored
// row-wise (ELL) and transposed on the device with a counting pass and a
// per-row sort (deterministic).  J^T = [[A, B], [C, 0]] (momentum rows, then
// continuity rows) is solved by right-preconditioned FGMRES with a block
// factorisation preconditioner
//   x* = A~^-1 f_m,  y = S~^-1 (C x* - f_p),  x = x* - A~^-1 B y.
// Three forms:
//   LSC_ADI (default) A~^-1 = line Jacobi on A: pentadiagonal solves along
//                     grid rows (x-lines) and columns (y-lines), an x pass,
//                     a y pass, then damped rounds of both; S~^-1 the
//                     least-squares-commutator approximation
//                     L^-1 (C D^-1 A D^-1 B) L^-1 of S = C A^-1 B, D = diag(A),
//                     L = C D^-1 B, which is exactly the SIMPLE
//                     pressure-correction operator with dE = hy / D_u,
//                     dN = hx / D_v, solved by a few geometric-multigrid PCG
//                     iterations;
//   LSC_LINE          the same with x-lines only (fewer passes, but its
//                     sweeps diverge where the flow turns across the rows);
//   SIMPLE            A~ = diag(aP0) (aP0 taken from the Jacobian kernel),
//                     S~ = C diag(aP0)^-1 B (the same multigrid operator with
//                     aP0 in place of D): 5-20x more iterations.
// Vectors are three padded fields [u | v | p]; rows that are not DOFs are
// identity rows.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "cps_mg.hpp"
#include "cps_simple.hpp"

namespace cps {
namespace adj {

constexpr int K = 16;          // dual slots: distinct DOFs one equation reads
constexpr int BS = 256;        // threads per block of the vector kernels
constexpr double kVelTol = (double)1e-8f;
constexpr double kAP0Floor = (double)1e-10f;

struct Dual {
  double v;
  double d[K];
};

__device__ inline Dual dconst(double v) {
  Dual r;
  r.v = v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = 0.0;
  return r;
}
__device__ inline Dual operator+(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
__device__ inline Dual operator-(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
__device__ inline Dual operator-(const Dual& a) { return dconst(0.0) - a; }
__device__ inline Dual operator*(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.v * b.d[k] + b.v * a.d[k];
  return r;
}
__device__ inline Dual operator*(double s, const Dual& a) {
  Dual r;
  r.v = s * a.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = s * a.d[k];
  return r;
}
__device__ inline Dual operator+(const Dual& a, double s) {
  Dual r = a;
  r.v += s;
  return r;
}
// max(0, a): derivative of the selected branch (0 at the tie, like fmax)
__device__ inline Dual pos(const Dual& a) { return a.v > 0.0 ? a : dconst(0.0); }
__device__ inline Dual clampd(const Dual& a, double m) {
  if (a.v > m) return dconst(m);
  if (a.v < -m) return dconst(-m);
  return a;
}

// slot table: global unknown ids of the DOFs an equation reads
struct Tape {
  int id[K];
  int n;
  int overflow;
};

__device__ inline Dual var(Tape& t, int gid, double val) {
  int s = -1;
  for (int k = 0; k < t.n; ++k)
    if (t.id[k] == gid) s = k;
  if (s < 0) {
    if (t.n == K) {
      t.overflow = 1;
      return dconst(val);
    }
    s = t.n;
    t.id[t.n++] = gid;
  }
  Dual r;
  r.v = val;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = k == s ? 1.0 : 0.0;
  return r;
}

struct View {
  Geo g;
  size_t S;  // elements per padded field; unknown id = comp * S + i * P + j
  const double *u, *v, *p, *gam, *alp, *inlet, *dE, *dN;
};

__device__ inline size_t at(const Geo& g, int i, int j) { return (size_t)i * g.P + j; }
__device__ inline bool u_dof(const Geo& g, int i, int j) { return i >= 1 && i <= g.M - 1 && j >= 1 && j <= g.N - 2; }
__device__ inline bool v_dof(const Geo& g, int i, int j) { return i >= 1 && i <= g.M - 2 && j >= 1 && j <= g.N - 1; }
__device__ inline bool p_dof(const Geo& g, int i, int j) { return i >= 1 && i <= g.M - 1 && j >= 1 && j <= g.N - 2; }

// u(i, j) through the boundary rules: inlet column (prescribed), outlet
// column N-1 = u(i, N-2), walls i = 0, M
__device__ inline Dual readU(const View& V, Tape& t, int i, int j) {
  const Geo& g = V.g;
  if (j == 0) return dconst(V.inlet[i]);
  if (i == 0 || i == g.M) return dconst(0.0);
  if (j == g.N - 1) j = g.N - 2;
  return var(t, (int)(0 * V.S + at(g, i, j)), V.u[at(g, i, j)]);
}
// v(i, j): walls i = 0, M-1; v(i, 0) = 0; v(i, N) = v(i, N-1)
__device__ inline Dual readV(const View& V, Tape& t, int i, int j) {
  const Geo& g = V.g;
  if (i == 0 || i == g.M - 1 || j == 0) return dconst(0.0);
  if (j == g.N) j = g.N - 1;
  return var(t, (int)(1 * V.S + at(g, i, j)), V.v[at(g, i, j)]);
}
// p(i, j): columns N-1 (never updated) and N (outlet) are 0; the top, bottom
// and left ghosts copy their inner neighbour
__device__ inline Dual readP(const View& V, Tape& t, int i, int j) {
  const Geo& g = V.g;
  if (j >= g.N - 1) return dconst(0.0);
  if (j == 0) j = 1;
  if (i == 0) i = 1;
  if (i == g.M) i = g.M - 1;
  return var(t, (int)(2 * V.S + at(g, i, j)), V.p[at(g, i, j)]);
}

// momentum residual of one U (IsU) or V DOF, mirroring the assembly of the
// SIMPLE step
template <bool IsU>
__device__ Dual mom_residual(const View& V, Tape& t, int i, int j, double* ap0) {
  const Geo& g = V.g;
  const int M = g.M, N = g.N;
  auto U = [&](int a, int b) { return readU(V, t, a, b); };
  auto W = [&](int a, int b) { return readV(V, t, a, b); };
  auto P = [&](int a, int b) { return readP(V, t, a, b); };
  auto A = [&](int a, int b) { return V.alp[at(g, a, b)]; };
  auto G = [&](int a, int b) { return V.gam[at(g, a, b)]; };
  auto Ur = [&](int a, int b) { return V.u[at(g, a, b)]; };
  auto Vr = [&](int a, int b) { return V.v[at(g, a, b)]; };
  const double mv = g.maxVel;
  double alphaLocal;
  Dual ue, uw, vn, vs;
  if (IsU) {
    alphaLocal = 0.5 * (A(i - 1, j - 1) + A(i - 1, j));
    ue = clampd(0.5 * (U(i, j) + U(i, j + 1)), mv);
    uw = clampd(0.5 * (U(i, j - 1) + U(i, j)), mv);
    vn = clampd(0.5 * (W(i, j) + W(i, j + 1)), mv);
    vs = clampd(0.5 * (W(i - 1, j) + W(i - 1, j + 1)), mv);
  } else {
    alphaLocal = 0.5 * (A(i - 1, j - 1) + A(i, j - 1));
    ue = clampd(0.5 * (U(i, j) + U(i + 1, j)), mv);
    uw = clampd(0.5 * (U(i, j - 1) + U(i + 1, j - 1)), mv);
    vn = clampd(0.5 * (W(i, j) + W(i + 1, j)), mv);
    vs = clampd(0.5 * (W(i - 1, j) + W(i, j)), mv);
  }
  const Dual Fe = (g.convScale * g.rho * g.hy) * ue, Fw = (g.convScale * g.rho * g.hy) * uw;
  const Dual Fn = (g.convScale * g.rho * g.hx) * vn, Fs = (g.convScale * g.rho * g.hx) * vs;
  Dual aE = pos(-Fe) + g.De;
  Dual aW = pos(Fw) + g.De;
  Dual aN = pos(-Fn) + g.Dn;
  Dual aS = pos(Fs) + g.Dn;
  if (IsU) {
    if (i == 1 && face_fluid(G(0, j - 1), G(0, j)) && fabs(Ur(0, j)) <= kVelTol &&
        fabs(Vr(0, j)) <= kVelTol && fabs(Vr(0, j + 1)) <= kVelTol)
      aS = aS + g.Dn;
    if (i == M - 1 && face_fluid(G(M - 2, j - 1), G(M - 2, j)) && fabs(Ur(M, j)) <= kVelTol &&
        fabs(Vr(M - 1, j)) <= kVelTol && fabs(Vr(M - 1, j + 1)) <= kVelTol)
      aN = aN + g.Dn;
  } else {
    if (j == 1 && face_fluid(G(i - 1, 0), G(i, 0)) && !(fabs(g.targetVel) > kVelTol) &&
        fabs(Vr(i, 0)) <= kVelTol && fabs(Ur(i, 0)) <= kVelTol && fabs(Ur(i + 1, 0)) <= kVelTol)
      aW = aW + g.De;
    if (j == N - 1 && face_fluid(G(i - 1, N - 2), G(i, N - 2)) &&
        !(fabs(V.p[at(g, i, N)]) <= kVelTol) && fabs(Vr(i, N)) <= kVelTol && fabs(Ur(i, N - 1)) <= kVelTol &&
        fabs(Ur(i + 1, N - 1)) <= kVelTol)
      aE = aE + g.De;
  }
  Dual aP0 = (aE + aW + aN + aS) + (Fe - Fw + Fn - Fs) + (g.sinkDiag + alphaLocal * g.vol);
  if (!(aP0.v > kAP0Floor)) aP0 = dconst(kAP0Floor);
  *ap0 = aP0.v;
  Dual xP, xE, xW, xN, xS, Sp;
  if (IsU) {
    xP = U(i, j); xE = U(i, j + 1); xW = U(i, j - 1); xN = U(i + 1, j); xS = U(i - 1, j);
    Sp = g.hy * (P(i, j) - P(i, j + 1));
  } else {
    xP = W(i, j); xE = W(i, j + 1); xW = W(i, j - 1); xN = W(i + 1, j); xS = W(i - 1, j);
    Sp = g.hx * (P(i, j) - P(i + 1, j));
  }
  if (g.sou) {
    // second-order upwind deferred correction (convection.cpp:53-186)
    Dual Sdc = dconst(0.0);
    auto X = [&](int a, int b) { return IsU ? U(a, b) : W(a, b); };
    auto corr = [&](const Dual& up, const Dual& upup) { return (1.5 * up - 0.5 * upup) - up; };
    if (IsU) {
      if (Fe.v >= 0.0) { if (j >= 2) Sdc = Sdc - Fe * corr(X(i, j), X(i, j - 1)); }
      else { if (j + 2 < N) Sdc = Sdc - Fe * corr(X(i, j + 1), X(i, j + 2)); }
      if (Fw.v >= 0.0) { if (j >= 2) Sdc = Sdc + Fw * corr(X(i, j - 1), X(i, j - 2)); }
      else { if (j + 1 < N) Sdc = Sdc + Fw * corr(X(i, j), X(i, j + 1)); }
      if (Fn.v >= 0.0) { if (i >= 1) Sdc = Sdc - Fn * corr(X(i, j), X(i - 1, j)); }
      else { if (i + 2 <= M) Sdc = Sdc - Fn * corr(X(i + 1, j), X(i + 2, j)); }
      if (Fs.v >= 0.0) { if (i >= 2) Sdc = Sdc + Fs * corr(X(i - 1, j), X(i - 2, j)); }
      else { if (i + 1 <= M) Sdc = Sdc + Fs * corr(X(i, j), X(i + 1, j)); }
    } else {
      if (Fe.v >= 0.0) { if (j >= 2) Sdc = Sdc - Fe * corr(X(i, j), X(i, j - 1)); }
      else { if (j + 2 <= N) Sdc = Sdc - Fe * corr(X(i, j + 1), X(i, j + 2)); }
      if (Fw.v >= 0.0) { if (j >= 2) Sdc = Sdc + Fw * corr(X(i, j - 1), X(i, j - 2)); }
      else { if (j + 1 <= N) Sdc = Sdc + Fw * corr(X(i, j), X(i, j + 1)); }
      if (Fn.v >= 0.0) { if (i >= 1) Sdc = Sdc - Fn * corr(X(i, j), X(i - 1, j)); }
      else { if (i + 2 < M) Sdc = Sdc - Fn * corr(X(i + 1, j), X(i + 2, j)); }
      if (Fs.v >= 0.0) { if (i >= 2) Sdc = Sdc + Fs * corr(X(i - 1, j), X(i - 2, j)); }
      else { if (i + 1 < M) Sdc = Sdc + Fs * corr(X(i, j), X(i + 1, j)); }
    }
    Sp = Sp + Sdc;
  }
  return aP0 * xP - (aE * xE + aW * xW + aN * xN + aS * xS) - Sp;
}

// mass residual of a pressure DOF (pressure_solver.cpp:470-496)
__device__ Dual mass_residual(const View& V, Tape& t, int i, int j) {
  const Geo& g = V.g;
  const Dual mE = (g.rho * g.hy) * readU(V, t, i, j), mW = (g.rho * g.hy) * readU(V, t, i, j - 1);
  const Dual mN = (g.rho * g.hx) * readV(V, t, i, j), mS = (g.rho * g.hx) * readV(V, t, i - 1, j);
  return mW - mE + mS - mN;
}

// one equation per thread; J stored slot-major: col[s * n + row], val[...]
// (aP0 of each momentum equation into dinv as 1/aP0; 1 on other rows)
__global__ __launch_bounds__(BS) void k_rows(View V, int n, int* col, double* val, double* res, int* flag,
                                            double* dinv) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  const Geo& g = V.g;
  const int comp = r / (int)V.S, q = r % (int)V.S, i = q / g.P, j = q % g.P;
  Tape t;
  t.n = 0;
  t.overflow = 0;
  Dual R;
  bool dof = false;
  double ap0 = 1.0;
  if (comp == 0 && u_dof(g, i, j)) { R = mom_residual<true>(V, t, i, j, &ap0); dof = true; }
  else if (comp == 1 && v_dof(g, i, j)) { R = mom_residual<false>(V, t, i, j, &ap0); dof = true; }
  else if (comp == 2 && p_dof(g, i, j)) { R = mass_residual(V, t, i, j); dof = true; }
  if (!dof) {  // identity row
    t.n = 1;
    t.id[0] = r;
    R = dconst(0.0);
    R.d[0] = 1.0;
  }
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const bool used = s < t.n && R.d[s] != 0.0;
    col[(size_t)s * n + r] = used ? t.id[s] : -1;
    val[(size_t)s * n + r] = used ? R.d[s] : 0.0;
  }
  res[r] = dof ? R.v : 0.0;
  flag[r] = t.overflow;
  dinv[r] = 1.0 / ap0;
}

__global__ __launch_bounds__(BS) void k_count(int n, const int* col, int* cnt) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  for (int s = 0; s < K; ++s) {
    const int c = col[(size_t)s * n + r];
    if (c >= 0) atomicAdd(&cnt[c], 1);
  }
}

// max of cnt[0..n) and of flag[0..n): block maxima folded into out[0..1]
// (zeroed beforehand) with integer atomics
__global__ __launch_bounds__(BS) void k_imax(int n, const int* cnt, const int* flag, int* out) {
  __shared__ int sh[BS / kWave + 1];
  int a = 0, b = 0;
  for (int r = blockIdx.x * BS + threadIdx.x; r < n; r += gridDim.x * BS) {
    a = max(a, cnt[r]);
    b = max(b, flag[r]);
  }
  a = block_max<BS>(a, sh);
  b = block_max<BS>(b, sh);
  if (threadIdx.x == 0) {
    atomicMax(&out[0], a);
    atomicMax(&out[1], b);
  }
}

__global__ __launch_bounds__(BS) void k_scatter(int n, const int* col, const double* val, int* cur, int WT,
                                               int* colT, double* valT) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  for (int s = 0; s < K; ++s) {
    const int c = col[(size_t)s * n + r];
    if (c < 0) continue;
    const int pos = atomicAdd(&cur[c], 1);
    if (pos < WT) {
      colT[(size_t)pos * n + c] = r;
      valT[(size_t)pos * n + c] = val[(size_t)s * n + r];
    }
  }
}

// orders each row of J^T by source equation (fixes the summation order)
__global__ __launch_bounds__(BS) void k_sort(int n, int WT, const int* cnt, int* colT, double* valT) {
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= n) return;
  const int m = min(cnt[c], WT);
  for (int a = 1; a < m; ++a) {
    const int kc = colT[(size_t)a * n + c];
    const double kv = valT[(size_t)a * n + c];
    int b = a - 1;
    while (b >= 0 && colT[(size_t)b * n + c] > kc) {
      colT[(size_t)(b + 1) * n + c] = colT[(size_t)b * n + c];
      valT[(size_t)(b + 1) * n + c] = valT[(size_t)b * n + c];
      --b;
    }
    colT[(size_t)(b + 1) * n + c] = kc;
    valT[(size_t)(b + 1) * n + c] = kv;
  }
}

// y = J x (row form)
__global__ __launch_bounds__(BS) void k_spmv_rows(int n, const int* col, const double* val, const double* x,
                                                 double* y) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  double a = 0.0;
  for (int s = 0; s < K; ++s) {
    const int c = col[(size_t)s * n + r];
    if (c >= 0) a += val[(size_t)s * n + r] * x[c];
  }
  y[r] = a;
}

// y = J^T x on rows [0, rows) (n: the slot stride, all rows)
__global__ __launch_bounds__(BS) void k_spmv(int n, int rows, int WT, const int* cnt, const int* colT,
                                            const double* valT, const double* x, double* y) {
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= rows) return;
  const int m = min(cnt[c], WT);
  double a = 0.0;
  for (int s = 0; s < m; ++s) a += valT[(size_t)s * n + c] * x[colT[(size_t)s * n + c]];
  y[c] = a;
}

// ---------------------------------------------------------------------------
// Krylov helpers (grid-stride, fixed-order block partials)
constexpr int kMaxBasis = 64;
struct Coefs {
  double c[kMaxBasis];
};

// part[j * nb + b] = block b's partial of V_j . w, j < nv
__global__ __launch_bounds__(BS) void k_mdot(int n, const double* Vb, int nv, const double* w, double* part) {
  __shared__ double sh[BS / kWave + 1];
  for (int jv = 0; jv < nv; ++jv) {
    const double* vj = Vb + (size_t)jv * n;
    double a = 0.0;
    for (int r = blockIdx.x * BS + threadIdx.x; r < n; r += gridDim.x * BS) a += vj[r] * w[r];
    a = block_sum<BS>(a, sh);
    if (threadIdx.x == 0) part[(size_t)jv * gridDim.x + blockIdx.x] = a;
  }
}

// out[s] = sum of slot s's nb partials in fixed order, one block per slot
__global__ __launch_bounds__(BS) void k_psum(const double* part, int nb, double* out) {
  __shared__ double sh[BS / kWave + 1];
  const int s = blockIdx.x;
  const double a = parts_sum<BS>(part + (size_t)s * nb, nb, sh);
  if (threadIdx.x == 0) out[s] = a;
}

// w = a0 * w + sum_j c_j V_j
__global__ __launch_bounds__(BS) void k_mcomb(int n, double a0, double* w, const double* Vb, int nv, Coefs c) {
  for (int r = blockIdx.x * BS + threadIdx.x; r < n; r += gridDim.x * BS) {
    double a = a0 == 0.0 ? 0.0 : a0 * w[r];
    for (int jv = 0; jv < nv; ++jv) a += c.c[jv] * Vb[(size_t)jv * n + r];
    w[r] = a;
  }
}

// w -= sum_j c_j V_j with the coefficients on the device (same arithmetic
// as k_mcomb with a0 = 1 and c = -coef)
__global__ __launch_bounds__(BS) void k_mcomb_dev(int n, double* w, const double* Vb, int nv, const double* coef) {
  for (int r = blockIdx.x * BS + threadIdx.x; r < n; r += gridDim.x * BS) {
    double a = w[r];
    for (int jv = 0; jv < nv; ++jv) a += (-coef[jv]) * Vb[(size_t)jv * n + r];
    w[r] = a;
  }
}
// y = x / sqrt(*ss) when *ss > 0 (the next Arnoldi vector)
__global__ __launch_bounds__(BS) void k_scale_rn(int n, const double* x, const double* ss, double* y) {
  const double q = *ss;
  if (!(q > 0.0)) return;
  const double a = 1.0 / sqrt(q);
  for (int r = blockIdx.x * BS + threadIdx.x; r < n; r += gridDim.x * BS) y[r] = a * x[r];
}

__global__ __launch_bounds__(BS) void k_fill(int n, double a, double* y) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r < n) y[r] = a;
}
__global__ __launch_bounds__(BS) void k_scale(int n, const double* x, double a, double* y) {
  for (int r = blockIdx.x * BS + threadIdx.x; r < n; r += gridDim.x * BS) y[r] = a * x[r];
}

// y = b - y
__global__ __launch_bounds__(BS) void k_rsub(int n, const double* b, double* y) {
  for (int r = blockIdx.x * BS + threadIdx.x; r < n; r += gridDim.x * BS) y[r] = b[r] - y[r];
}

// ---------------------------------------------------------------------------
// preconditioner pieces
// faces of the Schur operator in the pressure-correction layout: dE = hy / aP0
// on U DOFs, dN = hx / aP0 on V DOFs, 0 elsewhere
__global__ __launch_bounds__(BS) void k_faces(View V, const double* dinv, double* dEa, double* dNa) {
  const int q = blockIdx.x * BS + threadIdx.x;
  if (q >= (int)V.S) return;
  const Geo& g = V.g;
  const int i = q / g.P, j = q % g.P;
  dEa[q] = u_dof(g, i, j) ? g.hy * dinv[q] : 0.0;
  dNa[q] = v_dof(g, i, j) ? g.hx * dinv[V.S + q] : 0.0;
}

// stage 1: t = [dinv f_m ; 0] (momentum part), and z_nondof = f on non-DOF rows
__global__ __launch_bounds__(BS) void k_pre1(View V, int n, const double* dinv, const double* f, double* t) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  const int comp = r / (int)V.S;
  t[r] = comp < 2 ? dinv[r] * f[r] : 0.0;
}
// stage 2: pressure rhs = (J^T t)_p - f_p (f may be null) on pressure DOFs (field layout), x_p = 0
__global__ __launch_bounds__(BS) void k_pre2(View V, const double* Jt, const double* f, double* rp, double* xp) {
  const int q = blockIdx.x * BS + threadIdx.x;
  if (q >= (int)V.S) return;
  const Geo& g = V.g;
  const int i = q / g.P, j = q % g.P;
  const size_t r = 2 * V.S + q;
  rp[q] = p_dof(g, i, j) ? Jt[r] - (f ? f[r] : 0.0) : 0.0;
  xp[q] = 0.0;
}
// stage 3: s = [0 ; y] for the second product
__global__ __launch_bounds__(BS) void k_pre3(View V, int n, const double* y, double* s) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  const Geo& g = V.g;
  const int comp = r / (int)V.S, q = r % (int)V.S, i = q / g.P, j = q % g.P;
  s[r] = comp == 2 && p_dof(g, i, j) ? y[q] : 0.0;
}
// stage 4: z = [x* - dinv (J^T s)_m ; y] on DOFs, f on the identity rows
__global__ __launch_bounds__(BS) void k_pre4(View V, int n, const double* dinv, const double* f, const double* t,
                                            const double* Js, const double* y, double* z) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  const Geo& g = V.g;
  const int comp = r / (int)V.S, q = r % (int)V.S, i = q / g.P, j = q % g.P;
  double out;
  if (comp == 0) out = u_dof(g, i, j) ? t[r] - dinv[r] * Js[r] : f[r];
  else if (comp == 1) out = v_dof(g, i, j) ? t[r] - dinv[r] * Js[r] : f[r];
  else out = p_dof(g, i, j) ? y[q] : f[r];
  z[r] = out;
}

// ---------------------------------------------------------------------------
// Line preconditioners of the momentum block A of J^T.  The momentum
// equations couple a DOF to its neighbours two cells away in each direction
// (second-order upwinding), so every grid row (x-line) and every grid column
// (y-line) of u (i = 1..M-1, j = 1..N-2) and of v (i = 1..M-2, j = 1..N-1) is
// a pentadiagonal system; couplings off the line and between u and v stay
// explicit.  A sweep is x += omega Dl^-1 (f - A x).  Dropping the +-2 entries
// (tridiagonal lines) makes the sweeps diverge on SOU states; undamped x-line
// sweeps alone diverge where the flow turns across the lines (the v DOFs of
// an inlet bend carry an eigenvalue -0.53 +- 1.0i of I - Dl^-1 A), which
// damping and alternating x / y lines cure.
// Lines of one direction: Y = false rows (x-lines), Y = true columns.
template <bool Y>
__device__ inline int n_lines(const Geo& g) {
  return Y ? (g.N - 2) + (g.N - 1) : (g.M - 1) + (g.M - 2);
}
// first row (in the [u | v | p] numbering), length and stride of line l
template <bool Y>
__device__ inline void line_span(const View& V, int l, size_t* base, int* len, int* step) {
  const Geo& g = V.g;
  if (!Y) {
    *step = 1;
    if (l < g.M - 1) {
      *base = (size_t)(l + 1) * g.P + 1;
      *len = g.N - 2;
    } else {
      *base = V.S + (size_t)(l - (g.M - 1) + 1) * g.P + 1;
      *len = g.N - 1;
    }
  } else {
    *step = g.P;
    if (l < g.N - 2) {
      *base = (size_t)g.P + (l + 1);
      *len = g.M - 1;
    } else {
      *base = V.S + (size_t)g.P + (l - (g.N - 2) + 1);
      *len = g.M - 2;
    }
  }
}

// band[k * 2S + r] = entry (r, r + (k - 2) step) of J^T on momentum DOF rows
// (same component, same line); dj = 1 / diagonal
template <bool Y>
__global__ __launch_bounds__(BS) void k_band(View V, int n, int WT, const int* cnt, const int* colT,
                                            const double* valT, double* band, double* dj) {
  const int r = blockIdx.x * BS + threadIdx.x;
  const size_t S2 = 2 * V.S;
  if (r >= (int)S2) return;
  const Geo& g = V.g;
  const int comp = r / (int)V.S, q = r % (int)V.S, i = q / g.P, j = q % g.P;
  double b[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const bool dof = comp == 0 ? u_dof(g, i, j) : v_dof(g, i, j);
  if (dof) {
    const int m = min(cnt[r], WT);
    for (int s = 0; s < m; ++s) {
      const int c = colT[(size_t)s * n + r];
      if (c / (int)V.S != comp) continue;
      const int qc = c % (int)V.S, ic = qc / g.P, jc = qc % g.P;
      const int off = Y ? ic - i : jc - j;
      const bool same_line = Y ? jc == j : ic == i;
      if (same_line && off >= -2 && off <= 2) b[off + 2] += valT[(size_t)s * n + r];
    }
  } else {
    b[2] = 1.0;
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) band[k * S2 + r] = b[k];
  if (dj) dj[r] = b[2] != 0.0 ? 1.0 / b[2] : 1.0;
}

// in-place banded LU without pivoting, one thread per line:
// band slots become [l2, l1, 1/u0, u1, u2]
template <bool Y>
__global__ __launch_bounds__(64) void k_line_factor(View V, double* band, int* bad) {
  const int l = blockIdx.x * 64 + threadIdx.x;
  if (l >= n_lines<Y>(V.g)) return;
  size_t base;
  int len, step;
  line_span<Y>(V, l, &base, &len, &step);
  const size_t S2 = 2 * V.S;
  double* A2 = band;
  double* A1 = band + S2;
  double* D = band + 2 * S2;
  double* C1 = band + 3 * S2;
  double* C2 = band + 4 * S2;
  // U rows j-1 (p) and j-2 (q): (u0, u1, u2)
  double p0 = 1.0, p1 = 0.0, p2 = 0.0, q0 = 1.0, q1 = 0.0, q2 = 0.0;
  int fail = 0;
  for (int j = 0; j < len; ++j) {
    const size_t r = base + (size_t)j * step;
    const double a2 = j >= 2 ? A2[r] : 0.0, a1 = j >= 1 ? A1[r] : 0.0;
    const double c1 = j + 1 < len ? C1[r] : 0.0, c2 = j + 2 < len ? C2[r] : 0.0;
    const double l2 = a2 / q0;
    const double l1 = (a1 - l2 * q1) / p0;
    const double u0 = D[r] - l2 * q2 - l1 * p1;
    const double u1 = c1 - l1 * p2;
    if (!(std::fabs(u0) > 0.0) || !std::isfinite(u0)) fail = 1;
    A2[r] = l2;
    A1[r] = l1;
    D[r] = 1.0 / u0;
    C1[r] = u1;
    C2[r] = c2;
    q0 = p0; q1 = p1; q2 = p2;
    p0 = u0; p1 = u1; p2 = c2;
  }
  if (fail) atomicOr(bad, 1);
}

// 2 x 2 affine map s -> M s + c of the substitution state s = (x_j, x_j-1)
struct Aff {
  double m00, m01, m10, m11, c0, c1;
};
__device__ inline Aff aff_id() { return {1.0, 0.0, 0.0, 1.0, 0.0, 0.0}; }
// one recurrence step x' = beta - p x1 - q x2 applied after map a
__device__ inline Aff aff_step(const Aff& a, double p, double q, double beta) {
  return {-p * a.m00 - q * a.m10, -p * a.m01 - q * a.m11, a.m00, a.m01, beta - p * a.c0 - q * a.c1, a.c0};
}
// k o j (j first)
__device__ inline Aff aff_compose(const Aff& k, const Aff& j) {
  return {k.m00 * j.m00 + k.m01 * j.m10, k.m00 * j.m01 + k.m01 * j.m11,
          k.m10 * j.m00 + k.m11 * j.m10, k.m10 * j.m01 + k.m11 * j.m11,
          k.m00 * j.c0 + k.m01 * j.c1 + k.c0, k.m10 * j.c0 + k.m11 * j.c1 + k.c1};
}
template <bool Up>
__device__ inline Aff aff_shfl(const Aff& a, int off) {
  if (Up)
    return {__shfl_up(a.m00, off), __shfl_up(a.m01, off), __shfl_up(a.m10, off), __shfl_up(a.m11, off),
            __shfl_up(a.c0, off), __shfl_up(a.c1, off)};
  return {__shfl_down(a.m00, off), __shfl_down(a.m01, off), __shfl_down(a.m10, off), __shfl_down(a.m11, off),
          __shfl_down(a.c0, off), __shfl_down(a.c1, off)};
}

// x = Dl^-1 f (Ax null) or x += omega Dl^-1 (f - Ax) on momentum DOF rows,
// one wavefront per line.  Both substitutions of the factored line are
// second-order linear recurrences; a line is cut into tiles of 64 x kSeg
// elements, lane k owning kSeg consecutive ones.  Per tile each lane folds
// its segment into one affine map of the state (x_j, x_j-1), a 6-step
// shuffle scan composes the maps of the lanes before it (after it for the
// backward substitution), and each lane then replays its segment from the
// state entering it.  The forward sweep goes to tmp; every element is written
// and read back by the same lane.  4 lines per 256-thread block; y-line
// elements lie a row pitch apart, so the y-line blocks run in XCD-aware
// order (xcd_band): the four blocks of 16 adjacent columns, which share
// every row's cache line, run on one XCD and meet in its L2.
constexpr int kSeg = 8;
template <bool Y>
__global__ __launch_bounds__(256) void k_line_solve(View V, const double* band, const double* f, const double* Ax,
                                                   double omega, double* x, double* tmp) {
  const int lane = threadIdx.x & (kWave - 1);
  const int l = (Y ? xcd_band() : (int)blockIdx.x) * (256 / kWave) + threadIdx.x / kWave;
  if (l >= n_lines<Y>(V.g)) return;  // uniform per wavefront
  size_t base;
  int len, step;
  line_span<Y>(V, l, &base, &len, &step);
  const bool acc = Ax != nullptr;
  const size_t S2 = 2 * V.S;
  const double* L2 = band;
  const double* L1 = band + S2;
  const double* U0 = band + 2 * S2;
  const double* U1 = band + 3 * S2;
  const double* U2 = band + 4 * S2;
  constexpr int TILE = kWave * kSeg;
  // forward substitution y_j = b_j - l1_j y_j-1 - l2_j y_j-2
  double s0 = 0.0, s1 = 0.0;  // state carried into the tile
  for (int t0 = 0; t0 < len; t0 += TILE) {
    const int j0 = t0 + lane * kSeg;
    double bb[kSeg], p[kSeg], q[kSeg];
#pragma unroll
    for (int k = 0; k < kSeg; ++k) {
      const int j = j0 + k;
      if (j < len) {
        const size_t r = base + (size_t)j * step;
        bb[k] = acc ? f[r] - Ax[r] : f[r];
        p[k] = L1[r];
        q[k] = L2[r];
      } else {
        bb[k] = p[k] = q[k] = 0.0;
      }
    }
    Aff a = aff_id();
#pragma unroll
    for (int k = 0; k < kSeg; ++k)
      if (j0 + k < len) a = aff_step(a, p[k], q[k], bb[k]);
    for (int off = 1; off < kWave; off <<= 1) {
      const Aff o = aff_shfl<true>(a, off);
      if (lane >= off) a = aff_compose(a, o);
    }
    const Aff before = aff_shfl<true>(a, 1);  // lanes 0..k-1
    double y1 = s0, y2 = s1;
    if (lane > 0) {
      y1 = before.m00 * s0 + before.m01 * s1 + before.c0;
      y2 = before.m10 * s0 + before.m11 * s1 + before.c1;
    }
#pragma unroll
    for (int k = 0; k < kSeg; ++k) {
      const int j = j0 + k;
      if (j < len) {
        const double yj = bb[k] - p[k] * y1 - q[k] * y2;
        tmp[base + (size_t)j * step] = yj;
        y2 = y1;
        y1 = yj;
      }
    }
    s0 = __shfl(y1, kWave - 1);
    s1 = __shfl(y2, kWave - 1);
  }
  // backward substitution x_j = (y_j - u1_j x_j+1 - u2_j x_j+2) / u0_j
  s0 = 0.0;
  s1 = 0.0;
  const int ntiles = (len + TILE - 1) / TILE;
  for (int t = ntiles - 1; t >= 0; --t) {
    const int j0 = t * TILE + lane * kSeg;
    double yy[kSeg], d[kSeg], p[kSeg], q[kSeg], xo[kSeg];
#pragma unroll
    for (int k = 0; k < kSeg; ++k) {
      const int j = j0 + k;
      if (j < len) {
        const size_t r = base + (size_t)j * step;
        yy[k] = tmp[r];
        d[k] = U0[r];
        p[k] = U1[r] * d[k];
        q[k] = U2[r] * d[k];
        xo[k] = acc ? x[r] : 0.0;
      } else {
        yy[k] = d[k] = p[k] = q[k] = xo[k] = 0.0;
      }
    }
    Aff a = aff_id();
#pragma unroll
    for (int k = kSeg - 1; k >= 0; --k)
      if (j0 + k < len) a = aff_step(a, p[k], q[k], yy[k] * d[k]);
    for (int off = 1; off < kWave; off <<= 1) {
      const Aff o = aff_shfl<false>(a, off);
      if (lane + off < kWave) a = aff_compose(a, o);
    }
    const Aff after = aff_shfl<false>(a, 1);  // lanes k+1..63
    double x1 = s0, x2 = s1;
    if (lane < kWave - 1) {
      x1 = after.m00 * s0 + after.m01 * s1 + after.c0;
      x2 = after.m10 * s0 + after.m11 * s1 + after.c1;
    }
#pragma unroll
    for (int k = kSeg - 1; k >= 0; --k) {
      const int j = j0 + k;
      if (j < len) {
        const double xj = yy[k] * d[k] - p[k] * x1 - q[k] * x2;
        x[base + (size_t)j * step] = acc ? xo[k] + omega * xj : xj;
        x2 = x1;
        x1 = xj;
      }
    }
    s0 = __shfl(x1, 0);
    s1 = __shfl(x2, 0);
  }
}

// y = [D^-1 x_m ; 0] on momentum DOF rows, 0 elsewhere
__global__ __launch_bounds__(BS) void k_dscale(View V, int n, const double* dj, const double* x, double* y) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  const Geo& g = V.g;
  const int comp = r / (int)V.S, q = r % (int)V.S, i = q / g.P, j = q % g.P;
  const bool dof = comp == 0 ? u_dof(g, i, j) : (comp == 1 ? v_dof(g, i, j) : false);
  y[r] = dof ? dj[r] * x[r] : 0.0;
}

// z = [xs - xb ; y] on DOFs, f on the identity rows
__global__ __launch_bounds__(BS) void k_lsc_out(View V, int n, const double* f, const double* xs, const double* xb,
                                               const double* y, double* z) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  const Geo& g = V.g;
  const int comp = r / (int)V.S, q = r % (int)V.S, i = q / g.P, j = q % g.P;
  double out;
  if (comp == 0) out = u_dof(g, i, j) ? xs[r] - xb[r] : f[r];
  else if (comp == 1) out = v_dof(g, i, j) ? xs[r] - xb[r] : f[r];
  else out = p_dof(g, i, j) ? y[q] : f[r];
  z[r] = out;
}

// objective: mean of p(i, 1) over inlet rows whose cell has gamma >= 1 - tol
// (outlet mean is 0); g = dJ/dw; out = {J, rows}
__global__ void k_objective(View V, double tol, int n, double* gvec, double* out) {
  __shared__ double sh[BS / kWave + 1];
  const Geo& g = V.g;
  double cnt = 0.0, sum = 0.0;
  for (int i = 1 + threadIdx.x; i < g.M; i += BS)
    if (pure_fluid(V.gam[at(g, i - 1, 0)], tol)) {
      cnt += 1.0;
      sum += V.p[at(g, i, 1)];
    }
  cnt = block_sum<BS>(cnt, sh);
  sum = block_sum<BS>(sum, sh);
  if (threadIdx.x == 0) sh[0] = cnt, sh[1] = sum;
  __syncthreads();
  cnt = sh[0];
  sum = sh[1];
  for (int r = threadIdx.x; r < n; r += BS) gvec[r] = 0.0;
  __syncthreads();
  if (cnt > 0)
    for (int i = 1 + threadIdx.x; i < g.M; i += BS)
      if (pure_fluid(V.gam[at(g, i - 1, 0)], tol)) gvec[2 * V.S + at(g, i, 1)] = 1.0 / cnt;
  if (threadIdx.x == 0) {
    out[0] = cnt > 0 ? sum / cnt : 0.0;
    out[1] = cnt;
  }
}

// dJ/dgamma of each physical cell: - alpha'(gamma) * sum of lambda_r * 0.5 vol x_r over
// the momentum equations whose face average uses one of the padded cells
// that copy this physical cell
__global__ __launch_bounds__(BS) void k_grad(View V, const double* lam, double alphaMax, double q, double eps,
                                            double* grad) {
  const Geo& g = V.g;
  const int ny = g.M - 1, nx = g.N - 1;
  const int t = blockIdx.x * BS + threadIdx.x;
  if (t >= nx * ny) return;
  const int rr = t / nx, cc = t % nx;
  double acc = 0.0;
  for (int ca = 0; ca < 4; ++ca) {
    int a = rr, b = cc;
    if (ca & 1) { if (rr != ny - 1) continue; a = g.M - 1; }
    if (ca & 2) { if (cc != nx - 1) continue; b = g.N - 1; }
    // U equations (a+1, b+1) and (a+1, b); V equations (a+1, b+1) and (a, b+1)
    const int ui[2] = {a + 1, a + 1}, uj[2] = {b + 1, b};
    for (int k = 0; k < 2; ++k)
      if (u_dof(g, ui[k], uj[k])) acc += lam[at(g, ui[k], uj[k])] * V.u[at(g, ui[k], uj[k])];
    const int vi[2] = {a + 1, a}, vj[2] = {b + 1, b + 1};
    for (int k = 0; k < 2; ++k)
      if (v_dof(g, vi[k], vj[k])) acc += lam[V.S + at(g, vi[k], vj[k])] * V.v[at(g, vi[k], vj[k])];
  }
  const double gm = V.gam[at(g, rr, cc)];
  const double den = 1.0 + q * gm + eps;
  const double dalpha = -alphaMax * (1.0 + q + eps) / (den * den);
  grad[t] = -dalpha * 0.5 * g.vol * acc;
}

}  // namespace adj

// ===========================================================================
class FlowAdjoint {
 public:
  Ctx* ctx;
  Geo g;
  size_t S;
  int n = 0, WT = 0, nb = 0;
  DevBuf<int> col, colT, cnt, cur, flag, iout, bad;
  DevBuf<double> val, valT, res, dinv, gvec, lam, r, Vb, Zb, w, t1, t2, t3, t4, ta, lt, y, rp, xp, dEa, dNa, part,
      sums, grad, obj, bandx, bandy, dj;
  // preconditioner: CPS_ADJ_PRECOND_LSC_ADI (default), _LSC_LINE or _SIMPLE
  int kind = CPS_ADJ_PRECOND_LSC_ADI;
  int sweeps = 2;       // line sweeps per momentum-block solve after the first
  double omega = 0.8;   // damping of the sweeps
  // multigrid-PCG iterations per Schur-complement solve (rtol 1e-3): FGMRES
  // is flexible and, with the multigrid's outlet-scaled coarse levels, its
  // iteration count on 512 x 256 fields is the same from 2 to 30 (34 at 1),
  // so the inner solves stay short
  int schur_iters = 2;
  DevBuf<IterState> st;
  std::unique_ptr<PressureMG> mg;
  KernelTimer timer;
  // Krylov basis per cycle: the SIMPLE-preconditioned adjoint of SOU states
  // stagnates under short restarts (right-preconditioned GMRES(40) makes no
  // progress on 32 x 16 fields; 160-200 converges); the LSC forms converge
  // in 20-60 iterations on 512 x 256 fields.  The basis (restart + 1 vectors
  // of 3 fields) is allocated for the form in use.
  int restart() const { return kind == CPS_ADJ_PRECOND_SIMPLE ? 200 : 80; }
  void ensure_basis(int m) {
    if (Vb.n >= (size_t)(m + 1) * n) return;
    Vb.alloc((size_t)(m + 1) * n);
    Zb.alloc((size_t)m * n);
    part.alloc((size_t)(m + 1) * nb);
    sums.alloc(2 * (m + 1) + 1);
  }

  FlowAdjoint(Ctx* c, const Geo& geo, size_t fieldS) : ctx(c), g(geo), S(fieldS) {
    n = (int)(3 * S);
    hipStream_t s = ctx->stream;
    col.alloc((size_t)adj::K * n);
    val.alloc((size_t)adj::K * n);
    for (auto* b : {&cnt, &cur, &flag}) b->alloc(n);
    iout.alloc(2);
    for (auto* b : {&res, &dinv, &gvec, &lam, &r, &w, &t1, &t2, &t3, &t4, &ta, &lt}) b->alloc(n);
    bandx.alloc(5 * 2 * S);
    bandy.alloc(5 * 2 * S);
    dj.alloc(2 * S);
    bad.alloc(1);
    y.alloc(S);
    dEa.alloc(S);
    dNa.alloc(S);
    rp.alloc(S);
    xp.alloc(S);
    nb = std::min(ceil_div(n, adj::BS), 1024);
    grad.alloc((size_t)(g.M - 1) * (g.N - 1));
    obj.alloc(2);
    st.alloc(1);
    CPS_HIP(hipMemsetAsync(st.p, 0, sizeof(IterState), s));
    CPS_HIP(hipMemsetAsync(y.p, 0, y.bytes(), s));
    timer.stream = s;
    mg.reset(new PressureMG(ctx, g));
  }

  int blocks(int count) const { return ceil_div(count, adj::BS); }

  // builds J (at the current state) and J^T
  void build(const adj::View& V) {
    hipStream_t s = ctx->stream;
    adj::k_rows<<<blocks(n), adj::BS, 0, s>>>(V, n, col.p, val.p, res.p, flag.p, dinv.p);
    CPS_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), s));
    adj::k_count<<<blocks(n), adj::BS, 0, s>>>(n, col.p, cnt.p);
    CPS_HIP(hipMemsetAsync(iout.p, 0, 2 * sizeof(int), s));
    adj::k_imax<<<std::min(blocks(n), 1024), adj::BS, 0, s>>>(n, cnt.p, flag.p, iout.p);
    int h[2];
    CPS_HIP(hipMemcpyAsync(h, iout.p, sizeof(h), hipMemcpyDeviceToHost, s));
    CPS_HIP(hipStreamSynchronize(s));
    CPS_CHECK(h[1] == 0, CPS_ERR_STATE, "adjoint: an equation reads more DOFs than the dual slots hold");
    if (h[0] > WT) {
      WT = h[0];
      colT.alloc((size_t)WT * n);
      valT.alloc((size_t)WT * n);
    }
    CPS_HIP(hipMemsetAsync(cur.p, 0, cur.bytes(), s));
    adj::k_scatter<<<blocks(n), adj::BS, 0, s>>>(n, col.p, val.p, cur.p, WT, colT.p, valT.p);
    adj::k_sort<<<blocks(n), adj::BS, 0, s>>>(n, WT, cnt.p, colT.p, valT.p);
    if (kind != CPS_ADJ_PRECOND_SIMPLE) {
      const int b2 = blocks((int)(2 * S));
      CPS_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
      adj::k_band<false><<<b2, adj::BS, 0, s>>>(V, n, WT, cnt.p, colT.p, valT.p, bandx.p, dj.p);
      adj::k_line_factor<false><<<ceil_div(nlines(false), 64), 64, 0, s>>>(V, bandx.p, bad.p);
      if (kind == CPS_ADJ_PRECOND_LSC_ADI) {
        adj::k_band<true><<<b2, adj::BS, 0, s>>>(V, n, WT, cnt.p, colT.p, valT.p, bandy.p, nullptr);
        adj::k_line_factor<true><<<ceil_div(nlines(true), 64), 64, 0, s>>>(V, bandy.p, bad.p);
      }
      int hb = 0;
      CPS_HIP(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
      CPS_HIP(hipStreamSynchronize(s));
      CPS_CHECK(hb == 0, CPS_ERR_DIVERGED, "adjoint: zero pivot in a line factorisation of the momentum block");
    }
    adj::k_faces<<<blocks((int)S), adj::BS, 0, s>>>(V, kind == CPS_ADJ_PRECOND_SIMPLE ? dinv.p : dj.p, dEa.p,
                                                   dNa.p);
    CPS_HIP(hipGetLastError());
    schur_ready = false;  // new L: the next Schur solve rebuilds its hierarchy
  }

  void spmv(const double* x, double* out) {
    adj::k_spmv<<<blocks(n), adj::BS, 0, ctx->stream>>>(n, n, WT, cnt.p, colT.p, valT.p, x, out);
  }
  // the momentum rows (u and v blocks, [0, 2 S)) of J^T x only: what a line
  // pass reads (its x is zero on the pressure block, its lines lie in u, v)
  void spmv_mom(const double* x, double* out) {
    const int rows = (int)(2 * S);
    adj::k_spmv<<<blocks(rows), adj::BS, 0, ctx->stream>>>(n, rows, WT, cnt.p, colT.p, valT.p, x, out);
  }

  int nlines(bool y) const { return y ? (g.N - 2) + (g.N - 1) : (g.M - 1) + (g.M - 2); }

  // one line pass: x = Dl^-1 f (first) or x += omega Dl^-1 (f - A x)
  void line_pass(const adj::View& V, bool y, const double* f, double* x, bool first) {
    hipStream_t s = ctx->stream;
    const double* Ax = nullptr;
    if (!first) {
      spmv_mom(x, ta.p);
      Ax = ta.p;
    }
    const int lpb = 256 / kWave;  // lines per block (one wavefront each)
    if (y) adj::k_line_solve<true><<<ceil_div(nlines(true), lpb), 256, 0, s>>>(V, bandy.p, f, Ax, omega, x, lt.p);
    else adj::k_line_solve<false><<<ceil_div(nlines(false), lpb), 256, 0, s>>>(V, bandx.p, f, Ax, omega, x, lt.p);
  }

  // x = A~^-1 f_m by line Jacobi (x: momentum DOF rows, 0 elsewhere): an
  // x-line solve, then (ADI) a y-line pass, then `sweeps` rounds of x (and y)
  // passes
  void line_solve(const adj::View& V, const double* f, double* x) {
    const bool adi = kind == CPS_ADJ_PRECOND_LSC_ADI;
    adj::k_fill<<<blocks(n), adj::BS, 0, ctx->stream>>>(n, 0.0, x);
    line_pass(V, false, f, x, true);
    if (adi) line_pass(V, true, f, x, false);
    for (int k = 0; k < sweeps; ++k) {
      line_pass(V, false, f, x, false);
      if (adi) line_pass(V, true, f, x, false);
    }
  }

  // xp = L^-1 rp by multigrid-PCG; the hierarchy of L (fixed from build()
  // on) is set up by the first solve after each build only
  bool schur_ready = false;
  void schur_solve(const cps_simple_config& pcfg) {
    // FP32 multigrid inside one FP64 refinement sweep: the Schur solve is
    // a preconditioner of flexible GMRES, so its precision sets only the
    // iteration count (unchanged at 512 x 256), not the gradient's accuracy
    mg->solve(dEa.p, dNa.p, rp.p, xp.p, st.p, pcfg, false, &timer, schur_ready);
    schur_ready = true;
  }

  void precond(const adj::View& V, const double* f, double* z, const cps_simple_config& pcfg) {
    if (kind == CPS_ADJ_PRECOND_SIMPLE) precond_simple(V, f, z, pcfg);
    else precond_lsc(V, f, z, pcfg);
  }

  // z = M^-1 f, block factorisation with the x-line momentum solve and the
  // least-squares-commutator Schur approximation
  void precond_lsc(const adj::View& V, const double* f, double* z, const cps_simple_config& pcfg) {
    hipStream_t s = ctx->stream;
    const int sb = blocks((int)S);
    line_solve(V, f, t1.p);  // x* = A~^-1 f_m
    spmv(t1.p, t2.p);
    adj::k_pre2<<<sb, adj::BS, 0, s>>>(V, t2.p, f, rp.p, xp.p);  // C x* - f_p
    schur_solve(pcfg);  // L^-1
    adj::k_pre3<<<blocks(n), adj::BS, 0, s>>>(V, n, xp.p, t3.p);
    spmv(t3.p, t2.p);  // B y1
    adj::k_dscale<<<blocks(n), adj::BS, 0, s>>>(V, n, dj.p, t2.p, t3.p);
    spmv(t3.p, t2.p);  // A D^-1 B y1
    adj::k_dscale<<<blocks(n), adj::BS, 0, s>>>(V, n, dj.p, t2.p, t3.p);
    spmv(t3.p, t2.p);  // C D^-1 A D^-1 B y1
    adj::k_pre2<<<sb, adj::BS, 0, s>>>(V, t2.p, nullptr, rp.p, xp.p);
    schur_solve(pcfg);  // y = L^-1 (...)
    adj::k_pre3<<<blocks(n), adj::BS, 0, s>>>(V, n, xp.p, t3.p);
    spmv(t3.p, t2.p);  // B y
    line_solve(V, t2.p, t4.p);
    adj::k_lsc_out<<<blocks(n), adj::BS, 0, s>>>(V, n, f, t1.p, t4.p, xp.p, z);
    CPS_HIP(hipGetLastError());
  }

  // z = M^-1 f (SIMPLE block preconditioner)
  void precond_simple(const adj::View& V, const double* f, double* z, const cps_simple_config& pcfg) {
    hipStream_t s = ctx->stream;
    adj::k_pre1<<<blocks(n), adj::BS, 0, s>>>(V, n, dinv.p, f, t1.p);
    spmv(t1.p, t2.p);
    adj::k_pre2<<<blocks((int)S), adj::BS, 0, s>>>(V, t2.p, f, rp.p, xp.p);
    schur_solve(pcfg);
    adj::k_pre3<<<blocks(n), adj::BS, 0, s>>>(V, n, xp.p, t3.p);
    spmv(t3.p, t2.p);
    adj::k_pre4<<<blocks(n), adj::BS, 0, s>>>(V, n, dinv.p, f, t1.p, t2.p, xp.p, z);
    CPS_HIP(hipGetLastError());
  }

  // nv dot products V_j . x (j < nv), returned on the host
  // out[j] = V_j . x on the device, j < nv
  void mdot_dev(const double* Vbase, int nv, const double* x, double* out) {
    hipStream_t s = ctx->stream;
    adj::k_mdot<<<nb, adj::BS, 0, s>>>(n, Vbase, nv, x, part.p);
    adj::k_psum<<<nv, adj::BS, 0, s>>>(part.p, nb, out);
  }
  void mdot(const double* Vbase, int nv, const double* x, double* h) {
    hipStream_t s = ctx->stream;
    mdot_dev(Vbase, nv, x, sums.p);
    CPS_HIP(hipMemcpyAsync(h, sums.p, nv * sizeof(double), hipMemcpyDeviceToHost, s));
    CPS_HIP(hipStreamSynchronize(s));
  }
  double norm(const double* x) {
    double h;
    mdot(x, 1, x, &h);
    return std::sqrt(h);
  }
  void comb(double a0, double* wv, const double* Vbase, const std::vector<double>& c) {
    int done = 0;
    while (done < (int)c.size()) {  // chunks of kMaxBasis coefficients
      const int m = std::min((int)c.size() - done, adj::kMaxBasis);
      adj::Coefs cc;
      for (int k = 0; k < m; ++k) cc.c[k] = c[done + k];
      adj::k_mcomb<<<nb, adj::BS, 0, ctx->stream>>>(n, done == 0 ? a0 : 1.0, wv, Vbase + (size_t)done * n, m, cc);
      done += m;
    }
  }

  // FGMRES(restart()) on J^T lam = g, from lam = 0; returns iterations
  int fgmres(const adj::View& V, double rtol, int max_iter, const cps_simple_config& pcfg, double* relres) {
    hipStream_t s = ctx->stream;
    const int m = restart();
    ensure_basis(m);
    CPS_HIP(hipMemsetAsync(lam.p, 0, lam.bytes(), s));
    const double gnorm = norm(gvec.p);
    if (gnorm == 0.0) {
      *relres = 0.0;
      return 0;
    }
    const double tol = rtol * gnorm;
    CPS_HIP(hipMemcpyAsync(r.p, gvec.p, r.bytes(), hipMemcpyDeviceToDevice, s));
    double beta = gnorm;
    int it = 0;
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), e(m + 1), h(m + 1), hh(2 * (m + 1) + 1);
    while (it < max_iter && beta > tol) {
      adj::k_scale<<<nb, adj::BS, 0, s>>>(n, r.p, 1.0 / beta, Vb.p);
      std::fill(e.begin(), e.end(), 0.0);
      e[0] = beta;
      int k = 0;
      for (; k < m && it < max_iter; ++k, ++it) {
        double* vk = Vb.p + (size_t)k * n;
        double* zk = Zb.p + (size_t)k * n;
        precond(V, vk, zk, pcfg);
        spmv(zk, w.p);
        // classical Gram-Schmidt, twice, and the next basis vector, with the
        // coefficients kept on the device: one read-back per step
        double* hd = sums.p;               // h: first pass
        double* h2d = sums.p + (m + 1);    // h2: second pass
        double* nd = sums.p + 2 * (m + 1); // ||w||^2
        mdot_dev(Vb.p, k + 1, w.p, hd);
        adj::k_mcomb_dev<<<nb, adj::BS, 0, s>>>(n, w.p, Vb.p, k + 1, hd);
        mdot_dev(Vb.p, k + 1, w.p, h2d);
        adj::k_mcomb_dev<<<nb, adj::BS, 0, s>>>(n, w.p, Vb.p, k + 1, h2d);
        mdot_dev(w.p, 1, w.p, nd);
        adj::k_scale_rn<<<nb, adj::BS, 0, s>>>(n, w.p, nd, Vb.p + (size_t)(k + 1) * n);
        CPS_HIP(hipMemcpyAsync(hh.data(), sums.p, (2 * (m + 1) + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
        CPS_HIP(hipStreamSynchronize(s));
        for (int q = 0; q <= k; ++q) h[q] = hh[q] + hh[m + 1 + q];
        const double hn = std::sqrt(hh[2 * (m + 1)]);
        for (int q = 0; q <= k; ++q) H[(size_t)q * m + k] = h[q];
        H[(size_t)(k + 1) * m + k] = hn;
        // Givens rotations
        for (int q = 0; q < k; ++q) {
          const double a = H[(size_t)q * m + k], b = H[(size_t)(q + 1) * m + k];
          H[(size_t)q * m + k] = cs[q] * a + sn[q] * b;
          H[(size_t)(q + 1) * m + k] = -sn[q] * a + cs[q] * b;
        }
        const double a = H[(size_t)k * m + k], b = H[(size_t)(k + 1) * m + k];
        const double rr = std::hypot(a, b);
        cs[k] = rr > 0 ? a / rr : 1.0;
        sn[k] = rr > 0 ? b / rr : 0.0;
        H[(size_t)k * m + k] = rr;
        H[(size_t)(k + 1) * m + k] = 0.0;
        e[k + 1] = -sn[k] * e[k];
        e[k] = cs[k] * e[k];
        if (std::fabs(e[k + 1]) <= tol || hn == 0.0) {
          ++k;
          ++it;
          break;
        }
      }
      // back substitution, lam += Z y
      std::vector<double> yv(k, 0.0);
      for (int q = k - 1; q >= 0; --q) {
        double a = e[q];
        for (int p2 = q + 1; p2 < k; ++p2) a -= H[(size_t)q * m + p2] * yv[p2];
        yv[q] = H[(size_t)q * m + q] != 0.0 ? a / H[(size_t)q * m + q] : 0.0;
      }
      comb(1.0, lam.p, Zb.p, yv);
      // true residual
      spmv(lam.p, r.p);
      adj::k_rsub<<<nb, adj::BS, 0, s>>>(n, gvec.p, r.p);
      beta = norm(r.p);
      if (!std::isfinite(beta)) throw Error{CPS_ERR_DIVERGED, "adjoint solve produced a non-finite residual"};
    }
    *relres = beta / gnorm;
    return it;
  }

  void run(const SimpleView& sv, double rtol, int max_iter, double* dp, double* hgrad, cps_solve_stats* stats) {
    hipStream_t s = ctx->stream;
    adj::View V{sv.g, sv.S, sv.u, sv.v, sv.p, sv.gam, sv.alp, sv.inlet, sv.dE, sv.dN};
    build(V);
    adj::k_objective<<<1, adj::BS, 0, s>>>(V, sv.cfg.dp_fluid_tol, n, gvec.p, obj.p);
    cps_simple_config pcfg = sv.cfg;
    pcfg.p_tol = 0.0;
    pcfg.p_rtol = 1e-3;
    pcfg.max_pressure_iter = schur_iters;
    pcfg.max_refine = 1;
    double relres = 0.0;
    const int its = fgmres(V, rtol, max_iter, pcfg, &relres);
    const double alphaMax = sv.cfg.eta / std::max(sv.cfg.brinkman_kmin, (double)1e-20f);
    const int ncell = (g.M - 1) * (g.N - 1);
    adj::k_grad<<<blocks(ncell), adj::BS, 0, s>>>(V, lam.p, alphaMax, sv.cfg.brinkman_q, (double)1e-10f, grad.p);
    CPS_HIP(hipGetLastError());
    double ho[2];
    CPS_HIP(hipMemcpyAsync(ho, obj.p, sizeof(ho), hipMemcpyDeviceToHost, s));
    if (hgrad) CPS_HIP(hipMemcpyAsync(hgrad, grad.p, grad.bytes(), hipMemcpyDeviceToHost, s));
    CPS_HIP(hipStreamSynchronize(s));
    if (dp) *dp = ho[0];
    if (stats) {
      stats->iterations = its;
      stats->refinements = 0;
      stats->rel_residual = relres;
      stats->converged = relres <= rtol ? 1 : 0;
    }
  }

  // d J / d gamma of a caller's objective J(u, v, p) of the converged state:
  // dJ/dw (host, public shapes, null = 0) is folded through the boundary
  // rules the equations read the ghost values by (readU / readV / readP:
  // prescribed and zero values drop out, copies add to the DOF they copy),
  // then J^T lambda = dJ/dw and dJ/dgamma = -lambda^T dR/dgamma as for dP.
  void run_vjp(const SimpleView& sv, const double* hu, const double* hv, const double* hp, double rtol,
               int max_iter, double* hgrad, cps_solve_stats* stats) {
    hipStream_t s = ctx->stream;
    adj::View V{sv.g, sv.S, sv.u, sv.v, sv.p, sv.gam, sv.alp, sv.inlet, sv.dE, sv.dN};
    build(V);
    const int M = g.M, N = g.N;
    const size_t P = (size_t)g.P;
    std::vector<double> hg((size_t)n, 0.0);
    if (hu)
      for (int i = 0; i <= M; ++i)
        for (int j = 0; j < N; ++j) {
          const double x = hu[(size_t)i * N + j];
          if (x == 0.0 || j == 0 || i == 0 || i == M) continue;  // inlet, walls
          hg[(size_t)i * P + (j == N - 1 ? N - 2 : j)] += x;       // outlet copy
        }
    if (hv)
      for (int i = 0; i < M; ++i)
        for (int j = 0; j <= N; ++j) {
          const double x = hv[(size_t)i * (N + 1) + j];
          if (x == 0.0 || i == 0 || i == M - 1 || j == 0) continue;
          hg[S + (size_t)i * P + (j == N ? N - 1 : j)] += x;
        }
    if (hp)
      for (int i = 0; i <= M; ++i)
        for (int j = 0; j <= N; ++j) {
          const double x = hp[(size_t)i * (N + 1) + j];
          if (x == 0.0 || j >= N - 1) continue;  // columns N-1, N are 0
          const int ii = std::min(std::max(i, 1), M - 1), jj = j == 0 ? 1 : j;
          hg[2 * S + (size_t)ii * P + jj] += x;
        }
    CPS_HIP(hipMemcpyAsync(gvec.p, hg.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
    cps_simple_config pcfg = sv.cfg;
    pcfg.p_tol = 0.0;
    pcfg.p_rtol = 1e-3;
    pcfg.max_pressure_iter = schur_iters;
    pcfg.max_refine = 1;
    double relres = 0.0;
    const int its = fgmres(V, rtol, max_iter, pcfg, &relres);
    const double alphaMax = sv.cfg.eta / std::max(sv.cfg.brinkman_kmin, (double)1e-20f);
    const int ncell = (g.M - 1) * (g.N - 1);
    adj::k_grad<<<blocks(ncell), adj::BS, 0, s>>>(V, lam.p, alphaMax, sv.cfg.brinkman_q, (double)1e-10f, grad.p);
    CPS_HIP(hipGetLastError());
    if (hgrad) CPS_HIP(hipMemcpyAsync(hgrad, grad.p, grad.bytes(), hipMemcpyDeviceToHost, s));
    CPS_HIP(hipStreamSynchronize(s));
    if (stats) {
      stats->iterations = its;
      stats->refinements = 0;
      stats->rel_residual = relres;
      stats->converged = relres <= rtol ? 1 : 0;
    }
  }

  // y = J x or J^T x on host field arrays [u | v | p] in their public shapes
  void apply(const adj::View& V, bool transpose, const double* hx, double* hy) {
    hipStream_t s = ctx->stream;
    build(V);
    const int M = g.M, N = g.N;
    const int rows[3] = {M + 1, M, M + 1}, cols[3] = {N, N + 1, N + 1};
    CPS_HIP(hipMemsetAsync(t1.p, 0, t1.bytes(), s));
    size_t off = 0;
    for (int c = 0; c < 3; ++c) {
      CPS_HIP(hipMemcpy2DAsync(t1.p + c * S, g.P * sizeof(double), hx + off, cols[c] * sizeof(double),
                               cols[c] * sizeof(double), rows[c], hipMemcpyHostToDevice, s));
      off += (size_t)rows[c] * cols[c];
    }
    if (transpose) spmv(t1.p, t2.p);
    else adj::k_spmv_rows<<<blocks(n), adj::BS, 0, s>>>(n, col.p, val.p, t1.p, t2.p);
    CPS_HIP(hipGetLastError());
    off = 0;
    for (int c = 0; c < 3; ++c) {
      CPS_HIP(hipMemcpy2DAsync(hy + off, cols[c] * sizeof(double), t2.p + c * S, g.P * sizeof(double),
                               cols[c] * sizeof(double), rows[c], hipMemcpyDeviceToHost, s));
      off += (size_t)rows[c] * cols[c];
    }
    CPS_HIP(hipStreamSynchronize(s));
  }

  // max |R| of the momentum and continuity equations at the current state
  // (how converged the state the adjoint linearises about is)
  void state_residual(double* out3) {
    std::vector<double> h(n);
    CPS_HIP(hipMemcpyAsync(h.data(), res.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CPS_HIP(hipStreamSynchronize(ctx->stream));
    out3[0] = out3[1] = out3[2] = 0.0;
    for (int r2 = 0; r2 < n; ++r2) {
      const int comp = r2 / (int)S;
      out3[comp] = std::max(out3[comp], std::fabs(h[r2]));
    }
  }
};

FlowAdjoint* flow_adjoint_new(Ctx* ctx, const Geo& g, size_t S) { return new FlowAdjoint(ctx, g, S); }
void flow_adjoint_free(FlowAdjoint* a) { delete a; }
void flow_adjoint_run(FlowAdjoint* a, const SimpleView& sv, double rtol, int max_iter, double* dp, double* grad,
                      cps_solve_stats* stats) {
  a->run(sv, rtol, max_iter, dp, grad, stats);
}
void flow_adjoint_vjp(FlowAdjoint* a, const SimpleView& sv, const double* du, const double* dv, const double* dp,
                      double rtol, int max_iter, double* grad, cps_solve_stats* stats) {
  a->run_vjp(sv, du, dv, dp, rtol, max_iter, grad, stats);
}
void flow_adjoint_apply(FlowAdjoint* a, const SimpleView& sv, int transpose, const double* x, double* y) {
  adj::View V{sv.g, sv.S, sv.u, sv.v, sv.p, sv.gam, sv.alp, sv.inlet, sv.dE, sv.dN};
  a->apply(V, transpose != 0, x, y);
}
void flow_adjoint_options(FlowAdjoint* a, int kind, int sweeps, double omega, int schur_iters) {
  a->kind = kind;
  a->sweeps = sweeps;
  a->omega = omega;
  a->schur_iters = schur_iters;
}
void flow_adjoint_residual(FlowAdjoint* a, const SimpleView& sv, double* out3) {
  adj::View V{sv.g, sv.S, sv.u, sv.v, sv.p, sv.gam, sv.alp, sv.inlet, sv.dE, sv.dN};
  a->build(V);
  a->state_residual(out3);
}

}  // namespace cps
