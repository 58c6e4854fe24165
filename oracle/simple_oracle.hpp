// ============================================================================
// oracle/simple_oracle.hpp -- CPU restatement of the reference SIMPLE path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle (and the
// timed CPU baseline) for the HIP SIMPLE path.  Only tests/, the smoke check
// and bench.py's cpu_baseline leg may load it; the product library never does.
//
// It restates, templated on the real type R (float = the reference's own
// precision, double = the FP64 oracle), the algorithm of:
//   SIMPLE.cpp:53-54,78-79        grid padding M=Ny+1, N=Nx+1, alpha_max
//   SIMPLE.cpp:667-676,707-729    cellType/gamma padding of the last row/col
//   SIMPLE.cpp:766-771            Brinkman RAMP alpha(gamma)
//   SIMPLE.h:642-686,740-778      alphaAtU/V, gammaAtU/V face averages
//   boundaries.cpp:50-146         velocity / pressure boundary conditions
//   boundaries.cpp:148-206        half-cell external no-slip flags
//   convection.cpp:53-186         SOU deferred correction
//   momentum_solver.cpp:52-159    DOF mapping + sorted-column CSR structure
//   momentum_solver.cpp:164-355   momentum assembly, solve, scatter+clamp
//   jacobi_solver.cpp:20-75       point Jacobi, sum r^2 < tol^2 n each sweep
//   pressure_solver.cpp:425-458   pressure DOF map (j = 1..N-2)
//   pressure_solver.cpp:462-496   mass residual over j = 1..N-1
//   pressure_solver.cpp:649-742   pressure CSR values + RHS
//   pressure_solver.cpp:121-184   Jacobi-PCG with absolute 2-norm tolerance
//   pressure_solver.cpp:819-919   red-black SOR fallback
//   iterations.cpp:82-659         one SIMPLE step (implicit momentum branch)
//   postprocessing.cpp:62-243     samplePlaneAtX (core-plane dP)
//   postprocessing.cpp:245-327    sampleBoundaryPatch (full-system dP)
//   SIMPLE.cpp:310-491            outer loop: inlet ramp, stop on error <=
//                                 epsilon or on the core-dP window (+ slope)
//
// Extensions the reference does not have (documented conventions):
//   * an optional per-u-row inlet profile (parabolic inlet); the reference's
//     gamma > 0.5 inlet mask still applies on top of it;
//   * a configurable pure-fluid tolerance for the boundary-patch dP.
//
// Mask decisions (half-cell wall flags, the inlet mask, the pure-fluid rows
// of the dP samples) are taken in float arithmetic in both instantiations,
// as the reference writes them on its float gamma (SIMPLE.h:479-520,
// 740-778; boundaries.cpp:82-100,148-206; postprocessing.cpp:155-160,268),
// so the FP64 oracle makes the reference's discrete choices bit for bit.
//
// ORACLE_MUTANT (default 0) plants one known mistake in the momentum
// assembly, the Brinkman field or the inlet mask; the CPU tests build each
// mutant and check that the closed-form pins of those terms reject it.
//
// Layout: row-major, element (i,j) of an R x C array at [i*C + j]; the
// reference's Eigen arrays are column-major but index semantics are the same.
// ============================================================================
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/coldplate_hip.h"

#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

namespace oracle {

// the planted mistake of a mutant build (0 = the faithful oracle)
constexpr int kMutant = ORACLE_MUTANT;

#ifndef ORACLE_OMP_MIN_LOOP_DEFINED
#define ORACLE_OMP_MIN_LOOP_DEFINED
// loops shorter than this run on one thread: on small grids the fork/join
// and barrier cost of a parallel region exceeds the loop's work, badly so
// when other processes hold the cores
constexpr int kOmpMinLoop = 16384;
#endif

template <class R>
struct Csr {
  int n = 0;
  std::vector<int> row_ptr, col_idx;
  std::vector<R> values, rhs, sol;
  std::vector<int> diag, east, west, north, south;  // value offsets or -1
};

struct StepStats {
  double error = 0, resid_u = 0, resid_v = 0, resid_mass = 0;
  double resid_u_max = 0, resid_v_max = 0, resid_mass_max = 0;
  int u_iters = 0, v_iters = 0, p_iters = 0;
};

struct PatchMetrics {
  double area = 0, mass_flux = 0, avg_static = 0, avg_dynamic = 0, avg_total = 0;
  int valid = 0;
};

template <class R>
class Simple {
 public:
  // ---- parameters (reference names) ----
  int M = 0, N = 0;
  R hx = 0, hy = 0, rho = R(997.0f), eta = R(0.00089f);
  R targetVel = 0, inletVelocity = 0;
  int convectionScheme = 0;
  bool enable25D = false;
  R Ht = 0, convFactor25D = R(1.2f), sinkMult25D = R(1.0f);
  R brinkmanKMin = R(1e-14f), brinkmanQ = R(1.0f), brinkmanAlphaMax = 0;
  R uvAlpha = R(0.7f), pAlpha = R(0.3f);
  R momentumTol = R(1e-6f);
  int maxMomentumIter = 400;
  R pTol = R(1e-6f);
  int maxPressureIter = 30;
  int pressureSolverType = 1;  // 0 = SOR, 1 = Jacobi-PCG
  R epsilon = R(1e-7f);
  int maxIterations = 1000;
  R dpFluidTol = R(1e-6f);
  bool hasInletProfile = false;
  // outer loop (SIMPLE.h:88-103, 385-386)
  R xCoreIn = R(0.01f), xCoreOut = R(0.05f);
  bool usePressureDropConvergence = true;
  int dpConvergenceWindow = 500;
  R dpConvergencePercent = R(1.0f);
  bool usePressureDropSlopeGate = false;
  int dpSlopeWindowIters = 1000;
  R dpSlopeMaxDelta = R(10.0f);
  bool enableInletRamp = false;
  int rampSteps = 1000;
  R inletScale = R(1);  // ramp factor of the inlet profile

  // ---- fields ----
  std::vector<R> u, v, p, uStar, vStar, pStar, uOld, vOld, dE, dN;
  std::vector<R> gamma, alpha, inletProfile;

  // ---- CSR systems (persistent, as in the reference) ----
  Csr<R> csrU, csrV, csrP;
  std::vector<std::pair<int, int>> mapU, mapV, mapP;
  std::vector<int> linU, linV, linP;

  // ---- reported residuals ----
  R residU = 1, residV = 1, residMass = 1;
  R residU_RMS = 0, residV_RMS = 0, residMass_RMS = 0;
  R residU_max = 0, residV_max = 0, residMass_max = 0;
  int lastUIters = 0, lastVIters = 0;

  // index helpers
  R& U(int i, int j) { return u[(size_t)i * N + j]; }
  R& V(int i, int j) { return v[(size_t)i * (N + 1) + j]; }
  R& P(int i, int j) { return p[(size_t)i * (N + 1) + j]; }
  R G(int i, int j) const { return gamma[(size_t)i * N + j]; }
  R A(int i, int j) const { return alpha[(size_t)i * N + j]; }

  // cfg: plain numbers; gamma_phys: Ny x Nx (1 = fluid), row-major, values
  // are clamped like loadDensityField (SIMPLE.cpp:707-717).  inlet: M+1
  // per-u-row inlet speeds or nullptr for the uniform reference inlet.
  void setup(const cps_simple_config& c, const double* gamma_phys, const double* inlet) {
    M = c.ny + 1;
    N = c.nx + 1;
    hx = R(c.hx);
    hy = R(c.hy);
    rho = R(c.rho);
    eta = R(c.eta);
    targetVel = R(c.u_in);
    inletVelocity = targetVel;
    convectionScheme = c.convection_scheme;
    Ht = R(c.ht);
    enable25D = c.enable_25d != 0 && Ht > R(0);  // SIMPLE.cpp:67-73
    convFactor25D = R(c.conv_factor_25d);
    sinkMult25D = R(c.sink_mult_25d);
    brinkmanKMin = R(c.brinkman_kmin);
    brinkmanQ = R(c.brinkman_q);
    uvAlpha = R(c.uv_alpha);
    pAlpha = R(c.p_alpha);
    momentumTol = R(c.momentum_tol);
    maxMomentumIter = c.max_momentum_iter;
    pTol = R(c.p_tol);
    maxPressureIter = c.max_pressure_iter;
    pressureSolverType = c.pressure_solver == 0 ? 0 : 1;
    epsilon = R(c.epsilon);
    maxIterations = c.max_iterations;
    dpFluidTol = R(c.dp_fluid_tol);
    xCoreIn = R(c.x_core_in);
    xCoreOut = R(c.x_core_out);
    usePressureDropConvergence = c.use_dp_convergence != 0;
    dpConvergenceWindow = c.dp_window;
    dpConvergencePercent = R(c.dp_percent);
    usePressureDropSlopeGate = c.use_dp_slope_gate != 0;
    dpSlopeWindowIters = c.dp_slope_window;
    dpSlopeMaxDelta = R(c.dp_slope_max_delta);
    enableInletRamp = c.inlet_ramp != 0;
    rampSteps = c.ramp_steps;

    // SIMPLE.cpp:78-79
    R K_min = std::max(brinkmanKMin, R(1e-20f));
    brinkmanAlphaMax = eta / K_min;

    const size_t nu = (size_t)(M + 1) * N, nv = (size_t)M * (N + 1), np = (size_t)(M + 1) * (N + 1);
    u.assign(nu, R(0)); uStar.assign(nu, R(0)); uOld.assign(nu, R(0)); dE.assign(nu, R(0));
    v.assign(nv, R(0)); vStar.assign(nv, R(0)); vOld.assign(nv, R(0)); dN.assign(nv, R(0));
    p.assign(np, R(0)); pStar.assign(np, R(0));
    gamma.assign((size_t)M * N, R(1));
    alpha.assign((size_t)M * N, R(0));

    // loadDensityField: gamma = clamp, then pad last row/col (SIMPLE.cpp:707-729)
    const int physRows = M - 1, physCols = N - 1;
    if (gamma_phys) {
      for (int i = 0; i < physRows; ++i)
        for (int j = 0; j < physCols; ++j) {
          R g = R(gamma_phys[(size_t)i * physCols + j]);
          gamma[(size_t)i * N + j] = std::max(R(0), std::min(R(1), g));
        }
    }
    for (int i = 0; i < physRows; ++i) gamma[(size_t)i * N + physCols] = gamma[(size_t)i * N + physCols - 1];
    for (int j = 0; j < N; ++j) {
      int src = std::min(j, physCols - 1);
      gamma[(size_t)physRows * N + j] = gamma[(size_t)(physRows - 1) * N + src];
    }
    // buildAlphaFromDensity (SIMPLE.cpp:766-771)
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < N; ++j) {
        R g = gamma[(size_t)i * N + j];
        const R q = kMutant == 10 ? R(1.0f) : brinkmanQ;  // mutant 10: q ignored
        alpha[(size_t)i * N + j] = brinkmanAlphaMax * (R(1.0f) - g) / (R(1.0f) + q * g + R(1e-10f));
      }

    hasInletProfile = inlet != nullptr;
    inletProfile.assign(M + 1, targetVel);
    if (inlet)
      for (int i = 0; i <= M; ++i) inletProfile[i] = R(inlet[i]);

    buildMaps();
  }

  // ------------------------------------------------------------------------
  // face interpolation helpers (SIMPLE.h:642-686, 740-778)
  R alphaAtU(int i, int j) const {
    if (i < 1 || i >= M || j < 1 || j >= N - 1) return R(0);
    return R(0.5f) * (A(i - 1, j - 1) + A(i - 1, j));
  }
  R alphaAtV(int i, int j) const {
    if (i < 1 || i >= M - 1 || j < 1 || j >= N) return R(0);
    return R(0.5f) * (A(i - 1, j - 1) + A(i, j - 1));
  }
  // float, as the reference evaluates it: these feed mask decisions only
  float gammaAtU(int i, int j) const {
    if (i < 1 || i >= M || j < 1 || j >= N - 1) return 1.0f;
    return 0.5f * ((float)G(i - 1, j - 1) + (float)G(i - 1, j));
  }
  float gammaAtV(int i, int j) const {
    if (i < 1 || i >= M - 1 || j < 1 || j >= N) return 1.0f;
    return 0.5f * ((float)G(i - 1, j - 1) + (float)G(i, j - 1));
  }
  // pure-fluid test of the dP samples (postprocessing.cpp:155-160, 268)
  bool pureFluid(int i, int j) const { return !((float)G(i, j) < 1.0f - (float)dpFluidTol); }

  // ------------------------------------------------------------------------
  // boundaries.cpp:50-118
  void setVelocityBC(std::vector<R>& uu, std::vector<R>& vv) {
    auto Ui = [&](int i, int j) -> R& { return uu[(size_t)i * N + j]; };
    auto Vi = [&](int i, int j) -> R& { return vv[(size_t)i * (N + 1) + j]; };
    for (int j = 0; j < N; ++j) Ui(0, j) = R(0);
    for (int j = 0; j < N + 1; ++j) Vi(0, j) = R(0);
    for (int j = 0; j < N; ++j) Ui(M, j) = R(0);
    for (int j = 0; j < N + 1; ++j) Vi(M - 1, j) = R(0);
    for (int i = 0; i <= M; ++i) Ui(i, 0) = inletFluidFromURow(i) ? inletSpeed(i) : R(0);
    for (int i = 0; i < M; ++i) Vi(i, 0) = R(0);
    for (int i = 0; i < M + 1; ++i) Ui(i, N - 1) = Ui(i, N - 2);
    for (int i = 0; i < M; ++i) Vi(i, N) = Vi(i, N - 1);
  }
  bool inletFluidFromURow(int uRow) const {  // boundaries.cpp:82-96
    int cellRow;
    if (uRow <= 0) cellRow = kMutant == 9 ? 1 : 0;  // mutant 9: wrong corner row
    else if (uRow >= M) cellRow = std::max(0, M - 2);
    else cellRow = uRow - 1;
    cellRow = std::max(0, std::min(cellRow, M - 1));
    if (kMutant == 8) return (float)G(cellRow, 0) >= 0.5f;  // mutant 8: >= for >
    return (float)G(cellRow, 0) > 0.5f;
  }
  R inletSpeed(int uRow) const {
    // uniform reference inlet uses the (possibly ramped) inletVelocity
    return hasInletProfile ? inletScale * inletProfile[uRow] : inletVelocity;
  }
  // boundaries.cpp:123-146
  void setPressureBC(std::vector<R>& pp) {
    auto Pi = [&](int i, int j) -> R& { return pp[(size_t)i * (N + 1) + j]; };
    for (int j = 1; j < N; ++j) {
      Pi(0, j) = Pi(1, j);
      Pi(M, j) = Pi(M - 1, j);
    }
    for (int i = 1; i < M; ++i) Pi(i, 0) = Pi(i, 1);
    for (int i = 1; i < M; ++i) Pi(i, N) = R(0);
  }

  // boundaries.cpp:148-206
  bool southU(int i, int j) {
    if (i - 1 != 0) return false;
    if (j < 1 || j > N - 2) return false;
    if (gammaAtU(i, j) < 0.999f) return false;
    const R t = R(1e-8f);
    return std::abs(U(0, j)) <= t && std::abs(V(0, j)) <= t && std::abs(V(0, j + 1)) <= t;
  }
  bool northU(int i, int j) {
    if (i + 1 != M) return false;
    if (j < 1 || j > N - 2) return false;
    if (gammaAtU(i, j) < 0.999f) return false;
    const R t = R(1e-8f);
    return std::abs(U(M, j)) <= t && std::abs(V(M - 1, j)) <= t && std::abs(V(M - 1, j + 1)) <= t;
  }
  bool westV(int i, int j) {
    if (j - 1 != 0) return false;
    if (i < 1 || i > M - 2) return false;
    if (gammaAtV(i, j) < 0.999f) return false;
    const R t = R(1e-8f);
    if (std::abs(targetVel) > t) return false;
    return std::abs(V(i, 0)) <= t && std::abs(U(i, 0)) <= t && std::abs(U(i + 1, 0)) <= t;
  }
  bool eastV(int i, int j) {
    if (j + 1 != N) return false;
    if (i < 1 || i > M - 2) return false;
    if (gammaAtV(i, j) < 0.999f) return false;
    const R t = R(1e-8f);
    if (std::abs(P(i, N)) <= t) return false;
    return std::abs(V(i, N)) <= t && std::abs(U(i, N - 1)) <= t && std::abs(U(i + 1, N - 1)) <= t;
  }

  // convection.cpp:53-118
  R souU(int i, int j, R Fe, R Fw, R Fn, R Fs) {
    R Sdc = 0;
    const R a = R(1.5f), b = R(0.5f);
    if (Fe >= R(0)) {
      if (j >= 2) Sdc -= Fe * ((a * U(i, j) - b * U(i, j - 1)) - U(i, j));
    } else {
      if (j + 2 < N) Sdc -= Fe * ((a * U(i, j + 1) - b * U(i, j + 2)) - U(i, j + 1));
    }
    if (Fw >= R(0)) {
      if (j >= 2) Sdc += Fw * ((a * U(i, j - 1) - b * U(i, j - 2)) - U(i, j - 1));
    } else {
      if (j + 1 < N) Sdc += Fw * ((a * U(i, j) - b * U(i, j + 1)) - U(i, j));
    }
    if (Fn >= R(0)) {
      if (i >= 1) Sdc -= Fn * ((a * U(i, j) - b * U(i - 1, j)) - U(i, j));
    } else {
      if (i + 2 <= M) Sdc -= Fn * ((a * U(i + 1, j) - b * U(i + 2, j)) - U(i + 1, j));
    }
    if (Fs >= R(0)) {
      if (i >= 2) Sdc += Fs * ((a * U(i - 1, j) - b * U(i - 2, j)) - U(i - 1, j));
    } else {
      if (i + 1 <= M) Sdc += Fs * ((a * U(i, j) - b * U(i + 1, j)) - U(i, j));
    }
    return Sdc;
  }
  // convection.cpp:121-186
  R souV(int i, int j, R Fe, R Fw, R Fn, R Fs) {
    R Sdc = 0;
    const R a = R(1.5f), b = R(0.5f);
    if (Fe >= R(0)) {
      if (j >= 2) Sdc -= Fe * ((a * V(i, j) - b * V(i, j - 1)) - V(i, j));
    } else {
      if (j + 2 <= N) Sdc -= Fe * ((a * V(i, j + 1) - b * V(i, j + 2)) - V(i, j + 1));
    }
    if (Fw >= R(0)) {
      if (j >= 2) Sdc += Fw * ((a * V(i, j - 1) - b * V(i, j - 2)) - V(i, j - 1));
    } else {
      if (j + 1 <= N) Sdc += Fw * ((a * V(i, j) - b * V(i, j + 1)) - V(i, j));
    }
    if (Fn >= R(0)) {
      if (i >= 1) Sdc -= Fn * ((a * V(i, j) - b * V(i - 1, j)) - V(i, j));
    } else {
      if (i + 2 < M) Sdc -= Fn * ((a * V(i + 1, j) - b * V(i + 2, j)) - V(i + 1, j));
    }
    if (Fs >= R(0)) {
      if (i >= 2) Sdc += Fs * ((a * V(i - 1, j) - b * V(i - 2, j)) - V(i - 1, j));
    } else {
      if (i + 1 < M) Sdc += Fs * ((a * V(i, j) - b * V(i + 1, j)) - V(i, j));
    }
    return Sdc;
  }

  // ------------------------------------------------------------------------
  // DOF maps + CSR structure (momentum_solver.cpp:52-159, pressure_solver.cpp:438-458)
  void buildMaps() {
    mapU.clear(); mapV.clear(); mapP.clear();
    linU.assign((size_t)(M + 1) * N, -1);
    linV.assign((size_t)M * (N + 1), -1);
    linP.assign((size_t)(M + 1) * (N + 1), -1);
    for (int i = 1; i < M; ++i)
      for (int j = 1; j < N - 1; ++j) { linU[(size_t)i * N + j] = (int)mapU.size(); mapU.push_back({i, j}); }
    for (int i = 1; i < M - 1; ++i)
      for (int j = 1; j < N; ++j) { linV[(size_t)i * (N + 1) + j] = (int)mapV.size(); mapV.push_back({i, j}); }
    for (int i = 1; i < M; ++i)
      for (int j = 1; j < N - 1; ++j) { linP[(size_t)i * (N + 1) + j] = (int)mapP.size(); mapP.push_back({i, j}); }
    buildCsr(csrU, mapU, linU, M + 1, N);
    buildCsr(csrV, mapV, linV, M, N + 1);
    buildCsr(csrP, mapP, linP, M + 1, N + 1);
  }
  // 5-point pattern with sorted columns; neighbor order N,S,W,E,diag as in
  // momentum_solver.cpp:98-101, then sorted by column.
  void buildCsr(Csr<R>& c, const std::vector<std::pair<int, int>>& map, const std::vector<int>& lin,
                int rows, int cols) {
    const int n = (int)map.size();
    c.n = n;
    c.row_ptr.assign(n + 1, 0);
    c.col_idx.clear();
    c.diag.assign(n, -1); c.east.assign(n, -1); c.west.assign(n, -1);
    c.north.assign(n, -1); c.south.assign(n, -1);
    const int di[5] = {1, -1, 0, 0, 0}, dj[5] = {0, 0, -1, 1, 0}, ty[5] = {3, 4, 2, 1, 0};
    int nnz = 0;
    std::vector<std::pair<int, int>> e;
    for (int row = 0; row < n; ++row) {
      int i = map[row].first, j = map[row].second;
      c.row_ptr[row] = nnz;
      e.clear();
      for (int k = 0; k < 5; ++k) {
        int ni = i + di[k], nj = j + dj[k];
        bool inb = ni >= 0 && ni < rows && nj >= 0 && nj < cols;
        if (inb && (k == 4 || lin[(size_t)ni * cols + nj] >= 0))
          e.push_back({k == 4 ? row : lin[(size_t)ni * cols + nj], ty[k]});
      }
      std::sort(e.begin(), e.end());
      for (size_t q = 0; q < e.size(); ++q) {
        int off = nnz + (int)q;
        c.col_idx.push_back(e[q].first);
        switch (e[q].second) {
          case 0: c.diag[row] = off; break;
          case 1: c.east[row] = off; break;
          case 2: c.west[row] = off; break;
          case 3: c.north[row] = off; break;
          case 4: c.south[row] = off; break;
        }
      }
      nnz += (int)e.size();
    }
    c.row_ptr[n] = nnz;
    c.values.assign(nnz, R(0));
    c.rhs.assign(n, R(0));
    c.sol.assign(n, R(0));
  }

  // ------------------------------------------------------------------------
  // momentum_solver.cpp:164-314 (assembly only; fills csrU / csrV, dE / dN)
  template <bool IsU>
  void assembleMomentum() {
    Csr<R>& csr = IsU ? csrU : csrV;
    const auto& map = IsU ? mapU : mapV;
    const R vol = hx * hy;
    const R maxVel = R(3.0f) * std::max(std::abs(inletVelocity), R(0.1f));
    // mutants 1 (6 mu), 4 (no vol), 2 (factor dropped)
    const R sinkCoeff = enable25D ? sinkMult25D * (R(kMutant == 1 ? 6.0f : 12.0f) * eta / (Ht * Ht)) : R(0);
    const R sinkDiag = kMutant == 4 ? sinkCoeff : sinkCoeff * vol;
    const R DeScale = hy / hx, DnScale = hx / hy;
    const R convScale = enable25D && kMutant != 2 ? convFactor25D : R(1.0f);
    const int n = csr.n;
#pragma omp parallel for schedule(static) if(n >= kOmpMinLoop)
    for (int row = 0; row < n; ++row) {
      const int i = map[row].first, j = map[row].second;
      R alphaLocal = IsU ? alphaAtU(i, j) : alphaAtV(i, j);
      R ue, uw, vn, vs;
      if (IsU) {
        ue = clampv(R(0.5f) * (U(i, j) + U(i, j + 1)), maxVel);
        uw = clampv(R(0.5f) * (U(i, j - 1) + U(i, j)), maxVel);
        vn = clampv(R(0.5f) * (V(i, j) + V(i, j + 1)), maxVel);
        vs = clampv(R(0.5f) * (V(i - 1, j) + V(i - 1, j + 1)), maxVel);
      } else {
        ue = clampv(R(0.5f) * (U(i, j) + U(i + 1, j)), maxVel);
        uw = clampv(R(0.5f) * (U(i, j - 1) + U(i + 1, j - 1)), maxVel);
        vn = clampv(R(0.5f) * (V(i, j) + V(i + 1, j)), maxVel);
        vs = clampv(R(0.5f) * (V(i - 1, j) + V(i, j)), maxVel);
      }
      R Fe = convScale * rho * hy * ue, Fw = convScale * rho * hy * uw;
      R Fn = convScale * rho * hx * vn, Fs = convScale * rho * hx * vs;
      R De = eta * DeScale, Dn = eta * DnScale;
      R aE = De + std::max(R(0), -Fe);
      R aW = De + std::max(R(0), Fw);
      R aN = Dn + std::max(R(0), -Fn);
      R aS = Dn + std::max(R(0), Fs);
      if (IsU) {
        if (southU(i, j)) aS += Dn;
        if (northU(i, j)) aN += Dn;
      } else {
        if (westV(i, j)) aW += De;
        if (eastV(i, j)) aE += De;
      }
      R sumA = aE + aW + aN + aS;
      R brinkmanDrag = kMutant == 3 ? alphaLocal : alphaLocal * vol;  // mutant 3: no vol
      R aP0 = sumA + (Fe - Fw + Fn - Fs) + R(0) + sinkDiag + brinkmanDrag;
      aP0 = std::max(aP0, R(1e-10f));
      R aP = aP0 / uvAlpha;
      R Sp, velCur;
      if (IsU) {
        Sp = (P(i, j) - P(i, j + 1)) * (kMutant == 6 ? hx : hy);  // mutant 6: wrong face
        velCur = U(i, j);
      } else {
        Sp = (P(i, j) - P(i + 1, j)) * hx;
        velCur = V(i, j);
      }
      if (kMutant == 5) Sp += (R(1.0f) - uvAlpha) * aP0 * velCur;  // mutant 5: relaxation
      else Sp += (R(1.0f) - uvAlpha) / uvAlpha * aP0 * velCur;
      if (convectionScheme == 1) Sp += IsU ? souU(i, j, Fe, Fw, Fn, Fs) : souV(i, j, Fe, Fw, Fn, Fs);
      if (IsU) dE[(size_t)i * N + j] = (kMutant == 11 ? hx : hy) / aP0;  // mutant 11: wrong face
      else dN[(size_t)i * (N + 1) + j] = hx / aP0;
      csr.values[csr.diag[row]] = aP;
      auto upd = [&](int off, R coeff, R nb) {
        if (off >= 0) csr.values[off] = -coeff;
        else if (kMutant != 7) Sp += coeff * nb;  // mutant 7: boundary values not folded
      };
      if (IsU) {
        upd(csr.east[row], aE, U(i, j + 1));
        upd(csr.west[row], aW, U(i, j - 1));
        upd(csr.north[row], aN, U(i + 1, j));
        upd(csr.south[row], aS, U(i - 1, j));
        csr.sol[row] = U(i, j);
      } else {
        upd(csr.east[row], aE, V(i, j + 1));
        upd(csr.west[row], aW, V(i, j - 1));
        upd(csr.north[row], aN, V(i + 1, j));
        upd(csr.south[row], aS, V(i - 1, j));
        csr.sol[row] = V(i, j);
      }
      csr.rhs[row] = Sp;
    }
  }
  static R clampv(R x, R m) { return std::min(std::max(x, -m), m); }

  // momentum_solver.cpp:317-355 (CPU Jacobi backend + scatter with clamp)
  template <bool IsU>
  R solveMomentum() {
    assembleMomentum<IsU>();
    Csr<R>& csr = IsU ? csrU : csrV;
    const auto& map = IsU ? mapU : mapV;
    int it = jacobiCsr(csr.n, csr.row_ptr, csr.col_idx, csr.values, csr.rhs, csr.sol, csr.diag, momentumTol,
                       maxMomentumIter);
    if (IsU) lastUIters = it; else lastVIters = it;
    const R maxVel = R(3.0f) * std::max(std::abs(inletVelocity), R(0.1f));
    R resid = 0;
    for (int row = 0; row < csr.n; ++row) {
      int i = map[row].first, j = map[row].second;
      R nv = clampv(csr.sol[row], maxVel);
      if (IsU) { resid = std::max(resid, std::abs(nv - U(i, j))); uStar[(size_t)i * N + j] = nv; }
      else { resid = std::max(resid, std::abs(nv - V(i, j))); vStar[(size_t)i * (N + 1) + j] = nv; }
    }
    return resid;
  }

  // jacobi_solver.cpp:20-75
  static int jacobiCsr(int n, const std::vector<int>& rp, const std::vector<int>& ci, const std::vector<R>& val,
                       const std::vector<R>& b, std::vector<R>& x, const std::vector<int>& diag, R tol, int maxIter) {
    if (n == 0) return 0;
    std::vector<R> buf(n);
    std::vector<R>* xc = &x;
    std::vector<R>* xn = &buf;
    R tol_sq = tol * tol;
    int iter = 0;
    for (; iter < maxIter; ++iter) {
      R total = 0;
#pragma omp parallel for reduction(+ : total) schedule(static) if(n >= kOmpMinLoop)
      for (int i = 0; i < n; ++i) {
        R Ax = 0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) Ax += val[k] * (*xc)[ci[k]];
        R r = b[i] - Ax;
        total += r * r;
        (*xn)[i] = (*xc)[i] + r / val[diag[i]];
      }
      std::swap(xc, xn);
      if (total < tol_sq * R(n)) { iter++; break; }
    }
    if (xc == &buf) std::copy(buf.begin(), buf.end(), x.begin());
    return iter;
  }

  // pressure_solver.cpp:108-184
  static int pcgCsr(int n, const std::vector<int>& rp, const std::vector<int>& ci, const std::vector<R>& val,
                    const std::vector<R>& b, std::vector<R>& x, R tol, int maxIter) {
    if (n == 0) return 0;
    std::vector<R> r(n), z(n), pv(n), Ap(n), diag(n);
    for (int i = 0; i < n; ++i) {
      R d = R(1.0f);
      for (int k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] == i) { d = val[k]; break; }
      diag[i] = d;
    }
    auto spmv = [&](const std::vector<R>& xx, std::vector<R>& y) {
#pragma omp parallel for schedule(static) if(n >= kOmpMinLoop)
      for (int i = 0; i < n; ++i) {
        R s = 0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) s += val[k] * xx[ci[k]];
        y[i] = s;
      }
    };
    spmv(x, Ap);
    for (int i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
    auto jac = [&](const std::vector<R>& rv, std::vector<R>& zv) {
      for (int i = 0; i < n; ++i) zv[i] = (std::abs(diag[i]) > R(1e-20f)) ? rv[i] / diag[i] : rv[i];
    };
    jac(r, z);
    pv = z;
    R rz = 0;
#pragma omp parallel for reduction(+ : rz) if(n >= kOmpMinLoop)
    for (int i = 0; i < n; ++i) rz += r[i] * z[i];
    int iter = 0;
    for (; iter < maxIter; ++iter) {
      spmv(pv, Ap);
      R pAp = 0;
#pragma omp parallel for reduction(+ : pAp) if(n >= kOmpMinLoop)
      for (int i = 0; i < n; ++i) pAp += pv[i] * Ap[i];
      if (std::abs(pAp) < R(1e-20f)) break;
      R a = rz / pAp;
#pragma omp parallel for if(n >= kOmpMinLoop)
      for (int i = 0; i < n; ++i) {
        x[i] += a * pv[i];
        r[i] -= a * Ap[i];
      }
      R err = 0;
#pragma omp parallel for reduction(+ : err) if(n >= kOmpMinLoop)
      for (int i = 0; i < n; ++i) err += r[i] * r[i];
      if (std::sqrt(err) < tol) break;
      jac(r, z);
      R rz_prev = rz;
      rz = 0;
#pragma omp parallel for reduction(+ : rz) if(n >= kOmpMinLoop)
      for (int i = 0; i < n; ++i) rz += r[i] * z[i];
      R beta = rz / rz_prev;
#pragma omp parallel for if(n >= kOmpMinLoop)
      for (int i = 0; i < n; ++i) pv[i] = z[i] + beta * pv[i];
    }
    return iter;
  }

  // pressure_solver.cpp:470-496 : max and RMS of the mass imbalance over
  // i = 1..M-1, j = 1..N-1 (column N-1 included)
  void massResidual() {
    R mx = 0;
    double ss = 0;
    long long cnt = 0;
    for (int i = 1; i < M; ++i)
      for (int j = 1; j < N; ++j) {
        R mE = rho * uStar[(size_t)i * N + j] * hy;
        R mW = rho * uStar[(size_t)i * N + j - 1] * hy;
        R mN = rho * vStar[(size_t)i * (N + 1) + j] * hx;
        R mS = rho * vStar[(size_t)(i - 1) * (N + 1) + j] * hx;
        R b = std::abs(mW - mE + mS - mN);
        mx = std::max(mx, b);
        ss += (double)b * (double)b;
        cnt++;
      }
    residMass_max = mx;
    residMass_RMS = cnt > 0 ? R(std::sqrt(ss / (double)cnt)) : R(0);
  }

  // pressure_solver.cpp:703-742 : values + RHS of the pressure system
  void assemblePressure() {
    Csr<R>& c = csrP;
#pragma omp parallel for schedule(static) if(c.n >= kOmpMinLoop)
    for (int row = 0; row < c.n; ++row) {
      int i = mapP[row].first, j = mapP[row].second;
      R aE = (j < N - 1) ? rho * dE[(size_t)i * N + j] * hy : R(0);
      R aW = (j > 1) ? rho * dE[(size_t)i * N + j - 1] * hy : R(0);
      R aN = (i < M - 1) ? rho * dN[(size_t)i * (N + 1) + j] * hx : R(0);
      R aS = (i > 1) ? rho * dN[(size_t)(i - 1) * (N + 1) + j] * hx : R(0);
      R aP = aE + aW + aN + aS;
      if (aP < R(1e-20f)) aP = R(1.0f);
      c.values[c.diag[row]] = aP;
      if (c.east[row] >= 0) c.values[c.east[row]] = -aE;
      if (c.west[row] >= 0) c.values[c.west[row]] = -aW;
      if (c.north[row] >= 0) c.values[c.north[row]] = -aN;
      if (c.south[row] >= 0) c.values[c.south[row]] = -aS;
      R mE = rho * uStar[(size_t)i * N + j] * hy;
      R mW = rho * uStar[(size_t)i * N + j - 1] * hy;
      R mN = rho * vStar[(size_t)i * (N + 1) + j] * hx;
      R mS = rho * vStar[(size_t)(i - 1) * (N + 1) + j] * hx;
      c.rhs[row] = mW - mE + mS - mN;
      c.sol[row] = pStar[(size_t)i * (N + 1) + j];
    }
  }

  // pressure_solver.cpp:462-922
  int solvePressure() {
    std::fill(pStar.begin(), pStar.end(), R(0));
    massResidual();
    int iters = 0;
    if (pressureSolverType >= 1) {
      assemblePressure();
      Csr<R>& c = csrP;
      iters = pcgCsr(c.n, c.row_ptr, c.col_idx, c.values, c.rhs, c.sol, pTol, maxPressureIter);
      for (int row = 0; row < c.n; ++row) {
        int i = mapP[row].first, j = mapP[row].second;
        pStar[(size_t)i * (N + 1) + j] = c.sol[row];
      }
      for (int i = 1; i < M; ++i) pStar[(size_t)i * (N + 1) + N - 1] = R(0);
    } else {
      iters = solveSor();
    }
    return iters;
  }

  // pressure_solver.cpp:819-919 (red-black SOR; each colour pass in order)
  int solveSor() {
    int maxDim = std::max(M, N);
    R sr = R(std::cos(3.14159265358979323846 / maxDim));
    R omega = R(2.0f) / (R(1.0f) + std::sqrt(R(1.0f) - sr * sr));
    omega = std::max(R(1.0f), std::min(R(1.95f), omega));
    int it = 0;
    auto PS = [&](int i, int j) -> R& { return pStar[(size_t)i * (N + 1) + j]; };
    for (int pIter = 0; pIter < maxPressureIter; ++pIter) {
      R maxChange = 0;
      it = pIter + 1;
      for (int colour = 0; colour < 2; ++colour) {
        for (int i = 1; i < M; ++i)
          for (int j = 1; j < N; ++j) {
            if ((i + j) % 2 != colour) continue;
            R aE = (j < N - 1) ? rho * dE[(size_t)i * N + j] * hy : R(0);
            R aW = (j > 1) ? rho * dE[(size_t)i * N + j - 1] * hy : R(0);
            R aN = (i < M - 1) ? rho * dN[(size_t)i * (N + 1) + j] * hx : R(0);
            R aS = (i > 1) ? rho * dN[(size_t)(i - 1) * (N + 1) + j] * hx : R(0);
            R aP = aE + aW + aN + aS;
            if (j == N - 1 || aP < R(1e-20f)) { PS(i, j) = R(0); continue; }
            R mE = rho * uStar[(size_t)i * N + j] * hy;
            R mW = rho * uStar[(size_t)i * N + j - 1] * hy;
            R mN = rho * vStar[(size_t)i * (N + 1) + j] * hx;
            R mS = rho * vStar[(size_t)(i - 1) * (N + 1) + j] * hx;
            R b = mW - mE + mS - mN;
            R pE = (j < N - 1) ? PS(i, j + 1) : R(0);
            R pW = (j > 1) ? PS(i, j - 1) : R(0);
            R pN = (i < M - 1) ? PS(i + 1, j) : R(0);
            R pS = (i > 1) ? PS(i - 1, j) : R(0);
            R pGS = (aE * pE + aW * pW + aN * pN + aS * pS + b) / aP;
            R pNew = PS(i, j) + omega * (pGS - PS(i, j));
            pNew = std::max(R(-10000.0f), std::min(R(10000.0f), pNew));
            maxChange = std::max(maxChange, std::abs(pNew - PS(i, j)));
            PS(i, j) = pNew;
          }
      }
      if (maxChange < pTol) break;
    }
    return it;
  }

  // ------------------------------------------------------------------------
  // iterations.cpp:82-659 (momentumSolverType = 1, pseudo-time off,
  // residual viscosity off, raw RMS residual reporting)
  StepStats calculateStep() {
    StepStats st;
    const R maxVel = R(3.0f) * std::max(std::abs(inletVelocity), R(0.1f));
    setVelocityBC(u, v);
    setPressureBC(p);
    uOld = u;
    vOld = v;
    residU = solveMomentum<true>();
    residV = solveMomentum<false>();
    setVelocityBC(uStar, vStar);
    residU_max = residU;
    residV_max = residV;
    {
      double su = 0, sv = 0;
      long long cu = 0, cv = 0;
      for (int i = 1; i < M; ++i)
        for (int j = 1; j < N - 1; ++j) {
          R d = uStar[(size_t)i * N + j] - uOld[(size_t)i * N + j];
          su += (double)d * (double)d;
          cu++;
        }
      for (int i = 1; i < M - 1; ++i)
        for (int j = 1; j < N; ++j) {
          R d = vStar[(size_t)i * (N + 1) + j] - vOld[(size_t)i * (N + 1) + j];
          sv += (double)d * (double)d;
          cv++;
        }
      residU_RMS = cu > 0 ? R(std::sqrt(su / (double)cu)) : R(0);
      residV_RMS = cv > 0 ? R(std::sqrt(sv / (double)cv)) : R(0);
    }
    st.p_iters = solvePressure();
    // pressure update (iterations.cpp:558-567)
    for (int i = 1; i < M; ++i)
      for (int j = 1; j < N; ++j) P(i, j) += pAlpha * pStar[(size_t)i * (N + 1) + j];
    setPressureBC(p);
    // velocity correction (iterations.cpp:575-599)
    for (int i = 1; i < M; ++i)
      for (int j = 1; j < N - 1; ++j) {
        R du = dE[(size_t)i * N + j] * (pStar[(size_t)i * (N + 1) + j] - pStar[(size_t)i * (N + 1) + j + 1]);
        du = clampv(du, maxVel);
        R nu = uStar[(size_t)i * N + j] + du;
        U(i, j) = clampv(nu, maxVel);
      }
    for (int i = 1; i < M - 1; ++i)
      for (int j = 1; j < N; ++j) {
        R dv = dN[(size_t)i * (N + 1) + j] * (pStar[(size_t)i * (N + 1) + j] - pStar[(size_t)(i + 1) * (N + 1) + j]);
        dv = clampv(dv, maxVel);
        R nv = vStar[(size_t)i * (N + 1) + j] + dv;
        V(i, j) = clampv(nv, maxVel);
      }
    setVelocityBC(u, v);
    residU = residU_RMS;
    residV = residV_RMS;
    residMass = residMass_RMS;
    st.resid_u = residU;
    st.resid_v = residV;
    st.resid_mass = residMass;
    st.resid_u_max = residU_max;
    st.resid_v_max = residV_max;
    st.resid_mass_max = residMass_max;
    st.u_iters = lastUIters;
    st.v_iters = lastVIters;
    st.error = std::max(residU, std::max(residV, residMass));
    return st;
  }

  // postprocessing.cpp:245-327
  PatchMetrics boundaryPatch(bool inlet) {
    PatchMetrics m;
    const int MCells = M - 1, NCells = N - 1;
    const int uCol = inlet ? 0 : N - 1, pCol = inlet ? 0 : N, vCol = inlet ? 0 : N;
    const int adj = inlet ? 0 : NCells - 1;
    R sA = 0, dA = 0, tw = 0, area = 0, flux = 0;
    for (int i = 1; i < M; ++i) {
      const int cr = i - 1;
      if (cr < 0 || cr >= MCells) continue;
      if (!pureFluid(cr, adj)) continue;
      const R fa = hy;
      area += fa;
      const int vTop = std::max(0, std::min(i - 1, M - 1));
      const int vBot = std::max(0, std::min(i, M - 1));
      const R sp = P(i, pCol);
      const R un = U(i, uCol);
      const R vt = R(0.5f) * (V(vTop, vCol) + V(vBot, vCol));
      const R vm = std::sqrt(un * un + vt * vt);
      const R dyn = R(0.5f) * rho * vm * vm;
      const R tot = sp + dyn;
      sA += sp * fa;
      dA += dyn * fa;
      const R fl = std::abs(rho * un * fa);
      if (fl > R(0)) { flux += fl; tw += tot * fl; }
    }
    m.area = area;
    m.mass_flux = flux;
    if (area > R(0)) { m.avg_static = sA / area; m.avg_dynamic = dA / area; m.valid = 1; }
    if (flux > R(0)) m.avg_total = tw / flux;
    else if (area > R(0)) m.avg_total = m.avg_static + m.avg_dynamic;
    return m;
  }

  // postprocessing.cpp:62-96: the cell columns cl, cl + 1 whose centres
  // (c + 1/2) hx bracket x (clamped to the first / last centre), t in [0, 1].
  void bracket(R x, int& cl, R& t) const {
    const int NCells = N - 1;
    cl = 0;
    t = R(0);
    if (NCells <= 1) return;
    const R xc = std::max(R(0.5f) * hx, std::min(x, (R(NCells) - R(0.5f)) * hx));
    cl = (int)std::floor(xc / hx - R(0.5f));
    cl = std::max(0, std::min(cl, NCells - 2));
    const R xl = (R(cl) + R(0.5f)) * hx, xr = (R(cl + 1) + R(0.5f)) * hx;
    t = std::abs(xr - xl) < R(1e-12f) ? R(0) : (xc - xl) / (xr - xl);
    t = std::max(R(0), std::min(R(1), t));
  }

  // postprocessing.cpp:102-243 (samplePlaneAtX): p at columns cl+1, cl+2,
  // u at cl, cl+1, v at cl, cl+1 (rows i-1, i), each lerped by t; rows where
  // both bracketing cells are pure fluid.
  PatchMetrics planeAtX(R x) {
    PatchMetrics m;
    const int MCells = M - 1, NCells = N - 1;
    if (MCells <= 0 || NCells <= 0 || !(hx > R(0)) || !(hy > R(0))) return m;
    int cl;
    R t;
    bracket(x, cl, t);
    const int cr = NCells > 1 ? cl + 1 : cl;
    const R w = R(1.0f) - t;
    R sA = 0, dA = 0, tw = 0, area = 0, flux = 0;
    for (int i = 1; i < M; ++i) {
      const int row = i - 1;
      if (!pureFluid(row, cl) || !pureFluid(row, cr)) continue;
      const R fa = hy;
      area += fa;
      const R sp = w * P(i, cl + 1) + t * P(i, cr + 1);
      const R un = w * U(i, cl) + t * U(i, cr);
      const int vTop = std::max(0, std::min(i - 1, M - 1));
      const int vBot = std::max(0, std::min(i, M - 1));
      const R vt = w * V(vTop, cl) + t * V(vTop, cr);
      const R vb = w * V(vBot, cl) + t * V(vBot, cr);
      const R vtan = R(0.5f) * (vt + vb);
      const R vm = std::sqrt(un * un + vtan * vtan);
      const R dyn = R(0.5f) * rho * vm * vm;
      sA += sp * fa;
      dA += dyn * fa;
      const R fl = std::abs(rho * un * fa);
      if (fl > R(0)) { flux += fl; tw += (sp + dyn) * fl; }
    }
    m.area = area;
    m.mass_flux = flux;
    if (area > R(0)) { m.avg_static = sA / area; m.avg_dynamic = dA / area; m.valid = 1; }
    if (flux > R(0)) m.avg_total = tw / flux;
    else if (area > R(0)) m.avg_total = m.avg_static + m.avg_dynamic;
    return m;
  }

  struct DropStats {
    double core_total = 0, core_static = 0, full_total = 0, full_static = 0;
  };
  // SIMPLE.cpp:421-439
  DropStats drops() {
    const PatchMetrics ci = planeAtX(xCoreIn), co = planeAtX(xCoreOut);
    const PatchMetrics fi = boundaryPatch(true), fo = boundaryPatch(false);
    DropStats d;
    d.core_total = (double)(R(ci.avg_total) - R(co.avg_total));
    d.core_static = (double)(R(ci.avg_static) - R(co.avg_static));
    d.full_total = (double)(R(fi.avg_total) - R(fo.avg_total));
    d.full_static = (double)(R(fi.avg_static) - R(fo.avg_static));
    return d;
  }

  // SIMPLE.cpp:347-366: least-squares slope of (iteration, core dP) pairs,
  // iterations taken relative to the first.
  static R dpSlope(const std::vector<std::pair<int, R>>& buf) {
    if (buf.size() < 2) return R(0);
    const R x0 = R(buf.front().first), n = R(buf.size());
    R sx = 0, sy = 0, sxx = 0, sxy = 0;
    for (const auto& e : buf) {
      const R x = R(e.first) - x0, y = e.second;
      sx += x;
      sy += y;
      sxx += x * x;
      sxy += x * y;
    }
    const R den = n * sxx - sx * sx;
    if (std::abs(den) < R(1e-12f)) return R(0);
    return (n * sxy - sx * sy) / den;
  }

  // SIMPLE.cpp:310-491 (runIterations) without logging, checkpoints and the
  // pseudo-time controller (enablePseudoTimeStepping off): inlet ramp, one
  // calculateStep, the four dPs, the core-dP window and slope gate.
  // trace(iter, stats, drops) is called after every iteration.
  template <class F>
  int runIterations(StepStats& st, DropStats& dl, bool& dpConverged, F&& trace) {
    int iter = 0;
    R error = R(1.0f);
    const int win = std::max(1, dpConvergenceWindow);
    std::vector<R> hist(win, R(0));
    int idx = 0;
    bool filled = false;
    dpConverged = false;
    std::vector<std::pair<int, R>> slopeBuf;
    while (iter < maxIterations && error > epsilon && !dpConverged) {
      inletScale = (enableInletRamp && rampSteps > 0 && iter < rampSteps) ? R(iter + 1) / R(rampSteps) : R(1);
      inletVelocity = inletScale * targetVel;
      st = calculateStep();
      error = R(st.error);
      dl = drops();
      const R core = R(dl.core_total);
      if (usePressureDropSlopeGate && dpSlopeWindowIters > 0) {
        slopeBuf.emplace_back(iter, core);
        if (slopeBuf.size() > (size_t)dpSlopeWindowIters) slopeBuf.erase(slopeBuf.begin());
      }
      if (usePressureDropConvergence && iter >= rampSteps) {
        hist[idx] = core;
        idx = (idx + 1) % win;
        if (idx == 0) filled = true;
        if (filled) {
          double sum = 0.0;
          for (int k = 0; k < win; ++k) sum += (double)hist[k];
          const double mean = sum / double(win);
          double var = 0.0;
          for (int k = 0; k < win; ++k) {
            const double d = (double)hist[k] - mean;
            var += d * d;
          }
          const double sd = std::sqrt(var / double(win));
          const double den = std::max(std::abs(mean), 1.0);
          const double stdPct = sd / den * 100.0;
          const double latest = (double)hist[(idx - 1 + win) % win];
          const double lastPct = std::abs(latest - mean) / den * 100.0;
          bool slopeOk = true;
          if (usePressureDropSlopeGate && dpSlopeWindowIters > 0 && slopeBuf.size() >= (size_t)dpSlopeWindowIters)
            slopeOk = std::abs(dpSlope(slopeBuf)) <= dpSlopeMaxDelta / R(dpSlopeWindowIters);
          if (stdPct < (double)dpConvergencePercent && lastPct < (double)dpConvergencePercent && slopeOk)
            dpConverged = true;
        }
      }
      trace(iter, st, dl);
      iter++;
    }
    return iter;
  }
};

}  // namespace oracle
