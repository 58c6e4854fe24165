// ============================================================================
// oracle/thermal_oracle.hpp -- CPU restatement of the reference 3-D
// conjugate-heat-transfer assembly (FP64).
//
// TEST INFRASTRUCTURE ONLY (oracle + CPU baseline; never in the product path).
//
// Follows ThermalSolver/thermal_solver.cpp:
//   :235      idx(i,j,k) = k*Nx*Ny + j*Nx + i  (x fastest)
//   :236      harm(a,b) = 2ab/(a+b+1e-30)
//   :240-244  region: fluid iff channel layer and gamma > 0.01
//   :271-300  kval: solid k_s; channel RAMP / harmonic / linear in gamma
//   :317-321  dk/dgamma of the RAMP law
//   :323-327  pois(k) = 4z(1-z), z = (k - nz_s + 0.5) dz_f / H_t
//   :437-528  7-point assembly: harmonic-k diffusion, upwind advection only
//             between two fluid cells, inlet Dirichlet at i = 0, outflow at
//             i = Nx-1 and j = Ny-1 (none at j = 0), series z-conductance,
//             bottom heat flux, adiabatic top, diag += 1e-12
//   :88-133   duplicate COO entries summed (CSR with sorted columns)
//   :592-620  energy balance Q_in vs mass-averaged outlet enthalpy
//   thermal_output.cpp:202  T_avg_base (volume-weighted mean over base layers)
//   thermal_solver.cpp:345-358  p-norm over layer k = 0
//
// Extension (not in the reference): the layer stack is a list of layers,
// each solid (k_s, no advection) or channel (gamma-interpolated k, advected);
// solid layers above the channel are a lid.  With layers = nz_s solid then
// nz_f channel, this is exactly the reference stack.
// ============================================================================
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

namespace oracle {

#ifndef ORACLE_OMP_MIN_LOOP_DEFINED
#define ORACLE_OMP_MIN_LOOP_DEFINED
// loops shorter than this run on one thread: on small grids the fork/join
// and barrier cost of a parallel region exceeds the loop's work, badly so
// when other processes hold the cores
constexpr int kOmpMinLoop = 16384;
#endif

struct ThermalProblem {
  int nx = 0, ny = 0, nz = 0;
  double dx = 0, dy = 0;
  std::vector<double> dz;         // nz layer thicknesses
  std::vector<int> channel;       // nz flags: 1 = channel layer
  double ht = 0;                  // channel height for the z-profile
  double k_s = 400.0, k_f = 0.6, rho = 998.0, cp = 4180.0, t_in = 30.0;
  int k_mode = 1;                 // 1 RAMP, 2 harmonic, 0 linear
  double qk = 1.0;
  std::vector<double> gamma;      // ny*nx, 1 = fluid
  std::vector<double> u, v;       // ny*nx cell-centred velocities
  std::vector<double> q0;         // ny*nx bottom heat flux [W/m^2]

  int n() const { return nx * ny * nz; }
  int idx(int i, int j, int k) const { return k * nx * ny + j * nx + i; }
  static double harm(double a, double b) { return 2 * a * b / (a + b + 1e-30); }
  int firstChannel() const {
    for (int k = 0; k < nz; ++k)
      if (channel[k]) return k;
    return nz;
  }
  int region(int i, int j, int k) const {
    if (!channel[k]) return 0;
    return gamma[j * nx + i] > 0.01 ? 1 : 0;
  }
  double kval(int i, int j, int k) const {
    if (!channel[k]) return k_s;
    double g = std::max(0.0, std::min(1.0, gamma[j * nx + i]));
    if (k_mode == 1) return k_f + (k_s - k_f) * (1.0 - g) / (1.0 + qk * g);
    if (k_mode == 2) return k_s * k_f / ((1.0 - g) * k_f + g * k_s + 1e-30);
    return k_s * (1.0 - g) + k_f * g;
  }
  double dkdg(double g) const {
    g = std::max(0.0, std::min(1.0, g));
    double den = 1.0 + qk * g;
    return (k_s - k_f) * (-(1.0 + qk)) / (den * den);
  }
  // z-profile: channel layers are measured from the first channel layer
  double pois(int k) const {
    if (!channel[k]) return 0.0;
    int k0 = firstChannel();
    double zb = 0;
    for (int q = k0; q < k; ++q) zb += dz[q];
    double z = (zb + 0.5 * dz[k]) / ht;
    return 4.0 * z * (1 - z);
  }
};

struct ThermalCsr {
  int n = 0;
  std::vector<int> row_ptr, col_idx;
  std::vector<double> values, b;
};

// Assemble A T = b (thermal_solver.cpp:437-530).  Each row's entries are
// accumulated per neighbour and emitted in sorted column order.
inline ThermalCsr assembleThermal(const ThermalProblem& P) {
  const int nx = P.nx, ny = P.ny, nz = P.nz, n = P.n();
  const double dx = P.dx, dy = P.dy, rc = P.rho * P.cp, Tin = P.t_in;
  ThermalCsr A;
  A.n = n;
  A.row_ptr.assign(n + 1, 0);
  A.b.assign(n, 0.0);
  std::vector<std::vector<std::pair<int, double>>> rows(n);
#pragma omp parallel for schedule(static) if((long)nz * nx * ny >= kOmpMinLoop)
  for (int k = 0; k < nz; ++k) {
    const double dzk = P.dz[k], Ax = dy * dzk, Ay = dx * dzk, Az = dx * dy;
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int p = P.idx(i, j, k);
        double diag = 0, kP = P.kval(i, j, k), po = P.pois(k);
        double g = P.channel[k] ? std::max(0.0, std::min(1.0, P.gamma[j * nx + i])) : 0.0;
        double uP = g > 0.01 ? P.u[j * nx + i] * po : 0.0;
        double vP = g > 0.01 ? P.v[j * nx + i] * po : 0.0;
        const bool fl = P.region(i, j, k) == 1;
        double cW = 0, cE = 0, cS = 0, cN = 0, cB = 0, cT = 0;
        if (i > 0) {
          double D = P.harm(kP, P.kval(i - 1, j, k)) * Ax / dx;
          diag += D; cW -= D;
          if (fl && P.region(i - 1, j, k) == 1) {
            double F = rc * 0.5 * (uP + P.u[j * nx + i - 1] * po) * Ax;
            diag += std::max(-F, 0.0); cW -= std::max(F, 0.0);
          }
        } else if (fl) {
          double Dbc = 2.0 * kP * Ax / dx, Fbc = rc * uP * Ax;
          diag += Dbc; A.b[p] += Dbc * Tin;
          if (Fbc > 0) A.b[p] += Fbc * Tin; else diag += -Fbc;
        }
        if (i < nx - 1) {
          double D = P.harm(kP, P.kval(i + 1, j, k)) * Ax / dx;
          diag += D; cE -= D;
          if (fl && P.region(i + 1, j, k) == 1) {
            double F = rc * 0.5 * (uP + P.u[j * nx + i + 1] * po) * Ax;
            diag += std::max(F, 0.0); cE -= std::max(-F, 0.0);
          }
        } else if (fl && uP > 0) diag += rc * uP * Ax;
        if (j > 0) {
          double D = P.harm(kP, P.kval(i, j - 1, k)) * Ay / dy;
          diag += D; cS -= D;
          if (fl && P.region(i, j - 1, k) == 1) {
            double F = rc * 0.5 * (vP + P.v[(j - 1) * nx + i] * po) * Ay;
            diag += std::max(-F, 0.0); cS -= std::max(F, 0.0);
          }
        }
        if (j < ny - 1) {
          double D = P.harm(kP, P.kval(i, j + 1, k)) * Ay / dy;
          diag += D; cN -= D;
          if (fl && P.region(i, j + 1, k) == 1) {
            double F = rc * 0.5 * (vP + P.v[(j + 1) * nx + i] * po) * Ay;
            diag += std::max(F, 0.0); cN -= std::max(-F, 0.0);
          }
        } else if (fl && vP > 0) diag += rc * vP * Ay;
        if (k > 0) {
          double kN = P.kval(i, j, k - 1), dP = P.dz[k], dN = P.dz[k - 1];
          double kf = (dP + dN) / (dP / kP + dN / kN);
          double D = kf * Az / (0.5 * (dP + dN));
          diag += D; cB -= D;
        } else A.b[p] += P.q0[j * nx + i] * Az;
        if (k < nz - 1) {
          double kN = P.kval(i, j, k + 1), dP = P.dz[k], dN = P.dz[k + 1];
          double kf = (dP + dN) / (dP / kP + dN / kN);
          double D = kf * Az / (0.5 * (dP + dN));
          diag += D; cT -= D;
        }
        auto& r = rows[p];
        // sorted column order: below, south, west, diag, east, north, above
        if (k > 0) r.push_back({P.idx(i, j, k - 1), cB});
        if (j > 0) r.push_back({P.idx(i, j - 1, k), cS});
        if (i > 0) r.push_back({P.idx(i - 1, j, k), cW});
        r.push_back({p, diag + 1e-12});
        if (i < nx - 1) r.push_back({P.idx(i + 1, j, k), cE});
        if (j < ny - 1) r.push_back({P.idx(i, j + 1, k), cN});
        if (k < nz - 1) r.push_back({P.idx(i, j, k + 1), cT});
      }
  }
  for (int p = 0; p < n; ++p) A.row_ptr[p + 1] = A.row_ptr[p] + (int)rows[p].size();
  A.col_idx.resize(A.row_ptr[n]);
  A.values.resize(A.row_ptr[n]);
  for (int p = 0; p < n; ++p) {
    int o = A.row_ptr[p];
    for (auto& e : rows[p]) { A.col_idx[o] = e.first; A.values[o] = e.second; ++o; }
  }
  return A;
}

// Jacobi-preconditioned BiCGStab on the CSR system (CPU baseline solver;
// the reference uses AMGCL IDR(8), which is absent).  Relative tolerance on
// ||b - A x|| / ||b||.  Returns iterations (negative if not converged).
inline int bicgstabCsr(const ThermalCsr& A, std::vector<double>& x, double rtol, int maxit, double* relres) {
  const int n = A.n;
  std::vector<double> r(n), rh(n), p(n, 0), vv(n, 0), s(n), t(n), ph(n), sh(n), dinv(n);
  for (int i = 0; i < n; ++i) {
    double d = 1;
    for (int q = A.row_ptr[i]; q < A.row_ptr[i + 1]; ++q)
      if (A.col_idx[q] == i) d = A.values[q];
    dinv[i] = 1.0 / d;
  }
  auto spmv = [&](const std::vector<double>& xx, std::vector<double>& y) {
#pragma omp parallel for schedule(static) if(n >= kOmpMinLoop)
    for (int i = 0; i < n; ++i) {
      double acc = 0;
      for (int q = A.row_ptr[i]; q < A.row_ptr[i + 1]; ++q) acc += A.values[q] * xx[A.col_idx[q]];
      y[i] = acc;
    }
  };
  auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
    double acc = 0;
#pragma omp parallel for reduction(+ : acc) schedule(static) if(n >= kOmpMinLoop)
    for (int i = 0; i < n; ++i) acc += a[i] * b[i];
    return acc;
  };
  spmv(x, r);
  for (int i = 0; i < n; ++i) r[i] = A.b[i] - r[i];
  double bn = std::sqrt(dot(A.b, A.b));
  if (bn == 0) bn = 1;
  rh = r;
  double rho = 1, alpha = 1, omega = 1;
  double res = std::sqrt(dot(r, r)) / bn;
  int it = 0;
  for (; it < maxit && res > rtol; ++it) {
    double rho1 = dot(rh, r);
    double beta = (rho1 / rho) * (alpha / omega);
    rho = rho1;
#pragma omp parallel for schedule(static) if(n >= kOmpMinLoop)
    for (int i = 0; i < n; ++i) {
      p[i] = r[i] + beta * (p[i] - omega * vv[i]);
      ph[i] = dinv[i] * p[i];
    }
    spmv(ph, vv);
    alpha = rho / dot(rh, vv);
#pragma omp parallel for schedule(static) if(n >= kOmpMinLoop)
    for (int i = 0; i < n; ++i) {
      s[i] = r[i] - alpha * vv[i];
      sh[i] = dinv[i] * s[i];
    }
    spmv(sh, t);
    omega = dot(t, s) / dot(t, t);
#pragma omp parallel for schedule(static) if(n >= kOmpMinLoop)
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * ph[i] + omega * sh[i];
      r[i] = s[i] - omega * t[i];
    }
    res = std::sqrt(dot(r, r)) / bn;
  }
  if (relres) *relres = res;
  return res <= rtol ? it : -it;
}

}  // namespace oracle
