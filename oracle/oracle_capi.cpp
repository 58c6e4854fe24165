// ============================================================================
// oracle/oracle_capi.cpp -- C ABI of the CPU oracle (liboracle.so).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and bench.py's
// cpu_baseline leg as the checker; never by the product library.
// Every entry point takes/returns doubles; prec = 32 runs the float
// instantiation (the reference's own precision), prec = 64 the FP64 oracle.
// ============================================================================
#include <chrono>
#include <memory>

#include "simple_oracle.hpp"
#include "thermal_oracle.hpp"

using oracle::Simple;

namespace {

struct Handle {
  int prec = 64;
  std::unique_ptr<Simple<float>> f;
  std::unique_ptr<Simple<double>> d;
};

template <class F>
auto with(Handle* h, F&& fn) {
  if (h->prec == 32) return fn(*h->f);
  return fn(*h->d);
}

template <class R>
std::vector<R>* fieldPtr(Simple<R>& s, int field) {
  switch (field) {
    case CPS_FIELD_U: return &s.u;
    case CPS_FIELD_V: return &s.v;
    case CPS_FIELD_P: return &s.p;
    case CPS_FIELD_USTAR: return &s.uStar;
    case CPS_FIELD_VSTAR: return &s.vStar;
    case CPS_FIELD_PSTAR: return &s.pStar;
    case CPS_FIELD_DE: return &s.dE;
    case CPS_FIELD_DN: return &s.dN;
    case CPS_FIELD_GAMMA: return &s.gamma;
    case CPS_FIELD_ALPHA: return &s.alpha;
  }
  return nullptr;
}

void toStats(const oracle::StepStats& st, int iteration, cps_step_stats* out) {
  if (!out) return;
  out->error = st.error;
  out->resid_u = st.resid_u;
  out->resid_v = st.resid_v;
  out->resid_mass = st.resid_mass;
  out->resid_u_max = st.resid_u_max;
  out->resid_v_max = st.resid_v_max;
  out->resid_mass_max = st.resid_mass_max;
  out->u_iters = st.u_iters;
  out->v_iters = st.v_iters;
  out->p_iters = st.p_iters;
  out->iteration = iteration;
}

template <class D>
void toDrops(const D& d, int converged, cps_step_stats* out) {
  if (!out) return;
  out->dp_core_total = d.core_total;
  out->dp_core_static = d.core_static;
  out->dp_full_total = d.full_total;
  out->dp_full_static = d.full_static;
  out->dp_converged = converged;
}

// Stand-alone CSR solvers on double inputs (prec 32 runs them in float).
template <class R>
int jacobiT(int n, const int* rp, const int* ci, const double* val, const double* b, double* x, double tol,
                   int maxit) {
  std::vector<int> vrp(rp, rp + n + 1), vci(ci, ci + rp[n]), diag(n, -1);
  std::vector<R> vv(val, val + rp[n]), vb(b, b + n), vx(x, x + n);
  for (int i = 0; i < n; ++i)
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] == i) diag[i] = k;
  int it = Simple<R>::jacobiCsr(n, vrp, vci, vv, vb, vx, diag, R(tol), maxit);
  for (int i = 0; i < n; ++i) x[i] = vx[i];
  return it;
}
template <class R>
int pcgT(int n, const int* rp, const int* ci, const double* val, const double* b, double* x, double tol,
                int maxit) {
  std::vector<int> vrp(rp, rp + n + 1), vci(ci, ci + rp[n]);
  std::vector<R> vv(val, val + rp[n]), vb(b, b + n), vx(x, x + n);
  int it = Simple<R>::pcgCsr(n, vrp, vci, vv, vb, vx, R(tol), maxit);
  for (int i = 0; i < n; ++i) x[i] = vx[i];
  return it;
}

}  // namespace

extern "C" {

void* orc_simple_create(int prec, const cps_simple_config* cfg, const double* gamma, const double* inlet) {
  auto* h = new Handle;
  h->prec = prec == 32 ? 32 : 64;
  if (h->prec == 32) {
    h->f.reset(new Simple<float>);
    h->f->setup(*cfg, gamma, inlet);
  } else {
    h->d.reset(new Simple<double>);
    h->d->setup(*cfg, gamma, inlet);
  }
  return h;
}

void orc_simple_destroy(void* hp) { delete static_cast<Handle*>(hp); }

// n SIMPLE steps; stats of the last one; returns the number of steps done.
int orc_simple_step(void* hp, int n, cps_step_stats* out, double* seconds) {
  auto* h = static_cast<Handle*>(hp);
  oracle::StepStats st;
  auto t0 = std::chrono::steady_clock::now();
  for (int k = 0; k < n; ++k) st = with(h, [](auto& s) { return s.calculateStep(); });
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  toStats(st, n, out);
  with(h, [&](auto& s) { toDrops(s.drops(), 0, out); return 0; });
  return n;
}

// runIterations (SIMPLE.cpp:310-491); history / dp_history receive the
// error and the core total-pressure drop of the first history_len iterations.
int orc_simple_run(void* hp, cps_step_stats* out, double* history, double* dp_history, int history_len) {
  auto* h = static_cast<Handle*>(hp);
  return with(h, [&](auto& s) {
    oracle::StepStats st;
    typename std::remove_reference_t<decltype(s)>::DropStats dl;
    bool conv = false;
    const int iters = s.runIterations(st, dl, conv, [&](int it, const oracle::StepStats& x, const auto& d) {
      if (it < history_len) {
        if (history) history[it] = x.error;
        if (dp_history) dp_history[it] = d.core_total;
      }
    });
    toStats(st, iters, out);
    toDrops(dl, conv ? 1 : 0, out);
    return iters;
  });
}

// samplePlaneAtX (postprocessing.cpp:102-243) at the current fields.
int orc_simple_plane(void* hp, double x, cps_patch* out) {
  auto* h = static_cast<Handle*>(hp);
  oracle::PatchMetrics m = with(h, [&](auto& s) {
    using R = std::remove_reference_t<decltype(s.u[0])>;
    return s.planeAtX(R(x));
  });
  out->area = m.area;
  out->mass_flux = m.mass_flux;
  out->avg_static = m.avg_static;
  out->avg_dynamic = m.avg_dynamic;
  out->avg_total = m.avg_total;
  out->valid = m.valid;
  return 0;
}

int orc_simple_get_field(void* hp, int field, double* out) {
  auto* h = static_cast<Handle*>(hp);
  return with(h, [&](auto& s) {
    auto* f = fieldPtr(s, field);
    if (!f) return 1;
    for (size_t k = 0; k < f->size(); ++k) out[k] = (double)(*f)[k];
    return 0;
  });
}

int orc_simple_set_field(void* hp, int field, const double* in) {
  auto* h = static_cast<Handle*>(hp);
  return with(h, [&](auto& s) {
    auto* f = fieldPtr(s, field);
    if (!f) return 1;
    using R = std::remove_reference_t<decltype((*f)[0])>;
    for (size_t k = 0; k < f->size(); ++k) (*f)[k] = (R)in[k];
    return 0;
  });
}

int orc_simple_boundary_patch(void* hp, int inlet, cps_patch* out) {
  auto* h = static_cast<Handle*>(hp);
  oracle::PatchMetrics m = with(h, [&](auto& s) { return s.boundaryPatch(inlet != 0); });
  out->area = m.area;
  out->mass_flux = m.mass_flux;
  out->avg_static = m.avg_static;
  out->avg_dynamic = m.avg_dynamic;
  out->avg_total = m.avg_total;
  out->valid = m.valid;
  return 0;
}

// Apply the velocity + pressure boundary conditions to u, v, p in place.
int orc_simple_apply_bc(void* hp) {
  auto* h = static_cast<Handle*>(hp);
  return with(h, [&](auto& s) {
    s.setVelocityBC(s.u, s.v);
    s.setPressureBC(s.p);
    return 0;
  });
}

// which: 0 = U momentum, 1 = V momentum, 2 = pressure.  Assembles the system
// from the current fields (momentum: u, v, p, gamma; pressure: uStar, vStar,
// dE, dN) without solving it.
int orc_simple_assemble(void* hp, int which) {
  auto* h = static_cast<Handle*>(hp);
  return with(h, [&](auto& s) {
    if (which == 0) s.template assembleMomentum<true>();
    else if (which == 1) s.template assembleMomentum<false>();
    else { s.massResidual(); s.assemblePressure(); }
    return 0;
  });
}

int orc_simple_csr_size(void* hp, int which, int* n, int* nnz) {
  auto* h = static_cast<Handle*>(hp);
  return with(h, [&](auto& s) {
    auto& c = which == 0 ? s.csrU : which == 1 ? s.csrV : s.csrP;
    *n = c.n;
    *nnz = c.row_ptr.empty() ? 0 : c.row_ptr[c.n];
    return 0;
  });
}

// dof_ij: 2*n ints (i, j) of each DOF row.
int orc_simple_csr_get(void* hp, int which, int* row_ptr, int* col_idx, double* values, double* rhs,
                       double* sol, int* dof_ij) {
  auto* h = static_cast<Handle*>(hp);
  return with(h, [&](auto& s) {
    auto& c = which == 0 ? s.csrU : which == 1 ? s.csrV : s.csrP;
    auto& map = which == 0 ? s.mapU : which == 1 ? s.mapV : s.mapP;
    for (int k = 0; k <= c.n; ++k) row_ptr[k] = c.row_ptr[k];
    for (int k = 0; k < c.row_ptr[c.n]; ++k) { col_idx[k] = c.col_idx[k]; values[k] = c.values[k]; }
    for (int k = 0; k < c.n; ++k) {
      rhs[k] = c.rhs[k];
      sol[k] = c.sol[k];
      dof_ij[2 * k] = map[k].first;
      dof_ij[2 * k + 1] = map[k].second;
    }
    return 0;
  });
}

// Reported residuals after the last step or assembly: [residU_RMS,
// residV_RMS, residMass_RMS, residU_max, residV_max, residMass_max].
int orc_simple_residuals(void* hp, double* out6) {
  auto* h = static_cast<Handle*>(hp);
  return with(h, [&](auto& s) {
    out6[0] = s.residU_RMS; out6[1] = s.residV_RMS; out6[2] = s.residMass_RMS;
    out6[3] = s.residU_max; out6[4] = s.residV_max; out6[5] = s.residMass_max;
    return 0;
  });
}

// kind: 0 alphaAtU, 1 alphaAtV, 2 gammaAtU, 3 gammaAtV, 4 southU, 5 northU,
// 6 westV, 7 eastV (flags as 0/1), evaluated on the current fields.
double orc_simple_face(void* hp, int kind, int i, int j) {
  auto* h = static_cast<Handle*>(hp);
  return with(h, [&](auto& s) -> double {
    switch (kind) {
      case 0: return s.alphaAtU(i, j);
      case 1: return s.alphaAtV(i, j);
      case 2: return s.gammaAtU(i, j);
      case 3: return s.gammaAtV(i, j);
      case 4: return s.southU(i, j);
      case 5: return s.northU(i, j);
      case 6: return s.westV(i, j);
      case 7: return s.eastV(i, j);
    }
    return 0.0;
  });
}

double orc_simple_sou(void* hp, int isU, int i, int j, double Fe, double Fw, double Fn, double Fs) {
  auto* h = static_cast<Handle*>(hp);
  return with(h, [&](auto& s) -> double {
    using R = std::remove_reference_t<decltype(s.u[0])>;
    return isU ? s.souU(i, j, R(Fe), R(Fw), R(Fn), R(Fs)) : s.souV(i, j, R(Fe), R(Fw), R(Fn), R(Fs));
  });
}

int orc_jacobi_csr(int prec, int n, const int* rp, const int* ci, const double* val, const double* b, double* x,
                   double tol, int maxit) {
  return prec == 32 ? jacobiT<float>(n, rp, ci, val, b, x, tol, maxit) : jacobiT<double>(n, rp, ci, val, b, x, tol, maxit);
}
int orc_pcg_csr(int prec, int n, const int* rp, const int* ci, const double* val, const double* b, double* x,
                double tol, int maxit) {
  return prec == 32 ? pcgT<float>(n, rp, ci, val, b, x, tol, maxit) : pcgT<double>(n, rp, ci, val, b, x, tol, maxit);
}

// ---------------------------------------------------------------------------
// Thermal (thermal_solver.cpp:422-553): assemble once, solve on demand.
struct ThermalHandle {
  oracle::ThermalProblem P;
  oracle::ThermalCsr A;
};

void* orc_thermal_create(int nx, int ny, int nz, double dx, double dy, const double* dz, const int* channel, double ht,
                         double k_s, double k_f, double rho, double cp, double t_in, int k_mode, double qk,
                         const double* gamma, const double* u, const double* v, const double* q0) {
  auto* h = new ThermalHandle;
  auto& P = h->P;
  P.nx = nx; P.ny = ny; P.nz = nz; P.dx = dx; P.dy = dy;
  P.dz.assign(dz, dz + nz);
  P.channel.assign(channel, channel + nz);
  P.ht = ht; P.k_s = k_s; P.k_f = k_f; P.rho = rho; P.cp = cp; P.t_in = t_in;
  P.k_mode = k_mode; P.qk = qk;
  const int m = nx * ny;
  P.gamma.assign(gamma, gamma + m);
  P.u.assign(u, u + m);
  P.v.assign(v, v + m);
  P.q0.assign(q0, q0 + m);
  h->A = oracle::assembleThermal(P);
  return h;
}

void orc_thermal_destroy(void* hp) { delete static_cast<ThermalHandle*>(hp); }

int orc_thermal_csr_size(void* hp, int* n, int* nnz) {
  auto* h = static_cast<ThermalHandle*>(hp);
  *n = h->A.n;
  *nnz = h->A.row_ptr[h->A.n];
  return 0;
}

int orc_thermal_csr_get(void* hp, int* rp, int* ci, double* val, double* b) {
  auto* h = static_cast<ThermalHandle*>(hp);
  const auto& A = h->A;
  for (int k = 0; k <= A.n; ++k) rp[k] = A.row_ptr[k];
  for (int k = 0; k < A.row_ptr[A.n]; ++k) { ci[k] = A.col_idx[k]; val[k] = A.values[k]; }
  for (int k = 0; k < A.n; ++k) b[k] = A.b[k];
  return 0;
}

// BiCGStab(Jacobi) from the initial guess in T; returns iterations (<0 if
// not converged within maxit).
int orc_thermal_solve(void* hp, double* T, double rtol, int maxit, double* relres, double* seconds) {
  auto* h = static_cast<ThermalHandle*>(hp);
  std::vector<double> x(T, T + h->A.n);
  auto t0 = std::chrono::steady_clock::now();
  int it = oracle::bicgstabCsr(h->A, x, rtol, maxit, relres);
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  for (int k = 0; k < h->A.n; ++k) T[k] = x[k];
  return it;
}

}  // extern "C"
