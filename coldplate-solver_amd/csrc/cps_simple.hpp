// cps_simple.hpp -- device-resident SIMPLE solver (host side).
#pragma once

#include <memory>
#include <vector>

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "cps_common.hpp"

namespace cps {

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int* pinned_i = nullptr;  // small pinned host scratch for flag reads
  // class-batched stepping: launches cover nz fields whose buffers sit zs
  // bytes apart (nz = 1: this solver's own field only)
  int nz = 1;
  size_t zs = 0;
  dim3 zg(dim3 g) const { return dim3(g.x, g.y, (unsigned)nz); }
};

// Scalars shared by kernels of one SIMPLE iteration (device memory).
// Written by block 0 of the kernel that finalises a reduction; read by later
// kernels on the same stream and by the host.
struct IterState {
  int jac_active[2];   // component still iterating
  int jac_iters[2];    // sweeps done
  double rz0;          // r.z of the PCG start
  double rzs[2];       // r.z of PCG iteration it, slot it & 1
  double err;          // ||r||^2 after the last PCG update
  int pcg_stop, pcg_iters;
  double su, sv, umax, vmax;  // momentum change sums / max
  double mass_max, mass_ss;
  double bnorm2;              // ||rhs||^2 of the pressure system
  double resid_u, resid_v, resid_mass, error;
  int ir_outer;               // refinement sweeps done (mixed pressure)
  int pad_;
  double mon[4][5];           // k_sample sums: core in / out planes, inlet / outlet patches
};

// Partial-sum slots (each NB doubles) of the SIMPLE step kernels.
enum PartSlot {
  SLOT_JAC = 0,     // 4 slots: U/V sum r^2, by sweep parity
  SLOT_FIN = 4,     // 4 slots: sum du^2, sum dv^2, max |du|, max |dv|
  SLOT_RHS = 8,     // 4 slots: max |b|, sum b^2 (j <= N-1), r.z, ||b||^2 (DOFs)
  SLOT_PCG = 12,    // 3 slots: pAp, ||r||^2, r.z
  SLOT_SOR = 16,    // 4 slots: max |change| per colour, by iteration parity
  SLOT_JAC4 = 20,   // 8 slots: U/V sum r^2 of the paired sweeps, by sweep index mod 4
  SLOT_COUNT = 28
};

// Geometry + physics scalars passed to kernels by value.
struct Geo {
  int M, N, P;        // padded sizes and pitch (elements)
  int gx, gy, NB;     // capped launch grid of the field kernels, NB = gx*gy
  int nU, nV, nP;     // DOF counts
  double hx, hy, rho, eta, vol;
  double maxVel, convScale, sinkDiag, De, Dn;
  double uvAlpha, pAlpha;
  double targetVel;
  int sou;
  size_t zs;  // byte stride between the fields of a class-batched launch
};

class SimpleSolver;
class PressureMG;

// Kernels whose launches can be timed with HIP events on the solver stream.
enum TimedKernel {
  KT_NONE = -1,
  KT_MG_SMOOTH0 = 0,   // fine-level multigrid Jacobi smoother sweep (pressure)
  KT_MOM_JACOBI = 1,   // fused U+V momentum Jacobi sweep
  KT_PCG_STENCIL = 2,  // fine-level PCG p-update + 5-point apply
};

// Times the launches of one kernel: the selected kernel is dispatched with
// hipExtLaunchKernel, whose start / stop events are written by the dispatch
// itself (kernel execution only, like a profiler's kernel trace); drain()
// sums the pairs.  A kernel that can exit early on a device stop flag
// (KT_MG_SMOOTH0: the fine down kernel once its PCG has converged) sets its
// sample's `ran` slot when it passes its stop tests, and only those samples
// count, so the average is the duration of launches that did the work.
struct KernelTimer {
  int which = KT_NONE;
  hipStream_t stream = nullptr;
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  size_t launches = 0;  // counted samples (that did the work)
  size_t skipped = 0;   // samples that exited on the stop flag
  double total_ms = 0.0;
  double bytes_per_launch = 0.0;
  // one launch in kSample of the selected kernel carries the events, so the
  // timing barely perturbs the run it measures
  static constexpr size_t kSample = 8;
  // ran flags of the samples between two drains (device, zeroed by reset and drain)
  static constexpr size_t kRanSlots = 4096;
  int* ran_dev = nullptr;
  std::vector<int> ran_host;
  size_t seen = 0;
  bool tracks_ran() const { return which == KT_MG_SMOOTH0; }
  bool active(int id) {
    if (which != id || (seen++ % kSample) != 0) return false;
    return !tracks_ran() || (ran_dev && used / 2 < kRanSlots);
  }
  int* ran_slot() { return tracks_ran() && ran_dev ? ran_dev + used / 2 : nullptr; }
  hipEvent_t next_event() {
    if (used == ev.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      ev.push_back(e);
    }
    return ev[used++];
  }
  template <class F, class... Args>
  void ext_launch(F kernel, dim3 grid, dim3 block, unsigned shm, hipStream_t s, Args... args) {
    hipEvent_t a = next_event(), b = next_event();
    hipExtLaunchKernelGGL(kernel, grid, block, shm, s, a, b, 0, args...);
  }
  // adds the elapsed time of the recorded (begin, end) pairs that did the
  // work to total_ms
  void drain() {
    const size_t pairs = used / 2;
    if (used >= 2) (void)hipEventSynchronize(ev[used - 1]);
    const bool rt = tracks_ran() && ran_dev && pairs > 0;
    if (rt) {
      ran_host.assign(pairs, 0);
      if (hipMemcpy(ran_host.data(), ran_dev, pairs * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        ran_host.assign(pairs, 1);
      (void)hipMemset(ran_dev, 0, pairs * sizeof(int));
    }
    for (size_t k = 0; k < pairs; ++k) {
      if (rt && !ran_host[k]) {
        ++skipped;
        continue;
      }
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]) == hipSuccess) {
        total_ms += ms;
        ++launches;
      }
    }
    used = 0;
  }
  void reset(int id) {
    drain();
    which = id;
    launches = 0;
    skipped = 0;
    seen = 0;
    total_ms = 0.0;
    if (tracks_ran() && !ran_dev && hipMalloc((void**)&ran_dev, kRanSlots * sizeof(int)) != hipSuccess) ran_dev = nullptr;
    if (ran_dev) (void)hipMemset(ran_dev, 0, kRanSlots * sizeof(int));
  }
  ~KernelTimer() {
    for (auto e : ev) (void)hipEventDestroy(e);
    if (ran_dev) (void)hipFree(ran_dev);
  }
};

// Device state of a SIMPLE solver as seen by the adjoint (cps_adjoint.hip).
struct SimpleView {
  Geo g;
  size_t S;  // elements per padded field
  cps_simple_config cfg;
  const double *u, *v, *p, *gam, *alp, *inlet, *dE, *dN;
};

class FlowAdjoint;
FlowAdjoint* flow_adjoint_new(Ctx* ctx, const Geo& g, size_t S);
void flow_adjoint_free(FlowAdjoint* a);
void flow_adjoint_run(FlowAdjoint* a, const SimpleView& sv, double rtol, int max_iter, double* dp, double* grad,
                      cps_solve_stats* stats);
void flow_adjoint_residual(FlowAdjoint* a, const SimpleView& sv, double* out3);
void flow_adjoint_apply(FlowAdjoint* a, const SimpleView& sv, int transpose, const double* x, double* y);
void flow_adjoint_vjp(FlowAdjoint* a, const SimpleView& sv, const double* du, const double* dv, const double* dp,
                      double rtol, int max_iter, double* grad, cps_solve_stats* stats);
void flow_adjoint_options(FlowAdjoint* a, int kind, int sweeps, double omega, int schur_iters);

}  // namespace cps

namespace cps {
void csr_free(void* p);
// looks up the device address of the allocation-free drop-in's scratch on
// `device` (current) once per process (cps_csr.hip)
void csr_scratch_init(int device);
}

struct cps_ctx {
  cps::Ctx c;
  void* csr = nullptr;  // CSR drop-in work buffers (cps_csr.hip)
};
