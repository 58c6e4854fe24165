// cps_mg.hpp -- geometric-multigrid preconditioned CG for the SIMPLE
// pressure-correction system (replaces the AMGCL solve of
// pressure_solver.cpp:747-766 / pressure_solver.cu:81-163).
#pragma once

#include <memory>
#include <vector>

#include "cps_mg_coarse.hpp"
#include "cps_simple.hpp"

namespace cps {

class PressureMG {
 public:
  PressureMG(Ctx* ctx, const Geo& g);
  ~PressureMG();
  // Solves A x = b for the pressure correction on the fine grid (field
  // layout, FP64).  r holds b on entry and the final FP64 residual on exit
  // (mixed) ; x is zero on entry.  FP64: MG-PCG in double.  Mixed: FP64
  // iterative refinement around an FP32 MG-PCG.  same_operator: dE, dN are
  // unchanged since the previous solve in this precision, so the hierarchy
  // built then is reused.
  void solve(const double* dE, const double* dN, double* r, double* x, IterState* st, const cps_simple_config& cfg,
             bool fp64, KernelTimer* timer, bool same_operator = false);
  // Builds the coming solve's hierarchy ahead of it, once dE and dN are
  // final.  prebuild(j) hands the grid levels' faces to the kernel that
  // forms dE, dN (the assembly, on the level-0 tiles; false where deeper
  // levels need launches of their own), prebuild_launch() launches them;
  // coarse_job() then hands the coarse solver's set-up to a later kernel
  // (one extra block beside that kernel's own work).  The next solve in
  // that precision builds nothing.  coarse_job() returns false when no
  // prebuilt hierarchy is pending.
  bool prebuild(mgc::HierJob<double>* j);
  bool prebuild(mgc::HierJob<float>* j);
  void prebuild_launch(const double* dE, const double* dN, bool fp64);
  bool coarse_job(mgc::CoarseJob<double>* j, size_t* smem);
  bool coarse_job(mgc::CoarseJob<float>* j, size_t* smem);
  // The start of the coming solve (r = b, x = 0, flags cleared, FP64 b
  // kept), for the kernel that forms b; j->on = 0 where the solve must start
  // itself (host polling).  The next solve in that precision then skips its
  // start kernel.
  void start_job(const cps_simple_config& cfg, mgc::StartJob<double>* j);
  void start_job(const cps_simple_config& cfg, mgc::StartJob<float>* j);
  // Mixed solves with one refinement sweep leave their closing check (the
  // pressure statistics in IterState) until the host reads the statistics:
  // no later kernel of the step reads them.  Launches it if one is pending.
  void finish(IterState* st);
  // A replayed graph of captured steps runs none of the host code that
  // leaves the check pending: after replays, rearm() restores the pending
  // check as the captured solve (`captured`, possibly another field's solver
  // of the same size and settings) left it.
  void rearm(const PressureMG& captured) { pending_ = captured.armed_; }

 private:
  Ctx* ctx_;
  Geo g_;
  struct PendingCheck {
    bool on = false;
    double abs_tol = 0.0, rtol = 0.0;
    int endK = 0;
  } pending_, armed_;
  void* impl64_ = nullptr;  // MGImpl<double>
  void* impl32_ = nullptr;  // MGImpl<float>
  void* b64_ = nullptr;     // FP64 rhs kept across refinement sweeps
  double* b64();            // allocated on first use
};

}  // namespace cps
