// cps_csr.hip -- CSR drop-ins for the reference GPU backend ABI
// (cuda_solver.hpp:24-57): host CSR in, solution in/out, iterations out.
//
//   cps_pressure_solve_csr   Jacobi-PCG with the reference CPU semantics of
//                            solvePCG_CSR (pressure_solver.cpp:121-184)
//   cps_momentum_jacobi_csr  point Jacobi with the CPU stop rule of
//                            solveMomentumJacobi (jacobi_solver.cpp:20-75)
//
// Vector arithmetic is FP32 like the reference; dot products and residual
// sums accumulate in FP64 (wave shuffle + fixed-order block partials).
#include <mutex>
#include <vector>

#include "cps_simple.hpp"

namespace cps {

constexpr int CT = 256;

struct CsrDev {
  int n;
  const int *rp, *ci;
  const float* val;
};

struct CsrState {
  double rzs[2], err, rsq;
  int stop, iters, active, pad_;
};

// partial slots (each nb doubles): r.z start, pAp, ||r||^2, r.z, Jacobi parity 0/1
enum CsrSlot { C_RZ0 = 0, C_PAP = 1, C_ERR = 2, C_RZ = 3, C_JAC = 4, C_COUNT = 6 };

__device__ inline float csr_row_dot(const CsrDev& A, int i, const float* x) {
  float s = 0.0f;
  for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) s += A.val[k] * x[A.ci[k]];
  return s;
}

#define ROW_LOOP(A, i) for (int i = blockIdx.x * CT + threadIdx.x; i < (A).n; i += gridDim.x * CT)

// diag (1 if absent), r = b - A x, z = r/diag, p_old = 0, partial r.z
__global__ __launch_bounds__(CT) void k_csr_pcg_init(CsrDev A, const float* b, const float* x, float* diag, float* r,
                                                    float* z, float* pa, float* pb, CsrState* st, double* part) {
  __shared__ double sh[CT / kWave + 1];
  double rz = 0.0;
  ROW_LOOP(A, i) {
    float d = 1.0f;
    for (int k = A.rp[i]; k < A.rp[i + 1]; ++k)
      if (A.ci[k] == i) { d = A.val[k]; break; }
    diag[i] = d;
    const float ri = b[i] - csr_row_dot(A, i, x);
    r[i] = ri;
    const float zi = (fabsf(d) > 1e-20f) ? ri / d : ri;
    z[i] = zi;
    pa[i] = 0.0f;
    pb[i] = 0.0f;
    rz += (double)ri * (double)zi;
  }
  const double s = block_sum<CT>(rz, sh);
  if (threadIdx.x == 0) part[(size_t)C_RZ0 * gridDim.x + blockIdx.x] = s;
  if (blockIdx.x == 0 && threadIdx.x == 0) { st->stop = 0; st->iters = 0; }
}

// stop test of iteration it-1, beta, p = z + beta p_old (gathered per
// column), Ap, partial pAp
__global__ __launch_bounds__(CT) void k_csr_pcg_a(CsrDev A, const float* z, float* pa, float* pb, float* Ap, int it,
                                                 float tol, CsrState* st, double* part) {
  __shared__ double sh[CT / kWave + 1];
  __shared__ int shi;
  if (block_read(&st->stop, &shi)) return;
  const int nb = gridDim.x;
  const bool b0 = blockIdx.x == 0 && threadIdx.x == 0;
  double rzc, beta = 0.0;
  if (it == 0) {
    rzc = parts_sum<CT>(part + (size_t)C_RZ0 * nb, nb, sh);
  } else {
    const double err = parts_sum<CT>(part + (size_t)C_ERR * nb, nb, sh);
    if (sqrt(err) < (double)tol) {  // pressure_solver.cpp:171
      if (b0) { st->stop = 1; st->iters = it - 1; st->err = err; }
      return;
    }
    rzc = parts_sum<CT>(part + (size_t)C_RZ * nb, nb, sh);
    beta = rzc / st->rzs[(it - 1) & 1];
  }
  if (b0) { st->rzs[it & 1] = rzc; st->iters = it; }
  const float fb = (float)beta;
  const float* po = (it & 1) ? pb : pa;
  float* pn = (it & 1) ? pa : pb;
  double pap = 0.0;
  ROW_LOOP(A, i) {
    float s = 0.0f;
    for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
      const int c = A.ci[k];
      s += A.val[k] * (z[c] + fb * po[c]);
    }
    const float pc = z[i] + fb * po[i];
    pn[i] = pc;
    Ap[i] = s;
    pap += (double)pc * (double)s;
  }
  const double s = block_sum<CT>(pap, sh);
  if (threadIdx.x == 0) part[(size_t)C_PAP * nb + blockIdx.x] = s;
}

__global__ __launch_bounds__(CT) void k_csr_pcg_b(CsrDev A, float* x, float* r, float* z, const float* diag,
                                                 const float* pa, const float* pb, const float* Ap, int it,
                                                 CsrState* st, double* part) {
  __shared__ double sh[CT / kWave + 1];
  __shared__ int shi;
  if (block_read(&st->stop, &shi)) return;
  const int nb = gridDim.x;
  const double pap = parts_sum<CT>(part + (size_t)C_PAP * nb, nb, sh);
  if (fabs(pap) < (double)1e-20f) {  // pressure_solver.cpp:158
    if (blockIdx.x == 0 && threadIdx.x == 0) { st->stop = 1; st->iters = it; }
    return;
  }
  const float alpha = (float)(st->rzs[it & 1] / pap);
  const float* pn = (it & 1) ? pa : pb;
  double e = 0.0, rz = 0.0;
  ROW_LOOP(A, i) {
    x[i] += alpha * pn[i];
    const float ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    const float d = diag[i];
    const float zi = (fabsf(d) > 1e-20f) ? ri / d : ri;
    z[i] = zi;
    e += (double)ri * (double)ri;
    rz += (double)ri * (double)zi;
  }
  const double a = block_sum<CT>(e, sh), b = block_sum<CT>(rz, sh);
  if (threadIdx.x == 0) {
    part[(size_t)C_ERR * nb + blockIdx.x] = a;
    part[(size_t)C_RZ * nb + blockIdx.x] = b;
  }
}

// after K launched iterations (one block; nb = partial count)
__global__ __launch_bounds__(CT) void k_csr_pcg_end(int nb, int K, float tol, CsrState* st, const double* part) {
  __shared__ double sh[CT / kWave + 1];
  __shared__ int shi;
  if (block_read(&st->stop, &shi)) return;
  const double err = parts_sum<CT>(part + (size_t)C_ERR * nb, nb, sh);
  if (threadIdx.x == 0) {
    st->err = err;
    if (sqrt(err) < (double)tol) { st->stop = 1; st->iters = K - 1; }
    else st->iters = K;
  }
}

// Jacobi stop rule (jacobi_solver.cpp:63) on the partials of sweep k-1; the
// reference compares a float sum with tol^2 n in float
__device__ inline int csr_jac_detect(int nb, int k, float tol2n, CsrState* st, const double* part, int active,
                                     double* sh, bool b0) {
  if (!active || k <= 0) return active;
  const double a = parts_sum<CT>(part + (size_t)(C_JAC + ((k - 1) & 1)) * nb, nb, sh);
  if ((float)a < tol2n) {
    if (b0) { st->active = 0; st->iters = k; st->rsq = a; }
    return 0;
  }
  return 1;
}

__global__ __launch_bounds__(CT) void k_csr_jacobi(CsrDev A, const float* b, const float* x, float* y, int sweep,
                                                  float tol2n, CsrState* st, double* part) {
  __shared__ double sh[CT / kWave + 1];
  __shared__ int shi;
  const int nb = gridDim.x;
  int active = block_read(&st->active, &shi);
  active = csr_jac_detect(nb, sweep, tol2n, st, part, active, sh, blockIdx.x == 0 && threadIdx.x == 0);
  if (!active) return;
  double rr = 0.0;
  ROW_LOOP(A, i) {
    float ax = 0.0f, d = 1.0f;
    for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
      const int c = A.ci[k];
      ax += A.val[k] * x[c];
      if (c == i) d = A.val[k];
    }
    const float r = b[i] - ax;
    rr += (double)r * (double)r;
    y[i] = x[i] + r / d;
  }
  const double s = block_sum<CT>(rr, sh);
  if (threadIdx.x == 0) part[(size_t)(C_JAC + (sweep & 1)) * nb + blockIdx.x] = s;
}

__global__ __launch_bounds__(CT) void k_csr_jac_end(int nb, int K, float tol2n, CsrState* st, const double* part) {
  __shared__ double sh[CT / kWave + 1];
  __shared__ int shi;
  int active = block_read(&st->active, &shi);
  active = csr_jac_detect(nb, K, tol2n, st, part, active, sh, threadIdx.x == 0);
  if (threadIdx.x == 0 && active) st->iters = K;
}

__global__ void k_csr_state(CsrState* st, int active) {
  if (threadIdx.x == 0) {
    st->active = active;
    st->iters = 0;
    st->stop = 0;
  }
}

static int blocks(int n) { return std::max(1, std::min(ceil_div(n, CT), 1024)); }

struct CsrWork {
  DevBuf<int> rp, ci;
  DevBuf<float> val, b, x, y, r, z, pa, pb, Ap, diag;
  DevBuf<double> part;
  DevBuf<CsrState> st;
  void ensure(int n, int nnz) {
    if ((int)rp.n < n + 1) rp.alloc(n + 1);
    if ((int)ci.n < nnz) ci.alloc(nnz);
    if ((int)val.n < nnz) val.alloc(nnz);
    for (auto* q : {&b, &x, &y, &r, &z, &pa, &pb, &Ap, &diag})
      if ((int)q->n < n) q->alloc(n);
    const int nb = blocks(n);
    if ((int)part.n < C_COUNT * nb) part.alloc((size_t)C_COUNT * nb);
    if (!st.p) st.alloc(1);
  }
};

static void upload(hipStream_t s, CsrWork& w, int n, const int* rp, const int* ci, const float* val, const float* b,
                   const float* x) {
  const int nnz = rp[n];
  CPS_CHECK(nnz >= 0 && rp[0] == 0, CPS_ERR_ARG, "row_ptr must start at 0 and be non-decreasing");
  w.ensure(n, nnz > 0 ? nnz : 1);
  CPS_HIP(hipMemcpyAsync(w.rp.p, rp, (n + 1) * sizeof(int), hipMemcpyHostToDevice, s));
  if (nnz > 0) {
    CPS_HIP(hipMemcpyAsync(w.ci.p, ci, nnz * sizeof(int), hipMemcpyHostToDevice, s));
    CPS_HIP(hipMemcpyAsync(w.val.p, val, nnz * sizeof(float), hipMemcpyHostToDevice, s));
  }
  CPS_HIP(hipMemcpyAsync(w.b.p, b, n * sizeof(float), hipMemcpyHostToDevice, s));
  CPS_HIP(hipMemcpyAsync(w.x.p, x, n * sizeof(float), hipMemcpyHostToDevice, s));
}

static void validate_csr(int n, const int* rp, const int* ci) {
  for (int i = 0; i < n; ++i) CPS_CHECK(rp[i + 1] >= rp[i], CPS_ERR_ARG, "row_ptr must be non-decreasing");
  for (int k = 0; k < rp[n]; ++k) CPS_CHECK(ci[k] >= 0 && ci[k] < n, CPS_ERR_ARG, "col_idx out of range");
}

static CsrWork& work(cps_ctx* ctx) {
  if (!ctx->csr) ctx->csr = new CsrWork;
  return *static_cast<CsrWork*>(ctx->csr);
}

void csr_free(void* p) { delete static_cast<CsrWork*>(p); }

// Device buffers of one Jacobi-PCG solve.
struct PcgView {
  int *rp, *ci;
  float *val, *b, *x, *r, *z, *pa, *pb, *Ap, *diag;
  double* part;
  CsrState* st;
};

static void upload_view(hipStream_t s, const PcgView& w, int n, const int* rp, const int* ci, const float* val,
                        const float* b, const float* x) {
  const int nnz = rp[n];
  CPS_HIP(hipMemcpyAsync(w.rp, rp, (n + 1) * sizeof(int), hipMemcpyHostToDevice, s));
  if (nnz > 0) {
    CPS_HIP(hipMemcpyAsync(w.ci, ci, nnz * sizeof(int), hipMemcpyHostToDevice, s));
    CPS_HIP(hipMemcpyAsync(w.val, val, nnz * sizeof(float), hipMemcpyHostToDevice, s));
  }
  CPS_HIP(hipMemcpyAsync(w.b, b, n * sizeof(float), hipMemcpyHostToDevice, s));
  CPS_HIP(hipMemcpyAsync(w.x, x, n * sizeof(float), hipMemcpyHostToDevice, s));
}

// solvePCG_CSR (pressure_solver.cpp:121-184) on uploaded buffers: the start
// kernel writes every vector and the state the loop reads, then iterations
// run in chunks of 16 with one stop-flag read per chunk.
static void pcg_run(hipStream_t s, const PcgView& w, int n, float tolerance, int max_iterations, float* solution,
                    int* iterations) {
  CsrDev A{n, w.rp, w.ci, w.val};
  const int nb = blocks(n);
  k_csr_pcg_init<<<nb, CT, 0, s>>>(A, w.b, w.x, w.diag, w.r, w.z, w.pa, w.pb, w.st, w.part);
  CPS_HIP(hipGetLastError());
  const int chunk = 16;
  for (int k0 = 0; k0 < max_iterations; k0 += chunk) {
    const int k1 = std::min(max_iterations, k0 + chunk);
    for (int k = k0; k < k1; ++k) {
      k_csr_pcg_a<<<nb, CT, 0, s>>>(A, w.z, w.pa, w.pb, w.Ap, k, tolerance, w.st, w.part);
      k_csr_pcg_b<<<nb, CT, 0, s>>>(A, w.x, w.r, w.z, w.diag, w.pa, w.pb, w.Ap, k, w.st, w.part);
    }
    k_csr_pcg_end<<<1, CT, 0, s>>>(nb, k1, tolerance, w.st, w.part);
    CPS_HIP(hipGetLastError());
    if (k1 >= max_iterations) break;
    int stop = 0;
    CPS_HIP(hipMemcpyAsync(&stop, &w.st->stop, sizeof(int), hipMemcpyDeviceToHost, s));
    CPS_HIP(hipStreamSynchronize(s));
    if (stop) break;
  }
  CsrState h{};
  CPS_HIP(hipMemcpyAsync(&h, w.st, sizeof(CsrState), hipMemcpyDeviceToHost, s));
  CPS_HIP(hipMemcpyAsync(solution, w.x, n * sizeof(float), hipMemcpyDeviceToHost, s));
  CPS_HIP(hipStreamSynchronize(s));
  if (iterations) *iterations = h.iters;
}

// ---------------------------------------------------------------------------
// Buffers of cps_pressure_solve_csr_noalloc: one file-scope device array,
// the same on every call on a GPU (each GPU holds its own copy until the
// process ends).  Layout, each piece 256-B aligned: row_ptr, col_idx,
// values, then the vectors b, x, r, z, p (two), A p and the diagonal, the
// block partials and the solver state.  Limits: 2^19 rows (the pressure
// system of a 1024 x 512 grid has 523,776) and 5 * 2^19 nonzeros.
constexpr int kNoAllocMaxRows = 1 << 19;
constexpr long long kNoAllocMaxNnz = 5LL << 19;
constexpr int kNoAllocMaxBlocks = 1024;  // blocks(n) never exceeds it
constexpr size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kScrCi = al256((size_t)(kNoAllocMaxRows + 1) * sizeof(int));
constexpr size_t kScrVal = kScrCi + al256((size_t)kNoAllocMaxNnz * sizeof(int));
constexpr size_t kScrVec = kScrVal + al256((size_t)kNoAllocMaxNnz * sizeof(float));
constexpr size_t kScrVecBytes = al256((size_t)kNoAllocMaxRows * sizeof(float));
constexpr size_t kScrPart = kScrVec + 8 * kScrVecBytes;
constexpr size_t kScrState = kScrPart + al256((size_t)C_COUNT * kNoAllocMaxBlocks * sizeof(double));
constexpr size_t kScratchBytes = kScrState + al256(sizeof(CsrState));
constexpr int kScratchDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each GPU, looked up once per device by
// cps_create (with that device current).  The host-side name `scratch` is a
// stand-in, not a device address: only these pointers reach kernels and
// copies.
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kScratchDevices] = {};

void csr_scratch_init(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  CPS_CHECK(device >= 0 && device < kScratchDevices, CPS_ERR_ARG, "device index beyond the scratch table");
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  CPS_CHECK(e == hipSuccess && p != nullptr, CPS_ERR_HIP,
            std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = static_cast<unsigned char*>(p);
}

static PcgView scratch_view(int device) {
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = device >= 0 && device < kScratchDevices ? g_scratch_ptr[device] : nullptr;
  }
  CPS_CHECK(base != nullptr, CPS_ERR_STATE, "scratch of this device was not set up (cps_create)");
  PcgView w;
  w.rp = reinterpret_cast<int*>(base);
  w.ci = reinterpret_cast<int*>(base + kScrCi);
  w.val = reinterpret_cast<float*>(base + kScrVal);
  float** slots[8] = {&w.b, &w.x, &w.r, &w.z, &w.pa, &w.pb, &w.Ap, &w.diag};
  for (int k = 0; k < 8; ++k) *slots[k] = reinterpret_cast<float*>(base + kScrVec + (size_t)k * kScrVecBytes);
  w.part = reinterpret_cast<double*>(base + kScrPart);
  w.st = reinterpret_cast<CsrState*>(base + kScrState);
  return w;
}

}  // namespace cps

extern "C" {

int cps_pressure_solve_csr(cps_ctx* ctx, int n, const int* row_ptr, const int* col_idx, const float* values,
                           const float* rhs, float* solution, float tolerance, int max_iterations, int* iterations) {
  using namespace cps;
  return guard([&] {
    CPS_CHECK(ctx && row_ptr && rhs && solution, CPS_ERR_ARG, "null argument");
    CPS_CHECK(n >= 0 && max_iterations >= 0, CPS_ERR_ARG, "negative size");
    if (iterations) *iterations = 0;
    if (n == 0) return;  // pressure_solver.cpp:124
    CPS_CHECK(row_ptr[n] == 0 || (col_idx && values), CPS_ERR_ARG, "null CSR arrays");
    validate_csr(n, row_ptr, col_idx);
    CPS_HIP(hipSetDevice(ctx->c.device));
    hipStream_t s = ctx->c.stream;
    CsrWork& w = work(ctx);
    upload(s, w, n, row_ptr, col_idx, values, rhs, solution);
    const PcgView v{w.rp.p, w.ci.p, w.val.p, w.b.p, w.x.p, w.r.p, w.z.p, w.pa.p, w.pb.p, w.Ap.p, w.diag.p,
                    w.part.p, w.st.p};
    pcg_run(s, v, n, tolerance, max_iterations, solution, iterations);
  });
}

int cps_pressure_solve_csr_noalloc(cps_ctx* ctx, int n, const int* row_ptr, const int* col_idx,
                                   const float* values, const float* rhs, float* solution, float tolerance,
                                   int max_iterations, int* iterations) {
  using namespace cps;
  return guard([&] {
    CPS_CHECK(ctx && row_ptr && rhs && solution, CPS_ERR_ARG, "null argument");
    CPS_CHECK(n >= 0 && max_iterations >= 0, CPS_ERR_ARG, "negative size");
    if (iterations) *iterations = 0;
    if (n == 0) return;  // pressure_solver.cpp:124
    CPS_CHECK(n <= kNoAllocMaxRows, CPS_ERR_ARG,
              "cps_pressure_solve_csr_noalloc: n exceeds its scratch limit of 524288 rows");
    CPS_CHECK(row_ptr[0] == 0 && row_ptr[n] >= 0, CPS_ERR_ARG, "row_ptr must start at 0 and be non-decreasing");
    CPS_CHECK((long long)row_ptr[n] <= kNoAllocMaxNnz, CPS_ERR_ARG,
              "cps_pressure_solve_csr_noalloc: nnz exceeds its scratch limit of 2621440 entries");
    CPS_CHECK(row_ptr[n] == 0 || (col_idx && values), CPS_ERR_ARG, "null CSR arrays");
    validate_csr(n, row_ptr, col_idx);
    CPS_HIP(hipSetDevice(ctx->c.device));
    hipStream_t s = ctx->c.stream;
    const PcgView v = scratch_view(ctx->c.device);
    upload_view(s, v, n, row_ptr, col_idx, values, rhs, solution);
    pcg_run(s, v, n, tolerance, max_iterations, solution, iterations);
  });
}

int cps_momentum_jacobi_csr(cps_ctx* ctx, int n, const int* row_ptr, const int* col_idx, const float* values,
                            const float* rhs, float* solution, float tolerance, int max_iterations,
                            int* iterations) {
  using namespace cps;
  return guard([&] {
    CPS_CHECK(ctx && row_ptr && rhs && solution, CPS_ERR_ARG, "null argument");
    CPS_CHECK(n >= 0 && max_iterations >= 0, CPS_ERR_ARG, "negative size");
    if (iterations) *iterations = 0;
    if (n == 0) return;  // jacobi_solver.cpp:31
    CPS_CHECK(row_ptr[n] == 0 || (col_idx && values), CPS_ERR_ARG, "null CSR arrays");
    validate_csr(n, row_ptr, col_idx);
    CPS_HIP(hipSetDevice(ctx->c.device));
    hipStream_t s = ctx->c.stream;
    CsrWork& w = work(ctx);
    upload(s, w, n, row_ptr, col_idx, values, rhs, solution);
    CsrDev A{n, w.rp.p, w.ci.p, w.val.p};
    const int nb = blocks(n);
    const float tol2n = tolerance * tolerance * (float)n;
    k_csr_state<<<1, 64, 0, s>>>(w.st.p, 1);
    const int chunk = 8;
    for (int k0 = 0; k0 < max_iterations; k0 += chunk) {
      const int k1 = std::min(max_iterations, k0 + chunk);
      for (int k = k0; k < k1; ++k) {
        const float* xin = (k & 1) ? w.y.p : w.x.p;
        float* xout = (k & 1) ? w.x.p : w.y.p;
        k_csr_jacobi<<<nb, CT, 0, s>>>(A, w.b.p, xin, xout, k, tol2n, w.st.p, w.part.p);
      }
      k_csr_jac_end<<<1, CT, 0, s>>>(nb, k1, tol2n, w.st.p, w.part.p);
      CPS_HIP(hipGetLastError());
      if (k1 >= max_iterations) break;
      int active = 1;
      CPS_HIP(hipMemcpyAsync(&active, &w.st.p->active, sizeof(int), hipMemcpyDeviceToHost, s));
      CPS_HIP(hipStreamSynchronize(s));
      if (!active) break;
    }
    CsrState h{};
    CPS_HIP(hipMemcpyAsync(&h, w.st.p, sizeof(CsrState), hipMemcpyDeviceToHost, s));
    CPS_HIP(hipStreamSynchronize(s));
    const float* res = (h.iters & 1) ? w.y.p : w.x.p;
    CPS_HIP(hipMemcpyAsync(solution, res, n * sizeof(float), hipMemcpyDeviceToHost, s));
    CPS_HIP(hipStreamSynchronize(s));
    if (iterations) *iterations = h.iters;
  });
}

}  // extern "C"
