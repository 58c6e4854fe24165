// cps_capi.hip -- context lifetime and error reporting of the C ABI.
#include <cstring>
#include <string>

#include "cps_simple.hpp"

namespace cps {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

}  // namespace cps

extern "C" {

const char* cps_last_error(void) { return cps::g_last_error.c_str(); }

int cps_abi_version(void) { return 2; }

int cps_create(int device, cps_ctx** out) {
  return cps::guard([&] {
    CPS_CHECK(out != nullptr, CPS_ERR_ARG, "out is null");
    int n = 0;
    CPS_HIP(hipGetDeviceCount(&n));
    CPS_CHECK(n > 0, CPS_ERR_HIP, "no HIP device");
    CPS_CHECK(device >= 0 && device < n, CPS_ERR_ARG, "device index out of range");
    CPS_HIP(hipSetDevice(device));
    cps::csr_scratch_init(device);
    auto* c = new cps_ctx;
    c->c.device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      throw cps::Error{CPS_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)};
    }
    *out = c;
  });
}

void cps_destroy(cps_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  if (ctx->c.stream) {
    (void)hipStreamSynchronize(ctx->c.stream);
    (void)hipStreamDestroy(ctx->c.stream);
  }
  cps::csr_free(ctx->csr);
  delete ctx;
}

int cps_synchronize(cps_ctx* ctx) {
  return cps::guard([&] {
    CPS_CHECK(ctx != nullptr, CPS_ERR_ARG, "ctx is null");
    CPS_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

}  // extern "C"
