// An entry point that allocates nothing, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops): the
// volume-weighted residual norm of a device vector, with the reduction's
// scratch (first-stage partial sums and the two totals) in a file-scope
// device array instead of the context's allocated `red` / `scal` buffers.
//
// `scratch` below is a host-side stand-in in host code: its device address
// differs per GPU and is taken once per device by noalloc_init (called from
// cfd_ctx_create with that device current) through hipGetSymbolAddress.
// Host code only ever passes that address, never the array's name.
#include <mutex>

#include "ctx.hpp"

namespace {

// 2 x RED_BLOCKS first-stage partials, then the two sums (16.1 KiB per GPU,
// held until the process ends).  Every partial is written by its block on
// each call before it is read, so nothing needs clearing between calls.
constexpr int kScratchDoubles = 2 * cfd::RED_BLOCKS + 2;
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_mu;

}  // namespace

namespace cfd {

int noalloc_init(cfd_ctx* c) {
  if (c->device < 0 || c->device >= kMaxDevices) {
    c->err = "device ordinal beyond the library's per-device scratch table";
    return CFD_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[c->device]) return CFD_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    c->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return CFD_ERR_HIP;
  }
  g_scratch[c->device] = static_cast<double*>(p);
  return CFD_OK;
}

}  // namespace cfd

extern "C" int cfd_dev_norm_vw_noalloc(cfd_ctx* c, const double* R, double* out) {
  if (!c || !R || !out) return CFD_ERR_ARG;
  // the reduction kernel indexes the vector with 32-bit integers
  if ((long long)c->P.nvar * c->pitch > 2147483647LL) {
    c->err = "cfd_dev_norm_vw_noalloc: nvar x pitch exceeds 2^31 - 1 elements";
    return CFD_ERR_ARG;
  }
  double* s = nullptr;
  if (c->device >= 0 && c->device < kMaxDevices) {
    std::lock_guard<std::mutex> lock(g_mu);
    s = g_scratch[c->device];
  }
  if (!s) {
    c->err = "cfd_dev_norm_vw_noalloc: no scratch address for the context's device";
    return CFD_ERR_STATE;
  }
  cfd::launch_norm_vw_into(c, R, s, s + 2 * cfd::RED_BLOCKS, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    c->err = std::string("cfd_dev_norm_vw_noalloc: ") + hipGetErrorString(e);
    return CFD_ERR_HIP;
  }
  return CFD_OK;
}
