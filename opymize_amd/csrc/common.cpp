// Error state and shared host helpers of the opymize_amd C-ABI.
#include "opy_common.h"

#include <cstdarg>
#include <cstring>

namespace {
thread_local char g_last_error[1024] = {0};
}

namespace opy {

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    return fail(kHipError, "%s: %s (%d)", what, hipGetErrorString(e), static_cast<int>(e));
  }
  return kOk;
}

int make_grid(const opy_grid_t* g, Grid* out) {
  if (g == nullptr) return fail(kInvalid, "grid descriptor is null");
  if (g->ndims < 1 || g->ndims > kMaxDims)
    return fail(kInvalid, "ndims=%d outside [1,%d]", g->ndims, kMaxDims);
  if (g->nchannels < 1) return fail(kInvalid, "nchannels=%d < 1", g->nchannels);
  out->ndims = g->ndims;
  out->nchan = g->nchannels;
  int64_t n = 1;
  for (int t = 0; t < kMaxDims; ++t) {
    out->dims[t] = 1;
    out->skips[t] = 0;
    out->h[t] = 1.0;
  }
  for (int t = 0; t < g->ndims; ++t) {
    if (g->dims[t] < 1) return fail(kInvalid, "dims[%d]=%lld < 1", t, (long long)g->dims[t]);
    if (!(g->h[t] > 0.0)) return fail(kInvalid, "h[%d]=%g must be positive", t, g->h[t]);
    out->dims[t] = g->dims[t];
    out->h[t] = g->h[t];
    n *= g->dims[t];
  }
  int64_t s = 1;
  for (int t = g->ndims - 1; t >= 0; --t) {
    out->skips[t] = s;
    s *= g->dims[t];
  }
  out->npoints = n;
  return kOk;
}

}  // namespace opy

OPY_API const char* opy_last_error(void) { return g_last_error; }

OPY_API int opy_abi_version(void) { return OPY_ABI_VERSION; }

OPY_API int opy_device_sync(void) {
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return opy::fail(opy::kHipError, "hipDeviceSynchronize: %s", hipGetErrorString(e));
  return opy::kOk;
}
