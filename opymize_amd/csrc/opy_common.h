// Shared definitions for the opymize_amd HIP library (gfx950 / CDNA4).
//
// Every exported entry point returns an int status (0 = ok).  On failure the
// message is kept in a thread-local buffer that opy_last_error() returns.
// Entry points never allocate device memory on the hot path; work buffers are
// passed in by the caller.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#include "../../include/opymize_amd.h"

#define OPY_API extern "C" __attribute__((visibility("default")))

namespace opy {

constexpr int kMaxDims = OPY_MAX_DIMS;
constexpr int kBlock = 256;  // 4 wavefronts of 64 lanes

// status codes shared with the Python wrapper
enum Status : int {
  kOk = 0,
  kInvalid = 1,   // bad argument (shape, pointer, size)
  kHipError = 2,  // launch or runtime failure
};

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int check_launch(const char* what);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

// Grid description passed by value to kernels.  Points are numbered in
// C order over dims[0..ndims-1]; skips[t] is the linear stride of axis t.
struct Grid {
  int ndims;
  int nchan;
  int64_t npoints;
  int64_t dims[kMaxDims];
  int64_t skips[kMaxDims];
  double h[kMaxDims];
};

int make_grid(const opy_grid_t* g, Grid* out);

// FP64 reduction helpers -------------------------------------------------

// Sum over a 256-thread block; result valid in thread 0.
template <typename T>
__device__ inline T block_sum(T v, T* smem) {
  // wavefront reduction with cross-lane shuffles (64 lanes)
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (lane == 0) smem[wave] = v;
  __syncthreads();
  T r = T(0);
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) r += smem[w];
  }
  __syncthreads();
  return r;
}

template <typename T>
__device__ inline T block_max(T v, T* smem) {
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_down(v, off, 64);
    v = v > o ? v : o;
  }
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (lane == 0) smem[wave] = v;
  __syncthreads();
  T r = v;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    r = smem[0];
    for (int w = 1; w < nw; ++w) r = r > smem[w] ? r : smem[w];
  }
  __syncthreads();
  return r;
}

}  // namespace opy
