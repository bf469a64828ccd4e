// Matrix-free finite-difference operators on a Cartesian grid:
// gradient (forward / centered scheme), its adjoint (negative divergence) and
// the Laplacian with Neumann / curvature / second-order boundary conditions.
//
// Layouts follow the reference's Variables: an image is (npoints, C) and a
// vector field is (npoints, D, C), both flat in C order
// (reference: opymize/linear/diff.py:117-118, diff.cu:74).
#include "opy_common.h"

namespace opy {
namespace {

__device__ inline void point_coords(const Grid& g, int64_t i, int64_t* c) {
  for (int t = 0; t < g.ndims; ++t) {
    c[t] = i / g.skips[t];
    i -= c[t] * g.skips[t];
  }
}

// Offset of the m-th corner of the unit cell spanned by all axes except t.
// Bit j of m selects +1 along the j-th such axis (axes in increasing order).
__device__ inline int64_t cell_offset(const Grid& g, int t, int m, int64_t* bits) {
  int64_t off = 0;
  int j = 0;
  for (int a = 0; a < g.ndims; ++a) {
    bits[a] = 0;
    if (a == t) continue;
    bits[a] = (m >> j) & 1;
    off += bits[a] * g.skips[a];
    ++j;
  }
  return off;
}

// y (N,D,C) (+)= grad x, x (N,C); one thread per (point, channel).
// forward: reference diff.cu:64-71; centered: diff.cu:53-63.
template <typename T, bool kCentered>
__global__ void __launch_bounds__(kBlock) grad_kernel(const T* __restrict__ x, T* __restrict__ y,
                                                      Grid g, const double* __restrict__ w, int add) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int C = g.nchan;
  if (idx >= g.npoints * C) return;
  const int64_t i = idx / C;
  const int k = (int)(idx - i * C);
  const int D = g.ndims;
  int64_t c[kMaxDims];
  point_coords(g, i, c);
  const double wk = w ? w[k] : 1.0;
  bool br = false;
  if (kCentered) {
    for (int a = 0; a < D; ++a) br |= (c[a] == g.dims[a] - 1);
  }
  const int navg = 1 << (D - 1);
  for (int t = 0; t < D; ++t) {
    T val = T(0);
    if (kCentered) {
      const T fac = (T)(wk / (g.h[t] * (double)navg));
      if (!br) {
        T acc = T(0);
        int64_t bits[kMaxDims];
        for (int m = 0; m < navg; ++m) {
          const int64_t base = i + cell_offset(g, t, m, bits);
          acc += x[(base + g.skips[t]) * C + k] - x[base * C + k];
        }
        val = fac * acc;
      }
    } else {
      const T fac = (T)(wk / g.h[t]);
      if (c[t] < g.dims[t] - 1) val = fac * (x[(i + g.skips[t]) * C + k] - x[i * C + k]);
    }
    T* out = y + (i * D + t) * C + k;
    *out = add ? *out + val : val;
  }
}

// x (N,C) (+)= -div p, p (N,D,C): the adjoint of grad_kernel.
// forward: reference diff.cu:114-123; centered: diff.cu:96-113.
template <typename T, bool kCentered>
__global__ void __launch_bounds__(kBlock) div_kernel(const T* __restrict__ p, T* __restrict__ x,
                                                     Grid g, const double* __restrict__ w, int add) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int C = g.nchan;
  if (idx >= g.npoints * C) return;
  const int64_t i = idx / C;
  const int k = (int)(idx - i * C);
  const int D = g.ndims;
  int64_t c[kMaxDims];
  point_coords(g, i, c);
  const double wk = w ? w[k] : 1.0;
  T acc = T(0);
  T fac;
  if (kCentered) {
    const int navg = 1 << (D - 1);
    fac = (T)(wk / (double)navg);
    int64_t bits[kMaxDims];
    for (int t = 0; t < D; ++t) {
      const T ih = (T)(1.0 / g.h[t]);
      for (int m = 0; m < navg; ++m) {
        const int64_t off = cell_offset(g, t, m, bits);
        bool ok = true;
        for (int a = 0; a < D; ++a) {
          if (a == t) continue;
          const int64_t ca = c[a] - bits[a];
          ok &= (ca >= 0) && (ca != g.dims[a] - 1);
        }
        if (!ok) continue;
        const int64_t base = i - off;
        if (c[t] < g.dims[t] - 1) acc -= p[(base * D + t) * C + k] * ih;
        if (c[t] > 0) acc += p[((base - g.skips[t]) * D + t) * C + k] * ih;
      }
    }
  } else {
    fac = (T)wk;
    for (int t = 0; t < D; ++t) {
      const T ih = (T)(1.0 / g.h[t]);
      if (c[t] < g.dims[t] - 1) acc -= p[(i * D + t) * C + k] * ih;
      if (c[t] > 0) acc += p[((i - g.skips[t]) * D + t) * C + k] * ih;
    }
  }
  const T val = fac * acc;
  T* out = x + i * C + k;
  *out = add ? *out + val : val;
}

// The forward gradient and its adjoint on a 2-D grid (H x W, C channels) with
// one thread per (column, channel) of row blockIdx.y: the arithmetic of
// grad_kernel<T, false> / div_kernel<T, false> in the same order, without
// the per-element coordinate divisions of the n-D kernels.  CT > 0 fixes C;
// for C = 1 the pair (d/drow, d/dcol) of a point is one 2-vector access.
template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

template <typename T, int CT>
__global__ void __launch_bounds__(kBlock) grad2f_kernel(const T* __restrict__ x, T* __restrict__ y, int H, int W,
                                                        int Cr, const double* __restrict__ w, double h0, double h1,
                                                        int add) {
  const int C = CT > 0 ? CT : Cr;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= W * C) return;
  const int r = blockIdx.y;
  const int c = e / C;
  const int k = e - c * C;
  const int i = r * W + c;
  const double wk = w ? w[k] : 1.0;
  const T x0 = x[i * C + k];
  T v0 = T(0), v1 = T(0);
  if (r < H - 1) v0 = (T)(wk / h0) * (x[(i + W) * C + k] - x0);
  if (c < W - 1) v1 = (T)(wk / h1) * (x[(i + 1) * C + k] - x0);
  if constexpr (CT == 1) {
    using V2 = typename Vec2<T>::type;
    V2* out = reinterpret_cast<V2*>(y) + i;
    V2 o;
    if (add) {
      o = *out;
      o.x += v0;
      o.y += v1;
    } else {
      o.x = v0;
      o.y = v1;
    }
    *out = o;
  } else {
    T* out = y + (i * 2) * C + k;
    out[0] = add ? out[0] + v0 : v0;
    out[C] = add ? out[C] + v1 : v1;
  }
}

template <typename T, int CT>
__global__ void __launch_bounds__(kBlock) div2f_kernel(const T* __restrict__ p, T* __restrict__ x, int H, int W,
                                                       int Cr, const double* __restrict__ w, double h0, double h1,
                                                       int add) {
  const int C = CT > 0 ? CT : Cr;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= W * C) return;
  const int r = blockIdx.y;
  const int c = e / C;
  const int k = e - c * C;
  const int i = r * W + c;
  const double wk = w ? w[k] : 1.0;
  const T ih0 = (T)(1.0 / h0), ih1 = (T)(1.0 / h1);
  T p0, p1;
  if constexpr (CT == 1) {
    using V2 = typename Vec2<T>::type;
    const V2 pv = reinterpret_cast<const V2*>(p)[i];
    p0 = pv.x;
    p1 = pv.y;
  } else {
    p0 = p[(i * 2) * C + k];
    p1 = p[(i * 2 + 1) * C + k];
  }
  T acc = T(0);
  if (r < H - 1) acc -= p0 * ih0;
  if (r > 0) acc += p[((i - W) * 2) * C + k] * ih0;
  if (c < W - 1) acc -= p1 * ih1;
  if (c > 0) acc += p[((i - 1) * 2 + 1) * C + k] * ih1;
  const T val = (T)wk * acc;
  T* out = x + i * C + k;
  *out = add ? *out + val : val;
}

// The same on a 3-D grid (d0 x d1 x d2): a thread per (i1, i2, channel) of
// slab i0 = blockIdx.y, with 32-bit index arithmetic (the lifting model's
// (H, W, labels) grid, whose label axis is short).
template <typename T, int CT>
__global__ void __launch_bounds__(kBlock) grad3f_kernel(const T* __restrict__ x, T* __restrict__ y, int d0, int d1,
                                                        int d2, int Cr, const double* __restrict__ w, double h0,
                                                        double h1, double h2, int add) {
  const int C = CT > 0 ? CT : Cr;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = d2 * C;
  if (e >= d1 * row) return;
  const int i0 = blockIdx.y;
  const int i1 = e / row;
  const int rem = e - i1 * row;
  const int i2 = rem / C;
  const int k = rem - i2 * C;
  const int i = (i0 * d1 + i1) * d2 + i2;
  const double wk = w ? w[k] : 1.0;
  const T x0 = x[i * C + k];
  T v0 = T(0), v1 = T(0), v2 = T(0);
  if (i0 < d0 - 1) v0 = (T)(wk / h0) * (x[(i + d1 * d2) * C + k] - x0);
  if (i1 < d1 - 1) v1 = (T)(wk / h1) * (x[(i + d2) * C + k] - x0);
  if (i2 < d2 - 1) v2 = (T)(wk / h2) * (x[(i + 1) * C + k] - x0);
  T* out = y + (i * 3) * C + k;
  out[0] = add ? out[0] + v0 : v0;
  out[C] = add ? out[C] + v1 : v1;
  out[2 * C] = add ? out[2 * C] + v2 : v2;
}

template <typename T, int CT>
__global__ void __launch_bounds__(kBlock) div3f_kernel(const T* __restrict__ p, T* __restrict__ x, int d0, int d1,
                                                       int d2, int Cr, const double* __restrict__ w, double h0,
                                                       double h1, double h2, int add) {
  const int C = CT > 0 ? CT : Cr;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = d2 * C;
  if (e >= d1 * row) return;
  const int i0 = blockIdx.y;
  const int i1 = e / row;
  const int rem = e - i1 * row;
  const int i2 = rem / C;
  const int k = rem - i2 * C;
  const int i = (i0 * d1 + i1) * d2 + i2;
  const double wk = w ? w[k] : 1.0;
  const T ih0 = (T)(1.0 / h0), ih1 = (T)(1.0 / h1), ih2 = (T)(1.0 / h2);
  T acc = T(0);
  if (i0 < d0 - 1) acc -= p[(i * 3) * C + k] * ih0;
  if (i0 > 0) acc += p[((i - d1 * d2) * 3) * C + k] * ih0;
  if (i1 < d1 - 1) acc -= p[(i * 3 + 1) * C + k] * ih1;
  if (i1 > 0) acc += p[((i - d2) * 3 + 1) * C + k] * ih1;
  if (i2 < d2 - 1) acc -= p[(i * 3 + 2) * C + k] * ih2;
  if (i2 > 0) acc += p[((i - 1) * 3 + 2) * C + k] * ih2;
  const T val = (T)wk * acc;
  T* out = x + i * C + k;
  *out = add ? *out + val : val;
}

inline bool use_3d_forward(const Grid& g) {
  return g.ndims == 3 && g.dims[0] <= 65535 && g.npoints * g.nchan * 3 < (int64_t(1) << 31);
}

#define OPY_LAUNCH_3F(KERN, T, in, out, g, w, add, st)                                                        \
  do {                                                                                                      \
    const int d0_ = (int)g.dims[0], d1_ = (int)g.dims[1], d2_ = (int)g.dims[2], C_ = g.nchan;                \
    const dim3 grid_((unsigned)((d1_ * d2_ * C_ + kBlock - 1) / kBlock), (unsigned)d0_);                     \
    if (C_ == 1)                                                                                            \
      hipLaunchKernelGGL((KERN<T, 1>), grid_, dim3(kBlock), 0, st, in, out, d0_, d1_, d2_, C_, w, g.h[0],     \
                         g.h[1], g.h[2], add);                                                              \
    else                                                                                                    \
      hipLaunchKernelGGL((KERN<T, 0>), grid_, dim3(kBlock), 0, st, in, out, d0_, d1_, d2_, C_, w, g.h[0],     \
                         g.h[1], g.h[2], add);                                                              \
  } while (0)

// The 2-D kernels serve forward 2-D grids whose indices fit 32 bits and whose
// rows fit the grid's second dimension.
inline bool use_2d_forward(const Grid& g) {
  return g.ndims == 2 && g.dims[0] >= 1 && g.dims[0] <= 65535 && g.npoints * g.nchan * 2 < (int64_t(1) << 31);
}

// vec: the field (the gradient's output, the divergence's input) is aligned for 2-vector accesses
#define OPY_LAUNCH_2F(KERN, T, in, out, g, w, add, st, vec)                                                      \
  do {                                                                                                         \
    const int H_ = (int)g.dims[0], W_ = (int)g.dims[1], C_ = g.nchan;                                          \
    const dim3 grid_((unsigned)((W_ * C_ + kBlock - 1) / kBlock), (unsigned)H_);                                \
    if (C_ == 1 && (vec))                                                                                      \
      hipLaunchKernelGGL((KERN<T, 1>), grid_, dim3(kBlock), 0, st, in, out, H_, W_, C_, w, g.h[0], g.h[1], add); \
    else if (C_ == 2)                                                                                          \
      hipLaunchKernelGGL((KERN<T, 2>), grid_, dim3(kBlock), 0, st, in, out, H_, W_, C_, w, g.h[0], g.h[1], add); \
    else if (C_ == 3)                                                                                          \
      hipLaunchKernelGGL((KERN<T, 3>), grid_, dim3(kBlock), 0, st, in, out, H_, W_, C_, w, g.h[0], g.h[1], add); \
    else                                                                                                       \
      hipLaunchKernelGGL((KERN<T, 0>), grid_, dim3(kBlock), 0, st, in, out, H_, W_, C_, w, g.h[0], g.h[1], add); \
  } while (0)

// Row-wise l^p norms |A_i|_p^p of the gradient (rows = (point, t, channel)).
template <typename T, bool kCentered>
__global__ void __launch_bounds__(kBlock) grad_rowwise_kernel(T* __restrict__ y, Grid g,
                                                              const double* __restrict__ w, double p,
                                                              int add) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int C = g.nchan;
  if (idx >= g.npoints * C) return;
  const int64_t i = idx / C;
  const int k = (int)(idx - i * C);
  const int D = g.ndims;
  int64_t c[kMaxDims];
  point_coords(g, i, c);
  const double wk = fabs(w ? w[k] : 1.0);
  bool br = false;
  if (kCentered)
    for (int a = 0; a < D; ++a) br |= (c[a] == g.dims[a] - 1);
  const int navg = 1 << (D - 1);
  for (int t = 0; t < D; ++t) {
    double v = 0.0;
    if (kCentered) {
      if (!br) v = 2.0 * navg * pow(wk / (g.h[t] * navg), p);
    } else if (c[t] < g.dims[t] - 1) {
      v = 2.0 * pow(wk / g.h[t], p);
    }
    T* out = y + (i * D + t) * C + k;
    *out = add ? (T)((double)*out + v) : (T)v;
  }
}

// Row-wise l^p norms of the divergence (= column sums of |grad|^p).
template <typename T, bool kCentered>
__global__ void __launch_bounds__(kBlock) div_rowwise_kernel(T* __restrict__ x, Grid g,
                                                             const double* __restrict__ w, double p,
                                                             int add) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int C = g.nchan;
  if (idx >= g.npoints * C) return;
  const int64_t i = idx / C;
  const int k = (int)(idx - i * C);
  const int D = g.ndims;
  int64_t c[kMaxDims];
  point_coords(g, i, c);
  const double wk = fabs(w ? w[k] : 1.0);
  double v = 0.0;
  if (kCentered) {
    const int navg = 1 << (D - 1);
    int64_t bits[kMaxDims];
    for (int t = 0; t < D; ++t) {
      const double e = pow(wk / (g.h[t] * navg), p);
      for (int m = 0; m < navg; ++m) {
        cell_offset(g, t, m, bits);
        bool ok = true;
        for (int a = 0; a < D; ++a) {
          if (a == t) continue;
          const int64_t ca = c[a] - bits[a];
          ok &= (ca >= 0) && (ca != g.dims[a] - 1);
        }
        if (!ok) continue;
        if (c[t] < g.dims[t] - 1) v += e;
        if (c[t] > 0) v += e;
      }
    }
  } else {
    for (int t = 0; t < D; ++t) {
      const double e = pow(wk / g.h[t], p);
      if (c[t] < g.dims[t] - 1) v += e;
      if (c[t] > 0) v += e;
    }
  }
  T* out = x + i * C + k;
  *out = add ? (T)((double)*out + v) : (T)v;
}

// y (N,C) (+)= Laplacian(x) with boundary handling per reference diff.cu:131-243.
// bc: 'N' neumann, 'C' curvature, 'S' second-order; adj selects the adjoint.
template <typename T>
__global__ void __launch_bounds__(kBlock) laplacian_kernel(const T* __restrict__ x, T* __restrict__ y,
                                                           Grid g, int bc, int adj, int add) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int C = g.nchan;
  if (idx >= g.npoints * C) return;
  const int64_t i = idx / C;
  const int k = (int)(idx - i * C);
  const int D = g.ndims;
  int64_t c[kMaxDims];
  point_coords(g, i, c);
  const T xi = x[i * C + k];
  T v = T(0);
  if (bc == 'C') {
    if (adj) {
      int bdim = -1;
      bool corner = false;
      for (int t = 0; t < D; ++t) {
        if (c[t] == 0 || c[t] == g.dims[t] - 1) {
          if (bdim >= 0) corner = true;
          else bdim = t;
        }
      }
      if (!corner) {
        if (bdim >= 0) {
          const int t = bdim;
          const T ih2 = (T)(1.0 / (g.h[t] * g.h[t]));
          // an axis of length 1 has no neighbour to read
          if (g.dims[t] > 1) {
            if (c[t] == 0) v += ih2 * x[(i + g.skips[t]) * C + k];
            else v += ih2 * x[(i - g.skips[t]) * C + k];
          }
        } else {
          for (int t = 0; t < D; ++t) {
            const T ih2 = (T)(1.0 / (g.h[t] * g.h[t]));
            v -= T(2) * ih2 * xi;
            if (c[t] - 1 > 0) v += ih2 * x[(i - g.skips[t]) * C + k];
            if (c[t] + 1 < g.dims[t] - 1) v += ih2 * x[(i + g.skips[t]) * C + k];
          }
        }
      }
    } else {
      bool bnd = false;
      for (int t = 0; t < D; ++t) bnd |= (c[t] == 0 || c[t] == g.dims[t] - 1);
      if (!bnd) {
        for (int t = 0; t < D; ++t) {
          const T ih2 = (T)(1.0 / (g.h[t] * g.h[t]));
          v += ih2 * (x[(i - g.skips[t]) * C + k] - xi) + ih2 * (x[(i + g.skips[t]) * C + k] - xi);
        }
      }
    }
  } else if (bc == 'S') {
    for (int t = 0; t < D; ++t) {
      const T ih2 = (T)(1.0 / (g.h[t] * g.h[t]));
      const bool inner = c[t] > 0 && c[t] < g.dims[t] - 1;
      if (adj) {
        if (inner) v -= ih2 * T(2) * xi;
        if (c[t] > 1) v += ih2 * x[(i - g.skips[t]) * C + k];
        if (c[t] < g.dims[t] - 2) v += ih2 * x[(i + g.skips[t]) * C + k];
      } else if (inner) {
        v += ih2 * (x[(i - g.skips[t]) * C + k] - T(2) * xi + x[(i + g.skips[t]) * C + k]);
      }
    }
  } else {  // 'N'
    for (int t = 0; t < D; ++t) {
      const T ih2 = (T)(1.0 / (g.h[t] * g.h[t]));
      if (c[t] > 0) v += ih2 * (x[(i - g.skips[t]) * C + k] - xi);
      if (c[t] < g.dims[t] - 1) v += ih2 * (x[(i + g.skips[t]) * C + k] - xi);
    }
  }
  T* out = y + i * C + k;
  *out = add ? *out + v : v;
}

// Row-wise l^p norms of the Laplacian rows (for diagonal preconditioning).
template <typename T>
__global__ void __launch_bounds__(kBlock) laplacian_rowwise_kernel(T* __restrict__ y, Grid g, int bc,
                                                                   int adj, double p, int add) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int C = g.nchan;
  if (idx >= g.npoints * C) return;
  const int64_t i = idx / C;
  const int k = (int)(idx - i * C);
  const int D = g.ndims;
  int64_t c[kMaxDims];
  point_coords(g, i, c);
  // Accumulate the coefficient of every column touched by row i, merging the
  // diagonal entry, then sum |coef|^p.
  double diag = 0.0, off = 0.0;
  if (bc == 'C') {
    if (adj) {
      int bdim = -1;
      bool corner = false;
      for (int t = 0; t < D; ++t)
        if (c[t] == 0 || c[t] == g.dims[t] - 1) {
          if (bdim >= 0) corner = true;
          else bdim = t;
        }
      if (!corner) {
        if (bdim >= 0) {
          off += pow(1.0 / (g.h[bdim] * g.h[bdim]), p);
        } else {
          for (int t = 0; t < D; ++t) {
            const double ih2 = 1.0 / (g.h[t] * g.h[t]);
            diag -= 2.0 * ih2;
            if (c[t] - 1 > 0) off += pow(ih2, p);
            if (c[t] + 1 < g.dims[t] - 1) off += pow(ih2, p);
          }
        }
      }
    } else {
      bool bnd = false;
      for (int t = 0; t < D; ++t) bnd |= (c[t] == 0 || c[t] == g.dims[t] - 1);
      if (!bnd)
        for (int t = 0; t < D; ++t) {
          const double ih2 = 1.0 / (g.h[t] * g.h[t]);
          diag -= 2.0 * ih2;
          off += 2.0 * pow(ih2, p);
        }
    }
  } else if (bc == 'S') {
    for (int t = 0; t < D; ++t) {
      const double ih2 = 1.0 / (g.h[t] * g.h[t]);
      const bool inner = c[t] > 0 && c[t] < g.dims[t] - 1;
      if (adj) {
        if (inner) diag -= 2.0 * ih2;
        if (c[t] > 1) off += pow(ih2, p);
        if (c[t] < g.dims[t] - 2) off += pow(ih2, p);
      } else if (inner) {
        diag -= 2.0 * ih2;
        off += 2.0 * pow(ih2, p);
      }
    }
  } else {
    for (int t = 0; t < D; ++t) {
      const double ih2 = 1.0 / (g.h[t] * g.h[t]);
      if (c[t] > 0) {
        diag -= ih2;
        off += pow(ih2, p);
      }
      if (c[t] < g.dims[t] - 1) {
        diag -= ih2;
        off += pow(ih2, p);
      }
    }
  }
  const double v = off + (diag != 0.0 ? pow(fabs(diag), p) : 0.0);
  T* out = y + i * C + k;
  *out = add ? (T)((double)*out + v) : (T)v;
}

template <typename T>
int grad_impl(const T* x, T* y, const opy_grid_t* gd, const double* w, int scheme, int add, void* s) {
  Grid g;
  if (int rc = make_grid(gd, &g)) return rc;
  if (!x || !y) return fail(kInvalid, "gradient: null pointer");
  const int64_t n = g.npoints * g.nchan;
  if (scheme == OPY_SCHEME_CENTERED)
    hipLaunchKernelGGL((grad_kernel<T, true>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), x, y, g, w, add);
  else if (scheme == OPY_SCHEME_FORWARD && use_2d_forward(g)) {
    if (n > 0) OPY_LAUNCH_2F(grad2f_kernel, T, x, y, g, w, add, as_stream(s),
                             reinterpret_cast<uintptr_t>(y) % (2 * sizeof(T)) == 0);
  } else if (scheme == OPY_SCHEME_FORWARD && use_3d_forward(g)) {
    if (n > 0) OPY_LAUNCH_3F(grad3f_kernel, T, x, y, g, w, add, as_stream(s));
  } else if (scheme == OPY_SCHEME_FORWARD)
    hipLaunchKernelGGL((grad_kernel<T, false>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), x, y, g, w, add);
  else
    return fail(kInvalid, "gradient: unknown scheme %d", scheme);
  return check_launch("gradient");
}

template <typename T>
int div_impl(const T* p, T* x, const opy_grid_t* gd, const double* w, int scheme, int add, void* s) {
  Grid g;
  if (int rc = make_grid(gd, &g)) return rc;
  if (!x || !p) return fail(kInvalid, "divergence: null pointer");
  const int64_t n = g.npoints * g.nchan;
  if (scheme == OPY_SCHEME_CENTERED)
    hipLaunchKernelGGL((div_kernel<T, true>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), p, x, g, w, add);
  else if (scheme == OPY_SCHEME_FORWARD && use_2d_forward(g)) {
    if (n > 0) OPY_LAUNCH_2F(div2f_kernel, T, p, x, g, w, add, as_stream(s),
                             reinterpret_cast<uintptr_t>(p) % (2 * sizeof(T)) == 0);
  } else if (scheme == OPY_SCHEME_FORWARD && use_3d_forward(g)) {
    if (n > 0) OPY_LAUNCH_3F(div3f_kernel, T, p, x, g, w, add, as_stream(s));
  } else if (scheme == OPY_SCHEME_FORWARD)
    hipLaunchKernelGGL((div_kernel<T, false>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), p, x, g, w, add);
  else
    return fail(kInvalid, "divergence: unknown scheme %d", scheme);
  return check_launch("divergence");
}

template <typename T>
int diff_rowwise_impl(T* out, const opy_grid_t* gd, const double* w, int scheme, int adjoint, double p,
                      int add, void* s) {
  Grid g;
  if (int rc = make_grid(gd, &g)) return rc;
  if (!out) return fail(kInvalid, "diff rowwise: null pointer");
  if (!(p > 0)) return fail(kInvalid, "diff rowwise: p=%g must be positive", p);
  const int64_t n = g.npoints * g.nchan;
  const bool cen = scheme == OPY_SCHEME_CENTERED;
  if (scheme != OPY_SCHEME_CENTERED && scheme != OPY_SCHEME_FORWARD)
    return fail(kInvalid, "diff rowwise: unknown scheme %d", scheme);
  dim3 gr(grid_for(n)), bl(kBlock);
  hipStream_t st = as_stream(s);
  if (adjoint) {
    if (cen) hipLaunchKernelGGL((div_rowwise_kernel<T, true>), gr, bl, 0, st, out, g, w, p, add);
    else hipLaunchKernelGGL((div_rowwise_kernel<T, false>), gr, bl, 0, st, out, g, w, p, add);
  } else {
    if (cen) hipLaunchKernelGGL((grad_rowwise_kernel<T, true>), gr, bl, 0, st, out, g, w, p, add);
    else hipLaunchKernelGGL((grad_rowwise_kernel<T, false>), gr, bl, 0, st, out, g, w, p, add);
  }
  return check_launch("diff rowwise");
}

int check_bc(int bc) {
  if (bc != 'N' && bc != 'C' && bc != 'S') return fail(kInvalid, "laplacian: unknown boundary '%c'", (char)bc);
  return kOk;
}

template <typename T>
int laplacian_impl(const T* x, T* y, const opy_grid_t* gd, int bc, int adj, int add, void* s) {
  Grid g;
  if (int rc = make_grid(gd, &g)) return rc;
  if (int rc = check_bc(bc)) return rc;
  if (!x || !y) return fail(kInvalid, "laplacian: null pointer");
  const int64_t n = g.npoints * g.nchan;
  hipLaunchKernelGGL((laplacian_kernel<T>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), x, y, g, bc, adj,
                     add);
  return check_launch("laplacian");
}

template <typename T>
int laplacian_rowwise_impl(T* y, const opy_grid_t* gd, int bc, int adj, double p, int add, void* s) {
  Grid g;
  if (int rc = make_grid(gd, &g)) return rc;
  if (int rc = check_bc(bc)) return rc;
  if (!y) return fail(kInvalid, "laplacian rowwise: null pointer");
  const int64_t n = g.npoints * g.nchan;
  hipLaunchKernelGGL((laplacian_rowwise_kernel<T>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), y, g, bc,
                     adj, p, add);
  return check_launch("laplacian rowwise");
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_grad_f32(const float* x, float* y, const opy_grid_t* g, const double* w, int scheme, int add,
                         void* s) {
  return grad_impl<float>(x, y, g, w, scheme, add, s);
}
OPY_API int opy_grad_f64(const double* x, double* y, const opy_grid_t* g, const double* w, int scheme, int add,
                         void* s) {
  return grad_impl<double>(x, y, g, w, scheme, add, s);
}
OPY_API int opy_div_f32(const float* p, float* x, const opy_grid_t* g, const double* w, int scheme, int add,
                        void* s) {
  return div_impl<float>(p, x, g, w, scheme, add, s);
}
OPY_API int opy_div_f64(const double* p, double* x, const opy_grid_t* g, const double* w, int scheme, int add,
                        void* s) {
  return div_impl<double>(p, x, g, w, scheme, add, s);
}
OPY_API int opy_diff_rowwise_f32(float* out, const opy_grid_t* g, const double* w, int scheme, int adjoint,
                                 double p, int add, void* s) {
  return diff_rowwise_impl<float>(out, g, w, scheme, adjoint, p, add, s);
}
OPY_API int opy_diff_rowwise_f64(double* out, const opy_grid_t* g, const double* w, int scheme, int adjoint,
                                 double p, int add, void* s) {
  return diff_rowwise_impl<double>(out, g, w, scheme, adjoint, p, add, s);
}
OPY_API int opy_laplacian_f32(const float* x, float* y, const opy_grid_t* g, int bc, int adj, int add, void* s) {
  return laplacian_impl<float>(x, y, g, bc, adj, add, s);
}
OPY_API int opy_laplacian_f64(const double* x, double* y, const opy_grid_t* g, int bc, int adj, int add,
                              void* s) {
  return laplacian_impl<double>(x, y, g, bc, adj, add, s);
}
OPY_API int opy_laplacian_rowwise_f32(float* y, const opy_grid_t* g, int bc, int adj, double p, int add,
                                      void* s) {
  return laplacian_rowwise_impl<float>(y, g, bc, adj, p, add, s);
}
OPY_API int opy_laplacian_rowwise_f64(double* y, const opy_grid_t* g, int bc, int adj, double p, int add,
                                      void* s) {
  return laplacian_rowwise_impl<double>(y, g, bc, adj, p, add, s);
}
