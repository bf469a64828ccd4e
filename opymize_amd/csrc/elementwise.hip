// Pointwise kernels: the trivial linear operators (scale, identity, nihil,
// diagonal multiplication), the affine prox maps (shift-scale, SSD prox,
// constants, masks, positivity), and the elementwise PDHG steps.
#include "opy_common.h"

namespace opy {
namespace {

// Step size: per-element array, device scalar, or host scalar (in that order).
template <typename T>
struct Step {
  const T* arr;
  const double* dev;
  double host;
  __device__ inline double at(int64_t i) const {
    if (arr) return (double)arr[i];
    if (dev) return *dev;
    return host;
  }
};

template <typename T>
Step<T> make_step(const opy_step_t* s) {
  Step<T> r{nullptr, nullptr, 0.0};
  if (s) {
    r.arr = static_cast<const T*>(s->arr);
    r.dev = s->dev;
    r.host = s->host;
  }
  return r;
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ y, int64_t n, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

template <typename T>
__global__ void copy_kernel(const T* __restrict__ x, T* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}

// Copy with 16-byte accesses when both arrays are 16-byte aligned: thread i
// moves chunk i of the nv whole chunks, and the first n - head threads also
// one element of the tail [head, n) (fewer than 16 bytes).
template <typename T>
__global__ void copy16_kernel(const uint4* __restrict__ x, uint4* __restrict__ y, int64_t nv,
                              const T* __restrict__ xt, T* __restrict__ yt, int64_t head, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) y[i] = x[i];
  if (i < n - head) yt[head + i] = xt[head + i];
}

// y (+)= fact * x (reference: ScaleOp, scale.py:48-62; IdentityOp with fact 1)
template <typename T>
__global__ void scale_kernel(const T* __restrict__ x, T* __restrict__ y, int64_t n, const T* __restrict__ fa,
                             T fs, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T f = fa ? fa[i] : fs;
  const T v = f * x[i];
  y[i] = add ? y[i] + v : v;
}

// y[keep] (+)= x[keep], y[~keep] = 0 unless add (reference: NihilOp, scale.py:96-115)
template <typename T>
__global__ void nihil_kernel(const T* __restrict__ x, T* __restrict__ y, int64_t n,
                             const uint8_t* __restrict__ keep, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (keep[i]) y[i] = add ? y[i] + x[i] : x[i];
  else if (!add) y[i] = T(0);
}

// y[i,k] (+)= x[i,k] * A[k] (reference: DiagMatrixMultR, einsum.py:104-128)
template <typename T>
__global__ void diag_matmul_r_kernel(const T* __restrict__ x, T* __restrict__ y, int64_t n, int K,
                                     const T* __restrict__ A, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T v = x[i] * A[i % K];
  y[i] = add ? y[i] + v : v;
}

// y (+)= a*(x + b*shift); a, b, shift each array or scalar
// (reference: ShiftScaleOp, affine.py:112-146)
template <typename T>
__global__ void shiftscale_kernel(const T* __restrict__ x, T* __restrict__ y, int64_t n,
                                  const T* __restrict__ sa, T ss, const T* __restrict__ aa, T as,
                                  const T* __restrict__ ba, T bs, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T sh = sa ? sa[i] : ss;
  const T a = aa ? aa[i] : as;
  const T b = ba ? ba[i] : bs;
  const T v = a * (x[i] + b * sh);
  y[i] = add ? y[i] + v : v;
}

// prox of G(x) = 0.5*sum_k vol_k (x_ik - f_ik)^2 on unmasked points:
// y = (x + t*vol*f) / (1 + t*vol)   (reference: SSD.prox, quad.py:45-55 with
// ShiftScaleOp a = 1/(1+t*vol), b = t*vol)
template <typename T>
__global__ void ssd_prox_kernel(const T* __restrict__ x, T* __restrict__ y, int64_t n, int C,
                                const T* __restrict__ f, const double* __restrict__ vol,
                                const uint8_t* __restrict__ mask, Step<T> step, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t p = i / C;
  const int k = (int)(i - p * C);
  const bool on = mask ? (mask[p] != 0) : true;
  const double tv = on ? step.at(i) * vol[k] : 0.0;
  const T a = (T)(1.0 / (1.0 + tv));
  const T b = (T)tv;
  const T sh = on ? f[i] : T(0);
  const T v = a * (x[i] + b * sh);
  y[i] = add ? y[i] + v : v;
}

// prox of G(x) = sum_k lbd_k |x_ik - f_ik| (soft threshold about f):
// y = f + shrink(x - f, t*lbd)
template <typename T>
__global__ void l1data_prox_kernel(const T* __restrict__ x, T* __restrict__ y, int64_t n, int C,
                                   const T* __restrict__ f, const double* __restrict__ lbd, Step<T> step,
                                   int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = (int)(i % C);
  const T thr = (T)(step.at(i) * lbd[k]);
  const T d = x[i] - f[i];
  const T s = d > thr ? d - thr : (d < -thr ? d + thr : T(0));
  const T v = f[i] + s;
  y[i] = add ? y[i] + v : v;
}

// prox of the conjugate of the L1 data term: z -> clip(z - t*f, -t*... )
// G*(z) = sum_k [ delta_{|z_ik| <= lbd_k} + z_ik f_ik ], prox_{tG*}(z) = clip(z - t f, -lbd, lbd)
template <typename T>
__global__ void l1data_conj_prox_kernel(const T* __restrict__ x, T* __restrict__ y, int64_t n, int C,
                                        const T* __restrict__ f, const double* __restrict__ lbd, Step<T> step,
                                        int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = (int)(i % C);
  const T l = (T)lbd[k];
  const T z = x[i] - (T)step.at(i) * f[i];
  const T v = z > l ? l : (z < -l ? -l : z);
  y[i] = add ? y[i] + v : v;
}

// max(0,x) / min(0,x)   (reference: PosProj / NegProj, proj.py:19-85)
template <typename T, bool kPos>
__global__ void signproj_kernel(const T* __restrict__ x, T* __restrict__ y, int64_t n, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T xi = x[i];
  const T v = kPos ? (xi < T(0) ? T(0) : xi) : (xi > T(0) ? T(0) : xi);
  y[i] = add ? y[i] + v : v;
}

// y (+)= const   (reference: ConstOp, affine.py:15-46)
template <typename T>
__global__ void const_kernel(T* __restrict__ y, int64_t n, const T* __restrict__ c, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] = add ? y[i] + c[i] : c[i];
}

// y (+)= mask[i/M] ? c : x   (reference: ConstrainOp, affine.py:48-83)
template <typename T>
__global__ void constrain_kernel(const T* __restrict__ x, T* __restrict__ y, int64_t n, int64_t M,
                                 const T* __restrict__ c, const uint8_t* __restrict__ mask, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T v = mask[i / M] ? c[i] : x[i];
  y[i] = add ? y[i] + v : v;
}

// z_next = z + dir * step * g    (reference: take_step, pdhg.py:62-65)
template <typename T>
__global__ void take_step_kernel(T* __restrict__ zn, const T* __restrict__ z, int64_t n, int dir, Step<T> step,
                                 const T* __restrict__ g) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T s = (T)step.at(i);
  zn[i] = z[i] + (T)dir * s * g[i];
}

// y_next = y + step * ((1+theta) Ax_next - theta Ax)
// (reference: overrelax + take_step, pdhg.py:148-151, kernels :66-69, :62-65)
template <typename T>
__global__ void dual_step_kernel(T* __restrict__ yn, const T* __restrict__ y, int64_t n, T theta, Step<T> step,
                                 const T* __restrict__ axn, const T* __restrict__ ax) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T ybar = (T(1) + theta) * axn[i] - theta * ax[i];
  const T s = (T)step.at(i);
  yn[i] = y[i] + s * ybar;
}

// y[i] (+)= v[(i / div) % nv]: broadcast a short vector along a flat array
template <typename T>
__global__ void tile_kernel(T* __restrict__ y, int64_t n, const T* __restrict__ v, int64_t nv, int64_t div, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T e = v[(i / div) % nv];
  y[i] = add ? y[i] + e : e;
}

template <typename T>
__global__ void abs_kernel(const T* __restrict__ x, T* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = fabs(x[i]);
}

// x = 1/x where x > thr, unchanged elsewhere (diagonal preconditioners,
// reference pdhg.py:197-200)
template <typename T>
__global__ void recip_gt_kernel(T* __restrict__ x, int64_t n, double thr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T v = x[i];
  if ((double)v > thr) x[i] = T(1) / v;
}

template <typename T>
int check_n(const void* a, const void* b, int64_t n, const char* what) {
  if (n < 0) return fail(kInvalid, "%s: negative size", what);
  if (n > 0 && (!a || !b)) return fail(kInvalid, "%s: null pointer", what);
  return kOk;
}

}  // namespace
}  // namespace opy

using namespace opy;

#define OPY_LAUNCH(kern, n, s, ...)                                                                  \
  do {                                                                                               \
    if ((n) > 0) hipLaunchKernelGGL(kern, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), __VA_ARGS__); \
  } while (0)

// ---- fill / copy ---------------------------------------------------------
template <typename T>
static int fill_impl(T* y, int64_t n, double v, void* s) {
  if (n > 0 && !y) return fail(kInvalid, "fill: null pointer");
  OPY_LAUNCH((fill_kernel<T>), n, s, y, n, (T)v);
  return check_launch("fill");
}
OPY_API int opy_fill_f32(float* y, int64_t n, double v, void* s) { return fill_impl<float>(y, n, v, s); }
OPY_API int opy_fill_f64(double* y, int64_t n, double v, void* s) { return fill_impl<double>(y, n, v, s); }

template <typename T>
static int copy_impl(const T* x, T* y, int64_t n, void* s) {
  if (int rc = check_n<T>(x, y, n, "copy")) return rc;
  if (x == y || n == 0) return kOk;
  if (reinterpret_cast<uintptr_t>(x) % 16 == 0 && reinterpret_cast<uintptr_t>(y) % 16 == 0) {
    constexpr int64_t per = 16 / sizeof(T);
    const int64_t nv = n / per;
    const int64_t head = nv * per;
    OPY_LAUNCH((copy16_kernel<T>), (nv > 0 ? nv : 1), s, reinterpret_cast<const uint4*>(x),
               reinterpret_cast<uint4*>(y), nv, x, y, head, n);
  } else {
    OPY_LAUNCH((copy_kernel<T>), n, s, x, y, n);
  }
  return check_launch("copy");
}
OPY_API int opy_copy_f32(const float* x, float* y, int64_t n, void* s) { return copy_impl<float>(x, y, n, s); }
OPY_API int opy_copy_f64(const double* x, double* y, int64_t n, void* s) { return copy_impl<double>(x, y, n, s); }

template <typename T>
static int tile_impl(T* y, int64_t n, const T* v, int64_t nv, int64_t div, int add, void* s) {
  if (int rc = check_n<T>(y, v, n, "tile")) return rc;
  if (nv < 1 || div < 1) return fail(kInvalid, "tile: nv and div must be positive");
  OPY_LAUNCH((tile_kernel<T>), n, s, y, n, v, nv, div, add);
  return check_launch("tile");
}
OPY_API int opy_tile_f32(float* y, int64_t n, const float* v, int64_t nv, int64_t div, int add, void* s) {
  return tile_impl<float>(y, n, v, nv, div, add, s);
}
OPY_API int opy_tile_f64(double* y, int64_t n, const double* v, int64_t nv, int64_t div, int add, void* s) {
  return tile_impl<double>(y, n, v, nv, div, add, s);
}

template <typename T>
static int abs_impl(const T* x, T* y, int64_t n, void* s) {
  if (int rc = check_n<T>(x, y, n, "abs")) return rc;
  OPY_LAUNCH((abs_kernel<T>), n, s, x, y, n);
  return check_launch("abs");
}
OPY_API int opy_abs_f32(const float* x, float* y, int64_t n, void* s) { return abs_impl<float>(x, y, n, s); }
OPY_API int opy_abs_f64(const double* x, double* y, int64_t n, void* s) { return abs_impl<double>(x, y, n, s); }

template <typename T>
static int recip_gt_impl(T* x, int64_t n, double thr, void* s) {
  if (n > 0 && !x) return fail(kInvalid, "recip_gt: null pointer");
  OPY_LAUNCH((recip_gt_kernel<T>), n, s, x, n, thr);
  return check_launch("recip_gt");
}
OPY_API int opy_recip_gt_f32(float* x, int64_t n, double thr, void* s) { return recip_gt_impl<float>(x, n, thr, s); }
OPY_API int opy_recip_gt_f64(double* x, int64_t n, double thr, void* s) { return recip_gt_impl<double>(x, n, thr, s); }

// ---- scale / nihil / diag ------------------------------------------------
template <typename T>
static int scale_impl(const T* x, T* y, int64_t n, const T* fa, double fs, int add, void* s) {
  if (int rc = check_n<T>(x, y, n, "scale")) return rc;
  OPY_LAUNCH((scale_kernel<T>), n, s, x, y, n, fa, (T)fs, add);
  return check_launch("scale");
}
OPY_API int opy_scale_f32(const float* x, float* y, int64_t n, const float* fa, double fs, int add, void* s) {
  return scale_impl<float>(x, y, n, fa, fs, add, s);
}
OPY_API int opy_scale_f64(const double* x, double* y, int64_t n, const double* fa, double fs, int add, void* s) {
  return scale_impl<double>(x, y, n, fa, fs, add, s);
}

template <typename T>
static int nihil_impl(const T* x, T* y, int64_t n, const uint8_t* keep, int add, void* s) {
  if (int rc = check_n<T>(x, y, n, "nihil")) return rc;
  if (n > 0 && !keep) return fail(kInvalid, "nihil: null mask");
  OPY_LAUNCH((nihil_kernel<T>), n, s, x, y, n, keep, add);
  return check_launch("nihil");
}
OPY_API int opy_nihil_f32(const float* x, float* y, int64_t n, const uint8_t* keep, int add, void* s) {
  return nihil_impl<float>(x, y, n, keep, add, s);
}
OPY_API int opy_nihil_f64(const double* x, double* y, int64_t n, const uint8_t* keep, int add, void* s) {
  return nihil_impl<double>(x, y, n, keep, add, s);
}

template <typename T>
static int diag_impl(const T* x, T* y, int64_t N, int K, const T* A, int add, void* s) {
  const int64_t n = N * (int64_t)K;
  if (int rc = check_n<T>(x, y, n, "diag_matmul_r")) return rc;
  if (K < 1 || !A) return fail(kInvalid, "diag_matmul_r: bad diagonal");
  OPY_LAUNCH((diag_matmul_r_kernel<T>), n, s, x, y, n, K, A, add);
  return check_launch("diag_matmul_r");
}
OPY_API int opy_diag_matmul_r_f32(const float* x, float* y, int64_t N, int K, const float* A, int add, void* s) {
  return diag_impl<float>(x, y, N, K, A, add, s);
}
OPY_API int opy_diag_matmul_r_f64(const double* x, double* y, int64_t N, int K, const double* A, int add,
                                  void* s) {
  return diag_impl<double>(x, y, N, K, A, add, s);
}

// ---- affine prox maps ----------------------------------------------------
template <typename T>
static int shiftscale_impl(const T* x, T* y, int64_t n, const T* sa, double ss, const T* aa, double as,
                           const T* ba, double bs, int add, void* s) {
  if (int rc = check_n<T>(x, y, n, "shiftscale")) return rc;
  OPY_LAUNCH((shiftscale_kernel<T>), n, s, x, y, n, sa, (T)ss, aa, (T)as, ba, (T)bs, add);
  return check_launch("shiftscale");
}
OPY_API int opy_shiftscale_f32(const float* x, float* y, int64_t n, const float* sa, double ss, const float* aa,
                               double as, const float* ba, double bs, int add, void* s) {
  return shiftscale_impl<float>(x, y, n, sa, ss, aa, as, ba, bs, add, s);
}
OPY_API int opy_shiftscale_f64(const double* x, double* y, int64_t n, const double* sa, double ss,
                               const double* aa, double as, const double* ba, double bs, int add, void* s) {
  return shiftscale_impl<double>(x, y, n, sa, ss, aa, as, ba, bs, add, s);
}

template <typename T>
static int ssd_prox_impl(const T* x, T* y, int64_t N, int C, const T* f, const double* vol, const uint8_t* mask,
                         const opy_step_t* st, int add, void* s) {
  const int64_t n = N * (int64_t)C;
  if (int rc = check_n<T>(x, y, n, "ssd_prox")) return rc;
  if (n > 0 && (!f || !vol)) return fail(kInvalid, "ssd_prox: null data or volume");
  OPY_LAUNCH((ssd_prox_kernel<T>), n, s, x, y, n, C, f, vol, mask, make_step<T>(st), add);
  return check_launch("ssd_prox");
}
OPY_API int opy_ssd_prox_f32(const float* x, float* y, int64_t N, int C, const float* f, const double* vol,
                             const uint8_t* mask, const opy_step_t* st, int add, void* s) {
  return ssd_prox_impl<float>(x, y, N, C, f, vol, mask, st, add, s);
}
OPY_API int opy_ssd_prox_f64(const double* x, double* y, int64_t N, int C, const double* f, const double* vol,
                             const uint8_t* mask, const opy_step_t* st, int add, void* s) {
  return ssd_prox_impl<double>(x, y, N, C, f, vol, mask, st, add, s);
}

template <typename T>
static int l1data_prox_impl(const T* x, T* y, int64_t N, int C, const T* f, const double* lbd, int conj,
                            const opy_step_t* st, int add, void* s) {
  const int64_t n = N * (int64_t)C;
  if (int rc = check_n<T>(x, y, n, "l1data_prox")) return rc;
  if (n > 0 && (!f || !lbd)) return fail(kInvalid, "l1data_prox: null data or weights");
  if (conj) OPY_LAUNCH((l1data_conj_prox_kernel<T>), n, s, x, y, n, C, f, lbd, make_step<T>(st), add);
  else OPY_LAUNCH((l1data_prox_kernel<T>), n, s, x, y, n, C, f, lbd, make_step<T>(st), add);
  return check_launch("l1data_prox");
}
OPY_API int opy_l1data_prox_f32(const float* x, float* y, int64_t N, int C, const float* f, const double* lbd,
                                int conj, const opy_step_t* st, int add, void* s) {
  return l1data_prox_impl<float>(x, y, N, C, f, lbd, conj, st, add, s);
}
OPY_API int opy_l1data_prox_f64(const double* x, double* y, int64_t N, int C, const double* f, const double* lbd,
                                int conj, const opy_step_t* st, int add, void* s) {
  return l1data_prox_impl<double>(x, y, N, C, f, lbd, conj, st, add, s);
}

template <typename T>
static int signproj_impl(const T* x, T* y, int64_t n, int positive, int add, void* s) {
  if (int rc = check_n<T>(x, y, n, "sign projection")) return rc;
  if (positive) OPY_LAUNCH((signproj_kernel<T, true>), n, s, x, y, n, add);
  else OPY_LAUNCH((signproj_kernel<T, false>), n, s, x, y, n, add);
  return check_launch("sign projection");
}
OPY_API int opy_signproj_f32(const float* x, float* y, int64_t n, int positive, int add, void* s) {
  return signproj_impl<float>(x, y, n, positive, add, s);
}
OPY_API int opy_signproj_f64(const double* x, double* y, int64_t n, int positive, int add, void* s) {
  return signproj_impl<double>(x, y, n, positive, add, s);
}

template <typename T>
static int const_impl(T* y, int64_t n, const T* c, int add, void* s) {
  if (int rc = check_n<T>(y, c, n, "const")) return rc;
  OPY_LAUNCH((const_kernel<T>), n, s, y, n, c, add);
  return check_launch("const");
}
OPY_API int opy_const_f32(float* y, int64_t n, const float* c, int add, void* s) {
  return const_impl<float>(y, n, c, add, s);
}
OPY_API int opy_const_f64(double* y, int64_t n, const double* c, int add, void* s) {
  return const_impl<double>(y, n, c, add, s);
}

template <typename T>
static int constrain_impl(const T* x, T* y, int64_t n, int64_t M, const T* c, const uint8_t* mask, int add,
                          void* s) {
  if (int rc = check_n<T>(x, y, n, "constrain")) return rc;
  if (M < 1 || (n > 0 && (!c || !mask))) return fail(kInvalid, "constrain: bad mask or constant");
  OPY_LAUNCH((constrain_kernel<T>), n, s, x, y, n, M, c, mask, add);
  return check_launch("constrain");
}
OPY_API int opy_constrain_f32(const float* x, float* y, int64_t n, int64_t M, const float* c,
                              const uint8_t* mask, int add, void* s) {
  return constrain_impl<float>(x, y, n, M, c, mask, add, s);
}
OPY_API int opy_constrain_f64(const double* x, double* y, int64_t n, int64_t M, const double* c,
                              const uint8_t* mask, int add, void* s) {
  return constrain_impl<double>(x, y, n, M, c, mask, add, s);
}

// ---- PDHG elementwise steps ------------------------------------------------
template <typename T>
static int take_step_impl(T* zn, const T* z, int64_t n, int dir, const opy_step_t* st, const T* g, void* s) {
  if (int rc = check_n<T>(zn, z, n, "take_step")) return rc;
  if (n > 0 && !g) return fail(kInvalid, "take_step: null direction");
  OPY_LAUNCH((take_step_kernel<T>), n, s, zn, z, n, dir, make_step<T>(st), g);
  return check_launch("take_step");
}
OPY_API int opy_pdhg_take_step_f32(float* zn, const float* z, int64_t n, int dir, const opy_step_t* st,
                                   const float* g, void* s) {
  return take_step_impl<float>(zn, z, n, dir, st, g, s);
}
OPY_API int opy_pdhg_take_step_f64(double* zn, const double* z, int64_t n, int dir, const opy_step_t* st,
                                   const double* g, void* s) {
  return take_step_impl<double>(zn, z, n, dir, st, g, s);
}

template <typename T>
static int dual_step_impl(T* yn, const T* y, int64_t n, double theta, const opy_step_t* st, const T* axn,
                          const T* ax, void* s) {
  if (int rc = check_n<T>(yn, y, n, "dual_step")) return rc;
  if (n > 0 && (!axn || !ax)) return fail(kInvalid, "dual_step: null operand");
  OPY_LAUNCH((dual_step_kernel<T>), n, s, yn, y, n, (T)theta, make_step<T>(st), axn, ax);
  return check_launch("dual_step");
}
OPY_API int opy_pdhg_dual_step_f32(float* yn, const float* y, int64_t n, double theta, const opy_step_t* st,
                                   const float* axn, const float* ax, void* s) {
  return dual_step_impl<float>(yn, y, n, theta, st, axn, ax, s);
}
OPY_API int opy_pdhg_dual_step_f64(double* yn, const double* y, int64_t n, double theta, const opy_step_t* st,
                                   const double* axn, const double* ax, void* s) {
  return dual_step_impl<double>(yn, y, n, theta, st, axn, ax, s);
}
