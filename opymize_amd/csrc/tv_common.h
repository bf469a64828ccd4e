// Shared helpers of the fused TV kernels (tv_passes.hip, tv_iter.hip,
// tv_multi.hip, tv_control.hip): configuration, tile map, vector loads and streaming
// stores, per-pixel stencils, cross-lane moves and the host-side launch
// helpers.  Everything sits in an unnamed namespace: each translation
// unit gets its own copy.
//
// Fused PDHG iterations for TV-regularised 2-D denoising, batched over many
// instances of uneven size.
//
// Model per instance b (image H x W with C channels, forward differences):
//   min_x  G_b(x) + tvw_b * sum_p |grad x (p)|_F
//   G_b = 0.5*lam_b*|x - f|^2 (SSD, reference functionals/quad.py) or
//         lam_b*|x - f|_1     (TV-L1 data term)
// i.e. PDHG(G, L1Norms(N, (C, 2), tvw), GradientOp(scheme="forward")) of the
// reference (pdhg.py:124-187, diff.cu:64-71, :114-123, proj.cu:34-44).
//
// An iteration is one pass over HBM (tv_iter.hip) or two (tv_passes.hip), or
// K iterations share one pass (tv_multi.hip); the two passes are:
//   pass 1 (primal): x+ = prox_{tau G}(x - tau * A^T y)           reads x, y, f; writes x+
//   pass 2 (dual):   y+ = proj_{|.|<=tvw}(y + sigma * A(2x+ - x))  reads x, x+, y; writes y+
// The over-relaxed A x_bar = 2 A x+ - A x of the reference is applied as
// A(2x+ - x) (equal by linearity).  With adaptive steps pass 2 also forms the
// four residual norms of pdhg.py:155-164.
//
// Thread layout of the two-pass kernels: a tile is tw columns x (256/tw)*RPT
// rows (the single-pass and multi-iteration kernels use tw = 64); a thread owns one
// column segment of RPT consecutive rows and marches down it, carrying the
// row above / below in registers.  Horizontal neighbours come from the
// neighbouring lane by cross-lane shuffles; only the first/last lane of a row
// segment of a wavefront loads (or, for y+, recomputes) its neighbour.
//
// The kernels walk the tiles of the still-running instances with a grid
// stride: active_inst / act_off hold the running instances and the exclusive
// prefix of their tile counts, and block k takes tiles k, k + grid, ...  The
// solver launches one block per running tile (rounded up), so the hardware
// dispatcher keeps the tiles in flight contiguous in memory (DRAM page
// locality, halos shared through L2); a persistent grid of resident blocks,
// whose blocks drift apart, measured 14% slower with 1.34x the reads.
// Blocks are renumbered so that consecutive tiles run on the same XCD (the
// dispatcher deals blocks round-robin to the 8 XCDs, each with its own L2).
// The single-pass kernel (tv_iter_kernel) forms its halos "transposed", and
// the multi-iteration kernel (tv_multi_kernel) stages a tile with a K-pixel
// halo in LDS; see their comments.  All per-instance reductions (residuals,
// gap) go through per-tile FP64 partials that a per-instance kernel folds in
// a fixed order (reproducible results).
#pragma once

#include "opy_common.h"

#include <cmath>
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

namespace opy {
namespace {

constexpr int kTB = 256;
constexpr double kEps = 2.220446049250313e-16;

struct Cfg {
  int C, tw, rpt, steps;
  double ih0, ih1;
  double w[4];
};

inline Cfg make_cfg(const opy_tv_cfg_t* c) {
  Cfg r;
  r.C = c->C;
  r.tw = c->tw;
  r.rpt = c->rpt;
  r.steps = c->steps;
  r.ih0 = 1.0 / c->h0;
  r.ih1 = 1.0 / c->h1;
  for (int k = 0; k < 4; ++k) r.w[k] = c->w[k];
  return r;
}

struct TileRef {
  int b;  // instance
  int j;  // tile index within the instance
  int r0, c0;
};

__device__ inline int xcd_block() {
  // blocks are dealt round-robin to the 8 XCDs: give each XCD a contiguous range
  const int g = gridDim.x;
  return (g % 8 == 0) ? (blockIdx.x % 8) * (g / 8) + blockIdx.x / 8 : blockIdx.x;
}

// tile g of the running instances -> (instance, tile index, first row, first
// column): search of g in act_off; tiles are tw image columns by th rows,
// numbered row by row within an instance
__device__ inline TileRef locate(int g, const int32_t* __restrict__ active_inst, const int32_t* __restrict__ act_off,
                                 int nact, const opy_tv_inst_t* __restrict__ insts, int tw, int th) {
  // largest a with act_off[a] <= g: first the proportional guess (exact when
  // the running instances have equal tile counts; its two loads are
  // independent), else a binary search (dependent loads)
  int lo = min((int)((int64_t)g * nact / act_off[nact]), nact - 1);
  if (!(act_off[lo] <= g && g < act_off[lo + 1])) {
    int hi = nact - 1;
    lo = 0;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (act_off[mid] <= g) lo = mid;
      else hi = mid - 1;
    }
  }
  TileRef t;
  t.b = active_inst[lo];
  t.j = g - act_off[lo];
  const int W = insts[t.b].W;
  const int ntx = (W + tw - 1) / tw;
  const int ty = t.j / ntx;
  t.r0 = ty * th;
  t.c0 = (t.j - ty * ntx) * tw;
  return t;
}

template <typename T>
__device__ inline T shfl_up1(T v) { return __shfl_up(v, 1, 64); }
template <typename T>
__device__ inline T shfl_down1(T v) { return __shfl_down(v, 1, 64); }

template <typename T>
__device__ inline T precond_tau(int r, int c, int H, int W, int k, const Cfg& cfg) {
  // 1 / column sum of |A| (reference pdhg.py:197-200 via rowwise_lp of the divergence)
  const double s = cfg.w[k] * (((r < H - 1) + (r > 0)) * cfg.ih0 + ((c < W - 1) + (c > 0)) * cfg.ih1);
  return (T)(s > kEps ? 1.0 / s : s);
}

template <typename T>
__device__ inline T precond_sigma(bool inner, int k, double ih, const Cfg& cfg) {
  const double s = inner ? 2.0 * cfg.w[k] * ih : 0.0;
  return (T)(s > kEps ? 1.0 / s : s);
}

// Over-relaxation of the iteration: theta = 1/sqrt(1 + 2 lam tau) with
// accelerated steps on the lam-strongly convex SSD term (Chambolle & Pock
// 2011, Algorithm 2), else 1 (the reference's theta, pdhg.py:145-153).
__device__ inline double accel_theta(int steps, const opy_tv_inst_t& I, double tau) {
  return (steps == OPY_TV_STEPS_ACCEL && I.data == OPY_TV_DATA_L2) ? 1.0 / sqrt(1.0 + 2.0 * I.lam * tau) : 1.0;
}

// Products whose rounded value is an iterate (x+ of the SSD prox, the
// projected y+) are formed with contraction off, so that the compiler never
// fuses them into a later addition: every fused kernel rounds the same values
// the same way, also where several iterations share one basic block.
template <typename T>
__device__ inline T mul_rn(T a, T b) {
#pragma clang fp contract(off)
  return a * b;
}

// soft threshold d - clamp(d, -b, b) (b >= 0): d - b above b, d + b below
// -b, +0 between, without branches (for finite d the same values as the
// three-way select)
template <typename T>
__device__ inline T soft_rest(T d, T b) {
  if constexpr (sizeof(T) == 4) return d - __builtin_fminf(__builtin_fmaxf(d, -b), b);
  else return d - __builtin_fmin(__builtin_fmax(d, -b), b);
}

// |v|^2 of a vector of 2C components (the axis-0 ones first), in one fixed
// order of roundings: per channel k the fused v_k^2 + round(v_{C+k}^2),
// then the channels' sum
template <typename T, int C>
__device__ inline T sqnorm2(const T* v) {
  T n2 = T(0);
#pragma unroll
  for (int k = 0; k < C; ++k) {
    const T t = sizeof(T) == 4 ? (T)__builtin_fmaf((float)v[k], (float)v[k], (float)mul_rn(v[C + k], v[C + k]))
                               : (T)__builtin_fma((double)v[k], (double)v[k], (double)mul_rn(v[C + k], v[C + k]));
    n2 = k == 0 ? t : n2 + t;
  }
  return n2;
}

// x_bar = (1 + theta) x+ - theta x, as one fused multiply-add of theta1 =
// 1 + theta and x+ with the rounded product theta x
template <typename T>
__device__ inline T overrelax(T theta1, T theta, T xp, T x) {
  if constexpr (sizeof(T) == 4) return __builtin_fmaf(theta1, xp, -mul_rn(theta, x));
  else return __builtin_fma(theta1, xp, -mul_rn(theta, x));
}

// x+ = prox_{tau G}(x - tau (A^T y)) for the data terms: SSD gives
// (v + tl f) / (1 + tl) with tl = tau lam, formed as a * (v + b f) with
// a = 1/(1 + tl) and b = tl rounded to T; L1 gives f + soft(v - f, tl).
// The coefficients depend on tau only, so they are formed once per tile:
// for the instance's scalar step, or with preconditioned steps for an
// interior pixel (both neighbours present along both axes, the same tau for
// all of them); other pixels of a preconditioned run form theirs (an FP64
// division each) where they occur.
// PC: preconditioned steps (a template parameter, so that the scalar-step
// kernels carry none of their selects).
template <typename T, int C, bool PC>
struct PrimalStep {
  T tau[C], a[C], b[C];
  double lam;
  int data;
  __device__ PrimalStep(double tau_s, const opy_tv_inst_t& I, const Cfg& cfg) : lam(I.lam), data(I.data) {
#pragma unroll
    for (int k = 0; k < C; ++k) coef(PC ? precond_tau<T>(1, 1, 3, 3, k, cfg) : (T)tau_s, tau[k], a[k], b[k]);
  }
  __device__ void coef(T t, T& tk, T& ak, T& bk) const {
    const double tl = (double)t * lam;
    tk = t;
    bk = (T)tl;
    ak = data == OPY_TV_DATA_L1 ? T(1) : (T)(1.0 / (1.0 + tl));
  }
  // x+ at pixel (r, c), channel k, from x, (A^T y) and f there; L1: the
  // data kind as a compile-time constant; M = false: the pixel is known to
  // be interior
  template <bool L1, bool M = true>
  __device__ T apply(T x, T aty, T f, int r, int c, int H, int W, int k, const Cfg& cfg) const {
    T t = tau[k], ak = a[k], bk = b[k];
    if (PC && M && !(r > 0 && r < H - 1 && c > 0 && c < W - 1)) coef(precond_tau<T>(r, c, H, W, k, cfg), t, ak, bk);
    const T v = x - t * aty;
    if constexpr (L1) {
      const T d = v - f;
      return f + soft_rest(d, bk);
    } else {
      return mul_rn(ak, v + bk * f);
    }
  }
  __device__ T operator()(T x, T aty, T f, int r, int c, int H, int W, int k, const Cfg& cfg) const {
    return data == OPY_TV_DATA_L1 ? apply<true>(x, aty, f, r, c, H, W, k, cfg)
                                  : apply<false>(x, aty, f, r, c, H, W, k, cfg);
  }
};

// dual step per component e of y (e < C: along axis 0, else axis 1): sigma,
// or the preconditioner of a pixel with a neighbour along that axis (0 at
// the last row / column), formed once per tile
template <typename T, int C, bool PC>
struct DualStep {
  T s[2 * C];
  __device__ DualStep(T sigma, const Cfg& cfg) {
#pragma unroll
    for (int k = 0; k < C; ++k) {
      s[k] = PC ? precond_sigma<T>(true, k, cfg.ih0, cfg) : sigma;
      s[C + k] = PC ? precond_sigma<T>(true, k, cfg.ih1, cfg) : sigma;
    }
  }
  template <bool M = true>
  __device__ T at(int e, int r, int c, int H, int W) const {
    return (PC && M && !(e < C ? r < H - 1 : c < W - 1)) ? T(0) : s[e];
  }
};

// tvw / |v| for |v|^2 = n2 > tvw^2 > 0: one hardware reciprocal square root
// (v_rsq_f32, within 1 ulp) in FP32, the correctly rounded quotient in FP64
__device__ inline float ball_scale(float tvw, float n2) { return tvw * __builtin_amdgcn_rsqf(n2); }
__device__ inline double ball_scale(double tvw, double n2) { return tvw / sqrt(n2); }

// N consecutive values from an address aligned to min(N, 4) * sizeof(T) for
// even N (per-pixel vectors inside instance blocks whose offsets are
// multiples of the vector length): one vector load instead of N scalar ones.
template <typename T, int N>
__device__ inline void load_c(const T* __restrict__ p, T* v) {
  if constexpr (sizeof(T) == 4 && N % 4 == 0) {
#pragma unroll
    for (int k = 0; k < N; k += 4) {
      const float4 q = *reinterpret_cast<const float4*>(p + k);
      v[k] = q.x;
      v[k + 1] = q.y;
      v[k + 2] = q.z;
      v[k + 3] = q.w;
    }
  } else if constexpr (N % 2 == 0) {
    using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
#pragma unroll
    for (int k = 0; k < N; k += 2) {
      const V2 q = *reinterpret_cast<const V2*>(p + k);
      v[k] = q.x;
      v[k + 1] = q.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = p[k];
  }
}

// Streaming stores of the new iterates (nontemporal: at batch sizes beyond
// the last-level cache they are not re-read from it), vectorised like load_c.
template <typename T, int N>
__device__ inline void store_c(T* __restrict__ p, const T* v) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  typedef T v4 __attribute__((ext_vector_type(4)));
  if constexpr (sizeof(T) == 4 && N % 4 == 0) {
#pragma unroll
    for (int k = 0; k < N; k += 4) {
      const v4 q = {v[k], v[k + 1], v[k + 2], v[k + 3]};
      __builtin_nontemporal_store(q, reinterpret_cast<v4*>(p + k));
    }
  } else if constexpr (N % 2 == 0) {
#pragma unroll
    for (int k = 0; k < N; k += 2) {
      const v2 q = {v[k], v[k + 1]};
      __builtin_nontemporal_store(q, reinterpret_cast<v2*>(p + k));
    }
  } else {
#pragma unroll
    for (int k = 0; k < N; ++k) __builtin_nontemporal_store(v[k], p + k);
  }
}

// sum of four doubles over a 256-thread block, one LDS round; sm is used by
// tile parity so that consecutive tiles need one barrier each
__device__ inline void block_sum4(double (&s)[4], double (*sm)[4], int parity) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j] += __shfl_down(s[j], off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* buf = sm[parity * (kTB / 64) + wave];
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < 4; ++j) buf[j] = s[j];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double a = 0.0;
      for (int w = 0; w < kTB / 64; ++w) a += sm[parity * (kTB / 64) + w][j];
      s[j] = a;
    }
  }
}

// (A^T y)_k at a pixel from y(p) (t=0: y0, t=1: y1), y0 of the pixel above
// and y1 of the pixel to the left.  M = false: the pixel is known to have
// all four neighbours (the same operations without the boundary selects).
template <typename T, int C, bool M = true>
__device__ inline void adjoint_px(const T* y0, const T* y1, const T* yup0, const T* yleft1, int r, int c, int H, int W,
                                  const Cfg& cfg, T* out) {
  const T ih0 = (T)cfg.ih0, ih1 = (T)cfg.ih1;
#pragma unroll
  for (int k = 0; k < C; ++k) {
    T d = T(0);
    if (!M || r < H - 1) d -= y0[k] * ih0;
    if (!M || r > 0) d += yup0[k] * ih0;
    if (!M || c < W - 1) d -= y1[k] * ih1;
    if (!M || c > 0) d += yleft1[k] * ih1;
    out[k] = (T)cfg.w[k] * d;
  }
}

// gradient at a pixel from its value, the value below and the value to the right
template <typename T, int C, bool M = true>
__device__ inline void grad_px(const T* v, const T* vdown, const T* vright, int r, int c, int H, int W, const Cfg& cfg,
                               T* g) {
#pragma unroll
  for (int k = 0; k < C; ++k) {
    g[k] = (!M || r < H - 1) ? (T)(cfg.w[k] * cfg.ih0) * (vdown[k] - v[k]) : T(0);
    g[C + k] = (!M || c < W - 1) ? (T)(cfg.w[k] * cfg.ih1) * (vright[k] - v[k]) : T(0);
  }
}

// y+ = proj_{|.| <= tvw}(y + sigma * grad(xbar)) at one pixel; g receives grad(xbar)
template <typename T, int C, bool PC, bool M = true>
__device__ inline void yplus_px(const T* xb, const T* xbd, const T* xbr, const T* y, int r, int c, int H, int W,
                                const DualStep<T, C, PC>& ds, T tvw, const Cfg& cfg, T* yp, T* g) {
  grad_px<T, C, M>(xb, xbd, xbr, r, c, H, W, cfg, g);
#pragma unroll
  for (int e = 0; e < 2 * C; ++e) yp[e] = y[e] + ds.template at<M>(e, r, c, H, W) * g[e];
  const T n2 = sqnorm2<T, C>(yp);
  if constexpr (sizeof(T) == 4) {
    // branch-free: the scale is 1 (exact) inside the ball
    const T sc = n2 > tvw * tvw ? ball_scale(tvw, n2) : T(1);
#pragma unroll
    for (int e = 0; e < 2 * C; ++e) yp[e] = mul_rn(yp[e], sc);
  } else if (n2 > tvw * tvw) {
    const T sc = ball_scale(tvw, n2);
#pragma unroll
    for (int e = 0; e < 2 * C; ++e) yp[e] = mul_rn(yp[e], sc);
  }
}

// Scalar steps of one iteration (constant and accelerated step rules) in the
// form the folded stencils below use: the prox coefficients a, b of
// PrimalStep (from tau rounded to T), theta, and the step sizes folded into
// the difference weights, p = tau w_k / h and q = sigma w_k / h per axis,
// each rounded once from FP64.  Then x - tau A^T y costs one subtraction and
// one fused multiply-add per difference, and so does y + sigma A xbar, where
// PrimalStep/DualStep with adjoint_px/grad_px spend a further product on the
// weight and one on the step (preconditioned steps vary per pixel and
// adaptive steps need A^T y and A xbar for their residuals, so those keep
// the unfolded form).
template <typename T, int C>
struct FoldedStep {
  T a, b, theta;
  T p[2][C], q[2][C];
};

// tau, sigma: this iteration's primal and dual steps (sigma already divided by theta)
template <typename T, int C>
__device__ inline FoldedStep<T, C> fold_step(double tau, double sigma, double theta, const opy_tv_inst_t& I,
                                             const Cfg& cfg) {
  FoldedStep<T, C> s;
  const PrimalStep<T, C, false> ps(tau, I, cfg);
  s.a = ps.a[0];
  s.b = ps.b[0];
  s.theta = (T)theta;
#pragma unroll
  for (int k = 0; k < C; ++k) {
    s.p[0][k] = (T)(tau * cfg.w[k] * cfg.ih0);
    s.p[1][k] = (T)(tau * cfg.w[k] * cfg.ih1);
    s.q[0][k] = (T)(sigma * cfg.w[k] * cfg.ih0);
    s.q[1][k] = (T)(sigma * cfg.w[k] * cfg.ih1);
  }
  return s;
}

// x+ = prox_{tau G}(x - tau A^T y) at pixel (r, c), channel k, from x, f and
// y = (y0, y1) there, y0 of the pixel above and y1 of the pixel to the left
// (the terms adjoint_px masks at the image boundary are dropped by selects:
// the row tests when MR, the column tests when MC); L1: the data kind
template <typename T, int C, bool L1, bool MR, bool MC>
__device__ inline T primal_px_rc(T x, T f, T y0, T y1, T yup0, T yleft1, int r, int c, int H, int W, int k,
                                 const FoldedStep<T, C>& s) {
  if (MR) {
    if (!(r < H - 1)) y0 = T(0);
    if (!(r > 0)) yup0 = T(0);
  }
  if (MC) {
    if (!(c < W - 1)) y1 = T(0);
    if (!(c > 0)) yleft1 = T(0);
  }
  const T v = x + (y0 - yup0) * s.p[0][k] + (y1 - yleft1) * s.p[1][k];
  if constexpr (L1) {
    const T d = v - f;
    return f + soft_rest(d, s.b);
  } else {
    return mul_rn(s.a, v + s.b * f);
  }
}
template <typename T, int C, bool L1, bool M>
__device__ inline T primal_px(T x, T f, T y0, T y1, T yup0, T yleft1, int r, int c, int H, int W, int k,
                              const FoldedStep<T, C>& s) {
  return primal_px_rc<T, C, L1, M, M>(x, f, y0, y1, yup0, yleft1, r, c, H, W, k, s);
}

// y+ = proj_{|.| <= tvw}(y + sigma A xbar) at pixel (r, c) from xbar there,
// below and to the right (differences beyond the image are 0: down when MR,
// right when MC)
template <typename T, int C, bool MR, bool MC>
__device__ inline void dual_px_rc(const T* xb, const T* xbd, const T* xbr, const T* y, int r, int c, int H, int W,
                                  const FoldedStep<T, C>& s, T tvw, T* yp) {
#pragma unroll
  for (int k = 0; k < C; ++k) {
    T gd = xbd[k] - xb[k], gr = xbr[k] - xb[k];
    if (MR && !(r < H - 1)) gd = T(0);
    if (MC && !(c < W - 1)) gr = T(0);
    yp[k] = y[k] + gd * s.q[0][k];
    yp[C + k] = y[C + k] + gr * s.q[1][k];
  }
  const T n2 = sqnorm2<T, C>(yp);
  if constexpr (sizeof(T) == 4) {
    // min(1, tvw / |v|): 1 inside the ball (and for v = 0, where the
    // reciprocal square root is +inf)
    const T sc = __builtin_fminf(T(1), ball_scale(tvw, n2));
#pragma unroll
    for (int e = 0; e < 2 * C; ++e) yp[e] = mul_rn(yp[e], sc);
  } else if (n2 > tvw * tvw) {
    const T sc = ball_scale(tvw, n2);
#pragma unroll
    for (int e = 0; e < 2 * C; ++e) yp[e] = mul_rn(yp[e], sc);
  }
}
template <typename T, int C, bool M>
__device__ inline void dual_px(const T* xb, const T* xbd, const T* xbr, const T* y, int r, int c, int H, int W,
                               const FoldedStep<T, C>& s, T tvw, T* yp) {
  dual_px_rc<T, C, M, M>(xb, xbd, xbr, y, r, c, H, W, s, tvw, yp);
}

// ---- one pass per iteration -------------------------------------------------
// Cross-lane moves.  DPP wave shifts (one VALU instruction): shr1 gives lane
// L the value of lane L-1 and lane 0 `edge`; shl1 gives lane L the value of
// lane L+1 and lane 63 `edge`.  readlane broadcasts lane l's value.
__device__ inline int dpp_shr1_i(int edge, int v) { return __builtin_amdgcn_update_dpp(edge, v, 0x138, 0xF, 0xF, false); }
__device__ inline int dpp_shl1_i(int edge, int v) { return __builtin_amdgcn_update_dpp(edge, v, 0x130, 0xF, 0xF, false); }
__device__ inline float dpp_shr1(float edge, float v) { return __int_as_float(dpp_shr1_i(__float_as_int(edge), __float_as_int(v))); }
__device__ inline float dpp_shl1(float edge, float v) { return __int_as_float(dpp_shl1_i(__float_as_int(edge), __float_as_int(v))); }
__device__ inline double dpp_shr1(double edge, double v) {
  return __hiloint2double(dpp_shr1_i(__double2hiint(edge), __double2hiint(v)),
                          dpp_shr1_i(__double2loint(edge), __double2loint(v)));
}
__device__ inline double dpp_shl1(double edge, double v) {
  return __hiloint2double(dpp_shl1_i(__double2hiint(edge), __double2hiint(v)),
                          dpp_shl1_i(__double2loint(edge), __double2loint(v)));
}
__device__ inline float readlane(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ inline double readlane(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

// ---- host wrappers ---------------------------------------------------------
#define KF(...) reinterpret_cast<const void*>(&__VA_ARGS__)

// Blocks of `kernel` resident at once on the current device (occupancy x
// CUs), cached per kernel and device: the persistent loops need no more.
int resident_blocks(const void* kernel) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = -1;
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_pair(kernel, dev);
  const auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kTB, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  if (dev < 0 || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 256;
  const int g = per_cu * cus;
  cache[key] = g;
  return g;
}

// grid > 0: as given; grid == 0: the resident block count of the kernel
inline int launch_grid(const void* kernel, int grid) { return grid > 0 ? grid : resident_blocks(kernel); }

int check_cfg(const opy_tv_cfg_t* c) {
  if (!c) return fail(kInvalid, "tv: null config");
  if (c->C < 1 || c->C > 4) return fail(kInvalid, "tv: C=%d outside [1,4]", c->C);
  if (c->tw < 1 || c->tw > kTB || kTB % c->tw) return fail(kInvalid, "tv: tile width %d must divide %d", c->tw, kTB);
  if (c->rpt != 1 && c->rpt != 2 && c->rpt != 4 && c->rpt != 6 && c->rpt != 8 && c->rpt != 12 && c->rpt != 16)
    return fail(kInvalid, "tv: rows per thread %d not in {1,2,4,6,8,12,16}", c->rpt);
  if (c->C == 4 && c->rpt == 16)
    return fail(kInvalid, "tv: C=4 with 16 rows per thread exceeds the register budget (at most 12)");
  if (c->steps != OPY_TV_STEPS_SCALAR && c->steps != OPY_TV_STEPS_PRECOND && c->steps != OPY_TV_STEPS_ACCEL)
    return fail(kInvalid, "tv: bad step mode");
  if (!(c->h0 > 0) || !(c->h1 > 0)) return fail(kInvalid, "tv: grid spacing must be positive");
  return kOk;
}

// dispatch a kernel template over C (1..4) and RPT (1,2,4,6,8,12,16)
#define OPY_TV_DISPATCH_C(KERNEL_EXPR_C_R, CC)       \
  switch (cfg->rpt) {                                \
    case 1: KERNEL_EXPR_C_R(CC, 1); break;           \
    case 2: KERNEL_EXPR_C_R(CC, 2); break;           \
    case 4: KERNEL_EXPR_C_R(CC, 4); break;           \
    case 6: KERNEL_EXPR_C_R(CC, 6); break;           \
    case 8: KERNEL_EXPR_C_R(CC, 8); break;           \
    case 12: KERNEL_EXPR_C_R(CC, 12); break;         \
    default: KERNEL_EXPR_C_R(CC, 16); break;         \
  }
#define OPY_TV_DISPATCH(KERNEL_EXPR_C_R)                     \
  switch (cfg->C) {                                          \
    case 1: OPY_TV_DISPATCH_C(KERNEL_EXPR_C_R, 1); break;    \
    case 2: OPY_TV_DISPATCH_C(KERNEL_EXPR_C_R, 2); break;    \
    case 3: OPY_TV_DISPATCH_C(KERNEL_EXPR_C_R, 3); break;    \
    default: OPY_TV_DISPATCH_C(KERNEL_EXPR_C_R, 4); break;   \
  }

}  // namespace
}  // namespace opy
