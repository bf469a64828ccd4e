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
// One iteration is two passes over HBM:
//   pass 1 (primal): x+ = prox_{tau G}(x - tau * A^T y)           reads x, y, f; writes x+
//   pass 2 (dual):   y+ = proj_{|.|<=tvw}(y + sigma * A(2x+ - x))  reads x, x+, y; writes y+
// The over-relaxed A x_bar = 2 A x+ - A x of the reference is applied as
// A(2x+ - x) (equal by linearity).  With adaptive steps pass 2 also forms the
// four residual norms of pdhg.py:155-164.
//
// Thread layout: a tile is tw columns x (256/tw)*RPT rows; a thread owns one
// column segment of RPT consecutive rows and marches down it, carrying the
// row above / below in registers.  Horizontal neighbours come from the
// neighbouring lane by cross-lane shuffles; only the first/last lane of a row
// segment of a wavefront loads (or, for y+, recomputes) its neighbour.
//
// A persistent grid walks the tiles of the still-running instances:
// active_inst / act_off hold the running instances and the exclusive prefix
// of their tile counts; block k takes tiles k, k + grid, k + 2 grid, ... so
// that the blocks running at one time work on neighbouring tiles (DRAM page
// locality, shared halos; contiguous chunks per block measured slower).
// Blocks are renumbered so that consecutive tiles run on the same XCD (the
// dispatcher deals blocks round-robin to the 8 XCDs, each with its own L2),
// which keeps the halo rows between tiles in one L2.
// The single-pass kernel (tv_iter_kernel) forms its halos "transposed", see
// its comment.  All
// per-instance reductions (residuals, gap) go through per-tile FP64 partials
// that a per-instance kernel folds in a fixed order (reproducible results).
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
// column): binary search of g in act_off; tiles are tw image columns by th
// rows, numbered row by row within an instance
__device__ inline TileRef locate(int g, const int32_t* __restrict__ active_inst, const int32_t* __restrict__ act_off,
                                 int nact, const opy_tv_inst_t* __restrict__ insts, int tw, int th) {
  int lo = 0, hi = nact - 1;  // largest a with act_off[a] <= g
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (act_off[mid] <= g) lo = mid;
    else hi = mid - 1;
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

template <typename T>
__device__ inline T prox_data(T v, T f, double tl, int data) {
  if (data == OPY_TV_DATA_L1) {
    const T thr = (T)tl;
    const T d = v - f;
    return f + (d > thr ? d - thr : (d < -thr ? d + thr : T(0)));
  }
  const T a = (T)(1.0 / (1.0 + tl));
  const T b = (T)tl;
  return a * (v + b * f);
}

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

// (A^T y)_k at a pixel from y(p) (t=0: y0, t=1: y1), y0 of the pixel above and y1 of the pixel to the left
template <typename T, int C>
__device__ inline void adjoint_px(const T* y0, const T* y1, const T* yup0, const T* yleft1, int r, int c, int H, int W,
                                  const Cfg& cfg, T* out) {
  const T ih0 = (T)cfg.ih0, ih1 = (T)cfg.ih1;
#pragma unroll
  for (int k = 0; k < C; ++k) {
    T d = T(0);
    if (r < H - 1) d -= y0[k] * ih0;
    if (r > 0) d += yup0[k] * ih0;
    if (c < W - 1) d -= y1[k] * ih1;
    if (c > 0) d += yleft1[k] * ih1;
    out[k] = (T)cfg.w[k] * d;
  }
}

// gradient at a pixel from its value, the value below and the value to the right
template <typename T, int C>
__device__ inline void grad_px(const T* v, const T* vdown, const T* vright, int r, int c, int H, int W, const Cfg& cfg,
                               T* g) {
#pragma unroll
  for (int k = 0; k < C; ++k) {
    g[k] = (r < H - 1) ? (T)(cfg.w[k] * cfg.ih0) * (vdown[k] - v[k]) : T(0);
    g[C + k] = (c < W - 1) ? (T)(cfg.w[k] * cfg.ih1) * (vright[k] - v[k]) : T(0);
  }
}

// y+ = proj_{|.| <= tvw}(y + sigma * grad(xbar)) at one pixel; g receives grad(xbar)
template <typename T, int C>
__device__ inline void yplus_px(const T* xb, const T* xbd, const T* xbr, const T* y, int r, int c, int H, int W, T sigma,
                                T tvw, const Cfg& cfg, T* yp, T* g) {
  grad_px<T, C>(xb, xbd, xbr, r, c, H, W, cfg, g);
  T n2 = T(0);
#pragma unroll
  for (int e = 0; e < 2 * C; ++e) {
    T s = sigma;
    if (cfg.steps == OPY_TV_STEPS_PRECOND)
      s = (e < C) ? precond_sigma<T>(r < H - 1, e, cfg.ih0, cfg) : precond_sigma<T>(c < W - 1, e - C, cfg.ih1, cfg);
    yp[e] = y[e] + s * g[e];
    n2 += yp[e] * yp[e];
  }
  if (n2 > tvw * tvw) {
    const T sc = tvw / sqrt(n2);
#pragma unroll
    for (int e = 0; e < 2 * C; ++e) yp[e] *= sc;
  }
}

// ---- pass 1 ---------------------------------------------------------------
template <typename T, int C, int RPT>
__global__ void __launch_bounds__(kTB) tv_primal_kernel(const T* __restrict__ xin, T* __restrict__ xout,
                                                        const T* __restrict__ y, const T* __restrict__ f,
                                                        const opy_tv_inst_t* __restrict__ insts,
                                                        const opy_adapt_state_t* __restrict__ st,
                                                        const int32_t* __restrict__ active_inst,
                                                        const int32_t* __restrict__ act_off,
                                                        const int32_t* __restrict__ nact_p, Cfg cfg) {
  const int nact = *nact_p;
  if (nact <= 0) return;
  const int total = act_off[nact];
  const int tw = cfg.tw;
  const int th = (kTB / tw) * RPT;
  const int lane = threadIdx.x & 63;
  const int tx = threadIdx.x % tw, ty = threadIdx.x / tw;
  const bool lhalo = (tx == 0) || (lane == 0);
  for (int g = xcd_block(); g < total; g += gridDim.x) {
    const TileRef tr = locate(g, active_inst, act_off, nact, insts, tw, th);
    const opy_tv_inst_t I = insts[tr.b];
    const int H = I.H, W = I.W;
    const double tau_s = st[tr.b].tau;
    const int c = tr.c0 + tx;
    const bool col = c < W;
    const int rs = tr.r0 + ty * RPT;
    const T* xb = xin + I.off;
    const T* fb = f + I.off;
    const T* yb = y + 2 * I.off;
    T* ob = xout + I.off;
    // every load of the segment is issued before any arithmetic, so that a
    // wavefront keeps 3*RPT+1 (+RPT halo) loads in flight
    T yu[C];            // t=0 part of y in the row above the segment
    T yv[RPT][2 * C];   // y of the own rows
    T xv[RPT][C], fv[RPT][C];
    T yl[RPT][C];       // t=1 part of y of the left neighbour (halo lanes)
#pragma unroll
    for (int k = 0; k < C; ++k) yu[k] = T(0);
#pragma unroll
    for (int i = 0; i < RPT; ++i)
#pragma unroll
      for (int k = 0; k < C; ++k) yv[i][k] = yv[i][C + k] = xv[i][k] = fv[i][k] = yl[i][k] = T(0);
    if (col && rs > 0 && rs < H) load_c<T, C>(yb + ((int64_t)(rs - 1) * W + c) * 2 * C, yu);
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int64_t pix = (int64_t)(rs + i) * W + c;
      if (col && rs + i < H) {
        load_c<T, 2 * C>(yb + pix * 2 * C, yv[i]);
        load_c<T, C>(xb + pix * C, xv[i]);
        load_c<T, C>(fb + pix * C, fv[i]);
      }
    }
    if (lhalo && col && c > 0) {
#pragma unroll
      for (int i = 0; i < RPT; ++i)
        if (rs + i < H) load_c<T, C>(yb + ((int64_t)(rs + i) * W + c - 1) * 2 * C + C, yl[i]);
    }
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = rs + i;
      const bool ok = col && r < H;
      const int64_t pix = (int64_t)r * W + c;
      T yl1[C];
#pragma unroll
      for (int k = 0; k < C; ++k) {
        const T s = shfl_up1(yv[i][C + k]);
        yl1[k] = lhalo ? yl[i][k] : s;
      }
      if (ok) {
        T d[C], o[C];
        adjoint_px<T, C>(yv[i], yv[i] + C, i == 0 ? yu : yv[i - 1], yl1, r, c, H, W, cfg, d);
#pragma unroll
        for (int k = 0; k < C; ++k) {
          const T tau = cfg.steps == OPY_TV_STEPS_PRECOND ? precond_tau<T>(r, c, H, W, k, cfg) : (T)tau_s;
          o[k] = prox_data<T>(xv[i][k] - tau * d[k], fv[i][k], (double)tau * I.lam, I.data);
        }
        store_c<T, C>(ob + pix * C, o);
      }
    }
  }
}

// ---- pass 2 ---------------------------------------------------------------
template <typename T, int C, int RPT, bool kAdapt>
__global__ void __launch_bounds__(kTB) tv_dual_kernel(const T* __restrict__ xo, const T* __restrict__ xn,
                                                      const T* __restrict__ yin, T* __restrict__ yout,
                                                      const opy_tv_inst_t* __restrict__ insts,
                                                      const opy_adapt_state_t* __restrict__ st,
                                                      const int32_t* __restrict__ active_inst,
                                                      const int32_t* __restrict__ act_off,
                                                      const int32_t* __restrict__ nact_p, Cfg cfg,
                                                      double* __restrict__ partials) {
  __shared__ double sm[2 * (kTB / 64)][4];
  int parity = 0;
  const int nact = *nact_p;
  if (nact <= 0) return;
  const int total = act_off[nact];
  const int tw = cfg.tw;
  const int th = (kTB / tw) * RPT;
  const int lane = threadIdx.x & 63;
  const int tx = threadIdx.x % tw, ty = threadIdx.x / tw;
  const bool lhalo = (tx == 0) || (lane == 0);
  const bool rhalo = (tx == tw - 1) || (lane == 63);
  for (int g = xcd_block(); g < total; g += gridDim.x) {
    const TileRef tr = locate(g, active_inst, act_off, nact, insts, tw, th);
    const opy_tv_inst_t I = insts[tr.b];
    const int H = I.H, W = I.W;
    const T sigma = (T)st[tr.b].sigma;
    const T tau = (T)st[tr.b].tau;
    const T tvw = (T)I.tvw;
    const int c = tr.c0 + tx;
    const bool col = c < W;
    const int rs = tr.r0 + ty * RPT;
    const T* xob = xo + I.off;
    const T* xnb = xn + I.off;
    const T* yib = yin + 2 * I.off;
    T* yob = yout + 2 * I.off;
    // Every load of the segment is issued before any arithmetic.  Row index
    // j of the arrays is image row rs-1+j (row rs-1 only with adaptive steps,
    // whose residual needs y+ of the row above); the left-neighbour arrays
    // start at row rs.
    constexpr int NR = RPT + 2;
    T ax[NR][C], bx[NR][C];             // x_k, x_k+1 -> later x_k - x_k+1, 2x_k+1 - x_k
    T yk[RPT + 1][2 * C];               // y_k
    T axr[RPT + 1][C], bxr[RPT + 1][C];  // right neighbour (right halo lanes), rows rs-1..
    T axl[RPT + 1][C], bxl[RPT + 1][C];  // left neighbour (left halo lanes, adaptive), rows rs..
    T ykl[RPT][2 * C];                  // y_k of the left neighbour (left halo lanes, adaptive)
#pragma unroll
    for (int j = 0; j < NR; ++j)
#pragma unroll
      for (int k = 0; k < C; ++k) ax[j][k] = bx[j][k] = T(0);
#pragma unroll
    for (int j = 0; j <= RPT; ++j)
#pragma unroll
      for (int k = 0; k < C; ++k) {
        yk[j][k] = yk[j][C + k] = T(0);
        axr[j][k] = bxr[j][k] = axl[j][k] = bxl[j][k] = T(0);
        if (j < RPT) ykl[j][k] = ykl[j][C + k] = T(0);
      }
    constexpr int j0 = kAdapt ? 0 : 1;
#pragma unroll
    for (int j = j0; j < NR; ++j) {
      const int r = rs - 1 + j;
      if (col && r >= 0 && r < H) {
        const int64_t p = (int64_t)r * W + c;
        load_c<T, C>(xob + p * C, ax[j]);
        load_c<T, C>(xnb + p * C, bx[j]);
        if (j <= RPT) load_c<T, 2 * C>(yib + p * 2 * C, yk[j]);
      }
    }
    if (rhalo && col && c + 1 < W) {
#pragma unroll
      for (int j = j0; j <= RPT; ++j) {
        const int r = rs - 1 + j;
        if (r >= 0 && r < H) {
          const int64_t p = (int64_t)r * W + c + 1;
          load_c<T, C>(xob + p * C, axr[j]);
          load_c<T, C>(xnb + p * C, bxr[j]);
        }
      }
    }
    if (kAdapt && lhalo && col && c > 0) {
#pragma unroll
      for (int j = 0; j <= RPT; ++j) {
        const int r = rs + j;
        if (r < H) {
          const int64_t p = (int64_t)r * W + c - 1;
          load_c<T, C>(xob + p * C, axl[j]);
          load_c<T, C>(xnb + p * C, bxl[j]);
          if (j < RPT) load_c<T, 2 * C>(yib + p * 2 * C, ykl[j]);
        }
      }
    }
    // (x_k, x_k+1) -> (xd, xbar) = (x_k - x_k+1, 2 x_k+1 - x_k); right neighbours
    // from the next lane unless loaded by a halo lane
#pragma unroll
    for (int j = j0; j < NR; ++j)
#pragma unroll
      for (int k = 0; k < C; ++k) {
        const T a = ax[j][k], b = bx[j][k];
        ax[j][k] = a - b;
        bx[j][k] = T(2) * b - a;
      }
#pragma unroll
    for (int j = j0; j <= RPT; ++j)
#pragma unroll
      for (int k = 0; k < C; ++k) {
        const T a = axr[j][k], b = bxr[j][k];
        const T sd = shfl_down1(ax[j][k]), sb = shfl_down1(bx[j][k]);
        axr[j][k] = rhalo ? a - b : sd;
        bxr[j][k] = rhalo ? T(2) * b - a : sb;
        if (kAdapt) {
          const T al = axl[j][k], bl = bxl[j][k];
          axl[j][k] = al - bl;
          bxl[j][k] = T(2) * bl - al;
        }
      }
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const T isig = T(1) / sigma, itau = T(1) / tau;
    T ypu[C];  // t=0 part of y+ in the row above (adaptive)
#pragma unroll
    for (int k = 0; k < C; ++k) ypu[k] = T(0);
    if (kAdapt) {
      T yp[2 * C], g[2 * C];
      yplus_px<T, C>(bx[0], bx[1], bxr[0], yk[0], rs - 1, c, H, W, sigma, tvw, cfg, yp, g);
#pragma unroll
      for (int k = 0; k < C; ++k) ypu[k] = yp[k];
    }
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = rs + i;
      const bool ok = col && r < H;
      const int64_t pix = (int64_t)r * W + c;
      T yp[2 * C], g[2 * C];
      yplus_px<T, C>(bx[i + 1], bx[i + 2], bxr[i + 1], yk[i + 1], r, c, H, W, sigma, tvw, cfg, yp, g);
      if (ok) store_c<T, 2 * C>(yob + pix * 2 * C, yp);
      if (kAdapt) {
        // t=1 parts of y_k and y+ of the left neighbour
        T ykl1[C], ypl1[C];
#pragma unroll
        for (int k = 0; k < C; ++k) {
          ykl1[k] = shfl_up1(yk[i + 1][C + k]);
          ypl1[k] = shfl_up1(yp[C + k]);
        }
        if (lhalo && ok && c > 0) {
          T ypl[2 * C], gl[2 * C];
          yplus_px<T, C>(bxl[i], bxl[i + 1], bx[i + 1], ykl[i], r, c - 1, H, W, sigma, tvw, cfg, ypl, gl);
#pragma unroll
          for (int k = 0; k < C; ++k) {
            ykl1[k] = ykl[i][C + k];
            ypl1[k] = ypl[C + k];
          }
        }
        if (ok) {
          // primal residual (x_k - x_k+1)/tau - A^T(y_k - y+) and its reference term,
          // dual residual (y_k - y+)/sigma - A(x_k - x_k+1) and (y_k - y+)/sigma + A xbar
          T at_old[C], at_new[C], gxd[2 * C];
          adjoint_px<T, C>(yk[i + 1], yk[i + 1] + C, yk[i], ykl1, r, c, H, W, cfg, at_old);
          adjoint_px<T, C>(yp, yp + C, ypu, ypl1, r, c, H, W, cfg, at_new);
          grad_px<T, C>(ax[i + 1], ax[i + 2], axr[i + 1], r, c, H, W, cfg, gxd);
#pragma unroll
          for (int k = 0; k < C; ++k) {
            const T q = ax[i + 1][k] * itau - at_old[k];
            const T rp = q + at_new[k];
            s[0] += (double)rp * (double)rp;
            s[1] += (double)q * (double)q;
          }
#pragma unroll
          for (int e = 0; e < 2 * C; ++e) {
            const T q = (yk[i + 1][e] - yp[e]) * isig;
            const T rd = q - gxd[e];
            const T gd = q + g[e];
            s[2] += (double)rd * (double)rd;
            s[3] += (double)gd * (double)gd;
          }
        }
#pragma unroll
        for (int k = 0; k < C; ++k) ypu[k] = yp[k];
      }
    }
    if (kAdapt) {
      block_sum4(s, sm, parity);
      parity ^= 1;
      if (threadIdx.x == 0) {
        double* p = partials + 4 * (int64_t)(I.tile0 + tr.j);
        p[0] = s[0];
        p[1] = s[1];
        p[2] = s[2];
        p[3] = s[3];
      }
    }
  }
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

// x+ = prox_{tau G}(x - tau A^T y) and y+ = proj(y + sigma A(2x+ - x)) in one
// pass: reads x, f, y, writes x+, y+ (28 B per pixel at C=1 FP32 against 44 B
// for the two passes).  A wavefront owns 64 aligned columns, lane L column
// c0 + L, and a thread RPT rows; it forms x+ on its rows, the row below and
// (adaptive steps) the row above, and y+ on its rows and (adaptive) the row
// above.  What the edge lanes need from the neighbouring columns is formed
// "transposed": lane k works on row k of the column right of the segment
// (x+ at c0+64, for y+ of lane 63) and, with adaptive steps, of the column
// left of it (y+ at c0-1, for the residual of lane 0); one vector load per
// array brings those columns in, and the edge lanes pick their row j up with
// readlane plus a DPP shift.  Every access of the main data is aligned and
// no lane runs code of its own.
template <typename T, int C, int RPT, bool kAdapt>
__global__ void __launch_bounds__(kTB) tv_iter_kernel(const T* __restrict__ xin, T* __restrict__ xout,
                                                      const T* __restrict__ yin, T* __restrict__ yout,
                                                      const T* __restrict__ f, const opy_tv_inst_t* __restrict__ insts,
                                                      const opy_adapt_state_t* __restrict__ st,
                                                      const int32_t* __restrict__ active_inst,
                                                      const int32_t* __restrict__ act_off,
                                                      const int32_t* __restrict__ nact_p, Cfg cfg,
                                                      double* __restrict__ partials) {
  constexpr int OX = kAdapt ? 1 : 0;  // x row j <-> image row rs - OX + j
  constexpr int NX = RPT + 1 + OX;    // rows with x+
  constexpr int NY = NX + 1;          // y row j <-> image row rs - OX - 1 + j
  constexpr int NP = RPT + OX;        // rows with y+ (image rows rs - OX ..)
  static_assert(NY < 64, "the transposed halo needs a lane per row");
  __shared__ double sm[2 * (kTB / 64)][4];
  int parity = 0;
  const int nact = *nact_p;
  if (nact <= 0) return;
  const int total = act_off[nact];
  constexpr int th = (kTB / 64) * RPT;
  const int lane = threadIdx.x & 63;
  const int ty = threadIdx.x >> 6;
  for (int g = xcd_block(); g < total; g += gridDim.x) {
    const TileRef tr = locate(g, active_inst, act_off, nact, insts, 64, th);
    const opy_tv_inst_t I = insts[tr.b];
    const int H = I.H, W = I.W;
    const double tau_s = st[tr.b].tau;
    const T sigma = (T)st[tr.b].sigma;
    const T tvw = (T)I.tvw;
    const int c0 = tr.c0;
    const int c = c0 + lane;
    const bool col = c < W;
    const int rs = tr.r0 + ty * RPT;
    const T* xb = xin + I.off;
    const T* fb = f + I.off;
    const T* yb = yin + 2 * I.off;
    T* xob = xout + I.off;
    T* yob = yout + 2 * I.off;
    // ---- loads: the segment, then the transposed halo columns (lane k = row k)
    T xv[NX][C], fv[NX][C], yv[NY][2 * C];
#pragma unroll
    for (int j = 0; j < NX; ++j)
#pragma unroll
      for (int k = 0; k < C; ++k) xv[j][k] = fv[j][k] = T(0);
#pragma unroll
    for (int j = 0; j < NY; ++j)
#pragma unroll
      for (int e = 0; e < 2 * C; ++e) yv[j][e] = T(0);
#pragma unroll
    for (int j = 0; j < NY; ++j) {
      const int r = rs - OX - 1 + j;
      if (col && r >= 0 && r < H) load_c<T, 2 * C>(yb + ((int64_t)r * W + c) * 2 * C, yv[j]);
    }
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      const int r = rs - OX + j;
      if (col && r >= 0 && r < H) {
        const int64_t p = (int64_t)r * W + c;
        load_c<T, C>(xb + p * C, xv[j]);
        load_c<T, C>(fb + p * C, fv[j]);
      }
    }
    const int cr = c0 + 64, cl = c0 - 1;  // halo columns
    const int ry = rs - OX - 1 + lane;    // this lane's halo y row
    const int rx = rs - OX + lane;        // this lane's halo x row
    T yrh[2 * C], xrh[C], frh[C], ylh[2 * C], yllh[C], xlh[C], flh[C];
#pragma unroll
    for (int e = 0; e < 2 * C; ++e) yrh[e] = ylh[e] = T(0);
#pragma unroll
    for (int k = 0; k < C; ++k) xrh[k] = frh[k] = yllh[k] = xlh[k] = flh[k] = T(0);
    if (lane < NY && ry >= 0 && ry < H) {
      if (cr < W) load_c<T, 2 * C>(yb + ((int64_t)ry * W + cr) * 2 * C, yrh);
      if (cl >= 0) load_c<T, 2 * C>(yb + ((int64_t)ry * W + cl) * 2 * C, ylh);
      if (kAdapt && cl >= 1) load_c<T, C>(yb + ((int64_t)ry * W + cl - 1) * 2 * C + C, yllh);
    }
    if (lane < NX && rx >= 0 && rx < H) {
      if (cr < W) {
        load_c<T, C>(xb + ((int64_t)rx * W + cr) * C, xrh);
        load_c<T, C>(fb + ((int64_t)rx * W + cr) * C, frh);
      }
      if (kAdapt && cl >= 0) {
        load_c<T, C>(xb + ((int64_t)rx * W + cl) * C, xlh);
        load_c<T, C>(fb + ((int64_t)rx * W + cl) * C, flh);
      }
    }
    // ---- x+ on the segment rows; A^T y_k at row j (recomputed by the residual)
    auto aty_row = [&](int j, T* out) {
      T yl1[C];
#pragma unroll
      for (int k = 0; k < C; ++k) yl1[k] = dpp_shr1(readlane(ylh[C + k], j + 1), yv[j + 1][C + k]);
      adjoint_px<T, C>(yv[j + 1], yv[j + 1] + C, yv[j], yl1, rs - OX + j, c, H, W, cfg, out);
    };
    T xp[NX][C];
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      const int r = rs - OX + j;
      T aty[C];
      aty_row(j, aty);
#pragma unroll
      for (int k = 0; k < C; ++k) {
        const T tau = cfg.steps == OPY_TV_STEPS_PRECOND ? precond_tau<T>(r, c, H, W, k, cfg) : (T)tau_s;
        xp[j][k] = prox_data<T>(xv[j][k] - tau * aty[k], fv[j][k], (double)tau * I.lam, I.data);
      }
    }
    // ---- right halo, transposed: lane k forms x+ at (x row k, cr)
    T xbh[C], xdh[C];
    {
      T yo[2 * C], yl1[C], ah[C];
#pragma unroll
      for (int e = 0; e < 2 * C; ++e) yo[e] = dpp_shl1(T(0), yrh[e]);  // y row k+1 from lane k+1
#pragma unroll
      for (int k = 0; k < C; ++k) {
        T sel = T(0);  // y(x row k, c0+63), t=1, held by lane 63 as yv[k+1]
#pragma unroll
        for (int j = 0; j < NX; ++j) {
          const T v = readlane(yv[j + 1][C + k], 63);
          sel = lane == j ? v : sel;
        }
        yl1[k] = sel;
      }
      adjoint_px<T, C>(yo, yo + C, yrh, yl1, rx, cr, H, W, cfg, ah);
#pragma unroll
      for (int k = 0; k < C; ++k) {
        const T tau = cfg.steps == OPY_TV_STEPS_PRECOND ? precond_tau<T>(rx, cr, H, W, k, cfg) : (T)tau_s;
        const T p = prox_data<T>(xrh[k] - tau * ah[k], frh[k], (double)tau * I.lam, I.data);
        xbh[k] = T(2) * p - xrh[k];
        xdh[k] = xrh[k] - p;
      }
    }
    // xbar = 2x+ - x and xd = x - x+ (and their right neighbours) are formed
    // where they are used, which keeps fewer values live
    auto xbar = [&](int j, int k) { return T(2) * xp[j][k] - xv[j][k]; };
    // ---- left halo (adaptive), transposed: lane k forms y+ at (y+ row k, cl), t=1 part
    T yplh[C];
#pragma unroll
    for (int k = 0; k < C; ++k) yplh[k] = T(0);
    if (kAdapt) {
      T yo[2 * C], yl1[C], ah[C], xbl[C], xbld[C], xbc[C], ypl[2 * C], gl[2 * C];
#pragma unroll
      for (int e = 0; e < 2 * C; ++e) yo[e] = dpp_shl1(T(0), ylh[e]);  // y(row k, cl) from lane k+1
#pragma unroll
      for (int k = 0; k < C; ++k) yl1[k] = dpp_shl1(T(0), yllh[k]);    // y(row k, cl - 1), t=1
      adjoint_px<T, C>(yo, yo + C, ylh, yl1, rx, cl, H, W, cfg, ah);
#pragma unroll
      for (int k = 0; k < C; ++k) {
        const T p = prox_data<T>(xlh[k] - (T)tau_s * ah[k], flh[k], tau_s * I.lam, I.data);
        xbl[k] = T(2) * p - xlh[k];
      }
#pragma unroll
      for (int k = 0; k < C; ++k) {
        xbld[k] = dpp_shl1(T(0), xbl[k]);  // x row k+1
        T sel = T(0);                      // xbar(x row k, c0), formed by lane 0
#pragma unroll
        for (int j = 0; j < NX; ++j) {
          const T v = readlane(xbar(j, k), 0);
          sel = lane == j ? v : sel;
        }
        xbc[k] = sel;
      }
      yplus_px<T, C>(xbl, xbld, xbc, yo, rx, cl, H, W, sigma, tvw, cfg, ypl, gl);
#pragma unroll
      for (int k = 0; k < C; ++k) yplh[k] = ypl[C + k];
    }
    // ---- y+ on the segment rows (and the row above), stores, residuals
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const T isig = T(1) / sigma, itau = T(1) / (T)tau_s;
    T ypu[C];  // t=0 part of y+ in the row above (adaptive)
#pragma unroll
    for (int k = 0; k < C; ++k) ypu[k] = T(0);
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int r = rs - OX + i;
      const bool mine = i >= OX;
      const bool ok = col && r >= 0 && r < H;
      T yp[2 * C], gb[2 * C], xbi[C], xbd[C], xbr[C];
#pragma unroll
      for (int k = 0; k < C; ++k) {
        xbi[k] = xbar(i, k);
        xbd[k] = xbar(i + 1, k);
        xbr[k] = dpp_shl1(readlane(xbh[k], i), xbi[k]);
      }
      yplus_px<T, C>(xbi, xbd, xbr, yv[i + 1], r, c, H, W, sigma, tvw, cfg, yp, gb);
      if (mine && ok) {
        const int64_t p = (int64_t)r * W + c;
        store_c<T, C>(xob + p * C, xp[i]);
        store_c<T, 2 * C>(yob + p * 2 * C, yp);
      }
      if (kAdapt) {
        T ypl1[C];
#pragma unroll
        for (int k = 0; k < C; ++k) ypl1[k] = dpp_shr1(readlane(yplh[k], i), yp[C + k]);
        if (mine && ok) {
          // primal residual (x - x+)/tau - A^T(y - y+) with its reference term,
          // dual residual (y - y+)/sigma - A(x - x+) and (y - y+)/sigma + A xbar
          T atn[C], gxd[2 * C], xdi[C], xdd[C], xdr[C];
#pragma unroll
          for (int k = 0; k < C; ++k) {
            xdi[k] = xv[i][k] - xp[i][k];
            xdd[k] = xv[i + 1][k] - xp[i + 1][k];
          }
          T aty[C];
          aty_row(i, aty);
          adjoint_px<T, C>(yp, yp + C, ypu, ypl1, r, c, H, W, cfg, atn);
#pragma unroll
          for (int k = 0; k < C; ++k) xdr[k] = dpp_shl1(readlane(xdh[k], i), xdi[k]);
          grad_px<T, C>(xdi, xdd, xdr, r, c, H, W, cfg, gxd);
#pragma unroll
          for (int k = 0; k < C; ++k) {
            const T q = xdi[k] * itau - aty[k];
            const T rp = q + atn[k];
            s[0] += (double)rp * (double)rp;
            s[1] += (double)q * (double)q;
          }
#pragma unroll
          for (int e = 0; e < 2 * C; ++e) {
            const T q = (yv[i + 1][e] - yp[e]) * isig;
            const T rd = q - gxd[e];
            const T gd = q + gb[e];
            s[2] += (double)rd * (double)rd;
            s[3] += (double)gd * (double)gd;
          }
        }
#pragma unroll
        for (int k = 0; k < C; ++k) ypu[k] = yp[k];
      }
    }
    if (kAdapt) {
      block_sum4(s, sm, parity);
      parity ^= 1;
      if (threadIdx.x == 0) {
        double* p = partials + 4 * (int64_t)(I.tile0 + tr.j);
        p[0] = s[0];
        p[1] = s[1];
        p[2] = s[2];
        p[3] = s[3];
      }
    }
  }
}

// ---- per-instance adaptive step rule (reference pdhg.py:155-183) ------------
__global__ void __launch_bounds__(kTB) tv_adapt_kernel(const opy_tv_inst_t* __restrict__ insts,
                                                       opy_adapt_state_t* __restrict__ st,
                                                       const opy_tv_check_t* __restrict__ chk,
                                                       const double* __restrict__ partials, int C,
                                                       opy_adapt_params_t p) {
  __shared__ double sm[kTB / 64];
  const int b = blockIdx.x;
  if (chk[b].status != 0) return;
  const opy_tv_inst_t I = insts[b];
  double s[4] = {0, 0, 0, 0};
  for (int t = threadIdx.x; t < I.ntiles; t += blockDim.x)
    for (int j = 0; j < 4; ++j) s[j] += partials[4 * (int64_t)(I.tile0 + t) + j];
  for (int j = 0; j < 4; ++j) s[j] = block_sum(s[j], sm);
  if (threadIdx.x != 0) return;
  opy_adapt_state_t& S = st[b];
  const double n = (double)I.H * I.W * C;
  const double resp = sqrt(s[0]), gnp = sqrt(s[1]), resd = sqrt(s[2]), gnd = sqrt(s[3]);
  const double epsp = sqrt(n) * p.eps_absp + p.eps_relp * gnp;
  const double epsd = sqrt(2.0 * n) * p.eps_absd + p.eps_reld * gnd;
  const double scale = epsp / epsd;
  S.sums[0] = s[0];
  S.sums[1] = s[1];
  S.sums[2] = s[2];
  S.sums[3] = s[3];
  S.resp = resp;
  S.resd = resd;
  S.epsp = epsp;
  S.epsd = epsd;
  if (resp > scale * resd * p.Delta) {
    const double a = S.alpha;
    S.tau /= 1.0 - a;
    S.sigma *= 1.0 - a;
    S.alpha = a * p.eta;
    S.nchanges += 1;
  }
  if (resp < scale * resd / p.Delta) {
    const double a = S.alpha;
    S.tau *= 1.0 - a;
    S.sigma /= 1.0 - a;
    S.alpha = a * p.eta;
    S.nchanges += 1;
  }
}

// ---- primal-dual gap (reference pdhg.py:23-44) ------------------------------
// per tile partials: [P_data, P_tv, D_lin = sum z f, D_quad = sum z^2, max |y_p|, max |z|] with z = A^T y
template <typename T, int C, int RPT>
__global__ void __launch_bounds__(kTB) tv_gap_kernel(const T* __restrict__ x, const T* __restrict__ y,
                                                     const T* __restrict__ f, const opy_tv_inst_t* __restrict__ insts,
                                                     const int32_t* __restrict__ active_inst,
                                                     const int32_t* __restrict__ act_off,
                                                     const int32_t* __restrict__ nact_p, Cfg cfg,
                                                     double* __restrict__ partials) {
  __shared__ double sm[kTB / 64];
  const int nact = *nact_p;
  if (nact <= 0) return;
  const int total = act_off[nact];
  const int tw = cfg.tw;
  const int th = (kTB / tw) * RPT;
  const int lane = threadIdx.x & 63;
  const int tx = threadIdx.x % tw, ty = threadIdx.x / tw;
  const bool lhalo = (tx == 0) || (lane == 0);
  const bool rhalo = (tx == tw - 1) || (lane == 63);
  for (int g = xcd_block(); g < total; g += gridDim.x) {
    const TileRef tr = locate(g, active_inst, act_off, nact, insts, tw, th);
    const opy_tv_inst_t I = insts[tr.b];
    const int H = I.H, W = I.W;
    const int c = tr.c0 + tx;
    const bool col = c < W;
    const int rs = tr.r0 + ty * RPT;
    const T* xb = x + I.off;
    const T* fb = f + I.off;
    const T* yb = y + 2 * I.off;
    double v[6] = {0, 0, 0, 0, 0, 0};
    T xc[C], yup0[C];
#pragma unroll
    for (int k = 0; k < C; ++k) xc[k] = yup0[k] = T(0);
    if (col && rs < H) load_c<T, C>(xb + ((int64_t)rs * W + c) * C, xc);
    if (col && rs > 0 && rs < H) load_c<T, C>(yb + ((int64_t)(rs - 1) * W + c) * 2 * C, yup0);
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = rs + i;
      const bool ok = col && r < H;
      const int64_t pix = (int64_t)r * W + c;
      T xd[C], yv[2 * C], fv[C];
#pragma unroll
      for (int k = 0; k < C; ++k) xd[k] = fv[k] = yv[k] = yv[C + k] = T(0);
      if (ok && r + 1 < H) load_c<T, C>(xb + (pix + W) * C, xd);
      if (ok) {
        load_c<T, 2 * C>(yb + pix * 2 * C, yv);
        load_c<T, C>(fb + pix * C, fv);
      }
      T xr[C], yl1[C];
#pragma unroll
      for (int k = 0; k < C; ++k) {
        xr[k] = shfl_down1(xc[k]);
        yl1[k] = shfl_up1(yv[C + k]);
      }
      if (rhalo && ok && c + 1 < W) load_c<T, C>(xb + (pix + 1) * C, xr);
      if (lhalo && ok && c > 0) load_c<T, C>(yb + (pix - 1) * 2 * C + C, yl1);
      if (ok) {
        T gx[2 * C], z[C];
        grad_px<T, C>(xc, xd, xr, r, c, H, W, cfg, gx);
        adjoint_px<T, C>(yv, yv + C, yup0, yl1, r, c, H, W, cfg, z);
        double gn2 = 0.0, yn2 = 0.0;
#pragma unroll
        for (int e = 0; e < 2 * C; ++e) {
          gn2 += (double)gx[e] * (double)gx[e];
          yn2 += (double)yv[e] * (double)yv[e];
        }
        v[1] += sqrt(gn2);
        v[4] = fmax(v[4], sqrt(yn2));
#pragma unroll
        for (int k = 0; k < C; ++k) {
          const double d = (double)xc[k] - (double)fv[k];
          v[0] += I.data == OPY_TV_DATA_L1 ? fabs(d) : 0.5 * d * d;
          const double zk = (double)z[k];
          v[2] += zk * (double)fv[k];
          v[3] += zk * zk;
          v[5] = fmax(v[5], fabs(zk));
        }
      }
#pragma unroll
      for (int k = 0; k < C; ++k) {
        xc[k] = xd[k];
        yup0[k] = yv[k];
      }
    }
    for (int j = 0; j < 4; ++j) v[j] = block_sum(v[j], sm);
    v[4] = block_max(v[4], sm);
    v[5] = block_max(v[5], sm);
    if (threadIdx.x == 0) {
      double* p = partials + 6 * (int64_t)(I.tile0 + tr.j);
      for (int j = 0; j < 6; ++j) p[j] = v[j];
    }
  }
}

// Folds the gap partials of each running instance, records the objective
// values and marks converged / non-finite instances.
__global__ void __launch_bounds__(kTB) tv_check_kernel(const opy_tv_inst_t* __restrict__ insts,
                                                       opy_tv_check_t* __restrict__ chk,
                                                       const double* __restrict__ partials, double tol_gap,
                                                       double tol_infeas, int certify,
                                                       const int64_t* __restrict__ iter_counter, int iters_in_block) {
  __shared__ double sm[kTB / 64];
  const int b = blockIdx.x;
  if (chk[b].status != 0) return;
  const opy_tv_inst_t I = insts[b];
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int t = threadIdx.x; t < I.ntiles; t += blockDim.x) {
    const double* p = partials + 6 * (int64_t)(I.tile0 + t);
    for (int j = 0; j < 4; ++j) v[j] += p[j];
    v[4] = fmax(v[4], p[4]);
    v[5] = fmax(v[5], p[5]);
  }
  for (int j = 0; j < 4; ++j) v[j] = block_sum(v[j], sm);
  v[4] = block_max(v[4], sm);
  v[5] = block_max(v[5], sm);
  if (threadIdx.x != 0) return;
  opy_tv_check_t& R = chk[b];
  const double P = I.lam * v[0] + I.tvw * v[1];
  double D, infeas_d, Dc;
  const double ball = fmax(0.0, v[4] - I.tvw);
  if (I.data == OPY_TV_DATA_L1) {
    D = v[2];
    infeas_d = ball + fmax(0.0, v[5] - I.lam);
    const double s = v[5] > I.lam ? I.lam / v[5] : 1.0;  // the rescaled dual point is feasible
    Dc = s * v[2];
  } else {
    D = v[2] - 0.5 * v[3] / I.lam;
    infeas_d = ball;
    Dc = D;
  }
  R.objp = P;
  R.objd = D;
  R.gap = P - D;
  R.relgap = (P - D) / fmax(kEps, fabs(D));
  R.infeasp = 0.0;
  R.infeasd = infeas_d;
  R.objd_cert = Dc;
  R.relgap_cert = (P - Dc) / fmax(kEps, fabs(Dc));
  R.iters = (int32_t)(*iter_counter + iters_in_block);
  if (!isfinite(P) || !isfinite(D)) {
    R.status = 2;
    return;
  }
  const bool done = certify ? (fabs(R.relgap_cert) < tol_gap && ball < tol_infeas)
                            : (fabs(R.relgap) < tol_gap && infeas_d < tol_infeas);
  if (done) R.status = 1;
}

// Rebuilds the list of running instances and the prefix of their tile counts;
// advances the iteration counter by the block just run.
__global__ void __launch_bounds__(kTB) tv_compact_kernel(const opy_tv_inst_t* __restrict__ insts,
                                                         const opy_tv_check_t* __restrict__ chk, int B,
                                                         int32_t* __restrict__ active_inst,
                                                         int32_t* __restrict__ act_off, int32_t* __restrict__ nact_p,
                                                         int64_t* __restrict__ iter_counter, int iters_in_block) {
  __shared__ int cnt[kTB];
  __shared__ int tiles[kTB];
  const int per = (B + kTB - 1) / kTB;
  const int lo = threadIdx.x * per;
  const int hi = min(B, lo + per);
  int n = 0, nt = 0;
  for (int b = lo; b < hi; ++b)
    if (chk[b].status == 0) {
      ++n;
      nt += insts[b].ntiles;
    }
  cnt[threadIdx.x] = n;
  tiles[threadIdx.x] = nt;
  __syncthreads();
  if (threadIdx.x == 0) {  // exclusive scan over 256 entries
    int a = 0, t = 0;
    for (int i = 0; i < kTB; ++i) {
      const int ci = cnt[i], ti = tiles[i];
      cnt[i] = a;
      tiles[i] = t;
      a += ci;
      t += ti;
    }
    *nact_p = a;
    act_off[a] = t;
    *iter_counter += iters_in_block;
  }
  __syncthreads();
  int pos = cnt[threadIdx.x], off = tiles[threadIdx.x];
  for (int b = lo; b < hi; ++b)
    if (chk[b].status == 0) {
      active_inst[pos] = b;
      act_off[pos] = off;
      ++pos;
      off += insts[b].ntiles;
    }
}

// Cold restart of flagged instances: zero their x and y (buffer 0) and mark
// them running again.  One block per instance.
template <typename T>
__global__ void __launch_bounds__(kTB) tv_reset_kernel(const opy_tv_inst_t* __restrict__ insts,
                                                       const int32_t* __restrict__ flags, int C, T* __restrict__ x,
                                                       T* __restrict__ y, opy_tv_check_t* __restrict__ chk) {
  const int b = blockIdx.x;
  if (!flags[b]) return;
  const opy_tv_inst_t I = insts[b];
  const int64_t n = (int64_t)I.H * I.W * C;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) x[I.off + i] = T(0);
  for (int64_t i = threadIdx.x; i < 2 * n; i += blockDim.x) y[2 * I.off + i] = T(0);
  if (threadIdx.x == 0) chk[b].status = 0;
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
  if (c->steps != OPY_TV_STEPS_SCALAR && c->steps != OPY_TV_STEPS_PRECOND) return fail(kInvalid, "tv: bad step mode");
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

template <typename T>
int primal_impl(const T* xin, T* xout, const T* y, const T* f, const opy_tv_inst_t* insts, const opy_adapt_state_t* st,
                const int32_t* ai, const int32_t* ao, const int32_t* na, const opy_tv_cfg_t* cfg, int grid, void* s) {
  if (int rc = check_cfg(cfg)) return rc;
  if (!xin || !xout || !y || !f || !insts || !st || !ai || !ao || !na || grid < 0)
    return fail(kInvalid, "tv_primal: null pointer or negative grid");
  if (xin == xout) return fail(kInvalid, "tv_primal: x_in and x_out must differ");
  Cfg c = make_cfg(cfg);
  hipStream_t stm = as_stream(s);
#define OPY_L(CC, RR) \
  hipLaunchKernelGGL((tv_primal_kernel<T, CC, RR>), dim3(launch_grid(KF(tv_primal_kernel<T, CC, RR>), grid)), dim3(kTB), 0, \
                     stm, xin, xout, y, f, insts, st, ai, ao, na, c)
  OPY_TV_DISPATCH(OPY_L)
#undef OPY_L
  return check_launch("tv_primal");
}

template <typename T>
int dual_impl(const T* xo, const T* xn, const T* yi, T* yo, const opy_tv_inst_t* insts, const opy_adapt_state_t* st,
              const int32_t* ai, const int32_t* ao, const int32_t* na, const opy_tv_cfg_t* cfg, int adapt,
              double* partials, int grid, void* s) {
  if (int rc = check_cfg(cfg)) return rc;
  if (!xo || !xn || !yi || !yo || !insts || !st || !ai || !ao || !na || grid < 0)
    return fail(kInvalid, "tv_dual: null pointer or negative grid");
  if (yi == yo || xo == xn) return fail(kInvalid, "tv_dual: ping-pong buffers must differ");
  if (adapt && !partials) return fail(kInvalid, "tv_dual: adaptive steps need a partials buffer");
  if (adapt && cfg->steps == OPY_TV_STEPS_PRECOND) return fail(kInvalid, "tv_dual: adaptive and precond exclude each other");
  Cfg c = make_cfg(cfg);
  hipStream_t stm = as_stream(s);
  if (adapt) {
#define OPY_L(CC, RR) \
  hipLaunchKernelGGL((tv_dual_kernel<T, CC, RR, true>), dim3(launch_grid(KF(tv_dual_kernel<T, CC, RR, true>), grid)), \
                     dim3(kTB), 0, stm, xo, xn, yi, yo, insts, st, ai, ao, na, c, partials)
    OPY_TV_DISPATCH(OPY_L)
#undef OPY_L
  } else {
#define OPY_L(CC, RR) \
  hipLaunchKernelGGL((tv_dual_kernel<T, CC, RR, false>), dim3(launch_grid(KF(tv_dual_kernel<T, CC, RR, false>), grid)), \
                     dim3(kTB), 0, stm, xo, xn, yi, yo, insts, st, ai, ao, na, c, partials)
    OPY_TV_DISPATCH(OPY_L)
#undef OPY_L
  }
  return check_launch("tv_dual");
}

template <typename T>
int iter_impl(const T* xi, T* xo, const T* yi, T* yo, const T* f, const opy_tv_inst_t* insts,
              const opy_adapt_state_t* st, const int32_t* ai, const int32_t* ao, const int32_t* na,
              const opy_tv_cfg_t* cfg, int adapt, double* partials, int grid, void* s) {
  if (int rc = check_cfg(cfg)) return rc;
  if (cfg->tw != 64) return fail(kInvalid, "tv_iter: needs tw == 64 (got %d)", cfg->tw);
  if (cfg->rpt > 16) return fail(kInvalid, "tv_iter: at most 16 rows per thread");
  if (!xi || !xo || !yi || !yo || !f || !insts || !st || !ai || !ao || !na || grid < 0)
    return fail(kInvalid, "tv_iter: null pointer or negative grid");
  if (xi == xo || yi == yo) return fail(kInvalid, "tv_iter: ping-pong buffers must differ");
  if (adapt && !partials) return fail(kInvalid, "tv_iter: adaptive steps need a partials buffer");
  if (adapt && cfg->steps == OPY_TV_STEPS_PRECOND) return fail(kInvalid, "tv_iter: adaptive and precond exclude each other");
  Cfg c = make_cfg(cfg);
  hipStream_t stm = as_stream(s);
  if (adapt) {
#define OPY_L(CC, RR) \
  hipLaunchKernelGGL((tv_iter_kernel<T, CC, RR, true>), dim3(launch_grid(KF(tv_iter_kernel<T, CC, RR, true>), grid)), \
                     dim3(kTB), 0, stm, xi, xo, yi, yo, f, insts, st, ai, ao, na, c, partials)
    OPY_TV_DISPATCH(OPY_L)
#undef OPY_L
  } else {
#define OPY_L(CC, RR) \
  hipLaunchKernelGGL((tv_iter_kernel<T, CC, RR, false>), dim3(launch_grid(KF(tv_iter_kernel<T, CC, RR, false>), grid)), \
                     dim3(kTB), 0, stm, xi, xo, yi, yo, f, insts, st, ai, ao, na, c, partials)
    OPY_TV_DISPATCH(OPY_L)
#undef OPY_L
  }
  return check_launch("tv_iter");
}

template <typename T>
int gap_impl(const T* x, const T* y, const T* f, const opy_tv_inst_t* insts, const int32_t* ai, const int32_t* ao,
             const int32_t* na, const opy_tv_cfg_t* cfg, double* partials, int grid, void* s) {
  if (int rc = check_cfg(cfg)) return rc;
  if (!x || !y || !f || !insts || !ai || !ao || !na || !partials || grid < 0)
    return fail(kInvalid, "tv_gap: null pointer or negative grid");
  Cfg c = make_cfg(cfg);
  hipStream_t stm = as_stream(s);
#define OPY_L(CC, RR) \
  hipLaunchKernelGGL((tv_gap_kernel<T, CC, RR>), dim3(launch_grid(KF(tv_gap_kernel<T, CC, RR>), grid)), dim3(kTB), 0, \
                     stm, x, y, f, insts, ai, ao, na, c, partials)
  OPY_TV_DISPATCH(OPY_L)
#undef OPY_L
  return check_launch("tv_gap");
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_tv_primal_f32(const float* xin, float* xout, const float* y, const float* f,
                              const opy_tv_inst_t* insts, const opy_adapt_state_t* st, const int32_t* active_inst,
                              const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int grid,
                              void* s) {
  return primal_impl<float>(xin, xout, y, f, insts, st, active_inst, act_off, nact, cfg, grid, s);
}
OPY_API int opy_tv_primal_f64(const double* xin, double* xout, const double* y, const double* f,
                              const opy_tv_inst_t* insts, const opy_adapt_state_t* st, const int32_t* active_inst,
                              const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int grid,
                              void* s) {
  return primal_impl<double>(xin, xout, y, f, insts, st, active_inst, act_off, nact, cfg, grid, s);
}
OPY_API int opy_tv_dual_f32(const float* xo, const float* xn, const float* yi, float* yo, const opy_tv_inst_t* insts,
                            const opy_adapt_state_t* st, const int32_t* active_inst, const int32_t* act_off,
                            const int32_t* nact, const opy_tv_cfg_t* cfg, int adapt, double* partials, int grid,
                            void* s) {
  return dual_impl<float>(xo, xn, yi, yo, insts, st, active_inst, act_off, nact, cfg, adapt, partials, grid, s);
}
OPY_API int opy_tv_dual_f64(const double* xo, const double* xn, const double* yi, double* yo,
                            const opy_tv_inst_t* insts, const opy_adapt_state_t* st, const int32_t* active_inst,
                            const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int adapt,
                            double* partials, int grid, void* s) {
  return dual_impl<double>(xo, xn, yi, yo, insts, st, active_inst, act_off, nact, cfg, adapt, partials, grid, s);
}
OPY_API int opy_tv_iter_f32(const float* xi, float* xo, const float* yi, float* yo, const float* f,
                            const opy_tv_inst_t* insts, const opy_adapt_state_t* st, const int32_t* active_inst,
                            const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int adapt,
                            double* partials, int grid, void* s) {
  return iter_impl<float>(xi, xo, yi, yo, f, insts, st, active_inst, act_off, nact, cfg, adapt, partials, grid, s);
}
OPY_API int opy_tv_iter_f64(const double* xi, double* xo, const double* yi, double* yo, const double* f,
                            const opy_tv_inst_t* insts, const opy_adapt_state_t* st, const int32_t* active_inst,
                            const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int adapt,
                            double* partials, int grid, void* s) {
  return iter_impl<double>(xi, xo, yi, yo, f, insts, st, active_inst, act_off, nact, cfg, adapt, partials, grid, s);
}
OPY_API int opy_tv_adapt(const opy_tv_inst_t* insts, opy_adapt_state_t* st, const opy_tv_check_t* chk,
                         const double* partials, int B, int C, const opy_adapt_params_t* p, void* s) {
  if (!insts || !st || !chk || !partials || !p || B < 1) return fail(kInvalid, "tv_adapt: bad arguments");
  hipLaunchKernelGGL(tv_adapt_kernel, dim3(B), dim3(kTB), 0, as_stream(s), insts, st, chk, partials, C, *p);
  return check_launch("tv_adapt");
}
OPY_API int opy_tv_gap_f32(const float* x, const float* y, const float* f, const opy_tv_inst_t* insts,
                           const int32_t* active_inst, const int32_t* act_off, const int32_t* nact,
                           const opy_tv_cfg_t* cfg, double* partials, int grid, void* s) {
  return gap_impl<float>(x, y, f, insts, active_inst, act_off, nact, cfg, partials, grid, s);
}
OPY_API int opy_tv_gap_f64(const double* x, const double* y, const double* f, const opy_tv_inst_t* insts,
                           const int32_t* active_inst, const int32_t* act_off, const int32_t* nact,
                           const opy_tv_cfg_t* cfg, double* partials, int grid, void* s) {
  return gap_impl<double>(x, y, f, insts, active_inst, act_off, nact, cfg, partials, grid, s);
}
OPY_API int opy_tv_check(const opy_tv_inst_t* insts, opy_tv_check_t* chk, const double* partials, int B,
                         double tol_gap, double tol_infeas, int certify, const int64_t* iter_counter,
                         int iters_in_block, void* s) {
  if (!insts || !chk || !partials || !iter_counter || B < 1) return fail(kInvalid, "tv_check: bad arguments");
  hipLaunchKernelGGL(tv_check_kernel, dim3(B), dim3(kTB), 0, as_stream(s), insts, chk, partials, tol_gap, tol_infeas,
                     certify, iter_counter, iters_in_block);
  return check_launch("tv_check");
}
template <typename T>
static int reset_impl(const opy_tv_inst_t* insts, const int32_t* flags, int B, int C, T* x, T* y, opy_tv_check_t* chk,
                      void* s) {
  if (!insts || !flags || !x || !y || !chk || B < 1 || C < 1 || C > 4) return fail(kInvalid, "tv_reset: bad arguments");
  hipLaunchKernelGGL((tv_reset_kernel<T>), dim3(B), dim3(kTB), 0, as_stream(s), insts, flags, C, x, y, chk);
  return check_launch("tv_reset");
}
OPY_API int opy_tv_reset_f32(const opy_tv_inst_t* insts, const int32_t* flags, int B, int C, float* x, float* y,
                             opy_tv_check_t* chk, void* s) {
  return reset_impl<float>(insts, flags, B, C, x, y, chk, s);
}
OPY_API int opy_tv_reset_f64(const opy_tv_inst_t* insts, const int32_t* flags, int B, int C, double* x, double* y,
                             opy_tv_check_t* chk, void* s) {
  return reset_impl<double>(insts, flags, B, C, x, y, chk, s);
}

OPY_API int opy_tv_compact(const opy_tv_inst_t* insts, const opy_tv_check_t* chk, int B, int32_t* active_inst,
                           int32_t* act_off, int32_t* nact, int64_t* iter_counter, int iters_in_block, void* s) {
  if (!insts || !chk || !active_inst || !act_off || !nact || !iter_counter || B < 1)
    return fail(kInvalid, "tv_compact: bad arguments");
  hipLaunchKernelGGL(tv_compact_kernel, dim3(1), dim3(kTB), 0, as_stream(s), insts, chk, B, active_inst, act_off, nact,
                     iter_counter, iters_in_block);
  return check_launch("tv_compact");
}
