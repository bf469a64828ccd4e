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
//   pass 1 (primal): x+ = prox_{tau G}(x - tau * A^T y)      reads x, y, f; writes x+
//   pass 2 (dual):   y+ = proj_{|.|<=tvw}(y + sigma * A(2x+ - x))  reads x, x+, y; writes y+
// The over-relaxed A x_bar = 2 A x+ - A x of the reference is applied as
// A(2x+ - x) (equal by linearity).  With adaptive steps pass 2 also forms the
// four residual norms of pdhg.py:155-164 by recomputing y+ on a one-pixel halo
// (so that A^T y+ is available) -- no extra sweep over memory.
//
// Work is split into tiles of tw x (256/tw*rpt) pixels.  A persistent grid
// walks the tiles of the still-running instances: active_inst / act_off hold
// the running instances and the exclusive prefix of their tile counts, a tile
// index is mapped to its instance by binary search.  All per-instance
// reductions (residuals, gap) go through per-tile FP64 partials that a
// per-instance kernel folds in a fixed order, so results are reproducible.
#include "opy_common.h"

#include <cmath>

namespace opy {
namespace {

constexpr int kTB = 256;

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
  int b;    // instance
  int j;    // tile index within the instance
  int r0, c0;
};

__device__ inline TileRef locate(int g, const int32_t* __restrict__ active_inst, const int32_t* __restrict__ act_off,
                                 int nact, const opy_tv_inst_t* __restrict__ insts, const Cfg& cfg) {
  // largest a with act_off[a] <= g
  int lo = 0, hi = nact - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (act_off[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  TileRef t;
  t.b = active_inst[lo];
  t.j = g - act_off[lo];
  const int W = insts[t.b].W;
  const int ntx = (W + cfg.tw - 1) / cfg.tw;
  const int th = (kTB / cfg.tw) * cfg.rpt;
  t.r0 = (t.j / ntx) * th;
  t.c0 = (t.j - (t.j / ntx) * ntx) * cfg.tw;
  return t;
}

template <typename T>
__device__ inline T step_precond_tau(int r, int c, int H, int W, int k, const Cfg& cfg) {
  // 1 / column sum of |A| (reference pdhg.py:197-200 via rowwise_lp of the divergence)
  const double s = cfg.w[k] * (((r < H - 1) + (r > 0)) * cfg.ih0 + ((c < W - 1) + (c > 0)) * cfg.ih1);
  return (T)(s > 2.220446049250313e-16 ? 1.0 / s : s);
}

template <typename T>
__device__ inline T step_precond_sigma(bool inner, int k, double ih, const Cfg& cfg) {
  const double s = inner ? 2.0 * cfg.w[k] * ih : 0.0;
  return (T)(s > 2.220446049250313e-16 ? 1.0 / s : s);
}

// A^T y at pixel (r, c), channel k: w_k * (-y0(r,c)/h0 [r<H-1] + y0(r-1,c)/h0 [r>0]
//                                          -y1(r,c)/h1 [c<W-1] + y1(r,c-1)/h1 [c>0])
template <typename T, int C>
__device__ inline void adjoint_at(const T* __restrict__ yb, int64_t pix, int r, int c, int H, int W, const Cfg& cfg,
                                  T* out) {
  const T* yp = yb + pix * 2 * C;
  const T ih0 = (T)cfg.ih0, ih1 = (T)cfg.ih1;
#pragma unroll
  for (int k = 0; k < C; ++k) {
    T d = T(0);
    if (r < H - 1) d -= yp[k] * ih0;
    if (r > 0) d += yp[k - (int64_t)2 * C * W] * ih0;
    if (c < W - 1) d -= yp[C + k] * ih1;
    if (c > 0) d += yp[C + k - 2 * C] * ih1;
    out[k] = (T)cfg.w[k] * d;
  }
}

// grad of (2*xn - xo) (or of one array when xo == nullptr and two == false) at (r, c): g[0..C) axis 0, g[C..2C) axis 1
template <typename T, int C>
__device__ inline void grad_at(const T* __restrict__ xa, const T* __restrict__ xb, int64_t pix, int r, int c, int H,
                               int W, const Cfg& cfg, T* g) {
  // value of the combined field at an offset: xb ? 2*xa - xb : xa
  auto val = [&](int64_t p, int k) -> T { return xb ? T(2) * xa[p * C + k] - xb[p * C + k] : xa[p * C + k]; };
#pragma unroll
  for (int k = 0; k < C; ++k) {
    const T v = val(pix, k);
    g[k] = (r < H - 1) ? (T)(cfg.w[k] * cfg.ih0) * (val(pix + W, k) - v) : T(0);
    g[C + k] = (c < W - 1) ? (T)(cfg.w[k] * cfg.ih1) * (val(pix + 1, k) - v) : T(0);
  }
}

// y+ at one pixel: projection of y + sigma*grad(2xn - xo) onto the ball of radius tvw
template <typename T, int C>
__device__ inline void dual_at(const T* __restrict__ xo, const T* __restrict__ xn, const T* __restrict__ yin, int64_t pix,
                               int r, int c, int H, int W, T sigma, T tvw, const Cfg& cfg, T* yv) {
  T g[2 * C];
  grad_at<T, C>(xn, xo, pix, r, c, H, W, cfg, g);
  const T* yp = yin + pix * 2 * C;
  T n2 = T(0);
#pragma unroll
  for (int e = 0; e < 2 * C; ++e) {
    T s = sigma;
    if (cfg.steps == OPY_TV_STEPS_PRECOND) {
      const int k = e % C;
      s = (e < C) ? step_precond_sigma<T>(r < H - 1, k, cfg.ih0, cfg) : step_precond_sigma<T>(c < W - 1, k, cfg.ih1, cfg);
    }
    yv[e] = yp[e] + s * g[e];
    n2 += yv[e] * yv[e];
  }
  if (n2 > tvw * tvw) {
    const T sc = tvw / sqrt(n2);
#pragma unroll
    for (int e = 0; e < 2 * C; ++e) yv[e] *= sc;
  }
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

// ---- pass 1 ---------------------------------------------------------------
template <typename T, int C>
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
  const int nty = kTB / cfg.tw;
  const int tx = threadIdx.x % cfg.tw, ty = threadIdx.x / cfg.tw;
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const TileRef tr = locate(g, active_inst, act_off, nact, insts, cfg);
    const opy_tv_inst_t I = insts[tr.b];
    const double tau_s = st[tr.b].tau;
    const int c = tr.c0 + tx;
    if (c >= I.W) continue;
    const T* xb = xin + I.off;
    const T* fb = f + I.off;
    const T* yb = y + 2 * I.off;
    T* ob = xout + I.off;
    for (int rr = 0; rr < cfg.rpt; ++rr) {
      const int r = tr.r0 + ty + rr * nty;
      if (r >= I.H) break;
      const int64_t pix = (int64_t)r * I.W + c;
      T d[C];
      adjoint_at<T, C>(yb, pix, r, c, I.H, I.W, cfg, d);
#pragma unroll
      for (int k = 0; k < C; ++k) {
        const T tau = cfg.steps == OPY_TV_STEPS_PRECOND ? step_precond_tau<T>(r, c, I.H, I.W, k, cfg) : (T)tau_s;
        const T v = xb[pix * C + k] - tau * d[k];
        ob[pix * C + k] = prox_data<T>(v, fb[pix * C + k], (double)tau * I.lam, I.data);
      }
    }
  }
}

// ---- pass 2 ---------------------------------------------------------------
template <typename T, int C, bool kAdapt>
__global__ void __launch_bounds__(kTB) tv_dual_kernel(const T* __restrict__ xo, const T* __restrict__ xn,
                                                      const T* __restrict__ yin, T* __restrict__ yout,
                                                      const opy_tv_inst_t* __restrict__ insts,
                                                      const opy_adapt_state_t* __restrict__ st,
                                                      const int32_t* __restrict__ active_inst,
                                                      const int32_t* __restrict__ act_off,
                                                      const int32_t* __restrict__ nact_p, Cfg cfg,
                                                      double* __restrict__ partials) {
  __shared__ double sm[kTB / 64];
  const int nact = *nact_p;
  if (nact <= 0) return;
  const int total = act_off[nact];
  const int nty = kTB / cfg.tw;
  const int tx = threadIdx.x % cfg.tw, ty = threadIdx.x / cfg.tw;
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const TileRef tr = locate(g, active_inst, act_off, nact, insts, cfg);
    const opy_tv_inst_t I = insts[tr.b];
    const T sigma = (T)st[tr.b].sigma;
    const T tau = (T)st[tr.b].tau;
    const T tvw = (T)I.tvw;
    const int c = tr.c0 + tx;
    const T* xob = xo + I.off;
    const T* xnb = xn + I.off;
    const T* yib = yin + 2 * I.off;
    T* yob = yout + 2 * I.off;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (c < I.W) {
      for (int rr = 0; rr < cfg.rpt; ++rr) {
        const int r = tr.r0 + ty + rr * nty;
        if (r >= I.H) break;
        const int64_t pix = (int64_t)r * I.W + c;
        T yv[2 * C];
        dual_at<T, C>(xob, xnb, yib, pix, r, c, I.H, I.W, sigma, tvw, cfg, yv);
#pragma unroll
        for (int e = 0; e < 2 * C; ++e) yob[pix * 2 * C + e] = yv[e];
        if (kAdapt) {
          // y+ of the upper and left neighbours (recomputed) for A^T y+ at (r, c)
          T yu[2 * C], yl[2 * C];
          if (r > 0) dual_at<T, C>(xob, xnb, yib, pix - I.W, r - 1, c, I.H, I.W, sigma, tvw, cfg, yu);
          if (c > 0) dual_at<T, C>(xob, xnb, yib, pix - 1, r, c - 1, I.H, I.W, sigma, tvw, cfg, yl);
          const T ih0 = (T)cfg.ih0, ih1 = (T)cfg.ih1;
          T at_old[C];
          adjoint_at<T, C>(yib, pix, r, c, I.H, I.W, cfg, at_old);
          T gx_old[2 * C], gx_new[2 * C];
          grad_at<T, C>(xob, nullptr, pix, r, c, I.H, I.W, cfg, gx_old);
          grad_at<T, C>(xnb, nullptr, pix, r, c, I.H, I.W, cfg, gx_new);
          const T* yp = yib + pix * 2 * C;
#pragma unroll
          for (int k = 0; k < C; ++k) {
            T dn = T(0);
            if (r < I.H - 1) dn -= yv[k] * ih0;
            if (r > 0) dn += yu[k] * ih0;
            if (c < I.W - 1) dn -= yv[C + k] * ih1;
            if (c > 0) dn += yl[C + k] * ih1;
            const T at_new = (T)cfg.w[k] * dn;
            const T q = (xob[pix * C + k] - xnb[pix * C + k]) / tau - at_old[k];
            const T rp = q + at_new;
            s0 += (double)rp * (double)rp;
            s1 += (double)q * (double)q;
          }
#pragma unroll
          for (int e = 0; e < 2 * C; ++e) {
            const T q = (yp[e] - yv[e]) / sigma - gx_old[e];
            const T rd = q + gx_new[e];
            const T gd = q + T(2) * gx_new[e];
            s2 += (double)rd * (double)rd;
            s3 += (double)gd * (double)gd;
          }
        }
      }
    }
    if (kAdapt) {
      s0 = block_sum(s0, sm);
      s1 = block_sum(s1, sm);
      s2 = block_sum(s2, sm);
      s3 = block_sum(s3, sm);
      if (threadIdx.x == 0) {
        double* p = partials + 4 * (int64_t)(I.tile0 + tr.j);
        p[0] = s0;
        p[1] = s1;
        p[2] = s2;
        p[3] = s3;
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
template <typename T, int C>
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
  const int nty = kTB / cfg.tw;
  const int tx = threadIdx.x % cfg.tw, ty = threadIdx.x / cfg.tw;
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const TileRef tr = locate(g, active_inst, act_off, nact, insts, cfg);
    const opy_tv_inst_t I = insts[tr.b];
    const int c = tr.c0 + tx;
    const T* xb = x + I.off;
    const T* fb = f + I.off;
    const T* yb = y + 2 * I.off;
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (c < I.W) {
      for (int rr = 0; rr < cfg.rpt; ++rr) {
        const int r = tr.r0 + ty + rr * nty;
        if (r >= I.H) break;
        const int64_t pix = (int64_t)r * I.W + c;
        T gx[2 * C], z[C];
        grad_at<T, C>(xb, nullptr, pix, r, c, I.H, I.W, cfg, gx);
        adjoint_at<T, C>(yb, pix, r, c, I.H, I.W, cfg, z);
        double gn2 = 0.0, yn2 = 0.0;
        for (int e = 0; e < 2 * C; ++e) {
          gn2 += (double)gx[e] * (double)gx[e];
          const double ye = (double)yb[pix * 2 * C + e];
          yn2 += ye * ye;
        }
        v[1] += sqrt(gn2);
        v[4] = fmax(v[4], sqrt(yn2));
        for (int k = 0; k < C; ++k) {
          const double d = (double)xb[pix * C + k] - (double)fb[pix * C + k];
          v[0] += I.data == OPY_TV_DATA_L1 ? fabs(d) : 0.5 * d * d;
          const double zk = (double)z[k];
          v[2] += zk * (double)fb[pix * C + k];
          v[3] += zk * zk;
          v[5] = fmax(v[5], fabs(zk));
        }
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
  const double eps = 2.220446049250313e-16;
  const double P = I.lam * v[0] + I.tvw * v[1];
  double D, infeas_d, Dc;
  const double ball = fmax(0.0, v[4] - I.tvw);
  if (I.data == OPY_TV_DATA_L1) {
    D = v[2];
    infeas_d = ball + fmax(0.0, v[5] - I.lam);
    const double s = v[5] > I.lam ? I.lam / v[5] : 1.0;  // rescaled dual point is feasible
    Dc = s * v[2];
  } else {
    D = v[2] - 0.5 * v[3] / I.lam;
    infeas_d = ball;
    Dc = D;
  }
  R.objp = P;
  R.objd = D;
  R.gap = P - D;
  R.relgap = (P - D) / fmax(eps, fabs(D));
  R.infeasp = 0.0;
  R.infeasd = infeas_d;
  R.objd_cert = Dc;
  R.relgap_cert = (P - Dc) / fmax(eps, fabs(Dc));
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

// ---- host wrappers ---------------------------------------------------------
int check_cfg(const opy_tv_cfg_t* c) {
  if (!c) return fail(kInvalid, "tv: null config");
  if (c->C < 1 || c->C > 4) return fail(kInvalid, "tv: C=%d outside [1,4]", c->C);
  if (c->tw < 1 || c->tw > kTB || kTB % c->tw) return fail(kInvalid, "tv: tile width %d must divide %d", c->tw, kTB);
  if (c->rpt < 1 || c->rpt > 64) return fail(kInvalid, "tv: rows per thread %d outside [1,64]", c->rpt);
  if (c->steps != OPY_TV_STEPS_SCALAR && c->steps != OPY_TV_STEPS_PRECOND) return fail(kInvalid, "tv: bad step mode");
  if (!(c->h0 > 0) || !(c->h1 > 0)) return fail(kInvalid, "tv: grid spacing must be positive");
  return kOk;
}

template <typename T>
int primal_impl(const T* xin, T* xout, const T* y, const T* f, const opy_tv_inst_t* insts, const opy_adapt_state_t* st,
                const int32_t* ai, const int32_t* ao, const int32_t* na, const opy_tv_cfg_t* cfg, int grid, void* s) {
  if (int rc = check_cfg(cfg)) return rc;
  if (!xin || !xout || !y || !f || !insts || !st || !ai || !ao || !na || grid < 1)
    return fail(kInvalid, "tv_primal: null pointer or empty grid");
  if (xin == xout) return fail(kInvalid, "tv_primal: x_in and x_out must differ");
  Cfg c = make_cfg(cfg);
  hipStream_t stm = as_stream(s);
  switch (cfg->C) {
    case 1: hipLaunchKernelGGL((tv_primal_kernel<T, 1>), dim3(grid), dim3(kTB), 0, stm, xin, xout, y, f, insts, st, ai, ao, na, c); break;
    case 2: hipLaunchKernelGGL((tv_primal_kernel<T, 2>), dim3(grid), dim3(kTB), 0, stm, xin, xout, y, f, insts, st, ai, ao, na, c); break;
    case 3: hipLaunchKernelGGL((tv_primal_kernel<T, 3>), dim3(grid), dim3(kTB), 0, stm, xin, xout, y, f, insts, st, ai, ao, na, c); break;
    default: hipLaunchKernelGGL((tv_primal_kernel<T, 4>), dim3(grid), dim3(kTB), 0, stm, xin, xout, y, f, insts, st, ai, ao, na, c); break;
  }
  return check_launch("tv_primal");
}

template <typename T, bool A>
void launch_dual(int C, int grid, hipStream_t stm, const T* xo, const T* xn, const T* yi, T* yo,
                 const opy_tv_inst_t* insts, const opy_adapt_state_t* st, const int32_t* ai, const int32_t* ao,
                 const int32_t* na, const Cfg& c, double* part) {
  switch (C) {
    case 1: hipLaunchKernelGGL((tv_dual_kernel<T, 1, A>), dim3(grid), dim3(kTB), 0, stm, xo, xn, yi, yo, insts, st, ai, ao, na, c, part); break;
    case 2: hipLaunchKernelGGL((tv_dual_kernel<T, 2, A>), dim3(grid), dim3(kTB), 0, stm, xo, xn, yi, yo, insts, st, ai, ao, na, c, part); break;
    case 3: hipLaunchKernelGGL((tv_dual_kernel<T, 3, A>), dim3(grid), dim3(kTB), 0, stm, xo, xn, yi, yo, insts, st, ai, ao, na, c, part); break;
    default: hipLaunchKernelGGL((tv_dual_kernel<T, 4, A>), dim3(grid), dim3(kTB), 0, stm, xo, xn, yi, yo, insts, st, ai, ao, na, c, part); break;
  }
}

template <typename T>
int dual_impl(const T* xo, const T* xn, const T* yi, T* yo, const opy_tv_inst_t* insts, const opy_adapt_state_t* st,
              const int32_t* ai, const int32_t* ao, const int32_t* na, const opy_tv_cfg_t* cfg, int adapt,
              double* partials, int grid, void* s) {
  if (int rc = check_cfg(cfg)) return rc;
  if (!xo || !xn || !yi || !yo || !insts || !st || !ai || !ao || !na || grid < 1)
    return fail(kInvalid, "tv_dual: null pointer or empty grid");
  if (yi == yo || xo == xn) return fail(kInvalid, "tv_dual: ping-pong buffers must differ");
  if (adapt && !partials) return fail(kInvalid, "tv_dual: adaptive steps need a partials buffer");
  if (adapt && cfg->steps == OPY_TV_STEPS_PRECOND) return fail(kInvalid, "tv_dual: adaptive and precond exclude each other");
  Cfg c = make_cfg(cfg);
  hipStream_t stm = as_stream(s);
  if (adapt) launch_dual<T, true>(cfg->C, grid, stm, xo, xn, yi, yo, insts, st, ai, ao, na, c, partials);
  else launch_dual<T, false>(cfg->C, grid, stm, xo, xn, yi, yo, insts, st, ai, ao, na, c, partials);
  return check_launch("tv_dual");
}

template <typename T>
int gap_impl(const T* x, const T* y, const T* f, const opy_tv_inst_t* insts, const int32_t* ai, const int32_t* ao,
             const int32_t* na, const opy_tv_cfg_t* cfg, double* partials, int grid, void* s) {
  if (int rc = check_cfg(cfg)) return rc;
  if (!x || !y || !f || !insts || !ai || !ao || !na || !partials || grid < 1)
    return fail(kInvalid, "tv_gap: null pointer or empty grid");
  Cfg c = make_cfg(cfg);
  hipStream_t stm = as_stream(s);
  switch (cfg->C) {
    case 1: hipLaunchKernelGGL((tv_gap_kernel<T, 1>), dim3(grid), dim3(kTB), 0, stm, x, y, f, insts, ai, ao, na, c, partials); break;
    case 2: hipLaunchKernelGGL((tv_gap_kernel<T, 2>), dim3(grid), dim3(kTB), 0, stm, x, y, f, insts, ai, ao, na, c, partials); break;
    case 3: hipLaunchKernelGGL((tv_gap_kernel<T, 3>), dim3(grid), dim3(kTB), 0, stm, x, y, f, insts, ai, ao, na, c, partials); break;
    default: hipLaunchKernelGGL((tv_gap_kernel<T, 4>), dim3(grid), dim3(kTB), 0, stm, x, y, f, insts, ai, ao, na, c, partials); break;
  }
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
OPY_API int opy_tv_compact(const opy_tv_inst_t* insts, const opy_tv_check_t* chk, int B, int32_t* active_inst,
                           int32_t* act_off, int32_t* nact, int64_t* iter_counter, int iters_in_block, void* s) {
  if (!insts || !chk || !active_inst || !act_off || !nact || !iter_counter || B < 1)
    return fail(kInvalid, "tv_compact: bad arguments");
  hipLaunchKernelGGL(tv_compact_kernel, dim3(1), dim3(kTB), 0, as_stream(s), insts, chk, B, active_inst, act_off, nact,
                     iter_counter, iters_in_block);
  return check_launch("tv_compact");
}
