// The single-pass fused TV iteration; see tv_common.h for the model, the
// layout and the tile map.
#include "tv_common.h"

namespace opy {
namespace {

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
template <typename T, int C, int RPT, bool kAdapt, bool PC>
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
  constexpr bool kFold = !kAdapt && !PC;  // scalar steps, no residuals: folded stencils
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
    const double th_d = accel_theta(cfg.steps, I, tau_s);
    const T theta = (T)th_d;                      // 1 unless accelerated
    const T sigma = (T)(st[tr.b].sigma / th_d);
    const PrimalStep<T, C, PC> prox(tau_s, I, cfg);
    const DualStep<T, C, PC> ds(sigma, cfg);
    const FoldedStep<T, C> fs = fold_step<T, C>(tau_s, st[tr.b].sigma / th_d, th_d, I, cfg);
    const T tvw = (T)I.tvw;
    const int c0 = tr.c0;
    const int c = c0 + lane;
    const bool col = c < W;
    const int rs = tr.r0 + ty * RPT;
    // with adaptive steps (whose residuals make the kernel issue-bound), a
    // tile whose rows and columns, with two to spare on every side, lie inside
    // the image runs the stencils without boundary masks; the other rules are
    // bound by HBM, and the unmasked variant measured no faster there
    const bool inner = kAdapt && tr.r0 >= 3 && tr.r0 + th + 3 <= H && tr.c0 >= 2 && tr.c0 + 66 <= W;
    auto body = [&](auto m_tag) {
      constexpr bool M = decltype(m_tag)::value;
      // x+ at (r, cc), channel k, through the prox of the instance's data term
      auto pstep = [&](T x, T aty, T fx, int r, int cc, int k) {
        return I.data == OPY_TV_DATA_L1 ? prox.template apply<true, M>(x, aty, fx, r, cc, H, W, k, cfg)
                                        : prox.template apply<false, M>(x, aty, fx, r, cc, H, W, k, cfg);
      };
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
        adjoint_px<T, C, M>(yv[j + 1], yv[j + 1] + C, yv[j], yl1, rs - OX + j, c, H, W, cfg, out);
      };
      // scalar steps (constant, accelerated): the folded stencil
      auto fold_prox = [&](T x, T fx, const T* y, const T* yup, const T* yl1, int r, int cc, int k) {
        return I.data == OPY_TV_DATA_L1
                   ? primal_px<T, C, true, M>(x, fx, y[k], y[C + k], yup[k], yl1[k], r, cc, H, W, k, fs)
                   : primal_px<T, C, false, M>(x, fx, y[k], y[C + k], yup[k], yl1[k], r, cc, H, W, k, fs);
      };
      T xp[NX][C];
#pragma unroll
      for (int j = 0; j < NX; ++j) {
        const int r = rs - OX + j;
        if constexpr (kFold) {
          T yl1[C];
#pragma unroll
          for (int k = 0; k < C; ++k) yl1[k] = dpp_shr1(readlane(ylh[C + k], j + 1), yv[j + 1][C + k]);
#pragma unroll
          for (int k = 0; k < C; ++k) xp[j][k] = fold_prox(xv[j][k], fv[j][k], yv[j + 1], yv[j], yl1, r, c, k);
        } else {
          T aty[C];
          aty_row(j, aty);
#pragma unroll
          for (int k = 0; k < C; ++k) xp[j][k] = pstep(xv[j][k], aty[k], fv[j][k], r, c, k);
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
        if (!kFold) adjoint_px<T, C, M>(yo, yo + C, yrh, yl1, rx, cr, H, W, cfg, ah);
#pragma unroll
        for (int k = 0; k < C; ++k) {
          const T p = kFold ? fold_prox(xrh[k], frh[k], yo, yrh, yl1, rx, cr, k)
                            : pstep(xrh[k], ah[k], frh[k], rx, cr, k);
          xbh[k] = overrelax(T(1) + theta, theta, p, xrh[k]);
          xdh[k] = xrh[k] - p;
        }
      }
      // xbar = 2x+ - x and xd = x - x+ (and their right neighbours) are formed
      // where they are used, which keeps fewer values live
      auto xbar = [&](int j, int k) { return overrelax(T(1) + theta, theta, xp[j][k], xv[j][k]); };
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
        adjoint_px<T, C, M>(yo, yo + C, ylh, yl1, rx, cl, H, W, cfg, ah);
#pragma unroll
        for (int k = 0; k < C; ++k) {
          const T p = pstep(xlh[k], ah[k], flh[k], rx, cl, k);
          xbl[k] = overrelax(T(1) + theta, theta, p, xlh[k]);
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
        yplus_px<T, C, PC, M>(xbl, xbld, xbc, yo, rx, cl, H, W, ds, tvw, cfg, ypl, gl);
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
        if constexpr (kFold) dual_px<T, C, M>(xbi, xbd, xbr, yv[i + 1], r, c, H, W, fs, tvw, yp);
        else yplus_px<T, C, PC, M>(xbi, xbd, xbr, yv[i + 1], r, c, H, W, ds, tvw, cfg, yp, gb);
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
            adjoint_px<T, C, M>(yp, yp + C, ypu, ypl1, r, c, H, W, cfg, atn);
#pragma unroll
            for (int k = 0; k < C; ++k) xdr[k] = dpp_shl1(readlane(xdh[k], i), xdi[k]);
            grad_px<T, C, M>(xdi, xdd, xdr, r, c, H, W, cfg, gxd);
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
    };
    if (inner) body(std::false_type{});
    else body(std::true_type{});
  }
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
  if (adapt && cfg->steps != OPY_TV_STEPS_SCALAR)
    return fail(kInvalid, "tv_iter: adaptive steps exclude precond and accelerated steps");
  Cfg c = make_cfg(cfg);
  hipStream_t stm = as_stream(s);
#define OPY_L(CC, RR)                                                                                       \
  hipLaunchKernelGGL((tv_iter_kernel<T, CC, RR, AD, PC>), dim3(launch_grid(KF(tv_iter_kernel<T, CC, RR, AD, PC>), grid)), \
                     dim3(kTB), 0, stm, xi, xo, yi, yo, f, insts, st, ai, ao, na, c, partials)
  if (adapt) {
    constexpr bool AD = true, PC = false;
    OPY_TV_DISPATCH(OPY_L)
  } else if (cfg->steps == OPY_TV_STEPS_PRECOND) {
    constexpr bool AD = false, PC = true;
    OPY_TV_DISPATCH(OPY_L)
  } else {
    constexpr bool AD = false, PC = false;
    OPY_TV_DISPATCH(OPY_L)
  }
#undef OPY_L
  return check_launch("tv_iter");
}

}  // namespace
}  // namespace opy

using namespace opy;

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
