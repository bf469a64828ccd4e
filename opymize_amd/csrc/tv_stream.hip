// K fused TV iterations per pass over HBM, streamed down column strips
// (temporal blocking without LDS or barriers); see tv_common.h for the
// model, the per-pixel arithmetic and the running-instance lists.
#include "tv_common.h"

namespace opy {
namespace {

// Work unit: one wavefront per (strip, chunk) of an instance.  A strip is
// 64 - 2K image columns (the exact part), a chunk R image rows.  The 64
// lanes hold the strip and K columns on each side, one column per lane (the
// first and last strips of an instance hold K more exact columns instead of
// the halo beyond the image edge: stream_strips); the wavefront walks down
// the chunk and K rows above and below it (R + 2K rows), loading each row of
// x, f and y once.
//
// The K iterations run as a pipeline skewed by one row per level: after row
// n is loaded, level i (0-based) forms x^(i+1) at row n - i and y^(i+1) at
// row n - i - 1, from
//   x^i, y^i at row n - i (level i-1's outputs of this step; the loaded row
//   for i = 0), y^i at row n - i - 1 (kept from the previous step), f at row
//   n - i (a delay line of K rows), and x_bar^(i+1) at row n - i - 1 (kept).
// The left neighbour's y (for A^T y) and the right neighbour's x_bar (for
// A x_bar) come from the adjacent lanes by DPP wave shifts; the row above /
// below are the kept rows.  Level K-1's outputs at row n - K are the results
// (x^K one step older), stored for the strip's columns.  Exactness shrinks by
// one lane / row per level on each side exactly as in the tile kernel, so the
// strip's 64 - 2K middle lanes and the chunk's R middle rows are exact.
//
// Per pixel the arithmetic is that of the single-iteration kernels
// (primal_px, overrelax, dual_px with the folded scalar steps), so the
// iterates equal those of K single launches bit for bit; the products
// whose rounding matters are pinned in tv_common.h (mul_rn, overrelax,
// sqnorm2), since here the K iterations share one basic block.
//
// Loads run D rows ahead of the pipeline (D-row register ring).  Rows and
// columns beyond the image load as 0 (edge units only) and the boundary
// masks of the stencils keep them out of every exact value, as in the tile
// kernel.  A unit whose rows and columns, with a pixel to spare, lie inside
// the image runs without masks or bounds tests.
//
// Steps: as the tile kernel (the state is read, not advanced); lane i < K
// of each wavefront forms iteration i's folded steps in FP64 and the
// wavefront takes them from those lanes.
template <typename T, int C>
__device__ inline FoldedStep<T, C> bcast_step(const FoldedStep<T, C>& s, int l) {
  FoldedStep<T, C> o;
  o.a = readlane(s.a, l);
  o.b = readlane(s.b, l);
  o.theta = readlane(s.theta, l);
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int k = 0; k < C; ++k) {
      o.p[t][k] = readlane(s.p[t][k], l);
      o.q[t][k] = readlane(s.q[t][k], l);
    }
  return o;
}

// rows loaded ahead of the pipeline: 1 (no register ring to shift; 1-3% faster than 2 or 3 at
// 6 and 8 iterations per launch, equal below, where other wavefronts hide the loads' latency)
constexpr int kStreamD = 1;

// p + a 32-bit byte offset: a uniform base and a zero-extended vector offset
template <typename T>
__device__ inline T* byte_off(T* p, uint32_t bytes) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(p) + (size_t)bytes);
}
template <typename T>
__device__ inline const T* byte_off(const T* p, uint32_t bytes) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + (size_t)bytes);
}

// Strips of an instance W columns wide.  Plainly, strip s holds the exact
// columns [s sw, (s + 1) sw) (sw = 64 - 2K) on lanes K .. 63 - K, lane 0 on
// column s sw - K: ceil(W / sw) strips.  Extended, when that takes fewer
// strips: the first strip needs no halo beyond the left image edge, whose
// boundary the stencils' masks close, so lane 0 sits on column 0 and lanes
// 0 .. 63 - K are exact; the last (the one whose lanes reach column W - 1)
// needs none beyond the right edge, so its exact lanes run to 63; lane 0 of
// strip s sits on column s sw, and n strips cover 64 + (n - 1) sw columns.
// The plain layout keeps the strip edges where they are better aligned for
// the row loads and stores (3-7% faster at equal strip counts), which wins
// where HBM bounds the launch: the extended one is taken only from K = 6
// (sw <= 52), where instruction issue does (1024 columns at K = 8: 21
// strips instead of 22, 3.5% faster; at K = 2: 17 instead of 18, but 3.5%
// slower).
__host__ __device__ inline int stream_strips_plain(int W, int sw) { return (W + sw - 1) / sw; }
__host__ __device__ inline int stream_strips_ext(int W, int sw) { return W <= 64 ? 1 : (W - 64 + sw - 1) / sw + 1; }
__host__ __device__ inline bool stream_extended(int W, int sw) {
  return sw <= 52 && stream_strips_ext(W, sw) < stream_strips_plain(W, sw);
}
__host__ __device__ inline int stream_strips(int W, int sw) {
  return stream_extended(W, sw) ? stream_strips_ext(W, sw) : stream_strips_plain(W, sw);
}

// plain vector stores (the rows of adjacent strips share cache lines, which
// the L2 merges; measured no slower than streaming stores here)
template <typename T, int N>
__device__ inline void store_v(T* __restrict__ p, const T* v) {
  if constexpr (sizeof(T) == 4 && N % 2 == 0) {
#pragma unroll
    for (int k = 0; k < N; k += 2) *reinterpret_cast<float2*>(p + k) = make_float2(v[k], v[k + 1]);
  } else {
#pragma unroll
    for (int k = 0; k < N; ++k) p[k] = v[k];
  }
}

template <typename T, int C, int K>
__global__ void __launch_bounds__(kTB) tv_stream_kernel(const T* __restrict__ xin, T* __restrict__ xout,
                                                        const T* __restrict__ yin, T* __restrict__ yout,
                                                        const T* __restrict__ f, const opy_tv_inst_t* __restrict__ insts,
                                                        const opy_adapt_state_t* __restrict__ st,
                                                        const int32_t* __restrict__ active_inst,
                                                        const int32_t* __restrict__ unit_off,
                                                        const int32_t* __restrict__ nact_p, Cfg cfg, int R) {
  constexpr int D = kStreamD;
  constexpr int SW = 64 - 2 * K;  // exact columns per strip
  const int nact = *nact_p;
  if (nact <= 0) return;
  const int total = unit_off[nact];
  const int NSTEP = R + 2 * K;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int u = xcd_block() * (kTB / 64) + wave; u < total; u += gridDim.x * (kTB / 64)) {
    TileRef tr = locate(u, active_inst, unit_off, nact, insts, SW, R);
    {
      const int ns = stream_strips(insts[tr.b].W, SW), ty = tr.j / ns;
      tr.r0 = ty * R;
      tr.c0 = (tr.j - ty * ns) * SW;
    }
    const opy_tv_inst_t I = insts[tr.b];
    const int H = I.H, W = I.W;
    const bool ext = stream_extended(W, SW);
    const int r0 = tr.r0 - K, c0 = ext ? tr.c0 : tr.c0 - K;  // image row of pipeline row 0, column of lane 0
    // exact lanes, stored: [slo, shi)
    const int slo = (ext && c0 == 0) ? 0 : K, shi = (ext && c0 + 64 >= W) ? 64 : 64 - K;
    const int c = c0 + lane;
    FoldedStep<T, C> fs[K];
    {
      FoldedStep<T, C> mine = {};
      if (lane < K) {  // iteration `lane`: advance the state by `lane` iterations, then fold
        double tau_d = st[tr.b].tau, sig_d = st[tr.b].sigma;
        for (int m = 0; m < lane; ++m) {
          const double th_m = accel_theta(cfg.steps, I, tau_d);
          tau_d *= th_m;
          sig_d /= th_m;
        }
        const double th_d = accel_theta(cfg.steps, I, tau_d);
        mine = fold_step<T, C>(tau_d, sig_d / th_d, th_d, I, cfg);
      }
#pragma unroll
      for (int i = 0; i < K; ++i) fs[i] = bcast_step(mine, i);
    }
    const T tvw = (T)I.tvw;
    // the unit's columns reach past the image (the first and last strips)
    const bool cedge = !(c0 >= 1 && c0 + 64 < W);
    const T* xb = xin + I.off;
    const T* fb = f + I.off;
    const T* yb = yin + 2 * I.off;
    T* xob = xout + I.off;
    T* yob = yout + 2 * I.off;
    // L1: data kind; MC: column masks and column bounds tests (uniform over the unit)
    auto run = [&](auto l1_tag, auto mc_tag) {
      constexpr bool L1 = decltype(l1_tag)::value, MC = decltype(mc_tag)::value;
      // per level: x^(i+1) and x_bar^(i+1) of its last row, y^i of the row
      // above its current one, f delayed by i rows
      T Xo[K][C], XB[K][C], YL[K][2 * C], F[K][C];
#pragma unroll
      for (int i = 0; i < K; ++i) {
#pragma unroll
        for (int k = 0; k < C; ++k) Xo[i][k] = XB[i][k] = F[i][k] = T(0);
#pragma unroll
        for (int e = 0; e < 2 * C; ++e) YL[i][e] = T(0);
      }
      T nx[D][C], nf[D][C], ny[D][2 * C];
      // LB: bounds tests (the row's, uniform, and the lane's column when MC)
      auto load_row = [&](auto lb_tag, int n, T* x, T* fv, T* y) {
        constexpr bool LB = decltype(lb_tag)::value;
        const int r = r0 + n;
        if (!LB || (r >= 0 && r < H && (!MC || (c >= 0 && c < W)))) {
          const uint32_t p = (uint32_t)(r * W) + (uint32_t)c;
          load_c<T, 2 * C>(byte_off(yb, p * (uint32_t)(2 * C * sizeof(T))), y);
          load_c<T, C>(byte_off(xb, p * (uint32_t)(C * sizeof(T))), x);
          load_c<T, C>(byte_off(fb, p * (uint32_t)(C * sizeof(T))), fv);
        } else {
#pragma unroll
          for (int k = 0; k < C; ++k) x[k] = fv[k] = T(0);
#pragma unroll
          for (int e = 0; e < 2 * C; ++e) y[e] = T(0);
        }
      };
      // one pipeline step; MR: row masks and row bounds tests of the load
      auto step = [&](auto mr_tag, int n) {
        constexpr bool MR = decltype(mr_tag)::value;
        T xcur[C], ycur[2 * C];
#pragma unroll
        for (int i = K - 1; i > 0; --i)
#pragma unroll
          for (int k = 0; k < C; ++k) F[i][k] = F[i - 1][k];
#pragma unroll
        for (int k = 0; k < C; ++k) {
          xcur[k] = nx[0][k];
          F[0][k] = nf[0][k];
        }
#pragma unroll
        for (int e = 0; e < 2 * C; ++e) ycur[e] = ny[0][e];
#pragma unroll
        for (int d = 0; d + 1 < D; ++d) {
#pragma unroll
          for (int k = 0; k < C; ++k) {
            nx[d][k] = nx[d + 1][k];
            nf[d][k] = nf[d + 1][k];
          }
#pragma unroll
          for (int e = 0; e < 2 * C; ++e) ny[d][e] = ny[d + 1][e];
        }
        if (n + D < NSTEP) load_row(mr_tag, n + D, nx[D - 1], nf[D - 1], ny[D - 1]);
#pragma unroll
        for (int i = 0; i < K; ++i) {
          const int rx = r0 + n - i, ry = rx - 1;  // rows of x^(i+1) and y^(i+1) of this step
          const FoldedStep<T, C>& s = fs[i];
          const T theta1 = T(1) + s.theta;
          T xnew[C], xbar[C], xbr[C], yp[2 * C];
#pragma unroll
          for (int k = 0; k < C; ++k) {
            const T yl1 = dpp_shr1(T(0), ycur[C + k]);  // y1 of the left neighbour
            xnew[k] = primal_px_rc<T, C, L1, MR, MC>(xcur[k], F[i][k], ycur[k], ycur[C + k], YL[i][k], yl1, rx, c,
                                                     H, W, k, s);
            xbar[k] = overrelax(theta1, s.theta, xnew[k], xcur[k]);
            xbr[k] = dpp_shl1(T(0), XB[i][k]);  // x_bar of the right neighbour, row ry
          }
          dual_px_rc<T, C, MR, MC>(XB[i], xbar, xbr, YL[i], ry, c, H, W, s, tvw, yp);
#pragma unroll
          for (int e = 0; e < 2 * C; ++e) {
            YL[i][e] = ycur[e];
            ycur[e] = yp[e];
          }
#pragma unroll
          for (int k = 0; k < C; ++k) {
            XB[i][k] = xbar[k];
            const T xo = Xo[i][k];
            Xo[i][k] = xnew[k];
            xcur[k] = xo;
          }
        }
        // xcur = x^K, ycur = y^K at pipeline row n - K
        const int m = n - K;
        const int r = r0 + m;
        if (m >= K && r < H && lane >= slo && lane < shi && c < W) {
          const uint32_t p = (uint32_t)(r * W) + (uint32_t)c;
          store_v<T, C>(byte_off(xob, p * (uint32_t)(C * sizeof(T))), xcur);
          store_v<T, 2 * C>(byte_off(yob, p * (uint32_t)(2 * C * sizeof(T))), ycur);
        }
      };
      using TT = std::true_type;
      using FF = std::false_type;
#pragma unroll
      for (int d = 0; d < D; ++d) load_row(TT{}, d, nx[d], nf[d], ny[d]);
      // Steps whose rows (every level's, and the row loaded) lie inside the
      // image with a row to spare run without row masks and row bounds
      // tests: steps [nA, nB).  Level i's rows are r0 + n - i and the one
      // above, so n >= K - r0 and n <= H - 2 - r0; the row loaded is
      // r0 + n + D when n + D < NSTEP.  Units away from the top and bottom
      // edges run every step there.
      int nA = NSTEP, nB = NSTEP;
      if (!MC) {
        nA = min(NSTEP, max(0, K - r0));
        nB = min(NSTEP, H - 1 - r0);
        if (r0 + NSTEP > H) nB = min(nB, H - r0 - D);
        nB = max(nA, nB);
      }
      int n = 0;
#pragma unroll 1
      for (; n < nA; ++n) step(TT{}, n);
#pragma unroll 1
      for (; n < nB; ++n) step(FF{}, n);
#pragma unroll 1
      for (; n < NSTEP; ++n) step(TT{}, n);
    };
    using TT = std::true_type;
    using FF = std::false_type;
    if (I.data == OPY_TV_DATA_L1) {
      if (cedge) run(TT{}, TT{});
      else run(TT{}, FF{});
    } else {
      if (cedge) run(FF{}, TT{});
      else run(FF{}, FF{});
    }
  }
}

// unit_off[a] = units of the running instances before a (exclusive prefix),
// unit_off[nact] = all of them; one block
__global__ void __launch_bounds__(kTB) tv_stream_units_kernel(const opy_tv_inst_t* __restrict__ insts,
                                                              const int32_t* __restrict__ active_inst,
                                                              const int32_t* __restrict__ nact_p,
                                                              int32_t* __restrict__ unit_off, int sw, int rows) {
  __shared__ int part[kTB];
  const int nact = *nact_p;
  const int per = (nact + kTB - 1) / kTB;
  const int lo = threadIdx.x * per;
  const int hi = min(nact, lo + per);
  int n = 0;
  for (int a = lo; a < hi; ++a) {
    const opy_tv_inst_t& I = insts[active_inst[a]];
    n += stream_strips(I.W, sw) * ((I.H + rows - 1) / rows);
  }
  part[threadIdx.x] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int i = 0; i < kTB; ++i) {
      const int v = part[i];
      part[i] = t;
      t += v;
    }
    unit_off[nact > 0 ? nact : 0] = t;
  }
  __syncthreads();
  int t = part[threadIdx.x];
  for (int a = lo; a < hi; ++a) {
    unit_off[a] = t;
    const opy_tv_inst_t& I = insts[active_inst[a]];
    t += stream_strips(I.W, sw) * ((I.H + rows - 1) / rows);
  }
}

template <typename T, int C, int K>
void launch_stream(const T* xi, T* xo, const T* yi, T* yo, const T* f, const opy_tv_inst_t* insts,
                   const opy_adapt_state_t* st, const int32_t* ai, const int32_t* uo, const int32_t* na, const Cfg& c,
                   int rows, int grid, hipStream_t stm) {
  hipLaunchKernelGGL((tv_stream_kernel<T, C, K>), dim3(launch_grid(KF(tv_stream_kernel<T, C, K>), grid)), dim3(kTB),
                     0, stm, xi, xo, yi, yo, f, insts, st, ai, uo, na, c, rows);
}

constexpr int kStreamMaxRows = 1 << 20;
constexpr int kStreamMaxK = 8;  // iterations per launch (k > 4 for C <= 2: registers)

template <typename T>
int stream_impl(const T* xi, T* xo, const T* yi, T* yo, const T* f, const opy_tv_inst_t* insts,
                const opy_adapt_state_t* st, const int32_t* ai, const int32_t* uo, const int32_t* na,
                const opy_tv_cfg_t* cfg, int k, int rows, int grid, void* s) {
  if (int rc = check_cfg(cfg)) return rc;
  if (cfg->steps == OPY_TV_STEPS_PRECOND)
    return fail(kInvalid, "tv_stream: preconditioned steps vary per pixel; use opy_tv_multi");
  if (k < 2 || k > kStreamMaxK) return fail(kInvalid, "tv_stream: k = %d not in [2, %d]", k, kStreamMaxK);
  if (k > 4 && cfg->C > 2) return fail(kInvalid, "tv_stream: k = %d > 4 needs C <= 2 (got C = %d)", k, cfg->C);
  if (rows < 1 || rows > kStreamMaxRows) return fail(kInvalid, "tv_stream: rows = %d not in [1, %d]", rows, kStreamMaxRows);
  if (!xi || !xo || !yi || !yo || !f || !insts || !st || !ai || !uo || !na || grid < 0)
    return fail(kInvalid, "tv_stream: null pointer or negative grid");
  if (xi == xo || yi == yo) return fail(kInvalid, "tv_stream: ping-pong buffers must differ");
  Cfg c = make_cfg(cfg);
  hipStream_t stm = as_stream(s);
#define OPY_S(CC, KK) launch_stream<T, CC, KK>(xi, xo, yi, yo, f, insts, st, ai, uo, na, c, rows, grid, stm)
#define OPY_SK(CC)                  \
  switch (k) {                      \
    case 2: OPY_S(CC, 2); break;    \
    case 3: OPY_S(CC, 3); break;    \
    default: OPY_S(CC, 4); break;   \
  }
#define OPY_SK8(CC)                 \
  switch (k) {                      \
    case 2: OPY_S(CC, 2); break;    \
    case 3: OPY_S(CC, 3); break;    \
    case 4: OPY_S(CC, 4); break;    \
    case 5: OPY_S(CC, 5); break;    \
    case 6: OPY_S(CC, 6); break;    \
    case 7: OPY_S(CC, 7); break;    \
    default: OPY_S(CC, 8); break;   \
  }
  switch (cfg->C) {
    case 1: OPY_SK8(1); break;
    case 2: OPY_SK8(2); break;
    case 3: OPY_SK(3); break;
    default: OPY_SK(4); break;
  }
#undef OPY_SK8
#undef OPY_SK
#undef OPY_S
  return check_launch("tv_stream");
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_tv_stream_units(const opy_tv_inst_t* insts, const int32_t* active_inst, const int32_t* nact,
                                int32_t* unit_off, int k, int rows, void* s) {
  if (!insts || !active_inst || !nact || !unit_off) return fail(kInvalid, "tv_stream_units: null pointer");
  if (k < 2 || k > kStreamMaxK) return fail(kInvalid, "tv_stream_units: k = %d not in [2, %d]", k, kStreamMaxK);
  if (rows < 1 || rows > kStreamMaxRows)
    return fail(kInvalid, "tv_stream_units: rows = %d not in [1, %d]", rows, kStreamMaxRows);
  hipLaunchKernelGGL(tv_stream_units_kernel, dim3(1), dim3(kTB), 0, as_stream(s), insts, active_inst, nact, unit_off,
                     64 - 2 * k, rows);
  return check_launch("tv_stream_units");
}

OPY_API int opy_tv_stream_f32(const float* xi, float* xo, const float* yi, float* yo, const float* f,
                              const opy_tv_inst_t* insts, const opy_adapt_state_t* st, const int32_t* active_inst,
                              const int32_t* unit_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int k, int rows,
                              int grid, void* s) {
  return stream_impl<float>(xi, xo, yi, yo, f, insts, st, active_inst, unit_off, nact, cfg, k, rows, grid, s);
}
OPY_API int opy_tv_stream_f64(const double* xi, double* xo, const double* yi, double* yo, const double* f,
                              const opy_tv_inst_t* insts, const opy_adapt_state_t* st, const int32_t* active_inst,
                              const int32_t* unit_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int k, int rows,
                              int grid, void* s) {
  return stream_impl<double>(xi, xo, yi, yo, f, insts, st, active_inst, unit_off, nact, cfg, k, rows, grid, s);
}
