// FP64 reductions: the PDHG residual norms that drive adaptive steps, the
// device-side step-size rule, and the functional values / infeasibilities
// that make up the primal-dual gap.
//
// Reductions are two-stage with a fixed grid, so a given input always sums in
// the same order (bitwise reproducible).  Stage 1 writes per-block partials to
// the caller's work buffer (OPY_RED_BLOCKS * 3 doubles), stage 2 folds them.
#include "opy_common.h"

#include <algorithm>
#include <cmath>
#include <mutex>

namespace opy {
namespace {

constexpr int kRedBlocks = OPY_RED_BLOCKS;

template <typename T>
struct StepR {
  const T* arr;
  const double* dev;
  double host;
  __device__ inline double at(int64_t i) const {
    if (arr) return (double)arr[i];
    if (dev) return *dev;
    return host;
  }
};

// Stage 1: every thread accumulates (sum, max1, max2) over a grid-stride
// range, the block folds them and writes one triple of partials.
template <typename Op>
__global__ void __launch_bounds__(kBlock) red_stage1(Op op, int64_t n, double* __restrict__ work) {
  __shared__ double sm[kBlock / 64];
  double s = 0.0, m1 = 0.0, m2 = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) op(i, s, m1, m2);
  s = block_sum(s, sm);
  m1 = block_max(m1, sm);
  m2 = block_max(m2, sm);
  if (threadIdx.x == 0) {
    work[3 * blockIdx.x + 0] = s;
    work[3 * blockIdx.x + 1] = m1;
    work[3 * blockIdx.x + 2] = m2;
  }
}

// Stage 2: fold nb partial triples.  mode: 0 overwrite, 1 accumulate into out.
__global__ void __launch_bounds__(kBlock) red_stage2(const double* __restrict__ work, int nb,
                                                     double* __restrict__ out, int nout, int accumulate) {
  __shared__ double sm[kBlock / 64];
  double s = 0.0, m1 = 0.0, m2 = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    s += work[3 * b + 0];
    m1 = fmax(m1, work[3 * b + 1]);
    m2 = fmax(m2, work[3 * b + 2]);
  }
  s = block_sum(s, sm);
  m1 = block_max(m1, sm);
  m2 = block_max(m2, sm);
  if (threadIdx.x == 0) {
    if (accumulate) {
      out[0] += s;
      if (nout > 1) out[1] = fmax(out[1], m1);
      if (nout > 2) out[2] = fmax(out[2], m2);
    } else {
      out[0] = s;
      if (nout > 1) out[1] = m1;
      if (nout > 2) out[2] = m2;
    }
  }
}

template <typename Op>
int run_reduction(const Op& op, int64_t n, double* work, double* out, int nout, int accumulate, void* s,
                  const char* what) {
  if (!work || !out) return fail(kInvalid, "%s: null work/out buffer", what);
  int64_t nb = (n + kBlock - 1) / kBlock;
  if (nb > kRedBlocks) nb = kRedBlocks;
  if (nb < 1) nb = 1;
  hipStream_t st = as_stream(s);
  hipLaunchKernelGGL(red_stage1<Op>, dim3((unsigned)nb), dim3(kBlock), 0, st, op, n, work);
  hipLaunchKernelGGL(red_stage2, dim3(1), dim3(kBlock), 0, st, work, (int)nb, out, nout, accumulate);
  return check_launch(what);
}

// ---- dot product in FP64 (the conjugate gradients, Armijo slopes and
// residual norms of semismooth_newton.py:107-114, :291, :314, :322) -------
template <typename T>
struct DotOp {
  const T* x;
  const T* y;
  __device__ inline void operator()(int64_t i, double& s, double& m1, double& m2) const {
    s += (double)x[i] * (double)y[i];
  }
};

template <typename T>
int dot_impl(const T* x, const T* y, int64_t n, double* work, double* out, void* s) {
  if (n < 0) return fail(kInvalid, "dot: negative size");
  if (n > 0 && (!x || !y)) return fail(kInvalid, "dot: null operand");
  return run_reduction(DotOp<T>{x, y}, n, work, out, 1, 0, s, "dot");
}

// ---- PDHG residuals (reference pdhg.py:74-78, :111-116, :155-164) -------
// sums of ((z - zn)/a - b*g + c1*gn)^2 and ((z - zn)/a - b*g + c2*gn)^2
template <typename T>
struct Residual2Op {
  StepR<T> a;
  const T* z;
  const T* zn;
  const T* g;
  const T* gn;
  T b, c1, c2;
  __device__ inline void operator()(int64_t i, double& s, double& m1, double& m2) const {
    const T ai = (T)a.at(i);
    const T d = (z[i] - zn[i]) / ai - b * g[i];
    const T r1 = d + c1 * gn[i];
    const T r2 = d + c2 * gn[i];
    s += (double)r1 * (double)r1;
    m1 += (double)r2 * (double)r2;  // second sum carried in the m1 slot (no max here)
  }
};

// stage 1 for residuals needs two sums; reuse the triple but fold slot 1 with +
template <typename T>
__global__ void __launch_bounds__(kBlock) residual2_stage1(Residual2Op<T> op, int64_t n, double* __restrict__ work) {
  __shared__ double sm[kBlock / 64];
  double s = 0.0, s2 = 0.0, unused = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) op(i, s, s2, unused);
  s = block_sum(s, sm);
  s2 = block_sum(s2, sm);
  if (threadIdx.x == 0) {
    work[2 * blockIdx.x + 0] = s;
    work[2 * blockIdx.x + 1] = s2;
  }
}

__global__ void __launch_bounds__(kBlock) sum2_stage2(const double* __restrict__ work, int nb, double* __restrict__ out) {
  __shared__ double sm[kBlock / 64];
  double s = 0.0, s2 = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    s += work[2 * b + 0];
    s2 += work[2 * b + 1];
  }
  s = block_sum(s, sm);
  s2 = block_sum(s2, sm);
  if (threadIdx.x == 0) {
    out[0] = s;
    out[1] = s2;
  }
}

// Device-side adaptive step rule (reference pdhg.py:166-183).
__global__ void adapt_kernel(opy_adapt_state_t* st, opy_adapt_params_t p) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double resp = sqrt(st->sums[0]);
  const double gnp = sqrt(st->sums[1]);
  const double resd = sqrt(st->sums[2]);
  const double gnd = sqrt(st->sums[3]);
  const double epsp = p.sqrt_n * p.eps_absp + p.eps_relp * gnp;
  const double epsd = p.sqrt_m * p.eps_absd + p.eps_reld * gnd;
  const double scale = epsp / epsd;
  st->resp = resp;
  st->resd = resd;
  st->epsp = epsp;
  st->epsd = epsd;
  // The two tests exclude each other because Delta > 1.
  if (resp > scale * resd * p.Delta) {
    const double a = st->alpha;
    st->tau /= 1.0 - a;
    st->sigma *= 1.0 - a;
    st->alpha = a * p.eta;
    st->nchanges += 1;
  }
  if (resp < scale * resd / p.Delta) {
    const double a = st->alpha;
    st->tau *= 1.0 - a;
    st->sigma /= 1.0 - a;
    st->alpha = a * p.eta;
    st->nchanges += 1;
  }
}

// ---- functional values for the primal-dual gap --------------------------
// Each evaluation yields (sum of values, max infeasibility 1, max infeasibility 2).
template <typename T>
struct FnOp {
  int kind;
  const T* x;
  int width;  // values per point
  opy_fn_params_t p;
  __device__ inline void operator()(int64_t i, double& s, double& m1, double& m2) const {
    const T* xi = x + i * width;
    switch (kind) {
      case OPY_FN_SSD: {  // 0.5 sum_k vol_k (x - f)^2 on unmasked points
        if (p.mask && !p.mask[i]) return;
        const T* fi = static_cast<const T*>(p.f) + i * width;
        double v = 0.0;
        for (int k = 0; k < width; ++k) {
          const double d = (double)xi[k] - (double)fi[k];
          v += d * d * p.vol[k];
        }
        s += 0.5 * v;
        return;
      }
      case OPY_FN_L1DATA: {  // sum_k lbd_k |x - f|
        const T* fi = static_cast<const T*>(p.f) + i * width;
        double v = 0.0;
        for (int k = 0; k < width; ++k) v += p.vol[k] * fabs((double)xi[k] - (double)fi[k]);
        s += v;
        return;
      }
      case OPY_FN_L1DATA_CONJ: {  // sum z f, infeas max(0, |z| - lbd)
        const T* fi = static_cast<const T*>(p.f) + i * width;
        double v = 0.0, inf = 0.0;
        for (int k = 0; k < width; ++k) {
          const double z = (double)xi[k];
          v += z * (double)fi[k];
          inf = fmax(inf, fabs(z) - p.vol[k]);
        }
        s += v;
        m1 = fmax(m1, inf);
        return;
      }
      case OPY_FN_NORMS: {  // per-point matrix norms (M1 x M2): sum of dual norm, max(0, norm - lbd)
        // p.m1, p.m2: matrix shape; p.matrixnorm: 'F' frobenius / 'N' nuclear (sum) + spectral (infeas)
        double fro2 = 0.0;
        for (int k = 0; k < width; ++k) fro2 += (double)xi[k] * (double)xi[k];
        double sumnorm, ballnorm;
        if (p.matrixnorm == 'F' || p.m1 == 1 || p.m2 == 1) {
          sumnorm = ballnorm = sqrt(fro2);
        } else {
          // singular values of the (M1 x M2) block (stored column by column, side of length 2)
          // from its 2x2 Gram matrix; same line convention as the spectral projection kernel
          double c11 = 0.0, c12 = 0.0, c22 = 0.0;
          const bool two_rows = (p.m1 <= p.m2);
          const int lim = two_rows ? p.m2 : p.m1;
          const int step_m = two_rows ? p.m1 : 1;
          const int step_j = two_rows ? 1 : p.m1;
          for (int r = 0; r < lim; ++r) {
            const double a0 = (double)xi[r * step_m];
            const double a1 = (double)xi[r * step_m + step_j];
            c11 += a0 * a0;
            c12 += a0 * a1;
            c22 += a1 * a1;
          }
          const double tr = c11 + c22;
          const double dd = sqrt(fmax(0.0, 0.25 * tr * tr - (c11 * c22 - c12 * c12)));
          const double smax = sqrt(fmax(0.0, 0.5 * tr + dd));
          const double smin = sqrt(fmax(0.0, 0.5 * tr - dd));
          sumnorm = smax + smin;  // nuclear
          ballnorm = smax;        // spectral
        }
        s += sumnorm;
        m1 = fmax(m1, ballnorm - p.lbd);
        return;
      }
      case OPY_FN_DOT: {  // sum c*x (c array or scalar)
        double v = 0.0;
        for (int k = 0; k < width; ++k) {
          const double c = p.c ? (double)static_cast<const T*>(p.c)[i * width + k] : p.cs;
          v += c * (double)xi[k];
        }
        s += v;
        return;
      }
      case OPY_FN_INDICATOR: {  // max |x - c| over (masked) rows
        if (p.mask && !p.mask[i]) return;
        double inf = 0.0;
        for (int k = 0; k < width; ++k) {
          const double c = p.c ? (double)static_cast<const T*>(p.c)[i * width + k] : p.cs;
          inf = fmax(inf, fabs((double)xi[k] - c));
        }
        m1 = fmax(m1, inf);
        return;
      }
      case OPY_FN_MASKED_AFFINE: {  // sum_{mask} c x ; max_{~mask} |x|
        const T* c = static_cast<const T*>(p.c) + i * width;
        if (p.mask[i]) {
          double v = 0.0;
          for (int k = 0; k < width; ++k) v += (double)c[k] * (double)xi[k];
          s += v;
        } else {
          double inf = 0.0;
          for (int k = 0; k < width; ++k) inf = fmax(inf, fabs((double)xi[k]));
          m1 = fmax(m1, inf);
        }
        return;
      }
      case OPY_FN_POS: {  // max(0, -x)
        double inf = 0.0;
        for (int k = 0; k < width; ++k) inf = fmax(inf, -(double)xi[k]);
        m1 = fmax(m1, inf);
        return;
      }
      case OPY_FN_NEG: {  // max(0, x)
        double inf = 0.0;
        for (int k = 0; k < width; ++k) inf = fmax(inf, (double)xi[k]);
        m1 = fmax(m1, inf);
        return;
      }
      case OPY_FN_QUADEPI_IND: {  // max(0, 0.5 a |x1|^2 + <b,x1> + c - x2)
        const int M = width - 1;
        double fx = p.qc ? p.qc[i] : 0.0;
        for (int k = 0; k < M; ++k) {
          const double x1 = (double)xi[k];
          fx += 0.5 * p.qa * x1 * x1 + (p.qb ? p.qb[i * M + k] : 0.0) * x1;
        }
        m1 = fmax(m1, fx - (double)xi[M]);
        return;
      }
      case OPY_FN_QUADEPI_SUPP: {  // perspective of f_i on x2 < -1e-8, else infeasible by |x2|
        const int M = width - 1;
        const double x2 = -(double)xi[M];
        if ((double)xi[M] < -1e-8) {
          double v = 0.0;
          for (int k = 0; k < M; ++k) {
            const double x1 = (double)xi[k];
            v += 0.5 * p.qa * x1 * x1 / x2 + (p.qb ? p.qb[i * M + k] : 0.0) * x1;
          }
          v += x2 * (p.qc ? p.qc[i] : 0.0);
          s += v;
        } else {
          m1 = fmax(m1, fabs((double)xi[M]));
        }
        return;
      }
      case OPY_FN_HUBER: {  // Huber perspective (reference epigraph.py:224-236)
        const int M = width - 1;
        double n2 = 0.0;
        for (int k = 0; k < M; ++k) n2 += (double)xi[k] * (double)xi[k];
        const double nrm = sqrt(n2);
        const double x2 = -(double)xi[M];
        const bool q = nrm < p.qa * x2;
        if (q) s += p.lbd * 0.5 / p.qa * n2 / x2;
        else if (x2 > -1e-4) s += p.lbd * (nrm - x2 * p.qa / 2.0);
        m1 = fmax(m1, -x2);
        return;
      }
      case OPY_FN_TRUNCQUADEPI: {  // two infeasibilities (reference epigraph.py:256-270)
        const int M = width - 1;
        double n2 = 0.0;
        for (int k = 0; k < M; ++k) n2 += (double)xi[k] * (double)xi[k];
        const double nrm = sqrt(n2);
        m1 = fmax(m1, nrm - p.lbd);
        const double tn = fmin(p.lbd, nrm);
        m2 = fmax(m2, 0.5 * p.qa * tn * tn - (double)xi[M]);
        return;
      }
      case OPY_FN_EPIGRAPH_IND: {  // max(0, <v_l, x[:-1]> - x[-1] - b[fun, l]) over the region's points
        const int nd = width - 1;
        const int64_t j = i / p.nfuns, f = i - j * p.nfuns;
        const int32_t* jr = p.ej + j * p.nsub;
        double inf = 0.0;
        for (int k = 0; k < p.nsub; ++k) {
          const int l = jr[k];
          if (!p.ei[f * p.npoints + l]) continue;
          double a = -(double)xi[nd] - p.eb[f * p.npoints + l];
          for (int d = 0; d < nd; ++d) a += p.ev[(int64_t)l * nd + d] * (double)xi[d];
          inf = fmax(inf, a);
        }
        m1 = fmax(m1, inf);
        return;
      }
      case OPY_FN_EPIGRAPH_SUPP: {  // xt * max_faces <eq, (x[:-1]/xt, 1)>, xt = max(eps, |x[-1]|)
        const int D = width;
        const int nd = D - 1;
        const int64_t j = i / p.nfuns;
        const double* C = p.checks + j * D * D;
        double inf = 0.0;
        for (int r = 0; r < D; ++r) {
          double c = 0.0;
          for (int d = 0; d < D; ++d) c += C[r * D + d] * (double)xi[d];
          inf = fmax(inf, -c);
        }
        m1 = fmax(m1, inf);
        const double xt = fmax(2.220446049250313e-16, fabs((double)xi[nd]));
        const int64_t f0 = p.eq_off[i], f1 = p.eq_off[i + 1];
        if (f1 > f0) {
          double best = -INFINITY;
          for (int64_t q = f0; q < f1; ++q) {
            const double* e = p.eq + q * D;
            double v = e[nd];
            for (int d = 0; d < nd; ++d) v += e[d] * ((double)xi[d] / xt);
            best = fmax(best, v);
          }
          s += xt * best;
        }
        return;
      }
      case OPY_FN_BOX_IND: {  // max(0, lo - x, x - hi) with lo = p.cs, hi = p.lbd
        double inf = 0.0;
        for (int k = 0; k < width; ++k) {
          const double v = (double)xi[k];
          inf = fmax(inf, fmax(p.cs - v, v - p.lbd));
        }
        m1 = fmax(m1, inf);
        return;
      }
      case OPY_FN_BOX_SUPP: {  // support function of [lo, hi]^n: sum max(lo z, hi z)
        double v = 0.0;
        for (int k = 0; k < width; ++k) {
          const double z = (double)xi[k];
          v += fmax(p.cs * z, p.lbd * z);
        }
        s += v;
        return;
      }
      case OPY_FN_SIMPLEX_IND: {  // per point: max(|sum x - 1|, max(0, -x))
        double sum = 0.0, inf = 0.0;
        for (int k = 0; k < width; ++k) {
          const double v = (double)xi[k];
          sum += v;
          inf = fmax(inf, -v);
        }
        m1 = fmax(m1, fmax(inf, fabs(sum - 1.0)));
        return;
      }
      case OPY_FN_SIMPLEX_SUPP: {  // support function of the simplex per point: max_k z_k
        double mx = (double)xi[0];
        for (int k = 1; k < width; ++k) mx = fmax(mx, (double)xi[k]);
        s += mx;
        return;
      }
      default:
        return;
    }
  }
};

template <typename T>
int fn_eval_impl(int kind, const T* x, int64_t npts, int width, const opy_fn_params_t* prm, double* work,
                 double* out3, int accumulate, void* s) {
  if (npts < 0 || width < 1) return fail(kInvalid, "fn_eval: bad shape");
  if (npts > 0 && !x) return fail(kInvalid, "fn_eval: null input");
  if (!prm) return fail(kInvalid, "fn_eval: null params");
  switch (kind) {
    case OPY_FN_SSD:
    case OPY_FN_L1DATA:
    case OPY_FN_L1DATA_CONJ:
      if (!prm->f || !prm->vol) return fail(kInvalid, "fn_eval: data term needs f and weights");
      break;
    case OPY_FN_MASKED_AFFINE:
      if (!prm->c || !prm->mask) return fail(kInvalid, "fn_eval: masked affine needs c and mask");
      break;
    case OPY_FN_NORMS:
      if (prm->m1 * prm->m2 != width) return fail(kInvalid, "fn_eval: norm shape %dx%d != width %d", prm->m1, prm->m2, width);
      if (prm->matrixnorm != 'F' && prm->m1 != 1 && prm->m2 != 1 && prm->m1 != 2 && prm->m2 != 2)
        return fail(kInvalid, "fn_eval: spectral/nuclear norms need a side of length <= 2");
      break;
    case OPY_FN_DOT:
    case OPY_FN_INDICATOR:
    case OPY_FN_POS:
    case OPY_FN_NEG:
    case OPY_FN_QUADEPI_IND:
    case OPY_FN_QUADEPI_SUPP:
    case OPY_FN_HUBER:
    case OPY_FN_TRUNCQUADEPI:
    case OPY_FN_SIMPLEX_IND:
    case OPY_FN_SIMPLEX_SUPP:
      break;
    case OPY_FN_BOX_IND:
    case OPY_FN_BOX_SUPP:
      if (!(prm->cs <= prm->lbd)) return fail(kInvalid, "fn_eval: box needs lo <= hi");
      break;
    case OPY_FN_EPIGRAPH_IND:
      if (!prm->ei || !prm->ej || !prm->ev || !prm->eb || prm->nfuns < 1 || prm->npoints < 1 || prm->nsub < 1)
        return fail(kInvalid, "fn_eval: epigraph indicator needs I, J, v, b and their sizes");
      if (width < 2 || width > 4) return fail(kInvalid, "fn_eval: epigraph width %d outside [2,4]", width);
      if (npts % prm->nfuns) return fail(kInvalid, "fn_eval: %lld points not a multiple of %d functions",
                                         (long long)npts, prm->nfuns);
      break;
    case OPY_FN_EPIGRAPH_SUPP:
      if (!prm->eq || !prm->eq_off || !prm->checks || prm->nfuns < 1)
        return fail(kInvalid, "fn_eval: epigraph support needs pieces, offsets and checks");
      if (width < 2 || width > 4) return fail(kInvalid, "fn_eval: epigraph width %d outside [2,4]", width);
      if (npts % prm->nfuns) return fail(kInvalid, "fn_eval: %lld points not a multiple of %d functions",
                                         (long long)npts, prm->nfuns);
      break;
    default:
      return fail(kInvalid, "fn_eval: unknown kind %d", kind);
  }
  FnOp<T> op{kind, x, width, *prm};
  return run_reduction(op, npts, work, out3, 3, accumulate, s, "fn_eval");
}

template <typename T>
int residual2_impl(const opy_step_t* a, const T* z, const T* zn, double b, const T* g, double c1, double c2,
                   const T* gn, int64_t n, double* work, double* out2, void* s) {
  if (n < 0) return fail(kInvalid, "residual: negative size");
  if (n > 0 && (!z || !zn || !g || !gn)) return fail(kInvalid, "residual: null operand");
  if (!work || !out2) return fail(kInvalid, "residual: null work/out");
  StepR<T> st{nullptr, nullptr, 1.0};
  if (a) {
    st.arr = static_cast<const T*>(a->arr);
    st.dev = a->dev;
    st.host = a->host;
  }
  Residual2Op<T> op{st, z, zn, g, gn, (T)b, (T)c1, (T)c2};
  int64_t nb = (n + kBlock - 1) / kBlock;
  if (nb > kRedBlocks) nb = kRedBlocks;
  if (nb < 1) nb = 1;
  hipStream_t stm = as_stream(s);
  hipLaunchKernelGGL((residual2_stage1<T>), dim3((unsigned)nb), dim3(kBlock), 0, stm, op, n, work);
  hipLaunchKernelGGL(sum2_stage2, dim3(1), dim3(kBlock), 0, stm, work, (int)nb, out2);
  return check_launch("residual");
}

// ---- residuals without a caller-supplied work buffer ---------------------
// The partials of the two-stage residual reduction live in a file-scope
// device array instead of a workspace argument, for callers that cannot
// allocate (code under graph capture, serving loops).  Every device holds
// its own copy at its own address; the host never passes the array's name
// to a kernel, only the address hipGetSymbolAddress returns for the current
// device.  Stage 1 writes every slot stage 2 reads, so nothing needs
// clearing between calls; calls that overlap on one device race on it.
constexpr int kScratchDoubles = 2 * kRedBlocks;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxScratchDevices] = {nullptr};

int scratch_for_current_device(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(kHipError, "residual (no-alloc): hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxScratchDevices)
    return fail(kInvalid, "residual (no-alloc): device %d beyond the %d supported", dev, kMaxScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(kHipError, "residual (no-alloc): hipGetSymbolAddress on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}

}  // namespace
}  // namespace opy

using namespace opy;

OPY_API int opy_pdhg_residual2_noalloc_f32(const opy_step_t* a, const float* z, const float* zn, double b,
                                           const float* g, double c1, double c2, const float* gn, int64_t n,
                                           double* out2, void* s) {
  if (n < 0) return fail(kInvalid, "residual (no-alloc): negative size");
  // the reduction grid is clamped to kRedBlocks blocks, whose partials fill the array exactly
  const int64_t nb = std::min<int64_t>(std::max<int64_t>((n + kBlock - 1) / kBlock, 1), kRedBlocks);
  if (2 * nb > kScratchDoubles)
    return fail(kInvalid, "residual (no-alloc): %lld partials exceed the %d-double scratch array",
                (long long)(2 * nb), kScratchDoubles);
  double* work = nullptr;
  int rc = scratch_for_current_device(&work);
  if (rc != kOk) return rc;
  return residual2_impl<float>(a, z, zn, b, g, c1, c2, gn, n, work, out2, s);
}

OPY_API int opy_pdhg_residual2_f32(const opy_step_t* a, const float* z, const float* zn, double b, const float* g,
                                   double c1, double c2, const float* gn, int64_t n, double* work, double* out2,
                                   void* s) {
  return residual2_impl<float>(a, z, zn, b, g, c1, c2, gn, n, work, out2, s);
}
OPY_API int opy_pdhg_residual2_f64(const opy_step_t* a, const double* z, const double* zn, double b,
                                   const double* g, double c1, double c2, const double* gn, int64_t n,
                                   double* work, double* out2, void* s) {
  return residual2_impl<double>(a, z, zn, b, g, c1, c2, gn, n, work, out2, s);
}

OPY_API int opy_dot_f32(const float* x, const float* y, int64_t n, double* work, double* out, void* s) {
  return dot_impl<float>(x, y, n, work, out, s);
}
OPY_API int opy_dot_f64(const double* x, const double* y, int64_t n, double* work, double* out, void* s) {
  return dot_impl<double>(x, y, n, work, out, s);
}

OPY_API int opy_pdhg_adapt(opy_adapt_state_t* st, const opy_adapt_params_t* p, void* s) {
  if (!st || !p) return fail(kInvalid, "adapt: null state/params");
  hipLaunchKernelGGL(adapt_kernel, dim3(1), dim3(64), 0, as_stream(s), st, *p);
  return check_launch("adapt");
}

OPY_API int opy_fn_eval_f32(int kind, const float* x, int64_t npts, int width, const opy_fn_params_t* prm,
                            double* work, double* out3, int accumulate, void* s) {
  return fn_eval_impl<float>(kind, x, npts, width, prm, work, out3, accumulate, s);
}
OPY_API int opy_fn_eval_f64(int kind, const double* x, int64_t npts, int width, const opy_fn_params_t* prm,
                            double* work, double* out3, int accumulate, void* s) {
  return fn_eval_impl<double>(kind, x, npts, width, prm, work, out3, accumulate, s);
}
