/*
 * opymize_amd C-ABI: hand-written HIP kernels for the PDHG hot path of
 * Opymize (PDHG(G, F, A).solve on the GPU), built for MI355X (gfx950).
 *
 * Conventions
 *  - Every entry point returns an int status: 0 = ok, 1 = invalid argument,
 *    2 = HIP launch/runtime error.  opy_last_error() returns a thread-local
 *    message for the last failure.
 *  - Array arguments are device pointers (flat, C order).  "_f32"/"_f64"
 *    select the element type; FP64 accumulators are always double.
 *  - `stream` is a hipStream_t passed as void*; every entry point only
 *    enqueues work on it (no host synchronisation), so sequences of calls can
 *    be captured into a hipGraph.
 *  - `add` = 1 accumulates into the output (the reference's add=True),
 *    0 overwrites it (the reference's fill(0) followed by +=).
 *  - Outputs may alias inputs only where a function says so.
 *
 * The reference's device boundary is Operator.__call__ / LinOp.__call__
 * dispatching to _call_gpu (opymize/__init__.py:64-120), which launches
 * PyCUDA kernels compiled from the .cu files of opymize/linear and opymize/operators and the
 * string kernels of opymize/solvers/pdhg.py:61-79.  Each declaration below
 * names the reference code it replaces.
 */
#ifndef OPYMIZE_AMD_H
#define OPYMIZE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OPY_ABI_VERSION 3
#define OPY_MAX_DIMS 6
#define OPY_RED_BLOCKS 1024   /* reductions need work buffers of 3*OPY_RED_BLOCKS doubles */
#define OPY_QUADEPI_MAXM 16

#define OPY_SCHEME_FORWARD 0
#define OPY_SCHEME_CENTERED 1

/* Cartesian grid of a finite-difference operator (reference diff.py:44-65). */
typedef struct {
  int32_t ndims;
  int32_t nchannels;
  int64_t dims[OPY_MAX_DIMS];
  double h[OPY_MAX_DIMS];
} opy_grid_t;

/* Step size: per-element array `arr` (element type of the call), else the
 * double at device address `dev`, else `host`. */
typedef struct {
  const void* arr;
  const double* dev;
  double host;
} opy_step_t;

/* Device-resident adaptive step state (reference pdhg.py:212-214, :155-183). */
typedef struct {
  double tau, sigma, alpha;
  double resp, resd, epsp, epsd;
  double sums[4]; /* squared sums: res_p, gnorm_p, res_d, gnorm_d */
  int32_t nchanges;
  int32_t pad;
} opy_adapt_state_t;

typedef struct {
  double eta, Delta;
  double sqrt_n, sqrt_m;
  double eps_absp, eps_relp, eps_absd, eps_reld;
} opy_adapt_params_t;

/* Functional evaluation kinds for the FP64 primal-dual gap. */
enum {
  OPY_FN_SSD = 1,            /* SSD value                   (reference functionals/quad.py:33-44)      */
  OPY_FN_L1DATA = 2,         /* sum lbd |x-f| (TV-L1 data term, new)                                 */
  OPY_FN_L1DATA_CONJ = 3,    /* its conjugate: sum z f, infeas max(|z|-lbd)                          */
  OPY_FN_NORMS = 4,          /* L1Norms value / L1NormsConj infeas (reference l1norms.py:14-70)     */
  OPY_FN_DOT = 5,            /* AffineFct value             (reference functionals/affine.py:17-33) */
  OPY_FN_INDICATOR = 6,      /* Indicator / Constrain infeas (reference indicators.py:17-46)        */
  OPY_FN_MASKED_AFFINE = 7,  /* MaskedAffineFct             (reference functionals/affine.py:60-72) */
  OPY_FN_POS = 8,            /* PositivityFct infeas        (reference indicators.py:57-63)         */
  OPY_FN_NEG = 9,            /* NegativityFct infeas        (reference indicators.py:77-83)         */
  OPY_FN_QUADEPI_IND = 10,   /* QuadEpiInd infeas           (reference epigraph.py:187-197)         */
  OPY_FN_QUADEPI_SUPP = 11,  /* QuadEpiSupp value/infeas    (reference epigraph.py:154-166)         */
  OPY_FN_HUBER = 12,         /* HuberPerspective            (reference epigraph.py:224-236)         */
  OPY_FN_TRUNCQUADEPI = 13,  /* TruncQuadEpiInd (2 infeas)  (reference epigraph.py:256-270)         */
  OPY_FN_EPIGRAPH_IND = 14,  /* EpigraphInd infeas          (reference epigraph.py:42-47)           */
  OPY_FN_EPIGRAPH_SUPP = 15, /* EpigraphSupp value/infeas   (reference epigraph.py:115-126)         */
  /* no reference counterpart: the box and simplex constraints of the multi-label lifting model */
  OPY_FN_BOX_IND = 16,       /* infeas max(0, lo - x, x - hi), lo = cs, hi = lbd                     */
  OPY_FN_BOX_SUPP = 17,      /* its conjugate: sum max(lo z, hi z)                                   */
  OPY_FN_SIMPLEX_IND = 18,   /* per point of width K: infeas max(|sum x - 1|, max(0, -x))            */
  OPY_FN_SIMPLEX_SUPP = 19   /* its conjugate: sum over points of max_k z_k                          */
};

typedef struct {
  const void* f;         /* data term (element type of the call), per point x width */
  const double* vol;     /* per-channel weights (SSD volume / L1 lambda), width entries */
  const uint8_t* mask;   /* per-point mask or NULL */
  const void* c;         /* per-entry constant (element type of the call) or NULL */
  double cs;             /* scalar constant when c is NULL */
  double lbd;            /* ball radius / Huber lambda */
  int32_t m1, m2;        /* per-point matrix shape for OPY_FN_NORMS */
  int32_t matrixnorm;    /* 'F' frobenius, 'N' nuclear (value) + spectral (infeas) */
  int32_t pad;
  double qa;             /* quadratic coefficient a (QuadEpi) or alph (Huber / TruncQuadEpi) */
  const double* qb;      /* linear coefficients b (N x M) or NULL */
  const double* qc;      /* constants c (N) or NULL */
  /* polyhedral epigraphs, points numbered t = j*nfuns + i, width ndim+1 */
  const uint8_t* ei;     /* I (nfuns x npoints) */
  const int32_t* ej;     /* J (nregions x nsub) */
  const double* ev;      /* v (npoints x ndim) */
  const double* eb;      /* b (nfuns x npoints) */
  const double* eq;      /* affine pieces (sum of nfaces) x (ndim+1), EpigraphSupp */
  const int64_t* eq_off; /* piece offsets per point, nregions*nfuns+1 entries */
  const double* checks;  /* nregions x (ndim+1) x (ndim+1) feasibility rows */
  int32_t nfuns, npoints, nsub, epad;
} opy_fn_params_t;

/* ---- library ---------------------------------------------------------- */
const char* opy_last_error(void);
int opy_abi_version(void);
int opy_device_sync(void);

/* ---- finite differences (reference opymize/linear/diff.py, diff.cu) ----- */
/* y (N,D,C) (+)= grad x with x (N,C); GradientOp._call_gpu, diff.py:141-144, diff.cu:37-75 */
int opy_grad_f32(const float* x, float* y, const opy_grid_t* g, const double* w, int scheme, int add, void* stream);
int opy_grad_f64(const double* x, double* y, const opy_grid_t* g, const double* w, int scheme, int add, void* stream);
/* x (N,C) (+)= -div p with p (N,D,C); DivergenceOp._call_gpu, diff.py:180-183, diff.cu:77-127 */
int opy_div_f32(const float* p, float* x, const opy_grid_t* g, const double* w, int scheme, int add, void* stream);
int opy_div_f64(const double* p, double* x, const opy_grid_t* g, const double* w, int scheme, int add, void* stream);
/* row-wise l^p norms of grad (adjoint=0) or div (adjoint=1); LinOp.rowwise_lp, __init__.py:122-142 */
int opy_diff_rowwise_f32(float* out, const opy_grid_t* g, const double* w, int scheme, int adjoint, double p,
                         int add, void* stream);
int opy_diff_rowwise_f64(double* out, const opy_grid_t* g, const double* w, int scheme, int adjoint, double p,
                         int add, void* stream);
/* y (+)= Laplacian x, bc in {'N','C','S'}; LaplacianOp._call_gpu, diff.py:246-249, diff.cu:131-243 */
int opy_laplacian_f32(const float* x, float* y, const opy_grid_t* g, int bc, int adj, int add, void* stream);
int opy_laplacian_f64(const double* x, double* y, const opy_grid_t* g, int bc, int adj, int add, void* stream);
int opy_laplacian_rowwise_f32(float* y, const opy_grid_t* g, int bc, int adj, double p, int add, void* stream);
int opy_laplacian_rowwise_f64(double* y, const opy_grid_t* g, int bc, int adj, double p, int add, void* stream);

/* ---- dense / sparse products (reference opymize/linear/einsum.py, einsum.cu) ---- */
/* MatrixMult: y[j,i] (+)= sum_k A(j,k) x[k,i]; einsum.py:17-56, einsum.cu:3-28 */
int opy_matmul_f32(const float* x, float* y, const float* A, int64_t N, int J, int K, int trans, int add, void* stream);
int opy_matmul_f64(const double* x, double* y, const double* A, int64_t N, int J, int K, int trans, int add, void* stream);
/* MatrixMultR: y[i,j] (+)= sum_k x[i,k] A(k,j); einsum.py:58-97, einsum.cu:32-57 */
int opy_matmul_r_f32(const float* x, float* y, const float* A, int64_t N, int J, int K, int trans, int add, void* stream);
int opy_matmul_r_f64(const double* x, double* y, const double* A, int64_t N, int J, int K, int trans, int add, void* stream);
/* DiagMatrixMultR: y[i,k] (+)= x[i,k] A[k]; einsum.py:104-128 */
int opy_diag_matmul_r_f32(const float* x, float* y, int64_t N, int K, const float* A, int add, void* stream);
int opy_diag_matmul_r_f64(const double* x, double* y, int64_t N, int K, const double* A, int add, void* stream);
/* lifting operators.  IndexedMult (adj=0): y[j,i,m] (+)= -sum_l B[j,m,l] x[i,P[j,l]], x (N,K) -> y (J,N,M);
 * IndexedMultAdj (adj=1): y[i,P[j,l]] (+)= -sum_m B[j,m,l] x[j,i,m], x (J,N,M) -> y (N,K).  B (J,M,L) in the
 * element type, P (J,L) int32 in [0,K) (entries outside are skipped).
 * IndexedMult / IndexedMultAdj, indexed.py:34-94, indexed.cu:1-48 */
int opy_indexedmult_f32(const float* x, float* y, const float* B, const int32_t* P, int64_t N, int J, int M, int L,
                        int K, int adj, int add, void* stream);
int opy_indexedmult_f64(const double* x, double* y, const double* B, const int32_t* P, int64_t N, int J, int M, int L,
                        int K, int adj, int add, void* stream);
/* TangledMatrixMultR: y[m,i,k] (+)= sum_jl x[j,i,l] A(j,l,m,k), x (J,N,L) -> y (M,N,K); A is J x L x M x K, or
 * M x K x J x L with trans=1 (the adjoint); einsum.py:130-173, einsum.cu:90-123 */
int opy_tangled_matmul_r_f32(const float* x, float* y, const float* A, int64_t N, int J, int L, int M, int K, int trans,
                             int add, void* stream);
int opy_tangled_matmul_r_f64(const double* x, double* y, const double* A, int64_t N, int J, int L, int M, int K,
                             int trans, int add, void* stream);
/* MatrixMultRBatched: y[j,i,k] (+)= sum_m x[j,i,m] A(j,m,k), x (J,N,M) -> y (J,N,K); A is J x M x K, or J x K x M
 * with trans=1 (the adjoint); einsum.py:175-214, einsum.cu:60-89 */
int opy_matmul_r_batched_f32(const float* x, float* y, const float* A, int64_t N, int J, int M, int K, int trans,
                             int add, void* stream);
int opy_matmul_r_batched_f64(const double* x, double* y, const double* A, int64_t N, int J, int M, int K, int trans,
                             int add, void* stream);
/* CSR product for LinOps that only have a matrix (LinOp._call_cpu spmat.dot, __init__.py:95-108) */
int opy_csr_spmv_f32(int64_t nrows, const int64_t* indptr, const int64_t* indices, const float* data, const float* x,
                     float* y, int add, void* stream);
int opy_csr_spmv_f64(int64_t nrows, const int64_t* indptr, const int64_t* indices, const double* data,
                     const double* x, double* y, int add, void* stream);
int opy_csr_rowwise_f32(int64_t nrows, const int64_t* indptr, const float* data, double p, float* y, int add,
                        void* stream);
int opy_csr_rowwise_f64(int64_t nrows, const int64_t* indptr, const double* data, double p, double* y, int add,
                        void* stream);

/* ---- trivial operators (reference opymize/linear/scale.py) ------------- */
int opy_fill_f32(float* y, int64_t n, double value, void* stream);
int opy_fill_f64(double* y, int64_t n, double value, void* stream);
int opy_copy_f32(const float* x, float* y, int64_t n, void* stream);
int opy_copy_f64(const double* x, double* y, int64_t n, void* stream);
/* y[i] (+)= v[(i/div) % nv]: broadcast of per-row constants (rowwise norms of dense products) */
int opy_tile_f32(float* y, int64_t n, const float* v, int64_t nv, int64_t div, int add, void* stream);
int opy_tile_f64(double* y, int64_t n, const double* v, int64_t nv, int64_t div, int add, void* stream);
/* y = |x| (normest start vector, reference linear/__init__.py:23) */
int opy_abs_f32(const float* x, float* y, int64_t n, void* stream);
int opy_abs_f64(const double* x, double* y, int64_t n, void* stream);
/* x = 1/x where x > thr (diagonal preconditioners, reference pdhg.py:197-200) */
int opy_recip_gt_f32(float* x, int64_t n, double thr, void* stream);
int opy_recip_gt_f64(double* x, int64_t n, double thr, void* stream);
/* ScaleOp / IdentityOp: y (+)= fact*x, fact array or scalar; scale.py:34-79 */
int opy_scale_f32(const float* x, float* y, int64_t n, const float* fact, double fact_s, int add, void* stream);
int opy_scale_f64(const double* x, double* y, int64_t n, const double* fact, double fact_s, int add, void* stream);
/* NihilOp: y = keep ? x : 0; scale.py:96-115 */
int opy_nihil_f32(const float* x, float* y, int64_t n, const uint8_t* keep, int add, void* stream);
int opy_nihil_f64(const double* x, double* y, int64_t n, const uint8_t* keep, int add, void* stream);

/* ---- prox maps (reference opymize/operators) ---------------------------- */
/* ShiftScaleOp: y (+)= a*(x + b*shift), each of a, b, shift array or scalar; affine.py:85-146 */
int opy_shiftscale_f32(const float* x, float* y, int64_t n, const float* shift, double shift_s, const float* a,
                       double a_s, const float* b, double b_s, int add, void* stream);
int opy_shiftscale_f64(const double* x, double* y, int64_t n, const double* shift, double shift_s, const double* a,
                       double a_s, const double* b, double b_s, int add, void* stream);
/* SSD.prox(tau): y = (x + t vol f)/(1 + t vol) on masked points; quad.py:45-55 (+ affine.py:120-123).
 * The step comes from the device so adaptive steps need no host round trip. */
int opy_ssd_prox_f32(const float* x, float* y, int64_t N, int C, const float* f, const double* vol,
                     const uint8_t* mask, const opy_step_t* step, int add, void* stream);
int opy_ssd_prox_f64(const double* x, double* y, int64_t N, int C, const double* f, const double* vol,
                     const uint8_t* mask, const opy_step_t* step, int add, void* stream);
/* TV-L1 data term prox (conj=0: soft threshold about f) or its conjugate's (conj=1: clip(x - t f)).
 * No reference counterpart: the reference has no L1 data functional. */
int opy_l1data_prox_f32(const float* x, float* y, int64_t N, int C, const float* f, const double* lbd, int conj,
                        const opy_step_t* step, int add, void* stream);
int opy_l1data_prox_f64(const double* x, double* y, int64_t N, int C, const double* f, const double* lbd, int conj,
                        const opy_step_t* step, int add, void* stream);
/* PosProj / NegProj: proj.py:19-85 */
int opy_signproj_f32(const float* x, float* y, int64_t n, int positive, int add, void* stream);
int opy_signproj_f64(const double* x, double* y, int64_t n, int positive, int add, void* stream);
/* ConstOp: affine.py:15-46 */
int opy_const_f32(float* y, int64_t n, const float* c, int add, void* stream);
int opy_const_f64(double* y, int64_t n, const double* c, int add, void* stream);
/* ConstrainOp: y = mask[i/M] ? c : x; affine.py:48-83 */
int opy_constrain_f32(const float* x, float* y, int64_t n, int64_t M, const float* c, const uint8_t* mask, int add,
                      void* stream);
int opy_constrain_f64(const double* x, double* y, int64_t n, int64_t M, const double* c, const uint8_t* mask,
                      int add, void* stream);
/* L1NormsProj (in place, per-point M1 x M2 blocks stored column-major): proj.py:170-226, proj.cu:3-107 */
int opy_l1norms_proj_f32(float* x, int64_t N, int M1, int M2, double lbd, int spectral, void* stream);
int opy_l1norms_proj_f64(double* x, int64_t N, int M1, int M2, double lbd, int spectral, void* stream);
/* QuadEpiProj (in place; shift may be NULL; lbd = +inf for no truncation): proj.py:278-374, proj.cu:110-219 */
int opy_quadepi_proj_f32(float* x, int64_t N, int M, double alph, const float* shift, double lbd, void* stream);
int opy_quadepi_proj_f64(double* x, int64_t N, int M, double alph, const double* shift, double lbd, void* stream);
/* x (nregions, nfuns, ndim+1) <- projection onto the epigraphs of the pieces
 * f[i]_j* (active-set QP, at most ndim+1 active constraints, ndim <= 3);
 * I uint8 (nfuns, npoints), J int32 (nregions, nsub), v (npoints, ndim),
 * b (nfuns, npoints); EpigraphProj._call_gpu, proj.py:388-470, proj.cu:222-831 */
int opy_epigraph_proj_f32(float* x, const uint8_t* I, const int32_t* J, const float* v, const float* b, int ndim,
                          int nfuns, int nregions, int npoints, int nsub, int maxiter, double tol, void* stream);
int opy_epigraph_proj_f64(double* x, const uint8_t* I, const int32_t* J, const double* v, const double* b, int ndim,
                          int nfuns, int nregions, int npoints, int nsub, int maxiter, double tol, void* stream);

/* ---- lifting-model projections and block preconditioners (no reference counterpart) ---------
 * The multi-label lifting model keeps its lifted variable in a box and its per-point piece weights
 * in the probability simplex; the reference has neither functional. */
/* y = min(max(x, lo), hi) (x == y allowed) */
int opy_clip_f32(const float* x, float* y, int64_t n, double lo, double hi, void* stream);
int opy_clip_f64(const double* x, double* y, int64_t n, double lo, double hi, void* stream);
/* in place: each of the N points' K <= 32 values onto {v >= 0, sum v = 1} (Michelot's algorithm) */
int opy_simplex_proj_f32(float* x, int64_t N, int K, void* stream);
int opy_simplex_proj_f64(double* x, int64_t N, int K, void* stream);
/* in place: each block of `width` step sizes takes its smallest positive entry (0 if none).  The
 * diagonal preconditioner of pdhg.py:189-200 gives per-element steps; a prox that is a projection onto
 * a non-separable set (an epigraph, a simplex) is only the exact prox for a step that is constant on
 * the set's block, and the smaller step keeps the convergence condition. */
int opy_block_min_f32(float* x, int64_t nblocks, int width, void* stream);
int opy_block_min_f64(double* x, int64_t nblocks, int width, void* stream);

/* ---- PDHG step kernels (reference opymize/solvers/pdhg.py) ------------- */
/* take_step: zn = z + dir*step*g; pdhg.py:62-65 */
int opy_pdhg_take_step_f32(float* zn, const float* z, int64_t n, int dir, const opy_step_t* step, const float* g,
                           void* stream);
int opy_pdhg_take_step_f64(double* zn, const double* z, int64_t n, int dir, const opy_step_t* step, const double* g,
                           void* stream);
/* overrelax + dual take_step fused: yn = y + step*((1+theta) axn - theta ax); pdhg.py:66-69, :148-151 */
int opy_pdhg_dual_step_f32(float* yn, const float* y, int64_t n, double theta, const opy_step_t* step,
                           const float* axn, const float* ax, void* stream);
int opy_pdhg_dual_step_f64(double* yn, const double* y, int64_t n, double theta, const opy_step_t* step,
                           const double* axn, const double* ax, void* stream);
/* residual: out2 = [sum((z-zn)/a - b g + c1 gn)^2, sum((z-zn)/a - b g + c2 gn)^2] in FP64;
 * the two reductions of pdhg.py:74-78 per side, in one pass (work: 2*OPY_RED_BLOCKS doubles) */
int opy_pdhg_residual2_f32(const opy_step_t* a, const float* z, const float* zn, double b, const float* g, double c1,
                           double c2, const float* gn, int64_t n, double* work, double* out2, void* stream);
int opy_pdhg_residual2_f64(const opy_step_t* a, const double* z, const double* zn, double b, const double* g,
                           double c1, double c2, const double* gn, int64_t n, double* work, double* out2,
                           void* stream);
/* The same residual sums as opy_pdhg_residual2_f32 with no work argument: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under hipGraph capture, serving
 * loops).  The 2*OPY_RED_BLOCKS partials (16 KiB) sit in a device array of the library, one per
 * GPU, held until the process ends; every n fits, since the reduction grid is clamped to
 * OPY_RED_BLOCKS blocks.  Calls that overlap on one GPU race on that array.  Allocates nothing. */
int opy_pdhg_residual2_noalloc_f32(const opy_step_t* a, const float* z, const float* zn, double b, const float* g,
                                   double c1, double c2, const float* gn, int64_t n, double* out2, void* stream);
/* adaptive step rule on the device; pdhg.py:166-183 */
int opy_pdhg_adapt(opy_adapt_state_t* state, const opy_adapt_params_t* params, void* stream);

/* ---- FP64 functional evaluation for the gap (reference pdhg.py:23-44) --- */
/* out3 = [sum of values, max infeasibility, second max infeasibility]; accumulate=1 adds/maxes into out3 */
int opy_fn_eval_f32(int kind, const float* x, int64_t npts, int width, const opy_fn_params_t* prm, double* work,
                    double* out3, int accumulate, void* stream);
int opy_fn_eval_f64(int kind, const double* x, int64_t npts, int width, const opy_fn_params_t* prm, double* work,
                    double* out3, int accumulate, void* stream);

/* ---- semismooth Newton continuation (reference opymize/solvers/semismooth_newton.py) ---- */
/* out[0] = sum_i x_i y_i in FP64 (the np.einsum('i,i->') products of semismooth_newton.py:107-114,
 * :291, :314, :322; work: OPY_RED_BLOCKS*3 doubles) */
int opy_dot_f32(const float* x, const float* y, int64_t n, double* work, double* out, void* stream);
int opy_dot_f64(const double* x, const double* y, int64_t n, double* work, double* out, void* stream);
/* z = a*x + b*y, z may be x or y (the vector updates of CG_solve, semismooth_newton.py:102-112, of the
 * Newton system, :211-219, :240-248, of the residual, :276-284, and of armijo, :39) */
int opy_axpby_f32(const float* x, const float* y, float* z, int64_t n, double a, double b, void* stream);
int opy_axpby_f64(const double* x, const double* y, double* z, int64_t n, double a, double b, void* stream);
/* Jacobian of PosProj (positive=1) / NegProj at xbar applied to x: y (+)= (xbar > 0 / < 0) ? x : 0;
 * proj.py:25-50, :58-85 (NihilOp with keep from the prox input) */
int opy_signproj_jac_f32(const float* xbar, const float* x, float* y, int64_t n, int positive, int add,
                         void* stream);
int opy_signproj_jac_f64(const double* xbar, const double* x, double* y, int64_t n, int positive, int add,
                         void* stream);
/* Jacobian of the Frobenius L1NormsProj at xbar (N points of M entries) applied to x: y (+)= x inside the
 * ball, (lbd/|xbar|)(x - <u,x> u) with u = xbar/|xbar| outside; L12ProjJacobian, proj.py:228-275 */
int opy_l12proj_jac_f32(const float* xbar, const float* x, float* y, int64_t N, int M, double lbd, int add,
                        void* stream);
int opy_l12proj_jac_f64(const double* xbar, const double* x, double* y, int64_t N, int M, double lbd, int add,
                        void* stream);

/* ---- fused, batched TV iterations ---------------------------------------
 * PDHG for G + tvw*|grad x|_{2,1} with the forward-difference gradient on 2-D
 * images of C <= 4 channels, many instances of uneven size per launch.  These
 * entry points replace the chain of generic launches of one PDHG iteration
 * (take_step, gprox, linop, overrelax, take_step, fconjprox, linop.adjoint,
 * 4 residual reductions; reference pdhg.py:124-187) for that model. */
#define OPY_TV_DATA_L2 0      /* G = 0.5*lam*|x - f|^2  (SSD, reference functionals/quad.py) */
#define OPY_TV_DATA_L1 1      /* G = lam*|x - f|_1      (TV-L1 data term, new) */
#define OPY_TV_STEPS_SCALAR 0 /* tau, sigma from the instance's opy_adapt_state_t (constant or adaptive) */
#define OPY_TV_STEPS_PRECOND 1/* diagonal preconditioners of the stencil (reference pdhg.py:189-200) */
/* accelerated steps for the lam-strongly convex SSD term (Chambolle & Pock 2011,
 * Algorithm 2, no reference counterpart): per instance theta = 1/sqrt(1 + 2 lam tau)
 * (1 for TV-L1), xbar = x+ + theta (x+ - x), dual step sigma/theta; opy_tv_accel
 * then sets tau <- theta tau, sigma <- sigma/theta */
#define OPY_TV_STEPS_ACCEL 2

typedef struct {
  int64_t off;    /* element offset of the instance's x and f (pixels*C); y starts at 2*off */
  int32_t H, W;
  int32_t data;   /* OPY_TV_DATA_* */
  int32_t tile0;  /* first per-tile partial slot of the instance */
  int32_t ntiles; /* ceil(H/th)*ceil(W/tw) */
  int32_t pad;
  double lam;     /* data weight */
  double tvw;     /* TV weight = radius of the dual ball */
} opy_tv_inst_t;

typedef struct {
  int32_t C;      /* channels, 1..4 */
  int32_t tw;     /* threads per tile row; divides 256 */
  int32_t rpt;    /* rows per thread; tile height th = 256/tw*rpt */
  int32_t steps;  /* OPY_TV_STEPS_* */
  double h0, h1;  /* grid spacing */
  double w[4];    /* channel weights of the gradient */
} opy_tv_cfg_t;

typedef struct {
  double objp, objd, gap, relgap, infeasp, infeasd;
  double objd_cert, relgap_cert; /* dual value of the rescaled feasible dual point (L1) */
  int32_t status;                /* 0 running, 1 converged, 2 non-finite */
  int32_t iters;                 /* iterations run when last checked */
} opy_tv_check_t;

/* The fused kernels walk the tiles of the running instances with a grid
 * stride: `grid` > 0 launches that many 256-thread blocks, 0 launches the
 * number of blocks resident at once on the device for that kernel.  One block
 * per running tile (act_off[nact], read by the caller) streams fastest; blocks
 * beyond the running tiles return at once. */
/* pass 1: x_out = prox_{tau G}(x_in - tau A^T y) on the tiles of the running instances */
int opy_tv_primal_f32(const float* x_in, float* x_out, const float* y, const float* f, const opy_tv_inst_t* insts,
                      const opy_adapt_state_t* steps, const int32_t* active_inst, const int32_t* act_off,
                      const int32_t* nact, const opy_tv_cfg_t* cfg, int grid, void* stream);
int opy_tv_primal_f64(const double* x_in, double* x_out, const double* y, const double* f, const opy_tv_inst_t* insts,
                      const opy_adapt_state_t* steps, const int32_t* active_inst, const int32_t* act_off,
                      const int32_t* nact, const opy_tv_cfg_t* cfg, int grid, void* stream);
/* pass 2: y_out = proj(y_in + sigma A(2 x_new - x_old)); adapt=1 also writes 4 residual partials per tile */
int opy_tv_dual_f32(const float* x_old, const float* x_new, const float* y_in, float* y_out,
                    const opy_tv_inst_t* insts, const opy_adapt_state_t* steps, const int32_t* active_inst,
                    const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int adapt,
                    double* partials, int grid, void* stream);
int opy_tv_dual_f64(const double* x_old, const double* x_new, const double* y_in, double* y_out,
                    const opy_tv_inst_t* insts, const opy_adapt_state_t* steps, const int32_t* active_inst,
                    const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int adapt,
                    double* partials, int grid, void* stream);
/* one whole iteration in one pass (cfg->tw = 64, cfg->rpt <= 16):
 * x_out = prox_{tau G}(x_in - tau A^T y_in), y_out = proj(y_in + sigma A(2 x_out - x_in));
 * reads x, f, y and writes x+, y+ (PDHG.iteration_step, pdhg.py:124-153, for TV models);
 * adapt=1 also writes the 4 residual partials per tile (pdhg.py:155-164) */
int opy_tv_iter_f32(const float* x_in, float* x_out, const float* y_in, float* y_out, const float* f,
                    const opy_tv_inst_t* insts, const opy_adapt_state_t* steps, const int32_t* active_inst,
                    const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int adapt,
                    double* partials, int grid, void* stream);
int opy_tv_iter_f64(const double* x_in, double* x_out, const double* y_in, double* y_out, const double* f,
                    const opy_tv_inst_t* insts, const opy_adapt_state_t* steps, const int32_t* active_inst,
                    const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int adapt,
                    double* partials, int grid, void* stream);
/* k in {2,3,4} whole iterations in one pass over HBM (cfg->tw = 64, cfg->rpt in {2,4,6,8}; no
 * adaptive steps, whose rule needs every iteration's residuals): the same iterates as k
 * opy_tv_iter launches with steps SCALAR / PRECOND / ACCEL (PDHG.iteration_step, pdhg.py:124-153,
 * k times).  Each block loads its tile and k pixels around it into LDS and iterates there.  The
 * step state is read, not advanced: with OPY_TV_STEPS_ACCEL call opy_tv_accel_n with count k after.
 * Fails with an error when the region of a tile (3*C values per pixel) exceeds 64 KiB of LDS. */
int opy_tv_multi_f32(const float* x_in, float* x_out, const float* y_in, float* y_out, const float* f,
                     const opy_tv_inst_t* insts, const opy_adapt_state_t* steps, const int32_t* active_inst,
                     const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int k, int grid,
                     void* stream);
int opy_tv_multi_f64(const double* x_in, double* x_out, const double* y_in, double* y_out, const double* f,
                     const opy_tv_inst_t* insts, const opy_adapt_state_t* steps, const int32_t* active_inst,
                     const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int k, int grid,
                     void* stream);
/* k in 2..8 iterations (k > 4 for C <= 2) in one pass over HBM with steps SCALAR or ACCEL, the
 * iterates of opy_tv_multi / k single launches bit for bit, by a different work split: one wavefront
 * per unit of 64-2k image columns by `rows` image rows, walking down the unit and k rows above and
 * below it with the k iterations pipelined one row apart (no LDS).  unit_off (B+1 entries) is the
 * running instances' exclusive prefix of units, from opy_tv_stream_units with the same k and rows;
 * `grid` as above (one block of 4 wavefronts per 4 units streams fastest).  PRECOND steps are
 * refused (use opy_tv_multi).  Pixels are addressed by 32-bit byte offsets from each instance's
 * base, so every instance's y block (H*W*2C values) must be below 4 GiB.
 * No reference counterpart beyond the iteration itself (PDHG.iteration_step, pdhg.py:124-153). */
int opy_tv_stream_f32(const float* x_in, float* x_out, const float* y_in, float* y_out, const float* f,
                      const opy_tv_inst_t* insts, const opy_adapt_state_t* steps, const int32_t* active_inst,
                      const int32_t* unit_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int k, int rows,
                      int grid, void* stream);
int opy_tv_stream_f64(const double* x_in, double* x_out, const double* y_in, double* y_out, const double* f,
                      const opy_tv_inst_t* insts, const opy_adapt_state_t* steps, const int32_t* active_inst,
                      const int32_t* unit_off, const int32_t* nact, const opy_tv_cfg_t* cfg, int k, int rows,
                      int grid, void* stream);
/* unit_off[a] = sum over the running instances a' < a (active_inst, nact as from opy_tv_compact) of
 * strips(W) * ceil(H / rows); unit_off[nact] = the total.  One block.  strips(W) = ceil(W / sw),
 * sw = 64 - 2k, or for k >= 6, when that is fewer, the strips of the layout whose first and last
 * strips carry no halo beyond the image edge: 1 for W <= 64, else ceil((W - 64) / sw) + 1. */
int opy_tv_stream_units(const opy_tv_inst_t* insts, const int32_t* active_inst, const int32_t* nact,
                        int32_t* unit_off, int k, int rows, void* stream);
/* adaptive step rule per running instance from the residual partials (reference pdhg.py:155-183) */
int opy_tv_adapt(const opy_tv_inst_t* insts, opy_adapt_state_t* steps, const opy_tv_check_t* checks,
                 const double* partials, int B, int C, const opy_adapt_params_t* params, void* stream);
/* accelerated step update of the running instances after an iteration (OPY_TV_STEPS_ACCEL) */
int opy_tv_accel(const opy_tv_inst_t* insts, opy_adapt_state_t* steps, const opy_tv_check_t* checks, int B,
                 void* stream);
/* the same for `count` consecutive iterations in one launch (after opy_tv_multi_* with k = count) */
int opy_tv_accel_n(const opy_tv_inst_t* insts, opy_adapt_state_t* steps, const opy_tv_check_t* checks, int B,
                   int count, void* stream);
/* Restarts of the accelerated schedule (no reference counterpart; O'Donoghue & Candes 2015 restart
 * accelerated methods, Applegate et al. 2021 restart PDHG).  After opy_tv_check: every running
 * TV-L2 instance whose certified relative gap g = |relgap_cert| has fallen below beta times its
 * value at the last restart, or has risen above its smallest value since that restart, with at
 * least min_iters iterations run since, returns to the initial steps tau0, sigma0 (x and y stay).
 * One record per instance; initialise ref = best = +inf, last = 0 (the iteration count of the
 * start). */
typedef struct {
  double ref;   /* g at the last restart (+inf: none recorded yet) */
  double best;  /* smallest g since the last restart (+inf: none yet) */
  int32_t last; /* iterations run at the last restart */
  int32_t count;/* restarts so far */
} opy_tv_restart_t;
int opy_tv_restart(const opy_tv_inst_t* insts, opy_adapt_state_t* steps, const opy_tv_check_t* checks,
                   opy_tv_restart_t* records, int B, double beta, int min_iters, double tau0, double sigma0,
                   void* stream);
/* FP64 gap partials per tile (reference pdhg.py:23-44) */
int opy_tv_gap_f32(const float* x, const float* y, const float* f, const opy_tv_inst_t* insts,
                   const int32_t* active_inst, const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg,
                   double* partials, int grid, void* stream);
int opy_tv_gap_f64(const double* x, const double* y, const double* f, const opy_tv_inst_t* insts,
                   const int32_t* active_inst, const int32_t* act_off, const int32_t* nact, const opy_tv_cfg_t* cfg,
                   double* partials, int grid, void* stream);
/* per-instance gap, relative gap and termination test (reference pdhg.py:299-305) */
int opy_tv_check(const opy_tv_inst_t* insts, opy_tv_check_t* checks, const double* partials, int B, double tol_gap,
                 double tol_infeas, int certify, const int64_t* iter_counter, int iters_in_block, void* stream);
/* cold restart of the instances with flags[b] != 0: x = y = 0 in the given buffers, status running.
 * No reference counterpart (the fallback of the learned warm start). */
int opy_tv_reset_f32(const opy_tv_inst_t* insts, const int32_t* flags, int B, int C, float* x, float* y,
                     opy_tv_check_t* checks, void* stream);
int opy_tv_reset_f64(const opy_tv_inst_t* insts, const int32_t* flags, int B, int C, double* x, double* y,
                     opy_tv_check_t* checks, void* stream);
/* rebuild the running-instance list (active_inst, act_off: B+1 entries) and advance the iteration counter */
int opy_tv_compact(const opy_tv_inst_t* insts, const opy_tv_check_t* checks, int B, int32_t* active_inst,
                   int32_t* act_off, int32_t* nact, int64_t* iter_counter, int iters_in_block, void* stream);

/* ---- fused, batched multi-label lifting --------------------------------
 * PDHG with preconditioned steps on the sublabel-accurate TV lifting of many images of uneven size
 * and label count (opymize_amd/lifting.py): the generic chain of PDHG.solve(steps="precond") on that
 * model (reference pdhg.py:124-153 over GradientOp, MatrixMultR, L1NormsProj, QuadEpiProj and the
 * box / simplex projections) as two passes per iteration.  Instance b keeps its k = L-1 pieces as
 * planes of H*W values: U and r at offk + i*H*W, p and z at 2*offk + (2i + c)*H*W (c = 0 the row
 * difference / nu, c = 1 the column difference / s), f at off.  Tiles are 64 x 4 pixels. */
typedef struct {
  int64_t off;    /* pixel offset of the instance's f */
  int64_t offk;   /* plane offset of U and r (p and z at twice it) */
  int32_t H, W;
  int32_t K;      /* pieces (labels - 1), 1..32 */
  int32_t tile0;  /* first per-tile partial slot */
  int32_t ntiles; /* ceil(H/4) * ceil(W/64) */
  int32_t pad;
  double lam;     /* TV weight */
  double g0, h;   /* first label, label spacing */
  double tau_r;   /* primal step of the piece weights: min over l of 1/(|g_l| + h l + 1) */
  double omega;   /* primal weight: primal steps times omega, dual steps over omega (1: the reference's) */
} opy_lift_inst_t;
/* primal pass, in place: U <- clip(U - tau K^T y), r <- simplex(r - tau K^T y); Ub, rb <- 2 x+ - x.
 * kmax = the largest K of the batch; `grid` blocks walk the running tiles. */
int opy_lift_primal_f32(float* U, float* r, float* Ub, float* rb, const float* p, const float* z,
                        const opy_lift_inst_t* insts, const int32_t* active_inst, const int32_t* act_off,
                        const int32_t* nact, int kmax, int grid, void* stream);
int opy_lift_primal_f64(double* U, double* r, double* Ub, double* rb, const double* p, const double* z,
                        const opy_lift_inst_t* insts, const int32_t* active_inst, const int32_t* act_off,
                        const int32_t* nact, int kmax, int grid, void* stream);
/* dual pass, in place: p <- ball projection, z <- epigraph projection of y + sigma K x_bar */
int opy_lift_dual_f32(const float* Ub, const float* rb, float* p, float* z, const float* f,
                      const opy_lift_inst_t* insts, const int32_t* active_inst, const int32_t* act_off,
                      const int32_t* nact, int kmax, int grid, void* stream);
int opy_lift_dual_f64(const double* Ub, const double* rb, double* p, double* z, const double* f,
                      const opy_lift_inst_t* insts, const int32_t* active_inst, const int32_t* act_off,
                      const int32_t* nact, int kmax, int grid, void* stream);
/* FP64 gap partials, 8 per tile (reference pdhg.py:23-44 for this model) */
int opy_lift_gap_f32(const float* U, const float* r, const float* p, const float* z, const float* f,
                     const opy_lift_inst_t* insts, const int32_t* active_inst, const int32_t* act_off,
                     const int32_t* nact, int kmax, double* partials, int grid, void* stream);
int opy_lift_gap_f64(const double* U, const double* r, const double* p, const double* z, const double* f,
                     const opy_lift_inst_t* insts, const int32_t* active_inst, const int32_t* act_off,
                     const int32_t* nact, int kmax, double* partials, int grid, void* stream);
/* per-instance objectives, gap and termination (reference pdhg.py:299-305); B blocks */
int opy_lift_check(const opy_lift_inst_t* insts, opy_tv_check_t* checks, const double* partials, int B,
                   double tol_gap, double tol_infeas, const int64_t* iter_counter, int iters_in_block, void* stream);
/* running-instance list and iteration counter, as opy_tv_compact */
int opy_lift_compact(const opy_lift_inst_t* insts, const opy_tv_check_t* checks, int B, int32_t* active_inst,
                     int32_t* act_off, int32_t* nact, int64_t* iter_counter, int iters_in_block, void* stream);

/* ---- warm-start network: 3x3 convolutions, 32 -> 32 channels, NHWC float32 ----
 * No counterpart in the reference (its solvers always start from zero or a
 * user-supplied continue_at, pdhg.py:236-241); these kernels train and run the
 * network that predicts continue_at for the batched solver.  Matrix-core
 * BF16 operands with FP32 accumulation. */
/* y = act(conv(x, W) + b); wt from opy_conv3x3_prep_weights (9*32*32 bf16); act 0 none, 1 ReLU */
int opy_conv3x3_fwd_bf16(const float* x, float* y, const void* wt, const float* bias, int N, int H, int W, int act,
                         int grid, void* stream);
/* the same with the activations' element type chosen per side: x_bf16 / y_bf16 = 1 for BF16 NHWC (the
 * operands are rounded to BF16 for the matrix cores in any case, so BF16 activations between layers give the
 * same results at half the traffic), 0 for float32 */
int opy_conv3x3_fwd_io(const void* x, void* y, const void* wt, const float* bias, int N, int H, int W, int act,
                       int x_bf16, int y_bf16, int grid, void* stream);
/* per-block partials (grid x (9*32*32 + 32) floats) of dW[tap][co][ci] and dbias[co] */
/* The warm start's first layer straight from packed images (no counterpart): input channel 0 =
 * f - 0.5, channel 1 = loglam[n], channels 2..31 = 0 for image n at f + n*stride (H*W floats);
 * output NHWC BF16, 32 channels, ReLU. */
int opy_conv3x3_first(const float* f, int64_t stride, const float* loglam, void* y, const void* wt,
                      const float* bias, int N, int H, int W, int grid, void* stream);
/* The last layer into the packed solver start (no counterpart): with o = conv(x) + b (BF16 NHWC
 * input, no activation), x_out = ((f - 0.5) + 0.5) + o0 and y_out = (o1, o2) per pixel, image n at
 * f / x_out + n*stride and y_out + 2n*stride; x 16-byte and y_out 8-byte aligned. */
int opy_conv3x3_last(const void* x, const float* f, int64_t stride, float* x_out, float* y_out, const void* wt,
                     const float* bias, int N, int H, int W, int grid, void* stream);
int opy_conv3x3_wgrad_bf16(const float* x, const float* g, float* partials, int N, int H, int W, int grid,
                           void* stream);
/* out[i] = sum_b partials[b*len + i] in block order */
int opy_sum_partials_f32(const float* partials, int nb, int len, float* out, void* stream);
/* master weights W[tap][ci][co] (float) -> bf16 [tap][co][ci]; flip=1 gives the backward-data weights */
int opy_conv3x3_prep_weights(const float* w, void* out, int flip, void* stream);
/* g *= (y > 0) */
int opy_relu_bwd_f32(float* g, const float* y, int64_t n, void* stream);
/* weighted squared error over the first nc <= 3 channels: gradient tensor and per-block FP64 loss partials */
int opy_mse3_f32(const float* out, const float* tgt, float* grad, int64_t npix, int nc, double w0, double w1,
                 double w2, double* partials, int grid, void* stream);
/* Adam update of n FP32 parameters (bias-corrected, step >= 1) */
int opy_adam_f32(float* p, const float* g, float* m, float* v, int n, double lr, double b1, double b2, double eps,
                 int step, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* OPYMIZE_AMD_H */
