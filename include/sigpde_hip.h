/*
 * sigpde_hip.h -- C ABI of the MI355X (gfx950) signature-kernel PDE solver, libsigpde_hip.so.
 *
 * Conventions
 *   - Every pointer is a device pointer owned by the caller, except where stated. A call whose
 *     result is empty (a batch of 0) returns 0 at once; its pointers may then be NULL.
 *   - Arrays are C-contiguous unless a stride argument is given. Strides count elements.
 *   - dtype: SIGPDE_F32 (float storage, the reference's float32 semantics) or SIGPDE_F64.
 *   - Work is enqueued on `stream` (a hipStream_t; NULL = the default stream) and is
 *     stream-ordered; the call returns without synchronising.
 *   - Return value: 0 on success, otherwise a SIGPDE_E* code; sigpde_last_error() then holds a
 *     message for the calling thread.
 *   - Refined lengths follow the reference: Lref = ((L - 1) << order) + 1
 *     (sigpde/utils.py:6-7); with m = Lx - 1 and n = Ly - 1 increments the refined grid of one
 *     pair has (m << order) x (n << order) cells. Supported: (m << order) <= 2^20 refined
 *     rows and (n << order) <= 2^22 refined columns, with (rows / 16 + 1) x (columns + 48)
 *     <= 2^30, in the entry points that take increments (past 17695 SIGPDE_F32 / 14927
 *     SIGPDE_F64 refined columns the band-boundary lines go to global memory, allocated
 *     stream-ordered by the call: hipMallocAsync / hipFreeAsync on `stream`); the fused linear
 *     entry points take up to 9679 refined columns (4559 in SIGPDE_PREC_FP64), and keep their
 *     lines in a stream-ordered global block too when lines that long in LDS would hold far
 *     fewer waves per CU (the same results). Larger inputs return SIGPDE_EINVAL naming the limit.
 *
 * The reference has no C interface: its "FFI" is the Python/Numba launch layer
 * sigpde/PDESolvers.py, whose kernel launches these entry points replace one for one.
 */
#ifndef SIGPDE_HIP_H
#define SIGPDE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum sigpde_dtype { SIGPDE_F32 = 0, SIGPDE_F64 = 1 };

enum sigpde_status {
    SIGPDE_OK = 0,
    SIGPDE_EINVAL = 1,   /* bad argument (shape, dtype, null pointer, unsupported size) */
    SIGPDE_EDEVICE = 2,  /* HIP runtime error (launch failure, no device) */
    SIGPDE_ENOMEM = 3,   /* device allocation failed */
    SIGPDE_ETIMEOUT = 4  /* a bounded wait between the workgroups of a spread pair gave up; the
                            call's outputs are invalid (sigpde_set_spread) */
};

/* Message describing the last failure on the calling thread ("" if none). */
const char *sigpde_last_error(void);

/* Library version string. */
const char *sigpde_version(void);

/*
 * Wavefront from precomputed increments, one PDE per pair.
 * Replaces PairwisePDESolver.solve / solve_scaled (sigpde/PDESolvers.py:56-84), which launch
 * sigpde_pairwise / sigpde_pairwise_scaled (sigpde/cuda_pairwise_kernels.py:8-85).
 *   inc : (B, m, n) increments of dtype, pair b at inc + b * inc_stride_b, row stride n
 *   sx, sy : (B,) scales or NULL; when either is given c = inc * (sy[b] * sx[b]) in dtype
 *   out : (B,) result k[m<<order][n<<order]
 */
int sigpde_pde_pairwise(int dtype, const void *inc, int64_t B, int64_t m, int64_t n,
                        int64_t inc_stride_b, int order, const void *sx, const void *sy,
                        void *out, void *stream);
/* Few long pairs: when the GPU holds G workgroups per pair at once and G of them shorten a pair's
 * sequential chain, sigpde_pde_pairwise and the robust root-finds spread each pair over G
 * workgroups (bands handed between workgroups through global memory, with bounded waits). Such a
 * call allocates a small stream-ordered workspace and SYNCHRONISES `stream` before it returns (to
 * check that no wait gave up: SIGPDE_ETIMEOUT); spread calls run one at a time in the process, and
 * a stream under graph capture never takes the route. Results equal the other routes' bit for bit.
 *
 * sigpde_set_spread: no counterpart in the reference (this route's settings, process-wide).
 *   groups: 0 automatic (default), 1 never spread, 2..32 that many workgroups per pair whenever
 *           they are all resident (tests use it to run the route on small grids);
 *   wait_seconds: wall-clock limit of one wait (default 2, at most 600);
 *   test_hook: 0, or 1 = workgroup 0 of each pair publishes no progress, so the waits on it give
 *           up (a test of the failure path: the call returns SIGPDE_ETIMEOUT). */
int sigpde_set_spread(int groups, double wait_seconds, int test_hook);

/*
 * Wavefront from a precomputed increment tile, one PDE per (x, y) pair.
 * Replaces GramPDESolver.solve / solve_scaled and SymmetricGramPDESolver.solve / solve_scaled
 * (sigpde/PDESolvers.py:160-190, 209-245), which launch sigpde_gram[_scaled] and
 * sigpde_gram_symmetric[_scaled] (sigpde/cuda_gram_kernels.py:8-165).
 *   inc : tile (bx, by, m, n); pair (i, j) at inc + i*stride_x + j*stride_y, row stride n
 *   sx : (bx,) or NULL, sy : (by,) or NULL; c = inc * (sx[i] * sy[j]) in dtype
 *   out : result written at out[(off_x + i) * ldo + (off_y + j)]
 *   symmetric != 0: pairs with off_y + j > off_x + i are skipped and every solved value is
 *   also written to out[(off_y + j) * ldo + (off_x + i)] (sigpde/cuda_gram_kernels.py:85-86,
 *   110-113).
 */
int sigpde_pde_gram(int dtype, const void *inc, int64_t bx, int64_t by, int64_t m, int64_t n,
                    int64_t stride_x, int64_t stride_y, int order, const void *sx,
                    const void *sy, int symmetric, int64_t off_x, int64_t off_y, void *out,
                    int64_t ldo, void *stream);

/*
 * Robust normalisation: per-path Chandrupatla root-find of PDE(lambda^2 * Delta) = target.
 * Replaces RobustSigPDE._norm_factors (sigpde/torch/kernel.py:180-195) with its launches of
 * sigpde_pairwise and sigpde_pairwise_norm_chandraputla (sigpde/cuda_pairwise_kernels.py:
 * 8-39, 88-166) through PairwisePDESolver.solve / solve_norms (sigpde/PDESolvers.py:56-101).
 *   inc : (B, m, m) increments of each path with itself, path b at inc + b * inc_stride_b
 *   normalizer: SIGPDE_NORM_LOG  -- knorm = K(x,x) is solved here, NaN -> +inf, and the target is
 *                                  2 - 1/(1 + log knorm) in dtype (sigpde/torch/kernel.py:23-24);
 *                                  knorm (B,) receives K(x,x);
 *               SIGPDE_NORM_GIVEN -- knorm (B,) holds K(x,x) and target (B,) the normaliser's
 *                                  value, both computed by the caller.
 *   lambda_out : (B,) the root, set to 0 where knorm < 1 (sigpde/torch/kernel.py:195)
 *   iters_out : (B,) int32 wavefront evaluations of the root-find, or NULL
 */
enum sigpde_normalizer {
    SIGPDE_NORM_LOG = 0,
    SIGPDE_NORM_GIVEN = 1,
    SIGPDE_NORM_TENSOR = 2,  /* built-in, sigpde_robust_norm_builtin */
    SIGPDE_NORM_CLIP = 3     /* built-in, sigpde_robust_norm_builtin */
};

int sigpde_robust_norm(int dtype, const void *inc, int64_t B, int64_t m, int64_t inc_stride_b,
                       int order, int normalizer, const void *target, void *knorm, int maxit,
                       double tol, void *lambda_out, int32_t *iters_out, void *stream);

/*
 * The same root-find with a built-in normaliser evaluated on the device between the K(x,x) solve
 * and the root-find (no host round trip; the reference takes any Python callable there,
 * sigpde/torch/kernel.py:180-195, which sigpde_robust_norm serves with SIGPDE_NORM_GIVEN):
 *   SIGPDE_NORM_LOG    2 - 1/(1 + log x), the reference's default (p0, p1 unused);
 *   SIGPDE_NORM_TENSOR x for x <= C, else C + C^(1+a) (C^-a - x^-a) / a with C = p0 >= 1,
 *                      a = p1 > 0 (increasing, at most x, bounded by C (1 + 1/a));
 *   SIGPDE_NORM_CLIP   min(x, C) with C = p0 >= 1.
 * Evaluated in dtype, x = K(x,x) after NaN -> +inf; knorm (B,) receives K(x,x).
 */
int sigpde_robust_norm_builtin(int dtype, const void *inc, int64_t B, int64_t m,
                               int64_t inc_stride_b, int order, int normalizer, double p0,
                               double p1, void *knorm, int maxit, double tol, void *lambda_out,
                               int32_t *iters_out, void *stream);

/*
 * Fused linear static kernel (k(x, y) = kscale * <x, y>, sigpde/torch/static_kernels.py:3-11):
 * increments are formed on the device from the paths and multiplied on MFMA straight into the
 * wavefront, so no (Bx, By, m, n) increment tensor exists. These replace the chain
 * LinearKernel.gram / .pairwise -> gram_inner_product / pairwise_inner_product (torch) ->
 * GramPDESolver / PairwisePDESolver launches (sigpde/torch/kernel.py:13-21, 57-61, 100-122,
 * 145-167) for the linear kernel.
 *   precision SIGPDE_PREC_FP16X2: increments split into per-path-scaled fp16 hi + lo, products
 *     accumulated in fp32 on MFMA; the wavefront runs in double; results in dtype.
 *   precision SIGPDE_PREC_BF16X3: the same scaled increments split into three bf16 parts, the six
 *     products of magnitude >= 2^-16 of hi*hi accumulated in fp32 on bf16 MFMA; double wavefront.
 *   precision SIGPDE_PREC_FP32: plain FP32 -- increments rounded to float, products on fp32
 *     MFMA, the wavefront in the reference's float32 typing (c, c^2 in float, the rest in double,
 *     cells stored as float); results in dtype.
 *   precision SIGPDE_PREC_FP64: increments in double on fp64 MFMA, the wavefront in double (the
 *     reference scheme in float64, without the increment tensor); d <= 64; results in dtype.
 *   precision SIGPDE_PREC_FP16X2F: the fp16x2 increments feeding a float32 wavefront (float
 *     coefficients and state); dyadic order <= 2 and up to 512 refined rows and columns (other
 *     grids: SIGPDE_EINVAL), where it stays within 1e-5 of float64; ~14% faster than
 *     SIGPDE_PREC_FP16X2 at 255 x 255.
 *   X : (Bx, Lx, d), Y : (By, Ly, d) paths of dtype, C-contiguous; 1 <= d <= 128
 *   sx (Bx,), sy (By,) : optional scales of dtype (c *= sx[i] * sy[j]), or NULL
 *   gram: out[i * ldo + j] = K(x_i, y_j); symmetric != 0 requires Y == X and writes the lower
 *     triangle and its mirror.
 *   pairwise: out[b] = K(x_b, y_b) for Bx == By == B.
 *   workspace : device buffer of at least sigpde_linear_workspace_bytes(precision, ...) bytes
 *     (the device-side increments and per-path scales); same_paths != 0 when Y is X (one split
 *     serves both). Returns 0 for invalid sizes or an unknown precision.
 */
enum sigpde_precision {
    SIGPDE_PREC_FP16X2 = 1,
    SIGPDE_PREC_FP32 = 2,
    SIGPDE_PREC_BF16X3 = 3,
    SIGPDE_PREC_FP64 = 4,
    SIGPDE_PREC_FP16X2F = 5
};

size_t sigpde_linear_workspace_bytes(int precision, int64_t Bx, int64_t Lx, int64_t By, int64_t Ly,
                                     int64_t d, int same_paths);

int sigpde_linear_gram(int dtype, int precision, const void *X, int64_t Bx, int64_t Lx,
                       const void *Y, int64_t By, int64_t Ly, int64_t d, double kscale, int order,
                       const void *sx, const void *sy, int symmetric, void *out, int64_t ldo,
                       void *workspace, size_t workspace_bytes, void *stream);

int sigpde_linear_pairwise(int dtype, int precision, const void *X, const void *Y, int64_t B,
                           int64_t Lx, int64_t Ly, int64_t d, double kscale, int order,
                           const void *sx, const void *sy, void *out, void *workspace,
                           size_t workspace_bytes, void *stream);

/*
 * Up to 64 Grams of one shape class (dtype, precision, lengths Lx / Ly, d, order, kscale) in one
 * launch: several calls of SigPDE.gram (sigpde/torch/kernel.py:66-124 for the linear kernel),
 * coalesced by a serving loop. X holds nx paths and Y ny paths, C-contiguous; requests is a HOST
 * array of nreq rows (x0, bx, y0, by, symmetric): request r is the Gram of X paths x0 .. x0+bx-1
 * against Y paths y0 .. y0+by-1 (symmetric != 0: those Y paths equal the X paths, bx == by; the
 * lower triangle is solved and mirrored). Results go to out, request after request, each a
 * row-major (bx, by) block (request r at the sum of the earlier bx * by). sx (nx,), sy (ny,):
 * optional scales. Workspace: sigpde_linear_workspace_bytes(precision, nx, Lx, ny, Ly, d, 0).
 * Each pair is solved exactly as sigpde_linear_gram solves it: results equal that call's bit
 * for bit.
 */
int sigpde_linear_gram_batch(int dtype, int precision, const void *X, int64_t nx, int64_t Lx,
                             const void *Y, int64_t ny, int64_t Ly, int64_t d, double kscale,
                             int order, const int64_t *requests, int64_t nreq, const void *sx,
                             const void *sy, void *out, void *workspace, size_t workspace_bytes,
                             void *stream);

/*
 * Increments of the non-linear static kernels in one pass. These replace the chain
 * RBFKernel / L1RBFKernel .pairwise / .gram (sigpde/torch/static_kernels.py:13-54) followed by
 * pairwise_inner_product / gram_inner_product (sigpde/torch/kernel.py:13-21):
 *   out[a][b][i][j] = (((k[i+1][j+1] + k[i][j]) - k[i+1][j]) - k[i][j+1]) / 4^order,
 *   i < m, j < n, out C-contiguous (A, B, m, n); pairwise calls use B = 1.
 * sigpde_rbf_increments: k[i][j] = exp(-((-2 ip) + (xx + yy)) * (1 / sigma)) in dtype, where
 *   ip (a, b, i, j) = <x_i, y_j> sits at a*s[0] + b*s[1] + i*s[2] + j*s[3] (ip_strides, elements;
 *   points i <= m, j <= n), xx (a, i) at a*xx_stride_a + i, yy (a, b, j) at a*yy_stride_a +
 *   b*yy_stride_b + j: the reference's RBF formula operation for operation, from the caller's
 *   GEMM and row sums.
 * sigpde_l1rbf_increments: k[i][j] = exp(-(sum_k |x_ik - y_jk|) * (1 / sigma)) from the paths
 *   X (A, m+1, d), Y (B, n+1, d), C-contiguous; pairwise != 0 pairs X_a with Y_a (B = A).
 */
int sigpde_rbf_increments(int dtype, const void *ip, int64_t A, int64_t B, int64_t m, int64_t n,
                          const int64_t *ip_strides, const void *xx, int64_t xx_stride_a,
                          const void *yy, int64_t yy_stride_a, int64_t yy_stride_b, double sigma,
                          int order, void *out, void *stream);

int sigpde_l1rbf_increments(int dtype, const void *X, int64_t A, const void *Y, int64_t B,
                            int64_t m, int64_t n, int64_t d, int pairwise, double sigma, int order,
                            void *out, void *stream);

/* The same for the linear static kernel, k[i][j] = kscale * ip (a, b, i, j) with ip laid out as
 * for sigpde_rbf_increments: the reference scheme's increments (LinearKernel.gram / .pairwise then
 * the second difference, sigpde/torch/static_kernels.py:3-11, sigpde/torch/kernel.py:13-21) in one
 * pass, operation for operation. */
int sigpde_linear_increments(int dtype, const void *ip, int64_t A, int64_t B, int64_t m, int64_t n,
                             const int64_t *ip_strides, double kscale, int order, void *out,
                             void *stream);

/*
 * Reference-scheme Gram of a static kernel with the increments formed inside the solve (no
 * increment tensor). Replaces SigPDE._gram / _gram_sym (sigpde/torch/kernel.py:74-171): the point
 * values of LinearKernel / RBFKernel / L1RBFKernel (sigpde/torch/static_kernels.py:3-54), their
 * 4-term second difference / 4^order (kernel.py:18-21) and the Gram solvers
 * (sigpde/cuda_gram_kernels.py:8-160, scaled: c = inc * (sx[i] * sy[j]) in dtype), in one launch.
 *   kernel: SIGPDE_STATIC_LINEAR (param = scale), SIGPDE_STATIC_RBF or SIGPDE_STATIC_L1RBF
 *     (param = sigma); point values in dtype as the reference's torch expressions evaluate them
 *     (x / sigma as x * (1 / sigma)); inner products and distance sums in another summation order
 *     than torch's (equal when exact).
 *   X (Bx, Lx, d), Y (By, Ly, d) C-contiguous; out[i * ldo + j] = K(X_i, Y_j); symmetric != 0
 *     (Y == X) solves j <= i and mirrors. sx, sy: (Bx,), (By,) scales or NULL.
 *   Limits: dyadic order <= 3; d <= 64 (float32) / 32 (float64) for the linear and RBF kernels,
 *     d <= 16 (float32) / 8 (float64) for L1-RBF; refined columns whose lines fit one wave's LDS. Outside them the call
 *     returns SIGPDE_EINVAL (sigpde_static_gram_supported says so beforehand); the increments
 *     entry points above plus sigpde_pde_gram take every shape.
 *   workspace: sigpde_static_workspace_bytes(...) bytes of device memory (same_paths: Y == X).
 */
enum sigpde_static_kernel {
    SIGPDE_STATIC_LINEAR = 0,
    SIGPDE_STATIC_RBF = 1,
    SIGPDE_STATIC_L1RBF = 2
};

int sigpde_static_gram_supported(int dtype, int kernel, int64_t Lx, int64_t Ly, int64_t d,
                                 int order);

size_t sigpde_static_workspace_bytes(int dtype, int kernel, int64_t Bx, int64_t Lx, int64_t By,
                                     int64_t Ly, int64_t d, int same_paths);

int sigpde_static_gram(int dtype, int kernel, const void *X, int64_t Bx, int64_t Lx,
                       const void *Y, int64_t By, int64_t Ly, int64_t d, double param, int order,
                       const void *sx, const void *sy, int symmetric, void *out, int64_t ldo,
                       void *workspace, size_t workspace_bytes, void *stream);

/* sigpde_issue_rates: no counterpart in the reference (a measurement helper for the roofline of
 * the fused kernels, which issue MFMA and VALU work on the same SIMDs). Runs timing kernels
 * synchronously on the current device and writes
 *   out[0] clock64 ticks per wave64 FP32 FMA per SIMD,   out[1] the same for FP64 FMA,
 *   out[2] ticks per v_mfma_f32_16x16x32_f16 per SIMD (these three: two one-wave blocks per
 *          SIMD of every CU),
 *   out[3] time of MFMA and FP32 FMA issued by the same waves / (MFMA alone + FMA alone)
 *          (1: the two serialize, max/sum: they overlap),
 *   out[4] the device's reported clock in Hz, out[5] CUs, out[6] blocks per rate probe;
 * with n >= 12 also (waves of one role each, MFMA only or FMA only, on the same SIMDs)
 *   out[7] one MFMA wave beside one FMA wave per SIMD: the slower role's time / (the two roles'
 *          times alone), out[8] the full-overlap value of that ratio (max / sum of the times
 *          alone), out[9] the ratio with two waves of each role per SIMD,
 *   out[10] ticks per FP32 add with a DPP row_shr:1 operand per SIMD (two waves per SIMD),
 *   out[11] dependent-issue latency of FP32 FMA in ticks (one wave per SIMD, one chain),
 *   out[12] (n >= 13) the same for the DPP add.
 * n >= 7. Allocates and frees its own small device buffers. */
int sigpde_issue_rates(double *out, int n);

#ifdef __cplusplus
}
#endif

#endif /* SIGPDE_HIP_H */
