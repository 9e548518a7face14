/*
 * CPU restatement of the SigPDE signature-kernel scheme -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle that the HIP kernels are checked against. Only the tests, the
 * smoke check and the benchmark's CPU-baseline leg load it; no product code path calls it.
 *
 * Two arithmetic modes restate the reference's device arithmetic:
 *   FP64 "as written": every operation in double, in Python's left-to-right order,
 *     no FMA contraction (built with -ffp-contract=off).
 *   FP32 "ref": float storage of the three live anti-diagonals, per-cell arithmetic in double
 *     except inc*inc, which the reference's Numba typing keeps in float32; the increment tensor
 *     and the scale product are float32; every stored cell is rounded to float.
 *
 * Two contraction settings of pde_step (orc_set_contract), for both modes:
 *   0 (default): as written, every product and sum rounded separately;
 *   1: the FMA contraction NVVM applies to the expression on an NVIDIA GPU (the reference's
 *      target; Numba leaves NVVM's contraction on, sigpde never sets fastmath):
 *        A = fma(c2, 1/12, fma(c, 0.5, 1)),  B = fma(c2, -1/12, 1),
 *        k = fma(k01 + k10, A, -(k00 * B)).
 *      fma(c, 0.5, 1) equals 1 + 0.5 c (0.5 c is exact). The root-find's bracket arithmetic
 *      stays as written in both settings.
 *
 * Reference files followed (stojl/SigPDE):
 *   sigpde/cuda_device_functions.py:3-5    pde_step
 *   sigpde/cuda_pairwise_kernels.py:8-85   wavefront, boundary k=1, output k[M][N], scaled variant
 *   sigpde/cuda_pairwise_kernels.py:88-166 Chandrupatla root-find
 *   sigpde/cuda_gram_kernels.py:8-165      Gram / symmetric Gram / scaled variants
 *   sigpde/torch/kernel.py:13-21           4-term second difference of the point Gram, / 4^order
 *   sigpde/torch/static_kernels.py:3-11    linear static kernel scale*<x_s, y_t>
 *   sigpde/utils.py:3-7                    refined length ((L-1)<<order)+1
 *
 * Any traversal that respects the dependencies k[i-1][j-1], k[i-1][j], k[i][j-1] gives the same
 * values as the reference's anti-diagonal sweep, because each cell reads only those three
 * neighbours; this restatement sweeps row by row.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* Contraction setting of pde_step (see the header): set once before a call, read by every
 * thread. */
static int g_contract = 0;

void orc_set_contract(int on) { g_contract = on ? 1 : 0; }
int orc_get_contract(void) { return g_contract; }

/* pde_step, sigpde/cuda_device_functions.py:3-5, on double c and c2 = c**2:
 *   (k01 + k10) * (1. + 0.5 * inc + (1./12) * inc**2) - k00 * (1. - (1./12) * inc**2) */
static inline double step_core(double k00, double k01, double k10, double c, double c2)
{
    if (g_contract) {
        double a = fma(c2, 1.0 / 12, fma(c, 0.5, 1.0));
        double b = fma(c2, -1.0 / 12, 1.0);
        return fma(k01 + k10, a, -(k00 * b));
    }
    double a = (1.0 + 0.5 * c) + (1.0 / 12) * c2;
    double b = 1.0 - (1.0 / 12) * c2;
    return (k01 + k10) * a - k00 * b;
}

/* FP64: every operand double. */
static inline double step_f64(double k00, double k01, double k10, double c)
{
    return step_core(k00, k01, k10, c, c * c);
}

/* pde_step under the reference's FP32 typing: inc and inc**2 are float32, the literals and the
 * loaded k values promote everything else to float64 (sigpde/cuda_device_functions.py:5). */
static inline double step_f32ref(double k00, double k01, double k10, float c)
{
    float c2 = c * c;
    return step_core(k00, k01, k10, (double)c, (double)c2);
}

/* One pair, FP64. inc is m x n with row stride ld; c = inc * scale (sigpde/cuda_pairwise_kernels.py:69).
 * Multiplying by scale = 1.0 is exact, so the unscaled kernels are the scale = 1 case. */
static double pair_f64(const double *inc, int64_t ld, int64_t m, int64_t n, int order, double scale,
                       double *buf)
{
    const int64_t M = m << order, N = n << order;
    double *prev = buf, *cur = buf + (N + 1);
    for (int64_t j = 0; j <= N; ++j) prev[j] = 1.0;
    for (int64_t i = 1; i <= M; ++i) {
        const double *row = inc + ((i - 1) >> order) * ld;
        cur[0] = 1.0;
        for (int64_t j = 1; j <= N; ++j) {
            double c = row[(j - 1) >> order] * scale;
            cur[j] = step_f64(prev[j - 1], prev[j], cur[j - 1], c);
        }
        double *t = prev; prev = cur; cur = t;
    }
    return prev[N];
}

/* One pair, FP32-ref: float storage, c = float(inc * scale). */
static float pair_f32ref(const float *inc, int64_t ld, int64_t m, int64_t n, int order, float scale,
                         float *buf)
{
    const int64_t M = m << order, N = n << order;
    float *prev = buf, *cur = buf + (N + 1);
    for (int64_t j = 0; j <= N; ++j) prev[j] = 1.0f;
    for (int64_t i = 1; i <= M; ++i) {
        const float *row = inc + ((i - 1) >> order) * ld;
        cur[0] = 1.0f;
        for (int64_t j = 1; j <= N; ++j) {
            float c = row[(j - 1) >> order] * scale;
            cur[j] = (float)step_f32ref(prev[j - 1], prev[j], cur[j - 1], c);
        }
        float *t = prev; prev = cur; cur = t;
    }
    return prev[N];
}

/* One pair with a double increment and float storage: the root-find evaluates
 * inc * xt**2 in double even for float32 paths (sigpde/cuda_pairwise_kernels.py:114). */
static float pair_f32store_f64c(const float *inc, int64_t ld, int64_t m, int64_t n, int order,
                                double scale, float *buf)
{
    const int64_t M = m << order, N = n << order;
    float *prev = buf, *cur = buf + (N + 1);
    for (int64_t j = 0; j <= N; ++j) prev[j] = 1.0f;
    for (int64_t i = 1; i <= M; ++i) {
        const float *row = inc + ((i - 1) >> order) * ld;
        cur[0] = 1.0f;
        for (int64_t j = 1; j <= N; ++j) {
            double c = (double)row[(j - 1) >> order] * scale;
            cur[j] = (float)step_f64(prev[j - 1], prev[j], cur[j - 1], c);
        }
        float *t = prev; prev = cur; cur = t;
    }
    return prev[N];
}

/* ---------------------------------------------------------------- pairwise
 * sigpde/cuda_pairwise_kernels.py:8-85: out[b] = k_b[M][N] for Delta_b of shape (m, n).
 * sx/sy may be NULL (unscaled). Scaled: scale = sy[b] * sx[b] in the input dtype (:61). */
void orc_pairwise_f64(const double *inc, int64_t B, int64_t m, int64_t n, int order,
                      const double *sx, const double *sy, double *out)
{
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 2 * ((n << order) + 1));
#pragma omp for schedule(dynamic, 1)
        for (int64_t b = 0; b < B; ++b) {
            double s = (sx || sy) ? (sy ? sy[b] : 1.0) * (sx ? sx[b] : 1.0) : 1.0;
            out[b] = pair_f64(inc + b * m * n, n, m, n, order, s, buf);
        }
        free(buf);
    }
}

void orc_pairwise_f32(const float *inc, int64_t B, int64_t m, int64_t n, int order,
                      const float *sx, const float *sy, float *out)
{
#pragma omp parallel
    {
        float *buf = (float *)malloc(sizeof(float) * 2 * ((n << order) + 1));
#pragma omp for schedule(dynamic, 1)
        for (int64_t b = 0; b < B; ++b) {
            float s = (sx || sy) ? (sy ? sy[b] : 1.0f) * (sx ? sx[b] : 1.0f) : 1.0f;
            out[b] = pair_f32ref(inc + b * m * n, n, m, n, order, s, buf);
        }
        free(buf);
    }
}

/* The root-find's function evaluation for float32 paths (sigpde/cuda_pairwise_kernels.py:114):
 * float storage, c = inc * s[b] in double (s = xt**2, a double), c**2 in double. */
void orc_pairwise_f32_dscale(const float *inc, int64_t B, int64_t m, int64_t n, int order,
                             const double *s, float *out)
{
#pragma omp parallel
    {
        float *buf = (float *)malloc(sizeof(float) * 2 * ((n << order) + 1));
#pragma omp for schedule(dynamic, 1)
        for (int64_t b = 0; b < B; ++b)
            out[b] = pair_f32store_f64c(inc + b * m * n, n, m, n, order, s[b], buf);
        free(buf);
    }
}

/* ---------------------------------------------------------------- Gram
 * sigpde/cuda_gram_kernels.py:8-76: out[bx*ldo + by] = K(x_bx, y_by) from the increment tensor
 * (Bx, By, m, n). Scaled: scale = sx[bx] * sy[by] (:52).
 * symmetric != 0 (sigpde/cuda_gram_kernels.py:79-165): only by <= bx is solved and the value is
 * mirrored to out[by*ldo + bx]; requires Bx == By and m == n. */
void orc_gram_f64(const double *inc, int64_t Bx, int64_t By, int64_t m, int64_t n, int order,
                  const double *sx, const double *sy, int symmetric, double *out, int64_t ldo)
{
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 2 * ((n << order) + 1));
#pragma omp for schedule(dynamic, 1) collapse(2)
        for (int64_t bx = 0; bx < Bx; ++bx)
            for (int64_t by = 0; by < By; ++by) {
                if (symmetric && by > bx) continue;
                double s = (sx || sy) ? (sx ? sx[bx] : 1.0) * (sy ? sy[by] : 1.0) : 1.0;
                double r = pair_f64(inc + (bx * By + by) * m * n, n, m, n, order, s, buf);
                out[bx * ldo + by] = r;
                if (symmetric && bx != by) out[by * ldo + bx] = r;
            }
        free(buf);
    }
}

void orc_gram_f32(const float *inc, int64_t Bx, int64_t By, int64_t m, int64_t n, int order,
                  const float *sx, const float *sy, int symmetric, float *out, int64_t ldo)
{
#pragma omp parallel
    {
        float *buf = (float *)malloc(sizeof(float) * 2 * ((n << order) + 1));
#pragma omp for schedule(dynamic, 1) collapse(2)
        for (int64_t bx = 0; bx < Bx; ++bx)
            for (int64_t by = 0; by < By; ++by) {
                if (symmetric && by > bx) continue;
                float s = (sx || sy) ? (sx ? sx[bx] : 1.0f) * (sy ? sy[by] : 1.0f) : 1.0f;
                float r = pair_f32ref(inc + (bx * By + by) * m * n, n, m, n, order, s, buf);
                out[bx * ldo + by] = r;
                if (symmetric && bx != by) out[by * ldo + bx] = r;
            }
        free(buf);
    }
}

/* ---------------------------------------------------------------- linear increments
 * sigpde/torch/static_kernels.py:7-11 and sigpde/torch/kernel.py:13-21:
 *   ip[s][t] = kscale * <x_s, y_t>          (bmm / einsum, then the scalar multiply)
 *   D[s][t]  = ((ip[s+1][t+1] + ip[s][t]) - ip[s+1][t]) - ip[s][t+1]
 *   D       /= 4^order
 * The point Gram is accumulated sequentially over the feature axis. */
void orc_linear_inc_f64(const double *x, const double *y, int64_t Lx, int64_t Ly, int64_t d,
                        double kscale, int order, double *inc)
{
    double *ip = (double *)malloc(sizeof(double) * Lx * Ly);
    for (int64_t s = 0; s < Lx; ++s)
        for (int64_t t = 0; t < Ly; ++t) {
            double acc = 0.0;
            for (int64_t k = 0; k < d; ++k) acc += x[s * d + k] * y[t * d + k];
            ip[s * Ly + t] = kscale * acc;
        }
    const double div = (double)((int64_t)1 << (2 * order));
    for (int64_t s = 0; s + 1 < Lx; ++s)
        for (int64_t t = 0; t + 1 < Ly; ++t) {
            double v = ((ip[(s + 1) * Ly + t + 1] + ip[s * Ly + t]) - ip[(s + 1) * Ly + t]) -
                       ip[s * Ly + t + 1];
            inc[s * (Ly - 1) + t] = v / div;
        }
    free(ip);
}

void orc_linear_inc_f32(const float *x, const float *y, int64_t Lx, int64_t Ly, int64_t d,
                        double kscale, int order, float *inc)
{
    float *ip = (float *)malloc(sizeof(float) * Lx * Ly);
    const float ks = (float)kscale;
    for (int64_t s = 0; s < Lx; ++s)
        for (int64_t t = 0; t < Ly; ++t) {
            float acc = 0.0f;
            for (int64_t k = 0; k < d; ++k) acc += x[s * d + k] * y[t * d + k];
            ip[s * Ly + t] = ks * acc;
        }
    const float div = (float)((int64_t)1 << (2 * order));
    for (int64_t s = 0; s + 1 < Lx; ++s)
        for (int64_t t = 0; t + 1 < Ly; ++t) {
            float v = ((ip[(s + 1) * Ly + t + 1] + ip[s * Ly + t]) - ip[(s + 1) * Ly + t]) -
                      ip[s * Ly + t + 1];
            inc[s * (Ly - 1) + t] = v / div;
        }
    free(ip);
}

/* Fused linear Gram from paths: X (Bx, Lx, d), Y (By, Ly, d), per pair the increment matrix above
 * then the PDE solve. Used for large sampled checks and the CPU baseline, where materialising the
 * (Bx, By, m, n) increment tensor would not fit. sx/sy nullable. */
void orc_linear_gram_f64(const double *X, const double *Y, int64_t Bx, int64_t By, int64_t Lx,
                         int64_t Ly, int64_t d, double kscale, int order, const double *sx,
                         const double *sy, double *out, int64_t ldo)
{
    const int64_t m = Lx - 1, n = Ly - 1;
#pragma omp parallel
    {
        double *inc = (double *)malloc(sizeof(double) * m * n);
        double *buf = (double *)malloc(sizeof(double) * 2 * ((n << order) + 1));
#pragma omp for schedule(dynamic, 1) collapse(2)
        for (int64_t bx = 0; bx < Bx; ++bx)
            for (int64_t by = 0; by < By; ++by) {
                orc_linear_inc_f64(X + bx * Lx * d, Y + by * Ly * d, Lx, Ly, d, kscale, order, inc);
                double s = (sx || sy) ? (sx ? sx[bx] : 1.0) * (sy ? sy[by] : 1.0) : 1.0;
                out[bx * ldo + by] = pair_f64(inc, n, m, n, order, s, buf);
            }
        free(buf);
        free(inc);
    }
}

void orc_linear_gram_f32(const float *X, const float *Y, int64_t Bx, int64_t By, int64_t Lx,
                         int64_t Ly, int64_t d, double kscale, int order, const float *sx,
                         const float *sy, float *out, int64_t ldo)
{
    const int64_t m = Lx - 1, n = Ly - 1;
#pragma omp parallel
    {
        float *inc = (float *)malloc(sizeof(float) * m * n);
        float *buf = (float *)malloc(sizeof(float) * 2 * ((n << order) + 1));
#pragma omp for schedule(dynamic, 1) collapse(2)
        for (int64_t bx = 0; bx < Bx; ++bx)
            for (int64_t by = 0; by < By; ++by) {
                orc_linear_inc_f32(X + bx * Lx * d, Y + by * Ly * d, Lx, Ly, d, kscale, order, inc);
                float s = (sx || sy) ? (sx ? sx[bx] : 1.0f) * (sy ? sy[by] : 1.0f) : 1.0f;
                out[bx * ldo + by] = pair_f32ref(inc, n, m, n, order, s, buf);
            }
        free(buf);
        free(inc);
    }
}

/* ---------------------------------------------------------------- Chandrupatla
 * sigpde/cuda_pairwise_kernels.py:88-166. For path b, with T = target[b] (the normalizer applied
 * to K(x,x)) and K1 = knorm[b] = K(x,x), find lambda in [0, 1] with
 *   F(lambda) = PDE(lambda^2 * Delta_b) - T = 0,
 * Delta_b of shape (m, m). Every bracket/update operation is double (Python float literals);
 * F(xt) is evaluated with c = Delta * xt**2 in double.
 * f32 storage mode (f32 != 0): inc/target/knorm/out are float; fb and ft are float differences
 * promoted to double, as the reference's typing gives for float32 arrays (:97, :132).
 * iters[b] (nullable) receives the number of wavefront evaluations performed.
 * fixed[b] (nullable) receives the first evaluation count after which the iteration left its
 * state (a, b, c, fa, fb, fc, t) unchanged, or -1: from there on every iteration repeats (the
 * reference keeps iterating to maxit; the output does not change). Reporting it changes nothing
 * else. */
static double chandrupatla_one(const void *inc, int64_t m, int order, double T, double K1,
                               int f32, int maxit, double tol, void *buf, int *iters, int *fixed)
{
    double a = 0.0, c = 0.0, b = 1.0;
    double fa = 1.0 - T;
    double fb = f32 ? (double)((float)K1 - (float)T) : K1 - T;
    double fc = fa, t = 0.5;
    double outv = 0.0; /* the reference's output buffer starts at zero (kernel.py:212,249) */
    int it = 0;
    if (fixed) *fixed = -1;
    for (; it < maxit;) {
        const double a0 = a, b0 = b, c0 = c, fa0 = fa, fb0 = fb, fc0 = fc, t0 = t;
        double xt = a + t * (b - a);
        double s = xt * xt;
        double sol;
        ++it;
        if (f32)
            sol = (double)pair_f32store_f64c((const float *)inc, m, m, m, order, s, (float *)buf);
        else
            sol = pair_f64((const double *)inc, m, m, m, order, s, (double *)buf);
        double ft;
        if (isnan(sol))
            ft = INFINITY;
        else
            ft = f32 ? (double)((float)sol - (float)T) : sol - T;
        if (ft * fa >= 0) {
            c = a;
            fc = fa;
        } else {
            c = b;
            b = a;
            fc = fb;
            fb = fa;
        }
        a = xt;
        fa = ft;
        if (fabs(fa) < fabs(fb)) {
            outv = a;
            if (fabs(fa) < tol) break;
        } else {
            outv = b;
            if (fabs(fb) < tol) break;
        }
        double xi = (a - b) / (c - b);
        double phi = (fa - fb) / (fc - fb);
        if (phi * phi < xi && (1 - phi) * (1 - phi) < 1 - xi)
            t = fa / (fb - fa) * fc / (fb - fc) + (c - a) / (b - a) * fa / (fc - fa) * fb / (fc - fb);
        else
            t = 0.5;
        if (!(0 < t && t < 1)) t = 0.5;
        if (fixed && *fixed < 0 && a == a0 && b == b0 && c == c0 && fa == fa0 && fb == fb0 &&
            fc == fc0 && t == t0)
            *fixed = it;
    }
    if (iters) *iters = it;
    return outv;
}

void orc_chandrupatla_f64(const double *inc, int64_t B, int64_t m, int order, const double *target,
                          const double *knorm, int maxit, double tol, double *out, int *iters,
                          int *fixed)
{
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 2 * ((m << order) + 1));
#pragma omp for schedule(dynamic, 1)
        for (int64_t b = 0; b < B; ++b)
            out[b] = chandrupatla_one(inc + b * m * m, m, order, target[b], knorm[b], 0, maxit, tol,
                                      buf, iters ? iters + b : NULL, fixed ? fixed + b : NULL);
        free(buf);
    }
}

void orc_chandrupatla_f32(const float *inc, int64_t B, int64_t m, int order, const float *target,
                          const float *knorm, int maxit, double tol, float *out, int *iters,
                          int *fixed)
{
#pragma omp parallel
    {
        float *buf = (float *)malloc(sizeof(float) * 2 * ((m << order) + 1));
#pragma omp for schedule(dynamic, 1)
        for (int64_t b = 0; b < B; ++b)
            out[b] = (float)chandrupatla_one(inc + b * m * m, m, order, (double)target[b],
                                             (double)knorm[b], 1, maxit, tol, buf,
                                             iters ? iters + b : NULL, fixed ? fixed + b : NULL);
        free(buf);
    }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
