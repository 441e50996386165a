/*
 * CPU oracle for RELION's coarse diff2 search.  TEST INFRASTRUCTURE ONLY:
 * this file is the checker used by tests/, the smoke check and the CPU
 * baseline leg of bench.py.  Nothing in the product library links it.
 *
 * Two restatements of the reference arithmetic, both plain C:
 *
 *  oracle_diff2_f64  -- float64 everywhere, exact trilinear interpolation
 *                       with clamp-to-edge addressing.  The parity target.
 *      projection   acc_projectorkernel_impl.h:178-252 (rotate by e0,e1,e3,
 *                   e4,e6,e7, scale by the padding factor, int-truncated r^2
 *                   against maxR2_padded, Hermitian flip for xp < 0, sample
 *                   at (xp, yp - initY, zp - initZ)); clamp addressing from
 *                   acc_projector_impl.h:98-103; trilinear blend order from
 *                   cuda_device_utils.cuh:74-108.
 *      phase shift  cuda_device_utils.cuh:132-151 (translatePixelDouble:
 *                   phi = x*tx + y*ty, Xs = X * (cos phi + i sin phi)).
 *      score        diff2.cuh:205-365 (cuda_kernel_diff2_coarse_double:
 *                   diff2[o*T+t] += sum_p corr_p/2 * |P_op - Xs_tp|^2).
 *
 *  oracle_diff2_f32  -- the same loop in float32 with libm sincosf and a
 *                       float32 trilinear, the arithmetic of the reference's
 *                       CUDA-core kernel cuda_kernel_diff2_coarse
 *                       (diff2.cuh:369-537).  Timed as the CPU baseline;
 *                       OpenMP over orientations.
 *
 * Pixels [0, n_proj) are projected; pixels [n_proj, n) are treated as outside
 * the projector radius (P = 0), which is how the filtered ("in-radius first")
 * inputs of the optimised kernels account for the tail
 * (coarse_splitImg_BCF_PMO_TMDB_NFUSE.cuh:463-467).
 *
 * The CUDA texture unit quantises interpolation weights to 8 bits; this
 * oracle interpolates exactly, so its projection differs from the
 * reference GPU's at the 1e-3 level of P (the contract is formula parity).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    const float *re, *im;
    int mdlX, mdlY, mdlZ, initY, initZ;
} vol_t;

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* exact trilinear, clamp-to-edge, planar real/imag, x fastest */
static void tri_f64(const vol_t *v, double xp, double yp, double zp, double *re, double *im)
{
    double fx0 = floor(xp), fy0 = floor(yp), fz0 = floor(zp);
    double fx = xp - fx0, fy = yp - fy0, fz = zp - fz0;
    int x0 = (int)fx0, y0 = (int)fy0, z0 = (int)fz0;
    int xa = clampi(x0, 0, v->mdlX - 1), xb = clampi(x0 + 1, 0, v->mdlX - 1);
    int ya = clampi(y0, 0, v->mdlY - 1), yb = clampi(y0 + 1, 0, v->mdlY - 1);
    int za = clampi(z0, 0, v->mdlZ - 1), zb = clampi(z0 + 1, 0, v->mdlZ - 1);
    size_t sx = (size_t)v->mdlX, sxy = (size_t)v->mdlX * (size_t)v->mdlY;
    size_t i000 = za * sxy + ya * sx + xa, i001 = za * sxy + ya * sx + xb;
    size_t i010 = za * sxy + yb * sx + xa, i011 = za * sxy + yb * sx + xb;
    size_t i100 = zb * sxy + ya * sx + xa, i101 = zb * sxy + ya * sx + xb;
    size_t i110 = zb * sxy + yb * sx + xa, i111 = zb * sxy + yb * sx + xb;
    const float *planes[2] = {v->re, v->im};
    double out[2];
    for (int c = 0; c < 2; ++c) {
        const float *m = planes[c];
        double dx00 = m[i000] + (m[i001] - (double)m[i000]) * fx;
        double dx01 = m[i100] + (m[i101] - (double)m[i100]) * fx;
        double dx10 = m[i010] + (m[i011] - (double)m[i010]) * fx;
        double dx11 = m[i110] + (m[i111] - (double)m[i110]) * fx;
        double dxy0 = dx00 + (dx10 - dx00) * fy;
        double dxy1 = dx01 + (dx11 - dx01) * fy;
        out[c] = dxy0 + (dxy1 - dxy0) * fz;
    }
    *re = out[0];
    *im = out[1];
}

static void project_f64(const vol_t *v, double pad, int maxR2_padded, const float *e,
                        double x, double y, double *re, double *im)
{
    double xp = ((double)e[0] * x + (double)e[1] * y) * pad;
    double yp = ((double)e[3] * x + (double)e[4] * y) * pad;
    double zp = ((double)e[6] * x + (double)e[7] * y) * pad;
    int r2 = (int)(xp * xp + yp * yp + zp * zp);
    if (r2 > maxR2_padded) { *re = 0.0; *im = 0.0; return; }
    int conj = xp < 0;
    if (conj) { xp = -xp; yp = -yp; zp = -zp; }
    tri_f64(v, xp, yp - v->initY, zp - v->initZ, re, im);
    if (conj) *im = -*im;
}

/* Sample the projection of one orientation at n pixels (float64). */
void oracle_project_f64(const float *vre, const float *vim, int mdlX, int mdlY, int mdlZ,
                        int initY, int initZ, float pad, int maxR2_padded, const float *e9,
                        int n, const double *xs, const double *ys, double *out_re, double *out_im)
{
    vol_t v = {vre, vim, mdlX, mdlY, mdlZ, initY, initZ};
    for (int p = 0; p < n; ++p)
        project_f64(&v, (double)pad, maxR2_padded, e9, xs[p], ys[p], out_re + p, out_im + p);
}

/* Projections of O orientations at n pixels (float64, OpenMP over the
 * orientations): out[o*n + p], the per-orientation loop of oracle_diff2_f64
 * without the score.  Used by the expanded-form evaluator in oracle.py. */
void oracle_project_many_f64(const float *vre, const float *vim, int mdlX, int mdlY, int mdlZ,
                             int initY, int initZ, float pad, int maxR2_padded,
                             const float *eulers, int O, int n, const double *xs,
                             const double *ys, double *out_re, double *out_im)
{
    vol_t v = {vre, vim, mdlX, mdlY, mdlZ, initY, initZ};
#pragma omp parallel for schedule(dynamic, 16)
    for (int o = 0; o < O; ++o)
        for (int p = 0; p < n; ++p)
            project_f64(&v, (double)pad, maxR2_padded, eulers + 9 * (size_t)o, xs[p], ys[p],
                        out_re + (size_t)o * n + p, out_im + (size_t)o * n + p);
}

/* Phase-shift one image by one translation (float64). */
void oracle_translate_f64(int n, const double *xs, const double *ys, float tx, float ty,
                          const float *re, const float *im, double *out_re, double *out_im)
{
    for (int p = 0; p < n; ++p) {
        double phi = xs[p] * (double)tx + ys[p] * (double)ty;
        double s = sin(phi), c = cos(phi);
        out_re[p] = c * re[p] - s * im[p];
        out_im[p] = c * im[p] + s * re[p];
    }
}

/* Coarse diff2 of one particle image against O orientations x T translations.
 * out[o*T+t] = (init ? init[o*T+t] : 0) + sum_p corr_p/2 |P_op - Xs_tp|^2     */
void oracle_diff2_f64(const float *vre, const float *vim, int mdlX, int mdlY, int mdlZ,
                      int initY, int initZ, float pad, int maxR2_padded,
                      const float *eulers, int O, const float *trans_x, const float *trans_y,
                      int T, const float *re, const float *im, const float *corr,
                      const double *xs, const double *ys, int n, int n_proj,
                      const double *init, double *out)
{
    vol_t v = {vre, vim, mdlX, mdlY, mdlZ, initY, initZ};
#pragma omp parallel
    {
        double *pr = (double *)__builtin_alloca(sizeof(double) * (size_t)n);
        double *pi = (double *)__builtin_alloca(sizeof(double) * (size_t)n);
#pragma omp for schedule(dynamic, 4)
        for (int o = 0; o < O; ++o) {
            const float *e = eulers + 9 * (size_t)o;
            for (int p = 0; p < n; ++p) {
                pr[p] = pi[p] = 0.0;
                if (p < n_proj)
                    project_f64(&v, (double)pad, maxR2_padded, e, xs[p], ys[p], pr + p, pi + p);
            }
            for (int t = 0; t < T; ++t) {
                double acc = init ? init[(size_t)o * T + t] : 0.0;
                double tx = trans_x[t], ty = trans_y[t];
                for (int p = 0; p < n; ++p) {
                    double phi = xs[p] * tx + ys[p] * ty;
                    double s = sin(phi), c = cos(phi);
                    double xr = c * re[p] - s * im[p];
                    double xi = c * im[p] + s * re[p];
                    double dr = pr[p] - xr, di = pi[p] - xi;
                    acc += (dr * dr + di * di) * ((double)corr[p] / 2.0);
                }
                out[(size_t)o * T + t] = acc;
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* float32 restatement of the CUDA-core kernel (CPU baseline)             */
/* ---------------------------------------------------------------------- */
static void tri_f32(const vol_t *v, float xp, float yp, float zp, float *re, float *im)
{
    float fx0 = floorf(xp), fy0 = floorf(yp), fz0 = floorf(zp);
    float fx = xp - fx0, fy = yp - fy0, fz = zp - fz0;
    int x0 = (int)fx0, y0 = (int)fy0, z0 = (int)fz0;
    int xa = clampi(x0, 0, v->mdlX - 1), xb = clampi(x0 + 1, 0, v->mdlX - 1);
    int ya = clampi(y0, 0, v->mdlY - 1), yb = clampi(y0 + 1, 0, v->mdlY - 1);
    int za = clampi(z0, 0, v->mdlZ - 1), zb = clampi(z0 + 1, 0, v->mdlZ - 1);
    size_t sx = (size_t)v->mdlX, sxy = (size_t)v->mdlX * (size_t)v->mdlY;
    size_t i000 = za * sxy + ya * sx + xa, i001 = za * sxy + ya * sx + xb;
    size_t i010 = za * sxy + yb * sx + xa, i011 = za * sxy + yb * sx + xb;
    size_t i100 = zb * sxy + ya * sx + xa, i101 = zb * sxy + ya * sx + xb;
    size_t i110 = zb * sxy + yb * sx + xa, i111 = zb * sxy + yb * sx + xb;
    const float *planes[2] = {v->re, v->im};
    float out[2];
    for (int c = 0; c < 2; ++c) {
        const float *m = planes[c];
        float dx00 = m[i000] + (m[i001] - m[i000]) * fx;
        float dx01 = m[i100] + (m[i101] - m[i100]) * fx;
        float dx10 = m[i010] + (m[i011] - m[i010]) * fx;
        float dx11 = m[i110] + (m[i111] - m[i110]) * fx;
        float dxy0 = dx00 + (dx10 - dx00) * fy;
        float dxy1 = dx01 + (dx11 - dx01) * fy;
        out[c] = dxy0 + (dxy1 - dxy0) * fz;
    }
    *re = out[0];
    *im = out[1];
}

/* Orientations are processed 16 at a time with the 16 projections of a pixel
 * held together and one sincosf per (translation, pixel) per group, the
 * work split of cuda_kernel_diff2_coarse (D2C_EULERS_PER_BLOCK_REF3D = 16,
 * common.h:43-44; diff2.cuh:441-530).  Adds into out[o*T+t].           */
#define EPB 16
int oracle_diff2_f32(const float *vre, const float *vim, int mdlX, int mdlY, int mdlZ,
                     int initY, int initZ, float pad, int maxR2_padded,
                     const float *eulers, int O, const float *trans_x, const float *trans_y,
                     int T, const float *re, const float *im, const float *corr,
                     const float *coor_xy, int n, int n_proj, float *out, int n_threads)
{
    vol_t v = {vre, vim, mdlX, mdlY, mdlZ, initY, initZ};
    int used = 1;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
    {
#pragma omp single
        used = omp_get_num_threads();
    }
#endif
    int groups = (O + EPB - 1) / EPB;
#pragma omp parallel
    {
        float *pr = (float *)__builtin_alloca(sizeof(float) * (size_t)n * EPB);
        float *pi = (float *)__builtin_alloca(sizeof(float) * (size_t)n * EPB);
#pragma omp for schedule(dynamic, 1)
        for (int g = 0; g < groups; ++g) {
            int o0 = g * EPB, ne = O - o0 < EPB ? O - o0 : EPB;
            for (int p = 0; p < n; ++p) {
                float x = coor_xy[2 * p], y = coor_xy[2 * p + 1];
                for (int j = 0; j < EPB; ++j) {
                    float *r = pr + (size_t)p * EPB + j, *q = pi + (size_t)p * EPB + j;
                    *r = *q = 0.f;
                    if (p >= n_proj || j >= ne) continue;
                    const float *e = eulers + 9 * (size_t)(o0 + j);
                    float xp = (e[0] * x + e[1] * y) * pad;
                    float yp = (e[3] * x + e[4] * y) * pad;
                    float zp = (e[6] * x + e[7] * y) * pad;
                    int r2 = (int)(xp * xp + yp * yp + zp * zp);
                    if (r2 > maxR2_padded) continue;
                    int cj = xp < 0;
                    if (cj) { xp = -xp; yp = -yp; zp = -zp; }
                    tri_f32(&v, xp, yp - (float)initY, zp - (float)initZ, r, q);
                    if (cj) *q = -*q;
                }
            }
            for (int t = 0; t < T; ++t) {
                float acc[EPB];
                for (int j = 0; j < EPB; ++j) acc[j] = 0.f;
                float tx = trans_x[t], ty = trans_y[t];
                for (int p = 0; p < n; ++p) {
                    float x = coor_xy[2 * p], y = coor_xy[2 * p + 1];
                    float s, c;
                    sincosf(x * tx + y * ty, &s, &c);
                    float xr = c * re[p] - s * im[p];
                    float xi = c * im[p] + s * re[p];
                    float w = corr[p] / 2.f;
                    for (int j = 0; j < EPB; ++j) {
                        float dr = pr[(size_t)p * EPB + j] - xr;
                        float di = pi[(size_t)p * EPB + j] - xi;
                        acc[j] += (dr * dr + di * di) * w;
                    }
                }
                for (int j = 0; j < ne; ++j) out[(size_t)(o0 + j) * T + t] += acc[j];
            }
        }
    }
    return used;
}

/* ---------------------------------------------------------------------- */
/* cross-correlation score (float64)                                      */
/* ---------------------------------------------------------------------- */
/* cuda_kernel_diff2_CC_coarse (diff2.cuh:691-809):
 *   out[o*T+t] = - sum_p corr_p Re(P_op conj(Xs_tp)) / sqrt(sum_p corr_p |P_op|^2)
 * (overwrites; the full corr, not corr/2).  Pixel coordinates follow the CC
 * kernel's own rule: y > maxR maps to y - imgY when y >= imgY - maxR, else
 * x = maxR (diff2.cuh:765-771); out[] is written, not accumulated. */
void oracle_cc_f64(const float *vre, const float *vim, int mdlX, int mdlY, int mdlZ, int initY,
                   int initZ, float pad, int maxR2_padded, int imgX, int imgY, int maxR,
                   const float *eulers, int O, const float *trans_x, const float *trans_y, int T,
                   const float *re, const float *im, const float *corr, int image_size,
                   double *out)
{
    vol_t v = {vre, vim, mdlX, mdlY, mdlZ, initY, initZ};
#pragma omp parallel
    {
        double *pr = (double *)__builtin_alloca(sizeof(double) * (size_t)image_size);
        double *pi = (double *)__builtin_alloca(sizeof(double) * (size_t)image_size);
        double *xs = (double *)__builtin_alloca(sizeof(double) * (size_t)image_size);
        double *ys = (double *)__builtin_alloca(sizeof(double) * (size_t)image_size);
        for (int p = 0; p < image_size; ++p) {
            int x = p % imgX, y = p / imgX;
            if (y > maxR) {
                if (y >= imgY - maxR)
                    y -= imgY;
                else
                    x = maxR;
            }
            xs[p] = x;
            ys[p] = y;
        }
#pragma omp for schedule(dynamic, 4)
        for (int o = 0; o < O; ++o) {
            const float *e = eulers + 9 * (size_t)o;
            double norm = 0.0;
            for (int p = 0; p < image_size; ++p) {
                project_f64(&v, (double)pad, maxR2_padded, e, xs[p], ys[p], pr + p, pi + p);
                norm += (pr[p] * pr[p] + pi[p] * pi[p]) * (double)corr[p];
            }
            for (int t = 0; t < T; ++t) {
                double w = 0.0;
                for (int p = 0; p < image_size; ++p) {
                    double phi = xs[p] * (double)trans_x[t] + ys[p] * (double)trans_y[t];
                    double s = sin(phi), c = cos(phi);
                    double xr = c * re[p] - s * im[p];
                    double xi = c * im[p] + s * re[p];
                    w += (pr[p] * xr + pi[p] * xi) * (double)corr[p];
                }
                out[(size_t)o * T + t] = -(w / sqrt(norm));
            }
        }
    }
}

/* Fine-pass diff2 over job lists (cuda_kernel_diff2_fine, diff2.cuh:540-682):
 * job b scores orientation rot_idx[job_idx[b]] against translations
 * trans_idx[job_idx[b]] + j, j < job_num[b]:
 *   out[job_idx[b] + j] = sum_init + sum_p corr_p/2 |P(p) - X_p e^{i phi}|^2
 * over the unfiltered image, with that kernel's coordinate rule (rows between
 * the half-planes parked at x = maxR, diff2.cuh:593-617).  float64. */
void oracle_fine_f64(const float *vre, const float *vim, int mdlX, int mdlY, int mdlZ, int initY,
                     int initZ, float pad, int maxR2_padded, int imgX, int imgY, int maxR,
                     const float *eulers, const float *trans_x, const float *trans_y,
                     const float *re, const float *im, const float *corr, int image_size,
                     double sum_init, int n_jobs, const long long *rot_idx,
                     const long long *trans_idx, const long long *job_idx,
                     const long long *job_num, double *out)
{
    vol_t v = {vre, vim, mdlX, mdlY, mdlZ, initY, initZ};
#pragma omp parallel
    {
        double *pr = (double *)__builtin_alloca(sizeof(double) * (size_t)image_size);
        double *pi = (double *)__builtin_alloca(sizeof(double) * (size_t)image_size);
        double *xs = (double *)__builtin_alloca(sizeof(double) * (size_t)image_size);
        double *ys = (double *)__builtin_alloca(sizeof(double) * (size_t)image_size);
        for (int p = 0; p < image_size; ++p) {
            int x = p % imgX, y = p / imgX;
            if (y > maxR) {
                if (y >= imgY - maxR)
                    y -= imgY;
                else
                    x = maxR;
            }
            xs[p] = x;
            ys[p] = y;
        }
#pragma omp for schedule(dynamic, 1)
        for (int b = 0; b < n_jobs; ++b) {
            const long long j0 = job_idx[b];
            const float *e = eulers + 9 * (size_t)rot_idx[j0];
            for (int p = 0; p < image_size; ++p)
                project_f64(&v, (double)pad, maxR2_padded, e, xs[p], ys[p], pr + p, pi + p);
            for (long long j = 0; j < job_num[b]; ++j) {
                const long long it = trans_idx[j0] + j;
                double acc = 0.0;
                for (int p = 0; p < image_size; ++p) {
                    double phi = xs[p] * (double)trans_x[it] + ys[p] * (double)trans_y[it];
                    double s = sin(phi), c = cos(phi);
                    double dr = pr[p] - (c * re[p] - s * im[p]);
                    double di = pi[p] - (c * im[p] + s * re[p]);
                    acc += (dr * dr + di * di) * 0.5 * (double)corr[p];
                }
                out[j0 + j] = acc + sum_init;
            }
        }
    }
}
