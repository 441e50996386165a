// Auxiliary gfx950 kernels of the coarse diff2 search: image preparation,
// projection precompute, |P|^2 norms, per-particle argmin and the direct-form
// float32 kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "device_common.hpp"

namespace rc {

// ---------------------------------------------------------------------------
// image preparation: one block per particle
// ---------------------------------------------------------------------------
// xw[i][p]        = corr_p * X_p            (p < npix, zero to kpad)
// half_corr[i][p] = corr_p / 2              (p < npix, zero to kpad)
// coords[p]       = (x_p, y_p)              (written by block 0)
// xnorm[i]        = sum_{p < full_npix} corr_p/2 |X_p|^2   (float64)
// ascale[i]       = power-of-two scale of max |corr X| (FP16 form)
// Pixel coordinates come from coor_xy, or from the pixel index
// (x = i % imgX, y = i / imgX, y -= imgY if y > maxR; diff2.cuh:182-188).
__global__ __launch_bounds__(256) void prep_kernel(int npix, int full_npix, int kpad,
                                                   const float *__restrict__ real,
                                                   const float *__restrict__ imag,
                                                   const float *__restrict__ corr,
                                                   const float *__restrict__ coor_xy, int imgX,
                                                   int imgY, int maxR, float2 *__restrict__ xw,
                                                   float2 *__restrict__ coords,
                                                   double *__restrict__ xnorm,
                                                   float *__restrict__ half_corr,
                                                   float *__restrict__ ascale, int cc_coords)
{
    const int i = blockIdx.x;
    const size_t src = (size_t)i * full_npix;
    const size_t dst = (size_t)i * kpad;
    double acc = 0.0;
    float amax = 0.f;
    for (int p = threadIdx.x; p < kpad; p += blockDim.x) {
        float2 w = make_float2(0.f, 0.f);
        float hc = 0.f;
        if (p < npix) {
            const float xr = real[src + p], xi = imag[src + p], c = corr[src + p];
            w = make_float2(c * xr, c * xi);
            hc = c * 0.5f;
            acc += (double)hc * ((double)xr * xr + (double)xi * xi);
            amax = fmaxf(amax, fmaxf(fabsf(w.x), fabsf(w.y)));
        }
        xw[dst + p] = w;
        if (half_corr) half_corr[dst + p] = hc;
        if (blockIdx.x == 0) {
            float2 cxy = make_float2(0.f, 0.f);
            if (p < npix) {
                if (coor_xy) {
                    cxy = make_float2(coor_xy[2 * p], coor_xy[2 * p + 1]);
                } else {
                    int x = p % imgX, y = p / imgX;
                    if (y > maxR) {
                        // the CC kernel parks rows between the two half-planes at
                        // x = maxR (diff2.cuh:765-771); the diff2 kernels wrap them
                        if (!cc_coords || y >= imgY - maxR)
                            y -= imgY;
                        else
                            x = maxR;
                    }
                    cxy = make_float2((float)x, (float)y);
                }
            }
            coords[p] = cxy;
        }
    }
    for (int p = npix + threadIdx.x; p < full_npix; p += blockDim.x) {
        const float xr = real[src + p], xi = imag[src + p], c = corr[src + p];
        acc += (double)(c * 0.5f) * ((double)xr * xr + (double)xi * xi);
    }
    __shared__ double s_acc[256];
    __shared__ float s_max[256];
    s_acc[threadIdx.x] = acc;
    s_max[threadIdx.x] = amax;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s_acc[threadIdx.x] += s_acc[threadIdx.x + w];
            s_max[threadIdx.x] = fmaxf(s_max[threadIdx.x], s_max[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        xnorm[i] = s_acc[0];
        if (ascale) ascale[i] = pow2_scale(s_max[0]);
    }
}

hipError_t launch_prep(int n_particles, int npix, int full_npix, int kpad, const float *real,
                       const float *imag, const float *corr, const float *coor_xy, int imgX,
                       int imgY, int maxR, float2 *xw, float2 *coords, double *xnorm,
                       float *half_corr, float *ascale, hipStream_t s, bool cc_coords)
{
    hipLaunchKernelGGL(prep_kernel, dim3(n_particles), dim3(256), 0, s, npix, full_npix, kpad,
                       real, imag, corr, coor_xy, imgX, imgY, maxR, xw, coords, xnorm, half_corr,
                       ascale, (int)cc_coords);
    return hipGetLastError();
}

// Single unfiltered image of the drop-in: the same outputs as prep_kernel for
// the image's pixels reordered so that those a rotation can bring inside the
// projector's mask come first (index order), n_in of them; the others are
// outside the mask for every rotation (their projections are 0, so they add
// corr/2 |X|^2 to the score and nothing else) and enter xnorm only.  A pixel
// is kept when (float)(x^2 + y^2) * pad2 <= thr, thr = (maxR2_padded + 1)
// (1 + 2^-12): int(r^2) <= maxR2_padded needs r^2 < maxR2_padded + 1, and the
// rotated float32 coordinates change r^2 by far less than 2^-12 of it.  The
// host counts the same float expression to size the search.  One workgroup:
// each round of 1024 pixels is compacted by wave ballots and a scan of the
// 16 waves' counts.
constexpr int kPartT = 1024;

__global__ __launch_bounds__(kPartT) void prep_partition_kernel(
    int image_size, int kpad, const float *__restrict__ real, const float *__restrict__ imag,
    const float *__restrict__ corr, int imgX, int imgY, int maxR, float pad2, float thr,
    float2 *__restrict__ xw, float2 *__restrict__ coords, double *__restrict__ xnorm,
    float *__restrict__ half_corr, float *__restrict__ ascale, int cc_coords)
{
    __shared__ int s_cnt[kPartT / 64];
    __shared__ double s_acc[kPartT];
    __shared__ float s_max[kPartT];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double acc = 0.0;
    float amax = 0.f;
    int base = 0;
    for (int r0 = 0; r0 < image_size; r0 += kPartT) {
        const int p = r0 + tid;
        bool in = false;
        float2 w = make_float2(0.f, 0.f), cxy = make_float2(0.f, 0.f);
        float hc = 0.f;
        if (p < image_size) {
            int x = p % imgX, y = p / imgX;
            if (y > maxR) {
                if (!cc_coords || y >= imgY - maxR)
                    y -= imgY;
                else
                    x = maxR;
            }
            in = (float)(x * x + y * y) * pad2 <= thr;
            const float xr = real[p], xi = imag[p], c = corr[p];
            hc = c * 0.5f;
            acc += (double)hc * ((double)xr * xr + (double)xi * xi);
            w = make_float2(c * xr, c * xi);
            cxy = make_float2((float)x, (float)y);
        }
        const unsigned long long m = __ballot(in);
        if (lane == 0) s_cnt[wave] = __popcll(m);
        __syncthreads();
        int before = 0, total = 0;
        for (int j = 0; j < kPartT / 64; ++j) {
            const int c = s_cnt[j];
            before += j < wave ? c : 0;
            total += c;
        }
        // (dst < kpad always holds: the host sized kpad by the same count;
        // the test keeps a disagreement from writing past the buffers)
        const int dst = base + before + __popcll(m & ((1ull << lane) - 1ull));
        if (in && dst < kpad) {
            xw[dst] = w;
            coords[dst] = cxy;
            half_corr[dst] = hc;
            amax = fmaxf(amax, fmaxf(fabsf(w.x), fabsf(w.y)));
        }
        base += total;
        __syncthreads();   // s_cnt is rewritten by the next round
    }
    for (int p = base + tid; p < kpad; p += kPartT) {
        xw[p] = make_float2(0.f, 0.f);
        coords[p] = make_float2(0.f, 0.f);
        half_corr[p] = 0.f;
    }
    s_acc[tid] = acc;
    s_max[tid] = amax;
    __syncthreads();
    for (int w = kPartT / 2; w > 0; w >>= 1) {
        if (tid < w) {
            s_acc[tid] += s_acc[tid + w];
            s_max[tid] = fmaxf(s_max[tid], s_max[tid + w]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        xnorm[0] = s_acc[0];
        if (ascale) ascale[0] = pow2_scale(s_max[0]);
    }
}

hipError_t launch_prep_partition(int image_size, int kpad, const float *real, const float *imag,
                                 const float *corr, int imgX, int imgY, int maxR, float pad2,
                                 float thr, float2 *xw, float2 *coords, double *xnorm,
                                 float *half_corr, float *ascale, hipStream_t s, bool cc_coords)
{
    hipLaunchKernelGGL(prep_partition_kernel, dim3(1), dim3(kPartT), 0, s, image_size, kpad, real,
                       imag, corr, imgX, imgY, maxR, pad2, thr, xw, coords, xnorm, half_corr,
                       ascale, (int)cc_coords);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// projection precompute: block = 8 orientations x one 32-pixel chunk.
// One-dimensional grid, XCD-aware as project_wide_kernel's: block b runs on
// XCD b % 8, which takes whole 8-orientation groups (b % 8, b % 8 + 8, ...)
// with all their chunks, so a group's central sections are gathered
// through one L2.
// ---------------------------------------------------------------------------
// proj[(c*O + o)*nkc + kc] = encoded 32 samples; pabs2[(c*O+o)*kpad + p] = |P|^2
template <int FORM>
__global__ __launch_bounds__(256) void project_kernel(const ProjView *__restrict__ vols,
                                                      const float *__restrict__ bscale,
                                                      const float *__restrict__ eulers, int O,
                                                      int n_classes,
                                                      const float2 *__restrict__ coords, int npix,
                                                      int kpad, unsigned char *__restrict__ proj,
                                                      float *__restrict__ pabs2)
{
    constexpr int CB = chunk_bytes(FORM);
    constexpr int GB = group_bytes(FORM);
    __shared__ __attribute__((aligned(16))) unsigned char s_out[8 * CB];
    const int nkc = kpad / kChunk;
    const int n_og = (O + 7) / 8;   // 8-orientation groups per class
    const int jx = (int)(blockIdx.x >> 3);
    const int G = (int)(blockIdx.x & 7) + 8 * (jx / nkc);
    if (G >= n_og * n_classes) return;
    const int kc = jx % nkc;
    const int cls = G / n_og;
    const int by = G - cls * n_og;
    const int ol = threadIdx.x >> 5;
    const int pl = threadIdx.x & 31;
    const int o = by * 8 + ol;
    const int p = kc * kChunk + pl;
    const ProjView v = vols[cls];
    float2 val = make_float2(0.f, 0.f);
    if (o < O && p < npix) {
        const float *e = eulers + 9 * (size_t)o;
        const float2 xy = coords[p];
        val = project_sample(v, e[0], e[1], e[3], e[4], e[6], e[7], xy.x, xy.y);
    }
    const float sc = (FORM == kFP16) ? bscale[cls] : 1.f;
    store_encoded<FORM>(s_out + ol * CB + (pl >> 3) * GB, pl & 7, val.x, val.y, sc);
    if (o < O) pabs2[((size_t)cls * O + o) * kpad + p] = val.x * val.x + val.y * val.y;
    __syncthreads();
    constexpr int units = 8 * CB / 16;
    for (int u = threadIdx.x; u < units; u += 256) {
        const int oo = by * 8 + u / (CB / 16);
        if (oo < O) {
            const int seg = u % (CB / 16);
            uint4 d = reinterpret_cast<const uint4 *>(s_out)[u];
            reinterpret_cast<uint4 *>(proj + (((size_t)cls * O + oo) * nkc + kc) * CB)[seg] = d;
        }
    }
}

hipError_t launch_project(int form, const ProjView *vols, const float *bscale, int n_classes,
                          const float *eulers, int O, const float2 *coords, int npix, int kpad,
                          unsigned char *proj, float *pabs2, hipStream_t s)
{
    const long long groups = (long long)((O + 7) / 8) * n_classes;
    const dim3 grid((unsigned)(8 * ((groups + 7) / 8) * (kpad / kChunk)));
    switch (form) {
    case kFP32:
        hipLaunchKernelGGL(project_kernel<kFP32>, grid, dim3(256), 0, s, vols, bscale, eulers, O,
                           n_classes, coords, npix, kpad, proj, pabs2);
        break;
    case kBF16X3:
        hipLaunchKernelGGL(project_kernel<kBF16X3>, grid, dim3(256), 0, s, vols, bscale, eulers,
                           O, n_classes, coords, npix, kpad, proj, pabs2);
        break;
    default:
        hipLaunchKernelGGL(project_kernel<kFP16>, grid, dim3(256), 0, s, vols, bscale, eulers, O,
                           n_classes, coords, npix, kpad, proj, pabs2);
        break;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// |P|^2 norms: pnorm[i][n] = sum_p half_corr[i][p] * pabs2[n][p]
// ---------------------------------------------------------------------------
// A small GEMM on the BF16 matrix cores with float32 operands carried
// exactly: each float32 value is split into three bfloat16 pieces,
// v = v1 + v2 + v3 (8 significant bits each, RNE; the remainders are exact
// in float32, so the split loses nothing), and the product a b is summed as
// the six piece products with i + j <= 4: a3 b1 + a1 b3 + a2 b2 + a2 b1 +
// a1 b2 + a1 b1 (smallest first; the omitted a2 b3, a3 b2, a3 b3 are below
// 2^-25 of |a b|).  Every piece product is exact and accumulated in float32
// by v_mfma_f32_16x16x32_bf16: six 16-cycle MFMAs per 16 x 16 x 32 block
// instead of eight 32-cycle FP32 ones.
// Block tile 64 particles x 64 columns, 4 waves of 32 x 32 (2 x 2 fragments);
// k staged through LDS in 32-pixel chunks as [piece][row][32 bf16] (rows
// padded to 80 bytes), the next chunk's global loads issued before the
// current chunk's MFMAs.  A[i][k] = half_corr, B[k][n] = |P|^2; kpad is a
// multiple of 32.
// Summation: float32 within a 32-pixel chunk, chunk sums added in float64
// (a cascaded sum), so the rounding error grows with the chunk length, not
// with the pixel count.
constexpr int kPnK = 32;
constexpr int kPnRowB = kPnK * 2 + 16;   // bytes per staged row (64 + 16 pad)

typedef float pn_f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 pn_bf16x2 __attribute__((ext_vector_type(2)));

// the three bfloat16 pieces of two float32 values, as three packed words
__device__ __forceinline__ void split3_bf16(float x, float y, uint32_t &w1, uint32_t &w2,
                                            uint32_t &w3)
{
    const pn_f32x2 v = {x, y};
    const pn_bf16x2 h1 = __builtin_convertvector(v, pn_bf16x2);
    const pn_f32x2 r1 = v - __builtin_convertvector(h1, pn_f32x2);
    const pn_bf16x2 h2 = __builtin_convertvector(r1, pn_bf16x2);
    const pn_f32x2 r2 = r1 - __builtin_convertvector(h2, pn_f32x2);
    const pn_bf16x2 h3 = __builtin_convertvector(r2, pn_bf16x2);
    w1 = __builtin_bit_cast(uint32_t, h1);
    w2 = __builtin_bit_cast(uint32_t, h2);
    w3 = __builtin_bit_cast(uint32_t, h3);
}

__global__ __launch_bounds__(256) void pnorm_mfma_kernel(int P, int N, int kpad,
                                                         const float *__restrict__ hc,
                                                         const float *__restrict__ pabs2,
                                                         float *__restrict__ pnorm)
{
    __shared__ __attribute__((aligned(16))) unsigned char sA[3][64 * kPnRowB];
    __shared__ __attribute__((aligned(16))) unsigned char sB[3][64 * kPnRowB];
    // XCD-aware order: block b runs on XCD b % 8; each XCD takes a contiguous
    // range of tiles in which the particle tiles of one column tile are
    // adjacent, so that column tile's |P|^2 rows come from its L2
    const int nrow = (P + 63) / 64, ncol = (N + 63) / 64;
    const int total = nrow * ncol;
    const int per_xcd = (total + 7) / 8;
    const int L = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (L >= total) return;
    const int i0 = (L % nrow) * 64, n0 = (L / nrow) * 64;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
    const int r16 = lane & 15, q = lane >> 4;
    // this thread's share of a chunk: row lr, k offsets lk .. lk + 7
    const int lr = tid >> 2, lk = (tid & 3) * 8;
    const bool a_ok = i0 + lr < P, b_ok = n0 + lr < N;
    const float *ap = hc + (size_t)(a_ok ? i0 + lr : 0) * kpad + lk;
    const float *bp = pabs2 + (size_t)(b_ok ? n0 + lr : 0) * kpad + lk;
    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 ra[2], rb[2];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            ra[u] = zero;
            rb[u] = zero;
            if (a_ok) ra[u] = *reinterpret_cast<const float4 *>(ap + k0 + 4 * u);
            if (b_ok) rb[u] = *reinterpret_cast<const float4 *>(bp + k0 + 4 * u);
        }
    };
    // the 8 values of a thread's share, as the three pieces' 16-byte segments
    auto stage = [&](const float4 (&r)[2], unsigned char (*dst)[64 * kPnRowB]) {
        uint32_t w[3][4];
        split3_bf16(r[0].x, r[0].y, w[0][0], w[1][0], w[2][0]);
        split3_bf16(r[0].z, r[0].w, w[0][1], w[1][1], w[2][1]);
        split3_bf16(r[1].x, r[1].y, w[0][2], w[1][2], w[2][2]);
        split3_bf16(r[1].z, r[1].w, w[0][3], w[1][3], w[2][3]);
#pragma unroll
        for (int pc = 0; pc < 3; ++pc)
            *reinterpret_cast<uint4 *>(&dst[pc][lr * kPnRowB + lk * 2]) =
                make_uint4(w[pc][0], w[pc][1], w[pc][2], w[pc][3]);
    };
    double dacc[2][2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) dacc[x][y][v] = 0.0;
    fetch(0);
    for (int k0 = 0; k0 < kpad; k0 += kPnK) {
        __syncthreads();
        stage(ra, sA);
        stage(rb, sB);
        __syncthreads();
        if (k0 + kPnK < kpad) fetch(k0 + kPnK);
        bf16x8 fa[2][3], fb[2][3];
#pragma unroll
        for (int pc = 0; pc < 3; ++pc)
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                fa[x][pc] = *reinterpret_cast<const bf16x8 *>(
                    &sA[pc][(wr + 16 * x + r16) * kPnRowB + q * 16]);
                fb[x][pc] = *reinterpret_cast<const bf16x8 *>(
                    &sB[pc][(wc + 16 * x + r16) * kPnRowB + q * 16]);
            }
        floatx4 acc[2][2];
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) acc[x][y] = floatx4{0.f, 0.f, 0.f, 0.f};
        // piece pairs (i, j), i + j <= 4, smallest first
        constexpr int kPi[6] = {2, 0, 1, 1, 0, 0};
        constexpr int kPj[6] = {0, 2, 1, 0, 1, 0};
#pragma unroll
        for (int t = 0; t < 6; ++t)
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        fa[x][kPi[t]], fb[y][kPj[t]], acc[x][y], 0, 0, 0);
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int v = 0; v < 4; ++v) dacc[x][y][v] += (double)acc[x][y][v];
    }
    // D[row = 4q + v][col = r16] of each 16 x 16 fragment
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + wr + 16 * x + 4 * q + v, n = n0 + wc + 16 * y + r16;
                if (i < P && n < N) pnorm[(size_t)i * N + n] = (float)dacc[x][y][v];
            }
}

hipError_t launch_pnorm(int n_particles, int n_cols, int kpad, const float *half_corr,
                        const float *pabs2, float *pnorm, hipStream_t s)
{
    const int total = ((n_cols + 63) / 64) * ((n_particles + 63) / 64);
    dim3 grid(((total + 7) / 8) * 8);
    hipLaunchKernelGGL(pnorm_mfma_kernel, grid, dim3(256), 0, s, n_particles, n_cols, kpad,
                       half_corr, pabs2, pnorm);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// split-K assembly of one image (single-image entry points)
// ---------------------------------------------------------------------------
// out[o*T+t] = init + pn + xn - c   (or the CC score -c / sqrt(2 pn)), where
// c = sum_s part_c[s][o][t] and pn = sum_s part_pn[s][o] are summed in float64
// over the slices in order s = 0, 1, ..., so the result depends on the slice
// count only, never on the order the slices finished in.
__global__ __launch_bounds__(256) void splitk_assemble_kernel(int O, int T, int S,
                                                              const float *__restrict__ part_c,
                                                              const float *__restrict__ part_pn,
                                                              const double *__restrict__ xnorm,
                                                              const float *init, float *out,
                                                              int cc)
{
    const size_t n = (size_t)O * T;
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int o = (int)(k / T);
    double c = 0.0, pn = 0.0;
    for (int sl = 0; sl < S; ++sl) {
        c += (double)part_c[(size_t)sl * n + k];
        pn += (double)part_pn[(size_t)sl * O + o];
    }
    double d;
    if (cc) {
        d = -c / sqrt(2.0 * pn);
    } else {
        d = pn + xnorm[0] - c;
        if (init) d += (double)init[k];
    }
    out[k] = (float)d;
}

// One workgroup: compares the n floats bit for bit (a reused buffer with
// new contents is a new key), then refreshes the copy if anything differs.
// The check runs over many blocks, each comparing a 1024-float slice of the
// eulers with the cached copy; a block whose slice differs (or every block,
// when forced) refreshes its slice and writes this call's generation into
// the flag.  The flag is never cleared: the fill compares it with the same
// generation, so a call whose eulers all match leaves an older value there
// and the fill does nothing.
constexpr int kCheckSlice = 1024;

__global__ __launch_bounds__(256) void cache_check_kernel(const float *__restrict__ eulers,
                                                          float *__restrict__ copy, int n,
                                                          int force, unsigned gen,
                                                          unsigned *flag)
{
    __shared__ int s_diff;
    if (threadIdx.x == 0) s_diff = force;
    __syncthreads();
    const int i0 = blockIdx.x * kCheckSlice, i1 = min(n, i0 + kCheckSlice);
    int diff = 0;
    for (int i = i0 + threadIdx.x; i < i1; i += 256)
        diff |= __float_as_uint(eulers[i]) != __float_as_uint(copy[i]);
    if (diff) s_diff = 1;
    __syncthreads();
    if (s_diff) {
        for (int i = i0 + threadIdx.x; i < i1; i += 256) copy[i] = eulers[i];
        if (threadIdx.x == 0) *flag = gen;
    }
}

hipError_t launch_cache_check(const float *eulers, float *copy, int n, int force, unsigned gen,
                              unsigned *flag, hipStream_t s)
{
    const int blocks = std::max(1, (n + kCheckSlice - 1) / kCheckSlice);
    hipLaunchKernelGGL(cache_check_kernel, dim3(blocks), dim3(256), 0, s, eulers, copy, n, force,
                       gen, flag);
    return hipGetLastError();
}

// A grid of at most kFillBlocks blocks striding over (orientation, pixel):
// when the flag does not hold this call's generation every block returns at
// once.  The sample is the fused kernels' gather (rows e0, e1, e3, e4, e6,
// e7, pixel coordinates from prep).  512 blocks (two per CU) fill 4608 x
// 2112 samples in 10 % less time than 256, 1024 or 2048 did (a miss of the
// 113 x 4608 x 2112 drop-in call: 192 against 206-212 us per call), and
// giving each XCD a contiguous eighth of the orientations takes that call
// to 177 us; a hit costs the same.
constexpr int kFillBlocks = 512;

__global__ __launch_bounds__(256) void cache_fill_kernel(const unsigned *flag, unsigned gen,
                                                         const ProjView v,
                                                         const float *__restrict__ eulers,
                                                         const float2 *__restrict__ coords,
                                                         int O, int npix, int kpad, float2 *P)
{
    if (*flag != gen) return;
    // XCD-aware: the blocks on XCD x (x = block % 8) stride over the x-th
    // eighth of the (orientation, pixel) range, so neighbouring orientations
    // -- in-plane rotations of one central section -- gather through one L2
    const size_t total = (size_t)O * npix;
    const size_t per = (total + 7) / 8;
    const size_t k0 = (size_t)(blockIdx.x & 7) * per;
    const size_t k1 = min(total, k0 + per);
    const size_t nb = gridDim.x >> 3;
    for (size_t k = k0 + (size_t)(blockIdx.x >> 3) * 256 + threadIdx.x; k < k1; k += nb * 256) {
        const int o = (int)(k / npix), p = (int)(k - (size_t)o * npix);
        const float *e = eulers + 9 * (size_t)o;
        const float2 xy = coords[p];
        P[(size_t)o * kpad + p] =
            project_sample(v, e[0], e[1], e[3], e[4], e[6], e[7], xy.x, xy.y);
    }
}

hipError_t launch_cache_fill(const unsigned *flag, unsigned gen, const ProjView &v,
                             const float *eulers, int O, const float2 *coords, int npix, int kpad,
                             float2 *P, hipStream_t s)
{
    const size_t total = (size_t)O * npix;
    // a multiple of 8 (every XCD takes an equal share of the blocks)
    const int blocks = 8 * (int)std::min<size_t>(kFillBlocks / 8, (total + 2047) / 2048);
    hipLaunchKernelGGL(cache_fill_kernel, dim3(std::max(blocks, 8)), dim3(256), 0, s, flag, gen, v,
                       eulers, coords, O, npix, kpad, P);
    return hipGetLastError();
}

hipError_t launch_splitk_assemble(int O, int T, int S, const float *part_c, const float *part_pn,
                                  const double *xnorm, const float *init, float *out, int cc,
                                  hipStream_t s)
{
    const size_t n = (size_t)O * T;
    hipLaunchKernelGGL(splitk_assemble_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       O, T, S, part_c, part_pn, xnorm, init, out, cc);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// per-particle min / argmin over the cross kernel's partials (float64)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void argmin_kernel(const int *__restrict__ group_begin,
                                                     int parts_per_group,
                                                     const double *__restrict__ part_min,
                                                     const long long *__restrict__ part_idx,
                                                     double *__restrict__ best,
                                                     long long *__restrict__ best_idx)
{
    const int i = blockIdx.x;
    const size_t a = (size_t)group_begin[i] * parts_per_group;
    const size_t b = (size_t)group_begin[i + 1] * parts_per_group;
    double m = DBL_MAX;
    long long id = 0x7fffffffffffffffLL;
    for (size_t u = a + threadIdx.x; u < b; u += blockDim.x) {
        const double v = part_min[u];
        const long long k = part_idx[u];
        if (v < m || (v == m && k < id)) {
            m = v;
            id = k;
        }
    }
    __shared__ double sm[256];
    __shared__ long long si[256];
    sm[threadIdx.x] = m;
    si[threadIdx.x] = id;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const double v = sm[threadIdx.x + w];
            const long long k = si[threadIdx.x + w];
            if (v < sm[threadIdx.x] || (v == sm[threadIdx.x] && k < si[threadIdx.x])) {
                sm[threadIdx.x] = v;
                si[threadIdx.x] = k;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (best) best[i] = sm[0];
        if (best_idx) best_idx[i] = si[0];
    }
}

hipError_t launch_argmin(int n_particles, const int *group_begin, int parts_per_group,
                         const double *part_min, const long long *part_idx, double *best,
                         long long *best_idx, hipStream_t s)
{
    hipLaunchKernelGGL(argmin_kernel, dim3(n_particles), dim3(256), 0, s, group_begin,
                       parts_per_group, part_min, part_idx, best, best_idx);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// direct form, float32 vector ALU
// ---------------------------------------------------------------------------
// The arithmetic of cuda_kernel_diff2_coarse (diff2.cuh:369-537): per
// (orientation, translation) accumulate corr/2 * |P - Xs|^2 over all pixels,
// with one sincos per (translation, pixel) shared by a block's orientations.
// Block = 16 orientations x 64 pixel lanes... laid out as 256 threads:
// thread t handles translations t, t+256, ...; projections of the block's
// 16 orientations for a 64-pixel chunk are staged in LDS.  Deterministic
// (no atomics): one block owns its 16 orientations for all translations.
constexpr int kDirOrient = 16;
constexpr int kDirPix = 64;
constexpr int kDirMaxTPerThread = 4;

__global__ __launch_bounds__(256) void direct_kernel(ProjView v, const float *__restrict__ eulers,
                                                     int O, const float *__restrict__ trans_x,
                                                     const float *__restrict__ trans_y, int T,
                                                     const float2 *__restrict__ coords,
                                                     const float *__restrict__ real,
                                                     const float *__restrict__ imag,
                                                     const float *__restrict__ corr, int npix,
                                                     int full_npix, const float *init,
                                                     float *out)
{
    __shared__ float2 s_p[kDirPix][kDirOrient];
    __shared__ float4 s_x[kDirPix];   // (xr, xi, corr/2, unused)
    __shared__ float2 s_c[kDirPix];
    __shared__ float s_e[kDirOrient][6];
    const int o0 = blockIdx.x * kDirOrient;
    if (threadIdx.x < kDirOrient * 6) {
        const int j = threadIdx.x / 6, k = threadIdx.x % 6;
        const int src[6] = {0, 1, 3, 4, 6, 7};
        s_e[j][k] = (o0 + j < O) ? eulers[9 * (size_t)(o0 + j) + src[k]] : 0.f;
    }
    float acc[kDirMaxTPerThread][kDirOrient];
    float txv[kDirMaxTPerThread], tyv[kDirMaxTPerThread];
#pragma unroll
    for (int u = 0; u < kDirMaxTPerThread; ++u) {
        const int t = threadIdx.x + 256 * u;
        txv[u] = t < T ? trans_x[t] : 0.f;
        tyv[u] = t < T ? trans_y[t] : 0.f;
#pragma unroll
        for (int j = 0; j < kDirOrient; ++j) acc[u][j] = 0.f;
    }
    for (int p0 = 0; p0 < full_npix; p0 += kDirPix) {
        __syncthreads();
        for (int u = threadIdx.x; u < kDirPix * kDirOrient; u += 256) {
            const int pl = u / kDirOrient, j = u % kDirOrient;
            const int p = p0 + pl;
            float2 val = make_float2(0.f, 0.f);
            if (p < npix && o0 + j < O) {
                const float2 xy = coords[p];
                val = project_sample(v, s_e[j][0], s_e[j][1], s_e[j][2], s_e[j][3], s_e[j][4],
                                     s_e[j][5], xy.x, xy.y);
            }
            s_p[pl][j] = val;
        }
        if (threadIdx.x < kDirPix) {
            const int p = p0 + threadIdx.x;
            s_x[threadIdx.x] = p < full_npix
                                   ? make_float4(real[p], imag[p], corr[p] * 0.5f, 0.f)
                                   : make_float4(0.f, 0.f, 0.f, 0.f);
            s_c[threadIdx.x] = p < npix ? coords[p] : make_float2(0.f, 0.f);
        }
        __syncthreads();
        const int np = min(kDirPix, full_npix - p0);
#pragma unroll
        for (int u = 0; u < kDirMaxTPerThread; ++u) {
            if (threadIdx.x + 256 * u >= T) break;
            for (int pl = 0; pl < np; ++pl) {
                const float4 xv = s_x[pl];
                const float2 xy = s_c[pl];
                float s, c;
                sincosf(xy.x * txv[u] + xy.y * tyv[u], &s, &c);
                const float xr = c * xv.x - s * xv.y;
                const float xi = c * xv.y + s * xv.x;
#pragma unroll
                for (int j = 0; j < kDirOrient; ++j) {
                    const float2 pv = s_p[pl][j];
                    const float dr = pv.x - xr, di = pv.y - xi;
                    acc[u][j] = fmaf(dr * dr + di * di, xv.z, acc[u][j]);
                }
            }
        }
    }
#pragma unroll
    for (int u = 0; u < kDirMaxTPerThread; ++u) {
        const int t = threadIdx.x + 256 * u;
        if (t >= T) break;
#pragma unroll
        for (int j = 0; j < kDirOrient; ++j) {
            if (o0 + j < O) {
                const size_t k = (size_t)(o0 + j) * T + t;
                out[k] = (init ? init[k] : 0.f) + acc[u][j];
            }
        }
    }
}

hipError_t launch_direct(const ProjView &v, const float *eulers, int O, const float *trans_x,
                         const float *trans_y, int T, const float2 *coords, const float *real,
                         const float *imag, const float *corr, int npix, int full_npix,
                         const float *init, float *out, hipStream_t s)
{
    if (T > 256 * kDirMaxTPerThread) return hipErrorInvalidValue;
    dim3 grid((O + kDirOrient - 1) / kDirOrient);
    hipLaunchKernelGGL(direct_kernel, grid, dim3(256), 0, s, v, eulers, O, trans_x, trans_y, T,
                       coords, real, imag, corr, npix, full_npix, init, out);
    return hipGetLastError();
}

}  // namespace rc
