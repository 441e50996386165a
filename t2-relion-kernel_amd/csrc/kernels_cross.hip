// Cross-term kernel of the coarse diff2 search on gfx950 MFMA.
//
//   C[m, n] = sum_p Re( P_n(p) * conj(A_m(p)) ),  A_m(p) = corr_p X_p e^{i phi_m(p)}
//   diff2[m, n] = init + sum_p corr_p/2 |P_n|^2 + sum_p corr_p/2 |X|^2 - C[m, n]
//
// Rows m are translations of particles, packed as 16-row groups that never
// straddle a particle; columns n are orientations of one class.  The pixel
// axis is walked in 32-pixel chunks.  Per chunk:
//   * B (projections, 32 pixels x tile_n orientations) is staged in LDS,
//     either copied from the precomputed, pre-encoded projection buffer or
//     (fused variant) gathered from the Fourier volume by trilinear
//     interpolation right here;
//   * every lane builds its A operand in registers: 8 pixels of one
//     translation row, phase factor e^{i phi} from an exactly reduced
//     argument, multiplied into corr * X, encoded for the MFMA form;
//   * 16x16 MFMAs accumulate Re*Re and Im*Im products into FP32.
// Lane l of a wave owns row (l % 16) of a 16-row fragment and the pixel
// group q = l / 16 (pixels 8q .. 8q+7 of the chunk); the same lane reads
// column (l % 16) / pixel group q of B, which is exactly the k-slot pairing
// of v_mfma_f32_16x16x4_f32 (k = q, one MFMA per pixel j of the group) and
// of v_mfma_f32_16x16x32_{bf16,f16} (k = 8q + j).
// The epilogue assembles diff2 in float64, stores it, and reduces a
// float64 (min, argmin) partial per (row group, column tile, wave column).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "device_common.hpp"

namespace rc {

constexpr double kInv2Pi = 0.15915494309189533577;

// Score of one (row, column) in float64 from the cross term c = sum w
// Re(P conj(Xs)) and the norms pn = sum w/2 |P|^2, xn = sum w/2 |X|^2:
//   diff2:  init + pn + xn - c   (coarse_splitImg_BCF_PMO_TMDB_NFUSE.cuh:673-685)
//   CC:     -c / sqrt(2 pn), written, not accumulated (diff2.cuh:797-805)
__device__ __forceinline__ double assemble(const CrossArgs &a, float pn, double xn, float c,
                                           size_t k)
{
    if (a.cc) return -(double)c / sqrt(2.0 * (double)pn);
    double d = (double)pn + xn - (double)c;
    if (a.diff2_init) d += (double)a.diff2_init[k];
    return d;
}

template <int FORM, int WM, int WN, int FM, int FN, bool FUSED>
struct CrossCfg {
    static constexpr int kThreads = 64 * WM * WN;
    static constexpr int kTileM = WM * FM * 16;
    static constexpr int kTileN = WN * FN * 16;
    static constexpr int kGB = group_bytes(FORM);
    static constexpr int kCB = 4 * kGB;
    static constexpr int kRow = kCB + 16;          // LDS row stride (bank spread)
    static constexpr int kLds = kTileN * kRow;
};

template <int FORM, int WM, int WN, int FM, int FN, bool FUSED>
__global__ __launch_bounds__(64 * WM * WN) void cross_kernel(const CrossArgs a)
{
    using Cfg = CrossCfg<FORM, WM, WN, FM, FN, FUSED>;
    constexpr int NT = Cfg::kThreads;
    constexpr int TM = Cfg::kTileM;
    constexpr int TN = Cfg::kTileN;
    constexpr int GB = Cfg::kGB;
    constexpr int CB = Cfg::kCB;
    constexpr int ROW = Cfg::kRow;

    __shared__ __attribute__((aligned(16))) unsigned char sB[Cfg::kLds];
    __shared__ float sE[FUSED ? TN * 6 : 1];
    __shared__ float sPn[FUSED ? TN : 1];

    // ---- tile of this block: XCD-aware (blocks b, b+8, ... share an XCD and
    // get consecutive tiles, which share the column tile and so its B data)
    const int n_ntiles = a.n_classes * a.ntiles_per_class;
    const int total = a.n_mtiles * n_ntiles;
    const int per_xcd = (total + 7) >> 3;
    const int L = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (L >= total) return;
    const int ntile = L / a.n_mtiles;
    const int mtile = L - ntile * a.n_mtiles;
    const int cls = ntile / a.ntiles_per_class;
    const int o0 = (ntile - cls * a.ntiles_per_class) * TN;
    const int O = a.orientation_num;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WM;
    const int wn = wave / WM;
    const int r16 = lane & 15;
    const int q = lane >> 4;
    const int nkc = a.kpad / kChunk;
    // split-K (fused kernels): slice blockIdx.y covers chunks [kc0, kc1)
    int kc0 = 0, kc1 = nkc;
    if constexpr (FUSED) {
        if (a.ksplit > 1) {
            kc0 = (int)blockIdx.y * a.kc_per_split;
            kc1 = min(nkc, kc0 + a.kc_per_split);
        }
    }

    // ---- rows owned by this lane (one per fragment)
    int g_part[FM], g_t0[FM], g_T[FM], g_base[FM];
    bool row_ok[FM];
    double txr[FM], tyr[FM];
    float asc[FM];
#pragma unroll
    for (int fa = 0; fa < FM; ++fa) {
        const int g = mtile * (TM / 16) + wm * FM + fa;
        int part = 0, t0 = 0, T = 0, base = 0;
        if (g < a.n_groups) {
            if (a.groups) {
                const int2 gi = a.groups[g];
                part = gi.x;
                t0 = gi.y;
            } else {
                t0 = 16 * g;
            }
            if (a.trans_offset) {
                base = a.trans_offset[part];
                T = a.trans_offset[part + 1] - base;
            } else {
                T = a.n_trans;
            }
        }
        g_part[fa] = part;
        g_t0[fa] = t0;
        g_T[fa] = T;
        g_base[fa] = base;
        const int t = t0 + r16;
        row_ok[fa] = t < T;
        txr[fa] = row_ok[fa] ? (double)a.trans_x[base + t] * kInv2Pi : 0.0;
        tyr[fa] = row_ok[fa] ? (double)a.trans_y[base + t] * kInv2Pi : 0.0;
        asc[fa] = (FORM == kFP16) ? a.ascale[part] : 1.f;
    }

    // ---- accumulators
    floatx4 acc[FM][FN];
#pragma unroll
    for (int fa = 0; fa < FM; ++fa)
#pragma unroll
        for (int fb = 0; fb < FN; ++fb) acc[fa][fb] = floatx4{0.f, 0.f, 0.f, 0.f};

    // ---- B staging
    constexpr int UNITS = TN * (CB / 16);
    constexpr int UPT = (UNITS + NT - 1) / NT;
    constexpr int SPT = (TN * kChunk) / NT;            // fused: samples per thread
    uint4 pre[FUSED ? 1 : UPT];
    float pn_acc[FUSED ? SPT : 1];
    const float bsc = (FORM == kFP16) ? a.bscale[cls] : 1.f;

    if constexpr (FUSED) {
        for (int u = tid; u < TN * 6; u += NT) {
            const int col = u / 6, k = u % 6;
            const int src = (k >> 1) * 3 + (k & 1);   // 0,1,3,4,6,7
            sE[u] = (o0 + col < O) ? a.eulers[9 * (size_t)(o0 + col) + src] : 0.f;
        }
#pragma unroll
        for (int s = 0; s < SPT; ++s) pn_acc[s] = 0.f;
    } else {
        const unsigned char *src = a.proj + (size_t)cls * O * nkc * CB;
#pragma unroll
        for (int u = 0; u < UPT; ++u) {
            const int idx = tid + u * NT;
            const int col = idx / (CB / 16), seg = idx % (CB / 16);
            pre[u] = make_uint4(0, 0, 0, 0);
            if (idx < UNITS && o0 + col < O)
                pre[u] = *reinterpret_cast<const uint4 *>(
                    src + ((size_t)(o0 + col) * nkc + 0) * CB + seg * 16);
        }
    }

    // Fused kernels with cached projections (a.pcache): a chunk's values --
    // thread tid holds pixel tid % 32 of columns tid / 32 + 8 s -- and its
    // pixel weight are loaded a chunk ahead, under the previous chunk's MFMAs.
    static_assert(!FUSED || NT % kChunk == 0, "one pixel per thread in every chunk");
    float2 pv[FUSED ? SPT : 1];
    float hcv = 0.f;
    auto load_cached = [&](int kc) {
        const int pl = tid & (kChunk - 1);
        const int p = kc * kChunk + pl;
        hcv = a.half_corr[p];
#pragma unroll
        for (int s = 0; s < SPT; ++s) {
            const int col = (tid + s * NT) >> 5;
            pv[s] = (o0 + col < O && p < a.npix) ? a.pcache[(size_t)(o0 + col) * a.kpad + p]
                                                 : make_float2(0.f, 0.f);
        }
    };
    if constexpr (FUSED) {
        if (a.pcache && kc0 < kc1) load_cached(kc0);
    }

    for (int kc = kc0; kc < kc1; ++kc) {
        __syncthreads();
        if constexpr (FUSED) {
            if (a.pcache) {
#pragma unroll
                for (int s = 0; s < SPT; ++s) {
                    const int idx = tid + s * NT;
                    const int col = idx >> 5, pl = idx & 31;
                    const float2 val = pv[s];
                    pn_acc[s] = fmaf(hcv, val.x * val.x + val.y * val.y, pn_acc[s]);
                    store_encoded<FORM>(sB + col * ROW + (pl >> 3) * GB, pl & 7, val.x, val.y,
                                        bsc);
                }
            } else {
                const ProjView vol = a.vol[0];
#pragma unroll 2
                for (int s = 0; s < SPT; ++s) {
                    const int idx = tid + s * NT;
                    const int col = idx >> 5, pl = idx & 31;
                    const int p = kc * kChunk + pl;
                    float2 val = make_float2(0.f, 0.f);
                    if (o0 + col < O && p < a.npix) {
                        const float2 xy = a.coords[p];
                        const float *e = sE + col * 6;
                        val = project_sample(vol, e[0], e[1], e[2], e[3], e[4], e[5], xy.x, xy.y);
                    }
                    pn_acc[s] = fmaf(a.half_corr[p], val.x * val.x + val.y * val.y, pn_acc[s]);
                    store_encoded<FORM>(sB + col * ROW + (pl >> 3) * GB, pl & 7, val.x, val.y,
                                        bsc);
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < UPT; ++u) {
                const int idx = tid + u * NT;
                if (idx < UNITS) {
                    const int col = idx / (CB / 16), seg = idx % (CB / 16);
                    *reinterpret_cast<uint4 *>(sB + col * ROW + seg * 16) = pre[u];
                }
            }
        }
        __syncthreads();
        if constexpr (FUSED) {
            if (a.pcache && kc + 1 < kc1) load_cached(kc + 1);
        }
        if constexpr (!FUSED) {
            if (kc + 1 < nkc) {
                const unsigned char *src = a.proj + (size_t)cls * O * nkc * CB;
#pragma unroll
                for (int u = 0; u < UPT; ++u) {
                    const int idx = tid + u * NT;
                    const int col = idx / (CB / 16), seg = idx % (CB / 16);
                    pre[u] = make_uint4(0, 0, 0, 0);
                    if (idx < UNITS && o0 + col < O)
                        pre[u] = *reinterpret_cast<const uint4 *>(
                            src + ((size_t)(o0 + col) * nkc + kc + 1) * CB + seg * 16);
                }
            }
        }

        // ---- A operands in registers
        const int p0 = kc * kChunk + q * kGroupPix;
        const float4 *cp = reinterpret_cast<const float4 *>(a.coords + p0);
        const float4 c01 = cp[0], c23 = cp[1], c45 = cp[2], c67 = cp[3];
        const float cx[8] = {c01.x, c01.z, c23.x, c23.z, c45.x, c45.z, c67.x, c67.z};
        const float cy[8] = {c01.y, c01.w, c23.y, c23.w, c45.y, c45.w, c67.y, c67.w};

        float aRf[FORM == kFP32 ? FM : 1][8], aIf[FORM == kFP32 ? FM : 1][8];
        bf16x8 aRh[FORM == kBF16X3 ? FM : 1], aRl[FORM == kBF16X3 ? FM : 1];
        bf16x8 aIh[FORM == kBF16X3 ? FM : 1], aIl[FORM == kBF16X3 ? FM : 1];
        halfx8 aRs[FORM == kFP16 ? FM : 1], aIs[FORM == kFP16 ? FM : 1];
#pragma unroll
        for (int fa = 0; fa < FM; ++fa) {
            const float4 *xp =
                reinterpret_cast<const float4 *>(a.xw + (size_t)g_part[fa] * a.kpad + p0);
            const float4 x01 = xp[0], x23 = xp[1], x45 = xp[2], x67 = xp[3];
            const float xr[8] = {x01.x, x01.z, x23.x, x23.z, x45.x, x45.z, x67.x, x67.z};
            const float xi[8] = {x01.y, x01.w, x23.y, x23.w, x45.y, x45.w, x67.y, x67.w};
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                float s, c;
                phase_factor(cx[j], cy[j], txr[fa], tyr[fa], s, c);
                float ar = c * xr[j] - s * xi[j];
                float ai = c * xi[j] + s * xr[j];
                if (!row_ok[fa]) {
                    ar = 0.f;
                    ai = 0.f;
                }
                if constexpr (FORM == kFP32) {
                    aRf[fa][j] = ar;
                    aIf[fa][j] = ai;
                } else if constexpr (FORM == kBF16X3) {
                    __bf16 h, l;
                    split_bf16(ar, h, l);
                    aRh[fa][j] = h;
                    aRl[fa][j] = l;
                    split_bf16(ai, h, l);
                    aIh[fa][j] = h;
                    aIl[fa][j] = l;
                } else {
                    aRs[fa][j] = (_Float16)(ar * asc[fa]);
                    aIs[fa][j] = (_Float16)(ai * asc[fa]);
                }
            }
        }

        // ---- MFMAs over the column fragments
#pragma unroll
        for (int fb = 0; fb < FN; ++fb) {
            const int col = wn * FN * 16 + fb * 16 + r16;
            const unsigned char *bp = sB + col * ROW + q * GB;
            if constexpr (FORM == kFP32) {
                const float4 b0 = reinterpret_cast<const float4 *>(bp)[0];
                const float4 b1 = reinterpret_cast<const float4 *>(bp)[1];
                const float4 b2 = reinterpret_cast<const float4 *>(bp)[2];
                const float4 b3 = reinterpret_cast<const float4 *>(bp)[3];
                const float bR[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
                const float bI[8] = {b2.x, b2.y, b2.z, b2.w, b3.x, b3.y, b3.z, b3.w};
#pragma unroll
                for (int fa = 0; fa < FM; ++fa) {
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        acc[fa][fb] = __builtin_amdgcn_mfma_f32_16x16x4f32(aRf[fa][j], bR[j],
                                                                           acc[fa][fb], 0, 0, 0);
                        acc[fa][fb] = __builtin_amdgcn_mfma_f32_16x16x4f32(aIf[fa][j], bI[j],
                                                                           acc[fa][fb], 0, 0, 0);
                    }
                }
            } else if constexpr (FORM == kBF16X3) {
                const bf16x8 bRh = reinterpret_cast<const bf16x8 *>(bp)[0];
                const bf16x8 bRl = reinterpret_cast<const bf16x8 *>(bp)[1];
                const bf16x8 bIh = reinterpret_cast<const bf16x8 *>(bp)[2];
                const bf16x8 bIl = reinterpret_cast<const bf16x8 *>(bp)[3];
#pragma unroll
                for (int fa = 0; fa < FM; ++fa) {
                    floatx4 c = acc[fa][fb];
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aRl[fa], bRh, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aRh[fa], bRl, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aIl[fa], bIh, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aIh[fa], bIl, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aRh[fa], bRh, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aIh[fa], bIh, c, 0, 0, 0);
                    acc[fa][fb] = c;
                }
            } else {
                const halfx8 bR = reinterpret_cast<const halfx8 *>(bp)[0];
                const halfx8 bI = reinterpret_cast<const halfx8 *>(bp)[1];
#pragma unroll
                for (int fa = 0; fa < FM; ++fa) {
                    acc[fa][fb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aRs[fa], bR,
                                                                         acc[fa][fb], 0, 0, 0);
                    acc[fa][fb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aIs[fa], bI,
                                                                         acc[fa][fb], 0, 0, 0);
                }
            }
        }
    }

    // ---- fused variant: per-column |P|^2 norms from the gather partials
    if constexpr (FUSED) {
        __syncthreads();
#pragma unroll
        for (int s = 0; s < SPT; ++s) {
            float v = pn_acc[s];
            v += __shfl_xor(v, 16, 32);
            v += __shfl_xor(v, 8, 32);
            v += __shfl_xor(v, 4, 32);
            v += __shfl_xor(v, 2, 32);
            v += __shfl_xor(v, 1, 32);
            if ((tid & 31) == 0) sPn[(tid + s * NT) >> 5] = v;
        }
        __syncthreads();
        // split-K: this slice's partial sums, in true units (the FP16 scales
        // divided out), for the fixed-order float64 assembly pass
        if (a.part_c) {
            const size_t slice = (size_t)blockIdx.y;
            for (int c = tid; c < TN; c += NT)
                if (o0 + c < O) a.part_pn[slice * O + o0 + c] = sPn[c];
#pragma unroll
            for (int fa = 0; fa < FM; ++fa) {
                const int T = g_T[fa];
                const float inv = (FORM == kFP16) ? 1.f / (asc[fa] * bsc) : 1.f;
#pragma unroll
                for (int fb = 0; fb < FN; ++fb) {
                    const int o = o0 + wn * FN * 16 + fb * 16 + r16;
                    if (o >= O) continue;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int t = g_t0[fa] + 4 * q + v;
                        if (t < T) a.part_c[(slice * O + o) * T + t] = acc[fa][fb][v] * inv;
                    }
                }
            }
            return;
        }
    }

    // ---- epilogue
    const int C = a.n_classes;
#pragma unroll
    for (int fa = 0; fa < FM; ++fa) {
        const int part = g_part[fa];
        const int T = g_T[fa];
        const size_t out_base = (size_t)C * O * g_base[fa] + (size_t)cls * O * T;
        const double xn = a.xnorm[part];
        const float inv = (FORM == kFP16) ? 1.f / (asc[fa] * bsc) : 1.f;
        double best = DBL_MAX;
        long long best_k = 0x7fffffffffffffffLL;
#pragma unroll
        for (int fb = 0; fb < FN; ++fb) {
            const int col = wn * FN * 16 + fb * 16 + r16;
            const int o = o0 + col;
            if (o >= O) continue;
            float pn;
            if constexpr (FUSED) {
                pn = sPn[col];
            } else {
                pn = a.pnorm[((size_t)part * C + cls) * O + o];
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int t = g_t0[fa] + 4 * q + v;
                if (t >= T) continue;
                const size_t k = out_base + (size_t)o * T + t;
                const double d = assemble(a, pn, xn, acc[fa][fb][v] * inv, k);
                a.diff2[k] = (float)d;
                const long long flat = (long long)(cls * O + o) * T + t;
                if (d < best || (d == best && flat < best_k)) {
                    best = d;
                    best_k = flat;
                }
            }
        }
        if (a.part_min) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off, 64);
                const long long ok = __shfl_xor(best_k, off, 64);
                if (ob < best || (ob == best && ok < best_k)) {
                    best = ob;
                    best_k = ok;
                }
            }
            const int g = mtile * (TM / 16) + wm * FM + fa;
            if (lane == 0 && g < a.n_groups) {
                const size_t u = ((size_t)g * n_ntiles + ntile) * WN + wn;
                a.part_min[u] = best;
                a.part_idx[u] = best_k;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Pipelined variant (precomputed projections).
//
// B chunks go global -> LDS by LDS-DMA (global_load_lds_dwordx4: each lane
// names its own 16-byte source, the wave's 64 segments land contiguously at
// an LDS base) into two buffers, so the copy of chunk k+1 runs under the
// MFMAs of chunk k and one barrier per chunk suffices.  Inside a column's
// CB-byte LDS row the 16-byte segment s sits at position s ^ (col % SEGS):
// the 16 lanes that read the same segment of 16 consecutive columns then
// hit 16 different bank groups.  The A operands of chunk k+1 are generated
// while the MFMAs of chunk k issue (an even share of the values per column
// fragment), so the VALU work of the phase factors overlaps the matrix core.
// ---------------------------------------------------------------------------
template <int FORM, int FM>
struct AOps {
    float rf[FORM == kFP32 ? FM : 1][8], if_[FORM == kFP32 ? FM : 1][8];
    bf16x8 rh[FORM == kBF16X3 ? FM : 1], rl[FORM == kBF16X3 ? FM : 1];
    bf16x8 ih[FORM == kBF16X3 ? FM : 1], il[FORM == kBF16X3 ? FM : 1];
    halfx8 rs[FORM == kFP16 ? FM : 1], is[FORM == kFP16 ? FM : 1];
};

template <int FORM, int FM, int FA, int J>
__device__ __forceinline__ void a_store(AOps<FORM, FM> &A, float ar, float ai, float sc)
{
    if constexpr (FORM == kFP32) {
        A.rf[FA][J] = ar;
        A.if_[FA][J] = ai;
    } else if constexpr (FORM == kBF16X3) {
        __bf16 h, l;
        split_bf16(ar, h, l);
        A.rh[FA][J] = h;
        A.rl[FA][J] = l;
        split_bf16(ai, h, l);
        A.ih[FA][J] = h;
        A.il[FA][J] = l;
    } else {
        A.rs[FA][J] = (_Float16)(ar * sc);
        A.is[FA][J] = (_Float16)(ai * sc);
    }
}

// Generate A values V0 .. V0+COUNT-1 (value v = fragment v / 8, pixel v % 8).
template <int FORM, int FM, int V0, int COUNT>
__device__ __forceinline__ void a_gen(AOps<FORM, FM> &A, const float4 (&xw)[FM][4],
                                      const float4 (&cxy)[4], const double (&txr)[FM],
                                      const double (&tyr)[FM], const bool (&ok)[FM],
                                      const float (&asc)[FM])
{
    if constexpr (COUNT > 0 && V0 < FM * 8) {
        constexpr int FA = V0 / 8, J = V0 % 8;
        const float4 c = cxy[J >> 1];
        const float x = (J & 1) ? c.z : c.x, y = (J & 1) ? c.w : c.y;
        const float4 w = xw[FA][J >> 1];
        const float xr = (J & 1) ? w.z : w.x, xi = (J & 1) ? w.w : w.y;
        float s, co;
        phase_factor(x, y, txr[FA], tyr[FA], s, co);
        float ar = co * xr - s * xi;
        float ai = co * xi + s * xr;
        if (!ok[FA]) {
            ar = 0.f;
            ai = 0.f;
        }
        a_store<FORM, FM, FA, J>(A, ar, ai, asc[FA]);
        a_gen<FORM, FM, V0 + 1, COUNT - 1>(A, xw, cxy, txr, tyr, ok, asc);
    }
}

template <int FORM, int FM, int FB, int FN>
__device__ __forceinline__ void mfma_frag(floatx4 (&acc)[FM][FN], const AOps<FORM, FM> &A,
                                          const unsigned char *__restrict__ bp)
{
    if constexpr (FORM == kFP32) {
        const float4 b0 = reinterpret_cast<const float4 *>(bp)[0];
        const float4 b1 = reinterpret_cast<const float4 *>(bp)[1];
        const float4 b2 = reinterpret_cast<const float4 *>(bp)[2];
        const float4 b3 = reinterpret_cast<const float4 *>(bp)[3];
        const float bR[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
        const float bI[8] = {b2.x, b2.y, b2.z, b2.w, b3.x, b3.y, b3.z, b3.w};
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int fa = 0; fa < FM; ++fa) {
                acc[fa][FB] = __builtin_amdgcn_mfma_f32_16x16x4f32(A.rf[fa][j], bR[j],
                                                                   acc[fa][FB], 0, 0, 0);
                acc[fa][FB] = __builtin_amdgcn_mfma_f32_16x16x4f32(A.if_[fa][j], bI[j],
                                                                   acc[fa][FB], 0, 0, 0);
            }
    } else if constexpr (FORM == kBF16X3) {
        const bf16x8 bRh = reinterpret_cast<const bf16x8 *>(bp)[0];
        const bf16x8 bRl = reinterpret_cast<const bf16x8 *>(bp)[1];
        const bf16x8 bIh = reinterpret_cast<const bf16x8 *>(bp)[2];
        const bf16x8 bIl = reinterpret_cast<const bf16x8 *>(bp)[3];
#pragma unroll
        for (int fa = 0; fa < FM; ++fa)
            acc[fa][FB] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.rl[fa], bRh, acc[fa][FB], 0, 0, 0);
#pragma unroll
        for (int fa = 0; fa < FM; ++fa)
            acc[fa][FB] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.rh[fa], bRl, acc[fa][FB], 0, 0, 0);
#pragma unroll
        for (int fa = 0; fa < FM; ++fa)
            acc[fa][FB] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.il[fa], bIh, acc[fa][FB], 0, 0, 0);
#pragma unroll
        for (int fa = 0; fa < FM; ++fa)
            acc[fa][FB] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.ih[fa], bIl, acc[fa][FB], 0, 0, 0);
#pragma unroll
        for (int fa = 0; fa < FM; ++fa)
            acc[fa][FB] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.rh[fa], bRh, acc[fa][FB], 0, 0, 0);
#pragma unroll
        for (int fa = 0; fa < FM; ++fa)
            acc[fa][FB] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.ih[fa], bIh, acc[fa][FB], 0, 0, 0);
    } else {
        const halfx8 bR = reinterpret_cast<const halfx8 *>(bp)[0];
        const halfx8 bI = reinterpret_cast<const halfx8 *>(bp)[1];
#pragma unroll
        for (int fa = 0; fa < FM; ++fa) {
            acc[fa][FB] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.rs[fa], bR, acc[fa][FB], 0, 0, 0);
            acc[fa][FB] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.is[fa], bI, acc[fa][FB], 0, 0, 0);
        }
    }
}

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) const void gbl_void_t;

template <int FORM, int WM, int WN, int FM, int FN>
struct PipeCfg {
    static constexpr int kThreads = 64 * WM * WN;
    static constexpr int kTileM = WM * FM * 16;
    static constexpr int kTileN = WN * FN * 16;
    static constexpr int kCB = chunk_bytes(FORM);
    static constexpr int kSegs = kCB / 16;
    static constexpr int kBuf = kTileN * kCB;
    static constexpr int kLds = 2 * kBuf;
};

// One column fragment FB of the chunk: read B from LDS, issue the MFMAs,
// then generate this fragment's share of the next chunk's A values.
template <int FORM, int FM, int FN, int FB, int SEGS, int CB>
__device__ __forceinline__ void pipe_step(floatx4 (&acc)[FM][FN], const AOps<FORM, FM> &A,
                                          AOps<FORM, FM> &An, const unsigned char *buf,
                                          int col0, int r16, int q, bool more,
                                          const float4 (&xw)[FM][4], const float4 (&cxy)[4],
                                          const double (&txr)[FM], const double (&tyr)[FM],
                                          const bool (&ok)[FM], const float (&asc)[FM])
{
    if constexpr (FB < FN) {
        const int col = col0 + FB * 16 + r16;
        // segments of pixel group q, de-swizzled: FP32/BF16X3 4 per group, FP16 2
        constexpr int GS = SEGS / 4;
        __attribute__((aligned(16))) unsigned char tmp[4 * 16];
#pragma unroll
        for (int u = 0; u < GS; ++u) {
            const int pos = (q * GS + u) ^ (col & (SEGS - 1));
            *reinterpret_cast<uint4 *>(tmp + 16 * u) =
                *reinterpret_cast<const uint4 *>(buf + col * CB + pos * 16);
        }
        mfma_frag<FORM, FM, FB, FN>(acc, A, tmp);
        constexpr int PER = (FM * 8 + FN - 1) / FN;
        a_gen<FORM, FM, FB * PER, PER>(An, xw, cxy, txr, tyr, ok, asc);
        pipe_step<FORM, FM, FN, FB + 1, SEGS, CB>(acc, A, An, buf, col0, r16, q, more, xw, cxy,
                                                   txr, tyr, ok, asc);
    }
}

// One (row tile, column tile) of the batched product; L indexes tiles with
// the row tile fastest.
template <int FORM, int WM, int WN, int FM, int FN>
__device__ __forceinline__ void pipe_tile(const CrossArgs &a, int L, unsigned char *sB)
{
    using Cfg = PipeCfg<FORM, WM, WN, FM, FN>;
    constexpr int NT = Cfg::kThreads;
    constexpr int TM = Cfg::kTileM;
    constexpr int TN = Cfg::kTileN;
    constexpr int CB = Cfg::kCB;
    constexpr int SEGS = Cfg::kSegs;
    constexpr int BUF = Cfg::kBuf;
    constexpr int UNITS = TN * SEGS;
    static_assert(UNITS % NT == 0, "tile must split evenly into 16-byte loads");
    constexpr int LPT = UNITS / NT;

    const int n_ntiles = a.n_classes * a.ntiles_per_class;
    const int ntile = L / a.n_mtiles;
    const int mtile = L - ntile * a.n_mtiles;
    const int cls = ntile / a.ntiles_per_class;
    const int o0 = (ntile - cls * a.ntiles_per_class) * TN;
    const int O = a.orientation_num;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WM;
    const int wn = wave / WM;
    const int r16 = lane & 15;
    const int q = lane >> 4;
    const int nkc = a.kpad / kChunk;

    int g_part[FM], g_t0[FM], g_T[FM], g_base[FM];
    bool row_ok[FM];
    double txr[FM], tyr[FM];
    float asc[FM];
    const float2 *xw_row[FM];
#pragma unroll
    for (int fa = 0; fa < FM; ++fa) {
        const int g = mtile * (TM / 16) + wm * FM + fa;
        int part = 0, t0 = 0, T = 0, base = 0;
        if (g < a.n_groups) {
            if (a.groups) {
                const int2 gi = a.groups[g];
                part = gi.x;
                t0 = gi.y;
            } else {
                t0 = 16 * g;
            }
            if (a.trans_offset) {
                base = a.trans_offset[part];
                T = a.trans_offset[part + 1] - base;
            } else {
                T = a.n_trans;
            }
        }
        g_part[fa] = part;
        g_t0[fa] = t0;
        g_T[fa] = T;
        g_base[fa] = base;
        const int t = t0 + r16;
        row_ok[fa] = t < T;
        txr[fa] = row_ok[fa] ? (double)a.trans_x[base + t] * kInv2Pi : 0.0;
        tyr[fa] = row_ok[fa] ? (double)a.trans_y[base + t] * kInv2Pi : 0.0;
        asc[fa] = (FORM == kFP16) ? a.ascale[part] : 1.f;
        xw_row[fa] = a.xw + (size_t)part * a.kpad + q * kGroupPix;
    }
    const float2 *c_row = a.coords + q * kGroupPix;

    floatx4 acc[FM][FN];
#pragma unroll
    for (int fa = 0; fa < FM; ++fa)
#pragma unroll
        for (int fb = 0; fb < FN; ++fb) acc[fa][fb] = floatx4{0.f, 0.f, 0.f, 0.f};

    const unsigned char *cls_proj = a.proj + (size_t)cls * O * nkc * CB;
    auto issue_B = [&](int kc, int buf) {
#pragma unroll
        for (int it = 0; it < LPT; ++it) {
            const int u = it * NT + tid;
            const int col = u / SEGS, pos = u % SEGS;
            const int seg = pos ^ (col & (SEGS - 1));
            const int o = min(o0 + col, O - 1);
            const unsigned char *src = cls_proj + ((size_t)o * nkc + kc) * CB + seg * 16;
            unsigned char *dst = sB + buf * BUF + (it * NT + wave * 64) * 16;
            __builtin_amdgcn_global_load_lds((gbl_void_t *)src, (lds_void_t *)dst, 16, 0, 0);
        }
    };
    auto load_a_src = [&](int kc, float4 (&xw)[FM][4], float4 (&cxy)[4]) {
        const int p0 = kc * kChunk;
#pragma unroll
        for (int fa = 0; fa < FM; ++fa) {
            const float4 *xp = reinterpret_cast<const float4 *>(xw_row[fa] + p0);
#pragma unroll
            for (int u = 0; u < 4; ++u) xw[fa][u] = xp[u];
        }
        const float4 *cp = reinterpret_cast<const float4 *>(c_row + p0);
#pragma unroll
        for (int u = 0; u < 4; ++u) cxy[u] = cp[u];
    };

    AOps<FORM, FM> A, An;
    float4 xw[FM][4], cxy[4];
    issue_B(0, 0);
    load_a_src(0, xw, cxy);
    a_gen<FORM, FM, 0, FM * 8>(A, xw, cxy, txr, tyr, row_ok, asc);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();

    for (int kc = 0; kc < nkc; ++kc) {
        const int buf = kc & 1;
        const bool more = kc + 1 < nkc;
        if (more) issue_B(kc + 1, buf ^ 1);
        // the last chunk regenerates its own A (unused) so the MFMA stream
        // and the VALU stream stay in one basic block
        load_a_src(more ? kc + 1 : kc, xw, cxy);
        pipe_step<FORM, FM, FN, 0, SEGS, CB>(acc, A, An, sB + buf * BUF, wn * FN * 16, r16, q,
                                             more, xw, cxy, txr, tyr, row_ok, asc);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        A = An;
    }

    // ---- epilogue (as in the staged kernel)
    const int C = a.n_classes;
#pragma unroll
    for (int fa = 0; fa < FM; ++fa) {
        const int part = g_part[fa];
        const int T = g_T[fa];
        const size_t out_base = (size_t)C * O * g_base[fa] + (size_t)cls * O * T;
        const double xn = a.xnorm[part];
        const float inv = (FORM == kFP16) ? 1.f / (asc[fa] * a.bscale[cls]) : 1.f;
        double best = DBL_MAX;
        long long best_k = 0x7fffffffffffffffLL;
#pragma unroll
        for (int fb = 0; fb < FN; ++fb) {
            const int o = o0 + wn * FN * 16 + fb * 16 + r16;
            if (o >= O) continue;
            const float pn = a.pnorm[((size_t)part * C + cls) * O + o];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int t = g_t0[fa] + 4 * q + v;
                if (t >= T) continue;
                const size_t k = out_base + (size_t)o * T + t;
                const double d = assemble(a, pn, xn, acc[fa][fb][v] * inv, k);
                a.diff2[k] = (float)d;
                const long long flat = (long long)(cls * O + o) * T + t;
                if (d < best || (d == best && flat < best_k)) {
                    best = d;
                    best_k = flat;
                }
            }
        }
        if (a.part_min) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off, 64);
                const long long ok = __shfl_xor(best_k, off, 64);
                if (ob < best || (ob == best && ok < best_k)) {
                    best = ob;
                    best_k = ok;
                }
            }
            const int g = mtile * (TM / 16) + wm * FM + fa;
            if (lane == 0 && g < a.n_groups) {
                const size_t u = ((size_t)g * n_ntiles + ntile) * WN + wn;
                a.part_min[u] = best;
                a.part_idx[u] = best_k;
            }
        }
    }
}

// XCD-aware tile order: the dispatcher places block b on XCD b % 8; XCD x
// owns tiles [x * per_xcd, (x + 1) * per_xcd), so the blocks running together
// on one XCD work on consecutive row tiles of the same column tile and share
// its B data in that XCD's L2.  PERSIST: a grid of a few blocks per CU walks
// its XCD's range with stride (grid / 8).
template <int FORM, int WM, int WN, int FM, int FN, bool PERSIST>
__global__ __launch_bounds__(64 * WM * WN) void cross_kernel_pipe(const CrossArgs a)
{
    using Cfg = PipeCfg<FORM, WM, WN, FM, FN>;
    __shared__ __attribute__((aligned(16))) unsigned char sB[Cfg::kLds];
    const int total = a.n_mtiles * a.n_classes * a.ntiles_per_class;
    const int per_xcd = (total + 7) >> 3;
    const int x = (int)(blockIdx.x & 7), j = (int)(blockIdx.x >> 3);
    if constexpr (!PERSIST) {
        const int L = x * per_xcd + j;
        if (L < total) pipe_tile<FORM, WM, WN, FM, FN>(a, L, sB);
    } else {
        const int stride = (int)(gridDim.x >> 3);
        const int end = min(total, (x + 1) * per_xcd);
        for (int L = x * per_xcd + j; L < end; L += stride) {
            pipe_tile<FORM, WM, WN, FM, FN>(a, L, sB);
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// variant table
// ---------------------------------------------------------------------------
template <int FORM, int WM, int WN, int FM, int FN, bool FUSED>
static hipError_t launch_one(const CrossArgs &a, hipStream_t s)
{
    using Cfg = CrossCfg<FORM, WM, WN, FM, FN, FUSED>;
    const int total = a.n_mtiles * a.n_classes * a.ntiles_per_class;
    const int grid = ((total + 7) / 8) * 8;
    const int slices = FUSED && a.ksplit > 1 ? a.ksplit : 1;
    hipLaunchKernelGGL((cross_kernel<FORM, WM, WN, FM, FN, FUSED>), dim3(grid, slices),
                       dim3(Cfg::kThreads), 0, s, a);
    return hipGetLastError();
}

template <int FORM, int WM, int WN, int FM, int FN, bool PERSIST>
static hipError_t launch_pipe(const CrossArgs &a, hipStream_t s)
{
    using Cfg = PipeCfg<FORM, WM, WN, FM, FN>;
    const int total = a.n_mtiles * a.n_classes * a.ntiles_per_class;
    int grid = ((total + 7) / 8) * 8;
    if (PERSIST) grid = std::min(grid, 8 * ((a.persist_blocks + 7) / 8));
    hipLaunchKernelGGL((cross_kernel_pipe<FORM, WM, WN, FM, FN, PERSIST>), dim3(grid),
                       dim3(Cfg::kThreads), 0, s, a);
    return hipGetLastError();
}

template <int FORM, int WM, int WN, int FM, int FN, bool FUSED>
static VariantInfo info_one()
{
    using Cfg = CrossCfg<FORM, WM, WN, FM, FN, FUSED>;
    return VariantInfo{FORM, Cfg::kTileM, Cfg::kTileN, Cfg::kThreads, WN, FUSED, Cfg::kLds,
                       false};
}

template <int FORM, int WM, int WN, int FM, int FN, bool PERSIST>
static VariantInfo info_pipe()
{
    using Cfg = PipeCfg<FORM, WM, WN, FM, FN>;
    return VariantInfo{FORM,      Cfg::kTileM, Cfg::kTileN, Cfg::kThreads, WN, false,
                       Cfg::kLds, PERSIST,     false,       "cross_kernel_pipe"};
}

#define RC_VARIANT(F, WM, WN, FM, FN, FU) \
    VariantEntry{info_one<F, WM, WN, FM, FN, FU>(), &launch_one<F, WM, WN, FM, FN, FU>}
#define RC_PIPE(F, WM, WN, FM, FN, PE) \
    VariantEntry{info_pipe<F, WM, WN, FM, FN, PE>(), &launch_pipe<F, WM, WN, FM, FN, PE>}

// Variants 0-3: staged (register prefetch, single LDS buffer); 4-7:
// pipelined (LDS-DMA, double buffer, A generation under the MFMAs); 8-9:
// the pipelined kernel as a persistent grid.
#define RC_FORM_VARIANTS(F)                                                           \
    static const VariantEntry k_tab_##F[] = {                                          \
        RC_VARIANT(F, 4, 1, 2, 8, false),  /* 128 x 128 */                             \
        RC_VARIANT(F, 2, 2, 2, 8, false),  /*  64 x 256 */                             \
        RC_VARIANT(F, 4, 1, 2, 16, false), /* 128 x 256 */                             \
        RC_VARIANT(F, 2, 2, 1, 8, false),  /*  32 x 256 */                             \
        RC_PIPE(F, 4, 1, 2, 8, false),     /* 128 x 128 */                             \
        RC_PIPE(F, 4, 1, 2, 16, false),    /* 128 x 256 */                             \
        RC_PIPE(F, 2, 2, 2, 8, false),     /*  64 x 256 */                             \
        RC_PIPE(F, 4, 1, 1, 16, false),    /*  64 x 256 */                             \
        RC_PIPE(F, 4, 1, 2, 8, true),      /* 128 x 128, persistent */                 \
        RC_PIPE(F, 4, 1, 2, 16, true),     /* 128 x 256, persistent */                 \
    };                                                                                 \
    static const VariantEntry k_fused_##F[] = {                                        \
        RC_VARIANT(F, 4, 1, 2, 8, true),   /* 128 x 128 */                             \
        RC_VARIANT(F, 2, 2, 2, 4, true),   /*  64 x 128 */                             \
        RC_VARIANT(F, 1, 4, 2, 4, true),   /*  32 x 256 */                             \
    };

RC_FORM_VARIANTS(kFP32)
RC_FORM_VARIANTS(kBF16X3)
RC_FORM_VARIANTS(kFP16)

static const VariantEntry *table(int form, bool fused, int *n)
{
    constexpr int nt = sizeof(k_tab_kFP32) / sizeof(k_tab_kFP32[0]);
    constexpr int nf = sizeof(k_fused_kFP32) / sizeof(k_fused_kFP32[0]);
    *n = fused ? nf : nt;
    switch (form) {
    case kFP32: return fused ? k_fused_kFP32 : k_tab_kFP32;
    case kBF16X3: return fused ? k_fused_kBF16X3 : k_tab_kBF16X3;
    case kFP16: return fused ? k_fused_kFP16 : k_tab_kFP16;
    default:
        *n = 0;
        return nullptr;
    }
}

// Non-fused indices continue past the table with the wide kernels.
static const VariantEntry *entry(int form, bool fused, int idx)
{
    int n = 0;
    const VariantEntry *t = table(form, fused, &n);
    if (!t || idx < 0) return nullptr;
    if (idx < n) return &t[idx];
    if (fused) return nullptr;
    int nw = 0, nd = 0;
    const VariantEntry *w = wide_table(form, &nw);
    if (idx - n < nw) return &w[idx - n];
    const VariantEntry *d = direct_table(form, &nd);
    return idx - n - nw < nd ? &d[idx - n - nw] : nullptr;
}

int variant_count(int form, bool fused)
{
    int n = 0, nw = 0, nd = 0;
    table(form, fused, &n);
    if (!fused && n) {
        wide_table(form, &nw);
        direct_table(form, &nd);
    }
    return n + nw + nd;
}

const VariantInfo *variant_info(int form, bool fused, int idx)
{
    const VariantEntry *e = entry(form, fused, idx);
    return e ? &e->info : nullptr;
}

hipError_t launch_cross(int form, bool fused, int idx, const CrossArgs &a, hipStream_t s)
{
    const VariantEntry *e = entry(form, fused, idx);
    return e ? e->launch(a, s) : hipErrorInvalidValue;
}

}  // namespace rc
