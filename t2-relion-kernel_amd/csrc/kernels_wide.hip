// "Wide" cross-term kernel: both MFMA operands staged in LDS, the
// translation operand built once per block by all its threads.
//
// The pixel axis is walked in 16-pixel chunks whose real and imaginary parts
// are interleaved along the MFMA k axis (k = 2 * pixel + component), so one
// 16x16x32 BF16/FP16 MFMA (or eight 16x16x4 FP32 MFMAs) consumes a whole
// chunk: Re(P conj A) = sum_k P_k A_k.  Per chunk:
//   * B (projections): LDS-DMA from the pre-encoded projection buffer;
//   * A (translated, weighted images): each thread produces 8 (or 4)
//     consecutive pixels of one row -- phase factor from an exactly reduced
//     argument, times corr*X, encoded -- and stores them to LDS;
//   * every wave reads its row and column fragments from LDS and issues the
//     MFMAs.  A for chunk k+1 and the B copy of chunk k+1 are produced while
//     the MFMAs of chunk k run; both buffers are double-buffered, one
//     barrier per chunk.
// Rows are packed: row R is translation R of the batch's CSR translation
// arrays, so particles are not padded to a multiple of 16 rows.
// Row/column r of a buffer holds KB bytes (128: FP32/BF16X3, 64: FP16) as
// 16-byte segments; segment s sits at position s ^ ((r >> FSH) & (SEGS-1)),
// which spreads the 16 rows read by one k-group of a wave over all banks.
//
// Operand encodings of one row/column for one chunk (pixel p, component c):
//   BF16X3: hi[32] bf16 at k = 2p + c (bytes 0..63), lo[32] (bytes 64..127)
//   FP16:   [32] f16 at k = 2p + c, scaled by a power of two
//   FP32:   value v = 2p + c at float (v % 4) * 8 + v / 4, so lane group q
//           reads floats 8q .. 8q+7 and MFMA m pairs k-slot q with v = 4m+q.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <type_traits>

#include "wide_common.hpp"

namespace rc {

// ---------------------------------------------------------------------------
// projection precompute in the wide encoding: block = 16 orientations x
// kProjChunks 16-pixel chunks; each thread samples one orientation at one
// pixel of every chunk, so its orientation rows are loaded once and the
// gathers of its samples are in flight together.  One-dimensional grid,
// XCD-aware: block b runs on XCD b % 8, which takes whole 16-orientation
// groups (b % 8, b % 8 + 8, ...), all their pixel chunks together, so the
// voxels of a group's central sections -- shared by its in-plane (psi)
// rotations -- are gathered through one L2 (142 -> 126 us for 4608
// orientations x 1648 pixels, against spreading a group's chunks over all
// eight XCDs).
// ---------------------------------------------------------------------------
constexpr int kProjChunks = 2;

template <int FORM, bool DIRECT>
__global__ __launch_bounds__(256) void project_wide_kernel(const ProjView *__restrict__ vols,
                                                           const float *__restrict__ bscale,
                                                           const float *__restrict__ eulers,
                                                           int O, int n_classes,
                                                           const float2 *__restrict__ coords,
                                                           int npix, int kpad,
                                                           unsigned char *__restrict__ proj,
                                                           float *__restrict__ pabs2)
{
    constexpr int KB = WideEnc<FORM>::kBytes;
    __shared__ __attribute__((aligned(16))) unsigned char s_out[kProjChunks][16 * KB];
    const int nkc = kpad / 16;
    const int n_pairs = (nkc + kProjChunks - 1) / kProjChunks;   // blocks per group
    const int n_og = (O + 15) / 16;                               // groups per class
    const int jx = (int)(blockIdx.x >> 3);
    const int G = (int)(blockIdx.x & 7) + 8 * (jx / n_pairs);    // (class, group)
    if (G >= n_og * n_classes) return;
    const int bx = jx % n_pairs;
    const int cls = G / n_og;
    const int by = G - cls * n_og;
    const int ol = threadIdx.x >> 4, pl = threadIdx.x & 15;
    const int o = by * 16 + ol;
    const ProjView v = vols[cls];
    float2 val[kProjChunks];
    float e[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (o < O) {
        const float *eo = eulers + 9 * (size_t)o;
        e[0] = eo[0];
        e[1] = eo[1];
        e[2] = eo[3];
        e[3] = eo[4];
        e[4] = eo[6];
        e[5] = eo[7];
    }
#pragma unroll
    for (int h = 0; h < kProjChunks; ++h) {
        const int p = (bx * kProjChunks + h) * 16 + pl;
        val[h] = make_float2(0.f, 0.f);
        if (o < O && p < npix) {
            const float2 xy = coords[p];
            val[h] = project_sample(v, e[0], e[1], e[2], e[3], e[4], e[5], xy.x, xy.y);
        }
    }
    const float sc = (FORM == kFP16) ? bscale[cls] : 1.f;
#pragma unroll
    for (int h = 0; h < kProjChunks; ++h) {
        const int kc = bx * kProjChunks + h;
        wide_store_value<FORM>(s_out[h] + ol * KB, pl, val[h].x, val[h].y, sc);
        if (o < O && kc < nkc)
            pabs2[((size_t)cls * O + o) * kpad + kc * 16 + pl] =
                val[h].x * val[h].x + val[h].y * val[h].y;
    }
    __syncthreads();
    constexpr int units = 16 * KB / 16;
#pragma unroll
    for (int h = 0; h < kProjChunks; ++h) {
        const int kc = bx * kProjChunks + h;
        if (kc >= nkc) break;
        if constexpr (DIRECT) {
            // the whole group, orientations past O included (zeros)
            unsigned char *grp =
                proj + (((size_t)cls * nkc + kc) * n_og + by) * (16 * KB);
            for (int u = threadIdx.x; u < units; u += 256) {
                const int oll = u / (KB / 16), seg = u % (KB / 16);
                *reinterpret_cast<uint4 *>(grp + direct_seg_offset<FORM>(oll, seg)) =
                    reinterpret_cast<const uint4 *>(s_out[h])[u];
            }
        } else {
            for (int u = threadIdx.x; u < units; u += 256) {
                const int oo = by * 16 + u / (KB / 16);
                if (oo < O) {
                    const int seg = u % (KB / 16);
                    const uint4 d = reinterpret_cast<const uint4 *>(s_out[h])[u];
                    reinterpret_cast<uint4 *>(proj + (((size_t)cls * nkc + kc) * O + oo) * KB)[seg] =
                        d;
                }
            }
        }
    }
}

size_t projection_bytes(int form, bool direct, int n_classes, int O, int kpad)
{
    const size_t kb = form == kFP16 ? 64 : 128;   // bytes per orientation per 16-pixel chunk
    const size_t cols = direct ? (size_t)(O + 15) / 16 * 16 : (size_t)O;
    return (size_t)n_classes * cols * (size_t)(kpad / 16) * kb;
}

hipError_t launch_project_wide(int form, const ProjView *vols, const float *bscale, int n_classes,
                               const float *eulers, int O, const float2 *coords, int npix,
                               int kpad, unsigned char *proj, float *pabs2, bool direct,
                               hipStream_t s)
{
    const long long groups = (long long)((O + 15) / 16) * n_classes;
    const long long pairs = (kpad / 16 + kProjChunks - 1) / kProjChunks;
    const dim3 grid((unsigned)(8 * ((groups + 7) / 8) * pairs));
#define RC_PW(F, D)                                                                         \
    hipLaunchKernelGGL((project_wide_kernel<F, D>), grid, dim3(256), 0, s, vols, bscale, eulers, \
                       O, n_classes, coords, npix, kpad, proj, pabs2)
    switch (form) {
    case kFP32: if (direct) RC_PW(kFP32, true); else RC_PW(kFP32, false); break;
    case kBF16X3: if (direct) RC_PW(kBF16X3, true); else RC_PW(kBF16X3, false); break;
    default: if (direct) RC_PW(kFP16, true); else RC_PW(kFP16, false); break;
    }
#undef RC_PW
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// precomputed translation operand (schedule 8): block = one row tile x one
// 16-pixel chunk, built exactly as the in-kernel build does into an LDS image
// of the tile's A buffer, then written out as one contiguous block
// ---------------------------------------------------------------------------
template <int FORM, int TM, int NT>
__global__ __launch_bounds__(NT) void encode_a_kernel(const CrossArgs a,
                                                      unsigned char *__restrict__ out)
{
    constexpr int KB = WideEnc<FORM>::kBytes;
    constexpr int TPR = NT / TM;
    constexpr int PPT = 16 / TPR;
    static_assert(TPR * TM == NT && (TPR == 2 || TPR == 4), "2 or 4 threads per A row");
    __shared__ __attribute__((aligned(16))) unsigned char s_tile[TM * KB];
    const int kc = blockIdx.x, mtile = blockIdx.y;
    const int tid = threadIdx.x;
    const int gr = tid / TPR, gh = tid % TPR;
    const ARow ar = make_arow<FORM, PPT>(a, mtile * TM + gr, gh);
    float4 xw[PPT / 2], cxy[PPT / 2];
    const float4 *xp = reinterpret_cast<const float4 *>(ar.xw + kc * 16);
    const float4 *cp = reinterpret_cast<const float4 *>(a.coords + PPT * gh + kc * 16);
#pragma unroll
    for (int u = 0; u < PPT / 2; ++u) {
        xw[u] = xp[u];
        cxy[u] = cp[u];
    }
    a_build<FORM, PPT>(ar, xw, cxy, gh, gr, &s_tile[gr * KB]);
    __syncthreads();
    uint4 *dst = reinterpret_cast<uint4 *>(out + ((size_t)mtile * a.a_chunks + kc) * (TM * KB));
    for (int u = tid; u < TM * KB / 16; u += NT) dst[u] = reinterpret_cast<const uint4 *>(s_tile)[u];
}

hipError_t launch_encode_a(int form, int tile_m, const CrossArgs &a, unsigned char *a_enc,
                           hipStream_t s)
{
    const dim3 grid((unsigned)a.a_chunks, (unsigned)a.n_mtiles);
    if (tile_m != 128 && tile_m != 256) return hipErrorInvalidValue;
#define RC_ENC(F, TM) hipLaunchKernelGGL((encode_a_kernel<F, TM, 512>), grid, dim3(512), 0, s, a, a_enc)
    switch (form) {
    case kFP32: if (tile_m == 128) RC_ENC(kFP32, 128); else RC_ENC(kFP32, 256); break;
    case kBF16X3: if (tile_m == 128) RC_ENC(kBF16X3, 128); else RC_ENC(kBF16X3, 256); break;
    default: if (tile_m == 128) RC_ENC(kFP16, 128); else RC_ENC(kFP16, 256); break;
    }
#undef RC_ENC
    return hipGetLastError();
}

template <int FORM, int WM, int WN, int FM, int FN>
struct WideCfg {
    static constexpr int kThreads = 64 * WM * WN;
    static constexpr int kTileM = WM * FM * 16;
    static constexpr int kTileN = WN * FN * 16;
    static constexpr int kKB = WideEnc<FORM>::kBytes;
    static constexpr int kBufA = kTileM * kKB;
    static constexpr int kBufB = kTileN * kKB;
    static constexpr int kLds = 2 * (kBufA + kBufB);
};

// STAGGER 0: every wave runs [MFMAs of chunk k, build A of chunk k+1].
// STAGGER 1/2: the waves with bit 2 (1) or bit 0 (2) of their index set run
// [build A of chunk k+1, MFMAs of chunk k] instead, so that the two waves a
// SIMD holds keep its matrix core and its vector ALU busy at the same time.
// STAGGER 8: the translation operand is precomputed (encode_a_kernel) and
// copied global -> LDS like B, one chunk ahead.
// STAGGER 4: each column slice's MFMAs share a fenced scheduling segment
// with one pixel of the next chunk's A build.

template <int FORM, int WM, int WN, int FM, int FN, int STAGGER>
__global__ __launch_bounds__(64 * WM * WN) void cross_kernel_wide(const CrossArgs a)
{
    using Cfg = WideCfg<FORM, WM, WN, FM, FN>;
    constexpr int NT = Cfg::kThreads;
    constexpr int TM = Cfg::kTileM;
    constexpr int TN = Cfg::kTileN;
    constexpr int KB = Cfg::kKB;
    constexpr int SEGS = WideEnc<FORM>::kSegs;
    constexpr int TPR = NT / TM;   // threads building one row of a chunk
    constexpr int PPT = 16 / TPR;  // pixels each of them builds
    static_assert(TPR * TM == NT && (TPR == 2 || TPR == 4), "2 or 4 threads per A row");
    constexpr int UNITS = TN * SEGS;
    static_assert(UNITS % NT == 0, "B tile must split evenly into 16-byte loads");
    constexpr int LPT = UNITS / NT;

    __shared__ __attribute__((aligned(16))) unsigned char sA[2][Cfg::kBufA];
    __shared__ __attribute__((aligned(16))) unsigned char sB[2][Cfg::kBufB];

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
    const int wm = wave % WM, wn = wave / WM;
    const int r16 = lane & 15, q = lane >> 4;
    const int nkc = a.kpad / 16;

    // ---- this thread's A row (row gr of the tile, pixels 8 gh .. 8 gh + 7 of a chunk)
    const int gr = tid / TPR, gh = tid % TPR;
    // Rows are packed: row R of the batch is translation R of the CSR
    // translation arrays (particle row_part[R]), so no particle is padded to a
    // multiple of 16 rows.  Rows past the last one are built from particle 0
    // with zero shift and never stored.
    const ARow ar = make_arow<FORM, PPT>(a, mtile * TM + gr, gh);
    const float2 *c_row = a.coords + PPT * gh;

    floatx4 acc[FN][FM];
#pragma unroll
    for (int fb = 0; fb < FN; ++fb)
#pragma unroll
        for (int fa = 0; fa < FM; ++fa) acc[fb][fa] = floatx4{0.f, 0.f, 0.f, 0.f};

    // projections: [class][chunk][orientation][KB]; the chunk of a column
    // tile is one contiguous run of TN * KB bytes
    const unsigned char *cls_proj = a.proj + (size_t)cls * nkc * O * KB;
    // Copy `it` of a thread fetches 16 bytes of column it * CPI + tid / SEGS:
    // the swizzled segment does not depend on `it` (CPI >> kShift is a
    // multiple of SEGS), so each copy is a uniform base plus one per-thread
    // 32-bit offset.  Columns past the last orientation read the padding
    // after the projections (or the next chunk's first columns): never stored.
    constexpr int CPI = NT / SEGS;   // columns per copy instruction
    static_assert(NT % SEGS == 0 && ((CPI >> WideEnc<FORM>::kShift) % SEGS) == 0,
                  "copy offsets must repeat across copies");
    const uint32_t b_off =
        (uint32_t)((tid / SEGS) * KB + wide_swz<FORM>(tid / SEGS, tid % SEGS) * 16);
    // uniform parts of the copies' addresses, in SGPRs
    const unsigned sB_lds = uniform32(
        (unsigned)(size_t)(__attribute__((address_space(3))) void *)&sB[0][0] + wave * 64 * 16);
    const unsigned long long b_base = uniform64((unsigned long long)(size_t)(cls_proj + (size_t)o0 * KB));
    const unsigned b_stride = uniform32((unsigned)(O * KB));   // bytes per chunk of one class
    auto issue_B = [&](int kc, int buf) {
        const unsigned long long chunk = b_base + (unsigned long long)(unsigned)kc * b_stride;
#pragma unroll
        for (int it = 0; it < LPT; ++it)
            lds_copy16(chunk + (unsigned long long)it * CPI * KB, b_off,
                       sB_lds + buf * Cfg::kBufB + it * NT * 16);
    };
    auto load_src = [&](int kc, float4 (&xw)[PPT / 2], float4 (&cxy)[PPT / 2]) {
        const float4 *xp = reinterpret_cast<const float4 *>(ar.xw + kc * 16);
        const float4 *cp = reinterpret_cast<const float4 *>(c_row + kc * 16);
#pragma unroll
        for (int u = 0; u < PPT / 2; ++u) {
            xw[u] = xp[u];
            cxy[u] = cp[u];
        }
    };

    const int ra0 = wm * FM * 16, rb0 = wn * FN * 16;
    auto mfma_chunk = [&](int buf) {
        const unsigned char *pa = sA[buf];
        const unsigned char *pb = sB[buf];
        WideFrag<FORM> fa_reg[FM];
#pragma unroll
        for (int fa = 0; fa < FM; ++fa) {
            const int ra = ra0 + fa * 16 + r16;
            wide_load<FORM>(fa_reg[fa], pa + ra * KB, ra, q);
        }
#pragma unroll
        for (int fb = 0; fb < FN; ++fb) {
            const int rb = rb0 + fb * 16 + r16;
            WideFrag<FORM> fb_reg;
            wide_load<FORM>(fb_reg, pb + rb * KB, rb, q);
            wide_mfma<FORM, FM>(acc[fb], fa_reg, fb_reg);
        }
    };

    // chunks past the projected pixels hold zeros: skip them
    const int nrun = (a.npix + 15) / 16;
    if constexpr (STAGGER == 8) {
        // Translation operand precomputed for the whole launch: each chunk
        // of this row tile is one contiguous block holding the exact LDS
        // image of sA, copied like B (LPA 16-byte copies per thread).
        constexpr int LPA = Cfg::kBufA / (NT * 16);
        static_assert(LPA * NT * 16 == Cfg::kBufA, "A tile must split evenly into 16-byte loads");
        const unsigned sA_lds = uniform32(
            (unsigned)(size_t)(__attribute__((address_space(3))) void *)&sA[0][0] + wave * 64 * 16);
        const unsigned long long a_base = uniform64(
            (unsigned long long)(size_t)(a.a_enc + (size_t)mtile * a.a_chunks * Cfg::kBufA));
        const uint32_t a_off = (uint32_t)(tid * 16);
        auto issue_A = [&](int kc, int buf) {
            const unsigned long long chunk =
                a_base + (unsigned long long)(unsigned)kc * (unsigned)Cfg::kBufA;
#pragma unroll
            for (int it = 0; it < LPA; ++it)
                lds_copy16(chunk + (unsigned long long)it * NT * 16, a_off,
                           sA_lds + buf * Cfg::kBufA + it * NT * 16);
        };
        issue_B(0, 0);
        issue_A(0, 0);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        for (int kc = 0; kc < nrun; ++kc) {
            const int buf = kc & 1;
            const int kn = min(kc + 1, nrun - 1);
            // fragment reads first, so that the LDS serves them before the
            // next chunk's copies land
            const unsigned char *pa = sA[buf];
            const unsigned char *pb = sB[buf];
            WideFrag<FORM> fa_reg[FM];
#pragma unroll
            for (int fa = 0; fa < FM; ++fa) {
                const int ra = ra0 + fa * 16 + r16;
                wide_load<FORM>(fa_reg[fa], pa + ra * KB, ra, q);
            }
            WideFrag<FORM> fb_reg[2];
            wide_load<FORM>(fb_reg[0], pb + (rb0 + r16) * KB, rb0 + r16, q);
            issue_B(kn, buf ^ 1);
            issue_A(kn, buf ^ 1);
            __builtin_amdgcn_s_setprio(1);
#pragma unroll
            for (int fb = 0; fb < FN; ++fb) {
                if (fb + 1 < FN) {
                    const int rb = rb0 + (fb + 1) * 16 + r16;
                    wide_load<FORM>(fb_reg[(fb + 1) & 1], pb + rb * KB, rb, q);
                }
                wide_mfma<FORM, FM>(acc[fb], fa_reg, fb_reg[fb & 1]);
            }
            __builtin_amdgcn_s_setprio(0);
            // this step's copies land before the barrier
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
        }
    } else {
        const bool lead = STAGGER == 0 || STAGGER == 4 ||
                          (((STAGGER == 1 ? (wave >> 2) : wave) & 1) == 0);
        float4 xw[PPT / 2], cxy[PPT / 2];
        issue_B(0, 0);
        load_src(0, xw, cxy);
        a_build<FORM, PPT>(ar, xw, cxy, gh, gr, &sA[0][gr * KB]);
        if (!lead || STAGGER == 4) load_src(min(1, nrun - 1), xw, cxy);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();

        for (int kc = 0; kc < nrun; ++kc) {
            const int buf = kc & 1;
            // the last step fetches its own chunk again into the idle buffers,
            // which keeps each path through the loop body one straight block
            const int kn = min(kc + 1, nrun - 1);
            if (STAGGER != 4) issue_B(kn, buf ^ 1);
            if (STAGGER == 4) {
                // one column slice of MFMAs and one pixel of the next chunk's A
                // per fenced segment; the image data for the next build were
                // loaded at the top of the previous step.  The fragment reads go
                // first, so that the LDS serves them before the next chunk's
                // copies land.
                const unsigned char *pa = sA[buf];
                const unsigned char *pb = sB[buf];
                WideFrag<FORM> fa_reg[FM];
    #pragma unroll
                for (int fa = 0; fa < FM; ++fa) {
                    const int ra = ra0 + fa * 16 + r16;
                    wide_load<FORM>(fa_reg[fa], pa + ra * KB, ra, q);
                }
                WideFrag<FORM> fb_reg[2];
                wide_load<FORM>(fb_reg[0], pb + (rb0 + r16) * KB, rb0 + r16, q);
                // the copies below are invisible to the compiler's wait counting:
                // have it wait for this step's image registers (loaded in the last
                // step) before they are issued, not after
    #pragma unroll
                for (int u = 0; u < PPT / 2; ++u)
                    asm volatile("" : "+v"(xw[u].x), "+v"(xw[u].y), "+v"(xw[u].z), "+v"(xw[u].w),
                                 "+v"(cxy[u].x), "+v"(cxy[u].y), "+v"(cxy[u].z), "+v"(cxy[u].w));
                issue_B(kn, buf ^ 1);
                // 4-pixel threads load the image data of the step after this one
                // a whole step early, into a second register set (8-pixel
                // threads have no registers to spare: after the segments)
                float4 xw2[PPT / 2], cxy2[PPT / 2];
                if constexpr (PPT == 4) load_src(min(kn + 1, nrun - 1), xw2, cxy2);
                float2 v[PPT];
                uint32_t hw[PPT], lw[PPT];   // BF16x3, 4 pixels: each encoded in its own segment
                // pixel j goes to segment j * FN / PPT + kPixOff: the segments
                // that wait for the first fragments carry no pixel work
                constexpr int kPixOff = FN > PPT ? FN / PPT - 1 : 0;
                // raised issue priority while this wave has MFMAs to issue: the
                // SIMD's other wave, past its MFMAs, yields to it (2-3 % here)
                __builtin_amdgcn_s_setprio(1);
                __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
                for (int fb = 0; fb < FN; ++fb) {
                    if (fb + 1 < FN) {
                        const int rb = rb0 + (fb + 1) * 16 + r16;
                        wide_load<FORM>(fb_reg[(fb + 1) & 1], pb + rb * KB, rb, q);
                    }
                    wide_mfma<FORM, FM>(acc[fb], fa_reg, fb_reg[fb & 1]);
    #pragma unroll
                    for (int j = 0; j < PPT; ++j)
                        if ((j * FN) / PPT + kPixOff == fb) {
                            v[j] = a_pixel<PPT>(ar, xw, cxy, j);
                            if constexpr (FORM == kBF16X3 && PPT == 4)
                                encode_bf16x3(v[j], hw[j], lw[j]);
                        }
                    __builtin_amdgcn_sched_barrier(0);
                }
                __builtin_amdgcn_s_setprio(0);
                if constexpr (FORM == kBF16X3 && PPT == 4)
                    a_store_words<PPT>(hw, lw, gh, gr, &sA[buf ^ 1][gr * KB]);
                else
                    a_store<FORM, PPT>(ar, v, gh, gr, &sA[buf ^ 1][gr * KB]);
                if constexpr (PPT == 4) {
    #pragma unroll
                    for (int u = 0; u < PPT / 2; ++u) {
                        xw[u] = xw2[u];
                        cxy[u] = cxy2[u];
                    }
                } else {
                    load_src(min(kn + 1, nrun - 1), xw, cxy);
                }
            } else if (lead) {
                load_src(kn, xw, cxy);
                mfma_chunk(buf);
                a_build<FORM, PPT>(ar, xw, cxy, gh, gr, &sA[buf ^ 1][gr * KB]);
            } else {
                a_build<FORM, PPT>(ar, xw, cxy, gh, gr, &sA[buf ^ 1][gr * KB]);
                load_src(min(kn + 1, nrun - 1), xw, cxy);
                mfma_chunk(buf);
            }
            // the next chunk's B copies (and the image loads) land before the barrier
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
        }

    }

    wide_epilogue<FORM, FM, FN, WN, TM>(a, acc, mtile, ntile, n_ntiles, cls, o0, wm, wn, rb0,
                                         r16, q, EpiGlobal{a, cls, o0, rb0, r16});
}

// ---------------------------------------------------------------------------
// wide variant table
// ---------------------------------------------------------------------------
template <int FORM, int WM, int WN, int FM, int FN, int ST>
static hipError_t launch_wide_one(const CrossArgs &a, hipStream_t s)
{
    using Cfg = WideCfg<FORM, WM, WN, FM, FN>;
    const int total = a.n_mtiles * a.n_classes * a.ntiles_per_class;
    const int grid = ((total + 7) / 8) * 8;
    hipLaunchKernelGGL((cross_kernel_wide<FORM, WM, WN, FM, FN, ST>), dim3(grid),
                       dim3(Cfg::kThreads), 0, s, a);
    return hipGetLastError();
}

template <int FORM, int WM, int WN, int FM, int FN, int ST>
static VariantEntry wide_one()
{
    using Cfg = WideCfg<FORM, WM, WN, FM, FN>;
    return VariantEntry{VariantInfo{FORM, Cfg::kTileM, Cfg::kTileN, Cfg::kThreads, WN, false,
                                    Cfg::kLds, false, true, "cross_kernel_wide", ST},
                        &launch_wide_one<FORM, WM, WN, FM, FN, ST>};
}

#define RC_WIDE_TABLE(F)                                                              \
    static const VariantEntry k_wide_##F[] = {                                         \
        wide_one<F, 4, 2, 4, 8, 0>(), /* 256 x 256, 512 threads */                     \
        wide_one<F, 2, 2, 4, 8, 0>(), /* 128 x 256, 256 threads */                     \
        wide_one<F, 4, 2, 4, 4, 0>(), /* 256 x 128, 512 threads */                     \
        wide_one<F, 2, 2, 4, 4, 0>(), /* 128 x 128, 256 threads */                     \
        wide_one<F, 4, 2, 4, 8, 1>(), /* 256 x 256, staggered by wave bit 2 */         \
        wide_one<F, 4, 2, 4, 8, 2>(), /* 256 x 256, staggered by wave bit 0 */         \
        wide_one<F, 4, 2, 4, 4, 1>(), /* 256 x 128, staggered by wave bit 2 */         \
        wide_one<F, 4, 2, 4, 8, 4>(), /* 256 x 256, fenced segments */                 \
        wide_one<F, 4, 2, 4, 4, 4>(), /* 256 x 128, fenced segments */                 \
        wide_one<F, 2, 4, 4, 8, 0>(), /* 128 x 512, 512 threads */                     \
        wide_one<F, 2, 4, 4, 8, 1>(), /* 128 x 512, staggered by wave bit 2 */         \
        wide_one<F, 2, 4, 4, 8, 4>(), /* 128 x 512, fenced segments */                 \
        wide_one<F, 2, 4, 4, 8, 8>(), /* 128 x 512, translation operand precomputed */ \
        wide_one<F, 4, 2, 4, 8, 8>(), /* 256 x 256, translation operand precomputed */ \
    };

RC_WIDE_TABLE(kFP32)
RC_WIDE_TABLE(kBF16X3)
RC_WIDE_TABLE(kFP16)

const VariantEntry *wide_table(int form, int *n)
{
    *n = (int)(sizeof(k_wide_kFP32) / sizeof(k_wide_kFP32[0]));
    switch (form) {
    case kFP32: return k_wide_kFP32;
    case kBF16X3: return k_wide_kBF16X3;
    case kFP16: return k_wide_kFP16;
    default:
        *n = 0;
        return nullptr;
    }
}

}  // namespace rc
