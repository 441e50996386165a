// "Direct" cross-term kernel: the projections go from global memory straight
// into the MFMA operand registers; only the translation operand is staged
// in LDS.
//
// Block = 128 rows (packed translations) x 512 orientations, 4 waves, one
// wave per SIMD.  Wave w owns all 128 rows and orientations 128 w .. 128 w +
// 127 of the tile: 8 x 8 accumulators of 16x16 (256 accumulator registers),
// so no two waves need the same projection fragment and none is shared
// through LDS.  Per 16-pixel chunk:
//   * B: each wave loads its 8 column slices of the chunk from the direct
//     projection layout (one 16-orientation group = the fragments in lane
//     order, so each load is one contiguous 1 KB per wave).  Slice fb of
//     chunk k+1 is loaded into slice fb's registers as soon as chunk k's
//     MFMAs of that slice have been issued: about one chunk ahead, in the
//     registers of one chunk.
//   * A: all 256 threads build the chunk two ahead (k+2) -- two threads per
//     row, one pixel per column slice -- and store it at the end of the
//     step into a double-buffered, swizzled LDS tile.  The next chunk's
//     fragments are read at the end of each step, as soon as the current
//     ones have been used for the last time, so a step starts with its
//     operands in registers.
//   * one barrier per chunk, in the middle of the step (after column slice
//     3): it publishes the chunk stored at the end of the previous step
//     before any wave reads it at the end of this one, and it separates the
//     reads of a buffer from its next overwrite.  Nothing is outstanding on
//     the LDS at that point, so the barrier waits only for the other waves.
// Per chunk and CU this moves 64 KB of projections (as the LDS-staged wide
// kernels do) but only 16 KB of LDS writes and 64 KB of LDS reads (those
// kernels: 80 KB and 192 KB).
//
// PRE instances take the translation operand precomputed (encode_a_kernel:
// each row tile's chunk is one contiguous 16 KB block holding the exact LDS
// image of the A buffer) and copy it global -> LDS with LDS-DMA, two chunks
// ahead: no image staging, no phase arithmetic in the loop.  The copy of
// A(k+2) goes out right after step k's barrier (which frees A(k)'s buffer)
// and is waited for at step k+1's barrier, one and a half steps later.
//
// The MFMA sequence of every accumulator -- chunk by chunk, and within a
// chunk (lo x hi, hi x lo, hi x hi) for BF16X3 -- and the operand values are
// those of the wide kernels, so the results are bit-identical to theirs.
#include <hip/hip_runtime.h>

#include "wide_common.hpp"

namespace rc {

namespace {

constexpr int kDirTM = 128;      // rows per block
constexpr int kDirWN = 4;        // waves (column groups of 128 orientations)
constexpr int kDirFM = 8;        // row slices per wave
constexpr int kDirFN = 8;        // column slices per wave
constexpr int kDirThreads = 64 * kDirWN;
constexpr int kStageSlots = 4;   // image-data ring: chunk c in slot c % 4, staged 4 chunks ahead
constexpr int kStageBytes = (1 + kDirectMaxTileParts) * 128;   // coords + particles, 16 pixels

// One column slice (16 orientations) of one chunk, as lane l = 16 q + r16
// of the wave reads it.
template <int FORM>
__device__ __forceinline__ void direct_load_b(WideFrag<FORM> &f, const unsigned char *group,
                                              int lane)
{
    if constexpr (FORM == kBF16X3) {
        f.hi = *reinterpret_cast<const bf16x8 *>(group + lane * 16);
        f.lo = *reinterpret_cast<const bf16x8 *>(group + 1024 + lane * 16);
    } else if constexpr (FORM == kFP16) {
        f.v = *reinterpret_cast<const halfx8 *>(group + lane * 16);
    } else {
        const float4 x0 = *reinterpret_cast<const float4 *>(group + lane * 32);
        const float4 x1 = *reinterpret_cast<const float4 *>(group + lane * 32 + 16);
        f.v[0] = x0.x;
        f.v[1] = x0.y;
        f.v[2] = x0.z;
        f.v[3] = x0.w;
        f.v[4] = x1.x;
        f.v[5] = x1.y;
        f.v[6] = x1.z;
        f.v[7] = x1.w;
    }
}

// Inputs of a direct tile's epilogue, staged in LDS by the kernel's
// prologue: a row's key is the local index of its particle among the
// tile's particles; the column norms of particle k are kept lane-major, the
// eight column slices of lane r16 of wave wn at k * 512 + wn * 128 + r16 * 8.
struct EpiTables {
    const int *pidx;          // [128] per tile row
    const int *first, *last;  // [parts] trans_offset[p], trans_offset[p + 1]
    const double *xnorm;      // [parts]
    const float *ainv;        // [parts] FP16 form: 1 / (image scale x volume scale)
    const float *pnorm;       // [parts][512]
    int R0, col;              // the tile's first row; wn * 128 + r16 * 8
    __device__ __forceinline__ int key(int R) const { return pidx[R - R0]; }
    __device__ __forceinline__ int b0(int k) const { return first[k]; }
    __device__ __forceinline__ int end(int k) const { return last[k]; }
    __device__ __forceinline__ double xn(int k) const { return xnorm[k]; }
    __device__ __forceinline__ float inv(int k) const { return ainv[k]; }
    __device__ __forceinline__ float pn(int k, int fb) const { return pnorm[k * 512 + col + fb]; }
};

}  // namespace

// s_waitcnt vmcnt(N), then the block barrier (LDS reads drained first).
template <int N>
__device__ __forceinline__ void wait_vm_barrier()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    asm volatile("s_waitcnt vmcnt(%0)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::"n"(N) : "memory");
}

template <int FORM, bool PRE>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1, 1)))
void cross_kernel_direct(const CrossArgs a)
{
    constexpr int TM = kDirTM, FM = kDirFM, FN = kDirFN, WN = kDirWN;
    constexpr int TN = WN * FN * 16;
    constexpr int NT = kDirThreads;
    constexpr int KB = WideEnc<FORM>::kBytes;
    constexpr int GB = 16 * KB;          // one 16-orientation group of one chunk
    constexpr int TPR = NT / TM;         // threads building one A row
    constexpr int PPT = 16 / TPR;        // pixels each of them builds per chunk
    static_assert(TPR == 2 && PPT == FN, "one pixel of the build per column slice");

    __shared__ __attribute__((aligned(16))) unsigned char sA[2][TM * KB];
    __shared__ __attribute__((aligned(16))) unsigned char sImg[kStageSlots][PRE ? 16 : kStageBytes];
    // epilogue inputs (EpiTables)
    __shared__ __attribute__((aligned(16))) float s_pn[kDirectMaxTileParts * TN];
    __shared__ double s_xn[kDirectMaxTileParts];
    __shared__ int s_first[kDirectMaxTileParts], s_last[kDirectMaxTileParts], s_pidx[TM];
    __shared__ float s_inv[kDirectMaxTileParts];

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
    const int wn = (int)__builtin_amdgcn_readfirstlane((unsigned)(tid >> 6));
    const int r16 = lane & 15, q = lane >> 4;
    const int nkc = a.kpad / 16;
    const int n_groups = (O + 15) / 16;
    const int nrun = (a.npix + 15) / 16;   // chunks past the projected pixels are zero

    // ---- this thread's A row: row gr of the tile, pixels PPT gh .. PPT gh + PPT - 1
    const int gr = tid / TPR, gh = tid % TPR;
    const ARow ar = make_arow<FORM, PPT>(a, mtile * TM + gr, gh);

    // ---- image data: a chunk's pixel coordinates and the corr * X of each
    // particle of the tile's rows (at most kDirectMaxTileParts, checked on
    // the host) are staged in LDS once per block, kStageSlots chunks ahead,
    // by one 16-byte load per thread (block sb: 0 coordinates, 1 + j particle
    // part0 + j; unit su: pixels 2 su, 2 su + 1).  The builds read them there
    // instead of each thread loading its own pixels.
    const int R0 = mtile * TM;
    const int part0 = a.row_part[R0];
    const int n_parts = a.row_part[min(R0 + TM - 1, a.n_rows - 1)] - part0 + 1;
    // (Threads past the tile's particles repeat the last particle into
    // blocks nobody reads: no branch around the load or the store.)
    const int sb = tid >> 3, su = tid & 7;
    const float4 *stage_src = reinterpret_cast<const float4 *>(
        sb == 0 ? a.coords : a.xw + (size_t)(part0 + min(sb, n_parts) - 1) * a.kpad);
    auto stage_load = [&](int c) { return stage_src[min(c, nrun - 1) * 8 + su]; };
    auto stage_store = [&](int c, float4 v) {
        *reinterpret_cast<float4 *>(&sImg[c % kStageSlots][sb * 128 + su * 16]) = v;
    };
    const int pidx =
        R0 + gr < a.n_rows ? min(a.row_part[R0 + gr] - part0, kDirectMaxTileParts - 1) : 0;

    // ---- epilogue inputs into LDS, published by the prologue's barrier:
    // nothing in the epilogue waits on a chain of global loads
    if (tid < TM)
        s_pidx[tid] = R0 + tid < a.n_rows
                          ? min(a.row_part[R0 + tid] - part0, kDirectMaxTileParts - 1)
                          : 0;
    if (tid < min(n_parts, kDirectMaxTileParts)) {
        const int p = part0 + tid;
        s_first[tid] = a.trans_offset[p];
        s_last[tid] = a.trans_offset[p + 1];
        s_xn[tid] = a.xnorm[p];
        s_inv[tid] = FORM == kFP16 ? 1.f / (a.ascale[p] * a.bscale[cls]) : 1.f;
    }
    for (int j = 0; j < min(n_parts, kDirectMaxTileParts); ++j)
        for (int c = tid; c < TN; c += NT) {
            const float v =
                a.pnorm[((size_t)(part0 + j) * a.n_classes + cls) * O + min(o0 + c, O - 1)];
            // column c = 128 w + 16 f + r of the tile -> lane-major slot 128 w + 8 r + f
            s_pn[j * TN + (c & ~127) + (c & 15) * 8 + ((c >> 4) & 7)] = v;
        }

    // ---- B: projections in the direct layout, [class][chunk][group][GB]
    const unsigned char *b_wave =
        a.proj + ((size_t)cls * nkc * n_groups + (size_t)(o0 / 16 + wn * FN)) * GB;
    const size_t b_chunk = (size_t)n_groups * GB;

    floatx4 acc[FN][FM];
#pragma unroll
    for (int fb = 0; fb < FN; ++fb)
#pragma unroll
        for (int fa = 0; fa < FM; ++fa) acc[fb][fa] = floatx4{0.f, 0.f, 0.f, 0.f};

    WideFrag<FORM> aR[FM], bR[FN];
    float4 xw[PPT / 2], cxy[PPT / 2];
    APix<FORM> v[PPT];

    // pixels 2u, 2u + 1 of the thread's eight in staged chunk c
    auto read_img = [&](int c, int u) {
        const unsigned char *slot = sImg[c % kStageSlots];
        xw[u] = *reinterpret_cast<const float4 *>(slot + (1 + pidx) * 128 + gh * 64 + u * 16);
        cxy[u] = *reinterpret_cast<const float4 *>(slot + gh * 64 + u * 16);
    };
    auto read_a = [&](const unsigned char *buf, int fa) {
        const int ra = fa * 16 + r16;
        wide_load<FORM>(aR[fa], buf + ra * KB, ra, q);
    };

    // ---- PRE: A chunk kc of this row tile, copied into A buffer buf.
    // Thread t copies bytes 16 t + 4096 i of the 16 KB block (i < LPA): each
    // wave writes 1 KB runs of LDS from its own base, as the copy requires.
    constexpr int LPA = TM * KB / (NT * 16);
    static_assert(LPA * NT * 16 == TM * KB, "A chunk splits evenly into 16-byte copies");
    constexpr int LB = FORM == kFP16 ? 1 : 2;   // global loads per column slice of B
    const unsigned sA_lds = uniform32(
        (unsigned)(size_t)(__attribute__((address_space(3))) void *)&sA[0][0] + wn * 64 * 16);
    const unsigned long long a_tile = uniform64(
        (unsigned long long)(size_t)(a.a_enc + (size_t)mtile * a.a_chunks * (TM * KB)));
    const uint32_t a_off = (uint32_t)(tid * 16);
    auto issue_A = [&](int kc, int buf) {
        const unsigned long long chunk =
            a_tile + (unsigned long long)(unsigned)kc * (unsigned)(TM * KB);
#pragma unroll
        for (int it = 0; it < LPA; ++it)
            lds_copy16(chunk + (unsigned long long)it * NT * 16, a_off,
                       sA_lds + buf * (TM * KB) + it * NT * 16);
    };

    // ---- prologue: image data of chunks 0-3 staged; A(0) and A(1) built
    // and stored (PRE: copied); B(0) loaded; A(0) read after the barrier.
    // The loads and the LDS reads go out in the order the loop issues them,
    // so that the waits the compiler places at the top of the loop (valid
    // on entry and on every later step) are those of the steady state.
    auto load_b0 = [&]() {
        static_for<0, FN>([&](auto fb_c) {
            constexpr int fb = decltype(fb_c)::value;
            direct_load_b<FORM>(bR[fb], b_wave + fb * GB, lane);
            __builtin_amdgcn_sched_barrier(0);
        });
    };
    float4 stg{};
    if constexpr (PRE) {
        issue_A(0, 0);
        issue_A(min(1, nrun - 1), 1);
        __builtin_amdgcn_sched_barrier(0);
        load_b0();
        // A(0) landed: everything but A(1) and B(0); the barrier also
        // publishes the epilogue tables
        wait_vm_barrier<LPA + FN * LB>();
    } else {
        {
            float4 st[kStageSlots];
#pragma unroll
            for (int c = 0; c < kStageSlots; ++c) st[c] = stage_load(c);
#pragma unroll
            for (int c = 0; c < kStageSlots; ++c) stage_store(c, st[c]);
        }
        __syncthreads();
        for (int c = 0; c < 2; ++c) {
#pragma unroll
            for (int u = 0; u < PPT / 2; ++u) read_img(c, u);
            a_build<FORM, PPT>(ar, xw, cxy, gh, gr, &sA[c][gr * KB]);
        }
        __builtin_amdgcn_sched_barrier(0);
        load_b0();
        stg = stage_load(kStageSlots);
        read_img(2, 0);
        read_img(2, 1);
        __syncthreads();
    }
    if constexpr (FORM == kBF16X3) {
#pragma unroll
        for (int fa = 0; fa < FM; ++fa) {
            const int ra = fa * 16 + r16;
            aR[fa].lo = *reinterpret_cast<const bf16x8 *>(sA[0] + ra * KB +
                                                          wide_swz<FORM>(ra, 4 + q) * 16);
        }
#pragma unroll
        for (int fa = 0; fa < FM; ++fa) {
            const int ra = fa * 16 + r16;
            aR[fa].hi = *reinterpret_cast<const bf16x8 *>(sA[0] + ra * KB +
                                                          wide_swz<FORM>(ra, q) * 16);
        }
    } else {
#pragma unroll
        for (int fa = 0; fa < FM; ++fa) read_a(sA[0], fa);
    }
    __builtin_amdgcn_sched_barrier(0);

    for (int k = 0; k < nrun; ++k) {
        // chunks past the last one are clamped to it: their loads stay in
        // bounds and their results are never used
        const int kb = min(k + 1, nrun - 1);     // B loaded in this step
        const unsigned char *b_next = b_wave + (size_t)kb * b_chunk;
        unsigned char *s_build = sA[k & 1];              // A(k+2)
        const unsigned char *s_next = sA[(k + 1) & 1];   // A(k+1)
        __builtin_amdgcn_sched_barrier(0);
        static_for<0, FN>([&](auto fb_c) {
            constexpr int fb = decltype(fb_c)::value;
            constexpr bool last = fb == FN - 1;
            // image data of the pixel pair after the next, two column slices
            // ahead of its use: pixels fb + 2, fb + 3 of chunk k+2 (staged two
            // steps ago, published by the last step's barrier); in slice 6
            // pixels 0, 1 of chunk k+3 (published by this step's barrier)
            if constexpr (!PRE && (fb & 1) == 0) {
                if constexpr (fb == FN - 2) read_img(k + 3, 0);
                else read_img(k + 2, (fb + 2) >> 1);
            }
            if constexpr (FORM == kBF16X3) {
#pragma unroll
                for (int fa = 0; fa < FM; ++fa)
                    acc[fb][fa] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aR[fa].lo, bR[fb].hi,
                                                                          acc[fb][fa], 0, 0, 0);
                if constexpr (last) {
                    // last use of the lo parts: the next chunk's go in
#pragma unroll
                    for (int fa = 0; fa < FM; ++fa) {
                        const int ra = fa * 16 + r16;
                        aR[fa].lo = *reinterpret_cast<const bf16x8 *>(
                            s_next + ra * KB + wide_swz<FORM>(ra, 4 + q) * 16);
                    }
                }
#pragma unroll
                for (int fa = 0; fa < FM; ++fa)
                    acc[fb][fa] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aR[fa].hi, bR[fb].lo,
                                                                          acc[fb][fa], 0, 0, 0);
#pragma unroll
                for (int fa = 0; fa < FM; ++fa) {
                    acc[fb][fa] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aR[fa].hi, bR[fb].hi,
                                                                          acc[fb][fa], 0, 0, 0);
                    if constexpr (last) {
                        const int ra = fa * 16 + r16;
                        aR[fa].hi = *reinterpret_cast<const bf16x8 *>(
                            s_next + ra * KB + wide_swz<FORM>(ra, q) * 16);
                    }
                }
            } else if constexpr (FORM == kFP16) {
#pragma unroll
                for (int fa = 0; fa < FM; ++fa) {
                    acc[fb][fa] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aR[fa].v, bR[fb].v,
                                                                         acc[fb][fa], 0, 0, 0);
                    if constexpr (last) read_a(s_next, fa);
                }
            } else {
#pragma unroll
                for (int m = 0; m < 8; ++m)
#pragma unroll
                    for (int fa = 0; fa < FM; ++fa)
                        acc[fb][fa] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                            aR[fa].v[m], bR[fb].v[m], acc[fb][fa], 0, 0, 0);
                if constexpr (last) {
#pragma unroll
                    for (int fa = 0; fa < FM; ++fa) read_a(s_next, fa);
                }
            }
            // pixel fb of chunk k+2, encoded
            if constexpr (!PRE) v[fb] = a_encode<FORM>(ar, a_pixel<PPT>(ar, xw, cxy, fb));
            // PRE: A(k+2) into A(k)'s buffer, freed by this step's barrier;
            // issued before slice 4's B load, so that the B loads the
            // compiler waits for are never younger than a copy it cannot see
            if constexpr (PRE && fb == 4) issue_A(min(k + 2, nrun - 1), k & 1);
            // chunk k+1 of this column slice, into the registers just used
            direct_load_b<FORM>(bR[fb], b_next + fb * GB, lane);
            if constexpr (fb == 3) {
                // publishes A(k+1) and the image data of chunk k+4; every wave
                // has finished reading A(k), whose buffer receives A(k+2) at
                // the end of this step.  PRE: A(k+1) was copied at slice 4 of
                // the previous step; eight column slices of B loads followed it.
                if constexpr (PRE) wait_vm_barrier<FN * LB>();
                else asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
            }
            if constexpr (FORM == kBF16X3 && !PRE) {
                // One vector-ALU instruction after each MFMA: the build's
                // float64 phase, sin/cos and encode run under the slice's 24
                // MFMAs instead of bunching between its first few (and the
                // accumulators stay put: no AGPR copies in the loop).  FP16
                // (8 MFMAs per slice) is faster left to the scheduler.
                static_for<0, 3 * FM>([&](auto) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
                    __builtin_amdgcn_sched_group_barrier(0x002, 1, 0);   // VALU
                });
            }
            __builtin_amdgcn_sched_barrier(0);
        });
        if constexpr (!PRE) {
            a_store_enc<FORM, PPT>(v, gh, gr, &s_build[gr * KB]);
            // chunk k+4 into the slot of chunk k (read two steps ago), the
            // next stage load in flight for a step
            stage_store(k + kStageSlots, stg);
            stg = stage_load(k + 1 + kStageSlots);
        }
    }

    static_assert(TN == 512, "EpiTables assumes 512 columns");
    const EpiTables tables{s_pidx, s_first, s_last, s_xn, s_inv, s_pn, R0, wn * 128 + r16 * 8};
    wide_epilogue<FORM, FM, FN, WN, TM>(a, acc, mtile, ntile, n_ntiles, cls, o0, 0, wn,
                                        wn * FN * 16, r16, q, tables);
}

template <int FORM, bool PRE>
static hipError_t launch_direct_one(const CrossArgs &a, hipStream_t s)
{
    const int total = a.n_mtiles * a.n_classes * a.ntiles_per_class;
    const int grid = ((total + 7) / 8) * 8;
    if (PRE && (a.a_enc == nullptr || a.a_chunks != (a.npix + 15) / 16)) return hipErrorInvalidValue;
    hipLaunchKernelGGL((cross_kernel_direct<FORM, PRE>), dim3(grid), dim3(kDirThreads), 0, s, a);
    return hipGetLastError();
}

// LDS bytes of one block: the double-buffered translation operand, the
// image-data ring (not PRE) and the epilogue tables.
template <int FORM, bool PRE>
constexpr int direct_lds_bytes()
{
    return 2 * kDirTM * WideEnc<FORM>::kBytes + kStageSlots * (PRE ? 16 : kStageBytes) +
           kDirectMaxTileParts * (kDirWN * kDirFN * 16 * 4 + 8 + 3 * 4) + kDirTM * 4;
}

template <int FORM, bool PRE>
static VariantEntry direct_one()
{
    return VariantEntry{VariantInfo{FORM, kDirTM, kDirWN * kDirFN * 16, kDirThreads, kDirWN,
                                    false, direct_lds_bytes<FORM, PRE>(), false, true,
                                    "cross_kernel_direct",
                                    PRE ? kScheduleDirectPre : kScheduleDirect,
                                    kDirectMaxTileParts},
                        &launch_direct_one<FORM, PRE>};
}

static const VariantEntry k_direct_fp32[] = {direct_one<kFP32, false>(), direct_one<kFP32, true>()};
static const VariantEntry k_direct_bf16x3[] = {direct_one<kBF16X3, false>(),
                                               direct_one<kBF16X3, true>()};
static const VariantEntry k_direct_fp16[] = {direct_one<kFP16, false>(), direct_one<kFP16, true>()};

const VariantEntry *direct_table(int form, int *n)
{
    *n = 2;
    switch (form) {
    case kFP32: return k_direct_fp32;
    case kBF16X3: return k_direct_bf16x3;
    case kFP16: return k_direct_fp16;
    default:
        *n = 0;
        return nullptr;
    }
}

}  // namespace rc
