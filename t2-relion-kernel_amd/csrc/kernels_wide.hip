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

#include "device_common.hpp"

namespace rc {

constexpr double kInv2PiW = 0.15915494309189533577;

template <int FORM>
struct WideEnc {
    static constexpr int kBytes = FORM == kFP16 ? 64 : 128;
    static constexpr int kSegs = kBytes / 16;
    static constexpr int kShift = kSegs == 8 ? 1 : 2;   // rows per 256-byte bank row
};

template <int FORM>
__device__ __forceinline__ int wide_swz(int r, int s)
{
    return s ^ ((r >> WideEnc<FORM>::kShift) & (WideEnc<FORM>::kSegs - 1));
}

// Encode one complex value (pixel p in [0,16), parts re/im) into an unswizzled
// row image `row` of WideEnc<FORM>::kBytes bytes (used by the projection
// precompute, one value per thread).
template <int FORM>
__device__ __forceinline__ void wide_store_value(unsigned char *row, int p, float re, float im,
                                                 float scale)
{
    if constexpr (FORM == kBF16X3) {
        __bf16 *b = reinterpret_cast<__bf16 *>(row);
        __bf16 h, l;
        split_bf16(re, h, l);
        b[2 * p] = h;
        b[32 + 2 * p] = l;
        split_bf16(im, h, l);
        b[2 * p + 1] = h;
        b[32 + 2 * p + 1] = l;
    } else if constexpr (FORM == kFP16) {
        _Float16 *h = reinterpret_cast<_Float16 *>(row);
        h[2 * p] = (_Float16)(re * scale);
        h[2 * p + 1] = (_Float16)(im * scale);
    } else {
        float *f = reinterpret_cast<float *>(row);
        const int v0 = 2 * p, v1 = 2 * p + 1;
        f[(v0 & 3) * 8 + (v0 >> 2)] = re;
        f[(v1 & 3) * 8 + (v1 >> 2)] = im;
    }
}

// ---------------------------------------------------------------------------
// projection precompute in the wide encoding: block = 16 orientations x one
// 16-pixel chunk
// ---------------------------------------------------------------------------
template <int FORM>
__global__ __launch_bounds__(256) void project_wide_kernel(const ProjView *__restrict__ vols,
                                                           const float *__restrict__ bscale,
                                                           const float *__restrict__ eulers,
                                                           int O, const float2 *__restrict__ coords,
                                                           int npix, int kpad,
                                                           unsigned char *__restrict__ proj,
                                                           float *__restrict__ pabs2)
{
    constexpr int KB = WideEnc<FORM>::kBytes;
    __shared__ __attribute__((aligned(16))) unsigned char s_out[16 * KB];
    const int nkc = kpad / 16;
    const int kc = blockIdx.x, cls = blockIdx.z;
    const int ol = threadIdx.x >> 4, pl = threadIdx.x & 15;
    const int o = blockIdx.y * 16 + ol;
    const int p = kc * 16 + pl;
    const ProjView v = vols[cls];
    float2 val = make_float2(0.f, 0.f);
    if (o < O && p < npix) {
        const float *e = eulers + 9 * (size_t)o;
        const float2 xy = coords[p];
        val = project_sample(v, e[0], e[1], e[3], e[4], e[6], e[7], xy.x, xy.y);
    }
    const float sc = (FORM == kFP16) ? bscale[cls] : 1.f;
    wide_store_value<FORM>(s_out + ol * KB, pl, val.x, val.y, sc);
    if (o < O) pabs2[((size_t)cls * O + o) * kpad + p] = val.x * val.x + val.y * val.y;
    __syncthreads();
    constexpr int units = 16 * KB / 16;
    for (int u = threadIdx.x; u < units; u += 256) {
        const int oo = blockIdx.y * 16 + u / (KB / 16);
        if (oo < O) {
            const int seg = u % (KB / 16);
            const uint4 d = reinterpret_cast<const uint4 *>(s_out)[u];
            reinterpret_cast<uint4 *>(proj + (((size_t)cls * nkc + kc) * O + oo) * KB)[seg] = d;
        }
    }
}

hipError_t launch_project_wide(int form, const ProjView *vols, const float *bscale, int n_classes,
                               const float *eulers, int O, const float2 *coords, int npix,
                               int kpad, unsigned char *proj, float *pabs2, hipStream_t s)
{
    dim3 grid(kpad / 16, (O + 15) / 16, n_classes);
    switch (form) {
    case kFP32:
        hipLaunchKernelGGL(project_wide_kernel<kFP32>, grid, dim3(256), 0, s, vols, bscale,
                           eulers, O, coords, npix, kpad, proj, pabs2);
        break;
    case kBF16X3:
        hipLaunchKernelGGL(project_wide_kernel<kBF16X3>, grid, dim3(256), 0, s, vols, bscale,
                           eulers, O, coords, npix, kpad, proj, pabs2);
        break;
    default:
        hipLaunchKernelGGL(project_wide_kernel<kFP16>, grid, dim3(256), 0, s, vols, bscale,
                           eulers, O, coords, npix, kpad, proj, pabs2);
        break;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// A generation: 8 or 4 consecutive pixels of one row per thread
// ---------------------------------------------------------------------------
// Rows past the batch's last translation keep finite values -- zero shift,
// particle 0's image -- and need no masking: row R of the product depends on
// row R of A only, and the epilogue never stores or ranks such rows.
struct ARow {
    const float2 *xw;      // this row's particle: corr * X, padded to kpad
    double txr, tyr;       // translation in revolutions per pixel
    float scale;           // FP16 form: power-of-two image scale
};

__device__ __forceinline__ float2 a_value(const ARow &r, float x, float y, float2 w)
{
    float s, c;
    phase_factor(x, y, r.txr, r.tyr, s, c);
    return make_float2(c * w.x - s * w.y, c * w.y + s * w.x);
}

// Value j of the thread's PPT consecutive pixels of a chunk (source data
// as PPT/2 float4: two pixels each).
template <int PPT>
__device__ __forceinline__ float2 a_pixel(const ARow &ar, const float4 (&xw)[PPT / 2],
                                          const float4 (&cxy)[PPT / 2], int j)
{
    const float4 c = cxy[j >> 1], w = xw[j >> 1];
    const float x = (j & 1) ? c.z : c.x, y = (j & 1) ? c.w : c.y;
    const float2 ww = (j & 1) ? make_float2(w.z, w.w) : make_float2(w.x, w.y);
    return a_value(ar, x, y, ww);
}

// BF16x3 words of one value: hi = (bf16 re, bf16 im), lo = the remainders.
// (Two-lane vector conversions: one packed round-to-nearest-even conversion
// per word, the same values as split_bf16 on each component.)
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void encode_bf16x3(float2 z, uint32_t &hi, uint32_t &lo)
{
    const f32x2 v = {z.x, z.y};
    const bf16x2 h = __builtin_convertvector(v, bf16x2);
    const f32x2 hv = __builtin_convertvector(h, f32x2);
    const f32x2 r = {v.x - hv.x, v.y - hv.y};
    hi = __builtin_bit_cast(uint32_t, h);
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2));
}

// Stores the BF16x3 words of pixels PPT*h .. PPT*h+PPT-1 into row r:
// group g of 4 pixels goes to hi segment g and lo segment 4 + g.
template <int PPT>
__device__ __forceinline__ void a_store_words(const uint32_t (&hi)[PPT], const uint32_t (&lo)[PPT],
                                              int h, int r, unsigned char *dst)
{
    constexpr int G = PPT / 4;
#pragma unroll
    for (int half = 0; half < G; ++half) {
        const int g = G * h + half;
        *reinterpret_cast<uint4 *>(dst + wide_swz<kBF16X3>(r, g) * 16) =
            make_uint4(hi[4 * half], hi[4 * half + 1], hi[4 * half + 2], hi[4 * half + 3]);
        *reinterpret_cast<uint4 *>(dst + wide_swz<kBF16X3>(r, 4 + g) * 16) =
            make_uint4(lo[4 * half], lo[4 * half + 1], lo[4 * half + 2], lo[4 * half + 3]);
    }
}

// Encodes the thread's PPT (8 or 4) values -- pixels PPT*h .. PPT*h+PPT-1 of
// the chunk -- into LDS row buffer `dst` of row r.
template <int FORM, int PPT>
__device__ __forceinline__ void a_store(const ARow &ar, const float2 (&v)[PPT], int h, int r,
                                        unsigned char *dst)
{
    static_assert(PPT == 8 || PPT == 4, "8 or 4 pixels per thread");
    constexpr int G = PPT / 4;   // 4-pixel groups (one 16-byte segment each)
    if constexpr (FORM == kBF16X3) {
        uint32_t hi[PPT], lo[PPT];
#pragma unroll
        for (int j = 0; j < PPT; ++j) encode_bf16x3(v[j], hi[j], lo[j]);
        a_store_words<PPT>(hi, lo, h, r, dst);
    } else if constexpr (FORM == kFP16) {
#pragma unroll
        for (int half = 0; half < G; ++half) {
            uint32_t w[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float2 z = v[4 * half + i];
                const _Float16 re = (_Float16)(z.x * ar.scale), im = (_Float16)(z.y * ar.scale);
                w[i] = (uint32_t)__builtin_bit_cast(uint16_t, re) |
                       ((uint32_t)__builtin_bit_cast(uint16_t, im) << 16);
            }
            *reinterpret_cast<uint4 *>(dst + wide_swz<FORM>(r, G * h + half) * 16) =
                make_uint4(w[0], w[1], w[2], w[3]);
        }
    } else if constexpr (PPT == 8) {
        // segment 2q + h holds component q % 2 of pixels 2i + q / 2, i = 0..3
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float f[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float2 z = v[2 * i + (q >> 1)];
                f[i] = (q & 1) ? z.y : z.x;
            }
            *reinterpret_cast<float4 *>(dst + wide_swz<FORM>(r, 2 * q + h) * 16) =
                make_float4(f[0], f[1], f[2], f[3]);
        }
    } else {
        // floats q*8 + 2h, q*8 + 2h + 1 hold component q % 2 of pixels q / 2
        // and 2 + q / 2: 8 bytes in segment 2q + h/2 at byte (h & 1) * 8
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float2 z0 = v[q >> 1], z1 = v[2 + (q >> 1)];
            const float2 f = (q & 1) ? make_float2(z0.y, z1.y) : make_float2(z0.x, z1.x);
            *reinterpret_cast<float2 *>(dst + wide_swz<FORM>(r, 2 * q + (h >> 1)) * 16 +
                                        (h & 1) * 8) = f;
        }
    }
}

// Builds the thread's part h of a chunk into LDS row buffer `dst` (row r).
template <int FORM, int PPT>
__device__ __forceinline__ void a_build(const ARow &ar, const float4 (&xw)[PPT / 2],
                                        const float4 (&cxy)[PPT / 2], int h, int r,
                                        unsigned char *dst)
{
    float2 v[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) v[j] = a_pixel<PPT>(ar, xw, cxy, j);
    a_store<FORM, PPT>(ar, v, h, r, dst);
}

// The A row of packed row R for a thread building pixels PPT*gh ..
// PPT*gh+PPT-1 of each chunk.  Rows past the batch's last one are built from
// particle 0 with zero shift (finite, never stored or ranked).
template <int FORM, int PPT>
__device__ __forceinline__ ARow make_arow(const CrossArgs &a, int R, int gh)
{
    ARow ar;
    const bool ok = R < a.n_rows;
    const int part = ok ? a.row_part[R] : 0;
    ar.txr = ok ? (double)a.trans_x[R] * kInv2PiW : 0.0;
    ar.tyr = ok ? (double)a.trans_y[R] * kInv2PiW : 0.0;
    ar.scale = (FORM == kFP16) ? a.ascale[part] : 1.f;
    ar.xw = a.xw + (size_t)part * a.kpad + PPT * gh;
    return ar;
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

// Operand registers of one 16-row (or 16-column) slice for one chunk.
template <int FORM>
struct WideFrag {
    bf16x8 hi, lo;
};
template <>
struct WideFrag<kFP16> {
    halfx8 v;
};
template <>
struct WideFrag<kFP32> {
    float v[8];
};

// Lane (r16, q) reads row `r` (base `row`) of a staged buffer.
template <int FORM>
__device__ __forceinline__ void wide_load(WideFrag<FORM> &f, const unsigned char *row, int r,
                                          int q)
{
    if constexpr (FORM == kBF16X3) {
        f.hi = *reinterpret_cast<const bf16x8 *>(row + wide_swz<FORM>(r, q) * 16);
        f.lo = *reinterpret_cast<const bf16x8 *>(row + wide_swz<FORM>(r, 4 + q) * 16);
    } else if constexpr (FORM == kFP16) {
        f.v = *reinterpret_cast<const halfx8 *>(row + wide_swz<FORM>(r, q) * 16);
    } else {
        const float4 x0 = *reinterpret_cast<const float4 *>(row + wide_swz<FORM>(r, 2 * q) * 16);
        const float4 x1 =
            *reinterpret_cast<const float4 *>(row + wide_swz<FORM>(r, 2 * q + 1) * 16);
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

// One column slice against all FM row slices.  The accumulators of one
// column slice are independent, so consecutive MFMAs never wait on each
// other's results.
template <int FORM, int FM>
__device__ __forceinline__ void wide_mfma(floatx4 (&acc)[FM], const WideFrag<FORM> (&a)[FM],
                                          const WideFrag<FORM> &b)
{
    if constexpr (FORM == kBF16X3) {
#pragma unroll
        for (int fa = 0; fa < FM; ++fa)
            acc[fa] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[fa].lo, b.hi, acc[fa], 0, 0, 0);
#pragma unroll
        for (int fa = 0; fa < FM; ++fa)
            acc[fa] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[fa].hi, b.lo, acc[fa], 0, 0, 0);
#pragma unroll
        for (int fa = 0; fa < FM; ++fa)
            acc[fa] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[fa].hi, b.hi, acc[fa], 0, 0, 0);
    } else if constexpr (FORM == kFP16) {
#pragma unroll
        for (int fa = 0; fa < FM; ++fa)
            acc[fa] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[fa].v, b.v, acc[fa], 0, 0, 0);
    } else {
#pragma unroll
        for (int m = 0; m < 8; ++m)
#pragma unroll
            for (int fa = 0; fa < FM; ++fa)
                acc[fa] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[fa].v[m], b.v[m], acc[fa], 0, 0,
                                                               0);
    }
}

typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));   // dword-aligned 16 B
typedef __attribute__((address_space(3))) void lds_void_w;
typedef __attribute__((address_space(1))) const void gbl_void_w;

// 16-byte copy global -> LDS: lane l reads base + off (base uniform, off its
// own 32-bit offset) and writes LDS byte address lds_addr + 16 l (lds_addr
// uniform over the wave).  Written as asm so that the compiler does not see an LDS store: it
// would otherwise make every later LDS read of the other buffer wait for all
// copies.  The caller waits (vmcnt(0)) before the data are read.
__device__ __forceinline__ void lds_copy16(unsigned long long base, unsigned off,
                                           unsigned lds_addr)
{
    asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, %1"
                 :
                 : "v"(off), "s"(base), "s"(lds_addr)
                 : "memory", "m0");
}

// A wave-uniform value, marked as such (so that it is kept in SGPRs).
__device__ __forceinline__ unsigned uniform32(unsigned v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ unsigned long long uniform64(unsigned long long v)
{
    return ((unsigned long long)uniform32((unsigned)(v >> 32)) << 32) | uniform32((unsigned)v);
}

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// Lane i of each 16-lane row receives lane (i + N) mod 16 of the same row.
template <int N>
__device__ __forceinline__ int ror16(int v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x120 + N, 0xf, 0xf, false);   // row_ror:N
}
template <int N>
__device__ __forceinline__ long long ror16(long long v)
{
    const int lo = ror16<N>((int)(unsigned long long)v);
    const int hi = ror16<N>((int)((unsigned long long)v >> 32));
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
template <int N>
__device__ __forceinline__ double ror16(double v)
{
    return __builtin_bit_cast(double, ror16<N>(__builtin_bit_cast(long long, v)));
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
// epilogue output forms
constexpr int kEpiPlain = 0;   // diff2 from the two norms and the cross term
constexpr int kEpiInit = 1;    // the same, added to diff2_init
constexpr int kEpiCC = 2;      // cross-correlation score

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

    // ---- epilogue: float64 assembly, stores, per-row (min, argmin) partials.
    // Lane (r16, q) holds, for column slice fb, orientation o0 + rb0 + 16 fb +
    // r16 and the four consecutive rows R = R0 + 4q + v of fragment row fa.
    // When the four rows belong to one particle they are four consecutive
    // translations, i.e. 16 contiguous output bytes: one store (global
    // stores need only dword alignment).  A lane visits its columns in
    // increasing orientation, so a strict "<" keeps the lowest flat index
    // among equal minima of a row.
    const int C = a.n_classes;
    // Global loads cannot move past the stores (the compiler must assume
    // they alias), so every load of the common case -- the four rows of a
    // fragment row in one particle -- is issued ahead of the stores, for a
    // group of fragment rows, in two dependent rounds; the group's stores
    // then follow without waits.
    // (Two fragment rows at a time: all four would need more registers.)
    constexpr int EG = (FM >= 2 && FORM != kFP16) ? 2 : 1;   // fragment rows per load group
    // The three output forms -- diff2, diff2 added to diff2_init, and the
    // cross-correlation score -- are separate compile-time instances chosen
    // once per tile.  A per-element test would put every form's arithmetic
    // (the CC score's float64 division and square root among it) into the
    // instruction stream of the common form; every wave fetches the whole
    // epilogue once per tile, so its size is paid for on each tile.
    auto epilogue = [&](auto mode_c) {
        constexpr int MODE = decltype(mode_c)::value;
        int e_part[EG], e_b0[EG], e_T[EG];
        bool e_same[EG];
        double e_xn[EG];
        float e_inv[FORM == kFP16 ? EG : 1];
        float e_pn[EG][FN];
        auto epilogue_loads = [&](auto g_c) {
          static_for<0, EG>([&](auto e_c) {
            constexpr int e = decltype(e_c)::value;
            constexpr int fa = decltype(g_c)::value * EG + e;
            const int R0 = mtile * TM + (wm * FM + fa) * 16 + 4 * q;
            const int part = R0 < a.n_rows ? a.row_part[R0] : 0;
            const int part3 = R0 + 3 < a.n_rows ? a.row_part[R0 + 3] : -1;
            e_part[e] = part;
            e_same[e] = part3 == part;   // rows are sorted by particle
          });
          static_for<0, EG>([&](auto e_c) {
            constexpr int e = decltype(e_c)::value;
            const int part = e_part[e];
            e_b0[e] = a.trans_offset[part];
            e_T[e] = a.trans_offset[part + 1];
            e_xn[e] = a.xnorm[part];
            if constexpr (FORM == kFP16) e_inv[e] = 1.f / (a.ascale[part] * a.bscale[cls]);
#pragma unroll
            for (int fb = 0; fb < FN; ++fb) {
                const int o = min(o0 + rb0 + fb * 16 + r16, O - 1);
                e_pn[e][fb] = a.pnorm[((size_t)part * C + cls) * O + o];
            }
          });
        };
        // one float64 score
        auto score = [&](float cval, double pn, double xn, float init) -> double {
            if constexpr (MODE == kEpiCC) {
                return -(double)cval / sqrt(2.0 * pn);
            } else {
                double d = pn + xn - (double)cval;
                if constexpr (MODE == kEpiInit) d += (double)init;
                return d;
            }
        };
        // one call per group of EG fragment rows, with the fragment rows
        // compile-time constants so that the accumulators stay in registers
        auto epilogue_group = [&](auto g_c) {
            constexpr int g = decltype(g_c)::value;
            epilogue_loads(g_c);
            int R0[EG], rT[EG][4], rt[EG][4];
            bool rok[EG][4];
            double best[EG][4];
            int best_fb[EG];   // byte v: the column slice of row v's minimum, 0xff for none
            static_for<0, EG>([&](auto e_c) {
                constexpr int e = decltype(e_c)::value;
                R0[e] = mtile * TM + (wm * FM + g * EG + e) * 16 + 4 * q;
                const int T = e_T[e] - e_b0[e];
                best_fb[e] = -1;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    rok[e][v] = true;
                    rT[e][v] = T;
                    rt[e][v] = R0[e] + v - e_b0[e];
                    best[e][v] = DBL_MAX;
                }
            });
            auto set_fb = [](int &w, int v, int fb) { w = (w & ~(0xff << (8 * v))) | (fb << (8 * v)); };
            // Common case: the lane's four rows are four consecutive
            // translations of one particle, 16 contiguous output bytes.
            // Orientation-major, so that the group's fragment rows of one
            // orientation -- adjacent runs of one output row -- are stored
            // back to back.
            // output offset of (orientation o0 + rb0 + r16, the lane's first
            // row); column slice fb adds 16 fb orientations of T translations
            size_t k00[EG];
            static_for<0, EG>([&](auto e_c) {
                constexpr int e = decltype(e_c)::value;
                const int T = rT[e][0];
                k00[e] = (size_t)C * O * e_b0[e] + (size_t)cls * O * T +
                         (size_t)(o0 + rb0 + r16) * T + rt[e][0];
            });
#pragma unroll
            for (int fb = 0; fb < FN; ++fb) {
                const int o = o0 + rb0 + fb * 16 + r16;
                static_for<0, EG>([&](auto e_c) {
                    constexpr int e = decltype(e_c)::value;
                    constexpr int fa = g * EG + e;
                    if (o < O && e_same[e]) {
                        const size_t k0 = k00[e] + (size_t)(fb * 16 * rT[e][0]);
                        float init[4] = {0.f, 0.f, 0.f, 0.f};
                        if constexpr (MODE == kEpiInit) {
                            const f4u x = *reinterpret_cast<const f4u *>(a.diff2_init + k0);
                            init[0] = x.x;
                            init[1] = x.y;
                            init[2] = x.z;
                            init[3] = x.w;
                        }
                        const double pn = (double)e_pn[e][fb];
                        float out[4];
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const float cval =
                                acc[fb][fa][v] * (FORM == kFP16 ? e_inv[FORM == kFP16 ? e : 0] : 1.f);
                            const double d = score(cval, pn, e_xn[e], init[v]);
                            out[v] = (float)d;
                            if (d < best[e][v]) {
                                best[e][v] = d;
                                set_fb(best_fb[e], v, fb);
                            }
                        }
                        *reinterpret_cast<f4u *>(a.diff2 + k0) = f4u{out[0], out[1], out[2], out[3]};
                    }
                });
            }
            // A particle boundary (or the end of the batch) inside the lane's
            // four rows: row by row.
            static_for<0, EG>([&](auto e_c) {
                constexpr int e = decltype(e_c)::value;
                constexpr int fa = g * EG + e;
                if (e_same[e]) return;
                int rpart[4];
                size_t rb[4];
                double rxn[4];
                float rinv[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int R = R0[e] + v;
                    rok[e][v] = R < a.n_rows;
                    const int part = rok[e][v] ? a.row_part[R] : 0;
                    const int b0 = a.trans_offset[part];
                    rpart[v] = part;
                    rT[e][v] = a.trans_offset[part + 1] - b0;
                    rt[e][v] = R - b0;
                    rb[v] = (size_t)C * O * b0 + (size_t)cls * O * rT[e][v];
                    rxn[v] = a.xnorm[part];
                    rinv[v] = (FORM == kFP16) ? 1.f / (a.ascale[part] * a.bscale[cls]) : 1.f;
                }
#pragma unroll
                for (int fb = 0; fb < FN; ++fb) {
                    const int o = o0 + rb0 + fb * 16 + r16;
                    if (o >= O || !rok[e][0]) continue;
                    double pn[4];
                    float init[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        pn[v] = rok[e][v] ? (double)a.pnorm[((size_t)rpart[v] * C + cls) * O + o] : 0.0;
                        if (MODE == kEpiInit && rok[e][v])
                            init[v] = a.diff2_init[rb[v] + (size_t)o * rT[e][v] + rt[e][v]];
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const double d = score(acc[fb][fa][v] * rinv[v], pn[v], rxn[v], init[v]);
                        if (rok[e][v]) {
                            a.diff2[rb[v] + (size_t)o * rT[e][v] + rt[e][v]] = (float)d;
                            if (d < best[e][v]) {
                                best[e][v] = d;
                                set_fb(best_fb[e], v, fb);
                            }
                        }
                    }
                }
            });
            if (!a.part_min) return;
            // Per row, over the 16 lanes that share q (the columns): the
            // minimum value, then the lowest orientation among the lanes
            // holding it -- within a row the flat index grows with the
            // orientation, so this is the lowest index among equal minima.
            // Rotations 8, 4, 2, 1 within each 16-lane row (DPP moves on the
            // vector ALU) leave the result in every lane.
            static_for<0, EG>([&](auto e_c) {
                constexpr int e = decltype(e_c)::value;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    double bv = best[e][v];
                    static_for<0, 4>([&](auto st_c) {
                        const double ob = ror16<(8 >> decltype(st_c)::value)>(bv);
                        bv = ob < bv ? ob : bv;
                    });
                    const int bfb = (best_fb[e] >> (8 * v)) & 0xff;
                    int bo = bfb != 0xff && best[e][v] == bv ? rb0 + bfb * 16 + r16 : 0x7fffffff;
                    static_for<0, 4>([&](auto st_c) {
                        bo = min(bo, ror16<(8 >> decltype(st_c)::value)>(bo));
                    });
                    if (r16 == 0 && rok[e][v]) {
                        const size_t u = ((size_t)(R0[e] + v) * n_ntiles + ntile) * WN + wn;
                        a.part_min[u] = bv;
                        a.part_idx[u] = bo == 0x7fffffff
                                            ? 0x7fffffffffffffffLL
                                            : (long long)(cls * O + o0 + bo) * rT[e][v] + rt[e][v];
                    }
                }
            });
        };
        static_for<0, FM / EG>(epilogue_group);
    };
    if (a.cc)
        epilogue(std::integral_constant<int, kEpiCC>{});
    else if (a.diff2_init)
        epilogue(std::integral_constant<int, kEpiInit>{});
    else
        epilogue(std::integral_constant<int, kEpiPlain>{});
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
