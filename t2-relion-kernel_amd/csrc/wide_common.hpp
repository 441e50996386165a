// Helpers shared by the wide cross-term kernels (kernels_wide.hip,
// kernels_direct.hip): operand encodings and LDS swizzle, the translation
// operand build, fragment loads, the MFMA step of one column slice and the
// float64 epilogue.
#pragma once

#include <hip/hip_runtime.h>

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

// Direct layout (projections of the direct kernels): byte offset, inside the
// block of one 16-orientation group of one chunk, of 16-byte segment s of
// orientation ol.  The block holds the group's MFMA fragments as the lanes
// l = 16 q + r16 of a wave read them, 16 bytes per lane at l * 16 (BF16X3:
// the hi block, then the lo block; FP32: 32 bytes per lane at l * 32).
template <int FORM>
__host__ __device__ constexpr int direct_seg_offset(int ol, int s)
{
    return FORM == kBF16X3 ? ((s >> 2) * 64 + (s & 3) * 16 + ol) * 16
           : FORM == kFP16 ? (s * 16 + ol) * 16
                           : (((s >> 1) * 16 + ol) * 2 + (s & 1)) * 16;
}

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

// One translation-operand value in the words its form stores: BF16X3 the
// (hi, lo) words of (re, im); FP16 the scaled (re, im) halves; FP32 the
// floats themselves.  Encoding a value as soon as it is made (rather than
// a chunk's values at once) lets the conversions run under the MFMAs.
template <int FORM>
struct APix {
    float2 z;
};
template <>
struct APix<kBF16X3> {
    uint32_t hi, lo;
};
template <>
struct APix<kFP16> {
    uint32_t w;
};

template <int FORM>
__device__ __forceinline__ APix<FORM> a_encode(const ARow &ar, float2 z)
{
    APix<FORM> e;
    if constexpr (FORM == kBF16X3) {
        encode_bf16x3(z, e.hi, e.lo);
    } else if constexpr (FORM == kFP16) {
        const _Float16 re = (_Float16)(z.x * ar.scale), im = (_Float16)(z.y * ar.scale);
        e.w = (uint32_t)__builtin_bit_cast(uint16_t, re) |
              ((uint32_t)__builtin_bit_cast(uint16_t, im) << 16);
    } else {
        (void)ar;
        e.z = z;
    }
    return e;
}

// Stores the thread's PPT (8 or 4) encoded values -- pixels PPT*h ..
// PPT*h+PPT-1 of the chunk -- into LDS row buffer `dst` of row r.
template <int FORM, int PPT>
__device__ __forceinline__ void a_store_enc(const APix<FORM> (&e)[PPT], int h, int r,
                                            unsigned char *dst)
{
    static_assert(PPT == 8 || PPT == 4, "8 or 4 pixels per thread");
    constexpr int G = PPT / 4;   // 4-pixel groups (one 16-byte segment each)
    if constexpr (FORM == kBF16X3) {
        uint32_t hi[PPT], lo[PPT];
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            hi[j] = e[j].hi;
            lo[j] = e[j].lo;
        }
        a_store_words<PPT>(hi, lo, h, r, dst);
    } else if constexpr (FORM == kFP16) {
#pragma unroll
        for (int half = 0; half < G; ++half)
            *reinterpret_cast<uint4 *>(dst + wide_swz<FORM>(r, G * h + half) * 16) =
                make_uint4(e[4 * half].w, e[4 * half + 1].w, e[4 * half + 2].w,
                           e[4 * half + 3].w);
    } else if constexpr (PPT == 8) {
        // segment 2q + h holds component q % 2 of pixels 2i + q / 2, i = 0..3
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float f[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float2 z = e[2 * i + (q >> 1)].z;
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
            const float2 z0 = e[q >> 1].z, z1 = e[2 + (q >> 1)].z;
            const float2 f = (q & 1) ? make_float2(z0.y, z1.y) : make_float2(z0.x, z1.x);
            *reinterpret_cast<float2 *>(dst + wide_swz<FORM>(r, 2 * q + (h >> 1)) * 16 +
                                        (h & 1) * 8) = f;
        }
    }
}

// Encodes the thread's PPT values and stores them (a_encode + a_store_enc).
template <int FORM, int PPT>
__device__ __forceinline__ void a_store(const ARow &ar, const float2 (&v)[PPT], int h, int r,
                                        unsigned char *dst)
{
    APix<FORM> e[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) e[j] = a_encode<FORM>(ar, v[j]);
    a_store_enc<FORM, PPT>(e, h, r, dst);
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

// Epilogue output forms.
constexpr int kEpiPlain = 0;   // diff2 from the two norms and the cross term
constexpr int kEpiInit = 1;    // the same, added to diff2_init
constexpr int kEpiCC = 2;      // cross-correlation score

// Inputs of the epilogue, read from the plan's arrays in global memory.
// A row's key is its particle; pn(key, fb) is sum corr/2 |P|^2 of the
// lane's orientation o0 + rb0 + 16 fb + r16 (clamped to the last one).
struct EpiGlobal {
    const CrossArgs &a;
    int cls, o0, rb0, r16;
    __device__ __forceinline__ int key(int R) const { return a.row_part[R]; }
    __device__ __forceinline__ int b0(int k) const { return a.trans_offset[k]; }
    __device__ __forceinline__ int end(int k) const { return a.trans_offset[k + 1]; }
    __device__ __forceinline__ double xn(int k) const { return a.xnorm[k]; }
    __device__ __forceinline__ float inv(int k) const
    {
        return 1.f / (a.ascale[k] * a.bscale[cls]);
    }
    __device__ __forceinline__ float pn(int k, int fb) const
    {
        const int O = a.orientation_num;
        const int o = min(o0 + rb0 + fb * 16 + r16, O - 1);
        return a.pnorm[((size_t)k * a.n_classes + cls) * O + o];
    }
};

// Epilogue of one wave of a wide tile: float64 assembly, stores, per-row
// (min, argmin) partials.  The wave holds FM x FN 16x16 accumulators:
// rows mtile * TM + (wm * FM + fa) * 16 + 0..15, columns o0 + rb0 +
// fb * 16 + 0..15 of class cls; its partials go to slot wn of WN.  SRC
// provides the per-row and per-column inputs (EpiGlobal, or tables).
template <int FORM, int FM, int FN, int WN, int TM, typename SRC>
__device__ __forceinline__ void wide_epilogue(const CrossArgs &a, floatx4 (&acc)[FN][FM],
                                              int mtile, int ntile, int n_ntiles, int cls,
                                              int o0, int wm, int wn, int rb0, int r16,
                                              int q, const SRC &src)
{
    const int O = a.orientation_num;
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
// Row partials: after the 16-lane reduction every lane of a row group holds
// the row's (min, index); row j = 4 fa + v of lane group q is stored by lane
// r16 = j % 16 from its slot j / 16, so that a wave stores its partials with
// NS full-wave stores instead of one four-lane store per row.
constexpr int NS = (4 * FM + 15) / 16;
auto epilogue = [&](auto mode_c) {
    constexpr int MODE = decltype(mode_c)::value;
    double mine_min[NS];
    long long mine_idx[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        mine_min[s] = DBL_MAX;
        mine_idx[s] = 0x7fffffffffffffffLL;
    }
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
        const int part = R0 < a.n_rows ? src.key(R0) : 0;
        const int part3 = R0 + 3 < a.n_rows ? src.key(R0 + 3) : -1;
        e_part[e] = part;
        e_same[e] = part3 == part;   // rows are sorted by particle
      });
      static_for<0, EG>([&](auto e_c) {
        constexpr int e = decltype(e_c)::value;
        const int part = e_part[e];
        e_b0[e] = src.b0(part);
        e_T[e] = src.end(part);
        e_xn[e] = src.xn(part);
        if constexpr (FORM == kFP16) e_inv[e] = src.inv(part);
#pragma unroll
        for (int fb = 0; fb < FN; ++fb) e_pn[e][fb] = src.pn(part, fb);
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
        int best_fb[EG][4];   // the column slice of row v's minimum, -1 for none
        static_for<0, EG>([&](auto e_c) {
            constexpr int e = decltype(e_c)::value;
            R0[e] = mtile * TM + (wm * FM + g * EG + e) * 16 + 4 * q;
            const int T = e_T[e] - e_b0[e];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                rok[e][v] = true;
                rT[e][v] = T;
                rt[e][v] = R0[e] + v - e_b0[e];
                best[e][v] = DBL_MAX;
                best_fb[e][v] = -1;
            }
        });
        // Common case: the lane's four rows are four consecutive
        // translations of one particle, 16 contiguous output bytes.
        // Orientation-major, so that the group's fragment rows of one
        // orientation -- adjacent runs of one output row -- are stored
        // back to back.
        // output address of (orientation o0 + rb0 + r16, the lane's first
        // row); column slice fb adds 16 fb orientations of T translations
        float *p00[EG];
        int fb_step[EG];
        static_for<0, EG>([&](auto e_c) {
            constexpr int e = decltype(e_c)::value;
            const int T = rT[e][0];
            p00[e] = a.diff2 + ((size_t)C * O * e_b0[e] + (size_t)cls * O * T +
                                (size_t)(o0 + rb0 + r16) * T + rt[e][0]);
            fb_step[e] = 16 * T;
        });
#pragma unroll
        for (int fb = 0; fb < FN; ++fb) {
            const int o = o0 + rb0 + fb * 16 + r16;
            static_for<0, EG>([&](auto e_c) {
                constexpr int e = decltype(e_c)::value;
                constexpr int fa = g * EG + e;
                if (o < O && e_same[e]) {
                    float *const dst = p00[e] + fb * fb_step[e];
                    float init[4] = {0.f, 0.f, 0.f, 0.f};
                    if constexpr (MODE == kEpiInit) {
                        const f4u x = *reinterpret_cast<const f4u *>(
                            a.diff2_init + (dst - a.diff2));
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
                        const bool lt = d < best[e][v];
                        best[e][v] = lt ? d : best[e][v];
                        best_fb[e][v] = lt ? fb : best_fb[e][v];
                    }
                    *reinterpret_cast<f4u *>(dst) = f4u{out[0], out[1], out[2], out[3]};
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
                const int part = rok[e][v] ? src.key(R) : 0;
                const int b0 = src.b0(part);
                rpart[v] = part;
                rT[e][v] = src.end(part) - b0;
                rt[e][v] = R - b0;
                rb[v] = (size_t)C * O * b0 + (size_t)cls * O * rT[e][v];
                rxn[v] = src.xn(part);
                rinv[v] = (FORM == kFP16) ? src.inv(part) : 1.f;
            }
#pragma unroll
            for (int fb = 0; fb < FN; ++fb) {
                const int o = o0 + rb0 + fb * 16 + r16;
                if (o >= O || !rok[e][0]) continue;
                double pn[4];
                float init[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    pn[v] = rok[e][v] ? (double)src.pn(rpart[v], fb) : 0.0;
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
                            best_fb[e][v] = fb;
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
                const int bfb = best_fb[e][v];
                int bo = bfb >= 0 && best[e][v] == bv ? rb0 + bfb * 16 + r16 : 0x7fffffff;
                static_for<0, 4>([&](auto st_c) {
                    bo = min(bo, ror16<(8 >> decltype(st_c)::value)>(bo));
                });
                constexpr int j0 = 4 * (g * EG + e);   // row j0 + v of the lane group
                const long long idx = bo == 0x7fffffff
                                          ? 0x7fffffffffffffffLL
                                          : (long long)(cls * O + o0 + bo) * rT[e][v] + rt[e][v];
                const bool mine = r16 == ((j0 + v) & 15);
                mine_min[(j0 + v) >> 4] = mine ? bv : mine_min[(j0 + v) >> 4];
                mine_idx[(j0 + v) >> 4] = mine ? idx : mine_idx[(j0 + v) >> 4];
            }
        });
    };
    static_for<0, FM / EG>(epilogue_group);
    if (!a.part_min) return;
#pragma unroll
    for (int sl = 0; sl < NS; ++sl) {
        const int j = sl * 16 + r16;   // row 4 fa + v of the lane group
        const int R = mtile * TM + (wm * FM + (j >> 2)) * 16 + 4 * q + (j & 3);
        if (j < 4 * FM && R < a.n_rows) {
            const size_t u = ((size_t)R * n_ntiles + ntile) * WN + wn;
            a.part_min[u] = mine_min[sl];
            a.part_idx[u] = mine_idx[sl];
        }
    }
};
if (a.cc)
    epilogue(std::integral_constant<int, kEpiCC>{});
else if (a.diff2_init)
    epilogue(std::integral_constant<int, kEpiInit>{});
else
    epilogue(std::integral_constant<int, kEpiPlain>{});
}

}  // namespace rc
