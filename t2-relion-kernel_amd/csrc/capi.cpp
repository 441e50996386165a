// Host runtime behind the C ABI of include/relion_coarse.h: projector
// objects, argument validation, workspace layout, kernel selection and the
// batched plans.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <tuple>
#include <mutex>
#include <string>
#include <vector>

#include "plan_internal.hpp"
#include "rc_internal.hpp"

namespace rc {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

static float pow2_scale_host(float m)
{
    if (!(m > 0.f)) return 1.f;
    int e;
    std::frexp(m, &e);
    return std::ldexp(1.f, -e);
}

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Bump allocator over one device buffer.
struct Carve {
    unsigned char *base;
    size_t off = 0;
    template <class T>
    T *take(size_t count)
    {
        off = align_up(off, 256);
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off += count * sizeof(T);
        return p;
    }
};

static ProjView view_of(const rc_projector_kernel &k)
{
    ProjView v;
    v.vol = k.proj->vol;
    v.mdlX = k.proj->xdim;
    v.mdlY = k.proj->ydim;
    v.mdlZ = k.proj->zdim;
    v.mdlXY = k.proj->xdim * k.proj->ydim;
    v.initY = k.proj->inity;
    v.initZ = k.proj->initz;
    v.maxR2_padded = k.maxR2_padded;
    v.pad = k.padding_factor;
    return v;
}

static int check_kernel_view(const rc_projector_kernel *k)
{
    RC_REQUIRE(k && k->proj && k->proj->vol, RC_ERR_INVALID, "projector kernel view is NULL");
    RC_REQUIRE(k->imgX > 0 && k->imgY > 0, RC_ERR_INVALID, "image dimensions must be positive");
    return RC_OK;
}

// ---------------------------------------------------------------------------
// single-image workspace and split-K
// ---------------------------------------------------------------------------
// One image alone gives few tiles (C1: 3, a 113-translation particle at 4608
// orientations: 36), far fewer than the 256 CUs.  The fused kernel then
// splits the pixel axis into slices that each write float32 partial sums
// (cross term per (orientation, translation), |P|^2 norm per orientation)
// to the workspace, and a second kernel adds the slices in a fixed order in
// float64.  The slice count follows from the shape alone -- never from the
// device or from timing -- so a shape always gives the same bits.
// The slice count is the largest that keeps the launch within
// kSplitTargetBlocks workgroups -- two per CU of the 256 -- rather than the
// smallest that reaches it: at one 113 x 4608 x 2112 particle (72 tiles) 7
// slices (504 workgroups) take 76 us per call where 8 (576: a third
// workgroup on some CUs) take 83, 14 (1008) 86 and 3 (216) 90.
constexpr int kSplitTargetBlocks = 512;            // workgroups the split stays within
constexpr size_t kSplitMaxBytes = 32ull << 20;     // cap of the cross-term partials

struct SingleShape {
    int variant;             // fused cross-kernel variant
    int tiles;               // row tiles x column tiles
    int nkc;                 // 32-pixel chunks
    int ksplit, kc_per_split;
};

// Fused tile of the single-image calls, by measurement at 21 x 576 x 544 and
// 113 x 4608 x 2112 px: 64 x 128 for the FP32 form (the drop-in's, whose
// projections usually come from the cache, 1.4x faster than 128 x 128 at
// 113 translations) and for up to 64 translations; 128 x 128 for the split
// forms beyond, where the fused gather of each row tile's projections is
// what costs (BF16x3: 287 against 386 us).
static int pick_fused_variant(int form, int T)
{
    if (form == RC_FORM_FP32 || T <= 64) return 1;   // 64 x 128
    return 0;                                        // 128 x 128
}

static SingleShape single_shape(int T, int O, int npix, int form)
{
    SingleShape sh{};
    sh.nkc = (int)(align_up((size_t)std::max(npix, 1), kChunk) / kChunk);
    sh.ksplit = 1;
    sh.kc_per_split = sh.nkc;
    if (form == RC_FORM_FP32_DIRECT || T <= 0 || O <= 0) return sh;
    sh.variant = pick_fused_variant(form, T);
    const VariantInfo *info = variant_info(form, true, sh.variant);
    const long long rows = (long long)((T + 15) / 16) * 16;
    const long long tiles = ((rows + info->tile_m - 1) / info->tile_m) *
                            (((long long)O + info->tile_n - 1) / info->tile_n);
    sh.tiles = (int)std::min<long long>(tiles, 1 << 30);
    long long want = kSplitTargetBlocks / tiles;
    const long long by_bytes = (long long)(kSplitMaxBytes / ((size_t)T * O * sizeof(float)));
    want = std::max(1ll, std::min({want, (long long)sh.nkc, by_bytes}));
    sh.kc_per_split = (int)((sh.nkc + want - 1) / want);
    sh.ksplit = (sh.nkc + sh.kc_per_split - 1) / sh.kc_per_split;   // no empty slice
    return sh;
}

// Device buffers of the drop-in's projection cache (see ScratchEntry).
struct ProjCacheView {
    float *eulers;   // copy of the eulers the projections were made from
    unsigned *flag;  // set to gen by the check kernel when they must be made again
    unsigned gen;    // this call's generation (one more than the last call's)
    float2 *P;       // [O][kpad]
};

struct SingleWs {
    float2 *xw, *coords;
    double *xnorm;
    float *half_corr, *ascale;
    float *part_c, *part_pn;     // split-K partials (NULL without a split)
    size_t bytes;
};

static SingleWs single_layout(unsigned char *base, int npix, int T, int O,
                              const SingleShape &sh)
{
    const int kpad = (int)align_up((size_t)std::max(npix, 1), kChunk);
    Carve c{base};
    SingleWs w;
    w.xw = c.take<float2>(kpad);
    w.coords = c.take<float2>(kpad);
    w.xnorm = c.take<double>(1);
    w.half_corr = c.take<float>(kpad);
    w.ascale = c.take<float>(1);
    w.part_c = w.part_pn = nullptr;
    if (sh.ksplit > 1) {
        w.part_c = c.take<float>((size_t)sh.ksplit * O * T);
        w.part_pn = c.take<float>((size_t)sh.ksplit * O);
    }
    w.bytes = align_up(c.off, 256);
    return w;
}

static int run_single(int T, int O, int npix, int full_npix, int form, const float *eulers,
                      const float *trans_x, const float *trans_y, const float *real,
                      const float *imag, const rc_projector_kernel *pk, const float *corr,
                      const float *init, float *dest, hipStream_t stream, const float *coor_xy,
                      void *ws, size_t ws_bytes, bool cc = false,
                      const ProjCacheView *cache = nullptr, int cache_force = 0,
                      bool partition = false)
{
    if (int st = check_kernel_view(pk)) return st;
    RC_REQUIRE(T > 0 && O > 0 && npix > 0, RC_ERR_INVALID,
               "translation_num, orientation_num and image_size must be positive (got %d, %d, %d)",
               T, O, npix);
    RC_REQUIRE(full_npix >= npix, RC_ERR_INVALID, "full_image_size %d < image_size %d",
               full_npix, npix);
    RC_REQUIRE(npix <= kMaxPix && T <= kMaxTrans, RC_ERR_LIMIT,
               "image_size %d or translation_num %d beyond limits (%d, %d)", npix, T, kMaxPix,
               kMaxTrans);
    RC_REQUIRE(form >= RC_FORM_FP32 && form <= RC_FORM_FP32_DIRECT, RC_ERR_INVALID,
               "unknown form %d", form);
    RC_REQUIRE(eulers && trans_x && trans_y && real && imag && corr && dest, RC_ERR_INVALID,
               "NULL device pointer");
    const SingleShape sh = single_shape(T, O, npix, form);
    SingleWs w = single_layout((unsigned char *)ws, npix, T, O, sh);
    RC_REQUIRE(ws && ws_bytes >= w.bytes, RC_ERR_INVALID,
               "workspace of %zu bytes is smaller than the %zu bytes needed", ws_bytes, w.bytes);
    const int kpad = (int)align_up((size_t)npix, kChunk);
    ProjView v = view_of(*pk);
    if (partition) {
        // npix = the pixels a rotation can bring inside the mask, of the
        // full_npix of the unfiltered image
        float pad2 = 0.f, thr = 0.f;
        RC_REQUIRE(dropin_pixels(*pk, full_npix, cc, &pad2, &thr) == npix, RC_ERR_INVALID,
                   "partitioned pixel count %d does not match the geometry", npix);
        RC_HIP(launch_prep_partition(full_npix, kpad, real, imag, corr, pk->imgX, pk->imgY,
                                     pk->maxR, pad2, thr, w.xw, w.coords, w.xnorm, w.half_corr,
                                     w.ascale, stream, cc));
    } else {
        RC_HIP(launch_prep(1, npix, full_npix, kpad, real, imag, corr, coor_xy, pk->imgX, pk->imgY,
                           pk->maxR, w.xw, w.coords, w.xnorm, w.half_corr, w.ascale, stream, cc));
    }
    if (form == RC_FORM_FP32_DIRECT) {
        RC_REQUIRE(!cc, RC_ERR_UNSUPPORTED, "the direct form has no cross-correlation mode");
        RC_REQUIRE(T <= 1024, RC_ERR_LIMIT, "direct form supports at most 1024 translations");
        RC_HIP(launch_direct(v, eulers, O, trans_x, trans_y, T, w.coords, real, imag, corr, npix,
                             full_npix, init, dest, stream));
        return RC_OK;
    }
    const int vi = sh.variant;
    const VariantInfo *info = variant_info(form, true, vi);
    CrossArgs a{};
    a.groups = nullptr;
    a.n_groups = (T + 15) / 16;
    a.trans_offset = nullptr;
    a.n_trans = T;
    a.trans_x = trans_x;
    a.trans_y = trans_y;
    a.xw = w.xw;
    a.coords = w.coords;
    a.xnorm = w.xnorm;
    a.ascale = w.ascale;
    a.kpad = kpad;
    a.npix = npix;
    a.bscale = pk->proj->d_scale;
    a.half_corr = w.half_corr;
    a.eulers = eulers;
    a.vol[0] = v;
    a.n_classes = 1;
    a.orientation_num = O;
    a.ntiles_per_class = (O + info->tile_n - 1) / info->tile_n;
    a.n_mtiles = (a.n_groups * 16 + info->tile_m - 1) / info->tile_m;
    a.diff2_init = init;
    a.diff2 = dest;
    a.part_min = nullptr;
    a.part_idx = nullptr;
    a.cc = cc ? 1 : 0;
    a.ksplit = sh.ksplit;
    a.kc_per_split = sh.kc_per_split;
    a.part_c = w.part_c;
    a.part_pn = w.part_pn;
    a.pcache = nullptr;
    if (cache) {
        RC_HIP(launch_cache_check(eulers, cache->eulers, 9 * O, cache_force, cache->gen,
                                  cache->flag, stream));
        RC_HIP(launch_cache_fill(cache->flag, cache->gen, v, eulers, O, w.coords, npix, kpad,
                                 cache->P, stream));
        a.pcache = cache->P;
    }
    RC_HIP(launch_cross(form, true, vi, a, stream));
    if (sh.ksplit > 1)
        RC_HIP(launch_splitk_assemble(O, T, sh.ksplit, w.part_c, w.part_pn, w.xnorm, init, dest,
                                      cc ? 1 : 0, stream));
    return RC_OK;
}

size_t single_workspace_bytes(int T, int O, int npix, int form)
{
    if (npix <= 0 || T <= 0 || O <= 0) return 0;
    return single_layout(nullptr, npix, T, O, single_shape(T, O, npix, form)).bytes;
}

int single_split(int T, int O, int npix, int form, int *ksplit, int *kc_per_split)
{
    const SingleShape sh = single_shape(T, O, npix, form);
    if (ksplit) *ksplit = sh.ksplit;
    if (kc_per_split) *kc_per_split = sh.kc_per_split;
    return RC_OK;
}

int run_single_ws(int T, int O, int npix, int full_npix, int form, const float *eulers,
                  const float *trans_x, const float *trans_y, const float *real, const float *imag,
                  const rc_projector_kernel *pk, const float *corr, const float *init, float *dest,
                  hipStream_t stream, const float *coor_xy, void *ws, size_t ws_bytes, bool cc,
                  bool partition)
{
    return run_single(T, O, npix, full_npix, form, eulers, trans_x, trans_y, real, imag, pk, corr,
                      init, dest, stream, coor_xy, ws, ws_bytes, cc, nullptr, 0, partition);
}

int dropin_pixels(const rc_projector_kernel &k, int image_size, bool cc, float *pad2,
                  float *thr)
{
    const float p2 = k.padding_factor * k.padding_factor;
    const float t = ((float)k.maxR2_padded + 1.0f) * (1.0f + 1.0f / 4096.0f);
    *pad2 = p2;
    *thr = t;
    // counted once per geometry (an image of 2^20 pixels takes a millisecond)
    struct Key {
        int imgX, imgY, maxR, n, cc;
        float p2, t;
        bool operator<(const Key &o) const
        {
            return std::tie(imgX, imgY, maxR, n, cc, p2, t) <
                   std::tie(o.imgX, o.imgY, o.maxR, o.n, o.cc, o.p2, o.t);
        }
    };
    static std::mutex mu;
    static std::map<Key, int> counts;
    const Key key{k.imgX, k.imgY, k.maxR, image_size, cc ? 1 : 0, p2, t};
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = counts.find(key);
        if (it != counts.end()) return it->second;
    }
    int n = 0;
    for (int p = 0; p < image_size; ++p) {
        int x = p % k.imgX, y = p / k.imgX;
        if (y > k.maxR) {
            if (!cc || y >= k.imgY - k.maxR)
                y -= k.imgY;
            else
                x = k.maxR;
        }
        const volatile float r2 = (float)(x * x + y * y) * p2;   // one float multiply, as on the device
        n += r2 <= t;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (counts.size() > 256) counts.clear();
    counts[key] = n;
    return n;
}

// Scratch of rc_run_diff2_kernel_coarse, one buffer per (device, stream):
// calls on different streams never share memory, and calls on one stream are
// ordered by that stream.  The entry's mutex is held while a call enqueues,
// so a call that grows the buffer cannot free it under another thread that
// is still launching with it; the old buffer is freed only after the stream
// has drained.
// The entry also keeps the drop-in's projection cache: RELION calls the
// drop-in once per particle with one projector and one orientation set, so
// the projections of a call are those of the last one unless the key
// changed.  The host compares what it can see (projector, its generation,
// sizes, image geometry, the eulers pointer); a device kernel compares the
// eulers' contents with a copy, so a buffer refilled with other
// orientations is caught without synchronising; the fill kernel recomputes
// the projections only when either differs.
struct ProjCacheKey {
    const rc_projector *proj = nullptr;
    unsigned long long generation = 0;
    const float *eulers = nullptr;
    int O = 0, npix = 0, full = 0, imgX = 0, imgY = 0, maxR = 0, maxR2 = 0;
    bool operator==(const ProjCacheKey &o) const
    {
        return proj == o.proj && generation == o.generation && eulers == o.eulers && O == o.O &&
               npix == o.npix && full == o.full && imgX == o.imgX && imgY == o.imgY &&
               maxR == o.maxR && maxR2 == o.maxR2;
    }
};
struct ScratchEntry {
    std::mutex mu;
    void *ptr = nullptr;
    size_t bytes = 0;
    // projection cache: [eulers copy][flag][P: O x kpad float2]
    void *cache = nullptr;
    size_t cache_bytes = 0;
    ProjCacheKey key;
    bool cache_valid = false;
    unsigned cache_gen = 0;   // generation of the last call (see ProjCacheView)
};
// Largest projection cache per stream (float2 per orientation and pixel);
// larger searches gather their projections every call.
constexpr size_t kProjCacheMaxBytes = 1ull << 30;
static std::mutex g_scratch_mu;
static std::map<std::pair<int, hipStream_t>, std::unique_ptr<ScratchEntry>> g_scratch;

static int scratch_for(hipStream_t stream, size_t bytes, std::unique_lock<std::mutex> &held,
                       void **out, ScratchEntry **entry = nullptr)
{
    int dev = 0;
    RC_HIP(hipGetDevice(&dev));
    ScratchEntry *e;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        auto &slot = g_scratch[{dev, stream}];
        if (!slot) slot.reset(new ScratchEntry);
        e = slot.get();
    }
    held = std::unique_lock<std::mutex>(e->mu);
    if (e->bytes < bytes) {
        if (e->ptr) {
            RC_HIP(hipStreamSynchronize(stream));
            RC_HIP(hipFree(e->ptr));
            e->ptr = nullptr;
            e->bytes = 0;
        }
        const size_t want = align_up(bytes, 1 << 20);
        void *p = nullptr;
        if (hipMalloc(&p, want) != hipSuccess) {
            set_error("hipMalloc of %zu scratch bytes failed", want);
            return RC_ERR_NO_MEMORY;
        }
        e->ptr = p;
        e->bytes = want;
    }
    *out = e->ptr;
    if (entry) *entry = e;
    return RC_OK;
}

// Projection-cache buffers of a held scratch entry for O orientations of
// kpad pixels (grown like the scratch: the old buffer is freed after the
// stream drains).  *force is set when the cache holds nothing for this key.
static int proj_cache_for(ScratchEntry *e, hipStream_t stream, const ProjCacheKey &key, int kpad,
                          ProjCacheView *view, int *force)
{
    const size_t eb = align_up((size_t)9 * key.O * sizeof(float), 256);
    const size_t need = eb + 256 + (size_t)key.O * kpad * sizeof(float2);
    if (e->cache_bytes < need) {
        if (e->cache) {
            RC_HIP(hipStreamSynchronize(stream));
            RC_HIP(hipFree(e->cache));
            e->cache = nullptr;
            e->cache_bytes = 0;
        }
        const size_t want = align_up(need, 1 << 20);
        void *p = nullptr;
        if (hipMalloc(&p, want) != hipSuccess) {
            set_error("hipMalloc of %zu projection-cache bytes failed", want);
            return RC_ERR_NO_MEMORY;
        }
        e->cache = p;
        e->cache_bytes = want;
        e->cache_valid = false;
    }
    unsigned char *b = static_cast<unsigned char *>(e->cache);
    view->eulers = reinterpret_cast<float *>(b);
    view->flag = reinterpret_cast<unsigned *>(b + eb);
    view->gen = ++e->cache_gen;
    view->P = reinterpret_cast<float2 *>(b + eb + 256);
    *force = !(e->cache_valid && e->key == key);
    e->key = key;
    e->cache_valid = true;
    return RC_OK;
}

static int scratch_release(hipStream_t stream)
{
    int dev = 0;
    RC_HIP(hipGetDevice(&dev));
    std::unique_ptr<ScratchEntry> e;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        auto it = g_scratch.find({dev, stream});
        if (it == g_scratch.end()) return RC_OK;
        e = std::move(it->second);
        g_scratch.erase(it);
    }
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->ptr || e->cache) RC_HIP(hipStreamSynchronize(stream));
    if (e->ptr) RC_HIP(hipFree(e->ptr));
    if (e->cache) RC_HIP(hipFree(e->cache));
    return RC_OK;
}

int check_drop_in(const rc_projector_kernel *pk, unsigned long O, unsigned long T,
                  unsigned long image_size, int data_is_3D)
{
    RC_REQUIRE(!data_is_3D, RC_ERR_UNSUPPORTED, "3D data (data_is_3D) is not implemented");
    if (int st = check_kernel_view(pk)) return st;
    RC_REQUIRE(O > 0 && T > 0 && image_size > 0 && O < (1ul << 30) &&
                   T <= (unsigned long)kMaxTrans && image_size <= (unsigned long)kMaxPix,
               RC_ERR_LIMIT, "sizes out of range");
    RC_REQUIRE(image_size <= (unsigned long)pk->imgX * pk->imgY, RC_ERR_INVALID,
               "image_size %lu exceeds imgX*imgY", image_size);
    return RC_OK;
}

// Host statement of the in-radius rule of FilterCoarseParamImage
// (main.cu:88-151): int (x*x + y*y), then float products with the padding
// factor, compared with maxR2_padded as a float.
static bool filter_inside_host(int i, const rc_projector_kernel &k, int &x, int &y)
{
    x = i % k.imgX;
    y = i / k.imgX;
    if (y > k.maxR) y -= k.imgY;
    const float r2 = (float)(x * x + y * y) * k.padding_factor * k.padding_factor;
    return r2 <= (float)k.maxR2_padded;
}

// ---------------------------------------------------------------------------
// batched plans
// ---------------------------------------------------------------------------
// Variant forced on every later rc_plan_create (rc_set_variant_override);
// -1 = automatic.  Atomic: plans may be created from several host threads.
static std::atomic<int> g_variant_override{-1};

struct TuneKey {
    int P, C, O, npix, form, groups;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(P, C, O, npix, form, groups) <
               std::tie(o.P, o.C, o.O, o.npix, o.form, o.groups);
    }
};
static std::mutex g_tune_mu;
static std::map<TuneKey, int> g_tuned;

static int count_groups(const rc_batch_desc *d)
{
    int g = 0;
    for (int i = 0; i < d->n_particles; ++i)
        g += (d->host_trans_offset[i + 1] - d->host_trans_offset[i] + 15) / 16;
    return g;
}

static TuneKey tune_key(const rc_batch_desc *d)
{
    return TuneKey{d->n_particles, d->n_classes, d->orientation_num, d->npix, d->form,
                   count_groups(d)};
}

}  // namespace rc


namespace rc {

// Variant used when the autotuner has not seen a shape.  Measured on MI355X
// at 512 particles x 113 translations x 4608 orientations x 1636 pixels:
// BF16X3 and FP16 run fastest on the direct 128 x 512 kernel (projections
// loaded into registers; BF16X3 3.8-3.9 ms, 2-4 % ahead of the wide
// 128 x 512 kernels; FP16 2.0 ms, 12 % ahead), which takes batches whose
// 128-row tiles hold at most kDirectMaxTileParts particles, else the wide
// 128 x 512 kernel with fenced segments; FP32 on the wide 256 x 256 kernel
// with the translation operand precomputed (12.4 vs 13.2 ms: its longer
// MFMA phase hides the operand copies, while the VALU build competes with
// its FP32 MFMAs).  These need enough tiles to fill the device and, for
// 512 columns, little column waste; then the wide 256 x 256 kernel with
// staggered waves; smaller searches use the pipelined 128 x 128 kernel
// (FP16: 64 x 256), whose finer tiles waste less padding.  rc_autotune
// overrides this per shape.
static int default_variant(int form, int n_rows, int n_classes, int O, int tile128_parts)
{
    constexpr int kWide256 = 14;     // wide, 256 x 256, staggered waves
    constexpr int kWide512 = 21;     // wide, 128 x 512, fenced segments
    constexpr int kWide256Pre = 23;  // wide, 256 x 256, translation operand precomputed
    constexpr int kDirect = 24;      // direct, 128 x 512
    const long long rows = n_rows;   // the wide kernels pack the rows
    const long long t512 = ((rows + 127) / 128) * n_classes * ((O + 511) / 512);
    const long long t256 = ((rows + 255) / 256) * n_classes * ((O + 255) / 256);
    const bool narrow_waste = (long long)((O + 511) / 512) * 512 > O + O / 8;
    if (form == kFP32 && t256 >= 1024) return kWide256Pre;
    if (t512 >= 1024 && !narrow_waste)
        return tile128_parts <= kDirectMaxTileParts ? kDirect : kWide512;
    if (t256 >= 512) return kWide256;
    return form == kFP16 ? 7 : 4;
}

int validate_batch(const rc_batch_desc *d, bool need_ptrs)
{
    RC_REQUIRE(d, RC_ERR_INVALID, "NULL batch descriptor");
    RC_REQUIRE(d->n_particles > 0 && d->n_classes > 0 && d->orientation_num > 0 && d->npix > 0,
               RC_ERR_INVALID, "particles, classes, orientations and npix must be positive");
    RC_REQUIRE(d->full_npix >= d->npix, RC_ERR_INVALID, "full_npix < npix");
    RC_REQUIRE(d->npix <= kMaxPix, RC_ERR_LIMIT, "npix %d beyond %d", d->npix, kMaxPix);
    RC_REQUIRE(d->form == RC_FORM_FP32 || d->form == RC_FORM_BF16X3 || d->form == RC_FORM_FP16,
               RC_ERR_UNSUPPORTED, "batched path supports FP32, BF16X3 and FP16 forms (got %d)",
               d->form);
    RC_REQUIRE(d->projectors && d->host_trans_offset, RC_ERR_INVALID,
               "projectors and host_trans_offset are required");
    for (int c = 0; c < d->n_classes; ++c)
        if (int st = check_kernel_view(&d->projectors[c])) return st;
    RC_REQUIRE(d->host_trans_offset[0] == 0, RC_ERR_INVALID, "trans_offset[0] must be 0");
    for (int i = 0; i < d->n_particles; ++i) {
        const int T = d->host_trans_offset[i + 1] - d->host_trans_offset[i];
        RC_REQUIRE(T > 0 && T <= kMaxTrans, RC_ERR_INVALID,
                   "particle %d has %d translations (must be 1..%d)", i, T, kMaxTrans);
    }
    if (!need_ptrs) return RC_OK;
    RC_REQUIRE(d->d_trans_offset && d->d_trans_x && d->d_trans_y && d->d_eulers && d->d_real &&
                   d->d_imag && d->d_corr && d->d_coor_xy && d->d_diff2,
               RC_ERR_INVALID, "NULL device pointer in batch descriptor");
    return RC_OK;
}

size_t plan_bind(rc_plan *p, unsigned char *base)
{
    Carve c{base};
    const size_t nt_total = (size_t)p->C * p->ntiles_per_class;
    p->groups = c.take<int2>(p->n_groups);
    p->group_begin = c.take<int>(p->P + 1);
    p->row_part = c.take<int>(p->n_rows);
    p->vols = c.take<ProjView>(p->C);
    p->bscale = c.take<float>(p->C);
    p->xw = c.take<float2>((size_t)p->P * p->kpad);
    p->coords = c.take<float2>(p->kpad);
    p->xnorm = c.take<double>(p->P);
    p->half_corr = c.take<float>((size_t)p->P * p->kpad);
    p->ascale = c.take<float>(p->P);
    // padding: the wide kernels' copies of a last, partial column tile run
    // up to one tile (512 columns x 128 bytes) past the last projection
    p->proj = c.take<unsigned char>(
        (p->info.wide ? projection_bytes(p->form, (p->info.schedule & kScheduleDirect) != 0, p->C,
                                         p->O, p->kpad)
                      : (size_t)p->C * p->O * p->nkc * chunk_bytes(p->form)) +
        kProjPad);
    p->pabs2 = c.take<float>((size_t)p->C * p->O * p->kpad);
    p->pnorm = c.take<float>((size_t)p->P * p->C * p->O);
    const size_t prows = p->packed ? (size_t)p->n_rows : (size_t)p->n_groups;
    p->part_min = c.take<double>(prows * nt_total * p->parts_per_group);
    p->part_idx = c.take<long long>(prows * nt_total * p->parts_per_group);
    // schedule 8 (wide or direct): the translation operand of every row
    // tile and chunk, in the LDS image of the kernel's A buffer (tile_m rows
    // x row bytes)
    p->a_chunks = (p->npix + 15) / 16;
    p->a_enc = nullptr;
    if (p->info.wide && (p->info.schedule & kSchedulePrecomputedA))
        p->a_enc = c.take<unsigned char>((size_t)p->n_mtiles * p->a_chunks * p->info.tile_m *
                                         (p->form == kFP16 ? 64 : 128));
    return align_up(c.off, 256);
}

}  // namespace rc

namespace rc {

// Largest number of particles among the row tiles of `tile_m` packed rows.
static int max_tile_particles(const std::vector<int> &row_part, int tile_m)
{
    int worst = 0;
    const int n = (int)row_part.size();
    for (int r0 = 0; r0 < n; r0 += tile_m)
        worst = std::max(worst, row_part[std::min(r0 + tile_m, n) - 1] - row_part[r0] + 1);
    return worst;
}

int plan_setup(rc_plan *p, const rc_batch_desc *d, PlanTables &tab, int forced_variant)
{
    if (int st = validate_batch(d, false)) return st;
    p->P = d->n_particles;
    p->C = d->n_classes;
    p->O = d->orientation_num;
    p->npix = d->npix;
    p->full_npix = d->full_npix;
    p->form = d->form;
    p->kpad = (int)align_up((size_t)d->npix, kChunk);
    p->nkc = p->kpad / kChunk;
    p->trans_offset.assign(d->host_trans_offset, d->host_trans_offset + d->n_particles + 1);
    tab.groups.clear();
    tab.group_begin.assign(p->P + 1, 0);
    for (int i = 0; i < p->P; ++i) {
        const int T = p->trans_offset[i + 1] - p->trans_offset[i];
        tab.group_begin[i] = (int)tab.groups.size();
        for (int t0 = 0; t0 < T; t0 += 16) tab.groups.push_back(make_int2(i, t0));
    }
    tab.group_begin[p->P] = (int)tab.groups.size();
    p->n_groups = (int)tab.groups.size();
    p->n_rows = p->trans_offset[p->P];
    tab.row_part.resize(p->n_rows);
    for (int i = 0; i < p->P; ++i)
        for (int r = p->trans_offset[i]; r < p->trans_offset[i + 1]; ++r) tab.row_part[r] = i;
    int vi = forced_variant >= 0 ? forced_variant : g_variant_override.load();
    if (vi < 0) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(tune_key(d));
        vi = it != g_tuned.end()
                 ? it->second
                 : default_variant(p->form, p->trans_offset[p->P], p->C, p->O,
                                   max_tile_particles(tab.row_part, 128));
    }
    const VariantInfo *info = variant_info(p->form, false, vi);
    RC_REQUIRE(info, RC_ERR_INVALID, "variant %d does not exist for form %d", vi, d->form);
    if (info->max_tile_parts > 0) {
        // rows are packed in tiles of tile_m; the kernel stages the image
        // data of every particle of a tile
        const int worst = max_tile_particles(tab.row_part, info->tile_m);
        RC_REQUIRE(worst <= info->max_tile_parts, RC_ERR_UNSUPPORTED,
                   "variant %d takes at most %d particles per %d-row tile; this batch puts %d "
                   "in one",
                   vi, info->max_tile_parts, info->tile_m, worst);
    }
    p->variant = vi;
    p->info = *info;
    p->packed = info->wide;
    const long long rows = p->packed ? (long long)p->n_rows : (long long)p->n_groups * 16;
    p->n_mtiles = (int)((rows + info->tile_m - 1) / info->tile_m);
    p->ntiles_per_class = (p->O + info->tile_n - 1) / info->tile_n;
    p->parts_per_group = info->waves_n;
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 256;
        const int per_cu = std::max(1, std::min(2, (160 * 1024) / std::max(info->lds_bytes, 1)));
        p->persist_blocks = cus * per_cu;
    }
    p->proj_dims.clear();
    tab.views.resize(p->C);
    tab.scales.resize(p->C);
    for (int c = 0; c < p->C; ++c) {
        const rc_projector *pr = d->projectors[c].proj;
        p->proj_dims.insert(p->proj_dims.end(), {pr->xdim, pr->ydim, pr->zdim, pr->inity,
                                                 pr->initz, d->projectors[c].maxR2_padded,
                                                 (int)(pr->pad * 1000)});
        tab.views[c] = view_of(d->projectors[c]);
        tab.scales[c] = pow2_scale_host(pr->max_abs);
    }
    return RC_OK;
}

size_t plan_table_bytes(const PlanTables &tab)
{
    return align_up(tab.groups.size() * sizeof(int2), 256) +
           align_up(tab.group_begin.size() * sizeof(int), 256) +
           align_up(tab.row_part.size() * sizeof(int), 256) +
           align_up(tab.views.size() * sizeof(ProjView), 256) +
           align_up(tab.scales.size() * sizeof(float), 256);
}

int plan_upload_async(rc_plan *p, const PlanTables &tab, unsigned char *staging, hipStream_t s)
{
    size_t off = 0;
    auto put = [&](void *dst, const void *src, size_t n) -> hipError_t {
        memcpy(staging + off, src, n);
        hipError_t e = hipMemcpyAsync(dst, staging + off, n, hipMemcpyHostToDevice, s);
        off += align_up(n, 256);
        return e;
    };
    RC_HIP(put(p->groups, tab.groups.data(), tab.groups.size() * sizeof(int2)));
    RC_HIP(put(p->group_begin, tab.group_begin.data(), tab.group_begin.size() * sizeof(int)));
    RC_HIP(put(p->row_part, tab.row_part.data(), tab.row_part.size() * sizeof(int)));
    RC_HIP(put(p->vols, tab.views.data(), tab.views.size() * sizeof(ProjView)));
    RC_HIP(put(p->bscale, tab.scales.data(), tab.scales.size() * sizeof(float)));
    return RC_OK;
}

int plan_run(rc_plan *p, const rc_batch_desc *d, hipStream_t s)
{
    RC_REQUIRE(p, RC_ERR_INVALID, "NULL plan");
    if (int st = validate_batch(d, true)) return st;
    RC_REQUIRE(d->n_particles == p->P && d->n_classes == p->C && d->orientation_num == p->O &&
                   d->npix == p->npix && d->full_npix == p->full_npix && d->form == p->form,
               RC_ERR_INVALID, "batch shape differs from the plan");
    RC_REQUIRE(std::equal(p->trans_offset.begin(), p->trans_offset.end(), d->host_trans_offset),
               RC_ERR_INVALID, "per-particle translation counts differ from the plan");
    for (int c = 0; c < p->C; ++c) {
        const rc_projector *pr = d->projectors[c].proj;
        const int *dims = &p->proj_dims[7 * c];
        RC_REQUIRE(dims[0] == pr->xdim && dims[1] == pr->ydim && dims[2] == pr->zdim &&
                       dims[3] == pr->inity && dims[4] == pr->initz &&
                       dims[5] == d->projectors[c].maxR2_padded,
                   RC_ERR_INVALID, "projector %d differs from the plan", c);
    }
    const rc_projector_kernel &k0 = d->projectors[0];
    RC_HIP(launch_prep(p->P, p->npix, p->full_npix, p->kpad, d->d_real, d->d_imag, d->d_corr,
                       d->d_coor_xy, k0.imgX, k0.imgY, k0.maxR, p->xw, p->coords, p->xnorm,
                       p->half_corr, p->ascale, s, d->do_cc != 0));
    if (p->info.wide)
        RC_HIP(launch_project_wide(p->form, p->vols, p->bscale, p->C, d->d_eulers, p->O,
                                   p->coords, p->npix, p->kpad, p->proj, p->pabs2,
                                   (p->info.schedule & kScheduleDirect) != 0, s));
    else
        RC_HIP(launch_project(p->form, p->vols, p->bscale, p->C, d->d_eulers, p->O, p->coords,
                              p->npix, p->kpad, p->proj, p->pabs2, s));
    RC_HIP(launch_pnorm(p->P, p->C * p->O, p->kpad, p->half_corr, p->pabs2, p->pnorm, s));
    CrossArgs a{};
    a.groups = p->groups;
    a.n_groups = p->n_groups;
    a.row_part = p->row_part;
    a.n_rows = p->n_rows;
    a.trans_offset = d->d_trans_offset;
    a.n_trans = 0;
    a.trans_x = d->d_trans_x;
    a.trans_y = d->d_trans_y;
    a.xw = p->xw;
    a.coords = p->coords;
    a.xnorm = p->xnorm;
    a.ascale = p->ascale;
    a.kpad = p->kpad;
    a.npix = p->npix;
    a.proj = p->proj;
    a.pnorm = p->pnorm;
    a.bscale = p->bscale;
    a.half_corr = p->half_corr;
    a.eulers = d->d_eulers;
    a.n_classes = p->C;
    a.orientation_num = p->O;
    a.ntiles_per_class = p->ntiles_per_class;
    a.n_mtiles = p->n_mtiles;
    a.diff2_init = d->d_diff2_init;
    a.diff2 = d->d_diff2;
    const bool want_best = d->d_best_diff2 || d->d_best_index;
    a.part_min = want_best ? p->part_min : nullptr;
    a.part_idx = want_best ? p->part_idx : nullptr;
    a.persist_blocks = p->persist_blocks;
    a.cc = d->do_cc ? 1 : 0;
    a.a_enc = p->a_enc;
    a.a_chunks = p->a_chunks;
    // the timing slots bracket the whole cross term: the translation
    // operand's precompute (schedule 8) and the cross kernel
    const int nslot = p->capturing ? 0 : (int)p->ev.size() / 2;
    if (nslot) RC_HIP(hipEventRecord(p->ev[2 * (p->launches % nslot)], s));
    if (p->a_enc)
        RC_HIP(launch_encode_a(p->form, p->info.tile_m, a, p->a_enc, s));
    RC_HIP(launch_cross(p->form, false, p->variant, a, s));
    if (nslot) {
        RC_HIP(hipEventRecord(p->ev[2 * (p->launches % nslot) + 1], s));
        ++p->launches;
    }
    if (want_best)
        RC_HIP(launch_argmin(p->P, p->packed ? d->d_trans_offset : p->group_begin,
                             p->C * p->ntiles_per_class * p->parts_per_group, p->part_min,
                             p->part_idx, d->d_best_diff2, (long long *)d->d_best_index, s));
    return RC_OK;
}

}  // namespace rc

using namespace rc;

extern "C" {

const char *rc_last_error(void) { return g_last_error.c_str(); }

const char *rc_version(void) { return "relion_coarse_amd 0.1 (gfx950)"; }

int rc_projector_create(int xdim, int ydim, int zdim, int inity, int initz, int maxr,
                        float padding_factor, const float *host_real, const float *host_imag,
                        rc_projector **out)
{
    RC_REQUIRE(out, RC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    RC_REQUIRE(zdim != 1 && zdim != 0, RC_ERR_UNSUPPORTED,
               "2D references (zdim == 1) are not on the coarse 3D path");
    RC_REQUIRE(xdim >= 2 && ydim > 0 && zdim > 0, RC_ERR_INVALID,
               "bad volume dimensions (xdim must be at least 2)");
    RC_REQUIRE(padding_factor > 0.f && maxr > 0, RC_ERR_INVALID, "bad maxr or padding factor");
    RC_REQUIRE(host_real && host_imag, RC_ERR_INVALID, "NULL volume data");
    const size_t n = (size_t)xdim * ydim * zdim;
    std::vector<float2> tmp(n);
    float m = 0.f;
    for (size_t i = 0; i < n; ++i) {
        tmp[i] = make_float2(host_real[i], host_imag[i]);
        m = std::max(m, std::max(std::fabs(host_real[i]), std::fabs(host_imag[i])));
    }
    static std::atomic<unsigned long long> s_generation{0};
    rc_projector *p = new rc_projector{};
    p->generation = ++s_generation;
    p->xdim = xdim;
    p->ydim = ydim;
    p->zdim = zdim;
    p->inity = inity;
    p->initz = initz;
    p->maxr = maxr;
    p->pad = padding_factor;
    p->max_abs = m;
    p->bytes = n * sizeof(float2);
    if (hipGetDevice(&p->device) != hipSuccess ||
        hipMalloc(&p->vol, p->bytes + 256) != hipSuccess) {
        delete p;
        set_error("hipMalloc of %zu bytes for the projector volume failed", n * sizeof(float2));
        return RC_ERR_NO_MEMORY;
    }
    p->d_scale = reinterpret_cast<float *>(reinterpret_cast<unsigned char *>(p->vol) +
                                           align_up(p->bytes, 256));
    float sc = pow2_scale_host(m);
    if (hipMemcpy(p->vol, tmp.data(), p->bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_scale, &sc, sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p->vol);
        delete p;
        set_error("copy of the projector volume to the device failed");
        return RC_ERR_DEVICE;
    }
    *out = p;
    return RC_OK;
}

int rc_projector_create_from_file(const char *path, rc_projector **out)
{
    RC_REQUIRE(path && out, RC_ERR_INVALID, "NULL argument");
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    RC_REQUIRE(f.is_open(), RC_ERR_IO, "cannot open projector file '%s'", path);
    int32_t h[6];
    float pad = 0.f;
    f.read(reinterpret_cast<char *>(h), sizeof(h));
    f.read(reinterpret_cast<char *>(&pad), sizeof(pad));
    RC_REQUIRE(f.good(), RC_ERR_IO, "projector file '%s' has a truncated header", path);
    int zdim = h[2] == 1 ? 0 : h[2];
    RC_REQUIRE(h[0] > 0 && h[1] > 0 && zdim > 0, RC_ERR_UNSUPPORTED,
               "projector file '%s': dims %d x %d x %d unsupported", path, h[0], h[1], h[2]);
    const size_t n = (size_t)h[0] * h[1] * zdim;
    std::vector<float> re(n), im(n);
    f.read(reinterpret_cast<char *>(re.data()), n * sizeof(float));
    f.read(reinterpret_cast<char *>(im.data()), n * sizeof(float));
    RC_REQUIRE(f.good(), RC_ERR_IO, "projector file '%s' is truncated", path);
    return rc_projector_create(h[0], h[1], zdim, h[3], h[4], h[5], pad, re.data(), im.data(), out);
}

int rc_coarse_data_load(const char *path, rc_coarse_data *out)
{
    RC_REQUIRE(path && out, RC_ERR_INVALID, "NULL argument");
    memset(out, 0, sizeof(*out));
    std::ifstream f(path, std::ios::binary);
    RC_REQUIRE(f.is_open(), RC_ERR_IO, "cannot open coarse data file '%s'", path);
    int32_t h[7];
    f.read(reinterpret_cast<char *>(h), sizeof(h));
    RC_REQUIRE(f.good(), RC_ERR_IO, "coarse data file '%s' has a truncated header", path);
    RC_REQUIRE(h[4] > 0 && h[5] > 0 && h[6] > 0 && h[4] <= kMaxTrans && h[6] <= kMaxPix &&
                   h[5] < (1 << 24),
               RC_ERR_IO, "coarse data file '%s': implausible sizes T=%d O=%d npix=%d", path,
               h[4], h[5], h[6]);
    out->xdim = h[0];
    out->ydim = h[1];
    out->zdim = h[2];
    out->rmax = h[3];
    out->translation_num = h[4];
    out->orientation_num = h[5];
    out->image_size = h[6];
    const size_t O = (size_t)h[5], T = (size_t)h[4], n = (size_t)h[6];
    auto take = [&](size_t count) {
        float *p = static_cast<float *>(malloc(std::max<size_t>(count, 1) * sizeof(float)));
        if (p) f.read(reinterpret_cast<char *>(p), count * sizeof(float));
        return p;
    };
    out->eulers = take(9 * O);
    out->trans_x = take(T);
    out->trans_y = take(T);
    out->real = take(n);
    out->imag = take(n);
    out->corr = take(n);
    out->diff2s = take(O * T);
    const bool ok = f.good() && out->eulers && out->trans_x && out->trans_y && out->real &&
                    out->imag && out->corr && out->diff2s;
    if (!ok) {
        rc_coarse_data_free(out);
        set_error("coarse data file '%s' is truncated", path);
        return RC_ERR_IO;
    }
    return RC_OK;
}

void rc_coarse_data_free(rc_coarse_data *d)
{
    if (!d) return;
    free(d->eulers);
    free(d->trans_x);
    free(d->trans_y);
    free(d->real);
    free(d->imag);
    free(d->corr);
    free(d->diff2s);
    memset(d, 0, sizeof(*d));
}

int rc_projector_destroy(rc_projector *p)
{
    if (!p) return RC_OK;
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != p->device) (void)hipSetDevice(p->device);
    hipError_t e = hipFree(p->vol);
    if (cur != p->device) (void)hipSetDevice(cur);
    delete p;
    RC_REQUIRE(e == hipSuccess, RC_ERR_DEVICE, "hipFree of the projector volume failed");
    return RC_OK;
}

int rc_projector_dims(const rc_projector *p, int *xdim, int *ydim, int *zdim, int *inity,
                      int *initz, int *maxr, float *padding_factor)
{
    RC_REQUIRE(p, RC_ERR_INVALID, "NULL projector");
    if (xdim) *xdim = p->xdim;
    if (ydim) *ydim = p->ydim;
    if (zdim) *zdim = p->zdim;
    if (inity) *inity = p->inity;
    if (initz) *initz = p->initz;
    if (maxr) *maxr = p->maxr;
    if (padding_factor) *padding_factor = p->pad;
    return RC_OK;
}

int rc_projector_make_kernel(const rc_projector *p, int imgX, int imgY, int imgZ, int imgMaxR,
                             rc_projector_kernel *out)
{
    RC_REQUIRE(p && out, RC_ERR_INVALID, "NULL argument");
    RC_REQUIRE(imgX > 0 && imgY > 0 && imgMaxR >= 0, RC_ERR_INVALID, "bad image geometry");
    out->proj = p;
    out->imgX = imgX;
    out->imgY = imgY;
    out->imgZ = imgZ;
    out->maxR = p->maxr >= imgMaxR ? imgMaxR : p->maxr;
    out->padding_factor = p->pad;
    out->maxR2_padded = (int)(out->maxR * out->maxR * p->pad * p->pad);
    return RC_OK;
}

int rc_run_diff2_kernel_coarse(const rc_projector_kernel *projector, const float *d_trans_x,
                               const float *d_trans_y, const float *d_trans_z,
                               const float *d_corr_img, const float *d_fimg_real,
                               const float *d_fimg_imag, const float *d_eulers, float *d_diff2s,
                               float local_sqrtXi2, unsigned long orientation_num,
                               unsigned long translation_num, unsigned long image_size,
                               void *stream, int do_CC, int data_is_3D)
{
    (void)d_trans_z;
    (void)local_sqrtXi2;
    if (int st = check_drop_in(projector, orientation_num, translation_num, image_size,
                               data_is_3D))
        return st;
    // the search runs over the pixels a rotation can bring inside the mask
    // (the others only add corr/2 |X|^2); all of them when none can
    float pad2, thr;
    const int n_in = dropin_pixels(*projector, (int)image_size, do_CC != 0, &pad2, &thr);
    const bool part = n_in > 0;
    const int npix = part ? n_in : (int)image_size;
    const size_t need = single_workspace_bytes((int)translation_num, (int)orientation_num, npix,
                                               RC_FORM_FP32);
    std::unique_lock<std::mutex> held;
    void *ws = nullptr;
    ScratchEntry *entry = nullptr;
    if (int st = scratch_for((hipStream_t)stream, need, held, &ws, &entry)) return st;
    const int kpad = (int)align_up((size_t)npix, kChunk);
    ProjCacheView cache{};
    int force = 0;
    const bool use_cache = (size_t)orientation_num * kpad * sizeof(float2) <= kProjCacheMaxBytes &&
                           orientation_num <= 65535 && d_eulers;
    if (use_cache) {
        ProjCacheKey key;
        key.proj = projector->proj;
        key.generation = projector->proj->generation;
        key.eulers = d_eulers;
        key.O = (int)orientation_num;
        key.npix = npix;
        key.full = (int)image_size;
        key.imgX = projector->imgX;
        key.imgY = projector->imgY;
        key.maxR = projector->maxR;
        key.maxR2 = projector->maxR2_padded;
        if (int st = proj_cache_for(entry, (hipStream_t)stream, key, kpad, &cache, &force))
            return st;
    }
    return run_single((int)translation_num, (int)orientation_num, npix, (int)image_size,
                      RC_FORM_FP32, d_eulers, d_trans_x, d_trans_y, d_fimg_real, d_fimg_imag,
                      projector, d_corr_img, d_diff2s, d_diff2s, (hipStream_t)stream, nullptr, ws,
                      need, do_CC != 0, use_cache ? &cache : nullptr, force, part);
}

int rc_scratch_release(void *stream) { return scratch_release((hipStream_t)stream); }

size_t rc_coarse_workspace_bytes(int translation_num, int orientation_num, int image_size,
                                 int full_image_size, int form)
{
    (void)full_image_size;
    return single_workspace_bytes(translation_num, orientation_num, image_size, form);
}

int rc_coarse_run(int translation_num, int orientation_num, int image_size, int full_image_size,
                  int form, const float *d_eulers, const float *d_trans_x,
                  const float *d_trans_y, const float *d_real, const float *d_imag,
                  const rc_projector_kernel *projector, const float *d_corr,
                  const float *d_diff2s, float *d_diff2s_dest, void *stream,
                  const float *d_coor_xy, void *d_workspace, size_t workspace_bytes)
{
    int full = full_image_size;
    if (!d_coor_xy || full < image_size) full = image_size;
    return run_single(translation_num, orientation_num, image_size, full, form, d_eulers,
                      d_trans_x, d_trans_y, d_real, d_imag, projector, d_corr, d_diff2s,
                      d_diff2s_dest, (hipStream_t)stream, d_coor_xy, d_workspace,
                      workspace_bytes);
}

// Times `timed` launches of variant v on desc (after `warm` untimed ones)
// with HIP events on the stream; returns their median in *ms.
static int time_variant(const rc_batch_desc *d, hipStream_t stream, int v, int warm, int timed,
                        float *ms)
{
    rc_plan *p = nullptr;
    int st = plan_create_internal(d, v, &p);
    if (st == RC_OK) st = rc_plan_timing(p, timed);
    for (int k = 0; st == RC_OK && k < warm; ++k) st = rc_plan_execute(p, d, stream);
    if (st == RC_OK && hipStreamSynchronize(stream) != hipSuccess) {
        set_error("stream synchronisation failed while tuning");
        st = RC_ERR_DEVICE;
    }
    if (st == RC_OK) p->launches = 0;
    for (int k = 0; st == RC_OK && k < timed; ++k) st = rc_plan_execute(p, d, stream);
    std::vector<float> t(timed);
    int got = 0;
    if (st == RC_OK) st = rc_plan_timing_read(p, t.data(), timed, &got);
    if (p) rc_plan_destroy(p);
    if (st != RC_OK) return st;
    // median: robust to the odd slow launch (clock ramps, other streams)
    std::sort(t.begin(), t.begin() + got);
    *ms = got == 0 ? 0.f : (got & 1) ? t[got / 2] : 0.5f * (t[got / 2 - 1] + t[got / 2]);
    return RC_OK;
}

int rc_autotune(const rc_batch_desc *d, void *stream, int iters, int *best_variant,
                float *best_ms)
{
    if (int st = validate_batch(d, true)) return st;
    RC_REQUIRE(iters > 0 && iters <= 1000, RC_ERR_INVALID, "iters must be 1..1000");
    const hipStream_t s = (hipStream_t)stream;
    const int n = variant_count(d->form, false);
    // Each plan is made with its variant named explicitly, so plans created
    // meanwhile by other threads are not affected by the sweep.
    // Screen: one untimed and one timed launch per variant.
    // Variants that cannot take this batch (RC_ERR_UNSUPPORTED from plan
    // creation) are left out.
    std::vector<float> first(n, -1.f);
    float fastest = -1.f;
    for (int v = 0; v < n; ++v) {
        const int st = time_variant(d, s, v, 1, 1, &first[v]);
        if (st == RC_ERR_UNSUPPORTED) {
            first[v] = -1.f;
            continue;
        }
        if (st) return st;
        if (fastest < 0.f || first[v] < fastest) fastest = first[v];
    }
    RC_REQUIRE(fastest >= 0.f, RC_ERR_UNSUPPORTED, "no variant takes this batch");
    // Refine the variants within 25 % of the fastest screen: two sweeps of
    // two untimed and `iters` timed launches, each keeping its better
    // median (a variant timed early is not penalised by a clock still
    // ramping up).
    std::vector<int> cand;
    for (int v = 0; v < n; ++v)
        if (first[v] >= 0.f && first[v] <= 1.25f * fastest) cand.push_back(v);
    std::vector<float> vt(n, -1.f);
    for (int sweep = 0; sweep < 2; ++sweep)
        for (int v : cand) {
            float t = 0.f;
            if (int st = time_variant(d, s, v, 2, iters, &t)) return st;
            if (vt[v] < 0.f || t < vt[v]) vt[v] = t;
        }
    int best = -1;
    float best_t = 0.f;
    for (int v : cand)
        if (best < 0 || vt[v] < best_t) {
            best = v;
            best_t = vt[v];
        }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[tune_key(d)] = best;
    }
    if (best_variant) *best_variant = best;
    if (best_ms) *best_ms = best_t;
    return RC_OK;
}

int rc_set_variant_override(int variant)
{
    RC_REQUIRE(variant >= -1, RC_ERR_INVALID, "variant override %d (must be >= -1)", variant);
    g_variant_override.store(variant);
    return RC_OK;
}

int rc_variant_count(int form) { return variant_count(form, false); }

int rc_variant_info(int form, int variant, int *tile_m, int *tile_n, int *fused_gather)
{
    const VariantInfo *v = variant_info(form, false, variant);
    RC_REQUIRE(v, RC_ERR_INVALID, "no variant %d for form %d", variant, form);
    if (tile_m) *tile_m = v->tile_m;
    if (tile_n) *tile_n = v->tile_n;
    if (fused_gather) *fused_gather = v->fused ? 1 : 0;
    return RC_OK;
}

int rc_run_diff2_kernel_fine(const rc_projector_kernel *projector, const float *d_eulers,
                             const float *d_real, const float *d_imag, const float *d_trans_x,
                             const float *d_trans_y, const float *d_trans_z, const float *d_corr,
                             float *d_diff2s, unsigned long image_size, float sum_init,
                             unsigned long orientation_num, unsigned long translation_num,
                             unsigned long todo_blocks, const unsigned long *d_rot_idx,
                             const unsigned long *d_trans_idx, const unsigned long *d_job_idx,
                             const unsigned long *d_job_num, int data_is_3D, void *stream)
{
    (void)d_trans_z;
    RC_REQUIRE(!data_is_3D, RC_ERR_UNSUPPORTED, "3D data (data_is_3D) is not implemented");
    if (int st = check_kernel_view(projector)) return st;
    RC_REQUIRE(image_size > 0 && image_size <= (unsigned long)kMaxPix, RC_ERR_LIMIT,
               "image_size %lu out of range", image_size);
    RC_REQUIRE(image_size <= (unsigned long)projector->imgX * projector->imgY, RC_ERR_INVALID,
               "image_size %lu exceeds imgX*imgY", image_size);
    RC_REQUIRE(orientation_num > 0 && translation_num > 0, RC_ERR_INVALID,
               "orientation_num and translation_num must be positive");
    // one 256-thread workgroup per job: the grid's work-items stay below 2^32
    RC_REQUIRE(todo_blocks <= (1ul << 24) - 8, RC_ERR_LIMIT, "too many jobs");
    if (todo_blocks == 0) return RC_OK;
    RC_REQUIRE(d_eulers && d_real && d_imag && d_trans_x && d_trans_y && d_corr && d_diff2s &&
                   d_rot_idx && d_trans_idx && d_job_idx && d_job_num,
               RC_ERR_INVALID, "NULL device pointer");
    FineArgs a{};
    a.v = view_of(*projector);
    a.imgX = projector->imgX;
    a.imgY = projector->imgY;
    a.maxR = projector->maxR;
    a.eulers = d_eulers;
    a.real = d_real;
    a.imag = d_imag;
    a.corr = d_corr;
    a.trans_x = d_trans_x;
    a.trans_y = d_trans_y;
    a.diff2 = d_diff2s;
    a.image_size = (int)image_size;
    a.sum_init = sum_init;
    a.orientation_num = orientation_num;
    a.translation_num = translation_num;
    a.rot_idx = d_rot_idx;
    a.trans_idx = d_trans_idx;
    a.job_idx = d_job_idx;
    a.job_num = d_job_num;
    RC_HIP(launch_fine(a, todo_blocks, (hipStream_t)stream));
    return RC_OK;
}

int rc_coarse_weights(const rc_weights_desc *d, void *stream)
{
    RC_REQUIRE(d, RC_ERR_INVALID, "NULL weights descriptor");
    RC_REQUIRE(d->n_particles > 0 && d->n_classes > 0 && d->orientation_num > 0,
               RC_ERR_INVALID, "particles, classes and orientations must be positive");
    RC_REQUIRE(d->adaptive_fraction > 0.0 && d->adaptive_fraction <= 1.0, RC_ERR_INVALID,
               "adaptive_fraction %g outside (0, 1]", d->adaptive_fraction);
    RC_REQUIRE(d->host_trans_offset && d->d_trans_offset && d->d_diff2 && d->d_min_diff2 &&
                   d->d_weights && d->d_sum_weight && d->d_significant_weight &&
                   d->d_n_significant,
               RC_ERR_INVALID, "NULL pointer in weights descriptor");
    RC_REQUIRE(d->log_prior_orient_stride >= 0, RC_ERR_INVALID, "negative prior stride");
    for (int i = 0; i < d->n_particles; ++i) {
        const long long T = (long long)d->host_trans_offset[i + 1] - d->host_trans_offset[i];
        RC_REQUIRE(T > 0 && T <= kMaxTrans, RC_ERR_INVALID,
                   "particle %d has %lld translations", i, T);
        RC_REQUIRE((long long)d->n_classes * d->orientation_num * T < (1ll << 31),
                   RC_ERR_LIMIT, "particle %d has more than 2^31 samples", i);
    }
    WeightsArgs a{};
    a.n_particles = d->n_particles;
    a.n_classes = d->n_classes;
    a.orientation_num = d->orientation_num;
    a.trans_offset = d->d_trans_offset;
    a.diff2 = d->d_diff2;
    a.min_diff2 = d->d_min_diff2;
    a.lp_orient = d->d_log_prior_orient;
    a.orient_stride = d->log_prior_orient_stride;
    a.lp_offset = d->d_log_prior_offset;
    a.fraction = d->adaptive_fraction;
    a.weights = d->d_weights;
    a.sum_weight = d->d_sum_weight;
    a.significant_weight = d->d_significant_weight;
    a.n_significant = d->d_n_significant;
    a.mask = d->d_significant;
    a.max_log_weight = d->d_max_log_weight;
    RC_HIP(launch_weights(a, (hipStream_t)stream));
    return RC_OK;
}

const char *rc_variant_name(int form, int variant)
{
    const VariantInfo *v = variant_info(form, false, variant);
    if (!v) return nullptr;
    static const char *forms[] = {"fp32", "bf16x3", "fp16"};
    static thread_local char buf[160];
    snprintf(buf, sizeof(buf), "%s<%s> %dx%d, %d threads, %d B LDS%s%s", v->family,
             forms[v->form], v->tile_m, v->tile_n, v->threads, v->lds_bytes,
             v->persistent ? ", persistent" : "",
             !v->wide             ? ""
             : v->schedule == 1 ? ", staggered waves (bit 2)"
             : v->schedule == 2 ? ", staggered waves (bit 0)"
             : v->schedule == 4 ? ", fenced segments"
             : v->schedule == 8 ? ", translation operand precomputed"
             : v->schedule == kScheduleDirect ? ", projections loaded into registers"
             : v->schedule == kScheduleDirectPre
                 ? ", projections loaded into registers, translation operand precomputed"
                                : "");
    return buf;
}

int rc_plan_create(const rc_batch_desc *d, rc_plan **out) { return plan_create_internal(d, -1, out); }

}  // extern "C"

namespace rc {

int plan_create_internal(const rc_batch_desc *d, int forced_variant, rc_plan **out)
{
    RC_REQUIRE(out, RC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    rc_plan *p = new rc_plan{};
    PlanTables tab;
    if (int st = plan_setup(p, d, tab, forced_variant)) {
        delete p;
        return st;
    }
    p->bytes = plan_bind(p, nullptr);
    if (hipMalloc(&p->mem, p->bytes) != hipSuccess) {
        const size_t b = p->bytes;
        delete p;
        set_error("hipMalloc of %zu bytes for the batch plan failed", b);
        return RC_ERR_NO_MEMORY;
    }
    plan_bind(p, p->mem);
    if (hipMemcpy(p->groups, tab.groups.data(), tab.groups.size() * sizeof(int2),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->group_begin, tab.group_begin.data(), tab.group_begin.size() * sizeof(int),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->row_part, tab.row_part.data(), tab.row_part.size() * sizeof(int),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->vols, tab.views.data(), tab.views.size() * sizeof(ProjView),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->bscale, tab.scales.data(), tab.scales.size() * sizeof(float),
                  hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p->mem);
        delete p;
        set_error("upload of the batch plan tables failed");
        return RC_ERR_DEVICE;
    }
    *out = p;
    return RC_OK;
}

}  // namespace rc

extern "C" {

int rc_plan_execute(rc_plan *p, const rc_batch_desc *d, void *stream)
{
    return plan_run(p, d, (hipStream_t)stream);
}

static void drop_graph(rc_plan *p)
{
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->graph_exec = nullptr;
    p->graph = nullptr;
}

int rc_plan_capture(rc_plan *p, const rc_batch_desc *d, void *stream)
{
    RC_REQUIRE(p, RC_ERR_INVALID, "NULL plan");
    hipStream_t s = (hipStream_t)stream;
    RC_REQUIRE(s != nullptr, RC_ERR_INVALID, "capture needs a non-default stream");
    drop_graph(p);
    RC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    p->capturing = true;
    const int st = rc_plan_execute(p, d, stream);
    p->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (st != RC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    RC_REQUIRE(e == hipSuccess && g, RC_ERR_DEVICE, "stream capture failed: %s",
               hipGetErrorString(e));
    p->graph = g;
    RC_HIP(hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0));
    return RC_OK;
}

int rc_plan_replay(rc_plan *p, void *stream)
{
    RC_REQUIRE(p && p->graph_exec, RC_ERR_INVALID, "plan has no captured graph");
    RC_HIP(hipGraphLaunch(p->graph_exec, (hipStream_t)stream));
    return RC_OK;
}

static void drop_events(rc_plan *p)
{
    for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
    p->ev.clear();
    p->launches = 0;
}

int rc_plan_timing(rc_plan *p, int n_slots)
{
    RC_REQUIRE(p && n_slots >= 0 && n_slots <= 4096, RC_ERR_INVALID, "bad timing request");
    drop_events(p);
    for (int i = 0; i < 2 * n_slots; ++i) {
        hipEvent_t e;
        RC_HIP(hipEventCreate(&e));
        p->ev.push_back(e);
    }
    return RC_OK;
}

int rc_plan_timing_read(rc_plan *p, float *ms, int n, int *n_written)
{
    RC_REQUIRE(p && ms && n_written, RC_ERR_INVALID, "NULL argument");
    const int nslot = (int)p->ev.size() / 2;
    const long long have = std::min<long long>(p->launches, nslot);
    int k = 0;
    for (long long j = p->launches - have; j < p->launches && k < n; ++j, ++k) {
        const int s = (int)(j % nslot);
        RC_HIP(hipEventSynchronize(p->ev[2 * s + 1]));
        RC_HIP(hipEventElapsedTime(&ms[k], p->ev[2 * s], p->ev[2 * s + 1]));
    }
    *n_written = k;
    return RC_OK;
}

int rc_plan_destroy(rc_plan *p)
{
    if (!p) return RC_OK;
    drop_events(p);
    drop_graph(p);
    hipError_t e = p->mem ? hipFree(p->mem) : hipSuccess;
    delete p;
    RC_REQUIRE(e == hipSuccess, RC_ERR_DEVICE, "hipFree of the plan failed");
    return RC_OK;
}

size_t rc_plan_device_bytes(const rc_plan *p) { return p ? p->bytes : 0; }

int rc_plan_tile(const rc_plan *p, int *variant, int *tile_m, int *tile_n)
{
    RC_REQUIRE(p, RC_ERR_INVALID, "NULL plan");
    if (variant) *variant = p->variant;
    if (tile_m) *tile_m = p->info.tile_m;
    if (tile_n) *tile_n = p->info.tile_n;
    return RC_OK;
}

int rc_coarse_split(int translation_num, int orientation_num, int image_size, int form,
                    int *ksplit, int *kc_per_split)
{
    RC_REQUIRE(translation_num > 0 && orientation_num > 0 && image_size > 0, RC_ERR_INVALID,
               "sizes must be positive");
    RC_REQUIRE(form >= RC_FORM_FP32 && form <= RC_FORM_FP32_DIRECT, RC_ERR_INVALID,
               "unknown form %d", form);
    return single_split(translation_num, orientation_num, image_size, form, ksplit, kc_per_split);
}

int rc_filter_coarse_order(const rc_projector_kernel *k, int image_size, int *host_order,
                           int *n_inside)
{
    RC_REQUIRE(k && n_inside, RC_ERR_INVALID, "NULL argument");
    RC_REQUIRE(k->imgX > 0 && k->imgY > 0 && image_size > 0 &&
                   (long long)image_size <= (long long)k->imgX * k->imgY,
               RC_ERR_INVALID, "image_size %d does not fit %d x %d", image_size, k->imgX,
               k->imgY);
    int head = 0, tail = image_size - 1;
    for (int i = 0; i < image_size; ++i) {
        int x, y;
        const bool in = filter_inside_host(i, *k, x, y);
        const int d = in ? head++ : tail--;
        if (host_order) host_order[d] = i;
    }
    *n_inside = head;
    return RC_OK;
}

int rc_filter_coarse_images(const rc_projector_kernel *k, int image_size, int n_images,
                            const float *d_fimg, const float *d_corr, float *d_real,
                            float *d_imag, float *d_corr_out, float *d_coor_xy,
                            int *d_filtered_size, void *stream)
{
    RC_REQUIRE(k, RC_ERR_INVALID, "NULL projector kernel view");
    RC_REQUIRE(k->imgX > 0 && k->imgY > 0 && image_size > 0 &&
                   (long long)image_size <= (long long)k->imgX * k->imgY,
               RC_ERR_INVALID, "image_size %d does not fit %d x %d", image_size, k->imgX,
               k->imgY);
    RC_REQUIRE(n_images > 0 && n_images < (1 << 30), RC_ERR_INVALID, "n_images %d", n_images);
    RC_REQUIRE(d_fimg && d_corr && d_real && d_imag && d_corr_out, RC_ERR_INVALID,
               "NULL device pointer");
    RC_HIP(launch_filter(image_size, n_images, k->imgX, k->imgY, k->maxR, k->padding_factor,
                         k->maxR2_padded, d_fimg, d_corr, d_real, d_imag, d_corr_out, d_coor_xy,
                         d_filtered_size, (hipStream_t)stream));
    return RC_OK;
}

}  // extern "C"
