// Internal declarations shared by the host runtime and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "relion_coarse.h"

namespace rc {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

#define RC_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            rc::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                          __LINE__);                                                      \
            return RC_ERR_DEVICE;                                                         \
        }                                                                                 \
    } while (0)

#define RC_REQUIRE(cond, code, ...)      \
    do {                                 \
        if (!(cond)) {                   \
            rc::set_error(__VA_ARGS__);  \
            return (code);               \
        }                                \
    } while (0)

// ---------------------------------------------------------------------------
// device-side views
// ---------------------------------------------------------------------------
// Padded Fourier volume: one float2 (re, im) per voxel, x fastest, then y,
// then z -- the planar pair of AccProjector interleaved so one 8-byte load
// fetches both parts of a voxel.
struct ProjView {
    const float2 *vol;
    int mdlX, mdlY, mdlZ, mdlXY;
    int initY, initZ;
    int maxR2_padded;
    float pad;
};

// Limits of this build (checked on the host before any launch).
constexpr int kMaxPix = 1 << 20;          // pixels per image
constexpr int kMaxTrans = 1 << 16;        // translations per particle
constexpr size_t kProjPad = 512 * 128;    // bytes after the projections: one wide column tile
constexpr int kChunk = 32;                // pixels per MFMA k-chunk
constexpr int kGroupPix = 8;              // pixels per lane per chunk

// Cross-term arithmetic (values match rc_form).
enum Form : int { kFP32 = 0, kBF16X3 = 1, kFP16 = 2 };

// Bytes of one encoded 8-pixel group of a projection (or of a translated
// image row): FP32 re[8] im[8]; BF16X3 re_hi[8] re_lo[8] im_hi[8] im_lo[8];
// FP16 re[8] im[8].
constexpr int group_bytes(int form) { return form == kFP16 ? 32 : 64; }
constexpr int chunk_bytes(int form) { return 4 * group_bytes(form); }

// Arguments of the cross-term kernel (one struct, passed by value).
struct CrossArgs {
    // rows: row group g covers translations t0 .. t0+15 of one particle
    const int2 *groups;          // [n_groups] (particle, t0); NULL => particle 0, t0 = 16 g
    int n_groups;
    // packed rows (wide kernels): row R is translation R of the CSR arrays
    const int *row_part;         // [n_rows] particle of each row
    int n_rows;
    const int *trans_offset;     // [P+1]; NULL => single particle with T = n_trans
    int n_trans;
    const float *trans_x, *trans_y;
    // A side
    const float2 *xw;            // [P][kpad] corr * X (complex), zero padded
    const float2 *coords;        // [kpad] (x, y) pixel coordinates
    const double *xnorm;         // [P] sum corr/2 |X|^2 over all pixels (float64)
    const float *ascale;         // [P] FP16 form: power-of-two image scale
    int kpad;                    // pixels padded to a multiple of kChunk
    int npix;                    // projected pixels (the rest of kpad is zero)
    // B side
    const unsigned char *proj;   // [C][O][kpad/32][chunk_bytes] encoded projections
    const float *pnorm;          // [P][C][O] sum corr/2 |P|^2
    const float *bscale;         // [C] FP16 form: power-of-two volume scale
    const float *half_corr;      // fused gather: [P][kpad] corr/2 (zero padded)
    const float *eulers;         // fused gather: [O][9]
    ProjView vol[1];             // fused gather: the (single) class volume
    int n_classes, orientation_num;
    int ntiles_per_class;        // ceil(O / tile_n)
    int n_mtiles;
    int persist_blocks;          // persistent variants: blocks in the grid
    int cc;                      // cross-correlation score instead of diff2
    // split-K (fused single-image kernels only): blockIdx.y = slice s covers
    // chunks [s * kc_per_split, min(nkc, (s + 1) * kc_per_split)) and writes
    // its partial sums instead of diff2; ksplit <= 1 = one slice, normal epilogue
    int ksplit, kc_per_split;
    float *part_c;               // [ksplit][O][n_trans] cross-term partials
    float *part_pn;              // [ksplit][O] sum corr/2 |P|^2 partials
    // fused kernels: projections computed by an earlier call (drop-in
    // projection cache), [O][kpad] float2, read instead of the trilinear
    // gather; NULL = gather
    const float2 *pcache;
    // wide kernels with the translation operand precomputed (schedule 8):
    // [row tile][chunk][tile_m rows][row bytes], each chunk block the exact
    // (swizzled) LDS image of the tile's A buffer; a_chunks = chunks per row tile
    const unsigned char *a_enc;
    int a_chunks;
    // outputs
    const float *diff2_init;     // same layout as diff2, may be NULL, may alias diff2
    float *diff2;
    double *part_min;            // [n_groups][C*ntiles_per_class][waves_n] or NULL
    long long *part_idx;
};

// Description of one compiled cross-term variant.
struct VariantInfo {
    int form;
    int tile_m, tile_n;
    int threads;
    int waves_n;
    bool fused;
    int lds_bytes;
    bool persistent;
    bool wide = false;     // 16-pixel chunks, wide projection encoding
    const char *family = "cross_kernel";   // kernel template name
    int schedule = 0;      // wide kernels: wave schedule (0 plain, 1/2 staggered, 4 fenced,
                           // 8 translation operand precomputed and copied like B,
                           // kScheduleDirect projections loaded straight into registers,
                           // kScheduleDirectPre both)
    int max_tile_parts = 0;  // > 0: at most this many particles in one row tile
};

// Schedule of the direct kernels (kernels_direct.hip), whose projections use
// the direct layout (launch_project_wide, direct = true).
constexpr int kScheduleDirect = 16;
// Bit of the schedule: the translation operand is precomputed (a_enc).
constexpr int kSchedulePrecomputedA = 8;
constexpr int kScheduleDirectPre = kScheduleDirect | kSchedulePrecomputedA;
// Particles one row tile of the direct kernels may hold (their image data
// is staged in LDS per tile).
constexpr int kDirectMaxTileParts = 31;

struct VariantEntry {
    VariantInfo info;
    hipError_t (*launch)(const CrossArgs &, hipStream_t);
};

// Wide kernels (both operands staged in LDS, 16-pixel chunks).
const VariantEntry *wide_table(int form, int *n);
// Direct kernels (projections loaded into registers, 16-pixel chunks).
const VariantEntry *direct_table(int form, int *n);

int variant_count(int form, bool fused);
const VariantInfo *variant_info(int form, bool fused, int idx);
hipError_t launch_cross(int form, bool fused, int idx, const CrossArgs &a, hipStream_t s);

// Prep: corr*X, coordinates, |X|^2 (float64), FP16 image scale.
hipError_t launch_prep(int n_particles, int npix, int full_npix, int kpad, const float *real,
                       const float *imag, const float *corr, const float *coor_xy, int imgX,
                       int imgY, int maxR, float2 *xw, float2 *coords, double *xnorm,
                       float *half_corr, float *ascale, hipStream_t s, bool cc_coords = false);
// prep of one unfiltered drop-in image with the pixels a rotation can bring
// inside the mask moved first: (float)(x^2 + y^2) * pad2 <= thr (see
// dropin_pixels, which counts them on the host)
hipError_t launch_prep_partition(int image_size, int kpad, const float *real, const float *imag,
                                 const float *corr, int imgX, int imgY, int maxR, float pad2,
                                 float thr, float2 *xw, float2 *coords, double *xnorm,
                                 float *half_corr, float *ascale, hipStream_t s, bool cc_coords);

// Projections of every (class, orientation, pixel) into the encoded layout.
hipError_t launch_project(int form, const ProjView *vols, const float *bscale, int n_classes,
                          const float *eulers, int O, const float2 *coords, int npix, int kpad,
                          unsigned char *proj, float *pabs2, hipStream_t s);

// Translation operand of the wide kernels with precomputed A (schedule 8):
// a.a_enc[row tile][chunk] = the LDS image of the tile's A chunk, for rows of
// tile_m (128 or 256) built by the same arithmetic as the in-kernel build.
hipError_t launch_encode_a(int form, int tile_m, const CrossArgs &a, unsigned char *a_enc,
                           hipStream_t s);

// Projections in the wide kernels' encoding: [C][kpad/16][O][16-pixel chunk]
// or, direct = true, [C][kpad/16][ceil(O/16)][16-orientation group], each
// group block holding the group's MFMA fragments in lane order
// (direct_seg_offset; orientations past O are zero).
hipError_t launch_project_wide(int form, const ProjView *vols, const float *bscale, int n_classes,
                               const float *eulers, int O, const float2 *coords, int npix,
                               int kpad, unsigned char *proj, float *pabs2, bool direct,
                               hipStream_t s);
// Bytes of the projection buffer of a plan (without the trailing kProjPad).
size_t projection_bytes(int form, bool direct, int n_classes, int O, int kpad);

// pnorm[i][c*O+o] = sum_p half_corr[i][p] * pabs2[c*O+o][p]
hipError_t launch_pnorm(int n_particles, int n_cols, int kpad, const float *half_corr,
                        const float *pabs2, float *pnorm, hipStream_t s);

// Split-K assembly of one image: out[o*T+t] = score from the fixed-order
// float64 sums over the S slices of part_c and part_pn (see CrossArgs).
// Drop-in projection cache.  cache_check: when forced, or where the eulers
// differ bit for bit from the cached copy, the copy is refreshed and the
// flag set to gen (the call's generation; the flag is never cleared).
// cache_fill: when the flag holds gen, P[o][p] = the projection of
// orientation o at pixel p (p < npix; the fused gather's arithmetic), else
// nothing.
hipError_t launch_cache_check(const float *eulers, float *copy, int n, int force, unsigned gen,
                              unsigned *flag, hipStream_t s);
hipError_t launch_cache_fill(const unsigned *flag, unsigned gen, const ProjView &v,
                             const float *eulers, int O, const float2 *coords, int npix, int kpad,
                             float2 *P, hipStream_t s);
hipError_t launch_splitk_assemble(int O, int T, int S, const float *part_c, const float *part_pn,
                                  const double *xnorm, const float *init, float *out, int cc,
                                  hipStream_t s);

// In-radius partition of n_images unfiltered images (one workgroup each).
hipError_t launch_filter(int image_size, int n_images, int imgX, int imgY, int maxR, float pad,
                         int maxR2_padded, const float *fimg, const float *corr, float *out_re,
                         float *out_im, float *out_corr, float *coor_xy, int *filtered_size,
                         hipStream_t s);

// Single-image path (capi.cpp), shared with the allocation-free entry point:
// workspace bytes of a shape and the run on a caller-provided workspace.
size_t single_workspace_bytes(int T, int O, int npix, int form);
int single_split(int T, int O, int npix, int form, int *ksplit, int *kc_per_split);
int run_single_ws(int T, int O, int npix, int full_npix, int form, const float *eulers,
                  const float *trans_x, const float *trans_y, const float *real, const float *imag,
                  const rc_projector_kernel *pk, const float *corr, const float *init, float *dest,
                  hipStream_t stream, const float *coor_xy, void *ws, size_t ws_bytes, bool cc,
                  bool partition);
// Pixels of an unfiltered drop-in image of image_size pixels that a rotation
// can bring inside the projector's mask (the search runs over these, moved
// first by launch_prep_partition); 0 when none is.  *pad2 and *thr receive
// the test's float constants.
int dropin_pixels(const rc_projector_kernel &k, int image_size, bool cc, float *pad2,
                  float *thr);
// Argument checks of rc_run_diff2_kernel_coarse.
int check_drop_in(const rc_projector_kernel *pk, unsigned long O, unsigned long T,
                  unsigned long image_size, int data_is_3D);

// Per-particle min / argmin from the cross kernel's partials.
hipError_t launch_argmin(int n_particles, const int *group_begin, int parts_per_group,
                         const double *part_min, const long long *part_idx, double *best,
                         long long *best_idx, hipStream_t s);

// Coarse weights and significance of a batch (one workgroup per particle).
struct WeightsArgs {
    int n_particles, n_classes, orientation_num;
    const int *trans_offset;     // [P+1] device
    const float *diff2;          // batch layout of the search output
    const double *min_diff2;     // [P]
    const float *lp_orient;      // [C][O] per particle at i * orient_stride, or NULL
    long long orient_stride;
    const float *lp_offset;      // [C][T_i] at C * trans_offset[i], or NULL
    double fraction;             // adaptive fraction of the weight kept significant
    float *weights;              // batch layout (may alias diff2)
    double *sum_weight;          // [P]
    float *significant_weight;   // [P]
    int *n_significant;          // [P]
    unsigned char *mask;         // batch layout or NULL
    double *max_log_weight;      // [P] or NULL
};
hipError_t launch_weights(const WeightsArgs &a, hipStream_t s);

// Fine-pass diff2 over job lists (one workgroup per job).
struct FineArgs {
    ProjView v;
    int imgX, imgY, maxR;
    const float *eulers;
    const float *real, *imag, *corr;
    const float *trans_x, *trans_y;
    float *diff2;
    int image_size;
    float sum_init;
    unsigned long orientation_num, translation_num;
    const unsigned long *rot_idx, *trans_idx, *job_idx, *job_num;
};
hipError_t launch_fine(const FineArgs &a, unsigned long n_jobs, hipStream_t s);

// Direct-form float32 vector kernel (the original kernel's arithmetic).
hipError_t launch_direct(const ProjView &v, const float *eulers, int O, const float *trans_x,
                         const float *trans_y, int T, const float2 *coords, const float *real,
                         const float *imag, const float *corr, int npix, int full_npix,
                         const float *init, float *out, hipStream_t s);

}  // namespace rc

// opaque handle of the C ABI
struct rc_projector {
    float2 *vol;        // device, interleaved complex
    int xdim, ydim, zdim, inity, initz, maxr;
    float pad;
    float max_abs;      // max |re|, |im| over the volume
    float *d_scale;     // device float: power-of-two scale of max_abs (FP16 form)
    int device;
    size_t bytes;
    unsigned long long generation;   // unique per created projector (cache keys)
};
