// A C-ABI-only caller of the library (no Python): the flow a RELION-side
// integration takes for a batch of particles.
//
//   1. a projector from a host Fourier volume (rc_projector_create) and its
//      view for a 32^2 coarse image (rc_projector_make_kernel);
//   2. unfiltered half-plane images, as the reference's fimg (real block,
//      then imaginary block) and corr, partitioned in-radius-first on the
//      device (rc_filter_coarse_images; the in-radius count from
//      rc_filter_coarse_order);
//   3. one batched search over every particle, with a different number of
//      translations per particle (rc_plan_create / rc_plan_execute), in the
//      FP32 and BF16x3 forms;
//   4. a check of each particle's best (orientation, translation) against the
//      single-image drop-in (rc_run_diff2_kernel_coarse) run on the same
//      unfiltered image: the plan's best sample must hold the drop-in's
//      minimum to float32 rounding.
//
// Inputs are deterministic synthetic data.  Prints one line per form and
// exits 0 when every check holds.  Build: `make -C t2-relion-kernel_amd
// examples` (the binary looks for the library in the directory above it).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "relion_coarse.h"

namespace {

int fail(const char *what, int st)
{
    std::fprintf(stderr, "coarse_search: %s failed (%d): %s\n", what, st, rc_last_error());
    return 1;
}

#define CHECK_RC(call)                       \
    do {                                     \
        const int st_ = (call);              \
        if (st_ != RC_OK) return fail(#call, st_); \
    } while (0)
#define CHECK_HIP(call)                                                               \
    do {                                                                              \
        const hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                       \
            std::fprintf(stderr, "coarse_search: %s: %s\n", #call, hipGetErrorString(e_)); \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

// Deterministic pseudo-random numbers in [0, 1).
struct Lcg {
    uint64_t s;
    double next()
    {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(s >> 11) * 0x1p-53;
    }
};

template <class T>
T *to_device(const std::vector<T> &h)
{
    T *d = nullptr;
    if (hipMalloc(&d, h.size() * sizeof(T)) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    return d;
}

}  // namespace

int main()
{
    // ---- projector: a smooth half-complex volume, x in [0, 34), y and z in
    // [-33, 33], model radius 16 (unpadded), padding factor 2
    const int X = 34, Y = 67, Z = 67, initY = -33, initZ = -33, maxr = 16;
    const float pad = 2.f;
    std::vector<float> vre((size_t)X * Y * Z), vim(vre.size());
    for (int z = 0; z < Z; ++z)
        for (int y = 0; y < Y; ++y)
            for (int x = 0; x < X; ++x) {
                const double kx = x, ky = y + initY, kz = z + initZ;
                const double r2 = kx * kx + ky * ky + kz * kz;
                const double env = std::exp(-r2 / (2.0 * 14.0 * 14.0));
                const size_t i = ((size_t)z * Y + y) * X + x;
                vre[i] = (float)(env * std::cos(0.31 * kx - 0.17 * ky + 0.23 * kz));
                vim[i] = (float)(env * std::sin(0.19 * kx + 0.29 * ky - 0.13 * kz));
            }
    rc_projector *proj = nullptr;
    CHECK_RC(rc_projector_create(X, Y, Z, initY, initZ, maxr, pad, vre.data(), vim.data(), &proj));
    rc_projector_kernel view;
    const int imgX = 17, imgY = 32, imgMaxR = 16, image_size = imgX * imgY;
    CHECK_RC(rc_projector_make_kernel(proj, imgX, imgY, 1, imgMaxR, &view));

    // ---- orientations: 8 tilts x 12 in-plane angles (ZYZ rotation matrices)
    std::vector<float> eul;
    for (int a = 0; a < 8; ++a)
        for (int p = 0; p < 12; ++p) {
            const double rot = 0.7 * a, tilt = 0.35 * a + 0.1, psi = 2.0 * M_PI * p / 12.0;
            const double ca = std::cos(rot), sa = std::sin(rot), cb = std::cos(tilt),
                         sb = std::sin(tilt), cg = std::cos(psi), sg = std::sin(psi);
            const double cc = cb * ca, cs = cb * sa;
            const double m[9] = {cg * cc - sg * sa, cg * cs + sg * ca, -cg * sb,
                                 -sg * cc - cg * sa, -sg * cs + cg * ca, sg * sb,
                                 sb * ca, sb * sa, cb};
            for (double v : m) eul.push_back((float)v);
        }
    const int O = (int)eul.size() / 9;

    // ---- particles: unfiltered images and per-particle translation sets
    const int P = 6;
    const int radius2[P] = {2, 5, 8, 5, 2, 4};   // integer shifts with dx^2 + dy^2 <= r2
    std::vector<int> toff(1, 0);
    std::vector<float> tx, ty;
    for (int i = 0; i < P; ++i) {
        for (int dy = -3; dy <= 3; ++dy)
            for (int dx = -3; dx <= 3; ++dx)
                if (dx * dx + dy * dy <= radius2[i]) {
                    tx.push_back((float)(2.0 * M_PI * dx / imgY));
                    ty.push_back((float)(2.0 * M_PI * dy / imgY));
                }
        toff.push_back((int)tx.size());
    }
    Lcg rng{12345};
    std::vector<float> fimg((size_t)P * 2 * image_size), corr((size_t)P * image_size);
    for (float &v : fimg) v = (float)(rng.next() - 0.5);
    for (float &v : corr) v = (float)(0.5 + rng.next());

    hipStream_t s;
    CHECK_HIP(hipStreamCreate(&s));
    float *d_fimg = to_device(fimg), *d_corr_in = to_device(corr);
    float *d_eul = to_device(eul), *d_tx = to_device(tx), *d_ty = to_device(ty);
    int *d_toff = to_device(toff);
    if (!d_fimg || !d_corr_in || !d_eul || !d_tx || !d_ty || !d_toff) {
        std::fprintf(stderr, "coarse_search: device upload failed\n");
        return 1;
    }

    // ---- in-radius partition on the device
    int n_inside = 0;
    CHECK_RC(rc_filter_coarse_order(&view, image_size, nullptr, &n_inside));
    float *d_re, *d_im, *d_corr, *d_xy;
    CHECK_HIP(hipMalloc(&d_re, sizeof(float) * P * image_size));
    CHECK_HIP(hipMalloc(&d_im, sizeof(float) * P * image_size));
    CHECK_HIP(hipMalloc(&d_corr, sizeof(float) * P * image_size));
    CHECK_HIP(hipMalloc(&d_xy, sizeof(float) * 2 * image_size));
    CHECK_RC(rc_filter_coarse_images(&view, image_size, P, d_fimg, d_corr_in, d_re, d_im, d_corr,
                                     d_xy, nullptr, s));

    // ---- the single-image drop-in on each unfiltered image (reference answer)
    const int total = toff[P];
    std::vector<float> drop((size_t)O * total);
    float *d_drop;
    CHECK_HIP(hipMalloc(&d_drop, sizeof(float) * O * total));
    CHECK_HIP(hipMemsetAsync(d_drop, 0, sizeof(float) * O * total, s));
    for (int i = 0; i < P; ++i) {
        const int T = toff[i + 1] - toff[i];
        const float *re = d_fimg + (size_t)i * 2 * image_size;
        CHECK_RC(rc_run_diff2_kernel_coarse(&view, d_tx + toff[i], d_ty + toff[i], nullptr,
                                            d_corr_in + (size_t)i * image_size, re,
                                            re + image_size, d_eul, d_drop + (size_t)O * toff[i],
                                            1.f, O, T, image_size, s, 0, 0));
    }
    CHECK_HIP(hipMemcpyAsync(drop.data(), d_drop, sizeof(float) * drop.size(),
                             hipMemcpyDeviceToHost, s));
    CHECK_HIP(hipStreamSynchronize(s));

    // ---- one batched search per form
    float *d_diff2;
    double *d_best;
    int64_t *d_idx;
    CHECK_HIP(hipMalloc(&d_diff2, sizeof(float) * O * total));
    CHECK_HIP(hipMalloc(&d_best, sizeof(double) * P));
    CHECK_HIP(hipMalloc(&d_idx, sizeof(int64_t) * P));
    int bad = 0;
    for (int form : {RC_FORM_FP32, RC_FORM_BF16X3}) {
        rc_batch_desc d = {};
        d.n_particles = P;
        d.n_classes = 1;
        d.orientation_num = O;
        d.npix = n_inside;
        d.full_npix = image_size;
        d.form = form;
        d.projectors = &view;
        d.host_trans_offset = toff.data();
        d.d_trans_offset = d_toff;
        d.d_trans_x = d_tx;
        d.d_trans_y = d_ty;
        d.d_eulers = d_eul;
        d.d_real = d_re;
        d.d_imag = d_im;
        d.d_corr = d_corr;
        d.d_coor_xy = d_xy;
        d.d_diff2 = d_diff2;
        d.d_best_diff2 = d_best;
        d.d_best_index = d_idx;
        rc_plan *plan = nullptr;
        CHECK_RC(rc_plan_create(&d, &plan));
        CHECK_RC(rc_plan_execute(plan, &d, s));
        std::vector<double> best(P);
        std::vector<int64_t> idx(P);
        CHECK_HIP(hipMemcpyAsync(best.data(), d_best, sizeof(double) * P, hipMemcpyDeviceToHost, s));
        CHECK_HIP(hipMemcpyAsync(idx.data(), d_idx, sizeof(int64_t) * P, hipMemcpyDeviceToHost, s));
        CHECK_HIP(hipStreamSynchronize(s));
        int variant = -1, tm = 0, tn = 0;
        CHECK_RC(rc_plan_tile(plan, &variant, &tm, &tn));
        int agree = 0;
        double worst = 0.0;
        for (int i = 0; i < P; ++i) {
            const int T = toff[i + 1] - toff[i];
            const float *ref = drop.data() + (size_t)O * toff[i];
            int k = 0;
            for (int j = 1; j < O * T; ++j)
                if (ref[j] < ref[k]) k = j;
            const bool in_range = idx[i] >= 0 && idx[i] < (int64_t)O * T;
            // the plan's best sample holds the drop-in's minimum to float32
            // rounding, and the plan's minimum is that value
            const double at = in_range ? ref[idx[i]] : INFINITY;
            const double rel = std::fabs(at - ref[k]) / std::fabs(ref[k]);
            const double rel_best = std::fabs(best[i] - ref[k]) / std::fabs(ref[k]);
            worst = std::fmax(worst, std::fmax(rel, rel_best));
            agree += idx[i] == k;
            if (!in_range || rel > 1e-5 || rel_best > 1e-5) ++bad;
        }
        std::printf("coarse_search %s: %d particles x %d orientations x %d-%d translations x "
                    "%d/%d px, variant %d (%dx%d): arg-min equal to the drop-in's for %d/%d, "
                    "max rel gap %.2e\n",
                    form == RC_FORM_FP32 ? "fp32" : "bf16x3", P, O, toff[1] - toff[0],
                    toff[3] - toff[2], n_inside, image_size, variant, tm, tn, agree, P, worst);
        CHECK_RC(rc_plan_destroy(plan));
    }
    for (void *p : {(void *)d_fimg, (void *)d_corr_in, (void *)d_eul, (void *)d_tx, (void *)d_ty,
                    (void *)d_toff, (void *)d_re, (void *)d_im, (void *)d_corr, (void *)d_xy,
                    (void *)d_drop, (void *)d_diff2, (void *)d_best, (void *)d_idx})
        CHECK_HIP(hipFree(p));
    CHECK_RC(rc_scratch_release(s));
    CHECK_HIP(hipStreamDestroy(s));
    CHECK_RC(rc_projector_destroy(proj));
    if (bad) {
        std::fprintf(stderr, "coarse_search: %d particle checks failed\n", bad);
        return 1;
    }
    std::printf("coarse_search ok\n");
    return 0;
}
