// rc_run_diff2_kernel_coarse_noalloc: the single-image drop-in with its
// scratch in a file-scope device array, for callers that cannot allocate
// (code under hipGraph capture, tight serving loops).  No call allocates or
// frees device memory or synchronises the host.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "rc_internal.hpp"

namespace rc {

// 24 MiB: image staging (20 bytes per pixel) plus the split-K partial sums,
// which the slice rule's 512-workgroup bound keeps under 17.1 MB (a
// 113-translation, 4608-orientation, 2112-pixel particle needs 14.7 MB in
// all).  Each GPU the library runs on holds one copy until the process ends.
constexpr size_t kScratchBytes = 24ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device, resolved once per device with
// that device current.  The host-side name is a stand-in, never a device
// address: only these pointers reach kernels.
static std::mutex g_scratch_mu;
static std::vector<void *> g_scratch_ptr;

static int scratch_ptr(void **out)
{
    int dev = 0;
    RC_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        RC_REQUIRE(e == hipSuccess && p, RC_ERR_DEVICE,
                   "hipGetSymbolAddress of the static scratch failed on device %d: %s", dev,
                   hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return RC_OK;
}

}  // namespace rc

using namespace rc;

extern "C" int rc_run_diff2_kernel_coarse_noalloc(
    const rc_projector_kernel *projector, const float *d_trans_x, const float *d_trans_y,
    const float *d_trans_z, const float *d_corr_img, const float *d_fimg_real,
    const float *d_fimg_imag, const float *d_eulers, float *d_diff2s, float local_sqrtXi2,
    unsigned long orientation_num, unsigned long translation_num, unsigned long image_size,
    void *stream, int do_CC, int data_is_3D)
{
    (void)d_trans_z;
    (void)local_sqrtXi2;
    if (int st = check_drop_in(projector, orientation_num, translation_num, image_size,
                               data_is_3D))
        return st;
    float pad2, thr;
    const int n_in = dropin_pixels(*projector, (int)image_size, do_CC != 0, &pad2, &thr);
    const bool part = n_in > 0;
    const int npix = part ? n_in : (int)image_size;
    const size_t need = single_workspace_bytes((int)translation_num, (int)orientation_num, npix,
                                               RC_FORM_FP32);
    RC_REQUIRE(need <= kScratchBytes, RC_ERR_LIMIT,
               "this shape needs %zu bytes of scratch, beyond the %zu-byte static scratch of "
               "the allocation-free entry point (use rc_run_diff2_kernel_coarse)",
               need, kScratchBytes);
    void *ws = nullptr;
    if (int st = scratch_ptr(&ws)) return st;
    return run_single_ws((int)translation_num, (int)orientation_num, npix, (int)image_size,
                         RC_FORM_FP32, d_eulers, d_trans_x, d_trans_y, d_fimg_real, d_fimg_imag,
                         projector, d_corr_img, d_diff2s, d_diff2s, (hipStream_t)stream, nullptr,
                         ws, need, do_CC != 0, part);
}
