// In-radius partition of coarse images on the device: the pixel order of
// FilterCoarseParamImage (main.cu:88-151).  Pixel i of the half-plane image
// has integer coordinates x = i % imgX, y = i / imgX, y -= imgY if y > maxR
// (pixel_index2coor, diff2.cuh:182-188); it is inside when the float
// r2 = (x*x + y*y) * pad * pad is at most maxR2_padded.  Inside pixels keep
// their index order at the head; outside pixels fill the tail from the last
// position down, so the first outside pixel ends up last.
//
// One workgroup per image: the image is walked in 1024-pixel steps; a wave
// ballot gives each pixel's rank among the inside pixels of its wave, a
// 16-entry scan the wave offsets, and a running count the step offsets.
#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace rc {

constexpr int kFilterThreads = 1024;

__device__ __forceinline__ bool filter_inside(int i, int imgX, int imgY, int maxR, float pad,
                                              int maxR2_padded, int &x, int &y)
{
    x = i % imgX;
    y = i / imgX;
    if (y > maxR) y -= imgY;
    // (x*x + y*y) is an int, then float products, as in the reference
    const float r2 = (float)(x * x + y * y) * pad * pad;
    return r2 <= (float)maxR2_padded;
}

__global__ __launch_bounds__(kFilterThreads) void filter_kernel(
    int image_size, int imgX, int imgY, int maxR, float pad, int maxR2_padded,
    const float *__restrict__ fimg, const float *__restrict__ corr, float *__restrict__ out_re,
    float *__restrict__ out_im, float *__restrict__ out_corr, float *__restrict__ coor_xy,
    int *__restrict__ filtered_size)
{
    __shared__ int s_wave[kFilterThreads / 64];
    __shared__ int s_base;
    const int img = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t src = (size_t)img * 2 * image_size;
    const size_t dst = (size_t)img * image_size;
    if (threadIdx.x == 0) s_base = 0;
    __syncthreads();
    for (int i0 = 0; i0 < image_size; i0 += kFilterThreads) {
        const int i = i0 + (int)threadIdx.x;
        int x = 0, y = 0;
        const bool in = i < image_size &&
                        filter_inside(i, imgX, imgY, maxR, pad, maxR2_padded, x, y);
        const unsigned long long m = __ballot(in);
        const int below = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
        if (lane == 0) s_wave[wave] = __popcll(m);
        __syncthreads();
        int before = s_base;   // inside pixels before this step
        for (int w = 0; w < wave; ++w) before += s_wave[w];
        int total = 0;
        for (int w = 0; w < kFilterThreads / 64; ++w) total += s_wave[w];
        if (i < image_size) {
            const int rank_in = before + below;     // inside pixels with a lower index
            const int d = in ? rank_in : image_size - 1 - (i - rank_in);
            out_re[dst + d] = fimg[src + i];
            out_im[dst + d] = fimg[src + image_size + i];
            out_corr[dst + d] = corr[dst + i];
            if (coor_xy && img == 0) {
                coor_xy[2 * (size_t)d] = (float)x;
                coor_xy[2 * (size_t)d + 1] = (float)y;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) s_base += total;
        __syncthreads();
    }
    if (threadIdx.x == 0 && filtered_size) filtered_size[img] = s_base;
}

hipError_t launch_filter(int image_size, int n_images, int imgX, int imgY, int maxR, float pad,
                         int maxR2_padded, const float *fimg, const float *corr, float *out_re,
                         float *out_im, float *out_corr, float *coor_xy, int *filtered_size,
                         hipStream_t s)
{
    hipLaunchKernelGGL(filter_kernel, dim3(n_images), dim3(kFilterThreads), 0, s, image_size,
                       imgX, imgY, maxR, pad, maxR2_padded, fimg, corr, out_re, out_im, out_corr,
                       coor_xy, filtered_size);
    return hipGetLastError();
}

}  // namespace rc
