// Fine-pass diff2 over job lists: after the coarse pass marks significant
// (orientation, translation) samples, the fine pass scores their
// oversampled neighbours, grouped into jobs.  Job b scores orientation
// rot_idx[job_idx[b]] against translations trans_idx[job_idx[b]] + j,
// j < job_num[b], writing diff2[job_idx[b] + j] = sum_init +
// sum_p corr_p/2 |P(p) - X_p e^{i phi}|^2 over the whole (unfiltered) image
// (cuda_kernel_diff2_fine, diff2.cuh:540-682).  Pixel coordinates follow that
// kernel: x = p % imgX, y = p / imgX, and a row y > maxR is wrapped to
// y - imgY when y >= imgY - maxR, else parked at x = maxR (diff2.cuh:593-617).
//
// One workgroup per job; each thread projects its pixels once per group of
// up to 16 translations and keeps one float accumulator per translation (only
// the job's own translations are scored: jobs of 4 do a quarter of the phase
// work of jobs of 16); the workgroup sums them in float64 in a fixed order.
#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace rc {

constexpr int kFT = 256;       // threads per job
constexpr int kFChunk = 16;    // translations per pass over the image
constexpr double kInv2PiF = 0.15915494309189533577;

__global__ __launch_bounds__(kFT) void fine_kernel(const FineArgs a, unsigned n_jobs)
{
    __shared__ double s_red[kFChunk][kFT / 64];
    // XCD-aware job order: block i runs on XCD i % 8, which takes the
    // contiguous job range [x * per, (x + 1) * per), so consecutive jobs --
    // neighbouring orientations, whose central sections share voxels --
    // gather through one L2
    const unsigned per = (n_jobs + 7) / 8;
    const unsigned bj = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (bj >= n_jobs) return;
    const int b = (int)bj;
    const int tid = threadIdx.x;
    const unsigned long j0 = a.job_idx[b];
    const unsigned long ix = a.rot_idx[j0];
    const unsigned long it0 = a.trans_idx[j0];
    const unsigned long nj = a.job_num[b];
    // a job naming an orientation or translation outside the arrays is
    // skipped (its outputs are left as they were)
    if (ix >= a.orientation_num || it0 >= a.translation_num || nj > a.translation_num - it0)
        return;
    const int n = (int)nj;
    const float *e = a.eulers + 9 * ix;
    const float e0 = e[0], e1 = e[1], e3 = e[3], e4 = e[4], e6 = e[6], e7 = e[7];

    for (int c0 = 0; c0 < n; c0 += kFChunk) {
        const int cn = min(kFChunk, n - c0);
        double txr[kFChunk], tyr[kFChunk];
#pragma unroll
        for (int j = 0; j < kFChunk; ++j) {
            const unsigned long it = it0 + c0 + min(j, cn - 1);
            txr[j] = (double)a.trans_x[it] * kInv2PiF;
            tyr[j] = (double)a.trans_y[it] * kInv2PiF;
        }
        float acc[kFChunk];
#pragma unroll
        for (int j = 0; j < kFChunk; ++j) acc[j] = 0.f;
        for (int p = tid; p < a.image_size; p += kFT) {
            int x = p % a.imgX, y = p / a.imgX;
            if (y > a.maxR) {
                if (y >= a.imgY - a.maxR)
                    y -= a.imgY;
                else
                    x = a.maxR;
            }
            const float fx = (float)x, fy = (float)y;
            const float2 P = project_sample(a.v, e0, e1, e3, e4, e6, e7, fx, fy);
            const float xr = a.real[p], xi = a.imag[p], hw = 0.5f * a.corr[p];
#pragma unroll
            for (int j = 0; j < kFChunk; ++j) {
                if (j < cn) {     // uniform over the workgroup
                    float s, c;
                    phase_factor(fx, fy, txr[j], tyr[j], s, c);
                    const float dr = P.x - (c * xr - s * xi);
                    const float di = P.y - (c * xi + s * xr);
                    acc[j] += (dr * dr + di * di) * hw;
                }
            }
        }
        // fixed-order block sums in float64: wave butterflies, then the waves
        const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
        for (int j = 0; j < kFChunk; ++j) {
            if (j < cn) {
                double v = (double)acc[j];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
                if (lane == 0) s_red[j][wave] = v;
            }
        }
        __syncthreads();
        if (tid < cn) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < kFT / 64; ++w) t += s_red[tid][w];
            a.diff2[j0 + c0 + tid] = (float)(t + (double)a.sum_init);
        }
        __syncthreads();
    }
}

hipError_t launch_fine(const FineArgs &a, unsigned long n_jobs, hipStream_t s)
{
    const unsigned per = (unsigned)((n_jobs + 7) / 8);
    hipLaunchKernelGGL(fine_kernel, dim3(8 * per), dim3(kFT), 0, s, a, (unsigned)n_jobs);
    return hipGetLastError();
}

}  // namespace rc
