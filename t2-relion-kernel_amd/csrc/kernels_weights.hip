// Coarse-pass weights and significance: the step after the diff2 search.
//
// Per particle i (one workgroup each), over its samples k = (c*O + o)*T_i + t:
//   lw_k = lp_orient[c][o] + lp_offset[c][t] + (min_i - diff2_k)     (float64)
//   w_k  = exp(lw_k - max_k lw_k)              (float64 exp, stored as float)
// samples with a -inf log prior, or with diff2 below the minimum, get w = 0
// (the coarse weights exponent and exponentiate steps, utilities.h:405-459);
//   S    = sum_k w_k                                                     (float64)
// and the significance threshold of the adaptive-oversampling pass: with the
// weights sorted ascending, s[0..n-1], and c[j] = s[0] + ... + s[j],
//   m  = #{ j : c[j] <= (1 - fraction) * S },   w* = s[min(m, n-1)],
// a sample is significant when w_k >= w* (find_threshold_idx_in_cumulative and
// array_over_threshold, cuda_device_utils.cuh:238-267).
//
// w* is found without sorting: it is the smallest float v with
// sum_{w <= v} w > (1 - fraction) * S, and non-negative floats order as
// their bit patterns, so each pass splits the candidate bit range into 16
// power-of-two sub-ranges, sums the weights and counts of each in float64,
// and keeps the sub-range where the running sum crosses the bound.  Every
// sum is formed in a fixed order (per-thread slots in LDS, then a fixed
// tree), so the result does not depend on timing or on the other particles.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "rc_internal.hpp"

namespace rc {

constexpr int kWT = 256;   // threads per particle
constexpr int kWB = 16;    // sub-ranges per selection pass

// Walks this thread's samples k = tid, tid + kWT, ... keeping (row, t) and
// the class of the row up to date without divisions in the loop.
struct SampleWalk {
    long long k;
    int row, t, c, o;
    int drow, dt;     // kWT = drow * T + dt
    int T, O;
    __device__ void init(int tid, int T_, int O_)
    {
        T = T_;
        O = O_;
        k = tid;
        row = tid / T;
        t = tid - row * T;
        c = row / O;
        o = row - c * O;
        drow = kWT / T;
        dt = kWT - drow * T;
    }
    __device__ void next()
    {
        k += kWT;
        t += dt;
        int dr = drow;
        if (t >= T) {
            t -= T;
            ++dr;
        }
        row += dr;
        o += dr;
        while (o >= O) {
            o -= O;
            ++c;
        }
    }
};

// Fixed-order sum of one value per thread; every thread gets the result.
__device__ double block_sum(double v, double *red)
{
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = kWT / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__device__ double block_max(double v, double *red)
{
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = kWT / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] = fmax(red[tid], red[tid + s]);
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kWT) void weights_kernel(const WeightsArgs a)
{
    __shared__ double s_red[kWT];
    __shared__ double s_acc[kWB][kWT];
    __shared__ int s_cnt[kWB][kWT];
    __shared__ double s_tot[kWB];
    __shared__ long long s_totc[kWB];

    const int i = blockIdx.x;
    const int tid = threadIdx.x;
    const int base = a.trans_offset[i];
    const int T = a.trans_offset[i + 1] - base;
    const int C = a.n_classes, O = a.orientation_num;
    const long long n = (long long)C * O * T;
    const size_t off = (size_t)C * O * base;
    const float *d = a.diff2 + off;
    float *w = a.weights + off;
    const double mind = a.min_diff2[i];
    const float minf = (float)mind;   // the arg-min sample's stored value
    const float *lpo = a.lp_orient ? a.lp_orient + (size_t)i * a.orient_stride : nullptr;
    const float *lpt = a.lp_offset ? a.lp_offset + (size_t)C * base : nullptr;

    auto log_weight = [&](const SampleWalk &s, float dk, bool &ok) {
        const float po = lpo ? lpo[s.row] : 0.f;
        const float pt = lpt ? lpt[s.c * T + s.t] : 0.f;
        ok = !(dk < minf) && po != -INFINITY && pt != -INFINITY;
        return (double)po + (double)pt + (mind - (double)dk);
    };

    // ---- maximum log weight
    double lmax = -DBL_MAX;
    {
        SampleWalk s;
        s.init(tid, T, O);
        for (; s.k < n; s.next()) {
            bool ok;
            const double lw = log_weight(s, d[s.k], ok);
            if (ok) lmax = fmax(lmax, lw);
        }
    }
    lmax = block_max(lmax, s_red);
    const bool any = lmax > -DBL_MAX;

    // ---- weights and their sum
    double sum = 0.0;
    {
        SampleWalk s;
        s.init(tid, T, O);
        for (; s.k < n; s.next()) {
            bool ok;
            const double lw = log_weight(s, d[s.k], ok);
            const float wk = (ok && any) ? (float)exp(lw - lmax) : 0.f;
            w[s.k] = wk;
            sum += (double)wk;
        }
    }
    __syncthreads();   // the weights of this block are in memory for the passes
    sum = block_sum(sum, s_red);
    const double bound = (1.0 - a.fraction) * sum;

    // ---- threshold selection over the bit patterns of w in [0, 1]
    unsigned lo = 0u, hi = 0x3f800000u;   // bits of 0.0f .. 1.0f
    double below = 0.0;                   // sum of weights with bits < lo
    long long nbelow = 0;                 // their count
    bool crossed = false;                 // a pass has found the crossing
    for (;;) {
        const unsigned span = hi - lo;    // bit range [lo, hi]
        int shift = 0;
        while ((span >> shift) >= (unsigned)kWB) ++shift;   // (span >> shift) < 16
        for (int j = 0; j < kWB; ++j) {
            s_acc[j][tid] = 0.0;
            s_cnt[j][tid] = 0;
        }
        for (long long k = tid; k < n; k += kWT) {
            const float wk = w[k];
            const unsigned b = __float_as_uint(wk);
            if (b >= lo && b <= hi) {
                const int j = (int)((b - lo) >> shift);
                s_acc[j][tid] += (double)wk;
                s_cnt[j][tid] += 1;
            }
        }
        __syncthreads();
        if (tid < kWB) {
            double t = 0.0;
            long long c = 0;
            for (int u = 0; u < kWT; ++u) {
                t += s_acc[tid][u];
                c += s_cnt[tid][u];
            }
            s_tot[tid] = t;
            s_totc[tid] = c;
        }
        __syncthreads();
        // every thread walks the 16 totals the same way
        int sel = -1, first = -1, last = -1;
        double cum = below;
        long long ccum = nbelow;
        double cum_sel = below;
        long long ccum_sel = nbelow;
        for (int j = 0; j < kWB; ++j) {
            if (s_totc[j] > 0) {
                if (first < 0) first = j;
                last = j;
            }
            if (sel < 0 && cum + s_tot[j] > bound) {
                sel = j;
                cum_sel = cum;
                ccum_sel = ccum;
            }
            cum += s_tot[j];
            ccum += s_totc[j];
        }
        if (sel < 0 && crossed) {
            // the parent range crossed the bound but its parts, summed in
            // another grouping, round just below it: the crossing is at the
            // top of the range
            sel = last < 0 ? 0 : last;
            cum_sel = below;
            ccum_sel = nbelow;
            for (int j = 0; j < sel; ++j) {
                cum_sel += s_tot[j];
                ccum_sel += s_totc[j];
            }
        } else if (sel < 0) {
            // the bound is not below the total (only through rounding): the
            // reference's scan then finds no pair and keeps index 0, the
            // smallest weight -- follow the first non-empty sub-range down
            sel = first < 0 ? 0 : first;
            cum_sel = below;
            ccum_sel = nbelow;
        } else {
            crossed = true;
        }
        __syncthreads();
        const unsigned nlo = lo + ((unsigned)sel << shift);
        const unsigned nhi = min(hi, nlo + ((1u << shift) - 1u));
        below = cum_sel;
        nbelow = ccum_sel;
        lo = nlo;
        hi = nhi;
        if (shift == 0) break;
    }
    const float wstar = __uint_as_float(lo);

    if (tid == 0) {
        a.sum_weight[i] = sum;
        a.significant_weight[i] = wstar;
        a.n_significant[i] = (int)(n - nbelow);
        if (a.max_log_weight) a.max_log_weight[i] = any ? lmax : -INFINITY;
    }
    if (a.mask) {
        unsigned char *m = a.mask + off;
        for (long long k = tid; k < n; k += kWT) m[k] = w[k] >= wstar ? 1 : 0;
    }
}

hipError_t launch_weights(const WeightsArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(weights_kernel, dim3(a.n_particles), dim3(kWT), 0, s, a);
    return hipGetLastError();
}

}  // namespace rc
