// Coarse-pass weights and significance: the step after the diff2 search.
//
// Per particle i (one workgroup each), over its samples k = (c*O + o)*T_i + t:
//   lw_k = lp_orient[c][o] + lp_offset[c][t] + (min_i - diff2_k)     (float64)
//   w_k  = exp(lw_k - max_k lw_k)              (float64 exp, stored as float)
// samples with a -inf log prior, or with diff2 below the minimum, get w = 0
// (the coarse weights exponent and exponentiate steps, utilities.h:405-459);
//   S    = sum_k w_k
// and the significance threshold of the adaptive-oversampling pass: with the
// weights sorted ascending, s[0..n-1], and c[j] = s[0] + ... + s[j],
//   m  = #{ j : c[j] <= (1 - fraction) * S },   w* = s[min(m, n-1)],
// a sample is significant when w_k >= w* (find_threshold_idx_in_cumulative and
// array_over_threshold, cuda_device_utils.cuh:238-267).
//
// w* is found without sorting: it is the smallest float v with
// sum_{w <= v} w > (1 - fraction) * S, and non-negative floats order as
// their bit patterns.  Three radix levels (bits 29..20, 19..10, 9..0 of the
// pattern; weights lie in [0, 1], patterns up to 0x3f800000) histogram the
// candidate range in LDS and keep the bucket where the running sum crosses
// the bound.  Sums are exact: each weight is split into three 32-bit fixed-
// point parts (units 2^-32, 2^-64, 2^-96; exact for w >= 2^-72, off by less
// than 2^-96 below), accumulated as 64-bit integers with atomics, so the
// totals are the same whatever the order -- the result does not depend on
// timing or on the other particles.  (A bound of exactly 0, fraction 1, is
// crossed by the first positive weight, however small.)
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "rc_internal.hpp"

namespace rc {

constexpr int kWT = 1024;        // threads per particle
constexpr int kWBins = 1024;     // buckets per radix level
constexpr int kWLevels = 3;
__device__ constexpr int kWShift[kWLevels] = {20, 10, 0};

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

// Fixed-point parts of a weight in [0, 1]: w ~ p0 2^-32 + p1 2^-64 + p2 2^-96
// (every step is exact: power-of-two scalings and differences of doubles
// that hold at most 24 significant bits).
struct Fixed {
    unsigned long long p[3];
};

__device__ __forceinline__ Fixed fixed_parts(float w)
{
    Fixed f;
    double d = (double)w;
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        d *= 4294967296.0;
        const double h = floor(d);
        f.p[u] = (unsigned long long)h;
        d -= h;
    }
    return f;
}

__device__ __forceinline__ double fixed_value(const unsigned long long (&p)[3])
{
    return (double)p[0] * 2.3283064365386962890625e-10 +
           (double)p[1] * 5.42101086242752217e-20 + (double)p[2] * 1.2621774483536189e-29;
}

__global__ __launch_bounds__(kWT) void weights_kernel(const WeightsArgs a)
{
    __shared__ unsigned long long h_p[3][kWBins];
    __shared__ unsigned h_cnt[kWBins];
    __shared__ double s_red[kWT / 64];
    __shared__ unsigned long long s_sum[3];
    __shared__ unsigned long long s_zero;
    __shared__ int s_sel;

    const int i = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
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
    auto clear_hist = [&]() {
        for (int b = tid; b < kWBins; b += kWT) {
            h_p[0][b] = h_p[1][b] = h_p[2][b] = 0ull;
            h_cnt[b] = 0u;
        }
    };
    auto add_bucket = [&](int b, const Fixed &f) {
        atomicAdd(&h_p[0][b], f.p[0]);
        atomicAdd(&h_p[1][b], f.p[1]);
        atomicAdd(&h_p[2][b], f.p[2]);
        atomicAdd(&h_cnt[b], 1u);
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
    for (int o2 = 32; o2 > 0; o2 >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, o2, 64));
    if (lane == 0) s_red[wave] = lmax;
    if (tid == 0) {
        s_sum[0] = s_sum[1] = s_sum[2] = 0ull;
        s_zero = 0ull;
    }
    clear_hist();
    __syncthreads();
    lmax = s_red[0];
    for (int u = 1; u < kWT / 64; ++u) lmax = fmax(lmax, s_red[u]);
    const bool any = lmax > -DBL_MAX;

    // ---- weights, their exact sum, and the first radix level
    {
        unsigned long long ps[3] = {0ull, 0ull, 0ull}, zeros = 0ull;
        SampleWalk s;
        s.init(tid, T, O);
        for (; s.k < n; s.next()) {
            bool ok;
            const double lw = log_weight(s, d[s.k], ok);
            const float wk = (ok && any) ? (float)exp(lw - lmax) : 0.f;
            w[s.k] = wk;
            if (wk > 0.f) {
                const Fixed f = fixed_parts(wk);
                ps[0] += f.p[0];
                ps[1] += f.p[1];
                ps[2] += f.p[2];
                add_bucket((int)(__float_as_uint(wk) >> kWShift[0]), f);
            } else {
                ++zeros;
            }
        }
        atomicAdd(&s_sum[0], ps[0]);
        atomicAdd(&s_sum[1], ps[1]);
        atomicAdd(&s_sum[2], ps[2]);
        atomicAdd(&s_zero, zeros);
    }
    __syncthreads();
    const unsigned long long tot[3] = {s_sum[0], s_sum[1], s_sum[2]};
    const double sum = fixed_value(tot);
    const double bound = (1.0 - a.fraction) * sum;
    const long long nzero = (long long)s_zero;

    // ---- radix levels over the positive weights: the candidate patterns
    // are [lo, lo + 2^(shift of the previous level)); below_* is the exact sum
    // and count of the positive weights under lo
    unsigned lo = 0u;
    unsigned long long below[3] = {0ull, 0ull, 0ull};
    long long nbelow = 0;
    bool crossed = false, all_pass = false;
    for (int level = 0; level < kWLevels; ++level) {
        const int shift = kWShift[level];
        if (level > 0) {
            __syncthreads();
            clear_hist();
            __syncthreads();
            const unsigned top = lo + ((1u << kWShift[level - 1]) - 1u);
            for (long long k = tid; k < n; k += kWT) {
                const float wk = w[k];
                const unsigned bits = __float_as_uint(wk);
                if (wk > 0.f && bits >= lo && bits <= top)
                    add_bucket((int)((bits - lo) >> shift), fixed_parts(wk));
            }
            __syncthreads();
        }
        // one thread walks the buckets; the bucket totals are exact, so the
        // walk is the same on every run
        if (tid == 0) {
            int sel = -1, first = -1, last = -1;
            unsigned long long c[3] = {below[0], below[1], below[2]};
            for (int b = 0; b < kWBins; ++b) {
                if (h_cnt[b] == 0u) continue;
                if (first < 0) first = b;
                last = b;
                c[0] += h_p[0][b];
                c[1] += h_p[1][b];
                c[2] += h_p[2][b];
                if (sel < 0 && (bound <= 0.0 || fixed_value(c) > bound)) sel = b;
            }
            if (sel < 0) sel = crossed ? last : (first < 0 ? -1 : -2 - first);
            s_sel = sel;
        }
        __syncthreads();
        int sel = s_sel;
        if (sel == -1) {          // no positive weight at all: every w is 0
            all_pass = true;
            break;
        }
        if (sel <= -2) {
            // The bound is not below the total (only through rounding): the
            // reference's scan finds no pair and keeps the smallest weight --
            // 0 if any weight is 0, else follow the first bucket down.
            if (level == 0 && nzero > 0) {
                all_pass = true;
                break;
            }
            sel = -2 - sel;
        }
        for (int b = 0; b < sel; ++b) {
            below[0] += h_p[0][b];
            below[1] += h_p[1][b];
            below[2] += h_p[2][b];
            nbelow += h_cnt[b];
        }
        const unsigned long long with_sel[3] = {below[0] + h_p[0][sel], below[1] + h_p[1][sel],
                                                below[2] + h_p[2][sel]};
        if (h_cnt[sel] > 0u && (bound <= 0.0 || fixed_value(with_sel) > bound)) crossed = true;
        lo += (unsigned)sel << shift;
    }
    // lo is now the pattern of w*
    const float wstar = all_pass ? 0.f : __uint_as_float(lo);
    const long long n_under = all_pass ? 0 : nzero + nbelow;

    if (tid == 0) {
        a.sum_weight[i] = sum;
        a.significant_weight[i] = wstar;
        a.n_significant[i] = (int)(n - n_under);
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
