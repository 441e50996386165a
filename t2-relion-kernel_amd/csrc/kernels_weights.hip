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
// sum_{w <= v} w > (1 - fraction) * S, and non-negative floats order as their
// bit patterns.  Three radix levels (bits 29..20, 19..10, 9..0 of the pattern;
// weights lie in [0, 1], patterns up to 0x3f800000) histogram the candidate
// range in LDS and keep the bucket where the running sum crosses the bound.
//
// Sums are exact.  The first level's buckets fix the exponent (bits 29..23),
// so a bucket's weights are m * 2^(e - 150) with one e and 24-bit integer
// significands m: the bucket adds the m as one 64-bit integer.  Bucket totals
// are then placed in a 192-bit fixed-point number with unit 2^-160 (every
// float weight, denormals included, is a whole multiple of 2^-149) and the
// buckets' running sums come from one block-wide prefix scan.  Integer sums do
// not depend on the order of the atomics, so the results are the same on
// every run and do not depend on the other particles in the batch.  (A bound
// of exactly 0, fraction 1, is crossed by the first positive weight.)
//
// Memory passes over the particle's samples: (1) read diff2 for the maximum
// -- skipped without priors, see below; (2) read diff2, write the weights and
// a zero mask byte, first-level histogram, and every positive weight kept in
// an LDS list.  The posterior is sharp, so most weights round to 0 and the
// list (at most kWCap entries) usually holds all positive ones: the last two
// radix levels and the mask bytes of the positive samples are then finished
// from LDS, 9 bytes of memory traffic per sample (13 with priors).  With more
// positive weights: (3) read the weights: the samples outside the selected
// first-level bucket get their mask byte, those inside (at most kWCap) go to
// LDS; a bucket holding more than kWCap samples takes two more histogram
// passes and a mask pass instead.
//
// Without priors every counted sample has lw = min_i - diff2_k <= min_i -
// minf (samples below minf, the float the search stores for the minimum, get
// w = 0), with equality at the minimum's own sample: pass 2 takes that bound
// as the maximum, checks it against the largest lw it sees, and is repeated
// with the true maximum when no sample attains the bound (only when the
// weights do not overwrite diff2; in place, pass 1 runs).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "rc_internal.hpp"

namespace rc {

constexpr int kWT = 1024;        // threads per particle
constexpr int kWBins = 1024;     // buckets per radix level (one per thread)
constexpr int kWCap = 4096;      // first-level bucket size finished in LDS
// bit shift of radix level 0, 1, 2
__host__ __device__ constexpr int wshift(int level) { return level == 0 ? 20 : level == 1 ? 10 : 0; }
static_assert(kWT == kWBins, "the selection scan gives each thread one bucket");

// 192-bit unsigned fixed point, value (w2 2^128 + w1 2^64 + w0) * 2^-160.
struct U192 {
    unsigned long long w0, w1, w2;
};

__device__ __forceinline__ U192 u192_zero() { return U192{0ull, 0ull, 0ull}; }

__device__ __forceinline__ U192 u192_add(const U192 &a, const U192 &b)
{
    U192 r;
    r.w0 = a.w0 + b.w0;
    unsigned long long c = r.w0 < a.w0;
    r.w1 = a.w1 + b.w1 + c;
    c = (r.w1 < a.w1) || (c && r.w1 == a.w1);
    r.w2 = a.w2 + b.w2 + c;
    return r;
}

// m * 2^(e - 150) for 1 <= e <= 127 (a bucket total, below 2^32): m shifted
// left by e + 10.
__device__ __forceinline__ U192 u192_bucket(unsigned long long m, int e)
{
    const int sh = e + 10;        // 11 .. 137
    if (sh < 64) return U192{m << sh, m >> (64 - sh), 0ull};
    if (sh < 128) {
        const int s = sh - 64;
        return U192{0ull, m << s, s ? m >> (64 - s) : 0ull};
    }
    return U192{0ull, 0ull, m << (sh - 128)};
}

__device__ __forceinline__ double u192_value(const U192 &v)
{
    return ((double)v.w2 * 0x1p-32 + (double)v.w1 * 0x1p-96) + (double)v.w0 * 0x1p-160;
}

__device__ __forceinline__ U192 shfl_up(const U192 &v, int off)
{
    return U192{__shfl_up(v.w0, off, 64), __shfl_up(v.w1, off, 64), __shfl_up(v.w2, off, 64)};
}

// Integer significand of a positive float and its exponent (denormals: 1).
__device__ __forceinline__ unsigned long long significand(unsigned bits)
{
    return (bits >> 23) ? ((bits & 0x7fffffu) | 0x800000u) : (bits & 0x7fffffu);
}
__device__ __forceinline__ int exponent_of(unsigned bits) { return max((int)(bits >> 23), 1); }

// Visits the n floats at src: f4(k, x) for each 16-byte-aligned group of four
// (k = index of its first element), f1(k, x) for the up to three samples
// before the first 16-byte boundary and the up to three after the last full
// group.  Four groups per thread are loaded before any is used.
template <class F4, class F1>
__device__ __forceinline__ void sweep(const float *src, long long n, F4 &&f4, F1 &&f1)
{
    const int tid = threadIdx.x;
    long long h = (long long)(((16u - ((unsigned)(uintptr_t)src & 15u)) & 15u) >> 2);
    if (h > n) h = n;
    if (tid < h) f1((long long)tid, src[tid]);
    const long long nv = (n - h) >> 2;
    const float4 *s4 = reinterpret_cast<const float4 *>(src + h);
    for (long long v0 = tid; v0 < nv; v0 += 4 * kWT) {
        float4 x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (v0 + u * kWT < nv) x[u] = s4[v0 + u * kWT];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (v0 + u * kWT < nv) f4(h + 4 * (v0 + u * kWT), x[u]);
    }
    const long long t = h + 4 * nv + tid;
    if (tid < 3 && t < n) f1(t, src[t]);
}

// Stores four bytes at m + k, as one word when m + k is 4-byte aligned.
__device__ __forceinline__ void store_mask4(unsigned char *m, long long k, bool vec, unsigned char b0,
                                            unsigned char b1, unsigned char b2, unsigned char b3)
{
    if (vec) {
        *reinterpret_cast<uchar4 *>(m + k) = make_uchar4(b0, b1, b2, b3);
    } else {
        m[k] = b0;
        m[k + 1] = b1;
        m[k + 2] = b2;
        m[k + 3] = b3;
    }
}

// Selection state carried from one radix level to the next.
struct SelState {
    U192 below;                  // exact sum of the positive weights under lo
    unsigned long long nbelow;   // their number
    bool crossed;                // the kept bucket took the running sum past the bound
    unsigned lo;                 // lowest pattern of the kept bucket
};

struct SelShared {
    U192 wtot[kWT / 64];
    unsigned long long wcnt[kWT / 64];
    int sel, first, last;
    U192 excl;
    unsigned long long excl_cnt;
    int cond;
    unsigned pick_cnt;
};

// One radix level: thread b holds bucket b (count hc, significand sum hs; the
// weights in it are hs * 2^(e - 150) with e = max(b >> 3, 1) on the first
// level, eff_e on the later ones).  A prefix scan gives each bucket the exact
// running sum up to and including it; the kept bucket is the first that
// crosses the bound (on the first level the bound is set here from the
// total).  When none crosses: the last non-empty bucket if the previous level
// crossed, else the first non-empty one with *fallback set (the bound is not
// below the total), else -1 (no positive weight).  Kept out of line: three
// inlined copies of the scan raise the kernel's register count past 128.
__device__ __noinline__ int select_level(int level, unsigned long long hs, unsigned hc, int eff_e,
                            double fraction, double &bound, U192 &total,
                            unsigned long long &total_cnt, const SelState &st, SelShared &ss,
                            bool &fallback)
{
    const int b = threadIdx.x;
    const int lane = b & 63, wave = b >> 6;
    const U192 v = hc ? u192_bucket(hs, level == 0 ? max(b >> 3, 1) : eff_e) : u192_zero();
    U192 inc = v;
    unsigned long long ic = hc;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const U192 y = shfl_up(inc, off);
        const unsigned long long yc = __shfl_up(ic, off, 64);
        if (lane >= off) {
            inc = u192_add(inc, y);
            ic += yc;
        }
    }
    if (lane == 63) {
        ss.wtot[wave] = inc;
        ss.wcnt[wave] = ic;
    }
    if (b == 0) {
        ss.sel = kWBins;
        ss.first = kWBins;
        ss.last = -1;
    }
    __syncthreads();
    U192 pre = st.below;
    unsigned long long pc = st.nbelow;
    for (int u = 0; u < wave; ++u) {
        pre = u192_add(pre, ss.wtot[u]);
        pc += ss.wcnt[u];
    }
    if (level == 0) {
        U192 t = u192_zero();
        unsigned long long tc = 0;
        for (int u = 0; u < kWT / 64; ++u) {
            t = u192_add(t, ss.wtot[u]);
            tc += ss.wcnt[u];
        }
        total = t;
        total_cnt = tc;
        bound = (1.0 - fraction) * u192_value(t);
    }
    U192 exw = shfl_up(inc, 1);
    unsigned long long excw = __shfl_up(ic, 1, 64);
    if (lane == 0) {
        exw = u192_zero();
        excw = 0;
    }
    const U192 excl = u192_add(pre, exw);
    const unsigned long long excl_c = pc + excw;
    const bool cond = hc > 0 && (bound <= 0.0 || u192_value(u192_add(excl, v)) > bound);
    if (cond) atomicMin(&ss.sel, b);
    if (hc > 0) {
        atomicMin(&ss.first, b);
        atomicMax(&ss.last, b);
    }
    __syncthreads();
    int pick;
    fallback = false;
    if (ss.sel < kWBins)
        pick = ss.sel;
    else if (st.crossed)
        pick = ss.last;
    else if (ss.first == kWBins)
        pick = -1;
    else {
        pick = ss.first;
        fallback = true;
    }
    if (b == pick) {
        ss.excl = excl;
        ss.excl_cnt = excl_c;
        ss.cond = cond;
        ss.pick_cnt = hc;
    }
    __syncthreads();
    return pick;
}

__global__ __launch_bounds__(kWT, 2) void weights_kernel(const WeightsArgs a)
{
    __shared__ unsigned long long h_sum[kWBins];
    __shared__ unsigned h_cnt[kWBins];
    __shared__ float cap_val[kWCap];
    __shared__ int cap_idx[kWCap];
    __shared__ SelShared ss;
    __shared__ double s_red[kWT / 64];
    __shared__ unsigned s_ncap;

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
    unsigned char *mk = a.mask ? a.mask + off : nullptr;
    const double mind = a.min_diff2[i];
    const float minf = (float)mind;   // the arg-min sample's stored value
    const float *lpo = a.lp_orient ? a.lp_orient + (size_t)i * a.orient_stride : nullptr;
    const float *lpt = a.lp_offset ? a.lp_offset + (size_t)C * base : nullptr;
    const bool priors = lpo || lpt;
    const double invT = 1.0 / (double)T, invO = 1.0 / (double)O;

    auto clear_hist = [&]() {
        for (int b = tid; b < kWBins; b += kWT) {
            h_sum[b] = 0ull;
            h_cnt[b] = 0u;
        }
    };
    // q = floor(x / y) for 0 <= x < 2^31 from a float64 reciprocal, corrected
    // by one step (the product is within 2^-21 of x / y)
    auto quot = [](long long x, int y, double inv) {
        long long q = (long long)(((double)x + 0.5) * inv);
        if (q * y > x) --q;
        else if ((q + 1) * y <= x) ++q;
        return q;
    };
    auto log_weight = [&](long long k, float dk, bool &ok) {
        float po = 0.f, pt = 0.f;
        if (priors) {
            const long long row = quot(k, T, invT);
            if (lpo) po = lpo[row];
            if (lpt) {
                const long long t = k - row * T;
                pt = lpt[quot(row, O, invO) * T + t];
            }
        }
        ok = !(dk < minf) && po != -INFINITY && pt != -INFINITY;
        return (double)po + (double)pt + (mind - (double)dk);
    };

    // block-wide maximum of one value per thread
    auto block_max = [&](double v) {
        for (int o2 = 32; o2 > 0; o2 >>= 1) v = fmax(v, __shfl_xor(v, o2, 64));
        if (lane == 0) s_red[wave] = v;
        __syncthreads();
        v = s_red[0];
        for (int u = 1; u < kWT / 64; ++u) v = fmax(v, s_red[u]);
        __syncthreads();
        return v;
    };

    // ---- pass 1: maximum log weight (with priors, or in place)
    const bool bound_max = !priors && (const void *)a.weights != (const void *)a.diff2;
    double lmax = bound_max ? mind - (double)minf : -DBL_MAX;
    if (!bound_max) {
        auto one = [&](long long k, float dk) {
            bool ok;
            const double lw = log_weight(k, dk, ok);
            if (ok) lmax = fmax(lmax, lw);
        };
        sweep(d, n,
              [&](long long k, float4 x) {
                  one(k, x.x);
                  one(k + 1, x.y);
                  one(k + 2, x.z);
                  one(k + 3, x.w);
              },
              one);
        lmax = block_max(lmax);
    }
    bool any = lmax > -DBL_MAX;

    // ---- pass 2: weights, zero mask bytes, the first-level histogram and
    // the list of positive weights; repeated once when the bound taken as
    // the maximum is not attained
    for (int attempt = 0;; ++attempt) {
        clear_hist();
        if (tid == 0) s_ncap = 0u;
        __syncthreads();
        double seen = -DBL_MAX;   // largest log weight of a counted sample
        // exp(x) rounds to float 0 for x < -104 (below 2^-150)
        auto weight = [&](long long k, float dk) {
            bool ok;
            const double lw = log_weight(k, dk, ok);
            if (ok) seen = fmax(seen, lw);
            const double x = lw - lmax;
            return (ok && any) ? (x < -104.0 ? 0.f : (float)exp(x)) : 0.f;
        };
        auto count = [&](long long k, float wk) {
            if (wk > 0.f) {
                const unsigned bits = __float_as_uint(wk);
                atomicAdd(&h_sum[bits >> wshift(0)], significand(bits));
                atomicAdd(&h_cnt[bits >> wshift(0)], 1u);
                // once the list has overflowed (dense weights) the shared
                // counter is left alone: it only records the overflow
                if (*(volatile unsigned *)&s_ncap <= (unsigned)kWCap) {
                    const unsigned slot = atomicAdd(&s_ncap, 1u);
                    if (slot < (unsigned)kWCap) {
                        cap_val[slot] = wk;
                        cap_idx[slot] = (int)k;
                    }
                }
            }
        };
        long long hw = (long long)(((16u - ((unsigned)(uintptr_t)d & 15u)) & 15u) >> 2);
        const bool wvec = (((uintptr_t)(w + hw)) & 15u) == 0;
        const bool mvec = mk && (((uintptr_t)(mk + hw)) & 3u) == 0;
        sweep(d, n,
              [&](long long k, float4 x) {
                  const float4 r = make_float4(weight(k, x.x), weight(k + 1, x.y),
                                               weight(k + 2, x.z), weight(k + 3, x.w));
                  if (wvec) {
                      *reinterpret_cast<float4 *>(w + k) = r;
                  } else {
                      w[k] = r.x;
                      w[k + 1] = r.y;
                      w[k + 2] = r.z;
                      w[k + 3] = r.w;
                  }
                  if (mk) store_mask4(mk, k, mvec, 0, 0, 0, 0);
                  count(k, r.x);
                  count(k + 1, r.y);
                  count(k + 2, r.z);
                  count(k + 3, r.w);
              },
              [&](long long k, float dk) {
                  const float r = weight(k, dk);
                  w[k] = r;
                  if (mk) mk[k] = 0;
                  count(k, r);
              });
        // Each wave's stores of this pass complete before the barriers below:
        // a significant sample's mask byte is written again, possibly by
        // another wave, after them.
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        seen = block_max(seen);   // (its barriers also publish the histogram)
        if (!bound_max || attempt > 0 || seen == lmax) break;
        lmax = seen;
        any = lmax > -DBL_MAX;
    }
    const bool listed = s_ncap <= (unsigned)kWCap;   // every positive weight is in the list

    // ---- first level: total, bound, kept bucket
    SelState st{u192_zero(), 0ull, false, 0u};
    double bound = 0.0;
    U192 total = u192_zero();
    unsigned long long npos = 0;
    bool fallback;
    bool all_pass = false;
    int pick = select_level(0, h_sum[tid], h_cnt[tid], 1, a.fraction, bound, total, npos, st, ss,
                            fallback);
    const long long nzero = n - (long long)npos;
    unsigned cnt0 = 0;
    if (pick < 0 || (fallback && nzero > 0)) {
        // no positive weight, or the bound is not below the total (only
        // through rounding): the reference's scan finds no pair and keeps the
        // smallest weight, 0
        all_pass = true;
    } else {
        st.below = ss.excl;
        st.nbelow = ss.excl_cnt;
        st.crossed = ss.cond;
        st.lo = (unsigned)pick << wshift(0);
        cnt0 = ss.pick_cnt;
    }
    const int eff_e = exponent_of(st.lo);

    if (!all_pass && listed) {
        // ---- the last two levels and the positive samples' mask bytes from
        // the list (every other sample has weight 0 < w*, mask 0 from pass 2)
        const int ncap = (int)s_ncap;
        for (int level = 1; level < 3; ++level) {
            const int shift = wshift(level);
            const unsigned top = st.lo + ((1u << wshift(level - 1)) - 1u);
            clear_hist();
            __syncthreads();
            for (int j = tid; j < ncap; j += kWT) {
                const unsigned bits = __float_as_uint(cap_val[j]);
                if (bits >= st.lo && bits <= top) {
                    atomicAdd(&h_sum[(bits - st.lo) >> shift], significand(bits));
                    atomicAdd(&h_cnt[(bits - st.lo) >> shift], 1u);
                }
            }
            __syncthreads();
            pick = select_level(level, h_sum[tid], h_cnt[tid], eff_e, a.fraction, bound, total,
                                npos, st, ss, fallback);
            st.below = ss.excl;
            st.nbelow = ss.excl_cnt;
            st.crossed = ss.cond;
            st.lo += (unsigned)pick << shift;
        }
        if (mk) {
            const float ws = __uint_as_float(st.lo);
            for (int j = tid; j < ncap; j += kWT)
                if (cap_val[j] >= ws) mk[cap_idx[j]] = 1;
        }
    } else if (!all_pass && cnt0 <= (unsigned)kWCap) {
        // ---- pass 3: mask bytes of the samples outside the kept bucket;
        // the samples inside it go to LDS
        const unsigned lo = st.lo, hi = st.lo + (1u << wshift(0));
        const float fhi = __uint_as_float(hi);
        if (tid == 0) s_ncap = 0u;
        __syncthreads();
        auto take = [&](long long k, float wk) -> unsigned char {
            const unsigned bits = __float_as_uint(wk);
            if (wk > 0.f && bits >= lo && bits < hi) {
                const unsigned slot = atomicAdd(&s_ncap, 1u);
                if (slot < (unsigned)kWCap) {
                    cap_val[slot] = wk;
                    cap_idx[slot] = (int)k;
                }
                return 0;
            }
            return wk >= fhi ? 1 : 0;
        };
        long long hw = (long long)(((16u - ((unsigned)(uintptr_t)w & 15u)) & 15u) >> 2);
        const bool mvec = mk && (((uintptr_t)(mk + hw)) & 3u) == 0;
        sweep(w, n,
              [&](long long k, float4 x) {
                  const unsigned char b0 = take(k, x.x), b1 = take(k + 1, x.y),
                                      b2 = take(k + 2, x.z), b3 = take(k + 3, x.w);
                  if (mk) store_mask4(mk, k, mvec, b0, b1, b2, b3);
              },
              [&](long long k, float wk) {
                  const unsigned char b0 = take(k, wk);
                  if (mk) mk[k] = b0;
              });
        __syncthreads();
        const int ncap = (int)min(s_ncap, (unsigned)kWCap);
        for (int level = 1; level < 3; ++level) {
            const int shift = wshift(level);
            const unsigned top = st.lo + ((1u << wshift(level - 1)) - 1u);
            clear_hist();
            __syncthreads();
            for (int j = tid; j < ncap; j += kWT) {
                const unsigned bits = __float_as_uint(cap_val[j]);
                if (bits >= st.lo && bits <= top) {
                    atomicAdd(&h_sum[(bits - st.lo) >> shift], significand(bits));
                    atomicAdd(&h_cnt[(bits - st.lo) >> shift], 1u);
                }
            }
            __syncthreads();
            pick = select_level(level, h_sum[tid], h_cnt[tid], eff_e, a.fraction, bound, total,
                                npos, st, ss, fallback);
            st.below = ss.excl;
            st.nbelow = ss.excl_cnt;
            st.crossed = ss.cond;
            st.lo += (unsigned)pick << shift;
        }
        if (mk) {
            const float ws = __uint_as_float(st.lo);
            for (int j = tid; j < ncap; j += kWT) mk[cap_idx[j]] = cap_val[j] >= ws ? 1 : 0;
        }
    } else {
        if (!all_pass) {
            // ---- passes 3 and 4: the last two levels from the weights
            for (int level = 1; level < 3; ++level) {
                const int shift = wshift(level);
                const unsigned lo = st.lo, top = st.lo + ((1u << wshift(level - 1)) - 1u);
                clear_hist();
                __syncthreads();
                auto add = [&](float wk) {
                    const unsigned bits = __float_as_uint(wk);
                    if (wk > 0.f && bits >= lo && bits <= top) {
                        atomicAdd(&h_sum[(bits - lo) >> shift], significand(bits));
                        atomicAdd(&h_cnt[(bits - lo) >> shift], 1u);
                    }
                };
                sweep(w, n,
                      [&](long long, float4 x) {
                          add(x.x);
                          add(x.y);
                          add(x.z);
                          add(x.w);
                      },
                      [&](long long, float wk) { add(wk); });
                __syncthreads();
                pick = select_level(level, h_sum[tid], h_cnt[tid], eff_e, a.fraction, bound,
                                    total, npos, st, ss, fallback);
                st.below = ss.excl;
                st.nbelow = ss.excl_cnt;
                st.crossed = ss.cond;
                st.lo += (unsigned)pick << shift;
            }
        }
        if (mk) {
            // ---- last pass: mask bytes
            const float ws = all_pass ? 0.f : __uint_as_float(st.lo);
            long long hw = (long long)(((16u - ((unsigned)(uintptr_t)w & 15u)) & 15u) >> 2);
            const bool mvec = (((uintptr_t)(mk + hw)) & 3u) == 0;
            sweep(w, n,
                  [&](long long k, float4 x) {
                      store_mask4(mk, k, mvec, x.x >= ws, x.y >= ws, x.z >= ws, x.w >= ws);
                  },
                  [&](long long k, float wk) { mk[k] = wk >= ws ? 1 : 0; });
        }
    }

    if (tid == 0) {
        const long long n_under = all_pass ? 0 : nzero + (long long)st.nbelow;
        a.sum_weight[i] = u192_value(total);
        a.significant_weight[i] = all_pass ? 0.f : __uint_as_float(st.lo);
        a.n_significant[i] = (int)(n - n_under);
        if (a.max_log_weight) a.max_log_weight[i] = any ? lmax : -INFINITY;
    }
}

hipError_t launch_weights(const WeightsArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(weights_kernel, dim3(a.n_particles), dim3(kWT), 0, s, a);
    return hipGetLastError();
}

}  // namespace rc
