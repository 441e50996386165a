// Device helpers shared by the gfx950 kernels: projection sampling, phase
// factors, operand encodings.
#pragma once

#include <hip/hip_runtime.h>

#include "rc_internal.hpp"

namespace rc {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// Two adjacent float2 voxels (8-byte aligned, as the volume's float2 are).
typedef float vox_pair __attribute__((ext_vector_type(4), aligned(8)));

// Central-slice sample of the padded volume for image pixel (x, y) and the
// orientation rows (e0, e1), (e3, e4), (e6, e7): the arithmetic of
// AccProjectorKernel::project3Dmodel (acc_projectorkernel_impl.h:178-252)
// with the texture fetch replaced by an explicit trilinear blend under
// clamp-to-edge addressing (acc_projector_impl.h:98-103) in the blend order
// of no_tex3D (cuda_device_utils.cuh:74-108).  Float32 throughout.
__device__ __forceinline__ float2 project_sample(const ProjView &v, float e0, float e1, float e3,
                                                 float e4, float e6, float e7, float x, float y)
{
    float xp = (e0 * x + e1 * y) * v.pad;
    float yp = (e3 * x + e4 * y) * v.pad;
    float zp = (e6 * x + e7 * y) * v.pad;
    int r2 = (int)(xp * xp + yp * yp + zp * zp);
    if (r2 > v.maxR2_padded) return make_float2(0.f, 0.f);
    const bool neg = xp < 0.f;
    if (neg) {
        xp = -xp;
        yp = -yp;
        zp = -zp;
    }
    yp -= (float)v.initY;
    zp -= (float)v.initZ;
    const float fx0 = floorf(xp), fy0 = floorf(yp), fz0 = floorf(zp);
    const float fx = xp - fx0, fy = yp - fy0, fz = zp - fz0;
    const int x0 = (int)fx0, y0 = (int)fy0, z0 = (int)fz0;
    // The two x neighbours are adjacent voxels: one 16-byte load per (y, z)
    // corner pair, at xa (neighbours xa, xa + 1) or, where x is clamped at
    // the last voxel (xa = xb = mdlX - 1), at mdlX - 2 with its second voxel
    // used twice.  (mdlX >= 2 for every volume the library accepts.)
    const int xa = clampi(x0, 0, v.mdlX - 1);
    const int xl = min(xa, v.mdlX - 2);
    const bool edge = xl != xa;
    const int ya = clampi(y0, 0, v.mdlY - 1) * v.mdlX, yb = clampi(y0 + 1, 0, v.mdlY - 1) * v.mdlX;
    const size_t za = (size_t)clampi(z0, 0, v.mdlZ - 1) * v.mdlXY + xl;
    const size_t zb = (size_t)clampi(z0 + 1, 0, v.mdlZ - 1) * v.mdlXY + xl;
    auto pair = [&](size_t i, float2 &lo, float2 &hi) {
        const vox_pair w = *reinterpret_cast<const vox_pair *>(v.vol + i);
        hi = make_float2(w.z, w.w);
        lo = edge ? hi : make_float2(w.x, w.y);
    };
    float2 d000, d001, d010, d011, d100, d101, d110, d111;
    pair(za + ya, d000, d001);
    pair(za + yb, d010, d011);
    pair(zb + ya, d100, d101);
    pair(zb + yb, d110, d111);
    float2 r;
    {
        float dx00 = d000.x + (d001.x - d000.x) * fx;
        float dx01 = d100.x + (d101.x - d100.x) * fx;
        float dx10 = d010.x + (d011.x - d010.x) * fx;
        float dx11 = d110.x + (d111.x - d110.x) * fx;
        float dxy0 = dx00 + (dx10 - dx00) * fy;
        float dxy1 = dx01 + (dx11 - dx01) * fy;
        r.x = dxy0 + (dxy1 - dxy0) * fz;
    }
    {
        float dx00 = d000.y + (d001.y - d000.y) * fx;
        float dx01 = d100.y + (d101.y - d100.y) * fx;
        float dx10 = d010.y + (d011.y - d010.y) * fx;
        float dx11 = d110.y + (d111.y - d110.y) * fx;
        float dxy0 = dx00 + (dx10 - dx00) * fy;
        float dxy1 = dx01 + (dx11 - dx01) * fy;
        r.y = dxy0 + (dxy1 - dxy0) * fz;
    }
    if (neg) r.y = -r.y;
    return r;
}

// sin and cos of 2*pi*r for r in [-0.5, 0.5] with the CDNA transcendental
// unit: v_sin_f32 / v_cos_f32 take their argument in revolutions.  On
// gfx950 their error over [-0.5, 0.5] measures 1.24e-7 absolute (max) and
// 2.8e-8 (mean) against float64 over 2^20 uniformly spaced arguments.
__device__ __forceinline__ void sincos_2pi(float r, float &s, float &c)
{
    s = __builtin_amdgcn_sinf(r);
    c = __builtin_amdgcn_cosf(r);
}

// The same by polynomial: quadrant reduction u = r - q/4 (exact),
// theta = 2*pi*u in [-pi/4, pi/4], Taylor series of sin to theta^9 and of
// cos to theta^10 (truncation below 2e-9 there), fix-up sin(theta + q*pi/2).
__device__ __forceinline__ void sincos_2pi_poly(float r, float &s, float &c)
{
    const float q = rintf(r * 4.0f);
    const float u = fmaf(q, -0.25f, r);
    const float th = u * 6.28318530717958647692f;
    const float t2 = th * th;
    float ps = fmaf(t2, 2.75573192239858907e-6f, -1.98412698412698413e-4f);
    ps = fmaf(t2, ps, 8.33333333333333333e-3f);
    ps = fmaf(t2, ps, -1.66666666666666667e-1f);
    ps = th * fmaf(t2, ps, 1.0f);
    float pc = fmaf(t2, -2.75573192239858907e-7f, 2.48015873015873016e-5f);
    pc = fmaf(t2, pc, -1.38888888888888889e-3f);
    pc = fmaf(t2, pc, 4.16666666666666667e-2f);
    pc = fmaf(t2, pc, -0.5f);
    pc = fmaf(t2, pc, 1.0f);
    const int qi = ((int)q) & 3;
    const float sa = (qi & 1) ? pc : ps;
    const float ca = (qi & 1) ? ps : pc;
    s = (qi & 2) ? -sa : sa;
    c = ((qi + 1) & 2) ? -ca : ca;
}

// Phase factor e^{i phi}, phi = x*tx + y*ty, with tx, ty given in
// revolutions per pixel (float64) so the argument is reduced exactly.
__device__ __forceinline__ void phase_factor(float x, float y, double txr, double tyr, float &s,
                                             float &c)
{
    const double ph = fma((double)x, txr, (double)y * tyr);
    const float r = (float)(ph - rint(ph));
    sincos_2pi(r, s, c);
}

// ---------------------------------------------------------------------------
// operand encodings
// ---------------------------------------------------------------------------
__device__ __forceinline__ void split_bf16(float v, __bf16 &hi, __bf16 &lo)
{
    hi = (__bf16)v;
    lo = (__bf16)(v - (float)hi);
}

// Write one complex value (pixel j of the group at `g`) in the encoding of FORM.
template <int FORM>
__device__ __forceinline__ void store_encoded(unsigned char *g, int j, float re, float im,
                                              float scale)
{
    if constexpr (FORM == kFP32) {
        reinterpret_cast<float *>(g)[j] = re;
        reinterpret_cast<float *>(g)[8 + j] = im;
    } else if constexpr (FORM == kBF16X3) {
        __bf16 rh, rl, ih, il;
        split_bf16(re, rh, rl);
        split_bf16(im, ih, il);
        __bf16 *b = reinterpret_cast<__bf16 *>(g);
        b[j] = rh;
        b[8 + j] = rl;
        b[16 + j] = ih;
        b[24 + j] = il;
    } else {
        _Float16 *h = reinterpret_cast<_Float16 *>(g);
        h[j] = (_Float16)(re * scale);
        h[8 + j] = (_Float16)(im * scale);
    }
}

// Power-of-two scale putting max-abs m at or below 1 (1 if m == 0).
__host__ __device__ inline float pow2_scale(float m)
{
    if (!(m > 0.f)) return 1.f;
    int e;
    frexpf(m, &e);     // m = f * 2^e, f in [0.5, 1)
    return ldexpf(1.f, -e);
}

}  // namespace rc
