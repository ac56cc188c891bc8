// lp_devmath.hpp -- fp32 math core of the photon kernel (gfx950).
//
// Every function is a fixed sequence of IEEE-754 single operations
// (built with -ffp-contract=off and correctly rounded divide / sqrt), so a
// photon's trajectory is a deterministic function of its random numbers.
//
// The reference kernel (src/Simulation/MCLightPropagation.cl) leaves these
// to OpenCL built-ins whose results the language does not fix; this file
// fixes them:
//   dot(a,b)     = (a.x*b.x + a.y*b.y) + a.z*b.z
//   normalize(p) = p * (1/sqrt(dot(p,p))), p itself when dot(p,p) == 0
//   sign(x)      = OpenCL sign()
//   log, cos, acos from the published series named below.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lp {

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 mul(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ f3 normalize(f3 p)
{
    float d = dot(p, p);
    if (d == 0.0f) return p;
    float inv = 1.0f / __builtin_sqrtf(d);
    return mul(inv, p);
}
// Correctly rounded square root for x = 0 or x >= 2^-96 (and +inf, NaN):
// the hardware square root (within 1 ulp) corrected by the signs of the
// residuals x - s' s of its two neighbours s' -- the sequence the compiler
// emits for a correctly rounded sqrt, less its scaling of inputs below
// 2^-96 and its pass-through of 0 and inf, which the residual tests already
// leave unchanged (for 0 the lower neighbour is NaN and the upper one's
// residual is -0).  Used where the argument is 1 - c^2 of a float c in
// [-1, 1] (0 or at least 2^-25) or the squared norm of a vector of norm
// near 1.
__device__ __forceinline__ float sqrt_n(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __int_as_float(__float_as_int(s) - 1);
    const float sp = __int_as_float(__float_as_int(s) + 1);
    const float rm = __builtin_fmaf(-sm, s, x);
    const float rp = __builtin_fmaf(-sp, s, x);
    float r = rm <= 0.0f ? sm : s;
    r = rp > 0.0f ? sp : r;
    return r;
}
// Correctly rounded 1 / y for y in [2^-60, 2^60]: the compiler's sequence
// for a correctly rounded division (reciprocal seed, one Newton step on the
// reciprocal, two corrections of the quotient) without the range scaling and
// the special-case fixup around it, which leave such y and the numerator 1
// unchanged -- the same bits in fewer instructions.
__device__ __forceinline__ float recip_n(float y)
{
    float r = __builtin_amdgcn_rcpf(y);
    const float e = __builtin_fmaf(-y, r, 1.0f);
    r = __builtin_fmaf(e, r, r);
    float q = r;  // 1 * r
    float rem = __builtin_fmaf(-y, q, 1.0f);
    q = __builtin_fmaf(rem, r, q);
    rem = __builtin_fmaf(-y, q, 1.0f);
    return __builtin_fmaf(rem, r, q);
}
// normalize() for a vector whose squared norm is 0 or in [2^-96, 2^96]
__device__ __forceinline__ f3 normalize_n(f3 p)
{
    float d = dot(p, p);
    if (d == 0.0f) return p;
    float inv = recip_n(sqrt_n(d));
    return mul(inv, p);
}
// 1 - c^2 of two unit vectors is negative only by rounding; the reference
// takes the square root anyway (MCLightPropagation.cl:520, 541, 880) and the
// photon turns NaN.  Clamped at 0 here; finite trajectories are unchanged.
__device__ __forceinline__ float clamp0(float q) { return q < 0.0f ? 0.0f : q; }

__device__ __forceinline__ float sign_cl(float x)
{
    if (x > 0.0f) return 1.0f;
    if (x < 0.0f) return -1.0f;
    if (x == 0.0f) return x;
    return 0.0f;
}

// Natural logarithm: Cody & Waite reduction x = 2^e m, m in [sqrt2/2, sqrt2];
// log m = 2 atanh(s), s = (m-1)/(m+1), Taylor series of atanh through s^9
// (Abramowitz & Stegun 4.1.27).  ln2 split so that e*LN2_HI is exact.
__device__ __forceinline__ float logf_lp(float x)
{
    if (x != x) return x;
    if (x < 0.0f) return __int_as_float(0x7fc00000);
    if (x == 0.0f) return -__builtin_inff();
    if (x == __builtin_inff()) return x;
    int e = 0;
    uint32_t ix = __float_as_uint(x);
    if (ix < 0x00800000u) {
        x = x * 8388608.0f;
        ix = __float_as_uint(x);
        e = -23;
    }
    e += (int)(ix >> 23) - 127;
    float m = __uint_as_float((ix & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421354f) {
        m = m * 0.5f;
        e = e + 1;
    }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float p = z * (0.333333343f + z * (0.200000003f + z * (0.142857149f + z * 0.111111112f)));
    float logm = 2.0f * s + (2.0f * s) * p;
    float fe = (float)e;
    return fe * 0.693115234375f + (fe * 3.19461833e-05f + logm);
}

// Natural logarithm of x in [2^-126, 1] (a uniform draw; 0 is redrawn by the
// caller): logf_lp without its special cases, which such x never meets.
__device__ __forceinline__ float logf_u01(float x)
{
    int e = 0;
    const uint32_t ix = __float_as_uint(x);
    e += (int)(ix >> 23) - 127;
    float m = __uint_as_float((ix & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421354f) {
        m = m * 0.5f;
        e = e + 1;
    }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float p = z * (0.333333343f + z * (0.200000003f + z * (0.142857149f + z * 0.111111112f)));
    float logm = 2.0f * s + (2.0f * s) * p;
    float fe = (float)e;
    return fe * 0.693115234375f + (fe * 3.19461833e-05f + logm);
}

// Cosine, |x| < 2^10: Cody & Waite reduction by a three-part pi/2, Taylor
// series of cos (through r^10) and sin (through r^9) on |r| <= pi/4.
__device__ __forceinline__ float cos_poly(float r)
{
    float z = r * r;
    return 1.0f + z * (-0.5f + z * (0.0416666679f + z * (-0.00138888892f +
                 z * (2.48015876e-05f + z * (-2.75573188e-07f)))));
}
__device__ __forceinline__ float sin_poly(float r)
{
    float z = r * r;
    return r + (r * z) * (-0.166666672f + z * (0.00833333377f + z * (-0.000198412701f +
                 z * 2.75573188e-06f)));
}
__device__ __forceinline__ float cosf_lp(float x)
{
    if (x != x) return x;
    float ax = __builtin_fabsf(x);
    if (!(ax < 1024.0f)) return __int_as_float(0x7fc00000);
    int k = (int)(ax * 0.636619747f + 0.5f);
    float fk = (float)k;
    float r = ((ax - fk * 1.5703125f) - fk * 0.000483751297f) - fk * 7.54978995e-08f;
    // quadrant k & 3: cos, -sin, -cos, sin -- both series evaluated and the
    // one wanted selected (lanes in different quadrants share one pass)
    const float c = cos_poly(r), sn = sin_poly(r);
    const float v = (k & 1) ? sn : c;
    return ((k + 1) & 2) ? -v : v;
}

// Arc cosine: Maclaurin series of asin on |x| <= 1/2 (A&S 4.4.40, n <= 10)
// and the half-angle identities outside.
__device__ __forceinline__ float asin_poly(float x)
{
    float z = x * x;
    float p = 0.166666672f + z * (0.075000003f + z * (0.0446428582f + z * (0.0303819440f +
              z * (0.0223721601f + z * (0.0173527617f + z * (0.0139648430f + z * (0.0115518477f +
              z * (0.00976160448f + z * 0.00839033350f))))))));
    return x + (x * z) * p;
}
__device__ __forceinline__ float acosf_lp(float x)
{
    if (x != x) return x;
    if (x > 1.0f || x < -1.0f) return __int_as_float(0x7fc00000);
    if (x > 0.5f) return 2.0f * asin_poly(__builtin_sqrtf((1.0f - x) * 0.5f));
    if (x < -0.5f)
        return 3.14159274f - (2.0f * asin_poly(__builtin_sqrtf((1.0f + x) * 0.5f)) - (-8.74227766e-08f));
    return 1.57079637f - (asin_poly(x) - (-4.37113883e-08f));
}

// x86 cvtt semantics of the reference host's int(): truncation toward zero;
// INT_MIN for NaN and out-of-range values.
__device__ __forceinline__ int32_t cvt_i32(double x)
{
    if (!(x > -2147483649.0 && x < 2147483648.0)) return INT32_MIN;
    return (int32_t)x;
}
__device__ __forceinline__ int32_t cvt_i32f(float x)
{
    if (!(x > -2147483648.0f && x < 2147483648.0f)) return INT32_MIN;
    return (int32_t)x;
}

// Fixed point: 2^32 (weights) and 2^24 (weight / mua) per unit, truncated.
__device__ __forceinline__ unsigned long long fx32(float w)
{
    return (unsigned long long)(long long)(w * 4294967296.0f);
}
// fx32 of a weight w >= 0: one unsigned conversion below 2^32 (w < 1, the
// common case), the general form otherwise -- in a branch the wavefront
// takes only when one of its lanes needs it; the same value as fx32.
__device__ __forceinline__ unsigned long long fx32_nn(float w)
{
    const float x = w * 4294967296.0f;
    const bool small = x >= 0.0f && x < 4294967296.0f;
    unsigned long long q = (unsigned long long)(uint32_t)(small ? x : 0.0f);
    if (!__all(small)) {
        asm volatile("");  // keeps the branch: the compiler would otherwise evaluate both forms
        if (!small) q = (unsigned long long)(long long)x;
    }
    return q;
}
__device__ __forceinline__ unsigned long long fx24(float w)
{
    return (unsigned long long)(long long)(w * 16777216.0f);
}

// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11).  Each round's two
// 32 x 32 -> 64-bit products are written as one 64-bit product apiece, which
// gfx950 computes in one v_mad_u64_u32 instead of a v_mul_hi_u32 and a
// v_mul_lo_u32.
__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

}  // namespace lp
