// bp_detmath.h -- deterministic single-precision transcendentals.
//
// The world step needs cos/sin (rigid body rotation, force fields), exp (velocity drag
// for gamma > 0) and pow (field falloff). libm on the host and the device math library
// on gfx950 round these differently in the last place, which would make CPU and GPU
// runs of a many-body step drift apart. Both sides therefore evaluate the same
// published algorithms with plain IEEE-754 float operations:
//
//   sin/cos : S. L. Moshier, Cephes Math Library, sinf.c/cosf.c (3-part Cody-Waite
//             reduction by pi/4, degree-7 sine and degree-8 cosine polynomials).
//   exp     : Cephes expf.c (reduction by ln 2, degree-6 polynomial, exact scaling).
//   log     : Cephes logf.c (frexp reduction to [sqrt(1/2), sqrt(2)), degree-9 polynomial).
//
// Every operation is written out; the translation units that include this header are
// compiled with floating-point contraction disabled, so each line rounds exactly once
// on both host and device. Accuracy is a few ulp against libm for |x| < 8192 (checked
// by the test suite); the argument ranges of the step are far inside that.
#ifndef BP_DETMATH_H
#define BP_DETMATH_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define BP_HD __host__ __device__ inline
#else
#define BP_HD inline
#endif

namespace bpm {

BP_HD uint32_t float_bits(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
BP_HD float bits_float(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

// Exact power of two 2^n for n in [-126, 127].
BP_HD float pow2i(int n) { return bits_float((uint32_t)(n + 127) << 23); }

// x * 2^n, exact while the result is normal; splits large |n| into two factors.
BP_HD float scale2(float x, int n) {
    if (n > 127) { x = x * pow2i(127); n -= 127; }
    if (n < -126) { x = x * pow2i(-126); n += 126; }
    return x * pow2i(n < -126 ? -126 : (n > 127 ? 127 : n));
}

// Reduction of |x| by pi/4; returns the octant j and the reduced argument.
BP_HD float reduce_quarter_pi(float ax, int &octant) {
    const float four_over_pi = 1.27323954473516f;
    const float dp1 = 0.78515625f, dp2 = 2.4187564849853515625e-4f, dp3 = 3.77489497744594108e-8f;
    int j = (int)(four_over_pi * ax);
    float y = (float)j;
    if (j & 1) { j += 1; y = y + 1.0f; }
    octant = j & 7;
    float r = ax - y * dp1;
    r = r - y * dp2;
    r = r - y * dp3;
    return r;
}

BP_HD float sin_poly(float r) {
    float z = r * r;
    float p = -1.9515295891e-4f * z;
    p = p + 8.3321608736e-3f;
    p = p * z;
    p = p + -1.6666654611e-1f;
    p = p * z;
    p = p * r;
    return p + r;
}

BP_HD float cos_poly(float r) {
    float z = r * r;
    float p = 2.443315711809948e-5f * z;
    p = p + -1.388731625493765e-3f;
    p = p * z;
    p = p + 4.166664568298827e-2f;
    p = p * z;
    p = p * z;
    float half = 0.5f * z;
    p = p - half;
    return p + 1.0f;
}

BP_HD float sinf_det(float x) {
    if (!(x == x) || x - x != 0.0f) return x - x;   // NaN for NaN and +-inf
    float sign = 1.0f, ax = x;
    if (x < 0.0f) { sign = -1.0f; ax = -x; }
    int j;
    float r = reduce_quarter_pi(ax, j);
    if (j > 3) { sign = -sign; j -= 4; }
    float y = (j == 1 || j == 2) ? cos_poly(r) : sin_poly(r);
    return sign < 0.0f ? -y : y;
}

BP_HD float cosf_det(float x) {
    if (!(x == x) || x - x != 0.0f) return x - x;
    float ax = x < 0.0f ? -x : x;
    int j;
    float r = reduce_quarter_pi(ax, j);
    float sign = 1.0f;
    if (j > 3) { j -= 4; sign = -sign; }
    if (j > 1) sign = -sign;
    float y = (j == 1 || j == 2) ? sin_poly(r) : cos_poly(r);
    return sign < 0.0f ? -y : y;
}

BP_HD float expf_det(float x) {
    if (!(x == x)) return x;
    if (x > 88.72283905206835f) return bits_float(0x7f800000u);
    if (x < -103.278929903431851103f) return 0.0f;
    const float log2e = 1.44269504088896341f;
    float t = log2e * x;
    t = t + 0.5f;
    float fz = (float)(int)t;
    if (fz > t) fz = fz - 1.0f;                 // floor for negative arguments
    int n = (int)fz;
    float r = x - fz * 0.693359375f;
    r = r - fz * -2.12194440e-4f;
    float z = r * r;
    float p = 1.9875691500e-4f * r;
    p = p + 1.3981999507e-3f;
    p = p * r;
    p = p + 8.3334519073e-3f;
    p = p * r;
    p = p + 4.1665795894e-2f;
    p = p * r;
    p = p + 1.6666665459e-1f;
    p = p * r;
    p = p + 5.0000001201e-1f;
    p = p * z;
    p = p + r;
    p = p + 1.0f;
    return scale2(p, n);
}

// Natural logarithm for x > 0 (NaN for x < 0, -inf for 0).
BP_HD float logf_det(float x) {
    if (!(x == x)) return x;
    if (x < 0.0f) return bits_float(0x7fc00000u);
    if (x == 0.0f) return bits_float(0xff800000u);
    if (x - x != 0.0f) return x;                // +inf
    uint32_t u = float_bits(x);
    int e = 0;
    if ((u & 0x7f800000u) == 0) {               // subnormal: normalise
        x = x * 16777216.0f;
        u = float_bits(x);
        e = -24;
    }
    e += (int)((u >> 23) & 0xffu) - 126;
    float m = bits_float((u & 0x007fffffu) | 0x3f000000u);   // m in [0.5, 1)
    if (m < 0.707106781186547524f) { e -= 1; m = m + m; m = m - 1.0f; }
    else m = m - 1.0f;
    float z = m * m;
    float p = 7.0376836292e-2f * m;
    p = p + -1.1514610310e-1f; p = p * m;
    p = p + 1.1676998740e-1f;  p = p * m;
    p = p + -1.2420140846e-1f; p = p * m;
    p = p + 1.4249322787e-1f;  p = p * m;
    p = p + -1.6668057665e-1f; p = p * m;
    p = p + 2.0000714765e-1f;  p = p * m;
    p = p + -2.4999993993e-1f; p = p * m;
    p = p + 3.3333331174e-1f;  p = p * m;
    p = p * z;
    float fe = (float)e;
    p = p + fe * -2.12194440e-4f;
    p = p - 0.5f * z;
    float r = m + p;
    return r + fe * 0.693359375f;
}

// base^exponent for base >= 0 (the field falloff profile); exp(e * log(b)).
BP_HD float powf_det(float base, float exponent) {
    if (exponent == 0.0f) return 1.0f;
    if (base == 0.0f) return exponent > 0.0f ? 0.0f : bits_float(0x7f800000u);
    return expf_det(exponent * logf_det(base));
}

}  // namespace bpm

#endif  // BP_DETMATH_H
