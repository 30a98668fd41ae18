// Host/device inline math of the solver.
//
// Floating-point contract: the library is compiled with -ffp-contract=off, so
// only the fmaf() calls written here fuse; division and sqrtf are IEEE-correct
// (hipcc's default).  With that, every integer decision of a ray (cell
// sequence, step count, end type) and every (band, g) draw is reproducible
// bit for bit by a CPU restatement that follows the same operations.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DN_HD __host__ __device__ __forceinline__

namespace dn {

// 2^-24: descriptor direction fields hold 24-bit integers m, u = (m + off) 2^-24.
constexpr float kTwoM24 = 5.9604644775390625e-08f;

// sin and cos of 2*pi*u for u in [0, 1]: quadrant reduction on u (exact in
// FP32), reflection to [0, pi/4], Taylor polynomials to x^9 / x^8 evaluated
// with fmaf (truncation < 3e-8 on [0, pi/4]).
DN_HD void sincos2pi(float u, float* s, float* c)
{
    const float S1 = -1.66666672e-01f, S2 = 8.33333377e-03f, S3 = -1.98412701e-04f, S4 = 2.75573188e-06f;
    const float C1 = -5.00000000e-01f, C2 = 4.16666679e-02f, C3 = -1.38888892e-03f, C4 = 2.48015876e-05f;
    float q4 = u * 4.0f;
    float qf = floorf(q4);
    int q = (int)qf;
    float r = q4 - qf;
    bool swap = r > 0.5f;
    float rr = swap ? (1.0f - r) : r;
    float x = rr * 1.57079637f;
    float x2 = x * x;
    float ps = fmaf(x2, S4, S3);
    ps = fmaf(x2, ps, S2);
    ps = fmaf(x2, ps, S1);
    float sn = fmaf(x * x2, ps, x);
    float pc = fmaf(x2, C4, C3);
    pc = fmaf(x2, pc, C2);
    pc = fmaf(x2, pc, C1);
    float cs = fmaf(x2, pc, 1.0f);
    float a = swap ? cs : sn;   // sin of the reduced angle
    float b = swap ? sn : cs;   // cos of the reduced angle
    int qq = q & 3;
    *s = (qq == 0) ? a : (qq == 1) ? b : (qq == 2) ? -a : -b;
    *c = (qq == 0) ? b : (qq == 1) ? -a : (qq == 2) ? -b : a;
}

// Inverse-CDF bin search with the reference's semantics (code/mc_gpu.cu:840-853,
// :881-893): the bin is the number of CDF entries <= R, R < cdf[0] gives 0,
// R == 1 gives n-1; results past the table are clamped to n-1.
DN_HD int cdf_search(const float* cdf, int n, float R)
{
    int lo = 0, w = n;
    while (w > lo + 1) {
        int d = (lo + w) >> 1;
        float v = cdf[d];
        if (R < v) w = d;
        if (R >= v) lo = d;
    }
    lo += 1;
    if (R < cdf[0]) lo = 0;
    if (R == 1.0f) lo = n - 1;
    return lo > n - 1 ? n - 1 : lo;
}

// Philox4x32-10 (Salmon et al., SC'11).
struct u32x4 { uint32_t x, y, z, w; };
DN_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// (m + 1/2) 2^-23 with m the top 23 bits: exact in FP32, strictly inside (0, 1).
DN_HD float u01_open(uint32_t x) { return ((float)(x >> 9) + 0.5f) * 1.1920928955078125e-07f; }
DN_HD float u01_incl(uint32_t x) { return ((float)(x >> 8) + 1.0f) * kTwoM24; }

// XORWOW recurrence (Marsaglia 2003, "Xorshift RNGs", p.5).
struct Xorwow { uint32_t x0, x1, x2, x3, x4, d; };
DN_HD uint32_t xorwow_next(Xorwow& s)
{
    uint32_t t = s.x0 ^ (s.x0 >> 2);
    s.x0 = s.x1; s.x1 = s.x2; s.x2 = s.x3; s.x3 = s.x4;
    s.x4 = (s.x4 ^ (s.x4 << 4)) ^ (t ^ (t << 1));
    s.d += 362437u;
    return s.x4 + s.d;
}

// Unsigned 32-bit division by a runtime constant d >= 1 with a multiply and
// shifts (Granlund & Montgomery 1994, "round-up" method): l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1, q = (t + ((n - t) >> 1)) >> (l - 1),
// t = mulhi(m, n); exact for every n < 2^32.
struct FastDiv {
    uint32_t m, s, d;
};
inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f{0u, 0u, d};
    if (d <= 1) return f;
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    f.s = l;
    f.m = (uint32_t)((((1ull << l) - d) << 32) / d + 1ull);
    return f;
}
DN_HD uint32_t fdiv(uint32_t n, const FastDiv& f)
{
    if (f.d <= 1) return n;
    uint32_t t = (uint32_t)(((uint64_t)f.m * n) >> 32);
    return (t + ((n - t) >> 1)) >> (f.s - 1);
}

// Ray descriptor packing (16 bytes, four SoA words):
//   w0 = m_theta (24 bit) | band bits 0..7  << 24
//   w1 = m_phi   (24 bit) | band bits 8..9  << 24 | g (6 bit) << 26
//   w, winv: bundle weight and weight / Ib(T_emitter)
DN_HD void desc_pack(uint32_t mt, uint32_t mp, int band, int g, uint32_t* w0, uint32_t* w1)
{
    *w0 = (mt & 0xFFFFFFu) | ((uint32_t)(band & 0xFF) << 24);
    *w1 = (mp & 0xFFFFFFu) | ((uint32_t)((band >> 8) & 0x3) << 24) | ((uint32_t)(g & 0x3F) << 26);
}
DN_HD int desc_band(uint32_t w0, uint32_t w1) { return (int)((w0 >> 24) | (((w1 >> 24) & 0x3u) << 8)); }
DN_HD int desc_g(uint32_t w1) { return (int)(w1 >> 26); }

}  // namespace dn
