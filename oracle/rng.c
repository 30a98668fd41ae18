/*
 * Oracle RNGs (test infrastructure only; see oracle.h).
 *
 * Reference call sites: code/mc_gpu.cu:68,79-82,99 (curand_init per (cell,
 * block)), :593-599 (Sobol32 direction vectors), :815-833,878 (draws).
 *
 * Pins:
 *  - Philox4x32-10: known-answer vectors of the published Random123
 *    distribution (tests/test_oracle_rng.py).
 *  - XORWOW recurrence: Marsaglia (2003) p.5, checked by brute force; the
 *    GF(2) jump is checked against brute-force stepping for small n and for
 *    composition (jump(a)+jump(b) == jump(a+b)).
 *  - Sobol32: (0,m,2)-net property of the first 2^m points (every elementary
 *    interval holds exactly one point), which fails for any wrong direction
 *    number.
 *  - The seed scramble and the float conversion are this project's own
 *    definitions: parity with cuRAND's streams is UNPINNED (no CUDA here and
 *    cuRAND internals may not be reproduced).
 */
#include "oracle.h"
#include <string.h>
#include <stdlib.h>
#include <math.h>

/* ------------------------------------------------------------- Philox ---- */
static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(M0, c0, &hi0, &lo0);
        mulhilo32(M1, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += W0; k1 += W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

float orc_u01_incl(uint32_t x) { return ((float)(x >> 8) + 1.0f) * 5.9604644775390625e-08f; }
/* (m + 1/2) 2^-23, m = top 23 bits: exact in FP32 and strictly inside (0, 1). */
float orc_u01_open(uint32_t x) { return ((float)(x >> 9) + 0.5f) * 1.1920928955078125e-07f; }

/* ------------------------------------------------------------- XORWOW ---- */
uint32_t orc_xorwow_next(orc_xorwow *s)
{
    uint32_t t = s->x[0] ^ (s->x[0] >> 2);
    s->x[0] = s->x[1];
    s->x[1] = s->x[2];
    s->x[2] = s->x[3];
    s->x[3] = s->x[4];
    s->x[4] = (s->x[4] ^ (s->x[4] << 4)) ^ (t ^ (t << 1));
    s->d += 362437u;
    return s->x[4] + s->d;
}

static uint64_t splitmix64(uint64_t *st)
{
    uint64_t z = (*st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void orc_xorwow_seed(orc_xorwow *s, uint64_t seed)
{
    uint64_t st = seed;
    uint64_t a = splitmix64(&st), b = splitmix64(&st), c = splitmix64(&st);
    s->x[0] = (uint32_t)a; s->x[1] = (uint32_t)(a >> 32);
    s->x[2] = (uint32_t)b; s->x[3] = (uint32_t)(b >> 32);
    s->x[4] = (uint32_t)c; s->d = (uint32_t)(c >> 32);
    if ((s->x[0] | s->x[1] | s->x[2] | s->x[3] | s->x[4]) == 0) s->x[0] = 1u;
}

void orc_xorwow_skip_brute(orc_xorwow *s, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i) (void)orc_xorwow_next(s);
}

/* 160x160 GF(2) matrices stored as images of the basis vectors:
 * col[b][w] = word w of M*e_b, bit index b = 32*word + bit. */
typedef struct { uint32_t col[160][5]; } gf2m;

static void gf2_apply(const gf2m *A, const uint32_t in[5], uint32_t out[5])
{
    uint32_t r[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < 160; ++b)
        if ((in[b >> 5] >> (b & 31)) & 1u)
            for (int w = 0; w < 5; ++w) r[w] ^= A->col[b][w];
    memcpy(out, r, sizeof r);
}

static void gf2_mul(const gf2m *A, const gf2m *B, gf2m *C) /* C = A*B */
{
    gf2m T;
    for (int b = 0; b < 160; ++b) gf2_apply(A, B->col[b], T.col[b]);
    *C = T;
}

static void gf2_step_matrix(gf2m *M)
{
    for (int b = 0; b < 160; ++b) {
        orc_xorwow s;
        memset(&s, 0, sizeof s);
        s.x[b >> 5] = 1u << (b & 31);
        (void)orc_xorwow_next(&s);
        memcpy(M->col[b], s.x, sizeof s.x);
    }
}

/* pow2[k] = M^(2^k) for k in [0, 128) */
static gf2m *g_pow2 = NULL;
static void gf2_tables(void)
{
    if (g_pow2) return;
    gf2m *p = (gf2m *)malloc(sizeof(gf2m) * 128);
    gf2_step_matrix(&p[0]);
    for (int k = 1; k < 128; ++k) gf2_mul(&p[k - 1], &p[k - 1], &p[k]);
    g_pow2 = p;
}

static void xorwow_jump_pow2(orc_xorwow *s, int k) /* advance 2^k steps */
{
    uint32_t out[5];
    gf2_apply(&g_pow2[k], s->x, out);
    memcpy(s->x, out, sizeof out);
    if (k < 32) s->d += 362437u * (uint32_t)(1u << k);   /* 362437*2^k mod 2^32 */
}

void orc_xorwow_init(orc_xorwow *s, uint64_t seed, uint64_t subsequence, uint64_t offset)
{
    gf2_tables();
    orc_xorwow_seed(s, seed);
    for (int j = 0; j < 60; ++j)
        if ((subsequence >> j) & 1ull) xorwow_jump_pow2(s, 67 + j);
    for (int j = 0; j < 64; ++j)
        if ((offset >> j) & 1ull) xorwow_jump_pow2(s, j);
}

/* --------------------------------------------------------------- Sobol ---- */
/* Joe & Kuo (2008) "Constructing Sobol sequences with better two-dimensional
 * projections": dimension 0 has all m_k = 1; dimension 1 is s=1, a=0, m_1=1,
 * whose recurrence m_k = 2 m_{k-1} XOR m_{k-1} follows from the primitive
 * polynomial x+1. V_k = m_k << (32-k). */
void orc_sobol_dirs(int dim, uint32_t v[32])
{
    if (dim == 0) {
        for (int k = 1; k <= 32; ++k) v[k - 1] = 1u << (32 - k);
        return;
    }
    uint32_t m = 1u;
    for (int k = 1; k <= 32; ++k) {
        if (k > 1) m = (m << 1) ^ m;
        v[k - 1] = m << (32 - k);
    }
}

uint32_t orc_sobol_point(const uint32_t v[32], uint64_t n)
{
    uint64_t g = n ^ (n >> 1);          /* Gray code of n */
    uint32_t x = 0;
    for (int k = 0; k < 32 && g; ++k, g >>= 1)
        if (g & 1ull) x ^= v[k];
    return x;
}

/* ---------------------------------------------------------- own sincos ---- */
void orc_sincos2pi(float u, float *s, float *c)
{
    const float S1 = -1.66666672e-01f, S2 = 8.33333377e-03f, S3 = -1.98412701e-04f, S4 = 2.75573188e-06f;
    const float C1 = -5.00000000e-01f, C2 = 4.16666679e-02f, C3 = -1.38888892e-03f, C4 = 2.48015876e-05f;
    float q4 = u * 4.0f;
    float qf = floorf(q4);
    int q = (int)qf;
    float r = q4 - qf;
    int swap = r > 0.5f;
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
    if (swap) { float t = sn; sn = cs; cs = t; }
    switch (q & 3) {
    case 0: *s = sn;  *c = cs;  break;
    case 1: *s = cs;  *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
    }
}

/* Reference binary search (code/mc_gpu.cu:840-853 for the band,
 * :881-893 for g): returns the number of cdf entries <= R, with R < cdf[0] -> 0
 * and R == 1 -> n-1. */
int orc_cdf_search(const float *cdf, int n, float R)
{
    int lo = 0, w = n;
    while (w > lo + 1) {
        int d = (lo + w) / 2;
        if (R < cdf[d]) w = d;
        if (R >= cdf[d]) lo = d;
    }
    lo = lo + 1;
    if (R < cdf[0]) lo = 0;
    if (R == 1.0f) lo = n - 1;
    /* The reference indexes past the table when cdf[n-1] < R < 1 (rounding of
     * the last CDF entry); both this oracle and the kernels clamp instead. */
    if (lo > n - 1) lo = n - 1;
    return lo;
}
