// Kernel argument blocks shared by kernels.hip and solver.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmath.h"
#include "mlp.h"

namespace dn {

constexpr int kLevWords = 16;   // per-level parameter words staged in LDS
constexpr int kRing = 8;        // tally ring entries per wave (power of 2)
constexpr int kLRowFloats = 2048;   // a learned proposal row in a wave's row buffer (4 chunks x 64 lanes x 8)

// Per-level geometry, staged into LDS as ints (floats for nothing).
struct LevParams {
    int n[3];
    int foff[3];      // offsets of the face arrays in the packed face buffer
    int maxs;
    int nbx, nby;        // 4x4x4 bricks along x and y
    int ooff[3];         // offsets of the brick-offset tables in the packed table buffer
    long long tstride;   // floats of this level per snapshot (bricked, padded)
};

// 4x4x4-bricked field index of 0-based cell (x, y, z): bricks in (z, y, x)
// order, 64 contiguous floats per brick.  A ray moving along any axis stays
// inside the same 256-byte brick for 3 of 4 steps.  Inside a brick: x bit 0,
// y, z, then x bit 1, so each 128-byte line holds 2 x 4 x 4 cells (x, y, z):
// of the three ways to halve a brick this one crosses the fewest line
// boundaries per unit path on the channel grids, whose cells are thinnest in
// y, then z, then x (crossings ~ sum over axes of 1 / (cells per line x cell
// size)); measured 0.75 % faster than 4 x 4 x 2 lines.
__host__ __device__ __forceinline__ int brick_index(int x, int y, int z, int nbx, int nby)
{
    return ((((z >> 2) * nby + (y >> 2)) * nbx + (x >> 2)) << 6) | (((x >> 1) & 1) << 5) | ((z & 3) << 3) |
           ((y & 3) << 1) | (x & 1);
}

enum { kSpecTable = 0, kSpecGreyT4 = 1 };

// Where the traversal reads its per-step table values from:
//   kTabGlobal  kappa rows and Ib from global memory (L2),
//   kTabIbLds   Ib table in LDS, kappa rows from global memory,
//   kTabSliced  kappa rows and Ib of the temperature nodes the fields span
//               ([tab_lo, tab_lo + tab_nw)) in LDS, 1024-thread blocks.
enum { kTabGlobal = 0, kTabIbLds = 1, kTabSliced = 2 };

struct TraceArgs {
    LevParams lev[6];
    const float* T[6];          // level fields, all snapshots
    int nlev;
    int bc_periodic;            // bit (axis*2 + side) set if periodic
    int polar;
    float tau_cut;
    float u_off;
    int nB, nQ, nT;
    float inv_dT, mT0;
    float kappa_grey;
    const float* kq;            // kappa rows [(b*nQ+g)*nT + t]
    const float* Ib;            // Planck table [b*nT + t]
    const float* Ibw;           // wall Planck [(axis*2+side)*nB + b]
    const float* faces;         // all levels packed
    int nfaces;
    const int* offs;            // per level and axis: brick-layout offset of 0-based cell index c
    int noffs;
    const float* centres;       // level 0, x|y|z packed, ghost-inclusive
    int coff[3];
    int ncent;
    int tab_lo, tab_nw;         // sliced tables: first node and node count held in LDS
    // rays (SoA)
    const uint32_t* w0;
    const uint32_t* w1;
    const float* w;
    const float* winv;
    long long nrays;
    int rays_per_slot;
    int slots_per_emitter;
    FastDiv fd_R;                         // division by rays per slot
    int refill_k;                         // start rays once this many lanes of a wave are idle
    int super;                            // rays per wave super-chunk (multiple of 64)
    const int4* slotinfo;       // processing slot -> {tally slot, cell, bricked index, i|j<<10|k<<20}
    float2* pray;               // per-ray (E, A) for the deterministic tally modes
    unsigned* claim;            // 8 super-chunk counters (one per XCD, 32 words apart), zero at launch
    long long slot0;            // first slot of this batch
    int snap;
    // tallies
    int tally;                  // DNSMC_TALLY_*
    double* E;
    double* A;
    double* Q2;
    unsigned long long* nsteps;  // [0] segments, [1] rays traced from the long-ray lists
    // several snapshots in one launch (fused generation): rays of snapshot
    // P.snap + s are [0, nrays) again; tally slot = slotinfo.x + s * tally_stride
    int nsnap_launch;
    int tally_stride;
    // fused Philox generation (plain estimator, code/mc_gpu.cu:93-144 semantics
    // of the plain draw): counter (r, cell, snapshot, 0), key seed
    unsigned long long seed;
    const float* cdf_band;
    const float* cdf_g;
    const float* inv_emax;
    float Ibmax;
    int g_search;
    int gen_lds;                // bit 0: cdf_g staged in LDS too; bit 1: inv_emax too
    // long rays first (GEN == 1): the rays that the launch would start last
    // (last snapshot of the launch, final part of each XCD's range: offset >=
    // cand_from within its cand_fd.d rays) whose polar direction cosine is
    // near 0 (m_theta within cand_w of 2^23, rays along the walls that may run
    // to the step cap) are listed by cand_scan; the waves trace the list
    // first and skip those rays where their super-chunks come up
    const uint32_t* cand;       // listed ray indices, one list per XCD range (nullptr: off)
    unsigned* cand_cnt;         // list x: [64 x] rays found (more than cand_cap: off), [64 x + 32] claims
    uint32_t cand_cap;          // capacity of each list
    uint32_t cand_w;
    uint32_t cand_from;
    FastDiv cand_fd;            // division by the rays of one XCD range (per_x * super)
    int desc_ratio;             // learned descriptors: w holds p/q, the weight is computed at ray start
    // FP64 traversal extras
    const double* inv_emax_d;
    double Ibmax_d;
    float Tmax;
    // per-ray records
    long long nrec;
    int* rec_steps;
    int* rec_end;
    uint32_t* rec_hash;
    float* rec_D;
    float* rec_G;
    // learned proposal drawn inside the traversal (GEN == 2): the network,
    // the plain bin probabilities and CDF, the mixture weight, the features'
    // inputs (level-0 temperatures of snapshot 0, feat_stride floats per
    // snapshot); with log_rows set, every proposal row the waves evaluate
    // (cumulative weights U, noutp floats per processing slot) and every
    // drawn ray (descriptor words, p/q in log_w) are written out
    MlpWeights mlpw;
    const float* p_bg;
    const float* pcdf;
    const float* pdir;          // the direction head's plain CDF, (d + 1) / 64
    float mix_eps;
    FeatArgs feat;
    long long feat_stride;
    float* lrow_cache;          // per wave: the row buffer (kLRowFloats floats)
    float* log_rows;
    uint32_t* log_w0;
    uint32_t* log_w1;
    float* log_w;
};

struct GenArgs {
    // geometry of level 0 (emitter temperature)
    const float* T0;
    long long tstride0;
    int nB, nQ, nT;
    float inv_dT, mT0;
    float kappa_grey;
    int spec;
    const float* kq;
    const float* Ib;
    const float* cdf_band;
    const float* cdf_g;
    const float* inv_emax;
    float Ibmax;
    int g_search;
    // batch
    const int4* slotinfo;
    long long slot0;
    long long nslots;
    int rays_per_slot;
    int slots_per_emitter;
    int snap;
    unsigned long long seed;
    // compat
    uint32_t xw_seed[6];
    const uint32_t* xw_jump_sub;   // 60 matrices (2^(67+j))
    const uint32_t* xw_jump_off;   // 64 matrices (2^j)
    const uint32_t* sobol_v;       // 2 x 32
    int use_sobol;
    int draws_per_bundle;
    int sub_len;                   // bundles per generator thread
    int n0, n1, n2;                // level-0 sizes (reference stream decomposition)
    int num_streams, num_gpu, gpu, p_row, ystart;
    const int* band_order;         // sorted mode: band at each position of the kavg order
    int draws_find;                // sorted mode: XORWOW uniforms per (band, g) draw
    // learned proposal
    MlpWeights mlpw;
    FeatArgs feat;
    const float* p_bg;             // plain probability of each (band, g) bin
    float mix_eps;                 // defensive mixture weight of the plain distribution
    // outputs
    uint32_t* w0;
    uint32_t* w1;
    float* w;
    float* winv;
};

struct SplineArgs {
    const float* in;
    float* out;
    double* work;
    int nlines;
    int nf, nc;
    long long A, sa_in, sb_in, st_in, sa_out, sb_out, st_out;
    const double* sig;
    const double* invp;
    const double* y2f;
    const double* c6;
    const double* ihl;
    const double* ihh;
    const int* klo;
    const double* ca;
    const double* cb;
    const double* cA;
    const double* cB;
    long long snap_in, snap_out;   // per-snapshot strides
    int brick_out, m0, nbx, nby;   // z pass: write the bricked layout of the coarse level
    int in_bricked, fn1, fnbx, fnby;   // x pass: read the bricked fine field (line = j + n1 k)
};

// ------------------------------------------------ shared device helpers ----
// FP32 -> BF16, round to nearest even (NaN kept quiet).
__device__ __forceinline__ uint16_t f2bf(float x)
{
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Features of 0-based cell (i, j, k): position along the wall-normal axis in
// [-1, 1], T/Tmax, the temperature gradient (central differences, periodic
// wrap or one-sided at walls) scaled by (L/2)/Tmax, mean wall T / Tmax.
__device__ __forceinline__ void cell_features(const FeatArgs& f, int i, int j, int k, float* out)
{
    const int c[3] = {i, j, k};
    const float T = f.T0[brick_index(i, j, k, f.nbx, f.nby)];
    float g[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int n = f.n[ax];
        const bool per = (f.bc_periodic >> (2 * ax)) & 1;
        int cm = c[ax] - 1, cp = c[ax] + 1;      // neighbour cells; centres index +1 (ghost-inclusive)
        int tm = cm, tp = cp;
        if (per) {
            tm = cm < 0 ? n - 1 : cm;
            tp = cp >= n ? 0 : cp;
        } else {
            cm = tm = cm < 0 ? 0 : cm;
            cp = tp = cp >= n ? n - 1 : cp;
        }
        if (n < 2) {
            g[ax] = 0.f;
            continue;
        }
        int im[3] = {i, j, k}, ip[3] = {i, j, k};
        im[ax] = tm;
        ip[ax] = tp;
        const float Tm = f.T0[brick_index(im[0], im[1], im[2], f.nbx, f.nby)];
        const float Tp = f.T0[brick_index(ip[0], ip[1], ip[2], f.nbx, f.nby)];
        const float xm = f.centres[f.coff[ax] + cm + 1], xp = f.centres[f.coff[ax] + cp + 1];
        g[ax] = (Tp - Tm) / (xp - xm) * (0.5f * f.Lpolar) * f.inv_Tmax;
    }
    const int cpol = f.polar == 0 ? i : (f.polar == 1 ? j : k);
    const int opol = f.polar == 0 ? f.coff[0] : (f.polar == 1 ? f.coff[1] : f.coff[2]);
    const float y = f.centres[opol + cpol + 1];
    out[0] = 2.0f * y / f.Lpolar - 1.0f;
    out[1] = T * f.inv_Tmax;
    out[2] = g[0];
    out[3] = g[1];
    out[4] = g[2];
    out[5] = f.Tw_rel;
}


// Reference-style bin search (code/mc_gpu.cu:842-853 semantics) on the
// mixture CDF F(k) = (1 - eps) U[k] / U[n-1] + eps P[k], evaluated per probe;
// returns the bin and its probability F(k) - F(k-1).
__device__ __forceinline__ int mix_search(const float* Uh, const float* P, int n, float eps, float R, float* qk)
{
    const float s = (1.0f - eps) / Uh[n - 1];
    auto F = [&](int k) { return k < 0 ? 0.f : (k == n - 1 ? 1.0f : fminf(fmaf(Uh[k], s, eps * P[k]), 1.0f)); };
    int lo = 0, w = n;
    while (w > lo + 1) {
        const int d = (lo + w) >> 1;
        const float fd = F(d);
        if (R < fd) w = d;
        if (R >= fd) lo = d;
    }
    int k = lo + 1;
    if (R < F(0)) k = 0;
    if (R == 1.0f) k = n - 1;
    if (k > n - 1) k = n - 1;
    *qk = F(k) - F(k - 1);
    return k;
}

// mix_search on both heads of a row at once: the two binary searches advance
// in the same loop (their probes are independent loads), each with exactly
// mix_search's probe sequence and special cases, so the bins and
// probabilities are mix_search's.
__device__ __forceinline__ void mix_search2(const float* Ud, const float* Pd, int nd, const float* Ub,
                                            const float* Pb, int nb, float eps, float Rd, float Rb, int* kd,
                                            float* qd, int* kb, float* qb)
{
    const float sd = (1.0f - eps) / Ud[nd - 1], sb = (1.0f - eps) / Ub[nb - 1];
    auto Fd = [&](int k) { return k < 0 ? 0.f : (k == nd - 1 ? 1.0f : fminf(fmaf(Ud[k], sd, eps * Pd[k]), 1.0f)); };
    auto Fb = [&](int k) { return k < 0 ? 0.f : (k == nb - 1 ? 1.0f : fminf(fmaf(Ub[k], sb, eps * Pb[k]), 1.0f)); };
    int lod = 0, wd = nd, lob = 0, wb = nb;
    while (wd > lod + 1 || wb > lob + 1) {
        const bool ad = wd > lod + 1, ab = wb > lob + 1;
        const int dd = (lod + wd) >> 1, db = (lob + wb) >> 1;
        const float fd = ad ? Fd(dd) : 0.f;
        const float fb = ab ? Fb(db) : 0.f;
        if (ad) {
            if (Rd < fd) wd = dd;
            if (Rd >= fd) lod = dd;
        }
        if (ab) {
            if (Rb < fb) wb = db;
            if (Rb >= fb) lob = db;
        }
    }
    int k = lod + 1;
    if (Rd < Fd(0)) k = 0;
    if (Rd == 1.0f) k = nd - 1;
    if (k > nd - 1) k = nd - 1;
    *qd = Fd(k) - Fd(k - 1);
    *kd = k;
    k = lob + 1;
    if (Rb < Fb(0)) k = 0;
    if (Rb == 1.0f) k = nb - 1;
    if (k > nb - 1) k = nb - 1;
    *qb = Fb(k) - Fb(k - 1);
    *kb = k;
}

}  // namespace dn
