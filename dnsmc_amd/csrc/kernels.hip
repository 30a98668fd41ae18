// GPU kernels of the Monte Carlo radiative-transfer solver (gfx950).
//
//   raygen_plain   Philox rays: (band, g) from the Tmax proposal, isotropic
//                  direction uniforms, bundle weights -> SoA descriptors.
//   raygen_compat  reference sampling: XORWOW stream per (cell, block) for
//                  (band, g), Sobol32 dims 0/1 for directions.
//   trace          persistent cell-by-cell traversal (code/mc_gpu.cu:228-438)
//                  with FP64 per-slot tallies.
//   reduce_rays    deterministic per-slot sum of per-ray tallies (FP64, Kahan
//                  FP32 or plain FP32).
//   spline_pass    one axis of the natural-spline coarse-grid interpolation
//                  (code/spline.cpp:35-116).
//   finalize       Q = E - A, standard error (code/mc_gpu.cu:212-222).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmath.h"
#include "kernels.h"
#include "../../include/dnsmc_rad.h"

namespace dn {

typedef __attribute__((address_space(1))) const float gfloat;

// Emitter weights shared by the generators: Ibmax kappa_i Ib_i / (kappa_max Ib_max)
// (code/mc_gpu.cu:63,:163-169).
__device__ __forceinline__ void bundle_weight(const GenArgs& a, const float* ib_tab, float Ti, int band, int g,
                                              float* w, float* winv)
{
    float kap, ib;
    int bg = band * a.nQ + g;
    if (a.spec == kSpecGreyT4) {
        float t2 = Ti * Ti;
        kap = a.kappa_grey;
        ib = t2 * t2;
    } else {
        float pos = fmaf(Ti, a.inv_dT, a.mT0);
        pos = fminf(fmaxf(pos, 0.0f), (float)(a.nT - 1));
        int t0 = (int)pos;
        t0 = t0 > a.nT - 2 ? a.nT - 2 : t0;
        float fr = pos - (float)t0;
        const float* kr = a.kq + (size_t)bg * a.nT;
        float k0 = kr[t0], k1 = kr[t0 + 1];
        kap = fmaf(fr, k1 - k0, k0);
        const float* ir = ib_tab + (size_t)band * a.nT;
        float i0 = ir[t0], i1 = ir[t0 + 1];
        ib = fmaf(fr, i1 - i0, i0);
    }
    float ww = a.Ibmax * (kap * ib) * a.inv_emax[bg];
    *w = ww;
    *winv = ww * (1.0f / ib);
}

// ------------------------------------------------------- raygen_plain ----
// One thread per ray.  Counter (r, cell, snapshot, 0), key = seed: the draw of
// a ray depends only on its identity, not on the launch shape.
__global__ __launch_bounds__(256) void raygen_plain(GenArgs a)
{
    extern __shared__ float lds[];
    float* s_cdfb = lds;
    float* s_cdfg = s_cdfb + a.nB;
    float* s_ib = s_cdfg + a.nB * a.nQ;
    const bool ib_lds = a.spec == kSpecTable && a.nB * a.nT <= 16384;
    for (int i = threadIdx.x; i < a.nB; i += blockDim.x) s_cdfb[i] = a.cdf_band[i];
    for (int i = threadIdx.x; i < a.nB * a.nQ; i += blockDim.x) s_cdfg[i] = a.cdf_g[i];
    if (ib_lds)
        for (int i = threadIdx.x; i < a.nB * a.nT; i += blockDim.x) s_ib[i] = a.Ib[i];
    __syncthreads();
    const float* ibt = ib_lds ? s_ib : a.Ib;
    const long long nr = a.nslots * (long long)a.rays_per_slot;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nr; q += (long long)gridDim.x * blockDim.x) {
        long long sl = q / a.rays_per_slot;
        int r = (int)(q - sl * a.rays_per_slot);
        const int4 si = a.slotinfo[a.slot0 + sl];
        const int cell = si.y;
        u32x4 x = philox4x32_10(u32x4{(uint32_t)r, (uint32_t)cell, (uint32_t)a.snap, 0u}, (uint32_t)a.seed,
                                (uint32_t)(a.seed >> 32));
        int band = cdf_search(s_cdfb, a.nB, u01_open(x.x));
        int g = a.g_search ? cdf_search(s_cdfg + band * a.nQ, a.nQ, u01_open(x.y)) : 0;
        g = g > a.nQ - 1 ? a.nQ - 1 : g;
        float Ti = a.T0[(size_t)a.snap * a.tstride0 + si.z];
        float w, winv;
        bundle_weight(a, ibt, Ti, band, g, &w, &winv);
        uint32_t w0, w1;
        desc_pack(x.z >> 8, x.w >> 8, band, g, &w0, &w1);
        a.w0[q] = w0;
        a.w1[q] = w1;
        a.w[q] = w;
        a.winv[q] = winv;
    }
}

// ------------------------------------------------------ raygen_compat ----
// 160-bit GF(2) matrix (160 columns x 5 words) applied to an XORWOW state.
__device__ __forceinline__ void gf2_apply(const uint32_t* __restrict__ M, Xorwow& s)
{
    uint32_t in[5] = {s.x0, s.x1, s.x2, s.x3, s.x4};
    uint32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0;
#pragma unroll
    for (int w = 0; w < 5; ++w) {
        uint32_t v = in[w];
        for (int bit = 0; bit < 32; ++bit) {
            if ((v >> bit) & 1u) {
                const uint32_t* c = M + (size_t)(w * 32 + bit) * 5;
                r0 ^= c[0]; r1 ^= c[1]; r2 ^= c[2]; r3 ^= c[3]; r4 ^= c[4];
            }
        }
    }
    s.x0 = r0; s.x1 = r1; s.x2 = r2; s.x3 = r3; s.x4 = r4;
}

__device__ __forceinline__ uint32_t sobol_at(const uint32_t* v, uint64_t n)
{
    uint64_t gc = n ^ (n >> 1);
    uint32_t x = 0;
    for (int k = 0; k < 32; ++k)
        if ((gc >> k) & 1ull) x ^= v[k];
    return x;
}

// XORWOW stream of a compat slot (cell, block v), seeded as
// curand_init(1234, subsequence, 0) is in code/mc_gpu.cu:99: own seed
// scramble, subsequence = the reference's product
// (idx+1)(stream+1)(v+1)(1+gpu)ystart over its stream decomposition of the
// cell (code/mc_gpu.cu:84-87), jump 2^67 per subsequence.
__device__ Xorwow compat_stream(const GenArgs& a, int cell, int v)
{
    int ci = cell % a.n0, cj = (cell / a.n0) % a.n1, ck = cell / (a.n0 * a.n1);
    int per_stream = a.n0 / a.num_streams;
    int stream = ci / per_stream;
    int jloc = cj - (a.ystart - 1);
    int kloc = ck - a.gpu * (a.n2 / a.num_gpu);
    int idx = (ci - stream * per_stream) * (a.n2 / a.num_gpu * (a.n1 / a.p_row)) + jloc * (a.n2 / a.num_gpu) + kloc;
    int32_t prod = (int32_t)((uint32_t)(idx + 1) * (uint32_t)(stream + 1) * (uint32_t)(v + 1) *
                             (uint32_t)(1 + a.gpu) * (uint32_t)a.ystart);
    uint64_t subseq = (uint64_t)(int64_t)prod;
    Xorwow s{a.xw_seed[0], a.xw_seed[1], a.xw_seed[2], a.xw_seed[3], a.xw_seed[4], a.xw_seed[5]};
    for (int j = 0; j < 60; ++j)
        if ((subseq >> j) & 1ull) gf2_apply(a.xw_jump_sub + (size_t)j * 800, s);
    return s;
}

// Advance an XORWOW state by `off` draws (the Weyl counter d by 362437 per draw).
__device__ void xorwow_skip(const GenArgs& a, Xorwow& s, uint64_t off)
{
    for (int j = 0; j < 64; ++j)
        if ((off >> j) & 1ull) {
            gf2_apply(a.xw_jump_off + (size_t)j * 800, s);
            if (j < 32) s.d += 362437u * (1u << j);
        }
}

// One thread per (slot, run of sub_len bundles).  The stream of a slot (cell c,
// block v) is XORWOW seeded as curand_init(1234, subsequence, 0) is in
// code/mc_gpu.cu:99 (own seed scramble, subsequence = the reference's product
// (idx+1)(stream+1)(v+1)(1+gpu)ystart, jump 2^67 per subsequence); the run
// starting at bundle h0 jumps h0*draws_per_bundle further.  Sobol points are
// v*photmax + h (the Sobol states restart per cell, :79-82).
__global__ __launch_bounds__(256) void raygen_compat(GenArgs a)
{
    extern __shared__ float lds[];
    float* s_cdfb = lds;
    float* s_cdfg = s_cdfb + a.nB;
    float* s_ib = s_cdfg + a.nB * a.nQ;
    uint32_t* s_sob = (uint32_t*)(s_ib + ((a.spec == kSpecTable && a.nB * a.nT <= 16384) ? a.nB * a.nT : 0));
    const bool ib_lds = a.spec == kSpecTable && a.nB * a.nT <= 16384;
    for (int i = threadIdx.x; i < a.nB; i += blockDim.x) s_cdfb[i] = a.cdf_band[i];
    for (int i = threadIdx.x; i < a.nB * a.nQ; i += blockDim.x) s_cdfg[i] = a.cdf_g[i];
    if (ib_lds)
        for (int i = threadIdx.x; i < a.nB * a.nT; i += blockDim.x) s_ib[i] = a.Ib[i];
    for (int i = threadIdx.x; i < 64; i += blockDim.x) s_sob[i] = a.sobol_v[i];
    __syncthreads();
    const float* ibt = ib_lds ? s_ib : a.Ib;
    const int runs = (a.rays_per_slot + a.sub_len - 1) / a.sub_len;
    const long long nthreads = a.nslots * runs;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nthreads; t += (long long)gridDim.x * blockDim.x) {
        long long sl = t / runs;
        int run = (int)(t - sl * runs);
        const int4 si = a.slotinfo[a.slot0 + sl];
        const int cell = si.y;
        const int v = si.x % a.slots_per_emitter;
        Xorwow s = compat_stream(a, cell, v);
        int h0 = run * a.sub_len;
        xorwow_skip(a, s, (uint64_t)h0 * (uint64_t)a.draws_per_bundle);
        int h1 = h0 + a.sub_len < a.rays_per_slot ? h0 + a.sub_len : a.rays_per_slot;
        uint64_t n = (uint64_t)v * (uint64_t)a.rays_per_slot + (uint64_t)h0;
        uint32_t sx0 = sobol_at(s_sob, n), sx1 = sobol_at(s_sob + 32, n);
        float Ti = a.T0[(size_t)a.snap * a.tstride0 + si.z];
        for (int h = h0; h < h1; ++h) {
            uint32_t mp, mt;
            if (a.use_sobol) {
                mp = sx0 >> 8;
                mt = sx1 >> 8;
                int c = __builtin_ctzll(~n);
                sx0 ^= s_sob[c];
                sx1 ^= s_sob[32 + c];
                ++n;
            } else {
                mp = xorwow_next(s) >> 8;
                mt = xorwow_next(s) >> 8;
            }
            int band = cdf_search(s_cdfb, a.nB, u01_incl(xorwow_next(s)));
            int g = 0;
            if (a.g_search) g = cdf_search(s_cdfg + band * a.nQ, a.nQ, u01_incl(xorwow_next(s)));
            g = g > a.nQ - 1 ? a.nQ - 1 : g;
            float w, winv;
            bundle_weight(a, ibt, Ti, band, g, &w, &winv);
            uint32_t w0, w1;
            desc_pack(mt, mp, band, g, &w0, &w1);
            long long q = sl * a.rays_per_slot + h;
            a.w0[q] = w0;
            a.w1[q] = w1;
            a.w[q] = w;
            a.winv[q] = winv;
        }
    }
}

// Sorted compat generation (srt == 1: code/mc_gpu.cu:100-102, :138-139).  The
// reference draws all photmax (band, g) pairs of the slot's stream first and
// counts them (kernel_find, :786-812), then replays the counts position by
// position of the kavg band order, g ascending inside a band (wave_find,
// :897-930), drawing each bundle's direction in the photon loop.  Bundles with
// the same (band, g) are interchangeable until a direction is attached, so
// the replay is a histogram and a scan: one block per slot,
//  (1) each thread jumps to its run of (band, g) draws and counts them in LDS,
//  (2) the histogram is scanned in replay order (position p, then g),
//  (3) bundle h takes the bin where the running count first exceeds h, and its
//      direction is Sobol point v*photmax + h, or XORWOW uniforms
//      draws_find*photmax + 2h (phi) and + 2h + 1 (theta) of the stream.
// The reference permutes its tables into the sorted order, so position p
// carries band band_order[p]; the descriptor holds that band.
__global__ __launch_bounds__(256) void raygen_compat_sorted(GenArgs a)
{
    extern __shared__ float lds[];
    __shared__ uint32_t s_part[256];
    const int nbin = a.nB * a.nQ;
    float* s_cdfb = lds;
    float* s_cdfg = s_cdfb + a.nB;
    uint32_t* s_sob = (uint32_t*)(s_cdfg + nbin);
    int* s_ord = (int*)(s_sob + 64);
    uint32_t* s_hist = (uint32_t*)(s_ord + a.nB);
    uint32_t* s_cum = s_hist + nbin;
    const int tid = threadIdx.x, nth = blockDim.x;
    for (int i = tid; i < a.nB; i += nth) {
        s_cdfb[i] = a.cdf_band[i];
        s_ord[i] = a.band_order[i];
    }
    for (int i = tid; i < nbin; i += nth) {
        s_cdfg[i] = a.cdf_g[i];
        s_hist[i] = 0u;
    }
    for (int i = tid; i < 64; i += nth) s_sob[i] = a.sobol_v[i];
    __syncthreads();
    const int R = a.rays_per_slot;
    const int L = (R + nth - 1) / nth;
    const int h0 = tid * L < R ? tid * L : R;
    const int h1 = h0 + L < R ? h0 + L : R;
    for (long long sl = blockIdx.x; sl < a.nslots; sl += gridDim.x) {
        const int4 si = a.slotinfo[a.slot0 + sl];
        const int cell = si.y;
        const int v = si.x % a.slots_per_emitter;
        const Xorwow s0 = compat_stream(a, cell, v);
        // (1) histogram of the (band, g) draws h0..h1-1
        if (h0 < h1) {
            Xorwow s = s0;
            xorwow_skip(a, s, (uint64_t)h0 * (uint64_t)a.draws_find);
            for (int h = h0; h < h1; ++h) {
                int band = cdf_search(s_cdfb, a.nB, u01_incl(xorwow_next(s)));
                int g = 0;
                if (a.g_search) g = cdf_search(s_cdfg + band * a.nQ, a.nQ, u01_incl(xorwow_next(s)));
                g = g > a.nQ - 1 ? a.nQ - 1 : g;
                atomicAdd(&s_hist[band * a.nQ + g], 1u);
            }
        }
        __syncthreads();
        // (2) inclusive scan in replay order: entry k = p*nQ + g counts bin order[p]*nQ + g
        const int per = (nbin + nth - 1) / nth;
        const int k0 = tid * per < nbin ? tid * per : nbin;
        const int k1 = k0 + per < nbin ? k0 + per : nbin;
        uint32_t run = 0;
        for (int k = k0; k < k1; ++k) {
            const int p = k / a.nQ, g = k - p * a.nQ;
            run += s_hist[s_ord[p] * a.nQ + g];
        }
        s_part[tid] = run;
        __syncthreads();
        for (int o = 1; o < nth; o <<= 1) {
            uint32_t add = tid >= o ? s_part[tid - o] : 0u;
            __syncthreads();
            s_part[tid] += add;
            __syncthreads();
        }
        run = tid > 0 ? s_part[tid - 1] : 0u;
        for (int k = k0; k < k1; ++k) {
            const int p = k / a.nQ, g = k - p * a.nQ;
            run += s_hist[s_ord[p] * a.nQ + g];
            s_cum[k] = run;
        }
        __syncthreads();
        // (3) bundles h0..h1-1
        if (h0 < h1) {
            int lo = 0, hi = nbin - 1;             // first k with s_cum[k] > h0 (s_cum[nbin-1] == R)
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (s_cum[mid] > (uint32_t)h0) hi = mid; else lo = mid + 1;
            }
            int k = lo;
            Xorwow s = s0;
            uint64_t n = (uint64_t)v * (uint64_t)R + (uint64_t)h0;
            uint32_t sx0 = 0, sx1 = 0;
            if (a.use_sobol) {
                sx0 = sobol_at(s_sob, n);
                sx1 = sobol_at(s_sob + 32, n);
            } else {
                xorwow_skip(a, s, (uint64_t)R * (uint64_t)a.draws_find + 2ull * (uint64_t)h0);
            }
            const float Ti = a.T0[(size_t)a.snap * a.tstride0 + si.z];
            for (int h = h0; h < h1; ++h) {
                uint32_t mp, mt;
                if (a.use_sobol) {
                    mp = sx0 >> 8;
                    mt = sx1 >> 8;
                    int c = __builtin_ctzll(~n);
                    sx0 ^= s_sob[c];
                    sx1 ^= s_sob[32 + c];
                    ++n;
                } else {
                    mp = xorwow_next(s) >> 8;
                    mt = xorwow_next(s) >> 8;
                }
                while (s_cum[k] <= (uint32_t)h) ++k;
                const int p = k / a.nQ, g = k - p * a.nQ;
                const int band = s_ord[p];
                float w, winv;
                bundle_weight(a, a.Ib, Ti, band, g, &w, &winv);
                uint32_t w0, w1;
                desc_pack(mt, mp, band, g, &w0, &w1);
                const long long q = sl * R + h;
                a.w0[q] = w0;
                a.w1[q] = w1;
                a.w[q] = w;
                a.winv[q] = winv;
            }
        }
        // the next slot reuses the histogram
        __syncthreads();
        for (int i = tid; i < nbin; i += nth) s_hist[i] = 0u;
        __syncthreads();
    }
}

// -------------------------------------------------------------- trace ----
__device__ __forceinline__ gfloat* lev_T(const int* L)
{
    unsigned long long u = ((unsigned long long)(uint32_t)L[8] << 32) | (uint32_t)L[7];
    return (gfloat*)u;
}

// Temperature of bricked cell index idx: a 32-bit byte offset from the level
// base (fields stay below 4 GiB), so a uniform base addresses in one VGPR.
__device__ __forceinline__ float load_T(gfloat* base, int idx)
{
    typedef __attribute__((address_space(1))) const char gchar;
    return *(gfloat*)((gchar*)base + (uint32_t)idx * 4u);
}

// Wave-aggregated tally atomics: the lanes with `need` set hand their running
// slot sums over, grouped by slot; each group is summed across the wave and
// its first lane adds the group total with three FP64 atomics.  Called by all
// 64 lanes (wave-uniform control flow).
__device__ __forceinline__ void flush_wave(const TraceArgs& P, bool need, int aslot, double aE, double aA,
                                           double aQ, int lane)
{
    unsigned long long m = __ballot(need);
    while (m) {
        const int leader = __ffsll((long long)m) - 1;
        const int s = __shfl(aslot, leader);
        const bool mine = need && aslot == s;
        double e = mine ? aE : 0.0, a = mine ? aA : 0.0, q = mine ? aQ : 0.0;
        for (int off = 32; off > 0; off >>= 1) {
            e += __shfl_xor(e, off);
            a += __shfl_xor(a, off);
            q += __shfl_xor(q, off);
        }
        if (lane == leader) {
            unsafeAtomicAdd(&P.E[s], e);
            unsafeAtomicAdd(&P.A[s], a);
            unsafeAtomicAdd(&P.Q2[s], q);
        }
        need = need && !mine;
        m = __ballot(need);
    }
}

// Tally one finished ray into the lane's running slot sums (TMODE 0; a lane
// whose sums belong to another slot has handed them over with flush_wave
// first) or the per-ray buffer (TMODE 1).
template <int TMODE>
__device__ __forceinline__ void tally_ray(const TraceArgs& P, float er, float ar, int rslot, uint32_t q, int& aslot,
                                          double& aE, double& aA, double& aQ)
{
    if (TMODE == 1) {
        P.pray[q] = make_float2(er, ar);
        return;
    }
    if (rslot != aslot) {
        aslot = rslot;
        aE = aA = aQ = 0.0;
    }
    const double de = (double)er, da = (double)ar, x = de - da;
    aE += de;
    aA += da;
    aQ += x * x;
}

// A drawn ray, as the refill hands it to a lane: bricked index and packed
// (i, j, k) of the emitting cell, its tally slot, band | g << 16 | sign flags
// << 24, its processing slot in the launch, emitter temperature, direction and
// its reciprocals, weights, emission point (the cell centre).
struct RayStart {
    int idx, cell, six, bg, ps;
    float Tc, sx, sy, sz, ix, iy, iz, rw, rwi, px, py, pz;
};

// Lane `src`'s ray (all 64 lanes call it).
__device__ __forceinline__ RayStart pull_ray(const RayStart& r, int src)
{
    RayStart o;
    o.idx = __shfl(r.idx, src);
    o.cell = __shfl(r.cell, src);
    o.six = __shfl(r.six, src);
    o.bg = __shfl(r.bg, src);
    o.ps = __shfl(r.ps, src);
    o.Tc = __shfl(r.Tc, src);
    o.sx = __shfl(r.sx, src);
    o.sy = __shfl(r.sy, src);
    o.sz = __shfl(r.sz, src);
    o.ix = __shfl(r.ix, src);
    o.iy = __shfl(r.iy, src);
    o.iz = __shfl(r.iz, src);
    o.rw = __shfl(r.rw, src);
    o.rwi = __shfl(r.rwi, src);
    o.px = __shfl(r.px, src);
    o.py = __shfl(r.py, src);
    o.pz = __shfl(r.pz, src);
    return o;
}

// ------------------------------------------ learned draw inside the traversal
// The learned proposal of one emitting cell is evaluated by the whole wave
// whose queue draws that cell's rays: the network (features, two tanh
// layers, the output layer) and the running sums U of exp(clamp(z, -60, 60))
// over each head, written to the wave's row buffer in the layout of
// mlp_proposal's rows ([64 direction bins | nbin (band, g) bins | padding]).
// Each ray then searches the defensive mixture of both heads in that row
// (mix_search, as raygen_learned does from its rows).  It replaces, for the
// learned estimator, the draw from the fixed proposal that the reference
// makes where a bundle starts (code/mc_gpu.cu:138-144, find_band :830-896).
// Lane L computes columns col = 512 j + 8 L + c (j < 4, c < 8): the output
// layer's weights are then read 1 KB-contiguous per load instruction.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// c + a.lo b.lo + a.hi b.hi on BF16 pairs, FP32 accumulation (v_dot2_f32_bf16).
__device__ __forceinline__ float dot2bf(uint32_t a, uint32_t b, float c)
{
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, a), __builtin_bit_cast(bf16x2, b), c, false);
}

__device__ __forceinline__ uint32_t readlane_u(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }

// The trace kernel's argument block (its only argument, at offset 0 of the
// kernel argument segment) through a pointer the compiler cannot see through:
// fields read through it are loaded (scalar loads from the constant segment)
// where they are used instead of being hoisted to the kernel's start and held
// in scalar registers through the traversal loop.
typedef __attribute__((address_space(4))) const TraceArgs KArgs;
__device__ __forceinline__ KArgs* kernel_args()
{
    KArgs* k = (KArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(k));
    return k;
}

// The row of emitting cell ijk (i | j << 10 | k << 20) of snapshot
// `snapshot` into U (all 64 lanes call it): H1 = tanh(X W1 + b1), H2 =
// tanh(H1 W2 + b2), Z = H2 W3 + b3, H1 and H2 rounded to BF16 as the training
// kernels round them, the products in FP32 from BF16 pairs; U per head from
// 8-column sums within a lane and exclusive lane scans across (the direction
// head is lanes 0..7 of chunk 0).
// An out-of-line call: the registers it needs are saved around the call,
// once per row, instead of pressing on the traversal loop's registers.
// Arguments are plain values (the network's pointers and sizes, the feature
// inputs), so the kernel's argument block stays in its constant segment.
__device__ __noinline__ void learned_row(const uint16_t* W1, const float* b1, const uint16_t* W2, const float* b2,
                                         const uint16_t* W3, const float* b3, int nbin, int noutp, const float* T0,
                                         int n0, int n1, int n2, int nbx, int nby, const float* centres, int co0,
                                         int co1, int co2, int bcp, int polar, float Lpolar, float inv_Tmax,
                                         float Tw_rel, int ijk, int lane, float* U)
{
    FeatArgs f;
    f.T0 = T0;
    f.n[0] = n0; f.n[1] = n1; f.n[2] = n2;
    f.nbx = nbx; f.nby = nby;
    f.centres = centres;
    f.coff[0] = co0; f.coff[1] = co1; f.coff[2] = co2;
    f.bc_periodic = bcp;
    f.polar = polar;
    f.Lpolar = Lpolar;
    f.inv_Tmax = inv_Tmax;
    f.Tw_rel = Tw_rel;
    float x[kMlpIn];
    cell_features(f, ijk & 1023, (ijk >> 10) & 1023, (ijk >> 20) & 1023, x);
    MlpWeights w;
    w.W1 = W1; w.b1 = b1; w.W2 = W2; w.b2 = b2; w.W3 = W3; w.b3 = b3;
    w.nbin = nbin; w.noutp = noutp;
    w.W2T = nullptr; w.W3T = nullptr;
    float a1 = w.b1[lane];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) a1 = fmaf(x[i], bf2f(w.W1[i * kMlpHid + lane]), a1);
    const uint32_t h1 = f2bf(tanhf(a1));
    float a2 = w.b2[lane];
    for (int kp = 0; kp < kMlpHid / 2; ++kp) {
        const uint32_t hp = readlane_u(h1, 2 * kp) | (readlane_u(h1, 2 * kp + 1) << 16);
        const uint32_t wp = (uint32_t)w.W2[(2 * kp) * kMlpHid + lane] | ((uint32_t)w.W2[(2 * kp + 1) * kMlpHid + lane] << 16);
        a2 = dot2bf(hp, wp, a2);
    }
    const uint32_t h2 = f2bf(tanhf(a2));
    const int nout = kMlpDir + w.nbin;
    float base = 0.f;   // the (band, g) head's sum over the previous chunks
    for (int j = 0; j < 4; ++j) {
        const int col = 512 * j + 8 * lane;
        float a[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) a[c] = 0.f;
        if (col < noutp) {
            const uint16_t* r = w.W3 + col;
#pragma unroll 2
            for (int kp = 0; kp < kMlpHid / 2; ++kp) {
                const uint32_t hp = readlane_u(h2, 2 * kp) | (readlane_u(h2, 2 * kp + 1) << 16);
                const uint4 u0 = *(const uint4*)(r + (size_t)(2 * kp) * noutp);
                const uint4 u1 = *(const uint4*)(r + (size_t)(2 * kp + 1) * noutp);
                a[0] = dot2bf(hp, __builtin_amdgcn_perm(u1.x, u0.x, 0x05040100u), a[0]);
                a[1] = dot2bf(hp, __builtin_amdgcn_perm(u1.x, u0.x, 0x07060302u), a[1]);
                a[2] = dot2bf(hp, __builtin_amdgcn_perm(u1.y, u0.y, 0x05040100u), a[2]);
                a[3] = dot2bf(hp, __builtin_amdgcn_perm(u1.y, u0.y, 0x07060302u), a[3]);
                a[4] = dot2bf(hp, __builtin_amdgcn_perm(u1.z, u0.z, 0x05040100u), a[4]);
                a[5] = dot2bf(hp, __builtin_amdgcn_perm(u1.z, u0.z, 0x07060302u), a[5]);
                a[6] = dot2bf(hp, __builtin_amdgcn_perm(u1.w, u0.w, 0x05040100u), a[6]);
                a[7] = dot2bf(hp, __builtin_amdgcn_perm(u1.w, u0.w, 0x07060302u), a[7]);
            }
        }
        // e = exp(clamp(z)), running sums within the lane's 8 columns
        float run = 0.f;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (col + c < nout) run += __expf(fminf(fmaxf(a[c] + w.b3[col + c], -60.f), 60.f));
            a[c] = run;
        }
        // exclusive scan of the lanes' sums, the direction head on its own
        const int seg = (j == 0 && lane >= 8) ? 8 : 0;
        float incl = run;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const float t = __shfl_up(incl, o);
            if (lane - o >= seg) incl += t;
        }
        float excl = __shfl_up(incl, 1);
        if (lane == seg) excl = 0.f;
        const float sb = (j == 0 && lane < 8) ? excl : base + excl;
        if (col < noutp) {
            float4* d = (float4*)(U + col);
            d[0] = make_float4(sb + a[0], sb + a[1], sb + a[2], sb + a[3]);
            d[1] = make_float4(sb + a[4], sb + a[5], sb + a[6], sb + a[7]);
        }
        base = base + __shfl(incl, 63);
    }
}

// Whether ray q of snapshot qs of a fused launch, with polar uniform bits mt
// (m_theta of the plain draw), is a listed long-ray candidate
// (TraceArgs::cand): one integer test for cand_scan and the traversal's draw.
template <typename A>
__device__ __forceinline__ bool long_candidate(const A& P, uint32_t q, uint32_t qs, uint32_t mt)
{
    FastDiv fd;   // (field by field: P may be the kernel argument segment)
    fd.m = P.cand_fd.m;
    fd.s = P.cand_fd.s;
    fd.d = P.cand_fd.d;
    const uint32_t x = fdiv(q, fd);
    return qs + 1u == (uint32_t)P.nsnap_launch && q - x * fd.d >= P.cand_from &&
           mt - (1u << 23) + P.cand_w < 2u * P.cand_w;
}

// The long-ray candidates of a fused launch: each ray of the final part of
// the 8 XCD ranges of the launch's last snapshot (nfin rays per range from
// cand_from on, those below nrays) gets the plain draw's Philox block, and
// the rays that pass long_candidate go to their range's list (list x at
// x * cand_cap, its count at cand_cnt[64 x]; every hit is counted, also past
// cand_cap).  A block takes tiles of kCandTile rays of one range and stages
// the tile's hits in LDS: one global atomic per tile.
constexpr uint32_t kCandTile = 4096;
__global__ __launch_bounds__(256) void cand_scan(TraceArgs P, uint32_t nfin, uint32_t* list)
{
    __shared__ uint32_t s_q[kCandTile];
    __shared__ uint32_t s_n, s_base;
    const uint32_t qs = (uint32_t)P.nsnap_launch - 1u;
    const uint32_t tpr = (nfin + kCandTile - 1u) / kCandTile;   // tiles per range
    for (uint32_t t = blockIdx.x; t < 8u * tpr; t += gridDim.x) {
        const uint32_t x = t / tpr;
        const uint32_t j0 = (t - x * tpr) * kCandTile;
        const uint32_t jn = min(kCandTile, nfin - j0);
        const uint32_t q0 = x * P.cand_fd.d + P.cand_from + j0;
        if (threadIdx.x == 0) s_n = 0u;
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < jn; j += blockDim.x) {
            const uint32_t q = q0 + j;
            if (q < (uint32_t)P.nrays) {
                const uint32_t slot = fdiv(q, P.fd_R);
                const uint32_t r = q - slot * (uint32_t)P.rays_per_slot;
                const int4 si = P.slotinfo[P.slot0 + slot];
                const u32x4 h = philox4x32_10(u32x4{r, (uint32_t)si.y, (uint32_t)(P.snap + (int)qs), 0u},
                                              (uint32_t)P.seed, (uint32_t)(P.seed >> 32));
                if (long_candidate(P, q, qs, h.z >> 8)) s_q[atomicAdd(&s_n, 1u)] = q;
            }
        }
        __syncthreads();
        const uint32_t n = s_n;
        if (threadIdx.x == 0 && n) s_base = atomicAdd(P.cand_cnt + 64u * x, n);
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < n; k += blockDim.x)
            if (s_base + k < P.cand_cap) list[(size_t)x * P.cand_cap + s_base + k] = s_q[k];
        __syncthreads();   // the tile's hits are out before the next tile's reset
    }
}

// Persistent cell-by-cell traversal (code/mc_gpu.cu:228-438).
//
// Work: the launch's rays (in 3-D brick order of their emitting cells) are
// cut into super-chunks of `super` consecutive rays.  Blocks are dispatched
// round-robin over the 8 XCDs, so XCD x owns the contiguous super-chunk range
// [x*N/8, (x+1)*N/8); its waves claim super-chunks of that range in order from
// a per-XCD counter (one atomic per super-chunk), so the rays in flight on an
// XCD stay inside a compact region its L2 holds and the 64 lanes of a wave
// trace consecutive rays of the same one to three cells.  A wave whose XCD
// range is used up claims from the next XCD's range, so no XCD idles while
// another still has work.  The grid is sized to the resident capacity.
//
// Lanes whose ray ended stay idle until `refill_k` lanes of the wave are idle
// (or none is alive); one wave-uniform refill block then tallies the finished
// rays and starts new rays on every idle lane, in lane order.  With atomic
// tallies and no records (QUEUE) the new rays come from the wave's queue of 64
// drawn rays (lane j holds ray gbase + j), redrawn 64 at a time on the whole
// wave; otherwise each idle lane draws its own ray.
//
// Tallies: QUEUE keeps a ring of kRing slot accumulators per wave in LDS
// (FP64 E, A, sum x^2): the processing slots of a drawn queue get entries
// (slot mod kRing; an entry still holding an older slot first adds its sums
// to that slot with three device FP64 atomics), and a finished ray adds to
// its entry with LDS atomics while the entry holds its slot, else straight to
// the slot.  TMODE 0 with records keeps a running FP64 sum per lane, handed
// over with wave-aggregated atomics (flush_wave) when the lane moves to
// another slot.  TMODE 1 writes each ray's (E, A) for the deterministic
// reduce_rays pass.
//
// One step: the kappa gather of the current cell is issued first, the
// geometry (distances, face snap, move, boundary) runs while it is in flight,
// the next cell's field load is issued at the end of the geometry, and the
// optical depth / transmissivity / tally update consumes kappa last.  With
// the kappa/Ib slice in LDS the field holds table coordinates (t_to_u).
//
// GEN: the rays are drawn in the kernel instead of read from descriptors
// (1: the plain estimator's Philox draw, bit-identical to raygen_plain; 2:
// the learned proposal, evaluated per emitting cell by the wave that draws
// its rays, see learned_row), and one launch covers every snapshot: XCD x's
// claims walk snapshot by snapshot through its 1/8 of each snapshot's
// super-chunks, so all XCDs work on the same snapshot at a time.
//
// Long rays first (GEN == 1, a list from cand_scan): a ray along the walls
// of an optically thin bin can run to the step cap, tens of ms on its own;
// started among the launch's last rays it would hold the launch open after
// all other work is done.  The waves draw their queues from the list first
// and skip the listed rays (the same test) where their super-chunks come up:
// every ray is still traced once, with the same arithmetic.
template <int SPEC, int TAB, int TMODE, bool REC, bool AMC, int GEN>
__global__ __launch_bounds__(TAB == kTabSliced ? 1024 : 256) void trace(TraceArgs P)
{
    extern __shared__ float lds[];
    // the wave's tally ring (atomic tallies, no records), first in LDS:
    // kRing entries of FP64 (E, A, sum x^2), the processing slot each holds
    // (-1: free) and its tally slot
    constexpr bool QUEUE = TMODE == 0 && !REC;
    const int nwv = QUEUE ? (int)(blockDim.x >> 6) : 0;
    const int rb = (int)(threadIdx.x >> 6) * kRing;   // this wave's first entry
    double* s_rv = (double*)lds;
    int* s_rt = (int*)(s_rv + 3 * kRing * nwv);
    int* s_rg = s_rt + kRing * nwv;
    int* s_lev = s_rg + kRing * nwv;            // 6 * kLevWords
    float* s_face = (float*)(s_lev + 6 * kLevWords);
    int* s_off = (int*)(s_face + P.nfaces);     // brick offsets per level and axis
    float* s_ibw = (float*)(s_off + P.noffs);   // 6 * nB wall Planck values
    // face i of the face table at byte offset fb, brick offset of 1-based
    // cell c of the table at byte offset ob (both offsets from the level block)
    auto face_at = [&](int fb, int i) -> float { return *(const float*)((const char*)s_face + fb + (i << 2)); };
    auto off_at = [&](int ob, int c) -> int { return *(const int*)((const char*)s_off + ob + (c << 2)); };
    float* s_ib = s_ibw + 6 * P.nB;             // Ib rows (nT or tab_nw nodes)
    const int nw = P.tab_nw;
    float* s_kq = s_ib + P.nB * nw;             // sliced kappa rows (kTabSliced)
    // GEN: band CDF after the tables that are in LDS
    float* s_cdfb = TAB == kTabSliced ? s_kq + P.nB * P.nQ * nw : (TAB == kTabIbLds ? s_ib + P.nB * P.nT : s_ib);
    float* s_cdfg = s_cdfb + P.nB;              // GEN, gen_lds bit 0: g CDF rows; bit 1: 1/(k Ib)(Tmax)
    float* s_iem = s_cdfg + P.nB * P.nQ;
    for (int i = threadIdx.x; i < P.nlev; i += blockDim.x) {
        int* L = s_lev + i * kLevWords;
        L[0] = P.lev[i].n[0]; L[1] = P.lev[i].n[1]; L[2] = P.lev[i].n[2];
        // face and brick-offset table positions as byte offsets into s_face
        // and s_off (the offsets' -1 of the 1-based cell index folded in)
        L[3] = 4 * P.lev[i].foff[0]; L[4] = 4 * P.lev[i].foff[1]; L[5] = 4 * P.lev[i].foff[2];
        L[6] = P.lev[i].maxs;
        const float* tb = P.T[i] + (size_t)P.snap * P.lev[i].tstride;
        unsigned long long u = (unsigned long long)tb;
        L[7] = (int)(uint32_t)u;
        L[8] = (int)(uint32_t)(u >> 32);
        L[9] = P.lev[i].nbx;
        L[10] = P.lev[i].nby;
        L[11] = 4 * (P.lev[i].ooff[0] - 1);
        L[12] = 4 * (P.lev[i].ooff[1] - 1);
        L[13] = 4 * (P.lev[i].ooff[2] - 1);
        L[14] = (int)P.lev[i].tstride;
    }
    for (int i = threadIdx.x; i < P.nfaces; i += blockDim.x) s_face[i] = P.faces[i];
    for (int i = threadIdx.x; i < P.noffs; i += blockDim.x) s_off[i] = P.offs[i];
    for (int i = threadIdx.x; i < 6 * P.nB; i += blockDim.x) s_ibw[i] = P.Ibw[i];
    if (QUEUE && (threadIdx.x & 63) < kRing) {
        const int e = rb + (int)(threadIdx.x & 63);
        s_rt[e] = -1;
        s_rg[e] = 0;
        s_rv[3 * e] = s_rv[3 * e + 1] = s_rv[3 * e + 2] = 0.0;
    }
    if (GEN) {
        for (int i = threadIdx.x; i < P.nB; i += blockDim.x) s_cdfb[i] = P.cdf_band[i];
        if (P.gen_lds & 1)
            for (int i = threadIdx.x; i < P.nB * P.nQ; i += blockDim.x) s_cdfg[i] = P.cdf_g[i];
        if (P.gen_lds & 2)
            for (int i = threadIdx.x; i < P.nB * P.nQ; i += blockDim.x) s_iem[i] = P.inv_emax[i];
    }
    if (SPEC == kSpecTable && TAB == kTabIbLds)
        for (int i = threadIdx.x; i < P.nB * P.nT; i += blockDim.x) s_ib[i] = P.Ib[i];
    if (SPEC == kSpecTable && TAB == kTabSliced) {
        for (int i = threadIdx.x; i < P.nB * nw; i += blockDim.x) {
            const int r = i / nw;
            s_ib[i] = P.Ib[r * P.nT + P.tab_lo + (i - r * nw)];
        }
        for (int i = threadIdx.x; i < P.nB * P.nQ * nw; i += blockDim.x) {
            const int r = i / nw;
            s_kq[i] = P.kq[r * P.nT + P.tab_lo + (i - r * nw)];
        }
    }
    __syncthreads();
    const float* ibt = TAB == kTabGlobal ? P.Ib : s_ib;
    gfloat* kq = (gfloat*)P.kq;
    const int rowlen = TAB == kTabSliced ? nw : P.nT;   // nodes per table row

    const int lane = threadIdx.x & 63;
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const uint32_t nrays = (uint32_t)P.nrays;   // rays per snapshot of the launch
    const uint32_t S = (uint32_t)P.super;
    const uint32_t cps = (nrays + S - 1) / S;   // super-chunks per snapshot
    const bool split = (nb % 8u) == 0u;
    const uint32_t xcd = split ? b % 8u : 0u;
    const uint32_t nparts = split ? 8u : 1u;
    const uint32_t per_x = (cps + nparts - 1u) / nparts;        // per XCD and snapshot
    const uint32_t nclaim = per_x * (uint32_t)P.nsnap_launch;   // claims of one XCD counter
    const int nT = P.nT;

    // wave-uniform cursor over the claimed super-chunk (rays cur..cend of snapshot csnap)
    uint32_t xi = 0, cur = 0, cend = 0, csnap = 0;
    bool more = true;
    // lane state
    bool alive = false, fin = false;
    uint32_t qcur = 0;
    int rslot = -1, aslot = -1;
    double aE = 0.0, aA = 0.0, aQ = 0.0;
    uint32_t steps_lane = 0;
    int endt = 0;
    float Dfin = 1.0f;
    float px = 0.f, py = 0.f, pz = 0.f, sx = 1.f, sy = 1.f, sz = 1.f, ix = 1.f, iy = 1.f, iz = 1.f;
    float fx = 0.f, fy = 0.f, fz = 0.f, tau = 0.f, tra = 1.f, G = 0.f, rw = 0.f, rwi = 0.f, Tc = 0.f;
    int cx = 1, cy = 1, cz = 1, fl = 0, krow = 0, ibrow = 0, band = 0, lev = 0, cnt = 0, idx = 0, rsteps = 0;
    uint32_t hash = 2166136261u;
    int n0 = s_lev[0], n1 = s_lev[1], n2 = s_lev[2];
    int fo0 = s_lev[3], fo1 = s_lev[4], fo2 = s_lev[5], maxs = s_lev[6];
    int oo0 = s_lev[11], oo1 = s_lev[12], oo2 = s_lev[13];
    if (!AMC) {   // one level: its parameters are the same for every lane (scalar registers)
        n0 = __builtin_amdgcn_readfirstlane(n0); n1 = __builtin_amdgcn_readfirstlane(n1);
        n2 = __builtin_amdgcn_readfirstlane(n2);
        fo0 = __builtin_amdgcn_readfirstlane(fo0); fo1 = __builtin_amdgcn_readfirstlane(fo1);
        fo2 = __builtin_amdgcn_readfirstlane(fo2); maxs = __builtin_amdgcn_readfirstlane(maxs);
        oo0 = __builtin_amdgcn_readfirstlane(oo0); oo1 = __builtin_amdgcn_readfirstlane(oo1);
        oo2 = __builtin_amdgcn_readfirstlane(oo2);
    }
    gfloat* Tl = lev_T(s_lev);
    gfloat* T0 = Tl;
    uint32_t rsnap = 0;
    // the wave's queue of drawn rays (lane j: ray gbase + j, j < gn); the
    // per-ray outputs (TMODE 1) and records index rays by lane, so they keep
    // the direct hand-out
    uint32_t gn = 0, ghead = 0, gsnap = 0;
    int pslot = 0;                        // processing slot of the lane's ray (tally ring tag)
    const int nslot_l = (int)(nrays / (uint32_t)P.rays_per_slot);   // processing slots per snapshot
    // the finished ray's (e, a): into its ring entry while the entry holds its
    // slot, else (evicted) straight to the tally
    auto ring_tally = [&](float er, float ar) {
        const double de = (double)er, da = (double)ar, x = de - da;
        const int e = rb + (pslot & (kRing - 1));
        if (s_rt[e] == pslot) {
            __hip_atomic_fetch_add(&s_rv[3 * e], de, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(&s_rv[3 * e + 1], da, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(&s_rv[3 * e + 2], x * x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            unsafeAtomicAdd(&P.E[rslot], de);
            unsafeAtomicAdd(&P.A[rslot], da);
            unsafeAtomicAdd(&P.Q2[rslot], x * x);
        }
    };
    // ring entry e's sums to its tally slot (one lane)
    auto ring_flush = [&](int e) {
        const int t = s_rg[e];
        unsafeAtomicAdd(&P.E[t], s_rv[3 * e]);
        unsafeAtomicAdd(&P.A[t], s_rv[3 * e + 1]);
        unsafeAtomicAdd(&P.Q2[t], s_rv[3 * e + 2]);
    };
    RayStart Qr = {};
    // learned draw (GEN == 2): this lane's drawn bins and p/q for draw_ray, and
    // the wave's cached proposal row (processing slot and snapshot)
    int lr_band = 0, lr_g = 0;
    uint32_t lr_mt = 0, lr_mp = 0;
    float lr_ratio = 0.f;
    uint32_t row_slot = 0xFFFFFFFFu, row_snap = 0xFFFFFFFFu;
    const uint32_t wave_id = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    // long rays first, one scalar: bit 8 on (off when a list found more rays
    // than it holds; with none found no ray passes the test either), bit 31
    // while the wave still draws from the lists, bits 0..3 the lists it has
    // used up (its XCD's first, then the following XCDs')
    uint32_t cand = 0;
    if (GEN == 1 && QUEUE && P.cand != nullptr) {
        uint32_t tot = 0;
        bool fits = true;
        for (int x = 0; x < 8; ++x) {
            const uint32_t n = __builtin_amdgcn_readfirstlane(P.cand_cnt[64 * x]);
            fits = fits && n <= P.cand_cap;
            tot += n;
        }
        cand = fits && tot > 0u ? 0x80000100u : 0u;
    }

    // ray q of snapshot qs of the launch: drawn here (GEN, the plain estimator's
    // Philox draw, bit-identical to raygen_plain) or read from its descriptors.
    // A listed long ray drawn from its super-chunk (not from the list) comes
    // back with bit 31 of bg set: it is not started there.
    auto draw_ray = [&](uint32_t q, uint32_t qs, bool listed) -> RayStart {
        RayStart rd;
        const uint32_t slot = fdiv(q, P.fd_R);
        const int4 si = P.slotinfo[P.slot0 + slot];
        rd.six = si.x;
        rd.idx = si.z;
        rd.cell = si.w;
        rd.ps = (int)slot + (int)qs * nslot_l;
        rd.px = P.centres[P.coff[0] + (si.w & 1023) + 1];
        rd.py = P.centres[P.coff[1] + ((si.w >> 10) & 1023) + 1];
        rd.pz = P.centres[P.coff[2] + ((si.w >> 20) & 1023) + 1];
        gfloat* Te = lev_T(s_lev) + (GEN ? (size_t)qs * (uint32_t)s_lev[14] : (size_t)0);
        const float Tq = Te[si.z];
        rd.Tc = Tq;
        int bnd, g;
        uint32_t mt, mp;
        float ratio = 1.0f;       // p/q of a learned descriptor
        bool weigh = GEN;         // emission weight computed here (tables in LDS)
        bool skip = false;        // a listed long ray, traced from the list
        if (GEN == 1) {
            const uint32_t r = q - slot * (uint32_t)P.rays_per_slot;
            const u32x4 x = philox4x32_10(u32x4{r, (uint32_t)si.y, (uint32_t)(P.snap + (int)qs), 0u},
                                          (uint32_t)P.seed, (uint32_t)(P.seed >> 32));
            bnd = cdf_search(s_cdfb, P.nB, u01_open(x.x));
            if (!P.g_search)
                g = 0;
            else if (P.gen_lds & 1)
                g = cdf_search(s_cdfg + bnd * P.nQ, P.nQ, u01_open(x.y));
            else
                g = cdf_search(P.cdf_g + bnd * P.nQ, P.nQ, u01_open(x.y));
            g = g > P.nQ - 1 ? P.nQ - 1 : g;
            mt = x.z >> 8;
            mp = x.w >> 8;
            // (the test's arguments read from the kernel argument segment here)
            skip = (cand & 0x100u) && !listed && long_candidate(*kernel_args(), q, qs, mt);
        } else if (GEN == 2) {
            bnd = lr_band;
            g = lr_g;
            mt = lr_mt;
            mp = lr_mp;
            ratio = lr_ratio;
        } else {
            const uint32_t a0 = P.w0[q], a1 = P.w1[q];
            bnd = desc_band(a0, a1);
            g = desc_g(a1);
            mt = a0 & 0xFFFFFFu;
            mp = a1 & 0xFFFFFFu;
            if (P.desc_ratio) {
                ratio = P.w[q];
                weigh = true;
            } else {
                rd.rw = P.w[q];
                rd.rwi = P.winv[q];
            }
        }
        if (weigh) {
            // Ibmax (k_i Ib_i) / (k_max Ib_max) (code/mc_gpu.cu:63, :163-169),
            // the arithmetic of the generators' bundle weights
            float kap, ib;
            if (SPEC == kSpecGreyT4) {
                const float t2 = Tq * Tq;
                kap = P.kappa_grey;
                ib = t2 * t2;
            } else {
                // sliced tables: Tq is the table coordinate within the slice
                int t;
                float fr;
                if (TAB == kTabSliced) {
                    t = (int)Tq;
                    t = t > nw - 2 ? nw - 2 : t;
                    fr = Tq - (float)t;
                } else {
                    float pos = fmaf(Tq, P.inv_dT, P.mT0);
                    pos = fminf(fmaxf(pos, 0.0f), (float)(nT - 1));
                    t = (int)pos;
                    t = t > nT - 2 ? nT - 2 : t;
                    fr = pos - (float)t;
                }
                const int kr = (bnd * P.nQ + g) * rowlen, ir = bnd * rowlen;
                const float k0 = TAB == kTabSliced ? s_kq[kr + t] : kq[kr + t];
                const float k1 = TAB == kTabSliced ? s_kq[kr + t + 1] : kq[kr + t + 1];
                const float i0 = ibt[ir + t], i1 = ibt[ir + t + 1];
                kap = fmaf(fr, k1 - k0, k0);
                ib = fmaf(fr, i1 - i0, i0);
            }
            const float iem = (P.gen_lds & 2) ? s_iem[bnd * P.nQ + g] : P.inv_emax[bnd * P.nQ + g];
            const float w = P.Ibmax * (kap * ib) * iem;
            const float wi = w * (1.0f / ib);
            rd.rw = GEN == 1 ? w : w * ratio;
            rd.rwi = GEN == 1 ? wi : wi * ratio;
        }
        const float ut = ((float)mt + P.u_off) * kTwoM24;
        const float up = ((float)mp + P.u_off) * kTwoM24;
        const float cth = 1.0f - 2.0f * ut;
        const float sth = 2.0f * sqrtf(ut * (1.0f - ut));
        float sp, cp;
        sincos2pi(up, &sp, &cp);
        const float s1 = sth * cp, s2 = sth * sp;
        float dx = P.polar == 0 ? cth : (P.polar == 1 ? s2 : s1);
        float dy = P.polar == 0 ? s1 : (P.polar == 1 ? cth : s2);
        float dz = P.polar == 0 ? s2 : (P.polar == 1 ? s1 : cth);
        if (dx == 0.0f) dx = 1e-10f;
        if (dy == 0.0f) dy = 1e-10f;
        if (dz == 0.0f) dz = 1e-10f;
        const int f = (dx < 0.0f ? 1 : 0) | (dy < 0.0f ? 2 : 0) | (dz < 0.0f ? 4 : 0);
        rd.sx = dx; rd.sy = dy; rd.sz = dz;
        rd.ix = 1.0f / dx;
        rd.iy = 1.0f / dy;
        rd.iz = 1.0f / dz;
        rd.bg = bnd | (g << 16) | (f << 24) | (skip ? (int)(1u << 31) : 0);
        return rd;
    };

    for (;;) {
        const bool idle = !alive;
        const unsigned long long m_idle = __ballot(idle);
        const int ni = __popcll(m_idle);
        const int na = 64 - ni;
        const bool have_rays = more || cur < cend || ghead < gn || (int)cand < 0;
        if (na == 0 && !have_rays) break;
        if (have_rays && (ni >= P.refill_k || na == 0)) {
            // ---- tally the rays finished since the last refill
            if (!QUEUE && TMODE == 0) flush_wave(P, fin && rslot != aslot && aslot >= 0, aslot, aE, aA, aQ, lane);
            if (fin) {
                if (QUEUE)
                    ring_tally(rw * Dfin, rwi * G);
                else
                    tally_ray<TMODE>(P, rw * Dfin, rwi * G, rslot, qcur, aslot, aE, aA, aQ);
                if (REC && qcur < P.nrec) {
                    P.rec_steps[qcur] = rsteps;
                    P.rec_end[qcur] = endt;
                    P.rec_hash[qcur] = hash;
                    P.rec_D[qcur] = Dfin;
                    P.rec_G[qcur] = G;
                }
                steps_lane += (uint32_t)rsteps;
                fin = false;
            }
            // idle lanes below this one (mbcnt: no lane mask held in registers)
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m_idle >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)m_idle, 0u));
            // the drawn ray rd of snapshot rsn becomes this lane's ray
            auto start_ray = [&](const RayStart& rd, uint32_t rsn) {
                rsnap = rsn;
                rslot = rd.six + (int)rsnap * P.tally_stride;
                if (GEN) T0 = lev_T(s_lev) + (size_t)rsnap * (uint32_t)s_lev[14];
                Tc = rd.Tc;
                band = rd.bg & 0xFFFF;
                fl = (rd.bg >> 24) & 7;
                sx = rd.sx; sy = rd.sy; sz = rd.sz;
                ix = rd.ix; iy = rd.iy; iz = rd.iz;
                rw = rd.rw;
                rwi = rd.rwi;
                const int g = (rd.bg >> 16) & 0xFF;
                cx = (rd.cell & 1023) + 1;
                cy = ((rd.cell >> 10) & 1023) + 1;
                cz = ((rd.cell >> 20) & 1023) + 1;
                px = rd.px;
                py = rd.py;
                pz = rd.pz;
                pslot = rd.ps;
                if (AMC && lev != 0) {
                    lev = 0;
                    n0 = s_lev[0]; n1 = s_lev[1]; n2 = s_lev[2];
                    fo0 = s_lev[3]; fo1 = s_lev[4]; fo2 = s_lev[5]; maxs = s_lev[6];
                    oo0 = s_lev[11]; oo1 = s_lev[12]; oo2 = s_lev[13];
                }
                Tl = T0;
                fx = face_at(fo0, cx - (fl & 1));
                fy = face_at(fo1, cy - ((fl >> 1) & 1));
                fz = face_at(fo2, cz - ((fl >> 2) & 1));
                idx = rd.idx;
                tau = 0.0f;
                tra = 1.0f;
                G = 0.0f;
                krow = (band * P.nQ + g) * rowlen;
                ibrow = band * rowlen;
                cnt = 0;
                hash = 2166136261u;
                rsteps = 0;
                alive = true;
            };
            if (QUEUE) {
                // ---- from the wave's queue of drawn rays: queue lane j holds ray
                // gbase + j of snapshot gsnap (j < gn), and the idle lane of rank r
                // takes queue lane ghead + r.  An exhausted queue is redrawn (or
                // re-read from the descriptors) for the next (up to) 64 rays of the
                // claimed super-chunk, one ray per lane, so that work runs on whole
                // waves rather than on the idle lanes of each refill.
                const uint32_t t1 = min((uint32_t)ni, gn - ghead);
                {
                    const bool s1 = idle && rank < t1;
                    const RayStart rd = pull_ray(Qr, (int)((ghead + (s1 ? rank : 0u)) & 63u));
                    if (s1 && (GEN != 1 || rd.bg >= 0)) start_ray(rd, gsnap);
                }
                ghead += t1;
                if (t1 < (uint32_t)ni) {
                    bool drawn = false;
                    if (GEN == 1 && (int)cand < 0) {
                        // the next (up to) 64 listed long rays, all of the last
                        // snapshot, from this XCD's list first; no ring entries (their
                        // slots are scattered): they tally to their slots, or to an
                        // entry that holds the same slot
                        const KArgs* K = kernel_args();
                        while ((int)cand < 0) {
                            const uint32_t X = (xcd + (cand & 15u)) & 7u;
                            const uint32_t n = __builtin_amdgcn_readfirstlane(K->cand_cnt[64u * X]);
                            uint32_t c0 = 0;
                            if (lane == 0) c0 = atomicAdd(K->cand_cnt + 64u * X + 32u, 64u);
                            c0 = __builtin_amdgcn_readfirstlane(c0);
                            if (c0 >= n) {   // used up: the next list, or none left
                                cand = (cand & 15u) == 7u ? (cand & 0x7FFFFFFFu) : cand + 1u;
                                continue;
                            }
                            gn = min(64u, n - c0);
                            gsnap = (uint32_t)K->nsnap_launch - 1u;
                            if (lane == 0) atomicAdd(K->nsteps + 1, (unsigned long long)gn);   // rays listed
                            if ((uint32_t)lane < gn)
                                Qr = draw_ray(K->cand[(size_t)X * K->cand_cap + c0 + (uint32_t)lane], gsnap, true);
                            drawn = true;
                            break;
                        }
                    }
                    if (!drawn && cur >= cend && more) {
                        // claim the next super-chunk, from this XCD's share first,
                        // then from the following XCDs' shares.  Claim c of XCD X is
                        // chunk X*per_x + c % per_x of snapshot c / per_x.
                        bool got = false;
                        while (xi < nparts) {
                            const uint32_t X = (xcd + xi) % nparts;
                            uint32_t c = 0;
                            if (lane == 0) c = atomicAdd(P.claim + 32u * X, 1u);
                            c = __builtin_amdgcn_readfirstlane(c);
                            if (c >= nclaim) {
                                ++xi;
                                continue;
                            }
                            const uint32_t sn = c / per_x;
                            const uint32_t lc = X * per_x + (c - sn * per_x);
                            if (lc >= cps) continue;   // past the snapshot's last chunk
                            csnap = sn;
                            cur = lc * S;
                            cend = cur + S < nrays ? cur + S : nrays;
                            got = true;
                            break;
                        }
                        if (!got) more = false;
                    }
                    if (drawn || cur < cend) {
                      if (!drawn) {
                        gn = min(64u, cend - cur);
                        gsnap = csnap;
                        if (GEN == 2) {
                            // the learned draws of rays cur .. cur + gn - 1: for each
                            // processing slot among them (wave-uniform) its proposal
                            // row in the wave's row buffer (evaluated unless it is the
                            // row the buffer holds), then the rays of that slot search it
                            const uint32_t q = cur + min((uint32_t)lane, gn - 1u);
                            const uint32_t slot = fdiv(q, P.fd_R);
                            const uint32_t r = q - slot * (uint32_t)P.rays_per_slot;
                            const int4 si = P.slotinfo[P.slot0 + slot];
                            const uint32_t snapshot = (uint32_t)P.snap + gsnap;
                            const u32x4 x = philox4x32_10(u32x4{r, (uint32_t)si.y, snapshot, 1u}, (uint32_t)P.seed,
                                                          (uint32_t)(P.seed >> 32));
                            const uint32_t s0 = readlane_u(slot, 0), s1 = readlane_u(slot, (int)gn - 1);
                            lr_band = 0; lr_g = 0; lr_mt = 0; lr_mp = 0; lr_ratio = 0.f;
                            // the network's arguments are read from the kernel
                            // argument segment here, not kept in registers through
                            // the traversal
                            const KArgs* K = kernel_args();
                            float* const row_buf = K->lrow_cache + (size_t)wave_id * (size_t)kLRowFloats;
                            for (uint32_t sl = s0; sl <= s1; ++sl) {
                                if (sl != row_slot || gsnap != row_snap) {
                                    learned_row(K->mlpw.W1, K->mlpw.b1, K->mlpw.W2, K->mlpw.b2, K->mlpw.W3, K->mlpw.b3,
                                                K->mlpw.nbin, K->mlpw.noutp,
                                                K->feat.T0 + (size_t)snapshot * (size_t)K->feat_stride, K->feat.n[0],
                                                K->feat.n[1], K->feat.n[2], K->feat.nbx, K->feat.nby, K->feat.centres,
                                                K->feat.coff[0], K->feat.coff[1], K->feat.coff[2], K->feat.bc_periodic,
                                                K->feat.polar, K->feat.Lpolar, K->feat.inv_Tmax, K->feat.Tw_rel,
                                                P.slotinfo[P.slot0 + sl].w, lane, row_buf);
                                    row_slot = __builtin_amdgcn_readfirstlane(sl);
                                    row_snap = __builtin_amdgcn_readfirstlane(gsnap);
                                    // the row's columns were written by every lane of the wave
                                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                                    if (K->log_rows) {
                                        const int no = K->mlpw.noutp;
                                        float* lr = K->log_rows + (size_t)(P.slot0 + sl) * (size_t)no;
                                        for (int cc = lane; cc < no; cc += 64) lr[cc] = row_buf[cc];
                                    }
                                }
                                if (slot == sl) {
                                    float qd, qb;
                                    int d, bg;
                                    mix_search2(row_buf, K->pdir, kMlpDir, row_buf + kMlpDir, K->pcdf, K->mlpw.nbin,
                                                K->mix_eps, u01_open(x.x), u01_open(x.y), &d, &qd, &bg, &qb);
                                    lr_band = bg / P.nQ;
                                    lr_g = bg - lr_band * P.nQ;
                                    lr_ratio = (qb > 0.f && qd > 0.f) ? (K->p_bg[bg] / qb) * ((1.0f / kMlpDir) / qd) : 0.f;
                                    // direction bin d = 8 cos(theta) bins x 8 phi bins: the bin in
                                    // the top 3 of the 24 uniform bits, the Philox bits below
                                    lr_mt = ((uint32_t)(d >> 3) << 21) | (x.z >> 11);
                                    lr_mp = ((uint32_t)(d & 7) << 21) | (x.w >> 11);
                                }
                            }
                            if (K->log_w0 && (uint32_t)lane < gn) {
                                uint32_t w0, w1;
                                desc_pack(lr_mt, lr_mp, lr_band, lr_g, &w0, &w1);
                                const size_t o = (size_t)gsnap * (size_t)P.nrays + (cur + (uint32_t)lane);
                                K->log_w0[o] = w0;
                                K->log_w1[o] = w1;
                                K->log_w[o] = lr_ratio;
                            }
                        }
                        if ((uint32_t)lane < gn) Qr = draw_ray(cur + (uint32_t)lane, gsnap, false);
                        cur += gn;
                        // ring entries for the batch's processing slots (consecutive:
                        // lane 0's to lane gn-1's); an entry still holding an older
                        // slot hands its sums over first (its rays still in flight
                        // then tally straight to the tally slot)
                        const int ps0 = __shfl(Qr.ps, 0), ps1 = __shfl(Qr.ps, (int)gn - 1);
                        for (int sp = ps0; sp <= ps1; ++sp) {
                            const int e = rb + (sp & (kRing - 1));
                            const int held = s_rt[e];
                            if (held == sp) continue;
                            const unsigned long long mine = __ballot((uint32_t)lane < gn && Qr.ps == sp);
                            const int tgt = __shfl(Qr.six, __ffsll((long long)mine) - 1) + (int)gsnap * P.tally_stride;
                            if (lane == 0) {
                                if (held >= 0) ring_flush(e);
                                s_rt[e] = sp;
                                s_rg[e] = tgt;
                                s_rv[3 * e] = s_rv[3 * e + 1] = s_rv[3 * e + 2] = 0.0;
                            }
                        }
                      }
                        const bool s2 = idle && rank >= t1 && rank - t1 < gn;
                        const RayStart rd = pull_ray(Qr, (int)(s2 ? rank - t1 : 0u));
                        if (s2 && (GEN != 1 || rd.bg >= 0)) start_ray(rd, gsnap);
                        ghead = min((uint32_t)ni - t1, gn);
                    }
                }
            } else {
                // ---- hand out consecutive rays to the idle lanes, in lane order
                const uint32_t avail = cend - cur;
                uint32_t q = 0xFFFFFFFFu, qs = csnap;
                if (idle && rank < avail) q = cur + rank;
                if (avail < (uint32_t)ni && more) {
                    // the current super-chunk runs out: claim the next one (as above)
                    bool got = false;
                    uint32_t nstart = 0, nend = 0, nsn = 0;
                    while (xi < nparts) {
                        const uint32_t X = (xcd + xi) % nparts;
                        uint32_t c = 0;
                        if (lane == 0) c = atomicAdd(P.claim + 32u * X, 1u);
                        c = __builtin_amdgcn_readfirstlane(c);
                        if (c >= nclaim) {
                            ++xi;
                            continue;
                        }
                        const uint32_t sn = c / per_x;
                        const uint32_t lc = X * per_x + (c - sn * per_x);
                        if (lc >= cps) continue;   // past the snapshot's last chunk
                        nsn = sn;
                        nstart = lc * S;
                        nend = nstart + S < nrays ? nstart + S : nrays;
                        got = true;
                        break;
                    }
                    if (got) {
                        if (idle && rank >= avail && nstart + (rank - avail) < nend) {
                            q = nstart + (rank - avail);
                            qs = nsn;
                        }
                        const uint32_t used = (uint32_t)ni - avail;
                        cur = nstart + used < nend ? nstart + used : nend;
                        cend = nend;
                        csnap = nsn;
                    } else {
                        more = false;
                        cur = cend;
                    }
                } else {
                    cur += avail < (uint32_t)ni ? avail : (uint32_t)ni;
                }
                // ---- start the handed-out rays
                if (q != 0xFFFFFFFFu) {
                    qcur = q;
                    start_ray(draw_ray(q, qs, false), qs);
                }
            }
        }
        if (!alive) continue;

        // ---------------------------------------------------- one crossing
        float k0 = 0.f, k1 = 0.f, i0 = 0.f, i1 = 0.f, fr = 0.f;
        if (SPEC == kSpecTable) {
            if (TAB == kTabSliced) {
                // Tc is the cell's table coordinate within the slice (t_to_u):
                // node w of the slice and the lerp fraction
                int w = (int)Tc;
                w = w > nw - 2 ? nw - 2 : w;
                fr = Tc - (float)w;
                k0 = s_kq[krow + w];
                k1 = s_kq[krow + w + 1];
                i0 = s_ib[ibrow + w];
                i1 = s_ib[ibrow + w + 1];
            } else {
                float pos = fmaf(Tc, P.inv_dT, P.mT0);
                pos = fminf(fmaxf(pos, 0.0f), (float)(nT - 1));
                int t0 = (int)pos;
                t0 = t0 > nT - 2 ? nT - 2 : t0;
                fr = pos - (float)t0;
                k0 = kq[krow + t0];
                k1 = kq[krow + t0 + 1];
                i0 = ibt[ibrow + t0];
                i1 = ibt[ibrow + t0 + 1];
            }
        }
        if (REC) {
            const int lin = (cx - 1) + n0 * ((cy - 1) + n1 * (cz - 1));
            hash = (hash ^ (((uint32_t)lev << 28) ^ (uint32_t)lin)) * 16777619u;
        }
        const float d0 = (fx - px) * ix;
        const float d1 = (fy - py) * iy;
        const float d2 = (fz - pz) * iz;
        float ds = d0;
        ds = (ds > d1) ? d1 : ds;
        ds = (ds > d2) ? d2 : ds;
        px = (ds == d0) ? fx : fmaf(ds, sx, px);
        py = (ds == d1) ? fy : fmaf(ds, sy, py);
        pz = (ds == d2) ? fz : fmaf(ds, sz, pz);
        const int ax = (d0 < d1) ? ((d0 < d2) ? 0 : 2) : ((d1 < d2) ? 1 : 2);
        const bool m0 = ax == 0, m1 = ax == 1, m2 = ax == 2;
        const int fla = (fl >> ax) & 1;
        const int stp = fla ? -1 : 1;
        int cm = (m0 ? cx : (m1 ? cy : cz)) + stp;
        const int nm = m0 ? n0 : (m1 ? n1 : n2);
        const int fo = m0 ? fo0 : (m1 ? fo1 : fo2);
        // the domain boundary (cm = 0 or nm + 1): a periodic wrap, or a wall
        // hit that keeps the old cell (the ray ends there).  Rare per lane, so
        // a branch: most steps of a wave have no lane at the boundary.
        int wall = -1;
        if ((uint32_t)(cm - 1) >= (uint32_t)nm) {
            const int side = cm != 0;
            const int bcbit = ax * 2 + side;
            if ((P.bc_periodic >> bcbit) & 1) {
                const float pw = face_at(fo, side ? 0 : nm);
                cm = side ? 1 : nm;
                px = m0 ? pw : px;
                py = m1 ? pw : py;
                pz = m2 ? pw : pz;
            } else {
                wall = bcbit;
                cm -= stp;
            }
        }
        const float fnew = face_at(fo, cm - fla);
        cx = m0 ? cm : cx;
        cy = m1 ? cm : cy;
        cz = m2 ? cm : cz;
        fx = m0 ? fnew : fx;
        fy = m1 ? fnew : fy;
        fz = m2 ? fnew : fz;
        idx = off_at(oo0, cx) + off_at(oo1, cy) + off_at(oo2, cz);
        const float Tn = load_T(Tl, idx);
        float kap, ibj;
        if (SPEC == kSpecGreyT4) {
            const float t2 = Tc * Tc;
            kap = P.kappa_grey;
            ibj = t2 * t2;
        } else {
            kap = fmaf(fr, k1 - k0, k0);
            ibj = fmaf(fr, i1 - i0, i0);
        }
        ++cnt;
        ++rsteps;
        Tc = Tn;   // (read on every path, so the load is issued here, not sunk)
        // transport, branch-free on the common path: past the cut-off the
        // remaining energy is deposited here (tra), else the segment absorbs
        // tra - tra1; a wall absorbs what is left; the step cap ends the ray
        // (or moves it to the next coarser level)
        const float tau1 = fmaf(kap, ds, tau);
        const bool cut = tau1 >= P.tau_cut;
        const float tra1 = __expf(-tau1);
        G = fmaf(cut ? tra : tra - tra1, ibj, G);
        tau = tau1;
        tra = cut ? tra : tra1;
        const bool hitw = !cut && wall >= 0;
        if (hitw) G = fmaf(tra, s_ibw[wall * P.nB + band], G);
        const bool cap = !cut && !hitw && cnt >= maxs;
        if (cut || hitw) {
            alive = false; fin = true; endt = cut ? 0 : 1; Dfin = 1.0f;
        } else if (cap) {
            if (AMC && lev + 1 < P.nlev) {
                // adaptive mesh coarsening (code/mc_gpu.cu:193-207)
                ++lev;
                const int* L = s_lev + lev * kLevWords;
                n0 = L[0]; n1 = L[1]; n2 = L[2];
                fo0 = L[3]; fo1 = L[4]; fo2 = L[5]; maxs = L[6];
                oo0 = L[11]; oo1 = L[12]; oo2 = L[13];
                Tl = lev_T(L);
                if (GEN) Tl += (size_t)rsnap * (uint32_t)L[14];
                // ic = (ic + r - 1) / r with r = 2: every level halves each axis
                // (level g has n >> g cells)
                cx = (cx + 1) >> 1;
                cy = (cy + 1) >> 1;
                cz = (cz + 1) >> 1;
                fx = face_at(fo0, cx - (fl & 1));
                fy = face_at(fo1, cy - ((fl >> 1) & 1));
                fz = face_at(fo2, cz - ((fl >> 2) & 1));
                idx = off_at(oo0, cx) + off_at(oo1, cy) + off_at(oo2, cz);
                Tc = load_T(Tl, idx);
                cnt = 0;
            } else {
                alive = false; fin = true; endt = 2; Dfin = 1.0f - tra;
            }
        }
    }
    // a ray that ended after the last refill (nothing was left to hand out)
    if (!QUEUE && TMODE == 0) flush_wave(P, fin && rslot != aslot && aslot >= 0, aslot, aE, aA, aQ, lane);
    if (fin) {
        if (QUEUE)
            ring_tally(rw * Dfin, rwi * G);
        else
            tally_ray<TMODE>(P, rw * Dfin, rwi * G, rslot, qcur, aslot, aE, aA, aQ);
        if (REC && qcur < P.nrec) {
            P.rec_steps[qcur] = rsteps;
            P.rec_end[qcur] = endt;
            P.rec_hash[qcur] = hash;
            P.rec_D[qcur] = Dfin;
            P.rec_G[qcur] = G;
        }
        steps_lane += (uint32_t)rsteps;
    }
    if (!QUEUE && TMODE == 0) flush_wave(P, aslot >= 0, aslot, aE, aA, aQ, lane);
    // the ring's remaining sums
    if (QUEUE && lane < kRing && s_rt[rb + lane] >= 0) ring_flush(rb + lane);
    // one step-count atomic per wave
    unsigned long long st = steps_lane;
    for (int off = 32; off > 0; off >>= 1) st += __shfl_xor(st, off);
    if (lane == 0) atomicAdd(P.nsteps, st);
}

// ------------------------------------------------------------ trace_f64 ----
// FP64 traversal (precision 1, plain estimator): every ray in double, as the
// FP64 oracle computes it.  Direction from the descriptor's 24-bit uniforms,
// table lookups interpolated in FP64 from the FP32 tables, FP64 optical depth,
// transmissivity and gathered emission, bundle weights recomputed in FP64:
// w = Ibmax (k_i Ib_i) / (k_max Ib_max), E = w D, A = w G / Ib_i.  One thread
// per ray (grid-stride); each ray's (E, A) goes to pray64 for reduce_rays64.
__device__ __forceinline__ void lookup_f64(const TraceArgs& P, double Tj, int bg, int band, double* kap, double* ib)
{
    double pos = Tj * (double)P.inv_dT + (double)P.mT0;
    pos = pos < 0.0 ? 0.0 : pos;
    pos = pos > (double)(P.nT - 1) ? (double)(P.nT - 1) : pos;
    int t0 = (int)pos;
    t0 = t0 > P.nT - 2 ? P.nT - 2 : t0;
    const double fr = pos - (double)t0;
    const float* kr = P.kq + (size_t)bg * P.nT;
    const float* ir = P.Ib + (size_t)band * P.nT;
    *kap = (double)kr[t0] + fr * ((double)kr[t0 + 1] - (double)kr[t0]);
    *ib = (double)ir[t0] + fr * ((double)ir[t0 + 1] - (double)ir[t0]);
}

template <int SPEC, bool REC>
__global__ __launch_bounds__(256) void trace_f64(TraceArgs P, double2* pray64)
{
    extern __shared__ float lds[];
    int* s_lev = (int*)lds;
    float* s_face = lds + 6 * kLevWords;
    int* s_off = (int*)(s_face + P.nfaces);
    float* s_cent = (float*)(s_off + P.noffs);
    float* s_ibw = s_cent + P.ncent;
    for (int i = threadIdx.x; i < P.nlev; i += blockDim.x) {
        int* L = s_lev + i * kLevWords;
        L[0] = P.lev[i].n[0]; L[1] = P.lev[i].n[1]; L[2] = P.lev[i].n[2];
        L[3] = P.lev[i].foff[0]; L[4] = P.lev[i].foff[1]; L[5] = P.lev[i].foff[2];
        L[6] = P.lev[i].maxs;
        L[11] = P.lev[i].ooff[0]; L[12] = P.lev[i].ooff[1]; L[13] = P.lev[i].ooff[2];
    }
    for (int i = threadIdx.x; i < P.nfaces; i += blockDim.x) s_face[i] = P.faces[i];
    for (int i = threadIdx.x; i < P.noffs; i += blockDim.x) s_off[i] = P.offs[i];
    for (int i = threadIdx.x; i < P.ncent; i += blockDim.x) s_cent[i] = P.centres[i];
    for (int i = threadIdx.x; i < 6 * P.nB; i += blockDim.x) s_ibw[i] = P.Ibw[i];
    __syncthreads();
    const double Tm = (double)P.Tmax;
    unsigned long long steps_t = 0;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < P.nrays;
         q += (long long)gridDim.x * blockDim.x) {
        const uint32_t a0 = P.w0[q], a1 = P.w1[q];
        const int4 si = P.slotinfo[P.slot0 + fdiv((uint32_t)q, P.fd_R)];
        const int band = desc_band(a0, a1), g = desc_g(a1);
        const int bg = band * P.nQ + g;
        const double ut = ((double)(a0 & 0xFFFFFFu) + (double)P.u_off) * 5.9604644775390625e-08;
        const double up = ((double)(a1 & 0xFFFFFFu) + (double)P.u_off) * 5.9604644775390625e-08;
        double s[3], inv[3], p[3];
        {
            const int a = P.polar;
            const double cth = 1.0 - 2.0 * ut, st = 2.0 * sqrt(ut * (1.0 - ut));
            const double ph = 6.283185307179586476925 * up;
            const double c1 = st * cos(ph), c2 = st * sin(ph);
            s[0] = a == 0 ? cth : (a == 1 ? c2 : c1);
            s[1] = a == 0 ? c1 : (a == 1 ? cth : c2);
            s[2] = a == 0 ? c2 : (a == 1 ? c1 : cth);
        }
        int f[3], c[3];
        for (int a = 0; a < 3; ++a) {
            if (s[a] == 0.0) s[a] = 1e-10;
            f[a] = s[a] < 0.0;
            inv[a] = 1.0 / s[a];
        }
        c[0] = (si.w & 1023) + 1;
        c[1] = ((si.w >> 10) & 1023) + 1;
        c[2] = ((si.w >> 20) & 1023) + 1;
        for (int a = 0; a < 3; ++a) p[a] = (double)s_cent[P.coff[a] + c[a]];
        const double Ti = (double)P.T[0][(size_t)P.snap * P.lev[0].tstride + si.z];
        double ki, ii, kmx, imx;
        if (SPEC == kSpecGreyT4) {
            ki = kmx = (double)P.kappa_grey;
            ii = Ti * Ti * Ti * Ti;
            imx = Tm * Tm * Tm * Tm;
        } else {
            lookup_f64(P, Ti, bg, band, &ki, &ii);
            lookup_f64(P, Tm, bg, band, &kmx, &imx);
        }
        const double w = P.Ibmax_d * (ki * ii) / (kmx * imx);
        double tau = 0.0, tra = 1.0, G = 0.0, D = 1.0;
        const double tau_cut = (double)P.tau_cut;
        int steps = 0, endt = 2;
        uint32_t hash = 2166136261u;
        bool done = false;
        for (int lev = 0; lev < P.nlev && !done; ++lev) {
            const int* L = s_lev + lev * kLevWords;
            const int n[3] = {L[0], L[1], L[2]};
            const int fo[3] = {L[3], L[4], L[5]};
            const int oo[3] = {L[11], L[12], L[13]};
            const float* T = P.T[lev] + (size_t)P.snap * P.lev[lev].tstride;
            if (lev > 0) {
                const int* Lp = s_lev + (lev - 1) * kLevWords;
                for (int a = 0; a < 3; ++a) {
                    const int r = Lp[a] / n[a];
                    c[a] = (c[a] + r - 1) / r;
                }
            }
            int cnt = 0;
            while (cnt < L[6]) {
                double fx[3], d[3];
                for (int a = 0; a < 3; ++a) {
                    fx[a] = (double)s_face[fo[a] + c[a] - f[a]];
                    d[a] = (fx[a] - p[a]) * inv[a];
                }
                double ds = d[0];
                ds = ds > d[1] ? d[1] : ds;
                ds = ds > d[2] ? d[2] : ds;
                for (int a = 0; a < 3; ++a) p[a] = (ds == d[a]) ? fx[a] : p[a] + ds * s[a];
                if (REC) {
                    const int lin = (c[0] - 1) + n[0] * ((c[1] - 1) + n[1] * (c[2] - 1));
                    hash = (hash ^ (((uint32_t)lev << 28) ^ (uint32_t)lin)) * 16777619u;
                }
                const int idx = s_off[oo[0] + c[0] - 1] + s_off[oo[1] + c[1] - 1] + s_off[oo[2] + c[2] - 1];
                const double Tj = (double)T[idx];
                double kap, ibj;
                if (SPEC == kSpecGreyT4) {
                    kap = (double)P.kappa_grey;
                    ibj = Tj * Tj * Tj * Tj;
                } else {
                    lookup_f64(P, Tj, bg, band, &kap, &ibj);
                }
                ++steps;
                ++cnt;
                const double tau1 = tau + kap * ds;
                if (tau1 >= tau_cut) {
                    G += tra * ibj;
                    endt = 0;
                    done = true;
                    break;
                }
                const double tra1 = exp(-tau1);
                G += (tra - tra1) * ibj;
                tau = tau1;
                tra = tra1;
                const int ax = (d[0] < d[1]) ? ((d[0] < d[2]) ? 0 : 2) : ((d[1] < d[2]) ? 1 : 2);
                c[ax] += f[ax] ? -1 : 1;
                if (c[ax] == 0 || c[ax] == n[ax] + 1) {
                    const int side = c[ax] != 0;
                    if ((P.bc_periodic >> (ax * 2 + side)) & 1) {
                        c[ax] = side ? 1 : n[ax];
                        p[ax] = (double)s_face[fo[ax] + (side ? 0 : n[ax])];
                    } else {
                        c[ax] -= f[ax] ? -1 : 1;
                        G += tra * (double)s_ibw[(ax * 2 + side) * P.nB + band];
                        endt = 1;
                        done = true;
                        break;
                    }
                }
            }
        }
        if (!done) D = 1.0 - tra;
        pray64[q] = make_double2(w * D, w * G / ii);
        steps_t += (unsigned long long)steps;
        if (REC && q < P.nrec) {
            P.rec_steps[q] = steps;
            P.rec_end[q] = endt;
            P.rec_hash[q] = hash;
            P.rec_D[q] = (float)D;
            P.rec_G[q] = (float)G;
        }
    }
    for (int off = 32; off > 0; off >>= 1) steps_t += __shfl_xor(steps_t, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(P.nsteps, steps_t);
}
template __global__ void trace_f64<kSpecTable, false>(TraceArgs, double2*);
template __global__ void trace_f64<kSpecTable, true>(TraceArgs, double2*);
template __global__ void trace_f64<kSpecGreyT4, false>(TraceArgs, double2*);
template __global__ void trace_f64<kSpecGreyT4, true>(TraceArgs, double2*);

// Per-slot sums of the FP64 traversal's per-ray (E, A), in ray order, in the
// configured tally arithmetic (FP64, Kahan-compensated FP32 or plain FP32).
__global__ void reduce_rays64(const double2* pray, int tally, long long nslots, int R, long long slot0,
                              const int4* slotinfo, double* E, double* A, double* Q2)
{
    long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const long long ts = slotinfo[slot0 + s].x;
    const double2* p = pray + s * R;
    if (tally == DNSMC_TALLY_KAHAN_F32) {
        float e = 0, a = 0, q = 0, ce = 0, ca = 0, cq = 0;
        for (int r = 0; r < R; ++r) {
            const float pe = (float)p[r].x, pa = (float)p[r].y;
            float x = pe - pa;
            float y = pe - ce, t = e + y; ce = (t - e) - y; e = t;
            y = pa - ca; t = a + y; ca = (t - a) - y; a = t;
            y = x * x - cq; t = q + y; cq = (t - q) - y; q = t;
        }
        E[ts] += e; A[ts] += a; Q2[ts] += q;
    } else if (tally == DNSMC_TALLY_F32) {
        float e = 0, a = 0, q = 0;
        for (int r = 0; r < R; ++r) {
            const float pe = (float)p[r].x, pa = (float)p[r].y;
            float x = pe - pa;
            e += pe; a += pa; q += x * x;
        }
        E[ts] += e; A[ts] += a; Q2[ts] += q;
    } else {
        double e = 0, a = 0, q = 0;
        for (int r = 0; r < R; ++r) {
            const double x = p[r].x - p[r].y;
            e += p[r].x; a += p[r].y; q += x * x;
        }
        E[ts] += e; A[ts] += a; Q2[ts] += q;
    }
}

#define DN_TRACE_INST(S, I, T, R, A, G) template __global__ void trace<S, I, T, R, A, G>(TraceArgs);
#define DN_TRACE_INST2(S, I, T, R, G) DN_TRACE_INST(S, I, T, R, false, G) DN_TRACE_INST(S, I, T, R, true, G)
#define DN_TRACE_INST3(S, I) DN_TRACE_INST2(S, I, 0, false, 0) DN_TRACE_INST2(S, I, 0, true, 0) \
    DN_TRACE_INST2(S, I, 1, false, 0) DN_TRACE_INST2(S, I, 0, false, 1) DN_TRACE_INST2(S, I, 0, false, 2)
DN_TRACE_INST3(kSpecTable, kTabSliced)
DN_TRACE_INST3(kSpecTable, kTabIbLds)
DN_TRACE_INST3(kSpecTable, kTabGlobal)
DN_TRACE_INST3(kSpecGreyT4, kTabGlobal)

// Range of a bricked level field as order-preserving keys of its FP32 values
// (lohi = {min key, max key}; the host starts them at {~0u, 0}).
__global__ void t_range(const float* T, int n0, int n1, int n2, int nbx, int nby, unsigned* lohi)
{
    const long long n = (long long)n0 * n1 * n2;
    unsigned lo = 0xFFFFFFFFu, hi = 0u;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % n0), j = (int)((c / n0) % n1), k = (int)(c / ((long long)n0 * n1));
        const unsigned u = __float_as_uint(T[brick_index(i, j, k, nbx, nby)]);
        const unsigned key = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
        lo = key < lo ? key : lo;
        hi = key > hi ? key : hi;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned a = __shfl_xor(lo, off), b = __shfl_xor(hi, off);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&lohi[0], lo);
        atomicMax(&lohi[1], hi);
    }
}

// ------------------------------------------------------ reduce_rays ----
// Deterministic tallies: processing slot s owns rays [s*R, (s+1)*R) of the
// batch; their (E, A) are summed in ray order in FP64, Kahan-compensated FP32
// or plain FP32, and added to the slot's tally.
__global__ void reduce_rays(const float2* pray, int tally, long long nslots, int R, long long slot0,
                            const int4* slotinfo, double* E, double* A, double* Q2)
{
    long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const long long ts = slotinfo[slot0 + s].x;
    const float2* p = pray + s * R;
    if (tally == DNSMC_TALLY_F64) {
        double e = 0, a = 0, q = 0;
        for (int r = 0; r < R; ++r) {
            double de = p[r].x, da = p[r].y, x = de - da;
            e += de; a += da; q += x * x;
        }
        E[ts] += e; A[ts] += a; Q2[ts] += q;
    } else if (tally == DNSMC_TALLY_KAHAN_F32) {
        float e = 0, a = 0, q = 0, ce = 0, ca = 0, cq = 0;
        for (int r = 0; r < R; ++r) {
            float x = p[r].x - p[r].y;
            float y = p[r].x - ce, t = e + y; ce = (t - e) - y; e = t;
            y = p[r].y - ca; t = a + y; ca = (t - a) - y; a = t;
            y = x * x - cq; t = q + y; cq = (t - q) - y; q = t;
        }
        E[ts] += e; A[ts] += a; Q2[ts] += q;
    } else {
        float e = 0, a = 0, q = 0;
        for (int r = 0; r < R; ++r) {
            float x = p[r].x - p[r].y;
            e += p[r].x; a += p[r].y; q += x * x;
        }
        E[ts] += e; A[ts] += a; Q2[ts] += q;
    }
}

// ------------------------------------------------------------ finalize ----
// compat: nvar block estimates De_v = E_v - A_v; Q = mean; reference variance
// sum (De_v - Q)^2 / ((n - 0.5)(n - 1)) (code/mc_gpu.cu:217-222), sdev = sqrt.
// plain: Q = E - A, var(Q) = (R sum x^2 - Q^2)/(R - 1) from per-ray x = e - a.
// Eo and Ao (the mean emitted and absorbed parts) may be null.
__global__ void finalize(const double* E, const double* A, const double* Q2, long long nemit, int spe, int R,
                         int compat, double* Q, double* Eo, double* Ao, double* sd)
{
    long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= nemit) return;
    if (compat) {
        double m = 0, me = 0, ma = 0;
        for (int v = 0; v < spe; ++v) {
            m += E[e * spe + v] - A[e * spe + v];
            me += E[e * spe + v];
            ma += A[e * spe + v];
        }
        m /= spe; me /= spe; ma /= spe;
        double var = 0;
        for (int v = 0; v < spe; ++v) {
            double d = (E[e * spe + v] - A[e * spe + v]) - m;
            var += 1.0 / (spe - 0.5) * 1.0 / (spe - 1) * d * d;
        }
        Q[e] = m; sd[e] = sqrt(var);
        if (Eo) Eo[e] = me;
        if (Ao) Ao[e] = ma;
    } else {
        double q = E[e] - A[e];
        double v = R > 1 ? ((double)R * Q2[e] - q * q) / (double)(R - 1) : 0.0;
        Q[e] = q; sd[e] = sqrt(v > 0 ? v : 0.0);
        if (Eo) Eo[e] = E[e];
        if (Ao) Ao[e] = A[e];
    }
}

// --------------------------------------------------------- spline pass ----
// One thread per grid line: natural-spline second derivatives through the
// fine values (forward sweep + back substitution, FP64), then evaluation at
// the coarse centres.  The grid-only factors come precomputed from the host.
// The x pass reads the bricked fine field; the z pass writes the bricked
// layout of the coarse level.
__global__ void spline_pass(SplineArgs s)
{
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= s.nlines) return;
    double* u = s.work + (size_t)l * s.nf;
    const int nf = s.nf;
    const float* in = s.in + (l % s.A) * s.sa_in + (l / s.A) * s.sb_in;
    const int bj = s.in_bricked ? l % s.fn1 : 0, bk = s.in_bricked ? l / s.fn1 : 0;
    auto rd = [&](int i) -> double {
        if (s.in_bricked) return (double)s.in[brick_index(i, bj, bk, s.fnbx, s.fnby)];
        return (double)in[(long long)i * s.st_in];
    };
    u[0] = 0.0;
    double ym = rd(0), y0 = rd(1);
    for (int i = 1; i < nf - 1; ++i) {
        double yp = rd(i + 1);
        double dd = (yp - y0) * s.ihh[i] - (y0 - ym) * s.ihl[i];
        u[i] = (s.c6[i] * dd - s.sig[i] * u[i - 1]) * s.invp[i];
        ym = y0;
        y0 = yp;
    }
    u[nf - 1] = 0.0;
    for (int k = nf - 2; k >= 1; --k) u[k] = s.y2f[k] * u[k + 1] + u[k];
    u[0] = 0.0;
    for (int q = 0; q < s.nc; ++q) {
        int k = s.klo[q];
        double v = s.ca[q] * rd(k) + s.cb[q] * rd(k + 1) + s.cA[q] * u[k] + s.cB[q] * u[k + 1];
        long long o;
        if (s.brick_out) o = brick_index(l % s.m0, l / s.m0, q, s.nbx, s.nby);   // line l = i + m0 j
        else o = (l % s.A) * s.sa_out + (l / s.A) * s.sb_out + (long long)q * s.st_out;
        s.out[o] = (float)v;
    }
}

// Plain streaming copy (16 B per thread and iteration); the reference point
// for calibrating HBM byte counters.
__global__ void copy_f32x4(const float4* __restrict__ src, float4* __restrict__ dst, long long n4)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Streaming kernels of the achievable-bandwidth ceiling: every lane moves U
// 16-byte vectors per iteration (grid-stride, loads issued before stores),
// optionally with non-temporal hints.  stream_read sums what it reads and
// writes one float per thread only when the sum is the impossible value -1,
// so the loads are not removed and nothing else is written.
typedef float v4f __attribute__((ext_vector_type(4)));

template <int U, bool NT>
__global__ __launch_bounds__(256) void stream_copy(const float4* __restrict__ src4, float4* __restrict__ dst4,
                                                   long long n4)
{
    const v4f* src = (const v4f*)src4;
    v4f* dst = (v4f*)dst4;
    const long long st = (long long)gridDim.x * blockDim.x;
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; i + (U - 1) * st < n4; i += U * st) {
        v4f v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = NT ? __builtin_nontemporal_load(src + i + u * st) : src[i + u * st];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (NT) __builtin_nontemporal_store(v[u], dst + i + u * st);
            else dst[i + u * st] = v[u];
        }
    }
    for (; i < n4; i += st) dst[i] = src[i];
}

template <int U, bool NT>
__global__ __launch_bounds__(256) void stream_read(const float4* __restrict__ src4, long long n4, float* sink)
{
    const v4f* src = (const v4f*)src4;
    const long long st = (long long)gridDim.x * blockDim.x;
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    float acc = 0.f;
    for (; i + (U - 1) * st < n4; i += U * st) {
        v4f v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = NT ? __builtin_nontemporal_load(src + i + u * st) : src[i + u * st];
#pragma unroll
        for (int u = 0; u < U; ++u) acc += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    }
    for (; i < n4; i += st) acc += src[i].x;
    if (acc == -1.0f) sink[blockIdx.x * (long long)blockDim.x + threadIdx.x] = acc;
}

template __global__ void stream_copy<1, false>(const float4*, float4*, long long);
template __global__ void stream_copy<4, false>(const float4*, float4*, long long);
template __global__ void stream_copy<4, true>(const float4*, float4*, long long);
template __global__ void stream_read<1, false>(const float4*, long long, float*);
template __global__ void stream_read<4, false>(const float4*, long long, float*);
template __global__ void stream_read<4, true>(const float4*, long long, float*);

// ------------------------------------------------- DNS coupling slabs ----
// fix_temp (code/fix_var.f:49-73) for one rank's slab, written straight into
// the bricked device field instead of summed into a full-size host buffer and
// all-reduced: T(i,j,k) = cnew * Tc / Tplus + Tc in FP64 (left to right, as
// the Fortran expression), rounded to FP32 as the real(4) TMC is.  cnew is the
// rank's array cnew(0:n0+1, nj, nk), i fastest; global (i, j, k) of the slab
// (1-based, i0..i0+ni-1 etc.) reads cnew(i - i0 + 1, j - j0 + 1, k - k0 + 1).
// Ghost indices (0 and n+1) are not part of the bricked field and are skipped.
__global__ void slab_to_field(const double* cnew, int n0, int n1, int n2, int i0, int j0, int k0, int ni, int nj,
                              int nk, double Tc, double Tplus, float* T, int nbx, int nby)
{
    const long long n = (long long)ni * nj * nk;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        const int a = (int)(c % ni), b = (int)((c / ni) % nj), d = (int)(c / ((long long)ni * nj));
        const int i = i0 + a, j = j0 + b, k = k0 + d;
        if (i < 1 || i > n0 || j < 1 || j > n1 || k < 1 || k > n2) continue;
        const double v = cnew[(long long)(a + 1) + (long long)(n0 + 2) * (b + (long long)nj * d)];
        T[brick_index(i - 1, j - 1, k - 1, nbx, nby)] = (float)(v * Tc / Tplus + Tc);
    }
}

// get_results_ (code/mc_gpu.cu:759-776): Q and its standard error of the
// emitters in the j slab [jlo, jlo + nj) (0-based) into real(4) arrays
// (0:n0+1, 0:nj+1, 0:n2+1), i fastest; other entries keep their values.
__global__ void results_to_fortran(const int* emit, long long nemit, const double* Q, const double* sd, int n0,
                                   int n1, int jlo, int nj, int n2, float* resf, float* varf)
{
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= nemit) return;
    const int cell = emit[e];
    const int i = cell % n0, j = (cell / n0) % n1, k = cell / (n0 * n1);
    const int jl = j - jlo;
    if (jl < 0 || jl >= nj || k >= n2) return;
    const long long o = (long long)(i + 1) + (long long)(n0 + 2) * ((jl + 1) + (long long)(nj + 2) * (k + 1));
    resf[o] = (float)Q[e];
    varf[o] = (float)sd[e];
}

// fix_rad (code/fix_var.f:75-121) for one rank's k range, from the device
// results: divQ(i, jl, k - k0 + 1) = Q / stefan / Tc^4 and vcore likewise from
// the standard error, for the emitters in the solver's j slab [jlo, jlo + nj)
// (0-based) and k in [k0, k0 + nk) (0-based); out arrays (0:n0+1, nj, nk).
__global__ void field_to_slab(const int* emit, long long nemit, const double* Q, const double* sd, int n0, int n1,
                              int jlo, int nj, int k0, int nk, double stefan, double tc4, double* divq, double* vcore)
{
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= nemit) return;
    const int cell = emit[e];
    const int i = cell % n0, j = (cell / n0) % n1, k = cell / (n0 * n1);
    const int jl = j - jlo, kl = k - k0;
    if (jl < 0 || jl >= nj || kl < 0 || kl >= nk) return;
    const long long o = (long long)(i + 1) + (long long)(n0 + 2) * (jl + (long long)nj * kl);
    divq[o] = Q[e] / stefan / tc4;
    vcore[o] = sd[e] / stefan / tc4;
}

// A temperature field in the reference's layout, T(0:n0+1, 0:n1+1, 0:n2+1) with
// i fastest (code/param.h:104-105), into the 4x4x4-bricked layout: interior
// cells only (the ghost layers are not part of the bricked field).
__global__ void ghost_to_bricks(const float* Tg, int n0, int n1, int n2, float* T, int nbx, int nby)
{
    const long long n = (long long)n0 * n1 * n2;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % n0), j = (int)((c / n0) % n1), k = (int)(c / ((long long)n0 * n1));
        T[brick_index(i, j, k, nbx, nby)] = Tg[(long long)(i + 1) + (long long)(n0 + 2) * ((j + 1) + (long long)(n1 + 2) * (k + 1))];
    }
}

// Table-coordinate field of the sliced-table traversal: u = clamp(T/dT - T0/dT,
// 0, nT-1) - lo, the position of each cell's temperature among the nodes held
// in LDS.  The traversal's lerp from u is bit-identical to the one from T (the
// subtraction of the integer lo from the clamped position is exact).
__global__ void t_to_u(const float* T, float* U, long long n, float inv_dT, float mT0, float top, float lo)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float pos = fmaf(T[i], inv_dT, mT0);
        pos = fminf(fmaxf(pos, 0.0f), top);
        U[i] = pos - lo;
    }
}

__global__ void fill_u32(unsigned* p, long long n, unsigned v)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_f64(double* p, long long n, double v)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace dn
