// Kernel argument blocks shared by kernels.hip and solver.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmath.h"

namespace dn {

constexpr int kLevWords = 12;   // per-level parameter words staged in LDS

// Per-level geometry, staged into LDS as ints (floats for nothing).
struct LevParams {
    int n[3];
    int foff[3];      // offsets of the face arrays in the packed face buffer
    int maxs;
    int pad;
    long long tstride;   // cells of this level per snapshot
};

enum { kSpecTable = 0, kSpecGreyT4 = 1 };

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
    const float2* kq2;          // kappa rows {k, dk}
    const float* Ib;            // Planck table [b*nT + t]
    const float* Ibw;           // wall Planck [(axis*2+side)*nB + b]
    const float* faces;         // all levels packed
    int nfaces;
    const float* centres;       // level 0, x|y|z packed, ghost-inclusive
    int coff[3];
    int ncent;
    int ib_in_lds;
    // rays (SoA)
    const uint32_t* w0;
    const uint32_t* w1;
    const float* w;
    const float* winv;
    long long nrays;
    int chunk;
    int rays_per_slot;
    int slots_per_emitter;
    FastDiv fd_R, fd_spe, fd_nx, fd_ny;   // divisions by rays/slot, slots/emitter, n0, n1
    int refill_k;                         // start rays once this many lanes of a wave are idle
    const int* emitters;        // emitter -> level-0 cell (i fastest, 0-based)
    const int* order;           // processing position -> emitter (3-D brick order)
    long long slot0;            // first slot of this batch
    int snap;
    // tallies
    int tally;                  // DNSMC_TALLY_*
    double* E;
    double* A;
    double* Q2;
    double* part;               // per-chunk partials (deterministic modes), 3 per chunk
    float* partf;               // FP32 per-chunk partials (Kahan / naive)
    unsigned long long* nsteps;
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
};

struct GenArgs {
    // geometry of level 0 (emitter temperature)
    const float* T0;
    long long tstride0;
    int nB, nQ, nT;
    float inv_dT, mT0;
    float kappa_grey;
    int spec;
    const float2* kq2;
    const float* Ib;
    const float* cdf_band;
    const float* cdf_g;
    const float* inv_emax;
    float Ibmax;
    int g_search;
    // batch
    const int* emitters;
    const int* order;
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
    // learned proposal
    const void* mlp;
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
};

}  // namespace dn
