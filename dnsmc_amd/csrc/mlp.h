// Learned proposal network: shapes, parameter packing and kernel argument
// blocks shared by mlp.hip and solver.cpp.
//
// Network (per emitting cell): X[6] -> H1 = tanh(X W1 + b1)[64]
//   -> H2 = tanh(H1 W2 + b2)[64] -> Z = H2 W3 + b3, Z = [64 direction-bin
//   logits | nbin (band, g) logits | padding to a multiple of 64].
// Weights are BF16 (row-major [in][out]); biases, activations' accumulation
// and the optimiser state are FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dn {

constexpr int kMlpIn = 6;
constexpr int kMlpHid = 64;
constexpr int kMlpDir = 64;   // 8 cos(theta) bins x 8 phi bins (equal area)

__host__ __device__ inline int mlp_noutp(int nbin) { return ((kMlpDir + nbin + 63) / 64) * 64; }

// Offsets of the parameters in the packed FP32 vector theta.
struct MlpLayout {
    int noutp, oW1, ob1, oW2, ob2, oW3, ob3, P;
};
__host__ __device__ inline MlpLayout mlp_layout(int nbin)
{
    MlpLayout L;
    L.noutp = mlp_noutp(nbin);
    L.oW1 = 0;
    L.ob1 = L.oW1 + kMlpIn * kMlpHid;
    L.oW2 = L.ob1 + kMlpHid;
    L.ob2 = L.oW2 + kMlpHid * kMlpHid;
    L.oW3 = L.ob2 + kMlpHid;
    L.ob3 = L.oW3 + kMlpHid * L.noutp;
    L.P = L.ob3 + L.noutp;
    return L;
}

struct MlpWeights {
    const uint16_t* W1;   // [6][64] bf16
    const float* b1;      // [64]
    const uint16_t* W2;   // [64][64] bf16
    const float* b2;      // [64]
    const uint16_t* W3;   // [64][noutp] bf16
    const float* b3;      // [noutp]
    int nbin, noutp;
    // the same W2 and W3 with the input index contiguous ([64][64] and
    // [noutp][64]), for the proposal kernel's MFMA operand tiles (set by the
    // solver; the training kernels do not use them)
    const uint16_t* W2T;
    const uint16_t* W3T;
};

// Row length (elements) of the proposal kernel's LDS tiles: 68 BF16 = 34
// banks, so the 16 rows one MFMA operand fragment reads start in 16
// different even banks, and 68 floats put rows 4 apart 16 banks apart.
constexpr int kPropStride = 68;
__host__ __device__ inline size_t mlp_proposal_lds()
{
    return sizeof(float) * (64 * kMlpIn + 64 * kPropStride) + 2 * 64 * kPropStride * sizeof(uint16_t);
}

struct MlpTrainArgs {
    const float* X;       // [n][6]
    const int* td;        // direction bin of the sample
    const int* tb;        // (band, g) bin of the sample
    const float* wt;      // sample weight (|e - a| of the bundle)
    int n;
    const float* wsum;    // device scalar: sum of wt over the n samples
    MlpWeights w;
    float* part;          // [gridDim.x][P] per-block gradient partials
    float* loss_part;     // [gridDim.x]
};

// Per-cell features of the proposal (shared by ray generation and training).
struct FeatArgs {
    const float* T0;       // bricked level-0 field of the snapshot
    int n[3];
    int nbx, nby;
    const float* centres;  // level-0 centres, x|y|z packed, ghost-inclusive
    int coff[3];
    int bc_periodic;       // bit (axis*2+side)
    int polar;             // wall-normal axis (the polar axis of the directions)
    float Lpolar;          // domain length along it
    float inv_Tmax;
    float Tw_rel;          // mean wall temperature on the polar axis / Tmax
};

}  // namespace dn
