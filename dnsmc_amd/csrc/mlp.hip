// Learned importance-sampling proposal (gfx950): BF16 MFMA forward/backward
// training kernels, the per-cell features, the training samples taken from a
// plain run, and the ray generator that samples from the network.
//
// No counterpart in the reference: its proposal is the fixed Tmax
// distribution (code/mc_gpu.cu:584-590, :830-896).  Every learned ray carries
// the factor p/q (plain over learned probability of its direction bin and
// (band, g) bin), so the estimator's mean is the plain estimator's mean; a
// defensive mixture q = (1 - eps) q_net + eps p bounds p/q by 1/eps.
//
// Matrix products use v_mfma_f32_16x16x16_bf16 (FP32 accumulate).  Operand
// layout of one 16x16x16 step for lane l (0..63), element i (0..3):
//   A(row = l % 16, k = 4 (l / 16) + i),  B(k = 4 (l / 16) + i, col = l % 16),
//   D(row = 4 (l / 16) + i, col = l % 16).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmath.h"
#include "kernels.h"
#include "mlp.h"

namespace dn {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));


// Fragment of a 16-row panel starting at row r0, k-block k0: element (r, k)
// of the operand is M[r * rs + k * cs].
__device__ __forceinline__ s4v frag(const uint16_t* M, int rs, int cs, int r0, int k0, int lane)
{
    const int r = r0 + (lane & 15), k = k0 + 4 * (lane >> 4);
    const uint16_t* p = M + r * rs + k * cs;
    s4v v;
    if (cs == 1) {   // four consecutive BF16 (8-byte aligned): one LDS read
        const uint2 u = *(const uint2*)p;
        v.x = (short)(u.x & 0xFFFFu);
        v.y = (short)(u.x >> 16);
        v.z = (short)(u.y & 0xFFFFu);
        v.w = (short)(u.y >> 16);
        return v;
    }
    v.x = (short)p[0];
    v.y = (short)p[cs];
    v.z = (short)p[2 * cs];
    v.w = (short)p[3 * cs];
    return v;
}

// 64x64 output per 256-thread block (each wave a 32x32 quadrant of 2x2
// 16x16 tiles): acc += A[64 x K] B[K x 64], A(r, k) = A[r*ars + k*acs],
// B(k, n) = B[k*brs + n*bcs].  acc[ti][tj][i] holds element
// (row = qr + 16 ti + 4 (lane/16) + i, col = qc + 16 tj + lane % 16).
__device__ __forceinline__ void gemm64(const uint16_t* A, int ars, int acs, const uint16_t* B, int brs, int bcs, int K,
                                       f4v acc[2][2])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int qr = (w >> 1) * 32, qc = (w & 1) * 32;
    for (int k0 = 0; k0 < K; k0 += 16) {
        const s4v a0 = frag(A, ars, acs, qr, k0, lane), a1 = frag(A, ars, acs, qr + 16, k0, lane);
        const s4v b0 = frag(B, bcs, brs, qc, k0, lane), b1 = frag(B, bcs, brs, qc + 16, k0, lane);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b1, acc[1][1], 0, 0, 0);
    }
}

__device__ __forceinline__ void acc_zero(f4v acc[2][2])
{
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = f4v{0.f, 0.f, 0.f, 0.f};
}

// Visit every accumulator element with its (row, col) in the 64x64 tile.
template <class F>
__device__ __forceinline__ void acc_each(const f4v acc[2][2], F f)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int qr = (w >> 1) * 32, qc = (w & 1) * 32;
    for (int ti = 0; ti < 2; ++ti)
        for (int tj = 0; tj < 2; ++tj)
            for (int i = 0; i < 4; ++i) f(qr + 16 * ti + 4 * (lane >> 4) + i, qc + 16 * tj + (lane & 15), acc[ti][tj][i]);
}

// H1 = tanh(X W1 + b1) for rows < nrow (zero rows after), as FP32 and BF16.
__device__ __forceinline__ void layer1(const float* sX, const MlpWeights& w, int nrow, float* H1f, uint16_t* H1b,
                                       int hs = kMlpHid)
{
    for (int e = threadIdx.x; e < 64 * kMlpHid; e += blockDim.x) {
        const int s = e / kMlpHid, j = e - s * kMlpHid;
        float h = 0.f;
        if (s < nrow) {
            float acc = w.b1[j];
            for (int i = 0; i < kMlpIn; ++i) acc = fmaf(sX[s * kMlpIn + i], bf2f(w.W1[i * kMlpHid + j]), acc);
            h = tanhf(acc);
        }
        if (H1f) H1f[e] = h;
        H1b[s * hs + j] = f2bf(h);
    }
}

// Stage column tile jt of W3 ([64][noutp] BF16, row-major) into dst[64][64]:
// 64 rows of 128 contiguous bytes, 16 B per load.
__device__ __forceinline__ void stage_w3_tile(const uint16_t* W3, int noutp, int jt, uint16_t* dst)
{
    const uint4* src = (const uint4*)W3;
    uint4* d = (uint4*)dst;
    for (int e = threadIdx.x; e < 512; e += blockDim.x) {
        const int h = e >> 3, c8 = e & 7;
        d[e] = src[((size_t)h * noutp + jt * 64) / 8 + c8];
    }
}

// Barrier for the block's LDS traffic only: each wave's LDS operations are
// complete before it, its global stores (the rows, read by no other thread
// of this kernel) stay in flight.
__device__ __forceinline__ void lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Tile jt of W3^T (rows jt*64 .. jt*64 + 63 of the [noutp][64] copy, 128 B
// each) for a 256-thread block in two steps: its two 16-byte pieces per
// thread into registers (issued one tile ahead), then into the padded LDS
// rows (8-byte writes: the padded rows are 8-byte aligned).
__device__ __forceinline__ void load_w3_tile(const uint16_t* W3T, int jt, uint4& r0, uint4& r1)
{
    const uint4* src = (const uint4*)(W3T + (size_t)jt * 64 * 64);
    r0 = src[threadIdx.x];
    r1 = src[threadIdx.x + 256];
}
__device__ __forceinline__ void put_w3_tile(const uint4& r0, const uint4& r1, uint16_t* dst)
{
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int e = (int)threadIdx.x + 256 * i, n = e >> 3, k8 = e & 7;
        const uint4 r = i ? r1 : r0;
        uint2* d = (uint2*)(dst + n * kPropStride + k8 * 8);
        d[0] = make_uint2(r.x, r.y);
        d[1] = make_uint2(r.z, r.w);
    }
}

// ------------------------------------------------------------- training ----
// One block per 64 samples: forward, weighted cross-entropy of both heads,
// backward; the block's gradient of (sum_s wt_s CE_s) / sum wt goes to its row
// of `part`.  LDS (dynamic): see the carve-up below (about 91 KB).
__global__ __launch_bounds__(256) void mlp_train_tile(MlpTrainArgs a)
{
    extern __shared__ float lds[];
    float* sX = lds;                                  // 64*6
    float* sH1f = sX + 64 * kMlpIn;                   // 64*64
    float* sH2f = sH1f + 4096;                        // 64*64 (later dA1)
    float* sZ = sH2f + 4096;                          // 64*64 logits / dZ / dA2
    float* sw = sZ + 4096;                            // 64 sample scale
    float* sm_d = sw + 64;                            // log-sum-exp per row, direction head
    float* sm_b = sm_d + 64;                          // log-sum-exp per row, (band, g) head
    float* sloss = sm_b + 64;                         // 64
    int* std_ = (int*)(sloss + 64);                   // 64
    int* stb_ = std_ + 64;                            // 64
    uint16_t* sH1b = (uint16_t*)(stb_ + 64);          // 64*64
    uint16_t* sH2b = sH1b + 4096;
    uint16_t* sW2b = sH2b + 4096;
    uint16_t* sW3b = sW2b + 4096;
    uint16_t* sdZb = sW3b + 4096;

    const MlpLayout L = mlp_layout(a.w.nbin);
    const int noutp = a.w.noutp, nbin = a.w.nbin;
    const int tid = threadIdx.x;
    const long long r0 = (long long)blockIdx.x * 64;
    const int nrow = (int)((a.n - r0) < 64 ? (a.n - r0) : 64);
    float* part = a.part + (size_t)blockIdx.x * L.P;
    const float inv_w = 1.0f / *a.wsum;

    for (int e = tid; e < 64 * kMlpIn; e += blockDim.x) {
        const int s = e / kMlpIn;
        sX[e] = s < nrow ? a.X[(r0 + s) * kMlpIn + (e - s * kMlpIn)] : 0.f;
    }
    if (tid < 64) {
        const bool v = tid < nrow;
        sw[tid] = v ? a.wt[r0 + tid] * inv_w : 0.f;
        std_[tid] = v ? a.td[r0 + tid] : 0;
        stb_[tid] = v ? a.tb[r0 + tid] : 0;
    }
    for (int e = tid; e < 4096; e += blockDim.x) sW2b[e] = a.w.W2[e];
    __syncthreads();
    layer1(sX, a.w, nrow, sH1f, sH1b);
    __syncthreads();
    {
        f4v acc[2][2];
        acc_zero(acc);
        gemm64(sH1b, 64, 1, sW2b, 64, 1, 64, acc);
        acc_each(acc, [&](int r, int c, float v) {
            const float h = r < nrow ? tanhf(v + a.w.b2[c]) : 0.f;
            sH2f[r * 64 + c] = h;
            sH2b[r * 64 + c] = f2bf(h);
        });
    }
    __syncthreads();

    const int ntile = noutp / 64;
    // pass 1: row-wise log-sum-exp of both heads, online over the column
    // tiles; four threads per row, 16 columns each, merged at the end
    {
        const int row = tid >> 2, part = tid & 3;
        const int td = std_[row], tb = stb_[row];
        float md = -3.0e38f, sd = 0.f, mb = -3.0e38f, sb = 0.f, ztd = 0.f, ztb = 0.f;
        for (int jt = 0; jt < ntile; ++jt) {
            stage_w3_tile(a.w.W3, noutp, jt, sW3b);
            __syncthreads();
            f4v acc[2][2];
            acc_zero(acc);
            gemm64(sH2b, 64, 1, sW3b, 64, 1, 64, acc);
            acc_each(acc, [&](int r, int c, float v) { sZ[r * 64 + c] = v + a.w.b3[jt * 64 + c]; });
            __syncthreads();
            for (int c = part * 16; c < part * 16 + 16; ++c) {
                const int gc = jt * 64 + c;
                const float z = sZ[row * 64 + c];
                if (gc < kMlpDir) {
                    if (z > md) { sd = sd * __expf(md - z) + 1.f; md = z; }
                    else sd += __expf(z - md);
                    if (gc == td) ztd = z;
                } else if (gc < kMlpDir + nbin) {
                    if (z > mb) { sb = sb * __expf(mb - z) + 1.f; mb = z; }
                    else sb += __expf(z - mb);
                    if (gc - kMlpDir == tb) ztb = z;
                }
            }
            __syncthreads();
        }
        for (int o = 1; o < 4; o <<= 1) {   // merge the four partial (max, sum) pairs of the row
            const float md2 = __shfl_xor(md, o, 4), sd2 = __shfl_xor(sd, o, 4);
            const float mb2 = __shfl_xor(mb, o, 4), sb2 = __shfl_xor(sb, o, 4);
            const float mdn = fmaxf(md, md2), mbn = fmaxf(mb, mb2);
            sd = sd * __expf(md - mdn) + sd2 * __expf(md2 - mdn);
            sb = sb * __expf(mb - mbn) + sb2 * __expf(mb2 - mbn);
            md = mdn;
            mb = mbn;
            ztd += __shfl_xor(ztd, o, 4);
            ztb += __shfl_xor(ztb, o, 4);
        }
        if (part == 0) {   // log-sum-exp per head; the sample's loss
            sm_d[row] = md + __logf(sd);
            sm_b[row] = mb + __logf(sb);
            sloss[row] = sw[row] * ((sm_d[row] - ztd) + (sm_b[row] - ztb));
        }
    }
    __syncthreads();

    // pass 2: dZ = w (softmax - onehot), gradients of W3, b3 and dH2
    f4v accH[2][2];
    acc_zero(accH);
    for (int jt = 0; jt < ntile; ++jt) {
        stage_w3_tile(a.w.W3, noutp, jt, sW3b);
        __syncthreads();
        {
            f4v acc[2][2];
            acc_zero(acc);
            gemm64(sH2b, 64, 1, sW3b, 64, 1, 64, acc);
            acc_each(acc, [&](int r, int c, float v) {
                const int gc = jt * 64 + c;
                const float z = v + a.w.b3[gc];
                float dz = 0.f;
                if (gc < kMlpDir) dz = sw[r] * (__expf(z - sm_d[r]) - (gc == std_[r] ? 1.f : 0.f));
                else if (gc < kMlpDir + nbin) dz = sw[r] * (__expf(z - sm_b[r]) - (gc - kMlpDir == stb_[r] ? 1.f : 0.f));
                sZ[r * 64 + c] = dz;
                sdZb[r * 64 + c] = f2bf(dz);
            });
        }
        __syncthreads();
        if (tid < 64) {   // db3 (column sums, FP32)
            float s = 0.f;
            for (int r = 0; r < 64; ++r) s += sZ[r * 64 + tid];
            part[L.ob3 + jt * 64 + tid] = s;
        }
        {   // dW3 tile = H2^T dZ
            f4v acc[2][2];
            acc_zero(acc);
            gemm64(sH2b, 1, 64, sdZb, 64, 1, 64, acc);
            acc_each(acc, [&](int h, int c, float v) { part[L.oW3 + (size_t)h * noutp + jt * 64 + c] = v; });
        }
        // dH2 += dZ W3_tile^T
        gemm64(sdZb, 64, 1, sW3b, 1, 64, 64, accH);
        __syncthreads();
    }
    // layer 2 backward: dA2 = dH2 (1 - H2^2)
    acc_each(accH, [&](int r, int c, float v) {
        const float h = sH2f[r * 64 + c];
        const float d = v * (1.f - h * h);
        sZ[r * 64 + c] = d;
        sdZb[r * 64 + c] = f2bf(d);
    });
    __syncthreads();
    if (tid < 64) {
        float s = 0.f;
        for (int r = 0; r < 64; ++r) s += sZ[r * 64 + tid];
        part[L.ob2 + tid] = s;
    }
    {   // dW2 = H1^T dA2
        f4v acc[2][2];
        acc_zero(acc);
        gemm64(sH1b, 1, 64, sdZb, 64, 1, 64, acc);
        acc_each(acc, [&](int h, int c, float v) { part[L.oW2 + h * 64 + c] = v; });
    }
    {   // dA1 = (dA2 W2^T)(1 - H1^2), kept in sH2f
        f4v acc[2][2];
        acc_zero(acc);
        gemm64(sdZb, 64, 1, sW2b, 1, 64, 64, acc);
        __syncthreads();   // every wave has read sH2f (dA2 epilogue) before it is overwritten
        acc_each(acc, [&](int r, int c, float v) {
            const float h = sH1f[r * 64 + c];
            sH2f[r * 64 + c] = v * (1.f - h * h);
        });
    }
    __syncthreads();
    if (tid < 64) {
        float s = 0.f;
        for (int r = 0; r < 64; ++r) s += sH2f[r * 64 + tid];
        part[L.ob1 + tid] = s;
    }
    for (int e = tid; e < kMlpIn * kMlpHid; e += blockDim.x) {   // dW1 = X^T dA1 (FP32)
        const int i = e / kMlpHid, j = e - i * kMlpHid;
        float s = 0.f;
        for (int r = 0; r < 64; ++r) s = fmaf(sX[r * kMlpIn + i], sH2f[r * 64 + j], s);
        part[L.oW1 + e] = s;
    }
    if (tid == 0) {
        float s = 0.f;
        for (int r = 0; r < 64; ++r) s += sloss[r];
        a.loss_part[blockIdx.x] = s;
    }
}

// grad[p] = sum over blocks of part[b][p]; loss = sum of loss_part.
__global__ void mlp_reduce(const float* part, const float* loss_part, int nblk, int P, float* grad, float* loss)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P) {
        float s = 0.f;
        for (int b = 0; b < nblk; ++b) s += part[(size_t)b * P + p];
        grad[p] = s;
    }
    if (p == 0) {
        float s = 0.f;
        for (int b = 0; b < nblk; ++b) s += loss_part[b];
        *loss = s;
    }
}

__global__ void mlp_wsum(const float* wt, long long n, float* out)
{
    __shared__ float s[256];
    float v = 0.f;
    for (long long i = threadIdx.x; i < n; i += blockDim.x) v += wt[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0] > 0.f ? s[0] : 1.f;
}

// Adam on the FP32 master parameters; BF16 copies of the weight matrices.
__global__ void mlp_adam(float* theta, float* m, float* v, const float* grad, int P, MlpLayout L, float lr, float b1,
                         float b2, float eps, float c1, float c2, uint16_t* W1b, uint16_t* W2b, uint16_t* W3b)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float t = theta[p] - lr * (mm * c1) / (sqrtf(vv * c2) + eps);
    theta[p] = t;
    if (p >= L.oW1 && p < L.ob1) W1b[p - L.oW1] = f2bf(t);
    else if (p >= L.oW2 && p < L.ob2) W2b[p - L.oW2] = f2bf(t);
    else if (p >= L.oW3 && p < L.ob3) W3b[p - L.oW3] = f2bf(t);
}

// BF16 copies of the weight matrices of theta (after loading parameters).
__global__ void mlp_to_bf16(const float* theta, int P, MlpLayout L, uint16_t* W1b, uint16_t* W2b, uint16_t* W3b)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const float t = theta[p];
    if (p >= L.oW1 && p < L.ob1) W1b[p - L.oW1] = f2bf(t);
    else if (p >= L.oW2 && p < L.ob2) W2b[p - L.oW2] = f2bf(t);
    else if (p >= L.oW3 && p < L.ob3) W3b[p - L.oW3] = f2bf(t);
}

// Forward pass for n rows (64 per block): logits Z[n][noutp] (tests).
__global__ __launch_bounds__(256) void mlp_forward_rows(const float* X, long long n, MlpWeights w, float* Z)
{
    extern __shared__ float lds[];
    float* sX = lds;
    uint16_t* sH1b = (uint16_t*)(sX + 64 * kMlpIn);
    uint16_t* sH2b = sH1b + 4096;
    uint16_t* sW2b = sH2b + 4096;
    uint16_t* sW3b = sW2b + 4096;
    const long long r0 = (long long)blockIdx.x * 64;
    const int nrow = (int)((n - r0) < 64 ? (n - r0) : 64);
    for (int e = threadIdx.x; e < 64 * kMlpIn; e += blockDim.x) {
        const int s = e / kMlpIn;
        sX[e] = s < nrow ? X[(r0 + s) * kMlpIn + (e - s * kMlpIn)] : 0.f;
    }
    for (int e = threadIdx.x; e < 4096; e += blockDim.x) sW2b[e] = w.W2[e];
    __syncthreads();
    layer1(sX, w, nrow, nullptr, sH1b);
    __syncthreads();
    {
        f4v acc[2][2];
        acc_zero(acc);
        gemm64(sH1b, 64, 1, sW2b, 64, 1, 64, acc);
        acc_each(acc, [&](int r, int c, float v) { sH2b[r * 64 + c] = f2bf(r < nrow ? tanhf(v + w.b2[c]) : 0.f); });
    }
    __syncthreads();
    for (int jt = 0; jt < w.noutp / 64; ++jt) {
        for (int e = threadIdx.x; e < 4096; e += blockDim.x) sW3b[e] = w.W3[(size_t)(e >> 6) * w.noutp + jt * 64 + (e & 63)];
        __syncthreads();
        f4v acc[2][2];
        acc_zero(acc);
        gemm64(sH2b, 64, 1, sW3b, 64, 1, 64, acc);
        acc_each(acc, [&](int r, int c, float v) {
            if (r < nrow) Z[(r0 + r) * w.noutp + jt * 64 + c] = v + w.b3[jt * 64 + c];
        });
        __syncthreads();
    }
}

// ------------------------------------------------------------ features ----
// Direction bin of descriptor words: 8 equal-area cos(theta) bins x 8 phi bins.
__device__ __forceinline__ int desc_dir_bin(uint32_t w0, uint32_t w1)
{
    return (int)(((w0 & 0xFFFFFFu) >> 21) * 8u + ((w1 & 0xFFFFFFu) >> 21));
}

// Training samples from the last plain run (records on, one batch): per ray
// its cell's features, direction bin, (band, g) bin and weight
// |x| / sqrt(steps + cost0), x = w D - winv G the ray's net contribution.  With
// cost0 < 0 the weight is |x|.  The |x|-weighted fit targets the
// zero-variance density p|x|; dividing by the square root of the ray's cost
// targets p|x|/sqrt(c), the density of least variance x cost.
__global__ void mlp_samples(FeatArgs f, const int4* slotinfo, long long slot0, int R, long long nrays, int nQ,
                            const uint32_t* w0, const uint32_t* w1, const float* wv, const float* winv,
                            const float* recD, const float* recG, const int* recS, float cost0, float* X, int* td,
                            int* tb, float* wt)
{
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nrays; q += (long long)gridDim.x * blockDim.x) {
        const int4 si = slotinfo[slot0 + q / R];
        const int i = si.w & 1023, j = (si.w >> 10) & 1023, k = (si.w >> 20) & 1023;
        float x[kMlpIn];
        cell_features(f, i, j, k, x);
        for (int t = 0; t < kMlpIn; ++t) X[q * kMlpIn + t] = x[t];
        td[q] = desc_dir_bin(w0[q], w1[q]);
        tb[q] = desc_band(w0[q], w1[q]) * nQ + desc_g(w1[q]);
        const float xr = fabsf(wv[q] * recD[q] - winv[q] * recG[q]);
        wt[q] = cost0 < 0.f ? xr : xr * rsqrtf((float)recS[q] + cost0);
    }
}

// Features of the first n slots' cells (tests).
__global__ void mlp_features(FeatArgs f, const int4* slotinfo, long long n, float* X)
{
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        const int4 si = slotinfo[q];
        float x[kMlpIn];
        cell_features(f, si.w & 1023, (si.w >> 10) & 1023, (si.w >> 20) & 1023, x);
        for (int t = 0; t < kMlpIn; ++t) X[q * kMlpIn + t] = x[t];
    }
}

// -------------------------------------------------- learned ray generation --
// mlp_proposal: one block per 64 consecutive slots, one pass over the output
// tiles.  The network's logits of the 64 cells (MFMA), e = exp(clamp(z)) and
// the running (unnormalised) sums per head written to U[slot][noutp]:
// columns [0, 64) the direction head, [64, 64 + nbin) the (band, g) head.
// The sampler normalises by the head's last entry and forms the defensive
// mixture on the fly, so no second pass over the logits is needed.
__global__ __launch_bounds__(256) void mlp_proposal(GenArgs a, float* U)
{
    // LDS (mlp_proposal_lds()): rows of kPropStride elements, so that MFMA
    // operand fragments (four consecutive BF16 of one row, one 8-byte read)
    // and the row scans are free of bank conflicts
    extern __shared__ float lds[];
    constexpr int PS = kPropStride;
    float* sX = lds;                                  // 64 x 6 features
    float* sZ = sX + 64 * kMlpIn;                     // 64 x PS: exp(z) of a column tile
    uint16_t* sW2T = (uint16_t*)sZ;                   //   (W2^T during layer 2)
    uint16_t* sH1b = (uint16_t*)(sZ + 64 * PS);       // 64 x PS: H1, then the W3^T tile
    uint16_t* sW3T = sH1b;
    uint16_t* sH2b = sH1b + 64 * PS;                  // 64 x PS: H2
    const int noutp = a.mlpw.noutp, nbin = a.mlpw.nbin;
    const int tid = threadIdx.x;
    FeatArgs f = a.feat;
    f.T0 = a.T0 + (size_t)a.snap * a.tstride0;
    const int row = tid >> 2, part = tid & 3;         // 4 threads per row, 16 columns each
    const long long ngroups = (a.nslots + 63) / 64;
    for (long long grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        const long long s0 = grp * 64;
        const int ns = (int)((a.nslots - s0) < 64 ? (a.nslots - s0) : 64);
        for (int e = tid; e < 64 * kMlpIn; e += blockDim.x) sX[e] = 0.f;
        // W2^T (64 rows of 64 BF16) into the padded rows of the sZ region
        for (int e = tid; e < 512; e += blockDim.x) {
            const uint4 r = ((const uint4*)a.mlpw.W2T)[e];
            uint2* d = (uint2*)(sW2T + (e >> 3) * PS + (e & 7) * 8);
            d[0] = make_uint2(r.x, r.y);
            d[1] = make_uint2(r.z, r.w);
        }
        __syncthreads();
        if (tid < ns) {
            const int4 si = a.slotinfo[a.slot0 + s0 + tid];
            cell_features(f, si.w & 1023, (si.w >> 10) & 1023, (si.w >> 20) & 1023, sX + tid * kMlpIn);
        }
        __syncthreads();
        layer1(sX, a.mlpw, ns, nullptr, sH1b, PS);
        __syncthreads();
        {
            f4v acc[2][2];
            acc_zero(acc);
            gemm64(sH1b, PS, 1, sW2T, 1, PS, 64, acc);
            acc_each(acc, [&](int r, int c, float v) { sH2b[r * PS + c] = f2bf(r < ns ? tanhf(v + a.mlpw.b2[c]) : 0.f); });
        }
        __syncthreads();
        float run = 0.f;                              // this row's running sum (held by all 4 lanes)
        // the W3 tile and this lane's two b3 columns of it (qc + lane % 16 and
        // 16 more) are loaded one tile ahead; the tile loop's barriers wait
        // for LDS traffic only
        const int lane = tid & 63, bc = ((tid >> 6) & 1) * 32 + (lane & 15);
        uint4 w3a, w3b;
        load_w3_tile(a.mlpw.W3T, 0, w3a, w3b);
        float b3a = a.mlpw.b3[bc], b3b = a.mlpw.b3[bc + 16];
        for (int jt = 0; jt < noutp / 64; ++jt) {
            put_w3_tile(w3a, w3b, sW3T);
            const float ba = b3a, bb = b3b;
            lds_sync();
            const int jn = min(jt + 1, noutp / 64 - 1);   // (the last tile again at the end)
            load_w3_tile(a.mlpw.W3T, jn, w3a, w3b);
            b3a = a.mlpw.b3[jn * 64 + bc];
            b3b = a.mlpw.b3[jn * 64 + bc + 16];
            {
                f4v acc[2][2];
                acc_zero(acc);
                gemm64(sH2b, PS, 1, sW3T, 1, PS, 64, acc);
                acc_each(acc, [&](int r, int c, float v) {
                    const int gc = jt * 64 + c;
                    const float z = fminf(fmaxf(v + ((c & 16) ? bb : ba), -60.f), 60.f);
                    sZ[r * PS + c] = gc < kMlpDir + nbin ? __expf(z) : 0.f;
                });
            }
            lds_sync();
            if (jt == 1) run = 0.f;                   // the (band, g) head starts
            const float* zr = sZ + row * PS + part * 16;
            float loc[16];
            float acc = 0.f;
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                acc += zr[c];
                loc[c] = acc;
            }
            float excl = acc;                          // scan of the 4 segment totals
            for (int o = 1; o < 4; o <<= 1) {
                const float t = __shfl_up(excl, o, 4);
                if (part >= o) excl += t;
            }
            const float total = __shfl(excl, 3, 4);
            excl -= acc;
            if (row < ns) {
                float4* dst = (float4*)(U + (s0 + row) * (size_t)noutp + jt * 64 + part * 16);
                const float b0 = run + excl;
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    dst[c] = make_float4(b0 + loc[4 * c], b0 + loc[4 * c + 1], b0 + loc[4 * c + 2], b0 + loc[4 * c + 3]);
            }
            run += total;
            lds_sync();
        }
    }
}

// One learned ray: Philox (r, cell, snapshot, 1), the direction bin and the
// (band, g) bin from the defensive mixtures on the cell's row Ud (mix_search
// on each head), the direction uniform inside its bin, and p/q into the
// descriptor words of ray q.
__device__ __forceinline__ void learned_ray(const GenArgs& a, const float* Ud, const float* sPd, const float* sP,
                                            uint32_t r, uint32_t cell, long long q)
{
    const u32x4 x = philox4x32_10(u32x4{r, cell, (uint32_t)a.snap, 1u}, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
    float qd, qb;
    const int d = mix_search(Ud, sPd, kMlpDir, a.mix_eps, u01_open(x.x), &qd);
    const int bg = mix_search(Ud + kMlpDir, sP, a.mlpw.nbin, a.mix_eps, u01_open(x.y), &qb);
    const int band = bg / a.nQ, g = bg - band * a.nQ;
    // (a selected bin always has a positive step; the guard keeps a
    // degenerate row from producing inf * 0)
    const float ratio = (qb > 0.f && qd > 0.f) ? (a.p_bg[bg] / qb) * ((1.0f / kMlpDir) / qd) : 0.f;
    const uint32_t mt = ((uint32_t)(d >> 3) << 21) | (x.z >> 11);
    const uint32_t mp = ((uint32_t)(d & 7) << 21) | (x.w >> 11);
    uint32_t w0, w1;
    desc_pack(mt, mp, band, g, &w0, &w1);
    a.w0[q] = w0;
    a.w1[q] = w1;
    a.w[q] = ratio;
}

// raygen_learned: one thread per ray.  Each ray draws Philox (r, cell,
// snapshot, 1), its direction bin and (band, g) bin from the defensive
// mixtures by binary search on its cell's row of U (global memory: the R rays
// of a cell run in neighbouring lanes, so their first probes coincide and the
// rest hit the caches) and its direction uniform inside the bin.  The plain
// (band, g) CDF is in LDS.  The descriptor's w word carries p/q; the
// traversal multiplies it into the emission weight it computes from its LDS
// tables at ray start.
__global__ __launch_bounds__(256) void raygen_learned(GenArgs a, const float* U, const float* pcdf)
{
    extern __shared__ float lds[];
    const int noutp = a.mlpw.noutp, nbin = a.mlpw.nbin;
    float* sP = lds;
    for (int e = threadIdx.x; e < nbin; e += blockDim.x) sP[e] = pcdf[e];
    __shared__ float sPd[kMlpDir];
    for (int e = threadIdx.x; e < kMlpDir; e += blockDim.x) sPd[e] = (float)(e + 1) * (1.0f / kMlpDir);
    __syncthreads();
    const int R = a.rays_per_slot;
    const long long nr = a.nslots * (long long)R;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nr; q += (long long)gridDim.x * blockDim.x) {
        const long long sl = q / R;
        learned_ray(a, U + sl * (size_t)noutp, sPd, sP, (uint32_t)(q - sl * R), (uint32_t)a.slotinfo[a.slot0 + sl].y, q);
    }
}

// raygen_learned_lds: the same draws with the rows of G consecutive slots
// staged in LDS per block (G = 1..8 by the host: at 200 rays per cell and
// more the staging is shared by enough rays that the LDS probes' short
// latency pays for it).  Rays t of the group: slot t / R, ray t % R.
__global__ __launch_bounds__(256) void raygen_learned_lds(GenArgs a, const float* U, const float* pcdf, int G)
{
    extern __shared__ float lds[];
    const int noutp = a.mlpw.noutp, nbin = a.mlpw.nbin;
    float* sP = lds + G * noutp + 4;                  // (after the rows and one 16-byte pad slot)
    for (int e = threadIdx.x; e < nbin; e += blockDim.x) sP[e] = pcdf[e];
    __shared__ float sPd[kMlpDir];
    for (int e = threadIdx.x; e < kMlpDir; e += blockDim.x) sPd[e] = (float)(e + 1) * (1.0f / kMlpDir);
    __shared__ int s_cell[8];
    const int R = a.rays_per_slot;
    const long long ngroups = (a.nslots + G - 1) / G;
    const int npad = G * noutp / 4;                   // the pad slot (float4 index)
    for (long long grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        const long long s0 = grp * G;
        const int ns = (int)((a.nslots - s0) < G ? (a.nslots - s0) : G);
        // the rows into LDS, eight 16-byte loads in flight per thread (no
        // branch between them: loads past the rows read the last element
        // again, their stores go to the pad slot, which nothing reads)
        const float4* src = (const float4*)(U + s0 * (size_t)noutp);
        float4* dst = (float4*)lds;
        const int nv = ns * noutp / 4;
        for (int e0 = (int)threadIdx.x; e0 < nv; e0 += 8 * (int)blockDim.x) {
            float4 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = src[min(e0 + u * (int)blockDim.x, nv - 1)];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + u * (int)blockDim.x;
                dst[e < nv ? e : npad] = v[u];
            }
        }
        if ((int)threadIdx.x < ns) s_cell[threadIdx.x] = a.slotinfo[a.slot0 + s0 + threadIdx.x].y;
        __syncthreads();
        for (int t = threadIdx.x; t < ns * R; t += blockDim.x) {
            const int sl = t / R, r = t - sl * R;
            learned_ray(a, lds + sl * noutp, sPd, sP, (uint32_t)r, (uint32_t)s_cell[sl], (s0 + sl) * (long long)R + r);
        }
        __syncthreads();
    }
}

}  // namespace dn
