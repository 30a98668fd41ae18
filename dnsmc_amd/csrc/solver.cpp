// Solver orchestration and the C ABI (include/dnsmc_rad.h).
//
// Reference counterparts: code/mc_gpu.cu:440-638 (mc_gpu_: grids, T copy,
// interp3D, Tmax, readT, uploads, launches) and :640-784 (get_results_:
// synchronise, gather, reorder into the Fortran layout, sdev = sqrt(var)).
// Differences by design: tables and device state persist across calls (the
// reference re-reads 121 text files and resets the device every call), the
// coarse-grid spline runs on the GPU, and errors are returned as codes by the
// extended API (the two legacy symbols still abort with a message).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.h"
#include "solver.h"

namespace dn {
// kernels.hip
template <int SPEC, int TAB, int TMODE, bool REC, bool AMC, int GEN>
__global__ void trace(TraceArgs P);
__global__ void t_range(const float* T, int n0, int n1, int n2, int nbx, int nby, unsigned* lohi);
__global__ void raygen_plain(GenArgs a);
__global__ void raygen_compat(GenArgs a);
__global__ void raygen_compat_sorted(GenArgs a);
template <int SPEC, bool REC>
__global__ void trace_f64(TraceArgs P, double2* pray64);
__global__ void reduce_rays64(const double2* pray, int tally, long long nslots, int R, long long slot0,
                              const int4* slotinfo, double* E, double* A, double* Q2);
__global__ void finalize(const double* E, const double* A, const double* Q2, long long nemit, int spe, int R,
                         int compat, double* Q, double* Eo, double* Ao, double* sd);
__global__ void spline_pass(SplineArgs s);
__global__ void reduce_rays(const float2* pray, int tally, long long nslots, int R, long long slot0,
                            const int4* slotinfo, double* E, double* A, double* Q2);
__global__ void fill_f64(double* p, long long n, double v);
__global__ void cand_scan(TraceArgs P, uint32_t nfin, uint32_t* list);
// mlp.hip
__global__ void mlp_proposal(GenArgs a, float* cdf);
__global__ void raygen_learned(GenArgs a, const float* U, const float* pcdf);
__global__ void raygen_learned_lds(GenArgs a, const float* U, const float* pcdf, int G);
__global__ void mlp_train_tile(MlpTrainArgs a);
__global__ void mlp_reduce(const float* part, const float* loss_part, int nblk, int P, float* grad, float* loss);
__global__ void mlp_wsum(const float* wt, long long n, float* out);
__global__ void mlp_adam(float* theta, float* m, float* v, const float* grad, int P, MlpLayout L, float lr, float b1,
                         float b2, float eps, float c1, float c2, uint16_t* W1b, uint16_t* W2b, uint16_t* W3b);
__global__ void mlp_forward_rows(const float* X, long long n, MlpWeights w, float* Z);
__global__ void mlp_samples(FeatArgs f, const int4* slotinfo, long long slot0, int R, long long nrays, int nQ,
                            const uint32_t* w0, const uint32_t* w1, const float* wv, const float* winv,
                            const float* recD, const float* recG, const int* recS, float cost0, float* X, int* td,
                            int* tb, float* wt);
__global__ void mlp_features(FeatArgs f, const int4* slotinfo, long long n, float* X);
__global__ void fill_u32(unsigned* p, long long n, unsigned v);
__global__ void t_to_u(const float* T, float* U, long long n, float inv_dT, float mT0, float top, float lo);
__global__ void ghost_to_bricks(const float* Tg, int n0, int n1, int n2, float* T, int nbx, int nby);
__global__ void results_to_fortran(const int* emit, long long nemit, const double* Q, const double* sd, int n0,
                                   int n1, int jlo, int nj, int n2, float* resf, float* varf);
__global__ void slab_to_field(const double* cnew, int n0, int n1, int n2, int i0, int j0, int k0, int ni, int nj,
                              int nk, double Tc, double Tplus, float* T, int nbx, int nby);
__global__ void field_to_slab(const int* emit, long long nemit, const double* Q, const double* sd, int n0, int n1,
                              int jlo, int nj, int k0, int nk, double stefan, double tc4, double* divq, double* vcore);
__global__ void copy_f32x4(const float4* src, float4* dst, long long n4);
template <int U, bool NT>
__global__ void stream_copy(const float4* src, float4* dst, long long n4);
template <int U, bool NT>
__global__ void stream_read(const float4* src, long long n4, float* sink);
}  // namespace dn

using namespace dn;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(DNSMC_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

// Allocation-free solve (dnsmc_run_noalloc): the fused traversal's tallies,
// work-claim counters and step counter live in this file-scope device array,
// not in solver buffers.  Each device holds its own copy for the life of the
// process.  Host code addresses it only through the per-device pointer that
// hipGetSymbolAddress returns (scratch_base), never through the array's name.
// Layout: bytes [0, 1024) the 8 per-XCD claim counters (32 words apart),
// [1024, 1032) the step counter, from byte 4096 the FP64 tallies E, A and
// sum x^2, kScratchSlots slot-snapshots each.
constexpr long long kScratchBytes = 64ll << 20;
constexpr long long kScratchHead = 4096;
constexpr long long kScratchSlots = (kScratchBytes - kScratchHead) / (3 * (long long)sizeof(double));
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];
constexpr int kMaxDevices = 64;
// long rays first: the window of m_theta (24-bit polar uniform) about 2^23
// listed by cand_scan, |cos theta| < 0.02
constexpr uint32_t kCandW = 167772;
static std::mutex g_scratch_mu;
static char* g_scratch[kMaxDevices] = {};

// The scratch array's address on device `dev` (current), looked up once.
static hipError_t scratch_base(int dev, char** out)
{
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = (char*)ptr;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

// Where an allocation-free run keeps its counters and tallies and writes its
// results (device pointers).
struct NoAlloc {
    char* base;          // this device's scratch array
    double* Q;           // caller's results [nsnap][nemit]
    double* sd;
};

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    bool fits(size_t cnt) const { return p != nullptr && cnt <= n; }
    hipError_t ensure(size_t cnt)
    {
        if (cnt <= n && p) return hipSuccess;
        release();
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(cnt, 1) * sizeof(T));
        if (e == hipSuccess) n = cnt;
        return e;
    }
    hipError_t upload(const std::vector<T>& v, hipStream_t s)
    {
        hipError_t e = ensure(v.size());
        if (e != hipSuccess) return e;
        if (v.empty()) return hipSuccess;
        return hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    }
};

struct EventPair {
    hipEvent_t a, b;
    int kind;   // 0 raygen, 1 trace, 2 other
};

struct dnsmc_solver {
    dnsmc_config cfg;
    Tables tab;
    bool have_tables = false;
    std::vector<Level> lev;
    Derived der;
    float der_Tmax = -1.f;
    int dev = 0, ncu = 256;
    size_t lds_max = 65536;   // LDS bytes one block may allocate
    hipStream_t stream = nullptr;

    // fields
    std::vector<float> Tmax_snap;
    std::vector<float> Tlo_snap, Thi_snap;   // range over all levels of each snapshot
    std::vector<char> field_set;
    std::vector<char> slab_open;   // slabs written since the last commit
    DevBuf<unsigned> d_lohi;
    DevBuf<double> d_slab;         // staging of the DNS coupling slabs
    DevBuf<float> d_stage;         // staging of host temperature fields
    DevBuf<float> d_T[kMaxLev];
    DevBuf<float> d_U[kMaxLev];    // table coordinates of the sliced-table traversal
    long long ncell[kMaxLev] = {0, 0, 0, 0, 0, 0};     // cells per level
    long long tstride[kMaxLev] = {0, 0, 0, 0, 0, 0};   // bricked floats per level and snapshot
    int nbx[kMaxLev] = {0, 0, 0, 0, 0, 0}, nby[kMaxLev] = {0, 0, 0, 0, 0, 0};
    std::vector<float> faces_packed, centres_packed;
    std::vector<int> offs_packed;   // brick-layout offset of each cell index, per level and axis
    int foff[kMaxLev][3], coff[3], ooff[kMaxLev][3];
    DevBuf<int> d_offs;
    DevBuf<float> d_faces, d_centres;
    bool geom_uploaded = false;
    // spline
    std::vector<SplineAxis> sp[kMaxLev][3];
    DevBuf<double> d_spw;
    std::vector<DevBuf<double>*> sp_dbufs;
    std::vector<DevBuf<int>*> sp_ibufs;
    SplineArgs spa[kMaxLev][3];
    bool spline_ready = false;
    // tables
    DevBuf<float> d_kq, d_Ib, d_Ibw, d_cdfb, d_cdfg, d_inv_emax;
    DevBuf<int> d_band_order;   // sorted compat mode
    DevBuf<double> d_inv_emax_d;
    // compat
    DevBuf<uint32_t> d_xw_sub, d_xw_off, d_sobol;
    bool compat_ready = false;
    // emitters
    std::vector<int> emit;
    std::vector<int> order;     // processing position -> emitter, 3-D brick order
    DevBuf<int> d_emit, d_order;
    DevBuf<int4> d_slotinfo;    // processing slot -> {tally slot, cell, bricked index, packed i|j|k}
    bool emit_dirty = true;
    bool order_yfast = false;      // the slot tables' emitter order (brick_order)
    bool emit_is_slab = false;     // emit holds the j slab from ystart (dnsmc_set_emitters_slab)
    // ray batches are pipelined over kPipes streams, each with its own buffers,
    // so the tail of one traversal launch overlaps the next batch
    struct Pipe {
        hipStream_t stream = nullptr;
        DevBuf<uint32_t> d_w0, d_w1;
        DevBuf<float> d_w, d_winv;
        DevBuf<float2> d_pray;   // per-ray (E, A) for the deterministic tallies
        DevBuf<double2> d_pray64;   // per-ray (E, A) of the FP64 traversal
        DevBuf<float> d_qcdf;    // learned proposal: per-slot CDF rows of the batch
        DevBuf<unsigned> d_claim;   // per-XCD super-chunk counters of the traversal
        DevBuf<uint32_t> d_cand;    // fused launches: the listed long rays
        DevBuf<unsigned> d_ccnt;    // their count and list claim counter
        hipEvent_t done = nullptr;
    };
    static constexpr int kPipes = 3;
    Pipe pipe[kPipes];
    hipEvent_t ev_ready = nullptr;
    hipEvent_t ev_gen[2] = {nullptr, nullptr}, ev_tr[2] = {nullptr, nullptr};   // snapshot-wise learned path
    // tallies and results
    DevBuf<double> d_E, d_A, d_Q2, d_Q, d_Eo, d_Ao, d_sd;
    DevBuf<unsigned long long> d_nsteps;
    // learned proposal (BF16 weights, FP32 biases), plain bin probabilities
    bool have_mlp = false;
    int mlp_nbin = 0;
    float mix_eps = 0.1f;
    DevBuf<uint16_t> d_W1b, d_W2b, d_W3b;
    DevBuf<uint16_t> d_W2T, d_W3T;   // W2, W3 with the input index contiguous (the proposal kernel's operands)
    DevBuf<float> d_b1, d_b2, d_b3, d_pbg, d_pcdf;
    DevBuf<float> d_lcache;    // learned draw in the traversal: each wave's row buffer
    DevBuf<float> d_pdir;      // the direction head's plain CDF
    DevBuf<float> d_lrows;     // logged proposal rows of the last learned run (dnsmc_set_draw_log)
    int draw_log = 0;
    // records
    int record = 0;
    long long nrec = 0;
    DevBuf<int> d_rsteps, d_rend;
    DevBuf<uint32_t> d_rhash;
    DevBuf<float> d_rD, d_rG;
    long long first_batch_rays = 0;
    // stats
    std::vector<EventPair> evs;
    size_t ev_used = 0;
    dnsmc_stats st;
    std::chrono::steady_clock::time_point t0;
    bool running = false;
    int last_R = 0, last_spe = 1;
    long long last_nemit = 0;
    int ystart = 1;
    int last_ystart = 1;   // emitter j slab of the last run (its results' placement)
    char* scratch = nullptr;                       // this device's allocation-free scratch array
    unsigned long long* nsteps_src = nullptr;      // step counter of the last run
};

// Grow a solver buffer, or, in an allocation-free run, check that it already
// holds what the run needs.
#define NEED(buf, cnt)                                                                                   \
    do {                                                                                                 \
        if (na) {                                                                                        \
            if (!(buf).fits((size_t)(cnt)))                                                              \
                return fail(DNSMC_ESTATE, std::string("dnsmc_run_noalloc: ") + #buf +                    \
                                              " would have to grow: run dnsmc_run once with this "       \
                                              "configuration, emitters and temperature range first");    \
        } else                                                                                           \
            HIPCHK((buf).ensure((size_t)(cnt)));                                                         \
    } while (0)
#define PUT(buf, vec)                                                                                    \
    do {                                                                                                 \
        NEED(buf, (vec).size());                                                                         \
        if (!(vec).empty())                                                                              \
            HIPCHK(hipMemcpyAsync((buf).p, (vec).data(), (vec).size() * sizeof((vec)[0]),               \
                                  hipMemcpyHostToDevice, s->stream));                                    \
    } while (0)

// ------------------------------------------------------------------ util ----
static int spe_of(const dnsmc_config& c) { return c.mode == DNSMC_MODE_COMPAT ? c.nvar : 1; }
static int rays_per_slot_of(const dnsmc_config& c) { return c.mode == DNSMC_MODE_COMPAT ? c.photmax : c.rays_per_cell; }

static int check_config(const dnsmc_config& c)
{
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s)
            if (c.bc[a][s] != DNSMC_BC_PERIODIC && c.bc[a][s] != DNSMC_BC_BLACK)
                return fail(DNSMC_EINVAL, "bc must be 1 (periodic) or 2 (black wall)");
    if (c.polar_axis < 0 || c.polar_axis > 2) return fail(DNSMC_EINVAL, "polar_axis must be 0, 1 or 2");
    if (c.mode < 0 || c.mode > 2) return fail(DNSMC_EINVAL, "unknown mode");
    if (c.spectral < 0 || c.spectral > 3) return fail(DNSMC_EINVAL, "unknown spectral model");
    if (c.mode == DNSMC_MODE_COMPAT && (c.nvar < 2 || c.photmax < 1))
        return fail(DNSMC_EINVAL, "compat mode needs nvar >= 2 and photmax >= 1");
    if (c.mode != DNSMC_MODE_COMPAT && c.rays_per_cell < 2) return fail(DNSMC_EINVAL, "rays_per_cell must be >= 2");
    if (!(c.toll > 0.f && c.toll < 1.f)) return fail(DNSMC_EINVAL, "toll must be in (0, 1)");
    if (c.nsnap < 1) return fail(DNSMC_EINVAL, "nsnap must be >= 1");
    if (c.tally < 0 || c.tally > 3) return fail(DNSMC_EINVAL, "unknown tally mode");
    if (c.precision < 0 || c.precision > 1) return fail(DNSMC_EINVAL, "precision must be 0 (FP32) or 1 (FP64)");
    if (c.precision == 1 && c.mode != DNSMC_MODE_PLAIN)
        return fail(DNSMC_EINVAL, "precision 1 (FP64 traversal) runs the plain estimator only");
    for (int a = 0; a < 3; ++a)
        if (c.n[a] < 1 || c.n[a] > 1023) return fail(DNSMC_EINVAL, "grid sizes must be in [1, 1023]");
    if (c.batch_rays < 1024) return fail(DNSMC_EINVAL, "batch_rays must be >= 1024");
    if (c.spectral == DNSMC_SPEC_GREY_T4 && !(c.kappa_grey > 0.f)) return fail(DNSMC_EINVAL, "kappa_grey must be > 0");
    if (c.p_row < 1 || c.n[1] % c.p_row) return fail(DNSMC_EINVAL, "p_row must divide n[1]");
    if (c.mode == DNSMC_MODE_COMPAT) {
        if (c.num_streams < 1 || c.n[0] % c.num_streams) return fail(DNSMC_EINVAL, "num_streams must divide n[0]");
        if (c.num_gpu < 1 || c.n[2] % c.num_gpu) return fail(DNSMC_EINVAL, "num_gpu must divide n[2]");
    }
    return DNSMC_OK;
}

static EventPair& next_event(dnsmc_solver* s, int kind)
{
    if (s->ev_used == s->evs.size()) {
        EventPair e;
        (void)hipEventCreate(&e.a);
        (void)hipEventCreate(&e.b);
        s->evs.push_back(e);
    }
    EventPair& e = s->evs[s->ev_used++];
    e.kind = kind;
    return e;
}

// ---------------------------------------------------------- geometry ----
static int upload_geometry(dnsmc_solver* s)
{
    if (s->geom_uploaded) return DNSMC_OK;
    s->faces_packed.clear();
    for (size_t l = 0; l < s->lev.size(); ++l)
        for (int a = 0; a < 3; ++a) {
            s->foff[l][a] = (int)s->faces_packed.size();
            s->faces_packed.insert(s->faces_packed.end(), s->lev[l].faces[a].begin(), s->lev[l].faces[a].end());
        }
    s->centres_packed.clear();
    for (int a = 0; a < 3; ++a) {
        s->coff[a] = (int)s->centres_packed.size();
        s->centres_packed.insert(s->centres_packed.end(), s->lev[0].centres[a].begin(), s->lev[0].centres[a].end());
    }
    HIPCHK(s->d_faces.upload(s->faces_packed, s->stream));
    HIPCHK(s->d_centres.upload(s->centres_packed, s->stream));
    for (size_t l = 0; l < s->lev.size(); ++l) {
        const int* n = s->lev[l].n;
        s->ncell[l] = (long long)n[0] * n[1] * n[2];
        s->nbx[l] = (n[0] + 3) / 4;
        s->nby[l] = (n[1] + 3) / 4;
        s->tstride[l] = (long long)s->nbx[l] * s->nby[l] * ((n[2] + 3) / 4) * 64;
        HIPCHK(s->d_T[l].ensure((size_t)s->tstride[l] * s->cfg.nsnap));
    }
    // brick_index(x, y, z) = ox[x] + oy[y] + oz[z] (disjoint bit fields), tabulated
    s->offs_packed.clear();
    for (size_t l = 0; l < s->lev.size(); ++l) {
        const int* n = s->lev[l].n;
        for (int a = 0; a < 3; ++a) {
            s->ooff[l][a] = (int)s->offs_packed.size();
            for (int c = 0; c < n[a]; ++c) {
                const int x = a == 0 ? c : 0, y = a == 1 ? c : 0, z = a == 2 ? c : 0;
                s->offs_packed.push_back(brick_index(x, y, z, s->nbx[l], s->nby[l]));
            }
        }
    }
    HIPCHK(s->d_offs.upload(s->offs_packed, s->stream));
    s->geom_uploaded = true;
    return DNSMC_OK;
}

static int prepare_spline(dnsmc_solver* s)
{
    if (s->spline_ready || s->lev.size() < 2) return DNSMC_OK;
    const Level& L0 = s->lev[0];
    size_t work = 0;
    for (size_t l = 1; l < s->lev.size(); ++l) {
        const Level& Lc = s->lev[l];
        int nf[3] = {L0.n[0], L0.n[1], L0.n[2]}, nc[3] = {Lc.n[0], Lc.n[1], Lc.n[2]};
        for (int a = 0; a < 3; ++a) {
            SplineAxis ax;
            spline_axis_coeffs(L0.centres[a], nf[a], Lc.centres[a], nc[a], ax);
            SplineArgs& A = s->spa[l][a];
            memset(&A, 0, sizeof A);
            auto up_d = [&](const std::vector<double>& v) {
                DevBuf<double>* b = new DevBuf<double>();
                (void)b->upload(v, s->stream);
                s->sp_dbufs.push_back(b);
                return (const double*)b->p;
            };
            DevBuf<int>* kb = new DevBuf<int>();
            HIPCHK(kb->upload(ax.klo, s->stream));
            s->sp_ibufs.push_back(kb);
            A.sig = up_d(ax.sig); A.invp = up_d(ax.invp); A.y2f = up_d(ax.y2f);
            A.c6 = up_d(ax.c6); A.ihl = up_d(ax.ihl); A.ihh = up_d(ax.ihh);
            A.ca = up_d(ax.a); A.cb = up_d(ax.b); A.cA = up_d(ax.A); A.cB = up_d(ax.B);
            A.klo = kb->p;
            A.nf = nf[a];
            A.nc = nc[a];
            if (a == 0) {          // (n0,n1,n2) -> (m0,n1,n2): line l = j + n1 k
                A.nlines = nf[1] * nf[2];
                A.A = 1ll << 40; A.sa_in = nf[0]; A.sb_in = 0; A.st_in = 1;
                A.sa_out = nc[0]; A.sb_out = 0; A.st_out = 1;
            } else if (a == 1) {   // (m0,n1,n2) -> (m0,m1,n2): line l = i + m0 k
                A.nlines = nc[0] * nf[2];
                A.A = nc[0]; A.sa_in = 1; A.sb_in = (long long)nc[0] * nf[1]; A.st_in = nc[0];
                A.sa_out = 1; A.sb_out = (long long)nc[0] * nc[1]; A.st_out = nc[0];
            } else {               // (m0,m1,n2) -> (m0,m1,m2): line l = i + m0 j
                A.nlines = nc[0] * nc[1];
                A.A = 1ll << 40; A.sa_in = 1; A.sb_in = 0; A.st_in = (long long)nc[0] * nc[1];
                A.sa_out = 1; A.sb_out = 0; A.st_out = (long long)nc[0] * nc[1];
            }
            work = std::max(work, (size_t)A.nlines * A.nf);
        }
    }
    HIPCHK(s->d_spw.ensure(work));
    s->spline_ready = true;
    return DNSMC_OK;
}

// Coarse temperature levels of one snapshot (code/spline.cpp:35-116 on the GPU).
static int run_spline(dnsmc_solver* s, int snap)
{
    if (s->lev.size() < 2) return DNSMC_OK;
    int rc = prepare_spline(s);
    if (rc) return rc;
    const Level& L0 = s->lev[0];
    static DevBuf<float> tmpx, tmpy;   // per-process scratch for the intermediate passes
    for (size_t l = 1; l < s->lev.size(); ++l) {
        const Level& Lc = s->lev[l];
        size_t nx = (size_t)Lc.n[0] * L0.n[1] * L0.n[2], ny = (size_t)Lc.n[0] * Lc.n[1] * L0.n[2];
        HIPCHK(tmpx.ensure(nx));
        HIPCHK(tmpy.ensure(ny));
        for (int a = 0; a < 3; ++a) {
            SplineArgs A = s->spa[l][a];
            A.work = s->d_spw.p;
            if (a == 0) {   // bricked fine field in, line l = j + n1 k
                A.in = s->d_T[0].p + (size_t)snap * s->tstride[0];
                A.out = tmpx.p;
                A.in_bricked = 1;
                A.fn1 = L0.n[1];
                A.fnbx = s->nbx[0];
                A.fnby = s->nby[0];
            } else if (a == 1) {
                A.in = tmpx.p;
                A.out = tmpy.p;
            } else {        // bricked coarse level out, line l = i + m0 j
                A.in = tmpy.p;
                A.out = s->d_T[l].p + (size_t)snap * s->tstride[l];
                A.brick_out = 1;
                A.m0 = Lc.n[0];
                A.nbx = s->nbx[l];
                A.nby = s->nby[l];
            }
            A.snap_in = A.snap_out = 0;
            int nb = (A.nlines + 127) / 128;
            hipLaunchKernelGGL(spline_pass, dim3(nb), dim3(128), 0, s->stream, A);
            HIPCHK(hipGetLastError());
        }
    }
    return DNSMC_OK;
}

// ------------------------------------------------------------ tables ----
static int ensure_derived(dnsmc_solver* s, float Tmax, const NoAlloc* na)
{
    if (s->der_Tmax == Tmax) return DNSMC_OK;
    if (!s->have_tables) return fail(DNSMC_ESTATE, "spectral tables not loaded (dnsmc_load_tables)");
    std::string err;
    if (s->tab.synth_grey) synth_grey_tables(s->tab, Tmax);
    int rc = build_derived(s->cfg, s->tab, Tmax, s->der, &err);
    if (rc) return fail(rc, err);
    const Derived& d = s->der;
    PUT(s->d_kq, d.kq);
    PUT(s->d_Ib, d.Ib);
    PUT(s->d_Ibw, d.Ibw);
    PUT(s->d_cdfb, d.cdf_band);
    PUT(s->d_cdfg, d.cdf_g);
    PUT(s->d_inv_emax, d.inv_emax);
    if (d.sorted) PUT(s->d_band_order, d.band_order);
    {   // plain probability of each (band, g) bin: the CDF steps the plain draw uses
        std::vector<float> pbg((size_t)d.nB * d.nQ, 0.f);
        double pb0 = 0.0;
        for (int b = 0; b < d.nB; ++b) {
            const double pb = (double)d.cdf_band[b] - pb0;
            pb0 = d.cdf_band[b];
            double pg0 = 0.0;
            for (int g = 0; g < d.nQ; ++g) {
                const double pg = d.g_search ? (double)d.cdf_g[(size_t)b * d.nQ + g] - pg0 : (g == 0 ? 1.0 : 0.0);
                if (d.g_search) pg0 = d.cdf_g[(size_t)b * d.nQ + g];
                pbg[(size_t)b * d.nQ + g] = (float)(pb * pg);
            }
        }
        PUT(s->d_pbg, pbg);
        std::vector<float> pcdf(pbg.size());   // inclusive CDF of pbg, last entry exactly 1
        double acc = 0.0;
        for (size_t i = 0; i < pbg.size(); ++i) {
            acc += pbg[i];
            pcdf[i] = (float)std::min(acc, 1.0);
        }
        if (!pcdf.empty()) pcdf.back() = 1.0f;
        PUT(s->d_pcdf, pcdf);
    }
    s->der_Tmax = Tmax;
    return DNSMC_OK;
}

static int ensure_compat(dnsmc_solver* s)
{
    if (s->compat_ready) return DNSMC_OK;
    std::vector<uint32_t> sub, off;
    xorwow_jump_tables(sub);
    // offsets: M^(2^j), j < 64
    {
        std::vector<uint32_t> m(160 * 5), t(160 * 5);
        for (int b = 0; b < 160; ++b) {
            Xorwow x{0, 0, 0, 0, 0, 0};
            uint32_t* w = &x.x0;
            w[b >> 5] = 1u << (b & 31);
            (void)xorwow_next(x);
            m[b * 5 + 0] = x.x0; m[b * 5 + 1] = x.x1; m[b * 5 + 2] = x.x2; m[b * 5 + 3] = x.x3; m[b * 5 + 4] = x.x4;
        }
        off.assign((size_t)64 * 800, 0u);
        for (int j = 0; j < 64; ++j) {
            std::copy(m.begin(), m.end(), off.begin() + (size_t)j * 800);
            for (int b = 0; b < 160; ++b) {   // square
                uint32_t r[5] = {0, 0, 0, 0, 0};
                for (int c = 0; c < 160; ++c)
                    if ((m[b * 5 + (c >> 5)] >> (c & 31)) & 1u)
                        for (int w = 0; w < 5; ++w) r[w] ^= m[c * 5 + w];
                for (int w = 0; w < 5; ++w) t[b * 5 + w] = r[w];
            }
            m.swap(t);
        }
    }
    std::vector<uint32_t> sob(64);
    for (int k = 1; k <= 32; ++k) sob[k - 1] = 1u << (32 - k);   // dimension 0 (van der Corput)
    uint32_t mk = 1u;                                           // dimension 1: s=1, a=0, m1=1
    for (int k = 1; k <= 32; ++k) {
        if (k > 1) mk = (mk << 1) ^ mk;
        sob[32 + k - 1] = mk << (32 - k);
    }
    HIPCHK(s->d_xw_sub.upload(sub, s->stream));
    HIPCHK(s->d_xw_off.upload(off, s->stream));
    HIPCHK(s->d_sobol.upload(sob, s->stream));
    s->compat_ready = true;
    return DNSMC_OK;
}

// ---------------------------------------------------------------- run ----
// Processing order of the emitters: 8x8x8 blocks of cells, the cells inside
// a block in (k, j, i) order, the blocks in (k, i, j) order (yfast: j, the
// channel's wall-normal axis, fastest) or in (k, j, i) order.  Rays in flight
// then cover compact 3-D regions (few L2 lines, few TLB pages) instead of
// thin planes.  Measured on 384x256x256 at 100 rays/cell, j fastest against
// i fastest: plain 406 -> 403 ms per snapshot, the 8-snapshot launch 3176 ->
// 3156 ms, learned 1833 -> 1826 ms, but learned with coarsening 263 -> 273
// ms of traversal (8x16x8, 16x8x16, 8x32x8, 4x32x4 blocks no better); so the
// plain estimator without records walks j fastest and the other runs i
// fastest (the recorded runs that train the learned proposal included).
static void brick_order(const std::vector<int>& emit, int n0, int n1, int n2, bool yfast, std::vector<int>& order)
{
    const int B = 8;
    const long long ntot = (long long)n0 * n1 * n2;
    order.resize(emit.size());
    bool all = (long long)emit.size() == ntot;
    for (size_t e = 0; all && e < emit.size(); ++e) all = emit[e] == (int)e;
    const long long nbx = (n0 + B - 1) / B, nby = (n1 + B - 1) / B;
    if (all) {   // every cell in natural order: enumerate the blocks directly
        size_t p = 0;
        for (int bk = 0; bk < n2; bk += B)
            for (long long b = 0; b < nbx * nby; ++b) {
                const int bi = (int)(yfast ? b / nby : b % nbx) * B, bj = (int)(yfast ? b % nby : b / nbx) * B;
                for (int k = bk; k < std::min(bk + B, n2); ++k)
                    for (int j = bj; j < std::min(bj + B, n1); ++j)
                        for (int i = bi; i < std::min(bi + B, n0); ++i)
                            order[p++] = i + n0 * (j + n1 * k);
            }
        return;
    }
    std::vector<std::pair<long long, int>> key(emit.size());
    for (size_t e = 0; e < emit.size(); ++e) {
        int c = emit[e];
        int i = c % n0, j = (c / n0) % n1, k = c / (n0 * n1);
        long long block = (long long)(k / B) * nbx * nby + (yfast ? (long long)(i / B) * nby + (j / B)
                                                                  : (long long)(j / B) * nbx + (i / B));
        key[e] = {block * (B * B * B) + ((k % B) * B + (j % B)) * B + (i % B), (int)e};
    }
    std::sort(key.begin(), key.end());
    for (size_t p = 0; p < emit.size(); ++p) order[p] = key[p].second;
}

// Per processing slot: {tally slot, level-0 cell, bricked field index of the
// cell, packed (i-1) | (j-1) << 10 | (k-1) << 20}.
static void build_slotinfo(const std::vector<int>& emit, const std::vector<int>& order, int spe, const int* n,
                           int nbx, int nby, std::vector<int4>& out)
{
    out.resize(order.size() * (size_t)spe);
    for (size_t pe = 0; pe < order.size(); ++pe) {
        const int e = order[pe], cell = emit[e];
        const int i = cell % n[0], j = (cell / n[0]) % n[1], k = cell / (n[0] * n[1]);
        for (int v = 0; v < spe; ++v)
            out[pe * spe + v] = make_int4(e * spe + v, cell, brick_index(i, j, k, nbx, nby), i | (j << 10) | (k << 20));
    }
}

// Persistent traversal launch: the grid is the resident capacity of the
// device for this kernel variant (blocks per CU from the occupancy query,
// times the CU count), capped by `max_blocks`, a multiple of 8 (one share per
// XCD).  The sliced-table variant runs one 1024-thread block per CU.
template <int SPEC, int TAB, int TM, bool REC, int GEN = 0>
static hipError_t launch_trace_t(bool amc, int ncu, int max_blocks, size_t lds, hipStream_t st, const TraceArgs& P)
{
    void (*k)(TraceArgs) = amc ? trace<SPEC, TAB, TM, REC, true, GEN> : trace<SPEC, TAB, TM, REC, false, GEN>;
    int threads = TAB == kTabSliced ? 1024 : 256;
    if (TAB == kTabSliced)
        if (const char* be = getenv("DNSMC_BLOCK")) threads = std::max(64, std::min(1024, atoi(be) / 64 * 64));
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, threads, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    int nb = per_cu * ncu;
    nb = std::min(nb, TAB == kTabSliced ? std::max(8, max_blocks / 4) : max_blocks);
    nb = std::max(8, nb / 8 * 8);
    hipLaunchKernelGGL(k, dim3(nb), dim3(threads), lds, st, P);
    return hipGetLastError();
}

// gen: 0 descriptors, 1 the plain Philox draw, 2 the learned draw, in the kernel
template <int SPEC, int TAB>
static hipError_t launch_trace_m(int gen, bool det, bool rec, bool amc, int ncu, int max_blocks, size_t lds,
                                 hipStream_t st, const TraceArgs& P)
{
    if (gen == 2) return launch_trace_t<SPEC, TAB, 0, false, 2>(amc, ncu, max_blocks, lds, st, P);
    if (gen) return launch_trace_t<SPEC, TAB, 0, false, 1>(amc, ncu, max_blocks, lds, st, P);
    if (rec) return launch_trace_t<SPEC, TAB, 0, true>(amc, ncu, max_blocks, lds, st, P);
    if (det) return launch_trace_t<SPEC, TAB, 1, false>(amc, ncu, max_blocks, lds, st, P);
    return launch_trace_t<SPEC, TAB, 0, false>(amc, ncu, max_blocks, lds, st, P);
}

static hipError_t launch_trace(int spec, int tab, int gen, bool det, bool rec, bool amc, int ncu, int max_blocks,
                               size_t lds, hipStream_t st, const TraceArgs& P)
{
    if (spec == kSpecGreyT4)
        return launch_trace_m<kSpecGreyT4, kTabGlobal>(gen, det, rec, amc, ncu, max_blocks, lds, st, P);
    if (tab == kTabSliced) return launch_trace_m<kSpecTable, kTabSliced>(gen, det, rec, amc, ncu, max_blocks, lds, st, P);
    if (tab == kTabIbLds) return launch_trace_m<kSpecTable, kTabIbLds>(gen, det, rec, amc, ncu, max_blocks, lds, st, P);
    return launch_trace_m<kSpecTable, kTabGlobal>(gen, det, rec, amc, ncu, max_blocks, lds, st, P);
}

static MlpWeights mlp_weights(dnsmc_solver* s)
{
    MlpWeights w;
    w.W1 = s->d_W1b.p; w.b1 = s->d_b1.p;
    w.W2 = s->d_W2b.p; w.b2 = s->d_b2.p;
    w.W3 = s->d_W3b.p; w.b3 = s->d_b3.p;
    w.nbin = s->mlp_nbin;
    w.noutp = mlp_noutp(s->mlp_nbin);
    w.W2T = s->d_W2T.p;
    w.W3T = s->d_W3T.p;
    return w;
}

// The learned draws of a batch of nsl slots from their rows U: the rows of G
// slots staged in LDS per block (G = 4 below 200 rays per slot,
// 3 from 200 on), or one thread per ray searching the rows in global memory
// (DNSMC_RAYGEN_SLOTS = G forces a staging width, 0 the global search).  Both
// draw the same rays.  Measured, one coarsened 384x256x256 learned solve:
// 50 rays/cell 257 ms global, 257 / 235 / 249 / 255 ms for G = 2 / 4 / 6 /
// 8; 100: 411 global, 388 / 375 / 374 / 395 for G = 3 / 4 / 5 / 6; 200:
// 639 / 636 / 647 for G = 2 / 3 / 4; 400: 1198 / 1173 / 1163 for G = 1 / 2 / 3.
static hipError_t launch_raygen_learned(dnsmc_solver* s, const GenArgs& G, long long nsl, int R, hipStream_t st,
                                        const float* U, const float* pcdf)
{
    const int noutp = G.mlpw.noutp, nbin = G.mlpw.nbin;
    int gs = R < 200 ? 4 : 3;
    if (const char* e = getenv("DNSMC_RAYGEN_SLOTS")) gs = std::max(0, std::min(8, atoi(e)));
    const size_t lds_g = sizeof(float) * ((size_t)gs * noutp + 4 + nbin);   // rows, a 16-byte pad slot, the plain CDF
    if (gs > 0 && lds_g <= s->lds_max) {
        const int nb = (int)std::min<long long>((nsl + gs - 1) / gs, (long long)s->ncu * 8);
        hipLaunchKernelGGL(raygen_learned_lds, dim3(nb), dim3(256), lds_g, st, G, U, pcdf, gs);
        return hipGetLastError();
    }
    const int nb = (int)std::min<long long>((nsl * R + 255) / 256, (long long)s->ncu * 16);
    hipLaunchKernelGGL(raygen_learned, dim3(nb), dim3(256), sizeof(float) * (size_t)nbin, st, G, U, pcdf);
    return hipGetLastError();
}

// Feature arguments of the learned proposal (the field pointer is set per snapshot).
static FeatArgs feat_args(dnsmc_solver* s, const Derived& d)
{
    const dnsmc_config& c = s->cfg;
    FeatArgs f;
    memset(&f, 0, sizeof f);
    f.T0 = s->d_T[0].p;
    for (int a = 0; a < 3; ++a) {
        f.n[a] = c.n[a];
        f.coff[a] = s->coff[a];
    }
    f.nbx = s->nbx[0];
    f.nby = s->nby[0];
    f.centres = s->d_centres.p;
    for (int a = 0; a < 3; ++a)
        for (int sd = 0; sd < 2; ++sd)
            if (c.bc[a][sd] == DNSMC_BC_PERIODIC) f.bc_periodic |= 1 << (a * 2 + sd);
    f.polar = c.polar_axis;
    f.Lpolar = c.L[c.polar_axis];
    f.inv_Tmax = 1.0f / d.Tmax;
    f.Tw_rel = 0.5f * (c.Tw[c.polar_axis][0] + c.Tw[c.polar_axis][1]) / d.Tmax;
    return f;
}

// The whole solve, queued.  With `na` (dnsmc_run_noalloc) nothing is
// allocated: the fused plain path only, counters and tallies in the scratch
// array, results into the caller's arrays, no timing events.
static int do_run(dnsmc_solver* s, const NoAlloc* na)
{
    const dnsmc_config& c = s->cfg;
    int rc = check_config(c);
    if (rc) return rc;
    if (na) {
        if (c.mode != DNSMC_MODE_PLAIN || c.precision != 0 || c.tally != DNSMC_TALLY_F64_ATOMIC || s->record)
            return fail(DNSMC_EINVAL, "dnsmc_run_noalloc runs the plain estimator only (mode 0, precision 0, "
                                      "tally 0, records off)");
        if ((long long)s->emit.size() * c.nsnap > kScratchSlots)
            return fail(DNSMC_EINVAL, "dnsmc_run_noalloc: " + std::to_string((long long)s->emit.size()) +
                                          " emitters x " + std::to_string(c.nsnap) + " snapshots exceed the " +
                                          std::to_string(kScratchSlots) + " tally slots of its 64 MiB scratch");
    }
    for (int i = 0; i < c.nsnap; ++i)
        if (!s->field_set[i]) return fail(DNSMC_ESTATE, "temperature field of snapshot " + std::to_string(i) + " not set");
    if (s->emit.empty()) return fail(DNSMC_ESTATE, "no emitting cells");
    // tally slots, processing-slot tags and slot-table entries are 32-bit
    // indices over every snapshot of the run
    if ((long long)s->emit.size() * spe_of(c) * (long long)c.nsnap > 0x7FFFFFFFll)
        return fail(DNSMC_EINVAL, "emitters x blocks x snapshots must stay below 2^31 (split the run)");
    float Tmax = 0.f;
    for (int i = 0; i < c.nsnap; ++i) Tmax = std::max(Tmax, s->Tmax_snap[i]);
    for (int a = 0; a < 3; ++a)
        for (int sd = 0; sd < 2; ++sd) Tmax = std::max(Tmax, c.Tw[a][sd]);   // code/mc_gpu.cu:481-486
    rc = ensure_derived(s, Tmax, na);
    if (rc) return rc;
    if (c.mode == DNSMC_MODE_COMPAT && (rc = ensure_compat(s))) return rc;
    const int spe = spe_of(c), R = rays_per_slot_of(c);
    // the processing order for this run (brick_order): j-fastest blocks for the
    // plain estimator without records, i-fastest otherwise
    const bool yfast = c.mode == DNSMC_MODE_PLAIN && !s->record;
    if (s->emit_dirty || yfast != s->order_yfast) {
        s->order_yfast = yfast;
        brick_order(s->emit, c.n[0], c.n[1], c.n[2], yfast, s->order);
        std::vector<int4> si;
        build_slotinfo(s->emit, s->order, spe, c.n, s->nbx[0], s->nby[0], si);
        PUT(s->d_emit, s->emit);
        PUT(s->d_order, s->order);
        PUT(s->d_slotinfo, si);
        s->emit_dirty = false;
    }
    const Derived& d = s->der;
    const long long nemit = (long long)s->emit.size();
    const long long nslots = nemit * spe;
    const bool det = c.tally != DNSMC_TALLY_F64_ATOMIC;
    // ray batches: up to 2^30 rays (16 GB of descriptors) per launch; the
    // deterministic tallies keep 8 more bytes per ray, so they use 2^27
    const bool f64 = c.precision == 1;
    long long cap = f64 ? (1ll << 26) : (det ? (1ll << 27) : (1ll << 30));
    // learned batches: the per-slot CDF rows take at most 8 GB per pipeline;
    // when 80 % of the device memory the pipelines may use holds it, 32 GB and
    // up to 2^31 rays per batch instead (about 64 GB per pipeline at 400
    // rays/cell on the 288 GB MI355X): fewer, longer launches, 3.6 % faster
    long long cdf_bytes = 8ll << 30;
    const long long row_bytes = 4ll * mlp_noutp(s->mlp_nbin);
    if (c.mode == DNSMC_MODE_LEARNED && s->have_mlp && !f64 && !det) {
        const long long big_slots = std::max<long long>(1, std::min<long long>(std::min<long long>(c.batch_rays, 1ll << 31) / R,
                                                                                (32ll << 30) / row_bytes));
        const size_t need = (size_t)dnsmc_solver::kPipes * (size_t)big_slots * (size_t)(16ll * R + row_bytes);
        size_t have = 0, fr = 0, tot = 0;
        for (int p = 0; p < dnsmc_solver::kPipes; ++p)
            have += (size_t)s->pipe[p].d_w0.n * 16 + (size_t)s->pipe[p].d_qcdf.n * 4;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess && (double)need <= 0.8 * (double)(fr + have)) {
            cap = 1ll << 31;
            cdf_bytes = 32ll << 30;
        }
    }
    if (const char* bc = getenv("DNSMC_BATCH_CAP")) cap = std::max(1ll, std::min(atoll(bc), (1ll << 32) - (1ll << 20)));
    if (const char* cg = getenv("DNSMC_CDF_GB")) cdf_bytes = std::max(1ll, atoll(cg)) << 30;
    long long slots_per_batch = std::max<long long>(1, std::min<long long>(c.batch_rays, cap) / R);
    if (c.mode == DNSMC_MODE_LEARNED && s->have_mlp)
        slots_per_batch = std::max<long long>(1, std::min<long long>(slots_per_batch, cdf_bytes / row_bytes));
    // plain estimator with atomic tallies and no records: rays are drawn inside
    // the traversal, one launch for all snapshots (DNSMC_FUSED=0: batched);
    // more than 2^32 - 2^20 rays per snapshot take several launches over
    // consecutive slot ranges (ray indices within a launch are 32-bit)
    bool fused = c.mode == DNSMC_MODE_PLAIN && c.precision == 0 && c.tally == DNSMC_TALLY_F64_ATOMIC && !s->record;
    if (const char* fe = getenv("DNSMC_FUSED")) fused = fused && (atoi(fe) != 0 || na);
    // learned estimator, atomic tallies, no records, DNSMC_LEARNED_FUSED=1: the
    // proposal is evaluated and the rays are drawn inside the traversal too,
    // one launch for every snapshot (at most 1984 (band, g) bins, a wave's row
    // buffer).  Not the default: each wave evaluates one cell's row at a time,
    // so the output layer (64 x noutp BF16, 213 KB at 1600 bins) streams from
    // L2 once per cell, where the snapshot-wise generation below reuses each
    // of its tiles across 64 cells (MFMA); measured 2.11 s against 1.98 s for
    // 2.5e9 learned rays over 8 snapshots of a 384x256x256 channel
    bool lfused = c.mode == DNSMC_MODE_LEARNED && s->have_mlp && c.precision == 0 &&
                  c.tally == DNSMC_TALLY_F64_ATOMIC && !s->record && s->mlp_nbin <= 1984;
    {
        const char* le = getenv("DNSMC_LEARNED_FUSED");
        lfused = lfused && le && atoi(le) != 0;
    }
    if (lfused) fused = true;
    if (fused) slots_per_batch = nslots;
    // learned estimator, atomic tallies, no records: the descriptors of a whole
    // snapshot are generated (batch by batch, on a second stream) before one
    // traversal launch covers them, and the next snapshot is generated while
    // this one is traced (two snapshot-sized descriptor sets; DNSMC_SNAPWISE=0:
    // batched)
    bool snapwise = !lfused && c.mode == DNSMC_MODE_LEARNED && s->have_mlp && c.precision == 0 &&
                    c.tally == DNSMC_TALLY_F64_ATOMIC && !s->record && nslots * (long long)R <= (1ll << 32) - (1ll << 20);
    if (const char* se = getenv("DNSMC_SNAPWISE")) snapwise = snapwise && atoi(se) != 0;
    if (snapwise) {
        size_t fr = 0, tot = 0;
        const size_t need = (c.nsnap < 2 ? 1 : 2) * (size_t)(nslots * R) * 12 + (size_t)std::min(slots_per_batch, nslots) *
                                                               mlp_noutp(s->mlp_nbin) * 4;
        const size_t have = (size_t)s->pipe[0].d_w0.n * 12 + (size_t)s->pipe[1].d_w0.n * 12 +
                            (size_t)s->pipe[1].d_qcdf.n * 4;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess || need + ((size_t)4 << 30) > fr + have) snapwise = false;
    }
    const long long batch_rays = fused ? 0 : (snapwise ? nslots * R : std::min(slots_per_batch, nslots) * R);
    // fused: slots per traversal launch (32-bit ray indices per snapshot)
    long long fused_part = std::max<long long>(1, ((1ll << 32) - (1ll << 20)) / R);
    if (const char* fp = getenv("DNSMC_FUSED_SLOTS")) fused_part = std::max<long long>(1, std::min(fused_part, atoll(fp)));
    const long long nfused = (nslots + fused_part - 1) / fused_part;
    int pipes_max = dnsmc_solver::kPipes;   // DNSMC_PIPES=1 serialises the launches (standalone kernel timing)
    if (const char* pe = getenv("DNSMC_PIPES")) pipes_max = std::max(1, std::min(dnsmc_solver::kPipes, atoi(pe)));
    // Consecutive traversal launches of one solve alternate over two streams,
    // each with its own claim counters: a launch's last long rays (up to the
    // step cap, tens of ms) hold only their own CUs while the next launch fills
    // the others.  Fused: streams 0 and 1 (the in-kernel learned draw keeps one:
    // its row cache is indexed by wave).  Snapshot-wise: traversals on streams
    // 0 and 2, generation on 1.  Batched: up to kPipes pipelined streams.
    const int npipes = snapwise ? (c.nsnap > 1 && pipes_max >= 3 ? 3 : 2)
                       : fused  ? (int)std::min<long long>(lfused ? 1 : std::min(pipes_max, 2), nfused)
                                : (int)std::min<long long>(pipes_max, ((nslots + slots_per_batch - 1) / slots_per_batch) * c.nsnap);
    for (int p = 0; p < npipes && !na; ++p) {
        dnsmc_solver::Pipe& pp = s->pipe[p];
        // one snapshot needs only descriptor set 0; stream 2 of the
        // snapshot-wise path only traces (from the sets of streams 0 and 1)
        const long long nset = snapwise && ((p == 1 && c.nsnap < 2) || p == 2) ? 0 : batch_rays;
        HIPCHK(pp.d_w0.ensure(nset));
        HIPCHK(pp.d_w1.ensure(nset));
        HIPCHK(pp.d_w.ensure(nset));
        HIPCHK(pp.d_winv.ensure(snapwise ? 0 : batch_rays));
        if (det && !f64) HIPCHK(pp.d_pray.ensure(batch_rays));
        if (f64) HIPCHK(pp.d_pray64.ensure(batch_rays));
        if (c.mode == DNSMC_MODE_LEARNED && s->have_mlp && (!snapwise || p == 1))
            HIPCHK(pp.d_qcdf.ensure((size_t)std::min(slots_per_batch, nslots) * mlp_noutp(s->mlp_nbin)));
        HIPCHK(pp.d_claim.ensure(256));
    }
    // tallies, results and the step counter: solver buffers, or the scratch
    // array and the caller's arrays (allocation-free run)
    double *tE, *tA, *tQ2;
    unsigned long long* nsteps;
    unsigned* claim0 = nullptr;   // allocation-free run: its claim counters
    if (na) {
        claim0 = (unsigned*)na->base;
        nsteps = (unsigned long long*)(na->base + 1024);
        tE = (double*)(na->base + kScratchHead);
        tA = tE + kScratchSlots;
        tQ2 = tA + kScratchSlots;
    } else {
        HIPCHK(s->d_E.ensure(nslots * c.nsnap));
        HIPCHK(s->d_A.ensure(nslots * c.nsnap));
        HIPCHK(s->d_Q2.ensure(nslots * c.nsnap));
        HIPCHK(s->d_Q.ensure(nemit * c.nsnap));
        HIPCHK(s->d_Eo.ensure(nemit * c.nsnap));
        HIPCHK(s->d_Ao.ensure(nemit * c.nsnap));
        HIPCHK(s->d_sd.ensure(nemit * c.nsnap));
        HIPCHK(s->d_nsteps.ensure(2));
        tE = s->d_E.p;
        tA = s->d_A.p;
        tQ2 = s->d_Q2.p;
        nsteps = s->d_nsteps.p;
    }
    if (s->record) {
        s->nrec = std::min<long long>(batch_rays, 1ll << 22);
        HIPCHK(s->d_rsteps.ensure(s->nrec));
        HIPCHK(s->d_rend.ensure(s->nrec));
        HIPCHK(s->d_rhash.ensure(s->nrec));
        HIPCHK(s->d_rD.ensure(s->nrec));
        HIPCHK(s->d_rG.ensure(s->nrec));
    }
    s->ev_used = 0;
    memset(&s->st, 0, sizeof s->st);
    s->t0 = std::chrono::steady_clock::now();
    // counters are cleared by fill kernels (stream-ordered, capture-safe)
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, s->stream, (unsigned*)nsteps, 4ll, 0u);
    {
        long long n = nslots * c.nsnap;
        int nb = (int)((n + 255) / 256);
        hipLaunchKernelGGL(fill_f64, dim3(nb), dim3(256), 0, s->stream, tE, n, 0.0);
        hipLaunchKernelGGL(fill_f64, dim3(nb), dim3(256), 0, s->stream, tA, n, 0.0);
        hipLaunchKernelGGL(fill_f64, dim3(nb), dim3(256), 0, s->stream, tQ2, n, 0.0);
    }
    HIPCHK(hipEventRecord(s->ev_ready, s->stream));
    for (int p = 1; p < npipes; ++p) HIPCHK(hipStreamWaitEvent(s->pipe[p].stream, s->ev_ready, 0));
    // shared kernel arguments
    TraceArgs P;
    memset(&P, 0, sizeof P);
    P.nlev = (int)s->lev.size();
    for (int l = 0; l < P.nlev; ++l) {
        for (int a = 0; a < 3; ++a) {
            P.lev[l].n[a] = s->lev[l].n[a];
            P.lev[l].foff[a] = s->foff[l][a];
        }
        P.lev[l].maxs = s->lev[l].maxs;
        P.lev[l].tstride = s->tstride[l];
        P.lev[l].nbx = s->nbx[l];
        P.lev[l].nby = s->nby[l];
        for (int a = 0; a < 3; ++a) P.lev[l].ooff[a] = s->ooff[l][a];
        P.T[l] = s->d_T[l].p;
    }
    P.bc_periodic = 0;
    for (int a = 0; a < 3; ++a)
        for (int sd = 0; sd < 2; ++sd)
            if (c.bc[a][sd] == DNSMC_BC_PERIODIC) P.bc_periodic |= 1 << (a * 2 + sd);
    P.polar = c.polar_axis;
    P.tau_cut = (float)(-log((double)c.toll));
    P.u_off = d.u_off;
    P.nB = d.nB; P.nQ = d.nQ; P.nT = d.nT;
    P.inv_dT = d.inv_dT; P.mT0 = d.mT0;
    P.kappa_grey = c.kappa_grey;
    P.kq = s->d_kq.p;
    P.Ib = s->d_Ib.p;
    P.Ibw = s->d_Ibw.p;
    P.faces = s->d_faces.p;
    P.nfaces = (int)s->faces_packed.size();
    P.offs = s->d_offs.p;
    P.noffs = (int)s->offs_packed.size();
    P.centres = s->d_centres.p;
    for (int a = 0; a < 3; ++a) P.coff[a] = s->coff[a];
    P.ncent = (int)s->centres_packed.size();
    const int spec = c.spectral == DNSMC_SPEC_GREY_T4 ? kSpecGreyT4 : kSpecTable;
    // table placement: the kappa rows and Ib of the temperature nodes the
    // fields span go to LDS when they fit one block's share; else Ib alone
    // when small; else everything stays in global memory
    int tab = kTabGlobal;
    P.tab_lo = 0;
    P.tab_nw = d.nT;
    // the atomic-tally launches keep a tally ring per wave first in LDS (FP64
    // E, A, sum x^2 + two ints per entry, room for 16 waves); the cell centres
    // are read at ray start from global memory
    const size_t ring_bytes = (!det && !s->record) ? (size_t)16 * kRing * (3 * sizeof(double) + 2 * sizeof(int)) : 0;
    size_t lds_trace = ring_bytes + sizeof(float) * (6 * kLevWords + P.nfaces + P.noffs + 6 * (size_t)d.nB +
                                                     (fused ? (size_t)d.nB : 0));
    if (spec == kSpecTable) {
        float tlo = s->Tlo_snap[0], thi = s->Thi_snap[0];
        for (int i = 1; i < c.nsnap; ++i) {
            tlo = std::min(tlo, s->Tlo_snap[i]);
            thi = std::max(thi, s->Thi_snap[i]);
        }
        auto node = [&](float T) {   // the kernel's t0 for temperature T
            float pos = fmaf(T, d.inv_dT, d.mT0);
            pos = std::min(std::max(pos, 0.0f), (float)(d.nT - 1));
            int t0 = (int)pos;
            return std::min(t0, d.nT - 2);
        };
        const int lo = node(tlo), hi = node(thi);
        const int nw = hi + 2 - lo;
        const size_t sliced = lds_trace + sizeof(float) * (size_t)d.nB * (d.nQ + 1) * nw;
        const char* force = getenv("DNSMC_TABLES");   // "global": keep the tables out of LDS
        const bool allow = !(force && !strcmp(force, "global"));
        if (allow && sliced <= s->lds_max) {
            tab = kTabSliced;
            P.tab_lo = lo;
            P.tab_nw = nw;
            lds_trace = sliced;
            if (!f64) {
                // the FP32 traversal reads each cell's table coordinate within the
                // slice instead of its temperature (t_to_u), every level and snapshot
                for (size_t l = 0; l < s->lev.size(); ++l) {
                    const long long nl = s->tstride[l] * c.nsnap;
                    NEED(s->d_U[l], nl);
                    const int nb = (int)std::min<long long>((nl + 255) / 256, (long long)s->ncu * 16);
                    hipLaunchKernelGGL(t_to_u, dim3(nb), dim3(256), 0, s->stream, (const float*)s->d_T[l].p,
                                       s->d_U[l].p, nl, d.inv_dT, d.mT0, (float)(d.nT - 1), (float)lo);
                    HIPCHK(hipGetLastError());
                    P.T[l] = s->d_U[l].p;
                }
                // the pipeline streams start after the coordinates are written
                HIPCHK(hipEventRecord(s->ev_ready, s->stream));
                for (int p = 1; p < npipes; ++p) HIPCHK(hipStreamWaitEvent(s->pipe[p].stream, s->ev_ready, 0));
            }
        } else if (allow && (size_t)d.nB * d.nT <= 12288) {
            tab = kTabIbLds;
            lds_trace += sizeof(float) * (size_t)d.nB * d.nT;
        }
    }
    // fused generation: the g CDF rows, then 1/(k Ib)(Tmax), join the band CDF
    // in LDS when the block's share still holds them
    P.gen_lds = 0;
    P.desc_ratio = c.mode == DNSMC_MODE_LEARNED ? 1 : 0;
    P.inv_emax = s->d_inv_emax.p;
    P.Ibmax = d.Ibmax;
    const size_t gtab = sizeof(float) * (size_t)d.nB * d.nQ;
    if (fused && spec == kSpecTable && lds_trace + 2 * gtab <= s->lds_max) {
        P.gen_lds = 3;
        lds_trace += 2 * gtab;
    } else if (fused && spec == kSpecTable && lds_trace + gtab <= s->lds_max) {
        P.gen_lds = 1;
        lds_trace += gtab;
    }
    P.rays_per_slot = R;
    P.slots_per_emitter = spe;
    P.fd_R = make_fastdiv((uint32_t)R);
    {
        // measured defaults: launches with the ray queue (atomic tallies, no
        // records) refill at 24 idle lanes (the learned estimator's long rays
        // on the fine grid alone at 40: one 384x256x256 snapshot at 100
        // rays/cell traced in 1696 / 1671 / 1653 / 1710 ms at 24 / 32 / 40 /
        // 48, against 401 / 403 / 438 ms plain at 24 / 32 / 40 and 263 / 271
        // ms learned coarsened at 24 / 32), the others at 16; the fused and
        // snapshot-wise launches hand out 512-ray super-chunks, the batched 256
        const char* rk = getenv("DNSMC_REFILL_K");
        const int rk_queue = c.mode == DNSMC_MODE_LEARNED && P.nlev == 1 ? 40 : 24;
        P.refill_k = rk ? atoi(rk) : (!det && !s->record ? rk_queue : 16);
        if (P.refill_k < 1) P.refill_k = 1;
        if (P.refill_k > 64) P.refill_k = 64;
        const char* sc = getenv("DNSMC_SUPER");
        P.super = sc ? atoi(sc) : (fused || snapwise ? 512 : 256);
        P.super = std::max(64, (P.super + 63) / 64 * 64);
    }
    P.slotinfo = s->d_slotinfo.p;
    P.tally = c.tally;
    P.nsteps = nsteps;
    P.nrec = s->record ? s->nrec : 0;
    P.rec_steps = s->d_rsteps.p; P.rec_end = s->d_rend.p; P.rec_hash = s->d_rhash.p;
    P.rec_D = s->d_rD.p; P.rec_G = s->d_rG.p;
    P.Tmax = d.Tmax;
    P.Ibmax_d = d.Ibmax_d;
    const bool amc = P.nlev > 1;

    GenArgs G;
    memset(&G, 0, sizeof G);
    G.T0 = s->d_T[0].p;
    G.tstride0 = s->tstride[0];
    G.nB = d.nB; G.nQ = d.nQ; G.nT = d.nT;
    G.inv_dT = d.inv_dT; G.mT0 = d.mT0;
    G.kappa_grey = c.kappa_grey;
    G.spec = spec;
    G.kq = s->d_kq.p;
    G.Ib = s->d_Ib.p;
    G.cdf_band = s->d_cdfb.p;
    G.cdf_g = s->d_cdfg.p;
    G.inv_emax = s->d_inv_emax.p;
    G.Ibmax = d.Ibmax;
    G.g_search = d.g_search;
    G.slotinfo = s->d_slotinfo.p;
    G.rays_per_slot = R;
    G.slots_per_emitter = spe;
    G.seed = c.seed;

    if (c.mode == DNSMC_MODE_COMPAT) {
        xorwow_seed_state(c.seed, G.xw_seed);
        G.xw_jump_sub = s->d_xw_sub.p;
        G.xw_jump_off = s->d_xw_off.p;
        G.sobol_v = s->d_sobol.p;
        G.use_sobol = c.sobol;
        G.draws_per_bundle = (c.sobol ? 0 : 2) + 1 + (d.g_search ? 1 : 0);
        G.sub_len = 256;
        G.n0 = c.n[0]; G.n1 = c.n[1]; G.n2 = c.n[2];
        G.num_streams = c.num_streams; G.num_gpu = c.num_gpu; G.gpu = c.gpu; G.p_row = c.p_row;
        G.ystart = s->ystart;
        G.band_order = s->d_band_order.p;
        G.draws_find = 1 + (d.g_search ? 1 : 0);
    }
    const bool gib = spec == kSpecTable && (size_t)d.nB * d.nT <= 16384;
    const size_t lds_gen = sizeof(float) * ((size_t)d.nB + (size_t)d.nB * d.nQ + (gib ? (size_t)d.nB * d.nT : 0) + 64);
    const bool sorted = c.mode == DNSMC_MODE_COMPAT && d.sorted;
    const size_t lds_sorted = sizeof(float) * (2 * (size_t)d.nB + 3 * (size_t)d.nB * d.nQ + 64);
    if (sorted && lds_sorted + 1024 > s->lds_max)
        return fail(DNSMC_EINVAL, "sorted mode: too many (band, g) bins for the per-block histogram");
    size_t lds_learn = 0;
    if (c.mode == DNSMC_MODE_LEARNED) {
        if (!s->have_mlp) return fail(DNSMC_ESTATE, "learned mode needs dnsmc_set_mlp first");
        if (s->mlp_nbin != d.nB * d.nQ) return fail(DNSMC_EINVAL, "proposal network does not match the tables' bins");
        G.mlpw = mlp_weights(s);
        G.feat = feat_args(s, d);
        G.p_bg = s->d_pbg.p;
        G.mix_eps = s->mix_eps;
        lds_learn = sizeof(float) * (size_t)G.mlpw.nbin;   // raygen_learned: the plain (band, g) CDF
        if (lds_learn > s->lds_max) return fail(DNSMC_EINVAL, "too many (band, g) bins for the learned proposal");
    }

    const int trace_blocks = s->ncu * 16;
    s->first_batch_rays = 0;
    P.nsnap_launch = 1;
    P.tally_stride = 0;
    if (fused) {
        // one persistent launch: every snapshot, rays drawn in the refill
        dnsmc_solver::Pipe& pp = s->pipe[0];
        hipStream_t st = pp.stream;
        P.nsnap_launch = c.nsnap;
        P.tally_stride = (int)nslots;
        P.snap = 0;
        P.seed = c.seed;
        P.cdf_band = s->d_cdfb.p;
        P.cdf_g = s->d_cdfg.p;
        P.inv_emax = s->d_inv_emax.p;
        P.Ibmax = d.Ibmax;
        P.g_search = d.g_search;
        P.w0 = nullptr; P.w1 = nullptr; P.w = nullptr; P.winv = nullptr;
        P.pray = nullptr;
        P.E = tE;
        P.A = tA;
        P.Q2 = tQ2;
        P.nrec = 0;
        P.claim = na ? claim0 : pp.d_claim.p;
        if (lfused) {
            P.mlpw = mlp_weights(s);
            P.p_bg = s->d_pbg.p;
            P.pcdf = s->d_pcdf.p;
            {   // the direction head's plain CDF, as raygen_learned forms it
                std::vector<float> pd(kMlpDir);
                for (int e = 0; e < kMlpDir; ++e) pd[e] = (float)(e + 1) * (1.0f / kMlpDir);
                PUT(s->d_pdir, pd);
                P.pdir = s->d_pdir.p;
            }
            P.mix_eps = s->mix_eps;
            P.feat = feat_args(s, d);
            P.feat_stride = s->tstride[0];
            // a row per resident wave (at most 32 waves per CU)
            HIPCHK(s->d_lcache.ensure((size_t)s->ncu * 32 * kLRowFloats));
            P.lrow_cache = s->d_lcache.p;
            P.log_rows = nullptr;
            P.log_w0 = nullptr; P.log_w1 = nullptr; P.log_w = nullptr;
            if (s->draw_log) {
                const long long nr = nslots * (long long)R * c.nsnap;
                HIPCHK(s->d_lrows.ensure((size_t)nslots * mlp_noutp(s->mlp_nbin)));
                HIPCHK(pp.d_w0.ensure(nr));
                HIPCHK(pp.d_w1.ensure(nr));
                HIPCHK(pp.d_w.ensure(nr));
                HIPCHK(pp.d_winv.ensure(nr));   // not written: dnsmc_get_rays reads it
                P.log_rows = s->d_lrows.p;
                P.log_w0 = pp.d_w0.p; P.log_w1 = pp.d_w1.p; P.log_w = pp.d_w.p;
                s->first_batch_rays = nslots * (long long)R;
            }
        }
        // long rays first (the plain in-kernel draw; DNSMC_LONG_FIRST=0: off)
        bool long_first = !lfused;
        if (const char* lf = getenv("DNSMC_LONG_FIRST")) long_first = long_first && atoi(lf) != 0;
        // the listed part, the final 1/long_part of each XCD range of the last
        // snapshot: a half in launches of several snapshots, a quarter in
        // one-snapshot launches (measured on 384x256x256: 8 snapshots 3200 /
        // 3211 / 3205 ms for a half / quarter / all, one snapshot 411 / 409 / 417)
        long long long_part = c.nsnap > 1 ? 2 : 4;
        if (const char* lp = getenv("DNSMC_LONG_PART")) long_part = std::max(1, atoi(lp));
        long long li = 0;
        for (long long p0 = 0; p0 < nslots; p0 += fused_part, ++li) {
            // launch li on stream li % npipes with that stream's claim counters
            // (allocation-free run: 256 counters each at words 0 and 512 of the
            // scratch head)
            const int pi = (int)(li % npipes);
            hipStream_t ls = s->pipe[pi].stream;
            P.claim = na ? claim0 + 512 * pi : s->pipe[pi].d_claim.p;
            P.slot0 = p0;
            P.nrays = std::min(fused_part, nslots - p0) * R;
            hipLaunchKernelGGL(fill_u32, dim3(1), dim3(256), 0, ls, P.claim, 256ll, 0u);
            HIPCHK(hipGetLastError());
            // long rays first: list this launch's candidates, the rays of the
            // last snapshot in the final part of each of the 8 XCD ranges
            // that the traversal forms (its grid is a multiple of 8 blocks),
            // with |cos theta| < 0.02 about the polar axis (a sorting aid only:
            // any list that matches long_candidate traces every ray once)
            P.cand = nullptr;
            P.cand_cnt = nullptr;
            const uint32_t S = (uint32_t)P.super;
            const long long per_x = ((P.nrays + S - 1) / S + 7) / 8;
            if (long_first && !na && 8 * per_x * S < (1ll << 32)) {
                const long long clo = per_x - (per_x + long_part - 1) / long_part;
                const long long nfin = (per_x - clo) * S;
                P.cand_w = kCandW;
                P.cand_from = (uint32_t)(clo * S);
                P.cand_fd = make_fastdiv((uint32_t)(per_x * S));
                // per-range list capacity: the expected count (2 % of nfin) + 25 % + 512
                // (DNSMC_LONG_CAP=n: at most n, to exercise the overflow path)
                long long cap = (long long)((double)nfin * (2.0 * kCandW / 16777216.0) * 1.25) + 512;
                if (const char* lc = getenv("DNSMC_LONG_CAP")) cap = std::max(1ll, std::min(cap, atoll(lc)));
                dnsmc_solver::Pipe& cp = s->pipe[pi];
                HIPCHK(cp.d_cand.ensure(8 * cap));
                HIPCHK(cp.d_ccnt.ensure(512));
                hipLaunchKernelGGL(fill_u32, dim3(2), dim3(256), 0, ls, cp.d_ccnt.p, 512ll, 0u);
                P.cand = cp.d_cand.p;
                P.cand_cnt = cp.d_ccnt.p;
                P.cand_cap = (uint32_t)cap;
                const int nb = (int)std::min<long long>(8 * ((nfin + 4095) / 4096), (long long)s->ncu * 8);
                hipLaunchKernelGGL(cand_scan, dim3(nb), dim3(256), 0, ls, P, (uint32_t)nfin, cp.d_cand.p);
                HIPCHK(hipGetLastError());
            }
            const long long nsc = ((P.nrays + P.super - 1) / P.super) * c.nsnap;
            const int tb_cap = (int)std::min<long long>(trace_blocks, std::max<long long>(8, (nsc + 15) / 16));
            EventPair* et = na ? nullptr : &next_event(s, 1);
            if (et) HIPCHK(hipEventRecord(et->a, ls));
            HIPCHK(launch_trace(spec, tab, lfused ? 2 : 1, false, false, amc, s->ncu, tb_cap, lds_trace, ls, P));
            if (et) HIPCHK(hipEventRecord(et->b, ls));
            s->st.n_trace_launches++;
        }
    }
    if (snapwise) {
        // generation on pipe 1 into descriptor set (snap % 2), one traversal
        // launch per snapshot on the main stream
        hipStream_t gs = s->pipe[1].stream;
        const size_t lds_prop = mlp_proposal_lds();
        for (int snap = 0; snap < c.nsnap; ++snap) {
            const int b = snap & 1;
            dnsmc_solver::Pipe& buf = s->pipe[b];
            // odd snapshots trace on stream 2 when there is one
            dnsmc_solver::Pipe& tp = s->pipe[b && npipes > 2 ? 2 : 0];
            hipStream_t ts = tp.stream;
            if (snap >= 2) HIPCHK(hipStreamWaitEvent(gs, s->ev_tr[b], 0));   // trace(snap - 2) done with the set
            EventPair& eg = next_event(s, 0);
            HIPCHK(hipEventRecord(eg.a, gs));
            for (long long b0 = 0; b0 < nslots; b0 += slots_per_batch) {
                const long long nsl = std::min(slots_per_batch, nslots - b0);
                G.slot0 = b0;
                G.nslots = nsl;
                G.snap = snap;
                G.w0 = buf.d_w0.p + b0 * R;
                G.w1 = buf.d_w1.p + b0 * R;
                G.w = buf.d_w.p + b0 * R;
                G.winv = nullptr;
                int nb = (int)std::min<long long>((nsl + 63) / 64, (long long)s->ncu * 3);
                hipLaunchKernelGGL(mlp_proposal, dim3(nb), dim3(256), lds_prop, gs, G, s->pipe[1].d_qcdf.p);
                HIPCHK(hipGetLastError());
                HIPCHK(launch_raygen_learned(s, G, nsl, R, gs, (const float*)s->pipe[1].d_qcdf.p,
                                             (const float*)s->d_pcdf.p));
            }
            HIPCHK(hipEventRecord(eg.b, gs));
            HIPCHK(hipEventRecord(s->ev_gen[b], gs));
            HIPCHK(hipStreamWaitEvent(ts, s->ev_gen[b], 0));
            P.w0 = buf.d_w0.p; P.w1 = buf.d_w1.p; P.w = buf.d_w.p; P.winv = nullptr;
            P.pray = nullptr;
            P.nrays = nslots * R;
            P.slot0 = 0;
            P.snap = snap;
            P.E = s->d_E.p + (size_t)snap * nslots;
            P.A = s->d_A.p + (size_t)snap * nslots;
            P.Q2 = s->d_Q2.p + (size_t)snap * nslots;
            P.nrec = 0;
            P.claim = tp.d_claim.p;
            hipLaunchKernelGGL(fill_u32, dim3(1), dim3(256), 0, ts, tp.d_claim.p, 256ll, 0u);
            HIPCHK(hipGetLastError());
            const long long nsc = (P.nrays + P.super - 1) / P.super;
            const int tb_cap = (int)std::min<long long>(trace_blocks, std::max<long long>(8, (nsc + 15) / 16));
            EventPair& et = next_event(s, 1);
            HIPCHK(hipEventRecord(et.a, ts));
            HIPCHK(launch_trace(spec, tab, false, false, false, amc, s->ncu, tb_cap, lds_trace, ts, P));
            HIPCHK(hipEventRecord(et.b, ts));
            HIPCHK(hipEventRecord(s->ev_tr[b], ts));
            s->st.n_trace_launches++;
        }
    }
    long long bi = 0;
    for (int snap = 0; snap < c.nsnap && !fused && !snapwise; ++snap) {
        for (long long b0 = 0; b0 < nslots; b0 += slots_per_batch, ++bi) {
            dnsmc_solver::Pipe& pp = s->pipe[bi % npipes];
            hipStream_t st = pp.stream;
            long long nsl = std::min(slots_per_batch, nslots - b0);
            long long nr = nsl * R;
            if (snap == 0 && b0 == 0) s->first_batch_rays = nr;
            G.slot0 = b0;
            G.nslots = nsl;
            G.snap = snap;
            G.w0 = pp.d_w0.p; G.w1 = pp.d_w1.p; G.w = pp.d_w.p; G.winv = pp.d_winv.p;

            EventPair& eg = next_event(s, 0);
            HIPCHK(hipEventRecord(eg.a, st));
            if (c.mode == DNSMC_MODE_LEARNED) {
                const size_t lds_prop = mlp_proposal_lds();
                int nb = (int)std::min<long long>((nsl + 63) / 64, (long long)s->ncu * 3);
                hipLaunchKernelGGL(mlp_proposal, dim3(nb), dim3(256), lds_prop, st, G, pp.d_qcdf.p);
                HIPCHK(hipGetLastError());
                HIPCHK(launch_raygen_learned(s, G, nsl, R, st, (const float*)pp.d_qcdf.p, (const float*)s->d_pcdf.p));
            } else if (sorted) {
                int nb = (int)std::min<long long>(nsl, (long long)s->ncu * 8);
                hipLaunchKernelGGL(raygen_compat_sorted, dim3(nb), dim3(256), lds_sorted, st, G);
            } else if (c.mode == DNSMC_MODE_COMPAT) {
                long long nth = nsl * ((R + G.sub_len - 1) / G.sub_len);
                int nb = (int)std::min<long long>((nth + 255) / 256, 1 << 20);
                hipLaunchKernelGGL(raygen_compat, dim3(nb), dim3(256), lds_gen, st, G);
            } else {
                int nb = (int)std::min<long long>((nr + 255) / 256, (long long)s->ncu * 64);
                hipLaunchKernelGGL(raygen_plain, dim3(nb), dim3(256), lds_gen, st, G);
            }
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(eg.b, st));
            P.w0 = pp.d_w0.p; P.w1 = pp.d_w1.p; P.w = pp.d_w.p; P.winv = pp.d_winv.p;

            P.pray = pp.d_pray.p;
            P.nrays = nr;
            P.slot0 = b0;
            P.snap = snap;
            P.E = s->d_E.p + (size_t)snap * nslots;
            P.A = s->d_A.p + (size_t)snap * nslots;
            P.Q2 = s->d_Q2.p + (size_t)snap * nslots;
            P.nrec = (s->record && snap == 0 && b0 == 0) ? s->nrec : 0;
            bool rec = P.nrec > 0;
            // persistent grid, shrunk for small launches (about 4 super-chunks per wave)
            const long long nsc = (nr + P.super - 1) / P.super;
            const int tb_cap = (int)std::min<long long>(trace_blocks, std::max<long long>(8, (nsc + 15) / 16));
            P.claim = pp.d_claim.p;
            hipLaunchKernelGGL(fill_u32, dim3(1), dim3(256), 0, st, pp.d_claim.p, 256ll, 0u);
            HIPCHK(hipGetLastError());
            EventPair& et = next_event(s, 1);
            HIPCHK(hipEventRecord(et.a, st));
            if (f64) {
                const size_t lds64 = sizeof(float) * (6 * kLevWords + P.nfaces + P.noffs + P.ncent + 6 * (size_t)d.nB);
                const int nb = (int)std::min<long long>((nr + 255) / 256, (long long)s->ncu * 16);
                void (*k)(TraceArgs, double2*) =
                    spec == kSpecGreyT4 ? (rec ? trace_f64<kSpecGreyT4, true> : trace_f64<kSpecGreyT4, false>)
                                        : (rec ? trace_f64<kSpecTable, true> : trace_f64<kSpecTable, false>);
                hipLaunchKernelGGL(k, dim3(nb), dim3(256), lds64, st, P, pp.d_pray64.p);
            } else {
                HIPCHK(launch_trace(spec, tab, false, det, rec, amc, s->ncu, tb_cap, lds_trace, st, P));
            }
            HIPCHK(hipEventRecord(et.b, st));
            s->st.n_trace_launches++;
            if (f64) {
                int nb = (int)((nsl + 255) / 256);
                hipLaunchKernelGGL(reduce_rays64, dim3(nb), dim3(256), 0, st, (const double2*)pp.d_pray64.p, c.tally,
                                   nsl, R, b0, (const int4*)s->d_slotinfo.p, P.E, P.A, P.Q2);
                HIPCHK(hipGetLastError());
            } else if (det && !rec) {   // record launches tally with FP64 atomics
                int nb = (int)((nsl + 255) / 256);
                hipLaunchKernelGGL(reduce_rays, dim3(nb), dim3(256), 0, st, (const float2*)pp.d_pray.p, c.tally, nsl, R,
                                   b0, (const int4*)s->d_slotinfo.p, P.E, P.A, P.Q2);
                HIPCHK(hipGetLastError());
            }
        }
    }
    for (int p = 1; p < npipes; ++p) {
        HIPCHK(hipEventRecord(s->pipe[p].done, s->pipe[p].stream));
        HIPCHK(hipStreamWaitEvent(s->stream, s->pipe[p].done, 0));
    }
    for (int snap = 0; snap < c.nsnap; ++snap) {
        int nb = (int)((nemit + 255) / 256);
        const size_t o = (size_t)snap * nemit;
        hipLaunchKernelGGL(finalize, dim3(nb), dim3(256), 0, s->stream, tE + (size_t)snap * nslots,
                           tA + (size_t)snap * nslots, tQ2 + (size_t)snap * nslots, nemit, spe, R,
                           c.mode == DNSMC_MODE_COMPAT ? 1 : 0, na ? na->Q + o : s->d_Q.p + o,
                           na ? nullptr : s->d_Eo.p + o, na ? nullptr : s->d_Ao.p + o, na ? na->sd + o : s->d_sd.p + o);
        HIPCHK(hipGetLastError());
    }
    s->st.n_rays = nslots * (long long)R * c.nsnap;
    s->st.table_mode = tab;
    s->st.table_lo = P.tab_lo;
    s->st.table_nw = P.tab_nw;
    s->st.n_pipes = npipes;
    s->st.Tmax = d.Tmax;
    s->st.kappamax = d.kappamax;
    s->st.Ibmax = d.Ibmax_d;
    s->last_R = R;
    s->last_spe = spe;
    s->last_nemit = na ? 0 : nemit;   // an allocation-free run's results are the caller's only
    s->last_ystart = s->ystart;
    s->nsteps_src = nsteps;
    s->running = true;
    return DNSMC_OK;
}

static int do_wait(dnsmc_solver* s)
{
    if (!s->running) return DNSMC_OK;
    HIPCHK(hipStreamSynchronize(s->stream));
    s->running = false;
    auto t1 = std::chrono::steady_clock::now();
    s->st.ms_total = std::chrono::duration<double, std::milli>(t1 - s->t0).count();
    for (size_t i = 0; i < s->ev_used; ++i) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, s->evs[i].a, s->evs[i].b));
        if (s->evs[i].kind == 0) s->st.ms_raygen += ms;
        else if (s->evs[i].kind == 1) s->st.ms_trace += ms;
        else s->st.ms_other += ms;
    }
    unsigned long long ns[2] = {0, 0};
    HIPCHK(hipMemcpy(ns, s->nsteps_src, sizeof ns, hipMemcpyDeviceToHost));
    s->st.n_steps = (long long)ns[0];
    s->st.n_listed = (long long)ns[1];
    return DNSMC_OK;
}

// Rays and records of the first batch are stored in processing order; both
// getters return them in tally-slot order (emitter-major, then block, then ray).
static int slot_perm(dnsmc_solver* s, long long c, std::vector<long long>& perm)
{
    const int spe = spe_of(s->cfg), R = rays_per_slot_of(s->cfg);
    perm.resize(c);
    for (long long q = 0; q < c; ++q) {
        long long ps = q / R, r = q - ps * R;
        long long pe = ps / spe, v = ps - pe * spe;
        perm[q] = ((long long)s->order[pe] * spe + v) * R + r;
    }
    return DNSMC_OK;
}

template <class T>
static int fetch_perm(const T* dev, long long c, const std::vector<long long>& perm, T* out)
{
    if (!out) return DNSMC_OK;
    std::vector<T> tmp(c);
    HIPCHK(hipMemcpy(tmp.data(), dev, c * sizeof(T), hipMemcpyDeviceToHost));
    for (long long q = 0; q < c; ++q)
        if (perm[q] < c) out[perm[q]] = tmp[q];
    return DNSMC_OK;
}

// After the fine field of snapshot `snap` is on the device: coarse levels by
// the spline passes, the temperature range of every level (LDS table slice),
// Tmax and the positivity check (code/mc_gpu.cu:456-480).
static int commit_field(dnsmc_solver* s, int snap)
{
    int rc = run_spline(s, snap);
    if (rc) return rc;
    // temperature range over every level (the spline can leave the fine range)
    HIPCHK(s->d_lohi.ensure(2 * s->lev.size()));
    for (size_t l = 0; l < s->lev.size(); ++l) {
        hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, s->stream, s->d_lohi.p + 2 * l, 1ll, 0xFFFFFFFFu);
        hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, s->stream, s->d_lohi.p + 2 * l + 1, 1ll, 0u);
        const int* ln = s->lev[l].n;
        const long long ncl = (long long)ln[0] * ln[1] * ln[2];
        const int nb = (int)std::min<long long>((ncl + 255) / 256, (long long)s->ncu * 16);
        hipLaunchKernelGGL(t_range, dim3(nb), dim3(256), 0, s->stream, s->d_T[l].p + (size_t)snap * s->tstride[l],
                           ln[0], ln[1], ln[2], s->nbx[l], s->nby[l], s->d_lohi.p + 2 * l);
        HIPCHK(hipGetLastError());
    }
    std::vector<unsigned> lohi(2 * s->lev.size());
    HIPCHK(hipMemcpyAsync(lohi.data(), s->d_lohi.p, lohi.size() * sizeof(unsigned), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    auto key_float = [](unsigned k) {
        unsigned u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
        float f;
        memcpy(&f, &u, sizeof f);
        return f;
    };
    float lo = 0.f, hi = 0.f;
    for (size_t l = 0; l < s->lev.size(); ++l) {
        const float a = key_float(lohi[2 * l]), b = key_float(lohi[2 * l + 1]);
        lo = l == 0 ? a : std::min(lo, a);
        hi = l == 0 ? b : std::max(hi, b);
    }
    s->Tlo_snap[snap] = lo;
    s->Thi_snap[snap] = hi;
    s->Tmax_snap[snap] = key_float(lohi[1]);   // level-0 maximum
    s->slab_open[snap] = 0;
    // fine field: positive and finite everywhere (a NaN shows in the level-0 bounds)
    if (!(key_float(lohi[0]) > 0.f) || !std::isfinite(key_float(lohi[1]))) {
        s->field_set[snap] = 0;
        return fail(DNSMC_ERANGE, "temperature field holds a non-positive or non-finite value");
    }
    s->field_set[snap] = 1;
    return DNSMC_OK;
}

// ================================================================ C ABI ====
extern "C" {

void dnsmc_default_config(dnsmc_config* cfg) { default_config(cfg); }

int dnsmc_parse_config(const char* text, dnsmc_config* cfg)
{
    std::string err;
    int rc = parse_config(text, cfg, &err);
    if (rc) return fail(rc, err);
    return DNSMC_OK;
}

const char* dnsmc_last_error(void) { return g_err.c_str(); }

int dnsmc_create(const dnsmc_config* cfg, dnsmc_solver** out)
{
    if (!cfg || !out) return fail(DNSMC_EINVAL, "null argument");
    *out = nullptr;
    int rc = check_config(*cfg);
    if (rc) return rc;
    dnsmc_solver* s = new dnsmc_solver();
    s->cfg = *cfg;
    std::string err;
    rc = build_levels(s->cfg, s->lev, &err);
    if (rc) { delete s; return fail(rc, err); }
    if (hipGetDevice(&s->dev) != hipSuccess) { delete s; return fail(DNSMC_EDEVICE, "no HIP device"); }
    {   // the allocation-free run's scratch array on this device (looked up once per device)
        const hipError_t e = scratch_base(s->dev, &s->scratch);
        if (e != hipSuccess) {
            delete s;
            return fail(DNSMC_EDEVICE, std::string("scratch array address: ") + hipGetErrorString(e));
        }
    }
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, s->dev) == hipSuccess && ncu > 0) s->ncu = ncu;
    int lmax = 0;
    if (hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, s->dev) == hipSuccess && lmax > 0)
        s->lds_max = (size_t)lmax;
    if (hipStreamCreate(&s->stream) != hipSuccess) { delete s; return fail(DNSMC_EDEVICE, "hipStreamCreate failed"); }
    s->pipe[0].stream = s->stream;
    for (int p = 1; p < dnsmc_solver::kPipes; ++p)
        if (hipStreamCreate(&s->pipe[p].stream) != hipSuccess) { dnsmc_destroy(s); return fail(DNSMC_EDEVICE, "hipStreamCreate failed"); }
    for (int p = 0; p < dnsmc_solver::kPipes; ++p)
        if (hipEventCreateWithFlags(&s->pipe[p].done, hipEventDisableTiming) != hipSuccess) { dnsmc_destroy(s); return fail(DNSMC_EDEVICE, "hipEventCreate failed"); }
    if (hipEventCreateWithFlags(&s->ev_ready, hipEventDisableTiming) != hipSuccess) { dnsmc_destroy(s); return fail(DNSMC_EDEVICE, "hipEventCreate failed"); }
    for (int b = 0; b < 2; ++b)
        if (hipEventCreateWithFlags(&s->ev_gen[b], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_tr[b], hipEventDisableTiming) != hipSuccess) {
            dnsmc_destroy(s);
            return fail(DNSMC_EDEVICE, "hipEventCreate failed");
        }
    s->Tmax_snap.assign(cfg->nsnap, 0.f);
    s->Tlo_snap.assign(cfg->nsnap, 0.f);
    s->Thi_snap.assign(cfg->nsnap, 0.f);
    s->field_set.assign(cfg->nsnap, 0);
    s->slab_open.assign(cfg->nsnap, 0);
    rc = upload_geometry(s);
    if (rc) { dnsmc_destroy(s); return rc; }
    // default emitters: every cell
    long long n = (long long)cfg->n[0] * cfg->n[1] * cfg->n[2];
    s->emit.resize(n);
    for (long long i = 0; i < n; ++i) s->emit[i] = (int)i;
    *out = s;
    return DNSMC_OK;
}

void dnsmc_destroy(dnsmc_solver* s)
{
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto* b : s->sp_dbufs) delete b;
    for (auto* b : s->sp_ibufs) delete b;
    for (auto& e : s->evs) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (int p = 1; p < dnsmc_solver::kPipes; ++p)
        if (s->pipe[p].stream) {
            (void)hipStreamSynchronize(s->pipe[p].stream);
            (void)hipStreamDestroy(s->pipe[p].stream);
        }
    for (int p = 0; p < dnsmc_solver::kPipes; ++p)
        if (s->pipe[p].done) (void)hipEventDestroy(s->pipe[p].done);
    if (s->ev_ready) (void)hipEventDestroy(s->ev_ready);
    for (int b = 0; b < 2; ++b) {
        if (s->ev_gen[b]) (void)hipEventDestroy(s->ev_gen[b]);
        if (s->ev_tr[b]) (void)hipEventDestroy(s->ev_tr[b]);
    }
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int dnsmc_probe_tables(const dnsmc_config* cfg, int* nB, int* nQ, int* nT)
{
    if (!cfg) return fail(DNSMC_EINVAL, "null argument");
    Tables t;
    std::string err;
    int rc = read_tables(*cfg, t, &err);
    if (rc) return fail(rc, err);
    if (nB) *nB = t.synth_grey ? 1000 : t.nB;
    if (nQ) *nQ = t.synth_grey ? 1 : t.nQ;
    if (nT) *nT = t.synth_grey ? 55 : t.nT;
    return DNSMC_OK;
}

int dnsmc_load_tables(dnsmc_solver* s)
{
    if (!s) return fail(DNSMC_EINVAL, "null solver");
    std::string err;
    int rc = read_tables(s->cfg, s->tab, &err);
    if (rc) return fail(rc, err);
    s->have_tables = true;
    s->der_Tmax = -1.f;
    return DNSMC_OK;
}

int dnsmc_set_tables(dnsmc_solver* s, int nB, int nQ, int nT, const float* Tnb, const float* kP, const float* wvl,
                     const float* wvr, const float* wvc, const float* kq, const float* prob, const float* probg,
                     const float* kavg)
{
    if (!s || nB < 1 || nQ < 1 || nT < 2 || !Tnb || !kP || !wvl || !wvr || !wvc || !kq || !prob || !probg)
        return fail(DNSMC_EINVAL, "dnsmc_set_tables: bad arguments");
    if (nB > 1024 || nQ > 64) return fail(DNSMC_EINVAL, "at most 1024 bands and 64 quadrature points");
    Tables& t = s->tab;
    t = Tables();
    t.nB = nB; t.nQ = nQ; t.nT = nT;
    t.Tnb.assign(Tnb, Tnb + nT);
    t.kP.assign(kP, kP + nT);
    t.wvl.assign(wvl, wvl + nB);
    t.wvr.assign(wvr, wvr + nB);
    t.wvc.assign(wvc, wvc + nB);
    t.kq.assign(kq, kq + (size_t)nB * nT * nQ);
    t.prob.assign(prob, prob + (size_t)nT * nB);
    t.probg.assign(probg, probg + (size_t)nT * nB * nQ);
    t.kavg.assign(nB, 0.f);
    for (int b = 0; b < nB; ++b) {
        if (kavg) {
            t.kavg[b] = kavg[b];
        } else {
            double m = 0.0;
            for (int g = 0; g < nQ; ++g) m += t.kq[((size_t)b * nT + nT - 1) * nQ + g];
            t.kavg[b] = (float)(m / nQ);
        }
    }
    s->have_tables = true;
    s->der_Tmax = -1.f;
    return DNSMC_OK;
}

int dnsmc_set_field(dnsmc_solver* s, int snap, const float* T, int is_device)
{
    if (!s || !T) return fail(DNSMC_EINVAL, "null argument");
    if (snap < 0 || snap >= s->cfg.nsnap) return fail(DNSMC_EINVAL, "snapshot index out of range");
    if (s->running) {
        int rc = do_wait(s);
        if (rc) return rc;
    }
    const int n0 = s->cfg.n[0], n1 = s->cfg.n[1], n2 = s->cfg.n[2];
    const size_t gsz = (size_t)(n0 + 2) * (n1 + 2) * (n2 + 2);
    // the field as given (ghost layers included) to the device, then its
    // interior into the bricked layout by a kernel; the positivity / NaN check
    // and Tmax come from the commit's range pass (code/mc_gpu.cu:456-480)
    const float* src = T;
    if (!is_device) {
        HIPCHK(s->d_stage.ensure(gsz));
        HIPCHK(hipMemcpyAsync(s->d_stage.p, T, gsz * sizeof(float), hipMemcpyHostToDevice, s->stream));
        src = s->d_stage.p;
    }
    const long long nc = (long long)n0 * n1 * n2;
    const int nb = (int)std::min<long long>((nc + 255) / 256, (long long)s->ncu * 16);
    hipLaunchKernelGGL(ghost_to_bricks, dim3(nb), dim3(256), 0, s->stream, src, n0, n1, n2,
                       s->d_T[0].p + (size_t)snap * s->tstride[0], s->nbx[0], s->nby[0]);
    HIPCHK(hipGetLastError());
    s->slab_open[snap] = 0;
    return commit_field(s, snap);
}

int dnsmc_fix_temp_slab(dnsmc_solver* s, int snap, const double* cnew, const int* startc, const int* endc, double Tc,
                        double Tplus)
{
    if (!s || !cnew || !startc || !endc) return fail(DNSMC_EINVAL, "null argument");
    if (snap < 0 || snap >= s->cfg.nsnap) return fail(DNSMC_EINVAL, "snapshot index out of range");
    if (!(Tplus != 0.0) || !std::isfinite(Tplus) || !std::isfinite(Tc)) return fail(DNSMC_EINVAL, "Tc and Tplus must be finite, Tplus nonzero");
    const int* n = s->cfg.n;
    int len[3];
    for (int a = 0; a < 3; ++a) {
        if (startc[a] < 1 || endc[a] > n[a] || endc[a] < startc[a])
            return fail(DNSMC_EINVAL, "slab index range outside 1..n or empty");
        len[a] = endc[a] - startc[a] + 1;
    }
    if (s->running) {
        int rc = do_wait(s);
        if (rc) return rc;
    }
    float* T = s->d_T[0].p + (size_t)snap * s->tstride[0];
    if (!s->slab_open[snap]) {
        // first slab since the last commit: every cell starts uncovered (NaN),
        // so a field the slabs do not cover fails the commit's check
        const long long nf = s->tstride[0];
        hipLaunchKernelGGL(fill_u32, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s->stream, (unsigned*)T, nf,
                           0x7FC00000u);
        HIPCHK(hipGetLastError());
        s->slab_open[snap] = 1;
        s->field_set[snap] = 0;
    }
    // cnew(0:n0+1, nj, nk) of the rank, i fastest
    const size_t cnt = (size_t)(n[0] + 2) * len[1] * len[2];
    HIPCHK(s->d_slab.ensure(cnt));
    HIPCHK(hipMemcpyAsync(s->d_slab.p, cnew, cnt * sizeof(double), hipMemcpyHostToDevice, s->stream));
    const long long nc = (long long)len[0] * len[1] * len[2];
    const int nb = (int)std::min<long long>((nc + 255) / 256, (long long)s->ncu * 16);
    hipLaunchKernelGGL(slab_to_field, dim3(nb), dim3(256), 0, s->stream, (const double*)s->d_slab.p, n[0], n[1], n[2],
                       startc[0], startc[1], startc[2], len[0], len[1], len[2], Tc, Tplus, T, s->nbx[0], s->nby[0]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));   // cnew and the staging buffer are free again
    return DNSMC_OK;
}

int dnsmc_commit_field(dnsmc_solver* s, int snap)
{
    if (!s) return fail(DNSMC_EINVAL, "null solver");
    if (snap < 0 || snap >= s->cfg.nsnap) return fail(DNSMC_EINVAL, "snapshot index out of range");
    if (!s->slab_open[snap]) return fail(DNSMC_ESTATE, "no slab written since the last commit");
    if (s->running) {
        int rc = do_wait(s);
        if (rc) return rc;
    }
    return commit_field(s, snap);
}

int dnsmc_fix_rad_slab(dnsmc_solver* s, int snap, const int* startc, const int* endc, double Tc, double* divQ,
                       double* vcore)
{
    if (!s || !startc || !endc || !divQ || !vcore) return fail(DNSMC_EINVAL, "null argument");
    if (!(Tc > 0.0) || !std::isfinite(Tc)) return fail(DNSMC_EINVAL, "Tc must be positive and finite");
    const int* n = s->cfg.n;
    if (startc[2] < 1 || endc[2] > n[2] || endc[2] < startc[2]) return fail(DNSMC_EINVAL, "k range outside 1..n2 or empty");
    int rc = do_wait(s);
    if (rc) return rc;
    if (snap < 0 || snap >= s->cfg.nsnap || s->last_nemit == 0) return fail(DNSMC_ESTATE, "no results for snapshot");
    const int nj = n[1] / s->cfg.p_row, nk = endc[2] - startc[2] + 1;
    const size_t cnt = (size_t)(n[0] + 2) * nj * nk;   // divQ(0:n0+1, nj, nk)
    HIPCHK(s->d_slab.ensure(2 * cnt));
    double* dq = s->d_slab.p;
    double* dv = s->d_slab.p + cnt;
    hipLaunchKernelGGL(fill_f64, dim3((unsigned)((2 * cnt + 255) / 256)), dim3(256), 0, s->stream, dq, (long long)(2 * cnt),
                       0.0);
    const long long ne = s->last_nemit;
    const double stefan = 5.670373e-08;   // code/param.txt:19
    const double tc4 = Tc * Tc * Tc * Tc;
    hipLaunchKernelGGL(field_to_slab, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s->stream,
                       (const int*)s->d_emit.p, ne, (const double*)(s->d_Q.p + (size_t)snap * ne),
                       (const double*)(s->d_sd.p + (size_t)snap * ne), n[0], n[1], s->last_ystart - 1, nj, startc[2] - 1, nk,
                       stefan, tc4, dq, dv);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(divQ, dq, cnt * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(vcore, dv, cnt * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return DNSMC_OK;
}

int dnsmc_set_mlp(dnsmc_solver* s, const unsigned short* W1, const float* b1, const unsigned short* W2, const float* b2,
                  const unsigned short* W3, const float* b3, float mix_eps)
{
    if (!s || !W1 || !b1 || !W2 || !b2 || !W3 || !b3) return fail(DNSMC_EINVAL, "null argument");
    if (!s->have_tables) return fail(DNSMC_ESTATE, "load the spectral tables before the proposal network");
    if (!(mix_eps > 0.f && mix_eps <= 1.f)) return fail(DNSMC_EINVAL, "mix_eps must be in (0, 1]");
    if (s->running) {
        int rc = do_wait(s);
        if (rc) return rc;
    }
    const int nbin = s->cfg.spectral == DNSMC_SPEC_GREY_T4 ? 1 : s->tab.nB * s->tab.nQ;
    const int noutp = mlp_noutp(nbin);
    auto up16 = [&](DevBuf<uint16_t>& d, const unsigned short* h, size_t n) {
        return d.upload(std::vector<uint16_t>(h, h + n), s->stream);
    };
    auto up32 = [&](DevBuf<float>& d, const float* h, size_t n) {
        return d.upload(std::vector<float>(h, h + n), s->stream);
    };
    HIPCHK(up16(s->d_W1b, W1, (size_t)kMlpIn * kMlpHid));
    HIPCHK(up32(s->d_b1, b1, kMlpHid));
    HIPCHK(up16(s->d_W2b, W2, (size_t)kMlpHid * kMlpHid));
    HIPCHK(up32(s->d_b2, b2, kMlpHid));
    HIPCHK(up16(s->d_W3b, W3, (size_t)kMlpHid * noutp));
    HIPCHK(up32(s->d_b3, b3, noutp));
    {   // the transposed copies: [out][in]
        std::vector<uint16_t> t2((size_t)kMlpHid * kMlpHid), t3((size_t)noutp * kMlpHid);
        for (int k = 0; k < kMlpHid; ++k)
            for (int n = 0; n < kMlpHid; ++n) t2[(size_t)n * kMlpHid + k] = W2[(size_t)k * kMlpHid + n];
        for (int k = 0; k < kMlpHid; ++k)
            for (int n = 0; n < noutp; ++n) t3[(size_t)n * kMlpHid + k] = W3[(size_t)k * noutp + n];
        HIPCHK(s->d_W2T.upload(t2, s->stream));
        HIPCHK(s->d_W3T.upload(t3, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));   // (t2, t3 are read until the copies end)
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    s->have_mlp = true;
    s->mlp_nbin = nbin;
    s->mix_eps = mix_eps;
    return DNSMC_OK;
}

int dnsmc_set_emitters_slab(dnsmc_solver* s, int ystart)
{
    if (!s) return fail(DNSMC_EINVAL, "null solver");
    const int n0 = s->cfg.n[0], n1 = s->cfg.n[1], n2 = s->cfg.n[2];
    const int nj = n1 / s->cfg.p_row;
    if (ystart < 1 || ystart + nj - 1 > n1) return fail(DNSMC_EINVAL, "ystart out of range");
    if (s->emit_is_slab && s->ystart == ystart && !s->emit.empty()) return DNSMC_OK;   // unchanged: keep the slot tables
    s->emit.clear();
    for (int k = 0; k < n2; ++k)
        for (int j = ystart - 1; j < ystart - 1 + nj; ++j)
            for (int i = 0; i < n0; ++i) s->emit.push_back(i + n0 * (j + n1 * k));
    s->emit_dirty = true;
    s->emit_is_slab = true;
    s->ystart = ystart;
    return DNSMC_OK;
}

int dnsmc_set_emitters_list(dnsmc_solver* s, const int* cells, int ncells)
{
    if (!s || !cells || ncells < 1) return fail(DNSMC_EINVAL, "bad emitter list");
    long long ntot = (long long)s->cfg.n[0] * s->cfg.n[1] * s->cfg.n[2];
    for (int i = 0; i < ncells; ++i)
        if (cells[i] < 0 || cells[i] >= ntot) return fail(DNSMC_EINVAL, "emitter cell out of range");
    s->emit.assign(cells, cells + ncells);
    s->emit_dirty = true;
    s->emit_is_slab = false;
    s->ystart = 1;
    return DNSMC_OK;
}

int dnsmc_run(dnsmc_solver* s)
{
    if (!s) return fail(DNSMC_EINVAL, "null solver");
    if (s->running) {
        int rc = do_wait(s);
        if (rc) return rc;
    }
    return do_run(s, nullptr);
}

int dnsmc_run_noalloc(dnsmc_solver* s, double* Q, double* sdev)
{
    if (!s || !Q || !sdev) return fail(DNSMC_EINVAL, "null argument");
    if (s->running) {
        int rc = do_wait(s);
        if (rc) return rc;
    }
    NoAlloc na{s->scratch, Q, sdev};
    return do_run(s, &na);
}

int dnsmc_wait(dnsmc_solver* s)
{
    if (!s) return fail(DNSMC_EINVAL, "null solver");
    return do_wait(s);
}

int dnsmc_get_results(dnsmc_solver* s, int snap, double* Q, double* E, double* A, double* sdev)
{
    if (!s) return fail(DNSMC_EINVAL, "null solver");
    int rc = do_wait(s);
    if (rc) return rc;
    if (snap < 0 || snap >= s->cfg.nsnap || s->last_nemit == 0) return fail(DNSMC_ESTATE, "no results for snapshot");
    size_t n = (size_t)s->last_nemit, o = (size_t)snap * n;
    if (Q) HIPCHK(hipMemcpy(Q, s->d_Q.p + o, n * sizeof(double), hipMemcpyDeviceToHost));
    if (E) HIPCHK(hipMemcpy(E, s->d_Eo.p + o, n * sizeof(double), hipMemcpyDeviceToHost));
    if (A) HIPCHK(hipMemcpy(A, s->d_Ao.p + o, n * sizeof(double), hipMemcpyDeviceToHost));
    if (sdev) HIPCHK(hipMemcpy(sdev, s->d_sd.p + o, n * sizeof(double), hipMemcpyDeviceToHost));
    return DNSMC_OK;
}

// code/mc_gpu.cu:762-776: results into (0:imax+1, 0:jmax/p_row+1, 0:kmax+1).
int dnsmc_get_results_fortran(dnsmc_solver* s, int snap, float* resfort, float* varfort)
{
    if (!s || !resfort || !varfort) return fail(DNSMC_EINVAL, "null argument");
    int rc = do_wait(s);
    if (rc) return rc;
    if (snap < 0 || snap >= s->cfg.nsnap || s->last_nemit == 0) return fail(DNSMC_ESTATE, "no results for snapshot");
    const int n0 = s->cfg.n[0], n1 = s->cfg.n[1], n2 = s->cfg.n[2];
    const int nj = n1 / s->cfg.p_row;
    const size_t fsz = (size_t)(n0 + 2) * (nj + 2) * (n2 + 2);
    // the caller's arrays round-trip through the device so that the entries
    // the results do not cover (ghost layers) keep their values; the interior
    // is written by a kernel from the FP64 results
    HIPCHK(s->d_stage.ensure(2 * fsz));
    float* dq = s->d_stage.p;
    float* dv = s->d_stage.p + fsz;
    HIPCHK(hipMemcpyAsync(dq, resfort, fsz * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(dv, varfort, fsz * sizeof(float), hipMemcpyHostToDevice, s->stream));
    const long long ne = s->last_nemit;
    hipLaunchKernelGGL(results_to_fortran, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s->stream,
                       (const int*)s->d_emit.p, ne, (const double*)(s->d_Q.p + (size_t)snap * ne),
                       (const double*)(s->d_sd.p + (size_t)snap * ne), n0, n1, s->last_ystart - 1, nj, n2, dq, dv);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(resfort, dq, fsz * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(varfort, dv, fsz * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return DNSMC_OK;
}

int dnsmc_get_stats(dnsmc_solver* s, dnsmc_stats* st)
{
    if (!s || !st) return fail(DNSMC_EINVAL, "null argument");
    int rc = do_wait(s);
    if (rc) return rc;
    *st = s->st;
    return DNSMC_OK;
}

int dnsmc_get_array(dnsmc_solver* s, const char* name, int level, float* out, long long n)
{
    if (!s || !name) return fail(DNSMC_EINVAL, "null argument");
    int rc = do_wait(s);
    if (rc) return rc;
    std::string nm = name;
    std::vector<float> v;
    if (nm == "faces0" || nm == "faces1" || nm == "faces2") {
        if (level < 0 || level >= (int)s->lev.size()) return fail(DNSMC_EINVAL, "level out of range");
        v = s->lev[level].faces[nm[5] - '0'];
    } else if (nm == "centres0" || nm == "centres1" || nm == "centres2") {
        v = s->lev[level].centres[nm[7] - '0'];
    } else if (nm == "T") {
        if (level < 0 || level >= (int)s->lev.size()) return fail(DNSMC_EINVAL, "level out of range");
        // linear (i fastest) per snapshot, from the bricked device layout
        std::vector<float> b((size_t)s->tstride[level] * s->cfg.nsnap);
        HIPCHK(hipMemcpy(b.data(), s->d_T[level].p, b.size() * sizeof(float), hipMemcpyDeviceToHost));
        const int* ln = s->lev[level].n;
        v.resize((size_t)s->ncell[level] * s->cfg.nsnap);
        size_t o = 0;
        for (int sn = 0; sn < s->cfg.nsnap; ++sn)
            for (int k = 0; k < ln[2]; ++k)
                for (int j = 0; j < ln[1]; ++j)
                    for (int i = 0; i < ln[0]; ++i)
                        v[o++] = b[(size_t)sn * s->tstride[level] + brick_index(i, j, k, s->nbx[level], s->nby[level])];
    } else if (nm == "qcdf") {
        // the learned proposal's cumulative weights of the last batch, one row
        // of mlp_noutp(nB*nQ) per slot, rows in emitter order (the batch must
        // cover every slot: one snapshot, batched generation)
        const dnsmc_solver::Pipe& p = s->pipe[0].d_qcdf.n ? s->pipe[0] : s->pipe[1];
        const long long nsl = (long long)s->emit.size() * spe_of(s->cfg);
        const long long row = mlp_noutp(s->mlp_nbin);
        if (!s->have_mlp || s->emit_dirty || (long long)p.d_qcdf.n < nsl * row)
            return fail(DNSMC_ESTATE, "no learned batch covering every slot");
        std::vector<float> b((size_t)(nsl * row));
        HIPCHK(hipMemcpy(b.data(), p.d_qcdf.p, b.size() * sizeof(float), hipMemcpyDeviceToHost));
        v.resize(b.size());
        for (long long ps = 0; ps < nsl; ++ps)
            memcpy(v.data() + (size_t)s->order[ps] * row, b.data() + (size_t)ps * row, (size_t)row * sizeof(float));
    } else if (nm == "lrows") {
        // the learned proposal rows the traversal evaluated in the last run
        // with the draw log on (cumulative weights U, mlp_noutp(nbin) floats
        // per slot), rows in emitter order
        const long long nsl = (long long)s->emit.size() * spe_of(s->cfg);
        const long long row = mlp_noutp(s->mlp_nbin);
        if (!s->have_mlp || s->emit_dirty || (long long)s->d_lrows.n < nsl * row)
            return fail(DNSMC_ESTATE, "no logged learned run covering every slot (dnsmc_set_draw_log)");
        std::vector<float> b((size_t)(nsl * row));
        HIPCHK(hipMemcpy(b.data(), s->d_lrows.p, b.size() * sizeof(float), hipMemcpyDeviceToHost));
        v.resize(b.size());
        for (long long ps = 0; ps < nsl; ++ps)
            memcpy(v.data() + (size_t)s->order[ps] * row, b.data() + (size_t)ps * row, (size_t)row * sizeof(float));
    } else if (nm == "features") {
        // the learned proposal's input features of every slot of the last run
        // (kMlpIn floats per slot), rows in emitter order
        if (s->der_Tmax <= 0.f || s->emit_dirty) return fail(DNSMC_ESTATE, "run the solver first");
        const long long nsl = (long long)s->emit.size() * spe_of(s->cfg);
        DevBuf<float> dX;
        HIPCHK(dX.ensure((size_t)(nsl * kMlpIn)));
        const FeatArgs f = feat_args(s, s->der);
        const int nb = (int)std::min<long long>((nsl + 255) / 256, (long long)s->ncu * 64);
        hipLaunchKernelGGL(mlp_features, dim3(nb), dim3(256), 0, 0, f, (const int4*)s->d_slotinfo.p, nsl, dX.p);
        HIPCHK(hipGetLastError());
        std::vector<float> b((size_t)(nsl * kMlpIn));
        HIPCHK(hipMemcpy(b.data(), dX.p, b.size() * sizeof(float), hipMemcpyDeviceToHost));
        v.resize(b.size());
        for (long long ps = 0; ps < nsl; ++ps)
            memcpy(v.data() + (size_t)s->order[ps] * kMlpIn, b.data() + (size_t)ps * kMlpIn, kMlpIn * sizeof(float));
    } else if (nm == "p_bg" || nm == "pcdf") {
        const DevBuf<float>& db = nm == "p_bg" ? s->d_pbg : s->d_pcdf;
        v.resize(db.n);
        if (db.n) HIPCHK(hipMemcpy(v.data(), db.p, db.n * sizeof(float), hipMemcpyDeviceToHost));
    } else if (nm == "cdf_band") v = s->der.cdf_band;
    else if (nm == "cdf_g") v = s->der.cdf_g;
    else if (nm == "inv_emax") v = s->der.inv_emax;
    else if (nm == "Ib") v = s->der.Ib;
    else if (nm == "Ibw") v = s->der.Ibw;
    else if (nm == "kq") v = s->der.kq;
    else if (nm == "scalars") {
        const Derived& d = s->der;
        v = {d.Tmax, d.kappamax, d.Ibmax, (float)d.nB, (float)d.nQ, (float)d.nT, d.inv_dT, d.mT0,
             (float)d.g_search, d.u_off, (float)(-log((double)s->cfg.toll))};
    } else if (nm == "Ibmax_d") {
        double x = s->der.Ibmax_d;
        if (n >= 2) memcpy(out, &x, sizeof x);
        return 2;
    } else return fail(DNSMC_EINVAL, "unknown array " + nm);
    long long c = std::min<long long>(n, (long long)v.size());
    if (out && c > 0) memcpy(out, v.data(), c * sizeof(float));
    return (int)v.size();
}

long long dnsmc_get_rays(dnsmc_solver* s, unsigned int* w0, unsigned int* w1, float* w, float* winv, long long n)
{
    if (!s) return fail(DNSMC_EINVAL, "null solver");
    int rc = do_wait(s);
    if (rc) return rc;
    long long c = s->first_batch_rays;
    if (c <= 0) return 0;
    if (s->cfg.nsnap > 1 || (long long)s->emit.size() * spe_of(s->cfg) * rays_per_slot_of(s->cfg) > c)
        return fail(DNSMC_ESTATE, "ray buffer holds the last batch only; use one batch for ray inspection");
    if (n < c) return fail(DNSMC_EINVAL, "output arrays must hold every ray of the batch");
    std::vector<long long> perm;
    slot_perm(s, c, perm);
    const dnsmc_solver::Pipe& p = s->pipe[0];
    if ((rc = fetch_perm<uint32_t>(p.d_w0.p, c, perm, w0))) return rc;
    if ((rc = fetch_perm<uint32_t>(p.d_w1.p, c, perm, w1))) return rc;
    if ((rc = fetch_perm<float>(p.d_w.p, c, perm, w))) return rc;
    if ((rc = fetch_perm<float>(p.d_winv.p, c, perm, winv))) return rc;
    return c;
}

long long dnsmc_get_ray_records(dnsmc_solver* s, int* steps, int* end_type, unsigned int* hash, float* D, float* G,
                                long long n)
{
    if (!s) return fail(DNSMC_EINVAL, "null solver");
    int rc = do_wait(s);
    if (rc) return rc;
    if (!s->record) return fail(DNSMC_ESTATE, "records are off (dnsmc_set_record)");
    long long c = std::min(s->nrec, s->first_batch_rays);
    if (c <= 0) return 0;
    if (c < s->first_batch_rays || (long long)s->emit.size() * spe_of(s->cfg) * rays_per_slot_of(s->cfg) > c)
        return fail(DNSMC_ESTATE, "records cover the whole solve only when it fits one batch of <= 2^22 rays");
    if (n < c) return fail(DNSMC_EINVAL, "output arrays must hold every record");
    std::vector<long long> perm;
    slot_perm(s, c, perm);
    if ((rc = fetch_perm<int>(s->d_rsteps.p, c, perm, steps))) return rc;
    if ((rc = fetch_perm<int>(s->d_rend.p, c, perm, end_type))) return rc;
    if ((rc = fetch_perm<uint32_t>(s->d_rhash.p, c, perm, (uint32_t*)hash))) return rc;
    if ((rc = fetch_perm<float>(s->d_rD.p, c, perm, D))) return rc;
    if ((rc = fetch_perm<float>(s->d_rG.p, c, perm, G))) return rc;
    return c;
}

int dnsmc_set_draw_log(dnsmc_solver* s, int on)
{
    if (!s) return fail(DNSMC_EINVAL, "null solver");
    s->draw_log = on ? 1 : 0;
    return DNSMC_OK;
}

int dnsmc_set_record(dnsmc_solver* s, int on)
{
    if (!s) return fail(DNSMC_EINVAL, "null solver");
    s->record = on ? 1 : 0;
    return DNSMC_OK;
}

int dnsmc_bench_copy(long long nbytes, int reps, double* gbps)
{
    if (nbytes < 16 || reps < 1 || !gbps) return fail(DNSMC_EINVAL, "bad arguments");
    long long n4 = nbytes / 16;
    float4 *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc((void**)&a, n4 * 16));
    HIPCHK(hipMalloc((void**)&b, n4 * 16));
    HIPCHK(hipMemset(a, 0, n4 * 16));
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    hipLaunchKernelGGL(copy_f32x4, dim3(ncu * 8), dim3(256), 0, 0, a, b, n4);   // warm-up
    HIPCHK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(copy_f32x4, dim3(ncu * 8), dim3(256), 0, 0, a, b, n4);
    HIPCHK(hipEventRecord(e1, 0));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *gbps = 2.0 * (double)n4 * 16.0 * reps / (ms * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return DNSMC_OK;
}

int dnsmc_bench_stream(long long nbytes, int reps, int kind, double* gbps, int* best)
{
    if (nbytes < 16 || reps < 1 || !gbps || kind < 0 || kind > 1) return fail(DNSMC_EINVAL, "bad arguments");
    const long long n4 = nbytes / 16;
    float4 *a = nullptr, *b = nullptr;
    float* sink = nullptr;
    HIPCHK(hipMalloc((void**)&a, n4 * 16));
    if (kind == 0) HIPCHK(hipMalloc((void**)&b, n4 * 16));
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int per_cu[4] = {4, 8, 16, 32};
    HIPCHK(hipMalloc((void**)&sink, (size_t)ncu * 32 * 256 * sizeof(float)));
    HIPCHK(hipMemset(a, 0, n4 * 16));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // the variants: 16 B x {1, 4} per lane and iteration, plain or non-temporal,
    // 4 to 32 blocks of 256 threads per CU; the best rate is the ceiling
    double top = 0.0;
    int top_v = -1;
    for (int v = 0; v < 12; ++v) {
        const int shape = v % 3, nbk = ncu * per_cu[v / 3];
        auto launch = [&]() {
            if (kind == 0) {
                if (shape == 0) hipLaunchKernelGGL((stream_copy<1, false>), dim3(nbk), dim3(256), 0, 0, a, b, n4);
                else if (shape == 1) hipLaunchKernelGGL((stream_copy<4, false>), dim3(nbk), dim3(256), 0, 0, a, b, n4);
                else hipLaunchKernelGGL((stream_copy<4, true>), dim3(nbk), dim3(256), 0, 0, a, b, n4);
            } else {
                if (shape == 0) hipLaunchKernelGGL((stream_read<1, false>), dim3(nbk), dim3(256), 0, 0, a, n4, sink);
                else if (shape == 1) hipLaunchKernelGGL((stream_read<4, false>), dim3(nbk), dim3(256), 0, 0, a, n4, sink);
                else hipLaunchKernelGGL((stream_read<4, true>), dim3(nbk), dim3(256), 0, 0, a, n4, sink);
            }
        };
        launch();   // warm-up
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(e0, 0));
        for (int r = 0; r < reps; ++r) launch();
        HIPCHK(hipEventRecord(e1, 0));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        const double g = (kind == 0 ? 2.0 : 1.0) * (double)n4 * 16.0 * reps / (ms * 1e-3) / 1e9;
        if (g > top) {
            top = g;
            top_v = v;
        }
    }
    *gbps = top;
    if (best) *best = top_v;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    if (b) (void)hipFree(b);
    (void)hipFree(sink);
    return DNSMC_OK;
}

int dnsmc_sync(void)
{
    HIPCHK(hipDeviceSynchronize());
    return DNSMC_OK;
}

// ------------------------------------------------------ device buffers ----
// Plain device allocations for callers that hold training data on the GPU
// without another runtime (bindings, tests).
int dnsmc_dev_alloc(long long bytes, void** out)
{
    if (!out || bytes < 0) return fail(DNSMC_EINVAL, "bad arguments");
    *out = nullptr;
    HIPCHK(hipMalloc(out, (size_t)std::max<long long>(bytes, 1)));
    return DNSMC_OK;
}

int dnsmc_dev_free(void* p)
{
    if (p) HIPCHK(hipFree(p));
    return DNSMC_OK;
}

// kind: 1 host -> device, 2 device -> host, 3 device -> device (synchronous)
int dnsmc_dev_copy(void* dst, const void* src, long long bytes, int kind)
{
    if (!dst || !src || bytes < 0 || kind < 1 || kind > 3) return fail(DNSMC_EINVAL, "bad arguments");
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    HIPCHK(hipMemcpy(dst, src, (size_t)bytes, k));
    return DNSMC_OK;
}

// ------------------------------------------- learned proposal training ----
// All pointers below are device pointers; the calls run on the null stream
// and return when their results are ready.

long long dnsmc_mlp_nparams(int nbin) { return nbin < 1 ? -1 : mlp_layout(nbin).P; }

int dnsmc_mlp_samples(dnsmc_solver* s, float* X, int* td, int* tb, float* wt, long long n, float cost0)
{
    if (!s || !X || !td || !tb || !wt) return fail(DNSMC_EINVAL, "null argument");
    int rc = do_wait(s);
    if (rc) return rc;
    const dnsmc_config& c = s->cfg;
    if (!s->record) return fail(DNSMC_ESTATE, "samples need a run with records on (dnsmc_set_record)");
    const long long nr = s->first_batch_rays;
    if (nr <= 0 || nr > s->nrec || c.nsnap != 1 ||
        (long long)s->emit.size() * spe_of(c) * rays_per_slot_of(c) != nr)
        return fail(DNSMC_ESTATE, "samples need the last run to be one snapshot in one batch of <= 2^22 rays");
    if (n < nr) return fail(DNSMC_EINVAL, "sample arrays must hold every ray of the run");
    const dnsmc_solver::Pipe& p = s->pipe[0];
    FeatArgs f = feat_args(s, s->der);
    int nb = (int)std::min<long long>((nr + 255) / 256, (long long)s->ncu * 64);
    hipLaunchKernelGGL(mlp_samples, dim3(nb), dim3(256), 0, 0, f, (const int4*)s->d_slotinfo.p, 0ll,
                       rays_per_slot_of(c), nr, s->der.nQ, (const uint32_t*)p.d_w0.p, (const uint32_t*)p.d_w1.p,
                       (const float*)p.d_w.p, (const float*)p.d_winv.p, (const float*)s->d_rD.p,
                       (const float*)s->d_rG.p, (const int*)s->d_rsteps.p, cost0, X, td, tb, wt);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return (int)std::min<long long>(nr, 0x7FFFFFFF);
}

int dnsmc_mlp_features(dnsmc_solver* s, float* X, long long n)
{
    if (!s || !X) return fail(DNSMC_EINVAL, "null argument");
    int rc = do_wait(s);
    if (rc) return rc;
    if (s->der_Tmax <= 0.f || s->emit_dirty) return fail(DNSMC_ESTATE, "run the solver first");
    const long long ns = (long long)s->order.size() * spe_of(s->cfg);
    n = std::min(n, ns);
    FeatArgs f = feat_args(s, s->der);
    int nb = (int)std::min<long long>((n + 255) / 256, (long long)s->ncu * 64);
    hipLaunchKernelGGL(mlp_features, dim3(nb), dim3(256), 0, 0, f, (const int4*)s->d_slotinfo.p, n, X);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return (int)n;
}

static MlpWeights mlp_weights_of(int nbin, const unsigned short* W1b, const unsigned short* W2b,
                                 const unsigned short* W3b, const float* theta)
{
    const MlpLayout L = mlp_layout(nbin);
    MlpWeights w;
    w.W1 = W1b; w.W2 = W2b; w.W3 = W3b;
    w.b1 = theta + L.ob1; w.b2 = theta + L.ob2; w.b3 = theta + L.ob3;
    w.nbin = nbin;
    w.noutp = L.noutp;
    w.W2T = nullptr;   // (the training and row-forward kernels read W2, W3)
    w.W3T = nullptr;
    return w;
}

// Loss and gradient of the weighted cross-entropy over n samples (weights as
// BF16 copies, biases inside theta).  work: at least
// ceil(n/64) * (P + 1) + P + 2 floats.
int dnsmc_mlp_grad(int nbin, const float* theta, const unsigned short* W1b, const unsigned short* W2b,
                   const unsigned short* W3b, const float* X, const int* td, const int* tb, const float* wt, long long n,
                   float* work, long long nwork, float* grad, float* loss)
{
    if (nbin < 1 || !theta || !W1b || !W2b || !W3b || !X || !td || !tb || !wt || !work || !grad || n < 1)
        return fail(DNSMC_EINVAL, "bad arguments");
    const MlpLayout L = mlp_layout(nbin);
    const long long nblk = (n + 63) / 64;
    if (nblk > 0x7FFFFFFF || nwork < nblk * (L.P + 1) + L.P + 2) return fail(DNSMC_EINVAL, "work buffer too small");
    float* part = work;
    float* lpart = part + nblk * L.P;
    float* wsum = lpart + nblk;
    float* dloss = wsum + 1;
    MlpTrainArgs a;
    a.X = X; a.td = td; a.tb = tb; a.wt = wt; a.n = (int)std::min<long long>(n, 0x7FFFFFFF);
    a.wsum = wsum;
    a.w = mlp_weights_of(nbin, W1b, W2b, W3b, theta);
    a.part = part;
    a.loss_part = lpart;
    hipLaunchKernelGGL(mlp_wsum, dim3(1), dim3(256), 0, 0, wt, n, wsum);
    const size_t lds = sizeof(float) * (64 * kMlpIn + 3 * 4096 + 8 * 64 + 2 * 64) + 5 * 4096 * sizeof(uint16_t);
    hipLaunchKernelGGL(mlp_train_tile, dim3((unsigned)nblk), dim3(256), lds, 0, a);
    hipLaunchKernelGGL(mlp_reduce, dim3((L.P + 255) / 256), dim3(256), 0, 0, part, lpart, (int)nblk, L.P, grad, dloss);
    HIPCHK(hipGetLastError());
    if (loss) HIPCHK(hipMemcpy(loss, dloss, sizeof(float), hipMemcpyDeviceToHost));
    else HIPCHK(hipDeviceSynchronize());
    return DNSMC_OK;
}

// One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on theta; refreshes the
// BF16 weight copies.  step counts from 1.
int dnsmc_mlp_adam(int nbin, float* theta, float* m, float* v, const float* grad, float lr, int step,
                   unsigned short* W1b, unsigned short* W2b, unsigned short* W3b)
{
    if (nbin < 1 || !theta || !m || !v || !grad || !W1b || !W2b || !W3b || step < 1)
        return fail(DNSMC_EINVAL, "bad arguments");
    const MlpLayout L = mlp_layout(nbin);
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = (float)(1.0 / (1.0 - pow((double)b1, step))), c2 = (float)(1.0 / (1.0 - pow((double)b2, step)));
    hipLaunchKernelGGL(mlp_adam, dim3((L.P + 255) / 256), dim3(256), 0, 0, theta, m, v, grad, L.P, L, lr, b1, b2, 1e-8f,
                       c1, c2, W1b, W2b, W3b);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return DNSMC_OK;
}

// Logits Z[n][noutp] of the network for inputs X[n][6].
int dnsmc_mlp_forward(int nbin, const float* theta, const unsigned short* W1b, const unsigned short* W2b,
                      const unsigned short* W3b, const float* X, long long n, float* Z)
{
    if (nbin < 1 || !theta || !W1b || !W2b || !W3b || !X || !Z || n < 1) return fail(DNSMC_EINVAL, "bad arguments");
    const MlpWeights w = mlp_weights_of(nbin, W1b, W2b, W3b, theta);
    const size_t lds = sizeof(float) * 64 * kMlpIn + 4 * 4096 * sizeof(uint16_t);
    hipLaunchKernelGGL(mlp_forward_rows, dim3((unsigned)((n + 63) / 64)), dim3(256), lds, 0, X, n, w, Z);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return DNSMC_OK;
}

// ------------------------------------------------------ legacy boundary ----
static dnsmc_solver* g_legacy = nullptr;
static std::mutex g_legacy_mu;

static void legacy_die(const char* where)
{
    fprintf(stderr, "Fatal error: %s (%s)\n*** FAILED - ABORTING\n", where, g_err.c_str());
    exit(1);
}

static dnsmc_solver* legacy_solver()
{
    if (g_legacy) return g_legacy;
    dnsmc_config c;
    default_config(&c);
    const char* path = getenv("DNSMC_CONFIG");
    if (path && *path) {
        FILE* fp = fopen(path, "r");
        if (!fp) {
            g_err = std::string("cannot open DNSMC_CONFIG file ") + path;
            legacy_die("mc_gpu_");
        }
        std::string text;
        char buf[4096];
        size_t k;
        while ((k = fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, k);
        fclose(fp);
        if (dnsmc_parse_config(text.c_str(), &c)) legacy_die("mc_gpu_: config");
    }
    if (dnsmc_create(&c, &g_legacy)) legacy_die("mc_gpu_: create");
    if (dnsmc_load_tables(g_legacy)) legacy_die("mc_gpu_: tables");
    return g_legacy;
}

void mc_gpu_(float* Tfort, int* ystart)
{
    std::lock_guard<std::mutex> lk(g_legacy_mu);
    dnsmc_solver* s = legacy_solver();
    int ys = ystart ? *ystart : 1;
    // the DNS passes the full field; emitters are this node's j-slab
    if (dnsmc_set_field(s, 0, Tfort, 0)) legacy_die("mc_gpu_: field");
    if (dnsmc_set_emitters_slab(s, ys)) legacy_die("mc_gpu_: ystart");
    if (dnsmc_run(s)) legacy_die("mc_gpu_: launch");
}

void get_results_(float* resfort, float* varfort)
{
    std::lock_guard<std::mutex> lk(g_legacy_mu);
    if (!g_legacy) {
        g_err = "get_results_ called before mc_gpu_";
        legacy_die("get_results_");
    }
    if (dnsmc_get_results_fortran(g_legacy, 0, resfort, varfort)) legacy_die("get_results_");
}

// The DNS coupling without the host all-reduces, in the legacy conventions
// (Fortran-callable, arguments by reference, exit(1) with a message on error),
// on the solver mc_gpu_ uses: fix_temp_gpu_ per rank's slab, then
// mc_gpu_commit_ in place of mc_gpu_; fix_rad_gpu_ per rank after the solve.
void fix_temp_gpu_(double* cnew, int* startc, int* endc, double* Tc, double* Tplus)
{
    std::lock_guard<std::mutex> lk(g_legacy_mu);
    dnsmc_solver* s = legacy_solver();
    if (!cnew || !startc || !endc || !Tc || !Tplus) {
        g_err = "null argument";
        legacy_die("fix_temp_gpu_");
    }
    if (dnsmc_fix_temp_slab(s, 0, cnew, startc, endc, *Tc, *Tplus)) legacy_die("fix_temp_gpu_");
}

void mc_gpu_commit_(int* ystart)
{
    std::lock_guard<std::mutex> lk(g_legacy_mu);
    dnsmc_solver* s = legacy_solver();
    if (dnsmc_commit_field(s, 0)) legacy_die("mc_gpu_commit_: field");
    if (dnsmc_set_emitters_slab(s, ystart ? *ystart : 1)) legacy_die("mc_gpu_commit_: ystart");
    if (dnsmc_run(s)) legacy_die("mc_gpu_commit_: launch");
}

void fix_rad_gpu_(double* divQ, double* vcore, int* startc, int* endc, double* Tc)
{
    std::lock_guard<std::mutex> lk(g_legacy_mu);
    if (!g_legacy) {
        g_err = "fix_rad_gpu_ called before a solve";
        legacy_die("fix_rad_gpu_");
    }
    if (!Tc) {
        g_err = "null argument";
        legacy_die("fix_rad_gpu_");
    }
    if (dnsmc_fix_rad_slab(g_legacy, 0, startc, endc, *Tc, divQ, vcore)) legacy_die("fix_rad_gpu_");
}

void mc_gpu(float* Tfort, int* ystart) { mc_gpu_(Tfort, ystart); }
void get_results(float* resfort, float* varfort) { get_results_(resfort, varfort); }

}  // extern "C"
