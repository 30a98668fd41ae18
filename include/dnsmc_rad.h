/*
 * dnsmc_rad.h -- C ABI of the MI355X Monte Carlo radiative-transfer solver.
 *
 * The channel DNS (Fortran) computes the radiative source term Q_rad [W/m^3]
 * from its temperature field through two calls.  This library exports exactly
 * those symbols, with the same argument meaning, array layout and
 * asynchronous contract, so it replaces the reference's mc_gpu.o, memory_copy.o,
 * read_tables.o and spline.o plus -lcudart -lcurand on the link line
 * (code/make_test:22-23, code/makefile:21-22).
 *
 * Everything in the signatures is a plain pointer or size; nothing depends on
 * torch or on any C++ type.  All extended entry points return 0 on success and
 * a negative DNSMC_E* code on failure; dnsmc_last_error() gives the message.
 */
#ifndef DNSMC_RAD_H
#define DNSMC_RAD_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Reference boundary (Fortran implicit interface: lower case + trailing _).
 * ------------------------------------------------------------------------ */

/* Replaces code/mc_gpu.cu:440  extern "C" void mc_gpu_(float *Tfort, int *ystart).
 * Tfort: real(4) T(0:imax+1, 0:jmax+1, 0:kmax+1), column-major (i fastest),
 *        ghost layers included, in K.  Copied before return: the caller may
 *        overwrite it immediately.
 * ystart: 1-based first j of this node's slab (xstart(2) in 2DECOMP).
 * Returns as soon as the work is queued on the GPU (the DNS overlaps several
 * time steps with it, code/main_comp.f:151-208).  Sizes, domain and physics
 * come from the runtime configuration (DNSMC_CONFIG file; defaults equal the
 * reference's shipped code/param.h).  Errors abort with a message, as the
 * reference's cudaCheckErrors does (code/param.h:85-95). */
void mc_gpu_(float *Tfort, int *ystart);

/* Replaces code/mc_gpu.cu:640  extern "C" void get_results_(float *resfort, float *varfort).
 * Waits for the work queued by mc_gpu_, then writes interior cells of
 * resfort = Q_rad [W/m^3] (positive = net emission) and varfort = MC standard
 * error of Q_rad, both real(4) (0:imax+1, 0:jmax/p_row+1, 0:kmax+1),
 * column-major; ghost entries are left untouched. */
void get_results_(float *resfort, float *varfort);

/* BIND(C, NAME='mc_gpu') / NAME='get_results' spellings of the same calls:
 * the reference declares (but does not use) this iso_c_binding interface in
 * code/rad_main.f:2-9. */
void mc_gpu(float *Tfort, int *ystart);
void get_results(float *resfort, float *varfort);

/* Replace the host all-reduces of the DNS coupling, code/fix_var.f:49-73
 * (subroutine fix_temp) and :75-121 (subroutine fix_rad), called from
 * code/main_comp.f:154,196,199, on the solver mc_gpu_ uses (same legacy
 * conventions: by reference, abort with a message on error).
 * fix_temp_gpu_(cnew, startc, endc, Tc, Tplus): one rank's slab, real(8)
 *   cnew(0:imax+1, jmax/p_row, kmax/p_col) with the rank's xstart / xend,
 *   into the solver's temperature field (see dnsmc_fix_temp_slab);
 * mc_gpu_commit_(ystart): after every rank's slab has been written, queue the
 *   solve as mc_gpu_(TMC, ystart) would;
 * fix_rad_gpu_(divQ, vcore, startc, endc, Tc): one rank's real(8)
 *   divQ / vcore(0:imax+1, jmax/p_row, kmax/p_col) from the last solve (see
 *   dnsmc_fix_rad_slab); waits for the solve like get_results_. */
void fix_temp_gpu_(double *cnew, int *startc, int *endc, double *Tc, double *Tplus);
void mc_gpu_commit_(int *ystart);
void fix_rad_gpu_(double *divQ, double *vcore, int *startc, int *endc, double *Tc);

/* ------------------------------------------------------------------------
 * Extended API.  No counterpart in the reference (its configuration is
 * compile-time macros in code/param.h and code/NarrowBand.h); the two calls
 * above are thin wrappers over it.
 * ------------------------------------------------------------------------ */

enum {
    DNSMC_OK = 0,
    DNSMC_EINVAL = -1,      /* bad argument / configuration */
    DNSMC_EIO = -2,         /* table file missing or malformed */
    DNSMC_ERANGE = -3,      /* temperature outside the spectral table, NaN in T */
    DNSMC_EDEVICE = -4,     /* HIP runtime error */
    DNSMC_ESTATE = -5       /* calls out of order */
};

enum { DNSMC_BC_PERIODIC = 1, DNSMC_BC_BLACK = 2 };
enum { DNSMC_STRETCH_UNIFORM = 0, DNSMC_STRETCH_REF_QUAD = 1, DNSMC_STRETCH_TANH = 2 };
/* spectral models: code/NarrowBand.h grey==0/1/2, plus a constant-kappa grey gas */
enum { DNSMC_SPEC_GAS_CK = 0, DNSMC_SPEC_GREY_REF = 1, DNSMC_SPEC_PARTICLE = 2, DNSMC_SPEC_GREY_T4 = 3 };
/* estimators */
enum {
    DNSMC_MODE_PLAIN = 0,    /* Philox rays, proposal built exactly at Tmax */
    DNSMC_MODE_COMPAT = 1,   /* reference sampling: XORWOW per (cell, block), Sobol directions,
                                proposal CDF at the table node nearest Tmax */
    DNSMC_MODE_LEARNED = 2   /* MLP proposal over direction bins and (band, g), p/q reweighted */
};
enum { DNSMC_TALLY_F64_ATOMIC = 0, DNSMC_TALLY_F64 = 1, DNSMC_TALLY_KAHAN_F32 = 2, DNSMC_TALLY_F32 = 3 };

typedef struct dnsmc_config {
    int n[3];                /* interior cells imax, jmax, kmax */
    float L[3];              /* Lx, Ly, Lz [m] */
    int stretch[3];          /* face generator per axis (DNSMC_STRETCH_*) */
    float stretch_beta[3];   /* tanh factor for DNSMC_STRETCH_TANH */
    int bc[3][2];            /* per axis, low/high side: DNSMC_BC_* */
    float Tw[3][2];          /* wall temperatures [K] */
    int polar_axis;          /* axis of cos(theta) for emitted directions */
    int nvar;                /* variance blocks (compat) */
    int photmax;             /* bundles per block and cell (compat) */
    int rays_per_cell;       /* plain / learned estimators */
    int grid_num;            /* adaptive mesh coarsening levels (1 = off) */
    int maxs[6];             /* step cap per level */
    float toll;              /* transmissivity cut-off */
    int spectral;            /* DNSMC_SPEC_* */
    float kappa_grey;        /* DNSMC_SPEC_GREY_T4 absorption coefficient [1/m] */
    int mode;                /* DNSMC_MODE_* */
    int sobol;               /* compat: Sobol directions (1) or XORWOW (0) */
    int num_streams;         /* compat seeding: reference stream count (code/param.h:40) */
    int num_gpu, gpu, p_row; /* compat seeding and slab decomposition */
    unsigned long long seed;
    int nsnap;               /* temperature snapshots batched in one run */
    int precision;           /* 0 FP32 traversal, 1 FP64 traversal */
    int tally;               /* DNSMC_TALLY_* */
    int srt;                 /* compat mode, gas tables: replay each block's drawn (band, g) multiset in
                                ascending-kavg band order (code/param.h:47 srt; forced off for grey 1/2) */
    long long batch_rays;    /* rays generated per traversal launch */
    char table_dir[512];     /* directory of the spectral tables */
} dnsmc_config;

typedef struct dnsmc_solver dnsmc_solver;

typedef struct dnsmc_stats {
    double ms_total;          /* last run, host wall clock */
    double ms_trace;          /* sum of traversal kernel time (HIP events) */
    double ms_raygen;         /* sum of ray generation kernel time */
    double ms_other;          /* spline, finalize, reductions */
    long long n_rays;
    long long n_steps;        /* cell crossings (segments) */
    int n_trace_launches;
    float Tmax;
    float kappamax;
    double Ibmax;
    int table_mode;           /* per-step table reads: 0 global, 1 Ib in LDS, 2 kappa+Ib slice in LDS */
    int table_lo, table_nw;   /* temperature nodes held in LDS (mode 2) */
    int n_pipes;              /* batches in flight at once (one stream each); their traversal launches overlap */
    long long n_listed;       /* rays traced first from the long-ray lists (fused plain solves) */
} dnsmc_stats;

/* Defaults equal to the reference's shipped code/param.h + NarrowBand.h. */
void dnsmc_default_config(dnsmc_config *cfg);
/* Parse "key = value" lines; keys are the field names of dnsmc_config, lists
 * space-separated ("n = 384 256 256", "bc = 1 1 2 2 1 1"), '#' starts a comment. */
int dnsmc_parse_config(const char *text, dnsmc_config *cfg);

int dnsmc_create(const dnsmc_config *cfg, dnsmc_solver **out);
void dnsmc_destroy(dnsmc_solver *s);
const char *dnsmc_last_error(void);

/* Spectral tables.  dnsmc_load_tables reads the reference text formats from
 * cfg->table_dir (code/read_tables.cpp:20-92 gas, :187-241 particle;
 * grey==1 and grey-T4 tables are synthesised).  dnsmc_set_tables takes the
 * same data as arrays: Tnb[nT], kP[nT], wvl/wvr/wvc[nB], kq[nB][nT][nQ],
 * prob[nT][nB], probg[nT][nB][nQ], and kavg[nB], the band's kavg column of
 * its last temperature row in NarrBand%d.txt (the sorted mode's band key,
 * code/read_tables.cpp:50; NULL: the mean over g of kq at the last row). */
int dnsmc_load_tables(dnsmc_solver *s);
/* Reads the tables cfg names (as dnsmc_load_tables does) on the host only, no
 * device and no solver needed, and reports their sizes (nB, nQ, nT; any
 * pointer may be NULL).  The reference's readT (code/read_tables.cpp:20-92)
 * reads with unchecked fopen/fscanf; here a missing file, a short or
 * malformed one, or a probability file whose (T, band) rows do not match the
 * band files returns DNSMC_EIO with the file named in dnsmc_last_error(). */
int dnsmc_probe_tables(const dnsmc_config *cfg, int *nB, int *nQ, int *nT);
int dnsmc_set_tables(dnsmc_solver *s, int nB, int nQ, int nT, const float *Tnb, const float *kP,
                     const float *wvl, const float *wvr, const float *wvc, const float *kq,
                     const float *prob, const float *probg, const float *kavg);

/* Temperature field of snapshot `snap`, Fortran layout with ghosts
 * (n0+2)(n1+2)(n2+2), host or device pointer (is_device).  A field with a
 * non-positive or non-finite interior value is rejected (DNSMC_ERANGE) and
 * leaves the snapshot unset. */
int dnsmc_set_field(dnsmc_solver *s, int snap, const float *T, int is_device);
/* DNS coupling without the host all-reduces (code/fix_var.f:49-121, called
 * from code/main_comp.f around mc_gpu_ / get_results_).
 * dnsmc_fix_temp_slab replaces fix_temp (code/fix_var.f:49-73): one rank's
 * slab of the nondimensional temperature goes straight into snapshot `snap`
 * of the device field, T = cnew * Tc / Tplus + Tc evaluated in FP64 left to
 * right and rounded to FP32 (the value the reference's MPI_SUM all-reduce of
 * disjoint slabs leaves in the real(4) TMC).  cnew is the rank's x-pencil
 * array cnew(0:n0+1, nj, nk), i fastest, host memory, with nj, nk the extents
 * of [startc, endc] (1-based inclusive global indices, the 2DECOMP
 * xstart/xend; interior only).  The first slab after a commit marks every
 * cell uncovered; dnsmc_commit_field, after the last slab, builds the coarse
 * levels and the temperature range and fails with DNSMC_ERANGE if any cell
 * was left uncovered or is non-positive.  A node's GPU owner calls it for its
 * own slab and for the slabs the other ranks send it (a gather of slabs
 * instead of the all-reduce of the full field).
 * dnsmc_fix_rad_slab replaces fix_rad (code/fix_var.f:75-121) for one rank:
 * divQ(i, jl, k - startc[2] + 1) = Q / stefan / Tc^4 and vcore likewise from
 * the standard error, stefan = 5.670373e-08 (code/param.txt:19), Tc^4 as
 * Tc*Tc*Tc*Tc, all FP64 and from the FP64 results (the reference divides the
 * real(4) copy); i = 1..n0, jl = 1..n1/p_row over the emitter slab
 * (dnsmc_set_emitters_slab), k = startc[2]..endc[2]; host arrays
 * (0:n0+1, n1/p_row, endc[2]-startc[2]+1), zero outside those cells. */
int dnsmc_fix_temp_slab(dnsmc_solver *s, int snap, const double *cnew, const int *startc, const int *endc,
                        double Tc, double Tplus);
int dnsmc_commit_field(dnsmc_solver *s, int snap);
int dnsmc_fix_rad_slab(dnsmc_solver *s, int snap, const int *startc, const int *endc, double Tc, double *divQ,
                       double *vcore);
/* Learned proposal (no counterpart in the reference): network 6 -> 64 -> 64 ->
 * (64 direction bins | nbin = nB*nQ (band, g) bins, padded to noutp, a multiple
 * of 64).  Weights as BF16 bits, row-major [in][out] (W1 6x64, W2 64x64,
 * W3 64 x noutp), biases FP32 (64, 64, noutp); host pointers; the tables must
 * be loaded.  mix_eps in (0, 1]: the rays are drawn from
 * (1 - mix_eps) q_net + mix_eps p_plain and reweighted by p/q. */
int dnsmc_set_mlp(dnsmc_solver *s, const unsigned short *W1, const float *b1, const unsigned short *W2,
                  const float *b2, const unsigned short *W3, const float *b3, float mix_eps);

/* Emitting cells: the j-slab [ystart, ystart + n1/p_row) (1-based) as in the
 * reference, or an explicit list of 0-based linear cells (i fastest). */
int dnsmc_set_emitters_slab(dnsmc_solver *s, int ystart);
int dnsmc_set_emitters_list(dnsmc_solver *s, const int *cells, int ncells);

/* Queue the whole radiation solve (all snapshots) and return.
 * The plain estimator with FP64 atomic tallies and records off draws its
 * rays inside the traversal, one launch for all snapshots (one per range of
 * consecutive slots beyond 2^32 - 2^20 rays per snapshot, alternating over
 * two streams), the rays along the walls that it would start last traced
 * first (dnsmc_stats.n_listed); the learned
 * estimator with FP64 atomic tallies and records off generates the
 * descriptors of a whole snapshot on a second stream while the previous
 * snapshot is traced (one traversal launch per snapshot, two snapshot-sized
 * descriptor sets when free device memory allows); the other modes generate
 * descriptors in batches of batch_rays on three pipeline streams.
 * Environment overrides, read at each run (measurement and diagnosis; the
 * defaults are the measured best):
 *   DNSMC_FUSED=0     batched generation for the plain estimator too
 *   DNSMC_FUSED_SLOTS=n  at most n slots per fused launch
 *   DNSMC_LONG_FIRST=0   fused launches trace in claim order only (no
 *                     list of the rays along the walls traced first)
 *   DNSMC_LONG_PART=k    list those rays from the final 1/k of each XCD's
 *                     share of the last snapshot (default 2 with several
 *                     snapshots per launch, 4 with one)
 *   DNSMC_LONG_CAP=n     at most n rays per list (a list that finds more
 *                     turns the listing off for the launch: plain order)
 *   DNSMC_SNAPWISE=0  batched generation for the learned estimator too
 *   DNSMC_PIPES=1..3  pipeline streams of the batched modes (1: the
 *                     launches of a solve also stay on one stream)
 *   DNSMC_TABLES=global  keep kappa / Planck tables out of LDS
 *   DNSMC_REFILL_K=k  idle lanes of a wave before it starts new rays
 *   DNSMC_SUPER=n     rays per claimed super-chunk (multiple of 64)
 *   DNSMC_BLOCK=n     threads per block of the traversal with LDS table
 *                     slices (default 1024: one block and 16 waves per CU)
 *   DNSMC_BATCH_CAP=n rays per batched launch (at most 2^32 - 2^20)
 *   DNSMC_CDF_GB=g    GB of learned CDF rows per pipeline (default 8, or 32
 *                     when the device memory holds it) */
int dnsmc_run(dnsmc_solver *s);
/* Allocation-free variant of dnsmc_run.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It queues
 * the plain estimator's solve of every snapshot as dnsmc_run does (rays
 * drawn inside the traversal; mode 0, precision 0, tally 0, records off,
 * else DNSMC_EINVAL), in plain claim order (the long-ray list needs a buffer:
 * the same rays, the FP64 sums in another order), and calls no hipMalloc,
 * hipFree or other allocating routine: the tallies, work-claim counters
 * (two sets, for launches alternating over two streams) and the step
 * counter live in a static
 * 64 MiB device array of the library (one per GPU, held until the process
 * ends), which limits a call to 2,796,032 emitters x snapshots (DNSMC_EINVAL
 * beyond).  Results go to the caller's device arrays Q[nsnap][nemit] and
 * sdev[nsnap][nemit] (FP64, emitter order), written when the solve completes
 * (dnsmc_wait); dnsmc_get_results then has nothing to return.  Every other
 * device buffer the solve reads must already exist: the field (dnsmc_set_field),
 * and the slot, table and table-coordinate buffers that an earlier dnsmc_run
 * with this configuration and at least as many emitters sized; otherwise
 * DNSMC_ESTATE names the buffer.  All calls on one GPU share its scratch
 * array, so two allocation-free solves that overlap on one GPU (other solvers,
 * other streams, other threads) race and must be serialised by the caller. */
int dnsmc_run_noalloc(dnsmc_solver *s, double *Q, double *sdev);
/* Wait for the queued solve. */
int dnsmc_wait(dnsmc_solver *s);
/* Results of snapshot `snap` for the emitters, in emitter order (FP64):
 * Q = E - A, E, A and the MC standard error; any pointer may be NULL. */
int dnsmc_get_results(dnsmc_solver *s, int snap, double *Q, double *E, double *A, double *sdev);
/* Fortran-layout FP32 results (interior written, ghosts untouched). */
int dnsmc_get_results_fortran(dnsmc_solver *s, int snap, float *resfort, float *varfort);
int dnsmc_get_stats(dnsmc_solver *s, dnsmc_stats *st);

/* Test and training hooks. */
/* Derived arrays: name in {"faces0","faces1",..,"centres","T1".."T5","cdf_band","cdf_g",
 * "inv_emax","Ib","kq","Ibw","scalars","p_bg","pcdf","qcdf","lrows","features"}; copies up to n
 * floats, returns count.  "qcdf": the learned proposal's cumulative weights of the
 * last batch, a row per slot in emitter order (the batch must cover every slot);
 * "features": the proposal's 6 input features per slot of the last run, emitter order. */
int dnsmc_get_array(dnsmc_solver *s, const char *name, int level, float *out, long long n);
/* Ray descriptors of the last run's first batch (SoA words) for parity tests.
 * The plain estimator draws its rays inside the traversal unless records are
 * on (dnsmc_set_record) or a deterministic tally is chosen, so descriptors
 * exist only for those runs and for the compat, sorted and learned modes; a
 * learned descriptor's w holds p/q (the emission weight is applied at ray
 * start) and its winv is not written. */
long long dnsmc_get_rays(dnsmc_solver *s, unsigned int *w0, unsigned int *w1, float *w, float *winv,
                         long long n);
/* Per-ray record of the last run's first batch: steps, end type, hash, D, G. */
long long dnsmc_get_ray_records(dnsmc_solver *s, int *steps, int *end_type, unsigned int *hash,
                                float *D, float *G, long long n);
/* Switch on per-ray records (costs bandwidth; off by default). */
int dnsmc_set_record(dnsmc_solver *s, int on);
/* Test hook of the learned estimator's in-traversal draw (no reference
 * counterpart): with the log on, the run writes every proposal row the waves
 * evaluate (dnsmc_get_array "lrows": cumulative weights U, emitter order) and
 * every drawn ray (dnsmc_get_rays: descriptor words, p/q in w), so that the
 * draws can be replayed from the same rows. */
int dnsmc_set_draw_log(dnsmc_solver *s, int on);
/* Training hooks of the learned proposal (no counterpart in the reference; used
 * by dnsmc_amd/learned.py).  Pointers are device pointers.  Parameters are one
 * FP32 vector theta of dnsmc_mlp_nparams(nbin) floats: W1, b1, W2, b2, W3, b3
 * in that order with the shapes above; W1b/W2b/W3b are its BF16 copies. */
long long dnsmc_mlp_nparams(int nbin);
/* Samples of the last run (one snapshot, one batch of <= 2^22 rays, records
 * on): per ray its cell's 6 features, direction bin, (band, g) bin and weight
 * |x| / sqrt(steps + cost0) with x the ray's net contribution (|x| when
 * cost0 < 0): the fit then targets the proposal of least variance x cost. */
int dnsmc_mlp_samples(dnsmc_solver *s, float *X, int *dir_bin, int *bin, float *weight, long long n,
                      float cost0);
/* Features of the first n emitting slots (processing order) of the last run. */
int dnsmc_mlp_features(dnsmc_solver *s, float *X, long long n);
/* Loss (sum w CE / sum w over both heads) and its gradient over n samples;
 * work holds ceil(n/64)*(P+1) + P + 2 floats; loss is a host pointer (may be NULL). */
int dnsmc_mlp_grad(int nbin, const float *theta, const unsigned short *W1b, const unsigned short *W2b,
                   const unsigned short *W3b, const float *X, const int *dir_bin, const int *bin,
                   const float *weight, long long n, float *work, long long nwork, float *grad, float *loss);
/* One Adam step (0.9, 0.999, 1e-8) on theta, refreshing the BF16 copies. */
int dnsmc_mlp_adam(int nbin, float *theta, float *m, float *v, const float *grad, float lr, int step,
                   unsigned short *W1b, unsigned short *W2b, unsigned short *W3b);
/* Logits Z[n][noutp] for inputs X[n][6]. */
int dnsmc_mlp_forward(int nbin, const float *theta, const unsigned short *W1b, const unsigned short *W2b,
                      const unsigned short *W3b, const float *X, long long n, float *Z);

/* Device buffers for bindings that keep data on the GPU (training samples):
 * allocate, free, synchronous copy (kind 1 host->device, 2 device->host,
 * 3 device->device). */
int dnsmc_dev_alloc(long long bytes, void **out);
int dnsmc_dev_free(void *p);
int dnsmc_dev_copy(void *dst, const void *src, long long bytes, int kind);

/* Device synchronise (bench timing). */
int dnsmc_sync(void);
/* Measurement hook: `reps` launches of a plain device-to-device copy of
 * `nbytes` (kernel copy_f32x4); returns read+write GB/s.  Used as the
 * achievable-bandwidth reference and to calibrate HBM byte counters. */
int dnsmc_bench_copy(long long nbytes, int reps, double *gbps);
/* Measurement hook: the achievable HBM bandwidth, the best of twelve streaming
 * variants (16 B per lane x 1 or 4 per iteration, plain or non-temporal, 4 to
 * 32 blocks of 256 threads per CU) over `nbytes` (use several GiB), `reps`
 * timed launches each.  kind 0: copy (read + write GB/s), kind 1: read only.
 * best (may be NULL) receives the winning variant (index: per-CU blocks
 * 4 << (v / 3), shape v % 3 = 1x16 B, 4x16 B, 4x16 B non-temporal). */
int dnsmc_bench_stream(long long nbytes, int reps, int kind, double *gbps, int *best);

#ifdef __cplusplus
}
#endif
#endif
