/*
 * CPU oracle for the Monte Carlo radiative-transfer hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product (dnsmc_amd, libdnsmc_rad.so) never
 * links, loads or calls anything here.
 *
 * It is an independent plain-C restatement of the algorithm in the reference's
 * code/mc_gpu.cu (kernel_fluid :44-226, march_ray :228-438, find_band :830-896,
 * emiss_ang/emiss_angSobol :813-828), code/read_tables.cpp (:20-92, :104-183,
 * :187-241), code/spline.cpp (:35-170), code/definitions.h (:60-160) and
 * code/memory_copy.cu (I_blackC :291-300).  Each function cites what it follows.
 *
 * Floating-point contract shared with the GPU kernels (so that every integer
 * decision -- cell sequence, step count, end type, band/g index -- is bit-exact):
 *   - FP32 ray state; fmaf() written explicitly wherever the kernel uses an FMA;
 *     the library is built with -ffp-contract=off so nothing else contracts;
 *   - IEEE-correct division and sqrtf (hipcc's default for HIP);
 *   - sin/cos of 2*pi*u from orc_sincos2pi() (own polynomial, same ops on GPU);
 *   - the ray's cut-off is decided in optical-depth space (tau >= -ln(toll)),
 *     so the exp() rounding (glibc vs v_exp_f32) only moves tally values.
 *
 * Parity pins (what checks each function) are listed in the header comment of
 * every source file; functions without one are marked "parity unpinned".
 */
#ifndef DNSMC_ORACLE_H
#define DNSMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_LEVELS 6

/* ---------------------------------------------------------------- RNG ---- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* XORWOW (Marsaglia 2003, "Xorshift RNGs", J. Stat. Softw. 8(14), p.5). */
typedef struct { uint32_t x[5]; uint32_t d; } orc_xorwow;
uint32_t orc_xorwow_next(orc_xorwow *s);
/* Seed -> state: own SplitMix64-based scramble (NOT cuRAND's; parity unpinned). */
void orc_xorwow_seed(orc_xorwow *s, uint64_t seed);
/* Jump the linear xorshift part by n steps (brute force) -- test helper. */
void orc_xorwow_skip_brute(orc_xorwow *s, uint64_t n);
/* Jump by subsequence * 2^67 + offset using GF(2) matrix powers built here. */
void orc_xorwow_init(orc_xorwow *s, uint64_t seed, uint64_t subsequence, uint64_t offset);
/* 32-bit draw -> float uniform: (m+1) 2^-24 on the top 24 bits, in (0,1] (compat
 * modes); (m+1/2) 2^-23 on the top 23 bits, strictly in (0,1) (Philox mode). */
float orc_u01_incl(uint32_t x);
float orc_u01_open(uint32_t x);

/* Sobol32, Antonov-Saleev Gray-code order, Joe-Kuo direction numbers. */
void orc_sobol_dirs(int dim, uint32_t v[32]);
uint32_t orc_sobol_point(const uint32_t v[32], uint64_t n);

/* ---------------------------------------------------- own elementary fns -- */
void orc_sincos2pi(float u, float *s, float *c);

/* Reference search (mc_gpu.cu:842-853 / :881-893): bin index for R on cdf[n]. */
int orc_cdf_search(const float *cdf, int n, float R);

/* --------------------------------------------------------------- grids ---- */
/* Reference quadratic wall stretch (definitions.h:108-134; rad_main.f:191-219). */
void orc_grid_ref_quad(int im, int im0, float L, float *faces /* im+1 */, float *centres /* im+2 */);
/* Uniform faces i*(L/n). */
void orc_grid_uniform(int n, float L, float *faces, float *centres);
/* Channel tanh stretch: y_j = L/2 (1 + tanh(b(2j/n-1))/tanh(b)). */
void orc_grid_tanh(int n, float L, double beta, float *faces, float *centres);

/* ------------------------------------------------------------- spectra ---- */
float orc_planck_band(float T, float nu);            /* memory_copy.cu:291-300 */

/* ------------------------------------------------------------- spline ---- */
/* Natural cubic spline of the fine interior field onto coarse centres, x then
 * y then z (spline.cpp:35-116), FP64 throughout. */
void orc_interp3d(const int nf[3], const float *xf, const float *yf, const float *zf,
                  const float *Tf /* interior nf0*nf1*nf2, i fastest */,
                  const int nc[3], const float *xc, const float *yc, const float *zc,
                  float *Tc /* interior nc0*nc1*nc2 */);

/* ------------------------------------------------------- closed forms ---- */
double orc_expint_En(int n, double x);
/* Plane-parallel grey slab, cell-centre OERM expectation for a 1-D profile.
 * Returns Q at every cell centre (W/m^3). */
void orc_slab_grey_Q(int ny, const double *faces /* ny+1 */, const double *T /* ny */,
                     double kappa, double Tw_lo, double Tw_hi, double *Q);
/* One spectral bin, per-layer kappa and Ib: expected exchange fraction F_i. */
void orc_slab_exchange(int ny, const double *faces, const double *kappa, const double *Ib,
                       double Ib_lo, double Ib_hi, double *F);
/* Same with the estimator's optical-depth cut-off, by mu quadrature. */
void orc_slab_exchange_cut(int ny, const double *faces, const double *kappa, const double *Ib,
                           double Ib_lo, double Ib_hi, double tau_cut, int npanel, double *F);

/* ---------------------------------------------------------- MC problem ---- */
enum { ORC_SPEC_TABLE = 0, ORC_SPEC_GREY_T4 = 1 };
enum { ORC_BC_PERIODIC = 1, ORC_BC_BLACK = 2 };
enum { ORC_END_CUTOFF = 0, ORC_END_WALL = 1, ORC_END_CAP = 2 };

typedef struct {
    /* geometry */
    int nlev;
    int n[ORC_MAX_LEVELS][3];
    const float *face[ORC_MAX_LEVELS][3];   /* n+1 each */
    const float *centre[3];                 /* level-0 centres, n0+2 each */
    const float *T[ORC_MAX_LEVELS];         /* interior fields, i fastest */
    int maxs[ORC_MAX_LEVELS];
    int bc[3][2];
    int polar_axis;                         /* axis of cos(theta) */
    float tau_cut;                          /* -ln(toll) */
    /* spectral */
    int spectral;
    int nB, nQ, nT;
    float inv_dT, mT0;                      /* pos = fmaf(T, inv_dT, mT0) */
    const float *kq;                        /* [(b*nQ+g)*nT + t] */
    const float *Ib;                        /* [b*nT + t] */
    const float *Ibw;                       /* [(axis*2+side)*nB + b] */
    float kappa_grey;
    /* proposal + weights */
    const float *cdf_band;                  /* nB */
    const float *cdf_g;                     /* nB*nQ */
    const float *inv_emax;                  /* nB*nQ : 1/(kappa(Tmax)*Ib(Tmax)) */
    float Ibmax;                            /* emission weight scale per bundle */
    double Ibmax_d;                         /* same in FP64 (all-FP64 runs) */
    float Tmax;
    int g_search;                           /* 1: reference g search (mc_gpu.cu:878-893), 0: g=0 */
    float u_off;                            /* direction uniform u = (m + u_off) 2^-24 */
} orc_problem;

typedef struct {
    int steps;            /* total steps over all levels */
    int end_type;
    int end_cell[3];
    int end_level;
    uint32_t hash;        /* hash of the visited (level, cell) sequence */
    float D, G;           /* deposited fraction; sum seg*Ib_j (+ wall) */
} orc_ray_out;

/* One ray from level-0 cell (i,j,k) (1-based) with direction uniforms and bin. */
void orc_trace_ray_f32(const orc_problem *P, int i, int j, int k, float u_theta, float u_phi,
                       int band, int g, orc_ray_out *out);
void orc_trace_ray_f64(const orc_problem *P, int i, int j, int k, double u_theta, double u_phi,
                       int band, int g, double *D, double *G, int *steps, int *end_type);

/* Ray descriptors (16 B each, same packing as the GPU generator writes). */
typedef struct { uint32_t w0, w1; float w, winv; } orc_ray_desc;

/* Generate the descriptors of rays [r0, r0+nr) of cell `cell` (0-based linear,
 * i fastest) with Philox: key = seed, ctr = (r, cell, snap, 0). */
void orc_gen_philox(const orc_problem *P, uint64_t seed, uint32_t snap, int cell,
                    int r0, int nr, orc_ray_desc *out);
/* Compat generation for one (cell, block v) stream of photmax bundles:
 * XORWOW(seed, subseq) for band/g (and directions when sobol==0), Sobol dims
 * 0/1 at points v*photmax+h for directions when sobol==1 (mc_gpu.cu:93-144). */
void orc_gen_compat(const orc_problem *P, uint64_t seed, uint64_t subseq, int use_sobol,
                    uint64_t sobol_offset, int cell, int nr, orc_ray_desc *out);
/* Sorted mode (srt == 1): band order by kavg (sort_idx, mc_gpu.cu:941-957) and
 * the count-and-replay generation of kernel_find / wave_find (:786-812, :897-930). */
void orc_sort_idx(const float *kavg, int nB, int *order);
void orc_gen_compat_sorted(const orc_problem *P, uint64_t seed, uint64_t subseq, int use_sobol,
                           uint64_t sobol_offset, int cell, int nr, const int *order, orc_ray_desc *out);

/* Decode helpers. */
void orc_desc_decode(const orc_ray_desc *d, int nQ, int *band, int *g, float *u_theta, float *u_phi);

/* Trace a list of descriptors for one slot (cell) and return FP64 tallies. */
void orc_trace_descs(const orc_problem *P, int cell, const orc_ray_desc *d, int nr,
                     double *E, double *A, double *Q2, long *nsteps);

/* Whole-field plain MC (OpenMP over cells), Philox mode: R rays per cell.
 * Writes E, A, Q2 per cell for the cells listed (cells==NULL: all). */
void orc_mc_philox(const orc_problem *P, uint64_t seed, uint32_t snap, int R,
                   const int *cells, int ncells, double *E, double *A, double *Q2,
                   long *nsteps, int nthreads);
/* All-FP64 variant (precision study / converged oracle). */
void orc_mc_philox_f64(const orc_problem *P, uint64_t seed, uint32_t snap, int R,
                       const int *cells, int ncells, double *E, double *A, double *Q2,
                       long *nsteps, int nthreads);

/* ----------------------------------------------------------- proposal ---- */
float orc_lerp_f32(const float *row, int nT, float inv_dT, float mT0, float T);
/* Test hook: build the plain proposal with a planted mistake (proposal.c). */
void orc_plant_mistake(int k);
int orc_build_proposal(int mode, int nB, int nQ, int nT, const float *Tnb, const float *kP,
                       const float *wvl, const float *wvr, const float *wvc, const float *kq,
                       const float *prob, const float *probg, const double *wq, float Tmax, int R,
                       float *Ib, float *kqrow, float *cdf_band, float *cdf_g, float *inv_emax,
                       double *Ibmax_d, float *kappamax);

/* --------------------------------------------------------------- MLP ---- */
/* FP64 MLP 6->64->64->(64+1600) forward and backward (test checker). */
void orc_mlp_forward(int nb, const double *X, const double *W1, const double *b1,
                     const double *W2, const double *b2, const double *W3, const double *b3,
                     int nin, int nh, int nout, double *H1, double *H2, double *Z);
/* The learned proposal's draw (the generator of the learned estimator; no
 * reference counterpart): for each of nslots slots (emitting cell cells[sl],
 * row U[sl*noutp ...] of unnormalised inclusive cumulative weights, ndir
 * direction bins then nbin (band, g) bins) and each ray r < R: Philox
 * (r, cell, snap, 1) -> direction bin and (band, g) bin from the defensive
 * mixtures (1 - eps) q + eps p, descriptor words and the ratio p/q. */
void orc_learned_draw(int nslots, int R, int ndir, int nbin, int nQ, int noutp, const float *U, const float *pcdf,
                      const float *p_bg, const int *cells, uint32_t snap, uint64_t seed, float eps, uint32_t *w0,
                      uint32_t *w1, float *ratio);
void orc_mlp_loss_grad(int nb, const double *X, const double *W1, const double *b1,
                       const double *W2, const double *b2, const double *W3, const double *b3,
                       int nin, int nh, int nout, int ndir,
                       const int *tgt_dir, const int *tgt_bin, const double *wt,
                       double *loss, double *gW1, double *gb1, double *gW2, double *gb2,
                       double *gW3, double *gb3);
/* The same two with the kernels' BF16 rounding of the product operands
 * (H1, H2 forward; dZ, dA2 backward; see mlp.c), and the rounding itself. */
void orc_mlp_forward_bf16(int nb, const double *X, const double *W1, const double *b1,
                          const double *W2, const double *b2, const double *W3, const double *b3,
                          int nin, int nh, int nout, double *H1, double *H2, double *Z);
void orc_mlp_loss_grad_bf16(int nb, const double *X, const double *W1, const double *b1,
                            const double *W2, const double *b2, const double *W3, const double *b3,
                            int nin, int nh, int nout, int ndir,
                            const int *tgt_dir, const int *tgt_bin, const double *wt,
                            double *loss, double *gW1, double *gb1, double *gW2, double *gb2,
                            double *gW3, double *gb3);
double orc_bf16_round(double x);

#ifdef __cplusplus
}
#endif
#endif
