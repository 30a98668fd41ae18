/*
 * Oracle Monte Carlo (test infrastructure only; see oracle.h).
 *
 * Restates the OERM estimator of the reference's code/mc_gpu.cu:
 *   - bundle start at the emitting cell centre (:111-116);
 *   - isotropic direction, polar axis = wall-normal axis (:813-828), zero
 *     components replaced by 1e-10 (:127-132), sign flags (:133-135);
 *   - (band, g) from the proposal CDFs by the reference search (:830-896);
 *   - bundle weight Ibmax * kappa_i Ib_i / (kappa_max Ib_max) (:63,:163-169);
 *   - cell-by-cell march (:228-438): distances to the next face on every axis,
 *     the smallest wins, the winning coordinate snaps to the face, periodic
 *     wrap or black-wall absorption, cut-off when the transmitted fraction
 *     drops below toll (deposited in the current cell), step cap per level;
 *   - adaptive mesh coarsening: after maxs[lev] steps the indices coarsen as
 *     ic = (ic + r - 1) / r (:193-207);
 *   - tallies kept as emitted E and reciprocally absorbed A in FP64, Q = E-A.
 *
 * Pins: brute-force FP64 ray/box intersection on tiny grids (cell sequence),
 * the isothermal-enclosure invariant Q == 0, the plane-parallel E2 closed form
 * and its cut-off quadrature (orc_slab_exchange_cut), and code/resfortm.txt.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline size_t lin3(const int n[3], const int c[3])
{
    return (size_t)(c[0] - 1) + (size_t)n[0] * ((size_t)(c[1] - 1) + (size_t)n[1] * (size_t)(c[2] - 1));
}

static inline void lookup_f32(const orc_problem *P, float Tj, int bg, int band, float *kap, float *ib)
{
    if (P->spectral == ORC_SPEC_GREY_T4) {
        float t2 = Tj * Tj;
        *kap = P->kappa_grey;
        *ib = t2 * t2;
        return;
    }
    float pos = fmaf(Tj, P->inv_dT, P->mT0);
    pos = fminf(fmaxf(pos, 0.0f), (float)(P->nT - 1));
    int t0 = (int)pos;
    if (t0 > P->nT - 2) t0 = P->nT - 2;
    float fr = pos - (float)t0;
    const float *kr = P->kq + (size_t)bg * P->nT;
    const float *ir = P->Ib + (size_t)band * P->nT;
    *kap = fmaf(fr, kr[t0 + 1] - kr[t0], kr[t0]);
    *ib = fmaf(fr, ir[t0 + 1] - ir[t0], ir[t0]);
}

static inline void lookup_f64(const orc_problem *P, double Tj, int bg, int band, double *kap, double *ib)
{
    if (P->spectral == ORC_SPEC_GREY_T4) {
        *kap = P->kappa_grey;
        *ib = Tj * Tj * Tj * Tj;
        return;
    }
    double pos = Tj * (double)P->inv_dT + (double)P->mT0;
    if (pos < 0.0) pos = 0.0;
    if (pos > P->nT - 1) pos = P->nT - 1;
    int t0 = (int)pos;
    if (t0 > P->nT - 2) t0 = P->nT - 2;
    double fr = pos - t0;
    const float *kr = P->kq + (size_t)bg * P->nT;
    const float *ir = P->Ib + (size_t)band * P->nT;
    *kap = kr[t0] + fr * ((double)kr[t0 + 1] - kr[t0]);
    *ib = ir[t0] + fr * ((double)ir[t0 + 1] - ir[t0]);
}

static inline float wall_ib(const orc_problem *P, int ax, int side, int band)
{
    return P->Ibw[(ax * 2 + side) * P->nB + band];
}

static void direction_f32(int a, float ut, float up, float s[3])
{
    float c = 1.0f - 2.0f * ut;
    float st = 2.0f * sqrtf(ut * (1.0f - ut));
    float sp, cp;
    orc_sincos2pi(up, &sp, &cp);
    s[a] = c;
    s[(a + 1) % 3] = st * cp;
    s[(a + 2) % 3] = st * sp;
    for (int d = 0; d < 3; ++d)
        if (s[d] == 0.0f) s[d] = 1e-10f;
}

void orc_trace_ray_f32(const orc_problem *P, int i, int j, int k, float ut, float up,
                       int band, int g, orc_ray_out *o)
{
    float s[3], inv[3], p[3];
    int f[3], c[3] = {i, j, k};
    direction_f32(P->polar_axis, ut, up, s);
    for (int a = 0; a < 3; ++a) {
        f[a] = s[a] < 0.0f;
        inv[a] = 1.0f / s[a];
    }
    p[0] = P->centre[0][i];
    p[1] = P->centre[1][j];
    p[2] = P->centre[2][k];
    int bg = band * P->nQ + g;
    float tau = 0.0f, tra = 1.0f, G = 0.0f, D = 1.0f;
    int steps = 0, end = ORC_END_CAP, lev = 0;
    uint32_t h = 2166136261u;
    for (lev = 0; lev < P->nlev; ++lev) {
        const int *n = P->n[lev];
        const float *T = P->T[lev];
        const float *F0 = P->face[lev][0], *F1 = P->face[lev][1], *F2 = P->face[lev][2];
        const float *Fa[3] = {F0, F1, F2};
        if (lev > 0)
            for (int a = 0; a < 3; ++a) {
                int r = P->n[lev - 1][a] / n[a];
                c[a] = (c[a] + r - 1) / r;
            }
        int cnt = 0;
        while (cnt < P->maxs[lev]) {
            float fx[3], d[3];
            for (int a = 0; a < 3; ++a) {
                fx[a] = Fa[a][c[a] - f[a]];
                d[a] = (fx[a] - p[a]) * inv[a];
            }
            float ds = d[0];
            if (ds > d[1]) ds = d[1];
            if (ds > d[2]) ds = d[2];
            for (int a = 0; a < 3; ++a) p[a] = (ds == d[a]) ? fx[a] : fmaf(ds, s[a], p[a]);
            size_t idx = lin3(n, c);
            float kap, ibj;
            lookup_f32(P, T[idx], bg, band, &kap, &ibj);
            h = (h ^ (((uint32_t)lev << 28) ^ (uint32_t)idx)) * 16777619u;
            ++steps;
            ++cnt;
            float tau1 = fmaf(kap, ds, tau);
            if (tau1 >= P->tau_cut) {
                G = fmaf(tra, ibj, G);
                end = ORC_END_CUTOFF;
                goto done;
            }
            float tra1 = expf(-tau1);
            G = fmaf(tra - tra1, ibj, G);
            tau = tau1;
            tra = tra1;
            int ax = (d[0] < d[1]) ? ((d[0] < d[2]) ? 0 : 2) : ((d[1] < d[2]) ? 1 : 2);
            c[ax] += f[ax] ? -1 : 1;
            if (c[ax] == 0 || c[ax] == n[ax] + 1) {
                int side = (c[ax] != 0);
                if (P->bc[ax][side] == ORC_BC_PERIODIC) {
                    c[ax] = side ? 1 : n[ax];
                    p[ax] = side ? Fa[ax][0] : Fa[ax][n[ax]];
                } else {
                    G = fmaf(tra, wall_ib(P, ax, side, band), G);
                    end = ORC_END_WALL;
                    goto done;
                }
            }
        }
    }
    lev = P->nlev - 1;
    D = 1.0f - tra;
done:
    o->steps = steps;
    o->end_type = end;
    o->end_cell[0] = c[0];
    o->end_cell[1] = c[1];
    o->end_cell[2] = c[2];
    o->end_level = lev;
    o->hash = h;
    o->D = D;
    o->G = G;
}

void orc_trace_ray_f64(const orc_problem *P, int i, int j, int k, double ut, double up,
                       int band, int g, double *Dout, double *Gout, int *steps_out, int *end_out)
{
    double s[3], inv[3], p[3];
    int f[3], c[3] = {i, j, k};
    {
        int a = P->polar_axis;
        double cth = 1.0 - 2.0 * ut, st = 2.0 * sqrt(ut * (1.0 - ut));
        double ph = 6.283185307179586476925 * up;
        s[a] = cth;
        s[(a + 1) % 3] = st * cos(ph);
        s[(a + 2) % 3] = st * sin(ph);
        for (int d = 0; d < 3; ++d)
            if (s[d] == 0.0) s[d] = 1e-10;
    }
    for (int a = 0; a < 3; ++a) {
        f[a] = s[a] < 0.0;
        inv[a] = 1.0 / s[a];
    }
    p[0] = P->centre[0][i];
    p[1] = P->centre[1][j];
    p[2] = P->centre[2][k];
    int bg = band * P->nQ + g;
    double tau = 0.0, tra = 1.0, G = 0.0, D = 1.0;
    double tau_cut = (double)P->tau_cut;
    int steps = 0, end = ORC_END_CAP;
    for (int lev = 0; lev < P->nlev; ++lev) {
        const int *n = P->n[lev];
        const float *T = P->T[lev];
        const float *Fa[3] = {P->face[lev][0], P->face[lev][1], P->face[lev][2]};
        if (lev > 0)
            for (int a = 0; a < 3; ++a) {
                int r = P->n[lev - 1][a] / n[a];
                c[a] = (c[a] + r - 1) / r;
            }
        int cnt = 0;
        while (cnt < P->maxs[lev]) {
            double fx[3], d[3];
            for (int a = 0; a < 3; ++a) {
                fx[a] = Fa[a][c[a] - f[a]];
                d[a] = (fx[a] - p[a]) * inv[a];
            }
            double ds = d[0];
            if (ds > d[1]) ds = d[1];
            if (ds > d[2]) ds = d[2];
            for (int a = 0; a < 3; ++a) p[a] = (ds == d[a]) ? fx[a] : p[a] + ds * s[a];
            size_t idx = lin3(n, c);
            double kap, ibj;
            lookup_f64(P, (double)T[idx], bg, band, &kap, &ibj);
            ++steps;
            ++cnt;
            double tau1 = tau + kap * ds;
            if (tau1 >= tau_cut) {
                G += tra * ibj;
                end = ORC_END_CUTOFF;
                goto done;
            }
            double tra1 = exp(-tau1);
            G += (tra - tra1) * ibj;
            tau = tau1;
            tra = tra1;
            int ax = (d[0] < d[1]) ? ((d[0] < d[2]) ? 0 : 2) : ((d[1] < d[2]) ? 1 : 2);
            c[ax] += f[ax] ? -1 : 1;
            if (c[ax] == 0 || c[ax] == n[ax] + 1) {
                int side = (c[ax] != 0);
                if (P->bc[ax][side] == ORC_BC_PERIODIC) {
                    c[ax] = side ? 1 : n[ax];
                    p[ax] = side ? Fa[ax][0] : Fa[ax][n[ax]];
                } else {
                    G += tra * (double)wall_ib(P, ax, side, band);
                    end = ORC_END_WALL;
                    goto done;
                }
            }
        }
    }
    D = 1.0 - tra;
done:
    *Dout = D;
    *Gout = G;
    *steps_out = steps;
    *end_out = end;
}

/* ------------------------------------------------------ descriptors ---- */
static inline void desc_pack(orc_ray_desc *d, int band, int g, uint32_t mt, uint32_t mp, float w, float winv)
{
    /* w0 = m_theta | band[0:8] << 24 ; w1 = m_phi | band[8:10] << 24 | g << 26 */
    d->w0 = (mt & 0xFFFFFFu) | ((uint32_t)(band & 0xFF) << 24);
    d->w1 = (mp & 0xFFFFFFu) | ((uint32_t)((band >> 8) & 0x3) << 24) | ((uint32_t)(g & 0x3F) << 26);
    d->w = w;
    d->winv = winv;
}

void orc_desc_decode(const orc_ray_desc *d, int nQ, int *band, int *g, float *ut, float *up)
{
    (void)nQ;
    *band = (int)((d->w0 >> 24) | (((d->w1 >> 24) & 0x3u) << 8));
    *g = (int)(d->w1 >> 26);
    *ut = (float)(d->w0 & 0xFFFFFFu);
    *up = (float)(d->w1 & 0xFFFFFFu);
}

static inline void cell_ijk(const orc_problem *P, int cell, int *i, int *j, int *k)
{
    int nx = P->n[0][0], ny = P->n[0][1];
    *i = cell % nx + 1;
    *j = (cell / nx) % ny + 1;
    *k = cell / (nx * ny) + 1;
}

/* Bundle weight: Ibmax * (kappa_i Ib_i) / (kappa_max Ib_max) and the factor
 * that turns sum seg*Ib_j into reciprocally absorbed energy. */
static inline void weights_f32(const orc_problem *P, float Ti, int bg, int band, float *w, float *winv)
{
    float kap, ib;
    lookup_f32(P, Ti, bg, band, &kap, &ib);
    float ww = P->Ibmax * (kap * ib) * P->inv_emax[bg];
    *w = ww;
    *winv = ww * (1.0f / ib);
}

static inline int sample_g(const orc_problem *P, int band, float R)
{
    int g = orc_cdf_search(P->cdf_g + (size_t)band * P->nQ, P->nQ, R);
    return g;
}

void orc_gen_philox(const orc_problem *P, uint64_t seed, uint32_t snap, int cell, int r0, int nr,
                    orc_ray_desc *out)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    float Ti = P->T[0][cell];
    for (int r = 0; r < nr; ++r) {
        uint32_t ctr[4] = {(uint32_t)(r0 + r), (uint32_t)cell, snap, 0u}, x[4];
        orc_philox4x32_10(ctr, key, x);
        int band = orc_cdf_search(P->cdf_band, P->nB, orc_u01_open(x[0]));
        int g = P->g_search ? sample_g(P, band, orc_u01_open(x[1])) : 0;
        if (g > P->nQ - 1) g = P->nQ - 1;
        int bg = band * P->nQ + g;
        float w, winv;
        weights_f32(P, Ti, bg, band, &w, &winv);
        desc_pack(&out[r], band, g, x[2] >> 8, x[3] >> 8, w, winv);
    }
}

void orc_gen_compat(const orc_problem *P, uint64_t seed, uint64_t subseq, int use_sobol,
                    uint64_t sobol_offset, int cell, int nr, orc_ray_desc *out)
{
    orc_xorwow st;
    orc_xorwow_init(&st, seed, subseq, 0);
    uint32_t v0[32], v1[32];
    orc_sobol_dirs(0, v0);
    orc_sobol_dirs(1, v1);
    float Ti = P->T[0][cell];
    for (int h = 0; h < nr; ++h) {
        uint32_t mp, mt;
        if (use_sobol) {
            mp = orc_sobol_point(v0, sobol_offset + (uint64_t)h) >> 8;
            mt = orc_sobol_point(v1, sobol_offset + (uint64_t)h) >> 8;
        } else {
            mp = orc_xorwow_next(&st) >> 8;
            mt = orc_xorwow_next(&st) >> 8;
        }
        int band = orc_cdf_search(P->cdf_band, P->nB, orc_u01_incl(orc_xorwow_next(&st)));
        int g = 0;
        if (P->g_search) g = sample_g(P, band, orc_u01_incl(orc_xorwow_next(&st)));
        if (g > P->nQ - 1) g = P->nQ - 1;
        int bg = band * P->nQ + g;
        float w, winv;
        weights_f32(P, Ti, bg, band, &w, &winv);
        desc_pack(&out[h], band, g, mt, mp, w, winv);
    }
}

/* Band order of the sorted mode (code/mc_gpu.cu:941-957 sort_idx): order[p] is
 * the band at position p when the bands are ranked by kavg ascending.  The
 * reference uses std::sort (ties in unspecified order); ties here keep the
 * lower band index first. */
void orc_sort_idx(const float *kavg, int nB, int *order)
{
    for (int b = 0; b < nB; ++b) order[b] = b;
    for (int i = 1; i < nB; ++i) {          /* insertion sort: stable */
        int x = order[i], j = i - 1;
        while (j >= 0 && kavg[order[j]] > kavg[x]) { order[j + 1] = order[j]; --j; }
        order[j + 1] = x;
    }
}

/* Sorted compat generation (srt == 1; code/mc_gpu.cu:99-102, :138-139,
 * kernel_find :786-812, wave_find :897-930).  kernel_find draws nr (band, g)
 * pairs from the slot's XORWOW stream with find_band (Rwave, then Rwave2 when
 * g is searched) and counts them per band (nb_cnt) and per (g, band) (g_cnt).
 * The photon loop then draws each bundle's direction (Sobol point, or two
 * more XORWOW uniforms phi then theta) and takes its (band, g) from wave_find:
 * positions p = 0.. of the kavg order, each band's count replayed with g
 * ascending.  The reference's tables are permuted into that order, so the
 * bundle at position p carries band order[p]. */
void orc_gen_compat_sorted(const orc_problem *P, uint64_t seed, uint64_t subseq, int use_sobol,
                           uint64_t sobol_offset, int cell, int nr, const int *order, orc_ray_desc *out)
{
    orc_xorwow st;
    orc_xorwow_init(&st, seed, subseq, 0);
    uint32_t v0[32], v1[32];
    orc_sobol_dirs(0, v0);
    orc_sobol_dirs(1, v1);
    const int nB = P->nB, nQ = P->nQ;
    int *nb_cnt = (int *)calloc((size_t)nB, sizeof(int));
    int *g_cnt = (int *)calloc((size_t)nB * nQ, sizeof(int));    /* [g][band] */
    for (int h = 0; h < nr; ++h) {                               /* kernel_find */
        int band = orc_cdf_search(P->cdf_band, nB, orc_u01_incl(orc_xorwow_next(&st)));
        int g = 0;
        if (P->g_search) g = sample_g(P, band, orc_u01_incl(orc_xorwow_next(&st)));
        if (g > nQ - 1) g = nQ - 1;
        nb_cnt[band] += 1;
        g_cnt[(size_t)g * nB + band] += 1;
    }
    float Ti = P->T[0][cell];
    int nb = 0, g = 0, countnb = 0, countg = 0;
    for (int h = 0; h < nr; ++h) {
        uint32_t mp, mt;
        if (use_sobol) {
            mp = orc_sobol_point(v0, sobol_offset + (uint64_t)h) >> 8;
            mt = orc_sobol_point(v1, sobol_offset + (uint64_t)h) >> 8;
        } else {
            mp = orc_xorwow_next(&st) >> 8;
            mt = orc_xorwow_next(&st) >> 8;
        }
        /* wave_find */
        int more = 1;
        while (more) {
            if (countnb < nb_cnt[order[nb]]) { countnb += 1; more = 0; }
            else { nb += 1; countnb = 0; countg = 0; g = 0; more = 1; }
        }
        while (!more) {
            if (countg < g_cnt[(size_t)g * nB + order[nb]]) { countg += 1; more = 1; }
            else { g += 1; countg = 0; more = 0; }
        }
        int band = order[nb];
        int bg = band * nQ + g;
        float w, winv;
        weights_f32(P, Ti, bg, band, &w, &winv);
        desc_pack(&out[h], band, g, mt, mp, w, winv);
    }
    free(nb_cnt);
    free(g_cnt);
}

void orc_trace_descs(const orc_problem *P, int cell, const orc_ray_desc *d, int nr,
                     double *E, double *A, double *Q2, long *nsteps)
{
    int i, j, k;
    cell_ijk(P, cell, &i, &j, &k);
    double e = 0.0, a = 0.0, q2 = 0.0;
    long ns = 0;
    for (int r = 0; r < nr; ++r) {
        int band, g;
        float mt, mp;
        orc_desc_decode(&d[r], P->nQ, &band, &g, &mt, &mp);
        float ut = (mt + P->u_off) * 5.9604644775390625e-08f;
        float up = (mp + P->u_off) * 5.9604644775390625e-08f;
        orc_ray_out o;
        orc_trace_ray_f32(P, i, j, k, ut, up, band, g, &o);
        double er = (double)(d[r].w * o.D), ar = (double)(d[r].winv * o.G);
        e += er;
        a += ar;
        q2 += (er - ar) * (er - ar);
        ns += o.steps;
    }
    *E = e;
    *A = a;
    *Q2 = q2;
    if (nsteps) *nsteps = ns;
}

void orc_mc_philox(const orc_problem *P, uint64_t seed, uint32_t snap, int R, const int *cells,
                   int ncells, double *E, double *A, double *Q2, long *nsteps, int nthreads)
{
    int ntot = P->n[0][0] * P->n[0][1] * P->n[0][2];
    int nc = cells ? ncells : ntot;
    long ns_total = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        orc_ray_desc *buf = (orc_ray_desc *)malloc(sizeof(orc_ray_desc) * 256);
        long ns_local = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int q = 0; q < nc; ++q) {
            int cell = cells ? cells[q] : q;
            double e = 0, a = 0, q2 = 0;
            for (int r0 = 0; r0 < R; r0 += 256) {
                int nr = R - r0 < 256 ? R - r0 : 256;
                double e1, a1, q21;
                long ns1;
                orc_gen_philox(P, seed, snap, cell, r0, nr, buf);
                orc_trace_descs(P, cell, buf, nr, &e1, &a1, &q21, &ns1);
                e += e1; a += a1; q2 += q21; ns_local += ns1;
            }
            E[q] = e; A[q] = a; Q2[q] = q2;
        }
#ifdef _OPENMP
#pragma omp atomic
#endif
        ns_total += ns_local;
        free(buf);
    }
    if (nsteps) *nsteps = ns_total;
}

void orc_mc_philox_f64(const orc_problem *P, uint64_t seed, uint32_t snap, int R, const int *cells,
                       int ncells, double *E, double *A, double *Q2, long *nsteps, int nthreads)
{
    int ntot = P->n[0][0] * P->n[0][1] * P->n[0][2];
    int nc = cells ? ncells : ntot;
    long ns_total = 0;
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    double Tm = P->Tmax;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        long ns_local = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int q = 0; q < nc; ++q) {
            int cell = cells ? cells[q] : q;
            int i, j, k;
            cell_ijk(P, cell, &i, &j, &k);
            double Ti = P->T[0][cell];
            double e = 0, a = 0, q2 = 0;
            for (int r = 0; r < R; ++r) {
                uint32_t ctr[4] = {(uint32_t)r, (uint32_t)cell, snap, 0u}, x[4];
                orc_philox4x32_10(ctr, key, x);
                int band = orc_cdf_search(P->cdf_band, P->nB, orc_u01_open(x[0]));
                int g = P->g_search ? sample_g(P, band, orc_u01_open(x[1])) : 0;
                if (g > P->nQ - 1) g = P->nQ - 1;
                int bg = band * P->nQ + g;
                double ki, ii, kmx, imx;
                lookup_f64(P, Ti, bg, band, &ki, &ii);
                lookup_f64(P, Tm, bg, band, &kmx, &imx);
                double w = P->Ibmax_d * (ki * ii) / (kmx * imx);
                double ut = ((double)(x[2] >> 8) + 0.5) * 5.9604644775390625e-08;
                double up = ((double)(x[3] >> 8) + 0.5) * 5.9604644775390625e-08;
                double D, G;
                int st, en;
                orc_trace_ray_f64(P, i, j, k, ut, up, band, g, &D, &G, &st, &en);
                double er = w * D, ar = w * G / ii;
                e += er; a += ar; q2 += (er - ar) * (er - ar);
                ns_local += st;
            }
            E[q] = e; A[q] = a; Q2[q] = q2;
        }
#ifdef _OPENMP
#pragma omp atomic
#endif
        ns_total += ns_local;
    }
    if (nsteps) *nsteps = ns_total;
}
