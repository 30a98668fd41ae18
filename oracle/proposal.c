/*
 * Oracle restatement of the emission proposal and bundle weights (test
 * infrastructure only; see oracle.h).
 *
 * Reference: code/mc_gpu.cu:63 (Ibmax), :163-169 (ratioI), :541-551 (Planck
 * table), :584-590 and :837-858 (proposal rows at the table node nearest
 * Tmax), code/read_tables.cpp:34-36 (kappamax), code/memory_copy.cu:27-51
 * (wall emission).
 *
 * Two proposals:
 *   compat: CDF rows of the node nearest Tmax, Ibmax = 4 kappamax sigma Tmax^4 / photmax;
 *   plain : CDF of e(b,g) = kappa(Tmax,b,g) Ib(Tmax,b) w_g dnu_b built at
 *           Tmax itself with the kernels' interpolation, Ibmax = 4 pi sum e / R,
 *           so that the estimator is unbiased for the tabulated physics.
 * Pins (tests/test_oracle_proposal.py): the emission identity
 * R sum_bg p(bg) w(bg; T_i) = 4 pi sum_b dnu_b sum_g w_g kappa(T_i,b,g) Ib(T_i,b)
 * against a direct numpy sum over the parsed table files (unequal band widths,
 * asymmetric g weights), its grey limit 4 kappa sigma T^4, and the optically
 * thin closed form through the full estimator (orc_mc_philox); CDF
 * monotonicity and last entry 1; the compat rows equal the table rows.
 *
 * orc_plant_mistake(k) makes the plain proposal wrong on purpose (k = 1: the
 * width of band b+1 used for band b; 2: the g weights read in reverse; 3: the
 * g CDF shifted by one entry), so that the tests can show each check fails on
 * such a mistake; 0 restores the correct build.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>

static int g_plant = 0;

void orc_plant_mistake(int k) { g_plant = k; }

float orc_lerp_f32(const float *row, int nT, float inv_dT, float mT0, float T)
{
    float pos = fmaf(T, inv_dT, mT0);
    pos = fminf(fmaxf(pos, 0.0f), (float)(nT - 1));
    int t0 = (int)pos;
    if (t0 > nT - 2) t0 = nT - 2;
    float fr = pos - (float)t0;
    return fmaf(fr, row[t0 + 1] - row[t0], row[t0]);
}

/* kq: [b][t][g] (reference NarrowBand layout); outputs:
 * Ib[b][t], kqrow[(b*nQ+g)*nT + t], cdf_band[nB], cdf_g[nB*nQ], inv_emax[nB*nQ];
 * returns Ibmax in *Ibmax_d. mode 0 plain, 1 compat. */
int orc_build_proposal(int mode, int nB, int nQ, int nT, const float *Tnb, const float *kP,
                       const float *wvl, const float *wvr, const float *wvc, const float *kq,
                       const float *prob, const float *probg, const double *wq, float Tmax, int R,
                       float *Ib, float *kqrow, float *cdf_band, float *cdf_g, float *inv_emax,
                       double *Ibmax_d, float *kappamax_out)
{
    const double sigma = 5.670373e-08;
    float inv_dT = 1.0f / (Tnb[1] - Tnb[0]);
    float mT0 = -Tnb[0] * inv_dT;
    for (int b = 0; b < nB; ++b)
        for (int t = 0; t < nT; ++t) Ib[b * nT + t] = orc_planck_band(Tnb[t], wvc[b]);
    for (int b = 0; b < nB; ++b)
        for (int g = 0; g < nQ; ++g)
            for (int t = 0; t < nT; ++t) kqrow[(b * nQ + g) * nT + t] = kq[(b * nT + t) * nQ + g];
    int tt = (int)((Tmax - Tnb[0]) / (Tnb[1] - Tnb[0]));
    float kappamax = (kP[tt + 1] - kP[tt]) / (Tnb[tt + 1] - Tnb[tt]) * (Tmax - Tnb[tt]) + kP[tt];
    *kappamax_out = kappamax;
    double etot = 0.0;
    double *e = (double *)malloc(sizeof(double) * nB * nQ);
    for (int b = 0; b < nB; ++b) {
        float ibm = orc_lerp_f32(Ib + b * nT, nT, inv_dT, mT0, Tmax);
        const int bw = g_plant == 1 ? (b + 1) % nB : b;
        double dnu = ((double)wvr[bw] - (double)wvl[bw]) * 100.0;
        for (int g = 0; g < nQ; ++g) {
            float km = orc_lerp_f32(kqrow + (b * nQ + g) * nT, nT, inv_dT, mT0, Tmax);
            double em = (double)km * (double)ibm;
            inv_emax[b * nQ + g] = em > 0.0 ? (float)(1.0 / em) : 0.0f;
            const double wg = g_plant == 2 ? wq[nQ - 1 - g] : wq[g];
            e[b * nQ + g] = 4.0 * 3.14159265358979323846 * em * wg * dnu;
            etot += e[b * nQ + g];
        }
    }
    if (mode == 1) {
        int tm = tt;
        int t = ((Tmax - Tnb[tm]) < (Tnb[tm + 1] - Tmax)) ? tm : tm + 1;
        for (int b = 0; b < nB; ++b) {
            cdf_band[b] = prob[t * nB + b];
            for (int g = 0; g < nQ; ++g) cdf_g[b * nQ + g] = probg[(t * nB + b) * nQ + g];
        }
        float k4 = 4.0f * kappamax * Tmax * Tmax * Tmax * Tmax;
        *Ibmax_d = (double)k4 * sigma / (double)R;
    } else {
        double cum = 0.0;
        for (int b = 0; b < nB; ++b) {
            double eb = 0.0;
            for (int g = 0; g < nQ; ++g) eb += e[b * nQ + g];
            double cg = 0.0;
            for (int g = 0; g < nQ; ++g) {
                if (g_plant != 3) cg += e[b * nQ + g];
                cdf_g[b * nQ + g] = eb > 0.0 ? (float)(cg / eb) : (float)(g + 1) / nQ;
                if (g_plant == 3) cg += e[b * nQ + g];
            }
            cdf_g[b * nQ + nQ - 1] = 1.0f;
            cum += eb;
            cdf_band[b] = (float)(cum / etot);
        }
        cdf_band[nB - 1] = 1.0f;
        *Ibmax_d = etot / (double)R;
    }
    free(e);
    return 0;
}
