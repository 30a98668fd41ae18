/*
 * Oracle grids, Planck function, natural cubic spline and plane-parallel
 * closed forms (test infrastructure only; see oracle.h).
 *
 * Pins:
 *  - orc_grid_ref_quad: the reference-held fixtures code/grid.txt (im=190,
 *    L=2, faces and centres) and code/resfortm.txt column 1 (im=64, L=1).
 *  - orc_grid_tanh / orc_grid_uniform: closed forms (end faces exactly 0 and L,
 *    symmetry f[n-j] = L - f[j]).
 *  - orc_planck_band: band sum x width against sigma*T^4/pi (scipy quad).
 *  - orc_interp3d: scipy.interpolate.CubicSpline(bc_type='natural') applied
 *    per axis in the same order; exact for data linear in each coordinate.
 *  - orc_expint_En: scipy.special.expn; orc_slab_*: the isothermal E2 form and
 *    brute-force FP64 ray marching on small slabs.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* code/definitions.h:108-134 (Gridn::mk_grid, x faces) with the reference's
 * mixed float/double expression types kept: dx = 0.5-1.45(r-0.5)^2 on the
 * first half, normalised to L/2, mirrored; coarse faces subsample the fine
 * ones xu[i] = xu0[i*im0/im]; centres are face midpoints. */
void orc_grid_ref_quad(int im, int im0, float L, float *xu, float *x)
{
    float *xu0 = (float *)malloc(sizeof(float) * (size_t)(im0 + 1));
    xu0[0] = 0.0f;
    for (int i = 1; i < im0 / 2 + 1; ++i) {
        float r = (1.0f * (float)i) / (float)im0;
        float dx = (float)(0.5 - 1.45 * ((double)r - 0.5) * ((double)r - 0.5));
        xu0[i] = xu0[i - 1] + dx;
    }
    float xnorm = xu0[im0 / 2];
    for (int i = 1; i < im0 / 2 + 1; ++i)
        xu0[i] = (float)((double)L / 2.0 * (double)xu0[i] / (double)xnorm);
    for (int i = im0; i > im0 / 2; --i)
        xu0[i] = L - xu0[im0 - i];
    xu[0] = 0.0f;
    for (int i = 1; i < im + 1; ++i) xu[i] = xu0[(long)i * im0 / im];
    for (int i = 1; i < im + 1; ++i) x[i] = (xu[i] + xu[i - 1]) * 0.5f;
    x[0] = -x[1];
    x[im + 1] = xu[im] + (xu[im] - x[im]);
    free(xu0);
}

/* code/definitions.h:85-105 (y, z faces i*dy, centres midpoints). */
void orc_grid_uniform(int n, float L, float *f, float *c)
{
    float d = L / (float)n;
    for (int i = 0; i < n + 1; ++i) f[i] = (float)i * d;
    c[0] = -d / 2.0f;
    for (int i = 1; i < n + 1; ++i) c[i] = (f[i - 1] + f[i]) / 2.0f;
    c[n + 1] = c[n] + d;
}

/* Wall-normal tanh stretch of the channel workloads. */
void orc_grid_tanh(int n, float L, double beta, float *f, float *c)
{
    double tb = tanh(beta);
    for (int j = 0; j < n + 1; ++j) {
        double e = 2.0 * (double)j / (double)n - 1.0;
        f[j] = (float)(0.5 * (double)L * (1.0 + tanh(beta * e) / tb));
    }
    f[0] = 0.0f;
    f[n] = L;
    for (int i = 1; i < n + 1; ++i) c[i] = (f[i - 1] + f[i]) * 0.5f;
    c[0] = -c[1];
    c[n + 1] = f[n] + (f[n] - c[n]);
}

/* code/memory_copy.cu:291-300 I_blackC with its expression types: the pow3
 * macro expands textually to nu*100*nu*100*nu*100 (float chain). */
float orc_planck_band(float T, float nu)
{
    const float C1 = 3.741771790075259e-16f;
    const float C2 = 0.014387741858429f;
    float p3 = nu * 100.0f;
    p3 = p3 * nu;
    p3 = p3 * 100.0f;
    p3 = p3 * nu;
    p3 = p3 * 100.0f;
    float den = expf(C2 * nu * 100.0f / T) - 1.0f;
    return (float)(1.0 / 3.14159265358979323846 * (double)C1 * (double)p3 / (double)den);
}

/* ------------------------------------------------------------- spline ---- */
/* Natural cubic spline second derivatives (Numerical Recipes "spline" with
 * yp1 = ypn = 1e30, as called in code/spline.cpp:55,77,98), 0-based here. */
static void spline_d2(const double *x, const double *y, int n, double *y2, double *u)
{
    y2[0] = u[0] = 0.0;
    for (int i = 1; i < n - 1; ++i) {
        double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        u[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
        u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    y2[n - 1] = 0.0;
    for (int k = n - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
}

static double splint_eval(const double *xa, const double *ya, const double *y2, int n, double x)
{
    int klo = 0, khi = n - 1;
    while (khi - klo > 1) {
        int k = (khi + klo) >> 1;
        if (xa[k] > x) khi = k; else klo = k;
    }
    double h = xa[khi] - xa[klo];
    double a = (xa[khi] - x) / h, b = (x - xa[klo]) / h;
    return a * ya[klo] + b * ya[khi] + ((a * a * a - a) * y2[klo] + (b * b * b - b) * y2[khi]) * (h * h) / 6.0;
}

/* code/spline.cpp:35-116: x-lines, then y-lines, then z-lines. Centres are
 * given ghost-inclusive (index 1..n are the interior). */
void orc_interp3d(const int nf[3], const float *xf, const float *yf, const float *zf,
                  const float *Tf, const int nc[3], const float *xc, const float *yc,
                  const float *zc, float *Tc)
{
    int nfx = nf[0], nfy = nf[1], nfz = nf[2];
    int ncx = nc[0], ncy = nc[1], ncz = nc[2];
    int nmax = nfx > nfy ? nfx : nfy;
    if (nfz > nmax) nmax = nfz;
    double *ax = (double *)malloc(sizeof(double) * nmax), *ay = (double *)malloc(sizeof(double) * nmax);
    double *a2 = (double *)malloc(sizeof(double) * nmax), *au = (double *)malloc(sizeof(double) * nmax);
    double *tx = (double *)malloc(sizeof(double) * (size_t)ncx * nfy * nfz);
    double *txy = (double *)malloc(sizeof(double) * (size_t)ncx * ncy * nfz);
    for (int k = 0; k < nfz; ++k)
        for (int j = 0; j < nfy; ++j) {
            for (int i = 0; i < nfx; ++i) { ax[i] = xf[i + 1]; ay[i] = Tf[i + (size_t)nfx * (j + (size_t)nfy * k)]; }
            spline_d2(ax, ay, nfx, a2, au);
            for (int i = 0; i < ncx; ++i)
                tx[i + (size_t)ncx * (j + (size_t)nfy * k)] = splint_eval(ax, ay, a2, nfx, xc[i + 1]);
        }
    for (int k = 0; k < nfz; ++k)
        for (int i = 0; i < ncx; ++i) {
            for (int j = 0; j < nfy; ++j) { ax[j] = yf[j + 1]; ay[j] = tx[i + (size_t)ncx * (j + (size_t)nfy * k)]; }
            spline_d2(ax, ay, nfy, a2, au);
            for (int j = 0; j < ncy; ++j)
                txy[i + (size_t)ncx * (j + (size_t)ncy * k)] = splint_eval(ax, ay, a2, nfy, yc[j + 1]);
        }
    for (int j = 0; j < ncy; ++j)
        for (int i = 0; i < ncx; ++i) {
            for (int k = 0; k < nfz; ++k) { ax[k] = zf[k + 1]; ay[k] = txy[i + (size_t)ncx * (j + (size_t)ncy * k)]; }
            spline_d2(ax, ay, nfz, a2, au);
            for (int k = 0; k < ncz; ++k)
                Tc[i + (size_t)ncx * (j + (size_t)ncy * k)] = (float)splint_eval(ax, ay, a2, nfz, zc[k + 1]);
        }
    free(ax); free(ay); free(a2); free(au); free(tx); free(txy);
}

/* ------------------------------------------------------- closed forms ---- */
/* Exponential integral E_n(x), n >= 0, x > 0 (series for x <= 1, Lentz
 * continued fraction otherwise; Numerical Recipes 6.3 method). */
double orc_expint_En(int n, double x)
{
    const double EULER = 0.57721566490153286061;
    const int MAXIT = 1000;
    const double EPS = 1e-16, BIG = 1e300;
    if (n == 0) return exp(-x) / x;
    if (x == 0.0) return 1.0 / (double)(n - 1);
    if (x > 1.0) {
        double b = x + n, c = BIG, d = 1.0 / b, h = d;
        for (int i = 1; i <= MAXIT; ++i) {
            double an = -(double)i * (double)(n - 1 + i);
            b += 2.0;
            d = 1.0 / (an * d + b);
            c = b + an / c;
            double del = c * d;
            h *= del;
            if (fabs(del - 1.0) < EPS) break;
        }
        return h * exp(-x);
    }
    int nm1 = n - 1;
    double ans = (nm1 != 0) ? 1.0 / nm1 : -log(x) - EULER;
    double fact = 1.0;
    for (int i = 1; i <= MAXIT; ++i) {
        fact *= -x / i;
        double del;
        if (i != nm1) {
            del = -fact / (i - nm1);
        } else {
            double psi = -EULER;
            for (int ii = 1; ii <= nm1; ++ii) psi += 1.0 / ii;
            del = fact * (-log(x) + psi);
        }
        ans += del;
        if (fabs(del) < fabs(ans) * EPS) break;
    }
    return ans;
}

/* Expected OERM tally of a point emitter at every layer centre of a
 * plane-parallel medium (layers j = 0..ny-1 between faces[j] and
 * faces[j+1], black walls at faces[0] and faces[ny]), for one spectral bin
 * with per-layer kappa and Ib:
 *   F_i = sum_j P_ij (1 - Ib_j/Ib_i) + P_i,lo (1 - Ib_lo/Ib_i) + P_i,hi (...)
 * with P_ij = 1/2 [E2(tau_near) - E2(tau_far)] (isotropic emission, optical
 * depth measured along the wall normal). No cut-off. */
void orc_slab_exchange(int ny, const double *f, const double *kap, const double *Ib,
                       double Ib_lo, double Ib_hi, double *F)
{
    double *tau = (double *)malloc(sizeof(double) * (ny + 1)); /* optical depth of faces from f[0] */
    tau[0] = 0.0;
    for (int j = 0; j < ny; ++j) tau[j + 1] = tau[j] + kap[j] * (f[j + 1] - f[j]);
    for (int i = 0; i < ny; ++i) {
        double ci = 0.5 * (f[i] + f[i + 1]);
        double ti = tau[i] + kap[i] * (ci - f[i]);
        double acc = 0.0;
        for (int j = 0; j < ny; ++j) {
            if (j == i) continue;
            double tn, tf;
            if (j > i) { tn = tau[j] - ti; tf = tau[j + 1] - ti; }
            else { tn = ti - tau[j + 1]; tf = ti - tau[j]; }
            double P = 0.5 * (orc_expint_En(2, tn) - orc_expint_En(2, tf));
            acc += P * (1.0 - Ib[j] / Ib[i]);
        }
        acc += 0.5 * orc_expint_En(2, ti) * (1.0 - Ib_lo / Ib[i]);
        acc += 0.5 * orc_expint_En(2, tau[ny] - ti) * (1.0 - Ib_hi / Ib[i]);
        F[i] = acc;
    }
    free(tau);
}

/* Grey slab (kappa constant, Ib ~ T^4): Q_i = 4 kappa sigma T_i^4 F_i. */
void orc_slab_grey_Q(int ny, const double *faces, const double *T, double kappa,
                     double Tw_lo, double Tw_hi, double *Q)
{
    const double sigma = 5.670373e-08;
    if (ny < 1) return;
    double *kap = (double *)malloc(sizeof(double) * ny), *Ib = (double *)malloc(sizeof(double) * ny);
    for (int j = 0; j < ny; ++j) { kap[j] = kappa; Ib[j] = T[j] * T[j] * T[j] * T[j]; }
    orc_slab_exchange(ny, faces, kap, Ib, pow(Tw_lo, 4), pow(Tw_hi, 4), Q);
    for (int j = 0; j < ny; ++j) Q[j] *= 4.0 * kappa * sigma * Ib[j];
    free(kap); free(Ib);
}

/* Same expectation but with the estimator's optical-depth cut-off: once the
 * optical path from the emitter reaches tau_cut, the remaining transmitted
 * fraction exp(-tau_cut) is deposited in the layer reached.  Integrated over
 * mu = cos(theta) by composite Gauss-Legendre (4 points per panel). */
void orc_slab_exchange_cut(int ny, const double *f, const double *kap, const double *Ib,
                           double Ib_lo, double Ib_hi, double tau_cut, int npanel, double *F)
{
    static const double gx[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
    static const double gw[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};
    for (int i = 0; i < ny; ++i) {
        double ci = 0.5 * (f[i] + f[i + 1]);
        double acc = 0.0;
        for (int dir = 0; dir < 2; ++dir) {
            for (int p = 0; p < npanel; ++p) {
                double a = (double)p / npanel, b = (double)(p + 1) / npanel;
                for (int q = 0; q < 4; ++q) {
                    double mu = 0.5 * (a + b) + 0.5 * (b - a) * gx[q];
                    double wq = 0.5 * (b - a) * gw[q];
                    /* march from ci along +/- wall normal with path factor 1/mu */
                    double tra = 1.0, tau = 0.0, val = 0.0;
                    int j = i;
                    double pos = ci;
                    for (;;) {
                        double face = dir ? f[j + 1] : f[j];
                        double len = fabs(face - pos) / mu;
                        double dt = kap[j] * len;
                        double r = 1.0 - Ib[j] / Ib[i];
                        if (tau + dt >= tau_cut) { val += tra * r; break; }
                        double tra1 = exp(-(tau + dt));
                        val += (tra - tra1) * r;
                        tra = tra1; tau += dt; pos = face;
                        j += dir ? 1 : -1;
                        if (j < 0) { val += tra * (1.0 - Ib_lo / Ib[i]); break; }
                        if (j >= ny) { val += tra * (1.0 - Ib_hi / Ib[i]); break; }
                    }
                    acc += 0.5 * wq * val;
                }
            }
        }
        F[i] = acc;
    }
}
