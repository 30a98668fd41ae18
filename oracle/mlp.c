/*
 * FP64 reference of the proposal network (test infrastructure only).
 *
 * Network: X[nb x nin] -> H1 = tanh(X W1 + b1) -> H2 = tanh(H1 W2 + b2)
 *          -> Z = H2 W3 + b3, Z = [direction logits (ndir) | bin logits].
 * Loss (weighted maximum likelihood of the tallied bundles):
 *   L = sum_s wt_s [-log softmax(Z_dir)[d_s] - log softmax(Z_bin)[b_s]] / sum_s wt_s
 * The GPU forward/backward kernels are checked against these gradients, and
 * these are checked against central finite differences of orc_mlp_loss_grad's
 * loss (tests/test_oracle_mlp.py), so a dropped term or transposed operand
 * fails.  No reference counterpart (the learned proposal is new).
 *
 * The _bf16 variants round the matrix-product operands to BF16 where the
 * kernels do (FP64 -> FP32 -> BF16, round to nearest even at each step):
 * H1 and H2 as operands of the next layer's product and of the weight
 * gradients, dZ and dA2 as operands of the backward products; the tanh
 * derivatives, the bias gradients and dW1 = X^T dA1 use the unrounded values.
 * The rounding itself is pinned against numpy (tests/test_oracle_mlp.py).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

/* BF16 value of x as the kernels form it: x -> FP32 (nearest even), then the
 * FP32 bits rounded to the top 16 (nearest even; non-finite values kept). */
double orc_bf16_round(double x)
{
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, sizeof u);
    if ((u & 0x7F800000u) != 0x7F800000u) u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    memcpy(&f, &u, sizeof f);
    return (double)f;
}

static void round_all(double *a, size_t n, int q)
{
    if (!q) return;
    for (size_t i = 0; i < n; ++i) a[i] = orc_bf16_round(a[i]);
}

static void dense(int nb, int nin, int nout, const double *X, const double *W, const double *b,
                  double *Y, int act)
{
    for (int s = 0; s < nb; ++s)
        for (int o = 0; o < nout; ++o) {
            double acc = b[o];
            for (int i = 0; i < nin; ++i) acc += X[(size_t)s * nin + i] * W[(size_t)i * nout + o];
            Y[(size_t)s * nout + o] = act ? tanh(acc) : acc;
        }
}

/* Forward; with q the products read BF16-rounded H1 / H2 (H1q, H2q), while
 * H1, H2 keep the unrounded activations. */
static void forward_q(int nb, const double *X, const double *W1, const double *b1, const double *W2,
                      const double *b2, const double *W3, const double *b3, int nin, int nh, int nout,
                      double *H1, double *H2, double *Z, double *H1q, double *H2q, int q)
{
    dense(nb, nin, nh, X, W1, b1, H1, 1);
    memcpy(H1q, H1, sizeof(double) * (size_t)nb * nh);
    round_all(H1q, (size_t)nb * nh, q);
    dense(nb, nh, nh, H1q, W2, b2, H2, 1);
    memcpy(H2q, H2, sizeof(double) * (size_t)nb * nh);
    round_all(H2q, (size_t)nb * nh, q);
    dense(nb, nh, nout, H2q, W3, b3, Z, 0);
}

static void forward_api(int nb, const double *X, const double *W1, const double *b1, const double *W2,
                        const double *b2, const double *W3, const double *b3, int nin, int nh, int nout,
                        double *H1, double *H2, double *Z, int q)
{
    double *H1q = (double *)malloc(sizeof(double) * (size_t)nb * nh);
    double *H2q = (double *)malloc(sizeof(double) * (size_t)nb * nh);
    forward_q(nb, X, W1, b1, W2, b2, W3, b3, nin, nh, nout, H1, H2, Z, H1q, H2q, q);
    free(H1q);
    free(H2q);
}

void orc_mlp_forward(int nb, const double *X, const double *W1, const double *b1, const double *W2,
                     const double *b2, const double *W3, const double *b3, int nin, int nh, int nout,
                     double *H1, double *H2, double *Z)
{
    forward_api(nb, X, W1, b1, W2, b2, W3, b3, nin, nh, nout, H1, H2, Z, 0);
}

void orc_mlp_forward_bf16(int nb, const double *X, const double *W1, const double *b1, const double *W2,
                          const double *b2, const double *W3, const double *b3, int nin, int nh, int nout,
                          double *H1, double *H2, double *Z)
{
    forward_api(nb, X, W1, b1, W2, b2, W3, b3, nin, nh, nout, H1, H2, Z, 1);
}

static double softmax_ce(const double *z, int n, int t, double scale, double *dz)
{
    double m = z[0];
    for (int i = 1; i < n; ++i) m = z[i] > m ? z[i] : m;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += exp(z[i] - m);
    double lse = m + log(s);
    for (int i = 0; i < n; ++i) dz[i] = scale * (exp(z[i] - lse) - (i == t ? 1.0 : 0.0));
    return lse - z[t];
}

static void loss_grad_q(int nb, const double *X, const double *W1, const double *b1, const double *W2,
                        const double *b2, const double *W3, const double *b3, int nin, int nh, int nout,
                        int ndir, const int *td, const int *tb, const double *wt, double *loss,
                        double *gW1, double *gb1, double *gW2, double *gb2, double *gW3, double *gb3, int q)
{
    double *H1 = (double *)malloc(sizeof(double) * (size_t)nb * nh);
    double *H2 = (double *)malloc(sizeof(double) * (size_t)nb * nh);
    double *H1q = (double *)malloc(sizeof(double) * (size_t)nb * nh);
    double *H2q = (double *)malloc(sizeof(double) * (size_t)nb * nh);
    double *Z = (double *)malloc(sizeof(double) * (size_t)nb * nout);
    double *dZ = (double *)malloc(sizeof(double) * (size_t)nb * nout);
    double *dZq = (double *)malloc(sizeof(double) * (size_t)nb * nout);
    double *dA2 = (double *)malloc(sizeof(double) * (size_t)nb * nh);
    double *dA2q = (double *)malloc(sizeof(double) * (size_t)nb * nh);
    double *dA1 = (double *)malloc(sizeof(double) * (size_t)nb * nh);
    forward_q(nb, X, W1, b1, W2, b2, W3, b3, nin, nh, nout, H1, H2, Z, H1q, H2q, q);
    double wsum = 0.0;
    for (int s = 0; s < nb; ++s) wsum += wt[s];
    double L = 0.0;
    for (int s = 0; s < nb; ++s) {
        double sc = wt[s] / wsum;
        const double *z = Z + (size_t)s * nout;
        double *dz = dZ + (size_t)s * nout;
        L += sc * softmax_ce(z, ndir, td[s], sc, dz);
        L += sc * softmax_ce(z + ndir, nout - ndir, tb[s], sc, dz + ndir);
    }
    *loss = L;
    memcpy(dZq, dZ, sizeof(double) * (size_t)nb * nout);
    round_all(dZq, (size_t)nb * nout, q);
    memset(gW3, 0, sizeof(double) * (size_t)nh * nout);
    memset(gb3, 0, sizeof(double) * nout);
    for (int s = 0; s < nb; ++s)
        for (int o = 0; o < nout; ++o) {
            gb3[o] += dZ[(size_t)s * nout + o];
            double d = dZq[(size_t)s * nout + o];
            for (int i = 0; i < nh; ++i) gW3[(size_t)i * nout + o] += H2q[(size_t)s * nh + i] * d;
        }
    for (int s = 0; s < nb; ++s)
        for (int i = 0; i < nh; ++i) {
            double acc = 0.0;
            for (int o = 0; o < nout; ++o) acc += dZq[(size_t)s * nout + o] * W3[(size_t)i * nout + o];
            double h = H2[(size_t)s * nh + i];
            dA2[(size_t)s * nh + i] = acc * (1.0 - h * h);
        }
    memcpy(dA2q, dA2, sizeof(double) * (size_t)nb * nh);
    round_all(dA2q, (size_t)nb * nh, q);
    memset(gW2, 0, sizeof(double) * (size_t)nh * nh);
    memset(gb2, 0, sizeof(double) * nh);
    for (int s = 0; s < nb; ++s)
        for (int o = 0; o < nh; ++o) {
            gb2[o] += dA2[(size_t)s * nh + o];
            double d = dA2q[(size_t)s * nh + o];
            for (int i = 0; i < nh; ++i) gW2[(size_t)i * nh + o] += H1q[(size_t)s * nh + i] * d;
        }
    for (int s = 0; s < nb; ++s)
        for (int i = 0; i < nh; ++i) {
            double acc = 0.0;
            for (int o = 0; o < nh; ++o) acc += dA2q[(size_t)s * nh + o] * W2[(size_t)i * nh + o];
            double h = H1[(size_t)s * nh + i];
            dA1[(size_t)s * nh + i] = acc * (1.0 - h * h);
        }
    memset(gW1, 0, sizeof(double) * (size_t)nin * nh);
    memset(gb1, 0, sizeof(double) * nh);
    for (int s = 0; s < nb; ++s)
        for (int o = 0; o < nh; ++o) {
            double d = dA1[(size_t)s * nh + o];
            gb1[o] += d;
            for (int i = 0; i < nin; ++i) gW1[(size_t)i * nh + o] += X[(size_t)s * nin + i] * d;
        }
    free(H1); free(H2); free(H1q); free(H2q); free(Z); free(dZ); free(dZq); free(dA2); free(dA2q); free(dA1);
}

void orc_mlp_loss_grad(int nb, const double *X, const double *W1, const double *b1, const double *W2,
                       const double *b2, const double *W3, const double *b3, int nin, int nh, int nout,
                       int ndir, const int *td, const int *tb, const double *wt, double *loss,
                       double *gW1, double *gb1, double *gW2, double *gb2, double *gW3, double *gb3)
{
    loss_grad_q(nb, X, W1, b1, W2, b2, W3, b3, nin, nh, nout, ndir, td, tb, wt, loss, gW1, gb1, gW2, gb2, gW3,
                gb3, 0);
}

void orc_mlp_loss_grad_bf16(int nb, const double *X, const double *W1, const double *b1, const double *W2,
                            const double *b2, const double *W3, const double *b3, int nin, int nh, int nout,
                            int ndir, const int *td, const int *tb, const double *wt, double *loss,
                            double *gW1, double *gb1, double *gW2, double *gb2, double *gW3, double *gb3)
{
    loss_grad_q(nb, X, W1, b1, W2, b2, W3, b3, nin, nh, nout, ndir, td, tb, wt, loss, gW1, gb1, gW2, gb2, gW3,
                gb3, 1);
}

/* ---------------------------------------------------- learned draw ---- */
/* Defensive-mixture CDF search over one head: F(k) = min(fma(U[k], s, eps P[k]), 1)
 * with s = (1 - eps) / U[n-1] (U: the row's unnormalised inclusive cumulative
 * weights, P: the plain proposal's inclusive CDF), F(n-1) = 1, F(-1) = 0; the
 * bin is found with the search semantics of orc_cdf_search on F, and q = F(k) -
 * F(k-1) is its mixture probability. */
static float mix_F(const float *U, const float *P, int n, float s, float eps, int k)
{
    if (k < 0) return 0.0f;
    if (k == n - 1) return 1.0f;
    return fminf(fmaf(U[k], s, eps * P[k]), 1.0f);
}

static int mix_search(const float *U, const float *P, int n, float eps, float R, float *qk)
{
    const float s = (1.0f - eps) / U[n - 1];
    int lo = 0, w = n;
    while (w > lo + 1) {
        const int d = (lo + w) >> 1;
        const float fd = mix_F(U, P, n, s, eps, d);
        if (R < fd) w = d;
        if (R >= fd) lo = d;
    }
    int k = lo + 1;
    if (R < mix_F(U, P, n, s, eps, 0)) k = 0;
    if (R == 1.0f) k = n - 1;
    if (k > n - 1) k = n - 1;
    *qk = mix_F(U, P, n, s, eps, k) - mix_F(U, P, n, s, eps, k - 1);
    return k;
}

void orc_learned_draw(int nslots, int R, int ndir, int nbin, int nQ, int noutp, const float *U, const float *pcdf,
                      const float *p_bg, const int *cells, uint32_t snap, uint64_t seed, float eps, uint32_t *w0,
                      uint32_t *w1, float *ratio)
{
    float *Pd = (float *)malloc(sizeof(float) * (size_t)ndir);
    for (int e = 0; e < ndir; ++e) Pd[e] = (float)(e + 1) * (1.0f / (float)ndir);
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int sl = 0; sl < nslots; ++sl) {
        const float *Ud = U + (size_t)sl * noutp;
        for (int r = 0; r < R; ++r) {
            const uint32_t ctr[4] = {(uint32_t)r, (uint32_t)cells[sl], snap, 1u};
            uint32_t x[4];
            orc_philox4x32_10(ctr, key, x);
            float qd, qb;
            const int d = mix_search(Ud, Pd, ndir, eps, orc_u01_open(x[0]), &qd);
            const int bg = mix_search(Ud + ndir, pcdf, nbin, eps, orc_u01_open(x[1]), &qb);
            const int band = bg / nQ, g = bg - band * nQ;
            const float rt = (qb > 0.f && qd > 0.f) ? (p_bg[bg] / qb) * ((1.0f / (float)ndir) / qd) : 0.f;
            /* direction bin d = 8 cos(theta) bins x 8 phi bins: the bin index in
             * the top 3 of the 24 uniform bits, the Philox bits below it */
            const uint32_t mt = ((uint32_t)(d >> 3) << 21) | (x[2] >> 11);
            const uint32_t mp = ((uint32_t)(d & 7) << 21) | (x[3] >> 11);
            const size_t q = (size_t)sl * R + r;
            w0[q] = (mt & 0xFFFFFFu) | ((uint32_t)(band & 0xFF) << 24);
            w1[q] = (mp & 0xFFFFFFu) | ((uint32_t)((band >> 8) & 0x3) << 24) | ((uint32_t)(g & 0x3F) << 26);
            ratio[q] = rt;
        }
    }
    free(Pd);
}
