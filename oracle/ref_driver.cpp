// Known-answer vectors from the reference's own CPU sources.
//
// TEST INFRASTRUCTURE ONLY (checker of the oracle; never part of the product).
//
// Built by `make -C oracle ref` against code/spline.cpp and code/read_tables.cpp
// as they lie in the reference tree (with its param.h defaults: imax=jmax=kmax=64,
// grid_num=5, grey=2), and run from oracle/_ref/code so that the reference's
// relative table paths ../distr/tables/particles/... resolve to the particle
// tables kept under tests/golden/.  Writes, as text, "name count" headers
// followed by the values:
//   - the grid of every level, Var_CPU::mk_grid (code/definitions.h:180-247);
//   - the coarse-level temperatures from interp3D (code/spline.cpp:35-116) of a
//     smooth analytic level-0 field;
//   - readT's particle-model outputs (code/read_tables.cpp:187-241) at Tmax=750 K.
// The driver is the caller that code/mc_gpu.cu would be: it defines the
// proposal arrays mc_gpu.cu defines and calls the reference functions with the
// arguments mc_gpu_ passes them (code/mc_gpu.cu:453,468,503).
#include <math.h>
#include <stdio.h>

#include "param.h"
#include "definitions.h"
#include "functions.h"

myfloat prob[nT][nB], probg[nT][nB][nQ];

static void dump(const char* name, const myfloat* v, int n)
{
    printf("%s %d\n", name, n);
    for (int i = 0; i < n; ++i) printf("%.9e\n", (double)v[i]);
}

int main()
{
    Var_CPU g[grid_num];
    for (int l = 0; l < grid_num; ++l) g[l].mk_grid(maxi[l], maxj[l], maxk[l], maxi[0]);
    char nm[64];
    for (int l = 0; l < grid_num; ++l) {
        snprintf(nm, sizeof nm, "x%d", l);  dump(nm, g[l].x, g[l].im + 2);
        snprintf(nm, sizeof nm, "xu%d", l); dump(nm, g[l].xu, g[l].im + 1);
        snprintf(nm, sizeof nm, "y%d", l);  dump(nm, g[l].y, g[l].jm + 2);
        snprintf(nm, sizeof nm, "yu%d", l); dump(nm, g[l].yu, g[l].jm + 1);
        snprintf(nm, sizeof nm, "z%d", l);  dump(nm, g[l].z, g[l].km + 2);
        snprintf(nm, sizeof nm, "zu%d", l); dump(nm, g[l].zu, g[l].km + 1);
    }
    // smooth level-0 field on the interior cell centres
    for (int k = 1; k <= kmax; ++k)
        for (int j = 1; j <= jmax; ++j)
            for (int i = 1; i <= imax; ++i) {
                double x = g[0].x[i], y = g[0].y[j], z = g[0].z[k];
                g[0].T[idx_T(i, j, k, imax, jmax)] =
                    (myfloat)(800.0 + 120.0 * sin(6.283185307179586 * x) * cos(3.141592653589793 * y) +
                              60.0 * z * z - 40.0 * x * y + 25.0 * cos(9.42477796076938 * z));
            }
    {
        myfloat* t0 = new myfloat[imax * jmax * kmax];
        int o = 0;
        for (int k = 1; k <= kmax; ++k)
            for (int j = 1; j <= jmax; ++j)
                for (int i = 1; i <= imax; ++i) t0[o++] = g[0].T[idx_T(i, j, k, imax, jmax)];
        dump("T0", t0, o);
        delete[] t0;
    }
    interp3D(g);
    for (int l = 1; l < grid_num; ++l) {
        const int im = g[l].im, jm = g[l].jm, km = g[l].km;
        myfloat* tl = new myfloat[im * jm * km];
        int o = 0;
        for (int k = 1; k <= km; ++k)
            for (int j = 1; j <= jm; ++j)
                for (int i = 1; i <= im; ++i) tl[o++] = g[l].T[idx_T(i, j, k, im, jm)];
        snprintf(nm, sizeof nm, "T%d", l);
        dump(nm, tl, o);
        delete[] tl;
    }
    // particle tables
    static NarrowBand nb[nB];
    static myfloat Tnb[nT], kP[nT], prob_h[nT * nB], probg_h[nT * nB * nQ];
    myfloat kappamax = 0;
    const myfloat Tmax = 750.0f;
    readT(nb, Tnb, kP, prob_h, probg_h, Tmax, &kappamax);
    dump("Tnb", Tnb, nT);
    dump("kP", kP, nT);
    dump("kappamax", &kappamax, 1);
    myfloat wvc[nB], kq00[nB];
    for (int b = 0; b < nB; ++b) {
        wvc[b] = nb[b].wvc;
        kq00[b] = nb[b].kq[0][0];
    }
    dump("wvc", wvc, nB);
    dump("kq", kq00, nB);
    dump("prob", prob_h, nT * nB);
    return 0;
}
