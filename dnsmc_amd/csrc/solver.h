// Internal types of the MI355X radiation solver (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/dnsmc_rad.h"
#include "dmath.h"

namespace dn {

constexpr int kMaxLev = 6;

// Spectral tables as read (reference layouts, code/NarrowBand.h:3-8 and
// code/read_tables.cpp): kq[b][t][g], prob[t][b] band CDF, probg[t][b][g] g-CDF.
struct Tables {
    int nB = 0, nQ = 0, nT = 0;
    std::vector<float> Tnb, kP, wvl, wvr, wvc, kq, prob, probg, kavg;
    bool synth_grey = false;   // grey==1 tables depend on Tmax (code/read_tables.cpp:104-183)
};

struct Level {
    int n[3] = {0, 0, 0};
    std::vector<float> faces[3];    // n+1 each
    std::vector<float> centres[3];  // n+2 each (ghost-inclusive)
    int maxs = 100000;
};

// Everything the kernels need that is derived on the host from tables + Tmax.
struct Derived {
    int nB = 1, nQ = 1, nT = 2;
    float inv_dT = 0.f, mT0 = 0.f;
    std::vector<float> kq;       // kappa rows [(b*nQ+g)*nT + t]
    std::vector<float> Ib;       // [b*nT + t]
    std::vector<float> Ibw;      // [(axis*2+side)*nB + b]
    std::vector<float> cdf_band; // nB
    std::vector<float> cdf_g;    // nB*nQ
    std::vector<float> inv_emax; // nB*nQ
    std::vector<double> wq;      // quadrature weights (nQ)
    float Tmax = 0.f, kappamax = 0.f, Ibmax = 0.f;
    double Ibmax_d = 0.0;
    int g_search = 1;
    float u_off = 0.5f;
    int sorted = 0;              // compat srt == 1 on gas tables
    std::vector<int> band_order; // sorted: band at each position of the ascending-kavg order
};

// Host-side builders (host.cpp).
void default_config(dnsmc_config* c);
int parse_config(const char* text, dnsmc_config* c, std::string* err);
int build_levels(const dnsmc_config& c, std::vector<Level>& lev, std::string* err);
void grid_axis(int stretch, int n, int n0, float L, float beta, std::vector<float>& f, std::vector<float>& cc);
float planck_band(float T, float nu);
int read_tables(const dnsmc_config& c, Tables& t, std::string* err);
void synth_grey_tables(Tables& t, float Tmax);
int build_derived(const dnsmc_config& c, const Tables& t, float Tmax, Derived& d, std::string* err);
void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w);

// Natural-spline coefficients for one axis: fine centres -> coarse centres.
struct SplineAxis {
    int nf = 0, nc = 0;
    // forward sweep per fine node i=1..nf-2: sig, 1/p, y2 factor, 6/(x[i+1]-x[i-1]), 1/h_lo, 1/h_hi
    std::vector<double> sig, invp, y2f, c6, ihl, ihh;
    // evaluation per coarse node: klo, a, b, A=(a^3-a)h^2/6, B=(b^3-b)h^2/6
    std::vector<int> klo;
    std::vector<double> a, b, A, B;
};
void spline_axis_coeffs(const std::vector<float>& xf_centres, int nf, const std::vector<float>& xc_centres,
                        int nc, SplineAxis& s);

// XORWOW jump matrices: M^(2^(67+j)) for j in [0, 60), each 160 columns x 5 words.
void xorwow_jump_tables(std::vector<uint32_t>& out);
void xorwow_seed_state(uint64_t seed, uint32_t st[6]);

}  // namespace dn
