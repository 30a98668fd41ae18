// Host-side preparation: runtime configuration, grids, spectral tables, the
// emission proposal, spline coefficients and XORWOW jump matrices.
//
// Reference counterparts: code/param.h (compile-time configuration),
// code/definitions.h:60-160 (Gridn::mk_grid), code/read_tables.cpp (readT),
// code/memory_copy.cu:27-87,291-300 (wall emission, I_blackC),
// code/mc_gpu.cu:470-503,584-590 (Tmax, kappamax, proposal rows).
#include <dirent.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <sstream>

#include "solver.h"

namespace dn {

// ------------------------------------------------------------- config ----
// Defaults = the reference's shipped code/param.h and code/NarrowBand.h.
void default_config(dnsmc_config* c)
{
    memset(c, 0, sizeof(*c));
    c->n[0] = c->n[1] = c->n[2] = 64;
    c->L[0] = c->L[1] = c->L[2] = 1.0f;
    c->stretch[0] = DNSMC_STRETCH_REF_QUAD;
    c->stretch[1] = c->stretch[2] = DNSMC_STRETCH_UNIFORM;
    c->stretch_beta[0] = c->stretch_beta[1] = c->stretch_beta[2] = 1.5f;
    c->bc[0][0] = c->bc[0][1] = DNSMC_BC_BLACK;      // bdw = bde = 2
    c->bc[1][0] = c->bc[1][1] = DNSMC_BC_PERIODIC;   // bds = bdn = 1
    c->bc[2][0] = c->bc[2][1] = DNSMC_BC_PERIODIC;   // bdb = bdt = 1
    c->Tw[0][0] = c->Tw[0][1] = 500.0f;              // Tww = Twe = 500
    c->polar_axis = 0;
    c->nvar = 3;
    c->photmax = 20000;
    c->rays_per_cell = 100;
    c->grid_num = 5;
    c->maxs[0] = c->maxs[1] = c->maxs[2] = c->maxs[3] = 5;
    c->maxs[4] = 100000;
    c->maxs[5] = 100000;
    c->toll = 1.0e-3f;
    c->spectral = DNSMC_SPEC_PARTICLE;               // grey = 2
    c->kappa_grey = 1.0f;
    c->mode = DNSMC_MODE_COMPAT;
    c->sobol = 1;
    c->num_streams = 16;
    c->num_gpu = 1;
    c->gpu = 0;
    c->p_row = 1;
    c->seed = 1234;
    c->nsnap = 1;
    c->precision = 0;
    c->tally = DNSMC_TALLY_F64_ATOMIC;
    c->srt = 1;                                      // code/param.h:47
    c->batch_rays = 1ll << 26;
    snprintf(c->table_dir, sizeof(c->table_dir), "%s", "../distr/tables");
}

static bool parse_list(const std::string& v, std::vector<double>& out)
{
    out.clear();
    std::string t = v;
    for (char& ch : t)
        if (ch == ',') ch = ' ';
    std::istringstream is(t);
    double x;
    while (is >> x) out.push_back(x);
    return !out.empty();
}

int parse_config(const char* text, dnsmc_config* c, std::string* err)
{
    std::istringstream in(text ? text : "");
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        size_t h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        size_t eq = line.find('=');
        if (eq == std::string::npos) {
            if (line.find_first_not_of(" \t\r") != std::string::npos) {
                *err = "config line " + std::to_string(lineno) + ": expected key = value";
                return DNSMC_EINVAL;
            }
            continue;
        }
        auto trim = [](std::string s) {
            size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
            return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
        };
        std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
        std::vector<double> L;
        auto need = [&](size_t n) {
            if (!parse_list(v, L) || L.size() != n) {
                *err = "config key '" + k + "' expects " + std::to_string(n) + " number(s)";
                return false;
            }
            return true;
        };
        if (k == "n") { if (!need(3)) return DNSMC_EINVAL; for (int a = 0; a < 3; ++a) c->n[a] = (int)L[a]; }
        else if (k == "L") { if (!need(3)) return DNSMC_EINVAL; for (int a = 0; a < 3; ++a) c->L[a] = (float)L[a]; }
        else if (k == "stretch") { if (!need(3)) return DNSMC_EINVAL; for (int a = 0; a < 3; ++a) c->stretch[a] = (int)L[a]; }
        else if (k == "stretch_beta") { if (!need(3)) return DNSMC_EINVAL; for (int a = 0; a < 3; ++a) c->stretch_beta[a] = (float)L[a]; }
        else if (k == "bc") { if (!need(6)) return DNSMC_EINVAL; for (int a = 0; a < 6; ++a) c->bc[a / 2][a % 2] = (int)L[a]; }
        else if (k == "Tw") { if (!need(6)) return DNSMC_EINVAL; for (int a = 0; a < 6; ++a) c->Tw[a / 2][a % 2] = (float)L[a]; }
        else if (k == "maxs") {
            if (!parse_list(v, L) || L.size() > (size_t)kMaxLev) { *err = "maxs: 1..6 numbers"; return DNSMC_EINVAL; }
            for (size_t a = 0; a < L.size(); ++a) c->maxs[a] = (int)L[a];
        }
        else if (k == "table_dir") { snprintf(c->table_dir, sizeof(c->table_dir), "%s", v.c_str()); }
        else {
            if (!need(1)) return DNSMC_EINVAL;
            double x = L[0];
            if (k == "polar_axis") c->polar_axis = (int)x;
            else if (k == "nvar") c->nvar = (int)x;
            else if (k == "photmax") c->photmax = (int)x;
            else if (k == "rays_per_cell") c->rays_per_cell = (int)x;
            else if (k == "grid_num") c->grid_num = (int)x;
            else if (k == "toll") c->toll = (float)x;
            else if (k == "spectral") c->spectral = (int)x;
            else if (k == "kappa_grey") c->kappa_grey = (float)x;
            else if (k == "mode") c->mode = (int)x;
            else if (k == "sobol") c->sobol = (int)x;
            else if (k == "num_streams") c->num_streams = (int)x;
            else if (k == "num_gpu") c->num_gpu = (int)x;
            else if (k == "gpu") c->gpu = (int)x;
            else if (k == "p_row") c->p_row = (int)x;
            else if (k == "seed") c->seed = (unsigned long long)x;
            else if (k == "nsnap") c->nsnap = (int)x;
            else if (k == "precision") c->precision = (int)x;
            else if (k == "tally") c->tally = (int)x;
            else if (k == "srt") c->srt = (int)x;
            else if (k == "batch_rays") c->batch_rays = (long long)x;
            else { *err = "unknown config key '" + k + "'"; return DNSMC_EINVAL; }
        }
    }
    return DNSMC_OK;
}

// -------------------------------------------------------------- grids ----
// code/definitions.h:108-134, expression types kept (float variables, double
// literals): dx = 0.5 - 1.45 (r - 0.5)^2 on the first half, normalised to L/2,
// mirrored; coarse faces subsample the finest ones xu[i] = xu0[i*im0/im].
static void faces_ref_quad(int im0, float L, std::vector<float>& xu0)
{
    xu0.assign(im0 + 1, 0.0f);
    for (int i = 1; i < im0 / 2 + 1; ++i) {
        float r = (1.0f * (float)i) / (float)im0;
        float dx = (float)(0.5 - 1.45 * ((double)r - 0.5) * ((double)r - 0.5));
        xu0[i] = xu0[i - 1] + dx;
    }
    float xnorm = xu0[im0 / 2];
    for (int i = 1; i < im0 / 2 + 1; ++i) xu0[i] = (float)((double)L / 2.0 * (double)xu0[i] / (double)xnorm);
    for (int i = im0; i > im0 / 2; --i) xu0[i] = L - xu0[im0 - i];
}

void grid_axis(int stretch, int n, int n0, float L, float beta, std::vector<float>& f, std::vector<float>& cc)
{
    std::vector<float> f0;
    if (stretch == DNSMC_STRETCH_REF_QUAD) {
        faces_ref_quad(n0, L, f0);
    } else if (stretch == DNSMC_STRETCH_TANH) {
        f0.resize(n0 + 1);
        double tb = tanh((double)beta);
        for (int j = 0; j <= n0; ++j) {
            double e = 2.0 * (double)j / (double)n0 - 1.0;
            f0[j] = (float)(0.5 * (double)L * (1.0 + tanh((double)beta * e) / tb));
        }
        f0[0] = 0.0f;
        f0[n0] = L;
    } else {
        f0.resize(n0 + 1);
        float d = L / (float)n0;   // code/definitions.h:85-91: yu[i] = i*dy
        for (int i = 0; i <= n0; ++i) f0[i] = (float)i * d;
    }
    f.resize(n + 1);
    f[0] = 0.0f;
    for (int i = 1; i <= n; ++i) f[i] = f0[(long)i * n0 / n];
    cc.resize(n + 2);
    for (int i = 1; i <= n; ++i) cc[i] = (f[i] + f[i - 1]) * 0.5f;
    if (stretch == DNSMC_STRETCH_UNIFORM) {
        float d = f[1] - f[0];
        cc[0] = -d / 2.0f;
        cc[n + 1] = cc[n] + d;
    } else {
        cc[0] = -cc[1];
        cc[n + 1] = f[n] + (f[n] - cc[n]);
    }
}

int build_levels(const dnsmc_config& c, std::vector<Level>& lev, std::string* err)
{
    if (c.grid_num < 1 || c.grid_num > kMaxLev) { *err = "grid_num must be in [1, 6]"; return DNSMC_EINVAL; }
    for (int a = 0; a < 3; ++a) {
        if (c.n[a] < 1) { *err = "cell counts must be positive"; return DNSMC_EINVAL; }
        if (c.n[a] % (1 << (c.grid_num - 1)) != 0) {
            *err = "cell counts must be divisible by 2^(grid_num-1)";
            return DNSMC_EINVAL;
        }
        if (c.stretch[a] == DNSMC_STRETCH_REF_QUAD && c.n[a] % 2) { *err = "stretched axis needs even n"; return DNSMC_EINVAL; }
        if (!(c.L[a] > 0.f)) { *err = "domain lengths must be positive"; return DNSMC_EINVAL; }
    }
    lev.assign(c.grid_num, Level());
    for (int g = 0; g < c.grid_num; ++g) {
        for (int a = 0; a < 3; ++a) {
            lev[g].n[a] = c.n[a] >> g;   // code/param.h:21-23: maxi = {imax, imax/2, ...}
            grid_axis(c.stretch[a], lev[g].n[a], c.n[a], c.L[a], c.stretch_beta[a], lev[g].faces[a], lev[g].centres[a]);
        }
        lev[g].maxs = c.maxs[g] > 0 ? c.maxs[g] : 100000;
    }
    return DNSMC_OK;
}

// ---------------------------------------------------------- spectra ----
// code/memory_copy.cu:291-300 (I_blackC) with its expression types: the pow3
// macro expands textually to the float chain nu*100*nu*100*nu*100.
float planck_band(float T, float nu)
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

// Gauss-Legendre nodes and weights mapped to [0, 1] (code/NarrowBand.h:61-67
// lists the 16-point set with 8 digits).
void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w)
{
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            double z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) < 1e-15) break;
        }
        x[n - 1 - i] = 0.5 * (1.0 - z) ;
        w[n - 1 - i] = 1.0 / ((1.0 - z * z) * pp * pp);
    }
    // ascending order in [0,1]
    std::vector<std::pair<double, double>> p(n);
    for (int i = 0; i < n; ++i) p[i] = {x[i], w[i]};
    std::sort(p.begin(), p.end());
    for (int i = 0; i < n; ++i) { x[i] = p[i].first; w[i] = p[i].second; }
}

static int count_lines(FILE* fp)
{
    int n = 0, ch, last = '\n';
    while ((ch = fgetc(fp)) != EOF) {
        if (ch == '\n' && last != '\n') ++n;
        last = ch;
    }
    if (last != '\n') ++n;
    rewind(fp);
    return n;
}

static int count_bands(const std::string& dir)
{
    int n = 0;
    for (;;) {
        std::string p = dir + "/NarrBand" + std::to_string(n) + ".txt";
        FILE* fp = fopen(p.c_str(), "r");
        if (!fp) break;
        fclose(fp);
        ++n;
    }
    return n;
}

static bool read_floats_line(FILE* fp, std::vector<double>& v)
{
    v.clear();
    char buf[1 << 16];
    if (!fgets(buf, sizeof buf, fp)) return false;
    char* p = buf;
    for (;;) {
        char* e;
        double x = strtod(p, &e);
        if (e == p) break;
        v.push_back(x);
        p = e;
    }
    return true;
}

// code/read_tables.cpp:20-92 (grey==0, gas c-k) and :187-241 (grey==2,
// particle box model).  Sizes nB, nQ, nT are found from the files (the
// reference fixes them at compile time in code/NarrowBand.h).
int read_tables(const dnsmc_config& c, Tables& t, std::string* err)
{
    t = Tables();
    if (c.spectral == DNSMC_SPEC_GREY_T4) {
        t.nB = 1; t.nQ = 1; t.nT = 2;
        return DNSMC_OK;
    }
    if (c.spectral == DNSMC_SPEC_GREY_REF) {
        t.synth_grey = true;   // built at run time from Tmax
        return DNSMC_OK;
    }
    const bool gas = c.spectral == DNSMC_SPEC_GAS_CK;
    std::string dir = c.table_dir;
    std::string pm = dir + "/planck-mean.txt";
    FILE* fp = fopen(pm.c_str(), "r");
    if (!fp) { *err = "cannot open " + pm; return DNSMC_EIO; }
    int nT = count_lines(fp);
    std::vector<double> v;
    for (int i = 0; i < nT; ++i) {
        if (!read_floats_line(fp, v) || v.size() < 2) { fclose(fp); *err = "malformed " + pm; return DNSMC_EIO; }
        t.Tnb.push_back((float)v[0]);
        t.kP.push_back((float)v[1]);
    }
    fclose(fp);
    t.nT = nT;
    if (nT < 2) { *err = "planck-mean table needs >= 2 temperatures"; return DNSMC_EIO; }
    int nB = count_bands(dir);
    if (nB < 1) { *err = "no NarrBand*.txt in " + dir; return DNSMC_EIO; }
    if (nB > 1024) { *err = "at most 1024 bands are supported"; return DNSMC_EINVAL; }
    t.nB = nB;
    int nQ = 1;
    for (int b = 0; b < nB; ++b) {
        std::string p = dir + "/NarrBand" + std::to_string(b) + ".txt";
        fp = fopen(p.c_str(), "r");
        if (!fp) { *err = "cannot open " + p; return DNSMC_EIO; }
        if (!read_floats_line(fp, v) || v.size() < 3) { fclose(fp); *err = "malformed header in " + p; return DNSMC_EIO; }
        t.wvl.push_back((float)v[0]);
        t.wvr.push_back((float)v[1]);
        t.wvc.push_back((float)v[2]);
        if (!gas) {
            if (v.size() < 4) { fclose(fp); *err = "particle band file needs kappa in column 4: " + p; return DNSMC_EIO; }
            float k = (float)v[3];
            if (b == 0) t.kq.clear();
            for (int r = 0; r < nT; ++r) t.kq.push_back(k);
            t.kavg.push_back(k);
        } else {
            float kavg = 0.f;
            for (int r = 0; r < nT; ++r) {
                if (!read_floats_line(fp, v) || v.size() < 4) { fclose(fp); *err = "malformed row in " + p; return DNSMC_EIO; }
                int q = (int)v.size() - 3;
                if (b == 0 && r == 0) nQ = q;
                if (q != nQ) { fclose(fp); *err = "inconsistent quadrature size in " + p; return DNSMC_EIO; }
                kavg = (float)v[2];   // the last row's value is kept (code/read_tables.cpp:50)
                for (int g = 0; g < nQ; ++g) t.kq.push_back((float)v[3 + g]);
            }
            t.kavg.push_back(kavg);
        }
        fclose(fp);
    }
    if (nQ > 64) { *err = "at most 64 quadrature points are supported"; return DNSMC_EINVAL; }
    t.nQ = nQ;
    std::string pf = dir + (gas ? "/prob_new2.txt" : "/prob.txt");
    fp = fopen(pf.c_str(), "r");
    if (!fp) { *err = "cannot open " + pf; return DNSMC_EIO; }
    t.prob.assign((size_t)nT * nB, 0.f);
    t.probg.assign((size_t)nT * nB * nQ, 0.f);
    for (int i = 0; i < nT; ++i)
        for (int j = 0; j < nB; ++j) {
            if (!read_floats_line(fp, v) || v.size() < (size_t)(3 + (gas ? nQ : 0))) {
                fclose(fp);
                *err = "malformed or short " + pf;
                return DNSMC_EIO;
            }
            // each row names its temperature and band: a band file missing
            // from the NarrBand sequence shows here as a row mismatch
            if ((int)v[1] != j || fabs(v[0] - (double)t.Tnb[i]) > 1e-3 * (1.0 + fabs(v[0]))) {
                fclose(fp);
                *err = "row of " + pf + " does not match (T = " + std::to_string(t.Tnb[i]) + ", band " +
                       std::to_string(j) + "): NarrBand files and probability file disagree";
                return DNSMC_EIO;
            }
            t.prob[(size_t)i * nB + j] = (float)v[2];
            for (int g = 0; g < nQ; ++g)
                t.probg[((size_t)i * nB + j) * nQ + g] = gas ? (float)v[3 + g] : (float)g;   // :234 probg_h = g
        }
    // no rows beyond nT x nB (more bands in the probability file than band files)
    while (read_floats_line(fp, v))
        if (!v.empty()) {
            fclose(fp);
            *err = pf + " holds more rows than " + std::to_string(nT) + " temperatures x " + std::to_string(nB) +
                   " bands";
            return DNSMC_EIO;
        }
    fclose(fp);
    return DNSMC_OK;
}

// code/read_tables.cpp:104-183 (grey==1): nB=1000 bands up to 15000 1/cm,
// kappa = 1, temperature grid spanning [0.9*500, 1.1*Tmax).
void synth_grey_tables(Tables& t, float Tmax)
{
    const int nB = 1000, nT = 55;
    const float Tmins = 500.f, wvmax = 15000.f;
    t.nB = nB; t.nQ = 1; t.nT = nT;
    t.Tnb.resize(nT); t.kP.assign(nT, 1.0f);
    for (int i = 0; i < nT; ++i)
        t.Tnb[i] = Tmins * 0.9f + (Tmax * 1.1f - Tmins * 0.9f) * ((float)i) / ((float)nT);
    t.wvl.resize(nB); t.wvr.resize(nB); t.wvc.resize(nB);
    t.wvl[0] = 0.f;
    for (int i = 1; i < nB; ++i) t.wvl[i] = t.wvl[i - 1] + wvmax / nB;
    t.wvr[0] = wvmax / nB;
    for (int i = 1; i < nB; ++i) t.wvr[i] = t.wvr[i - 1] + wvmax / nB;
    for (int i = 0; i < nB; ++i) t.wvc[i] = (float)((t.wvl[i] + t.wvr[i]) / 2.);
    t.kq.assign((size_t)nB * nT, 1.0f);
    t.kavg.assign(nB, 1.0f);
    t.prob.assign((size_t)nT * nB, 0.f);
    t.probg.assign((size_t)nT * nB, 0.f);
    for (int i = 0; i < nT; ++i) {
        float* p = &t.prob[(size_t)i * nB];
        p[0] = planck_band(t.Tnb[i], t.wvc[0]) * (t.wvr[0] - t.wvl[0]) * 100;
        for (int j = 1; j < nB; ++j) p[j] = p[j - 1] + planck_band(t.Tnb[i], t.wvc[j]) * (t.wvr[j] - t.wvl[j]) * 100;
        float last = p[nB - 1];
        for (int j = 0; j < nB; ++j) p[j] = p[j] / last;
    }
}

static inline void lerp_tab(const float* row, float pos, int nT, float* v)
{
    pos = fminf(fmaxf(pos, 0.0f), (float)(nT - 1));
    int t0 = (int)pos;
    if (t0 > nT - 2) t0 = nT - 2;
    float fr = pos - (float)t0;
    *v = fmaf(fr, row[t0 + 1] - row[t0], row[t0]);
}

// Everything derived from the tables and Tmax (code/mc_gpu.cu:470-503,
// :541-561, :584-590; code/read_tables.cpp:34-36; code/memory_copy.cu:27-87).
int build_derived(const dnsmc_config& c, const Tables& t, float Tmax, Derived& d, std::string* err)
{
    d = Derived();
    d.Tmax = Tmax;
    const double sigma = 5.670373e-08;
    const int R = (c.mode == DNSMC_MODE_COMPAT) ? c.photmax : c.rays_per_cell;
    if (R < 1) { *err = "rays per cell must be positive"; return DNSMC_EINVAL; }
    if (c.spectral == DNSMC_SPEC_GREY_T4) {
        d.nB = 1; d.nQ = 1; d.nT = 2;
        d.kq.assign(2, 0.f);
        d.Ib.assign(2, 0.f);
        d.Ibw.resize(6);
        for (int a = 0; a < 3; ++a)
            for (int s = 0; s < 2; ++s) {
                float tw = c.Tw[a][s], t2 = tw * tw;
                d.Ibw[a * 2 + s] = t2 * t2;
            }
        d.cdf_band.assign(1, 1.0f);
        d.cdf_g.assign(1, 1.0f);
        float tm2 = Tmax * Tmax;
        d.inv_emax.assign(1, (float)(1.0 / ((double)c.kappa_grey * (double)(tm2 * tm2))));
        d.kappamax = c.kappa_grey;
        d.Ibmax_d = 4.0 * (double)c.kappa_grey * sigma * pow((double)Tmax, 4) / R;
        d.Ibmax = (float)d.Ibmax_d;
        d.g_search = 0;
        d.u_off = (c.mode == DNSMC_MODE_COMPAT) ? 1.0f : 0.5f;
        d.wq.assign(1, 1.0);
        return DNSMC_OK;
    }
    const int nB = t.nB, nQ = t.nQ, nT = t.nT;
    d.nB = nB; d.nQ = nQ; d.nT = nT;
    const float dT = t.Tnb[1] - t.Tnb[0];
    d.inv_dT = 1.0f / dT;
    d.mT0 = -t.Tnb[0] * d.inv_dT;
    if (!(Tmax >= t.Tnb[0]) || !(Tmax < t.Tnb[nT - 1])) {
        char b[200];
        snprintf(b, sizeof b, "Tmax = %g K outside the spectral table [%g, %g) K", Tmax, t.Tnb[0], t.Tnb[nT - 1]);
        *err = b;
        return DNSMC_ERANGE;
    }
    // kappa rows, row index (b*nQ + g), temperature fastest
    d.kq.assign((size_t)nB * nQ * nT, 0.f);
    for (int b = 0; b < nB; ++b)
        for (int g = 0; g < nQ; ++g)
            for (int tt = 0; tt < nT; ++tt)
                d.kq[((size_t)b * nQ + g) * nT + tt] = t.kq[((size_t)b * nT + tt) * nQ + g];
    // Planck table Ib[b][t] (code/mc_gpu.cu:541-551) and wall emission (black_copy)
    d.Ib.resize((size_t)nB * nT);
    for (int b = 0; b < nB; ++b)
        for (int tt = 0; tt < nT; ++tt) d.Ib[(size_t)b * nT + tt] = planck_band(t.Tnb[tt], t.wvc[b]);
    d.Ibw.resize((size_t)6 * nB);
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s)
            for (int b = 0; b < nB; ++b)
                d.Ibw[(size_t)(a * 2 + s) * nB + b] = c.Tw[a][s] > 0.f ? planck_band(c.Tw[a][s], t.wvc[b]) : 0.f;
    // kappamax: kP interpolated at Tmax (code/read_tables.cpp:34-36)
    {
        int tt = (int)((Tmax - t.Tnb[0]) / (t.Tnb[1] - t.Tnb[0]));
        d.kappamax = (t.kP[tt + 1] - t.kP[tt]) / (t.Tnb[tt + 1] - t.Tnb[tt]) * (Tmax - t.Tnb[tt]) + t.kP[tt];
        if (c.spectral == DNSMC_SPEC_GREY_REF) d.kappamax = 1.0f;
    }
    // per-bin emission at Tmax with the kernels' own interpolation
    float posm = fmaf(Tmax, d.inv_dT, d.mT0);
    std::vector<double> xg, wg;
    if (nQ > 1) gauss_legendre01(nQ, xg, wg); else wg.assign(1, 1.0);
    d.wq = wg;
    std::vector<double> e((size_t)nB * nQ);
    d.inv_emax.assign((size_t)nB * nQ, 0.f);
    double etot = 0.0;
    for (int b = 0; b < nB; ++b) {
        float ibm;
        lerp_tab(&d.Ib[(size_t)b * nT], posm, nT, &ibm);
        double dnu = ((double)t.wvr[b] - (double)t.wvl[b]) * 100.0;
        for (int g = 0; g < nQ; ++g) {
            std::vector<float> row(nT);
            for (int tt = 0; tt < nT; ++tt) row[tt] = t.kq[((size_t)b * nT + tt) * nQ + g];
            float km;
            lerp_tab(row.data(), posm, nT, &km);
            double em = (double)km * (double)ibm;
            d.inv_emax[(size_t)b * nQ + g] = em > 0.0 ? (float)(1.0 / em) : 0.f;
            e[(size_t)b * nQ + g] = 4.0 * M_PI * em * wg[g] * dnu;
            etot += e[(size_t)b * nQ + g];
        }
    }
    d.cdf_band.assign(nB, 0.f);
    d.cdf_g.assign((size_t)nB * nQ, 0.f);
    if (c.mode == DNSMC_MODE_COMPAT) {
        // proposal rows of the table node nearest Tmax (code/mc_gpu.cu:584-590, :837-858)
        int tm = (int)((Tmax - t.Tnb[0]) / (t.Tnb[1] - t.Tnb[0]));
        int tt = ((Tmax - t.Tnb[tm]) < (t.Tnb[tm + 1] - Tmax)) ? tm : tm + 1;
        for (int b = 0; b < nB; ++b) {
            d.cdf_band[b] = t.prob[(size_t)tt * nB + b];
            for (int g = 0; g < nQ; ++g) d.cdf_g[(size_t)b * nQ + g] = t.probg[((size_t)tt * nB + b) * nQ + g];
        }
        // Ibmax = 4 kappamax Tmax^4 sigma / photmax (code/mc_gpu.cu:63): the pow4
        // macro expands to a float chain, the sigma literal is double.
        float k4 = 4.0f * d.kappamax * Tmax * Tmax * Tmax * Tmax;
        d.Ibmax_d = (double)k4 * sigma / (double)R;
        d.Ibmax = (float)d.Ibmax_d;
        d.g_search = (c.spectral == DNSMC_SPEC_GREY_REF) ? 0 : 1;
        d.u_off = 1.0f;
        // sorted mode: gas tables only (code/NarrowBand.h:81-90 force srt 0 for grey 1 and 2);
        // band order by kavg ascending (code/mc_gpu.cu:941-957), ties keep the lower band first
        if (c.srt == 1 && c.spectral == DNSMC_SPEC_GAS_CK) {
            d.sorted = 1;
            d.band_order.resize(nB);
            for (int b = 0; b < nB; ++b) d.band_order[b] = b;
            std::stable_sort(d.band_order.begin(), d.band_order.end(),
                             [&](int x, int y) { return t.kavg[x] < t.kavg[y]; });
        }
    } else {
        if (!(etot > 0.0)) { *err = "spectral table emits nothing at Tmax"; return DNSMC_EINVAL; }
        double cum = 0.0;
        for (int b = 0; b < nB; ++b) {
            double eb = 0.0;
            for (int g = 0; g < nQ; ++g) eb += e[(size_t)b * nQ + g];
            double cg = 0.0;
            for (int g = 0; g < nQ; ++g) {
                cg += e[(size_t)b * nQ + g];
                d.cdf_g[(size_t)b * nQ + g] = eb > 0.0 ? (float)(cg / eb) : (float)(g + 1) / nQ;
            }
            d.cdf_g[(size_t)b * nQ + nQ - 1] = 1.0f;
            cum += eb;
            d.cdf_band[b] = (float)(cum / etot);
        }
        d.cdf_band[nB - 1] = 1.0f;
        d.Ibmax_d = etot / (double)R;
        d.Ibmax = (float)d.Ibmax_d;
        d.g_search = 1;
        d.u_off = 0.5f;
    }
    return DNSMC_OK;
}

// ------------------------------------------------------------- spline ----
// Natural cubic spline (code/spline.cpp:119-170, Numerical Recipes form)
// through the fine centres; the tridiagonal factors and the evaluation
// intervals depend only on the grids, so they are precomputed here.
void spline_axis_coeffs(const std::vector<float>& xf, int nf, const std::vector<float>& xc, int nc, SplineAxis& s)
{
    s = SplineAxis();
    s.nf = nf;
    s.nc = nc;
    std::vector<double> x(nf);
    for (int i = 0; i < nf; ++i) x[i] = xf[i + 1];
    s.sig.assign(nf, 0.0); s.invp.assign(nf, 0.0); s.y2f.assign(nf, 0.0);
    s.c6.assign(nf, 0.0); s.ihl.assign(nf, 0.0); s.ihh.assign(nf, 0.0);
    double y2prev = 0.0;
    for (int i = 1; i < nf - 1; ++i) {
        double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        double p = sig * y2prev + 2.0;
        double y2 = (sig - 1.0) / p;
        s.sig[i] = sig;
        s.invp[i] = 1.0 / p;
        s.y2f[i] = y2;
        s.c6[i] = 6.0 / (x[i + 1] - x[i - 1]);
        s.ihl[i] = 1.0 / (x[i] - x[i - 1]);
        s.ihh[i] = 1.0 / (x[i + 1] - x[i]);
        y2prev = y2;
    }
    s.klo.resize(nc); s.a.resize(nc); s.b.resize(nc); s.A.resize(nc); s.B.resize(nc);
    for (int q = 0; q < nc; ++q) {
        double xv = xc[q + 1];
        int klo = 0, khi = nf - 1;
        while (khi - klo > 1) {
            int k = (khi + klo) >> 1;
            if (x[k] > xv) khi = k; else klo = k;
        }
        double h = x[khi] - x[klo];
        double a = (x[khi] - xv) / h, b = (xv - x[klo]) / h;
        s.klo[q] = klo;
        s.a[q] = a;
        s.b[q] = b;
        s.A[q] = (a * a * a - a) * (h * h) / 6.0;
        s.B[q] = (b * b * b - b) * (h * h) / 6.0;
    }
}

// --------------------------------------------------------------- XORWOW ----
static uint64_t splitmix64(uint64_t* st)
{
    uint64_t z = (*st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Own seed scramble (SplitMix64); this is not cuRAND's seeding.
void xorwow_seed_state(uint64_t seed, uint32_t st[6])
{
    uint64_t s = seed;
    uint64_t a = splitmix64(&s), b = splitmix64(&s), c = splitmix64(&s);
    st[0] = (uint32_t)a; st[1] = (uint32_t)(a >> 32);
    st[2] = (uint32_t)b; st[3] = (uint32_t)(b >> 32);
    st[4] = (uint32_t)c; st[5] = (uint32_t)(c >> 32);
    if ((st[0] | st[1] | st[2] | st[3] | st[4]) == 0) st[0] = 1u;
}

namespace {
struct M160 { uint32_t c[160][5]; };
void m_apply(const M160& A, const uint32_t in[5], uint32_t out[5])
{
    uint32_t r[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < 160; ++b)
        if ((in[b >> 5] >> (b & 31)) & 1u)
            for (int w = 0; w < 5; ++w) r[w] ^= A.c[b][w];
    for (int w = 0; w < 5; ++w) out[w] = r[w];
}
}  // namespace

// Linear part of one XORWOW step as a GF(2) 160x160 matrix; squared 67+j
// times gives the jump by 2^(67+j) steps used for subsequences.
void xorwow_jump_tables(std::vector<uint32_t>& out)
{
    M160* m = new M160;
    for (int b = 0; b < 160; ++b) {
        Xorwow s{0, 0, 0, 0, 0, 0};
        uint32_t* x = &s.x0;
        x[b >> 5] = 1u << (b & 31);
        (void)xorwow_next(s);
        m->c[b][0] = s.x0; m->c[b][1] = s.x1; m->c[b][2] = s.x2; m->c[b][3] = s.x3; m->c[b][4] = s.x4;
    }
    M160* t = new M160;
    out.assign((size_t)60 * 160 * 5, 0u);
    for (int k = 1; k < 67 + 60; ++k) {
        for (int b = 0; b < 160; ++b) m_apply(*m, m->c[b], t->c[b]);
        std::swap(m, t);
        if (k >= 67)
            for (int b = 0; b < 160; ++b)
                for (int w = 0; w < 5; ++w) out[((size_t)(k - 67) * 160 + b) * 5 + w] = m->c[b][w];
    }
    delete m;
    delete t;
}

}  // namespace dn
