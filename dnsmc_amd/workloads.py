"""Synthetic inputs of the channel workloads: configurations, correlated-k
tables in the reference text formats, and seeded temperature snapshots.

Workloads (wall-normal axis y, periodic x/z, black walls; tanh-stretched y):
  grey32     32x32x32, Lx=2pi Ly=2 Lz=pi, grey kappa=1, T=1000 K + N(0,5 K), Tw=300 K, 100 rays/cell
  ck64       64x96x64, parabolic T(y) 573->955->573 K + N(0,10 K), 100x16x69 c-k table, 500 rays/cell
  ck256      256x128x128, profile + Fourier modes (|k|<=8, 50 K), 200 rays/cell
  ck384x8    384x256x256, 8 snapshots (profile + Fourier modes) in one launch
Seeds: grey32 field 1; table 11; ck64 field 12; ck256 21; learned snapshots
31-50 (train) and 51-55 (held out); ck384x8 snapshots 41-48.
"""
import math
import os

import numpy as np

from .solver import config_from_dict

SIGMA = 5.670373e-08
TWO_PI = 2.0 * math.pi


def _rng(seed):
    return np.random.Generator(np.random.Philox(key=int(seed)))


def channel_config(n, rays_per_cell=100, spectral=0, Tw=573.0, table_dir="", nsnap=1, **kw):
    d = dict(
        n=n, L=(TWO_PI, 2.0, math.pi), stretch=(0, 2, 0), stretch_beta=(1.5, 1.5, 1.5),
        bc=((1, 1), (2, 2), (1, 1)), Tw=((0.0, 0.0), (Tw, Tw), (0.0, 0.0)), polar_axis=1,
        nvar=3, photmax=20000, rays_per_cell=rays_per_cell, grid_num=1, maxs=(100000, 100000, 100000, 100000, 100000, 100000),
        toll=1.0e-3, spectral=spectral, kappa_grey=1.0, mode=0, sobol=1, num_streams=16, num_gpu=1, gpu=0,
        p_row=1, seed=20240601, nsnap=nsnap, precision=0, tally=0, srt=0, batch_rays=1 << 31,
        table_dir=table_dir,
    )
    d.update(kw)
    return config_from_dict(d)


def y_centres(ny, Ly=2.0, beta=1.5):
    j = np.arange(ny + 1)
    f = 0.5 * Ly * (1.0 + np.tanh(beta * (2.0 * j / ny - 1.0)) / np.tanh(beta))
    f[0], f[-1] = 0.0, Ly
    f = f.astype(np.float32)
    return f, 0.5 * (f[1:] + f[:-1])


def profile(y, Tc=955.0, Tw=573.0):
    return Tw + (Tc - Tw) * (1.0 - (y - 1.0) ** 2)


def ghosted(Tint):
    """Interior (n0,n1,n2) -> ghost-inclusive Fortran array (n0+2,n1+2,n2+2)."""
    n0, n1, n2 = Tint.shape
    T = np.empty((n0 + 2, n1 + 2, n2 + 2), np.float32, order="F")
    T[1:-1, 1:-1, 1:-1] = Tint
    T[0, :, :] = T[1, :, :] if n0 else 0
    T[-1, :, :] = T[-2, :, :]
    T[:, 0, :] = T[:, 1, :]
    T[:, -1, :] = T[:, -2, :]
    T[:, :, 0] = T[:, :, 1]
    T[:, :, -1] = T[:, :, -2]
    return T


def field_grey32(seed=1, n=(32, 32, 32), T0=1000.0, sd=5.0):
    g = _rng(seed)
    return (T0 + sd * g.standard_normal(n)).astype(np.float32)


def field_profile_noise(seed=12, n=(64, 96, 64), sd=10.0):
    _, yc = y_centres(n[1])
    g = _rng(seed)
    T = profile(yc.astype(np.float64))[None, :, None] + sd * g.standard_normal(n)
    return T.astype(np.float32)


def field_fourier(seed, n, amp=50.0, kmax=8):
    """Profile + sum of modes sin(pi ky y/Ly) exp(i 2pi (kx x/Lx + kz z/Lz)),
    kx^2+ky^2+kz^2 <= kmax^2, ky >= 1, random complex amplitudes; scaled so
    that max |dT| = amp."""
    n0, n1, n2 = n
    _, yc = y_centres(n1)
    g = _rng(seed)
    dT = np.zeros(n, np.float64)
    for ky in range(1, kmax + 1):
        spec = np.zeros((n0, n2), np.complex128)
        for kx in range(-kmax, kmax + 1):
            for kz in range(-kmax, kmax + 1):
                if kx * kx + ky * ky + kz * kz <= kmax * kmax:
                    spec[kx % n0, kz % n2] = g.standard_normal() + 1j * g.standard_normal()
        F = np.real(np.fft.ifft2(spec)) * (n0 * n2)
        dT += np.sin(math.pi * ky * yc.astype(np.float64) / 2.0)[None, :, None] * F[:, None, :]
    dT *= amp / np.abs(dT).max()
    T = profile(yc.astype(np.float64))[None, :, None] + dT
    return T.astype(np.float32)


# ------------------------------------------------------------ c-k table ----
def planck_band64(T, nu):
    """Spectral radiance per unit wavenumber (code/memory_copy.cu:291-300), FP64."""
    C1 = 3.741771790075259e-16
    C2 = 0.014387741858429
    nu_m = nu * 100.0
    return C1 / math.pi * nu_m ** 3 / np.expm1(C2 * nu_m / T)


def gauss_legendre01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def synth_ck_table(seed=11, nB=100, nQ=16, T0=300.0, T1=2000.0, dT=25.0, nu0=150.0, nu1=7500.0, sigma_g=2.0):
    """Log-normal, g-monotone correlated-k table: kappa(T,b,g) =
    exp(mu_b(T) + sigma_g Phi^-1(gq_g)), mu_b(T) = m_b + c_b ln(T/1000),
    m_b ~ U(ln 0.01, ln 10), c_b ~ U(-1, 0.5).  Band CDF Planck-weighted."""
    from scipy.special import ndtri
    g = _rng(seed)
    Tnb = np.arange(T0, T1 + 0.5 * dT, dT)
    nT = len(Tnb)
    dnu = (nu1 - nu0) / nB
    wvl = nu0 + dnu * np.arange(nB)
    wvr = wvl + dnu
    wvc = 0.5 * (wvl + wvr)
    gq, wq = gauss_legendre01(nQ)
    m = g.uniform(math.log(0.01), math.log(10.0), nB)
    c = g.uniform(-1.0, 0.5, nB)
    mu = m[:, None] + c[:, None] * np.log(Tnb[None, :] / 1000.0)          # [b][t]
    kq = np.exp(mu[:, :, None] + sigma_g * ndtri(gq)[None, None, :])       # [b][t][g]
    Ib = planck_band64(Tnb[None, :], wvc[:, None])                          # [b][t]
    e = kq * wq[None, None, :] * (Ib * dnu * 100.0)[:, :, None]             # [b][t][g]
    kP = math.pi * e.sum(axis=(0, 2)) / (SIGMA * Tnb ** 4)
    eb = e.sum(axis=2)                                                      # [b][t]
    prob = np.cumsum(eb, axis=0) / eb.sum(axis=0)[None, :]                  # [b][t]
    probg = np.cumsum(kq * wq[None, None, :], axis=2)
    probg /= probg[:, :, -1:]
    kavg = (kq * wq[None, None, :]).sum(axis=2)                             # [b][t]
    return dict(Tnb=Tnb, kP=kP, wvl=wvl, wvr=wvr, wvc=wvc, kq=kq, prob=prob.T.copy(),
                probg=probg.transpose(1, 0, 2).copy(), kavg=kavg, gq=gq, wq=wq)


def write_reference_tables(t, dirname):
    """Write a table dict in the formats of distr/write_tables.cpp:23-107:
    planck-mean.txt, NarrBand%d.txt, prob.txt, prob_new2.txt."""
    os.makedirs(dirname, exist_ok=True)
    Tnb, kq, prob, probg = t["Tnb"], t["kq"], t["prob"], t["probg"]
    nB, nT, nQ = kq.shape
    with open(os.path.join(dirname, "planck-mean.txt"), "w") as f:
        for i in range(nT):
            f.write("%f %e %e %e\n" % (Tnb[i], t["kP"][i], t["kP"][i], t["kP"][i]))
    with open(os.path.join(dirname, "prob.txt"), "w") as f:
        for i in range(nT):
            for j in range(nB):
                f.write("%f \t %-5d \t%30.20e \n" % (Tnb[i], j, prob[i, j]))
    with open(os.path.join(dirname, "prob_new2.txt"), "w") as f:
        for i in range(nT):
            for j in range(nB):
                f.write("%f \t %-5d \t%30.20e \t" % (Tnb[i], j, prob[i, j]))
                f.write("".join("%30.20e\t" % probg[i, j, q] for q in range(nQ)))
                f.write("\n")
    for j in range(nB):
        with open(os.path.join(dirname, "NarrBand%d.txt" % j), "w") as f:
            f.write("%e\t%e\t%e \n" % (t["wvl"][j], t["wvr"][j], t["wvc"][j]))
            for i in range(nT):
                f.write(" %f\t %e\t %e\t" % (Tnb[i], prob[i, j], t["kavg"][j, i]))
                f.write("".join(" %e\t" % kq[j, i, q] for q in range(nQ)))
                f.write("\n")
    return dirname


def table_arrays(t):
    """Dict accepted by Solver.set_tables (float32 arrays in reference layouts);
    kavg is the last temperature row's value, as NarrBand%d.txt holds it."""
    d = {k: np.asarray(t[k], np.float32) for k in ("Tnb", "kP", "wvl", "wvr", "wvc", "kq", "prob", "probg")}
    d["kavg"] = np.asarray(t["kavg"][:, -1], np.float32)
    return d


WORKLOADS = {
    "grey32": dict(n=(32, 32, 32), rays=100, spectral=3, Tw=300.0),
    "ck64": dict(n=(64, 96, 64), rays=500, spectral=0, Tw=573.0),
    "ck256": dict(n=(256, 128, 128), rays=200, spectral=0, Tw=573.0),
    "ck384x8": dict(n=(384, 256, 256), rays=100, spectral=0, Tw=573.0, nsnap=8),
}


def workload_fields(name, nsnap=None):
    w = WORKLOADS[name]
    if name == "grey32":
        return [field_grey32(1, w["n"])]
    if name == "ck64":
        return [field_profile_noise(12, w["n"])]
    if name == "ck256":
        return [field_fourier(21, w["n"])]
    if name == "ck384x8":
        k = nsnap or w["nsnap"]
        return [field_fourier(41 + s, w["n"]) for s in range(k)]
    raise KeyError(name)
