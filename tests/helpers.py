"""Test helpers: build the oracle's problem from a solver's derived arrays, and
brute-force references."""
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def oracle_problem_dict(S, cfg, snap=0):
    """Everything the oracle MC needs, read back from the solver (grids, fields
    per level, tables, proposal, scalars)."""
    nlev = cfg.grid_num
    levels = []
    for l in range(nlev):
        faces = [S.array("faces%d" % a, l) for a in range(3)]
        n = [len(f) - 1 for f in faces]
        nc = n[0] * n[1] * n[2]
        T = S.array("T", l)[snap * nc:(snap + 1) * nc]
        lev = dict(n=n, faces=faces, T=T, maxs=cfg.maxs[l])
        levels.append(lev)
    levels[0]["centres"] = [S.array("centres%d" % a, 0) for a in range(3)]
    sc = S.array("scalars")
    kq1 = S.array("kq")
    return dict(
        levels=levels,
        bc=[[cfg.bc[a][s] for s in range(2)] for a in range(3)],
        polar_axis=cfg.polar_axis,
        tau_cut=float(sc[10]),
        spectral=1 if cfg.spectral == 3 else 0,
        nB=int(sc[3]), nQ=int(sc[4]), nT=int(sc[5]),
        inv_dT=float(sc[6]), mT0=float(sc[7]),
        kq=kq1.copy(),
        Ib=S.array("Ib"), Ibw=S.array("Ibw"),
        kappa_grey=float(cfg.kappa_grey),
        cdf_band=S.array("cdf_band"), cdf_g=S.array("cdf_g"), inv_emax=S.array("inv_emax"),
        Ibmax=float(sc[2]), Ibmax_d=S.Ibmax_d(), Tmax=float(sc[0]),
        g_search=int(sc[8]), u_off=float(sc[9]),
    )


def exact_cells(origin, s, faces, bc, max_steps=64):
    """FP64 exact cell sequence of a ray through a rectilinear grid (periodic
    or absorbing walls): list of 0-based (i, j, k) visited, from the start cell."""
    n = [len(f) - 1 for f in faces]
    L = [f[-1] for f in faces]
    p = np.array(origin, float)
    cell = [int(np.searchsorted(faces[a], p[a], side="right") - 1) for a in range(3)]
    out = [tuple(cell)]
    for _ in range(max_steps):
        tn = []
        for a in range(3):
            if s[a] > 0:
                tn.append((faces[a][cell[a] + 1] - p[a]) / s[a])
            else:
                tn.append((faces[a][cell[a]] - p[a]) / s[a])
        a = int(np.argmin(tn))
        t = tn[a]
        p = p + t * np.asarray(s)
        step = 1 if s[a] > 0 else -1
        cell[a] += step
        if cell[a] < 0 or cell[a] >= n[a]:
            side = 1 if cell[a] >= n[a] else 0
            if bc[a][side] == 2:
                return out, "wall"
            cell[a] = 0 if side else n[a] - 1
            p[a] = 0.0 if side else L[a]
        out.append(tuple(cell))
    return out, "cap"


def ref_subsequence(i, j, k, v, n, num_streams=16, num_gpu=1, gpu=0, p_row=1, ystart=1):
    """Subsequence of cell (i,j,k) (0-based) and block v in the reference's
    seeding curand_init(1234, (idx+1)(stream+1)(v+1)(1+gpu)ystart, 0)
    (code/mc_gpu.cu:84-87, :99), int32 product sign-extended to 64 bits."""
    n0, n1, n2 = n
    per = n0 // num_streams
    stream = i // per
    jloc = j - (ystart - 1)
    kloc = k - gpu * (n2 // num_gpu)
    idx = (i - stream * per) * ((n2 // num_gpu) * (n1 // p_row)) + jloc * (n2 // num_gpu) + kloc
    prod = ((idx + 1) * (stream + 1) * (v + 1) * (1 + gpu) * ystart) & 0xFFFFFFFF
    if prod >= 1 << 31:
        prod -= 1 << 32
    return prod & 0xFFFFFFFFFFFFFFFF


def read_reference_tables(dirname, gas=True):
    """Independent numpy parser of the reference table formats
    (code/read_tables.cpp:20-92 gas, :187-241 particle)."""
    pm = np.loadtxt(os.path.join(dirname, "planck-mean.txt"), ndmin=2)
    Tnb, kP = pm[:, 0], pm[:, 1]
    nT = len(Tnb)
    nB = 0
    while os.path.exists(os.path.join(dirname, "NarrBand%d.txt" % nB)):
        nB += 1
    wvl, wvr, wvc, kq, kavg = [], [], [], [], []
    for b in range(nB):
        with open(os.path.join(dirname, "NarrBand%d.txt" % b)) as f:
            head = [float(x) for x in f.readline().split()]
            wvl.append(head[0]); wvr.append(head[1]); wvc.append(head[2])
            if gas:
                rows = np.array([[float(x) for x in f.readline().split()] for _ in range(nT)])
                kq.append(rows[:, 3:])
                kavg.append(rows[-1, 2])          # the last row's value is kept (code/read_tables.cpp:50)
            else:
                kq.append(np.full((nT, 1), head[3]))
                kavg.append(head[3])
    kq = np.array(kq)                       # [b][t][g]
    nQ = kq.shape[2]
    p = np.loadtxt(os.path.join(dirname, "prob_new2.txt" if gas else "prob.txt"), ndmin=2)
    prob = p[:, 2].reshape(nT, nB)
    if gas:
        probg = p[:, 3:3 + nQ].reshape(nT, nB, nQ)
    else:
        probg = np.zeros((nT, nB, 1))
    return {k: np.asarray(v, np.float32) for k, v in
            dict(Tnb=Tnb, kP=kP, wvl=wvl, wvr=wvr, wvc=wvc, kq=kq, prob=prob, probg=probg, kavg=kavg).items()}


def read_resfortm():
    return np.loadtxt(os.path.join(GOLDEN, "resfortm.txt"))


def particle_table_dir():
    return os.path.join(GOLDEN, "ref_particle_tables")


def z_scores(Q, Qref, sd, sd_ref=None):
    s2 = sd ** 2 + (0 if sd_ref is None else sd_ref ** 2)
    return (Q - Qref) / np.sqrt(np.maximum(s2, 1e-300))


def planck_mean_closed(T):
    return 5.670373e-08 * T ** 4 / math.pi


def legacy_conf_text():
    """DNSMC_CONFIG of the tests that drive the legacy entry points.  Those use
    one process-wide solver, created from this text at the first call, so
    every such test shares it."""
    return ("n = 16, 16, 16\nphotmax = 64\ngrid_num = 2\nmaxs = 5, 100000\ntally = 1\n"
            "table_dir = %s\n" % particle_table_dir())
