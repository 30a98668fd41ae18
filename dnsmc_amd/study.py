"""Accuracy studies of the radiation solve on the channel workloads.

Reference values come from the library itself: the FP64 traversal
(precision 1: FP64 directions, lookups, transport, weights and tallies) run
with many rays on a seeded subset of emitting cells.  It is the same
estimator as the CPU FP64 oracle, to 1e-12 on the same rays (the GPU parity
tests check exactly that), on the GPU's throughput.

  python -m dnsmc_amd.study errors       RMS / max error of Q_rad per workload
  python -m dnsmc_amd.study precision    FP32 vs FP64 traversal x FP64 vs
                                         Kahan-FP32 tallies, 256x128x128
  python -m dnsmc_amd.study heldout      learned proposal on held-out 64x96x64
                                         snapshots: variance ratio, errors
  python -m dnsmc_amd.study convergence  error vs rays per cell (R^-1/2)
  python -m dnsmc_amd.study legs         one 384x256x256 snapshot: plain, plain
                                         coarsened, learned (for counter runs)

Errors over the chosen cells: eps_rms = sqrt(sum (Q - Q_ref)^2 - sum var_ref)
/ sqrt(sum Q_ref^2) (the reference's own variance removed), eps_max =
max |Q - Q_ref| / max |Q_ref|.
"""
import json
import math
import os
import sys
import time

import numpy as np

from . import workloads
from .solver import Solver


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_solver(workload, rays, nsnap=1, mode=0, precision=0, tally=0, seed=None, cells=None, mlp=None,
                mix_eps=0.1, batch=None, fields=None, **cfg_kw):
    """Solver of a named channel workload with its table and seeded snapshots
    (cfg_kw: further configuration fields, e.g. grid_num and maxs)."""
    w = workloads.WORKLOADS[workload]
    kw = dict(cfg_kw, mode=mode, precision=precision, tally=tally)
    if seed is not None:
        kw["seed"] = seed
    if batch:
        kw["batch_rays"] = int(batch)
    cfg = workloads.channel_config(w["n"], rays_per_cell=rays, spectral=w["spectral"], Tw=w["Tw"], nsnap=nsnap,
                                   **kw)
    S = Solver(cfg)
    if w["spectral"] == 3:
        S.load_tables()
    else:
        S.set_tables(workloads.table_arrays(workloads.synth_ck_table(11)))
    if mlp is not None:
        S.set_mlp(*mlp, mix_eps=mix_eps)
    if cells is not None:
        S.set_emitters(cells)
    if fields is None:
        fields = workloads.workload_fields(workload, nsnap)
    for s in range(nsnap):
        S.set_field(s, workloads.ghosted(fields[s]))
    return S, cfg


def sample_cells(workload, k, key=11):
    n = workloads.WORKLOADS[workload]["n"]
    ncell = n[0] * n[1] * n[2]
    rng = np.random.Generator(np.random.Philox(key=key))
    return np.sort(rng.choice(ncell, size=min(k, ncell), replace=False)).astype(np.int32)


def reference(workload, cells, rays_ref, seed=987654321, fields=None, **cfg_kw):
    """Converged FP64 reference MC of snapshot 0 on `cells`: (Q, var(Q), seconds)."""
    t0 = time.time()
    S, _ = make_solver(workload, rays_ref, precision=1, tally=1, seed=seed, cells=cells,
                       fields=None if fields is None else fields[:1], **cfg_kw)
    S.run()
    r = S.results(0)
    S.close()
    return r["Q"], r["sdev"] ** 2, time.time() - t0


def errors(Q, Qref, var_ref=None):
    d2 = float(np.sum((Q - Qref) ** 2))
    if var_ref is not None:
        d2 = max(d2 - float(np.sum(var_ref)), 0.0)
    return dict(rms=math.sqrt(d2 / float(np.sum(Qref ** 2))),
                max=float(np.max(np.abs(Q - Qref)) / np.max(np.abs(Qref))))


def rays_to_target(workload, cells, Qref, vref, fields, mode=0, mlp=None, mix_eps=0.1, target=0.01, R0=100,
                   Rmax=1 << 23, seed=None, **cfg_kw):
    """Rays per cell to reach `target` RMS error against (Qref, vref), measured
    by doubling: the estimator is run on the emitters `cells` of snapshot 0
    (their errors are those of the full-grid solve: every cell's estimate is
    independent of which other cells emit) at R0, 2 R0, ... until the error
    is at most `target` or R exceeds Rmax.  Every point's error has the
    reference's own variance removed.  Returns the points, the first R that
    reached the target with its error, the R at the target interpolated
    between the two points that bracket it, and the least-squares fit of
    eps^2 = b^2 + c / R over the points (b: the estimator's bias against the
    reference, e.g. mesh coarsening; R_fit = c / (target^2 - b^2))."""
    pts = []
    R = R0
    first = None
    while R <= Rmax:
        S, _ = make_solver(workload, R, mode=mode, mlp=mlp, mix_eps=mix_eps, cells=cells, fields=fields[:1],
                           seed=seed, **cfg_kw)
        t0 = time.time()
        S.run()
        r = S.results(0)
        dt = time.time() - t0
        S.close()
        e = errors(r["Q"], Qref, vref)
        pts.append(dict(rays_per_cell=R, rms=e["rms"], max=e["max"], seconds=round(dt, 3)))
        log("rays_to_target", workload, "mode", mode, cfg_kw.get("grid_num", 1), pts[-1])
        if e["rms"] <= target:
            first = pts[-1]
            break
        R *= 2
    # eps^2 = b^2 + c/R, least squares over the points with the noise-dominated
    # ones weighted alike (relative residuals)
    Rs = np.array([p["rays_per_cell"] for p in pts], float)
    e2 = np.array([p["rms"] ** 2 for p in pts], float)
    A = np.stack([np.ones_like(Rs), 1.0 / Rs], axis=1) / e2[:, None]
    coef, *_ = np.linalg.lstsq(A, np.ones_like(Rs), rcond=None)
    b2, c = float(max(coef[0], 0.0)), float(max(coef[1], 0.0))
    R_fit = c / (target ** 2 - b2) if b2 < target ** 2 and c > 0 else None
    # R at the target between the two measured points that bracket it
    # (log-log interpolation: a local power law between neighbours)
    R_int = None
    if first is not None and len(pts) >= 2:
        (Ra, ea), (Rb, eb) = [(p["rays_per_cell"], p["rms"]) for p in pts[-2:]]
        if ea > target >= eb and ea > eb:
            R_int = Ra * (Rb / Ra) ** (math.log(ea / target) / math.log(ea / eb))
    elif first is not None:
        R_int = float(first["rays_per_cell"])
    return dict(points=pts, target=target, reached=first is not None,
                rays_per_cell=first["rays_per_cell"] if first else None, rms_at_R=first["rms"] if first else None,
                rays_per_cell_interp=R_int, fit=dict(bias_rms=math.sqrt(b2), c=c, rays_per_cell=R_fit))


def study_errors(workloads_=("grey32", "ck64", "ck256"), ncells=1024, rays_ref=65536):
    out = {}
    for wl in workloads_:
        w = workloads.WORKLOADS[wl]
        cells = sample_cells(wl, ncells)
        Qref, vref, tref = reference(wl, cells, rays_ref)
        S, cfg = make_solver(wl, w["rays"])
        t0 = time.time()
        S.run()
        st = S.stats()
        dt = time.time() - t0
        r = S.results(0)
        S.close()
        e = errors(r["Q"][cells], Qref, vref)
        e_est = math.sqrt(float(np.sum(r["sdev"][cells] ** 2)) / float(np.sum(Qref ** 2)))
        out[wl] = dict(grid=list(cfg.n), rays_per_cell=w["rays"], rms=e["rms"], max=e["max"],
                       rms_from_own_variance=e_est, solve_s=dt, rays_per_s=st["n_rays"] / dt, cells=len(cells),
                       reference=dict(rays_per_cell=rays_ref, seconds=round(tref, 2)))
        log(wl, out[wl])
    return out


def study_precision(ncells=1024, rays_ref=65536):
    """256x128x128 (profile + Fourier modes), 200 rays/cell, the four
    arithmetic variants on the same Philox rays.  'same_rays' compares with
    the FP64 traversal + FP64 tallies on the same rays (the arithmetic error
    alone); 'converged' with the converged FP64 reference (arithmetic plus
    Monte Carlo error)."""
    wl = "ck256"
    R = workloads.WORKLOADS[wl]["rays"]
    cells = sample_cells(wl, ncells)
    fields = workloads.workload_fields(wl, 1)
    Qref, vref, tref = reference(wl, cells, rays_ref, fields=fields)
    runs = {}
    for prec in (0, 1):
        for tally, tname in ((1, "fp64"), (2, "kahan_fp32")):
            S, _ = make_solver(wl, R, precision=prec, tally=tally, fields=fields)
            t0 = time.time()
            S.run()
            S.wait()
            dt = time.time() - t0
            r = S.results(0)
            S.close()
            runs[("fp32" if prec == 0 else "fp64", tname)] = (r, dt)
    base = runs[("fp64", "fp64")][0]
    out = {}
    for (trav, tname), (r, dt) in runs.items():
        Qs = base["Q"]
        same = errors(r["Q"], Qs)
        same_E = float(np.max(np.abs(r["E"] - base["E"]) / np.abs(base["E"])))
        conv = errors(r["Q"][cells], Qref, vref)
        out["%s_traversal_%s_tally" % (trav, tname)] = dict(
            same_rays_rms=same["rms"], same_rays_max=same["max"], same_rays_max_rel_E=same_E,
            converged_rms=conv["rms"], converged_max=conv["max"], solve_s=round(dt, 3))
        log(trav, tname, out["%s_traversal_%s_tally" % (trav, tname)])
    return dict(workload=wl, rays_per_cell=R, cells_for_converged=len(cells),
                reference=dict(rays_per_cell=rays_ref, seconds=round(tref, 2)), variants=out)


def study_convergence(workload="ck384x8", rays=(100, 200, 400, 800), ncells=1024, rays_ref=65536, mlp=None,
                      mix_eps=0.05):
    """Error against the converged reference as rays per cell double
    (snapshot 0 only); the fitted exponent of eps ~ R^p should be -1/2."""
    cells = sample_cells(workload, ncells)
    fields = workloads.workload_fields(workload, 1)
    Qref, vref, _ = reference(workload, cells, rays_ref, fields=fields)
    res = {}
    legs = [("plain", 0, None)] + ([("learned", 2, mlp)] if mlp is not None else [])
    for name, mode, W in legs:
        pts = []
        for R in rays:
            S, _ = make_solver(workload, R, mode=mode, mlp=W, mix_eps=mix_eps, fields=fields)
            S.run()
            S.wait()
            t0 = time.time()
            S.run()
            S.wait()
            dt = time.time() - t0
            r = S.results(0)
            S.close()
            e = errors(r["Q"][cells], Qref, vref)
            pts.append(dict(rays_per_cell=R, rms=e["rms"], max=e["max"], seconds=dt))
            log(name, pts[-1])
        lr = np.log([p["rays_per_cell"] for p in pts])
        le = np.log([p["rms"] for p in pts])
        slope = float(np.polyfit(lr, le, 1)[0])
        # seconds to 1% RMS from the fitted power law (time proportional to rays)
        c = np.polyfit(lr, le, 1)
        R1 = math.exp((math.log(0.01) - c[1]) / c[0])
        sec_per_ray = np.mean([p["seconds"] / p["rays_per_cell"] for p in pts])
        res[name] = dict(points=pts, exponent=slope, rays_per_cell_to_1pct=R1, seconds_to_1pct=R1 * sec_per_ray)
    if "learned" in res:
        res["speedup_learned"] = res["plain"]["seconds_to_1pct"] / res["learned"]["seconds_to_1pct"]
    return dict(workload=workload, snapshot=0, cells=len(cells), reference_rays_per_cell=rays_ref, legs=res)


def study_heldout(rays=(100, 1000), snaps=range(51, 56), mix_eps=0.05, **cfg_kw):
    """The learned proposal on held-out channel snapshots (64x96x64, parabolic
    profile + N(0, 10 K), seeds 51-55; trained on seeds 31-50): variance
    ratio vs the plain estimator at equal rays per cell (sums of the per-cell
    variances), and both estimators' RMS error against a converged FP64
    reference on 1024 cells of each snapshot."""
    from .learned import train_channel_proposal
    W, info = train_channel_proposal(**cfg_kw)
    wl = "ck64"
    cells = sample_cells(wl, 1024)
    out = dict(train=info, mix_eps=mix_eps, snapshots={})
    for seed in snaps:
        fields = [workloads.field_profile_noise(seed, workloads.WORKLOADS[wl]["n"])]
        Qref, vref, _ = reference(wl, cells, 65536, fields=fields, **cfg_kw)
        row = {}
        for R in rays:
            res = {}
            for name, mode, Wm in (("plain", 0, None), ("learned", 2, W)):
                S, _ = make_solver(wl, R, mode=mode, mlp=Wm, mix_eps=mix_eps, fields=fields, **cfg_kw)
                S.run()
                res[name] = S.results(0)
                S.close()
            vr = float(np.sum(res["plain"]["sdev"] ** 2) / np.sum(res["learned"]["sdev"] ** 2))
            row[str(R)] = dict(variance_ratio=vr,
                               rms_plain=errors(res["plain"]["Q"][cells], Qref, vref)["rms"],
                               rms_learned=errors(res["learned"]["Q"][cells], Qref, vref)["rms"])
            log("snapshot", seed, "R", R, row[str(R)])
        out["snapshots"][str(seed)] = row
    return out


def study_legs(workload="ck384x8", rays=100, mix_eps=0.1):
    """One snapshot of the workload solved five ways (plain on the fine grid,
    plain with the reference's default coarsening, learned on the fine grid
    with the proposal evaluated by the generation kernels and with the draw
    inside the traversal (DNSMC_LEARNED_FUSED=1), learned with coarsening),
    each after one warm-up solve: seconds, segments and rays of each solve
    (the kernel-trace and counter profiles of the traversal kernels are taken
    around this command)."""
    from .learned import train_channel_proposal
    fields = workloads.workload_fields(workload, 1)
    amc = dict(grid_num=5, maxs=(5, 5, 5, 5, 100000, 100000))
    W, _ = train_channel_proposal()
    Wc, _ = train_channel_proposal(**amc)
    out = {}
    for name, mode, kw, w, fused in (("plain_fine", 0, {}, None, None), ("plain_coarsened", 0, amc, None, None),
                                     ("learned_fine", 2, {}, W, None), ("learned_fine_in_traversal", 2, {}, W, "1"),
                                     ("learned_coarsened", 2, amc, Wc, None)):
        if fused:
            os.environ["DNSMC_LEARNED_FUSED"] = fused
        try:
            S, _ = make_solver(workload, rays, mode=mode, mlp=w, mix_eps=mix_eps, fields=fields, **kw)
            S.run()
            S.wait()
            t0 = time.time()
            S.run()
            S.wait()
            st = S.stats()
            out[name] = dict(seconds=time.time() - t0, segments=st["n_steps"], rays=st["n_rays"],
                             trace_launches=st["n_trace_launches"], ms_raygen=st["ms_raygen"])
            log(name, out[name])
            S.close()
        finally:
            os.environ.pop("DNSMC_LEARNED_FUSED", None)
    return out


def main(argv):
    what = argv[1] if len(argv) > 1 else "errors"
    if what == "errors":
        out = study_errors()
    elif what == "precision":
        out = study_precision()
    elif what == "heldout":
        out = study_heldout(mix_eps=float(argv[argv.index("--mix-eps") + 1]) if "--mix-eps" in argv else 0.05)
    elif what == "legs":
        out = study_legs()
    elif what == "convergence":
        mlp = None
        if "--learned" in argv:
            from .learned import train_channel_proposal
            mlp, _ = train_channel_proposal()
        eps = float(argv[argv.index("--mix-eps") + 1]) if "--mix-eps" in argv else 0.05
        out = study_convergence(mlp=mlp, mix_eps=eps)
    else:
        raise SystemExit("usage: python -m dnsmc_amd.study errors|precision|heldout [--mix-eps E]|"
                         "convergence [--learned] [--mix-eps E]|legs")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main(sys.argv)
