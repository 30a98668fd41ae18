#!/usr/bin/env python3
"""Benchmark of the Monte Carlo radiation solve (one GPU).

A step is one radiation solve: every emitting cell of every snapshot traced
with R bundles (ray generation + traversal + tallies + finalisation), inputs
resident in HBM.  Default workload: ck384x8 = 384x256x256 channel, 8
temperature snapshots (profile + Fourier modes) in one launch, synthetic
100x16x69 correlated-k table, 100 rays/cell (2.0e10 rays per step).

Prints ONE JSON line: rays/s (whole job); the traversal kernel's algorithmic
HBM bandwidth against the MI355X peak and against the achievable streaming
bandwidth measured in the same run, with the HBM traffic and L2 hit rate
rocprof measured for the same command (profiles/); the CPU baseline (the
oracle's restatement of the reference algorithm on this host's cores: a
bounded sample of the workload, and full runs of the two small
configurations); the accuracy legs (plain and learned estimators, fine grid
and the reference's default coarsening, all against one converged fine-grid
FP64 reference): RMS / max error, rays per cell to 1% RMS measured by
doubling, the coarsening's bias and the wall-clock to 1% RMS; and one key per
other configuration (grey 32^3, 64x96x64, 256x128x128 with its precision
study).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rays/s and wall-clock to 1% RMS Q_rad error vs FP64 oracle (plain vs learned)"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(workload, rays, nsnap=None, mode=0, batch=None):
    import dnsmc_amd
    from dnsmc_amd import workloads
    w = workloads.WORKLOADS[workload]
    R = rays or w["rays"]
    ns = nsnap or w.get("nsnap", 1)
    extra = {"batch_rays": int(batch)} if batch else {}
    cfg = workloads.channel_config(w["n"], rays_per_cell=R, spectral=w["spectral"], Tw=w["Tw"], nsnap=ns, mode=mode,
                                   **extra)
    S = dnsmc_amd.Solver(cfg)
    t0 = time.time()
    if w["spectral"] == 3:
        S.load_tables()
        table = None
    else:
        table = workloads.synth_ck_table(11)
        S.set_tables(workloads.table_arrays(table))
    fields = workloads.workload_fields(workload, ns)
    for s, T in enumerate(fields):
        S.set_field(s, workloads.ghosted(T))
    log("setup %s: %.1f s (fields %s x %d)" % (workload, time.time() - t0, w["n"], ns))
    return S, cfg, fields, table, R


def oracle_problem(S, cfg, Tint):
    n = list(cfg.n)
    faces = [S.array("faces%d" % a, 0) for a in range(3)]
    sc = S.array("scalars")
    d = dict(levels=[dict(n=n, faces=faces, T=np.ascontiguousarray(Tint.transpose(2, 1, 0)).ravel(),
                          maxs=cfg.maxs[0], centres=[S.array("centres%d" % a, 0) for a in range(3)])],
             bc=[[cfg.bc[a][s] for s in range(2)] for a in range(3)], polar_axis=cfg.polar_axis,
             tau_cut=float(sc[10]), spectral=1 if cfg.spectral == 3 else 0, nB=int(sc[3]), nQ=int(sc[4]),
             nT=int(sc[5]), inv_dT=float(sc[6]), mT0=float(sc[7]), kq=S.array("kq").copy(),
             Ib=S.array("Ib"), Ibw=S.array("Ibw"), kappa_grey=float(cfg.kappa_grey), cdf_band=S.array("cdf_band"),
             cdf_g=S.array("cdf_g"), inv_emax=S.array("inv_emax"), Ibmax=float(sc[2]), Ibmax_d=S.Ibmax_d(),
             Tmax=float(sc[0]), g_search=int(sc[8]), u_off=float(sc[9]))
    return d


def cpu_baseline(S, cfg, fields, R, budget_s):
    """Oracle (FP32 rays, FP64 tallies, OpenMP over cells) on a seeded sample
    of emitting cells of snapshot 0; rays/s on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.get()
    P = o.problem(oracle_problem(S, cfg, fields[0]))
    ncell = int(np.prod(cfg.n))
    rng = np.random.Generator(np.random.Philox(key=7))
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, 64))
    k = 4 * threads
    cells = rng.choice(ncell, size=k, replace=False).astype(np.int32)
    t0 = time.time()
    o.mc_philox(P, cfg.seed, 0, R, cells=cells, nthreads=threads)
    dt = time.time() - t0
    k2 = int(min(ncell, max(k, k * budget_s / max(dt, 1e-3))))
    cells = rng.choice(ncell, size=k2, replace=False).astype(np.int32)
    t0 = time.time()
    _, _, _, nsteps = o.mc_philox(P, cfg.seed, 0, R, cells=cells, nthreads=threads)
    dt = time.time() - t0
    rays = k2 * R
    out = dict(value=rays / dt, unit="rays/s", cores=threads, kind="port",
               sample="%d random emitting cells of snapshot 0 x %d rays (%.1f s, %.3g segments/s)"
                      % (k2, R, dt, nsteps / dt))
    out["full_runs"] = cpu_full_runs(o, threads)
    return out


def cpu_full_runs(o, threads):
    """The oracle over every cell of the grey 32^3 (100 rays/cell) and the
    64x96x64 correlated-k (500 rays/cell) workloads on this host's cores, and
    the largest relative difference of its emitted energies from the GPU
    solve of the same rays (the estimator's arithmetic, FP32 rays)."""
    from dnsmc_amd import study, workloads
    out = {}
    for wl in ("grey32", "ck64"):
        R = workloads.WORKLOADS[wl]["rays"]
        fields = workloads.workload_fields(wl, 1)
        S, cfg = study.make_solver(wl, R, fields=fields)
        S.run()
        r = S.results(0)
        P = o.problem(oracle_problem(S, cfg, fields[0]))
        S.close()
        t0 = time.time()
        E, A, _, ns = o.mc_philox(P, cfg.seed, 0, R, nthreads=threads)
        dt = time.time() - t0
        ncell = int(np.prod(cfg.n))
        out[wl] = dict(value=ncell * R / dt, unit="rays/s", cores=threads, seconds=round(dt, 2),
                       rays=ncell * R, segments_per_s=ns / dt,
                       max_rel_diff_E_vs_gpu=float(np.max(np.abs(E - r["E"]) / np.abs(r["E"]))),
                       max_rel_diff_A_vs_gpu=float(np.max(np.abs(A - r["A"]) / np.abs(r["A"]))))
        log("cpu full", wl, out[wl])
    return out


def measured_traffic(args, cfg, R):
    """HBM bytes per traversal launch measured by rocprofv3 (FETCH_SIZE and
    WRITE_SIZE, each scaled by its reading for a copy kernel of known size) on
    this same workload; the summary in profiles/ records the command."""
    path = os.path.join(ROOT, "profiles", "trace_%s_traffic.txt" % args.workload)
    if not os.path.exists(path):
        return None
    kv = {}
    with open(path) as f:
        for line in f:
            if "=" in line and not line.startswith("#"):
                k, v = line.split("=", 1)
                kv[k.strip()] = v.strip()
    want = "%s rays=%d nsnap=%d batch=%s" % (args.workload, R, cfg.nsnap, args.batch)
    if kv.get("workload") != want or "traffic_bytes_per_launch" not in kv:
        return None
    out = {"traffic_bytes_per_launch": float(kv["traffic_bytes_per_launch"]), "l2_hit_rate": float(kv["l2_hit_rate"]),
           "source": os.path.relpath(path, ROOT)}
    # the co-limits the same counter runs show: VALU instructions issued per
    # CU-cycle (at most 1: four 16-lane SIMDs), texture-data unit busy, L2
    # requests and misses per segment
    for k in ("valu_issue_per_cu_cycle", "td_busy_per_cu_cycle", "l2_requests_per_segment", "l2_misses_per_segment"):
        if k in kv:
            out[k] = float(kv[k])
    return out


def timed_solve(S, steps):
    import dnsmc_amd
    S.run()
    S.wait()
    dnsmc_amd.sync()
    st_all = []
    t0 = time.perf_counter()
    for _ in range(steps):
        S.run()
        st_all.append(S.stats())
    dnsmc_amd.sync()
    return (time.perf_counter() - t0) / steps, st_all


def shard_leg(S, cfg, steps, full_rays_per_s):
    """One of 8 shards of the solve, as a multi-GPU run would hand it to one
    GPU: the emitting cells with i < n0/8 (a slab of the periodic x axis; all
    j and k), every snapshot, tracing through the full replicated field.  The
    solver keeps its fields; the emitters are restored to the whole grid."""
    import dnsmc_amd
    n0, n1, n2 = cfg.n
    nx = n0 // 8
    i, j, k = np.meshgrid(np.arange(nx), np.arange(n1), np.arange(n2), indexing="ij")
    cells = (i + n0 * (j + n1 * k)).ravel().astype(np.int32)
    S.set_emitters(cells)
    S.run()
    S.wait()
    dnsmc_amd.sync()
    nrays = nsteps = 0
    ms_trace = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        S.run()
        st = S.stats()
        nrays += st["n_rays"]
        nsteps += st["n_steps"]
        ms_trace += st["ms_trace"]
    dnsmc_amd.sync()
    dt = time.perf_counter() - t0
    S.set_emitters_slab(1)
    v = nrays / dt
    return {"emitters": "i in [0, %d) of %d, all j and k: %d cells x %d snapshots" % (nx, n0, len(cells), cfg.nsnap),
            "rays_per_solve": nrays // steps, "ms_per_solve": 1000.0 * dt / steps,
            "ms_trace_per_solve": ms_trace / steps, "rays_per_s": v, "segments_per_ray": nsteps / max(nrays, 1),
            "rays_per_s_over_whole_grid": v / full_rays_per_s,
            "note": "one GPU's share of an 8-GPU run sharded along x with the field replicated (the field "
                    "all-gather before and the Q gather after are not part of it)"}


def solve_cost(workload, R, fields, mode=0, mlp=None, mix_eps=0.1, **cfg_kw):
    """Full-grid solves of snapshot 0 at R, 4R and 8R rays per cell (each after
    a warm-up solve): the fixed part a and the per-ray-per-cell part b of
    t(R') = a + b R' fitted on 4R and 8R, the three times, and the stats and
    results of the R solve.  The fit uses the larger two because the rays per
    cell that 1 % RMS needs (10^5 and more) run the way those do: above
    2^32 - 2^20 rays per 384x256x256 snapshot (170 rays/cell) a solve is cut
    into several launches, and the learned generation overlaps the traversal
    of the previous batch instead of preceding a snapshot's one launch."""
    from dnsmc_amd import study
    ts = []
    keep = None
    for RR in (R, 4 * R, 8 * R):
        S1, _ = study.make_solver(workload, RR, mode=mode, mlp=mlp, mix_eps=mix_eps, fields=fields[:1], **cfg_kw)
        S1.run()
        S1.wait()
        t0 = time.perf_counter()
        S1.run()
        S1.wait()
        ts.append(time.perf_counter() - t0)
        if keep is None:
            keep = (S1.stats(), S1.results(0))
        S1.close()
    b = (ts[2] - ts[1]) / (4 * R)
    return ts[1] - b * 4 * R, b, ts, keep[0], keep[1]


def learned_draw_paths(workload, R, fields, W, mix_eps):
    """One learned solve of snapshot 0 at R with each of the two ways of
    drawing learned rays: the default (proposal rows of 64 cells per MFMA pass
    and the draws in generation kernels, overlapped with the traversal of the
    previous snapshot) and the draw inside the traversal (each wave evaluates
    its cells' rows itself, DNSMC_LEARNED_FUSED=1)."""
    from dnsmc_amd import study
    out = {}
    for key, env in (("separate_generation", None), ("in_traversal", "1")):
        if env:
            os.environ["DNSMC_LEARNED_FUSED"] = env
        try:
            S1, _ = study.make_solver(workload, R, mode=2, mlp=W, mix_eps=mix_eps, fields=fields[:1])
            S1.run()
            S1.wait()
            t0 = time.perf_counter()
            S1.run()
            S1.wait()
            out[key] = time.perf_counter() - t0
            st = S1.stats()
            out[key + "_trace_launches"] = st["n_trace_launches"]
            S1.close()
        finally:
            os.environ.pop("DNSMC_LEARNED_FUSED", None)
    return out


def accuracy_legs(args, cfg, R, fields):
    """Plain and learned estimators on the fine grid and with the reference's
    default mesh coarsening, all scored against ONE converged FP64 reference
    on the fine grid (1024 sampled cells of snapshot 0), so the coarsening's
    interpolation bias counts: RMS / max error at the bench's R, rays per cell
    to 1% RMS measured by doubling R, and the wall-clock of a full solve at
    that R (the workload's snapshots, fixed + per-ray cost measured on the
    full grid)."""
    from dnsmc_amd import learned, study
    wl = args.workload
    cells = study.sample_cells(wl, args.ref_cells)
    Qref, vref, t_ref = study.reference(wl, cells, args.ref_rays, fields=fields)
    ref_noise = math.sqrt(float(np.sum(vref)) / float(np.sum(Qref ** 2)))
    log("reference: %d rays/cell on %d cells, %.1f s, RMS noise %.4f" % (args.ref_rays, len(cells), t_ref, ref_noise))
    amc = dict(grid_num=5, maxs=(5, 5, 5, 5, 100000, 100000))
    legs = {}
    weights = {}
    for grid, kw in (("fine", {}), ("amc", amc)):
        W, info = learned.train_channel_proposal(args.train_snaps, args.train_rays, args.cost0, args.train_epochs, **kw)
        weights[grid] = W
        log("proposal (%s): %s" % (grid, info))
        for est, mode in (("plain", 0), ("learned", 2)):
            name = "%s_%s" % (est, grid)
            Wm = W if mode == 2 else None
            # snapshot 0 on the full grid at R, 4R and 8R: cost model, rays/s, results
            a, b, ts, st0, res = solve_cost(wl, R, fields, mode=mode, mlp=Wm, mix_eps=args.mix_eps, **kw)
            e = study.errors(res["Q"][cells], Qref, vref)
            leg = dict(rms_at_bench_R=e["rms"], max_at_bench_R=e["max"],
                       rms_from_own_variance=math.sqrt(float(np.sum(res["sdev"][cells] ** 2)) /
                                                       float(np.sum(Qref ** 2))),
                       rays_per_s_snapshot0=st0["n_rays"] / ts[0],
                       segments_per_ray=st0["n_steps"] / max(st0["n_rays"], 1),
                       solve_s_snapshot0={str(R): ts[0], str(4 * R): ts[1], str(8 * R): ts[2]},
                       cost_model_s=dict(fixed=a, per_ray_per_cell=b))
            if mode == 2:
                leg["proposal"] = info
                leg["variance_ratio_vs_plain"] = float(np.sum(legs["plain_" + grid]["_sd2"]) / np.sum(res["sdev"] ** 2))
                if grid == "fine":
                    leg["draw_paths_s_snapshot0"] = learned_draw_paths(wl, R, fields, Wm, args.mix_eps)
            leg["_sd2"] = res["sdev"] ** 2
            # rays per cell to 1% RMS, measured by doubling R on the sampled cells
            tt = study.rays_to_target(wl, cells, Qref, vref, fields, mode=mode, mlp=Wm, mix_eps=args.mix_eps,
                                      R0=R, Rmax=args.rmax, **kw)
            leg["to_1pct"] = tt
            R1 = tt["rays_per_cell_interp"]
            leg["time_to_1pct_s"] = cfg.nsnap * (a + b * R1) if R1 else None
            R2 = tt["rays_per_cell"]
            leg["time_to_1pct_s_first_doubling"] = cfg.nsnap * (a + b * R2) if R2 else None
            R1f = tt["fit"]["rays_per_cell"]
            leg["time_to_1pct_s_fit"] = cfg.nsnap * (a + b * R1f) if R1f else None
            legs[name] = leg
            log(name, {k: v for k, v in leg.items() if k not in ("_sd2", "to_1pct", "proposal")})
    for leg in legs.values():
        leg.pop("_sd2", None)

    def ratio(x, y):
        return x / y if (x and y) else None
    out = dict(reference=dict(rays_per_cell=args.ref_rays, cells=len(cells), seconds=round(t_ref, 1),
                              rms_noise=ref_noise, grid="fine (grid_num 1)", snapshot=0,
                              kind="converged FP64 MC (FP64 traversal and tallies, other seed), the oracle's "
                                   "estimator; its own variance is removed from every error"),
               legs=legs)
    out["time_to_1pct_rms_s"] = dict(
        plain=legs["plain_fine"]["time_to_1pct_s"], learned=legs["learned_fine"]["time_to_1pct_s"],
        speedup_learned=ratio(legs["plain_fine"]["time_to_1pct_s"], legs["learned_fine"]["time_to_1pct_s"]),
        plain_amc=legs["plain_amc"]["time_to_1pct_s"], learned_amc=legs["learned_amc"]["time_to_1pct_s"],
        speedup_learned_amc=ratio(legs["plain_amc"]["time_to_1pct_s"], legs["learned_amc"]["time_to_1pct_s"]),
        speedup_learned_amc_vs_plain_fine=ratio(legs["plain_fine"]["time_to_1pct_s"],
                                                legs["learned_amc"]["time_to_1pct_s"]),
        rays_per_cell_to_1pct={k: v["to_1pct"]["rays_per_cell_interp"] for k, v in legs.items()},
        amc_bias_rms=dict(plain=legs["plain_amc"]["to_1pct"]["fit"]["bias_rms"],
                          learned=legs["learned_amc"]["to_1pct"]["fit"]["bias_rms"]),
        note="rays per cell to 1%% RMS measured by doubling R from %d on %d cells of snapshot 0 against the "
             "fine-grid FP64 reference, interpolated between the two doublings that bracket 1%% (each leg's "
             "to_1pct holds the points, the first R below 1%% with its error, and the fit eps^2 = b^2 + c/R whose b "
             "is the coarsening bias); time = %d snapshots x (fixed + per-ray cost of a full-grid solve, "
             "fitted on full-grid solves at 4R and 8R, the regime of the large rays per cell)" % (R, len(cells), cfg.nsnap))
    return out


def other_configs(args):
    """The other configurations of the workload list, one key each: rays/s of
    the full solve at the configuration's rays per cell and its RMS / max error
    against a converged FP64 reference on 1024 sampled cells (grey 32^3 also
    isothermal against the plane-parallel E2 solution); the 256x128x128
    precision study (FP32 / FP64 traversal x FP64 / Kahan-FP32 tallies)."""
    from dnsmc_amd import study, workloads
    out = {}
    for key, wl in (("grey32", "grey32"), ("ck64", "ck64"), ("ck256", "ck256")):
        w = workloads.WORKLOADS[wl]
        R = w["rays"]
        fields = workloads.workload_fields(wl, 1)
        cells = study.sample_cells(wl, args.ref_cells)
        Qref, vref, t_ref = study.reference(wl, cells, args.cfg_ref_rays, fields=fields)
        S, cfg = study.make_solver(wl, R, fields=fields)
        S.run()
        S.wait()
        t0 = time.perf_counter()
        S.run()
        st = S.stats()
        dt = time.perf_counter() - t0
        r = S.results(0)
        S.close()
        e = study.errors(r["Q"][cells], Qref, vref)
        d = dict(grid=list(cfg.n), rays_per_cell=R, rays_per_s=st["n_rays"] / dt, segments_per_s=st["n_steps"] / dt,
                 solve_s=dt, rms=e["rms"], max=e["max"],
                 reference=dict(rays_per_cell=args.cfg_ref_rays, cells=len(cells), seconds=round(t_ref, 2),
                                rms_noise=math.sqrt(float(np.sum(vref)) / float(np.sum(Qref ** 2)))))
        if wl == "grey32":
            d["isothermal_vs_E2"] = grey_isothermal(R)
        if wl == "ck256":
            d["precision"] = precision_variants(wl, R, fields, cells, Qref, vref)
        out[key] = d
        log(key, {k: v for k, v in d.items() if k != "precision"})
    return out


def grey_isothermal(R):
    """32^3 grey channel at exactly 1000 K, walls 300 K, kappa = 1: Q(y) =
    2 kappa sigma (T^4 - Tw^4) (E2(kappa y) + E2(kappa (L - y))), the exact
    expectation of the cell-centre estimator but for the transmissivity
    cut-off (toll 1e-3), which moves it by less than 1e-3 relative."""
    from scipy.special import expn
    from dnsmc_amd import study, workloads
    n = workloads.WORKLOADS["grey32"]["n"]
    S, cfg = study.make_solver("grey32", R, fields=[np.full(n, 1000.0, np.float32)])
    S.run()
    r = S.results(0)
    fy = S.array("faces1", 0).astype(np.float64)
    S.close()
    y = 0.5 * (fy[1:] + fy[:-1])
    L = fy[-1]
    want = 2.0 * 5.670373e-08 * (1000.0 ** 4 - 300.0 ** 4) * (expn(2, y) + expn(2, L - y))
    Q = r["Q"].reshape(n[2], n[1], n[0])
    diff = Q - want[None, :, None]
    return dict(rms=float(np.sqrt(np.sum(diff ** 2) / (np.sum(want ** 2) * n[0] * n[2]))),
                max=float(np.max(np.abs(diff)) / np.max(np.abs(want))),
                plane_mean_max_rel=float(np.max(np.abs(Q.mean(axis=(0, 2)) - want) / np.abs(want))))


def precision_variants(wl, R, fields, cells, Qref, vref):
    """The same Philox rays four ways: against the FP64-traversal / FP64-tally
    run (arithmetic error alone, all cells) and the converged reference."""
    from dnsmc_amd import study
    runs = {}
    for prec, tr in ((0, "fp32"), (1, "fp64")):
        for tally, tn in ((1, "fp64"), (2, "kahan_fp32")):
            S, _ = study.make_solver(wl, R, precision=prec, tally=tally, fields=fields)
            S.run()
            S.wait()
            t0 = time.perf_counter()
            S.run()
            S.wait()
            dt = time.perf_counter() - t0
            runs[(tr, tn)] = (S.results(0), dt)
            S.close()
    base = runs[("fp64", "fp64")][0]
    out = {}
    for (tr, tn), (r, dt) in runs.items():
        same = study.errors(r["Q"], base["Q"])
        conv = study.errors(r["Q"][cells], Qref, vref)
        out["%s_traversal_%s_tally" % (tr, tn)] = dict(same_rays_rms=same["rms"], same_rays_max=same["max"],
                                                       converged_rms=conv["rms"], converged_max=conv["max"],
                                                       solve_s=dt)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ck384x8")
    ap.add_argument("--rays", type=int, default=None)
    ap.add_argument("--nsnap", type=int, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", type=float, default=None, help="rays per traversal launch")
    ap.add_argument("--copy-gib", type=float, default=8.0, help="GiB of the achievable-bandwidth streams (0: skip)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal leg and the error study")
    ap.add_argument("--no-configs", action="store_true", help="skip the grey 32^3, 64x96x64 and 256x128x128 keys")
    ap.add_argument("--no-shard", action="store_true", help="skip the one-of-8 x-shard solve")
    ap.add_argument("--rmax", type=int, default=1 << 23, help="largest rays per cell of the doubling study")
    ap.add_argument("--train-snaps", type=int, default=20)
    ap.add_argument("--train-rays", type=int, default=8)
    ap.add_argument("--train-epochs", type=int, default=2)
    ap.add_argument("--cost0", type=float, default=16.0)
    ap.add_argument("--mix-eps", type=float, default=0.1)
    ap.add_argument("--ref-cells", type=int, default=1024)
    ap.add_argument("--ref-rays", type=int, default=6000000)
    ap.add_argument("--cfg-ref-rays", type=int, default=65536)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(emitting cells are independent units of work: a multi-GPU solve shards them, replicating the field)", file=sys.stderr)
        sys.exit(2)

    import dnsmc_amd
    S, cfg, fields, table, R = setup(args.workload, args.rays, args.nsnap, batch=args.batch)
    for _ in range(args.warmup):
        S.run()
        S.wait()
    dnsmc_amd.sync()
    ms_trace = 0.0
    ms_gen = 0.0
    nrays = 0
    nsteps = 0
    launches = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        S.run()
        st = S.stats()   # waits; HIP-event times of this step's kernels
        ms_trace += st["ms_trace"]
        ms_gen += st["ms_raygen"]
        nrays += st["n_rays"]
        nsteps += st["n_steps"]
        launches += st["n_trace_launches"]
        tables = ("global", "Ib in LDS", "kappa+Ib slice in LDS (%d nodes)" % st["table_nw"])[st["table_mode"]]
    dnsmc_amd.sync()
    t1 = time.perf_counter()
    dt = t1 - t0
    ms_step = 1000.0 * dt / args.steps
    ncell = int(np.prod(cfg.n)) * cfg.nsnap
    # algorithmic bytes of the traversal: 4 B per segment (cell T), 4 B per ray
    # (emitter T), 16 B per cell and snapshot (FP64 E and A written once)
    b_alg = 4.0 * nsteps + 4.0 * nrays + 16.0 * ncell * args.steps
    achieved = b_alg / (ms_trace * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": nrays / dt,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "grid": list(cfg.n), "snapshots": cfg.nsnap, "rays_per_cell": R,
                   "table": "grey kappa=1" if table is None else "synthetic c-k 100x16x69",
                   "estimator": "plain (Tmax proposal, Philox)", "tally": "fp64", "tables": tables,
                   "coarsening": "none (grid_num 1); accuracy.legs repeat the comparison with the reference's "
                                 "default 5-level coarsening",
                   "timed_region": "fields already in HBM: dnsmc_set_field (host copy, bricking, coarse levels, "
                                   "about 4 ms per 384x256x256 snapshot) runs before it"},
        "segments_per_s": nsteps / dt,
        "segments_per_ray": nsteps / max(nrays, 1),
        "ms_trace_per_step": ms_trace / args.steps,
        "ms_raygen_per_step": ms_gen / args.steps,
        "trace_launches_per_step": launches // args.steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "trace", "algorithmic_bytes_per_launch": b_alg / max(launches, 1),
                     "launch_ms_avg": ms_trace / max(launches, 1),
                     "concurrent_launches": "up to %d (one per pipeline stream)" % st["n_pipes"],
                     "achieved_whole_device": b_alg / dt / 1e9},
        "cpu_baseline": None,
    }
    prof = measured_traffic(args, cfg, R)
    if prof:
        tb = prof["traffic_bytes_per_launch"]
        out["roofline"]["traffic"] = tb
        out["roofline"]["traffic_source"] = prof["source"]
        out["roofline"]["l2_hit_rate"] = prof["l2_hit_rate"]
        for k in ("valu_issue_per_cu_cycle", "td_busy_per_cu_cycle", "l2_requests_per_segment", "l2_misses_per_segment"):
            if k in prof:
                out["roofline"][k] = prof[k]
        # measured HBM traffic (cache-line granular gathers, L2 misses) per second
        out["roofline"]["traffic_GBs_per_launch"] = tb / (ms_trace / max(launches, 1) * 1e-3) / 1e9
        out["roofline"]["traffic_GBs_whole_device"] = tb * launches / dt / 1e9
    if args.copy_gib > 0:
        import ctypes
        from dnsmc_amd._lib import check, lib
        g = ctypes.c_double(0.0)
        check(lib().dnsmc_bench_copy(int(2 ** 30), 5, ctypes.byref(g)))
        out["roofline"]["copy_f32x4_1GiB_GBs"] = g.value   # the byte counters' calibration kernel
        nb = int(args.copy_gib * 2 ** 30)
        best = ctypes.c_int(-1)
        check(lib().dnsmc_bench_stream(nb, 5, 1, ctypes.byref(g), ctypes.byref(best)))
        rd = g.value
        check(lib().dnsmc_bench_stream(nb, 5, 0, ctypes.byref(g), ctypes.byref(best)))
        cp = g.value
        r = out["roofline"]
        r["achievable_read_GBs"] = rd
        r["achievable_copy_GBs"] = cp
        r["achievable_note"] = ("best of 12 streaming variants over %.0f GiB (16 B per lane, 1 or 4 per iteration, "
                                "plain or non-temporal, 4 to 32 blocks per CU); the traversal reads "
                                "(writes are 0.5 %% of its traffic), so the read-only rate is its ceiling"
                                % args.copy_gib)
        r["frac_of_achievable"] = achieved / rd
        if prof:
            r["traffic_frac_of_achievable_whole_device"] = r["traffic_GBs_whole_device"] / rd
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(S, cfg, fields, R, args.cpu_seconds)
        except Exception as e:  # reported, never fatal to the GPU number
            log("cpu baseline failed: %r" % (e,))
    if not args.no_shard and cfg.n[0] % 8 == 0:
        out["one_shard_of_8"] = shard_leg(S, cfg, args.steps, out["value"])
        log("shard", out["one_shard_of_8"])
    S.close()
    if not args.no_learned:
        out["accuracy"] = accuracy_legs(args, cfg, R, fields)
        out["time_to_1pct_rms_s"] = out["accuracy"].pop("time_to_1pct_rms_s")
    if not args.no_configs:
        out["configs"] = other_configs(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
