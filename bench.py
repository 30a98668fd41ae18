#!/usr/bin/env python3
"""Benchmark of the Monte Carlo radiation solve (one GPU).

A step is one radiation solve: every emitting cell of every snapshot traced
with R bundles (ray generation + traversal + tallies + finalisation), inputs
resident in HBM.  Default workload: ck384x8 = 384x256x256 channel, 8
temperature snapshots (profile + Fourier modes) in one launch, synthetic
100x16x69 correlated-k table, 100 rays/cell (2.0e10 rays per step).

Prints ONE JSON line with rays/s (whole job), the traversal kernel's
algorithmic HBM bandwidth against the MI355X roofline (with the HBM traffic
and L2 hit rate measured by rocprof for the same command, profiles/), the
CPU baseline (the oracle's restatement of the reference algorithm on this
host's cores, timed on a bounded sample of emitting cells), and the learned
proposal against the plain estimator: RMS / max error of Q_rad against a
converged FP64 reference and the time to 1% RMS, on the fine grid and with
the reference's default mesh coarsening.
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
    return dict(value=rays / dt, unit="rays/s", cores=threads, kind="port",
                sample="%d random emitting cells of snapshot 0 x %d rays (%.1f s, %.3g segments/s)"
                       % (k2, R, dt, nsteps / dt))


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
    return {"traffic_bytes_per_launch": float(kv["traffic_bytes_per_launch"]), "l2_hit_rate": float(kv["l2_hit_rate"]),
            "source": os.path.relpath(path, ROOT)}


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


def learned_leg(args, cfg, R, fields, res_p, t_plain, cfg_kw):
    """Train the proposal, run the learned solve of the workload, and compare
    both estimators with a converged FP64 reference on 1024 cells of snapshot
    0: RMS / max error at R rays per cell and the time to 1% RMS."""
    from dnsmc_amd import learned, study
    cells = study.sample_cells(args.workload, args.ref_cells)
    Qref, vref, t_ref = study.reference(args.workload, cells, args.ref_rays, fields=fields, **cfg_kw)
    weights, info = learned.train_channel_proposal(args.train_snaps, args.train_rays, args.cost0, args.train_epochs,
                                                   **cfg_kw)
    log("proposal: %s" % (info,))
    SL, _ = study.make_solver(args.workload, R, nsnap=cfg.nsnap, mode=2, mlp=weights, mix_eps=args.mix_eps,
                              fields=fields, **cfg_kw)
    dt_l, st_l = timed_solve(SL, 1)
    res_l = SL.results(0)
    SL.close()
    err_p = study.errors(res_p["Q"][cells], Qref, vref)
    err_l = study.errors(res_l["Q"][cells], Qref, vref)
    eps_p, eps_l = err_p["rms"], err_l["rms"]
    t_p = t_plain * (eps_p / 0.01) ** 2
    t_l = dt_l * (eps_l / 0.01) ** 2
    var_ratio = float(np.sum(res_p["sdev"] ** 2) / np.sum(res_l["sdev"] ** 2))
    out = {}
    out["learned"] = dict(info, mix_eps=args.mix_eps, ms_per_step=dt_l * 1000.0,
                          rays_per_s=st_l[0]["n_rays"] / dt_l,
                          segments_per_ray=st_l[0]["n_steps"] / max(st_l[0]["n_rays"], 1),
                          variance_ratio_vs_plain=var_ratio)
    out["rms_error_vs_fp64_oracle"] = dict(
        plain=eps_p, learned=eps_l, max_plain=err_p["max"], max_learned=err_l["max"], rays_per_cell=R,
        snapshot=0, cells=int(len(cells)),
        reference=dict(rays_per_cell=args.ref_rays, seconds=round(t_ref, 1),
                       kind="converged FP64 MC (FP64 traversal and tallies, other seed), the oracle's estimator"),
        estimate_from_variance=dict(plain=math.sqrt(np.sum(res_p["sdev"] ** 2) / np.sum(res_p["Q"] ** 2)),
                                    learned=math.sqrt(np.sum(res_l["sdev"] ** 2) / np.sum(res_l["Q"] ** 2))))
    # rays per cell for 1% RMS (error ~ R^-1/2) and the cost of a solve as a
    # function of R: fixed part (per-cell proposal, launch, tails) plus a
    # marginal cost per ray per cell, measured on snapshot 0 at R and 4R and
    # scaled to the workload's snapshots
    R1_p = R * (eps_p / 0.01) ** 2
    R1_l = R * (eps_l / 0.01) ** 2
    f0 = fields[:1]
    lin = {}
    for name, mode, W in (("plain", 0, None), ("learned", 2, weights)):
        ts = []
        for RR in (R, 4 * R):
            S1, _ = study.make_solver(args.workload, RR, mode=mode, mlp=W, mix_eps=args.mix_eps, fields=f0, **cfg_kw)
            S1.run()
            S1.wait()
            reps = []
            for _ in range(3):   # median of three solves
                t0 = time.perf_counter()
                S1.run()
                S1.wait()
                reps.append(time.perf_counter() - t0)
            ts.append(float(np.median(reps)))
            S1.close()
        b = (ts[1] - ts[0]) / (3 * R)
        a = ts[0] - b * R
        lin[name] = (a, b, ts)
    ns = cfg.nsnap
    T1_p = ns * (lin["plain"][0] + lin["plain"][1] * R1_p)
    T1_l = ns * (lin["learned"][0] + lin["learned"][1] * R1_l)
    out["time_to_1pct_rms_s"] = dict(
        plain=T1_p, learned=T1_l, speedup_learned=T1_p / T1_l,
        rays_per_cell_to_1pct=dict(plain=R1_p, learned=R1_l),
        solve_s_snapshot0=dict(plain={str(R): lin["plain"][2][0], str(4 * R): lin["plain"][2][1]},
                               learned={str(R): lin["learned"][2][0], str(4 * R): lin["learned"][2][1]}),
        at_bench_rays=dict(plain=t_p, learned=t_l, speedup_learned=t_p / t_l),
        note="rays per cell to 1%% RMS from the error at %d rays/cell (error ~ R^-1/2); solve time at that "
             "R from the measured fixed + per-ray cost on snapshot 0 (R and 4R, median of three solves each) times %d snapshots; "
             "at_bench_rays scales the whole measured solve linearly instead" % (R, ns))
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
    ap.add_argument("--copy-gib", type=float, default=1.0, help="size of the achievable-bandwidth copy (0: skip)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal leg and the error study")
    ap.add_argument("--no-amc", action="store_true", help="skip the leg with the reference's default mesh coarsening")
    ap.add_argument("--train-snaps", type=int, default=20)
    ap.add_argument("--train-rays", type=int, default=8)
    ap.add_argument("--train-epochs", type=int, default=2)
    ap.add_argument("--cost0", type=float, default=16.0)
    ap.add_argument("--mix-eps", type=float, default=0.1)
    ap.add_argument("--ref-cells", type=int, default=1024)
    ap.add_argument("--ref-rays", type=int, default=65536)
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
                   "coarsening": "none (grid_num 1); reference_amc repeats the comparison with the reference's "
                                 "default 5-level coarsening"},
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
        # measured HBM traffic (cache-line granular gathers, L2 misses) per second
        out["roofline"]["traffic_GBs_per_launch"] = tb / (ms_trace / max(launches, 1) * 1e-3) / 1e9
        out["roofline"]["traffic_GBs_whole_device"] = tb * launches / dt / 1e9
    if args.copy_gib > 0:
        import ctypes
        from dnsmc_amd._lib import check, lib
        g = ctypes.c_double(0.0)
        check(lib().dnsmc_bench_copy(int(args.copy_gib * 2 ** 30), 5, ctypes.byref(g)))
        out["roofline"]["achievable_copy_GBs"] = g.value
        out["roofline"]["frac_of_achievable"] = achieved / g.value
        if prof:
            out["roofline"]["traffic_frac_of_achievable_whole_device"] = out["roofline"]["traffic_GBs_whole_device"] / g.value
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(S, cfg, fields, R, args.cpu_seconds)
        except Exception as e:  # reported, never fatal to the GPU number
            log("cpu baseline failed: %r" % (e,))
    if not args.no_learned:
        res_p = S.results(0)
        S.close()
        out.update(learned_leg(args, cfg, R, fields, res_p, ms_step / 1000.0, {}))
        if not args.no_amc:
            # the same comparison with the reference's default adaptive mesh
            # coarsening (code/param.h: 5 levels, 5 steps on each of the first four)
            amc = dict(grid_num=5, maxs=(5, 5, 5, 5, 100000, 100000))
            from dnsmc_amd import study
            SA, _ = study.make_solver(args.workload, R, nsnap=cfg.nsnap, fields=fields, **amc)
            dt_a, st_a = timed_solve(SA, 1)
            res_a = SA.results(0)
            SA.close()
            leg = learned_leg(args, cfg, R, fields, res_a, dt_a, amc)
            out["reference_amc"] = dict(
                config="grid_num 5, maxs 5 5 5 5 100000 (the reference's code/param.h defaults), else as above",
                plain=dict(rays_per_s=st_a[0]["n_rays"] / dt_a, ms_per_step=dt_a * 1000.0,
                           segments_per_ray=st_a[0]["n_steps"] / max(st_a[0]["n_rays"], 1)),
                **leg)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
