#!/usr/bin/env python3
"""Throughput benchmark of the coupled MRT-LBM step on one GPU.

Default workload: the 3D cubic cavity, 256^3 nodes, D3Q19 + D3Q7, Ra = 1e5,
Pr = 0.71, erf grid with A = 3.2, FP32 shifted populations, starting from
the reference initial condition (rest, linear temperature).  A "step" is one
time step of the whole lattice (one fused kernel launch).

Prints ONE JSON line: MLUPS (million lattice-node updates per second) with a
roofline block for the step kernel (algorithmic bytes per launch divided by
the average launch time measured with HIP events on the solver's own stream)
and a CPU baseline (the FP64 CPU oracle, a bounded sample of the same
workload on the host cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mpi-multi-process-version_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
DEFAULT_STEPS = 500

WORKLOADS = {
    # check_steps: the reference's own itc_max for the case
    # (natural-convection_3dim_mono.f90:31, natural-convection_2dim_mono.f90:31),
    # the step count at which the FP32 fields are compared with the FP64 path
    "3d256": dict(dim=3, n=256, Ra=1e5, check_steps=2000,
                  label="cavity3d_256_ra1e5_d3q19_d3q7"),
    # the reference's own default 3D grid (natural-convection_3dim_mono.f90:6)
    "3d257": dict(dim=3, n=257, Ra=1e5, check_steps=2000,
                  label="cavity3d_257_ra1e5_d3q19_d3q7"),
    "2d1025": dict(dim=2, n=1025, Ra=1e6, check_steps=20000,
                   label="cavity2d_1025_ra1e6_d2q9_d2q5"),
    # a 2D grid whose step (235 MB of populations written) exceeds the MALL
    "2d2049": dict(dim=2, n=2049, Ra=1e6, check_steps=2000,
                   label="cavity2d_2049_ra1e6_d2q9_d2q5"),
    "batch64": dict(dim=2, count=64, seed=20261203,
                    label="batch_64_cavities2d_n65-513_A1-2.5_Ra1e3-1e6_Pr0.7-7"),
    # 256 seeded 64^2 cavities (the warm-start data generation's cavities),
    # the reference's 20 000 steps (natural-convection_2dim_mono.f90:31)
    "batch256x64": dict(dim=2, count=256, n=64, seed=20261203, steps=20000,
                        label="batch_256_cavities2d_n64_A1-2.5_Ra1e3-1e6_Pr0.7-7"),
    # the 3D stopping rule (natural-convection_3dim_mono.f90:373, 1e-8) on a
    # 64^3 cube: FP64 from rest, and FP32 through the transient then FP64
    "steady3d64": dict(dim=3, n=64, Ra=1e5, steady=True,
                       label="cavity3d_64_ra1e5_to_steady_state"),
    "steady129": dict(dim=2, n=129, Ra=1e4, steady=True,
                      label="cavity2d_129_ra1e4_to_steady_state"),
    "warm64": dict(dim=2, n=64, held_out=100, seed=7, weights="warm64.npz", itc_max=600000,
                   label="warm_start_64x64_cavities_Ra1e3-1e6_Pr0.7-7_A1-2.5"),
    "warm257": dict(dim=2, n=257, held_out=100, seed=7, weights="warm257w28.npz", itc_max=2000000,
                    label="warm_start_257x257_cavities_Ra1e3-1e6_Pr0.7-7_A1-2.5"),
    "warm3d128": dict(dim=3, n=128, held_out=1, cube=dict(Ra=1e5, Pr=0.71, stretch=3.2),
                      weights="warm257w28.npz", itc_max=1200000, eps=2e-6,
                      label="warm_start_128^3_cube_Ra1e5_Pr0.71_A3.2_slice_extruded_64^3_presolve"),
}


def batch_cavities(count, seed, n=None):
    """The batched sweep's cavities (lbmhip.sweep_cavities), or, given n,
    count cavities on n x n grids with the warm start's parameter ranges
    (surrogate.sample_cavities)."""
    if n is not None:
        import surrogate
        return surrogate.sample_cavities(count, seed, n=n)
    import lbmhip
    return lbmhip.sweep_cavities(count, seed)


def traffic_from_profiles(workload, precision, launch):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by the copy-kernel calibration), if a summary for this
    workload, precision and launch shape has been committed under profiles/."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    shape = ",".join(str(launch[k]) for k in ("bx", "by", "remap", "occupancy", "variant"))
    return d.get("bytes_per_launch") if d.get("launch") == shape else None


def cpu_baseline(cfg, threads, sample_steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lbm_oracle
    lbm_oracle.set_threads(threads)
    o = lbm_oracle.Oracle(cfg["dim"], cfg["n"], cfg["Ra"])
    o.init_rest()
    o.step(1)  # first touch of every array
    t0 = time.perf_counter()
    o.step(sample_steps)
    dt = time.perf_counter() - t0
    nodes = cfg["n"] ** cfg["dim"]
    o.close()
    return {"value": nodes * sample_steps / dt / 1e6, "unit": "MLUPS",
            "cores": threads, "kind": "port",
            "sample": f"{sample_steps} steps of the same {cfg['n']}^{cfg['dim']} cavity "
                      f"(FP64 CPU oracle, {threads} OpenMP threads, {dt:.1f} s)"}


BATCH_LAUNCHES = {
    0: "single GPU, one launch per step for all cavities",
    1: "single GPU, one launch per step for all cavities (y-march strips)",
    3: "single GPU, one workgroup per cavity, up to 4096 steps per launch",
    4: "single GPU, one launch per two steps for all cavities",
    5: "single GPU, one cooperative launch per call (up to 4096 steps) for all cavities, "
       "grid-wide barrier between steps",
}


def run_batch(args, cfg):
    """The batched sweep: every cavity advanced by one launch per step."""
    import numpy as np
    import lbmhip
    cavs = batch_cavities(cfg["count"], cfg["seed"], cfg.get("n"))
    steps = cfg.get("steps") if args.steps == DEFAULT_STEPS and cfg.get("steps") else args.steps
    b = lbmhip.Batch(cavs, device=0)
    b.init_rest()
    if args.warmup:
        b.step(args.warmup)
    if args.launch:
        vals = [int(v) for v in args.launch.split(",")] + [0, 0, 0]
        b.set_launch(vals[0], vals[1], variant=vals[4])
    elif not args.no_autotune:
        b.autotune(5)   # restores the state it started from
    launch = b.get_launch()
    _, nodes = b.active()
    t0 = time.perf_counter()
    ms = b.time_steps(steps)
    wall = time.perf_counter() - t0
    args.steps = steps
    sec = ms / 1e3
    mlups = nodes * args.steps / sec / 1e6
    bpu = 2 * 4 * (9 + 5)
    achieved = bpu * nodes / (sec / args.steps) / 1e9
    # per-cavity Nusselt numbers against the FP64 path after the same steps
    nu = np.array(b.nusselt()[0])
    errs = []
    for i, c in enumerate([] if args.no_fp64_check else cavs):
        s = lbmhip.Solver(2, c["n"], c["Ra"], Pr=c["Pr"], stretch=c["stretch"],
                          precision=lbmhip.PREC_FP64)
        s.init_rest()
        s.step(b[i].itc)
        ref = s.nusselt()[0]
        errs.append(abs(nu[i] - ref) / abs(ref))
        s.close()
    errs = np.array(errs)
    sizes = {k: sum(1 for c in cavs if c["n"] == k) for k in sorted({c["n"] for c in cavs})}
    out = {
        "metric": "MLUPS (aggregate over the batch) + % HBM peak (roofline.frac), 1 GPU",
        "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded cavity parameters, reference initial condition",
        "config": {"workload": cfg["label"], "cavities": len(cavs), "nodes": nodes,
                   "edges": sizes, "seed": cfg["seed"], "launch": launch,
                   "parallelism": BATCH_LAUNCHES.get(launch["variant"], BATCH_LAUNCHES[0])},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "algorithmic_bytes_per_launch": bpu * nodes},
        "wall_s": round(wall, 4),
        "check": {"itc": int(b[0].itc),
                  "nu_rel_err_vs_fp64_max": float(errs.max()) if errs.size else None,
                  "nu_rel_err_vs_fp64_median": float(np.median(errs)) if errs.size else None},
    }
    b.close()
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import lbm_oracle
        lbm_oracle.set_threads(args.cpu_threads)
        steps = args.cpu_steps or 20
        t = 0.0
        for c in cavs:
            o = lbm_oracle.Oracle(2, c["n"], c["Ra"], Pr=c["Pr"], A=c["stretch"])
            o.init_rest()
            o.step(1)
            t0 = time.perf_counter()
            o.step(steps)
            t += time.perf_counter() - t0
            o.close()
        out["cpu_baseline"] = {
            "value": nodes * steps / t / 1e6, "unit": "MLUPS", "cores": args.cpu_threads,
            "kind": "port",
            "sample": f"{steps} steps of each of the {len(cavs)} cavities, one after another "
                      f"(FP64 CPU oracle, {args.cpu_threads} OpenMP threads, {t:.1f} s)"}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def run_steady3d(args, cfg):
    """The 3D cube to the reference's stopping rule (both residuals <= 1e-8,
    checking every 2000 steps): FP64 from rest, and FP32 shifted (the
    z-march) to a residual of 1e-4, then FP64 from the promoted state
    (lbm_copy_state).  The value is the FP64 wall time from lbm_create to
    the converged check."""
    import lbmhip
    n, Ra = cfg["n"], cfg["Ra"]
    t0 = time.perf_counter()
    s = lbmhip.Solver(3, n, Ra, precision=lbmhip.PREC_FP64, epsU=1e-8, epsT=1e-8,
                      itc_max=4000000, device=0)
    s.init_rest()
    st, itc, eu, et = s.run()
    wall = time.perf_counter() - t0
    nu = s.nusselt()
    s.close()
    fp64 = {"status": st, "itc": itc, "errU": eu, "errT": et, "wall_s": round(wall, 3),
            "nu_hot": nu[0], "nu_cold": nu[1]}
    t0 = time.perf_counter()
    a = lbmhip.Solver(3, n, Ra, epsU=1e-4, epsT=1e-4, itc_max=4000000, device=0)
    a.init_rest()
    st32, itc32, _, _ = a.run()
    b = lbmhip.Solver(3, n, Ra, precision=lbmhip.PREC_FP64, epsU=1e-8, epsT=1e-8,
                      itc_max=4000000, device=0)
    b.copy_state_from(a)
    a.close()
    stm, itcm, eum, etm = b.run()
    wallm = time.perf_counter() - t0
    num = b.nusselt()
    b.close()
    mixed = {"fp32_to": 1e-4, "fp32_status": st32, "fp32_itc": itc32, "status": stm, "itc": itcm,
             "errU": eum, "errT": etm, "wall_s": round(wallm, 3), "nu_hot": num[0],
             "nu_rel_err_vs_fp64": abs(num[0] / nu[0] - 1.0)}
    nodes = n ** 3
    out = {
        "metric": "time to steady state (wall, lbm_create to the converged check), 1 GPU",
        "value": fp64["wall_s"], "unit": "s", "n_gpus": 1, "steps": itc, "warmup": 0,
        "ms_per_step": round(1e3 * wall / itc, 5), "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: reference initial condition (rest, linear T)",
        "config": {"workload": cfg["label"], "nodes": nodes, "grid": "erf A=3.2", "Ra": Ra,
                   "Pr": 0.71, "eps": 1e-8, "check_every": 2000, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(432 * nodes * itc / wall / 1e9, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(432 * nodes * itc / wall / 1e9 / HBM_PEAK_GBS, 5),
                     "traffic": None,
                     "note": "262 144 nodes: check intervals replay as HIP graphs; the whole "
                             "run, checks included"},
        "check": {"fp64": fp64, "fp32_then_fp64": mixed},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_steady(args, cfg):
    """The reference's own case run to steady state: 2D 129^2, Ra = 1e4,
    Pr = 0.71, from the reference IC to both residuals <= 1e-8 in FP64,
    checking every 2000 steps, then the FP32 shifted path for the same number
    of steps with its Nu against the FP64 one; FP32 through the transient
    (to a residual of 1e-4, 1e-5 or 2e-6) and then FP64 from the promoted
    state (lbm_copy_state) to the same rule; and the learned warm start
    against the cold start (both FP64, same rule).  The value is the FP64
    cold wall time from lbm_create to the converged check."""
    import lbmhip
    runs = {}
    itc_max = 2000000
    for name, prec, eps in (("fp64", lbmhip.PREC_FP64, 1e-8),
                            ("fp32_same_steps", lbmhip.PREC_FP32_SHIFTED, 1e-8)):
        t0 = time.perf_counter()
        s = lbmhip.Solver(2, cfg["n"], cfg["Ra"], precision=prec, epsU=eps, epsT=eps,
                          itc_max=itc_max, device=0)
        s.init_rest()
        status, itc, eu, et = s.run()
        wall = time.perf_counter() - t0
        nu = s.nusselt()
        s.close()
        runs[name] = {"eps": eps, "status": status, "itc": itc, "errU": eu, "errT": et,
                      "wall_s": round(wall, 4), "nu_hot": nu[0], "nu_cold": nu[1]}
        # FP32 shifted populations level off near 1e-6 in the velocity residual,
        # so the FP32 leg runs the FP64 run's step count and is judged by its Nu
        itc_max = itc
    r = runs["fp64"]
    f = runs["fp32_same_steps"]
    f["nu_rel_err_vs_fp64"] = max(abs(f["nu_hot"] - r["nu_hot"]) / abs(r["nu_hot"]),
                                  abs(f["nu_cold"] - r["nu_cold"]) / abs(r["nu_cold"]))
    nodes = cfg["n"] ** 2
    out = {
        "metric": "time to steady state (wall, lbm_create to the converged check), 1 GPU",
        "value": r["wall_s"], "unit": "s", "n_gpus": 1, "steps": r["itc"], "warmup": 0,
        "ms_per_step": round(1e3 * r["wall_s"] / r["itc"], 5), "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: reference initial condition (rest, linear T)",
        "config": {"workload": cfg["label"], "nodes": nodes, "grid": "erf A=3.2",
                   "Ra": cfg["Ra"], "Pr": 0.71, "check_every": 2000,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(240 * nodes * r["itc"] / r["wall_s"] / 1e9, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(240 * nodes * r["itc"] / r["wall_s"] / 1e9 / HBM_PEAK_GBS, 5),
                     "traffic": None,
                     "note": "16 641 nodes: each step is one short kernel (its latency, not HBM, "
                             "bounds it); check intervals replay as HIP graphs"},
        "check": runs,
    }
    # FP32 through the transient, then FP64 (lbm_copy_state) to the same rule
    mixed = []
    for switch in (1e-4, 1e-5, 2e-6):
        t0 = time.perf_counter()
        a = lbmhip.Solver(2, cfg["n"], cfg["Ra"], epsU=switch, epsT=switch, itc_max=2000000,
                          device=0)
        a.init_rest()
        st32, itc32, _, _ = a.run()
        b = lbmhip.Solver(2, cfg["n"], cfg["Ra"], precision=lbmhip.PREC_FP64, epsU=1e-8,
                          epsT=1e-8, itc_max=2000000, device=0)
        b.copy_state_from(a)
        a.close()
        st, itc, eu, et = b.run()
        wall = time.perf_counter() - t0
        nu = b.nusselt()
        b.close()
        mixed.append({"fp32_to": switch, "fp32_status": st32, "fp32_itc": itc32, "status": st,
                      "itc": itc, "errU": eu, "errT": et, "wall_s": round(wall, 4),
                      "nu_rel_err_vs_fp64": abs(nu[0] / r["nu_hot"] - 1.0)})
    out["fp32_then_fp64"] = mixed

    # the learned warm start on the same case, FP64 to the same rule
    import surrogate
    wname = args.weights or "warm257w28.npz"
    weights = os.path.join(os.path.dirname(surrogate.WEIGHTS), wname)
    net = surrogate.load_net(weights)
    cav = dict(n=cfg["n"], Ra=cfg["Ra"], Pr=0.71, stretch=3.2, precision=lbmhip.PREC_FP64)
    rows, sm = surrogate.compare(net, [cav], eps=1e-8, itc_max=2000000)
    net.close()
    w = rows[0]
    out["warm_start"] = {
        "weights": "data/" + wname, "precision": "f64", "eps": 1e-8,
        "cold_itc": w["cold_itc"], "warm_itc": w["warm_itc"],
        "cold_s": round(w["cold_s"], 4), "warm_s": round(w["warm_s"], 4),
        "stop_wall_ratio": round(w["cold_s"] / w["warm_s"], 3),
        "nu_cold": w["cold_nu"], "nu_warm": w["warm_nu"], "nu_rel_dev": w.get("nu_rel_dev"),
        "nu_band_itc": [w.get("cold_band_itc"), w.get("warm_band_itc")],
        "nu_band_s": [w.get("cold_band_s"), w.get("warm_band_s")],
        "nu_band_wall_ratio": sm.get("band_wall_ratio"), "fallback": w.get("fallback"),
        "note": "traced runs (a check and a Nusselt number every 2000 steps, on the host); "
                "the warm time includes inference, resampling and lbm_init_fields"}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import lbm_oracle
        lbm_oracle.set_threads(args.cpu_threads)
        steps = args.cpu_steps or 20000
        o = lbm_oracle.Oracle(2, cfg["n"], cfg["Ra"])
        o.init_rest()
        o.step(1)
        t = time.perf_counter()
        o.step(steps)
        dt = time.perf_counter() - t
        o.close()
        out["cpu_baseline"] = {
            "value": round(dt / steps * r["itc"], 2), "unit": "s", "cores": args.cpu_threads,
            "kind": "port",
            "sample": f"{steps} steps of the same cavity (FP64 CPU oracle, {args.cpu_threads} "
                      f"OpenMP threads, {dt:.1f} s), scaled to the {r['itc']} steps to steady "
                      f"state"}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def _training_data(record):
    """One line on the steady fields a network was trained on, from the
    record surrogate.save_weights() stored with it."""
    parts = record.get("data")
    if not parts:
        return f"{record.get('count')} seeded {record.get('n', 64)}^2 cavities"
    n = sum(p.get("kept", 0) for p in parts)
    sizes = sorted({p.get("n") for p in parts})
    text = f"{n} seeded " + "/".join(f"{k}^2" for k in sizes) + " cavities"
    if record.get("init"):
        text += f", starting from {record['init']}"
    return text


def run_warm(args, cfg):
    """The learned warm start against the cold start on held-out cavities:
    each cavity runs from rest and from the trained network's prediction to
    the same stopping rule; the value is the ratio of the wall times until
    the hot-wall Nusselt number stays within 0.1 % of the cold run's
    converged value (network inference and initialisation included).
    Cavities on grids other than the network's own get its output
    resampled onto their nodes; the 3D cube gets the 2D prediction of its
    (y, z) cross-section extruded along x.  The cube's rule is eps = 2e-6:
    at 128^3 the FP32 velocity residual wanders between 1e-6 and 2e-6 for
    hundreds of thousands of steps before it settles below 1e-6."""
    import numpy as np
    import lbmhip
    import surrogate
    wname = args.weights or cfg["weights"]
    weights = os.path.join(os.path.dirname(surrogate.WEIGHTS), wname)
    net = surrogate.load_net(weights)
    net_grid, net_width, net_params = net.grid, net.width, net.param_count
    with np.load(weights, allow_pickle=False) as z:
        record = json.loads(str(z["record"]))
    dim, n = cfg["dim"], cfg["n"]
    if "cube" in cfg:
        cavs = [dict(dim=3, n=n, **cfg["cube"])]
    else:
        cavs = surrogate.sample_cavities(cfg["held_out"], cfg["seed"], n=n)
    t0 = time.perf_counter()
    eps = cfg.get("eps", surrogate.EPS)
    traces = [] if args.save_traces else None
    rows, s = surrogate.compare(net, cavs, eps=eps, itc_max=cfg["itc_max"],
                                log=lambda r: print(r, file=sys.stderr, flush=True),
                                traces=traces)
    if traces is not None:
        np.savez_compressed(args.save_traces,
                            **{f"{k}{i}": t[j] for i, t in enumerate(traces)
                               for j, k in enumerate(("cold", "warm"))},
                            cold_nu=np.array([r["cold_nu"] for r in rows]),
                            cold_status=np.array([r["cold_status"] for r in rows]))
    wall = time.perf_counter() - t0
    net.close()
    steps = sum(r["cold_itc"] + r["warm_itc"] for r in rows)
    probe = lbmhip.Solver(dim, n, 1e5, stretch=2.0)
    bpu, nodes = probe.bytes_per_update(), n ** dim
    probe.close()
    solver_s = s["cold_s"] + s["warm_s"]
    achieved = bpu * nodes * steps / solver_s / 1e9
    out = {
        "metric": "wall-time speedup of the learned warm start over the cold start to the "
                  "steady hot-wall Nu (within 0.1 %), 1 GPU",
        "value": None if s.get("band_wall_ratio") is None else round(s["band_wall_ratio"], 3),
        "unit": "x (cold time / warm time)",
        "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(1e3 * solver_s / steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded held-out cavity parameters; network trained on the "
                "solver's own steady fields ({})".format(_training_data(record)),
        "config": {"workload": cfg["label"], "held_out": len(cavs), "seed": cfg.get("seed"),
                   "eps": eps, "check_every": surrogate.CHECK_EVERY,
                   "itc_max": cfg["itc_max"],
                   "network": {"weights": "data/" + wname, "grid": net_grid,
                               "width": net_width, "params": net_params,
                               "train_samples": record.get("train_samples"),
                               "epochs": record.get("epochs"),
                               "val_rms": record.get("final", {}).get("val_rms")},
                   "parallelism": "single GPU, one cavity at a time"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "note": "whole solver runs (checks included); grids up to 257^2 are "
                             "launch-bound, not HBM-bound"},
        "wall_s": round(wall, 3),
        "check": {k: s.get(k) for k in ("judged", "fallbacks", "nu_rel_dev_max",
                                        "band_wall_ratio", "band_itc_ratio",
                                        "band_wall_ratio_median", "stop_wall_ratio",
                                        "stop_itc_ratio", "stop_wall_ratio_median", "cold_s",
                                        "warm_s", "cold_band_s", "warm_band_s",
                                        "settle_cavities", "settle_by_nu", "settle_wall_ratio",
                                        "settle_itc_ratio", "settle_wall_ratio_median",
                                        "settle_nu_dev_max", "settle_nu_dev_over_1e3",
                                        "cold_settle_s", "warm_settle_s")},
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import lbm_oracle
        lbm_oracle.set_threads(args.cpu_threads)
        steps_cpu = args.cpu_steps or int(min(20000, max(50, 2e8 / nodes)))
        c = cavs[0]
        o = lbm_oracle.Oracle(dim, n, c["Ra"], Pr=c["Pr"], A=c["stretch"])
        o.init_rest()
        o.step(1)
        t = time.perf_counter()
        o.step(steps_cpu)
        dt = time.perf_counter() - t
        o.close()
        cold_itc = rows[0]["cold_itc"]
        out["cpu_baseline"] = {
            "value": round(nodes * steps_cpu / dt / 1e6, 2), "unit": "MLUPS", "cores": args.cpu_threads,
            "kind": "port",
            "sample": f"{steps_cpu} steps of the first held-out cavity, {n}^{dim} (FP64 CPU "
                      f"oracle, {args.cpu_threads} OpenMP threads, {dt:.1f} s); its cold start "
                      f"to the stopping rule ({cold_itc} steps) would take about "
                      f"{dt * cold_itc / steps_cpu:.0f} s on these cores"}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def copy_ceiling(args):
    """The fastest a plain copy moves the 256^3 FP32 step's bytes (26
    population planes of 256^3 floats read and 26 written, 3.49 GB) on this
    GPU: torch's device copy, timed with HIP events over args.steps copies."""
    import torch
    n = 26 * 256 ** 3
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(max(1, args.warmup)):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(args.steps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    moved = 2.0 * n * 4
    print(json.dumps({"metric": "plain device copy of the 256^3 step's bytes", "value":
                      round(moved / ms / 1e6, 1), "unit": "GB/s", "ms_per_copy": round(ms, 5),
                      "bytes_moved": moved, "frac_of_8TBs": round(moved / ms / 1e6 / HBM_PEAK_GBS, 4),
                      "steps": args.steps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=DEFAULT_STEPS)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="3d256")
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="oracle steps for the CPU baseline (0: automatic)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--weights", default="",
                    help="warm-start workloads: network weights under data/ instead of the "
                         "workload's own")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--save-traces", default="",
                    help="warm-start workloads: write every run's (itc, wall, Nu) checks to this .npz")
    ap.add_argument("--no-fp64-check", action="store_true",
                    help="skip the FP64 GPU run that the FP32 fields and Nu are compared with")
    ap.add_argument("--launch", default="",
                    help="fixed launch shape bx,by,remap,occupancy,variant (implies no autotune)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (5 times), the "
                         "reference kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE")
    ap.add_argument("--copy-ceiling", action="store_true",
                    help="time a plain device copy of the 256^3 step's traffic (1.745 GB "
                         "read and written): the streaming ceiling the roofline fraction "
                         "is read against")
    args = ap.parse_args()
    if args.copy_ceiling:
        copy_ceiling(args)
        return
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": a.numel() * 4, "launches": 5}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs one cavity on one GPU; multi-GPU runs "
              "(--gpus > 1) are not supported", file=sys.stderr)
        sys.exit(2)

    import lbmhip
    cfg = WORKLOADS[args.workload]
    if "held_out" in cfg:
        run_warm(args, cfg)
        return
    if cfg.get("steady"):
        (run_steady3d if cfg["dim"] == 3 else run_steady)(args, cfg)
        return
    if "count" in cfg:
        if args.precision != "fp32":
            print("bench.py: the batch workloads run FP32", file=sys.stderr)
            sys.exit(2)
        run_batch(args, cfg)
        return
    prec = lbmhip.PREC_FP64 if args.precision == "fp64" else lbmhip.PREC_FP32_SHIFTED
    s = lbmhip.Solver(cfg["dim"], cfg["n"], cfg["Ra"], precision=prec, device=0)
    s.init_rest()
    if args.launch:
        vals = [int(v) for v in args.launch.split(",")] + [0]
        s.set_launch(vals[0], vals[1], remap=vals[2], occupancy=vals[3], variant=vals[4])
    elif not args.no_autotune:
        s.autotune(10)   # restores the state it started from
    if args.warmup:
        s.step(args.warmup)
    launch = s.get_launch()
    s.synchronize()
    t0 = time.perf_counter()
    ms_events = s.time_steps(args.steps)   # HIP events on the solver's stream
    s.synchronize()
    wall = time.perf_counter() - t0
    nodes = cfg["n"] ** cfg["dim"]
    sec = ms_events / 1e3
    mlups = nodes * args.steps / sec / 1e6
    bpu = s.bytes_per_update()
    launch_s = sec / args.steps
    achieved = bpu * nodes / launch_s / 1e9
    eu, et = s.check(allow_diverged=True)
    nu, nu_steps = s.nusselt(), s.itc
    vs_fp64 = None
    if prec == lbmhip.PREC_FP32_SHIFTED and not args.no_fp64_check:
        # the FP32 run continued (untimed) to the case's check_steps, and the
        # FP64 GPU path after the same number of steps from the same initial
        # condition: field and Nusselt errors of the FP32 run
        import numpy as np
        if s.itc < cfg.get("check_steps", 0):
            s.step(cfg["check_steps"] - s.itc)
            nu, nu_steps = s.nusselt(), s.itc
        itc = s.itc
        f32 = s.fields()
        s.close()
        d = lbmhip.Solver(cfg["dim"], cfg["n"], cfg["Ra"], precision=lbmhip.PREC_FP64, device=0)
        d.init_rest()
        d.step(itc)
        f64, nu64 = d.fields(), d.nusselt()
        d.close()
        vscale = max(np.abs(f64[k]).max() for k in range(1, 1 + cfg["dim"]))
        vs_fp64 = {"steps": itc,
                   "velocity_max_abs_err_rel_to_max_speed": float(max(
                       np.abs(f32[k] - f64[k]).max() for k in range(1, 1 + cfg["dim"])) / vscale),
                   "T_max_abs_err": float(np.abs(f32[4] - f64[4]).max()),
                   "nu_hot_rel_err": abs(nu[0] - nu64[0]) / abs(nu64[0]),
                   "nu_cold_rel_err": abs(nu[1] - nu64[1]) / abs(nu64[1])}
    else:
        s.close()

    traffic = traffic_from_profiles(args.workload, args.precision, launch)
    out = {
        "metric": "MLUPS + % HBM peak (roofline.frac), fused MRT-LBM step, 1 GPU",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_events / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == lbmhip.PREC_FP32_SHIFTED else "f64",
        "data": "synthetic: reference initial condition (rest, linear T)",
        "config": {"workload": cfg["label"], "nodes": nodes, "grid": "erf A=3.2",
                   "Ra": cfg["Ra"], "Pr": 0.71, "precision": args.precision,
                   "launch": launch,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "algorithmic_bytes_per_launch": bpu * nodes},
        "wall_s": round(wall, 4),
        "check": {"errU": eu, "errT": et, "nu_steps": nu_steps, "nu_hot": nu[0], "nu_cold": nu[1],
                  "vs_fp64_gpu": vs_fp64},
    }
    if not args.no_cpu_baseline:
        # about 20-30 s of CPU work: 20 steps of 256^3, 300 of 1025^2
        steps = args.cpu_steps or (20 if cfg["dim"] == 3 else 300)
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_threads, steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
