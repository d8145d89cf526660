"""Learned warm start of the steady 2D cavity: data from the batched solver,
training of the HIP network (lbmnn.Net), and the warm-versus-cold
comparison.

The reference always starts from rest with the linear conduction profile
(natural-convection_2dim_mono.f90:199-240) and marches to its stopping rule
(errU, errT < eps on snapshots `check_every` steps apart,
natural-convection_2dim_mono.f90:202-206,327).  Here the network predicts
the steady (u, v, T) of a cavity from (Ra, Pr, grid) and lbm_init_fields()
starts the solver from the equilibrium of that prediction; the solver then
runs to the same stopping rule, so it remains the judge of the result and
the cold start stays the fallback.

Every step runs on the GPU: the batched solver generates the steady fields
(one launch per time step for all cavities), the network trains with its
own forward, backward and Adam kernels, and inference is one forward call.

Grids other than the network's G x G (2D cavities of other sizes, the 3D
cube) get the network's output on G erf nodes resampled onto their nodes,
and a cube the resulting plane extruded along its adiabatic x axis
(warm_fields).  Training data from larger grids is sampled at the network's
nodes the same way (generate).

Command line (prints one JSON record per stage):
    python surrogate.py generate --out data.npz --n 257 --grid 128 [--count 256 ...]
    python surrogate.py train --out weights.npz [--count 2048 ...]
    python surrogate.py train --out w.npz --data data.npz [--width 28 | --init warm64.npz]
    python surrogate.py compare --weights weights.npz [--n 257 --count 100 ...]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)

import lbmhip   # noqa: E402
import lbmnn    # noqa: E402

GRID = 64
WIDTH = 16
LOG_RA = (3.0, 6.0)
PR = (0.7, 7.0)
STRETCH = (1.0, 2.5)
EPS = 1e-6
CHECK_EVERY = 2000
# The Nusselt stopping rule (settle_estimate): SETTLE_CHECKS hot-wall Nu
# values, SETTLE_SPACING checks apart, ending at the latest check, bound the
# drift still to come by SETTLE_TOL (relative), on SETTLE_HOLD checks in a
# row.  Chosen on the traces of 108 held-out 257^2 cavities and the 128^3
# cube, warm and cold: it never stopped more than 1.9e-4 from the converged
# Nu.  The hold matters most: with a hold of two checks, a cold run whose Nu
# flattened for about 20 checks between two modes was stopped 5.9e-3 away.
SETTLE_TOL = 3e-5
SETTLE_CHECKS = 4
SETTLE_SPACING = 8
SETTLE_SPREAD = 1.5
SETTLE_HOLD = 8
WEIGHTS = os.path.join(_HERE, "data", "warm64.npz")
# A cube larger than this starts from a coarse cube of this edge, itself
# started from the extruded prediction and run to COARSE_EPS (warm_fields).
COARSE_3D = 64
COARSE_EPS = 1e-5
# Targets in data files: int16 multiples of 2^-13 (|y| < 4, rounding <= 6.1e-5,
# far below the network's error).
TARGET_STEP = 2.0 ** -13


def sample_cavities(count, seed, n=GRID, log_ra=LOG_RA, pr=PR, stretch=STRETCH):
    """Seeded cavity parameters: log10 Ra, Pr and the erf stretch factor
    uniform in their ranges, all on an n x n grid."""
    rng = np.random.default_rng(seed)
    lr = rng.uniform(*log_ra, count)
    p = rng.uniform(*pr, count)
    a = rng.uniform(*stretch, count)
    return [dict(n=n, Ra=float(10.0 ** lr[i]), Pr=float(p[i]), stretch=float(a[i]))
            for i in range(count)]


def _batch(cavities, eps, itc_max):
    return lbmhip.Batch([dict(c, epsU=eps, epsT=eps, itc_max=itc_max) for c in cavities],
                        check_every=CHECK_EVERY)


def net_inputs(cavity, s, G=GRID):
    """The network input [4][G][G] for cavity (Ra, Pr, stretch) on Solver s,
    and the nodes x/H it is given on: None when those are the solver's own
    nodes (a 2D grid of G nodes), else the G nodes of the erf grid with the
    cavity's stretch clipped to the trained range, between which the
    network's output is resampled onto the solver's grid (lbmnn.resample)."""
    if s.dim == 2 and s.n == G:
        return lbmnn.inputs(cavity["Ra"], cavity["Pr"], s.grid(), s.info()["H"]), None
    xc, Hc = lbmnn.erf_nodes(G, min(max(cavity["stretch"], STRETCH[0]), STRETCH[1]))
    return lbmnn.inputs(cavity["Ra"], cavity["Pr"], xc, Hc), xc / Hc


def network_inputs(batch, cavities, G=GRID):
    """x[b][4][G][G] of the cavities of a Batch and their node lists (see
    net_inputs)."""
    got = [net_inputs(c, batch[i], G) for i, c in enumerate(cavities)]
    return np.stack([g[0] for g in got]), [g[1] for g in got]


def inputs_of(params, G):
    """Network inputs x[k][4][G][G] of cavities given as rows (Ra, Pr,
    stretch) on grids other than G x G (net_inputs' erf nodes of G)."""
    out = np.zeros((len(params), 4, G, G), np.float32)
    for i, (ra, pr, a) in enumerate(np.asarray(params, np.float64)):
        xc, Hc = lbmnn.erf_nodes(G, min(max(a, STRETCH[0]), STRETCH[1]))
        out[i] = lbmnn.inputs(ra, pr, xc, Hc)
    return out


def solve(cavities, eps=EPS, itc_max=600000, net=None, keep_fields=False, G=GRID):
    """Runs every cavity to the stopping rule, from rest (net None) or from
    the network's prediction.  Returns a dict of per-cavity arrays (itc,
    status, errU, errT, Nu on the hot wall) and wall times in seconds:
    `predict` (inputs, forward, lbm_init_fields), `run` (the time loop).
    G is the grid of the network inputs and kept fields (net.grid when a
    network is given)."""
    if net is not None:
        G = net.grid
    b = _batch(cavities, eps, itc_max)
    t0 = time.perf_counter()
    x, nodes = network_inputs(b, cavities, G)
    if net is None:
        b.init_rest()
    else:
        y = predict(net, x)
        for i in range(len(cavities)):
            yi = y[i] if nodes[i] is None else lbmnn.resample(
                y[i], nodes[i], b[i].grid() / b[i].info()["H"])
            u, v, T = lbmnn.fields(yi, lbmnn.velocity_scale(b[i].info()))
            b[i].init_fields(u, v, None, T)
    b[0].synchronize()
    t1 = time.perf_counter()
    # the batch's launch (one step or two steps per launch, or resident
    # cavities for edges up to 70), chosen by timing; the results do not
    # depend on it
    b.autotune(2)
    itc, eu, et, st = b.run()
    t2 = time.perf_counter()
    out = dict(itc=itc, status=st, errU=eu, errT=et, nu=np.array(b.nusselt()[0]),
               predict_s=t1 - t0, run_s=t2 - t1, x=x)
    if keep_fields:
        t = np.zeros((len(cavities), 3, G, G), dtype=np.float32)
        for i in range(len(cavities)):
            _, u, v, _, T = b[i].fields()
            ti = lbmnn.targets(u, v, T, lbmnn.velocity_scale(b[i].info()))
            # fields of a larger grid are sampled at the network's nodes
            t[i] = ti if nodes[i] is None else lbmnn.resample(
                ti, b[i].grid() / b[i].info()["H"], nodes[i])
        out["y"] = t
    b.close()
    return out


def predict(net, x):
    """Network output for any number of samples, in chunks of batch_max."""
    return np.concatenate([net.forward(x[i:i + net.batch_max])
                           for i in range(0, len(x), net.batch_max)])


def generate(count, seed, eps=EPS, itc_max=600000, per_batch=256, keep_tol=1e-4, n=GRID,
             stretch=STRETCH, G=GRID, log=print):
    """Steady fields of `count` seeded cavities on n x n grids, as network
    inputs and targets on the network's G x G nodes (sampled there from a
    larger grid, see net_inputs).  A cavity is kept if it reached the
    stopping rule, or ran to itc_max with errU and errT below keep_tol (a
    slowly settling cavity, already steady to that tolerance); diverged
    cavities are dropped.  Returns (x, y, record); record["cavities"] lists
    (Ra, Pr, stretch) of the kept cavities."""
    cavs = sample_cavities(count, seed, n=n, stretch=stretch)
    xs, ys, keep, itcs = [], [], [], []
    t0 = time.perf_counter()
    for s in range(0, count, per_batch):
        part = cavs[s:s + per_batch]
        r = solve(part, eps, itc_max, keep_fields=True, G=G)
        ok = (r["status"] == lbmhip.LBM_OK) | ((r["status"] == lbmhip.LBM_ERR_NOT_CONVERGED) &
                                              (r["errU"] <= keep_tol) & (r["errT"] <= keep_tol))
        xs.append(r["x"][ok])
        ys.append(r["y"][ok])
        keep += [s + i for i in np.nonzero(ok)[0]]
        itcs.append(r["itc"])
        log(json.dumps({"stage": "generate", "done": s + len(part), "of": count,
                        "converged": int(ok.sum()), "max_itc": int(r["itc"].max()),
                        "run_s": round(r["run_s"], 2)}))
    itcs = np.concatenate(itcs)
    rec = dict(count=count, seed=seed, kept=len(keep), eps=eps, itc_max=itc_max, keep_tol=keep_tol,
               itc_median=float(np.median(itcs)), generate_seconds=time.perf_counter() - t0,
               cavities=[[cavs[i]["Ra"], cavs[i]["Pr"], cavs[i]["stretch"]] for i in keep])
    return np.concatenate(xs), np.concatenate(ys), rec


def train(net, x, y, epochs, batch=32, lr=2e-3, lr_min=2e-5, seed=0, val=None, log=print,
          time_budget=None):
    """Adam over shuffled mini-batches with a cosine learning-rate decay;
    returns the per-epoch mean training loss (and validation loss).  With
    a time_budget (seconds), the first epoch's duration caps the number of
    epochs so that training ends within the budget; the decay then spans
    the capped count."""
    rng = np.random.default_rng(seed)
    hist = []
    per_epoch = (len(x) + batch - 1) // batch
    steps = epochs * per_epoch
    k = 0
    t0 = time.perf_counter()
    e = 0
    while e < epochs:
        order = rng.permutation(len(x))
        tot = 0.0
        for s in range(0, len(x), batch):
            idx = np.sort(order[s:s + batch])
            rate = lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * k / steps))
            tot += net.train_step(x[idx], y[idx], rate) * len(idx)
            k += 1
        if e == 0 and time_budget:
            fit = max(2, int(time_budget / (time.perf_counter() - t0)))
            if fit < epochs:
                epochs, steps = fit, fit * per_epoch
                log(json.dumps(dict(stage="train", epochs_capped_to=fit)))
        entry = {"epoch": e + 1, "loss": tot / len(x)}
        if val is not None and ((e + 1) % 10 == 0 or e + 1 == epochs):
            d = predict(net, val[0]) - val[1]
            entry["val_loss"] = float(np.mean(d ** 2))
            entry["val_rms"] = np.sqrt(np.mean(d ** 2, axis=(0, 2, 3))).round(5).tolist()
        hist.append(entry)
        if (e + 1) % 10 == 0 or e + 1 == epochs:
            log(json.dumps(dict(stage="train", seconds=round(time.perf_counter() - t0, 1),
                                **entry)))
        e += 1
    return hist


def save_weights(path, net, record):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=net.get_params(), grid=np.int32(net.grid),
             width=np.int32(net.width), record=np.array(json.dumps(record)))


def load_net(path=WEIGHTS, batch_max=32, device=0):
    """A network with the parameters saved by save_weights() (files without
    a width hold the base width 16)."""
    with np.load(path, allow_pickle=False) as z:
        width = int(z["width"]) if "width" in z.files else WIDTH
        net = lbmnn.Net(grid=int(z["grid"]), batch_max=batch_max, device=device, width=width)
        net.set_params(z["params"])
    return net


def warm_fields(net, cavity, s, coarse=COARSE_3D):
    """(u, v, w, T) for lbm_init_fields() of the Solver s from the network's
    prediction for cavity (Ra, Pr, stretch), on the nodes net_inputs()
    names and resampled onto the solver's grid where those differ; a 3D cube
    gets that plane extruded along x (lbmnn.extrude), the 2D cavity being
    the cube's (y, z) cross-section.  A cube with more than `coarse` nodes
    per edge gets, instead, the fields of a coarse cube of `coarse` nodes
    (the same cavity, FP32, the z-march) started from the extruded
    prediction and run to errU, errT <= COARSE_EPS (at most 400 000 steps),
    resampled onto its nodes (lbmnn.resample3): the extruded plane has the
    2D flow's structure, and the coarse run gives it the third dimension
    (the no-slip x walls)."""
    if s.dim == 3 and coarse and s.n > coarse:
        c = lbmhip.Solver(3, coarse, cavity["Ra"], Pr=cavity["Pr"], stretch=cavity["stretch"],
                          epsU=COARSE_EPS, epsT=COARSE_EPS, check_every=CHECK_EVERY,
                          itc_max=400000)
        c.init_fields(*warm_fields(net, dict(cavity, n=coarse), c, coarse=None))
        c.run()
        _, u, v, w, T = c.fields()
        xc = c.grid() / c.info()["H"]
        c.close()
        return lbmnn.resample3(u, v, w, T, xc, s.grid() / s.info()["H"])
    info = s.info()
    x, nodes = net_inputs(cavity, s, net.grid)
    y = net.forward(x[None])[0]
    if nodes is not None:
        y = lbmnn.resample(y, nodes, s.grid() / info["H"])
    u, v, T = lbmnn.fields(y, lbmnn.velocity_scale(info))
    if s.dim == 2:
        return u, v, None, T
    return lbmnn.extrude(u, v, T, s.n)


def settle_estimate(nus):
    """Bound on the relative drift |Nu_inf - Nu_k| / |Nu_k| still to come,
    from the Nu values of a run's checks so far (equally spaced in steps),
    or inf.  It reads SETTLE_CHECKS values SETTLE_SPACING checks apart,
    ending at the latest.

    Near a steady state the slowest mode dominates: successive differences
    d_i of the samples keep one sign and shrink by a common ratio
    rho = d_(i+1) / d_i < 1, and the drift still to come is the geometric
    tail |d_last| * rho / (1 - rho).  The bound uses the largest ratio of
    the window, and only when every ratio is positive and they agree within
    a factor SETTLE_SPREAD.  Otherwise the run is still in its transient,
    at a turning point or in an oscillation, and the bound is inf."""
    v = np.asarray(nus, dtype=np.float64)[::-1][::SETTLE_SPACING][:SETTLE_CHECKS][::-1]
    if v.size < SETTLE_CHECKS or not np.all(np.isfinite(v)) or v[-1] == 0.0:
        return np.inf
    d = np.diff(v)
    if np.all(d == 0.0):
        return 0.0
    if np.any(d[:-1] == 0.0):
        return np.inf
    rho = d[1:] / d[:-1]
    if np.any(rho <= 0.0) or rho.max() > SETTLE_SPREAD * rho.min():
        return np.inf
    r = float(rho.max())
    if r >= 1.0:
        return np.inf
    return float(abs(d[-1]) * r / (1.0 - r) / abs(v[-1]))


# checks settle_entry needs to decide at the latest one
SETTLE_SPAN = (SETTLE_CHECKS - 1) * SETTLE_SPACING + SETTLE_HOLD


def settle_entry(trace, tol=SETTLE_TOL):
    """(itc, wall, Nu) of the first check at which the Nusselt stopping rule
    holds: settle_estimate <= tol on SETTLE_HOLD checks in a row.  The rule
    reads only the checks up to the one it stops at, so a run stopped by it
    ends there; None if it never holds."""
    held = 0
    for k in range(len(trace)):
        held = held + 1 if settle_estimate(trace[:k + 1, 2]) <= tol else 0
        if held >= SETTLE_HOLD:
            return trace[k, 0], trace[k, 1], trace[k, 2]
    return None


def traced_run(cavity, net=None, eps=EPS, itc_max=600000, log=None, stop="residual"):
    """One cavity to the stopping rule (the loop of lbm_run(): check every
    CHECK_EVERY steps, stop when errU and errT are both <= eps), from rest
    or from the network's prediction (FP32 shifted storage unless the
    cavity names a `precision`), recording after every check the step
    count, the wall time since the start (inference and lbm_init_fields
    included) and the hot-wall Nusselt number; log (if given) receives a
    line every 50 checks.  stop="nu" stops instead at the Nusselt rule
    (settle_entry), or at the residual rule if that comes first."""
    c = dict(cavity)
    s = lbmhip.Solver(c.pop("dim", 2), c.pop("n"), c.pop("Ra"), Pr=c.pop("Pr"),
                      stretch=c.pop("stretch"), epsU=eps, epsT=eps, check_every=CHECK_EVERY,
                      itc_max=itc_max,
                      precision=c.pop("precision", lbmhip.PREC_FP32_SHIFTED))
    if s.dim == 3:
        s.set_launch(64, 4, remap=2, variant=1)   # the LDS z-march
    t0 = time.perf_counter()
    if net is None:
        s.init_rest()
    else:
        s.init_fields(*warm_fields(net, cavity, s))
    status, trace = lbmhip.LBM_ERR_NOT_CONVERGED, []
    while s.itc < itc_max:
        s.step(min(CHECK_EVERY, itc_max - s.itc))
        eu, et = s.check(allow_diverged=True)
        if not (np.isfinite(eu) and np.isfinite(et)):
            trace.append((s.itc, time.perf_counter() - t0, np.nan))
            status = lbmhip.LBM_ERR_DIVERGED
            break
        nu = s.nusselt()[0]
        trace.append((s.itc, time.perf_counter() - t0, nu))
        if log and len(trace) % 50 == 0:
            log(json.dumps(dict(stage="run", start="cold" if net is None else "warm",
                                itc=s.itc, errU=eu, errT=et, nu=nu)))
        if eu <= eps and et <= eps:
            status = lbmhip.LBM_OK
            break
        if stop == "nu" and settle_entry(np.array(trace[-SETTLE_SPAN:])):
            status = lbmhip.LBM_OK
            break
    s.close()
    return dict(status=status, itc=trace[-1][0], wall=trace[-1][1], nu=trace[-1][2],
                trace=np.array(trace))


def band_entry(trace, nu_ref, tol=1e-3):
    """(itc, wall) of the first check from which on every recorded Nusselt
    number stays within tol (relative) of nu_ref; None if the last one is
    outside."""
    dev = np.abs(trace[:, 2] / nu_ref - 1.0)
    if not np.isfinite(dev[-1]) or dev[-1] > tol:
        return None
    outside = np.nonzero(~(dev <= tol))[0]
    k = 0 if outside.size == 0 else outside[-1] + 1
    return trace[k, 0], trace[k, 1]


def compare(net, cavities, eps=EPS, itc_max=600000, tol=1e-3, log=None, traces=None):
    """Cold and warm start of each cavity to the same stopping rule.

    The cold run is the judge: only cavities whose cold run reaches the
    stopping rule count.  A warm run that diverges or stops early falls
    back to the cold start, and its time is charged in full (warm attempt
    plus cold run).  Two times are compared: to the stopping rule, and to
    the check from which on the hot-wall Nusselt number stays within tol
    of the cold run's converged value.  A third time needs no hindsight:
    the first check at which the Nusselt stopping rule (settle_entry) holds,
    or the residual rule's stop if that comes first, read from the same
    traces (the rule uses only earlier checks), with the Nu there compared
    against the cold run's converged value."""
    net.forward(np.zeros((1, 4, net.grid, net.grid), np.float32))   # load the kernels once
    rows = []
    for i, cav in enumerate(cavities):
        cold = traced_run(cav, None, eps, itc_max, log)
        warm = traced_run(cav, net, eps, itc_max, log)
        if traces is not None:   # (itc, wall, Nu) of every check, cold and warm
            traces.append((cold["trace"], warm["trace"]))
        row = dict(Ra=cav["Ra"], Pr=cav["Pr"], stretch=cav["stretch"],
                   cold_status=cold["status"], warm_status=warm["status"],
                   cold_itc=int(cold["itc"]), warm_itc=int(warm["itc"]),
                   cold_s=cold["wall"], warm_s=warm["wall"],
                   cold_nu=cold["nu"], warm_nu=warm["nu"])
        if cold["status"] == lbmhip.LBM_OK:
            cb = band_entry(cold["trace"], cold["nu"], tol)
            wb = band_entry(warm["trace"], cold["nu"], tol)
            fallback = warm["status"] != lbmhip.LBM_OK
            if fallback:
                row["warm_s"] = warm["wall"] + cold["wall"]
                row["warm_itc"] = int(warm["itc"] + cold["itc"])
                wb = (warm["itc"] + cb[0], warm["wall"] + cb[1])
            row.update(fallback=fallback, nu_rel_dev=abs(warm["nu"] / cold["nu"] - 1.0)
                       if not fallback else 0.0,
                       cold_band_itc=int(cb[0]), cold_band_s=cb[1],
                       warm_band_itc=int(wb[0]) if wb else None, warm_band_s=wb[1] if wb else None)
            # a run that the Nusselt rule never stops ends at the residual
            # rule, as traced_run(stop="nu") does
            cs = settle_entry(cold["trace"])
            ws = settle_entry(warm["trace"])
            row["settled_by_nu"] = cs is not None and ws is not None and not fallback
            cs = cs if cs is not None else tuple(cold["trace"][-1])
            ws = ws if ws is not None else tuple(warm["trace"][-1])
            if fallback:
                ws = (warm["itc"] + cs[0], warm["wall"] + cs[1], cs[2])
            for key, st in (("cold", cs), ("warm", ws)):
                row[f"{key}_settle_itc"] = int(st[0])
                row[f"{key}_settle_s"] = float(st[1])
                row[f"{key}_settle_nu_dev"] = float(abs(st[2] / cold["nu"] - 1.0))
        rows.append(row)
        if log:
            log(json.dumps(dict(stage="compare", cavity=i, **row)))
    return rows, summarise(rows)


def summarise(rows):
    """Aggregate ratios over the cavities whose cold run converged."""
    ok = [r for r in rows if r["cold_status"] == lbmhip.LBM_OK]
    if not ok:
        return dict(cavities=len(rows), judged=0)
    banded = [r for r in ok if r["warm_band_s"] is not None]
    settled = [r for r in ok if r.get("cold_settle_s") is not None
               and r.get("warm_settle_s") is not None]
    tot = lambda k, rs: float(sum(r[k] for r in rs))
    devs = [r[k] for r in settled for k in ("cold_settle_nu_dev", "warm_settle_nu_dev")]
    return dict(
        cavities=len(rows), judged=len(ok),
        fallbacks=sum(r["fallback"] for r in ok),
        nu_rel_dev_max=max(r["nu_rel_dev"] for r in ok),
        stop_wall_ratio=tot("cold_s", ok) / tot("warm_s", ok),
        stop_itc_ratio=tot("cold_itc", ok) / tot("warm_itc", ok),
        stop_wall_ratio_median=float(np.median([r["cold_s"] / r["warm_s"] for r in ok])),
        band_cavities=len(banded),
        band_wall_ratio=tot("cold_band_s", banded) / tot("warm_band_s", banded) if banded else None,
        band_itc_ratio=tot("cold_band_itc", banded) / tot("warm_band_itc", banded) if banded else None,
        band_wall_ratio_median=float(np.median([r["cold_band_s"] / r["warm_band_s"]
                                                for r in banded])) if banded else None,
        cold_s=tot("cold_s", ok), warm_s=tot("warm_s", ok),
        cold_band_s=tot("cold_band_s", banded), warm_band_s=tot("warm_band_s", banded),
        settle_cavities=len(settled),
        settle_by_nu=sum(bool(r.get("settled_by_nu")) for r in settled),
        settle_wall_ratio=(tot("cold_settle_s", settled) / tot("warm_settle_s", settled)
                           if settled else None),
        settle_itc_ratio=(tot("cold_settle_itc", settled) / tot("warm_settle_itc", settled)
                          if settled else None),
        settle_wall_ratio_median=float(np.median([r["cold_settle_s"] / r["warm_settle_s"]
                                                  for r in settled])) if settled else None,
        settle_nu_dev_max=max(devs) if devs else None,
        settle_nu_dev_over_1e3=sum(d > 1e-3 for d in devs),
        cold_settle_s=tot("cold_settle_s", settled), warm_settle_s=tot("warm_settle_s", settled))


def save_data(path, y, record):
    """Targets y[k][3][G][G] as int16 multiples of TARGET_STEP and the
    kept cavities' (Ra, Pr, stretch), from which load_data() rebuilds the
    inputs."""
    y = np.asarray(y, np.float64)
    if not np.all(np.abs(y) < 32767 * TARGET_STEP):
        raise ValueError("target outside the int16 range of save_data")
    rec = dict(record)
    cav = np.asarray(rec.pop("cavities"), np.float64)
    assert len(cav) == len(y)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez_compressed(path, yq=np.rint(y / TARGET_STEP).astype(np.int16), cavities=cav,
                        record=np.array(json.dumps(rec)))


def load_data(paths):
    """(x, y, records) of the files save_data() wrote, concatenated."""
    xs, ys, recs = [], [], []
    for p in paths:
        with np.load(p, allow_pickle=False) as z:
            y = z["yq"].astype(np.float32) * np.float32(TARGET_STEP)
            xs.append(inputs_of(z["cavities"], y.shape[-1]))
            ys.append(y)
            recs.append(json.loads(str(z["record"])))
    return np.concatenate(xs), np.concatenate(ys), recs


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate")
    g.add_argument("--out", required=True)
    g.add_argument("--n", type=int, default=GRID)
    g.add_argument("--grid", type=int, default=GRID, help="network grid of the kept fields")
    g.add_argument("--count", type=int, default=256)
    g.add_argument("--seed", type=int, default=20261203)
    g.add_argument("--itc-max", type=int, default=600000)
    g.add_argument("--per-batch", type=int, default=256)
    g.add_argument("--keep-tol", type=float, default=1e-4)
    g.add_argument("--stretch", type=float, nargs=2, default=list(STRETCH),
                   help="range of the erf stretch factor")
    t = sub.add_parser("train")
    t.add_argument("--out", default=WEIGHTS)
    t.add_argument("--data", nargs="*", default=None,
                   help="train on these generate files instead of generating")
    t.add_argument("--init", default=None, help="start from these weights")
    t.add_argument("--width", type=int, default=WIDTH, help="base width of a new network")
    t.add_argument("--n", type=int, default=GRID)
    t.add_argument("--count", type=int, default=2048)
    t.add_argument("--seed", type=int, default=20261203)
    t.add_argument("--epochs", type=int, default=200)
    t.add_argument("--batch", type=int, default=32)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--itc-max", type=int, default=600000)
    t.add_argument("--per-batch", type=int, default=256)
    t.add_argument("--keep-tol", type=float, default=1e-4)
    t.add_argument("--time-budget", type=float, default=None,
                   help="seconds of training; caps the epochs after the first")
    c = sub.add_parser("compare")
    c.add_argument("--weights", default=WEIGHTS)
    c.add_argument("--n", type=int, default=GRID)
    c.add_argument("--count", type=int, default=100)
    c.add_argument("--seed", type=int, default=7)
    c.add_argument("--itc-max", type=int, default=600000)
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    if a.cmd == "generate":
        x, y, rec = generate(a.count, a.seed, itc_max=a.itc_max, per_batch=a.per_batch,
                             keep_tol=a.keep_tol, n=a.n, stretch=tuple(a.stretch), G=a.grid,
                             log=log)
        rec.update(n=a.n, stretch=a.stretch, grid=a.grid)
        save_data(a.out, y, rec)
        log(json.dumps(dict(stage="saved", path=a.out,
                            **{k: v for k, v in rec.items() if k != "cavities"})))
    elif a.cmd == "train":
        if a.data:
            x, y, recs = load_data(a.data)
            rec = dict(data=[dict(path=os.path.basename(p), **r) for p, r in zip(a.data, recs)])
        else:
            x, y, rec = generate(a.count, a.seed, itc_max=a.itc_max, per_batch=a.per_batch,
                                 keep_tol=a.keep_tol, n=a.n, log=log)
            rec.pop("cavities")
            rec.update(n=a.n)
        nval = max(1, len(x) // 16)
        if a.init:
            net = load_net(a.init, batch_max=a.batch)
            rec.update(init=os.path.basename(a.init))
        else:
            net = lbmnn.Net(grid=x.shape[-1], batch_max=a.batch, seed=a.seed, width=a.width)
        t0 = time.perf_counter()
        hist = train(net, x[nval:], y[nval:], a.epochs, a.batch, a.lr,
                     val=(x[:nval], y[:nval]), log=log, time_budget=a.time_budget)
        rec.update(epochs=len(hist), grid=net.grid, width=net.width, params=net.param_count,
                   batch=a.batch, lr=a.lr, train_samples=len(x) - nval,
                   val_samples=nval, final=hist[-1], train_seconds=time.perf_counter() - t0)
        save_weights(a.out, net, rec)
        log(json.dumps(dict(stage="saved", path=a.out, **rec)))
    else:
        net = load_net(a.weights)
        _, summary = compare(net, sample_cavities(a.count, a.seed, n=a.n), itc_max=a.itc_max,
                             log=log)
        log(json.dumps(dict(stage="summary", seed=a.seed, **summary)))


if __name__ == "__main__":
    main()
