#!/usr/bin/env python3
"""Benchmark: pseudo-time iteration throughput (cell-updates/s) on one MI355X.

Workload (default): full physics on a 2048 x 2048 tanh-clustered mesh — QUICK, porous
middle-third block, Boussinesq Ri=2, species slot, Sc=0.7, mu(T) polynomial; FP64 drop-in.
A "step" is one pseudo-step (three SSP-RK3 stages of u, v, p, T, C with their boundary
conditions and the residual maxima) over the whole grid; one cell-update is one pseudo-step
at one p-cell, so value = imax * jmax * steps / seconds.

Prints one JSON line. `roofline` prices the dominant kernel (the fused u/v/p kernel) against
HBM (this path uses no MFMA): achieved = its algorithmic bytes per launch (88 B per cell:
every field it reads once, every output once, times imax*jmax cells) / its mean launch
duration from HIP events recorded around it on the context's stream. `step_roofline` does
the same for the whole pseudo-step (128 B per cell). `cpu_baseline` times the reference's own
Fortran hot path (OpenMp_Multicore's solve_U..solve_C, built from its sources with the
harness in oracle/ref into oracle/_ref/ref_b2048) on a bounded sample of the same grid on
this host; `cpu_port` times the C restatement (the oracle, OpenMP) on the same workload. Where
the reference build is absent, the C restatement is the baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "solver-cfd_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from cfd_amd import lib as L  # noqa: E402
from cfd_amd import cases as C  # noqa: E402
from cfd_amd.params import PREC_FP64, PREC_FP32  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per cell (each field read once, each output written once; 1-D geometry
# is negligible). Whole pseudo-step (FP64, full physics): reads um vm um_tau vm_tau p T T_tau
# C C_tau eps lp (11 x 8) + writes um_n vm_n pn T_n C_n (5 x 8) = 128. Per kernel:
#   u/v/p kernel: reads um vm um_tau vm_tau p T eps lp, writes um_n vm_n pn  -> 88
#   T/C kernel:   reads um_n vm_n T T_tau C C_tau eps lp, writes T_n C_n     -> 80
# (the kernels share um_n/vm_n/T/eps/lp, so the per-kernel sums exceed the step's 128).
BYTES_PER_CELL = {
    ("step", PREC_FP64): 128, ("k1", PREC_FP64): 88, ("k2", PREC_FP64): 80,
    ("step", PREC_FP32): 72,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


REF_BENCH = os.path.join(ROOT, "oracle", "_ref", "ref_b2048")


def cpu_reference(cfg, steps):
    """Time the reference's own Fortran hot path (OpenMp_Multicore solve_U, solve_V, solve_P,
    solve_Z, solve_C per pseudo-step, through the harness's mode-2 loop) on this host, on the
    same 2048 x 2048 mesh, porosity and initial state. Its physics is the reference's own
    (compile-time constants Re 200, Pe 140, Sc 0.7, Da 1e-3, Cf 0.55; its BC set and buoyancy
    form; constant viscosity), which is what the drop-in replaces. Seconds per step come from
    two runs (1 and 1 + steps pseudo-steps) whose difference cancels start-up and file I/O."""
    from oracle import refharness as H
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    p = cfg["params"]
    im, jm = p.imax, p.jmax
    pay = H.geometry_payload(cfg["x"], cfg["y"], cfg["eps"], cfg["lp"]) + H.state_payload(cfg["state"], im, jm)
    per, times = H.time_pseudo_steps(REF_BENCH, [p.dtau, p.dt, p.beta, 0.0, p.tinf, 0.0], pay, 1, 1 + steps, threads)
    return dict(value=im * jm / per, unit="cell-updates/s", cores=threads, kind="reference",
                sample="%d pseudo-steps of the Fortran reference's own hot path (OpenMP, flang -O2) on the "
                       "same %dx%d mesh and state with its own physics: %.1f s against %.1f s for 1 step"
                       % (steps, im, jm, times[1], times[0]))


def cpu_baseline(cfg, seconds):
    """Time the CPU oracle on the same workload: whole pseudo-steps until `seconds` pass."""
    import oracle as O
    O.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = O.Oracle(cfg["params"], cfg["x"], cfg["y"], cfg["eps"], cfg["lp"])
    s, cs = o.make_state(cfg["state"])
    n, t0 = 0, time.perf_counter()
    while True:
        o.pseudo_steps(s, cs, 1, history=False)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds and n >= 1:
            break
    cells = cfg["params"].imax * cfg["params"].jmax
    return dict(value=cells * n / el, unit="cell-updates/s", cores=threads, kind="port",
                sample="%d pseudo-step(s) of the same %dx%d workload, %.1f s, C oracle -O2 OpenMP"
                       % (n, cfg["params"].imax, cfg["params"].jmax, el))


def _build_batch(cs, prec):
    b = L.Batch([c["params"] for c in cs], precision=prec)
    for i, c in enumerate(cs):
        b.set_mesh(i, c["x"], c["y"])
        b.upload(i, "eps", c["eps"])
        b.upload(i, "lp", c["lp"])
        b.set_state(i, c["state"])
    b.apply_bc()
    return b


def run_plain(cases, tol, itc_max, log=log, check_every=500):
    """The plain FP64 path, run for real: one context per case from its initial state, its
    tile shape autotuned (cached per grid class for the process), cfd_run_to_tol to tol or
    itc_max. Returns per case (pseudo-steps, seconds of device time between the run's HIP
    events, converged)."""
    out = []
    for c in cases:
        ctx = L.Context(c["params"], c["x"], c["y"], c["eps"], c["lp"])
        ctx.set_state(c["state"])
        ctx.autotune(steps=10)
        code, it, res = ctx.run_to_tol(tol, itc_max, check_every=check_every, raise_on_fail=False)
        secs = ctx.last_ms_per_step() * it * 1e-3 if code == 0 else ctx.last_ms_per_step() * itc_max * 1e-3
        out.append((int(it), secs, code == 0))
        ctx.close()
        if log:
            log("  plain %s: %d steps, %.2f s, %s" % (c["name"], it, secs, "converged" if code == 0 else "not converged"))
    return out


def bench_batch(args):
    """Time to residual `tol` for every case of the seeded mixed batch (grids {64..512} x
    {64..256}, Re, Da, porosity, Ri drawn per case; steady channel flow from zero fields).

    plain FP64 path: one context per case, cfd_run_to_tol, autotuned tiles; the first case of
    each grid class runs for real to the same tolerance and cap, and the class's other cases
    are priced at its measured time per step and their own FP64-batch step counts (the batch
    and a context step the same arithmetic; the sample reports how many of its step counts
    equal the batch's). batched: FP32 batch (persistent launches, graph replay) down to
    max(tol, 5e-7) (the FP32 storage floor is ~2e-7), then the FP64 batch from the converted
    state down to tol, both timed with events. Cases that do not converge run to itc_max on
    both paths. Steady fields of cases converged on both paths are compared per field."""
    from cfd_amd.params import PREC_FP32 as P32, PREC_FP64 as P64
    from cfd_amd import warmstart as WS
    cs = C.case_mixed_batch(n=args.batch_cases, beta=args.batch_beta)
    cells = np.array([c["params"].imax * c["params"].jmax for c in cs], dtype=np.int64)
    log("batch: %d cases, %d cells, tol %.0e, itc_max %d" % (len(cs), int(cells.sum()), args.batch_tol, args.batch_itc))
    # FP64 batch: the plain path's step counts (and its fields for the comparison)
    b64 = _build_batch(cs, P64)
    tiles64, _ = b64.autotune(steps=20)
    (code, it64, res64, st64), ms64 = b64.run_sliced(args.batch_tol, args.batch_itc, log=log)
    conv = st64 == 1
    for i in np.nonzero((st64 == 2) | (st64 == 3))[0]:
        c = cs[i]
        log("  case %s failed (status %d at step %d): Re %.1f Da %.2e Ri %.2f dt %.3e"
            % (c["name"], st64[i], it64[i], c["re"], c["da"], c["ri"], c["params"].dt))
    log("FP64 batch: %d/%d converged, %.1f s, tiles %s" % (int(conv.sum()), len(cs), ms64 * 1e-3, tiles64))
    # plain path, run for real on a stratified sample: the first case of each grid class,
    # one context each, to the same tolerance and cap; a class's other cases are priced at
    # its sample's measured time per step times their own FP64 step counts
    t_class, sample = {}, []
    for i, c in enumerate(cs):
        k = (c["params"].imax, c["params"].jmax)
        if k not in t_class:
            t_class[k] = None
            sample.append(i)
    runs = run_plain([cs[i] for i in sample], args.batch_tol, args.batch_itc, log=log)
    for i, (it, secs, ok) in zip(sample, runs):
        t_class[(cs[i]["params"].imax, cs[i]["params"].jmax)] = secs / max(it, 1) * 1e3
    same_steps = sum(1 for i, (it, _, _) in zip(sample, runs) if it == int(it64[i]))
    t_plain = sum(int(it64[i]) * t_class[(c["params"].imax, c["params"].jmax)] for i, c in enumerate(cs)) * 1e-3
    t_sample = sum(r[1] for r in runs)
    log("plain FP64 path: sample of %d cases (one per grid class) %.1f s measured; all cases %.1f s"
        % (len(sample), t_sample, t_plain))
    # batched FP32 (+ FP64 finish below the FP32 floor)
    sol = WS.solve(cs, args.batch_tol, args.batch_itc, log=log)
    stb = sol["status"]
    convb = stb == 1
    both = np.nonzero(conv & convb)[0]
    errs = {k: 0.0 for k in ("um_tau", "vm_tau", "p", "T_tau")}
    for i in both[:32]:
        for k in errs:
            a, b = sol["batch"].download(int(i), k), b64.download(int(i), k)
            errs[k] = max(errs[k], float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)))
    t_batch = sol["seconds"]
    log("batched: %.1f s (FP32 %.1f s, FP64 finish %.1f s): %d/%d converged; field error vs FP64 %s"
        % (t_batch, sol["seconds32"], sol["seconds64"], int(convb.sum()), len(cs), errs))
    plain_cu = sum(int(it64[i]) * int(cells[i]) for i in range(len(cs))) / t_plain
    out = {
        "metric": "time-to-residual speedup, batched FP32 (+FP64 finish) vs plain FP64 path", "value": t_plain / t_batch,
        "unit": "x", "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic (seeded mixed batch: %d cases)" % len(cs),
        "config": {"workload": "mixed batch of steady channel cases, grids {64..512}x{64..256}",
                   "cases": len(cs), "cells": int(cells.sum()), "tol": args.batch_tol, "itc_max": args.batch_itc,
                   "beta": args.batch_beta},
        "plain_fp64_s": t_plain, "batched_s": t_batch, "batched_fp32_s": sol["seconds32"],
        "batched_fp64_finish_s": sol["seconds64"], "fp64_batch_s": ms64 * 1e-3,
        "converged_fp64": int(conv.sum()), "converged_batched": int(convb.sum()), "converged_both": int(both.size),
        "steps_fp64_median": float(np.median(it64)), "steps_batched_fp32_median": float(np.median(sol["iters32"])),
        "field_max_rel_err_vs_fp64": errs,
        "cell_updates_per_s": {"plain_fp64": plain_cu,
                               "fp64_batch": float(np.sum(it64 * cells)) / (ms64 * 1e-3),
                               "fp32_batch": float(np.sum(sol["iters32"] * cells)) / max(sol["seconds32"], 1e-9)},
        "plain_ms_per_step": {"%dx%d" % k: v for k, v in sorted(t_class.items())},
        "plain_sample": {"cases": len(sample), "seconds_measured": t_sample,
                         "converged": int(sum(1 for r in runs if r[2])),
                         "steps_equal_to_fp64_batch": same_steps,
                         "note": "one case per grid class run for real; the rest priced at their class's measured "
                                 "time per step and their own FP64-batch step counts"},
    }
    sol["batch"].close()
    b64.close()
    print(json.dumps(out), flush=True)


def bench_batch_steps(args):
    """Throughput of batched pseudo-steps on the seeded 256-case mix (no convergence test
    stops a case: tol 0): cell-updates/s of the persistent kernels. --batch-eager issues the
    launches directly instead of replaying a hipGraph (for kernel tracing)."""
    prec = PREC_FP64 if args.precision == "fp64" else PREC_FP32
    cs = C.case_mixed_batch(n=args.batch_cases, beta=args.batch_beta)
    cells = int(sum(c["params"].imax * c["params"].jmax for c in cs))
    b = _build_batch(cs, prec)
    b.set_graph(not args.batch_eager)
    if args.tiles:
        b.set_tiles(*(int(v) for v in args.tiles.split(",")))
    elif not args.no_autotune:
        b.autotune(steps=20)
    tiles = b.get_tiles()
    log("batch of %d cases, %d cells, %s, tiles %s, %s" % (len(cs), cells, args.precision, tiles,
                                                       "eager launches" if args.batch_eager else "graph replays"))
    spg = min(50, max(1, args.steps))
    b.run(0.0, max(args.warmup, 1), steps_per_graph=min(spg, max(args.warmup, 1)))
    t0 = time.perf_counter()
    code, it, res, st = b.run(0.0, args.steps, steps_per_graph=spg)
    wall = time.perf_counter() - t0
    ms, steps, cu = b.last_timing()
    if not np.all(np.isfinite(res)):
        raise SystemExit("non-finite residual in the batch")
    b.close()
    out = {"metric": "cell-updates/s", "value": cu / wall, "unit": "cell-updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if prec == PREC_FP64 else "f32+f64",
           "data": "synthetic (seeded mixed batch: %d cases)" % len(cs),
           "config": {"workload": "mixed batch of %d steady channel cases, grids {64..512}x{64..256}, %s"
                                  % (len(cs), args.precision), "cells": cells, "tiles": list(tiles),
                      "graph": not args.batch_eager},
           "event_ms_per_step": ms / max(steps, 1)}
    print(json.dumps(out), flush=True)


def bench_warmstart(args):
    """Learned warm start on held-out cases: solver-generated steady states (one grid size)
    train the U-Net with the library's own kernels; the held-out cases are then solved from
    the network's prediction (with fallback to zero fields) and from zero fields, and the
    plain FP64 path (one context per case from zero fields) runs for real on an evenly
    spaced sample of the held-out cases (--ws-plain-sample), whose mean time per case
    stands for all of them."""
    from cfd_amd import warmstart as WS
    n = args.ws_train + args.ws_test
    H = args.ws_size
    tol = args.batch_tol
    cs = WS.steady_cases(n, seed=5, size=H, re=(10.0, 60.0), da=(3e-3, 1e-1), ri=(0.0, 1.0))
    log("warm start: %d cases on %dx%d, training data to %.0e, held-out solves to %.0e"
        % (n, H, H, args.ws_train_tol, tol))
    sol0 = WS.solve(cs, args.ws_train_tol, args.ws_itc, log=log)
    conv = sol0["status"] == 1
    log("training-data solve: %d/%d converged, %.1f s" % (int(conv.sum()), n, sol0["seconds"]))
    X = np.stack([WS.features(c) for c in cs])
    Y = np.stack([WS.targets(sol0["batch"], i) for i in range(n)])
    sol0["batch"].close()
    train_idx = [i for i in range(args.ws_train) if conv[i]]
    test_idx = [i for i in range(args.ws_train, n) if conv[i]]
    mu = Y[train_idx].mean(axis=(0, 1, 2)); sd = Y[train_idx].std(axis=(0, 1, 2)) + 1e-6
    Yn = ((Y - mu) / sd).astype(np.float32)
    net = L.UNet(H, H, 16, seed=5)
    t1 = time.perf_counter()
    hist = WS.train(net, X[train_idx], Yn[train_idx], args.ws_epochs, 16, 1e-3, log=log)
    t_train = time.perf_counter() - t1
    pred = np.concatenate([net.predict(X[test_idx[k:k + 16]]) for k in range(0, len(test_idx), 16)])
    pred = pred * sd + mu
    pred_err = float(np.sqrt(np.mean((pred - Y[test_idx]) ** 2)) / np.sqrt(np.mean(Y[test_idx] ** 2)))
    test_cases = [cs[i] for i in test_idx]
    states = [WS.state_from_prediction(pred[k], H, H) for k in range(len(test_idx))]
    # held-out cases from the prediction (fallback to zero fields) and from zero fields
    stall = (args.ws_stall_window, args.ws_stall_ratio) if args.ws_stall_window > 0 else None
    solw = WS.solve(test_cases, tol, args.ws_itc, states=states, fallback_factor=10.0, stall=stall, log=log)
    solz = WS.solve(test_cases, tol, args.ws_itc, log=log)
    itw = solw["iters32"] + solw["iters64"]; itz = solz["iters32"] + solz["iters64"]
    # the plain FP64 path on the held-out cases: one context per case from zero fields, run
    # for real on an evenly spaced sample of them (to the same tolerance and cap); the whole
    # set is the sample's mean measured time per case times the number of cases. The FP64
    # batch from zero fields gives the step counts and convergence of every case.
    b64 = WS.build_batch(test_cases, PREC_FP64)
    (_, it64, _, st64), ms64 = b64.run_sliced(tol, args.ws_itc, log=log, tail=WS.TAIL[PREC_FP64])
    ns = min(args.ws_plain_sample, len(test_cases))
    sidx = sorted(set(int(round(v)) for v in np.linspace(0, len(test_cases) - 1, ns)))
    runs = run_plain([test_cases[k] for k in sidx], tol, args.ws_itc, log=log)
    t_sample = sum(r[1] for r in runs)
    t_plain = t_sample / len(sidx) * len(test_cases)
    same_steps = sum(1 for k, r in zip(sidx, runs) if r[0] == int(it64[k]))
    t_ctx = t_sample / max(sum(r[0] for r in runs), 1) * 1e3
    log("plain FP64 path: %d sampled cases run for real, %.1f s (%.4f ms per step); all %d cases %.1f s; "
        "warm-started batch %.1f s" % (len(sidx), t_sample, t_ctx, len(test_cases), t_plain, solw["seconds"]))

    def field_diffs(sol_a, b_ref, ok):
        e = {k: 0.0 for k in ("um_tau", "vm_tau", "p", "T_tau")}
        for k in np.nonzero(ok)[0][:32]:
            for f in e:
                a, b = sol_a.download(int(k), f), b_ref.download(int(k), f)
                e[f] = max(e[f], float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)))
        return e
    errs = field_diffs(solw["batch"], solz["batch"], (solw["status"] == 1) & (solz["status"] == 1))
    errs_plain = field_diffs(solw["batch"], b64, (solw["status"] == 1) & (st64 == 1))
    solw["batch"].close(); solz["batch"].close(); b64.close()
    out = {
        "metric": "warm-start speedup (held-out batch time to residual, zero init / U-Net init)",
        "value": solz["seconds"] / solw["seconds"], "unit": "x", "n_gpus": 1, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (U-Net) + f32/f64 (solver)",
        "data": "synthetic (solver-generated steady states, seeded cases)",
        "config": {"workload": "U-Net warm start, %dx%d grids" % (H, H), "train": len(train_idx),
                   "held_out": len(test_idx), "tol": tol, "train_tol": args.ws_train_tol, "epochs": args.ws_epochs,
                   "stall_test": list(stall) if stall else None},
        "train_loss_first_last": [hist[0], hist[-1]], "train_s": t_train,
        "prediction_rel_rms_err": pred_err,
        "iters_median_zero": float(np.median(itz)), "iters_median_warm": float(np.median(itw)),
        "iters_total_zero": int(itz.sum()), "iters_total_warm": int(itw.sum()),
        "fallbacks": int(solw["fallback"].sum()),
        "converged_zero": int((solz["status"] == 1).sum()), "converged_warm": int((solw["status"] == 1).sum()),
        "solve_s_zero": solz["seconds"], "solve_s_warm": solw["seconds"],
        "steady_field_max_rel_diff_warm_vs_zero": errs,
        "plain_fp64": {"seconds": t_plain, "ms_per_step": t_ctx, "iters_total": int(it64.sum()),
                       "sample": {"cases": len(sidx), "seconds_measured": t_sample,
                                  "steps": [r[0] for r in runs], "converged": int(sum(1 for r in runs if r[2])),
                                  "steps_equal_to_fp64_batch": same_steps},
                       "converged": int((st64 == 1).sum()), "fp64_batch_s": ms64 * 1e-3,
                       "speedup_warm_batched": t_plain / solw["seconds"],
                       "steady_field_max_rel_diff_warm_batched_vs_plain": errs_plain},
    }
    net.close()
    print(json.dumps(out), flush=True)


def bench_fp32_steady(args):
    """FP32 storage against FP64 on steady fields: the same seeded steady cases solved by the
    FP64 batch to a tight tolerance, by the FP32 batch alone (stopped at its residual floor or
    the step cap), and by the FP32 batch with the FP64 finish (hand-over below 5e-7)."""
    from cfd_amd import warmstart as WS
    n, H, itc = args.acc_cases, args.ws_size, args.acc_itc
    cs = WS.steady_cases(n, seed=7, size=H, re=(10.0, 60.0), da=(3e-3, 1e-1), ri=(0.0, 1.0))
    log("FP32 vs FP64 steady fields: %d cases on %dx%d" % (n, H, H))
    b64 = WS.build_batch(cs, PREC_FP64)
    (_, it64, _, st64), ms64 = b64.run_sliced(args.acc_tol, itc, log=log, tail=WS.TAIL[PREC_FP64])
    b32 = WS.build_batch(cs, PREC_FP32)
    (_, it32, res32, st32), ms32 = b32.run_sliced(2.5e-7, itc, log=log, tail=WS.TAIL[PREC_FP32])
    hand = WS.solve(cs, args.acc_tol, itc, log=log)
    fields = ("um_tau", "vm_tau", "p", "T_tau")
    ok = np.nonzero((st64 == 1) & (hand["status"] == 1))[0]

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))
    e32 = np.array([[rel(b32.download(int(i), f), b64.download(int(i), f)) for f in fields] for i in ok])
    eh = np.array([[rel(hand["batch"].download(int(i), f), b64.download(int(i), f)) for f in fields] for i in ok])
    worst32 = e32.max(axis=1) if len(ok) else np.zeros(0)
    out = {
        "metric": "FP32-storage steady fields vs FP64 (max relative difference per case, worst field)",
        "value": float(np.median(worst32)) if len(ok) else None, "unit": "relative", "n_gpus": 1,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64 vs f64",
        "data": "synthetic (seeded steady channel cases with porous blocks)",
        "config": {"workload": "steady cases, %dx%d grids" % (H, H), "cases": n, "fp64_tol": args.acc_tol,
                   "itc_max": itc},
        "cases_compared": int(len(ok)),
        "fp32_alone": {"median": float(np.median(worst32)) if len(ok) else None,
                       "max": float(worst32.max()) if len(ok) else None,
                       "within_1e-5": int((worst32 <= 1e-5).sum()),
                       "per_field_max": dict(zip(fields, [float(v) for v in e32.max(axis=0)])) if len(ok) else None,
                       "final_residual_median": float(np.median(np.max(res32, axis=1)))},
        "fp32_then_fp64": {"max": float(eh.max()) if len(ok) else None,
                           "steps_fp32_median": float(np.median(hand["iters32"][ok])) if len(ok) else None,
                           "steps_fp64_median": float(np.median(hand["iters64"][ok])) if len(ok) else None,
                           "seconds": hand["seconds"]},
        "fp64_alone": {"steps_median": float(np.median(it64[ok])) if len(ok) else None, "seconds": ms64 * 1e-3},
    }
    for b in (b64, b32, hand["batch"]):
        b.close()
    print(json.dumps(out), flush=True)


# resolution divisor of each U-Net convolution (the layer table is L.UNET_LAYERS)
UNET_RES = [1, 1, 2, 2, 4, 4, 8, 8, 4, 4, 2, 2, 1, 1, 1]


def unet_flops(H, W, B):
    """Multiply-add FLOPs (x2) of one forward pass of the U-Net on B H x W inputs."""
    return sum(2.0 * B * (H // r) * (W // r) * 9 * cin * cout for (cin, cout, _), r in zip(L.UNET_LAYERS, UNET_RES))


def conv_kernel_ms(H, W, B, cin, cout, reps=20):
    """Mean device time (ms) of the forward and weight-gradient convolution kernels of one
    layer on device buffers, between HIP events on the NULL stream, which the library's nn
    operations use."""
    from cfd_amd.hipmem import DeviceArray, time_null_stream
    rng = np.random.default_rng(3)

    def bf16_bits(a):
        return (np.asarray(a, np.float32).view(np.uint32) >> 16).astype(np.uint16)
    x = DeviceArray(bf16_bits(rng.standard_normal((B, H, W, cin))))
    w = DeviceArray(bf16_bits(rng.standard_normal((cout, 9 * cin)) / np.sqrt(9 * cin)))
    bias = DeviceArray(np.zeros(cout, np.float32))
    y = DeviceArray(np.zeros((B, H, W, cout), np.uint16))
    dy = DeviceArray(bf16_bits(rng.standard_normal((B, H, W, cout))))
    dw = DeviceArray(np.zeros((cout, 9 * cin), np.float32))
    db = DeviceArray(np.zeros(cout, np.float32))
    lib = L.load()

    def fwd():
        L.check(lib.cfd_nn_conv3x3(x.ptr, B, H, W, cin, w.ptr, bias.ptr, cout, 1, y.ptr, None))

    def wgrad():
        L.check(lib.cfd_nn_conv3x3_wgrad(x.ptr, dy.ptr, B, H, W, cin, cout, dw.ptr, db.ptr))
    fwd(); wgrad()
    out = time_null_stream(fwd, reps), time_null_stream(wgrad, reps)
    for a in (x, w, bias, y, dy, dw, db):
        a.free()
    return out


def bench_unet(args):
    """U-Net training throughput with the library's kernels: Adam steps on a synthetic batch.
    One step = forward + input gradients + weight gradients (~3x the forward FLOPs), priced
    against MI355X's dense BF16 MFMA peak on the device time of the step (HIP events around
    the forward, backward and update; the host-to-device copies of the batch are reported
    separately as the wall time)."""
    H = W = args.unet_size
    B = 16
    net = L.UNet(H, W, B, seed=1)
    rng = np.random.default_rng(0)
    x = rng.standard_normal((B, H, W, 5)).astype(np.float32)
    t = rng.standard_normal((B, H, W, 4)).astype(np.float32)
    for _ in range(5):
        net.train_step(x, t, 1e-4)
    n = max(args.steps // 4, 20)
    dev = np.zeros((n, 3))
    t0 = time.perf_counter()
    for k in range(n):
        net.train_step(x, t, 1e-4)
        dev[k] = net.last_timing()
    wall = (time.perf_counter() - t0) / n
    net.close()
    dev_ms = float(dev.sum(axis=1).mean())
    fl = 3.0 * unet_flops(H, W, B)
    # the kernels that dominate the step: forward and weight gradient of the widest
    # full-resolution layer (concat 96 -> 32 channels)
    cin, cout = 96, 32
    f_ms, w_ms = conv_kernel_ms(H, W, B, cin, cout)
    lf = 2.0 * B * H * W * 9 * cin * cout
    kern = {"layer": "%d->%d at %dx%d, batch %d" % (cin, cout, H, W, B), "flops_per_launch": lf,
            "forward_ms": f_ms, "forward_tflops": lf / f_ms / 1e9, "wgrad_ms": w_ms, "wgrad_tflops": lf / w_ms / 1e9}
    dom = "wgrad" if w_ms >= f_ms else "forward"
    ach = kern[dom + "_tflops"]
    out = {"metric": "U-Net training step throughput", "value": fl / dev_ms / 1e9, "unit": "TFLOP/s", "n_gpus": 1,
           "steps": n, "ms_per_step": dev_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16 (fp32 accumulate)", "data": "synthetic",
           "config": {"workload": "U-Net 32/64/128/256, %dx%d, batch %d" % (H, W, B)},
           "device_ms_forward_backward_update": [float(v) for v in dev.mean(axis=0)],
           "wall_ms_per_step": wall * 1e3,
           "roofline": {"bound": "mfma", "achieved": ach, "peak": 2500.0, "unit": "TFLOP/s", "frac": ach / 2500.0,
                        "traffic": None, "kernel": "k_conv3x3_wgrad_t" if dom == "wgrad" else "k_conv3x3_t",
                        "flops_per_launch": lf},
           "kernels": kern}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--case", default="full", choices=["full", "porous"],
                    help="full: full physics on a tanh mesh; porous: uniform mesh, porous block, buoyancy")
    ap.add_argument("--imax", type=int, default=2048)
    ap.add_argument("--jmax", type=int, default=2048)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--kernel", type=int, default=1, help="0: per-operator path, 1: fused")
    ap.add_argument("--graph", type=int, default=100, help="steps per captured hipGraph (0: eager launches)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ref-steps", type=int, default=60,
                    help="pseudo-steps of the reference's Fortran hot path timed for cpu_baseline")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic_full_physics_2048x2048_fp64.json"))
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in tile shapes")
    ap.add_argument("--tiles", default="",
                    help="rows_uvp,rows_tc[,waves_uvp,waves_tc]: set the kernel shapes instead of autotuning")
    ap.add_argument("--kernel-steps", type=int, default=20, help="steps timed per kernel with events")
    ap.add_argument("--pipeline", type=int, default=1, help="two-stream pipelined pseudo-steps (1) or one stream (0)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="end with plain copy kernels of known size (memory-counter calibration)")
    ap.add_argument("--batch-eager", action="store_true", help="batch launches issued directly (no hipGraph)")
    ap.add_argument("--workload", default="grid", choices=["grid", "batch", "batchsteps", "warmstart", "unet",
                                                           "fp32steady"],
                    help="grid: one large grid (the cell-update rate); batch: time to residual of the "
                         "256-case mixed batch against the plain FP64 path")
    ap.add_argument("--batch-tol", type=float, default=1e-6)
    ap.add_argument("--batch-itc", type=int, default=50000)
    ap.add_argument("--batch-cases", type=int, default=256)
    ap.add_argument("--batch-beta", type=float, default=200.0)
    ap.add_argument("--acc-cases", type=int, default=64)
    ap.add_argument("--acc-tol", type=float, default=1e-10)
    ap.add_argument("--acc-itc", type=int, default=300000)
    ap.add_argument("--ws-train", type=int, default=2048)
    ap.add_argument("--ws-test", type=int, default=256)
    ap.add_argument("--ws-size", type=int, default=64, help="grid of the warm-start and FP32-accuracy workloads")
    ap.add_argument("--ws-stall-window", type=int, default=0,
                    help="warm start: restart from zero fields a case whose residual has not fallen below "
                         "--ws-stall-ratio x its value this many pseudo-steps earlier (0: off)")
    ap.add_argument("--ws-stall-ratio", type=float, default=0.5)
    ap.add_argument("--unet-size", type=int, default=256, help="input grid of the U-Net training-step workload")
    ap.add_argument("--ws-epochs", type=int, default=30)
    ap.add_argument("--ws-itc", type=int, default=60000)
    ap.add_argument("--ws-train-tol", type=float, default=1e-5)
    ap.add_argument("--ws-plain-sample", type=int, default=16,
                    help="held-out cases run for real on the plain FP64 path (one context each)")
    args = ap.parse_args()
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "batchsteps":
        return bench_batch_steps(args)
    if args.workload == "warmstart":
        return bench_warmstart(args)
    if args.workload == "unet":
        return bench_unet(args)
    if args.workload == "fp32steady":
        return bench_fp32_steady(args)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the pseudo-time "
              "iteration runs on one MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    prec = PREC_FP64 if args.precision == "fp64" else PREC_FP32
    if args.case == "full":
        cfg = C.case_full_physics(args.imax, args.jmax)
    else:
        cfg = C.case_porous_buoyant(args.imax)
    prm = cfg["params"]
    cells = prm.imax * prm.jmax
    log("workload %s %dx%d %s, building context" % (cfg["name"], prm.imax, prm.jmax, args.precision))
    ctx = L.Context(prm, cfg["x"], cfg["y"], cfg["eps"], cfg["lp"], precision=prec)
    ctx.set_kernel(args.kernel)
    ctx.set_pipeline(args.pipeline)
    ctx.set_state(cfg["state"])
    tiles = None
    if args.tiles:
        tt = tuple(int(v) for v in args.tiles.split(","))
        ctx.set_tiles(*tt)
        tiles = ctx.get_tiles()
    elif args.kernel == 1 and not args.no_autotune:
        tiles, tms = ctx.autotune(steps=5)
        log("autotune: rows u/v/p %d, rows T/C %d, waves u/v/p %d, waves T/C %d (0 = built-in), %.4f ms/step"
            % (tiles + (tms,)))
    if tiles is not None:
        # waves 0 means the built-in budget, which is the one-wave / as-compiled variant
        tiles = dict(zip(("rows_uvp", "rows_tc", "waves_uvp", "waves_tc"),
                         (tiles[0], tiles[1], tiles[2] or 1, tiles[3] or 1)))
    ctx.pseudo_steps(max(args.warmup, 1))
    log("warmup done (%d steps), timing %d steps" % (args.warmup, args.steps))

    t0 = time.perf_counter()
    if args.graph and args.steps >= args.graph:
        reps = args.steps // args.graph
        res = ctx.pseudo_steps_graph(args.graph, reps)
        steps = reps * args.graph
    else:
        res = ctx.pseudo_steps(args.steps)
        steps = args.steps
    wall = time.perf_counter() - t0
    ms_event = ctx.last_ms_per_step()
    n_fused, n_exact = ctx.step_counts()
    if not np.all(np.isfinite(res)):
        raise SystemExit("non-finite residual %s" % res)
    log("timed: %.3f s wall, %.4f ms/step (events), residuals %s" % (wall, ms_event, res))

    value = cells * steps / wall
    Bstep = BYTES_PER_CELL.get(("step", prec), 128)
    step_gbs = Bstep * cells / (ms_event * 1e-3) / 1e9
    kern = None
    if args.kernel == 1 and args.kernel_steps > 0:
        _, kms = ctx.pseudo_steps_timed(args.kernel_steps)
        kern = {"uvp_ms": kms[0], "tc_ms": kms[1], "bc_ms": kms[2], "step_ms": kms[3]}
        log("per kernel (events, mean of %d): u/v/p %.4f ms, T/C %.4f ms, boundary %.4f ms"
            % (args.kernel_steps, kms[0], kms[1], kms[2]))
    if kern is not None and prec == PREC_FP64:
        # dominant kernel: the fused u/v/p kernel
        B1 = BYTES_PER_CELL[("k1", prec)]
        achieved = B1 * cells / (kern["uvp_ms"] * 1e-3) / 1e9
        rname, rB, rms = "k_uvp (fused u/v/p pseudo-step kernel)", B1, kern["uvp_ms"]
    else:
        achieved, rname, rB, rms = step_gbs, "whole pseudo-step", Bstep, ms_event
    traffic = None
    if os.path.exists(args.traffic):
        with open(args.traffic) as f:
            t = json.load(f)
        tt = t.get("tiles") or {}
        # the recorded figure is the u/v/p kernel's: it depends on that kernel's shape only
        if (t.get("imax") == prm.imax and t.get("jmax") == prm.jmax and t.get("kernel") == args.kernel
                and tiles is not None and prec == PREC_FP64
                and tt.get("rows_uvp") == tiles["rows_uvp"] and tt.get("waves_uvp") == tiles["waves_uvp"]):
            traffic = t.get("uvp_bytes_per_launch")
    if args.copy_probe:
        # plain copy kernels of known size (physical level -> pseudo-time level of u, v, T,
        # C), so that memory counters can be scaled against the bytes they must move
        ctx.sync_levels()
        nu, nv, npc = (prm.imax + 2) * (prm.jmax + 1), (prm.imax + 1) * (prm.jmax + 2), (prm.imax + 1) * (prm.jmax + 1)
        log("copy probe: k_copy launches of %d, %d, %d, %d doubles (read once, written once)" % (nu, nv, npc, npc))

    out = {
        "metric": "cell-updates/s", "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == PREC_FP64 else "f32+f64",
        "data": "synthetic (seeded SplitMix64 inputs: tanh mesh, porous block, perturbed state)",
        "config": {"workload": "%s fused pseudo-step, %s" % (cfg["name"], args.precision),
                   "imax": prm.imax, "jmax": prm.jmax, "kernel": args.kernel,
                   "graph_steps": args.graph, "tiles": tiles, "pipeline": args.pipeline,
                   "steps_fused": n_fused, "steps_exact": n_exact},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": rname, "bytes_per_cell": rB, "cells_per_launch": cells, "kernel_ms": rms},
        "step_roofline": {"bytes_per_cell": Bstep, "ms": ms_event, "achieved": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS},
        "kernel_ms": kern,
    }
    ctx.close()
    # measured copy bandwidth of this GPU (1 GiB buffers), reported beside the nominal peak
    copy_gbs = L.hbm_copy_bandwidth(1 << 30, 20)
    log("stream copy: %.0f GB/s (read + write) against the nominal %.0f GB/s" % (copy_gbs, HBM_PEAK_GBS))
    out["roofline"]["measured_copy_gbs"] = copy_gbs
    out["roofline"]["frac_of_measured_copy"] = achieved / copy_gbs
    out["step_roofline"]["frac_of_measured_copy"] = step_gbs / copy_gbs
    if not args.no_cpu:
        log("cpu baseline (%.0f s sample)" % args.cpu_seconds)
        port = cpu_baseline(cfg, args.cpu_seconds)
        if os.path.exists(REF_BENCH) and (prm.imax, prm.jmax) == (2048, 2048):
            log("reference CPU path (%d steps)" % args.ref_steps)
            out["cpu_baseline"] = cpu_reference(cfg, args.ref_steps)
            out["cpu_port"] = port
        else:
            out["cpu_baseline"] = port
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
