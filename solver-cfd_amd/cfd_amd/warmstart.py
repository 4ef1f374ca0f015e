"""Learned warm start for batched steady solves.

Pipeline (all GPU work through libcfd_amd.so: the batched solver and the U-Net kernels):
  1. cases of one grid size (the U-Net's H x W) with seeded parameters and porous blocks;
  2. steady states from the batched solver (FP32 storage until its floor, FP64 finish);
  3. U-Net training on (mask, porosity, log Re, log Da, Ri) -> (u, v, p, T);
  4. inference gives each new case an initial state; the solver iterates from it to the
     tolerance, and a case whose residual grows past `fallback_factor` x its best (the
     batch's divergence test), or optionally stalls (no fall below a ratio of its value a
     window of steps earlier, tested between slices), is restarted from the zero initial
     state.

Field mapping on the p-cell index grid (i = 1..H, j = 1..W): channel 0 is um(i, j), 1 is
vm(i, j), 2 is p(i, j), 3 is T(i, j); um(imax+1, .) and vm(., jmax+1) copy their neighbours,
and the channel BC set is applied to the whole state before the solve.
"""
import math

import numpy as np

from . import cases as C
from . import lib as L
from .params import PREC_FP32, PREC_FP64

RE_RANGE = (10.0, 500.0)
DA_RANGE = (1e-4, 1e-1)
RI_RANGE = (0.0, 5.0)


def features(case):
    """U-Net input [H, W, 5] of one case."""
    P = case["params"]
    H, W = P.imax, P.jmax
    x = np.zeros((H, W, 5), np.float32)
    x[..., 0] = case["lp"][1:, 1:]
    x[..., 1] = case["eps"][1:, 1:]
    x[..., 2] = (math.log(P.re) - math.log(RE_RANGE[0])) / (math.log(RE_RANGE[1]) - math.log(RE_RANGE[0]))
    x[..., 3] = (math.log(P.darcy) - math.log(DA_RANGE[0])) / (math.log(DA_RANGE[1]) - math.log(DA_RANGE[0]))
    x[..., 4] = P.ri / RI_RANGE[1]
    return x


def targets(batch, c):
    """Steady fields of case c of a batch as [H, W, 4] (um, vm, p, T on the p-cell grid)."""
    um = batch.download(c, "um_tau"); vm = batch.download(c, "vm_tau")
    p = batch.download(c, "p"); T = batch.download(c, "T_tau")
    return np.stack([um[1:-1, 1:], vm[1:, 1:-1], p[1:, 1:], T[1:, 1:]], axis=-1).astype(np.float32)


def state_from_prediction(y, imax, jmax):
    """Initial iterates (C layouts) from a U-Net output [H, W, 4]."""
    um = C.zeros("u", imax, jmax); vm = C.zeros("v", imax, jmax)
    p = C.zeros("p", imax, jmax); T = C.zeros("p", imax, jmax); Z = C.zeros("p", imax, jmax)
    um[1:imax + 1, 1:] = y[..., 0]; um[imax + 1, 1:] = y[-1, :, 0]
    vm[1:, 1:jmax + 1] = y[..., 1]; vm[1:, jmax + 1] = y[:, -1, 1]
    p[1:, 1:] = y[..., 2]; T[1:, 1:] = y[..., 3]
    return {"um": um, "vm": vm, "p": p, "T": T, "C": Z}


# Row blocks for the end of a solve, by precision and the interior cells of the cases still
# running (same kernel variants as the built-in shape, so the same results). The built-in
# blocks are sized for a full batch and leave most of the GPU idle once few cells run; a
# resumed run lists only the running cases' tiles, and a shorter block pays while those
# tiles fit one round of the resident workgroups (about 1024 in FP64, twice that in FP32).
# Measured per pseudo-step on 64x64 cases, built-in / 16-row / 8-row blocks: FP64 136 cases
# (523 K cells) 0.175 / 0.202 / 0.210 ms, 100 cases 0.172 / 0.125 / 0.148, 40 cases 0.151 /
# 0.117 / 0.086; FP32 136 cases 0.149 / 0.116 / 0.129, 100 cases 0.147 / 0.106 / 0.087.
TAIL = {PREC_FP64: [(400_000, (16, 16, 1, 1)), (200_000, (8, 8, 1, 1))],
        PREC_FP32: [(960_000, (16, 16, 1, 1)), (400_000, (8, 8, 1, 1))]}


def build_batch(cases, prec, states=None):
    b = L.Batch([c["params"] for c in cases], precision=prec)
    for i, c in enumerate(cases):
        b.set_mesh(i, c["x"], c["y"])
        b.upload(i, "eps", c["eps"])
        b.upload(i, "lp", c["lp"])
        b.set_state(i, states[i] if states is not None else c["state"])
    b.apply_bc()
    return b


def solve(cases, tol, itc_max, states=None, switch_tol=5e-7, fallback_factor=0.0, stall=None, log=None):
    """Steady solve of every case: FP32 batch down to max(tol, switch_tol), FP64 finish to
    tol when tol is below what FP32 storage resolves (switch_tol). With fallback_factor > 0,
    cases started from `states` whose residual grows past that factor x their best, or (stall
    = (window, ratio)) does not fall below ratio x its value `window` steps earlier, are
    restarted from their zero initial state. Returns a dict with per-case iterations (FP32 +
    FP64), status, event time (s), and the final FP64 batch."""
    b32 = build_batch(cases, PREC_FP32, states)
    warm = fallback_factor > 0 and states is not None
    (code, it32, res32, st32), ms32 = b32.run_sliced(max(tol, switch_tol), itc_max, diverge_factor=fallback_factor,
                                                    log=log, stall=stall if warm else None, tail=TAIL[PREC_FP32])
    fallback = np.zeros(len(cases), bool)
    ms_fb = 0.0
    if warm:
        bad = np.nonzero((st32 == 2) | (st32 == 3) | (st32 == 5))[0]
        if bad.size:
            fallback[bad] = True
            sub = [cases[i] for i in bad]
            bfb = build_batch(sub, PREC_FP32, None)
            (c2, it2, r2, s2), ms_fb = bfb.run_sliced(max(tol, switch_tol), itc_max, log=log, tail=TAIL[PREC_FP32])
            # hand the restarted cases' states back into the FP32 batch
            for k, i in enumerate(bad):
                for f in ("um_tau", "vm_tau", "p", "T_tau", "C_tau"):
                    b32.upload(int(i), f, bfb.download(k, f))
                it32[i] += it2[k]
                st32[i] = s2[k]
            bfb.close()
    if tol >= switch_tol:          # FP32 storage resolves tol: no FP64 finish
        return dict(iters32=it32, iters64=np.zeros_like(it32), status=st32, res=res32, fallback=fallback,
                    seconds=(ms32 + ms_fb) * 1e-3, seconds32=(ms32 + ms_fb) * 1e-3, seconds64=0.0, batch=b32)
    b64 = L.Batch([c["params"] for c in cases], precision=PREC_FP64)
    for i, c in enumerate(cases):
        b64.set_mesh(i, c["x"], c["y"])
    b64.copy_state_from(b32)
    b32.close()
    (code, it64, res64, st64), ms64 = b64.run_sliced(tol, itc_max, log=log, tail=TAIL[PREC_FP64])
    return dict(iters32=it32, iters64=it64, status=st64, res=res64, fallback=fallback,
                seconds=(ms32 + ms_fb + ms64) * 1e-3, seconds32=(ms32 + ms_fb) * 1e-3, seconds64=ms64 * 1e-3,
                batch=b64)


def steady_cases(n, seed, size=64, re=(10.0, 100.0), da=(1e-3, 1e-1), ri=(0.0, 2.0)):
    """Cases on one size x size grid with parameters drawn from the given ranges (the
    steady regime of the mixed-batch distribution)."""
    cs = C.case_mixed_batch(n=n, seed=seed, sizes_i=(size,), sizes_j=(size,), beta=200.0)
    r = C.uniform(seed * 31 + 7, 3 * n).reshape(n, 3)
    out = []
    for c, u in zip(cs, r):
        P = c["params"]
        P.re = re[0] * (re[1] / re[0]) ** u[0]
        P.darcy = da[0] * (da[1] / da[0]) ** u[1]
        P.ri = ri[0] + (ri[1] - ri[0]) * u[2]
        P.pe = 0.7 * P.re
        P.dt = C.steady_dt(c["x"], c["y"], P.re, P.pe, P.darcy, P.cf, 0.3)
        c["re"], c["da"], c["ri"] = P.re, P.darcy, P.ri
        out.append(c)
    return out


def train(net, X, Y, epochs, batch, lr, seed=0, log=None):
    """Adam over shuffled mini-batches; returns the mean loss per epoch."""
    rng = np.random.default_rng(seed)
    hist = []
    n = X.shape[0]
    for ep in range(epochs):
        perm = rng.permutation(n)
        tot, cnt = 0.0, 0
        for k in range(0, n - batch + 1, batch):
            idx = perm[k:k + batch]
            tot += net.train_step(X[idx], Y[idx], lr)
            cnt += 1
        hist.append(tot / max(cnt, 1))
        if log:
            log("  epoch %d: loss %.4e" % (ep + 1, hist[-1]))
    return hist
