"""GPU: FP32 storage mode and the batched persistent path.

FP32 contexts keep u, v, T, C (all levels) and eps, L in single precision with FP64
pressure, continuity update and residual maxima; their trajectories are compared with the
FP64 CPU oracle at 1e-5 max relative difference per field (normwise), the bar for FP32
fields. Batches run the same tile routines as single contexts, so a batch case reproduces a
context of the same case bit for bit (FP64 and FP32), including the exit step of the
convergence test, which the batch evaluates on the device.
"""
import sys

import numpy as np
import pytest

import oracle as O
from cfd_amd import lib as L
from cfd_amd import cases as C
from cfd_amd.params import make_params, BC_CHANNEL, BUOY_BOUSSINESQ, MU_POLY4, PREC_FP32, PREC_FP64
from helpers import max_rel

pytestmark = pytest.mark.gpu
TOL32 = 1e-5
FIELDS = ("um_tau", "vm_tau", "p", "T_tau", "C_tau")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _setup():
    O.build()
    L.load()


def _case(imax, jmax, seed, steady=0, re=100.0, da=1e-2, ri=1.0, mu=False, beta=1.0, amp=5e-2):
    x = C.tanh_nodes(imax, 0.0, 1.0, 0.3, 2.0)
    y = C.tanh_nodes(jmax, 0.0, 1.0, 0.3, 2.0)
    eps, lp = C.porous_block(x, y, 0.3, 0.6, 0.25, 0.7, 0.4, 0.9, seed=seed)
    dt = C.stable_dt(x, y, 2.0 / re, 1.0 / (0.7 * re), 1.5, 0.05)
    kw = dict(re=re, pe=0.7 * re, sc=0.7, beta=beta, darcy=da, cf=0.55, g=0.0, dt=dt, bc=BC_CHANNEL,
              buoyancy=BUOY_BOUSSINESQ, ri=ri, t_ref=0.5, invfr2=0.0, u_in=1.0, t_bottom=1.0, t_top=0.0,
              slot_lo=0.45, slot_hi=0.55, slot_value=1.0, steady=steady)
    if mu:
        kw.update(mu_model=MU_POLY4, mu_coef=C.mu_poly_coefficients(3))
    prm = make_params(imax, jmax, **kw)
    st = C.perturbed_state(imax, jmax, seed=seed, amp=amp, u0=0.4, v0=0.05, p0=0.2, T0=0.6, C0=0.1,
                           perturb_scalars=True)
    return prm, x, y, eps, lp, st


# ---- FP32 contexts --------------------------------------------------------------------
@pytest.mark.parametrize("shape,steady,mu", [((140, 520), 0, True), ((300, 260), 1, False), ((64, 48), 0, False)])
def test_fp32_trajectory_close_to_fp64_oracle(shape, steady, mu):
    say("case: FP32 trajectory", shape, "steady", steady, "mu", mu)
    prm, x, y, eps, lp, st = _case(*shape, seed=21, steady=steady, mu=mu)
    o = O.Oracle(prm, x, y, eps, lp)
    s, cs = o.make_state(st)
    h_ref = o.pseudo_steps(s, cs, 60)
    ctx = L.Context(prm, x, y, eps, lp, precision=PREC_FP32)
    ctx.set_state(st)
    res, hist = ctx.pseudo_steps(60, history=True)
    nf, ne = ctx.step_counts()
    errs = {k: max_rel(ctx.download(k), s[k]) for k in FIELDS}
    say(shape, "hist", max_rel(hist, h_ref), errs, nf, ne)
    assert (nf, ne) == (59, 1)
    assert max(errs.values()) <= TOL32
    assert max_rel(hist, h_ref) <= 1e-3        # residual maxima: differences of FP32 fields
    ctx.close()


def test_fp32_graph_replays_match_eager():
    say("case: FP32 graph vs eager")
    prm, x, y, eps, lp, st = _case(100, 300, seed=22)
    outs = []
    for mode in ("eager", "graph"):
        ctx = L.Context(prm, x, y, eps, lp, precision=PREC_FP32)
        ctx.set_state(st)
        ctx.pseudo_steps(1)
        if mode == "eager":
            ctx.pseudo_steps(24)
        else:
            for _ in range(3):
                ctx.pseudo_steps_graph(8, 1)
                _ = ctx.download("p")
        outs.append({k: ctx.download(k) for k in FIELDS})
        ctx.close()
    for k in FIELDS:
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_fp32_physical_steps_close_to_fp64():
    """Physical steps (the reference loop with its exit without update) in FP32 and FP64:
    the FP32 context converts at the path changes and stays within the FP32 bar."""
    say("case: FP32 physical steps")
    prm, x, y, eps, lp, st = _case(48, 40, seed=23)
    out = []
    for prec in (PREC_FP64, PREC_FP32):
        ctx = L.Context(prm, x, y, eps, lp, precision=prec)
        ctx.set_state(st)
        its = []
        for k in range(3):
            it, res = ctx.physical_step(2e-3, 60, k == 0)
            its.append(it)
        out.append(({f: ctx.download(f) for f in ("um", "vm", "p", "T", "C")}, its))
        ctx.close()
    (f64, it64), (f32, it32) = out
    errs = {k: max_rel(f32[k], f64[k]) for k in f64}
    say("iters", it64, it32, errs)
    assert max(errs.values()) <= TOL32
    assert all(abs(a - b) <= 1 for a, b in zip(it64, it32))


# ---- batches ----------------------------------------------------------------------------
BATCH_SHAPES = [(40, 70, 31, 80.0), (96, 33, 32, 300.0), (150, 260, 33, 25.0), (24, 20, 34, 150.0)]


def _bc_consistent(prec, steady, shapes=BATCH_SHAPES):
    """Cases stepped once through a context (exact first step): BC-consistent iterates."""
    cases = []
    for (im, jm, seed, re) in shapes:
        prm, x, y, eps, lp, st = _case(im, jm, seed=seed, steady=steady, re=re, da=10 ** (-1 - seed % 3))
        ctx = L.Context(prm, x, y, eps, lp, precision=prec)
        ctx.set_state(st)
        ctx.pseudo_steps(1)
        st1 = {k: ctx.download(k) for k in FIELDS}
        if not steady:
            st1.update({k: ctx.download(k) for k in ("um", "vm", "T", "C")})
        cases.append((prm, x, y, eps, lp, st1, ctx))
    return cases


def _make_batch(cases, prec):
    b = L.Batch([c[0] for c in cases], precision=prec)
    for i, (prm, x, y, eps, lp, st1, ctx) in enumerate(cases):
        b.set_mesh(i, x, y)
        b.upload(i, "eps", eps)
        b.upload(i, "lp", lp)
        b.set_state(i, st1)
    return b


@pytest.mark.parametrize("prec,steady,waves,waves_tc", [(PREC_FP64, 1, 1, 1), (PREC_FP64, 0, 1, 1), (PREC_FP32, 1, 1, 1),
                                                        (PREC_FP64, 1, 3, 1), (PREC_FP32, 0, 3, 1), (PREC_FP64, 0, 1, 2),
                                                        (PREC_FP32, 1, 3, 2)])
def test_batch_matches_single_contexts_bitwise(prec, steady, waves, waves_tc):
    say("case: batch vs contexts", "fp32" if prec == PREC_FP32 else "fp64", "steady", steady, "waves", waves, waves_tc)
    cases = _bc_consistent(prec, steady)
    b = _make_batch(cases, prec)
    b.set_tiles(24, 12, waves, waves_tc)
    code, iters, res, status = b.run(0.0, 25, steps_per_graph=10)
    say("batch", code, iters, status)
    assert code == 0 and list(iters) == [25] * len(cases) and list(status) == [4] * len(cases)
    for i, (prm, x, y, eps, lp, st1, ctx) in enumerate(cases):
        ctx.set_tiles(24, 12, waves, waves_tc)
        r = ctx.pseudo_steps(25)
        for k in FIELDS:
            a, c = b.download(i, k), ctx.download(k)
            assert np.array_equal(a, c), (i, k, max_rel(a, c))
        assert np.array_equal(res[i], r), (i, res[i], r)
        ctx.close()
    b.close()


@pytest.mark.parametrize("prec,steady", [(PREC_FP64, 1), (PREC_FP64, 0), (PREC_FP32, 1), (PREC_FP32, 0)])
def test_row_heights_do_not_change_bits(prec, steady):
    """Blocks of 5 and 3 rows (odd, shorter than the stage lags) in the contexts against 24
    and 12 rows in the batch: each cell's arithmetic does not depend on the block it falls
    in (rows run in pairs from even rows; guarded and unguarded row forms round alike)."""
    say("case: row heights, contexts 5/3 vs batch 24/12", "fp32" if prec == PREC_FP32 else "fp64", "steady", steady)
    cases = _bc_consistent(prec, steady)
    b = _make_batch(cases, prec)
    b.set_tiles(24, 12, 1, 1)
    code, iters, res, status = b.run(0.0, 25, steps_per_graph=10)
    assert code == 0 and list(iters) == [25] * len(cases)
    for i, (prm, x, y, eps, lp, st1, ctx) in enumerate(cases):
        ctx.set_tiles(5, 3, 1, 1)
        r = ctx.pseudo_steps(25)
        for k in FIELDS:
            a, c = b.download(i, k), ctx.download(k)
            assert np.array_equal(a, c), (i, k, max_rel(a, c))
        assert np.array_equal(res[i], r), (i, res[i], r)
        ctx.close()
    b.close()


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32])
def test_wide_batch_matches_single_contexts_bitwise(prec):
    """Cases 700 and 1030 columns wide: the batch takes 256-column workgroups (fewer lanes
    than 64-column ones for these widths) and still reproduces the contexts bit for bit."""
    say("case: wide batch vs contexts", "fp32" if prec == PREC_FP32 else "fp64")
    cases = _bc_consistent(prec, 1, shapes=[(40, 1030, 41, 80.0), (30, 700, 42, 120.0)])
    b = _make_batch(cases, prec)
    b.set_tiles(16, 8, 1, 1)
    code, iters, res, status = b.run(0.0, 20, steps_per_graph=10)
    assert code == 0 and list(iters) == [20] * len(cases)
    for i, (prm, x, y, eps, lp, st1, ctx) in enumerate(cases):
        ctx.set_tiles(16, 8, 1, 1)
        r = ctx.pseudo_steps(20)
        for k in FIELDS:
            a, c = b.download(i, k), ctx.download(k)
            assert np.array_equal(a, c), (i, k, max_rel(a, c))
        assert np.array_equal(res[i], r), (i, res[i], r)
        ctx.close()
    b.close()


def test_batch_exit_test_matches_run_to_tol():
    """Per-case stop of the device exit test = cfd_run_to_tol with check_every=1 (same step,
    same state, the step's output kept without the tau update)."""
    say("case: batch exit test")
    cases = _bc_consistent(PREC_FP64, 1)
    # a tolerance each case crosses at a different step
    tols = []
    for (prm, x, y, eps, lp, st1, ctx) in cases:
        c2 = L.Context(prm, x, y, eps, lp)
        c2.set_state(st1)
        c2.pseudo_steps(1)
        _, h = c2.pseudo_steps(300, history=True)
        tols.append(float(np.median(h[100:300].max(axis=1))))
        c2.close()
    tol = float(np.median(tols))
    b = _make_batch(cases, PREC_FP64)
    code, iters, res, status = b.run(tol, 2000, steps_per_graph=16)
    say("tol", tol, "batch", code, iters, status)
    for i, (prm, x, y, eps, lp, st1, ctx) in enumerate(cases):
        c, it, r = ctx.run_to_tol(tol, 2000, check_every=1, raise_on_fail=False)
        say(i, "context", c, it, "batch", status[i], iters[i])
        assert it == iters[i]
        assert (status[i] == 1) == (c == 0)
        if status[i] == 1:
            for k in FIELDS:
                assert np.array_equal(b.download(i, k), ctx.download(k)), (i, k)
        ctx.close()
    b.close()


def test_batch_nonfinite_case_stops_alone():
    say("case: batch with one non-finite case")
    cases = _bc_consistent(PREC_FP64, 1)
    b = _make_batch(cases, PREC_FP64)
    bad = cases[1][5]["um_tau"].copy()
    bad[10, 10] = np.nan
    b.upload(1, "um_tau", bad)
    code, iters, res, status = b.run(0.0, 12, steps_per_graph=4)
    say(code, iters, status)
    assert status[1] == 2 and iters[1] == 1
    assert all(status[i] == 4 and iters[i] == 12 for i in (0, 2, 3))
    assert code == -3          # CFD_ERR_NONFINITE
    for c in cases:
        c[6].close()
    b.close()


def test_batch_autotune_restores_state():
    say("case: batch autotune")
    cases = _bc_consistent(PREC_FP32, 1)
    b = _make_batch(cases, PREC_FP32)
    before = [{k: b.download(i, k) for k in FIELDS} for i in range(len(cases))]
    tiles, ms = b.autotune(steps=4)
    say("tuned", tiles, ms)
    assert ms > 0 and b.get_tiles() == tiles
    for i in range(len(cases)):
        for k in FIELDS:
            assert np.array_equal(b.download(i, k), before[i][k]), (i, k)
    # and the tuned shape still reproduces the contexts run with the same kernel variants
    code, iters, res, status = b.run(0.0, 8, steps_per_graph=8)
    for i, (prm, x, y, eps, lp, st1, ctx) in enumerate(cases):
        ctx.set_tiles(*tiles)
        ctx.pseudo_steps(8)
        for k in FIELDS:
            assert np.array_equal(b.download(i, k), ctx.download(k)), (i, k)
        ctx.close()
    b.close()


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32])
def test_sliced_run_with_stopped_cases_equals_one_run(prec):
    """A resumed run takes the cases that stay stopped out of the tile lists (the kernels'
    blocks take tiles by a fixed stride over them): the cases that go on give the steps,
    residuals, status and fields of one unsliced run, bit for bit."""
    from cfd_amd import warmstart as WS
    say("case: sliced run with stopped cases vs one run", "fp32" if prec == PREC_FP32 else "fp64")
    cs = WS.steady_cases(8, seed=11, size=32, re=(10.0, 60.0), da=(3e-3, 1e-1), ri=(0.0, 1.0))
    probe = WS.build_batch(cs, prec)
    _, _, r, _ = probe.run(0.0, 1500, steps_per_graph=100)
    probe.close()
    tol = float(np.sort(r.max(axis=1))[2]) * 1.0001       # about three cases stop by step 1500
    outs = []
    for sliced in (False, True):
        b = WS.build_batch(cs, prec)
        if sliced:
            (code, it, res, st), ms = b.run_sliced(tol, 6000, slice_steps=500, steps_per_graph=100)
        else:
            code, it, res, st = b.run(tol, 6000, steps_per_graph=100)
        outs.append((it, res, st, [{k: b.download(i, k) for k in FIELDS} for i in range(len(cs))]))
        b.close()
    (ia, ra, sa, fa), (ib, rb, sb, fb) = outs
    say("tol %.3e" % tol, "steps", ia.tolist(), "status", sa.tolist())
    assert len(set((ia // 500).tolist())) > 1                # cases stop in different slices
    assert np.array_equal(ia, ib) and np.array_equal(sa, sb) and np.array_equal(ra, rb)
    for i in range(len(cs)):
        for k in FIELDS:
            assert np.array_equal(fa[i][k], fb[i][k]), (i, k)


def test_batch_resume_equals_one_run():
    """A run sliced with resume (10 + 15 steps) gives the state and counts of one 25-step run."""
    say("case: batch resume")
    outs = []
    for sliced in (False, True):
        cases = _bc_consistent(PREC_FP64, 1)
        b = _make_batch(cases, PREC_FP64)
        if sliced:
            b.run(0.0, 10, steps_per_graph=5)
            code, it, res, st = b.run(0.0, 25, steps_per_graph=5, resume=True)
        else:
            code, it, res, st = b.run(0.0, 25, steps_per_graph=5)
        outs.append((list(it), [{k: b.download(i, k) for k in FIELDS} for i in range(len(cases))]))
        for c in cases:
            c[6].close()
        b.close()
    assert outs[0][0] == outs[1][0] == [25] * len(BATCH_SHAPES)
    for i in range(len(BATCH_SHAPES)):
        for k in FIELDS:
            assert np.array_equal(outs[0][1][i][k], outs[1][1][i][k]), (i, k)
