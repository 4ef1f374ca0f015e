"""GPU: the learned warm start end to end, on a problem small enough for a test.

Steady states of seeded 32 x 32 cases come from the batched solver; the U-Net is trained on
them with the library's own kernels; held-out cases are then solved from the network's
prediction and from the zero initial state. The trained network must save pseudo-steps,
both starts must reach the same steady state, and a start the solver cannot use (one that
diverges) must fall back to the zero initial state and still converge.
"""
import sys

import numpy as np
import pytest

from cfd_amd import lib as L
from cfd_amd import warmstart as WS
from cfd_amd.params import PREC_FP32, PREC_FP64

pytestmark = pytest.mark.gpu

H = 32
N_TRAIN, N_TEST = 256, 32
TOL_DATA, TOL_SOLVE, ITC = 1e-5, 1e-6, 100000
FIELDS = ("um_tau", "vm_tau", "p", "T_tau", "C_tau")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def trained():
    say("case: warm start data, %d cases on %dx%d" % (N_TRAIN + N_TEST, H, H))
    cs = WS.steady_cases(N_TRAIN + N_TEST, seed=9, size=H, re=(10.0, 60.0), da=(3e-3, 1e-1), ri=(0.0, 1.0))
    sol = WS.solve(cs, TOL_DATA, ITC)
    conv = sol["status"] == 1
    X = np.stack([WS.features(c) for c in cs])
    Y = np.stack([WS.targets(sol["batch"], i) for i in range(len(cs))])
    sol["batch"].close()
    tr = [i for i in range(N_TRAIN) if conv[i]]
    te = [i for i in range(N_TRAIN, N_TRAIN + N_TEST) if conv[i]]
    say("converged: train %d/%d, held out %d/%d" % (len(tr), N_TRAIN, len(te), N_TEST))
    assert len(tr) >= 0.5 * N_TRAIN and len(te) >= 12
    mu = Y[tr].mean(axis=(0, 1, 2)); sd = Y[tr].std(axis=(0, 1, 2)) + 1e-6
    net = L.UNet(H, H, 16, seed=5)
    hist = WS.train(net, X[tr], ((Y[tr] - mu) / sd).astype(np.float32), 20, 16, 1e-3)
    say("training loss", hist[0], "->", hist[-1])
    pred = np.concatenate([net.predict(X[te[k:k + 16]]) for k in range(0, len(te), 16)]) * sd + mu
    net.close()
    yield cs, te, pred, Y, hist


def test_trained_network_saves_pseudo_steps(trained):
    cs, te, pred, Y, hist = trained
    assert hist[-1] < 0.2 * hist[0]
    err = float(np.sqrt(np.mean((pred - Y[te]) ** 2)) / np.sqrt(np.mean(Y[te] ** 2)))
    cases = [cs[i] for i in te]
    states = [WS.state_from_prediction(pred[k], H, H) for k in range(len(te))]
    solw = WS.solve(cases, TOL_SOLVE, ITC, states=states, fallback_factor=10.0)
    solz = WS.solve(cases, TOL_SOLVE, ITC)
    itw = solw["iters32"] + solw["iters64"]
    itz = solz["iters32"] + solz["iters64"]
    say("prediction error %.3f; pseudo-steps from prediction %d (median %d), from zero %d (median %d); "
        "fallbacks %d" % (err, itw.sum(), np.median(itw), itz.sum(), np.median(itz), solw["fallback"].sum()))
    assert (solw["status"] == 1).all() and (solz["status"] == 1).all()
    assert err < 0.5
    assert itw.sum() < 0.9 * itz.sum()
    # both starts end at the same steady state, to the spread a 1e-6 stop leaves
    worst = 0.0
    for k in range(len(te)):
        for f in ("um_tau", "vm_tau", "T_tau"):
            a, b = solw["batch"].download(k, f), solz["batch"].download(k, f)
            worst = max(worst, float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)))
    say("steady fields, prediction start vs zero start: max rel diff %.2e" % worst)
    assert worst < 5e-2
    solw["batch"].close(); solz["batch"].close()


def test_unusable_start_falls_back_to_zero_init(trained):
    cs, te, pred, Y, hist = trained
    cases = [cs[i] for i in te[:8]]
    # velocities 1000x too large: the explicit pseudo-step diverges from such a start
    states = [WS.state_from_prediction(pred[k] * np.array([1e3, 1e3, 1.0, 1.0], np.float32), H, H)
              for k in range(len(cases))]
    sol = WS.solve(cases, TOL_SOLVE, ITC, states=states, fallback_factor=10.0)
    say("fallbacks %d/%d, status %s" % (sol["fallback"].sum(), len(cases), sol["status"]))
    assert sol["fallback"].all()
    assert (sol["status"] == 1).all()
    sol["batch"].close()


def test_stalled_start_falls_back_to_zero_init(trained):
    """The stall test of the warm start (run_sliced's stall=(window, ratio), applied between
    slices of 10 000 steps) with a ratio no residual can meet: every case still running after
    20 000 steps is stopped and restarted from its zero initial state. Started from that same
    zero state, a case that converges within 20 000 steps is untouched and takes exactly the
    zero-start steps; every other one takes exactly 20 000 + the zero-start steps (the batch
    kernels give each case the same bits whatever batch it is in)."""
    cs, te, pred, Y, hist = trained
    cases = [cs[i] for i in te[:12]]
    solz = WS.solve(cases, TOL_SOLVE, ITC)
    itz = solz["iters32"] + solz["iters64"]
    solz["batch"].close()
    sols = WS.solve(cases, TOL_SOLVE, ITC, states=[c["state"] for c in cases], fallback_factor=1e30,
                    stall=(10000, 1e-30))
    its = sols["iters32"] + sols["iters64"]
    say("zero-start steps %s; with the stall test %s; fallbacks %s" % (itz.tolist(), its.tolist(),
                                                                       sols["fallback"].astype(int).tolist()))
    assert (solz["status"] == 1).all() and (sols["status"] == 1).all()
    late = itz > 20000
    assert late.any()
    assert np.array_equal(sols["fallback"], late)
    assert np.array_equal(its, np.where(late, itz + 20000, itz))
    sols["batch"].close()


@pytest.mark.parametrize("stage", [0, 1])
@pytest.mark.parametrize("prec", [PREC_FP32, PREC_FP64])
def test_tail_tiles_keep_results_bitwise(prec, stage):
    """A solve switches its batch to shorter row blocks once few cells run (WS.TAIL: the same
    kernel variants as the built-in shape): each stage's switch changes no bit of any case."""
    tiles = WS.TAIL[prec][stage][1]
    say("case: tail tiles", tiles, "vs built-in tiles, 6 steady 48x48 cases, precision", prec)
    cs = WS.steady_cases(6, seed=4, size=48, re=(10.0, 60.0), da=(3e-3, 1e-1), ri=(0.0, 1.0))
    out = []
    for tail in (None, [(1 << 40, tiles)]):
        b = WS.build_batch(cs, prec)
        (code, it, res, st), ms = b.run_sliced(0.0, 3000, slice_steps=1000, steps_per_graph=100, tail=tail)
        out.append((it, res, st, [{k: b.download(i, k) for k in FIELDS} for i in range(len(cs))], b.get_tiles()))
        b.close()
    (ia, ra, sa, fa, ta), (ib, rb, sb, fb, tb) = out
    say("tiles", ta, tb, "steps", ia.tolist(), "max residual", ra.max(axis=1).tolist())
    assert ta == (0, 0, 0, 0) and tb == tiles
    assert np.array_equal(ia, ib) and np.array_equal(sa, sb) and np.array_equal(ra, rb)
    for i in range(len(cs)):
        for k in FIELDS:
            assert np.array_equal(fa[i][k], fb[i][k]), (i, k)
