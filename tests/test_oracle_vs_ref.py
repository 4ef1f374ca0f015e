"""The CPU oracle against the reference's own Fortran hot path on fresh inputs.

The golden vectors pin the oracle on the inputs they were made from. Where the reference's
Fortran path is built (oracle/_ref/, by __graft_entry__.build() in a checkout that has the
reference), these tests run it on new seeded states and parameters through the harness
(oracle/refharness.py) and require the oracle to reproduce it bit for bit: pseudo-step
trajectories in dual-time and steady mode, and physical steps with the reference's exit
test. Skipped where the reference build is absent.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
from oracle import refharness as H
from cfd_amd import cases as C

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
try:                            # the fixture generator (case geometry and constants); it is
    import make_golden as MG    # not shipped where the reference checkout is absent
except ImportError:
    MG = None

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "ref_t24")
pytestmark = pytest.mark.skipif(MG is None or not os.path.exists(REF),
                                reason="reference Fortran build (oracle/_ref) or its generator absent")


@pytest.fixture(scope="module", autouse=True)
def _built():
    O.build()


FIELDS = (("um_tau", "u"), ("vm_tau", "v"), ("p", "p"), ("T_tau", "p"), ("C_tau", "p"))


@pytest.mark.parametrize("seed,steady,b_art,dt", [(101, 0, 0.0, 0.005), (102, 1, 0.0, 0.03), (103, 0, 0.2, 0.004)])
def test_pseudo_steps_match_reference_on_new_inputs(seed, steady, b_art, dt):
    x, y, eps, lp = MG.grid24()
    im, jm = 24, 20
    st = C.perturbed_state(im, jm, seed=seed, amp=2e-2, p0=0.7, T0=1.1, C0=0.4, perturb_scalars=True)
    hd = dict(dtau=0.05, dt=dt, beta=1.1, b_art=b_art, tinf=1.0)
    n = 150
    buf = H.run(REF, [2, n, 0, 0, steady], [hd["dtau"], hd["dt"], hd["beta"], hd["b_art"], hd["tinf"], 0.0],
                H.geometry_payload(x, y, eps, lp) + H.state_payload(st, im, jm))
    r = H.Reader(buf)
    hist = np.ascontiguousarray(r.f64((3, n)).T)
    ref = {nm: r.field(kind, im, jm) for nm, kind in FIELDS}
    o = O.Oracle(MG.prm_from(MG.T24, steady=steady, **hd), x, y, eps, lp)
    s, cs = o.make_state(st)
    h = o.pseudo_steps(s, cs, n)
    assert np.array_equal(h, hist)
    for nm, _ in FIELDS:
        assert np.array_equal(s[nm], ref[nm]), nm


def test_physical_steps_match_reference_on_new_inputs():
    x, y, eps, lp = MG.grid24()
    im, jm = 24, 20
    st = C.perturbed_state(im, jm, seed=104, amp=2e-2, p0=1.0, T0=1.0, C0=0.5, perturb_scalars=True)
    hd = dict(dtau=0.05, dt=0.004, beta=1.2, b_art=0.0, tinf=1.0)
    nphys, itc_max, eps_conv = 4, 60, 1.5e-3          # the reference takes 59, 60 (cap), 40 and 9 steps
    buf = H.run(REF, [3, 0, nphys, itc_max, 0], [hd["dtau"], hd["dt"], hd["beta"], hd["b_art"], hd["tinf"], eps_conv],
                H.geometry_payload(x, y, eps, lp) + H.state_payload(st, im, jm))
    r = H.Reader(buf)
    taken = list(r.i32(nphys))
    ref = {nm: r.field(kind, im, jm) for nm, kind in (("um", "u"), ("vm", "v"), ("um_tau", "u"), ("vm_tau", "v"),
                                                       ("p", "p"), ("pn", "p"), ("T", "p"), ("T_tau", "p"),
                                                       ("C", "p"), ("C_tau", "p"))}
    o = O.Oracle(MG.prm_from(MG.T24, **hd), x, y, eps, lp)
    s, cs = o.make_state(st)
    got = [o.physical_step(s, cs, eps_conv, itc_max, t == 0)[0] for t in range(nphys)]
    assert got == taken
    assert max(taken) == itc_max and min(taken) < itc_max - 1     # both exits are exercised
    for nm in ref:
        assert np.array_equal(s[nm], ref[nm]), nm
