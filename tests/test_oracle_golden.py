"""The CPU oracle against the reference's own Fortran (golden vectors) and closed forms.

The golden vectors of tests/golden/ were produced by the flang-built OpenMp_Multicore hot
path (tests/golden/make_golden.py). Both sides compute with one IEEE rounding per
operation in the same order, so the oracle must match them exactly.
"""
import os
import numpy as np
import pytest

import oracle as O
from cfd_amd import cases as C
from cfd_amd.params import make_params, SCHEME_UPWIND, BC_CHANNEL, BUOY_BOUSSINESQ, MU_POLY4, K_POLY4
from helpers import load_golden, small_case, random_state


@pytest.fixture(scope="module", autouse=True)
def _built():
    O.build()


def _oracle(name):
    prm, x, y, eps, lp, st, out, meta = load_golden(name)
    return O.Oracle(prm, x, y, eps, lp), prm, st, out, meta


def test_operators_match_reference_exactly():
    o, prm, st, out, _ = _oracle("ops24")
    um, vm, p, T, Cc = (st[k] for k in ("um", "vm", "p", "T", "C"))
    assert np.array_equal(o.resu(um, vm, p, T), out["ru"])
    assert np.array_equal(o.resv(um, vm, p, T), out["rv"])
    assert np.array_equal(o.resz(um, vm, T), out["rz"])
    assert np.array_equal(o.resc(um, vm, Cc), out["rc"])
    assert np.array_equal(o.upwind_rows_u(um, vm, p, T), out["ru_up"])
    assert np.array_equal(o.upwind_rows_v(um, vm, p, T), out["rv_up"])


def test_boundary_conditions_match_reference_exactly():
    o, prm, st, out, _ = _oracle("ops24")
    u, v = st["um"].copy(), st["vm"].copy()
    o.bcuv(u, v)
    assert np.array_equal(u, out["bc_um"]) and np.array_equal(v, out["bc_vm"])
    for fn, key, src in ((o.bcp, "bc_p", "p"), (o.bcz, "bc_T", "T"), (o.bcc, "bc_C", "C")):
        a = st[src].copy()
        fn(a)
        assert np.array_equal(a, out[key]), key


def test_solves_match_reference_exactly():
    o, prm, st, out, _ = _oracle("ops24")
    s, _cs = o.make_state(st)
    r = o.solve_u(s["um"], s["um_tau"], s["vm_tau"], s["um_n_tau"], s["pn"], s["T"])
    assert r == out["su_res"][0]
    assert np.array_equal(s["um_n_tau"], out["su_um_n_tau"])
    assert np.array_equal(s["vm_tau"], out["su_vm_tau"])          # bcUV side effect
    r = o.solve_v(s["vm"], s["um_tau"], s["vm_tau"], s["vm_n_tau"], s["pn"], s["T"])
    assert r == out["sv_res"][0]
    assert np.array_equal(s["vm_n_tau"], out["sv_vm_n_tau"])
    assert np.array_equal(s["um_tau"], out["sv_um_tau"])
    r = o.solve_p(s["p"], s["um_n_tau"], s["vm_n_tau"], s["pn"])
    assert r == out["sp_res"][0] and np.array_equal(s["pn"], out["sp_pn"])
    o.solve_z(s["um_n_tau"], s["vm_n_tau"], s["T"], s["T_n_tau"], s["T_tau"])
    assert np.array_equal(s["T_n_tau"], out["sz_T_n_tau"])
    o.solve_c(s["um_n_tau"], s["vm_n_tau"], s["C"], s["C_n_tau"], s["C_tau"])
    assert np.array_equal(s["C_n_tau"], out["sc_C_n_tau"])


@pytest.mark.parametrize("name", ["steps24", "steady24", "uniform51_pe100", "quick_porous48", "quick_porous204"])
def test_trajectories_match_reference_exactly(name):
    o, prm, st, out, meta = _oracle(name)
    s, cs = o.make_state(st)
    h = o.pseudo_steps(s, cs, meta["nsteps"])
    assert np.array_equal(h, out["hist"])
    for k in ("um_tau", "vm_tau", "p", "T_tau", "C_tau"):
        assert np.array_equal(s[k], out[k]), k


def test_converged_steady_fixture_is_a_fixed_point_of_the_oracle():
    """steady51_converged: the oracle reproduces the reference's first 1000 steady-mode steps
    exactly, and one more step from the reference's converged fields keeps the residual (the
    largest increment of u, v and p) below the tolerance."""
    o, prm, st, out, meta = _oracle("steady51_converged")
    s, cs = o.make_state(st)
    h = o.pseudo_steps(s, cs, meta["every"])
    assert np.array_equal(h[-1], out["hist_sub"][0])
    conv = {"um": out["um_tau"], "vm": out["vm_tau"], "p": out["p"], "T": out["T_tau"], "C": out["C_tau"]}
    s, cs = o.make_state(conv)
    h = o.pseudo_steps(s, cs, 1)
    assert 0 < h[0].max() < meta["tol"] and h[0].max() <= out["hist_last"].max()


def test_physical_steps_match_reference_exactly():
    o, prm, st, out, meta = _oracle("phys24")
    s, cs = o.make_state(st)
    taken = [o.physical_step(s, cs, meta["eps_conv"], meta["itc_max"], t == 0)[0] for t in range(meta["nphys"])]
    assert taken == list(out["taken"].astype(int))
    assert min(taken) < meta["itc_max"] - 1      # the converged-exit path is exercised
    for k in ("um", "vm", "um_tau", "vm_tau", "p", "pn", "T", "T_tau", "C", "C_tau"):
        assert np.array_equal(s[k], out[k]), k


def test_mesh_matches_reference_exactly():
    prm, x, y, eps, lp, st, out, meta = load_golden("mesh51")
    X = O.mesh_ref_x(51, meta["Lhori"], meta["Px"], meta["Q"])
    assert np.array_equal(X[1:], out["x"][1:])
    # staggered midpoints and areas from the reference's own x, y
    o = O.Oracle(prm, out["x"], out["y"], out["eps"], out["lp"])
    g = o.geometry()
    for k in ("xm", "ym", "aun", "aue", "avn", "ave"):
        assert np.array_equal(g[k][1:], out[k][1:]), k


# ---- closed forms (App. B of the formulation; independent of the golden vectors) -------
def _uniform(n=12, m=10, **kw):
    x = C.uniform_nodes(n, 0.0, 1.1); y = C.uniform_nodes(m, 0.0, 0.9)
    eps = C.zeros("p", n, m); eps[1:, 1:] = 1.0
    lp = C.zeros("p", n, m)
    prm = make_params(n, m, **kw)
    return O.Oracle(prm, x, y, eps, lp), x, y


@pytest.mark.parametrize("scheme", [0, 1])
def test_uniform_flow_has_zero_momentum_residual(scheme):
    """u = U, v = V constant, p constant, no body force, no porosity: every convective
    and diffusive flux balances, so RU = RV = 0 exactly up to rounding."""
    o, x, y = _uniform(scheme=scheme, g=0.0, invfr2=0.0, re=7.0)
    n, m = 12, 10
    um = np.full(C.shape_of("u", n, m), 0.8); vm = np.full(C.shape_of("v", n, m), -0.3)
    p = np.full(C.shape_of("p", n, m), 2.0); T = np.full(C.shape_of("p", n, m), 1.0)
    assert np.abs(o.resu(um, vm, p, T)).max() < 1e-12
    assert np.abs(o.resv(um, vm, p, T)).max() < 1e-12


def test_linear_pressure_gives_exact_gradient():
    """u = v = 0, p = a x + b y: RU = -a (cell eps = 1), RV = -b, from the q_art term."""
    o, x, y = _uniform(g=0.0, invfr2=0.0)
    n, m = 12, 10
    a, b = 0.7, -1.9
    p = np.zeros(C.shape_of("p", n, m))
    p[1:, 1:] = a * x[1:, None] + b * y[None, 1:]
    um = np.zeros(C.shape_of("u", n, m)); vm = np.zeros(C.shape_of("v", n, m))
    T = np.ones(C.shape_of("p", n, m))
    ru = o.resu(um, vm, p, T); rv = o.resv(um, vm, p, T)
    assert np.allclose(ru[2:n + 1, 2:m], -a, rtol=0, atol=1e-12)
    assert np.allclose(rv[2:n, 2:m + 1], -b, rtol=0, atol=1e-12)


def test_body_force_and_buoyancy_forms():
    """Uniform state: RU = -g eps; RV = invfr2 (1 - 1/T) (reference) or Ri (T - Tref)."""
    o, x, y = _uniform(g=2.5, invfr2=0.64)
    n, m = 12, 10
    um = np.zeros(C.shape_of("u", n, m)); vm = np.zeros(C.shape_of("v", n, m))
    p = np.zeros(C.shape_of("p", n, m)); T = np.full(C.shape_of("p", n, m), 1.25)
    assert np.allclose(o.resu(um, vm, p, T)[2:n + 1, 2:m], -2.5, atol=1e-12)
    assert np.allclose(o.resv(um, vm, p, T)[2:n, 2:m + 1], 0.64 * (1 - 1 / 1.25), atol=1e-12)
    o2, _, _ = _uniform(g=0.0, buoyancy=BUOY_BOUSSINESQ, ri=3.0, t_ref=0.5)
    assert np.allclose(o2.resv(um, vm, p, T)[2:n, 2:m + 1], 3.0 * (1.25 - 0.5), atol=1e-12)


def test_quadratic_scalar_diffusion_is_exact():
    """u = v = 0, T = x^2 + 2 y^2 on a uniform grid: RZ = (2 + 4)/Pe at interior nodes."""
    o, x, y = _uniform(pe=4.0)
    n, m = 12, 10
    T = np.zeros(C.shape_of("p", n, m)); T[1:, 1:] = x[1:, None] ** 2 + 2 * y[None, 1:] ** 2
    um = np.zeros(C.shape_of("u", n, m)); vm = np.zeros(C.shape_of("v", n, m))
    rz = o.resz(um, vm, T)
    assert np.allclose(rz[2:n, 2:m], 6.0 / 4.0, rtol=1e-10)


def test_continuity_telescopes():
    """solve_P: sum of RP over the interior equals the net boundary flux, and the three
    RK stages give pn = p + dtau*beta*RP up to rounding (equations.f90:620-658)."""
    o, x, y = _uniform(beta=1.7, dtau=0.05)
    n, m = 12, 10
    st = random_state(n, m, seed=3)
    pn = np.zeros_like(st["p"])
    o.solve_p(st["p"], st["um"], st["vm"], pn)
    g = o.geometry()
    # interior rp from pn
    rp = (pn[2:n, 2:m] - st["p"][2:n, 2:m]) / (0.05 * 1.7)
    um, vm = st["um"], st["vm"]
    inflow = (um[2, 2:m] * g["aue"][2:m]).sum() - (um[n, 2:m] * g["aue"][2:m]).sum() \
        + (vm[2:n, 2] * g["avn"][2:n]).sum() - (vm[2:n, m] * g["avn"][2:n]).sum()
    assert abs(rp.sum() - inflow) < 1e-10


def test_rk3_of_zero_residual_is_identity():
    """R = 0 and phi^n = phi^tau: the three stages return phi^tau unchanged."""
    o, x, y = _uniform(g=0.0, invfr2=0.0)
    n, m = 12, 10
    T = np.full(C.shape_of("p", n, m), 1.0)
    um = np.zeros(C.shape_of("u", n, m)); vm = np.zeros(C.shape_of("v", n, m))
    out = np.zeros_like(T)
    o.solve_z(um, vm, T, out, T)
    assert np.array_equal(out[1:, 1:], T[1:, 1:])


def test_channel_bc_relations():
    """The channel set: inflow u_in, zero-gradient outflow, no-slip walls, p = 0 at the
    outlet, Dirichlet wall temperatures, species slot on the left wall."""
    n, m = 14, 12
    prm, x, y, eps, lp = small_case(n, m, bc=BC_CHANNEL, u_in=1.3, t_bottom=2.0, t_top=-1.0,
                                    slot_lo=0.3, slot_hi=0.6, slot_value=0.9)
    o = O.Oracle(prm, x, y, eps, lp)
    st = random_state(n, m, seed=9)
    u, v = st["um"].copy(), st["vm"].copy()
    o.bcuv(u, v)
    assert np.all(u[1:3, 1:m + 1] == 1.3)
    assert np.all(u[3:n, 1] == 0) and np.all(u[3:n, m] == 0)
    assert np.array_equal(u[n, 2:m], u[n - 1, 2:m]) and np.array_equal(u[n + 1, 2:m], u[n, 2:m])
    assert np.all(v[1:n + 1, 1:3] == 0) and np.all(v[2:n + 1, m:m + 2] == 0) and np.all(v[1, 1:] == 0)
    assert np.array_equal(v[n, 3:m], v[n - 1, 3:m])
    p = st["p"].copy(); o.bcp(p)
    assert np.all(p[n, 1:] == 0) and np.array_equal(p[1, 2:m], p[2, 2:m])
    T = st["T"].copy(); o.bcz(T)
    assert np.all(T[1:n + 1, 1] == 2.0) and np.all(T[1:n + 1, m] == -1.0)
    Cc = st["C"].copy(); o.bcc(Cc)
    inslot = (y[1:] >= 0.3) & (y[1:] <= 0.6)
    assert np.array_equal(Cc[1, 1:], np.where(inslot, 0.9, 0.0))
    # idempotent: applying twice changes nothing
    u2, v2 = u.copy(), v.copy(); o.bcuv(u2, v2)
    assert np.array_equal(u, u2) and np.array_equal(v, v2)


def test_global_upwind_equals_reference_rows():
    """CFD_SCHEME_UPWIND is the reference's upwind row operator on every row."""
    prm, x, y, eps, lp, st, out, meta = load_golden("ops24")
    p2 = make_params(prm.imax, prm.jmax, **{k: getattr(prm, k) for k, _ in type(prm)._fields_
                                           if k not in ("imax", "jmax", "scheme", "mu_coef")},
                     scheme=SCHEME_UPWIND)
    o = O.Oracle(p2, x, y, eps, lp)
    um, vm, pp, T = (st[k] for k in ("um", "vm", "p", "T"))
    assert np.array_equal(o.resu(um, vm, pp, T), out["ru_up"])
    assert np.array_equal(o.resv(um, vm, pp, T), out["rv_up"])


def test_mu_poly_reduces_to_constant():
    """mu(T) with c = (1, 0, 0, 0, 0) is the constant-viscosity reference operator; a
    constant mu = 2 doubles the viscous coefficients exactly as Re/2 does."""
    prm, x, y, eps, lp, st, out, meta = load_golden("ops24")
    base = {k: getattr(prm, k) for k, _ in type(prm)._fields_ if k not in ("imax", "jmax", "mu_coef", "mu_model")}
    um, vm, pp, T = (st[k] for k in ("um", "vm", "p", "T"))
    o1 = O.Oracle(make_params(prm.imax, prm.jmax, mu_model=MU_POLY4, mu_coef=(1, 0, 0, 0, 0), **base), x, y, eps, lp)
    assert np.array_equal(o1.resu(um, vm, pp, T), out["ru"])
    assert np.array_equal(o1.resv(um, vm, pp, T), out["rv"])
    b2 = dict(base); b2["re"] = base["re"] / 2
    o2 = O.Oracle(make_params(prm.imax, prm.jmax, mu_model=MU_POLY4, mu_coef=(2, 0, 0, 0, 0), **base), x, y, eps, lp)
    o3 = O.Oracle(make_params(prm.imax, prm.jmax, **b2), x, y, eps, lp)
    # Darcy drag also uses Re, so compare with drag off
    lp0 = np.zeros_like(lp)
    o2 = O.Oracle(make_params(prm.imax, prm.jmax, mu_model=MU_POLY4, mu_coef=(2, 0, 0, 0, 0), **base), x, y, eps, lp0)
    o3 = O.Oracle(make_params(prm.imax, prm.jmax, **b2), x, y, eps, lp0)
    assert np.allclose(o2.resu(um, vm, pp, T), o3.resu(um, vm, pp, T), rtol=1e-13, atol=1e-13)


def _ops24_without(prm, *drop):
    return {k: getattr(prm, k) for k, _ in type(prm)._fields_ if k not in ("imax", "jmax") + drop}


@pytest.mark.parametrize("k,t", [(0, 2.0), (1, 2.0), (2, 3.0), (3, 2.0), (4, 2.0), (1, 0.5)])
def test_mu_poly_each_power_scales_reynolds(k, t):
    """mu(T) = T^k (one coefficient set) at a uniform temperature t is the constant-viscosity
    operator at Re / t^k: a wrong Horner order, a dropped power or a swapped coefficient
    gives another power of t and fails."""
    prm, x, y, eps, lp, st, out, meta = load_golden("ops24")
    base = _ops24_without(prm, "mu_coef", "mu_model")
    lp0 = np.zeros_like(lp)                      # the Darcy drag also carries Re
    um, vm, pp = (st[q] for q in ("um", "vm", "p"))
    Tu = np.full_like(st["T"], t)
    coef = [0.0] * 5
    coef[k] = 1.0
    om = O.Oracle(make_params(prm.imax, prm.jmax, mu_model=MU_POLY4, mu_coef=coef, **base), x, y, eps, lp0)
    b2 = dict(base)
    b2["re"] = base["re"] / t ** k
    oc = O.Oracle(make_params(prm.imax, prm.jmax, **b2), x, y, eps, lp0)
    for f in ("resu", "resv"):
        a, b = getattr(om, f)(um, vm, pp, Tu), getattr(oc, f)(um, vm, pp, Tu)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.max(np.abs(b))), (f, k, t)
        if k > 0 and t != 1.0:
            # and it is not the unscaled operator
            c = getattr(O.Oracle(make_params(prm.imax, prm.jmax, **base), x, y, eps, lp0), f)(um, vm, pp, Tu)
            assert np.max(np.abs(a - c)) > 1e-6 * np.max(np.abs(c))


def test_mu_poly_reads_the_temperature_of_its_own_cell():
    """A temperature change in one cell (i0, j0) changes RU and RV at that cell only (the
    viscous factor of the u face (i, j) and of the v face (i, j) is mu(T(i, j))); with the
    buoyancy off nothing else reads T."""
    prm, x, y, eps, lp, st, out, meta = load_golden("ops24")
    base = _ops24_without(prm, "mu_coef", "mu_model", "invfr2")
    o = O.Oracle(make_params(prm.imax, prm.jmax, mu_model=MU_POLY4, mu_coef=(0.6, 0.3, 0.2, 0.1, 0.05),
                             invfr2=0.0, **base), x, y, eps, lp)
    um, vm, pp, T = (st[q] for q in ("um", "vm", "p", "T"))
    for (i0, j0) in ((7, 9), (12, 4), (20, 15)):
        T2 = T.copy()
        T2[i0, j0] *= 1.5
        du = np.argwhere(o.resu(um, vm, pp, T2) != o.resu(um, vm, pp, T)).tolist()
        dv = np.argwhere(o.resv(um, vm, pp, T2) != o.resv(um, vm, pp, T)).tolist()
        assert du == [[i0, j0]], (i0, j0, du)
        assert dv == [[i0, j0]], (i0, j0, dv)


def test_mu_poly_reads_the_physical_level_temperature():
    """In a dual-time pseudo-step mu(T) takes T of the physical level (the T that solve_V's
    buoyancy reads, main.f90:124), not the T iterate: changing T_tau leaves the u, v and p
    outputs unchanged, changing T does not."""
    prm, x, y, eps, lp = small_case()
    d = {k: getattr(prm, k) for k, _ in type(prm)._fields_ if k not in ("imax", "jmax", "mu_coef", "mu_model",
                                                                        "invfr2", "b_art")}
    o = O.Oracle(make_params(24, 20, mu_model=MU_POLY4, mu_coef=(0.6, 0.3, 0.2, 0.1, 0.05), invfr2=0.0,
                             b_art=0.0, **d), x, y, eps, lp)
    st = C.perturbed_state(24, 20, seed=31, amp=1e-2, p0=1.0, T0=1.0, C0=0.5, perturb_scalars=True)
    outs = []
    for which in (None, "T_tau", "T"):
        st2 = {k: v.copy() for k, v in st.items()}
        st2["T_tau"] = st["T"].copy()
        if which:
            st2[which] = st2[which] * 1.3
        s, cs = o.make_state(st2)
        o.pseudo_steps(s, cs, 1)
        outs.append({k: s[k].copy() for k in ("um_tau", "vm_tau", "p")})
    for k in ("um_tau", "vm_tau", "p"):
        assert np.array_equal(outs[0][k], outs[1][k]), k
    assert not np.array_equal(outs[0]["um_tau"], outs[2]["um_tau"])


# ---- compat = CUDA: the c_cuda differences, pinned against the reference's own rows and
# closed forms (the CUDA sources cannot be built here) ------------------------------------
def test_cuda_compat_resv_rows_are_reference_rows():
    """c_cuda's RESV (resv.cu QUICK interior i=3..imax-2, j=3..jmax-1, then the strips of
    upwind_V_pair.cu:103-130: rows j=2 and j=jmax-1 over i=2..imax-1, columns i=2 and
    imax-1 over j=2..jmax-1) assembled from the Fortran reference's own outputs: its QUICK
    RESV on rows 3..jmax-2, its upwind row operator upwind_Vi (equations.f90:423-542) on
    the strips, and nothing on row jmax."""
    prm, x, y, eps, lp, st, out, meta = load_golden("ops24")
    um, vm, p, T = (st[k] for k in ("um", "vm", "p", "T"))
    oc = O.Oracle(make_params(prm.imax, prm.jmax, compat=1, **_ops24_without(prm, "mu_coef", "compat")),
                  x, y, eps, lp)
    rv = oc.resv(um, vm, p, T).reshape(prm.imax + 1, prm.jmax + 2)
    quick = out["rv"].reshape(rv.shape)
    up = out["rv_up"].reshape(rv.shape)
    im, jm = prm.imax, prm.jmax
    exp = np.zeros_like(rv)
    exp[3:im - 1, 3:jm - 1] = quick[3:im - 1, 3:jm - 1]
    exp[2:im, 2] = up[2:im, 2]
    exp[2:im, jm - 1] = up[2:im, jm - 1]
    exp[2, 2:jm] = up[2, 2:jm]
    exp[im - 1, 2:jm] = up[im - 1, 2:jm]
    assert np.array_equal(rv, exp)
    # the Fortran operator differs on row jmax-1 (QUICK there) and row jmax (upwind strip)
    assert not np.array_equal(rv[3:im - 1, jm - 1], quick[3:im - 1, jm - 1])
    assert np.any(quick[2:im, jm] != 0) and np.all(rv[:, jm] == 0)


def test_cuda_compat_solve_p_stage_two_skips_row_two():
    """solve_P.cu:16-24 starts stage 2 at j=3, so row j=2 keeps the stage-1 value and the
    update there is pn = p + (4/3) a with a = dtau beta RP, against pn = p + a on every
    other row and in the Fortran stage order (equations.f90:607-665)."""
    prm, x, y, eps, lp, st, out, meta = load_golden("ops24")
    im, jm = prm.imax, prm.jmax
    base = _ops24_without(prm, "mu_coef", "compat")
    g = O.Oracle(prm, x, y, eps, lp).geometry()
    um, vm, p = st["um"], st["vm"], st["p"]
    U = um.reshape(im + 2, jm + 1); V = vm.reshape(im + 1, jm + 2)
    a = np.zeros((im + 1, jm + 1))
    for i in range(2, im):
        for j in range(2, jm):
            rp = -((U[i + 1, j] * g["aue"][j] - U[i, j] * g["aue"][j])
                   + (V[i, j + 1] * g["avn"][i] - V[i, j] * g["avn"][i]))
            a[i, j] = prm.dtau * prm.beta * rp
    P = p.reshape(im + 1, jm + 1)
    for compat in (0, 1):
        o = O.Oracle(make_params(im, jm, compat=compat, **base), x, y, eps, lp)
        pn = np.zeros_like(p)
        r = o.solve_p(p.copy(), um, vm, pn)
        pn = pn.reshape(P.shape)
        c2 = 4.0 / 3.0 if compat else 1.0
        exp = P + a
        exp[2:im, 2] = P[2:im, 2] + c2 * a[2:im, 2]
        scale = np.max(np.abs(P)) + np.max(np.abs(a))
        assert np.max(np.abs(pn[2:im, 2:jm] - exp[2:im, 2:jm])) <= 1e-14 * scale, compat
        if compat == 0:
            assert abs(r - np.max(np.abs(a[2:im, 2:jm]))) <= 1e-15 * np.max(np.abs(a))


def _res_region_case(kind, cell, compat):
    """One solve with dt = 0, where every RK stage's residual is exactly (phi - phi_tau) dtau
    (the operator is multiplied by dt), and phi differs from phi_tau by 1 in one cell."""
    prm, x, y, eps, lp = small_case()
    im, jm = prm.imax, prm.jmax
    d = {k: getattr(prm, k) for k, _ in type(prm)._fields_ if k not in ("imax", "jmax", "mu_coef", "compat", "dt")}
    o = O.Oracle(make_params(im, jm, compat=compat, dt=0.0, **d), x, y, eps, lp)
    st = random_state(im, jm, seed=41)
    um_tau, vm_tau = st["um"].copy(), st["vm"].copy()
    if kind == "u":
        um = um_tau.copy(); um[cell] += 1.0
        return o.solve_u(um, um_tau, vm_tau, np.zeros_like(um), st["p"], st["T"]), prm
    vm = vm_tau.copy(); vm[cell] += 1.0
    return o.solve_v(vm, um_tau, vm_tau, np.zeros_like(vm), st["p"], st["T"]), prm


@pytest.mark.parametrize("kind,n_i,n_j,ld,cells", [
    # c_cuda/solve_U.cu:58 max_reduce(dev_res_u, imax-1, jmax-2), u layout i*(jmax+1)+j
    ("u", lambda im: im - 1, lambda jm: jm - 2, lambda jm: jm + 1, [(3, 5), (19, 8), (20, 12), (23, 10)]),
    # c_cuda/solve_V.cu:52 max_reduce(dev_res_v, imax-1, jmax), v layout i*(jmax+2)+j
    ("v", lambda im: im - 1, lambda jm: jm, lambda jm: jm + 2, [(2, 3), (20, 19), (21, 3), (22, 10), (23, 4)]),
])
def test_cuda_compat_residual_scans_the_first_linear_entries(kind, n_i, n_j, ld, cells):
    """c_cuda reduces max|res| over the first n_i*n_j linear entries of the residual array
    (max_reduce.cu:14-41), not over the update region: a residual in a cell past that count
    is not seen (compat = CUDA), while the Fortran region sees every updated cell."""
    prm, _, _, _, _ = small_case()
    im, jm = prm.imax, prm.jmax
    n = n_i(im) * n_j(jm)
    for cell in cells:
        rf, _ = _res_region_case(kind, cell, 0)
        rc, _ = _res_region_case(kind, cell, 1)
        seen = cell[0] * ld(jm) + cell[1] < n
        assert rf == pytest.approx(prm.dtau, rel=1e-15), cell
        assert rc == (pytest.approx(prm.dtau, rel=1e-15) if seen else 0.0), (cell, rc, n)


def test_cuda_compat_solve_p_residual_scans_the_first_linear_entries():
    """solve_P.cu:50 max_reduce(dev_res_p, imax, jmax): a divergence only in row i=imax-1
    (linear entries from (imax-1)(jmax+1) on, past imax*jmax) is invisible to c_cuda."""
    prm, x, y, eps, lp = small_case()
    im, jm = prm.imax, prm.jmax
    assert (im - 1) * (jm + 1) + 2 >= im * jm
    d = {k: getattr(prm, k) for k, _ in type(prm)._fields_ if k not in ("imax", "jmax", "mu_coef", "compat")}
    for i0, seen in ((im - 1, False), (3, True)):
        um = C.zeros("u", im, jm)
        vm = C.zeros("v", im, jm)
        um[i0 + 1, 7] = 1.0           # RP(i0, 7) = -aue(7), RP(i0 + 1, 7) = +aue(7) if in range
        r = {}
        for compat in (0, 1):
            o = O.Oracle(make_params(im, jm, compat=compat, **d), x, y, eps, lp)
            r[compat] = o.solve_p(C.zeros("p", im, jm), um, vm, C.zeros("p", im, jm))
        a = prm.dtau * prm.beta * O.Oracle(prm, x, y, eps, lp).geometry()["aue"][7]
        assert r[0] == pytest.approx(a, rel=1e-14)
        assert r[1] == (pytest.approx(a, rel=1e-14) if seen else 0.0), (i0, r)


def test_steady_mode_is_dual_time_with_tracking_level():
    """steady=1 equals the dual-time step with the physical level set to the iterate."""
    prm, x, y, eps, lp = small_case()
    st = C.perturbed_state(24, 20, seed=21, amp=1e-2, p0=1.0, T0=1.0, C0=0.5, perturb_scalars=True)
    d = {k: getattr(prm, k) for k, _ in type(prm)._fields_ if k not in ("imax", "jmax", "mu_coef", "steady")}
    o1 = O.Oracle(make_params(24, 20, steady=1, **d), x, y, eps, lp)
    o0 = O.Oracle(make_params(24, 20, steady=0, **d), x, y, eps, lp)
    s1, c1 = o1.make_state(st)
    s0, c0 = o0.make_state(st)
    for _ in range(5):
        o1.pseudo_steps(s1, c1, 1)
        s0["um"][:] = s0["um_tau"]; s0["vm"][:] = s0["vm_tau"]; s0["T"][:] = s0["T_tau"]; s0["C"][:] = s0["C_tau"]
        o0.pseudo_steps(s0, c0, 1)
    for k in ("um_tau", "vm_tau", "p", "T_tau", "C_tau"):
        assert np.array_equal(s1[k], s0[k]), k


def test_residual_l2_closed_forms():
    """The oracle's L2 residual norms (an addition: the reference reports max|r3| only). With
    dt = 0 every RK stage's increment is (phi - phi_tau) dtau, so the u and v norms are
    dtau ||phi - phi_tau|| over the update regions (u: i=3..imax-1, j=2..jmax-1; v: i=2..imax-1,
    j=3..jmax-1); the p norm is ||pn - p|| over i, j = 2..max-1 (solve_P's three stages give
    pn = p + dtau beta RP). A wrong region, a missing square root or a dropped dtau fails."""
    prm, x, y, eps, lp = small_case()
    im, jm = prm.imax, prm.jmax
    d = {k: getattr(prm, k) for k, _ in type(prm)._fields_ if k not in ("imax", "jmax", "mu_coef", "dt")}
    o = O.Oracle(make_params(im, jm, dt=0.0, **d), x, y, eps, lp)
    st = random_state(im, jm, seed=43)
    phys = random_state(im, jm, seed=44)
    s, cs = o.make_state(dict(um=phys["um"], vm=phys["vm"], p=st["p"], T=st["T"], C=st["C"],
                              um_tau=st["um"], vm_tau=st["vm"]))
    p0 = s["p"].copy()
    h = o.pseudo_steps(s, cs, 1)
    l2 = o.last_res_l2()
    dtau = prm.dtau
    du = (phys["um"] - st["um"])[3:im, 2:jm]
    dv = (phys["vm"] - st["vm"])[2:im, 3:jm]
    assert l2[0] == pytest.approx(dtau * np.sqrt(np.sum(du * du)), rel=1e-13)
    assert l2[1] == pytest.approx(dtau * np.sqrt(np.sum(dv * dv)), rel=1e-13)
    # after the step p holds pn (tau update)
    dp = (s["p"] - p0)[2:im, 2:jm]
    assert l2[2] == pytest.approx(np.sqrt(np.sum(dp * dp)), rel=1e-12)
    # and the maxima of the same entries
    assert h[0, 0] == pytest.approx(dtau * np.max(np.abs(du)), rel=1e-15)
    assert h[0, 1] == pytest.approx(dtau * np.max(np.abs(dv)), rel=1e-15)


# ---- k(T): the energy diffusivity k(T)/Pe (k_model K_POLY4; the reference's RESZ has 1/Pe) --
def _kt_params(prm, **kw):
    base = {k: getattr(prm, k) for k, _ in type(prm)._fields_ if k not in ("imax", "jmax", "k_coef", "k_model")}
    base.update(kw)
    return make_params(prm.imax, prm.jmax, **base)


@pytest.mark.parametrize("k,t", [(0, 2.0), (1, 2.0), (2, 3.0), (3, 2.0), (4, 2.0), (1, 0.5)])
def test_conductivity_each_power_scales_peclet(k, t):
    """k(T) = T^k at a uniform temperature t is the constant-diffusivity RESZ at Pe / t^k; a
    wrong Horner order, a dropped power or the iterate in place of the physical-level
    temperature gives another power and fails."""
    prm, x, y, eps, lp, st, out, meta = load_golden("ops24")
    coef = [0.0] * 5
    coef[k] = 1.0
    ok = O.Oracle(_kt_params(prm, k_model=K_POLY4, k_coef=coef), x, y, eps, lp)
    oc = O.Oracle(_kt_params(prm, pe=prm.pe / t ** k), x, y, eps, lp)
    um, vm, T = st["um"], st["vm"], st["T"]
    Tu = np.full_like(T, t)
    a, b = ok.resz_kt(um, vm, T, Tu), oc.resz(um, vm, T)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.max(np.abs(b))), (k, t)
    # the reference's constant form where k(T) = 1
    one = O.Oracle(_kt_params(prm, k_model=K_POLY4, k_coef=(1.0, 0, 0, 0, 0)), x, y, eps, lp)
    assert np.array_equal(one.resz_kt(um, vm, T, T), O.Oracle(prm, x, y, eps, lp).resz(um, vm, T))


def test_conductivity_reads_the_temperature_of_its_own_cell():
    """Changing the temperature of the k(T) evaluation in one cell changes RZ there only."""
    prm, x, y, eps, lp, st, out, meta = load_golden("ops24")
    o = O.Oracle(_kt_params(prm, k_model=K_POLY4, k_coef=(0.8, 0.3, 0.1, 0.02, 0.001)), x, y, eps, lp)
    um, vm, T = st["um"], st["vm"], st["T"]
    for (i0, j0) in ((5, 7), (13, 3), (22, 18)):
        T2 = T.copy()
        T2[i0, j0] *= 1.7
        d = np.argwhere(o.resz_kt(um, vm, T, T2) != o.resz_kt(um, vm, T, T)).tolist()
        assert d == [[i0, j0]], (i0, j0, d)


def test_conductivity_takes_the_physical_level_in_solve_z():
    """In a dual-time pseudo-step k(T) takes the physical-level T (fixed through the stages,
    like mu(T)): solve_Z equals the three stages with RESZ at k(T^n), and changing T_tau
    changes the result only through the transported field."""
    prm, x, y, eps, lp = small_case()
    p2 = _kt_params(prm, k_model=K_POLY4, k_coef=(0.8, 0.3, 0.1, 0.02, 0.001), b_art=0.0)
    o = O.Oracle(p2, x, y, eps, lp)
    st = random_state(prm.imax, prm.jmax, seed=51)
    st2 = random_state(prm.imax, prm.jmax, seed=52)
    um, vm, Tn, Tt = st["um"], st["vm"], st["T"], st2["T"]
    zn = np.zeros_like(Tn)
    o.solve_z(um, vm, Tn, zn, Tt)
    # the same three stages by hand with RESZ at k(T^n)
    dt, dtau = p2.dt, p2.dtau
    im, jm = prm.imax, prm.jmax
    r = o.resz_kt(um, vm, Tt, Tn)
    zi = Tt.copy()
    zi[2:im, 2:jm] = Tt[2:im, 2:jm] + ((Tn - Tt) + r * dt)[2:im, 2:jm] * dtau
    o.bcz(zi)
    r = o.resz_kt(um, vm, zi, Tn)
    zi[2:im, 2:jm] = 0.75 * Tt[2:im, 2:jm] + 0.25 * (zi + ((Tn - Tt) + r * dt) * dtau)[2:im, 2:jm]
    o.bcz(zi)
    r = o.resz_kt(um, vm, zi, Tn)
    z3 = np.zeros_like(Tt)
    z3[2:im, 2:jm] = 1.0 / 3.0 * Tt[2:im, 2:jm] + 2.0 / 3.0 * (zi + ((Tn - Tt) + r * dt) * dtau)[2:im, 2:jm]
    o.bcz(z3)
    assert np.array_equal(zn[1:, 1:], z3[1:, 1:])


def test_conductivity_polynomial_matches_reference_properties():
    """cases.conductivity_poly_ch4_air restates the reference's mixture conductivity
    (properties_CH4.f90) as a polynomial in the nondimensional temperature; evaluated at
    t0 + dT T it equals the reference's own k_tot(Too) (the properties_ch4 fixture, made by
    the flang-built reference routine) over 250-900 K, normalised by k_tot(t0)."""
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "properties_ch4.npz"))
    temps, k_ref = z["temps"], z["props"][:, 0]
    for t0, dT in ((300.0, 300.0), (250.0, 650.0)):
        c = C.conductivity_poly_ch4_air(t0, dT)
        assert c[4] == 0.0 and c[0] == 1.0
        th = (temps - t0) / dT
        k0 = k_ref[np.argmin(np.abs(temps - t0))]
        assert abs(temps[np.argmin(np.abs(temps - t0))] - t0) < 1e-12
        got = k0 * sum(c[n] * th ** n for n in range(5))
        assert np.max(np.abs(got - k_ref) / k_ref) < 1e-13
