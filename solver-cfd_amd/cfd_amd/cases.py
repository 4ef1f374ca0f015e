"""Seeded inputs, array layouts and case builders for the pseudo-time iteration.

Layouts (include/cfd.h; the reference C layout of c_cuda/main.cu:59-89):
  p-like: 2-D array (imax+1, jmax+1), element (i, j) at [i, j], index 0 is padding
  u     : 2-D array (imax+2, jmax+1)
  v     : 2-D array (imax+1, jmax+2)
Fortran arrays (1:n, 1:m) are column-major; ``to_fortran``/``from_fortran`` convert.

The random generator is SplitMix64 (Steele, Lea, Flood 2014) mapped to doubles in [0, 1)
by the top 53 bits, so every input of the tests and benchmarks is reproducible from a seed
on any host.
"""
import math
import numpy as np

from .params import make_params, BC_CHANNEL, BUOY_BOUSSINESQ, SCHEME_UPWIND, MU_POLY4

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n):
    """n consecutive SplitMix64 outputs for the given seed (uint64 array)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, a=0.0, b=1.0):
    """n doubles uniform on [a, b) from SplitMix64(seed)."""
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return a + (b - a) * u


# ---- layouts --------------------------------------------------------------------------
def shape_of(kind, imax, jmax):
    return {"p": (imax + 1, jmax + 1), "u": (imax + 2, jmax + 1), "v": (imax + 1, jmax + 2)}[kind]


def fortran_shape(kind, imax, jmax):
    return {"p": (imax, jmax), "u": (imax + 1, jmax), "v": (imax, jmax + 1)}[kind]


def zeros(kind, imax, jmax, dtype=np.float64):
    return np.zeros(shape_of(kind, imax, jmax), dtype=dtype)


def to_fortran(a, kind, imax, jmax):
    n, m = fortran_shape(kind, imax, jmax)
    return np.asfortranarray(a[1:n + 1, 1:m + 1])


def from_fortran(f, kind, imax, jmax):
    n, m = fortran_shape(kind, imax, jmax)
    a = zeros(kind, imax, jmax)
    a[1:n + 1, 1:m + 1] = np.asarray(f).reshape((n, m), order="F")
    return a


FIELD_KIND = {"um": "u", "vm": "v", "um_tau": "u", "vm_tau": "v", "um_n_tau": "u",
              "vm_n_tau": "v", "p": "p", "pn": "p", "T": "p", "T_tau": "p", "T_n_tau": "p",
              "C": "p", "C_tau": "p", "C_n_tau": "p", "eps": "p", "lp": "p"}


# ---- geometry ---------------------------------------------------------------------------
def uniform_nodes(n, a=0.0, b=1.0):
    """x[1..n] = a + (i-1)(b-a)/(n-1); x[0] = 0 (padding)."""
    x = np.zeros(n + 1)
    x[1:] = a + (np.arange(n, dtype=np.float64)) * ((b - a) / (n - 1))
    return x


def tanh_nodes(n, a, b, P, Q):
    """Nodes on [a, b] clustered toward both ends by the reference stretch
    s = P*eta + (1-P)*(1 - tanh(Q(1-eta))/tanh(Q)) on each half
    (OpenMp_Multicore/nonsymetric_mesh.f90:17-29, restated for a general interval)."""
    x = uniform_nodes(n, a, b)
    c = n // 2 + 1
    xn, x0 = x[n], x[1]
    # x[c] is re-read each pass: the reference loops start at the split node and update it
    for i in range(c, n + 1):
        eta = (x[i] - xn) / (x[c] - xn)
        s = P * eta + (1.0 - P) * (1.0 - math.tanh(Q * (1.0 - eta)) / math.tanh(Q))
        x[i] = xn - s * (xn - x[c])
    for i in range(c, 0, -1):
        eta = (x[i] - x0) / (x[c] - x0)
        s = P * eta + (1.0 - P) * (1.0 - math.tanh(Q * (1.0 - eta)) / math.tanh(Q))
        x[i] = x0 + s * (x[c] - x0)
    return x


def derive_geometry(x, y):
    """Staggered midpoints and face areas (nonsymetric_mesh.f90:48-101), C-ABI sizes."""
    imax, jmax = len(x) - 1, len(y) - 1
    xm = np.zeros(imax + 2); ym = np.zeros(jmax + 2)
    xm[1] = -(x[2] + x[1]) * 0.5
    xm[imax + 1] = (x[imax] + x[imax - 1]) * 0.5 + (x[imax] - x[imax - 1])
    xm[2:imax + 1] = (x[2:imax + 1] + x[1:imax]) * 0.5
    ym[1] = y[1] - (y[2] - y[1]) * 0.5
    ym[jmax + 1] = (y[jmax] + y[jmax - 1]) * 0.5 + (y[jmax] - y[jmax - 1])
    ym[2:jmax + 1] = (y[2:jmax + 1] + y[1:jmax]) * 0.5
    aun = np.zeros(imax + 2); aue = np.zeros(jmax + 1)
    avn = np.zeros(imax + 1); ave = np.zeros(jmax + 2)
    aun[2:imax + 1] = x[2:imax + 1] - x[1:imax]
    aue[1:jmax + 1] = ym[2:jmax + 2] - ym[1:jmax + 1]
    avn[1:imax + 1] = xm[2:imax + 2] - xm[1:imax + 1]
    ave[2:jmax + 1] = y[2:jmax + 1] - y[1:jmax]
    return dict(xm=xm, ym=ym, aun=aun, aue=aue, avn=avn, ave=ave)


def porous_block(x, y, x0, x1, y0, y1, eps_lo, eps_hi, seed):
    """eps, lp (p layout): cells with x0<=x<=x1 and y0<=y<=y1 get lp=1 and a porosity
    drawn per cell from U[eps_lo, eps_hi); all other cells eps=1, lp=0."""
    imax, jmax = len(x) - 1, len(y) - 1
    eps = zeros("p", imax, jmax); lp = zeros("p", imax, jmax)
    eps[1:, 1:] = 1.0
    X, Y = np.meshgrid(x[1:], y[1:], indexing="ij")
    inside = (X >= x0) & (X <= x1) & (Y >= y0) & (Y <= y1)
    draws = uniform(seed, imax * jmax, eps_lo, eps_hi).reshape(imax, jmax)
    eps[1:, 1:] = np.where(inside, draws, 1.0)
    lp[1:, 1:] = np.where(inside, 1.0, 0.0)
    return eps, lp


# ---- state ------------------------------------------------------------------------------
def perturbed_state(imax, jmax, seed, amp=1e-3, u0=0.0, v0=0.0, p0=0.0, T0=1.0, C0=0.0,
                    perturb_scalars=False):
    """u=u0, v=v0, p=p0 (+/- amp uniform per DOF, seed), T=T0, C=C0 on every entry of the
    reference's (1-based) index range."""
    st = {}
    n = 0
    for name, kind, base in (("um", "u", u0), ("vm", "v", v0), ("p", "p", p0),
                             ("T", "p", T0), ("C", "p", C0)):
        a = zeros(kind, imax, jmax)
        fs = fortran_shape(kind, imax, jmax)
        cnt = fs[0] * fs[1]
        pert = name in ("um", "vm", "p") or perturb_scalars
        vals = base + (uniform(seed * 1000003 + n, cnt, -amp, amp) if pert else 0.0)
        a[1:fs[0] + 1, 1:fs[1] + 1] = np.reshape(vals, fs) if pert else base
        st[name] = a
        n += 1
    return st


def full_state(base):
    """Expand (um, vm, p, T, C) to the 10 input fields of a pseudo-step with the
    unset levels pinned: um_tau=um, vm_tau=vm, T_tau=T, C_tau=C (pn=p is implied)."""
    st = {k: np.array(v, copy=True) for k, v in base.items()}
    st["um_tau"] = st["um"].copy(); st["vm_tau"] = st["vm"].copy()
    st["T_tau"] = st["T"].copy(); st["C_tau"] = st["C"].copy()
    return st


def stable_dt(x, y, nu, kappa, umax, dtau, safety=0.4):
    """Physical step dt that keeps the explicit pseudo-time update stable.

    The pseudo-step advances d(phi)/d(tau) = (phi^n - phi) + dt R(phi) with SSP-RK3 over
    dtau, so dtau (1 + dt |lambda_max(R)|) must stay inside RK3's stability interval (~2.5).
    |lambda_max| is bounded by the diffusive 4 D (1/dx^2 + 1/dy^2) plus the convective
    2 |u| (1/dx + 1/dy) of the smallest cell."""
    dx = float(np.min(np.diff(x[1:]))); dy = float(np.min(np.diff(y[1:])))
    lam = 4.0 * max(nu, kappa) * (1.0 / dx ** 2 + 1.0 / dy ** 2) + 2.0 * umax * (1.0 / dx + 1.0 / dy)
    return (safety * 2.5 / dtau - 1.0) / lam


# ---- benchmark configurations ---------------------------------------------------------
def case_channel_upwind():
    """51x51 unit square, left inflow u=1, right outflow, no-slip walls, Re=100, beta=1,
    global first-order upwind, no porosity/buoyancy/species; seeded +/-1e-3 perturbation."""
    n = 51
    x = uniform_nodes(n); y = uniform_nodes(n)
    prm = make_params(n, n, scheme=SCHEME_UPWIND, bc=BC_CHANNEL, re=100.0, pe=100.0, beta=1.0,
                      g=0.0, invfr2=0.0, darcy=1.0, cf=0.0, u_in=1.0,
                      t_bottom=1.0, t_top=1.0, slot_lo=1.0, slot_hi=0.0)
    eps = zeros("p", n, n); eps[1:, 1:] = 1.0
    lp = zeros("p", n, n)
    st = perturbed_state(n, n, seed=1, amp=1e-3)
    return dict(name="channel_upwind_51x51", params=prm, x=x, y=y, eps=eps, lp=lp, state=st)


def ac_beta_limit(x, y, dt, dtau):
    """Largest artificial compressibility beta for which the fastest pressure-velocity mode
    of one SSP-RK3 pseudo-step stays inside the method's stability interval on the
    imaginary axis (|z| <= sqrt(3)). Per stage the u update carries dtau dt eps dp/dx
    (equations.f90:96, 286-297) and the p update dtau beta RP with RP ~ du * dy (the face
    areas, equations.f90:618-622), so the highest checkerboard mode has
    z^2 = 4 dtau^2 dt beta (dy/dx + dx/dy); with the widest aspect ratio of the mesh:
    beta_max = 3 / (4 dtau^2 dt (max dy / min dx + max dx / min dy))."""
    dx = np.diff(np.asarray(x, dtype=np.float64)[1:])
    dy = np.diff(np.asarray(y, dtype=np.float64)[1:])
    r = float(dy.max() / dx.min() + dx.max() / dy.min())
    return 3.0 / (4.0 * dtau * dtau * dt * r)


def case_porous_buoyant(n=204, beta=None):
    """204x204 uniform, QUICK, Re=200, porous middle third eps~U[0.4,0.9] (seed 2),
    Da=1e-3, Cf=0.55, Boussinesq Ri=1 with hot bottom (1) and cold top (0); dual time with
    the reference's dtau = 0.05, dt = 0.005, from zero velocities.

    beta (artificial compressibility): the continuity update pn = p + dtau beta RP carries
    the face areas (RP ~ div(u) dx dy), so the slowest pressure mode of the pseudo-time
    iteration decays at a rate ~ dtau dt beta pi^2 dx dy per step, while the fastest
    pressure-velocity mode stays stable only below ac_beta_limit (3e4 here, independent of
    the grid size). beta = 1 needs more than 2e6 pseudo-steps to reach a residual of 1e-8 at
    204x204 (measured: 6.3e-6 after 2e6); the default, half the limit, converges
    geometrically."""
    x = uniform_nodes(n); y = uniform_nodes(n)
    if beta is None:
        beta = 0.5 * ac_beta_limit(x, y, 0.005, 0.05)
    prm = make_params(n, n, bc=BC_CHANNEL, re=200.0, pe=140.0, beta=beta, g=0.0, darcy=1e-3, cf=0.55,
                      buoyancy=BUOY_BOUSSINESQ, ri=1.0, t_ref=0.5, invfr2=0.0,
                      u_in=1.0, t_bottom=1.0, t_top=0.0, slot_lo=1.0, slot_hi=0.0)
    eps, lp = porous_block(x, y, 1 / 3, 2 / 3, 1 / 3, 2 / 3, 0.4, 0.9, seed=2)
    st = perturbed_state(n, n, seed=2, amp=1e-3, T0=0.5)
    return dict(name="porous_buoyant_%dx%d" % (n, n), params=prm, x=x, y=y, eps=eps, lp=lp, state=st)


def mu_poly_coefficients(seed=3):
    """mu(T)/mu_ref = c0 + c1 T + ... + c4 T^4, a plausible gas fit (mu rises ~T^0.7 over
    T in [0,1] nondimensional), with seeded +/-2% perturbations of each coefficient."""
    base = np.array([1.0, 0.70, -0.12, 0.03, -0.004])
    return tuple(base * (1.0 + uniform(seed, 5, -0.02, 0.02)))


# Thermal conductivity fits of the reference's mixture properties (properties_CH4.f90:13-14,
# 34, 44): air (cubic) and methane (quadratic) in W/(m K) against T in K, mixed with the
# mole fractions X_met = 0.25, X_air = 0.75 (properties_CH4.f90:6). Kept as the reference
# states them: they are the data of that fit.
_K_AIR = (0.0009128206, 9.04946015844389e-5, -0.000000029, 4.63985599705762e-12)
_K_MET = (0.0037188796, 0.000064197, 1.30126895841081e-7)
_X_MET, _X_AIR = 0.25, 0.75


def conductivity_poly_ch4_air(t0=300.0, dT=300.0):
    """Coefficients c0..c4 of k(T)/k(t0) for the nondimensional temperature T, where the
    dimensional one is t0 + dT * T, from the reference's mixture conductivity
    k_tot = X_met k_met + X_air k_air (properties_CH4.f90, evaluated there once at the
    ambient temperature Too). For cfd_params.k_coef with k_model K_POLY4."""
    d = [_X_AIR * _K_AIR[n] + (_X_MET * _K_MET[n] if n < 3 else 0.0) for n in range(4)]
    # k_tot(t0 + dT T) = sum_n d_n (t0 + dT T)^n, expanded in powers of T
    c = [0.0] * 5
    for n, dn in enumerate(d):
        for m in range(n + 1):
            c[m] += dn * math.comb(n, m) * t0 ** (n - m) * dT ** m
    k0 = c[0]
    return tuple(v / k0 for v in c)


def case_full_physics(imax=2048, jmax=2048):
    """Full physics at scale: tanh-clustered mesh (Q=2) toward the walls, QUICK, porous
    middle-third block (eps~U[0.4,0.9]), Boussinesq Ri=2, species slot Z=1 of width 0.1 on
    the left wall, Sc=0.7, mu(T) polynomial (seed 3)."""
    x = tanh_nodes(imax, 0.0, 1.0, 0.0, 2.0)
    y = tanh_nodes(jmax, 0.0, 1.0, 0.0, 2.0)
    mu = mu_poly_coefficients(3)
    nu_max = sum(abs(c) for c in mu) / 200.0
    dt = stable_dt(x, y, nu_max, 1.0 / (200.0 * 0.7), 1.5, 0.05)
    prm = make_params(imax, jmax, dt=dt, bc=BC_CHANNEL, re=200.0, beta=1.0, g=0.0, darcy=1e-3,
                      cf=0.55, buoyancy=BUOY_BOUSSINESQ, ri=2.0, t_ref=0.5, invfr2=0.0,
                      sc=0.7, pe=200.0 * 0.7, u_in=1.0, t_bottom=1.0, t_top=0.0,
                      slot_lo=0.45, slot_hi=0.55, slot_value=1.0,
                      mu_model=MU_POLY4, mu_coef=mu)
    eps, lp = porous_block(x, y, 1 / 3, 2 / 3, 1 / 3, 2 / 3, 0.4, 0.9, seed=2)
    st = perturbed_state(imax, jmax, seed=3, amp=1e-3, T0=0.5)
    return dict(name="full_physics_tanh_%dx%d" % (imax, jmax), params=prm, x=x, y=y, eps=eps, lp=lp, state=st)


# ---- mixed batch of steady cases -----------------------------------------------------------
def steady_dt(x, y, re, pe, da, cf, eps_min, umax=1.5, dtau=0.05, safety=0.4):
    """Pseudo-time scale dt of a steady (steady=1) case: the update is dtau * dt * R(phi), so
    dtau * dt * |lambda_max| must stay inside SSP-RK3's stability interval (~2.5).
    |lambda_max| bounds diffusion 4 max(nu, kappa) (1/dx^2 + 1/dy^2), convection
    2 |u| (1/dx + 1/dy) (both divided by the smallest porosity: inside the porous block the
    transport terms are scaled by 1/eps) and the Darcy-Forchheimer drag eps/(Re Da) +
    Cf |u| / sqrt(Da);
    convective eigenvalues lie near the imaginary axis, where the interval is ~1.7, hence
    the default safety 0.4 (dtau dt |lambda_max| <= 1)."""
    dx = float(np.min(np.diff(x[1:]))); dy = float(np.min(np.diff(y[1:])))
    d = max(1.0 / re, 1.0 / pe)
    lam = 4.0 * d * (1.0 / dx ** 2 + 1.0 / dy ** 2) + 2.0 * umax * (1.0 / dx + 1.0 / dy)
    lam /= eps_min          # the scalar equations divide by L (eps - 1) + 1 = eps in the porous block
    lam += 1.0 / (re * da) + cf * umax / math.sqrt(da)
    return safety * 2.5 / (dtau * lam)


def case_mixed_batch(n=256, seed=4, sizes_i=(64, 128, 256, 512), sizes_j=(64, 128, 256), beta=1.0,
                     safety=0.4):
    """Independent steady cases for the batched solver, all seeded: grid imax from sizes_i,
    jmax from sizes_j (uniform choice), Re log-uniform [10, 500], Da log-uniform [1e-4, 1e-1],
    porosity per cell uniform [0.3, 0.95] inside one random porous rectangle covering 10-40 %
    of the unit square, Boussinesq buoyancy with Ri uniform [0, 5] (hot bottom, cold top),
    species slot Z = 1 of width 0.1 on the inflow wall, Sc = 0.7, channel BC set, zero
    initial fields. Returns a list of dicts (name, params, x, y, eps, lp, state).

    A larger artificial compressibility (beta up to a fraction of ac_beta_limit, which speeds
    up the pressure modes) does not make this mix converge: with beta at 0.03, 0.1 and 0.3 of
    each case's limit no case reaches a residual of 1e-8 within 1e5 pseudo-steps in FP64
    (the residuals stay at 1e-5 to 1: at Re up to 500 and Ri up to 5 with a heated bottom
    wall most cases have no steady state to converge to)."""
    r = uniform(seed, 12 * n).reshape(n, 12)
    out = []
    for c in range(n):
        u = r[c]
        im = int(sizes_i[min(int(u[0] * len(sizes_i)), len(sizes_i) - 1)])
        jm = int(sizes_j[min(int(u[1] * len(sizes_j)), len(sizes_j) - 1)])
        re = 10.0 * (50.0 ** u[2])
        da = 1e-4 * (1e3 ** u[3])
        ri = 5.0 * u[4]
        area = 0.1 + 0.3 * u[5]
        aspect = 0.5 + 1.5 * u[6]                      # width / height of the rectangle
        w = min(math.sqrt(area * aspect), 0.9); h = min(area / w, 0.9)
        x0 = 0.05 + (0.9 - w) * u[7]; y0 = 0.05 + (0.9 - h) * u[8]
        x = uniform_nodes(im); y = uniform_nodes(jm)
        eps, lp = porous_block(x, y, x0, x0 + w, y0, y0 + h, 0.3, 0.95, seed=seed * 100003 + c)
        pe = 0.7 * re
        dt = steady_dt(x, y, re, pe, da, 0.55, 0.3, safety=safety)
        prm = make_params(im, jm, steady=1, bc=BC_CHANNEL, re=re, pe=pe, sc=0.7, beta=beta, darcy=da, cf=0.55,
                          g=0.0, invfr2=0.0, buoyancy=BUOY_BOUSSINESQ, ri=ri, t_ref=0.5, u_in=1.0,
                          t_bottom=1.0, t_top=0.0, slot_lo=0.45, slot_hi=0.55, slot_value=1.0,
                          dt=dt, dtau=0.05)
        st = perturbed_state(im, jm, seed=seed * 7919 + c, amp=0.0, u0=0.0, v0=0.0, p0=0.0, T0=0.5, C0=0.0)
        out.append(dict(name="mixed%03d_%dx%d" % (c, im, jm), params=prm, x=x, y=y, eps=eps, lp=lp, state=st,
                        re=re, da=da, ri=ri))
    return out
