"""ctypes binding of the CPU oracle (liboracle_cfd.so). TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
as the checker or the timed CPU baseline; the product (libcfd_amd.so and its Python
binding) never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_cfd.so")

_dp = ctypes.POINTER(ctypes.c_double)


def build(force=False):
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous float64"
    return a.ctypes.data_as(_dp)


class OrcState(ctypes.Structure):
    _fields_ = [(n, _dp) for n in ("um", "vm", "um_tau", "vm_tau", "um_n_tau", "vm_n_tau", "p",
                                   "pn", "T", "T_tau", "T_n_tau", "C", "C_tau", "C_n_tau")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.orc_create_full.restype = vp
        L.orc_create_full.argtypes = [vp] + [_dp] * 14
        L.orc_destroy.argtypes = [vp]
        L.orc_set_params.argtypes = [vp, vp]
        L.orc_get_geometry.argtypes = [vp] + [_dp] * 6
        for n in ("orc_resu", "orc_resv"):
            getattr(L, n).argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
        L.orc_resz_kt.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
        for n in ("orc_resz", "orc_resc"):
            getattr(L, n).argtypes = [vp, _dp, _dp, _dp, _dp]
        for n in ("orc_upwind_ui", "orc_upwind_uj", "orc_upwind_vi", "orc_upwind_vj"):
            getattr(L, n).argtypes = [vp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_bcuv.argtypes = [vp, _dp, _dp]
        for n in ("orc_bcp", "orc_bcz", "orc_bcc"):
            getattr(L, n).argtypes = [vp, _dp]
        L.orc_solve_u.restype = ctypes.c_double
        L.orc_solve_u.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_solve_v.restype = ctypes.c_double
        L.orc_solve_v.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_solve_p.restype = ctypes.c_double
        L.orc_solve_p.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.orc_solve_z.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
        L.orc_solve_c.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
        L.orc_pseudo_step.restype = ctypes.c_double
        L.orc_pseudo_step.argtypes = [vp, ctypes.POINTER(OrcState), _dp, ctypes.c_int]
        L.orc_pseudo_steps.argtypes = [vp, ctypes.POINTER(OrcState), ctypes.c_int, _dp]
        L.orc_physical_step.restype = ctypes.c_long
        L.orc_physical_step.argtypes = [vp, ctypes.POINTER(OrcState), ctypes.c_double, ctypes.c_long,
                                        ctypes.c_int, _dp]
        L.orc_mesh_ref_x.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_maxabs.restype = ctypes.c_double
        L.orc_maxabs.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_last_res_l2.argtypes = [vp, _dp]
        _lib = L
    return _lib


class Oracle:
    """One oracle context: parameters + geometry + porosity."""

    def __init__(self, params, x, y, eps, lp, geometry=None):
        L = lib()
        self.params = params
        self.imax, self.jmax = params.imax, params.jmax
        x = np.ascontiguousarray(x, dtype=np.float64); y = np.ascontiguousarray(y, dtype=np.float64)
        self._keep = [x, y]
        eps = np.ascontiguousarray(eps, dtype=np.float64); lp = np.ascontiguousarray(lp, dtype=np.float64)
        if geometry is None:
            self.h = L.orc_create(ctypes.byref(params), _ptr(x), _ptr(y), _ptr(eps), _ptr(lp))
        else:
            g = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in geometry.items()}
            self.h = L.orc_create_full(ctypes.byref(params), _ptr(x), _ptr(y), _ptr(g["xm"]), _ptr(g["ym"]),
                                       _ptr(g["aun"]), _ptr(g["aun"]), _ptr(g["aue"]), _ptr(g["aue"]),
                                       _ptr(g["avn"]), _ptr(g["avn"]), _ptr(g["ave"]), _ptr(g["ave"]),
                                       _ptr(eps), _ptr(lp))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_params(self, params):
        self.params = params
        lib().orc_set_params(self.h, ctypes.byref(params))

    def geometry(self):
        im, jm = self.imax, self.jmax
        g = dict(xm=np.zeros(im + 2), ym=np.zeros(jm + 2), aun=np.zeros(im + 2),
                 aue=np.zeros(jm + 1), avn=np.zeros(im + 1), ave=np.zeros(jm + 2))
        lib().orc_get_geometry(self.h, *[_ptr(g[k]) for k in ("xm", "ym", "aun", "aue", "avn", "ave")])
        return g

    # operators (outputs zero-initialised, only the reference's written region changes)
    def resu(self, um, vm, p, T):
        ru = np.zeros_like(um); lib().orc_resu(self.h, _ptr(um), _ptr(vm), _ptr(p), _ptr(T), _ptr(ru)); return ru

    def resv(self, um, vm, p, T):
        rv = np.zeros_like(vm); lib().orc_resv(self.h, _ptr(um), _ptr(vm), _ptr(p), _ptr(T), _ptr(rv)); return rv

    def resz(self, um, vm, z):
        r = np.zeros_like(z); lib().orc_resz(self.h, _ptr(um), _ptr(vm), _ptr(z), _ptr(r)); return r

    def resz_kt(self, um, vm, z, T):
        """RESZ with the k(T) diffusivity at the temperature T (k_model K_POLY4)."""
        r = np.zeros_like(z); lib().orc_resz_kt(self.h, _ptr(um), _ptr(vm), _ptr(z), _ptr(T), _ptr(r)); return r

    def resc(self, um, vm, c):
        r = np.zeros_like(c); lib().orc_resc(self.h, _ptr(um), _ptr(vm), _ptr(c), _ptr(r)); return r

    def upwind_rows_u(self, um, vm, p, T):
        ru = np.zeros_like(um)
        for j in range(2, self.jmax):
            lib().orc_upwind_ui(self.h, _ptr(um), _ptr(vm), _ptr(p), _ptr(T), _ptr(ru), j)
        return ru

    def upwind_rows_v(self, um, vm, p, T):
        rv = np.zeros_like(vm)
        for j in range(2, self.jmax + 1):
            lib().orc_upwind_vi(self.h, _ptr(um), _ptr(vm), _ptr(p), _ptr(T), _ptr(rv), j)
        return rv

    def bcuv(self, um, vm):
        lib().orc_bcuv(self.h, _ptr(um), _ptr(vm))

    def bcp(self, a):
        lib().orc_bcp(self.h, _ptr(a))

    def bcz(self, a):
        lib().orc_bcz(self.h, _ptr(a))

    def bcc(self, a):
        lib().orc_bcc(self.h, _ptr(a))

    def solve_u(self, um, um_tau, vm_tau, um_n_tau, p, T):
        return lib().orc_solve_u(self.h, _ptr(um), _ptr(um_tau), _ptr(vm_tau), _ptr(um_n_tau), _ptr(p), _ptr(T))

    def solve_v(self, vm, um_tau, vm_tau, vm_n_tau, p, T):
        return lib().orc_solve_v(self.h, _ptr(vm), _ptr(um_tau), _ptr(vm_tau), _ptr(vm_n_tau), _ptr(p), _ptr(T))

    def solve_p(self, p, um_n, vm_n, pn):
        return lib().orc_solve_p(self.h, _ptr(p), _ptr(um_n), _ptr(vm_n), _ptr(pn))

    def solve_z(self, um_n, vm_n, z, z_n_tau, z_tau):
        lib().orc_solve_z(self.h, _ptr(um_n), _ptr(vm_n), _ptr(z), _ptr(z_n_tau), _ptr(z_tau))

    def solve_c(self, um_n, vm_n, c, c_n_tau, c_tau):
        lib().orc_solve_c(self.h, _ptr(um_n), _ptr(vm_n), _ptr(c), _ptr(c_n_tau), _ptr(c_tau))

    # driver
    def make_state(self, st):
        """st: dict with um, vm, p, T, C (+ optional um_tau, vm_tau, pn, T_tau, C_tau)."""
        s = {}
        for k in ("um", "vm", "p", "T", "C"):
            s[k] = np.ascontiguousarray(st[k], dtype=np.float64).copy()
        s["um_tau"] = np.array(st.get("um_tau", s["um"]), dtype=np.float64)
        s["vm_tau"] = np.array(st.get("vm_tau", s["vm"]), dtype=np.float64)
        s["pn"] = np.array(st.get("pn", s["p"]), dtype=np.float64)
        s["T_tau"] = np.array(st.get("T_tau", s["T"]), dtype=np.float64)
        s["C_tau"] = np.array(st.get("C_tau", s["C"]), dtype=np.float64)
        s["um_n_tau"] = np.zeros_like(s["um"]); s["vm_n_tau"] = np.zeros_like(s["vm"])
        s["T_n_tau"] = np.zeros_like(s["T"]); s["C_n_tau"] = np.zeros_like(s["C"])
        cs = OrcState(*[_ptr(s[n]) for n, _ in OrcState._fields_])
        return s, cs

    def last_res_l2(self):
        """L2 norms of the last solve_u / solve_v / solve_p final-stage residuals."""
        r = np.zeros(3)
        lib().orc_last_res_l2(self.h, _ptr(r))
        return r

    def pseudo_steps(self, s, cs, n, history=True):
        h = np.zeros((n, 3)) if history else None
        lib().orc_pseudo_steps(self.h, ctypes.byref(cs), n, _ptr(h) if history else None)
        return h

    def physical_step(self, s, cs, eps, itc_max, first):
        r = np.zeros(3)
        n = lib().orc_physical_step(self.h, ctypes.byref(cs), eps, itc_max, 1 if first else 0, _ptr(r))
        return n, r


def mesh_ref_x(n, L, P, Q):
    x = np.zeros(n + 1)
    lib().orc_mesh_ref_x(n, L, P, Q, _ptr(x))
    return x
