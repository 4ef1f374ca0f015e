"""ctypes binding of libcfd_amd.so (include/cfd.h, include/cfd_legacy.h, include/cfd_fortran.h).

The library is loaded from this package directory; a missing library is an error, never a
silent fallback. ``Context`` wraps one cfd_ctx: numpy arrays in the C layout go in and out,
all computation happens in the HIP kernels.
"""
import ctypes
import os

import numpy as np

from .params import CfdParams, LAYOUT_C, PREC_FP64, FIELD_NAMES

HERE = os.path.dirname(os.path.abspath(__file__))
# CFD_AMD_LIB names another build of the library (default: the in-tree one)
LIB_PATH = os.environ.get("CFD_AMD_LIB") or os.path.join(HERE, "libcfd_amd.so")

CFD_OK = 0
ERRORS = {-1: "CFD_ERR_ARG", -2: "CFD_ERR_HIP", -3: "CFD_ERR_NONFINITE", -4: "CFD_ERR_NOCONV",
          -5: "CFD_ERR_DIVERGED", -6: "CFD_ERR_STATE", -7: "CFD_ERR_UNSUPPORTED"}

# every symbol include/cfd.h and include/cfd_legacy.h declare
EXPORTS_CFD = [
    "cfd_params_default", "cfd_mesh_1d", "cfd_geometry_derive", "cfd_ctx_create", "cfd_ctx_destroy",
    "cfd_ctx_get_params", "cfd_ctx_set_params", "cfd_ctx_set_mesh", "cfd_ctx_upload", "cfd_ctx_download",
    "cfd_ctx_device_ptr", "cfd_ctx_sync_levels", "cfd_pseudo_steps", "cfd_run_to_tol", "cfd_physical_step",
    "cfd_pseudo_steps_graph", "cfd_ctx_last_timing", "cfd_ctx_last_residual_l2", "cfd_ctx_set_residual_l2", "cfd_ctx_set_kernel", "cfd_ctx_set_pipeline", "cfd_ctx_step_counts",
    "cfd_ctx_set_tiles", "cfd_ctx_get_tiles", "cfd_ctx_autotune", "cfd_pseudo_steps_timed",
    "cfd_batch_create", "cfd_batch_destroy", "cfd_batch_set_mesh", "cfd_batch_upload", "cfd_batch_download",
    "cfd_batch_apply_bc", "cfd_batch_set_tiles", "cfd_batch_get_tiles", "cfd_batch_autotune", "cfd_batch_run",
    "cfd_batch_last_timing", "cfd_batch_copy_state", "cfd_batch_resume", "cfd_batch_set_graph", "cfd_batch_stop_case", "cfd_mesh_reference", "cfd_porous_flags_reference",
    "cfd_unet_create", "cfd_unet_destroy", "cfd_unet_num_params", "cfd_unet_get_params", "cfd_unet_set_params",
    "cfd_unet_predict", "cfd_unet_loss_grad", "cfd_unet_train_step", "cfd_unet_last_timing", "cfd_nn_conv3x3", "cfd_nn_conv3x3_wgrad",
    "cfd_nn_flip_weights", "cfd_nn_maxpool2", "cfd_nn_maxpool2_bwd", "cfd_nn_upsample2", "cfd_nn_upsample2_bwd",
    "cfd_nodal_velocity", "cfd_write_state_table", "cfd_read_state_table", "cfd_write_restart", "cfd_read_restart",
    "cfd_write_vtk", "cfd_hbm_copy_bandwidth",
    "cfd_last_error", "cfd_version",
]
EXPORTS_LEGACY = [
    "RESU", "RESV", "RESZ", "RESC", "upwind_U_pair", "upwind_V_pair", "solve_U", "solve_V", "solve_P",
    "solve_Z", "solve_C", "bcUV", "bcP", "bcZ", "bcC", "max_reduce", "alocar_globais", "desalocar_globais",
    "cfd_legacy_init", "cfd_legacy_set_params", "cfd_legacy_set_mesh", "cfd_legacy_context",
    "max_reduce_noalloc",
]
EXPORTS_FORTRAN = [
    "cfd_fortran_setup", "cfd_fortran_finalize", "solve_u_", "solve_v_", "solve_p_", "solve_z_", "solve_c_",
    "resu_", "resv_", "resz_", "resc_", "bcuv_", "bcp_", "bcz_", "bcc_",
]


# U-Net convolutions (input channels, output channels, ReLU), in the library's order
# (cfd_unet.hip build_convs): encoder 1-3, bottom, decoder 3-1 (after concatenation), head
UNET_LAYERS = [(5, 32, 1), (32, 32, 1), (32, 64, 1), (64, 64, 1), (64, 128, 1), (128, 128, 1), (128, 256, 1),
               (256, 256, 1), (384, 128, 1), (128, 128, 1), (192, 64, 1), (64, 64, 1), (96, 32, 1), (32, 32, 1),
               (32, 4, 0)]


class Tiles(ctypes.Structure):
    """cfd_tiles of include/cfd.h."""
    _fields_ = [("rows_uvp", ctypes.c_int32), ("rows_tc", ctypes.c_int32),
                ("waves_uvp", ctypes.c_int32), ("waves_tc", ctypes.c_int32)]

    def as_tuple(self):
        return (self.rows_uvp, self.rows_tc, self.waves_uvp, self.waves_tc)


class CfdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERRORS.get(code, code), msg))
        self.code = code


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


def load():
    """Load libcfd_amd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError("libcfd_amd.so not built: run `python -m cfd_amd.build` in solver-cfd_amd/ "
                                "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    i32, i64, dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    P = ctypes.POINTER(CfdParams)
    sig = {
        "cfd_params_default": (None, [P, i32, i32]),
        "cfd_mesh_1d": (ctypes.c_int, [i32, dbl, dbl, dbl, dbl, _dp]),
        "cfd_geometry_derive": (ctypes.c_int, [i32, i32, _dp, _dp] + [_dp] * 10),
        "cfd_ctx_create": (ctypes.c_int, [P, i32, ctypes.POINTER(_vp)]),
        "cfd_ctx_destroy": (ctypes.c_int, [_vp]),
        "cfd_ctx_get_params": (ctypes.c_int, [_vp, P]),
        "cfd_ctx_set_params": (ctypes.c_int, [_vp, P]),
        "cfd_ctx_set_mesh": (ctypes.c_int, [_vp, _dp, _dp]),
        "cfd_ctx_upload": (ctypes.c_int, [_vp, i32, _dp, i32]),
        "cfd_ctx_download": (ctypes.c_int, [_vp, i32, _dp, i32]),
        "cfd_ctx_device_ptr": (ctypes.c_int, [_vp, i32, ctypes.POINTER(_vp)]),
        "cfd_ctx_sync_levels": (ctypes.c_int, [_vp]),
        "cfd_pseudo_steps": (ctypes.c_int, [_vp, i32, _dp, _dp]),
        "cfd_run_to_tol": (ctypes.c_int, [_vp, dbl, i64, i32, dbl, ctypes.POINTER(i64), _dp]),
        "cfd_physical_step": (ctypes.c_int, [_vp, dbl, i64, i32, ctypes.POINTER(i64), _dp]),
        "cfd_pseudo_steps_graph": (ctypes.c_int, [_vp, i32, i32, _dp]),
        "cfd_ctx_last_timing": (ctypes.c_int, [_vp, _dp]),
        "cfd_ctx_last_residual_l2": (ctypes.c_int, [_vp, _dp]),
        "cfd_ctx_set_residual_l2": (ctypes.c_int, [_vp, i32]),
        "cfd_ctx_set_kernel": (ctypes.c_int, [_vp, i32]),
        "cfd_ctx_set_pipeline": (ctypes.c_int, [_vp, i32]),
        "cfd_ctx_step_counts": (ctypes.c_int, [_vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "cfd_ctx_set_tiles": (ctypes.c_int, [_vp, ctypes.POINTER(Tiles)]),
        "cfd_pseudo_steps_timed": (ctypes.c_int, [_vp, i32, _dp, _dp]),
        "cfd_ctx_get_tiles": (ctypes.c_int, [_vp, ctypes.POINTER(Tiles)]),
        "cfd_ctx_autotune": (ctypes.c_int, [_vp, i32, ctypes.POINTER(Tiles), _dp]),
        "cfd_batch_create": (ctypes.c_int, [i32, P, i32, ctypes.POINTER(_vp)]),
        "cfd_batch_destroy": (ctypes.c_int, [_vp]),
        "cfd_batch_set_mesh": (ctypes.c_int, [_vp, i32, _dp, _dp]),
        "cfd_batch_upload": (ctypes.c_int, [_vp, i32, i32, _dp]),
        "cfd_batch_download": (ctypes.c_int, [_vp, i32, i32, _dp]),
        "cfd_batch_apply_bc": (ctypes.c_int, [_vp]),
        "cfd_batch_set_tiles": (ctypes.c_int, [_vp, ctypes.POINTER(Tiles)]),
        "cfd_batch_get_tiles": (ctypes.c_int, [_vp, ctypes.POINTER(Tiles)]),
        "cfd_batch_autotune": (ctypes.c_int, [_vp, i32, ctypes.POINTER(Tiles), _dp]),
        "cfd_batch_run": (ctypes.c_int, [_vp, dbl, i64, i32, dbl, ctypes.POINTER(i64), _dp, ctypes.POINTER(i32)]),
        "cfd_batch_last_timing": (ctypes.c_int, [_vp, _dp, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "cfd_batch_copy_state": (ctypes.c_int, [_vp, _vp]),
        "cfd_batch_set_graph": (ctypes.c_int, [_vp, i32]),
        "cfd_batch_stop_case": (ctypes.c_int, [_vp, i32, i32]),
        "cfd_batch_resume": (ctypes.c_int, [_vp, dbl, i64, i32, dbl, ctypes.POINTER(i64), _dp, ctypes.POINTER(i32)]),
        "cfd_mesh_reference": (ctypes.c_int, [i32, i32, dbl, dbl, dbl, dbl, dbl, _dp, _dp]),
        "cfd_porous_flags_reference": (ctypes.c_int, [i32, i32, _dp, _dp, dbl, dbl, dbl, _dp, _dp,
                                                      ctypes.POINTER(i32)]),
        "cfd_nodal_velocity": (ctypes.c_int, [i32, i32, _dp, _dp, _dp, _dp]),
        "cfd_write_state_table": (ctypes.c_int, [ctypes.c_char_p, i32, i32] + [_dp] * 7),
        "cfd_read_state_table": (ctypes.c_int, [ctypes.c_char_p, i32, i32] + [_dp] * 7),
        "cfd_write_restart": (ctypes.c_int, [ctypes.c_char_p, i32, i32] + [_dp] * 5),
        "cfd_read_restart": (ctypes.c_int, [ctypes.c_char_p, i32, i32, i32] + [_dp] * 5),
        "cfd_write_vtk": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, i32, i32] + [_dp] * 6),
        "cfd_unet_create": (ctypes.c_int, [i32, i32, i32, ctypes.c_uint64, ctypes.POINTER(_vp)]),
        "cfd_unet_destroy": (ctypes.c_int, [_vp]),
        "cfd_unet_num_params": (ctypes.c_int, [_vp, ctypes.POINTER(i64)]),
        "cfd_unet_get_params": (ctypes.c_int, [_vp, _fp]),
        "cfd_unet_set_params": (ctypes.c_int, [_vp, _fp]),
        "cfd_unet_predict": (ctypes.c_int, [_vp, _fp, i32, _fp]),
        "cfd_unet_loss_grad": (ctypes.c_int, [_vp, _fp, _fp, i32, _dp, _fp]),
        "cfd_unet_train_step": (ctypes.c_int, [_vp, _fp, _fp, i32, ctypes.c_float, _dp]),
        "cfd_unet_last_timing": (ctypes.c_int, [_vp, _dp]),
        "cfd_nn_conv3x3": (ctypes.c_int, [_vp, i32, i32, i32, i32, _vp, _vp, i32, i32, _vp, _vp]),
        "cfd_nn_conv3x3_wgrad": (ctypes.c_int, [_vp, _vp, i32, i32, i32, i32, i32, _vp, _vp]),
        "cfd_nn_flip_weights": (ctypes.c_int, [_vp, i32, i32, _vp]),
        "cfd_nn_maxpool2": (ctypes.c_int, [_vp, i32, i32, i32, i32, _vp]),
        "cfd_nn_maxpool2_bwd": (ctypes.c_int, [_vp, _vp, _vp, i32, i32, i32, i32, _vp]),
        "cfd_nn_upsample2": (ctypes.c_int, [_vp, i32, i32, i32, i32, _vp, i32, i32]),
        "cfd_nn_upsample2_bwd": (ctypes.c_int, [_vp, i32, i32, i32, i32, i32, i32, _vp]),
        "cfd_hbm_copy_bandwidth": (ctypes.c_int, [i64, i32, _dp]),
        "cfd_last_error": (ctypes.c_char_p, []),
        "cfd_version": (ctypes.c_char_p, []),
        "cfd_legacy_init": (ctypes.c_int, [P]),
        "cfd_legacy_set_params": (ctypes.c_int, [P]),
        "cfd_legacy_set_mesh": (ctypes.c_int, [_dp, _dp, _dp, _dp]),
        "cfd_legacy_context": (_vp, []),
        "RESU": (None, [_vp] * 4), "RESV": (None, [_vp] * 5), "RESZ": (None, [_vp] * 4),
        "RESC": (None, [_vp] * 4), "upwind_U_pair": (None, [_vp] * 4), "upwind_V_pair": (None, [_vp] * 5),
        "solve_U": (dbl, [_vp] * 10), "solve_V": (dbl, [_vp] * 11), "solve_P": (dbl, [_vp] * 4),
        "solve_Z": (None, [_vp] * 6), "solve_C": (None, [_vp] * 6),
        "bcUV": (None, [_vp] * 2), "bcP": (None, [_vp]), "bcZ": (None, [_vp]), "bcC": (None, [_vp]),
        "max_reduce": (dbl, [_vp, ctypes.c_int, ctypes.c_int]),
        "max_reduce_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]),
        "alocar_globais": (None, []), "desalocar_globais": (None, []),
        # Fortran entry points: every argument by reference, host arrays in column-major order
        "cfd_fortran_setup": (ctypes.c_int, [P, _dp, _dp, _dp, _dp]), "cfd_fortran_finalize": (None, []),
        "solve_u_": (None, [_dp] * 8), "solve_v_": (None, [_dp] * 9), "solve_p_": (None, [_dp] * 5),
        "solve_z_": (None, [_dp] * 5), "solve_c_": (None, [_dp] * 5),
        "resu_": (None, [_dp] * 4), "resv_": (None, [_dp] * 5), "resz_": (None, [_dp] * 4), "resc_": (None, [_dp] * 4),
        "bcuv_": (None, [_dp] * 2), "bcp_": (None, [_dp]), "bcz_": (None, [_dp]), "bcc_": (None, [_dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def hbm_copy_bandwidth(nbytes=1 << 30, reps=20):
    """Measured copy bandwidth (GB/s, read + write) of the current device."""
    g = ctypes.c_double(0.0)
    check(load().cfd_hbm_copy_bandwidth(int(nbytes), int(reps), ctypes.byref(g)))
    return g.value


def last_error():
    return load().cfd_last_error().decode()


def check(code):
    if code != CFD_OK:
        raise CfdError(code, last_error())
    return code


def _ptr(a):
    assert isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class Context:
    """One solver context (cfd_ctx) on the current HIP device."""

    def __init__(self, params, x, y, eps=None, lp=None, precision=PREC_FP64):
        L = load()
        self.params = params
        self.imax, self.jmax = params.imax, params.jmax
        h = _vp()
        check(L.cfd_ctx_create(ctypes.byref(params), precision, ctypes.byref(h)))
        self.h = h
        x = np.ascontiguousarray(x, dtype=np.float64); y = np.ascontiguousarray(y, dtype=np.float64)
        check(L.cfd_ctx_set_mesh(self.h, _ptr(x), _ptr(y)))
        if eps is not None:
            self.upload("eps", eps)
        if lp is not None:
            self.upload("lp", lp)

    def close(self):
        if getattr(self, "h", None):
            load().cfd_ctx_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _shape(self, name):
        from .cases import shape_of, FIELD_KIND
        return shape_of(FIELD_KIND[name], self.imax, self.jmax)

    def upload(self, name, arr, layout=LAYOUT_C):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if layout == LAYOUT_C and a.shape != self._shape(name):
            raise ValueError("%s: shape %s, expected %s" % (name, a.shape, self._shape(name)))
        check(load().cfd_ctx_upload(self.h, FIELD_NAMES.index(name), _ptr(a), layout))

    def download(self, name, layout=LAYOUT_C):
        if layout == LAYOUT_C:
            out = np.zeros(self._shape(name))
        else:
            from .cases import fortran_shape, FIELD_KIND
            n, m = fortran_shape(FIELD_KIND[name], self.imax, self.jmax)
            out = np.zeros(n * m)
        check(load().cfd_ctx_download(self.h, FIELD_NAMES.index(name), _ptr(out), layout))
        return out

    def set_state(self, st):
        """st: dict of any of um, vm, um_tau, vm_tau, p, T, T_tau, C, C_tau (C layout).
        Missing iterates are pinned to their physical level (um_tau := um, ...)."""
        for k in ("um", "vm", "p", "T", "C"):
            if k in st:
                self.upload(k, st[k])
        for k, base in (("um_tau", "um"), ("vm_tau", "vm"), ("T_tau", "T"), ("C_tau", "C")):
            src = st.get(k, st.get(base))
            if src is not None:
                self.upload(k, src)

    def set_params(self, params):
        check(load().cfd_ctx_set_params(self.h, ctypes.byref(params)))
        self.params = params

    def set_kernel(self, k):
        check(load().cfd_ctx_set_kernel(self.h, int(k)))

    def set_pipeline(self, on):
        """Two-stream pipeline of the fused path (unsteady problems): on (default) or off."""
        check(load().cfd_ctx_set_pipeline(self.h, 1 if on else 0))

    def pseudo_steps(self, n, history=False):
        res = np.zeros(3)
        hist = np.zeros((n, 3)) if history else None
        check(load().cfd_pseudo_steps(self.h, int(n), _ptr(res), _ptr(hist) if history else None))
        return (res, hist) if history else res

    def pseudo_steps_graph(self, steps_per_graph, replays):
        res = np.zeros(3)
        check(load().cfd_pseudo_steps_graph(self.h, int(steps_per_graph), int(replays), _ptr(res)))
        return res

    def run_to_tol(self, tol, itc_max, check_every=100, diverge_factor=0.0, raise_on_fail=True):
        it = ctypes.c_int64(0)
        res = np.zeros(3)
        code = load().cfd_run_to_tol(self.h, float(tol), int(itc_max), int(check_every), float(diverge_factor),
                                     ctypes.byref(it), _ptr(res))
        if raise_on_fail:
            check(code)
        return code, it.value, res

    def physical_step(self, eps, itc_max, first):
        it = ctypes.c_int64(0)
        res = np.zeros(3)
        check(load().cfd_physical_step(self.h, float(eps), int(itc_max), 1 if first else 0, ctypes.byref(it),
                                       _ptr(res)))
        return it.value, res

    def last_ms_per_step(self):
        v = ctypes.c_double(0)
        check(load().cfd_ctx_last_timing(self.h, ctypes.byref(v)))
        return v.value

    def set_residual_l2(self, on):
        """Fold the L2 norms of the final-stage increments from the next step on."""
        check(load().cfd_ctx_set_residual_l2(self.h, 1 if on else 0))

    def last_residual_l2(self):
        """L2 norms of the last pseudo-step's final-stage increments of u, v, p."""
        r = np.zeros(3)
        check(load().cfd_ctx_last_residual_l2(self.h, _ptr(r)))
        return r

    def step_counts(self):
        """(fused, exact): pseudo-steps taken by each path so far."""
        f, e = ctypes.c_int64(0), ctypes.c_int64(0)
        check(load().cfd_ctx_step_counts(self.h, ctypes.byref(f), ctypes.byref(e)))
        return f.value, e.value

    def pseudo_steps_timed(self, n):
        """n pseudo-steps with per-launch events: returns (res, ms) where ms = mean ms of the
        u/v/p kernel, the T/C kernel, the boundary pass and the whole fused step."""
        res, ms = np.zeros(3), np.zeros(4)
        check(load().cfd_pseudo_steps_timed(self.h, int(n), _ptr(res), _ptr(ms)))
        return res, ms

    def sync_levels(self):
        """Copy the physical levels into the pseudo-time iterates (um_tau := um, vm_tau := vm,
        T_tau := T, C_tau := C); the next pseudo-step starts like the first one."""
        check(load().cfd_ctx_sync_levels(self.h))

    def set_tiles(self, rows_uvp=0, rows_tc=0, waves_uvp=0, waves_tc=0):
        """Shapes of the fused u/v/p and T/C kernels (zeros: built-in choice)."""
        t = Tiles(int(rows_uvp), int(rows_tc), int(waves_uvp), int(waves_tc))
        check(load().cfd_ctx_set_tiles(self.h, ctypes.byref(t)))

    def get_tiles(self):
        t = Tiles()
        check(load().cfd_ctx_get_tiles(self.h, ctypes.byref(t)))
        return t.as_tuple()

    def autotune(self, steps=10):
        """Time candidate shapes on a snapshot of the state (restored afterwards); returns
        ((rows_uvp, rows_tc, waves_uvp, waves_tc), best ms per step)."""
        t = Tiles()
        ms = np.zeros(1)
        check(load().cfd_ctx_autotune(self.h, int(steps), ctypes.byref(t), _ptr(ms)))
        return t.as_tuple(), float(ms[0])

    def device_ptr(self, name):
        p = _vp()
        check(load().cfd_ctx_device_ptr(self.h, FIELD_NAMES.index(name), ctypes.byref(p)))
        return p.value


from .cases import shape_of, FIELD_KIND  # noqa: E402


class Batch:
    """Many independent cases advanced together (include/cfd.h, cfd_batch_*)."""

    def __init__(self, params_list, precision=PREC_FP64):
        L = load()
        n = len(params_list)
        arr = (CfdParams * n)(*params_list)
        h = _vp()
        check(L.cfd_batch_create(n, arr, precision, ctypes.byref(h)))
        self.h = h
        self.params = list(params_list)
        self.n = n

    def set_mesh(self, c, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64); y = np.ascontiguousarray(y, dtype=np.float64)
        check(load().cfd_batch_set_mesh(self.h, int(c), _ptr(x), _ptr(y)))

    def upload(self, c, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        P = self.params[c]
        want = shape_of(FIELD_KIND[name], P.imax, P.jmax)
        if a.shape != want:
            raise ValueError("%s of case %d: shape %s, expected %s" % (name, c, a.shape, want))
        check(load().cfd_batch_upload(self.h, int(c), FIELD_NAMES.index(name), _ptr(a)))

    def download(self, c, name):
        P = self.params[c]
        a = np.zeros(shape_of(FIELD_KIND[name], P.imax, P.jmax))
        check(load().cfd_batch_download(self.h, int(c), FIELD_NAMES.index(name), _ptr(a)))
        return a

    def set_state(self, c, st):
        """Current iterates of case c from a dict (um, vm, p, T, C and/or the *_tau names)."""
        for k in ("um", "vm", "p", "T", "C"):
            if k in st:
                self.upload(c, k, st[k])
        for k, base in (("um_tau", "um"), ("vm_tau", "vm"), ("T_tau", "T"), ("C_tau", "C")):
            src = st.get(k, st.get(base))
            if src is not None:
                self.upload(c, k, src)

    def apply_bc(self):
        check(load().cfd_batch_apply_bc(self.h))

    def set_tiles(self, rows_uvp=0, rows_tc=0, waves_uvp=0, waves_tc=0):
        t = Tiles(int(rows_uvp), int(rows_tc), int(waves_uvp), int(waves_tc))
        check(load().cfd_batch_set_tiles(self.h, ctypes.byref(t)))

    def get_tiles(self):
        t = Tiles()
        check(load().cfd_batch_get_tiles(self.h, ctypes.byref(t)))
        return t.as_tuple()

    def autotune(self, steps=10):
        t = Tiles()
        ms = np.zeros(1)
        check(load().cfd_batch_autotune(self.h, int(steps), ctypes.byref(t), _ptr(ms)))
        return t.as_tuple(), float(ms[0])

    def run(self, tol, itc_max, steps_per_graph=50, diverge_factor=0.0, raise_on_fail=False, resume=False):
        """Returns (code, iters[n], res[n, 3], status[n]). resume=True continues the last run
        (cases stopped at itc_max go on up to the new total itc_max)."""
        it = (ctypes.c_int64 * self.n)()
        st = (ctypes.c_int32 * self.n)()
        res = np.zeros((self.n, 3))
        fn = load().cfd_batch_resume if resume else load().cfd_batch_run
        code = fn(self.h, float(tol), int(itc_max), int(steps_per_graph), float(diverge_factor), it, _ptr(res), st)
        if raise_on_fail:
            check(code)
        return code, np.array(it[:], dtype=np.int64), res, np.array(st[:], dtype=np.int32)

    def set_graph(self, on):
        """True (default): runs replay a captured hipGraph; False: direct launches."""
        check(load().cfd_batch_set_graph(self.h, 1 if on else 0))

    def stop_case(self, c, status=5):
        """Stops case c (status >= 5) until the next run(); resume() leaves it stopped."""
        check(load().cfd_batch_stop_case(self.h, int(c), int(status)))

    def copy_state_from(self, other):
        """Take every case's current iterates from another batch of the same cases
        (precision converted on the device)."""
        check(load().cfd_batch_copy_state(self.h, other.h))

    def run_sliced(self, tol, itc_max, slice_steps=10000, steps_per_graph=500, diverge_factor=0.0, log=None,
                   stall=None, tail=None):
        """run() in slices of slice_steps (progress through `log`); returns the final
        (code, iters, res, status) and the summed event time in ms. stall = (window, ratio):
        between slices, a running case whose max(res_u, res_v, res_p) has not fallen below
        ratio x its value `window` pseudo-steps earlier (window a multiple of slice_steps) is
        stopped with status 5 (stalled). tail = [(cells, tiles), ...]: between slices, once
        the running cases hold no more than `cells` interior cells in all, the batch switches
        to that stage's `tiles` (rows_uvp, rows_tc, waves_uvp, waves_tc), the stage with the
        smallest such limit winning; a stage is taken once, and never after a smaller one:
        the cases left fill the GPU only with shorter row blocks than the built-in ones.
        Row heights alone do not change any case's results (rows are computed in pairs
        from even rows whatever the block), kernel variants (waves) can."""
        if stall is not None and (stall[0] <= 0 or stall[0] % slice_steps):
            raise ValueError("stall window %d must be a positive multiple of slice_steps %d" % (stall[0], slice_steps))
        total_ms, done, first = 0.0, 0, True
        seen = {}                       # slice end -> per-case max residual
        stages = sorted(tail or [], key=lambda s: -s[0])
        interior = np.array([(P.imax - 2) * (P.jmax - 2) for P in self.params], dtype=np.int64)
        while True:
            done = min(itc_max, done + slice_steps)
            out = self.run(tol, done, steps_per_graph, diverge_factor, resume=not first)
            first = False
            ms, steps, cells = self.last_timing()
            total_ms += ms
            code, it, res, st = out
            if stall is not None and done < itc_max:
                window, ratio = stall
                now = res.max(axis=1)
                seen[done] = now
                before = seen.get(done - window)
                if before is not None:
                    for c in np.nonzero((st == 4) & ~(now < ratio * before))[0]:
                        self.stop_case(int(c), 5)
                        st[c] = 5
            live = (st == 4) | (st == 0)
            running = int(live.sum())
            if stages and running and done < itc_max:
                now_cells = int(interior[live].sum())
                take = [k for k, (lim, _) in enumerate(stages) if now_cells <= lim]
                if take:
                    self.set_tiles(*stages[take[-1]][1])
                    stages = stages[take[-1] + 1:]
            if log:
                log("  %d steps: %d running, %d converged, %d failed, %d stalled, %.1f s"
                    % (done, running, int((st == 1).sum()), int(((st == 2) | (st == 3)).sum()),
                       int((st == 5).sum()), total_ms * 1e-3))
            if done >= itc_max or not running:
                return out, total_ms

    def last_timing(self):
        ms = ctypes.c_double(0); steps = ctypes.c_int64(0); cells = ctypes.c_int64(0)
        check(load().cfd_batch_last_timing(self.h, ctypes.byref(ms), ctypes.byref(steps), ctypes.byref(cells)))
        return ms.value, steps.value, cells.value

    def close(self):
        if self.h:
            load().cfd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- mesh, flags, nodal fields and state files --------------------------------------------
def mesh_reference(imax, jmax, lhori=12.0, y_down=5.0, px=1.0, py=1.6, q=1.0):
    """x[0..imax], y[0..jmax] of the reference mesh routine (index 0 unused)."""
    x = np.zeros(imax + 1); y = np.zeros(jmax + 1)
    check(load().cfd_mesh_reference(int(imax), int(jmax), float(lhori), float(y_down), float(px), float(py),
                                     float(q), _ptr(x), _ptr(y)))
    return x, y


def porous_flags_reference(x, y, x_edge=-4.0, y_edge=0.5, porosity=0.5):
    imax, jmax = len(x) - 1, len(y) - 1
    eps = np.zeros((imax + 1, jmax + 1)); lp = np.zeros_like(eps)
    fl = np.zeros((imax + 1, jmax + 1), dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64); y = np.ascontiguousarray(y, dtype=np.float64)
    check(load().cfd_porous_flags_reference(imax, jmax, _ptr(x), _ptr(y), float(x_edge), float(y_edge),
                                            float(porosity), _ptr(eps), _ptr(lp),
                                            fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    return eps, lp, fl


def nodal_velocity(um, vm):
    imax, jmax = um.shape[0] - 2, um.shape[1] - 1
    u = np.zeros((imax + 1, jmax + 1)); v = np.zeros_like(u)
    um = np.ascontiguousarray(um, dtype=np.float64); vm = np.ascontiguousarray(vm, dtype=np.float64)
    check(load().cfd_nodal_velocity(imax, jmax, _ptr(um), _ptr(vm), _ptr(u), _ptr(v)))
    return u, v


def write_state_table(path, x, y, u, v, p, T, C):
    imax, jmax = len(x) - 1, len(y) - 1
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, u, v, p, T, C)]
    check(load().cfd_write_state_table(path.encode(), imax, jmax, *[_ptr(a) for a in arrs]))


def read_state_table(path, imax, jmax):
    x = np.zeros(imax + 1); y = np.zeros(jmax + 1)
    f = [np.zeros((imax + 1, jmax + 1)) for _ in range(5)]
    check(load().cfd_read_state_table(path.encode(), imax, jmax, _ptr(x), _ptr(y), *[_ptr(a) for a in f]))
    return dict(x=x, y=y, u=f[0], v=f[1], p=f[2], T=f[3], C=f[4])


def write_restart(d, um, vm, p, T, C):
    imax, jmax = um.shape[0] - 2, um.shape[1] - 1
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (um, vm, p, T, C)]
    check(load().cfd_write_restart(d.encode(), imax, jmax, *[_ptr(a) for a in arrs]))


def write_vtk(path, x, y, u, v, p, T, title=None):
    """ParaView structured-grid file of nodal u, v, p, T (p-layout arrays), transient.f90's
    format."""
    imax, jmax = len(x) - 1, len(y) - 1
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, u, v, p, T)]
    check(load().cfd_write_vtk(path.encode(), title.encode() if title is not None else None, imax, jmax,
                               *[_ptr(a) for a in arrs]))


def read_restart(d, imax, jmax, order=0):
    um = np.zeros((imax + 2, jmax + 1)); vm = np.zeros((imax + 1, jmax + 2))
    p, T, C = (np.zeros((imax + 1, jmax + 1)) for _ in range(3))
    check(load().cfd_read_restart(d.encode(), imax, jmax, int(order), _ptr(um), _ptr(vm), _ptr(p), _ptr(T), _ptr(C)))
    return dict(um=um, vm=vm, p=p, T=T, C=C)


class UNet:
    """The warm-start U-Net (include/cfd.h, cfd_unet_*): host float32 arrays in and out."""

    def __init__(self, height, width, max_batch, seed=5):
        h = _vp()
        check(load().cfd_unet_create(int(height), int(width), int(max_batch), int(seed), ctypes.byref(h)))
        self.h = h
        self.H, self.W, self.max_batch = height, width, max_batch

    def num_params(self):
        n = ctypes.c_int64(0)
        check(load().cfd_unet_num_params(self.h, ctypes.byref(n)))
        return n.value

    def get_params(self):
        a = np.zeros(self.num_params(), dtype=np.float32)
        check(load().cfd_unet_get_params(self.h, a.ctypes.data_as(_fp)))
        return a

    def set_params(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        check(load().cfd_unet_set_params(self.h, a.ctypes.data_as(_fp)))

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(x.shape[:3] + (4,), dtype=np.float32)
        check(load().cfd_unet_predict(self.h, x.ctypes.data_as(_fp), x.shape[0], y.ctypes.data_as(_fp)))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32); t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.num_params(), dtype=np.float32)
        loss = ctypes.c_double(0)
        check(load().cfd_unet_loss_grad(self.h, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp), x.shape[0],
                                        ctypes.byref(loss), g.ctypes.data_as(_fp)))
        return loss.value, g

    def train_step(self, x, t, lr):
        x = np.ascontiguousarray(x, dtype=np.float32); t = np.ascontiguousarray(t, dtype=np.float32)
        loss = ctypes.c_double(0)
        check(load().cfd_unet_train_step(self.h, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp), x.shape[0],
                                         ctypes.c_float(lr), ctypes.byref(loss)))
        return loss.value

    def last_timing(self):
        """Device ms of the last train_step: (forward, backward, update)."""
        ms = (ctypes.c_double * 3)()
        check(load().cfd_unet_last_timing(self.h, ms))
        return tuple(ms)

    def close(self):
        if self.h:
            load().cfd_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
