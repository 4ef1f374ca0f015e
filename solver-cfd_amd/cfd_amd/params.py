"""ctypes mirror of ``cfd_params`` (include/cfd.h) and its reference defaults.

The defaults restate OpenMp_Multicore/comum.f90:29-115 (Re, Pe, Sc, Darcy number, Cf,
g, Fr, v_i), main.f90:38 (Tinf), main.f90:105-106 (dtau, dt) and
OpenMp_Multicore/input/iterations.dat (beta, b_art).
"""
import ctypes
import math

SCHEME_QUICK, SCHEME_UPWIND = 0, 1
BC_REFERENCE, BC_CHANNEL = 0, 1
BUOY_REFERENCE, BUOY_BOUSSINESQ = 0, 1
COMPAT_FORTRAN, COMPAT_CUDA = 0, 1
MU_CONST, MU_POLY4 = 0, 1
K_CONST, K_POLY4 = 0, 1
PREC_FP64, PREC_FP32 = 0, 1

# field ids (include/cfd.h)
F_UM, F_VM, F_UM_TAU, F_VM_TAU, F_P, F_T, F_T_TAU, F_C, F_C_TAU, F_EPS, F_LP, F_PN = range(12)
FIELD_NAMES = ["um", "vm", "um_tau", "vm_tau", "p", "T", "T_tau", "C", "C_tau", "eps", "lp", "pn"]
LAYOUT_C, LAYOUT_FORTRAN = 0, 1

CF_REFERENCE = 1.75 / math.sqrt(150.0 * 0.5 ** 3)   # comum.f90:31 with porosidade = 0.5


class CfdParams(ctypes.Structure):
    _fields_ = [
        ("imax", ctypes.c_int32), ("jmax", ctypes.c_int32),
        ("scheme", ctypes.c_int32), ("bc", ctypes.c_int32),
        ("buoyancy", ctypes.c_int32), ("compat", ctypes.c_int32),
        ("steady", ctypes.c_int32), ("mu_model", ctypes.c_int32),
        ("re", ctypes.c_double), ("pe", ctypes.c_double), ("sc", ctypes.c_double),
        ("darcy", ctypes.c_double), ("cf", ctypes.c_double), ("g", ctypes.c_double),
        ("invfr2", ctypes.c_double), ("ri", ctypes.c_double), ("t_ref", ctypes.c_double),
        ("v_i", ctypes.c_double), ("tinf", ctypes.c_double),
        ("dtau", ctypes.c_double), ("dt", ctypes.c_double),
        ("beta", ctypes.c_double), ("b_art", ctypes.c_double),
        ("u_in", ctypes.c_double), ("t_bottom", ctypes.c_double), ("t_top", ctypes.c_double),
        ("slot_lo", ctypes.c_double), ("slot_hi", ctypes.c_double), ("slot_value", ctypes.c_double),
        ("mu_coef", ctypes.c_double * 5),
        ("k_coef", ctypes.c_double * 5),
        ("k_model", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


DEFAULTS = dict(
    scheme=SCHEME_QUICK, bc=BC_REFERENCE, buoyancy=BUOY_REFERENCE, compat=COMPAT_FORTRAN,
    steady=0, mu_model=MU_CONST,
    re=1.0, pe=1.0, sc=1.0, darcy=1.0e-2, cf=CF_REFERENCE, g=9.80665, invfr2=1.0,
    ri=0.0, t_ref=0.0, v_i=0.5, tinf=1.0, dtau=5.0e-2, dt=0.5e-2, beta=1300.0, b_art=0.0,
    u_in=1.0, t_bottom=1.0, t_top=0.0, slot_lo=0.0, slot_hi=0.0, slot_value=1.0,
    mu_coef=(1.0, 0.0, 0.0, 0.0, 0.0),
    k_coef=(1.0, 0.0, 0.0, 0.0, 0.0), k_model=K_CONST, reserved=0,
)


def make_params(imax, jmax, **kw):
    """Build a CfdParams from the reference defaults overridden by keyword arguments."""
    d = dict(DEFAULTS)
    unknown = set(kw) - set(d)
    if unknown:
        raise KeyError("unknown cfd_params fields: %s" % sorted(unknown))
    d.update(kw)
    p = CfdParams()
    p.imax, p.jmax = int(imax), int(jmax)
    for k, v in d.items():
        if k in ("mu_coef", "k_coef"):
            for n in range(5):
                getattr(p, k)[n] = float(v[n])
        elif isinstance(getattr(p, k), int):
            setattr(p, k, int(v))
        else:
            setattr(p, k, float(v))
    return p


def params_to_dict(p):
    out = {}
    for name, _ in CfdParams._fields_:
        v = getattr(p, name)
        out[name] = list(v) if name in ("mu_coef", "k_coef") else v
    return out
