"""Input/output of the reference harness (oracle/ref/harness.f90).

TEST INFRASTRUCTURE ONLY: used by tests/golden/make_golden.py to produce golden vectors from
the reference's own Fortran hot path, and by bench.py's cpu_baseline to time that hot path
on the host. The executables are built from the reference's sources by
oracle/ref/build_ref.sh into oracle/_ref/; no reference source is kept in this repository.

The harness reads harness_in.bin from its working directory: int32 mode, nsteps, nphys,
itc_max, steady; float64 dtau, dt, beta, b_art, tinf, eps_conv; then the geometry and the
state as Fortran-ordered float64 arrays (see harness.f90), and writes harness_out.bin.
"""
import os
import resource
import subprocess
import tempfile
import time

import numpy as np

from cfd_amd import cases as C


def geometry_payload(x, y, eps, lp):
    """x(1:imax), y(1:jmax), xm, ym, face areas and porosity fields in the harness order."""
    im, jm = len(x) - 1, len(y) - 1
    g = C.derive_geometry(x, y)
    return [x[1:], y[1:], g["xm"][1:], g["ym"][1:], g["aun"][2:im + 1], g["aue"][1:jm + 1],
            g["avn"][1:im + 1], g["ave"][2:jm + 1],
            C.to_fortran(eps, "p", im, jm), C.to_fortran(lp, "p", im, jm)]


def state_payload(st, im, jm):
    return [C.to_fortran(st["um"], "u", im, jm), C.to_fortran(st["vm"], "v", im, jm),
            C.to_fortran(st["p"], "p", im, jm), C.to_fortran(st["T"], "p", im, jm),
            C.to_fortran(st["C"], "p", im, jm)]


def write_input(path, header_i, header_d, payload):
    with open(path, "wb") as f:
        np.asarray(header_i, dtype=np.int32).tofile(f)
        np.asarray(header_d, dtype=np.float64).tofile(f)
        for a in payload:
            np.asarray(a, dtype=np.float64).ravel(order="F").tofile(f)


def _big_stack():
    """The reference's routines keep grid-sized automatic arrays on the stack: lift the
    soft stack limit to the hard one for the child (the OpenMP workers get OMP_STACKSIZE)."""
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def run(binary, header_i, header_d, payload, threads=1):
    """Run the harness once; returns the bytes of harness_out.bin."""
    with tempfile.TemporaryDirectory() as d:
        os.makedirs(os.path.join(d, "data"), exist_ok=True)
        write_input(os.path.join(d, "harness_in.bin"), header_i, header_d, payload)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_STACKSIZE="1G")
        subprocess.run([binary], cwd=d, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                       env=env, preexec_fn=_big_stack)
        return open(os.path.join(d, "harness_out.bin"), "rb").read()


def properties(binary, temps):
    """The reference's mixture properties (properties_CH4.f90:8-58) at each ambient
    temperature Too in temps: an array [n, 6] of k_tot, rho_tot, cp_tot, nu_tot, alpha_tot, Pr."""
    temps = np.asarray(temps, dtype=np.float64)
    buf = run(binary, [5, len(temps), 0, 0, 0], [0.05, 0.005, 1.0, 0.0, 1.0, 0.0], [temps])
    return np.frombuffer(buf, dtype=np.float64).reshape(len(temps), 6)


def time_pseudo_steps(binary, header_d, payload, n_short, n_long, threads):
    """Seconds per pseudo-step of the harness's mode-2 loop (the reference's solve_U, solve_V,
    solve_P, solve_Z, solve_C per step): the difference of two runs of n_long and n_short
    steps, which cancels process start, input reading and output writing."""
    with tempfile.TemporaryDirectory() as d:
        os.makedirs(os.path.join(d, "data"), exist_ok=True)
        inp = os.path.join(d, "harness_in.bin")
        write_input(inp, [2, n_short, 0, 0, 0], header_d, payload)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_STACKSIZE="1G")
        times = []
        for n in (n_short, n_long):
            with open(inp, "r+b") as f:              # nsteps is the second int32 of the header
                f.seek(4)
                np.asarray([n], dtype=np.int32).tofile(f)
            t0 = time.perf_counter()
            subprocess.run([binary], cwd=d, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                           env=env, preexec_fn=_big_stack)
            times.append(time.perf_counter() - t0)
        return (times[1] - times[0]) / (n_long - n_short), times


class Reader:
    """Sequential reader of harness_out.bin."""

    def __init__(self, buf):
        self.buf, self.off = buf, 0

    def f64(self, shape):
        n = int(np.prod(shape))
        a = np.frombuffer(self.buf, dtype=np.float64, count=n, offset=self.off)
        self.off += 8 * n
        return a.reshape(shape, order="F")

    def i32(self, n):
        a = np.frombuffer(self.buf, dtype=np.int32, count=n, offset=self.off)
        self.off += 4 * n
        return a

    def field(self, kind, im, jm):
        return C.from_fortran(self.f64(C.fortran_shape(kind, im, jm)), kind, im, jm)
