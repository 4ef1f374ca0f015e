"""Build libcfd_amd.so in-tree: hipcc for the gfx950 kernels, gcc for the C host code.

    python -m cfd_amd.build          (from solver-cfd_amd/)  or  build() from Python

Objects go to solver-cfd_amd/_build/, the library to solver-cfd_amd/cfd_amd/libcfd_amd.so.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                      # solver-cfd_amd/
CSRC = os.path.join(ROOT, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libcfd_amd.so")
REPO = os.path.dirname(ROOT)
FORTRAN_MOD = os.path.join(ROOT, "fortran", "cfd_amd.f90")
FORTRAN_DRIVER = os.path.join(ROOT, "fortran", "cfd_fortran_driver.f90")
# the Fortran main program that runs main.f90's loop through the library (built next to it)
DRIVER = os.path.join(PKG, "cfd_fortran_driver")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("CFD_OFFLOAD_ARCH", "gfx950")

# packed FP32 instructions (v_pk_*_f32) are kept out of every kernel: some of their forms
# have given silently wrong results on MI355X. The host pass warns that it does not know
# the feature; that is harmless.
NO_PACKED = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HIP_COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-Wno-unused-command-line-argument",
              "-I" + os.path.join(REPO, "include")] + NO_PACKED

# (source, extra flags): the per-operator kernels keep one rounding per operation so that
# they reproduce the CPU oracle bit for bit; the fused kernels may contract to FMA.
HIP_SOURCES = [
    ("cfd_ops.hip", ["-ffp-contract=off"]),
    ("cfd_fused.hip", ["-ffp-contract=on"]),
    ("cfd_unet.hip", ["-ffp-contract=fast"]),
    ("cfd_noalloc.hip", ["-ffp-contract=off"]),
]
C_SOURCES = ["cfd_host.c", "cfd_batch.c", "cfd_mesh_io.c", "cfd_fortran.c"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: %s" % " ".join(cmd))
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, extra in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, "-c", os.path.join(CSRC, src), "-o", obj] + HIP_COMMON + extra)
        objs.append(obj)
    for src in C_SOURCES:
        obj = os.path.join(BUILD, src.replace(".c", ".o"))
        _run(["gcc", "-std=c11", "-O2", "-fPIC", "-pthread", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
              "-I" + os.path.join(ROCM, "include"), "-I" + os.path.join(REPO, "include"),
              "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs +
         ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-lm", "-lpthread", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, LIB)
    # the Fortran binding module and the Fortran driver (main.f90's loop calling the
    # library's Fortran entry points) are built when a Fortran compiler is present
    fc = os.path.join(ROCM, "lib", "llvm", "bin", "flang")
    if os.path.exists(FORTRAN_MOD) and os.path.exists(fc):
        modo = os.path.join(BUILD, "cfd_amd_f.o")
        _run([fc, "-c", FORTRAN_MOD, "-module-dir", BUILD, "-o", modo])
        tmpd = DRIVER + ".tmp"
        _run([fc, "-O2", "-module-dir", BUILD, "-o", tmpd, FORTRAN_DRIVER, modo, "-L" + PKG, "-lcfd_amd",
              "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
        os.replace(tmpd, DRIVER)
    if verbose:
        print("built", LIB)
    return LIB


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)
    for f in (LIB, DRIVER):
        if os.path.exists(f):
            os.remove(f)


if __name__ == "__main__":
    build(verbose=True)
