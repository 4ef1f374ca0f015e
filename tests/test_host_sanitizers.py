"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

Builds tests/host/mesh_io_check.c with cfd_mesh_io.c (the mesh, porous-flag, nodal-field and
state-file routines: plain C, no GPU code) using gcc's -fsanitize=address,undefined, and
runs it: every routine on several grid shapes, file round trips including records wrapped
over lines, and an invalid mesh request that must fail cleanly. Any sanitizer report or
leak fails the test. Skipped where gcc cannot build sanitized code.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_mesh_io_under_sanitizers(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc absent")
    exe = str(tmp_path / "mesh_io_check")
    cmd = [gcc, "-std=c11", "-O1", "-g", "-fno-omit-frame-pointer", "-Wall", "-Werror",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "solver-cfd_amd", "csrc"),
           os.path.join(ROOT, "tests", "host", "mesh_io_check.c"),
           os.path.join(ROOT, "solver-cfd_amd", "csrc", "cfd_mesh_io.c"), "-lm", "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True)
    if b.returncode != 0 and "sanitize" in b.stderr and "unrecognized" in b.stderr:
        pytest.skip("gcc without sanitizer support")
    assert b.returncode == 0, b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0 and "mesh_io_check ok" in r.stdout, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr
