"""CPU tests under AddressSanitizer + UndefinedBehaviorSanitizer: the oracle
(every entry point on small inputs), the C ABI's host-side paths that need
no GPU (the library built with its host half instrumented), and the C++
driver and VTK converter on the CPU paths of their own tests (ab.txt
parsing, -gen, flags, a missing input file, VTK conversion).  Each runs in
a child process with the sanitizer runtime preloaded; any report aborts it
(halt_on_error, -fno-sanitize-recover).  The instrumented builds come from
__graft_entry__.build() (oracle: make -C oracle san; host code:
scft_ver1.01_amd/build.py --san)."""
import glob
import os
import shutil
import subprocess
import sys

import pytest

from conftest import GOLDEN_DIR, ORACLE_DIR, PKG_DIR, ROOT

SAN_DIR = os.path.join(PKG_DIR, "_build", "san")
ORACLE_SAN = os.path.join(ORACLE_DIR, "libscft_oracle_san.so")
LIB_SAN = os.path.join(SAN_DIR, "libscft_hip_san.so")
DRIVER_SAN = os.path.join(SAN_DIR, "scft_san")
VTK_SAN = os.path.join(SAN_DIR, "scft_vtk_san")
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


def _gcc_asan():
    return subprocess.run(["gcc", "-print-file-name=libasan.so"], stdout=subprocess.PIPE, text=True).stdout.strip()


def _clang_asan():
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else ""


def _child(cmd, preload, cwd=None, extra=None):
    env = dict(os.environ, **SAN_ENV, LD_PRELOAD=preload, LD_LIBRARY_PATH=PKG_DIR)
    env.update(extra or {})
    return subprocess.run(cmd, cwd=cwd, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=600)


def _clean(r):
    return "ERROR: AddressSanitizer" not in r.stdout and "runtime error:" not in r.stdout


@pytest.mark.skipif(not os.path.exists(ORACLE_SAN), reason="sanitizer build of the oracle missing (make -C oracle san)")
def test_oracle_under_asan_ubsan():
    r = _child([sys.executable, os.path.join(ROOT, "tests", "san", "oracle_workout.py")], _gcc_asan(),
               extra={"SCFT_ORACLE_LIB": ORACLE_SAN})
    assert r.returncode == 0 and _clean(r) and "oracle workout done" in r.stdout, r.stdout[-3000:]


@pytest.mark.skipif(not (os.path.exists(LIB_SAN) and _clang_asan()),
                    reason="sanitizer build of the library missing (build.py --san)")
def test_abi_host_paths_under_asan_ubsan():
    r = _child([sys.executable, os.path.join(ROOT, "tests", "san", "abi_workout.py"), LIB_SAN], _clang_asan())
    assert r.returncode == 0 and _clean(r) and "abi workout done" in r.stdout, r.stdout[-3000:]


@pytest.mark.skipif(not (os.path.exists(DRIVER_SAN) and os.path.exists(VTK_SAN)),
                    reason="sanitizer build of the driver missing (build.py --san)")
def test_driver_cpu_paths_under_asan_ubsan(tmp_path):
    pre = _gcc_asan()
    shutil.copy(os.path.join(GOLDEN_DIR, "ab_reference.txt"), tmp_path / "ab.txt")
    r = _child([DRIVER_SAN, "-kernal=GPU", "-step=10"], pre, cwd=tmp_path)  # intag=1024 without phi_1.dat
    assert r.returncode == 1 and "phi_1.dat did not exist" in r.stdout and _clean(r), r.stdout[-3000:]
    with open(tmp_path / "ab.txt", "w") as f:
        f.write("intag=0\nxN=15.0\nfa=0.500\nlx=4.0, ly=3.0,  lz=2.0\nNx=8, Ny=6, Nz=4\nds0=0.01\nbatch=2\n"
                "field=gyroid\namp=1.0\nnoise=0.1\nseed=9\n")
    r = _child([DRIVER_SAN, "-gen"], pre, cwd=tmp_path)
    assert r.returncode == 0 and _clean(r), r.stdout[-3000:]
    r = _child([DRIVER_SAN, "-kernal=CPU"], pre, cwd=tmp_path)
    assert r.returncode == 2 and _clean(r), r.stdout[-3000:]
    r = _child([VTK_SAN, "phi_1.dat", "o.vtk"], pre, cwd=tmp_path)
    assert r.returncode == 0 and _clean(r), r.stdout[-3000:]
    assert open(tmp_path / "o.vtk").readline().strip() == "# vtk DataFile Version 3.0"
    r = _child([VTK_SAN, "missing.dat", "o2.vtk"], pre, cwd=tmp_path)
    assert r.returncode != 0 and _clean(r), r.stdout[-3000:]
