"""Build the HIP library (libscft_hip.so) for gfx950 in-tree.

    python scft_ver1.01_amd/build.py            # library + C++ driver
    python scft_ver1.01_amd/build.py --lib      # library only

Every translation unit is compiled with hipcc, in parallel, then linked into
scft_ver1.01_amd/libscft_hip.so.  Packed-FP32 instructions are disabled for
all device code (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libscft_hip.so")
DRIVER_SRC = os.path.join(HERE, "driver")
DRIVER_BIN = os.path.join(HERE, "scft")
VTK_BIN = os.path.join(HERE, "scft_vtk")

ARCH = os.environ.get("SCFT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{INCLUDE}", f"-I{CSRC}",
]

LIB_SOURCES = ["pass_f64.hip", "pass_f32.hip", "pass_f32x.hip", "scft_lib.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass warns about the device-only target feature; drop that line
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_lib(jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in LIB_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if out.strip():
                print(out)
    tmp = LIB + ".tmp"
    # -z defs: an object missing from the link is an error, not a load-time failure
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-z,defs", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_driver():
    """C++ host driver (main/init/io/vtk roles) linked against libscft_hip.so."""
    if not os.path.isdir(DRIVER_SRC):
        return None
    srcs = sorted(os.path.join(DRIVER_SRC, f) for f in os.listdir(DRIVER_SRC)
                  if f.endswith(".cpp") and f != "vtk_main.cpp")
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", "-o", DRIVER_BIN, *srcs,
          f"-L{HERE}", "-lscft_hip", "-Wl,-rpath,$ORIGIN"])
    vtk = os.path.join(DRIVER_SRC, "vtk_main.cpp")
    if os.path.exists(vtk):
        _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", "-o", VTK_BIN, vtk,
              os.path.join(DRIVER_SRC, "io.cpp")])
    return DRIVER_BIN


SAN = os.path.join(BUILD, "san")
SAN_LIB = os.path.join(SAN, "libscft_hip_san.so")
SAN_DRIVER = os.path.join(SAN, "scft_san")
SAN_VTK = os.path.join(SAN, "scft_vtk_san")
HOST_SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g"]


def build_sanitized():
    """AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code,
    for the CPU sanitizer tests (not loaded by anything else): the driver and
    the VTK converter (g++), and the library with scft_lib.hip's host half
    instrumented (device code unchanged; the pass objects of build_lib are
    reused).  Needs build_lib first."""
    os.makedirs(SAN, exist_ok=True)
    drv = sorted(os.path.join(DRIVER_SRC, f) for f in os.listdir(DRIVER_SRC)
                 if f.endswith(".cpp") and f != "vtk_main.cpp")
    _run(["g++", "-O1", "-std=c++17", *HOST_SAN, f"-I{INCLUDE}", "-o", SAN_DRIVER, *drv,
          f"-L{HERE}", "-lscft_hip", f"-Wl,-rpath,{HERE}"])
    _run(["g++", "-O1", "-std=c++17", *HOST_SAN, f"-I{INCLUDE}", "-o", SAN_VTK,
          os.path.join(DRIVER_SRC, "vtk_main.cpp"), os.path.join(DRIVER_SRC, "io.cpp")])
    # host-only sanitizers on a hipcc line: each flag behind -Xarch_host
    host_san = []
    for f in HOST_SAN:
        host_san += ["-Xarch_host", f]
    obj = os.path.join(SAN, "scft_lib_san.o")
    _run([HIPCC, *DEVICE_FLAGS, *host_san, "-c", os.path.join(CSRC, "scft_lib.hip"), "-o", obj])
    objs = [os.path.join(BUILD, os.path.splitext(f)[0] + ".o") for f in LIB_SOURCES if f != "scft_lib.hip"]
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-shared-libasan", *host_san, "-o", SAN_LIB,
          *objs, obj])
    return SAN_LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("--san", action="store_true", help="also the sanitizer builds of the host code")
    ap.add_argument("-j", type=int, default=4)
    a = ap.parse_args(argv)
    print("built", build_lib(a.j))
    if not a.lib:
        d = build_driver()
        if d:
            print("built", d)
    if a.san:
        print("built", build_sanitized())


if __name__ == "__main__":
    sys.exit(main())
