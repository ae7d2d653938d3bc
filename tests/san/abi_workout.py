"""Calls the host-side paths of the C ABI that need no GPU (parameter
validation, error messages, contour counts, the no-allocation propagate's
size check, NULL handling) on the library built with host-side
AddressSanitizer/UBSan.  The sanitizer test runs it in a child process with
the ASan runtime preloaded; argv[1] is the library."""
import ctypes
import importlib.util
import os
import sys

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..")
spec = importlib.util.spec_from_file_location("scft_amd", os.path.join(ROOT, "scft_ver1.01_amd", "__init__.py"),
                                              submodule_search_locations=[os.path.join(ROOT, "scft_ver1.01_amd")])
pkg = importlib.util.module_from_spec(spec)
sys.modules["scft_amd"] = pkg
spec.loader.exec_module(pkg)
pkg.LIB_PATH = sys.argv[1]
L = pkg.lib()
P = pkg.Params


def params(**kw):
    p = pkg.default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


a, b = ctypes.c_int(), ctypes.c_int()
assert L.scft_contour_counts(0.375, 0.01, ctypes.byref(a), ctypes.byref(b)) == 0 and (a.value, b.value) == (37, 62)
assert L.scft_contour_counts(0.5, 0.0, ctypes.byref(a), ctypes.byref(b)) != 0
bad = [dict(nx=0), dict(lx=-1.0), dict(ds=0.0), dict(fA=1.0), dict(batch=0), dict(precision=7), dict(mixing=5),
       dict(mixing=1, anderson_m=99), dict(nx=7), dict(nx=2 * 2053), dict(nx=1, ny=1, nz=1), dict(xy_plane=9),
       dict(tiling=5), dict(anderson_reg=-1.0), dict(fA=0.001, ds=0.01)]
for kw in bad:
    h = ctypes.c_void_p()
    rc = L.scft_create(ctypes.byref(params(**kw)), 0, ctypes.byref(h))
    assert rc != 0 and not h.value, kw
    assert len(L.scft_last_error(None)) > 0
# a valid context: on a host without a GPU the device query fails cleanly
h = ctypes.c_void_p()
rc = L.scft_create(ctypes.byref(params(nx=16, ny=16, nz=16)), 0, ctypes.byref(h))
assert rc in (0, pkg.SCFT_E_HIP), rc
if rc == 0:
    L.scft_destroy(h)
assert L.scft_destroy(None) == 0
assert L.scft_iterate(None, 1, 0.0, None, None) != 0
dummy = ctypes.c_void_p(16)
assert L.scft_propagate_noalloc(ctypes.byref(params(nx=64, ny=64, nz=64, batch=2)), 0, 0, 1, dummy, None, dummy,
                                None) == pkg.SCFT_E_INVALID
assert L.scft_propagate_noalloc(None, 0, 0, 1, dummy, None, dummy, None) == pkg.SCFT_E_INVALID
assert L.scft_propagate_noalloc(ctypes.byref(params()), 0, 2, 1, dummy, None, dummy, None) == pkg.SCFT_E_INVALID
inst = (pkg.Instance * 2)(pkg.Instance(16, 4.0, 15.0, 0.4), pkg.Instance(15, 4.0, 15.0, 0.4))
sw = ctypes.c_void_p()
assert L.scft_sweep_create(ctypes.byref(params()), inst, 2, 0, ctypes.byref(sw)) != 0
print("abi workout done")
