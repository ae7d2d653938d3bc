"""Host-side Python mirror of the SCFT solver interface (ctypes over libscft_hip.so).

The reference (ren5885a/SCFT_Ver1.01) drives its CUDA solver from C++ through
these functions; the methods of `Solver` keep their roles, argument meaning
and failure behaviour (an exception carrying the library's status instead of
`checkCudaErrors`' exit):

    Solver(params)            init_cuda + initialize_cufft      init_cuda.cu:227-546
    Solver.close()            finalize_cufft                    init_cuda.cu:551-628
    Solver.set_fields         phi_<n>.dat upload                init_cuda.cu:435-456, 508-513
    Solver.propagate          sovDifFft                         cuda_scft.cu:435-563
    Solver.concentrations     getConc                           cuda_scft.cu:61-138
    Solver.update_fields      phi_w                             cuda_aid.cu:382-416
    Solver.free_energy        F block of Free                   cuda_scft.cu:191-223
    Solver.iterate            Free's do-while loop              cuda_scft.cu:167-253
    Solver.time_iterations    fft_test                          cuda_scft.cu:382-432

All compute runs in the HIP library; there is no CPU fallback.  If the
library is missing, importing this package raises.
"""
import ctypes
import os

import numpy as np

from . import fields  # noqa: F401  (seeded field generators)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libscft_hip.so")

SCFT_OK = 0
SCFT_E_INVALID = 1
SCFT_E_UNSUPPORTED = 2
SCFT_E_HIP = 3
SCFT_E_NOMEM = 4
SCFT_E_NONFINITE = 5
SCFT_E_IO = 6
SCFT_E_STATE = 7

FP64 = 0
FP32 = 1
FP32X = 2  # FP32 storage and HBM bytes, FP64 transform arithmetic
MIX_SIMPLE = 0
MIX_ANDERSON = 1
REFERENCE_PI = 3.1415926535  # struct.h:11

_STATUS_NAMES = {
    1: "SCFT_E_INVALID", 2: "SCFT_E_UNSUPPORTED", 3: "SCFT_E_HIP", 4: "SCFT_E_NOMEM",
    5: "SCFT_E_NONFINITE", 6: "SCFT_E_IO", 7: "SCFT_E_STATE",
}

# every entry point declared in include/scft.h
ABI_SYMBOLS = (
    "scft_default_params", "scft_contour_counts", "scft_create", "scft_destroy",
    "scft_last_error", "scft_set_fields", "scft_get_fields", "scft_propagate",
    "scft_concentrations", "scft_update_fields", "scft_free_energy", "scft_iterate",
    "scft_time_iterations", "scft_time_pass", "scft_autotune", "scft_set_tiling", "scft_sizes",
    "scft_conc_adjoint", "scft_adjoint_loop",
    "scft_sweep_create", "scft_sweep_destroy", "scft_sweep_last_error", "scft_sweep_set_fields",
    "scft_sweep_get_fields", "scft_sweep_iterate", "scft_sweep_time_iterations", "scft_sweep_autotune",
    "scft_sweep_size", "scft_propagate_noalloc", "scft_set_precision", "scft_copy_bandwidth",
)


class ScftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    """scft_params of include/scft.h"""
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("chiN", ctypes.c_double), ("fA", ctypes.c_double), ("ds", ctypes.c_double),
        ("batch", ctypes.c_int), ("precision", ctypes.c_int), ("mixing", ctypes.c_int),
        ("wopt", ctypes.c_double), ("wcmp", ctypes.c_double),
        ("anderson_m", ctypes.c_int), ("anderson_mix", ctypes.c_double),
        ("pi", ctypes.c_double), ("use_graph", ctypes.c_int), ("xy_plane", ctypes.c_int),
        ("tiling", ctypes.c_int), ("checkpoint", ctypes.c_int), ("fuse_steps", ctypes.c_int),
        ("anderson_reg", ctypes.c_double),
    ]

    def replace(self, **kw):
        q = Params()
        ctypes.pointer(q)[0] = self
        for k, v in kw.items():
            setattr(q, k, v)
        return q

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Instance(ctypes.Structure):
    """scft_instance of include/scft.h"""
    _fields_ = [("n", ctypes.c_int), ("L", ctypes.c_double), ("chiN", ctypes.c_double), ("fA", ctypes.c_double)]


_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def lib():
    """Load libscft_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python scft_ver1.01_amd/build.py` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(Params)
        vp = ctypes.c_void_p
        sig = {
            "scft_default_params": (None, [P]),
            "scft_contour_counts": (ctypes.c_int, [ctypes.c_double, ctypes.c_double,
                                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
            "scft_create": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(vp)]),
            "scft_destroy": (ctypes.c_int, [vp]),
            "scft_last_error": (ctypes.c_char_p, [vp]),
            "scft_set_fields": (ctypes.c_int, [vp, _dp, _dp]),
            "scft_get_fields": (ctypes.c_int, [vp, _dp, _dp, _dp, _dp]),
            "scft_propagate": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
            "scft_concentrations": (ctypes.c_int, [vp, _dp]),
            "scft_update_fields": (ctypes.c_int, [vp]),
            "scft_free_energy": (ctypes.c_int, [vp, _dp]),
            "scft_iterate": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double,
                                            ctypes.POINTER(ctypes.c_int), _dp]),
            "scft_time_iterations": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
            "scft_time_pass": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
            "scft_conc_adjoint": (ctypes.c_int, [vp, _dp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp, _dp, _dp,
                                                 _dp, _dp]),
            "scft_adjoint_loop": (ctypes.c_int, [vp, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp]),
            "scft_sweep_create": (ctypes.c_int, [P, ctypes.POINTER(Instance), ctypes.c_int, ctypes.c_int,
                                                 ctypes.POINTER(vp)]),
            "scft_sweep_destroy": (ctypes.c_int, [vp]),
            "scft_sweep_last_error": (ctypes.c_char_p, [vp]),
            "scft_sweep_set_fields": (ctypes.c_int, [vp, ctypes.c_int, _dp, _dp]),
            "scft_sweep_get_fields": (ctypes.c_int, [vp, ctypes.c_int, _dp, _dp, _dp, _dp]),
            "scft_sweep_iterate": (ctypes.c_int, [vp, ctypes.c_int, _dp]),
            "scft_sweep_time_iterations": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
            "scft_sweep_autotune": (ctypes.c_int, [vp, ctypes.c_int]),
            "scft_sweep_size": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_long),
                                               ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
            "scft_autotune": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_float)]),
            "scft_set_tiling": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int)]),
            "scft_set_precision": (ctypes.c_int, [vp, ctypes.c_int]),
            "scft_copy_bandwidth": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                                   ctypes.POINTER(ctypes.c_double)]),
            "scft_sizes": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_long)]),
            # device pointers (w, q_init, q_out) and a hipStream_t
            "scft_propagate_noalloc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def default_params(**kw):
    p = Params()
    lib().scft_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def copy_bandwidth(nbytes=1 << 30, reps=10, device=0):
    """Attainable HBM bandwidth (GB/s, read + write) by the library's copy kernel."""
    g = ctypes.c_double()
    rc = lib().scft_copy_bandwidth(int(device), int(nbytes), int(reps), ctypes.byref(g))
    if rc:
        raise ScftError(rc, lib().scft_last_error(None).decode())
    return g.value


def contour_counts(fA, ds):
    """(NsA, NsB) truncated as init_cuda.cu:330-331."""
    a = ctypes.c_int()
    b = ctypes.c_int()
    rc = lib().scft_contour_counts(fA, ds, ctypes.byref(a), ctypes.byref(b))
    if rc:
        raise ScftError(rc, "bad fA/ds")
    return a.value, b.value


def _arr(a, n):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class Solver:
    """One solver context (batch of instances sharing params) on one HIP device."""

    def __init__(self, params, device=0):
        self.params = params
        self._h = ctypes.c_void_p()
        rc = lib().scft_create(ctypes.byref(params), device, ctypes.byref(self._h))
        if rc:
            msg = lib().scft_last_error(None).decode()
            raise ScftError(rc, msg)
        N = ctypes.c_long()
        a = ctypes.c_int()
        b = ctypes.c_int()
        nb = ctypes.c_long()
        lib().scft_sizes(self._h, ctypes.byref(N), ctypes.byref(a), ctypes.byref(b), ctypes.byref(nb))
        self.N, self.NsA, self.NsB, self.device_bytes = N.value, a.value, b.value, nb.value
        self.batch = params.batch

    # -- lifecycle --
    def close(self):
        if self._h:
            lib().scft_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise ScftError(rc, lib().scft_last_error(self._h).decode())

    # -- fields --
    def set_fields(self, wA, wB):
        n = self.batch * self.N
        wA = _arr(wA, n)
        wB = _arr(wB, n)
        self._check(lib().scft_set_fields(self._h, _ptr(wA), _ptr(wB)))

    def get_fields(self):
        n = self.batch * self.N
        out = [np.empty(n) for _ in range(4)]
        self._check(lib().scft_get_fields(self._h, *[_ptr(o) for o in out]))
        shp = (self.batch, self.N)
        return tuple(o.reshape(shp) for o in out)  # phiA, phiB, wA, wB

    # -- solver steps --
    def propagate(self, block, sign, q_init=None):
        """sovDifFft: returns q[batch][ns+1][N] (FP64 copy of the device slices)."""
        ns = self.NsA if block == 0 else self.NsB
        qi = None if q_init is None else _arr(q_init, self.batch * self.N)
        out = np.empty(self.batch * (ns + 1) * self.N)
        self._check(lib().scft_propagate(self._h, block, sign, _ptr(qi), _ptr(out)))
        return out.reshape(self.batch, ns + 1, self.N)

    def concentrations(self):
        """getConc: returns Q[batch]; phi readable with get_fields()."""
        Q = np.empty(self.batch)
        self._check(lib().scft_concentrations(self._h, _ptr(Q)))
        return Q

    def update_fields(self):
        self._check(lib().scft_update_fields(self._h))

    def free_energy(self):
        out = np.empty(5 * self.batch)
        self._check(lib().scft_free_energy(self._h, _ptr(out)))
        return out.reshape(self.batch, 5)

    def iterate(self, n, tol=0.0):
        """Free's loop: returns (iterations run, fe[batch][5] of the last one)."""
        done = ctypes.c_int()
        fe = np.empty(5 * self.batch)
        self._check(lib().scft_iterate(self._h, int(n), float(tol), ctypes.byref(done), _ptr(fe)))
        return done.value, fe.reshape(self.batch, 5)

    # -- adjoint --
    def conc_adjoint(self, uA, uB, lamA=None, lamB=None, mu=0.0):
        """L = sum(lamA phiA + lamB phiB) + mu ln Q at fields u (used as given).
        Returns (L[batch], gA, gB, phiA, phiB, Q[batch]), g = dL/du, arrays [batch][N]."""
        n = self.batch * self.N
        uA = _arr(uA, n)
        uB = _arr(uB, n)
        lA = None if lamA is None else _arr(lamA, n)
        lB = None if lamB is None else _arr(lamB, n)
        L = np.empty(self.batch)
        Q = np.empty(self.batch)
        out = [np.empty(n) for _ in range(4)]  # phiA, phiB, gA, gB
        self._check(lib().scft_conc_adjoint(self._h, _ptr(uA), _ptr(uB), _ptr(lA), _ptr(lB), float(mu), _ptr(L),
                                            _ptr(out[0]), _ptr(out[1]), _ptr(Q), _ptr(out[2]), _ptr(out[3])))
        shp = (self.batch, self.N)
        return L, out[2].reshape(shp), out[3].reshape(shp), out[0].reshape(shp), out[1].reshape(shp), Q

    def adjoint_loop(self, h, w0A, w0B, n_iter, target, grad=True):
        """loss = sum (phiA_last - target)^2 after n_iter chain evaluations of the
        SCFT loop with external field h on A; returns (loss[batch], grad_h[batch][N] or None, phiA_last)."""
        n = self.batch * self.N
        h = _arr(h, n)
        a = _arr(w0A, n)
        b = _arr(w0B, n)
        t = _arr(target, n)
        loss = np.empty(self.batch)
        g = np.empty(n) if grad else None
        ph = np.empty(n)
        self._check(lib().scft_adjoint_loop(self._h, _ptr(h), _ptr(a), _ptr(b), int(n_iter), _ptr(t), _ptr(loss),
                                            _ptr(g), _ptr(ph)))
        shp = (self.batch, self.N)
        return loss, (g.reshape(shp) if grad else None), ph.reshape(shp)

    # -- timing --
    def time_iterations(self, n):
        ms = ctypes.c_float()
        self._check(lib().scft_time_iterations(self._h, int(n), ctypes.byref(ms)))
        return ms.value

    def autotune(self, reps=20):
        """Pick the fastest pass tiling per axis (and plane fusion); returns
        (choice[4], ms[4]) = variants of x, y, z, plane flag; their times."""
        ch = (ctypes.c_int * 4)()
        ms = (ctypes.c_float * 4)()
        self._check(lib().scft_autotune(self._h, int(reps), ch, ms))
        return list(ch), list(ms)

    def set_tiling(self, choice):
        """Apply a tiling choice [x, y, z variant, plane 0/1] (as autotune returns)."""
        ch = (ctypes.c_int * 4)(*[int(c) for c in choice])
        self._check(lib().scft_set_tiling(self._h, ch))

    def set_precision(self, precision):
        """Switch an FP32-storage solver between FP32 and FP32X (fields, history kept)."""
        self._check(lib().scft_set_precision(self._h, int(precision)))

    def time_pass(self, axis, reps):
        ms = ctypes.c_float()
        self._check(lib().scft_time_pass(self._h, int(axis), int(reps), ctypes.byref(ms)))
        return ms.value


class Sweep:
    """Independent instances with their own grid, box, chiN and fA (scft_sweep_*)."""

    def __init__(self, base, instances, device=0):
        arr = (Instance * len(instances))(*[Instance(int(n), float(L), float(c), float(f))
                                             for (n, L, c, f) in instances])
        self._h = ctypes.c_void_p()
        rc = lib().scft_sweep_create(ctypes.byref(base), arr, len(instances), device, ctypes.byref(self._h))
        if rc:
            raise ScftError(rc, lib().scft_sweep_last_error(None).decode())
        self.count = len(instances)
        self.sizes = []
        for i in range(self.count):
            N = ctypes.c_long()
            a = ctypes.c_int()
            b = ctypes.c_int()
            lib().scft_sweep_size(self._h, i, ctypes.byref(N), ctypes.byref(a), ctypes.byref(b))
            self.sizes.append((N.value, a.value, b.value))

    def close(self):
        if self._h:
            lib().scft_sweep_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise ScftError(rc, lib().scft_sweep_last_error(self._h).decode())

    def set_fields(self, i, wA, wB):
        N = self.sizes[i][0]
        self._check(lib().scft_sweep_set_fields(self._h, i, _ptr(_arr(wA, N)), _ptr(_arr(wB, N))))

    def get_fields(self, i):
        N = self.sizes[i][0]
        out = [np.empty(N) for _ in range(4)]
        self._check(lib().scft_sweep_get_fields(self._h, i, *[_ptr(o) for o in out]))
        return tuple(out)

    def iterate(self, n):
        fe = np.empty(5 * self.count)
        self._check(lib().scft_sweep_iterate(self._h, int(n), _ptr(fe)))
        return fe.reshape(self.count, 5)

    def time_iterations(self, n):
        ms = ctypes.c_float()
        self._check(lib().scft_sweep_time_iterations(self._h, int(n), ctypes.byref(ms)))
        return ms.value

    def autotune(self, reps=10):
        self._check(lib().scft_sweep_autotune(self._h, int(reps)))
