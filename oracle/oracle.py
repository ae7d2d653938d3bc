"""ctypes binding of the CPU FP64 oracle (oracle/libscft_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in
__graft_entry__.py and by bench.py's cpu_baseline leg, and only to check results
(and, in bench.py, as the timed CPU baseline).
The product path (scft_ver1.01_amd) never imports this module.

Arrays are numpy float64, C-contiguous, grid index r = (k*Ny + j)*Nx + i.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SCFT_ORACLE_LIB selects another build of the same source (the sanitizer tests
# load libscft_oracle_san.so this way)
_LIB_PATH = os.environ.get("SCFT_ORACLE_LIB") or os.path.join(_HERE, "libscft_oracle.so")

REF_PI = 3.1415926535  # struct.h:11


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("chiN", ctypes.c_double), ("fA", ctypes.c_double), ("ds", ctypes.c_double),
        ("pi", ctypes.c_double), ("wopt", ctypes.c_double), ("wcmp", ctypes.c_double),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER(OrcParams)
        sig = {
            "orc_set_threads": (None, [ctypes.c_int]),
            "orc_get_threads": (ctypes.c_int, []),
            "orc_fft1d": (ctypes.c_int, [ctypes.c_int, _dp, _dp, ctypes.c_int]),
            "orc_fft3d": (ctypes.c_int, [ctypes.c_int] * 3 + [_dp, _dp, ctypes.c_int]),
            "orc_dft1d_brute": (None, [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]),
            "orc_contour_counts": (None, [ctypes.c_double, ctypes.c_double,
                                          ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
            "orc_ksq_axis": (None, [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]),
            "orc_ktable": (None, [P, _dp]),
            "orc_step": (None, [P, _dp, _dp, _dp, _dp]),
            "orc_march": (None, [P, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp]),
            "orc_minus_average": (None, [ctypes.c_long, _dp]),
            "orc_conc": (ctypes.c_double, [P, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_phi_w": (None, [P, ctypes.c_long, _dp, _dp, _dp, _dp]),
            "orc_free_energy": (None, [P, ctypes.c_long, _dp, _dp, _dp, _dp, ctypes.c_double, _dp]),
            "orc_iterate": (ctypes.c_double, [P, _dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_run": (None, [P, _dp, _dp, _dp, _dp, ctypes.c_int, _dp]),
            "orc_conc_adjoint": (ctypes.c_double, [P, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp,
                                                   _dp, _dp, _dp]),
            "orc_anderson_create": (ctypes.c_void_p, [ctypes.c_long, ctypes.c_int,
                                                      ctypes.c_double, ctypes.c_double]),
            "orc_anderson_destroy": (None, [ctypes.c_void_p]),
            "orc_anderson_set_ridge": (None, [ctypes.c_void_p, ctypes.c_double]),
            "orc_anderson_coef": (ctypes.c_int, [ctypes.c_void_p, _dp]),
            "orc_anderson_history": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp]),
            "orc_anderson_update": (ctypes.c_double, [ctypes.c_void_p, P, _dp, _dp, _dp, _dp]),
            "orc_iterate_anderson": (ctypes.c_double, [ctypes.c_void_p, P, _dp, _dp, _dp, _dp,
                                                       _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "need C-contiguous float64"
    return a.ctypes.data_as(_dp)


def params(nx, ny, nz, lx, ly, lz, chiN, fA, ds, pi=REF_PI, wopt=0.05, wcmp=0.10):
    return OrcParams(nx, ny, nz, lx, ly, lz, chiN, fA, ds, pi, wopt, wcmp)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def fft1d(x, sign=-1):
    re = np.ascontiguousarray(np.real(x), dtype=np.float64).copy()
    im = np.ascontiguousarray(np.imag(x), dtype=np.float64).copy()
    rc = lib().orc_fft1d(len(re), _p(re), _p(im), sign)
    if rc != 0:
        raise ValueError("unsupported length")
    return re + 1j * im


def fft3d(x, sign=-1):
    """x: complex array shaped [nz][ny][nx]."""
    nz, ny, nx = x.shape
    re = np.ascontiguousarray(np.real(x), dtype=np.float64).copy()
    im = np.ascontiguousarray(np.imag(x), dtype=np.float64).copy()
    rc = lib().orc_fft3d(nx, ny, nz, _p(re), _p(im), sign)
    if rc != 0:
        raise ValueError("unsupported shape")
    return re + 1j * im


def dft1d_brute(x, sign=-1):
    n = len(x)
    re = np.ascontiguousarray(np.real(x), dtype=np.float64)
    im = np.ascontiguousarray(np.imag(x), dtype=np.float64)
    ore = np.empty(n)
    oim = np.empty(n)
    lib().orc_dft1d_brute(n, _p(re), _p(im), _p(ore), _p(oim), sign)
    return ore + 1j * oim


def contour_counts(fA, ds):
    a = ctypes.c_int()
    b = ctypes.c_int()
    lib().orc_contour_counts(fA, ds, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def ksq_axis(n, length, pi=REF_PI):
    out = np.empty(n)
    lib().orc_ksq_axis(n, length, pi, _p(out))
    return out


def ktable(p):
    out = np.empty(p.nx * p.ny * p.nz)
    lib().orc_ktable(ctypes.byref(p), _p(out))
    return out


def step(p, kt, wdz, qin):
    out = np.empty_like(qin)
    lib().orc_step(ctypes.byref(p), _p(kt), _p(wdz), _p(qin), _p(out))
    return out


def march(p, kt, w, qinit, ns, sign):
    N = p.nx * p.ny * p.nz
    q = np.empty((ns + 1) * N)
    lib().orc_march(ctypes.byref(p), _p(kt), _p(w), _p(qinit), ns, sign, _p(q))
    return q.reshape(ns + 1, N)


def minus_average(w):
    lib().orc_minus_average(w.size, _p(w))


def conc(p, kt, wA, wB, keep_q=False):
    """getConc; wA is mean-subtracted in place.  Returns (phiA, phiB, Q[, qa, qca, qb, qcb])."""
    N = p.nx * p.ny * p.nz
    NsA, NsB = contour_counts(p.fA, p.ds)
    phiA = np.empty(N)
    phiB = np.empty(N)
    if keep_q:
        qa = np.empty((NsA + 1, N)); qca = np.empty((NsA + 1, N))
        qb = np.empty((NsB + 1, N)); qcb = np.empty((NsB + 1, N))
        Q = lib().orc_conc(ctypes.byref(p), _p(kt), _p(wA), _p(wB), _p(phiA), _p(phiB),
                           _p(qa), _p(qca), _p(qb), _p(qcb))
        return phiA, phiB, Q, qa, qca, qb, qcb
    Q = lib().orc_conc(ctypes.byref(p), _p(kt), _p(wA), _p(wB), _p(phiA), _p(phiB),
                       None, None, None, None)
    return phiA, phiB, Q


def conc_adjoint(p, kt, uA, uB, lamA, lamB, mu=0.0):
    """L = sum(lamA phiA + lamB phiB) + mu ln Q for fields used as given;
    returns (L, gA, gB, phiA, phiB, Q) with g = dL/du."""
    N = p.nx * p.ny * p.nz
    phiA = np.empty(N); phiB = np.empty(N); gA = np.empty(N); gB = np.empty(N)
    Q = ctypes.c_double()
    L = lib().orc_conc_adjoint(ctypes.byref(p), _p(kt), _p(np.ascontiguousarray(uA)), _p(np.ascontiguousarray(uB)),
                               _p(np.ascontiguousarray(lamA)), _p(np.ascontiguousarray(lamB)), mu,
                               _p(phiA), _p(phiB), ctypes.byref(Q), _p(gA), _p(gB))
    return L, gA, gB, phiA, phiB, Q.value


def phi_w(p, wA, wB, phiA, phiB):
    lib().orc_phi_w(ctypes.byref(p), wA.size, _p(wA), _p(wB), _p(phiA), _p(phiB))


def free_energy(p, wA, wB, phiA, phiB, Q):
    out = np.empty(5)
    lib().orc_free_energy(ctypes.byref(p), wA.size, _p(wA), _p(wB), _p(phiA), _p(phiB), Q, _p(out))
    return out


def iterate(p, kt, wA, wB, phiA, phiB):
    fe = np.empty(5)
    Q = lib().orc_iterate(ctypes.byref(p), _p(kt), _p(wA), _p(wB), _p(phiA), _p(phiB), _p(fe))
    return Q, fe


def run(p, wA, wB, iters):
    """Run `iters` reference iterations in place on wA, wB.  Returns (phiA, phiB, fe_hist)."""
    N = p.nx * p.ny * p.nz
    phiA = np.empty(N)
    phiB = np.empty(N)
    hist = np.empty((iters, 5))
    lib().orc_run(ctypes.byref(p), _p(wA), _p(wB), _p(phiA), _p(phiB), iters, _p(hist))
    return phiA, phiB, hist


class Anderson:
    def __init__(self, N, m, mix, wcmp=0.10, ridge=0.0):
        self.N = N
        self.h = lib().orc_anderson_create(N, m, mix, wcmp)
        if not self.h:
            raise ValueError("Anderson history must be 0..64")
        if ridge:
            lib().orc_anderson_set_ridge(self.h, float(ridge))

    def coef(self):
        """least-squares coefficients c_1..c_m' of the last update"""
        c = np.zeros(64)
        mp = lib().orc_anderson_coef(self.h, _p(c))
        return c[:mp].copy()

    def history(self, k):
        """entry k of the history (0 = the last update's input): (w, d), 2N each"""
        w = np.empty(2 * self.N)
        d = np.empty(2 * self.N)
        if lib().orc_anderson_history(self.h, k, _p(w), _p(d)) != 0:
            raise IndexError(k)
        return w, d

    def close(self):
        if self.h:
            lib().orc_anderson_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def update(self, p, wA, wB, phiA, phiB):
        return lib().orc_anderson_update(self.h, ctypes.byref(p), _p(wA), _p(wB), _p(phiA), _p(phiB))

    def iterate(self, p, kt, wA, wB, phiA, phiB):
        fe = np.empty(5)
        Q = lib().orc_iterate_anderson(self.h, ctypes.byref(p), _p(kt), _p(wA), _p(wB),
                                       _p(phiA), _p(phiB), _p(fe))
        return Q, fe
