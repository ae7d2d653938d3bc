"""GPU test of scft_propagate_noalloc (the variant of scft_propagate that
allocates nothing: its work buffers live in one static device array of the
library).  Called through the C ABI on device arrays the test allocates with
the HIP runtime; its output must equal scft_propagate's on the same inputs,
on consecutive calls whose second and third are smaller than the first (the
shared scratch keeps the previous call's contents)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Dev:
    """device buffers through the HIP runtime the library itself uses"""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(a.nbytes)) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, shape):
        out = np.empty(shape)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def test_propagate_noalloc_matches_propagate(pkg):
    L = pkg.lib()
    dev = _Dev()
    cases = [((32, 32, 32), 2, 0, 0, 1, True), ((16, 16, 16), 1, 0, 1, -1, True), ((24, 16, 8), 1, 1, 0, -1, False)]
    try:
        for k, (shape, batch, prec, block, sign, with_q0) in enumerate(cases):
            print(f"call {k}: shape {shape} batch {batch} precision {prec} block {block} sign {sign}", flush=True)
            p = pkg.default_params(nx=shape[0], ny=shape[1], nz=shape[2], lx=3.3, ly=4.1, lz=2.7, chiN=15.0,
                                   fA=0.4, ds=0.01, batch=batch, precision=prec)
            N = int(np.prod(shape))
            wA = np.concatenate([pkg.fields.seeded_fields(*shape, 3.3, 4.1, 2.7, "lamellar", 1.0, 0.3, 11 + b)[0]
                                 for b in range(batch)])
            wB = np.concatenate([pkg.fields.seeded_fields(*shape, 3.3, 4.1, 2.7, "lamellar", 1.0, 0.3, 11 + b)[1]
                                 for b in range(batch)])
            q0 = 0.5 + pkg.fields.splitmix_uniform(12, batch * N) if with_q0 else None
            with pkg.Solver(p) as s:
                s.set_fields(wA, wB)
                ns = s.NsA if block == 0 else s.NsB
                ref = s.propagate(block, sign, q0)
            w = wA if block == 0 else wB
            dw = dev.put(w)
            dq = dev.put(q0) if with_q0 else None
            dout = dev.put(np.full(batch * (ns + 1) * N, np.nan))
            rc = L.scft_propagate_noalloc(ctypes.byref(p), 0, block, sign, dw, dq, dout, None)
            assert rc == 0, L.scft_last_error(None)
            got = dev.get(dout, (batch, ns + 1, N))
            err = float(np.max(np.abs(got - ref.reshape(batch, ns + 1, N))) / np.max(np.abs(ref)))
            print(f"  rel diff vs scft_propagate {err:.2e}", flush=True)
            assert np.all(np.isfinite(got))
            assert err == 0.0 if prec == 0 else err < 1e-6
            dev.free()
    finally:
        dev.free()
