"""Runs every oracle entry point on small inputs; the sanitizer test runs it in
a child process with the AddressSanitizer/UBSan build of the oracle loaded
(SCFT_ORACLE_LIB) and the ASan runtime preloaded."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "oracle"))
import oracle as orc  # noqa: E402

rng = np.random.default_rng(0)
orc.set_threads(2)
for n in (8, 12, 30, 53, 64):
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(orc.fft1d(x), np.fft.fft(x))
    assert np.allclose(orc.dft1d_brute(x), np.fft.fft(x))
x = rng.standard_normal((4, 6, 10)) + 0j
assert np.allclose(orc.fft3d(x), np.fft.fftn(x))
for shape, fA, ds in (((8, 6, 4), 0.4, 0.02), ((6, 1, 1), 0.375, 0.01), ((10, 4, 1), 0.5, 0.05)):
    p = orc.params(*shape, 3.0, 2.5, 2.0, 14.0, fA, ds)
    kt = orc.ktable(p)
    N = int(np.prod(shape))
    wA = rng.uniform(-1, 1, N)
    wB = rng.uniform(-1, 1, N)
    pa, pb, Q, qa, qca, qb, qcb = orc.conc(p, kt, wA.copy(), wB.copy(), keep_q=True)
    assert Q > 0 and np.all(np.isfinite(pa))
    orc.run(p, wA.copy(), wB.copy(), 3)
    L, gA, gB, *_ = orc.conc_adjoint(p, kt, wA, wB, rng.standard_normal(N), rng.standard_normal(N), 0.5)
    assert np.isfinite(L)
    for m, ridge in ((0, 0.0), (3, 0.0), (4, 1e-2)):
        a = orc.Anderson(N, m, 0.1, 0.10, ridge=ridge)
        u, v = wA.copy(), wB.copy()
        for _ in range(6):
            a.iterate(p, kt, u, v, np.empty(N), np.empty(N))
        a.coef()
        a.history(0)
        a.close()
print("oracle workout done")
