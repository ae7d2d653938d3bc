"""GPU parity of every line length against the oracle, FP64 form.

The register tilings (2 .. 256 in the pass table) and the LDS Stockham
kernel that serves the other even 2-3-5-smooth lengths up to 2048 (e.g. 90,
300, 384, 480, 1000): one getConc, then two SCFT iterations, on grids with
the length along x, y or z, against the oracle to 1e-12 (Q, phi) and 1e-11
(F, phi after the iterations).  Then the reference's own sample
configuration (ab.txt: 480 x 480 x 1, lx = ly = 48.99, xN = 30, fa = 0.5,
ds0 = 0.01) for two iterations, in every form."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TABLE = [2, 4, 6, 8, 10, 12, 16, 20, 24, 32, 36, 40, 48, 60, 64, 72, 80, 96, 100, 128, 144, 160, 192, 256]
GENERIC = [18, 30, 50, 90, 120, 200, 240, 300, 384, 480, 512, 640, 1000, 2048]


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _case(pkg, orc, shape, L, prec=0, iters=2, fA=0.4, ds=0.02, chiN=15.0, **kw):
    p = pkg.default_params(nx=shape[0], ny=shape[1], nz=shape[2], lx=L[0], ly=L[1], lz=L[2], chiN=chiN, fA=fA,
                           ds=ds, precision=prec, **kw)
    wA, wB = pkg.fields.seeded_fields(*shape, *L, "lamellar", 1.0, 0.2, 97)
    with pkg.Solver(p) as s:
        s.set_fields(wA, wB)
        Q = s.concentrations()[0]
        phi0 = s.get_fields()[0][0].copy()
        _, fe = s.iterate(iters)
        phi = s.get_fields()[0][0].copy()
    op = orc.params(*shape, *L, chiN, fA, ds)
    kt = orc.ktable(op)
    a, b = wA.copy(), wB.copy()
    pa0, _, Qo = orc.conc(op, kt, a, b)
    pa, _, hist = orc.run(op, wA.copy(), wB.copy(), iters)
    return (abs(Q - Qo) / Qo, _rel(phi0, pa0), abs(fe[0][0] - hist[-1][0]) / abs(hist[-1][0]), _rel(phi, pa))


@pytest.mark.parametrize("n", [n for n in TABLE if n >= 36] + GENERIC)
def test_line_length_fp64_vs_oracle(pkg, orc, n):
    shapes = [(n, 6, 4)] + ([(6, n, 2), (4, 2, n)] if n not in TABLE or n >= 100 else [])
    for shape in shapes:
        if np.prod(shape) > 1 << 16:
            shape = tuple(min(s, 4) if s != n else s for s in shape)
        L = tuple(0.08 * s + 0.5 for s in shape)
        print("case", shape, flush=True)
        eQ, ephi0, eF, ephi = _case(pkg, orc, shape, L)
        print(f"  {shape}: relQ={eQ:.2e} phi={ephi0:.2e} F(2 it)={eF:.2e} phi(2 it)={ephi:.2e}")
        assert eQ < 1e-12 and ephi0 < 1e-12 and eF < 1e-11 and ephi < 1e-11, shape


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_reference_sample_config_480x480(pkg, orc, prec):
    """ab.txt of the reference (480 x 480 x 1): FP64 equals the oracle to 1e-11
    after two iterations; the FP32 forms are bounded against it (1e-5 FP32,
    1e-6 FP32X in F and phi)."""
    print("case 480x480x1 precision", prec, flush=True)
    shape, L = (480, 480, 1), (48.99, 48.99, 2.449)
    eQ, ephi0, eF, ephi = _case(pkg, orc, shape, L, prec=prec, fA=0.5, ds=0.01, chiN=30.0)
    print(f"  relQ={eQ:.2e} phi={ephi0:.2e} F(2 it)={eF:.2e} phi(2 it)={ephi:.2e}")
    tol = {0: 1e-11, 1: 1e-5, 2: 1e-6}[prec]
    assert eQ < tol and ephi0 < tol and eF < tol and ephi < tol
