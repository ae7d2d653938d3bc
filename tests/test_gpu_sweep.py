"""GPU tests of the heterogeneous sweep: every instance of a mixed-grid,
mixed-parameter sweep equals the same instance solved alone, and the oracle
(including instances whose truncated contour counts give Ns = 99)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("mixing,prec", [(0, 0), (1, 0), (0, 1), (0, 2)])
def test_sweep_matches_single_instances(pkg, mixing, prec):
    inst = pkg.fields.sweep_instances(count=9, seed=11, sizes=(16, 24, 32))
    fields = pkg.fields.sweep_fields(inst, noise=0.2, seed=500)
    base = pkg.default_params(ds=0.01, precision=prec, mixing=mixing, anderson_m=3, anderson_mix=0.2)
    with pkg.Sweep(base, inst) as sw:
        sw.autotune(3)
        for i, (a, b) in enumerate(fields):
            sw.set_fields(i, a, b)
        fe = sw.iterate(6)
        phis = [sw.get_fields(i)[0] for i in range(len(inst))]
        sizes = sw.sizes
    # a random f is never a multiple of ds: both counts truncate and a step is lost (Ns = 99)
    assert all(s[1] + s[2] == 99 for s in sizes) and len({s[1] for s in sizes}) > 1
    for i, (n, L, chi, f) in enumerate(inst):
        p = pkg.default_params(nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, chiN=chi, fA=f, ds=0.01, precision=prec,
                               mixing=mixing, anderson_m=3, anderson_mix=0.2)
        with pkg.Solver(p) as s:
            s.set_fields(*fields[i])
            _, fe1 = s.iterate(6)
            phi1 = s.get_fields()[0][0]
        # the sweep's autotuned kernels (plane forms, tilings) sum in another order than the single
        # instance's defaults: FP64 rounding differences, which 6 iterations of Anderson's least squares
        # (m = 3 from noise) amplify about 1e5-fold; simple mixing keeps them at rounding level
        # (FP32: the kernel forms' own rounding, ~1e-7 per evaluation, over 6 iterations)
        # (FP32X: one FP32 rounding per stored value, whose order the kernel forms change)
        tol = (1e-12 if mixing == 0 else 1e-9) if prec == 0 else (5e-6 if prec == 1 else 2e-6)
        assert np.max(np.abs(fe[i] - fe1[0])) < tol * max(1, np.max(np.abs(fe1[0]))), (i, fe[i], fe1[0])
        assert _rel(phis[i], phi1) < tol, i


def test_sweep_vs_oracle(pkg, orc):
    # f = 0.375 gives NsA + NsB = 99 (one contour step lost to truncation), 0.3 gives 100
    inst = [(16, 4.2, 17.0, 0.375), (24, 3.7, 22.0, 0.3), (16, 4.9, 13.0, 0.31)]
    fields = pkg.fields.sweep_fields(inst, noise=0.3, seed=600)
    base = pkg.default_params(ds=0.01)
    with pkg.Sweep(base, inst) as sw:
        for i, (a, b) in enumerate(fields):
            sw.set_fields(i, a, b)
        fe = sw.iterate(3)
        phis = [sw.get_fields(i)[0] for i in range(3)]
        assert sw.sizes[0][1:] == (37, 62)
    for i, (n, L, chi, f) in enumerate(inst):
        op = orc.params(n, n, n, L, L, L, chi, f, 0.01)
        pa, pb, hist = orc.run(op, fields[i][0].copy(), fields[i][1].copy(), 3)
        assert abs(fe[i][0] - hist[-1][0]) / abs(hist[-1][0]) < 1e-10, i
        assert _rel(phis[i], pa) < 1e-10, i
