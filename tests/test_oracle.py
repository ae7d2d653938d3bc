"""CPU tests that pin the FP64 oracle (no GPU).

Each check is chosen so that a plausible mistake in the oracle (a dropped term,
a wrong sign or index, a transposed axis) fails it: library FFT vectors
(numpy.fft) and a long-double brute-force DFT for the transform; closed forms
for the propagator step; invariants that only hold when march, junctions,
trapezoid weights and Q are all right (<q(s), qdag(s)> = N Q for every s,
mean(phi_A) = NsA ds, dlnQ/dw = -phi/N by finite differences, A<->B and
translation symmetry); closed-form KATs for the loop, the mixing update and F.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR

REF_PI = 3.1415926535


@pytest.fixture(scope="module")
def fftv():
    return np.load(os.path.join(GOLDEN_DIR, "fft_vectors.npz"))


# ---------------------------------------------------------------- FFT ------
def test_fft1d_matches_numpy_vectors(orc, fftv):
    for key in fftv.files:
        if not key.startswith("x1_"):
            continue
        n = int(key[3:])
        x = fftv[key]
        F = orc.fft1d(x, -1)
        I = orc.fft1d(x, +1)
        scale = max(1.0, np.max(np.abs(fftv[f"F1_{n}"])))
        assert np.max(np.abs(F - fftv[f"F1_{n}"])) / scale < 2e-15 * max(1, np.log2(max(n, 2))), n
        assert np.max(np.abs(I - fftv[f"I1_{n}"])) / scale < 2e-15 * max(1, np.log2(max(n, 2))), n


def test_fft3d_matches_numpy_vectors(orc, fftv):
    for key in fftv.files:
        if not key.startswith("x3_"):
            continue
        shp = key[3:]
        x = fftv[key]
        F = orc.fft3d(x.astype(complex), -1)
        ref = fftv[f"F3_{shp}"]
        assert np.max(np.abs(F - ref)) / np.max(np.abs(ref)) < 1e-14, shp
        # the half spectrum cuFFT's D2Z would return is a slice of the full one
        half = fftv[f"R3_{shp}"]
        assert np.max(np.abs(F[..., : half.shape[-1]] - half)) / np.max(np.abs(half)) < 1e-14


@pytest.mark.parametrize("n", list(range(1, 21)) + [24, 27, 32, 48, 49])
def test_fft1d_brute_force(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        F = orc.fft1d(x, sign)
        B = orc.dft1d_brute(x, sign)
        assert np.max(np.abs(F - B)) / max(1.0, np.max(np.abs(B))) < 5e-15 * max(1, np.log2(n + 1))


def test_fft_closed_forms(orc):
    n = 48
    d = np.zeros(n, complex)
    d[0] = 1.0
    assert np.allclose(orc.fft1d(d), np.ones(n), atol=0, rtol=0)  # delta -> constant, exactly
    j = np.arange(n)
    c = np.cos(2 * np.pi * 5 * j / n)
    F = orc.fft1d(c.astype(complex))
    expect = np.zeros(n)
    expect[5] = expect[n - 5] = n / 2
    assert np.max(np.abs(F - expect)) < 1e-13


# ---------------------------------------------------------- constants ------
def test_contour_counts_truncation(orc):
    assert orc.contour_counts(0.5, 0.01) == (50, 50)
    assert orc.contour_counts(0.3, 0.005) == (60, 140)
    assert orc.contour_counts(0.35, 0.005) == (70, 130)
    # f not a multiple of ds: both counts truncate and one contour step is lost
    a, b = orc.contour_counts(0.375, 0.01)
    assert (a, b) == (37, 62) and a + b == 99


def test_ksq_axis(orc):
    L = 3.7
    for n in (1, 2, 7, 8, 48):
        k = orc.ksq_axis(n, L, REF_PI)
        m = np.arange(n)
        m = np.where(m <= n // 2 - 1, m, m - n)
        assert np.allclose(k, (2 * REF_PI * m / L) ** 2, rtol=1e-14, atol=0)
        if n % 2 == 0:
            for i in range(1, n):
                assert k[i] == pytest.approx(k[n - i], rel=1e-14)  # even in k
        elif n > 1:
            # odd n: index n//2 gets frequency n//2 - n, so the table is not even in k
            assert k[n // 2] == pytest.approx((2 * REF_PI * (n // 2 - n) / L) ** 2, rel=1e-14)
            assert k[n // 2] != pytest.approx(k[n - n // 2], rel=1e-3)
    # the truncated pi differs from M_PI by 9e-11 -- visible at 1e-10 parity
    k = orc.ksq_axis(8, 1.0, REF_PI)
    kt = orc.ksq_axis(8, 1.0, np.pi)
    assert abs(k[1] / kt[1] - 1) > 5e-11
    # the reference's n = 1 table is (2 pi / L)^2, not 0 (init_cuda.cu:410-411)
    assert orc.ksq_axis(1, 2.0, REF_PI)[0] == pytest.approx((2 * REF_PI / 2.0) ** 2, rel=1e-15)


# -------------------------------------------------------------- step -------
def _grid(nx, ny, nz, lx, ly, lz):
    z, y, x = np.meshgrid(np.arange(nz) * lz / nz, np.arange(ny) * ly / ny, np.arange(nx) * lx / nx,
                          indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


@pytest.mark.parametrize("mode", [(1, 0, 0), (0, 2, 0), (0, 0, 3), (2, 1, 1)])
def test_step_closed_form_w0(orc, mode):
    """w = 0: the split step is exact, cos(k.r) decays by exp(-k^2 ds)."""
    nx, ny, nz, lx, ly, lz = 16, 12, 10, 3.3, 4.1, 2.7
    p = orc.params(nx, ny, nz, lx, ly, lz, 15.0, 0.5, 0.01)
    x, y, z = _grid(nx, ny, nz, lx, ly, lz)
    kx, ky, kz = (2 * REF_PI * mode[0] / lx, 2 * REF_PI * mode[1] / ly, 2 * REF_PI * mode[2] / lz)
    q0 = np.cos(2 * np.pi * (mode[0] * x / lx + mode[1] * y / ly + mode[2] * z / lz))
    N = nx * ny * nz
    q1 = orc.step(p, orc.ktable(p), np.ones(N), q0)
    expect = np.exp(-(kx * kx + ky * ky + kz * kz) * 0.01) * q0
    assert np.max(np.abs(q1 - expect)) < 1e-14


def test_step_uniform_field(orc):
    p = orc.params(8, 8, 8, 4.0, 4.0, 4.0, 15.0, 0.5, 0.01)
    c = 2.5
    wdz = np.full(512, np.exp(-c * 0.005))
    q1 = orc.step(p, orc.ktable(p), wdz, np.ones(512))
    assert np.max(np.abs(q1 - np.exp(-c * 0.01))) < 1e-15


def test_step_dense_operator(orc):
    """The step as dense DFT matrices (textbook restatement, independent of the
    oracle's FFT and index code): S = W Re(F^-1 diag(K) F) W on a 4x6x5 grid."""
    nx, ny, nz = 4, 6, 5
    lx, ly, lz = 2.0, 3.0, 2.5
    p = orc.params(nx, ny, nz, lx, ly, lz, 15.0, 0.5, 0.02)
    rng = np.random.default_rng(3)
    N = nx * ny * nz
    wdz = np.exp(-rng.uniform(-1, 1, N) * 0.01)
    q0 = rng.uniform(0.5, 1.5, N)

    def dft(n):
        j = np.arange(n)
        return np.exp(-2j * np.pi * np.outer(j, j) / n)

    def k2(n, L):
        m = np.arange(n)
        m = np.where(m <= n // 2 - 1, m, m - n)
        return (2 * REF_PI * m / L) ** 2

    F = np.kron(dft(nz), np.kron(dft(ny), dft(nx)))  # x fastest
    K = np.exp(-0.02 * (k2(nz, lz)[:, None, None] + k2(ny, ly)[None, :, None] + k2(nx, lx)[None, None, :])).ravel()
    Finv = np.conj(F)
    S = (Finv @ np.diag(K) @ F).real / N
    expect = wdz * (S @ (wdz * q0))
    got = orc.step(p, orc.ktable(p), wdz, q0)
    assert np.max(np.abs(got - expect)) < 1e-14


# ------------------------------------------------------------ chain --------
def _fields(n, amp, seed, N):
    rng = np.random.default_rng(seed)
    wA = amp * rng.uniform(-1, 1, N)
    wB = amp * rng.uniform(-1, 1, N)
    return wA - wA.mean(), wB - wB.mean()


@pytest.fixture(scope="module")
def chain(orc):
    p = orc.params(8, 10, 6, 3.0, 3.5, 2.5, 14.0, 0.4, 0.02)
    N = 480
    wA, wB = _fields(0, 1.5, 7, N)
    kt = orc.ktable(p)
    out = orc.conc(p, kt, wA.copy(), wB.copy(), keep_q=True)
    return p, kt, wA, wB, out


def test_qqdag_invariant(orc, chain):
    """<q(s), qdag(s)>/N = Q at every contour point (WKW is symmetric)."""
    p, kt, wA, wB, (phiA, phiB, Q, qa, qca, qb, qcb) = chain
    NsA, NsB = orc.contour_counts(p.fA, p.ds)
    for s in range(NsA + 1):
        assert np.mean(qa[s] * qca[s]) == pytest.approx(Q, rel=1e-13)
    for s in range(NsB + 1):
        assert np.mean(qb[s] * qcb[s]) == pytest.approx(Q, rel=1e-13)
    # junction hand-offs (qa_to_qInt, qa_to_qInt2)
    assert np.array_equal(qb[0], qa[NsA])
    assert np.array_equal(qca[NsA], qcb[0])
    assert np.all(qcb[NsB] == 1.0) and np.all(qa[0] == 1.0)


def test_phi_means(orc, chain):
    p, kt, wA, wB, (phiA, phiB, Q, *_) = chain
    NsA, NsB = orc.contour_counts(p.fA, p.ds)
    assert phiA.mean() == pytest.approx(NsA * p.ds, rel=1e-13)
    assert phiB.mean() == pytest.approx(NsB * p.ds, rel=1e-13)


def _logQ(orc, p, kt, wA, wB):
    NsA, NsB = orc.contour_counts(p.fA, p.ds)
    N = wA.size
    qa = orc.march(p, kt, wA, np.ones(N), NsA, 1)
    qb = orc.march(p, kt, wB, qa[NsA].copy(), NsB, 1)
    return np.log(qb[NsB].mean())


def test_dlnQ_dw_equals_minus_phi(orc, chain):
    """dlnQ/dw_A(r) = -phi_A(r)/N exactly for the discrete scheme (central FD)."""
    p, kt, wA, wB, (phiA, phiB, Q, *_) = chain
    N = wA.size
    eps = 1e-5
    for r in (0, 17, 123, 479):
        for w, phi, which in ((wA, phiA, "A"), (wB, phiB, "B")):
            wp = w.copy(); wp[r] += eps
            wm = w.copy(); wm[r] -= eps
            if which == "A":
                d = (_logQ(orc, p, kt, wp, wB) - _logQ(orc, p, kt, wm, wB)) / (2 * eps)
            else:
                d = (_logQ(orc, p, kt, wA, wp) - _logQ(orc, p, kt, wA, wm)) / (2 * eps)
            assert d == pytest.approx(-phi[r] / N, rel=2e-7), (r, which)


def test_ab_swap_symmetry(orc):
    """f = 1/2: reversing the chain swaps blocks, so (wA, wB) -> (wB, wA) swaps phiA, phiB."""
    p = orc.params(8, 8, 8, 3.0, 3.0, 3.0, 12.0, 0.5, 0.02)
    wA, wB = _fields(0, 1.0, 9, 512)
    kt = orc.ktable(p)
    pa, pb, Q = orc.conc(p, kt, wA.copy(), wB.copy())
    pa2, pb2, Q2 = orc.conc(p, kt, wB.copy(), wA.copy())
    assert Q2 == pytest.approx(Q, rel=1e-13)
    assert np.max(np.abs(pa2 - pb)) < 1e-13 and np.max(np.abs(pb2 - pa)) < 1e-13


def test_translation_covariance(orc):
    p = orc.params(8, 6, 4, 3.0, 2.0, 1.5, 12.0, 0.3, 0.05)
    wA, wB = _fields(0, 1.0, 10, 192)
    kt = orc.ktable(p)
    pa, pb, Q = orc.conc(p, kt, wA.copy(), wB.copy())
    shift = lambda f: np.roll(f.reshape(4, 6, 8), (1, 2, 3), axis=(0, 1, 2)).ravel()
    pa2, pb2, Q2 = orc.conc(p, kt, shift(wA), shift(wB))
    assert Q2 == pytest.approx(Q, rel=1e-13)
    assert np.max(np.abs(pa2 - shift(pa))) < 1e-13


# ------------------------------------------------------ loop KATs ----------
@pytest.mark.parametrize("chiN,f,ds", [(15.0, 0.5, 0.01), (20.0, 0.3, 0.005), (20.0, 0.35, 0.005)])
def test_disordered_free_energy(orc, chiN, f, ds):
    """w = 0: Q = 1, phi uniform; F is evaluated with the updated w (P4), which
    gives F = chiN fa fb - (wA' fa + wB' fb), = chiN f(1-f) + 0.1 chiN (1/2-f)^2
    when the counts are exact."""
    p = orc.params(8, 8, 8, 4.0, 4.0, 4.0, chiN, f, ds)
    NsA, NsB = orc.contour_counts(f, ds)
    fa, fb = NsA * ds, NsB * ds
    phiA, phiB, hist = orc.run(p, np.zeros(512), np.zeros(512), 3)
    assert np.max(np.abs(phiA - fa)) < 1e-13
    if abs(fa + fb - 1) < 1e-12:
        assert hist[-1][0] == pytest.approx(chiN * f * (1 - f) + 0.1 * chiN * (0.5 - f) ** 2, rel=1e-13)
    assert hist[-1][3] == pytest.approx(0.0, abs=1e-14)  # F_S = -ln Q = 0


def test_phi_w_fixed_point(orc):
    """At phiA + phiB = 1 and wA - wB = chiN (phiB - phiA) the update is the identity."""
    p = orc.params(4, 4, 4, 1.0, 1.0, 1.0, 17.0, 0.5, 0.01)
    rng = np.random.default_rng(5)
    phiA = rng.uniform(0.1, 0.9, 64)
    phiB = 1 - phiA
    wA = rng.uniform(-2, 2, 64)
    wB = wA - 17.0 * (phiB - phiA)
    a, b = wA.copy(), wB.copy()
    orc.phi_w(p, a, b, phiA, phiB)
    assert np.max(np.abs(a - wA)) < 1e-13 and np.max(np.abs(b - wB)) < 1e-13
    # an incompressibility defect psi = 1 - phiA - phiB moves both fields by -wopt*wcmp*psi
    pB = phiB - 0.01
    a, b = wA.copy(), wB.copy()
    orc.phi_w(p, a, b, phiA, pB)
    # psi = 0.01; eta unchanged; waDiff = chiN*(-0.01) - wcmp*0.01, wbDiff = -wcmp*0.01
    assert np.allclose(a - wA, 0.05 * (-17.0 * 0.01 - 0.1 * 0.01), rtol=1e-10, atol=1e-15)
    assert np.allclose(b - wB, 0.05 * (-0.1 * 0.01), rtol=1e-10, atol=1e-15)


def test_free_energy_uniform(orc):
    p = orc.params(4, 4, 4, 1.0, 1.0, 1.0, 10.0, 0.5, 0.01)
    N = 64
    out = orc.free_energy(p, np.full(N, 0.7), np.full(N, -0.2), np.full(N, 0.4), np.full(N, 0.55), 0.8)
    assert out[1] == pytest.approx(10.0 * 0.4 * 0.55)
    assert out[2] == pytest.approx(-(0.7 * 0.4 - 0.2 * 0.55))
    assert out[3] == pytest.approx(-np.log(0.8))
    assert out[0] == pytest.approx(out[1] + out[2] + out[3])
    assert out[4] == pytest.approx(0.05)


def test_c1_golden_consistency(orc):
    """The committed 300-iteration fixture starts where the oracle does (first 10 iterations)."""
    g = np.load(os.path.join(GOLDEN_DIR, "c1_oracle.npz"))
    import importlib.util
    from conftest import PKG_DIR
    spec = importlib.util.spec_from_file_location("scft_fields_t", os.path.join(PKG_DIR, "fields.py"))
    fields = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(fields)
    wA, wB = fields.workload_fields("c1_lamellar_32")
    assert np.array_equal(np.concatenate([wA[:16], wB[:16]]), g["input_head"])
    p = orc.params(16, 16, 16, 4.0, 4.0, 4.0, 15.0, 0.5, 0.01)  # cheap sanity of the loop at 16^3
    _, _, hist = orc.run(p, np.zeros(4096), np.zeros(4096), 1)
    assert hist[0][0] == pytest.approx(3.75, rel=1e-14)
    fe = g["fe_every10"]
    assert fe.shape == (30, 5) and np.all(np.isfinite(fe))
    assert np.all(np.abs(fe[:, 0] - fe[:, 1] - fe[:, 2] - fe[:, 3]) < 1e-12)


# ------------------------------------------------------------ Anderson -----
def test_anderson_m0_is_simple_mixing(orc):
    p = orc.params(8, 8, 8, 3.0, 3.0, 3.0, 14.0, 0.4, 0.02, wopt=0.07)
    kt = orc.ktable(p)
    wA, wB = _fields(0, 1.0, 12, 512)
    a1, b1 = wA.copy(), wB.copy()
    a2, b2 = wA.copy(), wB.copy()
    and_ = orc.Anderson(512, 0, 0.07, 0.10)
    pa = np.empty(512); pb = np.empty(512)
    for _ in range(5):
        orc.iterate(p, kt, a1, b1, pa, pb)
        and_.iterate(p, kt, a2, b2, pa, pb)
    assert np.max(np.abs(a1 - a2)) < 1e-13 and np.max(np.abs(b1 - b2)) < 1e-13


def test_anderson_reaches_simple_mixing_fixed_point(orc):
    """Anderson converges to the fixed point simple mixing finds: the disordered
    melt below the ODT (chiN = 8), and the lamellar phase at chiN = 14 when
    started from a partly converged field."""
    x = np.arange(8) * 0.5
    seed = np.cos(2 * np.pi * x / 4.0)
    for chiN, pre, m, mix in ((8.0, 0, 4, 0.2), (14.0, 150, 3, 0.1)):
        p = orc.params(8, 1, 1, 4.0, 1.0, 1.0, chiN, 0.5, 0.02)
        kt = orc.ktable(p)
        a, b = seed.copy(), -seed
        _, _, hs = orc.run(p, a, b, 4000)
        a0, b0 = seed.copy(), -seed
        if pre:
            orc.run(p, a0, b0, pre)
        and_ = orc.Anderson(8, m, mix, 0.10)
        pa = np.empty(8); pb = np.empty(8)
        for _ in range(150):
            Q, fe = and_.iterate(p, kt, a0, b0, pa, pb)
        assert fe[4] < 1e-9
        assert fe[0] == pytest.approx(hs[-1][0], abs=1e-9), chiN


def _anderson_run(orc, m, ridge, iters=8):
    """Anderson on an 8^3 melt; after each update yields (coefficients, history
    entries 0..m' of w and d, the updated fields)."""
    p = orc.params(8, 8, 8, 3.0, 3.0, 3.0, 14.0, 0.4, 0.02)
    kt = orc.ktable(p)
    wA, wB = _fields(0, 1.0, 21, 512)
    and_ = orc.Anderson(512, m, 0.1, 0.10, ridge=ridge)
    for _ in range(iters):
        orc.minus_average(wA)
        orc.minus_average(wB)
        pa, pb, _ = orc.conc(p, kt, wA, wB)
        and_.update(p, wA, wB, pa, pb)
        c = and_.coef()
        hist = [and_.history(k) for k in range(len(c) + 1)]
        yield c, hist, np.concatenate([wA, wB])


@pytest.mark.parametrize("ridge", [0.0, 1e-2])
def test_anderson_coefficients_are_least_squares(orc, ridge):
    """The coefficients solve min ||d0 - D c||^2 + lam ||c||^2 with D's columns
    d0 - d_i over the stored history, lam = ridge * tr(D^T D) / m' (numpy's SVD
    least squares on the stacked system, not the normal equations the oracle
    solves), and the update is the affine combination sum_j alpha_j (w_j + mix
    d_j) with alpha_0 = 1 - sum c, alpha_i = c_i."""
    mps = []
    for c, hist, wnew in _anderson_run(orc, 4, ridge):
        mp = len(c)
        mps.append(mp)
        w0, d0 = hist[0]
        if mp == 0:
            assert np.allclose(wnew, w0 + 0.1 * d0, rtol=0, atol=1e-14)
            continue
        D = np.column_stack([d0 - hist[i][1] for i in range(1, mp + 1)])
        lam = ridge * np.trace(D.T @ D) / mp
        Dx = np.vstack([D, np.sqrt(lam) * np.eye(mp)]) if ridge else D
        rhs = np.concatenate([d0, np.zeros(mp)]) if ridge else d0
        c_ls = np.linalg.lstsq(Dx, rhs, rcond=None)[0]
        assert np.max(np.abs(c - c_ls)) < 1e-8 * max(1.0, np.max(np.abs(c_ls))), (mp, c, c_ls)
        alpha = np.concatenate([[1.0 - c.sum()], c])
        w_aff = sum(alpha[j] * (hist[j][0] + 0.1 * hist[j][1]) for j in range(mp + 1))
        assert np.max(np.abs(wnew - w_aff)) < 1e-12 * max(1.0, np.max(np.abs(wnew)))
    assert mps == [0, 1, 2, 3, 4, 4, 4, 4]  # history grows to m, then the oldest entry drops


def test_anderson_ridge_changes_coefficients(orc):
    """the ridge is applied (a ridged solve is not the plain one)"""
    c0 = [c for c, _, _ in _anderson_run(orc, 4, 0.0, 5)][-1]
    c1 = [c for c, _, _ in _anderson_run(orc, 4, 1e-2, 5)][-1]
    assert np.max(np.abs(c0 - c1)) > 1e-6


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("npts", [2, 3])
def test_anderson_linear_map_terminates(orc, seed, npts):
    """Anderson without truncation is GMRES in disguise on an affine map (Walker
    & Ni, SIAM J. Numer. Anal. 49 (2011) 1715, Thm 2.2): with phi an affine
    function of w the residual d(w) is affine in the n = 2*npts unknowns, and
    the iterate after n + 1 updates is the fixed point up to rounding, while
    the first n updates are not.  Simple mixing (m = 0) is far from it."""
    n = 2 * npts
    rng = np.random.default_rng(seed)
    M = 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.uniform(0.2, 0.5, n)
    chiN, wcmp = 2.0, 0.10
    p = orc.params(npts, 1, 1, 1.0, 1.0, 1.0, chiN, 0.5, 0.01)

    def resid(w):
        f = M @ w + b
        wA, wB, pA, pB = w[:npts], w[npts:], f[:npts], f[npts:]
        eta = (wA + wB - chiN) / 2
        ps = 1 - pA - pB
        return np.concatenate([chiN * pB + eta - wA - wcmp * ps, chiN * pA + eta - wB - wcmp * ps])

    out = {}
    for m in (0, n + 2):
        and_ = orc.Anderson(npts, m, 0.5, wcmp)
        w = np.linspace(-1, 1, n)
        r0 = np.linalg.norm(resid(w))
        rel = []
        for _ in range(n + 2):
            f = M @ w + b
            wA, wB = w[:npts].copy(), w[npts:].copy()
            and_.update(p, wA, wB, f[:npts].copy(), f[npts:].copy())
            w = np.concatenate([wA, wB])
            rel.append(np.linalg.norm(resid(w)) / r0)
        out[m] = rel
    acc = out[n + 2]
    assert min(acc[:n]) > 1e-3, acc
    assert acc[n] < 1e-8 and acc[n + 1] < 1e-13, acc
    assert min(out[0]) > 1e-2, out[0]
