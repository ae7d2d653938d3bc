"""Property tests (hypothesis, derandomized so every run draws the same cases)
on tiny grids.

CPU: the oracle's step against numpy's FFT on any grid shape (odd and
length-1 axes included; the wavenumbers follow the reference's convention,
m <= n/2 - 1 positive, the rest negative, the truncated pi), and the chain
invariants <q(s), qdag(s)>/N = Q at every contour point and mean(phi_A) =
NsA ds, mean(phi_B) = NsB ds on random fields.

GPU: the HIP library in FP64 against the oracle on random even grids, box
lengths, chiN, f and ds (so odd and even Ns), with checkpointing and
step-boundary fusion each on or off: Q, phi and F after two iterations; the
chain adjoint against the oracle's adjoint on random cases (random lam, mu,
checkpointing on or off); the loop gradient against central differences of
its own forward on random small loops (simple or Anderson mixing)."""
import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

REF_PI = 3.1415926535

CPU = settings(max_examples=40, derandomize=True, deadline=None,
               suppress_health_check=[HealthCheck.function_scoped_fixture])
GPU = settings(max_examples=30, derandomize=True, deadline=None,
               suppress_health_check=[HealthCheck.function_scoped_fixture])


def _k2(n, L):
    m = np.arange(n)
    m = np.where(m <= n // 2 - 1, m, m - n)
    return (2 * REF_PI * m / L) ** 2


@CPU
@given(nx=st.integers(1, 9), ny=st.integers(1, 9), nz=st.integers(1, 9),
       lx=st.floats(1.5, 6.0), ly=st.floats(1.5, 6.0), lz=st.floats(1.5, 6.0),
       ds=st.sampled_from([0.005, 0.01, 0.02, 0.05]), seed=st.integers(0, 2**31 - 1))
def test_oracle_step_matches_numpy_fft(orc, nx, ny, nz, lx, ly, lz, ds, seed):
    """S q = W Re(ifftn(K fftn(W q))) with K = exp(-ds (kx^2 + ky^2 + kz^2))."""
    p = orc.params(nx, ny, nz, lx, ly, lz, 15.0, 0.5, ds)
    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    wdz = np.exp(-rng.uniform(-2, 2, N) * ds / 2)
    q0 = rng.uniform(0.1, 2.0, N)
    K = np.exp(-ds * (_k2(nz, lz)[:, None, None] + _k2(ny, ly)[None, :, None] + _k2(nx, lx)[None, None, :]))
    x = (wdz * q0).reshape(nz, ny, nx)
    expect = wdz * np.fft.ifftn(K * np.fft.fftn(x)).real.ravel()
    got = orc.step(p, orc.ktable(p), wdz, q0)
    assert np.max(np.abs(got - expect)) <= 1e-14 * max(1.0, np.max(np.abs(expect)))


@CPU
@given(nx=st.integers(1, 7), ny=st.integers(1, 7), nz=st.integers(1, 7), fA=st.floats(0.15, 0.85),
       ds=st.sampled_from([0.01, 0.02, 0.05]), amp=st.floats(0.0, 3.0), seed=st.integers(0, 2**31 - 1))
def test_oracle_chain_invariants(orc, nx, ny, nz, fA, ds, amp, seed):
    p = orc.params(nx, ny, nz, 3.0, 3.5, 2.5, 14.0, fA, ds)
    NsA, NsB = orc.contour_counts(fA, ds)
    N = nx * ny * nz
    rng = np.random.default_rng(seed)
    wA = amp * rng.uniform(-1, 1, N)
    wB = amp * rng.uniform(-1, 1, N)
    wA -= wA.mean()
    phiA, phiB, Q, qa, qca, qb, qcb = orc.conc(p, orc.ktable(p), wA.copy(), wB.copy(), keep_q=True)
    assert Q > 0 and np.all(phiA >= 0) and np.all(phiB >= 0)
    for s in range(NsA + 1):
        assert np.mean(qa[s] * qca[s]) == pytest.approx(Q, rel=1e-12)
    for s in range(NsB + 1):
        assert np.mean(qb[s] * qcb[s]) == pytest.approx(Q, rel=1e-12)
    assert phiA.mean() == pytest.approx(NsA * ds, rel=1e-12)
    assert phiB.mean() == pytest.approx(NsB * ds, rel=1e-12)


EVEN = st.sampled_from([2, 4, 6, 8, 10, 12, 16])


@pytest.mark.gpu
@GPU
@given(nx=EVEN, ny=EVEN, nz=st.sampled_from([1, 2, 4, 6, 8, 12]), lx=st.floats(2.0, 5.0), ly=st.floats(2.0, 5.0),
       lz=st.floats(2.0, 5.0), chiN=st.floats(8.0, 25.0), fA=st.floats(0.2, 0.8),
       ds=st.sampled_from([0.01, 0.015, 0.02, 0.03, 0.05]), ckpt=st.booleans(), fuse=st.booleans(),
       seed=st.integers(0, 2**20))
def test_gpu_matches_oracle_random_cases(pkg, orc, nx, ny, nz, lx, ly, lz, chiN, fA, ds, ckpt, fuse, seed):
    """FP64 library vs oracle: one getConc (Q rel 1e-12, phi 1e-12), then F and
    phi after two simple-mixing iterations (1e-11)."""
    print("case", (nx, ny, nz), (lx, ly, lz), chiN, fA, ds, "ckpt", ckpt, "fuse", fuse, "seed", seed, flush=True)
    p = pkg.default_params(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz, chiN=chiN, fA=fA, ds=ds, precision=0,
                           checkpoint=int(ckpt), fuse_steps=int(fuse))
    wA, wB = pkg.fields.seeded_fields(nx, ny, nz, lx, ly, lz, "lamellar", 1.0, 0.2, seed)
    op = orc.params(nx, ny, nz, lx, ly, lz, chiN, fA, ds, pi=p.pi, wopt=p.wopt, wcmp=p.wcmp)
    with pkg.Solver(p) as s:
        s.set_fields(wA, wB)
        Q = s.concentrations()[0]
        phiA, phiB, _, _ = s.get_fields()
        pa, pb, Qo = orc.conc(op, orc.ktable(op), wA.copy(), wB.copy())
        assert abs(Q - Qo) / Qo < 1e-12
        assert np.max(np.abs(phiA[0] - pa)) < 1e-12 and np.max(np.abs(phiB[0] - pb)) < 1e-12
        s.set_fields(wA, wB)
        _, fe = s.iterate(2)
        phiA, _, _, _ = s.get_fields()
    pa, _, hist = orc.run(op, wA.copy(), wB.copy(), 2)
    assert abs(fe[0][0] - hist[-1][0]) < 1e-11 * max(1.0, abs(hist[-1][0]))
    assert np.max(np.abs(phiA[0] - pa)) < 1e-11


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.gpu
@GPU
@given(nx=EVEN, ny=EVEN, nz=st.sampled_from([1, 2, 4, 6, 8]), fA=st.floats(0.2, 0.8),
       ds=st.sampled_from([0.01, 0.015, 0.02, 0.03, 0.05]), mu=st.sampled_from([0.0, 1.0, -0.7]), ckpt=st.booleans(),
       seed=st.integers(0, 2**20))
def test_gpu_chain_adjoint_matches_oracle_random_cases(pkg, orc, nx, ny, nz, fA, ds, mu, ckpt, seed):
    """L = sum(lam phi) + mu ln Q and dL/du by the library's adjoint march (its
    checkpointed pair walk for the random Ns and spacing when ckpt) against
    the oracle's independent FP64 adjoint: L 1e-11, gradient 1e-10."""
    print("case", (nx, ny, nz), fA, ds, "mu", mu, "ckpt", ckpt, "seed", seed, flush=True)
    p = pkg.default_params(nx=nx, ny=ny, nz=nz, lx=3.0, ly=3.5, lz=2.5, chiN=14.0, fA=fA, ds=ds, precision=0,
                           checkpoint=int(ckpt))
    N = nx * ny * nz
    rng = np.random.default_rng(seed)
    uA, uB = rng.uniform(-1.5, 1.5, N), rng.uniform(-1.5, 1.5, N)
    lA, lB = rng.standard_normal(N), rng.standard_normal(N)
    with pkg.Solver(p) as s:
        L, gA, gB, pa, pb, Q = s.conc_adjoint(uA, uB, lA, lB, mu)
    op = orc.params(nx, ny, nz, 3.0, 3.5, 2.5, 14.0, fA, ds, pi=p.pi, wopt=p.wopt, wcmp=p.wcmp)
    Lo, gAo, gBo, pao, pbo, Qo = orc.conc_adjoint(op, orc.ktable(op), uA, uB, lA, lB, mu)
    assert abs(L[0] - Lo) <= 1e-11 * max(1.0, abs(Lo))
    assert abs(Q[0] - Qo) / Qo < 1e-12 and _rel(pa[0], pao) < 1e-12
    assert _rel(gA[0], gAo) < 1e-10 and _rel(gB[0], gBo) < 1e-10


@pytest.mark.gpu
@settings(max_examples=12, derandomize=True, deadline=None,
          suppress_health_check=[HealthCheck.function_scoped_fixture])
@given(n=st.sampled_from([4, 6, 8]), fA=st.floats(0.3, 0.7), ds=st.sampled_from([0.02, 0.03, 0.05]),
       mixing=st.sampled_from([0, 1]), ckpt=st.booleans(), iters=st.integers(1, 6), seed=st.integers(0, 2**20))
def test_gpu_loop_gradient_matches_fd_random_cases(pkg, n, fA, ds, mixing, ckpt, iters, seed):
    """The loop adjoint's directional derivative along a random direction
    against central differences (eps = 1e-5) of its own forward: 1e-6."""
    print("case", n, fA, ds, "mixing", mixing, "ckpt", ckpt, "iters", iters, "seed", seed, flush=True)
    p = pkg.default_params(nx=n, ny=n, nz=n, lx=3.5, ly=3.5, lz=3.5, chiN=16.0, fA=fA, ds=ds, precision=0,
                           mixing=mixing, anderson_m=3, anderson_mix=0.2, anderson_reg=1e-3, checkpoint=int(ckpt))
    N = n ** 3
    rng = np.random.default_rng(seed)
    w0A, w0B = pkg.fields.seeded_fields(n, n, n, 3.5, 3.5, 3.5, "lamellar", 0.7, 0.1, seed)
    target = 0.5 + 0.3 * np.cos(2 * np.pi * pkg.fields.coords(n, n, n, 3.5, 3.5, 3.5)[0] / 3.5)
    h = 0.05 * rng.standard_normal(N)
    v = rng.standard_normal(N)
    eps = 1e-5
    with pkg.Solver(p) as s:
        _, g, _ = s.adjoint_loop(h, w0A, w0B, iters, target)
        lp = s.adjoint_loop(h + eps * v, w0A, w0B, iters, target, grad=False)[0][0]
        lm = s.adjoint_loop(h - eps * v, w0A, w0B, iters, target, grad=False)[0][0]
    fd = (lp - lm) / (2 * eps)
    ad = float(g[0] @ v)
    assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1e-3), (fd, ad)
