"""GPU tests of sqrt(Ns) checkpointing and the discrete adjoint.

Checks: checkpointed concentrations equal the full-store ones; the chain
adjoint equals the oracle's independent FP64 adjoint (1e-10) and central
finite differences; dlnQ/du = -phi/N; the unrolled-loop gradient (simple and
Anderson mixing, external field h on A) equals central finite differences of
the loop's own forward along random directions (eps = 1e-5); the FP32-form
gradient is stated against the FP64 form."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params(pkg, n, L, chiN, fA, ds, prec=0, **kw):
    nx, ny, nz = n if isinstance(n, tuple) else (n, n, n)
    return pkg.default_params(nx=nx, ny=ny, nz=nz, lx=L, ly=L, lz=L, chiN=chiN, fA=fA, ds=ds, precision=prec, **kw)


def _oparams(orc, p):
    return orc.params(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, p.chiN, p.fA, p.ds, pi=p.pi, wopt=p.wopt, wcmp=p.wcmp)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("shape,fA,ds,prec", [((16, 16, 16), 0.35, 0.01, 0), ((12, 20, 8), 0.3, 0.02, 0),
                                              ((32, 32, 32), 0.5, 0.01, 1), ((8, 8, 8), 0.45, 0.05, 0)])
def test_checkpoint_conc_matches_full(pkg, shape, fA, ds, prec):
    print("case", shape, fA, ds, prec, flush=True)
    out = []
    for ck in (0, 1):
        p = _params(pkg, shape, 4.0, 16.0, fA, ds, prec=prec, checkpoint=ck)
        wA, wB = pkg.fields.seeded_fields(p.nx, p.ny, p.nz, 4.0, 4.0, 4.0, "gyroid", 1.0, 0.2, 71)
        with pkg.Solver(p) as s:
            s.set_fields(wA, wB)
            Q = s.concentrations()[0]
            phiA, phiB, _, _ = s.get_fields()
            if ck and ds <= 0.02:
                # fewer slices held (for the 20-step chain the checkpoints, their
                # opened states and the working slices outnumber its 23 slices)
                assert s.device_bytes < out[0][3]
            out.append((Q, phiA, phiB, s.device_bytes))
    tol = 1e-13 if prec == 0 else 1e-6
    assert abs(out[0][0] - out[1][0]) / out[0][0] < tol
    assert _rel(out[1][1], out[0][1]) < tol and _rel(out[1][2], out[0][2]) < tol


@pytest.mark.parametrize("n,prec,tiling", [(32, 1, [0, 0, 0, 2]), (16, 1, [0, 0, 0, 4]), (12, 1, [0, 0, 0, 0]),
                                           (16, 1, [0, 0, 0, 0]), (32, 2, [0, 0, 0, 2]), (16, 2, [0, 0, 0, 0]),
                                           (16, 0, [0, 0, 0, 2]), (12, 0, [0, 0, 0, 0])])
def test_checkpoint_remarch_bit_identical(pkg, n, prec, tiling):
    """With step-boundary kernels the checkpointed evaluation keeps the march's
    opened state at its checkpoints and re-marches from it, so every slice it
    re-marches equals the full store's bit for bit, in every form and tiling:
    Q (from the same z(Ns)) is identical and phi differs only by the order of
    the FP64 contour sums (bound 1e-13)."""
    print("case", n, prec, tiling, flush=True)
    out = []
    for ck in (0, 1):
        p = _params(pkg, n, 4.0, 18.0, 0.5, 0.02, prec=prec, checkpoint=ck)
        wA, wB = pkg.fields.seeded_fields(n, n, n, 4.0, 4.0, 4.0, "lamellar", 0.5, 0.1, 73)
        with pkg.Solver(p) as s:
            s.set_tiling(tiling)
            s.set_fields(wA, wB)
            Q = s.concentrations()[0]
            phiA, phiB, _, _ = s.get_fields()
            out.append((Q, phiA.copy(), phiB.copy()))
    assert out[0][0] == out[1][0]
    assert _rel(out[1][1], out[0][1]) < 1e-13 and _rel(out[1][2], out[0][2]) < 1e-13


@pytest.mark.parametrize("ck", [0, 1])
@pytest.mark.parametrize("mu", [0.0, 1.0])
def test_conc_adjoint_vs_oracle(pkg, orc, ck, mu):
    shape = (12, 10, 8)
    p = _params(pkg, shape, 3.0, 14.0, 0.4, 0.02, checkpoint=ck)
    N = 960
    rng = np.random.default_rng(5)
    uA, uB = rng.uniform(-1.5, 1.5, N), rng.uniform(-1.5, 1.5, N)
    lA, lB = rng.standard_normal(N), rng.standard_normal(N)
    with pkg.Solver(p) as s:
        L, gA, gB, pa, pb, Q = s.conc_adjoint(uA, uB, lA, lB, mu)
    op = _oparams(orc, p)
    Lo, gAo, gBo, pao, pbo, Qo = orc.conc_adjoint(op, orc.ktable(op), uA, uB, lA, lB, mu)
    print(f"ck={ck} mu={mu}: relL={abs(L[0]-Lo)/abs(Lo):.2e} relgA={_rel(gA[0], gAo):.2e} relgB={_rel(gB[0], gBo):.2e}")
    assert abs(L[0] - Lo) / abs(Lo) < 1e-11
    assert abs(Q[0] - Qo) / Qo < 1e-12
    assert _rel(pa[0], pao) < 1e-12
    assert _rel(gA[0], gAo) < 1e-10 and _rel(gB[0], gBo) < 1e-10


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_dlnQ_du_is_minus_phi_over_N(pkg, prec):
    p = _params(pkg, (16, 16, 16), 4.0, 18.0, 0.5, 0.01, prec=prec)
    N = 4096
    wA, wB = pkg.fields.seeded_fields(16, 16, 16, 4.0, 4.0, 4.0, "lamellar", 1.5, 0.2, 72)
    with pkg.Solver(p) as s:
        L, gA, gB, pa, pb, Q = s.conc_adjoint(wA, wB, None, None, 1.0)
    tol = {0: 1e-12, 1: 1e-5, 2: 1e-6}[prec]
    eA, eB = _rel(gA[0], -pa[0] / N), _rel(gB[0], -pb[0] / N)
    print(f"prec={prec}: rel |dlnQ/du + phi/N| A={eA:.2e} B={eB:.2e}")
    assert eA < tol and eB < tol


def test_conc_adjoint_fd_gpu(pkg):
    p = _params(pkg, (16, 16, 8), 4.0, 16.0, 0.35, 0.02)
    N = 2048
    rng = np.random.default_rng(6)
    uA, uB = rng.uniform(-1, 1, N), rng.uniform(-1, 1, N)
    lA, lB = rng.standard_normal(N), rng.standard_normal(N)
    eps = 1e-5
    with pkg.Solver(p) as s:
        L, gA, gB, *_ = s.conc_adjoint(uA, uB, lA, lB, 0.3)
        for k in range(4):
            vA, vB = rng.standard_normal(N), rng.standard_normal(N)
            Lp = s.conc_adjoint(uA + eps * vA, uB + eps * vB, lA, lB, 0.3)[0][0]
            Lm = s.conc_adjoint(uA - eps * vA, uB - eps * vB, lA, lB, 0.3)[0][0]
            fd = (Lp - Lm) / (2 * eps)
            ad = float(gA[0] @ vA + gB[0] @ vB)
            assert abs(fd - ad) / abs(ad) < 1e-7, (k, fd, ad)


def _loop_case(pkg, mixing, prec=0, n=12, iters=8, **kw):
    p = _params(pkg, (n, n, n), 4.0, 18.0, 0.5, 0.02, prec=prec, mixing=mixing, anderson_m=4, anderson_mix=0.2, **kw)
    N = n ** 3
    w0A, w0B = pkg.fields.seeded_fields(n, n, n, 4.0, 4.0, 4.0, "lamellar", 0.5, 0.1, 73)
    x = pkg.fields.coords(n, n, n, 4.0, 4.0, 4.0)[0]
    target = 0.5 + 0.3 * np.cos(2 * np.pi * x / 4.0 + 0.4)
    h = 0.05 * pkg.fields.splitmix_uniform(74, N)
    return p, N, h, w0A, w0B, target, iters


@pytest.mark.parametrize("mixing,ridge", [(0, 0.0), (1, 0.0), (1, 1e-3)])
def test_adjoint_loop_fd(pkg, mixing, ridge):
    """Gradient of the loss through the unrolled loop vs central finite
    differences of the same forward along random directions (Anderson with and
    without the ridge on its least squares)."""
    p, N, h, w0A, w0B, target, iters = _loop_case(pkg, mixing, anderson_reg=ridge)
    rng = np.random.default_rng(8)
    # central differences: truncation error ~ eps^2, rounding noise of the
    # Anderson trajectory ~ 1/eps; measured relative mismatch against the
    # adjoint on a direction whose derivative is small (1e-2): 2.6e-6 at
    # eps = 1e-4, 2e-8 at 3e-5, 1.1e-6 at 1e-5 and 2.8e-6 at 3e-6, so eps sits at the crossover
    eps = 3e-5
    with pkg.Solver(p) as s:
        loss, g, _ = s.adjoint_loop(h, w0A, w0B, iters, target)
        worst = 0.0
        for k in range(6):
            v = rng.standard_normal(N)
            lp = s.adjoint_loop(h + eps * v, w0A, w0B, iters, target, grad=False)[0][0]
            lm = s.adjoint_loop(h - eps * v, w0A, w0B, iters, target, grad=False)[0][0]
            fd = (lp - lm) / (2 * eps)
            ad = float(g[0] @ v)
            worst = max(worst, abs(fd - ad) / max(abs(ad), 1e-12))
    print(f"mixing={mixing}: loss={loss[0]:.6e} worst rel FD mismatch={worst:.2e}")
    assert worst < 1e-6


def test_adjoint_loop_fp32_vs_fp64(pkg):
    res = {}
    for prec in (0, 1):
        p, N, h, w0A, w0B, target, iters = _loop_case(pkg, 1, prec=prec)
        with pkg.Solver(p) as s:
            res[prec] = s.adjoint_loop(h, w0A, w0B, iters, target)
    dl = abs(res[1][0][0] - res[0][0][0]) / res[0][0][0]
    dg = _rel(res[1][1][0], res[0][1][0])
    print(f"FP32 vs FP64 adjoint: rel dloss={dl:.2e} rel dgrad={dg:.2e}")
    assert dl < 1e-4 and dg < 1e-3


@pytest.mark.parametrize("mixing,ridge", [(0, 0.0), (1, 0.0), (1, 1e-3)])
def test_adjoint_loop_forward_matches_iterate(pkg, mixing, ridge):
    """The adjoint's forward (its own mixing kernels, least squares and
    ridge) follows the solver's own loop."""
    n = 12
    p = _params(pkg, (n, n, n), 4.0, 18.0, 0.5, 0.02, mixing=mixing, anderson_m=4, anderson_mix=0.2,
                anderson_reg=ridge)
    N = n ** 3
    w0A, w0B = pkg.fields.seeded_fields(n, n, n, 4.0, 4.0, 4.0, "lamellar", 0.5, 0.1, 75)
    with pkg.Solver(p) as s:
        _, _, phi_last = s.adjoint_loop(np.zeros(N), w0A, w0B, 6, np.zeros(N), grad=False)
        s.set_fields(w0A, w0B)
        s.iterate(5)
        # sixth chain evaluation of the solver's loop: subtract means, then getConc
        _, _, wA, wB = s.get_fields()
        wA = wA[0] - wA[0].mean()
        wB = wB[0] - wB[0].mean()
        _, _, _, pa, _, _ = s.conc_adjoint(wA, wB)
    assert _rel(phi_last[0], pa[0]) < (1e-11 if mixing == 0 else 1e-9)


@pytest.mark.parametrize("mixing", [0, 1])
def test_adjoint_loop_batch_matches_single(pkg, mixing):
    """Two instances in one context (their own start fields, h and target)
    give each instance's loss and gradient as a context of one would: the
    per-instance sums, least squares and chain adjoint do not mix instances."""
    p1, N, h, w0A, w0B, target, iters = _loop_case(pkg, mixing)
    n = p1.nx
    v0A, v0B = pkg.fields.seeded_fields(n, n, n, 4.0, 4.0, 4.0, "lamellar", 0.7, 0.15, 91)
    h2 = 0.05 * pkg.fields.splitmix_uniform(92, N)
    t2 = 0.5 + 0.2 * np.sin(2 * np.pi * pkg.fields.coords(n, n, n, 4.0, 4.0, 4.0)[1] / 4.0)
    single = []
    for a, b, hh, tt in ((w0A, w0B, h, target), (v0A, v0B, h2, t2)):
        with pkg.Solver(p1) as s:
            single.append(s.adjoint_loop(hh, a, b, iters, tt))
    p2, _, _, _, _, _, _ = _loop_case(pkg, mixing, batch=2)
    with pkg.Solver(p2) as s:
        loss, g, _ = s.adjoint_loop(np.concatenate([h, h2]), np.concatenate([w0A, v0A]), np.concatenate([w0B, v0B]),
                                    iters, np.concatenate([target, t2]))
    for b in range(2):
        dl = abs(loss[b] - single[b][0][0]) / abs(single[b][0][0])
        dg = _rel(g[b], single[b][1][0])
        print(f"mixing={mixing} instance {b}: rel dloss={dl:.2e} rel dgrad={dg:.2e}")
        assert dl < 1e-12 and dg < 1e-10


@pytest.mark.parametrize("mixing,prec", [(0, 0), (1, 0), (0, 1), (1, 1), (0, 2), (1, 2)])
def test_adjoint_loop_checkpoint_matches_full(pkg, mixing, prec):
    """sqrt(Ns) checkpointing in the loop adjoint (the reverse sweep restores
    each iteration's checkpoints and opened states instead of re-running its
    forward march) gives the full store's loss and gradient.  The forward
    evaluations and the reverse sweep's pair walk re-march from the march's
    opened states, so in the FP32-storage forms the loss is the full store's
    bit for bit, and so is the gradient through simple mixing.  Through
    Anderson mixing the gradients are equal for up to three iterations; from
    the fourth (m = 4) a difference of 2e-15 appears, which the Anderson loop
    amplifies through the FP32 rounding of the W factors to 3.0e-6 (FP32) and
    7.5e-7 (FP32X) after 8 iterations: bound 1e-5."""
    res = {}
    for ck in (0, 1):
        p, N, h, w0A, w0B, target, iters = _loop_case(pkg, mixing, prec=prec, checkpoint=ck)
        with pkg.Solver(p) as s:
            res[ck] = s.adjoint_loop(h, w0A, w0B, iters, target)
    dl = abs(res[1][0][0] - res[0][0][0]) / abs(res[0][0][0])
    dg = _rel(res[1][1][0], res[0][1][0])
    print(f"mixing={mixing} prec={prec}: checkpoint vs full rel dloss={dl:.2e} rel dgrad={dg:.2e}")
    if prec == 0:
        assert dl < 1e-12 and dg < 1e-10
    else:
        assert dl == 0.0 and (dg == 0.0 if mixing == 0 else dg < 1e-5)


def test_adjoint_loop_deterministic(pkg):
    """Two calls give the same bits: every device sum runs in a fixed order."""
    p, N, h, w0A, w0B, target, iters = _loop_case(pkg, 1, anderson_reg=1e-3)
    with pkg.Solver(p) as s:
        l1, g1, _ = s.adjoint_loop(h, w0A, w0B, iters, target)
        l2, g2, _ = s.adjoint_loop(h, w0A, w0B, iters, target)
    assert np.array_equal(l1, l2) and np.array_equal(g1, g2)


def test_adjoint_loop_long_history(pkg):
    """A 24-deep Anderson history (300 Gram entries per update, more than one
    per reduction thread): the adjoint's forward follows the solver's own
    Anderson loop, and its gradient matches central differences."""
    n, m, iters = 8, 24, 28
    p = _params(pkg, (n, n, n), 4.0, 16.0, 0.5, 0.02, mixing=1, anderson_m=m, anderson_mix=0.2, anderson_reg=1e-2)
    N = n ** 3
    w0A, w0B = pkg.fields.seeded_fields(n, n, n, 4.0, 4.0, 4.0, "lamellar", 0.5, 0.1, 93)
    x = pkg.fields.coords(n, n, n, 4.0, 4.0, 4.0)[0]
    target = 0.5 + 0.3 * np.cos(2 * np.pi * x / 4.0)
    h = 0.02 * pkg.fields.splitmix_uniform(94, N)
    rng = np.random.default_rng(9)
    eps = 1e-5
    with pkg.Solver(p) as s:
        _, _, phi_last = s.adjoint_loop(np.zeros(N), w0A, w0B, iters, target, grad=False)
        s.set_fields(w0A, w0B)
        s.iterate(iters - 1)
        _, _, wA, wB = s.get_fields()
        _, _, _, pa, _, _ = s.conc_adjoint(wA[0] - wA[0].mean(), wB[0] - wB[0].mean())
        dfwd = _rel(phi_last[0], pa[0])
        loss, g, _ = s.adjoint_loop(h, w0A, w0B, iters, target)
        worst = 0.0
        for _ in range(3):
            v = rng.standard_normal(N)
            lp = s.adjoint_loop(h + eps * v, w0A, w0B, iters, target, grad=False)[0][0]
            lm = s.adjoint_loop(h - eps * v, w0A, w0B, iters, target, grad=False)[0][0]
            fd = (lp - lm) / (2 * eps)
            ad = float(g[0] @ v)
            worst = max(worst, abs(fd - ad) / max(abs(ad), 1e-12))
    print(f"m={m}, {iters} iterations: forward vs iterate {dfwd:.2e}, worst rel FD mismatch {worst:.2e}")
    assert dfwd < 1e-10 and worst < 1e-6


def test_torch_autograd_wrapper(pkg):
    """torch.autograd.gradcheck (FP64 finite differences) through the wrapper on a 4^3 melt."""
    import importlib
    import torch
    tw = importlib.import_module("scft_amd.torch_adjoint")
    n = 4
    p = _params(pkg, (n, n, n), 2.0, 12.0, 0.5, 0.05)
    N = n ** 3
    w0A, w0B = pkg.fields.seeded_fields(n, n, n, 2.0, 2.0, 2.0, "lamellar", 0.5, 0.1, 76)
    target = 0.5 + 0.2 * np.cos(2 * np.pi * pkg.fields.coords(n, n, n, 2.0, 2.0, 2.0)[0] / 2.0)
    with pkg.Solver(p) as s:
        h = torch.tensor(0.05 * pkg.fields.splitmix_uniform(77, N), dtype=torch.float64, requires_grad=True)
        args = (s, torch.tensor(w0A), torch.tensor(w0B), 5, torch.tensor(target))
        loss = tw.SCFTLoopLoss.apply(h, *args)
        loss.backward()
        _, g, _ = s.adjoint_loop(h.detach().numpy(), w0A, w0B, 5, target)
        assert np.allclose(h.grad.numpy(), g[0], rtol=0, atol=0)
        assert torch.autograd.gradcheck(lambda x: tw.SCFTLoopLoss.apply(x, *args), (h,), eps=1e-6, atol=1e-6,
                                        rtol=1e-5)
