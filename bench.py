"""SCFT benchmark: iterations/s of the AB-diblock SCFT solver on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-128]

A "step" is one SCFT iteration (mean subtraction, the packed propagator march
over Ns = 200 contour steps, Q, the trapezoid contour integral, the Anderson
field update and the free energy) of the cylinder workload: AB diblock f=0.3,
chiN=20, 64^3 periodic grid, box 4.2 Rg, ds=0.005, Anderson history 10, seeded
hexagonal-star fields.  `value` is the FP32-propagator form (FP64 Q, phi, F,
fields and least squares): the median of 5 windows of K iterations (K = --steps;
by default the K that makes a window last at least 1 s), timed with HIP events
on the solver's stream after max(W, Anderson m) warm-up iterations.  Fields
are resident in HBM before the timed region.  Beside it: the FP64 and FP32X
forms, the 128^3 gyroid workload (sqrt(Ns) checkpointing, its storage mode,
and the full store), every form also timed in windows of at least 1 s
(median of 5), iteration HBM fractions against the nominal 8 TB/s and a
measured copy peak, the per-kernel roofline legs, the errors of both
FP32-storage forms against FP64 (one evaluation, along the timed trajectory
and after 500 iterations), the 256-instance sweep, the adjoint workload and
the CPU baseline (the oracle on the host cores).
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "scft_ver1.01_amd")
HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, published


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "scft_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["scft_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Ridge on the Anderson least squares (U += ridge * tr(U)/m * I) for both forms
# of the c2/c3 runs.  Without it the FP32 form's history of nearly collinear
# residual differences amplifies FP32 rounding, and at 128^3 with m = 20 the
# iteration leaves the basin for tens of iterations before it returns.
ANDERSON_RIDGE = 1e-2


def workload_params(pkg, name, prec, **kw):
    w = pkg.fields.WORKLOADS[name]
    n = w["n"]
    kw.setdefault("anderson_reg", ANDERSON_RIDGE)
    return pkg.default_params(nx=n, ny=n, nz=n, lx=w["L"], ly=w["L"], lz=w["L"], chiN=w["chiN"],
                              fA=w["fA"], ds=w["ds"], precision=prec, mixing=pkg.MIX_ANDERSON,
                              anderson_m=w["anderson_m"], anderson_mix=0.1, **kw)


def convergence(pkg, name, n_iter, choices):
    """The FP64, FP32 and FP32X forms from the same seeded fields for n_iter
    iterations (ridge ANDERSON_RIDGE): F and the incompressibility error
    max|1 - phiA - phiB| at iterations 10, 25, 50, 100, ... and |dF| of each
    FP32-storage form against FP64."""
    wA, wB = pkg.fields.workload_fields(name)
    marks = [k for k in (10, 25, 50, 100, 200, 400) if k <= n_iter]
    out = {}
    for prec, ch in ((pkg.FP64, choices[0]), (pkg.FP32, choices[1]), (pkg.FP32X, choices[2])):
        with pkg.Solver(workload_params(pkg, name, prec)) as s:
            s.set_tiling(ch)
            s.set_fields(wA, wB)
            done, hist = 0, {}
            for k in marks:
                _, fe = s.iterate(k - done)
                done = k
                hist[k] = (float(fe[0][0]), float(fe[0][4]))
            out[prec] = hist
    return {"iterations": marks, "ridge": ANDERSON_RIDGE,
            "F_fp64": [out[pkg.FP64][k][0] for k in marks], "err_fp64": [out[pkg.FP64][k][1] for k in marks],
            "F_fp32": [out[pkg.FP32][k][0] for k in marks], "err_fp32": [out[pkg.FP32][k][1] for k in marks],
            "F_fp32x": [out[pkg.FP32X][k][0] for k in marks], "err_fp32x": [out[pkg.FP32X][k][1] for k in marks],
            "dF": [abs(out[pkg.FP32][k][0] - out[pkg.FP64][k][0]) for k in marks],
            "dF_fp32x": [abs(out[pkg.FP32X][k][0] - out[pkg.FP64][k][0]) for k in marks]}


def iter_bytes(N, Ns, b):
    """Algorithmic HBM bytes of one SCFT iteration: N*[b*(6 Ns + 2 (Ns+2)) + 64]."""
    return N * (b * (6 * Ns + 2 * (Ns + 2)) + 64)


WINDOWS = 5        # timed windows per figure; the median is reported
MIN_WINDOW_S = 1.0  # a window of the long-window figure lasts at least this long


def windows_ms(time_fn, k, windows=WINDOWS):
    """elapsed ms of `windows` back-to-back windows of k iterations each"""
    return [time_fn(k) for _ in range(windows)]


def long_window_k(time_fn, min_s=MIN_WINDOW_S):
    """iterations per window so that one window lasts at least min_s"""
    ms = time_fn(5) / 5
    return max(5, int(np.ceil(min_s * 1e3 / ms)))


def run_form(pkg, name, prec, steps, warmup, choice=None, long_window=True, **kw):
    """One form of a workload: autotune (or the given tiling), warm-up of
    max(warmup, Anderson m) iterations so that the history is full, then the
    median of WINDOWS windows of `steps` iterations (steps <= 0: windows of at
    least MIN_WINDOW_S) and, with long_window, also the median of WINDOWS
    windows of at least MIN_WINDOW_S each.  Returns (solver, timing, fe, phiA)."""
    p = workload_params(pkg, name, prec, **kw)
    wA, wB = pkg.fields.workload_fields(name)
    s = pkg.Solver(p)
    form = {pkg.FP64: "fp64", pkg.FP32: "fp32", pkg.FP32X: "fp32x"}[prec]
    if choice is None:
        choice, tms = s.autotune(20)
        log(f"[bench] {name} {form} autotune choice={choice} pass_ms={tms}")
    else:
        s.set_tiling(choice)
        tms = None
        log(f"[bench] {name} {form} tiling set to {choice}")
    s.tuning = (choice, tms)
    s.set_fields(wA, wB)
    wrun = max(warmup, pkg.fields.WORKLOADS[name].get("anderson_m", 0))
    log(f"[bench] {name} {form} warmup {wrun}")
    s.iterate(wrun)
    k = steps if steps > 0 else long_window_k(s.time_iterations)
    log(f"[bench] {name} {form} timing {WINDOWS} windows of {k} iterations")
    w = windows_ms(s.time_iterations, k)
    t = {"k": k, "windows_ms": w, "ms_per_iter": float(np.median(w)) / k, "warmup_run": wrun}
    if long_window and k * t["ms_per_iter"] < MIN_WINDOW_S * 1e3:
        k1 = long_window_k(s.time_iterations)
        log(f"[bench] {name} {form} timing {WINDOWS} windows of {k1} iterations (>= {MIN_WINDOW_S} s each)")
        w1 = windows_ms(s.time_iterations, k1)
        t.update({"k_long": k1, "windows_long_ms": w1, "ms_per_iter_long": float(np.median(w1)) / k1})
    else:
        t.update({"k_long": k, "windows_long_ms": w, "ms_per_iter_long": t["ms_per_iter"]})
    fe = s.free_energy()
    phiA, _, _, _ = s.get_fields()
    return s, t, fe[0].copy(), phiA[0].copy()


def tol_overhead(pkg, name, prec, choice, n_iter=200):
    """Host-clock rate of scft_iterate(n) with a tolerance that is never met
    (the convergence check of every iteration, one iteration behind the GPU)
    against the same call without one."""
    wA, wB = pkg.fields.workload_fields(name)
    out = {}
    with pkg.Solver(workload_params(pkg, name, prec)) as s:
        s.set_tiling(choice)
        s.set_fields(wA, wB)
        s.iterate(20)
        for tol in (0.0, 1e-300, 0.0, 1e-300):
            t0 = time.perf_counter()
            s.iterate(n_iter, tol)
            out["with_tol" if tol else "without_tol"] = n_iter / (time.perf_counter() - t0)
    out["ratio"] = out["with_tol"] / out["without_tol"]
    out["iterations"] = n_iter
    return out


def precision_schedule(pkg, name, n_iter, switch_at, choices, ckpt):
    """A solve that runs its first switch_at iterations in the FP32 form and the
    rest in FP32X (scft_set_precision on the same context: fields, phi and the
    Anderson history carry over), against the FP64 form over the same n_iter
    iterations from the seeded fields: |dF|, max|dphiA| and the incompressibility
    error at the end, and the host-clock rate of each whole solve.  choices:
    tilings (fp64, fp32, fp32x)."""
    wA, wB = pkg.fields.workload_fields(name)
    res = {}
    with pkg.Solver(workload_params(pkg, name, pkg.FP64, checkpoint=ckpt)) as s:
        s.set_tiling(choices[0])
        s.set_fields(wA, wB)
        s.iterate(1)  # graph capture outside the timed solve
        s.set_fields(wA, wB)
        t0 = time.perf_counter()
        _, fe = s.iterate(n_iter)
        res["fp64"] = (time.perf_counter() - t0, fe[0].copy(), s.get_fields()[0][0].copy())
    with pkg.Solver(workload_params(pkg, name, pkg.FP32, checkpoint=ckpt)) as s:
        s.set_tiling(choices[1])
        s.set_fields(wA, wB)
        s.iterate(1)
        s.set_fields(wA, wB)
        t0 = time.perf_counter()
        s.iterate(switch_at)
        s.set_precision(pkg.FP32X)
        s.set_tiling(choices[2])
        _, fe = s.iterate(n_iter - switch_at)
        res["mixed"] = (time.perf_counter() - t0, fe[0].copy(), s.get_fields()[0][0].copy())
    (t64, f64, p64), (tm, fm, pm) = res["fp64"], res["mixed"]
    return {"iterations": n_iter, "switch_at": switch_at, "checkpoint": ckpt,
            "schedule": f"FP32 for {switch_at} iterations, then FP32X",
            "dF_vs_fp64": abs(float(fm[0] - f64[0])), "max_dphiA_vs_fp64": float(np.max(np.abs(pm - p64))),
            "incomp_fp64": float(f64[4]), "incomp_schedule": float(fm[4]), "F_fp64": float(f64[0]),
            "fp64_iters_per_s": n_iter / t64, "schedule_iters_per_s": n_iter / tm, "ratio": t64 / tm}


def trajectory_error(pkg, name, n_iter, choices):
    """The FP64, FP32 and FP32X forms from the same seeded fields for n_iter
    iterations: |dF|, max|dphiA|, max|dphiB| of each FP32-storage form against
    FP64 after the last one."""
    wA, wB = pkg.fields.workload_fields(name)
    res = {}
    for prec, ch in ((pkg.FP64, choices[0]), (pkg.FP32, choices[1]), (pkg.FP32X, choices[2])):
        with pkg.Solver(workload_params(pkg, name, prec)) as s:
            s.set_tiling(ch)
            s.set_fields(wA, wB)
            _, fe = s.iterate(n_iter)
            pa, pb, _, _ = s.get_fields()
            res[prec] = (fe[0].copy(), pa[0].copy(), pb[0].copy())
    f64, a64, b64 = res[pkg.FP64]
    out = {"iterations": n_iter, "F_fp64": float(f64[0]), "incomp_fp64": float(f64[4])}
    for prec, key in ((pkg.FP32, "fp32"), (pkg.FP32X, "fp32x")):
        f, a, b = res[prec]
        out[f"dF_{key}"] = abs(float(f[0] - f64[0]))
        out[f"max_dphiA_{key}"] = float(np.max(np.abs(a - a64)))
        out[f"max_dphiB_{key}"] = float(np.max(np.abs(b - b64)))
        out[f"incomp_{key}"] = float(f[4])
    return out


def single_eval_error(pkg, name, wA, wB, prec32=None):
    """One evaluation (getConc, field update, F) of an FP32-storage form
    (FP32 by default, or FP32X) and the FP64 form on identical fields: |dQ|/Q,
    |dF| and max|dphi| against the FP64 form (which matches the FP64 oracle to
    ~1e-12 on every tested grid)."""
    prec32 = pkg.FP32 if prec32 is None else prec32
    out = {}
    for prec in (pkg.FP64, prec32):
        p = workload_params(pkg, name, prec)
        with pkg.Solver(p) as s:
            s.set_fields(wA, wB)
            Q = s.concentrations()[0]
            phiA, phiB, _, _ = s.get_fields()
            s.update_fields()
            fe = s.free_energy()[0]
            out[prec] = (Q, fe[0], phiA[0].copy(), phiB[0].copy())
    q64, f64, a64, b64 = out[pkg.FP64]
    q32, f32, a32, b32 = out[prec32]
    return {"relQ": abs(q32 - q64) / q64, "dF": abs(f32 - f64), "F_fp64": f64,
            "max_dphiA": float(np.max(np.abs(a32 - a64))), "max_dphiB": float(np.max(np.abs(b32 - b64)))}


KERNEL_NAMES = {"x": "k_line_conv x pass", "y": "k_line_conv y pass", "z": "k_line_conv z pass",
                "xy_plane": "k_plane_xy", "xy_split": "k_plane_split",
                "fused_x": "k_line_fused x boundary", "fused_y": "k_line_fused y boundary",
                "fused_z": "k_line_fused z boundary", "fused_plane": "k_plane_xy boundary form",
                "fused_split": "k_plane_split boundary form",
                "fused_split3": "k_plane_split boundary form, three-level tiling",
                "fused_split_y": "k_plane_split boundary form, lane-local y lines",
                "fused_split3_y": "k_plane_split boundary form, three-level tiling, lane-local y lines"}
CONTOUR_CODE = 16  # scft_time_pass code of the contour kernel
CODES = {"x": 0, "y": 1, "z": 2, "xy_plane": 3, "xy_split": 4, "fused_x": 5, "fused_y": 6, "fused_z": 7,
         "fused_plane": 8, "fused_split": 9, "fused_split3": 11, "fused_split_y": 13, "fused_split3_y": 15}


def c5_adjoint(pkg, n_dirs=2):
    """c5: loss = ||phiA - phi_target||^2 after 50 unrolled Anderson iterations
    (64^3, f = 0.5, chiN = 18, ds = 0.01, sqrt(Ns) checkpointing), gradient with
    respect to an external field h on A by the discrete adjoint.  Reports the
    time per gradient in the FP64, FP32 and FP32X forms, each form's
    directional derivative against central differences (eps = 1e-5) of the
    FP64 loss along n_dirs random directions (worst and median relative
    mismatch), and the FP32 and FP32X gradients' distance from the FP64 one.
    Anderson m = 10 with the relative ridge 1e-2 on its least squares: without
    it the 50-iteration loss is not smooth on the finite-difference scale (the
    least squares turns ill-conditioned)."""
    w = pkg.fields.WORKLOADS["c5_adjoint_64"]
    n, L = w["n"], w["L"]
    N = n ** 3
    w0A, w0B = pkg.fields.workload_fields("c5_adjoint_64")
    x = pkg.fields.coords(n, n, n, L, L, L)[0]
    u = pkg.fields.splitmix_uniform(55, 2)
    target = 0.5 + (0.2 + 0.2 * u[0]) * np.cos(2 * np.pi * x / L + 2 * np.pi * u[1])
    h = np.zeros(N)
    out = {}
    grads = {}
    keys = {pkg.FP64: "fp64", pkg.FP32: "fp32", pkg.FP32X: "fp32x"}
    fd = []
    rng = np.random.default_rng(5)
    dirs = [rng.standard_normal(N) for _ in range(n_dirs)]
    for prec in (pkg.FP64, pkg.FP32, pkg.FP32X):
        p = pkg.default_params(nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, chiN=w["chiN"], fA=w["fA"], ds=w["ds"],
                               precision=prec, mixing=pkg.MIX_ANDERSON, anderson_m=10, anderson_mix=0.1,
                               checkpoint=1, anderson_reg=1e-2)
        with pkg.Solver(p) as s:
            s.autotune(5)
            s.adjoint_loop(h, w0A, w0B, 2, target)  # warm-up
            t0 = time.perf_counter()
            loss, g, _ = s.adjoint_loop(h, w0A, w0B, w["iters"], target)
            dt = time.perf_counter() - t0
            key = keys[prec]
            out[f"{key}_s_per_gradient"] = dt
            out[f"{key}_loss"] = float(loss[0])
            grads[key] = g[0].copy()
            if prec == pkg.FP64:
                eps = 1e-5
                for k, v in enumerate(dirs):
                    lp = s.adjoint_loop(h + eps * v, w0A, w0B, w["iters"], target, grad=False)[0][0]
                    lm = s.adjoint_loop(h - eps * v, w0A, w0B, w["iters"], target, grad=False)[0][0]
                    fd.append((lp - lm) / (2 * eps))
                    if (k + 1) % 8 == 0:
                        log(f"[bench] c5 finite differences: {k + 1}/{n_dirs} directions")
    out["fd_directions"] = n_dirs
    out["fd_eps"] = 1e-5
    for key in ("fp64", "fp32", "fp32x"):
        mm = [abs(f - float(grads[key] @ v)) / max(abs(f), 1e-300) for f, v in zip(fd, dirs)]
        out[f"{key}_fd_rel_mismatch"] = max(mm)
        out[f"{key}_fd_rel_mismatch_median"] = float(np.median(mm))
    for key in ("fp32", "fp32x"):
        out[f"{key}_vs_fp64_grad_rel"] = float(np.max(np.abs(grads[key] - grads["fp64"])) / np.max(np.abs(grads["fp64"])))
    out["iterations"] = w["iters"]
    out["anderson"] = {"m": 10, "mix": 0.1, "ridge": 1e-2}
    return out


def pass_roofline(s, prec, reps=50):
    """Per-launch duration (HIP events around `reps` back-to-back launches on
    the solver stream) of every kernel the fused march of this solver's tiling
    runs, each kernel's share of a step, and the dominant kernel's bandwidth.

    Steady-state step of the fused march (all three axes active):
      line form  : one middle y pass + one boundary kernel (x|x and z|z alternate)
      plane forms: one boundary kernel (plane|plane and z|z alternate)
    Algorithmic bytes per launch: a pass reads and writes the complex field
    (4bN); a boundary kernel reads the field, writes z(t) and the opened field
    (6bN) and reads the W factors (2bN when post = pre/N exactly, else 4bN)."""
    b = 4 if prec else 8
    plane = getattr(s, "tuning", ([0, 0, 0, 0], None))[0][3]
    N = s.N
    n = round(N ** (1 / 3))
    post_exact = (N & (N - 1)) == 0
    if plane == 0:
        kernels = {"y": 1.0, "fused_x": 0.5, "fused_z": 0.5}
    else:
        kernels = {{1: "fused_plane", 2: "fused_split", 3: "fused_split3", 4: "fused_split_y",
                    5: "fused_split3_y"}[plane]: 0.5, "fused_z": 0.5}
    times = {k: s.time_pass(CODES[k], reps) for k in kernels}
    share = {k: kernels[k] * times[k] for k in kernels}
    dom = max(share, key=share.get)
    # bytes each launch moves: a pass reads and writes the field (4bN); a
    # boundary kernel reads the field and the W factors and writes z(t) and the
    # opened field (8bN when post = pre/N exactly, else 10bN)
    bpl = {k: (4 * b * N if not k.startswith("fused") else (8 if post_exact else 10) * b * N) for k in kernels}
    # algorithmic bytes of the work one launch does: a boundary kernel advances
    # the march by one contour step (6bN: read z, read W, write z); a middle pass
    # does a third of a step's convolutions (2bN)
    alg = {k: (6 * b * N if k.startswith("fused") else 2 * b * N) for k in kernels}
    achieved = alg[dom] / (times[dom] * 1e-3) / 1e9
    # the contour kernel (once per iteration): reads every slice once (2 b N (Ns + 1)) and writes phiA, phiB
    tc = s.time_pass(CONTOUR_CODE, max(3, reps // 5))
    bc = 2 * b * N * (s.NsA + s.NsB + 1) + 16 * N
    return {"dominant": dom, "kernel": KERNEL_NAMES[dom], "ms_per_launch": times, "step_share_ms": share,
            "bytes_per_launch": bpl, "algorithmic_bytes_per_launch": alg, "achieved": achieved,
            "moved_GBps": bpl[dom] / (times[dom] * 1e-3) / 1e9, "grid": n,
            "contour": {"ms": tc, "bytes": bc, "GBps": bc / (tc * 1e-3) / 1e9, "frac": bc / (tc * 1e-3) / 1e9 / HBM_PEAK_GBPS}}


def calibration_copy(nbytes, reps):
    """Device-to-device copies of a known size (the runtime's blit kernel), so
    PMC byte counters can be scaled to bytes: factor = nbytes / counter."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src = ctypes.c_void_p()
    dst = ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) or \
            hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 0, ctypes.c_size_t(nbytes))
    for _ in range(reps):
        if hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"[bench] calibration copy: {reps} x {nbytes >> 20} MiB device to device")


def copy_peak(pkg, nbytes=1 << 30, reps=10):
    """Attainable HBM bandwidth on this GPU (GB/s, read + write): the
    library's 16-byte nontemporal copy kernel over two 1 GiB buffers
    (scft_copy_bandwidth), and, for comparison, the HIP runtime's
    device-to-device copy of the same size."""
    import ctypes
    kernel = pkg.copy_bandwidth(nbytes, reps)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) or \
            hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipEventCreate(ctypes.byref(ev0)) or hip.hipEventCreate(ctypes.byref(ev1)):
            raise RuntimeError("hipEventCreate failed")
        hip.hipMemset(src, 1, nbytes)
        best = 0.0
        for _ in range(3):
            hip.hipMemcpyAsync(dst, src, nbytes, 3, None)  # warm-up (hipMemcpyDeviceToDevice)
            hip.hipEventRecord(ev0, None)
            for _ in range(reps):
                if hip.hipMemcpyAsync(dst, src, nbytes, 3, None):
                    raise RuntimeError("hipMemcpyAsync failed")
            hip.hipEventRecord(ev1, None)
            hip.hipEventSynchronize(ev1)
            ms = ctypes.c_float()
            hip.hipEventElapsedTime(ctypes.byref(ms), ev0, ev1)
            best = max(best, 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9)
        return kernel, best
    finally:
        hip.hipDeviceSynchronize()
        if ev0:
            hip.hipEventDestroy(ev0)
        if ev1:
            hip.hipEventDestroy(ev1)
        hip.hipFree(src)
        hip.hipFree(dst)


def cpu_baseline(pkg, budget_s=10.0, max_iters=6):
    """The oracle (the C restatement of the reference method, OpenMP) on this
    host: the 64^3 cylinder workload with the host threads this process may
    use (OMP_NUM_THREADS, else every core) and with one thread, and the 128^3
    gyroid workload's time per iteration with the same threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    nproc = os.cpu_count() or 1
    threads = min(nproc, int(os.environ.get("OMP_NUM_THREADS", nproc)))

    def iters(name, nthreads, max_it, budget):
        w = pkg.fields.WORKLOADS[name]
        n = w["n"]
        orc.set_threads(nthreads)
        op = orc.params(n, n, n, w["L"], w["L"], w["L"], w["chiN"], w["fA"], w["ds"])
        kt = orc.ktable(op)
        wA, wB = pkg.fields.workload_fields(name)
        N = n ** 3
        phiA = np.empty(N)
        phiB = np.empty(N)
        anderson = orc.Anderson(N, w["anderson_m"], 0.1, 0.10, ridge=ANDERSON_RIDGE)
        it = 0
        t0 = time.perf_counter()
        while it < max_it:
            anderson.iterate(op, kt, wA, wB, phiA, phiB)
            it += 1
            if time.perf_counter() - t0 > budget:
                break
        return it, time.perf_counter() - t0

    log(f"[bench] cpu baseline: 64^3 with {threads} threads")
    iters("c2_cylinder_64", threads, 1, 0.0)  # warm-up (page-in, thread pool)
    i2, t2 = iters("c2_cylinder_64", threads, max_iters, budget_s)
    log("[bench] cpu baseline: 64^3 with 1 thread")
    i1, t1 = iters("c2_cylinder_64", 1, 1, 0.0)
    log(f"[bench] cpu baseline: 128^3 with {threads} threads")
    i3, t3 = iters("c3_gyroid_128", threads, 1, 0.0)
    orc.set_threads(threads)
    return {"value": i2 / t2, "unit": "iters/s", "cores": threads, "kind": "port",
            "sample": f"{i2} Anderson SCFT iterations (m=10, ridge {ANDERSON_RIDGE}) of the 64^3 cylinder "
                      f"workload (Ns=200, FP64) after 1 warm-up, oracle C restatement with OpenMP on "
                      f"{threads} threads",
            "host_nproc": nproc, "threads_used": threads,
            "threads_note": "OMP_NUM_THREADS (this process's CPU share) if set, else every core",
            "c2_1thread_iters_per_s": i1 / t1, "c2_1thread_sample": f"{i1} iteration",
            "c3_s_per_iter": t3 / i3, "c3_iters_per_s": i3 / t3,
            "c3_sample": f"{i3} Anderson iteration (m=20) of the 128^3 gyroid workload on {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # iterations per timed window (median of 5 windows); 0: windows of at least 1 s
    ap.add_argument("--steps", type=int, default=0)
    # warm-up >= the Anderson history (m = 10 at 64^3): every timed iteration
    # runs with the full history, as a long solve does
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-128", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--c5-dirs", type=int, default=64,
                    help="random directions of the c5 finite-difference check (the c5 spec: 64; about 0.3 s each)")
    ap.add_argument("--profile", choices=["c2_fp32", "c2_fp64", "c2_fp32x", "c3_fp32", "c3_fp64", "c3_fp32x"],
                    help="run only warmup+steps iterations of one workload/form (for rocprofv3)")
    ap.add_argument("--choice", default=None,
                    help="with --profile: tiling 'x,y,z,plane' to apply instead of autotuning "
                         "(a profiler distorts the autotuner's timings)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="with --profile and --choice: only the roofline leg (per-kernel launches, no iterations)")
    ap.add_argument("--fuse", type=int, default=1, choices=[0, 1],
                    help="with --profile: step-boundary kernels on (1, the default) or off")
    ap.add_argument("--ckpt", type=int, default=0, choices=[0, 1],
                    help="with --profile: sqrt(Ns) checkpointing of the slice store off (0, the default) or on")
    ap.add_argument("--calib-copy", action="store_true",
                    help="with --profile: also copy 64 MiB -> 64 MiB three times (PMC byte calibration)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this solver runs one GPU per process and this build measures one GPU only; "
              "independent SCFT instances shard across GPUs with no exchange. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    pkg = load_pkg()
    if a.profile:
        wl, form = a.profile.split("_")
        nm = {"c2": "c2_cylinder_64", "c3": "c3_gyroid_128"}[wl]
        prec = {"fp32": pkg.FP32, "fp32x": pkg.FP32X, "fp64": pkg.FP64}[form]
        choice = [int(c) for c in a.choice.split(",")] if a.choice else None
        if a.roofline_only:
            if choice is None:
                sys.exit("--roofline-only needs --choice")
            p = workload_params(pkg, nm, prec, fuse_steps=a.fuse)
            with pkg.Solver(p) as s:
                s.set_tiling(choice)
                s.tuning = (choice, None)
                wA, wB = pkg.fields.workload_fields(nm)
                s.set_fields(wA, wB)
                s.concentrations()  # W factors of the workload's fields
                print(json.dumps({"profile": a.profile, "tiling": choice, "roofline": pass_roofline(s, prec)}))
            if a.calib_copy:
                calibration_copy(64 << 20, 3)
            return
        s, t, fe, _ = run_form(pkg, nm, prec, max(1, a.steps), a.warmup, choice, long_window=False,
                               fuse_steps=a.fuse, checkpoint=a.ckpt)
        ms = t["ms_per_iter"]
        # the roofline leg times single kernels over the full slice store
        rl = pass_roofline(s, prec) if not a.ckpt else None
        print(json.dumps({"profile": a.profile, "tiling": s.tuning[0], "checkpoint": a.ckpt, "ms_per_iter": ms,
                          "iters_per_s": 1e3 / ms, "F": fe[0], "roofline": rl}))
        s.close()
        if a.calib_copy:
            calibration_copy(64 << 20, 3)
        return
    peak_meas, peak_runtime = copy_peak(pkg)
    log(f"[bench] measured copy peak {peak_meas:.0f} GB/s (runtime copy {peak_runtime:.0f}, nominal {HBM_PEAK_GBPS:.0f})")

    def form_block(name, ckpt=0, roof=True):
        """the three forms of one workload: rates (headline windows and >= 1 s
        windows, median of 5), ratios, iteration HBM fractions against the
        nominal and the measured peak, per-kernel roofline legs"""
        res = {}
        for prec, key in ((pkg.FP32, "fp32"), (pkg.FP64, "fp64"), (pkg.FP32X, "fp32x")):
            sv, t, fe, phi = run_form(pkg, name, prec, a.steps, a.warmup, checkpoint=ckpt)
            rl = pass_roofline(sv, prec, reps=50 if sv.N <= 1 << 18 else 20) if roof else None
            res[key] = (sv, t, fe, phi, rl)
        sv = res["fp32"][0]
        N, Ns = sv.N, sv.NsA + sv.NsB
        blk = {"checkpoint": ckpt, "grid": round(N ** (1 / 3)), "Ns": Ns, "windows": WINDOWS}
        for key, (sv, t, fe, phi, rl) in res.items():
            b = 8 if key == "fp64" else 4
            it, itl = 1e3 / t["ms_per_iter"], 1e3 / t["ms_per_iter_long"]
            blk[f"{key}_iters_per_s"] = it
            blk[f"{key}_iters_per_s_long"] = itl
            blk[f"{key}_k"] = t["k"]
            blk[f"{key}_k_long"] = t["k_long"]
            blk[f"{key}_windows_long_ms"] = t["windows_long_ms"]
            blk[f"{key}_iter_hbm_frac"] = iter_bytes(N, Ns, b) * itl / 1e9 / HBM_PEAK_GBPS
            blk[f"{key}_iter_frac_of_copy_peak"] = iter_bytes(N, Ns, b) * itl / 1e9 / peak_meas
            blk[f"tiling_{key}"] = sv.tuning[0]
            if rl is not None:
                blk[f"kernels_{key}"] = rl
        i64 = blk["fp64_iters_per_s_long"]
        blk["ratio_fp32_over_fp64"] = blk["fp32_iters_per_s_long"] / i64
        blk["ratio_fp32x_over_fp64"] = blk["fp32x_iters_per_s_long"] / i64
        f64, p64 = res["fp64"][2], res["fp64"][3]
        for key in ("fp32", "fp32x"):
            blk[f"trajectory_dF_{key}_vs_fp64"] = abs(res[key][2][0] - f64[0])
            blk[f"trajectory_max_dphiA_{key}_vs_fp64"] = float(np.max(np.abs(res[key][3] - p64)))
        blk["trajectory_iterations"] = res["fp64"][1]["warmup_run"] + WINDOWS * (
            res["fp64"][1]["k"] + (res["fp64"][1]["k_long"] if res["fp64"][1]["k_long"] != res["fp64"][1]["k"] else 0))
        _, _, wa64, wb64 = res["fp64"][0].get_fields()
        choices = (res["fp64"][0].tuning[0], res["fp32"][0].tuning[0], res["fp32x"][0].tuning[0])
        head = res["fp32"][1]
        rl32 = res["fp32"][4]
        for key in res:
            res[key][0].close()
        return blk, head, rl32, (wa64[0], wb64[0]), choices

    name = "c2_cylinder_64"
    c2, head, rl32, fin64, ch2 = form_block(name)
    dom = rl32["dominant"]
    forms = {"c2_64": c2}
    log("[bench] c2 one evaluation of the forms on the FP64 run's final fields and on the seeded fields")
    c2["single_eval_fp32_vs_fp64"] = single_eval_error(pkg, name, *fin64)
    c2["single_eval_fp32x_vs_fp64"] = single_eval_error(pkg, name, *fin64, pkg.FP32X)
    c2["single_eval_seeded_fields"] = {
        "fp32": single_eval_error(pkg, name, *pkg.fields.workload_fields(name)),
        "fp32x": single_eval_error(pkg, name, *pkg.fields.workload_fields(name), pkg.FP32X)}
    log("[bench] c2 iterate with and without a convergence tolerance")
    c2["tol_check_rate"] = tol_overhead(pkg, name, pkg.FP32, ch2[1])
    log("[bench] c2 500 iterations of the three forms from the seeded fields")
    c2["after_500_iterations"] = trajectory_error(pkg, name, 500, ch2)
    if not a.no_128:
        name3 = "c3_gyroid_128"
        # the 128^3 workload keeps q(s) with sqrt(Ns) checkpointing: that mode is its headline
        log("[bench] c3 with sqrt(Ns) checkpointing (the workload's storage mode)")
        c3, _, _, _, _ = form_block(name3, ckpt=1, roof=False)
        log("[bench] c3 200-iteration solve: FP32 then FP32X against FP64, sqrt(Ns) checkpointing")
        c3["precision_schedule"] = precision_schedule(
            pkg, name3, 200, 150, (c3["tiling_fp64"], c3["tiling_fp32"], c3["tiling_fp32x"]), 1)
        log("[bench] c3 with the full slice store")
        c3f, _, _, fin3, ch3 = form_block(name3, ckpt=0)
        c3["full_store"] = c3f
        forms["c3_128"] = c3
        log("[bench] c3 one evaluation of the forms on the FP64 run's final fields and on the seeded fields")
        c3["single_eval_fp32_vs_fp64"] = single_eval_error(pkg, name3, *fin3)
        c3["single_eval_fp32x_vs_fp64"] = single_eval_error(pkg, name3, *fin3, pkg.FP32X)
        c3["single_eval_seeded_fields"] = {
            "fp32": single_eval_error(pkg, name3, *pkg.fields.workload_fields(name3)),
            "fp32x": single_eval_error(pkg, name3, *pkg.fields.workload_fields(name3), pkg.FP32X)}
        log("[bench] c3 200 iterations of the three forms from the seeded fields")
        c3["convergence"] = convergence(pkg, name3, 200, ch3)
    if not a.no_sweep:
        inst = pkg.fields.sweep_instances(count=256, seed=4)
        pts = sum(n ** 3 for (n, L, c, f) in inst)
        forms["c4_sweep"] = {"instances": len(inst),
                             "grids": {str(g): sum(1 for x in inst if x[0] == g) for g in (32, 48, 64)}}
        for prec, key in ((pkg.FP64, "fp64"), (pkg.FP32, "fp32"), (pkg.FP32X, "fp32x")):
            log(f"[bench] c4 sweep: 256 instances, grids 32/48/64, {key}")
            base = pkg.default_params(ds=0.01, precision=prec)
            sw = pkg.Sweep(base, inst)
            sw.autotune(5)
            for i, (wa, wb) in enumerate(pkg.fields.sweep_fields(inst, noise=0.1, seed=400)):
                sw.set_fields(i, wa, wb)
            sw.iterate(1)
            k = max(3, int(np.ceil(MIN_WINDOW_S * 1e3 / (sw.time_iterations(2) / 2))))
            ms = float(np.median(windows_ms(sw.time_iterations, k))) / k
            nsum = sum(s_[1] + s_[2] for s_ in sw.sizes)
            b = 8 if prec == pkg.FP64 else 4
            forms["c4_sweep"][key] = {
                "ms_per_sweep_iteration": ms, "k": k, "instance_iters_per_s": len(inst) * 1e3 / ms,
                "grid_point_steps_per_s": pts * (nsum / len(inst)) * 1e3 / ms,
                "iter_hbm_frac": sum(iter_bytes(n ** 3, s_[1] + s_[2], b) for (n, L, c, f), s_ in zip(inst, sw.sizes))
                * 1e3 / ms / 1e9 / HBM_PEAK_GBPS,
            }
            sw.close()
    if not a.no_adjoint:
        log("[bench] c5 adjoint: 64^3, 50 unrolled Anderson iterations, sqrt(Ns) checkpointing")
        forms["c5_adjoint"] = c5_adjoint(pkg, a.c5_dirs)
    cpu = None
    if not a.no_cpu:
        log("[bench] cpu baseline (oracle, host cores)")
        cpu = cpu_baseline(pkg)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        traffic = tr.get(f"c2_fp32:{dom}")
    it32 = 1e3 / head["ms_per_iter"]
    out = {
        "metric": "SCFT iters/s at 64^3,128^3 Ns=200 (FP32 vs FP64 form), %HBM peak, dF vs FP64",
        "value": it32,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": head["k"],
        "warmup": a.warmup,
        "windows": WINDOWS,
        "timing": "median of the windows of `steps` iterations each, HIP events on the solver stream; "
                  f"{head['warmup_run']} warm-up iterations fill the Anderson history first",
        "ms_per_step": head["ms_per_iter"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 propagator, f64 accumulation",
        "data": "synthetic: seeded SplitMix64 noise + hexagonal-star field seed",
        "config": {"workload": "c2 cylinder: AB diblock f=0.3 chiN=20, 64^3, L=4.2 Rg, ds=0.005 "
                               "(NsA=60, NsB=140), Anderson m=10", "grid": [64, 64, 64], "Ns": c2["Ns"],
                   "batch": 1},
        "roofline": {"bound": "hbm", "kernel": rl32["kernel"] + " (fp32, 64^3)", "achieved": rl32["achieved"],
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": rl32["achieved"] / HBM_PEAK_GBPS,
                     "traffic": traffic, "measured_copy_peak": peak_meas,
                     "measured_runtime_copy_peak": peak_runtime,
                     "frac_of_measured_peak": rl32["achieved"] / peak_meas,
                     "algorithmic_bytes_per_launch": rl32["algorithmic_bytes_per_launch"][dom],
                     "bytes_per_launch": rl32["bytes_per_launch"][dom],
                     "ms_per_launch": rl32["ms_per_launch"][dom]},
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
