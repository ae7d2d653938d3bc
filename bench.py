"""SCFT benchmark: iterations/s of the AB-diblock SCFT solver on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-128]

A "step" is one SCFT iteration (mean subtraction, the packed propagator march
over Ns = 200 contour steps, Q, the trapezoid contour integral, the Anderson
field update and the free energy) of the cylinder workload: AB diblock f=0.3,
chiN=20, 64^3 periodic grid, box 4.2 Rg, ds=0.005, Anderson history 10, seeded
hexagonal-star fields.  `value` is the FP32-propagator form (FP64 Q, phi, F,
fields and least squares); the FP64 form, the FP32X form (FP32 storage and
HBM bytes, FP64 transform arithmetic), the 128^3 gyroid workload and the
errors of both FP32-storage forms against FP64 are reported beside it.
Fields are resident in HBM before the timed region; timing uses HIP events
on the solver's stream.
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


def run_form(pkg, name, prec, steps, warmup, choice=None, **kw):
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
    log(f"[bench] {name} {form} warmup {warmup}")
    s.iterate(warmup)
    log(f"[bench] {name} timing {steps} iterations")
    ms = s.time_iterations(steps)
    fe = s.free_energy()
    phiA, _, _, _ = s.get_fields()
    return s, ms / steps, fe[0].copy(), phiA[0].copy()


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
    bpl = {k: (4 * b * N if not k.startswith("fused") else (8 if post_exact else 10) * b * N) for k in kernels}
    achieved = bpl[dom] / (times[dom] * 1e-3) / 1e9
    # the contour kernel (once per iteration): reads every slice once (2 b N (Ns + 1)) and writes phiA, phiB
    tc = s.time_pass(CONTOUR_CODE, max(3, reps // 5))
    bc = 2 * b * N * (s.NsA + s.NsB + 1) + 16 * N
    return {"dominant": dom, "kernel": KERNEL_NAMES[dom], "ms_per_launch": times, "step_share_ms": share,
            "bytes_per_launch": bpl, "achieved": achieved, "grid": n,
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


def cpu_baseline(pkg, name, budget_s=12.0, max_iters=4):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    w = pkg.fields.WORKLOADS[name]
    n = w["n"]
    cores = min(16, os.cpu_count() or 1)
    orc.set_threads(cores)
    op = orc.params(n, n, n, w["L"], w["L"], w["L"], w["chiN"], w["fA"], w["ds"])
    kt = orc.ktable(op)
    wA, wB = pkg.fields.workload_fields(name)
    N = n ** 3
    phiA = np.empty(N)
    phiB = np.empty(N)
    anderson = orc.Anderson(N, w["anderson_m"], 0.1, 0.10)
    anderson.iterate(op, kt, wA, wB, phiA, phiB)  # warm-up
    it = 0
    t0 = time.perf_counter()
    while it < max_iters:
        anderson.iterate(op, kt, wA, wB, phiA, phiB)
        it += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": it / dt, "unit": "iters/s", "cores": orc.get_threads(), "kind": "port",
            "sample": f"{it} Anderson SCFT iterations of {name} (64^3, Ns=200, FP64) after 1 warm-up, "
                      f"oracle C restatement with OpenMP"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
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
        s, ms, fe, _ = run_form(pkg, nm, prec, a.steps, a.warmup, choice, fuse_steps=a.fuse, checkpoint=a.ckpt)
        # the roofline leg times single kernels over the full slice store
        rl = pass_roofline(s, prec) if not a.ckpt else None
        print(json.dumps({"profile": a.profile, "tiling": s.tuning[0], "checkpoint": a.ckpt, "ms_per_iter": ms,
                          "iters_per_s": 1e3 / ms, "F": fe[0], "roofline": rl}))
        s.close()
        if a.calib_copy:
            calibration_copy(64 << 20, 3)
        return
    name = "c2_cylinder_64"
    s32, ms32, fe32, phi32 = run_form(pkg, name, pkg.FP32, a.steps, a.warmup)
    s64, ms64, fe64, phi64 = run_form(pkg, name, pkg.FP64, a.steps, a.warmup)
    sx, msx, fex, phix = run_form(pkg, name, pkg.FP32X, a.steps, a.warmup)
    Ns = s32.NsA + s32.NsB
    rl32 = pass_roofline(s32, 1)
    rl64 = pass_roofline(s64, 0)
    rlx = pass_roofline(sx, 2)
    dom = rl32["dominant"]
    N = s32.N
    it32 = 1e3 / ms32
    it64 = 1e3 / ms64
    forms = {
        "c2_64": {
            "fp32_iters_per_s": it32, "fp64_iters_per_s": it64, "ratio_fp32_over_fp64": it32 / it64,
            "fp32_iter_hbm_frac": iter_bytes(N, Ns, 4) * it32 / 1e9 / HBM_PEAK_GBPS,
            "fp64_iter_hbm_frac": iter_bytes(N, Ns, 8) * it64 / 1e9 / HBM_PEAK_GBPS,
            "trajectory_dF_fp32_vs_fp64": abs(fe32[0] - fe64[0]),
            "trajectory_max_dphiA_fp32_vs_fp64": float(np.max(np.abs(phi32 - phi64))),
            "F_fp64": fe64[0], "iterations_compared": a.warmup + a.steps,
            "kernels_fp32": rl32, "kernels_fp64": rl64,
            "fp32x_iters_per_s": 1e3 / msx, "ratio_fp32x_over_fp64": ms64 / msx,
            "fp32x_iter_hbm_frac": iter_bytes(N, Ns, 4) * (1e3 / msx) / 1e9 / HBM_PEAK_GBPS,
            "trajectory_dF_fp32x_vs_fp64": abs(fex[0] - fe64[0]),
            "trajectory_max_dphiA_fp32x_vs_fp64": float(np.max(np.abs(phix - phi64))),
            "kernels_fp32x": rlx,
        }
    }
    _, _, wa64, wb64 = s64.get_fields()
    s32.close()
    s64.close()
    sx.close()
    forms["c2_64"]["single_eval_fp32_vs_fp64"] = single_eval_error(pkg, name, wa64[0], wb64[0])
    forms["c2_64"]["single_eval_fp32x_vs_fp64"] = single_eval_error(pkg, name, wa64[0], wb64[0], pkg.FP32X)
    forms["c2_64"]["single_eval_seeded_fields"] = {
        "fp32": single_eval_error(pkg, name, *pkg.fields.workload_fields(name)),
        "fp32x": single_eval_error(pkg, name, *pkg.fields.workload_fields(name), pkg.FP32X)}
    if not a.no_128:
        name3 = "c3_gyroid_128"
        # 128^3 warm-up: the Anderson history (m = 20) fills before the timed iterations
        w3 = pkg.fields.WORKLOADS[name3]["anderson_m"]
        t32, m32, f32, p32 = run_form(pkg, name3, pkg.FP32, max(5, a.steps // 2), w3)
        t64, m64, f64, p64 = run_form(pkg, name3, pkg.FP64, max(5, a.steps // 2), w3)
        tx, mx, fx, px = run_form(pkg, name3, pkg.FP32X, max(5, a.steps // 2), w3)
        rl3_32 = pass_roofline(t32, 1, reps=20)
        rl3_64 = pass_roofline(t64, 0, reps=20)
        rl3_x = pass_roofline(tx, 2, reps=20)
        _, _, wa64, wb64 = t64.get_fields()
        N3 = t32.N
        Ns3 = t32.NsA + t32.NsB
        forms["c3_128"] = {
            "fp32_iters_per_s": 1e3 / m32, "fp64_iters_per_s": 1e3 / m64, "ratio_fp32_over_fp64": m64 / m32,
            "fp32_iter_hbm_frac": iter_bytes(N3, Ns3, 4) * (1e3 / m32) / 1e9 / HBM_PEAK_GBPS,
            "fp64_iter_hbm_frac": iter_bytes(N3, Ns3, 8) * (1e3 / m64) / 1e9 / HBM_PEAK_GBPS,
            "trajectory_dF_fp32_vs_fp64": abs(f32[0] - f64[0]),
            "trajectory_max_dphiA_fp32_vs_fp64": float(np.max(np.abs(p32 - p64))),
            "kernels_fp32": rl3_32, "kernels_fp64": rl3_64,
            "tiling_fp32": t32.tuning[0], "tiling_fp64": t64.tuning[0],
            "fp32x_iters_per_s": 1e3 / mx, "ratio_fp32x_over_fp64": m64 / mx,
            "fp32x_iter_hbm_frac": iter_bytes(N3, Ns3, 4) * (1e3 / mx) / 1e9 / HBM_PEAK_GBPS,
            "trajectory_dF_fp32x_vs_fp64": abs(fx[0] - f64[0]),
            "trajectory_max_dphiA_fp32x_vs_fp64": float(np.max(np.abs(px - p64))),
            "kernels_fp32x": rl3_x, "tiling_fp32x": tx.tuning[0],
        }
        t32.close()
        t64.close()
        tx.close()
        log("[bench] c3 one evaluation of the forms on the FP64 run's final fields and on the seeded fields")
        forms["c3_128"]["single_eval_fp32_vs_fp64"] = single_eval_error(pkg, name3, wa64[0], wb64[0])
        forms["c3_128"]["single_eval_fp32x_vs_fp64"] = single_eval_error(pkg, name3, wa64[0], wb64[0], pkg.FP32X)
        forms["c3_128"]["single_eval_seeded_fields"] = {
            "fp32": single_eval_error(pkg, name3, *pkg.fields.workload_fields(name3)),
            "fp32x": single_eval_error(pkg, name3, *pkg.fields.workload_fields(name3), pkg.FP32X)}
        log("[bench] c3 200 iterations of the three forms from the seeded fields")
        forms["c3_128"]["convergence"] = convergence(pkg, name3, 200, (t64.tuning[0], t32.tuning[0], tx.tuning[0]))
        # the 128^3 workload keeps q(s) with sqrt(Ns) checkpointing: the three forms in that mode
        ck = {}
        for prec, key in ((pkg.FP64, "fp64"), (pkg.FP32, "fp32"), (pkg.FP32X, "fp32x")):
            log(f"[bench] c3 {key} with sqrt(Ns) checkpointing")
            tc, mc, fc, _ = run_form(pkg, name3, prec, 4, w3, checkpoint=1)
            ck[key] = (mc, fc[0])
            forms["c3_128"][f"{key}_ckpt_iters_per_s"] = 1e3 / mc
            tc.close()
        forms["c3_128"]["ratio_fp32_over_fp64_ckpt"] = ck["fp64"][0] / ck["fp32"][0]
        forms["c3_128"]["ratio_fp32x_over_fp64_ckpt"] = ck["fp64"][0] / ck["fp32x"][0]
        forms["c3_128"]["ckpt_trajectory_dF_fp32_vs_fp64"] = abs(ck["fp32"][1] - ck["fp64"][1])
        forms["c3_128"]["ckpt_trajectory_dF_fp32x_vs_fp64"] = abs(ck["fp32x"][1] - ck["fp64"][1])
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
            ms = sw.time_iterations(3) / 3
            nsum = sum(s_[1] + s_[2] for s_ in sw.sizes)
            b = 8 if prec == pkg.FP64 else 4
            forms["c4_sweep"][key] = {
                "ms_per_sweep_iteration": ms, "instance_iters_per_s": len(inst) * 1e3 / ms,
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
        cpu = cpu_baseline(pkg, name)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        traffic = tr.get(f"c2_fp32:{dom}")
    out = {
        "metric": "SCFT iters/s at 64^3,128^3 Ns=200 (FP32 vs FP64 form), %HBM peak, dF vs FP64",
        "value": it32,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 propagator, f64 accumulation",
        "data": "synthetic: seeded SplitMix64 noise + hexagonal-star field seed",
        "config": {"workload": "c2 cylinder: AB diblock f=0.3 chiN=20, 64^3, L=4.2 Rg, ds=0.005 "
                               "(NsA=60, NsB=140), Anderson m=10", "grid": [64, 64, 64], "Ns": Ns,
                   "batch": 1},
        "tiling": {"c2_fp32": s32.tuning[0], "c2_fp64": s64.tuning[0], "c2_fp32x": sx.tuning[0]},
        "roofline": {"bound": "hbm", "kernel": rl32["kernel"] + " (fp32, 64^3)", "achieved": rl32["achieved"],
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": rl32["achieved"] / HBM_PEAK_GBPS,
                     "traffic": traffic, "bytes_per_launch": rl32["bytes_per_launch"][dom],
                     "ms_per_launch": rl32["ms_per_launch"][dom]},
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
