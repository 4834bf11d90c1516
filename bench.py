#!/usr/bin/env python3
"""Throughput of the MI355X VVM timestep: grid-point-steps/s at 2048x256.

One step = one pass of vvm::Iteration::TimeMarching's loop body (advection,
buoyancy, forcing, elliptic solve for w and u, Kessler, turbulence, nudging)
over the config-2 grid (2048 x 256 interior, dx = dz = 100 m, dt = 1 s,
Gaussian theta noise, 100 W/m^2 surface heat flux, qv surface ramp), state
resident in HBM.  The solve stops at ||b - A x|| / ||b|| <= 1e-10.

Three paths are timed:
  fp64_plain     -- FP64 MG-PCG from a zero guess (the reference's guess);
  mixed          -- FP32 MG-PCG inside FP64 refinement from the linear
                    extrapolation of the last two solutions;
  mixed_quadratic -- the same solver from the quadratic extrapolation of
                    the last three solutions;
  mixed_learned  -- the same solver from the conv-net correction of the
                    quadratic extrapolation, whose weights come from --net or
                    are trained here first (untimed) on samples the model
                    records from 1024x128 runs; a guess that starts worse
                    than the extrapolation is dropped, and the network then
                    sits out 1, 4, 16, then 64 solves.
`value` is the mixed_learned path's throughput; all four are printed.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "2d-vector-vorticity-model_amd")]

import vvm_amd as V  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per interior point and launch: distinct 2-D arrays
# read + written (FP64), 1-D profiles and halo traffic excluded.
KERNEL_BYTES_PER_POINT = {
    "diffusion": 120.0,  # Mparam (RKM read, zeta read+write) 24 + Hparam x4 (RKH read, field read+write) 96
    "thermo4": 160.0,    # 4 now + u, w + 5 old AB2 slots read; 4 fut + 5 new slots written
    "pzeta": 96.0,       # zeta, u, w, th, qv, qc, qr + 2 old slots read; zetap + 2 new slots written
    "kessler": 96.0,     # th, qv, qc, qr read and written; 4 process diagnostics written
    "scale": 64.0,       # 4 fields read and written
    "turb": 40.0,        # u, w, thp read; RKM, RKH written
    "mean": 16.0,        # thp, qvp read
    "calu": 24.0,        # w, zetap read; u written
}


# Algorithmic bytes of a whole step (the same rule): the non-solver kernels
# above plus the solve, per PCG iteration and unknown ~220 B in FP64 (CG
# vectors 88 B + a V(2,2) cycle ~4/3 x 100 B) and ~110 B in FP32, plus
# ~100 B per unknown for each FP64 refinement (residual, conversion,
# correction).
STEP_NONSOLVER_BYTES_PER_POINT = sum(KERNEL_BYTES_PER_POINT.values()) + 64.0  # the second scale pass
SOLVER_BYTES_PER_ITER = {"f64": 220.0, "f32": 110.0}
REFINE_BYTES = 100.0


def step_roofline(ms_per_step: float, iters: float, refine: float, fp32: bool, pts: int, unknowns: int) -> dict:
    """The whole step as a fraction of HBM peak, from algorithmic bytes."""
    per_iter = SOLVER_BYTES_PER_ITER["f32" if fp32 else "f64"]
    b = STEP_NONSOLVER_BYTES_PER_POINT * pts + (iters * per_iter + refine * REFINE_BYTES) * unknowns
    gbs = b / (ms_per_step * 1e-3) / 1e9
    return {"bytes_per_step": b, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}


# kernels behind each timed phase (rows of the committed traffic table)
KERNEL_NAMES = {
    "diffusion": ["k_diff_coeffs", "k_diff_strips"], "thermo4": ["k_thermo4"], "pzeta": ["k_pzeta"],
    "kessler": ["k_kessler"], "scale": ["k_scale_forcing"], "turb": ["k_turb_coeffs"], "mean": ["k_update_mean"],
    "calu": ["k_calu_top", "k_calu_cols"],
}
TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "hbm_traffic_2048x256.md")


def measured_traffic(kernel: str) -> float | None:
    """HBM bytes per launch of a timed phase: the sum over its kernels of the
    calibrated FETCH_SIZE + WRITE_SIZE figures in the committed rocprofv3
    table (None if the table lacks a kernel)."""
    if not os.path.exists(TRAFFIC_TABLE):
        return None
    rows, col = {}, None
    for line in open(TRAFFIC_TABLE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if col is None:  # the header names the column of scaled HBM bytes
            col = next((n for n, c in enumerate(cells) if c.startswith("HBM bytes")), None)
            continue
        if len(cells) > col and cells[0].startswith("`"):
            try:
                rows[cells[0].strip("`").split("<")[0]] = float(cells[col])
            except ValueError:
                pass
    names = KERNEL_NAMES.get(kernel, [])
    if not names or any(n not in rows for n in names):
        return None
    return sum(rows[n] for n in names)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_path(case: V.Case, state: V.HostState, steps: int, warmup: int, kernel: str | None, net=None):
    with V.Device(case.cfg) as dev:
        dev.upload(state)
        if net is not None:
            dev.set_guess_net(net)
        if warmup:
            dev.step(warmup)
        if kernel:
            dev.kernel_timer(V.KT[kernel])
        t0 = time.perf_counter()
        stats = dev.step(steps)
        wall = time.perf_counter() - t0
        kavg, kn = dev.kernel_time() if kernel else (0.0, 0)
        final = state.copy()
        dev.download(final)
    return wall, stats, kavg, kn, final


def guess_network(path: str | None, train_steps: int):
    """The learned guess for the mixed_learned path: loaded from a file, or
    trained here on the GPU (untimed) on samples the model records from its
    own runs (1024x128, two seeds of 150 steps), evaluated on a held-out seed."""
    net = V.Net()
    if path:
        net.load(path)
        return net, {"source": path}
    net.init_params(1)
    t0 = time.perf_counter()
    for seed in (11, 12):
        c = V.grid_case(1024, 128, seed=seed, stop_mode=V.VVM_STOP_REL, tol_rel=1e-10, max_iter=400,
                        solver_mode=V.VVM_SOLVER_MIXED, guess_mode=V.VVM_GUESS_EXTRAP)
        with V.Device(c.cfg) as dev:
            dev.upload(V.init_state(c))
            dev.collect(150)
            dev.step(150)
            net.add_samples(dev)
    c = V.grid_case(1024, 128, seed=19, stop_mode=V.VVM_STOP_REL, tol_rel=1e-10, max_iter=400,
                    solver_mode=V.VVM_SOLVER_MIXED, guess_mode=V.VVM_GUESS_EXTRAP)
    held = V.Device(c.cfg)
    held.upload(V.init_state(c))
    held.collect(60)
    held.step(60)
    e0 = net.eval(held)
    ts = net.train(steps=train_steps, batch=16, crop=(64, 64), lr=2e-3, lr_final=5e-5, seed=5, log_every=500)
    e1 = net.eval(held)
    held.close()
    info = {"samples": net.pool_size(), "adam_steps": train_steps, "train_seconds": ts.seconds,
            "loss_first": ts.loss_first, "loss_last": ts.loss_last,
            "heldout_sq_error_vs_extrapolation": e1, "untrained": e0,
            "setup_seconds": time.perf_counter() - t0}
    log(f"[bench] guess network: {info}")
    return net, info


def cpu_baseline(case: V.Case, state: V.HostState) -> dict | None:
    """The reference compiled from its own sources (oracle/_ref), one thread,
    on one step of the same workload; the C oracle if that binary is absent."""
    import tempfile
    from oracle import oracle as O
    pts = (case.cfg.nx - 2) * (case.cfg.nz - 2)
    if O.ref_available():
        with tempfile.TemporaryDirectory() as tmp:
            pin, pout = os.path.join(tmp, "in.state"), os.path.join(tmp, "out.state")
            state.save(pin, case.cfg)
            t0 = time.perf_counter()
            O.run_ref("step", 1, pin, pout, timeout=900)
            wall = time.perf_counter() - t0
            r = V.HostState.read_records(pout)
        t_all = float(r["t_all"][0])
        return {"value": pts / t_all, "unit": "grid-point-steps/s", "cores": 1, "kind": "reference",
                "sample": "1 step of the 2048x256 workload, reference model compiled from its sources "
                          "(Eigen BiCGSTAB+ILUT, tol 1e-22), OMP_NUM_THREADS=1; value from its own t_all timer",
                "t_poisson_s": float(r["t_poisson"][0]), "t_all_s": t_all, "process_wall_s": wall}
    n = 3
    st = state.copy()
    t0 = time.perf_counter()
    O.step(case.cfg, st, n)
    wall = time.perf_counter() - t0
    return {"value": pts * n / wall, "unit": "grid-point-steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} steps of the 2048x256 workload, C restatement (direct FFT+Thomas solve), 1 thread"}


def k3_zeta(cfg, seed: int, rms: float = 1e-2) -> np.ndarray:
    """A seeded Gaussian random vorticity field with a k^-3 power spectrum
    (amplitude spectrum k^-1.5, periodic in x), RMS `rms` s^-1 over the
    interior, zero on rows k = 0, 1 and nz-1, periodic x ghosts; [i][k]."""
    nx, nz = cfg.nx, cfg.nz
    ni, nk = nx - 2, nz - 2
    rng = np.random.default_rng(seed)
    spec = np.fft.rfft2(rng.normal(size=(ni, nk)))
    kx = np.fft.fftfreq(ni)[:, None]
    kz = np.fft.rfftfreq(nk)[None, :]
    kk = np.sqrt(kx * kx + kz * kz)
    kk[0, 0] = np.inf
    z = np.fft.irfft2(spec * kk ** -1.5, s=(ni, nk))
    zeta = np.zeros((nx, nz))
    zeta[1:-1, 1:-1] = z
    zeta[:, :2] = 0.
    zeta[:, -1] = 0.
    zeta *= rms / np.sqrt(np.mean(zeta[1:-1, 1:-1] ** 2))
    zeta[0, :] = zeta[nx - 2, :]
    zeta[nx - 1, :] = zeta[1, :]
    return zeta


def calw_rhs(cfg, rhow: np.ndarray, zeta: np.ndarray) -> np.ndarray:
    """cal_w's right-hand side b = rhow^2 (zeta[i+1] - zeta[i]) dx for
    k = 2..nz-2 in the reference's row order (src/PoissonSolver.cpp:444-450)."""
    k = np.arange(2, cfg.nz - 1)
    d = zeta[2:cfg.nx, 2:cfg.nz - 1] - zeta[1:cfg.nx - 1, 2:cfg.nz - 1]
    return np.ascontiguousarray((rhow[k] ** 2 * d * cfg.dx).T).ravel()


def solve_bench(nfields: int = 3) -> dict:
    """The isolated elliptic solve on seeded k^-3 vorticity fields at
    2048x256: iterations and device time to ||b - A x|| / ||b|| <= 1e-10 of
    FP64 MG-PCG and of FP32 MG-PCG in FP64 refinement (both from a zero
    guess: a lone solve has no previous solutions to extrapolate or feed the
    network), the mixed solution's error against the FP64 solution iterated
    until its residual stops falling, and the reference's Eigen BiCGSTAB +
    ILUT iterations at the same 1e-10 (oracle/_ref, one thread) where that
    binary is present."""
    import tempfile
    from oracle import oracle as O
    base = V.case("c2")
    state = V.init_state(base)
    rhow = state["rhow"]
    n = (base.cfg.nx - 2) * (base.cfg.nz - 3)
    rows = []
    for seed in range(1, nfields + 1):
        zeta = k3_zeta(base.cfg, seed)
        b = calw_rhs(base.cfg, rhow, zeta)
        row = {"seed": seed}
        xs = {}
        for name, over in (("fp64_parity", dict(stop_mode=V.VVM_STOP_PARITY)),
                           ("fp64", dict(stop_mode=V.VVM_STOP_REL, tol_rel=1e-10)),
                           ("mixed", dict(stop_mode=V.VVM_STOP_REL, tol_rel=1e-10, solver_mode=V.VVM_SOLVER_MIXED))):
            c = V.case("c2", max_iter=400, **over)
            with V.Device(c.cfg) as dev:
                dev.upload(state)
                x = np.zeros(n)
                dev.poisson(b, x.copy())  # warm-up
                x = np.zeros(n)
                s = dev.poisson(b, x)
            xs[name] = x
            row[name] = {"iters": s.iters, "refine": s.refine, "rel_residual": s.rel_residual,
                         "ms": 1e3 * s.seconds}
        ref = xs["fp64_parity"]
        for name in ("fp64", "mixed"):
            row[name]["rel_err_vs_fp64"] = float(np.linalg.norm(xs[name] - ref) / np.linalg.norm(ref))
        if O.ref_available():
            st = state.copy()
            st["zetap"][:] = zeta
            with tempfile.TemporaryDirectory() as tmp:
                pin, pout = os.path.join(tmp, "in.state"), os.path.join(tmp, "out.state")
                st.save(pin, base.cfg)
                O.run_ref("solve", "1e-10", pin, pout, timeout=900)
                r = V.HostState.read_records(pout)
            row["reference_bicgstab_ilut"] = {"iters": int(r["solve_iterations"][0]),
                                              "spmv_equivalents": 2 * int(r["solve_iterations"][0]),
                                              "rel_residual": float(r["solve_rel_residual"][0]),
                                              "seconds_1_thread": float(r["wall_seconds"][0])}
        log(f"[bench] solve seed {seed}: {row}")
        rows.append(row)
    return {"workload": "isolated cal_w solve, 2048x256, k^-3 Gaussian zeta fields, RMS 1e-2 s^-1, zero guess, "
                        "stop at ||b - Ax||/||b|| <= 1e-10", "fields": rows}


def copy_calibration(nbytes: int = 256 << 20, reps: int = 3) -> None:
    """A device-to-device copy of known size (the runtime's blit kernel), so
    that FETCH_SIZE / WRITE_SIZE counters can be scaled to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] copy calibration: {reps} x {nbytes} bytes device-to-device")


def ensemble_bench(steps: int, net_path: str | None, seed: int = 4) -> dict:
    """Config 4: 64 members, nx in {256, 512, 1024, 2048}, nz in {64, 128,
    256}, dx = dz = 100 m, per-member seed, theta noise sigma U(0.05, 0.2) K,
    surface heat flux U(50, 200) W/m^2; mixed-precision solve (learned guess
    when a trained network is given, else the extrapolated guess), tuned
    launch shapes.  The same ensemble in FP64 (solve to round-off) gives the
    per-member error."""
    rng = np.random.default_rng(seed)
    draws = [(int(rng.choice([256, 512, 1024, 2048])), int(rng.choice([64, 128, 256])),
              float(rng.uniform(0.05, 0.2)), float(rng.uniform(50, 200))) for _ in range(64)]
    guess = V.VVM_GUESS_LEARNED if net_path else V.VVM_GUESS_EXTRAP2

    def cases(**over):
        return [V.grid_case(nx, nz, seed=1000 + k, sigma=sg, heat=hf, **over)
                for k, (nx, nz, sg, hf) in enumerate(draws)]

    fast = cases(stop_mode=V.VVM_STOP_REL, tol_rel=1e-10, solver_mode=V.VVM_SOLVER_MIXED, guess_mode=guess)
    net = None
    if net_path:
        net = V.Net()
        net.load(net_path)
    finals = {}
    out = {}
    for name, cs in (("mixed", fast), ("fp64", cases(stop_mode=V.VVM_STOP_PARITY))):
        with V.Ensemble(cs) as ens:
            if name == "mixed":
                if net is not None:
                    ens.set_guess_net(net)
                ens.tune(2)
            ens.step(2)  # warm-up
            log(f"[bench] ensemble {name}: 64 members, {steps} steps")
            s = ens.step(steps)
            out[name] = {"value": s.grid_point_steps / s.seconds, "seconds": s.seconds,
                         "solve_iters_per_member_step": s.solve_iters_total / (64 * steps),
                         "not_converged": s.not_converged}
            finals[name] = []
            for i, c in enumerate(cs):
                st = V.init_state(c)
                ens.member(i).download(st)
                finals[name].append(st)
    errs = []
    for a, b in zip(finals["mixed"], finals["fp64"]):
        errs.append(max(float(np.linalg.norm(a[f] - b[f]) / max(np.linalg.norm(b[f]), 1e-300))
                        for f in ("zeta", "th", "qv", "u", "w")))
    pts = sum(nx * nz for nx, nz, _, _ in draws)
    return {
        "metric": "ensemble grid-point-steps/s (64 uneven members, mixed precision)",
        "value": out["mixed"]["value"], "unit": "grid-point-steps/s", "n_gpus": 1, "steps": steps,
        "higher_is_better": True, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": "config 4: 64 members, nx in {256,512,1024,2048}, nz in {64,128,256}, "
                               "dx=dz=100 m", "grid_points_total": pts,
                   "guess": "learned" if net_path else "quadratic extrapolation"},
        "fp64_value": out["fp64"]["value"], "paths": out,
        "member_error_max_rel_l2": max(errs), "member_error_median_rel_l2": float(np.median(errs)),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ensemble", action="store_true",
                    help="time the 64-member ensemble of uneven grids instead of the 2048x256 step")
    ap.add_argument("--net", default=None, help="trained guess network (vvm_net_save file) for the mixed path")
    ap.add_argument("--solve", action="store_true",
                    help="also run the isolated elliptic solve on k^-3 vorticity fields (extra 'solve' key)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 256 MiB device-to-device copy (counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--kernel", default="diffusion", choices=sorted(KERNEL_BYTES_PER_POINT))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=30,
                    help="steps of the 64-member ensemble reported beside the 2048x256 paths (0: skip)")
    ap.add_argument("--paths", default="fp64_plain,mixed,mixed_quadratic,mixed_learned")
    ap.add_argument("--train-steps", type=int, default=6000,
                    help="Adam steps of the in-process training of the guess network (untimed)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the timestep is single-device; see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)
    if args.ensemble:
        print(json.dumps(ensemble_bench(args.steps, args.net)), flush=True)
        return

    paths = {
        "fp64_plain": dict(solver_mode=V.VVM_SOLVER_FP64, guess_mode=V.VVM_GUESS_ZERO),
        "mixed": dict(solver_mode=V.VVM_SOLVER_MIXED, guess_mode=V.VVM_GUESS_EXTRAP),
        "mixed_quadratic": dict(solver_mode=V.VVM_SOLVER_MIXED, guess_mode=V.VVM_GUESS_EXTRAP2),
        "mixed_learned": dict(solver_mode=V.VVM_SOLVER_MIXED, guess_mode=V.VVM_GUESS_LEARNED),
    }
    wanted = [p for p in args.paths.split(",") if p]
    if args.calibrate_copy:
        copy_calibration()
    base = V.case("c2")
    state = V.init_state(base)
    pts = (base.cfg.nx - 2) * (base.cfg.nz - 2)
    results, finals = {}, {}
    learned_info = None
    for name in wanted:
        c = V.case("c2", stop_mode=V.VVM_STOP_REL, tol_rel=1e-10, max_iter=400, **paths[name])
        log(f"[bench] path {name}: {args.warmup} warmup + {args.steps} timed steps of 2048x256")
        net = None
        if name == "mixed_learned":
            net, learned_info = guess_network(args.net, args.train_steps)
        wall, stats, kavg, kn, final = run_path(c, state, args.steps, args.warmup, args.kernel, net)
        finals[name] = final
        results[name] = {
            "value": pts * args.steps / wall, "ms_per_step": 1e3 * wall / args.steps,
            "solve_iters_per_step": stats.solve_iters_total / max(args.steps, 1),
            "refine_last": stats.refine_last, "rel_residual_last": stats.rel_residual_last,
            "not_converged": stats.not_converged,
            "guess_fallbacks": stats.guess_fallbacks, "guess_skipped": stats.guess_skipped,
            "phase_ms_last_step": {"advection": 1e3 * stats.t_advection, "poisson": 1e3 * stats.t_poisson,
                                   "microphysics": 1e3 * stats.t_microphysics,
                                   "diffusion": 1e3 * stats.t_diffusion},
            "kernel_avg_us": 1e6 * kavg, "kernel_launches": kn,
        }
        results[name]["step_roofline"] = step_roofline(
            results[name]["ms_per_step"], results[name]["solve_iters_per_step"], float(stats.refine_last),
            paths[name]["solver_mode"] == V.VVM_SOLVER_MIXED, pts, (base.cfg.nx - 2) * (base.cfg.nz - 3))
        log(f"[bench] {name}: {results[name]['value']:.4g} grid-point-steps/s, "
            f"{results[name]['ms_per_step']:.3f} ms/step, {results[name]['solve_iters_per_step']:.2f} iters/step")

    head = next((p for p in ("mixed_learned", "mixed") if p in results), wanted[0])
    R = results[head]
    kbytes = KERNEL_BYTES_PER_POINT[args.kernel] * pts
    achieved = kbytes / (R["kernel_avg_us"] * 1e-6) / 1e9 if R["kernel_avg_us"] > 0 else 0.0
    out = {
        "metric": "grid-point-steps/s at 2048x256, FP64 plain vs mixed+learned; Poisson iters to 1e-10",
        "value": R["value"], "unit": "grid-point-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": R["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": "config 2: 2048x256 interior, dx=dz=100 m, dt=1 s, Kessler, sigma=0.1 K theta "
                               "noise < 500 m, 100 W/m^2 surface heat flux, qv surface ramp; solve to 1e-10",
                   "grid_points": pts, "path": head},
        "paths": results,
        "roofline": {"kernel": args.kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "bytes_per_launch": kbytes, "avg_launch_us": R["kernel_avg_us"],
                     "traffic": measured_traffic(args.kernel), "traffic_source": os.path.relpath(TRAFFIC_TABLE, ROOT)},
    }
    for p in ("mixed", "mixed_quadratic", "mixed_learned"):
        if "fp64_plain" in results and p in results:
            out[f"speedup_{p}_vs_fp64"] = results[p]["value"] / results["fp64_plain"]["value"]
            a, b = finals[p], finals["fp64_plain"]
            out[f"{p}_vs_fp64_rel_l2"] = {f: float(np.linalg.norm(a[f] - b[f]) / max(np.linalg.norm(b[f]), 1e-300))
                                          for f in ("zeta", "th", "qv", "qc", "qr", "u", "w")}
    if "mixed_learned" in results:
        out["learned_guess"] = learned_info
    if args.solve:
        out["solve"] = solve_bench()
    if args.ensemble_steps > 0:
        e = ensemble_bench(args.ensemble_steps, args.net)
        out["ensemble"] = {k: e[k] for k in ("metric", "value", "unit", "steps", "fp64_value",
                                             "member_error_max_rel_l2", "member_error_median_rel_l2")}
        out["ensemble"]["config"] = e["config"]
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (one reference step at 2048x256, single thread)")
        cb = cpu_baseline(base, state)
        if cb:
            out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
