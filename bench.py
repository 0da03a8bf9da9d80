#!/usr/bin/env python3
"""FLIP step benchmark: 256^3 dam break on one MI355X.

Metric: FLIP steps/s with the mixed-precision PCG at relative residual 1e-6,
plus the PCG's HBM bandwidth and the mixed-versus-plain comparison.

Workload (dam break 256^3): fluid block x < 96, y < 160 (all z), solid walls on
all faces, 8 Philox-jittered particles per fluid cell (31,457,280 particles,
SoA float32), gravity 9.8 voxel/s^2 in -y, flip 0.95, cfl 1, 30 frames/s, the
reference's 2x2x2 partitions as 8 tiles of one domain.

A step is one FlipDriver::RunOneStep (advection + sort, clear, P2G, boundary
conditions, gravity, divergence, PCG solve, pressure update, velocity update,
extrapolation, G2P) with all data resident in HBM.

The timed loop is FlipDriver's frame loop on the device (fsl_advance_frame:
dt computed on the device, counts kept there, each step's front half
replayed from a HIP graph), until K steps ran. The same steps are then run
again from the same particles without the graph and as the host-driven
loop (fsl_compute_dt + fsl_step), then K more with the per-kernel timing
hook for the stage and kernel durations.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-compare]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fluid-sim-lb_amd"))
sys.path.insert(0, ROOT)

import fsl  # noqa: E402
from fsl import scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per fluid row (float vectors, 1-byte label; halo re-reads
# not counted). Pass A reads z, p, label and writes p, Ap; pass B reads x, p,
# r, Ap, label and writes x, r, z.
PASS_A_BYTES = {fsl.PCG_FP32: 17, fsl.PCG_MIXED: 17, fsl.PCG_FP64: 33}
PASS_B_BYTES = {fsl.PCG_FP32: 29, fsl.PCG_MIXED: 29, fsl.PCG_FP64: 57}
# Block IC(0) apply per fluid row: forward solve (read r, d, o, label; write
# y), backward solve (read y, d, o, label; write z), and the r.r / r.z pass
# (read r, z, label).
IC_BYTES = {fsl.PCG_FP32: 43, fsl.PCG_MIXED: 43, fsl.PCG_FP64: 83}
# Single-reduction pass (one kernel per iteration): reads r, p, label and
# writes r, p every iteration; reads and writes x every other iteration (the
# two-cell halo re-reads are not counted): 17 + 8/2 = 21 B per row with float
# vectors, 33 + 16/2 = 41 with double vectors.
SR_BYTES = {fsl.PCG_FP32: 21, fsl.PCG_MIXED: 21, fsl.PCG_FP64: 41}
PREC_NAME = {fsl.PCG_FP32: "fp32", fsl.PCG_FP64: "fp64", fsl.PCG_MIXED: "mixed"}
# Algorithmic bytes per particle: P2G reads position and velocity (24 B) and
# writes 12 B of faces + 1 B of label per cell (8 particles per cell); G2P
# reads position and velocity (24 B), writes velocity and grid velocity
# (24 B) and reads v and dv faces (24 B per cell / 8).
P2G_BYTES = 24 + 13 / 8
WORKLOAD_TEXT = {
    "dam_break_256": "fluid x<96 y<160",
    "multi_dam_512": "pool y<32 plus four 96x192x96 corner blocks",
    "sphere_drop_128": "pool y<32 plus sphere r=20 at (64,92,64)",
    "sphere_drop_64": "pool y<16 plus sphere r=10 at (32,46,32)",
    "dam_break_128": "fluid x<48 y<80",
}
G2P_BYTES = 48 + 24 / 8


def sync(ctx):
    """Waits for the context's stream (the library's own HIP stream)."""
    ctx.synchronize()


def setup(dims, tiles, mask, seed, ppc=8):
    ncell = int(mask.sum())
    ctx = fsl.Context(dims, tiles=tiles, capacity=ncell * ppc + 1024)
    ctx.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    ctx.set_box_boundary()
    ctx.seed_particles(mask, ppc, seed)
    return ctx


def pre_solve_stages(ctx, dt):
    """The stages of RunOneStep that precede the solve (flip_app_sim.cc:213-244)."""
    ctx.move_particles(dt)
    ctx.clear()
    ctx.p2g()
    ctx.apply_bc()
    ctx.save_velocities()
    ctx.delete_outside()
    ctx.add_gravity(dt)
    ctx.apply_bc()
    ctx.divergence()


def divergence_left(b, p, marker):
    """||b - A p|| / ||b|| over the FLUID rows, in FP64 from the stored FP32
    pressure: the divergence the pressure update leaves on the FLUID cells
    (div(u + G p) = b - A p there; A: diag = non-SOLID neighbours, -1 per
    FLUID neighbour). Box-format arrays."""
    # Crop to the FLUID cells' bounding box plus one cell (every row and its
    # six neighbours; the box has a cell of margin on every side).
    idx = np.argwhere(marker == 1)
    lo = np.maximum(idx.min(axis=0) - 1, 0)
    hi = np.minimum(idx.max(axis=0) + 2, marker.shape)
    sl = tuple(slice(int(a), int(c)) for a, c in zip(lo, hi))
    marker, b, p = marker[sl], b[sl], p[sl]
    fl = marker == 1
    ns = (marker != 2).astype(np.float64)
    pp = np.where(fl, p.astype(np.float64), 0.0)
    ap = np.zeros_like(pp)
    diag = np.zeros_like(pp)
    for ax in range(3):
        for sh in (-1, 1):
            diag += np.roll(ns, sh, axis=ax)
            ap -= np.roll(pp, sh, axis=ax)
    ap += diag * pp
    r = np.where(fl, b.astype(np.float64) - ap, 0.0)
    return float(np.linalg.norm(r) / np.linalg.norm(np.where(fl, b, 0.0)))


def pcg_compare(ctx, dims):
    """Solve the same system with each precision and iteration structure
    (x0 = 0 every time). "fp32" is the plain solve in the reference's
    precision and structure (FP32 vectors and dots, two passes per iteration).
    The "ic_" solves use the block IC(0) preconditioner (the reference's
    configured but inactive IC branch) in place of Jacobi. Each solve states
    its pressure error against FP64 solved to a relative residual of 1e-12
    and the divergence its pressure leaves; a sweep over the mixed solve's
    relative residual finds where its pressure error reaches 1e-6."""
    out = {}
    pressures = {}
    SR, TP = fsl.PCG_SINGLE_REDUCTION, fsl.PCG_TWO_PASS
    J, IC = fsl.PRECOND_JACOBI, fsl.PRECOND_IC
    runs = [
        ("mixed", fsl.PCG_MIXED, SR, fsl.STOP_REL, 1e-6, J),
        ("fp32", fsl.PCG_FP32, TP, fsl.STOP_REL, 1e-6, J),
        ("fp64", fsl.PCG_FP64, SR, fsl.STOP_REL, 1e-6, J),
        ("mixed_two_pass", fsl.PCG_MIXED, TP, fsl.STOP_REL, 1e-6, J),
        ("fp64_two_pass", fsl.PCG_FP64, TP, fsl.STOP_REL, 1e-6, J),
        ("fp32_reference_rule", fsl.PCG_FP32, TP, fsl.STOP_ABS_RR, None, J),
        ("ic_fp32_reference_rule", fsl.PCG_FP32, TP, fsl.STOP_ABS_RR, None, IC),
        ("ic_mixed", fsl.PCG_MIXED, TP, fsl.STOP_REL, 1e-6, IC),
        ("fp64_tight", fsl.PCG_FP64, SR, fsl.STOP_REL, 1e-12, J),
    ]
    for name, prec, var, stop, tol, pc in runs:
        prm = fsl.solve_params(precision=prec, stop_mode=stop, threshold=tol, dims=dims,
                               max_iterations=20000, variant=var, preconditioner=pc)
        ctx.set_timing(True)
        st = ctx.solve(prm)
        t = ctx.timing()
        ctx.set_timing(False)
        it = max(st["iterations"], 1)
        rows = st["n_rows"]
        rec = {
            "variant": "single_reduction" if var == SR else "two_pass",
            "preconditioner": "ic0" if pc == IC else "jacobi",
            "iterations": st["iterations"],
            "solve_s": st["seconds"],
            "ms_per_iteration": 1e3 * st["seconds"] / it,
            "true_rel_residual": st["true_rel_residual"],
        }
        if var == SR:
            bytes_it = SR_BYTES[prec] * rows
            rec["kernel_ms"] = t["fused"] / max(t["fused_n"], 1)
            rec["kernel_GBps"] = bytes_it / (rec["kernel_ms"] * 1e-3) / 1e9
        else:
            bytes_it = (PASS_A_BYTES[prec] + PASS_B_BYTES[prec]) * rows
            if pc == IC:
                bytes_it += IC_BYTES[prec] * rows
            rec["pass_a_ms"] = t["pass_a"] / max(t["pass_a_n"], 1)
            rec["pass_b_ms"] = t["pass_b"] / max(t["pass_b_n"], 1)
            rec["pass_a_GBps"] = PASS_A_BYTES[prec] * rows / (rec["pass_a_ms"] * 1e-3) / 1e9
            rec["pass_b_GBps"] = PASS_B_BYTES[prec] * rows / (rec["pass_b_ms"] * 1e-3) / 1e9
        rec["algorithmic_bytes_per_row_iteration"] = bytes_it / rows
        # The solve time includes the timing hook's events; GB/s over it is a
        # lower bound of the un-hooked rate.
        rec["pcg_GBps"] = bytes_it * it / st["seconds"] / 1e9 if st["seconds"] > 0 else None
        out[name] = rec
        pressures[name] = ctx.get_scalar(fsl.GRID_PRESSURE).astype(np.float64)
    # Un-hooked solve times (the timing hook's events slow the launches). Each
    # solve runs twice and the second is timed: a single-reduction solve
    # enqueues as many launches as the previous solve on the context took, so
    # the first run after a different solve (e.g. the 1e-12 one) would time
    # launches queued past convergence.
    for name, prec, var, stop, tol, pc in runs:
        prm = fsl.solve_params(precision=prec, stop_mode=stop, threshold=tol, dims=dims,
                               max_iterations=20000, variant=var, preconditioner=pc)
        ctx.solve(prm)
        out[name]["solve_s_unhooked"] = ctx.solve(prm)["seconds"]
    ref = pressures["fp64_tight"]
    den = np.linalg.norm(ref)
    b = ctx.get_scalar(fsl.GRID_DIVERGENCE)
    marker = ctx.get_int(fsl.GRID_MARKER)
    for name, *_ in runs:
        out[name]["divergence_left_rel"] = divergence_left(b, pressures[name], marker)
    for name, *_ in runs[:-1]:
        out[name]["pressure_rel_err_vs_fp64_1e-12"] = float(np.linalg.norm(pressures[name] - ref) / den)
    # The mixed (and, for reference, FP64) single-reduction solves at tighter
    # relative residuals: the pressure error is set by the stopping tolerance
    # and the system's conditioning, the same in both arithmetics.
    sweep = []
    for tol in (1e-6, 5e-7, 2e-7, 1e-7, 5e-8, 2e-8):
        row = {"threshold": tol}
        for name, prec in (("mixed", fsl.PCG_MIXED), ("fp64", fsl.PCG_FP64)):
            prm = fsl.solve_params(precision=prec, stop_mode=fsl.STOP_REL, threshold=tol,
                                   dims=dims, max_iterations=20000)
            ctx.solve(prm)
            st = ctx.solve(prm)  # timed with the launch count predicted right
            pr = ctx.get_scalar(fsl.GRID_PRESSURE).astype(np.float64)
            row[name] = {"iterations": st["iterations"], "solve_s": st["seconds"],
                         "true_rel_residual": st["true_rel_residual"],
                         "pressure_rel_err_vs_fp64_1e-12": float(np.linalg.norm(pr - ref) / den),
                         "divergence_left_rel": divergence_left(b, pr, marker)}
        sweep.append(row)
    out["mixed_tolerance_sweep"] = sweep
    ok = [r for r in sweep if r["mixed"]["pressure_rel_err_vs_fp64_1e-12"] <= 1e-6]
    if ok:
        r = ok[0]
        out["mixed_pressure_1e-6"] = {
            "threshold": r["threshold"], "iterations": r["mixed"]["iterations"],
            "solve_s": r["mixed"]["solve_s"],
            "pressure_rel_err_vs_fp64_1e-12": r["mixed"]["pressure_rel_err_vs_fp64_1e-12"],
            "speedup_vs_fp32_plain_at_1e-6": out["fp32"]["solve_s_unhooked"] / r["mixed"]["solve_s"],
        }
    f64 = pressures["fp64"]
    out["mixed"]["pressure_rel_err_vs_fp64_same_rule"] = float(
        np.linalg.norm(pressures["mixed"] - f64) / np.linalg.norm(f64))
    u = "solve_s_unhooked"
    out["speedup_mixed_vs_fp32_plain"] = out["fp32"][u] / out["mixed"][u]
    out["speedup_mixed_vs_fp64"] = out["fp64"][u] / out["mixed"][u]
    out["speedup_mixed_vs_fp64_two_pass"] = out["fp64_two_pass"][u] / out["mixed"][u]
    out["rows"] = rows
    return out


def _oracle_tile_step(pos, vel, gv, dims, variant, iters):
    """One oracle step on one partition (a 128^3 tile with its own particles),
    single-threaded as the reference runs a partition: (seconds of the step
    with `iters` PCG iterations, non-PCG seconds, assembly seconds, seconds
    per PCG iteration from a bounded 40-iteration solve, P2G seconds)."""
    from oracle.oracle import Oracle

    o = Oracle(dims, variant=variant, threads=1)
    o.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    o.set_box_boundary()
    if pos.shape[0]:
        o.set_particles(pos, vel, gv)
    dt, _ = o.compute_dt(1, 1.0 / 30, 0.0)
    t0 = time.perf_counter()
    o.move_particles(dt); o.clear()
    tp = time.perf_counter()
    o.p2g()
    tp = time.perf_counter() - tp
    o.apply_bc(); o.save_velocities()
    o.delete_outside(); o.add_gravity(dt); o.apply_bc(); o.divergence()
    t1 = time.perf_counter()
    o.solve(precision=0, stop_mode=1, threshold=1e-6, max_iter=0)
    t2 = time.perf_counter()
    _, it, _ = o.solve(precision=0, stop_mode=1, threshold=1e-6, max_iter=40)
    t3 = time.perf_counter()
    o.pressure_update(); o.velocity_update(); o.g2p()
    t4 = time.perf_counter()
    nonpcg, asm = (t1 - t0) + (t4 - t3), t2 - t1
    per_it = ((t3 - t2) - asm) / max(it, 1) if it else 0.0
    return nonpcg + asm + per_it * iters, nonpcg, asm, per_it, tp


def _host_cpu():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": usable, "cpu_count": os.cpu_count(), "model": model}


def cpu_baseline(iters_256, ppc=8):
    """The CPU oracle in the reference's execution model on the 256^3 dam
    break's own partitions: the 2x2x2 MetaGrid (metagrid.cc:25-35) cuts the
    domain into eight 128^3 boxes; each runs as one single-threaded task
    (tbb::task_scheduler_init(1), projects/flip-current-lb/water_app.cc:21),
    the eight concurrently on eight host threads, and the slowest partition
    sets the step. The partitions' fluid: x < 96 and y < 128 in the two
    lower boxes with x < 128 (1,572,864 rows and 12,582,912 particles each),
    x < 96 and 128 <= y < 160 in the two upper ones (393,216 rows, 3,145,728
    particles; run as the bottom 32 layers of a 128^3 tile), none in the
    other four. Each task runs one step's particle and grid stages and a
    bounded 40-iteration PCG; its step time is scaled to the GPU step's
    iteration count. The per-iteration all-reduces between partitions are
    not modelled (which favours the CPU). Built on this host with -O3
    -march=native (no contraction: the checker's bits) and timed again with
    -Ofast -march=native (the reference's flags, makeinclude/flags.mk:6)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as orc_mod

    orc_mod.build_baseline()
    dims = (128, 128, 128)
    tiles = []
    for y_end, copies in ((128, 2), (32, 2), (0, 4)):
        if y_end:
            mask = scenes.dam_break_mask(dims, 96, y_end)
            g = setup(dims, (1, 1, 1), mask, 3, ppc)
            pos, vel, gv = g.get_particles()
            g.close()
        else:
            pos = vel = gv = np.zeros((0, 3), np.float32)
        tiles += [(y_end, pos, vel, gv)] * copies
    out = {}
    for variant in ("o3", "ofast"):
        with ThreadPoolExecutor(len(tiles)) as ex:
            res = list(ex.map(lambda t: _oracle_tile_step(t[1], t[2], t[3], dims, variant,
                                                         iters_256), tiles))
        out[variant] = res
    res = out["o3"]
    step_n = max(r[0] for r in res)
    heavy = res[0]
    host = _host_cpu()
    return {
        "value": 1.0 / step_n,
        "unit": "FLIP steps/s",
        "cores": len(tiles),
        "kind": "port",
        "ofast_value": 1.0 / max(r[0] for r in out["ofast"]),
        "one_thread_value": 1.0 / sum(r[0] for r in res),
        "heavy_partition_ms_per_pcg_iteration": 1e3 * heavy[3],
        "heavy_partition_p2g_particles_per_s": tiles[0][1].shape[0] / heavy[4],
        "host": host,
        "sample": (f"oracle (CPU restatement of the reference step, FP32 Jacobi-PCG), built "
                   f"-O3 -march=native on this host ({host['model']}, {host['nproc']} usable "
                   f"cores): the 256^3 dam break's 8 partitions as 8 single-threaded tasks "
                   f"on {len(tiles)} threads (rows 1572864 x2, 393216 x2, 0 x4), one step "
                   f"each with a 40-iteration PCG sample, scaled to the GPU step's "
                   f"{iters_256} iterations; slowest partition: non-PCG stages "
                   f"{heavy[1]:.2f} s, assembly {heavy[2]:.2f} s, {heavy[3] * 1e3:.1f} ms per "
                   f"iteration; -Ofast build {1.0 / max(r[0] for r in out['ofast']):.4f} "
                   f"steps/s; partition all-reduces not modelled; one thread (the 8 "
                   f"partitions in turn) {1.0 / sum(r[0] for r in res):.4f} steps/s"),
    }


def profile_solve(dims, mask, seed, tiles, ic=False):
    """One mixed-precision solve of the 256^3 dam-break system (after one step
    of particle stages), plus a plain device copy of 1 GiB (read 1 GiB, write
    1 GiB) that calibrates the HBM byte counters of a profiler run. Uses no
    torch kernels (torch's kernels do not run under the profiler here).
    ic: the solve uses the block IC(0) preconditioner (two-pass)."""
    ctx = setup(dims, tiles, mask, seed)
    dt, _ = ctx.compute_dt(1, np.float32(1.0 / 30), 0.0)
    pre_solve_stages(ctx, dt)
    prm = fsl.solve_params(precision=fsl.PCG_MIXED, stop_mode=fsl.STOP_REL, threshold=1e-6,
                           dims=dims, max_iterations=20000,
                           preconditioner=fsl.PRECOND_IC if ic else fsl.PRECOND_JACOBI)
    st = ctx.solve(prm)
    ctx.close()
    # Plain device-to-device copy through the HIP runtime (a blit kernel).
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = 1 << 30
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    copy = {"bytes_read": n, "bytes_written": n}
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        copy = {"error": "hipMalloc failed"}
    else:
        hip.hipMemset(a, 1, ctypes.c_size_t(n))
        hip.hipDeviceSynchronize()
        hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
        hip.hipDeviceSynchronize()
        hip.hipFree(a)
        hip.hipFree(b)
    print(json.dumps({"solve": st, "copy_calibration": copy}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="dam_break_256")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-solve", action="store_true",
                    help="one mixed solve + a 1 GiB copy, for profiler counter runs")
    ap.add_argument("--ic", action="store_true",
                    help="with --profile-solve: the solve uses the block IC(0) preconditioner")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: runs on one GPU only; sharding the tiles over several GPUs is "
              "not implemented", file=sys.stderr)
        sys.exit(2)

    dims, mk, seed = scenes.WORKLOADS[a.workload]
    mask = mk(dims)
    tiles = (2, 2, 2) if dims[0] >= 256 else (1, 1, 1)
    if a.profile_solve:
        profile_solve(dims, mask, seed, tiles, ic=a.ic)
        return
    t_setup = time.perf_counter()
    ctx = setup(dims, tiles, mask, seed)
    n_particles = ctx.num_particles()
    t_setup = time.perf_counter() - t_setup
    prm = fsl.solve_params(precision=fsl.PCG_MIXED, stop_mode=fsl.STOP_REL, threshold=1e-6,
                           dims=dims, max_iterations=20000)
    frame = np.float32(1.0 / 30)
    compare = None
    if not a.no_compare:
        # A fixed system, the same for every command line: the pressure
        # system of the seeded scene's first step. The particles are then
        # seeded again, so the timed steps start from the seeded scene.
        dt, _ = ctx.compute_dt(1, frame, 0.0)
        pre_solve_stages(ctx, dt)
        compare = pcg_compare(ctx, dims)
        compare["system"] = "first step of the seeded scene (before the warm-up and timed steps)"
        ctx.seed_particles(mask, 8, seed)
        ctx.clear()
    step_t = 0.0

    def one_step():
        """The host-driven loop: dt from a host read of the CFL maximum
        (fsl_compute_dt), then fsl_step."""
        nonlocal step_t
        dt, step_t = ctx.compute_dt(1, frame, step_t)
        if step_t + 1e-5 >= frame:
            step_t = 0.0
        return ctx.step(dt, prm)

    def frames(k):
        """FlipDriver's frame loop on the device (fsl_advance_frame, dt
        computed there), until at least k steps ran: (steps, stats of each
        frame's last solve, PCG iterations)."""
        n, out, its = 0, [], []
        while n < k:
            fs = ctx.advance_frame(1, frame, prm)
            n += fs["steps"]
            out.append(fs["last_solve"])
            its.append(fs["pcg_iterations"])
        return n, out, its

    def timed(fn):
        sync(ctx)
        t = time.perf_counter()
        res = fn()
        sync(ctx)
        return res, time.perf_counter() - t

    # The headline: the device-resident frame loop with each step's front
    # half replayed from a HIP graph. Timed region: K steps, no measurement
    # hook (its per-kernel events slow the launches down), bracketed by
    # device synchronisations.
    ctx.set_step_graph(True)
    frames(a.warmup)
    snapshot = ctx.get_particles()
    (n_g, stats, iters), elapsed = timed(lambda: frames(a.steps))
    steps_per_s = n_g / elapsed

    def restart():
        """Back to the particles the timed loop started from: each loop
        below runs the same n_g steps (the same dt, hence the same PCG
        systems) as the headline loop."""
        nonlocal step_t
        ctx.set_particles(*snapshot)
        ctx.clear()
        step_t = 0.0

    # The same loop without the graph, and the host-driven loop.
    ctx.set_step_graph(False)
    restart()
    (n_ng, _, iters_ng), el_ng = timed(lambda: frames(n_g))
    restart()
    host_stats, el_host = timed(lambda: [one_step() for _ in range(n_g)])
    del snapshot
    variants = {
        "device_loop_graph": {"steps": n_g, "steps_per_s": steps_per_s,
                              "ms_per_step": 1e3 * elapsed / n_g,
                              "pcg_iterations_total": int(np.sum(iters))},
        "device_loop_no_graph": {"steps": n_ng, "steps_per_s": n_ng / el_ng,
                                 "ms_per_step": 1e3 * el_ng / n_ng,
                                 "pcg_iterations_total": int(np.sum(iters_ng))},
        "host_loop": {"steps": n_g, "steps_per_s": n_g / el_host, "ms_per_step": 1e3 * el_host / n_g,
                      "pcg_iterations_total": int(sum(s["iterations"] for s in host_stats))},
    }
    # Kernel durations: the next K steps with the hook on, which records each
    # PCG pass's own start and end on the library's stream.
    ctx.set_timing(True)
    sync(ctx)
    t1 = time.perf_counter()
    stats_t = [one_step() for _ in range(a.steps)]
    sync(ctx)
    hooked_ms = 1e3 * (time.perf_counter() - t1) / a.steps
    tm = ctx.timing()
    ctx.set_timing(False)
    iters_per_step = float(np.sum(iters)) / n_g
    iters_t = [s["iterations"] for s in stats_t]
    rows_t = [s["n_rows"] for s in stats_t]
    rows_w = float(np.average(rows_t, weights=np.maximum(iters_t, 1)))
    # Leaf occupancy: FLUID rows over the cells of the leaves holding them.
    occupancy = float(np.mean([s["n_rows"] / max(512 * s["n_fluid_leaves"], 1) for s in stats_t]))
    cg_ms = tm["fused"] / max(tm["fused_n"], 1)
    achieved = SR_BYTES[fsl.PCG_MIXED] * rows_w / (cg_ms * 1e-3) / 1e9
    pcg_time = tm["solve"] * 1e-3
    pcg_gbps = sum(SR_BYTES[fsl.PCG_MIXED] * r * i for r, i in zip(rows_t, iters_t)) / pcg_time / 1e9
    stage_ms = {k: tm[k] / a.steps for k in ("move", "clear", "p2g", "bc_save", "gravity_bc",
                                             "divergence", "solve", "update", "g2p")}
    p2g_rate = n_particles / (stage_ms["p2g"] * 1e-3)
    g2p_rate = n_particles / (stage_ms["g2p"] * 1e-3)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "pcg_cg_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            per_row = json.load(f).get("hbm_bytes_per_row")
        if per_row:
            traffic = per_row * rows_w

    ctx.close()

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(int(round(iters_per_step)))

    line = {
        "metric": ("FLIP steps/s at 256^3 (mixed-precision PCG, rel. residual 1e-6)"
                   if a.workload == "dam_break_256" else
                   f"FLIP steps/s, {a.workload} (mixed-precision PCG, rel. residual 1e-6)"),
        "value": steps_per_s,
        "unit": "FLIP steps/s",
        "n_gpus": 1,
        "steps": n_g,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / n_g,
        "step_loops": variants,
        "ms_per_step_with_timing_hook": hooked_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (stencil, axpys) + f64 (dot products, residual refresh)",
        "data": "synthetic (Philox-seeded dam break)",
        "config": {
            "workload": f"{a.workload}: FLIP step, {WORKLOAD_TEXT.get(a.workload, '')}, 8 ppc, "
                        f"tiles {'x'.join(str(t) for t in tiles)}",
            "dims": list(dims),
            "particles": n_particles,
            "fluid_rows_mean": rows_w,
            "fluid_leaf_occupancy": occupancy,
            "pcg_iterations_per_frame": [int(i) for i in iters],
            "pcg_iterations_per_step_mean": iters_per_step,
            "parallelism": "1 GPU, tiles as work units",
        },
        "pcg_GBps": pcg_gbps,
        "pcg_frac_of_hbm_peak": pcg_gbps / HBM_PEAK_GBS,
        "pcg_iterations_per_s": sum(iters_t) / pcg_time,
        "p2g_particles_per_s": p2g_rate,
        "p2g_frac_of_hbm_peak": p2g_rate * P2G_BYTES / (HBM_PEAK_GBS * 1e9),
        "g2p_particles_per_s": g2p_rate,
        "g2p_frac_of_hbm_peak": g2p_rate * G2P_BYTES / (HBM_PEAK_GBS * 1e9),
        "stage_ms_per_step": stage_ms,
        "roofline": {
            "kernel": "k_cg: one single-reduction PCG iteration (p, Ap, x, r, A M^-1 r from a "
                      "two-cell LDS halo; gamma, delta, r.r)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_row": SR_BYTES[fsl.PCG_MIXED],
            "rows": rows_w,
            "avg_launch_ms": cg_ms,
            "launches": tm["fused_n"],
        },
        "pcg_compare": compare,
        "cpu_baseline": cpu,
        "setup_s": t_setup,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
