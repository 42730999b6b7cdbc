#!/usr/bin/env python3
"""Throughput of the fused ZSC step on one MI355X.

Workload: a 512^3 periodic box with 128 bubbles (R ~ U[8,16], seed 2) and the
ZSC physics of lbm_bench_gpu_amd.configs -- FP32 populations for the headline
value, plus the same box with bf16 and fp16 deviation storage and with FP32
storage of g as deviations (f32d).  A step is one
reference time step (update, collide, stream), which is one launch of the
fused kernel.  Inputs are resident in HBM before the timed region.

Prints one JSON line.  `value` is MLUPS = cells * steps / wall time of the
timed steps.  `roofline` prices the fused kernel at its algorithmic bytes per
cell (26 populations read + written once) over its average launch duration
from HIP events on the launch stream.  `cpu_baseline` times the FP64 oracle
(the CPU restatement used as the parity checker) on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=512, help="cube edge (default: 512)")
    ap.add_argument("--variants", default="bf16,fp16,f32d",
                    help="extra storage precisions timed beside fp32")
    ap.add_argument("--ty", type=int, default=0, help="fused tile rows (0: library default)")
    ap.add_argument("--zchunk", type=int, default=-1, help="fused z chunk (-1: library default)")
    ap.add_argument("--flags", type=int, default=-1, help="fused flags (-1: library default)")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="steps of the FP32/bf16/fp16 vs FP64 accuracy runs (0 disables)")
    ap.add_argument("--adjoint-steps", type=int, default=400,
                    help="steps of the 128^3 adjoint timing (0 disables)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU work of the oracle baseline (0 disables)")
    ap.add_argument("--placement-trials", type=int, default=6,
                    help="lattice allocations tried per timed context, the fastest kept "
                         "(zsc_place_lattices; 0 disables)")
    ap.add_argument("--workloads", default="1,3,dropin,share",
                    help="extra workloads reported beside the headline: 1 = 256^3 FP64+FP32 "
                         "2000 steps, 3 = 1024^3 bf16 500 steps, dropin = the reference's "
                         "7-kernel step through the drop-in entry points at 256^3 FP32, "
                         "share = one GPU's z-slab of the 1024^3 box (1024x1024x128) at "
                         "FP32 and bf16 (empty disables)")
    ap.add_argument("--placement-study", type=int, default=0, metavar="K",
                    help="instead of the benchmark: K FP32 contexts of the box held at once "
                         "(K lattice allocations); per allocation the linear and z-march copy "
                         "rates between its two lattices and the fused step's time; one JSON line")
    return ap.parse_args()


# what each precision stores (compute is FP32 in all but f64)
STORAGE = {
    "f32": "f, g fp32 (208 B/cell)",
    "bf16": "g_i - w_i rho bf16, f_1..6 fp16, f_0 fp32 (mixed 16-bit storage, 108 B/cell)",
    "fp16": "g_i - w_i rho fp16, f_1..6 fp16, f_0 fp32 (108 B/cell)",
    "f32d": "f fp32, g_i - w_i rho fp32 (208 B/cell)",
}


def configure(ctx, ty, zchunk, flags):
    if ty > 0 or zchunk >= 0 or flags >= 0:
        ctx.set_fused_config(ty if ty > 0 else 4, zchunk if zchunk >= 0 else 16,
                             flags if flags >= 0 else 1)


def time_precision(prec, w, params, steps, warmup, ty, zchunk, flags, trials=0):
    """One context: `trials` lattice placements (zsc_place_lattices; 0 for
    none), the state, `warmup` untimed steps, then `steps` timed ones."""
    import torch
    from lbm_bench_gpu_amd import zsc
    nx, ny, nz = w.dims
    ctx = zsc.Context(nx, ny, nz, precision=prec)
    placement, placement_s = None, 0.0
    try:
        configure(ctx, ty, zchunk, flags)
        ctx.set_params(params)
        if trials > 0:
            tp = time.perf_counter()
            placement = [round(m, 4) for m in ctx.place_lattices(trials)]
            placement_s = time.perf_counter() - tp
        # read + write ceiling of the pages the timed lattices sit on
        ceiling = ctx.copy_ceiling()
        ctx.state_init_bubbles(w.bubbles)
        ctx.step(warmup + 1)  # the first step also runs the A-lattice prologue
        ctx.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev_ms = ctx.step_timed(steps)  # hipEvents on the launch stream around the launches
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        st = ctx.stats()
    finally:
        ctx.close()
    cells = nx * ny * nz
    bpc = zsc.bytes_per_cell(prec)
    per_launch = ev_ms / 1e3 / steps
    achieved = bpc * cells / per_launch / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None, "bytes_per_cell": bpc}
    if placement:
        # the step's rate at each tried allocation (4 steps each); the timed
        # run uses the fastest, and the median says what one allocation gives
        fr = sorted(bpc * cells / (m * 1e-3) / 1e9 / HBM_PEAK_GBS for m in placement if m > 0)
        roof["frac_placement_median"] = round(float(np.median(fr)), 4)
        roof["frac_placement_range"] = [round(fr[0], 4), round(fr[-1], 4)]
    for k in ("linear", "march"):
        gbs = ceiling[k]["GB/s"]
        roof[f"copy_{k}_GBs"] = gbs
        if gbs:
            roof[f"frac_of_copy_{k}"] = round(achieved / gbs, 4)
    return {
        "mlups": cells * steps / wall / 1e6,
        "ms_per_step": wall * 1e3 / steps,
        "kernel_ms": per_launch * 1e3,
        "roofline": roof,
        "sum_phi": st["sum_phi"],
        "finite": bool(np.isfinite(st["sum_phi"])),
        "placement_ms": placement,
        "placement_s": round(placement_s, 3),
    }


def accuracy(w, params, steps, precs, ty, zchunk, flags):
    """max and RMS error of phi, rho, |u| after `steps` steps of each precision
    against an FP64 run of the same box on the same GPU, the drift of total
    phi (FP64 device reductions) and the ratio of the phase volume (cells with
    phi >= 0) after and before, which the FP64 run gives for comparison."""
    from lbm_bench_gpu_amd import zsc

    def run(prec):
        ctx = zsc.Context(*w.dims, precision=prec)
        try:
            if prec != "f64":
                configure(ctx, ty, zchunk, flags)
            ctx.set_params(params)
            ctx.state_init_bubbles(w.bubbles)
            t0 = ctx.stats()
            ctx.step(steps)
            t1 = ctx.stats()
            phi, rho, ux, uy, uz = ctx.moments()
        finally:
            ctx.close()
        umag = np.sqrt(ux * ux + uy * uy + uz * uz)
        del ux, uy, uz
        return (phi, rho, umag, t0["sum_phi"], t1["sum_phi"],
                t1["bubble_volume_full"] / t0["bubble_volume_full"])

    ref = run("f64")
    out = {"steps": steps, "f64_sum_phi_drift": ref[4] - ref[3],
           "f64_phase_volume_ratio": ref[5]}
    for prec in precs:
        print(f"# accuracy {prec}", file=sys.stderr, flush=True)
        r = run(prec)
        e = {}
        for name, a, b in zip(("phi", "rho", "u"), r[:3], ref[:3]):
            d = a - b
            e[f"{name}_max"] = float(np.abs(d).max())
            e[f"{name}_rms"] = float(np.sqrt(np.mean(d * d)))
            del d
        e["sum_phi_drift"] = r[4] - r[3]
        e["sum_phi_drift_rel"] = (r[4] - r[3]) / abs(r[3])
        e["phase_volume_ratio"] = r[5]
        out[prec] = e
        del r
    return out


def adjoint_timing(steps: int, fwd_mlups_fp32: float):
    """MLUPS of the adjoint on the 128^3, 8-bubble workload (FP32): cells *
    steps / time of the backward pass with checkpoints every 20 steps (the
    workload as specified; the recomputation of every segment included,
    running on a second hardware queue beside the previous segment's
    adjoint steps), and of the adjoint steps alone (HIP events around each
    segment's backward steps; with the overlap these intervals and those of
    the recomputation run at the same time), after one untimed backward
    pass; beside it the same with the recomputation in line
    (`checkpoints_serial`, where the two kinds of interval add up to the
    pass) and with every step's state kept in HBM (`tape`: nothing to
    recompute; 90 GB at 400 steps).  Each mode is timed over three backward
    passes after the warm-up one; the median pass is reported.  All modes
    give the same gradients."""
    import torch
    from lbm_bench_gpu_amd import adjoint, configs, zsc
    w = configs.workload(4)
    p = w.params()
    phi0 = torch.from_numpy(zsc.init_phi_host(w.dims, w.bubbles, p.phi_star, p.width))
    theta = adjoint.theta_of(p)
    tt = [theta[0] * 1.2, theta[1] * 0.9, theta[2], theta[3], theta[4] * 1.5]
    # forward MLUPS of the fused FP32 step on the same box
    ctx = zsc.Context(*w.dims, precision="f32")
    ctx.set_params(p)
    ctx.state_init_bubbles(w.bubbles)
    ctx.step(5)
    ms = ctx.step_timed(steps)
    ctx.close()
    fused = w.cells * steps / (ms / 1e3) / 1e6
    out = {"workload": w.name, "steps": steps, "precision": "f32",
           "fused_forward_mlups": round(fused, 1)}
    grads = {}
    modes = (("tape", {"tape": "auto"}), ("checkpoints", {"tape": False}),
             ("checkpoints_serial", {"tape": False, "overlap": False}),
             # FP32 adjoint lattices over FP64 states: the FP64 forward's
             # gradients (1e-7 of central differences) at FP32 adjoint storage
             ("checkpoints_f32_over_f64_states", {"tape": False, "states": "f64"}),
             # every lattice FP32, g of the states stored as deviations from
             # w_i rho_mean: gradients within 1e-4 of central differences at
             # the FP32 modes' cost
             ("checkpoints_f32_over_f32d_states", {"tape": False, "states": "f32d"}))
    for mode, kw in modes:
        run = adjoint.AdjointRunner(w.dims, p, precision="f32", checkpoint_every=20, timing=True,
                                    **kw)
        tgt = run.forward(tt, phi0, steps, keep=False)[:7].sum(0).clone()
        t0 = time.perf_counter()
        a = run.forward(theta, phi0, steps)
        t1 = time.perf_counter()
        L, lam = run.loss(a, tgt)
        run.backward(lam.clone())  # warm-up (kernel loading, first-touch buffers)
        # three timed passes; the median one is reported
        passes = []
        for _ in range(3):
            t2 = time.perf_counter()
            grads[mode] = run.backward(lam.clone())
            t3 = time.perf_counter()
            passes.append((t3 - t2, run.last_timing))
        passes.sort(key=lambda x: x[0])
        dt, tm = passes[1]
        adj = w.cells * steps / dt / 1e6
        adj_steps = w.cells * steps / (tm["adjoint_steps_ms"] / 1e3) / 1e6
        r = {"states": kw.get("states", "f32"), "adjoint_lattices": "f32",
             "checkpoint_every": 1 if run.tape is not None else 20,
             "tape_GB": round(run.tape_bytes(steps) / 1e9, 1) if run.tape is not None else 0,
             "forward_mlups": round(w.cells * steps / (t1 - t0) / 1e6, 1),
             "adjoint_mlups": round(adj, 1), "adjoint_vs_forward": round(adj / fused, 3),
             "adjoint_step_mlups": round(adj_steps, 1),
             "adjoint_step_vs_forward": round(adj_steps / fused, 3),
             "recompute_ms": round(tm["recompute_ms"], 2),
             "adjoint_steps_ms": round(tm["adjoint_steps_ms"], 2),
             "recompute_overlapped": tm["overlapped"]}
        if mode == "checkpoints":
            # the workload as specified: checkpoints every 20 steps, recomputation
            out.update({"loss": L, "grad": dict(zip(adjoint.THETA, grads[mode])),
                        "adjoint_mlups": r["adjoint_mlups"],
                        "adjoint_vs_forward": r["adjoint_vs_forward"]})
        if mode == "checkpoints_serial":
            # the adjoint steps with nothing beside them
            out.update({"adjoint_step_mlups": r["adjoint_step_mlups"],
                        "adjoint_step_vs_forward": r["adjoint_step_vs_forward"]})
        if "states" in kw:
            r["grad"] = dict(zip(adjoint.THETA, grads[mode]))
        out[mode] = r
        run.close()
        del run, a, lam
        torch.cuda.empty_cache()
    out["grads_equal"] = grads["tape"] == grads["checkpoints"] == grads["checkpoints_serial"]
    return out


TRAFFIC_FILE = "profiles/k_fused_{}x{}x{}_traffic.txt"


def traffic_from_profile(dims):
    """HBM bytes per fused launch per precision as counted by rocprofv3
    (FETCH_SIZE + WRITE_SIZE of this box, scaled by a copy calibration), read
    from the committed summary: a profiler cannot run inside this process,
    so the figure is not measured in this run (empty when absent)."""
    path = os.path.join(ROOT, TRAFFIC_FILE.format(*dims))
    if not os.path.exists(path):
        return {}
    out = {}
    for line in open(path):
        f = line.split()
        if len(f) == 2 and not line.startswith("#"):
            out[f[0]] = float(f[1])
    return out


def placement_study(k: int, w, params, steps: int):
    """The step's dependence on the allocation behind its lattices: k FP32
    contexts held at once (k lattice pairs on different physical pages), and
    for each the read + write rates of a linear and of a z-march copy between
    its two lattices (zsc_copy_ceiling) and the fused step's device time on
    the same pair."""
    from lbm_bench_gpu_amd import zsc
    ctxs, out = [], []
    cells = w.cells
    try:
        for i in range(k):
            print(f"# allocation {i}", file=sys.stderr, flush=True)
            ctx = zsc.Context(*w.dims, precision="f32")
            ctxs.append(ctx)
            ctx.set_params(params)
            c = ctx.copy_ceiling()
            ctx.state_init_bubbles(w.bubbles)
            ctx.step(2)
            ms = ctx.step_timed(steps) / steps
            fused = zsc.bytes_per_cell("f32") * cells / (ms * 1e-3) / 1e9
            out.append({"copy_linear_GBs": c["linear"]["GB/s"], "copy_march_GBs": c["march"]["GB/s"],
                        "fused_ms": round(ms, 4), "fused_GBs": round(fused, 1),
                        "frac": round(fused / HBM_PEAK_GBS, 4),
                        "fused_over_march": round(fused / c["march"]["GB/s"], 4)})
    finally:
        for ctx in ctxs:
            ctx.close()
    return {"workload": w.name, "precision": "f32", "steps_timed": steps, "allocations": out}


def cpu_model() -> str:
    """The host CPU's model name (what lscpu prints), from /proc/cpuinfo."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def workload_1(steps: int):
    """256^3, 16 bubbles: FP64 and FP32 fused steps (MLUPS, roofline) and the
    FP32 field error against the FP64 run after `steps` steps."""
    from lbm_bench_gpu_amd import configs, zsc
    w = configs.workload(1)
    p = w.params()
    out = {"workload": w.name, "steps": steps}
    fields = {}
    for prec in ("f64", "f32"):
        ctx = zsc.Context(*w.dims, precision=prec)
        try:
            ctx.set_params(p)
            ctx.state_init_bubbles(w.bubbles)
            ctx.step(1)
            ms = ctx.step_timed(steps - 1)
            s = ctx.stats()
            fields[prec] = ctx.moments()
        finally:
            ctx.close()
        per = ms / 1e3 / (steps - 1)
        achieved = zsc.bytes_per_cell(prec) * w.cells / per / 1e9
        out[prec] = {"mlups": round(w.cells / per / 1e6, 1), "kernel_ms": round(per * 1e3, 4),
                     "roofline_frac": round(achieved / HBM_PEAK_GBS, 4),
                     "sum_phi": s["sum_phi"]}
    for name, a, b in zip(("phi", "rho", "ux", "uy", "uz"), fields["f32"], fields["f64"]):
        d = a - b
        out["f32"][f"{name}_max_err"] = float(np.abs(d).max())
        out["f32"][f"{name}_rms_err"] = float(np.sqrt(np.mean(d * d)))
    return out


def dropin_sequence(steps: int):
    """The reference's per-step launch sequence (update_phi, update_rho,
    update_velocity, collision_f, collision_g, stream_f, stream_g;
    src/main.cu:498-566) through the drop-in entry points, FP32, on the 256^3
    box: MLUPS over the wall time of `steps` steps, beside the fused step."""
    import torch
    from lbm_bench_gpu_amd import configs, zsc
    w = configs.workload(1)
    p = w.params()
    n = w.cells
    dev = "cuda"
    f = [torch.zeros(2 * n, dtype=torch.float32, device=dev) for _ in range(7)]
    g = [torch.zeros(2 * n, dtype=torch.float32, device=dev) for _ in range(19)]
    phi, rho, ux, uy, uz = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(5)]
    phi.copy_(torch.from_numpy(
        zsc.init_phi_host(w.dims, w.bubbles, p.phi_star, p.width).ravel()).float())
    rho.fill_(p.rho_mean)
    torch.cuda.synchronize()
    ctx = zsc.Context(*w.dims, precision="f32")
    try:
        ctx.set_params(p)
        ctx.init_f(phi, f)
        ctx.init_g(phi, rho, g)

        def step():
            ctx.update_phi(phi, f)
            ctx.update_rho(rho, g)
            ctx.update_velocity(rho, ux, uy, uz, g)
            ctx.collision_f(phi, rho, ux, uy, uz, f)
            ctx.collision_g(phi, rho, ux, uy, uz, g)
            ctx.stream_f(f)
            ctx.stream_g(g)
        for _ in range(3):
            step()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        ctx.sync()
        dt = time.perf_counter() - t0
    finally:
        ctx.close()
    return {"workload": w.name, "steps": steps, "precision": "f32",
            "kernels_per_step": 7, "mlups": round(n * steps / dt / 1e6, 1),
            "ms_per_step": round(dt * 1e3 / steps, 4)}


def workload_3(steps: int):
    """1024^3, 512 bubbles, bf16 storage: MLUPS, GB/s and conservation of
    the phase volume and of total phi over `steps` steps."""
    from lbm_bench_gpu_amd import configs, zsc
    w = configs.workload(3)
    p = w.params()
    ctx = zsc.Context(*w.dims, precision="bf16")
    try:
        ctx.set_params(p)
        ctx.state_init_bubbles(w.bubbles)
        s0 = ctx.stats()
        ctx.step(1)
        ms = ctx.step_timed(steps - 1)
        s1 = ctx.stats()
    finally:
        ctx.close()
    per = ms / 1e3 / (steps - 1)
    gbs = zsc.bytes_per_cell("bf16") * w.cells / per / 1e9
    return {"workload": w.name, "steps": steps, "precision": "bf16 storage, fp32 compute",
            "mlups": round(w.cells / per / 1e6, 1), "kernel_ms": round(per * 1e3, 4),
            "GB/s": round(gbs, 1), "roofline_frac": round(gbs / HBM_PEAK_GBS, 4),
            "phase_volume_ratio": s1["bubble_volume_full"] / s0["bubble_volume_full"],
            "sum_phi_drift_rel": (s1["sum_phi"] - s0["sum_phi"]) / abs(s0["sum_phi"]),
            "sum_rho_drift_rel": (s1["sum_rho"] - s0["sum_rho"]) / abs(s0["sum_rho"])}


def workload_share(steps: int, ty, zchunk, flags, trials):
    """One GPU's share of the 1024^3 box split into eight z-slabs (1024 x
    1024 x 128 and the bubbles inside it), FP32 and bf16 storage: MLUPS and
    roofline fraction of the step at that shape (periodic in z; the halo
    exchange that would join the slabs is not run)."""
    from lbm_bench_gpu_amd import configs
    w = configs.z_slab(3, 8)
    p = w.params()
    out = {"workload": w.name, "dims": list(w.dims), "bubbles": int(len(w.bubbles)),
           "steps": steps}
    for prec in ("f32", "bf16"):
        r = time_precision(prec, w, p, steps, 5, ty, zchunk, flags, trials)
        out[prec] = {"mlups": round(r["mlups"], 1), "kernel_ms": round(r["kernel_ms"], 4),
                     "roofline_frac": r["roofline"]["frac"], "finite": r["finite"],
                     "placement_ms": r["placement_ms"]}
    return out


def cpu_baseline(seconds: float, w):
    """FP64 oracle (C, OpenMP on all host cores) on a bounded sample of the
    timed workload: a 32-plane z-slab of its initial state (nx x ny x 32,
    periodic), stepped for about `seconds`.  MLUPS = slab cells * steps / time."""
    import oracle as O
    from lbm_bench_gpu_amd import zsc
    nx, ny, nz = w.dims
    nzs = min(nz, 32)
    p = w.params()
    c = O.ZoCoef()
    for k, _ in O.ZoCoef._fields_:
        setattr(c, k, p.values[k])
    # the slab's planes of the workload's initial phi (planes 0..nzs-1)
    b = np.array(w.bubbles, dtype=np.float64)
    phi0 = zsc.init_phi_host((nx, ny, nzs), b, p.phi_star, p.width)
    st = O.State((nzs, ny, nx), c).init(phi0)
    t0 = time.perf_counter()
    st.run(1)
    one = time.perf_counter() - t0
    k = max(1, int(seconds / max(one, 1e-3)))
    t0 = time.perf_counter()
    st.run(k)
    dt = time.perf_counter() - t0
    cells = nx * ny * nzs
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    del st
    # the same restatement evaluated in FP32 (the reference's precision), a
    # third of the time on the same slab
    s32 = O.State((nzs, ny, nx), c, dtype=np.float32).init(phi0)
    s32.run(1)
    k32 = max(1, k // 3)
    t0 = time.perf_counter()
    s32.run(k32)
    dt32 = time.perf_counter() - t0
    del s32
    return {"value": round(cells * k / dt / 1e6, 3), "unit": "MLUPS", "cores": cores,
            "cpu_model": cpu_model(), "kind": "port",
            "sample": f"FP64 oracle (C, OpenMP, {cores} threads) on a {nx}x{ny}x{nzs} z-slab of "
                      f"the {w.name} initial state, {k} steps in {dt:.1f} s",
            "fp32": {"value": round(cells * k32 / dt32 / 1e6, 3), "unit": "MLUPS",
                     "sample": f"the oracle's FP32 build (the reference's precision), same slab, "
                               f"{k32} steps in {dt32:.1f} s"},
            "workloads": cpu_side_by_side()}


def cpu_side_by_side():
    """The 64^3 parity box run in full (1000 FP64 steps) by the oracle on the
    host and by the fused step on the GPU, with their largest difference, and
    the 256^3 box for 20 oracle steps."""
    import oracle as O
    from lbm_bench_gpu_amd import configs, zsc
    out = {}
    w0 = configs.workload(0)
    p = w0.params()
    c = O.ZoCoef()
    for k, _ in O.ZoCoef._fields_:
        setattr(c, k, p.values[k])
    nx, ny, nz = w0.dims
    phi0 = O.init_phi((nz, ny, nx), w0.bubbles, p.phi_star, p.width)
    st = O.State((nz, ny, nx), c).init(phi0)
    t0 = time.perf_counter()
    st.run(w0.steps)
    dt = time.perf_counter() - t0
    ctx = zsc.Context(*w0.dims, precision="f64")
    try:
        ctx.set_params(p)
        ctx.state_init(phi0)
        ctx.step(1)
        ms = ctx.step_timed(w0.steps - 1)
        f, g = ctx.state_get()
    finally:
        ctx.close()

    def rel(a, b):
        return float(np.abs(a - b).max() / np.abs(b).max())
    out["0"] = {"workload": w0.name, "steps": w0.steps, "precision": "f64",
                "cpu_mlups": round(w0.cells * w0.steps / dt / 1e6, 2),
                "gpu_mlups": round(w0.cells * (w0.steps - 1) / (ms / 1e3) / 1e6, 1),
                "max_rel_diff_f": rel(f, st.f), "max_rel_diff_g": rel(g, st.g)}
    del st
    w1 = configs.workload(1)
    p = w1.params()
    for k, _ in O.ZoCoef._fields_:
        setattr(c, k, p.values[k])
    nx, ny, nz = w1.dims
    phi0 = O.init_phi((nz, ny, nx), w1.bubbles, p.phi_star, p.width)
    st = O.State((nz, ny, nx), c).init(phi0)
    st.run(1)
    t0 = time.perf_counter()
    st.run(20)
    dt = time.perf_counter() - t0
    out["1"] = {"workload": w1.name, "steps": 20, "precision": "f64",
                "cpu_mlups": round(w1.cells * 20 / dt / 1e6, 2)}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "lbm_bench_gpu_amd runs the step on one GPU; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    from lbm_bench_gpu_amd import configs
    from lbm_bench_gpu_amd.configs import Workload, physics
    w = configs.workload(2)
    if a.size != 512:
        from lbm_bench_gpu_amd import zsc
        dims = (a.size,) * 3
        w = Workload(f"zsc-{a.size}^3", dims, physics(*dims),
                     zsc.generate_bubbles(2, max(1, a.size ** 3 // 1048576), 8, 16, 4.0, 4.0, dims),
                     a.steps, ("f32",))
    params = w.params()
    if a.placement_study > 0:
        print(json.dumps(placement_study(a.placement_study, w, params, min(a.steps, 20))))
        return
    print(f"# workload {w.name}: fp32 {a.steps} steps", file=sys.stderr, flush=True)
    main_r = time_precision("f32", w, params, a.steps, a.warmup, a.ty, a.zchunk, a.flags,
                            a.placement_trials)
    extra = {}
    variants = [s for s in a.variants.split(",") if s]
    for v in variants:
        print(f"# {v} storage", file=sys.stderr, flush=True)
        extra[v] = time_precision(v, w, params, a.steps, a.warmup, a.ty, a.zchunk, a.flags,
                                  a.placement_trials)
    acc = None
    if a.accuracy_steps > 0:
        acc = accuracy(w, params, a.accuracy_steps, ["f32"] + variants, a.ty, a.zchunk, a.flags)
    adj = None
    if a.adjoint_steps > 0:
        print("# adjoint", file=sys.stderr, flush=True)
        adj = adjoint_timing(a.adjoint_steps, main_r["mlups"])
    extra_w = {}
    for wid in [s for s in a.workloads.split(",") if s]:
        print(f"# workload {wid}", file=sys.stderr, flush=True)
        if wid == "1":
            extra_w[wid] = workload_1(2000)
        elif wid == "3":
            extra_w[wid] = workload_3(500)
        elif wid == "dropin":
            extra_w[wid] = dropin_sequence(200)
        elif wid == "share":
            extra_w[wid] = workload_share(200, a.ty, a.zchunk, a.flags, a.placement_trials)
        else:
            raise SystemExit(f"no extra workload {wid}")
    cpu = cpu_baseline(a.cpu_seconds, w) if a.cpu_seconds > 0 else None
    traffic = traffic_from_profile(w.dims)
    for k, r in [("f32", main_r)] + list(extra.items()):
        r["roofline"]["traffic"] = traffic.get(k)
        if traffic.get(k) is not None:
            r["roofline"]["traffic_source"] = (TRAFFIC_FILE.format(*w.dims) + ": rocprofv3 --pmc "
                                               "FETCH_SIZE, WRITE_SIZE of this command, not this run")
    out = {
        "metric": "MLUPS + %HBM roofline, FP32 vs bf16-store (err vs FP64); adjoint MLUPS",
        "value": round(main_r["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(main_r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "cells": w.cells, "bubbles": int(len(w.bubbles)),
                   "precision": "fp32 compute; " + STORAGE["f32"], "gravity": "on"},
        "roofline": main_r["roofline"],
        "kernel_ms": round(main_r["kernel_ms"], 4),
        "placement_ms": main_r["placement_ms"],
        "placement_s": main_r["placement_s"],
        "variants": {k: {"storage": STORAGE.get(k, k),
                         "mlups": round(v["mlups"], 1), "ms_per_step": round(v["ms_per_step"], 4),
                         "speedup_vs_fp32": round(v["mlups"] / main_r["mlups"], 3),
                         "roofline": v["roofline"], "placement_ms": v["placement_ms"]}
                     for k, v in extra.items()},
        "accuracy_vs_fp64": acc,
        "adjoint": adj,
        "workloads": extra_w,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
