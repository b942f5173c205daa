"""Benchmark of the plain relaxing-TVD split MHD step on one MI355X.

Workload (default): one 1024^3 periodic FP32 box, seeded random-phase IC with
velocity amplitude 0.3 and vector-potential amplitude 0.05, CFL 0.7, 20
full steps.  A step is one full x,y,z,z,y,x split step (the reference's 16
operators: calcfl, 6 fluid sweeps, 6 CT field sweeps, 4 transposes; the
first half's two advectbyzx+transpose pairs, the second half's two
transpose+advectbyzx pairs and the middle's advectbyzx, advectbyzx, fluidx
run as one fused kernel each, and calcfl is fused into the last fluid sweep)
replayed from a captured HIP graph; state, t and dt stay in HBM.
`value` is measured in the FP32 tolerance mode (precision fast, checked
against the reference by tests/test_gpu_fast.py); the bitwise-exact mode is
timed beside it ("exact").  Prints one JSON line:

  value      cell-updates/s = N * steps / wall time of the timed steps
  roofline   the dominant kernel (largest time per step, by rocprofv3 kernel
             identity: both transposes are one transpose_kernel): its
             algorithmic bytes per launch / its mean launch duration (HIP
             events on the launch stream) vs 8 TB/s; traffic from the
             committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
  kernels    the same figures for every kernel of the step
  copy_GBs   the library's streaming copy (copy_matrix) on the state: the
             achievable ceiling next to the vendor peak
  step_roofline  792 B/cell (the reference decomposition) per step time
  ensemble   the 96-box batched workload, aggregate cell-updates/s
  corrector  learned-corrector errors and GPU times on held-out seeds
  sizes      the 64^3 and 256^3 workloads (v 0.1, A 0.05): GPU cell-updates/s
             of the plain step, exact and fast
  cpu_baseline  the oracle restatement (bitwise equal to the reference C
             path) timed on this host's cores on a bounded 256^3 sample
             (.small_64: the same oracle on the 64^3 workload, 10 steps,
             serial and on up to 16 threads)

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1024] [--precision fast|exact]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

PEAK_HBM_GBS = 8000.0
# compulsory bytes per cell per call at each operator's C-call boundary (FP32)
ALGO_BYTES = {"fluidx": 52, "advectbyzx": 32, "transpose12": 64, "transpose13": 64, "calcfl": 32,
              "advect_fluidx": 64, "advect2_fluidx": 64, "advect_transpose12": 64,
              "advect_transpose13": 64, "transpose12_advect": 64, "transpose13_advect": 64}


def calls_per_step(n: int) -> dict:
    """Launches per step of the plan on an n^3 box.  The first half's two
    advectbyzx+transpose pairs run fused; the middle (the first half's last
    advectbyzx and the second half's first advectbyzx+fluidx) is one kernel
    (advect2_fluidx).  When the partner axis tiles into 16-pencil blocks (n %
    16 == 0) the second half's two transpose+advectbyzx pairs run fused and
    their fluid sweeps stand alone (the last also reduces calcfl, so its
    launch takes somewhat longer than the plain fluidx timed here); otherwise
    the transposes stand alone and each advectbyzx+fluidx pair is one kernel,
    the last with calcfl."""
    if n % 16 == 0:
        return {"fluidx": 5, "advect_transpose12": 1, "advect_transpose13": 1, "advect2_fluidx": 1,
                "transpose13_advect": 1, "transpose12_advect": 1}
    return {"fluidx": 3, "advect_transpose12": 1, "advect_transpose13": 1, "transpose12": 1,
            "transpose13": 1, "advect_fluidx": 2, "advect2_fluidx": 1}
# the reference decomposition the step roofline is priced on (792 B/cell)
REF_CALLS = {"fluidx": 6, "advectbyzx": 6, "transpose12": 2, "transpose13": 2, "calcfl": 1}
# the kernel (as rocprofv3 names it) each operator launches, and that
# kernel's operator key in the committed traffic summary; one kernel can serve
# several operators (both transposes are one transpose_kernel)
KERNEL_OF = {"fluidx": "fluidx_kernel", "advectbyzx": "advect_kernel",
             "advect_fluidx": "advect_fluid_kernel (1 pass)",
             "advect2_fluidx": "advect_fluid_kernel (2 passes)",
             "transpose12": "transpose_kernel", "transpose13": "transpose_kernel",
             "advect_transpose12": "advect_transpose_kernel (12)",
             "advect_transpose13": "advect_transpose_kernel (13)",
             "transpose12_advect": "transpose_advect_kernel (12)",
             "transpose13_advect": "transpose_advect_kernel (13)", "calcfl": "cfl_kernel"}
TRAFFIC_KEY = {"fluidx_kernel": "fluidx", "advect_kernel": "advectbyzx",
               "advect_fluid_kernel (1 pass)": "advect_fluidx",
               "advect_fluid_kernel (2 passes)": "advect2_fluidx",
               "transpose_kernel": "transpose (u and b)", "cfl_kernel": "calcfl",
               "advect_transpose_kernel (12)": "advect_transpose12",
               "advect_transpose_kernel (13)": "advect_transpose13",
               "transpose_advect_kernel (12)": "transpose12_advect",
               "transpose_advect_kernel (13)": "transpose13_advect"}
PRECISION_INDEPENDENT = {"transpose_kernel", "cfl_kernel"}
METRIC = "cell-updates/s & % HBM roofline per 3D step"


def kernel_times(box, reps: int, precision: str = "exact"):
    """Mean device time (s) of each operator's launch, by HIP events recorded
    on the stream the kernels are launched on."""
    import torch
    from heterogeneous_mhd_amd import ops
    nx, ny, nz = box.nx, box.ny, box.nz
    u, b, u2, b2 = box.u, box.b, box.u2, box.b2
    dt = box.dt_tensor()
    calls = {
        "fluidx": lambda: ops.fluidx(u2, b, nx, ny, nz, dt, u, b, precision=precision),
        "advectbyzx": lambda: ops.advectbyzx(u, b2, nx, ny, nz, dt, u, b, precision=precision),
        "advect_fluidx": lambda: ops.advect_fluidx(u2, b2, nx, ny, nz, dt, u, b, precision=precision),
        "advect2_fluidx": lambda: ops.advect2_fluidx(u2, b2, nx, ny, nz, dt, u, b, precision=precision),
        "transpose12": lambda: ops.transpose12(u2, b2, u, b, nx, ny, nz),
        "transpose13": lambda: ops.transpose13(u2, b2, u, b, nx, ny, nz),
        "advect_transpose12": lambda: ops.advect_transpose12(u2, b2, nx, ny, nz, dt, u, b,
                                                             precision=precision),
        "advect_transpose13": lambda: ops.advect_transpose13(u2, b2, nx, ny, nz, dt, u, b,
                                                             precision=precision),
        "transpose12_advect": lambda: ops.transpose12_advect(u2, b2, nx, ny, nz, dt, u, b,
                                                             precision=precision),
        "transpose13_advect": lambda: ops.transpose13_advect(u2, b2, nx, ny, nz, dt, u, b,
                                                             precision=precision),
        "calcfl": lambda: ops.calcfl_device(u, b, nx, ny, nz),
    }
    if nx % 16:  # the transpose+advectbyzx kernels need 16-pencil blocks
        del calls["transpose12_advect"], calls["transpose13_advect"]
    out = {}
    with torch.cuda.stream(box.stream):
        for name, f in calls.items():
            if name == "advectbyzx":
                b2.copy_(b)  # advectbyzx reads main b(1) from its output buffer
            f()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
            for r in range(reps):
                ev[2 * r].record()
                f()
                ev[2 * r + 1].record()
            torch.cuda.synchronize()
            out[name] = float(np.mean([ev[2 * r].elapsed_time(ev[2 * r + 1]) for r in range(reps)])) / 1e3
    return out


def cpu_baseline(u, b, n: int, budget_s: float = 12.0):
    """Oracle (FP32, bitwise equal to the reference C path) on host cores,
    stepping the given n^3 state (host arrays) for about budget_s seconds."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import Oracle
    threads = min(16, os.cpu_count() or 1)
    o = Oracle("f32", threads=threads)
    u, b, _, _ = o.run(u, b, n, n, n, 1)  # warm-up step
    steps, t0 = 0, time.perf_counter()
    while True:
        u, b, _, _ = o.run(u, b, n, n, n, 1)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 50:
            break
    return {"value": n ** 3 * steps / el, "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle mhd_ref.c FP32 (bitwise equal to the reference C path), OpenMP over "
                      f"pencils on {threads} threads, {n}^3 box of the same IC generator, {steps} "
                      f"full steps after 1 warm-up, {el:.1f} s"}


def pmc_traffic(op: str, cells: int):
    """HBM bytes per launch of one operator's kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the same box size (counters
    calibrated on a copy of known size), or None if there is none."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"hbm_traffic_step_{round(cells ** (1 / 3))}.txt")
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith("json: "):
            d = json.loads(ln[6:])
            if d["cells"] == cells and op in d["ops"]:
                return d["ops"][op]["bytes_per_launch"]
    return None


def kernel_table(kt: dict, cells: int, precision: str, calls: dict):
    """Per kernel (rocprofv3 identity): launches per step, time per step, the
    mean launch, algorithmic bytes per launch and the achieved fraction of
    the HBM peak.  Operators that launch the same kernel are pooled."""
    groups = {}
    for op, t in kt.items():
        c = calls.get(op, 0)
        if not c:
            continue
        g = groups.setdefault(KERNEL_OF[op], {"ops": [], "launches": 0, "s": 0.0, "bytes": 0.0})
        g["ops"].append(op)
        g["launches"] += c
        g["s"] += c * t
        g["bytes"] += c * ALGO_BYTES[op] * cells
    out = {}
    for k, g in groups.items():
        gbs = g["bytes"] / g["s"] / 1e9
        key = TRAFFIC_KEY[k] + ("" if precision == "exact" or k in PRECISION_INDEPENDENT else " fast")
        out[k] = {"ops": g["ops"], "launches_per_step": g["launches"], "ms_per_step": 1e3 * g["s"],
                  "mean_launch_s": g["s"] / g["launches"], "bytes_per_launch": g["bytes"] / g["launches"],
                  "achieved_GBs": gbs, "frac": gbs / PEAK_HBM_GBS, "traffic": pmc_traffic(key, cells)}
    return out


def copy_probe(box, reps: int = 3):
    """Achieved bandwidth of the library's streaming copy (copy_matrix: u and
    b in one launch, 16-byte loads, nontemporal stores) on the state: 64
    bytes per cell moved.  Reported next to the vendor peak."""
    import torch
    from heterogeneous_mhd_amd import ops
    n = box.nx
    with torch.cuda.stream(box.stream):
        ops.copy_matrix(box.u, box.b, n, n, n, box.u2, box.b2)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            ops.copy_matrix(box.u, box.b, n, n, n, box.u2, box.b2)
        e1.record()
        torch.cuda.synchronize()
    s = e0.elapsed_time(e1) / 1e3 / reps
    return 64.0 * box.n / s / 1e9


def ensemble_bench(nbox: int = 96, steps: int = 50, precision: str = "exact"):
    """The batched-ensemble workload: nbox periodic boxes with sides drawn by
    seed from {32,48,64,96,128} and per-box amplitudes, autotuned launch shape,
    graph-replayed steps.  Returns aggregate cell-updates/s and details."""
    import torch
    from heterogeneous_mhd_amd import solver
    sides, specs = solver.ensemble_specs(nbox, seed=7)
    ens = solver.Ensemble(sides, cfl=0.7, precision=precision)
    ens.init(specs)
    best_ms = ens.tune(reps=2)
    ens.run(2, graph=True)
    ens.sync()
    t0 = time.perf_counter()
    ens.run(steps, graph=True)
    ens.sync()
    wall = time.perf_counter() - t0
    t, dt = ens.times()
    out = {"boxes": nbox, "cells": ens.cells, "steps": steps, "value": ens.cells * steps / wall,
           "unit": "cell-updates/s", "ms_per_step": 1e3 * wall / steps, "precision": precision,
           "config": ens.config(),
           "tune_best_ms": best_ms, "dt_range": [min(dt), max(dt)]}
    ens.close()
    torch.cuda.empty_cache()
    return out


def timed_steps(box, steps: int, warmup: int, graph: bool):
    import torch
    box.run(warmup, graph=graph)
    box.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    box.run(steps, graph=graph)
    box.sync()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def size_sweep(sizes=(64, 256), steps=(50, 20)):
    """The smaller workloads (64^3 and 256^3, v amp 0.1, A amp 0.05, CFL 0.7):
    GPU cell-updates/s of the plain step in both precisions, from HIP events
    on the box's stream around graph-replayed steps (state in HBM)."""
    import torch
    from heterogeneous_mhd_amd import solver
    out = {}
    for n, k in zip(sizes, steps):
        row = {}
        for prec in ("exact", "fast"):
            box = solver.Box(n, n, n, cfl=0.7, precision=prec)
            box.init(solver.ICSpec(seed=1, amp_v=0.1, amp_a=0.05))
            box.run(3, graph=True)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(box.stream)
            box.run(k, graph=True)
            e1.record(box.stream)
            e1.synchronize()
            ms = e0.elapsed_time(e1) / k
            row[prec] = {"value": n ** 3 / (ms / 1e3), "ms_per_step": ms}
            box.close()
        out[f"{n}^3"] = dict(row, steps=k)
    return out


def cpu_small(n: int = 64, steps: int = 10):
    """The oracle on the 64^3 workload (v amp 0.1, A amp 0.05, 10 steps),
    serial and on up to 16 host threads."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from heterogeneous_mhd_amd import solver
    from oracle import Oracle
    u0, b0 = solver.ic_host(n, n, n, solver.ICSpec(seed=1, amp_v=0.1, amp_a=0.05))
    out = {}
    for threads in (1, min(16, os.cpu_count() or 1)):
        o = Oracle("f32", threads=threads)
        t0 = time.perf_counter()
        o.run(u0, b0, n, n, n, steps)
        el = time.perf_counter() - t0
        out[f"threads_{threads}"] = {"value": n ** 3 * steps / el, "s": el}
    out["sample"] = f"{n}^3, {steps} full steps from the IC (oracle FP32, bitwise = reference C path)"
    return out


def corrector_bench(seeds: int, steps: int):
    """The learned-corrector workload: held-out seeds, 4x-downsampled 256^3
    runs as the reference solution; errors of plain 64^3, corrected 64^3 and plain 128^3
    at 50 and `steps` coarse steps, with their GPU times."""
    from heterogeneous_mhd_amd import corrector
    if not corrector.WEIGHTS.exists():
        return None
    net = corrector.CorrectorNet.load()
    rows = corrector.evaluate(net, list(range(5000, 5000 + seeds)), steps, checkpoints=(50, steps),
                              log=lambda m: print(m, file=sys.stderr, flush=True))
    out = corrector.summarize(rows)
    out["workload"] = (f"{seeds} held-out seeds, v amp 0.1, A amp 0.05, {steps} coarse steps; "
                       "error = mean over (rho, m, E, B) of ||x - ref|| / ||ref - mean(ref)||")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--amp-v", type=float, default=0.3)
    ap.add_argument("--amp-a", type=float, default=0.05)
    ap.add_argument("--cfl", type=float, default=0.7)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=5)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-exact", action="store_true", help="skip the bitwise-exact leg")
    ap.add_argument("--no-corrector", action="store_true")
    ap.add_argument("--no-sizes", action="store_true", help="skip the 64^3 / 256^3 throughputs")
    ap.add_argument("--corrector-seeds", type=int, default=8)
    ap.add_argument("--corrector-steps", type=int, default=200)
    ap.add_argument("--precision", choices=["exact", "fast"], default="fast")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="print only the 96-box ensemble measurement (for profiling it)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the z-slab "
              "decomposition with all-to-all transposes is designed, not built)", file=sys.stderr)
        sys.exit(2)

    if args.ensemble_only:
        print(json.dumps(ensemble_bench(precision=args.precision)), flush=True)
        return
    import torch
    from heterogeneous_mhd_amd import solver
    n = args.n
    N = n ** 3
    graph = not args.no_graph
    print(f"[bench] {n}^3 box, allocating state", file=sys.stderr, flush=True)
    box = solver.Box(n, n, n, cfl=args.cfl, precision=args.precision)
    box.init(solver.ICSpec(seed=1, amp_v=args.amp_v, amp_a=args.amp_a))
    print(f"[bench] IC ready; {args.warmup} warm-up + {args.steps} timed steps ({args.precision})",
          file=sys.stderr, flush=True)
    wall = timed_steps(box, args.steps, args.warmup, graph)
    t_sim, dt = box.time()
    ms_per_step = 1e3 * wall / args.steps
    value = N * args.steps / wall
    print(f"[bench] {ms_per_step:.2f} ms/step, t={t_sim:.4f} dt={dt:.4g}", file=sys.stderr, flush=True)

    kt = kernel_times(box, args.kernel_reps, args.precision)
    kernels = kernel_table(kt, N, args.precision, calls_per_step(n))
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    dk = kernels[dom]
    step_bytes = sum(ALGO_BYTES[k] * REF_CALLS[k] for k in REF_CALLS) * N

    def step_roofline(w):
        gbs = step_bytes / (w / args.steps) / 1e9
        return {"algo_bytes_per_cell_step": step_bytes / N, "achieved_GBs": gbs,
                "frac": gbs / PEAK_HBM_GBS}

    diag = box.diagnostics()
    copy_gbs = copy_probe(box)
    line = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded random-phase Fourier IC generated on the GPU)",
        "config": {"workload": f"plain relaxing-TVD split MHD step, {n}^3 periodic FP32 box, "
                               f"v amp {args.amp_v}, A amp {args.amp_a}, CFL {args.cfl}",
                   "nx": n, "ny": n, "nz": n, "graph": graph, "precision": args.precision},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": dk["achieved_GBs"], "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": dk["frac"], "traffic": dk["traffic"],
                     "algo_bytes_per_launch": dk["bytes_per_launch"], "mean_launch_s": dk["mean_launch_s"],
                     "launches_per_step": dk["launches_per_step"], "ops": dk["ops"]},
        "step_roofline": step_roofline(wall),
        "kernels": kernels,
        "kernel_ms": {k: 1e3 * v for k, v in kt.items()},
        "kernel_GBs": {k: ALGO_BYTES[k] * N / v / 1e9 for k, v in kt.items()},
        "copy_GBs": copy_gbs, "copy_frac": copy_gbs / PEAK_HBM_GBS,
        "divb_max": diag["divb_max"],
    }
    if args.precision == "fast" and not args.no_exact:
        # the bitwise-exact mode on the same box (the state just evolves further)
        print("[bench] exact leg", file=sys.stderr, flush=True)
        box.set_precision("exact")
        we = timed_steps(box, args.steps, 1, graph)
        line["exact"] = {"value": N * args.steps / we, "ms_per_step": 1e3 * we / args.steps,
                         "step_roofline": step_roofline(we)}
    del box
    torch.cuda.empty_cache()
    if not args.no_ensemble:
        print("[bench] batched ensemble", file=sys.stderr, flush=True)
        line["ensemble"] = ensemble_bench(precision=args.precision)
    if not args.no_sizes:
        print("[bench] 64^3 and 256^3 steps", file=sys.stderr, flush=True)
        line["sizes"] = size_sweep()
        torch.cuda.empty_cache()
    if not args.no_corrector:
        print("[bench] learned corrector", file=sys.stderr, flush=True)
        line["corrector"] = corrector_bench(args.corrector_seeds, args.corrector_steps)
    if not args.no_cpu:
        # the CPU leg steps a 256^3 box of the same generator (the GPU makes
        # the IC, bitwise equal to the host generator's)
        nc = min(n, 256)
        print(f"[bench] CPU baseline at {nc}^3", file=sys.stderr, flush=True)
        uc, bc = solver.ic_device(nc, nc, nc, solver.ICSpec(seed=1, amp_v=args.amp_v,
                                                            amp_a=args.amp_a))
        torch.cuda.synchronize()
        line["cpu_baseline"] = cpu_baseline(uc.cpu().numpy(), bc.cpu().numpy(), nc)
        print("[bench] CPU at 64^3", file=sys.stderr, flush=True)
        line["cpu_baseline"]["small_64"] = cpu_small()
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
