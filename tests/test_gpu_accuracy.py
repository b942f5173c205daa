"""Error of the GPU step against the FP64 build of the reference, per
workload, for the bitwise-exact mode (identical to the reference FP32 path)
and the FP32 tolerance mode.

Metric per field group (rho, momentum, E, B): relative L2 of the difference,
||x - x64|| / ||x64||, and the same normalised by the fluctuation,
||x - x64|| / ||x64 - mean(x64)||.  Each test prints a JSON line with the
numbers (the report); the assertion is that the fast mode's L2 error is at
most 3x the reference FP32 path's (measured about 1-2x) and that div b stays
at round-off.

The 256^3 x 100-step workload runs the FP64 oracle on the host for a few
minutes, so it is opt-in: MHD_ACCURACY_FULL=1.  1024^3 is not compared with
FP64: the host oracle would need about half an hour for its 20 steps.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _groups(u, b):
    u = np.asarray(u).reshape(-1, 5).astype(np.float64)
    b = np.asarray(b).reshape(-1, 3).astype(np.float64)
    return {"rho": u[:, 0], "mom": u[:, 1:4], "E": u[:, 4], "B": b}


def errors(x, ref):
    gx, gr = _groups(*x), _groups(*ref)
    out = {}
    for k in gx:
        d = np.linalg.norm(gx[k] - gr[k])
        out[k] = {"rel_l2": float(d / np.linalg.norm(gr[k])),
                  "rel_fluct_l2": float(d / np.linalg.norm(gr[k] - gr[k].mean(axis=0)))}
    return out


def _run_gpu(u0, b0, n, steps, precision):
    import torch
    from heterogeneous_mhd_amd import solver
    box = solver.Box(n, n, n, cfl=0.7, precision=precision)
    box.set_state(torch.from_numpy(u0), torch.from_numpy(b0))
    box.run(steps, graph=True)
    out = box.host_state()
    divb = box.diagnostics()["divb_max"]
    box.close()
    return out, divb


def _report(name, n, steps, spec, threads=16):
    from heterogeneous_mhd_amd import solver
    from oracle import Oracle
    u0, b0 = solver.ic_host(n, n, n, spec)

    def progress(s, u, b, dt):
        if (s + 1) % 5 == 0:
            print(f"  FP64 oracle {name}: step {s + 1}/{steps}", flush=True)
    r64 = Oracle("f64", threads=threads).run(u0.astype(np.float64), b0.astype(np.float64), n, n, n,
                                            steps, on_step=progress)
    exact, divb_e = _run_gpu(u0, b0, n, steps, "exact")
    fast, divb_f = _run_gpu(u0, b0, n, steps, "fast")
    rep = {"workload": name, "n": n, "steps": steps, "amp_v": spec.amp_v, "amp_a": spec.amp_a,
           "exact_vs_fp64": errors(exact, r64[:2]), "fast_vs_fp64": errors(fast, r64[:2]),
           "divb_max": {"exact": divb_e, "fast": divb_f}}
    print("ACCURACY " + json.dumps(rep), flush=True)
    for k in rep["exact_vs_fp64"]:
        e, f = rep["exact_vs_fp64"][k]["rel_l2"], rep["fast_vs_fp64"][k]["rel_l2"]
        assert f <= 3.0 * e + 1e-7, (k, e, f)
    assert divb_e <= 1e-5 and divb_f <= 1e-5
    return rep


def test_cpu_runnable_workload_64_10_steps(gpu):
    from heterogeneous_mhd_amd import solver
    print("accuracy 64^3 x 10 steps", flush=True)
    _report("64^3 periodic, 10 steps (CPU-runnable workload)", 64, 10,
            solver.ICSpec(seed=1, amp_v=0.1, amp_a=0.05))


def test_ensemble_boxes_50_steps(gpu):
    """Three box sizes of the ensemble workload with their own amplitudes."""
    from heterogeneous_mhd_amd import solver
    for n, av, aa in ((32, 0.4, 0.1), (48, 0.05, 0.01), (64, 0.25, 0.06)):
        print(f"accuracy ensemble box {n}^3 x 50 steps", flush=True)
        _report(f"ensemble box {n}^3, 50 steps", n, 50, solver.ICSpec(seed=n, amp_v=av, amp_a=aa))


@pytest.mark.skipif(os.environ.get("MHD_ACCURACY_FULL") != "1",
                    reason="several minutes of FP64 on the host; set MHD_ACCURACY_FULL=1")
def test_single_box_256_100_steps(gpu):
    from heterogeneous_mhd_amd import solver
    print("accuracy 256^3 x 100 steps", flush=True)
    _report("256^3 periodic, 100 steps", 256, 100, solver.ICSpec(seed=1, amp_v=0.1, amp_a=0.05))
