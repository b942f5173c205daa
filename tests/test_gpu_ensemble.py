"""Batched ensembles: every box of a mixed-size ensemble evolves bitwise as a
single-box run of the same box (state, t and dt per step), in eager and graph
mode, for every launch shape the autotuner may choose."""
import numpy as np
import pytest

from conftest import bits_equal

pytestmark = pytest.mark.gpu


def _single_runs(sides, specs, nsteps, cfl=0.7):
    import torch
    from heterogeneous_mhd_amd import solver
    out = []
    for n, sp in zip(sides, specs):
        box = solver.Box(n, n, n, cfl=cfl)
        box.init(sp)
        box.run(nsteps)
        t, dt = box.time()
        u, b = box.host_state()
        out.append((u, b, t, dt))
        box.close()
    return out


def test_ensemble_boxes_match_single_box_runs(gpu):
    from heterogeneous_mhd_amd import solver
    sides = [48, 16, 64, 24, 33, 32]
    _, specs = solver.ensemble_specs(len(sides), seed=3)
    print("sides", sides, flush=True)
    ref = _single_runs(sides, specs, 3)
    ens = solver.Ensemble(sides, cfl=0.7)
    ens.init(specs)
    ens.run(3)
    t, dt = ens.times()
    for i in range(len(sides)):
        u, b = ens.box_state(i)
        assert bits_equal(u, ref[i][0]) and bits_equal(b, ref[i][1]), f"box {i}"
        assert np.float32(t[i]) == np.float32(ref[i][2]) and np.float32(dt[i]) == np.float32(ref[i][3])
    ens.close()


def test_ensemble_graph_and_every_shape_agree(gpu):
    from heterogeneous_mhd_amd import lib, solver
    sides = [40, 64, 32, 16]
    _, specs = solver.ensemble_specs(len(sides), seed=5)
    ref = _single_runs(sides, specs, 2)
    nshape = lib().mhd_ensemble_num_configs() // 4
    tried = 0
    for shape in range(nshape):
        ens = solver.Ensemble(sides, cfl=0.7)
        try:
            ens.set_config(shape, 4 * 64)
        except Exception:
            ens.close()
            continue  # shape too small for the largest box
        tried += 1
        ens.init(specs)
        ens.run(2, graph=True)
        for i in range(len(sides)):
            u, b = ens.box_state(i)
            assert bits_equal(u, ref[i][0]) and bits_equal(b, ref[i][1]), f"shape {shape} box {i}"
        ens.close()
    assert tried >= 4


def test_ensemble_tune_keeps_state_and_results(gpu):
    from heterogeneous_mhd_amd import solver
    sides = [32, 48, 64]
    _, specs = solver.ensemble_specs(len(sides), seed=9)
    ref = _single_runs(sides, specs, 2)
    ens = solver.Ensemble(sides, cfl=0.7)
    ens.init(specs)
    ms = ens.tune(reps=1)
    assert ms > 0
    cfg = ens.config()
    print("tuned", cfg, ms, flush=True)
    ens.run(2, graph=True)
    t, _ = ens.times()
    for i in range(len(sides)):
        u, b = ens.box_state(i)
        assert bits_equal(u, ref[i][0]) and bits_equal(b, ref[i][1])
        assert np.float32(t[i]) == np.float32(ref[i][2])


@pytest.mark.parametrize("precision", ["exact", "fast"])
def test_ensemble_size_ranges_match_single_box_runs(gpu, precision):
    """With the default shape (2 cells x 64 threads per pencil) a sweep runs
    in ranges: boxes of side 128 on the compile-time pencil length, 96 on the
    generic two-cell form, 64 on the one-cell fixed form and 40, 32 on the
    one-cell generic form.  Every box still equals its single-box run."""
    import torch
    from heterogeneous_mhd_amd import solver
    sides = [64, 128, 32, 96, 40]
    _, specs = solver.ensemble_specs(len(sides), seed=11)
    print("ranges", sides, precision, flush=True)
    ref = []
    for n, sp in zip(sides, specs):
        box = solver.Box(n, n, n, cfl=0.7, precision=precision)
        box.init(sp)
        box.run(2)
        ref.append((*box.host_state(), *box.time()))
        box.close()
    ens = solver.Ensemble(sides, cfl=0.7, precision=precision)
    ens.init(specs)
    ens.run(2, graph=True)
    t, dt = ens.times()
    for i in range(len(sides)):
        u, b = ens.box_state(i)
        assert bits_equal(u, ref[i][0]) and bits_equal(b, ref[i][1]), f"box {i} ({sides[i]}^3)"
        assert np.float32(t[i]) == np.float32(ref[i][2]) and np.float32(dt[i]) == np.float32(ref[i][3])
    ens.close()
    torch.cuda.empty_cache()
