"""Writes the golden vectors under tests/golden/ from the reference's own C
kernels (oracle/_ref, compiled from /root/reference/C by ``make -C oracle
ref``).  Run in the build container only; the fixtures it writes are data
(inputs, outputs, digests) and travel without the reference.

    python tests/golden/make_golden.py

Inputs come from the library's seeded host initial condition
(heterogeneous_mhd_amd.solver.ic_host), optionally perturbed by seeded
noise so that every limiter branch and sign case is exercised.
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from heterogeneous_mhd_amd import solver  # noqa: E402
from oracle import RefC  # noqa: E402


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def rough(u, b, seed, scale=0.05):
    """IC plus seeded multiplicative noise on every component (rho stays > 0)."""
    rng = np.random.default_rng(seed)
    u = u * (1 + scale * rng.standard_normal(u.shape)).astype(np.float32)
    b = b + (scale * rng.standard_normal(b.shape)).astype(np.float32)
    return u.astype(np.float32), b.astype(np.float32)


def op_case(name, nx, ny, nz, spec, dt, noise_seed=None):
    u, b = solver.ic_host(nx, ny, nz, spec)
    if noise_seed is not None:
        u, b = rough(u, b, noise_seed)
    r32, r64 = RefC("f32"), RefC("f64")
    out = {"u": u, "b": b, "dims": np.array([nx, ny, nz], np.int32), "dt": np.float32(dt)}
    out["fluidx_u"] = r32.fluidx(u, b, nx, ny, nz, dt)
    out["advectbyzx_b"] = r32.advectbyzx(u, b, nx, ny, nz, dt)
    out["calcfl"] = np.float32(r32.calcfl(u, b, nx, ny, nz))
    dig = {}
    tu, tb = r32.transpose12(u, b, nx, ny, nz)
    dig["transpose12_u"], dig["transpose12_b"] = sha(tu), sha(tb)
    if nx == nz:
        tu, tb = r32.transpose13(u, b, nx, ny, nz)
        dig["transpose13_u"], dig["transpose13_b"] = sha(tu), sha(tb)
    # FP64 build on the same (float-exact) input
    u64, b64 = u.astype(np.float64), b.astype(np.float64)
    dig["f64_fluidx_u"] = sha(r64.fluidx(u64, b64, nx, ny, nz, float(np.float32(dt))))
    dig["f64_advectbyzx_b"] = sha(r64.advectbyzx(u64, b64, nx, ny, nz, float(np.float32(dt))))
    dig["f64_calcfl"] = float(r64.calcfl(u64, b64, nx, ny, nz))
    out["digests"] = np.array(json.dumps(dig))
    np.savez_compressed(HERE / f"ops_{name}.npz", **out)
    print("wrote", f"ops_{name}.npz", flush=True)


def traj_case(name, nx, ny, nz, spec, nsteps, cfl, keep_final=False):
    u, b = solver.ic_host(nx, ny, nz, spec)
    rec = {"dims": [nx, ny, nz], "spec": spec.__dict__, "nsteps": nsteps, "cfl": cfl,
           "ic_sha_u": sha(u), "ic_sha_b": sha(b)}
    for prec in ("f32", "f64"):
        steps = []

        def on_step(s, uu, bb, dt):
            steps.append({"dt": float(dt), "sha_u": sha(uu), "sha_b": sha(bb)})

        dtype = np.float32 if prec == "f32" else np.float64
        uf, bf, t, _ = RefC(prec).run(u.astype(dtype), b.astype(dtype), nx, ny, nz, nsteps,
                                      cfl=cfl, on_step=on_step)
        rec[prec] = {"t": float(t), "steps": steps}
        if keep_final and prec == "f32":
            np.savez_compressed(HERE / f"traj_{name}_final.npz", u=uf, b=bf)
    (HERE / f"traj_{name}.json").write_text(json.dumps(rec, indent=1))
    print("wrote", f"traj_{name}.json", flush=True)


def oracle_traj_case(name, n, spec, nsteps, cfl, threads=8):
    """A trajectory too large for the compiled reference's static buffers
    (dims <= 100): per-step SHA-256 and dt of the oracle's FP32 path, which
    equals the compiled reference bitwise on every case above, plus the FP32
    path's error against the FP64 oracle at the last step."""
    from oracle import Oracle
    u, b = solver.ic_host(n, n, n, spec)
    rec = {"dims": [n, n, n], "spec": spec.__dict__, "nsteps": nsteps, "cfl": cfl,
           "ic_sha_u": sha(u), "ic_sha_b": sha(b), "source": "oracle (pinned to the compiled reference)"}
    steps = []

    def on_step(s, uu, bb, dt):
        steps.append({"dt": float(dt), "sha_u": sha(uu), "sha_b": sha(bb)})
        if (s + 1) % 10 == 0:
            print(f"  {name} f32 step {s + 1}", flush=True)

    u32, b32, t32, _ = Oracle("f32", threads=threads).run(u, b, n, n, n, nsteps, cfl=cfl, on_step=on_step)
    rec["f32"] = {"t": float(t32), "steps": steps}
    u64, b64, t64, dts64 = Oracle("f64", threads=threads).run(u.astype(np.float64), b.astype(np.float64),
                                                             n, n, n, nsteps, cfl=cfl)
    g32 = {"rho": u32.reshape(-1, 5)[:, 0], "mom": u32.reshape(-1, 5)[:, 1:4], "E": u32.reshape(-1, 5)[:, 4],
           "B": b32.reshape(-1, 3)}
    g64 = {"rho": u64.reshape(-1, 5)[:, 0], "mom": u64.reshape(-1, 5)[:, 1:4], "E": u64.reshape(-1, 5)[:, 4],
           "B": b64.reshape(-1, 3)}
    err = {}
    for k in g32:
        d = np.linalg.norm(g32[k].astype(np.float64) - g64[k])
        err[k] = {"rel_l2": float(d / np.linalg.norm(g64[k])),
                  "rel_fluct_l2": float(d / np.linalg.norm(g64[k] - g64[k].mean(axis=0)))}
    rec["f64"] = {"t": float(t64), "dt_last": float(dts64[-1])}
    rec["f32_vs_f64"] = err
    (HERE / f"traj_{name}.json").write_text(json.dumps(rec, indent=1))
    print("wrote", f"traj_{name}.json", flush=True)


def main():
    smooth = solver.ICSpec(seed=7, amp_v=0.3, amp_a=0.1)
    op_case("16x16x16", 16, 16, 16, smooth, 0.1)
    op_case("24x16x16", 24, 16, 16, smooth, 0.1)
    op_case("rough_12x10x10", 12, 10, 10, solver.ICSpec(seed=11, amp_v=0.4, amp_a=0.1), 0.2,
            noise_seed=5)
    op_case("7x3x7", 7, 3, 7, solver.ICSpec(seed=3, amp_v=0.2, amp_a=0.05), 0.15, noise_seed=9)
    op_case("1x4x4", 1, 4, 4, solver.ICSpec(seed=4, amp_v=0.2, amp_a=0.05), 0.15, noise_seed=10)
    traj_case("16", 16, 16, 16, smooth, 10, 0.7, keep_final=True)
    traj_case("24x16x16", 24, 16, 16, solver.ICSpec(seed=2, amp_v=0.2, amp_a=0.08), 10, 0.7)
    traj_case("32", 32, 32, 32, solver.ICSpec(seed=5, amp_v=0.3, amp_a=0.1), 10, 0.7)
    # the CPU-runnable headline case: 64^3, amplitudes 0.1 / 0.05, 10 steps, CFL 0.7
    traj_case("64_c1", 64, 64, 64, solver.ICSpec(seed=1, amp_v=0.1, amp_a=0.05), 10, 0.7)
    # the single-box GPU workload: 256^3, same generator and parameters, 100 steps
    oracle_traj_case("256_c2", 256, solver.ICSpec(seed=1, amp_v=0.1, amp_a=0.05), 100, 0.7)


if __name__ == "__main__":
    if sys.argv[1:] == ["256_c2"]:
        oracle_traj_case("256_c2", 256, solver.ICSpec(seed=1, amp_v=0.1, amp_a=0.05), 100, 0.7)
    else:
        main()
