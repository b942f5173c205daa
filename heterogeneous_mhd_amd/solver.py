"""Device-resident driver of the full MHD step, the seeded initial condition
and the FP64 diagnostics.

``Box`` owns one periodic box on the GPU: the state (u, b) in the reference
layout, a second pair of buffers for the ping-pong that replaces the
reference's copy_matrix calls, and a native step plan (``mhd_plan``) that
keeps t and dt on the device.  ``Box.run(n)`` advances n full steps in the
sweep order of C/mhd.c:59-124, optionally by replaying a captured HIP graph.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import MHD_PREC_EXACT, MHD_PREC_FAST, MHDError, check, lib

GAMMA = 5.0 / 3.0


@dataclass(frozen=True)
class ICSpec:
    """Seeded random-phase initial condition (see include/mhd_hip.h)."""
    seed: int = 1
    amp_v: float = 0.1
    amp_a: float = 0.05
    rho0: float = 1.0
    p0: float = 0.6
    gamma: float = GAMMA
    kmax: int = 4


def ic_host(nx: int, ny: int, nz: int, spec: ICSpec = ICSpec()):
    """Initial condition on the host: returns (u, b) float32 numpy arrays of
    5*N and 3*N elements.  Bitwise equal to :func:`ic_device`."""
    u = np.empty(5 * nx * ny * nz, np.float32)
    b = np.empty(3 * nx * ny * nz, np.float32)
    check(lib().mhd_ic_generate_host(u.ctypes.data, b.ctypes.data, nx, ny, nz, spec.seed,
                                     spec.amp_v, spec.amp_a, spec.rho0, spec.p0, spec.gamma,
                                     spec.kmax), "ic_host")
    return u, b


def ic_device(nx: int, ny: int, nz: int, spec: ICSpec = ICSpec(), device="cuda", u=None, b=None):
    """Initial condition generated on the GPU into (u, b) (allocated if None)."""
    n = nx * ny * nz
    if u is None:
        u = torch.empty(5 * n, dtype=torch.float32, device=device)
    if b is None:
        b = torch.empty(3 * n, dtype=torch.float32, device=device)
    scratch = torch.empty(3 * n, dtype=torch.float64, device=u.device)
    check(lib().mhd_ic_generate(u.data_ptr(), b.data_ptr(), nx, ny, nz, spec.seed, spec.amp_v,
                                spec.amp_a, spec.rho0, spec.p0, spec.gamma, spec.kmax,
                                scratch.data_ptr(), torch.cuda.current_stream().cuda_stream),
          "ic_device")
    del scratch
    return u, b


def diagnostics(u: torch.Tensor, b: torch.Tensor, nx: int, ny: int, nz: int) -> dict:
    """FP64 conserved sums and div b on the device (deterministic)."""
    nbytes = lib().mhd_diag_workspace_bytes(nx, ny, nz)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=u.device)
    out = (C.c_double * 7)()
    check(lib().mhd_diagnostics(u.data_ptr(), b.data_ptr(), nx, ny, nz, ws.data_ptr(), out,
                                torch.cuda.current_stream().cuda_stream), "diagnostics")
    v = list(out)
    return {"mass": v[0], "mom": (v[1], v[2], v[3]), "energy": v[4], "divb_max": v[5],
            "divb_rms": v[6]}


def _prec(precision: str) -> int:
    modes = {"exact": MHD_PREC_EXACT, "fast": MHD_PREC_FAST}
    if precision not in modes:
        raise MHDError(f"precision must be one of {sorted(modes)}, not {precision!r}")
    return modes[precision]


class Box:
    """One periodic box advanced on the GPU by the native step driver."""

    def __init__(self, nx: int, ny: int, nz: int, cfl: float = 0.7, tf: float = 0.0,
                 device="cuda", stream: torch.cuda.Stream | None = None, precision: str = "exact"):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = nx * ny * nz
        self.device = torch.device(device)
        # a stream of its own (default priority): HIP graphs cannot be
        # captured on the null stream
        self.stream = stream if stream is not None else torch.cuda.Stream(self.device)
        self.u = torch.zeros(5 * self.n, dtype=torch.float32, device=self.device)
        self.b = torch.zeros(3 * self.n, dtype=torch.float32, device=self.device)
        self.u2 = torch.empty_like(self.u)
        self.b2 = torch.empty_like(self.b)
        h = C.c_void_p()
        check(lib().mhd_plan_create(C.byref(h), nx, ny, nz, float(cfl), float(tf),
                                    self.stream.cuda_stream), "plan_create")
        self._plan = h
        self.set_precision(precision)

    def set_precision(self, precision: str):
        """'exact' (bitwise the reference) or 'fast' (FP32 tolerance mode)."""
        check(lib().mhd_plan_set_precision(self._plan, _prec(precision)), "plan_set_precision")
        self.precision = precision

    def close(self):
        if getattr(self, "_plan", None):
            lib().mhd_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -------------------------------------------------------------
    def set_state(self, u, b, t: float = 0.0):
        self.stream.synchronize()  # steps still queued on the box's stream use the old state
        self.u.copy_(torch.as_tensor(u).reshape(-1))
        self.b.copy_(torch.as_tensor(b).reshape(-1))
        torch.cuda.current_stream(self.device).synchronize()
        check(lib().mhd_plan_set_time(self._plan, float(t)), "set_time")

    def init(self, spec: ICSpec = ICSpec()):
        with torch.cuda.stream(self.stream):
            ic_device(self.nx, self.ny, self.nz, spec, u=self.u, b=self.b)
        self.stream.synchronize()
        check(lib().mhd_plan_set_time(self._plan, 0.0), "set_time")

    def set_time(self, t: float):
        """Sets t; the next run recomputes dt from the state."""
        self.stream.synchronize()
        check(lib().mhd_plan_set_time(self._plan, float(t)), "set_time")

    def set_tf(self, tf: float):
        """End time of the clamp (0 disables it)."""
        check(lib().mhd_plan_set_tf(self._plan, float(tf)), "plan_set_tf")

    def state_changed(self):
        """Call after modifying u/b outside run(): the next step recomputes dt."""
        check(lib().mhd_plan_state_changed(self._plan), "plan_state_changed")

    def advance_to(self, t_end: float, max_steps: int = 1 << 20) -> int:
        """Steps until t reaches t_end exactly (the clamp makes the last step
        land on it); returns the number of steps taken."""
        self.set_tf(t_end)
        steps = 0
        while steps < max_steps:
            t, dt = self.time()
            if t >= t_end or (steps and dt == 0.0):
                break
            self.run(1, graph=False)
            steps += 1
        self.set_tf(0.0)
        return steps

    # -- stepping ----------------------------------------------------------
    def run(self, nsteps: int, graph: bool = False):
        check(lib().mhd_plan_run(self._plan, self.u.data_ptr(), self.b.data_ptr(),
                                 self.u2.data_ptr(), self.b2.data_ptr(), int(nsteps), int(graph)),
              "plan_run")

    def time(self):
        t, dt = C.c_float(), C.c_float()
        check(lib().mhd_plan_get_time(self._plan, C.byref(t), C.byref(dt)), "get_time")
        return t.value, dt.value

    def dt_tensor(self) -> torch.Tensor:
        """The current dt as a one-element device tensor (a copy)."""
        return torch.tensor([self.time()[1]], dtype=torch.float32, device=self.device)

    def sync(self):
        check(lib().mhd_plan_sync(self._plan), "sync")

    def host_state(self):
        """(u, b) as host numpy arrays, after the box's queued work finished."""
        self.sync()
        return self.u.cpu().numpy(), self.b.cpu().numpy()

    def diagnostics(self) -> dict:
        with torch.cuda.stream(self.stream):
            return diagnostics(self.u, self.b, self.nx, self.ny, self.nz)


class Ensemble:
    """Independent cubic periodic boxes of different sides advanced together
    on the GPU: one persistent launch per sweep for all boxes, per-box t and
    CFL dt, optional HIP-graph replay.  Each box evolves bitwise as a
    :class:`Box` of the same side would."""

    def __init__(self, sides, cfl: float = 0.7, tf: float = 0.0, device="cuda",
                 stream: torch.cuda.Stream | None = None, precision: str = "exact"):
        self.sides = [int(s) for s in sides]
        self.nbox = len(self.sides)
        self.device = torch.device(device)
        self.stream = stream if stream is not None else torch.cuda.Stream(self.device)
        arr = (C.c_int * self.nbox)(*self.sides)
        h = C.c_void_p()
        check(lib().mhd_ensemble_create(C.byref(h), self.nbox, arr, float(cfl), float(tf),
                                        self.stream.cuda_stream), "ensemble_create")
        self._e = h
        uf, bf = C.c_size_t(), C.c_size_t()
        uo = (C.c_longlong * self.nbox)()
        bo = (C.c_longlong * self.nbox)()
        check(lib().mhd_ensemble_sizes(h, C.byref(uf), C.byref(bf), uo, bo), "ensemble_sizes")
        self.u_off, self.b_off = list(uo), list(bo)
        self.u = torch.zeros(uf.value, dtype=torch.float32, device=self.device)
        self.b = torch.zeros(bf.value, dtype=torch.float32, device=self.device)
        self.u2 = torch.empty_like(self.u)
        self.b2 = torch.empty_like(self.b)
        check(lib().mhd_ensemble_bind(h, self.u.data_ptr(), self.b.data_ptr(), self.u2.data_ptr(),
                                      self.b2.data_ptr()), "ensemble_bind")
        self.set_precision(precision)

    def set_precision(self, precision: str):
        """'exact' (bitwise the reference) or 'fast' (FP32 tolerance mode)."""
        check(lib().mhd_ensemble_set_precision(self._e, _prec(precision)), "ensemble_set_precision")
        self.precision = precision

    def close(self):
        if getattr(self, "_e", None):
            lib().mhd_ensemble_destroy(self._e)
            self._e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def cells(self) -> int:
        return sum(n ** 3 for n in self.sides)

    def box_views(self, i: int):
        """(u, b) of box i: views into the packed device state."""
        n = self.sides[i]
        return (self.u[self.u_off[i]:self.u_off[i] + 5 * n ** 3],
                self.b[self.b_off[i]:self.b_off[i] + 3 * n ** 3])

    def init(self, specs):
        """Seeded IC per box (one ICSpec per box)."""
        nmax = max(self.sides)
        scratch = torch.empty(3 * nmax ** 3, dtype=torch.float64, device=self.device)
        for i, sp in enumerate(specs):
            check(lib().mhd_ensemble_init_box(self._e, i, sp.seed, sp.amp_v, sp.amp_a, sp.rho0, sp.p0,
                                              sp.gamma, sp.kmax, scratch.data_ptr()), "ensemble_init_box")
        self.stream.synchronize()
        check(lib().mhd_ensemble_reset_time(self._e), "ensemble_reset_time")

    def set_box_state(self, i: int, u, b):
        """Overwrite box i (host or device arrays); t of every box is reset."""
        uv, bv = self.box_views(i)
        self.stream.synchronize()  # steps still queued on the ensemble's stream
        uv.copy_(torch.as_tensor(u).reshape(-1))
        bv.copy_(torch.as_tensor(b).reshape(-1))
        torch.cuda.current_stream(self.device).synchronize()
        check(lib().mhd_ensemble_reset_time(self._e), "ensemble_reset_time")

    def run(self, nsteps: int, graph: bool = False):
        check(lib().mhd_ensemble_run(self._e, int(nsteps), int(graph)), "ensemble_run")

    def times(self):
        t = (C.c_float * self.nbox)()
        dt = (C.c_float * self.nbox)()
        check(lib().mhd_ensemble_get_times(self._e, t, dt), "ensemble_get_times")
        return list(t), list(dt)

    def sync(self):
        self.stream.synchronize()

    def box_state(self, i: int):
        self.sync()
        u, b = self.box_views(i)
        return u.cpu().numpy(), b.cpu().numpy()

    def config(self) -> dict:
        v = [C.c_int() for _ in range(5)]
        check(lib().mhd_ensemble_get_config(self._e, *[C.byref(x) for x in v]), "ensemble_get_config")
        return dict(zip(("shape", "grid", "block", "cells_per_thread", "threads_per_pencil"),
                        (x.value for x in v)))

    def set_config(self, shape: int, grid: int):
        check(lib().mhd_ensemble_set_config(self._e, int(shape), int(grid)), "ensemble_set_config")

    def tune(self, reps: int = 2) -> float:
        """Autotune the launch shape and persistent grid on the current state
        (which is restored); returns the best time per step in ms."""
        best = C.c_float()
        check(lib().mhd_ensemble_tune(self._e, int(reps), C.byref(best)), "ensemble_tune")
        return best.value

    def diagnostics(self, i: int) -> dict:
        u, b = self.box_views(i)
        n = self.sides[i]
        with torch.cuda.stream(self.stream):
            return diagnostics(u, b, n, n, n)


def ensemble_specs(nbox: int, seed: int = 7, sides=(32, 48, 64, 96, 128), amp_v=(0.05, 0.4),
                   amp_a=(0.01, 0.1)):
    """Box sides and per-box IC specs drawn from one seed (the batched
    ensemble workload: sides from `sides`, amplitudes uniform in the ranges)."""
    rng = np.random.default_rng(seed)
    s = [int(x) for x in rng.choice(np.asarray(sides), size=nbox)]
    specs = [ICSpec(seed=int(rng.integers(1, 2 ** 31)), amp_v=float(rng.uniform(*amp_v)),
                    amp_a=float(rng.uniform(*amp_a))) for _ in range(nbox)]
    return s, specs
