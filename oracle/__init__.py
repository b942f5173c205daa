"""CPU checkers for the MHD step.  TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's CPU-baseline leg use this
package, and only as the checker: the product (heterogeneous_mhd_amd) never
imports it.

* ``Oracle("f32"|"f64")`` -- the restatement ``mhd_ref.c`` (numpy in/out).
* ``RefC("f32"|"f64")`` -- the reference's own C kernels compiled from
  /root/reference/C into ``oracle/_ref`` (present only where that tree
  exists; used to write and re-check the golden vectors).

Every function of mhd_ref.c is pinned by the golden vectors of the compiled
reference (tests/golden) and by the closed-form checks listed in its header.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_vp = C.c_void_p


def _ptr(a: np.ndarray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


class Oracle:
    """ctypes view of libmhd_oracle_{f32,f64}.so."""

    def __init__(self, prec: str = "f32", threads: int = 1):
        self.prec = prec
        self.dtype = np.float32 if prec == "f32" else np.float64
        path = HERE / f"libmhd_oracle_{prec}.so"
        if not path.exists():
            raise FileNotFoundError(f"{path} missing: run make -C oracle")
        self.lib = C.CDLL(str(path))
        R = C.c_float if prec == "f32" else C.c_double
        self.R = R
        L = self.lib
        L.mhdref_set_threads.argtypes = [C.c_int]
        L.mhdref_set_cuda_semantics.argtypes = [C.c_int]
        L.mhdref_fluidx.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_int, R, _vp, _vp]
        L.mhdref_advectbyzx.argtypes = [_vp, C.c_int, C.c_int, C.c_int, R, _vp, _vp]
        L.mhdref_transpose12.argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int]
        L.mhdref_transpose13.argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int]
        L.mhdref_transpose13.restype = C.c_int
        L.mhdref_transpose13_any.argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int]
        L.mhdref_transpose13_any.restype = None
        L.mhdref_calcfl.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_int]
        L.mhdref_calcfl.restype = R
        L.mhdref_cfl_cell.argtypes = [_vp, _vp]
        L.mhdref_cfl_cell.restype = R
        L.mhdref_step.argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double, R,
                                  C.POINTER(R)]
        L.mhdref_step.restype = R
        L.mhdref_step_any.argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double, R,
                                  C.POINTER(R)]
        L.mhdref_step_any.restype = R
        L.mhdref_diagnostics.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_int,
                                         C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.mhdref_set_threads(threads)

    def set_threads(self, t: int):
        self.lib.mhdref_set_threads(int(t))

    def _a(self, x):
        return np.ascontiguousarray(x, dtype=self.dtype)

    def fluidx(self, ub, bb, nx, ny, nz, dt):
        """u after fluidx from update state (ub, bb)."""
        ub, bb = self._a(ub), self._a(bb)
        u = np.empty_like(ub)
        self.lib.mhdref_fluidx(_ptr(u), _ptr(bb), nx, ny, nz, self.R(dt), _ptr(ub), _ptr(bb))
        return u

    def advectbyzx(self, ub, bb, nx, ny, nz, dt, b_main=None):
        ub, bb = self._a(ub), self._a(bb)
        b = self._a(bb if b_main is None else b_main).copy()
        self.lib.mhdref_advectbyzx(_ptr(b), nx, ny, nz, self.R(dt), _ptr(ub), _ptr(bb))
        return b

    def transpose12(self, ub, bb, nx, ny, nz):
        ub, bb = self._a(ub), self._a(bb)
        u, b = np.empty_like(ub), np.empty_like(bb)
        self.lib.mhdref_transpose12(_ptr(u), _ptr(b), _ptr(ub), _ptr(bb), nx, ny, nz)
        return u, b

    def transpose13(self, ub, bb, nx, ny, nz):
        ub, bb = self._a(ub), self._a(bb)
        u, b = np.empty_like(ub), np.empty_like(bb)
        rc = self.lib.mhdref_transpose13(_ptr(u), _ptr(b), _ptr(ub), _ptr(bb), nx, ny, nz)
        if rc != 0:
            raise ValueError("transpose13 needs nx == nz")
        return u, b

    def set_cuda_semantics(self, on: bool):
        """Switch this library between the reference's C path (default) and
        the semantics of its CUDA path (the B1 neighbour index and the
        face-averaged CFL); process-wide for this precision's library."""
        self.lib.mhdref_set_cuda_semantics(1 if on else 0)

    def transpose13_any(self, ub, bb, nx, ny, nz):
        """x<->z swap for any nx, nz (the reference refuses nx != nz)."""
        ub, bb = self._a(ub), self._a(bb)
        u, b = np.empty_like(ub), np.empty_like(bb)
        self.lib.mhdref_transpose13_any(_ptr(u), _ptr(b), _ptr(ub), _ptr(bb), nx, ny, nz)
        return u, b

    def calcfl(self, u, b, nx, ny, nz):
        u, b = self._a(u), self._a(b)
        return self.dtype(self.lib.mhdref_calcfl(_ptr(u), _ptr(b), nx, ny, nz))

    def cfl_cell(self, q, bb):
        q, bb = self._a(q), self._a(bb)
        return self.dtype(self.lib.mhdref_cfl_cell(_ptr(q), _ptr(bb)))

    def run(self, u, b, nx, ny, nz, nsteps, cfl=0.7, tf=0.0, t=0.0, on_step=None, any_shape=False):
        """nsteps full steps; returns (u, b, t, [dt per step]).  any_shape:
        boxes with ny != nz, through transpose13_any."""
        u, b = self._a(u).copy(), self._a(b).copy()
        uu, bu = u.copy(), b.copy()
        tt = self.R(t)
        dts = []
        step = self.lib.mhdref_step_any if any_shape else self.lib.mhdref_step
        for s in range(nsteps):
            dt = step(_ptr(u), _ptr(b), _ptr(uu), _ptr(bu), nx, ny, nz, float(cfl), self.R(tf),
                      C.byref(tt))
            dts.append(self.dtype(dt))
            if on_step is not None:
                on_step(s, u, b, dts[-1])
        return u, b, self.dtype(tt.value), dts

    def diagnostics(self, u, b, nx, ny, nz):
        u, b = self._a(u), self._a(b)
        s = (C.c_double * 5)()
        d = (C.c_double * 2)()
        self.lib.mhdref_diagnostics(_ptr(u), _ptr(b), nx, ny, nz, s, d)
        return {"mass": s[0], "mom": (s[1], s[2], s[3]), "energy": s[4], "divb_max": d[0],
                "divb_rms": d[1]}


class RefC:
    """ctypes view of the compiled reference (oracle/_ref/libmhd_refc_*.so).

    Dimensions are limited to <= 100 by the reference's static buffers."""

    def __init__(self, prec: str = "f32"):
        self.dtype = np.float32 if prec == "f32" else np.float64
        path = HERE / "_ref" / f"libmhd_refc_{prec}.so"
        if not path.exists():
            raise FileNotFoundError(f"{path} missing: run make -C oracle ref")
        self.lib = C.CDLL(str(path))
        R = C.c_float if prec == "f32" else C.c_double
        self.R = R
        L = self.lib
        for f in ("refc_fluidx", "refc_advectbyzx"):
            getattr(L, f).argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_int, R, _vp, _vp]
        for f in ("refc_transpose12", "refc_transpose13"):
            getattr(L, f).argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int]
        L.refc_calcfl.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_int, C.POINTER(R)]
        L.refc_step.argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double, R,
                                C.POINTER(R), C.POINTER(R)]

    def _a(self, x):
        return np.ascontiguousarray(x, dtype=self.dtype)

    def fluidx(self, ub, bb, nx, ny, nz, dt):
        ub, bb = self._a(ub).copy(), self._a(bb).copy()
        u, b = ub.copy(), bb.copy()
        assert self.lib.refc_fluidx(_ptr(u), _ptr(b), nx, ny, nz, self.R(dt), _ptr(ub), _ptr(bb)) == 0
        return u

    def advectbyzx(self, ub, bb, nx, ny, nz, dt):
        ub, bb = self._a(ub).copy(), self._a(bb).copy()
        u, b = ub.copy(), bb.copy()
        assert self.lib.refc_advectbyzx(_ptr(u), _ptr(b), nx, ny, nz, self.R(dt), _ptr(ub), _ptr(bb)) == 0
        return b

    def transpose12(self, ub, bb, nx, ny, nz):
        ub, bb = self._a(ub).copy(), self._a(bb).copy()
        u, b = np.empty_like(ub), np.empty_like(bb)
        assert self.lib.refc_transpose12(_ptr(u), _ptr(b), _ptr(ub), _ptr(bb), nx, ny, nz) == 0
        return u, b

    def transpose13(self, ub, bb, nx, ny, nz):
        ub, bb = self._a(ub).copy(), self._a(bb).copy()
        u, b = np.empty_like(ub), np.empty_like(bb)
        assert self.lib.refc_transpose13(_ptr(u), _ptr(b), _ptr(ub), _ptr(bb), nx, ny, nz) == 0
        return u, b

    def calcfl(self, u, b, nx, ny, nz):
        u, b = self._a(u).copy(), self._a(b).copy()
        out = self.R()
        assert self.lib.refc_calcfl(_ptr(u), _ptr(b), nx, ny, nz, C.byref(out)) == 0
        return self.dtype(out.value)

    def run(self, u, b, nx, ny, nz, nsteps, cfl=0.7, tf=0.0, t=0.0, on_step=None):
        u, b = self._a(u).copy(), self._a(b).copy()
        uu, bu = u.copy(), b.copy()
        tt = self.R(t)
        dts = []
        for s in range(nsteps):
            dt = self.R()
            assert self.lib.refc_step(_ptr(u), _ptr(b), _ptr(uu), _ptr(bu), nx, ny, nz, float(cfl),
                                      self.R(tf), C.byref(tt), C.byref(dt)) == 0
            dts.append(self.dtype(dt.value))
            if on_step is not None:
                on_step(s, u, b, dts[-1])
        return u, b, self.dtype(tt.value), dts
