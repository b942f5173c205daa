"""The reference's operators over device tensors, mirroring C/spu_function.h.

Each function keeps the reference's name, argument order and meaning
(``fluidx(u, b, nx, ny, nz, dt, update_u, update_b, iter)`` etc.), takes
torch tensors that live on the GPU (float32, C-contiguous, u with 5*N and b
with 3*N elements in the reference's Fortran-order AoS layout) and enqueues
the HIP kernel on torch's current stream.  dt may be a Python float or a
one-element float32 device tensor (then the kernel reads it on the device).

There is no CPU path: host tensors are rejected, and a missing library
raises on first use.
"""
from __future__ import annotations

import torch

from ._lib import MHD_PREC_EXACT, MHD_PREC_FAST, check, lib


def _dev(t: torch.Tensor, n: int, what: str) -> int:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{what} must be a GPU tensor")
    if t.dtype != torch.float32 or not t.is_contiguous():
        raise TypeError(f"{what} must be contiguous float32")
    if t.numel() != n:
        raise ValueError(f"{what} has {t.numel()} elements, expected {n}")
    return t.data_ptr()


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _dt(dt):
    if isinstance(dt, torch.Tensor):
        if not dt.is_cuda or dt.dtype != torch.float32 or dt.numel() < 1:
            raise TypeError("dt tensor must be a float32 GPU tensor")
        return dt.data_ptr(), 0.0
    return None, float(dt)


def fluidx(u, b, nx, ny, nz, dt, update_u, update_b, iter=0, precision="exact"):
    """Fluid sweep along the first axis (C/fluidx.c:19-102); writes u.
    precision 'fast' selects the FP32 tolerance mode."""
    del b, iter
    n = nx * ny * nz
    dp, dv = _dt(dt)
    f = lib().mhd_fluidx_fast_async if precision == "fast" else lib().mhd_fluidx_async
    check(f(_dev(u, 5 * n, "u"), _dev(update_u, 5 * n, "update_u"),
            _dev(update_b, 3 * n, "update_b"), nx, ny, nz, dp, dv, _stream()), "fluidx")


def advectbyzx(u, b, nx, ny, nz, dt, update_u, update_b, precision="exact"):
    """CT advection of b2, b3 along the first axis (C/advectbyzx.c:14-23).

    Reads update_u, update_b and main b(1); writes b.  b must not alias
    update_b (the serial reference tolerates that, the parallel kernel does
    not; the C entry point ``advectbyzx`` handles it with a temporary)."""
    del u
    n = nx * ny * nz
    dp, dv = _dt(dt)
    pb = _dev(b, 3 * n, "b")
    f = lib().mhd_advectbyzx_fast_async if precision == "fast" else lib().mhd_advectbyzx_async
    check(f(pb, pb, _dev(update_u, 5 * n, "update_u"), _dev(update_b, 3 * n, "update_b"), nx, ny,
            nz, dp, dv, _stream()), "advectbyzx")


def advect_fluidx(u, b, nx, ny, nz, dt, update_u, update_b, precision="exact"):
    """advectbyzx then fluidx of the same sweep as one kernel (bitwise equal
    to the two calls): b <- CT update of update_b, u <- fluid sweep of
    update_u with the new b.  u, b must not alias update_u, update_b."""
    n = nx * ny * nz
    dp, dv = _dt(dt)
    check(lib().mhd_advect_fluidx_async(_dev(u, 5 * n, "u"), _dev(b, 3 * n, "b"),
                                        _dev(update_u, 5 * n, "update_u"),
                                        _dev(update_b, 3 * n, "update_b"), nx, ny, nz, dp, dv,
                                        MHD_PREC_FAST if precision == "fast" else MHD_PREC_EXACT,
                                        _stream()), "advect_fluidx")


def advect2_fluidx(u, b, nx, ny, nz, dt, update_u, update_b, precision="exact"):
    """advectbyzx twice with the same update_u, then fluidx, as one kernel
    (bitwise equal to the three calls; the intermediate field is not
    written): b <- CT update of the CT update of update_b, u <- fluid sweep
    of update_u with that b.  u, b must not alias update_u, update_b."""
    n = nx * ny * nz
    dp, dv = _dt(dt)
    check(lib().mhd_advect2_fluidx_async(_dev(u, 5 * n, "u"), _dev(b, 3 * n, "b"),
                                         _dev(update_u, 5 * n, "update_u"),
                                         _dev(update_b, 3 * n, "update_b"), nx, ny, nz, dp, dv,
                                         MHD_PREC_FAST if precision == "fast" else MHD_PREC_EXACT,
                                         _stream()), "advect2_fluidx")


def advect_transpose12(u, b, nx, ny, nz, dt, update_u, update_b, precision="exact"):
    """advectbyzx then transpose12 as one kernel (bitwise equal to the two
    calls, C/mhd.c:74, :78): u, b <- transpose12(update_u, CT update of
    update_b), dims (ny, nx, nz).  No output may overlap an input."""
    n = nx * ny * nz
    dp, dv = _dt(dt)
    check(lib().mhd_advect_transpose12_async(_dev(u, 5 * n, "u"), _dev(b, 3 * n, "b"),
                                             _dev(update_u, 5 * n, "update_u"),
                                             _dev(update_b, 3 * n, "update_b"), nx, ny, nz, dp, dv,
                                             MHD_PREC_FAST if precision == "fast" else MHD_PREC_EXACT,
                                             _stream()), "advect_transpose12")


def advect_transpose13(u, b, nx, ny, nz, dt, update_u, update_b, precision="exact"):
    """advectbyzx then the general transpose13 as one kernel (C/mhd.c:84,
    :88): out dims (nz, ny, nx), any nx, nz."""
    n = nx * ny * nz
    dp, dv = _dt(dt)
    check(lib().mhd_advect_transpose13_async(_dev(u, 5 * n, "u"), _dev(b, 3 * n, "b"),
                                             _dev(update_u, 5 * n, "update_u"),
                                             _dev(update_b, 3 * n, "update_b"), nx, ny, nz, dp, dv,
                                             MHD_PREC_FAST if precision == "fast" else MHD_PREC_EXACT,
                                             _stream()), "advect_transpose13")


def transpose12_advect(u, b, nx, ny, nz, dt, update_u, update_b, precision="exact"):
    """transpose12 then advectbyzx as one kernel (bitwise equal to the two
    calls, C/mhd.c:112, :115): u, b <- transpose12(update_u), CT update of
    transpose12(update_b) along its first axis; (nx, ny, nz) are the input
    dims, the output's are (ny, nx, nz).  Needs nx % 16 == 0 and ny % 4 == 0.
    No output may overlap an input."""
    n = nx * ny * nz
    dp, dv = _dt(dt)
    check(lib().mhd_transpose12_advect_async(_dev(u, 5 * n, "u"), _dev(b, 3 * n, "b"),
                                             _dev(update_u, 5 * n, "update_u"),
                                             _dev(update_b, 3 * n, "update_b"), nx, ny, nz, dp, dv,
                                             MHD_PREC_FAST if precision == "fast" else MHD_PREC_EXACT,
                                             _stream()), "transpose12_advect")


def transpose13_advect(u, b, nx, ny, nz, dt, update_u, update_b, precision="exact"):
    """The general transpose13 then advectbyzx as one kernel (C/mhd.c:102,
    :105): out dims (nz, ny, nx).  Needs nx % 16 == 0 and nz % 4 == 0."""
    n = nx * ny * nz
    dp, dv = _dt(dt)
    check(lib().mhd_transpose13_advect_async(_dev(u, 5 * n, "u"), _dev(b, 3 * n, "b"),
                                             _dev(update_u, 5 * n, "update_u"),
                                             _dev(update_b, 3 * n, "update_b"), nx, ny, nz, dp, dv,
                                             MHD_PREC_FAST if precision == "fast" else MHD_PREC_EXACT,
                                             _stream()), "transpose13_advect")


def transpose12(u, b, update_u, update_b, nx, ny, nz, iter=0):
    """x<->y swap of update_* into u, b (C/transpose12.c:7-47)."""
    del iter
    n = nx * ny * nz
    check(lib().mhd_transpose12_async(_dev(u, 5 * n, "u"), _dev(b, 3 * n, "b"),
                                      _dev(update_u, 5 * n, "update_u"),
                                      _dev(update_b, 3 * n, "update_b"), nx, ny, nz, _stream()),
          "transpose12")


def transpose13(u, b, update_u, update_b, nx, ny, nz):
    """x<->z swap of update_* into u, b (C/transpose13.c:9-62); needs nx == nz."""
    n = nx * ny * nz
    check(lib().mhd_transpose13_async(_dev(u, 5 * n, "u"), _dev(b, 3 * n, "b"),
                                      _dev(update_u, 5 * n, "update_u"),
                                      _dev(update_b, 3 * n, "update_b"), nx, ny, nz, _stream()),
          "transpose13")


def transpose13_any(u, b, update_u, update_b, nx, ny, nz):
    """x<->z swap of update_* into u, b for any nx, nz: out dims (nz, ny, nx),
    the component permutation of C/transpose13.c:24-31 (which itself refuses
    nx != nz)."""
    n = nx * ny * nz
    check(lib().mhd_transpose13_any_async(_dev(u, 5 * n, "u"), _dev(b, 3 * n, "b"),
                                          _dev(update_u, 5 * n, "update_u"),
                                          _dev(update_b, 3 * n, "update_b"), nx, ny, nz,
                                          _stream()), "transpose13_any")


def copy_matrix(u, b, nx, ny, nz, update_u, update_b):
    """update <- main (C/copy_matrix.c:9-38)."""
    n = nx * ny * nz
    check(lib().mhd_copy_matrix_async(_dev(u, 5 * n, "u"), _dev(b, 3 * n, "b"), nx, ny, nz,
                                      _dev(update_u, 5 * n, "update_u"),
                                      _dev(update_b, 3 * n, "update_b"), _stream()), "copy_matrix")


_cfl_ws = {}


def _cfl_workspace(nx, ny, nz, device):
    key = (device, nx * ny * nz)
    ws = _cfl_ws.get(key)
    if ws is None:
        nbytes = lib().mhd_calcfl_workspace_bytes(nx, ny, nz)
        ws = torch.zeros(nbytes, dtype=torch.uint8, device=device)
        _cfl_ws[key] = ws
    return ws


def calcfl_device(u, b, nx, ny, nz, out=None):
    """calcfl on the device (C/calcfl.c:13-63): returns a one-element float32
    tensor holding 1/c, without synchronising."""
    n = nx * ny * nz
    if out is None:
        out = torch.empty(1, dtype=torch.float32, device=u.device)
    ws = _cfl_workspace(nx, ny, nz, u.device)
    check(lib().mhd_calcfl_async(out.data_ptr(), _dev(u, 5 * n, "u"), _dev(b, 3 * n, "b"), nx, ny,
                                 nz, ws.data_ptr(), 0.0, None, None, 0.0, _stream()), "calcfl")
    return out


def calcfl(u, b, nx, ny, nz) -> float:
    """calcfl with the reference's return-by-value (synchronises)."""
    return float(calcfl_device(u, b, nx, ny, nz).item())
