"""mhd_calcfl_noalloc_async, the calcfl entry point that allocates nothing
(its workspace is a fixed device array of the library), against
mhd_calcfl_async with a workspace of the caller's, through the C ABI: the
same 1/c, dt and t, bitwise, on two calls in a row, the second on a smaller
box (so it reads nothing the first call left behind), and a third with a
NaN cell (the reference's restart rule).  Boxes above the fixed maximum are
refused before any GPU work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(nx, ny, nz, seed):
    from heterogeneous_mhd_amd import solver
    return solver.ic_host(nx, ny, nz, solver.ICSpec(seed=seed, amp_v=0.3, amp_a=0.1))


def _both(L, u, b, nx, ny, nz, cfl, tf, t0, stream):
    import torch
    dev = torch.device("cuda:0")
    out = {}
    for kind in ("async", "noalloc"):
        inv = torch.full((1,), np.nan, dtype=torch.float32, device=dev)
        dt = torch.full((1,), np.nan, dtype=torch.float32, device=dev)
        t = torch.full((1,), t0, dtype=torch.float32, device=dev)
        if kind == "async":
            ws = torch.zeros(L.mhd_calcfl_workspace_bytes(nx, ny, nz), dtype=torch.uint8, device=dev)
            rc = L.mhd_calcfl_async(inv.data_ptr(), u.data_ptr(), b.data_ptr(), nx, ny, nz, ws.data_ptr(), cfl,
                                    dt.data_ptr(), t.data_ptr(), tf, stream)
        else:
            rc = L.mhd_calcfl_noalloc_async(inv.data_ptr(), u.data_ptr(), b.data_ptr(), nx, ny, nz, cfl,
                                            dt.data_ptr(), t.data_ptr(), tf, stream)
        assert rc == 0, (kind, rc, L.mhd_last_error())
        torch.cuda.synchronize()
        out[kind] = np.concatenate([inv.cpu().numpy(), dt.cpu().numpy(), t.cpu().numpy()])
    return out


def test_calcfl_noalloc_matches_calcfl_async(gpu):
    import torch
    from heterogeneous_mhd_amd import lib
    from oracle import Oracle
    L = lib()
    stream = torch.cuda.current_stream().cuda_stream
    cases = [((192, 160, 128), 3, 0.0, 0.0), ((33, 17, 9), 4, 40.0, 39.5), ((64, 48, 40), 5, 0.0, 1.0)]
    for i, ((nx, ny, nz), seed, tf, t0) in enumerate(cases):
        print(f"noalloc calcfl call {i + 1}: {nx}x{ny}x{nz}", flush=True)
        u, b = _state(nx, ny, nz, seed)
        if i == 2:
            u = u.copy()
            u[5 * (nx * ny * 7 + nx * 3 + 11)] = np.float32(np.nan)  # a NaN cell mid-box
        ud, bd = torch.from_numpy(u).cuda(), torch.from_numpy(b).cuda()
        got = _both(L, ud, bd, nx, ny, nz, 0.7, tf, t0, stream)
        assert got["noalloc"].tobytes() == got["async"].tobytes(), (i, got)
        assert got["noalloc"][:1].tobytes() == Oracle("f32").calcfl(u, b, nx, ny, nz).reshape(1).tobytes()


def test_calcfl_noalloc_refuses_boxes_above_its_scratch(gpu):
    import torch
    from heterogeneous_mhd_amd import lib
    L = lib()
    dummy = torch.empty(4, dtype=torch.float32, device="cuda")
    rc = L.mhd_calcfl_noalloc_async(dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(), 1025, 1024, 1024, 0.7,
                                    None, None, 0.0, torch.cuda.current_stream().cuda_stream)
    assert rc == -7
    assert b"exceed the fixed scratch" in L.mhd_last_error()
