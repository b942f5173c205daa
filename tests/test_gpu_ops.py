"""Parity of every HIP operator with the reference, on the GPU.

The golden fixtures hold the compiled reference's outputs (oracle/_ref);
larger cases compare with the oracle restatement, which the CPU suite pins
bitwise to those fixtures.  Fluid, field, transposes and calcfl are all
checked bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import OPS_CASES, bits_equal, load_ops, sha

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _h(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("case", OPS_CASES)
def test_ops_match_reference_golden(gpu, case):
    import torch
    from heterogeneous_mhd_amd import ops
    d = load_ops(case)
    nx, ny, nz = d["dims"]
    print(f"case {case} dims {nx}x{ny}x{nz}", flush=True)
    u, b = _t(d["u"], gpu), _t(d["b"], gpu)
    out = torch.empty_like(u)
    ops.fluidx(out, b, nx, ny, nz, float(d["dt"]), u, b)
    assert bits_equal(_h(out), d["fluidx_u"])
    bo = b.clone()
    ops.advectbyzx(u, bo, nx, ny, nz, float(d["dt"]), u, b)
    assert bits_equal(_h(bo), d["advectbyzx_b"])
    tu, tb = torch.empty_like(u), torch.empty_like(b)
    ops.transpose12(tu, tb, u, b, nx, ny, nz)
    assert sha(_h(tu)) == d["digests"]["transpose12_u"]
    assert sha(_h(tb)) == d["digests"]["transpose12_b"]
    if nx == nz:
        ops.transpose13(tu, tb, u, b, nx, ny, nz)
        assert sha(_h(tu)) == d["digests"]["transpose13_u"]
        assert sha(_h(tb)) == d["digests"]["transpose13_b"]
    c = ops.calcfl(u, b, nx, ny, nz)
    assert np.float32(c).tobytes() == d["calcfl"].tobytes()


@pytest.mark.parametrize("case", ["16x16x16", "rough_12x10x10", "7x3x7"])
def test_reference_abi_host_pointers(gpu, case):
    """Part 1 of the C ABI with plain host arrays, as a reference C driver
    would call it: staged through device memory, results bitwise equal."""
    from heterogeneous_mhd_amd import lib
    L = lib()
    d = load_ops(case)
    nx, ny, nz = d["dims"]
    NX, NY, NZ = C.c_int(nx), C.c_int(ny), C.c_int(nz)
    dt = C.c_float(float(d["dt"]))
    uu, bu = d["u"].copy(), d["b"].copy()
    u = np.zeros_like(uu)
    b = bu.copy()
    p = lambda a: a.ctypes.data  # noqa: E731
    L.fluidx(p(u), p(b), C.byref(NX), C.byref(NY), C.byref(NZ), C.byref(dt), p(uu), p(bu), 1)
    assert bits_equal(u, d["fluidx_u"])
    # in place, as C/mhd.c does after copy_matrix (u aliases update_u)
    u2 = uu.copy()
    L.fluidx(p(u2), p(b), C.byref(NX), C.byref(NY), C.byref(NZ), C.byref(dt), p(u2), p(bu), 1)
    assert bits_equal(u2, d["fluidx_u"])
    # advectbyzx with b aliasing update_b (tolerated by the serial reference)
    b2 = bu.copy()
    L.advectbyzx(p(uu), p(b2), C.byref(NX), C.byref(NY), C.byref(NZ), C.byref(dt), p(uu), p(b2))
    assert bits_equal(b2, d["advectbyzx_b"])
    tu, tb = np.zeros_like(uu), np.zeros_like(bu)
    L.transpose12(p(tu), p(tb), p(uu), p(bu), C.byref(NX), C.byref(NY), C.byref(NZ), 1)
    assert sha(tu) == d["digests"]["transpose12_u"] and sha(tb) == d["digests"]["transpose12_b"]
    if nx == nz:
        L.transpose13(p(tu), p(tb), p(uu), p(bu), C.byref(NX), C.byref(NY), C.byref(NZ))
        assert sha(tu) == d["digests"]["transpose13_u"]
    c = L.calcfl(p(uu), p(bu), C.byref(NX), C.byref(NY), C.byref(NZ))
    assert np.float32(c).tobytes() == d["calcfl"].tobytes()
    cu, cb = np.zeros_like(uu), np.zeros_like(bu)
    L.copy_matrix(p(uu), p(bu), C.byref(NX), C.byref(NY), C.byref(NZ), p(cu), p(cb))
    assert bits_equal(cu, uu) and bits_equal(cb, bu)


def test_transpose13_rejects_nx_ne_nz(gpu, capfd):
    from heterogeneous_mhd_amd import lib, ops, MHDError
    import torch
    u = torch.zeros(5 * 4 * 3 * 5, device=gpu)
    b = torch.zeros(3 * 4 * 3 * 5, device=gpu)
    with pytest.raises(MHDError):
        ops.transpose13(torch.empty_like(u), torch.empty_like(b), u, b, 4, 3, 5)
    L = lib()
    uh = np.ones(60 * 5, np.float32)
    bh = np.ones(60 * 3, np.float32)
    uo, bo = np.zeros_like(uh), np.zeros_like(bh)
    nx, ny, nz = C.c_int(4), C.c_int(3), C.c_int(5)
    L.transpose13(uo.ctypes.data, bo.ctypes.data, uh.ctypes.data, bh.ctypes.data, C.byref(nx),
                  C.byref(ny), C.byref(nz))
    assert "nz<>nx not supported" in capfd.readouterr().out
    assert not uo.any() and not bo.any()  # nothing written, as in the reference


def _random_state(nx, ny, nz, seed, amp_v=0.3, amp_a=0.1):
    from heterogeneous_mhd_amd import solver
    return solver.ic_host(nx, ny, nz, solver.ICSpec(seed=seed, amp_v=amp_v, amp_a=amp_a))


# shapes covering every launch configuration of the sweeps: 1/2/4 cells per
# thread, 1-4 pencils per workgroup, 512-thread groups for nx > 1024, and
# one-pencil workgroups with several rows of pencils
SHAPES = [(64, 8, 8), (96, 6, 6), (128, 8, 8), (200, 4, 4), (256, 4, 4), (384, 3, 3), (512, 2, 2),
          (1000, 2, 2), (1024, 2, 2), (1536, 2, 2), (2048, 1, 1), (33, 17, 17), (512, 4, 4),
          (1024, 8, 4)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_sweeps_match_oracle_all_launch_shapes(gpu, shape):
    import torch
    from heterogeneous_mhd_amd import ops
    from oracle import Oracle
    nx, ny, nz = shape
    print(f"shape {nx}x{ny}x{nz}", flush=True)
    u, b = _random_state(nx, ny, nz, seed=nx + ny)
    O = Oracle("f32")
    dt = np.float32(0.05)
    ud, bd = _t(u, gpu), _t(b, gpu)
    out = torch.empty_like(ud)
    ops.fluidx(out, bd, nx, ny, nz, float(dt), ud, bd)
    assert bits_equal(_h(out), O.fluidx(u, b, nx, ny, nz, dt))
    bo = bd.clone()
    ops.advectbyzx(ud, bo, nx, ny, nz, float(dt), ud, bd)
    assert bits_equal(_h(bo), O.advectbyzx(u, b, nx, ny, nz, dt))
    assert np.float32(ops.calcfl(ud, bd, nx, ny, nz)).tobytes() == O.calcfl(u, b, nx, ny, nz).tobytes()


def test_fluidx_in_place_and_device_dt(gpu):
    import torch
    from heterogeneous_mhd_amd import ops
    from oracle import Oracle
    nx, ny, nz = 64, 32, 32
    u, b = _random_state(nx, ny, nz, seed=3)
    ud, bd = _t(u, gpu), _t(b, gpu)
    dt = torch.tensor([0.07], dtype=torch.float32, device=gpu)
    ops.fluidx(ud, bd, nx, ny, nz, dt, ud, bd)  # in place, dt read on the device
    assert bits_equal(_h(ud), Oracle("f32").fluidx(u, b, nx, ny, nz, np.float32(0.07)))


def test_transposes_large_match_permutation(gpu):
    import torch
    from heterogeneous_mhd_amd import ops
    nx, ny, nz = 100, 70, 100
    rng = np.random.default_rng(0)
    u = rng.standard_normal(5 * nx * ny * nz).astype(np.float32)
    b = rng.standard_normal(3 * nx * ny * nz).astype(np.float32)
    ud, bd = _t(u, gpu), _t(b, gpu)
    tu, tb = torch.empty_like(ud), torch.empty_like(bd)
    ops.transpose12(tu, tb, ud, bd, nx, ny, nz)
    U = u.reshape(nz, ny, nx, 5)
    B = b.reshape(nz, ny, nx, 3)
    assert bits_equal(_h(tu).reshape(nz, nx, ny, 5), U.transpose(0, 2, 1, 3)[..., [0, 2, 1, 3, 4]])
    assert bits_equal(_h(tb).reshape(nz, nx, ny, 3), B.transpose(0, 2, 1, 3)[..., [1, 0, 2]])
    ops.transpose13(tu, tb, ud, bd, nx, ny, nz)
    assert bits_equal(_h(tu).reshape(nx, ny, nz, 5), U.transpose(2, 1, 0, 3)[..., [0, 3, 2, 1, 4]])
    assert bits_equal(_h(tb).reshape(nx, ny, nz, 3), B.transpose(2, 1, 0, 3)[..., [2, 1, 0]])


def test_transposes_many_planes(gpu):
    """More than 32767 planes: the one launch for u and b then runs its
    planes on the grid's x dimension (the only one past 65535 workgroups);
    the result is still exactly the permutation."""
    import torch
    from heterogeneous_mhd_amd import ops
    rng = np.random.default_rng(3)
    for nx, ny, nz in ((2, 3, 40000), (3, 40000, 3)):
        print(f"transposes, {nx}x{ny}x{nz}", flush=True)
        u = rng.standard_normal(5 * nx * ny * nz).astype(np.float32)
        b = rng.standard_normal(3 * nx * ny * nz).astype(np.float32)
        ud, bd = _t(u, gpu), _t(b, gpu)
        tu, tb = torch.empty_like(ud), torch.empty_like(bd)
        U = u.reshape(nz, ny, nx, 5)
        B = b.reshape(nz, ny, nx, 3)
        ops.transpose12(tu, tb, ud, bd, nx, ny, nz)
        assert bits_equal(_h(tu).reshape(nz, nx, ny, 5), U.transpose(0, 2, 1, 3)[..., [0, 2, 1, 3, 4]])
        assert bits_equal(_h(tb).reshape(nz, nx, ny, 3), B.transpose(0, 2, 1, 3)[..., [1, 0, 2]])
        ops.transpose13_any(tu, tb, ud, bd, nx, ny, nz)  # out dims (nz, ny, nx)
        assert bits_equal(_h(tu).reshape(nx, ny, nz, 5), U.transpose(2, 1, 0, 3)[..., [0, 3, 2, 1, 4]])
        assert bits_equal(_h(tb).reshape(nx, ny, nz, 3), B.transpose(2, 1, 0, 3)[..., [2, 1, 0]])


def test_calcfl_nan_semantics(gpu):
    """The running max_2num of C/calcfl.c:55 restarts after a NaN cell; the
    device reduction reproduces the serial result, NaN positions included."""
    import torch
    from heterogeneous_mhd_amd import ops
    from oracle import Oracle
    O = Oracle("f32")
    nx, ny, nz = 64, 64, 8
    u, b = _random_state(nx, ny, nz, seed=21)
    n = nx * ny * nz
    for pos in ([5], [100, 30000], [n - 1], [n - 3000, 17], [2047, 2048], []):
        uu = u.copy()
        for p_ in pos:
            uu[5 * p_] = np.nan  # rho = NaN makes that cell's bound NaN
        got = ops.calcfl(_t(uu, gpu), _t(b, gpu), nx, ny, nz)
        want = O.calcfl(uu, b, nx, ny, nz)
        assert np.float32(got).tobytes() == want.tobytes(), pos


@pytest.mark.parametrize("case", OPS_CASES)
@pytest.mark.parametrize("precision", ["exact", "fast"])
def test_fused_advect_fluidx_equals_the_two_calls(gpu, case, precision):
    """advect_fluidx is bitwise advectbyzx followed by fluidx (C/mhd.c:95-97)."""
    import torch
    from heterogeneous_mhd_amd import ops
    d = load_ops(case)
    nx, ny, nz = d["dims"]
    dt = float(d["dt"])
    print(f"fused {case} {precision}", flush=True)
    u, b = _t(d["u"], gpu), _t(d["b"], gpu)
    b1 = b.clone()
    ops.advectbyzx(u, b1, nx, ny, nz, dt, u, b, precision=precision)
    u1 = torch.empty_like(u)
    ops.fluidx(u1, b1, nx, ny, nz, dt, u, b1, precision=precision)
    uf, bf = torch.empty_like(u), torch.empty_like(b)
    ops.advect_fluidx(uf, bf, nx, ny, nz, dt, u, b, precision=precision)
    assert bits_equal(_h(bf), _h(b1))
    assert bits_equal(_h(uf), _h(u1))
    if precision == "exact":
        assert bits_equal(_h(bf), d["advectbyzx_b"])


# launch shapes of the fused kernels, including one pencil per 256-, 512- and
# 1024-thread workgroup (nx = 512, 1024, 2048)
FUSED_SHAPES = [(64, 8, 8), (128, 8, 8), (200, 4, 4), (512, 4, 3), (1024, 3, 4), (2048, 2, 1),
                (33, 17, 17)]


@pytest.mark.parametrize("shape", FUSED_SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("precision", ["exact", "fast"])
def test_advect2_fluidx_equals_the_three_calls(gpu, shape, precision):
    """advect2_fluidx is bitwise advectbyzx, advectbyzx with the same u, then
    fluidx (C/mhd.c:93, :95-97); in exact mode also the oracle's three calls."""
    import torch
    from heterogeneous_mhd_amd import ops
    from oracle import Oracle
    nx, ny, nz = shape
    print(f"advect2_fluidx {nx}x{ny}x{nz} {precision}", flush=True)
    u, b = _random_state(nx, ny, nz, seed=nx + 2 * ny)
    dt = np.float32(0.05)
    ud, bd = _t(u, gpu), _t(b, gpu)
    b1 = bd.clone()
    ops.advectbyzx(ud, b1, nx, ny, nz, float(dt), ud, bd, precision=precision)
    b2 = b1.clone()
    ops.advectbyzx(ud, b2, nx, ny, nz, float(dt), ud, b1, precision=precision)
    u1 = torch.empty_like(ud)
    ops.fluidx(u1, b2, nx, ny, nz, float(dt), ud, b2, precision=precision)
    uf, bf = torch.empty_like(ud), torch.empty_like(bd)
    ops.advect2_fluidx(uf, bf, nx, ny, nz, float(dt), ud, bd, precision=precision)
    assert bits_equal(_h(bf), _h(b2))
    assert bits_equal(_h(uf), _h(u1))
    if precision == "exact":
        O = Oracle("f32")
        ob = O.advectbyzx(u, O.advectbyzx(u, b, nx, ny, nz, dt), nx, ny, nz, dt)
        assert bits_equal(_h(bf), ob)
        assert bits_equal(_h(uf), O.fluidx(u, ob, nx, ny, nz, dt))


# tiles of 56 cells x 16 pencils: full and partial tiles along both axes,
# pencils shorter than the 64-cell window (several wraps), partner extents
# below 16, a remaining axis of extent 1 (the walk's look-ahead row wraps
# twice), non-cubic boxes
AT_SHAPES = [(16, 16, 16), (24, 16, 16), (12, 10, 10), (7, 3, 7), (1, 4, 4), (3, 1, 2), (2, 1, 1), (5, 2, 1),
             (64, 48, 40),
             (100, 33, 17), (57, 16, 20), (130, 20, 3), (56, 32, 16), (112, 17, 5), (1024, 32, 3),
             (256, 5, 33)]


@pytest.mark.parametrize("shape", AT_SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("which", [12, 13])
@pytest.mark.parametrize("precision", ["exact", "fast"])
def test_advect_transpose_equals_the_two_calls(gpu, shape, which, precision):
    """advect_transpose12/13 is bitwise advectbyzx followed by transpose12
    (resp. the general transpose13): C/mhd.c:74, :78 and :84, :88; in exact
    mode also the oracle's two calls."""
    import torch
    from heterogeneous_mhd_amd import ops
    from oracle import Oracle
    nx, ny, nz = shape
    print(f"advect_transpose{which} {nx}x{ny}x{nz} {precision}", flush=True)
    u, b = _random_state(nx, ny, nz, seed=nx + 3 * ny + nz)
    dt = np.float32(0.05)
    ud, bd = _t(u, gpu), _t(b, gpu)
    b1 = bd.clone()
    ops.advectbyzx(ud, b1, nx, ny, nz, float(dt), ud, bd, precision=precision)
    tu, tb = torch.empty_like(ud), torch.empty_like(bd)
    if which == 12:
        ops.transpose12(tu, tb, ud, b1, nx, ny, nz)
    else:
        ops.transpose13_any(tu, tb, ud, b1, nx, ny, nz)
    fu, fb = torch.full_like(ud, np.nan), torch.full_like(bd, np.nan)
    getattr(ops, f"advect_transpose{which}")(fu, fb, nx, ny, nz, float(dt), ud, bd, precision=precision)
    assert bits_equal(_h(fb), _h(tb))
    assert bits_equal(_h(fu), _h(tu))
    if precision == "exact":
        O = Oracle("f32")
        ob = O.advectbyzx(u, b, nx, ny, nz, dt)
        ru, rb = O.transpose12(u, ob, nx, ny, nz) if which == 12 else O.transpose13_any(u, ob, nx, ny, nz)
        assert bits_equal(_h(fu), ru) and bits_equal(_h(fb), rb)


# input dims; the fused form needs nx % 16 == 0 and the output's first dim (ny
# for transpose12, nz for transpose13) a multiple of 4: full and partial
# tiles along the sweep axis, sweep axes shorter than the 64-cell window (4:
# many wraps), partner extents of one and several pencil blocks, a remaining
# axis of extent 1 (the look-ahead row wraps twice), non-cubic boxes
TA_SHAPES = [(16, 16, 16), (16, 24, 16), (32, 8, 4), (16, 4, 8), (48, 64, 20), (16, 8, 1), (16, 1, 8),
             (16, 100, 12), (64, 56, 56), (32, 4, 4), (16, 12, 100), (1024, 8, 4), (16, 20, 16)]
TA_CASES = [(s, w) for s in TA_SHAPES for w in (12, 13) if (s[1] if w == 12 else s[2]) % 4 == 0]


@pytest.mark.parametrize("shape,which", TA_CASES, ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple)
                         else str(v))
@pytest.mark.parametrize("precision", ["exact", "fast"])
def test_transpose_advect_equals_the_two_calls(gpu, shape, which, precision):
    """transpose12_advect / transpose13_advect is bitwise transpose12 (resp.
    the general transpose13) followed by advectbyzx of the transposed state:
    C/mhd.c:112, :115 and :102, :105; in exact mode also the oracle's two
    calls."""
    import torch
    from heterogeneous_mhd_amd import ops
    from oracle import Oracle
    nx, ny, nz = shape
    od = (ny, nx, nz) if which == 12 else (nz, ny, nx)
    print(f"transpose{which}_advect {nx}x{ny}x{nz} {precision}", flush=True)
    u, b = _random_state(nx, ny, nz, seed=2 * nx + ny + 5 * nz)
    dt = np.float32(0.05)
    ud, bd = _t(u, gpu), _t(b, gpu)
    tu, tb = torch.empty_like(ud), torch.empty_like(bd)
    if which == 12:
        ops.transpose12(tu, tb, ud, bd, nx, ny, nz)
    else:
        ops.transpose13_any(tu, tb, ud, bd, nx, ny, nz)
    b1 = tb.clone()
    ops.advectbyzx(tu, b1, *od, float(dt), tu, tb, precision=precision)
    fu, fb = torch.full_like(ud, np.nan), torch.full_like(bd, np.nan)
    getattr(ops, f"transpose{which}_advect")(fu, fb, nx, ny, nz, float(dt), ud, bd, precision=precision)
    assert bits_equal(_h(fu), _h(tu))
    assert bits_equal(_h(fb), _h(b1))
    if precision == "exact":
        O = Oracle("f32")
        ru, rb = O.transpose12(u, b, nx, ny, nz) if which == 12 else O.transpose13_any(u, b, nx, ny, nz)
        assert bits_equal(_h(fu), ru)
        assert bits_equal(_h(fb), O.advectbyzx(ru, rb, *od, dt))


def test_transpose_advect_rejects_unsupported_shapes_and_overlap(gpu):
    import torch
    from heterogeneous_mhd_amd import MHDError, ops
    for shape, which in [((8, 8, 8), 12), ((24, 16, 16), 13), ((16, 6, 16), 12), ((16, 16, 6), 13)]:
        nx, ny, nz = shape
        u, b = _random_state(nx, ny, nz, seed=3)
        ud, bd = _t(u, gpu), _t(b, gpu)
        with pytest.raises(MHDError):
            getattr(ops, f"transpose{which}_advect")(torch.empty_like(ud), torch.empty_like(bd), nx, ny, nz,
                                                     0.05, ud, bd)
    n = 16
    u, b = _random_state(n, n, n, seed=1)
    ud, bd = _t(u, gpu), _t(b, gpu)
    with pytest.raises(MHDError):
        ops.transpose12_advect(ud, torch.empty_like(bd), n, n, n, 0.05, ud, bd)
    with pytest.raises(MHDError):
        ops.transpose13_advect(torch.empty_like(ud), bd, n, n, n, 0.05, ud, bd)


def test_advect_transpose_rejects_overlap(gpu):
    import torch
    from heterogeneous_mhd_amd import MHDError, ops
    n = 8
    u, b = _random_state(n, n, n, seed=1)
    ud, bd = _t(u, gpu), _t(b, gpu)
    with pytest.raises(MHDError):
        ops.advect_transpose12(ud, torch.empty_like(bd), n, n, n, 0.05, ud, bd)
    with pytest.raises(MHDError):
        ops.advect_transpose13(torch.empty_like(ud), bd, n, n, n, 0.05, ud, bd)
