"""Learned coarse-grid corrector: a 4-layer periodic 3x3x3 convolution network
(32 channels, GELU) whose output corrects the plain 64^3 step toward the
4x-downsampled 256^3 solution.  Training and inference run on the GPU
through the library's own BF16 MFMA kernels (include/mhd_hip.h, Part 3);
torch only holds device memory and does the data bookkeeping.

Correction after each coarse step, with y the network output on the stepped
state:  u_c += alpha s_c (y_c - <y_c>) (c < 5; the box mean is removed
because the scheme and the fine solution conserve the totals of rho, m and
E), b += alpha s_B curl(y_5..7) -- the field is corrected through a vector
potential, so div b stays at round-off.  The relaxation alpha (and an
optional limit on the output values) is chosen on validation seeds by
`select` and stored with the weights.  The
correction is skipped (the plain step is kept) when the corrected state has
rho <= 0 or p <= 0 anywhere, or when it converts more than etol of the box's
total energy between heat and kinetic plus magnetic energy (the total of E
itself cannot change: the correction of u has zero box total).

Training pairs come from the solver itself: a 256^3 run of a seeded IC is
downsampled by 4 (block means of u, face means of b); one plain coarse step
from that state is compared with the downsampled fine state at the same
time (fine time = 4 x coarse time, since the scheme has unit cell size).

    python -m heterogeneous_mhd_amd.corrector train --out weights.npz [...]
    python -m heterogeneous_mhd_amd.corrector select --weights weights.npz [...]
    python -m heterogeneous_mhd_amd.corrector evaluate --weights weights.npz [...]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time
from pathlib import Path

import numpy as np
import torch

from . import solver
from ._lib import MHDError, check, lib

LAYERS = 4
CH = 32
WEIGHTS = Path(__file__).resolve().parent / "data" / "corrector_64.npz"


def _st() -> int:
    return torch.cuda.current_stream().cuda_stream


def _p(t: torch.Tensor) -> int:
    return t.data_ptr()


class Buffers:
    """Activation and gradient buffers for one grid shape."""

    def __init__(self, nx: int, ny: int, nz: int, device, train: bool):
        self.shape = (nx, ny, nz)
        v = nx * ny * nz
        if nx % 64 or v % 256:
            raise MHDError("the corrector needs nx a multiple of 64 and nx*ny*nz of 256")
        f32, i16 = torch.float32, torch.int16  # int16 holds bf16 bits
        self.x0 = torch.empty(v * CH, dtype=i16, device=device)
        # hidden pre-activations are kept only for a backward pass
        self.z = [torch.empty(v * CH, dtype=f32, device=device) if train or li == LAYERS - 1 else None
                  for li in range(LAYERS)]
        self.a = [torch.empty(v * CH, dtype=i16, device=device) for _ in range(LAYERS - 1)]
        if train:
            self.d = [torch.empty(v * CH, dtype=i16, device=device) for _ in range(2)]
            self.r = torch.empty(v * 8, dtype=f32, device=device)
            self.lossp = torch.empty(lib().mhd_corr_loss_partials(nx, ny, nz), dtype=torch.float64,
                                     device=device)
            nb = lib().mhd_corr_dw_workspace_bytes() // 4
            self.ws = [torch.empty(nb, dtype=f32, device=device) for _ in range(LAYERS)]
        self.u_cand = torch.empty(v * 5, dtype=f32, device=device)
        self.b_cand = torch.empty(v * 3, dtype=f32, device=device)
        self.check = torch.empty(lib().mhd_corr_check_bytes(nx, ny, nz), dtype=torch.uint8,
                                 device=device)


# network input channels per input kind (mhd_corr_pack_features): 0 the 8
# normalised fields; 1 plus their Laplacians; 2 plus their second differences
# along x, y and z; 3 plus the low parts their bf16 rounding drops
INPUT_CHANNELS = {0: 8, 1: 16, 2: 32, 3: 16}
DERIVED = {0: 0, 1: 8, 2: 24, 3: 0}


class CorrectorNet:
    """Parameters, optimizer state and the forward/backward/apply passes."""

    def __init__(self, device="cuda", seed: int = 0, inputs: int = 0):
        self.device = torch.device(device)
        if inputs not in INPUT_CHANNELS:
            raise MHDError(f"unknown input kind {inputs}")
        self.inputs = inputs
        cin = INPUT_CHANNELS[inputs]
        self.np = lib().mhd_corr_num_params()
        nw = self.np - CH
        rng = np.random.default_rng(seed)
        p = np.zeros((LAYERS, self.np), np.float32)
        for li in range(LAYERS):
            fan_in = 27 * (cin if li == 0 else CH)
            std = np.sqrt(2.0 / fan_in) if li < LAYERS - 1 else 0.01 / np.sqrt(fan_in)
            w = rng.standard_normal((27, CH, CH)) * std
            if li == 0:
                w[:, :, cin:] = 0.0        # padded input channels
            if li == LAYERS - 1:
                w[:, 8:, :] = 0.0          # padded output channels
            p[li, :nw] = w.reshape(-1)
        self.params = torch.from_numpy(p).to(self.device)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step_count = 0
        self.wf = torch.empty((LAYERS, nw), dtype=torch.int16, device=self.device)
        self.wb = torch.empty_like(self.wf)
        # the first layer's 8-channel input in the packed form (4 taps per K-step)
        self.wf0 = torch.empty(lib().mhd_corr_first_weight_elems(), dtype=torch.int16,
                               device=self.device)
        # input normalisation (means, inverse std devs, inverse std devs of
        # the derived channels) and output scales
        self.norm = torch.tensor([1, 0, 0, 0, 1, 0, 0, 0] + [1] * 32, dtype=torch.float32,
                                 device=self.device)
        self.scale = torch.ones(6, dtype=torch.float32, device=self.device)
        self.meta: dict = {"inputs": inputs}
        self.prep()

    # -- parameters ----------------------------------------------------------
    def prep(self):
        for li in range(LAYERS):
            check(lib().mhd_corr_weight_prep(_p(self.params[li]), _p(self.wf[li]), _p(self.wb[li]),
                                             _st()), "corr_weight_prep")
        check(lib().mhd_corr_weight_prep_first(_p(self.params[0]), _p(self.wf0), _st()),
              "corr_weight_prep_first")

    def bias(self, li: int) -> torch.Tensor:
        return self.params[li, self.np - CH:]

    def save(self, path):
        np.savez(path, params=self.params.cpu().numpy(), norm=self.norm.cpu().numpy(),
                 scale=self.scale.cpu().numpy(), meta=json.dumps(self.meta))

    @classmethod
    def load(cls, path=WEIGHTS, device="cuda"):
        z = np.load(path, allow_pickle=False)
        meta = json.loads(str(z["meta"]))
        net = cls(device, inputs=int(meta.get("inputs", 0)))
        net.params.copy_(torch.from_numpy(z["params"]))
        norm = torch.from_numpy(z["norm"])
        net.norm[:norm.numel()].copy_(norm)
        net.scale.copy_(torch.from_numpy(z["scale"]))
        net.meta = meta
        net.meta["inputs"] = net.inputs
        net.prep()
        return net

    # -- passes ----------------------------------------------------------------
    def forward(self, buf: Buffers, u: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
        """Network output (FP32 [V][32]) for state (u, b)."""
        nx, ny, nz = buf.shape
        L, s = lib(), _st()
        if self.inputs == 0:
            check(L.mhd_corr_pack(_p(u), _p(b), nx, ny, nz, _p(self.norm), _p(buf.x0), s), "corr_pack")
        else:
            check(L.mhd_corr_pack_features(_p(u), _p(b), nx, ny, nz, _p(self.norm), self.inputs,
                                           _p(buf.x0), s), "corr_pack_features")
        a = buf.x0
        first = (3, self.wf0) if self.inputs == 0 else (0, self.wf[0])  # 8 channels: packed taps
        for li in range(LAYERS):
            last = li == LAYERS - 1
            mode, w = (1, self.wf[li]) if last else first if li == 0 else (0, self.wf[li])
            check(L.mhd_corr_conv(_p(a), _p(w), _p(self.bias(li)), nx, ny, nz, mode,
                                  None if buf.z[li] is None else _p(buf.z[li]),
                                  None if last else _p(buf.a[li]),
                                  None, s), "corr_conv")
            if not last:
                a = buf.a[li]
        return buf.z[LAYERS - 1]

    def backward(self, buf: Buffers, target: torch.Tensor, accumulate: bool) -> torch.Tensor:
        """Loss of the last forward against target ([V][8]); adds the weight
        gradients to the per-layer workspaces.  Returns the loss (device)."""
        nx, ny, nz = buf.shape
        L, s = lib(), _st()
        check(L.mhd_corr_loss_grad(_p(buf.z[LAYERS - 1]), _p(target), nx, ny, nz, _p(buf.r),
                                   _p(buf.lossp), _p(buf.d[0]), s), "corr_loss_grad")
        d, nxt = buf.d[0], buf.d[1]
        for li in range(LAYERS - 1, -1, -1):
            a_in = buf.x0 if li == 0 else buf.a[li - 1]
            check(L.mhd_corr_conv_dw(_p(d), _p(a_in), nx, ny, nz, _p(buf.ws[li]), int(accumulate), s),
                  "corr_conv_dw")
            if li > 0:
                check(L.mhd_corr_conv(_p(d), _p(self.wb[li]), None, nx, ny, nz, 2, None, _p(nxt),
                                      _p(buf.z[li - 1]), s), "corr_conv_grad")
                d, nxt = nxt, d
        return buf.lossp.sum() / (nx * ny * nz)

    def gradients(self, buf: Buffers) -> torch.Tensor:
        """Folds the workspaces into self.grads (adds) and returns them."""
        for li in range(LAYERS):
            check(lib().mhd_corr_dw_reduce(_p(buf.ws[li]), _p(self.grads[li]), _st()), "corr_dw_reduce")
        return self.grads

    def optimizer_step(self, buf: Buffers, batch: int, lr: float):
        self.gradients(buf)
        self.step_count += 1
        check(lib().mhd_corr_adam(_p(self.params), _p(self.grads), _p(self.m), _p(self.v),
                                  self.params.numel(), lr, 0.9, 0.999, 1e-8, self.step_count,
                                  1.0 / batch, _st()), "corr_adam")
        self.prep()

    def apply(self, buf: Buffers, u, b, y, etol: float, flag_stats: torch.Tensor, scale=None,
              limit: float = 0.0):
        nx, ny, nz = buf.shape
        sc = self.scale if scale is None else scale
        check(lib().mhd_corr_apply(_p(u), _p(b), _p(y), nx, ny, nz, _p(sc), float(limit), float(etol),
                                   _p(buf.u_cand), _p(buf.b_cand), _p(buf.check), _p(flag_stats),
                                   _st()), "corr_apply")


# ---------------------------------------------------------------------------
# data
# ---------------------------------------------------------------------------

def downsample(u, b, n: int, r: int):
    """Block means by r of an n^3 state (device tensors) -> (uc, bc)."""
    m = (n // r) ** 3
    uc = torch.empty(5 * m, dtype=torch.float32, device=u.device)
    bc = torch.empty(3 * m, dtype=torch.float32, device=u.device)
    check(lib().mhd_downsample(_p(u), _p(b), n, n, n, r, _p(uc), _p(bc), _st()), "downsample")
    return uc, bc


def features(u, b, n):
    """[V][8] rho, m, E, cell-centred B (torch, for the normalisation stats)."""
    ug = u.view(n, n, n, 5)
    bg = b.view(n, n, n, 3)
    bc = 0.5 * (bg + torch.stack([bg[..., 0].roll(-1, 2), bg[..., 1].roll(-1, 1),
                                  bg[..., 2].roll(-1, 0)], dim=-1))
    return torch.cat([ug, bc], dim=-1).view(-1, 8)


def make_pairs(seeds, pairs_per_seed: int, fine: int = 256, r: int = 4, skip=(0.0, 0.0),
               amp_v=(0.1, 0.1), amp_a=(0.05, 0.05), seed: int = 11, log=print):
    """Training pairs (coarse-stepped u, b, raw correction [V][8]) from fine
    runs the solver generates.  skip: range of the coarse time a run reaches
    before its first pair (drawn uniformly per seed)."""
    rng = np.random.default_rng(seed)
    n = fine // r
    pairs = []
    fbox = solver.Box(fine, fine, fine, cfl=0.7)
    cbox = solver.Box(n, n, n, cfl=0.7)
    t0 = time.perf_counter()
    for si, s in enumerate(seeds):
        spec = solver.ICSpec(seed=int(s), amp_v=float(rng.uniform(*amp_v)),
                             amp_a=float(rng.uniform(*amp_a)))
        fbox.init(spec)
        tc = float(rng.uniform(*skip))
        if tc > 0.0:
            fbox.advance_to(r * tc)
        for k in range(pairs_per_seed):
            with torch.cuda.stream(fbox.stream):
                uc, bc = downsample(fbox.u, fbox.b, fine, r)
            fbox.stream.synchronize()
            cbox.u.copy_(uc)
            cbox.b.copy_(bc)
            cbox.set_time(tc)
            cbox.run(1)
            tc_new = cbox.time()[0]
            fbox.advance_to(r * tc_new)
            with torch.cuda.stream(fbox.stream):
                uf, bf = downsample(fbox.u, fbox.b, fine, r)
            fbox.stream.synchronize()
            cbox.sync()
            du = (uf - cbox.u).view(-1, 5)
            db = (bf - cbox.b).view(-1, 3)
            pairs.append((cbox.u.clone(), cbox.b.clone(), torch.cat([du, db], dim=1).contiguous()))
            tc = tc_new
        if log and (si + 1) % 10 == 0:
            log(f"[corrector] pairs from {si + 1}/{len(seeds)} seeds, {time.perf_counter() - t0:.1f} s")
    fbox.close()
    cbox.close()
    return pairs


def make_onpolicy_pairs(net: CorrectorNet, seeds, rollout: int, every: int, fine: int = 256,
                        r: int = 4, skip=(0.0, 0.0), amp_v=(0.1, 0.1), amp_a=(0.05, 0.05),
                        seed: int = 13, log=print):
    """Pairs along corrected rollouts: from the downsampled fine state at a
    random start time, the corrected coarse run takes `rollout` steps while
    the fine run tracks its times; every `every`-th step records (plain step
    of the corrected state, downsampled fine state - that step).  These teach
    the network to pull a drifted state back toward the fine solution."""
    rng = np.random.default_rng(seed)
    n = fine // r
    pairs = []
    fbox = solver.Box(fine, fine, fine, cfl=0.7)
    cb = CorrectedBox(net, n)
    bx = cb.box
    t0 = time.perf_counter()
    for si, s in enumerate(seeds):
        spec = solver.ICSpec(seed=int(s), amp_v=float(rng.uniform(*amp_v)),
                             amp_a=float(rng.uniform(*amp_a)))
        fbox.init(spec)
        tc = float(rng.uniform(*skip))
        if tc > 0.0:
            fbox.advance_to(r * tc)
        with torch.cuda.stream(fbox.stream):
            uc, bc = downsample(fbox.u, fbox.b, fine, r)
        fbox.stream.synchronize()
        bx.u.copy_(uc)
        bx.b.copy_(bc)
        bx.set_time(tc)
        for k in range(rollout):
            bx.run(1)
            tn = bx.time()[0]
            fbox.advance_to(r * tn)
            if k % every == 0:
                with torch.cuda.stream(fbox.stream):
                    uf, bf = downsample(fbox.u, fbox.b, fine, r)
                fbox.stream.synchronize()
                du = (uf - bx.u).view(-1, 5)
                db = (bf - bx.b).view(-1, 3)
                pairs.append((bx.u.clone(), bx.b.clone(), torch.cat([du, db], dim=1).contiguous()))
            with torch.cuda.stream(bx.stream):
                y = net.forward(cb.buf, bx.u, bx.b)
                net.apply(cb.buf, bx.u, bx.b, y, cb.etol, cb.stats, cb.scale, cb.limit)
            bx.state_changed()
        if log and (si + 1) % 10 == 0:
            log(f"[corrector] on-policy pairs from {si + 1}/{len(seeds)} seeds, "
                f"{time.perf_counter() - t0:.1f} s, fallbacks {cb.fallbacks()}")
    fbox.close()
    bx.close()
    return pairs


def second_differences(f, n):
    """[V][8] normalised fields -> [V][8][3]: (f(-e_d) + f(+e_d)) - 2 f along
    d = x, y, z (periodic), as mhd_corr_pack_features forms them."""
    g = f.view(n, n, n, 8)
    return torch.stack([(g.roll(1, a) + g.roll(-1, a)) - 2.0 * g for a in (2, 1, 0)],
                       dim=-1).view(-1, 8, 3)


def set_normalisation(net: CorrectorNet, pairs, n: int):
    f = torch.cat([features(u, b, n) for u, b, _ in pairs])
    mean, std = f.mean(0), f.std(0).clamp_min(1e-12)
    t = torch.cat([tg for _, _, tg in pairs])
    su = t[:, :5].std(0).clamp_min(1e-12)
    sb = t[:, 5:].std().clamp_min(1e-12)
    net.norm[:16].copy_(torch.cat([mean, 1.0 / std]))
    if DERIVED[net.inputs]:
        sd = torch.cat([second_differences((features(u, b, n) - mean) / std, n)
                        for u, b, _ in pairs[:64]])
        d = sd.sum(-1) if net.inputs == 1 else sd.reshape(-1, 24)  # Laplacian / per axis
        net.norm[16:16 + d.shape[1]].copy_(1.0 / d.std(0).clamp_min(1e-12))
    net.scale.copy_(torch.cat([su, sb.view(1)]))


def normalised_targets(net: CorrectorNet, pairs):
    s = torch.cat([net.scale[:5], net.scale[5].repeat(3)])
    return [tg / s for _, _, tg in pairs]


def train(net: CorrectorNet, pairs, n: int, epochs: int, batch: int = 8, lr: float = 1e-3,
          seed: int = 0, noise: float = 0.0, log=print):
    """Adam on the mean-squared correction error; returns the loss history.

    noise > 0 perturbs each input's u by Gaussian noise of `noise` times the
    standard deviation of the one-step correction and asks the network to
    remove it as well (the target becomes correction - perturbation), which
    keeps corrected rollouts from amplifying their own errors."""
    buf = Buffers(n, n, n, net.device, train=True)
    targets = normalised_targets(net, pairs)
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=net.device)
    gen.manual_seed(seed)
    out_scale = net.scale[:5].view(1, 5)
    hist = []
    nstep_total = epochs * (len(pairs) // batch)
    step = 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(len(pairs))
        losses = []
        for bi in range(len(pairs) // batch):
            lr_t = lr * 0.5 * (1.0 + np.cos(np.pi * step / max(1, nstep_total)))
            for q, idx in enumerate(order[bi * batch:(bi + 1) * batch]):
                u, b, _ = pairs[idx]
                tg = targets[idx]
                if noise > 0.0:
                    e = torch.randn(u.numel() // 5, 5, device=net.device, generator=gen) * (noise * out_scale)
                    u = (u.view(-1, 5) + e).view(-1)
                    tg = tg.clone()
                    tg[:, :5] -= e / out_scale
                net.forward(buf, u, b)
                losses.append(net.backward(buf, tg, accumulate=q > 0))
            net.optimizer_step(buf, batch, float(max(lr_t, lr * 0.02)))
            step += 1
        ep_loss = float(torch.stack(losses).mean())
        hist.append(ep_loss)
        if log:
            log(f"[corrector] epoch {ep + 1}/{epochs} loss {ep_loss:.4e} "
                f"({time.perf_counter() - t0:.1f} s)")
    return hist


# ---------------------------------------------------------------------------
# corrected runs and evaluation
# ---------------------------------------------------------------------------

class CorrectedBox:
    """A coarse Box whose every step is followed by the network correction
    (with the positivity / energy-drift fallback)."""

    def __init__(self, net: CorrectorNet, n: int, etol: float = 1e-3, cfl: float = 0.7,
                 alpha: float | None = None, limit: float | None = None,
                 alpha_b: float | None = None):
        self.net, self.n, self.etol = net, n, etol
        self.alpha = float(net.meta.get("alpha", 1.0) if alpha is None else alpha)
        # network outputs limited to +-limit (units of the training correction's std dev)
        self.limit = float(net.meta.get("limit", 0.0) if limit is None else limit)
        # relaxation of the field correction (default: the same as the fluid's)
        meta_b = net.meta.get("alpha_b")
        self.alpha_b = float(alpha_b if alpha_b is not None
                             else meta_b if meta_b is not None and alpha is None else self.alpha)
        self.scale = net.scale * self.alpha  # relaxation of the correction
        self.scale[5] = net.scale[5] * self.alpha_b
        self.box = solver.Box(n, n, n, cfl=cfl)
        self.buf = Buffers(n, n, n, net.device, train=False)
        self.stats = torch.zeros(3, dtype=torch.int32, device=net.device)
        self._graph = None

    def _correct(self):
        bx = self.box
        with torch.cuda.stream(bx.stream):
            y = self.net.forward(self.buf, bx.u, bx.b)
            self.net.apply(self.buf, bx.u, bx.b, y, self.etol, self.stats, self.scale, self.limit)

    def step(self, nsteps: int = 1, graph: bool = False):
        """nsteps corrected steps.  graph=True replays one HIP graph per step
        (the CFL of the corrected state, the step, the network and the
        checked correction: one launch instead of ~30 at these small sizes);
        the result is bitwise that of the eager form.  The graph is captured
        at the first graph=True call with the etol, limit and relaxation of
        that moment; call reset_graph() after changing them (weights updated
        in place are seen)."""
        bx = self.box
        if not graph:
            for _ in range(nsteps):
                bx.run(1, graph=True)
                self._correct()
                bx.state_changed()
            return
        if self._graph is None:
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=bx.stream):
                bx.state_changed()       # the step below starts with calcfl of this state
                bx.run(1, graph=False)   # enqueued on bx.stream, so captured
                self._correct()
            self._graph = g
        with torch.cuda.stream(bx.stream):
            for _ in range(nsteps):
                self._graph.replay()
        bx.state_changed()

    def reset_graph(self):
        """Drop the captured corrected step (the next graph=True call
        captures it again with the current settings)."""
        torch.cuda.synchronize()
        self._graph = None

    def fallbacks(self):
        s = self.stats.cpu().tolist()
        return {"accepted": s[1], "rejected": s[2]}


def rel_error(u, b, uref, bref, n):
    """Per group (rho, momentum, E, B): ||x - ref|| / ||ref - mean(ref)||, and their mean."""
    ug, rg = u.view(-1, 5).double(), uref.view(-1, 5).double()
    bg, sg = b.view(-1, 3).double(), bref.view(-1, 3).double()
    out = {}
    for name, x, r in (("rho", ug[:, 0], rg[:, 0]), ("mom", ug[:, 1:4], rg[:, 1:4]),
                       ("E", ug[:, 4], rg[:, 4]), ("B", bg, sg)):
        out[name] = float(torch.linalg.norm(x - r) / torch.linalg.norm(r - r.mean(0)))
    out["mean"] = float(np.mean([out[k] for k in ("rho", "mom", "E", "B")]))
    return out


def _timed(fn):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0, out


def _device_step_s(stream, run, k: int = 40, warm: int = 3) -> float:
    """Device time of one step: HIP events on `stream` around k graph-replayed
    steps (no host synchronisation between them)."""
    with torch.cuda.stream(stream):
        run(warm)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        run(k)
        e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / k


def evaluate(net: CorrectorNet, seeds, nsteps: int, fine: int = 256, amp_v=(0.1, 0.1),
             amp_a=(0.05, 0.05), seed: int = 99, etol: float = 1e-3, alpha: float | None = None,
             alpha_b: float | None = None,
             checkpoints=None, limit: float | None = None, log=print):
    """Per held-out seed: plain n^3, corrected n^3 (n = fine/4) and plain
    2n^3 runs against the 4x downsampled fine run, at the times the corrected
    run reaches after each checkpoint step count (default: nsteps); returns
    per-seed rows with the errors at each checkpoint and GPU times.

    GPU time (gpu_s) is the steps a run took times its device time per step,
    measured once (first seed) for each of the three runs with HIP events
    around graph-replayed steps, so the three are timed alike.  wall_s is the
    host-synchronous time of the evaluation loop itself (the plain runs
    synchronise every step to land on the checkpoint time)."""
    rng = np.random.default_rng(seed)
    n, r = fine // 4, 4
    cps = sorted(set(int(c) for c in (checkpoints or [nsteps]) if 0 < c <= nsteps) | {nsteps})
    res = []
    step_s = None
    fbox = solver.Box(fine, fine, fine, cfl=0.7)
    for s in seeds:
        spec = solver.ICSpec(seed=int(s), amp_v=float(rng.uniform(*amp_v)),
                             amp_a=float(rng.uniform(*amp_a)))
        fbox.init(spec)
        with torch.cuda.stream(fbox.stream):
            u4, b4 = downsample(fbox.u, fbox.b, fine, 4)
            u2, b2 = downsample(fbox.u, fbox.b, fine, 2)
        fbox.stream.synchronize()
        cb = CorrectedBox(net, n, etol=etol, alpha=alpha, limit=limit, alpha_b=alpha_b)
        cb.box.u.copy_(u4)
        cb.box.b.copy_(b4)
        cb.box.set_time(0.0)
        pc = solver.Box(n, n, n, cfl=0.7)
        pc.set_state(u4, b4)
        p2 = solver.Box(2 * n, 2 * n, 2 * n, cfl=0.7)
        p2.set_state(u2, b2)
        wall = {"plain": 0.0, "corrected": 0.0, "plain_2x": 0.0}
        steps = {"plain": 0, "corrected": 0, "plain_2x": 0}
        row = {"seed": int(s), "amp_v": spec.amp_v, "amp_a": spec.amp_a, "checkpoints": []}
        done = 0
        for c in cps:
            dt_s, _ = _timed(lambda: cb.step(c - done, graph=True))
            wall["corrected"] += dt_s
            steps["corrected"] += c - done
            done = c
            T = cb.box.time()[0]
            for key, bx, t_end in (("plain", pc, T), ("plain_2x", p2, 2 * T)):
                dt_s, m = _timed(lambda: bx.advance_to(t_end))
                wall[key] += dt_s
                steps[key] += m
            fbox.advance_to(4 * T)  # the reference solution
            with torch.cuda.stream(fbox.stream):
                uf, bf = downsample(fbox.u, fbox.b, fine, r)
            with torch.cuda.stream(p2.stream):
                u22, b22 = downsample(p2.u, p2.b, 2 * n, 2)
            torch.cuda.synchronize()
            row["checkpoints"].append({
                "steps": c, "T": T, "plain": rel_error(pc.u, pc.b, uf, bf, n),
                "corrected": rel_error(cb.box.u, cb.box.b, uf, bf, n),
                "plain_2x": rel_error(u22, b22, uf, bf, n), "steps_taken": dict(steps),
                "wall_s": dict(wall)})
        row["fallbacks"] = cb.fallbacks()
        if step_s is None:  # per-step device times, after this seed's error runs
            step_s = {"corrected": _device_step_s(cb.box.stream, lambda m: cb.step(m, graph=True)),
                      "plain": _device_step_s(pc.stream, lambda m: pc.run(m, graph=True)),
                      "plain_2x": _device_step_s(p2.stream, lambda m: p2.run(m, graph=True))}
        for q in row["checkpoints"]:
            q["gpu_s"] = {k: q["steps_taken"][k] * step_s[k] for k in step_s}
        row["step_ms"] = {k: 1e3 * v for k, v in step_s.items()}
        last = row["checkpoints"][-1]
        row.update({k: last[k] for k in ("T", "plain", "corrected", "plain_2x", "gpu_s", "wall_s",
                                         "steps_taken")})
        row["alpha"] = cb.alpha
        row["alpha_b"] = cb.alpha_b
        row["limit"] = cb.limit
        res.append(row)
        if log:
            log(f"[corrector] seed {s}: " + "; ".join(
                f"{q['steps']} steps T={q['T']:.1f} err plain {q['plain']['mean']:.3e} corrected "
                f"{q['corrected']['mean']:.3e} 2x {q['plain_2x']['mean']:.3e}"
                for q in row["checkpoints"]) + f"; gpu s {row['gpu_s']}; steps {steps}; "
                f"{row['fallbacks']}")
        cb.box.close()
        pc.close()
        p2.close()
    fbox.close()
    return res


def summarize(rows):
    keys = ("plain", "corrected", "plain_2x")
    out = {"seeds": len(rows), "alpha": rows[0]["alpha"] if rows else None,
           "alpha_b": rows[0].get("alpha_b") if rows else None,
           "limit": rows[0]["limit"] if rows else None,
           "err_mean": {k: float(np.mean([r[k]["mean"] for r in rows])) for k in keys},
           "gpu_s_mean": {k: float(np.mean([r["gpu_s"][k] for r in rows])) for k in keys},
           "wall_s_mean": {k: float(np.mean([r["wall_s"][k] for r in rows])) for k in keys},
           "steps_mean": {k: float(np.mean([r["steps_taken"][k] for r in rows])) for k in keys},
           "step_ms": rows[0]["step_ms"] if rows else None,
           "fallback_steps": int(sum(r["fallbacks"]["rejected"] for r in rows))}
    per = {}
    for r in rows:
        for q in r["checkpoints"]:
            per.setdefault(q["steps"], {k: [] for k in keys})
            for k in keys:
                per[q["steps"]][k].append(q[k]["mean"])
    out["err_mean_at_steps"] = {st: {k: float(np.mean(v)) for k, v in d.items()}
                                for st, d in sorted(per.items())}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--seeds", type=int, default=200)
    t.add_argument("--pairs", type=int, default=8)
    t.add_argument("--skip", type=float, default=120.0,
                   help="pairs start at a coarse time drawn from [0, skip]")
    t.add_argument("--epochs", type=int, default=20)
    t.add_argument("--batch", type=int, default=8)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--fine", type=int, default=256)
    t.add_argument("--out", default=str(WEIGHTS))
    t.add_argument("--eval-seeds", type=int, default=0)
    t.add_argument("--eval-steps", type=int, default=200)
    t.add_argument("--noise", type=float, default=1.0,
                   help="input perturbation in units of the one-step correction's std dev")
    t.add_argument("--alphas", default="0.5,1.0",
                   help="correction relaxations tried on validation seeds 4000.. after training")
    t.add_argument("--val-seeds", type=int, default=2)
    t.add_argument("--val-steps", type=int, default=100)
    t.add_argument("--inputs", type=int, default=0, choices=sorted(INPUT_CHANNELS),
                   help="network input kind (mhd_corr_pack_features)")
    f = sub.add_parser("finetune")
    f.add_argument("--init", default=str(WEIGHTS))
    f.add_argument("--seeds", type=int, default=100)
    f.add_argument("--rollout", type=int, default=60)
    f.add_argument("--every", type=int, default=6)
    f.add_argument("--skip", type=float, default=100.0)
    f.add_argument("--epochs", type=int, default=10)
    f.add_argument("--batch", type=int, default=8)
    f.add_argument("--lr", type=float, default=3e-4)
    f.add_argument("--noise", type=float, default=1.0)
    f.add_argument("--fine", type=int, default=256)
    f.add_argument("--out", default=str(WEIGHTS))
    f.add_argument("--rollout-alpha", type=float, default=None)
    f.add_argument("--alphas", default="0.5,1.0")
    f.add_argument("--val-seeds", type=int, default=2)
    f.add_argument("--val-steps", type=int, default=200)
    f.add_argument("--eval-seeds", type=int, default=0)
    f.add_argument("--eval-steps", type=int, default=200)
    sel = sub.add_parser("select", help="pick the correction relaxation and output limit on "
                                        "validation seeds 4000.. and store them in the weights")
    sel.add_argument("--weights", default=str(WEIGHTS))
    sel.add_argument("--out", default=None)
    sel.add_argument("--alphas", default="0.25,0.5,1.0")
    sel.add_argument("--limits", default="0,1.5,3")
    sel.add_argument("--seeds", type=int, default=2)
    sel.add_argument("--steps", type=int, default=200)
    sel.add_argument("--fine", type=int, default=256)
    st = sub.add_parser("steptime", help="device time of a plain and of a corrected n^3 step "
                                          "(graph replays)")
    st.add_argument("--weights", default=str(WEIGHTS))
    st.add_argument("--n", type=int, default=64)
    st.add_argument("--steps", type=int, default=50)
    e = sub.add_parser("evaluate")
    e.add_argument("--weights", default=str(WEIGHTS))
    e.add_argument("--seeds", type=int, default=50)
    e.add_argument("--steps", type=int, default=200)
    e.add_argument("--fine", type=int, default=256)
    e.add_argument("--json", default=None)
    e.add_argument("--alpha", type=float, default=None)
    e.add_argument("--alpha-b", type=float, default=None,
                   help="relaxation of the field correction (default: stored, else --alpha)")
    e.add_argument("--limit", type=float, default=None)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    if a.cmd == "train":
        n = a.fine // 4
        seeds = list(range(1000, 1000 + a.seeds))
        t0 = time.perf_counter()
        pairs = make_pairs(seeds, a.pairs, fine=a.fine, skip=(0.0, a.skip), log=log)
        t_data = time.perf_counter() - t0
        net = CorrectorNet(inputs=a.inputs)
        set_normalisation(net, pairs, n)
        t0 = time.perf_counter()
        hist = train(net, pairs, n, a.epochs, a.batch, a.lr, noise=a.noise, log=log)
        t_train = time.perf_counter() - t0
        del pairs
        torch.cuda.empty_cache()
        val = {}
        for al in [float(x) for x in a.alphas.split(",") if x]:
            rows = evaluate(net, list(range(4000, 4000 + a.val_seeds)), a.val_steps, fine=a.fine,
                            alpha=al, log=log)
            val[al] = summarize(rows)["err_mean"]["corrected"]
        alpha = min(val, key=val.get) if val else 1.0
        net.meta = {"inputs": a.inputs, "train_seeds": [seeds[0], seeds[-1]], "pairs_per_seed": a.pairs,
                    "skip_max": a.skip, "epochs": a.epochs, "batch": a.batch, "lr": a.lr,
                    "fine": a.fine, "coarse": n, "loss_history": hist, "data_s": t_data,
                    "train_s": t_train, "noise": a.noise, "alpha": alpha,
                    "validation": {str(k): v for k, v in val.items()}}
        Path(a.out).parent.mkdir(parents=True, exist_ok=True)
        net.save(a.out)
        log(f"[corrector] saved {a.out}: data {t_data:.1f} s, training {t_train:.1f} s, "
            f"alpha {alpha} (validation {val})")
        if a.eval_seeds:
            rows = evaluate(net, list(range(5000, 5000 + a.eval_seeds)), a.eval_steps, fine=a.fine,
                            checkpoints=(50, 100, a.eval_steps), log=log)
            log("[corrector] summary " + json.dumps(summarize(rows)))
    elif a.cmd == "finetune":
        n = a.fine // 4
        net = CorrectorNet.load(a.init)
        if a.rollout_alpha is not None:  # else the stored relaxation drives the rollouts
            net.meta["alpha"] = a.rollout_alpha
        seeds = list(range(2000, 2000 + a.seeds))
        t0 = time.perf_counter()
        pairs = make_onpolicy_pairs(net, seeds, a.rollout, a.every, fine=a.fine, skip=(0.0, a.skip),
                                    log=log)
        t_data = time.perf_counter() - t0
        t0 = time.perf_counter()
        hist = train(net, pairs, n, a.epochs, a.batch, a.lr, noise=a.noise, seed=1, log=log)
        t_train = time.perf_counter() - t0
        del pairs
        torch.cuda.empty_cache()
        val = {}
        for al in [float(x) for x in a.alphas.split(",") if x]:
            rows = evaluate(net, list(range(4000, 4000 + a.val_seeds)), a.val_steps, fine=a.fine,
                            alpha=al, log=log)
            val[al] = summarize(rows)["err_mean"]["corrected"]
        alpha = min(val, key=val.get) if val else 1.0
        net.meta.update({"finetune_seeds": [seeds[0], seeds[-1]], "rollout": a.rollout,
                         "every": a.every, "finetune_skip_max": a.skip,
                         "finetune_epochs": a.epochs, "finetune_lr": a.lr,
                         "finetune_loss_history": hist, "finetune_data_s": t_data,
                         "finetune_train_s": t_train, "alpha": alpha,
                         "validation": {str(k): v for k, v in val.items()}})
        Path(a.out).parent.mkdir(parents=True, exist_ok=True)
        net.save(a.out)
        log(f"[corrector] saved {a.out}: data {t_data:.1f} s, training {t_train:.1f} s, "
            f"alpha {alpha} (validation {val})")
        if a.eval_seeds:
            rows = evaluate(net, list(range(5000, 5000 + a.eval_seeds)), a.eval_steps, fine=a.fine,
                            checkpoints=(50, 100, a.eval_steps), log=log)
            log("[corrector] summary " + json.dumps(summarize(rows)))
    elif a.cmd == "select":
        net = CorrectorNet.load(a.weights)
        grid = {}
        for al in [float(x) for x in a.alphas.split(",") if x]:
            for lim in [float(x) for x in a.limits.split(",") if x]:
                rows = evaluate(net, list(range(4000, 4000 + a.seeds)), a.steps, fine=a.fine,
                                checkpoints=(50, 100, a.steps), alpha=al, limit=lim, log=None)
                sm = summarize(rows)
                # mean over the checkpoints: good early and still good late
                score = float(np.mean([v["corrected"] for v in sm["err_mean_at_steps"].values()]))
                grid[f"{al},{lim}"] = {"score": score, "err_mean_at_steps": sm["err_mean_at_steps"],
                                      "fallback_steps": sm["fallback_steps"]}
                log(f"[corrector] alpha {al} limit {lim}: score {score:.4f} "
                    f"{sm['err_mean_at_steps']} fallbacks {sm['fallback_steps']}")
        best = min(grid, key=lambda k: grid[k]["score"])
        al, lim = (float(x) for x in best.split(","))
        net.meta.update({"alpha": al, "limit": lim,
                         "selection": {"seeds": [4000, 4000 + a.seeds - 1], "steps": a.steps,
                                       "grid": grid}})
        net.save(a.out or a.weights)
        log(f"[corrector] selected alpha {al} limit {lim}; saved {a.out or a.weights}")
    elif a.cmd == "steptime":
        net = CorrectorNet.load(a.weights)
        spec = solver.ICSpec(seed=5000, amp_v=0.1, amp_a=0.05)
        cb = CorrectedBox(net, a.n)
        cb.box.init(spec)
        pc = solver.Box(a.n, a.n, a.n, cfl=0.7)
        pc.init(spec)
        out = {"n": a.n, "steps": a.steps,
               "plain_ms": 1e3 * _device_step_s(pc.stream, lambda m: pc.run(m, graph=True), a.steps),
               "corrected_ms": 1e3 * _device_step_s(cb.box.stream,
                                                    lambda m: cb.step(m, graph=True), a.steps),
               "fallbacks": cb.fallbacks()}
        log("[corrector] steptime " + json.dumps(out))
        cb.box.close()
        pc.close()
    else:
        net = CorrectorNet.load(a.weights)
        rows = evaluate(net, list(range(5000, 5000 + a.seeds)), a.steps, fine=a.fine,
                        checkpoints=(50, 100, a.steps), alpha=a.alpha, limit=a.limit,
                        alpha_b=a.alpha_b, log=log)
        summ = summarize(rows)
        log("[corrector] summary " + json.dumps(summ))
        if a.json:
            Path(a.json).write_text(json.dumps({"summary": summ, "rows": rows}, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
