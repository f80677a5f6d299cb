"""Differentiable walk: d Rd(r) / d optical properties on the GPU.

Two autograd paths, both behind torch.autograd.Function:
- `rd_of_label_optics(mua, mus)` : per-label absorption and scattering. The forward pass
  accumulates the perturbation-MC sums (libpmo pmo_walk_forward_diff), so backward is a
  product with the Jacobian it already holds.
- `rd_of_absorption_field(field)` : per-voxel absorption. Backward replays every detected
  packet from its Philox counter and deposits the gradient along its path
  (libpmo pmo_walk_adjoint); no path is stored.

`fit_layers` uses the first path with Adam to recover layer optical properties from a
measured Rd(r), the job the reference does with a ParameterGrid sweep of separate walks
(.ipynb_checkpoints/si_in_光学特性推定-checkpoint.ipynb:13945-13967).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _abi
from .geometry import Source
from .runtime import Walker


PmoDiff = _abi.PmoDiff


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class DiffWalk:
    """A label volume ready for differentiable walks of `n_packets` packets with Rd(r)
    detected in `nr` bins of `dr` mm."""

    def __init__(self, labels, pitch, optics, source: Source, n_packets=1e7, nr=40, dr=0.1, device="cuda:0",
                 lines="auto", skip=None):
        # skip: forward passes that need no replay (the per-label Jacobians) cross same-label
        # cubes in one step where the Walker finds that it pays (None) or always (True); the
        # replay, and every forward it may replay, steps voxel by voxel
        self.walker = Walker(labels, pitch, optics, device=device, skip=skip)
        self.device = self.walker.device
        self.base_optics = np.array(optics, np.float32)
        self.n_labels = self.base_optics.shape[0] - 1
        self.source = source
        self.n_packets = int(n_packets)
        self.nr, self.dr = int(nr), float(dr)
        self.area = torch.tensor([2 * np.pi * (i + 0.5) * self.dr ** 2 for i in range(self.nr)], dtype=torch.float64,
                                 device=self.device)
        self.lib = _abi.lib()
        self.last_mismatch = 0
        # True: replay deposits outside the LDS box gather per 2x2x2-voxel cube in LDS line
        # buffers and reach the gradient through a cube-major scratch (one atomic request per
        # cube line); False: one FP64 atomic per voxel deposit, beside a larger LDS box.
        # "auto": the first adjoint times both and keeps the faster (the line buffers pay
        # where many deposits fall outside the box; on the 64^3 grid at 0.1 mm the larger box
        # wins). Launches on a DiffWalk are ordered (Walker.launch_stream), so successive
        # replays share the scratch.
        if lines not in (True, False, "auto"):
            raise ValueError("lines must be True, False or 'auto'")
        self.lines = lines
        self._grad_lines = None

    @property
    def shape(self):
        return self.walker.shape

    def _structs(self, seed, stream_id, skip=False):
        v = self.walker.volume_struct(skip=skip)
        s = _abi.PmoSource()
        src = self.source
        s.kind = int(src.kind)
        s.ix, s.iy, s.iz = (int(x) for x in src.voxel)
        s.radius = float(src.radius)
        for i in range(3):
            s.dir[i] = float(src.dir[i])
        s.w0 = float(src.w0)
        r = _abi.PmoRun()
        r.n_packets = self.n_packets
        r.seed = int(seed) & 0xFFFFFFFF
        r.stream = int(stream_id) & 0xFFFFFFFF
        r.wth, r.roulette_m = 1e-4, 10.0
        return v, s, r

    def set_optics(self, optics):
        self.walker.set_optics(np.asarray(optics, np.float32))

    def forward(self, seed=0, stream_id=0, mua_field=None, jac=True, replay=True):
        """One forward pass. Returns Rd(r) per unit area and packet, the Jacobians
        dRd(r)/dmu_a and dRd(r)/dmu_s per label (when jac), and the exit record. replay=False
        (without mua_field) lets it cross same-label cubes in one step: the same draws and
        per-label sums up to rounding, faster, but its record cannot be replayed."""
        if jac and self.n_labels > 4:
            raise ValueError("per-label Jacobians support at most 4 labels")
        f64 = dict(dtype=torch.float64, device=self.device)
        scal = torch.zeros(_abi.PMO_S_COUNT, **f64)
        rd = torch.zeros(self.nr, **f64)
        J = torch.zeros(self.nr, self.n_labels, 2, **f64) if jac else None
        rec = torch.empty(2 * self.n_packets, dtype=torch.int32, device=self.device)
        field = None
        if mua_field is not None:
            field = mua_field.detach().to(self.device, torch.float32).contiguous()
            if tuple(field.shape) != tuple(self.shape):
                raise ValueError(f"mua_field must have the padded volume shape {self.shape}")
        d = PmoDiff(_p(field), self.nr, self.dr, _p(rd), _p(J), _p(rec), None, None, None, None)
        skip = not replay and field is None and self.walker.skip
        v, s, r = self._structs(seed, stream_id, skip=skip)
        with torch.cuda.device(self.device):
            st = self.walker.launch_stream()
            rc = self.lib.pmo_walk_forward_diff(C.byref(v), C.byref(s), C.byref(r), _p(scal), None, C.byref(d),
                                                _p(self.walker.workspace), C.c_void_p(st.cuda_stream))
            _abi.check(rc)
            self.walker.launched(st)
        norm = self.area * self.n_packets
        # the optics table this pass walked with: the replay must use the same one
        out = {"Rd_r": rd / norm, "scalars": scal, "exit_rec": rec, "field": field, "seed": seed,
               "stream": stream_id, "optics": self.walker.optics.clone(), "replayable": not skip}
        if jac:
            mus = torch.as_tensor(self.walker.optics[: self.n_labels, 1].to(torch.float64))
            wl, wn = J[..., 0], J[..., 1]
            out["J_mua"] = -wl / norm[:, None]
            out["J_mus"] = (wn / mus[None, :] - wl) / norm[:, None]
        return out

    def adjoint(self, fwd, grad_rd_r):
        """d loss / d mu_a per voxel (padded shape) for the forward pass `fwd`, given
        d loss / d Rd(r). The replay walks with the optics table of that forward pass, even
        if the walker's table changed since (another forward on this DiffWalk), and raises
        if any replayed packet left differently from its recorded exit: the gradient would
        then not be that of the forward pass."""
        if not fwd.get("replayable", True):
            raise ValueError("this forward pass crossed same-label cubes in one step (forward(replay=False)); "
                             "run forward(replay=True) for a record the adjoint can replay")
        if self.lines == "auto":
            # each mode once untimed (its kernel's first launch carries one-time costs that
            # can exceed the difference between the modes), then once timed
            self._line_scratch()
            times = {}
            for mode in (True, False):
                self._replay(fwd, grad_rd_r, mode)
            for mode in (True, False):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                grad = self._replay(fwd, grad_rd_r, mode)
                e1.record()
                e1.synchronize()
                times[mode] = e0.elapsed_time(e1)
            self.lines = min(times, key=times.get)
            return grad
        return self._replay(fwd, grad_rd_r, self.lines)

    def _line_scratch(self):
        if self._grad_lines is None:
            self._grad_lines = torch.empty(_abi.absorbed_lines_doubles(*self.shape), dtype=torch.float64,
                                           device=self.device)
        return self._grad_lines

    def _replay(self, fwd, grad_rd_r, lines):
        g = grad_rd_r.detach().to(self.device, torch.float64).reshape(-1)
        coef = (g / (self.area * self.n_packets)).contiguous()
        grad = torch.zeros(self.shape, dtype=torch.float64, device=self.device)
        mism = torch.zeros(1, dtype=torch.int64, device=self.device)
        d = PmoDiff(_p(fwd["field"]), self.nr, self.dr, None, None, _p(fwd["exit_rec"]), _p(coef), _p(grad),
                    _p(mism), _p(self._line_scratch()) if lines else None)
        v, s, r = self._structs(fwd["seed"], fwd["stream"])
        current = self.walker.optics.clone()
        self.walker.optics.copy_(fwd["optics"])              # stream-ordered before the replay
        try:
            with torch.cuda.device(self.device):
                st = self.walker.launch_stream()
                rc = self.lib.pmo_walk_adjoint(C.byref(v), C.byref(s), C.byref(r), C.byref(d),
                                               _p(self.walker.workspace), C.c_void_p(st.cuda_stream))
                _abi.check(rc)
                self.walker.launched(st)
        finally:
            self.walker.optics.copy_(current)
        self.last_mismatch = int(mism.item())
        if self.last_mismatch:
            raise RuntimeError(f"adjoint replay: {self.last_mismatch} packets left differently from the forward "
                               "pass (volume, source, field or packet count changed since); gradient discarded")
        return grad


class _LabelOptics(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mua, mus, dw: DiffWalk, seed, stream_id):
        opt = dw.base_optics.copy()
        L = dw.n_labels
        opt[:L, 0] = mua.detach().cpu().numpy()
        opt[:L, 1] = mus.detach().cpu().numpy()
        dw.set_optics(opt)
        # the backward is J^T g: no replay, so the forward may cross same-label cubes in one step
        out = dw.forward(seed=seed, stream_id=stream_id, jac=True, replay=False)
        ctx.save_for_backward(out["J_mua"], out["J_mus"])
        return out["Rd_r"].to(mua.dtype)

    @staticmethod
    def backward(ctx, g):
        J_mua, J_mus = ctx.saved_tensors
        g = g.to(torch.float64)
        return (J_mua.T @ g).to(g.dtype), (J_mus.T @ g).to(g.dtype), None, None, None


class _AbsorptionField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, field, dw: DiffWalk, seed, stream_id):
        out = dw.forward(seed=seed, stream_id=stream_id, mua_field=field, jac=False)
        ctx.dw = dw
        ctx.fwd = out
        return out["Rd_r"].to(field.dtype)

    @staticmethod
    def backward(ctx, g):
        grad = ctx.dw.adjoint(ctx.fwd, g)
        return grad.to(g.dtype), None, None, None


def rd_of_label_optics(mua, mus, dw: DiffWalk, seed=0, stream_id=0):
    """Rd(r) [nr] as a differentiable function of per-label mu_a [L] and mu_s [L] (device tensors)."""
    return _LabelOptics.apply(mua, mus, dw, seed, stream_id)


def rd_of_absorption_field(field, dw: DiffWalk, seed=0, stream_id=0):
    """Rd(r) [nr] as a differentiable function of a per-voxel mu_a field (padded shape)."""
    return _AbsorptionField.apply(field, dw, seed, stream_id)


def fit_layers(dw: DiffWalk, target_rd_r, g, mua0, musp0, iters=150, lr=0.05, seed0=1000, bins=None,
               callback=None, lr_final=None, average=0):
    """Adam on log mu_a and log mu_s' of each label so that the walk's Rd(r) matches
    `target_rd_r` (relative squared error over `bins`); g stays fixed, mu_s = mu_s'/(1-g).
    Each iteration walks with a fresh seed, so the iterates carry that walk's noise: with
    `lr_final` the step size decays geometrically from `lr` to it, and with `average` > 0
    the result is the mean of the last `average` iterates (in log space)."""
    dev = dw.device
    g = torch.as_tensor(np.asarray(g, np.float64), device=dev)
    la = torch.log(torch.as_tensor(np.asarray(mua0, np.float64), device=dev)).requires_grad_(True)
    ls = torch.log(torch.as_tensor(np.asarray(musp0, np.float64), device=dev)).requires_grad_(True)
    target = torch.as_tensor(np.asarray(target_rd_r, np.float64), device=dev)
    sel = torch.arange(target.numel(), device=dev) if bins is None else torch.as_tensor(bins, device=dev)
    opt = torch.optim.Adam([la, ls], lr=lr)
    sched = (torch.optim.lr_scheduler.ExponentialLR(opt, gamma=(lr_final / lr) ** (1.0 / max(iters - 1, 1)))
             if lr_final else None)
    hist = []
    acc_a, acc_s, n_acc = torch.zeros_like(la), torch.zeros_like(ls), 0
    for it in range(iters):
        opt.zero_grad()
        mua = torch.exp(la)
        mus = torch.exp(ls) / (1 - g)
        rd = rd_of_label_optics(mua, mus, dw, seed=seed0 + it)
        loss = (((rd[sel] - target[sel]) / target[sel]) ** 2).mean()
        loss.backward()
        opt.step()
        if sched is not None:
            sched.step()
        if average and it >= iters - average:
            acc_a += la.detach()
            acc_s += ls.detach()
            n_acc += 1
        hist.append(float(loss.detach()))
        if callback:
            callback(it, float(loss.detach()), torch.exp(la).detach().cpu().numpy(),
                     torch.exp(ls).detach().cpu().numpy())
    if n_acc:
        return torch.exp(acc_a / n_acc).cpu().numpy(), torch.exp(acc_s / n_acc).cpu().numpy(), hist
    return torch.exp(la).detach().cpu().numpy(), torch.exp(ls).detach().cpu().numpy(), hist
