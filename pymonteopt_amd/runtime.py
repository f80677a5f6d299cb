"""Host runtime of the HIP walk: device buffers (torch tensors used as plain device
memory), tally sets and the call into libpmo.so on torch's current HIP stream.

`Walker.walk` is the one product entry to the hot path; the model classes in voxel.py,
the batched sweep and the adjoint are built on it.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi
from .geometry import Source, validate_volume


# By default a walk skips same-label faces only where that pays. On the smallest volumes
# (the 12^3 padded 5 mm slab at 0.5 mm) little is skipped and the skipping walk's longer
# steps cost 6%; on a fragmented one (the 64^3 four-label grid at 0.1 mm, mean radius 0.24)
# it is 10% slower. Plates and the bone phantoms (mean radius 6-10) gain 1-25%, and fine
# grids up to 9.8x.
SKIP_MIN_VOXELS = 16 * 16 * 8
SKIP_MIN_MEAN_RADIUS = 1.0


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


@dataclass
class TallySpec:
    """Which exit/volume tallies a walk accumulates (sizes fix the LDS layout)."""
    nr: int = 0
    dr: float = 0.1
    radial: tuple = ("rd_r", "tt_r", "rd_top_r")
    na: int = 0
    nl: int = 0
    dl: float = 0.1
    line_y: float = 5.0
    rings: tuple = ()            # ((r_in, r_out), ...) mm, top-face annuli
    absorbed: bool = True
    lines: bool = True           # gather the absorbed map per 2x2x2 cube (64 B line) before its atomics
    exit_state: bool = False
    exit_voxel: bool = False     # the reference kernel's exit state: end voxel + in-voxel position
    events: bool = True
    rz: tuple = ()               # (nr, dr, nz, dz): r-z internal fluence bins, mm


class TallySet:
    """FP64 device tally buffers for one volume; zero() before reuse."""

    def __init__(self, spec: TallySpec, shape, n_packets: int, device):
        self.spec = spec
        f64 = dict(dtype=torch.float64, device=device)
        nvox = int(shape[0]) * int(shape[1]) * int(shape[2])
        self.scalars = torch.zeros(_abi.PMO_S_COUNT, **f64)
        self.events = torch.zeros(_abi.PMO_EV_COUNT, dtype=torch.int64, device=device) if spec.events else None
        self.absorbed = torch.zeros(nvox, **f64) if spec.absorbed else None
        # scratch of the line-gathered map (the library clears it and adds it into absorbed)
        self.absorbed_lines = (torch.empty(_abi.absorbed_lines_doubles(*shape), **f64)
                               if spec.absorbed and spec.lines else None)
        self.hist = {}
        if spec.nr > 0:
            for k in spec.radial:
                self.hist[k] = torch.zeros(spec.nr, **f64)
        if spec.na > 0:
            self.hist["rd_a"] = torch.zeros(spec.na, **f64)
            self.hist["tt_a"] = torch.zeros(spec.na, **f64)
        if spec.nl > 0:
            self.hist["line_z"] = torch.zeros(spec.nl, **f64)
        self.rings = None
        if spec.rings:
            r = np.asarray(spec.rings, np.float32).reshape(-1)
            self.rings = torch.from_numpy(r).to(device)
            self.hist["ring_w"] = torch.zeros(r.size // 2, **f64)
        if spec.rz:
            self.hist["fluence_rz"] = torch.zeros(int(spec.rz[0]) * int(spec.rz[2]), **f64)
        self.exit_state = (torch.zeros(7 * int(n_packets), dtype=torch.float32, device=device)
                           if spec.exit_state else None)
        self.exit_add = (torch.zeros(3 * int(n_packets), dtype=torch.int32, device=device)
                         if spec.exit_voxel else None)
        self.exit_p = (torch.zeros(3 * int(n_packets), dtype=torch.float32, device=device)
                       if spec.exit_voxel else None)
        self.n_exit = int(n_packets) if (spec.exit_state or spec.exit_voxel) else 0

    def zero(self):
        self.scalars.zero_()
        if self.events is not None:
            self.events.zero_()
        if self.absorbed is not None:
            self.absorbed.zero_()
        for v in self.hist.values():
            v.zero_()
        for t in (self.exit_state, self.exit_add, self.exit_p):
            if t is not None:
                t.zero_()

    def c_struct(self) -> _abi.PmoTallies:
        s = self.spec
        t = _abi.PmoTallies()
        t.scalars = _ptr(self.scalars)
        t.events = _ptr(self.events)
        t.absorbed = _ptr(self.absorbed)
        t.nr, t.dr = int(s.nr), float(s.dr)
        t.rd_r, t.tt_r, t.rd_top_r = (_ptr(self.hist.get(k)) for k in ("rd_r", "tt_r", "rd_top_r"))
        t.na = int(s.na)
        t.rd_a, t.tt_a = _ptr(self.hist.get("rd_a")), _ptr(self.hist.get("tt_a"))
        t.nl, t.dl, t.line_y = int(s.nl), float(s.dl), float(s.line_y)
        t.line_z = _ptr(self.hist.get("line_z"))
        t.n_rings = 0 if self.rings is None else int(self.rings.numel() // 2)
        t.rings = _ptr(self.rings)
        t.ring_w = _ptr(self.hist.get("ring_w"))
        t.exit_state = _ptr(self.exit_state)
        t.exit_add = _ptr(self.exit_add)
        t.exit_p = _ptr(self.exit_p)
        t.absorbed_lines = _ptr(self.absorbed_lines)
        if s.rz:
            t.rz_nr, t.rz_dr, t.rz_nz, t.rz_dz = int(s.rz[0]), float(s.rz[1]), int(s.rz[2]), float(s.rz[3])
            t.fluence_rz = _ptr(self.hist["fluence_rz"])
        return t

    def to_host(self, shape=None) -> dict:
        """Raw (unnormalised) sums, same layout as the oracle's walk()."""
        out = {"scalars": dict(zip(_abi.SCALAR_NAMES, self.scalars.cpu().tolist()))}
        if self.events is not None:
            out["events"] = dict(zip(_abi.EVENT_NAMES, [int(x) for x in self.events.cpu().tolist()]))
        if self.absorbed is not None:
            a = self.absorbed.cpu().numpy()
            out["absorbed"] = a.reshape(shape) if shape is not None else a
        for k, v in self.hist.items():
            out[k] = v.cpu().numpy()
        if self.spec.rz:
            out["fluence_rz"] = out["fluence_rz"].reshape(int(self.spec.rz[0]), int(self.spec.rz[2]))
        if self.exit_state is not None:
            out["exit_state"] = self.exit_state.cpu().numpy().reshape(7, self.n_exit)
        if self.exit_add is not None:
            out["exit_add"] = self.exit_add.cpu().numpy().reshape(3, self.n_exit)
            out["exit_p"] = self.exit_p.cpu().numpy().reshape(3, self.n_exit)
        return out


class Walker:
    """A label volume and optics table resident on the GPU, ready for walks."""

    def __init__(self, labels: np.ndarray, pitch: float, optics: np.ndarray, device="cuda:0",
                 validate=True, skip=None, skip_cap=15):
        """skip: cross the faces inside same-label regions in one step (pmo_skip_radius,
        radius capped at skip_cap voxels); False steps voxel by voxel as the reference does;
        None (default) skips on volumes of more than 16*16*8 voxels whose mean radius is at
        least SKIP_MIN_MEAN_RADIUS. Either way every packet makes the same draws, and the same
        events unless rounding flips a decision at a face."""
        labels = np.ascontiguousarray(labels)
        optics = np.ascontiguousarray(optics, dtype=np.float32)
        if optics.ndim != 2 or optics.shape[1] != 4:
            raise ValueError("optics must be float32 [n_labels+1][4]")
        self.n_labels = optics.shape[0] - 1
        if validate:
            # the checks of validate_volume; the full scans (largest label, the end shell) run
            # on the device after the upload
            validate_volume(labels, self.n_labels, shell=False)
        self.shape = tuple(int(s) for s in labels.shape)
        self.pitch = float(pitch)
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("the walk runs on a HIP device only (no CPU fallback)")
        self.lib = _abi.lib()
        self.labels = torch.from_numpy(labels.reshape(-1)).to(self.device)
        if validate:
            self._validate_on_device()
        self.optics = torch.from_numpy(optics).to(self.device)
        self.workspace = torch.zeros(int(self.lib.pmo_workspace_bytes()), dtype=torch.uint8,
                                     device=self.device)
        self._last_launch = None        # event after the latest launch that used the workspace
        self._inflight = []             # (event, device inputs) kept alive until the event
        if skip and not (1 <= int(skip_cap) <= _abi.PMO_SKIP_MAX_CAP):
            raise ValueError(f"skip_cap must be in [1, {_abi.PMO_SKIP_MAX_CAP}]")
        self.skip = self.nvox > SKIP_MIN_VOXELS if skip is None else bool(skip)
        self.skip_cap = int(skip_cap)
        self._skip_radius = None
        self._label_radius = None
        if self.skip:
            r = self.skip_radius()      # part of the geometry upload, outside any timed walk
            # an integer tensor's sum accumulates in int64 (an explicit dtype would cast a copy)
            if skip is None and int(r.sum()) < SKIP_MIN_MEAN_RADIUS * self.nvox:
                self.skip = False       # too fragmented to gain
                self._skip_radius = None
            else:
                self.label_radius()

    def skip_radius(self) -> torch.Tensor:
        """The per-voxel uniform-cube radius (uint8, the label volume's layout), computed on
        the device once; the computing stream is synchronised, so walks on any stream may
        read it."""
        if self._skip_radius is None:
            out = torch.empty_like(self.labels)
            scratch = torch.empty_like(self.labels)
            with torch.cuda.device(self.device):
                st = torch.cuda.current_stream(self.device)
                _abi.check(self.lib.pmo_skip_radius(_ptr(self.labels), *self.shape, self.skip_cap, _ptr(out),
                                                    _ptr(scratch), C.c_void_p(st.cuda_stream)))
                st.synchronize()
            self._skip_radius = out
        return self._skip_radius

    def label_radius(self) -> torch.Tensor:
        """Labels and uniform-cube radius packed per voxel (label | radius << 8, uint16 as
        int16 storage), computed on the device once: the skipping walk reads both with one
        2-byte load per crossing."""
        if self._label_radius is None:
            rad = self.skip_radius()
            out = torch.empty(self.nvox, dtype=torch.int16, device=self.device)
            with torch.cuda.device(self.device):
                st = torch.cuda.current_stream(self.device)
                _abi.check(self.lib.pmo_pack_label_radius(_ptr(self.labels), _ptr(rad), _ptr(out), self.nvox,
                                                          C.c_void_p(st.cuda_stream)))
                st.synchronize()
            self._label_radius = out
        return self._label_radius

    def _validate_on_device(self):
        v = self.labels.view(self.shape)
        end = self.n_labels
        top = int(v.max())
        if top > end:
            raise ValueError(f"label {top} exceeds the end label {end}")
        faces = (v[0], v[-1], v[:, 0], v[:, -1], v[:, :, 0], v[:, :, -1])
        if not bool(torch.stack([(f != end).any() for f in faces]).any()):
            return
        raise ValueError("the outer 1-voxel shell of the volume must hold the end label")

    @property
    def nvox(self):
        return self.shape[0] * self.shape[1] * self.shape[2]

    def set_optics(self, optics: np.ndarray):
        o = torch.as_tensor(np.ascontiguousarray(optics, dtype=np.float32))
        if tuple(o.shape) != tuple(self.optics.shape):
            raise ValueError("optics table shape cannot change")
        self.optics.copy_(o)

    def tallies(self, spec: TallySpec, n_packets: int = 0) -> TallySet:
        return TallySet(spec, self.shape, n_packets, self.device)

    def volume_struct(self, skip=False) -> _abi.PmoVolume:
        """The volume for the C ABI; skip=True adds the uniform-cube radius and its packed
        form with the labels (pmo_walk reads that; the differentiable walk the radius)."""
        v = _abi.PmoVolume()
        v.labels = _ptr(self.labels)
        v.nx, v.ny, v.nz = self.shape
        v.n_labels = self.n_labels
        v.pitch = self.pitch
        v.optics = _ptr(self.optics)
        v.skip = _ptr(self.skip_radius()) if skip else None
        v.label_radius = _ptr(self.label_radius()) if skip else None
        return v

    def _states(self, states, n):
        """Device copies (int32 add, float32 p, v, w) of a per-packet launch state, checked
        against n packets."""
        add, p, v, w = states
        dev = dict(device=self.device)
        out = (torch.as_tensor(add).to(dtype=torch.int32, **dev).contiguous(),
               torch.as_tensor(p).to(dtype=torch.float32, **dev).contiguous(),
               torch.as_tensor(v).to(dtype=torch.float32, **dev).contiguous(),
               torch.as_tensor(w).to(dtype=torch.float32, **dev).contiguous())
        m = out[3].numel()
        if out[3].dim() != 1 or m < n or any(tuple(t.shape) != (3, m) for t in out[:3]):
            raise ValueError(f"launch states must be add, p, v of shape (3, N) and w of shape (N,), N >= {n}")
        return out

    def c_args(self, src: Source, n_packets: int, seed: int, tallies: TallySet, *, stream_id=0,
               first_packet=0, wth=1e-4, roulette_m=10.0, fp64=False, normal_fresnel=False,
               block_size=0, grid_blocks=0, packets_per_lane=0, window=(0, 0, 0), quorum=0, refill=0):
        """The pmo_volume, pmo_source, pmo_run and pmo_tallies of one walk, plus the device
        launch states that must stay alive until the walk has read them."""
        if tallies.n_exit and tallies.n_exit != int(n_packets):
            raise ValueError("exit-state buffers were sized for a different packet count")
        vol = self.volume_struct(skip=self.skip)
        s = _abi.PmoSource()
        s.kind = int(src.kind)
        s.ix, s.iy, s.iz = (int(x) for x in src.voxel)
        s.radius = float(src.radius)
        for i in range(3):
            s.dir[i] = float(src.dir[i])
        s.w0 = float(src.w0)
        keep = ()
        if src.states is not None:
            keep = self._states(src.states, int(n_packets))
            s.kind = _abi.PMO_BEAM_STATES
            s.add, s.p, s.v, s.w = (t.data_ptr() for t in keep)
            s.states_stride = int(keep[3].numel())
        r = _abi.PmoRun()
        r.n_packets = int(n_packets)
        r.first_packet = int(first_packet)
        r.seed = int(seed) & 0xFFFFFFFF
        r.stream = int(stream_id) & 0xFFFFFFFF
        r.wth = float(wth)
        r.roulette_m = float(roulette_m)
        r.flags = (_abi.PMO_F_FP64 if fp64 else 0) | (_abi.PMO_F_NORMAL_FRESNEL if normal_fresnel else 0)
        r.block_size = int(block_size)
        r.grid_blocks = int(grid_blocks)
        r.packets_per_lane = int(packets_per_lane)
        for i in range(3):
            r.window[i] = int(window[i])
        r.quorum = int(quorum)
        r.refill = int(refill)
        return vol, s, r, tallies.c_struct(), keep

    def launch_stream(self, hip_stream=None):
        """The stream a launch that uses this Walker's workspace goes on (torch's current
        stream unless a HIP stream handle is given), made to wait for the Walker's previous
        launch: every launch claims packet ids from the one counter in the workspace, so
        launches of one Walker on different streams run one after the other."""
        st = (torch.cuda.current_stream(self.device) if hip_stream is None
              else torch.cuda.ExternalStream(hip_stream, device=self.device))
        if self._last_launch is not None:
            st.wait_event(self._last_launch)
        return st

    def launched(self, st, keep=()):
        """Record a launch made on `st`; `keep` (device inputs the kernel reads) stays alive
        until that launch has completed."""
        ev = torch.cuda.Event()
        ev.record(st)
        self._last_launch = ev
        self._inflight = [(e, k) for e, k in self._inflight if not e.query()]
        if keep:
            self._inflight.append((ev, keep))

    # launch shapes tried by tune(): the library default and the shapes that measured best
    # on the MI355X for the skin plate (one 1024-lane block per CU with a 24x24x16 window)
    # and the 256^3 bone phantom (five 256-lane blocks per CU with 9x9x21 windows)
    TUNE_CANDIDATES = ({}, {"block_size": 1024, "window": (24, 24, 16), "quorum": 40},
                       {"block_size": 256, "window": (9, 9, 21), "quorum": 32},
                       {"block_size": 256, "window": (12, 12, 12), "quorum": 40})

    def tune(self, src: Source, spec: TallySpec, n_packets=10_000_000, candidates=None, seed=0x7FFF0000):
        """Times one walk of `n_packets` packets with each launch shape of `candidates`
        (block size, LDS window, quorum) on this volume and tallies, HIP events on torch's
        current stream, and returns the fastest as walk() keywords with its time. Every shape
        walks the same packets with the same draws; only the speed differs."""
        cands = self.TUNE_CANDIDATES if candidates is None else candidates
        tal = self.tallies(spec, n_packets if (spec.exit_state or spec.exit_voxel) else 0)
        best = None
        for kw in cands:
            try:
                self.walk(src, max(1, n_packets // 10), seed, tal, **kw)     # code objects, LDS shape
            except (ValueError, _abi.PmoError):
                continue                                                     # shape does not fit
            tal.zero()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            st = torch.cuda.current_stream(self.device)
            e0.record(st)
            self.walk(src, n_packets, seed, tal, **kw)
            e1.record(st)
            e1.synchronize()
            t = e0.elapsed_time(e1) * 1e-3
            if best is None or t < best[1]:
                best = (dict(kw), t)
        if best is None:
            raise ValueError("no launch shape fits this walk")
        return best

    def walk(self, src: Source, n_packets: int, seed: int, tallies: TallySet, *, hip_stream=None, **run):
        """Launch one walk (asynchronous on torch's current stream unless a HIP stream handle
        is given; after this Walker's previous launch, whatever its stream). `run` takes the
        keywords of c_args: stream_id, first_packet, wth, roulette_m, fp64, normal_fresnel,
        block_size, grid_blocks, packets_per_lane, window, quorum, refill."""
        vol, s, r, t, keep = self.c_args(src, n_packets, seed, tallies, **run)
        with torch.cuda.device(self.device):
            st = self.launch_stream(hip_stream)
            rc = self.lib.pmo_walk(C.byref(vol), C.byref(s), C.byref(r), C.byref(t),
                                   _ptr(self.workspace), C.c_void_p(st.cuda_stream))
            _abi.check(rc)
            self.launched(st, keep)
