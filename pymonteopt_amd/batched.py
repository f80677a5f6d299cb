"""Batched sweeps: hundreds of plates with their own optics and packet budgets in one
persistent launch (libpmo pmo_walk_plates), optionally captured into a HIP graph.

This replaces the reference's way of sweeping optical properties, one VoxelPlateModel
build/start per parameter set (.ipynb_checkpoints/si_in_光学特性推定-checkpoint.ipynb:
13945-13967). Instance i with seed s gives exactly the tallies of a plain walk of the same
plate with (seed s, stream + i): the per-instance estimates are those of the sequential
runs, only computed together.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi
from .geometry import center_add, optics_table, plate_volume, specular


@dataclass
class Plate:
    """One instance: PlateModel parameters (pymopt/voxel/_classes.py:358-367) + budget."""
    thickness: list
    ma: list
    ms: list
    g: list
    n: list
    n_air: float = 1.0
    voxel_space: float = 0.05
    x_size: float = 5.05
    y_size: float = 5.05
    n_packets: int = 100000
    beam: str = "pencil"
    radius: float = 0.0

    def shape(self):
        l = self.voxel_space
        nx = int(np.int32(round(self.x_size / l))) + 2
        ny = int(np.int32(round(self.y_size / l))) + 2
        nz = int(np.int32(round(float(np.sum(self.thickness)) / l))) + 2
        return nx, ny, nz

    def z_labels(self):
        l = self.voxel_space
        layers = [int(round(t / l)) for t in self.thickness]
        nx, ny, nz = self.shape()
        if sum(layers) != nz - 2:
            raise ValueError("layer voxel counts do not add up to round(total/l)")
        end = len(self.thickness)
        col = np.full(nz, end, np.uint8)
        z = 1
        for i, k in enumerate(layers):
            col[z:z + k] = i
            z += k
        return col

    def layer_table(self):
        """(z_top, z_bottom, label, 0) of each run of equal labels in the interior slices,
        top to bottom, in mm below the top face (slice i spans (i-1)*l .. i*l)."""
        col = self.z_labels()
        l = self.voxel_space
        rows = []
        z, nz = 1, col.size
        while z < nz - 1:
            z0, lab = z, int(col[z])
            while z < nz - 1 and int(col[z]) == lab:
                z += 1
            rows.append((l * (z0 - 1), l * (z - 1), float(lab), 0.0))
        return np.array(rows, np.float32).reshape(-1, 4)

    def optics(self):
        return optics_table(self.ma, self.ms, self.g, self.n, self.n_air)

    def source(self):
        nx, ny, _ = self.shape()
        kind = {"pencil": _abi.PMO_BEAM_PENCIL, "flat": _abi.PMO_BEAM_FLAT, "gauss": _abi.PMO_BEAM_GAUSS}[self.beam]
        w0 = float(np.float32(1.0) - np.float32(specular(self.n_air, self.n[0])))
        return kind, (center_add(nx), center_add(ny), 1), w0

    def volume(self):
        return plate_volume(self.thickness, self.voxel_space, self.x_size, self.y_size)


def slab_sweep(n_inst=512, seed=0, pitch=0.05, lateral=101, budget=(1e5, 1e6), thickness=(0.5, 10.0)):
    """Random homogeneous slabs: thickness log-uniform (mm, rounded to whole voxels),
    mu_a log-uniform [0.005, 1] /mm, mu_s' uniform [0.5, 3] /mm, g uniform [0.7, 0.95],
    mu_s = mu_s'/(1-g), n uniform [1.33, 1.45], budgets log-uniform."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_inst):
        th = float(np.exp(rng.uniform(np.log(thickness[0]), np.log(thickness[1]))))
        nz = max(1, int(round(th / pitch)))
        ma = float(np.exp(rng.uniform(np.log(0.005), np.log(1.0))))
        msp = float(rng.uniform(0.5, 3.0))
        g = float(rng.uniform(0.7, 0.95))
        n = float(rng.uniform(1.33, 1.45))
        N = int(np.exp(rng.uniform(np.log(budget[0]), np.log(budget[1]))))
        out.append(Plate(thickness=[nz * pitch], ma=[ma], ms=[msp / (1 - g)], g=[g], n=[n], n_air=1.0,
                         voxel_space=pitch, x_size=lateral * pitch, y_size=lateral * pitch, n_packets=N))
    return out


class PlateSweep:
    """Device-resident instances, schedule and per-instance FP64 tallies of a sweep."""

    def __init__(self, plates, nr=40, dr=0.1, batch=256, device="cuda:0", layer_step=True):
        """layer_step: walk each plate from layer boundary to layer boundary and lateral wall
        (the default), or from voxel face to voxel face, which reproduces pmo_walk on the
        plate's voxel volume to the last bit. Both use the same draws."""
        self.plates = list(plates)
        self.layer_step = bool(layer_step)
        self.n_inst = len(self.plates)
        if self.n_inst < 1:
            raise ValueError("need at least one plate")
        self.device = torch.device(device)
        self.lib = _abi.lib()
        self.nr, self.dr = int(nr), float(dr)
        # packed z-label columns and optics tables
        cols = [p.z_labels() for p in self.plates]
        opts = [p.optics() for p in self.plates]
        lays = [p.layer_table() for p in self.plates]
        zoff = np.cumsum([0] + [c.size for c in cols])
        ooff = np.cumsum([0] + [o.shape[0] for o in opts])
        loff = np.cumsum([0] + [t.shape[0] for t in lays])
        self.z_labels = torch.from_numpy(np.concatenate(cols)).to(self.device)
        self.optics = torch.from_numpy(np.concatenate(opts).astype(np.float32)).to(self.device)
        self.layers = torch.from_numpy(np.concatenate(lays).astype(np.float32)).to(self.device)
        desc = (_abi.PmoPlateInstance * self.n_inst)()
        zb, ob, lb = self.z_labels.data_ptr(), self.optics.data_ptr(), self.layers.data_ptr()
        for i, p in enumerate(self.plates):
            d = desc[i]
            nx, ny, nz = p.shape()
            kind, vox, w0 = p.source()
            d.z_labels = zb + int(zoff[i])
            d.optics = ob + int(ooff[i]) * 16
            d.nx, d.ny, d.nz = nx, ny, nz
            d.n_labels = len(p.thickness)
            d.pitch = p.voxel_space
            d.src_kind = kind
            d.ix, d.iy, d.iz = vox
            d.radius = p.radius
            d.dir[0], d.dir[1], d.dir[2] = 0.0, 0.0, 1.0
            d.w0 = w0
            d.n_packets = int(p.n_packets)
            d.layers = lb + int(loff[i]) * 16
            d.n_layers = int(lays[i].shape[0])
        raw = np.frombuffer(bytes(desc), dtype=np.uint8).copy()
        self.desc = torch.from_numpy(raw).to(self.device)
        self.set_batch(batch)
        f64 = dict(dtype=torch.float64, device=self.device)
        self.scalars = torch.zeros(self.n_inst, _abi.PMO_S_COUNT, **f64)
        self.rd_r = torch.zeros(self.n_inst, max(1, self.nr), **f64)
        self.tt_r = torch.zeros(self.n_inst, max(1, self.nr), **f64)
        self.events = torch.zeros(_abi.PMO_EV_COUNT, dtype=torch.int64, device=self.device)
        self.workspace = torch.zeros(int(self.lib.pmo_workspace_bytes()), dtype=torch.uint8, device=self.device)
        self.key = torch.zeros(2, dtype=torch.int32, device=self.device)
        self.block_size = 0          # 0 = library default; autotune.tune_sweep sets the best
        self.quorum = 0              # lanes ready before a wave runs interactions; 0 = library default
        self.refill = 0              # idle lanes a wave gathers before starting packets; 0 = library default
        self._graph = None

    def set_batch(self, batch):
        """Packets per schedule entry (a wave claims one entry per work-counter atomic)."""
        self.batch = int(batch)
        budgets = (C.c_uint64 * self.n_inst)(*[int(p.n_packets) for p in self.plates])
        n = self.lib.pmo_plate_schedule(budgets, self.n_inst, self.batch, None, 0)
        _abi.check(int(n) if n < 0 else 0)
        sched = (C.c_uint32 * (4 * int(n)))()
        m = self.lib.pmo_plate_schedule(budgets, self.n_inst, self.batch, sched, int(n))
        _abi.check(int(m) if m < 0 else 0)
        self.n_batches = int(n)
        arr = np.frombuffer(bytes(sched), dtype=np.uint32).copy()
        self.schedule = torch.from_numpy(arr.view(np.int32)).to(self.device)
        self._graph = None

    @property
    def total_packets(self):
        return int(sum(p.n_packets for p in self.plates))

    def zero(self):
        self.scalars.zero_()
        self.rd_r.zero_()
        self.tt_r.zero_()
        self.events.zero_()

    def _launch(self, seed, stream_id, block_size, key_dev, hip_stream, wth=1e-4, roulette_m=10.0,
                normal_fresnel=False):
        r = _abi.PmoRun()
        r.seed = int(seed) & 0xFFFFFFFF
        r.stream = int(stream_id) & 0xFFFFFFFF
        r.wth, r.roulette_m = float(wth), float(roulette_m)
        r.flags = (_abi.PMO_F_NORMAL_FRESNEL if normal_fresnel else 0) | (_abi.PMO_F_LAYER_STEP if self.layer_step else 0)
        r.block_size = int(block_size)
        r.quorum = int(self.quorum)
        r.refill = int(self.refill)
        r.key_dev = C.c_void_p(self.key.data_ptr()) if key_dev else None
        t = _abi.PmoBatchTallies()
        t.scalars = C.c_void_p(self.scalars.data_ptr())
        t.events = C.c_void_p(self.events.data_ptr())
        t.nr, t.dr = self.nr, self.dr
        t.rd_r = C.c_void_p(self.rd_r.data_ptr()) if self.nr > 0 else None
        t.tt_r = C.c_void_p(self.tt_r.data_ptr()) if self.nr > 0 else None
        with torch.cuda.device(self.device):
            rc = self.lib.pmo_walk_plates(C.c_void_p(self.desc.data_ptr()), self.n_inst,
                                          C.c_void_p(self.schedule.data_ptr()), self.n_batches, C.byref(r),
                                          C.byref(t), C.c_void_p(self.workspace.data_ptr()),
                                          C.c_void_p(hip_stream))
        _abi.check(rc)

    def run(self, seed, stream_id=0, block_size=None, zero=True, **kw):
        """Eager sweep on torch's current stream (asynchronous)."""
        if zero:
            self.zero()
        bs = self.block_size if block_size is None else block_size
        self._launch(seed, stream_id, bs, False, torch.cuda.current_stream(self.device).cuda_stream, **kw)

    def capture(self, block_size=None):
        """Capture zero + sweep into a HIP graph; replay(seed) then runs it with a new key."""
        g = torch.cuda.CUDAGraph()
        self.key.copy_(torch.tensor([0, 0], dtype=torch.int32))
        torch.cuda.synchronize(self.device)
        bs = self.block_size if block_size is None else block_size
        with torch.cuda.graph(g):
            self.zero()
            self._launch(0, 0, bs, True, torch.cuda.current_stream(self.device).cuda_stream)
        self._graph = g
        return g

    def replay(self, seed, stream_id=0):
        if self._graph is None:
            self.capture()
        # the raw 32-bit key words, as run() passes them (the int32 tensor only carries the bits)
        key = np.array([int(seed) & 0xFFFFFFFF, int(stream_id) & 0xFFFFFFFF], np.uint32).view(np.int32)
        self.key.copy_(torch.from_numpy(key))
        self._graph.replay()

    def results(self):
        """Per-instance normalised tallies: Rd, Tt, A (per launched packet) and Rd(r), Tt(r)
        per unit area (pymopt/metrics/_scoring.py:16-25)."""
        sc = self.scalars.cpu().numpy()
        N = np.array([p.n_packets for p in self.plates], np.float64)
        out = {"Rd": sc[:, 0] / N, "Tt": sc[:, 1] / N, "A": sc[:, 3] / N, "scalars": sc,
               "events": dict(zip(_abi.EVENT_NAMES, self.events.cpu().tolist()))}
        if self.nr > 0:
            da = np.array([2 * np.pi * (i + 0.5) * self.dr ** 2 for i in range(self.nr)])
            out["r"] = np.arange(self.nr) * self.dr
            out["Rd_r"] = self.rd_r.cpu().numpy() / (da[None, :] * N[:, None])
            out["Tt_r"] = self.tt_r.cpu().numpy() / (da[None, :] * N[:, None])
            out["rd_r_raw"] = self.rd_r.cpu().numpy()
        return out
