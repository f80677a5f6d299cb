"""Turing-pattern trabecular bone model on the MI355X (pymopt/modeling_gpu/_classes.py:25-176).

`TuringPattern` keeps the reference's surface:
* the `params` keys (grid, dx, dt, du, dv, length, repetition, bv_tv, voxelsize, seed,
  ct_coef, tile_num_xz, tile_num_y) and `set_params`;
* `set_threshold_func_coef`, `set_activator_function`, `set_inhibitor_function`;
* `modeling(path)`, which returns the int8 bone/marrow label volume and writes
  `<path>/_info.json` with the microarchitecture (BV/TV, Tb.N*, Tb.Th, Tb.Sp, Dtorb).

The reaction-diffusion steps run in libpmo: `pmo_turing_step`, a fused reaction kernel, two
hipFFT R2C, a spectrum kernel with the implicit-diffusion symbol, and two C2R. So do the
threshold, the peak count and the binarisation. The initial fields are drawn on the host
from np.random as the reference draws them, so a given seed starts from the same fields.

Differences:
* Reaction terms are the reference's family f = u + dt (a1 u + a2 v + a3 u^3) and
  g = v + dt (b1 u + b2 v). `set_activator_function` and `set_inhibitor_function` accept a
  callable of that form and recover its coefficients; any other form raises ValueError.
* DICOM output (`save_dicom=True`) is written by `pymonteopt_amd.dicom`, this package's own
  reader and writer of uncompressed little-endian files, because pydicom is absent.
"""
from __future__ import annotations

import ctypes as C
import datetime
import json
import os

import numpy as np
import torch

from . import _abi
from .voxel import set_params

__all__ = ["TuringPattern"]


def _fit_reaction(fn, dt, with_cubic):
    """Coefficients (c_u, c_v[, c_u3]) of fn(u, v) = own + dt (c_u u + c_v v [+ c_u3 u^3])
    for the activator (own = u) or inhibitor (own = v); checked at random points."""
    rng = np.random.default_rng(0)
    u, v = rng.uniform(-2, 2, 12), rng.uniform(-2, 2, 12)
    base = u if with_cubic else v
    y = (np.asarray(fn(u, v), np.float64) - base) / dt
    A = np.stack([u, v, u ** 3] if with_cubic else [u, v], axis=1)
    coef, *_ = np.linalg.lstsq(A, y, rcond=None)
    if not np.allclose(A @ coef, y, rtol=1e-9, atol=1e-9):
        raise ValueError("the device model supports f = u + dt (a1 u + a2 v + a3 u^3) and "
                         "g = v + dt (b1 u + b2 v) only")
    return tuple(float(c) for c in coef)


class TuringPattern:
    """pymopt/modeling_gpu/_classes.py:25-176."""

    def __init__(self, device="cuda:0"):
        self.dtype = "float32"
        self.save_dicom = False
        self.params = {
            "grid": 40, "dx": 1 / 40, "dt": 1, "du": 0.0002, "dv": 0.01, "length": 13, "repetition": 100,
            "bv_tv": 0.138, "voxelsize": 0.0295, "seed": False, "ct_coef": 4.5e4, "tile_num_xz": 1, "tile_num_y": 1,
        }
        self.coef = np.array([-7.67681281, 1.65199492, -0.45314158, 0.60424417])
        self.keys_params = list(self.params.keys())
        self._setatter_params()
        self._a = (0.6, -1.0, -1.0)       # f(u, v) = u + dt (0.6 u - v - u^3)
        self._b = (1.5, -2.0)             # g(u, v) = v + dt (1.5 u - 2 v)
        self.device = torch.device(device)
        self.time_s = None

    # ----------------------------------------------------------------- parameters
    def set_threshold_func_coef(self, coef):
        self.coef = np.array(coef)

    def set_params(self, *initial_data, **kwargs):
        set_params(self.params, self.keys_params, *initial_data, **kwargs)
        self._setatter_params()

    def set_activator_function(self, f):
        self._a = _fit_reaction(f, float(self.dt), True)

    def set_inhibitor_function(self, g):
        self._b = _fit_reaction(g, float(self.dt), False)

    def _setatter_params(self):
        for k in self.params:
            setattr(self, k, self.params[k])
        self.gridsize = int(round(self.length * self.grid))

    # ----------------------------------------------------------------- model
    def _get_inital_vector(self):
        gs = self.gridsize
        if self.seed:
            np.random.seed(seed=int(self.seed))
        ui = 1 - np.random.rand(gs, gs, gs).astype(self.dtype)
        if self.seed:
            np.random.seed(seed=int(self.seed * 2))
        vi = 1 - np.random.rand(gs, gs, gs).astype(self.dtype)
        return ui, vi

    def _get_threshold(self):
        self.threshold = float(np.poly1d(self.coef)(self.bv_tv))
        return self.threshold

    def grow(self, u=None, v=None, steps=None):
        """The activator and inhibitor (device float32 [n][n][n]) after `steps` steps
        (default: repetition) from the given or the seeded initial fields."""
        n = self.gridsize
        if u is None or v is None:
            ui, vi = self._get_inital_vector()
            u = torch.from_numpy(ui).to(self.device)
            v = torch.from_numpy(vi).to(self.device)
        else:
            u = torch.as_tensor(u, dtype=torch.float32).to(self.device).contiguous().clone()
            v = torch.as_tensor(v, dtype=torch.float32).to(self.device).contiguous().clone()
        if tuple(u.shape) != (n, n, n) or tuple(v.shape) != (n, n, n):
            raise ValueError(f"fields must be ({n}, {n}, {n})")
        lib = _abi.lib()
        h = C.c_void_p()
        _abi.check(lib.pmo_turing_create(n, C.byref(h)))
        try:
            ws = torch.empty(int(lib.pmo_turing_workspace_bytes(h)), dtype=torch.uint8, device=self.device)
            prm = _abi.PmoTuringParams(float(self.du), float(self.dv), float(self.dx), float(self.dt), *self._a,
                                       *self._b)
            stream = torch.cuda.current_stream(self.device)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            with torch.cuda.device(self.device):
                rc = lib.pmo_turing_step(h, C.byref(prm), C.c_void_p(u.data_ptr()), C.c_void_p(v.data_ptr()),
                                         C.c_void_p(ws.data_ptr()), int(self.repetition if steps is None else steps),
                                         C.c_void_p(stream.cuda_stream))
            _abi.check(rc)
            e1.record(stream)
            e1.synchronize()
            self.time_s = e0.elapsed_time(e1) * 1e-3
        finally:
            lib.pmo_turing_destroy(h)
        return u, v

    def microarchitecture_of(self, u):
        """Thresholds u in place (u <= threshold -> 0) and returns BV/TV real and the peak
        count (strict maxima of order 5 along the last axis) on the device."""
        lib = _abi.lib()
        cnt = torch.zeros(2, dtype=torch.int64, device=self.device)
        stream = C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        n = u.numel()
        with torch.cuda.device(self.device):
            _abi.check(lib.pmo_threshold_count(C.c_void_p(u.data_ptr()), n, float(self._get_threshold()),
                                               C.c_void_p(cnt.data_ptr()), stream))
            _abi.check(lib.pmo_count_peaks(C.c_void_p(u.data_ptr()), n // u.shape[-1], int(u.shape[-1]), 5,
                                           C.c_void_p(cnt.data_ptr() + 8), stream))
        below, peaks = (int(x) for x in cnt.cpu().tolist())
        return 1 - below / n, peaks

    def modeling(self, path, save_dicom=False):
        self.save_dicom = save_dicom
        u, v = self.grow()
        del v
        self.model_shape = tuple(u.shape)
        if save_dicom:
            self._save_dicom(u.cpu().numpy(), path)
        print("Model Size: %s Mb" % (u.numel() * 4 / 1e6))
        self.bv_tv_real, peaks = self.microarchitecture_of(u)
        print("BV/TV = %.4f" % self.bv_tv_real)
        self._calc_microarchitecture(peaks)
        self._save_info(path)
        labels = torch.empty(u.shape, dtype=torch.int8, device=self.device)
        with torch.cuda.device(self.device):
            _abi.check(_abi.lib().pmo_binarize(C.c_void_p(u.data_ptr()), u.numel(), C.c_void_p(labels.data_ptr()),
                                               C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)))
        U = labels.cpu()
        if self.tile_num_xz != 0:
            # np.tile's layout, with torch's multi-threaded host copy
            U = U.repeat(self.tile_num_xz, self.tile_num_y, self.tile_num_xz)
        return U.numpy()

    def _save_dicom(self, field, path):
        """The field times ct_coef as 16-bit slices along the last axis, path/0000.dcm ...
        (TuringPattern._save_dicom, :178-183), written by pymonteopt_amd.dicom."""
        from .dicom import write_slice
        os.makedirs(path, exist_ok=True)
        px = (field.copy() * self.ct_coef).astype("uint16")
        for i in range(px.shape[2]):
            write_slice(os.path.join(path, f"{i:04}.dcm"), px[:, :, i], self.voxelsize)

    def _calc_microarchitecture(self, peaks):
        # per slice the reference averages the peaks over its rows, then sums the slices
        # (:140-150): total / gridsize
        self.num_peak = peaks / self.gridsize
        tbn = self.num_peak / (self.voxelsize * self.gridsize ** 2)
        self.microarchitecture = {
            "BV/TV": self.bv_tv,
            "BV/TV real": self.bv_tv_real,
            "Tb.N* [/mm]": tbn,
            "Tb.Th [mm]": (self.bv_tv_real / tbn) if tbn else float("inf"),
            "Tb.Sp [mm]": ((1 - self.bv_tv_real) / tbn) if tbn else float("inf"),
            "Dtorb [mg/cm3]": (self.bv_tv_real * 1200),
        }

    def _save_info(self, path, coment=""):
        info = {
            "Date": datetime.datetime.now().isoformat(),
            "coment": coment,
            "model_params": {k: (v.item() if hasattr(v, "item") else v) for k, v in self.params.items()},
            "microarchitecture": self.microarchitecture,
            "model_shape": list(self.model_shape),
        }
        os.makedirs(path, exist_ok=True)
        with open(os.path.join(path, "_info.json"), "w") as fp:
            json.dump(info, fp, indent=4)
