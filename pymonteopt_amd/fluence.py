"""Internal fluence in r-z and x-y-z bins: pymopt/fluence/intarnal_fluence.py on the GPU.

`Fluence2D` and `Fluence3D` keep the reference's constructor (nr, nz, dr, dz), `getArz`,
`getArrayZ`, `getArrayR` and `saveFluesnce(p, w)`. The binning runs in libpmo's
`pmo_fluence_bin`, and every weight stays with its own point. The reference sorts the
points by z but not the weights (:35-37, :113-115), and its bin search skips points without
advancing its weight index (:44-86, :122-151); both are fixed here. The bins are
left-closed, [i*dr, (i+1)*dr), as the reference's comparisons are. Fluence3D divides the
weights by the bin volume dz*dr^2, as its saveFluesnce does. Sums are FP64.

A walk can also tally the r-z fluence on the device as it goes:
`VoxelPlateModel(..., fluence_rz=(nr, dr, nz, dz))` gives `get_result()["fluence_rz"]`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _abi

__all__ = ["Fluence2D", "Fluence3D"]


class _FluenceBins:
    dims = 2

    def __init__(self, *, nr, nz, dr, dz, device="cuda:0"):
        self.nr, self.nz, self.dr, self.dz = int(nr), int(nz), float(dr), float(dz)
        self.device = torch.device(device)
        self.z = np.array([i * dz for i in range(self.nz + 1)])
        self.Arz = torch.zeros(self._shape(), dtype=torch.float64, device=self.device)
        print("Memory area size for fluence storage: %d Mbyte" % (self.Arz.numel() * 8 * 1e-6))

    def getArz(self):
        return self.Arz.cpu().numpy()

    def getArrayZ(self):
        return np.array([(i + 0.5) * self.dz for i in range(self.nz)])

    def _add(self, p, w, scale=1.0):
        p_t = torch.as_tensor(np.asarray(p) if not isinstance(p, torch.Tensor) else p)
        p_t = p_t.to(device=self.device, dtype=torch.float32).contiguous()
        w_t = torch.as_tensor(np.asarray(w) if not isinstance(w, torch.Tensor) else w)
        w_t = (w_t.to(device=self.device, dtype=torch.float32) * scale).contiguous()
        if p_t.dim() != 2 or p_t.shape[0] != 3 or w_t.shape != (p_t.shape[1],):
            raise ValueError("p must be (3, n) and w (n,)")
        n = p_t.shape[1]
        with torch.cuda.device(self.device):
            rc = _abi.lib().pmo_fluence_bin(C.c_void_p(p_t.data_ptr()), n, C.c_void_p(w_t.data_ptr()), n, self.dims,
                                            self.nr, self.dr, self.nz, self.dz, C.c_void_p(self.Arz.data_ptr()),
                                            C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream))
        _abi.check(rc)
        torch.cuda.current_stream(self.device).synchronize()


class Fluence2D(_FluenceBins):
    """pymopt/fluence/intarnal_fluence.py:92-152: bins of r = hypot(x, y) and z."""
    dims = 2

    def _shape(self):
        self.r = np.array([i * self.dr for i in range(self.nr + 1)])
        return (self.nr, self.nz)

    def getArrayR(self):
        return np.array([(i + 0.5) * self.dr for i in range(self.nr)])

    def saveFluesnce(self, p, w):
        self._add(p, w)


class Fluence3D(_FluenceBins):
    """pymopt/fluence/intarnal_fluence.py:12-90: x, y bins over [-nr*dr, nr*dr) and z bins;
    weights per unit bin volume."""
    dims = 3

    def _shape(self):
        self.r = np.array([i * self.dr for i in np.arange(-self.nr, self.nr + 1)])
        self.vol = self.dz * self.dr ** 2
        return (2 * self.nr, 2 * self.nr, self.nz)

    def getArrayR(self):
        return np.array([(i + 0.5) * self.dr for i in np.arange(-self.nr, self.nr)])

    def saveFluesnce(self, p, w):
        self._add(p, w, scale=1.0 / self.vol)
