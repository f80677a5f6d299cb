"""OBD receiver on the MI355X: the optical-beam-deflection set-up of pymopt/optics/_classes.py.

The set-up traces the packets a walk sends back out of the sample (its exit states p, v,
w) through a receiver scanned in focus offset Z:
* lens 1, a plano-convex lens with its plano face toward the sample;
* an annular slit, the LD mount, a second slit;
* lens 2, a plano-convex lens with its convex face toward the sample;
* a photodiode.

The result is I(Z), the weight reaching the photodiode. The ray trace runs in libpmo's
`pmo_obd_trace`, one lane per packet over all focus steps. The classes mirror the
reference's:
* `OBD` and `SideOBD` (:381-567) have the same parameter keys, `set_params` /
  `set_monte_data` / `load_file` / `start` / `get_result` / `save_result`, and the same
  DataFrame of Z, int and log(int);
* `Grass` (:223-293) gives the glass indices and the closed-form marginal ray.

Two differences:
* The Fresnel cut at each lens face draws from Philox (seed, stream) rather than
  np.random. `start(expect=True)` weights each ray by its transmission instead, which gives
  the cut's mean with less variance.
* Exactly normal incidence uses the Fresnel limit where the reference's sin/tan form gives
  NaN and drops the ray. That is a single direction.
"""
from __future__ import annotations

import bz2
import ctypes as C
import math
import pickle
import time

import numpy as np
import torch

from . import _abi
from .voxel import set_params

__all__ = ["Grass", "OBD", "SideOBD"]


class Grass:
    """Glass indices and the closed-form marginal ray of lens 1 (pymopt/optics/_classes.py:223-293)."""

    def __init__(self):
        self.supported_type = ["N-BK7", "N-SF11"]

    def get_inital_angle(self, slit_radius=10, lens_curvature_radius=51.68, wavelength=850, grass_type="N-BK7"):
        n1 = self.get_refractive_index(wavelength, grass_type)
        a1 = math.asin(slit_radius / lens_curvature_radius)
        a2 = math.asin(math.sin(a1) / n1)
        return math.asin(n1 * math.sin(a1 - a2))

    def get_real_bfl(self, slit_radius=10, lens_curvature_radius=51.68, lens_ct=10.0, wavelength=850,
                     grass_type="N-BK7"):
        n1 = self.get_refractive_index(wavelength, grass_type)
        r, h = lens_curvature_radius, slit_radius
        a1 = math.asin(h / r)
        a2 = math.asin(math.sin(a1) / n1)
        a3 = a1 - a2
        a4 = math.asin(n1 * math.sin(a3))
        length = math.sqrt(r * r - h * h) - r + lens_ct
        return (h - length * math.tan(a3)) / math.tan(a4)

    def get_refractive_index(self, ramda, grass_type="N-BK7"):
        if grass_type == "N-BK7":
            return self._NBK7(ramda)
        if grass_type == "N-SF11":
            return self._NSF11(ramda)
        return 1.517

    @staticmethod
    def _poly(x, c3, c2, c1, c0):
        return ((c3 * x + c2) * x + c1) * x + c0

    def _NBK7(self, ramda):
        x = ramda * 1e-3
        if 312.6 <= ramda <= 486.1:
            return self._poly(x, -2.4880, 3.5910, -1.8102, 1.8395)
        if 486.1 < ramda <= 852.1:
            return self._poly(x, -0.1605, 0.3866, -0.3307, 1.6102)
        if 852.1 < ramda <= 2325:
            return self._poly(x, -0.0029, 0.0125, -0.0304, 1.5284)
        raise ValueError("The wavelength is over the range that can be supported.")

    def _NSF11(self, ramda):
        x = ramda * 1e-3
        if 404.7 <= ramda <= 589.3:
            return self._poly(x, -4.4506, 7.8207, -4.7512, 2.7792)
        if 589.3 < ramda <= 852.1:
            return self._poly(x, -0.4659, 1.1864, -1.0622, 2.0937)
        if 852.1 < ramda <= 1060:
            return self._poly(x, 0.0, 0.0427, -0.1192, 1.8324)
        if 1060 < ramda <= 2325.4:
            return self._poly(x, -0.0046, 0.0257, -0.0648, 1.7993)
        raise ValueError("The wavelength is over the range that can be supported.")


def _as_device(a, device, dtype=torch.float32):
    t = a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a))
    return t.to(device=device, dtype=dtype).contiguous()


class OBD:
    """pymopt/optics/_classes.py:381-525."""

    def __init__(self, device="cuda:0"):
        self.params = {
            "start": -2.04, "end": 62.96, "split": 0.25, "wavelength": 850,
            "outerD_1": 50, "efl_1": 100, "bfl_1": 93.41, "ct_1": 10, "et_1": 3.553, "r_1": 51.68, "n_1": 1.517,
            "substrate_1": "N-BK7",
            "outerD_2": 50, "efl_2": 50, "bfl_2": 43.28, "ct_2": 12, "et_2": 3.01, "r_2": 39.24, "n_2": 1.758,
            "substrate_2": "N-SF11",
            "slit_outerD": 50, "slit_D": 20, "slit_width": 2, "slit_thickness": 5,
            "d_pd": 3, "ld_fix_part": True, "distance_2slits": 37, "pd_poit_correction": 0,
            "inversion": False, "side": False,
        }
        self._set_refrective_index()
        self.keys_params = list(self.params.keys())
        self.data = {"p": 0, "v": 0, "w": 0, "nPh": 1000, "z_max": 0}
        self.keys_data = list(self.data.keys())
        self.nPh = 1000
        self.device = torch.device(device)
        self.result = None
        self.time_s = None

    def _set_refrective_index(self):
        g = Grass()
        self.params["n_1"] = g.get_refractive_index(self.params["wavelength"], self.params["substrate_1"])
        self.params["n_2"] = g.get_refractive_index(self.params["wavelength"], self.params["substrate_2"])

    def set_params(self, *initial_data, **kwargs):
        set_params(self.params, self.keys_params, *initial_data, **kwargs)
        self._set_refrective_index()

    def set_monte_data(self, *initial_data, **kwargs):
        """Exit states p [3][n], v [3][n], w [n] and nPh. A dict argument may be a model's
        whole get_result(): only these keys are taken from it (this package's result dict
        also holds the tallies). Keywords must be known keys (KeyError otherwise)."""
        dicts = [{k: d[k] for k in d if k in self.keys_data} for d in initial_data]
        set_params(self.data, self.keys_data, *dicts, **kwargs)
        self.nPh = self.data["nPh"]

    def load_file(self, path):
        """Reads a `<name>_LID.pkl.bz2` written by a model's save_result."""
        with bz2.open(path, "rb") as fp:
            self.data = pickle.loads(fp.read())
        self.nPh = self.data["nPh"]

    # ----------------------------------------------------------------- scan set-up
    def steps(self):
        P = self.params
        return np.arange(P["start"], P["end"], P["split"])

    def _x_shift(self, Z):
        return np.zeros_like(Z)

    def step_table(self):
        """Planes and sphere centres of every focus step (OBD.opticalUnit, :454-460), FP64
        on the host, stored as the float rows pmo_obd_trace reads."""
        P = self.params
        Z = self.steps()
        zl1 = -P["bfl_1"] + Z
        zl2 = zl1 - P["ct_1"] - P["slit_thickness"] * 2 - P["ct_2"] - P["distance_2slits"]
        zs1 = zl1 - P["ct_1"]
        t = np.zeros((Z.size, _abi.PMO_OBD_STEP_FIELDS), np.float64)
        t[:, 0] = zl1
        t[:, 1] = zl1 - (P["ct_1"] - P["r_1"])
        t[:, 2] = zs1
        t[:, 3] = zs1 - P["slit_thickness"] - 10
        t[:, 4] = zl2 + P["ct_2"] + P["slit_thickness"]
        t[:, 5] = zl2
        t[:, 6] = zl2 + (P["ct_2"] - P["r_2"])
        t[:, 7] = zl2 - P["bfl_2"] - P["pd_poit_correction"]
        t[:, 8] = self._x_shift(Z)
        return Z, t.astype(np.float32)

    def _config(self, seed, stream, expect):
        P = self.params
        c = _abi.PmoObd()
        c.r_1, c.n_1, c.outer_r1 = P["r_1"], P["n_1"], P["outerD_1"] / 2
        c.r_2, c.n_2, c.outer_r2 = P["r_2"], P["n_2"], P["outerD_2"] / 2
        c.slit_r_in = P["slit_D"] / 2 - P["slit_width"] / 2
        c.slit_r_out = P["slit_D"] / 2 + P["slit_width"] / 2
        c.slit_thickness = P["slit_thickness"]
        c.pd_r = P["d_pd"] / 2
        c.ld_fix_part = 1 if P["ld_fix_part"] else 0
        c.lfp_angle, c.lfp_thickness, c.lfp_width = 68 / 180, 5.0, 6.0
        c.inversion, c.side = self._view()
        c.expect = 1 if expect else 0
        c.seed, c.stream = int(seed) & 0xFFFFFFFF, int(stream) & 0xFFFFFFFF
        return c

    def _view(self):
        P = self.params
        return (1 if P["inversion"] else 0), (float(P["side"]) if P["side"] else 0.0)

    # ----------------------------------------------------------------- device trace
    def trace(self, p, v, w, seed=0, stream=0, expect=False, hip_stream=None):
        """I(Z) [n_steps] (FP64 device tensor) for exit states p, v [3][n] and weights w [n]
        (numpy arrays or tensors; device tensors are used in place)."""
        dev = self.device
        p_t, v_t, w_t = _as_device(p, dev), _as_device(v, dev), _as_device(w, dev)
        if p_t.dim() != 2 or p_t.shape[0] != 3 or v_t.shape != p_t.shape or w_t.shape != (p_t.shape[1],):
            raise ValueError("p and v must be (3, n) and w (n,)")
        return self._launch(p_t.data_ptr(), v_t.data_ptr(), w_t.data_ptr(), p_t.shape[1], p_t.shape[1], seed,
                            stream, expect, hip_stream, keep=(p_t, v_t, w_t))

    def trace_exit_states(self, exit_state, seed=0, stream=0, expect=False, hip_stream=None):
        """I(Z) straight from a walk's device exit-state buffer [7][n] (no copy)."""
        es = exit_state.reshape(7, -1)
        if es.device != self.device or es.dtype != torch.float32 or not es.is_contiguous():
            raise ValueError("exit_state must be a contiguous float32 tensor on the OBD's device")
        n = es.shape[1]
        base = es.data_ptr()
        return self._launch(base, base + 3 * n * 4, base + 6 * n * 4, n, n, seed, stream, expect, hip_stream,
                            keep=(es,))

    def _device_steps(self):
        """The step table on the device, rebuilt only when the parameters change (a scan
        repeated over several walks reuses it without host work or a copy)."""
        key = repr(sorted((k, np.asarray(v).tolist()) for k, v in self.params.items()))
        cached = getattr(self, "_steps_cache", None)
        if cached is None or cached[0] != key:
            Z, table = self.step_table()
            cached = (key, Z, torch.from_numpy(table).to(self.device))
            self._steps_cache = cached
        return cached[1], cached[2]

    def _launch(self, pp, vp, wp, n, stride, seed, stream, expect, hip_stream, keep):
        Z, steps = self._device_steps()
        out = torch.zeros(Z.size, dtype=torch.float64, device=self.device)
        cfg = self._config(seed, stream, expect)
        lib = _abi.lib()
        hs = torch.cuda.current_stream(self.device).cuda_stream if hip_stream is None else hip_stream
        with torch.cuda.device(self.device):
            rc = lib.pmo_obd_trace(C.byref(cfg), C.c_void_p(pp), C.c_void_p(vp), C.c_void_p(wp), int(n), int(stride),
                                   C.c_void_p(steps.data_ptr()), int(Z.size), C.c_void_p(out.data_ptr()),
                                   C.c_void_p(hs))
        _abi.check(rc)
        self._keep = (steps, keep)          # alive until the asynchronous launch is done
        return out

    # ----------------------------------------------------------------- reference surface
    def start(self, show_graph=False, seed=None, expect=False):
        import pandas as pa
        if seed is None:
            seed = np.random.randint(1e5)
        t0 = time.time()
        intensity = self.trace(self.data["p"], self.data["v"], self.data["w"], seed=seed, expect=expect)
        I = intensity.cpu().numpy()
        self.time_s = time.time() - t0
        Z = self.steps()
        with np.errstate(divide="ignore"):
            logi = np.log10(I / self.nPh)
        if show_graph:
            import matplotlib.pyplot as plt
            plt.figure(figsize=(8, 6), dpi=90)
            plt.plot(Z, logi, "-", c="k")
            plt.xlabel("$Z\\ [mm]$")
            plt.ylabel("$log_{10}(I/I_0)$")
            plt.show()
        self.result = pa.DataFrame([Z, I], index=["Z", "int"]).T
        self.result["log(int)"] = logi
        return self

    def get_result(self):
        return self.result

    def save_result(self, path):
        self.result.to_csv(path + "_opt.csv", index=False)


class SideOBD(OBD):
    """pymopt/optics/_classes.py:527-567: the receiver looks at the side, the sample's
    x shifting by -(Z + dz) tan(theta0) per step, with theta0 and dz from Grass's marginal
    ray of lens 1 (N-BK7)."""

    def _view(self):
        return 0, 0.0

    def _x_shift(self, Z):
        P = self.params
        g = Grass()
        th0 = g.get_inital_angle(slit_radius=P["slit_D"] / 2, wavelength=P["wavelength"],
                                 lens_curvature_radius=P["r_1"], grass_type="N-BK7")
        z_rel = g.get_real_bfl(slit_radius=P["slit_D"] / 2, lens_curvature_radius=P["r_1"], lens_ct=P["ct_1"],
                               wavelength=P["wavelength"], grass_type="N-BK7")
        dz = z_rel - P["bfl_1"]
        return -(Z + dz) * math.tan(th0)
