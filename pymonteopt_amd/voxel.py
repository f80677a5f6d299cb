"""pymopt voxel-model API on the MI355X walk.

Mirrors pymopt/voxel_gpu/_classes.py (and the CPU twin pymopt/voxel/_classes.py):
`VoxelPlateModel(nPh=...).build(**params).start(rand_seed).get_result()`, with the same
parameter keys, the same KeyError contract of set_params (pymopt/utils/utilities.py:22-32)
and the same launch conventions. The walk itself runs in libpmo.so; the per-packet
post-processing of the reference (getRdTtRate, radialDistance, the virtualOBD detectors)
is replaced by FP64 tallies accumulated on the device, and get_result() returns those.
"""
from __future__ import annotations

import bz2
import datetime
import json
import pickle
import time

import numpy as np
import torch

from . import _abi
from .geometry import (Source, center_add, fresnel_oblique, optics_table, plate_volume, specular,
                       validate_volume)
from .runtime import TallySpec, Walker

__all__ = ["set_params", "PlateModel", "PlateExModel", "LabelVolumeModel", "BaseVoxelMonteCarlo",
           "VoxelPlateModel", "VoxelPlateExModel", "VoxelLabelModel", "VoxelTuringModel", "normalise"]


def set_params(data, keys, *initial_data, **kwargs):
    """Update `data` from dicts and keywords; unknown keys raise KeyError
    (pymopt/utils/utilities.py:22-32)."""
    for dictionary in initial_data:
        for key in dictionary:
            if key not in keys:
                raise KeyError(key)
            data[key] = dictionary[key]
    for key in kwargs:
        if key not in keys:
            raise KeyError(key)
        data[key] = kwargs[key]


class VoxelModel:
    def build(self):
        pass

    def set_params(self):
        pass

    def getModelSize(self):
        print("Memory area size for voxel storage: %0.3f Mbyte" % (self.voxel_model.nbytes * 1e-6))


class PlateModel(VoxelModel):
    """Layered plate (pymopt/voxel_gpu/_classes.py:396-468, pymopt/voxel/_classes.py:352-424)."""

    def __init__(self):
        self.model_name = "PlateModel"
        self.dtype_f = np.float32
        self.dtype = np.uint8
        self.params = {
            "x_size": 40, "y_size": 40,
            "voxel_space": 0.1,
            "thickness": [1.3, 40],
            "n": [1.5, 1.4],
            "n_air": 1.,
            "ma": [1e-5, 0.02374],
            "ms": [1e-5, 0.02374],
            "g": [0.9, 0.9],
        }
        self.keys = list(self.params.keys())
        self._param_instantiating()
        self.voxel_model = np.zeros((3, 3, 3), dtype=self.dtype)

    def _param_instantiating(self):
        f = self.dtype_f
        self.thickness = self.params["thickness"]
        self.n = np.array(list(self.params["n"]) + [self.params["n_air"]]).astype(f)
        self.ms = np.array(self.params["ms"]).astype(f)
        self.ma = np.array(self.params["ma"]).astype(f)
        self.g = np.array(self.params["g"]).astype(f)
        self.voxel_space = self.params["voxel_space"]
        self.x_size = np.int32(round(self.params["x_size"] / self.params["voxel_space"]))
        self.y_size = np.int32(round(self.params["y_size"] / self.params["voxel_space"]))
        self.z_size = np.int32(round(np.array(self.thickness).sum() / self.params["voxel_space"]))
        self.end_point = len(self.thickness)

    def build(self):
        self.voxel_model = plate_volume(self.thickness, self.voxel_space, self.params["x_size"],
                                        self.params["y_size"])
        self.end_point = len(self.thickness)
        self.getModelSize()

    def set_params(self, *initial_data, **kwargs):
        set_params(self.params, self.keys, *initial_data, **kwargs)
        self._param_instantiating()

    def get_params(self):
        return {"th": self.thickness, "ms": self.ms, "ma": self.ma, "n": self.n, "g": self.g}

    def optics(self):
        return optics_table(self.params["ma"], self.params["ms"], self.params["g"], self.params["n"],
                            self.params["n_air"])

    @property
    def n_ambient(self):
        return float(self.params["n_air"])

    @property
    def n_first(self):
        return float(self.params["n"][0])


class PlateExModel(VoxelModel):
    """Glass around an intralipid box (pymopt/voxel_gpu/_classes.py:470-551): label 1 fills
    int(x_size/l) x int(y_size/l) x int(thickness[1]/l) voxels, label 0 (glass) wraps it by
    int(thickness[0]/l) + 1 voxels on every side, and the outermost voxels are the end
    label 2. With first_layer_clear the beam starts in the first intralipid voxel."""

    def __init__(self):
        self.model_name = "PlateExModel"
        self.dtype_f = np.float32
        self.glass, self.intra, self.end_point = 0, 1, 2
        self.params = {
            "x_size": 40, "y_size": 40, "voxel_space": 0.1, "thickness": [1.3, 40], "n": [1.5, 1.4], "n_air": 1.,
            "ma": [1e-5, 0.02374], "ms": [1e-5, 0.02374], "g": [0.9, 0.9],
        }
        self.keys = list(self.params.keys())
        self._param_instantiating()
        self.voxel_model = np.zeros((3, 3, 3), dtype=np.uint8)

    def _param_instantiating(self):
        P, f = self.params, self.dtype_f
        self.thickness = P["thickness"]
        self.n = np.array(list(P["n"]) + [P["n_air"]]).astype(f)
        self.ms = np.array(P["ms"]).astype(f)
        self.ma = np.array(P["ma"]).astype(f)
        self.g = np.array(P["g"]).astype(f)
        self.voxel_space = P["voxel_space"]
        self.x_size = int(P["x_size"] / P["voxel_space"])
        self.y_size = int(P["y_size"] / P["voxel_space"])
        self.z_size = int(P["thickness"][1] / P["voxel_space"])
        self.num_pix = int(P["thickness"][0] / P["voxel_space"])

    def build(self):
        k = self.num_pix + 1
        vm = np.full((self.x_size, self.y_size, self.z_size), self.intra, np.uint8)
        vm = np.pad(vm, k, constant_values=self.glass)
        vm[[0, -1], :, :] = self.end_point
        vm[:, [0, -1], :] = self.end_point
        vm[:, :, [0, -1]] = self.end_point
        self.voxel_model = vm
        self.getModelSize()

    def set_params(self, *initial_data, **kwargs):
        set_params(self.params, self.keys, *initial_data, **kwargs)
        self._param_instantiating()

    def get_params(self):
        return {"th": self.thickness, "ms": self.ms, "ma": self.ma, "n": self.n, "g": self.g}

    def get_second_layer_addz(self):
        return self.num_pix + 1

    def optics(self):
        P = self.params
        return optics_table(P["ma"], P["ms"], P["g"], P["n"], P["n_air"])

    @property
    def n_ambient(self):
        return float(self.params["n_air"])

    @property
    def n_first(self):
        return float(self.params["n"][0])


class LabelVolumeModel(VoxelModel):
    """Arbitrary label volume with a per-label optics table (the role of the reference's
    TuringModel_* classes, pymopt/voxel_gpu/_classes.py:554-821, for volumes built
    elsewhere). `build(volume)` takes the padded uint8 volume whose shell is the end label;
    the beam enters at the top-face centre like the plate's."""

    def __init__(self):
        self.model_name = "LabelVolumeModel"
        self.params = {"voxel_space": 0.02, "ma": [0.1], "ms": [10.0], "g": [0.9], "n": [1.37],
                       "n_air": 1.0, "launch_z": 1}
        self.keys = list(self.params.keys())
        self.voxel_model = np.zeros((3, 3, 3), dtype=np.uint8)
        self._param_instantiating()

    def _param_instantiating(self):
        self.voxel_space = self.params["voxel_space"]
        self.n = np.array(list(self.params["n"]) + [self.params["n_air"]], np.float32)
        self.end_point = len(self.params["ma"])

    def set_params(self, *initial_data, **kwargs):
        set_params(self.params, self.keys, *initial_data, **kwargs)
        self._param_instantiating()

    def build(self, volume=None):
        if volume is not None:
            self.voxel_model = np.ascontiguousarray(volume, dtype=np.uint8)
        validate_volume(self.voxel_model, self.end_point)
        self.getModelSize()

    def get_params(self):
        return {k: self.params[k] for k in ("ma", "ms", "g", "n")}

    def optics(self):
        return optics_table(self.params["ma"], self.params["ms"], self.params["g"], self.params["n"],
                            self.params["n_air"])

    @property
    def n_ambient(self):
        return float(self.params["n_air"])

    @property
    def n_first(self):
        vm = self.voxel_model
        lab = int(vm[center_add(vm.shape[0]), center_add(vm.shape[1]), int(self.params["launch_z"])])
        return float(self.params["n"][lab])


def normalise(raw: dict, n_packets: int, spec: TallySpec, pitch: float, optics: np.ndarray,
              shape=None, interior=True) -> dict:
    """Turn raw FP64 sums into the reference's normalised quantities
    (getRdTtRate, radialDistance, angularyResolved, lineDistance; pymopt/metrics/_scoring.py)."""
    N = float(n_packets)
    s = raw["scalars"]
    out = {"nPh": int(n_packets), "Rd": s["Rd"] / N, "Tt": s["Tt"] / N, "A": s["A"] / N,
           "energy": {k: v / N for k, v in s.items()}}
    if "events" in raw:
        out["events"] = raw["events"]
    if spec.nr > 0:
        dr = spec.dr
        alpha = np.array([i * dr for i in range(spec.nr + 1)])
        da = np.array([2 * np.pi * (i + 0.5) * dr ** 2 for i in range(spec.nr)])
        out["r"] = alpha[:-1]
        for k, name in (("rd_r", "Rd_r"), ("tt_r", "Tt_r"), ("rd_top_r", "Rd_top_r")):
            if k in raw:
                out[name] = raw[k] / (da * N)
    if spec.na > 0:
        d = np.pi / (2 * spec.na)
        alpha = np.array([(i + 0.5) * d for i in range(spec.na + 1)])
        do = 4 * np.pi * np.sin(alpha) * np.sin(d / 2)
        out["alpha"] = alpha[:-1]
        out["Rd_a"] = raw["rd_a"] / (do[:-1] * N)
        out["Tt_a"] = raw["tt_a"] / (do[:-1] * N)
    if spec.nl > 0:
        out["z"] = np.array([i * spec.dl for i in range(spec.nl)])
        out["line_z"] = raw["line_z"] / (spec.dl * spec.line_y * 2 * N)
    if spec.rings:
        rings = np.asarray(spec.rings, np.float64).reshape(-1, 2)
        area = np.pi * (rings[:, 1] ** 2 - rings[:, 0] ** 2)
        out["rings"] = rings
        out["ring_R"] = raw["ring_w"] / (area * N)
    if spec.rz and "fluence_rz" in raw:
        # Fluence2D's bins (pymopt/fluence/intarnal_fluence.py:92-106): annulus volume
        # pi ((i+1)^2 - i^2) dr^2 dz, centres (i + 1/2) dr and (k + 1/2) dz
        rz_nr, rz_dr, rz_nz, rz_dz = spec.rz
        i = np.arange(int(rz_nr))
        vol = np.pi * ((i + 1) ** 2 - i ** 2) * rz_dr ** 2 * rz_dz
        out["fluence_rz"] = raw["fluence_rz"] / (vol[:, None] * N)
        out["rz_r"] = (i + 0.5) * rz_dr
        out["rz_z"] = (np.arange(int(rz_nz)) + 0.5) * rz_dz
    if "absorbed" in raw:
        a = raw["absorbed"]
        if shape is not None:
            a = a.reshape(shape)
        if interior:
            a = a[1:-1, 1:-1, 1:-1]
        out["absorbed"] = a / N
        out["_absorbed_raw_interior"] = a
    return out


class BaseVoxelMonteCarlo:
    """Walk driver with the reference's constructor/start/get_result surface
    (pymopt/voxel_gpu/_classes.py:41-382)."""

    def __init__(self, *, nPh, model, dtype_f=np.float32, dtype=np.int32, beam_type="TEM00", w_beam=0,
                 beam_angle=0, initial_refrect_by_angle=False, first_layer_clear=False, threadnum=128,
                 nr=50, dr=0.1, na=0, nl=0, dl=0.1, line_y=5.0, rings=(), fluence=True,
                 exit_state=False, fluence_rz=None, fp64=False, normal_fresnel=False, device="cuda:0"):
        beam_type_list = ["TEM00", "flat", False]
        if beam_type not in beam_type_list:
            raise ValueError("%s is not a permitted for factor. Please choose from %s." % (beam_type, beam_type_list))
        if threadnum > 1024:
            raise Exception("There are too many threads. threadnum < 1024.")
        if first_layer_clear and not hasattr(model, "get_second_layer_addz"):
            raise ValueError("first_layer_clear needs a model with a clear first layer (PlateExModel)")
        self.beam_type = beam_type
        self.dtype = dtype
        self.dtype_f = dtype_f
        self.nPh = nPh
        self.w_beam = w_beam
        self.beam_angle = beam_angle
        self.initial_refrect_by_angle = initial_refrect_by_angle
        self.first_layer_clear = first_layer_clear
        self.model = model
        self.threadnum = int(threadnum)
        self.spec = TallySpec(nr=int(nr), dr=float(dr), na=int(na), nl=int(nl), dl=float(dl),
                              line_y=float(line_y), rings=tuple(tuple(r) for r in rings),
                              absorbed=bool(fluence), exit_state=bool(exit_state),
                              rz=tuple(fluence_rz) if fluence_rz else ())
        self.fp64 = bool(fp64)
        self.normal_fresnel = bool(normal_fresnel)
        self.device = device
        self._walker = None
        self._result = None
        self.time_s = None

    # ----------------------------------------------------------------- set-up
    def build(self, *initial_data, **kwargs):
        if initial_data == () and kwargs == {}:
            pass
        else:
            self.model.set_params(*initial_data, **kwargs)
        self.model.build()
        self._walker = None
        return self

    def set_monte_params(self, *, nPh, model, dtype_f=np.float32, dtype=np.int32, w_beam=0):
        self.dtype = dtype
        self.dtype_f = dtype_f
        self.nPh = nPh
        self.w_beam = w_beam
        self.model = model
        self._walker = None

    def get_voxel_model(self):
        return self.model.voxel_model

    def get_model_params(self):
        return self.model.get_params()

    def _oblique(self):
        return self.beam_angle != 0 and self.w_beam == 0

    def source(self) -> Source:
        """Launch description, pymopt/voxel_gpu/_classes.py:144-231 (_set_inital_add,
        _set_inital_vector, _set_inital_w)."""
        vm = self.model.voxel_model
        ix, iy = center_add(vm.shape[0]), center_add(vm.shape[1])
        if self.first_layer_clear:
            iz = int(self.model.get_second_layer_addz())
        else:
            iz = int(getattr(self.model, "launch_z", getattr(self.model, "params", {}).get("launch_z", 1)))
        n1, n2 = self.model.n_ambient, self.model.n_first
        direction = (0.0, 0.0, 1.0)
        rsp = specular(n1, n2)
        if self._oblique():
            at, r = fresnel_oblique(float(self.beam_angle), n1, n2)
            direction = (float(np.sin(at)), 0.0, float(np.cos(at)))
            rsp = r if n1 != n2 else 0.0
            if self.initial_refrect_by_angle:
                rsp = 0.0              # the reflected share leaves as packets instead (start())
        elif self.first_layer_clear:
            # air | clear first layer (n[0]) | second layer (n[1]), with the multiple
            # reflections inside the clear layer (_set_inital_w, :205-218)
            n2, n3 = float(self.model.n[0]), float(self.model.n[1])
            rsp = 0.0
            if n1 != n2:
                r1, r2 = specular(n1, n2), specular(n2, n3)
                rsp = r1 + r2 * (1 - r1) ** 2 / (1 - r1 * r2)
        kind, radius = _abi.PMO_BEAM_PENCIL, 0.0
        if self.beam_type == "flat" and self.w_beam != 0:
            kind, radius = _abi.PMO_BEAM_FLAT, float(self.w_beam)
        elif self.beam_type == "TEM00" and self.w_beam != 0:
            kind, radius = _abi.PMO_BEAM_GAUSS, float(self.w_beam)
        return Source(voxel=(ix, iy, iz), w0=float(np.float32(1.0) - np.float32(rsp)), kind=kind,
                      radius=radius, dir=direction)

    def walker(self) -> Walker:
        if self._walker is None:
            self._walker = Walker(self.model.voxel_model, float(self.model.voxel_space), self.model.optics(),
                                  device=self.device)
        return self._walker

    # ----------------------------------------------------------------- run
    def start(self, rand_seed=None):
        if rand_seed is None:
            rand_seed = np.random.randint(1e5)
        self.nPh = int(self.nPh)
        self.rand_seed = int(rand_seed)
        wk = self.walker()
        # initial_refrect_by_angle (_set_inital_vector, :164-182): a share R(angle) of the
        # packets is reflected at entry with w = 1 and v = (sin a, 0, -cos a); they join the
        # results without walking
        n_refl = 0
        if self.initial_refrect_by_angle and self._oblique():
            _, r = fresnel_oblique(float(self.beam_angle), self.model.n_ambient, self.model.n_first)
            n_refl = int(np.random.default_rng([self.rand_seed, 0x5EF1]).binomial(self.nPh, r))
        n_walk = self.nPh - n_refl
        tal = wk.tallies(self.spec, n_walk)
        print("")
        print("###### Start (Random seed: %s) ######" % rand_seed)
        print("")
        # threadnum sized the reference's CUDA blocks (one packet per thread); the persistent
        # walk picks its own block size and LDS window from the run (block_size 0)
        bs = 0
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.time()
        ev0.record()
        if n_walk > 0:
            wk.walk(self.source(), n_walk, self.rand_seed, tal, fp64=self.fp64,
                    normal_fresnel=self.normal_fresnel, block_size=bs)
        ev1.record()
        ev1.synchronize()
        self.time_s = ev0.elapsed_time(ev1) * 1e-3
        raw = tal.to_host(shape=wk.shape)
        if n_refl:
            self._add_entry_reflections(raw, n_refl)
        self._raw = raw
        self._result = normalise(raw, self.nPh, self.spec, wk.pitch, self.model.optics(), shape=wk.shape)
        if "absorbed" in self._result:
            self._result["fluence"] = self._fluence(self._result["_absorbed_raw_interior"])
        self._result.pop("_absorbed_raw_interior", None)
        if self.spec.exit_state:
            es = raw["exit_state"]
            self._result["p"] = es[0:3].copy()
            self._result["v"] = es[3:6].copy()
            self._result["w"] = es[6].copy()
        print("###### End ######")
        self.getRdTtRate()
        print("Calculation time: %0.3f seconds (walk %0.3f s on the GPU)." % (time.time() - t0, self.time_s))
        return self

    def _add_entry_reflections(self, raw, n):
        """n packets of weight 1 leaving the top face centre at (sin a, 0, -cos a)."""
        a = float(self.beam_angle)
        s = raw["scalars"]
        for k in ("Rd", "face_z-", "w0"):
            s[k] += float(n)
        if "events" in raw:
            raw["events"]["packets"] += n
            raw["events"]["escaped"] += n
        if "rd_a" in raw and self.spec.na > 0:
            b = int(np.ceil(a / (np.pi / (2 * self.spec.na)))) - 1
            if 0 <= b < self.spec.na:
                raw["rd_a"][b] += float(n)
        if "exit_state" in raw:
            extra = np.zeros((7, n), np.float32)
            extra[3], extra[5], extra[6] = np.sin(a), -np.cos(a), 1.0
            raw["exit_state"] = np.concatenate([extra, raw["exit_state"]], axis=1)

    def _fluence(self, absorbed_raw_interior):
        """Phi[v] = A[v] / (mu_a[label v] * l^3 * nPh) in mm^-2 (zero where mu_a = 0)."""
        vm = self.model.voxel_model[1:-1, 1:-1, 1:-1]
        opt = self.model.optics()
        mua = opt[:, 0].astype(np.float64)[np.minimum(vm, opt.shape[0] - 1)]
        l3 = float(self.model.voxel_space) ** 3
        with np.errstate(divide="ignore", invalid="ignore"):
            phi = np.where(mua > 0, absorbed_raw_interior / (mua * l3 * self.nPh), 0.0)
        return phi

    # ----------------------------------------------------------------- results
    def get_result(self):
        if self._result is None:
            raise RuntimeError("start() has not been run")
        return self._result

    def getRdTtRate(self):
        res = self.get_result()
        self.Rdw = res["Rd"]
        self.Ttw = res["Tt"]
        print("######")
        print("Mean Rd %0.6f" % self.Rdw)
        print("Mean Td %0.6f" % self.Ttw)
        print()

    def getRdTtValues(self):
        return {"Rd": self.Rdw, "Td": self.Ttw}

    def save_result(self, fname, *, coment="", save_monte=True, save_params=True):
        """Writes <fname>_LID.pkl.bz2 (the get_result dict) and <fname>_info.json
        (pymopt/voxel/_classes.py:295-338)."""
        if save_monte:
            with bz2.open(fname + "_LID.pkl.bz2", "wb") as fp:
                fp.write(pickle.dumps(self.get_result()))
        if save_params:
            info = {
                "Date": datetime.datetime.now().isoformat(),
                "coment": coment,
                "number_of_photons": self.nPh,
                "calc_dtype": "64 bit" if self.fp64 else "32 bit",
                "model": {"model_name": self.model.model_name,
                          "model_params": {k: np.asarray(v).tolist() for k, v in self.model.get_params().items()},
                          "model_voxel_space": self.model.voxel_space},
                "w_beam": self.w_beam,
                "beam_angle": self.beam_angle,
                "beam_mode": self.beam_type,
                "rand_seed": getattr(self, "rand_seed", None),
            }
            with open(fname + "_info.json", "w") as fp:
                json.dump(info, fp, indent=4)


class VoxelPlateModel(BaseVoxelMonteCarlo):
    """pymopt/voxel_gpu/_classes.py:827-841."""

    def __init__(self, *, nPh=1000, dtype_f=np.float32, dtype=np.int32, beam_type="TEM00", w_beam=0,
                 beam_angle=0, initial_refrect_by_angle=False, first_layer_clear=False, threadnum=128,
                 **tally_kwargs):
        super().__init__(nPh=nPh, model=PlateModel(), dtype_f=dtype_f, dtype=dtype, w_beam=w_beam,
                         beam_angle=beam_angle, beam_type=beam_type,
                         initial_refrect_by_angle=initial_refrect_by_angle,
                         first_layer_clear=first_layer_clear, threadnum=threadnum, **tally_kwargs)


class VoxelLabelModel(BaseVoxelMonteCarlo):
    """Walk through an arbitrary label volume (the reference's VoxelTuringModel role,
    pymopt/voxel_gpu/_classes.py:888-1006)."""

    def __init__(self, *, nPh=1000, beam_type="TEM00", w_beam=0, beam_angle=0, threadnum=128, **tally_kwargs):
        super().__init__(nPh=nPh, model=LabelVolumeModel(), beam_type=beam_type, w_beam=w_beam,
                         beam_angle=beam_angle, threadnum=threadnum, **tally_kwargs)

    def build(self, volume=None, *initial_data, **kwargs):
        if initial_data or kwargs:
            self.model.set_params(*initial_data, **kwargs)
        self.model.build(volume)
        self._walker = None
        return self


class VoxelTuringModel(BaseVoxelMonteCarlo):
    """pymopt/voxel/_classes.py:847-1000: a walk through a Turing-pattern bone volume
    wrapped by TuringModel_Rectangular, _Cylinder or _RnC. `set_model(u)` takes the 0/1
    volume (TuringPattern.modeling's output), `build(**params)` wraps it. The reference's
    default name 'TuringModel' matches none of its classes (it fails there); here it
    selects the rectangular model."""

    def __init__(self, *, nPh=1000, beam_type="TEM00", w_beam=0, beam_angle=0, threadnum=128,
                 model_name="TuringModel_Rectangular", **tally_kwargs):
        from .bone_models import TuringModel_Cylinder, TuringModel_Rectangular, TuringModel_RnC
        models = {"TuringModel_Rectangular": TuringModel_Rectangular, "TuringModel": TuringModel_Rectangular,
                  "TuringModel_Cylinder": TuringModel_Cylinder, "TuringModel_RnC": TuringModel_RnC}
        if model_name not in models:
            raise ValueError("Invalid name: %s (choose from %s)" % (model_name, sorted(models)))
        super().__init__(nPh=nPh, model=models[model_name](), beam_type=beam_type, w_beam=w_beam,
                         beam_angle=beam_angle, threadnum=threadnum, **tally_kwargs)
        self.bone_model = None

    def set_model(self, u):
        self.bone_model = u

    def set_params(self, *initial_data, **kwargs):
        self.model.set_params(*initial_data, **kwargs)

    def build(self, *initial_data, **kwargs):
        if initial_data or kwargs:
            self.model.set_params(*initial_data, **kwargs)
        if self.bone_model is None and not self.model.params.get("dicom_path"):
            raise ValueError("set_model(u) first, or give dicom_path")
        self.model.build(self.bone_model)
        self.bone_model = None
        self._walker = None
        return self


class VoxelPlateExModel(BaseVoxelMonteCarlo):
    """pymopt/voxel_gpu/_classes.py:843-886: the glass-wrapped intralipid box, the beam
    starting under the clear glass by default."""

    def __init__(self, *, nPh=1000, dtype_f=np.float32, dtype=np.int32, beam_type="TEM00", w_beam=0, beam_angle=0,
                 initial_refrect_by_angle=False, first_layer_clear=True, threadnum=128, **tally_kwargs):
        super().__init__(nPh=nPh, model=PlateExModel(), dtype_f=dtype_f, dtype=dtype, w_beam=w_beam,
                         beam_angle=beam_angle, beam_type=beam_type,
                         initial_refrect_by_angle=initial_refrect_by_angle, first_layer_clear=first_layer_clear,
                         threadnum=threadnum, **tally_kwargs)

    def set_params(self, *initial_data, **kwargs):
        self.model.set_params(*initial_data, **kwargs)
