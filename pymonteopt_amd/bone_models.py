"""Layered bone geometries around a Turing-pattern volume (pymopt/voxel/_classes.py:510-781).

A binary trabecular volume (0 marrow/space, 1 bone, e.g. from
`pymonteopt_amd.modeling.TuringPattern`) is cropped and wrapped in cortical bone,
subcutis and dermis, then closed by the end-label shell:
* `TuringModel_Rectangular` (:510-623): flat layers on the four sides normal to x and z;
* `TuringModel_Cylinder` (:625-709): a bone cylinder of radius r_bone about the y axis in
  air, wrapped in cylindrical shells;
* `TuringModel_RnC` (:711-781): the rectangle and the cylinder intersected, wrapped in
  cylindrical shells.

Labels are 0 space, 1 trabecular, 2 cortical, 3 subcutis, 4 dermis, then air (5) for the
cylinders and the end label. The optics tables follow the reference's parameter names
(`n_space`, `ma_trabecular`, ...).

The layer thicknesses are `np.round(th / voxel_space)`, and the crop and radii use
Python's `round`, as the reference's do. The volumes are built on the host once per
model. `launch_z` is where the beam enters: z = 1, except for the cylinder, where it is
the first dermis voxel of the centre column (VoxelTuringModel._set_inital_add, :867-883).
"""
from __future__ import annotations

import numpy as np
import torch

from .geometry import center_add, optics_table

__all__ = ["TuringModel_Rectangular", "TuringModel_Cylinder", "TuringModel_RnC"]


def _pad(x, k, value, y_axis=False):
    """add_array (:585-601): k voxels of `value` on both sides of x and z (and y)."""
    if k <= 0:
        return x
    return np.pad(x, [(k, k), (k, k) if y_axis else (0, 0), (k, k)], constant_values=value)


def _disc(n0, n2, radius):
    """round_index (:664-671): cells of an (n0, n2) cross-section closer than `radius`
    voxels to (n0 // 2, n2 // 2)."""
    i = np.arange(n0)[:, None] - n0 // 2
    j = np.arange(n2)[None, :] - n2 // 2
    return np.sqrt(i * i + j * j) < radius


def _first_true(mask2d, n0, n2):
    """Leading False count along axis 0 of the column at z = n0 // 2 (:604-609)."""
    col = mask2d[:, min(n0 // 2, n2 - 1)]
    nz = np.flatnonzero(col)
    return int(nz[0]) if nz.size else col.size


class _TuringBase:
    tissues = ("_space", "_trabecular", "_cortical", "_subcutaneus", "_dermis")
    layer_labels = (2, 3, 4)

    def __init__(self):
        self.dtype_f = np.float32
        self.voxel_model = np.zeros((3, 3, 3), np.uint8)
        self.keys = list(self.params.keys())
        self.voxel_space = self.params["voxel_space"]
        self.launch_z = 1
        self._make_model_params()

    def set_params(self, *initial_data, **kwargs):
        from .voxel import set_params
        set_params(self.params, self.keys, *initial_data, **kwargs)
        self.voxel_space = self.params["voxel_space"]
        self._make_model_params()

    def _make_model_params(self):
        P = self.params
        self.ma = np.array([P["ma" + t] for t in self.tissues], np.float32)
        self.ms = np.array([P["ms" + t] for t in self.tissues], np.float32)
        self.g = np.array([P["g" + t] for t in self.tissues], np.float32)
        self.n = np.array([P["n" + t] for t in self.tissues] + [P["n_air"]], np.float32)
        self.end_point = len(self.tissues)

    def get_params(self):
        return {"ms": self.ms, "ma": self.ma, "n": self.n, "g": self.g}

    def optics(self):
        return optics_table(self.ma, self.ms, self.g, self.n[:-1], float(self.n[-1]))

    def getModelSize(self):
        print("Memory area size for voxel storage: %0.3f Mbyte" % (self.voxel_model.nbytes * 1e-6))

    @property
    def n_ambient(self):
        return float(self.params["n_air"])

    @property
    def n_first(self):
        vm = self.voxel_model
        return float(self.n[int(vm[center_add(vm.shape[0]), center_add(vm.shape[1]), self.launch_z])])

    def _layers(self):
        P = self.params
        return np.round(np.array([P["th_cortical"], P["th_subcutaneus"], P["th_dermis"]]) /
                        P["voxel_space"]).astype(np.int64)

    def _rect_crop(self, n0):
        P = self.params
        k = self._layers()
        c = int(n0 / 2 - round(P["xz_size"] / (P["voxel_space"] * 2))) + int(k[0])
        if c < 1 or 2 * c >= n0:
            raise ValueError("xz_size and th_cortical leave nothing of the bone volume to crop")
        return c

    def build(self, bone_model=None):
        if self.params.get("dicom_path"):
            # the stack of slice files replaces the volume, its spacing the voxel_space (:577-586)
            from .dicom import read_stack
            bone_model, vs = read_stack(self.params["dicom_path"])
            self.params["voxel_space"] = vs
            self.voxel_space = vs
        vm = np.ascontiguousarray(bone_model)
        if vm.ndim != 3:
            raise ValueError("the bone model must be a 3-D 0/1 volume")
        if vm.dtype in (np.int8, np.uint8, np.bool_):
            vm = vm.view(np.uint8)           # no copy; a negative int8 reads as >= 128
        elif vm.min(initial=0) < 0:
            raise ValueError("the bone model must be a 3-D 0/1 volume")
        else:
            vm = vm.astype(np.uint8)
        if vm.max(initial=0) > 1:
            raise ValueError("the bone model must be a 3-D 0/1 volume")
        self.voxel_model = self._make_voxel_model(vm)
        self.getModelSize()


class TuringModel_Rectangular(_TuringBase):
    """pymopt/voxel/_classes.py:510-623."""

    def __init__(self):
        self.model_name = "TuringModel_Rectangular"
        self.params = {
            "xz_size": 17.15, "voxel_space": 0.0245, "dicom_path": False, "bv_tv": 0.138,
            "th_cortical": 1., "th_subcutaneus": 2.6, "th_dermis": 1.4,
            "n_space": 1., "n_trabecular": 1.4, "n_cortical": 1.4, "n_subcutaneus": 1.4, "n_dermis": 1.4, "n_air": 1.,
            "ma_space": 1e-8, "ma_trabecular": 0.02374, "ma_cortical": 0.02374, "ma_subcutaneus": 0.011,
            "ma_dermis": 0.037,
            "ms_space": 1e-8, "ms_trabecular": 20.54, "ms_cortical": 17.67, "ms_subcutaneus": 20, "ms_dermis": 20,
            "g_space": 0.90, "g_trabecular": 0.90, "g_cortical": 0.90, "g_subcutaneus": 0.90, "g_dermis": .90,
        }
        super().__init__()

    def _make_voxel_model(self, vm):
        # the cropped core padded on x and z by cortical, subcutis and dermis, then by the end
        # label on every axis (add_array, :585-601), written into one allocation: the band
        # of each layer, then the core (a subject volume is ~1 GB, and each np.pad copies it)
        c = self._rect_crop(vm.shape[0])
        core = vm[c:-c, :, c:-c]
        k = [int(t) for t in self._layers()]
        K = sum(k)
        X, Y, Z = core.shape[0] + 2 * K + 2, core.shape[1] + 2, core.shape[2] + 2 * K + 2
        out = np.empty((X, Y, Z), np.uint8)
        T = torch.from_numpy(out)            # torch's host copies and fills use every core
        end = self.end_point
        for v in (T[0], T[-1], T[:, 0], T[:, -1], T[:, :, 0], T[:, :, -1]):
            v.fill_(end)
        ys = slice(1, Y - 1)
        o = 1
        for lab, t in reversed(list(zip(self.layer_labels, k))):     # outermost layer first
            if t > 0:
                T[o:o + t, ys, o:Z - o].fill_(lab)
                T[X - o - t:X - o, ys, o:Z - o].fill_(lab)
                T[o + t:X - o - t, ys, o:o + t].fill_(lab)
                T[o + t:X - o - t, ys, Z - o - t:Z - o].fill_(lab)
            o += t
        T[o:X - o, ys, o:Z - o].copy_(torch.from_numpy(core))       # core is a strided view of vm
        self.launch_z = 1
        return out


class TuringModel_Cylinder(_TuringBase):
    """pymopt/voxel/_classes.py:625-709."""
    tissues = ("_space", "_trabecular", "_cortical", "_subcutaneus", "_dermis", "_air")
    air = 5

    def __init__(self):
        self.model_name = "TuringModel_Cylinder"
        self.params = self._cyl_params()
        super().__init__()

    @staticmethod
    def _cyl_params():
        return {
            "r_bone": 9.14, "voxel_space": 0.0245, "dicom_path": False, "bv_tv": 0.138,
            "th_cortical": 1., "th_subcutaneus": 2.6, "th_dermis": 1.4,
            "n_space": 1., "n_trabecular": 1.4, "n_cortical": 1.4, "n_subcutaneus": 1.4, "n_dermis": 1.4, "n_air": 1.,
            "ma_space": 1e-8, "ma_trabecular": 0.02374, "ma_cortical": 0.02374, "ma_subcutaneus": 0.011,
            "ma_dermis": 0.037, "ma_air": 1e-5,
            "ms_space": 1e-8, "ms_trabecular": 20.54, "ms_cortical": 17.67, "ms_subcutaneus": 20, "ms_dermis": 20,
            "ms_air": 1e-5,
            "g_space": 0.90, "g_trabecular": 0.90, "g_cortical": 0.90, "g_subcutaneus": 0.90, "g_dermis": .90,
            "g_air": .90,
        }

    def _core_mask(self, n0, n2, radius):
        return _disc(n0, n2, radius)

    def _make_voxel_model(self, vm):
        n0, n2 = vm.shape[0], vm.shape[2]
        k = self._layers()
        radius = round(self.params["r_bone"] / self.params["voxel_space"]) - int(k[0])
        A = self._core_mask(n0, n2, radius)
        vm = vm.copy()
        vm[np.broadcast_to(~A[:, None, :], vm.shape)] = self.air
        x = _first_true(A, n0, n2)
        if x < 1 or 2 * x >= n0:
            raise ValueError("the bone cylinder does not fit the volume")
        A, vm = A[x:-x, x:-x], vm[x:-x, :, x:-x]
        for lab, kk in zip(self.layer_labels, k):
            kk = int(kk)
            B = np.pad(A, kk, constant_values=False)
            vm = _pad(vm, kk, self.air)
            radius += kk
            A = _disc(B.shape[0], B.shape[1], radius)
            vm[np.broadcast_to((A & ~B)[:, None, :], vm.shape)] = lab
        vm = _pad(vm, 1, self.end_point, y_axis=True)
        self.launch_z = self._first_dermis(vm)
        return vm

    def _first_dermis(self, vm):
        col = vm[center_add(vm.shape[0]), center_add(vm.shape[1])]
        hits = np.flatnonzero(col == self.end_point - 2)
        if not hits.size:
            raise ValueError("no dermis on the centre column")
        return int(hits[0])


class TuringModel_RnC(TuringModel_Cylinder):
    """pymopt/voxel/_classes.py:711-781: the rectangle of xz_size intersected with the bone
    cylinder, then cylindrical shells."""

    def __init__(self):
        self.model_name = "TuringModel_RnC"
        self.params = dict(self._cyl_params())
        self.params = {"r_bone": 9.14, "xz_size": 17.15, **{k: v for k, v in self.params.items() if k != "r_bone"}}
        _TuringBase.__init__(self)

    def _core_mask(self, n0, n2, radius):
        c = self._rect_crop(n0)
        R = np.zeros((n0, n2), bool)
        R[c:-c, c:-c] = True
        return R & _disc(n0, n2, radius)

    def _first_dermis(self, vm):
        return 1
