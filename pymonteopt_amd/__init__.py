"""pymonteopt_amd -- pyMonteOpt's voxel Monte Carlo photon walk, MI355X-native.

The walk runs in hand-written HIP for gfx950 (pymonteopt_amd/libpmo.so, C ABI in
include/pmo.h); this package is the host side:
* the pymopt voxel-model API (voxel.py), the geometry store (geometry.py) and the runtime
  (runtime.py);
* the batched sweep (batched.py) and the differentiable walk (adjoint.py);
* around the walk: the OBD receiver (optics.py), the fluence binning (fluence.py), the
  Turing-pattern bone model (modeling.py, bone_models.py), DICOM slices (dicom.py), the
  exit-dump detectors (metrics.py) and the virtual OBD pipeline (virtual_obd.py,
  param_generator.py).
"""
from .geometry import Source, bone_optics, bone_phantom, optics_table, plate_volume  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # the GPU-facing modules import torch; load them on first use
    if name in ("VoxelPlateModel", "VoxelLabelModel", "VoxelTuringModel", "VoxelPlateExModel", "PlateModel",
                "PlateExModel", "LabelVolumeModel", "set_params", "BaseVoxelMonteCarlo"):
        from . import voxel
        return getattr(voxel, name)
    if name in ("Walker", "TallySpec", "TallySet"):
        from . import runtime
        return getattr(runtime, name)
    if name == "TuringPattern":
        from .modeling import TuringPattern
        return TuringPattern
    if name in ("OBD", "SideOBD", "Grass"):
        from . import optics
        return getattr(optics, name)
    if name == "generate_variable_params":
        from .param_generator import generate_variable_params
        return generate_variable_params
    raise AttributeError(name)
