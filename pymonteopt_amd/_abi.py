"""ctypes mirror of include/pmo.h and the loader of pymonteopt_amd/libpmo.so.

The library is loaded from the package directory only. If it is missing the import of
the walk fails loudly: there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpmo.so")

ABI_VERSION = 11         # PMO_ABI_VERSION of include/pmo.h

PMO_OK = 0
PMO_ERR_INVALID = -1
PMO_ERR_LABEL = -2
PMO_ERR_HIP = -3
PMO_ERR_UNSUPPORTED = -4

PMO_BEAM_PENCIL = 0
PMO_BEAM_FLAT = 1
PMO_BEAM_GAUSS = 2
PMO_BEAM_STATES = 3

PMO_F_FP64 = 1 << 0
PMO_F_NORMAL_FRESNEL = 1 << 1
PMO_F_LAYER_STEP = 1 << 2

PMO_SKIP_MAX_CAP = 127

PMO_PROBE_F64, PMO_PROBE_F32, PMO_PROBE_U64, PMO_PROBE_U32, PMO_PROBE_F64_PER_XCD = 0, 1, 2, 3, 4
PMO_PROBE_F64_PER_XCD_WG, PMO_PROBE_F64_LINES = 5, 6
PMO_PROBE_F64_BLOCKS_32B, PMO_PROBE_F64_BLOCKS_64B, PMO_PROBE_F64_BLOCKS_128B = 7, 8, 9

SCALAR_NAMES = ["Rd", "Tt", "other", "A", "kill", "gain", "w0",
                "face_x-", "face_x+", "face_y-", "face_y+", "face_z-", "face_z+"]
PMO_S_COUNT = 13
EVENT_NAMES = ["packets", "interactions", "crossings", "fresnel", "roulette", "killed", "escaped",
               "global_deposits"]
PMO_EV_COUNT = 8


class PmoVolume(C.Structure):
    _fields_ = [("labels", C.c_void_p), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("n_labels", C.c_int32), ("pitch", C.c_float), ("optics", C.c_void_p), ("skip", C.c_void_p),
                ("label_radius", C.c_void_p)]


class PmoSource(C.Structure):
    _fields_ = [("kind", C.c_int32), ("ix", C.c_int32), ("iy", C.c_int32), ("iz", C.c_int32),
                ("radius", C.c_float), ("dir", C.c_float * 3), ("w0", C.c_float),
                ("add", C.c_void_p), ("p", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p),
                ("states_stride", C.c_uint64)]


class PmoRun(C.Structure):
    _fields_ = [("n_packets", C.c_uint64), ("first_packet", C.c_uint64),
                ("seed", C.c_uint32), ("stream", C.c_uint32),
                ("wth", C.c_float), ("roulette_m", C.c_float), ("flags", C.c_uint32),
                ("block_size", C.c_int32), ("grid_blocks", C.c_int32), ("packets_per_lane", C.c_int32),
                ("window", C.c_int32 * 3), ("key_dev", C.c_void_p), ("quorum", C.c_int32),
                ("refill", C.c_int32)]


class PmoTallies(C.Structure):
    _fields_ = [("scalars", C.c_void_p), ("events", C.c_void_p), ("absorbed", C.c_void_p),
                ("nr", C.c_int32), ("dr", C.c_float),
                ("rd_r", C.c_void_p), ("tt_r", C.c_void_p), ("rd_top_r", C.c_void_p),
                ("na", C.c_int32), ("rd_a", C.c_void_p), ("tt_a", C.c_void_p),
                ("nl", C.c_int32), ("dl", C.c_float), ("line_y", C.c_float), ("line_z", C.c_void_p),
                ("n_rings", C.c_int32), ("rings", C.c_void_p), ("ring_w", C.c_void_p),
                ("exit_state", C.c_void_p), ("rz_nr", C.c_int32), ("rz_nz", C.c_int32), ("rz_dr", C.c_float),
                ("rz_dz", C.c_float), ("fluence_rz", C.c_void_p), ("exit_add", C.c_void_p),
                ("exit_p", C.c_void_p), ("absorbed_lines", C.c_void_p)]


def absorbed_lines_doubles(nx, ny, nz):
    """PMO_ABSORBED_LINES_DOUBLES: the absorbed_lines scratch of an nx x ny x nz volume."""
    return 8 * ((nx + 1) // 2) * ((ny + 1) // 2) * ((nz + 1) // 2)


class PmoPlateInstance(C.Structure):
    _fields_ = [("z_labels", C.c_void_p), ("optics", C.c_void_p), ("nx", C.c_int32), ("ny", C.c_int32),
                ("nz", C.c_int32), ("n_labels", C.c_int32), ("pitch", C.c_float), ("src_kind", C.c_int32),
                ("ix", C.c_int32), ("iy", C.c_int32), ("iz", C.c_int32), ("radius", C.c_float),
                ("dir", C.c_float * 3), ("w0", C.c_float), ("n_packets", C.c_uint64), ("layers", C.c_void_p),
                ("n_layers", C.c_int32), ("reserved", C.c_int32)]


class PmoBatchTallies(C.Structure):
    _fields_ = [("scalars", C.c_void_p), ("events", C.c_void_p), ("nr", C.c_int32), ("dr", C.c_float),
                ("rd_r", C.c_void_p), ("tt_r", C.c_void_p)]


PMO_OBD_STEP_FIELDS = 12


class PmoObd(C.Structure):
    _fields_ = [("r_1", C.c_float), ("n_1", C.c_float), ("outer_r1", C.c_float), ("r_2", C.c_float),
                ("n_2", C.c_float), ("outer_r2", C.c_float), ("slit_r_in", C.c_float), ("slit_r_out", C.c_float),
                ("slit_thickness", C.c_float), ("pd_r", C.c_float), ("ld_fix_part", C.c_int32),
                ("lfp_angle", C.c_float), ("lfp_thickness", C.c_float), ("lfp_width", C.c_float),
                ("inversion", C.c_int32), ("side", C.c_float), ("expect", C.c_int32), ("seed", C.c_uint32),
                ("stream", C.c_uint32)]


class PmoTuringParams(C.Structure):
    _fields_ = [("du", C.c_float), ("dv", C.c_float), ("dx", C.c_float), ("dt", C.c_float), ("a1", C.c_float),
                ("a2", C.c_float), ("a3", C.c_float), ("b1", C.c_float), ("b2", C.c_float)]


class PmoDiff(C.Structure):
    _fields_ = [("mua_field", C.c_void_p), ("nr", C.c_int32), ("dr", C.c_float), ("rd_r", C.c_void_p),
                ("jac", C.c_void_p), ("exit_rec", C.c_void_p), ("coef", C.c_void_p), ("grad_mua", C.c_void_p),
                ("mismatch", C.c_void_p), ("grad_lines", C.c_void_p)]


# every symbol include/pmo.h declares, with its ctypes signature
SIGNATURES = {
    "pmo_abi_version": ([], C.c_int),
    "pmo_last_error": ([], C.c_char_p),
    "pmo_workspace_bytes": ([], C.c_size_t),
    "pmo_device_info": ([C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32], C.c_int),
    "pmo_walk": ([C.POINTER(PmoVolume), C.POINTER(PmoSource), C.POINTER(PmoRun), C.POINTER(PmoTallies),
                  C.c_void_p, C.c_void_p], C.c_int),
    "pmo_walk_noalloc": ([C.POINTER(PmoVolume), C.POINTER(PmoSource), C.POINTER(PmoRun), C.POINTER(PmoTallies),
                          C.c_void_p], C.c_int),
    "pmo_philox": ([C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_int64, C.c_void_p], C.c_int),
    "pmo_pack_label_radius": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p], C.c_int),
    "pmo_skip_radius": ([C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                         C.c_void_p], C.c_int),
    "pmo_walk_forward_diff": ([C.POINTER(PmoVolume), C.POINTER(PmoSource), C.POINTER(PmoRun), C.c_void_p,
                               C.c_void_p, C.POINTER(PmoDiff), C.c_void_p, C.c_void_p], C.c_int),
    "pmo_walk_adjoint": ([C.POINTER(PmoVolume), C.POINTER(PmoSource), C.POINTER(PmoRun), C.POINTER(PmoDiff),
                          C.c_void_p, C.c_void_p], C.c_int),
    "pmo_turing_create": ([C.c_int32, C.c_void_p], C.c_int),
    "pmo_turing_workspace_bytes": ([C.c_void_p], C.c_int64),
    "pmo_turing_destroy": ([C.c_void_p], C.c_int),
    "pmo_turing_step": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p], C.c_int),
    "pmo_threshold_count": ([C.c_void_p, C.c_int64, C.c_float, C.c_void_p, C.c_void_p], C.c_int),
    "pmo_count_peaks": ([C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p], C.c_int),
    "pmo_binarize": ([C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p], C.c_int),
    "pmo_fluence_bin": ([C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_int32,
                         C.c_float, C.c_void_p, C.c_void_p], C.c_int),
    "pmo_obd_trace": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int32,
                       C.c_void_p, C.c_void_p], C.c_int),
    "pmo_probe_atomics": ([C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_void_p], C.c_int),
    "pmo_probe_valu": ([C.c_void_p, C.c_int32, C.c_int32, C.c_void_p], C.c_int),
    "pmo_probe_philox": ([C.c_void_p, C.c_int32, C.c_int32, C.c_void_p], C.c_int),
    "pmo_plate_schedule": ([C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int64], C.c_int64),
    "pmo_walk_plates": ([C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.POINTER(PmoRun),
                         C.POINTER(PmoBatchTallies), C.c_void_p, C.c_void_p], C.c_int),
}

_lib = None


class PmoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libpmo error {code}: {msg}")
        self.code = code


def lib():
    """Load libpmo.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m pymonteopt_amd.build` "
                "(the walk has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        if L.pmo_abi_version() != ABI_VERSION:
            raise ImportError("libpmo ABI version mismatch")
        _lib = L
    return _lib


def check(rc: int):
    if rc != PMO_OK:
        msg = lib().pmo_last_error().decode(errors="replace")
        if rc == PMO_ERR_INVALID or rc == PMO_ERR_LABEL:
            raise ValueError(f"libpmo error {rc}: {msg}")
        raise PmoError(rc, msg)
