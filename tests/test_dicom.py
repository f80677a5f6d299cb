"""The build's own DICOM reader and writer (pymonteopt_amd.dicom): Part-10 layout checked
byte by byte, round trips, an explicit-VR file assembled by hand, readDicom's 8-bit
reduction, and bone geometries built from a DICOM stack (no GPU)."""
import struct

import numpy as np
import pytest

from pymonteopt_amd import dicom
from pymonteopt_amd.bone_models import TuringModel_Rectangular


def test_part10_layout_and_round_trip(tmp_path):
    a = (np.arange(12, dtype=np.uint16) * 1000).reshape(3, 4)
    f = tmp_path / "s.dcm"
    dicom.write_slice(str(f), a, 0.0295)
    b = f.read_bytes()
    assert b[:128] == b"\0" * 128 and b[128:132] == b"DICM"
    # file meta group: (0002,0000) UL 4 <length of the rest of the group>
    g, e, vr, ln, glen = struct.unpack_from("<HH2sHI", b, 132)
    assert (g, e, vr, ln) == (2, 0, b"UL", 4)
    end_meta = 144 + glen
    assert struct.unpack_from("<H", b, end_meta)[0] != 2          # the data set starts after it
    s = dicom.read_file(str(f))
    assert s.transfer_syntax == dicom.IMPLICIT_LE
    assert (s.Rows, s.Columns) == (4, 3) and s.PixelSpacing == [0.0295, 0.0295]
    # the reference's convention: Columns = shape[0], Rows = shape[1], bytes in C order
    assert np.array_equal(s.pixel_array, np.frombuffer(a.tobytes(), np.uint16).reshape(4, 3))
    assert s.pixel_array.dtype == np.uint16


def test_explicit_vr_file_assembled_by_hand(tmp_path):
    def el(g, e, vr, v):
        return struct.pack("<HH2sH", g, e, vr, len(v)) + v
    meta = el(2, 0x10, b"UI", dicom.EXPLICIT_LE.encode() + b"\0")
    meta = el(2, 0, b"UL", struct.pack("<I", len(meta))) + meta
    px = np.array([[1, 2], [3, 250]], np.uint8)
    ds = (el(0x28, 0x10, b"US", struct.pack("<H", 2)) + el(0x28, 0x11, b"US", struct.pack("<H", 2)) +
          el(0x28, 0x30, b"DS", b"0.5\\0.5 ") + el(0x28, 0x100, b"US", struct.pack("<H", 8)) +
          struct.pack("<HH2sHI", 0x7FE0, 0x10, b"OB", 0, 4) + px.tobytes())
    f = tmp_path / "e.dcm"
    f.write_bytes(b"\0" * 128 + b"DICM" + meta + ds)
    s = dicom.read_file(str(f))
    assert s.transfer_syntax == dicom.EXPLICIT_LE
    assert np.array_equal(s.pixel_array, px) and s.PixelSpacing == [0.5, 0.5]
    bad = tmp_path / "jpeg.dcm"
    meta2 = el(2, 0x10, b"UI", b"1.2.840.10008.1.2.4.50\0")
    bad.write_bytes(b"\0" * 128 + b"DICM" + el(2, 0, b"UL", struct.pack("<I", len(meta2))) + meta2)
    with pytest.raises(ValueError):
        dicom.read_file(str(bad))


def test_stack_round_trip_and_readdicom(tmp_path):
    rng = np.random.default_rng(1)
    vol = (rng.uniform(size=(5, 6, 7)) < 0.4).astype(np.int8)
    dicom.write_stack(str(tmp_path / "st"), vol, 0.02)
    back, vs = dicom.read_stack(str(tmp_path / "st"))
    assert np.array_equal(back, vol) and vs == 0.02
    arr, dims, spacing = dicom.readDicom(str(tmp_path / "st"), size_down=False)
    # each slice vol[k] is (Rows, Columns) = (6, 7); readDicom stacks them last
    assert dims == (6, 7, 5) and spacing == (0.02, 0.02, 0.02)
    assert np.array_equal(arr, np.moveaxis(vol, 0, -1))
    big = np.full((2, 3, 3), 1000, np.uint16)
    dicom.write_stack(str(tmp_path / "b"), big, 0.1)
    small, _, _ = dicom.readDicom(str(tmp_path / "b"))
    assert np.all(small == np.int8(round(1000 / 256)))
    assert np.array_equal(dicom.reConstArray(np.array([10, 40])), [0, 40])
    assert np.array_equal(dicom.reConstArray_8(np.array([30, 40])), [0, 40])   # 9500 / 257 -> 37


def test_bone_geometry_from_a_dicom_stack(tmp_path):
    vol = (np.random.default_rng(2).uniform(size=(40, 40, 40)) < 0.3).astype(np.int8)
    dicom.write_stack(str(tmp_path / "bone"), vol, 0.1)
    kw = dict(xz_size=3.0, th_cortical=0.2, th_subcutaneus=0.3, th_dermis=0.4)
    a = TuringModel_Rectangular()
    a.set_params(voxel_space=0.1, **kw)
    a.build(vol)
    b = TuringModel_Rectangular()
    b.set_params(voxel_space=0.3, dicom_path=str(tmp_path / "bone"), **kw)
    b.build()
    assert b.voxel_space == 0.1 and np.array_equal(a.voxel_model, b.voxel_model)
