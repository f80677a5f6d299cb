"""DICOM slices for the bone volumes, without pydicom.

The reference writes each z-slice of a Turing-pattern volume as one DICOM file
(TuringPattern._write_dicom, pymopt/modeling_gpu/_classes.py:185-222). It reads such stacks
back in TuringModel_*._read_dicom (pymopt/voxel/_classes.py:577-586) and in
`readDicom` (pymopt/utils/readDICOM.py:17-44). Both go through pydicom, which this package
does not use.

This module reads and writes the DICOM subset those files use:
* a Part-10 file: a 128-byte preamble, "DICM", then the file-meta group in explicit-VR
  little endian;
* a data set in implicit- or explicit-VR little endian, uncompressed;
* monochrome pixels of 8 or 16 bits.

Sequences of undefined length and compressed transfer syntaxes are refused with a
ValueError.
"""
from __future__ import annotations

import datetime
import os
import struct
import uuid

import numpy as np

__all__ = ["write_slice", "write_stack", "read_file", "read_stack", "readDicom", "reConstArray_8", "reConstArray"]

IMPLICIT_LE = "1.2.840.10008.1.2"
EXPLICIT_LE = "1.2.840.10008.1.2.1"
_LONG_VRS = {b"OB", b"OW", b"OF", b"SQ", b"UT", b"UN", b"OD", b"OL", b"UC", b"UR", b"OV"}


def _uid():
    return "2.25." + str(uuid.uuid4().int)


def _pad_even(b, fill=b" "):
    return b + fill if len(b) % 2 else b


def _elem_explicit(group, elem, vr, value):
    if vr in ("OB", "OW"):
        v = value if isinstance(value, bytes) else bytes(value)
        v = _pad_even(v, b"\0")
        return struct.pack("<HH2sHI", group, elem, vr.encode(), 0, len(v)) + v
    v = _pad_even(value.encode(), b"\0" if vr == "UI" else b" ") if isinstance(value, str) else value
    return struct.pack("<HH2sH", group, elem, vr.encode(), len(v)) + v


def _elem_implicit(group, elem, value):
    v = value
    if isinstance(v, str):
        v = _pad_even(v.encode())
    v = _pad_even(v, b"\0")
    return struct.pack("<HHI", group, elem, len(v)) + v


def write_slice(filename, pixel_array, voxelsize, modality="WSD"):
    """One slice as TuringPattern._write_dicom writes it: implicit-VR little endian,
    16-bit unsigned monochrome, Columns = shape[0], Rows = shape[1], PixelSpacing
    (voxelsize, voxelsize), pixel bytes in the array's C order."""
    a = np.ascontiguousarray(pixel_array, dtype=np.uint16)
    if a.ndim != 2:
        raise ValueError("a slice is a 2-D array")
    sop_class = _uid()
    sop_inst = _uid()
    meta = b"".join([
        _elem_explicit(0x0002, 0x0001, "OB", b"\x00\x01"),
        _elem_explicit(0x0002, 0x0002, "UI", sop_class),
        _elem_explicit(0x0002, 0x0003, "UI", sop_inst),
        _elem_explicit(0x0002, 0x0010, "UI", IMPLICIT_LE),
        _elem_explicit(0x0002, 0x0012, "UI", _uid()),
    ])
    meta = _elem_explicit(0x0002, 0x0000, "UL", struct.pack("<I", len(meta))) + meta
    us = lambda x: struct.pack("<H", int(x))
    ds = b"".join([
        _elem_implicit(0x0008, 0x0016, sop_class + "\0" if len(sop_class) % 2 else sop_class),
        _elem_implicit(0x0008, 0x0018, sop_inst + "\0" if len(sop_inst) % 2 else sop_inst),
        _elem_implicit(0x0008, 0x0023, str(datetime.date.today()).replace("-", "")),
        _elem_implicit(0x0008, 0x0033, "010101.000000"),
        _elem_implicit(0x0008, 0x0060, modality),
        _elem_implicit(0x0020, 0x000D, _uid()),
        _elem_implicit(0x0020, 0x000E, _uid()),
        _elem_implicit(0x0028, 0x0002, us(1)),
        _elem_implicit(0x0028, 0x0004, "MONOCHROME2"),
        _elem_implicit(0x0028, 0x0010, us(a.shape[1])),
        _elem_implicit(0x0028, 0x0011, us(a.shape[0])),
        _elem_implicit(0x0028, 0x0030, "%s\\%s" % (repr(float(voxelsize)), repr(float(voxelsize)))),
        _elem_implicit(0x0028, 0x0100, us(16)),
        _elem_implicit(0x0028, 0x0101, us(16)),
        _elem_implicit(0x0028, 0x0102, us(15)),
        _elem_implicit(0x0028, 0x0103, us(0)),
        _elem_implicit(0x0028, 0x0107, us(a.max(initial=0))),
        _elem_implicit(0x7FE0, 0x0010, a.tobytes()),
    ])
    with open(filename, "wb") as fp:
        fp.write(b"\0" * 128 + b"DICM" + meta + ds)


def write_stack(path, volume, voxelsize):
    """A label volume [n0][n1][n2] as files path/0000.dcm ... (one per index of axis 0)
    that read_stack returns unchanged. With the reference's Columns = shape[0], Rows =
    shape[1] convention, each slice is handed over as its bytes viewed (n2, n1)."""
    v = np.ascontiguousarray(volume)
    if v.ndim != 3 or v.min(initial=0) < 0 or v.max(initial=0) > 65535:
        raise ValueError("a 3-D volume of values in [0, 65535]")
    os.makedirs(path, exist_ok=True)
    for k in range(v.shape[0]):
        write_slice(os.path.join(path, f"{k:04}.dcm"), v[k].astype(np.uint16).reshape(v.shape[2], v.shape[1]),
                    voxelsize)


class DicomSlice:
    """The elements read from one file, by (group, element), plus the decoded attributes
    the bone models use."""

    def __init__(self, elements, syntax):
        self.elements = elements
        self.transfer_syntax = syntax

    def _us(self, tag):
        return struct.unpack("<H", self.elements[tag][:2])[0]

    @property
    def Rows(self):
        return self._us((0x0028, 0x0010))

    @property
    def Columns(self):
        return self._us((0x0028, 0x0011))

    @property
    def PixelSpacing(self):
        raw = self.elements.get((0x0028, 0x0030), b"1\\1").decode().strip("\0 ")
        return [float(x) for x in raw.split("\\")]

    @property
    def pixel_array(self):
        bits = self._us((0x0028, 0x0100)) if (0x0028, 0x0100) in self.elements else 16
        signed = (0x0028, 0x0103) in self.elements and self._us((0x0028, 0x0103)) == 1
        dt = {8: np.int8 if signed else np.uint8, 16: np.int16 if signed else np.uint16}.get(bits)
        if dt is None:
            raise ValueError("only 8- and 16-bit pixels are supported")
        data = self.elements[(0x7FE0, 0x0010)]
        n = self.Rows * self.Columns
        return np.frombuffer(data, dtype=np.dtype(dt).newbyteorder("<"), count=n).reshape(self.Rows, self.Columns)


def _parse(buf, pos, explicit, stop_group=None):
    out = {}
    end = len(buf)
    while pos + 8 <= end:
        group, elem = struct.unpack_from("<HH", buf, pos)
        if stop_group is not None and group != stop_group:
            break
        if explicit:
            vr = buf[pos + 4:pos + 6]
            if vr in _LONG_VRS:
                length = struct.unpack_from("<I", buf, pos + 8)[0]
                pos += 12
            else:
                length = struct.unpack_from("<H", buf, pos + 6)[0]
                pos += 8
        else:
            length = struct.unpack_from("<I", buf, pos + 4)[0]
            pos += 8
        if length == 0xFFFFFFFF:
            raise ValueError("elements of undefined length (sequences) are not supported")
        out[(group, elem)] = bytes(buf[pos:pos + length])
        pos += length
    return out, pos


def read_file(filename, force_implicit=False):
    """Reads one uncompressed little-endian DICOM file (with or without the Part-10
    preamble). force_implicit parses the data set as implicit VR whatever the meta says,
    as the reference's readers force ImplicitVRLittleEndian."""
    with open(filename, "rb") as fp:
        buf = fp.read()
    pos, syntax = 0, IMPLICIT_LE
    if len(buf) >= 132 and buf[128:132] == b"DICM":
        meta, pos = _parse(buf, 132, explicit=True, stop_group=0x0002)
        syntax = meta.get((0x0002, 0x0010), IMPLICIT_LE.encode()).decode().strip("\0 ")
    if syntax not in (IMPLICIT_LE, EXPLICIT_LE):
        raise ValueError("transfer syntax %s is not supported (uncompressed little endian only)" % syntax)
    explicit = syntax == EXPLICIT_LE and not force_implicit
    elements, _ = _parse(buf, pos, explicit=explicit)
    return DicomSlice(elements, syntax)


def read_stack(path, ext_name=".dcm"):
    """Sorted slice files of a directory as an int8 stack [n_files][Rows][Columns] and the
    pixel spacing (TuringModel_*._read_dicom, pymopt/voxel/_classes.py:577-586)."""
    files = sorted(f for f in os.listdir(path) if ext_name in f.lower())
    if not files:
        raise ValueError("no %s files in %s" % (ext_name, path))
    slices = [read_file(os.path.join(path, f)) for f in files]
    return np.array([s.pixel_array for s in slices]).astype(np.int8), round(float(slices[0].PixelSpacing[0]), 5)


def readDicom(path, *, size_down=True, ext_name=".dcm"):
    """pymopt/utils/readDICOM.py:17-44: the stack as [Rows][Columns][n_files], its
    dimensions and spacing; size_down maps 16-bit values to int8 by round(x / 2^8)."""
    files = []
    for d, _, names in os.walk(path):
        files += [os.path.join(d, f) for f in names if ext_name in f.lower()]
    files.sort()
    ref = read_file(files[0], force_implicit=True)
    dims = (int(ref.Rows), int(ref.Columns), len(files))
    sp = ref.PixelSpacing
    spacing = (float(sp[0]), float(sp[1]), float(sp[1]))
    arr = np.zeros(dims, dtype=ref.pixel_array.dtype)
    for k, f in enumerate(files):
        arr[:, :, k] = read_file(f, force_implicit=True).pixel_array
    if size_down:
        arr = np.round(arr.reshape(-1, dims[2]) / 2 ** 8).astype("int8").reshape(dims)
    print("ConstPixelDims: %s" % str(dims))
    print("ConstPixelSpacing: %s" % str(spacing))
    return arr, dims, spacing


def reConstArray_8(ArrayDicom, threshold=9500):
    threshold = round(threshold / (1 + 2 ** 8))
    return np.where(ArrayDicom < threshold, 0, ArrayDicom)


def reConstArray(ArrayDicom, threshold=37):
    threshold = round(threshold)
    return np.where(ArrayDicom < threshold, 0, ArrayDicom)
