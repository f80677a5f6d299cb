"""The C ABI library: loads without a GPU, exports every symbol include/pmo.h declares,
its ctypes mirror matches the C layout, and argument errors come back as codes."""
import ctypes as C
import os
import re
import subprocess

import pytest

from pymonteopt_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pmo.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pmo_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _abi.lib()
    names = declared_functions()
    assert "pmo_walk" in names and "pmo_philox" in names
    for n in names:
        assert hasattr(L, n), n
        assert n in _abi.SIGNATURES, n
    assert L.pmo_abi_version() == _abi.ABI_VERSION == 11
    assert L.pmo_workspace_bytes() >= 8


PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "pmo.h"
#define S(t) printf(#t " %zu\n", sizeof(t));
#define O(t, f) printf(#t "." #f " %zu\n", offsetof(t, f));
int main(void) {
  S(pmo_volume) S(pmo_source) S(pmo_run) S(pmo_tallies) S(pmo_plate_instance) S(pmo_batch_tallies) S(pmo_diff) S(pmo_obd)
  O(pmo_diff, mismatch) O(pmo_plate_instance, n_packets) O(pmo_plate_instance, layers) O(pmo_run, key_dev) O(pmo_run, quorum)
  O(pmo_volume, optics) O(pmo_volume, skip) O(pmo_volume, label_radius) O(pmo_source, w0) O(pmo_source, add) O(pmo_source, states_stride)
  O(pmo_run, packets_per_lane) O(pmo_tallies, exit_state)
  O(pmo_tallies, ring_w) O(pmo_tallies, line_z) O(pmo_tallies, fluence_rz) O(pmo_run, flags) O(pmo_run, window)
  O(pmo_tallies, exit_add) O(pmo_tallies, exit_p) O(pmo_tallies, absorbed_lines) O(pmo_run, refill)
  O(pmo_diff, grad_lines)
  printf("lines_3_5_4 %llu\n", PMO_ABSORBED_LINES_DOUBLES(3, 5, 4));
  O(pmo_obd, stream) O(pmo_obd, side)
  printf("PMO_S_COUNT %d\nPMO_EV_COUNT %d\nPMO_OBD_STEP_FIELDS %d\n", PMO_S_COUNT, PMO_EV_COUNT, PMO_OBD_STEP_FIELDS);
  printf("probes %d,%d,%d,%d,%d,%d,%d,%d,%d,%d\n", PMO_PROBE_F64, PMO_PROBE_F32, PMO_PROBE_U64, PMO_PROBE_U32,
         PMO_PROBE_F64_PER_XCD, PMO_PROBE_F64_PER_XCD_WG, PMO_PROBE_F64_LINES, PMO_PROBE_F64_BLOCKS_32B,
         PMO_PROBE_F64_BLOCKS_64B, PMO_PROBE_F64_BLOCKS_128B);
  return 0;
}
"""


def test_ctypes_layout_matches_header(tmp_path):
    src = tmp_path / "probe.c"
    exe = tmp_path / "probe"
    src.write_text(PROBE)
    subprocess.run(["gcc", "-std=c11", f"-I{os.path.dirname(HEADER)}", str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                              check=True).stdout.strip().split("\n"))
    for name, cls in (("pmo_volume", _abi.PmoVolume), ("pmo_source", _abi.PmoSource),
                      ("pmo_run", _abi.PmoRun), ("pmo_tallies", _abi.PmoTallies),
                      ("pmo_plate_instance", _abi.PmoPlateInstance), ("pmo_batch_tallies", _abi.PmoBatchTallies),
                      ("pmo_diff", _abi.PmoDiff), ("pmo_obd", _abi.PmoObd)):
        assert int(out[name]) == C.sizeof(cls), name
    assert int(out["pmo_volume.optics"]) == _abi.PmoVolume.optics.offset
    assert int(out["pmo_volume.skip"]) == _abi.PmoVolume.skip.offset
    assert int(out["pmo_volume.label_radius"]) == _abi.PmoVolume.label_radius.offset
    assert int(out["pmo_source.w0"]) == _abi.PmoSource.w0.offset
    assert int(out["pmo_source.add"]) == _abi.PmoSource.add.offset
    assert int(out["pmo_source.states_stride"]) == _abi.PmoSource.states_stride.offset
    assert int(out["pmo_run.packets_per_lane"]) == _abi.PmoRun.packets_per_lane.offset
    assert int(out["pmo_run.flags"]) == _abi.PmoRun.flags.offset
    assert int(out["pmo_run.window"]) == _abi.PmoRun.window.offset
    assert int(out["pmo_run.key_dev"]) == _abi.PmoRun.key_dev.offset
    assert int(out["pmo_run.quorum"]) == _abi.PmoRun.quorum.offset
    assert int(out["pmo_diff.mismatch"]) == _abi.PmoDiff.mismatch.offset
    assert int(out["pmo_plate_instance.n_packets"]) == _abi.PmoPlateInstance.n_packets.offset
    assert int(out["pmo_plate_instance.layers"]) == _abi.PmoPlateInstance.layers.offset
    assert int(out["pmo_tallies.exit_state"]) == _abi.PmoTallies.exit_state.offset
    assert int(out["pmo_tallies.ring_w"]) == _abi.PmoTallies.ring_w.offset
    assert int(out["pmo_tallies.line_z"]) == _abi.PmoTallies.line_z.offset
    assert int(out["pmo_tallies.fluence_rz"]) == _abi.PmoTallies.fluence_rz.offset
    assert int(out["pmo_tallies.exit_add"]) == _abi.PmoTallies.exit_add.offset
    assert int(out["pmo_tallies.exit_p"]) == _abi.PmoTallies.exit_p.offset
    assert int(out["pmo_tallies.absorbed_lines"]) == _abi.PmoTallies.absorbed_lines.offset
    assert int(out["pmo_run.refill"]) == _abi.PmoRun.refill.offset
    assert int(out["pmo_diff.grad_lines"]) == _abi.PmoDiff.grad_lines.offset
    assert int(out["lines_3_5_4"]) == _abi.absorbed_lines_doubles(3, 5, 4) == 8 * 2 * 3 * 2
    assert int(out["PMO_S_COUNT"]) == _abi.PMO_S_COUNT
    assert [int(x) for x in out["probes"].split(",")] == [
        _abi.PMO_PROBE_F64, _abi.PMO_PROBE_F32, _abi.PMO_PROBE_U64, _abi.PMO_PROBE_U32, _abi.PMO_PROBE_F64_PER_XCD,
        _abi.PMO_PROBE_F64_PER_XCD_WG, _abi.PMO_PROBE_F64_LINES, _abi.PMO_PROBE_F64_BLOCKS_32B,
        _abi.PMO_PROBE_F64_BLOCKS_64B, _abi.PMO_PROBE_F64_BLOCKS_128B]
    assert int(out["pmo_obd.stream"]) == _abi.PmoObd.stream.offset
    assert int(out["pmo_obd.side"]) == _abi.PmoObd.side.offset
    assert int(out["PMO_OBD_STEP_FIELDS"]) == _abi.PMO_OBD_STEP_FIELDS
    assert int(out["PMO_EV_COUNT"]) == _abi.PMO_EV_COUNT


def _valid_structs():
    v = _abi.PmoVolume(labels=1, nx=5, ny=5, nz=5, n_labels=1, pitch=0.1, optics=1)
    s = _abi.PmoSource(kind=0, ix=2, iy=2, iz=1, radius=0.0, w0=1.0)
    s.dir[2] = 1.0
    r = _abi.PmoRun(n_packets=10, seed=1, wth=1e-4, roulette_m=10.0)
    t = _abi.PmoTallies(scalars=1)
    return v, s, r, t


@pytest.mark.parametrize("field,value,msg", [
    ("labels", None, "labels"), ("n_labels", 0, "n_labels"), ("pitch", 0.0, "pitch"),
    ("ix", 0, "launch voxel"), ("kind", 9, "beam kind"), ("kind", 3, "per-packet launch states"),
    ("block_size", 100, "block_size"),
    ("roulette_m", 0.5, "roulette_m"), ("scalars", None, "scalars"), ("nx", 2, "extents"),
    ("skip", 1, "label_radius"),
])
def test_argument_errors_are_codes(field, value, msg):
    # every check below runs on the host before any HIP call
    L = _abi.lib()
    v, s, r, t = _valid_structs()
    for obj in (v, s, r, t):
        if field in [f[0] for f in obj._fields_]:
            setattr(obj, field, value)
    rc = L.pmo_walk(C.byref(v), C.byref(s), C.byref(r), C.byref(t), C.c_void_p(1), None)
    assert rc in (_abi.PMO_ERR_INVALID, _abi.PMO_ERR_LABEL)
    assert msg in L.pmo_last_error().decode()
    with pytest.raises(ValueError):
        _abi.check(rc)


def test_pack_label_radius_argument_errors():
    # checked on the host before any HIP call
    L = _abi.lib()
    assert L.pmo_pack_label_radius(None, C.c_void_p(1), C.c_void_p(1), 5, None) == _abi.PMO_ERR_INVALID
    assert L.pmo_pack_label_radius(C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), -1, None) == _abi.PMO_ERR_INVALID
    assert L.pmo_pack_label_radius(None, None, None, 0, None) == _abi.PMO_OK


@pytest.mark.parametrize("args,msg", [
    ((None, 5, 5, 5, 3, 2, 3), "non-null"), ((1, 5, 5, 5, 0, 2, 3), "cap"), ((1, 5, 5, 5, 128, 2, 3), "cap"),
    ((1, 0, 5, 5, 3, 2, 3), "extents"), ((1, 2048, 1024, 1024, 3, 2, 3), "exceeds"),
    ((1, 5, 5, 5, 3, 2, 2), "distinct"), ((1, 5, 5, 5, 3, 1, 2), "distinct"),
])
def test_skip_radius_argument_errors(args, msg):
    # checked on the host before any HIP call
    L = _abi.lib()
    lab, nx, ny, nz, cap, out, scratch = args
    with pytest.raises((ValueError, _abi.PmoError), match=msg):
        _abi.check(L.pmo_skip_radius(C.c_void_p(lab) if lab else None, nx, ny, nz, cap, C.c_void_p(out),
                                     C.c_void_p(scratch), None))


def test_philox_argument_errors():
    L = _abi.lib()
    assert L.pmo_philox(None, 0, 0, None, -1, None) == _abi.PMO_ERR_INVALID
    assert L.pmo_philox(None, 0, 0, None, 0, None) == _abi.PMO_OK
    assert L.pmo_philox(None, 0, 0, None, 4, None) == _abi.PMO_ERR_INVALID


def test_c_example_builds_against_the_header(tmp_path):
    """tests/c/walk_slab.c uses include/pmo.h and libpmo.so only (C11, no C++ or torch): it
    compiles with gcc and links against every symbol it calls."""
    import subprocess
    root = ROOT
    out = tmp_path / "walk_slab"
    cmd = ["gcc", "-O2", "-std=c11", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", f"-I{root}/include",
           "-I/opt/rocm/include", os.path.join(ROOT, "tests", "c", "walk_slab.c"), "-o", str(out),
           f"-L{root}/pymonteopt_amd", "-lpmo", "-L/opt/rocm/lib", "-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert out.exists()
