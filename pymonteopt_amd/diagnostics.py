"""Measurement helpers run from the command line.

    python -m pymonteopt_amd.diagnostics copy     # 3 copies of 256 MiB (counter calibration)
    python -m pymonteopt_amd.diagnostics peaks    # atomic-add rates (f64/f32/u64/u32, per-XCD
                                                  # copies) and the FP32 FMA issue rate

`copy` gives rocprofv3's FETCH_SIZE / WRITE_SIZE a kernel of known traffic (read 256 MiB,
write 256 MiB per dispatch), so walk-kernel counts can be scaled to bytes.
"""
from __future__ import annotations

import ctypes as C
import json
import sys


def copy_calibration(mib=256, reps=3):
    import torch
    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"bytes_read_per_dispatch": n * 4, "bytes_written_per_dispatch": n * 4, "dispatches": reps}


def peaks():
    import torch
    from . import _abi
    L = _abi.lib()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        fn()
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) * 1e-3

    out = {}
    blocks, iters = 4096, 256
    count = blocks * 256 * iters
    kinds = (("f64", _abi.PMO_PROBE_F64, 8), ("f32", _abi.PMO_PROBE_F32, 4), ("u64", _abi.PMO_PROBE_U64, 8),
             ("u32", _abi.PMO_PROBE_U32, 4))
    # footprints past the 256 MB Infinity Cache (MALL) as well: the device's atomics slow there
    for mib in (0.25, 2, 32, 256, 512, 1024, 2048, 4096):
        for name, kind, size in kinds:
            if name != "f64" and mib not in (2, 32):
                continue
            buf = torch.zeros(int(mib * 1024 * 1024 // size), dtype={8: torch.float64, 4: torch.float32}[size],
                              device="cuda:0")
            t = timed(lambda: _abi.check(L.pmo_probe_atomics(C.c_void_p(buf.data_ptr()), buf.numel(), blocks, iters,
                                                             kind, s)))
            out[f"{name}_atomics_per_s_{mib}MiB"] = count / t
        if mib in (2, 32):
            # 8 copies of `mib` MiB, each workgroup adding into the copy of its own XCD, at
            # device and at workgroup scope
            n = int(mib * 1024 * 1024 // 8)
            for tag, kind in (("", _abi.PMO_PROBE_F64_PER_XCD), ("_workgroup_scope", _abi.PMO_PROBE_F64_PER_XCD_WG)):
                buf = torch.zeros(8 * n, dtype=torch.float64, device="cuda:0")
                t = timed(lambda: _abi.check(L.pmo_probe_atomics(C.c_void_p(buf.data_ptr()), n, blocks, iters,
                                                                 kind, s)))
                out[f"f64_per_xcd_copy{tag}_atomics_per_s_{mib}MiB"] = count / t
                torch.cuda.synchronize()
                # every add of 1.0 must have landed: a lost update would show as a short sum
                out[f"f64_per_xcd_copy{tag}_sum_ok_{mib}MiB"] = bool(float(buf.sum()) == 2.0 * count)
            # each wave's adds on eight consecutive 64 B lines
            buf = torch.zeros(n, dtype=torch.float64, device="cuda:0")
            t = timed(lambda: _abi.check(L.pmo_probe_atomics(C.c_void_p(buf.data_ptr()), n, blocks, iters,
                                                             _abi.PMO_PROBE_F64_LINES, s)))
            out[f"f64_wave_on_8_lines_atomics_per_s_{mib}MiB"] = count / t
            torch.cuda.synchronize()
            out[f"f64_wave_on_8_lines_sum_ok_{mib}MiB"] = bool(float(buf.sum()) == 2.0 * count)
            # lanes in groups of 4 / 8 / 16 on their own random 32 / 64 / 128 B blocks: which
            # block size one atomic request covers
            for tag, kind in (("32B", _abi.PMO_PROBE_F64_BLOCKS_32B), ("64B", _abi.PMO_PROBE_F64_BLOCKS_64B),
                              ("128B", _abi.PMO_PROBE_F64_BLOCKS_128B)):
                buf.zero_()
                t = timed(lambda: _abi.check(L.pmo_probe_atomics(C.c_void_p(buf.data_ptr()), n, blocks, iters,
                                                                 kind, s)))
                out[f"f64_lanes_on_random_{tag}_blocks_atomics_per_s_{mib}MiB"] = count / t
                torch.cuda.synchronize()
                out[f"f64_lanes_on_random_{tag}_blocks_sum_ok_{mib}MiB"] = bool(float(buf.sum()) == 2.0 * count)
    o = torch.empty(2048 * 256, dtype=torch.float32, device="cuda:0")
    t = timed(lambda: _abi.check(L.pmo_probe_valu(C.c_void_p(o.data_ptr()), 2048, 8192, s)))
    out["valu_fp32_fma_lane_ops_per_s"] = 2048 * 256.0 * 8192 * 32 / t
    # the walk's draws: Philox4x32-10 blocks per second with every lane active
    w = torch.empty(2048 * 256, dtype=torch.int32, device="cuda:0")
    t = timed(lambda: _abi.check(L.pmo_probe_philox(C.c_void_p(w.data_ptr()), 2048, 1024, s)))
    out["philox4x32_10_blocks_per_s"] = 2048 * 256.0 * 1024 / t
    out["philox_fma_slots_per_block"] = out["valu_fp32_fma_lane_ops_per_s"] / out["philox4x32_10_blocks_per_s"]
    return out


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "peaks"
    res = copy_calibration() if what == "copy" else peaks()
    print(json.dumps(res), flush=True)
