"""Throughput of the MI355X voxel photon walk on the 3-layer skin plate.

One step = one walk of 1e8 packets (FP32 walk, FP64 tallies: Rd, Tt, Rd(r) 40 bins, the
per-voxel absorbed-weight/fluence map, the energy ledger) through the 201x201x140 @0.05 mm
3-layer skin plate, pencil beam. The geometry is uploaded before the timed region; the
timed region runs K walks (seeds 0..K-1) back to back. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "photons/s, 3-layer skin 1e8 ph; batched & adjoint vs plain, same GPU, 3σ"


def cpu_baseline(wl, budget_s=12.0):
    """The oracle (C restatement of the reference walk, float32 state like the reference's
    float32 arrays) on all host cores, on a bounded sample of the same workload."""
    from oracle import oracle as og
    og.build()
    ma, ms, g, n = wl.optics_columns()
    threads = len(os.sched_getaffinity(0))
    probe = 2000
    t0 = time.perf_counter()
    og.walk(wl.labels, wl.pitch, ma, ms, g, n, wl.source.as_dict(), probe, 12345, precision="f32",
            threads=threads, nr=40, dr=0.1)
    rate = probe / (time.perf_counter() - t0)
    n_s = int(max(probe, min(5e6, rate * budget_s)))
    t0 = time.perf_counter()
    og.walk(wl.labels, wl.pitch, ma, ms, g, n, wl.source.as_dict(), n_s, 777, precision="f32",
            threads=threads, nr=40, dr=0.1)
    dt = time.perf_counter() - t0
    return {"value": n_s / dt, "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": f"{n_s} packets of the same workload (f32 oracle, absorbed + Rd(r) tallies), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--packets", type=float, default=1e8)
    ap.add_argument("--beam", default="pencil", choices=["pencil", "flat"])
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--ppl", type=int, default=0)
    ap.add_argument("--fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-absorbed", action="store_true")
    ap.add_argument("--window", default="0,0,0", help="LDS tally window x,y,z (0 = default, -1 = none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported in this build "
              "(the walk shards by packet-id range; see the multi-GPU paragraph of DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

    from pymonteopt_amd import presets, roofline
    from pymonteopt_amd.runtime import TallySpec, Walker

    dev = torch.device("cuda:0")
    wl = presets.skin_plate(args.beam)
    wk = Walker(wl.labels, wl.pitch, wl.optics, device=dev)
    spec = TallySpec(**wl.tallies)
    if args.no_absorbed:
        spec.absorbed = False
    tal = wk.tallies(spec)
    N = int(args.packets)
    kw = dict(block_size=args.block, packets_per_lane=args.ppl, fp64=args.fp64,
              window=tuple(int(x) for x in args.window.split(",")))

    for i in range(args.warmup):
        tal.zero()
        wk.walk(wl.source, N, 1000 + i, tal, **kw)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream(dev)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    events_tot = None
    scal = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        tal.zero()
        evs[k][0].record(stream)
        wk.walk(wl.source, N, k, tal, **kw)
        evs[k][1].record(stream)
        # device-side copies of this step's scalars and event counters, read after timing
        ev = tal.events.clone()
        sc = tal.scalars.clone()
        scal.append((sc, ev))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    kernel_s = [a.elapsed_time(b) * 1e-3 for a, b in evs]
    walk_s = sum(kernel_s)
    ms_per_step = 1e3 * wall / args.steps
    value = N * args.steps / wall

    from pymonteopt_amd import _abi
    ev_sum = {k: 0 for k in _abi.EVENT_NAMES}
    rd, tt, ab, led = [], [], [], []
    for sc, ev in scal:
        s = dict(zip(_abi.SCALAR_NAMES, sc.cpu().tolist()))
        e = dict(zip(_abi.EVENT_NAMES, [int(x) for x in ev.cpu().tolist()]))
        for k in ev_sum:
            ev_sum[k] += e[k]
        rd.append(s["Rd"] / N)
        tt.append(s["Tt"] / N)
        ab.append(s["A"] / N)
        bal = s["Rd"] + s["Tt"] + s["other"] + s["A"] + s["kill"] - s["gain"]
        led.append(abs(s["w0"] - bal) / s["w0"])

    ops = roofline.ops_per_packet(ev_sum)
    mean_launch_s = walk_s / args.steps
    achieved = N * ops / mean_launch_s
    bytes_pp = roofline.algorithmic_bytes_per_packet(ev_sum, absorbed=spec.absorbed)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.fp64 else "f32",
        "data": "synthetic",
        "config": {
            "workload": wl.name + f", {N:.0e} packets per step",
            "packets_per_step": N,
            "tallies": "Rd, Tt, Rd(r) 40x0.1mm, absorbed/fluence map (FP64)" if spec.absorbed
            else "Rd, Tt, Rd(r) 40x0.1mm (FP64)",
            "walk_kernel_ms_mean": 1e3 * mean_launch_s,
            "events_per_packet": {k: ev_sum[k] / (N * args.steps) for k in ("interactions", "crossings", "fresnel",
                                                                              "roulette")},
            "Rd": [float(sum(rd) / len(rd)), float(_se(rd))],
            "Tt": [float(sum(tt) / len(tt)), float(_se(tt))],
            "A": [float(sum(ab) / len(ab)), float(_se(ab))],
            "energy_rel_err_max": max(led),
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved / 1e12,
            "peak": roofline.PEAK_VALU_LANE_OPS / 1e12,
            "unit": "Tlane-op/s",
            "frac": achieved / roofline.PEAK_VALU_LANE_OPS,
            "traffic": None,
            "ops_per_packet": ops,
            "hbm_algorithmic_GBps": N * bytes_pp / mean_launch_s / 1e9,
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl)
    print(json.dumps(out), flush=True)


def _se(x):
    n = len(x)
    if n < 2:
        return float("nan")
    m = sum(x) / n
    return math.sqrt(sum((v - m) ** 2 for v in x) / (n - 1) / n)


if __name__ == "__main__":
    main()
