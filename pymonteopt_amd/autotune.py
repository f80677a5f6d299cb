"""Launch-shape autotuner for the walk kernels.

Times short runs of a workload with HIP events over a grid of block sizes and
packets-per-lane (plain walk) or schedule batch sizes (batched sweep) and keeps the
fastest. Results are cached per (device, workload signature) for the life of the process
and, when a path is given, in a JSON file.
"""
from __future__ import annotations

import json
import os

import torch

_CACHE: dict = {}


def _time(fn, reps=2):
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()                                   # warm-up (code-object load, caches)
    best = float("inf")
    for _ in range(reps):
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
    return best


def _load(path):
    if path and os.path.exists(path):
        try:
            return json.load(open(path))
        except (OSError, ValueError):
            return {}
    return {}


def _store(path, key, value):
    if not path:
        return
    d = _load(path)
    d[key] = value
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(d, f, indent=1)
    os.replace(tmp, path)


def tune_walk(walker, source, spec, n_packets, *, sample=2_000_000, blocks=(128, 256, 512, 1024),
              ppls=(1, 2, 4, 8), seed=0, cache_path=None, **walk_kw):
    """Best (block_size, packets_per_lane) for `walker.walk` on this workload.

    The sample keeps the workload's packets-per-resident-wave regime: runs of n_packets
    are timed when they are small, else `sample` packets."""
    dev = torch.cuda.get_device_name(walker.device)
    key = f"walk|{dev}|{walker.shape}|{walker.pitch}|{walker.optics.cpu().numpy().tobytes().hex()[:64]}|" \
          f"{source.kind}|{int(n_packets)}|{spec.absorbed}"
    if key in _CACHE:
        return _CACHE[key]
    disk = _load(cache_path)
    if key in disk:
        _CACHE[key] = disk[key]
        return disk[key]
    n = int(min(n_packets, sample))
    tal = walker.tallies(spec)
    table = []
    for bs in blocks:
        for ppl in ppls:
            def run():
                tal.zero()
                walker.walk(source, n, seed, tal, block_size=bs, packets_per_lane=ppl, **walk_kw)
            t = _time(run)
            table.append({"block_size": bs, "packets_per_lane": ppl, "seconds": t, "photons_per_s": n / t})
    best = min(table, key=lambda r: r["seconds"])
    res = {"block_size": best["block_size"], "packets_per_lane": best["packets_per_lane"],
           "photons_per_s": best["photons_per_s"], "sample_packets": n, "table": table}
    _CACHE[key] = res
    _store(cache_path, key, res)
    return res


def tune_sweep(sweep, *, blocks=(128, 256, 512, 1024), batches=(64, 128, 256, 512, 1024), quorums=(32,),
               refills=(4,), seed=0, cache_path=None):
    """Best (block_size, batch, quorum, refill) for a PlateSweep, searched one after the
    other (block and batch together, then the quorum at the best pair, then the refill
    threshold); leaves the sweep set to the winner."""
    dev = torch.cuda.get_device_name(sweep.device)
    sig = tuple((tuple(p.thickness), tuple(p.ma), tuple(p.ms), p.n_packets) for p in sweep.plates[:8])
    key = f"sweep|{dev}|{sweep.n_inst}|{sweep.total_packets}|{sweep.layer_step}|{hash(sig)}"
    if key in _CACHE:
        res = _CACHE[key]
    else:
        res = _load(cache_path).get(key)
    if res is None:
        table = []
        orig_batch, orig_q, orig_r = sweep.batch, sweep.quorum, sweep.refill
        q0, r0 = quorums[0], refills[0]
        sweep.quorum, sweep.refill = q0, r0
        for b in batches:
            sweep.set_batch(b)
            for bs in blocks:
                t = _time(lambda: sweep.run(seed, block_size=bs), reps=1)
                table.append({"block_size": bs, "batch": b, "quorum": q0, "refill": r0, "seconds": t,
                              "photons_per_s": sweep.total_packets / t})
        best = min(table, key=lambda r: r["seconds"])
        sweep.set_batch(best["batch"])
        for q in quorums[1:]:
            sweep.quorum = q
            t = _time(lambda: sweep.run(seed, block_size=best["block_size"]), reps=1)
            table.append({"block_size": best["block_size"], "batch": best["batch"], "quorum": q, "refill": r0,
                          "seconds": t, "photons_per_s": sweep.total_packets / t})
        best = min(table, key=lambda r: r["seconds"])
        sweep.quorum = best["quorum"]
        for rf in refills[1:]:
            sweep.refill = rf
            t = _time(lambda: sweep.run(seed, block_size=best["block_size"]), reps=1)
            table.append({"block_size": best["block_size"], "batch": best["batch"], "quorum": best["quorum"],
                          "refill": rf, "seconds": t, "photons_per_s": sweep.total_packets / t})
        best = min(table, key=lambda r: r["seconds"])
        sweep.set_batch(orig_batch)
        sweep.quorum, sweep.refill = orig_q, orig_r
        res = {"block_size": best["block_size"], "batch": best["batch"], "quorum": best["quorum"],
               "refill": best["refill"], "photons_per_s": best["photons_per_s"], "table": table}
        _CACHE[key] = res
        _store(cache_path, key, res)
    sweep.set_batch(res["batch"])
    sweep.block_size = res["block_size"]
    sweep.quorum = res.get("quorum", 0)
    sweep.refill = res.get("refill", 0)
    return res
