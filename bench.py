"""Throughput of the MI355X voxel photon walk on the 3-layer skin plate.

One step = one walk of 1e8 packets (FP32 walk, FP64 tallies: Rd, Tt, Rd(r) 40 bins, the
per-voxel absorbed-weight/fluence map, the energy ledger) through the 201x201x140 @0.05 mm
3-layer skin plate, pencil beam. The geometry is uploaded before the timed region; the
timed region runs K walks (seeds 0..K-1) back to back. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import contextlib
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "photons/s, 3-layer skin 1e8 ph; batched & adjoint vs plain, same GPU, 3σ"


def _cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or "unknown"."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, budget_s=15.0, max_packets=50_000_000):
    """The oracle (C restatement of the reference walk, float32 state like the reference's
    float32 arrays) on all host cores, on a bounded sample of the same workload: a probe
    sizes the sample so that the timed run takes about budget_s seconds."""
    from oracle import oracle as og
    og.build()
    ma, ms, g, n = wl.optics_columns()
    threads = len(os.sched_getaffinity(0))
    kw = dict(precision="f32", threads=threads, nr=40, dr=0.1)
    n_s, dt, seed = 2000, 0.0, 777
    while True:
        # each run's time includes a fixed set-up cost, so scaling by budget / dt undershoots;
        # repeat (growth capped at x64) until a run takes most of the budget
        t0 = time.perf_counter()
        og.walk(wl.labels, wl.pitch, ma, ms, g, n, wl.source.as_dict(), n_s, seed, **kw)
        dt = time.perf_counter() - t0
        if dt >= 0.6 * budget_s or n_s >= max_packets:
            break
        n_s = int(min(max_packets, n_s * min(64.0, budget_s / dt)))
        seed += 1
    # the CPU-runnable slab (101 x 101 x 50 @0.1 mm, mu_a 0.1, mu_s 10, g 0.9, n 1.37): its
    # 1e5-packet seed batches, as many as fit in about 3 s
    from pymonteopt_amd import presets
    slab = presets.slab_plate()
    sma, sms, sg, sn = slab.optics_columns()
    n1, t1, k1 = 0, 0.0, 0
    while t1 < 3.0 and k1 < 10:
        t0 = time.perf_counter()
        og.walk(slab.labels, slab.pitch, sma, sms, sg, sn, slab.source.as_dict(), 100_000, k1, precision="f32",
                threads=threads, nr=50, dr=0.1)
        t1 += time.perf_counter() - t0
        n1 += 100_000
        k1 += 1
    return {"value": n_s / dt, "unit": "photons/s", "cores": threads, "kind": "port", "cpu_model": _cpu_model(),
            "sample": f"{n_s} packets of the same workload (f32 oracle, absorbed + Rd(r) tallies), {dt:.1f} s",
            "slab_101x101x50": {"value": n1 / t1, "unit": "photons/s",
                                "sample": f"seeds 0-{k1 - 1} x 1e5 packets (f32 oracle, absorbed + Rd(r)), {t1:.1f} s"}}


def cpu_baseline_batched(n_inst=512, budget_s=10.0, fraction=0.02):
    """The oracle (f32 state, all host cores) walking the sweep's plates one after another,
    each with `fraction` of its budget, in sweep order until about budget_s seconds: the
    sequential CPU sweep at a reduced budget, as photons/s."""
    import numpy as np
    from oracle import oracle as og
    from pymonteopt_amd.batched import slab_sweep
    og.build()
    threads = len(os.sched_getaffinity(0))
    plates = slab_sweep(n_inst, seed=0)
    done, t, k = 0, 0.0, 0
    for p in plates:
        kind, vox, w0 = p.source()
        opt = p.optics().astype(np.float64)
        L = opt.shape[0] - 1
        n = max(1000, int(p.n_packets * fraction))
        src = dict(kind=kind, voxel=tuple(int(x) for x in vox), w0=float(w0), radius=0.0, dir=(0.0, 0.0, 1.0))
        t0 = time.perf_counter()
        og.walk(p.volume(), p.voxel_space, opt[:L, 0], opt[:L, 1], opt[:L, 2], opt[:, 3], src, n, 2,
                precision="f32", threads=threads, absorbed=False, nr=40, dr=0.1)
        t += time.perf_counter() - t0
        done += n
        k += 1
        if t >= budget_s:
            break
    return {"value": done / t, "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": f"the first {k} of the {n_inst} plates, {fraction:.0%} of each budget ({done} packets, "
                      f"f32 oracle, Rd(r) tallies), {t:.1f} s"}


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def batched_leg(n_inst=512):
    """512 random slabs (uneven budgets) in one graph-captured persistent launch, against
    512 sequential plain walks of the same plates with the same draws (so each instance's
    estimate, and its standard error, is the same in both). The plain walks run as the
    plain walk defaults (same-label faces crossed in one step) and, beside that, voxel by
    voxel. Reported for layer stepping (the sweep's default) and for voxel stepping, which
    reproduces the voxel-by-voxel plain walks exactly."""
    import numpy as np
    import torch
    from pymonteopt_amd import autotune
    from pymonteopt_amd.batched import PlateSweep, slab_sweep
    from pymonteopt_amd.geometry import Source
    from pymonteopt_amd.runtime import TallySpec, Walker
    plates = slab_sweep(n_inst, seed=0)
    N = sum(p.n_packets for p in plates)

    def graph_time(sw):
        sw.capture()
        sw.replay(1)
        torch.cuda.synchronize()
        e0, e1 = _events()
        e0.record()
        sw.replay(2)
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) * 1e-3, sw.results()

    # layer stepping (the default) and voxel stepping, each autotuned, graph-captured
    sw = PlateSweep(plates, nr=40, dr=0.1, batch=256)
    tuned = autotune.tune_sweep(sw, blocks=(128, 256), batches=(128, 256, 512), quorums=(32, 24, 48), refills=(4, 2, 8))
    tb, rb = graph_time(sw)
    swv = PlateSweep(plates, nr=40, dr=0.1, batch=256, layer_step=False)
    tuned_v = autotune.tune_sweep(swv, blocks=(128, 256), batches=(128, 256, 512), quorums=(32, 24, 48),
                                  refills=(4, 2, 8))
    tv, rv = graph_time(swv)
    spec = TallySpec(nr=40, dr=0.1, absorbed=False)

    def sequential(skip):
        # the plain walk of each plate, one launch after another; skip=None is the plain
        # walk's default (same-label faces crossed in one step), False voxel by voxel
        walkers = []
        for p in plates:
            kind, vox, w0 = p.source()
            walkers.append((Walker(p.volume(), p.voxel_space, p.optics(), skip=skip), Source(voxel=vox, w0=w0, kind=kind)))
        tals = [w.tallies(spec) for w, _ in walkers]
        torch.cuda.synchronize()
        e0, e1 = _events()
        e0.record()
        for i, ((w, s), t) in enumerate(zip(walkers, tals)):
            w.walk(s, plates[i].n_packets, 2, t, stream_id=i)
        e1.record()
        e1.synchronize()
        rd = np.array([t.to_host()["scalars"]["Rd"] / p.n_packets for t, p in zip(tals, plates)])
        return e0.elapsed_time(e1) * 1e-3, rd

    tp, rd_p = sequential(None)
    tpv, rd_pv = sequential(False)
    strip = lambda r: {k: v for k, v in r.items() if k != "table"}
    ev = rb["events"]
    per_packet = {k: ev[k] / max(ev["packets"], 1) for k in ("interactions", "crossings", "fresnel")}
    return {"instances": n_inst, "packets": N, "stepping": "layer", "batched_s": tb, "batched_photons_per_s": N / tb,
            "events_per_packet": per_packet,
            "sequential_plain_s": tp, "sequential_photons_per_s": N / tp, "speedup": tp / tb,
            "max_abs_Rd_diff_same_draws": float(np.max(np.abs(rd_p - rb["Rd"]))), "tuned": strip(tuned),
            "sequential_plain_voxel_stepping_s": tpv, "speedup_vs_voxel_stepping_plain": tpv / tb,
            "voxel_stepping": {"batched_s": tv, "batched_photons_per_s": N / tv, "speedup": tp / tv,
                               "max_abs_Rd_diff_same_draws_vs_voxel_stepping_plain": float(np.max(np.abs(rd_pv - rv["Rd"]))),
                               "tuned": strip(tuned_v)},
            "graph": True}


def adjoint_leg(n_packets=10_000_000):
    """Gradients of the differentiable walk at 1e7 packets against the plain walk of the
    same packets (as the plain walk defaults, and voxel by voxel, which the replay follows):
    * skin plate, per-label d Rd(r) / d(mu_a, mu_s) as rd_of_label_optics computes it: one
      forward with Jacobians (the backward is J^T g). The per-voxel gradient of the same
      plate (rd_of_absorption_field: a forward with the absorption field, then the replay)
      is reported beside it;
    * 64^3 4-label grid, per-voxel d Rd(r) / d mu_a as rd_of_absorption_field computes it:
      a forward with the absorption field and the replay adjoint.
    The replay's deposits outside its LDS box go either through cube line buffers or one FP64
    atomic per voxel deposit beside a larger box; DiffWalk's first adjoint times both and keeps
    the faster (the timed calls run after that choice). Both times are reported."""
    import numpy as np
    import torch
    from pymonteopt_amd import presets
    from pymonteopt_amd.adjoint import DiffWalk
    from pymonteopt_amd.runtime import TallySpec, Walker
    out = {}

    def timed(fn):
        fn()
        e0, e1 = _events()
        e0.record()
        r = fn()
        e1.record()
        e1.synchronize()
        return r, e0.elapsed_time(e1) * 1e-3

    def replay_modes(fwd):
        # the replay with each deposit path, then back to the chosen one
        chosen, t = dw.lines, {}
        for mode, key in ((True, "line_buffers"), (False, "per_voxel_atomics")):
            dw.lines = mode
            t[key] = timed(lambda: dw.adjoint(fwd, dw.area.clone()))[1]
        dw.lines = chosen
        return dict(t, chosen="line_buffers" if chosen else "per_voxel_atomics")

    for name, wl in (("skin", presets.skin_plate()), ("grf64", presets.grf_grid(seed=0))):
        dw = DiffWalk(wl.labels, wl.pitch, wl.optics, wl.source, n_packets=n_packets, nr=40, dr=0.1)
        wk = Walker(wl.labels, wl.pitch, wl.optics)
        wkv = Walker(wl.labels, wl.pitch, wl.optics, skip=False)
        tal = wk.tallies(TallySpec(nr=40, dr=0.1, absorbed=False))
        _, t_plain = timed(lambda: wk.walk(wl.source, n_packets, 3, tal))
        _, t_plain_v = timed(lambda: wkv.walk(wl.source, n_packets, 3, tal))
        mua = torch.tensor(wl.optics[:, 0], device="cuda:0")
        field = mua[torch.as_tensor(np.minimum(wl.labels, wl.n_labels), device="cuda:0").long()].float()
        if name == "skin":
            # rd_of_label_optics: one forward with Jacobians, crossing same-label cubes in one
            # step (no replay follows). The per-voxel gradient of the same plate, as
            # rd_of_absorption_field computes it: a voxel-stepping forward with the absorption
            # field and no Jacobians, then the replay
            _, t_fwd = timed(lambda: dw.forward(seed=3, jac=True, replay=False))
            _, t_fwd_v = timed(lambda: dw.forward(seed=3, jac=True))
            fwd, t_fwd_f = timed(lambda: dw.forward(seed=3, mua_field=field, jac=False))
            _, t_adj = timed(lambda: dw.adjoint(fwd, dw.area.clone()))
            t_grad = t_fwd
            rec = {"gradient": "per-label mu_a, mu_s: forward with Jacobians", "forward_with_jacobian_s": t_fwd,
                   "forward_with_jacobian_voxel_stepping_s": t_fwd_v,
                   "per_voxel_forward_s": t_fwd_f, "per_voxel_replay_s": t_adj,
                   "per_voxel_replay_modes_s": replay_modes(fwd),
                   "per_voxel_gradient_cost_vs_plain": (t_fwd_f + t_adj) / t_plain,
                   "per_voxel_gradient_cost_vs_its_forward": (t_fwd_f + t_adj) / t_fwd_f}
        else:
            fwd, t_fwd = timed(lambda: dw.forward(seed=3, mua_field=field, jac=False))
            _, t_adj = timed(lambda: dw.adjoint(fwd, dw.area.clone()))
            t_grad = t_fwd + t_adj
            rec = {"gradient": "per-voxel mu_a: forward + replay adjoint", "forward_s": t_fwd,
                   "adjoint_replay_s": t_adj, "adjoint_replay_modes_s": replay_modes(fwd)}
        rec.update({"plain_walk_s": t_plain, "gradient_s": t_grad, "gradient_cost_vs_plain": t_grad / t_plain,
                    "plain_walk_voxel_stepping_s": t_plain_v,
                    "gradient_cost_vs_voxel_stepping_plain": t_grad / t_plain_v,
                    "replay_mismatches": int(dw.last_mismatch)})
        out[name] = rec
    return out


def obd_leg(n_packets=10_000_000, sample=60000):
    """OBD receiver scan (260 focus steps, reference defaults) of the exit states of a 1e7
    packet walk of the skin plate, traced on the device straight from the walk's buffer.
    Also returns (not printed) a host copy of the first exits, for the CPU baseline."""
    import torch
    from pymonteopt_amd import presets
    from pymonteopt_amd.optics import OBD
    from pymonteopt_amd.runtime import TallySpec, Walker
    wl = presets.skin_plate()
    wk = Walker(wl.labels, wl.pitch, wl.optics)
    tal = wk.tallies(TallySpec(absorbed=False, exit_state=True), n_packets)
    wk.walk(wl.source, n_packets, 7, tal)
    o = OBD()
    o.trace_exit_states(tal.exit_state, seed=1)          # warm-up
    torch.cuda.synchronize()
    e0, e1 = _events()
    e0.record()
    I = o.trace_exit_states(tal.exit_state, seed=2)
    e1.record()
    e1.synchronize()
    t = e0.elapsed_time(e1) * 1e-3
    es = tal.exit_state.reshape(7, n_packets)
    n_rd = int((es[5] < 0).sum())
    steps = o.steps().size
    res = {"packets_walked": n_packets, "exits_traced": n_rd, "focus_steps": steps, "seconds": t,
           "traces_per_s": n_rd * steps / t, "peak_Z": float(o.steps()[int(torch.argmax(I))])}
    return res, (o.params, es[:, :4 * sample].cpu().numpy(), sample)


def cpu_baseline_obd(params, host_es, sample):
    """The oracle's numpy receiver (float64, one process) on `sample` exits toward the
    receiver from the same walk, over the same 260 focus steps."""
    import numpy as np
    from oracle import obd as ob
    h = host_es.astype(np.float64)
    q, d, wt, idx = ob.select(params, h[0:3], h[3:6], h[6])
    q, d, wt, idx = q[:, :sample], d[:, :sample], wt[:sample], idx[:sample]
    Zs = ob.steps(params)
    t0 = time.perf_counter()
    for i, Z in enumerate(Zs):
        ob.trace_step(params, Z, q, d, wt, u=ob.step_uniforms(i, idx, 2, 0))
    tc = time.perf_counter() - t0
    return {"value": wt.size * Zs.size / tc, "unit": "traces/s", "cores": 1, "kind": "port",
            "sample": f"{wt.size} exits x {Zs.size} focus steps, numpy float64, {tc:.1f} s"}


def turing_leg():
    """The Turing-pattern bone model at the reference's defaults: 520^3 grid, 100 steps
    (reaction kernel, 2 R2C + 2 C2R hipFFT transforms, spectrum kernel per step)."""
    from pymonteopt_amd.modeling import TuringPattern
    t = TuringPattern()
    t.set_params(seed=1)
    u, v = t.grow(steps=2)                 # warm-up: plans, code objects
    del u, v
    u, v = t.grow()
    steps, n = int(t.repetition), t.gridsize
    bv, peaks = t.microarchitecture_of(u)
    return {"grid": n, "steps": steps, "seconds": t.time_s, "ms_per_step": 1e3 * t.time_s / steps,
            "cell_steps_per_s": n ** 3 * steps / t.time_s, "bv_tv_real": bv, "peaks": peaks}


def bone_leg(peaks, n_packets=100_000_000, wide=True):
    """The cancellous-bone phantom at 0.02 mm (GRF trabecular core, cortical shell,
    soft-tissue top): one launch of 1e8 packets with the absorbed map, Rd(r) and ten annular
    top-face detectors, after a 1e6-packet warm-up. Fresnel at every bone/marrow face makes
    this the crossing- and Fresnel-heavy case. wide: the 256^3 phantom's core tiled 5 x 5
    (1072 x 1072 x 256 voxels, 21.4 mm wide) with the rings centred at 1..10 mm; else the
    256^3 phantom itself with rings at 0.25..2.5 mm. The bone workload's outputs are the
    detectors, Rd(r) and the scalars; the walk's own per-voxel absorbed map (which the
    reference kernel does not tally) is timed with them, and the same launch without it is
    reported beside (`detectors_only`)."""
    from pymonteopt_amd import presets
    from pymonteopt_amd.runtime import Walker
    wl = presets.bone_wide_workload(seed=0) if wide else presets.bone_workload(seed=0)
    wk = Walker(wl.labels, wl.pitch, wl.optics)
    out = workload_leg(wl, peaks, n_packets, walker=wk)
    out["detectors_only"] = workload_leg(wl, peaks, n_packets, walker=wk, absorbed=False)
    return out


def workload_leg(wl, peaks, n_packets=100_000_000, walker=None, absorbed=True):
    """One launch of n_packets through a workload with all its tallies (and the absorbed
    map unless absorbed=False), at the launch shape Walker.tune picks, after a 1e6-packet
    warm-up; photons/s from HIP events around the launch, with its VALU and atomic roofline
    fractions."""
    import torch
    from pymonteopt_amd import _abi, roofline
    from pymonteopt_amd.runtime import TallySpec, Walker
    wk = walker if walker is not None else Walker(wl.labels, wl.pitch, wl.optics)
    spec = TallySpec(**wl.tallies)
    spec.absorbed = bool(absorbed)
    kw, _ = wk.tune(wl.source, spec, n_packets=5_000_000)
    tal = wk.tallies(spec)
    wk.walk(wl.source, 1_000_000, 99, tal, **kw)
    tal.zero()
    torch.cuda.synchronize()
    e0, e1 = _events()
    e0.record()
    wk.walk(wl.source, n_packets, 5, tal, **kw)
    e1.record()
    e1.synchronize()
    t = e0.elapsed_time(e1) * 1e-3
    s = dict(zip(_abi.SCALAR_NAMES, tal.scalars.cpu().tolist()))
    ev = dict(zip(_abi.EVENT_NAMES, [int(x) for x in tal.events.cpu().tolist()]))
    ops = roofline.ops_per_packet(ev)
    glob_pp = ev["global_deposits"] / n_packets
    out = {"workload": wl.name + f", {n_packets:.0e} packets", "seconds": t, "photons_per_s": n_packets / t,
           "launch": kw or "library default",
           "events_per_packet": {k: ev[k] / n_packets for k in ("interactions", "crossings", "fresnel", "roulette")},
           "Rd": s["Rd"] / n_packets, "Tt": s["Tt"] / n_packets, "A": s["A"] / n_packets,
           "valu_frac_published": n_packets * ops / t / roofline.PEAK_VALU_PUBLISHED_LANE_OPS,
           "atomic_frac_measured": glob_pp * n_packets / t / peaks["fp64_atomics_per_s"],
           "global_deposits_per_packet": glob_pp, "ops_per_packet": ops}
    if wl.tallies.get("rings"):
        out["rings_mm"] = [list(r) for r in wl.tallies["rings"]]
        out["ring_weight_per_packet"] = [float(x) for x in tal.hist["ring_w"].cpu().numpy() / n_packets]
    if absorbed and glob_pp > 0:
        # the atomic rate at this map's footprint: above the 256 MB of Infinity Cache the
        # device's FP64 atomics run slower than the 32 MiB probe of the headline roofline
        del tal
        torch.cuda.empty_cache()
        peak = atomic_peak_at(wk.nvox)
        out["atomic_peak_at_map_footprint"] = peak
        out["atomic_frac_at_map_footprint"] = glob_pp * n_packets / t / peak
    return out


def atomic_peak_at(n_doubles, blocks=4096, iters=256):
    """FP64 atomic adds per second (libpmo's probe, no return value) on pseudo-random
    addresses of a buffer of n_doubles, the best of two timed runs."""
    import ctypes as C
    import torch
    from pymonteopt_amd import _abi
    L = _abi.lib()
    n = int(min(max(n_doubles, 1), (1 << 32) - 1))
    buf = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    best = 0.0
    for _ in range(2):
        e0, e1 = _events()
        e0.record()
        _abi.check(L.pmo_probe_atomics(C.c_void_p(buf.data_ptr()), n, blocks, iters, _abi.PMO_PROBE_F64, s))
        e1.record()
        e1.synchronize()
        best = max(best, blocks * 256 * iters / (e0.elapsed_time(e1) * 1e-3))
    del buf
    torch.cuda.empty_cache()
    return best


def slab_leg(n_batches=10, n_packets=100_000):
    """The CPU-runnable slab (101 x 101 x 50 @0.1 mm, pencil beam, absorbed map and Rd(r) in
    50 bins of 0.1 mm) as its ten 1e5-packet seed batches (seeds 0-9), each a launch of its
    own: the batches' total GPU-event time, and Rd, Tt, A with their standard errors over the
    batches."""
    import numpy as np
    import torch
    from pymonteopt_amd import _abi, presets
    from pymonteopt_amd.runtime import TallySpec, Walker
    wl = presets.slab_plate()
    wk = Walker(wl.labels, wl.pitch, wl.optics)
    tal = wk.tallies(TallySpec(**wl.tallies))
    wk.walk(wl.source, n_packets, 99, tal)                  # code-object load
    torch.cuda.synchronize()
    t, res = 0.0, []
    for seed in range(n_batches):
        tal.zero()
        e0, e1 = _events()
        e0.record()
        wk.walk(wl.source, n_packets, seed, tal)
        e1.record()
        e1.synchronize()
        t += e0.elapsed_time(e1) * 1e-3
        s = dict(zip(_abi.SCALAR_NAMES, tal.scalars.cpu().tolist()))
        res.append([s["Rd"] / n_packets, s["Tt"] / n_packets, s["A"] / n_packets])
    r = np.array(res)
    m, se = r.mean(0), r.std(0, ddof=1) / math.sqrt(n_batches)
    return {"workload": wl.name + f", seeds 0-{n_batches - 1} x {n_packets:.0e} packets", "seconds": t,
            "photons_per_s": n_batches * n_packets / t, "ms_per_batch": 1e3 * t / n_batches,
            "Rd": [float(m[0]), float(se[0])], "Tt": [float(m[1]), float(se[1])], "A": [float(m[2]), float(se[2])]}


# The reference's own published benchmark: a homogeneous 5 mm slab, 5 x 5 mm wide (mu_s 10,
# mu_a 0.01 /mm, g 0.9, n 1.5, ambient 1.0), pencil beam, timed around start()
# (test_voxelcuda.ipynb:1409-1423 parameters; pymopt/voxel_gpu/_classes.py:88-127 timed region).
PUBLISHED_SLAB = dict(thickness=[5], ms=[10], ma=[0.01], g=[0.9], n=[1.5], n_air=1.0, x_size=5, y_size=5)
# 1e8 packets in 21.68, 21.33 and 22.06 s on an unstated NVIDIA GPU (test_voxelcuda.ipynb:1270-1273,
# :1403-1406 and cell 37's outputs); Rd, Td printed by the two 1e8 runs
PUBLISHED_GPU_S_1E8 = (21.680, 21.331, 22.06)
PUBLISHED_RDTD_1E8 = ((0.527456, 0.344113), (0.528290, 0.343399))
# pitch sweep: 5e6 packets per pitch, mean of 3 repeats (test_voxelcuda.ipynb:999-1001, cell 82)
PUBLISHED_PITCH_S = {5: 1.0833, 1: 0.969, 0.5: 1.0657, 0.1: 1.6243, 0.05: 2.182, 0.01: 6.938, 0.005: 12.2017}
# the Numba CPU kernel, serial, one core: 1e8 packets in 949.6 s (test_voxelcuda.ipynb:1585,1608)
PUBLISHED_CPU_PER_CORE = 1e8 / 949.6


def published_slab_leg(n_packets=100_000_000, sweep_packets=5_000_000):
    """The reference's published benchmark workloads through the model API, timed as the
    reference times them: wall time of start() on a freshly built model, which uploads the
    volume, walks, copies the tallies back and normalises them. The GPU-event time of the
    walk alone is reported beside it. The sweep repeats the reference's voxel-pitch series
    (5 mm down to 0.005 mm, a 1002^3 volume) at its packet count, without the absorbed map
    (the reference's sweep computed only Rd(r), 200 bins of 0.01 mm)."""
    import torch
    from pymonteopt_amd.voxel import VoxelPlateModel

    def run(nph, pitch, **kw):
        m = VoxelPlateModel(nPh=nph, **kw).build(**PUBLISHED_SLAB, voxel_space=pitch)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m.start(0)
        wall = time.perf_counter() - t0
        r = m.getRdTtValues()
        return m, wall, float(r["Rd"]), float(r["Td"])

    run(100_000, 0.5)                                       # code-object load, first allocations
    m, wall, rd, td = run(n_packets, 0.5, nr=100, dr=0.01)
    ref_s = statistics.mean(PUBLISHED_GPU_S_1E8)
    out = {"workload": "5 mm slab, 5 x 5 mm wide @0.5 mm (12^3 padded), pencil, "
                       f"{n_packets:.0e} packets, absorbed map + Rd(r) 100 x 0.01 mm",
           "start_wall_s": wall, "photons_per_s": n_packets / wall,
           "walk_s": m.time_s, "walk_photons_per_s": n_packets / m.time_s,
           "Rd": rd, "Td": td, "reference_Rd_Td_1e8": [list(x) for x in PUBLISHED_RDTD_1E8],
           "reference_gpu_photons_per_s": n_packets / ref_s,
           "speedup_vs_reference_gpu": ref_s / wall,
           "reference_hardware": "unstated NVIDIA GPU"}
    sweep = []
    for pitch, ref in PUBLISHED_PITCH_S.items():
        m, wall, rd, td = run(sweep_packets, pitch, fluence=False, nr=200, dr=0.01)
        shape = m.model.voxel_model.shape
        sweep.append({"pitch_mm": pitch, "volume": list(shape), "start_wall_s": wall,
                      "walk_s": m.time_s, "photons_per_s": sweep_packets / wall,
                      "reference_photons_per_s": sweep_packets / ref, "speedup": ref / wall,
                      "Rd": rd, "Td": td})
        del m
        torch.cuda.empty_cache()
    out["pitch_sweep"] = {"packets": sweep_packets, "points": sweep}
    return out


def cpu_anchor_published_slab(budget_s=8.0):
    """The oracle on one host thread, float32 state, on the published slab at 0.5 mm, beside
    the reference's published serial Numba rate (1.053e5 photons/s, unstated CPU): a sanity
    anchor for the CPU baseline, not a comparison of machines."""
    from oracle import oracle as og
    from pymonteopt_amd import presets
    og.build()
    wl = presets.plate("published slab", **PUBLISHED_SLAB, voxel_space=0.5)
    ma, ms, g, n = wl.optics_columns()
    n_s = 20_000
    while True:
        t0 = time.perf_counter()
        og.walk(wl.labels, wl.pitch, ma, ms, g, n, wl.source.as_dict(), n_s, 31, precision="f32",
                threads=1, nr=100, dr=0.01)
        dt = time.perf_counter() - t0
        if dt >= 0.6 * budget_s or n_s >= 20_000_000:
            break
        n_s = int(min(20_000_000, n_s * min(64.0, budget_s / dt)))
    return {"value": n_s / dt, "unit": "photons/s", "cores": 1, "kind": "port",
            "sample": f"{n_s} packets of the 5 mm slab @0.5 mm (f32 oracle, absorbed + Rd(r)), {dt:.1f} s",
            "reference_numba_per_core": PUBLISHED_CPU_PER_CORE}


def virtual_obd_leg(n_packets=10_000_000, seeds=(4, 5)):
    """Subjects of the virtual OBD pipeline at the script's sizes (pymopt/virtualOBD.py): a
    300^3 Turing volume grown 100 steps and tiled 2 x 4 x 2, wrapped in the subject's
    cortical, subcutis and dermis layers, 1e7 packets walked with the three detector
    tallies, the run and the three detector CSVs written. Times per stage, in seconds, for
    each subject; the first one includes hipFFT's one-time kernel builds for the 300^3
    transforms, so the second is the steady state of a long run."""
    import shutil
    import tempfile

    import numpy as np
    import torch
    from pymonteopt_amd import virtual_obd as vo
    from pymonteopt_amd.param_generator import generate_variable_params
    from pymonteopt_amd.voxel import VoxelTuringModel
    root = tempfile.mkdtemp(prefix="pmo_vobd_")
    subjects = []
    try:
        for seed in seeds:
            mp, mo = dict(vo.model_params), dict(vo.monte_params)
            np.random.seed(seed)
            t0 = time.perf_counter()
            g = generate_variable_params()
            g.set_params(vo.determining_params_range())
            vp = g.generate(1)
            u, bv = vo.generate_bone_model(vp["bv_tv"][0], os.path.join(root, "model"), mp)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            vp["bv_tv"][0] = bv
            vo.subject_params(vp, mo)
            m = VoxelTuringModel(nPh=n_packets, model_name="TuringModel", nr=vo.RD_BINS, dr=vo.RD_WIDTH,
                                 nl=vo.LINE_BINS, dl=vo.LINE_WIDTH, line_y=vo.LINE_Y, fluence=False)
            m.set_model(u)
            m.build(**{k: v for k, v in mo.items() if k not in ("symmetrization", "enclosure")})
            shape = m.model.voxel_model.shape
            t2 = time.perf_counter()
            m.walker()
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            m.start(seed)
            t4 = time.perf_counter()
            m.save_result(os.path.join(root, "monte"), coment="bench")
            vo.calc_ray_tracing(m.get_result(), mo, os.path.join(root, "opt"), "bench")
            t5 = time.perf_counter()
            res = m.get_result()
            subjects.append({"seed": seed, "voxels": list(shape), "packets": n_packets, "bv_tv_real": bv,
                             "bone_model_s": t1 - t0, "wrap_s": t2 - t1, "upload_s": t3 - t2, "walk_s": m.time_s,
                             "start_s": t4 - t3, "save_and_csv_s": t5 - t4, "subject_s": t5 - t0,
                             "walk_photons_per_s": n_packets / m.time_s, "Rd": res["Rd"], "Tt": res["Tt"]})
            del m, u
        # the script's loop as virtual_obd.run runs it: subject i walks on its own stream while
        # subject i + 1 is drawn, grown and wrapped
        np.random.seed(seeds[-1])
        t0 = time.perf_counter()
        done = vo.run(3, overlap=True, root=os.path.join(root, "loop"), nPh=n_packets, model_params=dict(vo.model_params),
                      monte_params=dict(vo.monte_params))
        loop_s = (time.perf_counter() - t0) / max(1, len(done))
    finally:
        shutil.rmtree(root, ignore_errors=True)
    return dict(subjects[-1], first_subject_s=subjects[0]["subject_s"], run_overlapped_s_per_subject=loop_s,
                subjects=subjects)


def cpu_baseline_turing(n=128, steps=4):
    """The oracle's numpy restatement (float64 FFTs, one process) on an n^3 grid."""
    from oracle import turing as tp
    u, v = (x.astype("float64") for x in tp.initial_fields(n, 1))
    ku, kv = tp.inverse_symbol(n, 0.0002, 1.0, 1 / 40), tp.inverse_symbol(n, 0.01, 1.0, 1 / 40)
    t0 = time.perf_counter()
    for _ in range(steps):
        u, v = tp.step(u, v, ku, kv, 1.0)
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * steps / dt, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"{n}^3 grid x {steps} steps, numpy float64 FFTs, {dt:.1f} s"}


def measure_peaks():
    """Re-measures the two peaks with libpmo's probes: FP64 global atomic adds per second
    (pseudo-random addresses in 32 MiB) and FP32 FMA lane-ops per second; and the rate of
    Philox4x32-10 blocks (the walk's draws) with every lane active."""
    import ctypes as C
    import torch
    from pymonteopt_amd import _abi
    L = _abi.lib()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    buf = torch.zeros(1 << 22, dtype=torch.float64, device="cuda:0")
    blocks, iters = 4096, 256
    out = {}
    for _ in range(2):
        e0, e1 = _events()
        e0.record()
        _abi.check(L.pmo_probe_atomics(C.c_void_p(buf.data_ptr()), buf.numel(), blocks, iters,
                                        _abi.PMO_PROBE_F64, s))
        e1.record()
        e1.synchronize()
        out["fp64_atomics_per_s"] = blocks * 256 * iters / (e0.elapsed_time(e1) * 1e-3)
    vb, vi = 2048, 8192
    o = torch.empty(vb * 256, dtype=torch.float32, device="cuda:0")
    for _ in range(2):
        e0, e1 = _events()
        e0.record()
        _abi.check(L.pmo_probe_valu(C.c_void_p(o.data_ptr()), vb, vi, s))
        e1.record()
        e1.synchronize()
        out["valu_lane_ops_per_s"] = vb * 256.0 * vi * 32 / (e0.elapsed_time(e1) * 1e-3)
    w = torch.empty(vb * 256, dtype=torch.int32, device="cuda:0")
    for _ in range(2):
        e0, e1 = _events()
        e0.record()
        _abi.check(L.pmo_probe_philox(C.c_void_p(w.data_ptr()), vb, 1024, s))
        e1.record()
        e1.synchronize()
        out["philox_blocks_per_s"] = vb * 256.0 * 1024 / (e0.elapsed_time(e1) * 1e-3)
    return out


def traffic_from_profile():
    """HBM bytes per walk launch measured by rocprofv3 (profiles/walk_skin_traffic.json)."""
    path = os.path.join(ROOT, "profiles", "walk_skin_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path))
    except ValueError:
        return None


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
    ap.add_argument("--no-extra", action="store_true", help="skip the batched and adjoint legs")
    ap.add_argument("--quorum", type=int, default=0, help="lanes ready before a wave runs interactions")
    ap.add_argument("--no-tune", action="store_true", help="take the library's default launch shape")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported. The walk would shard "
              "by packet-id range, one share per GPU, with one all-reduce of the FP64 tallies at the end.",
              file=sys.stderr)
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
              window=tuple(int(x) for x in args.window.split(",")), quorum=args.quorum)
    launch = "library default"
    if not args.no_tune and args.block == 0 and args.window == "0,0,0" and args.quorum == 0 and not args.fp64:
        # the launch shape (block size, LDS window, quorum) picked by timing a 1e7-packet walk
        # of each candidate, before the timed region
        tuned, _ = wk.tune(wl.source, spec, n_packets=min(N, 10_000_000))
        kw.update(tuned)
        launch = tuned or "library default"

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
            "launch": launch,
            "tallies": "Rd, Tt, Rd(r) 40x0.1mm, absorbed/fluence map (FP64)" if spec.absorbed
            else "Rd, Tt, Rd(r) 40x0.1mm (FP64)",
            "walk_kernel_ms_mean": 1e3 * mean_launch_s,
            "walk_kernel_ms_median": 1e3 * statistics.median(kernel_s),
            "photons_per_s_per_step_median": N / statistics.median(kernel_s),
            "photons_per_s_per_step_ci95": _ci95([N / t for t in kernel_s]),
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
            "peak": roofline.PEAK_VALU_PUBLISHED_LANE_OPS / 1e12,
            "unit": "Tlane-op/s",
            "frac": achieved / roofline.PEAK_VALU_PUBLISHED_LANE_OPS,
            "traffic": None,
            "peak_source": "MI355X published FP32 peak 157.3 TFLOP/s = 78.65e12 FMA lane-op slots/s",
            "ops_per_packet": ops,
            "hbm_algorithmic_GBps": N * bytes_pp / mean_launch_s / 1e9,
            "hbm_algorithmic_frac": N * bytes_pp / mean_launch_s / roofline.PEAK_HBM_BPS,
        },
    }
    peaks = measure_peaks()
    glob_pp = ev_sum.get("global_deposits", 0) / (N * args.steps)
    # the FP32 FMA issue rate re-measured by libpmo's probe (non-packed FP32, as the walk
    # issues it) and the one-op-per-lane-per-clock figure are kept beside the published peak
    out["roofline"]["peak_measured"] = peaks["valu_lane_ops_per_s"] / 1e12
    out["roofline"]["frac_of_measured_peak"] = achieved / peaks["valu_lane_ops_per_s"]
    out["roofline"]["peak_64_lanes_per_clock"] = roofline.PEAK_VALU_LANE_OPS / 1e12
    out["roofline_atomic"] = {
        "bound": "fp64_atomic",
        "achieved": glob_pp * N / mean_launch_s,
        "peak": peaks["fp64_atomics_per_s"],
        "unit": "atomics/s",
        "frac": glob_pp * N / mean_launch_s / peaks["fp64_atomics_per_s"],
        "global_deposits_per_packet": glob_pp,
        "peak_source": "libpmo's device-scope FP64 atomic-add probe (no published figure exists)",
    }
    # the resource that binds this walk, measured: the FP64 atomics of the absorbed map
    out["roofline"]["binding"] = "roofline_atomic"
    prof = traffic_from_profile()
    if prof is not None and int(prof.get("packets_per_launch", 0)) > 0:
        scale = N / float(prof["packets_per_launch"])
        out["roofline"]["traffic"] = prof["hbm_bytes_per_launch"] * scale
        out["roofline"]["traffic_source"] = prof["source"]
    # the legs' own progress prints (the model API prints as the reference's does) go to
    # stderr, so stdout carries the one JSON line only
    with contextlib.redirect_stdout(sys.stderr):
        if not args.no_extra:
            out["batched"] = batched_leg()
            # the sweep draws one Philox block per interaction: its rate of blocks against the
            # probe's rate with every lane active is the share of the device the draws take
            b = out["batched"]
            blocks_s = b["packets"] * b["events_per_packet"]["interactions"] / b["batched_s"]
            b["draws"] = {"philox_blocks_per_s": blocks_s, "probe_blocks_per_s": peaks["philox_blocks_per_s"],
                          "share_of_probe_rate": blocks_s / peaks["philox_blocks_per_s"]}
            out["adjoint"] = adjoint_leg()
            out["obd"], obd_sample = obd_leg()
            out["turing"] = turing_leg()
            out["virtual_obd"] = virtual_obd_leg()
            out["skin_flat"] = workload_leg(presets.skin_plate("flat"), peaks)
            out["slab"] = slab_leg()
            out["bone"] = bone_leg(peaks, wide=True)
            out["bone_256"] = bone_leg(peaks, wide=False)
            out["published_slab"] = published_slab_leg()
        if not args.no_cpu:
            out["cpu_baseline"] = cpu_baseline(wl)
            if not args.no_extra:
                out["slab"]["gpu_vs_cpu"] = out["slab"]["photons_per_s"] / out["cpu_baseline"]["slab_101x101x50"]["value"]
                out["obd"]["cpu_baseline"] = cpu_baseline_obd(*obd_sample)
                out["obd"]["gpu_vs_cpu"] = out["obd"]["traces_per_s"] / out["obd"]["cpu_baseline"]["value"]
                out["turing"]["cpu_baseline"] = cpu_baseline_turing()
                out["published_slab"]["cpu_anchor"] = cpu_anchor_published_slab()
                out["batched"]["cpu_baseline"] = cpu_baseline_batched()
                out["batched"]["gpu_vs_cpu"] = out["batched"]["batched_photons_per_s"] / out["batched"]["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


def _ci95(x):
    """Student-t 95% interval of the mean of x."""
    n = len(x)
    if n < 2:
        return None
    from scipy import stats
    m = sum(x) / n
    h = stats.t.ppf(0.975, n - 1) * _se(x)
    return [m - h, m + h]


def _se(x):
    n = len(x)
    if n < 2:
        return float("nan")
    m = sum(x) / n
    return math.sqrt(sum((v - m) ** 2 for v in x) / (n - 1) / n)


if __name__ == "__main__":
    main()
