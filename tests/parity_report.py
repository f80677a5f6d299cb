"""Parity report of the HIP walk against the oracle (the table of the design notes' parity
section). Not a pytest module: run it on a machine with an MI355X from the repo root,

    python -u tests/parity_report.py

For the slab, the 3-layer skin plate (pencil and 1 mm flat beam) and the 256^3 bone phantom
it walks 10 seed batches with the FP32 kernel and its FP64 build, and the oracle on
independent seeds with as many packets, then prints per output: the means, the pooled z-score of FP32 against
the oracle, the largest per-seed |FP32 - FP64| against the seed standard error, binned
z-scores of Rd(r), the absorbed map's FP32-FP64 L1 difference and the energy ledger.
"""
import math
import os
import sys
import time

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

from _parity import binned_z, gpu_run, ledger_error, oracle_run, zscore  # noqa: E402
from oracle import oracle as og  # noqa: E402
from pymonteopt_amd import presets  # noqa: E402
from pymonteopt_amd.runtime import TallySpec, Walker  # noqa: E402

def rep(name, wl, N, seeds, oracle_N=None, absorbed=True):
    wk = Walker(wl.labels, wl.pitch, wl.optics)
    oN = oracle_N or N
    G, D, O = [], [], []
    for s in seeds:
        t0 = time.time(); G.append(gpu_run(wk, wl, N, s)); t1 = time.time()
        D.append(gpu_run(wk, wl, N, s, fp64=True)); t2 = time.time()
        O.append(oracle_run(og, wl, oN, 1000 + s, absorbed=absorbed)); t3 = time.time()
        print(f'   {name} seed {s}: f32 {t1-t0:.2f}s f64 {t2-t1:.2f}s oracle {t3-t2:.2f}s', flush=True)
    print(f'== {name}: GPU f32 / GPU f64 at {N} x {len(seeds)} seeds; oracle f64 at {oN} x {len(seeds)} independent seeds', flush=True)
    for k in ('Rd', 'Tt', 'A'):
        a = np.array([g['scalars'][k] / N for g in G]); d = np.array([x['scalars'][k] / N for x in D])
        o = np.array([x['scalars'][k] / oN for x in O])
        se = np.std(a, ddof=1) / math.sqrt(len(a))
        print(f'  {k}: f32 {a.mean():.6f}  f64 {d.mean():.6f}  oracle {o.mean():.6f}  z(f32,oracle)={zscore(a,o):.2f}  '
              f'max|f32-f64| per seed {np.max(np.abs(a-d)):.2e} (seed SE {se:.1e})', flush=True)
    if 'rd_r' in G[0]:
        A = np.array([g['rd_r'] / N for g in G]); Dd = np.array([x['rd_r'] / N for x in D]); Oo = np.array([x['rd_r'] / oN for x in O])
        z = binned_z(A, Oo)
        m = Dd.mean(0) > 0
        rel = np.abs(A.mean(0) - Dd.mean(0))[m] / Dd.mean(0)[m]
        print(f'  Rd(r): bins with z>3: {int(np.sum(z>3))} of {z.size}, max z {np.max(z):.2f}; max rel |f32-f64| over bins {rel.max():.2e}', flush=True)
    if absorbed and 'absorbed' in G[0]:
        ga = np.mean([g['absorbed'] for g in G], 0); da = np.mean([x['absorbed'] for x in D], 0)
        print(f'  absorbed map: L1 |f32-f64| / L1 = {np.abs(ga-da).sum()/da.sum():.2e}', flush=True)
    print(f'  energy ledger max rel err: f32 {max(ledger_error(g["scalars"]) for g in G):.1e} f64 {max(ledger_error(x["scalars"]) for x in D):.1e}', flush=True)


def main():
    og.build()
    # 1.2e6 packets per batch runs the wide launch (1024-lane blocks, large LDS box)
    rep('slab 101x101x50 @0.1mm', presets.slab_plate(), 1_200_000, range(10))
    rep('3-layer skin pencil', presets.skin_plate(), 1_200_000, range(10))
    rep('3-layer skin flat 1mm', presets.skin_plate('flat'), 200000, range(10), absorbed=False)
    b = presets.bone_workload(seed=0)
    wk = Walker(b.labels, b.pitch, b.optics)
    spec = TallySpec(**b.tallies)
    tal = wk.tallies(spec)
    t0 = time.time(); wk.walk(b.source, 1_000_000, 0, tal); torch.cuda.synchronize(); t1 = time.time() - t0
    print(f'bone 1e6 packets: {t1:.2f} s', flush=True)
    Nb = int(min(1e8, max(1e6, 1e6 * 20.0 / max(t1, 1e-3))))
    for N in (Nb,):
        tal.zero()
        e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
        e0.record(); wk.walk(b.source, N, 1, tal); e1.record(); e1.synchronize()
        t = e0.elapsed_time(e1) / 1e3
        r = tal.to_host()
        print(f'== bone phantom 256^3 @0.02mm: {N:.0e} packets in {t:.2f} s = {N/t:.3e} ph/s; Rd {r["scalars"]["Rd"]/N:.5f} Tt {r["scalars"]["Tt"]/N:.5f} A {r["scalars"]["A"]/N:.5f}; events/packet', {k: v/N for k, v in r['events'].items()}, flush=True)
        area = np.pi * (np.array([x[1] for x in spec.rings])**2 - np.array([x[0] for x in spec.rings])**2)
        print('   ring detectors (SDS mm, R per mm^2):', [(round((a+b_)/2, 3), f'{w/(ar*N):.3e}') for (a, b_), w, ar in zip(spec.rings, r['ring_w'], area)], flush=True)
    rep('bone phantom 256^3 (parity)', b, 200000, range(10), absorbed=False)


if __name__ == "__main__":
    main()
