"""Helpers shared by the GPU parity tests: run a workload through the HIP walk and
through the oracle with the same seeds, and compare."""
import math

import numpy as np


def oracle_run(og, wl, n, seed, precision="f64", absorbed=True, **extra):
    ma, ms, g, nn = wl.optics_columns()
    t = dict(wl.tallies)
    t.update(extra)
    rings = t.pop("rings", None)
    return og.walk(wl.labels, wl.pitch, ma, ms, g, nn, wl.source.as_dict(), n, seed, absorbed=absorbed,
                   precision=precision, rings=rings, threads=0, **t)


def gpu_run(walker, wl, n, seed, spec=None, **kw):
    import torch
    from pymonteopt_amd.runtime import TallySpec
    spec = spec or TallySpec(**wl.tallies)
    tal = walker.tallies(spec, n if spec.exit_state else 0)
    walker.walk(wl.source, n, seed, tal, **kw)
    torch.cuda.synchronize()
    return tal.to_host(shape=walker.shape)


def zscore(a, b):
    """Pooled z of two sets of per-batch values."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    se = math.sqrt(a.var(ddof=1) / a.size + b.var(ddof=1) / b.size)
    d = a.mean() - b.mean()
    if se == 0:
        return 0.0 if d == 0 else math.inf
    return abs(d) / se


def binned_z(A, B):
    """Per-bin pooled z of two [batches, bins] arrays (bins empty on both sides give 0)."""
    A = np.asarray(A, float)
    B = np.asarray(B, float)
    se = np.sqrt(A.var(axis=0, ddof=1) / A.shape[0] + B.var(axis=0, ddof=1) / B.shape[0])
    d = np.abs(A.mean(0) - B.mean(0))
    with np.errstate(divide="ignore", invalid="ignore"):
        z = np.where(se > 0, d / se, np.where(d == 0, 0.0, np.inf))
    return z


def ledger_error(s):
    bal = s["Rd"] + s["Tt"] + s["other"] + s["A"] + s["kill"] - s["gain"]
    return abs(s["w0"] - bal) / s["w0"]


def packet_events(r):
    """The per-packet event counts of a run: every counter except the atomic-request count,
    which depends on how packets shared lanes (and their deposit buffers) in turn."""
    return {k: v for k, v in r["events"].items() if k != "global_deposits"}


def binned_equal_up_to_edges(x, w, edges, got, eps):
    """The device's per-bin sums `got` against host binning of the dumped values x (bins
    (edges[i], edges[i+1]]): equal except for the weight of packets within eps of an edge,
    which FP32 rounding on the device may put on either side. Returns the largest excess of
    |difference| over that slack, relative to the largest bin (<= 0 when they agree)."""
    x = np.asarray(x, np.float64)
    w = np.asarray(w, np.float64)
    edges = np.asarray(edges, np.float64)
    idx = np.searchsorted(edges, x, side="left") - 1
    ok = (idx >= 0) & (idx < edges.size - 1) & (x > edges[0])
    want = np.bincount(idx[ok], weights=w[ok], minlength=edges.size - 1)
    j = np.clip(np.searchsorted(edges, x), 0, edges.size - 1)
    near = (np.abs(x - edges[j]) < eps) | (np.abs(x - edges[np.maximum(j - 1, 0)]) < eps)
    slack = np.zeros(edges.size - 1)
    # a near-edge packet may move into either bin beside its edge
    ne = np.argmin(np.abs(x[near][:, None] - edges[None, :]), axis=1)
    for k, wk in zip(ne, w[near]):
        if k - 1 >= 0:
            slack[k - 1] += wk
        if k < edges.size - 1:
            slack[k] += wk
    scale = max(float(np.max(np.abs(want))), 1e-300)
    return float(np.max(np.abs(np.asarray(got, np.float64) - want) - slack - 1e-9 * scale) / scale)
