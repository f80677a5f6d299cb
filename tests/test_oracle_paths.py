"""Pins the oracle's per-voxel absorbed map and its per-label path sums (the perturbation
Monte Carlo statistics the adjoint is checked against) on the CPU.

* absorbed map, non-scattering slab: with mu_s = 0 the first interaction drops the whole
  weight (w -= w mu_a/mu_t, pymopt/voxel/_kernel.py:130), so the centre column holds
  w0 e^{-mu_a (k-1) l} (1 - e^{-mu_a l}) in its k-th tissue voxel and nothing else does;
* path sums, non-scattering slab with the normal-incidence Fresnel limit: a packet leaving
  through the top after k round trips has W = w0, L = 2kd, probability
  R^{2k-1} (1-R) e^{-2k mu_a d}; its interaction count is zero;
* path sums, scattering 3-layer slab: sum_j of w L_j and w n_j recounted packet by packet
  from the oracle's own event trace (segment lengths between consecutive events, in the
  label of the voxel the segment lies in; interactions per label), and the walk's event
  counters recounted from the same trace.
"""
import math

import numpy as np
import pytest

from pymonteopt_amd.geometry import center_add, optics_table, plate_volume, specular


def _cols(opt):
    L = opt.shape[0] - 1
    return opt[:L, 0], opt[:L, 1], opt[:L, 2], opt[:, 3]


def test_absorbed_map_non_scattering_closed_form(oracle):
    ma, d, l, n = 2.0, 0.5, 0.05, 1.4
    vol = plate_volume([d], l, 1.05, 1.05)
    opt = optics_table([ma], [0.0], [0.0], [n], 1.0)
    cx, cy = center_add(vol.shape[0]), center_add(vol.shape[1])
    w0 = 1 - specular(1.0, n)
    src = dict(kind=0, voxel=(cx, cy, 1), w0=w0)
    N = 40000
    out = oracle.walk(vol, l, *_cols(opt), src, N, 5, absorbed=True)
    A = out["absorbed"]
    col = A[cx, cy, 1:-1] / (N * w0)
    nz = vol.shape[2] - 2
    k = np.arange(nz)
    p = np.exp(-ma * k * l) * (1 - np.exp(-ma * l))
    se = np.sqrt(p * (1 - p) / N)
    assert np.all(np.abs(col - p) < 4 * se + 1e-12)
    off = A.copy()
    off[cx, cy, :] = 0
    assert off.sum() == 0.0
    assert A.sum() == pytest.approx(out["scalars"]["A"], rel=1e-12)


def test_path_sums_non_scattering_closed_form(oracle):
    ma, d, l, n = 1.0, 0.5, 0.05, 1.8
    vol = plate_volume([d], l, 1.05, 1.05)
    opt = optics_table([ma], [0.0], [0.0], [n], 1.0)
    cx, cy = center_add(vol.shape[0]), center_add(vol.shape[1])
    w0 = 1 - specular(1.0, n)
    # a 0.02 mm flat disc keeps every exit inside the first 0.1 mm radial bin and off r = 0,
    # which the (i dr, (i+1) dr] bins leave out
    src = dict(kind=1, voxel=(cx, cy, 1), radius=0.02, w0=w0)
    N = 60000
    out = oracle.walk(vol, l, *_cols(opt), src, N, 9, absorbed=False, nr=4, dr=0.1, path_stats=True,
                      flags=2)
    R = ((n - 1) / (n + 1)) ** 2
    q = math.exp(-2 * ma * d)
    ks = np.arange(1, 60)
    prob = R ** (2 * ks - 1) * (1 - R) * q ** ks
    EW = w0 * prob.sum()
    EWL = w0 * (prob * 2 * ks * d).sum()
    EWL2 = w0 ** 2 * (prob * (2 * ks * d) ** 2).sum()
    se = math.sqrt((EWL2 - EWL ** 2) / N)
    assert out["rd_r"][0] / N == pytest.approx(EW, abs=4 * math.sqrt(EW * w0 / N))
    assert out["rd_wl"][0, 0] / N == pytest.approx(EWL, abs=4 * se)
    assert out["rd_wl"][1:].sum() == 0.0
    assert out["rd_wn"].sum() == 0.0
    # the pMC identity dE[W]/dmu_a = -E[W L] holds for these closed forms
    h = 1e-6
    EW_h = lambda m: w0 * (R ** (2 * ks - 1) * (1 - R) * math.exp(-2 * m * d) ** ks).sum()
    assert -(EW_h(ma + h) - EW_h(ma - h)) / (2 * h) == pytest.approx(EWL, rel=1e-6)


def test_path_sums_match_the_event_trace(oracle):
    l = 0.05
    vol = plate_volume([0.1, 0.3, 0.6], l, 2.05, 2.05)
    opt = optics_table([0.5, 0.2, 1.5], [20.0, 10.0, 5.0], [0.8, 0.9, 0.7], [1.44, 1.38, 1.5], 1.0)
    cols = _cols(opt)
    cx, cy = center_add(vol.shape[0]), center_add(vol.shape[1])
    src = dict(kind=0, voxel=(cx, cy, 1), w0=1 - specular(1.0, 1.44))
    nr, dr, L = 20, 0.1, 3
    seed = 21
    checked = 0
    for pid in range(300):
        ev = oracle.trace(vol, l, *cols, src, pid, seed, cap=200000)
        assert ev[-1]["type"] in (4, 5)
        Lj, nj = np.zeros(L), np.zeros(L)
        for a, b in zip(ev[:-1], ev[1:]):
            lab = int(vol[a["a"]])
            ga = np.array(a["a"], float) * l + np.array(a["p"])
            # a transmission's event sits in the new voxel with p flipped onto the shared
            # face, so a*l + p is continuous and the segment lies in the earlier event's voxel
            gb = np.array(b["a"], float) * l + np.array(b["p"])
            Lj[lab] += float(np.linalg.norm(gb - ga))
            if b["type"] == 3:
                nj[int(vol[b["a"]])] += 1
        out = oracle.walk(vol, l, *cols, src, 1, seed, first_packet=pid, absorbed=False, nr=nr, dr=dr,
                          path_stats=True, exit_state=True)
        x, y, z, vx, vy, vz, _ = (float(c) for c in out["exit_state"][:, 0])
        w = ev[-1]["w"]            # the double weight (the dump holds float32)
        r = math.hypot(x, y)
        binned = ev[-1]["type"] == 4 and vz < 0 and 0 < r <= nr * dr
        wl, wn = out["rd_wl"].sum(axis=0), out["rd_wn"].sum(axis=0)
        # event counters: a kill ends an interaction before its (post-spin) trace event
        types = [e["type"] for e in ev]
        nidx = lambda a: float(opt[int(vol[a]), 3])
        fres = sum(1 for a, b in zip(ev[:-1], ev[1:])
                   if b["type"] == 2 or (b["type"] in (1, 4) and nidx(a["a"]) != nidx(b["a"])))
        gains = sum(1 for a, b in zip(ev[:-1], ev[1:]) if b["type"] == 3 and b["w"] > a["w"])
        assert out["events"] == dict(packets=1, interactions=types.count(3) + types.count(5),
                                     crossings=types.count(1) + types.count(2) + types.count(4),
                                     fresnel=fres, roulette=gains + types.count(5), killed=types.count(5),
                                     escaped=types.count(4))
        if binned:
            checked += 1
            assert wl == pytest.approx(w * Lj, rel=1e-9, abs=1e-12)
            assert wn == pytest.approx(w * nj, rel=1e-12, abs=1e-15)
        else:
            assert wl.sum() == 0 and wn.sum() == 0
    assert checked > 50


def test_rz_fluence_non_scattering_closed_form(oracle):
    # each packet drops all of w0 at one depth z ~ Exp(mu_a) on the axis; the r-z tally holds
    # w0/mu_a per packet in the bin of that depth. dz = 0.03 does not align with the 0.05 mm
    # voxels, so the tally must bin interaction positions, not voxels
    ma, d, l, n = 2.0, 0.5, 0.05, 1.4
    vol = plate_volume([d], l, 1.05, 1.05)
    opt = optics_table([ma], [0.0], [0.0], [n], 1.0)
    cx, cy = center_add(vol.shape[0]), center_add(vol.shape[1])
    w0 = 1 - specular(1.0, n)
    N, dz, nz = 40000, 0.03, 16
    out = oracle.walk(vol, l, *_cols(opt), dict(kind=0, voxel=(cx, cy, 1), w0=w0), N, 6, absorbed=False,
                      rz=(3, 0.1, nz, dz))
    F = out["fluence_rz"]
    assert F.shape == (3, nz) and F[1:].sum() == 0.0
    k = np.arange(nz)
    p = np.exp(-ma * k * dz) - np.exp(-ma * (k + 1) * dz)
    got = F[0] * ma / (w0 * N)
    assert np.all(np.abs(got - p) < 4 * np.sqrt(p * (1 - p) / N))


def test_rz_fluence_radial_closed_form_flat_beam(oracle):
    # a flat disc of radius R puts the (single) interaction of each packet at r with
    # P(r in [i dr, (i+1) dr)) = ((i+1)^2 - i^2) dr^2 / R^2
    ma, d, l, n = 1.0, 0.3, 0.05, 1.0
    vol = plate_volume([d], l, 2.05, 2.05)
    opt = optics_table([ma], [0.0], [0.0], [n], 1.0)
    cx, cy = center_add(vol.shape[0]), center_add(vol.shape[1])
    R, dr, nr, N = 0.6, 0.07, 10, 40000
    out = oracle.walk(vol, l, *_cols(opt), dict(kind=1, voxel=(cx, cy, 1), radius=R, w0=1.0), N, 8,
                      absorbed=False, rz=(nr, dr, 1, 1.0))
    F = out["fluence_rz"][:, 0] * ma                  # w0 = 1: one unit per interacting packet
    n_int = F.sum()
    frac = F / n_int
    i = np.arange(nr)
    p = np.clip(((i + 1) ** 2 - i ** 2) * dr ** 2 / R ** 2, 0, None)
    p[i * dr >= R] = 0.0
    last = int(R // dr)
    p[last] = 1 - p[:last].sum()                      # the bin holding the disc edge
    assert n_int == pytest.approx(N * (1 - np.exp(-ma * d)), rel=0.05)
    assert np.all(np.abs(frac - p) < 4 * np.sqrt(p * (1 - p) / n_int) + 1e-12)


def test_rz_fluence_times_mua_sums_to_the_absorbed_weight(oracle):
    # one label: every deposit lands in some bin when the bins cover the slab, so
    # mu_a * sum(fluence) equals the absorbed scalar A
    ma, d, l = 0.3, 1.0, 0.05
    vol = plate_volume([d], l, 2.05, 2.05)
    opt = optics_table([ma], [8.0], [0.8], [1.4], 1.0)
    cx, cy = center_add(vol.shape[0]), center_add(vol.shape[1])
    out = oracle.walk(vol, l, *_cols(opt), dict(kind=0, voxel=(cx, cy, 1), w0=0.97), 3000, 12, absorbed=False,
                      rz=(30, 0.05, 26, 0.04))
    # the table holds mu_a in float32, and the tally divides by that value
    assert out["fluence_rz"].sum() * float(opt[0, 0]) == pytest.approx(out["scalars"]["A"], rel=1e-12)
