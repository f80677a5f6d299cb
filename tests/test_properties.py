"""Property tests (hypothesis) of the oracle's per-event physics and of the host logic that
feeds the device: invariants that hold for every input, checked on generated ones.

* Fresnel (pymopt/voxel/_kernel.py:82-108): reflectance in [0, 1], 1 exactly under total
  internal reflection, and the same from either side of the face for a ray and its
  refracted partner (Stokes reciprocity), which a swapped index or angle breaks.
* Henyey-Greenstein cos(theta) (_kernel.py:12-21): -1 at u = 0, +1 at u = 1, and
  non-decreasing in u for every g.
* Spin (_kernel.py:143-165): unit result, cos of the turn equals cos(theta).
* radialDistance (pymopt/metrics/_scoring.py:16-25): every weight with 0 < r <= nn*dr lands
  in exactly one annulus, so the annuli times their areas sum to that weight.
* The batched schedule (pmo_plate_schedule): each instance's ids [0, budget) exactly once,
  in batches of at most `batch`.
* The layered plate (pymopt/voxel/_classes.py:372-415): round(th_i / l) slices per layer,
  in order, inside an end-label shell.
"""
import ctypes as C
import math

import numpy as np
import pytest
from hypothesis import assume, given, settings
from hypothesis import strategies as st

from pymonteopt_amd import _abi, metrics
from pymonteopt_amd.geometry import plate_volume

SET = settings(max_examples=200, deadline=None)
idx = st.floats(1.0, 2.0, allow_nan=False)


@SET
@given(ni=idx, nt=idx, ci=st.floats(0.0, 0.999, allow_nan=False))
def test_fresnel_range_tir_and_reciprocity(oracle, ni, nt, ci):
    # the walk evaluates Fresnel only where the index changes; the literal form loses its
    # digits to cancellation as ni -> nt, so the indices differ by a tissue's step or more
    assume(abs(ni - nt) >= 0.01)
    r, ct = oracle.fresnel(ni, nt, ci)
    # the literal sin/tan form rounds to a hair above 1 at grazing incidence (ci = 0); a
    # reflectance above 1 reflects for every draw, as 1 does
    assert 0.0 <= r <= 1.0 + 1e-12
    si = math.sqrt(1.0 - ci * ci)
    tir = ni > nt and ni / nt * si >= 1.0
    if tir:
        assert r == 1.0
        return
    # the refracted ray seen from the other side reflects the same fraction
    assume(ci > 1e-3 and ct > 1e-3)
    assert ct == pytest.approx(math.sqrt(max(0.0, 1.0 - (ni / nt * si) ** 2)), rel=1e-12, abs=1e-12)
    r2, ci2 = oracle.fresnel(nt, ni, ct)
    assert ci2 == pytest.approx(ci, rel=1e-9, abs=1e-9)
    assert r2 == pytest.approx(r, rel=1e-7, abs=1e-12)


@SET
@given(g=st.floats(-0.99, 0.99, allow_nan=False).filter(lambda g: abs(g) > 1e-3),
       u=st.lists(st.floats(0.0, 1.0, allow_nan=False), min_size=2, max_size=8))
def test_hg_cos_endpoints_and_monotone(oracle, g, u):
    assert oracle.hg_cos(g, 0.0) == pytest.approx(-1.0, abs=1e-9)
    assert oracle.hg_cos(g, 1.0) == pytest.approx(1.0, abs=1e-9)
    u = sorted(u)
    c = [oracle.hg_cos(g, x) for x in u]
    assert all(-1.0 <= x <= 1.0 for x in c)
    assert all(b >= a - 1e-12 for a, b in zip(c, c[1:]))


unit = st.tuples(st.floats(-1, 1), st.floats(-1, 1), st.floats(-1, 1)).filter(
    lambda v: 0.1 < math.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2) <= 1.0)


@SET
@given(v=unit, c=st.floats(-1.0, 1.0, allow_nan=False), phi=st.floats(0.0, 1.0, allow_nan=False))
def test_spin_keeps_unit_length_and_turns_by_theta(oracle, v, c, phi):
    v = np.array(v) / np.linalg.norm(v)
    w = oracle.spin(v, c, phi)
    assert np.linalg.norm(w) == pytest.approx(1.0, abs=1e-12)
    # above |v_z| = 0.99999 the reference turns v as if it were (0, 0, sign v_z)
    # (_kernel.py:149-153), off by at most the tilt it ignores
    tilt = 0.0 if abs(v[2]) <= 0.99999 else 2.0 * math.sqrt(max(0.0, 1.0 - v[2] ** 2))
    assert float(np.dot(v, w)) == pytest.approx(c, abs=1e-6 + tilt)


@SET
@given(seed=st.integers(0, 2**31 - 1), n=st.integers(1, 400), nn=st.integers(1, 60),
       dr=st.floats(0.01, 1.0, allow_nan=False))
def test_radial_distance_conserves_weight(seed, n, nn, dr):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, nn * dr * 0.6, size=(3, n))
    w = rng.random(n)
    _, rd = metrics.radialDistance(p, w, nn, dr, 1)
    area = np.array([2 * np.pi * (i + 0.5) * dr ** 2 for i in range(nn)])
    r = np.sqrt(p[0] ** 2 + p[1] ** 2)
    # the same (i dr, (i+1) dr] edges radialDistance forms
    edges = np.array([i * dr for i in range(nn + 1)])
    inside = (r > 0) & (r <= edges[-1])
    assert float((rd * area).sum()) == pytest.approx(float(w[inside].sum()), rel=1e-12, abs=1e-12)


@SET
@given(budgets=st.lists(st.integers(0, 5000), min_size=1, max_size=12), batch=st.integers(1, 700))
def test_plate_schedule_partitions_every_budget(budgets, batch):
    L = _abi.lib()
    b = (C.c_uint64 * len(budgets))(*budgets)
    n = L.pmo_plate_schedule(b, len(budgets), batch, None, 0)
    assert n == sum((x + batch - 1) // batch for x in budgets)
    s = (C.c_uint32 * max(4 * n, 4))()
    assert L.pmo_plate_schedule(b, len(budgets), batch, s, n) == n
    s = np.array(list(s), np.uint32)[:4 * n].reshape(-1, 4)
    assert np.all((s[:, 2] >= 1) & (s[:, 2] <= batch))
    for i, N in enumerate(budgets):
        rows = s[s[:, 0] == i]
        ids = np.concatenate([np.arange(f, f + c) for _, f, c, _ in rows]) if len(rows) else np.zeros(0)
        assert np.array_equal(np.sort(ids), np.arange(N))


@SET
@given(slices=st.lists(st.integers(1, 12), min_size=1, max_size=5), l=st.sampled_from([0.5, 0.1, 0.05, 0.02]),
       nx=st.integers(1, 9), ny=st.integers(1, 9))
def test_plate_volume_layers_in_order_inside_the_shell(slices, l, nx, ny):
    th = [k * l for k in slices]
    vol = plate_volume(th, l, nx * l, ny * l)
    end = len(slices)
    assert vol.shape == (nx + 2, ny + 2, sum(slices) + 2)
    interior = vol[1:-1, 1:-1, 1:-1]
    expect = np.repeat(np.arange(end, dtype=np.uint8), slices)
    assert np.all(interior == expect[None, None, :])
    for face in (vol[0], vol[-1], vol[:, 0], vol[:, -1], vol[:, :, 0], vol[:, :, -1]):
        assert np.all(face == end)
