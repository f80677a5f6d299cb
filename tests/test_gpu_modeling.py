"""GPU: the Turing-pattern model (pmo_turing_step and friends) against the numpy oracle
from the same seeded initial fields."""
import json

import numpy as np
import pytest

from oracle import turing as tp

pytestmark = pytest.mark.gpu


def model(n_grid=8, length=4, seed=3, reps=10):
    from pymonteopt_amd.modeling import TuringPattern
    t = TuringPattern()
    t.set_params(grid=n_grid, dx=1 / 40, length=length, seed=seed, repetition=reps)
    return t


@pytest.mark.parametrize("steps", [1, 10])
def test_fields_match_the_oracle(steps):
    t = model(reps=steps)
    u, v = t.grow()
    n = t.gridsize
    u0, v0 = tp.initial_fields(n, 3)
    ku, kv = tp.inverse_symbol(n, t.du, t.dt, t.dx), tp.inverse_symbol(n, t.dv, t.dt, t.dx)
    uo, vo = u0.astype(np.float64), v0.astype(np.float64)
    for _ in range(steps):
        uo, vo = tp.step(uo, vo, ku, kv, t.dt)
    ug, vg = u.cpu().numpy(), v.cpu().numpy()
    # float32 transforms against float64: relative to the field's spread
    assert np.abs(ug - uo).max() < 2e-5 * max(1.0, np.abs(uo).max()), np.abs(ug - uo).max()
    assert np.abs(vg - vo).max() < 2e-5 * max(1.0, np.abs(vo).max()), np.abs(vg - vo).max()


def test_threshold_peaks_and_labels_match_the_oracle():
    import torch
    t = model(reps=30)
    u, _ = t.grow()
    field = u.cpu().numpy().astype(np.float64)
    bv, peaks = t.microarchitecture_of(u)
    thr = np.float32(tp.threshold(t.bv_tv))
    below = field <= thr
    assert bv == pytest.approx(1 - below.mean(), abs=0)
    thr_field = np.where(below, 0.0, field)
    assert np.array_equal(u.cpu().numpy(), thr_field.astype(np.float32))
    assert peaks == tp.count_peaks_last_axis(thr_field.astype(np.float32), 5)
    assert 0 < peaks < field.size


def test_modeling_end_to_end(tmp_path):
    t = model(reps=40)
    t.set_params(tile_num_xz=2, tile_num_y=1)
    U = t.modeling(str(tmp_path))
    n = t.gridsize
    assert U.shape == (2 * n, n, 2 * n) and U.dtype == np.int8 and set(np.unique(U)) <= {0, 1}
    assert abs(U.mean() - t.bv_tv_real) < 1e-12
    info = json.load(open(tmp_path / "_info.json"))
    assert info["model_shape"] == [n, n, n] and info["microarchitecture"]["BV/TV real"] == pytest.approx(t.bv_tv_real)
    with pytest.raises(NotImplementedError):
        t.modeling(str(tmp_path), save_dicom=True)
