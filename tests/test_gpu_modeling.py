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
    assert np.array_equal(U[:n], U[n:]) and np.array_equal(U[:, :, :n], U[:, :, n:])   # np.tile's layout
    assert abs(U.mean() - t.bv_tv_real) < 1e-12
    info = json.load(open(tmp_path / "_info.json"))
    assert info["model_shape"] == [n, n, n] and info["microarchitecture"]["BV/TV real"] == pytest.approx(t.bv_tv_real)
    # save_dicom: one 16-bit slice per z level of the field times ct_coef
    from pymonteopt_amd import dicom
    t2 = model(reps=5)
    u, _ = t2.grow()
    field = u.cpu().numpy()
    t2.modeling(str(tmp_path / "dcm"), save_dicom=True)
    files = sorted(p.name for p in (tmp_path / "dcm").glob("*.dcm"))
    assert len(files) == n and files[0] == "0000.dcm"
    sl = dicom.read_file(str(tmp_path / "dcm" / files[3]))
    want = (field * t2.ct_coef).astype("uint16")[:, :, 3]
    assert np.array_equal(sl.pixel_array, np.frombuffer(want.tobytes(), np.uint16).reshape(n, n))
    assert sl.PixelSpacing[0] == pytest.approx(t2.voxelsize)


@pytest.mark.parametrize("name", ["TuringModel_Rectangular", "TuringModel_Cylinder", "TuringModel_RnC"])
def test_turing_volume_walk_matches_oracle(name, oracle):
    from pymonteopt_amd.voxel import VoxelTuringModel
    t = model(n_grid=8, length=5, reps=30)             # 40^3 bone volume
    u, _ = t.grow()
    t.microarchitecture_of(u)
    bone = (u.cpu().numpy() > 0).astype(np.int8)
    m = VoxelTuringModel(nPh=40000, model_name=name, nr=20, dr=0.1)
    m.set_model(bone)
    kw = dict(voxel_space=0.05, th_cortical=0.1, th_subcutaneus=0.15, th_dermis=0.1)
    kw.update(dict(xz_size=1.4) if name != "TuringModel_Cylinder" else {})
    kw.update(dict(r_bone=0.9) if name != "TuringModel_Rectangular" else {})
    m.build(**kw)
    vm = m.model.voxel_model
    src = m.source()
    assert vm[src.voxel] == (4 if name == "TuringModel_Cylinder" else vm[src.voxel])
    rd = []
    for s in range(6):
        m.start(rand_seed=s)
        r = m.get_result()
        rd.append(r["Rd"])
        assert abs(sum(r["energy"][k] for k in ("Rd", "Tt", "other", "A", "kill")) - r["energy"]["gain"]
                   - r["energy"]["w0"]) < 1e-9
    opt = m.model.optics().astype(np.float64)
    L = opt.shape[0] - 1
    ora = [oracle.walk(vm, 0.05, opt[:L, 0], opt[:L, 1], opt[:L, 2], opt[:, 3], src.as_dict(), 40000, 50 + s,
                       absorbed=False)["scalars"]["Rd"] / 40000 for s in range(6)]
    a, b = np.array(rd), np.array(ora)
    se = np.sqrt(a.var(ddof=1) / a.size + b.var(ddof=1) / b.size)
    assert abs(a.mean() - b.mean()) < 4 * se + 1e-6, (a.mean(), b.mean(), se)
