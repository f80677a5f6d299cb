"""Host logic: plate voxelisation, parameter contract, sources, phantoms, validation."""
import numpy as np
import pytest

from pymonteopt_amd import geometry, presets
from pymonteopt_amd.voxel import PlateModel, set_params


def test_plate_volume_layout():
    vol = geometry.plate_volume([0.1, 2.0, 4.9], 0.05, 10.05, 10.05)
    assert vol.shape == (203, 203, 142)
    assert vol.dtype == np.uint8
    # layer i fills round(th_i/l) slices from z = 1 (pymopt/voxel/_classes.py:403-407)
    col = vol[100, 100]
    assert list(col[1:3]) == [0, 0] and col[3] == 1 and col[42] == 1 and col[43] == 2 and col[140] == 2
    assert col[0] == 3 and col[141] == 3
    for face in (vol[0], vol[-1], vol[:, 0], vol[:, -1]):
        assert np.all(face == 3)
    assert np.bincount(vol[1:-1, 1:-1, 1:-1].reshape(-1)).tolist() == [201 * 201 * 2, 201 * 201 * 40,
                                                                         201 * 201 * 98]


def test_plate_volume_rejects_inconsistent_rounding():
    # per-layer rounding 1 + 1 = 2 slices vs round(0.3/0.2) = 2 is fine; 0.25+0.25 at 0.1: 2+2 vs 5
    with pytest.raises(ValueError):
        geometry.plate_volume([0.25, 0.25], 0.1, 1.0, 1.0)


def test_set_params_contract():
    m = PlateModel()
    m.set_params({"x_size": 3}, voxel_space=0.2)
    assert m.params["x_size"] == 3 and m.voxel_space == 0.2
    with pytest.raises(KeyError):
        m.set_params(xy_size=3)
    d = {"a": 1}
    with pytest.raises(KeyError):
        set_params(d, ["a"], {"b": 2})


def test_center_and_specular():
    assert geometry.center_add(103) == 51
    assert geometry.center_add(104) == 51
    assert geometry.specular(1.0, 1.37) == pytest.approx((0.37 / 2.37) ** 2)
    assert geometry.specular(1.0, 1.0) == 0.0
    at, r = geometry.fresnel_oblique(30 / 180, 1.0, 1.4)
    assert r == pytest.approx(0.02778833340245436, rel=1e-12)


def test_optics_table_and_validation():
    t = geometry.optics_table([0.1, 0.2], [10, 20], [0.9, 0.8], [1.4, 1.5], 1.0)
    assert t.shape == (3, 4) and t.dtype == np.float32
    assert t[2, 3] == 1.0 and t[1, 1] == 20
    with pytest.raises(ValueError):
        geometry.optics_table([0.0], [0.0], [0.9], [1.4], 1.0)
    with pytest.raises(ValueError):
        geometry.optics_table([0.1], [1.0], [1.0], [1.4], 1.0)
    vol = geometry.plate_volume([1.0], 0.5, 2.0, 2.0)
    geometry.validate_volume(vol, 1)
    bad = vol.copy()
    bad[0, 1, 1] = 0
    with pytest.raises(ValueError):
        geometry.validate_volume(bad, 1)
    bad = vol.copy()
    bad[2, 2, 2] = 7
    with pytest.raises(ValueError):
        geometry.validate_volume(bad, 1)


def test_bone_phantom_structure():
    vol = geometry.bone_phantom(seed=3, n=64, pitch=0.02, cortical=0.1, soft=0.2)
    assert vol.shape == (64, 64, 64)
    geometry.validate_volume(vol, 4)
    inner = vol[1:-1, 1:-1, 1:-1]
    assert np.all(inner[:, :, :10] == 3)          # soft tissue top (0.2 mm / 0.02)
    assert np.all(inner[:, :, 10:15] == 2)        # cortical plate under it
    assert np.all(inner[:5, :, 10:] == 2) and np.all(inner[-5:, :, 10:] == 2)
    assert np.all(inner[:, :5, 10:] == 2) and np.all(inner[:, :, -5:] == 2)
    core = inner[5:-5, 5:-5, 15:-5]
    assert set(np.unique(core)) <= {0, 1}
    porosity = (core == 1).mean()
    assert porosity == pytest.approx(0.7, abs=1e-3)
    again = geometry.bone_phantom(seed=3, n=64, pitch=0.02, cortical=0.1, soft=0.2)
    assert np.array_equal(vol, again)


def test_gaussian_random_field_correlation():
    f = geometry.gaussian_random_field((48, 48, 48), 0.02, 0.15, seed=1)
    assert f.mean() == pytest.approx(0.0, abs=1e-9) and f.std() == pytest.approx(1.0)
    # correlation at lag d falls like exp(-d^2 / (4 L^2)) for a Gaussian-filtered field
    lag = 5
    c = np.mean(f * np.roll(f, lag, axis=0))
    assert c == pytest.approx(np.exp(-(lag * 0.02) ** 2 / (4 * 0.15 ** 2)), abs=0.1)


def test_presets():
    w = presets.slab_plate()
    assert w.labels.shape == (103, 103, 52) and w.source.voxel == (51, 51, 1)
    assert w.source.w0 == pytest.approx(1 - (0.37 / 2.37) ** 2, rel=1e-6)
    s = presets.skin_plate("flat")
    assert s.labels.shape == (203, 203, 142) and s.source.radius == 1.0
    assert s.source.w0 == pytest.approx(1 - (0.44 / 2.44) ** 2, rel=1e-6)


def test_wide_bone_phantom_tiles_the_core():
    """bone_phantom_wide: the n^3 phantom's trabecular core tiled across x and y, the same
    layers in depth, and the centre voxel over the same core cell."""
    kw = dict(seed=3, n=64, pitch=0.02, cortical=0.1, soft=0.2)
    small = geometry.bone_phantom(**kw)
    wide = geometry.bone_phantom_wide(tiles=3, **kw)
    core = small[6:-6, 6:-6, :]                   # 1 end + 5 cortical voxels on each side
    m = core.shape[0]
    assert wide.shape == (3 * m + 12, 3 * m + 12, 64)
    geometry.validate_volume(wide, 4)
    assert np.array_equal(wide[6:-6, 6:-6, :], np.tile(core, (3, 3, 1)))
    assert np.all(wide[1:6, 1:-1, 11:-1] == 2) and np.all(wide[1:-1, 1:-1, 1:11] == 3)
    c_small, c_wide = geometry.center_add(64), geometry.center_add(wide.shape[0])
    assert np.array_equal(wide[c_wide, c_wide, :], small[c_small, c_small, :])
    with pytest.raises(ValueError):
        geometry.bone_phantom_wide(tiles=2, **kw)
    w = presets.bone_wide_workload(seed=0)
    assert w.labels.shape == (1072, 1072, 256) and w.source.voxel == (535, 535, 1)
    assert len(w.tallies["rings"]) == 10 and w.tallies["rings"][-1] == (9.75, 10.25)
