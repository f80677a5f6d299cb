"""GPU: internal fluence. Fluence2D / Fluence3D binning (pmo_fluence_bin) against numpy
histograms, and the walk's on-line r-z fluence tally against the oracle, a closed form and
the absorbed-weight identity."""
import numpy as np
import pytest

from _parity import binned_z

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def slab():
    from pymonteopt_amd import presets
    return presets.slab_plate()


def points(n, seed, dr, dz, nr, nz, lateral):
    """Random points kept at least 1e-4 of a bin away from every bin edge, so float32 and
    float64 binning agree."""
    rng = np.random.default_rng(seed)
    p = np.stack([rng.uniform(-lateral, lateral, n), rng.uniform(-lateral, lateral, n),
                  rng.uniform(-0.2, nz * dz * 1.1, n)]).astype(np.float32)
    q = p.astype(np.float64)
    r = np.hypot(q[0], q[1])
    ok = np.ones(n, bool)
    for v, d in ((r, dr), (q[0], dr), (q[1], dr), (q[2], dz)):
        f = v / d - np.floor(v / d)
        ok &= (f > 1e-4) & (f < 1 - 1e-4)
    w = rng.uniform(0.1, 1.0, n).astype(np.float32)
    return p[:, ok], w[ok]


def test_fluence2d_and_3d_binning_match_numpy():
    from pymonteopt_amd.fluence import Fluence2D, Fluence3D
    nr, nz, dr, dz = 20, 15, 0.1, 0.2
    f2 = Fluence2D(nr=nr, nz=nz, dr=dr, dz=dz)
    f3 = Fluence3D(nr=nr, nz=nz, dr=dr, dz=dz)
    want2 = np.zeros((nr, nz))
    want3 = np.zeros((2 * nr, 2 * nr, nz))
    for seed in (1, 2):                                  # two calls accumulate
        p, w = points(60000, seed, dr, dz, nr, nz, lateral=2.5)
        f2.saveFluesnce(p, w)
        f3.saveFluesnce(p, w)
        q = p.astype(np.float64)
        h, _, _ = np.histogram2d(np.hypot(q[0], q[1]), q[2], bins=[np.arange(nr + 1) * dr, np.arange(nz + 1) * dz],
                                 weights=w.astype(np.float64))
        want2 += h
        h3, _ = np.histogramdd(q.T, bins=[np.arange(-nr, nr + 1) * dr, np.arange(-nr, nr + 1) * dr,
                                          np.arange(nz + 1) * dz], weights=w.astype(np.float64))
        want3 += h3 / (dz * dr * dr)
    assert np.allclose(f2.getArz(), want2, rtol=1e-6, atol=1e-9)
    assert np.allclose(f3.getArz(), want3, rtol=1e-6, atol=1e-6)
    assert np.allclose(f2.getArrayR(), (np.arange(nr) + 0.5) * dr)
    assert np.allclose(f3.getArrayR(), (np.arange(-nr, nr) + 0.5) * dr)
    assert np.allclose(f2.getArrayZ(), (np.arange(nz) + 0.5) * dz)


def _walk(wl, N, seed, rz, absorbed=False):
    from pymonteopt_amd.runtime import TallySpec, Walker
    wk = Walker(wl.labels, wl.pitch, wl.optics)
    tal = wk.tallies(TallySpec(nr=0, absorbed=absorbed, rz=rz))
    wk.walk(wl.source, N, seed, tal)
    return tal.to_host()


def test_rz_fluence_matches_oracle_and_identity(slab, oracle):
    rz = (25, 0.1, 25, 0.1)
    ma, ms, g, n = slab.optics_columns()
    N, seeds = 50000, range(10)
    gpu = np.array([_walk(slab, N, s, rz)["fluence_rz"] for s in seeds])
    ora = np.array([oracle.walk(slab.labels, slab.pitch, ma, ms, g, n, slab.source.as_dict(), N, 100 + s,
                                absorbed=False, rz=rz)["fluence_rz"] for s in seeds])
    o_m = ora.mean(0)
    big = o_m > 0.02 * o_m.max()
    z = binned_z(gpu[:, big], ora[:, big])
    # hundreds of bins, each z from 10 + 10 batches (t-distributed): bound the share of
    # bins beyond 3 and the largest one
    assert big.sum() > 50
    assert np.mean(z > 3) < 0.03 and np.max(z) < 6, (np.mean(z > 3), np.max(z))
    # one label, bins covering the slab (7.5 mm > the 7.14 mm corner radius, 5.1 mm > 5 mm
    # depth): mu_a times the sum is the absorbed weight A (float32 1/mu_a in the kernel)
    full = _walk(slab, N, 3, (75, 0.1, 51, 0.1))
    mua = float(np.float32(ma[0]))
    assert full["fluence_rz"].sum() * mua == pytest.approx(full["scalars"]["A"], rel=1e-6)


def test_rz_fluence_beer_lambert_on_gpu():
    # non-scattering slab: every packet drops its whole weight at depth z ~ Exp(mu_a) on the
    # axis; dz = 0.03 mm does not align with the 0.05 mm voxels
    from pymonteopt_amd.geometry import Source, center_add, optics_table, plate_volume, specular
    from pymonteopt_amd.presets import Workload
    ma, d, l, n = 2.0, 0.5, 0.05, 1.4
    vol = plate_volume([d], l, 1.05, 1.05)
    w0 = float(np.float32(1) - np.float32(specular(1.0, n)))
    wl = Workload("beer", vol, l, optics_table([ma], [0.0], [0.0], [n], 1.0),
                  Source(voxel=(center_add(vol.shape[0]), center_add(vol.shape[1]), 1), w0=w0))
    N, dz, nz = 400000, 0.03, 16
    F = _walk(wl, N, 4, (3, 0.1, nz, dz))["fluence_rz"]
    assert F[1:].sum() == 0.0
    k = np.arange(nz)
    p = np.exp(-ma * k * dz) - np.exp(-ma * (k + 1) * dz)
    got = F[0] * float(np.float32(ma)) / (w0 * N)
    assert np.all(np.abs(got - p) < 4 * np.sqrt(p * (1 - p) / N) + 1e-6)


def test_rz_tally_leaves_the_walk_unchanged(slab):
    a = _walk(slab, 100000, 9, (), absorbed=True)
    b = _walk(slab, 100000, 9, (20, 0.1, 20, 0.1), absorbed=True)
    for k in ("Rd", "Tt", "A", "w0"):
        assert a["scalars"][k] == pytest.approx(b["scalars"][k], rel=1e-12, abs=1e-12)
    assert np.allclose(a["absorbed"], b["absorbed"], rtol=1e-12, atol=1e-12)


def test_model_api_returns_normalised_rz_fluence():
    from pymonteopt_amd.voxel import VoxelPlateModel
    m = VoxelPlateModel(nPh=100000, fluence_rz=(20, 0.1, 30, 0.1))
    m.build(thickness=[5.0], ma=[0.1], ms=[10.0], g=[0.9], n=[1.37], n_air=1.0, x_size=10.1, y_size=10.1,
            voxel_space=0.1)
    res = m.start(rand_seed=1).get_result()
    F = res["fluence_rz"]
    assert F.shape == (20, 30) and np.all(F >= 0) and F[0, 0] > 0
    i = np.arange(20)
    vol = np.pi * ((i + 1) ** 2 - i ** 2) * 0.01 * 0.1
    raw = m._raw["fluence_rz"]
    assert np.allclose(F, raw / (vol[:, None] * 100000))
    assert np.allclose(res["rz_r"], (i + 0.5) * 0.1) and res["rz_z"].size == 30


def test_binning_at_the_lds_table_limit():
    """Tables of up to 8192 bins are summed in an LDS copy per block, larger ones with one
    global add per point: 128 x 64 = 8192 r-z bins and 8193 x 1 both match numpy. With
    thousands of bins along r, float32 r / dr is off by up to ~5e-4 of a bin, so the points
    keep 1e-2 of a bin from every edge."""
    from pymonteopt_amd.fluence import Fluence2D
    rng = np.random.default_rng(3)
    for nr, nz, dr, dz in ((128, 64, 0.02, 0.05), (8193, 1, 0.0004, 5.0)):
        lat = nr * dr * 0.8
        p = np.stack([rng.uniform(-lat, lat, 200000), rng.uniform(-lat, lat, 200000),
                      rng.uniform(-0.2, nz * dz * 1.1, 200000)]).astype(np.float32)
        q = p.astype(np.float64)
        ok = np.ones(p.shape[1], bool)
        for v, d in ((np.hypot(q[0], q[1]), dr), (q[2], dz)):
            fr = v / d - np.floor(v / d)
            ok &= (fr > 1e-2) & (fr < 1 - 1e-2)
        p, q = p[:, ok], q[:, ok]
        w = rng.uniform(0.1, 1.0, p.shape[1]).astype(np.float32)
        f = Fluence2D(nr=nr, nz=nz, dr=dr, dz=dz)
        f.saveFluesnce(p, w)
        want, _, _ = np.histogram2d(np.hypot(q[0], q[1]), q[2],
                                    bins=[np.arange(nr + 1) * dr, np.arange(nz + 1) * dz], weights=w.astype(np.float64))
        print("bins", nr * nz, "points", w.size, flush=True)
        assert want.sum() > 0.5 * w.sum()
        assert np.allclose(f.getArz(), want, rtol=1e-6, atol=1e-9)
