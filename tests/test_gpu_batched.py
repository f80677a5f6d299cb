"""GPU: the batched sweep reproduces the plain walks it replaces, and graph replays
reproduce eager sweeps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def plates():
    from pymonteopt_amd.batched import Plate
    return [
        Plate([1.0], [0.1], [10.0], [0.9], [1.37], voxel_space=0.05, x_size=2.55, y_size=2.55, n_packets=20000),
        Plate([0.5], [1.0], [5.0], [0.7], [1.45], voxel_space=0.05, x_size=5.05, y_size=5.05, n_packets=7000),
        Plate([0.1, 0.4], [0.5, 0.05], [30.0, 20.0], [0.8, 0.9], [1.44, 1.40], voxel_space=0.05, x_size=3.05,
              y_size=3.05, n_packets=15000),
        Plate([2.0], [0.01], [12.0], [0.95], [1.33], voxel_space=0.05, x_size=5.05, y_size=5.05, n_packets=12345,
              beam="flat", radius=0.5),
        Plate([0.3], [0.2], [8.0], [0.0], [1.0], voxel_space=0.05, x_size=2.05, y_size=2.05, n_packets=3000),
    ]


@pytest.mark.parametrize("layer_step", [False, True])
def test_batched_matches_plain_walks(layer_step):
    """Voxel stepping reproduces each plain walk to rounding of the FP64 sums. Layer
    stepping uses the same draws and differs only by FP32 rounding of positions and of
    the optical depth left, which flips a packet's history only rarely: per-instance
    totals agree to a few 1e-4 of a packet's weight per packet."""
    import torch
    from pymonteopt_amd.batched import PlateSweep
    from pymonteopt_amd.geometry import Source
    from pymonteopt_amd.runtime import TallySpec, Walker
    ps = plates()
    sw = PlateSweep(ps, nr=20, dr=0.1, batch=128, layer_step=layer_step)
    sw.run(seed=5, stream_id=40)
    torch.cuda.synchronize()
    res = sw.results()
    assert res["events"]["packets"] == sum(p.n_packets for p in ps)
    for i, p in enumerate(ps):
        kind, vox, w0 = p.source()
        wk = Walker(p.volume(), p.voxel_space, p.optics(), skip=False)   # voxel by voxel, as the sweep
        tal = wk.tallies(TallySpec(nr=20, dr=0.1, absorbed=False))
        wk.walk(Source(voxel=vox, w0=w0, kind=kind, radius=p.radius), p.n_packets, 5, tal, stream_id=40 + i)
        torch.cuda.synchronize()
        r = tal.to_host()
        s = r["scalars"]
        N = p.n_packets
        for j, k in enumerate(["Rd", "Tt", "other", "A", "kill", "gain", "w0"]):
            if layer_step:
                assert abs(res["scalars"][i, j] - s[k]) / N < 3e-4, (i, k, res["scalars"][i, j], s[k])
            else:
                assert res["scalars"][i, j] == pytest.approx(s[k], rel=1e-9, abs=1e-9), (i, k)
        if layer_step:
            assert np.abs(res["rd_r_raw"][i] - r["rd_r"]).sum() / N < 1e-3
        else:
            assert np.allclose(res["rd_r_raw"][i], r["rd_r"], rtol=1e-9, atol=1e-12)
        bal = sum(res["scalars"][i, j] for j in (0, 1, 2, 3, 4)) - res["scalars"][i, 5]
        assert abs(bal - res["scalars"][i, 6]) < 1e-9 * res["scalars"][i, 6]


def test_graph_replay_matches_eager():
    import torch
    from pymonteopt_amd.batched import PlateSweep
    sw = PlateSweep(plates(), nr=20, dr=0.1, batch=64)
    ref = {}
    big = 2 ** 31 + 5            # key words >= 2^31 reach the kernel unmasked on both paths
    for seed in (3, 9, big):
        sw.run(seed=seed, stream_id=big if seed == big else 0)
        torch.cuda.synchronize()
        ref[seed] = sw.results()["scalars"].copy()
    sw.run(seed=big & 0x7FFFFFFF, stream_id=big & 0x7FFFFFFF)
    torch.cuda.synchronize()
    assert not np.allclose(ref[big], sw.results()["scalars"])   # the key is not masked to 31 bits
    sw.capture()
    for seed in (3, 9, big, 3, 9):
        sw.replay(seed, stream_id=big if seed == big else 0)
        torch.cuda.synchronize()
        got = sw.results()["scalars"]
        assert np.allclose(got, ref[seed], rtol=1e-12, atol=1e-12), seed
        # eager work between replays
        x = torch.ones(1000, device="cuda:0").sum()
        torch.cuda.synchronize()
        assert float(x) == 1000.0


@pytest.mark.parametrize("layer_step", [False, True])
def test_wave_lds_tallies_match_direct_adds(layer_step):
    """Each wave sums its current plate's tallies in LDS while they fit (13 + 2 nr doubles a
    wave, at most 32 KiB a block) and adds them to global memory otherwise. With 400 radial
    bins a 256-lane block takes the LDS path (26 KiB) and a 1024-lane block the direct one
    (104 KiB); the packets and their draws are the same, so every instance's sums agree to
    the rounding of their order. Also the case the LDS path is for: waves that all walk the
    same two plates."""
    import torch
    from pymonteopt_amd.batched import Plate, PlateSweep
    two = [Plate([1.0], [0.1], [10.0], [0.9], [1.37], voxel_space=0.05, x_size=2.55, y_size=2.55, n_packets=200000),
           Plate([0.5], [1.0], [5.0], [0.7], [1.45], voxel_space=0.05, x_size=5.05, y_size=5.05, n_packets=150000)]
    for ps in (plates(), two):
        got = {}
        for bs in (256, 1024):
            sw = PlateSweep(ps, nr=400, dr=0.01, batch=64, layer_step=layer_step)
            sw.block_size = bs
            print("plates", len(ps), "layer_step", layer_step, "block", bs, flush=True)
            sw.run(seed=7, stream_id=3)
            torch.cuda.synchronize()
            r = sw.results()
            assert r["events"]["packets"] == sum(p.n_packets for p in ps)
            got[bs] = (r["scalars"].copy(), r["rd_r_raw"].copy(), sw.tt_r.cpu().numpy().copy(), r["events"])
        (s1, rd1, tt1, e1), (s2, rd2, tt2, e2) = got[256], got[1024]
        assert e1 == e2
        assert np.allclose(s1, s2, rtol=1e-11, atol=1e-9)
        assert np.allclose(rd1, rd2, rtol=1e-11, atol=1e-9)
        assert np.allclose(tt1, tt2, rtol=1e-11, atol=1e-9)
        assert rd1.sum() > 0 and tt1.sum() > 0
