"""GPU: the walk at the edges of its inputs.

* no packets: the call succeeds and leaves every tally untouched;
* one packet: it launches, ends, and its energy balances;
* the smallest volume the walk takes (3 x 3 x 3: one interior voxel in the end-label shell)
  against the oracle on the same seeds;
* packet ids past 2^32 (the counter's high word, pymopt has no such limit since its arrays
  are per packet): a run whose ids straddle 2^32 equals its two halves and follows the
  oracle's draws for the same ids;
* the largest label table (255 labels plus the end label).
"""
import numpy as np
import pytest

from _parity import gpu_run, ledger_error, oracle_run, packet_events

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


@pytest.fixture(scope="module")
def slab(torch_cuda):
    from pymonteopt_amd import presets
    from pymonteopt_amd.runtime import Walker
    wl = presets.slab_plate()
    return wl, Walker(wl.labels, wl.pitch, wl.optics)


def test_zero_and_one_packet(slab):
    wl, wk = slab
    from pymonteopt_amd.runtime import TallySpec
    spec = TallySpec(**wl.tallies)
    tal = wk.tallies(spec)
    print("zero packets", flush=True)
    wk.walk(wl.source, 0, 3, tal)
    r = tal.to_host(shape=wk.shape)
    assert all(v == 0 for v in r["scalars"].values())
    assert all(v == 0 for v in r["events"].values())
    assert not r["absorbed"].any() and not r["rd_r"].any()
    print("one packet", flush=True)
    one = gpu_run(wk, wl, 1, 3)
    assert one["events"]["packets"] == 1
    assert one["events"]["escaped"] + one["events"]["killed"] == 1
    assert ledger_error(one["scalars"]) < 1e-12
    assert one["absorbed"].sum() == pytest.approx(one["scalars"]["A"], rel=1e-12, abs=1e-15)


def test_single_interior_voxel(torch_cuda, oracle):
    """A 3 x 3 x 3 volume: one tissue voxel (0.5 mm, mu_a 1, mu_s 20, g 0.8, n 1.4) inside
    the end-label shell, pencil beam into its top face, 2e5 packets x 6 seeds on the same
    draws as the oracle."""
    from pymonteopt_amd import presets
    from pymonteopt_amd.runtime import Walker
    wl = presets.plate("one voxel", [0.5], [1.0], [20.0], [0.8], [1.4], 1.0, 0.5, 0.5, 0.5, tallies=dict(nr=5, dr=0.1))
    assert wl.labels.shape == (3, 3, 3)
    wk = Walker(wl.labels, wl.pitch, wl.optics)
    N = 200000
    G, O = [], []
    for s in range(6):
        print("one voxel seed", s, flush=True)
        g = gpu_run(wk, wl, N, s)
        o = oracle_run(oracle, wl, N, s)
        assert g["events"]["packets"] == N
        assert ledger_error(g["scalars"]) < 1e-9
        G.append([g["scalars"][k] / N for k in ("Rd", "Tt", "A")])
        O.append([o["scalars"][k] / N for k in ("Rd", "Tt", "A")])
    G, O = np.array(G), np.array(O)
    spread = O.std(axis=0, ddof=1)
    # common random numbers: per-seed differences far below the seed-to-seed spread
    assert np.all(np.abs(G - O).max(axis=0) < spread), (G - O, spread)
    assert np.all(np.abs(G.mean(0) - O.mean(0)) < 3 * np.sqrt(2.0 / len(G)) * spread)


def test_packet_ids_past_2_32(slab, oracle):
    """Packets [2^32 - 5e4, 2^32 + 5e4): the run equals its halves below and above 2^32
    (every event count, sums to their order), and follows the oracle walking the same ids."""
    import torch
    from pymonteopt_amd.runtime import TallySpec
    wl, wk = slab
    base, half = (1 << 32) - 50000, 50000
    spec = TallySpec(**wl.tallies)
    print("ids across 2^32", flush=True)
    whole = gpu_run(wk, wl, 2 * half, 5, first_packet=base)
    tal = wk.tallies(spec)
    wk.walk(wl.source, half, 5, tal, first_packet=base)
    wk.walk(wl.source, half, 5, tal, first_packet=base + half)
    torch.cuda.synchronize()
    parts = tal.to_host(shape=wk.shape)
    assert packet_events(parts) == packet_events(whole)
    for k in ("Rd", "Tt", "A", "w0"):
        assert parts["scalars"][k] == pytest.approx(whole["scalars"][k], rel=1e-12), k
    G, O = [], []
    for s in range(5):
        g = gpu_run(wk, wl, 2 * half, s, first_packet=base)
        o = oracle_run(oracle, wl, 2 * half, s, first_packet=base)
        G.append([g["scalars"][k] for k in ("Rd", "Tt", "A")])
        O.append([o["scalars"][k] for k in ("Rd", "Tt", "A")])
    G, O = np.array(G), np.array(O)
    spread = O.std(axis=0, ddof=1)
    assert np.all(np.abs(G - O).max(axis=0) < spread), (G - O, spread)


def test_largest_label_table(torch_cuda):
    """255 tissue labels (the end label is 255): a column of 255 thin layers with indices
    stepping between 1.33 and 1.45 walks, every packet ends and the energy balances, with
    face skipping (label 255 counts as a boundary) and without."""
    from pymonteopt_amd import geometry
    from pymonteopt_amd.geometry import Source
    from pymonteopt_amd.runtime import TallySpec, Walker
    L = 255
    vol = np.empty((7, 7, L + 2), np.uint8)
    vol[:, :, 1:-1] = np.arange(L, dtype=np.uint8)[None, None, :]
    geometry.shell(vol, L)
    rng = np.random.default_rng(0)
    optics = geometry.optics_table(rng.uniform(0.01, 0.5, L), rng.uniform(1, 20, L), rng.uniform(0.5, 0.95, L),
                                   rng.uniform(1.33, 1.45, L), 1.0)
    src = Source(voxel=(3, 3, 1), w0=1.0)
    for skip in (True, False):
        wk = Walker(vol, 0.05, optics, skip=skip)
        tal = wk.tallies(TallySpec(nr=10, dr=0.05))
        print("255 labels, skip", skip, flush=True)
        wk.walk(src, 100000, 1, tal)
        torch_cuda.cuda.synchronize()
        r = tal.to_host(shape=wk.shape)
        assert r["events"]["packets"] == 100000
        assert r["events"]["escaped"] + r["events"]["killed"] == 100000
        assert r["events"]["fresnel"] > 0
        assert ledger_error(r["scalars"]) < 1e-9
