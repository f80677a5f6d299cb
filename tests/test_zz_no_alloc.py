"""GPU: pmo_walk_noalloc (the walk with its claim counters in libpmo's static device scratch)
gives pmo_walk's results on the same inputs, on two calls in a row, the second smaller.

Both entry points run the same kernels on the same packet histories, so the integer event
counts agree exactly and the FP64 sums to the summation order of the atomics."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _walk_both(wl, n, seed, spec):
    import torch
    from pymonteopt_amd import _abi
    from pymonteopt_amd.runtime import Walker
    wk = Walker(wl.labels, wl.pitch, wl.optics)
    ref = wk.tallies(spec)
    wk.walk(wl.source, n, seed, ref)
    got = wk.tallies(spec)
    vol, s, r, t, keep = wk.c_args(wl.source, n, seed, got)
    st = torch.cuda.current_stream()
    print(f"pmo_walk_noalloc: {wl.name}, {n} packets, seed {seed}", flush=True)
    _abi.check(_abi.lib().pmo_walk_noalloc(C.byref(vol), C.byref(s), C.byref(r), C.byref(t),
                                           C.c_void_p(st.cuda_stream)))
    torch.cuda.synchronize()
    return ref.to_host(shape=wk.shape), got.to_host(shape=wk.shape)


def _same(ref, got):
    # every per-packet event count is exact; the global-deposit count is not compared: a
    # lane merges deposits into one voxel across the packets it happens to run in turn
    for k, v in ref["events"].items():
        if k != "global_deposits":
            assert got["events"][k] == v, k
    for k, v in ref["scalars"].items():
        assert got["scalars"][k] == pytest.approx(v, rel=1e-12, abs=1e-12), k
    for k in ("rd_r", "tt_r", "rd_top_r", "ring_w"):
        if k in ref:
            assert np.allclose(got[k], ref[k], rtol=1e-12, atol=1e-12), k
    if "absorbed" in ref:
        a, b = ref["absorbed"], got["absorbed"]
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(a).max())
        assert b.sum() == pytest.approx(a.sum(), rel=1e-12)


def test_noalloc_walk_matches_pmo_walk_twice():
    import torch
    from pymonteopt_amd import presets
    from pymonteopt_amd.runtime import TallySpec
    assert torch.cuda.is_available()
    # first call: the 101x101x50 slab, 3e5 packets, absorbed map and Rd(r)
    wl = presets.slab_plate()
    ref, got = _walk_both(wl, 300000, 11, TallySpec(nr=50, dr=0.1))
    assert got["events"]["packets"] == 300000
    _same(ref, got)
    # second call, smaller: the reference's 5 mm slab at 0.5 mm (12^3 padded), 4e4 packets.
    # The counters left by the first call must have been cleared, or no packet would launch.
    small = presets.plate("slab 5 mm @0.5mm", [5.0], [0.01], [10.0], [0.9], [1.5], 1.0, 0.5, 5.0, 5.0,
                          tallies=dict(nr=20, dr=0.25))
    assert small.labels.size < wl.labels.size
    ref2, got2 = _walk_both(small, 40000, 3, TallySpec(nr=20, dr=0.25))
    assert got2["events"]["packets"] == 40000
    assert got2["events"]["escaped"] + got2["events"]["killed"] == 40000
    _same(ref2, got2)
    # a third call, a flat beam with per-packet source draws, through the same scratch
    flat = presets.plate("slab flat", [5.0], [0.1], [10.0], [0.9], [1.37], 1.0, 0.1, 4.1, 4.1,
                         beam="flat", radius=0.5)
    ref3, got3 = _walk_both(flat, 20000, 5, TallySpec(nr=20, dr=0.1, rings=((0.2, 0.4), (0.4, 0.8))))
    _same(ref3, got3)
