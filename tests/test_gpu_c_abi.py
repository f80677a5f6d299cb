"""GPU: a C program (tests/c/walk_slab.c) drives libpmo.so through include/pmo.h alone, with
its own hipMalloc'd buffers, and gets the tallies the Python layer gets for the same seed."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "c", "walk_slab")


def test_c_program_matches_the_python_layer():
    import torch
    from pymonteopt_amd import presets
    from pymonteopt_amd.runtime import TallySpec, Walker
    assert os.path.exists(EXE), "tests/c/walk_slab is built by __graft_entry__.build() (make -C tests/c)"
    N, seed = 200000, 5
    print(f"C example: slab, {N} packets, seed {seed}", flush=True)
    r = subprocess.run([EXE, str(N), str(seed)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    c = json.loads(r.stdout.strip().splitlines()[-1])
    wl = presets.slab_plate()
    wk = Walker(wl.labels, wl.pitch, wl.optics, skip=False)   # the C program passes no skip radius
    tal = wk.tallies(TallySpec(nr=50, dr=0.1))
    wk.walk(wl.source, N, seed, tal)
    torch.cuda.synchronize()
    p = tal.to_host(shape=wk.shape)
    # same volume, optics, launch and draws: the same packet histories, so the sums agree
    # to the FP64 atomics' summation order
    for k, ck in (("Rd", "Rd"), ("Tt", "Tt"), ("A", "A"), ("w0", "w0")):
        assert c[ck] == pytest.approx(p["scalars"][k], rel=1e-12), k
    assert c["absorbed_sum"] == pytest.approx(p["absorbed"].sum(), rel=1e-12)
    assert c["interactions"] == p["events"]["interactions"] and c["crossings"] == p["events"]["crossings"]
    assert np.allclose(c["rd_r"], p["rd_r"], rtol=1e-12, atol=0)
