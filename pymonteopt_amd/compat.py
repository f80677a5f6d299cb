"""Kernel-level drop-in for pyMonteOpt's walk functions.

`vmc_kernel` has the signature and return value of pymopt/voxel/_kernel.py:23-28 (the Numba
walk, also the argument list of the CuPy launch in pymopt/voxel_gpu/_classes.py:105-111),
so a pymopt model can run on the MI355X by swapping one import. It reads the launch state
the model prepared (add, p, v, w of every packet), runs libpmo's walk on the GPU and returns
per-packet exit states that the model's own _encooder / getRdTtRate
(pymopt/voxel/_classes.py:261-293) turn into the same results.

Any launch state is accepted. When all packets share one (the pencil beam, optionally
oblique) the device generates it; otherwise the arrays themselves are handed to the walk
(PMO_BEAM_STATES), as with the reference's Gaussian beam.
"""
from __future__ import annotations

import numpy as np

from .geometry import Source
from .runtime import TallySpec, Walker

__all__ = ["vmc_kernel"]


def vmc_kernel(add, p, v, w, ma, ms, n, g, voxel_model, l, nPh, end_point, rand_seed=None):
    """Returns (add, p, v, w) after the walk, as the reference kernel leaves them: the exit
    voxel's address (the end-label voxel entered) and the position in its frame, so that
    l*(add - centre) + p is the exit point; the exit direction and weight; zeros for packets
    ended by roulette."""
    add = np.asarray(add)
    p = np.asarray(p, np.float32)
    v = np.asarray(v, np.float32)
    w = np.asarray(w, np.float32)
    N = int(nPh)
    if add.shape != (3, N) or p.shape != (3, N) or v.shape != (3, N) or w.shape != (N,):
        raise ValueError("add, p, v must be (3, nPh) and w (nPh,)")
    identical = all(np.all(arr == arr[:, :1]) for arr in (add, p, v)) and np.all(w == w[0])
    vm = np.ascontiguousarray(voxel_model).astype(np.uint8)
    end = int(end_point)
    nl = int(np.asarray(ma).size)
    if end != nl:
        raise ValueError("end_point must equal the number of tissue labels (the label tables' length)")
    opt = np.zeros((nl + 1, 4), np.float32)
    opt[:nl, 0] = np.asarray(ma, np.float32)
    opt[:nl, 1] = np.asarray(ms, np.float32)
    opt[:nl, 2] = np.asarray(g, np.float32)
    opt[:, 3] = np.asarray(n, np.float32)[: nl + 1]
    centre = (vm.shape[0] - 1) // 2, (vm.shape[1] - 1) // 2, 1
    on_face = np.isclose(float(p[2, 0]), -float(l) / 2, rtol=1e-5) and p[0, 0] == 0 and p[1, 0] == 0
    if identical and on_face:
        src = Source(voxel=tuple(int(x) for x in add[:, 0]), w0=float(w[0]),
                     dir=tuple(float(x) for x in v[:, 0]))
    else:
        # the caller's arrays; the LDS tally window sits at the top-face centre
        src = Source(voxel=centre, w0=1.0, states=(add.astype(np.int32), p, v, w))
    seed = np.random.randint(1e5) if rand_seed is None else int(rand_seed)
    wk = Walker(vm, float(l), opt)
    tal = wk.tallies(TallySpec(absorbed=False, exit_state=True, exit_voxel=True), N)
    wk.walk(src, N, seed, tal)
    r = tal.to_host()
    # the end-label voxel each packet entered and its position there (_kernel.py:110-116),
    # zeros for roulette kills (:134-137), as the reference kernel leaves its arrays
    es = r["exit_state"]
    return (r["exit_add"].astype(np.int32), r["exit_p"].astype(np.float32), es[3:6].astype(np.float32),
            es[6].astype(np.float32))
