"""GPU: the differentiable walk. Jacobians against the oracle's path sums and central
finite differences with common random numbers; the replay adjoint against its own
forward, against finite differences, and its cost."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def skin():
    from pymonteopt_amd import presets
    return presets.skin_plate()


@pytest.fixture(scope="module")
def grid():
    from pymonteopt_amd import presets
    return presets.grf_grid(seed=0)


def label_field(wl):
    import torch
    mua = torch.tensor(wl.optics[:, 0], device="cuda:0")
    f = mua[torch.as_tensor(np.minimum(wl.labels, wl.n_labels), device="cuda:0").long()].float()
    return f


def test_path_sums_match_oracle(skin, oracle):
    import torch
    from pymonteopt_amd.adjoint import DiffWalk
    N = 20000
    dw = DiffWalk(skin.labels, skin.pitch, skin.optics, skin.source, n_packets=N, nr=40, dr=0.1)
    ma, ms, g, n = skin.optics_columns()
    area = dw.area.cpu().numpy()
    for s in range(3):
        f = dw.forward(seed=s)
        torch.cuda.synchronize()
        o = oracle.walk(skin.labels, skin.pitch, ma, ms, g, n, skin.source.as_dict(), N, s, nr=40, dr=0.1,
                        path_stats=True, absorbed=False)
        wl = -(f["J_mua"].cpu().numpy() * (area[:, None] * N)).sum(0)
        assert np.allclose(wl, o["rd_wl"].sum(0), rtol=0.01)
        rd = f["Rd_r"].cpu().numpy() * area * N
        assert rd.sum() == pytest.approx(o["rd_r"].sum(), rel=0.005)


def test_skipping_forward_matches_voxel_forward(skin):
    """forward(replay=False) crosses same-label cubes in one step: the same draws, so per
    seed its Rd(r) and per-label path sums differ from the voxel-stepping forward's by
    rounding only (a small fraction of the seed-to-seed spread); its record is refused by
    the adjoint."""
    import torch
    from pymonteopt_amd.adjoint import DiffWalk
    N = 200_000
    dw = DiffWalk(skin.labels, skin.pitch, skin.optics, skin.source, n_packets=N, nr=40, dr=0.1)
    assert dw.walker.skip                              # the skin plate's layers are uniform cubes
    area = dw.area.cpu().numpy()
    A, B = [], []
    for s in range(6):
        a = dw.forward(seed=s, jac=True, replay=False)
        b = dw.forward(seed=s, jac=True)
        torch.cuda.synchronize()
        assert a["replayable"] is False and b["replayable"] is True
        for out, acc in ((a, A), (b, B)):
            wl = -(out["J_mua"].cpu().numpy() * area[:, None]).sum(0)          # sum W L_j / N
            wn = (out["J_mus"].cpu().numpy() * area[:, None]).sum(0)           # sum W (n_j/mu_s - L_j) / N
            acc.append(np.concatenate([[float((out["Rd_r"].cpu().numpy() * area).sum())], wl, wn]))
    A, B = np.array(A), np.array(B)
    spread = B.std(0, ddof=1)
    assert np.all(np.abs(A - B) <= 0.5 * spread + 1e-12), (np.abs(A - B) / spread)
    with pytest.raises(ValueError, match="replay"):
        dw.adjoint(a, dw.area.clone())


def _fd(dw, base, j, col, rel, seeds):
    area = dw.area
    out = []
    for s in seeds:
        h = rel * base[j, col]
        vals = []
        for sign in (1, -1):
            op = base.copy()
            op[j, col] += sign * h
            dw.set_optics(op)
            vals.append(float((dw.forward(seed=s, jac=False)["Rd_r"] * area).sum()))
        out.append((vals[0] - vals[1]) / (2 * h))
    dw.set_optics(base)
    return float(np.mean(out))


def test_label_gradients_vs_central_fd(skin):
    """d(detected Rd)/d(mu_a, mu_s) of dermis and subcutis within 5% of central FD at the
    larger relative steps 1e-2 (mu_a) and 5e-2 (mu_s), where two seeds of 4e6 packets
    suffice (test_gradients_at_the_specified_fd_step checks the 1e-3 step)."""
    from pymonteopt_amd.adjoint import DiffWalk
    dw = DiffWalk(skin.labels, skin.pitch, skin.optics, skin.source, n_packets=4_000_000, nr=40, dr=0.1)
    base = dw.base_optics.copy()
    seeds = (7, 8)
    pm = {}
    for s in seeds:
        f = dw.forward(seed=s, jac=True)
        for name in ("J_mua", "J_mus"):
            pm.setdefault(name, []).append((f[name] * dw.area[:, None]).sum(0).cpu().numpy())
    for j in (1, 2):
        a = np.mean([x[j] for x in pm["J_mua"]])
        s_ = np.mean([x[j] for x in pm["J_mus"]])
        fa = _fd(dw, base, j, 0, 1e-2, seeds)
        fs = _fd(dw, base, j, 1, 5e-2, seeds)
        assert abs(a - fa) < 0.05 * abs(fa), (j, a, fa)
        assert abs(s_ - fs) < 0.05 * abs(fs), (j, s_, fs)


def _fd_seeds(evaluate, rel, seeds):
    """Central FD values of a scalar evaluate(sign * rel, seed), one per seed (common random
    numbers within a seed)."""
    return np.array([(evaluate(rel, s) - evaluate(-rel, s)) / (2 * rel) for s in seeds])


def test_gradients_at_the_specified_fd_step(skin, grid):
    """The gradients against central FD at relative step 1e-3 with common random numbers and
    1e7 packets per evaluation: per-label mu_a and mu_s of dermis and subcutis on the skin
    plate, and the per-voxel mu_a gradient summed over 6x6x4 and 20x20x10 blocks under the
    source on the 64^3 grid. Over 50 seeds the FD mean has a standard error below 2% of
    itself and lies within 5% of the pMC / replay gradient (mean of 4 seeds)."""
    import torch
    from pymonteopt_amd.adjoint import DiffWalk
    N, REL, seeds = 10_000_000, 1e-3, range(3000, 3050)
    dw = DiffWalk(skin.labels, skin.pitch, skin.optics, skin.source, n_packets=N, nr=40, dr=0.1)
    base = dw.base_optics.copy()
    area = dw.area
    pm = []
    for s in range(4):
        f = dw.forward(seed=100 + s, jac=True)
        pm.append([(f["J_mua"] * area[:, None]).sum(0).cpu().numpy(),
                   (f["J_mus"] * area[:, None]).sum(0).cpu().numpy()])
    pm = np.mean(pm, axis=0)

    def label_eval(j, col):
        def ev(rel, s):
            op = base.copy()
            op[j, col] *= 1 + rel
            dw.set_optics(op)
            return float((dw.forward(seed=s, jac=False)["Rd_r"] * area).sum())
        return ev

    for j in (1, 2):
        for col in (0, 1):
            fd = _fd_seeds(label_eval(j, col), REL, seeds) / base[j, col]
            dw.set_optics(base)
            m, se = fd.mean(), fd.std(ddof=1) / np.sqrt(fd.size)
            print(f"label {j} {'mu_a' if col == 0 else 'mu_s'}: FD {m:.6g} +- {se:.2g}, pMC {pm[col, j]:.6g}",
                  flush=True)
            assert se < 0.02 * abs(m), (j, col, m, se)
            assert abs(pm[col, j] - m) < 0.05 * abs(m), (j, col, pm[col, j], m)
    del dw
    dg = DiffWalk(grid.labels, grid.pitch, grid.optics, grid.source, n_packets=N, nr=40, dr=0.1)
    field = label_field(grid)
    gb = dg.area.clone()
    c = grid.source.voxel
    for blk in ((slice(c[0] - 3, c[0] + 3), slice(c[1] - 3, c[1] + 3), slice(1, 5)),
                (slice(c[0] - 10, c[0] + 10), slice(c[1] - 10, c[1] + 10), slice(1, 11))):
        pmc = []
        for s in range(4):
            grad = dg.adjoint(dg.forward(seed=200 + s, mua_field=field, jac=False), gb)
            pmc.append(float((grad[blk].double() * field[blk].double()).sum()))
        pmc = float(np.mean(pmc))

        def ev(rel, s):
            fp = field.clone()
            fp[blk] *= 1 + rel
            return float((dg.forward(seed=s, mua_field=fp, jac=False)["Rd_r"] * gb).sum())

        fd = _fd_seeds(ev, REL, seeds)           # d/d(log mu_a) of the block
        m, se = fd.mean(), fd.std(ddof=1) / np.sqrt(fd.size)
        print(f"block {[b.stop - b.start for b in blk]}: FD {m:.6g} +- {se:.2g}, replay {pmc:.6g}", flush=True)
        assert se < 0.02 * abs(m), (m, se)
        assert abs(pmc - m) < 0.05 * abs(m), (pmc, m)
    torch.cuda.synchronize()


def test_voxel_gradient_by_replay(grid):
    import torch
    from pymonteopt_amd.adjoint import DiffWalk, rd_of_absorption_field
    dg = DiffWalk(grid.labels, grid.pitch, grid.optics, grid.source, n_packets=2_000_000, nr=40, dr=0.1)
    field = label_field(grid)
    gb = dg.area.clone()                         # loss = detected weight within 4 mm
    fwd = dg.forward(seed=5, mua_field=field, jac=True)
    grad = dg.adjoint(fwd, gb)
    torch.cuda.synchronize()
    assert dg.last_mismatch == 0     # every replayed path ended as recorded
    gc = grad.cpu().numpy()
    Jl = (fwd["J_mua"] * gb[:, None]).sum(0).cpu().numpy()
    for j in range(grid.n_labels):
        # forward sums path lengths per packet in FP32, the replay in FP64
        assert gc[grid.labels == j].sum() == pytest.approx(Jl[j], rel=1e-6)
    assert np.all(gc[grid.labels == grid.n_labels] == 0)
    # central FD of a 20x20x10 block under the source, relative step 0.1, CRN
    c = grid.source.voxel
    blk = (slice(c[0] - 10, c[0] + 10), slice(c[1] - 10, c[1] + 10), slice(1, 11))
    pmc = float((grad[blk].double() * field[blk].double()).sum())
    fds = []
    for s in (5, 6):
        vals = []
        for sign in (1, -1):
            fp = field.clone()
            fp[blk] *= (1 + sign * 0.1)
            vals.append(float((dg.forward(seed=s, mua_field=fp, jac=False)["Rd_r"] * gb).sum()))
        fds.append((vals[0] - vals[1]) / 0.2)
    fd = float(np.mean(fds))
    assert abs(pmc - fd) < 0.05 * abs(fd), (pmc, fd)
    # the autograd path returns the same map
    fld = field.clone().double().requires_grad_(True)
    (rd_of_absorption_field(fld, dg, seed=5) * gb).sum().backward()
    assert torch.allclose(fld.grad, grad, rtol=1e-9, atol=1e-15)
    # a per-label forward on the same DiffWalk between the field's forward and its backward
    # rewrites the walker's optics table; the replay still walks with the forward's table
    from pymonteopt_amd.adjoint import rd_of_label_optics
    fld2 = field.clone().double().requires_grad_(True)
    y = (rd_of_absorption_field(fld2, dg, seed=5) * gb).sum()
    L = grid.n_labels
    mua = torch.tensor(grid.optics[:L, 0] * 1.3, dtype=torch.float64, device=dg.device)
    mus = torch.tensor(grid.optics[:L, 1] * 0.7, dtype=torch.float64, device=dg.device)
    rd_of_label_optics(mua, mus, dg, seed=9)
    y.backward()
    assert dg.last_mismatch == 0
    assert torch.allclose(fld2.grad, grad, rtol=1e-9, atol=1e-15)
    # a record replayed under another key diverges: the adjoint refuses the gradient
    with pytest.raises(RuntimeError, match="left differently"):
        dg.adjoint(dict(fwd, seed=6), gb)
    assert dg.last_mismatch > 0


@pytest.mark.parametrize("n_packets", [200_000, 1_200_000])
def test_replay_line_buffers_match_per_voxel_atomics(skin, n_packets):
    """The replay's line buffers (per-lane 2x2x2-voxel cubes handed over eight per atomic
    instruction) change only the order of the FP64 sums: the map equals the one from one
    atomic per voxel deposit. 2e5 packets take 256-lane blocks, 1.2e6 the wide launch."""
    import torch
    from pymonteopt_amd.adjoint import DiffWalk
    dw = DiffWalk(skin.labels, skin.pitch, skin.optics, skin.source, n_packets=n_packets, nr=40, dr=0.1)
    fwd = dw.forward(seed=11, mua_field=label_field(skin), jac=False)
    gb = dw.area.clone()
    dw.lines = True
    g_lines = dw.adjoint(fwd, gb)
    dw.lines = False
    g_vox = dw.adjoint(fwd, gb)
    torch.cuda.synchronize()
    assert dw.last_mismatch == 0
    assert int((g_vox != 0).sum()) > 1000            # deposits far outside the LDS box too
    assert torch.equal(g_lines != 0, g_vox != 0)
    assert torch.allclose(g_lines, g_vox, rtol=1e-9, atol=1e-15 * float(g_vox.abs().max()))
    # a second replay through the same scratch (cleared by the library) repeats the first
    dw.lines = True
    assert torch.allclose(dw.adjoint(fwd, gb), g_lines, rtol=1e-9, atol=1e-15 * float(g_vox.abs().max()))
    # "auto" times both on its first adjoint and keeps one of them
    dw.lines = "auto"
    assert torch.allclose(dw.adjoint(fwd, gb), g_lines, rtol=1e-9, atol=1e-15 * float(g_vox.abs().max()))
    assert dw.lines in (True, False)


def test_adjoint_cost_vs_forward(grid, skin):
    import torch
    from pymonteopt_amd.adjoint import DiffWalk

    def t(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        fn()
        e0.record()
        r = fn()
        e1.record()
        e1.synchronize()
        return r, e0.elapsed_time(e1)

    for wl, field in ((grid, label_field(grid)), (skin, None)):
        dw = DiffWalk(wl.labels, wl.pitch, wl.optics, wl.source, n_packets=2_000_000, nr=40, dr=0.1)
        fwd, tf = t(lambda: dw.forward(seed=3, mua_field=field, jac=False))
        _, ta = t(lambda: dw.adjoint(fwd, dw.area.clone()))
        assert ta <= 3.0 * tf, (wl.name, ta, tf)


def _log_param_sigma(dw, mua, musp, g, target, n_jac=5, n_var=24):
    """Predicted standard error of the fitted log mu_a, log mu_s' of each layer from the
    target's own noise (linearised least squares of the fit's relative-error loss): the
    Jacobian of Rd(r) at the truth from n_jac x 1e7 packets of pMC sums, the per-bin
    variance of a 1e7-packet Rd(r) from n_var 1e6-packet seeds."""
    opt = dw.base_optics.copy()
    opt[:3, 0] = mua
    opt[:3, 1] = musp / (1 - g)
    dw.set_optics(opt)
    dw.n_packets = 10_000_000
    Js = []
    for s in range(n_jac):
        f = dw.forward(seed=1000 + s, jac=True)
        Js.append(np.concatenate([f["J_mua"].cpu().numpy()[:, :3] * mua[None, :],
                                  f["J_mus"].cpu().numpy()[:, :3] * (musp / (1 - g))[None, :]], 1))
    J = np.mean(Js, 0)
    dw.n_packets = 1_000_000
    var = np.array([dw.forward(seed=200 + k, jac=False)["Rd_r"].cpu().numpy() for k in range(n_var)]).var(0, ddof=1)
    dw.set_optics(dw.base_optics)
    W = 1.0 / np.asarray(target) ** 2
    Ai = np.linalg.inv(J.T @ (W[:, None] * J))
    cov = Ai @ (J.T @ ((W * (var / 10) * W)[:, None] * J)) @ Ai
    return np.sqrt(np.diag(cov))


def test_adam_recovers_layer_optics(skin):
    """Inverse problem: 3-layer (mu_a, mu_s') drawn within +-30% of the skin values,
    target Rd(r) from a 1e7-packet walk, Adam from the nominal values. Every one of the six
    parameters must land within 4 predicted standard errors of its hidden value, the error
    Rd(r) at 1e7 packets leaves it (the Fisher matrix of the fit's loss at the truth): that
    is 0.5% and 1.5% for the epidermis and dermis mu_s', but 7-21% for the three mu_a and the
    subcutis mu_s', which Rd(r) over 4 mm only weakly separates (epidermis and dermis mu_a
    are anti-correlated at -0.99)."""
    from pymonteopt_amd.adjoint import DiffWalk, fit_layers
    dw = DiffWalk(skin.labels, skin.pitch, skin.optics, skin.source, n_packets=10_000_000, nr=40, dr=0.1)
    rng = np.random.default_rng(1234)
    g = np.array([0.8, 0.9, 0.9])
    mua0 = np.array([0.5, 0.05, 0.02])
    musp0 = np.array([30.0, 20.0, 12.0]) * (1 - g)
    mua_t = mua0 * rng.uniform(0.7, 1.3, 3)
    musp_t = musp0 * rng.uniform(0.7, 1.3, 3)
    opt = dw.base_optics.copy()
    opt[:3, 0] = mua_t
    opt[:3, 1] = musp_t / (1 - g)
    dw.set_optics(opt)
    target = dw.forward(seed=99, jac=False)["Rd_r"].cpu().numpy()
    dw.set_optics(dw.base_optics)
    sigma = _log_param_sigma(dw, mua_t, musp_t, g, target)
    assert np.all(sigma[3:5] < 0.03)              # the two well-determined mu_s'
    dw.n_packets = 2_000_000
    mua_f, musp_f, hist = fit_layers(dw, target, g, mua0, musp0, iters=600, lr=0.04, lr_final=0.004, average=100)
    assert hist[-1] < 0.05 * hist[0]
    err = np.log(np.concatenate([mua_f / mua_t, musp_f / musp_t]))
    assert np.all(np.abs(err) < 4 * sigma), (err, sigma)
    assert abs(err[3]) < 0.05 and abs(err[4]) < 0.05
