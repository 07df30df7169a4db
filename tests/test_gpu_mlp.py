"""The learned-guess network on the GPU, checked against the CPU oracle, and
the fallback Newton entry point."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from goma_amd import capi, continuation, learned, problems  # noqa: E402
from oracle import MlpOracle  # noqa: E402

SHIFT = [0.5, 0.5, 4.5]
SCALE = [2.0, 2.0, 1 / 1.5]


def _data(n, seed=1):
    rng = np.random.default_rng(seed)
    x = np.stack([rng.uniform(0, 1, n), rng.uniform(0, 1, n), rng.uniform(3, 6, n)], 1)
    t = rng.standard_normal((n, 4))
    m = (rng.uniform(size=(n, 4)) > 0.3).astype(np.float32)
    return x.astype(np.float32), t.astype(np.float32), m


def _layer_slices(n_in, n_out, hidden, width):
    out, off, k = [], 0, n_in
    for l in range(hidden + 1):
        n = n_out if l == hidden else width
        out.append((off, off + n * k + n))
        off += n * k + n
        k = n
    return out


@pytest.mark.parametrize("npts", [1000, 64])
def test_gradient_matches_oracle(npts):
    print(f"[case] mlp gradient npts={npts}", flush=True)
    net = capi.Mlp(3, 4, hidden=4, width=256, seed=7, in_shift=SHIFT, in_scale=SCALE)
    o = MlpOracle(3, 4, 4, 256, in_shift=SHIFT, in_scale=SCALE)
    p = net.get_params()
    assert p.size == o.num_params == 3 * 256 + 256 + 3 * (256 * 256 + 256) + 256 * 4 + 4
    x, t, m = _data(npts)
    lg, gg = net.loss_grad(x, t, m)
    lo, go = o.loss_grad(p, x, t, m, quant=True)
    assert abs(lg - lo) <= 1e-4 * lo, (lg, lo)
    for a, b in _layer_slices(3, 4, 4, 256):
        rel = np.linalg.norm(gg[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        assert rel < 5e-3, (a, b, rel)
    # and the BF16 gradient is close to the exact one (the rounding is benign)
    _, ge = o.loss_grad(p, x, t, m, quant=False)
    assert np.linalg.norm(gg - ge) / np.linalg.norm(ge) < 5e-2


def test_forward_matches_oracle_and_params_roundtrip():
    print("[case] mlp forward", flush=True)
    net = capi.Mlp(3, 4, hidden=4, width=256, seed=11, in_shift=SHIFT, in_scale=SCALE)
    o = MlpOracle(3, 4, 4, 256, in_shift=SHIFT, in_scale=SCALE)
    rng = np.random.default_rng(2)
    p = (net.get_params() + rng.standard_normal(o.num_params).astype(np.float32) * 0.01)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    x, _, _ = _data(3001, seed=4)
    y = net.predict(x)
    yo = o.forward(p, x, quant=True)
    assert np.max(np.abs(y - yo)) <= 2e-3 * np.max(np.abs(yo)), np.max(np.abs(y - yo))


def test_training_fits_and_is_deterministic():
    print("[case] mlp training", flush=True)
    rng = np.random.default_rng(0)
    n = 20000
    x = np.stack([rng.uniform(0, 1, n), rng.uniform(0, 1, n), rng.uniform(3, 6, n)], 1).astype(np.float32)
    s = (x[:, 2] - 3) / 3
    t = np.stack([np.sin(np.pi * x[:, 0]) * np.cos(np.pi * x[:, 1]) * (1 + s),
                  x[:, 0] * x[:, 1] - s, np.cos(2 * x[:, 1]) * s, x[:, 0] ** 2 + 0.3 * s], 1)
    m = np.ones_like(t, dtype=np.float32)
    runs = []
    for rep in range(2):
        net = capi.Mlp(3, 4, hidden=4, width=256, seed=3, in_shift=SHIFT, in_scale=SCALE)
        net.set_data(x, t, m)
        hist = net.train(epochs=40, batch=2048, lr=2e-3, lr_final=1e-4, seed=9)
        runs.append((hist, net.get_params()))
        net.close()
    hist = runs[0][0]
    assert np.all(np.isfinite(hist))
    assert hist[-1] < hist[0] / 20, hist[[0, -1]]
    assert np.array_equal(runs[0][1], runs[1][1]), "training is not deterministic"
    assert np.array_equal(runs[0][0], runs[1][0])


def _ctx(n=4):
    prob = problems.natural_convection(n=n, Ra=1e3, carreau=True, jitter=0.15)
    return prob, capi.Context(prob)


def test_guess_mapping_matches_host_restatement():
    print("[case] mlp guess mapping", flush=True)
    prob, ctx = _ctx(4)
    net = capi.Mlp(3, 4, hidden=2, width=64, seed=5, in_shift=SHIFT, in_scale=SCALE)
    scale, shift = [1.5, -2.0, 30.0, 0.25], [0.1, 0.2, -3.0, 0.05]
    q = capi.mlp_query([4.2], scale, shift)
    x = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    net.guess_dev(ctx, q, x.data_ptr())
    got = x.cpu().numpy()
    # host restatement: network at the sample points, affine map, P1 by differences
    pts = learned.sample_points(prob.mesh)
    y = net.predict(np.concatenate([pts, np.full((len(pts), 1), 4.2)], 1).astype(np.float32))
    y = y.astype(np.float64) * np.array(scale) + np.array(shift)
    fu = ctx.first_unknown()
    m = prob.mesh
    exp = np.zeros(ctx.N)
    f = fu[: m.num_nodes]
    exp[f], exp[f + 1], exp[f + 2] = y[: m.num_nodes, 0], y[: m.num_nodes, 1], y[: m.num_nodes, 3]
    pe = y[m.num_nodes:, 2].reshape(m.num_elems, 5)
    cen = fu[m.connect[:, 8]] + 3
    exp[cen], exp[cen + 1], exp[cen + 2] = pe[:, 0], pe[:, 1] - pe[:, 2], pe[:, 3] - pe[:, 4]
    mask, val = ctx.dirichlet()
    exp[mask > 0] = val[mask > 0]
    assert np.max(np.abs(got - exp)) < 1e-5 * np.max(np.abs(exp))
    # delta form: x = x_prev + G(q) - G(q_prev), Dirichlet values kept
    xp = torch.tensor(np.random.default_rng(1).standard_normal(ctx.N), device="cuda")
    q2 = capi.mlp_query([3.9], scale, shift)
    x2 = torch.zeros_like(x)
    net.guess_dev(ctx, q, x2.data_ptr(), q2, xp.data_ptr())
    x3 = torch.zeros_like(x)
    net.guess_dev(ctx, q2, x3.data_ptr())
    d = xp.cpu().numpy() + got - x3.cpu().numpy()
    d[mask > 0] = val[mask > 0]
    assert np.max(np.abs(x2.cpu().numpy() - d)) < 1e-9 * np.max(np.abs(d))
    ctx.close()


def test_fallback_restarts_from_previous_guess():
    """Forced from the host: an unreachable drop factor makes any guess that is
    not converged within 3 passes fall back."""
    print("[case] fallback", flush=True)
    prob, ctx = _ctx(6)
    opts = dict(precond=capi.GM_PC_ILU0, patch_bins=2, line_search=1, max_newton=30)
    x0 = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    _, rep0 = ctx.solve_dev(x0.data_ptr(), **opts)
    assert rep0.converged
    # a poor guess: the converged field scaled and shifted
    bad = x0 * 3.0 + 1.0
    prev = torch.zeros_like(x0)
    x = bad.clone()
    _, rep = ctx.solve_guess_dev(x.data_ptr(), prev.data_ptr(), watch_its=3, drop=1e30, **opts)
    assert rep.fell_back == 1 and rep.discarded_its == 3
    assert rep.converged and rep.newton_its == rep0.newton_its
    assert torch.max(torch.abs(x - x0)).item() < 1e-8 * torch.max(torch.abs(x0)).item()
    # a good guess (the solution itself, slightly perturbed) is kept
    x = x0 + 1e-6 * torch.sin(torch.arange(ctx.N, device="cuda", dtype=torch.float64))
    _, rep = ctx.solve_guess_dev(x.data_ptr(), prev.data_ptr(), watch_its=3, drop=10.0, **opts)
    assert rep.fell_back == 0 and rep.converged and rep.newton_its <= 3
    ctx.close()


def test_learned_guess_beats_previous_step_guess():
    """Small Carreau-Yasuda sweep (16^2, 12 Ra values in [1e3, 6e3]; the
    shear-thinning flow needs finer meshes beyond): train on the sweep,
    evaluate at the 6 held-out geometric midpoints; the learned (delta-form)
    guess starts closer and needs no more Newton passes in total than the
    previous-step guess."""
    print("[case] learned vs previous-step guess", flush=True)
    prob = problems.natural_convection(n=16, Ra=1e3, carreau=True, jitter=0.15)
    ctx = capi.Context(prob)
    opts = dict(precond=capi.GM_PC_ILU0, patch_bins=1, line_search=1, max_newton=30, kspace=50,
                max_its=500)
    ras = problems.sweep_ra(12, 1e3, 6e3)
    steps, sols = continuation.sweep(ctx, prob, ras, opts)
    assert all(s.converged for s in steps), [(s.Ra, s.newton_its, s.lin_its) for s in steps]
    lg = learned.LearnedGuess(prob, ctx, ras, sols, hidden=4, width=256, seed=1)
    hist = lg.train(epochs=300, batch=4096, lr=2e-3, lr_final=2e-5, seed=2)
    assert hist[-1] < hist[0] / 30
    assert lg.set_basis(sols) == len(sols) - 1
    base_its = learn_its = 0
    base_r0, learn_r0 = [], []
    for k in range(0, len(ras) - 1, 2):
        Ra = float(np.sqrt(ras[k] * ras[k + 1]))
        ctx.set_physics(problems.physics_at(prob, Ra))
        xprev = torch.tensor(sols[k], device="cuda")
        x = xprev.clone()
        _, rb = ctx.solve_dev(x.data_ptr(), **opts)
        x = torch.zeros_like(xprev)
        lg.guess(x.data_ptr(), Ra, float(ras[k]), xprev.data_ptr())
        _, rl = ctx.solve_guess_dev(x.data_ptr(), xprev.data_ptr(), 3, 10.0, **opts)
        assert rb.converged and rl.converged
        base_its += rb.newton_its
        learn_its += rl.newton_its + rl.discarded_its
        base_r0.append(rb.resid_L2[0])
        learn_r0.append(rl.resid_L2[0] if not rl.fell_back else rb.resid_L2[0])
        print(f"  Ra={Ra:.4g} previous-step: {rb.newton_its} its r0={rb.resid_L2[0]:.2e}; "
              f"learned: {rl.newton_its}+{rl.discarded_its} its r0={rl.resid_L2[0]:.2e}", flush=True)
    assert np.mean(np.log10(learn_r0)) < np.mean(np.log10(base_r0)) - 0.3
    assert learn_its <= base_its
    lg.close()
    ctx.close()
