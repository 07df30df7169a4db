"""Pins of the learned-guess network oracle (oracle/gm_mlp_oracle.c).

The oracle is checked against independent routines: torch's own bfloat16
conversion (round to nearest even), a torch float64 MLP and its autograd
gradients, and central finite differences of the loss.
"""
import numpy as np
import pytest
import torch

from oracle import MlpOracle, bf16, lib

SHAPE = dict(n_in=3, n_out=4, hidden=3, width=16)


def _params(o, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(o.num_params) * 0.4 * scale).astype(np.float32)


def _data(n, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    x[:, 2] = rng.uniform(3, 6, n)  # a log10 Ra-like input, normalised by the network
    t = rng.standard_normal((n, 4)).astype(np.float32)
    m = (rng.uniform(size=(n, 4)) > 0.3).astype(np.float32)
    return x, t, m


def _torch_layers(o, p):
    """Split the flat parameter vector into torch float64 (W, b) pairs."""
    out, off, k = [], 0, o.n_in
    for l in range(o.shape.hidden + 1):
        n = o.n_out if l == o.shape.hidden else o.shape.width
        W = torch.tensor(p[off:off + n * k].astype(np.float64).reshape(n, k), requires_grad=True)
        off += n * k
        b = torch.tensor(p[off:off + n].astype(np.float64), requires_grad=True)
        off += n
        out.append((W, b))
        k = n
    return out


def _torch_forward(layers, u):
    a = u
    for i, (W, b) in enumerate(layers):
        z = a @ W.T + b
        a = torch.tanh(z) if i < len(layers) - 1 else z
    return a


def _oracle():
    return MlpOracle(**SHAPE, in_shift=[0.0, 0.0, 4.5], in_scale=[1.0, 1.0, 1 / 1.5])


def test_num_params():
    o = _oracle()
    assert o.num_params == (3 * 16 + 16) + 2 * (16 * 16 + 16) + (16 * 4 + 4)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.standard_normal(4000).astype(np.float32) * 10 ** rng.uniform(-8, 8, 4000),
                        np.float32([0.0, -0.0, 1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, -(1 + 2 ** -8),
                                    3.4e38, 1e-40])]).astype(np.float32)
    ref = torch.tensor(v).to(torch.bfloat16).to(torch.float32).numpy()
    ours = np.array([lib().omlp_bf16(float(a)) for a in v], np.float32)
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(bf16(v).view(np.uint32), ref.view(np.uint32))
    # ties go to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert lib().omlp_bf16(1 + 2 ** -8) == 1.0


def test_forward_matches_torch_fp64():
    o = _oracle()
    p = _params(o)
    x, _, _ = _data(50)
    y = o.forward(p, x, quant=False)
    u = (torch.tensor(x, dtype=torch.float64) - torch.tensor(o.in_shift)) * torch.tensor(o.in_scale)
    ref = _torch_forward(_torch_layers(o, p), u).detach().numpy()
    assert np.max(np.abs(y - ref)) < 1e-12


def test_forward_bf16_rounding_points():
    """quant=1 rounds the normalised input, the weights and each hidden
    activation to bf16; biases and the output stay wide."""
    o = _oracle()
    p = _params(o)
    x, _, _ = _data(40)
    y = o.forward(p, x, quant=True)
    q = lambda a: torch.tensor(bf16(np.asarray(a, np.float32)), dtype=torch.float64)
    u = q((x - o.in_shift.astype(np.float32)) * o.in_scale.astype(np.float32))
    layers = _torch_layers(o, p)
    a = u
    for i, (W, b) in enumerate(layers):
        z = a @ q(W.detach().numpy()).T + b.detach()
        a = q(torch.tanh(z).numpy()) if i < len(layers) - 1 else z
    assert np.max(np.abs(y - a.numpy())) < 1e-9
    # and the rounding matters: the unrounded network differs visibly
    assert np.max(np.abs(y - o.forward(p, x, quant=False))) > 1e-4


def test_gradient_matches_autograd_fp64():
    o = _oracle()
    p = _params(o)
    x, t, m = _data(37)
    loss, g = o.loss_grad(p, x, t, m, quant=False)
    layers = _torch_layers(o, p)
    u = (torch.tensor(x, dtype=torch.float64) - torch.tensor(o.in_shift)) * torch.tensor(o.in_scale)
    y = _torch_forward(layers, u)
    L = (torch.tensor(m, dtype=torch.float64) * (y - torch.tensor(t, dtype=torch.float64)) ** 2).sum() \
        / (x.shape[0] * o.n_out)
    L.backward()
    ref = np.concatenate([np.concatenate([W.grad.numpy().ravel(), b.grad.numpy()]) for W, b in layers])
    assert abs(loss - L.item()) < 1e-12 * max(1.0, abs(L.item()))
    assert np.max(np.abs(g - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_gradient_matches_finite_differences():
    o = _oracle()
    p = _params(o).astype(np.float64)
    x, t, m = _data(20)
    _, g = o.loss_grad(p.astype(np.float32), x, t, m, quant=False)
    rng = np.random.default_rng(5)
    idx = rng.choice(o.num_params, 25, replace=False)
    # the oracle takes float32 parameters: perturb by exactly representable steps
    p32 = p.astype(np.float32)
    for i in idx:
        h = np.float32(2.0 ** -10)
        pp, pm = p32.copy(), p32.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = o.loss_grad(pp, x, t, m, quant=False)
        lm, _ = o.loss_grad(pm, x, t, m, quant=False)
        fd = (lp - lm) / (float(pp[i]) - float(pm[i]))
        assert abs(fd - g[i]) < 1e-5 * max(1e-3, abs(g[i])) + 1e-8, (i, fd, g[i])


def test_planted_sign_error_is_caught():
    """A flipped tanh' sign in the backward pass must fail the FD check."""
    o = _oracle()
    p = _params(o)
    x, t, m = _data(20)
    _, g = o.loss_grad(p, x, t, m, quant=False)
    # planted mistake: gradients of the first layer with the wrong sign
    n0 = o.n_in * o.shape.width + o.shape.width
    bad = g.copy()
    bad[:n0] = -bad[:n0]
    h = np.float32(2.0 ** -10)
    i = int(np.argmax(np.abs(g[:n0])))
    pp, pm = p.copy(), p.copy()
    pp[i] += h
    pm[i] -= h
    fd = (o.loss_grad(pp, x, t, m, quant=False)[0] - o.loss_grad(pm, x, t, m, quant=False)[0]) / (2 * h)
    assert abs(fd - g[i]) < 1e-5 * abs(g[i])
    assert abs(fd - bad[i]) > 0.5 * abs(g[i])


def test_quantised_gradient_close_to_exact():
    """BF16 rounding perturbs the gradient by a few tenths of a percent, not more."""
    o = _oracle()
    p = _params(o)
    x, t, m = _data(200)
    _, g0 = o.loss_grad(p, x, t, m, quant=False)
    _, g1 = o.loss_grad(p, x, t, m, quant=True)
    rel = np.linalg.norm(g1 - g0) / np.linalg.norm(g0)
    assert 1e-5 < rel < 3e-2, rel
