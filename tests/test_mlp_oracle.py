"""Pins of the MLP oracle (oracle/mlp_oracle.py) on the CPU: the gradient against central
finite differences of the unrounded FP64 network, the BF16 rounding against torch's
bfloat16 conversion, GELU against its closed form, and Adam against torch.optim.Adam."""
import math

import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo


def test_layout():
    assert mo.NPARAMS == 24864
    p = np.arange(mo.NPARAMS, dtype=np.float64)
    d = mo.unpack(p)
    assert d["W1"].shape == (128, 32) and d["b3"].shape == (32,)
    assert d["W1"][0, 1] == 1 and d["b1"][0] == 4096 and d["W2"][0, 0] == 4224
    assert d["b2"][0] == 20608 and d["W3"][0, 0] == 20736 and d["b3"][0] == 24832
    assert np.array_equal(mo.pack(d), p)


def test_bf16_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-6, 6, 20000),
                        [0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, 65504.0]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16(x), ref)


def test_gelu_closed_form():
    z = np.linspace(-6, 6, 101)
    ref = np.array([0.5 * v * (1 + math.erf(v / math.sqrt(2))) for v in z])
    assert np.allclose(mo.gelu(z), ref, rtol=0, atol=1e-15)
    # derivative against central differences
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.allclose(mo.gelu_grad(z), fd, rtol=1e-7, atol=1e-8)


def test_gradient_finite_differences():
    """Backward of the unrounded network against central differences: a dropped term, a
    transposed weight or a wrong GELU derivative fails this."""
    rng = np.random.default_rng(11)
    p = mo.init_params(5).astype(np.float64) + 0.01 * rng.standard_normal(mo.NPARAMS)
    x = rng.standard_normal((6, 32))
    t = rng.standard_normal((6, 32))
    loss, g = mo.loss_and_gradient(p, x, t, rounding=False)
    idx = rng.choice(mo.NPARAMS, 60, replace=False)
    # include every block and every bias
    idx = np.concatenate([idx, [0, 4096, 4224, 20608, 20736, 24832, 24863]])
    h = 1e-6
    for i in idx:
        e = np.zeros(mo.NPARAMS)
        e[i] = h
        lp, _ = mo.loss_and_gradient(p + e, x, t, rounding=False)
        lm, _ = mo.loss_and_gradient(p - e, x, t, rounding=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_rounded_forward_close_to_exact():
    rng = np.random.default_rng(1)
    p = mo.init_params(2)
    x = rng.standard_normal((50, 32))
    y1, _ = mo.forward(p, x, rounding=True)
    y0, _ = mo.forward(p, x, rounding=False)
    rel = np.linalg.norm(y1 - y0) / np.linalg.norm(y0)
    assert 1e-4 < rel < 2e-2  # BF16 operands: visible but small


def test_adam_matches_torch():
    rng = np.random.default_rng(4)
    p0 = rng.standard_normal(40)
    grads = [rng.standard_normal(40) for _ in range(5)]
    opt = mo.Adam(p0)
    tp = torch.nn.Parameter(torch.from_numpy(p0.copy()))
    topt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for g in grads:
        opt.update(g, 1e-2)
        tp.grad = torch.from_numpy(g.copy())
        topt.step()
    assert np.allclose(opt.p, tp.detach().numpy(), rtol=1e-12, atol=1e-14)
