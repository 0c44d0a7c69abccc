"""GPU parity of the learned column correction's MLP kernels (forward on BF16 MFMA, FP32
backward, Adam) against the CPU oracle, through the C ABI."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu

# BF16 layer inputs: a value within one FP32 ulp of a BF16 rounding boundary can round the
# other way on the GPU (FP32 sums in a different order than the oracle's FP64), moving one
# activation by 2^-8 relative; stated relative-L2 tolerances cover that.
TOL_Y = 2e-3
TOL_G = 5e-3


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def net(sw, gpu):
    mlp = sw.mlp
    p = mo.init_params(7)
    m = mlp.ColumnMLP(p, max_batch=300, device=gpu)
    yield mlp, m, p
    m.close()


def data(B, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((B, 32)).astype(np.float32), rng.standard_normal((B, 32)).astype(np.float32)


def test_nparams(net):
    mlp, m, p = net
    assert mlp.nparams() == mo.NPARAMS
    assert np.array_equal(m.params().numpy(), p)


@pytest.mark.parametrize("B", [1, 63, 64, 200, 300])
def test_forward(net, B):
    mlp, m, p = net
    x, _ = data(B, B)
    y = m.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    ref, _ = mo.forward(p, x, rounding=True)
    assert rel(y, ref) < TOL_Y
    # against the unrounded FP64 network: BF16-level agreement only
    exact, _ = mo.forward(p, x, rounding=False)
    assert rel(y, exact) < 3e-2


@pytest.mark.parametrize("B", [37, 200])
def test_gradient(net, B):
    mlp, m, p = net
    x, t = data(B, 100 + B)
    loss, g = m.gradient(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    loss = float(loss.item())
    g = g.cpu().numpy()
    lref, gref = mo.loss_and_gradient(p, x, t, rounding=True)
    assert abs(loss - lref) <= 1e-4 * lref
    blocks = mo.unpack(g.astype(np.float64))
    rblocks = mo.unpack(gref)
    for name in blocks:
        assert rel(blocks[name], rblocks[name]) < TOL_G, name


def test_gradient_deterministic(net):
    mlp, m, p = net
    x, t = data(200, 9)
    xs, ts = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    _, g1 = m.gradient(xs, ts)
    _, g2 = m.gradient(xs, ts)
    assert torch.equal(g1, g2)


def test_train_steps_match_oracle(sw, gpu):
    mlp = sw.mlp
    p = mo.init_params(3)
    m = mlp.ColumnMLP(p, max_batch=128, device=gpu)
    opt = mo.Adam(p)
    x, t = data(128, 42)
    xs, ts = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    losses = []
    for _ in range(5):
        losses.append(float(m.train_step(xs, ts, 1e-3).item()))
        _, g = mo.loss_and_gradient(opt.p.astype(np.float32), x, t, rounding=True)
        opt.update(g, 1e-3)
    got = m.params().numpy()
    assert rel(got, opt.p) < 1e-4
    assert losses[-1] < losses[0]
    m.close()


def test_bad_arguments(net):
    mlp, m, p = net
    x, t = data(301, 1)
    with pytest.raises(ValueError):
        m.gradient(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())  # above max_batch
    with pytest.raises(ValueError):
        m.forward(torch.from_numpy(x))  # host tensor
