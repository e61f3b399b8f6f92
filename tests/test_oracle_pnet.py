"""FP64 oracle of the displacement-proposal CNN: forward against torch's own
float64 conv3d/linear, gradients against torch autograd and central
finite differences."""
import numpy as np
import pytest

LAYERS = [(2, 16, 2), (16, 32, 2), (32, 32, 1), (32, 32, 2)]


def unpack(p):
    import torch
    out, o = [], 0
    for ci, co, _ in LAYERS:
        w = p[o:o + co * ci * 27].reshape(co, ci, 3, 3, 3)
        o += co * ci * 27
        b = p[o:o + co]
        o += co
        out.append((w, b))
    wf = p[o:o + 768].reshape(3, 256)
    bf = p[o + 768:o + 771]
    assert o + 771 == len(p)
    return out, wf, bf


def torch_forward(p, x):
    import torch
    import torch.nn.functional as F
    convs, wf, bf = unpack(p)
    h = x
    for (w, b), (_, _, s) in zip(convs, LAYERS):
        h = F.relu(F.conv3d(h, w, b, stride=s, padding=1))
    return F.linear(h.reshape(h.shape[0], -1), wf, bf)


@pytest.fixture(scope="module")
def case(orc):
    rng = np.random.default_rng(5)
    n = orc.pnet_param_count()
    p = rng.standard_normal(n) * 0.08
    x = rng.standard_normal((4, 2, 16, 16, 16))
    t = rng.standard_normal((4, 3)) * 3
    return p, x, t


def test_param_count(orc):
    assert orc.pnet_param_count() == sum(co * ci * 27 + co for ci, co, _ in LAYERS) + 771


def test_forward_matches_torch(orc, case):
    import torch
    p, x, _ = case
    want = torch_forward(torch.from_numpy(p), torch.from_numpy(x)).numpy()
    got = orc.pnet_forward(p, x)
    assert np.abs(got - want).max() < 1e-10 * max(1.0, np.abs(want).max())


def test_gradient_matches_autograd(orc, case):
    import torch
    p, x, t = case
    pt = torch.from_numpy(p.copy()).requires_grad_(True)
    y = torch_forward(pt, torch.from_numpy(x))
    loss = 0.5 * ((y - torch.from_numpy(t)) ** 2).sum() / x.shape[0]
    loss.backward()
    got_loss, got = orc.pnet_loss_grad(p, x, t)
    assert abs(got_loss - loss.item()) < 1e-10 * abs(loss.item())
    want = pt.grad.numpy()
    assert np.abs(got - want).max() < 1e-9 * np.abs(want).max()


def test_gradient_finite_differences(orc, case):
    p, x, t = case
    _, g = orc.pnet_loss_grad(p, x, t)
    rng = np.random.default_rng(9)
    n = len(p)
    # one coordinate from every parameter block, plus random ones
    idx = [0, 864, 880, 14704, 14740, 42412, 42450, 70060, 70100, n - 1] + list(rng.integers(0, n, 10))
    for i in idx:
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        fd = (orc.pnet_loss_grad(pp, x, t)[0] - orc.pnet_loss_grad(pm, x, t)[0]) / (2 * e)
        assert abs(g[i] - fd) <= 1e-6 * max(1.0, abs(fd)), (i, g[i], fd)
