"""Displacement-proposal CNN on the GPU: block extraction, training forward and
gradients against the FP64 oracle, fused inference against the training
path, Adam, a short training run, and learned block matching composed from
the oracle's windowed and full searches.

BF16 operands with FP32 accumulation: outputs and gradients are held to a
relative L2 error of 3e-2 against the FP64 oracle (4 layers of BF16 rounding
of activations, weights and gradients; a wrong index or a dropped term gives
errors of order 1)."""
import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

from cuda_ultrasound_toolkit_amd import api

pytestmark = pytest.mark.gpu

LAYERS = [(2, 16), (16, 32), (32, 32), (32, 32)]


def init_params(seed, n):
    rng = np.random.default_rng(seed)
    p = np.zeros(n, np.float32)
    o = 0
    for ci, co in LAYERS:
        p[o:o + co * ci * 27] = rng.standard_normal(co * ci * 27) * np.sqrt(2.0 / (ci * 27))
        o += co * ci * 27
        p[o:o + co] = rng.standard_normal(co) * 0.05
        o += co
    p[o:o + 768] = rng.standard_normal(768) * np.sqrt(1.0 / 256)
    p[o + 768:] = rng.standard_normal(3) * 0.1
    return p


def blocks_of(p):
    """(name, slice) of every parameter block."""
    out, o = [], 0
    for i, (ci, co) in enumerate(LAYERS):
        out.append((f"w{i + 1}", slice(o, o + co * ci * 27)))
        o += co * ci * 27
        out.append((f"b{i + 1}", slice(o, o + co)))
        o += co
    out.append(("wf", slice(o, o + 768)))
    out.append(("bf", slice(o + 768, o + 771)))
    return out


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _t(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def volume_pair(n, d, seed):
    rng = np.random.default_rng(seed)
    v = gaussian_filter(rng.standard_normal((n, n, n)), 1.2)
    s = np.roll(v, (d[2], d[1], d[0]), axis=(0, 1, 2)) + 0.05 * v.std() * rng.standard_normal(v.shape)
    return np.abs(v).astype(np.float32), np.abs(s).astype(np.float32)


def norm_block(vol, x0, y0, z0, flips):
    b = vol[z0:z0 + 16, y0:y0 + 16, x0:x0 + 16].astype(np.float64)
    if flips & 1:
        b = b[:, :, ::-1]
    if flips & 2:
        b = b[:, ::-1, :]
    if flips & 4:
        b = b[::-1, :, :]
    b = b - b.mean()
    sd = np.sqrt((b * b).mean())
    return b / sd if sd > 0 else b


@pytest.fixture(scope="module")
def pair(torch_dev):
    tpl, src = volume_pair(48, (2, -1, 3), 11)
    return tpl, src, _t(tpl, torch_dev), _t(src, torch_dev)


G = 5  # 48^3 at stride 8: 5 x 5 x 5 blocks at origins 0, 8, ..., 32
NB = G ** 3


def origins(ids):
    ids = np.asarray(ids)
    return np.stack([ids % G, (ids // G) % G, ids // (G * G)], 1) * 8


def gathered(torch_dev, pair, n, seed=0, offsets=False, flips=False):
    import torch
    tpl, src, dt, ds = pair
    rng = np.random.default_rng(seed)
    desc = np.zeros((n, 6), np.int32)
    desc[:, 1] = rng.integers(0, NB, n)
    if offsets:
        desc[:, 2:5] = rng.integers(-3, 4, (n, 3))
        g = origins(desc[:, 1])
        desc[:, 2:5] = np.clip(g + desc[:, 2:5], 0, 32) - g
    if flips:
        desc[:, 5] = rng.integers(0, 8, n)
    vols = torch.tensor([dt.data_ptr(), ds.data_ptr()], dtype=torch.int64, device=torch_dev)
    out = torch.zeros((n, 2, 16, 16, 16), dtype=torch.float32, device=torch_dev)
    assert api.pnet_gather(vols, dt.shape, 8, _t(desc, torch_dev), out)
    torch.cuda.synchronize()
    return desc, out


def pooled_region(vol, x0, y0, z0):
    """2x2x2 means of the 32^3 region [x0 - 8, x0 + 24) per axis, coordinates clamped."""
    n = vol.shape
    zi = np.clip(np.arange(z0 - 8, z0 + 24), 0, n[0] - 1)
    yi = np.clip(np.arange(y0 - 8, y0 + 24), 0, n[1] - 1)
    xi = np.clip(np.arange(x0 - 8, x0 + 24), 0, n[2] - 1)
    r = vol[np.ix_(zi, yi, xi)].astype(np.float64)
    return r.reshape(16, 2, 16, 2, 16, 2).mean(axis=(1, 3, 5))


def normalise(b, flips):
    if flips & 1:
        b = b[:, :, ::-1]
    if flips & 2:
        b = b[:, ::-1, :]
    if flips & 4:
        b = b[::-1, :, :]
    b = b - b.mean()
    sd = np.sqrt((b * b).mean())
    return b / sd if sd > 0 else b


@pytest.fixture(params=[0, 1])
def input_mode(request):
    assert api.pnet_set_input_mode(request.param)
    yield request.param
    assert api.pnet_set_input_mode(1)


def test_gather_normalises_blocks(torch_dev, pair, input_mode):
    tpl, src = pair[0], pair[1]
    desc, out = gathered(torch_dev, pair, 24, seed=1, offsets=True, flips=True)
    got = out.cpu().numpy()
    for i, (_, b, ox, oy, oz, fl) in enumerate(desc):
        x0, y0, z0 = origins([b])[0]
        if input_mode == 0:
            want_t = norm_block(tpl, x0, y0, z0, fl)
            want_s = norm_block(src, x0 + ox, y0 + oy, z0 + oz, fl)
        else:
            want_t = normalise(pooled_region(tpl, x0, y0, z0), fl)
            want_s = normalise(pooled_region(src, x0 + ox, y0 + oy, z0 + oz), fl)
        assert np.abs(got[i, 0] - want_t).max() < 2e-4
        assert np.abs(got[i, 1] - want_s).max() < 2e-4


@pytest.fixture(params=["fp32", "bf16"])
def precision(request):
    assert api.pnet_set_precision(request.param == "fp32")
    yield request.param
    assert api.pnet_set_precision(False)


# (forward, per-block gradient) relative L2 bounds: FP32 operands pin the layer
# structure; BF16 rounding of activations, weights and gradients through four
# layers gives 5-10 % gradient error against FP64 (torch's own BF16 autograd of
# the same network shows the same spread)
TOL = {"fp32": (1e-5, 1e-4), "bf16": (3e-2, 0.15)}


def test_training_forward_matches_oracle(torch_dev, orc, pair, precision):
    import torch
    n = api.pnet_param_count()
    p = init_params(3, n)
    _, x = gathered(torch_dev, pair, 40, seed=2)
    y = torch.zeros((40, 3), dtype=torch.float32, device=torch_dev)
    assert api.pnet_forward(_t(p, torch_dev), x, y)
    torch.cuda.synchronize()
    want = orc.pnet_forward(p.astype(np.float64), x.cpu().numpy())
    assert rel_l2(y.cpu().numpy(), want) < TOL[precision][0]


def test_gradients_match_oracle(torch_dev, orc, pair, precision):
    import torch
    n = api.pnet_param_count()
    p = init_params(4, n)
    _, x = gathered(torch_dev, pair, 32, seed=3, flips=True)
    t = np.random.default_rng(5).uniform(-6, 6, (32, 3)).astype(np.float32)
    g = torch.zeros(n, dtype=torch.float32, device=torch_dev)
    loss = torch.zeros(1, dtype=torch.float32, device=torch_dev)
    assert api.pnet_loss_grad(_t(p, torch_dev), x, _t(t, torch_dev), g, loss)
    torch.cuda.synchronize()
    want_loss, want = orc.pnet_loss_grad(p.astype(np.float64), x.cpu().numpy(), t.astype(np.float64))
    ftol, gtol = TOL[precision]
    assert abs(loss.item() - want_loss) < ftol * 3 * want_loss
    got = g.cpu().numpy().astype(np.float64)
    errs = {name: rel_l2(got[sl], want[sl]) for name, sl in blocks_of(p)}
    print(precision, {k: round(v, 6) for k, v in errs.items()})
    assert max(errs.values()) < gtol, errs
    cos = float(got @ want / (np.linalg.norm(got) * np.linalg.norm(want)))
    assert cos > (1 - 1e-6 if precision == "fp32" else 0.995)
    # deterministic: the same call gives the same bits
    g2 = torch.zeros_like(g)
    assert api.pnet_loss_grad(_t(p, torch_dev), x, _t(t, torch_dev), g2)
    torch.cuda.synchronize()
    assert torch.equal(g, g2)


def test_fused_inference_matches_training_path(torch_dev, orc, pair, input_mode):
    import torch
    n = api.pnet_param_count()
    p = init_params(6, n)
    dp = _t(p, torch_dev)
    tpl, src, dt, ds = pair
    vols = torch.tensor([dt.data_ptr(), ds.data_ptr()], dtype=torch.int64, device=torch_dev)
    desc = np.zeros((NB, 6), np.int32)
    desc[:, 1] = np.arange(NB)
    x = torch.zeros((NB, 2, 16, 16, 16), dtype=torch.float32, device=torch_dev)
    assert api.pnet_gather(vols, dt.shape, 8, _t(desc, torch_dev), x)
    y_train = torch.zeros((NB, 3), dtype=torch.float32, device=torch_dev)
    assert api.pnet_forward(dp, x, y_train)
    y_fused = torch.zeros((NB, 3), dtype=torch.float32, device=torch_dev)
    centre = torch.zeros((NB, 3), dtype=torch.int32, device=torch_dev)
    assert api.pnet_infer(dp, dt, ds, 8, y_fused, centre)
    # a block list: every third block
    ids = np.arange(0, NB, 3, dtype=np.int32)
    y_ids = torch.zeros((len(ids), 3), dtype=torch.float32, device=torch_dev)
    assert api.pnet_infer(dp, dt, ds, 8, y_ids, d_block_ids=_t(ids, torch_dev))
    torch.cuda.synchronize()
    yt, yf = y_train.cpu().numpy(), y_fused.cpu().numpy()
    assert rel_l2(yf, yt) < 5e-3
    assert rel_l2(yf, orc.pnet_forward(p.astype(np.float64), x.cpu().numpy())) < 3e-2
    assert np.array_equal(centre.cpu().numpy(), np.clip(np.rint(yf), -8, 8).astype(np.int32))
    assert np.array_equal(y_ids.cpu().numpy(), yf[ids])


def test_adam_step(torch_dev):
    import torch
    rng = np.random.default_rng(7)
    n = 1000
    p, g, m, v = (rng.standard_normal(n).astype(np.float32) for _ in range(4))
    v = np.abs(v)
    dp, dg, dm, dv = (_t(a, torch_dev) for a in (p, g, m, v))
    assert api.pnet_adam(dp, dg, dm, dv, lr=1e-2, step=3, wd=1e-3)
    torch.cuda.synchronize()
    b1, b2, eps = 0.9, 0.999, 1e-8
    m1 = b1 * m + (1 - b1) * g
    v1 = b2 * v + (1 - b2) * g * g
    p1 = p - 1e-2 * ((m1 / (1 - b1 ** 3)) / (np.sqrt(v1 / (1 - b2 ** 3)) + eps) + 1e-3 * p)
    assert np.allclose(dm.cpu().numpy(), m1, rtol=1e-6, atol=1e-7)
    assert np.allclose(dv.cpu().numpy(), v1, rtol=1e-6, atol=1e-7)
    assert np.allclose(dp.cpu().numpy(), p1, rtol=1e-5, atol=1e-6)


def test_training_reduces_loss(torch_dev, pair):
    """A student network fits a teacher's outputs with the library's own
    forward, backward and Adam kernels."""
    import torch
    n = api.pnet_param_count()
    _, x = gathered(torch_dev, pair, 128, seed=9, offsets=True, flips=True)
    teacher = _t(init_params(10, n), torch_dev)
    t = torch.zeros((128, 3), dtype=torch.float32, device=torch_dev)
    assert api.pnet_forward(teacher, x, t)
    p = _t(init_params(11, n), torch_dev)
    g = torch.zeros_like(p)
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    loss = torch.zeros(1, dtype=torch.float32, device=torch_dev)
    losses = []
    for step in range(1, 301):
        assert api.pnet_loss_grad(p, x, t, g, loss)
        assert api.pnet_adam(p, g, m, v, lr=2e-3, step=step)
        if step == 1 or step % 100 == 0:
            torch.cuda.synchronize()
            losses.append(loss.item())
    print("losses", losses)
    assert losses[-1] < 0.1 * losses[0]


@pytest.mark.parametrize("bias,radius", [((2.0, -1.0, 3.0), 2), ((-7.0, 7.0, -7.0), 2), ((1.6, -0.6, 2.4), 1)])
def test_learned_matching_composes_oracle(torch_dev, orc, pair, bias, radius):
    """A network whose only nonzero parameters are the output biases proposes
    the same displacement for every block; the learned search must equal the
    oracle's windowed search around it, with the oracle's full search for the
    blocks that fail the peak or window-face tests."""
    import torch
    tpl, src, dt, ds = pair
    n = api.pnet_param_count()
    p = np.zeros(n, np.float32)
    p[-3:] = bias
    nb = NB
    disp = torch.zeros((nb, 3), dtype=torch.int32, device=torch_dev)
    peak = torch.zeros(nb, dtype=torch.float32, device=torch_dev)
    path = torch.zeros(nb, dtype=torch.int32, device=torch_dev)
    pred = torch.zeros((nb, 3), dtype=torch.float32, device=torch_dev)
    thr = 0.6
    ok, nfb = api.motion3d_learned(_t(p, torch_dev), dt, ds, 8, disp, peak, radius=radius, threshold=thr,
                                   d_path=path, d_pred=pred, count_fallbacks=True)
    assert ok
    pr = pred.cpu().numpy()
    assert np.allclose(pr, np.array(bias, np.float32)[None, :], atol=1e-6)
    centre = np.clip(np.rint(pr), -8, 8).astype(np.int32)
    wd, wp = orc.ncc3d_full(tpl, src, centre=centre, radius=radius)
    fd, fp = orc.ncc3d_full(tpl, src)
    need = ~(wp >= thr)
    origin = origins(np.arange(nb))
    lo = np.maximum(-8, -origin)
    hi = np.minimum(8, 48 - 16 - origin)
    face = ((wd == centre - radius) & (wd > lo)) | ((wd == centre + radius) & (wd < hi))
    need |= face.any(axis=1)
    clear = np.abs(wp - thr) > 1e-4  # blocks whose peak test cannot flip on rounding
    got_path = path.cpu().numpy().astype(bool)
    assert np.array_equal(got_path[clear], need[clear])
    assert nfb == int(got_path.sum())
    want = np.where(got_path[:, None], fd, wd)
    assert np.array_equal(disp.cpu().numpy(), want)
    wantp = np.where(got_path, fp, wp)
    assert np.abs(peak.cpu().numpy() - wantp).max() < 2e-5
    print(f"bias {bias} r {radius}: {nfb}/{nb} full searches")
