"""3-D ZNCC block search on the GPU vs the FP64 oracle (full and windowed)."""
import numpy as np
import pytest
from scipy.ndimage import gaussian_filter, shift as nd_shift

from cuda_ultrasound_toolkit_amd import api

pytestmark = pytest.mark.gpu


def volume_pair(n, d, seed, noise=0.05, sub=(0.0, 0.0, 0.0)):
    rng = np.random.default_rng(seed)
    v = gaussian_filter(rng.standard_normal((n, n, n)), 1.2)
    s = np.roll(v, (d[2], d[1], d[0]), axis=(0, 1, 2))
    if any(sub):
        s = nd_shift(s, (sub[2], sub[1], sub[0]), order=3, mode="wrap")
    s = s + noise * v.std() * rng.standard_normal(v.shape)
    return np.abs(v).astype(np.float32), np.abs(s).astype(np.float32)


def _t(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


@pytest.mark.parametrize("n,d,stride", [(48, (3, -2, 5), 8), (40, (-6, 6, 0), 8), (64, (1, 1, -7), 16), (32, (0, 0, 0), 8)])
def test_full_search_matches_oracle(torch_dev, orc, n, d, stride):
    import torch
    print(f"ncc3d full n={n} d={d}", flush=True)
    tpl, src = volume_pair(n, d, n)
    want_d, want_p = orc.ncc3d_full(tpl, src, stride=stride)
    nb = len(want_p)
    disp = torch.zeros((nb, 3), dtype=torch.int32, device=torch_dev)
    peak = torch.zeros(nb, dtype=torch.float32, device=torch_dev)
    assert api.ncc3d_search(_t(tpl, torch_dev), _t(src, torch_dev), stride, disp, peak)
    torch.cuda.synchronize()
    got_d, got_p = disp.cpu().numpy(), peak.cpu().numpy()
    assert np.array_equal(got_d, want_d)
    assert np.abs(got_p - want_p).max() < 2e-5


def test_window_and_block_list_match_oracle(torch_dev, orc):
    import torch
    n = 48
    tpl, src = volume_pair(n, (2, -3, 1), 7, sub=(0.3, -0.4, 0.2))
    full_d, full_p = orc.ncc3d_full(tpl, src)
    rng = np.random.default_rng(1)
    centre = np.clip(full_d + rng.integers(-2, 3, size=full_d.shape), -8, 8).astype(np.int32)
    want_d, want_p = orc.ncc3d_full(tpl, src, centre=centre, radius=2)
    nb = len(full_p)
    disp = torch.zeros((nb, 3), dtype=torch.int32, device=torch_dev)
    peak = torch.zeros(nb, dtype=torch.float32, device=torch_dev)
    assert api.ncc3d_search(_t(tpl, torch_dev), _t(src, torch_dev), 8, disp, peak, d_centre=_t(centre, torch_dev),
                            radius=2)
    torch.cuda.synchronize()
    assert np.array_equal(disp.cpu().numpy(), want_d)
    assert np.abs(peak.cpu().numpy() - want_p).max() < 2e-5
    # full search on a subset of blocks
    ids = np.array([0, 5, 17, nb - 1], np.int32)
    d2 = torch.zeros((4, 3), dtype=torch.int32, device=torch_dev)
    p2 = torch.zeros(4, dtype=torch.float32, device=torch_dev)
    assert api.ncc3d_search(_t(tpl, torch_dev), _t(src, torch_dev), 8, d2, p2, d_block_ids=_t(ids, torch_dev),
                            n_blocks=4)
    torch.cuda.synchronize()
    assert np.array_equal(d2.cpu().numpy(), full_d[ids])


def test_envelope(torch_dev):
    import torch
    z = (np.random.default_rng(0).standard_normal(1000) + 1j * np.random.default_rng(1).standard_normal(1000)).astype(np.complex64)
    out = torch.zeros(1000, dtype=torch.float32, device=torch_dev)
    assert api.envelope(_t(z.view(np.float32), torch_dev), out)
    torch.cuda.synchronize()
    assert np.allclose(out.cpu().numpy(), np.abs(z), rtol=1e-6)
