"""Pins the 3-D block-search oracle: ZNCC definition on one shift, exact
recovery of integer rigid shifts, window search equals full search when the
window holds the full-search optimum, and exclusion of shifts that leave the
volume."""
import numpy as np
from scipy.ndimage import gaussian_filter


def test_zncc3d_definition(orc):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((24, 24, 24)).astype(np.float32)
    b = rng.standard_normal((24, 24, 24)).astype(np.float32)
    t = a[2:18, 3:19, 4:20].astype(np.float64)
    s = b[1:17, 5:21, 6:22].astype(np.float64)
    t, s = t - t.mean(), s - s.mean()
    want = (t * s).sum() / np.sqrt((t * t).sum() * (s * s).sum())
    # block at (bx, by, bz) = (4, 3, 2); shift (dx, dy, dz) = (2, 2, -1)
    assert abs(orc.zncc3d_at(a, b, 4, 3, 2, 16, 2, 2, -1) - want) < 1e-12
    assert orc.zncc3d_at(a, b, 4, 3, 2, 16, -5, 0, 0) == -np.inf  # leaves the volume


def test_rigid_shift_recovered_and_window(orc):
    rng = np.random.default_rng(1)
    v = gaussian_filter(rng.standard_normal((40, 40, 40)), 1.0).astype(np.float32)
    s = np.roll(v, (2, -1, 4), axis=(0, 1, 2))
    d, p = orc.ncc3d_full(v, s)
    # blocks whose true match stays inside the volume recover it exactly, ZNCC 1
    org = np.array([(x, y, z) for z in range(0, 25, 8) for y in range(0, 25, 8) for x in range(0, 25, 8)])
    inside = np.all((org + [4, -1, 2] >= 0) & (org + [4, -1, 2] + 16 <= 40), axis=1)
    ok = (d == np.array([4, -1, 2])).all(axis=1)
    assert inside.sum() == 27 and np.all(ok[inside]) and np.allclose(p[inside], 1.0, atol=1e-6)
    dw, pw = orc.ncc3d_full(v, s, centre=d, radius=2)
    assert np.array_equal(dw, d) and np.array_equal(pw, p)
