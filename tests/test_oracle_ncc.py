"""Pins the block-matching oracle (CPU only): ZNCC against its textbook
definition and invariances, the 3x3 least-squares fit table against the
pseudo-inverse of its design matrix, sub-pixel vertex recovery on exact
paraboloids, and whole-pipeline motion recovery on shifted speckle."""
import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

from cuda_ultrasound_toolkit_amd import api


def zncc_numpy(src, tpl):
    th, tw = tpl.shape
    out = np.zeros((src.shape[0] - th + 1, src.shape[1] - tw + 1))
    t = tpl - tpl.mean()
    for v in range(out.shape[0]):
        for u in range(out.shape[1]):
            w = src[v:v + th, u:u + tw]
            w = w - w.mean()
            d = np.sqrt((t * t).sum() * (w * w).sum())
            out[v, u] = (t * w).sum() / d if d > 0 else 0.0
    return out


def speckle(shape, seed, sigma=1.5):
    rng = np.random.default_rng(seed)
    return gaussian_filter(rng.standard_normal(shape), sigma)


def test_zncc_definition_and_invariance(orc):
    rng = np.random.default_rng(0)
    src = rng.standard_normal((12, 15))
    tpl = rng.standard_normal((4, 6))
    m = orc.zncc_valid(src, tpl)
    assert np.allclose(m, zncc_numpy(src, tpl), atol=1e-12)
    # affine changes of either image do not change the map
    assert np.allclose(orc.zncc_valid(3 * src + 7, tpl), m, atol=1e-12)
    assert np.allclose(orc.zncc_valid(src, -0.5 * tpl + 2) , -m, atol=1e-12)
    # a copy of a window correlates to exactly 1 at its own offset
    tpl2 = src[5:9, 3:9]
    m2 = orc.zncc_valid(src, tpl2)
    assert abs(m2[5, 3] - 1.0) < 1e-12 and np.argmax(m2) == 5 * m2.shape[1] + 3
    # a flat window gives 0 (zero denominator)
    flat = np.ones((10, 10))
    assert np.all(orc.zncc_valid(flat, tpl[:3, :3]) == 0.0)


def test_p3_is_least_squares_pseudo_inverse():
    xs = [(x, y) for y in (-1, 0, 1) for x in (-1, 0, 1)]
    A = np.array([[x * x, y * y, x * y, x, y, 1.0] for x, y in xs])
    P = np.linalg.pinv(A)
    P3 = np.array([
        [1 / 6, -1 / 3, 1 / 6] * 3,
        [1 / 6] * 3 + [-1 / 3] * 3 + [1 / 6] * 3,
        [0.25, 0, -0.25, 0, 0, 0, -0.25, 0, 0.25],
        [-1 / 6, 0, 1 / 6] * 3,
        [-1 / 6] * 3 + [0] * 3 + [1 / 6] * 3,
        [-1 / 9, 2 / 9, -1 / 9, 2 / 9, 5 / 9, 2 / 9, -1 / 9, 2 / 9, -1 / 9],
    ])
    assert np.allclose(P, P3, atol=1e-12)


def _motion(orc, frames, grid, spacing, patch, margins, **kw):
    F, H, W = frames.shape
    p = api.motion_params(W, H, F, grid, spacing, patch, margins, **kw)
    return orc.motion_detection(frames.astype(np.float32), p), p


def test_motion_recovers_integer_shifts(orc):
    base = speckle((80, 96), 1)
    shifts = [(0, 0), (2, -1), (-3, 2)]
    frames = np.stack([np.roll(base, (dy, dx), axis=(0, 1)) for dx, dy in shifts])
    maps, p = _motion(orc, frames, (6, 5), 12, 16, (5, 5), abs_thr=0.5, rel_thr=0.0, subpixel=False)
    assert np.all(maps[0] == 0)  # reference frame
    for f, (dx, dy) in enumerate(shifts[1:], start=1):
        inner = maps[f, 1:-1, 1:-1]  # grid points whose template and search window are fully inside
        assert np.all(inner[..., 0] == dx) and np.all(inner[..., 1] == dy), f
        assert np.all(inner[..., 2] > 0.999)


def test_motion_subpixel_on_paraboloid(orc):
    """A correlation surface that is locally an exact paraboloid gives its vertex."""
    yy, xx = np.mgrid[0:60, 0:60].astype(float)
    x0, y0 = 30.3, 29.6
    img = np.exp(-((xx - x0) ** 2 + (yy - y0) ** 2) / 40.0)
    frames = np.stack([img, np.roll(img, (1, 2), axis=(0, 1))])
    maps, _ = _motion(orc, frames, (1, 1), 1, 21, (4, 4), abs_thr=0.0, rel_thr=0.0)
    # one template centred at (0, 0) is clipped; use a grid point in the middle instead
    maps, _ = _motion(orc, frames, (4, 4), 10, 21, (4, 4), abs_thr=0.0, rel_thr=0.0)
    v = maps[1, 3, 3]
    assert abs(v[0] - 2.0) < 0.15 and abs(v[1] - 1.0) < 0.15
    maps0, _ = _motion(orc, frames, (4, 4), 10, 21, (4, 4), abs_thr=0.0, rel_thr=0.0, subpixel=False)
    assert maps0[1, 3, 3, 0] == 2.0 and maps0[1, 3, 3, 1] == 1.0


def test_motion_thresholds_reject_no_match(orc):
    """Uncorrelated frames: no peak beats the absolute threshold -> zero offset."""
    frames = np.stack([speckle((64, 64), 2), speckle((64, 64), 3)])
    maps, _ = _motion(orc, frames, (4, 4), 14, 12, (4, 4), abs_thr=0.9, rel_thr=0.0)
    assert np.all(maps[1, ..., 0] == 0) and np.all(maps[1, ..., 1] == 0)


def test_motion_neighbour_compare(orc):
    base = speckle((48, 48), 4)
    frames = np.stack([np.roll(base, (0, k), axis=(0, 1)) for k in range(4)])
    maps, _ = _motion(orc, frames, (3, 3), 12, 12, (3, 3), abs_thr=0.5, rel_thr=0.0, subpixel=False,
                      neighbour=True)
    assert np.all(maps[1:, 1, 1, 0] == 1.0)  # each frame moved 1 px from its neighbour
