"""Block matching on the GPU vs the FP64 oracle, through the reference's C++
entry points (cuda_toolkit::motion_detection / corr_images)."""
import ctypes as C

import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

from cuda_ultrasound_toolkit_amd import _lib as L, api

pytestmark = pytest.mark.gpu


def speckle(shape, seed, sigma=1.5):
    rng = np.random.default_rng(seed)
    return gaussian_filter(rng.standard_normal(shape), sigma).astype(np.float32)


def frames_with_shifts(shape, shifts, seed, noise=0.0):
    base = speckle(shape, seed)
    rng = np.random.default_rng(seed + 100)
    out = []
    for dx, dy in shifts:
        f = np.roll(base, (dy, dx), axis=(0, 1)) + noise * rng.standard_normal(shape).astype(np.float32)
        out.append(f)
    return np.ascontiguousarray(np.stack(out).astype(np.float32))


def compare(got, want):
    # the candidates' order is decided on float maps on both sides; offsets agree
    # to float rounding of the paraboloid vertex, peaks to float ZNCC rounding
    assert np.allclose(got[..., 2:], want[..., 2:], atol=2e-5), np.abs(got[..., 2:] - want[..., 2:]).max()
    assert np.allclose(got[..., :2], want[..., :2], atol=2e-3), np.abs(got[..., :2] - want[..., :2]).max()


@pytest.mark.parametrize("case", ["basic", "edges", "nosub", "neighbour", "rect", "noisy", "wide", "huge",
                                  "very_wide"])
def test_motion_detection_matches_oracle(lib, orc, case):
    print(f"motion {case}", flush=True)
    if case == "basic":
        imgs = frames_with_shifts((96, 128), [(0, 0), (3, -2), (-1, 4)], 1)
        p = api.motion_params(128, 96, 3, (8, 6), 14, 16, (6, 6), abs_thr=0.3, rel_thr=0.02)
    elif case == "edges":  # grid points whose templates are clipped by the image border
        imgs = frames_with_shifts((64, 80), [(1, 1), (0, 0)], 2)
        p = api.motion_params(80, 64, 2, (9, 7), 10, 12, (5, 4), abs_thr=0.2, rel_thr=0.0, ref=1)
    elif case == "nosub":
        imgs = frames_with_shifts((80, 80), [(0, 0), (2, 2)], 3)
        p = api.motion_params(80, 80, 2, (5, 5), 15, 20, (4, 4), abs_thr=0.3, rel_thr=0.05, subpixel=False)
    elif case == "neighbour":
        imgs = frames_with_shifts((64, 64), [(0, 0), (1, 0), (2, 1), (3, 1)], 4)
        p = api.motion_params(64, 64, 4, (4, 4), 14, 12, (3, 3), abs_thr=0.3, rel_thr=0.0, ref=0, neighbour=True)
    elif case == "rect":  # different x / y margins, large search
        imgs = frames_with_shifts((100, 90), [(0, 0), (-7, 3)], 5)
        p = api.motion_params(90, 100, 2, (5, 6), 16, 18, (9, 5), abs_thr=0.3, rel_thr=0.0)
    elif case == "wide":
        # (2*24+1)^2 = 2401 offsets: the peak search reads the first 2048 map
        # entries only (block_match_kernels.cu:7-64); the third frame's shift
        # (dy = +19) lies in the truncated rows
        imgs = frames_with_shifts((128, 136), [(0, 0), (3, -2), (5, 19)], 9)
        p = api.motion_params(136, 128, 3, (4, 4), 24, 16, (24, 24), abs_thr=0.3, rel_thr=0.0)
    elif case == "huge":
        # 97^2 = 9409 offsets: more than the LDS map prefix; the no-shift entry
        # (index 4656) lies beyond the entries the search keeps in LDS
        imgs = frames_with_shifts((176, 180), [(0, 0), (4, 3), (-6, 30)], 10)
        p = api.motion_params(180, 176, 3, (3, 3), 40, 16, (48, 48), abs_thr=0.3, rel_thr=0.0)
    elif case == "very_wide":
        # 7001 x 3 offsets: the first 2048 entries are in map row 0, their lower
        # neighbours (row 1) beyond the LDS prefix of 8192 entries
        imgs = frames_with_shifts((40, 7100), [(0, 0), (-2, 0)], 11)
        p = api.motion_params(7100, 40, 2, (3, 2), 1800, 16, (3500, 1), abs_thr=0.3, rel_thr=0.0)
    else:
        imgs = frames_with_shifts((72, 72), [(0, 0), (2, -3)], 6, noise=0.2)
        p = api.motion_params(72, 72, 2, (5, 5), 14, 16, (5, 5), abs_thr=0.2, rel_thr=0.01, min_var=0.01)
    F = p.frame_count
    got = np.zeros((F, p.motion_grid_dims[1], p.motion_grid_dims[0], 4), np.float32)
    assert api.motion_detection(imgs, got, p)
    want = orc.motion_detection(imgs, p)
    compare(got, want)


def test_motion_matlab_shim(lib, orc):
    imgs = frames_with_shifts((64, 64), [(0, 0), (1, 2)], 7)
    p = api.motion_params(64, 64, 2, (4, 4), 14, 14, (4, 4), abs_thr=0.3, rel_thr=0.0)
    got = np.zeros((2, 4, 4, 4), np.float32)
    lib.motion_detect_f32_matlab(imgs.ctypes.data_as(C.POINTER(C.c_float)), p,
                                 got.ctypes.data_as(C.POINTER(C.c_float)))
    compare(got, orc.motion_detection(imgs, p))


@pytest.mark.parametrize("tshape,sshape", [((16, 16), (64, 80)), ((7, 11), (20, 13)), ((32, 32), (32, 32))])
def test_corr_images_matches_oracle(lib, orc, tshape, sshape):
    src = speckle(sshape, 8) + 3.0
    tpl = src[2:2 + tshape[0], 1:1 + tshape[1]].copy() if tshape[0] + 2 <= sshape[0] and tshape[1] + 1 <= sshape[1] \
        else speckle(tshape, 9)
    out = np.zeros((sshape[0] - tshape[0] + 1, sshape[1] - tshape[1] + 1), np.float32)
    assert api.corr_images(np.ascontiguousarray(tpl), src, out)
    want = orc.zncc_valid(src, tpl)
    assert np.abs(out - want).max() < 1e-5
    # matlab shim
    par = L.CorrImagesParameters()
    par.template_dims[:] = (tshape[1], tshape[0])
    par.source_dims[:] = (sshape[1], sshape[0])
    out2 = np.zeros_like(out)
    lib.corr_images_f32_matlab(np.ascontiguousarray(tpl).ctypes.data_as(C.POINTER(C.c_float)),
                               src.ctypes.data_as(C.POINTER(C.c_float)), par,
                               out2.ctypes.data_as(C.POINTER(C.c_float)))
    assert np.array_equal(out, out2)
