"""The oracle's UFORCES and RCA_TPW restatements (beamformer_kernels.cu:94-172,
:489-546) pinned by hand-computed single-path cases and closed forms.

Hand-computed cases place a voxel at a known position relative to one
transmit and one receive element and read a ramp (Catmull-Rom reproduces a
linear row exactly, so the output equals the sample index the delay model
gives, times the weights).  Closed forms use rows of constant value, where
the output is a sum of apodization weights.  These pin the element positions
(including the reference's extra half pitch), the row mapping (UFORCES skips
the flash in row block 0), the 1/sqrt(128) weight of the first UFORCES
transmit, the plane-wave steering and the sign and frequency of the TPW phase
rotation."""
import math

import numpy as np
import pytest

from cuda_ultrasound_toolkit_amd import _lib as L, configs


def ramp(T, C, S, row_offset=0.0):
    s = np.arange(S, dtype=float)
    rows = np.arange(T * C, dtype=float).reshape(T, C, 1)
    return (s[None, None, :] + row_offset * rows) + 0j


def set_grid(bp, xs, zs, y=(0.0, 0.0)):
    bp.output_points[:] = (len(xs) if len(xs) > 1 else 2, 1, len(zs), 1)
    bp.output_min_coordinate[:] = (xs[0], y[0], zs[0], 0.0)
    bp.output_max_coordinate[:] = (xs[-1], y[1], zs[-1], 0.0)


def test_uforces_single_paths(orc):
    """Flash + 2 transmits from elements 0 and 4, one receive element, voxels
    straight below the first transmit element and 1 mm beside it."""
    acq = configs.uforces_2d(T=3, S=2048, nx=2, nz=2, fn_rx=1.0, generate=False)
    bp = acq.bp
    bp.dec_data_dim[1] = 1
    bp.sparse_elements[0], bp.sparse_elements[1] = 0, 4
    k0 = orc.das_constants(bp)
    p = np.float64(k0.pitch_x)
    ex = np.float64(k0.xdc_min_x) + p / 2  # the element the kernel places at x_min + p/2
    set_grid(bp, [float(np.float32(ex)), float(np.float32(ex + 1e-3))], [10e-3, 20e-3])
    k = orc.das_constants(bp)
    iq = ramp(3, 1, 2048, row_offset=1000.0)
    vol, act = orc.das(iq, bp, cubic=True)
    spm = np.float64(k.spm)
    for iz in range(2):
        for ix in range(2):
            vx = np.float64(k.min_x) + ix * np.float64(k.res_x)
            vz = np.float64(k.min_z) + iz * np.float64(k.res_z)
            d = ex - vx                       # receive element minus voxel
            rx = math.hypot(d, vz)
            apo = math.cos(math.pi * min(1.0 * abs(d) / vz / 2, 0.5)) ** 2
            want = 0.0
            for t, (shift, scale) in enumerate([(0.0, 1 / math.sqrt(128)), (4 * p, 1.0)]):
                tx = math.hypot(d + shift, vz)   # focus at depth 0 above element 0
                idx = (rx + tx) * spm + k.delay
                want += scale * apo * (idx + 1000.0 * (t + 1))   # row t + 1 (row 0 is the flash)
            assert abs(vol[iz, 0, ix] - want) < 1e-7 * abs(want), (iz, ix)
    assert act == 2 * 2 * 2


def test_uforces_constant_rows(orc):
    """Rows of constant value r[row]: out = sum_t sum_e s_t apo(e) r[(t+1) C + e]."""
    acq = configs.uforces_2d(T=6, S=1024, nx=7, nz=3, fn_rx=1.5, generate=False)
    bp = acq.bp
    set_grid(bp, list(np.linspace(-4e-3, 4e-3, 7)), [8e-3, 12e-3, 16e-3])
    k = orc.das_constants(bp)
    T, Cc, S = k.T, k.C, k.S
    rng = np.random.default_rng(3)
    r = rng.standard_normal(T * Cc) + 1j * rng.standard_normal(T * Cc)
    iq = np.broadcast_to(r.reshape(T, Cc, 1), (T, Cc, S)).copy()
    vol, act = orc.das(iq, bp, cubic=True)
    p = np.float64(k.pitch_x)
    ex = np.float64(k.xdc_min_x) + p / 2 + np.arange(Cc) * p
    for iz in range(k.Nz):
        vz = np.float64(k.min_z) + iz * np.float64(k.res_z)
        for ix in range(k.Nx):
            vx = np.float64(k.min_x) + ix * np.float64(k.res_x)
            a = np.minimum(k.fn_rx * np.abs(ex - vx) / vz / 2, 0.5)
            apo = np.cos(np.pi * a) ** 2
            want = sum((1 / math.sqrt(128) if t == 0 else 1.0) * np.sum(apo * r[(t + 1) * Cc:(t + 2) * Cc])
                       for t in range(T - 1))
            assert abs(vol[iz, 0, ix] - want) < 1e-9 * max(1.0, abs(want))


@pytest.mark.parametrize("orient", ["yz", "xz"])
def test_tpw_single_paths(orc, orient):
    """One receive element, two steering angles: out = sum_a idx_a e^{2 pi i idx_a fc/fs}
    with idx_a = (|rx| + lat sin(a) + z cos(a)) spm + delay."""
    acq = configs.tpw_3d(T=2, C=1, S=2048, n=2, generate=False)
    bp = acq.bp
    if orient == "xz":
        bp.rx_orientation = L.ORIENT_ROWS
    angles = [math.radians(-7.0), math.radians(12.0)]
    for t, a in enumerate(angles):
        bp.foci[3 * t + 2] = a
    k0 = orc.das_constants(bp)
    p = np.float64(k0.pitch_x)
    ex = np.float64(k0.xdc_min_x) + p / 2
    bp.output_points[:] = (2, 2, 2, 1)
    bp.output_min_coordinate[:] = (float(np.float32(ex)), -2e-3, 9e-3, 0.0)
    bp.output_max_coordinate[:] = (float(np.float32(ex + 1.5e-3)), 0.0, 13e-3, 0.0)
    k = orc.das_constants(bp)
    assert k.focus_type == (2 if orient == "yz" else 1)
    iq = ramp(2, 1, 2048)
    vol, act = orc.das(iq, bp, cubic=True)
    f = np.float64(np.float32(bp.center_frequency)) / np.float64(np.float32(bp.sampling_frequency))
    spm = np.float64(k.spm)
    for iz in range(2):
        for iy in range(2):
            for ix in range(2):
                vx = np.float64(k.min_x) + ix * np.float64(k.res_x)
                vy = np.float64(k.min_y) + iy * np.float64(k.res_y)
                vz = np.float64(k.min_z) + iz * np.float64(k.res_z)
                d = ex - vx
                apo = math.cos(math.pi * min(k.fn_rx * abs(d) / vz / 2, 0.5)) ** 2
                lat = vy if orient == "yz" else vx
                want = 0j
                if apo >= 0.1:
                    for a in angles:
                        idx = (math.hypot(d, vz) + lat * math.sin(np.float32(a)) + vz * math.cos(np.float32(a))) * spm + k.delay
                        want += apo * idx * np.exp(2j * math.pi * idx * f)
                assert abs(vol[iz, iy, ix] - want) < 1e-7 * max(1.0, abs(want)), (iz, iy, ix)


def test_tpw_constant_iq_is_sum_of_weights(orc):
    """fc = 0 (no rotation) and IQ = 1: out = T sum_c apo_c [apo_c >= 0.1]; the
    active-tap count is the number of such (voxel, channel) pairs times T."""
    acq = configs.tpw_3d(T=5, C=16, S=1024, n=6, generate=False)
    bp = acq.bp
    bp.center_frequency = 0.0
    k = orc.das_constants(bp)
    iq = np.ones((k.T, k.C, k.S), complex)
    vol, act = orc.das(iq, bp, cubic=True)
    p = np.float64(k.pitch_x)
    ex = np.float64(k.xdc_min_x) + p / 2 + np.arange(k.C) * p
    n_act = 0
    for iz in range(k.Nz):
        vz = np.float64(k.min_z) + iz * np.float64(k.res_z)
        for ix in range(k.Nx):
            vx = np.float64(k.min_x) + ix * np.float64(k.res_x)
            apo = np.cos(np.pi * np.minimum(k.fn_rx * np.abs(ex - vx) / vz / 2, 0.5)) ** 2
            m = apo >= 0.1
            n_act += int(m.sum()) * k.Ny
            assert np.allclose(vol[iz, :, ix], k.T * apo[m].sum(), atol=1e-9)
    assert act == n_act * k.T


def test_legacy_coherence_factor_bounds(orc):
    acq = configs.tpw_3d(T=3, C=8, S=512, n=4, generate=False, z_range=(2e-3, 6e-3))
    bp = acq.bp
    rng = np.random.default_rng(8)
    iq = rng.standard_normal((3, 8, 512)) + 1j * rng.standard_normal((3, 8, 512))
    v0, _ = orc.das(iq, bp)
    bp.coherency_weighting = 1.0
    v1, _ = orc.das(iq, bp)
    ratio = np.abs(v1) / np.abs(v0)
    assert np.all(ratio >= 0) and np.all(ratio <= 3 * 8 + 1e-9)


def test_legacy_rejections(orc):
    acq = configs.uforces_2d(T=4, S=256, nx=4, nz=4, generate=False)
    iq = np.zeros((4, 4, 256), complex)
    acq.bp.sparse_elements[0] = -1
    with pytest.raises(ValueError):
        orc.das(iq, acq.bp)
    acq = configs.uforces_2d(T=4, S=256, nx=4, nz=4, generate=False)
    acq.bp.tx_orientation = L.ORIENT_NONE
    with pytest.raises(ValueError):
        orc.das(iq, acq.bp)
    acq = configs.tpw_3d(T=2, C=4, S=256, n=4, generate=False)
    acq.bp.tx_orientation = L.ORIENT_NONE
    with pytest.raises(ValueError):
        orc.das(np.zeros((2, 4, 256), complex), acq.bp)
