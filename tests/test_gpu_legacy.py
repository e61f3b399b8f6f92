"""UFORCES and RCA_TPW on the GPU (beamformer_kernels.cu:94-172, :489-546)
against the FP64 oracle: the whole chain (convert, decode, Hilbert, DAS) and
the DAS alone, cubic (the reference's interpolation) and linear.

Tolerance: relative L2 1e-4 of the volume.  The kernels compute in FP32 with
sample indices of magnitude up to S; their rounding (about 3e-5 samples)
moves a band-limited analytic signal by ~1e-4 of its amplitude per tap, and
the taps' errors are independent, so the volume's error stays below that."""
import math

import numpy as np
import pytest

from cuda_ultrasound_toolkit_amd import _lib as L, api, configs

pytestmark = pytest.mark.gpu
TOL = 1e-4


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _t(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _vol(bp, dev):
    import torch
    n = bp.output_points
    return torch.zeros(n[0] * n[1] * n[2] * 2, dtype=torch.float32, device=dev)


def _beamform(acq, dev, cubic=True):
    import torch
    assert api.set_das_options(cubic=cubic)
    vol = _vol(acq.bp, dev)
    try:
        assert api.beamform_device(_t(acq.raw, dev), vol, acq.bp)
        torch.cuda.synchronize()
    finally:
        assert api.set_das_options(cubic=True)
    n = acq.bp.output_points
    return vol.cpu().numpy().view(np.complex64).reshape(n[2], n[1], n[0])


CASES = {
    "uforces_2d": lambda: configs.uforces_2d(seed=6),
    "uforces_2d_cw1_fn0": lambda: _with(configs.uforces_2d(seed=9, fn_rx=0.0), coherency_weighting=1.0),
    "tpw_3d_yz": lambda: configs.tpw_3d(seed=7),
    "tpw_3d_xz_cw05": lambda: _with(configs.tpw_3d(seed=8), rx_orientation=L.ORIENT_ROWS, coherency_weighting=0.5),
    # x and y not multiples of the 4 x 16 voxel patches the kernel's waves take
    "tpw_3d_ragged": lambda: _points(configs.tpw_3d(seed=10, n=21), (19, 13, 21)),
}


def _points(acq, n):
    acq.bp.output_points[:3] = n
    return acq


def _with(acq, **kw):
    for k, v in kw.items():
        setattr(acq.bp, k, v)
    return acq


@pytest.mark.parametrize("cubic", [True, False])
@pytest.mark.parametrize("case", list(CASES))
def test_pipeline_matches_oracle(torch_dev, orc, case, cubic):
    print(f"legacy pipeline {case} cubic={cubic}", flush=True)
    acq = CASES[case]()
    got = _beamform(acq, torch_dev, cubic)
    want = orc.pipeline(acq, cubic=cubic)
    err = rel_l2(got, want)
    print(f"  rel L2 {err:.3e}")
    assert np.all(np.isfinite(got))
    assert err < TOL


def test_das_reads_across_row_ends_like_the_reference(torch_dev, orc):
    """UFORCES indices in (0, 1) read the previous row's last sample (the
    reference's cubic spline indexes the flat buffer): a negative time offset
    puts shallow voxels there."""
    import torch
    print("legacy uforces flat reads", flush=True)
    acq = configs.uforces_2d(seed=6, T=8, S=512, nx=32, nz=48, generate=False)
    bp = acq.bp
    bp.output_min_coordinate[2], bp.output_max_coordinate[2] = 0.2e-3, 6e-3
    bp.time_offset = -12.0 / bp.sampling_frequency
    rng = np.random.default_rng(4)
    T, Cc, S = 8, 8, 512
    iq = (rng.standard_normal((T, Cc, S)) + 1j * rng.standard_normal((T, Cc, S))).astype(np.complex64)
    want, _ = orc.das(iq.astype(np.complex128), bp, cubic=True)
    vol = _vol(bp, torch_dev)
    d_iq = _t(iq.view(np.float32), torch_dev)
    assert api.das_device(d_iq, vol, bp)
    torch.cuda.synchronize()
    n = bp.output_points
    got = vol.cpu().numpy().view(np.complex64).reshape(n[2], n[1], n[0])
    err = rel_l2(got, want)
    print(f"  rel L2 {err:.3e}")
    assert err < 1e-4
    # the flat reads matter: zeroing the row-end samples changes the shallow voxels
    iq2 = iq.copy()
    iq2[:, :, -1] = 0
    want2, _ = orc.das(iq2.astype(np.complex128), bp, cubic=True)
    assert rel_l2(want2[:3], want[:3]) > 3e-3


def test_tpw_full_size_depth_planes(torch_dev, orc):
    """128 x 128 columns at full size (32 angles, 128 channels, 4096 samples),
    two depth planes, against the oracle on the same IQ; the whole 128^3
    volume is checked for linearity in the IQ (the DAS is linear when cw = 0)."""
    import torch
    print("legacy tpw full size", flush=True)
    acq = configs.tpw_3d(seed=2, T=32, C=128, S=4096, n=128, generate=False, z_range=(5e-3, 60e-3),
                         half_width=19e-3, box_half=19e-3)
    bp = acq.bp
    T, Cc, S = 32, 128, 4096
    g = torch.Generator(device=torch_dev).manual_seed(5)
    a = torch.randn(T * Cc * S * 2, device=torch_dev, generator=g)
    b = torch.randn(T * Cc * S * 2, device=torch_dev, generator=g)
    va, vb, vab = _vol(bp, torch_dev), _vol(bp, torch_dev), _vol(bp, torch_dev)
    assert api.das_device(a, va, bp) and api.das_device(b, vb, bp) and api.das_device(2.0 * a - 0.5 * b, vab, bp)
    torch.cuda.synchronize()
    lin = (vab - (2.0 * va - 0.5 * vb)).norm() / vab.norm()
    print(f"  linearity {lin.item():.3e}")
    assert lin.item() < 1e-5
    # two depth planes against the oracle
    nz = bp.output_points[2]
    zmin, zmax = bp.output_min_coordinate[2], bp.output_max_coordinate[2]
    res = np.float32((np.float32(zmax) - np.float32(zmin)) / np.float32(nz - 1))
    bp.output_points[2] = 2
    bp.output_min_coordinate[2] = float(np.float32(zmin) + np.float32(40) * res)
    bp.output_max_coordinate[2] = float(np.float32(zmin) + np.float32(100) * res)
    # smooth IQ: the analytic signal of band-limited noise
    from numpy.fft import fft, ifft
    rng = np.random.default_rng(6)
    x = rng.standard_normal((T * Cc, S))
    X = fft(x, axis=1)
    f = np.fft.fftfreq(S)
    X *= ((f > 0.1) & (f < 0.4))
    iq = ifft(X, axis=1).reshape(T, Cc, S).astype(np.complex64)
    want, _ = orc.das(iq.astype(np.complex128), bp, cubic=True)
    vol = _vol(bp, torch_dev)
    assert api.das_device(_t(iq.view(np.float32), torch_dev), vol, bp)
    torch.cuda.synchronize()
    got = vol.cpu().numpy().view(np.complex64).reshape(2, 128, 128)
    err = rel_l2(got, want)
    print(f"  depth planes rel L2 {err:.3e}")
    assert err < TOL


def test_legacy_rejections(torch_dev):
    import torch
    acq = configs.uforces_2d(seed=6, T=4, S=256, nx=8, nz=8, generate=False)
    acq.bp.sparse_elements[0] = -1
    vol = _vol(acq.bp, torch_dev)
    assert not api.beamform_device(_t(acq.raw, torch_dev), vol, acq.bp)
    acq = configs.tpw_3d(seed=7, T=4, C=8, S=256, n=8, generate=False)
    acq.bp.tx_orientation = L.ORIENT_NONE
    vol = _vol(acq.bp, torch_dev)
    assert not api.beamform_device(_t(acq.raw, torch_dev), vol, acq.bp)
    acq = configs.tpw_3d(seed=7, T=4, C=8, S=256, n=8, generate=False)
    assert api.set_das_options(iq_format=1)
    try:
        assert not api.beamform_device(_t(acq.raw, torch_dev), _vol(acq.bp, torch_dev), acq.bp)
    finally:
        assert api.set_das_options(iq_format=0)
