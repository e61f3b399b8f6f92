"""Bandpass input on the GPU (data_conversion_kernels.cuh:14-102): the fused
convert + demodulation + centred-FIR kernel against the FP64 oracle, and the
whole path (bandpass rows -> IQ -> DAS) through the beamform entry point.

Tolerances: the stage is FP32 with one rounding per product and at most 1024
accumulated taps, relative L2 < 1e-6; the whole path holds the DAS bar,
relative L2 < 1e-4."""
import numpy as np
import pytest

from cuda_ultrasound_toolkit_amd import _lib as L, api, configs

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _t(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def bandpass_rows(rng, Cc, T, S, dtype, pad=10, f=0.23):
    """(I, Q) rows [C][2 T S + pad]: a band-limited complex tone train plus noise."""
    n = np.arange(T * S)
    z = np.exp(2j * np.pi * f * n)[None, :] * (1 + 0.5 * np.sin(2 * np.pi * n / 97.0))[None, :]
    z = z * rng.uniform(0.3, 1.0, (Cc, 1)) + 0.2 * (rng.standard_normal((Cc, T * S)) + 1j * rng.standard_normal((Cc, T * S)))
    raw = np.zeros((Cc, 2 * T * S + pad))
    raw[:, 0:2 * T * S:2] = z.real * 9000
    raw[:, 1:2 * T * S:2] = -z.imag * 9000
    return raw.astype(dtype)


@pytest.mark.parametrize("dtype", [np.int16, np.float32])
@pytest.mark.parametrize("taps", [0, 1, 16, 63, 64, 1024])
def test_convert_demod_stage(torch_dev, orc, dtype, taps):
    import torch
    print(f"bandpass stage dtype={np.dtype(dtype).name} taps={taps}", flush=True)
    T, Cc, S = 3, 8, 2500
    rng = np.random.default_rng(taps + 7)
    raw = bandpass_rows(rng, Cc, T, S, dtype)
    cmap = rng.permutation(Cc).astype(np.int16)
    R0 = raw.shape[1]
    assert api.rf_init((R0, Cc), (S, Cc, T), L.NONE)
    assert api.rf_set_channel_mapping(cmap)
    h = rng.standard_normal(taps).astype(np.float32) if taps else None
    assert api.rf_set_match_filter(h)
    out = torch.zeros(T * Cc * S * 2, dtype=torch.float32, device=torch_dev)
    typ = L.TYPE_I16 if dtype == np.int16 else L.TYPE_F32
    fd, fs = np.float32(5.3e6), np.float32(20.8e6)   # fd/fs not a short binary fraction
    assert api.rf_convert_demod(_t(raw, torch_dev), typ, out, float(fd), float(fs))
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.complex64).reshape(T, Cc, S)
    want = orc.convert_demod(raw, typ, cmap, R0, T, Cc, S, float(fd), float(fs),
                             None if h is None else h.astype(np.float64))
    err = rel_l2(got, want)
    print(f"  rel L2 {err:.3e}")
    assert err < 1e-6
    assert api.rf_set_match_filter(None)


@pytest.mark.parametrize("dtype", [L.TYPE_I16, L.TYPE_F32])
@pytest.mark.parametrize("S,taps", [(1003, 64), (1003, 0), (1029, 1024), (17, 5), (5, 4), (2, 1)])
def test_convert_demod_ragged_rows(torch_dev, orc, S, taps, dtype):
    """Row lengths that are not multiples of the four outputs a thread
    computes, rows shorter than the filter's reach, and the widest filter on a
    row barely longer than it (nearly every output at a row end).  A padding
    of 3 samples makes rf_raw_dim[0] odd, so half of the rows start off the
    (I, Q) pair alignment (8 bytes for float32) and take the unpaired loads."""
    import torch
    print(f"bandpass ragged S={S} taps={taps} dtype={dtype}", flush=True)
    T, Cc = 2, 3
    rng = np.random.default_rng(S * 7 + taps)
    raw = bandpass_rows(rng, Cc, T, S, np.int16 if dtype == L.TYPE_I16 else np.float32, pad=3)
    assert raw.shape[1] % 2 == 1
    cmap = rng.permutation(Cc).astype(np.int16)
    R0 = raw.shape[1]
    assert api.rf_init((R0, Cc), (S, Cc, T), L.NONE)
    assert api.rf_set_channel_mapping(cmap)
    h = rng.standard_normal(taps).astype(np.float32) if taps else None
    assert api.rf_set_match_filter(h)
    out = torch.full((T * Cc * S * 2,), np.nan, dtype=torch.float32, device=torch_dev)
    assert api.rf_convert_demod(_t(raw, torch_dev), dtype, out, 5.3e6, 20.8e6)
    torch.cuda.synchronize()
    assert api.rf_set_match_filter(None)
    got = out.cpu().numpy().view(np.complex64).reshape(T, Cc, S)
    want = orc.convert_demod(raw, dtype, cmap, R0, T, Cc, S, 5.3e6, 20.8e6,
                             None if h is None else h.astype(np.float64))
    assert np.all(np.isfinite(got))
    err = rel_l2(got, want)
    print(f"  rel L2 {err:.3e}")
    assert err < 1e-6


def _bandpass_frame(acq, seed, taps):
    bp = acq.bp
    S, Cc, T = bp.dec_data_dim[0], bp.dec_data_dim[1], bp.dec_data_dim[2]
    rng = np.random.default_rng(seed)
    raw = bandpass_rows(rng, Cc, T, S, np.int16, f=0.25)
    bp.sample_type = L.SAMPLE_BANDPASS
    bp.decode = L.NONE
    bp.rf_raw_dim[0] = raw.shape[1]
    if taps:
        h = np.hanning(taps + 2)[1:-1].astype(np.float32)
        bp.filter_length = taps
        bp.rf_filter[:taps] = h.tolist()
    else:
        bp.filter_length = 0
    acq.raw = raw
    return acq


@pytest.mark.parametrize("case", ["forces_2d", "hercules_small", "tpw_small"])
def test_bandpass_beamform_matches_oracle(torch_dev, orc, case):
    import torch
    print(f"bandpass beamform {case}", flush=True)
    if case == "forces_2d":
        acq = _bandpass_frame(configs.linear_2d(seed=1, generate=False), 3, 0)
    elif case == "hercules_small":
        acq = _bandpass_frame(configs.rca_small(seed=3, generate=False), 4, 16)
    else:
        acq = _bandpass_frame(configs.tpw_3d(seed=7, generate=False), 5, 8)
    bp = acq.bp
    n = bp.output_points
    vol = torch.zeros(n[0] * n[1] * n[2] * 2, dtype=torch.float32, device=torch_dev)
    assert api.beamform_device(_t(acq.raw, torch_dev), vol, bp)
    torch.cuda.synchronize()
    got = vol.cpu().numpy().view(np.complex64).reshape(n[2], n[1], n[0])
    want = orc.pipeline(acq)
    err = rel_l2(got, want)
    print(f"  rel L2 {err:.3e}")
    assert err < 1e-4


def test_bandpass_rejections(torch_dev):
    import torch
    acq = _bandpass_frame(configs.linear_2d(seed=1, generate=False), 3, 0)
    bp = acq.bp
    n = bp.output_points
    vol = torch.zeros(n[0] * n[1] * n[2] * 2, dtype=torch.float32, device=torch_dev)
    bp.decode = L.HADAMARD
    assert not api.beamform_device(_t(acq.raw, torch_dev), vol, bp)
    bp.decode = L.NONE
    bp.rf_raw_dim[0] = bp.dec_data_dim[0] * bp.dec_data_dim[2]   # one value per sample: too short for pairs
    assert not api.beamform_device(_t(acq.raw, torch_dev), vol, bp)
