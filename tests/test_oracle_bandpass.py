"""The oracle's bandpass input path (data_conversion_kernels.cuh:14-102,
data_converter.cu:64-131) pinned by closed forms: the (I, Q)-pair index map
with the channel permutation, the conjugation and demodulation (a tone at the
demodulation frequency becomes a constant), and the centred FIR's tap order,
centring and valid-count normalisation (box filters on constants and ramps,
a shifted impulse)."""
import math

import numpy as np

from cuda_ultrasound_toolkit_amd import _lib as L


def _raw(T, Cc, S, fn, pad=6, dtype=np.float32):
    """Raw rows [C][2 T S + pad] with (I, Q) = fn(t, c_raw, s)."""
    raw = np.zeros((Cc, 2 * T * S + pad), dtype)
    for c in range(Cc):
        for t in range(T):
            for s in range(S):
                i, q = fn(t, c, s)
                raw[c, 2 * (t * S + s)] = i
                raw[c, 2 * (t * S + s) + 1] = q
    return raw


def test_index_map_and_conjugation(orc):
    T, Cc, S = 3, 4, 8
    cmap = np.array([2, 0, 3, 1], np.int16)
    raw = _raw(T, Cc, S, lambda t, c, s: (100 * t + 10 * c + s, -(1000 + 100 * t + 10 * c + s)))
    out = orc.convert_demod(raw, L.TYPE_F32, cmap, raw.shape[1], T, Cc, S, 0.0, 20e6)
    for t in range(T):
        for c in range(Cc):
            for s in range(S):
                rc = cmap[c]
                want = (100 * t + 10 * rc + s) + 1j * (1000 + 100 * t + 10 * rc + s)  # I - iQ
                assert out[t, c, s] == want


def test_tone_at_the_demodulation_frequency_becomes_constant(orc):
    T, Cc, S = 2, 2, 256
    fd, fs = 5e6, 20e6 * 0.98
    w = 2 * math.pi * fd / fs
    A = 0.7 - 0.4j

    def fn(t, c, s):
        z = A * (c + 1) * np.exp(1j * w * s)       # I - iQ = z
        return z.real, -z.imag
    raw = _raw(T, Cc, S, fn, dtype=np.float64).astype(np.float32)
    out = orc.convert_demod(raw, L.TYPE_F32, np.arange(Cc, dtype=np.int16), raw.shape[1], T, Cc, S, fd, fs)
    for c in range(Cc):
        assert np.allclose(out[:, c, :], A * (c + 1), atol=3e-7)


def test_box_filter_normalises_by_valid_taps(orc):
    T, Cc, S = 1, 1, 32
    cmap = np.zeros(1, np.int16)
    const = _raw(T, Cc, S, lambda t, c, s: (3.0, 0.0))
    ramp = _raw(T, Cc, S, lambda t, c, s: (float(s), 0.0))
    for Lh in (1, 4, 5, 9):
        h = np.ones(Lh)
        y = orc.convert_demod(const, L.TYPE_F32, cmap, const.shape[1], T, Cc, S, 0.0, 20e6, h)[0, 0]
        assert np.allclose(y, 3.0, atol=1e-12)           # edges too: divided by the valid count
        y = orc.convert_demod(ramp, L.TYPE_F32, cmap, ramp.shape[1], T, Cc, S, 0.0, 20e6, h)[0, 0].real
        half = Lh // 2
        for s in range(S):
            js = [s + i - half for i in range(Lh) if 0 <= s + i - half < S]
            assert abs(y[s] - sum(js) / len(js)) < 1e-12
        # interior: a centred box average of a ramp is the ramp shifted by (L - 1)/2 - L//2
        assert abs(y[S // 2] - (S // 2 + (Lh - 1) / 2 - half)) < 1e-12


def test_fir_tap_order(orc):
    """h = [1, 0, 0]: y[s] = v[s + 1] / 3 in the interior (taps applied reversed
    around the centre L // 2 = 1, normalised by the count of valid taps)."""
    T, Cc, S = 1, 1, 16
    rng = np.random.default_rng(2)
    vals = rng.standard_normal((S, 2))
    raw = _raw(T, Cc, S, lambda t, c, s: (vals[s, 0], vals[s, 1]))
    v = orc.convert_demod(raw, L.TYPE_F32, np.zeros(1, np.int16), raw.shape[1], T, Cc, S, 0.0, 20e6)[0, 0]
    y = orc.convert_demod(raw, L.TYPE_F32, np.zeros(1, np.int16), raw.shape[1], T, Cc, S, 0.0, 20e6,
                          np.array([1.0, 0.0, 0.0]))[0, 0]
    assert np.allclose(y[1:S - 1], v[2:S] / 3, atol=1e-12)
    assert abs(y[0] - v[1] / 2) < 1e-12           # s = 0: taps i = 1, 2 are valid
    assert abs(y[S - 1] - 0.0) < 1e-12            # s = S-1: the only nonzero tap reads past the row


def test_int16_rows(orc):
    T, Cc, S = 2, 3, 5
    raw = _raw(T, Cc, S, lambda t, c, s: (-32768 + 7 * s + t, 32767 - c), dtype=np.int16)
    out = orc.convert_demod(raw, L.TYPE_I16, np.array([1, 2, 0], np.int16), raw.shape[1], T, Cc, S, 0.0, 1.0)
    assert out[1, 0, 4] == (-32768 + 28 + 1) - 1j * (32767 - 1)
