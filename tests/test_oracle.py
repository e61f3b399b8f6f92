"""Pins the FP64 oracle (CPU only).

The reference ships no golden vectors except its Hadamard base tables, so the
oracle is pinned by that fixture, by textbook identities, independent library
routines (numpy / scipy) and brute force on small inputs.  Each check is
chosen so that a plausible slip (dropped term, wrong sign, transposed
operand, off-by-one) fails it.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from cuda_ultrasound_toolkit_amd import _lib as L

GOLDEN = Path(__file__).resolve().parent / "golden"


# ---- Hadamard ---------------------------------------------------------------
def test_base_tables_match_reference_fixture(orc):
    tables = json.loads((GOLDEN / "hadamard_base_tables.json").read_text())
    assert np.array_equal(orc.hadamard(12, L.HADAMARD), np.array(tables["h12"], float))
    assert np.array_equal(orc.hadamard(20, L.HADAMARD), np.array(tables["h20"], float))


@pytest.mark.parametrize("n", [1, 2, 4, 8, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 256])
def test_hadamard_orthogonal(orc, n):
    for order in (L.HADAMARD, L.WALSH):
        H = orc.hadamard(n, order)
        assert H is not None
        assert np.array_equal(np.abs(H), np.ones((n, n)))
        assert np.array_equal(H @ H.T, n * np.eye(n))


def test_hadamard_sylvester_doubling(orc):
    for base in (1, 12, 20):
        n = base
        while 2 * n <= 160:
            A, B = orc.hadamard(n, L.HADAMARD), orc.hadamard(2 * n, L.HADAMARD)
            assert np.array_equal(B, np.block([[A, A], [A, -A]]))
            n *= 2


def test_walsh_order_is_sequency(orc):
    for n in (2, 8, 64, 128):
        W = orc.hadamard(n, L.WALSH)
        changes = [(np.diff(np.sign(r)) != 0).sum() for r in W]
        assert changes == list(range(n))
        # same rows as the natural order, permuted
        H = orc.hadamard(n, L.HADAMARD)
        assert sorted(map(tuple, W)) == sorted(map(tuple, H))


@pytest.mark.parametrize("n", [3, 6, 36, 100, 2048])
def test_hadamard_rejects(orc, n):
    assert orc.hadamard(n, L.HADAMARD) is None


def test_decode_matches_matmul_and_inverts_encoding(orc):
    rng = np.random.default_rng(0)
    for T in (4, 12, 32):
        H = orc.hadamard(T, L.HADAMARD)
        x = rng.integers(-3000, 3000, size=(T, 5, 7)).astype(np.float64)
        dec = orc.decode(x, H)
        ref = np.einsum("jk,kcs->jcs", H, x) / T
        assert np.allclose(dec, ref, rtol=0, atol=1e-12)
        # encoding with H^T is undone exactly
        r = rng.standard_normal((T, 3, 4))
        enc = np.einsum("tk,tcs->kcs", H, r)
        assert np.allclose(orc.decode(enc, H), r, atol=1e-12)


def test_convert_index_map(orc):
    rng = np.random.default_rng(1)
    T, C, S, RC = 3, 4, 5, 6
    R0 = T * S + 7  # padded rows: padding must be ignored
    raw = rng.integers(-32768, 32767, size=(RC, R0)).astype(np.int16)
    cmap = np.array([5, 0, 3, 2], np.int16)
    out = orc.convert(raw, L.TYPE_I16, cmap, R0, T, C, S)
    for t in range(T):
        for c in range(C):
            for s in range(S):
                assert out[t, c, s] == raw[cmap[c], t * S + s]
    rawf = rng.standard_normal((RC, R0)).astype(np.float32)
    outf = orc.convert(rawf, L.TYPE_F32, cmap, R0, T, C, S)
    assert outf[2, 1, 4] == rawf[0, 2 * S + 4]


# ---- DFT / FFT / Hilbert -----------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 12, 13, 16, 31, 60, 64])
def test_fft_matches_bruteforce_dft(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        assert np.allclose(orc.fft(x, sign), orc.dft(x, sign), atol=1e-9 * max(1, n))


@pytest.mark.parametrize("n", [1000, 1024, 4096, 1031, 5000, 8191, 16384, 32768, 65536])
def test_fft_matches_numpy(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    tol = 1e-8 * max(1.0, np.sqrt(n / 4096))
    assert np.allclose(orc.fft(x, -1), np.fft.fft(x), atol=tol)
    assert np.allclose(orc.fft(x, 1), np.fft.ifft(x) * n, atol=tol)


def test_hilbert_closed_forms(orc):
    S = 64
    n = np.arange(S)
    k = 5
    w = 2 * np.pi * k / S
    y = orc.hilbert(np.cos(w * n)[None])[0]
    assert np.allclose(y, 0.75 * np.exp(1j * w * n), atol=1e-12)
    y = orc.hilbert(np.sin(w * n)[None])[0]
    assert np.allclose(y, 0.75 * -1j * np.exp(1j * w * n), atol=1e-12)
    # DC passes at 0.75
    assert np.allclose(orc.hilbert(np.full((1, S), 2.0))[0], 1.5, atol=1e-12)
    # the Nyquist term is weighted 1.5/S rather than 0.75/S
    alt = (-1.0) ** n
    assert np.allclose(orc.hilbert(alt[None])[0], 1.5 * alt, atol=1e-12)


def test_hilbert_matches_scipy_analytic_signal(orc):
    from scipy.signal import hilbert as sp_hilbert
    rng = np.random.default_rng(3)
    for S in (63, 64, 1024):
        x = rng.standard_normal((3, S))
        y = orc.hilbert(x)
        a = sp_hilbert(x, axis=-1)
        if S % 2 == 0:  # remove the Nyquist bin where the weights differ
            X = np.fft.fft(x, axis=-1)
            nyq = X[:, S // 2][:, None] * (-1.0) ** np.arange(S) / S
            y = y - 1.5 * nyq
            a = a - nyq
        assert np.allclose(y, 0.75 * a, atol=1e-10)


def test_hilbert_filter_is_circular_convolution(orc):
    rng = np.random.default_rng(4)
    S = 128
    x = rng.standard_normal((2, S))
    base = orc.hilbert(x)
    assert np.allclose(orc.hilbert(x, np.array([1.0])), base, atol=1e-10)
    for d in (1, 5):
        h = np.zeros(d + 1)
        h[d] = 1.0
        assert np.allclose(orc.hilbert(x, h), np.roll(base, d, axis=-1), atol=1e-10)
    h = rng.standard_normal(9)
    conv = np.real(np.fft.ifft(np.fft.fft(x, axis=-1) * np.fft.fft(h, S)))
    assert np.allclose(orc.hilbert(x, h), orc.hilbert(conv), atol=1e-9)
    F = orc.filter_spectrum(h, S)
    assert np.allclose(F, np.fft.fft(h, S), atol=1e-12)


# ---- DAS ----------------------------------------------------------------------
def _bp_single(seq, T=1, C=1, S=64, nx=4, nz=3):
    from cuda_ultrasound_toolkit_amd import configs
    if seq == L.HERCULES:
        acq = configs.rca_3d(T=T, S=S, n=nx, generate=False, filter_taps=0)
        acq.bp.output_points[:] = (nx, 2, nz, 1)
    else:
        acq = configs.linear_2d(T=T, S=S, nx=nx, nz=nz, generate=False)
    return acq.bp


def test_das_constants_closed_form(orc):
    from cuda_ultrasound_toolkit_amd import configs
    bp = configs.rca_3d(generate=False).bp
    k = orc.das_constants(bp)
    assert k.spm == np.float32(np.float32(20e6) / np.float32(1540.0))
    assert abs(k.xdc_min_x - (-19.2e-3 + 0.15e-3)) < 1e-9
    assert abs(k.res_x - 38e-3 / 127) < 1e-9 and k.res_y == k.res_x
    assert abs(k.res_z - 55e-3 / 127) < 1e-9
    assert k.focus_type == 2 and k.focal_z == 0.0  # inf focus -> YZ plane, f_z cleared
    assert k.delay == 63
    bp.foci[2] = 0.02
    assert orc.das_constants(bp).focus_type == 4
    bp.tx_orientation = L.ORIENT_COLS
    assert orc.das_constants(bp).focus_type == 3


def test_das_single_element_reads_the_model_delay(orc):
    """T = C = 1, no apodization: out = interp(iq, (z + |rx|) * spm + delay)."""
    bp = _bp_single(L.HERCULES, S=256)
    bp.fn_rx = 0.0
    rng = np.random.default_rng(5)
    iq = rng.standard_normal((1, 1, 256)) + 1j * rng.standard_normal((1, 1, 256))
    k = orc.das_constants(bp)
    vol, act = orc.das(iq, bp, cubic=False)
    for iz in range(k.Nz):
        for iy in range(k.Ny):
            for ix in range(k.Nx):
                vx = k.min_x + ix * np.float64(k.res_x)
                vy = k.min_y + iy * np.float64(k.res_y)
                vz = k.min_z + iz * np.float64(k.res_z)
                r = np.sqrt((k.xdc_min_x - vx) ** 2 + (k.xdc_min_y - vy) ** 2 + vz * vz)
                x = np.clip((vz + r) * np.float64(k.spm) + k.delay, 1, 254)
                i = int(np.floor(x))
                want = iq[0, 0, i] + (x - i) * (iq[0, 0, i + 1] - iq[0, 0, i])
                assert abs(vol[iz, iy, ix] - want) < 1e-9
    assert act == k.Nx * k.Ny * k.Nz


def test_das_cubic_reproduces_cubic_polynomials(orc):
    """Catmull-Rom is exact on quadratics; constant data gives sum of weights."""
    bp = _bp_single(L.FORCES, T=4, S=512, nx=5, nz=4)
    s = np.arange(512, dtype=float)
    iq = np.broadcast_to((2.0 + 0j) * np.ones(512), (4, 4, 512)).copy()
    vol, act = orc.das(iq, bp, cubic=True)
    assert np.allclose(vol, 2.0 * 16, atol=1e-9)  # no apodization: 16 taps of value 2
    # a quadratic row: interpolation is exact, so the sum equals the analytic value
    q = 1e-4 * s * s - 0.3 * s + 7
    iq = np.broadcast_to(q + 0j, (4, 4, 512)).copy()
    vol, _ = orc.das(iq, bp, cubic=True)
    vol_lin, _ = orc.das(iq, bp, cubic=False)
    k = orc.das_constants(bp)
    vx = k.min_x + 2 * np.float64(k.res_x)
    vz = k.min_z + 1 * np.float64(k.res_z)
    want = 0.0
    for c in range(4):
        ec = k.xdc_min_x + c * np.float64(k.pitch_x)
        for t in range(4):
            et = k.xdc_min_x + t * np.float64(k.pitch_x)
            x = (np.hypot(vx - et, vz) + np.hypot(ec - vx, vz)) * np.float64(k.spm) + k.delay
            x = min(max(x, 1), 510)
            want += 1e-4 * x * x - 0.3 * x + 7
    assert abs(vol[1, 0, 2].real - want) < 1e-8
    assert abs(vol_lin[1, 0, 2].real - want) > 1e-7  # linear is not exact on a quadratic


def test_das_point_scatterer_focuses(orc):
    """A single synthetic scatterer beamforms to its own voxel (HERCULES)."""
    from cuda_ultrasound_toolkit_amd import configs
    acq = configs.rca_small(n_scatterers=1, noise=0.0, seed=11)
    st = orc.pipeline(acq, cubic=True, return_stages=True)
    env = np.abs(st["vol"])
    iz, iy, ix = np.unravel_index(np.argmax(env), env.shape)
    k = orc.das_constants(acq.bp)
    # recover the scatterer position the synthesizer drew (same RNG stream)
    seed = acq.spec.seed
    M = (1 << 64) - 1
    s = (seed * 0x2545F4914F6CDD1D + 1) & M

    def nxt():
        nonlocal s
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return ((z ^ (z >> 31)) >> 11) / 9007199254740992.0
    bmin, bmax = acq.spec.box_min, acq.spec.box_max
    p = [bmin[i] + (bmax[i] - bmin[i]) * nxt() for i in range(3)]
    vox = (k.min_x + ix * k.res_x, k.min_y + iy * k.res_y, k.min_z + iz * k.res_z)
    assert abs(vox[0] - p[0]) <= 1.5 * k.res_x
    assert abs(vox[1] - p[1]) <= 1.5 * k.res_y
    assert abs(vox[2] - p[2]) <= 1.5 * k.res_z


def test_forces_coherence_factor(orc):
    bp = _bp_single(L.FORCES, T=2, S=128, nx=3, nz=2)
    rng = np.random.default_rng(6)
    iq = rng.standard_normal((2, 2, 128)) + 1j * rng.standard_normal((2, 2, 128))
    v0, _ = orc.das(iq, bp, cubic=True)
    bp.coherency_weighting = 1.0
    v1, _ = orc.das(iq, bp, cubic=True)
    # with cw = 1 the output is acc * |acc|^2 / sum|v|^2, so |v1| <= |v0| * N
    ratio = np.abs(v1) / np.abs(v0)
    assert np.all(ratio > 0) and np.all(ratio <= 4 + 1e-9)
    bp.coherency_weighting = 7.0  # clamped to 1
    v2, _ = orc.das(iq, bp, cubic=True)
    assert np.allclose(v1, v2)


def test_hann_apodization_threshold(orc):
    """HERCULES keeps exactly the taps with cos^2(pi*min(fn*rho/(2z), .5)) > 0.1."""
    from cuda_ultrasound_toolkit_amd import configs
    acq = configs.rca_small(generate=False)
    bp = acq.bp
    k = orc.das_constants(bp)
    iq = np.ones((k.T, k.C, k.S), complex)
    vol, act = orc.das(iq, bp, cubic=True)
    want = 0
    xs = k.min_x + np.arange(k.Nx) * np.float64(k.res_x)
    ys = k.min_y + np.arange(k.Ny) * np.float64(k.res_y)
    zs = k.min_z + np.arange(k.Nz) * np.float64(k.res_z)
    ex = k.xdc_min_x + np.arange(k.C) * np.float64(k.pitch_x)
    ey = k.xdc_min_y + np.arange(k.T) * np.float64(k.pitch_y)
    rho = np.sqrt((ex[None, :, None, None] - xs[None, None, None, :]) ** 2 +
                  (ey[:, None, None, None] - ys[None, None, :, None]) ** 2)  # [T][C][y][x]
    for iz, z in enumerate(zs):
        a = np.minimum(k.fn_rx * rho / z / 2, 0.5)
        apo = np.cos(np.pi * a) ** 2
        m = apo > 0.1
        want += m.sum()
        assert np.allclose(vol[iz].real, np.where(m, apo, 0).sum(axis=(0, 1)), atol=1e-9)
    assert act == want
