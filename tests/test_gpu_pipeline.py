"""HIP path vs the FP64 oracle, through the C ABI (runs on an MI355X).

Bars: decode bit-exact; convert exact; Hilbert within 1e-5 relative L2;
DAS and the whole beamform within 1e-4 relative L2 (FP32 arithmetic against
FP64, the tolerance set for beamformed IQ).
"""
import numpy as np
import pytest

from cuda_ultrasound_toolkit_amd import _lib as L, api, configs

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def _t(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _rf_setup(acq):
    bp = acq.bp
    S, Cc, T = bp.dec_data_dim[0], bp.dec_data_dim[1], bp.dec_data_dim[2]
    assert api.rf_init((bp.rf_raw_dim[0], bp.rf_raw_dim[1]), (S, Cc, T), bp.decode)
    assert api.rf_set_channel_mapping(np.array(bp.channel_mapping[:Cc], np.int16))
    return S, Cc, T


@pytest.mark.parametrize("T,S,encoding", [(32, 1024, L.HADAMARD), (128, 512, L.HADAMARD), (16, 256, L.WALSH),
                                          (64, 200, L.HADAMARD), (24, 128, L.HADAMARD), (40, 96, L.WALSH),
                                          # rows that do not start on 16 bytes (S not a multiple of 8)
                                          (32, 1001, L.HADAMARD), (128, 1027, L.HADAMARD), (16, 3001, L.WALSH),
                                          (64, 1500, L.HADAMARD), (128, 2050, L.HADAMARD)])
def test_decode_bit_exact(torch_dev, orc, T, S, encoding):
    import torch
    print(f"decode T={T} S={S} enc={encoding}", flush=True)
    acq = configs.linear_2d(T=T, S=S, encoding=encoding, seed=T + S)
    S_, Cc, T_ = _rf_setup(acq)
    raw = _t(acq.raw, torch_dev)
    H = orc.hadamard(T, encoding)
    conv = orc.convert(acq.raw, L.TYPE_I16, acq.channel_map, T * S, T, Cc, S)
    want = orc.decode(conv, H).astype(np.float32)
    pow2 = (T & (T - 1)) == 0
    outs = {}
    for mode in (1, 2):  # FP32 MFMA GEMM, int8 MFMA hi/lo
        assert api.set_decode_mode(mode)
        out = torch.full((T * Cc * S,), np.nan, dtype=torch.float32, device=torch_dev)
        assert api.rf_convert_decode(raw, L.TYPE_I16, out, 1, True)
        torch.cuda.synchronize()
        outs[mode] = out.cpu().numpy().reshape(T, Cc, S)
        if pow2:
            assert np.array_equal(outs[mode], want), f"mode {mode}"
        else:  # 1/T inexact: one rounding of sum * fl(1/T)
            assert np.allclose(outs[mode], want, rtol=2 ** -23, atol=0)
    assert np.array_equal(outs[1], outs[2])  # int8 path is bit-exact to the FP32 path
    assert api.set_decode_mode(0)
    # into the real parts of a complex buffer, imaginary parts untouched
    cplx = torch.full((T * Cc * S * 2,), 7.0, dtype=torch.float32, device=torch_dev)
    assert api.rf_convert_decode(raw, L.TYPE_I16, cplx, 2, True)
    torch.cuda.synchronize()
    c = cplx.cpu().numpy()
    assert np.array_equal(c[0::2].reshape(T, Cc, S), outs[2]) and np.all(c[1::2] == 7.0)


def test_default_stream_work_is_ordered_before_library_calls(torch_dev, orc):
    """The library's own stream is ordered with the default stream: an output
    filled on torch's default stream right before a device-pointer call is
    never written after the library's kernels (repeated, fresh buffers)."""
    import torch
    print("decode after default-stream fill", flush=True)
    T, S = 128, 512
    acq = configs.linear_2d(T=T, S=S, seed=T + S)
    S_, Cc, T_ = _rf_setup(acq)
    raw = _t(acq.raw, torch_dev)
    want = orc.decode(orc.convert(acq.raw, L.TYPE_I16, acq.channel_map, T * S, T, Cc, S),
                      orc.hadamard(T, L.HADAMARD)).astype(np.float32)
    assert api.set_decode_mode(2)
    for _ in range(8):
        out = torch.full((T * Cc * S,), np.nan, dtype=torch.float32, device=torch_dev)
        assert api.rf_convert_decode(raw, L.TYPE_I16, out, 1, True)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy().reshape(T, Cc, S), want)
    assert api.set_decode_mode(0)


def test_decode_full_int16_range(torch_dev, orc):
    """Extreme int16 values (the hi/lo split edges) decode exactly."""
    import torch
    T, S = 32, 64
    acq = configs.linear_2d(T=T, S=S, generate=False)
    rng = np.random.default_rng(9)
    raw = rng.choice(np.array([-32768, -32767, -129, -128, -1, 0, 1, 127, 128, 255, 256, 32639, 32640, 32767],
                              np.int16), size=(T, T * S)).astype(np.int16)
    acq.raw = raw
    _rf_setup(acq)
    H = orc.hadamard(T, L.HADAMARD)
    want = orc.decode(orc.convert(raw, L.TYPE_I16, acq.channel_map, T * S, T, T, S), H).astype(np.float32)
    for mode in (1, 2):
        assert api.set_decode_mode(mode)
        out = torch.zeros(T * T * S, dtype=torch.float32, device=torch_dev)
        assert api.rf_convert_decode(_t(raw, torch_dev), L.TYPE_I16, out, 1, True)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy().reshape(T, T, S), want)
    assert api.set_decode_mode(0)


def test_convert_f32_and_i16(torch_dev, orc):
    import torch
    T, S = 8, 300
    acq = configs.linear_2d(T=T, S=S, encoding=L.NONE, seed=5)
    acq.bp.decode = L.NONE
    _rf_setup(acq)
    for dtype, arr in ((L.TYPE_I16, acq.raw), (L.TYPE_F32, acq.raw.astype(np.float32) * 0.37)):
        out = torch.zeros(T * T * S, dtype=torch.float32, device=torch_dev)
        assert api.rf_convert_decode(_t(arr, torch_dev), dtype, out, 1, False)
        torch.cuda.synchronize()
        want = orc.convert(arr, dtype, acq.channel_map, T * S, T, T, S)
        assert np.array_equal(out.cpu().numpy().reshape(T, T, S), want.astype(np.float32))


@pytest.mark.parametrize("S,taps", [(4096, 0), (4096, 64), (1024, 0), (1024, 17), (2048, 5), (256, 0),
                                    (64, 3), (32, 0), (16, 1),
                                    # lengths that are not powers of two: chirp-z path
                                    (4095, 0), (3000, 64), (2049, 5), (1234, 17), (1000, 0), (100, 7),
                                    (17, 3), (12, 0), (3, 0), (1, 0)])
@pytest.mark.parametrize("vpt", [16, 8])
def test_hilbert_matches_oracle(torch_dev, orc, S, taps, vpt):
    """Both FFT tiles (8 or 16 values per thread) against the oracle."""
    import torch
    print(f"hilbert S={S} taps={taps} fft tile {vpt}", flush=True)
    assert api.set_fft_tile(vpt)
    T, Cc = 4, 3
    acq = configs.linear_2d(T=T, S=S, generate=False)
    acq.bp.dec_data_dim[1] = Cc
    assert api.rf_init((T * S, Cc), (S, Cc, T), L.HADAMARD)
    rng = np.random.default_rng(S + taps)
    x = rng.standard_normal((T * Cc, S)).astype(np.float32) * 1000
    h = rng.standard_normal(taps).astype(np.float32) if taps else None
    assert api.rf_set_match_filter(h)
    out = torch.zeros(T * Cc * S * 2, dtype=torch.float32, device=torch_dev)
    assert api.rf_hilbert(_t(x, torch_dev), 1, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.complex64).reshape(T * Cc, S)
    want = orc.hilbert(x.astype(np.float64), None if h is None else h.astype(np.float64))
    assert rel_l2(got, want) < 1e-5
    # strided input: real parts of a complex buffer
    xc = np.zeros((T * Cc, S), np.complex64)
    xc.real = x
    xc.imag = 123.0
    out2 = torch.zeros_like(out)
    assert api.rf_hilbert(_t(xc.view(np.float32), torch_dev), 2, out2)
    torch.cuda.synchronize()
    assert api.set_fft_tile(0)
    assert np.array_equal(out2.cpu().numpy(), out.cpu().numpy())
    assert api.rf_set_match_filter(None)


@pytest.mark.parametrize("S,taps", [(4096, 64), (2048, 0), (256, 5), (16, 3)])
def test_hilbert_odd_row_count(torch_dev, orc, S, taps):
    """Power-of-two rows are transformed in pairs: with an odd row count the
    last row pairs with zeros (and, for short rows, the last workgroup holds
    fewer pairs than it has room for)."""
    import torch
    print(f"hilbert odd rows S={S} taps={taps}", flush=True)
    T, Cc = 3, 5
    assert api.rf_init((T * S, Cc), (S, Cc, T), L.NONE)  # (no Hadamard matrix of order 3)
    rng = np.random.default_rng(S + 3 * taps + 1)
    # rows of very different scale: each pair shares one transform
    x = (rng.standard_normal((T * Cc, S)) * np.geomspace(1.0, 3000.0, T * Cc)[:, None]).astype(np.float32)
    h = rng.standard_normal(taps).astype(np.float32) if taps else None
    assert api.rf_set_match_filter(h)
    out = torch.full((T * Cc * S * 2,), np.nan, dtype=torch.float32, device=torch_dev)
    assert api.rf_hilbert(_t(x, torch_dev), 1, out)
    torch.cuda.synchronize()
    assert api.rf_set_match_filter(None)
    got = out.cpu().numpy().view(np.complex64).reshape(T * Cc, S)
    want = orc.hilbert(x.astype(np.float64), None if h is None else h.astype(np.float64))
    assert np.all(np.isfinite(got))
    assert rel_l2(got, want) < 1e-5
    # every row on its own scale (a row paired with a 3000x larger one included)
    worst = max(rel_l2(got[r], want[r]) for r in range(T * Cc))
    print(f"  worst row rel L2 {worst:.2e}")
    assert worst < 1e-4


@pytest.mark.parametrize("taps", [0, 64])
def test_hilbert_walk_many_pairs(torch_dev, orc, taps):
    """4096-point rows take a fixed grid of workgroups (two per CU) that walk
    the row pairs, each requesting the next pair's rows while it transforms
    the current one.  1517 rows = 759 pairs: workgroups walk one or two pairs
    and the last pair holds a lone row.  FP32 against the oracle, strided
    input bit-equal to contiguous, FP16 / BF16 output against the rounding of
    the FP32 output."""
    import torch
    S, T, Cc = 4096, 37, 41
    print(f"hilbert walk S={S} rows={T * Cc} taps={taps}", flush=True)
    assert api.rf_init((T * S, Cc), (S, Cc, T), L.NONE)
    rng = np.random.default_rng(41 + taps)
    x = (rng.standard_normal((T * Cc, S)) * np.geomspace(1.0, 300.0, T * Cc)[:, None]).astype(np.float32)
    h = rng.standard_normal(taps).astype(np.float32) if taps else None
    assert api.rf_set_match_filter(h)
    try:
        d_x = _t(x, torch_dev)
        out = torch.full((T * Cc * S * 2,), np.nan, dtype=torch.float32, device=torch_dev)
        assert api.rf_hilbert(d_x, 1, out)
        torch.cuda.synchronize()
        got = out.cpu().numpy().view(np.complex64).reshape(T * Cc, S)
        want = orc.hilbert(x.astype(np.float64), None if h is None else h.astype(np.float64))
        assert np.all(np.isfinite(got))
        err = rel_l2(got, want)
        worst = max(rel_l2(got[r], want[r]) for r in range(T * Cc))
        print(f"  rel L2 {err:.2e}, worst row {worst:.2e}")
        assert err < 1e-5
        assert worst < 1e-4
        xc = np.zeros((T * Cc, S), np.complex64)
        xc.real = x
        xc.imag = -7.0
        out2 = torch.zeros_like(out)
        assert api.rf_hilbert(_t(xc.view(np.float32), torch_dev), 2, out2)
        torch.cuda.synchronize()
        assert np.array_equal(out2.cpu().numpy(), out.cpu().numpy())
        g32 = out.cpu().numpy().reshape(-1, 2)
        for fmt in (1, 2):
            words = torch.zeros(T * Cc * S, dtype=torch.int32, device=torch_dev)
            ok, scale = api.rf_hilbert_format(d_x, 1, words, fmt)
            assert ok
            torch.cuda.synchronize()
            w = words.cpu().numpy().view(np.uint32)
            if fmt == 1:
                v = (g32 * np.float32(scale)).astype(np.float16).view(np.uint16).astype(np.uint32)
            else:
                u = g32.view(np.uint32)
                v = (u + np.uint32(0x7FFF) + ((u >> 16) & np.uint32(1))) >> 16  # nearest-even BF16
            want_w = v[:, 0] | (v[:, 1] << 16)
            same = float(np.mean(w == want_w))
            print(f"  format {fmt}: {same:.6f} of the words equal the FP32 output's rounding")
            assert same > 0.999
    finally:
        assert api.rf_set_match_filter(None)


@pytest.mark.parametrize("rows", [1, 2, 3, 1026])
def test_hilbert_walk_row_counts(torch_dev, orc, rows):
    """4096-point rows at the edges of the row-pair walk: one row (a lone row
    with its zero partner), two, three, and 513 pairs (one more than a
    512-workgroup grid holds, so the first workgroup walks two pairs while
    the rest take one); 64-tap filter, FP32 output against the oracle."""
    import torch
    S = 4096
    print(f"hilbert walk rows={rows}", flush=True)
    assert api.rf_init((rows * S, 1), (S, 1, rows), L.NONE)
    rng = np.random.default_rng(rows + 7)
    x = (rng.standard_normal((rows, S)) * 100).astype(np.float32)
    h = rng.standard_normal(64).astype(np.float32)
    assert api.rf_set_match_filter(h)
    try:
        out = torch.full((rows * S * 2 + 64,), np.nan, dtype=torch.float32, device=torch_dev)
        assert api.rf_hilbert(_t(x, torch_dev), 1, out)
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        assert np.all(np.isnan(o[rows * S * 2:]))  # nothing written past the last row
        got = o[: rows * S * 2].view(np.complex64).reshape(rows, S)
        want = orc.hilbert(x.astype(np.float64), h.astype(np.float64))
        assert np.all(np.isfinite(got))
        assert rel_l2(got, want) < 1e-5
        assert max(rel_l2(got[r], want[r]) for r in range(rows)) < 1e-4
    finally:
        assert api.rf_set_match_filter(None)


@pytest.mark.parametrize("dtype", [L.TYPE_I16, L.TYPE_F32])
def test_beamform_decode_none_many_rows(lib, orc, dtype):
    """decode = NONE at 4096 samples: the fused convert + Hilbert reads the
    int16 (or float32) rows through the channel map while walking several row
    pairs per workgroup; 47 x 47 = 2209 rows (odd)."""
    print(f"beamform decode none, 47 x 47 rows of 4096, data type {dtype}", flush=True)
    acq = configs.rca_small(seed=13, T=47, S=4096, encoding=L.NONE)
    acq.bp.decode = L.NONE
    if dtype == L.TYPE_F32:
        acq.raw = (acq.raw.astype(np.float32) * np.float32(0.37)).astype(np.float32)
        acq.bp.data_type = L.TYPE_F32
    got = api.beamform_volume(acq.raw, acq.bp)
    assert got is not None
    err = rel_l2(got, orc.pipeline(acq))
    print(f"  rel L2 {err:.2e}")
    assert err < 1e-4


def test_beamform_decode_none_odd_rows(lib, orc):
    """decode = NONE takes the fused convert + Hilbert kernel; 5 x 5 rows."""
    print("beamform decode none, 25 rows", flush=True)
    assert api.set_das_options(cubic=True, cf_hercules=False)
    acq = configs.rca_small(seed=5, T=5, encoding=L.NONE)
    acq.bp.decode = L.NONE
    got = api.beamform_volume(acq.raw, acq.bp)
    assert got is not None
    assert rel_l2(got, orc.pipeline(acq)) < 1e-4


def _das_case(acq, orc, torch_dev, cubic, tol, cf=False):
    import torch
    st = orc.pipeline(acq, cubic=cubic, return_stages=True)
    iq = st["iq"].astype(np.complex64)
    # reference volume from the same (float-rounded) IQ
    want, active = orc.das(iq.astype(np.complex128), acq.bp, cubic=cubic, cf_hercules=cf)
    assert api.set_das_options(cubic=cubic, cf_hercules=cf)
    vol = torch.zeros(want.size * 2, dtype=torch.float32, device=torch_dev)
    assert api.das_device(_t(iq.view(np.float32), torch_dev), vol, acq.bp)
    torch.cuda.synchronize()
    got = vol.cpu().numpy().view(np.complex64).reshape(want.shape)
    assert api.set_das_options(cubic=True, cf_hercules=False)
    err = rel_l2(got, want)
    print(f"  rel L2 {err:.3e}, active taps {active}")
    assert err < tol
    return got, want


@pytest.mark.parametrize("cubic", [True, False])
def test_das_hercules_matches_oracle(torch_dev, orc, cubic):
    print(f"das hercules cubic={cubic}", flush=True)
    acq = configs.rca_small(seed=21)
    _das_case(acq, orc, torch_dev, cubic, 1e-4)


def test_das_hercules_focused_and_ragged_grid(torch_dev, orc):
    print("das hercules focus ragged", flush=True)
    acq = configs.rca_small(seed=22, n=19, focus_z=8e-3)
    acq.bp.output_points[:] = (19, 13, 21, 1)
    _das_case(acq, orc, torch_dev, True, 1e-4)


def test_das_hercules_coherence_factor_extension(torch_dev, orc):
    print("das hercules cf", flush=True)
    acq = configs.rca_small(seed=24)
    acq.bp.coherency_weighting = 0.7
    _das_case(acq, orc, torch_dev, True, 1e-4, cf=True)


@pytest.mark.parametrize("cubic", [False, True])
@pytest.mark.parametrize("focus", [False, True])
def test_das_hercules_coherence_factor_loops(torch_dev, orc, cubic, focus):
    """The coherence factor in the classified, pipelined tap loops (wave-channels
    wholly inside the aperture and partly inside, the incoherent sum over the
    taps inside only): linear and cubic, plane and focused transmits, on a
    ragged grid."""
    print(f"das hercules cf loops cubic={cubic} focus={focus}", flush=True)
    acq = configs.rca_small(seed=31, n=21, **({"focus_z": 8e-3} if focus else {}))
    acq.bp.output_points[:] = (21, 17, 23, 1)
    acq.bp.coherency_weighting = 0.6
    _das_case(acq, orc, torch_dev, cubic, 1e-4, cf=True)


@pytest.mark.parametrize("cubic", [False, True])
def test_das_hercules_every_transmit_split(torch_dev, orc, tmp_path, cubic):
    """Every transmit split of the HERCULES launch (1, 2, 4, 8 and 16 workgroups
    per tile, forced through the tuning table for this shape) against the
    oracle, and all within 1e-6 of one another (only the order in which the
    partial sums are added differs)."""
    import torch
    print(f"das hercules every split cubic={cubic}", flush=True)
    acq = configs.rca_small(seed=29, T=128, n=16)
    bp = acq.bp
    Nx, Ny, Nz = bp.output_points[0], bp.output_points[1], bp.output_points[2]
    S, Cc, T = bp.dec_data_dim[0], bp.dec_data_dim[1], bp.dec_data_dim[2]
    iq = orc.pipeline(acq, cubic=cubic, return_stages=True)["iq"].astype(np.complex64)
    want, _ = orc.das(iq.astype(np.complex128), bp, cubic=cubic)
    assert api.set_das_options(cubic=cubic, cf_hercules=False)
    d_iq = _t(iq.view(np.float32), torch_dev)
    vols = {}
    try:
        for sp in (1, 2, 4, 8, 16):
            path = tmp_path / f"split{sp}.txt"
            # table 0 (HERCULES split), key Nx Ny Nz T C S G fmt cubic cf focus generic
            path.write_text(f"0 {Nx} {Ny} {Nz} {T} {Cc} {S} 1 0 {int(cubic)} 0 0 0 {sp}\n")
            assert api.tuning_load(path)
            vol = torch.zeros(want.size * 2, dtype=torch.float32, device=torch_dev)
            assert api.das_device(d_iq, vol, bp)
            torch.cuda.synchronize()
            vols[sp] = vol.cpu().numpy().view(np.complex64).reshape(want.shape)
            err = rel_l2(vols[sp], want)
            print(f"  split {sp}: rel L2 {err:.3e}")
            assert err < 1e-4
    finally:
        assert api.set_das_options(cubic=True, cf_hercules=False)
    for sp in (2, 4, 8, 16):
        assert rel_l2(vols[sp], vols[1]) < 1e-6


def test_das_hercules_single_tile_no_split(torch_dev, orc):
    """A grid of one tile with few transmits takes the unsplit path."""
    print("das hercules unsplit", flush=True)
    acq = configs.rca_small(seed=25, T=16, n=8)
    _das_case(acq, orc, torch_dev, True, 1e-4)


@pytest.mark.parametrize("cubic", [False, True])
@pytest.mark.parametrize("nz", [3, 12])
def test_das_hercules_long_windows(torch_dev, orc, cubic, nz):
    """Coarse depth grids over 3-14 mm: with 3 planes every window is longer
    than the LDS limit (all channels read HBM directly); with 12 the windows
    fit one by one but a pass's windows overflow the LDS pool, so some
    channels are staged and the rest read directly."""
    print(f"das hercules long windows nz={nz} cubic={cubic}", flush=True)
    acq = configs.rca_small(seed=26)
    acq.bp.output_points[2] = nz
    _das_case(acq, orc, torch_dev, cubic, 1e-4)


@pytest.mark.parametrize("cubic", [False, True])
@pytest.mark.parametrize("fn", [1.0, 0.0])
def test_das_hercules_above_array(torch_dev, orc, cubic, fn):
    """Depths from -3.125 to 8.875 mm: above the array the reference's Hann
    weight cos^2(pi min(fn rho / 2z, 1/2)) oscillates in |a| instead of
    falling to zero; such volumes take the exact per-tap weight.

    The weight jumps from 0.1 to 0 at the threshold, so a tap whose a lies
    within FP32 rounding of it (~1e-7 relative) is included by one
    arithmetic and not the other, moving its voxel by 0.1 x a sample.  This
    grid keeps every tap at least 2e-6 (relative) from the threshold (FP64
    scan of all taps; the -3..9 mm grid has taps 5e-9 from it, and one such
    tap put a single voxel 1.3 % off)."""
    print(f"das hercules above array fn={fn} cubic={cubic}", flush=True)
    acq = configs.rca_small(seed=27, z_range=(-3.125e-3, 8.875e-3))
    acq.bp.fn_rx = fn
    _das_case(acq, orc, torch_dev, cubic, 1e-4)


def test_das_hercules_above_array_cf_and_negative_fnumber(torch_dev, orc):
    print("das hercules above array cf / negative F-number", flush=True)
    # the grid of test_das_hercules_above_array (no tap within 8e-7 of the
    # apodization threshold for F-number 1 or -1)
    acq = configs.rca_small(seed=28, z_range=(-3.125e-3, 8.875e-3))
    acq.bp.coherency_weighting = 0.5
    _das_case(acq, orc, torch_dev, False, 1e-4, cf=True)
    # a negative F-number below the array (3-14 mm, |a| < 2.5): near z = 0 the
    # negative ratio a = fn rho / 2z is unbounded and cos^2(pi a) is beyond FP32
    # (on the grid above, a depth of 5 um makes |a| ~ 1300)
    acq = configs.rca_small(seed=29)
    acq.bp.fn_rx = -1.0
    _das_case(acq, orc, torch_dev, True, 1e-4)


@pytest.mark.parametrize("cubic,cf,focus", [(True, False, None), (False, False, None), (True, True, None),
                                            (False, True, None), (True, False, 8e-3), (False, True, 8e-3)])
def test_das_hercules_sin_apodization(torch_dev, orc, cubic, cf, focus):
    print(f"das hercules RX_TO_SIN cubic={cubic} cf={cf} focus={focus}", flush=True)
    acq = configs.rca_small(seed=23) if focus is None else configs.rca_small(seed=25, focus_z=focus)
    acq.bp.apo_type = L.RX_TO_SIN
    acq.bp.to_power = 2
    if cf:
        acq.bp.coherency_weighting = 0.5
    _das_case(acq, orc, torch_dev, cubic, 1e-4, cf=cf)


def test_das_hercules_sin_apodization_above_array(torch_dev, orc):
    """Sine apodization over a grid reaching z <= 0: the weight depends on the
    element alone, so the volume takes the same channel loop as below the array."""
    print("das hercules RX_TO_SIN above array", flush=True)
    acq = configs.rca_small(seed=24, z_range=(-3.125e-3, 8.875e-3))
    acq.bp.apo_type = L.RX_TO_SIN
    acq.bp.to_power = 2
    _das_case(acq, orc, torch_dev, False, 1e-4)


@pytest.mark.parametrize("G,gid,enc", [(2, 1, L.HADAMARD), (4, 2, L.HADAMARD), (4, 3, L.WALSH)])
def test_das_hercules_readi_groups(torch_dev, orc, G, gid, enc):
    """READI: G groups of the transmit aperture, the group's sign row from the
    order-G matrix (templated_beamformer.cuh:43-63)."""
    print(f"das hercules READI G={G} id={gid} enc={enc}", flush=True)
    acq = configs.rca_small(seed=23)
    acq.bp.readi_group_count = G
    acq.bp.readi_group_id = gid
    acq.bp.decode = enc
    _das_case(acq, orc, torch_dev, cubic=True, tol=1e-4)


def test_das_forces_readi_groups(torch_dev, orc):
    print("das forces READI G=2", flush=True)
    acq = configs.linear_2d(seed=4)
    acq.bp.readi_group_count = 2
    acq.bp.readi_group_id = 1
    _das_case(acq, orc, torch_dev, cubic=False, tol=1e-4)


@pytest.mark.parametrize("cubic", [True, False])
def test_das_forces_matches_oracle(torch_dev, orc, cubic):
    print(f"das forces cubic={cubic}", flush=True)
    acq = configs.linear_2d(seed=31)
    _das_case(acq, orc, torch_dev, cubic, 1e-4)


def test_das_forces_apodized_with_cf(torch_dev, orc):
    print("das forces fn=1 cw=0.5", flush=True)
    acq = configs.linear_2d(seed=32)
    acq.bp.fn_tx = 1.0
    acq.bp.fn_rx = 1.0
    acq.bp.coherency_weighting = 0.5
    _das_case(acq, orc, torch_dev, True, 1e-4)


def test_beamform_api_linear_2d(lib, orc):
    print("beamform linear_2d", flush=True)
    acq = configs.linear_2d(seed=41)
    got = api.beamform_volume(acq.raw, acq.bp)
    assert got is not None
    want = orc.pipeline(acq)
    assert rel_l2(got, want) < 1e-4


def test_beamform_api_rca_small(lib, orc):
    print("beamform rca_small", flush=True)
    acq = configs.rca_small(seed=42)
    got = api.beamform_volume(acq.raw, acq.bp)
    assert got is not None
    want = orc.pipeline(acq)
    assert rel_l2(got, want) < 1e-4
    # decode = NONE takes the fused convert+Hilbert kernel
    acq2 = configs.rca_small(seed=43, encoding=L.NONE)
    acq2.bp.decode = L.NONE
    got2 = api.beamform_volume(acq2.raw, acq2.bp)
    assert rel_l2(got2, orc.pipeline(acq2)) < 1e-4


def test_beamform_rca_small_odd_sample_count(lib, orc):
    """Whole path with S = 1000 (chirp-z Hilbert) against the oracle."""
    print("rca_small S=1000", flush=True)
    acq = configs.rca_small(seed=11, S=1000)
    got = api.beamform_volume(acq.raw, acq.bp)
    assert got is not None
    assert rel_l2(got, orc.pipeline(acq)) < 1e-4


def test_beamform_matlab_shim(lib, orc):
    import ctypes as C
    acq = configs.linear_2d(seed=44)
    out = np.zeros(acq.volume_shape, np.complex64)
    lib.beamform_i16_matlab(acq.raw.ctypes.data_as(C.POINTER(C.c_int16)), acq.bp,
                            out.ctypes.data_as(C.POINTER(C.c_float)))
    assert rel_l2(out, orc.pipeline(acq)) < 1e-4


def test_gl_api_through_device_buffers(torch_dev, orc):
    """cuda_decode / cuda_hilbert on registered device buffers (the GL seam)."""
    import torch
    print("gl api seam", flush=True)
    T, S = 32, 1024
    acq = configs.linear_2d(T=T, S=S, seed=51)
    Cc = T
    # raw buffer with a 64-byte header in front of the frame
    raw = np.zeros(32 + acq.raw.size, np.int16)
    raw[32:] = acq.raw.ravel()
    d_raw = _t(raw, torch_dev)
    bufs = [torch.zeros(T * Cc * S * 2, dtype=torch.float32, device=torch_dev) for _ in range(2)]
    assert api.init_cuda_configuration((T * S, Cc), (S, Cc, T))
    assert api.register_device_buffers(bufs, d_raw)
    assert api.cuda_set_channel_mapping(acq.channel_map)
    h = configs.chirp_filter(16)
    assert api.cuda_set_match_filter(h)
    assert api.cuda_decode(64, 0)
    assert api.cuda_hilbert(0, 1)
    torch.cuda.synchronize()
    conv = orc.convert(acq.raw, L.TYPE_I16, acq.channel_map, T * S, T, Cc, S)
    dec = orc.decode(conv, orc.hadamard(T, L.HADAMARD))
    got_dec = bufs[0].cpu().numpy()[0::2].reshape(T, Cc, S)
    assert np.array_equal(got_dec, dec.astype(np.float32))
    got = bufs[1].cpu().numpy().view(np.complex64).reshape(T * Cc, S)
    want = orc.hilbert(dec.reshape(T * Cc, S), h.astype(np.float64))
    assert rel_l2(got, want) < 1e-5
    assert api.cuda_hilbert(0, 5) is False  # unregistered index
    api.deinit_cuda_configuration()
    assert api.cuda_decode(0, 0) is False


@pytest.mark.parametrize("cubic", [False, True])
def test_full_size_rca3d_depth_planes(torch_dev, orc, cubic):
    """Full 128 tx x 128 ch x 4096 frame with the 64-tap filter, beamformed on
    4 depth planes (iz = 16, 48, 80, 112 of the 128^3 grid) by the GPU and by
    the FP64 oracle: relative L2 within 1e-4."""
    import torch
    print(f"full-size rca_3d planes cubic={cubic}", flush=True)
    acq = configs.rca_3d(seed=2)
    bp = acq.bp
    nz = bp.output_points[2]
    zmin, zmax = bp.output_min_coordinate[2], bp.output_max_coordinate[2]
    res = np.float32((np.float32(zmax) - np.float32(zmin)) / np.float32(nz - 1))
    bp.output_points[2] = 4
    bp.output_min_coordinate[2] = float(np.float32(zmin) + np.float32(16) * res)
    bp.output_max_coordinate[2] = float(np.float32(zmin) + np.float32(112) * res)
    assert api.set_das_options(cubic=cubic)
    d_raw = _t(acq.raw, torch_dev)
    vol = torch.zeros(128 * 128 * 4 * 2, dtype=torch.float32, device=torch_dev)
    assert api.beamform_device(d_raw, vol, bp)
    torch.cuda.synchronize()
    assert api.set_das_options(cubic=True)
    got = vol.cpu().numpy().view(np.complex64).reshape(4, 128, 128)
    want = orc.pipeline(acq, cubic=cubic)
    err = rel_l2(got, want)
    print(f"  rel L2 {err:.3e}; max |err| / max |v| {np.abs(got - want).max() / np.abs(want).max():.3e}")
    assert err < 1e-4


@pytest.mark.parametrize("T,S", [(32, 512), (128, 256), (24, 200)])
def test_decode_f32_input(torch_dev, orc, T, S):
    """TYPE_F32 raw input with Hadamard decode (the FP32 MFMA GEMM path; the
    reference's cuBLAS SGEMM, hadamard_decoder.cpp:49-78).  Integer-valued
    samples decode bit-exactly for T = 2^n (every partial sum is an integer
    below 2^24); arbitrary floats stay within the error bound of any FP32
    summation order: |got - exact| <= u * sum_k |x_k| + u * |exact|, u = 2^-24
    (the T-term sum errs by at most (T-1) u sum |x_k|, 1/T scales it down)."""
    import torch
    print(f"decode f32 input T={T} S={S}", flush=True)
    acq = configs.linear_2d(T=T, S=S, seed=T * 3 + S)
    S_, Cc, T_ = _rf_setup(acq)
    H = orc.hadamard(T, L.HADAMARD)
    pow2 = (T & (T - 1)) == 0
    rng = np.random.default_rng(T + S)
    for kind in ("integer", "fraction"):
        raw = acq.raw.astype(np.float32)
        if kind == "fraction":
            raw = (raw * np.float32(0.731) + rng.standard_normal(raw.shape).astype(np.float32) * 37.0).astype(np.float32)
        conv = orc.convert(raw, L.TYPE_F32, acq.channel_map, T * S, T, Cc, S)
        want = orc.decode(conv, H)
        out = torch.full((T * Cc * S,), np.nan, dtype=torch.float32, device=torch_dev)
        assert api.rf_convert_decode(_t(raw, torch_dev), L.TYPE_F32, out, 1, True)
        torch.cuda.synchronize()
        got = out.cpu().numpy().reshape(T, Cc, S).astype(np.float64)
        if kind == "integer" and pow2:
            assert np.array_equal(got, want.astype(np.float32).astype(np.float64))
            continue
        u = 2.0 ** -24
        bound = u * np.einsum("jk,kcs->jcs", np.abs(H), np.abs(conv)) + u * np.abs(want)
        if not pow2:  # one more rounding: sum * fl(1/T)
            bound += u * np.abs(want)
        excess = np.abs(got - want) / bound
        print(f"  {kind}: worst error / bound {excess.max():.3f}")
        assert np.all(np.abs(got - want) <= bound)


def test_beamform_f32_matlab_shim(lib, orc):
    """beamform_f32_matlab (cuda_toolkit_matlab.cpp:24-37): float32 raw frame
    with Hadamard decode, the whole path against the oracle."""
    import ctypes as C
    print("beamform f32 matlab shim", flush=True)
    for make in (lambda: configs.linear_2d(seed=45), lambda: configs.rca_small(seed=46)):
        acq = make()
        acq.raw = (acq.raw.astype(np.float32) * np.float32(0.37)).astype(np.float32)
        acq.bp.data_type = L.TYPE_F32
        out = np.zeros(acq.volume_shape, np.complex64)
        lib.beamform_f32_matlab(acq.raw.ctypes.data_as(C.POINTER(C.c_float)), acq.bp,
                                out.ctypes.data_as(C.POINTER(C.c_float)))
        err = rel_l2(out, orc.pipeline(acq))
        print(f"  rel L2 {err:.3e}")
        assert err < 1e-4


@pytest.mark.parametrize("S,taps,rows", [(4097, 0, 12), (5000, 64, 12), (6000, 0, 7), (8192, 64, 12), (8192, 0, 5),
                                         (12289, 17, 6), (16384, 64, 6), (32768, 0, 3), (65536, 64, 3),
                                         (32767, 3, 2),
                                         # 1030 rows of the 65536-point chirp-z length: two workspace chunks
                                         (20000, 5, 1030)])
def test_hilbert_long_rows(torch_dev, orc, S, taps, rows):
    """Rows longer than 4096 samples (the reference plans cuFFT for any length,
    hilbert_handler.cu:30-53): N = S (powers of two) or the chirp-z length as
    N/4096 x 4096 steps through HBM, against the FP64 oracle."""
    import torch
    print(f"hilbert long S={S} taps={taps} rows={rows}", flush=True)
    assert api.rf_init((rows * S, 1), (S, 1, rows), L.NONE)
    rng = np.random.default_rng(S + taps)
    x = rng.standard_normal((rows, S)).astype(np.float32) * 1000
    h = rng.standard_normal(taps).astype(np.float32) if taps else None
    assert api.rf_set_match_filter(h)
    out = torch.full((rows * S * 2,), np.nan, dtype=torch.float32, device=torch_dev)
    assert api.rf_hilbert(_t(x, torch_dev), 1, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.complex64).reshape(rows, S)
    want = orc.hilbert(x.astype(np.float64), None if h is None else h.astype(np.float64))
    err = rel_l2(got, want)
    worst = max(rel_l2(got[r], want[r]) for r in range(min(rows, 16)))
    print(f"  rel L2 {err:.2e}, worst row {worst:.2e}")
    assert np.all(np.isfinite(got))
    assert err < 1e-5 and worst < 1e-5
    if rows <= 12:  # strided input: real parts of a complex buffer
        xc = np.zeros((rows, S), np.complex64)
        xc.real = x
        xc.imag = -5.0
        out2 = torch.zeros_like(out)
        assert api.rf_hilbert(_t(xc.view(np.float32), torch_dev), 2, out2)
        torch.cuda.synchronize()
        assert np.array_equal(out2.cpu().numpy(), out.cpu().numpy())
    assert api.rf_set_match_filter(None)


@pytest.mark.parametrize("S", [8192, 6000])
@pytest.mark.parametrize("fmt,tol", [(1, 3e-4), (2, 2.5e-3)])
def test_hilbert_long_rows_staged_formats(torch_dev, orc, S, fmt, tol):
    """FP16 (x 2^-k) and BF16 IQ from long rows (the workspace path for powers
    of two too), against the oracle at the staged formats' bars."""
    import torch
    print(f"hilbert long staged S={S} fmt={fmt}", flush=True)
    rows = 6
    assert api.rf_init((rows * S, 1), (S, 1, rows), L.NONE)
    rng = np.random.default_rng(S + fmt)
    x = np.round(rng.standard_normal((rows, S)) * 3000).astype(np.float32)
    h = configs.chirp_filter(64)
    assert api.rf_set_match_filter(h)
    out = torch.zeros(rows * S, dtype=torch.int32, device=torch_dev)
    ok, scale = api.rf_hilbert_format(_t(x, torch_dev), 1, out, fmt)
    assert ok
    torch.cuda.synchronize()
    assert api.rf_set_match_filter(None)
    raw = out.cpu().numpy().view(np.uint32)
    if fmt == 1:
        got = raw.view(np.float16).astype(np.float32).reshape(rows, S, 2) / scale
    else:
        got = np.stack([(raw << 16).view(np.float32), (raw & 0xFFFF0000).view(np.float32)], -1).reshape(rows, S, 2)
    got = got[..., 0] + 1j * got[..., 1]
    want = orc.hilbert(x.astype(np.float64), h.astype(np.float64))
    err = rel_l2(got, want)
    print(f"  rel L2 {err:.2e}")
    assert err < tol


@pytest.mark.parametrize("S,encoding", [(6000, L.HADAMARD), (8192, L.HADAMARD), (8192, L.NONE), (5001, L.NONE)])
def test_beamform_long_rows(lib, orc, S, encoding):
    """The whole path with rows longer than 4096 samples (decode = NONE takes
    the convert-fused first step) against the oracle."""
    print(f"beamform rca_small S={S} encoding={encoding}", flush=True)
    acq = configs.rca_small(seed=12, S=S, encoding=encoding)
    acq.bp.decode = encoding
    got = api.beamform_volume(acq.raw, acq.bp)
    assert got is not None
    err = rel_l2(got, orc.pipeline(acq))
    print(f"  rel L2 {err:.2e}")
    assert err < 1e-4
