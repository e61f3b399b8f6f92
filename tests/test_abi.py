"""C-ABI checks that need no GPU: struct layout, exported symbols, host-side
validation (every error path beamform() takes before touching the device)."""
import ctypes as C
import subprocess

import numpy as np
import pytest

from cuda_ultrasound_toolkit_amd import _lib as L, api, configs


def test_struct_layout_matches_reference():
    BP = L.CudaBeamformerParameters
    assert C.sizeof(BP) == 8948 and C.alignment(BP) == 4
    off = {n: getattr(BP, n).offset for n, _ in BP._fields_}
    assert off["rf_raw_dim"] == 72 and off["decode"] == 104 and off["interpolate"] == 192
    assert off["coherency_weighting"] == 196 and off["channel_mapping"] == 200
    assert off["foci"] == 1224 and off["tx_orientation"] == 4296 and off["filter_length"] == 4832
    assert off["rf_filter"] == 4836 and off["sample_type"] == 8944
    assert C.sizeof(L.NccMotionParameters) == 72
    assert L.NccMotionParameters.use_subpixel.offset == 44
    assert L.NccMotionParameters.image_dims.offset == 60
    assert C.sizeof(L.CorrImagesParameters) == 16


def test_library_exports_every_declared_symbol(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], capture_output=True, text=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for sym in L.exported_symbols():
        assert sym in names, sym
        getattr(lib, sym)


def test_header_declarations_are_exported():
    """Every function declared in include/*.h is in the library."""
    import re
    from pathlib import Path
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], capture_output=True, text=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    inc = Path(__file__).resolve().parent.parent / "include"
    pending = set()
    for h in inc.glob("*.h"):
        for m in re.finditer(r"EXPORT\w*\s+[\w\s\*]+?\b(\w+)\s*\(", h.read_text()):
            if m.group(1) in pending:
                continue
            assert m.group(1) in names, (h.name, m.group(1))


def _expect_false(acq, mutate, out_elems=None, resize_raw=True):
    bp = acq.bp
    mutate(bp)
    n = bp.output_points[0] * bp.output_points[1] * bp.output_points[2]
    out = np.zeros(n if out_elems is None else out_elems, np.complex64)
    # raw input sized to whatever rf_raw_dim now says, so that only the
    # parameter under test is wrong (raw_short keeps the original array)
    raw = np.zeros((bp.rf_raw_dim[1], bp.rf_raw_dim[0]), np.int16) if resize_raw else acq.raw
    return api.beamform(raw, out, bp)


# Each case must fail for its own reason, checked by the message on stderr:
# with no GPU every call returns false, so the flag alone would pin nothing.
REJECTIONS = {
    "sample_type": (lambda b: setattr(b, "sample_type", L.SAMPLE_BANDPASS), "takes no transmit decoding"),
    "data_type": (lambda b: setattr(b, "data_type", L.TYPE_U8), "Unsupported input data type"),
    "bad_order": (lambda b: (b.dec_data_dim.__setitem__(2, 36), b.rf_raw_dim.__setitem__(0, 36 * 1024)),
                  "Size '36' not supported"),
    "map_range": (lambda b: b.channel_mapping.__setitem__(3, 999), "entry 3 = 999 outside"),
    "filter_long": (lambda b: setattr(b, "filter_length", 1025), "Filter length 1025 exceeds 1024"),
    "sequence": (lambda b: setattr(b, "das_shader_id", L.RCA_VLS), "Unsupported sequence ID 3"),
    "herc_focus": (lambda b: setattr(b, "tx_orientation", L.ORIENT_NONE), "Invalid focus for HERCULES"),
    "zero_dims": (lambda b: b.dec_data_dim.__setitem__(0, 0), "Invalid dimensions for RF raw data"),
    "raw_short": (lambda b: b.rf_raw_dim.__setitem__(1, 64), "rf_raw_dim needs"),
    "out_short": (lambda b: None, "Output span holds 80 bytes"),
    # rows up to 65536 samples (powers of two) or 32768 (other lengths) are supported
    "fft_len": (lambda b: (b.dec_data_dim.__setitem__(0, 40000), b.rf_raw_dim.__setitem__(0, 32 * 40000)),
                "Sample count 40000 is not supported"),
    "fft_len_pow2": (lambda b: (b.dec_data_dim.__setitem__(0, 131072), b.rf_raw_dim.__setitem__(0, 32 * 131072)),
                     "Sample count 131072 is not supported"),
}


@pytest.mark.parametrize("case", sorted(REJECTIONS))
def test_beamform_rejects_invalid_parameters(lib, capfd, case):
    acq = configs.rca_small(generate=False)
    mut, message = REJECTIONS[case]
    out_elems = 10 if case == "out_short" else None
    assert _expect_false(acq, mut, out_elems, resize_raw=case != "raw_short") is False
    err = capfd.readouterr().err
    assert message in err, err


def test_motion_detection_rejects_invalid_parameters(lib, capfd):
    imgs = np.zeros((3, 40, 50), np.float32)
    maps = np.zeros((3, 4, 5, 4), np.float32)
    ok = api.motion_params(50, 40, 3, (5, 4), 8, 8, (4, 4))
    bad = [
        (dict(frame_count=0), "empty images"),
        (dict(reference_frame=3), "reference frame out of range"),
        (dict(patch_size=0), "empty motion grid or patch"),
    ]
    for b, message in bad:
        p = api.motion_params(50, 40, 3, (5, 4), 8, 8, (4, 4))
        for k, v in b.items():
            setattr(p, k, v)
        assert api.motion_detection(imgs, maps, p) is False, b
        assert message in capfd.readouterr().err, b
    assert api.motion_detection(imgs[:1], maps, ok) is False           # image span too small
    assert "image span holds 8000 bytes" in capfd.readouterr().err
    assert api.motion_detection(imgs, maps[:1], ok) is False           # map span too small
    assert "motion map span holds 320 bytes" in capfd.readouterr().err
    # margins whose sign bit is set (uint fields) are refused, any search size is not
    p = api.motion_params(50, 40, 3, (5, 4), 8, 8, (4, 4))
    p.search_margins[0] = 0x80000000
    assert api.motion_detection(imgs, maps, p) is False
    assert "negative search margins" in capfd.readouterr().err


def test_corr_images_rejects_invalid_dims(lib, capfd):
    t = np.zeros((5, 5), np.float32)
    s = np.zeros((4, 9), np.float32)
    assert api.corr_images(t, s, np.zeros((1, 5), np.float32)) is False  # template taller than source
    assert "Template dimensions must be <= source dimensions" in capfd.readouterr().err
    assert api.corr_images(np.zeros((0, 3), np.float32), s, np.zeros((1, 1), np.float32)) is False
    assert "Invalid image dimensions" in capfd.readouterr().err
    assert api.corr_images(np.zeros((2, 3), np.float32), s, np.zeros((1, 1), np.float32)) is False  # map too small
    assert "corr_map span is too small" in capfd.readouterr().err


def test_c_api_call_order_errors(lib, capfd):
    # decode / hilbert before buffers are registered
    assert api.cuda_decode(0, 0) is False
    assert "OGL buffers not registered" in capfd.readouterr().err
    assert api.cuda_hilbert(0, 1) is False
    assert "OGL buffers not registered" in capfd.readouterr().err
    assert api.register_cuda_buffers(list(range(1, 18)), 99) is False  # > 16 buffers
    assert "Too many RF data buffers" in capfd.readouterr().err


def test_tuning_table_text_and_fft_tile_values(lib, capfd, tmp_path):
    """The tuning table's text form (no GPU needed): load, save, reject
    malformed lines; the FFT tile accepts 0, 8 and 16 only."""
    from cuda_ultrasound_toolkit_amd import api
    good = tmp_path / "good.txt"
    good.write_text("2 4096 16384 1 1 1 0 0 0 0 0 0 0 8\n0 128 128 128 128 128 4096 1 0 0 0 0 0 16\n")
    assert api.tuning_load(good)
    out = tmp_path / "out.txt"
    assert api.tuning_save(out)
    lines = set(out.read_text().splitlines())
    assert set(good.read_text().splitlines()) <= lines
    bad = tmp_path / "bad.txt"
    bad.write_text("7 1 2 3\n")
    assert api.tuning_load(bad) is False
    assert "malformed line 1" in capfd.readouterr().err
    assert api.tuning_load(tmp_path / "missing.txt") is False
    assert "cannot read" in capfd.readouterr().err
    for v in (0, 8, 16):
        assert api.set_fft_tile(v)
    assert api.set_fft_tile(4) is False
    assert "FFT tile must be 0" in capfd.readouterr().err
    assert api.set_fft_tile(0)
