"""Host-side logic without a GPU: committed network weights, workload
generators, parameter packing shared by the library and the oracle."""
import math

import numpy as np
import pytest

from cuda_ultrasound_toolkit_amd import _lib as L, api, configs, motion_data
from cuda_ultrasound_toolkit_amd import motion_learned as ml


def test_committed_weights_load_safely(orc):
    p, meta = ml.load_params()
    assert p.dtype == np.float32 and p.shape == (api.pnet_param_count(),)
    assert api.pnet_param_count() == orc.pnet_param_count()
    assert np.all(np.isfinite(p))
    assert 0.6 <= meta["threshold"] <= 0.9
    assert meta["input_mode"] == ml.INPUT_MODE
    ev = meta["eval_by_threshold"][f"{meta['threshold']:.2f}"]
    assert ev["agreement"] >= 0.99 and ev["fallback_rate"] < 1.0
    # the recorded baseline (proposal 0) falls back on nearly every block
    assert meta["baseline_zero_proposal"]["fallback_rate"] > 0.95


def test_init_params_layout(orc):
    p = ml.init_params(0)
    n = orc.pnet_param_count()
    assert p.shape == (n,)
    o = 0
    for ci, co in ml.LAYERS:
        w = p[o:o + co * ci * 27]
        assert abs(w.std() - math.sqrt(2.0 / (ci * 27))) < 0.2 * math.sqrt(2.0 / (ci * 27))
        o += co * ci * 27
        assert np.all(p[o:o + co] == 0)
        o += co
    assert o + 771 == n


def test_uneven_frames_ranges():
    fr = configs.uneven_frames(seed=4, n=16, generate=False)
    for a in fr:
        S, C, T = a.bp.dec_data_dim[0], a.bp.dec_data_dim[1], a.bp.dec_data_dim[2]
        assert T in (16, 32, 64, 128) and C == 128 and 1024 <= S <= 4096
        n = a.bp.output_points
        assert n[0] == n[1] == n[2] and 64 <= n[0] <= 160
        # the imaged depth stays within the recorded round trip
        zmax = a.bp.output_max_coordinate[2]
        assert 2 * zmax / 1540.0 * 20e6 <= S
    again = configs.uneven_frames(seed=4, n=16, generate=False)
    assert [a.bp.dec_data_dim[:3] for a in fr] == [b.bp.dec_data_dim[:3] for b in again]


def test_motion_voxel_pitch():
    a = configs.rca_3d(generate=False)
    px, py, pz = motion_data.voxel_pitch(a.bp)
    assert px == py == pytest.approx(38e-3 / 127)
    assert pz == pytest.approx(55e-3 / 127)
    s = motion_data.random_shifts(np.random.default_rng(0), 1000)
    assert s.shape == (1000, 3) and np.abs(s).max() <= 6.0


def test_library_exports_batch_and_motion_entry_points(lib):
    for name in ("cut_beamform_batch", "cut_set_autotune", "cut_motion3d_learned", "cut_pnet_infer",
                 "cut_rf_hilbert_format", "cut_das_device_format"):
        assert name in L.exported_symbols()
        assert hasattr(lib, name)
