"""Out-of-process round trips through cut_transfer_server and the C client
library against the in-process calls of the same entry points (same library,
same inputs).  The DAS launch shape is tuned by timing once per process, so
the server may sum the beamformer's terms in another order than the test
process: volumes are compared to 1e-6 relative L2; motion maps and
correlation maps (no tuned launches) bit for bit."""
import os

import numpy as np
import pytest

from cuda_ultrasound_toolkit_amd import _lib as L, api, configs, transfer as T

from test_gpu_ncc import frames_with_shifts, speckle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def server():
    name = f"cut_g{os.getpid()}"
    print(f"transfer server {name}", flush=True)
    srv = T.start_server(name, data_bytes=64 << 20, idle_timeout_ms=120000)
    cl = T.Client(name, timeout_ms=120000)
    yield cl
    cl.deinit()
    assert T.stop_server(srv) == 0


def test_transfer_beamform_i16_and_f32(server, lib):
    print("transfer beamform rca_small", flush=True)
    acq = configs.rca_small(seed=31)
    want = api.beamform_volume(acq.raw, acq.bp)
    got = server.beamform(acq.raw, acq.bp)
    assert got is not None and server.last_ok
    assert np.linalg.norm(got - want) <= 1e-6 * np.linalg.norm(want)
    print("transfer beamform linear_2d f32", flush=True)
    acq = configs.linear_2d(seed=32)
    acq.bp.data_type = L.TYPE_F32
    raw = acq.raw.astype(np.float32)
    want = api.beamform_volume(raw, acq.bp)
    got = server.beamform(raw, acq.bp)
    assert got is not None and np.linalg.norm(got - want) <= 1e-6 * np.linalg.norm(want)


def test_transfer_motion_and_corr(server, lib):
    print("transfer motion", flush=True)
    imgs = frames_with_shifts((96, 128), [(0, 0), (3, -2), (-1, 4)], 1)
    p = api.motion_params(128, 96, 3, (8, 6), 14, 16, (6, 6), abs_thr=0.3, rel_thr=0.02)
    want = np.zeros((3, 6, 8, 4), np.float32)
    assert api.motion_detection(imgs, want, p)
    got = server.motion_detect(imgs, p)
    assert got is not None and np.array_equal(got, want)
    print("transfer corr_images", flush=True)
    src = speckle((64, 80), 8) + 3.0
    tpl = np.ascontiguousarray(src[2:18, 1:17])
    want = np.zeros((64 - 16 + 1, 80 - 16 + 1), np.float32)
    assert api.corr_images(tpl, src, want)
    got = server.corr_images(tpl, src)
    assert got is not None and np.array_equal(got, want)
    # a request the library refuses comes back as an error, and the server keeps serving
    bad = api.motion_params(128, 96, 3, (8, 6), 14, 16, (6, 6))
    bad.reference_frame = 7
    assert server.motion_detect(imgs, bad) is None
    assert server.corr_images(tpl, src) is not None
