"""The benchmark's own launch at full size against the FP64 oracle.

The 128 tx x 128 ch x 4096 frame (64-tap match filter, linear interpolation,
F# 1.5 Hann) is beamformed to the whole 128^3 volume through
cut_beamform_device exactly as the benchmark does (16-way transmit split and
all); depth planes 16, 48, 80 and 112 of that volume are compared with the
oracle evaluated on the same planes of the same grid.  Bar: relative L2 below
1e-4 (FP32 against FP64) on every seed.
"""
import numpy as np
import pytest

from cuda_ultrasound_toolkit_amd import api, configs

pytestmark = pytest.mark.gpu

PLANES = [16, 48, 80, 112]


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("seed", [2, 5, 9])
def test_full_volume_planes_match_oracle(torch_dev, orc, seed):
    import torch
    print(f"full-size rca_3d seed {seed}: whole 128^3 launch vs oracle planes {PLANES}", flush=True)
    acq = configs.rca_3d(seed=seed)
    bp = acq.bp
    n = bp.output_points
    assert api.set_das_options(cubic=False)
    try:
        d_raw = torch.from_numpy(acq.raw).to(torch_dev)
        vol = torch.zeros(n[0] * n[1] * n[2] * 2, dtype=torch.float32, device=torch_dev)
        assert api.beamform_device(d_raw, vol, bp)
        assert api.beamform_device(d_raw, vol, bp)  # a second call, as the benchmark's timed steps
        torch.cuda.synchronize()
    finally:
        assert api.set_das_options(cubic=True)
    got = vol.cpu().numpy().view(np.complex64).reshape(n[2], n[1], n[0])
    assert np.all(np.isfinite(got))
    want = orc.pipeline(acq, cubic=False, planes=PLANES)
    errs = [rel_l2(got[p], want[i]) for i, p in enumerate(PLANES)]
    err = rel_l2(got[PLANES], want)
    print(f"  rel L2 {err:.3e} (per plane {', '.join(f'{e:.2e}' for e in errs)}); "
          f"max |err| / max |v| {np.abs(got[PLANES] - want).max() / np.abs(want).max():.3e}")
    assert err < 1e-4
