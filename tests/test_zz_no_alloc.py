"""cut_das_device_noalloc (workspace in the library's fixed device array)
against cut_das_device (workspace from the library's cached allocation),
through the C ABI.

Two calls in a row on each variant, the second with a smaller input: the full
128 x 128 x 4096 IQ frame to 128^3 (the split is limited to what the fixed
array holds, so the summation order differs from the tuned 16-way split:
relative L2 < 1e-5), then the 32 x 32 x 1024 frame to 24^3 (the same, and the
FP64 oracle within 1e-4).  The coherence-factor form (float4 partials) and
RCA_TPW (angle table in the array) are checked the same way.
"""
import numpy as np
import pytest

from cuda_ultrasound_toolkit_amd import api, configs

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def _iq(acq, dev, seed):
    import torch
    bp = acq.bp
    S, Cc, T = bp.dec_data_dim[0], bp.dec_data_dim[1], bp.dec_data_dim[2]
    g = torch.Generator(device=dev).manual_seed(seed)
    return torch.randn(T * Cc * S * 2, device=dev, generator=g)


def _both(acq, d_iq, dev):
    import torch
    n = acq.bp.output_points
    shape = (n[2], n[1], n[0])
    a = torch.full((n[0] * n[1] * n[2] * 2,), np.nan, dtype=torch.float32, device=dev)
    b = torch.full_like(a, np.nan)
    assert api.das_device(d_iq, a, acq.bp)
    assert api.das_device_noalloc(d_iq, b, acq.bp)
    torch.cuda.synchronize()
    return (a.cpu().numpy().view(np.complex64).reshape(shape), b.cpu().numpy().view(np.complex64).reshape(shape))


@pytest.mark.parametrize("cubic,cf", [(False, False), (True, True)])
def test_noalloc_matches_das_device_large_then_small(torch_dev, orc, cubic, cf):
    print(f"no-alloc DAS cubic={cubic} cf={cf}: 128^3 then 24^3", flush=True)
    assert api.set_das_options(cubic=cubic, cf_hercules=cf)
    try:
        big = configs.rca_3d(seed=2, generate=False)
        ref, got = _both(big, _iq(big, torch_dev, 1), torch_dev)
        assert np.all(np.isfinite(got))
        e_big = rel_l2(got, ref)
        small = configs.rca_small(seed=3, generate=False)
        d_small = _iq(small, torch_dev, 2)
        ref_s, got_s = _both(small, d_small, torch_dev)
    finally:
        assert api.set_das_options(cubic=True, cf_hercules=False)
    e_small = rel_l2(got_s, ref_s)
    iq = d_small.cpu().numpy().view(np.complex64).reshape(small.bp.dec_data_dim[2], small.bp.dec_data_dim[1], -1)
    want, _ = orc.das(iq.astype(np.complex128), small.bp, cubic=cubic, cf_hercules=cf)
    e_or = rel_l2(got_s, want)
    print(f"  128^3 vs cut_das_device {e_big:.2e}; 24^3 vs cut_das_device {e_small:.2e}, vs oracle {e_or:.2e}")
    assert e_big < 1e-5
    assert e_small < 1e-5
    assert e_or < 1e-4


def test_noalloc_rca_tpw_large_then_small(torch_dev, orc):
    print("no-alloc DAS RCA_TPW: 128^3 then 24^3", flush=True)
    big = configs.tpw_3d(seed=2, T=32, C=128, S=4096, n=128, generate=False, z_range=(5e-3, 60e-3),
                         half_width=19e-3, box_half=19e-3)
    ref, got = _both(big, _iq(big, torch_dev, 3), torch_dev)
    small = configs.tpw_3d(seed=7, generate=False)
    d_small = _iq(small, torch_dev, 4)
    ref_s, got_s = _both(small, d_small, torch_dev)
    iq = d_small.cpu().numpy().view(np.complex64).reshape(small.bp.dec_data_dim[2], small.bp.dec_data_dim[1], -1)
    want, _ = orc.das(iq.astype(np.complex128), small.bp)
    print(f"  128^3 vs cut_das_device {rel_l2(got, ref):.2e}; 24^3 {rel_l2(got_s, ref_s):.2e}, "
          f"vs oracle {rel_l2(got_s, want):.2e}")
    assert rel_l2(got, ref) < 1e-5
    assert rel_l2(got_s, ref_s) < 1e-5
    assert rel_l2(got_s, want) < 1e-4


def test_noalloc_rejects_null_arguments(lib, capfd):
    import ctypes as C
    acq = configs.rca_small(seed=3, generate=False)
    assert not lib.cut_das_device_noalloc(None, None, C.byref(acq.bp))
    assert "null argument" in capfd.readouterr().err
