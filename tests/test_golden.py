"""Committed oracle fixtures (tests/golden/oracle_fixtures.npz, written by
tests/golden/make_oracle_fixtures.py): the seeded generators must still make
the same raw input (SHA-256), the FP64 oracle must still produce the recorded
outputs, and the HIP path must match them to the same bars as against a live
oracle (1e-4 relative L2 for volumes, 1e-5 for the Hilbert transform)."""
import hashlib
import sys
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(HERE))

import make_oracle_fixtures as mk  # noqa: E402

FIX = np.load(HERE / "oracle_fixtures.npz")
CASES = {name: (acq, raw) for name, acq, raw in mk.cases()}


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.mark.parametrize("name", sorted(CASES))
def test_generator_reproduces_raw(name):
    assert mk.sha(CASES[name][1]) == str(FIX[f"{name}_raw_sha256"])


def test_hilbert_input_reproduces():
    x, _ = mk.hilbert_input()
    assert mk.sha(x) == str(FIX["hilbert_input_sha256"])


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_fixture(orc, name):
    acq, _ = CASES[name]
    got = orc.pipeline(acq, cubic=True)
    want = FIX[f"{name}_volume"]
    assert got.shape == want.shape
    assert rel_l2(got, want) < 1e-6


def test_oracle_hilbert_reproduces_fixture(orc):
    x, h = mk.hilbert_input()
    got = orc.hilbert(x.astype(np.float64), h.astype(np.float64))
    assert rel_l2(got, FIX["hilbert_iq"]) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_hip_pipeline_matches_fixture(lib, name):
    from cuda_ultrasound_toolkit_amd import api
    print(f"golden {name}", flush=True)
    acq, raw = CASES[name]
    assert api.set_das_options(cubic=True, cf_hercules=False)
    got = api.beamform_volume(raw, acq.bp)
    assert got is not None
    want = FIX[f"{name}_volume"]
    assert rel_l2(got.reshape(want.shape), want) < 1e-4


@pytest.mark.gpu
def test_hip_hilbert_matches_fixture(torch_dev):
    import torch
    from cuda_ultrasound_toolkit_amd import _lib as L, api, configs
    print("golden hilbert", flush=True)
    x, h = mk.hilbert_input()
    T, S = x.shape
    acq = configs.linear_2d(T=T, S=S, generate=False)
    acq.bp.dec_data_dim[1] = 1
    assert api.rf_init((T * S, 1), (S, 1, T), L.HADAMARD)
    assert api.rf_set_match_filter(h)
    out = torch.zeros(T * S * 2, dtype=torch.float32, device=torch_dev)
    assert api.rf_hilbert(torch.from_numpy(x).to(torch_dev), 1, out)
    torch.cuda.synchronize()
    assert api.rf_set_match_filter(None)
    got = out.cpu().numpy().view(np.complex64).reshape(T, S)
    assert rel_l2(got, FIX["hilbert_iq"]) < 1e-5
