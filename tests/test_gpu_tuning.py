"""Tuned launch shapes and reproducibility across processes.

By default the choices that change the rounding (HERCULES transmit split, FFT
tile) are fixed heuristics, so two fresh processes beamform the same frames to
the same bits.  With timed tuning on (CUT_AUTOTUNE=1), two fresh processes that
share a tuning file (CUT_TUNING_FILE: the first one tunes and writes it, the
second reads it) do too, through every tuned choice.
"""
import json
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent

CHILD = r"""
import sys, numpy as np, torch
sys.path.insert(0, sys.argv[1])
from cuda_ultrasound_toolkit_amd import api, configs
torch.cuda.set_device(0)
outs = []
for S in (1000, 1024):  # chirp-z and power-of-two Hilbert
    acq = configs.rca_small(seed=13, S=S)
    assert api.set_das_options(cubic=False)
    v = api.beamform_volume(acq.raw, acq.bp)
    assert v is not None
    outs.append(v)
np.save(sys.argv[2], np.stack(outs))
"""


def _clean_env(**kw):
    env = {k: v for k, v in os.environ.items() if k not in ("CUT_TUNING_FILE", "CUT_AUTOTUNE")}
    env.update(kw)
    return env


def test_default_processes_compute_the_same_bits(tmp_path):
    print("two processes, default settings", flush=True)
    vols = []
    for i in range(2):
        out = tmp_path / f"vol{i}.npy"
        r = subprocess.run([sys.executable, "-c", CHILD, str(ROOT), str(out)], env=_clean_env(), capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-3000:]
        vols.append(np.load(out))
    assert np.array_equal(vols[0], vols[1])


def test_processes_sharing_a_tuning_file_compute_the_same_bits(tmp_path):
    print("two processes, timed tuning, one tuning file", flush=True)
    tune = tmp_path / "tuning.txt"
    env = _clean_env(CUT_TUNING_FILE=str(tune), CUT_AUTOTUNE="1")
    vols = []
    for i in range(2):
        out = tmp_path / f"vol{i}.npy"
        r = subprocess.run([sys.executable, "-c", CHILD, str(ROOT), str(out)], env=env, capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-3000:]
        vols.append(np.load(out))
        if i == 0:
            lines = tune.read_text().splitlines()
            tables = {int(l.split()[0]) for l in lines}
            print(f"  tuning file: {len(lines)} choices, tables {sorted(tables)}")
            assert {0, 2} <= tables  # DAS split and FFT tile were tuned and written
    assert np.array_equal(vols[0], vols[1])


def test_tuning_save_and_load_round_trip(lib, tmp_path):
    from cuda_ultrasound_toolkit_amd import api
    a = tmp_path / "a.txt"
    assert api.tuning_save(a)
    assert api.tuning_load(a)
    b = tmp_path / "b.txt"
    assert api.tuning_save(b)
    assert a.read_text() == b.read_text()
