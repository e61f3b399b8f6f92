"""Writes oracle_fixtures.npz: outputs of the FP64 oracle on small seeded
workloads, each with the SHA-256 of its generated raw input, so the tests can
check that the generator, the oracle and the HIP path still agree with a
committed record.  The reference holds no golden vectors for this path (it
has no tests); these are the build's own, made by:

    python tests/golden/make_oracle_fixtures.py

Cases: FORCES 2-D (32 x 32 x 1024 -> 64 x 128, cubic), HERCULES 3-D
(32 x 32 x 1024 -> 16^3, cubic, 16-tap filter), and the Hilbert transform of
two 4096-sample rows through the 64-tap chirp filter."""
import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from cuda_ultrasound_toolkit_amd import configs  # noqa: E402
import oracle  # noqa: E402

OUT = Path(__file__).with_name("oracle_fixtures.npz")


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def cases():
    """(name, acquisition or None, raw array) in a fixed order."""
    a = configs.linear_2d(seed=1, nx=64, nz=128)
    yield "forces_2d", a, a.raw
    b = configs.rca_small(seed=3, n=16)
    yield "hercules_3d", b, b.raw


def hilbert_input():
    rng = np.random.default_rng(2024)
    x = np.clip(rng.standard_normal((2, 4096)) * 4000, -32768, 32767).astype(np.float32)
    return x, configs.chirp_filter(64)


def main():
    oracle.load()
    out = {}
    for name, acq, raw in cases():
        out[f"{name}_raw_sha256"] = np.array(sha(raw))
        out[f"{name}_volume"] = oracle.pipeline(acq, cubic=True).astype(np.complex64)
    x, h = hilbert_input()
    out["hilbert_input_sha256"] = np.array(sha(x))
    out["hilbert_iq"] = oracle.hilbert(x.astype(np.float64), h.astype(np.float64)).astype(np.complex64)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
