"""Envelope-volume pairs for 3-D motion estimation.

A pair is two 128^3 (by default) envelope volumes |IQ| beamformed by the
plain pipeline from row-column acquisitions of the same scatterers before and
after a rigid shift.  The shift is given in voxels (integer plus sub-voxel)
and converted to metres with the voxel pitch of the output grid; the second
frame redraws a fraction of every scatterer amplitude (speckle
decorrelation) and carries its own electronic noise.

Volumes are float32 [nz][ny][nx] torch tensors on the current device; every
step runs through the library (cut_synth_rf on the host, cut_beamform_device
and cut_envelope on the device).
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from . import api, configs


def voxel_pitch(bp: L.CudaBeamformerParameters) -> tuple[float, float, float]:
    """(x, y, z) voxel pitch [m] of the output grid (y uses the x pitch, as the DAS does)."""
    n = bp.output_points
    px = (bp.output_max_coordinate[0] - bp.output_min_coordinate[0]) / max(1, n[0] - 1)
    pz = (bp.output_max_coordinate[2] - bp.output_min_coordinate[2]) / max(1, n[2] - 1)
    return px, px, pz


def beamform_envelope(acq: configs.Acquisition, dev) -> "torch.Tensor":
    import torch
    n = acq.bp.output_points
    raw = torch.from_numpy(acq.raw).to(dev)
    iq = torch.empty((n[2], n[1], n[0], 2), dtype=torch.float32, device=dev)
    if not api.beamform_device(raw, iq, acq.bp):
        raise RuntimeError("beamform_device failed")
    env = torch.empty((n[2], n[1], n[0]), dtype=torch.float32, device=dev)
    if not api.envelope(iq, env):
        raise RuntimeError("envelope failed")
    return env


def envelope_pair(seed: int, shift_vox, dev, decorrelation: float = 0.02, **rca_kw):
    """(template, source) envelope volumes; the source's scatterers are moved
    by shift_vox = (dx, dy, dz) voxels.  Returns (tpl, src, bp)."""
    a = configs.rca_3d(seed=seed, **rca_kw)
    px, py, pz = voxel_pitch(a.bp)
    shift_m = (shift_vox[0] * px, shift_vox[1] * py, shift_vox[2] * pz)
    b = configs.rca_3d(seed=seed, shift=shift_m, decorrelation=decorrelation, amp_seed=seed * 7919 + 1, **rca_kw)
    tpl = beamform_envelope(a, dev)
    src = beamform_envelope(b, dev)
    return tpl, src, a.bp


def random_shifts(rng: np.random.Generator, n: int, limit: float = 6.0) -> np.ndarray:
    """n rigid shifts in voxels, U[-limit, limit] per axis (integer plus sub-voxel)."""
    return rng.uniform(-limit, limit, size=(n, 3))
