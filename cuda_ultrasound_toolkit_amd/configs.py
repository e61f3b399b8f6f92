"""Acquisition/processing set-ups used by the tests and benchmarks.

Each set-up returns a CudaBeamformerParameters block and the synthetic
acquisition that goes with it.

* ``linear_2d``: 32-element linear array, pitch 0.3 mm, c = 1540 m/s,
  fc = 5 MHz, fs = 20 MHz, Hadamard-32, 32 tx x 32 ch x 1024 samples int16,
  20 point scatterers (x in [-5, 5] mm, z in [5, 40] mm, amplitude U[0.2, 1]),
  1 % noise, +-16000 full scale, random channel permutation, no match filter,
  FORCES 2-D DAS to a 128 x 256 complex image, no apodization, cw = 0.
* ``rca_3d``: 128 x 128 row-column array, pitch 0.3 mm, 128 transmits
  (Hadamard-128) x 128 channels x 4096 samples int16 (128 MiB per frame), 500
  scatterers in a 38 x 38 x [5, 60] mm box, 3 % noise, channel permutation,
  64-tap 3-7 MHz linear-FM match filter (passed time reversed), HERCULES 3-D
  DAS to 128^3 over x, y in [-19, 19] mm, z in [5, 60] mm, F# 1.5 Hann
  receive apodization.  ``scale`` shrinks T, C, S and the grid together for
  tests that must finish in seconds on the FP64 oracle.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _lib as L


@dataclass
class Acquisition:
    bp: L.CudaBeamformerParameters
    raw: np.ndarray            # int16 [raw channels][T*S]
    channel_map: np.ndarray    # int16 [C]
    spec: L.CutSynthSpec

    @property
    def volume_shape(self):
        p = self.bp.output_points
        return (p[2], p[1], p[0])


def chirp_filter(taps: int = 64, f0: float = 3e6, f1: float = 7e6, fs: float = 20e6) -> np.ndarray:
    """Hann-tapered linear-FM chirp, unit energy, returned time reversed
    (the reference filters by circular convolution, so a reversed chirp
    correlates)."""
    t = np.arange(taps) / fs
    dur = taps / fs
    phase = 2 * np.pi * (f0 * t + 0.5 * (f1 - f0) / dur * t * t)
    h = np.sin(phase) * np.hanning(taps)
    h /= np.sqrt(np.sum(h * h))
    return h[::-1].astype(np.float32).copy()


def synth(spec: L.CutSynthSpec, raw_channels: int) -> tuple[np.ndarray, np.ndarray]:
    lib = L.load()
    raw = np.zeros((raw_channels, spec.T * spec.S), dtype=np.int16)
    cmap = np.zeros(spec.C, dtype=np.int16)
    ok = lib.cut_synth_rf(spec, raw.ctypes.data_as(L.C.POINTER(L.C.c_int16)),
                          cmap.ctypes.data_as(L.C.POINTER(L.C.c_int16)))
    if not ok:
        raise RuntimeError("cut_synth_rf failed")
    return raw, cmap


def _spec(geometry, T, C, S, n, pitch, box_min, box_max, noise, seed, encoding=L.HADAMARD,
          shift=(0.0, 0.0, 0.0), decorrelation=0.0, amp_seed=0) -> L.CutSynthSpec:
    s = L.CutSynthSpec()
    s.geometry = geometry
    s.T, s.C, s.S = T, C, S
    s.n_scatterers = n
    s.pitch = pitch
    s.c, s.fc, s.fs = 1540.0, 5e6, 20e6
    s.box_min[:] = box_min
    s.box_max[:] = box_max
    s.amp_min, s.amp_max = 0.2, 1.0
    s.noise_frac = noise
    s.peak = 16000.0
    s.encoding = encoding
    s.cycles = 2
    s.seed = seed
    s.shift[:] = shift
    s.decorrelation = decorrelation
    s.amp_seed = amp_seed
    return s


def linear_2d(seed: int = 1, T: int = 32, S: int = 1024, nx: int = 128, nz: int = 256,
              encoding: int = L.HADAMARD, generate: bool = True) -> Acquisition:
    C = T
    pitch = 0.3e-3
    spec = _spec(0, T, C, S, 20, pitch, (-5e-3, 0.0, 5e-3), (5e-3, 0.0, 40e-3), 0.01, seed, encoding)
    bp = L.CudaBeamformerParameters()
    bp.xdc_transform[0] = bp.xdc_transform[5] = bp.xdc_transform[10] = bp.xdc_transform[15] = 1.0
    bp.xdc_transform[12] = C * pitch / 2
    bp.xdc_transform[13] = 0.0
    bp.xdc_element_pitch[:] = (pitch, pitch)
    bp.rf_raw_dim[:] = (T * S, C, 0, 0)
    bp.dec_data_dim[:] = (S, C, T, 0)
    bp.decode = encoding
    bp.das_shader_id = L.FORCES
    bp.time_offset = 0.0
    bp.output_points[:] = (nx, 1, nz, 1)
    bp.output_min_coordinate[:] = (-5e-3, 0.0, 5e-3, 0.0)
    bp.output_max_coordinate[:] = (5e-3, 0.0, 40e-3, 0.0)
    bp.sampling_frequency, bp.center_frequency, bp.speed_of_sound = 20e6, 5e6, 1540.0
    bp.fn_tx = bp.fn_rx = 0.0
    bp.interpolate = True
    bp.coherency_weighting = 0.0
    bp.tx_orientation = L.ORIENT_NONE
    bp.rx_orientation = L.ORIENT_NONE
    bp.readi_group_count = 0
    bp.filter_length = 0
    bp.data_type = L.TYPE_I16
    bp.apo_type = L.RX_HANN
    bp.sample_type = L.SAMPLE_NORMAL
    for i in range(L.MAX_CHANNEL_COUNT):
        bp.sparse_elements[i] = -1
    raw = np.zeros((C, T * S), np.int16)
    cmap = np.arange(C, dtype=np.int16)
    if generate:
        raw, cmap = synth(spec, C)
    bp.channel_mapping[:C] = cmap.tolist()
    return Acquisition(bp, raw, cmap, spec)


def rca_3d(seed: int = 2, T: int = 128, S: int = 4096, n: int = 128, n_scatterers: int = 500,
           filter_taps: int = 64, fn_rx: float = 1.5, encoding: int = L.HADAMARD, generate: bool = True,
           shift=(0.0, 0.0, 0.0), decorrelation: float = 0.0, amp_seed: int = 0,
           z_range=(5e-3, 60e-3), half_width: float = 19e-3, box_half: float = 19e-3,
           noise: float = 0.03, focus_z: float = math.inf, C: int | None = None) -> Acquisition:
    C = T if C is None else C
    pitch = 0.3e-3
    spec = _spec(1, T, C, S, n_scatterers, pitch, (-box_half, -box_half, z_range[0]),
                 (box_half, box_half, z_range[1]), noise, seed, encoding, shift, decorrelation, amp_seed)
    bp = L.CudaBeamformerParameters()
    bp.xdc_transform[0] = bp.xdc_transform[5] = bp.xdc_transform[10] = bp.xdc_transform[15] = 1.0
    bp.xdc_transform[12] = C * pitch / 2
    bp.xdc_transform[13] = T * pitch / 2
    bp.xdc_element_pitch[:] = (pitch, pitch)
    bp.rf_raw_dim[:] = (T * S, C, 0, 0)
    bp.dec_data_dim[:] = (S, C, T, 0)
    bp.decode = encoding
    bp.das_shader_id = L.HERCULES
    bp.output_points[:] = (n, n, n, 1)
    bp.output_min_coordinate[:] = (-half_width, -half_width, z_range[0], 0.0)
    bp.output_max_coordinate[:] = (half_width, half_width, z_range[1], 0.0)
    bp.sampling_frequency, bp.center_frequency, bp.speed_of_sound = 20e6, 5e6, 1540.0
    bp.fn_tx = 0.0
    bp.fn_rx = fn_rx
    bp.interpolate = True
    bp.coherency_weighting = 0.0
    bp.foci[0], bp.foci[1], bp.foci[2] = 0.0, 0.0, focus_z
    bp.tx_orientation = L.ORIENT_ROWS
    bp.rx_orientation = L.ORIENT_COLS
    bp.readi_group_count = 0
    if filter_taps > 0:
        h = chirp_filter(filter_taps)
        bp.filter_length = filter_taps
        bp.rf_filter[:filter_taps] = h.tolist()
        # the correlation peak of a reversed L-tap chirp sits L-1 samples late
        bp.time_offset = (filter_taps - 1 + 0.5) / 20e6
    else:
        bp.filter_length = 0
        bp.time_offset = 0.0
    bp.data_type = L.TYPE_I16
    bp.apo_type = L.RX_HANN
    bp.sample_type = L.SAMPLE_NORMAL
    for i in range(L.MAX_CHANNEL_COUNT):
        bp.sparse_elements[i] = -1
    raw = np.zeros((C, T * S), np.int16)
    cmap = np.arange(C, dtype=np.int16)
    if generate:
        raw, cmap = synth(spec, C)
    bp.channel_mapping[:C] = cmap.tolist()
    return Acquisition(bp, raw, cmap, spec)


def uneven_frames(seed: int = 4, n: int = 64, small: bool = False, generate: bool = True) -> list:
    """Frames of uneven size for batching: transmits (Hadamard order) in
    {16, 32, 64, 128}, 128 receive channels, samples U{1024..4096}, 50-500
    scatterers, output cube side U{64..160} voxels, the rca_3d chain otherwise.
    The imaged depth follows the sample count (deepest point 90 % of the
    round-trip range).  small: a reduced set for tests (transmits {16, 32},
    32 channels, samples U{600..1100}, sides U{12..40}, 20-60 scatterers)."""
    return [_uneven(seed, i, d, generate) for i, d in enumerate(_uneven_draws(seed, n, small))]


def uneven_frame(seed: int, index: int, small: bool = False, generate: bool = True) -> Acquisition:
    """Frame `index` of uneven_frames(seed, ...), alone (the same sizes and data)."""
    return _uneven(seed, index, _uneven_draws(seed, index + 1, small)[index], generate)


def _uneven_draws(seed: int, n: int, small: bool) -> list:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        if small:
            T = int(rng.choice([16, 32]))
            Cc, S = 32, int(rng.integers(600, 1101))
            nsc, side = int(rng.integers(20, 61)), int(rng.integers(12, 41))
            taps = 16
        else:
            T = int(rng.choice([16, 32, 64, 128]))
            Cc, S = 128, int(rng.integers(1024, 4097))
            nsc, side = int(rng.integers(50, 501)), int(rng.integers(64, 161))
            taps = 64
        out.append((T, Cc, S, nsc, side, taps))
    return out


def _uneven(seed: int, i: int, draw, generate: bool) -> Acquisition:
    T, Cc, S, nsc, side, taps = draw
    half = 0.5 * Cc * 0.3e-3
    zmax = min(60e-3, 0.9 * S * 1540.0 / (2 * 20e6))
    zmin = min(5e-3, 0.25 * zmax)
    return rca_3d(seed=seed * 1000 + i, T=T, C=Cc, S=S, n=side, n_scatterers=nsc, filter_taps=taps,
                  z_range=(zmin, zmax), half_width=half, box_half=half, generate=generate)


def rca_small(seed: int = 3, **kw) -> Acquisition:
    """The rca_3d chain at 32 tx x 32 ch x 1024 samples -> 24 x 24 x 32 voxels
    (an aperture of 9.6 mm, box and grid scaled to it): oracle-checkable in seconds."""
    args = dict(T=32, S=1024, n=24, n_scatterers=40, filter_taps=16, z_range=(3e-3, 14e-3),
                half_width=4.5e-3, box_half=4.5e-3)
    args.update(kw)
    acq = rca_3d(seed=seed, **args)
    return acq


def uforces_2d(seed: int = 6, T: int = 16, S: int = 1024, nx: int = 128, nz: int = 256, step: int = 2,
               fn_rx: float = 1.0, generate: bool = True) -> Acquisition:
    """UFORCES on the linear_2d array: the first of T transmits is a flash
    that the beamformer discards, transmit t + 1 is fired from element
    ``step * t`` (sparse_elements), diverging from the array (transmit and
    receive on the same rows: focus at depth 0).  Hann receive apodization at
    F# ``fn_rx``.  The RF is the linear_2d synthesizer's (Hadamard-T)."""
    acq = linear_2d(seed=seed, T=T, S=S, nx=nx, nz=nz, generate=generate)
    bp = acq.bp
    bp.das_shader_id = L.UFORCES
    bp.tx_orientation = L.ORIENT_ROWS
    bp.rx_orientation = L.ORIENT_ROWS
    bp.foci[0], bp.foci[1], bp.foci[2] = 0.0, 0.0, 0.0
    bp.fn_rx = fn_rx
    for t in range(T):
        bp.sparse_elements[t] = step * t
    return acq


def tpw_3d(seed: int = 7, T: int = 16, C: int = 32, S: int = 1024, n: int = 24, max_angle_deg: float = 10.0,
           generate: bool = True, **kw) -> Acquisition:
    """RCA_TPW: T tilted plane waves from the rows (steering angles evenly
    spaced in [-max, max] degrees, passed in radians in foci[3t+2]), received
    on C columns, Hann receive apodization.  The RF is the rca_3d
    synthesizer's at the same sizes."""
    args = dict(T=T, C=C, S=S, n=n, n_scatterers=40, filter_taps=16, z_range=(3e-3, 14e-3),
                half_width=0.15e-3 * C, box_half=0.15e-3 * C)
    args.update(kw)
    acq = rca_3d(seed=seed, generate=generate, **args)
    bp = acq.bp
    bp.das_shader_id = L.RCA_TPW
    ang = np.deg2rad(np.linspace(-max_angle_deg, max_angle_deg, T))
    for t in range(T):
        bp.foci[3 * t + 0], bp.foci[3 * t + 1], bp.foci[3 * t + 2] = 0.0, 0.0, float(ang[t])
    return acq
