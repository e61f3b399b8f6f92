"""ctypes wrapper of the FP64 CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
cuda_ultrasound_toolkit_amd package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None

DP = C.POINTER(C.c_double)


class DasConsts(C.Structure):
    _fields_ = [
        ("S", C.c_int), ("C", C.c_int), ("T", C.c_int),
        ("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int),
        ("min_x", C.c_float), ("min_y", C.c_float), ("min_z", C.c_float),
        ("res_x", C.c_float), ("res_y", C.c_float), ("res_z", C.c_float),
        ("xdc_min_x", C.c_float), ("xdc_min_y", C.c_float), ("xdc_max_x", C.c_float), ("xdc_max_y", C.c_float),
        ("pitch_x", C.c_float), ("pitch_y", C.c_float),
        ("spm", C.c_float),
        ("delay", C.c_int),
        ("focal_z", C.c_float),
        ("focus_type", C.c_int),
        ("fn_tx", C.c_float), ("fn_rx", C.c_float), ("cw", C.c_float),
        ("readi_groups", C.c_int), ("readi_id", C.c_int), ("encoding", C.c_int),
        ("apo_type", C.c_int), ("to_power", C.c_int),
        ("sequence", C.c_int),
        ("fs", C.c_float), ("fc", C.c_float),
        ("tx_angle", C.c_float * 256),
        ("tx_element", C.c_int * 256),
    ]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    lib.or_hadamard.restype = C.c_int
    lib.or_hadamard.argtypes = [C.c_int, C.c_int, DP]
    lib.or_decode.argtypes = [DP, DP, C.c_int, C.c_size_t, DP]
    lib.or_convert.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int16), C.c_size_t, C.c_int, C.c_int, C.c_int, DP]
    lib.or_convert_demod.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int16), C.c_size_t, C.c_int, C.c_int,
                                     C.c_int, C.c_double, C.c_double, DP, C.c_int, DP]
    lib.or_dft.argtypes = [DP, DP, C.c_int, C.c_int, DP, DP]
    lib.or_fft.argtypes = [DP, DP, C.c_int, C.c_int]
    lib.or_filter_spectrum.argtypes = [DP, C.c_int, C.c_int, DP]
    lib.or_hilbert.argtypes = [DP, C.c_size_t, C.c_int, DP, C.c_int, DP]
    lib.or_das_constants.restype = C.c_int
    lib.or_das_constants.argtypes = [C.c_void_p, C.POINTER(DasConsts)]
    lib.or_das.restype = C.c_int
    lib.or_das.argtypes = [DP, C.POINTER(DasConsts), C.c_int, C.c_int, DP, C.POINTER(C.c_longlong)]
    lib.or_das_planes.restype = C.c_int
    lib.or_das_planes.argtypes = [DP, C.POINTER(DasConsts), C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int, DP,
                                  C.POINTER(C.c_longlong)]
    lib.or_zncc_valid.argtypes = [DP, C.c_int, C.c_int, C.c_int, DP, C.c_int, C.c_int, C.c_int, DP]
    lib.or_motion_detection.restype = C.c_int
    lib.or_motion_detection.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.POINTER(C.c_float)]
    FP = C.POINTER(C.c_float)
    IP = C.POINTER(C.c_int)
    lib.or_zncc3d_at.restype = C.c_double
    lib.or_zncc3d_at.argtypes = [FP, FP] + [C.c_int] * 10
    lib.or_ncc3d_full.argtypes = [FP, FP] + [C.c_int] * 6 + [IP, FP]
    lib.or_ncc3d_window.argtypes = [FP, FP] + [C.c_int] * 6 + [IP, C.c_int, IP, FP]
    lib.or_pnet_param_count.restype = C.c_int
    lib.or_pnet_forward.argtypes = [DP, DP, C.c_int, DP]
    lib.or_pnet_loss_grad.restype = C.c_double
    lib.or_pnet_loss_grad.argtypes = [DP, DP, DP, C.c_int, DP]
    _lib = lib
    return lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(DP)


def hadamard(n: int, ordering: int) -> np.ndarray | None:
    H = np.zeros((n, n), np.float64)
    return H if load().or_hadamard(n, ordering, _dp(H)) == 0 else None


def convert(raw: np.ndarray, dtype: int, cmap: np.ndarray, R0: int, T: int, Cc: int, S: int) -> np.ndarray:
    out = np.zeros((T, Cc, S), np.float64)
    m = np.ascontiguousarray(cmap, np.int16)
    raw = np.ascontiguousarray(raw)
    load().or_convert(raw.ctypes.data, dtype, m.ctypes.data_as(C.POINTER(C.c_int16)), R0, T, Cc, S, _dp(out))
    return out


def convert_demod(raw: np.ndarray, dtype: int, cmap: np.ndarray, R0: int, T: int, Cc: int, S: int,
                  fd: float, fs: float, h=None) -> np.ndarray:
    """Bandpass (I, Q)-pair rows -> demodulated, optionally FIR-filtered complex [T][C][S]."""
    out = np.zeros(2 * T * Cc * S, np.float64)
    m = np.ascontiguousarray(cmap, np.int16)
    raw = np.ascontiguousarray(raw)
    if h is None or len(h) == 0:
        hp, L = None, 0
    else:
        hh = np.ascontiguousarray(h, np.float64)
        hp, L = _dp(hh), len(hh)
    load().or_convert_demod(raw.ctypes.data, dtype, m.ctypes.data_as(C.POINTER(C.c_int16)), R0, T, Cc, S,
                            float(fd), float(fs), hp, L, _dp(out))
    return (out[0::2] + 1j * out[1::2]).reshape(T, Cc, S)


def decode(conv: np.ndarray, H: np.ndarray) -> np.ndarray:
    T = H.shape[0]
    conv = np.ascontiguousarray(conv, np.float64)
    cols = conv.size // T
    out = np.zeros_like(conv)
    load().or_decode(_dp(conv), _dp(np.ascontiguousarray(H)), T, cols, _dp(out))
    return out


def dft(x: np.ndarray, sign: int = -1) -> np.ndarray:
    x = np.asarray(x, np.complex128)
    re = np.ascontiguousarray(x.real)
    im = np.ascontiguousarray(x.imag)
    ore = np.zeros(len(x))
    oim = np.zeros(len(x))
    load().or_dft(_dp(re), _dp(im), len(x), sign, _dp(ore), _dp(oim))
    return ore + 1j * oim


def fft(x: np.ndarray, sign: int = -1) -> np.ndarray:
    x = np.asarray(x, np.complex128)
    re = np.ascontiguousarray(x.real.copy())
    im = np.ascontiguousarray(x.imag.copy())
    load().or_fft(_dp(re), _dp(im), len(x), sign)
    return re + 1j * im


def filter_spectrum(h: np.ndarray, S: int) -> np.ndarray:
    h = np.ascontiguousarray(h, np.float64)
    out = np.zeros(2 * S)
    load().or_filter_spectrum(_dp(h), len(h), S, _dp(out))
    return out[0::2] + 1j * out[1::2]


def hilbert(x: np.ndarray, h: np.ndarray | None = None) -> np.ndarray:
    """x: real rows [..., S] -> complex rows (reference Hilbert + match filter)."""
    x = np.ascontiguousarray(x, np.float64)
    S = x.shape[-1]
    rows = x.size // S
    out = np.zeros(2 * x.size)
    if h is None or len(h) == 0:
        load().or_hilbert(_dp(x), rows, S, None, 0, _dp(out))
    else:
        hh = np.ascontiguousarray(h, np.float64)
        load().or_hilbert(_dp(x), rows, S, _dp(hh), len(hh), _dp(out))
    return (out[0::2] + 1j * out[1::2]).reshape(x.shape)


def das_constants(bp) -> DasConsts:
    k = DasConsts()
    load().or_das_constants(C.addressof(bp), C.byref(k))
    return k


def das(iq: np.ndarray, bp, cubic: bool = True, cf_hercules: bool = False, planes=None):
    """iq: complex [T][C][S] -> (complex volume [Nz][Ny][Nx], active tap count).
    planes: depth-plane indices of the grid to evaluate (volume [len(planes)][Ny][Nx])."""
    k = das_constants(bp)
    iqd = np.ascontiguousarray(np.stack([iq.real, iq.imag], axis=-1), np.float64)
    nz = k.Nz if planes is None else len(planes)
    out = np.zeros(2 * k.Nx * k.Ny * nz)
    act = C.c_longlong(0)
    if planes is None:
        rc = load().or_das(_dp(iqd), C.byref(k), 0 if cubic else 1, 1 if cf_hercules else 0, _dp(out),
                           C.byref(act))
    else:
        pl = np.ascontiguousarray(planes, np.int32)
        rc = load().or_das_planes(_dp(iqd), C.byref(k), 0 if cubic else 1, 1 if cf_hercules else 0,
                                  pl.ctypes.data_as(C.POINTER(C.c_int)), len(pl), _dp(out), C.byref(act))
    if rc != 0:
        raise ValueError("oracle DAS rejected the parameters")
    vol = (out[0::2] + 1j * out[1::2]).reshape(nz, k.Ny, k.Nx)
    return vol, act.value


def pipeline(acq, cubic: bool = True, cf_hercules: bool = False, return_stages: bool = False, planes=None):
    """The whole reference chain on the CPU: convert -> decode -> Hilbert -> DAS
    (on the depth planes `planes` of the grid only, when given)."""
    bp = acq.bp
    S, Cc, T = bp.dec_data_dim[0], bp.dec_data_dim[1], bp.dec_data_dim[2]
    R0 = bp.rf_raw_dim[0]
    if bp.sample_type == 1:  # SAMPLE_BANDPASS: (I, Q) pairs -> demodulation + centred FIR, no Hilbert
        h = np.array(bp.rf_filter[: bp.filter_length], np.float64) if bp.filter_length else None
        iq = convert_demod(acq.raw, bp.data_type, np.array(bp.channel_mapping[:Cc], np.int16), R0, T, Cc, S,
                           bp.center_frequency, bp.sampling_frequency, h)
        vol, act = das(iq, bp, cubic, cf_hercules, planes)
        if return_stages:
            return dict(iq=iq, vol=vol, active=act)
        return vol
    conv = convert(acq.raw, bp.data_type, np.array(bp.channel_mapping[:Cc], np.int16), R0, T, Cc, S)
    if bp.decode:
        H = hadamard(T, bp.decode)
        dec = decode(conv, H)
    else:
        dec = conv
    h = np.array(bp.rf_filter[: bp.filter_length], np.float64) if bp.filter_length else None
    iq = hilbert(dec.reshape(T * Cc, S), h).reshape(T, Cc, S)
    vol, act = das(iq, bp, cubic, cf_hercules, planes)
    if return_stages:
        return dict(conv=conv, dec=dec, iq=iq, vol=vol, active=act)
    return vol


def zncc_valid(src: np.ndarray, tpl: np.ndarray) -> np.ndarray:
    """Valid-region ZNCC map [Hs-Ht+1][Ws-Wt+1] (NPP CrossCorrValid_NormLevel semantics)."""
    src = np.ascontiguousarray(src, np.float64)
    tpl = np.ascontiguousarray(tpl, np.float64)
    sh, sw = src.shape
    th, tw = tpl.shape
    out = np.zeros((sh - th + 1, sw - tw + 1))
    load().or_zncc_valid(_dp(src), sw, sh, sw, _dp(tpl), tw, th, tw, _dp(out))
    return out


def motion_detection(images: np.ndarray, params) -> np.ndarray:
    """images f32 [F][H][W] -> float4 maps [F][gy][gx][4] (reference block matching)."""
    images = np.ascontiguousarray(images, np.float32)
    F = params.frame_count
    out = np.zeros((F, params.motion_grid_dims[1], params.motion_grid_dims[0], 4), np.float32)
    load().or_motion_detection(images.ctypes.data_as(C.POINTER(C.c_float)), C.addressof(params),
                               out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def ncc3d_full(tpl: np.ndarray, src: np.ndarray, block=16, stride=8, margin=8, centre=None, radius=2):
    """Full (or windowed, when centre is given) 3-D ZNCC search: (disp [nb][3] int32 (dx,dy,dz), peak [nb])."""
    tpl = np.ascontiguousarray(tpl, np.float32)
    src = np.ascontiguousarray(src, np.float32)
    nz, ny, nx = tpl.shape
    nb = ((nx - block) // stride + 1) * ((ny - block) // stride + 1) * ((nz - block) // stride + 1)
    disp = np.zeros((nb, 3), np.int32)
    peak = np.zeros(nb, np.float32)
    if centre is None:
        load().or_ncc3d_full(_fp(tpl), _fp(src), nx, ny, nz, block, stride, margin, _ip(disp), _fp(peak))
    else:
        c = np.ascontiguousarray(centre, np.int32)
        load().or_ncc3d_window(_fp(tpl), _fp(src), nx, ny, nz, block, stride, margin, _ip(c), radius,
                               _ip(disp), _fp(peak))
    return disp, peak


def zncc3d_at(tpl, src, bx, by, bz, block, dx, dy, dz):
    tpl = np.ascontiguousarray(tpl, np.float32)
    src = np.ascontiguousarray(src, np.float32)
    nz, ny, nx = tpl.shape
    return load().or_zncc3d_at(_fp(tpl), _fp(src), nx, ny, nz, bx, by, bz, block, dx, dy, dz)


def pnet_param_count() -> int:
    return load().or_pnet_param_count()


def pnet_forward(params: np.ndarray, x: np.ndarray) -> np.ndarray:
    """x [B][2][16][16][16] -> [B][3] (FP64 CNN forward)."""
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    B = x.shape[0]
    out = np.zeros((B, 3))
    load().or_pnet_forward(_dp(p), _dp(x), B, _dp(out))
    return out


def pnet_loss_grad(params: np.ndarray, x: np.ndarray, target: np.ndarray):
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    t = np.ascontiguousarray(target, np.float64)
    g = np.zeros_like(p)
    loss = load().or_pnet_loss_grad(_dp(p), _dp(x), _dp(t), x.shape[0], _dp(g))
    return loss, g
