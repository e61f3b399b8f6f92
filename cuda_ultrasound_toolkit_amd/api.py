"""Python mirror of the toolkit's entry points.

Same names, argument meaning and error behaviour as the reference
(cuda_toolkit/src/public/cuda_toolkit.hpp:17-29, cuda_toolkit_ogl.h:37-75):
each call returns False (after the library prints the reason on stderr)
instead of raising.  Host arrays are numpy; device arrays are any object with
``data_ptr()`` (torch tensors on the GPU), used only as plumbing.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _span(a: np.ndarray) -> L.Span:
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays passed to the library must be C-contiguous")
    return L.Span(a.ctypes.data, a.nbytes)


def beamform(input_data: np.ndarray, output_data: np.ndarray, bp: L.CudaBeamformerParameters) -> bool:
    """cuda_toolkit::beamform: raw RF bytes in, complex64 volume bytes out (host memory)."""
    lib = L.load()
    return bool(getattr(lib, L.SYM_BEAMFORM)(_span(input_data), _span(output_data), C.byref(bp)))


def beamform_volume(raw: np.ndarray, bp: L.CudaBeamformerParameters) -> np.ndarray | None:
    """Convenience wrapper: returns the [Nz][Ny][Nx] complex64 volume or None."""
    nx, ny, nz = bp.output_points[0], bp.output_points[1], bp.output_points[2]
    out = np.zeros((nz, ny, nx), dtype=np.complex64)
    return out if beamform(np.ascontiguousarray(raw), out, bp) else None


def motion_detection(images: np.ndarray, motion_maps: np.ndarray, params: L.NccMotionParameters) -> bool:
    """cuda_toolkit::motion_detection: f32 frames [F][H][W] -> float4 maps [F][gy][gx][4]."""
    lib = L.load()
    return bool(getattr(lib, L.SYM_MOTION)(_span(images), _span(motion_maps), C.byref(params)))


def corr_images(template: np.ndarray, source: np.ndarray, corr_map: np.ndarray) -> bool:
    """cuda_toolkit::corr_images on f32 images [H][W]; corr_map [Hs-Ht+1][Ws-Wt+1]."""
    lib = L.load()
    td = (C.c_uint * 2)(template.shape[1], template.shape[0])
    sd = (C.c_uint * 2)(source.shape[1], source.shape[0])
    return bool(getattr(lib, L.SYM_CORR)(_span(template), _span(source), _span(corr_map), td, sd))


def motion_params(W, H, frames, grid, spacing, patch, margins, abs_thr=0.5, rel_thr=0.05, min_var=0.0,
                  subpixel=True, ref=0, neighbour=False) -> L.NccMotionParameters:
    p = L.NccMotionParameters()
    p.patch_size, p.motion_grid_spacing = patch, spacing
    p.motion_grid_dims[:] = grid
    p.search_margins[:] = margins
    p.abs_cor_threshold, p.rel_cor_threshold, p.min_patch_variance = abs_thr, rel_thr, min_var
    p.use_subpixel, p.reference_frame, p.neighbour_compare = subpixel, ref, neighbour
    p.frame_count = frames
    p.image_dims[:] = (W, H)
    p.data_type = L.TYPE_F32
    return p


# ---- C API over GL buffers --------------------------------------------------
def _uints(v):
    arr = (C.c_uint * len(v))(*v)
    return arr


def init_cuda_configuration(input_dims, decoded_dims) -> bool:
    return bool(L.load().init_cuda_configuration(_uints(input_dims), _uints(decoded_dims)))


def register_cuda_buffers(rf_ssbos, raw_ssbo: int) -> bool:
    return bool(L.load().register_cuda_buffers(_uints(rf_ssbos), len(rf_ssbos), raw_ssbo))


def cuda_set_channel_mapping(mapping: np.ndarray) -> bool:
    m = np.zeros(L.MAX_CHANNEL_COUNT, np.int16)
    m[: len(mapping)] = mapping
    return bool(L.load().cuda_set_channel_mapping(m.ctypes.data_as(C.POINTER(C.c_int16))))


def cuda_set_match_filter(taps: np.ndarray | None) -> bool:
    if taps is None or len(taps) == 0:
        return bool(L.load().cuda_set_match_filter(None, 0))
    h = np.ascontiguousarray(taps, dtype=np.float32)
    return bool(L.load().cuda_set_match_filter(h.ctypes.data_as(C.POINTER(C.c_float)), len(h)))


def cuda_decode(input_byte_offset: int, output_buffer_idx: int) -> bool:
    return bool(L.load().cuda_decode(input_byte_offset, output_buffer_idx))


def cuda_hilbert(input_buffer_idx: int, output_buffer_idx: int) -> bool:
    return bool(L.load().cuda_hilbert(input_buffer_idx, output_buffer_idx))


def deinit_cuda_configuration() -> None:
    L.load().deinit_cuda_configuration()


def register_device_buffers(rf_tensors, raw_tensor) -> bool:
    """GL seam: register device tensors in place of GL buffers."""
    n = len(rf_tensors)
    ptrs = (C.c_void_p * max(n, 1))(*[t.data_ptr() for t in rf_tensors])
    sizes = (C.c_size_t * max(n, 1))(*[t.numel() * t.element_size() for t in rf_tensors])
    return bool(L.load().cut_gl_register_device_buffers(ptrs, sizes, n, raw_tensor.data_ptr(),
                                                         raw_tensor.numel() * raw_tensor.element_size()))


# ---- extension entry points on device memory --------------------------------
def set_stream(stream_handle: int | None) -> None:
    L.load().cut_set_stream(stream_handle)


def synchronize() -> bool:
    return bool(L.load().cut_synchronize())


def set_decode_mode(mode: int) -> bool:
    return bool(L.load().cut_set_decode_mode(mode))


def set_das_options(cubic: bool = True, cf_hercules: bool = False, iq_format: int = 0) -> bool:
    return bool(L.load().cut_set_das_options(int(cubic), int(cf_hercules), iq_format))


def beamform_device(d_raw, d_volume, bp: L.CudaBeamformerParameters) -> bool:
    return bool(L.load().cut_beamform_device(d_raw.data_ptr(), d_raw.numel() * d_raw.element_size(),
                                             d_volume.data_ptr(), d_volume.numel() * d_volume.element_size(),
                                             C.byref(bp)))


def rf_init(raw_dims, dec_dims, encoding: int) -> bool:
    return bool(L.load().cut_rf_init(_uints(raw_dims), _uints(dec_dims), encoding))


def rf_set_channel_mapping(mapping: np.ndarray) -> bool:
    m = np.ascontiguousarray(mapping, dtype=np.int16)
    return bool(L.load().cut_rf_set_channel_mapping(m.ctypes.data_as(C.POINTER(C.c_int16)), len(m)))


def rf_set_match_filter(taps) -> bool:
    if taps is None or len(taps) == 0:
        return bool(L.load().cut_rf_set_match_filter(None, 0))
    h = np.ascontiguousarray(taps, dtype=np.float32)
    return bool(L.load().cut_rf_set_match_filter(h.ctypes.data_as(C.POINTER(C.c_float)), len(h)))


def rf_convert_decode(d_raw, dtype: int, d_out, out_stride: int = 1, decode: bool = True) -> bool:
    return bool(L.load().cut_rf_convert_decode(d_raw.data_ptr(), dtype, d_out.data_ptr(), out_stride, int(decode)))


def rf_convert_demod(d_raw, dtype: int, d_out, demod_freq: float, sample_freq: float) -> bool:
    """Bandpass (I, Q)-pair rows -> demodulated, FIR-filtered complex [T][C][S] (device tensors)."""
    return bool(L.load().cut_rf_convert_demod(d_raw.data_ptr(), dtype, d_out.data_ptr(), float(demod_freq),
                                              float(sample_freq)))


def rf_hilbert(d_in, in_stride: int, d_out) -> bool:
    return bool(L.load().cut_rf_hilbert(d_in.data_ptr(), in_stride, d_out.data_ptr()))


def rf_hilbert_format(d_in, in_stride: int, d_out, fmt: int) -> tuple[bool, float]:
    """Hilbert into IQ stored as fmt (0 complex f32, 1 FP16 x scale, 2 BF16); returns (ok, scale)."""
    sc = C.c_float(0.0)
    ok = bool(L.load().cut_rf_hilbert_format(d_in.data_ptr(), in_stride, d_out.data_ptr(), fmt, C.byref(sc)))
    return ok, sc.value


def das_device_format(d_iq, fmt: int, scale: float, d_volume, bp: L.CudaBeamformerParameters) -> bool:
    return bool(L.load().cut_das_device_format(d_iq.data_ptr(), fmt, scale, d_volume.data_ptr(), C.byref(bp)))


def das_device(d_iq, d_volume, bp: L.CudaBeamformerParameters) -> bool:
    return bool(L.load().cut_das_device(d_iq.data_ptr(), d_volume.data_ptr(), C.byref(bp)))


def das_device_noalloc(d_iq, d_volume, bp: L.CudaBeamformerParameters) -> bool:
    """cut_das_device with its workspace in the library's fixed device array (allocates nothing)."""
    return bool(L.load().cut_das_device_noalloc(d_iq.data_ptr(), d_volume.data_ptr(), C.byref(bp)))


def _ptr(t):
    return None if t is None else t.data_ptr()


def ncc3d_search(d_tpl, d_src, stride, d_disp, d_peak, d_block_ids=None, n_blocks=None, d_centre=None,
                 radius=2) -> bool:
    """3-D ZNCC block search on device volumes [nz][ny][nx] (16^3 blocks, |d| <= 8)."""
    nz, ny, nx = d_tpl.shape
    if n_blocks is None:
        n_blocks = ((nx - 16) // stride + 1) * ((ny - 16) // stride + 1) * ((nz - 16) // stride + 1)
    return bool(L.load().cut_ncc3d_search(d_tpl.data_ptr(), d_src.data_ptr(), nx, ny, nz, stride,
                                          _ptr(d_block_ids), n_blocks, _ptr(d_centre), radius,
                                          d_disp.data_ptr(), d_peak.data_ptr()))


def envelope(d_iq, d_out) -> bool:
    return bool(L.load().cut_envelope(d_iq.data_ptr(), d_out.numel(), d_out.data_ptr()))


# ---- learned displacement proposal ----------------------------------------
def pnet_param_count() -> int:
    return int(L.load().cut_pnet_param_count())


def pnet_set_precision(fp32: bool) -> bool:
    """Training-path products: BF16 MFMA (default) or FP32 (verification)."""
    return bool(L.load().cut_pnet_set_precision(1 if fp32 else 0))


def pnet_set_input_mode(mode: int) -> bool:
    """0: the 16^3 block pair; 1: the pooled 32^3 search regions."""
    return bool(L.load().cut_pnet_set_input_mode(mode))


def pnet_gather(d_vols, shape, stride, d_desc, d_out) -> bool:
    """d_vols: int64 device tensor of volume addresses [2 * pairs] (template,
    source); d_desc int32 [n][6]; d_out f32 [n][2][16][16][16]."""
    nz, ny, nx = shape
    return bool(L.load().cut_pnet_gather(d_vols.data_ptr(), nx, ny, nz, stride, d_desc.data_ptr(), d_desc.shape[0],
                                         d_out.data_ptr()))


def pnet_forward(d_params, d_x, d_y) -> bool:
    return bool(L.load().cut_pnet_forward(d_params.data_ptr(), d_x.data_ptr(), d_x.shape[0], d_y.data_ptr()))


def pnet_loss_grad(d_params, d_x, d_target, d_grad, d_loss=None) -> bool:
    return bool(L.load().cut_pnet_loss_grad(d_params.data_ptr(), d_x.data_ptr(), d_target.data_ptr(), d_x.shape[0],
                                            d_grad.data_ptr(), _ptr(d_loss)))


def pnet_adam(d_params, d_grad, d_m, d_v, lr, step, beta1=0.9, beta2=0.999, eps=1e-8, wd=0.0) -> bool:
    return bool(L.load().cut_pnet_adam(d_params.data_ptr(), d_grad.data_ptr(), d_m.data_ptr(), d_v.data_ptr(),
                                       d_params.numel(), lr, beta1, beta2, eps, step, wd))


def pnet_infer(d_params, d_tpl, d_src, stride, d_pred, d_centre=None, d_block_ids=None) -> bool:
    nz, ny, nx = d_tpl.shape
    return bool(L.load().cut_pnet_infer(d_params.data_ptr(), d_tpl.data_ptr(), d_src.data_ptr(), nx, ny, nz, stride,
                                        _ptr(d_block_ids), d_pred.shape[0], d_pred.data_ptr(), _ptr(d_centre)))


def motion3d_learned(d_params, d_tpl, d_src, stride, d_disp, d_peak, radius=2, threshold=0.6, boundary_check=True,
                     d_path=None, d_pred=None, count_fallbacks=False):
    """Learned 3-D block matching; returns (ok, number of full searches or None)."""
    nz, ny, nx = d_tpl.shape
    n = C.c_int(-1)
    ok = bool(L.load().cut_motion3d_learned(d_params.data_ptr(), d_tpl.data_ptr(), d_src.data_ptr(), nx, ny, nz,
                                            stride, radius, threshold, int(bool(boundary_check)), d_disp.data_ptr(),
                                            d_peak.data_ptr(), _ptr(d_path), _ptr(d_pred),
                                            C.byref(n) if count_fallbacks else None))
    return ok, (n.value if count_fallbacks else None)


# ---- batches of frames ------------------------------------------------------
def beamform_batch(frames, mode: int = 0, latencies: bool = False):
    """frames: [(d_raw, d_volume, bp)] device tensors and parameters.  mode 0:
    persistent DAS launch inside a replayed HIP graph; 1: frame by frame;
    2: as 0 with the work kept in frame order (lower per-frame latency).
    Returns (ok, per-frame completion times in ms from the batch start or None)."""
    n = len(frames)
    arr = (L.CutFrame * n)()
    keep = []
    for i, (d_raw, d_vol, bp) in enumerate(frames):
        arr[i].d_raw = d_raw.data_ptr()
        arr[i].raw_bytes = d_raw.numel() * d_raw.element_size()
        arr[i].d_volume = d_vol.data_ptr()
        arr[i].volume_bytes = d_vol.numel() * d_vol.element_size()
        arr[i].bp = C.pointer(bp)
        keep.append(bp)
    ms = (C.c_double * n)() if latencies else None
    ok = bool(L.load().cut_beamform_batch(arr, n, mode, ms))
    return ok, (list(ms) if latencies else None)


def set_autotune(on: bool) -> bool:
    return bool(L.load().cut_set_autotune(int(bool(on))))


def set_fft_tile(values_per_thread: int) -> bool:
    """Hilbert FFT tile: 0 tuned, 8 or 16 values per thread."""
    return bool(L.load().cut_set_fft_tile(values_per_thread))


def tuning_save(path) -> bool:
    return bool(L.load().cut_tuning_save(str(path).encode()))


def tuning_load(path) -> bool:
    return bool(L.load().cut_tuning_load(str(path).encode()))


def batch_launch_shape() -> tuple[int, int]:
    g, o = C.c_int(), C.c_int()
    L.load().cut_batch_launch_shape(C.byref(g), C.byref(o))
    return g.value, o.value
