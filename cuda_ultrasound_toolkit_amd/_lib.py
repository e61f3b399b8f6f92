"""ctypes view of libcuda_toolkit.so: struct mirrors and symbol prototypes.

The library is the product; this module only loads it.  It fails loudly when
the shared object is missing: there is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libcuda_toolkit.so"
# CUT_LIBRARY selects another build of the same library (the host-sanitizer
# build that the CPU suite runs its host-side tests against).
if os.environ.get("CUT_LIBRARY"):
    LIB_PATH = Path(os.environ["CUT_LIBRARY"]).resolve()

MAX_ACQ_COUNT = 256
MAX_CHANNEL_COUNT = 256
MAX_FILTER_LENGTH = 1024

# enums (cuda_toolkit/src/public/cuda_beamformer_parameters.h:15-89)
ORIENT_NONE, ORIENT_ROWS, ORIENT_COLS = 0, 1, 2
SAMPLE_NORMAL, SAMPLE_BANDPASS = 0, 1
FORCES, UFORCES, HERCULES, RCA_VLS, RCA_TPW = 0, 1, 2, 3, 4
TYPE_I16, TYPE_F32, TYPE_F32C, TYPE_U8 = 1, 2, 3, 4
NONE, HADAMARD, WALSH = 0, 1, 2
RX_HANN, RX_TO_SIN, TX_TO_SIN, BOTH_TO_SIN = 0, 1, 2, 3


class CudaBeamformerParameters(C.Structure):
    _fields_ = [
        ("xdc_transform", C.c_float * 16),
        ("xdc_element_pitch", C.c_float * 2),
        ("rf_raw_dim", C.c_uint * 4),
        ("dec_data_dim", C.c_uint * 4),
        ("decode", C.c_int),
        ("das_shader_id", C.c_int),
        ("time_offset", C.c_float),
        ("output_points", C.c_uint * 4),
        ("output_min_coordinate", C.c_float * 4),
        ("output_max_coordinate", C.c_float * 4),
        ("sampling_frequency", C.c_float),
        ("center_frequency", C.c_float),
        ("speed_of_sound", C.c_float),
        ("off_axis_pos", C.c_float),
        ("beamform_plane", C.c_int),
        ("fn_tx", C.c_float),
        ("fn_rx", C.c_float),
        ("interpolate", C.c_bool),
        ("coherency_weighting", C.c_float),
        ("channel_mapping", C.c_short * MAX_CHANNEL_COUNT),
        ("sparse_elements", C.c_short * MAX_CHANNEL_COUNT),
        ("foci", C.c_float * (MAX_ACQ_COUNT * 3)),
        ("tx_orientation", C.c_int),
        ("rx_orientation", C.c_int),
        ("readi_group_count", C.c_uint),
        ("readi_group_id", C.c_uint),
        ("mixes_count", C.c_uint),
        ("mixes_offset", C.c_uint),
        ("mixes_rows", C.c_uint * 128),
        ("filter_length", C.c_uint),
        ("rf_filter", C.c_float * MAX_FILTER_LENGTH),
        ("data_type", C.c_int),
        ("apo_type", C.c_int),
        ("to_power", C.c_int),
        ("sample_type", C.c_int),
    ]


class NccMotionParameters(C.Structure):
    _fields_ = [
        ("patch_size", C.c_uint),
        ("motion_grid_spacing", C.c_uint),
        ("motion_grid_dims", C.c_uint * 2),
        ("scale_dims", C.c_uint * 2),
        ("search_margins", C.c_uint * 2),
        ("abs_cor_threshold", C.c_float),
        ("rel_cor_threshold", C.c_float),
        ("min_patch_variance", C.c_float),
        ("use_subpixel", C.c_bool),
        ("reference_frame", C.c_uint),
        ("neighbour_compare", C.c_bool),
        ("frame_count", C.c_uint),
        ("image_dims", C.c_uint * 2),
        ("data_type", C.c_int),
    ]


class CorrImagesParameters(C.Structure):
    _fields_ = [("template_dims", C.c_uint * 2), ("source_dims", C.c_uint * 2)]


class CutFrame(C.Structure):
    _fields_ = [
        ("d_raw", C.c_void_p),
        ("raw_bytes", C.c_size_t),
        ("d_volume", C.c_void_p),
        ("volume_bytes", C.c_size_t),
        ("bp", C.POINTER(CudaBeamformerParameters)),
    ]


class CutSynthSpec(C.Structure):
    _fields_ = [
        ("geometry", C.c_int),
        ("T", C.c_int), ("C", C.c_int), ("S", C.c_int),
        ("n_scatterers", C.c_int),
        ("pitch", C.c_float),
        ("c", C.c_float), ("fc", C.c_float), ("fs", C.c_float),
        ("box_min", C.c_float * 3), ("box_max", C.c_float * 3),
        ("amp_min", C.c_float), ("amp_max", C.c_float),
        ("noise_frac", C.c_float),
        ("peak", C.c_float),
        ("encoding", C.c_int),
        ("cycles", C.c_int),
        ("seed", C.c_ulonglong),
        ("shift", C.c_float * 3),
        ("decorrelation", C.c_float),
        ("amp_seed", C.c_ulonglong),
    ]


class Span(C.Structure):
    """std::span<T> with dynamic extent as laid out by libstdc++: {pointer, size}."""
    _fields_ = [("data", C.c_void_p), ("size", C.c_size_t)]


SYM_BEAMFORM = "_ZN12cuda_toolkit8beamformESt4spanIKhLm18446744073709551615EES0_IhLm18446744073709551615EERK24CudaBeamformerParameters"
SYM_MOTION = "_ZN12cuda_toolkit16motion_detectionESt4spanIKhLm18446744073709551615EES0_IhLm18446744073709551615EERK19NccMotionParameters"
SYM_CORR = "_ZN12cuda_toolkit11corr_imagesESt4spanIKfLm18446744073709551615EES2_S0_IfLm18446744073709551615EEPKjS5_"

_lib = None


def _proto(lib, name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


def load() -> C.CDLL:
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or python -m "
                           "cuda_ultrasound_toolkit_amd.build) first; there is no fallback path")
    # torch ships its own libamdhip64.so.7 / libhsa-runtime64.so.1 under the
    # same sonames as /opt/rocm's, and whichever copy loads first serves the
    # whole process.  torch does not initialise on a newer runtime loaded
    # before it (torch.cuda.is_available() reads False), so when torch is
    # installed it is imported first and the library binds to its runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(LIB_PATH))
    B, V, P = C.c_bool, C.c_void_p, C.POINTER
    u = C.c_uint
    _proto(lib, SYM_BEAMFORM, B, [Span, Span, P(CudaBeamformerParameters)])
    _proto(lib, SYM_MOTION, B, [Span, Span, P(NccMotionParameters)])
    _proto(lib, SYM_CORR, B, [Span, Span, Span, P(u), P(u)])
    _proto(lib, "motion_detect_f32_matlab", None, [P(C.c_float), NccMotionParameters, P(C.c_float)])
    _proto(lib, "corr_images_f32_matlab", None, [P(C.c_float), P(C.c_float), CorrImagesParameters, P(C.c_float)])
    _proto(lib, "init_cuda_configuration", B, [P(u), P(u)])
    _proto(lib, "register_cuda_buffers", B, [P(u), u, u])
    _proto(lib, "cuda_set_channel_mapping", B, [P(C.c_int16)])
    _proto(lib, "cuda_set_match_filter", B, [P(C.c_float), u])
    _proto(lib, "cuda_hilbert", B, [u, u])
    _proto(lib, "cuda_decode", B, [C.c_size_t, u])
    _proto(lib, "deinit_cuda_configuration", None, [])
    _proto(lib, "beamform_i16_matlab", None, [P(C.c_int16), CudaBeamformerParameters, P(C.c_float)])
    _proto(lib, "beamform_f32_matlab", None, [P(C.c_float), CudaBeamformerParameters, P(C.c_float)])
    _proto(lib, "cut_set_stream", None, [V])
    _proto(lib, "cut_synchronize", B, [])
    _proto(lib, "cut_version", C.c_char_p, [])
    _proto(lib, "cut_set_decode_mode", B, [C.c_int])
    _proto(lib, "cut_set_das_options", B, [C.c_int, C.c_int, C.c_int])
    _proto(lib, "cut_beamform_device", B, [V, C.c_size_t, V, C.c_size_t, P(CudaBeamformerParameters)])
    _proto(lib, "cut_beamform_batch", B, [P(CutFrame), C.c_int, C.c_int, P(C.c_double)])
    _proto(lib, "cut_set_autotune", B, [C.c_int])
    _proto(lib, "cut_set_fft_tile", B, [C.c_int])
    _proto(lib, "cut_tuning_save", B, [C.c_char_p])
    _proto(lib, "cut_tuning_load", B, [C.c_char_p])
    _proto(lib, "cut_batch_launch_shape", B, [P(C.c_int), P(C.c_int)])
    _proto(lib, "cut_rf_init", B, [P(u), P(u), C.c_int])
    _proto(lib, "cut_rf_set_channel_mapping", B, [P(C.c_int16), u])
    _proto(lib, "cut_rf_set_match_filter", B, [P(C.c_float), u])
    _proto(lib, "cut_rf_convert_decode", B, [V, C.c_int, V, C.c_int, C.c_int])
    _proto(lib, "cut_rf_convert_demod", B, [V, C.c_int, V, C.c_float, C.c_float])
    _proto(lib, "cut_rf_hilbert", B, [V, C.c_int, V])
    _proto(lib, "cut_das_device", B, [V, V, P(CudaBeamformerParameters)])
    _proto(lib, "cut_das_device_noalloc", B, [V, V, P(CudaBeamformerParameters)])
    _proto(lib, "cut_rf_hilbert_format", B, [V, C.c_int, V, C.c_int, P(C.c_float)])
    _proto(lib, "cut_das_device_format", B, [V, C.c_int, C.c_float, V, P(CudaBeamformerParameters)])
    _proto(lib, "cut_das_active_taps", B, [P(CudaBeamformerParameters), P(C.c_ulonglong)])
    _proto(lib, "cut_ncc3d_search", B, [V, V, C.c_int, C.c_int, C.c_int, C.c_int, V, C.c_int, V, C.c_int, V, V])
    _proto(lib, "cut_envelope", B, [V, C.c_size_t, V])
    F = C.c_float
    _proto(lib, "cut_pnet_param_count", C.c_int, [])
    _proto(lib, "cut_pnet_set_precision", B, [C.c_int])
    _proto(lib, "cut_pnet_set_input_mode", B, [C.c_int])
    _proto(lib, "cut_pnet_gather", B, [V, C.c_int, C.c_int, C.c_int, C.c_int, V, C.c_int, V])
    _proto(lib, "cut_pnet_forward", B, [V, V, C.c_int, V])
    _proto(lib, "cut_pnet_loss_grad", B, [V, V, V, C.c_int, V, V])
    _proto(lib, "cut_pnet_adam", B, [V, V, V, V, C.c_int, F, F, F, F, C.c_int, F])
    _proto(lib, "cut_pnet_infer", B, [V, V, V, C.c_int, C.c_int, C.c_int, C.c_int, V, C.c_int, V, V])
    _proto(lib, "cut_motion3d_learned", B, [V, V, V, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, F, C.c_int, V, V,
                                            V, V, P(C.c_int)])
    _proto(lib, "cut_gl_register_device_buffers", B, [P(V), P(C.c_size_t), u, V, C.c_size_t])
    _proto(lib, "cut_synth_rf", B, [P(CutSynthSpec), P(C.c_int16), P(C.c_int16)])
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    """Names every header under include/ declares (checked by the CPU test suite)."""
    return [
        SYM_BEAMFORM, SYM_MOTION, SYM_CORR, "motion_detect_f32_matlab", "corr_images_f32_matlab",
        "init_cuda_configuration", "register_cuda_buffers", "cuda_set_channel_mapping",
        "cuda_set_match_filter", "cuda_hilbert", "cuda_decode", "deinit_cuda_configuration",
        "beamform_i16_matlab", "beamform_f32_matlab",
        "cut_set_stream", "cut_synchronize", "cut_version", "cut_set_decode_mode", "cut_set_das_options",
        "cut_beamform_device", "cut_rf_init", "cut_rf_set_channel_mapping", "cut_rf_set_match_filter",
        "cut_rf_convert_decode", "cut_rf_convert_demod", "cut_rf_hilbert", "cut_das_device", "cut_das_active_taps", "cut_gl_register_device_buffers",
        "cut_das_device_noalloc",
        "cut_ncc3d_search", "cut_envelope", "cut_rf_hilbert_format", "cut_das_device_format",
        "cut_beamform_batch", "cut_set_autotune", "cut_batch_launch_shape",
        "cut_set_fft_tile", "cut_tuning_save", "cut_tuning_load",
        "cut_pnet_param_count", "cut_pnet_set_precision", "cut_pnet_set_input_mode", "cut_pnet_gather", "cut_pnet_forward", "cut_pnet_loss_grad", "cut_pnet_adam",
        "cut_pnet_infer", "cut_motion3d_learned",
        "cut_synth_rf",
    ]
