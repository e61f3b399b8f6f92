"""Out-of-process use of the library through the transfer protocol
(include/cut_transfer.h; the Linux form of the reference harness's
test_app/src/communication_params.h and test_app/client_lib/cuda_transfer.h).

``start_server`` launches ``cut_transfer_server`` as a child process;
``Client`` calls the C client library ``libcut_transfer.so`` (which needs no
GPU) and returns numpy arrays.  The ctypes mirrors of the wire structs are
checked against the C layout in the tests.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from . import _lib as L

HERE = Path(__file__).resolve().parent
# CUT_TRANSFER_SERVER / CUT_TRANSFER_CLIENT select other builds of the two
# (the host-sanitizer build)
SERVER = Path(os.environ.get("CUT_TRANSFER_SERVER", HERE / "cut_transfer_server")).resolve()
CLIENT = Path(os.environ.get("CUT_TRANSFER_CLIENT", HERE / "libcut_transfer.so")).resolve()

# CudaCommand (communication_params.h:21-30)
ERR, ACK, SUCCESS, BEAMFORM_VOLUME, SVD_FILTER, NCC_MOTION_DETECT, CORR_IMAGES = range(7)


class CommandPipeMessage(C.Structure):
    _fields_ = [("opcode", C.c_int), ("data_size", C.c_ulonglong), ("frame_count", C.c_int)]


class SVDParameters(C.Structure):
    _fields_ = [("frame_count", C.c_int), ("filter_indicies", C.c_int * 256)]


class SharedMemoryParams(C.Structure):
    _fields_ = [
        ("beamformerParameters", L.CudaBeamformerParameters),
        ("svdParameters", SVDParameters),
        ("nccMotionParameters", L.NccMotionParameters),
        ("corrImagesParameters", L.CorrImagesParameters),
    ]


def start_server(name: str, data_bytes: int = 1 << 31, idle_timeout_ms: int = 0, max_commands: int = 0,
                 timeout_s: float = 30.0) -> subprocess.Popen:
    """Start cut_transfer_server and wait for its ready line."""
    if not SERVER.exists():
        raise RuntimeError(f"{SERVER} is missing: build the package first")
    cmd = [str(SERVER), "--name", name, "--data-bytes", str(int(data_bytes)),
           "--idle-timeout-ms", str(int(idle_timeout_ms)), "--max-commands", str(int(max_commands))]
    proc = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=None, text=True)
    line = proc.stdout.readline()
    if "ready" not in line:
        proc.kill()
        proc.wait(timeout=timeout_s)
        raise RuntimeError(f"cut_transfer_server did not start: {line!r}")
    return proc


def stop_server(proc: subprocess.Popen, timeout_s: float = 30.0) -> int:
    if proc.poll() is None:
        proc.terminate()
    return proc.wait(timeout=timeout_s)


class Client:
    """The C client library bound to one server name (CUT_TRANSFER_NAME)."""

    def __init__(self, name: str, timeout_ms: int | None = None):
        if not CLIENT.exists():
            raise RuntimeError(f"{CLIENT} is missing: build the package first")
        os.environ["CUT_TRANSFER_NAME"] = name
        if timeout_ms is not None:
            os.environ["CUT_TRANSFER_TIMEOUT_MS"] = str(int(timeout_ms))
        lib = C.CDLL(str(CLIENT))
        fp, P = C.POINTER(C.c_float), C.POINTER
        lib.beamform_i16.argtypes = [P(C.c_short), L.CudaBeamformerParameters, fp]
        lib.beamform_f32.argtypes = [fp, L.CudaBeamformerParameters, fp]
        lib.motion_detect_f32.argtypes = [fp, L.NccMotionParameters, fp]
        lib.corr_images_f32.argtypes = [fp, fp, L.CorrImagesParameters, fp]
        for f in ("beamform_i16", "beamform_f32", "motion_detect_f32", "corr_images_f32", "deinit"):
            getattr(lib, f).restype = None
        lib.cut_transfer_last_ok.restype = C.c_int
        self.lib = lib

    @property
    def last_ok(self) -> bool:
        return bool(self.lib.cut_transfer_last_ok())

    def beamform(self, raw: np.ndarray, bp: L.CudaBeamformerParameters) -> np.ndarray | None:
        """raw int16 or float32 [rf_raw_dim[1]][rf_raw_dim[0]] -> complex64 [Nz][Ny][Nx]."""
        nx, ny, nz = bp.output_points[0], bp.output_points[1], bp.output_points[2]
        out = np.zeros(nz * ny * nx * 2, np.float32)
        fp = out.ctypes.data_as(C.POINTER(C.c_float))
        if raw.dtype == np.int16:
            r = np.ascontiguousarray(raw)
            self.lib.beamform_i16(r.ctypes.data_as(C.POINTER(C.c_short)), bp, fp)
        else:
            r = np.ascontiguousarray(raw, dtype=np.float32)
            self.lib.beamform_f32(r.ctypes.data_as(C.POINTER(C.c_float)), bp, fp)
        return out.view(np.complex64).reshape(nz, ny, nx) if self.last_ok else None

    def motion_detect(self, images: np.ndarray, params: L.NccMotionParameters) -> np.ndarray | None:
        """f32 frames -> float4 maps [frames][grid_y][grid_x][4]."""
        gx, gy = params.motion_grid_dims[0], params.motion_grid_dims[1]
        out = np.zeros((params.frame_count, gy, gx, 4), np.float32)
        im = np.ascontiguousarray(images, dtype=np.float32)
        self.lib.motion_detect_f32(im.ctypes.data_as(C.POINTER(C.c_float)), params,
                                   out.ctypes.data_as(C.POINTER(C.c_float)))
        return out if self.last_ok else None

    def corr_images(self, template: np.ndarray, source: np.ndarray) -> np.ndarray | None:
        """2-D float32 images [rows][cols]; dims are passed as [width, height]."""
        t = np.ascontiguousarray(template, dtype=np.float32)
        s = np.ascontiguousarray(source, dtype=np.float32)
        p = L.CorrImagesParameters()
        p.template_dims[:] = (t.shape[1], t.shape[0])
        p.source_dims[:] = (s.shape[1], s.shape[0])
        out = np.zeros((max(s.shape[0] - t.shape[0] + 1, 1), max(s.shape[1] - t.shape[1] + 1, 1)), np.float32)
        self.lib.corr_images_f32(t.ctypes.data_as(C.POINTER(C.c_float)), s.ctypes.data_as(C.POINTER(C.c_float)), p,
                                 out.ctypes.data_as(C.POINTER(C.c_float)))
        return out if self.last_ok else None

    def deinit(self) -> None:
        self.lib.deinit()
