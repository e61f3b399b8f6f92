"""Transfer protocol on the CPU: wire layout, the server's answers to requests
it must refuse (no GPU work is reached), client behaviour without a server,
and clean shutdown.  The GPU round trips are in test_gpu_transfer.py."""
import ctypes as C
import os
import socket
import struct
import time

import numpy as np
import pytest

from cuda_ultrasound_toolkit_amd import _lib as L, transfer as T


def _name(tag):
    return f"cut_t{os.getpid()}_{tag}"


def test_wire_layout():
    # communication_params.h: enum (4) + pad (4) + u64 + int + pad = 24
    assert C.sizeof(T.CommandPipeMessage) == 24
    assert T.CommandPipeMessage.data_size.offset == 8
    assert T.CommandPipeMessage.frame_count.offset == 16
    assert C.sizeof(T.SVDParameters) == 1028
    assert C.sizeof(T.SharedMemoryParams) == 8948 + 1028 + 72 + 16
    assert T.SharedMemoryParams.svdParameters.offset == 8948
    assert T.SharedMemoryParams.nccMotionParameters.offset == 8948 + 1028
    assert T.SharedMemoryParams.corrImagesParameters.offset == 8948 + 1028 + 72
    assert (T.ERR, T.ACK, T.SUCCESS, T.BEAMFORM_VOLUME, T.SVD_FILTER, T.NCC_MOTION_DETECT, T.CORR_IMAGES) == \
        (0, 1, 2, 3, 4, 5, 6)


def test_client_exports_reference_api():
    lib = C.CDLL(str(T.CLIENT))
    for sym in ("beamform_i16", "beamform_f32", "motion_detect_f32", "corr_images_f32", "deinit",
                "cut_transfer_last_ok"):
        assert hasattr(lib, sym), sym


def _send(s, op, size):
    s.send(struct.pack("<iiQi4x", op, 0, size, 0))
    d = s.recv(64)
    assert len(d) == 24
    op, _, n, _ = struct.unpack("<iiQi4x", d)
    return op, n


def test_server_refuses_without_touching_the_gpu():
    name = _name("refuse")
    srv = T.start_server(name, data_bytes=1 << 20, idle_timeout_ms=20000)
    try:
        assert os.path.exists(f"/dev/shm/{name}_data") and os.path.getsize(f"/dev/shm/{name}_data") == 1 << 20
        assert os.path.getsize(f"/dev/shm/{name}_parameters") == C.sizeof(T.SharedMemoryParams)
        s = socket.socket(socket.AF_UNIX, socket.SOCK_SEQPACKET)
        s.settimeout(20)
        s.connect(f"/tmp/{name}_command.sock")
        assert _send(s, T.BEAMFORM_VOLUME, 2 << 20) == (T.ERR, 0)   # larger than the data segment
        assert _send(s, T.SVD_FILTER, 0) == (T.ERR, 0)              # not implemented by the reference
        assert _send(s, 99, 0) == (T.ERR, 0)                        # unknown opcode
        assert _send(s, T.CORR_IMAGES, 0) == (T.ERR, 0)             # zero dimensions
        # corr_images whose input size does not match the dimensions
        with open(f"/dev/shm/{name}_parameters", "r+b") as f:
            p = T.SharedMemoryParams()
            p.corrImagesParameters.template_dims[:] = (4, 4)
            p.corrImagesParameters.source_dims[:] = (8, 8)
            f.write(bytes(p))
        assert _send(s, T.CORR_IMAGES, 16) == (T.ERR, 0)
        s.send(b"xyz")                                              # malformed message
        op, _, _, _ = struct.unpack("<iiQi4x", s.recv(64))
        assert op == T.ERR
        s.close()
    finally:
        assert T.stop_server(srv) == 0
    # segments and socket are removed at exit
    assert not os.path.exists(f"/dev/shm/{name}_data")
    assert not os.path.exists(f"/dev/shm/{name}_parameters")
    assert not os.path.exists(f"/tmp/{name}_command.sock")


def test_client_reports_refusals_and_missing_server(capfd):
    name = _name("client")
    cl = T.Client(name, timeout_ms=20000)
    # no server: fails at once, without hanging
    t0 = time.time()
    assert cl.corr_images(np.ones((4, 4), np.float32), np.ones((8, 8), np.float32)) is None
    assert time.time() - t0 < 5
    assert "cut_transfer" in capfd.readouterr().err
    # invalid dimensions are refused by the client before any transfer
    assert cl.corr_images(np.ones((9, 4), np.float32), np.ones((8, 8), np.float32)) is None
    srv = T.start_server(name, data_bytes=1 << 16, idle_timeout_ms=20000)
    try:
        # input larger than the (small) data segment: refused on the client side
        bp = L.CudaBeamformerParameters()
        bp.rf_raw_dim[:] = (1 << 16, 2, 0, 0)
        bp.output_points[:] = (2, 2, 2, 1)
        assert cl.beamform(np.zeros((2, 1 << 16), np.int16), bp) is None
        assert not cl.last_ok
        # zero-size output
        bp.rf_raw_dim[:] = (16, 2, 0, 0)
        bp.output_points[:] = (0, 2, 2, 1)
        assert cl.beamform(np.zeros((2, 16), np.int16), bp) is None
    finally:
        cl.deinit()
        assert T.stop_server(srv) == 0


def test_server_idle_timeout_and_command_budget():
    name = _name("idle")
    srv = T.start_server(name, data_bytes=4096, idle_timeout_ms=300)
    assert srv.wait(timeout=20) == 0
    srv = T.start_server(name, data_bytes=4096, max_commands=1, idle_timeout_ms=20000)
    s = socket.socket(socket.AF_UNIX, socket.SOCK_SEQPACKET)
    s.settimeout(20)
    s.connect(f"/tmp/{name}_command.sock")
    assert _send(s, T.SVD_FILTER, 0) == (T.ERR, 0)
    s.close()
    assert srv.wait(timeout=20) == 0
    assert not os.path.exists(f"/dev/shm/{name}_data")
