"""The DAS work term (CPU only).

`cut_das_active_taps` is the host-side count of (voxel, transmit, channel)
taps whose apodization passes the sequence's threshold; the benchmark's
roofline figure is that count times the FLOP per tap.  It is computed by the
library independently of the oracle (a closed-form aperture test instead of
the oracle's per-tap weight), so the two counts agreeing pins it.
"""
import ctypes as C

import numpy as np
import pytest

from cuda_ultrasound_toolkit_amd import _lib as L, configs


def lib_taps(lib, bp):
    n = C.c_ulonglong()
    assert lib.cut_das_active_taps(C.byref(bp), C.byref(n))
    return n.value


def oracle_taps(orc, acq):
    bp = acq.bp
    S, Cc, T = bp.dec_data_dim[0], bp.dec_data_dim[1], bp.dec_data_dim[2]
    iq = np.zeros((T, Cc, S), np.complex128)
    _, act = orc.das(iq, bp, cubic=False)
    return act


def _hercules_plane():
    return configs.rca_small(seed=3, generate=False)


def _hercules_focus():
    return configs.rca_small(seed=3, generate=False, focus_z=8e-3)


def _hercules_sin():
    a = configs.rca_small(seed=3, generate=False)
    a.bp.apo_type = L.RX_TO_SIN
    a.bp.to_power = 2
    return a


def _hercules_readi(G, enc):
    def make():
        a = configs.rca_small(seed=3, generate=False, T=16)
        a.bp.readi_group_count = G
        a.bp.readi_group_id = 1
        a.bp.decode = enc
        return a
    return make


def _hercules_no_fnumber():
    a = configs.rca_small(seed=3, generate=False, T=8, n=12)
    a.bp.fn_rx = 0.0
    return a


def _hercules_above_array(fn):
    """Depths from -3.125 to 8.875 mm: voxels above the array (z < 0), where
    the reference's cos^2(pi min(fn rho / 2z, 1/2)) oscillates in |a| instead
    of falling to zero, and F-number 0 or 1."""
    def make():
        a = configs.rca_small(seed=3, generate=False, T=8, n=14, z_range=(-3.125e-3, 8.875e-3))
        a.bp.fn_rx = fn
        return a
    return make


def _hercules_negative_fnumber():
    a = configs.rca_small(seed=3, generate=False, T=8, n=12)
    a.bp.fn_rx = -1.0
    return a


def _hercules_rca3d_subgrid():
    """The benchmark geometry (128 x 128 aperture, 5-60 mm) on a coarse grid."""
    a = configs.rca_3d(seed=2, generate=False, T=128, n=12, S=4096)
    return a


def _forces(fn):
    def make():
        a = configs.linear_2d(seed=1, generate=False, nx=32, nz=48)
        a.bp.fn_rx = a.bp.fn_tx = fn
        return a
    return make


def _uforces():
    return configs.uforces_2d(seed=6, generate=False, nx=32, nz=48)


def _tpw():
    return configs.tpw_3d(seed=7, generate=False)


CASES = {
    "hercules_plane": _hercules_plane,
    "hercules_focus": _hercules_focus,
    "hercules_rx_to_sin": _hercules_sin,
    "hercules_readi_g2_hadamard": _hercules_readi(2, L.HADAMARD),
    "hercules_readi_g4_walsh": _hercules_readi(4, L.WALSH),
    "hercules_no_fnumber": _hercules_no_fnumber,
    "hercules_above_array_fn1": _hercules_above_array(1.0),
    "hercules_above_array_fn0": _hercules_above_array(0.0),
    "hercules_negative_fnumber": _hercules_negative_fnumber,
    "hercules_rca3d_geometry": _hercules_rca3d_subgrid,
    "forces_fn0": _forces(0.0),
    "forces_fn1": _forces(1.0),
    "uforces": _uforces,
    "rca_tpw": _tpw,
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_active_taps_equal_oracle_count(lib, orc, case):
    acq = CASES[case]()
    got, want = lib_taps(lib, acq.bp), oracle_taps(orc, acq)
    bp = acq.bp
    dense = (bp.output_points[0] * bp.output_points[1] * bp.output_points[2] * bp.dec_data_dim[1] *
             bp.dec_data_dim[2] * max(1, bp.readi_group_count))
    print(f"{case}: {got} active taps of {dense} dense")
    assert got == want
    assert 0 < got <= dense


def test_oracle_plane_subset_equals_full_volume_planes(orc):
    """or_das_planes evaluates the listed depth planes of the grid exactly as
    the whole-volume call does."""
    acq = configs.rca_small(seed=8, T=8, n=10)
    bp = acq.bp
    S, Cc, T = bp.dec_data_dim[0], bp.dec_data_dim[1], bp.dec_data_dim[2]
    rng = np.random.default_rng(3)
    iq = rng.standard_normal((T, Cc, S)) + 1j * rng.standard_normal((T, Cc, S))
    full, act_full = orc.das(iq, bp)
    planes = [7, 0, 9, 3]
    sub, act_sub = orc.das(iq, bp, planes=planes)
    assert np.array_equal(sub, full[planes])
    assert 0 < act_sub < act_full
    with pytest.raises(ValueError):
        orc.das(iq, bp, planes=[10])
