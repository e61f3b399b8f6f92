"""Pins the oracle's READI branch (CPU only).

With G groups, READI beamforms the same T decoded transmits G times, group g
fired from element rows g*T + t (HADAMARD) or t*G + g (WALSH), each group's
contribution signed by entry g of row `readi_group_id` of the order-G matrix,
WALSH reversing that entry's index on odd transmits
(templated_beamformer.cuh:43-63, beamformer.cu:238-252).  So the READI volume
of group id equals sum_g H_G[id][g] V_g, where V_g is a plain (non-READI)
volume whose element rows are moved to where group g fires them: an offset
of g*T pitches (HADAMARD), or a pitch of G elements offset by g, on even
transmits, and G-1-g, on odd ones (WALSH).  The plain volumes come from the
oracle's non-READI path and H_G from its Hadamard generator (pinned to the
reference's tables), so a wrong sign, group, row or parity rule in the READI
branch fails this.  Geometry in exact binary fractions (pitch 2^-12 m) so both
sides place the elements at identical coordinates.
"""
import numpy as np
import pytest

from cuda_ultrasound_toolkit_amd import _lib as L, configs

P = 2.0 ** -12


def _setup(T=8, C=8):
    acq = configs.rca_small(seed=31, T=T, C=C, n=10, generate=False)
    bp = acq.bp
    bp.xdc_element_pitch[:] = (P, P)
    bp.xdc_transform[12] = C * P / 2
    bp.xdc_transform[13] = T * P / 2
    bp.output_min_coordinate[0] = bp.output_min_coordinate[1] = -C * P / 2
    bp.output_max_coordinate[0] = bp.output_max_coordinate[1] = C * P / 2
    rng = np.random.default_rng(5)
    S = bp.dec_data_dim[0]
    iq = rng.standard_normal((T, C, S)) + 1j * rng.standard_normal((T, C, S))
    return acq, iq


def _plain(orc, bp0, iq, pitch_y, min_y_offset):
    """Non-READI volume with element row t at xdc_min_y + min_y_offset + t * pitch_y."""
    bp = L.CudaBeamformerParameters.from_buffer_copy(bp0)
    bp.readi_group_count = 0
    bp.readi_group_id = 0
    bp.xdc_element_pitch[1] = pitch_y
    # xdc_min_y = -xdc_transform[13] + pitch_y / 2 (beamformer.cu:17-98)
    min_y0 = -bp0.xdc_transform[13] + bp0.xdc_element_pitch[1] / 2
    bp.xdc_transform[13] = -(min_y0 + min_y_offset) + pitch_y / 2
    k = orc.das_constants(bp)
    assert k.xdc_min_y == np.float32(min_y0 + min_y_offset)
    return orc.das(iq, bp)[0]


@pytest.mark.parametrize("enc,G", [(L.HADAMARD, 2), (L.HADAMARD, 4), (L.WALSH, 2), (L.WALSH, 4)])
def test_readi_volume_is_signed_sum_of_group_volumes(orc, enc, G):
    acq, iq = _setup()
    bp0 = acq.bp
    T = bp0.dec_data_dim[2]
    H = orc.hadamard(G, enc)
    if enc == L.HADAMARD:
        V = [_plain(orc, bp0, iq, P, g * T * P) for g in range(G)]
    else:
        even, odd = iq.copy(), iq.copy()
        even[1::2] = 0
        odd[0::2] = 0
        V = [_plain(orc, bp0, even, G * P, g * P) + _plain(orc, bp0, odd, G * P, (G - 1 - g) * P) for g in range(G)]
    scale = max(np.abs(v).max() for v in V)
    for gid in range(G):
        bp = L.CudaBeamformerParameters.from_buffer_copy(bp0)
        bp.readi_group_count = G
        bp.readi_group_id = gid
        bp.decode = enc
        got = orc.das(iq, bp)[0]
        want = sum(H[gid, g] * V[g] for g in range(G))
        assert np.abs(got - want).max() <= 1e-12 * scale, (gid, np.abs(got - want).max() / scale)
    # the groups are distinguishable: a different row gives a different volume
    assert np.abs(sum(H[0, g] * V[g] for g in range(G)) - sum(H[1, g] * V[g] for g in range(G))).max() > 1e-3 * scale
