"""Encoding trajectories (host code in libhom.so, no GPU): get_kspace
(reference src/datatypes/HO_Sequence.jl:221-287) against closed forms on
discretized sequences: constant and ramp gradients integrate exactly
under the trapezoid rule, an excitation restarts the integral at the pulse,
a refocusing pulse negates it, skip_rf continues through it, ADC rows are the
values at the ADC nodes, shapes as test/unit/test_koma_compatibility.jl:46-53.
The RF kind and break point come from KomaMRIBase in the reference (not in
its tree): parity unpinned beyond these closed forms."""
import numpy as np
import pytest

from hom_amd import get_kspace, rf_flip_angles
from hom_amd.types import GAMMA, DimensionMismatch, HO_DiscreteSequence
from hom_amd.workloads import FreeBlock, RFBlock, build_sequence


def grad_const(g):
    def f(rel):
        return np.vstack([np.full(rel.size, g[0]), np.full(rel.size, g[1]), np.full(rel.size, g[2])])
    return f


def test_constant_and_ramp_gradients_integrate_exactly():
    g = (3e-3, -2e-3, 1e-3)
    adc = np.array([1e-4, 2.5e-4, 7e-4, 1e-3])
    seq = build_sequence([FreeBlock(1e-3, adc, grad_const(g))], dt_max=1.7e-4)
    kn, kna, kd, kda = get_kspace(seq)
    t_end = seq.t[1:]
    assert kn.shape == (seq.n_nodes - 1, 3) and kd.shape == (seq.n_nodes - 1, 16)
    assert kna.shape == (4, 3) and kda.shape == (4, 16)
    for j in range(3):
        np.testing.assert_allclose(kn[:, j], GAMMA * g[j] * t_end, rtol=1e-13, atol=1e-9)
    np.testing.assert_allclose(kd[:, 1:4], kn, rtol=1e-14)  # measured h1..h3 = nominal here
    assert np.all(kd[:, 0] == 0) and np.all(kd[:, 4:] == 0)
    np.testing.assert_allclose(kna, GAMMA * np.outer(adc, g), rtol=1e-12)
    np.testing.assert_allclose(kda[:, 1:4], kna, rtol=1e-14)

    a = 40.0  # T/m/s ramp in x
    ramp = build_sequence([FreeBlock(1e-3, adc, lambda rel: np.vstack([a * rel, 0 * rel, 0 * rel]))],
                          dt_max=1.3e-4)
    kn, kna, _, _ = get_kspace(ramp)
    np.testing.assert_allclose(kn[:, 0], GAMMA * a * ramp.t[1:] ** 2 / 2, rtol=1e-12)
    np.testing.assert_allclose(kna[:, 0], GAMMA * a * adc ** 2 / 2, rtol=1e-12)


def lobes(flip_deg, area_g=(2e-3, 1e-3, 0.0)):
    """Gradient lobe A (1 ms), hard RF of the given flip, lobe B (1 ms)."""
    tau = 1e-4
    b1 = flip_deg / 360.0 / (GAMMA * tau)
    gb = (-1e-3, 4e-3, 2e-3)
    return build_sequence([FreeBlock(1e-3, np.array([1e-3]), grad_const(area_g)),
                           RFBlock(tau, b1),
                           FreeBlock(1e-3, np.array([5e-4, 1e-3]), grad_const(gb))]), area_g, gb


def test_excitation_refocusing_and_skip():
    for flip, sign in ((90.0, 0.0), (45.0, 0.0), (180.0, -1.0), (120.0, -1.0)):
        seq, ga, gb = lobes(flip)
        np.testing.assert_allclose(rf_flip_angles(seq), [flip], rtol=1e-12)
        kn, kna, kd, kda = get_kspace(seq)
        for j in range(3):
            end_a = GAMMA * ga[j] * 1e-3
            expected = sign * end_a + GAMMA * gb[j] * np.array([5e-4, 1e-3])
            np.testing.assert_allclose(kna[1:, j], expected, rtol=1e-12, atol=1e-9)
            np.testing.assert_allclose(kna[0, j], end_a, rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(kda[:, 1 + j], kna[:, j], rtol=1e-14, atol=1e-12)
        kn2, kna2, _, _ = get_kspace(seq, skip_rf=[True])
        for j in range(3):
            expected = GAMMA * ga[j] * 1e-3 + GAMMA * gb[j] * np.array([5e-4, 1e-3])
            np.testing.assert_allclose(kna2[1:, j], expected, rtol=1e-12, atol=1e-9)
        # the ADC rows are the per-interval values at the ADC nodes
        nodes = np.flatnonzero(seq.adc)
        np.testing.assert_array_equal(kna, kn[nodes - 1])


def test_errors_and_zero_encoding():
    seq, _, _ = lobes(90.0)
    with pytest.raises(DimensionMismatch):
        get_kspace(seq, skip_rf=[True, False])
    one = HO_DiscreteSequence(t=[0.0], gx=[0.0], gy=[0.0], gz=[0.0], b1=[0j], df=[0.0],
                              psi=[0.0], adc=[False], excitation=np.zeros(0, bool),
                              h=np.zeros((16, 1)))
    with pytest.raises(ValueError):
        get_kspace(one)
    # no gradients and no fields: every encoding is zero (reference
    # test_koma_compatibility.jl:46-53)
    quiet = build_sequence([RFBlock(1e-4, 1.0 / (4 * GAMMA * 1e-4)),
                            FreeBlock(1e-3, np.linspace(2e-4, 1e-3, 5), None)])
    kn, kna, kd, kda = get_kspace(quiet)
    assert kna.shape == (5, 3) and kda.shape == (5, 16)
    assert not kn.any() and not kna.any() and not kd.any() and not kda.any()
