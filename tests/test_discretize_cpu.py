"""Sequence discretization (native hom_discretize, host only).  Invariants of
reference test/unit/test_koma_compatibility.jl:17-37 (node/interval counts,
nonnegative steps, a zero-length interval at a field jump, the trapezoid
integral of h0 across the jump, ADC times equal to the native times), step
limits, RF windows and flip angle, and agreement of the f64 oracle on the same
physical sequence discretized here and by the synthetic block builder.
Node placement beyond these rules is KomaMRIBase's in the reference (not in
its tree): parity unpinned."""
import numpy as np
import pytest

import oracle
from hom_amd import SeqBlock, discretize, get_kspace, rf_flip_angles
from hom_amd.types import GAMMA, BlochHighOrder, HO_Phantom
from hom_amd.workloads import FreeBlock, RFBlock, build_sequence


def test_field_jump_invariants():
    blocks = [SeqBlock(1e-3, fields={0: ([0, 1e-3], [1e-6, 1e-6])}),
              SeqBlock(1e-3, fields={0: ([0, 1e-3], [-2e-6, -2e-6])})]
    s = discretize(blocks, dt_max=2e-4)
    assert s.t[-1] == pytest.approx(2e-3)
    assert np.all(s.dt >= 0) and np.any(s.dt == 0)
    trapz = np.sum((s.h[0][1:] + s.h[0][:-1]) * s.dt / 2)
    assert abs(trapz - (-1e-9)) <= 1e-20
    assert s.t.size == s.dt.size + 1 and s.excitation.size == s.dt.size
    assert s.h.shape == (16, s.t.size)
    assert np.all(s.dt[s.dt > 0] <= 2e-4 * (1 + 1e-12))
    v = s.view(0, 2)
    assert v.t.size == 3 and v.dt.size == 2


def test_adc_times_steps_and_rf():
    adc = np.linspace(1e-5, 2e-3, 200)
    tau = 1e-4
    blocks = [SeqBlock.hard_rf(tau, 1.0 / (4 * GAMMA * tau)),
              SeqBlock(2e-3, gx=([0, 2e-4, 1.8e-3, 2e-3], [0, 5e-3, 5e-3, 0]), adc=adc),
              SeqBlock(5e-3)]
    s = discretize(blocks, dt_max=1e-3, dt_rf=2e-5)
    np.testing.assert_allclose(s.t[s.adc], tau + adc, rtol=0, atol=1e-15)
    np.testing.assert_allclose(rf_flip_angles(s), [90.0], rtol=1e-12)
    rf = s.excitation
    assert np.all(s.dt[rf] <= 2e-5 * (1 + 1e-12)) and s.dt[rf].sum() == pytest.approx(tau)
    assert np.all(s.dt[~rf] <= 1e-3 * (1 + 1e-12))
    # trapezoid gradient: the k-space at the end of the lobe is its area
    kn, kna, _, _ = get_kspace(s)
    area = 5e-3 * (2e-3 - 2e-4)
    assert kna[-1, 0] == pytest.approx(GAMMA * area, rel=1e-12)
    # measured first order: h1 = gx
    np.testing.assert_array_equal(s.h[1], s.gx)


def test_matches_block_builder_under_the_oracle():
    """The same RF + gradient + field sequence, discretized by hom_discretize
    and by the synthetic builder, gives the same f64 oracle signal."""
    rng = np.random.Generator(np.random.PCG64(4))
    n = 64
    obj = HO_Phantom(x=rng.uniform(-0.1, 0.1, n), y=rng.uniform(-0.1, 0.1, n),
                     z=rng.uniform(-0.03, 0.03, n), rho=rng.uniform(0.2, 1, n),
                     T1=rng.uniform(0.5, 2, n), T2=rng.uniform(0.02, 0.2, n),
                     dw=rng.normal(0, 2 * np.pi * 20, n))
    tau, dur, g = 1e-4, 3e-3, (2e-3, -1e-3, 5e-4)
    adc = np.arange(1, 301) * 1e-5
    h5 = 3e-4
    blocks = [SeqBlock.hard_rf(tau, 1.0 / (4 * GAMMA * tau), df=15.0, phase=0.3),
              SeqBlock(dur, gx=([0, dur], [g[0]] * 2), gy=([0, dur], [g[1]] * 2),
                       gz=([0, dur], [g[2]] * 2), adc=adc, fields={5: ([0, dur], [h5, h5])})]
    s1 = discretize(blocks, dt_max=1e-3, dt_rf=5e-5)

    def fields(t):
        h = np.zeros((16, t.size))
        h[5] = np.where(t >= tau, h5, 0.0)
        return h
    s2 = build_sequence([RFBlock(tau, 1.0 / (4 * GAMMA * tau), df=15.0, psi=0.3),
                         FreeBlock(dur, adc, lambda r: np.vstack([np.full(r.size, g[0]),
                                                                  np.full(r.size, g[1]),
                                                                  np.full(r.size, g[2])]))],
                        fields)
    m = BlochHighOrder("1111")
    a, _, _ = oracle.run_workload(obj, s1, m.flags, GAMMA)
    b, _, _ = oracle.run_workload(obj, s2, m.flags, GAMMA)
    assert a.shape == b.shape == (300, 1)
    assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)


def test_invalid_blocks():
    with pytest.raises(ValueError):
        discretize([SeqBlock(1e-3, adc=[2e-3])])        # ADC after the block
    with pytest.raises(ValueError):
        discretize([SeqBlock(1e-3, gx=([1e-3, 0.0], [1, 1]))])  # decreasing knots
    with pytest.raises(ValueError):
        discretize([SeqBlock(1e-3)], dt_max=0.0)
