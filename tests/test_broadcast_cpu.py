"""The reference CPU algorithm restated in broadcast form (oracle/broadcast.py,
the bench's reference-algorithm CPU leg) against the per-spin C oracle: the
two restate the same reference steps in different evaluation orders, so the
signal and the final state agree to rounding, including RF blocks, every
method flag and several thread parts."""
import numpy as np
import pytest

import oracle
from oracle import broadcast
from hom_amd.types import GAMMA, BlochHighOrder, get_sim_ranges
from hom_amd.workloads import high_order_workload, make_config
from helpers import nrmse


@pytest.mark.parametrize("index,stride,mask,threads", [
    (0, 7, "1111", 1),      # 2D, 1 coil, 90-degree pulse + spiral
    (1, 3000, "1111", 3),   # 3D, 8 coils, 4 x (RF + readout + idle)
    (1, 3000, "0110", 2),   # nominal gradients in place of order one
])
def test_broadcast_matches_oracle(index, stride, mask, threads):
    obj, seq, _ = make_config(index, spin_subsample=stride)
    parts = get_sim_ranges(seq, 32, 32)
    m = BlochHighOrder(mask)
    ref, mxy_ref, mz_ref = oracle.run(obj, seq, parts, m.flags, GAMMA)
    sig, mxy, mz = broadcast.run(obj, seq, parts, m.flags, GAMMA, n_threads=threads)
    assert sig.shape == ref.shape
    assert nrmse(sig, ref) <= 1e-13
    np.testing.assert_allclose(mxy, mxy_ref, rtol=0, atol=1e-13)
    np.testing.assert_allclose(mz, mz_ref, rtol=0, atol=1e-13)


@pytest.mark.parametrize("dwx,dwp", [(False, False), (True, True)])
def test_broadcast_blocks_from_a_state(dwx, dwp):
    """One precession block and one RF block from a nonzero state, every
    field order, both off-resonance flags (HighOrderGPU.jl:73-93, :203-208)."""
    for rf in (False, True):
        obj, seq, mxy0, mz0 = high_order_workload(nspins=61, nsteps=32, ncoils=3, rf=rf)
        parts = (np.array([0]), np.array([seq.n_nodes - 1]), np.array([int(rf)], np.uint8))
        for bits in (0b1111, 0b1010, 0b0001):
            m = BlochHighOrder(format(bits, "04b"), dwx, dwp)
            fn = oracle.excitation if rf else oracle.precession
            ref, mxy_ref, mz_ref = fn(obj, seq, 0, seq.n_nodes - 1, m.flags, GAMMA, mxy0, mz0)
            sig, mxy, mz = broadcast.run(obj, seq, parts, m.flags, GAMMA, mxy0, mz0, n_threads=2)
            np.testing.assert_allclose(sig, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())
            np.testing.assert_allclose(mxy, mxy_ref, rtol=0, atol=1e-13)
            np.testing.assert_allclose(mz, mz_ref, rtol=0, atol=1e-13)
