"""Golden vectors: the oracle reproduces its committed outputs on the restated
reference workload (CPU), and the f64 / mixed GPU paths match them (GPU)."""
from pathlib import Path

import numpy as np
import pytest

import oracle
from hom_amd.types import GAMMA, BlochHighOrder
from hom_amd.workloads import high_order_workload
from helpers import assert_isapprox, nrmse, print_case

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "high_order_workload.npz")


def test_oracle_reproduces_golden():
    obj, seq, mxy0, mz0 = high_order_workload(nspins=259, nsteps=32, ncoils=3)
    for bits in range(16):
        mask = format(bits, "04b")
        for dw in (0, 1):
            m = BlochHighOrder(mask, True, bool(dw))
            sig, mxy, mz = oracle.precession(obj, seq, 0, seq.n_nodes - 1, m.flags, GAMMA, mxy0, mz0)
            np.testing.assert_allclose(sig, GOLD[f"prec_{mask}_{dw}_sig"], rtol=1e-13, atol=1e-13)
            np.testing.assert_allclose(mxy, GOLD[f"prec_{mask}_{dw}_mxy"], rtol=1e-13, atol=1e-13)
    obj, seq, mxy0, mz0 = high_order_workload(nspins=259, nsteps=32, ncoils=3, rf=True)
    for dwx in (0, 1):
        m = BlochHighOrder("1111", bool(dwx), True)
        sig, mxy, mz = oracle.excitation(obj, seq, 0, seq.n_nodes - 1, m.flags, GAMMA, mxy0, mz0)
        np.testing.assert_allclose(sig, GOLD[f"exc_{dwx}_sig"], rtol=1e-13, atol=1e-13)


@pytest.mark.gpu
def test_gpu_matches_golden():
    from helpers import require_gpu
    from hom_amd.runtime import Plan
    require_gpu()
    obj, seq, mxy0, mz0 = high_order_workload(nspins=259, nsteps=32, ncoils=3)
    parts = (np.array([0]), np.array([seq.n_nodes - 1]), np.array([0], np.uint8))
    for precision in ("f64", "mixed"):
        for bits in range(16):
            mask = format(bits, "04b")
            for dw in (0, 1):
                print_case("golden", precision, mask, dw)
                m = BlochHighOrder(mask, True, bool(dw))
                with Plan(precision, len(obj), 3, seq.n_nodes, 32) as plan:
                    plan.upload_phantom(obj)
                    plan.upload_sequence(seq)
                    plan.set_state(mxy0, mz0)
                    plan.run(parts, m)
                    sig = plan.download_signal()
                    mxy, mz = plan.get_state()
                ref = GOLD[f"prec_{mask}_{dw}_sig"]
                if precision == "f64":
                    assert_isapprox(sig, ref, 1e-11, 1e-11, mask)
                else:
                    assert nrmse(sig, ref) <= 1e-5, mask
                assert_isapprox(mxy, GOLD[f"prec_{mask}_{dw}_mxy"], 1e-11, 1e-11, mask)
                assert_isapprox(mz, GOLD[f"prec_{mask}_{dw}_mz"], 1e-11, 1e-11, mask)
