"""The continuous-batching server at the serving stream's real shape: a
resident 1,048,576-spin, 32-coil phantom and a 20,000-sample readout (the
largest request the stream draws), served alone and in a mix with smaller
requests of other coil counts, field orders and readout lengths.

* The large request (1,048,576 spins x 20,000 samples x 32 coils, 2.1e10
  spin-samples) runs over several steps at the default step budget (2^33
  spin-samples) and gives the same bits alone and in the mix; so do the
  small ones.
* A spin-subsample request (2048 spins of the same resident phantom, all
  samples and coils) matches the f64 oracle at the mixed bar (NRMSE <= 1e-5);
  the oracle cannot run the full request in a test's time.
* The large request equals the sum of two half-phantom requests (the signal
  is a sum over spins; the halves get their own power-of-two scales):
  NRMSE <= 1e-6.
"""
import numpy as np
import pytest

import oracle
from hom_amd import workloads
from hom_amd.serving import SimulationServer
from hom_amd.types import GAMMA, BlochHighOrder, HO_Phantom
from helpers import nrmse, print_case, require_gpu

pytestmark = pytest.mark.gpu

NS, NC, NSAMP = 1 << 20, 32, 20_000
GB = 1 << 30
LARGE = (0, NS, NC, NSAMP, "1111")
SMALL = [(1000, 5000, 8, 3000, "1100"),
         (500_000, 20_000, 16, 12_000, "1110"),
         (NS - 1000, 1000, 1, NSAMP, "1111"),
         (123_456, 150_000, 32, 7000, "1111")]
SUB = (700_000, 2048, NC, NSAMP, "1111")


@pytest.fixture(scope="module")
def inputs():
    return workloads.serving_inputs(n_spins=NS, n_coils=NC, n_samples=NSAMP)


def serve(obj, seq, cases):
    with SimulationServer(memory_bytes=16 * GB, result_bytes=GB, threaded=False) as srv:
        ph, sq = srv.put_phantom(obj), srv.put_sequence(seq)
        reqs = [srv.submit(ph, sq, BlochHighOrder(c[4]), spins=(c[0], c[1]), coils=c[2],
                           n_samples=c[3]) for c in cases]
        srv.drain()
        stats = srv.stats()
        return [r.result() for r in reqs], stats


def test_serve_fullsize_alone_equals_mix_and_oracle(inputs):
    require_gpu()
    obj, seq = inputs
    print_case("serve-fullsize-alone", LARGE)
    (large_alone,), st = serve(obj, seq, [LARGE])
    assert st["steps"] >= 3, f"the large request ran in {st['steps']} steps"
    assert large_alone.shape == (NSAMP, NC) and np.isfinite(large_alone).all()
    small_alone = [serve(obj, seq, [c])[0][0] for c in SMALL + [SUB]]

    print_case("serve-fullsize-mix", len(SMALL) + 2)
    mixed, st = serve(obj, seq, [LARGE] + SMALL + [SUB])
    assert st["max_active_seen"] >= 2
    np.testing.assert_array_equal(mixed[0], large_alone)
    for got, want, case in zip(mixed[1:], small_alone, SMALL + [SUB]):
        np.testing.assert_array_equal(got, want, err_msg=str(case))

    first, count, coils, nsamp, order = SUB
    s = slice(first, first + count)
    sub = HO_Phantom(x=obj.x[s], y=obj.y[s], z=obj.z[s], rho=obj.rho[s], T1=obj.T1[s],
                     T2=obj.T2[s], dw=obj.dw[s], csm=np.asfortranarray(obj.csm[s, :coils]))
    ref, _, _ = oracle.run_workload(sub, seq, BlochHighOrder(order).flags, GAMMA, n_threads=16)
    e = nrmse(mixed[-1], ref)
    print_case("serve-fullsize-oracle", e)
    assert e <= 1e-5, f"spin-subsample request vs oracle: NRMSE {e:.3e}"


def test_serve_fullsize_additive_over_spins(inputs):
    require_gpu()
    obj, seq = inputs
    half = NS // 2
    print_case("serve-fullsize-halves")
    (whole, a, b), _ = serve(obj, seq, [LARGE, (0, half, NC, NSAMP, "1111"),
                                        (half, half, NC, NSAMP, "1111")])
    e = nrmse(a + b, whole)
    print_case("serve-fullsize-additivity", e)
    assert e <= 1e-6, f"halves vs whole: NRMSE {e:.3e}"
