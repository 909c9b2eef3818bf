"""Persisted launch shapes on the GPU.  A mixed plan takes the spin tile that
tuning.json lists for its shape class as the upper bound of a per-segment
choice (smaller tiles only when the work items would not fill the GPU's
workgroup slots), keeps a tile fixed explicitly, and the automatic choice is
exactly a fixed tile (bit-identical results).  The tile sets how many spins
one FP32 partial sum covers before the FP64 sum across tiles, so it moves the
result only by that accumulation error (at 40,000 spins and 32 coils about
1.2e-7, 3.5e-7 and 1.1e-6 NRMSE for 64-, 256- and 1024-group tiles): every
tile matches the f64 oracle at the mixed bar (NRMSE <= 1e-5) and the tiles
agree with one another to NRMSE <= 2e-6."""
import numpy as np
import pytest

import oracle
from hom_amd import tuning, workloads
from hom_amd.runtime import Plan
from hom_amd.types import GAMMA, BlochHighOrder, get_sim_ranges
from helpers import nrmse, print_case, require_gpu

pytestmark = pytest.mark.gpu

NS, NC, NSAMP = 40_000, 32, 2000


@pytest.fixture(scope="module")
def inputs():
    obj, seq = workloads.serving_inputs(n_spins=NS, n_coils=NC, n_samples=NSAMP)
    method = BlochHighOrder("1111")
    ref, _, _ = oracle.run_workload(obj, seq, method.flags, GAMMA, n_threads=16)
    return obj, seq, method, ref


def run(obj, seq, method, tile=None):
    """(signal, configured tile before the run, configured tile after, tile used)."""
    parts = get_sim_ranges(seq, 10 ** 9, 10 ** 9)
    nadc = max(seq.block_adc(a, b) for a, b in zip(parts[0], parts[1]))
    with Plan("mixed", len(obj), obj.n_coils, seq.n_nodes, nadc, max_rf_adc=0) as plan:
        plan.upload_phantom(obj)
        plan.upload_sequence(seq)
        before = plan.spin_tile
        if tile is not None:
            plan.spin_tile = tile
        plan.run(parts, method)
        sig = plan.download_signal()
        return sig, before, plan.spin_tile, plan.last_spin_tile


def bound(method):
    return tuning.plan_spin_tile(NC, method.flags) or tuning.DEFAULT_SPIN_TILE


def test_plan_applies_the_tuned_bound(inputs):
    require_gpu()
    obj, seq, method, _ = inputs
    print_case("tuning-applied")
    auto, before, after, used = run(obj, seq, method)
    assert before == tuning.DEFAULT_SPIN_TILE
    assert after == bound(method)
    assert used in tuning.SPIN_TILES and used <= after
    # 2,500 groups x 8 sample blocks: the bound leaves most slots idle
    assert used < after
    fixed, _, _, used_fixed = run(obj, seq, method, tile=used)
    assert used_fixed == used
    np.testing.assert_array_equal(auto, fixed)
    _, _, _, explicit = run(obj, seq, method, tile=1024)
    assert explicit == 1024


def test_large_problem_keeps_the_bound():
    require_gpu()
    obj, seq = workloads.serving_inputs(n_spins=1 << 20, n_coils=NC, n_samples=NSAMP)
    method = BlochHighOrder("1111")
    print_case("tuning-large")
    sig, _, after, used = run(obj, seq, method)
    assert used == after == bound(method)
    assert np.isfinite(sig).all()


def test_spin_tile_changes_only_the_partial_sums(inputs):
    require_gpu()
    obj, seq, method, ref = inputs
    sigs = {}
    for tile in (64, 256, 1024):
        print_case("tuning-tile", tile)
        sigs[tile], _, _, used = run(obj, seq, method, tile=tile)
        assert used == tile
        e = nrmse(sigs[tile], ref)
        print_case("tuning-tile-nrmse", tile, e)
        assert e <= 1e-5, f"tile {tile}: NRMSE {e:.3e} vs the f64 oracle"
    for tile in (64, 1024):
        d = nrmse(sigs[tile], sigs[256])
        assert d <= 2e-6, f"tile {tile} vs 256: NRMSE {d:.3e}"
