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
from hom_amd.types import GAMMA, BlochHighOrder, HO_Phantom, get_sim_ranges
from hom_amd.workloads import FreeBlock, RFBlock, build_sequence
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


def test_segments_of_one_run_take_their_own_tiles():
    """A short readout (320 samples) and a long one (6400) after their own
    pulses: at 40,000 spins the short segment's work items fill the GPU only
    with the smallest tile, the long one's already at 128 groups.  The run
    matches the f64 oracle at the mixed bar, and a run with the short
    segment's tile fixed for both agrees to NRMSE <= 2e-6."""
    require_gpu()
    rng = np.random.default_rng(7)
    ns, nc = NS, 8
    obj = HO_Phantom(x=rng.uniform(-0.1, 0.1, ns), y=rng.uniform(-0.1, 0.1, ns),
                     z=rng.uniform(-0.035, 0.035, ns), rho=rng.uniform(0.0, 1.0, ns),
                     T1=rng.uniform(0.5, 2.0, ns), T2=rng.uniform(0.02, 0.2, ns),
                     dw=2 * np.pi * rng.normal(0.0, 20.0, ns),
                     csm=np.asfortranarray(rng.normal(size=(ns, nc)) + 1j * rng.normal(size=(ns, nc))))
    tau = 1e-4
    b1 = 30.0 / 360.0 / (GAMMA * tau)

    def grad(rel):
        return np.vstack([5e-3 * np.sin(2 * np.pi * 700 * rel), 4e-3 * np.cos(2 * np.pi * 700 * rel),
                          np.full(rel.size, 1e-3)])
    blocks = []
    for n in (320, 6400):
        dur = n * 2e-6
        blocks += [RFBlock(tau, complex(b1)), FreeBlock(dur, np.arange(1, n + 1) * 2e-6, grad)]
    seq = build_sequence(blocks, dt_max=2e-4)
    method = BlochHighOrder("1111")
    parts = get_sim_ranges(seq, 10 ** 9, 10 ** 9)
    nadc = max(seq.block_adc(a, b) for a, b in zip(parts[0], parts[1]))
    ref, _, _ = oracle.run(obj, seq, parts, method.flags, GAMMA, n_threads=16)
    with Plan("mixed", ns, nc, seq.n_nodes, nadc, max_rf_adc=0) as plan:
        plan.upload_phantom(obj)
        plan.upload_sequence(seq)
        plan.run(parts, method)
        auto = plan.download_signal()
        long_tile = plan.last_spin_tile
    print_case("tuning-segments", long_tile)
    assert long_tile == 128
    e = nrmse(auto, ref)
    assert e <= 1e-5, f"NRMSE {e:.3e} vs the f64 oracle"
    with Plan("mixed", ns, nc, seq.n_nodes, nadc, max_rf_adc=0) as plan:
        plan.upload_phantom(obj)
        plan.upload_sequence(seq)
        plan.spin_tile = 16
        plan.run(parts, method)
        fixed = plan.download_signal()
    # the short readout's rows: the same tile (16) either way, the same bits
    np.testing.assert_array_equal(auto[:320], fixed[:320])
    d = nrmse(auto, fixed)
    assert 0 < d <= 2e-6, f"automatic vs fixed 16: NRMSE {d:.3e}"
