"""GPU parity of the mixed-precision segment-fused path against the f64 oracle.

Bar: signal NRMSE <= 1e-5 (||s - s_f64|| / ||s_f64|| over every sample and
coil, no fitted scale or phase); the end-of-run spin state is computed in
FP64 and must match at 1e-11 (atol = rtol, norm-wise) like the f64 drop-in.
"""
import numpy as np
import pytest

import oracle
from hom_amd.runtime import Plan
from hom_amd.types import GAMMA, BlochHighOrder, HO_Phantom, get_sim_ranges
from hom_amd.serving import SimulationServer
from hom_amd.workloads import FreeBlock, RFBlock, build_sequence, high_order_workload, make_config
from helpers import assert_isapprox, nrmse, print_case, require_gpu

pytestmark = pytest.mark.gpu

NRMSE_BAR = 1e-5


def run_mixed(obj, seq, parts, method, mxy0=None, mz0=None, precision="mixed"):
    nadc_max = max([seq.block_adc(a, b) for a, b in zip(parts[0], parts[1])] + [1])
    with Plan(precision, len(obj), obj.n_coils, seq.n_nodes, nadc_max) as plan:
        plan.upload_phantom(obj)
        plan.upload_sequence(seq)
        if mxy0 is not None:
            plan.set_state(mxy0, mz0)
        plan.run(parts, method)
        sig = plan.download_signal()
        mxy, mz = plan.get_state()
    return sig, mxy, mz


def whole(seq):
    return (np.array([0]), np.array([seq.n_nodes - 1]), np.array([0], np.uint8))


def test_mixed_all_masks_reference_workload():
    require_gpu()
    for mask in [format(m, "04b") for m in range(16)]:
        for dw in (True, False):
            print_case("mixed-mask", mask, dw)
            obj, seq, mxy0, mz0 = high_order_workload(nspins=259, ncoils=32)
            m = BlochHighOrder(mask, True, dw)
            ref, rmxy, rmz = oracle.precession(obj, seq, 0, seq.n_nodes - 1, m.flags, GAMMA, mxy0, mz0)
            sig, mxy, mz = run_mixed(obj, seq, whole(seq), m, mxy0, mz0)
            e = nrmse(sig, ref)
            assert e <= NRMSE_BAR, f"{mask} dw={dw}: NRMSE {e:.3e}"
            assert_isapprox(mxy, rmxy, 1e-11, 1e-11, "Mxy")
            assert_isapprox(mz, rmz, 1e-11, 1e-11, "Mz")


@pytest.mark.parametrize("nspins,ncoils,nsteps", [
    (1, 1, 1), (15, 3, 7), (17, 8, 33), (255, 32, 100), (4097, 40, 64), (9000, 16, 20)])
def test_mixed_shapes(nspins, ncoils, nsteps):
    """Partial 16-spin groups and spin tiles, coil counts needing 1..5 column
    blocks (40 coils = two MFMA column chunks), partial sample tiles."""
    require_gpu()
    print_case("mixed-shape", nspins, ncoils, nsteps)
    obj, seq, mxy0, mz0 = high_order_workload(nspins=nspins, nsteps=nsteps, ncoils=ncoils)
    m = BlochHighOrder("1111")
    ref, rmxy, rmz = oracle.precession(obj, seq, 0, seq.n_nodes - 1, m.flags, GAMMA, mxy0, mz0)
    sig, mxy, mz = run_mixed(obj, seq, whole(seq), m, mxy0, mz0)
    assert nrmse(sig, ref) <= NRMSE_BAR
    assert_isapprox(mxy, rmxy, 1e-11, 1e-11, "Mxy")


def test_mixed_rf_interleaves_per_interleave_error():
    """RF + spiral readout + idle to TR, 4 interleaves (the 350,000-spin 3D
    geometry, every 250th spin): NRMSE per interleave and overall."""
    require_gpu()
    obj, seq, cfg = make_config(1, spin_subsample=250)
    m = BlochHighOrder("1111")
    for L in (32, 10 ** 9):
        print_case("mixed-config1-sub", L)
        parts = get_sim_ranges(seq, L, L)
        ref, rmxy, rmz = oracle.run(obj, seq, parts, m.flags, GAMMA)
        sig, mxy, mz = run_mixed(obj, seq, parts, m)
        n = cfg.n_adc
        for il in range(cfg.n_interleaves):
            e = nrmse(sig[il * n:(il + 1) * n], ref[il * n:(il + 1) * n])
            print(f"interleave {il}: NRMSE {e:.3e}", flush=True)
            assert e <= NRMSE_BAR
        assert_isapprox(mxy, rmxy, 1e-11, 1e-11, "Mxy")
        assert_isapprox(mz, rmz, 1e-11, 1e-11, "Mz")


def test_mixed_adc_during_rf_and_empty_segments():
    require_gpu()
    obj, seq, mxy0, mz0 = high_order_workload(nspins=300, nsteps=32, ncoils=3)
    seq.excitation[:] = False
    seq.excitation[10:14] = True       # an RF block in the middle, acquiring
    seq.b1[10:14] = 5e-6
    seq.adc[:] = False
    seq.adc[[3, 11, 12, 14, 20, 21, 22]] = True   # segment 1, RF, segment 2 (30..32 idle)
    m = BlochHighOrder("1111")
    parts = get_sim_ranges(seq, 4, 2)
    ref, rmxy, rmz = oracle.run(obj, seq, parts, m.flags, GAMMA, mxy0, mz0)
    sig, mxy, mz = run_mixed(obj, seq, parts, m, mxy0, mz0)
    assert sig.shape == ref.shape == (7, 3)
    assert nrmse(sig, ref) <= NRMSE_BAR
    assert_isapprox(mxy, rmxy, 1e-11, 1e-11, "Mxy")


def test_mixed_column_scaling_extreme_magnitudes():
    """rho ~ 1e3 and tiny / huge coil sensitivities stay inside FP16 range."""
    require_gpu()
    obj, seq, mxy0, mz0 = high_order_workload(nspins=500, ncoils=4)
    obj.csm[:, 0] *= 1e-4
    obj.csm[:, 1] *= 3e3
    obj.csm[:, 3] = 0.0
    mxy0 *= 1e3
    mz0 *= 1e3
    m = BlochHighOrder("1111")
    ref, _, _ = oracle.precession(obj, seq, 0, seq.n_nodes - 1, m.flags, GAMMA, mxy0, mz0)
    sig, _, _ = run_mixed(obj, seq, whole(seq), m, mxy0, mz0)
    for c in range(3):
        assert nrmse(sig[:, c], ref[:, c]) <= NRMSE_BAR, c
    assert np.all(sig[:, 3] == 0)


def strong_field_sequence():
    """Two 300-sample readouts, each after a 90-degree hard pulse.  Third-order
    fields of 1 T/m^3 during the first readout give the edge spins about 18
    revolutions of higher-order phase inside one 256-sample anchor block, far
    beyond what the FP16 offsets hold, so that segment runs the all-FP64
    phase body; the second readout has 1e-3 T/m^3 and stays on the split
    path."""
    adc = np.linspace(1e-3 / 300, 1e-3, 300)

    def grad(rel):
        return np.vstack([np.full(rel.size, 2e-3), 1e-3 * np.sin(2e3 * rel), np.zeros(rel.size)])

    def fields(t):
        h = np.zeros((16, t.size))
        amp = np.where(t < 1.15e-3, 1.0, 1e-3)
        for j in range(9, 16):
            h[j] = amp * (0.5 + 0.5 * np.sin(2e3 * t + j))
        h[4:9] = 1e-4 * np.cos(3e3 * t)[None, :]
        return h

    pulse = RFBlock(1e-4, 90.0 / 360.0 / (GAMMA * 1e-4))
    blocks = [pulse, FreeBlock(1e-3, adc, grad), pulse, FreeBlock(1e-3, adc, grad)]
    return build_sequence(blocks, fields=fields, dt_max=5e-6)


def test_mixed_fp64_fallback_segment():
    """The all-FP64 phase body (and the split body in the same run) against
    the f64 oracle, through a plan and through the server."""
    require_gpu()
    obj, _, _, _ = high_order_workload(nspins=700, ncoils=8)
    seq = strong_field_sequence()
    m = BlochHighOrder("1111")
    parts = get_sim_ranges(seq, 32, 32)
    ref, rmxy, rmz = oracle.run(obj, seq, parts, m.flags, GAMMA)
    print_case("mixed-fp64-fallback", len(obj), seq.n_nodes)
    with Plan("mixed", len(obj), obj.n_coils, seq.n_nodes, 64) as plan:
        plan.upload_phantom(obj)
        plan.upload_sequence(seq)
        plan.run(parts, m)
        sig = plan.download_signal()
        mxy, mz = plan.get_state()
        assert plan.mixed_stats() == (2, 1)
    assert sig.shape == ref.shape == (600, 8)
    for half in (slice(0, 300), slice(300, 600)):
        assert nrmse(sig[half], ref[half]) <= NRMSE_BAR
    assert_isapprox(mxy, rmxy, 1e-11, 1e-11, "Mxy")
    assert_isapprox(mz, rmz, 1e-11, 1e-11, "Mz")
    print_case("serve-fp64-fallback")
    with SimulationServer(memory_bytes=1 << 28, result_bytes=1 << 24, threaded=False) as srv:
        req = srv.submit(srv.put_phantom(obj), srv.put_sequence(seq), m)
        srv.drain()
        out = req.result()
    assert out.shape == ref.shape
    for half in (slice(0, 300), slice(300, 600)):
        assert nrmse(out[half], ref[half]) <= NRMSE_BAR


def test_mixed_deterministic_and_matches_f64_plan():
    require_gpu()
    obj, seq, cfg = make_config(0, spin_subsample=20)
    m = BlochHighOrder("1111")
    parts = get_sim_ranges(seq, 32, 32)
    a, _, _ = run_mixed(obj, seq, parts, m)
    b, _, _ = run_mixed(obj, seq, parts, m)
    assert np.array_equal(a, b)
    f, _, _ = run_mixed(obj, seq, parts, m, precision="f64")
    assert nrmse(a, f) <= NRMSE_BAR


def scaled_gradients(seq, factor):
    """A copy of seq with the first-order waveforms (h1..h3 and the nominal
    gradients) scaled by factor."""
    import copy
    s = copy.deepcopy(seq)
    s.h = np.array(s.h, copy=True)
    s.h[1:4] *= factor
    s.gx, s.gy, s.gz = s.gx * factor, s.gy * factor, s.gz * factor
    return s


@pytest.mark.parametrize("factor,fp64_segments", [(4.0, 0), (9.0, 4)])
def test_mixed_strong_gradients_stay_split(factor, fp64_segments):
    """Strong gradients: the 350,000-spin, 8-coil, 4 x 4750-sample spiral
    workload (every 500th spin) with its gradients scaled 4x keeps all four
    readouts on the split-phase body (first-order offsets up to about 8
    revolutions per term within 8 samples); scaled 9x (about 18 revolutions)
    they run the all-FP64 body.  Both against the f64 oracle per readout."""
    require_gpu()
    obj, seq, cfg = make_config(1, spin_subsample=500)
    seq = scaled_gradients(seq, factor)
    m = BlochHighOrder("1111")
    parts = get_sim_ranges(seq, 32, 32)
    print_case("mixed-strong-gradients", factor, len(obj))
    ref, _, _ = oracle.run(obj, seq, parts, m.flags, GAMMA, n_threads=8)
    nadc = max(seq.block_adc(a, b) for a, b in zip(parts[0], parts[1]))
    with Plan("mixed", len(obj), obj.n_coils, seq.n_nodes, nadc) as plan:
        plan.upload_phantom(obj)
        plan.upload_sequence(seq)
        plan.run(parts, m)
        sig = plan.download_signal()
        segments, fp64 = plan.mixed_stats()
    assert (segments, fp64) == (cfg.n_interleaves, fp64_segments)
    n = cfg.n_adc
    for i in range(cfg.n_interleaves):
        assert nrmse(sig[i * n:(i + 1) * n], ref[i * n:(i + 1) * n]) <= NRMSE_BAR


@pytest.mark.parametrize("factor,fp64_segments", [(30.0, 0), (300.0, 4)])
def test_mixed_strong_higher_order_fields_stay_split(factor, fp64_segments):
    """Strong 2nd- and 3rd-order fields: the 350,000-spin, 8-coil,
    4 x 4750-sample workload (every 500th spin) with h4..h15 scaled 30x keeps
    all four readouts on the split-phase body (higher-order offsets of a few
    revolutions within a 256-sample anchor block); scaled 300x they run the
    all-FP64 body.  Both against the f64 oracle per readout."""
    import copy
    require_gpu()
    obj, seq, cfg = make_config(1, spin_subsample=500)
    seq = copy.deepcopy(seq)
    seq.h = np.array(seq.h, copy=True)
    seq.h[4:16] *= factor
    m = BlochHighOrder("1111")
    parts = get_sim_ranges(seq, 32, 32)
    print_case("mixed-strong-higher-order", factor, len(obj))
    ref, _, _ = oracle.run(obj, seq, parts, m.flags, GAMMA, n_threads=8)
    nadc = max(seq.block_adc(a, b) for a, b in zip(parts[0], parts[1]))
    with Plan("mixed", len(obj), obj.n_coils, seq.n_nodes, nadc) as plan:
        plan.upload_phantom(obj)
        plan.upload_sequence(seq)
        plan.run(parts, m)
        sig = plan.download_signal()
        segments, fp64 = plan.mixed_stats()
    assert (segments, fp64) == (cfg.n_interleaves, fp64_segments)
    n = cfg.n_adc
    for i in range(cfg.n_interleaves):
        assert nrmse(sig[i * n:(i + 1) * n], ref[i * n:(i + 1) * n]) <= NRMSE_BAR
