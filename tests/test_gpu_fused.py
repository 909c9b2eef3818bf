"""GPU parity of the fused f32/f64 precession + coil reception (coil modes
"fused_warp" and "fused_tiled") against the oracle and the per-block path.

Mirrors the reference's fused-coil suite (test/gpu/test_gpu_fused_coils.jl):
every mode gives the signal and state of the default path at the reference's
tolerances (f32 3e-5, f64 1e-11, atol = rtol, norm-wise), over masks, coil
counts that leave partial 8-coil blocks, spin counts that leave a partial
256-spin workgroup, and blocks whose ADC count leaves a partial 16-sample tile.
"""
import numpy as np
import pytest

import oracle
from hom_amd.runtime import Plan
from hom_amd.types import GAMMA, BlochHighOrder, DimensionMismatch, get_sim_ranges
from hom_amd.workloads import high_order_workload, make_config
from helpers import assert_isapprox, assert_signal, nrmse, print_case, require_gpu, round_f32, tol
from test_gpu_plain import make_plan, run_block

pytestmark = pytest.mark.gpu

MODES = ["fused_warp", "fused_tiled"]

# (spins, coils, intervals, mask): partial workgroups, partial coil blocks
# (1, 3, 9, 17, 40 coils: 4, 4, 2, 1 and 1 spin parts per workgroup) and
# partial sample tiles (15, 40 and 33 samples)
CASES = [(259, 32, 32, "1111"), (259, 1, 32, "1111"), (300, 3, 15, "0101"),
         (513, 17, 40, "1010"), (100, 40, 33, "1111"), (1, 8, 32, "1111"),
         (256, 9, 16, "0000"), (1000, 24, 17, "0111")]


def _inputs(precision, nspins, ncoils, nsteps, rf=False):
    obj, seq, mxy0, mz0 = high_order_workload(nspins=nspins, nsteps=nsteps, ncoils=ncoils, rf=rf)
    if precision == "f32":
        obj, seq = round_f32(obj, seq)
        mxy0 = mxy0.astype(np.complex64).astype(np.complex128)
        mz0 = mz0.astype(np.float32).astype(np.float64)
    return obj, seq, mxy0, mz0


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("mode", MODES)
def test_fused_precession_parity(precision, mode):
    require_gpu()
    t = tol(precision)
    for nspins, ncoils, nsteps, mask in CASES:
        for adc in (True, False):
            print_case("fused", precision, mode, nspins, ncoils, nsteps, mask, adc)
            obj, seq, mxy0, mz0 = _inputs(precision, nspins, ncoils, nsteps)
            if not adc:
                seq.adc[:] = False
            method = BlochHighOrder(mask, True, True)
            last = seq.n_nodes - 1
            ref, rmxy, rmz = oracle.precession(obj, seq, 0, last, method.flags, GAMMA, mxy0, mz0)
            nadc = seq.block_adc(0, last)
            with make_plan(precision, obj, seq, max(nadc, 1), max_rf_adc=0,
                           coil_combination=mode) as plan:
                plan.set_state(mxy0, mz0)
                sig = run_block(plan, "precession", 0, last, method, nadc, ncoils)
                mxy, mz = plan.get_state()
            if nadc:
                assert_signal(sig, ref, precision, "signal")
            assert_isapprox(mxy, rmxy, t, t, "Mxy")
            assert_isapprox(mz, rmz, t, t, "Mz")


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_fused_strided_output(precision):
    """Rows 16..16+n of a 100-row parent; the other rows keep their sentinel."""
    require_gpu()
    for nsteps in (1, 15, 32):
        print_case("fused-strided", precision, nsteps)
        obj, seq, mxy0, mz0 = _inputs(precision, 259, 3, nsteps)
        m = BlochHighOrder("1111")
        ref, _, _ = oracle.precession(obj, seq, 0, nsteps, m.flags, GAMMA, mxy0, mz0)
        with make_plan(precision, obj, seq, 32, coil_combination="fused_tiled") as plan:
            plan.set_state(mxy0, mz0)
            full = run_block(plan, "precession", 0, nsteps, m, nsteps, 3, ld=100, row0=16,
                             sentinel=complex(7, -3))
        assert_signal(full[16:16 + nsteps], ref, precision, "rows")
        assert np.all(full[:16] == complex(7, -3))
        assert np.all(full[16 + nsteps:] == complex(7, -3))


def test_fused_matches_gemm_and_is_deterministic():
    """f64: fused and per-block results agree to rounding; two fused runs are
    bit-identical (fixed partial rows, fixed finalize order)."""
    require_gpu()
    obj, seq, mxy0, mz0 = high_order_workload(nspins=5000, nsteps=37, ncoils=32)
    m = BlochHighOrder("1111")
    out = {}
    for mode in ("gemm", "fused_tiled", "fused_tiled"):
        print_case("fused-vs-gemm", mode)
        with make_plan("f64", obj, seq, 37, coil_combination=mode) as plan:
            plan.set_state(mxy0, mz0)
            sig = run_block(plan, "precession", 0, 37, m, 37, 32)
        out.setdefault(mode, []).append(sig)
    a, b = out["fused_tiled"]
    assert np.array_equal(a, b)
    assert np.linalg.norm(a - out["gemm"][0]) / np.linalg.norm(out["gemm"][0]) < 1e-13


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_fused_run_with_rf(precision):
    """hom_run in a fused mode: RF blocks (per-block path, states of RF capacity
    only) between fused precession blocks, configuration-0 geometry."""
    require_gpu()
    obj, seq, _ = make_config(0, spin_subsample=100)
    m = BlochHighOrder("1111")
    parts = get_sim_ranges(seq, 32, 32)
    ref, rmxy, rmz = oracle.run(obj, seq, parts, m.flags, GAMMA)
    counts = [seq.block_adc(a, b) for a, b in zip(parts[0], parts[1])]
    rf_counts = [c for c, r in zip(counts, parts[2]) if r]
    for mode in MODES:
        print_case("fused-run", precision, mode)
        with make_plan(precision, obj, seq, max(counts), max_rf_adc=max(rf_counts, default=0),
                       coil_combination=mode) as plan:
            plan.run(parts, m)
            sig = plan.download_signal()
            mxy, mz = plan.get_state()
        err = np.linalg.norm(sig - ref) / np.linalg.norm(ref)
        print(f"fused-run {precision} {mode} NRMSE {err:.3e}", flush=True)
        if precision == "f64":
            assert_signal(sig, ref, "f64", mode)
            assert_isapprox(mxy, rmxy, 1e-11, 1e-11, "Mxy")
            assert_isapprox(mz, rmz, 1e-11, 1e-11, "Mz")
        else:
            assert err < 1e-3


def test_fused_rf_adc_capacity_and_empty_phantom():
    require_gpu()
    obj, seq, mxy0, mz0 = high_order_workload(nspins=64, nsteps=8, ncoils=2, rf=True)
    m = BlochHighOrder("1111")
    with make_plan("f64", obj, seq, 8, max_rf_adc=4, coil_combination="fused_warp") as plan:
        with pytest.raises(DimensionMismatch):
            run_block(plan, "excitation", 0, 8, m, 8, 2)
    # RF-capacity ADC on a fused plan runs and matches the oracle
    ref, _, _ = oracle.excitation(obj, seq, 0, 8, m.flags, GAMMA, mxy0, mz0)
    with make_plan("f64", obj, seq, 8, max_rf_adc=8, coil_combination="fused_warp") as plan:
        plan.set_state(mxy0, mz0)
        sig = run_block(plan, "excitation", 0, 8, m, 8, 2)
    assert_signal(sig, ref, "f64", "RF ADC")
    # no spins: the fused block writes zeros
    from hom_amd.types import HO_Phantom
    _, seq0, _, _ = high_order_workload(nspins=4, nsteps=20, ncoils=3)
    empty = HO_Phantom(x=np.zeros(0), csm=np.zeros((0, 3), dtype=np.complex128, order="F"))
    with make_plan("f64", empty, seq0, 20, coil_combination="fused_tiled") as plan:
        sig = run_block(plan, "precession", 0, 20, m, 20, 3, sentinel=complex(7, -3))
    assert np.all(sig == 0)


def test_failed_allocation_leaves_no_stale_error():
    """A device allocation that fails (1 PiB) is reported as such and does not
    surface later as a launch error of the next block."""
    require_gpu()
    from hom_amd.runtime import HomDeviceError
    obj, seq, mxy0, mz0 = high_order_workload(nspins=300, nsteps=20, ncoils=3)
    m = BlochHighOrder("1111")
    ref, _, _ = oracle.precession(obj, seq, 0, 20, m.flags, GAMMA, mxy0, mz0)
    for mode in ("gemm", "fused_tiled"):
        print_case("failed-alloc", mode)
        with make_plan("f64", obj, seq, 20, coil_combination=mode) as plan:
            with pytest.raises(HomDeviceError):
                plan.buffer(1 << 50)
            plan.set_state(mxy0, mz0)
            sig = run_block(plan, "precession", 0, 20, m, 20, 3)
        assert_signal(sig, ref, "f64", mode)


def test_mixed_plan_state_buffer_is_rf_only():
    """A mixed plan sizes its ADC-state buffer for RF-block ADC only (its
    precession segments keep states in registers): 2^20 spins x 10^6 samples
    per block would need 16 TiB.  Per-block stepping of a precession block on
    a mixed plan (callbacks) runs that block as one mixed segment, with its
    states on chip: the mixed bar on the signal, FP64 parity on the state."""
    require_gpu()
    print_case("mixed-state-buffer")
    with Plan("mixed", 1 << 20, 32, 4, 10 ** 6, max_rf_adc=0):
        pass
    obj, seq, mxy0, mz0 = high_order_workload(nspins=300, nsteps=40, ncoils=3)
    m = BlochHighOrder("1111")
    ref, rmxy, _ = oracle.precession(obj, seq, 0, 40, m.flags, GAMMA, mxy0, mz0)
    with make_plan("mixed", obj, seq, 40, max_rf_adc=0) as plan:
        plan.set_state(mxy0, mz0)
        sig = run_block(plan, "precession", 0, 40, m, 40, 3)
        mxy, _ = plan.get_state()
        assert plan.mixed_stats() == (1, 0)
    assert nrmse(sig, ref) <= 1e-5, "per-block mixed plan"
    assert_isapprox(mxy, rmxy, 1e-11, 1e-11, "Mxy")
