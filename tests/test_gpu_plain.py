"""GPU parity of the per-block drop-in kernels (f32 / f64) against the oracle.

Mirrors the reference's CUDA suites:
- test/gpu/test_gpu_kernels.jl:16-85   precession (16 masks, coils, ADC, dw), analytic phase
- test/gpu/test_gpu_kernels.jl:90-109  excitation
- test/gpu/test_gpu_kernels.jl:113-138  packed output into strided signal views
- test/gpu/test_gpu_reduction.jl        deterministic coil reduction, guards, cancellation
- test/gpu/test_gpu_compatibility.jl:183-187  block-length invariance
Tolerances are the reference's: f32 3e-5, f64 1e-11 (atol = rtol, norm-wise).
"""
import numpy as np
import pytest

import oracle
from hom_amd.runtime import Plan
from hom_amd.types import GAMMA, BlochHighOrder, get_sim_ranges
from hom_amd.workloads import high_order_workload
from helpers import assert_isapprox, assert_signal, print_case, require_gpu, round_f32, tol

pytestmark = pytest.mark.gpu

MASKS = [format(m, "04b") for m in range(16)]


def make_plan(precision, obj, seq, max_adc, **kw):
    plan = Plan(precision, len(obj), obj.n_coils, seq.n_nodes, max_adc, **kw)
    plan.upload_phantom(obj)
    plan.upload_sequence(seq)
    return plan


def run_block(plan, kind, first, last, method, nadc, nc, ld=None, row0=0, rows=None, sentinel=None):
    """Runs one block into a device buffer (rows x nc, column stride ld) whose
    other entries hold a sentinel; returns the whole host buffer."""
    ld = nadc if ld is None else ld
    rows = ld if rows is None else rows
    host = np.zeros((ld, nc), dtype=plan.cdtype, order="F")
    if sentinel is not None:
        host[:] = sentinel
    buf = plan.buffer(max(host.nbytes, 16))
    buf.upload(np.asfortranarray(host).ravel(order="F"))
    fn = plan.precession_block if kind == "precession" else plan.excitation_block
    fn(first, last, method, buf if nadc else None, max(ld, 1), offset_elems=row0)
    flat = np.zeros(host.size, dtype=plan.cdtype)
    buf.download(flat)
    buf.free()
    return flat.reshape((ld, nc), order="F")


def check_precession(precision, mask, ncoils=32, adc=True, dw_prec=True):
    print_case("precession", precision, mask, ncoils, adc, dw_prec)
    obj, seq, mxy0, mz0 = high_order_workload(nspins=259, ncoils=ncoils)
    if not adc:
        seq.adc[:] = False
    if precision == "f32":
        obj, seq = round_f32(obj, seq)
        mxy0 = mxy0.astype(np.complex64).astype(np.complex128)
        mz0 = mz0.astype(np.float32).astype(np.float64)
    method = BlochHighOrder(mask, True, dw_prec)
    ref, rmxy, rmz = oracle.precession(obj, seq, 0, seq.n_nodes - 1, method.flags, GAMMA, mxy0, mz0)
    nadc = seq.block_adc(0, seq.n_nodes - 1)
    with make_plan(precision, obj, seq, max(nadc, 1)) as plan:
        plan.set_state(mxy0, mz0)
        sig = run_block(plan, "precession", 0, seq.n_nodes - 1, method, nadc, ncoils)
        mxy, mz = plan.get_state()
    t = tol(precision)
    if nadc:
        assert_signal(sig, ref, precision, "signal")
    assert_isapprox(mxy, rmxy, t, t, "Mxy")
    assert_isapprox(mz, rmz, t, t, "Mz")


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_precession_all_masks(precision):
    require_gpu()
    for mask in MASKS:
        check_precession(precision, mask)
    for nc in (1, 3):
        check_precession(precision, "1111", ncoils=nc)
    check_precession(precision, "1111", adc=False)
    check_precession(precision, "1111", dw_prec=False)


def test_analytic_phase_each_channel():
    require_gpu()
    obj, seq, mxy0, mz0 = high_order_workload(nspins=3, nsteps=4, ncoils=3)
    obj.T1[:] = np.inf
    obj.T2[:] = np.inf
    obj.dw[:] = 0
    x, y, z = obj.x, obj.y, obj.z
    r2 = x ** 2 + y ** 2 + z ** 2
    basis = np.array([np.ones(3), x, y, z, x * y, z * y, 3 * z ** 2 - r2, x * z, x ** 2 - y ** 2,
                      3 * y * x ** 2 - y ** 3, x * z * y, (5 * z ** 2 - r2) * y,
                      5 * z ** 3 - 3 * z * r2, (5 * z ** 2 - r2) * x, x ** 2 * z - y ** 2 * z,
                      x ** 3 - 3 * x * y ** 2])
    for term in range(16):
        print_case("analytic", term)
        seq.h[:] = 0
        seq.h[term] = 2e-5
        with make_plan("f64", obj, seq, 4) as plan:
            plan.set_state(mxy0, mz0)
            sig = run_block(plan, "precession", 0, 4, BlochHighOrder("1111"), 4, 3)
            mxy, _ = plan.get_state()
        phase = -2 * np.pi * GAMMA * np.outer(basis[term], 2e-5 * seq.t[1:])
        states = mxy0[:, None] * np.exp(1j * phase)
        assert_isapprox(sig, states.T @ obj.csm, 1e-11, 1e-11, f"term {term}")
        assert_isapprox(mxy, states[:, -1], 1e-11, 1e-11, f"term {term} state")


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_excitation(precision):
    require_gpu()
    for ncoils in (1, 3, 32):
        for off in (False, True):
            print_case("excitation", precision, ncoils, off)
            obj, seq, mxy0, mz0 = high_order_workload(nspins=259, ncoils=ncoils, rf=True)
            if precision == "f32":
                obj, seq = round_f32(obj, seq)
                mxy0 = mxy0.astype(np.complex64).astype(np.complex128)
                mz0 = mz0.astype(np.float32).astype(np.float64)
            method = BlochHighOrder("1111", off, True)
            ref, rmxy, rmz = oracle.excitation(obj, seq, 0, seq.n_nodes - 1, method.flags, GAMMA,
                                               mxy0, mz0)
            with make_plan(precision, obj, seq, 32) as plan:
                plan.set_state(mxy0, mz0)
                sig = run_block(plan, "excitation", 0, seq.n_nodes - 1, method, 32, ncoils)
                mxy, mz = plan.get_state()
            t = tol(precision)
            assert_signal(sig, ref, precision, "signal")
            assert_isapprox(mxy, rmxy, t, t, "Mxy")
            assert_isapprox(mz, rmz, t, t, "Mz")


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_strided_signal_views(precision):
    """Output into rows 17..16+nsteps of a 100 x 3 parent; other rows untouched."""
    require_gpu()
    for nsteps in (1, 15, 32):
        print_case("strided", precision, nsteps)
        obj, seq, mxy0, mz0 = high_order_workload(nspins=259, nsteps=nsteps, ncoils=3)
        if precision == "f32":
            obj, seq = round_f32(obj, seq)
            mxy0 = mxy0.astype(np.complex64).astype(np.complex128)
            mz0 = mz0.astype(np.float32).astype(np.float64)
        m = BlochHighOrder("1111")
        ref, _, _ = oracle.precession(obj, seq, 0, nsteps, m.flags, GAMMA, mxy0, mz0)
        with make_plan(precision, obj, seq, 32) as plan:
            plan.set_state(mxy0, mz0)
            full = run_block(plan, "precession", 0, nsteps, m, nsteps, 3, ld=100, row0=16,
                             sentinel=complex(7, -3))
        assert_signal(full[16:16 + nsteps], ref, precision, "rows")
        assert np.all(full[:16] == complex(7, -3))
        assert np.all(full[16 + nsteps:] == complex(7, -3))


def check_reduction(dtype, nspins, nadc, ncoils, seed=42):
    print_case("reduction", dtype, nspins, nadc, ncoils)
    rng = np.random.default_rng(seed)
    st = (rng.normal(size=(nspins, nadc)) + 1j * rng.normal(size=(nspins, nadc))).astype(dtype)
    cs = (rng.normal(size=(nspins, ncoils)) + 1j * rng.normal(size=(nspins, ncoils))).astype(dtype)
    ref = st.astype(np.complex128).T @ cs.astype(np.complex128)
    precision = "f32" if dtype == np.complex64 else "f64"
    from hom_amd.types import HO_Phantom
    obj = HO_Phantom(x=np.zeros(nspins), csm=np.asfortranarray(cs.astype(np.complex128)))
    plan = Plan(precision, nspins, ncoils, 2, max(nadc, 1))
    try:
        plan.upload_phantom(obj)
        sbuf = plan.buffer(max(st.nbytes, 16))
        sbuf.upload(np.asfortranarray(st).ravel(order="F"))
        host = np.full((nadc + 4, ncoils), complex(7, -3), dtype=dtype, order="F")
        obuf = plan.buffer(host.nbytes)
        obuf.upload(host.ravel(order="F"))
        plan.reduce_signal(sbuf, max(nspins, 1), nadc, obuf, nadc + 4, offset_elems=2)
        flat = np.zeros(host.size, dtype=dtype)
        obuf.download(flat)
        out = flat.reshape(host.shape, order="F")
    finally:
        plan.close()
    t = 5e-6 if precision == "f32" else 5e-13
    scale = max(np.linalg.norm(ref), np.sqrt(nspins * nadc * ncoils), 1)
    assert np.linalg.norm(out[2:nadc + 2] - ref) / scale < t
    assert np.all(np.isfinite(out))
    assert np.all(out[[0, 1, nadc + 2, nadc + 3]] == complex(7, -3))


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_coil_reduction(dtype):
    require_gpu()
    for nspins in (0, 1, 259, 2049, 32769):
        for nadc, nc in ((0, 3), (1, 1), (15, 3), (32, 32)):
            check_reduction(dtype, nspins, nadc, nc)
    # several sample and coil tiles over interleaved spin chunks
    check_reduction(dtype, 4097, 40, 33)
    check_reduction(dtype, 70001, 65, 47)
    check_reduction(np.complex64, 1, 257, 257)


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_reduction_cancellation(precision):
    """Alternating signs cancel; the remainder must survive (test_gpu_reduction.jl:55-67)."""
    require_gpu()
    dtype = np.complex64 if precision == "f32" else np.complex128
    st = np.full((4097, 3), 1 - 2j, dtype=dtype)
    st[1:4096:2] *= -1
    st[-1] = 1e-3 + 2e-3j
    cs = np.full((4097, 3), 0.5 + 0.25j, dtype=dtype)
    ref = st.astype(np.complex128).T @ cs.astype(np.complex128)
    from hom_amd.types import HO_Phantom
    obj = HO_Phantom(x=np.zeros(4097), csm=np.asfortranarray(cs.astype(np.complex128)))
    with Plan(precision, 4097, 3, 2, 3) as plan:
        plan.upload_phantom(obj)
        sbuf = plan.buffer(st.nbytes)
        sbuf.upload(np.asfortranarray(st).ravel(order="F"))
        obuf = plan.buffer(9 * np.dtype(dtype).itemsize)
        plan.reduce_signal(sbuf, 4097, 3, obuf, 3)
        out = np.zeros(9, dtype=dtype)
        obuf.download(out)
    atol = 1e-7 if precision == "f32" else 1e-14
    np.testing.assert_allclose(out.reshape(3, 3, order="F"), ref, atol=atol, rtol=0)


def test_block_length_invariance_and_run():
    """hom_run over 1/2/32/128-node blocks matches the oracle at 1e-11 (f64)."""
    require_gpu()
    obj, seq, _, _ = high_order_workload(nspins=517, nsteps=128, ncoils=4)
    seq.adc[:] = False
    seq.adc[5::3] = True
    m = BlochHighOrder("1111")
    ref, rmxy, rmz = oracle.run_workload(obj, seq, m.flags, GAMMA)
    for L in (1, 2, 32, 128):
        print_case("run", L)
        parts = get_sim_ranges(seq, L)
        nadc_max = max(seq.block_adc(a, b) for a, b in zip(parts[0], parts[1]))
        with make_plan("f64", obj, seq, nadc_max) as plan:
            plan.run(parts, m)
            sig = plan.download_signal()
            mxy, mz = plan.get_state()
        assert_signal(sig, ref, "f64", f"L={L}")
        assert_isapprox(mxy, rmxy, 1e-11, 1e-11, "Mxy")
        assert_isapprox(mz, rmz, 1e-11, 1e-11, "Mz")


def test_rf_and_precession_run_small_config():
    """RF + spiral readout (the 100 x 100 2D phantom, 1/100 of its spins)."""
    require_gpu()
    from hom_amd.workloads import make_config
    obj, seq, _ = make_config(0, spin_subsample=100)
    m = BlochHighOrder("1111")
    parts = get_sim_ranges(seq, 32, 32)
    ref, _, _ = oracle.run(obj, seq, parts, m.flags, GAMMA)
    nadc_max = max(seq.block_adc(a, b) for a, b in zip(parts[0], parts[1]))
    for precision in ("f64", "f32"):
        print_case("config0-sub", precision)
        with make_plan(precision, obj, seq, nadc_max) as plan:
            plan.run(parts, m)
            sig = plan.download_signal()
        err = np.linalg.norm(sig - ref) / np.linalg.norm(ref)
        print(f"config0-sub {precision} NRMSE {err:.3e}", flush=True)
        if precision == "f64":
            assert_signal(sig, ref, "f64", precision)
        else:
            assert err < 1e-3
