"""Public simulate / prepare_simulation surface on the GPU, against closed forms
and the oracle.  Mirrors test/unit/test_koma_compatibility.jl:55-135,
test/unit/test_phantom_construction.jl:123-150,
test/gpu/test_gpu_compatibility.jl:61-230 and test/gpu/test_prepared_simulation.jl."""
import numpy as np
import pytest

import hom_amd
import oracle
from hom_amd.raw import RawAcquisitionData
from hom_amd import BlochHighOrder, HO_Phantom, prepare_simulation, simulate
from hom_amd.types import GAMMA, get_sim_ranges
from hom_amd.workloads import FreeBlock, RFBlock, build_sequence
from helpers import assert_isapprox, assert_signal, nrmse, print_case, require_gpu

pytestmark = pytest.mark.gpu
TAU = 1e-4
TIMES = 1e-3 * np.arange(1, 6) / 5


def compat(h_readout=None, csm=None, rho=(1.0, 0.7), x=(0.0, 0.01), dw=None, inf=True):
    if csm is None:
        csm = np.array([[1, 1j], [0.5 - 0.2j, -0.3 + 0.1j]])
    obj = HO_Phantom(x=np.array(x), rho=np.array(rho), csm=csm, dw=dw,
                     T1=np.full(len(x), np.inf) if inf else None,
                     T2=np.full(len(x), np.inf) if inf else None)

    def fields(t):
        out = np.zeros((16, t.size))
        for j, v in (h_readout or {}).items():
            out[j, t >= TAU] = v
        return out

    seq = build_sequence([RFBlock(TAU, 1 / (4 * GAMMA * TAU)), FreeBlock(2e-3, TIMES)], fields)
    return obj, seq


def params(**kw):
    p = {"precision": "f64", "sim_method": BlochHighOrder("1111"), "return_type": "mat"}
    p.update(kw)
    return p


def test_legacy_api_state_ownership_and_kats():
    require_gpu()
    obj, seq = compat()
    density = obj.rho.copy()
    p = params(sim_method=BlochHighOrder("0000"), Nblocks=1)
    original = dict(p)
    sig = simulate(obj, seq, sim_params=p)
    expected = 1j * (density @ obj.csm)
    assert sig.shape == (5, 2, 1)
    np.testing.assert_allclose(sig[:, :, 0], np.tile(expected, (5, 1)), atol=1e-12)
    assert np.array_equal(obj.rho, density)
    assert p == original
    seen = []
    split = simulate(obj, seq, sim_params={**p, "Nblocks": 1000},
                     callbacks=[lambda prog, *_: seen.append(prog["samples"] - 1)])
    np.testing.assert_allclose(split, sig, atol=1e-12)
    assert seen == sorted(seen) and seen[-1] == 5
    single = simulate(obj, seq, sim_params={**p, "precision": "f32"})
    assert single.dtype == np.complex64
    assert_isapprox(single, sig, 2e-5, 2e-5)
    mixed = simulate(obj, seq, sim_params={**p, "precision": "mixed"})
    assert nrmse(mixed, sig) <= 1e-5
    state = simulate(obj, seq, sim_params={**p, "return_type": "state"})
    np.testing.assert_allclose(state.xy, 1j * density, atol=1e-12)
    raw = simulate(obj, seq, sim_params={**p, "return_type": "raw"})
    # reference test_koma_compatibility.jl:99-105
    assert isinstance(raw, RawAcquisitionData)
    assert sum(pr.data.shape[0] for pr in raw.profiles) == 5
    assert all(pr.data.shape[1] == 2 for pr in raw.profiles)
    assert raw.params["userParameters"]["Nblocks"] >= 2
    np.testing.assert_allclose(np.vstack([pr.data for pr in raw.profiles]),
                               sig[:, :, 0].astype(np.complex64), rtol=1e-6, atol=1e-7)


def test_zeroth_order_phase_and_rf_ignores_fields():
    require_gpu()
    obj, seq = compat(h_readout={0: 20 / GAMMA})
    for precision in ("f64", "mixed"):
        sig = simulate(obj, seq, sim_params=params(sim_method=BlochHighOrder("1000"), precision=precision))
        expected = 1j * (obj.rho @ obj.csm)
        ref = np.exp(-1j * 2 * np.pi * 20 * TIMES)[:, None] * expected
        if precision == "f64":
            np.testing.assert_allclose(sig[:, :, 0], ref, atol=1e-12)
        else:
            assert nrmse(sig[:, :, 0], ref) <= 1e-5


def test_ideal_receive_coil_and_weights():
    require_gpu()
    for precision, tol in (("f32", 20 * np.finfo(np.float32).eps), ("f64", 20 * np.finfo(float).eps)):
        print_case("ideal-coil", precision)
        obj, seq = compat(csm=np.ones((2, 1)))
        sig = simulate(obj, seq, sim_params=params(precision=precision))
        np.testing.assert_allclose(np.abs(sig[:, 0, 0]), obj.rho.sum(), rtol=tol)
        w = np.array([[0, 2 + 1j], [0, 2 + 1j]])
        objw, seqw = compat(csm=w)
        res = simulate(objw, seqw, sim_params=params(precision=precision))
        assert np.all(res[:, 0, 0] == 0)
        np.testing.assert_allclose(res[:, 1, 0], (2 + 1j) * sig[:, 0, 0], rtol=tol)


@pytest.mark.parametrize("precision", ["f32", "f64", "mixed"])
def test_high_order_compatibility_all_coil_modes(precision):
    """test_gpu_compatibility.jl:125-171: measured fields on every term, off-resonance,
    masks 0000/1000/1111, every coil_combination, block lengths 2/1."""
    require_gpu()
    for mask in ("0000", "1000", "1111"):
        obj, seq = compat(h_readout={j: (j + 1) * 1e-7 for j in range(16)}, dw=np.array([5.0, 25.0]),
                          inf=True)
        m = BlochHighOrder(mask)
        parts = get_sim_ranges(seq, 2, 1)
        ref, _, _ = oracle.run(obj, seq, parts, m.flags, GAMMA)
        for mode in ("gemm", "reduction", "fused_warp", "fused_tiled"):
            print_case("compat", precision, mask, mode)
            snaps = []

            def seen(progress, blocks, data, sim_params):
                # the rows of the blocks done so far, as the callback sees them
                snaps.append(data["sig"][:progress["samples"] - 1].copy())

            sig = simulate(obj, seq, sim_params=params(precision=precision, sim_method=m,
                                                       coil_combination=mode, max_block_length=2,
                                                       max_rf_block_length=1),
                           callbacks=[seen])
            assert len(snaps) > 0
            for rows in snaps:
                np.testing.assert_array_equal(rows, sig[:rows.shape[0], :, 0])
            if precision == "mixed":
                assert nrmse(sig[:, :, 0], ref) <= 1e-5
            else:
                assert_signal(sig[:, :, 0], ref, precision)
            assert np.all(np.isfinite(sig))


def test_block_length_invariance_public():
    require_gpu()
    obj, seq = compat(h_readout={j: (j + 1) * 1e-7 for j in range(16)})
    ref = simulate(obj, seq, sim_params=params(max_block_length=1))
    for L in (2, 32, 128):
        sig = simulate(obj, seq, sim_params=params(max_block_length=L))
        assert_isapprox(sig, ref, 1e-11, 1e-11)


def test_prepared_simulation_reuse_and_failures():
    require_gpu()
    obj, seq = compat(h_readout={j: (j + 1) * 1e-7 for j in range(16)}, inf=False)
    one_shot = simulate(obj, seq, sim_params=params())
    plan = prepare_simulation(obj, seq, sim_params=params())
    try:
        a = plan.simulate()
        b = plan.simulate()
        np.testing.assert_array_equal(a, b)
        assert_isapprox(a, one_shot, 1e-12, 1e-12)

        def failing(*_):
            raise RuntimeError("callback failure")
        with pytest.raises(RuntimeError, match="callback failure"):
            plan.simulate(callbacks=[failing])
        c = plan.simulate()
        np.testing.assert_array_equal(a, c)

        def recurse(*_):
            plan.simulate()
        with pytest.raises(ValueError, match="own callback"):
            plan.simulate(callbacks=[recurse])

        def zero_state(progress, blocks, data, p):
            if progress["block"] == 1:
                data["Xt"].xy[:] = 0
                data["Xt"].z[:] = 0
        edited = plan.simulate(callbacks=[zero_state])
        assert np.abs(edited).max() < np.abs(a).max()
    finally:
        plan.close()
        plan.close()
    with pytest.raises(ValueError):
        plan.simulate()


def test_mixed_per_block_stepping_runs_mixed_kernels():
    """A mixed plan stepped block by block (callbacks, synchronization="block")
    runs each free-precession block on the mixed segment kernels, not the FP64
    per-block path: hom_plan_mixed_stats counts the block as a mixed segment,
    and the ADC-state buffer is never grown to the block's sample count."""
    require_gpu()
    from hom_amd.runtime import Plan
    from hom_amd.workloads import make_config
    obj, seq, cfg = make_config(1, spin_subsample=500)
    m = BlochHighOrder("1111")
    parts = get_sim_ranges(seq, 10 ** 9, 10 ** 9)
    ref, rmxy, rmz = oracle.run(obj, seq, parts, m.flags, GAMMA)
    adc = [seq.block_adc(a, b) for a, b in zip(parts[0], parts[1])]
    total = int(sum(adc))
    print_case("mixed-per-block", len(obj), len(parts[0]))
    with Plan("mixed", len(obj), obj.n_coils, seq.n_nodes, max(adc), max_rf_adc=0) as plan:
        plan.upload_phantom(obj)
        plan.upload_sequence(seq)
        plan.reset_state()
        buf = plan.buffer(16 * total * obj.n_coils)
        row = 0
        for b in range(len(parts[0])):
            first, last, rf = int(parts[0][b]), int(parts[1][b]), bool(parts[2][b])
            fn = plan.excitation_block if rf else plan.precession_block
            fn(first, last, m, buf if adc[b] else None, total, offset_elems=row)
            if not rf and adc[b]:
                assert plan.mixed_stats() == (1, 0)
            row += adc[b]
        plan.synchronize()
        flat = np.zeros(total * obj.n_coils, dtype=np.complex128)
        buf.download(flat)
        sig = flat.reshape((total, obj.n_coils), order="F")
        xy, z = plan.get_state()
    assert nrmse(sig, ref) <= 1e-5
    assert_isapprox(xy, rmxy, 1e-11, 1e-11, "Mxy")
    assert_isapprox(z, rmz, 1e-11, 1e-11, "Mz")
