"""Acquisition during RF through the public simulate(), as the reference's
test/gpu/test_gpu_rf_buffers.jl:64-134 drives it: a 90 deg pulse with ADC
samples inside it, an optional readout, a 45 deg pulse with ADC inside it,
over three spins and three coils, for RF blocks of one node and of any
length, in f32, f64 and mixed precision and every coil combination.

* the signal matches the f64 oracle (f32 3e-5, f64 1e-11 and NRMSE <= 1e-12,
  mixed NRMSE <= 1e-5; exactly zero without RF);
* the block metadata callbacks see (Simulate.jl:173-175): max_rf_adc_per_block
  is min(max(RF ADC counts), RF block length) and the ADC-state buffer holds
  every block's samples for "gemm" / "reduction" but only the RF blocks' for
  the fused modes and the mixed path;
* the fused modes give the gemm signal and bit-identical end states;
* "raw" output carries every sample in its profiles.
The sequence is this package's own discretization of the reference's
rf_acquisition_sequence (test/support/compatibility_workload.jl:19-28);
KomaMRIBase's sample placement is not in the reference tree."""
import numpy as np
import pytest

import oracle
from hom_amd.raw import RawAcquisitionData
from hom_amd.simulate import simulate
from hom_amd.types import GAMMA, BlochHighOrder, HO_Phantom
from hom_amd.workloads import FreeBlock, RFBlock, build_sequence
from helpers import assert_signal, nrmse, print_case, require_gpu

pytestmark = pytest.mark.gpu

TAU = 120e-6
SCENARIOS = [(0, 0, 11, True), (3, 2, 11, True), (5, 2, 0, True), (0, 0, 11, False), (0, 0, 0, True)]


def phantom():
    csm = np.array([[1, 1j, 0.5], [0.5 - 0.2j, -0.3 + 0.1j, 0.2j], [0.7, 0.2j, 0.4 - 0.3j]],
                   dtype=np.complex128)
    return HO_Phantom(x=[-0.01, 0.0, 0.01], y=[0.02, -0.01, 0.0], rho=[1.0, 0.7, 0.2],
                      T1=[0.9, 1.0, 1.1], T2=[0.07, 0.08, 0.09], dw=[-20.0, 5.0, 25.0],
                      csm=np.asfortranarray(csm))


def rf_adc(n):
    if n == 0:
        return None
    return np.array([TAU / 2]) if n == 1 else TAU / 4 + TAU / 2 * np.arange(n) / (n - 1)


def sequence(nrf_first, nrf_last, npre, rf_on):
    b1 = 1.0 / (4 * GAMMA * TAU) if rf_on else 0.0
    pre = None if npre == 0 else 1e-4 + 1e-3 * np.arange(npre) / (npre - 1)
    zero = (lambda rel: np.zeros((3, rel.size)))
    return build_sequence([RFBlock(TAU, complex(b1), adc_times=rf_adc(nrf_first)),
                           FreeBlock(2e-3, pre, zero),
                           RFBlock(TAU, complex(b1 / 2), adc_times=rf_adc(nrf_last))])


@pytest.mark.parametrize("scenario", SCENARIOS)
@pytest.mark.parametrize("rf_limit", [1, np.inf])
def test_rf_acquisition_public_simulate(scenario, rf_limit):
    require_gpu()
    nrf_first, nrf_last, npre, rf_on = scenario
    obj, seq = phantom(), sequence(*scenario)
    method = BlochHighOrder("1111")
    ref, _, _ = oracle.run_workload(obj, seq, method.flags, GAMMA, max_block_length=32,
                                    max_rf_block_length=10 ** 9 if rf_limit == np.inf else rf_limit)
    total = nrf_first + nrf_last + npre
    assert ref.shape == (total, 3)
    expected_rf = min(max(nrf_first, nrf_last), rf_limit) if rf_on else 0
    for precision in ("f32", "f64", "mixed"):
        base = {"precision": precision, "sim_method": method, "return_type": "mat",
                "max_block_length": 32, "max_rf_block_length": rf_limit}
        reports, sigs, states = {}, {}, {}
        modes = ("gemm", "reduction") if precision == "mixed" else \
            ("gemm", "reduction", "fused_warp", "fused_tiled")
        for mode in modes:
            print_case("rf-buffers", scenario, rf_limit, precision, mode)
            params = dict(base, coil_combination=mode)
            owned = dict(params)

            def capture(progress, blocks, data, actual, mode=mode):
                reports.setdefault(mode, (actual["max_adc_per_block"], actual["max_rf_adc_per_block"],
                                          actual["adc_state_buffer_samples"], actual["implementation"]))
            sig = simulate(obj, seq, sim_params=params, callbacks=[capture])
            assert params == owned
            assert sig.shape == (total, 3, 1)
            assert np.isfinite(sig).all()
            sigs[mode] = sig[:, :, 0]
            states[mode] = simulate(obj, seq, sim_params=dict(params, return_type="state"))
            if not rf_on or total == 0:
                assert not sig.any()
            elif precision == "mixed":
                assert nrmse(sig[:, :, 0], ref) <= 1e-5
            else:
                assert_signal(sig[:, :, 0], ref, precision, f"{precision} {mode}")
            if expected_rf > 0:
                raw = simulate(obj, seq, sim_params=dict(params, return_type="raw"))
                assert isinstance(raw, RawAcquisitionData)
                assert sum(pr.data.shape[0] for pr in raw.profiles) == total
                assert all(pr.data.shape[1] == 3 for pr in raw.profiles)
        maximum, rf, capacity, impl = reports["gemm"]
        assert impl == "kernel"
        assert rf == expected_rf
        assert capacity == (expected_rf if precision == "mixed" else maximum)
        for mode in modes[1:]:
            m, r, c, _ = reports[mode]
            assert (m, r) == (maximum, expected_rf)
            fused = mode in ("fused_warp", "fused_tiled") or precision == "mixed"
            assert c == (expected_rf if fused else maximum)
            if total:
                assert_signal(sigs[mode], sigs["gemm"], "f32" if precision == "mixed" else precision,
                              f"{precision} {mode} vs gemm")
            np.testing.assert_array_equal(states[mode].xy, states["gemm"].xy)
            np.testing.assert_array_equal(states[mode].z, states["gemm"].z)
