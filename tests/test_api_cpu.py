"""Validation of the public surface before any device work
(src/simulation/Simulate.jl:70-135; test/gpu/test_gpu_compatibility.jl:202-223)."""
import numpy as np
import pytest

import hom_amd
from hom_amd import BlochHighOrder, HO_Phantom, setup_simulation
from hom_amd.workloads import FreeBlock, RFBlock, build_sequence, high_order_workload


def compat():
    obj = HO_Phantom(x=[0.0, 0.01], rho=[1.0, 0.7], T1=[np.inf] * 2, T2=[np.inf] * 2,
                     csm=np.array([[1, 1j], [0.5 - 0.2j, -0.3 + 0.1j]]))
    seq = build_sequence([RFBlock(1e-4, 1 / (4 * hom_amd.GAMMA * 1e-4)),
                          FreeBlock(2e-3, 1e-3 * np.arange(1, 6) / 5)])
    return obj, seq


@pytest.mark.parametrize("bad", [
    {"implementation": "unknown"}, {"implementation": "array"}, {"coil_combination": "unknown"},
    {"return_type": "image"}, {"precision": "f16"}, {"Nthreads": 0}, {"Nblocks": 0},
    {"synchronization": "never"}, {"max_block_length": 0}, {"sim_method": "Bloch"}])
def test_rejects_bad_params(bad):
    obj, seq = compat()
    with pytest.raises(ValueError):
        setup_simulation(obj, seq, bad)


def test_rejects_bad_inputs():
    obj, seq = compat()
    with pytest.raises(ValueError):
        setup_simulation(HO_Phantom(x=np.zeros(0)), seq, {})
    with pytest.raises(ValueError):
        setup_simulation(obj, seq, {}, gpus=[0, 1])
    with pytest.raises(ValueError):
        setup_simulation(obj, seq, {1: 2})
    obj.csm = np.ones((3, 1), complex)
    with pytest.raises(hom_amd.DimensionMismatch):
        setup_simulation(obj, seq, {})


def test_block_plan_and_legacy_nblocks():
    obj, seq = compat()
    s = setup_simulation(obj, seq, {"max_block_length": 2, "max_rf_block_length": 1})
    assert s.adc_counts.sum() == 5
    assert s.max_adc == 2
    assert s.max_rf_adc == 0
    assert np.all(s.parts[0][1:] == s.parts[1][:-1])
    legacy = setup_simulation(obj, seq, {"Nblocks": 3})
    assert legacy.params["max_block_length"] == -(-len(seq) // 3)
    params = {"precision": "f64"}
    setup_simulation(obj, seq, params)
    assert params == {"precision": "f64"}


def test_simulation_spin_ranges():
    # test/unit/test_execution_options.jl:4-7 (1-based 1:2, 3:5, 6:8, 9:11)
    ranges = hom_amd.simulation_spin_ranges(11, 4)
    assert [(r.start, r.stop) for r in ranges] == [(0, 2), (2, 5), (5, 8), (8, 11)]
    assert [i for r in ranges for i in r] == list(range(11))
    with pytest.raises(ValueError):
        hom_amd.simulation_spin_ranges(2, 3)
    with pytest.raises(ValueError):
        hom_amd.simulation_spin_ranges(5, 0)


@pytest.mark.parametrize("gpus", [[], [0, 0], [-1], [True], [0.5], 0, "0", [0, 1, 2]])
def test_rejects_bad_gpu_lists(gpus):
    # test/unit/test_execution_options.jl:19-23: checked before any device work
    obj, seq = compat()
    with pytest.raises(ValueError):
        setup_simulation(obj, seq, {"sim_method": BlochHighOrder("0000")}, gpus=gpus)


def test_more_gpus_than_spins_and_string_keys():
    obj, seq = compat()
    with pytest.raises(ValueError, match="at least one spin"):
        setup_simulation(obj, seq, {}, gpus=[0, 1, 2])
    with pytest.raises(ValueError, match="string keys"):
        setup_simulation(obj, seq, {"precision": "f64", 3: 1})
    s = setup_simulation(obj, seq, {"sim_method": BlochHighOrder("0000")})
    assert s.params["gpu"] is True and s.params["gpu_device"] == 0 and s.params["gpus"] == [0]
    assert s.params["implementation"] == "kernel"
    assert s.params["max_adc_per_block"] == s.max_adc
    assert s.params["adc_state_buffer_samples"] == s.max_adc  # f32 "gemm": every block's ADC
