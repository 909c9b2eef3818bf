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
