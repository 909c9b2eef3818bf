"""CPU-side checks of the boundary: the C ABI library loads and exports every
symbol include/hom_abi.h declares, the host mirror validates like the
reference (src/simulation/BlochHighOrder.jl:21-43, test/unit/test_phantom_interfaces.jl:88-103),
and invalid shapes are rejected before any device work."""
import re
from pathlib import Path

import numpy as np
import pytest

import hom_amd
from hom_amd import runtime as rt
from hom_amd.types import BlochHighOrder, DimensionMismatch, HO_DiscreteSequence, HO_Phantom

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    text = (ROOT / "include" / "hom_abi.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hom_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = rt.load_library()
    declared = header_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(rt.ABI_SYMBOLS)
    assert lib.hom_abi_version() == 1


def test_plan_create_without_device_or_bad_args_reports_error():
    lib = rt.load_library()
    # Argument checks run before any device query.
    with pytest.raises(ValueError, match="maximum ADC"):
        rt.Plan("f64", 4, 1, 3, -1)
    with pytest.raises(ValueError, match="RF ADC"):
        rt.Plan("f64", 4, 1, 3, 2, max_rf_adc=3)
    with pytest.raises(ValueError, match="coil_combination"):
        rt.Plan("f64", 4, 1, 3, 2, coil_combination="magic")
    with pytest.raises(ValueError):
        rt.Plan("f16", 4, 1, 3, 2)
    if rt.device_count() == 0:
        with pytest.raises(rt.HomDeviceError, match="No HIP device"):
            rt.Plan("f64", 4, 1, 3, 2)
    assert lib is rt.load_library()


def test_workspace_keyword_checks():
    """HighOrderGPUWorkspace's keyword checks (HighOrderGPU.jl:24-35) as the
    reference's own test drives them (test/gpu/test_gpu_fused_coils.jl:63-75):
    raised before any device work, so they run without a GPU."""
    with pytest.raises(ValueError, match="multiple of 32"):
        rt.Plan("f64", 1, 1, 3, 32, coil_combination="fused_warp", groupsize=48)
    for adc_tile, coil_tile in ((3, 1), (4.0, 1), (4, 0)):
        with pytest.raises(ValueError):
            rt.Plan("f64", 1, 1, 3, 32, coil_combination="fused_tiled", adc_tile=adc_tile,
                    coil_tile=coil_tile)
    # 1024 spins x 8 samples of Complex{Float64} exceed a workgroup's 64 KiB
    with pytest.raises(ValueError, match="shared-memory"):
        rt.Plan("f64", 1, 1, 3, 32, coil_combination="fused_tiled", groupsize=1024)
    for max_rf_adc in (-1, 33, 2.5):
        with pytest.raises(ValueError, match="RF ADC"):
            rt.Plan("f64", 1, 1, 3, 32, coil_combination="fused_tiled", max_rf_adc=max_rf_adc)
    # fused_warp fixes both tiles at 1; gemm ignores the knobs, as the reference does
    rt.check_workspace_options("f64", "fused_warp", 1024, 3, 0)
    rt.check_workspace_options("f32", "gemm", 48, 3, 0)
    rt.check_workspace_options("f32", "fused_tiled", 1024, 8, 4)  # 64 KiB exactly


def test_mask_consistency():
    for bits in range(16):
        name = format(bits, "04b")
        flags = tuple(c == "1" for c in name)
        for m in (BlochHighOrder(name), BlochHighOrder(name, False, True)):
            assert (m.ho0, m.ho1, m.ho2, m.ho3) == flags
            assert m.name == name
    for bad in ("", "000", "00000", "10x1", "111₂"):
        with pytest.raises(ValueError):
            BlochHighOrder(bad)
        with pytest.raises(ValueError):
            BlochHighOrder(bad, False, False)
    with pytest.raises(ValueError):
        BlochHighOrder("1111", ho0=False)
    assert BlochHighOrder(ho0=False).name == "0111"
    m = BlochHighOrder("1010", False, True)
    assert m.flags == (1 | 4 | 32)


def test_phantom_defaults_and_validation():
    p = HO_Phantom(x=[0.0, 0.01])
    assert np.all(p.rho == 1) and np.all(p.T1 == 1e6) and np.all(p.T2 == 1e6)
    assert p.csm.shape == (2, 1) and np.all(p.csm == 1) and p.csm.flags.f_contiguous
    p.csm = np.ones((3, 1), complex)
    with pytest.raises(DimensionMismatch):
        p.validate()


def test_sequence_validation():
    n = 4
    ok = dict(t=np.arange(n) * 1e-3, gx=np.zeros(n), gy=np.zeros(n), gz=np.zeros(n),
              b1=np.zeros(n), df=np.zeros(n), psi=np.zeros(n), adc=np.ones(n, bool),
              excitation=np.zeros(n - 1, bool), h=np.zeros((16, n)))
    s = HO_DiscreteSequence(**ok)
    assert len(s) == 3 and s.n_nodes == 4
    assert s.block_adc(0, 3) == 3
    v = s.view(1, 3)
    assert v.n_nodes == 3 and len(v) == 2
    with pytest.raises(DimensionMismatch):
        HO_DiscreteSequence(**{**ok, "h": np.zeros((15, n))})
    with pytest.raises(DimensionMismatch):
        HO_DiscreteSequence(**{**ok, "excitation": np.zeros(n, bool)})
    with pytest.raises(ValueError):
        HO_DiscreteSequence(**{**ok, "t": np.array([0, 1e-3, 0.5e-3, 2e-3])})


def test_package_surface():
    for name in ("BlochHighOrder", "HO_Phantom", "HO_DiscreteSequence", "get_sim_ranges", "GAMMA"):
        assert hasattr(hom_amd, name)


def serve_header_symbols():
    text = (ROOT / "include" / "hom_serve.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hom_server_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_serving_symbol():
    lib = rt.load_library()
    declared = serve_header_symbols()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(rt.SERVE_SYMBOLS)


def test_server_config_checks_before_device():
    for kw, msg in [({"max_active": 65}, "max_active"), ({"gpt": 100}, "gpt"),
                    ({"main_grid": 12}, "multiple of 8"), ({"step_budget": -1}, "step budget"),
                    ({"device_bytes": -5}, "Arena")]:
        with pytest.raises(ValueError, match=msg):
            rt.Server(**kw)
    if rt.device_count() == 0:
        with pytest.raises(rt.HomDeviceError, match="No HIP device"):
            rt.Server()


def test_request_stream_distribution():
    from hom_amd.serving import request_stream
    a = request_stream(3, 2000, rate=200.0, n_spins_max=1 << 20)
    b = request_stream(3, 2000, rate=200.0, n_spins_max=1 << 20)
    assert a == b
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0)
    assert abs(len(a) / arr[-1] - 200.0) < 20.0  # Poisson rate
    ns = np.array([r.spin_count for r in a])
    assert ns.min() >= 1000 and ns.max() <= 1_000_000
    # log-uniform: the median is near the geometric mean sqrt(1e3 * 1e6)
    assert 0.6 < np.median(ns) / np.sqrt(1e9) < 1.6
    assert all(0 <= r.spin_first <= (1 << 20) - r.spin_count for r in a)
    assert {r.n_coils for r in a} == {1, 8, 16, 32}
    assert {r.order for r in a} == {"1100", "1110", "1111"}
    assert min(r.n_samples for r in a) >= 1000 and max(r.n_samples for r in a) <= 20000
