"""The f32 and f64 drop-in (coil modes "gemm"
and "fused_tiled") on the 350,000-spin, 8-coil, 4 x 4750-sample 3D workload,
run through the public simulate() with the reference demos' 32-node blocks
(demo/SimulationDemoSupport.jl:114).

Against the f64 oracle on every 250th spin (1400 spins, all four interleaves):
the reference's GPU-vs-CPU tolerances (test/gpu/test_gpu_kernels.jl:31:
isapprox f32 3e-5, f64 1e-11), and for f64 NRMSE <= 1e-12.
f32 is compared with the oracle run on inputs rounded to f32, as the
reference converts its inputs before the f32 path runs.

On all 350,000 spins, through properties that hold at any size: two runs
give bit-identical signals (deterministic reductions), and the signal of the
phantom equals the sum of the signals of its two halves (the signal is a sum
over spins; only the rounding of the reduction order differs).
"""
import numpy as np
import pytest

import oracle
from hom_amd.simulate import simulate
from hom_amd.types import GAMMA, BlochHighOrder, get_sim_ranges
from hom_amd.workloads import make_config
from helpers import assert_isapprox, nrmse, print_case, require_gpu, round_f32, tol

pytestmark = pytest.mark.gpu

MODES = [("f64", "gemm"), ("f64", "fused_tiled"), ("f32", "gemm"), ("f32", "fused_tiled")]
BLOCK = 32


def dropin(obj, seq, precision, mode):
    p = {"precision": precision, "coil_combination": mode, "return_type": "mat",
         "sim_method": BlochHighOrder("1111"), "max_block_length": BLOCK,
         "max_rf_block_length": BLOCK}
    return simulate(obj, seq, sim_params=p)[:, :, 0]


@pytest.fixture(scope="module")
def config1():
    return make_config(1)


@pytest.fixture(scope="module")
def oracle_sub(config1):
    obj, seq, cfg = config1
    sub = obj[::250]
    parts = get_sim_ranges(seq, BLOCK, BLOCK)
    flags = BlochHighOrder("1111").flags
    ref64, _, _ = oracle.run(sub, seq, parts, flags, GAMMA, n_threads=16)
    o32, s32 = round_f32(sub, seq)
    ref32, _, _ = oracle.run(o32, s32, parts, flags, GAMMA, n_threads=16)
    return sub, ref64, ref32


@pytest.mark.parametrize("precision,mode", MODES)
def test_config1_dropin_vs_oracle(config1, oracle_sub, precision, mode):
    require_gpu()
    obj, seq, cfg = config1
    sub, ref64, ref32 = oracle_sub
    print_case("config1-dropin-sub", precision, mode, len(sub))
    sig = dropin(sub, seq, precision, mode)
    ref = ref64 if precision == "f64" else ref32
    assert sig.shape == ref.shape == (cfg.n_interleaves * cfg.n_adc, cfg.n_coils)
    t = tol(precision)
    n = cfg.n_adc
    for il in range(cfg.n_interleaves):
        rows = slice(il * n, (il + 1) * n)
        assert_isapprox(sig[rows], ref[rows], t, t, f"interleave {il}")
    e = nrmse(sig, ref)
    print_case("config1-dropin-nrmse", precision, mode, e)
    if precision == "f64":
        assert e <= 1e-12, f"f64 drop-in NRMSE {e:.3e} above the 1e-12 bar"
    else:
        assert e <= 3e-5


@pytest.mark.parametrize("precision,mode", MODES)
def test_config1_dropin_fullsize_properties(config1, precision, mode):
    require_gpu()
    obj, seq, cfg = config1
    print_case("config1-dropin-full", precision, mode, len(obj))
    sig = dropin(obj, seq, precision, mode)
    again = dropin(obj, seq, precision, mode)
    assert np.isfinite(sig).all() and np.abs(sig).max() > 0
    np.testing.assert_array_equal(sig, again)
    cut = len(obj) // 2
    a = dropin(obj[:cut], seq, precision, mode)
    b = dropin(obj[cut:], seq, precision, mode)
    e = nrmse(a.astype(np.complex128) + b, sig.astype(np.complex128))
    print_case("config1-dropin-additivity", precision, mode, e)
    assert e <= (3e-6 if precision == "f32" else 1e-13), f"additivity NRMSE {e:.3e}"
