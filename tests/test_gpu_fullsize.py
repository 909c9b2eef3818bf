"""Full-size GPU checks of the mixed path (configurations 2 and 4 of the
benchmark: 1,048,576 spins x 16 x 8000 samples and 4,000,000 spins x 32 x
(RF + 10000 samples), 32 coils, mask 1111), through properties that hold at
any size:

* determinism: two runs give bit-identical signals;
* linearity in rho: the whole state is linear in the equilibrium
  magnetization, and the state scale is a power of two of max |M0|, so
  doubling rho doubles the signal exactly (bitwise);
* additivity over spins: the signal of a phantom equals the sum of the
  signals of two halves split at a 4096-spin tile boundary (NRMSE <= 1e-6:
  the halves get their own power-of-two scales);
* parity on a spin subsample against the f64 oracle (NRMSE <= 1e-5), run
  with the full-size run's spin tile so each FP32 partial sum covers as many
  spins as there;
* the public simulate() with its default whole-run blocks gives the same
  signal as 32-node blocks (the mixed path fuses free precession anyway).
"""
import numpy as np
import pytest

import oracle
from hom_amd.runtime import Plan
from hom_amd.types import GAMMA, BlochHighOrder, HO_Phantom, get_sim_ranges
from hom_amd.workloads import make_config
from helpers import nrmse, print_case, require_gpu

pytestmark = pytest.mark.gpu


def run(obj, seq, parts, method, twice=False, tile=None):
    """(signal, signal of a second run or None, spin tile the run used)."""
    nadc = max(seq.block_adc(a, b) for a, b in zip(parts[0], parts[1]))
    with Plan("mixed", len(obj), obj.n_coils, seq.n_nodes, nadc) as plan:
        plan.upload_phantom(obj)
        plan.upload_sequence(seq)
        if tile is not None:
            plan.spin_tile = tile
        plan.run(parts, method)
        first = plan.download_signal()
        used = plan.last_spin_tile
        if not twice:
            return first, None, used
        plan.reset_state()
        plan.run(parts, method)
        return first, plan.download_signal(), used


def scaled(obj, k):
    return HO_Phantom(x=obj.x, y=obj.y, z=obj.z, rho=obj.rho * k, T1=obj.T1, T2=obj.T2,
                      dw=obj.dw, csm=obj.csm)


@pytest.mark.parametrize("config,subsample", [(2, 512), (4, 2000)])
def test_fullsize_properties(config, subsample):
    require_gpu()
    print_case("fullsize", config)
    obj, seq, cfg = make_config(config)
    method = BlochHighOrder("1111")
    parts = get_sim_ranges(seq, 32, 32)
    sig, again, tile = run(obj, seq, parts, method, twice=True)
    np.testing.assert_array_equal(sig, again)
    assert np.isfinite(sig).all() and np.abs(sig).max() > 0

    # the public entry point with its default whole-run blocks: each readout
    # is one block of n_adc samples, and the plan must not size an ADC-state
    # buffer for it (4M spins x 10000 samples would be 640 GB)
    from hom_amd.simulate import simulate
    mat = simulate(obj, seq, sim_params={"precision": "mixed", "return_type": "mat",
                                         "sim_method": method})
    e = nrmse(mat[:, :, 0], sig)
    assert e <= 1e-12, f"simulate() with whole-readout blocks: NRMSE {e:.3e}"

    sig2, _, _ = run(scaled(obj, 2.0), seq, parts, method)
    np.testing.assert_array_equal(sig2, 2.0 * sig)

    cut = (len(obj) // 2) // 4096 * 4096
    a, _, _ = run(obj[:cut], seq, parts, method)
    b, _, _ = run(obj[cut:], seq, parts, method)
    e = nrmse(a + b, sig)
    assert e <= 1e-6, f"additivity NRMSE {e:.3e}"

    sub = obj[::subsample]
    ref, _, _ = oracle.run(sub, seq, parts, method.flags, GAMMA, n_threads=16)
    s_sub, _, _ = run(sub, seq, parts, method, tile=tile)
    n = cfg.n_adc
    worst = max(nrmse(s_sub[i * n:(i + 1) * n], ref[i * n:(i + 1) * n])
                for i in range(cfg.n_interleaves))
    print_case("fullsize-nrmse", config, nrmse(s_sub, ref), worst)
    assert worst <= 1e-5, f"per-interleave NRMSE {worst:.3e}"
