"""The mixed path's segment loop on two streams, under short segments.

Each precession segment of a mixed run uses one of two operand slots (the
phase seeds, log-magnitudes, initial phases, |M0| maximum and spin-tile
partials).  Its main kernel runs on the plan's second stream while the RF
blocks and the other segments' prep, end-state and finalize kernels run on
the plan's stream.  Here many readouts of a few samples are separated by RF
pulses, and some readouts take no samples.  Every main kernel is then short
and the slots are reused quickly: a segment that read or wrote the wrong
slot, or one ordering event too few, would change the signal between two
runs of one plan or move it away from the f64 drop-in and the f64 oracle."""
import numpy as np
import pytest

import oracle
from hom_amd.runtime import Plan
from hom_amd.types import GAMMA, BlochHighOrder, HO_Phantom, get_sim_ranges
from hom_amd.workloads import FreeBlock, RFBlock, build_sequence
from helpers import assert_isapprox, nrmse, print_case, require_gpu

pytestmark = pytest.mark.gpu


def short_segments_case(seed=7, ns=30000, nc=8, nread=24):
    rng = np.random.default_rng(seed)
    obj = HO_Phantom(
        x=rng.uniform(-0.1, 0.1, ns), y=rng.uniform(-0.1, 0.1, ns), z=rng.uniform(-0.035, 0.035, ns),
        rho=rng.uniform(0.0, 1.0, ns), T1=rng.uniform(0.5, 2.0, ns), T2=rng.uniform(0.02, 0.2, ns),
        dw=2 * np.pi * rng.normal(0.0, 20.0, ns),
        csm=np.asfortranarray(0.5 * (rng.normal(size=(ns, nc)) + 1j * rng.normal(size=(ns, nc)))))
    blocks = []
    for r in range(nread):
        tau = 1e-4
        b1 = rng.uniform(10.0, 60.0) / 360.0 / (GAMMA * tau) * np.exp(1j * rng.uniform(0.0, 2 * np.pi))
        blocks.append(RFBlock(tau, complex(b1)))
        dur = float(rng.uniform(1e-4, 4e-4))
        n_adc = 0 if r % 5 == 3 else int(rng.integers(1, 49))
        adc = np.sort(rng.choice(np.arange(1, 1001), size=n_adc, replace=False)) / 1000.0 * dur
        g = rng.uniform(-5e-3, 5e-3, 3)

        def grad(rel, g=g):
            return np.vstack([np.full(rel.size, g[0]), np.full(rel.size, g[1]), np.full(rel.size, g[2])])
        blocks.append(FreeBlock(dur, adc, grad))
    amp = np.array([1e-6, 0, 0, 0] + [1e-4] * 5 + [1e-3] * 7)
    fq = rng.uniform(10.0, 500.0, 16)
    ph = rng.uniform(0.0, 2 * np.pi, 16)

    def fields(t):
        return amp[:, None] * np.sin(2 * np.pi * fq[:, None] * t[None, :] + ph[:, None])
    return obj, build_sequence(blocks, fields=fields, dt_max=5e-5)


def test_short_segments_two_streams():
    require_gpu()
    obj, seq = short_segments_case()
    parts = get_sim_ranges(seq, 32, 32)
    method = BlochHighOrder("1111")
    max_adc = max([seq.block_adc(a, b) for a, b in zip(parts[0], parts[1])] + [1])
    print_case("short segments", len(obj), obj.n_coils, seq.n_nodes, int(seq.adc.sum()), max_adc)
    with Plan("mixed", len(obj), obj.n_coils, seq.n_nodes, max_adc) as plan:
        plan.upload_phantom(obj)
        plan.upload_sequence(seq)
        runs = []
        for _ in range(3):
            plan.reset_state()
            plan.run(parts, method)
            runs.append((plan.download_signal(), *plan.get_state()))
        segments, fp64_segments = plan.mixed_stats()
    assert segments >= 15 and fp64_segments == 0  # sampled segments, each on the split body
    for sig, mxy, mz in runs[1:]:
        assert np.array_equal(sig, runs[0][0])
        assert np.array_equal(mxy, runs[0][1]) and np.array_equal(mz, runs[0][2])
    with Plan("f64", len(obj), obj.n_coils, seq.n_nodes, max_adc) as plan:
        plan.upload_phantom(obj)
        plan.upload_sequence(seq)
        plan.run(parts, method)
        ref64 = plan.download_signal()
    sig = runs[0][0]
    assert sig.shape == ref64.shape and np.linalg.norm(ref64) > 0
    assert nrmse(sig, ref64) <= 1e-5, f"mixed vs f64 drop-in NRMSE {nrmse(sig, ref64):.3e}"
    # the f64 oracle on every 50th spin (the signal is linear in the spins)
    sub = obj[np.arange(0, len(obj), 50)]
    ref, rmxy, rmz = oracle.run(sub, seq, parts, method.flags, GAMMA)
    with Plan("mixed", len(sub), sub.n_coils, seq.n_nodes, max_adc) as plan:
        plan.upload_phantom(sub)
        plan.upload_sequence(seq)
        plan.run(parts, method)
        s = plan.download_signal()
        mxy, mz = plan.get_state()
    assert nrmse(s, ref) <= 1e-5, f"mixed vs oracle NRMSE {nrmse(s, ref):.3e}"
    assert_isapprox(mxy, rmxy, 1e-11, 1e-11, "mixed Mxy")
    assert_isapprox(mz, rmz, 1e-11, 1e-11, "mixed Mz")
