"""Seeded randomized parity sweep of the mixed path and the f64 drop-in
against the f64 oracle.

Each case draws: a phantom of 1..3000 spins (partial 16-spin groups and spin
tiles) with the configuration-0 tissue distributions and 1..40 random complex
coils (partial MFMA column blocks, a second 32-coil chunk); 1..3 hard pulses
of random flip, phase, off-resonance and frame phase, some acquiring during
the pulse, each followed by a readout of 0..700 samples at random times under
sinusoidal gradients; 16 field terms of the configuration-0 amplitudes; a
random field-order mask and off-resonance flags; a random block length.

Strong cases scale the first-order waveforms (1x..8x) and the 2nd/3rd-order
fields (1x..200x), so that both phase bodies of the mixed path (split offsets
and the all-FP64 fallback) and the edges of the split body's bounds are hit.
HOM_RANDOM_SWEEP=N widens the strong cases to N seeds (4000 seeds ran with a
worst mixed NRMSE of 1.8e-6 and f64 5.2e-13).

Bars: mixed NRMSE <= 1e-5 on the signal and 1e-11 on the FP64 end state;
f64 drop-in 1e-11 (atol = rtol, norm-wise, as test/gpu/test_gpu_kernels.jl:31).
"""
import copy
import os

import numpy as np
import pytest

import oracle
from hom_amd.runtime import Plan
from hom_amd.types import GAMMA, BlochHighOrder, HO_Phantom, get_sim_ranges
from hom_amd.workloads import FreeBlock, RFBlock, build_sequence
from helpers import assert_isapprox, assert_signal, nrmse, print_case, require_gpu

pytestmark = pytest.mark.gpu

SEEDS = list(range(1000, 1032))
STRONG_SEEDS = list(range(10000, 10000 + int(os.environ.get("HOM_RANDOM_SWEEP", "24"))))


def random_case(seed):
    rng = np.random.default_rng(seed)
    ns = int(rng.integers(1, 3001))
    nc = int(rng.integers(1, 41))
    obj = HO_Phantom(
        x=rng.uniform(-0.1, 0.1, ns), y=rng.uniform(-0.1, 0.1, ns), z=rng.uniform(-0.035, 0.035, ns),
        rho=rng.uniform(0.0, 1.0, ns), T1=rng.uniform(0.5, 2.0, ns), T2=rng.uniform(0.02, 0.2, ns),
        dw=2 * np.pi * rng.normal(0.0, 20.0, ns),
        csm=np.asfortranarray(0.5 * (rng.normal(size=(ns, nc)) + 1j * rng.normal(size=(ns, nc)))))
    blocks = []
    for _ in range(int(rng.integers(1, 4))):
        tau = float(rng.uniform(5e-5, 2e-4))
        flip = float(rng.uniform(10.0, 180.0))
        b1 = flip / 360.0 / (GAMMA * tau) * np.exp(1j * rng.uniform(0.0, 2 * np.pi))
        rf_adc = np.sort(rng.uniform(0.1, 1.0, 2)) * tau if rng.random() < 0.3 else None
        blocks.append(RFBlock(tau, complex(b1), df=float(rng.normal(0.0, 50.0)),
                              psi=float(rng.uniform(0.0, 1.0)), adc_times=rf_adc))
        dur = float(rng.uniform(0.5e-3, 2e-3))
        n_adc = int(rng.integers(0, 701))
        adc = np.sort(rng.choice(np.arange(1, 10001), size=n_adc, replace=False)) / 10000.0 * dur
        g = rng.uniform(-5e-3, 5e-3, 3)
        f = float(rng.uniform(200.0, 2000.0))

        def grad(rel, g=g, f=f):
            return np.vstack([g[0] * np.sin(2 * np.pi * f * rel), g[1] * np.cos(2 * np.pi * f * rel),
                              np.full(rel.size, g[2])])
        blocks.append(FreeBlock(dur, adc, grad))
    amp = np.array([1e-6, 0, 0, 0] + [1e-4] * 5 + [1e-3] * 7)
    fq = rng.uniform(10.0, 500.0, (16, 4))
    ph = rng.uniform(0.0, 2 * np.pi, (16, 4))

    def fields(t):
        return amp[:, None] * 0.25 * np.sin(2 * np.pi * fq[:, :, None] * t[None, None, :]
                                             + ph[:, :, None]).sum(axis=1)
    seq = build_sequence(blocks, fields=fields, dt_max=float(rng.choice([5e-5, 2e-4])))
    method = BlochHighOrder(format(int(rng.integers(0, 16)), "04b"), bool(rng.random() < 0.5),
                            bool(rng.random() < 0.5))
    L = int(rng.choice([4, 32, 10 ** 9]))
    return obj, seq, method, get_sim_ranges(seq, L, L), L


def run(precision, obj, seq, parts, method):
    max_adc = max([seq.block_adc(a, b) for a, b in zip(parts[0], parts[1])] + [1])
    with Plan(precision, len(obj), obj.n_coils, seq.n_nodes, max_adc) as plan:
        plan.upload_phantom(obj)
        plan.upload_sequence(seq)
        plan.run(parts, method)
        sig = plan.download_signal()
        mxy, mz = plan.get_state()
    return sig, mxy, mz


def strengthen(seq, seed):
    """The strong variant of a case: first-order waveforms and gradients
    scaled by one of 1, 3, 8 and the 2nd/3rd-order fields by one of 1, 10,
    50, 200 (seeded)."""
    rng = np.random.default_rng(seed + 777)
    gs, fs = float(rng.choice([1, 1, 3, 8])), float(rng.choice([1, 1, 10, 50, 200]))
    seq = copy.deepcopy(seq)
    seq.h = np.array(seq.h, copy=True)
    seq.h[1:4] *= gs
    seq.gx, seq.gy, seq.gz = seq.gx * gs, seq.gy * gs, seq.gz * gs
    seq.h[4:16] *= fs
    return seq, gs, fs


@pytest.mark.parametrize("seed,strong", [(s, False) for s in SEEDS] + [(s, True) for s in STRONG_SEEDS])
def test_random_case(seed, strong):
    require_gpu()
    obj, seq, method, parts, L = random_case(seed)
    gs = fs = 1.0
    if strong:
        seq, gs, fs = strengthen(seq, seed)
    print_case("random", seed, len(obj), obj.n_coils, seq.n_nodes, int(seq.adc.sum()), method.name, L,
               f"gradients x{gs:g} fields x{fs:g}")
    ref, rmxy, rmz = oracle.run(obj, seq, parts, method.flags, GAMMA)
    sig, mxy, mz = run("mixed", obj, seq, parts, method)
    assert sig.shape == ref.shape
    if np.linalg.norm(ref) > 0:
        assert nrmse(sig, ref) <= 1e-5, f"mixed NRMSE {nrmse(sig, ref):.3e}"
    else:
        assert not sig.any()
    assert_isapprox(mxy, rmxy, 1e-11, 1e-11, "mixed Mxy")
    assert_isapprox(mz, rmz, 1e-11, 1e-11, "mixed Mz")
    sig, mxy, mz = run("f64", obj, seq, parts, method)
    assert_signal(sig, ref, "f64", "f64 signal")
    assert_isapprox(mxy, rmxy, 1e-11, 1e-11, "f64 Mxy")
    assert_isapprox(mz, rmz, 1e-11, 1e-11, "f64 Mz")


LARGE_SEEDS = [3001, 3003, 3016, 3020]


@pytest.mark.parametrize("seed", LARGE_SEEDS)
def test_random_case_large_phantom(seed):
    """The random sequences of test_random_case on 20,000-120,000 random spins
    (many spin tiles, automatic tiles above the smallest), mixed path against
    the f64 oracle."""
    require_gpu()
    obj, seq, method, parts, L = random_case(seed)
    rng = np.random.default_rng(seed + 99)
    ns, nc = int(rng.integers(20_000, 120_000)), obj.n_coils
    big = HO_Phantom(x=rng.uniform(-0.1, 0.1, ns), y=rng.uniform(-0.1, 0.1, ns),
                     z=rng.uniform(-0.035, 0.035, ns), rho=rng.uniform(0.0, 1.0, ns),
                     T1=rng.uniform(0.5, 2.0, ns), T2=rng.uniform(0.02, 0.2, ns),
                     dw=2 * np.pi * rng.normal(0.0, 20.0, ns),
                     csm=np.asfortranarray(0.5 * (rng.normal(size=(ns, nc)) + 1j * rng.normal(size=(ns, nc)))))
    print_case("random-large", seed, ns, nc, int(seq.adc.sum()), method.name, L)
    ref, rmxy, rmz = oracle.run(big, seq, parts, method.flags, GAMMA, n_threads=16)
    sig, mxy, mz = run("mixed", big, seq, parts, method)
    assert sig.shape == ref.shape
    assert nrmse(sig, ref) <= 1e-5, f"mixed NRMSE {nrmse(sig, ref):.3e}"
    assert_isapprox(mxy, rmxy, 1e-11, 1e-11, "mixed Mxy")
    assert_isapprox(mz, rmz, 1e-11, 1e-11, "mixed Mz")
