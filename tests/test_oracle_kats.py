"""Pins the CPU f64 oracle with closed forms, brute force and invariants.

Each check follows a reference test or a fact the mathematics fixes:
- per-channel analytic phase          test/gpu/test_gpu_kernels.jl:54-85
- zeroth-order readout phase          test/unit/test_koma_compatibility.jl:118-130
- RF ignores high-order fields        test/unit/test_koma_compatibility.jl:132-135
- 90-degree pulse, transpose CSM      test/unit/test_koma_compatibility.jl:55-70,97
- ideal coil and CSM linearity        test/unit/test_phantom_construction.jl:123-150
- block-length invariance             test/gpu/test_gpu_compatibility.jl:183-187
- mask semantics over all 16 masks    closed form of src/simulation/kernels/HighOrderGPU.jl:73-93
- spinor update                       brute force: Rodrigues rotation of the 3-vector,
                                      with df, dw (both Delta-w_excitation values), Gy, Gz
- trapezoid exactness, T1/T2, frame phase psi   closed forms
"""
import numpy as np
import pytest

import hom_amd
import oracle
from hom_amd.types import GAMMA, HO_DiscreteSequence, HO_Phantom, BlochHighOrder
from hom_amd.workloads import FreeBlock, RFBlock, build_sequence, high_order_workload

TWO_PI = 2 * np.pi


def basis(x, y, z):
    """The reference test's own basis expression (test_gpu_kernels.jl:72-76)."""
    r2 = x ** 2 + y ** 2 + z ** 2
    return np.array([np.ones_like(x), x, y, z, x * y, z * y, 3 * z ** 2 - r2, x * z, x ** 2 - y ** 2,
                     3 * y * x ** 2 - y ** 3, x * z * y, (5 * z ** 2 - r2) * y, 5 * z ** 3 - 3 * z * r2,
                     (5 * z ** 2 - r2) * x, x ** 2 * z - y ** 2 * z, x ** 3 - 3 * x * y ** 2])


def compatibility(rho=(1.0, 0.7), x=(0.0, 0.01), csm=None, h_readout=None, h_rf=None,
                  n_adc=5, tau=1e-4):
    """RF(1/(4 gamma tau), tau) then a 2 ms block acquiring n_adc samples over
    1 ms (test/support/compatibility_workload.jl:4-17), sampled by this
    package's own sampler."""
    if csm is None:
        csm = np.array([[1, 1j], [0.5 - 0.2j, -0.3 + 0.1j]])
    obj = HO_Phantom(x=np.array(x), rho=np.array(rho), T1=np.full(len(x), np.inf),
                     T2=np.full(len(x), np.inf), csm=csm)
    times = 1e-3 * np.arange(1, n_adc + 1) / n_adc

    def fields(t):
        out = np.zeros((16, t.size))
        for j, v in (h_readout or {}).items():
            out[j, t >= tau] = v
        for j, v in (h_rf or {}).items():
            out[j, t < tau] = v
        return out

    seq = build_sequence([RFBlock(tau, 1 / (4 * GAMMA * tau)), FreeBlock(2e-3, times)], fields)
    return obj, seq, times


def run(obj, seq, mask="1111", dw_exc=True, dw_prec=True, max_len=None, threads=1):
    m = BlochHighOrder(mask, dw_exc, dw_prec)
    return oracle.run_workload(obj, seq, m.flags, GAMMA, max_block_length=max_len,
                               n_threads=threads)


def test_per_channel_analytic_phase():
    obj, seq, mxy0, mz0 = high_order_workload(nspins=3, nsteps=4, ncoils=3)
    obj.T1[:] = np.inf
    obj.T2[:] = np.inf
    obj.dw[:] = 0
    b = basis(obj.x, obj.y, obj.z)
    for term in range(16):
        seq.h[:] = 0
        seq.h[term] = 2e-5
        sig, mxy, _ = oracle.precession(obj, seq, 0, seq.n_nodes - 1, BlochHighOrder("1111").flags,
                                        GAMMA, mxy0, mz0)
        phase = -TWO_PI * GAMMA * np.outer(b[term], 2e-5 * seq.t[1:])
        states = mxy0[:, None] * np.exp(1j * phase)
        np.testing.assert_allclose(sig, states.T @ obj.csm, rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(mxy, states[:, -1], rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("mask", [format(m, "04b") for m in range(16)])
@pytest.mark.parametrize("dw_prec", [False, True])
def test_mask_semantics_closed_form(mask, dw_prec):
    """Constant coefficients: phase = -2 pi gamma t (sum of enabled terms, with the
    nominal gradient in place of h1..h3 when order one is off) - dw t."""
    obj, seq, mxy0, mz0 = high_order_workload(nspins=7, nsteps=6, ncoils=2)
    obj.T1[:] = np.inf
    obj.T2[:] = np.inf
    c = 1e-6 * (1 + np.arange(16)) * np.array([1 if j % 2 else -1 for j in range(16)])
    seq.h[:] = c[:, None]
    g = np.array([3e-5, -2e-5, 1e-5])
    seq.gx[:], seq.gy[:], seq.gz[:] = g
    m = BlochHighOrder(mask, True, dw_prec)
    b = basis(obj.x, obj.y, obj.z)
    enabled = np.zeros(16, bool)
    enabled[0] = m.ho0
    enabled[1:4] = m.ho1
    enabled[4:9] = m.ho2
    enabled[9:16] = m.ho3
    bz = (c[enabled, None] * b[enabled]).sum(axis=0)
    if not m.ho1:
        bz = bz + g[0] * obj.x + g[1] * obj.y + g[2] * obj.z
    rate = -TWO_PI * GAMMA * bz - (obj.dw if dw_prec else 0.0)
    sig, mxy, mz = oracle.precession(obj, seq, 0, seq.n_nodes - 1, m.flags, GAMMA, mxy0, mz0)
    states = mxy0[:, None] * np.exp(1j * np.outer(rate, seq.t[1:]))
    np.testing.assert_allclose(sig, states.T @ obj.csm, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(mxy, states[:, -1], rtol=1e-11, atol=1e-12)
    np.testing.assert_array_equal(mz, mz0)   # T1 = Inf: no recovery


def test_trapezoid_exact_on_linear_ramp():
    """The trapezoid rule integrates a linear ramp exactly, on any node grid."""
    t = np.concatenate([[0.0], np.sort(np.random.default_rng(1).uniform(0, 1e-3, 9)), [1.2e-3]])
    n = t.size
    a = 40.0  # T/m/s
    h = np.zeros((16, n))
    h[1] = a * t
    seq = HO_DiscreteSequence(t=t, gx=np.zeros(n), gy=np.zeros(n), gz=np.zeros(n),
                              b1=np.zeros(n), df=np.zeros(n), psi=np.zeros(n),
                              adc=np.r_[False, np.ones(n - 1, bool)],
                              excitation=np.zeros(n - 1, bool), h=h)
    obj = HO_Phantom(x=np.array([-0.03, 0.02, 0.05]), T1=np.full(3, np.inf), T2=np.full(3, np.inf))
    mxy0 = np.ones(3, complex)
    sig, _, _ = oracle.precession(obj, seq, 0, n - 1, BlochHighOrder("0100").flags, GAMMA,
                                  mxy0, obj.rho.copy())
    phase = -TWO_PI * GAMMA * np.outer(obj.x, a * t[1:] ** 2 / 2)
    np.testing.assert_allclose(sig[:, 0], np.exp(1j * phase).sum(axis=0), rtol=1e-11, atol=1e-11)


def test_relaxation_closed_forms():
    obj, seq, mxy0, mz0 = high_order_workload(nspins=5, nsteps=8, ncoils=1)
    seq.h[:] = 0
    seq.gx[:] = seq.gy[:] = seq.gz[:] = 0
    sig, mxy, mz = oracle.precession(obj, seq, 0, seq.n_nodes - 1, BlochHighOrder("0000", True, False).flags,
                                     GAMMA, mxy0, mz0)
    T = seq.t[-1]
    np.testing.assert_allclose(mxy, mxy0 * np.exp(-T / obj.T2), rtol=1e-13)
    e1 = np.exp(-T / obj.T1)
    np.testing.assert_allclose(mz, mz0 * e1 + obj.rho * (1 - e1), rtol=1e-13)
    states = mxy0[:, None] * np.exp(-np.outer(1 / obj.T2, seq.t[1:]))
    np.testing.assert_allclose(sig, states.T @ obj.csm, rtol=1e-12)


def test_ninety_degree_pulse_transpose_csm():
    obj, seq, _ = compatibility()
    sig, mxy, _ = run(obj, seq, "0000")
    expected = 1j * (obj.rho @ obj.csm)
    assert sig.shape == (5, 2)
    np.testing.assert_allclose(sig, np.tile(expected, (5, 1)), atol=1e-12)
    np.testing.assert_allclose(mxy, 1j * obj.rho, atol=1e-12)
    # A conjugating combine would give conj(csm) here.
    assert not np.allclose(sig[0], 1j * (obj.rho @ obj.csm.conj()), atol=1e-6)


def test_zeroth_order_readout_phase_and_rf_ignores_fields():
    obj, seq, times = compatibility(h_readout={0: 20 / GAMMA})
    sig, _, _ = run(obj, seq, "1000")
    expected = 1j * (obj.rho @ obj.csm)
    np.testing.assert_allclose(sig, np.exp(-1j * TWO_PI * 20 * times)[:, None] * expected,
                               atol=1e-12)
    obj2, seq2, _ = compatibility(h_readout={0: 20 / GAMMA}, h_rf={0: 30 / GAMMA, 5: 1e-2, 12: 3.0})
    sig2, _, _ = run(obj2, seq2, "1111")
    sig1, _, _ = run(obj, seq, "1111")
    np.testing.assert_allclose(sig2, sig1, atol=1e-12)


def test_ideal_coil_and_csm_linearity():
    obj, seq, _ = compatibility(csm=np.ones((2, 1)))
    sig, _, _ = run(obj, seq, "1111")
    np.testing.assert_allclose(np.abs(sig[:, 0]), obj.rho.sum(), rtol=20 * np.finfo(float).eps)
    w = np.array([[0, 2 + 1j], [0, 2 + 1j]])
    objw, seqw, _ = compatibility(csm=w)
    sigw, _, _ = run(objw, seqw, "1111")
    assert np.all(sigw[:, 0] == 0)
    np.testing.assert_allclose(sigw[:, 1], (2 + 1j) * sig[:, 0], rtol=20 * np.finfo(float).eps)


def test_block_length_invariance():
    obj, seq, mxy0, mz0 = high_order_workload(nspins=33, nsteps=32, ncoils=3)
    ref, mxy_r, mz_r = run(obj, seq, "1111", max_len=1)
    for L in (2, 5, 32, None):
        sig, mxy, mz = run(obj, seq, "1111", max_len=L)
        np.testing.assert_allclose(sig, ref, rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(mxy, mxy_r, rtol=1e-11, atol=1e-12)
        np.testing.assert_allclose(mz, mz_r, rtol=1e-12)


def test_threads_partition_sum():
    obj, seq, _, _ = high_order_workload(nspins=101, nsteps=16, ncoils=4)
    a, ma, za = run(obj, seq, "1111", threads=1)
    b, mb, zb = run(obj, seq, "1111", threads=4)
    np.testing.assert_allclose(b, a, rtol=1e-13, atol=1e-13)
    np.testing.assert_array_equal(ma, mb)
    np.testing.assert_array_equal(za, zb)


def rodrigues(v, n, ang):
    return v * np.cos(ang) + np.cross(n, v) * np.sin(ang) + n * (n @ v) * (1 - np.cos(ang))


def test_spinor_matches_rodrigues_rotation():
    """One RF interval, no relaxation: the spinor update equals a right-handed
    rotation of (Mx, My, Mz) about (Re B1, Im B1, Bz)/|B| by -2 pi gamma |B| dt."""
    rng = np.random.default_rng(7)
    for _ in range(20):
        b1 = complex(*rng.normal(0, 5e-6, 2))
        gx = rng.normal(0, 1e-3)
        x = rng.uniform(-0.05, 0.05)
        dt = rng.uniform(1e-6, 2e-4)
        m0 = rng.normal(size=3)
        t = np.array([0.0, dt])
        seq = HO_DiscreteSequence(t=t, gx=[gx, 0], gy=[0, 0], gz=[0, 0], b1=[b1, 0], df=[0, 0],
                                  psi=[0, 0], adc=[False, True], excitation=[True],
                                  h=np.zeros((16, 2)))
        obj = HO_Phantom(x=[x], T1=[np.inf], T2=[np.inf])
        sig, mxy, mz = oracle.excitation(obj, seq, 0, 1, BlochHighOrder("1111").flags, GAMMA,
                                         np.array([m0[0] + 1j * m0[1]]), np.array([m0[2]]))
        bvec = np.array([b1.real, b1.imag, gx * x])
        B = np.linalg.norm(bvec)
        v = rodrigues(m0, bvec / B, -TWO_PI * GAMMA * B * dt)
        np.testing.assert_allclose([mxy[0].real, mxy[0].imag, mz[0]], v, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(sig[0, 0], v[0] + 1j * v[1], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dw_exc", [True, False])
def test_spinor_off_resonance_terms_match_rodrigues(dw_exc):
    """One RF interval with every term of the rotation axis nonzero: the axis is
    (Re B1, Im B1, Gx x + Gy y + Gz z + [dw_exc] dw/(2 pi gamma) - df/gamma)
    (BlochHighOrderSimulatitonMethod.jl:116-117, HighOrderGPU.jl:203-208).
    The off-resonance terms are sized like B1, so a flipped sign, a missing
    2 pi or an ignored Delta-w_excitation flag moves the axis by O(1)."""
    rng = np.random.default_rng(11 + dw_exc)
    flags = BlochHighOrder("1111", dw_exc, True).flags
    for _ in range(20):
        b1 = complex(*rng.normal(0, 5e-6, 2))
        g = rng.normal(0, 2e-4, 3)
        r = rng.uniform(-0.05, 0.05, 3)
        df = rng.normal(0, 200.0)                   # Hz: df / gamma ~ 5e-6 T
        dw = TWO_PI * rng.normal(0, 200.0)          # rad/s: dw / (2 pi gamma) ~ 5e-6 T
        dt = rng.uniform(1e-6, 1e-4)
        m0 = rng.normal(size=3)
        t = np.array([0.0, dt])
        seq = HO_DiscreteSequence(t=t, gx=[g[0], 0], gy=[g[1], 0], gz=[g[2], 0], b1=[b1, 0],
                                  df=[df, 0], psi=[0, 0], adc=[False, True], excitation=[True],
                                  h=rng.normal(0, 1e-3, (16, 2)))   # ignored during RF
        obj = HO_Phantom(x=[r[0]], y=[r[1]], z=[r[2]], dw=[dw], T1=[np.inf], T2=[np.inf])
        sig, mxy, mz = oracle.excitation(obj, seq, 0, 1, flags, GAMMA,
                                         np.array([m0[0] + 1j * m0[1]]), np.array([m0[2]]))
        bz = g @ r + (dw / (TWO_PI * GAMMA) if dw_exc else 0.0) - df / GAMMA
        bvec = np.array([b1.real, b1.imag, bz])
        B = np.linalg.norm(bvec)
        v = rodrigues(m0, bvec / B, -TWO_PI * GAMMA * B * dt)
        np.testing.assert_allclose([mxy[0].real, mxy[0].imag, mz[0]], v, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(sig[0, 0], v[0] + 1j * v[1], rtol=1e-12, atol=1e-12)
        # the same check fails for an axis with either term's sign flipped
        for wrong in (g @ r + (dw / (TWO_PI * GAMMA) if dw_exc else 0.0) + df / GAMMA,
                      g @ r - (dw / (TWO_PI * GAMMA) if dw_exc else -dw / (TWO_PI * GAMMA)) - df / GAMMA):
            bw = np.array([b1.real, b1.imag, wrong])
            vw = rodrigues(m0, bw / np.linalg.norm(bw), -TWO_PI * GAMMA * np.linalg.norm(bw) * dt)
            assert np.max(np.abs(vw - v)) > 1e-9


def test_excitation_frame_phase_and_relaxation():
    """B1 = 0, Bz = 0: only the frame phase psi and relaxation act."""
    n = 5
    t = np.linspace(0, 2e-4, n)
    psi = np.array([0.3, -0.1, 0.7, 1.1, -0.4])
    seq = HO_DiscreteSequence(t=t, gx=np.zeros(n), gy=np.zeros(n), gz=np.zeros(n),
                              b1=np.zeros(n), df=np.zeros(n), psi=psi,
                              adc=[False, True, False, True, True], excitation=np.ones(n - 1, bool),
                              h=np.zeros((16, n)))
    obj = HO_Phantom(x=[0.0, 0.01], rho=[1.0, 0.5], T1=[0.8, 1.2], T2=[0.05, 0.07])
    m0 = np.array([0.2 + 0.4j, -0.3 + 0.1j])
    z0 = np.array([0.6, 0.2])
    sig, mxy, mz = oracle.excitation(obj, seq, 0, n - 1, BlochHighOrder("1111", False, True).flags,
                                     GAMMA, m0, z0)
    for k, node in enumerate([1, 3, 4]):
        st = m0 * np.exp(-1j * psi[0]) * np.exp(-t[node] / obj.T2) * np.exp(1j * psi[node])
        np.testing.assert_allclose(sig[k, 0], st.sum(), rtol=1e-12)
    np.testing.assert_allclose(mxy, m0 * np.exp(1j * (psi[-1] - psi[0])) * np.exp(-t[-1] / obj.T2),
                               rtol=1e-12)
    e1 = np.exp(-t[-1] / obj.T1)
    np.testing.assert_allclose(mz, z0 * e1 + obj.rho * (1 - e1), rtol=1e-12)


def test_combine_transpose_no_conjugate():
    rng = np.random.default_rng(3)
    st = rng.normal(size=(37, 5)) + 1j * rng.normal(size=(37, 5))
    cs = rng.normal(size=(37, 4)) + 1j * rng.normal(size=(37, 4))
    np.testing.assert_allclose(oracle.combine(st, cs), st.T @ cs, rtol=1e-12, atol=1e-12)


def test_partition_matches_host_mirror():
    _, seq, _, _ = high_order_workload(nspins=3, nsteps=32, ncoils=1)
    exc = np.zeros(32, bool)
    exc[3:9] = True
    exc[20:21] = True
    seq.excitation = exc
    for mb, mr in ((1, 1), (4, 2), (32, 32), (7, 3)):
        a = oracle.partition(seq, mb, mr)
        b = hom_amd.get_sim_ranges(seq, mb, mr)
        for u, v in zip(a, b):
            np.testing.assert_array_equal(u, v)
        # every interval covered once, adjacent parts share a node
        assert a[0][0] == 0 and a[1][-1] == 32
        np.testing.assert_array_equal(a[0][1:], a[1][:-1])
