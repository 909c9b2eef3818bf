"""Shared test helpers: Julia-style isapprox, f32 input rounding, GPU gating."""
import os

import numpy as np
import pytest


def julia_isapprox(a, b, atol=0.0, rtol=0.0):
    """Julia's isapprox for arrays: norm(a - b) <= max(atol, rtol*max(norm(a), norm(b)))."""
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.linalg.norm((a - b).ravel())
    return d <= max(atol, rtol * max(np.linalg.norm(a.ravel()), np.linalg.norm(b.ravel())))


def assert_isapprox(a, b, atol=0.0, rtol=0.0, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} != {b.shape}"
    d = np.linalg.norm((a - b).ravel())
    scale = max(np.linalg.norm(a.ravel()), np.linalg.norm(b.ravel()))
    assert julia_isapprox(a, b, atol, rtol), \
        f"{what}: |a-b| = {d:.3e}, scale = {scale:.3e}, atol={atol}, rtol={rtol}"


def nrmse(a, ref):
    """||a - ref||_2 / ||ref||_2 over every entry, no fitted scale or phase."""
    ref = np.asarray(ref)
    return float(np.linalg.norm((np.asarray(a) - ref).ravel()) / np.linalg.norm(ref.ravel()))


def assert_signal(sig, ref, precision, what=""):
    """Signal against the f64 oracle: the reference's isapprox (f32 3e-5, f64
    1e-11, atol = rtol, norm-wise) and, for f64, NRMSE <= 1e-12 (the bar the
    f64 drop-in is held to on every workload)."""
    t = tol(precision)
    assert_isapprox(sig, ref, t, t, what)
    if precision == "f64" and np.linalg.norm(np.asarray(ref).ravel()) > 0:
        e = nrmse(sig, ref)
        assert e <= 1e-12, f"{what}: f64 NRMSE {e:.3e} above 1e-12"


def round_f32(obj, seq):
    """Round every input to float32 (the reference converts inputs with f32()
    before the f32 path runs), keeping float64 containers for the oracle."""
    import copy
    o = copy.deepcopy(obj)
    s = copy.deepcopy(seq)
    for nm in ("x", "y", "z", "rho", "T1", "T2", "dw"):
        setattr(o, nm, getattr(o, nm).astype(np.float32).astype(np.float64))
    o.csm = np.asfortranarray(o.csm.astype(np.complex64).astype(np.complex128))
    for nm in ("t", "dt", "gx", "gy", "gz", "df", "psi"):
        setattr(s, nm, getattr(s, nm).astype(np.float32).astype(np.float64))
    s.b1 = s.b1.astype(np.complex64).astype(np.complex128)
    s.h = s.h.astype(np.float32).astype(np.float64)
    return o, s


def tol(precision):
    """Reference tolerances (test/gpu/test_gpu_kernels.jl:31)."""
    return 3e-5 if precision == "f32" else 1e-11


def require_gpu():
    import hom_amd.runtime as rt
    if rt.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests must run on the GPU box")


def print_case(*parts):
    print("CASE", *parts, flush=True)
    os.sys.stdout.flush()
