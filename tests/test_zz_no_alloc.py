"""Allocation-free coil reduction: hom_reduce_signal_noalloc against
hom_reduce_signal (reference CoilReductionGPU.jl:79-100 for the operation).

Both entry points run the same two kernels in the same order, so the results
must agree bit for bit; the allocation-free one keeps its stage-1 partials in
the library's per-device scratch array, which the second (smaller) call finds
holding the first call's values.
"""
import numpy as np
import pytest

from helpers import print_case, require_gpu
from hom_amd.runtime import HOM_REDUCE_SCRATCH_BYTES, Plan
from hom_amd.types import DimensionMismatch, HO_Phantom

pytestmark = pytest.mark.gpu


def _inputs(nspins, nadc, ncoils, dtype, seed):
    rng = np.random.default_rng(seed)
    st = (rng.normal(size=(nspins, nadc)) + 1j * rng.normal(size=(nspins, nadc))).astype(dtype)
    cs = (rng.normal(size=(nspins, ncoils)) + 1j * rng.normal(size=(nspins, ncoils))).astype(dtype)
    return st, cs


def _reduce_both(precision, nspins, nadc, ncoils, seed):
    dtype = np.complex64 if precision == "f32" else np.complex128
    print_case("no-alloc reduction", precision, nspins, nadc, ncoils)
    st, cs = _inputs(nspins, nadc, ncoils, dtype, seed)
    obj = HO_Phantom(x=np.zeros(nspins), csm=np.asfortranarray(cs.astype(np.complex128)))
    outs = {}
    with Plan(precision, nspins, ncoils, 2, max(nadc, 1)) as plan:
        plan.upload_phantom(obj)
        sbuf = plan.buffer(max(st.nbytes, 16))
        sbuf.upload(np.asfortranarray(st).ravel(order="F"))
        for fn in ("reduce_signal_noalloc", "reduce_signal"):
            host = np.full((nadc + 3, ncoils), complex(5, -1), dtype=dtype, order="F")
            obuf = plan.buffer(host.nbytes)
            obuf.upload(host.ravel(order="F"))
            getattr(plan, fn)(sbuf, max(nspins, 1), nadc, obuf, nadc + 3, offset_elems=1)
            flat = np.zeros(host.size, dtype=dtype)
            obuf.download(flat)
            outs[fn] = flat.reshape(host.shape, order="F")
    ref = st.astype(np.complex128).T @ cs.astype(np.complex128)
    return outs["reduce_signal_noalloc"], outs["reduce_signal"], ref


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_noalloc_matches_allocating_reduction_twice(precision):
    require_gpu()
    # a large call, then a smaller one that finds the first call's partials
    # still in the scratch array
    for nspins, nadc, ncoils, seed in ((70001, 32, 32, 1), (4097, 7, 5, 2)):
        got, want, ref = _reduce_both(precision, nspins, nadc, ncoils, seed)
        # same kernels, same summation order: identical bits, rows outside the
        # block untouched
        assert np.array_equal(np.ascontiguousarray(got).view(np.uint8),
                              np.ascontiguousarray(want).view(np.uint8))
        assert np.all(got[[0, nadc + 1, nadc + 2]] == complex(5, -1))
        t = 5e-6 if precision == "f32" else 5e-13
        scale = max(np.linalg.norm(ref), 1.0)
        assert np.linalg.norm(got[1:nadc + 1] - ref) / scale < t


def test_noalloc_rejects_partials_beyond_scratch():
    require_gpu()
    # ceil(2049 / 2048) = 2 spin tiles; 2 * 1025 * 1024 complex doubles exceed
    # the 32 MiB array, which is refused before any launch
    nspins, nadc, ncoils = 2049, 1025, 1024
    assert 2 * nadc * ncoils * 16 > HOM_REDUCE_SCRATCH_BYTES
    print_case("no-alloc reduction over the limit", nspins, nadc, ncoils)
    obj = HO_Phantom(x=np.zeros(nspins), csm=np.ones((nspins, ncoils), dtype=np.complex128, order="F"))
    with Plan("f64", nspins, ncoils, 2, nadc) as plan:
        plan.upload_phantom(obj)
        sbuf = plan.buffer(16 * nspins * nadc)
        obuf = plan.buffer(16 * nadc * ncoils)
        with pytest.raises(DimensionMismatch, match="scratch"):
            plan.reduce_signal_noalloc(sbuf, nspins, nadc, obuf, nadc)
        # within the limit on the same plan: one sample
        plan.reduce_signal_noalloc(sbuf, nspins, 1, obuf, 1)
        plan.synchronize()
