"""Host<->device copies through the plan's pinned chunk buffers.

Copies of 4 MiB and more are split into 8 MiB chunks and packed or unpacked
by several host threads; these tests put sizes on both sides of those limits,
unaligned host addresses and strided (pitched) layouts through every entry
point that moves a large array, and check the bytes exactly."""
import ctypes as C

import numpy as np
import pytest

from hom_amd.runtime import Plan
from hom_amd.types import HO_Phantom, BlochHighOrder, get_sim_ranges
from hom_amd.workloads import high_order_workload
from helpers import print_case, require_gpu

pytestmark = pytest.mark.gpu
MIB = 1 << 20
_dp = C.POINTER(C.c_double)


@pytest.mark.parametrize("nbytes", [1, 4 * MIB - 1, 4 * MIB, 8 * MIB + 5, 3 * 8 * MIB + 12345, 5 * 8 * MIB])
def test_buffer_round_trip(nbytes):
    require_gpu()
    print_case("buffer round trip", nbytes)
    rng = np.random.default_rng(nbytes % 9973)
    store = rng.integers(0, 256, nbytes + 3, dtype=np.uint8)
    host = store[3:]  # an odd host address
    with Plan("f64", 1, 1, 2, 1) as plan:
        buf = plan.buffer(nbytes + 64)
        buf.upload(host)
        back = np.zeros(nbytes + 5, np.uint8)
        buf.download(back[5:])
        assert np.array_equal(back[5:], host)
        assert not back[:5].any()


def test_pitched_csm_upload():
    """A CSM with column stride ld > Ns (padding rows hold NaN, which must
    never reach the device) checked through the coil reduction against numpy."""
    require_gpu()
    ns, nc, nadc = 40007, 9, 6
    print_case("pitched csm", ns, nc, nadc)
    rng = np.random.default_rng(5)
    cs = rng.normal(size=(ns, nc)) + 1j * rng.normal(size=(ns, nc))
    st = rng.normal(size=(ns, nadc)) + 1j * rng.normal(size=(ns, nadc))
    ld = ns + 13
    padded = np.full((ld, nc), np.nan + 1j * np.nan, order="F")
    padded[:ns] = cs
    assert padded.nbytes >= 4 * MIB
    with Plan("f64", ns, nc, 2, nadc) as plan:
        plan.upload_phantom(HO_Phantom(x=np.zeros(ns), csm=np.asfortranarray(np.zeros((ns, nc), complex))))
        plan._check(plan.lib.hom_upload_csm(plan.handle, padded.ctypes.data_as(_dp), ld))
        sbuf = plan.buffer(st.nbytes)
        sbuf.upload(np.asfortranarray(st).ravel(order="F"))
        obuf = plan.buffer(16 * nadc * nc)
        plan.reduce_signal(sbuf, ns, nadc, obuf, nadc)
        out = obuf.download(np.zeros(nadc * nc, complex)).reshape((nadc, nc), order="F")
    ref = st.T @ cs
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-13


def test_pitched_signal_download():
    """hom_download_signal into ld > rows: the same values as a dense download,
    padding rows untouched."""
    require_gpu()
    obj, seq, mxy0, mz0 = high_order_workload(nspins=1500, nsteps=36000, ncoils=8)
    parts = get_sim_ranges(seq, 32, 32)
    nadc = max(seq.block_adc(a, b) for a, b in zip(parts[0], parts[1]))
    print_case("pitched signal", len(obj), obj.n_coils, seq.n_nodes, nadc)
    with Plan("mixed", len(obj), obj.n_coils, seq.n_nodes, nadc) as plan:
        plan.upload_phantom(obj)
        plan.upload_sequence(seq)
        plan.set_state(mxy0, mz0)
        plan.run(parts, BlochHighOrder("1111"))
        dense = plan.download_signal()
        rows = plan.signal_rows()
        assert dense.nbytes >= 4 * MIB
        ld = rows + 7
        sentinel = complex(7, -3)
        out = np.full((ld, obj.n_coils), sentinel, order="F")
        plan._check(plan.lib.hom_download_signal(plan.handle, out.ctypes.data, ld))
    assert np.all(np.isfinite(dense)) and np.abs(dense).max() > 0
    assert np.array_equal(out[:rows], dense)
    assert np.all(out[rows:] == sentinel)


@pytest.mark.parametrize("precision,ns", [("f64", 300001), ("f32", 600001)])
def test_state_round_trip(precision, ns):
    require_gpu()
    print_case("state round trip", precision, ns)
    rng = np.random.default_rng(ns)
    mxy = rng.normal(size=ns) + 1j * rng.normal(size=ns)
    mz = rng.normal(size=ns)
    if precision == "f32":
        mxy = mxy.astype(np.complex64).astype(np.complex128)
        mz = mz.astype(np.float32).astype(np.float64)
    with Plan(precision, ns, 1, 2, 1) as plan:
        plan.set_state(mxy, mz)
        a, b = plan.get_state()
    assert np.array_equal(a, mxy)
    assert np.array_equal(b, mz)
