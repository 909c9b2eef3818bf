"""GPU tests of the continuous-batching server (include/hom_serve.h).

Bars: every request's signal matches the f64 oracle at NRMSE <= 1e-5 (the
mixed-precision bar) and the single-plan mixed path (hom_run) at NRMSE <= 1e-12;
a request's result is bit-identical alone and in a mix, with and without
graph replay, and whatever the admission order and chunking.
"""
import numpy as np
import pytest

import oracle
from hom_amd import workloads
from hom_amd.runtime import HomDeviceError, Plan
from hom_amd.serving import SimulationServer, request_stream, run_stream
from hom_amd.types import GAMMA, BlochHighOrder, DimensionMismatch, HO_Phantom, get_sim_ranges
from helpers import nrmse, print_case, require_gpu

pytestmark = pytest.mark.gpu

NRMSE_BAR = 1e-5
MB = 1 << 20


@pytest.fixture(scope="module")
def inputs():
    obj, seq = workloads.serving_inputs(n_spins=40_000, n_coils=40, n_samples=2_500, seed=7)
    return obj, seq


def sub_phantom(obj, first, count, coils):
    s = slice(first, first + count)
    return HO_Phantom(x=obj.x[s], y=obj.y[s], z=obj.z[s], rho=obj.rho[s], T1=obj.T1[s],
                      T2=obj.T2[s], dw=obj.dw[s], csm=np.asfortranarray(obj.csm[s, :coils]))


def oracle_signal(obj, seq, first, count, coils, last_node, method):
    o = sub_phantom(obj, first, count, coils)
    sig, _, _ = oracle.run_workload(o, seq.view(0, last_node), method.flags, GAMMA)
    return sig


def plan_signal(obj, seq, first, count, coils, last_node, method, spin_tile=256):
    """The same request through the single-plan mixed path, with the server's
    partition (maximal RF / precession runs) and spin tile (the FP32 partial
    sums run over one spin tile)."""
    o = sub_phantom(obj, first, count, coils)
    v = seq.view(0, last_node)
    exc = np.asarray(v.excitation, bool)
    pf, pl, rf = [], [], []
    a = 0
    while a < exc.size:
        b = a + 1
        while b < exc.size and exc[b] == exc[a]:
            b += 1
        pf.append(a), pl.append(b), rf.append(int(exc[a]))
        a = b
    parts = (np.array(pf), np.array(pl), np.array(rf, np.uint8))
    nadc = max([v.block_adc(x, y) for x, y in zip(pf, pl)] + [1])
    with Plan("mixed", len(o), coils, v.n_nodes, nadc) as plan:
        plan.spin_tile = spin_tile
        plan.upload_phantom(o)
        plan.upload_sequence(v)
        plan.run(parts, method)
        return plan.download_signal()


CASES = [  # (first, count, coils, n_samples, order)
    (100, 3000, 8, 1500, "1111"),
    (0, 517, 1, 2500, "1100"),
    (5000, 2000, 32, 777, "1110"),
    (39_999, 1, 16, 300, "1111"),
    (1234, 9000, 40, 2500, "1111"),
    (20_000, 4097, 3, 1024, "0110"),
]


def submit_case(srv, ph, sq, case):
    first, count, coils, ns_, order = case
    return srv.submit(ph, sq, BlochHighOrder(order), spins=(first, count), coils=coils,
                      n_samples=ns_)


def test_serve_matches_oracle_and_plan(inputs):
    require_gpu()
    obj, seq = inputs
    with SimulationServer(memory_bytes=1 << 30, result_bytes=64 * MB, threaded=False) as srv:
        ph, sq = srv.put_phantom(obj), srv.put_sequence(seq)
        for case in CASES:
            print_case("serve-oracle", *case)
            req = submit_case(srv, ph, sq, case)
            srv.drain()
            sig = req.result()
            first, count, coils, ns_, order = case
            last = int(sq.adc_nodes[ns_ - 1])
            m = BlochHighOrder(order)
            assert sig.shape == (ns_, coils)
            ref = oracle_signal(obj, seq, first, count, coils, last, m)
            e = nrmse(sig, ref)
            assert e <= NRMSE_BAR, f"{case}: NRMSE vs oracle {e:.3e}"
            e2 = nrmse(sig, plan_signal(obj, seq, first, count, coils, last, m,
                                        srv.config()["gpt"]))
            assert e2 <= 1e-12, f"{case}: NRMSE vs plan {e2:.3e}"


def run_alone(obj, seq, cases, **kw):
    out = []
    with SimulationServer(memory_bytes=1 << 30, result_bytes=64 * MB, threaded=False, **kw) as srv:
        ph, sq = srv.put_phantom(obj), srv.put_sequence(seq)
        for case in cases:
            req = submit_case(srv, ph, sq, case)
            srv.drain()
            out.append(req.result())
    return out


def test_serve_alone_equals_mix(inputs):
    """Bit-identical results alone, in one mixed batch, with small steps
    (many chunks per request), without graph replay, and in reverse order."""
    require_gpu()
    obj, seq = inputs
    print_case("serve-alone")
    alone = run_alone(obj, seq, CASES)
    for kw, order in [({}, 1), ({"step_budget": 3_000_000}, 1), ({"use_graphs": False}, -1),
                      ({"step_budget": 1_000_000, "use_graphs": False, "max_active": 3}, -1)]:
        print_case("serve-mix", kw, order)
        with SimulationServer(memory_bytes=1 << 30, result_bytes=64 * MB, threaded=False,
                              **kw) as srv:
            ph, sq = srv.put_phantom(obj), srv.put_sequence(seq)
            idx = list(range(len(CASES)))[::order]
            reqs = {i: submit_case(srv, ph, sq, CASES[i]) for i in idx}
            srv.drain()
            st = srv.stats()
            for i in idx:
                np.testing.assert_array_equal(reqs[i].result(), alone[i], err_msg=f"{kw} case {i}")
            assert st["completed"] == len(CASES)
            assert st["max_active_seen"] == min(len(CASES), kw.get("max_active", 64))
            if kw.get("use_graphs", True):
                assert st["graph_launches"] > 0 and st["graphs_captured"] >= 1


def test_serve_memory_admission(inputs):
    """An arena that holds about two requests at a time: the rest queue and
    are admitted as memory frees; results are unchanged."""
    require_gpu()
    obj, seq = inputs
    cases = [(i * 3000, 8000, 32, 2500, "1111") for i in range(6)]
    alone = run_alone(obj, seq, cases)
    with SimulationServer(memory_bytes=1 << 30, threaded=False) as probe:
        ph, sq = probe.put_phantom(obj), probe.put_sequence(seq)
        fp = probe.footprint(ph, sq, BlochHighOrder("1111"), spins=(0, 8000), coils=32,
                             n_samples=2500)
    arena = int(2.5 * fp)
    print_case("serve-admission", fp, arena)
    with SimulationServer(memory_bytes=arena, result_bytes=64 * MB, threaded=False) as srv:
        ph, sq = srv.put_phantom(obj), srv.put_sequence(seq)
        reqs = [submit_case(srv, ph, sq, c) for c in cases]
        srv.drain()
        st = srv.stats()
        assert st["max_active_seen"] == 2
        assert st["device_bytes_peak"] <= arena
        for r, a in zip(reqs, alone):
            np.testing.assert_array_equal(r.result(), a)
        assert srv.stats()["device_bytes_used"] == 0


def test_serve_threaded_stream(inputs):
    """Threaded server fed by a real-time Poisson stream: every request
    completes with its stand-alone result; latency figures are sane."""
    require_gpu()
    obj, seq = inputs
    stream = request_stream(11, 24, rate=400.0, n_spins_max=len(obj), spins=(100, 20_000),
                            samples=(200, 2_500), coils=(1, 8, 16, 32))
    cases = [(r.spin_first, r.spin_count, r.n_coils, r.n_samples, r.order) for r in stream]
    alone = run_alone(obj, seq, cases)
    print_case("serve-stream", len(stream))
    with SimulationServer(memory_bytes=1 << 30, result_bytes=128 * MB, threaded=True,
                          step_budget=20_000_000) as srv:
        ph, sq = srv.put_phantom(obj), srv.put_sequence(seq)
        out = run_stream(srv, ph, sq, stream)
        for r, a in zip(out["requests_list"], alone):
            np.testing.assert_array_equal(r.result(), a)
        assert 0 < out["p50_ms"] <= out["p99_ms"] <= out["max_ms"]
        assert out["spin_samples"] == sum(c[1] * c[3] for c in cases)
        srv.drain()


def test_serve_rf_segments():
    """Requests spanning several RF pulses (ADC during RF, idle gaps, readouts
    in between) against the f64 oracle."""
    require_gpu()
    rng = np.random.Generator(np.random.PCG64(5))
    obj, _ = workloads.serving_inputs(n_spins=3000, n_coils=8, n_samples=400, seed=3)
    fields = workloads.sinusoid_fields(rng)
    g = workloads.spiral_gradient(400, 2e-6, 0.2, 2)
    blocks = []
    for il in range(3):
        gi = g * np.exp(1j * np.pi * il / 3)

        def grad(rel, gi=gi):
            idx = np.clip(np.rint(rel / 2e-6).astype(np.int64) - 1, 0, gi.size - 1)
            inside = (rel > 0) & (rel <= gi.size * 2e-6 * (1 + 1e-12))
            v = np.where(inside, gi[idx], 0.0)
            return np.vstack([v.real, v.imag, np.zeros(rel.size)])

        rf_adc = np.array([2e-5, 6e-5]) if il == 1 else None
        blocks.append(workloads.RFBlock(1e-4, 1.0 / (8 * GAMMA * 1e-4) * (1 + 0.3j), df=30.0,
                                        psi=0.4 * il, adc_times=rf_adc))
        blocks.append(workloads.FreeBlock(4e-3, 2e-6 * np.arange(1, 401), grad))
    seq = workloads.build_sequence(blocks, fields)
    method = BlochHighOrder("1111")
    ref, _, _ = oracle.run_workload(obj, seq, method.flags, GAMMA)
    with SimulationServer(memory_bytes=1 << 30, result_bytes=64 * MB, threaded=False,
                          step_budget=500_000) as srv:
        ph, sq = srv.put_phantom(obj), srv.put_sequence(seq)
        print_case("serve-rf", seq.n_nodes, int(seq.adc.sum()))
        req = srv.submit(ph, sq, method)
        srv.drain()
        sig = req.result()
    assert sig.shape == ref.shape
    e = nrmse(sig, ref)
    assert e <= NRMSE_BAR, f"NRMSE {e:.3e}"


def test_serve_autotune(inputs):
    require_gpu()
    obj, seq = inputs
    with SimulationServer(memory_bytes=1 << 30, result_bytes=64 * MB, threaded=False) as srv:
        ph, sq = srv.put_phantom(obj), srv.put_sequence(seq)
        print_case("serve-autotune")
        gpt, grid, secs = srv.autotune(ph, sq, BlochHighOrder("1111"), spins=(0, 40_000),
                                       coils=32, n_samples=2500, gpts=[64, 256],
                                       grids=[256, 768])
        assert gpt in (64, 256) and grid in (256, 768) and secs > 0
        cfg = srv.config()
        assert cfg["gpt"] == gpt and cfg["main_grid"] == grid
        case = (100, 3000, 8, 1500, "1111")
        req = submit_case(srv, ph, sq, case)
        srv.drain()
        sig = req.result()
    ref = oracle_signal(obj, seq, 100, 3000, 8, int(sq.adc_nodes[1499]), BlochHighOrder("1111"))
    assert nrmse(sig, ref) <= NRMSE_BAR


def test_serve_errors(inputs):
    require_gpu()
    obj, seq = inputs
    with SimulationServer(memory_bytes=16 * MB, result_bytes=8 * MB, threaded=False) as srv:
        ph, sq = srv.put_phantom(obj), srv.put_sequence(seq)
        m = BlochHighOrder("1111")
        with pytest.raises(DimensionMismatch):
            srv.submit(ph, sq, m, spins=(39_990, 20))
        with pytest.raises(DimensionMismatch):
            srv.submit(ph, sq, m, spins=(0, 10), coils=41)
        with pytest.raises(DimensionMismatch):
            srv.submit(ph, sq, m, spins=(0, 10), nodes=(5, 5))
        with pytest.raises(HomDeviceError):  # larger than the whole arena
            srv.submit(ph, sq, m, spins=(0, 40_000), coils=40)
        bad = srv.srv.request(ph.phantom_id, 0, 10, 999, 1, sq.sequence_id, 0, 10, m.flags)
        with pytest.raises(ValueError):
            srv.srv.submit(bad)
        req = srv.submit(ph, sq, m, spins=(0, 100), coils=2, n_samples=100)
        assert srv.srv.status(req.id) == "queued"
        with pytest.raises(HomDeviceError):
            srv.srv.result(req.id)  # not finished
        with pytest.raises(HomDeviceError):  # its CSM is still resident
            srv.srv.release("phantom", ph.phantom_id)
        srv.drain()
        assert req.result().shape == (100, 2)
        srv.release(ph)
        with pytest.raises(ValueError):
            srv.submit(ph, sq, m, spins=(0, 10))


def test_serve_block_description_sequence():
    """A request whose sequence arrives as block descriptions (native
    discretization) against the f64 oracle on the same discretized sequence."""
    require_gpu()
    from hom_amd import SeqBlock, discretize
    obj, _ = workloads.serving_inputs(n_spins=5000, n_coils=8, n_samples=400, seed=9)
    tau, dur = 1e-4, 4e-3
    adc = np.arange(1, 1201) * 2e-6
    t_knots = np.linspace(0, dur, 41)
    blocks = [SeqBlock.hard_rf(tau, 1.0 / (4 * GAMMA * tau)),
              SeqBlock(dur, gx=([0, 2e-4, 2.2e-3, 2.4e-3], [0, 8e-3, 8e-3, 0]),
                       gy=(t_knots, 3e-3 * np.sin(2 * np.pi * t_knots / dur)), adc=adc,
                       fields={0: (t_knots, 1e-6 * np.cos(t_knots * 900)),
                               6: (t_knots, 2e-4 * np.sin(t_knots * 700)),
                               12: (t_knots, 1e-3 * np.cos(t_knots * 300))})]
    seqd = discretize(blocks)
    method = BlochHighOrder("1111")
    ref, _, _ = oracle.run_workload(obj, seqd, method.flags, GAMMA)
    with SimulationServer(memory_bytes=1 << 30, result_bytes=64 * MB, threaded=False) as srv:
        ph = srv.put_phantom(obj)
        sq = srv.put_sequence_blocks(blocks)
        print_case("serve-blocks", sq.n_nodes)
        req = srv.submit(ph, sq, method)
        srv.drain()
        sig = req.result()
    assert sig.shape == ref.shape == (1200, 8)
    e = nrmse(sig, ref)
    assert e <= NRMSE_BAR, f"NRMSE {e:.3e}"


def test_device_error_fails_queued_requests(inputs, monkeypatch):
    """A step that ends in a device error stops the server: the active request
    and the queued ones fail (queued requests used to stay queued, so a waiter
    on one never returned), and later submissions and steps report the error.
    The error is raised on the host after step 2's kernels have finished
    normally (HOM_SERVE_TEST_FAIL_STEP, read when the server is created)."""
    require_gpu()
    obj, seq = inputs
    method = BlochHighOrder("1111")
    monkeypatch.setenv("HOM_SERVE_TEST_FAIL_STEP", "2")
    srv = SimulationServer(memory_bytes=512 * MB, result_bytes=64 * MB, max_active=1,
                           step_budget=1 << 23, threaded=False)
    monkeypatch.delenv("HOM_SERVE_TEST_FAIL_STEP")
    with srv:
        ph = srv.put_phantom(sub_phantom(obj, 0, 20_000, 8))
        sq = srv.put_sequence(seq)
        reqs = [srv.submit(ph, sq, method) for _ in range(3)]
        print_case(f"injected failure at step 2, {len(reqs)} requests of 20000 spins x 8 coils")
        assert srv.step()                       # step 1: admits the first request
        assert srv.srv.status(reqs[0].id) == "active"
        assert srv.srv.status(reqs[1].id) == "queued"
        with pytest.raises(HomDeviceError, match="injected"):
            srv.step()                          # step 2 fails
        assert [srv.srv.status(r.id) for r in reqs] == ["failed"] * 3
        assert [srv.srv.wait(r.id, 5.0) for r in reqs] == ["failed"] * 3
        st = srv.stats()
        assert st["failed"] == 3 and st["queued"] == 0 and st["active"] == 0
        with pytest.raises(HomDeviceError):
            srv.submit(ph, sq, method)
        with pytest.raises(HomDeviceError):
            srv.drain()
