"""Continuous-batching simulation serving over libhom's server (include/hom_serve.h).

A :class:`SimulationServer` keeps phantoms (with their sensitivity maps) and
discretized sequences resident in HBM.  A request names a resident phantom and
sequence plus a spin range, a coil count, a node range and a
``BlochHighOrder`` method; the server admits requests in arrival order as
device memory and batch slots free up, advances every admitted request by one
chunk per step in shared persistent launches (replayed from hipGraphs), and
returns each signal as soon as that request finishes.  A result does not
depend on the other requests of the batch.

Each request is simulated with the mixed-precision path of ``hom_run`` (FP64
phase, split-FP16 MFMA coil GEMM, FP64 signal) and returns the
``(sum(ADC), Nc)`` signal matrix of ``simulate(...; return_type="mat")``
(src/simulation/Simulate.jl:150,185) before ADC phase compensation, which is
the identity for sequences with zero ADC phase.

The Poisson request stream of the serving benchmark (``request_stream``,
``run_stream``) lives here too, so tests and the benchmark drive the same code.
"""
from __future__ import annotations

import math
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import runtime
from .types import GAMMA, BlochHighOrder, DimensionMismatch, HO_DiscreteSequence, HO_Phantom


@dataclass(frozen=True)
class ResidentPhantom:
    phantom_id: int
    csm_id: int
    n_spins: int
    n_coils: int


@dataclass(frozen=True)
class ResidentSequence:
    sequence_id: int
    n_nodes: int
    adc_nodes: np.ndarray = field(repr=False)  # node index of every ADC sample

    def nodes_for_samples(self, n_samples: int) -> tuple[int, int]:
        """Node range [0, node of the n-th ADC sample] (the leading RF included)."""
        if not 1 <= n_samples <= self.adc_nodes.size:
            raise DimensionMismatch("The sequence has fewer ADC samples than requested.")
        return 0, int(self.adc_nodes[n_samples - 1])


class Request:
    """Handle of a submitted request."""

    def __init__(self, server: "SimulationServer", rid: int, n_spins: int, n_samples: int,
                 n_coils: int):
        self.server = server
        self.id = rid
        self.n_spins = n_spins
        self.n_samples = n_samples
        self.n_coils = n_coils
        self._signal = None
        self.t_submit = self.t_admit = self.t_done = None

    @property
    def spin_samples(self) -> int:
        return self.n_spins * self.n_samples

    def done(self) -> bool:
        if self._signal is not None:
            return True
        return self.server.srv.status(self.id) in ("done", "failed")

    def result(self, timeout: float | None = None) -> np.ndarray:
        """Signal matrix (samples x coils, complex128); waits for completion."""
        if self._signal is None:
            st = self.server.srv.wait(self.id, -1.0 if timeout is None else timeout)
            if st not in ("done", "failed"):
                raise TimeoutError(f"request {self.id} is still {st}")
            self.t_submit, self.t_admit, self.t_done = self.server.srv.times(self.id)
            self._signal = self.server.srv.result(self.id)
        return self._signal

    @property
    def latency(self) -> float:
        """Seconds from submission to the signal being on the host."""
        self.result()
        return self.t_done - self.t_submit


class SimulationServer:
    """Resident inputs + continuous batching on one device (threaded by default).

    ``memory_bytes`` is the device arena requests are admitted against;
    ``step_budget`` bounds the spin-samples a request advances per step
    (smaller steps interleave requests more finely); ``gpt`` and
    ``main_grid`` are the fused kernel's spin-tile and grid shapes, which
    :meth:`autotune` can choose; left unset, they come from the persisted
    tuning (tuning.json) when it has a server entry."""

    def __init__(self, device: int = 0, memory_bytes: int = 8 << 30, result_bytes: int = 1 << 30,
                 max_active: int = 64, step_budget: int | None = None, use_graphs: bool = True,
                 gpt: int | None = None, main_grid: int | None = None, threaded: bool = True):
        if gpt is None and main_grid is None:
            from . import tuning
            gpt, main_grid = tuning.server_shape() or (None, None)
        self.srv = runtime.Server(device=device, device_bytes=memory_bytes, host_bytes=result_bytes,
                                  max_active=max_active, step_budget=step_budget or 0,
                                  use_graphs=use_graphs, gpt=gpt or 0, main_grid=main_grid or 0,
                                  threaded=threaded)
        self.threaded = threaded

    def close(self):
        self.srv.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- resident store -----------------------------------------------------
    def put_phantom(self, obj: HO_Phantom) -> ResidentPhantom:
        obj.validate()
        pid = self.srv.put_phantom(obj.x, obj.y, obj.z, obj.rho, obj.T1, obj.T2, obj.dw)
        cid = self.srv.put_csm(pid, obj.csm)
        return ResidentPhantom(pid, cid, len(obj), obj.n_coils)

    def put_sequence(self, seq: HO_DiscreteSequence, gamma: float = GAMMA) -> ResidentSequence:
        qid = self.srv.put_sequence(seq, gamma)
        return ResidentSequence(qid, seq.n_nodes, np.flatnonzero(np.asarray(seq.adc, bool)))

    def put_sequence_blocks(self, blocks, gamma: float = GAMMA, dt_max: float = 1e-3,
                            dt_rf: float = 5e-5,
                            measured_first_order: bool = True) -> ResidentSequence:
        """Discretizes a block description (``sequence.discretize``, native)
        and keeps the result resident."""
        from .sequence import discretize
        return self.put_sequence(discretize(blocks, dt_max, dt_rf, measured_first_order), gamma)

    def release(self, resident):
        if isinstance(resident, ResidentPhantom):
            self.srv.release("csm", resident.csm_id)
            self.srv.release("phantom", resident.phantom_id)
        elif isinstance(resident, ResidentSequence):
            self.srv.release("sequence", resident.sequence_id)
        else:
            raise TypeError("release takes a ResidentPhantom or a ResidentSequence")

    # -- requests -------------------------------------------------------------
    def _desc(self, phantom, sequence, sim_method, spins, coils, nodes, n_samples):
        sim_method = sim_method or BlochHighOrder()
        if spins is None:
            spins = (0, phantom.n_spins)
        first, count = spins
        coils = phantom.n_coils if coils is None else coils
        if nodes is None:
            nodes = (sequence.nodes_for_samples(n_samples) if n_samples is not None
                     else (0, sequence.n_nodes - 1))
        desc = runtime.Server.request(phantom.phantom_id, first, count, phantom.csm_id, coils,
                                      sequence.sequence_id, nodes[0], nodes[1], sim_method.flags)
        n_adc = int(np.count_nonzero((sequence.adc_nodes > nodes[0]) & (sequence.adc_nodes <= nodes[1])))
        return desc, count, n_adc, coils

    def submit(self, phantom: ResidentPhantom, sequence: ResidentSequence,
               sim_method: BlochHighOrder | None = None, *, spins: tuple[int, int] | None = None,
               coils: int | None = None, nodes: tuple[int, int] | None = None,
               n_samples: int | None = None) -> Request:
        """Queues one simulation: spins (first, count) of the phantom, its first
        ``coils`` coils, and the inclusive node range ``nodes`` (or the nodes up
        to the ``n_samples``-th ADC sample) of the sequence."""
        desc, count, n_adc, coils = self._desc(phantom, sequence, sim_method, spins, coils, nodes,
                                               n_samples)
        rid = self.srv.submit(desc)
        return Request(self, rid, count, n_adc, coils)

    def footprint(self, phantom, sequence, sim_method=None, *, spins=None, coils=None, nodes=None,
                  n_samples=None) -> int:
        desc, *_ = self._desc(phantom, sequence, sim_method, spins, coils, nodes, n_samples)
        return self.srv.footprint(desc)

    def drain(self):
        self.srv.drain()

    def step(self) -> bool:
        return self.srv.step()

    def stats(self) -> dict:
        return self.srv.stats()

    def config(self) -> dict:
        return self.srv.config()

    def autotune(self, phantom, sequence, sim_method=None, *, spins=None, coils=None, nodes=None,
                 n_samples=None, gpts=None, grids=None):
        """Times the fused kernels on a probe request for each (gpt, grid) pair,
        keeps the fastest and returns (gpt, grid, seconds)."""
        desc, *_ = self._desc(phantom, sequence, sim_method, spins, coils, nodes, n_samples)
        return self.srv.autotune(desc, gpts, grids)


# ------------------------------------------------------------ request stream --

ORDERS = ("1100", "1110", "1111")
COIL_COUNTS = (1, 8, 16, 32)


@dataclass(frozen=True)
class StreamRequest:
    arrival: float      # seconds after the stream starts
    spin_first: int
    spin_count: int
    n_samples: int
    n_coils: int
    order: str


def request_stream(seed: int, n_requests: int, rate: float = 200.0, n_spins_max: int = 1 << 20,
                   spins=(1_000, 1_000_000), samples=(1_000, 20_000), coils=COIL_COUNTS,
                   orders=ORDERS) -> list[StreamRequest]:
    """Poisson arrivals (mean ``rate`` per second); spin counts log-uniform
    in ``spins``, readout lengths uniform in ``samples``, coil counts and
    field orders uniform over the given choices; each request takes a random
    contiguous spin range of a resident phantom of ``n_spins_max`` spins."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for _ in range(n_requests):
        t += rng.exponential(1.0 / rate)
        ns = int(round(math.exp(rng.uniform(math.log(spins[0]), math.log(spins[1])))))
        ns = max(1, min(ns, n_spins_max))
        first = int(rng.integers(0, n_spins_max - ns + 1))
        out.append(StreamRequest(t, first, ns, int(rng.integers(samples[0], samples[1] + 1)),
                                 int(rng.choice(coils)), str(rng.choice(orders))))
    return out


def run_stream(server: SimulationServer, phantom: ResidentPhantom, sequence: ResidentSequence,
               stream: list[StreamRequest], real_time: bool = True,
               dw_excitation: bool = True, dw_precession: bool = True) -> dict:
    """Submits the stream (at its arrival times when ``real_time``) while a
    collector thread fetches results as they finish, and returns latency
    percentiles (submission to signal on the host, from the server's clock)
    and throughput."""
    reqs = [None] * len(stream)
    ready = queue.Queue()
    errors = []

    def collect():
        while True:
            q = ready.get()
            if q is None:
                return
            try:
                q.result()
            except Exception as e:  # surfaced after the stream
                errors.append(e)

    collector = threading.Thread(target=collect, daemon=True) if server.threaded else None
    if collector:
        collector.start()
    t0 = time.perf_counter()
    for i, r in enumerate(stream):
        if real_time:
            delay = t0 + r.arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
        method = BlochHighOrder(r.order, dw_excitation=dw_excitation, dw_precession=dw_precession)
        reqs[i] = server.submit(phantom, sequence, method, spins=(r.spin_first, r.spin_count),
                                coils=r.n_coils, n_samples=r.n_samples)
        if collector:
            ready.put(reqs[i])
    if collector:
        ready.put(None)
        collector.join()
    else:
        server.drain()
    if errors:
        raise errors[0]
    lat = np.array([q.latency for q in reqs])
    t1 = max(q.t_done for q in reqs)
    ts0 = min(q.t_submit for q in reqs)
    ss = float(sum(q.spin_samples for q in reqs))
    wait = np.array([q.t_admit - q.t_submit for q in reqs])
    return {
        "requests": len(reqs),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3),
        "max_ms": float(lat.max() * 1e3),
        "admission_wait_p99_ms": float(np.percentile(wait, 99) * 1e3),
        "spin_samples": ss,
        "wall_s": float(t1 - ts0),
        "throughput": ss / float(t1 - ts0),
        "offered_rate": len(reqs) / max(stream[-1].arrival, 1e-9),
        "requests_list": reqs,
    }
