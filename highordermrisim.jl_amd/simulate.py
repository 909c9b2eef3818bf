"""The reference's public simulation surface over the C ABI.

- :func:`simulate`             src/simulation/Simulate.jl:52-67 (one-shot run)
- :func:`setup_simulation`     src/simulation/Simulate.jl:70-135 (validation + block plan)
- :func:`prepare_simulation`   src/simulation/PreparedSimulation.jl:144-179 (resident plan)
- :class:`PreparedSimulation`  src/simulation/PreparedSimulation.jl:2-287 (reuse, close)

``sim_params`` keeps its keys and meanings: ``"sim_method"`` (BlochHighOrder),
``"precision"`` (``"f32"``, ``"f64"``, plus ``"mixed"`` for the fused
mixed-precision path), ``"return_type"`` (``"mat"``, ``"state"``, ``"raw"``),
``"implementation"``, ``"coil_combination"``, ``"synchronization"``,
``"max_block_length"``, ``"max_rf_block_length"``, legacy ``"Nblocks"``,
``"Nthreads"``.  The sequence argument is an already sampled
``HO_DiscreteSequence``: sampling a continuous sequence is KomaMRIBase's
``discretize`` and is outside this package.

Every run executes on the GPU through libhom.so; ``implementation="array"``
(the reference's CPU path) is rejected rather than emulated.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field

import numpy as np

from .raw import signal_to_raw_data
from .runtime import COIL_MODES, Plan, device_count
from .types import GAMMA, BlochHighOrder, DimensionMismatch, HO_DiscreteSequence, HO_Phantom, \
    get_sim_ranges


@dataclass
class Mag:
    """Spin state (KomaMRICore Mag): transverse xy (complex) and longitudinal z."""
    xy: np.ndarray
    z: np.ndarray


def default_sim_params() -> dict:
    return {"sim_method": BlochHighOrder("1111"), "precision": "f32", "return_type": "raw",
            "implementation": "auto", "coil_combination": "gemm", "synchronization": "auto",
            "max_block_length": np.inf, "max_rf_block_length": np.inf, "Nthreads": 1,
            "gpu": True, "gpu_device": 0, "gamma": GAMMA}


@dataclass
class SimulationSetup:
    obj: HO_Phantom
    seqd: HO_DiscreteSequence
    params: dict
    parts: tuple
    adc_counts: np.ndarray
    max_adc: int
    max_rf_adc: int
    t_sim_parts: np.ndarray
    device: int


def simulation_spin_ranges(nspins: int, nshards: int) -> list[range]:
    """Contiguous spin shards [floor((i-1) N / k), floor(i N / k)) for i = 1..k,
    0-based (src/simulation/DistributedSimulation.jl:28-31): how a multi-GPU
    run would split the phantom."""
    if not (isinstance(nspins, (int, np.integer)) and isinstance(nshards, (int, np.integer))
            and 1 <= nshards <= nspins):
        raise ValueError("Require 1 <= nshards <= nspins.")
    return [range((i - 1) * nspins // nshards, i * nspins // nshards) for i in range(1, nshards + 1)]


def select_device(params: dict, gpus, nspins: int) -> int:
    """The device checks of set_simulation_devices!
    (src/simulation/DistributedSimulation.jl:2-26) for this one-GPU build: gpus
    is None or a list/tuple of distinct nonnegative integer IDs (not bools), at
    most one per spin, each a visible device; more than one ID is refused (one
    plan runs on one GPU).  Records gpu, gpu_device and gpus in params."""
    if gpus is not None and not isinstance(gpus, (list, tuple)):
        raise ValueError("gpus must be a vector or tuple of visible-device IDs.")
    ids = list(gpus) if gpus is not None else [params.get("gpu_device") or 0]
    if not ids or not all(isinstance(i, (int, np.integer)) and not isinstance(i, (bool, np.bool_))
                          and i >= 0 for i in ids):
        raise ValueError("gpus must contain nonnegative integer visible-device IDs.")
    if len(set(ids)) != len(ids):
        raise ValueError("gpus contains duplicate IDs.")
    if len(ids) > nspins:
        raise ValueError("Each GPU needs at least one spin.")
    if len(ids) > 1:
        raise ValueError("This build runs one simulation on one GPU; pass gpus=[id].")
    if gpus is not None and ids[0] >= device_count():
        raise ValueError("gpus contains an unavailable device ID.")
    params["gpu"] = True
    params["gpu_device"] = int(ids[0])
    params["gpus"] = [int(ids[0])]
    return int(ids[0])


def setup_simulation(obj: HO_Phantom, seqd: HO_DiscreteSequence, sim_params=None, gpus=None):
    """Validation and block plan, in the reference's order
    (src/simulation/Simulate.jl:70-135); errors are raised before any device work."""
    sim_params = {} if sim_params is None else sim_params
    if not isinstance(sim_params, dict) or not all(isinstance(k, str) for k in sim_params):
        raise ValueError("sim_params must be a dictionary with string keys.")
    params = default_sim_params()
    params.update(sim_params)
    if params["synchronization"] not in ("auto", "block"):
        raise ValueError("synchronization must be auto or block.")
    if params["implementation"] not in ("auto", "array", "kernel"):
        raise ValueError("implementation must be auto, array, or kernel.")
    if params["implementation"] == "array":
        raise ValueError("The CPU array implementation is not part of this GPU package; "
                         "use implementation=auto or kernel.")
    if params["coil_combination"] not in COIL_MODES:
        raise ValueError(f"Unknown coil_combination: {params['coil_combination']}.")
    if not isinstance(params["sim_method"], BlochHighOrder):
        raise ValueError("HO_Phantom simulation here supports BlochHighOrder().")
    if params["return_type"] not in ("mat", "state", "raw"):
        raise ValueError("return_type must be mat, state, or raw.")
    if params["precision"] not in ("f32", "f64", "mixed"):
        raise ValueError("The high-order path supports f32, f64 and mixed precision.")
    nt = params["Nthreads"]
    if not (isinstance(nt, (int, np.integer)) and nt > 0):
        raise ValueError("Nthreads must be a positive integer.")
    obj.validate()
    if len(obj) == 0:
        raise ValueError("The phantom must contain at least one spin.")
    if obj.n_coils < 1:
        raise ValueError("The CSM must contain at least one coil.")
    seqd.validate()
    device = select_device(params, gpus, len(obj))
    if "Nblocks" in sim_params:
        nb = sim_params["Nblocks"]
        if not (isinstance(nb, (int, np.integer)) and nb > 0):
            raise ValueError("Nblocks must be a positive integer.")
        legacy = max(1, -(-len(seqd) // nb))
        if "max_block_length" not in sim_params:
            params["max_block_length"] = legacy
        if "max_rf_block_length" not in sim_params:
            params["max_rf_block_length"] = legacy
    for key in ("max_block_length", "max_rf_block_length"):
        v = params[key]
        if not (v == np.inf or (isinstance(v, (int, np.integer)) and v > 0)):
            raise ValueError(f"{key} must be a positive integer or Inf.")
    parts = get_sim_ranges(seqd, params["max_block_length"], params["max_rf_block_length"])
    if len(parts[0]) == 0:
        raise ValueError("The sequence must contain a simulation interval.")
    adc_counts = np.array([seqd.block_adc(a, b) for a, b in zip(parts[0], parts[1])], np.int64)
    max_adc = int(adc_counts.max())
    rf_counts = adc_counts[parts[2].astype(bool)]
    max_rf_adc = int(rf_counts.max()) if rf_counts.size else 0
    t_sim_parts = np.append(seqd.t[parts[0]], seqd.t[-1])
    setup = SimulationSetup(obj, seqd, params, parts, adc_counts, max_adc, max_rf_adc,
                            t_sim_parts, device)
    # what simulate_shard records before the block loop, so callbacks see it
    # (Simulate.jl:168-175); params is this call's own copy
    params["implementation"] = "kernel"
    params["max_adc_per_block"] = max_adc
    params["max_rf_adc_per_block"] = max_rf_adc
    params["adc_state_buffer_samples"] = _state_buffer_samples(setup)
    return setup


def _run_shard(plan: Plan, setup: SimulationSetup, callbacks=(), initial=None):
    """simulate_shard (src/simulation/Simulate.jl:138-193) on an existing plan."""
    params = setup.params
    method = params["sim_method"]
    if initial is None:
        plan.reset_state()
    else:
        plan.set_state(initial.xy, initial.z)
    t0 = time.perf_counter()
    fence = params["synchronization"] == "block" or len(callbacks) > 0
    if not fence:
        plan.run(setup.parts, method)
        plan.synchronize()
        sig = plan.download_signal()
    else:
        # Per-block stepping with completed data for callbacks (SimulatorCore.jl:109-149).
        # Each block writes its rows (column stride = its ADC count) into a
        # device buffer of max_adc x Nc, copied into the host signal's rows.
        total = int(setup.adc_counts.sum())
        nc = setup.obj.n_coils
        itemsize = np.dtype(plan.cdtype).itemsize
        buf = plan.buffer(max(max(int(setup.max_adc), 1) * nc * itemsize, 16))
        sig = np.zeros((total, nc), dtype=plan.cdtype, order="F")
        try:
            row = 0
            rfs = 0
            nparts = len(setup.parts[0])
            for b in range(nparts):
                first, last, rf = (int(setup.parts[0][b]), int(setup.parts[1][b]), bool(setup.parts[2][b]))
                nadc = int(setup.adc_counts[b])
                fn = plan.excitation_block if rf else plan.precession_block
                fn(first, last, method, buf if nadc else None, max(nadc, 1))
                if nadc:
                    rows = np.zeros(nadc * nc, dtype=plan.cdtype)
                    buf.download(rows)
                    sig[row:row + nadc] = rows.reshape((nadc, nc), order="F")
                rfs += rf
                row += nadc
                plan.synchronize()
                if callbacks:
                    xy, z = plan.get_state()
                    state = Mag(xy.copy(), z.copy())
                    progress = {"block": b + 1, "Nblocks": nparts, "rfs": rfs, "samples": row + 1}
                    blocks = {"excitation_bool": setup.parts[2].astype(bool), "parts": setup.parts}
                    data = {"obj": setup.obj, "seqd": setup.seqd, "sig": sig, "Xt": state}
                    for cb in callbacks:
                        cb(progress, blocks, data, params)
                    if not (np.array_equal(state.xy, xy) and np.array_equal(state.z, z)):
                        plan.set_state(state.xy, state.z)
        finally:
            buf.free()
    seconds = time.perf_counter() - t0
    xy, z = plan.get_state()
    return sig, Mag(xy, z), seconds


def _finalize(sig, state, seconds, setup: SimulationSetup):
    """finalize_simulation! (src/simulation/Simulate.jl:195-214).  ADC phase
    compensation: a per-sample phase array may be given as sim_params["adc_phase"]
    (radians); without it the compensation is the identity (ADC phase 0)."""
    params = setup.params
    phase = params.get("adc_phase")
    if phase is not None:
        phase = np.asarray(phase, dtype=np.float64).reshape(-1)
        if phase.size != sig.shape[0]:
            raise DimensionMismatch("adc_phase must contain one value per ADC sample.")
        sig = sig * np.exp(-1j * phase)[:, None].astype(sig.dtype)
    if params["return_type"] == "state":
        return state
    mat = sig.reshape(sig.shape[0], sig.shape[1], 1)
    if params["return_type"] == "mat":
        return mat
    raw_params = dict(params)
    raw_params.update({"sim_method": repr(params["sim_method"]), "gpu_device": setup.device,
                       "t_sim_parts": setup.t_sim_parts, "type_sim_parts": setup.parts[2].astype(bool),
                       "Nblocks": len(setup.parts[0]), "sim_time_sec": seconds,
                       "max_adc_per_block": setup.max_adc, "max_rf_adc_per_block": setup.max_rf_adc,
                       "adc_state_buffer_samples": _state_buffer_samples(setup)})
    return signal_to_raw_data(mat, setup.seqd, phantom_name=getattr(setup.obj, "name", ""),
                              sim_params=raw_params, gamma=float(params.get("gamma", GAMMA)))


def _state_buffer_samples(setup: SimulationSetup) -> int:
    """ADC samples per spin of the plan's state buffer as created: RF-block
    ADC only when precession blocks keep their states on chip (mixed
    precision, fused coil modes), every block's ADC otherwise
    (HighOrderGPU.jl:37).  A mixed run with callbacks steps its precession
    blocks one at a time, each as its own on-chip segment, so the buffer
    keeps that size."""
    p = setup.params
    on_chip = p["precision"] == "mixed" or p["coil_combination"] in ("fused_warp", "fused_tiled")
    return setup.max_rf_adc if on_chip else setup.max_adc


def _make_plan(setup: SimulationSetup) -> Plan:
    if device_count() < 1:
        raise RuntimeError("No HIP device is available (this package has no CPU fallback).")
    p = setup.params
    plan = Plan(p["precision"], len(setup.obj), setup.obj.n_coils, setup.seqd.n_nodes,
                setup.max_adc, setup.max_rf_adc, float(p.get("gamma", GAMMA)),
                p["coil_combination"], setup.device)
    plan.upload_phantom(setup.obj)
    plan.upload_sequence(setup.seqd)
    return plan


def simulate(obj: HO_Phantom, seqd: HO_DiscreteSequence, sys=None, *, sim_params=None, gpus=None,
             verbose=False, callbacks=()):
    """Runs the BlochHighOrder simulation on one GPU.  ``sys`` (the scanner)
    is accepted for signature compatibility; nothing on this path reads it."""
    setup = setup_simulation(obj, seqd, sim_params, gpus)
    with _make_plan(setup) as plan:
        sig, state, seconds = _run_shard(plan, setup, tuple(callbacks))
    return _finalize(sig, state, seconds, setup)


class PreparedSimulation:
    """Resident plan: phantom, CSM, sequence and block plan stay on the device
    across runs; each run restarts from Mxy = 0, Mz = rho (or the state it is
    given).  Runs are serialized by a lock and a run may not start from inside
    one of its own callbacks (src/simulation/PreparedSimulation.jl:190-194,224)."""

    def __init__(self, setup: SimulationSetup):
        self.setup = setup
        setup.params["prepared"] = True  # PreparedSimulation.jl:104
        self._plan = _make_plan(setup)
        self._lock = threading.Lock()
        self._running = threading.local()
        self.closed = False

    def simulate(self, callbacks=(), initial_state: Mag | None = None):
        if self.closed:
            raise ValueError("The prepared simulation has been closed.")
        if getattr(self._running, "active", False):
            raise ValueError("A prepared simulation cannot be run from its own callback.")
        with self._lock:
            self._running.active = True
            try:
                sig, state, seconds = _run_shard(self._plan, self.setup, tuple(callbacks),
                                                 initial_state)
            finally:
                self._running.active = False
        return _finalize(sig, state, seconds, self.setup)

    def close(self):
        if not self.closed:
            with self._lock:
                self._plan.close()
                self.closed = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def prepare_simulation(obj: HO_Phantom, seqd: HO_DiscreteSequence, sys=None, *, sim_params=None,
                       gpus=None, verbose=False) -> PreparedSimulation:
    return PreparedSimulation(setup_simulation(obj, seqd, sim_params, gpus))
