"""``"raw"`` simulation output: an ISMRMRD-shaped acquisition record.

``finalize_simulation!`` (reference src/simulation/Simulate.jl:195-214) turns
the signal into ``RawAcquisitionData`` with KomaMRI's ``signal_to_raw_data``,
which is not part of the reference tree.  This module builds the same kind of
record from the signal and the discretized sequence: one ``Profile`` per
readout (header, nominal trajectory, samples x channels data in complex64, as
ISMRMRD stores it) plus a parameter dictionary with the simulation metadata.

Readouts: consecutive ADC samples form one readout unless an RF interval lies
between them or the gap exceeds twice the readout's current sample spacing.
Trajectories are the nominal encoding at the samples (``get_kspace``) in
cycles/m.  Both are this package's conventions (the third-party layout they
replace is not pinned).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .trajectory import get_kspace
from .types import GAMMA, HO_DiscreteSequence


@dataclass
class AcquisitionHeader:
    scan_counter: int
    number_of_samples: int
    available_channels: int
    active_channels: int
    sample_time_us: float
    acquisition_time_stamp_us: float      # time of the first sample
    trajectory_dimensions: int = 3
    center_sample: int = 0
    discard_pre: int = 0
    discard_post: int = 0
    flags: int = 0
    idx: dict = field(default_factory=dict)


@dataclass
class Profile:
    head: AcquisitionHeader
    traj: np.ndarray   # trajectory_dimensions x number_of_samples (cycles/m)
    data: np.ndarray   # number_of_samples x channels, complex64


@dataclass
class RawAcquisitionData:
    params: dict
    profiles: list


def readouts(seqd: HO_DiscreteSequence) -> list[np.ndarray]:
    """ADC sample indices (into the signal rows) of each readout."""
    nodes = np.flatnonzero(np.asarray(seqd.adc, bool))
    if nodes.size == 0:
        return []
    exc = np.asarray(seqd.excitation, bool)
    rf_before = np.concatenate([[0], np.cumsum(exc)])  # RF intervals before each node
    out, cur = [], [0]
    for i in range(1, nodes.size):
        a, b = nodes[i - 1], nodes[i]
        gap = seqd.t[b] - seqd.t[a]
        spacing = (seqd.t[nodes[cur[-1]]] - seqd.t[nodes[cur[-2]]]) if len(cur) > 1 else gap
        if rf_before[b] != rf_before[a] or gap > 2.0 * spacing:
            out.append(np.array(cur))
            cur = [i]
        else:
            cur.append(i)
    out.append(np.array(cur))
    return out


def signal_to_raw_data(sig, seqd: HO_DiscreteSequence, *, phantom_name: str = "",
                       sys=None, sim_params: dict | None = None,
                       gamma: float = GAMMA) -> RawAcquisitionData:
    """Raw record of a (ΣADC, Nc[, 1]) signal acquired with ``seqd``."""
    sig = np.asarray(sig)
    if sig.ndim == 3:
        sig = sig[:, :, 0]
    nodes = np.flatnonzero(np.asarray(seqd.adc, bool))
    if sig.shape[0] != nodes.size:
        raise ValueError("The signal must have one row per ADC sample of the sequence.")
    ncoils = sig.shape[1]
    traj = (get_kspace(seqd, gamma)[1] if seqd.n_nodes > 1 and nodes.size
            else np.zeros((nodes.size, 3)))
    profiles = []
    for k, rows in enumerate(readouts(seqd)):
        times = seqd.t[nodes[rows]]
        dwell = float(np.median(np.diff(times))) if rows.size > 1 else 0.0
        head = AcquisitionHeader(scan_counter=k, number_of_samples=int(rows.size),
                                 available_channels=ncoils, active_channels=ncoils,
                                 sample_time_us=dwell * 1e6,
                                 acquisition_time_stamp_us=float(times[0]) * 1e6,
                                 center_sample=int(rows.size // 2),
                                 idx={"kspace_encode_step_1": k, "repetition": 0})
        profiles.append(Profile(head, np.ascontiguousarray(traj[rows].T),
                                np.ascontiguousarray(sig[rows].astype(np.complex64))))
    params = {
        "systemVendor": "HighOrderMRISim MI355X BlochHighOrder",
        "institutionName": "", "patientName": phantom_name,
        "trajectory": "other", "trajectory_units": "cycles/m",
        "encodedSize": [int(nodes.size), 1, 1], "receiverChannels": ncoils,
        "TE": None, "TR": None,
        "userParameters": dict(sim_params or {}),
    }
    if sys is not None:
        params["scanner"] = repr(sys)
    return RawAcquisitionData(params, profiles)
