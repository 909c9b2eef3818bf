"""Encoding trajectories: ``get_kspace`` on a discretized sequence.

``get_kspace`` mirrors reference src/datatypes/HO_Sequence.jl:221-287 and runs
in the native library (``hom_kspace``); it takes the discretized sequence (the
reference discretizes an ``HO_Sequence`` first with KomaMRIBase, which is
outside this package).
"""
from __future__ import annotations

import numpy as np

from . import runtime
from .types import GAMMA, HO_DiscreteSequence


def get_kspace(hoseqd: HO_DiscreteSequence, gamma: float = GAMMA, skip_rf=None):
    """Returns ``(K_nominal, K_nominal_adc, K_dfc, K_dfc_adc)``: the nominal
    encoding (3 columns, cycles/m) and the attached-field encoding (16
    columns, cycles/m^order), per node interval and at the ADC samples.  Each
    excitation (flip <= 90.01 degrees) restarts the integral, each refocusing
    pulse negates it, unless ``skip_rf`` (one flag per RF event) is set."""
    return runtime.kspace(hoseqd, gamma, skip_rf)


def rf_flip_angles(hoseqd: HO_DiscreteSequence, gamma: float = GAMMA) -> np.ndarray:
    """Flip angle in degrees of each RF event (run of excitation intervals)."""
    return runtime.rf_flip_angles(hoseqd, gamma)
