"""Encoding trajectories: ``get_kspace`` on a discretized sequence and the
``grad2traj`` / ``traj2grad`` helpers.

``get_kspace`` mirrors reference src/datatypes/HO_Sequence.jl:221-287 and runs
in the native library (``hom_kspace``); it takes the discretized sequence (the
reference discretizes an ``HO_Sequence`` first with KomaMRIBase, which is
outside this package).  ``grad2traj`` / ``traj2grad`` mirror
src/utils/conversion.jl:1-89.
"""
from __future__ import annotations

import math

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


def _check(x, dt, dim):
    if not (isinstance(dt, (int, float, np.floating, np.integer)) and math.isfinite(dt) and dt > 0):
        raise ValueError("The sample interval must be finite and positive.")
    if not 0 <= dim < x.ndim:
        raise ValueError("dim must identify an input array dimension.")


def grad2traj(x, dt, dim: int = 0):
    """Trapezoidal cumulative integral along ``dim`` (0-based), same size as
    ``x``, starting at zero; no gamma or unit conversion."""
    x = np.asarray(x)
    _check(x, dt, dim)
    if x.shape[dim] == 0:
        return np.zeros(x.shape, dtype=np.result_type(x.dtype, np.float64))
    a = np.moveaxis(x, dim, 0)
    cs = np.cumsum((a[:-1] + a[1:]) * (dt / 2), axis=0)
    out = np.concatenate([np.zeros((1,) + a.shape[1:], dtype=cs.dtype), cs], axis=0)
    return np.moveaxis(out, 0, dim)


def traj2grad(x, dt, dim: int = 0):
    """Consecutive differences along ``dim`` divided by ``dt`` (one fewer
    sample); not an exact inverse of ``grad2traj``."""
    x = np.asarray(x)
    _check(x, dt, dim)
    a = np.moveaxis(x, dim, 0)
    return np.moveaxis((a[1:] - a[:-1]) / dt, 0, dim)
