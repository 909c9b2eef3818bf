"""Block sequence descriptions and their discretization (native ``hom_discretize``).

The reference builds sequences from KomaMRI blocks and discretizes them with
``discretize(hoseq)`` (src/datatypes/HO_DiscreteSequence.jl:93-179).  Here a
block is a duration plus piecewise-linear waveforms (``(times, values)``
relative to the block start; a repeated time is a jump; zero outside the
knots): nominal gradients, complex B1 inside an RF window, and any of the
sixteen field coefficients h0..h15, with ADC sample times.  ``discretize``
returns the ``HO_DiscreteSequence`` the simulation consumes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import runtime
from .types import HO_DiscreteSequence


@dataclass
class SeqBlock:
    duration: float
    gx: tuple | None = None
    gy: tuple | None = None
    gz: tuple | None = None
    rf: tuple | None = None             # (times, complex B1 in T)
    rf_window: tuple | None = None      # (t0, t1); default: the RF knots' span
    rf_df: float = 0.0
    rf_phase: float = 0.0
    adc: np.ndarray | None = None       # sample times in (0, duration]
    fields: dict = field(default_factory=dict)  # term j -> (times, values)

    @staticmethod
    def hard_rf(duration: float, b1: complex, df: float = 0.0, phase: float = 0.0, adc=None):
        """Constant-amplitude pulse over the whole block."""
        return SeqBlock(duration, rf=([0.0, duration], [complex(b1), complex(b1)]),
                        rf_window=(0.0, duration), rf_df=df, rf_phase=phase, adc=adc)


def _wave(spec, keep):
    w = runtime.Wave(0, None, None)
    if spec is None:
        return w
    t = np.ascontiguousarray(np.asarray(spec[0], dtype=np.float64).reshape(-1))
    a = np.ascontiguousarray(np.asarray(spec[1], dtype=np.float64).reshape(-1))
    if t.size != a.size:
        raise ValueError("A waveform needs one value per knot.")
    keep += [t, a]
    return runtime.Wave(t.size, t.ctypes.data_as(C.POINTER(C.c_double)),
                        a.ctypes.data_as(C.POINTER(C.c_double)))


def discretize(blocks, dt_max: float = 1e-3, dt_rf: float = 5e-5,
               measured_first_order: bool = True) -> HO_DiscreteSequence:
    """Samples the blocks on one node grid (steps <= dt_max, <= dt_rf during
    RF; jumps keep both sides).  With ``measured_first_order`` the nominal
    gradients are added to h1..h3, so the simulated first-order field is the
    nominal gradient plus whatever the blocks put in h1..h3."""
    keep, descs = [], []
    for b in blocks:
        d = runtime.BlockDesc()
        d.duration = float(b.duration)
        d.gx, d.gy, d.gz = _wave(b.gx, keep), _wave(b.gy, keep), _wave(b.gz, keep)
        if b.rf is not None:
            tr = np.asarray(b.rf[0], dtype=np.float64)
            vr = np.asarray(b.rf[1], dtype=np.complex128)
            d.rf_re, d.rf_im = _wave((tr, vr.real), keep), _wave((tr, vr.imag), keep)
            w0, w1 = b.rf_window if b.rf_window is not None else (tr.min(), tr.max())
            d.rf_t0, d.rf_t1 = float(w0), float(w1)
        else:
            d.rf_re, d.rf_im = _wave(None, keep), _wave(None, keep)
            d.rf_t0 = d.rf_t1 = 0.0
        d.rf_df, d.rf_phase = float(b.rf_df), float(b.rf_phase)
        adc = np.ascontiguousarray(np.zeros(0) if b.adc is None else np.asarray(b.adc, np.float64))
        keep.append(adc)
        d.adc_n = adc.size
        d.adc_t = adc.ctypes.data_as(C.POINTER(C.c_double)) if adc.size else None
        for j in range(16):
            d.h[j] = _wave(b.fields.get(j), keep)
        descs.append(d)
    return runtime.discretize_blocks(descs, keep, dt_max, dt_rf, measured_first_order)
