"""ctypes binding of oracle/libhom_oracle.so (CPU f64 restatement).

TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.  Objects passed in are the
package's HO_Phantom / HO_DiscreteSequence (duck-typed: only attributes are
read), so this module does not import the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libhom_oracle.so"
_lib = None

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i64p = C.POINTER(C.c_int64)


class _Phantom(C.Structure):
    _fields_ = [("n_spins", C.c_int64), ("n_coils", C.c_int64),
                ("x", _dp), ("y", _dp), ("z", _dp), ("rho", _dp), ("T1", _dp), ("T2", _dp),
                ("dw", _dp), ("csm", _dp), ("ld_csm", C.c_int64)]


class _Sequence(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("t", _dp), ("dt", _dp), ("gx", _dp), ("gy", _dp),
                ("gz", _dp), ("b1", _dp), ("df", _dp), ("psi", _dp), ("adc", _u8p),
                ("excitation", _u8p), ("h", _dp * 16)]


def build() -> None:
    """Compiles libhom_oracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def available() -> bool:
    return LIB_PATH.exists()


def _load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = C.CDLL(str(LIB_PATH))
        P, S = C.POINTER(_Phantom), C.POINTER(_Sequence)
        lib.hom_oracle_block_adc.restype = C.c_int64
        lib.hom_oracle_block_adc.argtypes = [S, C.c_int64, C.c_int64]
        for fn in (lib.hom_oracle_precession, lib.hom_oracle_excitation):
            fn.restype = C.c_int
            fn.argtypes = [P, S, C.c_int64, C.c_int64, C.c_int, C.c_double, C.c_int64, C.c_int64,
                           _dp, _dp, _dp, C.c_int64]
        lib.hom_oracle_run.restype = C.c_int
        lib.hom_oracle_run.argtypes = [P, S, _i64p, _i64p, _u8p, C.c_int64, C.c_int, C.c_double,
                                       _dp, _dp, _dp, C.c_int64, C.c_int]
        lib.hom_oracle_combine.restype = None
        lib.hom_oracle_combine.argtypes = [_dp, C.c_int64, C.c_int64, C.c_int64, _dp, C.c_int64,
                                           C.c_int64, _dp, C.c_int64]
        lib.hom_oracle_partition.restype = C.c_int64
        lib.hom_oracle_partition.argtypes = [_u8p, C.c_int64, C.c_int64, C.c_int64, _i64p, _i64p,
                                             _u8p, C.c_int64]
        _lib = lib
    return _lib


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


class _Bound:
    """Keeps the numpy buffers alive for as long as the structs are in use."""

    def __init__(self, obj, seq):
        self.keep = []

        def f64(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self.keep.append(a)
            return _ptr(a)

        csm = np.asfortranarray(obj.csm, dtype=np.complex128)
        self.keep.append(csm)
        self.ph = _Phantom(len(obj.rho), csm.shape[1], f64(obj.x), f64(obj.y), f64(obj.z),
                           f64(obj.rho), f64(obj.T1), f64(obj.T2), f64(obj.dw),
                           csm.ctypes.data_as(_dp), csm.shape[0])
        h = np.ascontiguousarray(seq.h, dtype=np.float64)
        self.keep.append(h)
        hp = (_dp * 16)(*[C.cast(h.ctypes.data + j * h.strides[0], _dp) for j in range(16)])
        adc = np.ascontiguousarray(seq.adc, dtype=np.uint8)
        exc = np.ascontiguousarray(seq.excitation, dtype=np.uint8)
        b1 = np.ascontiguousarray(seq.b1, dtype=np.complex128)
        self.keep += [adc, exc, b1]
        self.seq = _Sequence(seq.t.size, f64(seq.t), f64(seq.dt), f64(seq.gx), f64(seq.gy),
                             f64(seq.gz), b1.ctypes.data_as(_dp), f64(seq.df), f64(seq.psi),
                             _ptr(adc, _u8p), _ptr(exc, _u8p), hp)


def _state(obj, mxy, mz):
    n = len(obj.rho)
    mxy = np.zeros(n, np.complex128) if mxy is None else np.array(mxy, dtype=np.complex128)
    mz = np.array(obj.rho, dtype=np.float64) if mz is None else np.array(mz, dtype=np.float64)
    return mxy, mz


def block_adc(seq, first, last):
    return int(np.count_nonzero(seq.adc[first + 1:last + 1]))


def _block(kind, obj, seq, first, last, flags, gamma, mxy, mz):
    lib = _load()
    b = _Bound(obj, seq)
    mxy, mz = _state(obj, mxy, mz)
    nadc = block_adc(seq, first, last)
    sig = np.zeros((nadc, obj.csm.shape[1]), dtype=np.complex128, order="F")
    fn = lib.hom_oracle_precession if kind == "precession" else lib.hom_oracle_excitation
    rc = fn(C.byref(b.ph), C.byref(b.seq), first, last, flags, gamma, 0, len(obj.rho),
            mxy.ctypes.data_as(_dp), mz.ctypes.data_as(_dp),
            sig.ctypes.data_as(_dp) if nadc else None, max(nadc, 1))
    if rc:
        raise RuntimeError(f"oracle {kind} failed with status {rc}")
    return sig, mxy, mz


def precession(obj, seq, first, last, flags, gamma, mxy=None, mz=None):
    """One free-precession block; returns (sig (Nadc x Nc), mxy, mz)."""
    return _block("precession", obj, seq, first, last, flags, gamma, mxy, mz)


def excitation(obj, seq, first, last, flags, gamma, mxy=None, mz=None):
    """One RF block; returns (sig (Nadc x Nc), mxy, mz)."""
    return _block("excitation", obj, seq, first, last, flags, gamma, mxy, mz)


def partition(seq, max_len, max_rf_len=None):
    lib = _load()
    exc = np.ascontiguousarray(seq.excitation, dtype=np.uint8)
    max_rf_len = max_len if max_rf_len is None else max_rf_len
    n = lib.hom_oracle_partition(_ptr(exc, _u8p), exc.size, max_len, max_rf_len,
                                 None, None, None, 0)
    pf, pl = np.zeros(n, np.int64), np.zeros(n, np.int64)
    rf = np.zeros(n, np.uint8)
    lib.hom_oracle_partition(_ptr(exc, _u8p), exc.size, max_len, max_rf_len,
                             _ptr(pf, _i64p), _ptr(pl, _i64p), _ptr(rf, _u8p), n)
    return pf, pl, rf


def run(obj, seq, parts, flags, gamma, mxy=None, mz=None, n_threads=1):
    """Whole block loop; returns (sig (SigmaADC x Nc), mxy, mz).  ``parts`` is
    (first, last, is_rf) arrays of 0-based inclusive node ranges."""
    lib = _load()
    b = _Bound(obj, seq)
    mxy, mz = _state(obj, mxy, mz)
    pf = np.ascontiguousarray(parts[0], dtype=np.int64)
    pl = np.ascontiguousarray(parts[1], dtype=np.int64)
    rf = np.ascontiguousarray(parts[2], dtype=np.uint8)
    total = sum(block_adc(seq, int(a), int(z)) for a, z in zip(pf, pl))
    sig = np.zeros((total, obj.csm.shape[1]), dtype=np.complex128, order="F")
    rc = lib.hom_oracle_run(C.byref(b.ph), C.byref(b.seq), _ptr(pf, _i64p), _ptr(pl, _i64p),
                            _ptr(rf, _u8p), pf.size, flags, gamma, mxy.ctypes.data_as(_dp),
                            mz.ctypes.data_as(_dp), sig.ctypes.data_as(_dp), max(total, 1),
                            int(n_threads))
    if rc:
        raise RuntimeError(f"oracle run failed with status {rc}")
    return sig, mxy, mz


def run_workload(obj, seq, method_flags, gamma, max_block_length=None, max_rf_block_length=None,
                 n_threads=1):
    """Initial state Mxy = 0, Mz = rho (BlochHighOrderSimulatitonMethod.jl:9-15),
    then the whole block loop."""
    n = len(seq.excitation)
    mb = n if max_block_length is None else max_block_length
    mr = mb if max_rf_block_length is None else max_rf_block_length
    parts = partition(seq, max(int(mb), 1), max(int(mr), 1))
    return run(obj, seq, parts, method_flags, gamma, n_threads=n_threads)


def combine(states, csm):
    """sig = transpose(states) * csm in f64 (no conjugation)."""
    lib = _load()
    st = np.asfortranarray(states, dtype=np.complex128)
    cs = np.asfortranarray(csm, dtype=np.complex128)
    ns, nadc = st.shape
    nc = cs.shape[1]
    sig = np.zeros((nadc, nc), dtype=np.complex128, order="F")
    lib.hom_oracle_combine(st.ctypes.data_as(_dp), max(ns, 1), ns, nadc, cs.ctypes.data_as(_dp),
                           max(cs.shape[0], 1), nc, sig.ctypes.data_as(_dp), max(nadc, 1))
    return sig


os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")
