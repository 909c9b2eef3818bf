"""The reference CPU algorithm in its own form: broadcast temporaries per block
and a complex GEMM for the coil combination -- TEST INFRASTRUCTURE ONLY.

Only tests/ and bench.py's ``cpu_baseline`` leg use this module.  It is the
CPU baseline that follows the reference's CPU path step by step (the C oracle
in hom_oracle.c computes the same values one spin at a time and is what the
parity tests compare against):

* spins split into contiguous parts, one per thread
  (``kfoldperm(length(obj), Nthreads)``, src/simulation/SimulatorCore.jl:23-35,
  56-70); each thread writes its own partial signal, summed in part order;
* free precession, one block at a time
  (src/simulation/Bloch/BlochHighOrderSimulatitonMethod.jl:43-94): all fifteen
  field terms as spins x nodes broadcast arrays (:54-68), selected by the
  method flags (:70-75), ``cumtrapz`` over the block's intervals (:78),
  ``Mxy = [M.xy M.xy .* exp.(1im .* phi .- tp' ./ T2)]`` (:85), the ADC
  columns times the sensitivities as one complex GEMM, transpose without
  conjugation (:17-20, :89-92), end state and T1 recovery (:86-87);
* RF blocks node by node, vectorised over spins, with the hard-pulse spinor
  (:100-134).

Motion (``ux, uy, uz``) is zero here, as on the accelerated path.
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
from threadpoolctl import threadpool_limits

_HO0, _HO1, _HO2, _HO3, _DWX, _DWP = 1, 2, 4, 8, 16, 32


def _kfoldperm(n: int, k: int) -> list[tuple[int, int]]:
    """Contiguous near-equal spin ranges, one per thread."""
    edges = np.linspace(0, n, max(1, min(k, n)) + 1).round().astype(np.int64)
    return [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:]) if b > a]


def _precession(p, seq, first, last, flags, gamma, mxy, mz):
    """One free-precession block over nodes first..last (inclusive)."""
    nodes = slice(first, last + 1)
    x, y, z = p["x"][:, None], p["y"][:, None], p["z"][:, None]
    h = [seq.h[j][nodes][None, :] for j in range(16)]
    r2 = x ** 2 + y ** 2 + z ** 2
    bzh1 = x * h[1]
    bzh2 = y * h[2]
    bzh3 = z * h[3]
    bzh4 = (x * y) * h[4]
    bzh5 = (z * y) * h[5]
    bzh6 = (3 * z ** 2 - r2) * h[6]
    bzh7 = (x * z) * h[7]
    bzh8 = (x ** 2 - y ** 2) * h[8]
    bzh9 = (3 * y * x ** 2 - y ** 3) * h[9]
    bzh10 = (x * z * y) * h[10]
    bzh11 = (5 * z ** 2 - r2) * y * h[11]
    bzh12 = (5 * z ** 3 - 3 * z * r2) * h[12]
    bzh13 = (5 * z ** 2 - r2) * x * h[13]
    bzh14 = (x ** 2 * z - y ** 2 * z) * h[14]
    bzh15 = (x ** 3 - 3 * x * y ** 2) * h[15]
    bz0 = h[0] if flags & _HO0 else 0.0
    if flags & _HO1:
        bz1 = bzh1 + bzh2 + bzh3
    else:
        bz1 = x * seq.gx[nodes][None, :] + y * seq.gy[nodes][None, :] + z * seq.gz[nodes][None, :]
    bz2 = bzh4 + bzh5 + bzh6 + bzh7 + bzh8 if flags & _HO2 else 0.0
    bz3 = bzh9 + bzh10 + bzh11 + bzh12 + bzh13 + bzh14 + bzh15 if flags & _HO3 else 0.0
    bz = bz0 + bz1 + bz2 + bz3
    if flags & _DWP:
        bz = bz + p["dw"][:, None] / (2 * np.pi * gamma)
    dt = seq.dt[first:last][None, :]
    # cumtrapz(dt', Bz): running sum of (B[n] + B[n+1]) dt[n] / 2
    phi = (-2 * np.pi * gamma) * np.cumsum(dt * (bz[:, 1:] + bz[:, :-1]) / 2, axis=1)
    tp = np.cumsum(seq.dt[first:last])
    dur = float(np.sum(seq.dt[first:last]))
    mxy_cols = mxy[:, None] * np.exp(1j * phi - tp[None, :] / p["T2"][:, None])
    adc = np.flatnonzero(seq.adc[first + 1:last + 1])
    sig = mxy_cols[:, adc].T @ p["csm"] if adc.size else None
    if last > first:
        mxy = mxy_cols[:, -1].copy()
    e1 = np.exp(-dur / p["T1"])
    mz = mz * e1 + p["rho"] * (1 - e1)
    return sig, mxy, mz


def _excitation(p, seq, first, last, flags, gamma, mxy, mz):
    """One RF block: spinor rotation per interval, then relaxation."""
    x, y, z = p["x"], p["y"], p["z"]
    mxy = mxy * np.exp(-1j * seq.psi[first])
    rows = []
    for n in range(first, last):
        dbz = (p["dw"] / (2 * np.pi * gamma) if flags & _DWX else 0.0) - seq.df[n] / gamma
        bz = (seq.gx[n] * x + seq.gy[n] * y + seq.gz[n] * z) + dbz
        b1 = complex(seq.b1[n])
        B = np.sqrt(abs(b1) ** 2 + np.abs(bz) ** 2)
        B[B == 0] = np.finfo(np.float64).eps
        phi = (-2 * np.pi * gamma) * (B * seq.dt[n])
        sn, cn = np.sin(phi / 2), np.cos(phi / 2)
        alpha = cn - 1j * (bz / B) * sn
        beta = -1j * (b1 / B) * sn
        ca = np.conj(alpha)
        mxy, mz = (2 * ca * beta * mz + ca ** 2 * mxy - beta ** 2 * np.conj(mxy),
                   (np.abs(alpha) ** 2 - np.abs(beta) ** 2) * mz
                   - 2 * np.real(alpha * beta * np.conj(mxy)))
        mxy = mxy * np.exp(-seq.dt[n] / p["T2"])
        e1 = np.exp(-seq.dt[n] / p["T1"])
        mz = mz * e1 + p["rho"] * (1 - e1)
        if seq.adc[n + 1]:
            rows.append((mxy * np.exp(1j * seq.psi[n + 1]))[None, :] @ p["csm"])
    mxy = mxy * np.exp(1j * seq.psi[last])
    return (np.concatenate(rows) if rows else None), mxy, mz


def _part(obj, a, b):
    return {"x": np.asarray(obj.x[a:b], np.float64), "y": np.asarray(obj.y[a:b], np.float64),
            "z": np.asarray(obj.z[a:b], np.float64), "rho": np.asarray(obj.rho[a:b], np.float64),
            "T1": np.asarray(obj.T1[a:b], np.float64), "T2": np.asarray(obj.T2[a:b], np.float64),
            "dw": np.asarray(obj.dw[a:b], np.float64),
            "csm": np.ascontiguousarray(np.asarray(obj.csm, np.complex128)[a:b])}


def run(obj, seq, parts, flags, gamma, mxy=None, mz=None, n_threads=1):
    """Whole block loop from the given state (default Mxy = 0, Mz = rho);
    returns (sig (SigmaADC x Nc), mxy, mz) like ``oracle.run``.  ``parts`` is
    (first, last, is_rf) arrays of 0-based inclusive node ranges."""
    ns = len(obj.rho)
    nc = np.asarray(obj.csm).shape[1]
    mxy = np.zeros(ns, np.complex128) if mxy is None else np.array(mxy, np.complex128)
    mz = np.array(obj.rho, np.float64) if mz is None else np.array(mz, np.float64)
    blocks = [(int(f), int(l), bool(r)) for f, l, r in zip(*parts)]
    total = sum(int(np.count_nonzero(seq.adc[f + 1:l + 1])) for f, l, _ in blocks)

    def work(rng):
        a, b = rng
        p = _part(obj, a, b)
        m, z = mxy[a:b].copy(), mz[a:b].copy()
        sig = np.zeros((total, nc), np.complex128)
        row = 0
        for f, l, rf in blocks:
            s, m, z = (_excitation if rf else _precession)(p, seq, f, l, flags, gamma, m, z)
            if s is not None:
                sig[row:row + s.shape[0]] = s
                row += s.shape[0]
        return sig, m, z

    ranges = _kfoldperm(ns, n_threads)
    # one BLAS thread per worker, as each reference thread runs its own GEMMs
    with threadpool_limits(limits=1, user_api="blas"):
        if len(ranges) > 1:
            with ThreadPoolExecutor(max_workers=len(ranges)) as ex:
                results = list(ex.map(work, ranges))
        else:
            results = [work(r) for r in ranges]
    sig = np.zeros((total, nc), np.complex128)
    for (a, b), (s, m, z) in zip(ranges, results):
        sig += s
        mxy[a:b], mz[a:b] = m, z
    return sig, mxy, mz
