"""Seeded synthetic inputs: the reference's deterministic test workload and
the benchmark configurations.

- :func:`high_order_workload` restates test/support/high_order_workload.jl:4-32
  (a DiscreteSequence-level closed form, no sampling step involved).
- :func:`build_sequence` is this package's own sampler for hard RF pulses and
  free-precession blocks.  It places nodes at block boundaries, at every ADC
  sample time and at most ``dt_max`` apart in between; RF blocks are cut into
  equal steps of at most ``dt_rf``.  KomaMRIBase's ``discretize`` (call site
  src/datatypes/HO_DiscreteSequence.jl:136-179) is a third-party sampler whose
  exact node placement is not reproduced: the GPU path and the oracle always
  consume the same node arrays, so parity does not depend on it.
- :func:`make_config` builds the five benchmark configurations: voxel-centre
  grids, uniform tissue draws, Gaussian receive maps on a ring, a 90-degree hard
  pulse (B1 = 1/(4 gamma tau), as in the reference README example) and slew-
  and amplitude-limited Archimedean spiral readouts with sinusoidal
  high-order field terms.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .types import GAMMA, HO_DiscreteSequence, HO_Phantom


# --------------------------------------------------------------------------
# test/support/high_order_workload.jl:4-32
# --------------------------------------------------------------------------
def high_order_workload(nspins=32768, nsteps=32, ncoils=32, rf=False):
    """Returns (obj, seqd, mxy0, mz0) with the reference's closed-form inputs:
    a 1-D curve of spins, 32 (nsteps) intervals over 2 ms (0.2 ms with RF),
    every node but the first acquired, and 16 sinusoidal field terms."""
    u = np.array([0.0]) if nspins == 1 else np.linspace(-1.0, 1.0, nspins)
    coils = np.arange(1, ncoils + 1)
    csm = np.exp(1j * 2 * np.pi * np.outer(u, coils) / ncoils) / (1 + coils / ncoils)
    obj = HO_Phantom(x=0.08 * u, y=0.06 * np.sin(3 * u), z=0.04 * np.cos(5 * u),
                     rho=0.8 + 0.2 * np.cos(u), T1=0.9 + 0.1 * np.cos(u),
                     T2=0.07 + 0.01 * np.cos(u), dw=80 * u, csm=np.asfortranarray(csm))
    t = np.linspace(0.0, 2e-4 if rf else 2e-3, nsteps + 1)
    q = np.linspace(0.0, 2 * np.pi, t.size)
    adc = np.ones(t.size, dtype=bool)
    adc[0] = False
    h = np.empty((16, t.size))
    for term in range(1, 17):
        scale = 1e-7 if term == 1 else 2e-4 if term <= 4 else 1e-3 if term <= 9 else 1e-2
        h[term - 1] = scale * np.sin(q + term / 4)
    seqd = HO_DiscreteSequence(
        t=t, gx=2e-4 * np.sin(q), gy=2e-4 * np.cos(q), gz=np.full(t.size, 1e-4),
        b1=(8e-6 * np.exp(1j * q / 4)) if rf else np.zeros(t.size, dtype=np.complex128),
        df=(100 * np.sin(q)) if rf else np.zeros(t.size),
        psi=(0.2 * q) if rf else np.zeros(t.size),
        adc=adc, excitation=np.full(nsteps, rf), h=h, dt=np.diff(t))
    mxy0 = (0.3 * obj.rho + 0.6j * obj.rho).astype(np.complex128)
    mz0 = 0.5 * obj.rho
    return obj, seqd, mxy0, mz0


# --------------------------------------------------------------------------
# Sequence construction
# --------------------------------------------------------------------------
@dataclass
class RFBlock:
    """Hard pulse of constant complex amplitude ``b1`` (T) for ``dur`` seconds."""
    dur: float
    b1: complex
    df: float = 0.0
    psi: float = 0.0
    adc_times: np.ndarray | None = None   # relative to block start, within (0, dur]


@dataclass
class FreeBlock:
    """Free precession for ``dur`` seconds.  ``grad(t_rel)`` returns (gx, gy, gz)
    arrays in T/m; ``adc_times`` are relative sample times within (0, dur]."""
    dur: float
    adc_times: np.ndarray | None = None
    grad: object = None


def build_sequence(blocks, fields=None, dt_max=1e-3, dt_rf=5e-5, measured_first_order=True):
    """Samples a list of RFBlock / FreeBlock onto one node grid.

    ``fields(t)`` returns a 16 x len(t) array of coefficient waveforms at
    absolute times t.  When ``measured_first_order`` is true, rows 1..3 get the
    nominal gradient added (so h1..h3 equal the applied gradient plus whatever
    ``fields`` puts there).  Adjacent blocks share their boundary node when
    their values agree there; where a waveform jumps, both sides are kept at
    the same time with a zero-length interval between them, as the
    reference's discretization keeps field discontinuities
    (src/datatypes/HO_DiscreteSequence.jl:85-120)."""
    ts, gx, gy, gz, b1, df, psi, adc, exc = [], [], [], [], [], [], [], [], []
    t0 = 0.0
    for bi, blk in enumerate(blocks):
        if isinstance(blk, RFBlock):
            m = max(1, int(math.ceil(blk.dur / dt_rf - 1e-9)))
            rel = np.linspace(0.0, blk.dur, m + 1)
            a = np.zeros(rel.size, dtype=bool)
            if blk.adc_times is not None and len(blk.adc_times):
                rel = np.union1d(rel, np.asarray(blk.adc_times, dtype=np.float64))
                a = np.isin(rel, blk.adc_times)
            g = np.zeros((3, rel.size))
            b = np.full(rel.size, complex(blk.b1))
            b[-1] = 0.0
            d = np.full(rel.size, blk.df)
            p = np.full(rel.size, blk.psi)
            e = np.ones(rel.size - 1, dtype=bool)
        else:
            adc_t = np.zeros(0) if blk.adc_times is None else np.asarray(blk.adc_times, np.float64)
            key = np.union1d(np.array([0.0, blk.dur]), adc_t)
            extra = []
            for lo, hi in zip(key[:-1], key[1:]):
                if hi - lo > dt_max:
                    k = int(math.ceil((hi - lo) / dt_max - 1e-9))
                    extra.append(lo + (hi - lo) * np.arange(1, k) / k)
            rel = np.union1d(key, np.concatenate(extra)) if extra else key
            a = np.isin(rel, adc_t)
            g = np.zeros((3, rel.size)) if blk.grad is None else np.asarray(blk.grad(rel), np.float64)
            b = np.zeros(rel.size, dtype=np.complex128)
            d = np.zeros(rel.size)
            p = np.zeros(rel.size)
            e = np.zeros(rel.size - 1, dtype=bool)
        start = 0 if bi == 0 else 1
        if bi > 0:
            same = (gx[-1][-1] == g[0][0] and gy[-1][-1] == g[1][0] and gz[-1][-1] == g[2][0]
                    and b1[-1][-1] == b[0] and df[-1][-1] == d[0] and psi[-1][-1] == p[0])
            if same:
                start = 1  # one shared boundary node
            else:
                # A jump keeps both sides at the same time: a zero-length interval.
                start = 0
                exc.append(np.zeros(1, dtype=bool))
        ts.append(t0 + rel[start:])
        gx.append(g[0][start:].copy()); gy.append(g[1][start:].copy()); gz.append(g[2][start:].copy())
        b1.append(b[start:].copy()); df.append(d[start:].copy()); psi.append(p[start:].copy())
        adc.append(a[start:])
        exc.append(e)
        t0 += blk.dur
    t = np.concatenate(ts)
    G = np.vstack([np.concatenate(gx), np.concatenate(gy), np.concatenate(gz)])
    h = np.zeros((16, t.size)) if fields is None else np.asarray(fields(t), np.float64).copy()
    if measured_first_order:
        h[1:4] += G
    return HO_DiscreteSequence(t=t, gx=G[0], gy=G[1], gz=G[2], b1=np.concatenate(b1),
                               df=np.concatenate(df), psi=np.concatenate(psi),
                               adc=np.concatenate(adc), excitation=np.concatenate(exc), h=h)


def spiral_gradient(n_samples, dwell, fov, n_interleaves=1, gmax=30e-3, smax=150.0,
                    gamma=GAMMA):
    """Archimedean spiral k(theta) = lam*theta*exp(i theta), lam =
    n_interleaves/(2 pi fov), advanced one dwell at a time at the largest
    angular rate allowed by the amplitude (gmax, T/m) and slew (smax, T/m/s)
    limits.  Returns the complex gradient Gx + iGy at n_samples + 1 points."""
    lam = n_interleaves / (2 * np.pi * fov)
    th = np.zeros(n_samples + 2)
    w = 0.0
    cap_s = gamma * smax / lam
    for i in range(1, n_samples + 2):
        theta = th[i - 1]
        w_amp = gamma * gmax / (lam * math.sqrt(1 + theta * theta))
        acc = (cap_s - w * w * math.sqrt(4 + theta * theta)) / math.sqrt(1 + theta * theta)
        w = min(w + max(acc, 0.0) * dwell, w_amp)
        th[i] = theta + w * dwell
    k = lam * th * np.exp(1j * th)
    return np.diff(k) / (gamma * dwell)


def sinusoid_fields(rng, amp0=1e-6, amp2=1e-4, amp3=1e-3, n_sin=4, fmin=10.0, fmax=500.0):
    """16 coefficient functions: h0 (T), h4..h8 (T/m^2) and h9..h15 (T/m^3),
    each amplitude * mean of n_sin sinusoids with random frequency and phase;
    h1..h3 are left at zero for the nominal gradient to be added."""
    freqs = rng.uniform(fmin, fmax, size=(16, n_sin))
    phases = rng.uniform(0.0, 2 * np.pi, size=(16, n_sin))
    amp = np.zeros(16)
    amp[0] = amp0
    amp[4:9] = amp2
    amp[9:16] = amp3

    def fields(t):
        t = np.asarray(t, dtype=np.float64)
        out = np.zeros((16, t.size))
        for j in range(16):
            if amp[j] == 0.0:
                continue
            acc = np.zeros(t.size)
            for f, p in zip(freqs[j], phases[j]):
                acc += np.sin(2 * np.pi * f * t + p)
            out[j] = amp[j] * acc / n_sin
        return out

    return fields


def grid_centres(n, fov):
    """Voxel centres of an n-point grid over fov (metres), symmetric about 0."""
    return -fov / 2 + (np.arange(n) + 0.5) * fov / n


def gaussian_csm(x, y, z, n_coils, rng, radius=0.15, sigma=0.08):
    """Gaussian receive maps centred on a ring of the given radius in z = 0,
    each with a random constant phase.  Returns a Fortran-ordered Ns x Nc map."""
    ang = 2 * np.pi * np.arange(n_coils) / max(n_coils, 1)
    cx, cy = radius * np.cos(ang), radius * np.sin(ang)
    ph = rng.uniform(0.0, 2 * np.pi, size=n_coils)
    csm = np.empty((x.size, n_coils), dtype=np.complex128, order="F")
    for c in range(n_coils):
        r2 = (x - cx[c]) ** 2 + (y - cy[c]) ** 2 + z ** 2
        csm[:, c] = np.exp(-r2 / (2 * sigma * sigma)) * np.exp(1j * ph[c])
    return csm


def tissue(rng, n):
    rho = rng.uniform(0.0, 1.0, n)
    T1 = rng.uniform(0.5, 2.0, n)
    T2 = rng.uniform(0.02, 0.2, n)
    dw = 2 * np.pi * rng.normal(0.0, 20.0, n)
    return rho, T1, T2, dw


CONFIG_SEEDS = (20260, 20261, 20262, 20263, 20264)


@dataclass
class Config:
    name: str
    grid: tuple
    fov: tuple
    n_coils: int
    n_interleaves: int
    n_adc: int
    tr: float | None
    dwell: float = 2e-6
    tau: float = 1e-4


CONFIGS = {
    0: Config("2d-100x100-1coil-spiral5000", (100, 100, 1), (0.2, 0.2, 0.0), 1, 1, 5000, None),
    1: Config("3d-100x100x35-8coil-4x4750", (100, 100, 35), (0.2, 0.2, 0.07), 8, 4, 4750, 0.5),
    2: Config("3d-128x128x64-32coil-16x8000", (128, 128, 64), (0.2, 0.2, 0.07), 32, 16, 8000, 0.5),
    4: Config("3d-200x200x100-32coil-32x10000", (200, 200, 100), (0.2, 0.2, 0.07), 32, 32, 10000, 0.5),
}


def config_sequence(cfg: Config, rng, gamma=GAMMA):
    """RF + spiral readout per interleave (then idle to TR when tr is set)."""
    fields = sinusoid_fields(rng)
    g = spiral_gradient(cfg.n_adc, cfg.dwell, cfg.fov[0], cfg.n_interleaves, gamma=gamma)
    blocks = []
    for il in range(cfg.n_interleaves):
        rot = np.exp(1j * 2 * np.pi * il / cfg.n_interleaves)
        gi = g * rot
        times = cfg.dwell * np.arange(1, cfg.n_adc + 1)
        dur = times[-1] if cfg.tr is None else cfg.tr - cfg.tau

        def grad(rel, gi=gi, n=cfg.n_adc, dwell=cfg.dwell):
            idx = np.clip(np.rint(rel / dwell).astype(np.int64) - 1, 0, n)
            inside = (rel > 0) & (rel <= n * dwell * (1 + 1e-12))
            gxy = np.where(inside, gi[idx], 0.0)
            return np.vstack([gxy.real, gxy.imag, np.zeros(rel.size)])

        blocks.append(RFBlock(cfg.tau, 1.0 / (4 * gamma * cfg.tau)))
        blocks.append(FreeBlock(dur, times, grad))
    return build_sequence(blocks, fields)


def make_config(index: int, spin_subsample: int | None = None, gamma=GAMMA):
    """Builds (phantom, sequence, config) for workload 0 (2D, 10,000 spins, 1
    coil, one 5000-sample spiral), 1 (350,000 spins, 8 coils, 4 x 4750), 2
    (1,048,576 spins, 32 coils, 16 x 8000) or 4 (4,000,000 spins, 32 coils,
    32 x 10000; RF every TR).  With
    ``spin_subsample`` the phantom keeps every k-th spin of the full grid
    (signal is additive over spins, so a subsample is an exact sub-problem)."""
    cfg = CONFIGS[index]
    rng = np.random.Generator(np.random.PCG64(CONFIG_SEEDS[index]))
    nx, ny, nz = cfg.grid
    xs, ys = grid_centres(nx, cfg.fov[0]), grid_centres(ny, cfg.fov[1])
    zs = np.zeros(1) if nz == 1 else grid_centres(nz, cfg.fov[2])
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    x, y, z = X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")
    rho, T1, T2, dw = tissue(rng, x.size)
    csm = (np.ones((x.size, 1), dtype=np.complex128, order="F") if cfg.n_coils == 1
           else gaussian_csm(x, y, z, cfg.n_coils, rng))
    seq = config_sequence(cfg, rng, gamma)
    obj = HO_Phantom(x=x, y=y, z=z, rho=rho, T1=T1, T2=T2, dw=dw, csm=csm, name=cfg.name)
    if spin_subsample:
        obj = obj[::spin_subsample]
    return obj, seq, cfg


def serving_inputs(n_spins: int = 1 << 20, n_coils: int = 32, n_samples: int = 20_000,
                   seed: int = CONFIG_SEEDS[3], gamma=GAMMA):
    """Resident inputs of the serving stream: a 3D phantom with config-0
    tissue distributions on a near-cubic grid over 0.2 x 0.2 x 0.07 m (the
    first ``n_spins`` grid points), ``n_coils`` Gaussian sensitivity maps, and
    one 90-degree hard RF followed by an ``n_samples`` spiral readout with
    config-0 field terms.  Requests take spin ranges, coil prefixes and
    readout prefixes of these."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nz = max(1, int(round((n_spins / 4) ** (1 / 3))))
    nxy = int(math.ceil(math.sqrt(n_spins / nz)))
    xs, ys = grid_centres(nxy, 0.2), grid_centres(nxy, 0.2)
    zs = grid_centres(nz, 0.07) if nz > 1 else np.zeros(1)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    x, y, z = (A.ravel(order="F")[:n_spins] for A in (X, Y, Z))
    rho, T1, T2, dw = tissue(rng, n_spins)
    csm = (np.ones((n_spins, 1), dtype=np.complex128, order="F") if n_coils == 1
           else gaussian_csm(x, y, z, n_coils, rng))
    cfg = Config("serving-resident", (nxy, nxy, nz), (0.2, 0.2, 0.07), n_coils, 1, n_samples, None)
    seq = config_sequence(cfg, rng, gamma)
    obj = HO_Phantom(x=x, y=y, z=z, rho=rho, T1=T1, T2=T2, dw=dw, csm=csm, name=cfg.name)
    return obj, seq
