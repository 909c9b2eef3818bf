"""Host-side mirror of the reference's data types for the BlochHighOrder path.

Each type keeps the reference's field names, defaults, argument meaning and
error behaviour (``ValueError`` stands in for Julia's ``ArgumentError``;
``DimensionMismatch`` is its own class):

- :class:`BlochHighOrder`       src/simulation/BlochHighOrder.jl:21-43
- :class:`HO_Phantom`           src/datatypes/HO_Phantom.jl:47-71, 97-109
- :class:`HO_DiscreteSequence`  src/datatypes/HO_DiscreteSequence.jl:31-80
- :func:`get_sim_ranges`        block partition consumed by
                                src/simulation/Simulate.jl:113-116

Arrays are numpy: real per-spin vectors are float64, the coil sensitivity map
is a Fortran-ordered (column-major, like Julia) complex128 ``Ns x Nc`` matrix,
and the 16 coefficient waveforms are one C-ordered ``16 x n`` float64 array.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

import numpy as np

# KomaMRIBase's gyromagnetic ratio in Hz/T.  Every entry point takes gamma as
# an argument, so a different constant can be passed through unchanged.
GAMMA = 42.5774688e6


class DimensionMismatch(ValueError):
    """Shape error, raised before any device work (Julia's DimensionMismatch)."""


# Flag bits shared with the C ABI (include/hom_abi.h).
HO0, HO1, HO2, HO3, DW_EXCITATION, DW_PRECESSION = (1 << i for i in range(6))


@dataclass(frozen=True)
class BlochHighOrder:
    """Field-order mask: ``BlochHighOrder("1111")``, ``BlochHighOrder("1100",
    False, True)`` or keyword flags.  Mirrors src/simulation/BlochHighOrder.jl:21-43:
    the name must be four binary digits, order zero first, and must agree
    with the flags."""

    ho0: bool = True
    ho1: bool = True
    ho2: bool = True
    ho3: bool = True
    dw_excitation: bool = True
    dw_precession: bool = True
    name: str = ""

    def __init__(self, name: str | None = None, dw_excitation: bool = True,
                 dw_precession: bool = True, *, ho0=None, ho1=None, ho2=None, ho3=None):
        kw = (ho0, ho1, ho2, ho3)
        if name is not None and not isinstance(name, str):
            raise ValueError("The field-order mask must be a string.")
        if name is not None and any(v is not None for v in kw):
            flags = tuple(True if v is None else bool(v) for v in kw)
            expected = "".join(str(int(f)) for f in flags)
            if name != expected:
                raise ValueError("The field-order name must match ho0 through ho3.")
        elif name is not None:
            if not (len(name) == 4 and re.fullmatch(r"[01]{4}", name)):
                raise ValueError("The field-order mask must contain four binary digits, "
                                 "from order zero to three.")
            flags = tuple(ch == "1" for ch in name)
        else:
            flags = tuple(True if v is None else bool(v) for v in kw)
            name = "".join(str(int(f)) for f in flags)
        object.__setattr__(self, "ho0", flags[0])
        object.__setattr__(self, "ho1", flags[1])
        object.__setattr__(self, "ho2", flags[2])
        object.__setattr__(self, "ho3", flags[3])
        object.__setattr__(self, "dw_excitation", bool(dw_excitation))
        object.__setattr__(self, "dw_precession", bool(dw_precession))
        object.__setattr__(self, "name", name)

    @property
    def flags(self) -> int:
        """Bit mask in the C ABI's layout (HOM_HO0 ... HOM_DW_PRECESSION)."""
        return ((HO0 if self.ho0 else 0) | (HO1 if self.ho1 else 0) | (HO2 if self.ho2 else 0)
                | (HO3 if self.ho3 else 0) | (DW_EXCITATION if self.dw_excitation else 0)
                | (DW_PRECESSION if self.dw_precession else 0))

    @property
    def n_terms(self) -> int:
        """Spatial basis terms used in free precession (1st order is always
        present: measured h1..h3 or the nominal gradients)."""
        return 3 + 5 * self.ho2 + 7 * self.ho3 + (1 if self.ho0 else 0)

    def __repr__(self) -> str:
        return (f"BlochHighOrder[[{self.name}] zero order={self.ho0} | first order={self.ho1} | "
                f"second order={self.ho2} | third order={self.ho3} | "
                f"Δw_excitation={self.dw_excitation} | Δw_precession={self.dw_precession}]")


def _vec(v, n, default):
    if v is None:
        return np.full(n, default, dtype=np.float64)
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))


@dataclass
class HO_Phantom:
    """Spin phantom with a receive-sensitivity map (src/datatypes/HO_Phantom.jl:47-71).

    Defaults follow the reference: y = z = 0, rho = 1, T1 = T2 = 1e6 s,
    dw = 0 rad/s and a single unit-sensitivity coil."""

    x: np.ndarray
    y: np.ndarray | None = None
    z: np.ndarray | None = None
    rho: np.ndarray | None = None
    T1: np.ndarray | None = None
    T2: np.ndarray | None = None
    dw: np.ndarray | None = None
    csm: np.ndarray | None = None
    name: str = "spins"

    def __post_init__(self):
        self.x = np.ascontiguousarray(np.asarray(self.x, dtype=np.float64).reshape(-1))
        n = self.x.size
        self.y = _vec(self.y, n, 0.0)
        self.z = _vec(self.z, n, 0.0)
        self.rho = _vec(self.rho, n, 1.0)
        self.T1 = _vec(self.T1, n, 1e6)
        self.T2 = _vec(self.T2, n, 1e6)
        self.dw = _vec(self.dw, n, 0.0)
        if self.csm is None:
            self.csm = np.ones((n, 1), dtype=np.complex128, order="F")
        else:
            csm = np.asarray(self.csm, dtype=np.complex128)
            if csm.ndim == 1:
                csm = csm.reshape(-1, 1)
            self.csm = np.asfortranarray(csm)

    def __len__(self) -> int:
        return self.rho.size

    @property
    def n_coils(self) -> int:
        return self.csm.shape[1]

    def validate(self) -> None:
        """src/datatypes/HO_Phantom.jl:97-109: one value per spin everywhere."""
        n = len(self)
        for nm in ("x", "y", "z", "rho", "T1", "T2", "dw"):
            if getattr(self, nm).size != n:
                raise DimensionMismatch(f"Phantom field {nm} must contain one value per spin.")
        if self.csm.shape[0] != n:
            raise DimensionMismatch("The CSM must have one row per spin.")

    def __getitem__(self, idx) -> "HO_Phantom":
        return HO_Phantom(x=self.x[idx], y=self.y[idx], z=self.z[idx], rho=self.rho[idx],
                          T1=self.T1[idx], T2=self.T2[idx], dw=self.dw[idx],
                          csm=self.csm[idx, :], name=self.name)


@dataclass
class HO_DiscreteSequence:
    """Sampled sequence: KomaMRI's DiscreteSequence node arrays plus the 16
    field-coefficient waveforms (src/datatypes/HO_DiscreteSequence.jl:31-49).

    Node arrays have length n; ``dt`` and ``excitation`` have length n-1.  ``h``
    is ``16 x n`` with row j holding h_j in the reference's order (h0 in T,
    h1..h3 in T/m, h4..h8 in T/m^2, h9..h15 in T/m^3)."""

    t: np.ndarray
    gx: np.ndarray
    gy: np.ndarray
    gz: np.ndarray
    b1: np.ndarray
    df: np.ndarray
    psi: np.ndarray
    adc: np.ndarray
    excitation: np.ndarray
    h: np.ndarray
    dt: np.ndarray = field(default=None)

    def __post_init__(self):
        f64 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
        self.t = f64(self.t)
        n = self.t.size
        self.dt = np.diff(self.t) if self.dt is None else f64(self.dt)
        self.gx, self.gy, self.gz = f64(self.gx), f64(self.gy), f64(self.gz)
        self.b1 = np.ascontiguousarray(np.asarray(self.b1, dtype=np.complex128).reshape(-1))
        self.df, self.psi = f64(self.df), f64(self.psi)
        self.adc = np.ascontiguousarray(np.asarray(self.adc, dtype=bool).reshape(-1))
        self.excitation = np.ascontiguousarray(np.asarray(self.excitation, dtype=bool).reshape(-1))
        h = np.asarray(self.h, dtype=np.float64)
        if h.shape != (16, n):
            raise DimensionMismatch("Sixteen coefficient waveforms with one value per node are required.")
        self.h = np.ascontiguousarray(h)
        self.validate(n)

    def validate(self, n=None) -> None:
        n = self.t.size if n is None else n
        for nm in ("gx", "gy", "gz", "b1", "df", "psi", "adc"):
            if getattr(self, nm).size != n:
                raise DimensionMismatch(f"Sequence field {nm} must contain one value per node.")
        if self.dt.size != max(n - 1, 0) or self.excitation.size != max(n - 1, 0):
            raise DimensionMismatch("Interval arrays must contain one value per interval.")
        if self.h.shape != (16, n):
            raise DimensionMismatch("Sixteen coefficient waveforms with one value per node are required.")
        if n and np.any(self.dt < 0):
            raise ValueError("Time steps must be nonnegative.")

    def __len__(self) -> int:
        """Number of intervals, like Base.length(::HO_DiscreteSequence)."""
        return max(self.t.size - 1, 0)

    @property
    def n_nodes(self) -> int:
        return self.t.size

    def block_adc(self, first: int, last: int) -> int:
        """ADC samples owned by the 0-based inclusive node range [first, last]:
        the shared first node belongs to the previous block
        (src/simulation/SimulatorCore.jl:115-116)."""
        return int(np.count_nonzero(self.adc[first + 1:last + 1]))

    def view(self, first: int, last: int) -> "HO_DiscreteSequence":
        """Nodes first..last inclusive (0-based), like view(hoseqd, first:last)."""
        sl = slice(first, last + 1)
        return HO_DiscreteSequence(t=self.t[sl], gx=self.gx[sl], gy=self.gy[sl], gz=self.gz[sl],
                                   b1=self.b1[sl], df=self.df[sl], psi=self.psi[sl],
                                   adc=self.adc[sl], excitation=self.excitation[first:last],
                                   h=self.h[:, sl], dt=self.dt[first:last])


def get_sim_ranges(seqd: HO_DiscreteSequence, max_block_length=np.inf,
                   max_rf_block_length=None):
    """Block partition: maximal runs of equal excitation flag over intervals,
    each cut into chunks of at most ``max_block_length`` (RF runs:
    ``max_rf_block_length``) intervals.  Returns (first, last, is_rf) arrays of
    0-based inclusive node ranges; adjacent ranges share a node.

    This restates the contract the reference relies on from
    KomaMRICore.get_sim_ranges (call site src/simulation/Simulate.jl:113-116):
    the exact chunk boundaries inside a run are KomaMRICore's own and are not
    reproduced; for free precession a different partition changes only
    rounding, and RF nodes are applied one at a time either way."""
    if max_rf_block_length is None:
        max_rf_block_length = max_block_length
    n = len(seqd)
    big = np.iinfo(np.int64).max // 4

    def lim(v):
        if v == np.inf:
            return big
        if not (isinstance(v, (int, np.integer)) and v > 0):
            raise ValueError("Block lengths must be positive integers or Inf.")
        return int(v)

    mb, mr = lim(max_block_length), lim(max_rf_block_length)
    first, last, is_rf = [], [], []
    i = 0
    exc = seqd.excitation
    while i < n:
        f = bool(exc[i])
        j = i
        while j < n and bool(exc[j]) == f:
            j += 1
        step = mr if f else mb
        a = i
        while a < j:
            b = min(a + step, j)
            first.append(a)
            last.append(b)
            is_rf.append(f)
            a = b
        i = j
    return (np.asarray(first, dtype=np.int64), np.asarray(last, dtype=np.int64),
            np.asarray(is_rf, dtype=np.uint8))
