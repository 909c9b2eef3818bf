"""Persisted launch shapes of the mixed-precision kernels (the autotuner's output).

``tuning.json`` (next to this file) is written by ``tune.py`` on the target
GPU and read by :class:`runtime.Plan` (before every mixed run) and by
:class:`serving.SimulationServer` (as its default configuration).  It holds

* ``plan``: one entry per shape class of the plan path -- coil blocks
  ``ncb`` = ceil(2 Nc / 16) of one 32-coil pass (1..4) and offset MFMAs
  ``nmh`` per tile (1 or 2, fixed by the field orders and the off-resonance
  flag) -- with the spin tile (16-spin groups per workgroup item) whose
  measured kernel time was lowest on a 1M-spin probe, and the times of every
  candidate.  A plan uses it as the upper bound of the library's per-segment
  choice, which takes smaller tiles when the work items would not fill the
  GPU;
* ``server``: the persistent kernel's spin tile and grid.

The reference has no counterpart: its CUDA launch shapes are fixed
(``HighOrderGPU.jl:16-35`` takes groupsize and tiles as constructor
arguments).  A missing file or a shape class it does not list leaves the
library defaults (256 groups; the server's own defaults).
"""
from __future__ import annotations

import json
import math
from pathlib import Path

SCHEMA = 1
PATH = Path(__file__).resolve().with_name("tuning.json")

HOM_HO2, HOM_HO3, HOM_DW_PRECESSION = 1 << 2, 1 << 3, 1 << 5
SPIN_TILES = (16, 32, 64, 128, 256, 512, 1024, 2048, 4096)
DEFAULT_SPIN_TILE = 256  # the library's own default (hom_plan_create)

_cache: dict = {}


def shape_class(n_coils: int, flags: int) -> tuple[int, int]:
    """(ncb, nmh) of a mixed run: the coil blocks of its largest 32-coil pass
    and its offset MFMAs per tile, as the library's term split counts them
    (three slots per higher-order term, six per first-order term, 32 per
    MFMA)."""
    ncb = min(4, max(1, math.ceil(2 * int(n_coils) / 16)))
    nho = (5 if flags & HOM_HO2 else 0) + (7 if flags & HOM_HO3 else 0)
    n1 = 3 + (1 if flags & HOM_DW_PRECESSION else 0)
    return ncb, -(-(3 * nho + 6 * n1) // 32)


def validate(doc: dict) -> dict:
    """Checks the file's schema; returns the document or raises ValueError."""
    if not isinstance(doc, dict) or doc.get("schema") != SCHEMA:
        raise ValueError(f"tuning file: schema {SCHEMA} expected")
    plan = doc.get("plan")
    if not isinstance(plan, list):
        raise ValueError("tuning file: 'plan' must be a list")
    seen = set()
    for e in plan:
        if not isinstance(e, dict):
            raise ValueError("tuning file: plan entries must be objects")
        ncb, nmh, tile = e.get("ncb"), e.get("nmh"), e.get("spin_tile")
        if ncb not in (1, 2, 3, 4) or nmh not in (1, 2) or tile not in SPIN_TILES:
            raise ValueError(f"tuning file: bad plan entry {e}")
        if (ncb, nmh) in seen:
            raise ValueError(f"tuning file: duplicate plan entry {(ncb, nmh)}")
        seen.add((ncb, nmh))
        times = e.get("kernel_ms", {})
        if not isinstance(times, dict) or not all(
                int(k) in SPIN_TILES and float(v) > 0 for k, v in times.items()):
            raise ValueError(f"tuning file: bad kernel_ms in {e}")
        if times and min(times, key=lambda k: float(times[k])) != str(tile):
            raise ValueError(f"tuning file: spin_tile {tile} is not the fastest of {times}")
    srv = doc.get("server")
    if srv is not None:
        if not isinstance(srv, dict) or srv.get("spin_tile") not in SPIN_TILES or \
                not (isinstance(srv.get("main_grid"), int) and srv["main_grid"] > 0):
            raise ValueError(f"tuning file: bad server entry {srv}")
    return doc


def load(path: Path | str | None = None) -> dict | None:
    """The tuning document (cached per path), or None when the file is absent."""
    p = Path(path) if path else PATH
    key = str(p)
    if key not in _cache:
        _cache[key] = validate(json.loads(p.read_text())) if p.exists() else None
    return _cache[key]


def plan_spin_tile(n_coils: int, flags: int, doc: dict | None = None) -> int | None:
    """Tuned spin tile of a mixed run of n_coils coils with these method flags."""
    doc = load() if doc is None else doc
    if not doc:
        return None
    ncb, nmh = shape_class(n_coils, flags)
    for e in doc["plan"]:
        if e["ncb"] == ncb and e["nmh"] == nmh:
            return int(e["spin_tile"])
    return None


def server_shape(doc: dict | None = None) -> tuple[int, int] | None:
    """Tuned (spin tile, grid) of the serving kernel."""
    doc = load() if doc is None else doc
    if not doc or not doc.get("server"):
        return None
    return int(doc["server"]["spin_tile"]), int(doc["server"]["main_grid"])
