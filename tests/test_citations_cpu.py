"""Every `file.jl:a-b` citation of the reference in the headers, the oracle, the
package, the tests and the two integration documents names a file of the
reference and lines that exist in it.  Reads the reference tree as text;
skipped where it is absent (it is not shipped with this repository)."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REFERENCE = Path("/root/reference")
CITE = re.compile(r"([\w/.\-]+\.(?:jl|toml)):(\d+)(?:-(\d+))?((?:,\s?\d+(?:-\d+)?)*)")
SCOPE = ["include", "oracle", "highordermrisim.jl_amd", "tests", "DESIGN.md", "INTEGRATION.md",
         "bench.py", "__graft_entry__.py"]
SUFFIXES = {".h", ".c", ".cpp", ".hip", ".py", ".jl", ".md"}


def _sources():
    for entry in SCOPE:
        p = ROOT / entry
        if p.is_file():
            yield p
        elif p.is_dir():
            for q in sorted(p.rglob("*")):
                if q.is_file() and q.suffix in SUFFIXES and "__pycache__" not in q.parts:
                    yield q


def _reference_lines():
    by_name = {}
    for q in REFERENCE.rglob("*"):
        if q.suffix in (".jl", ".toml") and q.is_file():
            by_name.setdefault(q.name, []).append(q)
    return by_name


@pytest.mark.skipif(not REFERENCE.is_dir(), reason="the reference tree is not present")
def test_every_reference_citation_is_in_range():
    by_name = _reference_lines()
    counts = {}
    bad, seen = [], 0
    for src in _sources():
        if src.name == Path(__file__).name:
            continue
        for ln, line in enumerate(src.read_text(errors="ignore").splitlines(), 1):
            for m in CITE.finditer(line):
                cited = m.group(1)
                cands = by_name.get(Path(cited).name, [])
                if "/" in cited:
                    cands = [c for c in cands if str(c).endswith(cited.lstrip("./"))] or cands
                if not cands:
                    bad.append(f"{src.relative_to(ROOT)}:{ln}: {m.group(0)} (no such file)")
                    continue
                n = max(counts.setdefault(c, len(c.read_text(errors="ignore").splitlines()))
                        for c in cands)
                ranges = [(int(m.group(2)), int(m.group(3) or m.group(2)))]
                for extra in filter(None, (m.group(4) or "").replace(" ", "").split(",")):
                    a, _, b = extra.partition("-")
                    ranges.append((int(a), int(b or a)))
                seen += 1
                for a, b in ranges:
                    if not 1 <= a <= b <= n:
                        bad.append(f"{src.relative_to(ROOT)}:{ln}: {m.group(0)} (file has {n} lines)")
    assert seen > 100, f"only {seen} citations found: the pattern no longer matches"
    assert not bad, "\n".join(bad)
