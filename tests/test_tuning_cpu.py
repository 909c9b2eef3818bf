"""The persisted launch shapes (tuning.json) and the shape classes they are
keyed by."""
import copy
import json

import pytest

from hom_amd import tuning
from hom_amd.types import BlochHighOrder


def test_shape_classes_follow_the_term_split():
    # offset K slots: 3 per higher-order term, 6 per first-order term, 32 per MFMA
    assert tuning.shape_class(32, BlochHighOrder("1111").flags) == (4, 2)  # 36 + 24 slots
    assert tuning.shape_class(32, BlochHighOrder("1110").flags) == (4, 2)  # 15 + 24
    assert tuning.shape_class(8, BlochHighOrder("1100").flags) == (1, 1)   # 0 + 24
    assert tuning.shape_class(9, BlochHighOrder("1100", True, False).flags) == (2, 1)  # 18
    assert tuning.shape_class(1, BlochHighOrder("0000").flags) == (1, 1)
    assert tuning.shape_class(40, BlochHighOrder("1111").flags)[0] == 4   # per 32-coil pass


def test_committed_tuning_file_is_valid():
    doc = tuning.load()
    assert doc is not None, "highordermrisim.jl_amd/tuning.json is missing"
    classes = {(e["ncb"], e["nmh"]) for e in doc["plan"]}
    assert classes == {(n, m) for n in (1, 2, 3, 4) for m in (1, 2)}
    assert tuning.plan_spin_tile(32, BlochHighOrder("1111").flags, doc) in tuning.SPIN_TILES
    assert tuning.server_shape(doc) is not None


def test_validate_rejects_malformed_documents():
    good = {"schema": 1, "plan": [{"ncb": 4, "nmh": 2, "spin_tile": 256,
                                   "kernel_ms": {"128": 7.4, "256": 7.3}}],
            "server": {"spin_tile": 256, "main_grid": 768}}
    tuning.validate(copy.deepcopy(good))
    bad = []
    for path, value in ((("schema",), 2), (("plan", 0, "ncb"), 5), (("plan", 0, "spin_tile"), 100),
                        (("plan", 0, "spin_tile"), 128), (("server", "main_grid"), 0)):
        d = copy.deepcopy(good)
        tgt = d
        for k in path[:-1]:
            tgt = tgt[k]
        tgt[path[-1]] = value
        bad.append(d)
    dup = copy.deepcopy(good)
    dup["plan"].append(dict(dup["plan"][0]))
    bad.append(dup)
    for d in bad:
        with pytest.raises(ValueError):
            tuning.validate(d)
    json.dumps(good)
