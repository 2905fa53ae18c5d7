"""The autotuner's JSON cache on the CPU: entries recorded and looked up by key, files replaced whole, and entries that
do not fit the library's variants (or unreadable files) ignored rather than trusted."""
import json
import os

import numpy as np

from curvedspacesim_amd import _native as nat


def test_record_and_lookup(tmp_path):
    path = str(tmp_path / "t.json")
    assert nat.lookup_tuning(path, "a") is None  # no file yet
    ms = np.arange(1.0, nat.LDS_VARIANTS + 1)
    nat.record_tuning(path, "a", 2, ms)
    nat.record_tuning(path, "b", 0, ms * 2)
    v, got = nat.lookup_tuning(path, "a")
    assert v == 2 and np.array_equal(got, ms)
    assert nat.lookup_tuning(path, "b")[0] == 0
    nat.record_tuning(path, "a", 1, ms)  # replaces the entry, keeps the others
    assert nat.lookup_tuning(path, "a")[0] == 1 and nat.lookup_tuning(path, "b")[0] == 0
    assert [f for f in os.listdir(tmp_path) if f.endswith(".tmp")] == []


def test_unfit_entries_are_ignored(tmp_path):
    path = str(tmp_path / "t.json")
    good = {"variant": 1, "ms": [1.0] * nat.LDS_VARIANTS, "variants": list(nat.LDS_VARIANT_NAMES)}
    table = {"ok": good, "range": dict(good, variant=nat.LDS_VARIANTS), "short": dict(good, ms=[1.0]),
             "renamed": dict(good, variants=["x"] * nat.LDS_VARIANTS), "junk": "x"}
    with open(path, "w") as fh:
        json.dump(table, fh)
    assert nat.lookup_tuning(path, "ok")[0] == 1
    for k in ("range", "short", "renamed", "junk", "missing"):
        assert nat.lookup_tuning(path, k) is None
    with open(path, "w") as fh:
        fh.write("{not json")
    assert nat.lookup_tuning(path, "ok") is None
    nat.record_tuning(path, "ok", 3, np.ones(nat.LDS_VARIANTS))  # an unreadable file is replaced
    assert nat.lookup_tuning(path, "ok")[0] == 3
