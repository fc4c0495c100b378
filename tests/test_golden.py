"""The FP64 oracle reproduces the committed golden fixtures exactly (tests/golden/, written by make_golden.py from
the seeded cases in tests/golden/cases.py), so any change to the oracle's semantics shows up here on the CPU."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from tests.golden import cases
from tests.helpers import run_oracle

GOLDEN = Path(__file__).resolve().parent / "golden"
META = json.loads((GOLDEN / "golden.json").read_text())


@pytest.mark.parametrize("name", list(cases.CASES))
def test_oracle_reproduces_golden(oracle, name):
    case, fields = cases.CASES[name]()
    out = run_oracle(oracle, case, fields=fields)
    g = np.load(GOLDEN / f"{name}.npz")
    n = META[name]["return_points"]
    np.testing.assert_array_equal(np.unpackbits(g["hit_bits"])[:n], out[cases.IS_HIT])
    np.testing.assert_array_equal(g["entity"], out[cases.ENT])
    np.testing.assert_array_equal(g["ambiguous"], out["ambiguous"])
    for f in fields:
        assert hashlib.sha256(np.ascontiguousarray(out[f]).tobytes()).hexdigest() == META[name]["sha256"][str(f)], f
        if f"f{f}" in g:
            assert np.asarray(out[f])[g["sample"]].tobytes() == g[f"f{f}"].tobytes(), f


def test_oracle_noise_matches_golden(oracle):
    for s in META["noise"]:
        got = oracle.normals(s["seed"], s["run"], s["tag"], 64).view(np.uint32)
        np.testing.assert_array_equal(got, np.array(s["bits"], np.uint32))
