"""The HIP path against the committed golden fixtures (oracle outputs of tests/golden/cases.py), to the parity bar:
hit masks equal except on beams the oracle marks geometrically ambiguous (< 1e-5 of rays, at most one ray on small
cases), entity ids equal except on ambiguous beams, and on the sampled return points of unambiguous beams:
distance within max(1 mm, 1e-5 relative), XYZ within 2 mm, normals within 1e-4, intensity within 1e-4 relative
+ 2e-3. Noise variates are bit-identical to the committed ones."""
import json
from pathlib import Path

import numpy as np
import pytest

from tests.golden import cases
from tests.helpers import run_gpu

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
META = json.loads((GOLDEN / "golden.json").read_text())
GEOM, DIST_TIE, INT_TIE, TEXEL = 1, 2, 4, 8  # oracle ambiguity bits


def blocking_bits(return_type: int, quantity: str) -> int:
    """Ambiguity bits under which a return point's `quantity` may legitimately differ from FP64. A return selected
    by distance (FIRST = 4, LAST = 2) keeps its distance under near ties but not its sample, hence not its
    intensity; STRONGEST (1) keeps its intensity but not its sample; SECOND (3) and SECOND_STRONGEST (5) follow
    order-dependent strict comparisons and change on any near tie of their key. Texel edges change intensities."""
    by_distance = return_type in (2, 3, 4)
    if quantity == "distance":
        if return_type in (2, 4):
            return GEOM
        if return_type == 3:
            return GEOM | DIST_TIE
        return GEOM | TEXEL | INT_TIE
    if by_distance:
        return GEOM | DIST_TIE | TEXEL
    return GEOM | TEXEL | (INT_TIE if return_type == 5 else 0)


@pytest.mark.parametrize("name", list(cases.CASES))
def test_gpu_matches_golden(rgl_lib, name):
    case, fields = cases.CASES[name]()
    gpu = run_gpu(rgl_lib, case, fields)
    g = np.load(GOLDEN / f"{name}.npz")
    n = META[name]["return_points"]
    beams = len(case.rays)
    R = n // beams
    amb = np.repeat(g["ambiguous"], R)
    hit = np.unpackbits(g["hit_bits"])[:n].astype(np.int32)
    assert len(gpu[cases.IS_HIT]) == n
    hit_mism = gpu[cases.IS_HIT] != hit
    assert not (hit_mism & ((amb & GEOM) == 0)).any(), np.nonzero(hit_mism & ((amb & GEOM) == 0))[0][:10]
    if R == 1:
        assert hit_mism.sum() <= max(1, int(1e-5 * n))
    ent_mism = (gpu[cases.IS_HIT] == 1) & (hit == 1) & (gpu[cases.ENT] != g["entity"])
    assert not (ent_mism & (amb == 0)).any(), np.nonzero(ent_mism & (amb == 0))[0][:10]

    idx = g["sample"]
    types = (case.return_mode >> (8 * (idx % R))) & 0xFF
    both = (hit[idx] == 1) & (gpu[cases.IS_HIT][idx] == 1)
    block_d = np.array([blocking_bits(int(t), "distance") for t in types])
    block_i = np.array([blocking_bits(int(t), "intensity") for t in types])
    sel = both & ((amb[idx] & block_d) == 0)
    sel_i = both & ((amb[idx] & block_i) == 0)
    print(f"{name}: {int(hit.sum())} golden hits, {int(hit_mism.sum())} mask differences, "
          f"{int((amb[::R] != 0).sum())} ambiguous beams, {int(sel.sum())} sampled points compared")
    assert sel.sum() >= (min(100, len(idx) // 4) if R == 1 else 50)
    d_o = g[f"f{cases.DIST}"][sel].astype(np.float64)
    d_g = gpu[cases.DIST][idx][sel].astype(np.float64)
    assert np.all(np.abs(d_g - d_o) <= np.maximum(1e-3, 1e-5 * d_o)), np.abs(d_g - d_o).max()
    exact = both & (amb[idx] == 0)  # unambiguous beams: the same sample is chosen
    if cases.XYZ in fields:
        assert np.abs(gpu[cases.XYZ][idx][exact] - g[f"f{cases.XYZ}"][exact]).max() <= 2e-3
    if cases.NORMAL in fields:
        assert np.abs(gpu[cases.NORMAL][idx][exact] - g[f"f{cases.NORMAL}"][exact]).max() <= 1e-4
    if cases.INTEN in fields:
        i_o = g[f"f{cases.INTEN}"][sel_i].astype(np.float64)
        i_g = gpu[cases.INTEN][idx][sel_i].astype(np.float64)
        bad = np.abs(i_g - i_o) > 1e-4 * i_o + 2e-3
        assert not bad.any(), (idx[sel_i][bad], i_g[bad], i_o[bad], amb[idx][sel_i][bad])
    non_hit = (hit[idx] == 0) & (gpu[cases.IS_HIT][idx] == 0)
    np.testing.assert_array_equal(gpu[cases.DIST][idx][non_hit], g[f"f{cases.DIST}"][non_hit])


def test_gpu_noise_matches_golden(rgl_lib):
    for s in META["noise"]:
        got = rgl_lib.noise_normals(s["seed"], s["run"], s["tag"], 64).view(np.uint32)
        np.testing.assert_array_equal(got, np.array(s["bits"], np.uint32))
