"""Writes the golden fixtures of tests/golden/ from the FP64 oracle (run from the repository root):

    python -m tests.golden.make_golden

For each case in cases.py: `<case>.npz` (hit mask bits and entity ids of every return point, ambiguity bits of
every beam, explicit field values for a seeded sample of return points) and an entry in `golden.json` (case
description, sample indices' seed, SHA-256 of every complete output array). Also the first 64 normal variates of
the noise streams in cases.NOISE_STREAMS, as float32 bit patterns.
"""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np

from oracle import oracle
from tests.golden import cases
from tests.helpers import run_oracle

HERE = Path(__file__).resolve().parent


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def sample_indices(n: int) -> np.ndarray:
    rng = np.random.default_rng(cases.SAMPLE_SEED)
    k = min(cases.SAMPLE_SIZE, n)
    return np.sort(rng.choice(n, k, replace=False)).astype(np.int32)


def write_case(name: str, meta: dict) -> None:
    case, fields = cases.CASES[name]()
    out = run_oracle(oracle, case, fields=fields)
    n_points = len(out[cases.IS_HIT])
    idx = sample_indices(n_points)
    arrays = {
        "hit_bits": np.packbits(out[cases.IS_HIT].astype(np.uint8)),
        "entity": out[cases.ENT].astype(np.int32),
        "ambiguous": out["ambiguous"].astype(np.int8),
        "sample": idx,
    }
    for f in fields:
        if f in (cases.IS_HIT, cases.ENT):
            continue
        arrays[f"f{f}"] = np.asarray(out[f])[idx]
    np.savez_compressed(HERE / f"{name}.npz", **arrays)
    meta[name] = {
        "description": " ".join(cases.CASES[name].__doc__.split()),
        "rays": int(len(case.rays)),
        "return_points": int(n_points),
        "hits": int(out[cases.IS_HIT].sum()),
        "fields": fields,
        "sha256": {str(f): digest(np.asarray(out[f])) for f in fields},
    }
    print(f"{name}: {n_points} points, {meta[name]['hits']} hits, {int((out['ambiguous'] != 0).sum())} ambiguous beams")


def main() -> None:
    oracle.load()
    meta = {}
    for name in cases.CASES:
        write_case(name, meta)
    meta["noise"] = [{"seed": s, "run": r, "tag": t,
                      "bits": oracle.normals(s, r, t, 64).view(np.uint32).tolist()} for s, r, t in cases.NOISE_STREAMS]
    (HERE / "golden.json").write_text(json.dumps(meta, indent=1) + "\n")


if __name__ == "__main__":
    main()
