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
from robotecgpulidar_amd import scenes as S
from tests.golden import cases
from tests.helpers import compose64, run_oracle

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


def write_c4(meta: dict) -> None:
    """The 64-sensor batch: the raytrace node's geometry fields for cases.C4_BEAMS seeded beams of every sensor."""
    specs, sensors = S.config4_workload()
    worlds = {}
    parts = {k: [] for k in ("beams", "hit", "entity", "ambiguous", "f8", "f1", "f22")}
    offsets = [0]
    for k, x in enumerate(sensors):
        beams = cases.c4_beams(k, len(x.lidar.rays))
        if x.scene not in worlds:
            worlds[x.scene] = oracle.World(specs[x.scene])
        out = oracle.raytrace(specs[x.scene], compose64(x.lidar.pose, x.lidar.rays[beams]),
                              ray_origin_to_world=x.lidar.pose, ranges=[x.lidar.range], world=worlds[x.scene],
                              fields=cases.GEOMETRY_FIELDS, threads=8)
        parts["beams"].append(beams.astype(np.int32))
        parts["hit"].append(out[cases.IS_HIT].astype(np.uint8))
        parts["entity"].append(out[cases.ENT].astype(np.int32))
        parts["ambiguous"].append(out["ambiguous"].astype(np.int8))
        parts["f8"].append(out[cases.DIST])
        parts["f1"].append(out[cases.XYZ][:cases.C4_EXPLICIT])
        parts["f22"].append(out[cases.NORMAL][:cases.C4_EXPLICIT])
        offsets.append(offsets[-1] + len(beams))
    for w in worlds.values():
        w.close()
    arrays = {k: np.concatenate(v) for k, v in parts.items()}
    hits = int(arrays["hit"].sum())
    arrays["hit_bits"] = np.packbits(arrays.pop("hit"))
    arrays["offsets"] = np.array(offsets, np.int64)
    np.savez_compressed(HERE / "c4_sample.npz", **arrays)
    meta["c4_sample"] = {
        "description": "64-lidar batch over 8 urban scenes (scenes.config4_workload): raytrace-node XYZ, DISTANCE, "
                       "ENTITY_ID, NORMAL, IS_HIT of a seeded beam subset of every sensor (explicit XYZ and normal for "
                       "the first 256 beams of each)",
        "sensors": len(sensors), "beams": int(offsets[-1]), "hits": hits,
        "ambiguous_beams": int((arrays["ambiguous"] != 0).sum()),
    }
    print(f"c4_sample: {offsets[-1]} beams, {hits} hits, {meta['c4_sample']['ambiguous_beams']} ambiguous")


def write_c5_radar(meta: dict, count: int = 3) -> None:
    """Radar clients of the serving workload: the oracle's hits binned into cells (scenes.RADAR_CELL)."""
    specs, clients = S.config5_workload()
    radars = [(k, c) for k, c in enumerate(clients) if c.kind == "radar"]
    chosen, seen = [], set()
    for k, c in radars:
        if c.scene not in seen and len(chosen) < count:
            chosen.append((k, c))
            seen.add(c.scene)
    arrays, info = {}, []
    cell = S.RADAR_CELL
    for k, c in chosen:
        out = oracle.raytrace(specs[c.scene], compose64(c.lidar.pose, c.lidar.rays), ray_origin_to_world=c.lidar.pose,
                              ranges=[c.lidar.range], fields=[cases.XYZ, cases.DIST, cases.INC, cases.IS_HIT], threads=8)
        inv = np.linalg.inv(np.vstack([c.lidar.pose.astype(np.float64), [0, 0, 0, 1]]))
        hit = out[cases.IS_HIT]
        xyz_s = np.where(hit[:, None] == 1, out[cases.XYZ].astype(np.float64), 0.0) @ inv[:3, :3].T + inv[:3, 3]
        cells, counts, rcs_db, alt = oracle.radar_bin(xyz_s, out[cases.DIST], out[cases.INC], hit, cell["range_step"],
                                                      cell["range_max"], cell["azimuth_step"], cell["azimuth_min"],
                                                      cell["azimuth_max"], S.RADAR_SOLID_ANGLE)
        # Cells that a point within FP32 reach of a cell edge (or an ambiguous beam) may land in or leave.
        ac = int(np.ceil((np.float32(cell["azimuth_max"]) - np.float32(cell["azimuth_min"])) /
                         np.float32(cell["azimuth_step"])))
        d = out[cases.DIST].astype(np.float64)
        rpos = d / np.float32(cell["range_step"])
        apos = (np.arctan2(xyz_s[:, 1], xyz_s[:, 0]) - np.float32(cell["azimuth_min"])) / np.float32(cell["azimuth_step"])
        own = np.floor(np.where(hit == 1, rpos, 0)).astype(np.int64) * ac + np.floor(np.where(hit == 1, apos, 0)).astype(np.int64)
        touched = set()
        amb = out["ambiguous"] != 0
        for i in range(len(hit)):
            if alt[i] or (amb[i] and hit[i]):
                touched.add(int(own[i]))
                touched.update(alt[i])
        arrays[f"cells{k}"] = cells.astype(np.int64)
        arrays[f"counts{k}"] = counts.astype(np.int64)
        arrays[f"rcs{k}"] = rcs_db.astype(np.float64)
        arrays[f"touched{k}"] = np.array(sorted(touched), np.int64)
        info.append({"client": k, "scene": c.scene, "triangles": specs[c.scene].triangle_count,
                     "hits": int(hit.sum()), "cells": int(len(cells)), "edge_cells": len(touched)})
        print(f"c5_radar client {k}: scene {c.scene}, {int(hit.sum())} hits, {len(cells)} cells, {len(touched)} edge")
    np.savez_compressed(HERE / "c5_radar.npz", **arrays)
    meta["c5_radar"] = {"description": "radar clients of the serving workload (scenes.config5_workload), oracle hits "
                                       "binned into 0.2 m x 1 deg cells with RCS", "radars": info}


def main() -> None:
    oracle.load()
    meta = {}
    for name in cases.CASES:
        write_case(name, meta)
    write_c4(meta)
    write_c5_radar(meta)
    meta["noise"] = [{"seed": s, "run": r, "tag": t,
                      "bits": oracle.normals(s, r, t, 64).view(np.uint32).tolist()} for s, r, t in cases.NOISE_STREAMS]
    (HERE / "golden.json").write_text(json.dumps(meta, indent=1) + "\n")


if __name__ == "__main__":
    main()
