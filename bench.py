"""Benchmark: served-batched vs plain lidar frames through the RGL graph on one MI355X.

Default workload (`--workload batch64`), one step = one frame of 64 lidars, all inputs resident in HBM:
64 lidars of uneven size (16-128 channels x 900-4096 azimuths, 10.1M rays in total) over 8 device-resident urban
scenes of 76k-3.6M triangles (8.0M in total). Each sensor's graph is
rays -> pose -> range -> raytrace -> points transform (to the sensor frame) -> distance noise -> compact -> yield.
`value` is Mrays/s (API beams completed through compaction, results readable on the host) with every frame run as
ONE batch (rgl_graph_run_batch); `plain` is the same frames with one rgl_graph_run per sensor (each graph on its
own stream). p50/p99 are frame latencies (submission until every sensor's kept count is readable).

`--workload urban`: one 128 x 2048 lidar frame on the ~2M-triangle urban scene with angular and distance noise and
velocity distortion (rays -> pose -> range -> time offsets -> angular noise -> raytrace -> distance noise ->
compact -> yield), run plain.

`roofline` is computed for the (batched) raytrace kernel from HIP-event durations on the stream it runs on: the
bytes it must move (ray poses, output fields, and the unique BVH and scene records its rays touch) over its time,
against 8 TB/s HBM. `accuracy` compares the benchmark frame's own hit distances with the FP64 oracle at 65,536
committed beams; the urban and multi-return lines carry the same figures for their configurations from the committed
golden fixtures (C2 / C1 and C3 at reduced ray counts). `cpu_baseline` times the FP32 CPU port of the same graph (oracle/cpu_port.cpp) on whole 64-lidar
frames on the host's CPU share.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from robotecgpulidar_amd import rgl  # noqa: E402
from robotecgpulidar_amd import scenes as S  # noqa: E402
from robotecgpulidar_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s + p50/p99 frame latency, served-batched vs plain, 1 MI355X"
FIELDS = W.FIELDS
GOLDEN = ROOT / "tests" / "golden"


kept_count = W.kept_count


def frame(entry, out) -> int:  # one plain frame
    rgl.check("rgl_graph_run", entry.h)
    return kept_count(out)


def canonical_bytes(rt, n_rays: int):
    """BVH traffic of one raytrace launch counted per ray on the canonical binary trees (64 B per node visit, 36 B per
    triangle test, 64 B per instance entry, every fetch of every ray counted): the survey's per-ray figure. Caches
    serve most of it, so bytes / time is an effective bandwidth, not DRAM traffic."""
    nodes = np.empty(n_rays, np.int32)
    tris = np.empty(n_rays, np.int32)
    insts = np.empty(n_rays, np.int32)
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
    rgl.check("rgl_node_raytrace_traversal_stats", rt.h, p(nodes), p(tris), p(insts), n_rays)
    sums = (int(nodes.sum(dtype=np.int64)), int(tris.sum(dtype=np.int64)), int(insts.sum(dtype=np.int64)))
    return 64 * sums[0] + 36 * sums[1] + 64 * sums[2], sums


def cpu_share() -> tuple[int, int]:
    """(threads used, host CPUs visible). The GPU box gives one GPU a share of 16 host cores whatever nproc shows;
    the CPU baseline runs on min(CPUs this process may use, 16) threads."""
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return min(visible, 16), os.cpu_count() or visible


def inv34(m):
    return np.linalg.inv(np.vstack([np.asarray(m, np.float64), [0, 0, 0, 1]]))[:3]


def cpu_baseline(specs, sensors, budget_s: float, label: str):
    """The FP32 CPU port of the same graph (oracle/cpu_port.cpp: canonical LBVH built on the CPU, FP32 two-level
    traversal, FP64 hit distances, points transform, the same seeded distance noise, stable compaction) over whole
    frames of every sensor, 64-ray tiles on std::threads; BVH build excluded. Up to 3 warm-up frames, then frames
    until `budget_s` (at least one, at most 20); the median frame time gives the rate."""
    from oracle import oracle  # the CPU port is the timed CPU baseline only here
    threads, visible = cpu_share()
    port = oracle.CpuPort(specs)
    sens = []
    for k, x in enumerate(sensors):
        to_sensor = inv34(x.lidar.pose)
        look = inv34(to_sensor) @ np.vstack([inv34(x.lidar.pose), [0, 0, 0, 1]])
        sens.append(dict(scene=x.scene, rays=x.lidar.rays, pose=x.lidar.pose, range=x.lidar.range,
                         to_sensor=S.f32(to_sensor), look_at=S.f32(look), look_at_inv=S.f32(inv34(look)),
                         noise=(0.0, 0.02, 0.001), seed=0xC0F40000 + k, run=0, tag=3))
    frame = port.prepare(sens)
    rays = sum(len(x.lidar.rays) for x in sensors)
    times = []
    t_start = time.perf_counter()
    for i in range(23):
        t0 = time.perf_counter()
        port.execute(frame, threads)
        el = time.perf_counter() - t0
        if i >= 3 or (time.perf_counter() - t_start) > budget_s / 2:
            times.append(el)
        if times and time.perf_counter() - t_start > budget_s:
            break
    kept = sum(r["kept"] for r in port.results(frame))
    port.close()
    med = float(np.median(times))
    return {"value": round(rays / med / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "host_cpus_visible": visible, "frame_p50_ms": round(med * 1e3, 2),
            "frame_p99_ms": round(float(np.percentile(times, 99)) * 1e3, 2), "kept_points_per_step": kept,
            "sample": f"{len(times)} whole frames of {label} ({rays} rays each) after warm-up, FP32 CPU port of the "
                      f"same graph on {threads} threads, BVH build excluded"}


def latency_stats(lat_s):
    ms = np.array(lat_s) * 1e3
    return round(float(np.percentile(ms, 50)), 4), round(float(np.percentile(ms, 99)), 4)


def roofline(b_io: int, footprint: int, b_canonical: int, kernel_ms: float, profile):
    """HBM roofline of the raytrace launch. `achieved` counts the bytes the launch must move at least once: ray poses
    in and every field of every return out (b_io) plus the unique bytes of the BVH nodes, leaf triangles, instance
    records and shading records its rays touch (footprint, each 64 B line once, rgl_graph_traversal_footprint).
    `traffic` is the DRAM bytes rocprof counted for the same kernel (FETCH_SIZE + WRITE_SIZE, scaled by a copy
    kernel of known size); `effective_gbs` is the canonical per-ray figure (every node and triangle fetch of every
    ray counted) over the same time; `limiter` the traversal's actual limiter from the committed counter run: the
    fraction of the launch the vector memory data return path (TD) is busy, most of it waiting on L1/L2 returns of
    divergent node fetches."""
    t = kernel_ms * 1e-3
    b_alg = b_io + footprint
    achieved = b_alg / t / 1e9
    traffic = profile.get("bytes_per_launch") if profile else None
    out = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "algorithmic_bytes_per_launch": b_alg,
           "io_bytes_per_launch": b_io, "footprint_bytes_per_launch": footprint,
           "effective_gbs": round((b_io + b_canonical) / t / 1e9, 2), "canonical_bytes_per_launch": b_io + b_canonical}
    if traffic:
        out["dram_frac"] = round(traffic / t / 1e9 / HBM_PEAK_GBS, 4)
    if profile and "td_busy" in profile:
        out["limiter"] = {"unit": "vector memory data path (TD) busy", "frac": profile["td_busy"],
                          "td_stalled_on_cache": profile["td_tc_stall"], "ta_busy": profile["ta_busy"],
                          "vector_loads_per_wave": profile["vector_loads_per_wave"],
                          "scalar_loads_per_wave": profile["scalar_loads_per_wave"],
                          "valu_instructions_per_wave": profile["valu_instructions_per_wave"],
                          "source": profile["source"]}
    return out


def kernel_profile(name: str):
    tf = ROOT / "profiles" / name
    return json.loads(tf.read_text()) if tf.exists() else None


def golden_accuracy(names):
    """Accuracy of a workload's configuration against the FP64 oracle, from the committed golden fixtures of the same
    configuration at reduced ray count (tests/golden/cases.py): each case runs once through the library in a scene
    of its own, and tests.helpers.check_golden reports hit-mask differences (all on beams the oracle marks
    geometrically ambiguous), the distances compared and the largest |d - d_FP64|; a case that misses the parity bar
    reports the failed check instead of numbers."""
    sys.path.insert(0, str(ROOT))
    from tests.golden import cases  # noqa: E402  (fixture definitions; no oracle code)
    from tests.helpers import Graph, check_golden  # noqa: E402
    meta = json.loads((GOLDEN / "golden.json").read_text())
    out = {}
    for name in names:
        rgl.cleanup()
        rgl.check("rgl_configure_logging", 6, None, False)
        case, fields = cases.CASES[name]()
        S.load_into_rgl(case.spec)
        gpu = Graph(rgl, case, fields).run()
        g = np.load(GOLDEN / f"{name}.npz")
        hit = np.unpackbits(g["hit_bits"])[:meta[name]["return_points"]].astype(np.int32)
        values = {int(k[1:]): g[k] for k in g.files if k.startswith("f")}
        try:
            stats = check_golden(gpu, hit, g["entity"], g["ambiguous"], g["sample"], values, case.return_mode, name,
                                 min_compared=0)
        except AssertionError as e:
            stats = {"parity": f"failed: {e}"}
        out[name] = {"rays": int(len(case.rays)), **stats, "reference": f"tests/golden/{name}.npz (FP64 oracle)"}
    rgl.cleanup()
    return out


def c4_accuracy(graphs):
    """The benchmark's own raytrace-node results at the 65,536 seeded beams of the committed FP64-oracle fixture
    (tests/golden/c4_sample.npz, 1,024 beams of each sensor): hit-mask and entity differences, beams the oracle marks
    geometrically ambiguous, and the largest |d - d_FP64| over the other hits."""
    g = np.load(GOLDEN / "c4_sample.npz")
    off = g["offsets"]
    hit = np.unpackbits(g["hit_bits"])[:off[-1]].astype(np.int32)
    mism = mism_unexplained = ent = ent_unexplained = compared = 0
    worst = 0.0
    for k, sg in enumerate(graphs):
        a, b = int(off[k]), int(off[k + 1])
        beams = g["beams"][a:b]
        gh = rgl.result(sg.raytrace, rgl.IS_HIT_I32)[beams]
        gd = rgl.result(sg.raytrace, rgl.DISTANCE_F32)[beams].astype(np.float64)
        ge = rgl.result(sg.raytrace, rgl.ENTITY_ID_I32)[beams]
        amb = (g["ambiguous"][a:b] & 1) != 0
        m = gh != hit[a:b]
        mism += int(m.sum())
        mism_unexplained += int((m & ~amb).sum())
        both = (gh == 1) & (hit[a:b] == 1)
        e = both & (ge != g["entity"][a:b])
        ent += int(e.sum())
        ent_unexplained += int((e & ~amb).sum())
        sel = both & ~amb
        compared += int(sel.sum())
        if sel.any():
            worst = max(worst, float(np.abs(gd[sel] - g["f8"][a:b][sel]).max()))
    amb_all = g["ambiguous"]
    return {"reference": "FP64 oracle, 1,024 seeded beams of each sensor (tests/golden/c4_sample.npz)",
            "beams": int(off[-1]), "hit_mask_mismatches": mism, "hit_mask_mismatches_unexplained": mism_unexplained,
            "entity_mismatches": ent, "entity_mismatches_unexplained": ent_unexplained,
            "ambiguous_beams": int(((amb_all & 1) != 0).sum()),
            "instance_tie_beams": int(((amb_all & 16) != 0).sum()),
            "distances_compared": compared, "max_dist_err_m": worst,
            "note": "ambiguous beams graze an edge or tie a hit of another instance within 1e-5 relative in FP64; "
                    "instance ties here are mostly exactly coplanar faces (a building's floor on the ground seen "
                    "from a sensor placed inside the building), which FP32 and FP64 round apart"}


def run_batch64(args):
    specs, sensors = S.config4_workload()
    scenes = [sc for sc, _ in W.load_scenes(specs)]
    graphs = [W.sensor_graph(x.lidar, scenes[x.scene], 0xC0F40000 + k) for k, x in enumerate(sensors)]
    entries = [g.rays for g in graphs]
    outs = [g.out for g in graphs]
    rays_per_frame = sum(len(x.lidar.rays) for x in sensors)

    counts = rgl.ResultSizes(outs, rgl.XYZ_VEC3_F32)  # every sensor's kept count in one call, in both modes
    run_batch = rgl.BatchRun(entries)

    def plain_frame():
        for e in entries:
            rgl.check("rgl_graph_run", e.h)
        return int(counts().sum())

    submit_s = []

    def batched_frame():
        t0 = time.perf_counter()
        run_batch()
        submit_s.append(time.perf_counter() - t0)
        return int(counts().sum())

    def timed(fn, steps):
        lat = []
        t0 = time.perf_counter()
        for _ in range(steps):
            f0 = time.perf_counter()
            kept = fn()
            lat.append(time.perf_counter() - f0)
        return time.perf_counter() - t0, lat, kept

    for _ in range(args.warmup):
        plain_frame()
        batched_frame()
    p_total, p_lat, p_kept = timed(plain_frame, args.steps)
    # Each timed leg follows two untimed frames of its own mode, so the switch from per-sensor runs to batches
    # (stream and buffer reuse patterns) is not charged to the first timed batch.
    batched_frame()
    batched_frame()
    caps0 = rgl.batch_capture()
    submit_s.clear()
    b_total, b_lat, b_kept = timed(batched_frame, args.steps)
    submit_p50 = latency_stats(submit_s)[0]
    caps1 = rgl.batch_capture()
    # The same batches replayed from captured HIP graphs (capture is off by default), for comparison.
    rgl.batch_capture(True)
    batched_frame()
    batched_frame()
    e_total, e_lat, _ = timed(batched_frame, args.steps)
    caps2 = rgl.batch_capture()
    rgl.batch_capture(False)
    # Kernel durations from HIP events, in separate frames so the timed frames carry no extra events.
    kframes = 3
    f0 = rgl.batch_fused_stages()
    batched_frame()
    fused_per_frame = rgl.batch_fused_stages() - f0
    for g in graphs:
        rgl.check("rgl_node_raytrace_set_kernel_timing", g.raytrace.h, True)
    timed(plain_frame, kframes)
    p_kernel = 0.0
    for g in graphs:
        ms, n = C.c_double(), C.c_int32()
        rgl.check("rgl_node_raytrace_get_kernel_timing", g.raytrace.h, C.byref(ms), C.byref(n))
        rgl.check("rgl_node_raytrace_set_kernel_timing", g.raytrace.h, False)
        p_kernel += ms.value / max(n.value, 1)
    rgl.batch_kernel_timing(True)
    timed(batched_frame, kframes)
    kms, launches = rgl.batch_kernel_timing()
    rgl.batch_kernel_timing(False)
    kernel_ms = kms / kframes  # batched raytrace time per frame
    launches_per_frame = launches / kframes
    # The benchmark frame's own hit distances against the FP64 oracle (results of the last batched frame).
    accuracy = c4_accuracy(graphs)

    # End to end with the points on the host: every frame also reads every yielded field of every sensor into host
    # arrays, either by copying from device memory at query time or from pinned memory the run's own kernel wrote.
    readback = {}
    bufs = [{f: np.empty((len(x.lidar.rays),) + ((3,) if f in (rgl.XYZ_VEC3_F32, rgl.NORMAL_VEC3_F32) else ()),
                         dtype=rgl.FIELD_DTYPE[f][0]) for f in FIELDS} for x in sensors]

    def read_all():
        total = 0
        for o, b in zip(outs, bufs):
            n = kept_count(o)
            for f in FIELDS:
                if n:
                    rgl.check("rgl_graph_get_result_data", o.h, int(f), b[f].ctypes.data)
                total += n * rgl.field_size(f)
        return total

    for name, prefetch in (("device_copy", False), ("prefetch", True)):
        for o in outs:
            rgl.check("rgl_node_points_yield_configure_prefetch", o.h, prefetch)
        for _ in range(2):
            rgl.graph_run_batch(entries)
            read_all()
        lat = []
        t0 = time.perf_counter()
        moved = 0
        for _ in range(args.steps):
            f0 = time.perf_counter()
            rgl.graph_run_batch(entries)
            moved = read_all()
            lat.append(time.perf_counter() - f0)
        tot = time.perf_counter() - t0
        r50, r99 = latency_stats(lat)
        readback[name] = {"mrays_per_s": round(args.steps * rays_per_frame / tot / 1e6, 3), "p50_ms": r50,
                          "p99_ms": r99, "bytes_to_host_per_step": moved}
    for o in outs:
        rgl.check("rgl_node_points_yield_configure_prefetch", o.h, False)

    rtv, cpv = C.c_int32(), C.c_int32()
    rgl.check("rgl_graph_batch_get_variants", C.byref(rtv), C.byref(cpv))
    tuned = {"raytrace_block_threads": {0: 128, 1: 64}.get(rtv.value, "tuning"),
             "compaction_points_per_thread": {0: 8, 1: 4}.get(cpv.value, "tuning")}

    point_bytes = sum(rgl.field_size(f) for f in FIELDS)
    b_trav = 0
    sums = np.zeros(3, np.int64)
    for g, x in zip(graphs, sensors):
        trav, s3 = canonical_bytes(g.raytrace, len(x.lidar.rays))
        b_trav += trav
        sums += np.array(s3, np.int64)
    footprint = rgl.traversal_footprint([g.raytrace for g in graphs])
    # Per ray the batched launch reads its 48 B pose and writes the requested fields (36 B), plus the folded points
    # transform (XYZ, 12 B) and distance noise (XYZ and DISTANCE, 16 B); the XYZ and DISTANCE those stages start from
    # are the launch's own writes.
    assert fused_per_frame == 2 * len(graphs), fused_per_frame
    b_io_batched = rays_per_frame * (48 + point_bytes + 12 + 16)
    value = args.steps * rays_per_frame / b_total / 1e6
    plain_value = args.steps * rays_per_frame / p_total / 1e6
    p50, p99 = latency_stats(b_lat)
    pp50, pp99 = latency_stats(p_lat)
    cpu = None
    if args.cpu_budget > 0:
        cpu = cpu_baseline(specs, sensors, args.cpu_budget, "the 64-lidar batch")
    return {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(b_total / args.steps * 1e3, 4),
        "p50_ms": p50,
        "p99_ms": p99,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "64-lidar mixed batch over 8 urban scenes, every sensor frame in one rgl_graph_run_batch "
                        "(rays->pose->range->raytrace->points transform->distance noise->compact->yield)",
            "sensors": len(sensors), "scenes": len(specs), "triangles": sum(s.triangle_count for s in specs),
            "scene_triangles": [s.triangle_count for s in specs], "rays_per_step": rays_per_frame,
            "kept_points_per_step": b_kept, "mode": "batched", "autotuned": tuned,
        },
        "plain": {"value": round(plain_value, 3), "ms_per_step": round(p_total / args.steps * 1e3, 4),
                  "p50_ms": pp50, "p99_ms": pp99, "raytrace_kernel_ms_per_step": round(p_kernel, 4),
                  "kept_points_per_step": p_kept, "mode": "one rgl_graph_run per sensor"},
        "batched_vs_plain": round(value / plain_value, 3),
        "host_submit_p50_ms": submit_p50,
        "graph_replays_per_step": round((caps1[1] - caps0[1]) / args.steps, 3),
        "batched_captured": {"value": round(args.steps * rays_per_frame / e_total / 1e6, 3),
                             "ms_per_step": round(e_total / args.steps * 1e3, 4),
                             "p50_ms": latency_stats(e_lat)[0], "p99_ms": latency_stats(e_lat)[1],
                             "graph_replays_per_step": round((caps2[1] - caps1[1]) / args.steps, 3),
                             "mode": "same batches replayed from captured HIP graphs (rgl_graph_batch_configure_capture)"},
        "with_readback": readback,
        "raytrace_kernel_ms": round(kernel_ms, 4),
        "point_stages_folded_per_step": fused_per_frame,
        "raytrace_launches_per_step": launches_per_frame,
        "traversal_per_ray": {"nodes": round(sums[0] / rays_per_frame, 2), "triangles": round(sums[1] / rays_per_frame, 2),
                              "instances": round(sums[2] / rays_per_frame, 2)},
        "accuracy": accuracy,
        "roofline": roofline(b_io_batched, footprint, b_trav, kernel_ms, kernel_profile("raytrace_batch64_traffic.json")),
        "cpu_baseline": cpu,
    }


def poisson_trace(rate_hz: float, duration_s: float, clients: int, seed: int):
    rng = np.random.default_rng(seed)
    n = int(rate_hz * duration_s * 1.5) + 16
    t = np.cumsum(rng.exponential(1.0 / rate_hz, n))
    t = t[t < duration_s]
    return t, rng.integers(0, clients, len(t))


def serve_open_loop(entries, rays, times, who, batching: bool, slots: int, budget: int):
    """Replays the arrival trace in real time through a server; returns its stats plus wall time and Mrays/s."""
    server = rgl.Server(slots=slots, max_batch_graphs=len(entries), hbm_budget=budget, batching=batching)
    k, n = 0, len(times)
    t0 = time.perf_counter()
    while k < n:
        now = time.perf_counter() - t0
        while k < n and times[k] <= now:
            server.submit(entries[who[k]])
            k += 1
        server.step()
    while server.stats()["completed"] < n:
        server.step()
    wall = time.perf_counter() - t0
    st = server.stats()
    server.destroy()
    st["wall_s"] = wall
    st["mrays_per_s"] = st["rays_completed"] / wall / 1e6
    return st


def serve_closed_loop(entries, batching: bool, slots: int, budget: int, duration_s: float):
    """Every client resubmits as soon as its previous frame is done: the server's saturated throughput."""
    server = rgl.Server(slots=slots, max_batch_graphs=len(entries), hbm_budget=budget, batching=batching)
    tickets = [server.submit(e) for e in entries]
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < duration_s:
        if server.step() == 0:
            continue  # nothing finished: no ticket changed
        for i, t in enumerate(tickets):
            if server.poll(t)[0] == rgl.Server.DONE:
                tickets[i] = server.submit(entries[i])
    while server.stats()["completed"] < server.stats()["submitted"]:
        server.step()
    wall = time.perf_counter() - t0
    st = server.stats()
    server.destroy()
    st["wall_s"] = wall
    st["mrays_per_s"] = st["rays_completed"] / wall / 1e6
    return st


def run_serve(args, rate_hz: float = 2000.0, duration_s: float = 2.0, budget: int = 32 << 30):
    specs, clients = S.config5_workload()
    scenes = []
    for spec in specs:
        sc = rgl.scene_create()
        S.load_into_rgl(spec, scene=sc)
        scenes.append(sc)
    graphs = []
    for k, c in enumerate(clients):
        if c.kind == "radar":
            graphs.append(W.radar_bin_graph(c.lidar, scenes[c.scene]))
        else:
            graphs.append(W.sensor_graph(c.lidar, scenes[c.scene], 0xC0F50000 + k))
    entries = [g.rays for g in graphs]
    rays = [len(c.lidar.rays) for c in clients]
    for e in entries:  # builds every scene's structures and every graph's buffers
        rgl.check("rgl_graph_run", e.h)
    rgl.batch_kernel_timing(False)
    times, who = poisson_trace(rate_hz, duration_s, len(clients), 0xC0F5)
    keys = ("completed", "batches", "rays_completed", "evictions", "deferred", "resident_bytes")
    out = {"arrival_rate_hz": rate_hz, "duration_s": duration_s, "requests": int(len(times)),
           "clients": len(clients), "lidars": sum(c.kind == "lidar" for c in clients),
           "radars": sum(c.kind == "radar" for c in clients), "scenes": len(specs),
           "scene_triangles": [s.triangle_count for s in specs], "hbm_budget_bytes": budget,
           "offered_mrays_per_s": round(sum(rays[w] for w in who) / duration_s / 1e6, 3)}
    for name, batching, slots in (("batched", True, 3), ("plain", False, len(clients))):
        st = serve_open_loop(entries, rays, times, who, batching, slots, budget)
        sat = serve_closed_loop(entries, batching, slots, budget, 1.0)
        out[name] = {"p50_ms": round(st["p50_ms"], 4), "p99_ms": round(st["p99_ms"], 4),
                     "max_ms": round(st["max_ms"], 4), "mrays_per_s": round(st["mrays_per_s"], 3),
                     "saturated_mrays_per_s": round(sat["mrays_per_s"], 3),
                     "saturated_p99_ms": round(sat["p99_ms"], 4), "slots": slots,
                     **{k: int(st[k]) for k in keys}}
    out["saturated_batched_vs_plain"] = round(out["batched"]["saturated_mrays_per_s"] /
                                              out["plain"]["saturated_mrays_per_s"], 3)
    return out


def run_multireturn(args, moving_fraction: float = 0.2, divergence: float = 0.13e-3):
    """C3: ~10M triangles in 2,001 instances with a 1024 x 1024 intensity texture; a 64 x 4096 lidar with 0.13 mrad
    divergence (58 samples per beam, the reference's beam model) traced twice per frame: FIRST_LAST and STRONGEST
    (the reference allows at most two returns per node), both graphs in one batch. Before every frame 20% of the
    instances move by U[0, 0.5] m; the TLAS is refitted or rebuilt."""
    spec, movable = S.config3_scene()
    lidar = S.config3_lidar()
    _, ents, _ = S.load_into_rgl(spec)
    graphs = [W.multireturn_graph(lidar, rgl.RGL_RETURN_FIRST_LAST, divergence),
              W.multireturn_graph(lidar, rgl.RGL_RETURN_STRONGEST, divergence)]
    entries = [g.rays for g in graphs]
    beams = 2 * len(lidar.rays)
    base = [spec.instances[i].transform.astype(np.float64) for i in movable]
    rng = np.random.default_rng(0xC0F3)
    pos = None

    def move():
        nonlocal pos
        pick = rng.choice(len(movable), int(moving_fraction * len(movable)), replace=False)
        step = rng.uniform(0.0, 0.5, len(pick))
        ang = rng.uniform(0.0, 2 * np.pi, len(pick))
        pos[pick, 0] += step * np.cos(ang)
        pos[pick, 1] += step * np.sin(ang)
        for i in pick:
            m = base[i].copy()
            m[:2, 3] = pos[i]
            rgl.entity_set_transform(ents[movable[i]], S.f32(m))

    def frame():
        rgl.graph_run_batch(entries)
        return sum(kept_count(g.out) for g in graphs)

    out = {"triangles": spec.triangle_count, "instances": len(spec.instances), "beams_per_step": beams,
           "samples_per_beam": 58, "divergence_rad": divergence, "moving_fraction": moving_fraction,
           "returns": "FIRST_LAST + STRONGEST (two graphs, one batch)", "texture": "1024 x 1024 uint8"}
    for mode in ("rebuild", "refit"):
        rgl.check("rgl_scene_set_tlas_refit", None, mode == "refit")
        pos = np.array([b[:2, 3] for b in base])
        for i, b in enumerate(base):
            rgl.entity_set_transform(ents[movable[i]], S.f32(b))
        for _ in range(args.warmup):
            move()
            frame()
        lat, upd = [], []
        kept = 0
        rgl.batch_kernel_timing(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            u0 = time.perf_counter()
            move()
            f0 = time.perf_counter()
            kept = frame()
            lat.append(time.perf_counter() - f0)
            upd.append(f0 - u0)
        total = time.perf_counter() - t0
        kms, launches = rgl.batch_kernel_timing()
        rgl.batch_kernel_timing(False)
        p50, p99 = latency_stats(lat)
        frame_s = sum(lat) / len(lat)
        out[mode] = {"mbeams_per_s": round(beams / frame_s / 1e6, 3),
                     "traced_mrays_per_s": round(58 * beams / frame_s / 1e6, 3),
                     "p50_ms": p50, "p99_ms": p99, "raytrace_kernel_ms": round(kms / args.steps, 4),
                     "host_pose_update_ms": round(1e3 * sum(upd) / len(upd), 4),
                     "wall_ms_per_step": round(total / args.steps * 1e3, 4), "kept_points_per_step": kept}
    return out


RADAR_SCOPES = [(0.0, 50.0, 0.5, 0.5, math.radians(1.0)), (50.0, 200.0, 1.0, 1.0, math.radians(1.5))]


def build_radar_post_graph(lidar: S.Lidar, scene, seed: int):
    """rays -> pose -> range -> raytrace(scene, sensor moving at 10 m/s) -> compact -> radar postprocess (the
    reference's clustering, RCS and link budget; 77 GHz, 31 dBm, 27 dBi) -> yield."""
    n = rgl.Node
    rays = rgl.node_rays_from_mat3x4f(n(), lidar.rays)
    pose = rgl.node_rays_transform(n(), lidar.pose)
    rng = rgl.node_rays_set_range(n(), np.array([lidar.range], np.float32))
    rt = rgl.node_raytrace(n(), scene)
    lin, ang = np.float32([10.0, 0.0, 0.0]), np.zeros(3, np.float32)
    rgl.check("rgl_node_raytrace_configure_velocity", rt.h, lin.ctypes.data, ang.ctypes.data)
    comp = rgl.node_points_compact_by_field(n(), rgl.IS_HIT_I32)
    radar = rgl.node_points_radar_postprocess(n(), RADAR_SCOPES, math.radians(120.0 / 256), math.radians(20.0 / 63),
                                              77e9, 31.0, 27.0, -90.0, 1.0)
    rgl.check("rgl_node_gaussian_noise_set_seed", radar.h, C.c_uint64(seed))
    out = rgl.node_points_yield(n(), [rgl.XYZ_VEC3_F32, rgl.DISTANCE_F32, rgl.RCS_F32, rgl.POWER_F32, rgl.SNR_F32])
    rgl.chain(rays, pose, rng, rt, comp, radar, out)
    return rays, comp, out


def run_radar(args, radars: int = 8):
    """Radar postprocessing with the reference's semantics: 8 radars of 256 x 64 rays (+-60 deg x +-10 deg) on the
    urban scene, moving entities so radial speeds are defined, one frame of all radars per step."""
    spec = S.config2_scene()
    sc = rgl.scene_create()
    _, ents, _ = S.load_into_rgl(spec, scene=sc)
    rng = np.random.default_rng(0xC0F5)
    sensors = []
    for k in range(radars):
        yaw = 360.0 * k / radars
        pose = S.f32(S.trs((rng.uniform(-40, 40), rng.uniform(-40, 40), 1.2), (0.0, 0.0, yaw)))
        sensors.append(S.Lidar(S.radar_rays(256, 64), pose, (0.5, 200.0)))
    graphs = [build_radar_post_graph(x, sc, 0xAADA0 + k) for k, x in enumerate(sensors)]
    entries = [g[0] for g in graphs]
    movers = list(range(1, len(ents), 7))
    t_ns = 0

    def move():
        nonlocal t_ns
        t_ns += 50_000_000
        rgl.check("rgl_scene_set_time", sc, C.c_uint64(t_ns))
        for j in movers:
            m = spec.instances[j].transform.astype(np.float64).copy()
            m[0, 3] += 0.5 * (t_ns / 50_000_000) * (1 if j % 2 else -1)
            rgl.entity_set_transform(ents[j], S.f32(m))

    def frame(batched):
        if batched:
            rgl.graph_run_batch(entries)
        else:
            for e in entries:
                rgl.check("rgl_graph_run", e.h)
        clusters = sum(kept_count(g[2]) for g in graphs)
        detections = sum(kept_count(g[1]) for g in graphs)
        return clusters, detections

    out = {"radars": radars, "rays_per_radar": 256 * 64, "triangles": spec.triangle_count,
           "scopes": [list(map(float, s)) for s in RADAR_SCOPES]}
    for mode in ("batched", "plain"):
        for _ in range(args.warmup):
            move()
            frame(mode == "batched")
        lat = []
        for _ in range(args.steps):
            move()
            f0 = time.perf_counter()
            clusters, detections = frame(mode == "batched")
            lat.append(time.perf_counter() - f0)
        p50, p99 = latency_stats(lat)
        out[mode] = {"p50_ms": p50, "p99_ms": p99, "frames_per_s": round(len(lat) / sum(lat), 2),
                     "mrays_per_s": round(radars * 256 * 64 * len(lat) / sum(lat) / 1e6, 3),
                     "detections_per_step": detections, "clusters_per_step": clusters}
    return out


def run_urban(args):
    spec, lidar = S.config2_scene(), S.config2_lidar()
    S.load_into_rgl(spec)
    g = W.urban_graph(lidar)
    entry, rt, out = g.rays, g.raytrace, g.out
    n_rays = len(lidar.rays)
    for _ in range(args.warmup):
        frame(entry, out)
    rgl.check("rgl_node_raytrace_set_kernel_timing", rt.h, True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f0 = time.perf_counter()
        hits = frame(entry, out)
        lat.append(time.perf_counter() - f0)
    total = time.perf_counter() - t0
    kms = C.c_double()
    launches = C.c_int32()
    rgl.check("rgl_node_raytrace_get_kernel_timing", rt.h, C.byref(kms), C.byref(launches))
    rgl.check("rgl_node_raytrace_set_kernel_timing", rt.h, False)
    kernel_ms = kms.value / max(launches.value, 1)
    point_bytes = sum(rgl.field_size(f) for f in FIELDS)
    b_trav, sums = canonical_bytes(rt, n_rays)
    b_io = n_rays * (48 + point_bytes)
    footprint = rgl.traversal_footprint([rt])
    p50, p99 = latency_stats(lat)
    return {
        "metric": METRIC,
        "value": round(args.steps * n_rays / total / 1e6, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 4),
        "p50_ms": p50,
        "p99_ms": p99,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "urban scene, plain single-sensor graph (rays->pose->range->offsets->angular noise->raytrace->"
                        "distance noise->compact->yield)",
            "triangles": spec.triangle_count, "instances": len(spec.instances), "rays": n_rays,
            "lidar": "128 ch x 2048 az, el [-25,15] deg, 200 m", "noise": "angular 0.05 deg, distance 0.02+0.001d",
            "distortion": "10 m/s", "hits_per_frame": hits, "mode": "plain",
        },
        "raytrace_kernel_ms": round(kernel_ms, 4),
        "traversal_per_ray": {"nodes": round(sums[0] / n_rays, 2), "triangles": round(sums[1] / n_rays, 2),
                              "instances": round(sums[2] / n_rays, 2)},
        "roofline": roofline(b_io, footprint, b_trav, kernel_ms, kernel_profile("raytrace_urban_traffic.json")),
        "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=8, help="untimed frames (the first six also settle the autotuner)")
    ap.add_argument("--workload", choices=["batch64", "urban", "serve", "multireturn", "radar"], default="batch64")
    ap.add_argument("--no-serve", action="store_true", help="batch64: skip the serving-loop measurement")
    ap.add_argument("--bvh-format", choices=["bvh2", "bvh4", "bvh4q", "bvh4h"], default="bvh4",
                    help="node format: binary FP32, 4-wide FP32, or 4-wide 8-bit quantized")
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="seconds of CPU-port frames for cpu_baseline (0 = skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copy kernels of MIB MiB (reference for rocprof byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch  # device memory plumbing for the calibration copy only
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": n * 4, "launches": 5}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this library runs every sensor of a frame on one MI355X and has no "
              "multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    rgl.load()
    rgl.check("rgl_configure_logging", 6, None, False)
    fmt = {"bvh2": rgl.BVH2, "bvh4": rgl.BVH4_F32, "bvh4q": rgl.BVH4_Q8, "bvh4h": rgl.BVH4_Q16}[args.bvh_format]
    rgl.set_bvh_format(fmt)
    if args.workload in ("serve", "multireturn", "radar"):
        run = {"serve": run_serve, "multireturn": run_multireturn, "radar": run_radar}[args.workload]
        result = run(args)
        if args.workload == "multireturn":
            result["accuracy"] = golden_accuracy(["c3_first_last", "c3_strongest"])
        print(json.dumps(result))
        rgl.cleanup()
        return
    result = run_batch64(args) if args.workload == "batch64" else run_urban(args)
    if args.workload == "urban":
        rgl.cleanup()
        result["accuracy"] = golden_accuracy(["urban", "c1"])
    if args.workload == "batch64" and not args.no_serve:
        rgl.cleanup()
        result["serving"] = run_serve(args)
    result["config"]["bvh_format"] = {rgl.BVH2: "binary FP32 (64 B nodes)", rgl.BVH4_F32: "4-wide FP32 (128 B nodes)",
                                      rgl.BVH4_Q8: "4-wide 8-bit quantized (64 B nodes)",
                                      rgl.BVH4_Q16: "4-wide 16-bit quantized (80 B nodes)"}[fmt]
    rgl.cleanup()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
