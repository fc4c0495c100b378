"""Benchmark: one lidar frame of the urban workload through the RGL graph on one MI355X.

Workload (one step = one frame, all inputs resident in HBM): ~2M-triangle urban scene (500 instanced icospheres,
boxes and 4-12 sided prisms on a ground grid), 128 x 2048 = 262,144-ray lidar, range 200 m, Gaussian angular ray
noise (sigma 0.05 deg) and distance noise (0.02 m + 0.001 d), velocity distortion at 10 m/s, non-hit compaction.
Graph: rays -> pose -> range -> time offsets -> angular noise -> raytrace -> distance noise -> compact -> yield.

Prints one JSON line. `value` = Mrays/s of API beams completed through compaction with a host wait per frame
(each frame's latency is measured from rgl_graph_run until the compacted count is readable on the host).
`roofline` is computed for the raytrace kernel from its HIP-event durations on the graph stream;
`cpu_baseline` times the FP64 CPU oracle (multithreaded) on the same frame for a bounded ~10 s sample.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from robotecgpulidar_amd import rgl  # noqa: E402
from robotecgpulidar_amd import scenes as S  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s + p50/p99 frame latency, served-batched vs plain, 1 MI355X"
FIELDS = [rgl.XYZ_VEC3_F32, rgl.DISTANCE_F32, rgl.ENTITY_ID_I32, rgl.NORMAL_VEC3_F32, rgl.IS_HIT_I32]
ANGULAR_SIGMA = float(np.float32(0.05 * np.pi / 180.0))


def build_graph(lidar: S.Lidar):
    n = rgl.Node
    rays = rgl.node_rays_from_mat3x4f(n(), lidar.rays)
    pose = rgl.node_rays_transform(n(), lidar.pose)
    rng = rgl.node_rays_set_range(n(), np.array([lidar.range], np.float32))
    offs = rgl.node_rays_set_time_offsets(n(), lidar.time_offsets)
    ang = n()
    rgl.check("rgl_node_gaussian_noise_angular_ray", ang.ref, 0.0, ANGULAR_SIGMA, rgl.RGL_AXIS_Z)
    rt = rgl.node_raytrace(n())
    lin = np.array([10.0, 0.0, 0.0], np.float32)
    angv = np.zeros(3, np.float32)
    rgl.check("rgl_node_raytrace_configure_velocity", rt.h, lin.ctypes.data, angv.ctypes.data)
    rgl.check("rgl_node_raytrace_configure_distortion", rt.h, True)
    dist = n()
    rgl.check("rgl_node_gaussian_noise_distance", dist.ref, 0.0, 0.02, 0.001)
    comp = rgl.node_points_compact_by_field(n(), rgl.IS_HIT_I32)
    out = rgl.node_points_yield(n(), FIELDS)
    rgl.chain(rays, pose, rng, offs, ang, rt, dist, comp, out)
    rgl.check("rgl_node_gaussian_noise_set_seed", ang.h, C.c_uint64(0xC0F2A))
    rgl.check("rgl_node_gaussian_noise_set_seed", dist.h, C.c_uint64(0xC0F2D))
    return rays, rt, out


def frame(entry, out) -> int:
    rgl.check("rgl_graph_run", entry.h)
    count = C.c_int32()
    rgl.check("rgl_graph_get_result_size", out.h, rgl.XYZ_VEC3_F32, C.byref(count), None)
    return count.value


def algorithmic_bytes(rt, n_rays: int, return_count: int, point_bytes: int, n_instances: int):
    """Bytes one raytrace launch must move: ray poses in, per-return fields out, instance records, plus BVH
    traffic counted per ray on this BVH (64 B per node visit, 36 B per triangle test, 64 B per instance entry)."""
    nodes = np.empty(n_rays, np.int32)
    tris = np.empty(n_rays, np.int32)
    insts = np.empty(n_rays, np.int32)
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
    rgl.check("rgl_node_raytrace_traversal_stats", rt.h, p(nodes), p(tris), p(insts), n_rays)
    b_io = n_rays * (48 + return_count * point_bytes) + 48 * n_instances
    b_trav = 64 * int(nodes.sum(dtype=np.int64)) + 36 * int(tris.sum(dtype=np.int64)) + 64 * int(insts.sum(dtype=np.int64))
    return b_io, b_trav, float(nodes.mean()), float(tris.mean()), float(insts.mean())


def cpu_baseline(spec, lidar, budget_s: float):
    from oracle import oracle  # the oracle is the timed CPU baseline only here
    threads = min(os.cpu_count() or 1, 16)
    world_rays = np.einsum("ij,njk->nik", lidar.pose[:, :3], lidar.rays).astype(np.float32)
    world_rays[:, :, 3] += lidar.pose[:, 3]
    w = oracle.World(spec)
    kw = dict(ray_origin_to_world=lidar.pose, ranges=[lidar.range], time_offsets=lidar.time_offsets, distortion=True,
              lin_vel=(10, 0, 0), eps=0.0, threads=threads, world=w, fields=FIELDS)
    frames, t0 = 0, time.perf_counter()
    while True:
        oracle.raytrace(spec, world_rays, **kw)
        frames += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    w.close()
    return {"value": round(frames * len(world_rays) / el / 1e6, 4), "unit": "Mrays/s", "cores": threads,
            "kind": "port",
            "sample": f"{frames} full frames ({len(world_rays)} rays each) of the same scene and lidar, FP64 oracle "
                      f"BVH traversal with distortion, BVH build excluded, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds of oracle work for cpu_baseline (0 = skip)")
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
    spec, lidar = S.config2_scene(), S.config2_lidar()
    S.load_into_rgl(spec)
    entry, rt, out = build_graph(lidar)
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
    b_io, b_trav, nodes_pr, tris_pr, insts_pr = algorithmic_bytes(rt, n_rays, 1, point_bytes, len(spec.instances))
    achieved = (b_io + b_trav) / (kernel_ms * 1e-3) / 1e9
    traffic = None
    tf = ROOT / "profiles" / "raytrace_config2_traffic.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("bytes_per_launch")

    lat_ms = np.array(lat) * 1e3
    value = args.steps * n_rays / total / 1e6
    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 4),
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 4),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "urban scene, plain per-sensor graph (rays->pose->range->offsets->angular noise->raytrace->"
                        "distance noise->compact->yield)",
            "triangles": spec.triangle_count, "instances": len(spec.instances), "rays": n_rays,
            "lidar": "128 ch x 2048 az, el [-25,15] deg, 200 m", "noise": "angular 0.05 deg, distance 0.02+0.001d",
            "distortion": "10 m/s", "hits_per_frame": hits, "mode": "plain",
        },
        "raytrace_kernel_ms": round(kernel_ms, 4),
        "traversal_per_ray": {"nodes": round(nodes_pr, 2), "triangles": round(tris_pr, 2), "instances": round(insts_pr, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": b_io + b_trav, "io_bytes_per_launch": b_io,
                     "io_only_frac": round(b_io / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)},
        "cpu_baseline": cpu_baseline(spec, lidar, args.cpu_budget) if args.cpu_budget > 0 else None,
    }
    rgl.cleanup()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
