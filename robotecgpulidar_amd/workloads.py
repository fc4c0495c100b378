"""Sensor graphs of the benchmark workloads, built through the RGL API (rgl.py), shared by bench.py and the tests
that check those same graphs.

  sensor_graph      rays -> pose -> range -> raytrace(scene) -> points transform (world -> sensor) -> distance noise
                    -> compact -> yield                      (the 64-lidar batch and the serving loop's lidars)
  urban_graph       rays -> pose -> range -> time offsets -> angular ray noise -> raytrace (velocity distortion)
                    -> distance noise -> compact -> yield    (the urban frame)
  radar_bin_graph   rays -> pose -> range -> raytrace(scene) -> radar range x azimuth binning
  multireturn_graph rays -> pose -> range -> raytrace (divergent beams, return mode) -> compact -> yield
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import rgl
from . import scenes as S

FIELDS = [rgl.XYZ_VEC3_F32, rgl.DISTANCE_F32, rgl.ENTITY_ID_I32, rgl.NORMAL_VEC3_F32, rgl.IS_HIT_I32]
MR_FIELDS = [rgl.XYZ_VEC3_F32, rgl.DISTANCE_F32, rgl.INTENSITY_F32, rgl.ENTITY_ID_I32, rgl.IS_HIT_I32,
             rgl.RETURN_TYPE_U8]
ANGULAR_SIGMA = float(np.float32(0.05 * np.pi / 180.0))


@dataclass
class SensorGraph:
    rays: "rgl.Node"       # entry node
    raytrace: "rgl.Node"
    out: "rgl.Node"        # yield node
    nodes: list            # every node of the chain, in order


_COUNT = C.c_int32()
_COUNT_REF = C.byref(_COUNT)


def kept_count(out) -> int:
    """Points the output node holds after the last run (waits for the run's counts, not for its data)."""
    rgl.check("rgl_graph_get_result_size", out.h, rgl.XYZ_VEC3_F32, _COUNT_REF, None)
    return _COUNT.value


def sensor_graph(lidar: S.Lidar, scene, seed: int) -> SensorGraph:
    n = rgl.Node
    rays = rgl.node_rays_from_mat3x4f(n(), lidar.rays)
    pose = rgl.node_rays_transform(n(), lidar.pose)
    rng = rgl.node_rays_set_range(n(), np.array([lidar.range], np.float32))
    rt = rgl.node_raytrace(n(), scene)
    to_sensor = S.f32(np.linalg.inv(np.vstack([lidar.pose.astype(np.float64), [0, 0, 0, 1]]))[:3])
    tf = rgl.node_points_transform(n(), to_sensor)
    dist = n()
    rgl.check("rgl_node_gaussian_noise_distance", dist.ref, 0.0, 0.02, 0.001)
    comp = rgl.node_points_compact_by_field(n(), rgl.IS_HIT_I32)
    out = rgl.node_points_yield(n(), FIELDS)
    chain = [rays, pose, rng, rt, tf, dist, comp, out]
    rgl.chain(*chain)
    rgl.check("rgl_node_gaussian_noise_set_seed", dist.h, C.c_uint64(seed))
    return SensorGraph(rays, rt, out, chain)


def urban_graph(lidar: S.Lidar, seeds=(0xC0F2A, 0xC0F2D)) -> SensorGraph:
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
    chain = [rays, pose, rng, offs, ang, rt, dist, comp, out]
    rgl.chain(*chain)
    rgl.check("rgl_node_gaussian_noise_set_seed", ang.h, C.c_uint64(seeds[0]))
    rgl.check("rgl_node_gaussian_noise_set_seed", dist.h, C.c_uint64(seeds[1]))
    return SensorGraph(rays, rt, out, chain)


def radar_bin_graph(lidar: S.Lidar, scene) -> SensorGraph:
    n = rgl.Node
    c = S.RADAR_CELL
    rays = rgl.node_rays_from_mat3x4f(n(), lidar.rays)
    pose = rgl.node_rays_transform(n(), lidar.pose)
    rng = rgl.node_rays_set_range(n(), np.array([lidar.range], np.float32))
    rt = rgl.node_raytrace(n(), scene)
    out = rgl.node_points_radar_bin(n(), c["range_step"], c["range_max"], c["azimuth_step"], c["azimuth_min"],
                                    c["azimuth_max"], S.RADAR_SOLID_ANGLE)
    chain = [rays, pose, rng, rt, out]
    rgl.chain(*chain)
    return SensorGraph(rays, rt, out, chain)


def multireturn_graph(lidar: S.Lidar, mode: int, divergence: float, scene=None) -> SensorGraph:
    n = rgl.Node
    rays = rgl.node_rays_from_mat3x4f(n(), lidar.rays)
    pose = rgl.node_rays_transform(n(), lidar.pose)
    rng = rgl.node_rays_set_range(n(), np.array([lidar.range], np.float32))
    rt = rgl.node_raytrace(n(), scene)
    rgl.check("rgl_node_raytrace_configure_return_mode", rt.h, mode)
    rgl.check("rgl_node_raytrace_configure_beam_divergence", rt.h, divergence, divergence)
    rgl.check("rgl_node_raytrace_configure_default_intensity", rt.h, 30.0)
    comp = rgl.node_points_compact_by_field(n(), rgl.IS_HIT_I32)
    out = rgl.node_points_yield(n(), MR_FIELDS)
    chain = [rays, pose, rng, rt, comp, out]
    rgl.chain(*chain)
    return SensorGraph(rays, rt, out, chain)


def load_scenes(specs) -> list:
    """Every spec in a scene of its own (rgl_scene_create); returns [(scene, entities)]."""
    out = []
    for spec in specs:
        sc = rgl.scene_create()
        _, ents, _ = S.load_into_rgl(spec, scene=sc)
        out.append((sc, ents))
    return out
