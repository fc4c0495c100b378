"""Shared test helpers: one description of a raytrace case, run either through the HIP library (C-ABI) or the
FP64 oracle, so that every known-answer test reads the same on both paths."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from robotecgpulidar_amd import scenes as S

FLT_MAX = float(np.finfo(np.float32).max)


@dataclass
class Case:
    spec: S.SceneSpec
    rays: np.ndarray                    # (N, 3, 4) sensor-frame rays given to rgl_node_rays_from_mat3x4f
    pose: np.ndarray | None = None      # rgl_node_rays_transform, if any
    ranges: np.ndarray | None = None    # rgl_node_rays_set_range, if any
    near: float = float("inf")
    far: float = float("inf")
    default_intensity: float = 0.0
    mask: np.ndarray | None = None
    return_mode: int = (1 << 24) | 4
    divergence: tuple = (0.0, 0.0)
    ring_ids: np.ndarray | None = None
    time_offsets: np.ndarray | None = None
    distortion: bool = False
    velocity: tuple | None = None       # ((vx,vy,vz), (wx,wy,wz))
    extra: dict = field(default_factory=dict)

    def world_rays(self) -> np.ndarray:
        if self.pose is None:
            return self.rays
        p = np.asarray(self.pose, np.float32)
        r = np.asarray(self.rays, np.float32)
        # TransformRaysNode: out = T * in, evaluated in FP32 like the kernel (src/gpu/nodeKernels.cu:45-49).
        out = np.einsum("ij,njk->nik", p[:, :3], r).astype(np.float32)
        out[:, :, 3] += p[:, 3]
        return out


def run_oracle(oracle, case: Case, fields=None, brute_force=False, eps=1e-5):
    ranges = case.ranges if case.ranges is not None else np.array([[0.0, FLT_MAX]], np.float32)
    vel = case.velocity or ((0, 0, 0), (0, 0, 0))
    return oracle.raytrace(case.spec, case.world_rays(),
                           ray_origin_to_world=case.pose if case.pose is not None else np.eye(3, 4),
                           ranges=ranges, ring_ids=case.ring_ids, time_offsets=case.time_offsets, mask=case.mask,
                           distortion=case.distortion, lin_vel=vel[0], ang_vel=vel[1], near=case.near, far=case.far,
                           default_intensity=case.default_intensity, divergence=case.divergence,
                           return_mode=case.return_mode, brute_force=brute_force, eps=eps, fields=fields)


class Graph:
    """rays -> [transform] -> [range] -> [ring ids] -> [time offsets] -> raytrace -> [compact] -> yield."""

    def __init__(self, rgl, case: Case, fields, compact=False, points_transform=None):
        self.rgl = rgl
        self.rays = rgl.node_rays_from_mat3x4f(rgl.Node(), case.rays)
        chain = [self.rays]
        if case.pose is not None:
            chain.append(rgl.node_rays_transform(rgl.Node(), case.pose))
        if case.ranges is not None:
            chain.append(rgl.node_rays_set_range(rgl.Node(), case.ranges))
        if case.ring_ids is not None:
            chain.append(rgl.node_rays_set_ring_ids(rgl.Node(), case.ring_ids))
        if case.time_offsets is not None:
            chain.append(rgl.node_rays_set_time_offsets(rgl.Node(), case.time_offsets))
        self.raytrace = rgl.node_raytrace(rgl.Node())
        rt = self.raytrace.h
        if case.near != float("inf") or case.far != float("inf"):
            rgl.check("rgl_node_raytrace_configure_non_hits", rt, case.near, case.far)
        if case.default_intensity:
            rgl.check("rgl_node_raytrace_configure_default_intensity", rt, case.default_intensity)
        if case.mask is not None:
            m = np.ascontiguousarray(case.mask, np.int8)
            rgl.check("rgl_node_raytrace_configure_mask", rt, m.ctypes.data, len(m))
        if case.return_mode != ((1 << 24) | 4):
            rgl.check("rgl_node_raytrace_configure_return_mode", rt, case.return_mode)
        if case.divergence != (0.0, 0.0):
            rgl.check("rgl_node_raytrace_configure_beam_divergence", rt, *case.divergence)
        if case.velocity is not None:
            lin = np.asarray(case.velocity[0], np.float32)
            ang = np.asarray(case.velocity[1], np.float32)
            rgl.check("rgl_node_raytrace_configure_velocity", rt, lin.ctypes.data, ang.ctypes.data)
        if case.distortion:
            rgl.check("rgl_node_raytrace_configure_distortion", rt, True)
        chain.append(self.raytrace)
        if compact:
            chain.append(rgl.node_points_compact_by_field(rgl.Node(), rgl.IS_HIT_I32))
        if points_transform is not None:
            chain.append(rgl.node_points_transform(rgl.Node(), points_transform))
        self.out = rgl.node_points_yield(rgl.Node(), fields)
        chain.append(self.out)
        rgl.chain(*chain)
        self.fields = list(fields)

    def run(self) -> dict:
        self.rgl.graph_run(self.rays)
        return {f: self.rgl.result(self.out, f) for f in self.fields}


def run_gpu(rgl, case: Case, fields, compact=False, points_transform=None) -> dict:
    S.load_into_rgl(case.spec)
    return Graph(rgl, case, fields, compact=compact, points_transform=points_transform).run()


# ------------------------------------------------------------------------------------------ reference KATs
def cube_scene(*transforms, ids=None) -> S.SceneSpec:
    spec = S.SceneSpec()
    m = spec.add_mesh(S.CUBE_VERTICES, S.CUBE_INDICES)
    for k, t in enumerate(transforms):
        spec.add_instance(m, t, entity_id=None if ids is None else ids[k])
    return spec


def identity_ray() -> np.ndarray:
    return S.f32(S.identity()[None])


DISTANCE_PARAMS = [23.1753849, 119.9884538, 333.3394629]  # test/src/graph/DistanceFieldTest.cpp:83


def f32sub(a, b) -> np.float32:
    return np.float32(np.float32(a) - np.float32(b))
