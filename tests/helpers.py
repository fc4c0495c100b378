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
        """World-frame rays for the oracle. With a pose, the exact FP64 product T * in of the FP32 pose and rays
        (TransformRaysNode, src/gpu/nodeKernels.cu:45-49), so that the library's own composition error counts
        against it; without one, the FP32 rays as given."""
        if self.pose is None:
            return self.rays
        return compose64(self.pose, self.rays)


def compose64(pose, rays) -> np.ndarray:
    """T * rays for (3, 4) T and (N, 3, 4) rays, in FP64 from their FP32 values."""
    p = np.asarray(pose, np.float32).astype(np.float64)
    r = np.asarray(rays, np.float32).astype(np.float64)
    out = np.matmul(p[None, :, :3], r)
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


# ------------------------------------------------------------------------------------------ golden comparison
GEOM, DIST_TIE, INT_TIE, TEXEL = 1, 2, 4, 8  # oracle ambiguity bits
G_XYZ, G_INTEN, G_IS_HIT, G_ENT, G_DIST, G_NORMAL = 1, 2, 4, 7, 8, 22


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


def check_golden(gpu: dict, hit, entity, ambiguous, idx, values: dict, return_mode: int, label: str,
                 min_compared: int, exact_idx=None) -> dict:
    """Checks the library's return points `gpu` (field -> array over every return point) against oracle outputs:
    `hit`, `entity` over every return point, per-beam `ambiguous` bits, and `values` (field -> oracle values at the
    return points `idx`). XYZ and normals are compared at `exact_idx` (positions into idx; default all of idx).
    Bar: hit masks equal except on geometrically ambiguous beams, and for single returns at most max(1, 1e-5 n)
    differences; entity ids equal on unambiguous beams; distance within max(1 mm, 1e-5 relative), XYZ within 2 mm,
    normals within 1e-4, intensity within 1e-4 relative + 2e-3 where no ambiguity bit blocks the quantity. At least
    `min_compared` distances and intensities together must be compared.
    Returns counts and the largest distance error."""
    n = len(hit)
    beams = len(ambiguous)
    R = n // beams
    amb = np.repeat(np.asarray(ambiguous), R)
    g_hit = gpu[G_IS_HIT]
    assert len(g_hit) == n, (label, len(g_hit), n)
    hit_mism = g_hit != hit
    unexplained = hit_mism & ((amb & GEOM) == 0)
    assert not unexplained.any(), (label, np.nonzero(unexplained)[0][:10])
    if R == 1:
        assert hit_mism.sum() <= max(1, int(1e-5 * n)), (label, int(hit_mism.sum()))
    ent_mism = (g_hit == 1) & (hit == 1) & (gpu[G_ENT] != entity)
    assert not (ent_mism & (amb == 0)).any(), (label, np.nonzero(ent_mism & (amb == 0))[0][:10])
    types = (return_mode >> (8 * (idx % R))) & 0xFF
    both = (hit[idx] == 1) & (g_hit[idx] == 1)
    block_d = np.array([blocking_bits(int(t), "distance") for t in types], np.int64)
    block_i = np.array([blocking_bits(int(t), "intensity") for t in types], np.int64)
    sel = both & ((amb[idx] & block_d) == 0)
    sel_i = both & ((amb[idx] & block_i) == 0)
    compared = int(sel.sum()) + (int(sel_i.sum()) if G_INTEN in values else 0)
    assert compared >= min_compared, (label, int(sel.sum()), int(sel_i.sum()))
    d_o = values[G_DIST][sel].astype(np.float64)
    d_g = gpu[G_DIST][idx][sel].astype(np.float64)
    err = np.abs(d_g - d_o)
    assert np.all(err <= np.maximum(1e-3, 1e-5 * d_o)), (label, err.max())
    exact = both & (amb[idx] == 0)
    ex = np.arange(len(idx)) if exact_idx is None else np.asarray(exact_idx)
    ex = ex[exact[ex]]
    if G_XYZ in values:
        assert np.abs(gpu[G_XYZ][idx][ex] - values[G_XYZ][ex]).max(initial=0.0) <= 2e-3, label
    if G_NORMAL in values:
        assert np.abs(gpu[G_NORMAL][idx][ex] - values[G_NORMAL][ex]).max(initial=0.0) <= 1e-4, label
    if G_INTEN in values:
        i_o = values[G_INTEN][sel_i].astype(np.float64)
        i_g = gpu[G_INTEN][idx][sel_i].astype(np.float64)
        bad = np.abs(i_g - i_o) > 1e-4 * i_o + 2e-3
        assert not bad.any(), (label, idx[sel_i][bad], i_g[bad], i_o[bad])
    non_hit = (hit[idx] == 0) & (g_hit[idx] == 0)
    np.testing.assert_array_equal(gpu[G_DIST][idx][non_hit], values[G_DIST][non_hit])
    return {"points": n, "hits": int(hit.sum()), "mask_differences": int(hit_mism.sum()),
            "ambiguous_beams": int((np.asarray(ambiguous) != 0).sum()), "compared_distances": int(sel.sum()),
            "compared_intensities": int(sel_i.sum()) if G_INTEN in values else 0,
            "max_dist_err_m": float(err.max(initial=0.0))}
