"""Deferred ray transforms: when the only consumers of a rays-transform node are raytrace nodes that do not output
ray poses, the transform is composed with each ray in FP64 inside the raytrace kernel instead of in a pass of its
own, and the FP64 origin is kept for the world -> object transforms and the hit distance. A materialized transform
(forced here by also yielding RAY_POSE) writes the same composition rounded to FP32 rays, so the two differ only by
that rounding: hit masks and entity ids agree except for at most a few grazing rays, distances within 1e-6
relative. Each form is bit-identical between solo runs and batches, with divergent beams and with velocity
distortion."""
import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S
from tests.test_gpu_batch import assert_same, build_world

pytestmark = pytest.mark.gpu

XYZ, IS_HIT, ENT, DIST, NORMAL, RAY_POSE = 1, 4, 7, 8, 22, 24
FIELDS = [XYZ, DIST, ENT, IS_HIT, NORMAL]


class Graph:
    """rays -> transform(pose) -> range -> [time offsets] -> raytrace -> yield (+ RAY_POSE if `with_pose`)."""

    def __init__(self, rgl, scene, ch, az, seed, with_pose, divergent=False, distortion=False):
        self.rgl = rgl
        lidar = S.spinning_lidar(ch, az, -25.0, 12.0)
        pose = S.f32(S.trs((0.7 + 0.3 * seed, -1.3, 1.9), (3.0 * seed, -2.0, 17.0 * seed), (1, 1, 1)))
        self.rays = rgl.node_rays_from_mat3x4f(rgl.Node(), lidar)
        chain = [self.rays, rgl.node_rays_transform(rgl.Node(), pose),
                 rgl.node_rays_set_range(rgl.Node(), np.array([[0.3, 120.0]], np.float32))]
        if distortion:
            chain.append(rgl.node_rays_set_time_offsets(rgl.Node(), np.linspace(0.0, 50.0, len(lidar), dtype=np.float32)))
        rt = rgl.node_raytrace(rgl.Node(), scene)
        rgl.check("rgl_node_raytrace_configure_non_hits", rt.h, 0.0, 120.0)
        if divergent:
            rgl.check("rgl_node_raytrace_configure_beam_divergence", rt.h, 2e-3, 2e-3)
            rgl.check("rgl_node_raytrace_configure_return_mode", rt.h, rgl.RGL_RETURN_FIRST_LAST)
        if distortion:
            lin = np.asarray([12.0, -3.0, 0.5], np.float32)
            ang = np.asarray([0.0, 0.1, 0.8], np.float32)
            rgl.check("rgl_node_raytrace_configure_velocity", rt.h, lin.ctypes.data, ang.ctypes.data)
            rgl.check("rgl_node_raytrace_configure_distortion", rt.h, True)
        chain.append(rt)
        self.fields = FIELDS + ([RAY_POSE] if with_pose else [])
        self.out = rgl.node_points_yield(rgl.Node(), self.fields)
        chain.append(self.out)
        rgl.chain(*chain)

    def results(self):
        return {f: self.rgl.result(self.out, f) for f in self.fields if f != RAY_POSE}


CONFIGS = [
    dict(ch=32, az=700),
    dict(ch=8, az=150, divergent=True),
    dict(ch=16, az=400, distortion=True),
]


def pairs(rgl, scenes):
    out = []
    for k, cfg in enumerate(CONFIGS):
        scene = scenes[k % 2]
        out.append((Graph(rgl, scene, seed=k, with_pose=False, **cfg), Graph(rgl, scene, seed=k, with_pose=True, **cfg)))
    return out


def assert_close(a, b, label):
    """Results of a deferred and a materialized transform: the FP32 rounding of the composed origin apart."""
    assert a[IS_HIT].shape == b[IS_HIT].shape, label
    mism = (a[IS_HIT] != b[IS_HIT]) | ((a[IS_HIT] == 1) & (a[ENT] != b[ENT]))
    assert mism.sum() <= 2, (label, int(mism.sum()))
    both = (a[IS_HIT] == 1) & (b[IS_HIT] == 1) & ~mism
    np.testing.assert_allclose(a[DIST][both], b[DIST][both], rtol=1e-6, atol=1e-6, err_msg=label)
    np.testing.assert_allclose(a[XYZ][both], b[XYZ][both], rtol=1e-6, atol=1e-5, err_msg=label)


def test_deferred_transform_matches_materialized(rgl_lib):
    rgl = rgl_lib
    _, scenes, _ = build_world(rgl)
    graphs = pairs(rgl, scenes)
    solo = []
    for k, (deferred, materialized) in enumerate(graphs):
        rgl.graph_run(deferred.rays)
        rgl.graph_run(materialized.rays)
        a, b = deferred.results(), materialized.results()
        assert_close(a, b, f"solo config {k}")
        assert int(a[IS_HIT].sum()) > 0, k
        solo.append((a, b))
    rgl.graph_run_batch([g.rays for pair in graphs for g in pair])
    for k, (deferred, materialized) in enumerate(graphs):
        assert_same(deferred.results(), solo[k][0], f"batched deferred config {k}")
        assert_same(materialized.results(), solo[k][1], f"batched materialized config {k}")


def test_ray_poses_follow_the_transform_when_requested(rgl_lib):
    """With RAY_POSE requested the transform is materialized; the poses are the transform applied to the input
    rays (float64 reference, within FP32 rounding of a 3x4 product)."""
    rgl = rgl_lib
    _, scenes, _ = build_world(rgl)
    lidar = S.spinning_lidar(8, 64, -10.0, 10.0)
    pose = S.f32(S.trs((1.0, 2.0, 3.0), (10.0, 20.0, 30.0), (1, 1, 1)))
    rays = rgl.node_rays_from_mat3x4f(rgl.Node(), lidar)
    tf = rgl.node_rays_transform(rgl.Node(), pose)
    rt = rgl.node_raytrace(rgl.Node(), scenes[1])
    out = rgl.node_points_yield(rgl.Node(), [XYZ, RAY_POSE])
    rgl.chain(rays, tf, rt, out)
    rgl.graph_run(rays)
    got = rgl.result(out, RAY_POSE).reshape(-1, 3, 4).astype(np.float64)
    p = pose.astype(np.float64)
    want = np.einsum("ij,njk->nik", p[:, :3], lidar.astype(np.float64))
    want[:, :, 3] += p[:, 3]
    np.testing.assert_allclose(got, want, rtol=0, atol=4e-6 * max(1.0, float(np.abs(want).max())))


def test_deferral_follows_graph_changes(rgl_lib):
    """The decision to defer a transform into the raytrace launch is cached per graph structure: adding a yield that
    requests RAY_POSE must materialize the transform at the next run (poses equal to the transform of the rays), and
    removing it must defer it again: the third run's points are bit-identical to the first's, and the materialized
    run differs from them only by the FP32 rounding of the composed origin."""
    rgl = rgl_lib
    _, scenes, _ = build_world(rgl)
    lidar = S.spinning_lidar(8, 64, -10.0, 10.0)
    pose = S.f32(S.trs((1.0, 2.0, 3.0), (10.0, 20.0, 30.0), (1, 1, 1)))
    rays = rgl.node_rays_from_mat3x4f(rgl.Node(), lidar)
    tf = rgl.node_rays_transform(rgl.Node(), pose)
    rt = rgl.node_raytrace(rgl.Node(), scenes[1])
    out = rgl.node_points_yield(rgl.Node(), [XYZ, DIST])
    rgl.chain(rays, tf, rt, out)
    rgl.graph_run(rays)
    first = {f: rgl.result(out, f) for f in (XYZ, DIST)}
    poses = rgl.node_points_yield(rgl.Node(), [RAY_POSE])
    rgl.add_child(rt, poses)
    rgl.graph_run(rays)
    got = rgl.result(poses, RAY_POSE).reshape(-1, 3, 4).astype(np.float64)
    p = pose.astype(np.float64)
    want = np.einsum("ij,njk->nik", p[:, :3], lidar.astype(np.float64))
    want[:, :, 3] += p[:, 3]
    np.testing.assert_allclose(got, want, rtol=0, atol=4e-6 * max(1.0, float(np.abs(want).max())))
    second = {f: rgl.result(out, f) for f in (XYZ, DIST)}
    rgl.check("rgl_graph_node_remove_child", rt.h, poses.h)
    rgl.graph_run(rays)
    third = {f: rgl.result(out, f) for f in (XYZ, DIST)}
    for f in (XYZ, DIST):
        assert first[f].tobytes() == third[f].tobytes(), f
        np.testing.assert_allclose(second[f], first[f], rtol=1e-5, atol=1e-4)
