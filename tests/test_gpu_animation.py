"""Scene update path: external vertex animation (rgl_entity_apply_external_animation) and skeletal animation
(rgl_mesh_set_bone_weights / rgl_mesh_set_restposes / rgl_entity_set_pose_world) with the animated entity's own BLAS
refitted each step, and the hit-point velocities that the vertex displacement contributes
(src/gpu/optixPrograms.cu:169-212, src/scene/Entity.cpp:88-117, src/gpu/sceneKernels.cu:18-49).

Pins: the reference's velocity known-answer tests (test/src/scene/animationVelocityTest.cpp), the FP64 oracle on
animated scenes with moving entities (hit masks, distances, normals, velocities), bit-identical results between a
refitted animated entity and a freshly built mesh holding the same vertices in every BVH format, and the
reference's argument errors."""
import ctypes as C

import numpy as np
import pytest

from robotecgpulidar_amd import rgl as R
from robotecgpulidar_amd import scenes as S
from tests.helpers import Case, Graph, cube_scene
from tests.test_gpu_features import mismatch_ok

pytestmark = pytest.mark.gpu

XYZ, IS_HIT, ENT, DIST, ABS_V, REL_V, RAD, NORMAL = 1, 4, 7, 8, 15, 16, 17, 22
SECOND = 1_000_000_000

WORLD_TFS = [S.identity(), S.translation(10.0, 20.0, 30.0), S.rotation_deg(10.0, 20.0, 30.0),
             S.trs((10.0, 20.0, 30.0), (10.0, 20.0, 30.0))]
ANIM_TFS = [S.identity(), S.translation(0.0, 0.0, 0.5), S.translation(0.0, 0.0, -0.5), S.translation(0.0, 0.5, 0.0),
            S.translation(0.0, -0.5, 0.0)]


def velocity_graph(rgl):
    rays = rgl.node_rays_from_mat3x4f(rgl.Node(), S.f32(S.identity()[None]))
    pose = rgl.node_rays_transform(rgl.Node(), S.f32(S.identity()))
    rt = rgl.node_raytrace(rgl.Node())
    out = rgl.node_points_yield(rgl.Node(), [XYZ, REL_V, ABS_V])
    rgl.chain(rays, pose, rt, out)
    return rays, pose, out


def check_velocity(rgl, graph, world_tf, expected_rel):
    rays, pose, out = graph
    rgl.node_rays_transform(pose, S.f32(world_tf))
    rgl.graph_run(rays)
    rel, ab = rgl.result(out, REL_V), rgl.result(out, ABS_V)
    assert rel.shape == (1, 3)
    # The reference's tolerance is 1e-6. Vertex positions are rounded to FP32 in world space (|x| < 64 here), so the
    # displacement of a vertex carries up to one unit in the last place of its coordinates: 2^-18 = 3.8e-6.
    tol = 1e-6 + 2.0 ** -18
    np.testing.assert_allclose(rel[0], expected_rel, atol=tol)
    np.testing.assert_allclose(ab[0], S.f32(world_tf)[:, :3] @ expected_rel, atol=tol)


def test_skeleton_animation_velocity_kat(rgl_lib):
    rgl = rgl_lib
    mesh = rgl.mesh_create(S.CUBE_VERTICES, S.CUBE_INDICES)
    ent = rgl.entity_create(mesh)
    # Vertices 0-3 follow bone 0, vertices 4-7 (the z = +1 face) follow bone 1.
    weights = np.tile([1.0, 0.0, 0.0, 0.0], (8, 1))
    bones = np.zeros((8, 4), np.int32)
    bones[4:, 0] = 1
    rgl.mesh_set_bone_weights(mesh, weights, bones)
    rest = [S.identity(), S.identity()]
    rgl.mesh_set_restposes(mesh, np.stack([S.f32(r) for r in rest]))
    graph = velocity_graph(rgl)
    frame = 0
    for wt in WORLD_TFS:
        for at in ANIM_TFS:
            rgl.scene_set_time(None, frame * SECOND)
            rgl.entity_set_pose_world(ent, np.stack([S.f32(S.mul(wt, r)) for r in rest]))
            frame += 1
            rgl.scene_set_time(None, frame * SECOND)
            rgl.entity_set_pose_world(ent, np.stack([S.f32(S.mul(wt, rest[0])), S.f32(S.mul(S.mul(wt, at), rest[1]))]))
            check_velocity(rgl, graph, wt, np.asarray(at)[:, 3].astype(np.float32))
            frame += 1


def test_external_animation_velocity_kat(rgl_lib):
    rgl = rgl_lib
    mesh = rgl.mesh_create(S.CUBE_VERTICES, S.CUBE_INDICES)
    ent = rgl.entity_create(mesh)
    graph = velocity_graph(rgl)
    frame = 0
    for wt in WORLD_TFS:
        for at in ANIM_TFS:
            rgl.scene_set_time(None, frame * SECOND)
            rgl.entity_apply_external_animation(ent, S.CUBE_VERTICES)
            rgl.entity_set_transform(ent, S.f32(wt))
            frame += 1
            rgl.scene_set_time(None, frame * SECOND)
            moved = S.CUBE_VERTICES.copy()
            moved[4:] = (np.asarray(at)[:, :3] @ moved[4:].T).T + np.asarray(at)[:, 3]
            rgl.entity_apply_external_animation(ent, moved)
            rgl.entity_set_transform(ent, S.f32(wt))
            check_velocity(rgl, graph, wt, np.asarray(at)[:, 3].astype(np.float32))
            frame += 1


def cloth(seed, t):
    """A 40 x 40 m sheet of 60 x 60 cells with a travelling wave: vertex positions at time t (seconds)."""
    v, i = S.grid_mesh(40.0, 40.0, 60, 60)
    v = v.astype(np.float64)
    rng = np.random.default_rng(seed)
    phase = rng.uniform(0, 2 * np.pi)
    v[:, 2] = 1.5 * np.sin(0.3 * v[:, 0] + 2.0 * t + phase) * np.cos(0.2 * v[:, 1] - t)
    return v.astype(np.float32), i


def jitter(v, seed, amp):
    rng = np.random.default_rng(seed)
    return (v + rng.uniform(-amp, amp, v.shape)).astype(np.float32)


@pytest.mark.parametrize("fmt", [R.BVH2, R.BVH4_F32, R.BVH4_Q8, R.BVH4_Q16])
def test_animated_scene_matches_oracle_with_velocities(rgl_lib, oracle, fmt):
    rgl = rgl_lib
    default_fmt = rgl.get_bvh_format()
    rgl.set_bvh_format(fmt)
    try:
        base = S.config1_scene()
        cloth0, cloth_idx = cloth(7, 0.0)
        base_meshes, ents, _ = S.load_into_rgl(base)
        cloth_mesh = rgl.mesh_create(cloth0, cloth_idx)
        cube_mesh = rgl.mesh_create(S.CUBE_VERTICES, S.CUBE_INDICES)
        cloth_tf = S.f32(S.trs((0.0, 25.0, 3.0), (0, 0, 20), (1, 1, 1)))
        cube_tf = S.f32(S.trs((8.0, -6.0, 2.0), (0, 0, 35), (2, 2, 2)))
        cloth_ent, cube_ent = rgl.entity_create(cloth_mesh), rgl.entity_create(cube_mesh)
        rgl.entity_set_id(cloth_ent, 500)
        rgl.entity_set_id(cube_ent, 501)
        rays = S.spinning_lidar(32, 900, -25.0, 15.0)
        pose = S.f32(S.translation(0.0, 0.0, 1.8))
        case = Case(base, rays, pose=pose, ranges=np.array([[0.1, 100.0]], np.float32),
                    velocity=((2.0, 0.0, 0.0), (0.0, 0.0, 0.3)))
        fields = [XYZ, DIST, ENT, IS_HIT, NORMAL, ABS_V, REL_V, RAD]
        g = Graph(rgl, case, fields)
        dt = 0.05
        prev = {"cloth": cloth0, "cube": S.CUBE_VERTICES, "cube_tf": cube_tf}
        for frame in range(4):
            rgl.scene_set_time(None, int(round(frame * dt * SECOND)))
            cloth_v, _ = cloth(7, frame * dt)
            cube_v = jitter(S.CUBE_VERTICES, 100 + frame, 0.3) if frame else S.CUBE_VERTICES
            moved_tf = S.f32(S.trs((8.0 + 0.4 * frame, -6.0, 2.0), (0, 0, 35 + 5 * frame), (2, 2, 2)))
            rgl.entity_apply_external_animation(cloth_ent, cloth_v)
            rgl.entity_apply_external_animation(cube_ent, cube_v)
            rgl.entity_set_transform(cloth_ent, cloth_tf)
            rgl.entity_set_transform(cube_ent, moved_tf)
            gpu = g.run()
            # The oracle's scene: the base scene plus the two animated entities as meshes of their current vertices.
            spec = S.SceneSpec(list(base.meshes), list(base.instances), list(base.textures))
            mc = spec.add_mesh(cloth_v, cloth_idx)
            mq = spec.add_mesh(cube_v, S.CUBE_INDICES)
            ic = spec.add_instance(mc, cloth_tf, entity_id=500)
            iq = spec.add_instance(mq, moved_tf, entity_id=501)
            prev_tfs = [i.transform for i in base.instances] + [cloth_tf, prev["cube_tf"] if frame else moved_tf]
            disp = {ic: cloth_v.astype(np.float64) - prev["cloth"], iq: cube_v.astype(np.float64) - prev["cube"]}
            orc = oracle.raytrace(spec, case.world_rays(), ray_origin_to_world=pose, ranges=case.ranges,
                                  lin_vel=case.velocity[0], ang_vel=case.velocity[1],
                                  scene_dt=dt if frame else 0.0, prev_transforms=np.stack(prev_tfs),
                                  displacements=disp if frame else None, fields=fields)
            ok = mismatch_ok(gpu, orc, len(rays))
            animated = ok & (gpu[ENT] >= 500)
            print(f"format {fmt} frame {frame}: {int(ok.sum())} hits compared, {int(animated.sum())} on animated "
                  f"entities, {int(orc['ambiguous'].astype(bool).sum())} ambiguous beams")
            assert animated.sum() > 300
            d_g, d_o = gpu[DIST][ok].astype(np.float64), orc[DIST][ok].astype(np.float64)
            assert np.all(np.abs(d_g - d_o) <= np.maximum(1e-3, 1e-5 * d_o))
            exact = ok & ~np.asarray(orc["ambiguous"], bool)
            assert np.abs(gpu[NORMAL][exact] - orc[NORMAL][exact]).max() <= 1e-4
            if frame:
                # Velocities: cloth points move up to ~3 m/s, cube points up to ~15 m/s (0.3 m jitter in 50 ms, x2).
                # FP32 positions differenced over 50 ms give about 1e-5 m / 0.05 s; the budget allows 100 times that.
                for f in (ABS_V, REL_V):
                    np.testing.assert_allclose(gpu[f][ok], orc[f][ok], atol=2e-2, rtol=1e-4)
                np.testing.assert_allclose(gpu[RAD][ok], orc[RAD][ok], atol=2e-2, rtol=1e-4)
                moving = np.linalg.norm(orc[ABS_V][animated], axis=1)
                assert (moving > 0.5).sum() > 100
            else:
                assert np.isnan(gpu[ABS_V][ok]).all()
            prev = {"cloth": cloth_v.astype(np.float64), "cube": cube_v.astype(np.float64), "cube_tf": moved_tf}
    finally:
        rgl.set_bvh_format(default_fmt)


@pytest.mark.parametrize("fmt", [R.BVH2, R.BVH4_F32, R.BVH4_Q8, R.BVH4_Q16])
def test_refitted_animation_equals_fresh_mesh(rgl_lib, fmt):
    """The animated entity's BLAS is refitted (the tree of the first vertices with boxes of the current ones), a
    fresh mesh is built from scratch: closest hits do not depend on the tree, so results are bit-identical."""
    rgl = rgl_lib
    default_fmt = rgl.get_bvh_format()
    rgl.set_bvh_format(fmt)
    try:
        v0, idx = cloth(3, 0.0)
        sphere_v, sphere_i = S.icosphere(3)
        tf = S.f32(S.trs((0.0, 15.0, 2.0), (0, 0, 10), (1, 1, 1)))
        stf = S.f32(S.trs((-4.0, 9.0, 3.0), (0, 0, 0), (2.5, 2.5, 2.5)))
        anim_scene, fresh_scene = rgl.scene_create(), rgl.scene_create()
        m_anim = rgl.mesh_create(v0, idx)
        m_sphere = rgl.mesh_create(sphere_v, sphere_i)
        a_cloth = rgl.entity_create(m_anim, anim_scene)
        a_sphere = rgl.entity_create(m_sphere, anim_scene)
        rgl.entity_set_transform(a_cloth, tf)
        rgl.entity_set_transform(a_sphere, stf)
        rays = S.spinning_lidar(64, 720, -30.0, 10.0)
        fields = [XYZ, DIST, ENT, IS_HIT, NORMAL]

        def graph(scene):
            n = rgl.Node
            src = rgl.node_rays_from_mat3x4f(n(), rays)
            pose = rgl.node_rays_transform(n(), S.f32(S.translation(0.0, 0.0, 1.5)))
            rt = rgl.node_raytrace(n(), scene)
            out = rgl.node_points_yield(n(), fields)
            rgl.chain(src, pose, rt, out)
            return src, out

        ga, gf = graph(anim_scene), graph(fresh_scene)
        fresh_meshes = []
        for step in range(4):
            v = cloth(3, 0.7 * step)[0]
            if step == 3:
                v = jitter(v, 9, 0.5)  # large non-smooth deformation: refitted boxes overlap heavily
            rgl.entity_apply_external_animation(a_cloth, v)
            # A fresh scene each step: the same vertices in a newly built mesh, the same sphere.
            for e in fresh_meshes:
                rgl.check("rgl_entity_destroy", e)
            fm = rgl.mesh_create(v, idx)
            fresh_meshes = [rgl.entity_create(fm, fresh_scene), rgl.entity_create(m_sphere, fresh_scene)]
            rgl.entity_set_transform(fresh_meshes[0], tf)
            rgl.entity_set_transform(fresh_meshes[1], stf)
            rgl.graph_run(ga[0])
            rgl.graph_run(gf[0])
            ra = {f: rgl.result(ga[1], f) for f in fields}
            rf = {f: rgl.result(gf[1], f) for f in fields}
            hits = int(ra[IS_HIT].sum())
            print(f"format {fmt} step {step}: {hits} hits")
            assert hits > 5000
            for f in fields:
                assert ra[f].tobytes() == rf[f].tobytes(), (step, f)
    finally:
        rgl.set_bvh_format(default_fmt)


def skin(rest, weights, bones, mats):
    """FP64 linear blend skinning: sum_k w_k (M[b_k] * v)."""
    out = np.zeros((len(rest), 3))
    for k in range(4):
        m = mats[bones[:, k]]  # (N, 3, 4)
        p = np.einsum("nij,nj->ni", m[:, :, :3], rest) + m[:, :, 3]
        out += weights[:, k:k + 1] * p
    return out


def test_skinned_scene_matches_oracle(rgl_lib, oracle):
    rgl = rgl_lib
    rest_v, idx = S.grid_mesh(30.0, 30.0, 48, 48)
    rest_v = rest_v.astype(np.float32)
    n = len(rest_v)
    rng = np.random.default_rng(11)
    # Three bones along x; each vertex blends the two nearest with weights summing to one.
    centers = np.array([-10.0, 0.0, 10.0])
    dist = np.abs(rest_v[:, :1] - centers[None, :])
    order = np.argsort(dist, axis=1)
    bones = np.zeros((n, 4), np.int32)
    bones[:, :2] = order[:, :2]
    w0 = np.clip(1.0 - dist[np.arange(n), order[:, 0]] / 10.0, 0.5, 1.0)
    weights = np.zeros((n, 4), np.float32)
    weights[:, 0] = w0
    weights[:, 1] = 1.0 - w0
    restposes = np.stack([S.f32(S.translation(-c, 0.0, 0.0)) for c in centers])  # bind: bone frame at its center
    mesh = rgl.mesh_create(rest_v, idx)
    rgl.mesh_set_bone_weights(mesh, weights, bones)
    rgl.mesh_set_restposes(mesh, restposes)
    ent = rgl.entity_create(mesh)
    rgl.entity_set_id(ent, 77)
    rgl.entity_set_transform(ent, S.f32(S.translation(0.0, 20.0, 4.0)))
    rays = S.spinning_lidar(48, 720, -40.0, 20.0)
    pose = S.f32(S.translation(0.0, 0.0, 1.8))
    case = Case(S.SceneSpec(), rays, pose=pose, ranges=np.array([[0.1, 100.0]], np.float32),
                velocity=((0.0, 0.0, 0.0), (0.0, 0.0, 0.0)))
    fields = [XYZ, DIST, ENT, IS_HIT, NORMAL, ABS_V]
    g = Graph(rgl, case, fields)
    dt = 0.1
    prev_skinned = None
    for frame in range(3):
        rgl.scene_set_time(None, int(round(frame * dt * SECOND)))
        poses = np.stack([S.f32(S.trs((c, 0.0, 0.6 * frame * (k - 1)), (15.0 * frame * (k - 1), 0, 0), (1, 1, 1)))
                          for k, c in enumerate(centers)])
        rgl.entity_set_pose_world(ent, poses)
        gpu = g.run()
        mats = np.stack([S.mul(poses[b].astype(np.float64), restposes[b].astype(np.float64)) for b in range(3)])
        skinned = skin(rest_v.astype(np.float64), weights.astype(np.float64), bones, mats)
        spec = S.SceneSpec()
        spec.add_instance(spec.add_mesh(skinned.astype(np.float32), idx), S.translation(0.0, 20.0, 4.0), entity_id=77)
        disp = {0: skinned - (prev_skinned if prev_skinned is not None else rest_v)}
        orc = oracle.raytrace(spec, case.world_rays(), ray_origin_to_world=pose, ranges=case.ranges,
                              scene_dt=dt if frame else 0.0, prev_transforms=np.stack([S.translation(0.0, 20.0, 4.0)]),
                              displacements=disp, fields=fields)
        ok = mismatch_ok(gpu, orc, len(rays))
        print(f"skinned frame {frame}: {int(ok.sum())} hits compared")
        assert ok.sum() > 1500
        d_g, d_o = gpu[DIST][ok].astype(np.float64), orc[DIST][ok].astype(np.float64)
        # Skinned vertices are FP32 on the GPU and FP64 in the oracle (a few ulp of |x| < 32 apart).
        assert np.all(np.abs(d_g - d_o) <= np.maximum(1e-3, 1e-5 * d_o))
        if frame:
            np.testing.assert_allclose(gpu[ABS_V][ok], orc[ABS_V][ok], atol=1e-3, rtol=1e-4)
            assert (np.linalg.norm(orc[ABS_V][ok], axis=1) > 1.0).sum() > 500
        prev_skinned = skinned


def test_animation_argument_errors(rgl_lib):
    rgl = rgl_lib
    mesh = rgl.mesh_create(S.CUBE_VERTICES, S.CUBE_INDICES)
    ent = rgl.entity_create(mesh)
    v = np.zeros((7, 3), np.float32)
    assert rgl.call("rgl_entity_apply_external_animation", ent, v.ctypes.data, 7) == rgl.RGL_INVALID_ARGUMENT
    assert "vertex counts do not match" in rgl.last_error()
    pose = np.stack([S.f32(S.identity())] * 2)
    assert rgl.call("rgl_entity_set_pose_world", ent, pose.ctypes.data, 2) == rgl.RGL_INVALID_ARGUMENT
    assert "no bone weights" in rgl.last_error()
    w = np.ascontiguousarray(np.concatenate([np.tile(np.float32([1, 0, 0, 0]), (7, 1)).view(np.uint32),
                                             np.zeros((7, 4), np.uint32)], axis=1))
    assert rgl.call("rgl_mesh_set_bone_weights", mesh, w.ctypes.data, 7) == rgl.RGL_INVALID_ARGUMENT
    assert "bone weights count" in rgl.last_error()
    bones = np.zeros((8, 4), np.int32)
    bones[5, 2] = 2
    rgl.mesh_set_bone_weights(mesh, np.tile([1.0, 0, 0, 0], (8, 1)), bones)
    assert rgl.call("rgl_entity_set_pose_world", ent, pose.ctypes.data, 2) == rgl.RGL_INVALID_ARGUMENT
    assert "no restposes" in rgl.last_error()
    rgl.mesh_set_restposes(mesh, pose)
    three = np.stack([S.f32(S.identity())] * 3)
    assert rgl.call("rgl_entity_set_pose_world", ent, three.ctypes.data, 3) == rgl.RGL_INVALID_ARGUMENT
    assert "bones count do not match" in rgl.last_error()
    assert rgl.call("rgl_entity_set_pose_world", ent, pose.ctypes.data, 2) == rgl.RGL_INVALID_ARGUMENT
    assert "refer to bone 2" in rgl.last_error()
    bones[5, 2] = -1
    neg = np.ascontiguousarray(np.concatenate([np.tile(np.float32([1, 0, 0, 0]), (8, 1)).view(np.uint32),
                                               bones.view(np.uint32)], axis=1))
    assert rgl.call("rgl_mesh_set_bone_weights", mesh, neg.ctypes.data, 8) == rgl.RGL_INVALID_ARGUMENT
    # A valid animation still works after the errors, and the entity keeps raytracing.
    bones[5, 2] = 1
    rgl.mesh_set_bone_weights(mesh, np.tile([1.0, 0, 0, 0], (8, 1)), bones)
    rgl.entity_set_pose_world(ent, pose)
    case = Case(cube_scene(), S.f32(S.identity()[None]))
    g = Graph(rgl, case, [XYZ, IS_HIT])
    res = g.run()
    assert res[IS_HIT][0] == 1 and abs(res[XYZ][0][2] - 1.0) < 1e-6


def test_animated_scene_batched_equals_solo(rgl_lib):
    from tests.test_gpu_batch import assert_same
    rgl = rgl_lib
    v0, idx = cloth(5, 0.0)
    m = rgl.mesh_create(v0, idx)
    e = rgl.entity_create(m)
    rgl.entity_set_transform(e, S.f32(S.translation(0.0, 18.0, 2.0)))
    rays = S.spinning_lidar(32, 512, -25.0, 10.0)
    fields = [XYZ, DIST, ENT, IS_HIT, NORMAL, ABS_V]
    mk = lambda: Graph(rgl, Case(S.SceneSpec(), rays, pose=S.f32(S.translation(0, 0, 1.6)),
                                 velocity=((0, 0, 0), (0, 0, 0))), fields)
    solo, batched = mk(), mk()
    for frame in range(4):
        rgl.scene_set_time(None, frame * SECOND // 20)
        rgl.entity_apply_external_animation(e, cloth(5, 0.05 * frame)[0])
        a = solo.run()
        rgl.graph_run_batch([batched.rays])
        b = {f: rgl.result(batched.out, f) for f in fields}
        assert_same(a, b, f"frame {frame}")
        assert int(a[IS_HIT].sum()) > 1000
