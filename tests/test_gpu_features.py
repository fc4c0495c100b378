"""GPU parity of the remaining per-ray features and point stages against the FP64 oracle / numpy restatements,
plus API error semantics of the reference (test/src/graph/nodes/*Test.cpp invalid_argument/valid_arguments pairs).
"""
import ctypes as C

import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S
from tests.helpers import Case, Graph, cube_scene, identity_ray, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

XYZ, INTEN, INTEN_U8, IS_HIT, RAY_IDX, ENT, DIST, AZ, EL, RING, RET, TS64, TS32 = 1, 2, 3, 4, 6, 7, 8, 9, 10, 11, 12, 13, 14
ABS_V, REL_V, RAD, NORMAL, INC, RETRO = 15, 16, 17, 22, 23, 25
FORMAT = 13842


def small_lidar(ch=16, az=360, el=(-15.0, 15.0)):
    return S.spinning_lidar(ch, az, *el)


def oblique_cube_uvs():
    """Texture coordinates that are an affine function of the cube vertex position, with no face on which a
    coordinate is constant: with CUBE_UVS, every face holds u or v at exactly 0 or 1, where the FP32 interpolation
    w*a + u*a + v*a lands on either side of the wrap and point sampling may pick the texel at the far edge, so the
    oracle (rightly) flags every textured hit as a texel-edge case."""
    v = S.CUBE_VERTICES.astype(np.float64)
    return np.stack([0.5 + 0.2 * v[:, 0] + 0.13 * v[:, 1] + 0.07 * v[:, 2],
                     0.5 + 0.05 * v[:, 0] - 0.17 * v[:, 1] + 0.21 * v[:, 2]], 1).astype(np.float32)


def textured_scene(seed=3):
    rng = np.random.default_rng(seed)
    spec = S.config1_scene()
    cube_uv = spec.add_mesh(S.CUBE_VERTICES, S.CUBE_INDICES, oblique_cube_uvs())
    spec.textures.append(rng.integers(0, 256, (64, 64)).astype(np.uint8))
    for k in range(8):
        a = k * np.pi / 4
        spec.add_instance(cube_uv, S.trs((12 * np.cos(a), 12 * np.sin(a), 2.0), (0, 0, 20 * k), (1.5, 1.5, 2.0)),
                          entity_id=100 + k, texture=0, laser_retro=0.5 + k)
    return spec


GEOM, DIST_TIE, INT_TIE, TEXEL = 1, 2, 4, 8  # oracle ambiguity bits


def mismatch_ok(gpu, orc, n, R=1, bits=GEOM | DIST_TIE | INT_TIE | TEXEL):
    amb = np.repeat((orc["ambiguous"] & bits) != 0, R)
    mism = (gpu[IS_HIT] != orc[IS_HIT]) | ((gpu[IS_HIT] == 1) & (gpu[ENT] != orc[ENT]))
    assert not (mism & ~amb).any(), np.nonzero(mism & ~amb)[0][:10]
    assert mism.sum() <= max(1, int(1e-5 * n))
    return (gpu[IS_HIT] == 1) & ~mism


def test_texture_intensity_laser_retro_and_u8(rgl_lib, oracle):
    spec = textured_scene()
    case = Case(spec, small_lidar(32, 720, (-20, 10)), pose=S.f32(S.translation(0, 0, 1.8)),
                ranges=np.array([[0.1, 100.0]], np.float32), default_intensity=50.0)
    fields = [XYZ, DIST, ENT, IS_HIT, INTEN, INTEN_U8, RETRO, INC, NORMAL]
    gpu = run_gpu(rgl_lib, case, fields)
    orc = run_oracle(oracle, case, fields=fields)
    ok = mismatch_ok(gpu, orc, len(case.rays), bits=GEOM)
    textured = ok & (gpu[ENT] >= 100)
    assert textured.sum() > 200
    # Point sampling may pick the neighbouring texel only where the oracle flags a texel edge; elsewhere the
    # intensity (texel * cos incidence) agrees to FP32 accuracy of the incidence factor.
    edge = (orc["ambiguous"] & TEXEL) != 0
    print(f"texel-edge beams: {int(edge.sum())} of {int(textured.sum())} textured hits")
    assert (textured & edge).sum() <= 0.01 * textured.sum()
    sel = ok & ~edge
    d = np.abs(gpu[INTEN][sel].astype(np.float64) - orc[INTEN][sel])
    assert np.all(d <= 1e-4 * orc[INTEN][sel] + 1e-3), d.max()
    np.testing.assert_array_equal(gpu[RETRO][ok], orc[RETRO][ok])
    np.testing.assert_allclose(gpu[INC][ok], orc[INC][ok], atol=2e-4)
    u8_expected = np.where(gpu[INTEN] < 255, np.round(gpu[INTEN]), 255).astype(np.uint8)
    np.testing.assert_array_equal(gpu[INTEN_U8], np.where(gpu[IS_HIT] == 1, u8_expected, 0))


@pytest.mark.parametrize("mode", ["FIRST_LAST", "STRONGEST_SECOND_STRONGEST", "LAST", "FIRST_SECOND"])
def test_multi_return_beams_match_oracle(rgl_lib, oracle, mode):
    from robotecgpulidar_amd import rgl
    rm = getattr(rgl, "RGL_RETURN_" + mode)
    R = rm >> 24
    spec = textured_scene()
    div = (2e-3, 2e-3)  # 58 samples per beam (both divergences > 0)
    case = Case(spec, small_lidar(8, 360, (-10, 5)), pose=S.f32(S.translation(0, 0, 1.8)),
                ranges=np.array([[0.1, 100.0]], np.float32), return_mode=rm, divergence=div, default_intensity=30.0)
    fields = [XYZ, DIST, ENT, IS_HIT, INTEN, RET]
    gpu = run_gpu(rgl_lib, case, fields)
    orc = run_oracle(oracle, case, fields=fields)
    n = len(case.rays)
    assert len(gpu[IS_HIT]) == n * R
    types = [(rm >> (8 * r)) & 0xFF for r in range(R)]
    np.testing.assert_array_equal(gpu[RET], np.tile(np.array(types, np.uint8), n))
    amb = orc["ambiguous"]
    # Hit masks agree on every beam without a geometric ambiguity (ties only change which sample is chosen).
    geom = np.repeat((amb & GEOM) != 0, R)
    assert np.all((gpu[IS_HIT] == orc[IS_HIT]) | geom)
    # FIRST / LAST / STRONGEST select an extreme, which is stable under near ties (only the chosen sample may
    # change, by no more than the tie): compared on every beam without a geometric ambiguity (and, for
    # intensities, without a sample at a texel edge, where point sampling jumps). SECOND and
    # SECOND_STRONGEST depend on the order of strict comparisons (nodeKernels.cu:369-397) and jump on a near tie:
    # compared on beams whose reduction has no near tie of the compared quantity. The 58 samples lie on symmetric
    # rings, so on a plane pairs of samples have equal distances and, on one texel, near-equal intensities: most
    # beams carry a tie, and the tracker rule itself is pinned in closed form on planes (tests/test_pins.py). On
    # every beam SECOND and SECOND_STRONGEST are also checked against the oracle's range over all tie resolutions.
    for r, t in enumerate(types):
        hit = (gpu[IS_HIT][r::R] == 1) & (orc[IS_HIT][r::R] == 1)
        if t in (2, 4, 3):  # LAST, FIRST, SECOND
            sel = hit & ((amb & (GEOM | (DIST_TIE if t == 3 else 0))) == 0)
            d_g, d_o = gpu[DIST][r::R][sel].astype(np.float64), orc[DIST][r::R][sel].astype(np.float64)
            assert np.all(np.abs(d_g - d_o) <= np.maximum(1e-3, 1e-5 * d_o))
        if t in (1, 5):  # STRONGEST, SECOND_STRONGEST
            sel = hit & ((amb & (GEOM | TEXEL | (INT_TIE if t == 5 else 0))) == 0)
            i_g, i_o = gpu[INTEN][r::R][sel].astype(np.float64), orc[INTEN][r::R][sel]
            assert np.all(np.abs(i_g - i_o) <= 1e-4 * i_o + 2e-3), np.abs(i_g - i_o).max()
        print(f"return {r} type {t}: {int(sel.sum())} of {n} beams compared")
        assert sel.sum() > (n // 4 if t in (1, 2, 4) else 100)
        if t in (3, 5):
            # Every beam: the value lies within the oracle's range of what the reduction can return when each
            # near-tied comparison may go either way (oracle second_range), to the same tolerances.
            lo, hi = (orc["second_range"][:, 0], orc["second_range"][:, 1]) if t == 3 else \
                (orc["second_range"][:, 2], orc["second_range"][:, 3])
            skip = GEOM | (TEXEL if t == 5 else 0)
            sel = hit & ((amb & skip) == 0) & np.isfinite(lo) & np.isfinite(hi)
            v = (gpu[DIST] if t == 3 else gpu[INTEN])[r::R][sel].astype(np.float64)
            lo, hi = lo[sel].astype(np.float64), hi[sel].astype(np.float64)
            tol_lo = np.maximum(1e-3, 1e-5 * np.abs(lo)) if t == 3 else 1e-4 * np.abs(lo) + 2e-3
            tol_hi = np.maximum(1e-3, 1e-5 * np.abs(hi)) if t == 3 else 1e-4 * np.abs(hi) + 2e-3
            assert np.all((v >= lo - tol_lo) & (v <= hi + tol_hi)), np.nonzero(~((v >= lo - tol_lo) & (v <= hi + tol_hi)))
            print(f"return {r} type {t}: {int(sel.sum())} of {n} beams within the tie-resolution range")
            assert sel.sum() > 0.7 * n
    nonhit_both = (gpu[IS_HIT] == 0) & (orc[IS_HIT] == 0)
    np.testing.assert_array_equal(gpu[DIST][nonhit_both], orc[DIST][nonhit_both])


def test_distortion_and_velocities_match_oracle(rgl_lib, oracle):
    """Velocity distortion (optixPrograms.cu:61-70) and hit-point velocities (:169-212) with moving entities."""
    from robotecgpulidar_amd import rgl
    spec = S.config1_scene()
    rays = small_lidar(16, 720, (-15, 5))
    offsets = np.tile(np.arange(720, dtype=np.float64) * (100.0 / 720), 16).astype(np.float32)
    lin, ang = (10.0, 2.0, 0.0), (0.0, 0.0, 0.5)
    pose = S.f32(S.translation(0, 0, 1.8))
    case = Case(spec, rays, pose=pose, ranges=np.array([[0.1, 100.0]], np.float32), time_offsets=offsets,
                distortion=True, velocity=(lin, ang))
    fields = [XYZ, DIST, ENT, IS_HIT, ABS_V, REL_V, RAD, TS64, TS32]
    # Two frames 50 ms apart; every box moves between them.
    rgl.check("rgl_scene_set_time", None, C.c_uint64(1_000_000_000))
    _, entities, _ = S.load_into_rgl(spec)
    prev = [i.transform.copy() for i in spec.instances]
    rgl.check("rgl_scene_set_time", None, C.c_uint64(1_050_000_000))
    moved = []
    for k, (e, inst) in enumerate(zip(entities, spec.instances)):
        t = inst.transform.copy()
        if k > 0:
            t[:, 3] += np.float32([0.3, -0.2, 0.0])
        rgl.entity_set_transform(e, t)
        moved.append(t)
    for inst, t in zip(spec.instances, moved):
        inst.transform = t
    gpu = Graph(rgl_lib, case, fields).run()
    orc = run_oracle(oracle, case, fields=fields) if False else oracle.raytrace(
        spec, case.world_rays(), ray_origin_to_world=pose, ranges=case.ranges, time_offsets=offsets, distortion=True,
        lin_vel=lin, ang_vel=ang, scene_dt=0.05, prev_transforms=np.stack(prev), fields=fields)
    ok = mismatch_ok(gpu, orc, len(rays))
    assert ok.sum() > 1000
    d_g, d_o = gpu[DIST][ok].astype(np.float64), orc[DIST][ok].astype(np.float64)
    assert np.all(np.abs(d_g - d_o) <= np.maximum(1e-3, 1e-5 * d_o))
    np.testing.assert_allclose(gpu[ABS_V][ok], orc[ABS_V][ok], atol=2e-3)
    np.testing.assert_allclose(gpu[REL_V][ok], orc[REL_V][ok], atol=2e-3)
    np.testing.assert_allclose(gpu[RAD][ok], orc[RAD][ok], atol=2e-3)
    moving = ok & (gpu[ENT] > 1)
    assert np.allclose(gpu[ABS_V][moving], [6.0, -4.0, 0.0], atol=1e-3)  # 0.3 m / 0.05 s, -0.2 m / 0.05 s
    np.testing.assert_array_equal(gpu[TS64], orc[TS64])
    np.testing.assert_array_equal(gpu[TS32], orc[TS32])


def test_beam_shared_fields(rgl_lib, oracle):
    rays = S.make_lidar3d_rays(180, 90, 5, 5)
    ring = np.arange(7, dtype=np.int32) + 3
    n = (len(rays) // 7) * 7
    case = Case(cube_scene(S.trs((0, 0, 6), (0, 0, 0), (3, 3, 1))), rays[:n], ring_ids=ring)
    fields = [XYZ, RAY_IDX, RING, AZ, EL]
    gpu = run_gpu(rgl_lib, case, fields)
    orc = run_oracle(oracle, case, fields=fields)
    np.testing.assert_array_equal(gpu[RAY_IDX], np.arange(n, dtype=np.uint32))
    np.testing.assert_array_equal(gpu[RING], orc[RING])
    np.testing.assert_allclose(gpu[AZ], orc[AZ], atol=2e-6)
    np.testing.assert_allclose(gpu[EL], orc[EL], atol=2e-6)


def _noise_graph(rgl, spec, rays, noise_kind, params, fields):
    S.load_into_rgl(spec)
    n = rgl.Node
    src = rgl.node_rays_from_mat3x4f(n(), rays)
    pose = rgl.node_rays_transform(n(), S.f32(S.trs((0.5, -1.0, 1.8), (0, 0, 30))))
    nodes = [src, pose]
    noise = n()
    if noise_kind == "ray":
        rgl.check("rgl_node_gaussian_noise_angular_ray", noise.ref, *params)
        nodes += [noise, rgl.node_raytrace(n())]
    else:
        nodes.append(rgl.node_raytrace(n()))
        if noise_kind == "hitpoint":
            rgl.check("rgl_node_gaussian_noise_angular_hitpoint", noise.ref, *params)
        else:
            rgl.check("rgl_node_gaussian_noise_distance", noise.ref, *params)
        nodes.append(noise)
    out = rgl.node_points_yield(n(), fields)
    nodes.append(out)
    rgl.chain(*nodes)
    rgl.check("rgl_node_gaussian_noise_set_seed", noise.h, C.c_uint64(0xABCDEF))
    return src, noise, out, nodes[1]


@pytest.mark.parametrize("kind", ["distance", "hitpoint", "ray"])
def test_noise_nodes_replay_on_cpu(rgl_lib, oracle, kind):
    rgl = rgl_lib
    spec = cube_scene(S.trs((0, 0, 0), (0, 0, 0), (20, 20, 20)))
    rays = S.make_lidar3d_rays(360, 180, 3, 3)
    fields = [XYZ, DIST, IS_HIT]
    params = {"distance": (0.05, 0.02, 0.001), "hitpoint": (0.0, 0.01, 3), "ray": (0.0, 0.01, 2)}[kind]
    src, noise, out, pose = _noise_graph(rgl, spec, rays, kind, params, fields)
    rgl.graph_run(src)
    got = {f: rgl.result(out, f) for f in fields}
    seed, run, tag = C.c_uint64(), C.c_uint32(), C.c_uint32()
    rgl.check("rgl_node_gaussian_noise_get_stream", noise.h, C.byref(seed), C.byref(run), C.byref(tag))
    normals = oracle.normals(seed.value, run.value - 1, tag.value, len(rays))
    pose_m = S.f32(S.trs((0.5, -1.0, 1.8), (0, 0, 30)))
    look_at = np.linalg.inv(np.vstack([pose_m.astype(np.float64), [0, 0, 0, 1]]))[:3]
    world = np.einsum("ij,njk->nik", pose_m[:, :3], rays).astype(np.float32)
    world[:, :, 3] += pose_m[:, 3]
    if kind == "ray":
        noisy = oracle.noise_angular_rays(world, normals, params[0], params[1], params[2], look_at).astype(np.float32)
        ref = oracle.raytrace(spec, noisy, ray_origin_to_world=pose_m, fields=fields)
        ok = (got[IS_HIT] == 1) & (ref[IS_HIT] == 1)
        assert ok.mean() > 0.99
        np.testing.assert_allclose(got[DIST][ok], ref[DIST][ok], rtol=1e-5, atol=1e-4)
        np.testing.assert_allclose(got[XYZ][ok], ref[XYZ][ok], atol=2e-4)
        return
    clean = oracle.raytrace(spec, world, ray_origin_to_world=pose_m, fields=fields)
    assert np.all(clean[IS_HIT] == 1)
    if kind == "distance":
        xyz, dist = oracle.noise_distance(clean[XYZ], clean[DIST], normals, *params, look_at)
    else:
        xyz, dist = oracle.noise_angular_hitpoint(clean[XYZ], normals, params[0], params[1], params[2], look_at)
    np.testing.assert_allclose(got[DIST], dist, rtol=2e-6, atol=2e-5)
    np.testing.assert_allclose(got[XYZ], xyz, atol=1e-4)


def test_points_transform_and_format(rgl_lib):
    rgl = rgl_lib
    spec = S.config1_scene()
    S.load_into_rgl(spec)
    lidar = S.config1_lidar()
    tf = S.f32(S.trs((1.0, 2.0, 3.0), (10, 20, 30)))
    n = rgl.Node
    src = rgl.node_rays_from_mat3x4f(n(), lidar.rays[::7])
    pose = rgl.node_rays_transform(n(), lidar.pose)
    rt = rgl.node_raytrace(n())
    comp = rgl.node_points_compact_by_field(n(), IS_HIT)
    yld = rgl.node_points_yield(n(), [XYZ, DIST, ENT])
    ptf = rgl.node_points_transform(n(), tf)
    fmt_fields = [XYZ, 1026, DIST, ENT]  # PADDING_32 between XYZ and DISTANCE
    fmt = rgl.node_points_format(n(), fmt_fields)
    rgl.chain(src, pose, rt, comp, yld, ptf, fmt)
    rgl.graph_run(src)
    xyz = rgl.result(yld, XYZ)
    moved = rgl.result(ptf, XYZ)
    expected = (xyz.astype(np.float64) @ tf[:, :3].T.astype(np.float64) + tf[:, 3]).astype(np.float32)
    np.testing.assert_allclose(moved, expected, atol=1e-4)
    raw = rgl.result(fmt, FORMAT)
    rec = np.dtype([("xyz", np.float32, 3), ("pad", np.uint32), ("d", np.float32), ("e", np.int32)])
    assert raw.shape == (len(xyz), rec.itemsize)
    packed = raw.view(rec).reshape(-1)
    np.testing.assert_array_equal(packed["xyz"], moved)
    np.testing.assert_array_equal(packed["pad"], 0)
    np.testing.assert_array_equal(packed["d"], rgl.result(yld, DIST))
    np.testing.assert_array_equal(packed["e"], rgl.result(yld, ENT))


def test_api_error_semantics(rgl_lib):
    rgl = rgl_lib
    node = rgl.Node()
    assert rgl.call("rgl_node_raytrace", None, None) == rgl.RGL_INVALID_ARGUMENT
    assert "node != nullptr" in rgl.last_error()
    assert rgl.call("rgl_node_points_compact_by_field", node.ref, XYZ) == rgl.RGL_INVALID_ARGUMENT
    rgl.node_points_compact_by_field(node, IS_HIT)
    assert rgl.call("rgl_graph_run", node.h) == rgl.RGL_INVALID_PIPELINE
    assert "looked for" in rgl.last_error() and "IPointsNode" in rgl.last_error()
    assert rgl.call("rgl_node_raytrace_configure_return_mode", node.h, rgl.RGL_RETURN_FIRST) == rgl.RGL_INVALID_API_OBJECT
    bad = C.c_void_p(0x1234)
    assert rgl.call("rgl_graph_run", bad) == rgl.RGL_INVALID_API_OBJECT
    rt = rgl.node_raytrace(rgl.Node())
    assert rgl.call("rgl_node_raytrace_configure_beam_divergence", rt.h, 1.0, 0.0) == rgl.RGL_INVALID_ARGUMENT
    assert rgl.call("rgl_node_raytrace_configure_default_intensity", rt.h, -1.0) == rgl.RGL_INVALID_ARGUMENT
    r = np.array([[2.0, 1.0]], np.float32)
    assert rgl.call("rgl_node_rays_set_range", rgl.Node().ref, r.ctypes.data, 1) == rgl.RGL_INVALID_ARGUMENT
    assert "minimum range must be lower" in rgl.last_error()
    assert rgl.call("rgl_node_points_radar_postprocess", rgl.Node().ref, None, 0, *([1.0] * 7)) == rgl.RGL_INVALID_ARGUMENT
    assert rgl.call("rgl_node_points_radar_track_objects", rgl.Node().ref, *([1.0] * 7)) == rgl.RGL_NOT_IMPLEMENTED
    # Results of a node that never ran.
    y = rgl.node_points_yield(rgl.Node(), [XYZ])
    assert rgl.call("rgl_graph_get_result_size", y.h, XYZ, None, None) == rgl.RGL_INVALID_PIPELINE
    # Entity id limits (src/scene/Entity.cpp:45-48).
    e = rgl.entity_create(rgl.mesh_create(S.CUBE_VERTICES, S.CUBE_INDICES))
    assert rgl.call("rgl_entity_set_id", e, 0) == rgl.RGL_INVALID_ARGUMENT
    assert rgl.call("rgl_entity_set_id", e, 1 << 28) == rgl.RGL_INVALID_ARGUMENT
    rgl.check("rgl_entity_set_id", e, (1 << 28) - 1)
    rgl.check("rgl_entity_destroy", e)
    assert rgl.call("rgl_entity_destroy", e) == rgl.RGL_INVALID_API_OBJECT
    alive = C.c_bool(True)
    rgl.check("rgl_entity_is_alive", e, C.byref(alive))
    assert not alive.value


def test_graph_reruns_after_reconfiguration(rgl_lib):
    """Changing rays, mask and return mode between runs is reflected in the next run (create-or-update nodes)."""
    rgl = rgl_lib
    S.load_into_rgl(cube_scene(S.translation(0, 0, 5)))
    case = Case(S.SceneSpec(), identity_ray())
    g = Graph(rgl, case, [DIST, IS_HIT])
    assert g.run()[DIST][0] == 4.0
    rgl.node_rays_from_mat3x4f(g.rays, S.f32(S.translation(0, 0, 1)[None]))
    assert g.run()[DIST][0] == 3.0
    m = np.zeros(1, np.int8)
    rgl.check("rgl_node_raytrace_configure_mask", g.raytrace.h, m.ctypes.data, 1)
    r = g.run()
    assert r[IS_HIT][0] == 0 and r[DIST][0] == np.inf
    rgl.check("rgl_node_raytrace_configure_return_mode", g.raytrace.h, rgl.RGL_RETURN_FIRST_LAST)
    m[0] = 1
    rgl.check("rgl_node_raytrace_configure_mask", g.raytrace.h, m.ctypes.data, 1)
    r = g.run()
    assert list(r[DIST]) == [3.0, 3.0]
