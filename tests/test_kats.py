"""Known-answer tests restated from the reference's own test suite, run on both the FP64 oracle (CPU) and the
HIP library (GPU, through the C-ABI). Expected values are the reference tests' closed forms:
  test/src/graph/DistanceFieldTest.cpp:84-276,329-362, test/src/apiReadmeExample.cpp:54-55,
  test/src/scene/incidentAngleTest.cpp:16-82, test/src/scene/entityIdTest.cpp:28-112,
  test/src/graph/MaskRaysTest.cpp:58-100, test/src/graph/nodes/RaytraceNodeTest.cpp:109-326.
"""
import math

import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S
from tests.helpers import Case, DISTANCE_PARAMS, FLT_MAX, cube_scene, f32sub, identity_ray, run_gpu, run_oracle

XYZ, IS_HIT, DIST, ENT, NORMAL, INC, INTEN = 1, 4, 8, 7, 22, 23, 2

BACKENDS = [pytest.param("oracle", id="oracle"), pytest.param("gpu", id="gpu", marks=pytest.mark.gpu)]


@pytest.fixture
def run(request, oracle):
    backend = request.param
    if backend == "oracle":
        return lambda case, fields, **kw: run_oracle(oracle, case, fields=fields)
    rgl = request.getfixturevalue("rgl_lib")
    return lambda case, fields, **kw: run_gpu(rgl, case, fields, **kw)


def length32(v):
    v = np.asarray(v, np.float32)
    return np.sqrt(np.float32(v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2]).astype(np.float32)


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
@pytest.mark.parametrize("z", DISTANCE_PARAMS)
def test_distance_single_object(run, z):
    case = Case(cube_scene(S.translation(0, 0, np.float32(z))), identity_ray(),
                ranges=np.array([[0.0, np.float32(z) + 10.0]], np.float32))
    r = run(case, [DIST, XYZ])
    assert r[DIST][0] == f32sub(z, 1.0)
    assert r[DIST][0] == length32(r[XYZ][0])


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
def test_distance_ray_origin_on_cube_face(run):
    case = Case(cube_scene(S.translation(0, 0, 1.0)), identity_ray(), ranges=np.array([[0.0, 11.0]], np.float32))
    r = run(case, [DIST, XYZ])
    assert r[DIST][0] == 2.0


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
def test_distance_ray_origin_inside_cube(run):
    z = np.float32(-0.001)
    case = Case(cube_scene(S.translation(0, 0, z)), identity_ray(), ranges=np.array([[0.0, z + 10.0]], np.float32))
    r = run(case, [DIST, XYZ])
    assert r[DIST][0] == np.float32(np.float32(1.0) + z)


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
def test_nonhit_for_object_off_ray_path(run):
    rays = S.f32(np.stack([S.identity(), S.rotation_rad(math.pi, 0, 0)]))
    case = Case(cube_scene(S.translation(0, 55.5, 0)), rays, ranges=np.array([[0.0, 10000.0]], np.float32))
    r = run(case, [DIST, XYZ])
    assert np.all(r[DIST] == np.inf)


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
@pytest.mark.parametrize("z", DISTANCE_PARAMS)
def test_nonhit_beyond_max_range(run, z):
    z = np.float32(z)
    case = Case(cube_scene(S.translation(0, 0, z)), identity_ray(), ranges=np.array([[0.0, z - 2.0]], np.float32))
    assert run(case, [DIST])[DIST][0] == np.inf


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
@pytest.mark.parametrize("z", DISTANCE_PARAMS)
def test_nonhit_below_min_range(run, z):
    z = np.float32(z)
    case = Case(cube_scene(S.translation(0, 0, z)), identity_ray(), ranges=np.array([[z + 2.0, z + 3.0]], np.float32))
    assert run(case, [DIST])[DIST][0] == np.inf


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
@pytest.mark.parametrize("z", DISTANCE_PARAMS)
def test_object_in_range_behind_object_below_min_range(run, z):
    """An object closer than the minimum range still occludes (src/gpu/optixPrograms.cu:132-136)."""
    z = np.float32(z)
    case = Case(cube_scene(S.translation(0, 0, z), S.translation(0, 0, z + 10.0)), identity_ray(),
                ranges=np.array([[z + 2.0, z + 12.0]], np.float32))
    assert run(case, [DIST])[DIST][0] == np.inf


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
@pytest.mark.parametrize("z", DISTANCE_PARAMS)
def test_distances_for_various_ray_angles(run, z):
    zf = np.float32(z)
    max_angle = math.atan(1.0 / (zf - 1.0))
    angles = np.random.default_rng(int(z * 1000)).uniform(-max_angle, max_angle, 100)
    rays = S.f32(np.stack([S.rotation_rad(a, 0, 0) for a in angles]))
    case = Case(cube_scene(S.translation(0, 0, zf)), rays, ranges=np.array([[0.0, zf + 10.0]], np.float32))
    r = run(case, [DIST, XYZ])
    expected = (zf - 1.0) / np.cos(angles)
    np.testing.assert_allclose(length32(r[XYZ]), r[DIST], atol=2e-4)
    np.testing.assert_allclose(expected, r[DIST], atol=2e-4)


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
def test_distance_from_ray_beginning(run):
    case = Case(cube_scene(S.translation(0, 0, 10.0)), S.f32(S.translation(0, 0, 5.0)[None]),
                ranges=np.array([[0.0, 20.0]], np.float32))
    r = run(case, [DIST, XYZ])
    assert r[DIST][0] == 4.0
    assert r[DIST][0] == length32(r[XYZ][0]) - 5.0


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
def test_readme_example(run):
    case = Case(cube_scene(S.translation(0, 0, 5)), identity_ray())
    r = run(case, [XYZ])
    assert r[XYZ].shape == (1, 3)
    assert r[XYZ][0, 2] == pytest.approx(4.0, abs=1e-6)


INCIDENT_PARAMS = [(sp, wo, sr, wr)
                   for sp in [(0, 0, 0), (1, 1, 1)]
                   for wo in [(0, 0, 0), (1, 1, 1)]
                   for sr in [(5, 0, 0), (0, -15, 0), (-25, 25, 0)]
                   for wr in [(-10, 0, 0), (0, 20, 0), (30, -40, 0)]]


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
@pytest.mark.parametrize("sensor_pos,wall_off,sensor_rot,wall_rot", INCIDENT_PARAMS)
def test_incident_angle_and_normal(run, sensor_pos, wall_off, sensor_rot, wall_rot):
    wall_pos = np.add((0, 0, 10.0), wall_off)
    k = math.pi / 180
    ray_dir = S.rotation_rad(*(np.multiply(sensor_rot, k)))[:, :3] @ np.array([0, 0, 1.0])
    normal = S.rotation_rad(*(np.multiply(wall_rot, k)))[:, :3] @ np.array([0, 0, -1.0])
    expected_angle = math.acos(-normal.dot(ray_dir))
    case = Case(cube_scene(S.trs(wall_pos, wall_rot, (200, 200, 0.1))),
                S.make_grid_of_parallel_rays((-1, -1), (1, 1), (10, 10)), pose=S.f32(S.trs(sensor_pos, sensor_rot)))
    r = run(case, [XYZ, INC, NORMAL, IS_HIT])
    assert np.all(r[IS_HIT] == 1)
    np.testing.assert_allclose(r[NORMAL], np.broadcast_to(normal, r[NORMAL].shape), atol=1e-6)
    np.testing.assert_allclose(r[INC], expected_angle, atol=4e-6)


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
def test_entity_ids(run):
    spec = cube_scene(S.trs((6, -5, 0)), S.trs((6, 0, 0)), S.trs((6, 5, 0)), ids=[1, 2, None])
    case = Case(spec, S.make_lidar3d_rays(360, 180, 0.72, 0.36))
    r = run(case, [XYZ, ENT, IS_HIT])
    hit = r[IS_HIT] != 0
    assert np.all(r[ENT][~hit] == 0)
    y = r[XYZ][:, 1]
    assert hit.sum() > 100
    assert np.all(r[ENT][hit & (y < -2.5)] == 1)
    assert np.all(r[ENT][hit & (y > -2.5) & (y < 2.5)] == 2)
    assert np.all(r[ENT][hit & (y > 2.5)] == 268435455)


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
def test_mask_removes_exactly_masked_hits(run):
    rays = S.make_lidar3d_rays(360, 180, 0.36 * 4, 0.18 * 4)
    spec = cube_scene(S.trs((0, 0, 0), (0, 0, 0), (10, 10, 10)))
    full = run(Case(spec, rays), [XYZ, IS_HIT])
    mask = np.ones(len(rays), np.int8)
    mask[:100] = 0
    masked = run(Case(spec, rays, mask=mask), [XYZ, IS_HIT])
    assert full[IS_HIT].sum() == len(rays)
    assert masked[IS_HIT].sum() == full[IS_HIT].sum() - 100


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
def test_non_hit_near_distance(run):
    rays = S.make_lidar3d_rays(360, 180, 30, 30)
    case = Case(cube_scene(S.identity()), rays, ranges=np.array([[2.0, 3.0]], np.float32), near=12.34, far=0.0)
    r = run(case, [XYZ, IS_HIT, DIST])
    o = rays[:, :, 3].astype(np.float64)
    d = (rays[:, :, 2] + rays[:, :, 3]) - rays[:, :, 3]
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    np.testing.assert_allclose(r[XYZ], o + d * np.float32(12.34), atol=1e-5)
    assert np.all(r[DIST] == np.float32(12.34))
    assert np.all(r[IS_HIT] == 0)


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
def test_non_hit_far_distance(run):
    rays = S.make_lidar3d_rays(360, 180, 30, 30)
    case = Case(S.SceneSpec(), rays, near=0.0, far=13.24)
    r = run(case, [XYZ, IS_HIT, DIST])
    assert np.all(r[DIST] == np.float32(13.24))
    assert np.all(r[IS_HIT] == 0)


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
def test_default_intensity(run):
    rays = S.f32(np.stack([S.identity(), S.translation(3, 0, 0)]))
    r = run(Case(cube_scene(S.identity()), rays, default_intensity=100.0), [IS_HIT, INTEN])
    assert list(r[IS_HIT]) == [1, 0]
    assert r[INTEN][0] == 100.0 and r[INTEN][1] == 0.0


@pytest.mark.parametrize("run", BACKENDS, indirect=True)
def test_uniform_texture_intensity(run):
    """A uniform texel value times cos(0) is returned exactly (test/src/scene/textureTest.cpp:41-88)."""
    spec = S.SceneSpec()
    m = spec.add_mesh(S.CUBE_VERTICES, S.CUBE_INDICES, S.CUBE_UVS)
    spec.textures.append(np.full((4, 4), 100, np.uint8))
    spec.add_instance(m, S.translation(0, 0, 5), texture=0)
    r = run(Case(spec, identity_ray()), [IS_HIT, INTEN])
    assert r[IS_HIT][0] == 1 and r[INTEN][0] == 100.0
