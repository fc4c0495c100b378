"""Closed-form pins of the parts of the oracle that the reference's tests do not pin, run on the oracle (CPU) and on
the library (GPU, through the C-ABI), each also run with a deliberate mistake planted in the oracle to show that
the check catches it (orc_plant_mistake / oracle.plant_mistake, CPU only):

  beam reduction     src/gpu/nodeKernels.cu:334-422 (the reference's multiReturnTest cases are all skipped): a
                     divergent beam on a large plane. Sample s leaves along Rx(v_s) Ry(h_s) z (src/gpu/
                     optixPrograms.cu:230-248), so its distance to the plane n.x = c is c / (n . d_s) and its
                     intensity I0 |n . d_s| in closed form. FIRST / LAST are the extremes, STRONGEST the largest
                     intensity; SECOND is the reference's tracker rule in closed form: with k the first sample at
                     the minimum, SECOND = d_0 if k = 0, else min(min(d_0..d_k-1), min{d_j : j > k, d_j > d_k});
                     SECOND_STRONGEST likewise on intensities (maxima).
  velocity distortion test/src/graph/VelocityDistortionTest.cpp:17-95 (exact hit point (0, 0, 1); rays past the
                     cube miss), plus a cube beside the rays that only the moved rays reach.
  noise geometry     sigma = 0, mean mu: distance noise gives d + mu and moves the point by mu along the ray from the
                     sensor; angular hitpoint noise rotates the point by mu about the sensor's axis; angular ray
                     noise equals the noise-free graph fed rays pre-rotated by mu; and
                     test/src/graph/gaussianPoseIndependentTest.cpp:94-134 (moving sensor and cube together
                     leaves sensor-frame results unchanged).
  texture texels     a 4 x 4 texture of distinct values hit at every texel centre (and through wrapped UVs).
  azimuth/elevation  src/math/Mat3x4f.hpp:146-175 on rays rotationDeg(x, y, 0) = Ry(y) Rx(x): azimuth =
                     atan2(sin y, cos x cos y), elevation = -x, whatever the sensor pose.
"""
import ctypes as C
import math

import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S
from tests.helpers import Case, compose64, run_gpu, run_oracle

XYZ, INTEN, IS_HIT, ENT, DIST, AZ, EL, RET, INC = 1, 2, 4, 7, 8, 9, 10, 12, 23
BACKENDS = [pytest.param("oracle", id="oracle"), pytest.param("gpu", id="gpu", marks=pytest.mark.gpu)]


@pytest.fixture
def backend(request, oracle):
    if request.param == "oracle":
        return ("oracle", oracle)
    return ("gpu", request.getfixturevalue("rgl_lib"))


@pytest.fixture
def planted(oracle):
    def plant(which):
        oracle.plant_mistake(which)
    yield plant
    oracle.plant_mistake(0)


def trace(backend, case, fields):
    kind, lib = backend
    if kind == "oracle":
        return run_oracle(lib, case, fields=fields)
    lib.check("rgl_cleanup")  # each case on an empty scene
    return run_gpu(lib, case, fields)


# ------------------------------------------------------------------------------------------ beam reduction
HALF_H, HALF_V = 0.06, 0.035  # half divergences (rad); the API takes full angles
PLANE_C = 10.0
# Plane normals as (tilt from +z, azimuth of the tilt) in rad. Neighbouring sample distances and intensities differ by
# far more than FP32 rounding near every selected value, so the closed forms hold under 2e-6 relative perturbations.
NORMALS = {"perpendicular": (0.0, 0.0), "tilted": (0.15, 0.3), "tilted_back": (0.1, 2.6)}


def plane_scene(n):
    """A 400 x 400 m square in the plane n.x = PLANE_C, centred on the ray axis' crossing point."""
    n = np.asarray(n, np.float64)
    u = np.cross(n, [0.0, 1.0, 0.0])
    u /= np.linalg.norm(u)
    w = np.cross(n, u)
    c = n * PLANE_C
    v = np.array([c - 200 * u - 200 * w, c + 200 * u - 200 * w, c + 200 * u + 200 * w, c - 200 * u + 200 * w])
    spec = S.SceneSpec()
    m = spec.add_mesh(v.astype(np.float32), np.array([[0, 1, 2], [0, 2, 3]], np.int32))
    spec.add_instance(m, S.identity(), entity_id=5)
    return spec


def sample_dirs():
    """The 58 sample directions of a beam along +z in closed form."""
    out = [(0.0, 0.0)]
    for layer in range(3):
        shrink = 1.0 - layer / 3.0
        for k in range(19):
            a = 2.0 * math.pi * k / 19
            out.append((HALF_H * shrink * math.cos(a), HALF_V * shrink * math.sin(a)))
    hv = np.array(out)
    h, v = hv[:, 0], hv[:, 1]
    return np.stack([np.sin(h), -np.sin(v) * np.cos(h), np.cos(v) * np.cos(h)], axis=1)


def tracker_second(keys, smaller):
    """The SECOND (smaller=True, on distances) or SECOND_STRONGEST (False, on intensities) return in closed form."""
    keys = np.asarray(keys) if smaller else -np.asarray(keys)
    k = int(np.argmin(keys))  # first sample at the extreme
    if k == 0:
        return keys[0] if smaller else -keys[0]
    later = keys[k + 1:][keys[k + 1:] > keys[k]]
    val = min(keys[:k].min(), later.min() if len(later) else np.inf)
    return val if smaller else -val


def expected_returns(n, intensity=50.0):
    d = sample_dirs()
    cos = d @ np.asarray(n)
    dist = PLANE_C / cos
    inten = intensity * np.abs(cos)
    return {4: dist.min(), 2: dist.max(), 3: tracker_second(dist, True), 1: inten.max(), 5: tracker_second(inten, False)}


def plane_normal(name):
    tilt, az = NORMALS[name]
    return np.array([math.sin(tilt) * math.cos(az), math.sin(tilt) * math.sin(az), math.cos(tilt)])


def check_beam_reduction(backend, name):
    import robotecgpulidar_amd.rgl as R
    n = plane_normal(name)
    exp = expected_returns(n)
    spec = plane_scene(n)
    for mode in (R.RGL_RETURN_FIRST_LAST, R.RGL_RETURN_FIRST_SECOND, R.RGL_RETURN_STRONGEST_SECOND_STRONGEST):
        case = Case(spec, S.f32(S.identity()[None]), divergence=(2 * HALF_H, 2 * HALF_V), return_mode=mode,
                    default_intensity=50.0)
        r = trace(backend, case, [DIST, INTEN, IS_HIT, RET])
        assert list(r[IS_HIT]) == [1, 1]
        for p in range(2):
            t = int(r[RET][p])
            assert t == (mode >> (8 * p)) & 0xFF
            if t in (4, 2, 3):
                assert abs(float(r[DIST][p]) - exp[t]) <= 1e-5 * exp[t], (name, t, float(r[DIST][p]), exp[t])
            else:
                assert abs(float(r[INTEN][p]) - exp[t]) <= 1e-5 * exp[t], (name, t, float(r[INTEN][p]), exp[t])


@pytest.mark.parametrize("backend", BACKENDS, indirect=True)
@pytest.mark.parametrize("name", list(NORMALS))
def test_beam_reduction_on_planes(backend, name):
    check_beam_reduction(backend, name)


def test_beam_reduction_closed_forms_are_distinct():
    """The geometries exercise different trackers: the minimum is not always the primary sample, SECOND differs
    from FIRST, and the two intensity returns differ."""
    e = {k: expected_returns(plane_normal(k)) for k in NORMALS}
    assert e["perpendicular"][3] == e["perpendicular"][4]  # primary nearest: SECOND stays on it
    assert e["tilted"][3] > e["tilted"][4] * (1 + 1e-4)
    assert e["tilted_back"][5] < e["tilted_back"][1] * (1 - 1e-4)


def second_ranges(oracle, name, mode):
    import robotecgpulidar_amd.rgl as R
    n = plane_normal(name)
    case = Case(plane_scene(n), S.f32(S.identity()[None]), divergence=(2 * HALF_H, 2 * HALF_V),
                return_mode=getattr(R, "RGL_RETURN_" + mode), default_intensity=50.0)
    return run_oracle(oracle, case, fields=[DIST, INTEN, IS_HIT]), expected_returns(n)


@pytest.mark.parametrize("name", list(NORMALS))
def test_tie_resolution_ranges_hold_the_closed_forms(oracle, name):
    """The oracle's range of SECOND / SECOND_STRONGEST over every resolution of near-tied comparisons (used by the
    GPU multi-return checks) contains the closed-form values on the planes, and so does the returned value."""
    for mode, t, col in (("FIRST_SECOND", 3, 0), ("STRONGEST_SECOND_STRONGEST", 5, 2)):
        r, exp = second_ranges(oracle, name, mode)
        lo, hi = float(r["second_range"][0, col]), float(r["second_range"][0, col + 1])
        got = float(r[DIST][1] if t == 3 else r[INTEN][1])
        assert lo - 1e-5 * abs(lo) <= exp[t] <= hi + 1e-5 * abs(hi), (name, t, lo, exp[t], hi)
        assert lo <= got <= hi, (name, t, lo, got, hi)


def test_tie_resolution_ranges_catch_planted_mistake(oracle, planted):
    """The ranges come from a tracker model of their own: with the SECOND trackers seeded wrongly (mistake 1),
    the returned values leave the ranges (on the perpendicular plane, where SECOND stays on the nearest, primary
    sample)."""
    planted(1)
    outside = 0
    for name in NORMALS:
        for mode, t, col in (("FIRST_SECOND", 3, 0), ("STRONGEST_SECOND_STRONGEST", 5, 2)):
            r, _ = second_ranges(oracle, name, mode)
            lo, hi = float(r["second_range"][0, col]), float(r["second_range"][0, col + 1])
            got = float(r[DIST][1] if t == 3 else r[INTEN][1])
            outside += not (lo <= got <= hi)
    assert outside > 0


@pytest.mark.parametrize("mistake", [1, 2, 6])
def test_beam_reduction_catches_planted_mistakes(oracle, planted, mistake):
    planted(mistake)
    with pytest.raises(AssertionError):
        for name in NORMALS:
            check_beam_reduction(("oracle", oracle), name)


# ------------------------------------------------------------------------------------------ velocity distortion
def check_distortion(backend):
    """VelocityDistortionTest.smoke_test: 10 identity rays fired 300 ms apart from a sensor moving at 1 m/s along
    y, inside a unit cube: rays whose moved origin left the cube (y > 1) miss; the others return exactly (0, 0, 1),
    the point at the hit distance along the undistorted ray. A cube at y in [1, 3] is reached by rays 4..9 only."""
    rays = S.f32(np.stack([S.identity()] * 10))
    offsets = np.arange(10, dtype=np.float32) * 300.0
    vel = ((0.0, 1.0, 0.0), (0.0, 0.0, 0.0))
    for cube_y, hits in ((0.0, [i * 0.3 < 1.0 for i in range(10)]), (2.0, [1.0 < i * 0.3 < 3.0 for i in range(10)])):
        spec = S.SceneSpec()
        m = spec.add_mesh(S.CUBE_VERTICES, S.CUBE_INDICES)
        spec.add_instance(m, S.translation(0, cube_y, 0))
        case = Case(spec, rays, time_offsets=offsets, distortion=True, velocity=vel)
        r = trace(backend, case, [XYZ, IS_HIT])
        assert list(r[IS_HIT] == 1) == hits, (cube_y, list(r[IS_HIT]))
        for i in np.nonzero(hits)[0]:
            assert list(r[XYZ][i]) == [0.0, 0.0, 1.0], (cube_y, i, r[XYZ][i])
        if cube_y == 0.0:  # without distortion every ray hits (0, 0, 1)
            r = trace(backend, Case(spec, rays, time_offsets=offsets, distortion=False, velocity=vel), [XYZ, IS_HIT])
            assert np.all(r[IS_HIT] == 1) and np.all(r[XYZ] == np.float32([0, 0, 1]))


@pytest.mark.parametrize("backend", BACKENDS, indirect=True)
def test_velocity_distortion_known_answers(backend):
    check_distortion(backend)


def test_distortion_catches_planted_mistake(oracle, planted):
    planted(3)
    with pytest.raises(AssertionError):
        check_distortion(("oracle", oracle))


# ------------------------------------------------------------------------------------------ noise geometry
POSE = S.f32(S.trs((3.0, -2.0, 1.5), (10.0, -5.0, 35.0)))
MU = 0.25


def noise_scene():
    """A 40 m cube around the sensor: every ray hits one of its inner faces."""
    spec = S.SceneSpec()
    m = spec.add_mesh(S.CUBE_VERTICES, S.CUBE_INDICES)
    spec.add_instance(m, S.trs((3.0, -2.0, 1.5), (0, 0, 0), (20, 20, 20)))
    return spec


def noise_rays():
    return S.make_lidar3d_rays(300, 120, 20, 20)


def gpu_noise_graph(rgl, rays, pose, kind, params, transform_back=None):
    """rays -> pose -> [angular ray noise] -> raytrace -> [hit-point noise] -> [points transform] -> yield."""
    n = rgl.Node
    chain = [rgl.node_rays_from_mat3x4f(n(), rays), rgl.node_rays_transform(n(), pose)]
    noise = None
    if kind is not None:
        noise = n()
        fn = {"ray": "rgl_node_gaussian_noise_angular_ray", "hitpoint": "rgl_node_gaussian_noise_angular_hitpoint",
              "distance": "rgl_node_gaussian_noise_distance"}[kind]
        rgl.check(fn, noise.ref, *params)
    if kind == "ray":
        chain.append(noise)
    chain.append(rgl.node_raytrace(n()))
    if kind in ("hitpoint", "distance"):
        chain.append(noise)
    if transform_back is not None:
        chain.append(rgl.node_points_transform(n(), transform_back))
    out = rgl.node_points_yield(n(), [XYZ, DIST, IS_HIT])
    chain.append(out)
    rgl.chain(*chain)
    rgl.graph_run(chain[0])
    return {f: rgl.result(out, f) for f in (XYZ, DIST, IS_HIT)}


def run_noise(backend, kind, params, rays=None, pose=POSE, spec=None):
    """Results of rays -> pose -> noise -> ... on either backend (the oracle applies its FP64 noise stages)."""
    which, lib = backend
    rays = noise_rays() if rays is None else rays
    spec = noise_scene() if spec is None else spec
    if which == "gpu":
        lib.check("rgl_cleanup")
        S.load_into_rgl(spec)
        return gpu_noise_graph(lib, rays, pose, kind, params)
    look_at = np.linalg.inv(np.vstack([pose.astype(np.float64), [0, 0, 0, 1]]))[:3]
    world = compose64(pose, rays)
    zeros = np.zeros(len(rays))
    if kind == "ray":
        world = lib.noise_angular_rays(world, zeros, params[0], params[1], params[2], look_at)
    r = lib.raytrace(spec, world, ray_origin_to_world=pose, fields=[XYZ, DIST, IS_HIT])
    if kind == "distance":
        r[XYZ], r[DIST] = lib.noise_distance(r[XYZ], r[DIST], zeros, *params, look_at)
    elif kind == "hitpoint":
        r[XYZ], r[DIST] = lib.noise_angular_hitpoint(r[XYZ], zeros, params[0], params[1], params[2], look_at)
    return r


def rot(axis, a):
    c, s = math.cos(a), math.sin(a)
    return {3: np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]]), 2: np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])}[axis]


def check_noise_closed_forms(backend):
    clean = run_noise(backend, None, None)
    hit = clean[IS_HIT] == 1
    assert hit.all()
    p = clean[XYZ].astype(np.float64)
    o = POSE[:, 3].astype(np.float64)
    R3 = POSE[:, :3].astype(np.float64)
    # distance noise: d + mu, point moved by mu along the ray from the sensor
    r = run_noise(backend, "distance", (MU, 0.0, 0.0))
    np.testing.assert_allclose(r[DIST], clean[DIST].astype(np.float64) + MU, rtol=1e-6, atol=1e-5)
    u = (p - o) / np.linalg.norm(p - o, axis=1, keepdims=True)
    np.testing.assert_allclose(r[XYZ], p + MU * u, atol=1e-4)
    # angular hit-point noise about the sensor's y axis: p' = o + R Ry(mu) R^T (p - o), distance kept
    r = run_noise(backend, "hitpoint", (MU, 0.0, 2))
    q = (p - o) @ R3 @ rot(2, MU).T @ R3.T + o
    np.testing.assert_allclose(r[XYZ], q, atol=1e-4)
    np.testing.assert_allclose(r[DIST], np.linalg.norm(p - o, axis=1), rtol=1e-5)
    # angular ray noise about z equals tracing rays pre-rotated by Rz(mu) in the sensor frame
    pre = S.f32(np.einsum("ij,njk->nik", rot(3, MU), noise_rays().astype(np.float64)))  # Rz(mu) * ray
    r = run_noise(backend, "ray", (MU, 0.0, 3))
    ref = run_noise(backend, None, None, rays=pre)
    same = (r[IS_HIT] == 1) & (ref[IS_HIT] == 1)
    assert same.mean() > 0.99
    np.testing.assert_allclose(r[DIST][same], ref[DIST][same], rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(r[XYZ][same], ref[XYZ][same], atol=1e-3)


@pytest.mark.parametrize("backend", BACKENDS, indirect=True)
def test_noise_geometry_closed_forms(backend):
    check_noise_closed_forms(backend)


@pytest.mark.parametrize("mistake", [7, 8])
def test_noise_geometry_catches_planted_mistakes(oracle, planted, mistake):
    planted(mistake)
    with pytest.raises(AssertionError):
        check_noise_closed_forms(("oracle", oracle))


@pytest.mark.gpu
@pytest.mark.parametrize("kind,params", [("distance", (MU, 0.0, 0.0)), ("hitpoint", (MU, 0.0, 2)),
                                         ("ray", (MU, 0.0, 2))])
def test_noise_pose_independent(rgl_lib, kind, params):
    """gaussianPoseIndependentTest.cpp:94-134: the cube and the lidar moved by the same transform give the same
    sensor-frame results (within 1e-4 m)."""
    rgl = rgl_lib
    rays = S.make_lidar3d_rays(360, 180, 10, 10)
    first = None
    for k, tf in enumerate([S.identity(), S.trs((10, -4, 2), (0, 0, 90)), S.trs((-50, 30, 5), (20, 30, 40))]):
        spec = S.SceneSpec([S.Mesh(S.CUBE_VERTICES, S.CUBE_INDICES)], [S.Instance(0, S.f32(tf))])
        S.load_into_rgl(spec)
        inv = S.f32(np.linalg.inv(np.vstack([tf, [0, 0, 0, 1]]))[:3])
        r = gpu_noise_graph(rgl, rays, S.f32(tf), kind, params, transform_back=inv)
        rgl.check("rgl_cleanup")
        if first is None:
            first = r
            continue
        np.testing.assert_allclose(r[XYZ], first[XYZ], atol=1e-4)
        np.testing.assert_allclose(r[DIST], first[DIST], atol=1e-4)


# ------------------------------------------------------------------------------------------ texels
TEXELS = (10 + 16 * np.arange(4)[:, None] + np.arange(4)[None, :]).astype(np.uint8)  # row j, column i


def check_texels(backend, uv_shift=(0.0, 0.0)):
    """A 4 x 4 quad at z = 5 with UV = (x / 4, y / 4) + shift; a ray at every texel centre returns that texel's
    value (point sampling, wrap addressing, cos 0 = 1): texel row = floor(v * 4), column = floor(u * 4)."""
    spec = S.SceneSpec()
    v = np.array([[0, 0, 5], [4, 0, 5], [4, 4, 5], [0, 4, 5]], np.float32)
    uv = v[:, :2] / 4.0 + np.float32(uv_shift)
    m = spec.add_mesh(v, np.array([[0, 1, 2], [0, 2, 3]], np.int32), uv)
    spec.textures.append(TEXELS)
    spec.add_instance(m, S.identity(), texture=0)
    ii, jj = np.meshgrid(np.arange(4), np.arange(4), indexing="xy")
    rays = S.f32(np.stack([S.translation(i + 0.5, j + 0.5, 0) for i, j in zip(ii.ravel(), jj.ravel())]))
    r = trace(backend, Case(spec, rays), [INTEN, IS_HIT])
    assert np.all(r[IS_HIT] == 1)
    np.testing.assert_array_equal(r[INTEN], TEXELS[jj.ravel(), ii.ravel()].astype(np.float32))


@pytest.mark.parametrize("backend", BACKENDS, indirect=True)
@pytest.mark.parametrize("shift", [(0.0, 0.0), (1.0, -2.0)])
def test_texel_centres(backend, shift):
    check_texels(backend, shift)


def test_texels_catch_planted_mistake(oracle, planted):
    planted(4)
    with pytest.raises(AssertionError):
        check_texels(("oracle", oracle))


# ------------------------------------------------------------------------------------------ azimuth / elevation
ANGLES = [(x, y) for x in (-40.0, -7.5, 0.0, 12.0, 33.0) for y in (-150.0, -60.0, -5.0, 0.0, 20.0, 89.0, 170.0)]


def check_angles(backend):
    rays = S.f32(np.stack([S.rotation_deg(x, y, 0.0) for x, y in ANGLES]))
    x = np.deg2rad([a for a, _ in ANGLES])
    y = np.deg2rad([b for _, b in ANGLES])
    spec = S.SceneSpec([S.Mesh(S.CUBE_VERTICES, S.CUBE_INDICES)], [S.Instance(0, S.f32(S.translation(0, 0, 50)))])
    for pose in (None, S.f32(S.trs((5.0, -3.0, 2.0), (15.0, -25.0, 70.0)))):
        r = trace(backend, Case(spec, rays, pose=pose), [AZ, EL, IS_HIT])
        np.testing.assert_allclose(r[AZ], np.arctan2(np.sin(y), np.cos(x) * np.cos(y)), atol=2e-6)
        np.testing.assert_allclose(r[EL], -x, atol=2e-6)


@pytest.mark.parametrize("backend", BACKENDS, indirect=True)
def test_azimuth_elevation_from_known_angles(backend):
    check_angles(backend)


def test_angles_catch_planted_mistake(oracle, planted):
    planted(5)
    with pytest.raises(AssertionError):
        check_angles(("oracle", oracle))
