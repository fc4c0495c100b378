"""Radar postprocessing (rgl_node_points_radar_postprocess, src/graph/RadarPostprocessPointsNode.cpp).

CPU: the oracle's clustering against the reference's own clustering known answers
(test/src/graph/nodes/RadarPostprocessPointsNodeTest.cpp), its energy computation against the physical-optics
flat-plate cross-section sigma = 4 pi A^2 / lambda^2 at normal incidence (invariant under rotation of the
polarisation about the ray), and its received power against the radar equation.
GPU: the library's node on the same known answers through rgl_node_points_from_array, and on raytraced radar frames
against the oracle: identical clusters (count, membership through the centre point's fields), RCS within 1e-3 dB,
power within 1e-3 dB plus the node's noise, which is replayed from the library's seeded normal stream; more clusters
than a wave has lanes (joins to clusters kept in LDS, NaN radial-speed bounds) against the oracle, and the node's
cluster limit reported as an error.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import radar_post as RP

XYZ, DIST, AZ, EL, RAD, POSE, NORMAL = 1, 8, 9, 10, 17, 24, 22
POWER, RCS, NOISE, SNR = 18, 19, 20, 21
REQUIRED = [DIST, AZ, EL, RAD, POSE, NORMAL, XYZ]

SCOPES = [(0.0, 10.0, 2.0, 2.0, 2.0), (10.0, 20.0, 1.0, 1.0, 1.0)]
# (distances, azimuths, radial speeds) -> expected cluster count (RadarPostprocessPointsNodeTest.cpp:56-126)
KATS = [
    (([100.0, 200.0], [10.0, 10.0], [10.0, 10.0]), 0),
    (([5.0, 6.0, 7.0], [10.0, 10.1, 10.2], [10.0, 10.1, 10.2]), 1),
    (([5.0, 10.0], [10.0, 10.0], [10.0, 10.0]), 2),
    (([5.0, 5.0], [10.0, 20.0], [10.0, 10.0]), 2),
    (([5.0, 5.0], [10.0, 10.0], [10.0, 20.0]), 2),
    (([0.0, 2.5, 1.0], [11.0, 12.5, 12.0], [11.0, 12.5, 12.0]), 1),
    (([0.0, 2.5, 1.0], [11.0, 11.5, 12.0], [11.0, 11.5, 12.0]), 1),
    (([0.0, 2.5, 1.0], [11.0, 0.0, 12.0], [11.0, 11.5, 12.0]), 2),
    (([0.0, 2.5, 1.0], [11.0, 11.5, 12.0], [11.0, 0.0, 12.0]), 2),
    (([6.5, 8.0, 9.5, 11.0], [10.0] * 4, [10.0] * 4), 2),
]


@pytest.mark.parametrize("case", range(len(KATS)))
def test_oracle_clustering_known_answers(case):
    (d, a, r), expected = KATS[case]
    cl = RP.cluster(d, a, r, np.zeros(len(d)), SCOPES)
    assert len(cl) == expected


def head_on(n, roll_deg=0.0):
    """n rays along +Z from the origin, rolled about the ray (polarisation turns), with a plate facing them."""
    c, s = math.cos(math.radians(roll_deg)), math.sin(math.radians(roll_deg))
    pose = np.zeros((n, 3, 4), np.float32)
    pose[:, :, :3] = [[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]]
    normal = np.tile(np.float32([0.0, 0.0, -1.0]), (n, 1))
    return pose, normal


@pytest.mark.parametrize("freq,roll", [(77e9, 0.0), (77e9, 30.0), (24e9, 90.0), (1e9, 45.0)])
def test_oracle_energy_flat_plate(freq, roll):
    """Normal incidence on a plate: each ray contributes i k A_ray / (2 pi); the coherent sum over the plate gives
    the physical-optics plate cross-section 4 pi A^2 / lambda^2, whatever the polarisation."""
    rng = np.random.default_rng(3)
    n = 37
    dist = rng.uniform(5.0, 40.0, n).astype(np.float32)
    az_step, el_step = math.radians(0.5), math.radians(0.8)
    pose, normal = head_on(n, roll)
    bu, br, fac = RP.energy(pose, np.eye(3, 4), dist, normal, az_step, el_step, freq)
    au, ar = np.sum(bu * fac), np.sum(br * fac)
    area = float(np.sum(dist.astype(np.float64) ** 2) * math.sin(el_step) * az_step)
    lam = RP.C0 / freq
    sigma = 4.0 * math.pi * (abs(au) ** 2 + abs(ar) ** 2)
    # The rolled poses are FP32 (cos and sin rounded), hence 1e-6 rather than FP64 agreement.
    assert sigma == pytest.approx(4.0 * math.pi * area ** 2 / lam ** 2, rel=1e-6)


def test_oracle_received_power_is_the_radar_equation():
    """P_r = P_t G^2 lambda^2 sigma / ((4 pi)^3 d^4), in dB."""
    pose, normal = head_on(1)
    d = np.float32([12.5])
    freq, pt, g = 77e9, 31.0, 30.0
    cl, centres, rcs, power = RP.postprocess(pose, np.eye(3, 4), d, np.zeros(1), np.zeros(1), np.zeros(1), normal,
                                             SCOPES, 0.01, 0.01, freq, pt, g)
    lam = RP.C0 / freq
    sigma = 10.0 ** (rcs[0] / 10.0)
    pr = 10.0 ** (pt / 10.0) * (10.0 ** (g / 10.0)) ** 2 * lam ** 2 * sigma / ((4 * math.pi) ** 3 * 12.5 ** 4)
    assert len(cl) == 1 and power[0] == pytest.approx(10.0 * math.log10(pr), abs=1e-9)


def test_oracle_merge_keeps_reference_bound_rule():
    """After a merge the radial-speed upper bound is the smaller of the two (RadarPostprocessPointsNode.cpp:380-383).
    Assignment gives A = {p0, p2} (r 10..10.5), B = {p1} (r 11, azimuth 0.5) and C = {p3} (r 11.3, azimuth -0.5);
    A and B merge with upper bound 10.5 (not 11), so C, 0.3 from 11 but 0.8 from 10.5, stays separate."""
    scopes = [(0.0, 50.0, 1.0, 0.6, 0.5)]
    d = [5.0, 5.0, 5.0, 5.0]
    a = [0.0, 0.5, 0.0, -0.5]
    r = [10.0, 11.0, 10.5, 11.3]
    cl = RP.cluster(d, a, r, np.zeros(4), scopes)
    assert [c.idx for c in cl] == [[0, 2, 1], [3]]
    assert cl[0].r[1] == np.float32(10.5)


# ---------------------------------------------------------------------------------------------------------------- GPU
def records(d, a, r, seed=0):
    """Dense input points with the fields the node requires: the given distance / azimuth / radial speed, elevation
    i / (i + 1) as the reference's test generator makes, random ray poses (rotations) and unit normals."""
    n = len(d)
    rng = np.random.default_rng(seed)
    from robotecgpulidar_amd import scenes as S
    dt = np.dtype([("d", "<f4"), ("a", "<f4"), ("e", "<f4"), ("r", "<f4"), ("pose", "<f4", (12,)),
                   ("n", "<f4", (3,)), ("xyz", "<f4", (3,))])
    rec = np.zeros(n, dt)
    rec["d"], rec["a"], rec["r"] = d, a, r
    rec["e"] = np.arange(n) / (np.arange(n) + 1.0)
    rec["pose"] = np.stack([S.f32(S.rotation_deg(*rng.uniform(-180, 180, 3))).ravel() for _ in range(n)])
    nv = rng.normal(size=(n, 3))
    rec["n"] = nv / np.linalg.norm(nv, axis=1, keepdims=True)
    rec["xyz"] = rng.uniform(-10, 10, (n, 3))
    return rec, [DIST, AZ, EL, RAD, POSE, NORMAL, XYZ]


PARAMS = (1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)  # as in RadarPostprocessPointsNodeTest.cpp:30


@pytest.mark.gpu
def test_gpu_clustering_known_answers(rgl_lib):
    rgl = rgl_lib
    radar = rgl.node_points_radar_postprocess(rgl.Node(), SCOPES, *PARAMS)
    for k, ((d, a, r), expected) in enumerate(KATS):
        rec, fields = records(d, a, r, seed=k)
        src = rgl.node_points_from_array(rgl.Node(), rec, fields)
        rgl.add_child(src, radar)
        rgl.graph_run(radar)
        n, _ = rgl.result_size(radar, DIST)
        assert n == expected, (k, n, expected)
        cl = RP.cluster(d, a, r, rec["e"], SCOPES)
        centres = [RP.centre_index(c, a, rec["e"]) for c in cl]
        if n:
            np.testing.assert_array_equal(rgl.result(radar, DIST), rec["d"][centres])
            np.testing.assert_array_equal(rgl.result(radar, XYZ), rec["xyz"][centres])
        rgl.check("rgl_graph_node_remove_child", src.h, radar.h)
        rgl.check("rgl_graph_destroy", src.h)


def many_clusters(k, seed):
    """k clusters far apart in azimuth (one point each, in order), then a second point for each in shuffled order;
    every 7th cluster starts with a NaN radial speed, which the joining point's speed replaces."""
    rng = np.random.default_rng(seed)
    first_r = np.where(np.arange(k) % 7 == 3, np.nan, 1.0)
    order = rng.permutation(k)
    d = np.concatenate([np.full(k, 5.0), np.full(k, 5.5)])
    a = np.concatenate([10.0 * np.arange(k), 10.0 * order + 0.5])
    r = np.concatenate([first_r, np.full(k, 1.2)])
    return d, a, r


@pytest.mark.gpu
@pytest.mark.parametrize("k", [63, 64, 65, 300])
def test_gpu_clustering_beyond_register_clusters(rgl_lib, k):
    """More clusters than one per lane: later points join clusters held in LDS, in the reference's order."""
    rgl = rgl_lib
    d, a, r = many_clusters(k, seed=k)
    rec, fields = records(d, a, r, seed=k)
    cl = RP.cluster(d, a, r, rec["e"], SCOPES)
    assert len(cl) == k and all(len(c.idx) == 2 for c in cl)
    centres = [RP.centre_index(c, a, rec["e"]) for c in cl]
    radar = rgl.node_points_radar_postprocess(rgl.Node(), SCOPES, *PARAMS)
    src = rgl.node_points_from_array(rgl.Node(), rec, fields)
    rgl.add_child(src, radar)
    rgl.graph_run(radar)
    assert rgl.result_size(radar, DIST)[0] == k
    np.testing.assert_array_equal(rgl.result(radar, DIST), rec["d"][centres])
    np.testing.assert_array_equal(rgl.result(radar, XYZ), rec["xyz"][centres])
    rgl.check("rgl_graph_node_remove_child", src.h, radar.h)
    rgl.check("rgl_graph_destroy", src.h)


def interleaved(n, seed, nan_speeds):
    """n points drawn from 12 objects in random order, spread over both scopes and up to the thresholds, so that
    clusters grow towards each other, points of one object land in several clusters, and a later point can be taken
    by a cluster that an earlier one widened; with nan_speeds, radial speeds are NaN for whole objects, for random
    points or for all."""
    rng = np.random.default_rng(seed)
    k = 12
    cd, ca, cr = rng.uniform(1.0, 19.0, k), rng.uniform(0.0, 90.0, k), rng.uniform(-6.0, 6.0, k)
    obj = rng.integers(0, k, n)
    if seed % 2:  # runs of one object, as a scan crosses it
        obj = np.repeat(obj[: (n + 15) // 16], 16)[:n]
    d = cd[obj] + rng.normal(0.0, 0.8, n)
    a = ca[obj] + rng.normal(0.0, 1.5, n)
    r = cr[obj] + rng.normal(0.0, 0.8, n)
    if nan_speeds == "objects":
        r[obj % 3 == 0] = np.nan
    elif nan_speeds == "points":  # clusters opened by a NaN speed, whose bounds a later finite speed replaces
        r[rng.random(n) < 0.3] = np.nan
    elif nan_speeds == "all":
        r[:] = np.nan
    return d.astype(np.float32), a.astype(np.float32), r.astype(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("seed,nan_speeds", [(0, None), (1, None), (2, None), (3, None), (4, "objects"),
                                             (5, "objects"), (6, "all"), (7, "all"), (8, "points"), (9, "points")])
def test_gpu_clustering_interleaved_matches_oracle(rgl_lib, seed, nan_speeds):
    """Interleaved, overlapping clusters: the same clusters (count, centres bit for bit) as the oracle."""
    rgl = rgl_lib
    d, a, r = interleaved(3000, seed, nan_speeds)
    rec, fields = records(d, a, r, seed=seed)
    cl = RP.cluster(d, a, r, rec["e"], SCOPES)
    centres = [RP.centre_index(c, a, rec["e"]) for c in cl]
    radar = rgl.node_points_radar_postprocess(rgl.Node(), SCOPES, *PARAMS)
    src = rgl.node_points_from_array(rgl.Node(), rec, fields)
    rgl.add_child(src, radar)
    rgl.graph_run(radar)
    print(f"seed {seed}: {len(cl)} clusters, sizes {sorted((len(c.idx) for c in cl), reverse=True)[:8]}")
    assert rgl.result_size(radar, DIST)[0] == len(cl)
    np.testing.assert_array_equal(rgl.result(radar, DIST), rec["d"][centres])
    np.testing.assert_array_equal(rgl.result(radar, XYZ), rec["xyz"][centres])
    rgl.check("rgl_graph_node_remove_child", src.h, radar.h)
    rgl.check("rgl_graph_destroy", src.h)


@pytest.mark.gpu
def test_gpu_clustering_cluster_limit(rgl_lib):
    """1025 clusters exceed the node's limit of 1024: the run reports it instead of returning a partial result."""
    rgl = rgl_lib
    k = 1025
    rec, fields = records(np.full(k, 5.0), 10.0 * np.arange(k), np.ones(k), seed=1)
    radar = rgl.node_points_radar_postprocess(rgl.Node(), SCOPES, *PARAMS)
    src = rgl.node_points_from_array(rgl.Node(), rec, fields)
    rgl.add_child(src, radar)
    st = rgl.call("rgl_graph_run", radar.h)
    if st == 0:
        cnt, size = C.c_int32(), C.c_int32()
        st = rgl.call("rgl_graph_get_result_size", radar.h, DIST, C.byref(cnt), C.byref(size))
    assert st == rgl.RGL_INVALID_PIPELINE, st
    assert "at most 1024 clusters" in rgl.last_error()
    rgl.check("rgl_graph_node_remove_child", src.h, radar.h)
    rgl.check("rgl_graph_destroy", src.h)


def radar_frame_graph(rgl, scene_spec, frames=2):
    """Radar rays -> pose -> raytrace (moving boxes, so radial speeds are defined) -> compact -> radar postprocess."""
    from robotecgpulidar_amd import scenes as S
    rgl.scene_set_time(None, 0)
    _, ents, _ = S.load_into_rgl(scene_spec)
    n = rgl.Node
    rays = rgl.node_rays_from_mat3x4f(n(), S.radar_rays(256, 64))
    pose = rgl.node_rays_transform(n(), S.f32(S.trs((0.0, 0.0, 1.5), (0, 0, 25.0))))
    rng_node = rgl.node_rays_set_range(n(), np.array([[0.5, 120.0]], np.float32))
    rt = rgl.node_raytrace(n())
    lin, ang = np.float32([3.0, 0.0, 0.0]), np.float32([0.0, 0.0, 0.1])
    rgl.check("rgl_node_raytrace_configure_velocity", rt.h, lin.ctypes.data, ang.ctypes.data)
    comp = rgl.node_points_compact_by_field(n(), rgl.IS_HIT_I32)
    inp = rgl.node_points_yield(n(), REQUIRED)
    radar = rgl.node_points_radar_postprocess(n(), [(0.0, 30.0, 1.0, 0.5, math.radians(2.0)),
                                                    (30.0, 120.0, 2.0, 1.0, math.radians(1.0))],
                                              math.radians(120.0 / 256), math.radians(20.0 / 63), 77e9, 31.0, 27.0,
                                              -90.0, 2.0)
    out = rgl.node_points_yield(n(), [XYZ, DIST, AZ, EL, RAD, RCS, POWER, NOISE, SNR])
    rgl.chain(rays, pose, rng_node, rt, comp, inp, radar, out)
    rgl.check("rgl_node_gaussian_noise_set_seed", radar.h, C.c_uint64(0xAADA2))
    for f in range(frames):
        rgl.scene_set_time(None, (f + 1) * 50_000_000)
        for k, e in enumerate(ents[1:], start=1):
            t = scene_spec.instances[k].transform.copy()
            t[:, 3] += np.float32([0.4 * (f + 1) * ((k % 3) - 1), 0.2 * (f + 1), 0.0])
            rgl.entity_set_transform(e, t)
    return rays, inp, radar, out


@pytest.mark.gpu
def test_gpu_radar_frame_matches_oracle(rgl_lib):
    from robotecgpulidar_amd import scenes as S
    rgl = rgl_lib
    spec = S.config1_scene()
    rays, inp, radar, out = radar_frame_graph(rgl, spec)
    scopes = [(0.0, 30.0, 1.0, 0.5, math.radians(2.0)), (30.0, 120.0, 2.0, 1.0, math.radians(1.0))]
    for frame in range(2):
        rgl.graph_run(rays)
        x = {f: rgl.result(inp, f) for f in REQUIRED}
        g = {f: rgl.result(out, f) for f in (XYZ, DIST, AZ, EL, RAD, RCS, POWER, NOISE, SNR)}
        look_at = np.linalg.inv(np.vstack([S.trs((0.0, 0.0, 1.5), (0, 0, 25.0)), [0, 0, 0, 1]]))[:3]
        cl, centres, rcs, power = RP.postprocess(x[POSE].reshape(-1, 3, 4), look_at, x[DIST], x[AZ], x[EL], x[RAD],
                                                 x[NORMAL], scopes, math.radians(120.0 / 256),
                                                 math.radians(20.0 / 63), 77e9, 31.0, 27.0)
        print(f"frame {frame}: {len(x[DIST])} detections, {len(cl)} clusters, radial speeds defined: "
              f"{int(np.isfinite(x[RAD]).sum())}")
        assert len(x[DIST]) > 3000 and len(cl) > 10
        assert len(g[DIST]) == len(cl)
        for f in (XYZ, DIST, AZ, EL, RAD):  # the centre point of every cluster, bit for bit
            np.testing.assert_array_equal(g[f], x[f][centres])
        # RCS: FP32 sums of the per-ray terms against FP64; the error bound is relative to the sum of magnitudes.
        bu, br, fac = RP.energy(x[POSE].reshape(-1, 3, 4), look_at, x[DIST], x[NORMAL], math.radians(120.0 / 256),
                                math.radians(20.0 / 63), 77e9)
        for k, c in enumerate(cl):
            mag = float(np.sum(np.abs(bu[c.idx] * fac[c.idx]) + np.abs(br[c.idx] * fac[c.idx])))
            lin_g = 10.0 ** (float(g[RCS][k]) / 10.0) / (4 * math.pi)
            lin_o = 10.0 ** (rcs[k] / 10.0) / (4 * math.pi)
            assert abs(math.sqrt(lin_g) - math.sqrt(lin_o)) <= 1e-4 * mag + 1e-30, (k, g[RCS][k], rcs[k])
        s, run, tag = C.c_uint64(), C.c_uint32(), C.c_uint32()
        rgl.check("rgl_node_gaussian_noise_get_stream", radar.h, C.byref(s), C.byref(run), C.byref(tag))
        z = rgl.noise_normals(s.value, run.value - 1, tag.value, len(cl)).astype(np.float64)
        np.testing.assert_allclose(g[NOISE], -90.0 + 2.0 * z, rtol=1e-6, atol=1e-5)
        good = np.abs(g[RCS] - rcs) < 0.01
        np.testing.assert_allclose((g[POWER] - g[NOISE])[good], power[good], atol=0.02)
        np.testing.assert_allclose(g[SNR], g[POWER] - g[NOISE], atol=1e-4)


@pytest.mark.gpu
def test_gpu_radar_postprocess_errors(rgl_lib):
    rgl = rgl_lib
    sc = np.float32([[0.0, 10.0, 1.0, 1.0, 1.0]])
    bad = [(sc, 0, 1.0), (np.float32([[5.0, 1.0, 1.0, 1.0, 1.0]]), 1, 1.0), (np.float32([[0.0, 1.0, -1.0, 1.0, 1.0]]), 1, 1.0),
           (sc, 1, 0.0)]
    for scopes, count, az_step in bad:
        h = rgl.Node()
        st = rgl.call("rgl_node_points_radar_postprocess", h.ref, scopes.ctypes.data, count, C.c_float(az_step),
                      C.c_float(1.0), C.c_float(1.0), C.c_float(1.0), C.c_float(1.0), C.c_float(1.0), C.c_float(1.0))
        assert st == rgl.RGL_INVALID_ARGUMENT
    # Input that is not dense (raytrace output before compaction) is rejected at run time.
    from robotecgpulidar_amd import scenes as S
    S.load_into_rgl(S.config1_scene())
    n = rgl.Node
    rays = rgl.node_rays_from_mat3x4f(n(), S.radar_rays(16, 8))
    rt = rgl.node_raytrace(n())
    radar = rgl.node_points_radar_postprocess(n(), SCOPES, *PARAMS)
    rgl.chain(rays, rt, radar)
    assert rgl.call("rgl_graph_run", rays.h) == rgl.RGL_INVALID_PIPELINE
    assert "requires dense input" in rgl.last_error()


@pytest.mark.gpu
def test_gpu_radar_batched_equals_solo(rgl_lib):
    from robotecgpulidar_amd import scenes as S
    from tests.test_gpu_batch import assert_same
    rgl = rgl_lib
    spec = S.config1_scene()
    a = radar_frame_graph(rgl, spec, frames=1)
    b_rays = rgl.node_rays_from_mat3x4f(rgl.Node(), S.radar_rays(256, 64))
    # The second graph is built the same way and traces the same default scene.
    n = rgl.Node
    pose = rgl.node_rays_transform(n(), S.f32(S.trs((0.0, 0.0, 1.5), (0, 0, 25.0))))
    rng_node = rgl.node_rays_set_range(n(), np.array([[0.5, 120.0]], np.float32))
    rt = rgl.node_raytrace(n())
    lin, ang = np.float32([3.0, 0.0, 0.0]), np.float32([0.0, 0.0, 0.1])
    rgl.check("rgl_node_raytrace_configure_velocity", rt.h, lin.ctypes.data, ang.ctypes.data)
    comp = rgl.node_points_compact_by_field(n(), rgl.IS_HIT_I32)
    radar = rgl.node_points_radar_postprocess(n(), [(0.0, 30.0, 1.0, 0.5, math.radians(2.0)),
                                                    (30.0, 120.0, 2.0, 1.0, math.radians(1.0))],
                                              math.radians(120.0 / 256), math.radians(20.0 / 63), 77e9, 31.0, 27.0,
                                              -90.0, 2.0)
    out = rgl.node_points_yield(n(), [XYZ, DIST, AZ, EL, RAD, RCS, POWER, NOISE, SNR])
    rgl.chain(b_rays, pose, rng_node, rt, comp, radar, out)
    rgl.check("rgl_node_gaussian_noise_set_seed", radar.h, C.c_uint64(0xAADA2))
    fields = [XYZ, DIST, AZ, EL, RAD, RCS, POWER, NOISE, SNR]
    for frame in range(3):
        rgl.graph_run(a[0])
        rgl.graph_run_batch([b_rays])
        ra = {f: rgl.result(a[3], f) for f in fields}
        rb = {f: rgl.result(out, f) for f in fields}
        assert_same(ra, rb, f"frame {frame}")
        assert len(ra[DIST]) > 10
