"""GPU parity: the HIP library (through the C-ABI) against the FP64 oracle and against its own brute force.

Bar (per the path's contract): hit mask and entity ids bit-exact except on beams the oracle marks ambiguous
(grazing an edge, a near tie between instances, or at a range limit) and those must stay below 1e-5 of the rays
(at most one ray for small sensors); distances within max(1 mm, 1e-5 relative); normals within 1e-4;
compaction order bit-exact.
"""
import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S
from tests.helpers import Case, Graph, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

XYZ, INTEN, IS_HIT, ENT, DIST, NORMAL, INC, RET = 1, 2, 4, 7, 8, 22, 23, 12
DIST_ABS, DIST_REL, NORMAL_TOL = 1e-3, 1e-5, 1e-4


def compare(gpu, orc, n_rays, label=""):
    amb = orc["ambiguous"].astype(bool)
    R = len(gpu[IS_HIT]) // n_rays
    amb_p = np.repeat(amb, R)
    mism = (gpu[IS_HIT] != orc[IS_HIT]) | ((gpu[IS_HIT] == 1) & (gpu[ENT] != orc[ENT]))
    unexplained = mism & ~amb_p
    print(f"{label}: rays={n_rays} hits={int(gpu[IS_HIT].sum())} mismatches={int(mism.sum())} "
          f"ambiguous={int(amb.sum())} unexplained={int(unexplained.sum())}")
    assert not unexplained.any(), np.nonzero(unexplained)[0][:20]
    assert mism.sum() <= max(1, int(1e-5 * n_rays))
    same = (gpu[IS_HIT] == 1) & (orc[IS_HIT] == 1) & ~mism
    d_g, d_o = gpu[DIST][same].astype(np.float64), orc[DIST][same].astype(np.float64)
    tol = np.maximum(DIST_ABS, DIST_REL * d_o)
    assert np.all(np.abs(d_g - d_o) <= tol), np.max(np.abs(d_g - d_o) - tol)
    if NORMAL in gpu:
        sel = same & ~amb_p
        assert np.all(np.abs(gpu[NORMAL][sel] - orc[NORMAL][sel]) <= NORMAL_TOL)
    nh = (gpu[IS_HIT] == 0) & (orc[IS_HIT] == 0)
    np.testing.assert_array_equal(gpu[DIST][nh], orc[DIST][nh])
    return same


def test_config1_frame_matches_oracle(rgl_lib, oracle):
    spec, lidar = S.config1_scene(), S.config1_lidar()
    case = Case(spec, lidar.rays, pose=lidar.pose, ranges=np.array([lidar.range], np.float32))
    fields = [XYZ, DIST, ENT, NORMAL, IS_HIT]
    gpu = run_gpu(rgl_lib, case, fields)
    orc = run_oracle(oracle, case, fields=fields)
    same = compare(gpu, orc, len(lidar.rays), "config1")
    err = np.abs(gpu[XYZ][same].astype(np.float64) - orc[XYZ][same])
    assert err.max() <= 2e-3


@pytest.mark.parametrize("offset", [(15000.0, -8000.0, 0.0), (-120000.0, 60000.0, 35.0)])
def test_large_world_coordinates_match_oracle(rgl_lib, oracle, offset):
    """The C1 scene and lidar moved far from the world origin (ulp of the coordinates 1-8 mm): the sensor pose is
    composed with the rays in FP64, instance entry maps the FP64 origin, and distances come from FP64 hit points, so
    distances, hit masks, entity ids and normals still meet the parity bar against the FP64 oracle fed the exact
    (FP64) world rays; XYZ, an FP32 field in world coordinates, is within 3 ulp of the coordinates."""
    base, lidar = S.config1_scene(), S.config1_lidar()
    shift = S.translation(*offset)
    spec = S.SceneSpec(list(base.meshes), [S.Instance(i.mesh, S.f32(S.mul(shift, i.transform)), i.entity_id)
                                           for i in base.instances], [])
    pose = S.f32(S.mul(shift, lidar.pose))
    case = Case(spec, lidar.rays, pose=pose, ranges=np.array([lidar.range], np.float32))
    fields = [XYZ, DIST, ENT, NORMAL, IS_HIT]
    gpu = run_gpu(rgl_lib, case, fields)
    orc = run_oracle(oracle, case, fields=fields)
    same = compare(gpu, orc, len(lidar.rays), f"C1 at {offset}")
    ulp = float(np.spacing(np.float32(max(abs(c) for c in offset))))
    assert np.abs(gpu[XYZ][same].astype(np.float64) - orc[XYZ][same]).max() <= 3 * ulp + 2e-3
    assert same.sum() > 30000


def test_config1_compaction_is_stable_and_exact(rgl_lib):
    spec, lidar = S.config1_scene(), S.config1_lidar()
    case = Case(spec, lidar.rays, pose=lidar.pose, ranges=np.array([lidar.range], np.float32))
    fields = [XYZ, DIST, ENT, NORMAL, IS_HIT]
    S.load_into_rgl(spec)
    full = Graph(rgl_lib, case, fields).run()
    compacted = Graph(rgl_lib, case, fields, compact=True).run()
    keep = full[IS_HIT] != 0
    assert len(compacted[IS_HIT]) == keep.sum()
    for f in fields:
        np.testing.assert_array_equal(compacted[f], full[f][keep])


@pytest.mark.parametrize("seed", [11, 12])
def test_bvh_traversal_equals_brute_force(rgl_lib, seed):
    """Same watertight test, all triangles vs LBVH traversal: results must be identical bit for bit."""
    rng = np.random.default_rng(seed)
    spec = S.SceneSpec()
    for m in range(4):
        n = int(rng.integers(1, 400))
        v = rng.uniform(-1, 1, (n * 3, 3)).astype(np.float32)
        spec.add_mesh(v, np.arange(n * 3, dtype=np.int32).reshape(-1, 3))
    for k in range(60):
        spec.add_instance(k % 4, S.trs(rng.uniform(-10, 10, 3), rng.uniform(0, 360, 3), rng.uniform(0.3, 3, 3)),
                          entity_id=k + 1)
    rays = S.make_lidar3d_rays(360, 180, 2, 2)
    case = Case(spec, rays)
    S.load_into_rgl(spec)
    g = Graph(rgl_lib, case, [DIST, IS_HIT, ENT])
    g.run()
    n = len(rays)
    t = np.empty(n, np.float32)
    inst = np.empty(n, np.int32)
    prim = np.empty(n, np.int32)
    import ctypes as C
    rgl_lib.check("rgl_node_raytrace_brute_force", g.raytrace.h, t.ctypes.data_as(C.POINTER(C.c_float)),
                  inst.ctypes.data_as(C.POINTER(C.c_int32)), prim.ctypes.data_as(C.POINTER(C.c_int32)), n)
    res = g.run()
    assert (inst >= 0).sum() > 500
    np.testing.assert_array_equal(res[IS_HIT] == 1, inst >= 0)
    np.testing.assert_array_equal(res[ENT][inst >= 0], inst[inst >= 0] + 1)


def test_bvh_traversal_equals_brute_force_large_trees(rgl_lib):
    """Trees above the agglomerative-build limit (16,384 leaves) take the radix-tree topology: a 20,000-triangle
    ground spanning the scene (sort-key bit for scene-spanning instances) under 16,500 small instances. Traversal
    must still equal brute force bit for bit."""
    rng = np.random.default_rng(77)
    spec = S.SceneSpec()
    ground = spec.add_mesh(*S.grid_mesh(120.0, 120.0, 100, 100))
    cube = spec.add_mesh(S.CUBE_VERTICES, S.CUBE_INDICES)
    spec.add_instance(ground, S.translation(0, 0, -1.0), entity_id=1)
    for k in range(16500):
        spec.add_instance(cube, S.trs((rng.uniform(-55, 55), rng.uniform(-55, 55), rng.uniform(-0.5, 4)),
                                      rng.uniform(0, 360, 3), rng.uniform(0.1, 0.8, 3)), entity_id=k + 2)
    rays = S.make_lidar3d_rays(360, 90, 4, 2)
    case = Case(spec, rays)
    S.load_into_rgl(spec)
    g = Graph(rgl_lib, case, [DIST, IS_HIT, ENT])
    g.run()
    n = len(rays)
    t = np.empty(n, np.float32)
    inst = np.empty(n, np.int32)
    prim = np.empty(n, np.int32)
    import ctypes as C
    rgl_lib.check("rgl_node_raytrace_brute_force", g.raytrace.h, t.ctypes.data_as(C.POINTER(C.c_float)),
                  inst.ctypes.data_as(C.POINTER(C.c_int32)), prim.ctypes.data_as(C.POINTER(C.c_int32)), n)
    res = g.run()
    assert (inst >= 0).sum() > n // 4
    np.testing.assert_array_equal(res[IS_HIT] == 1, inst >= 0)
    np.testing.assert_array_equal(res[ENT][inst >= 0], inst[inst >= 0] + 1)


@pytest.mark.parametrize("count", [1, 100, 20000])
def test_compaction_by_field_from_array(rgl_lib, count):
    """test/src/graph/nodes/CompactByFieldPointsNodeTest.cpp:86-177: random hits, exact stable output."""
    rng = np.random.default_rng(count)
    rec = np.zeros(count, dtype=[("xyz", np.float32, 3), ("hit", np.int32), ("ground", np.int32)])
    rec["xyz"] = rng.normal(size=(count, 3))
    rec["hit"] = rng.integers(0, 2, count)
    rec["ground"] = rng.integers(0, 7, count)
    src = rgl_lib.node_points_from_array(rgl_lib.Node(), rec, [XYZ, IS_HIT, 5])
    comp = rgl_lib.node_points_compact_by_field(rgl_lib.Node(), IS_HIT)
    rgl_lib.add_child(src, comp)
    rgl_lib.graph_run(src)
    keep = rec["hit"] != 0
    np.testing.assert_array_equal(rgl_lib.result(comp, XYZ), rec["xyz"][keep])
    np.testing.assert_array_equal(rgl_lib.result(comp, IS_HIT), rec["hit"][keep])
    np.testing.assert_array_equal(rgl_lib.result(comp, 5), rec["ground"][keep])


@pytest.mark.parametrize("count", [5000, 700001])
def test_compaction_mixed_field_sizes(rgl_lib, count):
    """Stable compaction of 12-, 8-, 4-, 2- and 1-byte fields; 700,001 points span more than 256 tiles (several rounds
    of the tile-offset scan and a partial last tile)."""
    rng = np.random.default_rng(count)
    rec = np.zeros(count, dtype=[("xyz", np.float32, 3), ("hit", np.int32), ("ring", np.uint16), ("rt", np.uint8),
                                 ("ts", np.float64)])
    rec["xyz"] = rng.normal(size=(count, 3))
    rec["hit"] = rng.integers(0, 2, count) * rng.integers(1, 5, count)
    rec["ring"] = rng.integers(0, 65535, count)
    rec["rt"] = rng.integers(0, 255, count)
    rec["ts"] = rng.normal(size=count)
    fields = [XYZ, IS_HIT, 11, 12, 13]  # RING_ID_U16, RETURN_TYPE_U8, TIME_STAMP_F64
    src = rgl_lib.node_points_from_array(rgl_lib.Node(), rec, fields)
    comp = rgl_lib.node_points_compact_by_field(rgl_lib.Node(), IS_HIT)
    rgl_lib.add_child(src, comp)
    rgl_lib.graph_run(src)
    keep = rec["hit"] != 0
    np.testing.assert_array_equal(rgl_lib.result(comp, XYZ), rec["xyz"][keep])
    np.testing.assert_array_equal(rgl_lib.result(comp, IS_HIT), rec["hit"][keep])
    np.testing.assert_array_equal(rgl_lib.result(comp, 11), rec["ring"][keep])
    np.testing.assert_array_equal(rgl_lib.result(comp, 12), rec["rt"][keep])
    np.testing.assert_array_equal(rgl_lib.result(comp, 13), rec["ts"][keep])


def test_compaction_all_non_hit_gives_empty_cloud(rgl_lib):
    rec = np.zeros(100, dtype=[("xyz", np.float32, 3), ("hit", np.int32)])
    src = rgl_lib.node_points_from_array(rgl_lib.Node(), rec, [XYZ, IS_HIT])
    comp = rgl_lib.node_points_compact_by_field(rgl_lib.Node(), IS_HIT)
    rgl_lib.add_child(src, comp)
    rgl_lib.graph_run(src)
    assert rgl_lib.result_size(comp, XYZ) == (0, 12)


def test_noise_stream_replays_on_cpu(rgl_lib, oracle):
    """GPU normals are bit-identical to the oracle's independent CPU replay of the same stream."""
    for seed, run, tag in [(0, 0, 1), (0xDEADBEEFCAFEF00D, 7, 3), (42, 123456, 2)]:
        g = rgl_lib.noise_normals(seed, run, tag, 100_000)
        c = oracle.normals(seed, run, tag, 100_000)
        np.testing.assert_array_equal(g.view(np.uint32), c.view(np.uint32))
