"""BVH node formats: the 4-wide FP32 and the 4-wide 8-bit- and 16-bit-quantized trees must give results
bit-identical to the binary FP32 tree (which the parity tests pin against the FP64 oracle and brute force), on random
triangle soups, flat and single-triangle meshes, far-from-origin coordinates, the C1 and urban scenes, batches, and
TLAS refits. Quantized child bounds lie outside the exact ones by a margin that covers the traversal's folded decode,
so no reachable triangle is culled.
"""
import numpy as np
import pytest

from robotecgpulidar_amd import rgl as R
from robotecgpulidar_amd import scenes as S
from tests.helpers import Case, Graph, run_oracle
from tests.test_gpu_parity import compare

pytestmark = pytest.mark.gpu

XYZ, IS_HIT, ENT, DIST, NORMAL = 1, 4, 7, 8, 22
FIELDS = [XYZ, DIST, ENT, NORMAL, IS_HIT]
FORMATS = [R.BVH2, R.BVH4_F32, R.BVH4_Q8, R.BVH4_Q16]


@pytest.fixture
def lib(rgl_lib):
    default = rgl_lib.get_bvh_format()
    yield rgl_lib
    rgl_lib.set_bvh_format(default)


def run_formats(rgl, graph):
    out = {}
    for f in FORMATS:
        rgl.set_bvh_format(f)
        assert rgl.get_bvh_format() == f
        out[f] = graph.run()
    return out


def assert_identical(res, label=""):
    base = res[R.BVH2]
    for f in FORMATS[1:]:
        for k in base:
            assert res[f][k].tobytes() == base[k].tobytes(), (label, f, k)


def soup(seed, offset=(0.0, 0.0, 0.0), n_inst=60):
    rng = np.random.default_rng(seed)
    spec = S.SceneSpec()
    for _ in range(4):
        n = int(rng.integers(1, 400))
        v = rng.uniform(-1, 1, (n * 3, 3)).astype(np.float32)
        spec.add_mesh(v, np.arange(n * 3, dtype=np.int32).reshape(-1, 3))
    spec.add_mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.int32))
    gv, gi = S.grid_mesh(20.0, 20.0, 7, 5)  # flat: zero extent along z
    spec.add_mesh(gv, gi)
    for k in range(n_inst):
        t = np.asarray(offset) + rng.uniform(-10, 10, 3)
        spec.add_instance(k % 6, S.trs(t, rng.uniform(0, 360, 3), rng.uniform(0.3, 3, 3)), entity_id=k + 1)
    return spec


@pytest.mark.parametrize("seed,offset,n_inst", [(11, (0, 0, 0), 60), (12, (5000.0, -3000.0, 200.0), 60),
                                                (13, (0, 0, 0), 1), (14, (-1e4, 0, 0), 7)])
def test_formats_identical_on_random_soups(lib, seed, offset, n_inst):
    spec = soup(seed, offset, n_inst)
    rays = S.make_lidar3d_rays(360, 180, 2, 2)
    pose = S.f32(S.translation(*offset)) if any(offset) else None
    S.load_into_rgl(spec)
    g = Graph(lib, Case(spec, rays, pose=pose), FIELDS)
    res = run_formats(lib, g)
    hits = int(res[R.BVH2][IS_HIT].sum())
    print(f"seed {seed}: {hits} hits of {len(rays)}")
    if n_inst > 1:
        assert hits > 500
    assert_identical(res, f"soup {seed}")


def test_formats_identical_and_oracle_on_c1(lib, oracle):
    spec, lidar = S.config1_scene(), S.config1_lidar()
    case = Case(spec, lidar.rays, pose=lidar.pose, ranges=np.array([lidar.range], np.float32))
    S.load_into_rgl(spec)
    res = run_formats(lib, Graph(lib, case, FIELDS))
    assert_identical(res, "c1")
    orc = run_oracle(oracle, case, fields=FIELDS)
    compare(res[R.BVH4_Q8], orc, len(lidar.rays), "c1 quantized")
    compare(res[R.BVH4_Q16], orc, len(lidar.rays), "c1 16-bit quantized")


def test_formats_identical_on_urban_scene(lib):
    spec = S.config2_scene()
    lidar = S.config2_lidar()
    rays = lidar.rays.reshape(128, 2048, 3, 4)[::4].reshape(-1, 3, 4)  # 32 of the 128 channels
    case = Case(spec, rays, pose=lidar.pose, ranges=np.array([[0.0, 200.0]], np.float32))
    S.load_into_rgl(spec)
    res = run_formats(lib, Graph(lib, case, FIELDS, compact=True))
    print(f"urban: {len(res[R.BVH2][XYZ])} kept points")
    assert len(res[R.BVH2][XYZ]) > 20000
    assert_identical(res, "urban")


def test_formats_identical_under_tlas_refit(lib):
    spec = soup(21, n_inst=40)
    rays = S.make_lidar3d_rays(360, 180, 2, 2)
    _, ents, _ = S.load_into_rgl(spec)
    lib.check("rgl_scene_set_tlas_refit", None, True)
    g = Graph(lib, Case(spec, rays), FIELDS)
    rng = np.random.default_rng(5)
    moves = [[S.f32(S.trs(rng.uniform(-10, 10, 3), rng.uniform(0, 360, 3), (1, 1, 1))) for _ in range(8)]
             for _ in range(3)]
    per_format = {}
    for f in FORMATS:
        lib.set_bvh_format(f)
        for k, inst in enumerate(spec.instances):  # restore the initial pose, then the same motion per format
            lib.entity_set_transform(ents[k], inst.transform)
        frames = [g.run()]
        for step in moves:
            for j, m in enumerate(step):
                lib.entity_set_transform(ents[3 * j], m)
            frames.append(g.run())
        per_format[f] = frames
    for i in range(len(moves) + 1):
        assert_identical({f: per_format[f][i] for f in FORMATS}, f"refit frame {i}")


def test_formats_identical_in_batches(lib):
    from tests.test_gpu_batch import assert_same, build_world, make_set
    _, scenes, _ = build_world(lib)
    graphs = make_set(lib, scenes)
    per_format = {}
    for f in FORMATS:
        lib.set_bvh_format(f)
        for g in graphs:
            for k, n in enumerate(g.noise):  # replay the same noise under every format
                lib.check("rgl_node_gaussian_noise_set_seed", n.h, __import__("ctypes").c_uint64(0x5EED + k))
        lib.graph_run_batch([g.rays for g in graphs])
        per_format[f] = [g.results() for g in graphs]
    for f in FORMATS[1:]:
        for k, (a, b) in enumerate(zip(per_format[R.BVH2], per_format[f])):
            assert_same(a, b, f"format {f} sensor {k}")
