"""Pins the FP64 oracle itself (CPU only): BVH traversal against brute force, the noise stream against published
Philox4x32-10 known answers and normal statistics, the single-precision log/sin/cos against libm, stable
compaction against its definition, and the ambiguity diagnostic against constructed grazing and tie rays."""
import math

import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S
from tests.helpers import Case, cube_scene, identity_ray, run_oracle

IS_HIT, DIST, ENT, NORMAL = 4, 8, 7, 22


def random_scene(seed: int, n_inst=20, n_tris=60):
    rng = np.random.default_rng(seed)
    spec = S.SceneSpec()
    for m in range(3):
        v = rng.uniform(-1, 1, (n_tris * 3, 3)).astype(np.float32)
        i = np.arange(n_tris * 3, dtype=np.int32).reshape(-1, 3)
        spec.add_mesh(v, i)
    for k in range(n_inst):
        spec.add_instance(k % 3, S.trs(rng.uniform(-8, 8, 3), rng.uniform(0, 360, 3), rng.uniform(0.5, 2, 3)),
                          entity_id=k + 1)
    return spec


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_bvh_equals_brute_force(oracle, seed):
    """The oracle's BVH must find exactly the all-triangle closest hit (no culling mistakes)."""
    spec = random_scene(seed)
    rays = S.make_lidar3d_rays(360, 180, 6, 6)
    case = Case(spec, rays)
    a = run_oracle(oracle, case, fields=[IS_HIT, DIST, ENT])
    b = run_oracle(oracle, case, fields=[IS_HIT, DIST, ENT], brute_force=True)
    assert a[IS_HIT].sum() > 100
    for k in (IS_HIT, DIST, ENT, "primaryInst", "primaryPrim"):
        np.testing.assert_array_equal(a[k], b[k])


def test_philox_known_answers(oracle):
    """Random123 kat_vectors for philox4x32 with 10 rounds."""
    assert list(oracle.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_normal_statistics(oracle):
    n = oracle.normals(0x1234, 0, 2, 1_000_000).astype(np.float64)
    assert abs(n.mean()) < 0.003 and abs(n.std() - 1.0) < 0.003
    # tail mass of a standard normal: P(|x| > 2) = 0.0455, P(|x| > 3) = 0.0027
    assert abs((np.abs(n) > 2).mean() - 0.0455) < 0.001
    assert abs((np.abs(n) > 3).mean() - 0.0027) < 0.0003
    # different runs / tags give different streams
    assert not np.array_equal(oracle.normals(0x1234, 1, 2, 64), oracle.normals(0x1234, 0, 2, 64))


def test_log_and_sincos_accuracy(oracle):
    rng = np.random.default_rng(5)
    xs = rng.uniform(2.0 ** -24, 1.0, 2000).astype(np.float32)
    for x in xs:
        exact = math.log(float(x))
        assert abs(oracle.logf(float(x)) - exact) <= 3e-7 * max(1.0, abs(exact))
    for u in rng.uniform(0, 1, 2000).astype(np.float32):
        s, c = oracle.sincos2pi(float(u))
        assert abs(s - math.sin(2 * math.pi * float(u))) < 3e-7
        assert abs(c - math.cos(2 * math.pi * float(u))) < 3e-7


def test_compaction_definition(oracle):
    flags = np.array([0, 1, 1, 0, 5, 0, -1], np.int32)
    vals = np.arange(7)
    (out,) = oracle.compact(flags, vals)
    assert list(out) == [1, 2, 4, 6]


def test_ambiguity_flags_edges_and_ties(oracle):
    # A ray through the shared edge of two coplanar triangles is grazing; one through a face centre is not.
    spec = cube_scene(S.translation(0, 0, 5))
    rays = S.f32(np.stack([S.identity(), S.translation(0.5, 0.25, 0)]))
    r = run_oracle(oracle, Case(spec, rays), fields=[DIST])
    assert list(r["ambiguous"]) == [1, 0]
    # Two coincident cubes with different ids tie at equal distance; the lower instance index wins.
    spec2 = cube_scene(S.translation(0, 0, 5), S.translation(0, 0, 5), ids=[7, 8])
    r2 = run_oracle(oracle, Case(spec2, S.f32(S.translation(0.5, 0.25, 0)[None])), fields=[ENT])
    assert r2["ambiguous"][0] == 1 | 16 and r2[ENT][0] == 7  # geometric ambiguity, of the instance-tie kind


def test_vertex_displacement_velocity_closed_form(oracle):
    """Vertex-animation velocity (src/gpu/optixPrograms.cu:186-196): for per-vertex displacements of an affine field
    d(v) = A v + b, barycentric interpolation is exact, so a hit at object point p moves at M3 (A p + b) / dt, where
    M3 is the linear part of the entity transform (scale included). Catches a wrong vertex index, a missing rotation
    or an interpolation weight on the wrong vertex."""
    rng = np.random.default_rng(4)
    tf = S.trs((1.0, 2.0, 3.0), (10.0, 20.0, 30.0), (2.0, 1.5, 1.0))
    spec = cube_scene(tf)
    A = rng.uniform(-0.2, 0.2, (3, 3))
    b = rng.uniform(-0.1, 0.1, 3)
    disp = S.CUBE_VERTICES.astype(np.float64) @ A.T + b
    dirs = rng.normal(size=(64, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rays = np.zeros((64, 3, 4), np.float32)
    for k, d in enumerate(dirs):  # rays from the entity's centre; local +Z of the ray frame along d
        z = d
        x = np.cross(z, [0.0, 0.0, 1.0] if abs(z[2]) < 0.9 else [1.0, 0.0, 0.0])
        x /= np.linalg.norm(x)
        y = np.cross(z, x)
        rays[k, :, :3] = np.stack([x, y, z], axis=1)
        rays[k, :, 3] = (1.0, 2.0, 3.0)
    dt = 0.1
    fields = [1, IS_HIT, 15]
    orc = oracle.raytrace(spec, rays, scene_dt=dt, prev_transforms=np.stack([tf]), displacements={0: disp},
                          fields=fields)
    assert orc[IS_HIT].all()
    m3, t = np.asarray(tf, np.float64)[:, :3], np.asarray(tf, np.float64)[:, 3]
    p = np.linalg.solve(m3, (orc[1].astype(np.float64) - t).T).T
    expected = (p @ A.T + b) @ m3.T / dt
    np.testing.assert_allclose(orc[15], expected, atol=2e-5)
    # Without displacements (and an unchanged transform) the hit points do not move.
    still = oracle.raytrace(spec, rays, scene_dt=dt, prev_transforms=np.stack([tf]), fields=fields)
    assert np.abs(still[15]).max() < 1e-5
