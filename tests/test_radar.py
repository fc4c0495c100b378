"""Radar binning: the oracle's restatement pinned by closed forms (CPU), and the HIP stage against the oracle and
the same closed forms (GPU)."""
import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S

XYZ, IS_HIT, DIST, AZ, RCS, INC = 1, 4, 8, 9, 19, 23
RANGE_STEP, RANGE_MAX, AZ_STEP = 0.2, 200.0, np.deg2rad(1.0)
AZ_MIN, AZ_MAX = -np.deg2rad(60.0), np.deg2rad(60.0)
OMEGA = 2.0 ** -10  # exactly representable, so d^2 * OMEGA is exact for the wall distance below


def f32(x):
    return float(np.float32(x))


# Azimuth cells of the grid as the stage sizes it from its FP32 parameters (121: 120 degrees / 1 degree rounds up).
AZ_CELLS = int(np.ceil((f32(AZ_MAX) - f32(AZ_MIN)) / f32(AZ_STEP)))


X_WALL = 10.0625  # 161/16: its square is exact in FP32 and it lies inside a 0.2 m range cell


def wall_rays(n_y=41, n_z=5, x_wall=X_WALL):
    """Parallel rays along +X from (0, y, z) hitting a wall at x = x_wall head-on; y spaced so every point's azimuth
    atan2(y, x_wall) lies away from 1-degree cell edges."""
    ys = x_wall * np.tan(np.deg2rad(np.arange(-20, 21)[:n_y] + 0.5))
    zs = np.linspace(-1.0, 1.0, n_z)
    rays = []
    for z in zs:
        for y in ys:
            m = S.mul(S.translation(0.0, y, z), S.rotation_deg(0.0, 90.0, 0.0))  # local +Z -> world +X
            rays.append(m)
    return S.f32(np.stack(rays)), ys, zs


def expected_wall_cells(ys, n_z, x_wall=X_WALL):
    """Closed form: every ray hits at d = x_wall with incidence 0, so it adds x_wall^2 * OMEGA m^2 to the cell of
    distance x_wall and azimuth atan2(y, x_wall); each y column holds n_z rays."""
    step, az_min = f32(RANGE_STEP), f32(AZ_MIN)
    rc_idx = int(np.floor(x_wall / step))
    ac = AZ_CELLS
    cells = {}
    for y in ys:
        a = int(np.floor((np.arctan2(y, x_wall) - az_min) / f32(AZ_STEP)))
        cells[rc_idx * ac + a] = cells.get(rc_idx * ac + a, 0) + n_z
    keys = np.array(sorted(cells))
    counts = np.array([cells[k] for k in keys])
    return keys, counts, 10.0 * np.log10(counts * x_wall ** 2 * OMEGA)


def test_oracle_radar_bin_closed_form(oracle):
    rays, ys, zs = wall_rays()
    n = len(rays)
    xyz = np.stack([np.full(n, X_WALL), np.tile(ys, len(zs)), np.repeat(zs, len(ys))], 1)
    cells, counts, rcs, amb = oracle.radar_bin(xyz, np.full(n, X_WALL), np.zeros(n), np.ones(n, np.int32), RANGE_STEP,
                                               RANGE_MAX, AZ_STEP, AZ_MIN, AZ_MAX, OMEGA)
    k, c, r = expected_wall_cells(ys, len(zs))
    np.testing.assert_array_equal(cells, k)
    np.testing.assert_array_equal(counts, c)
    np.testing.assert_allclose(rcs, r, atol=1e-9)
    assert not any(amb)


def test_oracle_radar_bin_drops_outside_and_non_hits(oracle):
    xyz = np.array([[10.0, 0.0, 0.0], [10.0, 5.0, 0.0], [10.0, 20.0, 0.0], [250.0, 0.0, 0.0], [10.0, 0.0, 0.0]])
    d = np.linalg.norm(xyz, axis=1)
    cells, counts, _, _ = oracle.radar_bin(xyz, d, np.zeros(5), np.array([1, 1, 1, 1, 0]), RANGE_STEP, RANGE_MAX,
                                           AZ_STEP, AZ_MIN, AZ_MAX, OMEGA)
    # kept: azimuth 0 and 26.6 deg; dropped: 63.4 deg (outside +-60), 250 m (beyond 200 m), and the non-hit
    assert counts.sum() == 2
    assert len(cells) == 2


def radar_graph(rgl, rays, pose, scene=None, far=None):
    n = rgl.Node
    src = rgl.node_rays_from_mat3x4f(n(), rays)
    chain = [src]
    if pose is not None:
        chain.append(rgl.node_rays_transform(n(), pose))
    rt = rgl.node_raytrace(n(), scene)
    chain.append(rt)
    radar = rgl.node_points_radar_bin(n(), RANGE_STEP, RANGE_MAX, AZ_STEP, AZ_MIN, AZ_MAX, OMEGA)
    chain.append(radar)
    rgl.chain(*chain)
    raw = rgl.node_points_yield(n(), [XYZ, DIST, INC, IS_HIT])
    rgl.add_child(rt, raw)
    return src, raw, radar


@pytest.mark.gpu
def test_gpu_radar_bin_closed_form(rgl_lib):
    rgl = rgl_lib
    spec = S.SceneSpec()
    gv, gi = S.grid_mesh(100.0, 100.0, 4, 4)
    wall = spec.add_mesh(gv, gi)
    spec.add_instance(wall, S.mul(S.translation(X_WALL, 0.0, 0.0), S.rotation_deg(0.0, 90.0, 0.0)), entity_id=5)
    S.load_into_rgl(spec)
    rays, ys, zs = wall_rays()
    src, raw, radar = radar_graph(rgl, rays, None)
    rgl.graph_run(src)
    d = rgl.result(raw, DIST)
    assert np.all(rgl.result(raw, IS_HIT) == 1) and np.all(d == np.float32(X_WALL))
    k, c, r = expected_wall_cells(ys, len(zs))
    dist = rgl.result(radar, DIST)
    az = rgl.result(radar, AZ)
    rcs = rgl.result(radar, RCS)
    assert len(rcs) == len(k)
    ac = AZ_CELLS
    np.testing.assert_allclose(dist, (k // ac + 0.5) * f32(RANGE_STEP), rtol=1e-6)
    np.testing.assert_allclose(az, f32(AZ_MIN) + (k % ac + 0.5) * f32(AZ_STEP), atol=1e-6)
    np.testing.assert_allclose(rcs, r, atol=1e-5)
    np.testing.assert_array_equal(rgl.result(radar, IS_HIT), 1)


@pytest.mark.gpu
def test_gpu_radar_bin_matches_oracle_on_c1(rgl_lib, oracle):
    """A 256 x 64 radar (+-60 deg azimuth, +-10 deg elevation) on the C1 scene: the binning of the GPU's own hits
    equals the oracle's binning of the same hits (cell sets and counts exact except for points within FP32 reach
    of a cell edge; RCS within 1e-4 dB on unaffected cells)."""
    rgl = rgl_lib
    S.load_into_rgl(S.config1_scene())
    # 256 azimuths over +-60 deg, offset so that no beam points exactly along a 1-degree cell edge
    az = -60.0 + (np.arange(256) + 0.37) * (120.0 / 256)
    el = np.linspace(-10.0, 10.0, 64)
    rays = []
    for e in el:
        for a in az:
            rays.append(S.mul(S.rotation_deg(0.0, 0.0, a), S.rotation_deg(0.0, 90.0 - e, 0.0)))
    rays = S.f32(np.stack(rays))
    pose = S.f32(S.trs((-30.0, 5.0, 1.0), (0, 0, 20.0)))
    src, raw, radar = radar_graph(rgl, rays, pose)
    rgl.graph_run(src)
    xyz, d, inc, hit = (rgl.result(raw, f) for f in (XYZ, DIST, INC, IS_HIT))
    assert hit.sum() > 5000
    p64 = np.vstack([pose.astype(np.float64), [0, 0, 0, 1]])
    inv = np.linalg.inv(p64)
    xyz_s = np.where(hit[:, None] == 1, xyz.astype(np.float64), 0.0) @ inv[:3, :3].T + inv[:3, 3]
    cells, counts, rcs_db, alt = oracle.radar_bin(xyz_s, d, inc, hit, RANGE_STEP, RANGE_MAX, AZ_STEP, AZ_MIN, AZ_MAX,
                                                  OMEGA)
    ac = AZ_CELLS
    # cell centres decode back to their cell indices
    g_cells = (np.floor(rgl.result(radar, DIST).astype(np.float64) / f32(RANGE_STEP)).astype(np.int64) * ac +
               np.floor((rgl.result(radar, AZ).astype(np.float64) - f32(AZ_MIN)) / f32(AZ_STEP)).astype(np.int64))
    g_rcs = rgl.result(radar, RCS)
    # Cells a point within FP32 reach of a cell edge may land in (its own and the neighbour) may differ.
    rpos = d.astype(np.float64) / f32(RANGE_STEP)
    apos = (np.arctan2(xyz_s[:, 1], xyz_s[:, 0]) - f32(AZ_MIN)) / f32(AZ_STEP)
    own = np.floor(rpos).astype(np.int64) * ac + np.floor(apos).astype(np.int64)
    touched = set()
    for i, cs in enumerate(alt):
        if cs:
            touched.add(int(own[i]))
            touched.update(cs)
    n_edge = sum(1 for cs in alt if cs)
    o_map = dict(zip(cells.tolist(), zip(counts.tolist(), rcs_db.tolist())))
    g_map = dict(zip(g_cells.tolist(), g_rcs.tolist()))
    clean = [c for c in o_map if c not in touched]
    print(f"radar: {int(hit.sum())} hits, {len(o_map)} oracle cells, {len(g_map)} GPU cells, "
          f"{n_edge} edge points, {len(clean)} cells compared")
    assert len(clean) > 0.9 * len(o_map)
    for c in clean:
        assert c in g_map, c
        assert abs(g_map[c] - o_map[c][1]) <= 1e-4, (c, g_map[c], o_map[c])
    assert set(g_map) - set(o_map) <= touched
    assert np.all(np.diff(g_cells) > 0)  # cell order
