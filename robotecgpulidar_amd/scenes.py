"""Seeded synthetic scenes and sensors, and loading them through the RGL API.

A SceneSpec is plain arrays (meshes, instances, textures) that both the HIP library (through rgl_* calls) and the
FP64 oracle consume, so parity tests feed identical inputs to both. Generators cover the unit-cube fixtures of the
reference's tests (test/include/helpers/geometryData.hpp:5-57, lidarHelpers.hpp:7-47) and the benchmark
workloads: a 57,600-ray lidar over a ground grid with 64 boxes, and a ~2M-triangle urban scene with a
262,144-ray lidar.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# ------------------------------------------------------------------------------------------ transforms


def identity() -> np.ndarray:
    return np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]], dtype=np.float64)


def translation(x, y, z) -> np.ndarray:
    m = identity()
    m[:, 3] = (x, y, z)
    return m


def scale(x, y, z) -> np.ndarray:
    return np.array([[x, 0, 0, 0], [0, y, 0, 0], [0, 0, z, 0]], dtype=np.float64)


def rotation_rad(x, y, z) -> np.ndarray:
    """Rz(z) * Ry(y) * Rx(x), the Euler composition of the reference (src/math/Mat3x4f.hpp:38-60)."""
    cr, sr, cp, sp, cy, sy = math.cos(x), math.sin(x), math.cos(y), math.sin(y), math.cos(z), math.sin(z)
    return np.array([
        [cp * cy, sr * sp * cy - cr * sy, sr * sy + cr * sp * cy, 0],
        [cp * sy, cr * cy + sr * sp * sy, cr * sp * sy - sr * cy, 0],
        [-sp, sr * cp, cr * cp, 0],
    ], dtype=np.float64)


def rotation_deg(x, y, z) -> np.ndarray:
    k = math.pi / 180.0
    return rotation_rad(x * k, y * k, z * k)


def mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    r = a[:, :3] @ b
    r[:, 3] += a[:, 3]
    return r


def trs(t=(0, 0, 0), rdeg=(0, 0, 0), s=(1, 1, 1)) -> np.ndarray:
    return mul(mul(translation(*t), rotation_deg(*rdeg)), scale(*s))


def f32(m) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(m, dtype=np.float32))


# ------------------------------------------------------------------------------------------ fixtures
# Unit cube with outward right-handed winding and per-vertex UVs (the reference's test fixture geometry).
CUBE_VERTICES = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                          [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float32)
CUBE_INDICES = np.array([[0, 3, 1], [3, 2, 1], [1, 2, 5], [2, 6, 5], [5, 6, 4], [6, 7, 4],
                         [4, 7, 0], [7, 3, 0], [3, 7, 2], [7, 6, 2], [4, 0, 5], [0, 1, 5]], dtype=np.int32)
CUBE_UVS = np.array([[0, 0], [1, 0], [1, 1], [0, 1], [0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float32)
CUBE_HALF_EDGE = 1.0


def make_lidar3d_rays(fov_x, fov_y, res_x=1.0, res_y=1.0) -> np.ndarray:
    """FOV grid of rotationDeg(x, y, 0) rays, accumulated in float32 like lidarHelpers.hpp:7-29."""
    fov_x = np.float32(min(max(0.0, fov_x), 360.0))
    fov_y = np.float32(min(max(0.0, fov_y), 360.0))
    rx, ry = np.float32(res_x), np.float32(res_y)
    sx, sy = np.float32(-1) * fov_x / np.float32(2), np.float32(-1) * fov_y / np.float32(2)
    rays = []
    ax = np.float32(0)
    while ax <= fov_x:
        rot_x = sx + ax
        ay = np.float32(0)
        while ay <= fov_y:
            rot_y = sy + ay
            rays.append(rotation_deg(float(rot_x), float(rot_y), 0.0))
            ay = np.float32(ay + ry)
        ax = np.float32(ax + rx)
    return f32(np.stack(rays))


def make_grid_of_parallel_rays(min_xy, max_xy, count) -> np.ndarray:
    sy = (max_xy[1] - min_xy[1]) / (count[1] + 1)
    sx = (max_xy[0] - min_xy[0]) / (count[0] + 1)
    rays = []
    for y in range(1, count[1] + 1):
        for x in range(1, count[0] + 1):
            rays.append(translation(min_xy[0] + sx * x, min_xy[1] + sy * y, 0.0))
    return f32(np.stack(rays))


# ------------------------------------------------------------------------------------------ scene spec
@dataclass
class Mesh:
    vertices: np.ndarray  # (V, 3) float32
    indices: np.ndarray   # (T, 3) int32
    uvs: np.ndarray | None = None  # (V, 2) float32


@dataclass
class Instance:
    mesh: int
    transform: np.ndarray  # (3, 4) float32
    entity_id: int | None = None  # None: library default id
    texture: int = -1
    laser_retro: float = 0.0


@dataclass
class SceneSpec:
    meshes: list[Mesh] = field(default_factory=list)
    instances: list[Instance] = field(default_factory=list)
    textures: list[np.ndarray] = field(default_factory=list)  # (H, W) uint8

    def add_mesh(self, v, i, uvs=None) -> int:
        self.meshes.append(Mesh(np.ascontiguousarray(v, dtype=np.float32), np.ascontiguousarray(i, dtype=np.int32),
                                None if uvs is None else np.ascontiguousarray(uvs, dtype=np.float32)))
        return len(self.meshes) - 1

    def add_instance(self, mesh, transform, entity_id=None, texture=-1, laser_retro=0.0) -> int:
        self.instances.append(Instance(mesh, f32(transform), entity_id, texture, laser_retro))
        return len(self.instances) - 1

    @property
    def triangle_count(self) -> int:
        return sum(len(self.meshes[i.mesh].indices) for i in self.instances)


@dataclass
class Lidar:
    rays: np.ndarray  # (N, 3, 4) float32, sensor frame
    pose: np.ndarray  # (3, 4) float32, applied with rgl_node_rays_transform
    range: tuple[float, float]
    time_offsets: np.ndarray | None = None  # ms per ray
    channels: int = 0
    azimuths: int = 0


def load_into_rgl(spec: SceneSpec, scene=None, meshes=None):
    """Creates meshes, textures and entities through the RGL API (in `scene`, None = the default scene).

    `meshes` reuses mesh handles created earlier for the same spec. Returns (meshes, entities, textures)."""
    from . import rgl
    if meshes is not None:
        meshes = list(meshes)
    else:
        meshes = []
        for m in spec.meshes:
            h = rgl.mesh_create(m.vertices, m.indices)
            if m.uvs is not None:
                rgl.mesh_set_texture_coords(h, m.uvs)
            meshes.append(h)
    return _load_entities(spec, scene, meshes)


def _load_entities(spec: SceneSpec, scene, meshes):
    from . import rgl
    textures = [rgl.texture_create(t) for t in spec.textures]
    entities = []
    for inst in spec.instances:
        e = rgl.entity_create(meshes[inst.mesh], scene)
        rgl.entity_set_transform(e, inst.transform)
        if inst.entity_id is not None:
            rgl.entity_set_id(e, inst.entity_id)
        if inst.texture >= 0:
            rgl.check("rgl_entity_set_intensity_texture", e, textures[inst.texture])
        if inst.laser_retro != 0.0:
            rgl.check("rgl_entity_set_laser_retro", e, float(inst.laser_retro))
        entities.append(e)
    return meshes, entities, textures


# ------------------------------------------------------------------------------------------ primitives
def grid_mesh(size_x: float, size_y: float, cells_x: int, cells_y: int):
    xs = np.linspace(-size_x / 2, size_x / 2, cells_x + 1)
    ys = np.linspace(-size_y / 2, size_y / 2, cells_y + 1)
    gx, gy = np.meshgrid(xs, ys, indexing="xy")
    v = np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)], axis=1).astype(np.float32)
    i0 = (np.arange(cells_y)[:, None] * (cells_x + 1) + np.arange(cells_x)[None, :]).ravel()
    a, b, c, d = i0, i0 + 1, i0 + cells_x + 2, i0 + cells_x + 1
    tris = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)], axis=0).astype(np.int32)
    return v, tris


def icosphere(subdivisions: int):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                p = v[i] + v[j]
                v.append(p / np.linalg.norm(p))
                cache[key] = len(v) - 1
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(v, dtype=np.float32), np.array(f, dtype=np.int32)


def prism(n: int):
    """Unit-radius n-gon prism from z=0 to z=1: 2n side triangles + 2(n-2) cap triangles = 4n-4."""
    ang = np.arange(n) * (2 * math.pi / n)
    bottom = np.stack([np.cos(ang), np.sin(ang), np.zeros(n)], 1)
    top = np.stack([np.cos(ang), np.sin(ang), np.ones(n)], 1)
    v = np.concatenate([bottom, top]).astype(np.float32)
    tris = []
    for k in range(n):
        k1 = (k + 1) % n
        tris += [(k, k1, n + k1), (k, n + k1, n + k)]
    for k in range(1, n - 1):
        tris += [(0, k + 1, k), (n, n + k, n + k + 1)]
    return v, np.array(tris, dtype=np.int32)


def spinning_lidar(channels: int, azimuths: int, el_min_deg: float, el_max_deg: float) -> np.ndarray:
    """Sensor-frame rays Rz(az) * Ry(90 deg - el): local +Z maps to the beam direction."""
    el = np.deg2rad(np.linspace(el_min_deg, el_max_deg, channels))
    az = np.arange(azimuths) * (2 * math.pi / azimuths)
    A, E = np.meshgrid(az, el, indexing="xy")  # (channels, azimuths)
    A, E = A.ravel(), E.ravel()
    th = np.pi / 2 - E
    ca, sa, ct, st = np.cos(A), np.sin(A), np.cos(th), np.sin(th)
    m = np.zeros((A.size, 3, 4))
    # Rz(a) * Ry(t)
    m[:, 0, 0], m[:, 0, 1], m[:, 0, 2] = ca * ct, -sa, ca * st
    m[:, 1, 0], m[:, 1, 1], m[:, 1, 2] = sa * ct, ca, sa * st
    m[:, 2, 0], m[:, 2, 1], m[:, 2, 2] = -st, 0.0, ct
    return f32(m)


# ------------------------------------------------------------------------------------------ workloads
def config1_scene(seed: int = 0xC0F1) -> SceneSpec:
    """200x200 m ground grid (2x100x100 triangles) plus 64 boxes, sizes U[1,5] m, positions U[-90,90] m."""
    rng = np.random.default_rng(seed)
    spec = SceneSpec()
    gv, gi = grid_mesh(200.0, 200.0, 100, 100)
    ground = spec.add_mesh(gv, gi)
    cube = spec.add_mesh(CUBE_VERTICES, CUBE_INDICES)
    spec.add_instance(ground, identity(), entity_id=1)
    for k in range(64):
        sx, sy, sz = rng.uniform(1.0, 5.0, 3)
        px, py = rng.uniform(-90.0, 90.0, 2)
        spec.add_instance(cube, mul(translation(px, py, sz / 2), scale(sx / 2, sy / 2, sz / 2)), entity_id=2 + k)
    return spec


def config1_lidar() -> Lidar:
    """32 channels over [-15, +15] deg x 1800 azimuths = 57,600 rays at (0, 0, 1.8), range [0.1, 100] m."""
    return Lidar(spinning_lidar(32, 1800, -15.0, 15.0), f32(translation(0, 0, 1.8)), (0.1, 100.0), channels=32,
                 azimuths=1800)


def config2_scene(seed: int = 0xC0F2, target_tris: int = 2_000_000, instances: int = 500) -> SceneSpec:
    """Urban scene: `instances` instanced icospheres (subdivision 3), boxes and 4-12 sided building prisms (heights
    U[3,60] m) under random yaw, on a ground grid sized so the total is about `target_tris` triangles."""
    rng = np.random.default_rng(seed)
    spec = SceneSpec()
    sv, si = icosphere(3)
    sphere = spec.add_mesh(sv, si)
    cube = spec.add_mesh(CUBE_VERTICES, CUBE_INDICES)
    prisms = {n: spec.add_mesh(*prism(n)) for n in range(4, 13)}
    kinds = rng.integers(0, 3, instances)
    inst_tris = 0
    placed = []
    for k in range(instances):
        yaw = rng.uniform(0.0, 360.0)
        px, py = rng.uniform(-190.0, 190.0, 2)
        if kinds[k] == 0:
            r = rng.uniform(0.5, 3.0)
            placed.append((sphere, trs((px, py, r), (0, 0, yaw), (r, r, r))))
            inst_tris += len(si)
        elif kinds[k] == 1:
            sx, sy, sz = rng.uniform(1.0, 5.0, 3)
            placed.append((cube, trs((px, py, sz / 2), (0, 0, yaw), (sx / 2, sy / 2, sz / 2))))
            inst_tris += 12
        else:
            n = int(rng.integers(4, 13))
            rad = rng.uniform(3.0, 15.0)
            h = rng.uniform(3.0, 60.0)
            placed.append((prisms[n], trs((px, py, 0.0), (0, 0, yaw), (rad, rad, h))))
            inst_tris += 4 * n - 4
    cells = max(1, int(round(math.sqrt(max(target_tris - inst_tris, 2) / 2))))
    gv, gi = grid_mesh(400.0, 400.0, cells, cells)
    ground = spec.add_mesh(gv, gi)
    spec.add_instance(ground, identity(), entity_id=1)
    for k, (mesh, tf) in enumerate(placed):
        spec.add_instance(mesh, tf, entity_id=2 + k)
    return spec


def config2_lidar() -> Lidar:
    """128 channels over [-25, +15] deg x 2048 azimuths = 262,144 rays at (0, 0, 1.8), range 200 m; firing time
    offsets 100 ms * j / 2048 per azimuth column j."""
    ch, az = 128, 2048
    offsets = np.tile(np.arange(az, dtype=np.float64) * (100.0 / az), ch).astype(np.float32)
    return Lidar(spinning_lidar(ch, az, -25.0, 15.0), f32(translation(0, 0, 1.8)), (0.0, 200.0), offsets, ch, az)


@dataclass
class BatchSensor:
    scene: int     # index into the workload's scene list
    lidar: Lidar


def config4_workload(seed: int = 0xC0F4, sensors: int = 64, scenes: int = 8):
    """Batched multi-sensor step: `sensors` lidars of uneven size (channels U{16..128}, azimuths U{900..4096}, so
    14.4k-524k rays each; elevations [-25, +15] deg, range 200 m) placed at random positions and yaws over `scenes`
    urban scenes of (stratified) log-uniform 50k-5M triangles. Returns (scene specs, sensors)."""
    rng = np.random.default_rng(seed)
    # Stratified log-uniform sizes: one draw per equal slice of [log 50k, log 5M], so the range is covered.
    lo, hi = math.log(5e4), math.log(5e6)
    targets = np.exp(lo + (np.arange(scenes) + rng.uniform(0.0, 1.0, scenes)) * (hi - lo) / scenes)
    specs = [config2_scene(seed=seed + 1 + k, target_tris=int(t), instances=int(np.clip(t / 4000, 24, 500)))
             for k, t in enumerate(targets)]
    out = []
    for _ in range(sensors):
        ch = int(rng.integers(16, 129))
        az = int(rng.integers(900, 4097))
        sc = int(rng.integers(0, scenes))
        px, py = rng.uniform(-60.0, 60.0, 2)
        yaw = rng.uniform(0.0, 360.0)
        pose = f32(trs((px, py, 1.8), (0, 0, yaw)))
        out.append(BatchSensor(sc, Lidar(spinning_lidar(ch, az, -25.0, 15.0), pose, (0.0, 200.0), None, ch, az)))
    return specs, out


def radar_rays(azimuths: int = 256, elevations: int = 64, az_half_deg: float = 60.0, el_half_deg: float = 10.0):
    """Sensor-frame radar rays over [-az_half, +az_half] x [-el_half, +el_half] degrees (z up, azimuth from +X), with
    azimuths offset by 0.37 of a step so that no beam lies exactly on a 1-degree cell edge."""
    az = -az_half_deg + (np.arange(azimuths) + 0.37) * (2 * az_half_deg / azimuths)
    el = np.linspace(-el_half_deg, el_half_deg, elevations)
    A, E = np.meshgrid(np.deg2rad(az), np.deg2rad(el), indexing="xy")
    A, E = A.ravel(), E.ravel()
    th = np.pi / 2 - E
    ca, sa, ct, st = np.cos(A), np.sin(A), np.cos(th), np.sin(th)
    m = np.zeros((A.size, 3, 4))
    m[:, 0, 0], m[:, 0, 1], m[:, 0, 2] = ca * ct, -sa, ca * st
    m[:, 1, 0], m[:, 1, 1], m[:, 1, 2] = sa * ct, ca, sa * st
    m[:, 2, 0], m[:, 2, 1], m[:, 2, 2] = -st, 0.0, ct
    return f32(m)


# Radar cells of the serving workload: 0.2 m range x 1 degree azimuth over [0, 200) m x [-60, 60) degrees; the
# solid angle of one radar ray is the azimuth step x elevation step of radar_rays().
RADAR_CELL = dict(range_step=0.2, range_max=200.0, azimuth_step=float(np.deg2rad(1.0)),
                  azimuth_min=float(np.deg2rad(-60.0)), azimuth_max=float(np.deg2rad(60.0)))
RADAR_SOLID_ANGLE = float(np.deg2rad(120.0 / 256) * np.deg2rad(20.0 / 63))


@dataclass
class Client:
    scene: int
    kind: str      # "lidar" or "radar"
    lidar: Lidar   # sensor rays and pose (for radars: 256 x 64 radar rays)


def config5_workload(seed: int = 0xC0F5, clients: int = 32, radars: int = 8, scenes: int = 16):
    """Serving loop: `clients` simulated clients (the last `radars` of them radars of 256 x 64 rays over +-60 deg
    azimuth and +-10 deg elevation, range 200 m; the others lidars drawn as in config4_workload) over `scenes` urban
    scenes of (stratified) log-uniform 50k-10M triangles. Returns (scene specs, clients)."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(5e4), math.log(1e7)
    targets = np.exp(lo + (np.arange(scenes) + rng.uniform(0.0, 1.0, scenes)) * (hi - lo) / scenes)
    specs = [config2_scene(seed=seed + 1 + k, target_tris=int(t), instances=int(np.clip(t / 4000, 24, 500)))
             for k, t in enumerate(targets)]
    out = []
    for c in range(clients):
        sc = int(rng.integers(0, scenes))
        px, py = rng.uniform(-60.0, 60.0, 2)
        pose = f32(trs((px, py, 1.8), (0, 0, rng.uniform(0.0, 360.0))))
        if c >= clients - radars:
            out.append(Client(sc, "radar", Lidar(radar_rays(), pose, (0.0, 200.0), None, 64, 256)))
        else:
            ch = int(rng.integers(16, 129))
            az = int(rng.integers(900, 4097))
            out.append(Client(sc, "lidar", Lidar(spinning_lidar(ch, az, -25.0, 15.0), pose, (0.0, 200.0), None, ch, az)))
    return specs, out


def config3_scene(seed: int = 0xC0F3, target_tris: int = 10_000_000, instances: int = 2000):
    """Multi-return dynamic scene: `instances` instanced icospheres, textured boxes and building prisms (as in
    config2_scene) on a textured ground grid sized so the total is about `target_tris` triangles. The ground and
    the boxes sample a seeded 1024 x 1024 uint8 intensity texture through their UVs (the ground tiles it 16 x 16).
    Returns (spec, indices of the movable instances = every instance but the ground)."""
    rng = np.random.default_rng(seed)
    spec = SceneSpec()
    spec.textures.append(rng.integers(0, 256, (1024, 1024)).astype(np.uint8))
    sv, si = icosphere(3)
    sphere = spec.add_mesh(sv, si)
    cube = spec.add_mesh(CUBE_VERTICES, CUBE_INDICES, CUBE_UVS)
    prisms = {n: spec.add_mesh(*prism(n)) for n in range(4, 13)}
    kinds = rng.integers(0, 3, instances)
    placed, inst_tris = [], 0
    for k in range(instances):
        yaw = rng.uniform(0.0, 360.0)
        px, py = rng.uniform(-190.0, 190.0, 2)
        if kinds[k] == 0:
            r = rng.uniform(0.5, 3.0)
            placed.append((sphere, trs((px, py, r), (0, 0, yaw), (r, r, r)), -1))
            inst_tris += len(si)
        elif kinds[k] == 1:
            sx, sy, sz = rng.uniform(1.0, 5.0, 3)
            placed.append((cube, trs((px, py, sz / 2), (0, 0, yaw), (sx / 2, sy / 2, sz / 2)), 0))
            inst_tris += 12
        else:
            n = int(rng.integers(4, 13))
            rad = rng.uniform(3.0, 15.0)
            h = rng.uniform(3.0, 60.0)
            placed.append((prisms[n], trs((px, py, 0.0), (0, 0, yaw), (rad, rad, h)), -1))
            inst_tris += 4 * n - 4
    cells = max(1, int(round(math.sqrt(max(target_tris - inst_tris, 2) / 2))))
    gv, gi = grid_mesh(400.0, 400.0, cells, cells)
    uv = (gv[:, :2] / 400.0 + 0.5) * 16.0
    ground = spec.add_mesh(gv, gi, uv)
    spec.add_instance(ground, identity(), entity_id=1, texture=0)
    movable = []
    for k, (mesh, tf, tex) in enumerate(placed):
        movable.append(spec.add_instance(mesh, tf, entity_id=2 + k, texture=tex, laser_retro=float(k % 7)))
    return spec, movable


def config3_lidar() -> Lidar:
    """64 channels over [-25, +15] deg x 4096 azimuths = 262,144 beams at (0, 0, 1.8), range 200 m."""
    return Lidar(spinning_lidar(64, 4096, -25.0, 15.0), f32(translation(0, 0, 1.8)), (0.0, 200.0), None, 64, 4096)
