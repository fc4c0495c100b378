"""Python access to the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module, as the checker (or the
timed CPU baseline). It wraps oracle/liboracle.so (oracle.cpp, FP64 restatement of the reference's per-ray
semantics; cpu_port.cpp, the FP32 multithreaded CPU port of the lidar graph timed as the CPU baseline) and restates
the point-cloud stages in numpy:
  compaction     src/graph/CompactByFieldPointsNode.cpp:33-105 (stable: keep elements whose field != 0)
  point transform src/gpu/nodeKernels.cu:51-56
  noise          src/gpu/gaussianNoiseKernels.cu:20-66 with the library's seeded stream (orc_normals)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)


class OrcScene(C.Structure):
    _fields_ = [
        ("nMeshes", C.c_int), ("meshVertOffset", i32p), ("meshTriOffset", i32p), ("vertices", f32p),
        ("indices", i32p), ("uvs", f32p), ("meshHasUV", i32p), ("nInstances", C.c_int), ("instMesh", i32p),
        ("instTransform", f32p), ("instPrevTransform", f32p), ("instHasPrev", i32p), ("instId", i32p),
        ("instRetro", f32p), ("instTexture", i32p), ("nTextures", C.c_int), ("texWidth", i32p), ("texHeight", i32p),
        ("texOffset", C.POINTER(C.c_int64)), ("texels", C.POINTER(C.c_uint8)),
        ("instDispOffset", C.POINTER(C.c_int64)), ("displacements", f32p),
    ]


class OrcRequest(C.Structure):
    _fields_ = [
        ("nRays", C.c_int), ("rays", f32p), ("rayOriginToWorld", C.c_float * 12), ("ranges", f32p), ("nRanges", C.c_int),
        ("ringIds", i32p), ("nRingIds", C.c_int), ("timeOffsetsMs", f32p), ("mask", C.POINTER(C.c_int8)),
        ("distortion", C.c_int), ("linVel", C.c_float * 3), ("angVel", C.c_float * 3), ("nearDistance", C.c_float),
        ("farDistance", C.c_float), ("defaultIntensity", C.c_float), ("sceneDeltaTime", C.c_float),
        ("hDivergence", C.c_float), ("vDivergence", C.c_float), ("returnMode", C.c_int32), ("bruteForce", C.c_int),
        ("threads", C.c_int), ("ambiguityEps", C.c_double), ("raysF64", C.POINTER(C.c_double)),
    ]


# (field name in OrcOutputs, rgl field id or None, dtype, components)
OUT_FIELDS = [
    ("xyz", 1, np.float32, 3), ("isHit", 4, np.int32, 1), ("rayIdx", 6, np.uint32, 1), ("ringId", 11, np.uint16, 1),
    ("returnType", 12, np.uint8, 1), ("distance", 8, np.float32, 1), ("intensityF32", 2, np.float32, 1),
    ("intensityU8", 3, np.uint8, 1), ("laserRetro", 25, np.float32, 1), ("timeStampF64", 13, np.float64, 1),
    ("timeStampU32", 14, np.uint32, 1), ("entityId", 7, np.int32, 1), ("absVelocity", 15, np.float32, 3),
    ("relVelocity", 16, np.float32, 3), ("radialSpeed", 17, np.float32, 1), ("azimuth", 9, np.float32, 1),
    ("elevation", 10, np.float32, 1), ("normal", 22, np.float32, 3), ("incidentAngle", 23, np.float32, 1),
]
_CTYPE = {np.float32: C.c_float, np.int32: C.c_int32, np.uint32: C.c_uint32, np.uint16: C.c_uint16,
          np.uint8: C.c_uint8, np.float64: C.c_double, np.int8: C.c_int8}


class OrcOutputs(C.Structure):
    _fields_ = [(n, C.POINTER(_CTYPE[dt])) for n, _, dt, _ in OUT_FIELDS] + [
        ("ambiguous", C.POINTER(C.c_int8)), ("primaryT", C.POINTER(C.c_double)), ("primaryInst", i32p),
        ("primaryPrim", i32p), ("secondRange", f32p),
    ]


class CpuFrame(C.Structure):
    _fields_ = [("nRays", C.c_int), ("rays", f32p), ("pose", C.c_float * 12), ("range", C.c_float * 2),
                ("toSensor", C.c_float * 12), ("lookAt", C.c_float * 12), ("lookAtInv", C.c_float * 12),
                ("noiseMean", C.c_float), ("noiseBase", C.c_float), ("noiseRise", C.c_float), ("seed", C.c_uint64),
                ("run", C.c_uint32), ("tag", C.c_uint32)]


class CpuOut(C.Structure):
    _fields_ = [("fullXyz", f32p), ("fullDistance", f32p), ("fullEntityId", i32p), ("fullNormal", f32p),
                ("fullIsHit", i32p), ("xyz", f32p), ("distance", f32p), ("entityId", i32p), ("normal", f32p),
                ("kept", C.c_int32)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        lib.orc_raytrace.argtypes = [C.POINTER(OrcScene), C.POINTER(OrcRequest), C.POINTER(OrcOutputs)]
        lib.orc_raytrace.restype = C.c_int
        lib.orc_world_create.argtypes = [C.POINTER(OrcScene), C.c_int]
        lib.orc_world_create.restype = C.c_void_p
        lib.orc_world_raytrace.argtypes = [C.c_void_p, C.POINTER(OrcRequest), C.POINTER(OrcOutputs)]
        lib.orc_world_raytrace.restype = C.c_int
        lib.orc_world_destroy.argtypes = [C.c_void_p]
        lib.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        lib.orc_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, f32p, C.c_int]
        lib.orc_logf.argtypes = [C.c_float]
        lib.orc_logf.restype = C.c_float
        lib.orc_sincos2pi.argtypes = [C.c_float, f32p, f32p]
        lib.orc_plant_mistake.argtypes = [C.c_int]
        lib.cpu_scene_create.argtypes = [C.POINTER(OrcScene)]
        lib.cpu_scene_create.restype = C.c_void_p
        lib.cpu_scene_destroy.argtypes = [C.c_void_p]
        lib.cpu_run.argtypes = [C.POINTER(C.c_void_p), C.POINTER(CpuFrame), C.POINTER(CpuOut), C.c_int, C.c_int]
        lib.cpu_run.restype = C.c_int
        _lib = lib
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


class _Keep:
    """Keeps numpy arrays alive while ctypes structures point into them."""

    def __init__(self):
        self.items = []

    def __call__(self, a):
        self.items.append(a)
        return a


def make_scene(spec, keep: _Keep, prev_transforms=None, displacements=None) -> OrcScene:
    """`displacements`: {instance index: (vertices, 3) object-space vertex displacement since the previous frame}."""
    vo, to = [0], [0]
    verts, idx, uvs, has_uv = [], [], [], []
    for m in spec.meshes:
        verts.append(m.vertices.reshape(-1, 3))
        idx.append(m.indices.reshape(-1, 3))
        uvs.append(m.uvs if m.uvs is not None else np.zeros((len(m.vertices), 2), np.float32))
        has_uv.append(1 if m.uvs is not None else 0)
        vo.append(vo[-1] + len(m.vertices))
        to.append(to[-1] + len(m.indices))
    verts = verts or [np.zeros((1, 3), np.float32)]
    idx = idx or [np.zeros((1, 3), np.int32)]
    uvs = uvs or [np.zeros((1, 2), np.float32)]
    V = keep(np.ascontiguousarray(np.concatenate(verts), dtype=np.float32))
    I = keep(np.ascontiguousarray(np.concatenate(idx), dtype=np.int32))
    UV = keep(np.ascontiguousarray(np.concatenate(uvs), dtype=np.float32))
    s = OrcScene()
    s.nMeshes = len(spec.meshes)
    s.meshVertOffset = _p(keep(np.array(vo, np.int32)), C.c_int32)
    s.meshTriOffset = _p(keep(np.array(to, np.int32)), C.c_int32)
    s.vertices = _p(V, C.c_float)
    s.indices = _p(I, C.c_int32)
    s.uvs = _p(UV, C.c_float)
    s.meshHasUV = _p(keep(np.array(has_uv, np.int32)), C.c_int32)
    n = len(spec.instances)
    s.nInstances = n
    s.instMesh = _p(keep(np.array([i.mesh for i in spec.instances], np.int32)), C.c_int32)
    tf = keep(np.ascontiguousarray(np.stack([i.transform for i in spec.instances] or [np.zeros((3, 4))])
                                   .astype(np.float32)))
    s.instTransform = _p(tf, C.c_float)
    if prev_transforms is not None:
        pv = keep(np.ascontiguousarray(np.asarray(prev_transforms, np.float32).reshape(n, 3, 4)))
        s.instPrevTransform = _p(pv, C.c_float)
        s.instHasPrev = _p(keep(np.ones(n, np.int32)), C.c_int32)
    if displacements:
        offs = np.full(n, -1, np.int64)
        chunks, at = [], 0
        for i, dv in sorted(displacements.items()):
            dv = np.ascontiguousarray(dv, np.float32).reshape(-1, 3)
            offs[i] = at
            chunks.append(dv.ravel())
            at += dv.size
        s.instDispOffset = _p(keep(offs), C.c_int64)
        s.displacements = _p(keep(np.concatenate(chunks)), C.c_float)
    s.instId = _p(keep(np.array([268435455 if i.entity_id is None else i.entity_id for i in spec.instances],
                                np.int32)), C.c_int32)
    s.instRetro = _p(keep(np.array([i.laser_retro for i in spec.instances], np.float32)), C.c_float)
    s.instTexture = _p(keep(np.array([i.texture for i in spec.instances], np.int32)), C.c_int32)
    s.nTextures = len(spec.textures)
    if spec.textures:
        offs = np.cumsum([0] + [t.size for t in spec.textures[:-1]]).astype(np.int64)
        s.texWidth = _p(keep(np.array([t.shape[1] for t in spec.textures], np.int32)), C.c_int32)
        s.texHeight = _p(keep(np.array([t.shape[0] for t in spec.textures], np.int32)), C.c_int32)
        s.texOffset = _p(keep(offs), C.c_int64)
        s.texels = _p(keep(np.ascontiguousarray(np.concatenate([t.ravel() for t in spec.textures]), np.uint8)),
                      C.c_uint8)
    return s


def raytrace(spec, rays_world, ray_origin_to_world=None, ranges=((0.0, 3.4028234663852886e38),), ring_ids=None,
             time_offsets=None, mask=None, distortion=False, lin_vel=(0, 0, 0), ang_vel=(0, 0, 0),
             near=float("inf"), far=float("inf"), default_intensity=0.0, scene_dt=0.0, divergence=(0.0, 0.0),
             return_mode=(1 << 24) | 4, brute_force=False, threads=None, eps=1e-5, prev_transforms=None,
             fields=None, world: "World | None" = None, displacements=None) -> dict:
    """Runs the raytrace node semantics on the CPU. rays_world: (N, 3, 4) rays after the rays chain; a float64
    array is traced as given (FP64 world rays, direction = third column), anything else as float32 rays with the
    reference's float rounding of the direction. Returns a dict of per-return-point arrays keyed by rgl field id,
    plus 'ambiguous', 'primaryT', 'primaryInst', 'primaryPrim' and 'second_range'."""
    lib = load()
    keep = _Keep()
    sc = make_scene(spec, keep, prev_transforms, displacements) if world is None else None
    q = OrcRequest()
    if isinstance(rays_world, np.ndarray) and rays_world.dtype == np.float64:
        rays64 = keep(np.ascontiguousarray(rays_world.reshape(-1, 3, 4)))
        q.raysF64 = _p(rays64, C.c_double)
        rays = keep(rays64.astype(np.float32))
    else:
        rays = keep(np.ascontiguousarray(np.asarray(rays_world, np.float32).reshape(-1, 3, 4)))
    n = len(rays)
    q.nRays = n
    q.rays = _p(rays, C.c_float)
    tw = np.asarray(ray_origin_to_world if ray_origin_to_world is not None else np.eye(3, 4), np.float32).ravel()
    for k in range(12):
        q.rayOriginToWorld[k] = float(tw[k])
    rg = keep(np.ascontiguousarray(np.asarray(ranges, np.float32).reshape(-1, 2)))
    q.ranges = _p(rg, C.c_float)
    q.nRanges = len(rg)
    if ring_ids is not None:
        ri = keep(np.ascontiguousarray(np.asarray(ring_ids, np.int32)))
        q.ringIds = _p(ri, C.c_int32)
        q.nRingIds = len(ri)
    if time_offsets is not None:
        q.timeOffsetsMs = _p(keep(np.ascontiguousarray(np.asarray(time_offsets, np.float32))), C.c_float)
    if mask is not None:
        q.mask = _p(keep(np.ascontiguousarray(np.asarray(mask, np.int8))), C.c_int8)
    q.distortion = 1 if distortion else 0
    for k in range(3):
        q.linVel[k] = float(lin_vel[k])
        q.angVel[k] = float(ang_vel[k])
    q.nearDistance, q.farDistance = near, far
    q.defaultIntensity = default_intensity
    q.sceneDeltaTime = scene_dt
    q.hDivergence, q.vDivergence = divergence
    q.returnMode = return_mode
    q.bruteForce = 1 if brute_force else 0
    q.threads = threads or min(8, os.cpu_count() or 1)
    q.ambiguityEps = eps
    R = (return_mode >> 24) & 0xff
    out = OrcOutputs()
    res = {}
    for name, fid, dt, k in OUT_FIELDS:
        if fields is not None and fid not in fields:
            continue
        a = np.empty((n * R, k) if k > 1 else (n * R,), dtype=dt)
        res[fid] = a
        setattr(out, name, _p(a, _CTYPE[dt]))
    for name, dt in (("ambiguous", np.int8), ("primaryT", np.float64), ("primaryInst", np.int32),
                     ("primaryPrim", np.int32)):
        a = np.empty(n, dtype=dt)
        res[name] = a
        setattr(out, name, _p(a, _CTYPE[dt]))
    # per beam: [min, max] SECOND distance and [min, max] SECOND_STRONGEST intensity over every resolution of the
    # reduction's near-tied comparisons (+-inf where unbounded)
    res["second_range"] = np.empty((n, 4), np.float32)
    out.secondRange = _p(res["second_range"], C.c_float)
    if world is not None:
        status = lib.orc_world_raytrace(world.handle, C.byref(q), C.byref(out))
    else:
        status = lib.orc_raytrace(C.byref(sc), C.byref(q), C.byref(out))
    if status != 0:
        raise RuntimeError("orc_raytrace failed")
    return res


class World:
    """Oracle scene with its BVHs built once (for timing the CPU baseline without the build)."""

    def __init__(self, spec, brute_force=False, prev_transforms=None):
        lib = load()
        self.keep = _Keep()
        self.scene = make_scene(spec, self.keep, prev_transforms)
        self.handle = lib.orc_world_create(C.byref(self.scene), 1 if brute_force else 0)

    def close(self):
        if self.handle:
            load().orc_world_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


# Deliberate mistake planted in the oracle (see orc_plant_mistake in oracle.h for 1-6); 7 and 8 live in the numpy
# stages below: 7 = distance noise moves points along their world position instead of their look-at position,
# 8 = angular noise rotates by the negated angle. Only tests set it, and reset it to 0.
MISTAKE = 0


def plant_mistake(which: int) -> None:
    global MISTAKE
    MISTAKE = which
    load().orc_plant_mistake(which if which <= 6 else 0)


class CpuPort:
    """The FP32 CPU port of the lidar graph rays -> pose -> range -> raytrace -> points transform -> distance noise
    -> compact over `specs` (one CPU scene each, LBVHs built once)."""

    def __init__(self, specs):
        lib = load()
        self.keep = _Keep()
        self.scenes = [lib.cpu_scene_create(C.byref(make_scene(sp, self.keep))) for sp in specs]

    def prepare(self, sensors) -> dict:
        """sensors: list of dicts(scene, rays, pose, range, to_sensor, look_at, look_at_inv, noise=(mean, base,
        rise), seed, run, tag). Allocates the frame's descriptors and output arrays once (execute() reuses them)."""
        n = len(sensors)
        frames = (CpuFrame * n)()
        outs = (CpuOut * n)()
        keep = _Keep()
        res = []
        for k, sn in enumerate(sensors):
            rays = keep(np.ascontiguousarray(np.asarray(sn["rays"], np.float32).reshape(-1, 3, 4)))
            m = len(rays)
            f = frames[k]
            f.nRays = m
            f.rays = _p(rays, C.c_float)
            for name, dst in (("pose", f.pose), ("to_sensor", f.toSensor), ("look_at", f.lookAt),
                              ("look_at_inv", f.lookAtInv)):
                arr = np.asarray(sn[name], np.float32).ravel()
                for i in range(12):
                    dst[i] = float(arr[i])
            f.range[0], f.range[1] = sn["range"]
            f.noiseMean, f.noiseBase, f.noiseRise = sn["noise"]
            f.seed, f.run, f.tag = sn["seed"], sn["run"], sn["tag"]
            r = {"full_xyz": np.empty((m, 3), np.float32), "full_distance": np.empty(m, np.float32),
                 "full_entity": np.empty(m, np.int32), "full_normal": np.empty((m, 3), np.float32),
                 "full_is_hit": np.empty(m, np.int32), "xyz": np.empty((m, 3), np.float32),
                 "distance": np.empty(m, np.float32), "entity": np.empty(m, np.int32),
                 "normal": np.empty((m, 3), np.float32)}
            for a in r.values():  # touch every page before any timing
                a.fill(0)
            o = outs[k]
            for field, key, ct in (("fullXyz", "full_xyz", C.c_float), ("fullDistance", "full_distance", C.c_float),
                                   ("fullEntityId", "full_entity", C.c_int32), ("fullNormal", "full_normal", C.c_float),
                                   ("fullIsHit", "full_is_hit", C.c_int32), ("xyz", "xyz", C.c_float),
                                   ("distance", "distance", C.c_float), ("entityId", "entity", C.c_int32),
                                   ("normal", "normal", C.c_float)):
                setattr(o, field, _p(r[key], ct))
            res.append(r)
        handles = (C.c_void_p * n)(*[self.scenes[sn["scene"]] for sn in sensors])
        return {"frames": frames, "outs": outs, "handles": handles, "keep": keep, "arrays": res, "count": n}

    def execute(self, frame: dict, threads: int) -> None:
        load().cpu_run(frame["handles"], frame["frames"], frame["outs"], frame["count"], threads)

    def results(self, frame: dict) -> list:
        """Per sensor: the kept (xyz, distance, entity, normal), `kept`, and the uncompacted arrays."""
        out = []
        for k, r in enumerate(frame["arrays"]):
            kept = frame["outs"][k].kept
            d = dict(r)
            d["kept"] = kept
            for key in ("xyz", "distance", "entity", "normal"):
                d[key] = r[key][:kept]
            out.append(d)
        return out

    def run(self, sensors, threads: int) -> list:
        frame = self.prepare(sensors)
        self.execute(frame, threads)
        return self.results(frame)

    def close(self):
        for h in self.scenes:
            if h:
                load().cpu_scene_destroy(h)
        self.scenes = []

    def __del__(self):
        self.close()


def philox(ctr, key) -> np.ndarray:
    lib = load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.orc_philox4x32_10(c, k, o)
    return np.array(list(o), dtype=np.uint32)


def normals(seed: int, run: int, tag: int, count: int) -> np.ndarray:
    out = np.empty(count, np.float32)
    load().orc_normals(C.c_uint64(seed), run, tag, out.ctypes.data_as(f32p), count)
    return out


def logf(x: float) -> float:
    return load().orc_logf(x)


def sincos2pi(u: float) -> tuple[float, float]:
    s, c = C.c_float(), C.c_float()
    load().orc_sincos2pi(u, C.byref(s), C.byref(c))
    return s.value, c.value


# ------------------------------------------------------------------------------------------ point stages (numpy)
def compact(flags: np.ndarray, *fields: np.ndarray):
    """Stable compaction: keep elements whose flag != 0, in input order."""
    keep = np.asarray(flags) != 0
    return tuple(np.asarray(f)[keep] for f in fields)


def _mat_apply(m, p):
    m = np.asarray(m, np.float64)
    return p @ m[:, :3].T + m[:, 3]


def _inverse(m):
    m = np.asarray(m, np.float64)
    r = np.linalg.inv(m[:, :3])
    return np.concatenate([r, -(r @ m[:, 3])[:, None]], axis=1)


def _rot_axis(axis: int, ang: np.ndarray):
    c, s = np.cos(ang), np.sin(ang)
    z, o = np.zeros_like(ang), np.ones_like(ang)
    if axis == 1:
        rows = [[o, z, z], [z, c, -s], [z, s, c]]
    elif axis == 2:
        rows = [[c, z, s], [z, o, z], [-s, z, c]]
    else:
        rows = [[c, -s, z], [s, c, z], [z, z, o]]
    return np.stack([np.stack(r, -1) for r in rows], -2)  # (n, 3, 3)


def noise_distance(xyz, distance, normals_, mean, base, rise, look_at):
    """src/gpu/gaussianNoiseKernels.cu:49-66 in FP64: sigma = base + d*rise; e = mean + n*sigma."""
    d = np.asarray(distance, np.float64)
    e = mean + np.asarray(normals_, np.float64) * (d * rise + base)
    if MISTAKE == 7:
        look_at = np.eye(3, 4)
    pr = _mat_apply(look_at, np.asarray(xyz, np.float64))
    with np.errstate(invalid="ignore", divide="ignore"):
        moved = pr + pr / np.linalg.norm(pr, axis=1, keepdims=True) * e[:, None]
    return _mat_apply(_inverse(look_at), moved), d + e


def noise_angular_hitpoint(xyz, normals_, mean, st_dev, axis, look_at):
    """src/gpu/gaussianNoiseKernels.cu:31-47 in FP64."""
    ang = (mean + np.asarray(normals_, np.float64) * st_dev) * (-1.0 if MISTAKE == 8 else 1.0)
    q = np.einsum("nij,nj->ni", _rot_axis(axis, ang), _mat_apply(look_at, np.asarray(xyz, np.float64)))
    return _mat_apply(_inverse(look_at), q), np.linalg.norm(q, axis=1)


def noise_angular_rays(rays, normals_, mean, st_dev, axis, look_at):
    """src/gpu/gaussianNoiseKernels.cu:20-29 in FP64: out = L^-1 * Rot(axis, e) * L * ray."""
    rays = np.asarray(rays, np.float64).reshape(-1, 3, 4)
    L = np.asarray(look_at, np.float64)
    Li = _inverse(L)
    ang = (mean + np.asarray(normals_, np.float64) * st_dev) * (-1.0 if MISTAKE == 8 else 1.0)
    R = _rot_axis(axis, ang)

    def mm(a, b):  # 3x4 affine products, batched
        r = a[..., :3] @ b
        r[..., 3] += a[..., 3]
        return r

    lr = mm(np.broadcast_to(L, rays.shape), rays)
    rr = np.concatenate([R, np.zeros((len(R), 3, 1))], axis=2)
    return mm(np.broadcast_to(Li, rays.shape), mm(rr, lr))


RCS_FIXED_SCALE = 2.0 ** 24


def radar_bin(xyz_sensor, distance, incident, is_hit, range_step, range_max, azimuth_step, azimuth_min, azimuth_max,
              ray_solid_angle, eps_range=1e-6, eps_azimuth=1e-4):
    """Range x azimuth binning with RCS (the build-defined radar binning stage), in FP64.

    A hit (is_hit != 0) at distance d with sensor-frame point p falls in cell (floor(d / range_step),
    floor((atan2(p.y, p.x) - azimuth_min) / azimuth_step)) when inside [0, range_max) x [azimuth_min, azimuth_max);
    its RCS contribution d^2 * ray_solid_angle * cos(incident) is rounded to 2^-24 m^2 and summed per cell.
    Returns (cells, counts, rcs_db, alt_cells): sorted non-empty cell indices (range-major), hit counts, 10 log10 of
    the summed contributions, and per input point the cells it could fall in under FP32 rounding other than its own
    (a list of up to three cell indices; empty when its range position is farther than eps_range (relative) and its
    azimuth position farther than eps_azimuth (in cells) from a cell edge)."""
    d = np.asarray(distance, np.float64)
    p = np.asarray(xyz_sensor, np.float64)
    inc = np.asarray(incident, np.float64)
    hit = np.asarray(is_hit) != 0
    # The stage receives its parameters as FP32 (rgl_node_points_radar_bin takes floats).
    range_step, range_max, azimuth_step, azimuth_min, azimuth_max, ray_solid_angle = (
        float(np.float32(v)) for v in (range_step, range_max, azimuth_step, azimuth_min, azimuth_max, ray_solid_angle))
    rc = int(np.ceil(range_max / range_step))
    ac = int(np.ceil((azimuth_max - azimuth_min) / azimuth_step))
    az = np.arctan2(p[:, 1], p[:, 0])
    d = np.where(hit, d, -1.0)  # non-hits carry infinite distances; they are dropped below
    rpos = d / range_step
    apos = (az - azimuth_min) / azimuth_step
    rf, af = np.floor(rpos), np.floor(apos)
    inside = hit & (rf >= 0) & (rf < rc) & (af >= 0) & (af < ac)
    r_near = hit & (np.abs(rpos - np.round(rpos)) <= eps_range * np.maximum(np.abs(rpos), 1.0))
    a_near = hit & (np.abs(apos - np.round(apos)) <= eps_azimuth)
    r_alt = np.where(rpos - rf < 0.5, rf - 1, rf + 1)
    a_alt = np.where(apos - af < 0.5, af - 1, af + 1)
    alt_cells = []
    for i in np.nonzero(r_near | a_near)[0]:
        rs = [rf[i]] + ([r_alt[i]] if r_near[i] else [])
        as_ = [af[i]] + ([a_alt[i]] if a_near[i] else [])
        alt_cells.append((i, [int(r * ac + a) for r in rs for a in as_ if (r, a) != (rf[i], af[i])]))
    cell = (rf * ac + af)[inside].astype(np.int64)
    sigma = d[inside] ** 2 * ray_solid_angle * np.cos(inc[inside])
    q = np.rint(sigma * RCS_FIXED_SCALE).astype(np.int64)
    cells, inv = np.unique(cell, return_inverse=True)
    counts = np.bincount(inv, minlength=len(cells)).astype(np.int64)
    sums = np.bincount(inv, weights=q.astype(np.float64), minlength=len(cells))
    with np.errstate(divide="ignore"):
        rcs_db = 10.0 * np.log10(sums / RCS_FIXED_SCALE)
    alt = [[] for _ in range(len(d))]
    for i, cs in alt_cells:
        alt[i] = cs
    return cells, counts, rcs_db, alt
