"""Seeded workloads whose oracle outputs are committed as golden fixtures (written by make_golden.py).

Every case is regenerated from its generator and seed, so a fixture holds only outputs: per-return hit mask and
entity ids for every ray, the oracle's per-beam ambiguity bits, explicit distances / points / normals / intensities
for a seeded sample of rays, and SHA-256 digests of the complete output arrays.
"""
from __future__ import annotations

import numpy as np

from robotecgpulidar_amd import rgl
from robotecgpulidar_amd import scenes as S
from tests.helpers import Case

XYZ, INTEN, IS_HIT, ENT, DIST, RET, NORMAL, INC = 1, 2, 4, 7, 8, 12, 22, 23
GEOMETRY_FIELDS = [XYZ, DIST, ENT, NORMAL, IS_HIT]
MULTI_FIELDS = [DIST, ENT, IS_HIT, INTEN, RET]


def textured_scene(seed=3) -> S.SceneSpec:
    """C1 scene plus 8 textured, retro-reflective boxes around the sensor (64 x 64 uint8 texture)."""
    rng = np.random.default_rng(seed)
    spec = S.config1_scene()
    cube_uv = spec.add_mesh(S.CUBE_VERTICES, S.CUBE_INDICES, S.CUBE_UVS)
    spec.textures.append(rng.integers(0, 256, (64, 64)).astype(np.uint8))
    for k in range(8):
        a = k * np.pi / 4
        spec.add_instance(cube_uv, S.trs((12 * np.cos(a), 12 * np.sin(a), 2.0), (0, 0, 20 * k), (1.5, 1.5, 2.0)),
                          entity_id=100 + k, texture=0, laser_retro=0.5 + k)
    return spec


def c1_case():
    """C1: 200 x 200 m ground + 64 boxes, 32 x 1800 lidar at (0, 0, 1.8), range [0.1, 100] m."""
    spec, lidar = S.config1_scene(), S.config1_lidar()
    return Case(spec, lidar.rays, pose=lidar.pose, ranges=np.array([lidar.range], np.float32)), GEOMETRY_FIELDS


def urban_case():
    """C2 geometry at reduced size: the 2.0M-triangle urban scene, every 64th ray of the 128 x 2048 lidar (4096
    rays), range 200 m, velocity distortion at 10 m/s with the lidar's firing time offsets."""
    spec, lidar = S.config2_scene(), S.config2_lidar()
    sel = np.arange(0, len(lidar.rays), 64)
    case = Case(spec, lidar.rays[sel], pose=lidar.pose, ranges=np.array([[0.0, 200.0]], np.float32),
                time_offsets=lidar.time_offsets[sel], distortion=True, velocity=((10.0, 0.0, 0.0), (0.0, 0.0, 0.0)))
    return case, GEOMETRY_FIELDS


def multireturn_case(mode: str):
    """C3 semantics at reduced size: textured scene, 8 x 128 lidar with 2 mrad divergence (58 samples per beam),
    return mode `mode`, default intensity 30."""
    rays = S.spinning_lidar(8, 128, -10.0, 5.0)
    case = Case(textured_scene(), rays, pose=S.f32(S.translation(0, 0, 1.8)), ranges=np.array([[0.1, 100.0]], np.float32),
                return_mode=getattr(rgl, "RGL_RETURN_" + mode), divergence=(2e-3, 2e-3), default_intensity=30.0)
    return case, MULTI_FIELDS


def c3_case(mode: str):
    """C3 at full scene size, reduced beam set: the ~10M-triangle textured scene of 2,001 instances, every 256th beam
    of the 64 x 4096 lidar (1,024 beams) with 0.13 mrad divergence (58 samples per beam), return mode `mode`,
    default intensity 30, range 200 m."""
    spec, _ = S.config3_scene()
    lidar = S.config3_lidar()
    sel = np.arange(0, len(lidar.rays), 256)
    case = Case(spec, lidar.rays[sel], pose=lidar.pose, ranges=np.array([lidar.range], np.float32),
                return_mode=getattr(rgl, "RGL_RETURN_" + mode), divergence=(0.13e-3, 0.13e-3), default_intensity=30.0)
    return case, MULTI_FIELDS


def _c3(mode: str):
    def make():
        return c3_case(mode)
    make.__doc__ = c3_case.__doc__.replace("return mode `mode`", f"return mode {mode}")
    return make


def _multireturn(mode: str):
    def make():
        return multireturn_case(mode)
    make.__doc__ = multireturn_case.__doc__.replace("return mode `mode`", f"return mode {mode}")
    return make


CASES = {
    "c1": c1_case,
    "urban": urban_case,
    "multireturn_first_last": _multireturn("FIRST_LAST"),
    "multireturn_strongest": _multireturn("STRONGEST_SECOND_STRONGEST"),
    "c3_first_last": _c3("FIRST_LAST"),
    "c3_strongest": _c3("STRONGEST"),
}

# C4 sample: for every sensor of scenes.config4_workload(), C4_BEAMS seeded beam indices (C4_SEED + sensor index);
# geometry fields of the raytrace node for those beams (explicit XYZ and normal for the first C4_EXPLICIT of them).
C4_BEAMS = 1024
C4_EXPLICIT = 256
C4_SEED = 0xC4C4


def c4_beams(sensor: int, rays: int) -> np.ndarray:
    rng = np.random.default_rng(C4_SEED + sensor)
    return np.sort(rng.choice(rays, min(C4_BEAMS, rays), replace=False)).astype(np.int64)

# (seed, run, tag) streams whose first 64 normal variates are committed.
NOISE_STREAMS = [(0, 0, 1), (0xC0F2D, 3, 3), (0xDEADBEEFCAFEF00D, 123456, 2)]
SAMPLE_SIZE = 1024
SAMPLE_SEED = 0x601D
