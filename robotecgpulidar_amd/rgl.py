"""ctypes binding of the RGL C API exported by libRobotecGPULidar.so (include/rgl/api/core.h).

Names, argument order and status codes mirror the reference API (RobotecGPULidar 0.19.0,
include/rgl/api/core.h:526-1198), so Python code reads like the reference's C++ tests. The library is the
HIP build in this package directory; loading fails loudly if it has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libRobotecGPULidar.so"

# ------------------------------------------------------------------------------------------ enums
RGL_SUCCESS = 0
RGL_INVALID_ARGUMENT = 1
RGL_INVALID_STATE = 2
RGL_LOGGING_ERROR = 3
RGL_INVALID_API_OBJECT = 4
RGL_INVALID_FILE_PATH = 5
RGL_TAPE_ERROR = 6
RGL_UDP_ERROR = 7
RGL_ROS2_ERROR = 8
RGL_INVALID_PIPELINE = 9
RGL_INITIALIZATION_ERROR = 10
RGL_NOT_IMPLEMENTED = 404
RGL_INTERNAL_EXCEPTION = 500

RGL_ENTITY_INVALID_ID = 0
RGL_DEFAULT_ENTITY_ID = 268435455

XYZ_VEC3_F32 = 1
INTENSITY_F32 = 2
INTENSITY_U8 = 3
IS_HIT_I32 = 4
IS_GROUND_I32 = 5
RAY_IDX_U32 = 6
ENTITY_ID_I32 = 7
DISTANCE_F32 = 8
AZIMUTH_F32 = 9
ELEVATION_F32 = 10
RING_ID_U16 = 11
RETURN_TYPE_U8 = 12
TIME_STAMP_F64 = 13
TIME_STAMP_U32 = 14
ABSOLUTE_VELOCITY_VEC3_F32 = 15
RELATIVE_VELOCITY_VEC3_F32 = 16
RADIAL_SPEED_F32 = 17
POWER_F32 = 18
RCS_F32 = 19
NOISE_F32 = 20
SNR_F32 = 21
NORMAL_VEC3_F32 = 22
INCIDENT_ANGLE_F32 = 23
RAY_POSE_MAT3x4_F32 = 24
LASER_RETRO_F32 = 25
PADDING_8 = 1024
PADDING_16 = 1025
PADDING_32 = 1026
DYNAMIC_FORMAT = 13842

# numpy dtype and per-point shape of each field (src/RGLFields.hpp:108-191)
FIELD_DTYPE = {
    XYZ_VEC3_F32: (np.float32, 3), INTENSITY_F32: (np.float32, 1), INTENSITY_U8: (np.uint8, 1),
    IS_HIT_I32: (np.int32, 1), IS_GROUND_I32: (np.int32, 1), RAY_IDX_U32: (np.uint32, 1),
    ENTITY_ID_I32: (np.int32, 1), DISTANCE_F32: (np.float32, 1), AZIMUTH_F32: (np.float32, 1),
    ELEVATION_F32: (np.float32, 1), RING_ID_U16: (np.uint16, 1), RETURN_TYPE_U8: (np.uint8, 1),
    TIME_STAMP_F64: (np.float64, 1), TIME_STAMP_U32: (np.uint32, 1),
    ABSOLUTE_VELOCITY_VEC3_F32: (np.float32, 3), RELATIVE_VELOCITY_VEC3_F32: (np.float32, 3),
    RADIAL_SPEED_F32: (np.float32, 1), POWER_F32: (np.float32, 1), RCS_F32: (np.float32, 1),
    NOISE_F32: (np.float32, 1), SNR_F32: (np.float32, 1), NORMAL_VEC3_F32: (np.float32, 3),
    INCIDENT_ANGLE_F32: (np.float32, 1), RAY_POSE_MAT3x4_F32: (np.float32, 12), LASER_RETRO_F32: (np.float32, 1),
    PADDING_8: (np.uint8, 1), PADDING_16: (np.uint16, 1), PADDING_32: (np.uint32, 1),
}


def field_size(f: int) -> int:
    dt, n = FIELD_DTYPE[f]
    return np.dtype(dt).itemsize * n


RETURN_MODE_SINGLE = 1 << 24
RETURN_MODE_DUAL = 2 << 24
RETURN_TYPE_UNKNOWN, RETURN_TYPE_STRONGEST, RETURN_TYPE_LAST, RETURN_TYPE_SECOND, RETURN_TYPE_FIRST, \
    RETURN_TYPE_SECOND_STRONGEST = range(6)
RGL_RETURN_FIRST = RETURN_MODE_SINGLE | RETURN_TYPE_FIRST
RGL_RETURN_SECOND = RETURN_MODE_SINGLE | RETURN_TYPE_SECOND
RGL_RETURN_LAST = RETURN_MODE_SINGLE | RETURN_TYPE_LAST
RGL_RETURN_STRONGEST = RETURN_MODE_SINGLE | RETURN_TYPE_STRONGEST
RGL_RETURN_LAST_STRONGEST = RETURN_MODE_DUAL | RETURN_TYPE_LAST | (RETURN_TYPE_STRONGEST << 8)
RGL_RETURN_FIRST_LAST = RETURN_MODE_DUAL | RETURN_TYPE_FIRST | (RETURN_TYPE_LAST << 8)
RGL_RETURN_FIRST_STRONGEST = RETURN_MODE_DUAL | RETURN_TYPE_FIRST | (RETURN_TYPE_STRONGEST << 8)
RGL_RETURN_STRONGEST_SECOND_STRONGEST = RETURN_MODE_DUAL | RETURN_TYPE_STRONGEST | (RETURN_TYPE_SECOND_STRONGEST << 8)
RGL_RETURN_FIRST_SECOND = RETURN_MODE_DUAL | RETURN_TYPE_FIRST | (RETURN_TYPE_SECOND << 8)

RGL_AXIS_X, RGL_AXIS_Y, RGL_AXIS_Z = 1, 2, 3

# ------------------------------------------------------------------------------------------ types
c_node = C.c_void_p
c_mesh = C.c_void_p
c_entity = C.c_void_p
c_texture = C.c_void_p
P = C.POINTER
f32p = P(C.c_float)
i32p = P(C.c_int32)

# (name, argtypes) of every exported function; restype is rgl_status_t unless listed in VOID_FUNCS.
_PROTOTYPES = [
    ("rgl_get_version_info", [i32p, i32p, i32p]),
    ("rgl_get_extension_info", [C.c_int32, i32p]),
    ("rgl_configure_logging", [C.c_int32, C.c_char_p, C.c_bool]),
    ("rgl_get_last_error_string", [P(C.c_char_p)]),
    ("rgl_cleanup", []),
    ("rgl_mesh_create", [P(c_mesh), C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]),
    ("rgl_mesh_set_texture_coords", [c_mesh, C.c_void_p, C.c_int32]),
    ("rgl_mesh_set_bone_weights", [c_mesh, C.c_void_p, C.c_int32]),
    ("rgl_mesh_set_restposes", [c_mesh, C.c_void_p, C.c_int32]),
    ("rgl_mesh_destroy", [c_mesh]),
    ("rgl_mesh_is_alive", [c_mesh, P(C.c_bool)]),
    ("rgl_entity_create", [P(c_entity), C.c_void_p, c_mesh]),
    ("rgl_entity_destroy", [c_entity]),
    ("rgl_entity_set_transform", [c_entity, C.c_void_p]),
    ("rgl_entity_set_pose_world", [c_entity, C.c_void_p, C.c_int32]),
    ("rgl_entity_set_id", [c_entity, C.c_int32]),
    ("rgl_entity_set_intensity_texture", [c_entity, c_texture]),
    ("rgl_entity_set_laser_retro", [c_entity, C.c_float]),
    ("rgl_entity_apply_external_animation", [c_entity, C.c_void_p, C.c_int32]),
    ("rgl_entity_is_alive", [c_entity, P(C.c_bool)]),
    ("rgl_texture_create", [P(c_texture), C.c_void_p, C.c_int32, C.c_int32]),
    ("rgl_texture_destroy", [c_texture]),
    ("rgl_texture_is_alive", [c_texture, P(C.c_bool)]),
    ("rgl_scene_set_time", [C.c_void_p, C.c_uint64]),
    ("rgl_node_rays_from_mat3x4f", [P(c_node), C.c_void_p, C.c_int32]),
    ("rgl_node_rays_set_ring_ids", [P(c_node), C.c_void_p, C.c_int32]),
    ("rgl_node_rays_set_range", [P(c_node), C.c_void_p, C.c_int32]),
    ("rgl_node_rays_set_time_offsets", [P(c_node), C.c_void_p, C.c_int32]),
    ("rgl_node_rays_transform", [P(c_node), C.c_void_p]),
    ("rgl_node_points_transform", [P(c_node), C.c_void_p]),
    ("rgl_node_raytrace", [P(c_node), C.c_void_p]),
    ("rgl_node_raytrace_configure_velocity", [c_node, C.c_void_p, C.c_void_p]),
    ("rgl_node_raytrace_configure_distortion", [c_node, C.c_bool]),
    ("rgl_node_raytrace_configure_non_hits", [c_node, C.c_float, C.c_float]),
    ("rgl_node_raytrace_configure_mask", [c_node, C.c_void_p, C.c_int32]),
    ("rgl_node_raytrace_configure_beam_divergence", [c_node, C.c_float, C.c_float]),
    ("rgl_node_raytrace_configure_default_intensity", [c_node, C.c_float]),
    ("rgl_node_raytrace_configure_return_mode", [c_node, C.c_int32]),
    ("rgl_node_points_format", [P(c_node), C.c_void_p, C.c_int32]),
    ("rgl_node_points_yield", [P(c_node), C.c_void_p, C.c_int32]),
    ("rgl_node_points_compact_by_field", [P(c_node), C.c_int32]),
    ("rgl_node_points_compact", [P(c_node)]),
    ("rgl_node_points_spatial_merge", [P(c_node), C.c_void_p, C.c_int32]),
    ("rgl_node_points_temporal_merge", [P(c_node), C.c_void_p, C.c_int32]),
    ("rgl_node_points_from_array", [P(c_node), C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]),
    ("rgl_node_points_radar_postprocess", [P(c_node), C.c_void_p, C.c_int32] + [C.c_float] * 7),
    ("rgl_node_points_radar_track_objects", [P(c_node)] + [C.c_float] * 7),
    ("rgl_node_points_radar_set_classes", [c_node, C.c_void_p, C.c_void_p, C.c_int32]),
    ("rgl_node_points_filter_ground", [P(c_node), C.c_void_p, C.c_float]),
    ("rgl_node_gaussian_noise_angular_ray", [P(c_node), C.c_float, C.c_float, C.c_int32]),
    ("rgl_node_gaussian_noise_angular_hitpoint", [P(c_node), C.c_float, C.c_float, C.c_int32]),
    ("rgl_node_gaussian_noise_distance", [P(c_node), C.c_float, C.c_float, C.c_float]),
    ("rgl_node_is_alive", [c_node, P(C.c_bool)]),
    ("rgl_graph_run", [c_node]),
    ("rgl_graph_destroy", [c_node]),
    ("rgl_graph_get_result_size", [c_node, C.c_int32, i32p, i32p]),
    ("rgl_graph_get_result_data", [c_node, C.c_int32, C.c_void_p]),
    ("rgl_graph_node_add_child", [c_node, c_node]),
    ("rgl_graph_node_remove_child", [c_node, c_node]),
    ("rgl_graph_node_set_priority", [c_node, C.c_int32]),
    ("rgl_graph_node_get_priority", [c_node, i32p]),
    # extensions (include/rgl/api/extensions/amd.h)
    ("rgl_graph_get_execution_order", [c_node, P(c_node), i32p]),
    ("rgl_node_gaussian_noise_set_seed", [c_node, C.c_uint64]),
    ("rgl_node_gaussian_noise_get_stream", [c_node, P(C.c_uint64), P(C.c_uint32), P(C.c_uint32)]),
    ("rgl_scene_set_tlas_refit", [C.c_void_p, C.c_bool]),
    ("rgl_node_raytrace_traversal_stats", [c_node, i32p, i32p, i32p, C.c_int32]),
    ("rgl_node_raytrace_brute_force", [c_node, f32p, i32p, i32p, C.c_int32]),
    ("rgl_graph_traversal_footprint", [P(c_node), C.c_int32, P(C.c_int64)]),
    ("rgl_noise_normals", [C.c_uint64, C.c_uint32, C.c_uint32, f32p, C.c_int32]),
    ("rgl_noise_normals_no_alloc", [C.c_uint64, C.c_uint32, C.c_uint32, f32p, C.c_int32]),
    ("rgl_node_raytrace_set_kernel_timing", [c_node, C.c_bool]),
    ("rgl_node_raytrace_get_kernel_timing", [c_node, P(C.c_double), i32p]),
    ("rgl_scene_create", [P(C.c_void_p)]),
    ("rgl_scene_destroy", [C.c_void_p]),
    ("rgl_graph_run_batch", [P(c_node), C.c_int32]),
    ("rgl_graph_get_result_sizes", [P(c_node), C.c_int32, C.c_int32, i32p]),
    ("rgl_graph_batch_set_kernel_timing", [C.c_bool]),
    ("rgl_graph_batch_get_kernel_timing", [P(C.c_double), i32p]),
    ("rgl_graph_batch_configure_autotune", [C.c_bool]),
    ("rgl_graph_batch_force_variants", [C.c_int32, C.c_int32]),
    ("rgl_graph_batch_get_variants", [i32p, i32p]),
    ("rgl_tape_record_begin", [C.c_char_p]),
    ("rgl_tape_record_end", []),
    ("rgl_tape_record_is_active", [P(C.c_bool)]),
    ("rgl_tape_play", [C.c_char_p]),
    ("rgl_graph_batch_configure_capture", [C.c_bool]),
    ("rgl_graph_batch_get_capture_stats", [P(C.c_int64), P(C.c_int64)]),
    ("rgl_graph_batch_get_fused_stages", [P(C.c_int64)]),
    ("rgl_configure_bvh_format", [C.c_int32]),
    ("rgl_get_bvh_format", [i32p]),
    ("rgl_node_points_radar_bin", [P(c_node)] + [C.c_float] * 6),
    ("rgl_node_points_yield_configure_prefetch", [c_node, C.c_bool]),
    ("rgl_server_create", [P(C.c_void_p), C.c_int32, C.c_int32, C.c_int64, C.c_bool]),
    ("rgl_server_destroy", [C.c_void_p]),
    ("rgl_server_submit", [C.c_void_p, c_node, P(C.c_int64)]),
    ("rgl_server_step", [C.c_void_p, i32p]),
    ("rgl_server_poll", [C.c_void_p, C.c_int64, i32p, P(C.c_double)]),
    ("rgl_server_wait", [C.c_void_p, C.c_int64, P(C.c_double)]),
    ("rgl_server_get_stats", [C.c_void_p, C.c_void_p]),
    ("rgl_server_reset_stats", [C.c_void_p]),
]
BVH2, BVH4_F32, BVH4_Q8, BVH4_Q16 = 0, 1, 2, 3
VOID_FUNCS = {"rgl_get_last_error_string"}

_lib = None


class RGLError(RuntimeError):
    def __init__(self, func: str, status: int, message: str):
        super().__init__(f"{func} failed with status {status}: {message}")
        self.status = status
        self.message = message


def load(path: os.PathLike | None = None) -> C.CDLL:
    """Loads the HIP library (once) and declares every prototype."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else Path(os.environ.get("RGL_LIBRARY", LIB_PATH))
    if not p.exists():
        raise FileNotFoundError(f"{p} is not built; run __graft_entry__.build() or make -C robotecgpulidar_amd/csrc")
    lib = C.CDLL(str(p))
    for name, argtypes in _PROTOTYPES:
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = None if name in VOID_FUNCS else C.c_int32
    _lib = lib
    return lib


def exported_names() -> list[str]:
    return [name for name, _ in _PROTOTYPES]


def last_error() -> str:
    s = C.c_char_p()
    load().rgl_get_last_error_string(C.byref(s))
    return (s.value or b"").decode()


_FNS: dict = {}


def _fn(name: str):
    fn = _FNS.get(name)
    if fn is None:
        fn = _FNS[name] = getattr(load(), name)
    return fn


def call(name: str, *args) -> int:
    """Calls an API function and returns its status (no exception)."""
    return _fn(name)(*args)


def check(name: str, *args) -> None:
    st = _fn(name)(*args)
    if st != RGL_SUCCESS:
        raise RGLError(name, st, last_error())


# ------------------------------------------------------------------------------------------ helpers
def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def mat(m) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(m, dtype=np.float32).reshape(3, 4))
    return a


def mesh_create(vertices, indices) -> C.c_void_p:
    v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float32).reshape(-1, 3))
    i = np.ascontiguousarray(np.asarray(indices, dtype=np.int32).reshape(-1, 3))
    h = C.c_void_p()
    check("rgl_mesh_create", C.byref(h), _ptr(v), len(v), _ptr(i), len(i))
    return h


def mesh_set_texture_coords(mesh, uvs) -> None:
    u = np.ascontiguousarray(np.asarray(uvs, dtype=np.float32).reshape(-1, 2))
    check("rgl_mesh_set_texture_coords", mesh, _ptr(u), len(u))


def entity_create(mesh, scene=None) -> C.c_void_p:
    h = C.c_void_p()
    check("rgl_entity_create", C.byref(h), scene, mesh)
    return h


def scene_create() -> C.c_void_p:
    h = C.c_void_p()
    check("rgl_scene_create", C.byref(h))
    return h


def scene_destroy(scene) -> None:
    check("rgl_scene_destroy", scene)


def entity_set_transform(entity, m) -> None:
    a = mat(m)
    check("rgl_entity_set_transform", entity, _ptr(a))


def entity_apply_external_animation(entity, vertices) -> None:
    """Replaces the entity's object-space vertices (one per mesh vertex)."""
    v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float32).reshape(-1, 3))
    check("rgl_entity_apply_external_animation", entity, _ptr(v), len(v))


def mesh_set_bone_weights(mesh, weights, bones) -> None:
    """Per-vertex skinning weights (N, 4) float32 and bone indices (N, 4) int32."""
    w = np.asarray(weights, np.float32).reshape(-1, 4)
    b = np.asarray(bones, np.int32).reshape(-1, 4)
    rec = np.ascontiguousarray(np.concatenate([w.view(np.uint32), b.view(np.uint32)], axis=1))
    check("rgl_mesh_set_bone_weights", mesh, _ptr(rec), len(rec))


def mesh_set_restposes(mesh, poses) -> None:
    a = np.ascontiguousarray(np.asarray(poses, np.float32).reshape(-1, 3, 4))
    check("rgl_mesh_set_restposes", mesh, _ptr(a), len(a))


def entity_set_pose_world(entity, poses) -> None:
    a = np.ascontiguousarray(np.asarray(poses, np.float32).reshape(-1, 3, 4))
    check("rgl_entity_set_pose_world", entity, _ptr(a), len(a))


def scene_set_time(scene, nanoseconds: int) -> None:
    check("rgl_scene_set_time", scene, C.c_uint64(int(nanoseconds)))


def entity_set_id(entity, i: int) -> None:
    check("rgl_entity_set_id", entity, int(i))


def texture_create(texels: np.ndarray) -> C.c_void_p:
    t = np.ascontiguousarray(texels, dtype=np.uint8)
    h = C.c_void_p()
    check("rgl_texture_create", C.byref(h), _ptr(t), t.shape[1], t.shape[0])
    return h


class Node:
    """Holder of an rgl_node_t handle (create-or-update semantics like the C API)."""

    def __init__(self):
        self.h = C.c_void_p()

    @property
    def ref(self):
        return C.byref(self.h)


def node_rays_from_mat3x4f(node: Node, rays: np.ndarray) -> Node:
    r = np.ascontiguousarray(np.asarray(rays, dtype=np.float32).reshape(-1, 3, 4))
    check("rgl_node_rays_from_mat3x4f", node.ref, _ptr(r), len(r))
    return node


def node_rays_set_range(node: Node, ranges) -> Node:
    r = np.ascontiguousarray(np.asarray(ranges, dtype=np.float32).reshape(-1, 2))
    check("rgl_node_rays_set_range", node.ref, _ptr(r), len(r))
    return node


def node_rays_set_ring_ids(node: Node, ids) -> Node:
    r = np.ascontiguousarray(np.asarray(ids, dtype=np.int32).reshape(-1))
    check("rgl_node_rays_set_ring_ids", node.ref, _ptr(r), len(r))
    return node


def node_rays_set_time_offsets(node: Node, offsets) -> Node:
    r = np.ascontiguousarray(np.asarray(offsets, dtype=np.float32).reshape(-1))
    check("rgl_node_rays_set_time_offsets", node.ref, _ptr(r), len(r))
    return node


def node_rays_transform(node: Node, m) -> Node:
    a = mat(m)
    check("rgl_node_rays_transform", node.ref, _ptr(a))
    return node


def node_points_transform(node: Node, m) -> Node:
    a = mat(m)
    check("rgl_node_points_transform", node.ref, _ptr(a))
    return node


def node_raytrace(node: Node, scene=None) -> Node:
    check("rgl_node_raytrace", node.ref, scene)
    return node


def node_points_yield(node: Node, fields) -> Node:
    f = np.ascontiguousarray(np.asarray(fields, dtype=np.int32))
    check("rgl_node_points_yield", node.ref, _ptr(f), len(f))
    return node


def node_points_format(node: Node, fields) -> Node:
    f = np.ascontiguousarray(np.asarray(fields, dtype=np.int32))
    check("rgl_node_points_format", node.ref, _ptr(f), len(f))
    return node


def node_points_compact_by_field(node: Node, field: int = IS_HIT_I32) -> Node:
    check("rgl_node_points_compact_by_field", node.ref, int(field))
    return node


def node_points_radar_bin(node: Node, range_step: float, range_max: float, azimuth_step: float, azimuth_min: float,
                          azimuth_max: float, ray_solid_angle: float) -> Node:
    check("rgl_node_points_radar_bin", node.ref, range_step, range_max, azimuth_step, azimuth_min, azimuth_max,
          ray_solid_angle)
    return node


def node_points_radar_postprocess(node: Node, scopes, azimuth_step: float, elevation_step: float, frequency: float,
                                  power_transmitted: float, device_gain: float, noise_mean: float,
                                  noise_st_dev: float) -> Node:
    """`scopes`: rows of (begin_distance, end_distance, distance_threshold, radial_speed_threshold,
    azimuth_threshold), the layout of rgl_radar_scope_t."""
    sc = np.ascontiguousarray(np.asarray(scopes, np.float32).reshape(-1, 5))
    check("rgl_node_points_radar_postprocess", node.ref, _ptr(sc), len(sc), azimuth_step, elevation_step, frequency,
          power_transmitted, device_gain, noise_mean, noise_st_dev)
    return node


def node_points_from_array(node: Node, records: np.ndarray, fields) -> Node:
    f = np.ascontiguousarray(np.asarray(fields, dtype=np.int32))
    buf = np.ascontiguousarray(records)
    count = buf.shape[0]
    check("rgl_node_points_from_array", node.ref, _ptr(buf), count, _ptr(f), len(f))
    return node


def add_child(parent: Node, child: Node) -> None:
    check("rgl_graph_node_add_child", parent.h, child.h)


def chain(*nodes: Node) -> None:
    for a, b in zip(nodes, nodes[1:]):
        add_child(a, b)


def execution_order(node: Node) -> list[int]:
    """Handles (as integers) of the nodes of `node`'s graph in execution order."""
    count = C.c_int32(0)
    check("rgl_graph_get_execution_order", node.h, None, C.byref(count))
    arr = (c_node * max(count.value, 1))()
    check("rgl_graph_get_execution_order", node.h, arr, C.byref(count))
    return [arr[i] for i in range(count.value)]


def traversal_footprint(raytrace_nodes) -> int:
    """Unique bytes of BVH and scene records the given raytrace nodes' launches read (rgl_graph_traversal_footprint)."""
    arr = (c_node * len(raytrace_nodes))(*[n.h.value for n in raytrace_nodes])
    out = C.c_int64()
    check("rgl_graph_traversal_footprint", arr, len(raytrace_nodes), C.byref(out))
    return out.value


def get_priority(node: Node) -> int:
    p = C.c_int32()
    check("rgl_graph_node_get_priority", node.h, C.byref(p))
    return p.value


def graph_run(node: Node) -> None:
    check("rgl_graph_run", node.h)


def graph_run_batch(nodes) -> None:
    """Runs the graphs of `nodes` (one node per graph) as one batch."""
    arr = (c_node * max(len(nodes), 1))(*[n.h.value for n in nodes])
    check("rgl_graph_run_batch", arr, len(nodes))


class BatchRun:
    """rgl_graph_run_batch over a fixed list of nodes, with the node array built once (a C++ caller passes a static
    array; rebuilding the ctypes array every frame costs the host tens of microseconds)."""

    def __init__(self, nodes):
        self.n = len(nodes)
        self.arr = (c_node * max(self.n, 1))(*[x.h.value for x in nodes])

    def __call__(self) -> None:
        check("rgl_graph_run_batch", self.arr, self.n)


class ResultSizes:
    """Point counts of one field at a fixed list of nodes, read with one rgl_graph_get_result_sizes call."""

    def __init__(self, nodes, field: int):
        self.n = len(nodes)
        self.arr = (c_node * max(self.n, 1))(*[x.h.value for x in nodes])
        self.out = np.zeros(max(self.n, 1), np.int32)
        self.ptr = self.out.ctypes.data_as(i32p)
        self.field = int(field)

    def __call__(self) -> np.ndarray:
        check("rgl_graph_get_result_sizes", self.arr, self.n, self.field, self.ptr)
        return self.out[: self.n]


def batch_kernel_timing(enable: bool | None = None):
    """Enables/disables batched raytrace timing, or (enable=None) returns (total_ms, launches)."""
    if enable is not None:
        check("rgl_graph_batch_set_kernel_timing", bool(enable))
        return None
    ms = C.c_double()
    n = C.c_int32()
    check("rgl_graph_batch_get_kernel_timing", C.byref(ms), C.byref(n))
    return ms.value, n.value


def batch_fused_stages() -> int:
    """Point stages folded into batched raytrace launches so far."""
    v = C.c_int64()
    check("rgl_graph_batch_get_fused_stages", C.byref(v))
    return v.value


def batch_capture(enable: bool | None = None):
    """Enables/disables HIP graph capture of batches, or (enable=None) returns (captures, replays)."""
    if enable is not None:
        check("rgl_graph_batch_configure_capture", bool(enable))
        return None
    cap = C.c_int64()
    rep = C.c_int64()
    check("rgl_graph_batch_get_capture_stats", C.byref(cap), C.byref(rep))
    return cap.value, rep.value


def result_size(node: Node, field: int) -> tuple[int, int]:
    n = C.c_int32()
    s = C.c_int32()
    check("rgl_graph_get_result_size", node.h, int(field), C.byref(n), C.byref(s))
    return n.value, s.value


def result(node: Node, field: int) -> np.ndarray:
    """Result array of one field: shape (count,) or (count, k); DYNAMIC_FORMAT gives raw bytes (count, size)."""
    n, s = result_size(node, field)
    if field == DYNAMIC_FORMAT:
        out = np.empty((n, s), dtype=np.uint8)
    else:
        dt, k = FIELD_DTYPE[field]
        out = np.empty((n, k) if k > 1 else (n,), dtype=dt)
    if n > 0:
        check("rgl_graph_get_result_data", node.h, int(field), _ptr(out))
    return out


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("batches", C.c_int64),
                ("rays_completed", C.c_int64), ("evictions", C.c_int64), ("deferred", C.c_int64), ("failed", C.c_int64),
                ("resident_bytes", C.c_int64), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("max_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Server:
    """rgl_server_t: continuous batching of sensor frames (see include/rgl/api/extensions/amd.h)."""
    QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3

    def __init__(self, slots: int = 2, max_batch_graphs: int = 64, hbm_budget: int = 0, batching: bool = True):
        self.h = C.c_void_p()
        check("rgl_server_create", C.byref(self.h), slots, max_batch_graphs, C.c_int64(hbm_budget), bool(batching))

    def submit(self, node: Node) -> int:
        t = C.c_int64()
        check("rgl_server_submit", self.h, node.h, C.byref(t))
        return t.value

    def step(self) -> int:
        n = C.c_int32()
        check("rgl_server_step", self.h, C.byref(n))
        return n.value

    def poll(self, ticket: int):
        st, ms = C.c_int32(), C.c_double()
        check("rgl_server_poll", self.h, C.c_int64(ticket), C.byref(st), C.byref(ms))
        return st.value, ms.value

    def wait(self, ticket: int) -> float:
        ms = C.c_double()
        check("rgl_server_wait", self.h, C.c_int64(ticket), C.byref(ms))
        return ms.value

    def stats(self) -> dict:
        st = ServerStats()
        check("rgl_server_get_stats", self.h, C.byref(st))
        return st.as_dict()

    def reset_stats(self) -> None:
        check("rgl_server_reset_stats", self.h)

    def destroy(self) -> None:
        if self.h:
            check("rgl_server_destroy", self.h)
            self.h = C.c_void_p()


def set_bvh_format(fmt: int) -> None:
    check("rgl_configure_bvh_format", int(fmt))


def get_bvh_format() -> int:
    v = C.c_int32()
    check("rgl_get_bvh_format", C.byref(v))
    return v.value


def cleanup() -> None:
    check("rgl_cleanup")


def noise_normals(seed: int, run: int, tag: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.float32)
    check("rgl_noise_normals", C.c_uint64(seed), run, tag, out.ctypes.data_as(f32p), count)
    return out


def noise_normals_no_alloc(seed: int, run: int, tag: int, count: int) -> np.ndarray:
    """rgl_noise_normals_no_alloc: the same variates staged without device allocation (count <= 2^20)."""
    out = np.empty(count, dtype=np.float32)
    check("rgl_noise_normals_no_alloc", C.c_uint64(seed), run, tag, out.ctypes.data_as(f32p), count)
    return out
