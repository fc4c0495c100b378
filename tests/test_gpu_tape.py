"""Tapes (the reference's tape extension: rgl_tape_record_begin / _end / _is_active, rgl_tape_play;
src/tape/TapeRecorder.hpp, src/tape/TapePlayer.cpp, test/src/TapeTest.cpp): the reference's format errors, recording
and replaying every recordable call, scene reconstruction from a tape (results bit-identical to the recorded
session's), and replay of a tape written independently in the reference's format (tests/golden/make_tape.py) whose
scene is checked in closed form."""
import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S
from tests.helpers import Case, Graph

pytestmark = pytest.mark.gpu

XYZ, IS_HIT, ENT, DIST, RETRO, NORMAL = 1, 4, 7, 8, 25, 22
GOLDEN = Path(__file__).resolve().parent / "golden"


def play(rgl, path):
    return rgl.call("rgl_tape_play", str(path).encode())


def write_tape(tmp_path, name, yaml_text):
    (tmp_path / f"{name}.yaml").write_text(yaml_text)
    (tmp_path / f"{name}.bin").write_bytes(b"")
    return tmp_path / name


def test_tape_format_errors(rgl_lib, tmp_path):
    rgl = rgl_lib
    cases = [
        ("empty", "", "Invalid Tape: Empty YAML file detected"),
        ("noversion", "- rgl_mesh_create: {t: 0, a: [0]}\n", "Unsupported Tape format: Missing version record in the Tape"),
        ("outdated", "- rgl_get_version_info: {t: 0, a: [0, 16, 2]}\n",
         "Unsupported Tape Format: Tape version is too old. Required version: 0.16.3."),
        ("structure", "rgl_version: {major: 0, minor: 16, patch: 2}\n", "Unsupported Tape format: Detected outdated format"),
        ("unknown", "- rgl_get_version_info: {t: 0, a: [0, 19, 0]}\n- rgl_no_such_call: {t: 1, a: []}\n",
         "unknown function to play: rgl_no_such_call"),
    ]
    for name, text, msg in cases:
        assert play(rgl, write_tape(tmp_path, name, text)) == rgl.RGL_TAPE_ERROR, name
        assert msg in rgl.last_error(), (name, rgl.last_error())
    assert play(rgl, tmp_path / "missing") == rgl.RGL_INVALID_FILE_PATH


def test_record_state_errors(rgl_lib, tmp_path):
    rgl = rgl_lib
    active = C.c_bool(True)
    rgl.check("rgl_tape_record_is_active", C.byref(active))
    assert not active.value
    assert rgl.call("rgl_tape_record_end") == rgl.RGL_TAPE_ERROR
    rgl.check("rgl_tape_record_begin", str(tmp_path / "a").encode())
    rgl.check("rgl_tape_record_is_active", C.byref(active))
    assert active.value
    assert rgl.call("rgl_tape_record_begin", str(tmp_path / "b").encode()) == rgl.RGL_TAPE_ERROR
    assert play(rgl, tmp_path / "a") == rgl.RGL_TAPE_ERROR  # not while recording
    rgl.check("rgl_tape_record_end")
    rgl.check("rgl_tape_record_is_active", C.byref(active))
    assert not active.value
    assert play(rgl, tmp_path / "a") == rgl.RGL_SUCCESS


def test_record_play_logging(rgl_lib, tmp_path):
    rgl = rgl_lib
    rgl.check("rgl_tape_record_begin", str(tmp_path / "logging").encode())
    log = str(tmp_path / "rgl.log").encode()
    rgl.check("rgl_configure_logging", 2, log, False)
    rgl.check("rgl_tape_record_end")
    text = (tmp_path / "logging.yaml").read_text()
    assert text.startswith("- rgl_get_version_info: {t: ") and "rgl_configure_logging" in text
    assert play(rgl, tmp_path / "logging") == rgl.RGL_SUCCESS, rgl.last_error()
    rgl.check("rgl_configure_logging", 6, None, False)


def test_record_play_all_calls(rgl_lib, tmp_path):
    """Every recordable call (test/src/TapeTest.cpp:166-380) records and replays."""
    rgl = rgl_lib
    rgl.check("rgl_tape_record_begin", str(tmp_path / "all").encode())
    ident = S.f32(S.identity())
    a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
    rgl.check("rgl_get_version_info", C.byref(a), C.byref(b), C.byref(c))
    mesh = rgl.mesh_create(S.CUBE_VERTICES, S.CUBE_INDICES)
    rgl.mesh_set_bone_weights(mesh, np.zeros((8, 4)), np.zeros((8, 4), np.int32))
    rgl.mesh_set_restposes(mesh, np.stack([ident, ident]))
    ent = rgl.entity_create(mesh)
    rgl.entity_set_transform(ent, ident)
    rgl.entity_set_pose_world(ent, np.stack([ident, ident]))
    rgl.entity_set_id(ent, 1)
    rgl.entity_apply_external_animation(ent, S.CUBE_VERTICES)
    tex = rgl.texture_create(np.tile(np.uint8([0, 255]), (64, 32)))
    rgl.check("rgl_mesh_set_texture_coords", mesh, S.CUBE_UVS.ctypes.data, 8)
    rgl.check("rgl_entity_set_intensity_texture", ent, tex)
    rgl.check("rgl_entity_set_laser_retro", ent, C.c_float(50.0))
    rgl.scene_set_time(None, 1_500_000_000)
    rays = rgl.node_rays_from_mat3x4f(rgl.Node(), np.stack([ident, ident]))
    rings = rgl.node_rays_set_ring_ids(rgl.Node(), np.array([0, 1], np.int32))
    offs = rgl.node_rays_set_time_offsets(rgl.Node(), np.array([1.0, 2.0], np.float32))
    rng = rgl.node_rays_set_range(rgl.Node(), np.array([[0.0, 100.0]], np.float32))
    tf = rgl.node_rays_transform(rgl.Node(), ident)
    rt = rgl.node_raytrace(rgl.Node())
    z3 = np.zeros(3, np.float32)
    rgl.check("rgl_node_raytrace_configure_velocity", rt.h, z3.ctypes.data, z3.ctypes.data)
    rgl.check("rgl_node_raytrace_configure_distortion", rt.h, False)
    rgl.check("rgl_node_raytrace_configure_non_hits", rt.h, 0.0, 100.0)
    mask = np.ones(2, np.int8)
    rgl.check("rgl_node_raytrace_configure_mask", rt.h, mask.ctypes.data, 2)
    rgl.check("rgl_node_raytrace_configure_beam_divergence", rt.h, 0.0, 0.0)
    rgl.check("rgl_node_raytrace_configure_default_intensity", rt.h, 10.0)
    rgl.check("rgl_node_raytrace_configure_return_mode", rt.h, rgl.RGL_RETURN_FIRST)
    noise_r = rgl.Node()
    rgl.check("rgl_node_gaussian_noise_angular_ray", noise_r.ref, 0.0, 0.0, 3)
    noise_h = rgl.Node()
    rgl.check("rgl_node_gaussian_noise_angular_hitpoint", noise_h.ref, 0.0, 0.0, 3)
    noise_d = rgl.Node()
    rgl.check("rgl_node_gaussian_noise_distance", noise_d.ref, 0.0, 0.0, 0.0)
    pt = rgl.node_points_transform(rgl.Node(), ident)
    comp = rgl.node_points_compact_by_field(rgl.Node(), IS_HIT)
    fmt = rgl.node_points_format(rgl.Node(), [XYZ, DIST])
    out = rgl.node_points_yield(rgl.Node(), [XYZ, DIST, IS_HIT])
    rgl.chain(rays, rings, offs, rng, tf, noise_r, rt, noise_h, noise_d, pt, comp, out)
    rgl.add_child(out, fmt)
    p = C.c_int32()
    rgl.check("rgl_graph_node_set_priority", out.h, 1)
    rgl.check("rgl_graph_node_get_priority", out.h, C.byref(p))
    rgl.graph_run(rays)
    rgl.result(out, XYZ)
    rgl.result(fmt, 13842)
    rgl.check("rgl_graph_node_remove_child", out.h, fmt.h)
    rgl.check("rgl_graph_destroy", fmt.h)
    rgl.check("rgl_graph_destroy", rays.h)
    rgl.check("rgl_entity_destroy", ent)
    rgl.check("rgl_texture_destroy", tex)
    rgl.check("rgl_mesh_destroy", mesh)
    rgl.check("rgl_tape_record_end")
    rgl.check("rgl_cleanup")
    text = (tmp_path / "all.yaml").read_text()
    for fn in ("rgl_mesh_create", "rgl_entity_set_pose_world", "rgl_node_points_format", "rgl_graph_get_result_data",
               "rgl_graph_destroy", "rgl_node_raytrace_configure_mask", "rgl_texture_create"):
        assert f"- {fn}: " in text, fn
    assert play(rgl, tmp_path / "all") == rgl.RGL_SUCCESS, rgl.last_error()


def test_scene_reconstruction(rgl_lib, tmp_path):
    """A tape of scene construction rebuilds the same scene: a graph traced after replay gives the recorded
    session's results bit for bit (test/src/TapeTest.cpp:421-437)."""
    rgl = rgl_lib
    spec = S.config1_scene()
    rgl.check("rgl_tape_record_begin", str(tmp_path / "scene").encode())
    _, ents, _ = S.load_into_rgl(spec)
    rgl.entity_set_transform(ents[3], S.f32(S.trs((4.0, -2.0, 0.5), (0, 0, 30), (2, 2, 2))))
    rgl.check("rgl_tape_record_end")
    lidar = S.config1_lidar()
    case = Case(spec, lidar.rays, pose=lidar.pose, ranges=np.array([lidar.range], np.float32))
    fields = [XYZ, DIST, ENT, NORMAL, IS_HIT]
    before = Graph(rgl, case, fields).run()
    rgl.check("rgl_cleanup")
    assert play(rgl, tmp_path / "scene") == rgl.RGL_SUCCESS, rgl.last_error()
    after = Graph(rgl, case, fields).run()
    assert int(before[IS_HIT].sum()) > 10000
    for f in fields:
        assert before[f].tobytes() == after[f].tobytes(), f


def test_golden_tape_in_reference_format(rgl_lib):
    """tests/golden/tape_cube_scene: two cubes, entity 42 at (0, 0, 5) and entity 7 at (0, 4, 0) turned 45 degrees
    about z. A ray up +Z hits the bottom face at distance 4; a ray along +Y meets the turned cube's edge at
    y = 4 - sqrt(2)."""
    rgl = rgl_lib
    assert play(rgl, GOLDEN / "tape_cube_scene") == rgl.RGL_SUCCESS, rgl.last_error()
    up = S.f32(S.identity())
    side = S.f32(S.rotation_deg(-90.0, 0.0, 0.0))  # local +Z -> world +Y
    case = Case(S.SceneSpec(), np.stack([up, side]))
    res = Graph(rgl, case, [XYZ, DIST, ENT, RETRO, IS_HIT]).run()
    assert res[IS_HIT].tolist() == [1, 1]
    assert res[ENT].tolist() == [42, 7]
    assert res[DIST][0] == np.float32(4.0)
    assert abs(res[DIST][1] - (4.0 - math.sqrt(2.0))) < 1e-5
    assert res[RETRO].tolist() == [0.0, 0.25]
