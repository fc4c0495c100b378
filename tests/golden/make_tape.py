"""Writes tests/golden/tape_cube_scene.{yaml,bin}: a tape in the reference's format (src/tape/TapeRecorder.hpp:27-151:
a YAML block sequence of `name: {t: ns, a: [args]}` maps as yaml-cpp emits them, arrays in the .bin file at 16-byte
aligned offsets, object handles as integers), written here independently of the library so that replaying it
checks the library's tape reader against the format itself (run from the repository root):

    python -m tests.golden.make_tape

The tape builds a scene of two unit cubes (half-edge 1): entity 42 at (0, 0, 5) and entity 7 at (0, 4, 0) rotated
45 degrees about z, then a one-ray graph (rays -> raytrace -> yield) that it runs once and reads back.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
CUBE_V = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                   [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], np.float32)
CUBE_I = np.array([[0, 3, 1], [3, 2, 1], [1, 2, 5], [2, 6, 5], [5, 6, 4], [6, 7, 4],
                   [4, 7, 0], [7, 3, 0], [3, 7, 2], [7, 6, 2], [4, 0, 5], [0, 1, 5]], np.int32)


class Writer:
    def __init__(self):
        self.lines, self.bin, self.t = [], bytearray(), 1000

    def array(self, a) -> int:
        data = np.ascontiguousarray(a).tobytes()
        off = len(self.bin)
        self.bin += data + bytes((16 - len(data) % 16) % 16)
        return off

    def call(self, fn, *args):
        self.t += 2500
        txt = ", ".join(str(a).lower() if isinstance(a, bool) else str(a) for a in args)
        self.lines.append(f"- {fn}: {{t: {self.t}, a: [{txt}]}}")


def main():
    w = Writer()
    w.call("rgl_get_version_info", 0, 19, 0)
    mesh, e1, e2 = 94000001, 94000101, 94000102
    w.call("rgl_mesh_create", mesh, w.array(CUBE_V), 8, w.array(CUBE_I), 12)
    w.call("rgl_entity_create", e1, 0, mesh)
    w.call("rgl_entity_set_transform", e1, w.array(np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 5]], np.float32)))
    w.call("rgl_entity_set_id", e1, 42)
    w.call("rgl_entity_create", e2, 0, mesh)
    c, s = math.cos(math.pi / 4), math.sin(math.pi / 4)
    w.call("rgl_entity_set_transform", e2, w.array(np.array([[c, -s, 0, 0], [s, c, 0, 4], [0, 0, 1, 0]], np.float32)))
    w.call("rgl_entity_set_id", e2, 7)
    w.call("rgl_entity_set_laser_retro", e2, 0.25)
    w.call("rgl_scene_set_time", 0, 1500000000)
    rays, rt, out = 95000001, 95000002, 95000003
    w.call("rgl_node_rays_from_mat3x4f", rays, w.array(np.eye(3, 4, dtype=np.float32)[None]), 1)
    w.call("rgl_node_raytrace", rt, 0)
    w.call("rgl_node_raytrace_configure_non_hits", rt, 0, ".inf")
    w.call("rgl_node_points_yield", out, w.array(np.array([1, 8, 7], np.int32)), 3)
    w.call("rgl_graph_node_add_child", rays, rt)
    w.call("rgl_graph_node_add_child", rt, out)
    w.call("rgl_graph_run", rays)
    w.call("rgl_graph_get_result_size", out, 8, 1, 4)
    w.call("rgl_graph_get_result_data", out, 8, 140000000)
    (HERE / "tape_cube_scene.yaml").write_text("\n".join(w.lines) + "\n")
    (HERE / "tape_cube_scene.bin").write_bytes(bytes(w.bin))


if __name__ == "__main__":
    main()
