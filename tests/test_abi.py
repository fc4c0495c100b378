"""The C-ABI library loads and exports every function the public headers declare (no GPU calls)."""
import ctypes
import re
from pathlib import Path

from robotecgpulidar_amd import rgl

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = []
    for h in sorted((ROOT / "include").rglob("*.h")):
        names += re.findall(r"RGL_API\s+(?:rgl_status_t|void)\s+(rgl_\w+)\s*\(", h.read_text())
    return names


def test_headers_declare_the_reference_api():
    names = set(declared_functions())
    # The drop-in set named for this path (include/rgl/api/core.h of RGL 0.19.0).
    for n in ["rgl_mesh_create", "rgl_entity_create", "rgl_node_rays_from_mat3x4f", "rgl_node_raytrace",
              "rgl_node_points_compact_by_field", "rgl_node_gaussian_noise_angular_ray",
              "rgl_node_gaussian_noise_angular_hitpoint", "rgl_node_gaussian_noise_distance", "rgl_graph_run",
              "rgl_graph_get_result_data"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(rgl.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_functions()) == set(rgl.exported_names())
