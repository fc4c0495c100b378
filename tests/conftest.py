import ctypes
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs through the C-ABI library")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (only for device discovery)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture
def rgl_lib():
    """The HIP library, with every API object released after the test (like the reference's RGLTest fixture,
    test/include/helpers/commonHelpers.hpp:32-38)."""
    from robotecgpulidar_amd import rgl
    lib = rgl.load()
    rgl.check("rgl_configure_logging", 6, None, False)
    rgl.set_bvh_format(rgl.BVH4_F32)  # process-wide settings a previous test may have changed
    rgl.check("rgl_scene_set_tlas_refit", None, False)
    yield rgl
    active = ctypes.c_bool(False)
    if rgl.call("rgl_tape_record_is_active", ctypes.byref(active)) == rgl.RGL_SUCCESS and active.value:
        rgl.call("rgl_tape_record_end")  # a failed tape test must not leave a recording running
    assert rgl.call("rgl_cleanup") == rgl.RGL_SUCCESS, rgl.last_error()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.load()
    return o
