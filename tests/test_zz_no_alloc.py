"""rgl_noise_normals_no_alloc (the allocation-free entry point) against rgl_noise_normals: two calls in a row, the
second with fewer variates, must give the allocating entry point's variates bit for bit; a count beyond the staging
array's 2^20 floats is RGL_INVALID_ARGUMENT."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_no_alloc_normals_equal_allocating_entry_point(rgl_lib):
    rgl = rgl_lib
    for seed, run, tag, count in ((0xC0FFEE, 3, 2, 1 << 20), (0x1234, 0, 1, 4097)):
        want = rgl.noise_normals(seed, run, tag, count)
        got = rgl.noise_normals_no_alloc(seed, run, tag, count)
        assert got.tobytes() == want.tobytes(), (seed, count)
    # the second call left the first call's tail in the staging array: a later small call reads only its own range
    assert rgl.noise_normals_no_alloc(7, 7, 7, 3).tobytes() == rgl.noise_normals(7, 7, 7, 3).tobytes()
    out = np.empty((1 << 20) + 1, np.float32)
    status = rgl.call("rgl_noise_normals_no_alloc", C.c_uint64(1), 0, 0, out.ctypes.data_as(C.POINTER(C.c_float)),
                      len(out))
    assert status == rgl.RGL_INVALID_ARGUMENT and "1048576" in rgl.last_error()
