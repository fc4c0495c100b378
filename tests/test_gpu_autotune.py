"""Launch-shape variants of rgl_graph_run_batch (raytrace block size or the persistent grid, compaction tile) never
change results, and the autotuner settles on one of each after timing them on live batches."""
import ctypes as C

import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S
from tests.test_gpu_batch import SensorGraph, assert_same, build_world, make_set

pytestmark = pytest.mark.gpu


@pytest.fixture
def lib(rgl_lib):
    yield rgl_lib
    rgl_lib.check("rgl_graph_batch_configure_autotune", True)


def variants(rgl):
    a, b = C.c_int32(), C.c_int32()
    rgl.check("rgl_graph_batch_get_variants", C.byref(a), C.byref(b))
    return a.value, b.value


@pytest.mark.parametrize("rt,cp", [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (2, 1)])
def test_forced_variants_equal_solo(lib, rt, cp):
    rgl = lib
    _, scenes, _ = build_world(rgl)
    solo, batched = make_set(rgl, scenes), make_set(rgl, scenes)
    rgl.check("rgl_graph_batch_force_variants", rt, cp)
    assert variants(rgl) == (rt, cp)
    for frame in range(2):
        for g in solo:
            rgl.graph_run(g.rays)
        rgl.graph_run_batch([g.rays for g in batched])
        for k, (a, b) in enumerate(zip(solo, batched)):
            assert_same(a.results(), b.results(), f"variants {rt},{cp} frame {frame} sensor {k}")


def test_autotune_settles_without_changing_results(lib):
    rgl = lib
    _, scenes, _ = build_world(rgl)
    # three 64 x 1024 lidars: more than 100k rays and points per batch, so every launch is timed while tuning
    mk = lambda: [SensorGraph(rgl, 64, 1024, scenes[k % 2], True, True, seed=100 + k) for k in range(3)]  # noqa
    solo, batched = mk(), mk()
    rgl.check("rgl_graph_batch_configure_autotune", True)
    assert variants(rgl) == (-1, -1)
    for frame in range(10):
        for g in solo:
            rgl.graph_run(g.rays)
        rgl.graph_run_batch([g.rays for g in batched])
        for k, (a, b) in enumerate(zip(solo, batched)):
            assert_same(a.results(), b.results(), f"tuning frame {frame} sensor {k}")
    rt, cp = variants(rgl)
    print(f"autotune chose raytrace variant {rt}, compaction variant {cp}")
    assert rt in (0, 1) and cp in (0, 1)
