"""Trace order: a batched raytrace launch of at least 2^20 rays traces the rays of every rays node of 4096 or more
rays in the order of their directions' (azimuth, elevation) tiles (thread t of the launch traces a permuted beam),
while a single graph's launch and smaller batches trace in index order. Every ray still writes its own index, so the
results must be bit-identical either way. The check runs a 1,049,600-ray sensor alone (index order) and as a batch
(direction order), with masks, ring ids, per-ray ranges, time offsets with velocity distortion, and compaction."""
import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S
from tests.helpers import Case, Graph

pytestmark = pytest.mark.gpu

XYZ, IS_HIT, RAY_IDX, ENT, DIST, AZ, EL, RING, NORMAL = 1, 4, 6, 7, 8, 9, 10, 11, 22
FIELDS = [XYZ, DIST, ENT, NORMAL, IS_HIT, RAY_IDX, RING, AZ, EL]


def sensor():
    rays = S.spinning_lidar(128, 8200, -22.0, 12.0)  # 1,049,600 rays, channel by channel
    n = len(rays)
    rng = np.random.default_rng(3)
    return Case(S.config1_scene(), rays, pose=S.f32(S.trs((1.5, -2.0, 1.8), (0, 0, 33.0), (1, 1, 1))),
                ranges=np.stack([rng.uniform(0.1, 2.0, n), rng.uniform(40.0, 100.0, n)], 1).astype(np.float32),
                mask=(rng.uniform(0, 1, n) > 0.1).astype(np.int8), ring_ids=np.arange(n, dtype=np.int32) % 128,
                time_offsets=np.linspace(0.0, 90.0, n, dtype=np.float32), distortion=True,
                velocity=((8.0, 1.0, 0.0), (0.0, 0.0, 0.4)), near=0.0, far=100.0)


@pytest.mark.parametrize("compact", [False, True])
def test_direction_order_gives_index_order_results(rgl_lib, compact):
    rgl = rgl_lib
    case = sensor()
    assert len(case.rays) >= 1 << 20
    S.load_into_rgl(case.spec)
    g = Graph(rgl, case, FIELDS, compact=compact)
    rgl.graph_run(g.rays)  # one graph: index order
    solo = {f: rgl.result(g.out, f) for f in FIELDS}
    rgl.graph_run_batch([g.rays])  # a batched launch of >= 2^20 rays: direction order
    batched = {f: rgl.result(g.out, f) for f in FIELDS}
    hits = int(solo[IS_HIT].sum())
    print(f"{len(solo[XYZ])} points, {hits} hits")
    assert 100000 < hits
    if not compact:
        assert solo[RAY_IDX].tolist() == list(range(len(case.rays)))
    for f in FIELDS:
        assert solo[f].tobytes() == batched[f].tobytes(), f
