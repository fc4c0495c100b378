"""Trace order: a rays node of at least 4096 rays is traced in the order of its rays' direction Morton codes (thread t
of the launch traces a permuted beam). Every ray still writes its own index, so results must be bit-identical to
tracing the same rays in index order, which rays nodes of fewer than 4096 rays keep. The checks split one sensor
into halves of 4095 rays (index order) and compare them with the whole sensor (direction order), solo and batched,
with masks, ring ids, per-ray ranges, time offsets with distortion, and compaction."""
import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S
from tests.helpers import Case, Graph

pytestmark = pytest.mark.gpu

XYZ, IS_HIT, RAY_IDX, ENT, DIST, AZ, EL, RING, NORMAL = 1, 4, 6, 7, 8, 9, 10, 11, 22
FIELDS = [XYZ, DIST, ENT, NORMAL, IS_HIT, RAY_IDX, RING, AZ, EL]
HALF = 4095


def sensor(n=2 * HALF):
    rays = S.spinning_lidar(45, 182, -22.0, 12.0)[:n]  # 8190 rays, channel by channel
    rng = np.random.default_rng(3)
    case = Case(S.config1_scene(), rays, pose=S.f32(S.trs((1.5, -2.0, 1.8), (0, 0, 33.0), (1, 1, 1))),
                ranges=np.stack([rng.uniform(0.1, 2.0, n), rng.uniform(40.0, 100.0, n)], 1).astype(np.float32),
                mask=(rng.uniform(0, 1, n) > 0.1).astype(np.int8), ring_ids=np.arange(n, dtype=np.int32) % 45,
                time_offsets=np.linspace(0.0, 90.0, n, dtype=np.float32), distortion=True,
                velocity=((8.0, 1.0, 0.0), (0.0, 0.0, 0.4)), near=0.0, far=100.0)
    return case


def halves(case):
    out = []
    for k in range(2):
        s = slice(k * HALF, (k + 1) * HALF)
        out.append(Case(case.spec, case.rays[s], pose=case.pose, ranges=case.ranges[s], mask=case.mask[s],
                        ring_ids=case.ring_ids[s], time_offsets=case.time_offsets[s], distortion=True,
                        velocity=case.velocity, near=case.near, far=case.far))
    return out


@pytest.mark.parametrize("batched", [False, True])
def test_direction_order_gives_index_order_results(rgl_lib, batched):
    rgl = rgl_lib
    case = sensor()
    S.load_into_rgl(case.spec)
    whole = Graph(rgl, case, FIELDS)
    parts = [Graph(rgl, c, FIELDS) for c in halves(case)]
    if batched:
        rgl.graph_run_batch([whole.rays] + [p.rays for p in parts])
    else:
        for g in [whole] + parts:
            rgl.graph_run(g.rays)
    a = {f: rgl.result(whole.out, f) for f in FIELDS}
    b = [{f: rgl.result(p.out, f) for f in FIELDS} for p in parts]
    hits = int(a[IS_HIT].sum())
    assert 1000 < hits < len(case.rays)
    for f in FIELDS:
        if f == RAY_IDX:
            assert a[f].tolist() == list(range(len(case.rays)))
            continue
        joined = np.concatenate([b[0][f], b[1][f]])
        assert a[f].tobytes() == joined.tobytes(), f


def test_direction_order_with_compaction(rgl_lib):
    rgl = rgl_lib
    case = sensor()
    S.load_into_rgl(case.spec)
    whole = Graph(rgl, case, FIELDS, compact=True)
    parts = [Graph(rgl, c, FIELDS, compact=True) for c in halves(case)]
    rgl.graph_run_batch([whole.rays] + [p.rays for p in parts])
    a = {f: rgl.result(whole.out, f) for f in FIELDS}
    b = [{f: rgl.result(p.out, f) for f in FIELDS} for p in parts]
    idx1 = b[1][RAY_IDX].astype(np.int64) + HALF  # the second half numbers its rays from 0
    assert a[RAY_IDX].tolist() == np.concatenate([b[0][RAY_IDX], idx1]).tolist()
    for f in (XYZ, DIST, ENT, NORMAL, RING, AZ, EL):
        assert a[f].tobytes() == np.concatenate([b[0][f], b[1][f]]).tobytes(), f
