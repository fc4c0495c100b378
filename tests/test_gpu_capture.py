"""HIP graph capture of batches (rgl_graph_batch_configure_capture): a batch whose kernel sequence repeats is
captured once and then replayed. Every replayed frame must equal the same graphs run alone, bit for bit and noise
included, while entity poses (hence scene structures), noise streams and node buffers change between frames and
eager work (solo graph runs, result reads) runs between the replays; alternating batch compositions each get
their own graph."""
import pytest

from robotecgpulidar_amd import scenes as S
from tests.test_gpu_batch import IS_HIT, assert_same, build_world, make_set

pytestmark = pytest.mark.gpu


@pytest.fixture
def lib(rgl_lib):
    rgl_lib.check("rgl_graph_batch_force_variants", 0, 0)  # no tuning launches, so every batch may be captured
    rgl_lib.batch_capture(True)
    yield rgl_lib
    rgl_lib.batch_capture(False)
    rgl_lib.check("rgl_graph_batch_configure_autotune", True)


def move(rgl, ents, frame):
    rgl.entity_set_transform(ents[0][3], S.f32(S.trs((1.5 * frame, 1.0, 0.0), (0, 0, 11 * frame), (1, 1, 1))))
    rgl.entity_set_transform(ents[1][5], S.f32(S.trs((-3.0, 2.0 * frame, 0.5), (7 * frame, 0, 0), (2, 2, 2))))


def test_replays_equal_solo(lib):
    rgl = lib
    _, scenes, ents = build_world(rgl)
    solo, batched = make_set(rgl, scenes), make_set(rgl, scenes)
    c0, r0 = rgl.batch_capture()
    frames = 7
    for frame in range(frames):
        if frame in (3, 4, 6):  # scene updates between replays: structures are rebuilt, descriptors change
            move(rgl, ents, frame)
        for g in solo:
            rgl.graph_run(g.rays)
        rgl.graph_run_batch([g.rays for g in batched])
        for k, (a, b) in enumerate(zip(solo, batched)):
            assert_same(a.results(), b.results(), f"frame {frame} sensor {k}")
            assert a.noise_runs() == b.noise_runs() == [frame + 1] * len(a.noise)
        assert sum(int(g.results()[IS_HIT].sum()) for g in batched) > 1000
    c1, r1 = rgl.batch_capture()
    print(f"captures {c1 - c0}, replays {r1 - r0}")
    assert c1 - c0 == 1 and r1 - r0 == frames - 2


def test_alternating_compositions_and_disable(lib):
    rgl = lib
    _, scenes, ents = build_world(rgl)
    solo, batched = make_set(rgl, scenes), make_set(rgl, scenes)
    halves = [list(range(0, len(solo), 2)), list(range(1, len(solo), 2))]
    c0, r0 = rgl.batch_capture()
    rounds = 4
    for frame in range(rounds):
        for part in halves:
            for k in part:
                rgl.graph_run(solo[k].rays)
            rgl.graph_run_batch([batched[k].rays for k in part])
            for k in part:
                assert_same(solo[k].results(), batched[k].results(), f"frame {frame} sensor {k}")
        if frame == 1:
            move(rgl, ents, 2)
    c1, r1 = rgl.batch_capture()
    print(f"captures {c1 - c0}, replays {r1 - r0}")
    assert c1 - c0 == 2 and r1 - r0 == 2 * (rounds - 2)
    rgl.batch_capture(False)
    for frame in range(3):
        rgl.graph_run_batch([g.rays for g in batched])
        for g in solo:
            rgl.graph_run(g.rays)
        for k, (a, b) in enumerate(zip(solo, batched)):
            assert_same(a.results(), b.results(), f"uncaptured frame {frame} sensor {k}")
    assert rgl.batch_capture() == (c1, r1)
