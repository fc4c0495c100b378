"""Serving (rgl_server_*): continuous batching must give every request exactly the results its graph gives when run
alone (noise included), whether it shares a batch, arrives while other batches run, waits behind an earlier frame
of the same graph, or has its scene evicted and rebuilt under the HBM budget; plain (unbatched) serving likewise."""
import ctypes as C

import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S
from tests.test_gpu_batch import SENSORS, SensorGraph, assert_same, build_world, make_set
from tests.test_radar import AZ_MAX, AZ_MIN, AZ_STEP, OMEGA, RANGE_MAX, RANGE_STEP

pytestmark = pytest.mark.gpu

XYZ, IS_HIT, DIST, AZ, RCS, INC = 1, 4, 8, 9, 19, 23


class RadarGraph:
    """256 x 64 radar over +-60 deg azimuth and +-10 deg elevation -> raytrace -> radar binning."""

    def __init__(self, rgl, scene, yaw):
        self.rgl = rgl
        az = np.linspace(-60.0, 60.0, 256)
        el = np.linspace(-10.0, 10.0, 64)
        rays = [S.mul(S.rotation_deg(0.0, 0.0, a), S.rotation_deg(0.0, 90.0 - e, 0.0)) for e in el for a in az]
        self.rays = rgl.node_rays_from_mat3x4f(rgl.Node(), S.f32(np.stack(rays)))
        pose = rgl.node_rays_transform(rgl.Node(), S.f32(S.trs((-20.0, 3.0, 1.0), (0, 0, yaw))))
        rt = rgl.node_raytrace(rgl.Node(), scene)
        self.out = rgl.node_points_radar_bin(rgl.Node(), RANGE_STEP, RANGE_MAX, AZ_STEP, AZ_MIN, AZ_MAX, OMEGA)
        rgl.chain(self.rays, pose, rt, self.out)
        self.noise = []

    def results(self):
        return {f: self.rgl.result(self.out, f) for f in (XYZ, DIST, AZ, RCS)}


def mixed_set(rgl, scenes):
    return make_set(rgl, scenes) + [RadarGraph(rgl, scenes[0], 15.0), RadarGraph(rgl, scenes[1], -40.0)]


def solo_frames(rgl, graphs, frames):
    out = []
    for _ in range(frames):
        for g in graphs:
            rgl.graph_run(g.rays)
        out.append([g.results() for g in graphs])
    return out


@pytest.mark.parametrize("batching", [True, False])
def test_served_results_equal_solo(rgl_lib, batching):
    rgl = rgl_lib
    _, scenes, _ = build_world(rgl)
    ref, served = mixed_set(rgl, scenes), mixed_set(rgl, scenes)
    expected = solo_frames(rgl, ref, 3)
    server = rgl.Server(slots=3, max_batch_graphs=4, batching=batching)
    for f in range(3):
        tickets = [server.submit(g.rays) for g in served]
        for t, g, exp in zip(tickets, served, expected[f]):
            server.wait(t)
        for k, (g, exp) in enumerate(zip(served, expected[f])):
            assert_same(g.results(), exp, f"frame {f} graph {k}")
    st = server.stats()
    print(f"batching={batching}: {st}")
    assert st["completed"] == 3 * len(served)
    assert st["batches"] >= (3 * len(served) + 3) // 4 if batching else st["batches"] == 3 * len(served)
    server.destroy()


def test_requests_arriving_during_flight_overlap_and_match(rgl_lib):
    rgl = rgl_lib
    _, scenes, _ = build_world(rgl)
    ref, served = make_set(rgl, scenes), make_set(rgl, scenes)
    expected = solo_frames(rgl, ref, 2)
    server = rgl.Server(slots=2, max_batch_graphs=64)
    # Frame 0 arrives one request at a time, each followed by a step: later requests find earlier ones in flight.
    tickets = []
    for g in served:
        tickets.append(server.submit(g.rays))
        server.step()
    # Frame 1 of every graph is queued while frame 0 may still run: it must wait for its own graph's frame 0.
    for t in tickets:
        server.wait(t)
    for k, (g, exp) in enumerate(zip(served, expected[0])):
        assert_same(g.results(), exp, f"frame 0 graph {k}")
    second = [server.submit(g.rays) for g in served]
    third_first = server.submit(served[0].rays)  # same graph twice in the queue: runs after its frame 1
    for t in second:
        server.wait(t)
    for k, (g, exp) in enumerate(zip(served[1:], expected[1][1:])):
        assert_same(g.results(), exp, f"frame 1 graph {k + 1}")
    server.wait(third_first)
    assert server.poll(third_first)[0] == rgl.Server.DONE
    st = server.stats()
    print(f"overlap: {st}")
    assert st["batches"] >= 3
    assert st["deferred"] >= 1
    server.destroy()


def test_hbm_budget_evicts_and_rebuilds(rgl_lib):
    rgl = rgl_lib
    _, scenes, _ = build_world(rgl)
    served = make_set(rgl, scenes)
    for g in served:  # frame 0: builds every acceleration structure, so the sizes below are the built ones
        rgl.graph_run(g.rays)
    # Learn each scene's structure size, then pick a budget that holds the larger one but not both:
    # alternating scenes forces evictions.
    probe = rgl.Server(slots=1, max_batch_graphs=1)
    sizes = []
    for g in (served[0], served[1]):  # sensors 0 and 1 raytrace scenes 0 and 1 (frame 1)
        probe.wait(probe.submit(g.rays))
        sizes.append(probe.stats()["resident_bytes"] - sum(sizes))
    probe.destroy()
    for g in served[2:]:  # frame 1 of the others
        rgl.graph_run(g.rays)
    expected = solo_frames(rgl, make_set(rgl, scenes), 4)[2:]  # frames 2 and 3 of identically seeded graphs
    budget = max(sizes) + min(sizes) // 2
    server = rgl.Server(slots=1, max_batch_graphs=1, hbm_budget=budget)
    for f in range(2):
        for k, g in enumerate(served):
            t = server.submit(g.rays)
            server.wait(t)
            assert_same(g.results(), expected[f][k], f"frame {f} graph {k}")
    st = server.stats()
    print(f"budget {budget} (scenes {sizes}): {st}")
    assert st["evictions"] >= 2
    assert st["resident_bytes"] <= budget
    tiny = rgl.Server(slots=1, max_batch_graphs=1, hbm_budget=min(sizes) // 2)
    t = tiny.submit(served[0].rays)
    tiny.step()
    assert tiny.poll(t)[0] == rgl.Server.FAILED and tiny.stats()["failed"] == 1
    tiny.destroy()
    server.destroy()


def test_server_latency_stats_and_failures(rgl_lib):
    rgl = rgl_lib
    _, scenes, _ = build_world(rgl)
    graphs = make_set(rgl, scenes)
    server = rgl.Server(slots=2, max_batch_graphs=8)
    tickets = [server.submit(g.rays) for g in graphs]
    lat = [server.wait(t) for t in tickets]
    st = server.stats()
    assert st["completed"] == len(graphs) and st["p99_ms"] >= st["p50_ms"] > 0
    assert st["rays_completed"] == sum(len(S.spinning_lidar(ch, az, -20.0, 10.0)) if ch > 1 else 1
                                       for ch, az, *_ in SENSORS)
    assert max(lat) <= st["max_ms"] + 1e-9
    server.reset_stats()
    assert server.stats()["completed"] == 0
    assert rgl.call("rgl_server_poll", server.h, C.c_int64(10 ** 9), C.byref(C.c_int32()),
                    C.byref(C.c_double())) == rgl.RGL_INVALID_ARGUMENT
    server.destroy()
