"""Scene updates while batches are in flight. rgl_entity_set_transform, rgl_entity_set_id and
rgl_entity_set_laser_retro change host state only and return without waiting for running graphs (the reference
drains every graph first, src/api/apiCore.cpp:333-425); the scene's next rebuild runs on its own stream after the
raytrace launches that still read its device arrays, from pinned staging arrays used in turn. Frames are enqueued
back to back with entities moved between them and nothing read until the last frame is enqueued; every frame's
results must equal those of the same graphs run alone, one at a time, under that frame's scene state, with the TLAS
rebuilt and refitted."""
import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S
from tests.test_gpu_batch import assert_same, build_world, make_set, second_scene_spec

pytestmark = pytest.mark.gpu

FRAMES = 4
INSTANCES = (len(S.config1_scene().instances), len(second_scene_spec().instances))


def frame_updates(rng, frame):
    """{(scene, instance, kind): value} updates of one frame: a few moved instances per scene, one new entity id
    (second scene) and one laser retro (default scene)."""
    ups = {}
    for si, count in enumerate(INSTANCES):
        for i in rng.choice(count, 3, replace=False):
            m = S.trs((rng.uniform(-6, 6), rng.uniform(-6, 6), rng.uniform(-0.5, 2.0)), rng.uniform(0, 360, 3),
                      rng.uniform(0.5, 2.0, 3))
            ups[(si, int(i), "transform")] = S.f32(m)
    ups[(1, 1 + frame, "id")] = 5000 + frame
    ups[(0, 2 + frame, "retro")] = 0.25 * (frame + 1)
    return ups


def initial_value(specs, key):
    si, i, kind = key
    inst = specs[si].instances[i]
    if kind == "transform":
        return S.f32(inst.transform)
    if kind == "id":
        return inst.entity_id
    return float(inst.laser_retro)


def apply(rgl, ents, state):
    for (si, i, kind), v in state.items():
        e = ents[si][i]
        if kind == "transform":
            rgl.entity_set_transform(e, v)
        elif kind == "id":
            rgl.check("rgl_entity_set_id", e, int(v))
        else:
            rgl.check("rgl_entity_set_laser_retro", e, float(v))


@pytest.mark.parametrize("refit", [False, True])
def test_updates_between_in_flight_batches_equal_solo_runs(rgl_lib, refit):
    rgl = rgl_lib
    specs, scenes, ents = build_world(rgl)
    assert tuple(len(e) for e in ents) == INSTANCES
    for sc in scenes:
        rgl.check("rgl_scene_set_tlas_refit", sc, refit)
    rng = np.random.default_rng(31 + int(refit))
    updates = [frame_updates(rng, f) for f in range(FRAMES)]
    # every touched value per frame: the frame's own scene state, restorable after later frames changed it
    current = {k: initial_value(specs, k) for u in updates for k in u}
    states = []
    for u in updates:
        current.update(u)
        states.append(dict(current))
    sets = [make_set(rgl, scenes) for _ in range(FRAMES)]  # one graph set per frame in flight
    for f in range(FRAMES):
        apply(rgl, ents, updates[f])  # the previous frames' batches may still run
        rgl.graph_run_batch([g.rays for g in sets[f]])
    got = [[g.results() for g in sets[f]] for f in range(FRAMES)]
    hits = 0
    for f in range(FRAMES):
        apply(rgl, ents, states[f])
        solo = make_set(rgl, scenes)  # fresh graphs: first run of every noise stream, as in the batch
        for g in solo:
            rgl.graph_run(g.rays)
            hits += int(g.results()[4].sum())
        for k, g in enumerate(solo):
            assert_same(g.results(), got[f][k], f"frame {f} sensor {k} refit {refit}")
    assert hits > 4000
    print(f"{FRAMES} frames enqueued back to back with {sum(len(u) for u in updates)} updates between them; "
          f"every result equals a solo run (refit {refit})", flush=True)
