"""Point stages folded into the batched raytrace launch (rgl_graph_batch_get_fused_stages): a points transform and a
distance noise that follow a raytrace node are computed inside its launch. Every node's results (the raytrace node's,
each folded node's and everything downstream) must equal the same graph run alone, bit for bit and noise included,
frame after frame; a chain that branches stops folding at the branch."""
import ctypes as C

import numpy as np
import pytest

from robotecgpulidar_amd import scenes as S
from tests.test_gpu_batch import assert_same

pytestmark = pytest.mark.gpu

XYZ, IS_HIT, DIST, ENT = 1, 4, 8, 7


class Chain:
    """rays -> pose -> range -> raytrace -> points transform -> distance noise -> compact -> yield; with
    `branch`, the transform has a second output (a yield), so only the transform can be folded; with `two_returns`,
    FIRST_LAST returns of undiverged beams (two return points per ray, both from the single sample)."""

    def __init__(self, rgl, seed, branch, two_returns=False):
        self.rgl = rgl
        n = rgl.Node
        rays = S.make_lidar3d_rays(360, 40, 1.0, 1.0)
        self.rays = rgl.node_rays_from_mat3x4f(n(), rays)
        pose = rgl.node_rays_transform(n(), S.f32(S.trs((1.0, -2.0, 1.5), (0, 0, 20))))
        rng = rgl.node_rays_set_range(n(), np.array([[0.3, 120.0]], np.float32))
        self.rt = rgl.node_raytrace(n())
        if two_returns:
            rgl.check("rgl_node_raytrace_configure_return_mode", self.rt.h, rgl.RGL_RETURN_FIRST_LAST)
        self.tf = rgl.node_points_transform(n(), S.f32(S.trs((0.25, 0.5, -1.0), (0, 5, 30))))
        self.noise = n()
        rgl.check("rgl_node_gaussian_noise_distance", self.noise.ref, 0.01, 0.02, 0.001)
        comp = rgl.node_points_compact_by_field(n(), IS_HIT)
        self.out = rgl.node_points_yield(n(), [XYZ, DIST, ENT])
        rgl.chain(self.rays, pose, rng, self.rt, self.tf, self.noise, comp, self.out)
        if branch:
            self.side = rgl.node_points_yield(n(), [XYZ, DIST])
            rgl.add_child(self.tf, self.side)
        rgl.check("rgl_node_gaussian_noise_set_seed", self.noise.h, C.c_uint64(0xF05E0000 + seed))

    def results(self):
        r = {}
        for name, node in (("rt", self.rt), ("tf", self.tf), ("noise", self.noise)):
            for f in (XYZ, DIST):
                r[(name, f)] = self.rgl.result(node, f)
        for f in (XYZ, DIST, ENT):
            r[("out", f)] = self.rgl.result(self.out, f)
        return r


@pytest.mark.parametrize("branch,two_returns", [(False, False), (True, False), (False, True)])
def test_folded_stages_equal_separate_launches(rgl_lib, branch, two_returns):
    rgl = rgl_lib
    spec = S.config1_scene()
    _, ents, _ = S.load_into_rgl(spec)
    solo = [Chain(rgl, k, branch, two_returns) for k in range(3)]
    batched = [Chain(rgl, k, branch, two_returns) for k in range(3)]
    for frame in range(3):
        if frame:
            rgl.entity_set_transform(ents[2], S.f32(S.trs((1.0 * frame, 0.5, 0.0), (0, 0, 12 * frame), (1, 1, 1))))
        for g in solo:
            rgl.graph_run(g.rays)
        f0 = rgl.batch_fused_stages()
        rgl.graph_run_batch([g.rays for g in batched])
        fused = rgl.batch_fused_stages() - f0
        assert fused == len(batched) * (1 if branch else 2), fused
        for k, (a, b) in enumerate(zip(solo, batched)):
            ra, rb = a.results(), b.results()
            assert_same(ra, rb, f"frame {frame} chain {k}")
            if branch:
                for f in (XYZ, DIST):
                    assert rgl.result(a.side, f).tobytes() == rgl.result(b.side, f).tobytes()
        kept = sum(len(g.results()[("out", DIST)]) for g in batched)
        moved = np.abs(batched[0].results()[("noise", DIST)] - batched[0].results()[("tf", DIST)])
        moved = moved[np.isfinite(moved)]  # non-hits sit at an infinite distance
        print(f"frame {frame}: folded {fused} stages, {kept} kept points, mean |noise| {moved.mean():.4f}")
        assert kept > 1000 and moved.mean() > 0.005
