"""The batched workload at its full size: 4,096 cloths, sides log-uniform in [8, 256] (19.75M
vertices), each with its own Config (one reference Simulation per cloth, springy/simulation.rs:
27-45), stepped from captured HIP graphs of 50 steps with the exact kernels."""
import numpy as np
import pytest

import feriphys_amd as fa
import pyoracle as orc
from feriphys_amd import workloads as wl
from gpu_common import announce, bits

pytestmark = pytest.mark.gpu

STEPS = 100


@pytest.fixture(scope="module")
def batch():
    sc = wl.config2(steps=STEPS)
    announce(f"config 2 full size: {len(sc.cloths)} cloths, {sum(c.rows * c.cols for c in sc.cloths)} vertices")
    g = wl.build_gpu(sc)
    g.set_kernel(fa.FP_KERNEL_AUTO, fa.FP_MATH_EXACT)
    before = [g.get_points(k) for k in range(len(sc.cloths))]
    g.step_graph(STEPS, 50)
    after = [g.get_points(k) for k in range(len(sc.cloths))]
    announce("config 2 full size: 100 graph-replayed steps done")
    return sc, before, after


def test_config2_finite_and_pins_held(batch):
    sc, before, after = batch
    for k, c in enumerate(sc.cloths):
        assert np.isfinite(after[k]).all(), k
        p = c.pins
        assert (bits(after[k][p, :7]) == bits(before[k][p, :7])).all(), k


def test_config2_graph_replay_equals_eager(batch):
    """A second simulation of the same batch stepped eagerly, one launch sequence per step: every
    cloth ends with the same bits as the graph-replayed run."""
    sc, _, after = batch
    e = wl.build_gpu(sc)
    e.set_kernel(fa.FP_KERNEL_AUTO, fa.FP_MATH_EXACT)
    e.step(STEPS)
    rng = np.random.default_rng(5)
    for k in sorted(rng.choice(len(sc.cloths), 256, replace=False)):
        assert (bits(e.get_points(int(k))) == bits(after[int(k)])).all(), int(k)


def test_config2_sampled_cloths_match_oracle(batch):
    """Eight cloths drawn from the batch, each against its own oracle Simulation: bit-identical."""
    sc, _, after = batch
    rng = np.random.default_rng(9)
    picks = sorted(int(k) for k in rng.choice(len(sc.cloths), 8, replace=False))
    sub = wl.Scene([sc.cloths[k] for k in picks], sc.config, [], sc.jitter, sc.seed, STEPS)
    sub.mesh_winds = [sc.mesh_winds[k] for k in picks]
    sims = orc.build_scene_per_mesh(sub)
    for j, k in enumerate(picks):
        # the oracle's per-cloth jitter seed is seed + index within its scene: use the batch's index
        c = sc.cloths[k]
        m = orc.Mesh.cloth(*c.args(), plane=c.plane)
        rec = wl.jitter_records(m.points(), c.pins, sc.jitter, sc.seed + k)
        sims[j].mesh(0).set_points(rec)
        sims[j].step(STEPS)
        ref = sims[j].mesh(0).points()
        announce(f"cloth {k} ({c.rows}x{c.cols}, pins {len(c.pins)}): identical "
                 f"{(bits(ref) == bits(after[k])).mean():.4f}")
        assert (bits(ref) == bits(after[k])).all(), k
