"""The exact GPU kernels against the committed golden digests of the 32 x 32 cloth workload
(tests/golden/config0.json): one accumulate_forces, 1 and 10 steps with the corner pins, 600 steps
with the top row pinned.  No oracle is run here."""
import json
import os
import sys

import numpy as np
import pytest

import feriphys_amd as fa
from feriphys_amd import workloads as wl
from gpu_common import announce

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

GOLD = json.load(open(os.path.join(HERE, "golden", "config0.json")))


def sim_of(sc):
    g = wl.build_gpu(sc)
    g.set_kernel(fa.FP_KERNEL_AUTO, fa.FP_MATH_EXACT)
    return g


def test_forces_and_steps_hash_like_golden():
    g = sim_of(wl.config0(steps=10))
    assert mg.digest(g.get_points(0).astype(np.float32)) == GOLD["initial_records"]
    g.accumulate_forces(0)
    f = g.get_points(0)[:, 7:10].astype(np.float32)
    assert mg.digest(f) == GOLD["forces_1"], f[100]
    for n in (1, 10):
        g = sim_of(wl.config0(steps=n))
        g.step(n)
        assert mg.digest(g.get_points(0).astype(np.float32)) == GOLD[f"records_{n}"], n
    g = sim_of(wl.config0(steps=600, pins="edge"))
    g.step(600)
    r = g.get_points(0).astype(np.float32)
    announce(f"golden: 600-step record 0 {r[0].tolist()}")
    assert mg.digest(r) == GOLD["records_600_edge"]
