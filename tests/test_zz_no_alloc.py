"""fp_sim_count_nonfinite_noalloc (no device allocation; counters in the library's fixed device array)
against fp_sim_count_nonfinite on the same states: two calls in a row, the second on a smaller
simulation, so counters left by the first call would show if the second did not clear them."""
import numpy as np
import pytest

from feriphys_amd import workloads as wl

pytestmark = pytest.mark.gpu


def plant_nan(sim, mesh, rows):
    rec = sim.get_points(mesh)
    rec[rows, 1 + (np.arange(len(rows)) % 6)] = np.nan
    sim.set_points(mesh, rec)


def test_count_nonfinite_noalloc_matches_allocating_entry_point():
    big = wl.build_gpu(wl.config2(n_cloths=12, smin=5, smax=40, steps=1))
    for m, k in ((0, 3), (4, 1), (7, 17), (11, 2)):
        n = big.mesh_counts(m)[0]
        plant_nan(big, m, list(range(0, n, max(1, n // k)))[:k])
    ref1 = big.count_nonfinite()
    got1 = big.count_nonfinite_noalloc()
    assert ref1.sum() > 0
    assert (got1 == ref1).all(), (got1, ref1)
    small = wl.build_gpu(wl.config2(n_cloths=3, smin=5, smax=20, steps=1, seed=4))
    plant_nan(small, 1, [0, 1])
    ref2 = small.count_nonfinite()
    got2 = small.count_nonfinite_noalloc()
    assert list(ref2) == [0, 2, 0]
    assert (got2 == ref2).all(), (got2, ref2)
    # and the first simulation again, after the second has used the counters
    assert (big.count_nonfinite_noalloc() == ref1).all()
