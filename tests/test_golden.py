"""Committed golden digests of the 32 x 32 cloth workload (tests/golden/config0.json, written by
tests/golden/make_golden.py from the oracle): the oracle must still produce them, and the library's
host-side tables must hash the same."""
import json
import os
import sys

import numpy as np

import feriphys_amd as fa
from feriphys_amd import workloads as wl

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

GOLD = json.load(open(os.path.join(HERE, "golden", "config0.json")))


def test_oracle_reproduces_golden():
    assert mg.golden() == GOLD


def test_counts_match_closed_forms():
    # 32 x 32: tensile 2*32*31 = 1984, shear 31*31 = 961 (one diagonal per quad), bend 2*30*30 = 1800
    assert GOLD["counts"] == [1024, 1984 + 961 + 1800, 2 * 31 * 31, 2]


def test_library_tables_hash_like_golden():
    cl = wl.config0().cloths[0]
    m = fa.Cloth(*cl.args(), plane=cl.plane).mesh
    v01, kcl, _ = m.struts()
    s3, v3 = m.faces()
    assert mg.digest(np.asarray(v01, np.uint64)) == GOLD["struts_v01"]
    assert mg.digest(np.asarray(kcl, np.float32)) == GOLD["struts_kcl"]
    assert mg.digest(np.asarray(s3, np.uint64)) == GOLD["faces_s3"]
    assert mg.digest(np.asarray(v3, np.uint64)) == GOLD["faces_v3"]
    assert mg.digest(np.asarray(m.pins(), np.uint64)) == GOLD["pins"]
