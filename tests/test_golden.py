"""The oracle and the library's host-side set-up against the frozen vectors
of tests/golden/oracle_small.npz (made by tests/golden/make_golden.py): the
oracle must reproduce them to 1e-13 (ran1 exactly), and the library's host
initial-field generators (no GPU needed) must match them."""
import os

import numpy as np

import speccodelib_amd as scl
from tests.golden import make_golden

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_small.npz")


def test_oracle_reproduces_golden_vectors():
    g = np.load(GOLD)
    new = make_golden.build()
    assert set(g.files) == set(new)
    for k in g.files:
        if k.startswith("ran1"):
            assert np.array_equal(g[k], new[k]), k
        else:
            a, b = np.asarray(new[k]), g[k]
            assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max(), k


def test_host_initial_fields_match_golden_vectors():
    g = np.load(GOLD)
    ic = np.stack(scl.initialize(16, 16, 16, -3, 4.0, 1.0, scl.DP))
    assert np.abs(ic - g["ic16"]).max() <= 1e-14 * np.abs(g["ic16"]).max()
