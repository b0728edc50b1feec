"""The oracle against its own regression fixtures (tests/golden/
oracle_fixtures.npz, written by tests/golden/make_oracle_fixtures.py): the
reconstructions of every supported order, ConstructL (2D Euler fifth and
seventh order with global_alpha, 1D advection) and three controlled SSP-RK3
steps, recomputed from the stored inputs and compared bit for bit."""
import os

import numpy as np
import pytest

import orc

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "oracle_fixtures.npz")


@pytest.fixture(scope="module")
def current():
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_oracle_fixtures
    return make_oracle_fixtures.fixtures()


def test_oracle_reproduces_its_fixtures(current):
    stored = np.load(FIX, allow_pickle=False)
    assert set(stored.files) == set(current)
    for k in stored.files:
        assert np.array_equal(stored[k], current[k]), k
