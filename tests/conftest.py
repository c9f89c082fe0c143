import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "treewaveoldg.jl_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def scrambled(n, seed=0):
    """Deterministic rough data, as the reference's tests use closed forms
    (refined_tests.jl:19-21): a fixed-seed normal sample."""
    import numpy as np
    return np.random.default_rng(seed).standard_normal(n)


@pytest.fixture
def rough():
    return scrambled
