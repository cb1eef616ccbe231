"""Shared fixtures.  Tests marked `gpu` need a GPU and the built HIP library;
the rest run on the CPU (oracle, fixtures, host logic, ABI exports)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X and the built HIP library")


@pytest.fixture(scope="session")
def ctx():
    import bonxai_amd as bx
    c = bx.Context(0)
    yield c
    c.close()


def sorted_cells(coords, vals=None):
    order = np.lexsort((coords[:, 2], coords[:, 1], coords[:, 0]))
    return coords[order], (vals[order] if vals is not None else None)


def assert_same_cells(gpu, ref, what=""):
    gc, gv = sorted_cells(*gpu)
    rc, rv = ref
    assert gc.shape == rc.shape, f"{what}: {len(gc)} cells on the GPU, {len(rc)} in the oracle"
    assert np.array_equal(gc, rc), f"{what}: allocated coordinate sets differ"
    if rv is not None:
        assert np.array_equal(gv, rv), f"{what}: cell values differ"
