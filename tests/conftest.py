import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def bits_equal(a, b):
    """Bit-exact equality of two arrays (floats compared by their bit patterns)."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    if a.dtype.kind == "f":
        # NaN payloads are not part of the contract: any NaN equals any NaN
        na, nb = np.isnan(a), np.isnan(b)
        if not np.array_equal(na, nb):
            return False
        ua = a.view(np.uint32 if a.dtype.itemsize == 4 else np.uint64)
        ub = b.view(np.uint32 if b.dtype.itemsize == 4 else np.uint64)
        return np.array_equal(ua[~na], ub[~nb])
    return np.array_equal(a, b)


def assert_same_surface(got, want, label=""):
    """The HIP surface equals the oracle's bit for bit: triangle count and
    order, connections, coordinates and normals."""
    for key in ("coords", "connections", "normals"):
        g, w = got[key], want[key]
        assert g.shape == w.shape, f"{label} {key}: shape {g.shape} != {w.shape}"
        if not bits_equal(g, w):
            if g.dtype.kind == "f":
                diff = np.flatnonzero((g.view(np.uint32) != w.view(np.uint32)) & ~(np.isnan(g) & np.isnan(w)))
            else:
                diff = np.flatnonzero(g != w)
            i = int(diff[0])
            raise AssertionError(f"{label} {key}: {diff.size} entries differ, first at {i}: {g[i]!r} != {w[i]!r}")
    assert got["n_triangles"] == want["n_triangles"], label
    assert got["normal_type"] == want["normal_type"], label
    assert got["polygon_type"] == want["polygon_type"], label
    assert got["color_type"] == want["color_type"], label
