"""kvs::SlicePlane on the device (kmc_slice_plane) against the reference's own
output (tests/golden/slice.npz) and the oracle, bit for bit: coordinates,
negated face normals and per-vertex colours, in the reference's triangle
order."""
import numpy as np
import pytest

import oracle
from kvs_amd import datasets, kmc
from tests import golden_io
from tests.conftest import bits_equal

pytestmark = pytest.mark.gpu


def table_of(kind):
    name = ("BrewerRdBu", "BrewerSingleHueGreys")[kind]
    return np.array(golden_io.meta()["colormap_tables"][f"{name}_256"], dtype=np.uint8)


def same(got, want, label):
    assert got["n_triangles"] == want["normals"].size // 3, label
    assert bits_equal(got["coords"], want["coords"]), label
    assert bits_equal(got["normals"], want["normals"]), label
    assert np.array_equal(got["colors"], want["colors"]), label


def test_slice_plane_golden():
    for case, v, (co, no, cl) in golden_io.slice_cases():
        mm = tuple(case["min_max"]) if case["min_max"] is not None else None
        got = kmc.slice_plane(v, case["coefficients"], table_of(case["tf_kind"]), min_max=mm, coords=case["coords"])
        label = f"{case['dtype']} {case['shape']} {case['plane']}"
        print("case", label, flush=True)
        same(got, {"coords": co, "normals": no, "colors": cl}, label)
        assert got["color_type"] == 0 and got["normal_type"] == 1 and got["polygon_type"] == 3


@pytest.mark.parametrize("nx", [63, 64, 65, 300, 1024, 2049])
def test_slice_plane_row_widths(nx):
    rng = np.random.default_rng(nx)
    v = datasets.turbulence((nx, 9, 7), seed=nx, period=max(nx, 64))
    for t in range(3):
        coef = [float(np.float32(x)) for x in rng.standard_normal(3)] + [float(np.float32(rng.uniform(-nx, 0)))]
        table = table_of(t % 2)
        label = f"nx={nx} plane {coef}"
        print("case", label, flush=True)
        same(kmc.slice_plane(v, coef, table), oracle.slice_plane(v, coef, table), label)


def test_slice_plane_blobs_128_oblique():
    v = datasets.blobs(128, seed=4, nblobs=5)
    table = table_of(0)
    coef = [0.31, -0.52, 0.79, -20.0]
    got = kmc.slice_plane(v, coef, table)
    same(got, oracle.slice_plane(v, coef, table), "blobs 128^3 oblique")
    assert got["n_triangles"] > 10000


def test_slice_plane_explicit_color_range_and_device_outputs():
    v = datasets.ct_like((40, 35, 30), seed=2, nellipsoids=6, dtype=np.uint16)
    table = table_of(1)
    coef = [0.0, 1.0, 0.2, -17.25]
    want = oracle.slice_plane(v, coef, table, color_range=(1000.0, 50000.0))
    same(kmc.slice_plane(v, coef, table, color_range=(1000.0, 50000.0)), want, "explicit range")
    same(kmc.slice_plane(v, coef, table, color_range=(1000.0, 50000.0), out_memory=kmc.MEMORY_DEVICE), want,
         "device outputs")


def test_slice_plane_missing_the_volume_is_empty():
    v = datasets.blobs(20, seed=1)
    got = kmc.slice_plane(v, [1.0, 0.0, 0.0, 100.0], table_of(0))
    assert got["n_triangles"] == 0 and got["coords"].size == 0 and got["colors"].size == 0


def test_slice_plane_both_device_paths(monkeypatch):
    # the threshold path (a bisection per node row, runs of cube indices per
    # cell row) and the classification of every node give the same surface;
    # axis-aligned, tiny-c0, negative-c0 and oblique planes, and planes whose
    # values overflow to infinity at the volume's corners (classification path)
    v = datasets.turbulence((70, 33, 29), seed=8, period=64)
    table = table_of(0)
    planes = [[0.0, 0.0, 1.0, -14.5], [0.0, 1.0, 0.0, -10.0], [1.0, 0.0, 0.0, -35.0], [1e-7, 1.0, 0.5, -20.0],
              [-1.3, 0.2, 0.7, 40.0], [0.3, -0.5, 0.8, -3.0], [2.0, 3.0, -1.0, -60.0], [3e38, 3e38, 0.0, -1.0]]
    for coef in planes:
        coef = [float(np.float32(c)) for c in coef]
        want = oracle.slice_plane(v, coef, table)
        monkeypatch.delenv("KMC_SLICE", raising=False)
        fast = kmc.slice_plane(v, coef, table)
        monkeypatch.setenv("KMC_SLICE", "classify")
        full = kmc.slice_plane(v, coef, table)
        print("case plane", coef, want["n_triangles"], flush=True)
        same(fast, want, f"threshold path {coef}")
        same(full, want, f"classification path {coef}")
    assert want["n_triangles"] >= 0


@pytest.mark.parametrize("res", [1, 2, 1024, 1025, 4000])
def test_slice_plane_color_map_resolutions(monkeypatch, res):
    # the threshold path keeps colour maps of up to 1024 entries in LDS and
    # reads larger ones from device memory; both paths against the reference's
    # own kvs::SlicePlane with a transfer function of that resolution
    v = datasets.turbulence((40, 30, 20), seed=3, period=32)
    coef = [0.3, -0.5, 0.8, -3.0]
    for kind in (0, 1):
        table = np.asarray(oracle.ref_tf_table(kind, res), dtype=np.uint8)
        want = oracle.ref_slice_plane(v, coef, tf_kind=kind, tf_resolution=res)
        print("case colour map", res, kind, flush=True)
        monkeypatch.delenv("KMC_SLICE", raising=False)
        same(kmc.slice_plane(v, coef, table), want, f"threshold path {res} {kind}")
        monkeypatch.setenv("KMC_SLICE", "classify")
        same(kmc.slice_plane(v, coef, table), want, f"classification path {res} {kind}")
