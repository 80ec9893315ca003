"""The unstructured branch of kvs::Isosurface (MarchingTetrahedra /
Hexahedra / Pyramid / Prism) and kvs::SlicePlane over unstructured cells on the
device, against the reference's own output (tests/golden/unstructured.npz) and
the oracle, bit for bit: triangle order, corners, normals, per-vertex colours,
the mapper's types, and min/max values and coordinates in connection order."""
import numpy as np
import pytest

import oracle
from kvs_amd import datasets, kmc
from tests import golden_io
from tests.conftest import bits_equal

pytestmark = pytest.mark.gpu

CELLS = [kmc.TETRAHEDRA, kmc.HEXAHEDRA, kmc.PYRAMID, kmc.PRISM]
MODES = [(kmc.POLYGON_NORMAL, True), (kmc.POLYGON_NORMAL, False), (kmc.VERTEX_NORMAL, False)]


def table_of(kind):
    name = ("BrewerRdBu", "BrewerSingleHueGreys")[kind]
    return np.array(golden_io.meta()["colormap_tables"][f"{name}_256"], dtype=np.uint8)


def test_unstructured_golden():
    for case, v, co, k, outs, (sc, sn, scl) in golden_io.unstructured_cases():
        ct = case["cell_type"]
        mm = tuple(case["min_max"]) if case["min_max"] is not None else None
        label = f"cell {ct} {case['dtype']} {case['dims']}"
        print("case", label, flush=True)
        for (normal, dup), (want_c, want_n) in outs.items():
            got = kmc.extract_unstructured(v, co, k, ct, case["iso"], normal, dup, min_max=mm)
            assert bits_equal(got["coords"], want_c), (label, normal, dup)
            assert bits_equal(got["normals"], want_n), (label, normal, dup)
            meta = case[f"n{normal}_d{int(dup)}"]
            assert (got["polygon_type"], got["normal_type"], got["color_type"]) == \
                (meta["polygon_type"], meta["normal_type"], meta["color_type"]), (label, normal, dup)
            assert got["min_coord"] == case["min_obj"] and got["max_coord"] == case["max_obj"], label
            if mm is None:
                assert [got["min_value"], got["max_value"]] == case["volume_min_max"], label
        got = kmc.slice_plane_unstructured(v, co, k, ct, case["slice"]["coefficients"], table_of(case["slice"]["tf_kind"]),
                                           min_max=mm)
        assert bits_equal(got["coords"], sc) and bits_equal(got["normals"], sn), label + " slice"
        assert np.array_equal(got["colors"], scl), label + " slice"
        assert got["color_type"] == 0


@pytest.mark.parametrize("ct", CELLS)
def test_unstructured_larger_meshes(ct):
    v, c, k = datasets.unstructured_mesh((40, 33, 29), ct, seed=ct, nblobs=6)
    for iso in (0.3, 0.55):
        got = kmc.extract_unstructured(v, c, k, ct, iso)
        want = oracle.unstructured_isosurface(v, c, k, ct, iso)
        assert bits_equal(got["coords"], want["coords"]) and bits_equal(got["normals"], want["normals"]), (ct, iso)
        assert got["n_triangles"] > 1000
    coef = [0.21, -0.37, 0.9, -14.0]
    table = table_of(ct % 2)
    got = kmc.slice_plane_unstructured(v, c, k, ct, coef, table)
    want = oracle.unstructured_slice_plane(v, c, k, ct, coef, table)
    assert bits_equal(got["coords"], want["coords"]) and np.array_equal(got["colors"], want["colors"]), ct


@pytest.mark.parametrize("dtype", [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64,
                                   np.float32, np.float64])
def test_unstructured_all_value_types(dtype):
    for ct in CELLS:
        v, c, k = datasets.unstructured_mesh((9, 8, 7), ct, seed=int(np.dtype(dtype).itemsize) + ct, dtype=dtype)
        hi = float(v.max())
        iso = 0.4 * hi if np.dtype(dtype).kind == "f" else float(int(0.4 * hi)) + 0.5
        got = kmc.extract_unstructured(v, c, k, ct, iso)
        want = oracle.unstructured_isosurface(v, c, k, ct, iso)
        assert bits_equal(got["coords"], want["coords"]) and bits_equal(got["normals"], want["normals"]), (dtype, ct)
        coef = [0.3, 0.1, -0.6, 2.0]
        table = table_of(0)
        got = kmc.slice_plane_unstructured(v, c, k, ct, coef, table)
        want = oracle.unstructured_slice_plane(v, c, k, ct, coef, table)
        assert np.array_equal(got["colors"], want["colors"]) and bits_equal(got["coords"], want["coords"]), (dtype, ct)


def test_unstructured_errors_and_empty_cases():
    v, c, k = datasets.unstructured_mesh((5, 5, 5), kmc.TETRAHEDRA, seed=1)
    with pytest.raises(kmc.KmcError) as e:
        kmc.extract_unstructured(v, c, k, 3, 0.5)             # quadratic tetrahedra: no marching table
    assert e.value.status == kmc.ERR_CELL_TYPE
    flat = np.full_like(v, 0.5)
    got = kmc.extract_unstructured(flat, c, k, kmc.TETRAHEDRA, 0.5)   # Math::Equal(min, max)
    assert got["empty_range"] == 1 and got["n_triangles"] == 0
    got = kmc.extract_unstructured(v, c, k, kmc.TETRAHEDRA, 0.5, kmc.VERTEX_NORMAL, False)
    assert got["n_triangles"] == 0 and got["polygon_type"] == 5


def test_unstructured_device_outputs():
    v, c, k = datasets.unstructured_mesh((12, 11, 10), kmc.PRISM, seed=4)
    want = oracle.unstructured_isosurface(v, c, k, kmc.PRISM, 0.4)
    got = kmc.extract_unstructured(v, c, k, kmc.PRISM, 0.4, out_memory=kmc.MEMORY_DEVICE)
    assert bits_equal(got["coords"], want["coords"]) and bits_equal(got["normals"], want["normals"])
