"""The oracle (oracle/mc_oracle.c) pinned: against the golden vectors written
from the reference itself, against the compiled reference directly where it
is built here, and against closed forms the mathematics fixes."""
import numpy as np
import pytest

import oracle
from kvs_amd import datasets
from tests import golden_io
from tests.conftest import bits_equal

MODES = golden_io.MODES


def test_tables_match_reference_hash():
    m = golden_io.meta()["tables"]
    tri = np.full((256, 16), -1, dtype="<i4")
    for ci in range(256):
        n, edges = oracle.case_edges(ci)
        tri[ci] = edges
        assert n == (np.array(edges) >= 0).sum() // 3
    assert golden_io.sha(tri) == m["TriangleID_sha256"]
    vert = np.zeros((12, 2, 3), dtype="<i4")
    corner = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    ends = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
    for e, (a, b) in enumerate(ends):
        vert[e, 0], vert[e, 1] = corner[a], corner[b]
    assert golden_io.sha(vert) == m["VertexID_sha256"]


def test_device_table_words_match_reference_hash():
    """KMC_TABLE_WORDS of kvs_amd/csrc/mc_tables.h -- the case table the kernels
    use, in its packed form (triangle count in bits 60..63, nibble j = j-th
    edge id in table order) -- decoded back into TriangleID[256][16] and
    pinned by the reference table's SHA-256 and per-case triangle counts."""
    import os
    import re
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kvs_amd", "csrc", "mc_tables.h")
    text = open(path).read()
    body = text[text.index("#define KMC_TABLE_WORDS"):text.index("static const uint64_t kTriangles")]
    words = [int(w, 16) for w in re.findall(r"0x([0-9a-fA-F]{16})ull", body)]
    assert len(words) == 256
    tri = np.full((256, 16), -1, dtype="<i4")
    for ci, w in enumerate(words):
        n = w >> 60
        assert (w >> (12 * n)) & ((1 << (60 - 12 * n)) - 1) == 0, ci    # nothing past the last triangle
        tri[ci, :3 * n] = [(w >> (4 * j)) & 0xF for j in range(3 * n)]
    assert golden_io.sha(tri) == golden_io.meta()["tables"]["TriangleID_sha256"]
    counts = np.load(os.path.join(golden_io.GOLDEN, "tables.npz"))["triangle_counts"]
    assert np.array_equal(counts, [w >> 60 for w in words])


def test_table_topology_invariants():
    # every triangle edge of case ci is an edge whose corner bits differ, and
    # every such edge is used: a plausible transcription slip breaks one of these
    corner_of = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
    for ci in range(256):
        n, edges = oracle.case_edges(ci)
        used = set(e for e in edges if e >= 0)
        crossing = set(e for e, (a, b) in enumerate(corner_of) if ((ci >> a) & 1) != ((ci >> b) & 1))
        assert used == crossing, ci
        assert 1 <= n <= 5 or ci in (0, 255)


def test_brute_force_cubes_against_golden():
    for v, iso, ci, outs in golden_io.cube_cases():
        for (normal, dup), (co, cn, no) in outs.items():
            r = oracle.extract(v, iso, normal, dup)
            assert bits_equal(r["coords"], co), (v.dtype, ci, normal, dup)
            assert bits_equal(r["connections"], cn), (v.dtype, ci, normal, dup)
            assert bits_equal(r["normals"], no), (v.dtype, ci, normal, dup)


def test_random_volumes_against_golden():
    for i, (v, iso, mm, coords, outs) in enumerate(golden_io.random_cases()):
        for (normal, dup), (co, cn, no, _) in outs.items():
            r = oracle.extract(v, iso, normal, dup, min_max_values=mm, coords=coords)
            assert bits_equal(r["coords"], co), (i, normal, dup)
            assert bits_equal(r["connections"], cn), (i, normal, dup)
            assert bits_equal(r["normals"], no), (i, normal, dup)


def test_large_64bit_integers_against_golden():
    # int64 / uint64 values beyond 2^53 (several share one double), isolevels
    # equal to or one double away from a data value, a clamped range
    for name, v, iso, mm, outs in golden_io.bigint_cases():
        for (normal, dup), (co, cn, no) in outs.items():
            r = oracle.extract(v, iso, normal, dup, min_max_values=mm)
            assert bits_equal(r["coords"], co), (name, iso, normal, dup)
            assert bits_equal(r["connections"], cn), (name, iso, normal, dup)
            assert bits_equal(r["normals"], no), (name, iso, normal, dup)
            assert r["n_triangles"] > 0, name


def test_hydrogen_against_golden_digests():
    h = golden_io.hydrogen()
    meta = golden_io.meta()["hydrogen64"]
    assert golden_io.sha(h) == meta["sha256"]
    for key, d in meta["levels"].items():
        iso, mode = key.split("_n")
        normal, dup = int(mode[0]), bool(int(mode.split("_d")[1]))
        r = oracle.extract(h, float(iso), normal, dup)
        assert golden_io.sha(r["coords"]) == d["coords"], key
        assert golden_io.sha(r["connections"]) == d["connections"], key
        assert golden_io.sha(r["normals"]) == d["normals"], key
        assert r["normal_type"] == d["normal_type"], key


def test_min_max_corner_cases():
    for c in golden_io.meta()["minmax"]:
        a = np.array([float(x) for x in c["values"]], dtype=c["dtype"]).reshape(1, 1, -1)
        lo, hi = oracle.min_max(a)
        assert repr(lo) == c["min"] and repr(hi) == c["max"], c["name"]


@pytest.mark.skipif(not oracle.ref_available(), reason="reference library not built here")
def test_oracle_equals_reference_on_fresh_random_volumes():
    rng = np.random.default_rng(77)
    dtypes = [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64, np.float32, np.float64]
    for t in range(30):
        dt = dtypes[t % 10]
        shape = tuple(int(s) for s in rng.integers(2, 20, 3))
        if np.dtype(dt).kind == "f":
            v = rng.standard_normal(shape).astype(dt)
            iso = float(rng.standard_normal() * 0.3)
        else:
            v = rng.integers(0, 100, shape).astype(dt)
            iso = float(rng.integers(0, 100)) + 0.5 * (t % 2)
        for normal, dup in MODES:
            a = oracle.extract(v, iso, normal, dup)
            b = oracle.ref_marching_cubes(v, iso, normal, dup)
            for k in ("coords", "connections", "normals"):
                assert bits_equal(a[k], b[k]), (t, k, normal, dup)
            assert a["normal_type"] == b["normal_type"]


@pytest.mark.skipif(not oracle.ref_available(), reason="reference library not built here")
def test_oracle_equals_reference_blobs_128():
    v = datasets.blobs(96, seed=3, nblobs=6)
    for normal, dup in MODES:
        a = oracle.extract(v, 0.4, normal, dup)
        b = oracle.ref_marching_cubes(v, 0.4, normal, dup)
        for k in ("coords", "connections", "normals"):
            assert bits_equal(a[k], b[k]), (k, normal, dup)


def test_linear_ramp_closed_form():
    # f(x,y,z) = x on an n^3 grid, isolevel k + 0.5: a plane of quads at
    # x = k + 0.5 exactly, two triangles per crossed cell, all normals along x
    n = 9
    x = np.arange(n, dtype=np.float32)
    v = np.broadcast_to(x[None, None, :], (n, n, n)).copy()
    for k in range(n - 1):
        iso = k + 0.5
        r = oracle.extract(v, iso, oracle.POLYGON_NORMAL, True)
        c = r["coords"].reshape(-1, 3)
        assert r["n_triangles"] == 2 * (n - 1) * (n - 1)
        assert np.all(c[:, 0] == np.float32(iso))
        nrm = r["normals"].reshape(-1, 3)
        assert np.all(nrm[:, 1] == 0) and np.all(nrm[:, 2] == 0) and np.all(nrm[:, 0] != 0)
        rn = oracle.extract(v, iso, oracle.POLYGON_NORMAL, False)
        assert rn["coords"].size // 3 == n * n          # one vertex per x-edge crossing
        assert rn["n_triangles"] == r["n_triangles"]


def test_mode_invariants_on_blobs():
    v = datasets.blobs(40, seed=8, nblobs=4)
    dup = oracle.extract(v, 0.45, oracle.POLYGON_NORMAL, True)
    nod = oracle.extract(v, 0.45, oracle.POLYGON_NORMAL, False)
    vn = oracle.extract(v, 0.45, oracle.VERTEX_NORMAL, False)
    assert dup["n_triangles"] == nod["n_triangles"] == vn["n_triangles"]
    nv = nod["coords"].size // 3
    assert nod["connections"].max() < nv
    assert np.array_equal(nod["connections"], vn["connections"])
    assert bits_equal(nod["coords"], vn["coords"])
    # every vertex is used by some triangle and the duplicated positions are the
    # shared ones (probe-observed equality of the two orientations holds here)
    assert np.unique(nod["connections"]).size == nv
    shared = nod["coords"].reshape(-1, 3)[nod["connections"].reshape(-1)]
    assert np.allclose(shared.reshape(-1), dup["coords"], rtol=0, atol=1e-5)
    # vertex normals are the sums of the incident polygon normals
    pn = nod["normals"].reshape(-1, 3).astype(np.float64)
    acc = np.zeros((nv, 3))
    for j in range(3):
        np.add.at(acc, nod["connections"].reshape(-1, 3)[:, j], pn)
    assert np.allclose(acc, vn["normals"].reshape(-1, 3), rtol=1e-4, atol=1e-6)


def test_constant_volume_short_circuit():
    v = np.full((5, 5, 5), 7, dtype=np.uint8)
    r = oracle.extract(v, 7.0, oracle.POLYGON_NORMAL, True)
    assert r["empty_range"] == 1 and r["coords"].size == 0


def test_fp64_reevaluation_error_is_small():
    v = datasets.blobs(32, seed=5)
    r = oracle.extract(v, 0.5, oracle.VERTEX_NORMAL, False, want_f64=True)
    err = np.abs(r["coords"].astype(np.float64) - r["coords_f64"])
    scale = np.maximum(np.abs(r["coords_f64"]), 1.0)
    assert (err / scale).max() < 1e-6


def _slice_color_inputs(case, v):
    """The colour map the fixture's transfer function used and its range."""
    table = np.array(golden_io.meta()["colormap_tables"][("BrewerRdBu", "BrewerSingleHueGreys")[case["tf_kind"]] + "_256"],
                     dtype=np.uint8)
    mm = tuple(case["min_max"]) if case["min_max"] is not None else None
    return table, oracle.tf_range(v, mm)


def test_slice_plane_against_golden():
    # kvs::SlicePlane restated (oracle/slice_oracle.c) against the reference's
    # own output: coordinates, negated face normals and per-vertex colours
    for case, v, (co, no, cl) in golden_io.slice_cases():
        table, rng = _slice_color_inputs(case, v)
        r = oracle.slice_plane(v, case["coefficients"], table, color_range=rng, coords=case["coords"])
        label = (case["dtype"], case["plane"])
        assert bits_equal(r["coords"], co), label
        assert bits_equal(r["normals"], no), label
        assert np.array_equal(r["colors"], cl), label
        assert case["color_type"] == 0 and case["normal_type"] == 1 and case["polygon_type"] == 3


def test_slice_plane_axis_aligned_closed_form():
    # the plane z = k + 0.5 through an n^3 grid: one quad (two triangles) per
    # cell of the slab, every vertex at z = k + 0.5, every normal along -z or +z
    n = 7
    v = datasets.blobs(n, seed=2)
    table = np.arange(3 * 256, dtype=np.uint8)
    for k in range(n - 1):
        r = oracle.slice_plane(v, [0.0, 0.0, 1.0, -(k + 0.5)], table)
        c = r["coords"].reshape(-1, 3)
        assert r["n_triangles"] == 2 * (n - 1) * (n - 1)
        assert np.all(c[:, 2] == np.float32(k + 0.5))
        nrm = r["normals"].reshape(-1, 3)
        assert np.all(nrm[:, 0] == 0) and np.all(nrm[:, 1] == 0) and np.all(nrm[:, 2] != 0)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference library not built here")
def test_slice_plane_equals_reference_fresh():
    rng = np.random.default_rng(17)
    for t in range(12):
        v = datasets.turbulence((int(rng.integers(5, 30)), int(rng.integers(5, 30)), int(rng.integers(5, 30))), seed=t)
        coef = [float(np.float32(x)) for x in rng.standard_normal(4)]
        coef[3] *= 10
        for kind in (0, 1):
            a = oracle.slice_plane(v, coef, oracle.ref_tf_table(kind))
            b = oracle.ref_slice_plane(v, coef, tf_kind=kind)
            assert bits_equal(a["coords"], b["coords"]) and bits_equal(a["normals"], b["normals"]), (t, kind)
            assert np.array_equal(a["colors"], b["colors"]), (t, kind)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference library not built here")
def test_slice_plane_color_map_resolutions():
    # ColorMap::at with one entry (both neighbours the same), two, and tables
    # larger than the default 256
    v = datasets.turbulence((40, 30, 20), seed=3, period=32)
    coef = [0.3, -0.5, 0.8, -3.0]
    for res in (1, 2, 1024, 1025, 4000):
        for kind in (0, 1):
            a = oracle.slice_plane(v, coef, oracle.ref_tf_table(kind, res))
            b = oracle.ref_slice_plane(v, coef, tf_kind=kind, tf_resolution=res)
            assert a["n_triangles"] > 0
            assert bits_equal(a["coords"], b["coords"]) and np.array_equal(a["colors"], b["colors"]), (res, kind)


CELL_NAMES = ("tetrahedra", "hexahedra", "pyramid", "prism")
CELL_SHORT = ("Tet", "Hex", "Pyr", "Prism")


def _decode_device_cell_tables():
    import os
    import re
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kvs_amd", "csrc", "cell_tables.h")
    text = open(path).read()
    out = {}
    for short in CELL_SHORT:
        body = text[text.index(f"k{short}Tri["):text.index(f"k{short}Edge[")]
        words = [int(w, 16) for w in re.findall(r"0x([0-9a-f]{16})ull", body)]
        edge_line = text[text.index(f"k{short}Edge["):].split("\n")[0]
        edges = [int(b, 16) for b in re.findall(r"0x([0-9a-f]{2})", edge_line)]
        out[short] = (words, edges)
    return out


def _decode_oracle_cell_tables():
    import os
    import re
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "cell_tables_oracle.h")
    text = open(path).read()
    out = {}
    for short in CELL_SHORT:
        body = text[text.index(f"kOracle{short}Cases["):text.index(f"kOracle{short}Edges")]
        cases = re.findall(r'"([0-9a-f]*)"', body)
        edges = re.search(rf'kOracle{short}Edges = "([0-9]*)"', text).group(1)
        out[short] = (cases, edges)
    return out


def test_unstructured_tables_match_reference_hash():
    """Both transcriptions of the unstructured case tables -- the device words of
    kvs_amd/csrc/cell_tables.h and the oracle's strings -- decode to the
    reference's TriangleID / VertexID arrays (pinned by SHA-256)."""
    meta = golden_io.meta()["cell_tables"]
    dev, orc = _decode_device_cell_tables(), _decode_oracle_cell_tables()
    for name, short in zip(CELL_NAMES, CELL_SHORT):
        m = meta[name]
        rows, width = m["TriangleID_shape"]
        words, edges = dev[short]
        tri = np.full((rows, width), -1, dtype="<i4")
        for i, w in enumerate(words):
            n = w >> 60
            assert (w >> (12 * n)) & ((1 << (60 - 12 * n)) - 1) == 0, (name, i)
            tri[i, :3 * n] = [(w >> (4 * j)) & 0xF for j in range(3 * n)]
        assert len(words) == rows
        assert golden_io.sha(tri) == m["TriangleID_sha256"], name
        vert = np.array([[e & 0xF, e >> 4] for e in edges], dtype="<i4")
        assert golden_io.sha(vert) == m["VertexID_sha256"], name
        cases, estr = orc[short]
        tri2 = np.full((rows, width), -1, dtype="<i4")
        for i, c in enumerate(cases):
            tri2[i, :len(c)] = [int(ch, 16) for ch in c]
        assert golden_io.sha(tri2) == m["TriangleID_sha256"], name
        vert2 = np.array([[int(estr[2 * e]), int(estr[2 * e + 1])] for e in range(len(estr) // 2)], dtype="<i4")
        assert golden_io.sha(vert2) == m["VertexID_sha256"], name


def test_unstructured_against_golden():
    # MarchingTetrahedra / Hexahedra / Pyramid / Prism and the unstructured
    # SlicePlane restated (oracle/unstructured_oracle.c) against the reference
    for case, v, co, k, outs, (sc, sn, scl) in golden_io.unstructured_cases():
        ct = case["cell_type"]
        mm = tuple(case["min_max"]) if case["min_max"] is not None else None
        label = (ct, case["dtype"], case["dims"])
        for (normal, dup), (want_c, want_n) in outs.items():
            r = oracle.unstructured_isosurface(v, co, k, ct, case["iso"], normal, dup, min_max_values=mm)
            assert bits_equal(r["coords"], want_c), label + (normal, dup)
            assert bits_equal(r["normals"], want_n), label + (normal, dup)
        assert outs[(1, True)][0].size > 0, label
        table = np.array(golden_io.meta()["colormap_tables"][("BrewerRdBu", "BrewerSingleHueGreys")[case["slice"]["tf_kind"]] + "_256"],
                         dtype=np.uint8)
        r = oracle.unstructured_slice_plane(v, co, k, ct, case["slice"]["coefficients"], table, min_max_values=mm)
        assert bits_equal(r["coords"], sc) and bits_equal(r["normals"], sn), label
        assert np.array_equal(r["colors"], scl), label
        assert list(oracle.unstructured_min_max(v, k)) == case["volume_min_max"] or mm is not None, label
        lo, hi = oracle.unstructured_min_max_coords(co, k)
        assert lo == case["min_obj"] and hi == case["max_obj"], label


def test_unstructured_min_max_connection_order():
    # seeded with values[0] even when node 0 is in no cell; NaN handling of the
    # serial Math::Min / Math::Max scan; coordinates restart after a NaN
    v = np.array([5.0, 1.0, 2.0, 3.0, 4.0], np.float32)
    k = np.array([1, 2, 3, 4], np.uint32)
    assert oracle.unstructured_min_max(v, k) == (1.0, 5.0)
    c = np.array([[0, 0, 0], [1, 2, 3], [np.nan, 5, 5], [0.5, 7, 1], [2, 1, 9]], np.float32)
    lo, hi = oracle.unstructured_min_max_coords(c, np.array([1, 2, 3, 4], np.uint32))
    assert lo[0] == 0.5 and hi[0] == 2.0          # x restarts after the NaN
    assert lo[1] == 1.0 and hi[1] == 7.0 and lo[2] == 1.0 and hi[2] == 9.0


@pytest.mark.skipif(not oracle.ref_available(), reason="reference library not built here")
def test_unstructured_equals_reference_fresh():
    for ct in (datasets.TETRAHEDRA, datasets.HEXAHEDRA, datasets.PYRAMID, datasets.PRISM):
        for seed in range(3):
            v, c, k = datasets.unstructured_mesh((7 + seed, 6, 5), ct, seed=seed + 50)
            for iso in (0.25, 0.5):
                a = oracle.unstructured_isosurface(v, c, k, ct, iso)
                b = oracle.ref_unstructured_isosurface(v, c, k, ct, iso)
                assert bits_equal(a["coords"], b["coords"]) and bits_equal(a["normals"], b["normals"]), (ct, seed, iso)


def test_stl_fixtures_hash_and_layout():
    # the committed STL files are the ones make_golden.py recorded; the binary
    # ones hold 80 + 4 + 50 * T bytes for the triangles of the oracle's surface
    import hashlib
    import os
    for case in golden_io.meta()["stl"]:
        data = open(os.path.join(golden_io.GOLDEN, "stl", case["file"]), "rb").read()
        assert hashlib.sha256(data).hexdigest() == case["sha256"], case["file"]
        vol = datasets.blobs(20, seed=6) if case["volume"] == "blobs20" else datasets.blobs(10, seed=7, nblobs=2)
        r = oracle.extract(vol, case["iso"], case["normal"], case["dup"])
        if not case["ascii"]:
            assert data[:16] == b"Generated by KVS" and len(data) == 84 + 50 * r["n_triangles"]
            assert int.from_bytes(data[80:84], "little") == r["n_triangles"]
        else:
            assert data.count(b"facet normal") == r["n_triangles"]
