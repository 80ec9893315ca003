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
