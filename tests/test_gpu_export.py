"""The output side on the device: kmc_surface_expand (the triangle soup of
PolygonExporter<kvs::Stl>) and kmc_export_stl against the STL files the
reference's own exporter and writer produced (tests/golden/stl/), and
kmc_export_kvsml against its KVSML polygon files in the three writing types
(tests/golden/kvsml_polygon/), byte for byte; kmc_export_ply against its PLY
files (tests/golden/ply/) and, for surfaces with connections, which its PLY
writer cannot write, against the exporter's arrays and its PLY reader."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from kvs_amd import datasets, kmc
from tests import golden_io
from tests.conftest import bits_equal

pytestmark = pytest.mark.gpu

SDIR = os.path.join(golden_io.GOLDEN, "stl")
VOLUMES = {"blobs20": lambda: datasets.blobs(20, seed=6), "blobs10": lambda: datasets.blobs(10, seed=7, nblobs=2)}


def expected_soup(d, normal_type):
    """The exporter's arrays restated in numpy float32 (PolygonExporter.cpp:91-160)."""
    if d["connections"].size == 0:
        return d["coords"], d["normals"]
    k = d["connections"].reshape(-1, 3).astype(np.int64)
    c = d["coords"].reshape(-1, 3)
    coords = c[k].reshape(-1)
    if normal_type == kmc.VERTEX_NORMAL:
        n = d["normals"].reshape(-1, 3)
        s = (n[k[:, 0]] + n[k[:, 1]]).astype(np.float32)
        normals = ((s + n[k[:, 2]]).astype(np.float32) / np.float32(3.0)).astype(np.float32).reshape(-1)
    else:
        normals = d["normals"]
    return coords, normals


@pytest.mark.parametrize("out_memory", [kmc.MEMORY_HOST, kmc.MEMORY_DEVICE])
def test_surface_expand(out_memory):
    v = datasets.blobs(48, seed=3, nblobs=5)
    for normal, dup in [(kmc.POLYGON_NORMAL, True), (kmc.POLYGON_NORMAL, False), (kmc.VERTEX_NORMAL, False)]:
        d = kmc.extract(v, 0.4, normal, dup)
        got = kmc.surface_expand(d, d["normal_type"], out_memory)
        coords, normals = expected_soup(d, d["normal_type"])
        assert bits_equal(got["coords"], coords), (normal, dup)
        assert bits_equal(got["normals"], normals), (normal, dup)
        assert got["n_triangles"] == d["n_triangles"] and got["connections"].size == 0


def test_export_stl_matches_reference_files(tmp_path):
    for case in golden_io.meta()["stl"]:
        v = VOLUMES[case["volume"]]()
        d = kmc.extract(v, case["iso"], case["normal"], case["dup"])
        out = tmp_path / case["file"]
        kmc.export_stl(d, d["normal_type"], out, ascii=case["ascii"])
        want = open(os.path.join(SDIR, case["file"]), "rb").read()
        assert out.read_bytes() == want, case["file"]


KPDIR = os.path.join(golden_io.GOLDEN, "kvsml_polygon")


def same_files(out_dir, case):
    stem = case["file"][:-len(".kvsml")]
    ours = sorted(f for f in os.listdir(out_dir) if f == case["file"] or f.startswith(stem + "_"))
    assert ours == case["files"], (case["file"], ours)
    for f in ours:
        with open(os.path.join(out_dir, f), "rb") as a, open(os.path.join(KPDIR, f), "rb") as b:
            assert a.read() == b.read(), f


def test_export_kvsml_equals_reference_files(tmp_path):
    # PolygonExporter<KVSMLPolygonObject> + write: marching-cubes surfaces in
    # the three modes and an empty one (host and device surfaces), a slice
    # plane with per-vertex colours, each in Ascii, ExternalAscii and
    # ExternalBinary
    vol = datasets.blobs(10, seed=7, nblobs=2)
    table = np.array(golden_io.meta()["colormap_tables"]["BrewerRdBu_256"], dtype=np.uint8)
    lib = kmc.load()
    for case in golden_io.meta()["kvsml_polygon"]:
        print("case", case["file"], flush=True)
        host, dev = tmp_path / ("host_" + case["file"]), tmp_path / ("dev_" + case["file"])
        host.mkdir()
        dev.mkdir()
        if case["kind"] == "slice":
            d = kmc.slice_plane(vol, case["coef"], table)
            kmc.export_kvsml(d, kmc.POLYGON_NORMAL, host / case["file"], case["writing"], opacity=case["opacity"])
            same_files(host, case)
            continue
        d = kmc.extract(vol, case["iso"], case["normal"], case["dup"])
        kmc.export_kvsml(d, case["normal"], host / case["file"], case["writing"], color=case["color"],
                         opacity=case["opacity"])
        same_files(host, case)
        v = kmc.make_volume(vol)
        s = kmc.Surface()
        assert lib.kmc_extract(ctypes.byref(v), case["iso"], case["normal"], int(case["dup"]), kmc.MEMORY_DEVICE,
                               None, ctypes.byref(s)) == 0
        try:
            c = (ctypes.c_uint8 * 3)(*case["color"]) if case["color"] else None
            assert lib.kmc_export_kvsml(ctypes.byref(s), c, case["opacity"], str(dev / case["file"]).encode(),
                                        case["writing"], None) == 0
        finally:
            lib.kmc_surface_release(ctypes.byref(s))
        same_files(dev, case)


def test_export_kvsml_refusals(tmp_path):
    d = kmc.extract(datasets.blobs(10, seed=7, nblobs=2), 0.4, kmc.POLYGON_NORMAL, True)
    for writing, opacity in ((3, 255), (-1, 255), (0, 256)):
        with pytest.raises(kmc.KmcError):
            kmc.export_kvsml(d, kmc.POLYGON_NORMAL, tmp_path / "x.kvsml", writing, opacity=opacity)
    with pytest.raises(kmc.KmcError):
        kmc.export_kvsml(d, kmc.POLYGON_NORMAL, tmp_path / "missing_dir" / "x.kvsml", kmc.KVSML_ASCII)


PDIR = os.path.join(golden_io.GOLDEN, "ply")


def test_export_ply_equals_reference_files(tmp_path):
    # PolygonExporter<kvs::Ply> + Ply::write on surfaces without connections
    # (the reference's writer fails on connections): a duplicated surface (its
    # PolygonNormal normals average to zero without connections), a
    # VertexNormal surface without its connections, an empty one and a slice
    # plane with per-vertex colours; ascii and binary; host and device surfaces
    vol = datasets.blobs(10, seed=7, nblobs=2)
    table = np.array(golden_io.meta()["colormap_tables"]["BrewerRdBu_256"], dtype=np.uint8)
    lib = kmc.load()
    for case in golden_io.meta()["ply"]:
        print("case", case["file"], flush=True)
        want = open(os.path.join(PDIR, case["file"]), "rb").read()
        mine = tmp_path / ("host_" + case["file"])
        if case["kind"] == "slice":
            kmc.export_ply(kmc.slice_plane(vol, case["coef"], table), kmc.POLYGON_NORMAL, mine, case["binary"])
            assert open(mine, "rb").read() == want, case["file"]
            continue
        d = kmc.extract(vol, case["iso"], case["normal"], case["dup"])
        if case.get("drop_connections"):
            d = dict(d, connections=np.zeros(0, np.uint32))
        kmc.export_ply(d, case["normal"], mine, case["binary"], color=case["color"])
        assert open(mine, "rb").read() == want, case["file"]
        if case.get("drop_connections"):
            continue
        v = kmc.make_volume(vol)
        s = kmc.Surface()
        assert lib.kmc_extract(ctypes.byref(v), case["iso"], case["normal"], int(case["dup"]), kmc.MEMORY_DEVICE,
                               None, ctypes.byref(s)) == 0
        try:
            c = (ctypes.c_uint8 * 3)(*case["color"]) if case["color"] else None
            dev = tmp_path / ("dev_" + case["file"])
            assert lib.kmc_export_ply(ctypes.byref(s), c, str(dev).encode(), int(case["binary"]), None) == 0
        finally:
            lib.kmc_surface_release(ctypes.byref(s))
        assert open(dev, "rb").read() == want, case["file"]


def ply_body(path, nv, nf, colors):
    """Vertex records and faces of a binary PLY file of our writer."""
    raw = open(path, "rb").read()
    body = raw[raw.index(b"end_header\n") + len(b"end_header\n"):]
    vt = [("p", "<f4", 3)] + ([("c", "u1", 3)] if colors else []) + [("n", "<f4", 3)]
    verts = np.frombuffer(body, dtype=np.dtype(vt), count=nv)
    faces = np.frombuffer(body, dtype=np.dtype([("k", "u1"), ("i", "<i4", 3)]), count=nf,
                          offset=verts.nbytes)
    return raw[:raw.index(b"end_header\n")].decode(), verts, faces


def test_export_ply_with_connections(tmp_path):
    # surfaces with connections, which the reference's writer cannot write:
    # the header declares the faces, the vertex records hold the exporter's
    # arrays (PolygonNormal normals summed over each vertex's triangles in
    # triangle order, divided by their count), the faces follow as a uchar
    # count and int indices, and the reference's own reader gets the
    # surface back; the ascii file is the reference's file of the same surface
    # without connections with the faces added
    vol = datasets.blobs(10, seed=7, nblobs=2)
    cases = {c["file"]: c for c in golden_io.meta()["ply"]}
    for normal in (kmc.POLYGON_NORMAL, kmc.VERTEX_NORMAL):
        d = kmc.extract(vol, 0.4, normal, False)
        conn = d["connections"].astype(np.int64)
        nv, nf = d["coords"].size // 3, conn.size // 3
        color = cases["mc_n1_d1_ascii.ply"]["color"]
        if normal == kmc.VERTEX_NORMAL:
            normals = d["normals"].reshape(-1, 3)
        else:
            acc = np.zeros((nv, 3), np.float32)
            np.add.at(acc, conn, np.repeat(d["normals"].reshape(-1, 3), 3, axis=0))
            cnt = np.bincount(conn, minlength=nv).astype(np.float32)
            normals = acc.copy()
            normals[cnt > 0] = (acc[cnt > 0] / cnt[cnt > 0, None]).astype(np.float32)
        path = tmp_path / f"n{normal}.ply"
        kmc.export_ply(d, normal, path, binary=True, color=color)
        head, verts, faces = ply_body(path, nv, nf, True)
        assert f"element vertex {nv}\n" in head and head.endswith(
            f"element face {nf}\nproperty list uchar int vertex_indices\n")
        assert bits_equal(verts["p"].reshape(-1), d["coords"]) and bits_equal(verts["n"], normals)
        assert (verts["c"] == np.array(color, np.uint8)).all()
        assert (faces["k"] == 3).all() and np.array_equal(faces["i"].reshape(-1), conn)
        back = oracle.ref_read_ply(path)
        assert bits_equal(back["coords"], d["coords"]) and np.array_equal(back["connections"], d["connections"])
        if normal == kmc.VERTEX_NORMAL:
            kmc.export_ply(d, normal, tmp_path / "vn_ascii", binary=False, color=color)
            ours = open(tmp_path / "vn_ascii.ply", "rb").read()
            ref = open(os.path.join(PDIR, "mc_n0_d0_noconn_ascii.ply"), "rb").read()
            ref = ref.replace(b"element face 0\n", f"element face {nf}\nproperty list uchar int vertex_indices\n"
                              .encode())
            faces_txt = "".join(f"3 {a} {b} {c} \n" for a, b, c in conn.reshape(-1, 3)).encode()
            assert ours == ref + faces_txt
