"""The file writers of the output side (kmc_export_kvsml, kmc_export_ply) are
host code: on host surfaces they run without a GPU.  Fed the oracle's surfaces
(bit-identical to the reference's own, which tests/test_oracle.py pins), they
must write the files the reference's PolygonExporter<KVSMLPolygonObject> and
PolygonExporter<Ply> wrote (tests/golden/kvsml_polygon/, tests/golden/ply/),
byte for byte."""
import os

import numpy as np
import pytest

import oracle
from kvs_amd import datasets, kmc
from tests import golden_io

GOLDEN = golden_io.GOLDEN
needs_lib = pytest.mark.skipif(not os.path.exists(kmc.LIB_PATH), reason="libkmc.so not built")


def surface(case, vol, table):
    if case["kind"] == "slice":
        return oracle.slice_plane(vol, case["coef"], table), kmc.POLYGON_NORMAL
    d = oracle.extract(vol, case["iso"], case["normal"], case["dup"])
    if case.get("drop_connections"):
        d = dict(d, connections=np.zeros(0, np.uint32))
    return d, case["normal"]


@needs_lib
def test_kvsml_writer_on_host_surfaces(tmp_path):
    vol = datasets.blobs(10, seed=7, nblobs=2)
    table = np.array(golden_io.meta()["colormap_tables"]["BrewerRdBu_256"], dtype=np.uint8)
    for case in golden_io.meta()["kvsml_polygon"]:
        d, normal = surface(case, vol, table)
        out = tmp_path / case["file"][:-len(".kvsml")]
        out.mkdir()
        kmc.export_kvsml(d, normal, out / case["file"], case["writing"], color=case["color"], opacity=case["opacity"])
        assert sorted(os.listdir(out)) == case["files"], case["file"]
        for f in case["files"]:
            with open(out / f, "rb") as a, open(os.path.join(GOLDEN, "kvsml_polygon", f), "rb") as b:
                assert a.read() == b.read(), f


@needs_lib
def test_ply_writer_on_host_surfaces(tmp_path):
    vol = datasets.blobs(10, seed=7, nblobs=2)
    table = np.array(golden_io.meta()["colormap_tables"]["BrewerRdBu_256"], dtype=np.uint8)
    for case in golden_io.meta()["ply"]:
        d, normal = surface(case, vol, table)
        kmc.export_ply(d, normal, tmp_path / case["file"], case["binary"], color=case["color"])
        with open(tmp_path / case["file"], "rb") as a, open(os.path.join(GOLDEN, "ply", case["file"]), "rb") as b:
            assert a.read() == b.read(), case["file"]
