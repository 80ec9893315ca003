"""The drop-in kvs::MarchingCubes / kvs::Isosurface (kvs_amd/dropin/
MarchingCubes.cpp, built against the reference's own headers and object
model by oracle/Makefile into oracle/_ref/dropin_test) on the GPU, against
the golden digests written from the reference and against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from kvs_amd import datasets, kmc
from tests import golden_io
from tests.conftest import bits_equal

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_test")


def run(tmp_path, mode, values, iso, normal, dup):
    raw = tmp_path / "v.raw"
    values.tofile(raw)
    nz, ny, nx = values.shape
    out = str(tmp_path / "out")
    print("dropin", mode, values.dtype, values.shape, iso, normal, dup, flush=True)
    subprocess.run([EXE, mode, str(raw), str(kmc.NUMPY_TYPES[values.dtype]), str(nx), str(ny), str(nz),
                    repr(float(iso)), str(int(normal)), str(int(dup)), out], check=True, timeout=120)
    meta = {}
    for line in open(out + ".meta"):
        k, *rest = line.split()
        meta[k] = [int(x) for x in rest]
    return {
        "coords": np.fromfile(out + ".coords", dtype=np.float32),
        "connections": np.fromfile(out + ".conn", dtype=np.uint32),
        "normals": np.fromfile(out + ".normals", dtype=np.float32),
        "meta": meta,
    }


needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in driver not built (make -C oracle dropin)")


@needs_exe
def test_dropin_hydrogen_matches_reference(tmp_path):
    h = golden_io.hydrogen()
    levels = golden_io.meta()["hydrogen64"]["levels"]
    for key, d in levels.items():
        iso, mode = key.split("_n")
        normal, dup = int(mode[0]), bool(int(mode.split("_d")[1]))
        r = run(tmp_path, "mc", h, float(iso), normal, dup)
        assert golden_io.sha(r["coords"]) == d["coords"], key
        assert golden_io.sha(r["connections"]) == d["connections"], key
        assert golden_io.sha(r["normals"]) == d["normals"], key
        m = r["meta"]
        assert m["colors"] == d["colors"] and m["opacities"] == d["opacities"], key
        assert m["polygon_type"][0] == d["polygon_type"] and m["normal_type"][0] == d["normal_type"], key
        assert m["color_type"][0] == d["color_type"] and m["success"][0] == d["success"], key


@needs_exe
def test_dropin_isosurface_default_level(tmp_path):
    h = golden_io.hydrogen()
    d = golden_io.meta()["hydrogen64"]["isosurface_default_level"]
    r = run(tmp_path, "isodefault", h, 0.0, kmc.POLYGON_NORMAL, True)
    assert golden_io.sha(r["coords"]) == d["coords"]
    assert golden_io.sha(r["normals"]) == d["normals"]
    assert r["meta"]["colors"] == d["colors"]


@needs_exe
def test_dropin_isosurface_and_types_against_oracle(tmp_path):
    rng = np.random.default_rng(31)
    for dtype in (np.float32, np.uint16, np.int32, np.float64):
        v = (datasets.blobs((21, 18, 15), seed=4, nblobs=4) * 1000).astype(dtype)
        iso = float(np.percentile(v, 70)) + 0.25
        for normal, dup in golden_io.MODES:
            for mode in ("mc", "iso"):
                r = run(tmp_path, mode, v, iso, normal, dup)
                want = oracle.extract(v, iso, normal, dup)
                assert bits_equal(r["coords"], want["coords"]), (dtype, mode, normal, dup)
                assert bits_equal(r["connections"], want["connections"]), (dtype, mode, normal, dup)
                assert bits_equal(r["normals"], want["normals"]), (dtype, mode, normal, dup)
                assert r["meta"]["opacities"] == [255]


@needs_exe
def test_dropin_error_paths(tmp_path):
    r = run(tmp_path, "null", np.zeros((2, 2, 2), np.float32), 0.0, kmc.POLYGON_NORMAL, True)
    assert r["meta"]["success"] == [0] and r["meta"]["returned"] == [0]
    # constant volume: no surface, opacity 255 from the PolygonObject constructor
    r = run(tmp_path, "mc", np.full((4, 4, 4), 2.0, np.float32), 2.0, kmc.POLYGON_NORMAL, True)
    assert r["coords"].size == 0 and r["meta"]["success"] == [1]
