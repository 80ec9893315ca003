"""The HIP path against the golden vectors written from the reference itself
(tests/golden/), through the C ABI."""
import numpy as np
import pytest

import oracle
from kvs_amd import datasets, kmc
from tests import golden_io
from tests.conftest import bits_equal

pytestmark = pytest.mark.gpu


def test_brute_force_cubes():
    for v, iso, ci, outs in golden_io.cube_cases():
        for (normal, dup), (co, cn, no) in outs.items():
            r = kmc.extract(v, iso, normal, dup)
            assert bits_equal(r["coords"], co), (v.dtype, ci, normal, dup)
            assert bits_equal(r["connections"], cn), (v.dtype, ci, normal, dup)
            assert bits_equal(r["normals"], no), (v.dtype, ci, normal, dup)


def test_random_volumes():
    for i, (v, iso, mm, coords, outs) in enumerate(golden_io.random_cases()):
        for (normal, dup), (co, cn, no, _) in outs.items():
            r = kmc.extract(v, iso, normal, dup, min_max=mm, coords=coords)
            assert bits_equal(r["coords"], co), (i, normal, dup)
            assert bits_equal(r["connections"], cn), (i, normal, dup)
            assert bits_equal(r["normals"], no), (i, normal, dup)


def test_large_64bit_integers():
    for name, v, iso, mm, outs in golden_io.bigint_cases():
        for (normal, dup), (co, cn, no) in outs.items():
            r = kmc.extract(v, iso, normal, dup, min_max=mm)
            assert bits_equal(r["coords"], co), (name, iso, normal, dup)
            assert bits_equal(r["connections"], cn), (name, iso, normal, dup)
            assert bits_equal(r["normals"], no), (name, iso, normal, dup)


def test_hydrogen64_digests():
    h = golden_io.hydrogen()
    meta = golden_io.meta()["hydrogen64"]
    for key, d in meta["levels"].items():
        iso, mode = key.split("_n")
        normal, dup = int(mode[0]), bool(int(mode.split("_d")[1]))
        r = kmc.extract(h, float(iso), normal, dup)
        assert golden_io.sha(r["coords"]) == d["coords"], key
        assert golden_io.sha(r["connections"]) == d["connections"], key
        assert golden_io.sha(r["normals"]) == d["normals"], key
        assert r["normal_type"] == d["normal_type"] and r["polygon_type"] == d["polygon_type"], key
        rgb = kmc.surface_color(kmc.COLORMAP_DEFAULT, 256, kmc.UINT8, float(h.min()), float(h.max()), float(iso))
        assert list(rgb) == d["colors"], key


def test_c0_configuration():
    # 64^3 float32 blobs, isolevel 0.5, PolygonNormal + duplication: the full
    # arrays of the reference's output
    v = datasets.blobs(64, seed=1000)
    meta = golden_io.meta()["c0"]
    r = kmc.extract(v, 0.5, kmc.POLYGON_NORMAL, True)
    if golden_io.sha(v) == meta["sha256"]:
        z = np.load(golden_io.GOLDEN + "/c0_dup.npz")
        assert bits_equal(r["coords"], z["coords"]) and bits_equal(r["normals"], z["normals"])
        for key, d in meta["levels"].items():
            normal, dup = int(key.split("_n")[1][0]), bool(int(key.split("_d")[1]))
            rr = kmc.extract(v, 0.5, normal, dup)
            assert golden_io.sha(rr["coords"]) == d["coords"], key
            assert golden_io.sha(rr["connections"]) == d["connections"], key
            assert golden_io.sha(rr["normals"]) == d["normals"], key
    else:
        # the host's libm rounded the generator differently: compare with the oracle
        want = oracle.extract(v, 0.5, kmc.POLYGON_NORMAL, True)
        assert bits_equal(r["coords"], want["coords"]) and bits_equal(r["normals"], want["normals"])
