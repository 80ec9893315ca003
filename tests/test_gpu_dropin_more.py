"""The drop-in kvs::SlicePlane and kvs::Marching{Tetrahedra,Hexahedra,Pyramid,
Prism} (kvs_amd/dropin, built against the reference's own headers and object
model by oracle/Makefile into oracle/_ref/dropin_more_test) on the GPU: the
reference's kvs::SlicePlane, kvs::Isosurface and mappers run unchanged on the
library, and their outputs equal the reference's own (golden fixtures written
from it), bit for bit, with the same types, colours and opacities."""
import os
import subprocess

import numpy as np
import pytest

from kvs_amd import datasets, kmc
from tests import golden_io
from tests.conftest import bits_equal

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_more_test")
needs_exe = pytest.mark.skipif(not os.path.exists(EXE),
                               reason="drop-in driver not built (make -C oracle dropin)")


def run(tmp_path, values, **kw):
    raw = tmp_path / "v.raw"
    np.ascontiguousarray(values).tofile(raw)
    out = str(tmp_path / "out")
    args = [EXE, f"values={raw}", f"type={kmc.NUMPY_TYPES[values.dtype]}", f"out={out}"]
    for k, v in kw.items():
        if isinstance(v, (list, tuple)):
            v = ",".join(repr(float(x)) for x in v)
        args.append(f"{k}={v}")
    print("dropin_more", " ".join(args[1:]), flush=True)
    subprocess.run(args, check=True, timeout=120)
    meta = {}
    for line in open(out + ".meta"):
        k, *rest = line.split()
        meta[k] = [float(x) for x in rest]
    return {
        "coords": np.fromfile(out + ".coords", dtype=np.float32),
        "normals": np.fromfile(out + ".normals", dtype=np.float32),
        "colors": np.fromfile(out + ".colors", dtype=np.uint8),
        "connections": np.fromfile(out + ".conn", dtype=np.uint32),
        "meta": meta,
    }


def unstructured_inputs(tmp_path, co, k):
    cpath, kpath = tmp_path / "c.raw", tmp_path / "k.raw"
    np.ascontiguousarray(co, dtype=np.float32).tofile(cpath)
    np.ascontiguousarray(k, dtype=np.uint32).tofile(kpath)
    return cpath, kpath


@needs_exe
def test_dropin_slice_plane_structured_golden(tmp_path):
    for case, v, (co, no, cl) in golden_io.slice_cases():
        nz, ny, nx = v.shape
        kw = dict(mode="slice", dims=(nx, ny, nz), coef=case["coefficients"], tf=case["tf_kind"])
        if case["min_max"] is not None:
            kw["minmax"] = case["min_max"]
        if case["coords"] is not None:
            kw["objcoords"] = list(case["coords"][0]) + list(case["coords"][1])
        r = run(tmp_path, v, **kw)
        label = (case["dtype"], case["plane"])
        assert bits_equal(r["coords"], co), label
        assert bits_equal(r["normals"], no), label
        assert np.array_equal(r["colors"], cl), label
        m = r["meta"]
        assert int(m["polygon_type"][0]) == case["polygon_type"] and int(m["color_type"][0]) == case["color_type"]
        assert int(m["normal_type"][0]) == case["normal_type"]
        assert [int(x) for x in m["opacities"]] == case["opacities"], label


@needs_exe
def test_dropin_unstructured_mappers_and_isosurface_golden(tmp_path):
    for case, v, co, k, outs, (sc, sn, scl) in golden_io.unstructured_cases():
        cpath, kpath = unstructured_inputs(tmp_path, co, k)
        ct = case["cell_type"]
        base = dict(coords=cpath, conn=kpath, nnodes=case["n_nodes"], ncells=k.size // kmc.CELL_NODES[ct], cell=ct)
        if case["min_max"] is not None:
            base["minmax"] = case["min_max"]
        for (normal, dup), (want_c, want_n) in outs.items():
            want = case[f"n{normal}_d{int(dup)}"]
            for mode in ("uiso", "umap"):
                r = run(tmp_path, v, mode=mode, iso=repr(case["iso"]), normal=normal, dup=int(dup), tf=0, **base)
                label = (ct, case["dtype"], case["dims"], normal, dup, mode)
                assert bits_equal(r["coords"], want_c), label
                assert bits_equal(r["normals"], want_n), label
                m = r["meta"]
                assert [int(x) for x in r["colors"]] == want["colors"], label
                assert int(m["polygon_type"][0]) == want["polygon_type"], label
                assert int(m["color_type"][0]) == want["color_type"], label
                assert int(m["normal_type"][0]) == want["normal_type"], label
                assert [int(x) for x in m["opacities"]] == want["opacities"], label
                assert m["volume_min_max"] == [float(x) for x in case["volume_min_max"]] or case["min_max"], label
                assert np.array_equal(np.float32(m["min_obj"]), np.float32(case["min_obj"])), label
                assert np.array_equal(np.float32(m["max_obj"]), np.float32(case["max_obj"])), label
        s = case["slice"]
        r = run(tmp_path, v, mode="uslice", coef=s["coefficients"], tf=s["tf_kind"], **base)
        assert bits_equal(r["coords"], sc) and bits_equal(r["normals"], sn), (ct, case["dtype"])
        assert np.array_equal(r["colors"], scl), (ct, case["dtype"])
        assert int(r["meta"]["color_type"][0]) == s["color_type"]


@needs_exe
def test_dropin_isosurface_on_a_larger_tetrahedral_mesh(tmp_path):
    # through kvs::Isosurface on the drop-in, against the library's own call
    v, c, k = datasets.unstructured_mesh((24, 20, 18), datasets.TETRAHEDRA, seed=77)
    cpath, kpath = unstructured_inputs(tmp_path, c, k)
    r = run(tmp_path, v, mode="uiso", iso="0.4", normal=kmc.POLYGON_NORMAL, dup=1, tf=0, coords=cpath, conn=kpath,
            nnodes=v.size, ncells=k.size // 4, cell=datasets.TETRAHEDRA)
    want = kmc.extract_unstructured(v, c, k, datasets.TETRAHEDRA, 0.4)
    assert r["coords"].size > 0
    assert bits_equal(r["coords"], want["coords"]) and bits_equal(r["normals"], want["normals"])
