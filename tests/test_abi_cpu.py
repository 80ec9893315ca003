"""C ABI checks that need no GPU: the library loads and exports every entry
point include/kmc.h declares, argument validation follows the reference's
error exits, and the host-side colour restatement matches the reference."""
import ctypes
import os
import re

import numpy as np
import pytest

from kvs_amd import kmc
from tests import golden_io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    src = open(os.path.join(ROOT, "include", "kmc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(kmc_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = kmc.load()
    names = declared_entry_points()
    assert names, "no declarations found"
    for name in names:
        assert hasattr(lib, name), name
    assert set(kmc.EXPORTED) <= set(names)


def _extract_status(v):
    lib = kmc.load()
    s = kmc.Surface()
    return lib.kmc_extract(ctypes.byref(v) if v is not None else None, 0.5, kmc.POLYGON_NORMAL, 1,
                           kmc.MEMORY_HOST, None, ctypes.byref(s))


def test_error_exits_before_any_device_work(monkeypatch):
    monkeypatch.setenv("KMC_QUIET", "1")
    values = np.zeros((4, 4, 4), np.float32)
    assert _extract_status(None) == kmc.ERR_NULL_INPUT                       # MarchingCubes.cpp:50-54
    v = kmc.make_volume(values, kind=kmc.UNSTRUCTURED)
    assert _extract_status(v) == kmc.ERR_NOT_STRUCTURED                      # :56-61
    v = kmc.make_volume(values, veclen=3)
    assert _extract_status(v) == kmc.ERR_NOT_SCALAR                          # :63-67
    v = kmc.make_volume(values)
    v.type = 10
    assert _extract_status(v) == kmc.ERR_UNSUPPORTED_TYPE                    # :125-129
    v = kmc.make_volume(values)
    v.resolution[1] = 0
    assert _extract_status(v) == kmc.ERR_INVALID_ARGUMENT


def test_status_strings_follow_reference_messages():
    lib = kmc.load()
    assert lib.kmc_status_string(kmc.ERR_NULL_INPUT) == b"Input object is NULL."
    assert lib.kmc_status_string(kmc.ERR_NOT_STRUCTURED) == b"Input object is not structured volume object."
    assert lib.kmc_status_string(kmc.ERR_NOT_SCALAR) == b"Input volume is not sclar field data."


@pytest.mark.parametrize("name,kind", [("BrewerRdBu", kmc.COLORMAP_DEFAULT), ("BrewerSingleHueGreys", kmc.COLORMAP_GREYS)])
@pytest.mark.parametrize("res", [256, 17])
def test_colormap_tables_match_reference(name, kind, res):
    want = golden_io.meta()["colormap_tables"][f"{name}_{res}"]
    got = kmc.colormap_table(kind, res)
    assert list(got) == want


def test_surface_color_at_matches_reference():
    lib = kmc.load()
    for c in golden_io.meta()["colors"]:
        # kmc_surface_color applies MapperBase::setRange from (type, min, max):
        # a float type passes the range through unchanged unless it is degenerate
        lo, hi = c["lo"], c["hi"]
        if np.float32(lo) == np.float32(hi) or abs(np.float32(lo) - np.float32(hi)) < 1e-6:
            continue
        got = kmc.surface_color(c["kind"], 256, kmc.REAL32, lo, hi, c["value"])
        assert list(got) == c["rgb"], c


def test_mapper_surface_colors_match_reference():
    for c in golden_io.meta()["surface_colors"]:
        t = kmc.NUMPY_TYPES[np.dtype(c["dtype"])]
        got = kmc.surface_color(kmc.COLORMAP_DEFAULT, 256, t, c["min"], c["max"], c["iso"])
        assert list(got) == c["colors"], c


def test_random_case_colors_match_reference():
    for v, iso, mm, coords, outs in golden_io.random_cases():
        co, _, _, colors = outs[(1, True)]
        if co.size == 0:
            continue
        lo, hi = (mm if mm is not None else (float(v.min()), float(v.max())))
        got = kmc.surface_color(kmc.COLORMAP_DEFAULT, 256, kmc.NUMPY_TYPES[v.dtype], lo, hi, iso)
        assert list(got) == list(colors)


def test_surface_color_with_a_one_value_float_range():
    # int64 data near 2^62: float(min) == float(max) and the +1.0f of
    # MapperBase::setRange is absorbed, so ColorMap::at divides 0 by 0; the
    # reference then indexes its table with size_t(NaN), kmc_surface_color
    # returns the table's first colour
    lo, hi = float(2 ** 62 - 3000 * 512), float(2 ** 62 + 3000 * 512)
    assert np.float32(lo) == np.float32(hi)
    got = kmc.surface_color(kmc.COLORMAP_DEFAULT, 256, kmc.INT64, lo, hi, float(2 ** 62))
    first = kmc.colormap_table(kmc.COLORMAP_DEFAULT, 256)[:3]
    assert list(got) == list(first)


def test_noalloc_refuses_work_beyond_its_scratch():
    # the work a 2048 x 2048 x 512 volume needs (48 bytes per 64-node chunk)
    # exceeds the fixed scratch: refused before any device call, so this holds
    # without a GPU; the values pointer is never dereferenced
    lib = kmc.load()
    v = kmc.make_volume(1, dims=(2048, 2048, 512), memory=kmc.MEMORY_DEVICE, dtype=np.float32, min_max=(0.0, 1.0))
    s = kmc.Surface()
    st = lib.kmc_extract_noalloc(ctypes.byref(v), 0.5, kmc.POLYGON_NORMAL, 1, None, None, 0, None, 0, None, 0,
                                 ctypes.byref(s))
    assert st == kmc.ERR_SCRATCH
    # host memory is refused: the call never copies (that would allocate)
    a = np.zeros((4, 4, 4), np.float32)
    st = lib.kmc_extract_noalloc(ctypes.byref(kmc.make_volume(a)), 0.5, kmc.POLYGON_NORMAL, 1, None, None, 0, None,
                                 0, None, 0, ctypes.byref(s))
    assert st == kmc.ERR_INVALID_ARGUMENT
