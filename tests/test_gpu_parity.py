"""HIP path vs the oracle, bit for bit, through the C ABI (kmc_extract /
kmc_extract_levels).  Seeded inputs at sizes the oracle finishes in seconds."""
import ctypes

import numpy as np
import pytest

import oracle
from kvs_amd import datasets, kmc
from tests.conftest import assert_same_surface

pytestmark = pytest.mark.gpu

MODES = [(kmc.POLYGON_NORMAL, True), (kmc.POLYGON_NORMAL, False), (kmc.VERTEX_NORMAL, False)]
DTYPES = [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64, np.float32, np.float64]


def random_volume(rng, dtype, shape):
    if np.dtype(dtype).kind == "f":
        v = rng.standard_normal(shape).astype(dtype)
        iso = float(np.float32(rng.standard_normal() * 0.5))
    else:
        info = np.iinfo(dtype)
        lo, hi = max(int(info.min), -100), min(int(info.max), 100)
        v = rng.integers(lo, hi + 1, shape).astype(dtype)
        iso = float(rng.integers(lo, hi)) + (0.5 if rng.random() < 0.5 else 0.0)
    return v, iso


def check(values, iso, normal, dup, **kw):
    label = f"{values.dtype} {values.shape} iso={iso!r} normal={normal} dup={dup}"
    print("case", label, flush=True)
    got = kmc.extract(values, iso, normal, dup, **kw)
    want = oracle.extract(values, iso, normal, dup,
                          min_max_values=kw.get("min_max"), coords=kw.get("coords"))
    assert_same_surface(got, want, label)
    return got


@pytest.mark.parametrize("dtype", DTYPES, ids=lambda d: np.dtype(d).name)
def test_random_small_all_types(dtype):
    rng = np.random.default_rng(100 + DTYPES.index(dtype))
    for trial in range(4):
        shape = tuple(int(s) for s in rng.integers(2, 30, 3))
        v, iso = random_volume(rng, dtype, shape)
        for normal, dup in MODES:
            check(v, iso, normal, dup)


def test_blobs_64_all_modes():
    v = datasets.blobs(64, seed=1000)
    for normal, dup in MODES:
        got = check(v, 0.5, normal, dup)
        assert got["n_triangles"] > 1000


def test_wide_rows_and_band_edges():
    # row lengths around the 64-node chunk size and band heights around R
    rng = np.random.default_rng(7)
    for nx in (1, 2, 63, 64, 65, 127, 128, 129, 200):
        for ny in (1, 2, 3, 4, 5, 9):
            v = rng.standard_normal((6, ny, nx)).astype(np.float32)
            for normal, dup in MODES:
                check(v, 0.1, normal, dup)


def test_many_slices_across_slabs():
    v = datasets.blobs((37, 29, 101), seed=5, nblobs=6)
    for normal, dup in MODES:
        check(v, 0.3, normal, dup)


def test_levels_fused_equal_separate():
    v = datasets.blobs(48, seed=9, nblobs=5)
    levels = [0.2, 0.35, 0.5, 0.65]
    for normal, dup in MODES:
        fused = kmc.extract_levels(v, levels, normal, dup)
        for lv, got in zip(levels, fused):
            want = oracle.extract(v, lv, normal, dup)
            assert_same_surface(got, want, f"fused level {lv} mode {normal},{dup}")


def test_fused_levels_with_empty_levels_between():
    # levels outside the value range give no items; the per-item kernels walk
    # the levels one after another and must skip them
    v = datasets.blobs((33, 20, 70), seed=12, nblobs=5)
    lo, hi = float(v.min()), float(v.max())
    levels = [hi + 1.0, 0.3, lo - 1.0, hi + 2.0, 0.6, 0.45, lo - 0.5]
    for normal, dup in MODES:
        outs = kmc.extract_levels(v, levels, normal, dup)
        for lv, got in zip(levels, outs):
            want = oracle.extract(v, lv, normal, dup)
            assert_same_surface(got, want, f"level {lv} mode {normal},{dup}")
        assert outs[0]["n_triangles"] == 0 and outs[1]["n_triangles"] > 0


def test_more_levels_than_one_pass():
    v = datasets.blobs(24, seed=11, nblobs=4)
    levels = [0.05 * (i + 1) for i in range(11)]
    outs = kmc.extract_levels(v, levels, kmc.VERTEX_NORMAL, False)
    for lv, got in zip(levels, outs):
        assert_same_surface(got, oracle.extract(v, lv, kmc.VERTEX_NORMAL, False), f"level {lv}")


@pytest.mark.parametrize("nlev", [5, 6, 7])
def test_pass_of_five_to_seven_levels(nlev):
    # passes of 5..7 levels run the 8-level classify instance
    v = datasets.blobs((70, 21, 13), seed=40 + nlev, nblobs=4)
    levels = list(np.linspace(0.15, 0.85, nlev))
    for normal, dup in MODES:
        for lv, got in zip(levels, kmc.extract_levels(v, levels, normal, dup)):
            assert_same_surface(got, oracle.extract(v, lv, normal, dup), f"{nlev} levels, {lv} mode {normal},{dup}")


def test_constant_volume_is_empty():
    v = np.full((8, 8, 8), 3.0, dtype=np.float32)
    got = kmc.extract(v, 3.0, kmc.POLYGON_NORMAL, True)
    assert got["empty_range"] == 1
    assert got["coords"].size == 0 and got["normals"].size == 0


def test_ties_count_as_below():
    # values equal to the isolevel classify as "not above" (strict >)
    v = np.zeros((5, 5, 5), dtype=np.uint8)
    v[2, 2, 2] = 1
    v[1:4, 1:4, 3] = 2
    for normal, dup in MODES:
        check(v, 1.0, normal, dup)


def test_nan_values_float():
    rng = np.random.default_rng(3)
    v = rng.standard_normal((12, 13, 14)).astype(np.float32)
    v[rng.random(v.shape) < 0.05] = np.nan
    v[0, 0, 0] = 0.25       # first value finite: NaNs are skipped by min/max
    for normal, dup in MODES:
        check(v, 0.0, normal, dup)


def test_user_min_max_and_coords_quirks():
    v = datasets.blobs(20, seed=21, nblobs=3)
    mm = (0.2, 0.7)                         # narrower than the data: clamps the ratios
    coords = ([1.5, -2.0, 3.0], [10.0, 4.0, 9.5])   # (p + min) * scale quirk
    for normal, dup in MODES:
        check(v, 0.45, normal, dup, min_max=mm, coords=coords)


def test_device_min_max_matches_reference_scan():
    rng = np.random.default_rng(4)
    for dtype in DTYPES:
        v, _ = random_volume(rng, dtype, (9, 10, 11))
        assert kmc.min_max(v) == oracle.min_max(v)
    f = rng.standard_normal((6, 7, 8)).astype(np.float32)
    f[0, 0, 0] = np.nan
    got = kmc.min_max(f)
    assert np.isnan(got[0]) and np.isnan(got[1])


def test_two_dimensional_volume_vertices_without_cells():
    rng = np.random.default_rng(8)
    v = rng.standard_normal((1, 17, 23)).astype(np.float32)
    for normal, dup in MODES:
        check(v, 0.0, normal, dup)


def test_rows_wider_than_fused_pass():
    # rows of more than 64 chunks (4096 nodes) take the separate classify + count path
    rng = np.random.default_rng(12)
    v = rng.standard_normal((3, 4, 4100)).astype(np.float32)
    for normal, dup in MODES:
        check(v, 0.2, normal, dup)


def test_many_rows_across_groups():
    # more rows than one fused-pass workgroup holds, and a partial last group
    v = datasets.blobs((40, 67, 9), seed=13, nblobs=5)
    for normal, dup in MODES:
        check(v, 0.3, normal, dup)


def test_tile_team_emit(monkeypatch):
    # the tile-team emit (used for volumes whose y or z exceed the flat records'
    # 20-bit fields, and an autotuner candidate of the batched form) on the same
    # cases: KMC_EMIT=team selects it per call
    monkeypatch.setenv("KMC_EMIT", "team")
    rng = np.random.default_rng(31)
    for dtype in (np.uint8, np.int16, np.float32, np.float64):
        shape = tuple(int(s) for s in rng.integers(2, 30, 3))
        v, iso = random_volume(rng, dtype, shape)
        for normal, dup in MODES:
            check(v, iso, normal, dup)
    for nx in (63, 64, 65, 129):
        v = rng.standard_normal((6, 5, nx)).astype(np.float32)
        for normal, dup in MODES:
            check(v, 0.1, normal, dup)
    v = datasets.blobs((37, 29, 101), seed=5, nblobs=6)
    for normal, dup in MODES:
        check(v, 0.3, normal, dup)
    v = rng.standard_normal((3, 4, 4100)).astype(np.float32)
    for normal, dup in MODES:
        check(v, 0.2, normal, dup)
    v = datasets.blobs(48, seed=77)
    levels = [0.2, 0.45, 0.7]
    for normal, dup in MODES:
        got = kmc.extract_levels(v, levels, normal, dup)
        for lv, g in zip(levels, got):
            assert_same_surface(g, oracle.extract(v, lv, normal, dup), f"team levels {lv} {normal} {dup}")


@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.uint16], ids=lambda d: np.dtype(d).name)
def test_small_types_rows_of_whole_dwords_and_not(dtype):
    """8- and 16-bit volumes classify one dword of nodes per lane when their
    rows are whole dwords, and node by node otherwise: row lengths on both
    sides of that rule and around the 64-dword segments and 64-node chunks."""
    rng = np.random.default_rng(300 + np.dtype(dtype).itemsize * 2 + int(np.dtype(dtype).kind == "u"))
    for nx in (2, 3, 4, 6, 60, 62, 64, 66, 68, 124, 126, 128, 130, 252, 256, 260, 510, 512, 516, 1030):
        v, iso = random_volume(rng, dtype, (5, 7, nx))
        for normal, dup in MODES:
            check(v, iso, normal, dup)


def test_small_types_values_off_dword_alignment():
    """Device values one element past a 4-byte boundary (a view into a
    larger array) classify node by node: the same surface as the oracle's."""
    lib = kmc.load()
    rng = np.random.default_rng(57)
    for dtype in (np.uint8, np.int16):
        v, iso = random_volume(rng, dtype, (6, 9, 64))
        flat = np.concatenate([np.zeros(1, dtype), v.ravel()])
        d = kmc.DeviceArray(flat)
        try:
            vol = kmc.make_volume(d.ptr.value + np.dtype(dtype).itemsize, dims=(64, 9, 6), memory=kmc.MEMORY_DEVICE,
                                  dtype=dtype)
            for normal, dup in MODES:
                label = f"{np.dtype(dtype).name} off by one element, normal={normal} dup={dup}"
                print("case", label, flush=True)
                s = kmc.Surface()
                st = lib.kmc_extract(ctypes.byref(vol), float(iso), int(normal), int(dup), kmc.MEMORY_HOST, None,
                                     ctypes.byref(s))
                assert st == kmc.OK, st
                try:
                    got = kmc.surface_to_numpy(s)
                finally:
                    lib.kmc_surface_release(ctypes.byref(s))
                assert_same_surface(got, oracle.extract(v, iso, normal, dup), label)
        finally:
            d.free()
