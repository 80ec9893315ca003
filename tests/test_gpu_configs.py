"""HIP path vs the oracle, bit for bit, on the geometries the benchmark runs.

The count and compaction passes split a wave's lanes into rows of CS chunk
lanes (CS = the power of two >= min(chunks per row, 64)) and scan the chunks
of a row with DPP row shifts (CS <= 16) or cross-lane shuffles (CS = 32, 64);
rows of at most 64 chunks keep both in-row chunk prefixes packed in one 32-bit
word.  The row widths below take every one of these paths, including the
widths of the 512^3 CT-like volume (8 chunks) and of the 1024^3 turbulence
field (16 chunks), which are also checked on slabs of those volumes with the
benchmark's isolevels and normal type."""
import numpy as np
import pytest

import oracle
from kvs_amd import datasets, kmc
from tests.conftest import assert_same_surface

pytestmark = pytest.mark.gpu

MODES = [(kmc.POLYGON_NORMAL, True), (kmc.POLYGON_NORMAL, False), (kmc.VERTEX_NORMAL, False)]
TURBULENCE_LEVELS = [-0.3, 0.0, 0.3]


def check_levels(values, levels, normal, dup, label):
    print("case", label, values.dtype, values.shape, levels, normal, dup, flush=True)
    got = kmc.extract_levels(values, levels, normal, dup)
    for lv, g in zip(levels, got):
        want = oracle.extract(values, lv, normal, dup)
        assert_same_surface(g, want, f"{label} {values.shape} level {lv} mode {normal},{dup}")
    return got


# chunks per row: 5 (CS 8), 9 and 16 (CS 16), 24 (CS 32), 33 and 64 (CS 64,
# packed prefixes), 65 (two prefix arrays)
ROW_WIDTHS = [300, 512, 513, 1024, 1500, 2049, 4096, 4097]


@pytest.mark.parametrize("nx", ROW_WIDTHS)
def test_row_widths_dense_crossings(nx):
    # white noise: nearly every chunk holds vertices and triangles
    rng = np.random.default_rng(nx)
    v = rng.standard_normal((4, 5, nx)).astype(np.float32)
    for normal, dup in MODES:
        check_levels(v, [0.1], normal, dup, f"noise nx={nx}")
        check_levels(v, TURBULENCE_LEVELS, normal, dup, f"noise nx={nx} fused")


@pytest.mark.parametrize("nx", ROW_WIDTHS)
def test_row_widths_smooth_field(nx):
    # a smooth field: sparse non-empty chunks at uneven places along the rows
    v = datasets.turbulence((nx, 7, 5), seed=nx, period=max(nx, 64))
    for normal, dup in MODES:
        check_levels(v, TURBULENCE_LEVELS, normal, dup, f"turbulence nx={nx}")


@pytest.mark.parametrize("nx", [1024, 4096])
def test_checkerboard_fills_packed_prefixes(nx):
    # every edge crosses and every cell holds four triangles: the in-row chunk
    # prefixes reach 63 * 192 vertices and 63 * 256 triangles before the last
    # chunk of a 64-chunk row, the largest values their 16-bit halves carry here
    z, y, x = np.meshgrid(np.arange(3), np.arange(3), np.arange(nx), indexing="ij")
    v = np.where((x + y + z) % 2 == 0, 1.0, -1.0).astype(np.float32)
    for normal, dup in MODES:
        got = check_levels(v, [0.0], normal, dup, f"checkerboard nx={nx}")
        assert got[0]["n_triangles"] == 4 * (nx - 1) * 2 * 2


@pytest.mark.parametrize("nx", [21000, 33000])
def test_rows_longer_than_a_tile_team_holds(nx):
    # rows of 329 and 516 chunks: the tile teams' LDS state does not fit, the
    # flat emit (which keeps none) serves them
    rng = np.random.default_rng(nx)
    v = rng.standard_normal((3, 3, nx)).astype(np.float32)
    for normal, dup in MODES:
        check_levels(v, [0.0], normal, dup, f"long rows nx={nx}")


def test_turbulence_1024_slab_three_levels_vertex_normal():
    # the 1024^3 turbulence field's geometry: 1024 x 1024 rows, its three fused
    # isolevels, VertexNormal, on a z-slab of the same field definition
    v = datasets.turbulence((1024, 1024, 10), seed=2, period=1024)
    got = check_levels(v, TURBULENCE_LEVELS, kmc.VERTEX_NORMAL, False, "turbulence 1024^2 slab")
    assert all(g["n_triangles"] > 10000 for g in got)


def test_turbulence_1024_slab_polygon_modes():
    v = datasets.turbulence((1024, 1024, 6), seed=2, period=1024)
    for normal, dup in MODES[:2]:
        check_levels(v, TURBULENCE_LEVELS, normal, dup, "turbulence 1024^2 slab")


def test_ct_like_512_slab_vertex_normal():
    # the 512^3 uint8 CT-like volume's geometry (512 x 512 rows), isolevel 100,
    # VertexNormal, on a z-slab of the same volume definition
    v = datasets.ct_like((512, 512, 24), seed=1)
    got = check_levels(v, [100.0], kmc.VERTEX_NORMAL, False, "ct-like 512^2 slab")
    assert got[0]["n_triangles"] > 10000


def test_batched_wide_rows():
    # the batched form over rows of 300..2049 nodes (CS 8..64) next to small volumes
    rng = np.random.default_rng(5)
    vols = [rng.standard_normal((3, 4, nx)).astype(np.float32) for nx in (300, 1024, 2049)]
    vols.append(datasets.blobs((40, 33, 21), seed=1, nblobs=5))
    vols.append(datasets.turbulence((1500, 6, 4), seed=3, period=1500))
    for normal, dup in MODES:
        specs = [(kmc.make_volume(v), [0.0, 0.4] if k % 2 else [0.2], normal, dup) for k, v in enumerate(vols)]
        outs = kmc.batch_extract(specs)
        for (vol, levels, n, d), v, job_outs in zip(specs, vols, outs):
            for lv, got in zip(levels, job_outs):
                want = oracle.extract(v, lv, normal, dup)
                assert_same_surface(got, want, f"batched {v.shape} level {lv} mode {normal},{dup}")


@pytest.mark.parametrize("nx", [21000, 33000])
def test_rows_too_long_for_a_tile_team(nx):
    # 329 and 516 chunks a row: no tile team's LDS state fits such rows, so
    # they run on the flat emit only -- plain, fused, batched and served
    rng = np.random.default_rng(nx)
    v = rng.standard_normal((3, 4, nx)).astype(np.float32)
    for normal, dup in MODES:
        check_levels(v, [0.1], normal, dup, f"long rows nx={nx}")
        check_levels(v, [-0.2, 0.3], normal, dup, f"long rows nx={nx} fused")
    specs = [(kmc.make_volume(v), [0.1], normal, dup) for normal, dup in MODES]
    for (vol, levels, normal, dup), job_outs in zip(specs, kmc.batch_extract(specs)):
        assert_same_surface(job_outs[0], oracle.extract(v, 0.1, normal, dup), f"batched nx={nx} mode {normal},{dup}")
    server = kmc.Server(budget_bytes=1 << 30, workers=2)
    try:
        vid = server.register(kmc.make_volume(v))
        ids = [server.submit(vid, 0.1, normal, dup) for normal, dup in MODES]
        for rid, (normal, dup) in zip(ids, MODES):
            assert_same_surface(server.wait(rid), oracle.extract(v, 0.1, normal, dup), f"served nx={nx} mode {normal},{dup}")
    finally:
        server.close()
