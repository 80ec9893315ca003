"""Batched form (kmc_batch_*): every surface of a batch equals the plain call
on the same volume and isolevel, bit for bit -- across value types, uneven
sizes, modes, more levels than one pass holds, empty ranges, user min/max and
coordinates, repeated runs over changed values (arena growth), autotuning."""
import numpy as np
import pytest

import oracle
from kvs_amd import datasets, kmc
from tests.conftest import assert_same_surface

pytestmark = pytest.mark.gpu

MODES = [(kmc.POLYGON_NORMAL, True), (kmc.POLYGON_NORMAL, False), (kmc.VERTEX_NORMAL, False)]


def mixed_specs(seed=11):
    """(values, levels, normal, dup, min_max, coords) of a mixed batch."""
    rng = np.random.default_rng(seed)
    out = []
    shapes = [(17, 23, 70), (40, 9, 130), (2, 2, 2), (31, 31, 31), (5, 64, 65), (12, 33, 1), (3, 7, 200)]
    for i, shape in enumerate(shapes):
        for dtype in (np.float32, np.uint8, np.int16, np.float64, np.uint16):
            if dtype == np.float32:
                v = datasets.blobs(shape[::-1], seed=int(rng.integers(1 << 30)), nblobs=int(rng.integers(1, 8)))
                v = v.astype(np.float32)
                levels = [float(x) for x in rng.uniform(0.2, 0.8, int(rng.integers(1, 5)))]
            elif np.dtype(dtype).kind == "f":
                v = rng.standard_normal(shape)
                levels = [float(x) for x in rng.standard_normal(int(rng.integers(1, 4))) * 0.5]
            else:
                v = rng.integers(0, 90, shape).astype(dtype)
                levels = [float(rng.integers(5, 80)) + 0.5 * float(rng.integers(0, 2))
                          for _ in range(int(rng.integers(1, 4)))]
            normal, dup = MODES[(i + len(out)) % 3]
            mm = None
            co = None
            if len(out) % 7 == 3:
                lo, hi = float(v.min()), float(v.max())
                mm = (lo + 0.2 * (hi - lo), hi - 0.2 * (hi - lo))
            if len(out) % 5 == 2:
                co = ([-1.0, 0.5, 2.0], [3.0, 4.5, 9.0])
            out.append((np.ascontiguousarray(v.astype(dtype)), levels, normal, dup, mm, co))
    # a constant volume (empty range), a job with no levels, and one with more levels than a pass holds
    out.append((np.full((4, 5, 6), 7.0, np.float32), [7.0, 6.5], kmc.POLYGON_NORMAL, True, None, None))
    out.append((datasets.blobs(16, seed=3), [], kmc.VERTEX_NORMAL, False, None, None))
    out.append((datasets.blobs((20, 18, 90), seed=4, nblobs=5), list(np.linspace(0.1, 0.9, 11)),
                kmc.VERTEX_NORMAL, False, None, None))
    return out


def plain(values, level, normal, dup, mm, co):
    return kmc.extract(values, level, normal, dup, min_max=mm, coords=co)


def test_batch_extract_equals_plain_calls():
    specs = mixed_specs()
    print("batch of", len(specs), "jobs", flush=True)
    vols = [kmc.make_volume(v, min_max=mm, coords=co) for (v, _, _, _, mm, co) in specs]
    res = kmc.batch_extract([(vol, lv, n, d) for vol, (v, lv, n, d, mm, co) in zip(vols, specs)])
    total = 0
    for j, ((v, levels, normal, dup, mm, co), surfs) in enumerate(zip(specs, res)):
        assert len(surfs) == len(levels)
        for lv, got in zip(levels, surfs):
            want = plain(v, lv, normal, dup, mm, co)
            label = f"job {j} {v.dtype} {v.shape} level {lv} mode {normal},{dup}"
            assert_same_surface(got, want, label)
            assert got["empty_range"] == want["empty_range"], label
            assert got["min_value"] == want["min_value"] and got["max_value"] == want["max_value"], label
            assert got["min_coord"] == want["min_coord"] and got["max_coord"] == want["max_coord"], label
            total += got["n_triangles"]
    assert total > 10000


def test_batch_against_oracle_directly():
    rng = np.random.default_rng(5)
    specs = []
    for k in range(6):
        shape = tuple(int(s) for s in rng.integers(3, 40, 3))
        v = datasets.blobs(shape, seed=100 + k, nblobs=int(rng.integers(1, 16)))
        specs.append((v, [float(x) for x in rng.uniform(0.2, 0.8, int(rng.integers(1, 5)))], *MODES[k % 3]))
    res = kmc.batch_extract(specs)
    for (v, levels, normal, dup), surfs in zip(specs, res):
        for lv, got in zip(levels, surfs):
            assert_same_surface(got, oracle.extract(v, lv, normal, dup), f"{v.shape} {lv} {normal},{dup}")


def test_batch_object_reruns_follow_changed_values():
    rng = np.random.default_rng(21)
    arrays = [datasets.blobs(tuple(int(s) for s in rng.integers(8, 48, 3)), seed=200 + k, nblobs=2)
              for k in range(9)]
    devs = [kmc.DeviceArray(a) for a in arrays]
    levels = [[0.5], [0.3, 0.6], [0.25, 0.5, 0.75]]
    specs = [(d.volume(), levels[k % 3], *MODES[k % 3]) for k, d in enumerate(devs)]
    planned = [kmc.min_max(a) for a in arrays]      # taken once, at creation
    b = kmc.Batch(specs)
    try:
        for rnd in range(3):
            if rnd > 0:
                # new values in place: more blobs -> more triangles (forces arena growth at least once)
                for k, d in enumerate(devs):
                    nz, ny, nx = arrays[k].shape
                    arrays[k] = datasets.blobs((nx, ny, nz), seed=300 + 10 * rnd + k, nblobs=4 + 4 * rnd)
                    d.upload(arrays[k])
            print("run", rnd, flush=True)
            b.run()
            views = b.surfaces()
            k = 0
            for j, a in enumerate(arrays):
                normal, dup = MODES[j % 3]
                mn, mx = planned[j]
                for level in levels[j % 3]:
                    got = kmc.device_surface_to_numpy(views[k])
                    want = kmc.extract(a, level, normal, dup, min_max=(mn, mx))
                    assert_same_surface(got, want, f"run {rnd} job {j} level {level}")
                    k += 1
        info = b.info()
        assert info["n_surfaces"] == 18 and info["n_jobs"] == 9
        assert info["flat_emit"] == 1           # the default emit form
        assert info["graph_nodes"] > 0
        assert info["recaptures"] >= 2          # the sizing capture plus at least one growth
    finally:
        b.close()
        for d in devs:
            d.free()


def test_batch_min_max_is_planned_once():
    """Min/max are taken at creation (as the reference caches them in the
    volume); reruns clamp with those values even when the values change."""
    a = datasets.blobs(24, seed=8, nblobs=3)
    d = kmc.DeviceArray(a)
    mn, mx = kmc.min_max(a)
    b = kmc.Batch([(d.volume(), [0.4], kmc.POLYGON_NORMAL, True)])
    try:
        scaled = (a * np.float32(0.9)).astype(np.float32)
        d.upload(scaled)
        b.run()
        got = kmc.device_surface_to_numpy(b.surfaces()[0])
        want = kmc.extract(scaled, 0.4, kmc.POLYGON_NORMAL, True, min_max=(mn, mx))
        assert_same_surface(got, want, "cached min/max")
    finally:
        b.close()
        d.free()


def test_autotuned_batch_matches_plain():
    rng = np.random.default_rng(33)
    arrays = [datasets.blobs(tuple(int(s) for s in rng.integers(16, 96, 3)), seed=400 + k,
                             nblobs=int(rng.integers(1, 16))) for k in range(12)]
    specs = [(kmc.make_volume(a), [0.3, 0.55], kmc.POLYGON_NORMAL, True) for a in arrays]
    b = kmc.Batch(specs, autotune=True)
    try:
        info = b.info()
        print("autotuned", info, flush=True)
        assert info["tile_rows"] in (2, 4, 8, 16) and info["queue_capacity"] in (64, 128, 256)
        assert info["flat_emit"] in (0, 1)
        b.run()
        views = b.surfaces()
        for j, a in enumerate(arrays):
            for i, lv in enumerate((0.3, 0.55)):
                got = kmc.device_surface_to_numpy(views[2 * j + i])
                assert_same_surface(got, kmc.extract(a, lv, kmc.POLYGON_NORMAL, True), f"job {j} {lv}")
    finally:
        b.close()


def test_batch_rejects_bad_job_like_plain_call():
    a = datasets.blobs(8, seed=1)
    good = kmc.make_volume(a)
    bad = kmc.make_volume(a, veclen=3)
    with pytest.raises(kmc.KmcError) as e:
        kmc.batch_extract([(good, [0.5], kmc.POLYGON_NORMAL, True), (bad, [0.5], kmc.POLYGON_NORMAL, True)])
    assert e.value.status == kmc.ERR_NOT_SCALAR


def test_empty_batch():
    assert kmc.batch_extract([]) == []


def test_batch_on_tile_team_emit(monkeypatch):
    # the batched form's tile-team emit (KMC_EMIT=team at creation) equals the plain calls
    monkeypatch.setenv("KMC_EMIT", "team")
    specs = mixed_specs(seed=29)
    print("batch of", len(specs), "jobs on the tile teams", flush=True)
    vols = [kmc.make_volume(v, min_max=mm, coords=co) for (v, _, _, _, mm, co) in specs]
    res = kmc.batch_extract([(vol, lv, n, d) for vol, (v, lv, n, d, mm, co) in zip(vols, specs)])
    monkeypatch.delenv("KMC_EMIT")
    for j, ((v, levels, normal, dup, mm, co), surfs) in enumerate(zip(specs, res)):
        assert len(surfs) == len(levels)
        for lv, got in zip(levels, surfs):
            assert_same_surface(got, plain(v, lv, normal, dup, mm, co), f"team job {j} level {lv}")


def test_batch_wide_and_short_rows():
    # rows of more than 64 chunks (4100 nodes) next to very short rows in one
    # batch: the classify runs chunks across row ends and 64-chunk windows
    rng = np.random.default_rng(41)
    specs = [
        (rng.standard_normal((3, 5, 4100)).astype(np.float32), [0.1, -0.2], kmc.VERTEX_NORMAL, False),
        (rng.standard_normal((9, 7, 3)).astype(np.float32), [0.0], kmc.POLYGON_NORMAL, True),
        (rng.standard_normal((4, 6, 130)).astype(np.float32), [0.3], kmc.POLYGON_NORMAL, False),
        (rng.standard_normal((2, 3, 4097)).astype(np.float32), [0.05], kmc.POLYGON_NORMAL, True),
    ]
    vols = [kmc.make_volume(v) for (v, _, _, _) in specs]
    res = kmc.batch_extract([(vol, lv, n, d) for vol, (v, lv, n, d) in zip(vols, specs)])
    for j, ((v, levels, normal, dup), surfs) in enumerate(zip(specs, res)):
        for lv, got in zip(levels, surfs):
            assert_same_surface(got, oracle.extract(v, lv, normal, dup), f"wide job {j} level {lv}")


def test_batch_small_types_rows_of_whole_dwords_against_oracle():
    """8- and 16-bit jobs whose rows are whole dwords classify a dword of
    nodes per lane inside the batch classification too (rows of one and of
    several 64-dword segments, and rows that are not whole dwords beside
    them); more than four levels take the wide instance."""
    rng = np.random.default_rng(77)
    specs = []
    shapes = [(6, 9, 256), (5, 7, 260), (3, 4, 1024), (4, 6, 66), (5, 5, 63), (7, 3, 512)]
    for k, shape in enumerate(shapes):
        dtype = (np.uint8, np.int8, np.uint16, np.int16)[k % 4]
        v = rng.integers(0, 90, shape).astype(dtype)
        nlev = 6 if k == 2 else int(rng.integers(1, 4))
        levels = [float(rng.integers(5, 80)) + 0.5 * float(rng.integers(0, 2)) for _ in range(nlev)]
        specs.append((v, levels, *MODES[k % 3]))
    res = kmc.batch_extract(specs)
    for (v, levels, normal, dup), surfs in zip(specs, res):
        for lv, got in zip(levels, surfs):
            assert_same_surface(got, oracle.extract(v, lv, normal, dup), f"{v.dtype} {v.shape} {lv} {normal},{dup}")
