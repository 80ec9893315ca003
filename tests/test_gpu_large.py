"""A volume of more than 2^32 nodes, checked through properties that hold at
any size: every node, row, chunk and record index past 32 bits.

The volume is a step between two z-slices that lie past node 2^32, so its
isosurface is a flat sheet: two triangles per cell of the sheet, one shared
vertex per z-edge in node scan order, every vertex at the interpolated height,
every face normal along z, and the same per-cell connection pattern the
oracle gives for the same step in a small volume."""
import numpy as np
import pytest

import oracle
from kvs_amd import kmc

pytestmark = pytest.mark.gpu

NX, NY, NZ, K = 4096, 1024, 1040, 1030   # 4.36e9 nodes; the sheet lies between slices K and K+1


def step_volume(nx, ny, nz, k):
    v = np.zeros((nz, ny, nx), np.uint8)
    v[k + 1:] = 200
    return v


def cell_pattern(conn, nx):
    """Corner offsets (dx, dy) of the first cell's two triangles."""
    ids = conn[:6].astype(np.int64)
    return [(int(i % nx), int(i // nx)) for i in ids]


def test_step_sheet_past_2_32_nodes():
    assert NX * NY * K > 2 ** 32
    v = step_volume(NX, NY, NZ, K)
    small = step_volume(3, 3, 3, 0)
    want_small = oracle.extract(small, 100.0, oracle.VERTEX_NORMAL, False)
    ncell = (NX - 1) * (NY - 1)

    print("case shared vertices, VertexNormal", v.shape, flush=True)
    got = kmc.extract(v, 100.0, kmc.VERTEX_NORMAL, False, min_max=(0.0, 200.0))
    assert got["n_triangles"] == 2 * ncell
    coords = got["coords"].reshape(-1, 3)
    assert coords.shape[0] == NX * NY
    i = np.arange(NX * NY)
    assert np.array_equal(coords[:, 0], (i % NX).astype(np.float32))
    assert np.array_equal(coords[:, 1], (i // NX).astype(np.float32))
    assert np.all(coords[:, 2] == np.float32(K + 0.5))
    conn = got["connections"].reshape(-1, 2, 3).astype(np.int64)
    c = np.arange(ncell)
    base = (c % (NX - 1)) + (c // (NX - 1)) * NX
    rel = conn - base[:, None, None]
    assert np.array_equal(rel, np.broadcast_to(rel[0], rel.shape))
    assert cell_pattern(got["connections"], NX) == cell_pattern(want_small["connections"], 3)
    normals = got["normals"].reshape(-1, 3)
    assert np.all(normals[:, :2] == 0)
    sign = np.sign(normals[0, 2])
    assert sign != 0 and np.all(np.sign(normals[:, 2]) == sign)
    assert np.sign(want_small["normals"].reshape(-1, 3)[0, 2]) == sign
    del got, coords, conn, rel, normals

    print("case duplicated vertices, PolygonNormal", v.shape, flush=True)
    got = kmc.extract(v, 100.0, kmc.POLYGON_NORMAL, True, min_max=(0.0, 200.0))
    assert got["n_triangles"] == 2 * ncell
    coords = got["coords"].reshape(-1, 3)
    assert coords.shape[0] == 6 * ncell
    assert np.all(coords[:, 2] == np.float32(K + 0.5))
    # the last cell's corners: x in {NX-2, NX-1}, y in {NY-2, NY-1}
    last = coords[-6:]
    assert set(last[:, 0].tolist()) == {NX - 2.0, NX - 1.0} and set(last[:, 1].tolist()) == {NY - 2.0, NY - 1.0}
    normals = got["normals"].reshape(-1, 3)
    assert np.all(normals[:, :2] == 0) and np.all(np.abs(normals[:, 2]) == 1)
