"""kmc_extract_noalloc (no counterpart in the reference; for callers that
cannot allocate device memory) against kmc_extract, bit for bit, through the
C ABI: device volume in, caller-owned device outputs, every work array in the
library's fixed scratch.  Two calls in a row, the second on a smaller volume,
reuse the same scratch."""
import ctypes

import numpy as np
import pytest

import oracle
from kvs_amd import datasets, kmc
from tests.conftest import assert_same_surface

MODES = [(kmc.POLYGON_NORMAL, True), (kmc.POLYGON_NORMAL, False), (kmc.VERTEX_NORMAL, False)]


class DeviceBuffer:
    def __init__(self, nbytes):
        self.hip = kmc._hip()
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), max(1, int(nbytes))) == 0

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def plain_device(values, iso, normal, dup, **kw):
    """kmc_extract on the same device volume, device outputs."""
    lib = kmc.load()
    d = kmc.DeviceArray(values)
    try:
        s = kmc.Surface()
        st = lib.kmc_extract(ctypes.byref(d.volume(**kw)), float(iso), normal, int(dup), kmc.MEMORY_DEVICE, None,
                             ctypes.byref(s))
        assert st == kmc.OK, st
        try:
            return kmc.device_surface_to_numpy(s)
        finally:
            lib.kmc_surface_release(ctypes.byref(s))
    finally:
        d.free()


def noalloc(values, iso, normal, dup, cap_scale=1.0, **kw):
    """kmc_extract_noalloc with outputs sized from the plain call (times
    cap_scale); returns (status, surface dict or None, sizes)."""
    lib = kmc.load()
    want = plain_device(values, iso, normal, dup, **kw)
    nc = int(want["coords"].size * cap_scale)
    nk = int(want["connections"].size * cap_scale)
    nn = int(want["normals"].size * cap_scale)
    d = kmc.DeviceArray(values)
    bufs = [DeviceBuffer(4 * max(1, n)) for n in (nc, nk, nn)]
    try:
        s = kmc.Surface()
        st = lib.kmc_extract_noalloc(ctypes.byref(d.volume(**kw)), float(iso), normal, int(dup), None,
                                     bufs[0].ptr, nc, bufs[1].ptr, nk, bufs[2].ptr, nn, ctypes.byref(s))
        sizes = (int(s.n_vertices), int(s.n_triangles), int(s.n_connections), int(s.n_normals))
        got = kmc.device_surface_to_numpy(s) if st == kmc.OK else None
        return st, got, want, sizes
    finally:
        for b in bufs:
            b.free()
        d.free()


@pytest.mark.gpu
def test_noalloc_equals_plain_call_twice_second_smaller():
    big = datasets.blobs(72, seed=31, nblobs=6)
    small = datasets.turbulence((40, 23, 17), seed=6)
    for normal, dup in MODES:
        for values, iso in ((big, 0.45), (small, 0.1)):
            label = f"noalloc {values.shape} mode {normal},{dup}"
            print("case", label, flush=True)
            st, got, want, _ = noalloc(values, iso, normal, dup)
            assert st == kmc.OK, (label, st)
            assert_same_surface(got, want, label)
            assert got["n_triangles"] > 100
            assert_same_surface(got, oracle.extract(values, iso, normal, dup), label + " (oracle)")


@pytest.mark.gpu
def test_noalloc_other_types_min_max_and_coords():
    rng = np.random.default_rng(9)
    v = rng.integers(0, 200, (19, 22, 31)).astype(np.uint16)
    for kw in ({}, {"min_max": (20.0, 150.0), "coords": ([1.0, 2.0, -3.0], [5.0, 9.0, 4.0])}):
        for normal, dup in MODES:
            st, got, want, _ = noalloc(v, 99.5, normal, dup, **kw)
            assert st == kmc.OK
            assert_same_surface(got, want, f"uint16 {kw} mode {normal},{dup}")


@pytest.mark.gpu
def test_noalloc_reports_sizes_when_outputs_are_too_small():
    v = datasets.blobs(40, seed=4)
    st, got, want, sizes = noalloc(v, 0.5, kmc.VERTEX_NORMAL, False, cap_scale=0.5)
    assert st == kmc.ERR_CAPACITY and got is None
    assert sizes[0] * 3 == want["coords"].size and sizes[1] == want["n_triangles"]
    # the next call with room succeeds on the same scratch
    st, got, want, _ = noalloc(v, 0.5, kmc.VERTEX_NORMAL, False)
    assert st == kmc.OK
    assert_same_surface(got, want, "after a refused call")


@pytest.mark.gpu
def test_noalloc_empty_range_and_no_surface():
    st, got, want, sizes = noalloc(np.full((6, 7, 8), 2.0, np.float32), 2.0, kmc.POLYGON_NORMAL, True)
    assert st == kmc.OK and got["empty_range"] == 1 and sizes == (0, 0, 0, 0)
    v = datasets.blobs(20, seed=2)
    st, got, want, sizes = noalloc(v, 5.0, kmc.POLYGON_NORMAL, False)   # level above every value
    assert st == kmc.OK and sizes == (0, 0, 0, 0)
