"""Readers of the golden fixtures in tests/golden/ (written by make_golden.py
from the reference compiled from its own sources)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODES = [(1, True), (1, False), (0, False)]


def meta():
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        return json.load(f)


class Packed:
    """Per-case slices of a packed fixture file."""

    def __init__(self, name):
        self.z = np.load(os.path.join(GOLDEN, name))

    def get(self, field, i):
        off = self.z[field + "_off"]
        return self.z[field][off[i]:off[i + 1]]


def cube_cases():
    """Yield (values, isolevel, {mode: (coords, connections, normals)}) for the
    2x2x2 brute-force set."""
    m = meta()["cubes"]
    p = Packed("cubes.npz")
    counters = {}
    for i, c in enumerate(m):
        dt = c["dtype"]
        j = counters.get(dt, 0)
        counters[dt] = j + 1
        v = p.get("in_" + dt, j).astype(np.dtype(dt)).reshape(2, 2, 2)
        outs = {}
        for normal, dup in MODES:
            mk = f"n{normal}_d{int(dup)}"
            outs[(normal, dup)] = (p.get(mk + "_coords", i).astype(np.float32),
                                   p.get(mk + "_conn", i).astype(np.uint32),
                                   p.get(mk + "_normals", i).astype(np.float32))
        yield v, c["iso"], c["cube_index"], outs


def random_cases():
    """Yield (values, isolevel, min_max, coords, {mode: (coords, conn, normals, colors)})."""
    m = meta()["random"]
    p = Packed("random.npz")
    for i, c in enumerate(m):
        dt = np.dtype(c["dtype"])
        v = np.frombuffer(p.get("in_bytes", i).tobytes(), dtype=dt).reshape(c["shape"])
        outs = {}
        for normal, dup in MODES:
            mk = f"n{normal}_d{int(dup)}"
            outs[(normal, dup)] = (p.get(mk + "_coords", i).astype(np.float32),
                                   p.get(mk + "_conn", i).astype(np.uint32),
                                   p.get(mk + "_normals", i).astype(np.float32),
                                   p.get(mk + "_colors", i).astype(np.uint8))
        mm = tuple(c["min_max"]) if c["min_max"] is not None else None
        co = c["coords"]
        yield v, c["iso"], mm, co, outs


def bigint_cases():
    """Yield (name, values, isolevel, min_max, {mode: (coords, conn, normals)}) for
    the 64-bit integer volumes whose values exceed 2^53."""
    m = meta()["bigint"]
    p = Packed("bigint.npz")
    for i, c in enumerate(m):
        dt = np.dtype(c["dtype"])
        v = np.frombuffer(p.get("in_bytes", i).tobytes(), dtype=dt).reshape(c["shape"])
        outs = {}
        for normal, dup in MODES:
            mk = f"n{normal}_d{int(dup)}"
            outs[(normal, dup)] = (p.get(mk + "_coords", i).astype(np.float32),
                                   p.get(mk + "_conn", i).astype(np.uint32),
                                   p.get(mk + "_normals", i).astype(np.float32))
        mm = tuple(c["min_max"]) if c["min_max"] is not None else None
        yield c["name"], v, c["iso"], mm, outs


def slice_cases():
    """Yield (case, values, (coords, normals, colors)) for the kvs::SlicePlane fixtures."""
    m = meta()["slice"]
    p = Packed("slice.npz")
    for i, c in enumerate(m):
        dt = np.dtype(c["dtype"])
        v = np.frombuffer(p.get("in_bytes", i).tobytes(), dtype=dt).reshape(c["shape"])
        yield c, v, (p.get("coords", i).astype(np.float32), p.get("normals", i).astype(np.float32),
                     p.get("colors", i).astype(np.uint8))


def unstructured_cases():
    """Yield (case, values, coords, connections, {mode: (coords, normals)},
    (slice coords, slice normals, slice colors)) for the unstructured fixtures."""
    m = meta()["unstructured"]
    p = Packed("unstructured.npz")
    for i, c in enumerate(m):
        v = np.frombuffer(p.get("values", i).tobytes(), dtype=np.dtype(c["dtype"]))
        co = p.get("coords", i).astype(np.float32).reshape(-1, 3)
        k = p.get("conn", i).astype(np.uint32)
        outs = {}
        for normal, dup in MODES:
            mk = f"n{normal}_d{int(dup)}"
            outs[(normal, dup)] = (p.get(mk + "_coords", i).astype(np.float32), p.get(mk + "_normals", i).astype(np.float32))
        sl = (p.get("slice_coords", i).astype(np.float32), p.get("slice_normals", i).astype(np.float32),
              p.get("slice_colors", i).astype(np.uint8))
        yield c, v, co, k, outs, sl


def hydrogen():
    return np.load(os.path.join(GOLDEN, "hydrogen64.npz"))["values"]


def sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
