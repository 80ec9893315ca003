"""ctypes binding of libkmc.so (include/kmc.h) for Python callers.

This is plumbing for tests and the benchmark: the product is the C ABI and
the HIP kernels behind it.  Loading fails loudly if the library is missing;
there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkmc.so")

# kvs::Type::TypeID order (Source/Core/Utility/Type.h:40-53)
INT8, INT16, INT32, INT64, UINT8, UINT16, UINT32, UINT64, REAL32, REAL64 = range(10)
NUMPY_TYPES = {
    np.dtype(np.int8): INT8, np.dtype(np.int16): INT16, np.dtype(np.int32): INT32, np.dtype(np.int64): INT64,
    np.dtype(np.uint8): UINT8, np.dtype(np.uint16): UINT16, np.dtype(np.uint32): UINT32,
    np.dtype(np.uint64): UINT64, np.dtype(np.float32): REAL32, np.dtype(np.float64): REAL64,
}
TYPE_NUMPY = {v: k for k, v in NUMPY_TYPES.items()}

VERTEX_NORMAL, POLYGON_NORMAL = 0, 1
MEMORY_HOST, MEMORY_DEVICE = 0, 1
STRUCTURED, UNSTRUCTURED = 0, 1
COLORMAP_DEFAULT, COLORMAP_GREYS = 0, 1

OK = 0
ERR_NULL_INPUT, ERR_NOT_STRUCTURED, ERR_NOT_SCALAR, ERR_UNSUPPORTED_TYPE = 1, 2, 3, 4
ERR_INVALID_ARGUMENT, ERR_OUT_OF_MEMORY, ERR_DEVICE, ERR_BUDGET, ERR_NOT_READY, ERR_NO_DEVICE = 5, 6, 7, 8, 9, 10


class Volume(ctypes.Structure):
    _fields_ = [
        ("values", ctypes.c_void_p),
        ("memory", ctypes.c_int32),
        ("type", ctypes.c_int32),
        ("resolution", ctypes.c_uint32 * 3),
        ("veclen", ctypes.c_uint32),
        ("kind", ctypes.c_int32),
        ("has_min_max_values", ctypes.c_int32),
        ("min_value", ctypes.c_double),
        ("max_value", ctypes.c_double),
        ("has_min_max_coords", ctypes.c_int32),
        ("min_coord", ctypes.c_float * 3),
        ("max_coord", ctypes.c_float * 3),
    ]


class Surface(ctypes.Structure):
    _fields_ = [
        ("isolevel", ctypes.c_double),
        ("n_vertices", ctypes.c_uint64),
        ("n_triangles", ctypes.c_uint64),
        ("n_connections", ctypes.c_uint64),
        ("n_normals", ctypes.c_uint64),
        ("coords", ctypes.POINTER(ctypes.c_float)),
        ("connections", ctypes.POINTER(ctypes.c_uint32)),
        ("normals", ctypes.POINTER(ctypes.c_float)),
        ("memory", ctypes.c_int32),
        ("polygon_type", ctypes.c_int32),
        ("normal_type", ctypes.c_int32),
        ("color_type", ctypes.c_int32),
        ("empty_range", ctypes.c_int32),
        ("min_value", ctypes.c_double),
        ("max_value", ctypes.c_double),
        ("min_coord", ctypes.c_float * 3),
        ("max_coord", ctypes.c_float * 3),
        ("internal", ctypes.c_void_p),
    ]


# Every symbol include/kmc.h declares (checked by the CPU tests).
EXPORTED = [
    "kmc_init", "kmc_status_string", "kmc_extract", "kmc_extract_levels", "kmc_surface_release",
    "kmc_volume_min_max", "kmc_surface_color", "kmc_colormap_table",
]

_lib = None


def load():
    """Load libkmc.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m kvs_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    lib.kmc_init.argtypes = [ctypes.c_int]
    lib.kmc_init.restype = ctypes.c_int
    lib.kmc_status_string.argtypes = [ctypes.c_int]
    lib.kmc_status_string.restype = ctypes.c_char_p
    lib.kmc_extract.argtypes = [ctypes.POINTER(Volume), ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(Surface)]
    lib.kmc_extract.restype = ctypes.c_int
    lib.kmc_extract_levels.argtypes = [ctypes.POINTER(Volume), ctypes.POINTER(ctypes.c_double), ctypes.c_uint32,
                                       ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                       ctypes.POINTER(Surface)]
    lib.kmc_extract_levels.restype = ctypes.c_int
    lib.kmc_surface_release.argtypes = [ctypes.POINTER(Surface)]
    lib.kmc_surface_release.restype = None
    lib.kmc_volume_min_max.argtypes = [ctypes.POINTER(Volume), ctypes.c_void_p,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.kmc_volume_min_max.restype = ctypes.c_int
    lib.kmc_surface_color.argtypes = [ctypes.c_int32, ctypes.c_uint32, ctypes.c_int32, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_uint8)]
    lib.kmc_surface_color.restype = ctypes.c_int
    lib.kmc_colormap_table.argtypes = [ctypes.c_int32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint8)]
    lib.kmc_colormap_table.restype = ctypes.c_int
    _lib = lib
    return lib


class KmcError(RuntimeError):
    def __init__(self, status):
        self.status = status
        msg = load().kmc_status_string(status).decode()
        super().__init__(f"kmc status {status}: {msg}")


def make_volume(values, dims=None, memory=MEMORY_HOST, min_max=None, coords=None, veclen=1, kind=STRUCTURED,
                dtype=None):
    """Build a kmc_volume.  values: a C-contiguous numpy array (host) laid out
    [z][y][x], or a device pointer (int) with dims and dtype given."""
    v = Volume()
    if memory == MEMORY_HOST:
        arr = np.ascontiguousarray(values)
        if dims is None:
            nz, ny, nx = arr.shape
            dims = (nx, ny, nz)
        v.values = arr.ctypes.data
        v.type = NUMPY_TYPES[arr.dtype]
        v._keep = arr
    else:
        v.values = int(values)
        v.type = NUMPY_TYPES[np.dtype(dtype)]
    v.memory = memory
    v.resolution[:] = [int(d) for d in dims]
    v.veclen = veclen
    v.kind = kind
    if min_max is not None:
        v.has_min_max_values = 1
        v.min_value, v.max_value = float(min_max[0]), float(min_max[1])
    if coords is not None:
        v.has_min_max_coords = 1
        v.min_coord[:] = [float(c) for c in coords[0]]
        v.max_coord[:] = [float(c) for c in coords[1]]
    return v


def surface_to_numpy(s):
    """Copy a host-memory surface into numpy arrays (coords (V,3), connections
    (T,3) or None, normals (N,3) or None) plus its metadata."""
    assert s.memory == MEMORY_HOST
    coords = np.ctypeslib.as_array(s.coords, shape=(int(s.n_vertices) * 3,)).copy() if s.n_vertices else np.zeros(0, np.float32)
    conn = np.ctypeslib.as_array(s.connections, shape=(int(s.n_connections),)).copy() if s.n_connections else np.zeros(0, np.uint32)
    normals = np.ctypeslib.as_array(s.normals, shape=(int(s.n_normals) * 3,)).copy() if s.n_normals else np.zeros(0, np.float32)
    return {
        "coords": coords, "connections": conn, "normals": normals,
        "n_triangles": int(s.n_triangles), "polygon_type": s.polygon_type, "normal_type": s.normal_type,
        "color_type": s.color_type, "empty_range": s.empty_range,
        "min_value": s.min_value, "max_value": s.max_value,
        "min_coord": list(s.min_coord), "max_coord": list(s.max_coord),
    }


def extract(values, isolevel, normal_type=POLYGON_NORMAL, duplication=True, min_max=None, coords=None):
    """Host-in, host-out extraction of one isolevel (kmc_extract)."""
    lib = load()
    v = make_volume(values, min_max=min_max, coords=coords)
    s = Surface()
    st = lib.kmc_extract(ctypes.byref(v), float(isolevel), int(normal_type), int(bool(duplication)),
                         MEMORY_HOST, None, ctypes.byref(s))
    if st != OK:
        raise KmcError(st)
    try:
        return surface_to_numpy(s)
    finally:
        lib.kmc_surface_release(ctypes.byref(s))


def extract_levels(values, isolevels, normal_type=POLYGON_NORMAL, duplication=True, min_max=None, coords=None):
    """Host-in, host-out extraction of several isolevels in one pass."""
    lib = load()
    v = make_volume(values, min_max=min_max, coords=coords)
    n = len(isolevels)
    levels = (ctypes.c_double * n)(*[float(x) for x in isolevels])
    outs = (Surface * n)()
    st = lib.kmc_extract_levels(ctypes.byref(v), levels, n, int(normal_type), int(bool(duplication)),
                                MEMORY_HOST, None, outs)
    if st != OK:
        raise KmcError(st)
    try:
        return [surface_to_numpy(outs[i]) for i in range(n)]
    finally:
        for i in range(n):
            lib.kmc_surface_release(ctypes.byref(outs[i]))


def min_max(values):
    lib = load()
    v = make_volume(values)
    mn, mx = ctypes.c_double(), ctypes.c_double()
    st = lib.kmc_volume_min_max(ctypes.byref(v), None, ctypes.byref(mn), ctypes.byref(mx))
    if st != OK:
        raise KmcError(st)
    return mn.value, mx.value


def surface_color(colormap, resolution, vtype, vmin, vmax, isolevel):
    lib = load()
    rgb = (ctypes.c_uint8 * 3)()
    st = lib.kmc_surface_color(int(colormap), int(resolution), int(vtype), float(vmin), float(vmax),
                               float(isolevel), rgb)
    if st != OK:
        raise KmcError(st)
    return tuple(rgb)


def colormap_table(colormap, resolution=256):
    lib = load()
    out = (ctypes.c_uint8 * (3 * resolution))()
    st = lib.kmc_colormap_table(int(colormap), int(resolution), out)
    if st != OK:
        raise KmcError(st)
    return np.frombuffer(bytes(out), dtype=np.uint8).copy()
