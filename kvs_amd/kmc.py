"""ctypes binding of libkmc.so (include/kmc.h) for Python callers.

This is plumbing for tests and the benchmark: the product is the C ABI and
the HIP kernels behind it.  Loading fails loudly if the library is missing;
there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# KMC_LIBRARY names another build of the library (a different install or a
# build of modified sources); by default the in-tree build.
LIB_PATH = os.environ.get("KMC_LIBRARY") or os.path.join(HERE, "libkmc.so")

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
ERR_SCRATCH, ERR_CAPACITY, ERR_CELL_TYPE, ERR_FILE, ERR_NOT_TRIANGLES = 11, 12, 13, 14, 15
TETRAHEDRA, HEXAHEDRA, PYRAMID, PRISM = 1, 2, 5, 7


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
        ("colors", ctypes.POINTER(ctypes.c_uint8)),
        ("n_colors", ctypes.c_uint64),
    ]


class UnstructuredVolume(ctypes.Structure):
    _fields_ = [
        ("values", ctypes.c_void_p),
        ("coords", ctypes.c_void_p),
        ("connections", ctypes.c_void_p),
        ("memory", ctypes.c_int32),
        ("type", ctypes.c_int32),
        ("n_nodes", ctypes.c_uint64),
        ("n_cells", ctypes.c_uint64),
        ("cell_type", ctypes.c_int32),
        ("veclen", ctypes.c_uint32),
        ("has_min_max_values", ctypes.c_int32),
        ("min_value", ctypes.c_double),
        ("max_value", ctypes.c_double),
        ("has_min_max_coords", ctypes.c_int32),
        ("min_coord", ctypes.c_float * 3),
        ("max_coord", ctypes.c_float * 3),
    ]


class Job(ctypes.Structure):
    _fields_ = [
        ("volume", Volume),
        ("isolevels", ctypes.POINTER(ctypes.c_double)),
        ("n_levels", ctypes.c_uint32),
        ("normal_type", ctypes.c_int32),
        ("duplication", ctypes.c_int32),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("n_jobs", ctypes.c_uint32),
        ("n_surfaces", ctypes.c_uint32),
        ("n_passes", ctypes.c_uint32),
        ("n_type_groups", ctypes.c_uint32),
        ("tile_rows", ctypes.c_uint32),
        ("queue_capacity", ctypes.c_uint32),
        ("graph_nodes", ctypes.c_uint32),
        ("recaptures", ctypes.c_uint32),
        ("nodes_classified", ctypes.c_uint64),
        ("n_triangles", ctypes.c_uint64),
        ("arena_bytes", ctypes.c_uint64),
        ("work_bytes", ctypes.c_uint64),
        ("last_run_ms", ctypes.c_double),
        ("flat_emit", ctypes.c_uint32),
    ]


BATCH_AUTOTUNE = 1


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("workers", ctypes.c_uint32),
        ("max_inflight", ctypes.c_uint32),
        ("max_batch", ctypes.c_uint32),
        ("budget_bytes", ctypes.c_uint64),
        ("out_memory", ctypes.c_int32),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("failed", ctypes.c_uint64),
        ("batches", ctypes.c_uint64),
        ("triangles", ctypes.c_uint64),
        ("peak_inflight", ctypes.c_uint32),
        ("deferred", ctypes.c_uint32),
        ("peak_budget_bytes", ctypes.c_uint64),
        ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double),
        ("mean_ms", ctypes.c_double),
        ("max_ms", ctypes.c_double),
        ("span_s", ctypes.c_double),
        ("triangles_per_s", ctypes.c_double),
        ("peak_batch_device_bytes", ctypes.c_uint64),
    ]

# Every symbol include/kmc.h declares (checked by the CPU tests).
EXPORTED = [
    "kmc_init", "kmc_status_string", "kmc_extract", "kmc_extract_levels", "kmc_surface_release",
    "kmc_volume_min_max", "kmc_surface_color", "kmc_colormap_table",
    "kmc_batch_create", "kmc_batch_run", "kmc_batch_surfaces", "kmc_batch_info_get", "kmc_batch_destroy",
    "kmc_batch_extract",
    "kmc_server_create", "kmc_server_register_volume", "kmc_server_submit", "kmc_server_poll", "kmc_server_wait",
    "kmc_server_get_stats", "kmc_server_reset_stats", "kmc_server_destroy",
    "kmc_set_timing", "kmc_last_timing", "kmc_extract_noalloc", "kmc_slice_plane",
    "kmc_extract_unstructured", "kmc_slice_plane_unstructured",
    "kmc_kvsml_read_volume", "kmc_avs_read_volume", "kmc_kvsml_volume_release",
    "kmc_surface_expand", "kmc_export_stl", "kmc_export_kvsml", "kmc_export_ply",
]


class Timing(ctypes.Structure):
    _fields_ = [
        ("classify_ms", ctypes.c_double),
        ("count_ms", ctypes.c_double),
        ("scan_ms", ctypes.c_double),
        ("emit_ms", ctypes.c_double),
        ("passes", ctypes.c_uint32),
        ("compact_ms", ctypes.c_double),
        ("vertices_ms", ctypes.c_double),
        ("triangles_ms", ctypes.c_double),
        ("vertex_normals_ms", ctypes.c_double),
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
    lib.kmc_extract_noalloc.argtypes = [ctypes.POINTER(Volume), ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                        ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(Surface)]
    lib.kmc_extract_noalloc.restype = ctypes.c_int
    lib.kmc_slice_plane.argtypes = [ctypes.POINTER(Volume), ctypes.POINTER(ctypes.c_float),
                                    ctypes.POINTER(ctypes.c_uint8), ctypes.c_uint32, ctypes.POINTER(ctypes.c_float),
                                    ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(Surface)]
    lib.kmc_slice_plane.restype = ctypes.c_int
    lib.kmc_extract_unstructured.argtypes = [ctypes.POINTER(UnstructuredVolume), ctypes.c_double, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(Surface)]
    lib.kmc_extract_unstructured.restype = ctypes.c_int
    lib.kmc_slice_plane_unstructured.argtypes = [ctypes.POINTER(UnstructuredVolume), ctypes.POINTER(ctypes.c_float),
                                                 ctypes.POINTER(ctypes.c_uint8), ctypes.c_uint32,
                                                 ctypes.POINTER(ctypes.c_float), ctypes.c_int32, ctypes.c_void_p,
                                                 ctypes.POINTER(Surface)]
    lib.kmc_slice_plane_unstructured.restype = ctypes.c_int
    lib.kmc_kvsml_read_volume.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(Volume)]
    lib.kmc_kvsml_read_volume.restype = ctypes.c_int
    lib.kmc_avs_read_volume.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(Volume)]
    lib.kmc_avs_read_volume.restype = ctypes.c_int
    lib.kmc_kvsml_volume_release.argtypes = [ctypes.POINTER(Volume)]
    lib.kmc_kvsml_volume_release.restype = None
    lib.kmc_surface_expand.argtypes = [ctypes.POINTER(Surface), ctypes.c_int32, ctypes.c_void_p,
                                       ctypes.POINTER(Surface)]
    lib.kmc_surface_expand.restype = ctypes.c_int
    lib.kmc_export_stl.argtypes = [ctypes.POINTER(Surface), ctypes.c_char_p, ctypes.c_int32, ctypes.c_void_p]
    lib.kmc_export_stl.restype = ctypes.c_int
    lib.kmc_export_kvsml.argtypes = [ctypes.POINTER(Surface), ctypes.POINTER(ctypes.c_uint8), ctypes.c_int32,
                                     ctypes.c_char_p, ctypes.c_int32, ctypes.c_void_p]
    lib.kmc_export_kvsml.restype = ctypes.c_int
    lib.kmc_export_ply.argtypes = [ctypes.POINTER(Surface), ctypes.POINTER(ctypes.c_uint8), ctypes.c_char_p,
                                   ctypes.c_int32, ctypes.c_void_p]
    lib.kmc_export_ply.restype = ctypes.c_int
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
    lib.kmc_batch_create.argtypes = [ctypes.POINTER(Job), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                     ctypes.POINTER(ctypes.c_void_p)]
    lib.kmc_batch_create.restype = ctypes.c_int
    lib.kmc_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.kmc_batch_run.restype = ctypes.c_int
    lib.kmc_batch_surfaces.argtypes = [ctypes.c_void_p, ctypes.POINTER(Surface), ctypes.c_uint32]
    lib.kmc_batch_surfaces.restype = ctypes.c_int
    lib.kmc_batch_info_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(BatchInfo)]
    lib.kmc_batch_info_get.restype = ctypes.c_int
    lib.kmc_batch_destroy.argtypes = [ctypes.c_void_p]
    lib.kmc_batch_destroy.restype = None
    lib.kmc_batch_extract.argtypes = [ctypes.POINTER(Job), ctypes.c_uint32, ctypes.c_int32, ctypes.c_void_p,
                                      ctypes.POINTER(Surface)]
    lib.kmc_batch_extract.restype = ctypes.c_int
    lib.kmc_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(ctypes.c_void_p)]
    lib.kmc_server_create.restype = ctypes.c_int
    lib.kmc_server_register_volume.argtypes = [ctypes.c_void_p, ctypes.POINTER(Volume), ctypes.POINTER(ctypes.c_uint32)]
    lib.kmc_server_register_volume.restype = ctypes.c_int
    lib.kmc_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double, ctypes.c_int32,
                                      ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64)]
    lib.kmc_server_submit.restype = ctypes.c_int
    lib.kmc_server_poll.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(Surface)]
    lib.kmc_server_poll.restype = ctypes.c_int
    lib.kmc_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(Surface)]
    lib.kmc_server_wait.restype = ctypes.c_int
    lib.kmc_server_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServerStats)]
    lib.kmc_server_get_stats.restype = ctypes.c_int
    lib.kmc_server_reset_stats.argtypes = [ctypes.c_void_p]
    lib.kmc_server_reset_stats.restype = ctypes.c_int
    lib.kmc_set_timing.argtypes = [ctypes.c_int32]
    lib.kmc_set_timing.restype = None
    lib.kmc_last_timing.argtypes = [ctypes.POINTER(Timing)]
    lib.kmc_last_timing.restype = ctypes.c_int
    lib.kmc_server_destroy.argtypes = [ctypes.c_void_p]
    lib.kmc_server_destroy.restype = None
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
    colors = (np.ctypeslib.as_array(s.colors, shape=(int(s.n_colors) * 3,)).copy() if s.n_colors and s.colors
              else np.zeros(0, np.uint8))
    return {
        "coords": coords, "connections": conn, "normals": normals, "colors": colors,
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


def slice_plane(values, coefficients, color_table, color_range=None, min_max=None, coords=None,
                out_memory=MEMORY_HOST):
    """kmc_slice_plane: kvs::SlicePlane over a host numpy volume [z][y][x];
    color_table: the colour map's table (3 bytes per entry)."""
    lib = load()
    v = make_volume(values, min_max=min_max, coords=coords)
    c = (ctypes.c_float * 4)(*[float(x) for x in coefficients])
    t = np.ascontiguousarray(color_table, dtype=np.uint8)
    rng = (ctypes.c_float * 2)(*[float(x) for x in color_range]) if color_range is not None else None
    s = Surface()
    st = lib.kmc_slice_plane(ctypes.byref(v), c, t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), t.size // 3,
                             rng, out_memory, None, ctypes.byref(s))
    if st != OK:
        raise KmcError(st)
    try:
        return surface_to_numpy(s) if s.memory == MEMORY_HOST else device_surface_to_numpy(s)
    finally:
        lib.kmc_surface_release(ctypes.byref(s))


CELL_NODES = {TETRAHEDRA: 4, HEXAHEDRA: 8, PYRAMID: 5, PRISM: 6}


def make_unstructured_volume(values, coords, connections, cell_type, min_max=None, min_max_coords=None, veclen=1):
    """A host kmc_unstructured_volume over numpy arrays (kept alive by the struct)."""
    v = UnstructuredVolume()
    vals = np.ascontiguousarray(values).ravel()
    co = np.ascontiguousarray(coords, dtype=np.float32).ravel()
    cn = np.ascontiguousarray(connections, dtype=np.uint32).ravel()
    v.values, v.coords, v.connections = vals.ctypes.data, co.ctypes.data, cn.ctypes.data
    v._keep = (vals, co, cn)
    v.memory = MEMORY_HOST
    v.type = NUMPY_TYPES[vals.dtype]
    v.n_nodes = vals.size // max(1, veclen)
    v.n_cells = cn.size // CELL_NODES.get(int(cell_type), 4)
    v.cell_type = int(cell_type)
    v.veclen = veclen
    if min_max is not None:
        v.has_min_max_values = 1
        v.min_value, v.max_value = float(min_max[0]), float(min_max[1])
    if min_max_coords is not None:
        v.has_min_max_coords = 1
        v.min_coord[:] = [float(c) for c in min_max_coords[0]]
        v.max_coord[:] = [float(c) for c in min_max_coords[1]]
    return v


def extract_unstructured(values, coords, connections, cell_type, isolevel, normal_type=POLYGON_NORMAL,
                         duplication=True, min_max=None, min_max_coords=None, out_memory=MEMORY_HOST):
    """kmc_extract_unstructured on host numpy arrays."""
    lib = load()
    v = make_unstructured_volume(values, coords, connections, cell_type, min_max, min_max_coords)
    s = Surface()
    st = lib.kmc_extract_unstructured(ctypes.byref(v), float(isolevel), int(normal_type), int(bool(duplication)),
                                      out_memory, None, ctypes.byref(s))
    if st != OK:
        raise KmcError(st)
    try:
        return surface_to_numpy(s) if s.memory == MEMORY_HOST else device_surface_to_numpy(s)
    finally:
        lib.kmc_surface_release(ctypes.byref(s))


def slice_plane_unstructured(values, coords, connections, cell_type, coefficients, color_table, color_range=None,
                             min_max=None, out_memory=MEMORY_HOST):
    """kmc_slice_plane_unstructured on host numpy arrays."""
    lib = load()
    v = make_unstructured_volume(values, coords, connections, cell_type, min_max)
    c = (ctypes.c_float * 4)(*[float(x) for x in coefficients])
    t = np.ascontiguousarray(color_table, dtype=np.uint8)
    rng = (ctypes.c_float * 2)(*[float(x) for x in color_range]) if color_range is not None else None
    s = Surface()
    st = lib.kmc_slice_plane_unstructured(ctypes.byref(v), c, t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                          t.size // 3, rng, out_memory, None, ctypes.byref(s))
    if st != OK:
        raise KmcError(st)
    try:
        return surface_to_numpy(s) if s.memory == MEMORY_HOST else device_surface_to_numpy(s)
    finally:
        lib.kmc_surface_release(ctypes.byref(s))


class KvsmlVolume:
    """A volume read by kmc_kvsml_read_volume (or, for an AVS field file .fld,
    kmc_avs_read_volume); .volume is the kmc_volume (values in host or device
    memory, owned until close())."""

    def __init__(self, filename, memory=MEMORY_DEVICE):
        self.lib = load()
        self.volume = Volume()
        read = self.lib.kmc_avs_read_volume if str(filename).lower().endswith(".fld") else self.lib.kmc_kvsml_read_volume
        st = read(str(filename).encode(), memory, None, ctypes.byref(self.volume))
        if st != OK:
            raise KmcError(st)

    def values(self):
        """The value array as numpy [z][y][x] (copied to the host)."""
        v = self.volume
        dtype = [k for k, t in NUMPY_TYPES.items() if t == v.type][0]
        n = int(v.resolution[0]) * int(v.resolution[1]) * int(v.resolution[2]) * int(v.veclen)
        out = np.zeros(n, dtype=dtype)
        if v.memory == MEMORY_HOST:
            ctypes.memmove(out.ctypes.data, v.values, out.nbytes)
        elif n:
            rc = _hip().hipMemcpy(out.ctypes.data, ctypes.c_void_p(v.values), out.nbytes, 2)
            if rc != 0:
                raise RuntimeError(f"hipMemcpy failed: {rc}")
        shape = (int(v.resolution[2]), int(v.resolution[1]), int(v.resolution[0])) + ((int(v.veclen),) if v.veclen > 1 else ())
        return out.reshape(shape)

    def close(self):
        if self.volume.values:
            self.lib.kmc_kvsml_volume_release(ctypes.byref(self.volume))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def host_surface(d, normal_type, polygon_type=3):
    """A host kmc_surface over numpy arrays (coords, connections, normals of a
    surface dict as extract returns); kept alive by the struct."""
    s = Surface()
    co = np.ascontiguousarray(d["coords"], dtype=np.float32)
    cn = np.ascontiguousarray(d.get("connections", np.zeros(0, np.uint32)), dtype=np.uint32)
    no = np.ascontiguousarray(d["normals"], dtype=np.float32)
    s._keep = (co, cn, no)
    s.coords = co.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    s.connections = cn.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    s.normals = no.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    s.n_vertices = co.size // 3
    s.n_connections = cn.size
    s.n_normals = no.size // 3
    s.n_triangles = cn.size // 3 if cn.size else co.size // 9
    s.memory = MEMORY_HOST
    s.polygon_type = polygon_type
    s.normal_type = normal_type
    s.color_type = 1
    return s


def surface_expand(d, normal_type, out_memory=MEMORY_HOST):
    """kmc_surface_expand of a host surface dict."""
    lib = load()
    s = host_surface(d, normal_type)
    o = Surface()
    st = lib.kmc_surface_expand(ctypes.byref(s), out_memory, None, ctypes.byref(o))
    if st != OK:
        raise KmcError(st)
    try:
        return surface_to_numpy(o) if o.memory == MEMORY_HOST else device_surface_to_numpy(o)
    finally:
        lib.kmc_surface_release(ctypes.byref(o))


def export_stl(d, normal_type, filename, ascii=False):
    """kmc_export_stl of a host surface dict."""
    lib = load()
    s = host_surface(d, normal_type)
    st = lib.kmc_export_stl(ctypes.byref(s), str(filename).encode(), int(bool(ascii)), None)
    if st != OK:
        raise KmcError(st)


KVSML_ASCII, KVSML_EXTERNAL_ASCII, KVSML_EXTERNAL_BINARY = range(3)


def export_kvsml(d, normal_type, filename, writing=KVSML_ASCII, color=None, opacity=255):
    """kmc_export_kvsml of a host surface dict.  A dict with per-vertex
    "colors" (kmc.slice_plane's) is a VertexColor polygon, else a PolygonColor
    one with the single `color` (3 bytes, or None)."""
    lib = load()
    s = host_surface(d, normal_type)
    cl = np.ascontiguousarray(d.get("colors", np.zeros(0, np.uint8)), dtype=np.uint8)
    if cl.size and cl.size == s.n_vertices * 3:
        s._keep = s._keep + (cl,)
        s.colors = cl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        s.n_colors = cl.size // 3
        s.color_type = 0
    c = (ctypes.c_uint8 * 3)(*color) if color is not None else None
    st = lib.kmc_export_kvsml(ctypes.byref(s), c, int(opacity), str(filename).encode(), int(writing), None)
    if st != OK:
        raise KmcError(st)


def export_ply(d, normal_type, filename, binary=False, color=None):
    """kmc_export_ply of a host surface dict (per-vertex "colors" when it has
    them, else the single `color`)."""
    lib = load()
    s = host_surface(d, normal_type)
    cl = np.ascontiguousarray(d.get("colors", np.zeros(0, np.uint8)), dtype=np.uint8)
    if cl.size and cl.size == s.n_vertices * 3:
        s._keep = s._keep + (cl,)
        s.colors = cl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        s.n_colors = cl.size // 3
        s.color_type = 0
    c = (ctypes.c_uint8 * 3)(*color) if color is not None else None
    st = lib.kmc_export_ply(ctypes.byref(s), c, str(filename).encode(), int(bool(binary)), None)
    if st != OK:
        raise KmcError(st)


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


def make_jobs(specs):
    """specs: list of (kmc_volume, isolevels, normal_type, duplication).
    Returns a ctypes Job array (keeping the level arrays alive)."""
    jobs = (Job * len(specs))()
    keep = []
    for i, (vol, levels, normal_type, dup) in enumerate(specs):
        lv = (ctypes.c_double * max(1, len(levels)))(*[float(x) for x in levels])
        keep.append((vol, lv))
        jobs[i].volume = vol
        jobs[i].isolevels = ctypes.cast(lv, ctypes.POINTER(ctypes.c_double))
        jobs[i].n_levels = len(levels)
        jobs[i].normal_type = int(normal_type)
        jobs[i].duplication = int(bool(dup))
    jobs._keep = keep
    return jobs


def batch_extract(specs):
    """Host-out one-shot batch (kmc_batch_extract); specs as for make_jobs with
    host numpy volumes or kmc_volume structs.  Returns one list of surfaces per job."""
    lib = load()
    vols = [sp[0] if isinstance(sp[0], Volume) else make_volume(sp[0]) for sp in specs]
    jobs = make_jobs([(v,) + tuple(sp[1:]) for v, sp in zip(vols, specs)])
    n = sum(len(sp[1]) for sp in specs)
    outs = (Surface * max(1, n))()
    st = lib.kmc_batch_extract(jobs, len(specs), MEMORY_HOST, None, outs)
    if st != OK:
        raise KmcError(st)
    try:
        res, k = [], 0
        for sp in specs:
            res.append([surface_to_numpy(outs[k + i]) for i in range(len(sp[1]))])
            k += len(sp[1])
        return res
    finally:
        for i in range(n):
            lib.kmc_surface_release(ctypes.byref(outs[i]))


class Batch:
    """A planned batch (kmc_batch_create) that can be run repeatedly."""

    def __init__(self, specs, autotune=False, stream=None):
        self.lib = load()
        self.jobs = make_jobs(specs)
        self.n_surfaces = sum(len(sp[1]) for sp in specs)
        h = ctypes.c_void_p()
        st = self.lib.kmc_batch_create(self.jobs, len(specs), BATCH_AUTOTUNE if autotune else 0, stream,
                                       ctypes.byref(h))
        if st != OK:
            raise KmcError(st)
        self.handle = h

    def run(self, stream=None):
        st = self.lib.kmc_batch_run(self.handle, stream)
        if st != OK:
            raise KmcError(st)

    def surfaces(self):
        outs = (Surface * max(1, self.n_surfaces))()
        st = self.lib.kmc_batch_surfaces(self.handle, outs, self.n_surfaces)
        if st != OK:
            raise KmcError(st)
        return [outs[i] for i in range(self.n_surfaces)]

    def info(self):
        i = BatchInfo()
        st = self.lib.kmc_batch_info_get(self.handle, ctypes.byref(i))
        if st != OK:
            raise KmcError(st)
        return {f: getattr(i, f) for f, _ in BatchInfo._fields_}

    def close(self):
        if self.handle:
            self.lib.kmc_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_hip_rt = None


def _hip():
    """The HIP runtime (libamdhip64), for copies of device-memory surfaces."""
    global _hip_rt
    if _hip_rt is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip_rt = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip_rt is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip_rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip_rt.hipMemcpy.restype = ctypes.c_int
        _hip_rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip_rt.hipMalloc.restype = ctypes.c_int
        _hip_rt.hipFree.argtypes = [ctypes.c_void_p]
        _hip_rt.hipFree.restype = ctypes.c_int
        _hip_rt.hipDeviceSynchronize.argtypes = []
        _hip_rt.hipDeviceSynchronize.restype = ctypes.c_int
    return _hip_rt


class DeviceArray:
    """A numpy array copied into device memory (hipMalloc), for device-resident
    volumes without a framework.  Only the HIP runtime libkmc itself uses is
    involved (a process must not mix it with another HIP runtime)."""

    def __init__(self, arr):
        load()
        self.hip = _hip()
        arr = np.ascontiguousarray(arr)
        self.shape, self.dtype, self.nbytes = arr.shape, arr.dtype, arr.nbytes
        self.ptr = ctypes.c_void_p()
        if self.hip.hipMalloc(ctypes.byref(self.ptr), max(1, arr.nbytes)) != 0:
            raise MemoryError("hipMalloc failed")
        self.upload(arr)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.shape == self.shape
        if self.hip.hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy to device failed")

    def volume(self, **kw):
        nz, ny, nx = self.shape
        return make_volume(self.ptr.value, dims=(nx, ny, nz), memory=MEMORY_DEVICE, dtype=self.dtype, **kw)

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_surface_to_numpy(s):
    """Copy a device-memory surface (or view) into numpy arrays, like surface_to_numpy."""
    assert s.memory == MEMORY_DEVICE
    hip = _hip()

    def fetch(ptr, n, dtype):
        out = np.zeros(n, dtype)
        if n:
            rc = hip.hipMemcpy(out.ctypes.data, ctypes.cast(ptr, ctypes.c_void_p), out.nbytes, 2)
            if rc != 0:
                raise RuntimeError(f"hipMemcpy failed: {rc}")
        return out

    return {
        "coords": fetch(s.coords, int(s.n_vertices) * 3, np.float32),
        "connections": fetch(s.connections, int(s.n_connections), np.uint32),
        "normals": fetch(s.normals, int(s.n_normals) * 3, np.float32),
        "colors": fetch(s.colors, int(s.n_colors) * 3 if s.colors else 0, np.uint8),
        "n_triangles": int(s.n_triangles), "polygon_type": s.polygon_type, "normal_type": s.normal_type,
        "color_type": s.color_type, "empty_range": s.empty_range,
        "min_value": s.min_value, "max_value": s.max_value,
        "min_coord": list(s.min_coord), "max_coord": list(s.max_coord),
    }


class Server:
    """A kmc_server: resident volumes and a continuously batched request loop."""

    def __init__(self, budget_bytes=0, workers=0, max_inflight=0, max_batch=0, out_memory=MEMORY_HOST, device=0):
        self.lib = load()
        cfg = ServerConfig(device, workers, max_inflight, max_batch, budget_bytes, out_memory)
        h = ctypes.c_void_p()
        st = self.lib.kmc_server_create(ctypes.byref(cfg), ctypes.byref(h))
        if st != OK:
            raise KmcError(st)
        self.handle = h
        self.out_memory = out_memory
        self._vols = []

    def register(self, vol):
        """vol: a kmc Volume (host or device values).  Returns the volume id."""
        vid = ctypes.c_uint32()
        st = self.lib.kmc_server_register_volume(self.handle, ctypes.byref(vol), ctypes.byref(vid))
        if st != OK:
            raise KmcError(st)
        self._vols.append(vol)
        return vid.value

    def submit(self, vid, isolevel, normal_type=POLYGON_NORMAL, duplication=True):
        rid = ctypes.c_uint64()
        st = self.lib.kmc_server_submit(self.handle, vid, float(isolevel), int(normal_type), int(bool(duplication)),
                                        ctypes.byref(rid))
        if st != OK:
            raise KmcError(st)
        return rid.value

    def wait(self, rid, timeout_ms=-1):
        """Wait for a request; returns its surface as numpy arrays (raises on failure)."""
        s = Surface()
        st = self.lib.kmc_server_wait(self.handle, rid, int(timeout_ms), ctypes.byref(s))
        if st != OK:
            raise KmcError(st)
        try:
            return surface_to_numpy(s) if s.memory == MEMORY_HOST else device_surface_to_numpy(s)
        finally:
            self.lib.kmc_surface_release(ctypes.byref(s))

    def wait_count(self, rid, timeout_ms=-1):
        """Wait for a request and release its surface; returns its triangle count."""
        s = Surface()
        st = self.lib.kmc_server_wait(self.handle, rid, int(timeout_ms), ctypes.byref(s))
        if st != OK:
            raise KmcError(st)
        n = int(s.n_triangles)
        self.lib.kmc_surface_release(ctypes.byref(s))
        return n

    def poll(self, rid):
        s = Surface()
        st = self.lib.kmc_server_poll(self.handle, rid, ctypes.byref(s))
        if st == ERR_NOT_READY:
            return None
        if st != OK:
            raise KmcError(st)
        try:
            return surface_to_numpy(s) if s.memory == MEMORY_HOST else device_surface_to_numpy(s)
        finally:
            self.lib.kmc_surface_release(ctypes.byref(s))

    def stats(self):
        st = ServerStats()
        rc = self.lib.kmc_server_get_stats(self.handle, ctypes.byref(st))
        if rc != OK:
            raise KmcError(rc)
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def reset_stats(self):
        self.lib.kmc_server_reset_stats(self.handle)

    def close(self):
        if self.handle:
            self.lib.kmc_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
