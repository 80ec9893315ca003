"""THE ORACLE -- test infrastructure only.

Python bindings of
  * liboracle.so   -- mc_oracle.c, the plain-C restatement of kvs::MarchingCubes;
  * _ref/libkvsref.so -- the reference's own marching-cubes path compiled from
    its sources by oracle/Makefile (present when it was built).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package; the product never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libkvsref.so")

NUMPY_TYPES = {
    np.dtype(np.int8): 0, np.dtype(np.int16): 1, np.dtype(np.int32): 2, np.dtype(np.int64): 3,
    np.dtype(np.uint8): 4, np.dtype(np.uint16): 5, np.dtype(np.uint32): 6, np.dtype(np.uint64): 7,
    np.dtype(np.float32): 8, np.dtype(np.float64): 9,
}
VERTEX_NORMAL, POLYGON_NORMAL = 0, 1


class _OracleMesh(ctypes.Structure):
    _fields_ = [
        ("coords", ctypes.POINTER(ctypes.c_float)), ("n_coords", ctypes.c_uint64),
        ("connections", ctypes.POINTER(ctypes.c_uint32)), ("n_connections", ctypes.c_uint64),
        ("normals", ctypes.POINTER(ctypes.c_float)), ("n_normals", ctypes.c_uint64),
        ("coords_f64", ctypes.POINTER(ctypes.c_double)),
        ("n_triangles", ctypes.c_uint64),
        ("polygon_type", ctypes.c_int32), ("normal_type", ctypes.c_int32),
        ("color_type", ctypes.c_int32), ("empty_range", ctypes.c_int32),
    ]


class _SliceMesh(ctypes.Structure):
    _fields_ = [
        ("coords", ctypes.POINTER(ctypes.c_float)), ("n_coords", ctypes.c_uint64),
        ("normals", ctypes.POINTER(ctypes.c_float)), ("n_normals", ctypes.c_uint64),
        ("colors", ctypes.POINTER(ctypes.c_uint8)), ("n_colors", ctypes.c_uint64),
        ("n_triangles", ctypes.c_uint64),
    ]


class _RefUVolume(ctypes.Structure):
    _fields_ = [
        ("values", ctypes.c_void_p), ("dtype", ctypes.c_int32),
        ("nnodes", ctypes.c_uint64), ("ncells", ctypes.c_uint64), ("cell_type", ctypes.c_int32),
        ("coords", ctypes.c_void_p), ("connections", ctypes.c_void_p),
        ("has_min_max_values", ctypes.c_int32), ("min_value", ctypes.c_double), ("max_value", ctypes.c_double),
    ]


class _RefReadVolume(ctypes.Structure):
    _fields_ = [
        ("values", ctypes.c_void_p), ("nbytes", ctypes.c_uint64), ("dtype", ctypes.c_int32),
        ("dims", ctypes.c_uint32 * 3), ("veclen", ctypes.c_uint32), ("grid_type", ctypes.c_int32),
        ("min_value", ctypes.c_double), ("max_value", ctypes.c_double),
        ("min_obj", ctypes.c_float * 3), ("max_obj", ctypes.c_float * 3),
    ]


class _RefVolume(ctypes.Structure):
    _fields_ = [
        ("values", ctypes.c_void_p), ("dtype", ctypes.c_int32), ("dims", ctypes.c_uint32 * 3),
        ("has_min_max_values", ctypes.c_int32), ("min_value", ctypes.c_double), ("max_value", ctypes.c_double),
        ("has_min_max_coords", ctypes.c_int32), ("min_coord", ctypes.c_float * 3), ("max_coord", ctypes.c_float * 3),
    ]


class _RefMesh(ctypes.Structure):
    _fields_ = [
        ("coords", ctypes.POINTER(ctypes.c_float)), ("n_coords", ctypes.c_uint64),
        ("connections", ctypes.POINTER(ctypes.c_uint32)), ("n_connections", ctypes.c_uint64),
        ("normals", ctypes.POINTER(ctypes.c_float)), ("n_normals", ctypes.c_uint64),
        ("colors", ctypes.POINTER(ctypes.c_uint8)), ("n_colors", ctypes.c_uint64),
        ("opacities", ctypes.POINTER(ctypes.c_uint8)), ("n_opacities", ctypes.c_uint64),
        ("polygon_type", ctypes.c_int32), ("normal_type", ctypes.c_int32), ("color_type", ctypes.c_int32),
        ("success", ctypes.c_int32),
        ("volume_min_value", ctypes.c_double), ("volume_max_value", ctypes.c_double),
        ("min_obj", ctypes.c_float * 3), ("max_obj", ctypes.c_float * 3),
        ("seconds", ctypes.c_double),
    ]


_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle oracle`")
        lib = ctypes.CDLL(ORACLE_SO)
        lib.oracle_extract.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32),
                                       ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_float),
                                       ctypes.POINTER(ctypes.c_float), ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.POINTER(_OracleMesh)]
        lib.oracle_extract.restype = ctypes.c_int
        lib.oracle_free.argtypes = [ctypes.POINTER(_OracleMesh)]
        lib.oracle_min_max.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        lib.oracle_min_max.restype = ctypes.c_int
        lib.oracle_case_edges.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        lib.oracle_case_edges.restype = ctypes.c_int
        lib.oracle_slice_plane.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32),
                                           ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint8),
                                           ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                           ctypes.POINTER(_SliceMesh)]
        lib.oracle_slice_plane.restype = ctypes.c_int
        lib.oracle_slice_free.argtypes = [ctypes.POINTER(_SliceMesh)]
        lib.oracle_unstructured_min_max.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64,
                                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        lib.oracle_unstructured_min_max.restype = ctypes.c_int
        lib.oracle_unstructured_min_max_coords.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                                           ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        lib.oracle_unstructured_min_max_coords.restype = None
        lib.oracle_unstructured_iso.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_int, ctypes.POINTER(_SliceMesh)]
        lib.oracle_unstructured_iso.restype = ctypes.c_int
        lib.oracle_unstructured_slice.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                                  ctypes.POINTER(ctypes.c_uint8), ctypes.c_uint32, ctypes.c_float,
                                                  ctypes.c_float, ctypes.POINTER(_SliceMesh)]
        lib.oracle_unstructured_slice.restype = ctypes.c_int
        lib.oracle_unstructured_free.argtypes = [ctypes.POINTER(_SliceMesh)]
        _oracle = lib
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"{REF_SO} missing: run `make -C oracle ref` where the reference sources exist")
        lib = ctypes.CDLL(REF_SO)
        lib.ref_marching_cubes.argtypes = [ctypes.POINTER(_RefVolume), ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, ctypes.POINTER(_RefMesh)]
        lib.ref_isosurface.argtypes = [ctypes.POINTER(_RefVolume), ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_RefMesh)]
        lib.ref_mesh_free.argtypes = [ctypes.POINTER(_RefMesh)]
        lib.ref_min_max.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32),
                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        lib.ref_tf_color.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.POINTER(ctypes.c_uint8)]
        lib.ref_tf_table.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint8)]
        lib.ref_tf_table.restype = ctypes.c_int
        lib.ref_mc_tables.argtypes = [ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
        lib.ref_hydrogen.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint8)]
        lib.ref_hydrogen.restype = ctypes.c_int
        lib.ref_slice_plane.argtypes = [ctypes.POINTER(_RefVolume), ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                        ctypes.c_int, ctypes.POINTER(_RefMesh)]
        lib.ref_slice_plane_coefficients.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                                     ctypes.POINTER(ctypes.c_float)]
        lib.ref_slice_plane_coefficients.restype = None
        lib.ref_isosurface_unstructured.argtypes = [ctypes.POINTER(_RefUVolume), ctypes.c_double, ctypes.c_int,
                                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_RefMesh)]
        lib.ref_slice_plane_unstructured.argtypes = [ctypes.POINTER(_RefUVolume), ctypes.POINTER(ctypes.c_float),
                                                     ctypes.c_int, ctypes.c_int, ctypes.POINTER(_RefMesh)]
        lib.ref_unstructured_tables.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                                                ctypes.POINTER(ctypes.c_int32)]
        lib.ref_unstructured_tables.restype = ctypes.c_int
        lib.ref_write_kvsml_volume.argtypes = [ctypes.POINTER(_RefVolume), ctypes.c_void_p, ctypes.c_char_p,
                                               ctypes.c_int, ctypes.c_int]
        lib.ref_write_kvsml_volume.restype = ctypes.c_int
        lib.ref_read_kvsml_volume.argtypes = [ctypes.c_char_p, ctypes.POINTER(_RefReadVolume)]
        lib.ref_read_kvsml_volume.restype = ctypes.c_int
        lib.ref_read_volume_free.argtypes = [ctypes.POINTER(_RefReadVolume)]
        lib.ref_write_avs_volume.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint32),
                                             ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_float),
                                             ctypes.c_char_p]
        lib.ref_write_avs_volume.restype = ctypes.c_int
        lib.ref_read_avs_volume.argtypes = [ctypes.c_char_p, ctypes.POINTER(_RefReadVolume),
                                            ctypes.POINTER(ctypes.c_float)]
        lib.ref_read_avs_volume.restype = ctypes.c_int
        lib.ref_export_stl.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64,
                                       ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_char_p]
        lib.ref_export_stl.restype = ctypes.c_int
        lib.ref_export_kvsml_polygon.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64,
                                                 ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_char_p]
        lib.ref_export_kvsml_polygon.restype = ctypes.c_int
        lib.ref_export_ply.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p,
                                       ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_char_p]
        lib.ref_export_ply.restype = ctypes.c_int
        lib.ref_read_ply.argtypes = [ctypes.c_char_p, ctypes.POINTER(_RefMesh)]
        lib.ref_read_ply.restype = ctypes.c_int
        _ref = lib
    return _ref


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(int(n),)).copy()


def _default_coords(dims):
    return [0.0, 0.0, 0.0], [float(np.float32(d) - np.float32(1)) for d in dims]


def min_max(values):
    """GetMinMaxValues<T> as restated by the oracle."""
    a = np.ascontiguousarray(values)
    mn, mx = ctypes.c_double(), ctypes.c_double()
    rc = oracle_lib().oracle_min_max(a.ctypes.data, NUMPY_TYPES[a.dtype], a.size, ctypes.byref(mn), ctypes.byref(mx))
    assert rc == 0
    return mn.value, mx.value


def extract(values, isolevel, normal_type=POLYGON_NORMAL, duplication=True, min_max_values=None, coords=None,
            want_f64=False):
    """kvs::MarchingCubes as restated by the oracle.  values: numpy [z][y][x]."""
    a = np.ascontiguousarray(values)
    nz, ny, nx = a.shape
    dims = (ctypes.c_uint32 * 3)(nx, ny, nz)
    if min_max_values is None:
        min_max_values = min_max(a)
    if coords is None:
        coords = _default_coords((nx, ny, nz))
    minc = (ctypes.c_float * 3)(*coords[0])
    maxc = (ctypes.c_float * 3)(*coords[1])
    m = _OracleMesh()
    rc = oracle_lib().oracle_extract(a.ctypes.data, NUMPY_TYPES[a.dtype], dims, float(min_max_values[0]),
                                     float(min_max_values[1]), minc, maxc, float(isolevel), int(normal_type),
                                     int(bool(duplication)), int(bool(want_f64)), ctypes.byref(m))
    if rc != 0:
        raise RuntimeError(f"oracle_extract failed: {rc}")
    try:
        out = {
            "coords": _arr(m.coords, m.n_coords, np.float32),
            "connections": _arr(m.connections, m.n_connections, np.uint32),
            "normals": _arr(m.normals, m.n_normals, np.float32),
            "n_triangles": int(m.n_triangles),
            "polygon_type": m.polygon_type, "normal_type": m.normal_type, "color_type": m.color_type,
            "empty_range": m.empty_range,
        }
        if want_f64:
            out["coords_f64"] = _arr(m.coords_f64, m.n_coords if m.coords_f64 else 0, np.float64)
        return out
    finally:
        oracle_lib().oracle_free(ctypes.byref(m))


def case_edges(ci):
    e = (ctypes.c_int * 16)()
    n = oracle_lib().oracle_case_edges(ci, e)
    return n, list(e)


def _ref_volume(a, min_max_values, coords):
    v = _RefVolume()
    nz, ny, nx = a.shape
    v.values = a.ctypes.data
    v.dtype = NUMPY_TYPES[a.dtype]
    v.dims[:] = [nx, ny, nz]
    if min_max_values is not None:
        v.has_min_max_values = 1
        v.min_value, v.max_value = float(min_max_values[0]), float(min_max_values[1])
    if coords is not None:
        v.has_min_max_coords = 1
        v.min_coord[:] = list(coords[0])
        v.max_coord[:] = list(coords[1])
    return v


def _ref_mesh_dict(m):
    return {
        "coords": _arr(m.coords, m.n_coords, np.float32),
        "connections": _arr(m.connections, m.n_connections, np.uint32),
        "normals": _arr(m.normals, m.n_normals, np.float32),
        "colors": _arr(m.colors, m.n_colors, np.uint8),
        "opacities": _arr(m.opacities, m.n_opacities, np.uint8),
        "polygon_type": m.polygon_type, "normal_type": m.normal_type, "color_type": m.color_type,
        "success": m.success, "volume_min_value": m.volume_min_value, "volume_max_value": m.volume_max_value,
        "min_obj": list(m.min_obj), "max_obj": list(m.max_obj), "seconds": m.seconds,
    }


def ref_marching_cubes(values, isolevel, normal_type=POLYGON_NORMAL, duplication=True, min_max_values=None,
                       coords=None, tf_kind=0, tf_resolution=256):
    """The reference's own kvs::MarchingCubes (oracle/_ref)."""
    a = np.ascontiguousarray(values)
    v = _ref_volume(a, min_max_values, coords)
    m = _RefMesh()
    ref_lib().ref_marching_cubes(ctypes.byref(v), float(isolevel), int(normal_type), int(bool(duplication)),
                                 int(tf_kind), int(tf_resolution), ctypes.byref(m))
    try:
        return _ref_mesh_dict(m)
    finally:
        ref_lib().ref_mesh_free(ctypes.byref(m))


def ref_isosurface(values, isolevel=None, normal_type=POLYGON_NORMAL, duplication=True, tf_kind=0,
                   tf_resolution=256, min_max_values=None, coords=None):
    """The reference's own kvs::Isosurface (oracle/_ref); isolevel None uses the
    (min+max)/2 constructor."""
    a = np.ascontiguousarray(values)
    v = _ref_volume(a, min_max_values, coords)
    m = _RefMesh()
    ref_lib().ref_isosurface(ctypes.byref(v), int(isolevel is None), float(isolevel or 0.0), int(normal_type),
                             int(bool(duplication)), int(tf_kind), int(tf_resolution), ctypes.byref(m))
    try:
        return _ref_mesh_dict(m)
    finally:
        ref_lib().ref_mesh_free(ctypes.byref(m))


def ref_min_max(values):
    a = np.ascontiguousarray(values)
    nz, ny, nx = a.shape
    dims = (ctypes.c_uint32 * 3)(nx, ny, nz)
    mn, mx = ctypes.c_double(), ctypes.c_double()
    ref_lib().ref_min_max(a.ctypes.data, NUMPY_TYPES[a.dtype], dims, ctypes.byref(mn), ctypes.byref(mx))
    return mn.value, mx.value


def ref_tf_color(tf_kind, resolution, range_min, range_max, value):
    rgb = (ctypes.c_uint8 * 3)()
    ref_lib().ref_tf_color(int(tf_kind), int(resolution), float(range_min), float(range_max), float(value), rgb)
    return tuple(rgb)


def ref_tf_table(tf_kind, resolution=256):
    out = (ctypes.c_uint8 * (3 * resolution))()
    n = ref_lib().ref_tf_table(int(tf_kind), int(resolution), out)
    return np.frombuffer(bytes(out), dtype=np.uint8)[:n].copy()


def ref_tables():
    tri = (ctypes.c_int32 * (256 * 16))()
    vert = (ctypes.c_int32 * (12 * 2 * 3))()
    ref_lib().ref_mc_tables(tri, vert)
    return np.array(tri, dtype=np.int32).reshape(256, 16), np.array(vert, dtype=np.int32).reshape(12, 2, 3)


def ref_hydrogen(dims):
    """kvs::HydrogenVolumeData(dims) values, [z][y][x] uint8 (oracle/_ref)."""
    nx, ny, nz = dims
    out = (ctypes.c_uint8 * (nx * ny * nz))()
    d = (ctypes.c_uint32 * 3)(nx, ny, nz)
    ref_lib().ref_hydrogen(d, out)
    return np.frombuffer(bytes(out), dtype=np.uint8).reshape(nz, ny, nx).copy()


def tf_range(values, min_max_values=None):
    """The colour range MapperBase::setRange gives a transfer function without
    a range of its own (MapperBase.cpp:54-83): [-128, 127] for int8, [0, 255]
    for uint8, else [float(min), float(max)], widened by 1 when Math::Equal."""
    a = np.asarray(values)
    if a.dtype == np.int8:
        return -128.0, 127.0
    if a.dtype == np.uint8:
        return 0.0, 255.0
    mn, mx = min_max_values if min_max_values is not None else min_max(a)
    lo, hi = np.float32(mn), np.float32(mx)
    if abs(float(np.float32(lo - hi))) < 1e-6:
        hi = np.float32(hi + np.float32(1.0))
    return float(lo), float(hi)


def slice_plane(values, coef, table, color_range=None, coords=None, min_max_values=None):
    """kvs::SlicePlane over a structured volume as restated by the oracle.
    table: the colour map's created table (3 bytes per entry); color_range:
    its [min, max] (default: tf_range of the volume)."""
    a = np.ascontiguousarray(values)
    nz, ny, nx = a.shape
    dims = (ctypes.c_uint32 * 3)(nx, ny, nz)
    if coords is None:
        coords = _default_coords((nx, ny, nz))
    if color_range is None:
        color_range = tf_range(a, min_max_values)
    minc = (ctypes.c_float * 3)(*coords[0])
    maxc = (ctypes.c_float * 3)(*coords[1])
    c = (ctypes.c_float * 4)(*[float(x) for x in coef])
    t = np.ascontiguousarray(table, dtype=np.uint8)
    m = _SliceMesh()
    rc = oracle_lib().oracle_slice_plane(a.ctypes.data, NUMPY_TYPES[a.dtype], dims, minc, maxc, c,
                                         t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), t.size // 3,
                                         float(color_range[0]), float(color_range[1]), ctypes.byref(m))
    if rc != 0:
        raise RuntimeError(f"oracle_slice_plane failed: {rc}")
    try:
        return {"coords": _arr(m.coords, m.n_coords, np.float32), "normals": _arr(m.normals, m.n_normals, np.float32),
                "colors": _arr(m.colors, m.n_colors, np.uint8), "n_triangles": int(m.n_triangles)}
    finally:
        oracle_lib().oracle_slice_free(ctypes.byref(m))


def ref_slice_plane(values, coef, tf_kind=0, tf_resolution=256, min_max_values=None, coords=None):
    """The reference's own kvs::SlicePlane over a structured volume (oracle/_ref)."""
    a = np.ascontiguousarray(values)
    v = _ref_volume(a, min_max_values, coords)
    c = (ctypes.c_float * 4)(*[float(x) for x in coef])
    m = _RefMesh()
    ref_lib().ref_slice_plane(ctypes.byref(v), c, int(tf_kind), int(tf_resolution), ctypes.byref(m))
    try:
        return _ref_mesh_dict(m)
    finally:
        ref_lib().ref_mesh_free(ctypes.byref(m))


def ref_slice_plane_coefficients(point, normal):
    """The coefficients kvs::SlicePlane::setPlane( point, normal ) stores."""
    p = (ctypes.c_float * 3)(*[float(x) for x in point])
    n = (ctypes.c_float * 3)(*[float(x) for x in normal])
    c = (ctypes.c_float * 4)()
    ref_lib().ref_slice_plane_coefficients(p, n, c)
    return [float(x) for x in c]


# kvs::UnstructuredVolumeObject::CellType and the nodes of each cell type
TETRAHEDRA, HEXAHEDRA, PYRAMID, PRISM = 1, 2, 5, 7
CELL_NODES = {TETRAHEDRA: 4, HEXAHEDRA: 8, PYRAMID: 5, PRISM: 6}


def _u_inputs(values, coords, connections):
    v = np.ascontiguousarray(values).ravel()
    c = np.ascontiguousarray(coords, dtype=np.float32).ravel()
    k = np.ascontiguousarray(connections, dtype=np.uint32).ravel()
    return v, c, k


def unstructured_min_max(values, connections):
    """GetMinMaxValues<T> of an unstructured volume (connection order), as restated."""
    v = np.ascontiguousarray(values).ravel()
    k = np.ascontiguousarray(connections, dtype=np.uint32).ravel()
    mn, mx = ctypes.c_double(), ctypes.c_double()
    assert oracle_lib().oracle_unstructured_min_max(v.ctypes.data, NUMPY_TYPES[v.dtype], k.ctypes.data, k.size,
                                                    ctypes.byref(mn), ctypes.byref(mx)) == 0
    return mn.value, mx.value


def unstructured_min_max_coords(coords, connections):
    c = np.ascontiguousarray(coords, dtype=np.float32).ravel()
    k = np.ascontiguousarray(connections, dtype=np.uint32).ravel()
    lo, hi = (ctypes.c_float * 3)(), (ctypes.c_float * 3)()
    oracle_lib().oracle_unstructured_min_max_coords(c.ctypes.data, k.ctypes.data, k.size, lo, hi)
    return list(lo), list(hi)


def _umesh_dict(m, colors):
    out = {"coords": _arr(m.coords, m.n_coords, np.float32), "normals": _arr(m.normals, m.n_normals, np.float32),
           "n_triangles": int(m.n_triangles)}
    if colors:
        out["colors"] = _arr(m.colors, m.n_colors, np.uint8)
    return out


def unstructured_isosurface(values, coords, connections, cell_type, isolevel, normal_type=POLYGON_NORMAL,
                            duplication=True, min_max_values=None):
    """The unstructured branch of kvs::Isosurface (MarchingTetrahedra / Hexahedra /
    Pyramid / Prism) as restated by the oracle."""
    v, c, k = _u_inputs(values, coords, connections)
    if min_max_values is None:
        min_max_values = unstructured_min_max(v, k)
    dup = bool(duplication) and normal_type != VERTEX_NORMAL
    m = _SliceMesh()
    rc = oracle_lib().oracle_unstructured_iso(v.ctypes.data, NUMPY_TYPES[v.dtype], c.ctypes.data, k.ctypes.data,
                                              k.size // CELL_NODES[cell_type], int(cell_type),
                                              float(min_max_values[0]), float(min_max_values[1]), float(isolevel),
                                              int(dup), ctypes.byref(m))
    if rc != 0:
        raise RuntimeError(f"oracle_unstructured_iso failed: {rc}")
    try:
        return _umesh_dict(m, False)
    finally:
        oracle_lib().oracle_unstructured_free(ctypes.byref(m))


def unstructured_slice_plane(values, coords, connections, cell_type, coef, table, color_range=None,
                             min_max_values=None):
    """kvs::SlicePlane over an unstructured volume as restated by the oracle."""
    v, c, k = _u_inputs(values, coords, connections)
    if color_range is None:
        if v.dtype in (np.int8, np.uint8):
            color_range = tf_range(v)
        else:
            mm = min_max_values if min_max_values is not None else unstructured_min_max(v, k)
            color_range = tf_range(v, mm)
    cf = (ctypes.c_float * 4)(*[float(x) for x in coef])
    t = np.ascontiguousarray(table, dtype=np.uint8)
    m = _SliceMesh()
    rc = oracle_lib().oracle_unstructured_slice(v.ctypes.data, NUMPY_TYPES[v.dtype], c.ctypes.data, k.ctypes.data,
                                                k.size // CELL_NODES[cell_type], int(cell_type), cf,
                                                t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), t.size // 3,
                                                float(color_range[0]), float(color_range[1]), ctypes.byref(m))
    if rc != 0:
        raise RuntimeError(f"oracle_unstructured_slice failed: {rc}")
    try:
        return _umesh_dict(m, True)
    finally:
        oracle_lib().oracle_unstructured_free(ctypes.byref(m))


def _ref_uvolume(v, c, k, cell_type, min_max_values):
    u = _RefUVolume()
    u.values = v.ctypes.data
    u.dtype = NUMPY_TYPES[v.dtype]
    u.nnodes = v.size
    u.ncells = k.size // CELL_NODES[cell_type]
    u.cell_type = int(cell_type)
    u.coords = c.ctypes.data
    u.connections = k.ctypes.data
    if min_max_values is not None:
        u.has_min_max_values = 1
        u.min_value, u.max_value = float(min_max_values[0]), float(min_max_values[1])
    return u


def ref_unstructured_isosurface(values, coords, connections, cell_type, isolevel, normal_type=POLYGON_NORMAL,
                                duplication=True, tf_kind=0, tf_resolution=256, min_max_values=None):
    """The reference's own kvs::Isosurface over an unstructured volume (oracle/_ref)."""
    v, c, k = _u_inputs(values, coords, connections)
    u = _ref_uvolume(v, c, k, cell_type, min_max_values)
    m = _RefMesh()
    ref_lib().ref_isosurface_unstructured(ctypes.byref(u), float(isolevel), int(normal_type), int(bool(duplication)),
                                          int(tf_kind), int(tf_resolution), ctypes.byref(m))
    try:
        return _ref_mesh_dict(m)
    finally:
        ref_lib().ref_mesh_free(ctypes.byref(m))


def ref_unstructured_slice_plane(values, coords, connections, cell_type, coef, tf_kind=0, tf_resolution=256,
                                 min_max_values=None):
    v, c, k = _u_inputs(values, coords, connections)
    u = _ref_uvolume(v, c, k, cell_type, min_max_values)
    cf = (ctypes.c_float * 4)(*[float(x) for x in coef])
    m = _RefMesh()
    ref_lib().ref_slice_plane_unstructured(ctypes.byref(u), cf, int(tf_kind), int(tf_resolution), ctypes.byref(m))
    try:
        return _ref_mesh_dict(m)
    finally:
        ref_lib().ref_mesh_free(ctypes.byref(m))


def ref_unstructured_tables(cell):
    """(TriangleID, VertexID) of cell 0 tetrahedra, 1 hexahedra, 2 pyramid, 3 prism."""
    shapes = {0: ((16, 7), (6, 2)), 1: ((256, 16), (12, 2)), 2: ((36, 13), (8, 2)), 3: ((64, 10), (9, 2))}
    ts, vs = shapes[cell]
    tri = (ctypes.c_int32 * (ts[0] * ts[1]))()
    vert = (ctypes.c_int32 * (vs[0] * vs[1]))()
    ref_lib().ref_unstructured_tables(int(cell), tri, vert)
    return np.array(tri, dtype=np.int32).reshape(ts), np.array(vert, dtype=np.int32).reshape(vs)


def ref_write_kvsml_volume(values, filename, ascii=False, external=True, rect_coords=None, min_max_values=None):
    """StructuredVolumeObject::write (KVSML) by the reference (oracle/_ref)."""
    a = np.ascontiguousarray(values)
    v = _ref_volume(a, min_max_values, None)
    rc_arr = np.ascontiguousarray(rect_coords, dtype=np.float32) if rect_coords is not None else None
    rc = ref_lib().ref_write_kvsml_volume(ctypes.byref(v), rc_arr.ctypes.data if rc_arr is not None else None,
                                          str(filename).encode(), int(bool(ascii)), int(bool(external)))
    if rc != 0:
        raise RuntimeError("reference KVSML write failed")


def ref_read_kvsml_volume(filename):
    """StructuredVolumeObject::read (KVSML) by the reference: values [z][y][x],
    min/max values and min/max object coordinates."""
    r = _RefReadVolume()
    if ref_lib().ref_read_kvsml_volume(str(filename).encode(), ctypes.byref(r)) != 0:
        raise RuntimeError("reference KVSML read failed")
    try:
        dtype = [k for k, t in NUMPY_TYPES.items() if t == r.dtype][0]
        buf = ctypes.string_at(r.values, r.nbytes)
        vals = np.frombuffer(buf, dtype=dtype).reshape(r.dims[2], r.dims[1], r.dims[0]).copy()
        return {"values": vals, "min_value": r.min_value, "max_value": r.max_value, "grid_type": r.grid_type,
                "min_obj": list(r.min_obj), "max_obj": list(r.max_obj)}
    finally:
        ref_lib().ref_read_volume_free(ctypes.byref(r))


def ref_write_avs_volume(values, filename, ext=None):
    """AVSField::write of a uniform field by the reference (oracle/_ref);
    values [z][y][x] of uint8 / int16 / int32 / float32 / float64; ext: the
    optional (min_ext, max_ext)."""
    a = np.ascontiguousarray(values)
    nz, ny, nx = a.shape
    dims = (ctypes.c_uint32 * 3)(nx, ny, nz)
    e = (ctypes.c_float * 6)(*(list(ext[0]) + list(ext[1]))) if ext is not None else (ctypes.c_float * 6)()
    if ref_lib().ref_write_avs_volume(a.ctypes.data, NUMPY_TYPES[a.dtype], dims, 1, int(ext is not None), e,
                                      str(filename).encode()) != 0:
        raise RuntimeError("reference AVS write failed")


def ref_read_avs_volume(filename):
    """StructuredVolumeImporter( filename ) of an AVS field by the reference:
    values [z][y][x], min/max values, min/max object and external coordinates."""
    r = _RefReadVolume()
    ext = (ctypes.c_float * 6)()
    if ref_lib().ref_read_avs_volume(str(filename).encode(), ctypes.byref(r), ext) != 0:
        raise RuntimeError("reference AVS read failed")
    try:
        dtype = [k for k, t in NUMPY_TYPES.items() if t == r.dtype][0]
        buf = ctypes.string_at(r.values, r.nbytes)
        vals = np.frombuffer(buf, dtype=dtype).reshape(r.dims[2], r.dims[1], r.dims[0]).copy()
        return {"values": vals, "min_value": r.min_value, "max_value": r.max_value, "grid_type": r.grid_type,
                "min_obj": list(r.min_obj), "max_obj": list(r.max_obj), "min_ext": list(ext[:3]),
                "max_ext": list(ext[3:])}
    finally:
        ref_lib().ref_read_volume_free(ctypes.byref(r))


def ref_export_kvsml_polygon(mesh, filename, writing_type=0):
    """PolygonExporter<kvs::KVSMLPolygonObject> + setWritingDataType(writing_type)
    + write by the reference (oracle/_ref): 0 Ascii, 1 ExternalAscii, 2 ExternalBinary.
    mesh: a dict as ref_marching_cubes / ref_slice_plane return (its colours,
    opacities and types are the polygon's)."""
    c = np.ascontiguousarray(mesh["coords"], dtype=np.float32)
    k = np.ascontiguousarray(mesh["connections"], dtype=np.uint32)
    n = np.ascontiguousarray(mesh["normals"], dtype=np.float32)
    cl = np.ascontiguousarray(mesh["colors"], dtype=np.uint8)
    op = np.asarray(mesh["opacities"], dtype=np.uint8)
    rc = ref_lib().ref_export_kvsml_polygon(
        c.ctypes.data if c.size else None, c.size, k.ctypes.data if k.size else None, k.size,
        n.ctypes.data if n.size else None, n.size, cl.ctypes.data if cl.size else None, cl.size,
        int(op[0]) if op.size else -1, int(mesh["color_type"]), int(mesh["normal_type"]), int(writing_type),
        str(filename).encode())
    if rc != 0:
        raise RuntimeError("reference KVSML polygon export failed")


def ref_export_ply(mesh, filename, binary=False):
    """PolygonExporter<kvs::Ply> + setFileType + write by the reference
    (oracle/_ref), for a mesh dict without connections (the reference's writer
    fails on connections: it stores face indices through a null pointer)."""
    c = np.ascontiguousarray(mesh["coords"], dtype=np.float32)
    n = np.ascontiguousarray(mesh["normals"], dtype=np.float32)
    cl = np.ascontiguousarray(mesh["colors"], dtype=np.uint8)
    k = np.asarray(mesh.get("connections", np.zeros(0)))
    rc = ref_lib().ref_export_ply(c.ctypes.data if c.size else None, c.size, k.size, n.ctypes.data if n.size else None,
                                  n.size, cl.ctypes.data if cl.size else None, cl.size, int(mesh["color_type"]),
                                  int(mesh["normal_type"]), int(bool(binary)), str(filename).encode())
    if rc != 0:
        raise RuntimeError(f"reference PLY export failed ({rc})")


def ref_read_ply(filename):
    """Ply::read by the reference: coords, colors, normals, connections and
    whether the file had colours."""
    m = _RefMesh()
    if ref_lib().ref_read_ply(str(filename).encode(), ctypes.byref(m)) != 0:
        raise RuntimeError("reference PLY read failed")
    try:
        return {"coords": _arr(m.coords, m.n_coords, np.float32), "colors": _arr(m.colors, m.n_colors, np.uint8),
                "normals": _arr(m.normals, m.n_normals, np.float32),
                "connections": _arr(m.connections, m.n_connections, np.uint32), "has_colors": bool(m.success)}
    finally:
        ref_lib().ref_mesh_free(ctypes.byref(m))


def ref_export_stl(mesh, normal_type, filename, ascii=False):
    """PolygonExporter<kvs::Stl> + Stl::write by the reference (oracle/_ref)."""
    c = np.ascontiguousarray(mesh["coords"], dtype=np.float32)
    k = np.ascontiguousarray(mesh["connections"], dtype=np.uint32)
    n = np.ascontiguousarray(mesh["normals"], dtype=np.float32)
    rc = ref_lib().ref_export_stl(c.ctypes.data, c.size, k.ctypes.data if k.size else None, k.size, n.ctypes.data,
                                  n.size, int(normal_type), int(bool(ascii)), str(filename).encode())
    if rc != 0:
        raise RuntimeError("reference STL export failed")
