"""Write the golden fixtures of tests/golden/ from the reference itself.

Run in the build container, where oracle/_ref/libkvsref.so -- the
reference's own kvs::MarchingCubes / kvs::Isosurface / ColorMap code compiled
from its sources by oracle/Makefile -- exists:

    make -C oracle && python tests/golden/make_golden.py

Every expected output below comes from that library; inputs are either the
reference's own HydrogenVolumeData or seeded synthetic volumes.  The files are
data only (inputs and expected outputs) and are what the GPU tests check
against, since the reference is not available where they run.
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from kvs_amd import datasets  # noqa: E402

MODES = [(1, True), (1, False), (0, False)]   # (normal_type, duplication)
DTYPES = [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64, np.float32, np.float64]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def digest(r):
    return {
        "n_coords": int(r["coords"].size), "n_connections": int(r["connections"].size),
        "n_normals": int(r["normals"].size),
        "coords": sha(r["coords"]), "connections": sha(r["connections"]), "normals": sha(r["normals"]),
        "colors": [int(c) for c in r["colors"]], "opacities": [int(c) for c in r["opacities"]],
        "polygon_type": r["polygon_type"], "normal_type": r["normal_type"], "color_type": r["color_type"],
        "success": r["success"],
    }


class Packer:
    """Concatenate many small arrays per field; offsets[i]..offsets[i+1] is case i."""

    def __init__(self):
        self.data = {}

    def add(self, field, arr):
        self.data.setdefault(field, []).append(np.ravel(arr))

    def save(self, path):
        out = {}
        for field, parts in self.data.items():
            out[field] = np.concatenate(parts) if parts else np.zeros(0)
            out[field + "_off"] = np.cumsum([0] + [p.size for p in parts]).astype(np.int64)
        np.savez_compressed(path, **out)


def random_case(rng, dtype, shape):
    if np.dtype(dtype).kind == "f":
        v = rng.standard_normal(shape).astype(dtype)
        iso = float(np.float32(rng.standard_normal() * 0.5))
    else:
        info = np.iinfo(dtype)
        lo, hi = max(int(info.min), -60), min(int(info.max), 60)
        v = rng.integers(lo, hi + 1, shape).astype(dtype)
        iso = float(rng.integers(lo, hi)) + (0.5 if rng.random() < 0.5 else 0.0)
    return v, iso


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libkvsref.so is missing: run `make -C oracle` first")
    meta = {"generator": "tests/golden/make_golden.py", "reference": "oracle/_ref/libkvsref.so"}

    # 1. the reference's case and edge tables, pinned by hash
    tri, vert = oracle.ref_tables()
    meta["tables"] = {"TriangleID_sha256": sha(tri.astype("<i4")), "VertexID_sha256": sha(vert.astype("<i4"))}
    np.savez_compressed(os.path.join(HERE, "tables.npz"), triangle_counts=(tri >= 0).sum(axis=1) // 3)

    # 2. brute force: every 2x2x2 sign pattern, four dtypes, three modes
    rng = np.random.default_rng(2024)
    pk = Packer()
    cases = []
    for dtype in (np.float32, np.float64, np.uint8, np.int16):
        for ci in range(256):
            if np.dtype(dtype).kind == "f":
                lo_v = rng.uniform(-1.0, -0.01, 8)
                hi_v = rng.uniform(0.01, 1.0, 8)
                iso = 0.0
            else:
                lo_v = rng.integers(-20 if np.dtype(dtype).kind == "i" else 0, 10, 8)
                hi_v = rng.integers(11, 40, 8)
                iso = 10.5
            # corner k of the single cell: (x, y, z) offsets of MarchingCubes.cpp:187-194
            corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
            v = np.zeros((2, 2, 2), dtype=dtype)
            for k, (x, y, z) in enumerate(corners):
                v[z, y, x] = hi_v[k] if (ci >> k) & 1 else lo_v[k]
            pk.add("in_" + np.dtype(dtype).name, v)
            for normal, dup in MODES:
                r = oracle.ref_marching_cubes(v, iso, normal, dup)
                mk = f"n{normal}_d{int(dup)}"
                pk.add(mk + "_coords", r["coords"])
                pk.add(mk + "_conn", r["connections"])
                pk.add(mk + "_normals", r["normals"])
            cases.append({"iso": iso, "dtype": np.dtype(dtype).name, "cube_index": ci})
    pk.save(os.path.join(HERE, "cubes.npz"))
    meta["cubes"] = cases

    # 3. random small volumes of every dtype, with and without user min/max and coords
    pk = Packer()
    cases = []
    for t in range(40):
        dtype = DTYPES[t % len(DTYPES)]
        shape = tuple(int(s) for s in rng.integers(2, 14, 3))
        v, iso = random_case(rng, dtype, shape)
        mm = None
        coords = None
        if t % 4 == 1:
            lo, hi = float(v.min()), float(v.max())
            mm = (lo + 0.2 * (hi - lo), hi - 0.25 * (hi - lo))
        if t % 4 == 2:
            coords = ([float(rng.uniform(-3, 3)) for _ in range(3)], [float(rng.uniform(4, 9)) for _ in range(3)])
        pk.add("in_bytes", np.frombuffer(v.tobytes(), dtype=np.uint8))
        for normal, dup in MODES:
            r = oracle.ref_marching_cubes(v, iso, normal, dup, min_max_values=mm, coords=coords)
            mk = f"n{normal}_d{int(dup)}"
            pk.add(mk + "_coords", r["coords"])
            pk.add(mk + "_conn", r["connections"])
            pk.add(mk + "_normals", r["normals"])
            pk.add(mk + "_colors", r["colors"])
        cases.append({"iso": iso, "dtype": np.dtype(dtype).name, "shape": list(shape), "min_max": mm,
                      "coords": coords})
    pk.save(os.path.join(HERE, "random.npz"))
    meta["random"] = cases

    # 4. HydrogenVolumeData 64^3 (the reference's own fixture), isolevels 37.5 and 100
    h = oracle.ref_hydrogen((64, 64, 64))
    np.savez_compressed(os.path.join(HERE, "hydrogen64.npz"), values=h)
    meta["hydrogen64"] = {"sha256": sha(h), "levels": {}}
    for iso in (37.5, 100.0):
        for normal, dup in MODES:
            r = oracle.ref_marching_cubes(h, iso, normal, dup)
            meta["hydrogen64"]["levels"][f"{iso}_n{normal}_d{int(dup)}"] = digest(r)
    r = oracle.ref_isosurface(h, None, 1, True)
    meta["hydrogen64"]["isosurface_default_level"] = digest(r)

    # 5. the 64^3 CPU configuration: 64^3 float32, 3 Gaussian blobs + noise,
    #    isolevel 0.5, PolygonNormal, duplication, default transfer function
    c0 = datasets.blobs(64, seed=1000)
    meta["c0"] = {"sha256": sha(c0), "levels": {}}
    for normal, dup in MODES:
        r = oracle.ref_marching_cubes(c0, 0.5, normal, dup)
        meta["c0"]["levels"][f"0.5_n{normal}_d{int(dup)}"] = digest(r)
        if normal == 1 and dup:
            np.savez_compressed(os.path.join(HERE, "c0_dup.npz"), coords=r["coords"], normals=r["normals"])

    # 6. colours: created tables and ColorMap::at after setRange
    meta["colormap_tables"] = {}
    for kind, name in ((0, "BrewerRdBu"), (1, "BrewerSingleHueGreys")):
        for res in (256, 17):
            meta["colormap_tables"][f"{name}_{res}"] = [int(x) for x in oracle.ref_tf_table(kind, res)]
    colors = []
    for kind in (0, 1):
        for (lo, hi) in ((0.0, 255.0), (-1.0, 1.0), (0.0, 1.0), (-128.0, 127.0), (3.5, 3.5000002)):
            for value in np.linspace(lo - 0.2 * abs(hi - lo + 1), hi + 0.2 * abs(hi - lo + 1), 23):
                rgb = oracle.ref_tf_color(kind, 256, lo, hi, float(np.float32(value)))
                colors.append({"kind": kind, "lo": lo, "hi": hi, "value": float(np.float32(value)), "rgb": list(rgb)})
    meta["colors"] = colors
    # the mapper's own colour choice: MapperBase::setRange from the dtype / volume range
    surf = []
    for dtype in (np.uint8, np.int8, np.float32, np.int16):
        v, iso = random_case(rng, dtype, (6, 7, 8))
        r = oracle.ref_marching_cubes(v, iso, 1, True)
        surf.append({"dtype": np.dtype(dtype).name, "min": float(v.min()), "max": float(v.max()), "iso": iso,
                     "colors": [int(c) for c in r["colors"]]})
    meta["surface_colors"] = surf

    # 7. GetMinMaxValues<T> corner cases: NaN first / inside, +-0
    mm = []
    for name, arr in (("nan_first", np.array([np.nan, 1.0, -2.0, 3.0], np.float32)),
                      ("nan_inside", np.array([0.5, np.nan, -2.0, 3.0], np.float32)),
                      ("zeros", np.array([0.0, -0.0, 0.0, -0.0], np.float32)),
                      ("neg_zero_first", np.array([-0.0, 0.0, 1.0, -1.0], np.float64))):
        a = arr.reshape(1, 1, 4)
        lo, hi = oracle.ref_min_max(a)
        mm.append({"name": name, "values": [repr(float(x)) for x in arr], "dtype": arr.dtype.name,
                   "min": repr(lo), "max": repr(hi)})
    meta["minmax"] = mm

    # 8. 64-bit integers beyond 2^53: double(v) rounds (several values share one
    #    double), isolevels equal to or one double away from data values, and
    #    uint64 values above 2^63 (MarchingCubes.cpp:299-316, :326-356)
    rng = np.random.default_rng(53)
    pk = Packer()
    cases = []
    shape = (6, 5, 7)
    n = int(np.prod(shape))
    i62, u63 = np.int64(1) << np.int64(62), np.uint64(1) << np.uint64(63)
    near = (i62 + rng.integers(-3000, 3000, n) * 512).astype(np.int64)
    # two far values keep float(min) != float(max): with a float range of one
    # value the reference's colour lookup divides 0 by 0 (MapperBase.cpp:69-78,
    # ColorMap.cpp:1045-1047) and indexes its table with size_t(NaN)
    near[0], near[-1] = i62 - (1 << 45), i62 + (1 << 45)
    big = [
        ("int64_near_2p62", near, None),
        ("int64_wide", rng.integers(-(1 << 62), 1 << 62, n, dtype=np.int64), None),
        ("int64_near_min", (np.int64(-(1 << 63)) + rng.integers(0, 1 << 41, n)).astype(np.int64), None),
        ("uint64_above_2p63", u63 + rng.integers(0, 1 << 62, n, dtype=np.int64).astype(np.uint64), None),
        ("uint64_near_max", np.uint64(0xFFFFFFFFFFFFFFFF) - rng.integers(0, 1 << 40, n, dtype=np.int64).astype(np.uint64),
         None),
        ("uint64_clamped", u63 + rng.integers(0, 1 << 62, n, dtype=np.int64).astype(np.uint64), 0.3),
    ]
    for name, flat, clamp in big:
        v = flat.reshape(shape)
        d = np.sort(v.astype(np.float64).ravel())
        mid = float(d[n // 2])
        mm = None
        if clamp is not None:
            lo, hi = float(d[0]), float(d[-1])
            mm = (lo + clamp * (hi - lo), hi - clamp * (hi - lo))
        for iso in (mid, float(np.nextafter(mid, np.inf)), float(np.nextafter(mid, -np.inf))):
            pk.add("in_bytes", np.frombuffer(v.tobytes(), dtype=np.uint8))
            for normal, dup in MODES:
                r = oracle.ref_marching_cubes(v, iso, normal, dup, min_max_values=mm)
                mk = f"n{normal}_d{int(dup)}"
                pk.add(mk + "_coords", r["coords"])
                pk.add(mk + "_conn", r["connections"])
                pk.add(mk + "_normals", r["normals"])
            cases.append({"name": name, "iso": iso, "dtype": v.dtype.name, "shape": list(shape), "min_max": mm})
    pk.save(os.path.join(HERE, "bigint.npz"))
    meta["bigint"] = cases

    # 9. kvs::SlicePlane over structured volumes: every dtype (the colour value
    #    follows C++'s promotions, unsigned differences wrap), planes given by
    #    coefficients and by point / normal, two colour maps, user min/max and
    #    object coordinates
    rng = np.random.default_rng(91)
    pk = Packer()
    cases = []
    for t in range(24):
        dtype = DTYPES[t % len(DTYPES)]
        shape = tuple(int(s) for s in rng.integers(3, 16, 3))
        if np.dtype(dtype).kind == "f":
            v = (rng.standard_normal(shape) * 40).astype(dtype)
        else:
            info = np.iinfo(dtype)
            v = rng.integers(max(int(info.min), -120), min(int(info.max), 250) + 1, shape).astype(dtype)
        nz, ny, nx = shape
        if t % 3 == 0:
            point = [float(np.float32(rng.uniform(0, nx - 1))), float(np.float32(rng.uniform(0, ny - 1))),
                     float(np.float32(rng.uniform(0, nz - 1)))]
            normal = [float(np.float32(x)) for x in rng.standard_normal(3)]
            coef = oracle.ref_slice_plane_coefficients(point, normal)
            plane = {"point": point, "normal": normal}
        elif t % 3 == 1:
            # an axis-aligned plane through nodes: ties (plane == 0) count as below
            coef = [0.0, 0.0, 1.0, -float(nz // 2)]
            plane = {"coefficients": coef}
        else:
            coef = [float(np.float32(x)) for x in rng.standard_normal(3)] + [float(np.float32(rng.uniform(-5, 5)))]
            plane = {"coefficients": coef}
        mm = None
        coords = None
        if t % 4 == 1:
            lo, hi = float(v.min()), float(v.max())
            mm = (lo + 0.1 * (hi - lo), hi - 0.1 * (hi - lo))
        if t % 4 == 2:
            coords = ([float(rng.uniform(-3, 3)) for _ in range(3)], [float(rng.uniform(4, 9)) for _ in range(3)])
        kind = t % 2
        r = oracle.ref_slice_plane(v, coef, tf_kind=kind, min_max_values=mm, coords=coords)
        pk.add("in_bytes", np.frombuffer(v.tobytes(), dtype=np.uint8))
        pk.add("coords", r["coords"])
        pk.add("normals", r["normals"])
        pk.add("colors", r["colors"])
        cases.append({"dtype": np.dtype(dtype).name, "shape": list(shape), "coefficients": [float(c) for c in coef],
                      "plane": plane, "tf_kind": kind, "min_max": mm, "coords": coords,
                      "color_type": r["color_type"], "normal_type": r["normal_type"],
                      "polygon_type": r["polygon_type"], "opacities": [int(o) for o in r["opacities"]],
                      "volume_min_max": [r["volume_min_value"], r["volume_max_value"]]})
    pk.save(os.path.join(HERE, "slice.npz"))
    meta["slice"] = cases

    # 10. the unstructured branch of kvs::Isosurface (MarchingTetrahedra /
    #     Hexahedra / Pyramid / Prism) and kvs::SlicePlane over unstructured
    #     cells; their case tables pinned by hash
    meta["cell_tables"] = {}
    for cell, name in enumerate(("tetrahedra", "hexahedra", "pyramid", "prism")):
        tri, vert = oracle.ref_unstructured_tables(cell)
        meta["cell_tables"][name] = {"TriangleID_sha256": sha(tri.astype("<i4")),
                                     "VertexID_sha256": sha(vert.astype("<i4")),
                                     "TriangleID_shape": list(tri.shape), "VertexID_shape": list(vert.shape)}
    pk = Packer()
    cases = []
    for ci, cell_type in enumerate((datasets.TETRAHEDRA, datasets.HEXAHEDRA, datasets.PYRAMID, datasets.PRISM)):
        for t, dtype in enumerate((np.float32, np.float64, np.uint8, np.int16, np.uint32)):
            dims = (5 + t, 4 + (t + ci) % 3, 6 - t % 2)
            v, c, k = datasets.unstructured_mesh(dims, cell_type, seed=100 * ci + t, dtype=dtype)
            hi = float(v.max())
            iso = 0.45 * hi if np.dtype(dtype).kind == "f" else float(int(0.45 * hi)) + 0.5
            mm = None
            if t == 3:
                lo = float(v.min())
                mm = (lo + 0.1 * (hi - lo), hi - 0.1 * (hi - lo))
            entry = {"cell_type": cell_type, "dtype": np.dtype(dtype).name, "dims": list(dims), "iso": iso,
                     "min_max": mm, "n_nodes": int(v.size), "n_conn": int(k.size)}
            pk.add("values", np.frombuffer(v.tobytes(), dtype=np.uint8))
            pk.add("coords", c)
            pk.add("conn", k)
            for normal, dup in MODES:
                r = oracle.ref_unstructured_isosurface(v, c, k, cell_type, iso, normal, dup, min_max_values=mm)
                mk = f"n{normal}_d{int(dup)}"
                pk.add(mk + "_coords", r["coords"])
                pk.add(mk + "_normals", r["normals"])
                entry[mk] = {"polygon_type": r["polygon_type"], "normal_type": r["normal_type"],
                             "color_type": r["color_type"], "colors": [int(x) for x in r["colors"]],
                             "opacities": [int(x) for x in r["opacities"]]}
            entry["volume_min_max"] = [r["volume_min_value"], r["volume_max_value"]]
            entry["min_obj"], entry["max_obj"] = r["min_obj"], r["max_obj"]
            coef = [0.3 - 0.1 * t, -0.2, 0.5 + 0.05 * ci, -0.4 * dims[2]]
            coef = [float(np.float32(x)) for x in coef]
            r = oracle.ref_unstructured_slice_plane(v, c, k, cell_type, coef, tf_kind=t % 2, min_max_values=mm)
            pk.add("slice_coords", r["coords"])
            pk.add("slice_normals", r["normals"])
            pk.add("slice_colors", r["colors"])
            entry["slice"] = {"coefficients": coef, "tf_kind": t % 2, "color_type": r["color_type"]}
            cases.append(entry)
    pk.save(os.path.join(HERE, "unstructured.npz"))
    meta["unstructured"] = cases

    # 11. KVSML volume files written by the reference's own writer
    #     (StructuredVolumeObject::write), and what its reader makes of them;
    #     two hand-edited variants: a big-endian copy and an object_coord attribute
    kdir = os.path.join(HERE, "kvsml")
    os.makedirs(kdir, exist_ok=True)
    rng = np.random.default_rng(111)
    files = []
    blob = datasets.blobs((23, 17, 11), seed=5)
    specs = [
        ("f32_binary", blob, False, True, None),
        ("u8_ascii_internal", rng.integers(0, 256, (6, 7, 9)).astype(np.uint8), True, False, None),
        ("i16_ascii_external", rng.integers(-300, 300, (5, 8, 7)).astype(np.int16), True, True, None),
        ("f64_binary", (rng.standard_normal((9, 6, 8)) * 3).astype(np.float64), False, True, None),
        ("f32_ascii_internal", (rng.standard_normal((4, 5, 6)) * 7).astype(np.float32), True, False, None),
        ("u16_binary_rectilinear", rng.integers(0, 4000, (7, 6, 5)).astype(np.uint16), False, True,
         np.concatenate([np.cumsum(rng.uniform(0.5, 2.0, n)) for n in (5, 6, 7)]).astype(np.float32)),
    ]
    for name, vol, ascii_, external, rect in specs:
        path = os.path.join(kdir, name + ".kvsml")
        oracle.ref_write_kvsml_volume(vol, path, ascii=ascii_, external=external, rect_coords=rect)
        files.append(name)
    # big-endian copy of f32_binary: byte-swapped data, endian="big"
    with open(os.path.join(kdir, "f32_binary.kvsml")) as f:
        text = f.read()
    dat = [t.split('"')[0] for t in text.split('file="')[1:]][0]
    raw = np.fromfile(os.path.join(kdir, dat), dtype="<f4")
    raw.astype(">f4").tofile(os.path.join(kdir, "f32_big_value.dat"))
    with open(os.path.join(kdir, "f32_big.kvsml"), "w") as f:
        f.write(text.replace(dat, "f32_big_value.dat").replace('endian="little"', 'endian="big"'))
    files.append("f32_big")
    # object_coord on the Object tag
    with open(os.path.join(kdir, "f32_ascii_internal.kvsml")) as f:
        text = f.read()
    with open(os.path.join(kdir, "f32_object_coord.kvsml"), "w") as f:
        f.write(text.replace('type="StructuredVolumeObject"',
                             'type="StructuredVolumeObject" object_coord="-1.5 0.25 2 4.5 8.75 9.125"', 1))
    files.append("f32_object_coord")
    pk = Packer()
    cases = []
    for name in files:
        r = oracle.ref_read_kvsml_volume(os.path.join(kdir, name + ".kvsml"))
        pk.add("values", np.frombuffer(r["values"].tobytes(), dtype=np.uint8))
        cases.append({"name": name, "dtype": r["values"].dtype.name, "shape": list(r["values"].shape),
                      "min_value": r["min_value"], "max_value": r["max_value"], "grid_type": r["grid_type"],
                      "min_obj": r["min_obj"], "max_obj": r["max_obj"]})
    pk.save(os.path.join(HERE, "kvsml.npz"))
    meta["kvsml"] = cases

    # 12. STL files the reference's PolygonExporter<kvs::Stl> + Stl::write make
    #     of marching-cubes surfaces in the three modes, binary and ascii
    sdir = os.path.join(HERE, "stl")
    os.makedirs(sdir, exist_ok=True)
    stl = []
    for vol_name, vol, iso, ascii_ in (("blobs20", datasets.blobs(20, seed=6), 0.45, False),
                                        ("blobs10", datasets.blobs(10, seed=7, nblobs=2), 0.4, True)):
        for normal, dup in MODES:
            r = oracle.ref_marching_cubes(vol, iso, normal, dup)
            name = f"{vol_name}_n{normal}_d{int(dup)}_{'ascii' if ascii_ else 'binary'}.stl"
            oracle.ref_export_stl(r, r["normal_type"], os.path.join(sdir, name), ascii=ascii_)
            stl.append({"file": name, "volume": vol_name, "iso": iso, "normal": normal, "dup": dup, "ascii": ascii_,
                        "sha256": hashlib.sha256(open(os.path.join(sdir, name), "rb").read()).hexdigest()})
    meta["stl"] = stl

    # 13. AVS field files written by the reference's AVSField::write, and what
    #     its StructuredVolumeImporter makes of them; hand-edited variants: no
    #     "field" line (the reader's default, uniform) with extra comments, and
    #     a rectilinear field (refused by the importer)
    os.environ.setdefault("USER", "kvs")   # AVSField::write puts getenv("USER") in the header
    adir = os.path.join(HERE, "avs")
    os.makedirs(adir, exist_ok=True)
    rng = np.random.default_rng(131)
    blob = datasets.blobs((19, 13, 11), seed=9)
    aspecs = [
        ("u8_ext", (blob * 255).astype(np.uint8), ((0.5, -1.0, 2.0), (10.0, 12.5, 20.25))),
        ("i16", rng.integers(-3000, 3000, (6, 9, 7)).astype(np.int16), None),
        ("i32", rng.integers(-10 ** 6, 10 ** 6, (5, 4, 8)).astype(np.int32), None),
        ("f32", blob.astype(np.float32), ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))),
        ("f64", (rng.standard_normal((7, 5, 6)) * 4).astype(np.float64), None),
    ]
    avs = []
    for name, vol, ext in aspecs:
        oracle.ref_write_avs_volume(vol, os.path.join(adir, name + ".fld"), ext=ext)
        avs.append(name)
    with open(os.path.join(adir, "f32.fld"), "rb") as f:
        raw = f.read()
    head, data = raw.split(b"\f\f", 1)
    with open(os.path.join(adir, "f32_nofield.fld"), "wb") as f:
        f.write(head.replace(b"field  = uniform\n", b"# a comment line = 7\n") + b"\f\f" + data)
    avs.append("f32_nofield")
    with open(os.path.join(adir, "f32_rectilinear.fld"), "wb") as f:
        f.write(head.replace(b"field  = uniform", b"field  = rectilinear") + b"\f\f" + data)
    pk = Packer()
    cases = []
    for name in avs:
        r = oracle.ref_read_avs_volume(os.path.join(adir, name + ".fld"))
        pk.add("values", np.frombuffer(r["values"].tobytes(), dtype=np.uint8))
        cases.append({"name": name, "dtype": r["values"].dtype.name, "shape": list(r["values"].shape),
                      "min_value": r["min_value"], "max_value": r["max_value"], "grid_type": r["grid_type"],
                      "min_obj": r["min_obj"], "max_obj": r["max_obj"], "min_ext": r["min_ext"],
                      "max_ext": r["max_ext"]})
    try:
        oracle.ref_read_avs_volume(os.path.join(adir, "f32_rectilinear.fld"))
        refused = False
    except RuntimeError:
        refused = True
    pk.save(os.path.join(HERE, "avs.npz"))
    meta["avs"] = {"cases": cases, "refused": {"f32_rectilinear": refused}}

    # 14. KVSML polygon files the reference's PolygonExporter<KVSMLPolygonObject>
    #     + write make, in its three writing types: marching-cubes surfaces in the
    #     three modes (one colour and opacity: <DataValue>s), an empty one, and a
    #     slice plane (per-vertex colours)
    kdir = os.path.join(HERE, "kvsml_polygon")
    os.makedirs(kdir, exist_ok=True)
    vol = datasets.blobs(10, seed=7, nblobs=2)
    coef = [0.3, -0.5, 0.8, -3.0]
    meshes = [(f"mc_n{n}_d{int(d)}", oracle.ref_marching_cubes(vol, 0.4, n, d), {"kind": "mc", "iso": 0.4,
               "normal": n, "dup": d}) for n, d in MODES]
    meshes.append(("mc_empty", oracle.ref_marching_cubes(vol, 5.0, 1, True),
                   {"kind": "mc", "iso": 5.0, "normal": 1, "dup": True}))
    meshes.append(("slice", oracle.ref_slice_plane(vol, coef), {"kind": "slice", "coef": coef, "tf_kind": 0}))
    kp = []
    for name, mesh, spec in meshes:
        for wt in (0, 1, 2):
            stem = f"{name}_w{wt}"
            oracle.ref_export_kvsml_polygon(mesh, os.path.join(kdir, stem + ".kvsml"), wt)
            files = sorted(f for f in os.listdir(kdir) if f == stem + ".kvsml" or f.startswith(stem + "_"))
            kp.append({"file": stem + ".kvsml", "writing": wt, "files": files,
                       "color": [int(x) for x in mesh["colors"][:3]] if mesh["colors"].size == 3 else None,
                       "opacity": int(mesh["opacities"][0]) if mesh["opacities"].size else -1, **spec})
    meta["kvsml_polygon"] = kp

    # 15. PLY files the reference's PolygonExporter<kvs::Ply> + Ply::write make
    #     of surfaces without connections, ascii and binary (with connections
    #     its writer stores face indices through a null pointer, Ply.cpp:475-478).
    #     Its ply_close also frees pointers the writing path never set
    #     (PlyFile.cpp:1345-1380): it runs in a child process whose allocator
    #     hands out zeroed blocks (no per-thread cache, perturb byte 255)
    env = dict(os.environ, GLIBC_TUNABLES="glibc.malloc.tcache_count=0:glibc.malloc.perturb=255")
    subprocess.run([sys.executable, os.path.abspath(__file__), "ply"], env=env, check=True)
    with open(os.path.join(HERE, "ply", "cases.json")) as f:
        meta["ply"] = json.load(f)
    os.remove(os.path.join(HERE, "ply", "cases.json"))

    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", sorted(os.listdir(HERE)))


def ply_goldens():
    """Section 15, run in its own process (see main)."""
    pdir = os.path.join(HERE, "ply")
    os.makedirs(pdir, exist_ok=True)
    vol = datasets.blobs(10, seed=7, nblobs=2)
    vn = oracle.ref_marching_cubes(vol, 0.4, 0, False)
    vn_noconn = dict(vn, connections=np.zeros(0, np.uint32))
    meshes = [("mc_n1_d1", oracle.ref_marching_cubes(vol, 0.4, 1, True), {"kind": "mc", "iso": 0.4, "normal": 1,
                                                                          "dup": True}),
              ("mc_n0_d0_noconn", vn_noconn, {"kind": "mc", "iso": 0.4, "normal": 0, "dup": False,
                                              "drop_connections": True}),
              ("mc_empty", oracle.ref_marching_cubes(vol, 5.0, 1, True), {"kind": "mc", "iso": 5.0, "normal": 1,
                                                                          "dup": True}),
              ("slice", oracle.ref_slice_plane(vol, [0.3, -0.5, 0.8, -3.0]), {"kind": "slice",
                                                                             "coef": [0.3, -0.5, 0.8, -3.0]})]
    cases = []
    for name, mesh, spec in meshes:
        for binary in (False, True):
            fn = f"{name}_{'binary' if binary else 'ascii'}.ply"
            oracle.ref_export_ply(mesh, os.path.join(pdir, fn), binary=binary)
            cases.append({"file": fn, "binary": binary,
                          "color": [int(x) for x in mesh["colors"]] if mesh["colors"].size == 3 else None,
                          "sha256": hashlib.sha256(open(os.path.join(pdir, fn), "rb").read()).hexdigest(), **spec})
    with open(os.path.join(pdir, "cases.json"), "w") as f:
        json.dump(cases, f)


if __name__ == "__main__" and len(sys.argv) > 1 and sys.argv[1] == "ply":
    ply_goldens()
    sys.exit(0)

if __name__ == "__main__":
    main()
