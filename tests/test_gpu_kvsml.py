"""Volume ingestion: KVSML (kmc_kvsml_read_volume) against what the
reference's own reader (StructuredVolumeObject::read) makes of files its
writer produced (tests/golden/kvsml/), and AVS fields (kmc_avs_read_volume)
against its StructuredVolumeImporter on files its AVSField::write produced
(tests/golden/avs/): values bit for bit, min/max values (a KVSML file's
rounded attributes when present) and min/max object coordinates, into device
and host memory; and the isosurface of the device-resident volume against the
oracle."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from kvs_amd import kmc
from tests import golden_io
from tests.conftest import assert_same_surface, bits_equal

pytestmark = pytest.mark.gpu

KDIR = os.path.join(golden_io.GOLDEN, "kvsml")


def cases():
    p = golden_io.Packed("kvsml.npz")
    for i, c in enumerate(golden_io.meta()["kvsml"]):
        v = np.frombuffer(p.get("values", i).tobytes(), dtype=np.dtype(c["dtype"])).reshape(c["shape"])
        yield c, v


@pytest.mark.parametrize("memory", [kmc.MEMORY_DEVICE, kmc.MEMORY_HOST])
def test_kvsml_read_matches_reference_reader(memory):
    for c, want in cases():
        print("case", c["name"], memory, flush=True)
        kv = kmc.KvsmlVolume(os.path.join(KDIR, c["name"] + ".kvsml"), memory=memory)
        try:
            v = kv.volume
            assert v.memory == memory and v.kind == kmc.STRUCTURED and v.veclen == 1
            assert list(v.resolution) == c["shape"][::-1]
            assert bits_equal(kv.values(), want), c["name"]
            assert (v.min_value, v.max_value) == (c["min_value"], c["max_value"]), c["name"]
            assert list(v.min_coord) == c["min_obj"] and list(v.max_coord) == c["max_obj"], c["name"]
        finally:
            kv.close()


@pytest.mark.parametrize("memory", [kmc.MEMORY_DEVICE, kmc.MEMORY_HOST])
def test_large_external_arrays_in_both_byte_orders(tmp_path, memory):
    # arrays of several 64 MiB staging buffers, each range read by several
    # threads that byte-swap their own parts; the small fixtures above pin the
    # byte-order rule against the reference's reader, this the split reads
    rng = np.random.default_rng(3)
    for dtype, name, (nz, ny, nx) in ((np.float64, "double", (97, 301, 331)), (np.int16, "short", (61, 1031, 1201))):
        v = (rng.standard_normal((nz, ny, nx)) * 1000).astype(dtype)
        for endian in ("little", "big"):
            dat = tmp_path / f"v_{endian}.dat"
            v.astype(v.dtype.newbyteorder("<" if endian == "little" else ">")).tofile(dat)
            path = tmp_path / f"v_{endian}.kvsml"
            path.write_text('<?xml version="1.0" ?>\n<KVSML>\n<Object type="StructuredVolumeObject">\n'
                            f'<StructuredVolumeObject grid_type="uniform" resolution="{nx} {ny} {nz}">\n'
                            '<Node>\n<Value veclen="1">\n'
                            f'<DataArray type="{name}" format="binary" file="{dat.name}" endian="{endian}" />\n'
                            '</Value>\n</Node>\n</StructuredVolumeObject>\n</Object>\n</KVSML>\n')
            print("case", name, endian, v.nbytes, memory, flush=True)
            kv = kmc.KvsmlVolume(str(path), memory=memory)
            try:
                got = kv.values().reshape(v.shape)
                assert got.dtype == v.dtype and np.array_equal(got, v), (name, endian)
                assert (kv.volume.min_value, kv.volume.max_value) == (float(v.min()), float(v.max())), (name, endian)
            finally:
                kv.close()


def test_isosurface_of_a_kvsml_volume_in_hbm():
    lib = kmc.load()
    for c, want in cases():
        kv = kmc.KvsmlVolume(os.path.join(KDIR, c["name"] + ".kvsml"))
        try:
            v = kv.volume
            lo, hi = c["min_value"], c["max_value"]
            iso = 0.5 * (lo + hi)
            if np.dtype(c["dtype"]).kind in "ui":
                iso = float(np.floor(iso)) + 0.5
            for normal, dup in [(kmc.POLYGON_NORMAL, True), (kmc.VERTEX_NORMAL, False)]:
                s = kmc.Surface()
                st = lib.kmc_extract(ctypes.byref(v), iso, normal, int(dup), kmc.MEMORY_HOST, None, ctypes.byref(s))
                assert st == kmc.OK
                try:
                    got = kmc.surface_to_numpy(s)
                finally:
                    lib.kmc_surface_release(ctypes.byref(s))
                exp = oracle.extract(want, iso, normal, dup, min_max_values=(lo, hi),
                                     coords=(c["min_obj"], c["max_obj"]))
                assert_same_surface(got, exp, f"{c['name']} mode {normal},{dup}")
        finally:
            kv.close()


def test_kvsml_errors():
    with pytest.raises(kmc.KmcError) as e:
        kmc.KvsmlVolume(os.path.join(KDIR, "missing.kvsml"))
    assert e.value.status == kmc.ERR_FILE
    with pytest.raises(kmc.KmcError) as e:
        kmc.KvsmlVolume(os.path.join(KDIR, "f32_binary_value.dat"))   # not a KVSML extension
    assert e.value.status == kmc.ERR_FILE


ADIR = os.path.join(golden_io.GOLDEN, "avs")


def avs_cases():
    p = golden_io.Packed("avs.npz")
    for i, c in enumerate(golden_io.meta()["avs"]["cases"]):
        v = np.frombuffer(p.get("values", i).tobytes(), dtype=np.dtype(c["dtype"])).reshape(c["shape"])
        yield c, v


@pytest.mark.parametrize("memory", [kmc.MEMORY_DEVICE, kmc.MEMORY_HOST])
def test_avs_read_matches_reference_importer(memory):
    # kmc_avs_read_volume against kvs::StructuredVolumeImporter on files the
    # reference's AVSField::write produced (and a copy without a field line)
    for c, want in avs_cases():
        print("avs case", c["name"], memory, flush=True)
        kv = kmc.KvsmlVolume(os.path.join(ADIR, c["name"] + ".fld"), memory=memory)
        try:
            v = kv.volume
            assert v.memory == memory and v.kind == kmc.STRUCTURED and v.veclen == 1
            assert list(v.resolution) == c["shape"][::-1]
            assert bits_equal(kv.values(), want), c["name"]
            assert (v.min_value, v.max_value) == (c["min_value"], c["max_value"]), c["name"]
            assert list(v.min_coord) == c["min_obj"] and list(v.max_coord) == c["max_obj"], c["name"]
        finally:
            kv.close()


def test_avs_refusals_and_isosurface():
    # the importer takes uniform fields only
    assert golden_io.meta()["avs"]["refused"]["f32_rectilinear"]
    with pytest.raises(kmc.KmcError) as e:
        kmc.KvsmlVolume(os.path.join(ADIR, "f32_rectilinear.fld"))
    assert e.value.status == kmc.ERR_FILE
    with pytest.raises(kmc.KmcError) as e:
        kmc.KvsmlVolume(os.path.join(ADIR, "missing.fld"))
    assert e.value.status == kmc.ERR_FILE
    # an isosurface of the device-resident volume equals the oracle's
    lib = kmc.load()
    c, want = next(x for x in avs_cases() if x[0]["name"] == "f32")
    kv = kmc.KvsmlVolume(os.path.join(ADIR, "f32.fld"))
    try:
        s = kmc.Surface()
        st = lib.kmc_extract(ctypes.byref(kv.volume), 0.5, kmc.VERTEX_NORMAL, 0, kmc.MEMORY_HOST, None, ctypes.byref(s))
        assert st == kmc.OK
        try:
            got = kmc.surface_to_numpy(s)
        finally:
            lib.kmc_surface_release(ctypes.byref(s))
        exp = oracle.extract(want, 0.5, kmc.VERTEX_NORMAL, False, min_max_values=(c["min_value"], c["max_value"]),
                             coords=(c["min_obj"], c["max_obj"]))
        assert_same_surface(got, exp, "avs f32")
    finally:
        kv.close()
