"""Write the case tables of the unstructured cell types from the reference's own
tables (MarchingTetrahedraTable.cpp, MarchingHexahedraTable.cpp,
MarchingPyramidTable.cpp, MarchingPrismTable.cpp, read through
oracle/_ref/libkvsref.so in the build container):

  kvs_amd/csrc/cell_tables.h        the device tables, one 64-bit word per case
                                     (triangle count in bits 60..63, nibble j =
                                     the j-th edge id in table order) and the
                                     edge endpoints;
  oracle/cell_tables_oracle.h        the oracle's tables, one hex string per case.

    make -C oracle && python tests/golden/make_cell_tables.py

tests/test_oracle.py pins both files against the SHA-256 of the reference
tables that make_golden.py records in golden.json.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

# cell index of ref_unstructured_tables: (name, TriangleID shape, VertexID shape)
CELLS = [("Tet", (16, 7), (6, 2)), ("Hex", (256, 16), (12, 2)), ("Pyr", (36, 13), (8, 2)), ("Prism", (64, 10), (9, 2))]


def reference_tables():
    lib = oracle.ref_lib()
    out = []
    for cell, (name, ts, vs) in enumerate(CELLS):
        tri = (ctypes.c_int32 * (ts[0] * ts[1]))()
        vert = (ctypes.c_int32 * (vs[0] * vs[1]))()
        lib.ref_unstructured_tables(cell, tri, vert)
        out.append((name, np.array(tri, dtype=np.int32).reshape(ts), np.array(vert, dtype=np.int32).reshape(vs)))
    return out


def word(row):
    edges = [int(e) for e in row if e >= 0]
    w = (len(edges) // 3) << 60
    for j, e in enumerate(edges):
        w |= e << (4 * j)
    return w


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libkvsref.so is missing: run `make -C oracle` first")
    tables = reference_tables()
    dev = ["/*",
           " * cell_tables.h -- case tables of the unstructured cell types, written by",
           " * tests/golden/make_cell_tables.py from the reference's tables",
           " * (Source/Core/Visualization/Mapper/Marching{Tetrahedra,Hexahedra,Pyramid,Prism}Table.cpp)",
           " * and pinned by their SHA-256 (tests/test_oracle.py).",
           " *",
           " * k<Cell>Tri[index]: triangle count in bits 60..63, nibble j = the j-th edge id",
           " * of the case in table order.  k<Cell>Edge[e]: the two local nodes of edge e,",
           " * first in the low nibble.",
           " */",
           "#pragma once",
           "#include <stdint.h>",
           "",
           "namespace kmc {",
           ""]
    orc = ["/*",
           " * cell_tables_oracle.h -- THE ORACLE's transcription of the unstructured case",
           " * tables (test infrastructure only), written by tests/golden/make_cell_tables.py",
           " * from the reference's Marching{Tetrahedra,Hexahedra,Pyramid,Prism}Table.cpp and",
           " * pinned by their SHA-256 (tests/test_oracle.py): per case, its edge ids in",
           " * table order as hex digits; per edge, its two local nodes as two digits.",
           " */",
           ""]
    for name, tri, vert in tables:
        dev.append(f"__constant__ const uint64_t k{name}Tri[{tri.shape[0]}] = {{")
        words = [word(r) for r in tri]
        for i in range(0, len(words), 4):
            dev.append("    " + ", ".join(f"0x{w:016x}ull" for w in words[i:i + 4]) + ",")
        dev.append("};")
        dev.append(f"__constant__ const uint8_t k{name}Edge[{vert.shape[0]}] = {{ "
                   + ", ".join(f"0x{int(b) << 4 | int(a):02x}" for a, b in vert) + " };")
        dev.append("")
        orc.append(f"static const char* const kOracle{name}Cases[{tri.shape[0]}] = {{")
        strs = ['"' + "".join(f"{int(e):x}" for e in r if e >= 0) + '"' for r in tri]
        for i in range(0, len(strs), 8):
            orc.append("    " + ", ".join(strs[i:i + 8]) + ",")
        orc.append("};")
        orc.append(f'static const char* const kOracle{name}Edges = "'
                   + "".join(f"{int(a)}{int(b)}" for a, b in vert) + '";')
        orc.append("")
    dev.append("}  // namespace kmc")
    with open(os.path.join(ROOT, "kvs_amd", "csrc", "cell_tables.h"), "w") as f:
        f.write("\n".join(dev) + "\n")
    with open(os.path.join(ROOT, "oracle", "cell_tables_oracle.h"), "w") as f:
        f.write("\n".join(orc) + "\n")
    print("wrote kvs_amd/csrc/cell_tables.h and oracle/cell_tables_oracle.h")


if __name__ == "__main__":
    main()
