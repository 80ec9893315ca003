/*
 * unstructured_oracle.c -- THE ORACLE (test infrastructure only, never part of the product).
 *
 * A plain-C, single-threaded restatement of the unstructured branch of
 * kvs::Isosurface (Isosurface.cpp:192-257) -- kvs::MarchingTetrahedra,
 * MarchingHexahedra, MarchingPyramid and MarchingPrism -- and of
 * kvs::SlicePlane over an unstructured volume, reproducing the reference bit
 * for bit.  Loaded only by tests/.
 *
 * Reference lines followed (Source/Core/Visualization/...):
 *   Mapper/MarchingTetrahedra.cpp:155-286 (duplicated extraction, table index, interpolate_vertex)
 *   Mapper/MarchingHexahedra.cpp:140-286, Mapper/MarchingPyramid.cpp:140-317,
 *   Mapper/MarchingPrism.cpp:140-283 (the same per cell type; the pyramid's
 *     ambiguous cases 10, 11, 20, 21 choose 32..35 by the sign of E, :289-317)
 *   Mapper/SlicePlane.cpp:322-800 extract_{tetrahedra,hexahedra,pyramid,prism}_plane
 *   Object/UnstructuredVolumeObject.cpp:82-165 GetMinMaxValues (connection order)
 *   Object/UnstructuredVolumeObject.cpp:396-460 updateMinMaxCoords (connection order)
 *   Matrix/Vector3.h:479-485 cross
 * Tables: cell_tables_oracle.h (Marching*Table.cpp).
 *
 * Differences from the reference, outside its defined domain: node indices
 * are 64-bit (the reference narrows them to int and computes 3 * index in int,
 * MarchingTetrahedra.cpp:262-268 and the same lines of the others).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "cell_tables_oracle.h"

enum { UO_I8, UO_I16, UO_I32, UO_I64, UO_U8, UO_U16, UO_U32, UO_U64, UO_F32, UO_F64 };
/* kvs::UnstructuredVolumeObject::CellType */
enum { UO_TETRA = 1, UO_HEXA = 2, UO_PYRAMID = 5, UO_PRISM = 7 };

typedef struct {
    float*   coords;  uint64_t n_coords;
    float*   normals; uint64_t n_normals;
    uint8_t* colors;  uint64_t n_colors;
    uint64_t n_triangles;
} UMesh;

static double uvalue(const void* p, int dt, size_t i)
{
    switch (dt) {
    case UO_I8:  return (double)((const int8_t*)p)[i];
    case UO_I16: return (double)((const int16_t*)p)[i];
    case UO_I32: return (double)((const int32_t*)p)[i];
    case UO_I64: return (double)((const int64_t*)p)[i];
    case UO_U8:  return (double)((const uint8_t*)p)[i];
    case UO_U16: return (double)((const uint16_t*)p)[i];
    case UO_U32: return (double)((const uint32_t*)p)[i];
    case UO_U64: return (double)((const uint64_t*)p)[i];
    case UO_F32: return (double)((const float*)p)[i];
    default:     return ((const double*)p)[i];
    }
}

typedef struct {
    int nodes;                  /* nodes per cell */
    int ncases;
    const char* const* cases;   /* edge ids per case */
    const char* edges;          /* two local nodes per edge */
    int swap12;                 /* 1: triangle edges (T[i], T[i+2], T[i+1]) */
    int all_ones;               /* the table index skipped besides 0 */
} CellKind;

static int cell_kind(int cell_type, int slice, CellKind* k)
{
    switch (cell_type) {
    case UO_TETRA:   *k = (CellKind){ 4, 16, kOracleTetCases, kOracleTetEdges, 0, 15 }; break;
    case UO_HEXA:    *k = (CellKind){ 8, 256, kOracleHexCases, kOracleHexEdges, 1, 255 }; break;
    case UO_PYRAMID: *k = (CellKind){ 5, 36, kOraclePyrCases, kOraclePyrEdges, 1, slice ? 31 : 36 }; break;
    case UO_PRISM:   *k = (CellKind){ 6, 64, kOraclePrismCases, kOraclePrismEdges, 0, 63 }; break;
    default: return 0;
    }
    if (slice) k->swap12 = 0;   /* SlicePlane takes (T[i], T[i+1], T[i+2]) for every cell type */
    return 1;
}

/* The cell's nodes in the mapper's local order: the hexahedra take
 * connections 4..7 as nodes 0..3 and 0..3 as 4..7 (MarchingHexahedra.cpp:163-170). */
static void local_nodes(int cell_type, const uint32_t* conn, uint64_t cell, int nodes, uint64_t out[8])
{
    const uint32_t* c = conn + cell * (uint64_t)nodes;
    for (int i = 0; i < nodes; ++i) out[i] = c[i];
    if (cell_type == UO_HEXA)
        for (int i = 0; i < 8; ++i) out[i] = c[(i + 4) & 7];
}

static int hexd(char c) { return (c >= '0' && c <= '9') ? c - '0' : c - 'a' + 10; }

#define UMINMAX_BODY(T)                                                    \
    {                                                                      \
        const T* v = (const T*)values;                                     \
        T mn = v[0], mx = v[0];                                            \
        for (uint64_t i = 0; i < nconn; ++i) {                             \
            const T x = v[conn[i]];                                        \
            mn = (x < mn) ? x : mn;                                        \
            mx = (x > mx) ? x : mx;                                        \
        }                                                                  \
        *out_min = (double)mn; *out_max = (double)mx;                      \
        return 0;                                                          \
    }

/* GetMinMaxValues<T> with connections (UnstructuredVolumeObject.cpp:123-141):
 * seeded with the first value of the array, then every connection entry in
 * order through Math::Min / Math::Max. */
int oracle_unstructured_min_max(const void* values, int dtype, const uint32_t* conn, uint64_t nconn,
                                double* out_min, double* out_max)
{
    switch (dtype) {
    case UO_I8:  UMINMAX_BODY(int8_t)
    case UO_I16: UMINMAX_BODY(int16_t)
    case UO_I32: UMINMAX_BODY(int32_t)
    case UO_I64: UMINMAX_BODY(int64_t)
    case UO_U8:  UMINMAX_BODY(uint8_t)
    case UO_U16: UMINMAX_BODY(uint16_t)
    case UO_U32: UMINMAX_BODY(uint32_t)
    case UO_U64: UMINMAX_BODY(uint64_t)
    case UO_F32: UMINMAX_BODY(float)
    case UO_F64: UMINMAX_BODY(double)
    default: return 4;
    }
}

/* updateMinMaxCoords with connections (UnstructuredVolumeObject.cpp:425-455):
 * seeded with the first connected node, Math::Min(running, x). */
void oracle_unstructured_min_max_coords(const float* coords, const uint32_t* conn, uint64_t nconn, float minc[3],
                                        float maxc[3])
{
    for (int k = 0; k < 3; ++k) minc[k] = maxc[k] = coords[3 * (uint64_t)conn[0] + k];
    for (uint64_t i = 0; i < nconn; ++i)
        for (int k = 0; k < 3; ++k) {
            const float x = coords[3 * (uint64_t)conn[i] + k];
            minc[k] = (minc[k] < x) ? minc[k] : x;
            maxc[k] = (maxc[k] > x) ? maxc[k] : x;
        }
}

typedef struct { float* f; size_t n, cap; } UVec;

static int upush(UVec* v, float x)
{
    if (v->n == v->cap) {
        const size_t c = v->cap ? 2 * v->cap : 1024;
        float* p = (float*)realloc(v->f, c * sizeof(float));
        if (!p) return 0;
        v->f = p;
        v->cap = c;
    }
    v->f[v->n++] = x;
    return 1;
}

/* (v1 - v0) x (v2 - v0), Vector3::cross. */
static void cross_of(const float v0[3], const float v1[3], const float v2[3], float n[3])
{
    const float ax = v1[0] - v0[0], ay = v1[1] - v0[1], az = v1[2] - v0[2];
    const float bx = v2[0] - v0[0], by = v2[1] - v0[1], bz = v2[2] - v0[2];
    n[0] = ay * bz - az * by;
    n[1] = az * bx - ax * bz;
    n[2] = ax * by - ay * bx;
}

/* The duplicated extraction of MarchingTetrahedra / Hexahedra / Pyramid /
 * Prism.  duplication: after VertexNormal forced it off; without it the
 * reference extracts nothing.  minv / maxv: the volume's min / max (clamp
 * range and the Math::Equal exit).  Returns 0, 5 for an unknown cell type, 6
 * when out of memory. */
int oracle_unstructured_iso(const void* values, int dtype, const float* coords, const uint32_t* conn,
                            uint64_t ncells, int cell_type, double minv, double maxv, double iso, int duplication,
                            UMesh* out)
{
    memset(out, 0, sizeof(*out));
    CellKind K;
    if (!cell_kind(cell_type, 0, &K)) return 5;
    const double d = minv - maxv;
    if ((d < 0 ? -d : d) < 1.e-6 || !duplication) return 0;   /* Math::Equal exit; no shared-vertex path */
    UVec co = { 0, 0, 0 }, no = { 0, 0, 0 };
    int ok = 1;
    for (uint64_t cell = 0; cell < ncells && ok; ++cell) {
        uint64_t li[8];
        local_nodes(cell_type, conn, cell, K.nodes, li);
        int ti = 0;
        for (int k = 0; k < K.nodes; ++k)
            if (uvalue(values, dtype, li[k]) > iso) ti |= 1 << k;
        if (cell_type == UO_PYRAMID && (ti == 10 || ti == 11 || ti == 20 || ti == 21)) {
            /* calculate_special_table_index (MarchingPyramid.cpp:289-317) */
            const double v1 = uvalue(values, dtype, li[1]), v2 = uvalue(values, dtype, li[2]);
            const double v3 = uvalue(values, dtype, li[3]), v4 = uvalue(values, dtype, li[4]);
            const double A = (v1 + v2 + v3 + v4) / 4;
            const double B = (-v1 + v2 + v3 - v4) / 4;
            const double C = (-v1 - v2 + v3 + v4) / 4;
            const double D = (v1 - v2 + v3 - v4) / 4;
            const double E = (iso - A) * D + B * C;
            if (ti == 10 && E < 0) ti = 32;
            else if (ti == 11 && E >= 0) ti = 33;
            else if (ti == 20 && E >= 0) ti = 34;
            else if (ti == 21 && E < 0) ti = 35;
        }
        if (ti == 0 || ti == K.all_ones) continue;
        const char* es = K.cases[ti];
        const int ne = (int)strlen(es);
        for (int i = 0; i + 2 < ne && ok; i += 3) {
            const int e[3] = { hexd(es[i]), hexd(es[i + (K.swap12 ? 2 : 1)]), hexd(es[i + (K.swap12 ? 1 : 2)]) };
            float v[3][3];
            for (int j = 0; j < 3; ++j) {
                const uint64_t a = li[K.edges[2 * e[j]] - '0'], b = li[K.edges[2 * e[j] + 1] - '0'];
                double va = uvalue(values, dtype, a), vb = uvalue(values, dtype, b);
                va = va <= minv ? minv : (va <= maxv ? va : maxv);
                vb = vb <= minv ? minv : (vb <= maxv ? vb : maxv);
                double r = (iso - va) / (vb - va);
                r = r < 0 ? -r : r;
                for (int k = 0; k < 3; ++k) {
                    const float ca = coords[3 * a + k], cb = coords[3 * b + k];
                    const float diff = cb - ca;
                    const double t = r * (double)diff;
                    v[j][k] = (float)((double)ca + t);
                }
            }
            for (int j = 0; j < 3 && ok; ++j)
                for (int k = 0; k < 3 && ok; ++k) ok = upush(&co, v[j][k]);
            float n[3];
            cross_of(v[0], v[1], v[2], n);
            ok = ok && upush(&no, n[0]) && upush(&no, n[1]) && upush(&no, n[2]);
        }
    }
    if (!ok) { free(co.f); free(no.f); return 6; }
    out->coords = co.f; out->n_coords = co.n;
    out->normals = no.f; out->n_normals = no.n;
    out->n_triangles = no.n / 3;
    return 0;
}

/* SlicePlane's plane equation on a vertex (SlicePlane.cpp:955-971). */
static float uplane(const float c[4], const float* v)
{
    float s = c[0] * v[0];
    s = s + c[1] * v[1];
    s = s + c[2] * v[2];
    return s + c[3];
}

#define UVALUE_BODY(T, DIFF_T, ACC)                                         \
    {                                                                       \
        const T* v = (const T*)values;                                      \
        const DIFF_T d = (DIFF_T)v[i1] - (DIFF_T)v[i0];                     \
        const ACC prod = (ACC)r * (ACC)d;                                   \
        return (double)((ACC)v[i0] + prod);                                 \
    }

/* interpolate_value<T>(unstructured) (SlicePlane.cpp:1036-1050). */
static double uslice_value(const void* values, int dt, uint64_t i0, uint64_t i1, float r)
{
    switch (dt) {
    case UO_I8:  UVALUE_BODY(int8_t, int32_t, float)
    case UO_I16: UVALUE_BODY(int16_t, int32_t, float)
    case UO_I32: UVALUE_BODY(int32_t, int32_t, float)
    case UO_I64: UVALUE_BODY(int64_t, int64_t, float)
    case UO_U8:  UVALUE_BODY(uint8_t, int32_t, float)
    case UO_U16: UVALUE_BODY(uint16_t, int32_t, float)
    case UO_U32: UVALUE_BODY(uint32_t, uint32_t, float)
    case UO_U64: UVALUE_BODY(uint64_t, uint64_t, float)
    case UO_F32: UVALUE_BODY(float, float, float)
    default:     UVALUE_BODY(double, double, double)
    }
}

static int uround(float a) { return a >= 0 ? (int)(a + 0.5) : -(int)(0.5 - a); }
static float uclamp(float a, float lo, float hi) { return a <= lo ? lo : (a <= hi ? a : hi); }

static void ucolor(const uint8_t* table, uint32_t res, float lo, float hi, float value, uint8_t rgb[3])
{
    const float v0 = uclamp(value, lo, hi);
    const float rr = (float)(res - 1);
    const float v = (v0 - lo) / (hi - lo) * rr;
    const size_t s0 = (size_t)v;
    const size_t s1 = (s0 + 1 < (size_t)(res - 1)) ? s0 + 1 : (size_t)(res - 1);
    const float t = v - (float)s0;
    for (int k = 0; k < 3; ++k) {
        const float a = (float)table[3 * s0 + k] / 255.0f;
        const float b = (float)table[3 * s1 + k] / 255.0f;
        rgb[k] = (uint8_t)uround(uclamp(a * (1 - t) + b * t, 0.0f, 1.0f) * 255.0f);
    }
}

/* extract_{tetrahedra,hexahedra,pyramid,prism}_plane<T> (SlicePlane.cpp:322-800). */
int oracle_unstructured_slice(const void* values, int dtype, const float* coords, const uint32_t* conn,
                              uint64_t ncells, int cell_type, const float coef[4], const uint8_t* table,
                              uint32_t res, float lo, float hi, UMesh* out)
{
    memset(out, 0, sizeof(*out));
    CellKind K;
    if (!cell_kind(cell_type, 1, &K)) return 5;
    UVec co = { 0, 0, 0 }, no = { 0, 0, 0 };
    uint8_t* colors = NULL;
    size_t ncol = 0, capcol = 0;
    int ok = 1;
    for (uint64_t cell = 0; cell < ncells && ok; ++cell) {
        uint64_t li[8];
        local_nodes(cell_type, conn, cell, K.nodes, li);
        int ti = 0;
        for (int k = 0; k < K.nodes; ++k)
            if (uplane(coef, coords + 3 * li[k]) > 0.0) ti |= 1 << k;
        if (ti == 0 || ti == K.all_ones) continue;
        const char* es = K.cases[ti];
        const int ne = (int)strlen(es);
        for (int i = 0; i + 2 < ne && ok; i += 3) {
            const int e[3] = { hexd(es[i]), hexd(es[i + 1]), hexd(es[i + 2]) };
            float v[3][3];
            uint8_t rgb[3][3];
            for (int j = 0; j < 3; ++j) {
                const uint64_t a = li[K.edges[2 * e[j]] - '0'], b = li[K.edges[2 * e[j] + 1] - '0'];
                const float* pa = coords + 3 * a;
                const float* pb = coords + 3 * b;
                const float f0 = uplane(coef, pa), f1 = uplane(coef, pb);
                float r = f0 / (f1 - f0);
                r = r < 0 ? -r : r;
                const float s = 1.0f - r;
                for (int k = 0; k < 3; ++k) {
                    const float p = s * pa[k];
                    const float q = r * pb[k];
                    v[j][k] = p + q;
                }
                ucolor(table, res, lo, hi, (float)uslice_value(values, dtype, a, b, r), rgb[j]);
            }
            for (int j = 0; j < 3 && ok; ++j)
                for (int k = 0; k < 3 && ok; ++k) ok = upush(&co, v[j][k]);
            if (ncol + 9 > capcol) {
                capcol = capcol ? 2 * capcol : 4096;
                uint8_t* p = (uint8_t*)realloc(colors, capcol);
                if (!p) { ok = 0; break; }
                colors = p;
            }
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) colors[ncol++] = rgb[j][k];
            /* -(vertex2 - vertex0).cross(vertex1 - vertex0) */
            float n[3];
            cross_of(v[0], v[2], v[1], n);
            ok = ok && upush(&no, -n[0]) && upush(&no, -n[1]) && upush(&no, -n[2]);
        }
    }
    if (!ok) { free(co.f); free(no.f); free(colors); return 6; }
    out->coords = co.f; out->n_coords = co.n;
    out->normals = no.f; out->n_normals = no.n;
    out->colors = colors; out->n_colors = ncol;
    out->n_triangles = no.n / 3;
    return 0;
}

void oracle_unstructured_free(UMesh* m)
{
    free(m->coords); free(m->normals); free(m->colors);
    memset(m, 0, sizeof(*m));
}
