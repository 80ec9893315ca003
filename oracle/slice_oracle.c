/*
 * slice_oracle.c -- THE ORACLE (test infrastructure only, never part of the product).
 *
 * A plain-C, single-threaded restatement of kvs::SlicePlane over a
 * kvs::StructuredVolumeObject (Source/Core/Visualization/Mapper/SlicePlane.cpp),
 * reproducing the reference bit for bit.  Loaded only by tests/.
 *
 * Reference lines followed (Source/Core/...):
 *   Visualization/Mapper/SlicePlane.cpp:165-280 extract_plane<T>(structured)
 *   Visualization/Mapper/SlicePlane.cpp:802-818 calculate_table_index
 *   Visualization/Mapper/SlicePlane.cpp:936-971 substitute_plane_equation (both forms)
 *   Visualization/Mapper/SlicePlane.cpp:981-990 interpolate_vertex
 *   Visualization/Mapper/SlicePlane.cpp:1001-1026 interpolate_value<T>(structured)
 *   Visualization/Mapper/ColorMap.cpp:1043-1054 ColorMap::at
 *   Image/RGBColor.cpp:23-31,125-134,170-175 Mix, from / to Vec3
 *   Matrix/Vector3.h:479-485 cross
 *   Visualization/Mapper/MarchingCubesTable.cpp (the same tables as mc_oracle.c)
 *
 * The colour table (ColorMap::table(), 3 bytes per entry) and its range (the
 * transfer function's, after MapperBase::setRange) are inputs: they come from
 * the caller exactly as the reference's ColorMap holds them.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int oracle_case_edges(int ci, int* edges16);   /* mc_oracle.c: TriangleID[ci] */

enum { SO_I8, SO_I16, SO_I32, SO_I64, SO_U8, SO_U16, SO_U32, SO_U64, SO_F32, SO_F64 };

typedef struct {
    float*   coords;  uint64_t n_coords;    /* 9 floats per triangle */
    float*   normals; uint64_t n_normals;   /* 3 floats per triangle */
    uint8_t* colors;  uint64_t n_colors;    /* 9 bytes per triangle (one RGB per vertex) */
    uint64_t n_triangles;
} SliceMesh;

/* Edge endpoints in table orientation (MarchingCubesTable.cpp:276-290). */
static const int kEdgeEnds[12][2][3] = {
    {{0,0,0},{1,0,0}}, {{1,0,0},{1,1,0}}, {{1,1,0},{0,1,0}}, {{0,1,0},{0,0,0}},
    {{0,0,1},{1,0,1}}, {{1,0,1},{1,1,1}}, {{1,1,1},{0,1,1}}, {{0,1,1},{0,0,1}},
    {{0,0,0},{0,0,1}}, {{1,0,0},{1,0,1}}, {{1,1,0},{1,1,1}}, {{0,1,0},{0,1,1}}
};

/* m_coef.x() * x + m_coef.y() * y + m_coef.z() * z + m_coef.w(), left to
 * right in float; the integer form converts size_t node indices to float. */
static float plane_at(const float c[4], float x, float y, float z)
{
    float s = c[0] * x;
    s = s + c[1] * y;
    s = s + c[2] * z;
    return s + c[3];
}

/* interpolate_vertex (:981-990): ratio = |f0 / (f1 - f0)|, (1 - r) * v0 + r * v1 in float. */
static void plane_vertex(const float c[4], const float a[3], const float b[3], float out[3], float* ratio)
{
    const float f0 = plane_at(c, a[0], a[1], a[2]);
    const float f1 = plane_at(c, b[0], b[1], b[2]);
    float r = f0 / (f1 - f0);
    r = r < 0 ? -r : r;
    const float s = 1.0f - r;
    for (int k = 0; k < 3; ++k) {
        const float p = s * a[k];
        const float q = r * b[k];
        out[k] = p + q;
    }
    *ratio = r;
}

/* interpolate_value<T> (:1001-1026): values[i0] + ratio * (values[i1] - values[i0])
 * with C++'s usual arithmetic conversions for T (integer promotion of the
 * difference, unsigned wrap-around, float product), then widened to double. */
#define VALUE_BODY(T, DIFF_T, ACC)                                          \
    {                                                                       \
        const T* v = (const T*)values;                                      \
        const DIFF_T d = (DIFF_T)v[i1] - (DIFF_T)v[i0];                     \
        const ACC prod = (ACC)r * (ACC)d;                                   \
        return (double)((ACC)v[i0] + prod);                                 \
    }

static double plane_value(const void* values, int dt, size_t i0, size_t i1, float r)
{
    switch (dt) {
    case SO_I8:  VALUE_BODY(int8_t, int32_t, float)
    case SO_I16: VALUE_BODY(int16_t, int32_t, float)
    case SO_I32: VALUE_BODY(int32_t, int32_t, float)
    case SO_I64: VALUE_BODY(int64_t, int64_t, float)
    case SO_U8:  VALUE_BODY(uint8_t, int32_t, float)
    case SO_U16: VALUE_BODY(uint16_t, int32_t, float)
    case SO_U32: VALUE_BODY(uint32_t, uint32_t, float)
    case SO_U64: VALUE_BODY(uint64_t, uint64_t, float)
    case SO_F32: VALUE_BODY(float, float, float)
    default:     VALUE_BODY(double, double, double)
    }
}

/* Math::Round (Utility/Math.h:110-114): half away from zero, in double. */
static int round_half_away(float a) { return a >= 0 ? (int)(a + 0.5) : -(int)(0.5 - a); }
static float clampf(float a, float lo, float hi) { return a <= lo ? lo : (a <= hi ? a : hi); }

/* ColorMap::at (ColorMap.cpp:1043-1054) with RGBColor::Mix (RGBColor.cpp:23-31):
 * both colours to [0,1] floats, Math::Mix x*(1-t) + y*t, back to bytes. */
static void color_at(const uint8_t* table, uint32_t res, float lo, float hi, float value, uint8_t rgb[3])
{
    const float v0 = clampf(value, lo, hi);
    const float rr = (float)(res - 1);
    const float v = (v0 - lo) / (hi - lo) * rr;
    const size_t s0 = (size_t)v;
    const size_t s1 = (s0 + 1 < (size_t)(res - 1)) ? s0 + 1 : (size_t)(res - 1);
    const float t = v - (float)s0;
    for (int k = 0; k < 3; ++k) {
        const float a = (float)table[3 * s0 + k] / 255.0f;
        const float b = (float)table[3 * s1 + k] / 255.0f;
        const float m = a * (1 - t) + b * t;
        rgb[k] = (uint8_t)round_half_away(clampf(m, 0.0f, 1.0f) * 255.0f);
    }
}

typedef struct {
    float* f; size_t n, cap;
} FVec;

static int fpush(FVec* v, float x)
{
    if (v->n == v->cap) {
        size_t c = v->cap ? 2 * v->cap : 1024;
        float* p = (float*)realloc(v->f, c * sizeof(float));
        if (!p) return 0;
        v->f = p;
        v->cap = c;
    }
    v->f[v->n++] = x;
    return 1;
}

/* extract_plane<T>(structured) (:165-280).  coef: the plane's coefficients;
 * minc / maxc: min / max object coordinates; table / res / lo / hi: the colour
 * map.  Returns 0, or 6 when out of memory. */
int oracle_slice_plane(const void* values, int dtype, const uint32_t dims[3], const float minc[3],
                       const float maxc[3], const float coef[4], const uint8_t* table, uint32_t res, float lo,
                       float hi, SliceMesh* out)
{
    memset(out, 0, sizeof(*out));
    const uint32_t ncx = dims[0] - 1, ncy = dims[1] - 1, ncz = dims[2] - 1;
    const size_t line = dims[0], slice = (size_t)dims[0] * dims[1];
    float sc[3];
    const uint32_t nc[3] = { ncx, ncy, ncz };
    for (int k = 0; k < 3; ++k) sc[k] = (maxc[k] - minc[k]) / (float)nc[k];
    FVec co = { 0, 0, 0 }, no = { 0, 0, 0 };
    uint8_t* colors = NULL;
    size_t ncol = 0, capcol = 0;
    int ok = 1;
    for (uint32_t z = 0; z < ncz && ok; ++z)
        for (uint32_t y = 0; y < ncy && ok; ++y)
            for (uint32_t x = 0; x < ncx && ok; ++x) {
                /* calculate_table_index (:802-818): plane > 0.0f at the eight corners */
                static const int corner[8][3] = { {0,0,0},{1,0,0},{1,1,0},{0,1,0},{0,0,1},{1,0,1},{1,1,1},{0,1,1} };
                int ti = 0;
                for (int k = 0; k < 8; ++k)
                    if (plane_at(coef, (float)(x + corner[k][0]), (float)(y + corner[k][1]), (float)(z + corner[k][2])) > 0.0f)
                        ti |= 1 << k;
                if (ti == 0 || ti == 255) continue;
                int edges[16];
                oracle_case_edges(ti, edges);
                for (int i = 0; edges[i] != -1; i += 3) {
                    const int es[3] = { edges[i], edges[i + 2], edges[i + 1] };
                    float vert[3][3];
                    uint8_t rgb[3][3];
                    for (int j = 0; j < 3; ++j) {
                        const int e = es[j];
                        const float a[3] = { (float)(x + kEdgeEnds[e][0][0]), (float)(y + kEdgeEnds[e][0][1]),
                                             (float)(z + kEdgeEnds[e][0][2]) };
                        const float b[3] = { (float)(x + kEdgeEnds[e][1][0]), (float)(y + kEdgeEnds[e][1][1]),
                                             (float)(z + kEdgeEnds[e][1][2]) };
                        float p[3], r;
                        plane_vertex(coef, a, b, p, &r);
                        /* scale_coord: (coord + min_coord) * scale_factor */
                        for (int k = 0; k < 3; ++k) {
                            const float q = p[k] + minc[k];
                            vert[j][k] = q * sc[k];
                        }
                        const double i0d = (double)a[0] + (double)a[1] * (double)line + (double)a[2] * (double)slice;
                        const double i1d = (double)b[0] + (double)b[1] * (double)line + (double)b[2] * (double)slice;
                        const double val = plane_value(values, dtype, (size_t)i0d, (size_t)i1d, r);
                        color_at(table, res, lo, hi, (float)val, rgb[j]);
                    }
                    for (int j = 0; j < 3 && ok; ++j)
                        for (int k = 0; k < 3 && ok; ++k) ok = fpush(&co, vert[j][k]);
                    if (ncol + 9 > capcol) {
                        capcol = capcol ? 2 * capcol : 4096;
                        uint8_t* p = (uint8_t*)realloc(colors, capcol);
                        if (!p) { ok = 0; break; }
                        colors = p;
                    }
                    for (int j = 0; j < 3; ++j)
                        for (int k = 0; k < 3; ++k) colors[ncol++] = rgb[j][k];
                    /* -(vertex2 - vertex0).cross(vertex1 - vertex0) */
                    float d1[3], d2[3];
                    for (int k = 0; k < 3; ++k) { d1[k] = vert[2][k] - vert[0][k]; d2[k] = vert[1][k] - vert[0][k]; }
                    const float nx = d1[1] * d2[2] - d1[2] * d2[1];
                    const float ny = d1[2] * d2[0] - d1[0] * d2[2];
                    const float nz = d1[0] * d2[1] - d1[1] * d2[0];
                    ok = ok && fpush(&no, -nx) && fpush(&no, -ny) && fpush(&no, -nz);
                }
            }
    if (!ok) {
        free(co.f); free(no.f); free(colors);
        return 6;
    }
    out->coords = co.f; out->n_coords = co.n;
    out->normals = no.f; out->n_normals = no.n;
    out->colors = colors; out->n_colors = ncol;
    out->n_triangles = no.n / 3;
    return 0;
}

void oracle_slice_free(SliceMesh* m)
{
    free(m->coords); free(m->normals); free(m->colors);
    memset(m, 0, sizeof(*m));
}
