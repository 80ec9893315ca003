/*
 * mc_oracle.c -- THE ORACLE (test infrastructure only, never part of the product).
 *
 * A plain-C, single-threaded restatement of kvs::MarchingCubes over a
 * kvs::StructuredVolumeObject, written to reproduce the reference bit for bit
 * on every input the reference itself handles.  Only tests/, the smoke check
 * in __graft_entry__.py and the cpu_baseline leg of bench.py may load it; the
 * HIP product path never links or calls it.
 *
 * Reference files followed (paths relative to Source/Core/):
 *   Visualization/Mapper/MarchingCubes.cpp      (whole file)
 *   Visualization/Mapper/MarchingCubesTable.cpp (triangle / edge tables)
 *   Visualization/Object/StructuredVolumeObject.cpp:66-111 (GetMinMaxValues)
 *   Utility/Math.h:26-66,128-138 (Min, Max, Abs, Clamp, Equal)
 *   Matrix/Vector3.h:479-485 (cross)
 *
 * Differences from the reference, all outside the reference's defined domain:
 *   - edge slots and vertex-map indices are 64-bit, so volumes above
 *     715,827,882 nodes (where MarchingCubes.cpp:508-510 narrows a size_t to
 *     int and reads out of bounds) produce the output the reference would
 *     produce without that narrowing;
 *   - an optional FP64 re-evaluation of vertex positions (no float rounding
 *     of the lerp, double (p+min)*scale) used only to report errors.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fPIC -shared (x86-64 SSE2 math,
 * the same IEEE double/float arithmetic as the reference's g++ -O3 build,
 * Makefile.def:596-606, which has no -march and so no FMA).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { OR_I8, OR_I16, OR_I32, OR_I64, OR_U8, OR_U16, OR_U32, OR_U64, OR_F32, OR_F64 };
enum { OR_VERTEX_NORMAL = 0, OR_POLYGON_NORMAL = 1 };
enum { OR_OK = 0, OR_ERR_DTYPE = 4, OR_ERR_ARG = 5, OR_ERR_OOM = 6 };

/* Triangle table, MarchingCubesTable.cpp:16-274: per cube index, the edge ids
 * of its triangles in table order, one hex digit per edge. */
static const char* const kCaseEdges[256] = {
    "", "083", "019", "183981", "12a", "08312a", "92a029", "2832a8a98",
    "3b2", "0b28b0", "19023b", "1b219b98b", "3a1ba3", "0a108a8ba", "3903b9ba9", "98aa8b",
    "478", "430734", "019847", "419471731", "12a847", "34730412a", "92a902847", "2a9297273794",
    "8473b2", "b47b24204", "90184723b", "47b94b9b2921", "3a13ba784", "1ba14b1047b4", "47890b9bab03", "47b4b99ba",
    "954", "954083", "054150", "854835315", "12a954", "30812a495", "52a542402", "2a5325354348",
    "95423b", "0b208b495", "05401523b", "21525828b485", "a3ba13954", "4950818a18ba", "54050b5bab03", "54858aa8b",
    "978579", "930953573", "078017157", "153357", "978957a12", "a12950530573", "802825857a52", "2a5253357",
    "7957893b2", "95797292027b", "23b018178157", "b21b17715", "958857a13a3b", "5705097b010aba0", "ba0b03a50807570", "ba57b5",
    "a65", "0835a6", "9015a6", "1831985a6", "165261", "165126308", "965906026", "598582526328",
    "23ba65", "b08b20a65", "01923b5a6", "5a61929b298b", "63b653513", "08b0b50515b6", "3b6036065059", "65969bb98",
    "5a6478", "43047365a", "1905a6847", "a65197173794", "612651478", "125526304347", "847905065026", "739794329596269",
    "3b2784a65", "5a647242027b", "01947823b5a6", "9219b294b7b45a6", "8473b53515b6", "51b5b610b7b404b", "059065036b63847", "65969b4797b9",
    "a4964a", "4a649a083", "a01a60640", "83181686461a", "149124264", "308129249264", "024426", "832824426",
    "a49a64b23", "08228b49a4a6", "3b201606461a", "64161a48121b8b1", "964936913b63", "8b1810b61914641", "3b6360064", "648b68",
    "7a678a89a", "0730a709a67a", "a671a7178180", "a67a71173", "126168189867", "269291679093739", "780706602", "732672",
    "23ba68a89867", "20727b09767a9a7", "1801781a767a23b", "b21b17a61671", "896867916b63136", "091b67", "7807063b0b60", "7b6",
    "76b", "308b76", "019b76", "819831b76", "a126b7", "12a3086b7", "2902a96b7", "6b72a3a83a98",
    "723627", "708760620", "276237019", "162186198876", "a76a17137", "a7617a187108", "03707a0a96a7", "76a7a88a9",
    "684b86", "36b306046", "86b846901", "946963931b36", "6846b82a1", "12a30b06b046", "4b846b0292a9", "a93a32943b36463",
    "823842462", "042462", "190234246438", "194142246", "8138618466a1", "a10a06604", "4634386a3039a93", "a946a4",
    "49576b", "083495b76", "50154076b", "b76834354315", "954a1276b", "6b712a083495", "76b54a42a402", "348354325a52b76",
    "723762549", "954086062687", "362376150540", "628687218485158", "954a16176137", "16a176107870954", "40a4a503a6a737a", "76a7a854a48a",
    "6956b9b89", "36b063056095", "0b805b01556b", "6b3635531", "12a95b9b8b56", "0b306b09656912a", "b85b56805a52025", "6b36352a3a53",
    "589528562382", "956960062", "158180568382628", "156216", "13616a386569896", "a10a06950560", "03856a", "a56",
    "b5a75b", "b5ab75830", "5b75ab190", "a75ab7981831", "b12b71751", "08312717572b", "9759279022b7", "75272b592328982",
    "25a235375", "820852875a25", "9015a35373a2", "982921872a25752", "135375", "087071175", "903935537", "987597",
    "5845a8ab8", "5045b05abb30", "01984a8aba45", "ab4a45b34941314", "2512852b8458", "04b0b345b2b151b", "0250592b5458b85", "9452b3",
    "25a352345384", "5a2524420", "3a235a385458019", "5a2524192942", "845853351", "045105", "845853905035", "945",
    "4b749b9ab", "0834979b79ab", "1ab1b414074b", "3143481a474bab4", "4b79b492b912", "9749b791b2b1083", "b74b42240", "b74b42834324",
    "29a279237749", "9a7974a27870207", "37a3a274a1a040a", "1a2874", "491417713", "491417081871", "403743", "487",
    "9a8ab8", "30939bb9a", "01a0a88ab", "31ab3a", "12b1b99b8", "30939b1292b9", "02b80b", "32b",
    "23828aa89", "9a2092", "23828a0181a8", "1a2", "138918", "091", "038", "",
};

/* Cube corner offsets (MarchingCubes.cpp:187-194) and the two corners of each
 * edge in table orientation (MarchingCubesTable.cpp:276-290). */
static const int kCorner[8][3] = {
    {0,0,0}, {1,0,0}, {1,1,0}, {0,1,0}, {0,0,1}, {1,0,1}, {1,1,1}, {0,1,1}
};
static const int kEdgeCorners[12][2] = {
    {0,1}, {1,2}, {2,3}, {3,0}, {4,5}, {5,6}, {6,7}, {7,4}, {0,4}, {1,5}, {2,6}, {3,7}
};

static int hexval(char c) { return (c >= '0' && c <= '9') ? c - '0' : c - 'a' + 10; }

static double load_value(const void* p, int dt, size_t i)
{
    switch (dt) {
    case OR_I8:  return (double)((const int8_t*)p)[i];
    case OR_I16: return (double)((const int16_t*)p)[i];
    case OR_I32: return (double)((const int32_t*)p)[i];
    case OR_I64: return (double)((const int64_t*)p)[i];
    case OR_U8:  return (double)((const uint8_t*)p)[i];
    case OR_U16: return (double)((const uint16_t*)p)[i];
    case OR_U32: return (double)((const uint32_t*)p)[i];
    case OR_U64: return (double)((const uint64_t*)p)[i];
    case OR_F32: return (double)((const float*)p)[i];
    default:     return ((const double*)p)[i];
    }
}

/* GetMinMaxValues<T>, StructuredVolumeObject.cpp:66-88: a serial scan in the
 * value type T with Math::Min(a,b) = a<b?a:b and Math::Max(a,b) = a>b?a:b,
 * the running extreme as the second argument (so a NaN first value sticks). */
#define MINMAX_BODY(T)                                                     \
    {                                                                      \
        const T* v = (const T*)values;                                     \
        T mn = v[0], mx = v[0];                                            \
        for (uint64_t i = 0; i < n; ++i) {                                 \
            mn = (v[i] < mn) ? v[i] : mn;                                  \
            mx = (v[i] > mx) ? v[i] : mx;                                  \
        }                                                                  \
        *out_min = (double)mn; *out_max = (double)mx;                      \
        return OR_OK;                                                      \
    }

int oracle_min_max(const void* values, int dtype, uint64_t n, double* out_min, double* out_max)
{
    if (!values || n == 0) return OR_ERR_ARG;
    switch (dtype) {
    case OR_I8:  MINMAX_BODY(int8_t)
    case OR_I16: MINMAX_BODY(int16_t)
    case OR_I32: MINMAX_BODY(int32_t)
    case OR_I64: MINMAX_BODY(int64_t)
    case OR_U8:  MINMAX_BODY(uint8_t)
    case OR_U16: MINMAX_BODY(uint16_t)
    case OR_U32: MINMAX_BODY(uint32_t)
    case OR_U64: MINMAX_BODY(uint64_t)
    case OR_F32: MINMAX_BODY(float)
    case OR_F64: MINMAX_BODY(double)
    default: return OR_ERR_DTYPE;
    }
}

typedef struct {
    float*    coords;        uint64_t n_coords;       /* floats (3 per vertex) */
    uint32_t* connections;   uint64_t n_connections;  /* uint32 (3 per triangle, 0 if duplicated) */
    float*    normals;       uint64_t n_normals;      /* floats (3 per vertex or per triangle) */
    double*   coords_f64;    /* optional FP64 re-evaluation, same layout as coords */
    uint64_t  n_triangles;
    int32_t   polygon_type;  /* 3 = Triangle */
    int32_t   normal_type;   /* 0 = VertexNormal, 1 = PolygonNormal */
    int32_t   color_type;    /* 1 = PolygonColor */
    int32_t   empty_range;   /* 1 if Math::Equal(min,max) short-circuited */
} oracle_mesh;

/* Growable arrays standing in for the reference's std::vector outputs. */
typedef struct { float* p; uint64_t n, cap; } fvec;
typedef struct { double* p; uint64_t n, cap; } dvec;
typedef struct { uint32_t* p; uint64_t n, cap; } uvec;

static int fpush(fvec* v, float x)
{
    if (v->n == v->cap) {
        uint64_t c = v->cap ? v->cap * 2 : 1024;
        float* q = (float*)realloc(v->p, c * sizeof(float));
        if (!q) return 0;
        v->p = q; v->cap = c;
    }
    v->p[v->n++] = x;
    return 1;
}
static int dpush(dvec* v, double x)
{
    if (v->n == v->cap) {
        uint64_t c = v->cap ? v->cap * 2 : 1024;
        double* q = (double*)realloc(v->p, c * sizeof(double));
        if (!q) return 0;
        v->p = q; v->cap = c;
    }
    v->p[v->n++] = x;
    return 1;
}
static int upush(uvec* v, uint32_t x)
{
    if (v->n == v->cap) {
        uint64_t c = v->cap ? v->cap * 2 : 1024;
        uint32_t* q = (uint32_t*)realloc(v->p, c * sizeof(uint32_t));
        if (!q) return 0;
        v->p = q; v->cap = c;
    }
    v->p[v->n++] = x;
    return 1;
}

typedef struct {
    const void* values; int dtype;
    uint64_t nx, ny, nz, line, slice;
    double minv, maxv, iso;
    float minc[3], sc[3];
} ctx_t;

/* Math::Clamp, Math.h:128-132. */
static double clampd(double a, double lo, double hi) { return a <= lo ? lo : (a <= hi ? a : hi); }

/* interpolate_vertex<T> (MarchingCubes.cpp:326-356) followed by the caller's
 * (p + min_coord) * scale_factor (:169-173 / :382-386).  a and b are integer
 * node coordinates; the reference holds them as floats, exact below 2^24. */
static void interp(const ctx_t* c, const uint64_t a[3], const uint64_t b[3], float out[3], double out64[3])
{
    const uint64_t ia = a[0] + a[1] * c->line + a[2] * c->slice;
    const uint64_t ib = b[0] + b[1] * c->line + b[2] * c->slice;
    double va = clampd(load_value(c->values, c->dtype, ia), c->minv, c->maxv);
    double vb = clampd(load_value(c->values, c->dtype, ib), c->minv, c->maxv);
    double r = (c->iso - va) / (vb - va);
    r = (r < 0) ? -r : r;                                   /* Math::Abs */
    for (int k = 0; k < 3; ++k) {
        const double pa = (double)(float)a[k], pb = (double)(float)b[k];
        const double t0 = (1.0 - r) * pa;
        const double t1 = r * pb;
        const float p = (float)(t0 + t1);
        const float q = p + c->minc[k];
        out[k] = q * c->sc[k];
        if (out64) out64[k] = ((t0 + t1) + (double)c->minc[k]) * (double)c->sc[k];
    }
}

/* Vector3::cross, Vector3.h:479-485, on d1 = v1 - v0 and d2 = v2 - v0. */
static void face_normal(const float* v0, const float* v1, const float* v2, float n[3])
{
    const float d1x = v1[0] - v0[0], d1y = v1[1] - v0[1], d1z = v1[2] - v0[2];
    const float d2x = v2[0] - v0[0], d2y = v2[1] - v0[1], d2z = v2[2] - v0[2];
    const float a = d1y * d2z, b = d1z * d2y;
    const float c = d1z * d2x, d = d1x * d2z;
    const float e = d1x * d2y, f = d1y * d2x;
    n[0] = a - b; n[1] = c - d; n[2] = e - f;
}

/* calculate_table_index<T>, MarchingCubes.cpp:299-316: bit k set iff
 * double(V[corner k]) > isolevel. */
static int cube_index(const ctx_t* c, uint64_t x, uint64_t y, uint64_t z)
{
    int ci = 0;
    for (int k = 0; k < 8; ++k) {
        const uint64_t i = (x + kCorner[k][0]) + (y + kCorner[k][1]) * c->line + (z + kCorner[k][2]) * c->slice;
        if (load_value(c->values, c->dtype, i) > c->iso) ci |= 1 << k;
    }
    return ci;
}

static int extract_dup(const ctx_t* c, oracle_mesh* out, int want64)
{
    /* extract_surfaces_with_duplication<T>, MarchingCubes.cpp:155-243. */
    fvec coords = {0}, normals = {0}; dvec c64 = {0};
    uint64_t ntri = 0;
    const uint64_t ncx = c->nx - 1, ncy = c->ny - 1, ncz = c->nz - 1;
    for (uint64_t z = 0; z < ncz; ++z)
    for (uint64_t y = 0; y < ncy; ++y)
    for (uint64_t x = 0; x < ncx; ++x) {
        const int ci = cube_index(c, x, y, z);
        if (ci == 0 || ci == 255) continue;
        const char* s = kCaseEdges[ci];
        for (int i = 0; s[i]; i += 3) {
            const int e[3] = { hexval(s[i]), hexval(s[i + 2]), hexval(s[i + 1]) };
            float w[3][3]; double w64[3][3];
            for (int j = 0; j < 3; ++j) {
                const int* ka = kCorner[kEdgeCorners[e[j]][0]];
                const int* kb = kCorner[kEdgeCorners[e[j]][1]];
                const uint64_t a[3] = { x + ka[0], y + ka[1], z + ka[2] };
                const uint64_t b[3] = { x + kb[0], y + kb[1], z + kb[2] };
                interp(c, a, b, w[j], want64 ? w64[j] : NULL);
            }
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    if (!fpush(&coords, w[j][k])) return OR_ERR_OOM;
                    if (want64 && !dpush(&c64, w64[j][k])) return OR_ERR_OOM;
                }
            float n[3];
            face_normal(w[0], w[1], w[2], n);
            for (int k = 0; k < 3; ++k) if (!fpush(&normals, n[k])) return OR_ERR_OOM;
            ++ntri;
        }
    }
    if (coords.n > 0) {   /* :236-242 */
        out->coords = coords.p; out->n_coords = coords.n;
        out->normals = normals.p; out->n_normals = normals.n;
        out->coords_f64 = c64.p;
    } else {
        free(coords.p); free(normals.p); free(c64.p);
    }
    out->n_triangles = ntri;
    return OR_OK;
}

static int extract_nodup(const ctx_t* c, int normal_type, oracle_mesh* out, int want64)
{
    /* extract_surfaces_without_duplication<T>, MarchingCubes.cpp:251-290. */
    const uint64_t nx = c->nx, ny = c->ny, nz = c->nz;
    const uint64_t ncx = nx - 1, ncy = ny - 1, ncz = nz - 1;
    const uint64_t nnodes = nx * ny * nz;
    uint32_t* vmap = (uint32_t*)calloc(3 * nnodes, sizeof(uint32_t));
    if (!vmap) return OR_ERR_OOM;
    fvec coords = {0}, normals = {0}; dvec c64 = {0}; uvec conn = {0};

    /* calculate_isopoints<T>, :365-446: nodes in z,y,x order; x-, y-, z-edge. */
    uint32_t nv = 0;
    uint64_t index = 0;
    for (uint64_t z = 0; z < nz; ++z)
    for (uint64_t y = 0; y < ny; ++y)
    for (uint64_t x = 0; x < nx; ++x, ++index) {
        const int b0 = load_value(c->values, c->dtype, index) > c->iso;
        const uint64_t a[3] = { x, y, z };
        for (int axis = 0; axis < 3; ++axis) {
            const uint64_t lim = axis == 0 ? ncx : axis == 1 ? ncy : ncz;
            const uint64_t pos = axis == 0 ? x : axis == 1 ? y : z;
            if (pos == lim) continue;
            const uint64_t step = axis == 0 ? 1 : axis == 1 ? c->line : c->slice;
            const int b1 = load_value(c->values, c->dtype, index + step) > c->iso;
            if (b0 == b1) continue;
            uint64_t b[3] = { x, y, z };
            b[axis] += 1;
            float w[3]; double w64[3];
            interp(c, a, b, w, want64 ? w64 : NULL);
            for (int k = 0; k < 3; ++k) {
                if (!fpush(&coords, w[k])) goto oom;
                if (want64 && !dpush(&c64, w64[k])) goto oom;
            }
            vmap[3 * index + axis] = nv++;
        }
    }

    /* connect_isopoints<T>, :455-521: edge slots of the cell (64-bit here). */
    for (uint64_t z = 0; z < ncz; ++z)
    for (uint64_t y = 0; y < ncy; ++y)
    for (uint64_t x = 0; x < ncx; ++x) {
        const int ci = cube_index(c, x, y, z);
        if (ci == 0 || ci == 255) continue;
        const uint64_t n0 = x + y * c->line + z * c->slice;
        uint64_t le[12];
        le[0] = 3 * n0;
        le[1] = le[0] + 3 + 1;
        le[2] = le[0] + 3 * c->line;
        le[3] = le[0] + 1;
        le[4] = le[0] + 3 * c->slice;
        le[5] = le[1] + 3 * c->slice;
        le[6] = le[2] + 3 * c->slice;
        le[7] = le[3] + 3 * c->slice;
        le[8] = le[0] + 2;
        le[9] = le[8] + 3;
        le[10] = le[8] + 3 + 3 * c->line;
        le[11] = le[8] + 3 * c->line;
        const char* s = kCaseEdges[ci];
        for (int i = 0; s[i]; i += 3) {
            if (!upush(&conn, vmap[le[hexval(s[i])]])) goto oom;
            if (!upush(&conn, vmap[le[hexval(s[i + 2])]])) goto oom;
            if (!upush(&conn, vmap[le[hexval(s[i + 1])]])) goto oom;
        }
    }
    free(vmap); vmap = NULL;

    if (coords.n > 0) {
        if (normal_type == OR_POLYGON_NORMAL) {
            /* calculate_normals_on_polygon, :531-559. */
            for (uint64_t t = 0; t < conn.n; t += 3) {
                float n[3];
                face_normal(coords.p + 3 * (uint64_t)conn.p[t], coords.p + 3 * (uint64_t)conn.p[t + 1],
                            coords.p + 3 * (uint64_t)conn.p[t + 2], n);
                for (int k = 0; k < 3; ++k) if (!fpush(&normals, n[k])) goto oom;
            }
        } else {
            /* calculate_normals_on_vertex, :569-605: serial float sums in
             * triangle order, starting from 0.0f. */
            normals.p = (float*)calloc(coords.n ? coords.n : 1, sizeof(float));
            if (!normals.p) goto oom;
            normals.n = normals.cap = coords.n;
            for (uint64_t t = 0; t < conn.n; t += 3) {
                const float* v0 = coords.p + 3 * (uint64_t)conn.p[t];
                const float* v1 = coords.p + 3 * (uint64_t)conn.p[t + 1];
                const float* v2 = coords.p + 3 * (uint64_t)conn.p[t + 2];
                float n[3];
                face_normal(v0, v1, v2, n);
                for (int j = 0; j < 3; ++j) {
                    float* dst = normals.p + 3 * (uint64_t)conn.p[t + j];
                    dst[0] += n[0]; dst[1] += n[1]; dst[2] += n[2];
                }
            }
        }
        out->coords = coords.p; out->n_coords = coords.n;
        out->connections = conn.p; out->n_connections = conn.n;
        out->normals = normals.p; out->n_normals = normals.n;
        out->coords_f64 = c64.p;
    } else {
        free(coords.p); free(conn.p); free(normals.p); free(c64.p);
    }
    out->n_triangles = conn.n / 3;
    return OR_OK;
oom:
    free(vmap); free(coords.p); free(conn.p); free(normals.p); free(c64.p);
    return OR_ERR_OOM;
}

/*
 * MarchingCubes::exec + mapping (MarchingCubes.cpp:46-132) for one isolevel.
 * min_value/max_value are the volume's cached min/max (clamp range of
 * interpolate_vertex and the Math::Equal short-circuit at :108-110);
 * min_coord/max_coord its min/max object coordinates.
 */
int oracle_extract(const void* values, int dtype, const uint32_t dims[3],
                   double min_value, double max_value,
                   const float min_coord[3], const float max_coord[3],
                   double isolevel, int normal_type, int duplication, int want_f64,
                   oracle_mesh* out)
{
    if (!out || !dims) return OR_ERR_ARG;
    memset(out, 0, sizeof(*out));
    if (dtype < OR_I8 || dtype > OR_F64) return OR_ERR_DTYPE;
    if (normal_type == OR_VERTEX_NORMAL) duplication = 0;              /* :69-73 */
    out->polygon_type = 3;
    out->normal_type = duplication ? OR_POLYGON_NORMAL : normal_type;   /* :94-104 */
    out->color_type = 1;
    {
        const double d = min_value - max_value;
        if ((d < 0 ? -d : d) < 1.e-6) { out->empty_range = 1; return OR_OK; }   /* :108-110 */
    }
    if (!values || dims[0] < 1 || dims[1] < 1 || dims[2] < 1) return OR_ERR_ARG;
    ctx_t c;
    c.values = values; c.dtype = dtype;
    c.nx = dims[0]; c.ny = dims[1]; c.nz = dims[2];
    c.line = c.nx; c.slice = c.nx * c.ny;
    c.minv = min_value; c.maxv = max_value; c.iso = isolevel;
    for (int k = 0; k < 3; ++k) {
        /* scale_factor = (max - min) / Vec3(ncells), :166-168 (float ops). */
        const float span = max_coord[k] - min_coord[k];
        c.minc[k] = min_coord[k];
        c.sc[k] = span / (float)(dims[k] - 1);
    }
    return duplication ? extract_dup(&c, out, want_f64) : extract_nodup(&c, normal_type, out, want_f64);
}

void oracle_free(oracle_mesh* m)
{
    if (!m) return;
    free(m->coords); free(m->connections); free(m->normals); free(m->coords_f64);
    memset(m, 0, sizeof(*m));
}

/* Table access for the table-pinning tests: number of triangles of a case and
 * its edge list in table order (-1 terminated, 16 entries like the reference). */
int oracle_case_edges(int cube_index, int edges_out[16])
{
    if (cube_index < 0 || cube_index > 255) return -1;
    const char* s = kCaseEdges[cube_index];
    int n = 0;
    for (; s[n]; ++n) edges_out[n] = hexval(s[n]);
    for (int i = n; i < 16; ++i) edges_out[i] = -1;
    return n / 3;
}
