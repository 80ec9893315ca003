/*
 * kmc_unstructured.hip -- the unstructured branch of kvs::Isosurface
 * (Isosurface.cpp:192-257: MarchingTetrahedra, MarchingHexahedra,
 * MarchingPyramid, MarchingPrism) and kvs::SlicePlane over an
 * UnstructuredVolumeObject (SlicePlane.cpp:288-800) on the device.
 *
 * Output order is the reference's: cells in connection order, triangles of a
 * cell in table order.  One thread per cell classifies its nodes (values
 * against the isolevel, or node coordinates against the plane), keeps the
 * case index and counts its triangles; a scan of the counts gives every
 * cell's first triangle; one thread per cell with triangles then writes them
 * at their ranks.  Only the duplicated extraction exists in the reference (its
 * shared-vertex branch is compiled out, MarchingTetrahedra.cpp:240-286), so
 * without duplication (and with VertexNormal, which forces it off) the
 * surface is empty, as there.
 */
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <vector>

#include "cell_tables.h"
#include "kmc.h"
#include "kmc_guard.h"
#include "kmc_internal.h"
#include "kmc_job.h"

#pragma clang fp contract(off)

namespace kmc {

namespace {

enum : int { kTet = 0, kHex = 1, kPyr = 2, kPrism = 3 };

struct UArgs {
    const void* values;
    const float* coords;          // 3 per node
    const uint32_t* conn;         // nodes per cell, per cell
    uint64_t ncells;
    double iso, minv, maxv;       // isosurface: level and clamp range
    float plane[4];               // SlicePlane: coefficients
    const uint8_t* ctable;        // SlicePlane: colour map
    uint32_t cres;
    float crange[2];
    uint32_t* counts;             // triangles per cell
    uint8_t* tidx;                // case index per cell
    const uint64_t* prefix;       // first triangle per cell
    float* out_coords;            // 9 per triangle
    float* out_normals;           // 3 per triangle
    uint8_t* out_colors;          // 9 per triangle (SlicePlane)
};

template <int C> struct Cell;
template <> struct Cell<kTet> {
    static constexpr int nodes = 4, all_ones = 15;
    static constexpr bool swap12 = false;   // MarchingTetrahedra.cpp:191-193
    static __device__ __forceinline__ uint64_t tri(uint32_t i) { return kTetTri[i]; }
    static __device__ __forceinline__ uint32_t edge(uint32_t e) { return kTetEdge[e]; }
};
template <> struct Cell<kHex> {
    static constexpr int nodes = 8, all_ones = 255;
    static constexpr bool swap12 = true;    // MarchingHexahedra.cpp:180-182
    static __device__ __forceinline__ uint64_t tri(uint32_t i) { return kHexTri[i]; }
    static __device__ __forceinline__ uint32_t edge(uint32_t e) { return kHexEdge[e]; }
};
template <> struct Cell<kPyr> {
    static constexpr int nodes = 5, all_ones = 31;
    static constexpr bool swap12 = true;    // MarchingPyramid.cpp:183-185
    static __device__ __forceinline__ uint64_t tri(uint32_t i) { return kPyrTri[i]; }
    static __device__ __forceinline__ uint32_t edge(uint32_t e) { return kPyrEdge[e]; }
};
template <> struct Cell<kPrism> {
    static constexpr int nodes = 6, all_ones = 63;
    static constexpr bool swap12 = false;   // MarchingPrism.cpp:163-165
    static __device__ __forceinline__ uint64_t tri(uint32_t i) { return kPrismTri[i]; }
    static __device__ __forceinline__ uint32_t edge(uint32_t e) { return kPrismEdge[e]; }
};

/* The cell's nodes in the mapper's local order: a hexahedron's connections
 * 4..7 are its local nodes 0..3 and 0..3 are 4..7 (MarchingHexahedra.cpp:163-170,
 * SlicePlane.cpp:462-469). */
template <int C>
__device__ __forceinline__ void cell_nodes(const UArgs& a, uint64_t cell, uint64_t li[8])
{
    constexpr int K = Cell<C>::nodes;
    const uint32_t* c = a.conn + cell * K;
#pragma unroll
    for (int i = 0; i < K; ++i) li[i] = c[C == kHex ? ((i + 4) & 7) : i];
}

/* SlicePlane::substitute_plane_equation on a node's coordinates (:955-971). */
__device__ __forceinline__ float uplane(const float c[4], const float* v)
{
    float s = c[0] * v[0];
    s = s + c[1] * v[1];
    s = s + c[2] * v[2];
    return s + c[3];
}

/* Case index of a cell: bit k set iff double(value of local node k) > iso
 * (MarchingTetrahedra.cpp:297-310 and the same in the others), with the
 * pyramid's ambiguous cases split by the sign of E (MarchingPyramid.cpp:289-317);
 * for SlicePlane, plane(node) > 0. */
template <typename T, int C, bool SLICE>
__device__ __forceinline__ uint32_t cell_case(const UArgs& a, const uint64_t li[8])
{
    constexpr int K = Cell<C>::nodes;
    const T* vol = (const T*)a.values;
    uint32_t ti = 0;
    if constexpr (SLICE) {
#pragma unroll
        for (int k = 0; k < K; ++k) ti |= (uplane(a.plane, a.coords + 3 * li[k]) > 0.0f ? 1u : 0u) << k;
    } else {
        double v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = (double)vol[li[k]];
#pragma unroll
        for (int k = 0; k < K; ++k) ti |= (v[k] > a.iso ? 1u : 0u) << k;
        if constexpr (C == kPyr) {
            if (ti == 10 || ti == 11 || ti == 20 || ti == 21) {
                const double A = (v[1] + v[2] + v[3] + v[4]) / 4;
                const double B = (-v[1] + v[2] + v[3] - v[4]) / 4;
                const double Cc = (-v[1] - v[2] + v[3] + v[4]) / 4;
                const double D = (v[1] - v[2] + v[3] - v[4]) / 4;
                const double p = (a.iso - A) * D;
                const double q = B * Cc;
                const double E = p + q;
                if (ti == 10 && E < 0) ti = 32;
                else if (ti == 11 && E >= 0) ti = 33;
                else if (ti == 20 && E >= 0) ti = 34;
                else if (ti == 21 && E < 0) ti = 35;
            }
        }
    }
    return ti;
}

template <typename T, int C, bool SLICE>
__global__ __launch_bounds__(kBlock) void k_ucount(UArgs a)
{
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t cell = (uint64_t)blockIdx.x * kBlock + threadIdx.x; cell < a.ncells; cell += step) {
        uint64_t li[8];
        cell_nodes<C>(a, cell, li);
        const uint32_t ti = cell_case<T, C, SLICE>(a, li);
        const bool skip = ti == 0 || ti == (uint32_t)Cell<C>::all_ones;
        a.tidx[cell] = (uint8_t)ti;
        a.counts[cell] = skip ? 0u : (uint32_t)(Cell<C>::tri(ti) >> 60);
    }
}

/* Math::Round (Utility/Math.h:110-114). */
__device__ __forceinline__ int uround(float x) { return x >= 0 ? (int)(x + 0.5) : -(int)(0.5 - x); }
__device__ __forceinline__ float uclamp(float x, float lo, float hi) { return x <= lo ? lo : (x <= hi ? x : hi); }

/* ColorMap::at with RGBColor::Mix (ColorMap.cpp:1043-1054, RGBColor.cpp:23-31). */
__device__ __forceinline__ void ucolor(const UArgs& a, float value, uint8_t rgb[3])
{
    const float lo = a.crange[0], hi = a.crange[1];
    const float v0 = uclamp(value, lo, hi);
    const float rr = (float)(a.cres - 1);
    float v = (v0 - lo) / (hi - lo) * rr;
    if (!(v >= 0.0f)) v = 0.0f;   // an empty range: the first colour
    const uint32_t s0 = (uint32_t)v;
    const uint32_t s1 = (s0 + 1 < a.cres - 1) ? s0 + 1 : a.cres - 1;
    const float t = v - (float)s0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float c0 = (float)a.ctable[3 * s0 + k] / 255.0f;
        const float c1 = (float)a.ctable[3 * s1 + k] / 255.0f;
        const float p = c0 * (1 - t);
        const float q = c1 * t;
        rgb[k] = (uint8_t)uround(uclamp(p + q, 0.0f, 1.0f) * 255.0f);
    }
}

/* interpolate_value<T>(unstructured) (SlicePlane.cpp:1036-1050): C++'s
 * promotions of values[i0] + r * (values[i1] - values[i0]). */
template <typename T>
__device__ __forceinline__ double uvalue_at(T va, T vb, float r)
{
    return va + r * (vb - va);
}

template <typename T, int C, bool SLICE>
__global__ __launch_bounds__(kBlock) void k_uemit(UArgs a)
{
    const T* vol = (const T*)a.values;
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t cell = (uint64_t)blockIdx.x * kBlock + threadIdx.x; cell < a.ncells; cell += step) {
        if (a.counts[cell] == 0) continue;
        const uint32_t ti = a.tidx[cell];
        const uint64_t w = Cell<C>::tri(ti);
        const uint32_t n = (uint32_t)(w >> 60);
        uint64_t li[8];
        cell_nodes<C>(a, cell, li);
        const uint64_t base = a.prefix[cell];
        for (uint32_t t = 0; t < n; ++t) {
            /* edges of triangle t: (T[i], T[i+1], T[i+2]), or (T[i], T[i+2], T[i+1]) */
            constexpr bool sw = SLICE ? false : Cell<C>::swap12;
            const uint32_t e[3] = { (uint32_t)((w >> (12 * t)) & 0xF),
                                    (uint32_t)((w >> (12 * t + (sw ? 8 : 4))) & 0xF),
                                    (uint32_t)((w >> (12 * t + (sw ? 4 : 8))) & 0xF) };
            float p[3][3];
            uint8_t rgb[3][3];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const uint32_t ends = Cell<C>::edge(e[j]);
                const uint64_t na = li[ends & 0xFu], nb = li[ends >> 4];
                const float* ca = a.coords + 3 * na;
                const float* cb = a.coords + 3 * nb;
                if constexpr (SLICE) {
                    /* interpolate_vertex (:981-990), float */
                    const float f0 = uplane(a.plane, ca), f1 = uplane(a.plane, cb);
                    float r = f0 / (f1 - f0);
                    r = r < 0 ? -r : r;
                    const float s = 1.0f - r;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float u = s * ca[k];
                        const float v = r * cb[k];
                        p[j][k] = u + v;
                    }
                    ucolor(a, (float)uvalue_at<T>(vol[na], vol[nb], r), rgb[j]);
                } else {
                    /* interpolate_vertex<T> (MarchingTetrahedra.cpp:320-340): FP64 clamp and
                     * ratio, coords[a] + ratio * (coords[b] - coords[a]) in double, then float */
                    double va = (double)vol[na], vb = (double)vol[nb];
                    va = va <= a.minv ? a.minv : (va <= a.maxv ? va : a.maxv);
                    vb = vb <= a.minv ? a.minv : (vb <= a.maxv ? vb : a.maxv);
                    double r = (a.iso - va) / (vb - va);
                    r = r < 0 ? -r : r;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float diff = cb[k] - ca[k];
                        const double dt = r * (double)diff;
                        p[j][k] = (float)((double)ca[k] + dt);
                    }
                }
            }
            /* isosurface: (v1 - v0) x (v2 - v0); SlicePlane: -((v2 - v0) x (v1 - v0)) */
            const int i1 = SLICE ? 2 : 1, i2 = SLICE ? 1 : 2;
            const float ax = p[i1][0] - p[0][0], ay = p[i1][1] - p[0][1], az = p[i1][2] - p[0][2];
            const float bx = p[i2][0] - p[0][0], by = p[i2][1] - p[0][1], bz = p[i2][2] - p[0][2];
            float nx = ay * bz - az * by;
            float ny = az * bx - ax * bz;
            float nz = ax * by - ay * bx;
            if constexpr (SLICE) { nx = -nx; ny = -ny; nz = -nz; }
            const uint64_t id = base + t;
            float* co = a.out_coords + 9 * id;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                co[3 * j + 0] = p[j][0];
                co[3 * j + 1] = p[j][1];
                co[3 * j + 2] = p[j][2];
            }
            a.out_normals[3 * id + 0] = nx;
            a.out_normals[3 * id + 1] = ny;
            a.out_normals[3 * id + 2] = nz;
            if constexpr (SLICE) {
                uint8_t* cl = a.out_colors + 9 * id;
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    cl[3 * j + 0] = rgb[j][0];
                    cl[3 * j + 1] = rgb[j][1];
                    cl[3 * j + 2] = rgb[j][2];
                }
            }
        }
    }
}

/* updateMinMaxCoords with connections (UnstructuredVolumeObject.cpp:425-455):
 * a serial scan running = Math::Min(running, x) over the coordinates of every
 * connection entry.  A NaN x makes running NaN and the next value restarts
 * it, so a range of the scan acts on an incoming running value r as
 * F(r) = nan ? t : (isnan(r) ? t : min(r, t)): t is the scan of the range
 * after its last NaN (ties take the later value), nan whether the range holds
 * a NaN.  Ranges compose in order, and the scan's result is t of the whole. */
struct CoordSpan {
    float mn[3], mx[3];
    uint32_t nan;     // bit k: the span holds a NaN on axis k
    uint32_t valid;
};

__device__ __forceinline__ void span_merge(CoordSpan& f, const CoordSpan& g)
{
    if (!g.valid) return;
    if (!f.valid) { f = g; return; }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if ((g.nan >> k) & 1u) {
            f.mn[k] = g.mn[k];
            f.mx[k] = g.mx[k];
            f.nan |= 1u << k;
        } else {
            f.mn[k] = f.mn[k] != f.mn[k] ? g.mn[k] : (f.mn[k] < g.mn[k] ? f.mn[k] : g.mn[k]);
            f.mx[k] = f.mx[k] != f.mx[k] ? g.mx[k] : (f.mx[k] > g.mx[k] ? f.mx[k] : g.mx[k]);
        }
    }
}

constexpr uint32_t kSpanParts = 4096;

/* Merge of the spans sh[0 .. kBlock) in order into sh[0]: a tree whose step
 * s merges sh[i] with sh[i + s] for i a multiple of 2s (span_merge is the
 * composition of the ranges' scans, so any bracketing of the ordered
 * sequence gives the same result). */
__device__ __forceinline__ void block_span_merge(CoordSpan* sh)
{
    for (uint32_t s = 1; s < kBlock; s <<= 1) {
        __syncthreads();
        if ((threadIdx.x & (2 * s - 1)) == 0) span_merge(sh[threadIdx.x], sh[threadIdx.x + s]);
    }
    __syncthreads();
}

/* Part b of kSpanParts: the span of connection entries [b * per, (b+1) * per),
 * each thread a contiguous run of them (its loads unrolled so that several
 * coordinate gathers are in flight), the runs merged in order. */
__global__ __launch_bounds__(kBlock) void k_coord_span_partial(const float* coords, const uint32_t* conn, uint64_t nconn,
                                                             CoordSpan* parts)
{
    __shared__ CoordSpan sh[kBlock];
    const uint64_t nth = (uint64_t)gridDim.x * kBlock;
    const uint64_t tid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t per = (nconn + nth - 1) / nth;
    const uint64_t b = tid * per < nconn ? tid * per : nconn, e = b + per < nconn ? b + per : nconn;
    CoordSpan s;
    s.valid = 0;
    s.nan = 0;
    constexpr uint32_t kU = 8;
    for (uint64_t i0 = b; i0 < e; i0 += kU) {
        float c[kU][3];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            const uint64_t i = i0 + u < e ? i0 + u : e - 1;
            const float* p = coords + 3 * (uint64_t)conn[i];
            c[u][0] = p[0];
            c[u][1] = p[1];
            c[u][2] = p[2];
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            if (i0 + u >= e) break;
            CoordSpan x;
            x.valid = 1;
            x.nan = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                x.mn[k] = x.mx[k] = c[u][k];
                x.nan |= (c[u][k] != c[u][k] ? 1u : 0u) << k;
            }
            span_merge(s, x);
        }
    }
    sh[threadIdx.x] = s;
    block_span_merge(sh);
    if (threadIdx.x == 0) parts[blockIdx.x] = sh[0];
}

/* The parts merged in order: each thread merges kSpanParts / kBlock
 * consecutive parts (loaded four at a time), then the tree. */
__global__ __launch_bounds__(kBlock) void k_coord_span_final(const CoordSpan* parts, float* out)
{
    __shared__ CoordSpan sh[kBlock];
    constexpr uint32_t kPer = kSpanParts / kBlock, kLoad = 4;
    static_assert(kPer % kLoad == 0, "whole groups of loads");
    CoordSpan r;
    r.valid = 0;
    r.nan = 0;
    for (uint32_t i0 = 0; i0 < kPer; i0 += kLoad) {
        CoordSpan loaded[kLoad];
#pragma unroll
        for (uint32_t i = 0; i < kLoad; ++i) loaded[i] = parts[threadIdx.x * kPer + i0 + i];
#pragma unroll
        for (uint32_t i = 0; i < kLoad; ++i) span_merge(r, loaded[i]);
    }
    sh[threadIdx.x] = r;
    block_span_merge(sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) { out[k] = sh[0].mn[k]; out[3 + k] = sh[0].mx[k]; }
}

static_assert(kSpanParts % kBlock == 0, "the final merge gives every thread the same number of parts");

template <typename T, int C, bool SLICE>
hipError_t launch_cells_t(const UArgs& a, bool emit, hipStream_t st)
{
    const uint64_t b = (a.ncells + kBlock - 1) / kBlock;
    const dim3 g((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 256 * 16))), blk(kBlock);
    if (emit) hipLaunchKernelGGL((k_uemit<T, C, SLICE>), g, blk, 0, st, a);
    else hipLaunchKernelGGL((k_ucount<T, C, SLICE>), g, blk, 0, st, a);
    return hipGetLastError();
}

template <typename T, bool SLICE>
hipError_t launch_cells_c(int cell, const UArgs& a, bool emit, hipStream_t st)
{
    switch (cell) {
    case kTet: return launch_cells_t<T, kTet, SLICE>(a, emit, st);
    case kHex: return launch_cells_t<T, kHex, SLICE>(a, emit, st);
    case kPyr: return launch_cells_t<T, kPyr, SLICE>(a, emit, st);
    default: return launch_cells_t<T, kPrism, SLICE>(a, emit, st);
    }
}

template <bool SLICE>
hipError_t launch_cells(int dtype, int cell, const UArgs& a, bool emit, hipStream_t st)
{
    switch (dtype) {
    case 0: return launch_cells_c<int8_t, SLICE>(cell, a, emit, st);
    case 1: return launch_cells_c<int16_t, SLICE>(cell, a, emit, st);
    case 2: return launch_cells_c<int32_t, SLICE>(cell, a, emit, st);
    case 3: return launch_cells_c<int64_t, SLICE>(cell, a, emit, st);
    case 4: return launch_cells_c<uint8_t, SLICE>(cell, a, emit, st);
    case 5: return launch_cells_c<uint16_t, SLICE>(cell, a, emit, st);
    case 6: return launch_cells_c<uint32_t, SLICE>(cell, a, emit, st);
    case 7: return launch_cells_c<uint64_t, SLICE>(cell, a, emit, st);
    case 8: return launch_cells_c<float, SLICE>(cell, a, emit, st);
    default: return launch_cells_c<double, SLICE>(cell, a, emit, st);
    }
}

/* kvs::UnstructuredVolumeObject::CellType -> cell kind (-1: no marching table). */
int cell_kind(int32_t cell_type)
{
    switch (cell_type) {
    case KMC_CELL_TETRAHEDRA: return kTet;
    case KMC_CELL_HEXAHEDRA: return kHex;
    case KMC_CELL_PYRAMID: return kPyr;
    case KMC_CELL_PRISM: return kPrism;
    default: return -1;
    }
}

int cell_nodes_of(int kind) { return kind == kTet ? 4 : kind == kHex ? 8 : kind == kPyr ? 5 : 6; }

/* The common part of both entry points: validation, device copies, min/max
 * values and coordinates, the count pass and the scan; then the emit into
 * library-owned outputs. */
kmc_status run_cells(const kmc_unstructured_volume* v, bool slice, double iso, int dup, const float* coef,
                     const uint8_t* table, uint32_t res, const float* color_range, int32_t out_memory, hipStream_t st,
                     kmc_surface* out, int normal_type)
{
    if (!v) return KMC_ERR_NULL_INPUT;
    if (v->veclen != 1) return KMC_ERR_NOT_SCALAR;
    if (type_size(v->type) == 0) return KMC_ERR_UNSUPPORTED_TYPE;
    const int kind = cell_kind(v->cell_type);
    if (kind < 0) return KMC_ERR_CELL_TYPE;
    if (!out || !v->values || !v->coords || !v->connections || v->n_nodes == 0) return KMC_ERR_INVALID_ARGUMENT;
    if (v->memory != KMC_MEMORY_HOST && v->memory != KMC_MEMORY_DEVICE) return KMC_ERR_INVALID_ARGUMENT;
    if (out_memory != KMC_MEMORY_HOST && out_memory != KMC_MEMORY_DEVICE) return KMC_ERR_INVALID_ARGUMENT;
    if (slice && (!coef || !table || res == 0)) return KMC_ERR_INVALID_ARGUMENT;
    kmc_status s = ensure_device();
    if (s != KMC_OK) return s;

    const int K = cell_nodes_of(kind);
    const uint64_t ncells = v->n_cells, nconn = ncells * (uint64_t)K;
    std::vector<void*> work;
    auto alloc = [&](size_t bytes) -> void* {
        void* p = device_alloc(bytes ? bytes : 1);
        if (p) work.push_back(p);
        return p;
    };
    auto finish = [&](kmc_status status) {
        (void)hipStreamSynchronize(st);
        for (void* p : work) device_free(p);
        return status;
    };
    /* inputs on the device */
    const size_t vb = v->n_nodes * type_size(v->type), cb = v->n_nodes * 12, kb = nconn * 4;
    const void* dvals = v->values;
    const float* dcoords = v->coords;
    const uint32_t* dconn = v->connections;
    if (v->memory == KMC_MEMORY_HOST) {
        void* a = alloc(vb);
        void* b = alloc(cb);
        void* c = alloc(kb);
        if (!a || !b || !c) return finish(KMC_ERR_OUT_OF_MEMORY);
        if (upload_host(a, v->values, vb, st) != hipSuccess || upload_host(b, v->coords, cb, st) != hipSuccess ||
            (kb && upload_host(c, v->connections, kb, st) != hipSuccess))
            return finish(KMC_ERR_DEVICE);
        dvals = a;
        dcoords = (const float*)b;
        dconn = (const uint32_t*)c;
    }
    /* MapperBase::setRange / setMinMaxCoords on an unstructured volume:
     * GetMinMaxValues and updateMinMaxCoords in connection order */
    double mn = v->min_value, mx = v->max_value;
    float minc[3], maxc[3];
    for (int k = 0; k < 3; ++k) { minc[k] = v->min_coord[k]; maxc[k] = v->max_coord[k]; }
    const bool need_values = !v->has_min_max_values, need_coords = !v->has_min_max_coords && nconn > 0;
    /* the count pass does not depend on the ranges, so it runs before their
     * read-back and one read-back brings ranges and triangle count together
     * (with an empty value range its count is discarded) */
    const bool count_now = ncells > 0 && (slice || dup);
    UArgs a;
    std::memset(&a, 0, sizeof(a));
    a.values = dvals;
    a.coords = dcoords;
    a.conn = dconn;
    a.ncells = ncells;
    a.iso = iso;
    if (slice)
        for (int k = 0; k < 4; ++k) a.plane[k] = coef[k];
    ReadPart parts[2];
    int nparts_back = 0;
    uint8_t ranges[40];
    if (need_values || need_coords) {
        uint8_t* scratch = (uint8_t*)alloc(minmax_scratch_bytes() + 64 + kSpanParts * sizeof(CoordSpan));
        if (!scratch) return finish(KMC_ERR_OUT_OF_MEMORY);
        double* dmm = (double*)(scratch + minmax_scratch_bytes());
        float* dcc = (float*)(scratch + minmax_scratch_bytes() + 16);
        CoordSpan* spans = (CoordSpan*)(scratch + minmax_scratch_bytes() + 64);
        bool ok = true;
        if (need_values) ok = launch_minmax_gather(v->type, dvals, dconn, nconn, scratch, dmm, st) == hipSuccess;
        if (ok && need_coords) {
            hipLaunchKernelGGL(k_coord_span_partial, dim3(kSpanParts), dim3(kBlock), 0, st, dcoords, dconn, nconn,
                               spans);
            hipLaunchKernelGGL(k_coord_span_final, dim3(1), dim3(kBlock), 0, st, spans, dcc);
            ok = hipGetLastError() == hipSuccess;
        }
        if (!ok) {
            (void)hipGetLastError();
            return finish(KMC_ERR_DEVICE);
        }
        /* the min/max (16 bytes) and the coordinate span (24) lie together */
        parts[nparts_back++] = ReadPart{ ranges, dmm, sizeof(ranges) };
    }
    const uint32_t nparts = count_now ? scan_partials_needed(ncells, 1) : 0;
    uint64_t* prefix = nullptr;
    uint64_t* aux = nullptr;
    uint64_t nt = 0;
    if (count_now) {
        a.counts = (uint32_t*)alloc(ncells * 4);
        a.tidx = (uint8_t*)alloc(ncells);
        prefix = (uint64_t*)alloc(ncells * 8);
        aux = (uint64_t*)alloc((nparts + 1) * 8);
        if (!a.counts || !a.tidx || !prefix || !aux) return finish(KMC_ERR_OUT_OF_MEMORY);
        a.prefix = prefix;
        const hipError_t e1 = slice ? launch_cells<true>(v->type, kind, a, false, st)
                                    : launch_cells<false>(v->type, kind, a, false, st);
        if (e1 != hipSuccess || launch_scan(a.counts, ncells, 1, aux, prefix, aux + nparts, st) != hipSuccess) {
            (void)hipGetLastError();
            return finish(KMC_ERR_DEVICE);
        }
        parts[nparts_back++] = ReadPart{ &nt, aux + nparts, 8 };
    }
    if (nparts_back && read_back_parts(parts, nparts_back, st) != hipSuccess) {
        (void)hipGetLastError();
        return finish(KMC_ERR_DEVICE);
    }
    if (need_values) {
        double hmm[2];
        std::memcpy(hmm, ranges, 16);
        mn = hmm[0];
        mx = hmm[1];
    }
    if (need_coords) {
        float hcc[6];
        std::memcpy(hcc, ranges + 16, 24);
        for (int k = 0; k < 3; ++k) { minc[k] = hcc[k]; maxc[k] = hcc[3 + k]; }
    }

    /* the surface as the mapper leaves it */
    init_surface(out, slice ? 0.0 : iso, normal_type, dup, mn, mx, minc, maxc);
    if (slice) {
        out->color_type = 0;                       // VertexColor (SlicePlane.cpp:338-340)
        out->normal_type = KMC_POLYGON_NORMAL;
    } else if (kind == kTet && !dup) {
        /* MarchingTetrahedra sets no types without duplication (:105-118): the
         * PolygonObject defaults stay, and the normal type it was given */
        out->polygon_type = 5;                     // UnknownPolygonType
        out->color_type = 2;                       // UnknownColorType
        out->normal_type = normal_type;
    } else {
        out->normal_type = KMC_POLYGON_NORMAL;    // set by mapping() whatever was asked (:100-102)
    }
    out->memory = out_memory;
    if (!slice) {
        const double d = mn - mx;
        if ((d < 0 ? -d : d) < 1.e-6) { out->empty_range = 1; return finish(KMC_OK); }   // Math::Equal exit
        if (!dup) return finish(KMC_OK);          // no shared-vertex extraction in the reference
    }
    if (ncells == 0) return finish(KMC_OK);

    a.minv = mn;
    a.maxv = mx;
    if (slice) {
        a.cres = res;
        if (color_range) { a.crange[0] = color_range[0]; a.crange[1] = color_range[1]; }
        else if (v->type == KMC_INT8) { a.crange[0] = -128; a.crange[1] = 127; }
        else if (v->type == KMC_UINT8) { a.crange[0] = 0; a.crange[1] = 255; }
        else {
            a.crange[0] = (float)mn;
            a.crange[1] = (float)mx;
            const float dd = a.crange[0] - a.crange[1];
            if ((double)(dd < 0 ? -dd : dd) < 1.e-6) a.crange[1] = a.crange[1] + 1.0f;
        }
        uint8_t* dt = (uint8_t*)alloc(3 * (size_t)res);
        if (!dt || hipMemcpyAsync(dt, table, 3 * (size_t)res, hipMemcpyHostToDevice, st) != hipSuccess)
            return finish(dt ? KMC_ERR_DEVICE : KMC_ERR_OUT_OF_MEMORY);
        a.ctable = dt;
    }
    if (nt == 0) return finish(KMC_OK);

    SurfaceStorage* store = new SurfaceStorage();
    store->memory = KMC_MEMORY_DEVICE;
    out->internal = store;
    out->memory = KMC_MEMORY_DEVICE;
    out->n_triangles = nt;
    out->n_vertices = 3 * nt;
    out->n_normals = nt;
    store->coords = device_alloc(9 * nt * 4);
    store->normals = device_alloc(3 * nt * 4);
    if (slice) store->colors = device_alloc(9 * nt);
    if (!store->coords || !store->normals || (slice && !store->colors)) {
        kmc_surface_release(out);
        return finish(KMC_ERR_OUT_OF_MEMORY);
    }
    out->coords = (float*)store->coords;
    out->normals = (float*)store->normals;
    out->colors = (uint8_t*)store->colors;
    out->n_colors = slice ? 3 * nt : 0;
    a.out_coords = out->coords;
    a.out_normals = out->normals;
    a.out_colors = out->colors;
    const hipError_t e2 = slice ? launch_cells<true>(v->type, kind, a, true, st)
                                : launch_cells<false>(v->type, kind, a, true, st);
    if (e2 != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
        (void)hipGetLastError();
        kmc_surface_release(out);
        return finish(KMC_ERR_DEVICE);
    }
    if (out_memory == KMC_MEMORY_HOST) {
        s = to_host(out, st);
        if (s != KMC_OK) kmc_surface_release(out);
    }
    return finish(s);
}

}  // namespace

}  // namespace kmc

extern "C" {

kmc_status kmc_extract_unstructured(const kmc_unstructured_volume* volume, double isolevel, int32_t normal_type,
                                    int32_t duplication, int32_t out_memory, void* stream, kmc_surface* out)
{
    return kmc::abi_guard([&]() -> kmc_status {
        kmc_status s = KMC_OK;
        if (normal_type != KMC_VERTEX_NORMAL && normal_type != KMC_POLYGON_NORMAL) s = KMC_ERR_INVALID_ARGUMENT;
        const int dup = (normal_type == KMC_VERTEX_NORMAL) ? 0 : (duplication ? 1 : 0);   // MarchingTetrahedra.cpp:78-81
        if (s == KMC_OK)
            s = kmc::run_cells(volume, false, isolevel, dup, nullptr, nullptr, 0, nullptr, out_memory, (hipStream_t)stream,
                               out, normal_type);
        kmc::report(s);
        return s;
    });
}

kmc_status kmc_slice_plane_unstructured(const kmc_unstructured_volume* volume, const float coefficients[4],
                                        const uint8_t* color_table, uint32_t color_resolution,
                                        const float* color_range, int32_t out_memory, void* stream, kmc_surface* out)
{
    return kmc::abi_guard([&]() -> kmc_status {
        const kmc_status s = kmc::run_cells(volume, true, 0.0, 1, coefficients, color_table, color_resolution, color_range,
                                            out_memory, (hipStream_t)stream, out, KMC_POLYGON_NORMAL);
        kmc::report(s);
        return s;
    });
}

}  // extern "C"
