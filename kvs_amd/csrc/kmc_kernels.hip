/*
 * kmc_kernels.hip -- marching-cubes kernels for gfx950 (CDNA4, wave64).
 *
 * The path is a memory-bound stencil followed by an ordered stream
 * compaction; nothing here is a contraction, so nothing goes to MFMA.
 *
 * Classification is bit-sliced: a wave loads 64 consecutive nodes of a row,
 * compares them against every isolevel of the pass and turns each level's
 * predicate into a 64-bit chunk mask with one ballot.  Edge crossings are XORs
 * of (shifted) chunk masks, cell activity is OR & ~AND of the eight corner
 * masks, and every rank needed for the ordered compaction is a popcount.
 * Dense per-node work is therefore one load, one compare per level and one
 * ballot per level; only crossing edges and active cells do more.
 *
 * Passes: classify (values -> masks), count (per-row counts and in-row chunk
 * prefixes), scan (row prefixes), then the flat emit: a compaction writes
 * every vertex's key and every triangle's record at its output rank, and one
 * thread per vertex / triangle / vertex normal produces the outputs.  The
 * tile-team emit (each wave marching a band of rows through z with LDS rings
 * of masks) remains for volumes whose rows exceed the records' fields.
 *
 * Arithmetic follows Source/Core/Visualization/Mapper/MarchingCubes.cpp bit for
 * bit: the comparison double(v) > isolevel (:299-316), the FP64 clamp / ratio /
 * lerp of interpolate_vertex (:326-356), the float (p + min) * scale of the
 * callers (:169-173, :382-386) and the float cross product of Vector3::cross
 * (Source/Core/Matrix/Vector3.h:479-485).  The file is compiled with
 * -ffp-contract=off so that no multiply-add is fused.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <map>
#include <mutex>
#include "kmc_job.h"
#include "mc_tables.h"

#pragma clang fp contract(off)

namespace kmc {

__constant__ uint64_t d_tri[256] = { KMC_TABLE_WORDS };

/* Emit modes, a template parameter so that each kernel carries only its own
 * path: duplicated vertices, shared vertices with PolygonNormal, shared
 * vertices with VertexNormal. */
enum : int { kModeDup = 0, kModePolygon = 1, kModeVertex = 2 };

__host__ __device__ inline int emit_mode(const JobArgs& a) { return a.dup ? kModeDup : (a.vnormal ? kModeVertex : kModePolygon); }

/* ------------------------------------------------------------------------- */
/* classification                                                              */
/* ------------------------------------------------------------------------- */

template <typename T>
struct Classify {
    /* bit l set iff double(v) > iso[l]  (MarchingCubes.cpp:306-313) */
    static __device__ __forceinline__ uint32_t bits(T v, const JobArgs& a)
    {
        const double d = (double)v;
        uint32_t r = 0;
#pragma unroll
        for (int l = 0; l < kMaxLevels; ++l)
            if (l < (int)a.nlev) r |= (d > a.iso[l] ? 1u : 0u) << l;
        return r;
    }
};

template <>
struct Classify<float> {
    /* For float data, double(v) > iso  <=>  v > thr with thr the largest
     * float not above iso (computed on the host; NaN for a NaN isolevel). */
    static __device__ __forceinline__ uint32_t bits(float v, const JobArgs& a)
    {
        uint32_t r = 0;
#pragma unroll
        for (int l = 0; l < kMaxLevels; ++l)
            if (l < (int)a.nlev) r |= (v > a.thr[l] ? 1u : 0u) << l;
        return r;
    }
};

__device__ __forceinline__ double iso_of(const JobArgs& a, int l)
{
    double r = a.iso[0];
#pragma unroll
    for (int k = 1; k < kMaxLevels; ++k)
        if (k == l) r = a.iso[k];
    return r;
}

/* ------------------------------------------------------------------------- */
/* interpolation and normals                                                   */
/* ------------------------------------------------------------------------- */

template <typename T>
__device__ __forceinline__ void interpolate_values(const JobArgs& a, double iso, T value_a, T value_b,
                                                   uint32_t ax, uint32_t ay, uint32_t az,
                                                   uint32_t bx, uint32_t by, uint32_t bz, float out[3]);

/* interpolate_vertex<T> + (p + min_coord) * scale_factor for the edge from
 * node a to node b (integer node coordinates). */
template <typename T>
__device__ __forceinline__ void interpolate(const JobArgs& a, const T* vol, double iso,
                                            uint32_t ax, uint32_t ay, uint32_t az,
                                            uint32_t bx, uint32_t by, uint32_t bz, float out[3])
{
    const uint64_t line = a.nx, slice = (uint64_t)a.nx * a.ny;
    const uint64_t ia = ax + ay * line + az * slice;
    const uint64_t ib = bx + by * line + bz * slice;
    interpolate_values<T>(a, iso, vol[ia], vol[ib], ax, ay, az, bx, by, bz, out);
}

/* The interpolation of interpolate() from the two node values. */
template <typename T>
__device__ __forceinline__ void interpolate_values(const JobArgs& a, double iso, T value_a, T value_b,
                                                   uint32_t ax, uint32_t ay, uint32_t az,
                                                   uint32_t bx, uint32_t by, uint32_t bz, float out[3])
{
    double va = (double)value_a;
    double vb = (double)value_b;
    va = va <= a.minv ? a.minv : (va <= a.maxv ? va : a.maxv);
    vb = vb <= a.minv ? a.minv : (vb <= a.maxv ? vb : a.maxv);
    double r = (iso - va) / (vb - va);
    r = r < 0 ? -r : r;
    const double s = 1.0 - r;
    const double pa[3] = { (double)(float)ax, (double)(float)ay, (double)(float)az };
    const double pb[3] = { (double)(float)bx, (double)(float)by, (double)(float)bz };
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double t0 = s * pa[k];
        const double t1 = r * pb[k];
        const float p = (float)(t0 + t1);
        const float q = p + a.minc[k];
        out[k] = q * a.sc[k];
    }
}

__device__ __forceinline__ void face_normal(const float v0[3], const float v1[3], const float v2[3], float n[3])
{
    const float d1x = v1[0] - v0[0], d1y = v1[1] - v0[1], d1z = v1[2] - v0[2];
    const float d2x = v2[0] - v0[0], d2y = v2[1] - v0[1], d2z = v2[2] - v0[2];
    const float p0 = d1y * d2z, p1 = d1z * d2y;
    const float p2 = d1z * d2x, p3 = d1x * d2z;
    const float p4 = d1x * d2y, p5 = d1y * d2x;
    n[0] = p0 - p1;
    n[1] = p2 - p3;
    n[2] = p4 - p5;
}

/* ------------------------------------------------------------------------- */
/* tile-team emit                                                              */
/* ------------------------------------------------------------------------- */

/* The emit pass runs one independent team per wave: a team marches its own
 * unit (level, band, slab) in its own LDS region, so nothing in the emit pass
 * waits for another wave.  The LDS operations of one wave complete in issue
 * order; wave_sync only keeps the compiler from moving memory operations
 * across the points where lanes start reading what other lanes wrote. */
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t uniform(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }

__device__ __forceinline__ uint64_t uniform64(uint64_t x)
{
    return (uint64_t)uniform((uint32_t)(x >> 32)) << 32 | uniform((uint32_t)x);
}

__device__ __forceinline__ uint64_t readlane64(uint64_t x, int k)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, k);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), k);
    return (uint64_t)hi << 32 | lo;
}

/* Set bits of the wave mask m on the lanes below this one (v_mbcnt). */
__device__ __forceinline__ uint32_t lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

/* LDS arrays of one team (layout of lds_layout) and the shared case table. */
struct Team {
    const uint64_t* tri;  // [256] case table
    uint64_t* bits;       // [nslots][nrb][NC]
    uint32_t* wpv;        // [2][R+2][NC]
    uint64_t* act;        // [ncs][R+1][NC]
    uint32_t* pt;         // [ncs][R+1][NC]
    uint32_t* ap;         // [ncs][R+1][NC]
    uint32_t* crec;       // [ncs][fn_cap]
    float* fn;            // [ncs][fn_cap][3]
    uint32_t* vq;         // [queue_cap]
    uint32_t* tq;         // [queue_cap]
    uint32_t nrb, R, NC, fn_cap;
};

__device__ __forceinline__ Team make_team(const JobArgs& a, uint8_t* smem, uint32_t wave)
{
    const LdsLayout L = lds_layout(a);
    uint8_t* base = smem + kTableBytes + wave * a.team_stride;
    Team v;
    v.tri = (const uint64_t*)smem;
    v.bits = (uint64_t*)(base + L.off_bits);
    v.wpv = (uint32_t*)(base + L.off_wpv);
    v.act = (uint64_t*)(base + L.off_act);
    v.pt = (uint32_t*)(base + L.off_pt);
    v.ap = (uint32_t*)(base + L.off_ap);
    v.crec = (uint32_t*)(base + L.off_crec);
    v.fn = (float*)(base + L.off_fn);
    v.vq = (uint32_t*)(base + L.off_vq);
    v.tq = (uint32_t*)(base + L.off_tq);
    v.nrb = L.nrb;
    v.R = a.R;
    v.NC = a.NC;
    v.fn_cap = a.fn_cap;
    return v;
}

/* The level a unit works on: its bit volume, prefix segments and outputs. */
struct Level {
    const uint64_t* bits;    // [nz][ny][NC] chunk masks of the level
    const uint64_t* pre_v;   // prefixes of per-node-row vertex counts
    const uint64_t* pre_t;   // prefixes of per-cell-row triangle counts
    uint64_t bv, bt;         // first prefix of each segment (outputs index from 0)
    double iso;
    float* coords;
    float* normals;
    uint32_t* conn;
    uint64_t cap_v, cap_t;
};

__device__ __forceinline__ Level level_of(const JobArgs& a, uint32_t l)
{
    Level L;
    L.bits = a.bitvol + (uint64_t)l * ((uint64_t)a.nz * a.ny * a.NC);
    L.pre_v = a.prefix + (uint64_t)(l * 2 + 0) * a.nrows;
    L.pre_t = a.prefix + (uint64_t)(l * 2 + 1) * a.nrows;
    L.bv = L.pre_v[0];
    L.bt = L.pre_t[0];
    L.iso = a.iso[0];
    L.coords = a.coords[0];
    L.normals = a.normals[0];
    L.conn = a.conn[0];
    L.cap_v = a.cap_vertices[0];
    L.cap_t = a.cap_triangles[0];
#pragma unroll
    for (int k = 1; k < kMaxLevels; ++k) {
        if (k == (int)l) {
            L.iso = a.iso[k];
            L.coords = a.coords[k];
            L.normals = a.normals[k];
            L.conn = a.conn[k];
            L.cap_v = a.cap_vertices[k];
            L.cap_t = a.cap_triangles[k];
        }
    }
    return L;
}

/* Ring slot of node slice s (s >= -2): any 3 consecutive slices get distinct
 * slots. */
template <int M>
__device__ __forceinline__ uint32_t bslot(int s) { return (uint32_t)(s + 3) % 3u; }

/* Ring slot of cell slice zc (VertexNormal keeps two cell slices). */
template <int M>
__device__ __forceinline__ uint32_t cslot(int zc) { return M == kModeVertex ? ((uint32_t)zc & 1u) : 0u; }

/* Row of chunk masks: slot sl, LDS row r (node row y = y0 - 1 + r). */
__device__ __forceinline__ const uint64_t* brow(const Team& v, uint32_t sl, uint32_t r)
{
    return v.bits + (sl * v.nrb + r) * v.NC;
}
/* Window vertex chunk prefixes of window slice ws (0: node slice z-1, 1: z)
 * of tile z, window row w.  The two slots hold node slices by parity, so
 * slice z's prefixes of step z are slice z-1's of step z+1. */
__device__ __forceinline__ uint32_t* wrow(const Team& v, int z, uint32_t ws, uint32_t w)
{
    const uint32_t slot = (uint32_t)(z - 1 + (int)ws) & 1u;
    return v.wpv + (slot * (v.R + 2) + w) * v.NC;
}
/* Cell data index: cell slot cs, cell row j (cell row y0 - 1 + j), chunk c. */
__device__ __forceinline__ uint32_t cidx(const Team& v, uint32_t cs, uint32_t j, uint32_t c)
{
    return (cs * (v.R + 1) + j) * v.NC + c;
}

/* Mask of the nodes x < limit inside chunk c. */
__device__ __forceinline__ uint64_t valid_mask(uint32_t c, uint32_t limit)
{
    const uint32_t x0 = c * 64u;
    if (limit <= x0) return 0ull;
    const uint32_t n = limit - x0;
    return n >= 64u ? ~0ull : ((1ull << n) - 1ull);
}

/* Bits of chunk c of a row of NC chunks shifted so that bit i holds node x+1. */
__device__ __forceinline__ uint64_t shift_next(const uint64_t* row, uint32_t c, uint32_t NC)
{
    return (row[c] >> 1) | ((c + 1 < NC ? row[c + 1] : 0ull) << 63);
}

struct Cross { uint64_t x, y, z; };

/* Crossing edges owned by the nodes of chunk c of window row w (node row
 * y0-1+w) of window slice ws (node slice z-1+ws).  z-edges reach slice
 * z+ws, whose masks the ring holds. */
template <int M>
__device__ __forceinline__ Cross window_cross(const JobArgs& a, const Team& v, int z, int y0,
                                              uint32_t ws, uint32_t w, uint32_t c)
{
    Cross k;
    const int s = z - 1 + (int)ws;
    const int y = y0 - 1 + (int)w;
    if (s < 0 || s >= (int)a.nz || y < 0 || y >= (int)a.ny) { k.x = k.y = k.z = 0; return k; }
    const uint64_t* row = brow(v, bslot<M>(s), w);
    const uint64_t b = row[c];
    k.x = (b ^ shift_next(row, c, a.NC)) & valid_mask(c, a.nx - 1);
    k.y = (y + 1 < (int)a.ny) ? ((b ^ brow(v, bslot<M>(s), w + 1)[c]) & valid_mask(c, a.nx)) : 0ull;
    k.z = (s + 1 < (int)a.nz) ? ((b ^ brow(v, bslot<M>(s + 1), w)[c]) & valid_mask(c, a.nx)) : 0ull;
    return k;
}

/* The eight corner masks of the cells of chunk c, cell row cy (node rows cy,
 * cy+1), cell slice cz (node slices cz, cz+1), shifted so that bit i is
 * corner k of cell x = 64c + i (corner order of MarchingCubes.cpp:187-194). */
struct Corners { uint64_t c[8]; };

template <int M>
__device__ __forceinline__ Corners corners_at(const JobArgs& a, const Team& v, int y0, uint32_t c, int cy, int cz)
{
    const uint32_t r = (uint32_t)(cy - y0 + 1);
    const uint64_t* r0 = brow(v, bslot<M>(cz), r);
    const uint64_t* r1 = brow(v, bslot<M>(cz), r + 1);
    const uint64_t* r2 = brow(v, bslot<M>(cz + 1), r);
    const uint64_t* r3 = brow(v, bslot<M>(cz + 1), r + 1);
    Corners k;
    k.c[0] = r0[c]; k.c[1] = shift_next(r0, c, a.NC); k.c[3] = r1[c]; k.c[2] = shift_next(r1, c, a.NC);
    k.c[4] = r2[c]; k.c[5] = shift_next(r2, c, a.NC); k.c[7] = r3[c]; k.c[6] = shift_next(r3, c, a.NC);
    return k;
}

__device__ __forceinline__ uint32_t corner_index(const Corners& k, uint32_t i)
{
    uint32_t ci = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) ci |= (uint32_t)((k.c[q] >> i) & 1ull) << q;
    return ci;
}

/* Cube index of cell (cx, cy, cz) from the mask ring (y0: first band row). */
template <int M>
__device__ __forceinline__ uint32_t cube_index_at(const Team& v, int y0, uint32_t cx, int cy, int cz)
{
    const uint32_t r = (uint32_t)(cy - y0 + 1);
    const uint64_t* b00 = brow(v, bslot<M>(cz), r);
    const uint64_t* b01 = brow(v, bslot<M>(cz), r + 1);
    const uint64_t* b10 = brow(v, bslot<M>(cz + 1), r);
    const uint64_t* b11 = brow(v, bslot<M>(cz + 1), r + 1);
    const uint32_t c0 = cx >> 6, i0 = cx & 63u;
    const uint32_t c1 = (cx + 1) >> 6, i1 = (cx + 1) & 63u;
    uint32_t ci = 0;
    ci |= (uint32_t)((b00[c0] >> i0) & 1ull) << 0;
    ci |= (uint32_t)((b00[c1] >> i1) & 1ull) << 1;
    ci |= (uint32_t)((b01[c1] >> i1) & 1ull) << 2;
    ci |= (uint32_t)((b01[c0] >> i0) & 1ull) << 3;
    ci |= (uint32_t)((b10[c0] >> i0) & 1ull) << 4;
    ci |= (uint32_t)((b10[c1] >> i1) & 1ull) << 5;
    ci |= (uint32_t)((b11[c1] >> i1) & 1ull) << 6;
    ci |= (uint32_t)((b11[c0] >> i0) & 1ull) << 7;
    return ci;
}

/* Cube index of cell (cx, cy, cz) read from the level's bit volume in memory
 * (for cells whose node slices have left the mask ring). */
__device__ __forceinline__ uint32_t cube_index_mem(const JobArgs& a, const uint64_t* bits, uint32_t cx, int cy, int cz)
{
    const uint64_t* b00 = bits + ((uint64_t)cz * a.ny + (uint64_t)cy) * a.NC;
    const uint64_t* b01 = b00 + a.NC;
    const uint64_t* b10 = b00 + (uint64_t)a.ny * a.NC;
    const uint64_t* b11 = b10 + a.NC;
    const uint32_t c0 = cx >> 6, i0 = cx & 63u;
    const uint32_t c1 = (cx + 1) >> 6, i1 = (cx + 1) & 63u;
    uint32_t ci = 0;
    ci |= (uint32_t)((b00[c0] >> i0) & 1ull) << 0;
    ci |= (uint32_t)((b00[c1] >> i1) & 1ull) << 1;
    ci |= (uint32_t)((b01[c1] >> i1) & 1ull) << 2;
    ci |= (uint32_t)((b01[c0] >> i0) & 1ull) << 3;
    ci |= (uint32_t)((b10[c0] >> i0) & 1ull) << 4;
    ci |= (uint32_t)((b10[c1] >> i1) & 1ull) << 5;
    ci |= (uint32_t)((b11[c1] >> i1) & 1ull) << 6;
    ci |= (uint32_t)((b11[c0] >> i0) & 1ull) << 7;
    return ci;
}

/* ------------------------------------------------------------------------- */
/* grid-stride ranges                                                          */
/* ------------------------------------------------------------------------- */

/* (Contiguous parts per XCD -- workgroups reach the eight XCDs round-robin,
 * each with an L2 of its own -- measured 3 % slower on the whole call than
 * this plain interleave: compaction and triangles lost more than count
 * gained.) */
struct Span { uint64_t lo, first, end, step; };

/* Grid-stride walk over [b, e) in units of `units` items per workgroup
 * (`unit`: this thread's or wave's place in its workgroup). */
__device__ __forceinline__ Span grid_span(uint64_t b, uint64_t e, uint32_t units, uint32_t unit)
{
    Span sp;
    sp.lo = b;
    sp.first = b + (uint64_t)blockIdx.x * units + unit;
    sp.end = e;
    sp.step = (uint64_t)gridDim.x * units;
    return sp;
}

/* ------------------------------------------------------------------------- */
/* wave scan                                                                   */
/* ------------------------------------------------------------------------- */

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(x, d);
        if (lane >= d) x += t;
    }
    return x;
}

/* In-place exclusive scan of n u32 values in LDS by one wave; returns the
 * total (the same in every lane). */
__device__ __forceinline__ uint32_t wave_exclusive_scan(uint32_t* p, uint32_t n)
{
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t carry = 0;
    for (uint32_t g = 0; g < n; g += 64) {
        const uint32_t i = g + lane;
        const uint32_t x = i < n ? p[i] : 0u;
        const uint32_t inc = wave_incl_scan(x);
        if (i < n) p[i] = carry + inc - x;
        carry += __shfl(inc, 63);
    }
    return carry;
}

/* ------------------------------------------------------------------------- */
/* slice loading                                                               */
/* ------------------------------------------------------------------------- */

/* Bring the chunk masks of node rows y0-1 .. y0+R+1 of node slices
 * s_lo .. s_hi of the level into their ring slots; rows and slices outside
 * the volume are zero.  For one slice the rows are contiguous both in the
 * bit volume and in LDS, so the copy is a flat range per slice; a lane's
 * loads are all issued before its stores. */
template <int M>
__device__ void load_slices(const JobArgs& a, const Team& v, const Level& L, int s_lo, int s_hi, int y0)
{
    if (s_hi < s_lo) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nseg = v.nrb * a.NC;              // elements of one slice
    const uint32_t total = (uint32_t)(s_hi - s_lo + 1) * nseg;
    const uint32_t e_lo = (y0 - 1 < 0) ? a.NC : 0u;
    const int rows_left = (int)a.ny - (y0 - 1);
    const uint32_t e_hi = (rows_left >= (int)v.nrb ? v.nrb : (uint32_t)(rows_left > 0 ? rows_left : 0)) * a.NC;
    const uint64_t slice_words = (uint64_t)a.ny * a.NC;
    const int64_t row0 = (int64_t)(y0 - 1) * a.NC;
    constexpr int kB = 4;
    /* element i = b0 + 64u + lane is element e of slice s_lo + si */
    uint32_t si = 0, e = lane;
    while (e >= nseg) { e -= nseg; ++si; }
    for (uint32_t b0 = 0; b0 < total; b0 += 64u * kB) {
        uint64_t val[kB];
        uint32_t dst[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const uint32_t i = b0 + 64u * u + lane;
            const int s = s_lo + (int)si;
            const bool ok = i < total && s >= 0 && s < (int)a.nz && e >= e_lo && e < e_hi;
            val[u] = ok ? L.bits[(uint64_t)((int64_t)s * (int64_t)slice_words + row0 + (int64_t)e)] : 0ull;
            dst[u] = i < total ? bslot<M>(s) * nseg + e : ~0u;
            e += 64u;
            while (e >= nseg) { e -= nseg; ++si; }
        }
#pragma unroll
        for (int u = 0; u < kB; ++u)
            if (dst[u] != ~0u) v.bits[dst[u]] = val[u];
    }
}

/* ------------------------------------------------------------------------- */
/* per-tile counts and prefixes (lanes on (row, chunk))                        */
/* ------------------------------------------------------------------------- */

/* Totals of a tile: vertices it emits, triangles of its cell slice (halo row
 * included) and triangles of the halo row. */
struct TileTotals { uint32_t nv, nt, halo; };

/* Tile z: window vertex counts per chunk of window slices 0, 1 (node slices
 * z-1, z), rows 1..R+1 (for output ranks), and active cells / triangle
 * counts of cell slice z-1, cell rows j0..R (j = 0: the halo row y0-1, only
 * with VertexNormal); then exclusive prefixes -- window prefixes per slice
 * and triangle prefixes.  ws_v: window slice whose vertices the tile emits.
 * Lanes cover 64 >> cshift rows of 1 << cshift chunk lanes. */
template <int M>
__device__ TileTotals tile_prefixes(const JobArgs& a, const Team& v, int z, int y0, uint32_t ws_v, bool ranks,
                                    bool first)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t sub = lane >> a.cshift, cl = lane & ((1u << a.cshift) - 1u);
    const uint32_t G = 64u >> a.cshift, CS = 1u << a.cshift;
    const uint32_t nwr = a.R + 2;
    constexpr bool halo = M == kModeVertex;
    constexpr bool dup = M == kModeDup;
    const uint32_t j0 = halo ? 0u : 1u;
    const int zc = z - 1;
    const uint32_t cs = cslot<M>(zc);
    const bool wranks = ranks && !dup;
    /* window rows 1..R+1 of window slice 1 (and of slice 0 at a unit's first
     * step; later it is the previous step's slice 1) */
    if (wranks) {
        for (uint32_t ws = first ? 0u : 1u; ws < 2u; ++ws) {
            for (uint32_t w = 1 + sub; w < nwr; w += G) {
                uint32_t* dst = wrow(v, z, ws, w);
                for (uint32_t c = cl; c < a.NC; c += CS) {
                    const Cross k = window_cross<M>(a, v, z, y0, ws, w, c);
                    dst[c] = (uint32_t)(__popcll(k.x) + __popcll(k.y) + __popcll(k.z));
                }
            }
        }
    }
    for (uint32_t j = sub; j <= a.R; j += G) {
        const int y = y0 - 1 + (int)j;
        const bool live = j >= j0 && zc >= 0 && zc + 1 < (int)a.nz && y >= 0 && y + 1 < (int)a.ny;
        for (uint32_t c = cl; c < a.NC; c += CS) {
            uint64_t act = 0;
            uint32_t cnt = 0;
            if (live) {
                const Corners k = corners_at<M>(a, v, y0, c, y, zc);
                uint64_t o = 0, an = ~0ull;
#pragma unroll
                for (int q = 0; q < 8; ++q) { o |= k.c[q]; an &= k.c[q]; }
                act = (o & ~an) & valid_mask(c, a.nx - 1);
                uint64_t m = act;
                while (m) {
                    const uint32_t i = (uint32_t)__ffsll((unsigned long long)m) - 1u;
                    m &= m - 1ull;
                    cnt += (uint32_t)(v.tri[corner_index(k, i)] >> 60);
                }
            }
            v.act[cidx(v, cs, j, c)] = act;
            v.pt[cidx(v, cs, j, c)] = cnt;
            if (halo) v.ap[cidx(v, cs, j, c)] = (uint32_t)__popcll(act);
        }
    }
    wave_sync();
    if (halo) (void)wave_exclusive_scan(v.ap + cidx(v, cs, 0, 0), (a.R + 1) * a.NC);
    if (wranks) {
        for (uint32_t ws = first ? 0u : 1u; ws < 2u; ++ws)
            (void)wave_exclusive_scan(wrow(v, z, ws, 1), (nwr - 1) * a.NC);
    }
    TileTotals t;
    t.nt = uniform(wave_exclusive_scan(v.pt + cidx(v, cs, 0, 0), (a.R + 1) * a.NC));
    wave_sync();
    t.nv = wranks ? uniform(wrow(v, z, ws_v, a.R + 1)[0] - wrow(v, z, ws_v, 1)[0]) : 0u;
    t.halo = halo ? uniform(v.pt[cidx(v, cs, 1, 0)]) : 0u;
    return t;
}

/* ------------------------------------------------------------------------- */
/* vertex coordinates and normals                                              */
/* ------------------------------------------------------------------------- */

/* Coordinates of the shared vertex on the edge of node (x,y,z) along axis,
 * interpolated low -> high as calculate_isopoints does (:406,420,434). */
template <typename T>
__device__ __forceinline__ void node_edge_coords(const JobArgs& a, const T* vol, double iso,
                                                 uint32_t x, uint32_t y, uint32_t z, uint32_t axis, float out[3])
{
    interpolate<T>(a, vol, iso, x, y, z, x + (axis == 0), y + (axis == 1), z + (axis == 2), out);
}

/* Owner node offset (dx, dy, dz) and axis of local edge e, packed as
 * dx | dy << 1 | dz << 2 | axis << 3 (connect_isopoints, :493-504), five
 * bits per edge. */
__device__ __forceinline__ uint32_t edge_owner(uint32_t e)
{
    constexpr uint64_t kOwner =
        (0x00ull << 0) | (0x09ull << 5) | (0x02ull << 10) | (0x08ull << 15) | (0x04ull << 20) | (0x0Dull << 25) |
        (0x06ull << 30) | (0x0Cull << 35) | (0x10ull << 40) | (0x11ull << 45) | (0x13ull << 50) | (0x12ull << 55);
    return (uint32_t)(kOwner >> (5 * e)) & 31u;
}

/* Table-orientation endpoints of local edge e as two corner offsets packed
 * bit0 x, bit1 y, bit2 z, first endpoint in the low nibble
 * (MarchingCubesTable.cpp:276-290), eight bits per edge. */
__device__ __forceinline__ uint32_t edge_ends(uint32_t e)
{
    constexpr uint64_t kLo = 0x4667755402233110ull;   // edges 0..7
    constexpr uint32_t kHi = 0x62735140u;             // edges 8..11
    return e < 8 ? (uint32_t)(kLo >> (8 * e)) & 0xFFu : (kHi >> (8 * (e - 8))) & 0xFFu;
}

/* Output bases of tile (z, band): vertex prefixes of node rows (z-1, y0),
 * (z-1, y0+R), (z, y0), (z, y0+R) and the triangle prefix of cell row
 * (y0, z-1), relative to the first row of their segment; one lane loads each
 * and every lane gets all five. */
struct Bases { uint64_t v[4]; uint64_t t; };

__device__ __forceinline__ Bases tile_bases(const JobArgs& a, const Level& L, int z, int y0)
{
    const uint32_t lane = threadIdx.x & 63u;
    const bool in = z < (int)a.nz, next = (uint32_t)y0 + a.R < a.ny;
    const uint64_t row = (uint64_t)z * a.ny + (uint64_t)y0;
    const uint64_t prow = (uint64_t)(z - 1) * a.ny + (uint64_t)y0;
    const uint64_t* p = nullptr;
    uint64_t b = 0;
    if (lane == 0 && z > 0) { p = L.pre_v + prow; b = L.bv; }
    if (lane == 1 && z > 0 && next) { p = L.pre_v + prow + a.R; b = L.bv; }
    if (lane == 2 && in) { p = L.pre_v + row; b = L.bv; }
    if (lane == 3 && in && next) { p = L.pre_v + row + a.R; b = L.bv; }
    if (lane == 4 && in) { p = L.pre_t + row; b = L.bt; }
    const uint64_t x = p ? *p - b : 0ull;
    Bases B;
    B.v[0] = readlane64(x, 0);
    B.v[1] = readlane64(x, 1);
    B.v[2] = readlane64(x, 2);
    B.v[3] = readlane64(x, 3);
    B.t = readlane64(x, 4);
    return B;
}

/* Rank within its tile row range of the vertex on (node x of band row k (0..R)
 * of node slice s in {z-1, z}, axis): rows 0..R-1 rank from the band's first
 * row, row R (first row of the next band) from its own start. */
template <int M>
__device__ __forceinline__ uint32_t tile_rank(const JobArgs& a, const Team& v, int z, int y0,
                                              uint32_t x, uint32_t k, int s, uint32_t axis)
{
    const uint32_t ws = (uint32_t)(s - z + 1), w = k + 1;
    const uint32_t c = x >> 6, i = x & 63u;
    const Cross e = window_cross<M>(a, v, z, y0, ws, w, c);
    const uint64_t lo = (1ull << i) - 1ull;
    const uint32_t start = wrow(v, z, ws, k == a.R ? a.R + 1 : 1)[0];
    uint32_t r = wrow(v, z, ws, w)[c] - start;
    r += __popcll(e.x & lo) + __popcll(e.y & lo) + __popcll(e.z & lo);
    if (axis >= 1) r += (uint32_t)((e.x >> i) & 1ull);
    if (axis >= 2) r += (uint32_t)((e.y >> i) & 1ull);
    return r;
}

/* The four cells around the edge of node (x, y, zv) along axis, in scan
 * order: offsets across the edge are -1/0 along the other two axes (low axis
 * first); the edge's local id in cell q is the q-th nibble of `around`. */
__device__ __forceinline__ void around_cell(uint32_t axis, int q, int x, int y, int zv, int& cx, int& cy, int& cz)
{
    const int lo = (q & 1) ? 0 : -1, hi = (q & 2) ? 0 : -1;
    cx = x + (axis == 0 ? 0 : lo);
    cy = y + (axis == 0 ? lo : (axis == 1 ? 0 : hi));
    cz = zv + (axis == 2 ? 0 : hi);
}

__device__ __forceinline__ uint32_t around_nibbles(uint32_t axis)
{
    return axis == 0 ? 0x0246u : (axis == 1 ? 0x3175u : 0x89BAu);
}

/* VertexNormal of the vertex on (x,y,zv,axis): the float sum, in triangle
 * order, of the unnormalized face normals of the triangles that use it --
 * the order of calculate_normals_on_vertex (:569-605).  Incident cells are
 * visited in cell scan order and triangles in table order; cube indices come
 * from the bit volume and every face normal is recomputed from interpolated
 * coordinates (used when the face normals of a cell slice did not fit in
 * LDS). */
template <typename T>
__device__ void vertex_normal_direct(const JobArgs& a, const Team& v, const Level& L, const T* vol,
                                     uint32_t x, uint32_t y, int zv, uint32_t axis, float n[3])
{
    n[0] = 0.0f; n[1] = 0.0f; n[2] = 0.0f;
    const uint32_t around = around_nibbles(axis);
    for (int q = 0; q < 4; ++q) {
        int cx, cy, cz;
        around_cell(axis, q, (int)x, (int)y, zv, cx, cy, cz);
        if (cx < 0 || cy < 0 || cz < 0) continue;
        if (cx + 1 >= (int)a.nx || cy + 1 >= (int)a.ny || cz + 1 >= (int)a.nz) continue;
        const uint32_t ci = cube_index_mem(a, L.bits, (uint32_t)cx, cy, cz);
        const uint64_t w = v.tri[ci];
        const int nt = (int)(w >> 60);
        const uint32_t me = (around >> (4 * q)) & 0xFu;
        for (int t = 0; t < nt; ++t) {
            const uint32_t e0 = (uint32_t)((w >> (12 * t)) & 0xF);
            const uint32_t e1 = (uint32_t)((w >> (12 * t + 8)) & 0xF);
            const uint32_t e2 = (uint32_t)((w >> (12 * t + 4)) & 0xF);
            if (e0 != me && e1 != me && e2 != me) continue;
            float p[3][3], fn[3];
            const uint32_t es[3] = { e0, e1, e2 };
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const uint32_t o = edge_owner(es[j]);
                node_edge_coords<T>(a, vol, L.iso, (uint32_t)cx + (o & 1u), (uint32_t)(cy + (int)((o >> 1) & 1u)),
                                    (uint32_t)(cz + (int)((o >> 2) & 1u)), o >> 3, p[j]);
            }
            face_normal(p[0], p[1], p[2], fn);
            n[0] = n[0] + fn[0];
            n[1] = n[1] + fn[1];
            n[2] = n[2] + fn[2];
        }
    }
}

/* The same sum from the face normals the triangle phase kept in LDS for cell
 * slices zv-1 and zv (band rows and the halo row above the band).  Every cell
 * around a crossing edge is active; its record, at its active-cell rank (the
 * active-cell prefix of its (cell row, chunk) + the active cells before it
 * in the chunk), holds its cube index and the rank of its first triangle. */
template <int M>
__device__ void vertex_normal_ring(const JobArgs& a, const Team& v, int y0, uint32_t x, uint32_t y, int zv,
                                   uint32_t axis, float n[3])
{
    n[0] = 0.0f; n[1] = 0.0f; n[2] = 0.0f;
    const uint32_t around = around_nibbles(axis);
    for (int q = 0; q < 4; ++q) {
        int cx, cy, cz;
        around_cell(axis, q, (int)x, (int)y, zv, cx, cy, cz);
        if (cx < 0 || cy < 0 || cz < 0) continue;
        if (cx + 1 >= (int)a.nx || cy + 1 >= (int)a.ny || cz + 1 >= (int)a.nz) continue;
        const uint32_t cs = (uint32_t)cz & 1u, j = (uint32_t)(cy - y0 + 1);
        const uint32_t c = (uint32_t)cx >> 6, i = (uint32_t)cx & 63u;
        const uint32_t d = cidx(v, cs, j, c);
        const uint32_t rank = v.ap[d] + (uint32_t)__popcll(v.act[d] & ((1ull << i) - 1ull));
        const uint32_t rc = v.crec[cs * v.fn_cap + rank];
        const uint32_t base = rc >> 8;
        const uint64_t w = v.tri[rc & 0xFFu];
        const int nt = (int)(w >> 60);
        const uint32_t me = (around >> (4 * q)) & 0xFu;
        const float* f = v.fn + (size_t)cs * v.fn_cap * 3u;
        for (int t = 0; t < nt; ++t) {
            const uint32_t e0 = (uint32_t)((w >> (12 * t)) & 0xF);
            const uint32_t e1 = (uint32_t)((w >> (12 * t + 8)) & 0xF);
            const uint32_t e2 = (uint32_t)((w >> (12 * t + 4)) & 0xF);
            if (e0 != me && e1 != me && e2 != me) continue;
            const float* g = f + 3u * (base + (uint32_t)t);
            n[0] = n[0] + g[0];
            n[1] = n[1] + g[1];
            n[2] = n[2] + g[2];
        }
    }
}

/* Emit tile (z, band): the vertices of window slice ws_v (shared-vertex
 * modes, when out_vert) and the triangles of cell slice z-1 (when out_tri),
 * through two LDS work queues whose entry index is the output rank.  With
 * VertexNormal the triangles of the halo row are processed too and, when
 * keep_fn, every face normal is kept in the cell slice's ring slot; ring_ok:
 * both slots hold all their face normals. */
template <typename T, int M>
__device__ void emit_tile(const JobArgs& a, const Team& v, const Level& L, const Bases& B, const T* vol, int z,
                          int y0, uint32_t ws_v, bool out_vert, bool out_tri, const TileTotals& tt, bool keep_fn,
                          bool ring_ok)
{
    constexpr bool vn = M == kModeVertex;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t sub = lane >> a.cshift, cl = lane & ((1u << a.cshift) - 1u);
    const uint32_t G = 64u >> a.cshift, CS = 1u << a.cshift;
    const uint32_t nv_all = out_vert ? tt.nv : 0u;
    const uint32_t nt_all = tt.nt;
    const uint32_t Q = a.queue_cap;
    const int zc = z - 1;
    const int zv = z - 1 + (int)ws_v;
    const uint32_t cs = cslot<M>(zc);

    /* triangles of cell slice z-1 (with VertexNormal first, so that their face
     * normals are in LDS before the vertices below read them) */
    for (uint32_t w0 = 0; w0 < nt_all; w0 += Q) {
        for (uint32_t j = sub; j <= a.R; j += G) {
            for (uint32_t c = cl; c < a.NC; c += CS) {
                uint64_t m = v.act[cidx(v, cs, j, c)];
                if (!m) continue;
                const Corners k = corners_at<M>(a, v, y0, c, y0 - 1 + (int)j, zc);
                uint32_t r = v.pt[cidx(v, cs, j, c)];
                /* with the face normals kept, each active cell also records its first
                 * triangle and cube index at its active-cell rank */
                const bool rec = vn && keep_fn && w0 == 0;
                uint32_t ai = rec ? v.ap[cidx(v, cs, j, c)] : 0u;
                while (m) {
                    const uint32_t i = (uint32_t)__ffsll((unsigned long long)m) - 1u;
                    m &= m - 1ull;
                    const uint32_t ci = corner_index(k, i);
                    const uint32_t nt = (uint32_t)(v.tri[ci] >> 60);
                    const uint32_t x = c * 64u + i;
                    if (rec) v.crec[cs * v.fn_cap + ai++] = r << 8 | ci;
                    for (uint32_t t = 0; t < nt; ++t, ++r)
                        if (r >= w0 && r - w0 < Q) v.tq[r - w0] = x | (j << 20) | (t << 25);
                }
            }
        }
        wave_sync();
        const uint32_t ntq = (nt_all - w0) < Q ? (nt_all - w0) : Q;
        for (uint32_t q = lane; q < ntq; q += 64u) {
            const uint32_t e = v.tq[q];
            const uint32_t x = e & 0xFFFFFu, j = (e >> 20) & 0x1Fu, t = (e >> 25) & 7u;
            const uint32_t rank = w0 + q;                           // rank in the tile, halo row included
            const uint64_t tri = B.t + (uint64_t)(rank - tt.halo);
            const bool write = out_tri && j >= 1 && tri < L.cap_t;
            if (!write && !keep_fn) continue;
            const int y = y0 - 1 + (int)j;
            const uint32_t ci = cube_index_at<M>(v, y0, x, y, zc);
            const uint64_t w = v.tri[ci];
            const uint32_t eds[3] = { (uint32_t)((w >> (12 * t)) & 0xF), (uint32_t)((w >> (12 * t + 8)) & 0xF),
                                      (uint32_t)((w >> (12 * t + 4)) & 0xF) };
            float p[3][3], nn[3];
            if constexpr (M == kModeDup) {
                /* extract_surfaces_with_duplication: table orientation (:206-228) */
#pragma unroll
                for (int jj = 0; jj < 3; ++jj) {
                    const uint32_t qq = edge_ends(eds[jj]);
                    interpolate<T>(a, vol, L.iso,
                                   x + (qq & 1u), (uint32_t)y + ((qq >> 1) & 1u), (uint32_t)zc + ((qq >> 2) & 1u),
                                   x + ((qq >> 4) & 1u), (uint32_t)y + ((qq >> 5) & 1u),
                                   (uint32_t)zc + ((qq >> 6) & 1u), p[jj]);
                }
                face_normal(p[0], p[1], p[2], nn);
                float* co = L.coords + 9 * tri;
#pragma unroll
                for (int jj = 0; jj < 3; ++jj) {
                    co[3 * jj + 0] = p[jj][0];
                    co[3 * jj + 1] = p[jj][1];
                    co[3 * jj + 2] = p[jj][2];
                }
                L.normals[3 * tri + 0] = nn[0];
                L.normals[3 * tri + 1] = nn[1];
                L.normals[3 * tri + 2] = nn[2];
                continue;
            }
            if (write) {
                /* connect_isopoints: vertex ids of the three edges (:493-514) */
                uint32_t id[3];
#pragma unroll
                for (int jj = 0; jj < 3; ++jj) {
                    const uint32_t o = edge_owner(eds[jj]);
                    const uint32_t dx = o & 1u, dy = (o >> 1) & 1u, dz = (o >> 2) & 1u, ax = o >> 3;
                    const uint32_t kk = j - 1 + dy;
                    const uint32_t rk = tile_rank<M>(a, v, z, y0, x + dx, kk, zc + (int)dz, ax);
                    const uint32_t sel = (dz << 1) | (kk == a.R ? 1u : 0u);
                    const uint64_t b = sel == 0 ? B.v[0] : sel == 1 ? B.v[1] : sel == 2 ? B.v[2] : B.v[3];
                    id[jj] = (uint32_t)(b + rk);
                }
                L.conn[3 * tri + 0] = id[0];
                L.conn[3 * tri + 1] = id[1];
                L.conn[3 * tri + 2] = id[2];
            }
            if (vn && !keep_fn) continue;
            /* calculate_normals_on_polygon (:531-559); with VertexNormal the same face
             * normal is what calculate_normals_on_vertex sums */
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) {
                const uint32_t o = edge_owner(eds[jj]);
                node_edge_coords<T>(a, vol, L.iso, x + (o & 1u), (uint32_t)(y + (int)((o >> 1) & 1u)),
                                    (uint32_t)(zc + (int)((o >> 2) & 1u)), o >> 3, p[jj]);
            }
            face_normal(p[0], p[1], p[2], nn);
            if constexpr (vn) {
                float* g = v.fn + ((size_t)cs * v.fn_cap + rank) * 3u;
                g[0] = nn[0]; g[1] = nn[1]; g[2] = nn[2];
            } else {
                L.normals[3 * tri + 0] = nn[0];
                L.normals[3 * tri + 1] = nn[1];
                L.normals[3 * tri + 2] = nn[2];
            }
        }
        wave_sync();
    }

    /* vertices of window slice ws_v, band rows */
    for (uint32_t w0 = 0; w0 < nv_all; w0 += Q) {
        const uint32_t start = wrow(v, z, ws_v, 1)[0];
        for (uint32_t k = sub; k < a.R && y0 + (int)k < (int)a.ny; k += G) {
            for (uint32_t c = cl; c < a.NC; c += CS) {
                const Cross e = window_cross<M>(a, v, z, y0, ws_v, k + 1, c);
                uint64_t m = e.x | e.y | e.z;
                uint32_t r = wrow(v, z, ws_v, k + 1)[c] - start;
                while (m) {
                    const uint32_t i = (uint32_t)__ffsll((unsigned long long)m) - 1u;
                    m &= m - 1ull;
                    const uint32_t base = (c * 64u + i) | (k << 20);
                    if ((e.x >> i) & 1ull) { if (r >= w0 && r - w0 < Q) v.vq[r - w0] = base; ++r; }
                    if ((e.y >> i) & 1ull) { if (r >= w0 && r - w0 < Q) v.vq[r - w0] = base | (1u << 26); ++r; }
                    if ((e.z >> i) & 1ull) { if (r >= w0 && r - w0 < Q) v.vq[r - w0] = base | (2u << 26); ++r; }
                }
            }
        }
        wave_sync();
        const uint32_t nvq = (nv_all - w0) < Q ? (nv_all - w0) : Q;
        for (uint32_t q = lane; q < nvq; q += 64u) {
            const uint32_t e = v.vq[q];
            const uint32_t x = e & 0xFFFFFu, k = (e >> 20) & 0x3Fu, axis = (e >> 26) & 3u;
            const uint64_t vid = (ws_v == 1 ? B.v[2] : B.v[0]) + (uint64_t)(w0 + q);
            if (vid >= L.cap_v) continue;
            const uint32_t y = (uint32_t)y0 + k;
            float p[3];
            node_edge_coords<T>(a, vol, L.iso, x, y, (uint32_t)zv, axis, p);
            L.coords[3 * vid + 0] = p[0];
            L.coords[3 * vid + 1] = p[1];
            L.coords[3 * vid + 2] = p[2];
            if constexpr (vn) {
                float nn[3];
                if (ring_ok) vertex_normal_ring<M>(a, v, y0, x, y, zv, axis, nn);
                else vertex_normal_direct<T>(a, v, L, vol, x, y, zv, axis, nn);
                L.normals[3 * vid + 0] = nn[0];
                L.normals[3 * vid + 1] = nn[1];
                L.normals[3 * vid + 2] = nn[2];
            }
        }
        wave_sync();
    }
}

/* Stage the case table in the workgroup's LDS (shared by its teams). */
__device__ void stage_table(uint8_t* smem)
{
    uint64_t* tri = (uint64_t*)smem;
    for (uint32_t i = threadIdx.x; i < 256; i += kEmitBlock) tri[i] = d_tri[i];
    __syncthreads();
}

/* ------------------------------------------------------------------------- */
/* classify pass: values -> bit volume                                         */
/* ------------------------------------------------------------------------- */

/* Predicate of level l for a value (double(v) > iso[l]). */
template <typename T>
__device__ __forceinline__ bool above(T v, const JobArgs& a, int l) { return (double)v > a.iso[l]; }
template <>
__device__ __forceinline__ bool above<float>(float v, const JobArgs& a, int l) { return v > a.thr[l]; }

/* Classification of node row `row` of a job (the single-volume pass). */
template <typename T, int NLEV>
__device__ __forceinline__ void classify_row(const JobArgs& a, uint64_t row)
{
    const uint32_t lane = threadIdx.x & 63u;
    const T* src = (const T*)a.values + row * a.nx;
    const uint64_t z = row / a.ny, y = row - z * a.ny;
    constexpr uint32_t kBatch = 16;
    for (uint32_t c0 = 0; c0 < a.NC; c0 += 64) {
        uint32_t lo[NLEV], hi[NLEV];
#pragma unroll
        for (int l = 0; l < NLEV; ++l) { lo[l] = 0; hi[l] = 0; }
        const uint32_t cend = (a.NC - c0) < 64u ? (a.NC - c0) : 64u;
        for (uint32_t cb = 0; cb < cend; cb += kBatch) {
            T val[kBatch];
#pragma unroll
            for (uint32_t i = 0; i < kBatch; ++i) {
                const uint32_t x = (c0 + cb + i) * 64u + lane;
                val[i] = (cb + i < cend && x < a.nx) ? __builtin_nontemporal_load(src + x) : T(0);   // read once
            }
#pragma unroll
            for (uint32_t i = 0; i < kBatch; ++i) {
                if (cb + i < cend) {
                    const bool in = (c0 + cb + i) * 64u + lane < a.nx;
                    const bool mine = lane == cb + i;
#pragma unroll
                    for (int l = 0; l < NLEV; ++l) {
                        const uint64_t m = __ballot(in && above<T>(val[i], a, l));
                        lo[l] = mine ? (uint32_t)m : lo[l];
                        hi[l] = mine ? (uint32_t)(m >> 32) : hi[l];
                    }
                }
            }
        }
        /* the 8-level instance also serves passes of 5..7 levels: store only
         * the pass's own levels.  Nontemporal stores: the masks' writes,
         * mixed into the volume's read stream, cost more than their bytes
         * (measured: 384 MiB of mask stores add ~180 us to a 4 GiB read, 25 us
         * less with the nontemporal hint) */
        if (lane < cend) {
#pragma unroll
            for (int l = 0; l < NLEV; ++l)
                if (NLEV < 8 || l < (int)a.nlev)
                    __builtin_nontemporal_store(((uint64_t)hi[l] << 32) | lo[l],
                                                &a.bitvol[(((uint64_t)l * a.nz + z) * a.ny + y) * a.NC + c0 + lane]);
        }
    }
}

/* Classification of chunks [q0, q1) of a job, q = row * NC + chunk: the
 * chunks of consecutive rows follow each other, so short rows share a wave's
 * batches of loads and its windows of 64 parked masks, which leave with one
 * coalesced store per level (the masks of consecutive rows are contiguous). */
template <typename T, int NLEV>
__device__ __forceinline__ void classify_chunks(const JobArgs& a, uint64_t q0, uint64_t q1)
{
    const uint32_t lane = threadIdx.x & 63u;
    const T* vals = (const T*)a.values;
    const uint64_t level_words = (uint64_t)a.nz * a.ny * a.NC;
    constexpr uint32_t kBatch = 16;
    for (uint64_t w0 = q0; w0 < q1; w0 += 64) {
        const uint32_t wn = (q1 - w0) < 64 ? (uint32_t)(q1 - w0) : 64u;
        uint32_t lo[NLEV], hi[NLEV];
#pragma unroll
        for (int l = 0; l < NLEV; ++l) { lo[l] = 0; hi[l] = 0; }
        uint64_t row = w0 / a.NC;
        uint32_t c = (uint32_t)(w0 - row * a.NC);
        for (uint32_t cb = 0; cb < wn; cb += kBatch) {
            T val[kBatch];
            uint32_t cs[kBatch];
#pragma unroll
            for (uint32_t i = 0; i < kBatch; ++i) {
                cs[i] = c;
                const uint32_t x = c * 64u + lane;
                val[i] = (cb + i < wn && x < a.nx) ? vals[row * a.nx + x] : T(0);
                if (++c == a.NC) { c = 0; ++row; }
            }
#pragma unroll
            for (uint32_t i = 0; i < kBatch; ++i) {
                if (cb + i < wn) {
                    const bool in = cs[i] * 64u + lane < a.nx;
                    const bool mine = lane == cb + i;
#pragma unroll
                    for (int l = 0; l < NLEV; ++l) {
                        const uint64_t m = __ballot(in && above<T>(val[i], a, l));
                        lo[l] = mine ? (uint32_t)m : lo[l];
                        hi[l] = mine ? (uint32_t)(m >> 32) : hi[l];
                    }
                }
            }
        }
        /* the 8-level instance also serves passes of 5..7 levels: store only
         * the pass's own levels */
        if (lane < wn) {
#pragma unroll
            for (int l = 0; l < NLEV; ++l)
                if (NLEV < 8 || l < (int)a.nlev)
                    a.bitvol[(uint64_t)l * level_words + w0 + lane] = ((uint64_t)hi[l] << 32) | lo[l];
        }
    }
}

/* Streams the volume once.  Each wave owns whole rows: its lanes load 64
 * consecutive nodes per chunk (one coalesced access), the comparison against
 * level l is a ballot (v_cmp into a lane mask), and the mask is parked in
 * lane c of a register by a lane-select; the masks of up to 64 chunks of a
 * row leave with one coalesced 64-bit store per level.  NLEV is the number of
 * levels, a compile-time constant so that nothing is guarded per level. */
template <typename T, int NLEV>
__global__ __launch_bounds__(kBlock) void k_classify(JobArgs a)
{
    const uint64_t nrows = (uint64_t)a.ny * a.nz;
    const uint64_t wave_id = (uint64_t)blockIdx.x * kWaves + uniform(threadIdx.x >> 6);   // uniform: scalar
    const uint64_t nwaves = ((uint64_t)gridDim.x * kBlock) >> 6;
    for (uint64_t row = wave_id; row < nrows; row += nwaves) classify_row<T, NLEV>(a, row);
}

/* Node q of a dword of 8- or 16-bit values (little-endian: the first node
 * in the low bits). */
template <typename T>
__device__ __forceinline__ T dword_node(uint32_t d, uint32_t q)
{
    if constexpr (sizeof(T) == 1) return (T)(uint8_t)(d >> (8u * q));
    else return (T)(uint16_t)(d >> (16u * q));
}

/* Classification of an 8- or 16-bit volume whose rows are whole dwords (nx
 * a multiple of P = 4 / sizeof(T), the values 4-byte aligned).  A lane loads
 * one dword, P nodes, so that a wave load moves 256 bytes as a float row's
 * does (one node a lane would move 64 or 128); the lane's P comparisons are
 * bits (lane mod LPC) * P + q of its chunk's mask word, which an OR over the
 * chunk's LPC = 64 / P lanes assembles.  Work items are (row, segment of 64
 * dwords), 16 consecutive ones per wave iteration, their loads in flight
 * together; the wave's item indices are scalars. */
/* Items [i0, i1) of a job in the packed classification, item = row * nseg +
 * segment: 16 consecutive ones at a time from i0 on, then every istep-th
 * batch of 16 (i0, i1 and istep the same on all of the wave's lanes). */
template <typename T, int NLEV>
__device__ __forceinline__ void classify_packed_items(const JobArgs& a, uint64_t i0, uint64_t i1, uint64_t istep)
{
    static_assert(sizeof(T) <= 2, "8- and 16-bit values");
    constexpr uint32_t P = 4 / sizeof(T), LPC = 64 / P;
    constexpr uint32_t kBatch = 16;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nd = a.nx / P;
    const uint32_t nseg = (nd + 63) / 64;
    const uint64_t level_words = (uint64_t)a.nz * a.ny * a.NC;
    const uint32_t* vals = (const uint32_t*)a.values;
    for (; i0 < i1; i0 += istep) {
        uint64_t row = i0 / nseg;
        uint32_t seg = (uint32_t)(i0 - row * nseg);
        uint32_t d[kBatch];
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i) {
            const uint32_t k = seg * 64u + lane;
            d[i] = (i0 + i < i1 && k < nd) ? __builtin_nontemporal_load(vals + row * nd + k) : 0u;   // read once
            if (++seg == nseg) { seg = 0; ++row; }
        }
        row = i0 / nseg;
        seg = (uint32_t)(i0 - row * nseg);
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i) {
            if (i0 + i < i1) {
                const bool valid = seg * 64u + lane < nd;
                const uint32_t c = seg * P + lane / LPC;
                const uint32_t sh = (lane % LPC) * P;
#pragma unroll
                for (int l = 0; l < NLEV; ++l) {
                    if (NLEV == 8 && l >= (int)a.nlev) break;   // the 8-level instance serves 5..7 levels
                    uint32_t bits = 0;
#pragma unroll
                    for (uint32_t q = 0; q < P; ++q) bits |= (valid && above<T>(dword_node<T>(d[i], q), a, l) ? 1u : 0u) << q;
                    /* OR over the chunk's lanes: each contributes its P bits */
                    uint32_t lo = sh < 32u ? bits << sh : 0u, hi = sh >= 32u ? bits << (sh - 32u) : 0u;
#pragma unroll
                    for (uint32_t m = 1; m < LPC; m <<= 1) {
                        lo |= (uint32_t)__shfl_xor((int)lo, (int)m);
                        hi |= (uint32_t)__shfl_xor((int)hi, (int)m);
                    }
                    if (lane % LPC == 0 && c < a.NC)
                        __builtin_nontemporal_store((uint64_t)hi << 32 | lo,
                                                    &a.bitvol[(uint64_t)l * level_words + row * a.NC + c]);
                }
            }
            if (++seg == nseg) { seg = 0; ++row; }
        }
    }
}

/* Whether job a's values load a dword of nodes per lane: 8- or 16-bit
 * values, rows of whole dwords, 4-byte aligned. */
template <typename T>
__device__ __host__ __forceinline__ bool packed_rows(const JobArgs& a)
{
    return sizeof(T) <= 2 && a.nx % (4 / sizeof(T)) == 0 && ((uintptr_t)a.values & 3u) == 0;
}

template <typename T, int NLEV>
__global__ __launch_bounds__(kBlock) void k_classify_packed(JobArgs a)
{
    const uint64_t nseg = (a.nx / (4 / sizeof(T)) + 63) / 64;
    const uint64_t items = (uint64_t)a.ny * a.nz * nseg;
    const uint64_t wave_id = (uint64_t)blockIdx.x * kWaves + uniform(threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * kWaves;
    classify_packed_items<T, NLEV>(a, wave_id * 16, items, nwaves * 16);
}

/* Index j with start[j] <= g < start[j+1] (start has n+1 ascending entries). */
__device__ __forceinline__ uint32_t find_job(const uint64_t* start, uint32_t n, uint64_t g)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (start[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

/* Classify pass of a batch of jobs sharing the value type T: the node rows of
 * all jobs form one sequence (row_start[j] = first row of job j), cut into one
 * contiguous range per wave so that the job changes rarely along a range; a
 * wave classifies its rows of a job as one run of chunks.  kWide: some job has
 * more than four levels (the eight-level instance's registers would otherwise
 * halve the waves per SIMD of every batch). */
template <typename T, bool kWide>
__global__ __launch_bounds__(kBlock) void k_classify_batch(const JobArgs* jobs, const uint64_t* row_start, uint32_t njobs)
{
    const uint64_t total = row_start[njobs];
    const uint64_t wave_id = (uint64_t)blockIdx.x * kWaves + uniform(threadIdx.x >> 6);   // uniform: scalar
    const uint64_t nwaves = ((uint64_t)gridDim.x * kBlock) >> 6;
    const uint64_t per = (total + nwaves - 1) / nwaves;
    uint64_t g = wave_id * per;
    const uint64_t end = g + per < total ? g + per : total;
    if (g >= end) return;
    uint32_t j = __builtin_amdgcn_readfirstlane(find_job(row_start, njobs, g));
    while (g < end) {
        while (g >= row_start[j + 1]) ++j;
        const JobArgs& a = jobs[j];
        const uint64_t je = row_start[j + 1] < end ? row_start[j + 1] : end;
        if constexpr (sizeof(T) <= 2) {
            if (packed_rows<T>(a)) {
                /* the wave's rows as (row, segment) items, 16 at a time */
                const uint64_t nseg = (a.nx / (4 / sizeof(T)) + 63) / 64;
                const uint64_t i0 = (g - row_start[j]) * nseg, i1 = (je - row_start[j]) * nseg;
                switch (a.nlev) {
                case 1: classify_packed_items<T, 1>(a, i0, i1, 16); break;
                case 2: classify_packed_items<T, 2>(a, i0, i1, 16); break;
                case 3: classify_packed_items<T, 3>(a, i0, i1, 16); break;
                case 4: classify_packed_items<T, 4>(a, i0, i1, 16); break;
                default:
                    if constexpr (kWide) classify_packed_items<T, 8>(a, i0, i1, 16);
                    break;
                }
                g = je;
                continue;
            }
        }
        const uint64_t q0 = (g - row_start[j]) * a.NC, q1 = (je - row_start[j]) * a.NC;
        switch (a.nlev) {
        case 1: classify_chunks<T, 1>(a, q0, q1); break;
        case 2: classify_chunks<T, 2>(a, q0, q1); break;
        case 3: classify_chunks<T, 3>(a, q0, q1); break;
        case 4: classify_chunks<T, 4>(a, q0, q1); break;
        default:
            if constexpr (kWide) classify_chunks<T, 8>(a, q0, q1);
            break;
        }
        g = je;
    }
}

/* ------------------------------------------------------------------------- */
/* count pass: per-row counts from the bit volume                              */
/* ------------------------------------------------------------------------- */

/* Crossing edges and active cells of chunk c of node row (y, z) of one level,
 * read straight from the bit volume `lev` (coalesced across the lanes of a
 * row, neighbours from L1/L2).  x/y/z: crossing edges owned by the row's
 * nodes; act: active cells of cell row (y, z-1); k0/k3/k4/k7: the masks of
 * the cells' four node rows (bit i of kq is corner q of cell x = 64c + i; the
 * x+1 corners 1, 2, 5, 6 are the next bits of the same masks); next: bit 0 of
 * the next chunk's word of those four rows (corners 1, 2, 5, 6 of cell 63). */
struct RowChunk {
    uint64_t x, y, z, act;
    uint64_t k0, k3, k4, k7;
    uint32_t next;
};

__device__ __forceinline__ void row_chunk(const JobArgs& a, const uint64_t* lev, uint32_t y, uint32_t z, uint32_t c,
                                          RowChunk& o)
{
    const uint64_t slice_words = (uint64_t)a.ny * a.NC;
    const bool has_up = y + 1 < a.ny, has_next = z + 1 < a.nz, has_cells = z >= 1 && has_up;
    /* Every load is issued (no branch around it): a neighbour row outside the
     * volume is read as the row itself, which makes its crossings (b ^ u,
     * b ^ n0) zero, and the cells are masked out below.  The next chunk of a
     * row's last chunk is read as the chunk itself: its bit 0 lands only on
     * bit 63 of the shifted words and of `next`, i.e. on node / cell 64c + 63,
     * which lies at or past nx - 1 in a row's last chunk, so valid_mask(c,
     * nx - 1) clears it from x and act (and corner_pairs reads `next` only for
     * an active cell 63). */
    const uint64_t* row = lev + (uint64_t)z * slice_words + (uint64_t)y * a.NC;
    const uint64_t* up = has_up ? row + a.NC : row;
    const uint64_t* nz = has_next ? row + slice_words : row;
    const uint64_t* pr = has_cells ? row - slice_words : row;
    const uint64_t* pu = has_cells ? pr + a.NC : row;
    const uint32_t cn = c + 1 < a.NC ? c + 1 : c;
    const uint64_t b = row[c], b1 = row[cn];
    const uint64_t u = up[c], u1 = up[cn];
    const uint64_t n0 = nz[c];
    const uint64_t p = pr[c], p1 = pr[cn];
    const uint64_t q = pu[c], q1 = pu[cn];
    const uint64_t c4 = b, c5 = (b >> 1) | (b1 << 63);
    const uint64_t c7 = u, c6 = (u >> 1) | (u1 << 63);
    const uint64_t c1 = (p >> 1) | (p1 << 63), c2 = (q >> 1) | (q1 << 63);
    const uint64_t vx = valid_mask(c, a.nx - 1), vn = valid_mask(c, a.nx);
    o.x = (b ^ c5) & vx;
    o.y = (b ^ u) & vn;
    o.z = (b ^ n0) & vn;
    const uint64_t or_ = p | c1 | c2 | q | c4 | c5 | c6 | c7;
    const uint64_t an = p & c1 & c2 & q & c4 & c5 & c6 & c7;
    o.act = has_cells ? (or_ & ~an) & vx : 0ull;
    o.k0 = p;
    o.k3 = q;
    o.k4 = c4;
    o.k7 = c7;
    o.next = (uint32_t)(p1 & 1ull) | (uint32_t)(q1 & 1ull) << 1 | (uint32_t)(b1 & 1ull) << 2 |
             (uint32_t)(u1 & 1ull) << 3;
}

/* (y, z) of node row g + sub of a wave that walks row groups g = g0, g0 +
 * step, ... (g0 and step are the same on all of the wave's lanes): the
 * group's first row advances by the step's quotient and remainder by ny,
 * instead of a division of every row index. */
struct RowWalk { uint32_t yg, zg, qs, rs; };

__device__ __forceinline__ RowWalk row_walk(const JobArgs& a, uint64_t g0, uint64_t step)
{
    RowWalk w;
    const uint64_t g = uniform64(g0), s = uniform64(step);
    w.zg = (uint32_t)(g / a.ny);
    w.yg = (uint32_t)(g - (uint64_t)w.zg * a.ny);
    w.qs = (uint32_t)(s / a.ny);
    w.rs = (uint32_t)(s - (uint64_t)w.qs * a.ny);
    return w;
}

__device__ __forceinline__ void row_at(const JobArgs& a, const RowWalk& w, uint32_t sub, uint32_t& y, uint32_t& z)
{
    y = w.yg + sub;
    z = w.zg;
    if (y >= a.ny) {
        /* sub < 64: one wrap when ny >= 64 */
        if (a.ny >= 64u) {
            y -= a.ny;
            ++z;
        } else {
            z += y / a.ny;
            y %= a.ny;
        }
    }
}

__device__ __forceinline__ void row_next(const JobArgs& a, RowWalk& w)
{
    w.yg += w.rs;
    w.zg += w.qs;
    if (w.yg >= a.ny) {
        w.yg -= a.ny;
        ++w.zg;
    }
}

/* The cube index with bits 2, 3 and bits 6, 7 swapped (an involution). */
__device__ __forceinline__ uint32_t swap_pairs(uint32_t c)
{
    return (c & 0x33u) | ((c & 0x44u) << 1) | ((c & 0x88u) >> 1);
}

/* swap_pairs(cube index) of cell i of a chunk: the corners (x, x+1) of each of
 * the cell's four node rows are two adjacent bits of that row's mask (k0, k3,
 * k4, k7), taken as one 2-bit window; for cell 63 the x+1 bits come from the
 * next chunk (o.next). */
__device__ __forceinline__ uint32_t corner_pairs(const RowChunk& o, uint32_t i)
{
    if (i == 63u) {
        return (uint32_t)(o.k0 >> 63) | (o.next & 1u) << 1 | (uint32_t)(o.k3 >> 63) << 2 |
               ((o.next >> 1) & 1u) << 3 | (uint32_t)(o.k4 >> 63) << 4 | ((o.next >> 2) & 1u) << 5 |
               (uint32_t)(o.k7 >> 63) << 6 | ((o.next >> 3) & 1u) << 7;
    }
    const uint32_t p0 = (uint32_t)(o.k0 >> i) & 3u, p3 = (uint32_t)(o.k3 >> i) & 3u;
    const uint32_t p4 = (uint32_t)(o.k4 >> i) & 3u, p7 = (uint32_t)(o.k7 >> i) & 3u;
    return p0 | p3 << 2 | p4 << 4 | p7 << 6;
}

/* x of the lane d below within its DPP row of 16 lanes (0 for the row's
 * first d lanes): row_shr:d. */
template <int D>
__device__ __forceinline__ uint32_t row_shr(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x110 + D, 0xf, 0xf, true);
}

/* Exclusive scan of x over the CS lanes of each row group (CS a power of two;
 * group cl = lane mod CS).  Groups of up to 16 lanes lie inside one DPP row:
 * the steps are row shifts, with no LDS round trip. */
__device__ __forceinline__ uint32_t group_excl_scan(uint32_t x, uint32_t cl, uint32_t CS, uint32_t& total)
{
    uint32_t inc = x;
    if (CS <= 16) {
        if (CS > 1) { const uint32_t t = row_shr<1>(inc); inc += cl >= 1 ? t : 0u; }
        if (CS > 2) { const uint32_t t = row_shr<2>(inc); inc += cl >= 2 ? t : 0u; }
        if (CS > 4) { const uint32_t t = row_shr<4>(inc); inc += cl >= 4 ? t : 0u; }
        if (CS > 8) { const uint32_t t = row_shr<8>(inc); inc += cl >= 8 ? t : 0u; }
    } else {
        for (uint32_t d = 1; d < CS; d <<= 1) {
            const uint32_t t = __shfl_up(inc, d, (int)CS);
            if (cl >= d) inc += t;
        }
    }
    total = __shfl(inc, (int)CS - 1, (int)CS);
    return inc - x;
}

/* Counts of node rows [r0, r1) of a job: one row group of 64 >> cshift rows
 * per wave iteration, one lane per 64-node chunk, no LDS window and no
 * barrier.  Per row: the crossing edges its nodes own and the triangles of
 * cell row (y, z-1); with a.vpc / a.tpc also each chunk's exclusive prefix of
 * both within its row.  ntri: triangle count per cube index (LDS). */
template <int NLEV>
__device__ __forceinline__ void count_rows_t(const JobArgs& a, const uint8_t* ntri, uint64_t r0, uint64_t r1,
                                             uint64_t wave_first, uint64_t wave_step)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t sub = lane >> a.cshift, cl = lane & ((1u << a.cshift) - 1u);
    const uint32_t G = 64u >> a.cshift, CS = 1u << a.cshift;
    const uint64_t level_words = (uint64_t)a.nz * a.ny * a.NC;
    const bool chunk_prefixes = a.vpc != nullptr;
    RowWalk rw = row_walk(a, r0 + wave_first * G, wave_step * G);
    for (uint64_t g = r0 + wave_first * G; g < r1; g += wave_step * G, row_next(a, rw)) {
        const uint64_t r = g + sub;
        const bool ok = r < r1;
        uint32_t y, z;
        row_at(a, rw, sub, y, z);
        const uint32_t nl = NLEV ? (uint32_t)NLEV : a.nlev;
#pragma unroll
        for (uint32_t l = 0; l < nl; ++l) {
            const uint64_t* lev = a.bitvol + (uint64_t)l * level_words;
            uint32_t vc = 0, tc = 0;
            for (uint32_t c0 = 0; c0 < a.NC; c0 += CS) {
                const uint32_t c = c0 + cl;
                uint32_t cv = 0, ct = 0;
                if (ok && c < a.NC) {
                    RowChunk o;
                    row_chunk(a, lev, y, z, c, o);
                    cv = (uint32_t)(__popcll(o.x) + __popcll(o.y) + __popcll(o.z));
                    uint64_t m = o.act;
                    while (m) {
                        const uint32_t i = (uint32_t)__ffsll((unsigned long long)m) - 1u;
                        m &= m - 1ull;
                        ct += ntri[corner_pairs(o, i)];
                    }
                }
                if (chunk_prefixes) {
                    /* both counts in one scan: a chunk holds at most 192
                     * crossing edges and 320 triangles, a group of at most 64
                     * chunks sums below 2^16 in each half */
                    uint32_t s;
                    const uint32_t e = group_excl_scan(cv | ct << 16, cl, CS, s);
                    if (ok && c < a.NC) {
                        const uint64_t w = (uint64_t)l * level_words + r * a.NC + c;
                        /* coalesced stream of prefixes: nontemporal (measured
                         * faster).  Rows of at most 64 chunks are one group,
                         * so the packed scan result is the chunk's in-row
                         * prefix pair (chunk_prefix) */
                        if (a.NC <= 64) {
                            __builtin_nontemporal_store(e, &a.vpc[w]);
                        } else {
                            __builtin_nontemporal_store(vc + (e & 0xFFFFu), &a.vpc[w]);
                            __builtin_nontemporal_store(tc + (e >> 16), &a.tpc[w]);
                        }
                    }
                    vc += s & 0xFFFFu;
                    tc += s >> 16;
                } else {
                    vc += cv;
                    tc += ct;
                }
            }
            if (!chunk_prefixes) {
                /* sum over the CS lanes of each row group */
                for (uint32_t d = CS >> 1; d >= 1; d >>= 1) {
                    vc += __shfl_xor(vc, (int)d);
                    tc += __shfl_xor(tc, (int)d);
                }
            }
            if (ok && cl == 0) {
                a.counts[(uint64_t)(l * 2 + 0) * a.nrows + r] = vc;
                a.counts[(uint64_t)(l * 2 + 1) * a.nrows + r] = tc;
            }
        }
    }
}

/* Level count as a compile-time constant where it is small, so that the loads
 * of all levels can be in flight together. */
__device__ __forceinline__ void count_rows(const JobArgs& a, const uint8_t* ntri, uint64_t r0, uint64_t r1,
                                           uint64_t wave_first, uint64_t wave_step)
{
    switch (a.nlev) {
    case 1: count_rows_t<1>(a, ntri, r0, r1, wave_first, wave_step); break;
    case 2: count_rows_t<2>(a, ntri, r0, r1, wave_first, wave_step); break;
    case 3: count_rows_t<3>(a, ntri, r0, r1, wave_first, wave_step); break;
    case 4: count_rows_t<4>(a, ntri, r0, r1, wave_first, wave_step); break;
    default: count_rows_t<0>(a, ntri, r0, r1, wave_first, wave_step); break;
    }
}

/* Triangle count per swap_pairs(cube index) (the index corner_pairs gives). */
__device__ __forceinline__ void stage_ntri(uint8_t* ntri)
{
    for (uint32_t i = threadIdx.x; i < 256; i += kBlock) ntri[i] = (uint8_t)(d_tri[swap_pairs(i)] >> 60);
    __syncthreads();
}

/* seven waves per SIMD (at most 72 VGPRs): the pass is bound by VALU issue
 * and load latency, and at eight waves (64 VGPRs) its spills cost more
 * issue than the eighth wave hides */
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_count(JobArgs a)
{
    __shared__ uint8_t ntri[256];
    stage_ntri(ntri);
    /* row groups of 64 >> cshift rows as the units */
    const uint64_t G = 64u >> a.cshift, groups = (a.nrows + G - 1) / G;
    const Span sp = grid_span(0, groups, kWaves, uniform(threadIdx.x >> 6));
    const uint64_t r1 = sp.end * G < a.nrows ? sp.end * G : a.nrows;
    count_rows(a, ntri, sp.lo * G, r1, sp.first - sp.lo, sp.step);
}

/* Count pass of a batch: the node rows of all jobs form one sequence, cut
 * into one contiguous range per wave. */
__global__ __launch_bounds__(kBlock) void k_count_batch(const JobArgs* jobs, const uint64_t* row_start, uint32_t njobs)
{
    __shared__ uint8_t ntri[256];
    stage_ntri(ntri);
    const uint64_t total = row_start[njobs];
    const uint64_t wave_id = (uint64_t)blockIdx.x * kWaves + uniform(threadIdx.x >> 6);   // uniform: scalar
    const uint64_t nwaves = ((uint64_t)gridDim.x * kBlock) >> 6;
    const uint64_t per = (total + nwaves - 1) / nwaves;
    uint64_t g = wave_id * per;
    const uint64_t end = g + per < total ? g + per : total;
    while (g < end) {
        const uint32_t j = __builtin_amdgcn_readfirstlane(find_job(row_start, njobs, g));
        const uint64_t je = row_start[j + 1] < end ? row_start[j + 1] : end;
        count_rows(jobs[j], ntri, g - row_start[j], je - row_start[j], 0, 1);
        g = je;
    }
}

/* ------------------------------------------------------------------------- */
/* tile-team emit pass                                                         */
/* ------------------------------------------------------------------------- */

/* One unit: level l of band b over steps zs .. ze-1 (with VertexNormal one
 * more step after the last node slice).  Tile z emits the triangles of cell
 * slice z-1 and the vertices of node slice z, or of node slice z-1 with
 * VertexNormal: their normals sum face normals of cell slices z-2 and z-1,
 * both then in the face-normal ring (a first unit step without output fills
 * in z-2).  The mask ring holds node slices z-1 .. z+1 and takes one new
 * slice per step. */
template <typename T, int M>
__device__ void emit_unit(const JobArgs& a, const Team& v, uint64_t unit)
{
    const T* vol = (const T*)a.values;
    const uint64_t per_level = (uint64_t)a.nb * a.nslabs;
    const uint32_t l = (uint32_t)(unit / per_level);
    const uint64_t rem = unit - (uint64_t)l * per_level;
    const uint32_t slab = (uint32_t)(rem / a.nb);
    const uint32_t band = (uint32_t)(rem - (uint64_t)slab * a.nb);
    const int y0 = (int)(band * a.R);
    const int zs = (int)(slab * a.slab);
    const int ze = min(zs + (int)a.slab, (int)a.nz);
    constexpr bool vn = M == kModeVertex;
    const int zend = (vn && ze == (int)a.nz) ? ze + 1 : ze;
    const uint32_t ws_v = vn ? 0u : 1u;
    const Level L = level_of(a, l);

    /* VertexNormal first prepares cell slice zs-2 (node slices zs-2, zs-1) */
    load_slices<M>(a, v, L, vn ? zs - 2 : zs - 1, vn ? zs : zs + 1, y0);
    uint32_t ovf = 0;   // bit cs: the face normals of cell slot cs did not fit
    wave_sync();
    if (vn && zs >= 2) {
        const TileTotals t = tile_prefixes<M>(a, v, zs - 1, y0, ws_v, false, true);
        const bool keep = t.nt <= v.fn_cap;
        ovf = keep ? 0u : (1u << cslot<M>(zs - 2));
        Bases B;
        B.v[0] = B.v[1] = B.v[2] = B.v[3] = 0;
        B.t = 0;
        emit_tile<T, M>(a, v, L, B, vol, zs - 1, y0, ws_v, false, false, t, keep, false);
        wave_sync();
    }
    for (int z = zs; z < zend; ++z) {
        if (vn || z > zs) load_slices<M>(a, v, L, z + 1, z + 1, y0);
        const Bases B = tile_bases(a, L, z, y0);
        wave_sync();
        const TileTotals t = tile_prefixes<M>(a, v, z, y0, ws_v, true, z == zs);
        bool keep = false;
        if (vn) {
            const uint32_t bit = 1u << cslot<M>(z - 1);
            keep = t.nt <= v.fn_cap;
            ovf = keep ? (ovf & ~bit) : (ovf | bit);
        }
        emit_tile<T, M>(a, v, L, B, vol, z, y0, ws_v, true, true, t, keep, ovf == 0);
        wave_sync();   // the next step overwrites a mask slot and the cell data
    }
}

/* Persistent workgroups of `teams` one-wave teams (the other waves only help
 * stage the table): team w of workgroup g takes units g * teams + w, then
 * strides by all teams of the grid; neighbouring teams take neighbouring
 * bands of the same slab and level. */
template <typename T, int M>
__global__ __launch_bounds__(kEmitBlock) void k_emit(JobArgs a)
{
    extern __shared__ __align__(16) uint8_t smem[];
    stage_table(smem);
    const uint32_t wave = uniform(threadIdx.x >> 6);
    if (wave >= a.teams) return;
    const Team v = make_team(a, smem, wave);
    const uint64_t step = (uint64_t)gridDim.x * a.teams;
    for (uint64_t u = (uint64_t)blockIdx.x * a.teams + wave; u < a.nunits; u += step) emit_unit<T, M>(a, v, u);
}

/* Emit pass of a batch of jobs sharing the value type T and the mode: the
 * units of all jobs in one sequence, taken by persistent teams.  Every job of
 * the launch has the same team_stride. */
template <typename T, int M>
__global__ __launch_bounds__(kEmitBlock) void k_emit_batch(const JobArgs* jobs, const uint64_t* unit_start, uint32_t njobs,
                                                            uint32_t teams)
{
    extern __shared__ __align__(16) uint8_t smem[];
    stage_table(smem);
    const uint32_t wave = uniform(threadIdx.x >> 6);
    if (wave >= teams) return;
    const uint64_t total = unit_start[njobs];
    const uint64_t step = (uint64_t)gridDim.x * teams;
    for (uint64_t g = (uint64_t)blockIdx.x * teams + wave; g < total; g += step) {
        const uint32_t j = uniform(find_job(unit_start, njobs, g));
        const JobArgs& a = jobs[j];
        const Team v = make_team(a, smem, wave);
        emit_unit<T, M>(a, v, g - unit_start[j]);
    }
}

/* ------------------------------------------------------------------------- */
/* flat emit: compaction records, then one thread per output item              */
/* ------------------------------------------------------------------------- */

/* Level of output item i given the per-level start offsets (nlev <= 8). */
__device__ __forceinline__ uint32_t level_at(const uint64_t* off, uint32_t nlev, uint64_t i)
{
    uint32_t l = 0;
    for (uint32_t k = 1; k < nlev; ++k) l += i >= off[k] ? 1u : 0u;
    return l;
}

/* In-row prefixes (vertices, triangles) of chunk word w.  Rows of at most 64
 * chunks keep both in one word of vpc (vertices in the low, triangles in the
 * high 16 bits: at most 63 chunks of 192 and 320 come before a chunk);
 * longer rows keep them in vpc and tpc. */
__device__ __forceinline__ void chunk_prefix(const JobArgs& a, uint64_t w, uint32_t& v, uint32_t& t)
{
    if (a.NC <= 64) {
        const uint32_t p = a.vpc[w];
        v = p & 0xFFFFu;
        t = p >> 16;
    } else {
        v = a.vpc[w];
        t = a.tpc[w];
    }
}

/* One non-empty chunk of the compaction, queued as y | z << 20 | level << 40
 * | chunk << 44. */
__device__ __forceinline__ void compact_chunk(const JobArgs& a, const uint8_t* ntri, uint64_t e)
{
    const uint32_t y = (uint32_t)(e & 0xFFFFFu), z = (uint32_t)((e >> 20) & 0xFFFFFu);
    const uint32_t l = (uint32_t)((e >> 40) & 0xFu), c = (uint32_t)(e >> 44);
    const uint64_t level_words = (uint64_t)a.nz * a.ny * a.NC;
    const uint64_t r = (uint64_t)z * a.ny + y;
    const uint64_t w = (uint64_t)l * level_words + r * a.NC + c;
    const uint64_t* pre_v = a.prefix + (uint64_t)(l * 2 + 0) * a.nrows;
    const uint64_t* pre_t = a.prefix + (uint64_t)(l * 2 + 1) * a.nrows;
    RowChunk o;
    row_chunk(a, a.bitvol + (uint64_t)l * level_words, y, z, c, o);
    uint32_t pv, pt;
    chunk_prefix(a, w, pv, pt);
    if (!a.dup) {
        uint64_t m = o.x | o.y | o.z;
        if (m) {
            ChunkCross cc;
            cc.base = (pre_v[r] - pre_v[0]) + pv;
            cc.x = o.x;
            cc.y = o.y;
            cc.z = o.z;
            a.cross[w] = cc;
            uint64_t id = a.voff[l] + cc.base;
            const uint64_t rowkey = (uint64_t)y << 20 | (uint64_t)z << 40;
            while (m) {
                const uint32_t i = (uint32_t)__ffsll((unsigned long long)m) - 1u;
                m &= m - 1ull;
                const uint64_t key = rowkey | (c * 64u + i);
                if ((o.x >> i) & 1ull) a.vkey[id++] = key;
                if ((o.y >> i) & 1ull) a.vkey[id++] = key | 1ull << 60;
                if ((o.z >> i) & 1ull) a.vkey[id++] = key | 2ull << 60;
            }
        }
    }
    uint64_t m = o.act;
    if (m) {
        uint64_t id = a.toff[l] + (pre_t[r] - pre_t[0]) + pt;
        const uint64_t cellkey = (uint64_t)y << 20 | (uint64_t)(z - 1) << 40;
        while (m) {
            const uint32_t i = (uint32_t)__ffsll((unsigned long long)m) - 1u;
            m &= m - 1ull;
            const uint32_t pi = corner_pairs(o, i);
            const uint32_t ci = swap_pairs(pi);
            const uint32_t nt = ntri[pi];
            const uint64_t key = cellkey | (c * 64u + i);
            for (uint32_t t = 0; t < nt; ++t, ++id) {
                a.trec[id] = key | (uint64_t)t << 60;
                a.tci[id] = (uint8_t)ci;
            }
        }
    }
}

/* Compaction of node rows [r0, r1): every crossing edge writes its key at its
 * vertex id and every triangle its cell and table index t (plus the cell's
 * cube index) at its triangle id -- row prefix + chunk prefix within the row
 * + rank within the chunk, i.e. vertices in node scan order x {x,y,z}-edge
 * and triangles in cell scan order x table order (calculate_isopoints
 * :365-437, connect_isopoints :439-521, extract_surfaces_with_duplication
 * :155-243); chunks with a vertex also leave their crossing record.  Lanes
 * first scan (row, chunk) pairs as the count pass does, NLEV levels at a time
 * (their loads in flight together; NLEV = 0: one level at a time, any level
 * count), and queue the non-empty chunks (the chunk prefixes tell) in the
 * wave's LDS queue q (kCompactQueue entries); the work on them runs 64 chunks
 * at a time, one per lane, taken from the queue's end.  Output positions come
 * from the prefixes alone, so the queue's order does not matter. */
constexpr uint32_t kCompactLevels = 4;
constexpr uint32_t kCompactQueue = 64 * (kCompactLevels + 1);

template <int NLEV>
__device__ __forceinline__ void compact_rows_t(const JobArgs& a, const uint8_t* ntri, uint64_t* q, uint64_t r0, uint64_t r1,
                                               uint64_t wave_first, uint64_t wave_step)
{
    constexpr uint32_t NL = NLEV ? (uint32_t)NLEV : 1u;
    static_assert(NL <= kCompactLevels, "queue holds 63 + 64 * NL entries");
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t sub = lane >> a.cshift, cl = lane & ((1u << a.cshift) - 1u);
    const uint32_t G = 64u >> a.cshift, CS = 1u << a.cshift;
    const uint64_t level_words = (uint64_t)a.nz * a.ny * a.NC;
    if (!a.trec) return;   // the records did not fit (batched form): nothing to write
    uint32_t qn = 0;
    RowWalk rw = row_walk(a, r0 + wave_first * G, wave_step * G);
    for (uint64_t g = r0 + wave_first * G; g < r1; g += wave_step * G, row_next(a, rw)) {
        const uint64_t r = g + sub;
        const bool ok = r < r1;
        uint32_t y, z;
        row_at(a, rw, sub, y, z);
        for (uint32_t l0 = 0; l0 < a.nlev; l0 += NL) {
            for (uint32_t c0 = 0; c0 < a.NC; c0 += CS) {
                const uint32_t c = c0 + cl;
                bool busy[NL];
#pragma unroll
                for (uint32_t k = 0; k < NL; ++k) {
                    busy[k] = false;
                    if (ok && c < a.NC) {
                        /* chunks without crossing edges or active cells (most of
                         * them) are skipped on the count pass's prefixes alone */
                        const uint32_t l = l0 + k;
                        if (!((a.rec_levels >> l) & 1u)) continue;   // a level without records
                        const uint64_t w = (uint64_t)l * level_words + r * a.NC + c;
                        uint32_t v0, t0, v1, t1;
                        chunk_prefix(a, w, v0, t0);
                        if (c + 1 < a.NC) {
                            chunk_prefix(a, w + 1, v1, t1);
                        } else {
                            v1 = a.counts[(uint64_t)(l * 2 + 0) * a.nrows + r];
                            t1 = a.counts[(uint64_t)(l * 2 + 1) * a.nrows + r];
                        }
                        busy[k] = (!a.dup && v1 != v0) || t1 != t0;
                    }
                }
#pragma unroll
                for (uint32_t k = 0; k < NL; ++k) {
                    const uint64_t bm = __ballot(busy[k]);
                    if (busy[k])
                        q[qn + lanes_below(bm)] =
                            (uint64_t)y | (uint64_t)z << 20 | (uint64_t)(l0 + k) << 40 | (uint64_t)c << 44;
                    qn += (uint32_t)__popcll(bm);
                }
                while (qn >= 64) {
                    wave_sync();
                    qn -= 64;
                    compact_chunk(a, ntri, q[qn + lane]);
                    wave_sync();
                }
            }
        }
    }
    wave_sync();
    if (lane < qn) compact_chunk(a, ntri, q[lane]);
    wave_sync();   // the queue is reused by the next call
}

__device__ __forceinline__ void compact_rows(const JobArgs& a, const uint8_t* ntri, uint64_t* q, uint64_t r0, uint64_t r1,
                                             uint64_t wave_first, uint64_t wave_step)
{
    switch (a.nlev) {
    case 1: compact_rows_t<1>(a, ntri, q, r0, r1, wave_first, wave_step); break;
    case 2: compact_rows_t<2>(a, ntri, q, r0, r1, wave_first, wave_step); break;
    case 3: compact_rows_t<3>(a, ntri, q, r0, r1, wave_first, wave_step); break;
    case 4: compact_rows_t<4>(a, ntri, q, r0, r1, wave_first, wave_step); break;
    default: compact_rows_t<0>(a, ntri, q, r0, r1, wave_first, wave_step); break;
    }
}

__global__ __launch_bounds__(kBlock) void k_compact(JobArgs a)
{
    __shared__ uint8_t ntri[256];
    __shared__ uint64_t queue[kWaves][kCompactQueue];
    stage_ntri(ntri);
    const uint64_t G = 64u >> a.cshift, groups = (a.nrows + G - 1) / G;
    const Span sp = grid_span(0, groups, kWaves, uniform(threadIdx.x >> 6));
    const uint64_t r1 = sp.end * G < a.nrows ? sp.end * G : a.nrows;
    compact_rows(a, ntri, queue[threadIdx.x >> 6], sp.lo * G, r1, sp.first - sp.lo, sp.step);
}

/* Compaction of a batch of jobs: the rows of all jobs in one sequence, split
 * as the count pass splits them. */
__global__ __launch_bounds__(kBlock) void k_compact_batch(const JobArgs* jobs, const uint64_t* row_start, uint32_t njobs)
{
    __shared__ uint8_t ntri[256];
    __shared__ uint64_t queue[kWaves][kCompactQueue];
    stage_ntri(ntri);
    const uint64_t total = row_start[njobs];
    const uint64_t nwaves = ((uint64_t)gridDim.x * kBlock) >> 6;
    const uint64_t wave_id = (uint64_t)blockIdx.x * kWaves + uniform(threadIdx.x >> 6);   // uniform: scalar
    const uint64_t per = (total + nwaves - 1) / nwaves;
    const uint64_t begin = wave_id * per;
    const uint64_t end = begin + per < total ? begin + per : total;
    uint64_t g = begin;
    while (g < end) {
        const uint32_t j = __builtin_amdgcn_readfirstlane(find_job(row_start, njobs, g));
        const uint64_t je = row_start[j + 1] < end ? row_start[j + 1] : end;
        compact_rows(jobs[j], ntri, queue[threadIdx.x >> 6], g - row_start[j], je - row_start[j], 0, 1);
        g = je;
    }
}

/* Vertex id (within its level) of the vertex on the edge along `axis` of node
 * i of a chunk, from the chunk's crossing record o: the chunk's first vertex id
 * (row prefix + chunk prefix within the row) + the crossing edges of the
 * chunk's nodes before i and of node i's lower axes. */
__device__ __forceinline__ uint64_t cross_rank(const ChunkCross& o, uint32_t i, uint32_t axis)
{
    const uint64_t lo = (1ull << i) - 1ull;
    uint64_t r = o.base;
    r += (uint64_t)(__popcll(o.x & lo) + __popcll(o.y & lo) + __popcll(o.z & lo));
    if (axis >= 1) r += (o.x >> i) & 1ull;
    if (axis >= 2) r += (o.y >> i) & 1ull;
    return r;
}

/* Node values at the two ends of the edge of vertex key `key`. */
template <typename T>
__device__ __forceinline__ void edge_values(const JobArgs& a, uint64_t key, T& va, T& vb)
{
    const T* vol = (const T*)a.values;
    const uint64_t x = key & 0xFFFFFu, y = (key >> 20) & 0xFFFFFu, z = (key >> 40) & 0xFFFFFu;
    const uint32_t axis = (uint32_t)(key >> 60);
    const uint64_t line = a.nx, slice = (uint64_t)a.nx * a.ny;
    const uint64_t ia = x + y * line + z * slice;
    va = vol[ia];
    vb = vol[ia + (axis == 0 ? 1ull : axis == 1 ? line : slice)];
}

/* Coordinates of vertex g (of level l, which has outputs) from its key and
 * its edge's node values (calculate_isopoints, :365-437). */
template <typename T>
__device__ __forceinline__ void vertex_item(const JobArgs& a, uint32_t l, uint64_t g, uint64_t key, T va, T vb)
{
    const uint32_t x = (uint32_t)(key & 0xFFFFFu), y = (uint32_t)((key >> 20) & 0xFFFFFu);
    const uint32_t z = (uint32_t)((key >> 40) & 0xFFFFFu), axis = (uint32_t)(key >> 60);
    float p[3];
    interpolate_values<T>(a, a.iso[l], va, vb, x, y, z, x + (axis == 0), y + (axis == 1), z + (axis == 2), p);
    float* co = a.coords[l];
    const uint64_t id = g - a.voff[l];
    co[3 * id + 0] = p[0];
    co[3 * id + 1] = p[1];
    co[3 * id + 2] = p[2];
}

/* Vertices g, g + step, ... < n of level l.  The level's isolevel, first item
 * and outputs are uniform (scalar loads, which leave the vector memory
 * operations in flight alone).  Software pipeline: while item g is finished,
 * the node values of the next item and the key of the one after it are in
 * flight (vector memory operations complete in issue order, so they are issued
 * before item g's stores; indices past the range's end are clamped to its
 * last item, so that every iteration issues the same loads). */
template <typename T>
__device__ __forceinline__ void vertices_level(const JobArgs& a, uint32_t l, uint64_t g, uint64_t n, uint64_t step)
{
    if (g >= n || !a.coords[l]) return;   // no items, or a level without outputs (not admitted)
    const uint64_t last = n - 1;
    /* two register sets take turns (unrolled by two), so that no loaded
     * value is copied -- a copy would wait for its load */
    uint64_t k0 = a.vkey[g], k1;
    T a0, b0, a1, b1;
    edge_values<T>(a, k0, a0, b0);
    uint64_t kn = a.vkey[g + step < last ? g + step : last];
    for (;;) {
        k1 = kn;
        edge_values<T>(a, k1, a1, b1);
        kn = a.vkey[g + 2 * step < last ? g + 2 * step : last];
        vertex_item<T>(a, l, g, k0, a0, b0);
        g += step;
        if (g >= n) break;
        k0 = kn;
        edge_values<T>(a, k0, a0, b0);
        kn = a.vkey[g + 2 * step < last ? g + 2 * step : last];
        vertex_item<T>(a, l, g, k1, a1, b1);
        g += step;
        if (g >= n) break;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_vertices(JobArgs a)
{
    for (uint32_t l = 0; l < a.nlev; ++l) {
        const Span sp = grid_span(a.voff[l], a.voff[l + 1], kBlock, threadIdx.x);
        vertices_level<T>(a, l, sp.first, sp.end, sp.step);
    }
}

__device__ __forceinline__ void stage_table_block(uint64_t* tri)
{
    for (uint32_t i = threadIdx.x; i < 256; i += kBlock) tri[i] = d_tri[i];
    __syncthreads();
}

/* Every triangle of shared vertices (duplicated ones: dup_finish): its
 * connectivity (connect_isopoints, :439-521) and the face normal of the
 * shared vertices' coordinates -- the PolygonNormal output
 * (calculate_normals_on_polygon, :531-559) or, with VertexNormal, kept in
 * a.fn for the per-vertex sums. */
template <typename T, int M>
__device__ __forceinline__ void triangle_item(const JobArgs& a, const uint64_t* tri, uint64_t g, uint64_t rec, uint32_t ci)
{
    const uint64_t level_words = (uint64_t)a.nz * a.ny * a.NC;
    const uint32_t l = level_at(a.toff, a.nlev, g);
    const uint32_t x = (uint32_t)(rec & 0xFFFFFu), y = (uint32_t)((rec >> 20) & 0xFFFFFu);
    const uint32_t zc = (uint32_t)((rec >> 40) & 0xFFFFFu), t = (uint32_t)(rec >> 60);
    const uint64_t w = tri[ci];
    const uint32_t eds[3] = { (uint32_t)((w >> (12 * t)) & 0xF), (uint32_t)((w >> (12 * t + 8)) & 0xF),
                              (uint32_t)((w >> (12 * t + 4)) & 0xF) };
    float* coords = nullptr;
    float* normals = nullptr;
    uint32_t* conn = nullptr;
#pragma unroll
    for (int k = 0; k < kMaxLevels; ++k)
        if (k == (int)l) { coords = a.coords[k]; normals = a.normals[k]; conn = a.conn[k]; }
    if (!coords) return;   // a level without outputs (not admitted)
    const uint64_t id = g - a.toff[l];
    float p[3][3], nn[3];
    static_assert(M != kModeDup, "duplicated triangles: dup_finish");
    const ChunkCross* cross_l = a.cross + (uint64_t)l * level_words;
    /* the owner chunks' crossing records of all three edges are loaded
     * before any is used, so the three loads are in flight together */
    const uint64_t vbase = M == kModeVertex ? a.voff[l] : 0ull;
    ChunkCross cc[3];
    uint32_t bit[3], axs[3];
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) {
        const uint32_t o = edge_owner(eds[jj]);
        const uint32_t vx = x + (o & 1u), vy = y + ((o >> 1) & 1u), vz = zc + ((o >> 2) & 1u);
        bit[jj] = vx & 63u;
        axs[jj] = o >> 3;
        cc[jj] = cross_l[((uint64_t)vz * a.ny + vy) * a.NC + (vx >> 6)];
    }
    uint64_t vid[3];
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) vid[jj] = cross_rank(cc[jj], bit[jj], axs[jj]);
    /* the vertices' coordinates are loaded before this item's stores
     * (vector memory operations complete in issue order) */
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) {
        p[jj][0] = coords[3 * vid[jj] + 0];
        p[jj][1] = coords[3 * vid[jj] + 1];
        p[jj][2] = coords[3 * vid[jj] + 2];
    }
    conn[3 * id + 0] = (uint32_t)vid[0];
    conn[3 * id + 1] = (uint32_t)vid[1];
    conn[3 * id + 2] = (uint32_t)vid[2];
    if constexpr (M == kModeVertex) {
        /* the first triangle of this cell that uses an edge tells the edge's
         * vertex where the cell's records start; the cell sits at position q
         * around the edge (around_cell) */
        uint32_t* slots = a.vslot + 4 * vbase;
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
            bool first = true;
            for (uint32_t u = 0; u < t; ++u) {
                const uint32_t f0 = (uint32_t)((w >> (12 * u)) & 0xF), f1 = (uint32_t)((w >> (12 * u + 4)) & 0xF);
                const uint32_t f2 = (uint32_t)((w >> (12 * u + 8)) & 0xF);
                first = first && f0 != eds[jj] && f1 != eds[jj] && f2 != eds[jj];
            }
            if (!first) continue;
            const uint32_t o = edge_owner(eds[jj]);
            const uint32_t dx = o & 1u, dy = (o >> 1) & 1u, dz = (o >> 2) & 1u, ax = o >> 3;
            const uint32_t q = ax == 0 ? ((dy ? 0u : 1u) | (dz ? 0u : 2u))
                             : ax == 1 ? ((dx ? 0u : 1u) | (dz ? 0u : 2u))
                                       : ((dx ? 0u : 1u) | (dy ? 0u : 2u));
            slots[4 * vid[jj] + q] = (uint32_t)(g - t);
        }
    }
    face_normal(p[0], p[1], p[2], nn);
    float* dst = M == kModeVertex ? a.fn + 3 * g : normals + 3 * id;
    dst[0] = nn[0];
    dst[1] = nn[1];
    dst[2] = nn[2];
}

/* A duplicated triangle in flight: its record and cube index, and the node
 * values at both ends of its three edges. */
template <typename T>
struct DupItem {
    uint64_t rec;
    uint32_t ci;
    T va[3], vb[3];
};

/* The three edges of triangle t of cube index ci, in table order. */
__device__ __forceinline__ void dup_edges(const uint64_t* tri, uint64_t rec, uint32_t ci, uint32_t eds[3])
{
    const uint32_t t = (uint32_t)(rec >> 60);
    const uint64_t w = tri[ci];
    eds[0] = (uint32_t)((w >> (12 * t)) & 0xF);
    eds[1] = (uint32_t)((w >> (12 * t + 8)) & 0xF);
    eds[2] = (uint32_t)((w >> (12 * t + 4)) & 0xF);
}

/* Issues the node-value loads of item d (its record and cube index set). */
template <typename T>
__device__ __forceinline__ void dup_load(const JobArgs& a, const uint64_t* tri, DupItem<T>& d)
{
    const T* vol = (const T*)a.values;
    const uint64_t line = a.nx, slice = (uint64_t)a.nx * a.ny;
    const uint64_t x = d.rec & 0xFFFFFu, y = (d.rec >> 20) & 0xFFFFFu, zc = (d.rec >> 40) & 0xFFFFFu;
    uint32_t eds[3];
    dup_edges(tri, d.rec, d.ci, eds);
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) {
        const uint32_t qq = edge_ends(eds[jj]);
        const uint64_t ia = (x + (qq & 1u)) + (y + ((qq >> 1) & 1u)) * line + (zc + ((qq >> 2) & 1u)) * slice;
        const uint64_t ib = (x + ((qq >> 4) & 1u)) + (y + ((qq >> 5) & 1u)) * line + (zc + ((qq >> 6) & 1u)) * slice;
        d.va[jj] = vol[ia];
        d.vb[jj] = vol[ib];
    }
}

/* Duplicated triangle g from its loaded item: its three vertices in table
 * orientation and its face normal (extract_surfaces_with_duplication,
 * :155-243), as triangle_item computes them. */
template <typename T>
__device__ __forceinline__ void dup_finish(const JobArgs& a, const uint64_t* tri, uint64_t g, const DupItem<T>& d)
{
    const uint32_t l = level_at(a.toff, a.nlev, g);
    float* coords = nullptr;
    float* normals = nullptr;
#pragma unroll
    for (int k = 0; k < kMaxLevels; ++k)
        if (k == (int)l) { coords = a.coords[k]; normals = a.normals[k]; }
    if (!coords) return;   // a level without outputs (not admitted)
    const uint64_t id = g - a.toff[l];
    const double iso = iso_of(a, (int)l);
    const uint32_t x = (uint32_t)(d.rec & 0xFFFFFu), y = (uint32_t)((d.rec >> 20) & 0xFFFFFu);
    const uint32_t zc = (uint32_t)((d.rec >> 40) & 0xFFFFFu);
    uint32_t eds[3];
    dup_edges(tri, d.rec, d.ci, eds);
    float p[3][3], nn[3];
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) {
        const uint32_t qq = edge_ends(eds[jj]);
        interpolate_values<T>(a, iso, d.va[jj], d.vb[jj], x + (qq & 1u), y + ((qq >> 1) & 1u), zc + ((qq >> 2) & 1u),
                              x + ((qq >> 4) & 1u), y + ((qq >> 5) & 1u), zc + ((qq >> 6) & 1u), p[jj]);
    }
    face_normal(p[0], p[1], p[2], nn);
    float* co = coords + 9 * id;
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) {
        co[3 * jj + 0] = p[jj][0];
        co[3 * jj + 1] = p[jj][1];
        co[3 * jj + 2] = p[jj][2];
    }
    normals[3 * id + 0] = nn[0];
    normals[3 * id + 1] = nn[1];
    normals[3 * id + 2] = nn[2];
}

/* Duplicated triangles g, g + step, ... < n, software-pipelined over two
 * register sets that take turns (a copy of a loaded value would wait for its
 * load): while item g is interpolated and stored, the node values of the
 * next item and the record of the one after it are in flight (vector memory
 * operations complete in issue order, so the wait for item g's values leaves
 * the newer loads outstanding); indices past the range's end are clamped to
 * its last item, so that every iteration issues the same loads. */
template <typename T>
__device__ __forceinline__ void triangles_range_dup(const JobArgs& a, const uint64_t* tri, uint64_t g, uint64_t n,
                                                    uint64_t step)
{
    if (g >= n) return;
    const uint64_t last = n - 1;
    DupItem<T> d0, d1;
    d0.rec = a.trec[g];
    d0.ci = a.tci[g];
    dup_load<T>(a, tri, d0);
    uint64_t gn = g + step < last ? g + step : last;
    uint64_t rn = a.trec[gn];
    uint32_t cn = a.tci[gn];
    for (;;) {
        d1.rec = rn;
        d1.ci = cn;
        dup_load<T>(a, tri, d1);
        gn = g + 2 * step < last ? g + 2 * step : last;
        rn = a.trec[gn];
        cn = a.tci[gn];
        dup_finish<T>(a, tri, g, d0);
        g += step;
        if (g >= n) break;
        d0.rec = rn;
        d0.ci = cn;
        dup_load<T>(a, tri, d0);
        gn = g + 2 * step < last ? g + 2 * step : last;
        rn = a.trec[gn];
        cn = a.tci[gn];
        dup_finish<T>(a, tri, g, d1);
        g += step;
        if (g >= n) break;
    }
}

/* Triangles g, g + step, ... < n: duplicated ones through the two-stage
 * pipeline above; with shared vertices the next item's record is loaded
 * while this one is worked on. */
template <typename T, int M>
__device__ __forceinline__ void triangles_range(const JobArgs& a, const uint64_t* tri, uint64_t g, uint64_t n,
                                                uint64_t step)
{
    if constexpr (M == kModeDup) {
        triangles_range_dup<T>(a, tri, g, n, step);
    } else {
        uint64_t rec = g < n ? a.trec[g] : 0ull;
        uint32_t ci = g < n ? a.tci[g] : 0u;
        for (; g < n; g += step) {
            const uint64_t cur = rec;
            const uint32_t cur_ci = ci;
            if (g + step < n) {
                rec = a.trec[g + step];
                ci = a.tci[g + step];
            }
            triangle_item<T, M>(a, tri, g, cur, cur_ci);
        }
    }
}

template <typename T, int M>
__global__ __launch_bounds__(kBlock) void k_triangles(JobArgs a)
{
    __shared__ uint64_t tri[256];
    stage_table_block(tri);
    const Span sp = grid_span(0, a.toff[a.nlev], kBlock, threadIdx.x);
    triangles_range<T, M>(a, tri, sp.first, sp.end, sp.step);
}

/* The cells around vertex key `key`'s edge that lie in the volume: each is
 * active, and the first of its triangles that uses the edge wrote the slot. */
__device__ __forceinline__ void vn_cells_in(const JobArgs& a, uint64_t key, bool in[4])
{
    const int x = (int)(key & 0xFFFFFu), y = (int)((key >> 20) & 0xFFFFFu);
    const int zv = (int)((key >> 40) & 0xFFFFFu);
    const uint32_t axis = (uint32_t)(key >> 60);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int cx, cy, cz;
        around_cell(axis, q, x, y, zv, cx, cy, cz);
        in[q] = cx >= 0 && cy >= 0 && cz >= 0 && cx + 1 < (int)a.nx && cy + 1 < (int)a.ny && cz + 1 < (int)a.nz;
    }
}

/* Cube indices of the cells around the edge (0 for cells outside the volume).
 * With `always`, cells outside load record 0 instead of nothing (the job has
 * triangles), so that every call issues the same loads. */
template <bool always>
__device__ __forceinline__ void vn_cube_indices(const JobArgs& a, uint64_t key, uint4 sl, uint32_t ci[4])
{
    bool in[4];
    vn_cells_in(a, key, in);
    const uint32_t first[4] = { sl.x, sl.y, sl.z, sl.w };
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if constexpr (always) {
            const uint32_t c = a.tci[in[q] ? first[q] : 0u];
            ci[q] = in[q] ? c : 0u;
        } else {
            ci[q] = in[q] ? a.tci[first[q]] : 0u;
        }
    }
}

/* Per cell around the edge, the mask of its triangles (table order) that use
 * the edge, and the face normals of the first two such triangles of every
 * cell, loaded together (most cells have one or two). */
struct VnGather {
    uint32_t use[4];
    uint32_t first[4];
    float f[4][2][3];
};

__device__ __forceinline__ void vn_gather(const JobArgs& a, const uint64_t* tri, uint64_t key, uint4 sl,
                                          const uint32_t ci[4], VnGather& o)
{
    bool in[4];
    vn_cells_in(a, key, in);
    const uint32_t around = around_nibbles((uint32_t)(key >> 60));
    o.first[0] = sl.x;
    o.first[1] = sl.y;
    o.first[2] = sl.z;
    o.first[3] = sl.w;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        o.use[q] = 0u;
        if (!in[q]) continue;
        const uint64_t w = tri[ci[q]];
        const int nt = (int)(w >> 60);
        const uint32_t me = (around >> (4 * q)) & 0xFu;
        for (int t = 0; t < nt; ++t) {
            const uint32_t e0 = (uint32_t)((w >> (12 * t)) & 0xF);
            const uint32_t e1 = (uint32_t)((w >> (12 * t + 8)) & 0xF);
            const uint32_t e2 = (uint32_t)((w >> (12 * t + 4)) & 0xF);
            if (e0 == me || e1 == me || e2 == me) o.use[q] |= 1u << t;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint32_t m = o.use[q];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            o.f[q][k][0] = o.f[q][k][1] = o.f[q][k][2] = 0.0f;
            if (m) {
                const uint32_t t = (uint32_t)__ffs(m) - 1u;
                m &= m - 1u;
                const float* p = a.fn + 3 * ((uint64_t)o.first[q] + t);
                o.f[q][k][0] = p[0];
                o.f[q][k][1] = p[1];
                o.f[q][k][2] = p[2];
            }
        }
    }
}

/* VertexNormal of vertex g (level l, which has outputs): the float sum, in
 * triangle order, of the face normals of the triangles that use it
 * (calculate_normals_on_vertex, :569-605) -- the four cells around the edge in
 * cell scan order and each cell's triangles that use the edge in table order.
 * A slot without a triangle holds +0, and adding +0 changes no partial sum:
 * the sum starts at +0 and a float sum is -0 only when both terms are, so no
 * partial sum is -0.  The starting +0 comes from an opaque move, so that the
 * compiler does not fold the first addition into the first load's branch
 * (which would wait for that load alone). */
__device__ __forceinline__ void vn_finish(const JobArgs& a, uint32_t l, uint64_t g, const VnGather& o)
{
    float zero;
    asm volatile("v_mov_b32 %0, 0" : "=v"(zero));
    float nv[3] = { zero, zero, zero };
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            nv[0] = nv[0] + o.f[q][k][0];
            nv[1] = nv[1] + o.f[q][k][1];
            nv[2] = nv[2] + o.f[q][k][2];
        }
        uint32_t m = o.use[q];
        m &= m - 1u;
        m &= m - 1u;
        while (m) {   // further triangles of the cell that use the edge (rare)
            const uint32_t t = (uint32_t)__ffs(m) - 1u;
            m &= m - 1u;
            const float* p = a.fn + 3 * ((uint64_t)o.first[q] + t);
            nv[0] = nv[0] + p[0];
            nv[1] = nv[1] + p[1];
            nv[2] = nv[2] + p[2];
        }
    }
    float* normals = a.normals[l];
    const uint64_t id = g - a.voff[l];
    normals[3 * id + 0] = nv[0];
    normals[3 * id + 1] = nv[1];
    normals[3 * id + 2] = nv[2];
}

__device__ __forceinline__ void vertex_normal_item(const JobArgs& a, const uint64_t* tri, uint32_t l, uint64_t g,
                                                   uint64_t key, uint4 sl)
{
    if (!a.normals[l]) return;   // a level without outputs (not admitted)
    uint32_t ci[4];
    vn_cube_indices<false>(a, key, sl, ci);
    VnGather o;
    vn_gather(a, tri, key, sl, ci, o);
    vn_finish(a, l, g, o);
}

/* eight waves per SIMD (64 VGPRs, a few bytes of spill) beat six without */
/* Vertex normals g, g + step, ... < n of level l (`triangles`: the job has a
 * triangle at all; without one no vertex has a cell in the volume around it).
 * Software pipeline over two register sets that take turns (a copy would
 * wait for its load): while vertex g is summed, the next vertex's cube
 * indices are in flight, and the key and slots of the one after it are
 * loaded before this vertex's store (vector memory operations complete in
 * issue order); indices past the range's end are clamped to its last item. */
__device__ __forceinline__ void vertex_normals_level(const JobArgs& a, const uint64_t* tri, uint32_t l, bool triangles,
                                                     uint64_t g, uint64_t n, uint64_t step)
{
    if (g >= n || !a.normals[l]) return;   // no items, or a level without outputs (not admitted)
    const uint64_t last = n - 1;
    if (!triangles) {
        for (; g < n; g += step) vertex_normal_item(a, tri, l, g, a.vkey[g], *(const uint4*)(a.vslot + 4 * g));
        return;
    }
    uint64_t k0 = a.vkey[g], k1;
    uint4 s0 = *(const uint4*)(a.vslot + 4 * g), s1;
    uint32_t c0[4], c1[4];
    vn_cube_indices<true>(a, k0, s0, c0);
    {
        const uint64_t g1 = g + step < last ? g + step : last;
        k1 = a.vkey[g1];
        s1 = *(const uint4*)(a.vslot + 4 * g1);
    }
    VnGather o;
    for (;;) {
        vn_cube_indices<true>(a, k1, s1, c1);
        vn_gather(a, tri, k0, s0, c0, o);
        {
            const uint64_t g2 = g + 2 * step < last ? g + 2 * step : last;
            k0 = a.vkey[g2];
            s0 = *(const uint4*)(a.vslot + 4 * g2);
        }
        vn_finish(a, l, g, o);
        g += step;
        if (g >= n) break;
        vn_cube_indices<true>(a, k0, s0, c0);
        vn_gather(a, tri, k1, s1, c1, o);
        {
            const uint64_t g2 = g + 2 * step < last ? g + 2 * step : last;
            k1 = a.vkey[g2];
            s1 = *(const uint4*)(a.vslot + 4 * g2);
        }
        vn_finish(a, l, g, o);
        g += step;
        if (g >= n) break;
    }
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_vertex_normals(JobArgs a)
{
    __shared__ uint64_t tri[256];
    stage_table_block(tri);
    const bool triangles = a.toff[a.nlev] > 0;
    /* level by level (uniform per-level values) */
    for (uint32_t l = 0; l < a.nlev; ++l) {
        const Span sp = grid_span(a.voff[l], a.voff[l + 1], kBlock, threadIdx.x);
        vertex_normals_level(a, tri, l, triangles, sp.first, sp.end, sp.step);
    }
}

/* Items of a batch of jobs: start[j] is job j's first item in one sequence
 * over the launch's jobs (start[njobs]: the end).  Persistent grid: each
 * workgroup takes one contiguous part of the sequence (whole multiples of its
 * size) and runs the plain kernels' loops over the pieces of it that fall in
 * one job and one level -- no per-item job or level search. */
struct BlockPart { uint64_t lo, hi; };

__device__ __forceinline__ BlockPart block_part(uint64_t n)
{
    const uint64_t per = ((n + gridDim.x - 1) / gridDim.x + kBlock - 1) / kBlock * kBlock;
    BlockPart p;
    p.lo = (uint64_t)blockIdx.x * per < n ? (uint64_t)blockIdx.x * per : n;
    p.hi = p.lo + per < n ? p.lo + per : n;
    return p;
}

/* Calls f(jobs[j], lo, hi) for each job j with items in this workgroup's part,
 * [lo, hi) being those items as the job's own item indices. */
template <typename F>
__device__ __forceinline__ void for_part_jobs(const JobArgs* jobs, const uint64_t* start, uint32_t njobs, F&& f)
{
    const uint64_t base = uniform64(start[0]);
    const BlockPart bp = block_part(uniform64(start[njobs]) - base);
    if (bp.lo >= bp.hi) return;
    for (uint32_t j = uniform(find_job(start, njobs, bp.lo + base)); j < njobs; ++j) {
        const uint64_t s0 = uniform64(start[j]) - base, s1 = uniform64(start[j + 1]) - base;
        if (s0 >= bp.hi) break;
        const uint64_t lo = (bp.lo > s0 ? bp.lo : s0) - s0, hi = (bp.hi < s1 ? bp.hi : s1) - s0;
        if (lo >= hi) continue;
        f(jobs[j], lo, hi);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_vertices_batch(const JobArgs* jobs, const uint64_t* start, uint32_t njobs)
{
    for_part_jobs(jobs, start, njobs, [&](const JobArgs& a, uint64_t lo, uint64_t hi) {
        for (uint32_t l = 0; l < a.nlev; ++l) {
            const uint64_t l0 = lo > a.voff[l] ? lo : a.voff[l], l1 = hi < a.voff[l + 1] ? hi : a.voff[l + 1];
            if (l0 < l1) vertices_level<T>(a, l, l0 + threadIdx.x, l1, kBlock);
        }
    });
}

template <typename T, int M>
__global__ __launch_bounds__(kBlock) void k_triangles_batch(const JobArgs* jobs, const uint64_t* start, uint32_t njobs)
{
    __shared__ uint64_t tri[256];
    stage_table_block(tri);
    for_part_jobs(jobs, start, njobs, [&](const JobArgs& a, uint64_t lo, uint64_t hi) {
        triangles_range<T, M>(a, tri, lo + threadIdx.x, hi, kBlock);
    });
}

__global__ __launch_bounds__(kBlock) void k_vertex_normals_batch(const JobArgs* jobs, const uint64_t* start, uint32_t njobs)
{
    __shared__ uint64_t tri[256];
    stage_table_block(tri);
    for_part_jobs(jobs, start, njobs, [&](const JobArgs& a, uint64_t lo, uint64_t hi) {
        const bool triangles = a.toff[a.nlev] > 0;
        for (uint32_t l = 0; l < a.nlev; ++l) {
            const uint64_t l0 = lo > a.voff[l] ? lo : a.voff[l], l1 = hi < a.voff[l + 1] ? hi : a.voff[l + 1];
            if (l0 < l1) vertex_normals_level(a, tri, l, triangles, l0 + threadIdx.x, l1, kBlock);
        }
    });
}


/* ------------------------------------------------------------------------- */
/* SlicePlane: classification by a plane, emit with per-vertex colours         */
/* ------------------------------------------------------------------------- */

/* The plane's value at (x, y, z) as SlicePlane::substitute_plane_equation
 * evaluates it (SlicePlane.cpp:936-971): c0*x + c1*y + c2*z + c3, left to right
 * in float, node indices converted to float. */
__device__ __forceinline__ float plane_at(const float c[4], float x, float y, float z)
{
    float s = c[0] * x;
    s = s + c[1] * y;
    s = s + c[2] * z;
    return s + c[3];
}

/* Bit i of chunk c of row (y, z): plane(64c+i, y, z) > 0.0f
 * (calculate_table_index, SlicePlane.cpp:802-818).  Nothing is read: one node
 * row per wave, one ballot per 64 nodes, the masks of up to 64 chunks parked
 * in lanes and stored together. */
__global__ __launch_bounds__(kBlock) void k_classify_plane(JobArgs a)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t nrows = (uint64_t)a.ny * a.nz;
    const uint64_t wave_id = (uint64_t)blockIdx.x * kWaves + uniform(threadIdx.x >> 6);   // uniform: scalar
    const uint64_t nwaves = ((uint64_t)gridDim.x * kBlock) >> 6;
    for (uint64_t row = wave_id; row < nrows; row += nwaves) {
        const uint64_t z = row / a.ny, y = row - z * a.ny;
        const float fy = (float)y, fz = (float)z;
        for (uint32_t c0 = 0; c0 < a.NC; c0 += 64) {
            const uint32_t cend = (a.NC - c0) < 64u ? (a.NC - c0) : 64u;
            uint64_t mine = 0;
            for (uint32_t i = 0; i < cend; ++i) {
                const uint32_t x = (c0 + i) * 64u + lane;
                const uint64_t m = __ballot(x < a.nx && plane_at(a.plane, (float)x, fy, fz) > 0.0f);
                mine = lane == i ? m : mine;
            }
            if (lane < cend) a.bitvol[row * a.NC + c0 + lane] = mine;
        }
    }
}

/* interpolate_value<T> (SlicePlane.cpp:1001-1026): the value at ratio r along
 * the edge, computed in the types C++ gives values[i0] + r * (values[i1] -
 * values[i0]) for T (integer promotion and unsigned wrap-around of the
 * difference, a float product, double only for double data). */
template <typename T>
__device__ __forceinline__ double slice_value(T va, T vb, float r)
{
    return va + r * (vb - va);
}

/* Math::Round (Utility/Math.h:110-114): half away from zero, in double. */
__device__ __forceinline__ int round_half_away(float a) { return a >= 0 ? (int)(a + 0.5) : -(int)(0.5 - a); }

__device__ __forceinline__ float clampf(float a, float lo, float hi) { return a <= lo ? lo : (a <= hi ? a : hi); }

/* Colour-map entry s as r | g << 8 | b << 16: from the device table (3 bytes
 * an entry) or from a packed copy in LDS. */
struct GlobalColors {
    const uint8_t* t;
    __device__ __forceinline__ uint32_t operator()(uint32_t s) const
    {
        return (uint32_t)t[3 * s] | (uint32_t)t[3 * s + 1] << 8 | (uint32_t)t[3 * s + 2] << 16;
    }
};
struct LdsColors {
    const uint32_t* t;
    __device__ __forceinline__ uint32_t operator()(uint32_t s) const { return t[s]; }
};

/* ColorMap::at (ColorMap.cpp:1043-1054): the two table entries around the
 * value, mixed by RGBColor::Mix (RGBColor.cpp:23-31) in [0,1] floats,
 * x*(1-t) + y*t, and rounded back to bytes.  A NaN position (an empty range)
 * takes the first entry. */
template <typename C>
__device__ __forceinline__ void color_at(const JobArgs& a, const C& tab, float value, uint8_t rgb[3])
{
    const float lo = a.crange[0], hi = a.crange[1];
    const float v0 = clampf(value, lo, hi);
    const float rr = (float)(a.cres - 1);
    float v = (v0 - lo) / (hi - lo) * rr;
    if (!(v >= 0.0f)) v = 0.0f;
    const uint32_t s0 = (uint32_t)v;
    const uint32_t s1 = (s0 + 1 < a.cres - 1) ? s0 + 1 : a.cres - 1;
    const float t = v - (float)s0;
    const uint32_t e0 = tab(s0), e1 = tab(s1);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float c0 = (float)((e0 >> (8 * k)) & 0xFFu) / 255.0f;
        const float c1 = (float)((e1 >> (8 * k)) & 0xFFu) / 255.0f;
        const float p = c0 * (1 - t);
        const float q = c1 * t;
        rgb[k] = (uint8_t)round_half_away(clampf(p + q, 0.0f, 1.0f) * 255.0f);
    }
}

/* One triangle of extract_plane<T> (SlicePlane.cpp:165-280): per edge in the
 * order (T[i], T[i+2], T[i+1]) its plane crossing in table orientation
 * (interpolate_vertex :981-990, float), (p + min) * scale, and the colour of
 * the value interpolated at the same ratio; the normal is
 * -(v2 - v0) x (v1 - v0). */
template <typename T, typename C, typename Sink>
__device__ __forceinline__ void slice_triangle_values(const JobArgs& a, const uint64_t* tri, uint64_t rec, uint32_t ci,
                                                      const C& colors, Sink& out)
{
    const T* vol = (const T*)a.values;
    const uint32_t x = (uint32_t)(rec & 0xFFFFFu), y = (uint32_t)((rec >> 20) & 0xFFFFFu);
    const uint32_t zc = (uint32_t)((rec >> 40) & 0xFFFFFu), t = (uint32_t)(rec >> 60);
    const uint64_t w = tri[ci];
    const uint32_t eds[3] = { (uint32_t)((w >> (12 * t)) & 0xF), (uint32_t)((w >> (12 * t + 8)) & 0xF),
                              (uint32_t)((w >> (12 * t + 4)) & 0xF) };
    const uint64_t line = a.nx, slice = (uint64_t)a.nx * a.ny;
    float p[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const uint32_t qq = edge_ends(eds[j]);
        const uint32_t ax = x + (qq & 1u), ay = y + ((qq >> 1) & 1u), az = zc + ((qq >> 2) & 1u);
        const uint32_t bx = x + ((qq >> 4) & 1u), by = y + ((qq >> 5) & 1u), bz = zc + ((qq >> 6) & 1u);
        const float fa[3] = { (float)ax, (float)ay, (float)az };
        const float fb[3] = { (float)bx, (float)by, (float)bz };
        const float f0 = plane_at(a.plane, fa[0], fa[1], fa[2]);
        const float f1 = plane_at(a.plane, fb[0], fb[1], fb[2]);
        float r = f0 / (f1 - f0);
        r = r < 0 ? -r : r;
        const float s = 1.0f - r;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float u = s * fa[k];
            const float v = r * fb[k];
            const float q = (u + v) + a.minc[k];
            p[j][k] = q * a.sc[k];
        }
        const T va = vol[ax + ay * line + az * slice];
        const T vb = vol[bx + by * line + bz * slice];
        uint8_t rgb[3];
        color_at(a, colors, (float)slice_value<T>(va, vb, r), rgb);
        out.vertex(j, p[j], rgb);
    }
    const float d1x = p[2][0] - p[0][0], d1y = p[2][1] - p[0][1], d1z = p[2][2] - p[0][2];
    const float d2x = p[1][0] - p[0][0], d2y = p[1][1] - p[0][1], d2z = p[1][2] - p[0][2];
    const float nn[3] = { -(d1y * d2z - d1z * d2y), -(d1z * d2x - d1x * d2z), -(d1x * d2y - d1y * d2x) };
    out.normal(nn);
}

/* Where a triangle's corners (coordinates, colour) and normal go: arrays of
 * 9, 9 and 3 elements a triangle, from element 0 of triangle 0. */
template <typename F, typename B>
struct SliceSink {
    F* co;
    F* no;
    B* cl;
    __device__ __forceinline__ void vertex(int j, const float p[3], const uint8_t rgb[3])
    {
        co[3 * j + 0] = p[0];
        co[3 * j + 1] = p[1];
        co[3 * j + 2] = p[2];
        cl[3 * j + 0] = rgb[0];
        cl[3 * j + 1] = rgb[1];
        cl[3 * j + 2] = rgb[2];
    }
    __device__ __forceinline__ void normal(const float n[3])
    {
        no[0] = n[0];
        no[1] = n[1];
        no[2] = n[2];
    }
};

/* Triangle `id` of the output, written by its own thread. */
template <typename T>
__device__ __forceinline__ void slice_triangle(const JobArgs& a, const uint64_t* tri, uint64_t id, uint64_t rec,
                                               uint32_t ci)
{
    SliceSink<float, uint8_t> out{ a.coords[0] + 9 * id, a.normals[0] + 3 * id, a.colors + 9 * id };
    slice_triangle_values<T>(a, tri, rec, ci, GlobalColors{ a.ctable }, out);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_slice_triangles(JobArgs a)
{
    __shared__ uint64_t tri[256];
    stage_table_block(tri);
    const Span sp = grid_span(0, a.toff[1], kBlock, threadIdx.x);
    for (uint64_t g = sp.first; g < sp.end; g += sp.step) slice_triangle<T>(a, tri, g, a.trec[g], a.tci[g]);
}

/* SlicePlane without classifying every node.  Along a node row (y, z) the
 * plane value plane_at(c, x, y, z) is monotone in x -- each float operation
 * is monotone in its operand, rounding included -- so the nodes above the
 * plane are x >= t (c0 > 0; c0 == 0 with t = 0 or nx) or x < t (c0 < 0) for
 * one threshold t per node row, found by bisection on the same float
 * expression.  A cell row's active cells are then the one run
 * [min t - 1, max t - 1] over its four node rows, cut by the points t - 1
 * and t of each row into at most nine runs of one cube index.  Counts per
 * cell row, a scan, then one thread per triangle.  The host takes this path
 * when the plane is finite at the volume's eight corners (and so, being
 * monotone in each coordinate, everywhere inside). */
__device__ __forceinline__ bool above_thr(int dir, int32_t t, int32_t x) { return dir < 0 ? x < t : x >= t; }

__global__ __launch_bounds__(kBlock) void k_slice_thresholds(JobArgs a, int32_t* thr)
{
    const int dir = a.plane[0] > 0.0f ? 1 : (a.plane[0] < 0.0f ? -1 : 0);
    const uint64_t nrows = (uint64_t)a.ny * a.nz, step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t row = (uint64_t)blockIdx.x * kBlock + threadIdx.x; row < nrows; row += step) {
        const uint32_t z = (uint32_t)(row / a.ny), y = (uint32_t)(row - (uint64_t)z * a.ny);
        const float fy = (float)y, fz = (float)z;
        int32_t t;
        if (dir == 0) {
            t = plane_at(a.plane, 0.0f, fy, fz) > 0.0f ? 0 : (int32_t)a.nx;
        } else {
            /* the first x whose node is above (c0 > 0) or not above (c0 < 0) */
            uint32_t lo = 0, hi = a.nx;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                const bool up = plane_at(a.plane, (float)mid, fy, fz) > 0.0f;
                if (up == (dir > 0)) hi = mid; else lo = mid + 1;
            }
            t = (int32_t)lo;
        }
        thr[row] = t;
    }
}

/* The active cells of cell row (y, z) as nine runs [b[i], b[i + 1]) of one
 * cube index each (ci(i)): the breakpoints t - 1 and t of the four node rows,
 * clamped to [lo, hi + 1] and sorted, so that coinciding or outside points
 * give empty runs.  Every index is static, so the runs stay in registers. */
struct SliceRuns {
    int32_t b[10];
    int32_t t[4];   // thresholds of node rows (y,z), (y+1,z), (y,z+1), (y+1,z+1)
    int dir;
    __device__ __forceinline__ uint32_t ci(int i) const;
};

__device__ __forceinline__ uint32_t slice_cube_index(int dir, const int32_t t[4], int32_t x)
{
    /* corners 0..7: (x,y,z) (x+1,y,z) (x+1,y+1,z) (x,y+1,z) and the same at z+1;
     * t: node rows (y,z), (y+1,z), (y,z+1), (y+1,z+1) */
    return (uint32_t)above_thr(dir, t[0], x) | (uint32_t)above_thr(dir, t[0], x + 1) << 1 |
           (uint32_t)above_thr(dir, t[1], x + 1) << 2 | (uint32_t)above_thr(dir, t[1], x) << 3 |
           (uint32_t)above_thr(dir, t[2], x) << 4 | (uint32_t)above_thr(dir, t[2], x + 1) << 5 |
           (uint32_t)above_thr(dir, t[3], x + 1) << 6 | (uint32_t)above_thr(dir, t[3], x) << 7;
}

__device__ __forceinline__ uint32_t SliceRuns::ci(int i) const { return slice_cube_index(dir, t, b[i]); }

__device__ __forceinline__ void order2(int32_t& p, int32_t& q)
{
    const int32_t lo = p < q ? p : q, hi = p < q ? q : p;
    p = lo;
    q = hi;
}

/* Returns false when the cell row has no active cell. */
__device__ __forceinline__ bool slice_runs(const JobArgs& a, const int32_t* thr, uint32_t y, uint32_t z, SliceRuns& r)
{
    r.dir = a.plane[0] < 0.0f ? -1 : 1;
    const uint64_t r0 = (uint64_t)z * a.ny + y;
    int32_t* t = r.t;
    t[0] = thr[r0];
    t[1] = thr[r0 + 1];
    t[2] = thr[r0 + a.ny];
    t[3] = thr[r0 + a.ny + 1];
    const int32_t tmin = min(min(t[0], t[1]), min(t[2], t[3]));
    const int32_t tmax = max(max(t[0], t[1]), max(t[2], t[3]));
    const int32_t lo = max(tmin - 1, 0), hi = min(tmax - 1, (int32_t)a.nx - 2);
    if (lo > hi) return false;
    int32_t* b = r.b;
    b[0] = lo;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        b[1 + 2 * k] = min(max(t[k] - 1, lo), hi + 1);
        b[2 + 2 * k] = min(max(t[k], lo), hi + 1);
    }
    b[9] = hi + 1;
    /* sorting network of b[1..8] (19 comparators) */
    order2(b[1], b[3]); order2(b[2], b[4]); order2(b[5], b[7]); order2(b[6], b[8]);
    order2(b[1], b[5]); order2(b[2], b[6]); order2(b[3], b[7]); order2(b[4], b[8]);
    order2(b[1], b[2]); order2(b[3], b[4]); order2(b[5], b[6]); order2(b[7], b[8]);
    order2(b[3], b[5]); order2(b[4], b[6]);
    order2(b[2], b[5]); order2(b[4], b[7]);
    order2(b[2], b[3]); order2(b[4], b[5]); order2(b[6], b[7]);
    return true;
}

__global__ __launch_bounds__(kBlock) void k_slice_row_counts(JobArgs a, const int32_t* thr, uint32_t* counts)
{
    __shared__ uint8_t ntri[256];
    for (uint32_t i = threadIdx.x; i < 256; i += kBlock) ntri[i] = (uint8_t)(d_tri[i] >> 60);
    __syncthreads();
    const uint64_t ncr = (uint64_t)(a.ny - 1) * (a.nz - 1), step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t cr = (uint64_t)blockIdx.x * kBlock + threadIdx.x; cr < ncr; cr += step) {
        const uint32_t z = (uint32_t)(cr / (a.ny - 1)), y = (uint32_t)(cr - (uint64_t)z * (a.ny - 1));
        SliceRuns r;
        uint32_t c = 0;
        if (slice_runs(a, thr, y, z, r)) {
#pragma unroll
            for (int i = 0; i < 9; ++i) c += (uint32_t)(r.b[i + 1] - r.b[i]) * ntri[r.ci(i)];
        }
        counts[cr] = c;
    }
}

/* The cell rows of the first and the last triangle of every unit of 64
 * triangles (unit u: triangles 64u .. min(64u + 63, n - 1)), written by the
 * rows that hold them: one writer each. */
__global__ __launch_bounds__(kBlock) void k_slice_unit_rows(uint64_t ncr, const uint32_t* counts, const uint64_t* prefix,
                                                           uint64_t n, uint64_t* first, uint64_t* last)
{
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t cr = (uint64_t)blockIdx.x * kBlock + threadIdx.x; cr < ncr; cr += step) {
        const uint32_t c = counts[cr];
        if (!c) continue;
        const uint64_t p = prefix[cr], e = p + c;   // the row's triangles [p, e)
        for (uint64_t u = p >> 6; (u << 6) < e; ++u) {
            const uint64_t f = u << 6, l = f + 63 < n ? f + 63 : n - 1;
            if (f >= p) first[u] = cr;
            if (l >= p && l < e) last[u] = cr;
        }
    }
}

/* Colour maps of up to this many entries are staged in LDS by the run emit. */
constexpr uint32_t kSliceLdsColors = 1024;

/* One thread per triangle.  A wave's 64 triangles lie in the cell rows
 * between those of its first and last (k_slice_unit_rows); each lane bisects
 * that short range, then walks its row's runs to its cell and table index.
 * The wave's 64 triangles are consecutive in the output: their coordinates,
 * normals and colours are assembled in LDS and written as whole 16-byte
 * words. */
template <typename T, bool kLdsColors>
__global__ __launch_bounds__(kBlock) void k_slice_triangles_runs(JobArgs a, const int32_t* thr, const uint64_t* prefix,
                                                                const uint64_t* unit_first, const uint64_t* unit_last)
{
    __shared__ uint64_t tri[256];
    __shared__ __align__(16) float stage[kWaves][9 * 64 + 3 * 64 + 9 * 64 / 4];   // coords, normals, colours
    extern __shared__ __align__(16) uint8_t smem[];
    uint32_t* ctab = (uint32_t*)smem;
    if (kLdsColors)
        for (uint32_t i = threadIdx.x; i < a.cres; i += kBlock) ctab[i] = GlobalColors{ a.ctable }(i);
    stage_table_block(tri);
    const uint64_t n = a.toff[1];
    const uint32_t lane = threadIdx.x & 63u;
    float* sc = stage[threadIdx.x >> 6];
    float* sn = sc + 9 * 64;
    uint8_t* scl = (uint8_t*)(sn + 3 * 64);
    const Span sp = grid_span(0, n, kBlock, threadIdx.x);
    for (uint64_t gw = sp.first - lane; gw < n; gw += sp.step) {   // wave-uniform
        uint64_t lo = unit_first[gw >> 6], hi = unit_last[gw >> 6];
        const uint64_t g = gw + lane;
        if (g < n) {
            while (lo < hi) {
                const uint64_t mid = (lo + hi + 1) >> 1;
                if (prefix[mid] <= g) lo = mid; else hi = mid - 1;
            }
            const uint32_t z = (uint32_t)(lo / (a.ny - 1)), y = (uint32_t)(lo - (uint64_t)z * (a.ny - 1));
            SliceRuns r;
            slice_runs(a, thr, y, z, r);   // the row has triangles, so active cells
            uint32_t k = (uint32_t)(g - prefix[lo]);
            uint32_t x = 0, t = 0, ci = 0;
            bool found = false;
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                const uint32_t ci_i = r.ci(i);
                const uint32_t nt = (uint32_t)(tri[ci_i] >> 60);
                const uint32_t m = (uint32_t)(r.b[i + 1] - r.b[i]) * nt;
                if (!found) {
                    if (k < m) {
                        x = (uint32_t)r.b[i] + k / nt;
                        t = k % nt;
                        ci = ci_i;
                        found = true;
                    } else {
                        k -= m;
                    }
                }
            }
            const uint64_t rec = (uint64_t)x | (uint64_t)y << 20 | (uint64_t)z << 40 | (uint64_t)t << 60;
            SliceSink<float, uint8_t> out{ sc + 9 * lane, sn + 3 * lane, scl + 9 * lane };
            if (kLdsColors)
                slice_triangle_values<T>(a, tri, rec, ci, LdsColors{ ctab }, out);
            else
                slice_triangle_values<T>(a, tri, rec, ci, GlobalColors{ a.ctable }, out);
        }
        wave_sync();
        /* the wave's output: triangles gw .. gw + cnt - 1 (9 * gw floats and
         * bytes from the array starts: 16-byte aligned when cnt is 64) */
        const uint32_t cnt = n - gw < 64 ? (uint32_t)(n - gw) : 64u;
        float* co = a.coords[0] + 9 * gw;
        float* no = a.normals[0] + 3 * gw;
        uint8_t* cl = a.colors + 9 * gw;
        if (cnt == 64) {
            for (uint32_t i = lane; i < 9 * 64 / 4; i += 64) ((float4*)co)[i] = ((const float4*)sc)[i];
            if (lane < 3 * 64 / 4) ((float4*)no)[lane] = ((const float4*)sn)[lane];
            if (lane < 9 * 64 / 16) ((uint4*)cl)[lane] = ((const uint4*)scl)[lane];
        } else {
            for (uint32_t i = lane; i < 9 * cnt; i += 64) co[i] = sc[i];
            for (uint32_t i = lane; i < 3 * cnt; i += 64) no[i] = sn[i];
            for (uint32_t i = lane; i < 9 * cnt; i += 64) cl[i] = scl[i];
        }
        wave_sync();
    }
}

/* ------------------------------------------------------------------------- */
/* row-count scan                                                              */
/* ------------------------------------------------------------------------- */

constexpr uint32_t kScanPer = 16;                      // elements per thread
constexpr uint32_t kScanTile = kScanPer * kBlock;      // elements per block

/* Sum of each kScanTile segment of each of the narr count arrays. */
__global__ __launch_bounds__(kBlock) void k_scan_partials(const uint32_t* counts, uint64_t n, uint32_t nseg, uint64_t* partials)
{
    __shared__ uint64_t sh[kWaves];
    const uint32_t arr = blockIdx.x / nseg, seg = blockIdx.x - arr * nseg;
    const uint32_t* src = counts + (uint64_t)arr * n;
    const uint64_t b = (uint64_t)seg * kScanTile;
    uint64_t s = 0;
    for (uint32_t i = threadIdx.x; i < kScanTile; i += kBlock) {
        const uint64_t g = b + i;
        if (g < n) s += src[g];
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += (uint64_t)__shfl_xor((unsigned long long)s, d);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < kWaves; ++w) t += sh[w];
        partials[blockIdx.x] = t;
    }
}

/* In place: partials of each array become their exclusive prefix, and the
 * grand total of array arr goes to totals[arr].  One workgroup. */
__global__ __launch_bounds__(kBlock) void k_scan_mid(uint64_t* partials, uint32_t nseg, uint32_t narr, uint64_t* totals)
{
    __shared__ uint64_t sh[kWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t arr = 0; arr < narr; ++arr) {
        uint64_t* p = partials + (uint64_t)arr * nseg;
        uint64_t carry = 0;
        for (uint32_t b0 = 0; b0 < nseg; b0 += kScanTile) {
            const uint32_t b = b0 + threadIdx.x * kScanPer;
            uint64_t vals[kScanPer];
            uint64_t s = 0;
#pragma unroll
            for (uint32_t i = 0; i < kScanPer; ++i) {
                vals[i] = b + i < nseg ? p[b + i] : 0ull;
                s += vals[i];
            }
            uint64_t incl = s;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint64_t t = (uint64_t)__shfl_up((unsigned long long)incl, d);
                if (lane >= d) incl += t;
            }
            if (lane == 63) sh[wave] = incl;
            __syncthreads();
            uint64_t run = carry;
            uint64_t all = 0;
            for (int w = 0; w < kWaves; ++w) {
                if (w < wave) run += sh[w];
                all += sh[w];
            }
            run += incl - s;
#pragma unroll
            for (uint32_t i = 0; i < kScanPer; ++i) {
                if (b + i < nseg) p[b + i] = run;
                run += vals[i];
            }
            carry += all;
            __syncthreads();
        }
        if (threadIdx.x == 0) totals[arr] = carry;
    }
}

/* Exclusive prefix of each count array, from the scanned partials. */
__global__ __launch_bounds__(kBlock) void k_scan_apply(const uint32_t* counts, uint64_t n, uint32_t nseg,
                                                       const uint64_t* partials, uint64_t* prefix)
{
    __shared__ uint64_t sh[kWaves];
    const uint32_t arr = blockIdx.x / nseg, seg = blockIdx.x - arr * nseg;
    const uint32_t* src = counts + (uint64_t)arr * n;
    uint64_t* dst = prefix + (uint64_t)arr * n;
    const uint64_t base = partials[blockIdx.x];
    const uint64_t b = (uint64_t)seg * kScanTile + (uint64_t)threadIdx.x * kScanPer;
    uint32_t vals[kScanPer];
    uint64_t s = 0;
    /* a thread's 16 counts are contiguous: 16-byte loads where they lie
     * inside the array (the array bases are 16-byte aligned when n is a
     * multiple of 4) */
    const bool whole = b + kScanPer <= n && (n & 3u) == 0;
    if (whole) {
#pragma unroll
        for (uint32_t i = 0; i < kScanPer; i += 4) {
            const uint4 v = *(const uint4*)(src + b + i);
            vals[i] = v.x;
            vals[i + 1] = v.y;
            vals[i + 2] = v.z;
            vals[i + 3] = v.w;
        }
    } else {
#pragma unroll
        for (uint32_t i = 0; i < kScanPer; ++i) {
            const uint64_t g = b + i;
            vals[i] = g < n ? src[g] : 0u;
        }
    }
#pragma unroll
    for (uint32_t i = 0; i < kScanPer; ++i) s += vals[i];
    /* exclusive scan of s across the block */
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t incl = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t t = (uint64_t)__shfl_up((unsigned long long)incl, d);
        if (lane >= d) incl += t;
    }
    if (lane == 63) sh[wave] = incl;
    __syncthreads();
    uint64_t run = base;
    for (int w = 0; w < wave; ++w) run += sh[w];
    run += incl - s;
    if (whole) {
#pragma unroll
        for (uint32_t i = 0; i < kScanPer; i += 2) {
            ulonglong2 o;
            o.x = run;
            run += vals[i];
            o.y = run;
            run += vals[i + 1];
            *(ulonglong2*)(dst + b + i) = o;
        }
    } else {
#pragma unroll
        for (uint32_t i = 0; i < kScanPer; ++i) {
            const uint64_t g = b + i;
            if (g < n) dst[g] = run;
            run += vals[i];
        }
    }
}

/* ------------------------------------------------------------------------- */
/* min / max                                                                   */
/* ------------------------------------------------------------------------- */

/* GetMinMaxValues<T> (StructuredVolumeObject.cpp:66-88) is a serial scan
 * min = v < min ? v : min seeded with the first value: NaNs never replace the
 * running extreme, and a NaN first value is never replaced.  The parallel form
 * reduces the non-NaN values and applies the first-value rule at the end. */
template <typename T>
__device__ __forceinline__ bool is_nan_val(T) { return false; }
template <> __device__ __forceinline__ bool is_nan_val<float>(float v) { return v != v; }
template <> __device__ __forceinline__ bool is_nan_val<double>(double v) { return v != v; }

/* Running extreme with the reference's tie rule: on equal values the earlier
 * index wins (only observable for float +0/-0). */
constexpr uint32_t kMinMaxParts = 2048;   // partial results of the first pass

template <typename T>
struct Extreme {
    T mn, mx;
    uint64_t imn, imx;
    uint32_t valid;
};

template <typename T>
__device__ __forceinline__ void merge_extreme(Extreme<T>& a, const Extreme<T>& b)
{
    if (!b.valid) return;
    if (!a.valid) { a = b; return; }
    if (b.mn < a.mn || (!(a.mn < b.mn) && b.imn < a.imn)) { a.mn = b.mn; a.imn = b.imn; }
    if (b.mx > a.mx || (!(a.mx > b.mx) && b.imx < a.imx)) { a.mx = b.mx; a.imx = b.imx; }
}

template <typename T>
__device__ __forceinline__ void take_value(Extreme<T>& e, T v, uint64_t i)
{
    if (is_nan_val<T>(v)) return;
    if (!e.valid) { e.mn = v; e.mx = v; e.imn = i; e.imx = i; e.valid = 1; return; }
    if (v < e.mn) { e.mn = v; e.imn = i; }
    if (v > e.mx) { e.mx = v; e.imx = i; }
}

/* Block-wide merge of the threads' extremes; thread 0 gets the result. */
template <typename T>
__device__ __forceinline__ Extreme<T> block_extreme(Extreme<T> e)
{
    __shared__ Extreme<T> sh[kBlock];
    sh[threadIdx.x] = e;
    __syncthreads();
    for (uint32_t s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            Extreme<T> x = sh[threadIdx.x];
            merge_extreme(x, sh[threadIdx.x + s]);
            sh[threadIdx.x] = x;
        }
        __syncthreads();
    }
    return sh[0];
}

/* Streams the volume with 16-byte loads, kU of them in flight per thread;
 * each thread visits its elements in increasing index order, so strict
 * comparisons keep the first occurrence, and merges break ties by index. */
template <typename T>
__global__ __launch_bounds__(kBlock) void k_minmax_partial(const T* vol, uint64_t n, Extreme<T>* parts)
{
    constexpr uint32_t V = 16 / sizeof(T);
    constexpr uint32_t kU = 4;
    Extreme<T> e;
    e.mn = T(0); e.mx = T(0); e.imn = 0; e.imx = 0; e.valid = 0;
    const uint64_t tid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t nth = (uint64_t)gridDim.x * kBlock;
    const bool aligned = ((uintptr_t)vol & 15u) == 0;
    const uint64_t nvec = aligned ? n / V : 0;
    const uint4* vp = (const uint4*)vol;
    for (uint64_t b = tid; b < nvec; b += nth * kU) {
        uint4 r[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            const uint64_t idx = b + u * nth;
            r[u] = idx < nvec ? vp[idx] : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            const uint64_t idx = b + u * nth;
            if (idx < nvec) {
                T vals[V];
                __builtin_memcpy(vals, &r[u], 16);
#pragma unroll
                for (uint32_t k = 0; k < V; ++k) take_value(e, vals[k], idx * V + k);
            }
        }
    }
    for (uint64_t i = nvec * V + tid; i < n; i += nth) take_value(e, vol[i], i);
    const Extreme<T> r = block_extreme(e);
    if (threadIdx.x == 0) parts[blockIdx.x] = r;
}

/* GetMinMaxValues<T> of an unstructured volume (UnstructuredVolumeObject.cpp:
 * 123-141): the first value of the array, then every connection entry in
 * order.  Position 0 is vol[0], position p >= 1 is vol[conn[p - 1]]; each
 * thread visits its positions in increasing order. */
template <typename T>
__global__ __launch_bounds__(kBlock) void k_minmax_gather_partial(const T* vol, const uint32_t* conn, uint64_t nconn,
                                                                  Extreme<T>* parts)
{
    Extreme<T> e;
    e.mn = T(0); e.mx = T(0); e.imn = 0; e.imx = 0; e.valid = 0;
    const uint64_t tid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t nth = (uint64_t)gridDim.x * kBlock;
    for (uint64_t p = tid; p <= nconn; p += nth) take_value(e, p == 0 ? vol[0] : vol[conn[p - 1]], p);
    const Extreme<T> r = block_extreme(e);
    if (threadIdx.x == 0) parts[blockIdx.x] = r;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_minmax_final(const T* vol, const Extreme<T>* parts, uint32_t nparts, double* out)
{
    Extreme<T> e;
    e.mn = T(0); e.mx = T(0); e.imn = 0; e.imx = 0; e.valid = 0;
    for (uint32_t i = threadIdx.x; i < nparts; i += kBlock) merge_extreme(e, parts[i]);
    const Extreme<T> r = block_extreme(e);
    if (threadIdx.x == 0) {
        Extreme<T> f = r;
        const T first = vol[0];
        if (is_nan_val<T>(first) || !f.valid) { f.mn = first; f.mx = first; }
        out[0] = (double)f.mn;
        out[1] = (double)f.mx;
    }
}

/* ------------------------------------------------------------------------- */
/* host-side launch wrappers                                                   */
/* ------------------------------------------------------------------------- */

/* Emit workgroups that fit on the device at once for this LDS size
 * (persistent grid); remembered per (kernel, LDS size). */
template <typename K>
uint32_t resident_blocks(K kernel, uint32_t lds_bytes)
{
    static std::mutex mu;
    static std::map<std::pair<const void*, uint32_t>, uint32_t> known;
    const std::pair<const void*, uint32_t> key((const void*)kernel, lds_bytes);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = known.find(key);
        if (it != known.end()) return it->second;
    }
    int cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) {
        (void)hipGetLastError();
        cus = 256;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kEmitBlock, lds_bytes) != hipSuccess || per_cu <= 0) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    const uint32_t n = (uint32_t)(cus * per_cu);
    std::lock_guard<std::mutex> lk(mu);
    known[key] = n;
    return n;
}

template <typename T>
hipError_t launch_emit_t(const JobArgs& a, uint32_t lds_bytes, hipStream_t st)
{
    const uint64_t teams = (a.nunits + a.teams - 1) / a.teams;
    switch (emit_mode(a)) {
    case kModeDup: {
        const uint64_t grid = std::min<uint64_t>(teams, resident_blocks(k_emit<T, kModeDup>, lds_bytes));
        hipLaunchKernelGGL((k_emit<T, kModeDup>), dim3((uint32_t)grid), dim3(kEmitBlock), lds_bytes, st, a);
        break;
    }
    case kModePolygon: {
        const uint64_t grid = std::min<uint64_t>(teams, resident_blocks(k_emit<T, kModePolygon>, lds_bytes));
        hipLaunchKernelGGL((k_emit<T, kModePolygon>), dim3((uint32_t)grid), dim3(kEmitBlock), lds_bytes, st, a);
        break;
    }
    default: {
        const uint64_t grid = std::min<uint64_t>(teams, resident_blocks(k_emit<T, kModeVertex>, lds_bytes));
        hipLaunchKernelGGL((k_emit<T, kModeVertex>), dim3((uint32_t)grid), dim3(kEmitBlock), lds_bytes, st, a);
        break;
    }
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_classify_packed_t(const JobArgs& a, hipStream_t st)
{
    constexpr uint32_t P = 4 / sizeof(T);
    const uint64_t items = (uint64_t)a.ny * a.nz * ((a.nx / P + 63) / 64);
    uint64_t blocks = (items + 16 * kWaves - 1) / (16 * kWaves);
    if (blocks > (1u << 20)) blocks = 1u << 20;
    const dim3 g((uint32_t)blocks), b(kBlock);
    switch (a.nlev) {
    case 1: hipLaunchKernelGGL((k_classify_packed<T, 1>), g, b, 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_classify_packed<T, 2>), g, b, 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_classify_packed<T, 3>), g, b, 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_classify_packed<T, 4>), g, b, 0, st, a); break;
    default: hipLaunchKernelGGL((k_classify_packed<T, 8>), g, b, 0, st, a); break;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_classify_t(const JobArgs& a, hipStream_t st)
{
    if constexpr (sizeof(T) <= 2) {
        /* rows of whole dwords: one dword of nodes per lane */
        if (packed_rows<T>(a)) return launch_classify_packed_t<T>(a, st);
    }
    /* one node row per wave: consecutive workgroups sweep the volume in
     * address order (measured faster than waves that take several rows) */
    const uint64_t rows = (uint64_t)a.ny * a.nz;
    uint64_t blocks = (rows + kWaves - 1) / kWaves;
    if (blocks > (1u << 20)) blocks = 1u << 20;
    const dim3 g((uint32_t)blocks), b(kBlock);
    switch (a.nlev) {
    case 1: hipLaunchKernelGGL((k_classify<T, 1>), g, b, 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_classify<T, 2>), g, b, 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_classify<T, 3>), g, b, 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_classify<T, 4>), g, b, 0, st, a); break;
    default: hipLaunchKernelGGL((k_classify<T, 8>), g, b, 0, st, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_classify(int dtype, const JobArgs& a, hipStream_t st)
{
    switch (dtype) {
    case 0: return launch_classify_t<int8_t>(a, st);
    case 1: return launch_classify_t<int16_t>(a, st);
    case 2: return launch_classify_t<int32_t>(a, st);
    case 3: return launch_classify_t<int64_t>(a, st);
    case 4: return launch_classify_t<uint8_t>(a, st);
    case 5: return launch_classify_t<uint16_t>(a, st);
    case 6: return launch_classify_t<uint32_t>(a, st);
    case 7: return launch_classify_t<uint64_t>(a, st);
    case 8: return launch_classify_t<float>(a, st);
    default: return launch_classify_t<double>(a, st);
    }
}

hipError_t launch_count(const JobArgs& a, hipStream_t st)
{
    const uint64_t G = 64u >> a.cshift;
    uint64_t blocks = ((a.nrows + G - 1) / G + kWaves - 1) / kWaves;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_count, dim3((uint32_t)blocks), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

/* The VertexNormal triangle kernel runs best at four waves per SIMD: more
 * waves in flight evict the cache lines its gathers reuse (measured on the
 * 1024^3 turbulence field: 0.609 ms at its register-bound five waves, 0.587
 * at four, 0.617 at three).  A launch with this much dynamic LDS per
 * 256-thread workgroup (the kernel uses none of it) leaves room for four
 * workgroups in a CU's 160 KiB. */
constexpr uint32_t kTriangleLdsPad = 38u << 10;

/* Grid of a flat per-item kernel: up to 8 resident workgroups per CU. */
static uint32_t flat_blocks(uint64_t items)
{
    const uint64_t b = (items + kBlock - 1) / kBlock;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 256 * 8));
}

template <typename T>
hipError_t launch_flat_t(const JobArgs& a, hipStream_t st, hipEvent_t* marks)
{
    const uint64_t nv = a.voff[a.nlev], nt = a.toff[a.nlev];
    if (!a.dup && nv) hipLaunchKernelGGL((k_vertices<T>), dim3(flat_blocks(nv)), dim3(kBlock), 0, st, a);
    if (marks) (void)hipEventRecord(marks[1], st);
    if (nt) {
        switch (emit_mode(a)) {
        case kModeDup: hipLaunchKernelGGL((k_triangles<T, kModeDup>), dim3(flat_blocks(nt)), dim3(kBlock), 0, st, a); break;
        case kModePolygon:
            hipLaunchKernelGGL((k_triangles<T, kModePolygon>), dim3(flat_blocks(nt)), dim3(kBlock), 0, st, a);
            break;
        default: hipLaunchKernelGGL((k_triangles<T, kModeVertex>), dim3(flat_blocks(nt)), dim3(kBlock), kTriangleLdsPad, st, a); break;
        }
    }
    if (marks) (void)hipEventRecord(marks[2], st);
    if (a.vnormal && nv) hipLaunchKernelGGL(k_vertex_normals, dim3(flat_blocks(nv)), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_compact_batch(const JobArgs* jobs, const uint64_t* row_start, uint32_t njobs, uint64_t rows,
                                hipStream_t st)
{
    uint64_t blocks = (rows / 4 + kWaves - 1) / kWaves + 1;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_compact_batch, dim3((uint32_t)blocks), dim3(kBlock), 0, st, jobs, row_start, njobs);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_flat_batch_t(int mode, const JobArgs* jobs, const uint64_t* vstart, const uint64_t* tstart,
                               uint32_t njobs, hipStream_t st)
{
    /* item counts live on the device: persistent grids */
    const dim3 g(256 * 8), b(kBlock);
    if (mode != kModeDup) hipLaunchKernelGGL((k_vertices_batch<T>), g, b, 0, st, jobs, vstart, njobs);
    switch (mode) {
    case kModeDup: hipLaunchKernelGGL((k_triangles_batch<T, kModeDup>), g, b, 0, st, jobs, tstart, njobs); break;
    case kModePolygon: hipLaunchKernelGGL((k_triangles_batch<T, kModePolygon>), g, b, 0, st, jobs, tstart, njobs); break;
    default: hipLaunchKernelGGL((k_triangles_batch<T, kModeVertex>), g, b, 0, st, jobs, tstart, njobs); break;
    }
    if (mode == kModeVertex) hipLaunchKernelGGL(k_vertex_normals_batch, g, b, 0, st, jobs, vstart, njobs);
    return hipGetLastError();
}

hipError_t launch_flat_items(int dtype, const JobArgs& a, hipStream_t st)
{
    switch (dtype) {
    case 0: return launch_flat_t<int8_t>(a, st, nullptr);
    case 1: return launch_flat_t<int16_t>(a, st, nullptr);
    case 2: return launch_flat_t<int32_t>(a, st, nullptr);
    case 3: return launch_flat_t<int64_t>(a, st, nullptr);
    case 4: return launch_flat_t<uint8_t>(a, st, nullptr);
    case 5: return launch_flat_t<uint16_t>(a, st, nullptr);
    case 6: return launch_flat_t<uint32_t>(a, st, nullptr);
    case 7: return launch_flat_t<uint64_t>(a, st, nullptr);
    case 8: return launch_flat_t<float>(a, st, nullptr);
    default: return launch_flat_t<double>(a, st, nullptr);
    }
}

hipError_t launch_flat_batch(int dtype, int mode, const JobArgs* jobs, const uint64_t* vstart, const uint64_t* tstart,
                             uint32_t njobs, hipStream_t st)
{
    switch (dtype) {
    case 0: return launch_flat_batch_t<int8_t>(mode, jobs, vstart, tstart, njobs, st);
    case 1: return launch_flat_batch_t<int16_t>(mode, jobs, vstart, tstart, njobs, st);
    case 2: return launch_flat_batch_t<int32_t>(mode, jobs, vstart, tstart, njobs, st);
    case 3: return launch_flat_batch_t<int64_t>(mode, jobs, vstart, tstart, njobs, st);
    case 4: return launch_flat_batch_t<uint8_t>(mode, jobs, vstart, tstart, njobs, st);
    case 5: return launch_flat_batch_t<uint16_t>(mode, jobs, vstart, tstart, njobs, st);
    case 6: return launch_flat_batch_t<uint32_t>(mode, jobs, vstart, tstart, njobs, st);
    case 7: return launch_flat_batch_t<uint64_t>(mode, jobs, vstart, tstart, njobs, st);
    case 8: return launch_flat_batch_t<float>(mode, jobs, vstart, tstart, njobs, st);
    default: return launch_flat_batch_t<double>(mode, jobs, vstart, tstart, njobs, st);
    }
}

hipError_t launch_flat_emit(int dtype, const JobArgs& a, hipStream_t st, hipEvent_t* marks)
{
    const uint64_t G = 64u >> a.cshift;
    uint64_t blocks = ((a.nrows + G - 1) / G + kWaves - 1) / kWaves;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_compact, dim3((uint32_t)blocks), dim3(kBlock), 0, st, a);
    if (marks) (void)hipEventRecord(marks[0], st);
    switch (dtype) {
    case 0: return launch_flat_t<int8_t>(a, st, marks);
    case 1: return launch_flat_t<int16_t>(a, st, marks);
    case 2: return launch_flat_t<int32_t>(a, st, marks);
    case 3: return launch_flat_t<int64_t>(a, st, marks);
    case 4: return launch_flat_t<uint8_t>(a, st, marks);
    case 5: return launch_flat_t<uint16_t>(a, st, marks);
    case 6: return launch_flat_t<uint32_t>(a, st, marks);
    case 7: return launch_flat_t<uint64_t>(a, st, marks);
    case 8: return launch_flat_t<float>(a, st, marks);
    default: return launch_flat_t<double>(a, st, marks);
    }
}

hipError_t launch_classify_plane(const JobArgs& a, hipStream_t st)
{
    const uint64_t rows = (uint64_t)a.ny * a.nz;
    uint64_t blocks = (rows + kWaves - 1) / kWaves;
    if (blocks > (1u << 20)) blocks = 1u << 20;
    hipLaunchKernelGGL(k_classify_plane, dim3((uint32_t)blocks), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_slice_emit_t(const JobArgs& a, hipStream_t st)
{
    const uint64_t G = 64u >> a.cshift;
    uint64_t blocks = ((a.nrows + G - 1) / G + kWaves - 1) / kWaves;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_compact, dim3((uint32_t)blocks), dim3(kBlock), 0, st, a);
    const uint64_t nt = a.toff[1];
    if (nt) hipLaunchKernelGGL((k_slice_triangles<T>), dim3(flat_blocks(nt)), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_slice_runs_front(const JobArgs& a, int32_t* thr, uint32_t* counts, hipStream_t st)
{
    const uint64_t rows = (uint64_t)a.ny * a.nz, ncr = (uint64_t)(a.ny - 1) * (a.nz - 1);
    hipLaunchKernelGGL(k_slice_thresholds, dim3(flat_blocks(rows)), dim3(kBlock), 0, st, a, thr);
    if (ncr) hipLaunchKernelGGL(k_slice_row_counts, dim3(flat_blocks(ncr)), dim3(kBlock), 0, st, a, thr, counts);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_slice_runs_emit_t(const JobArgs& a, const int32_t* thr, const uint32_t* counts,
                                    const uint64_t* prefix, uint64_t* units, hipStream_t st)
{
    const uint64_t nt = a.toff[1], ncr = (uint64_t)(a.ny - 1) * (a.nz - 1), nunits = (nt + 63) / 64;
    if (!nt) return hipSuccess;
    uint64_t* first = units;
    uint64_t* last = units + nunits;
    hipLaunchKernelGGL(k_slice_unit_rows, dim3(flat_blocks(ncr)), dim3(kBlock), 0, st, ncr, counts, prefix, nt, first,
                       last);
    if (a.cres <= kSliceLdsColors)
        hipLaunchKernelGGL((k_slice_triangles_runs<T, true>), dim3(flat_blocks(nt)), dim3(kBlock), 4 * a.cres, st, a,
                           thr, prefix, first, last);
    else
        hipLaunchKernelGGL((k_slice_triangles_runs<T, false>), dim3(flat_blocks(nt)), dim3(kBlock), 0, st, a, thr,
                           prefix, first, last);
    return hipGetLastError();
}

hipError_t launch_slice_runs_emit(int dtype, const JobArgs& a, const int32_t* thr, const uint32_t* counts,
                                  const uint64_t* prefix, uint64_t* units, hipStream_t st)
{
    switch (dtype) {
    case 0: return launch_slice_runs_emit_t<int8_t>(a, thr, counts, prefix, units, st);
    case 1: return launch_slice_runs_emit_t<int16_t>(a, thr, counts, prefix, units, st);
    case 2: return launch_slice_runs_emit_t<int32_t>(a, thr, counts, prefix, units, st);
    case 3: return launch_slice_runs_emit_t<int64_t>(a, thr, counts, prefix, units, st);
    case 4: return launch_slice_runs_emit_t<uint8_t>(a, thr, counts, prefix, units, st);
    case 5: return launch_slice_runs_emit_t<uint16_t>(a, thr, counts, prefix, units, st);
    case 6: return launch_slice_runs_emit_t<uint32_t>(a, thr, counts, prefix, units, st);
    case 7: return launch_slice_runs_emit_t<uint64_t>(a, thr, counts, prefix, units, st);
    case 8: return launch_slice_runs_emit_t<float>(a, thr, counts, prefix, units, st);
    default: return launch_slice_runs_emit_t<double>(a, thr, counts, prefix, units, st);
    }
}

hipError_t launch_slice_emit(int dtype, const JobArgs& a, hipStream_t st)
{
    switch (dtype) {
    case 0: return launch_slice_emit_t<int8_t>(a, st);
    case 1: return launch_slice_emit_t<int16_t>(a, st);
    case 2: return launch_slice_emit_t<int32_t>(a, st);
    case 3: return launch_slice_emit_t<int64_t>(a, st);
    case 4: return launch_slice_emit_t<uint8_t>(a, st);
    case 5: return launch_slice_emit_t<uint16_t>(a, st);
    case 6: return launch_slice_emit_t<uint32_t>(a, st);
    case 7: return launch_slice_emit_t<uint64_t>(a, st);
    case 8: return launch_slice_emit_t<float>(a, st);
    default: return launch_slice_emit_t<double>(a, st);
    }
}

hipError_t launch_emit(int dtype, const JobArgs& a, uint32_t lds_bytes, hipStream_t st)
{
    switch (dtype) {
    case 0: return launch_emit_t<int8_t>(a, lds_bytes, st);
    case 1: return launch_emit_t<int16_t>(a, lds_bytes, st);
    case 2: return launch_emit_t<int32_t>(a, lds_bytes, st);
    case 3: return launch_emit_t<int64_t>(a, lds_bytes, st);
    case 4: return launch_emit_t<uint8_t>(a, lds_bytes, st);
    case 5: return launch_emit_t<uint16_t>(a, lds_bytes, st);
    case 6: return launch_emit_t<uint32_t>(a, lds_bytes, st);
    case 7: return launch_emit_t<uint64_t>(a, lds_bytes, st);
    case 8: return launch_emit_t<float>(a, lds_bytes, st);
    default: return launch_emit_t<double>(a, lds_bytes, st);
    }
}

template <typename T>
hipError_t launch_classify_batch_t(const JobArgs* jobs, const uint64_t* row_start, uint32_t njobs, uint64_t rows,
                                   uint32_t max_levels, hipStream_t st)
{
    uint64_t blocks = (rows + kWaves - 1) / kWaves;
    if (blocks > 65536) blocks = 65536;
    if (max_levels > 4)
        hipLaunchKernelGGL((k_classify_batch<T, true>), dim3((uint32_t)blocks), dim3(kBlock), 0, st, jobs, row_start,
                           njobs);
    else
        hipLaunchKernelGGL((k_classify_batch<T, false>), dim3((uint32_t)blocks), dim3(kBlock), 0, st, jobs, row_start,
                           njobs);
    return hipGetLastError();
}

hipError_t launch_classify_batch(int dtype, const JobArgs* jobs, const uint64_t* row_start, uint32_t njobs,
                                 uint64_t rows, uint32_t max_levels, hipStream_t st)
{
    switch (dtype) {
    case 0: return launch_classify_batch_t<int8_t>(jobs, row_start, njobs, rows, max_levels, st);
    case 1: return launch_classify_batch_t<int16_t>(jobs, row_start, njobs, rows, max_levels, st);
    case 2: return launch_classify_batch_t<int32_t>(jobs, row_start, njobs, rows, max_levels, st);
    case 3: return launch_classify_batch_t<int64_t>(jobs, row_start, njobs, rows, max_levels, st);
    case 4: return launch_classify_batch_t<uint8_t>(jobs, row_start, njobs, rows, max_levels, st);
    case 5: return launch_classify_batch_t<uint16_t>(jobs, row_start, njobs, rows, max_levels, st);
    case 6: return launch_classify_batch_t<uint32_t>(jobs, row_start, njobs, rows, max_levels, st);
    case 7: return launch_classify_batch_t<uint64_t>(jobs, row_start, njobs, rows, max_levels, st);
    case 8: return launch_classify_batch_t<float>(jobs, row_start, njobs, rows, max_levels, st);
    default: return launch_classify_batch_t<double>(jobs, row_start, njobs, rows, max_levels, st);
    }
}

hipError_t launch_count_batch(const JobArgs* jobs, const uint64_t* row_start, uint32_t njobs, uint64_t rows,
                              hipStream_t st)
{
    uint64_t blocks = (rows / 4 + kWaves - 1) / kWaves + 1;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_count_batch, dim3((uint32_t)blocks), dim3(kBlock), 0, st, jobs, row_start, njobs);
    return hipGetLastError();
}

template <typename T, int M>
hipError_t launch_emit_batch_m(const JobArgs* jobs, const uint64_t* unit_start, uint32_t njobs, uint64_t units,
                               uint32_t lds_bytes, uint32_t teams, hipStream_t st)
{
    const uint64_t grid = std::min<uint64_t>((units + teams - 1) / teams,
                                             resident_blocks(k_emit_batch<T, M>, lds_bytes));
    hipLaunchKernelGGL((k_emit_batch<T, M>), dim3((uint32_t)grid), dim3(kEmitBlock), lds_bytes, st, jobs, unit_start,
                       njobs, teams);
    return hipGetLastError();
}

/* Every pass of a batch group shares the value type T and the emit mode. */
template <typename T>
hipError_t launch_emit_batch_t(int mode, const JobArgs* jobs, const uint64_t* unit_start, uint32_t njobs,
                               uint64_t units, uint32_t lds_bytes, uint32_t teams, hipStream_t st)
{
    switch (mode) {
    case kModeDup: return launch_emit_batch_m<T, kModeDup>(jobs, unit_start, njobs, units, lds_bytes, teams, st);
    case kModePolygon: return launch_emit_batch_m<T, kModePolygon>(jobs, unit_start, njobs, units, lds_bytes, teams, st);
    default: return launch_emit_batch_m<T, kModeVertex>(jobs, unit_start, njobs, units, lds_bytes, teams, st);
    }
}

hipError_t launch_emit_batch(int dtype, int mode, const JobArgs* jobs, const uint64_t* unit_start, uint32_t njobs,
                             uint64_t units, uint32_t lds_bytes, uint32_t teams, hipStream_t st)
{
    switch (dtype) {
    case 0: return launch_emit_batch_t<int8_t>(mode, jobs, unit_start, njobs, units, lds_bytes, teams, st);
    case 1: return launch_emit_batch_t<int16_t>(mode, jobs, unit_start, njobs, units, lds_bytes, teams, st);
    case 2: return launch_emit_batch_t<int32_t>(mode, jobs, unit_start, njobs, units, lds_bytes, teams, st);
    case 3: return launch_emit_batch_t<int64_t>(mode, jobs, unit_start, njobs, units, lds_bytes, teams, st);
    case 4: return launch_emit_batch_t<uint8_t>(mode, jobs, unit_start, njobs, units, lds_bytes, teams, st);
    case 5: return launch_emit_batch_t<uint16_t>(mode, jobs, unit_start, njobs, units, lds_bytes, teams, st);
    case 6: return launch_emit_batch_t<uint32_t>(mode, jobs, unit_start, njobs, units, lds_bytes, teams, st);
    case 7: return launch_emit_batch_t<uint64_t>(mode, jobs, unit_start, njobs, units, lds_bytes, teams, st);
    case 8: return launch_emit_batch_t<float>(mode, jobs, unit_start, njobs, units, lds_bytes, teams, st);
    default: return launch_emit_batch_t<double>(mode, jobs, unit_start, njobs, units, lds_bytes, teams, st);
    }
}


hipError_t launch_scan(const uint32_t* counts, uint64_t n, uint32_t narr, uint64_t* partials,
                       uint64_t* prefix, uint64_t* totals, hipStream_t st)
{
    const uint32_t nseg = (uint32_t)((n + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(k_scan_partials, dim3(nseg * narr), dim3(kBlock), 0, st, counts, n, nseg, partials);
    hipLaunchKernelGGL(k_scan_mid, dim3(1), dim3(kBlock), 0, st, partials, nseg, narr, totals);
    hipLaunchKernelGGL(k_scan_apply, dim3(nseg * narr), dim3(kBlock), 0, st, counts, n, nseg, partials, prefix);
    return hipGetLastError();
}

uint32_t scan_partials_needed(uint64_t n, uint32_t narr)
{
    return (uint32_t)((n + kScanTile - 1) / kScanTile) * narr;
}

template <typename T>
hipError_t launch_minmax_t(const void* vol, uint64_t n, void* scratch, double* out, hipStream_t st)
{
    const uint32_t nparts = kMinMaxParts;
    Extreme<T>* parts = (Extreme<T>*)scratch;
    hipLaunchKernelGGL((k_minmax_partial<T>), dim3(nparts), dim3(kBlock), 0, st, (const T*)vol, n, parts);
    hipLaunchKernelGGL((k_minmax_final<T>), dim3(1), dim3(kBlock), 0, st, (const T*)vol, parts, nparts, out);
    return hipGetLastError();
}

hipError_t launch_minmax(int dtype, const void* vol, uint64_t n, void* scratch, double* out, hipStream_t st)
{
    switch (dtype) {
    case 0: return launch_minmax_t<int8_t>(vol, n, scratch, out, st);
    case 1: return launch_minmax_t<int16_t>(vol, n, scratch, out, st);
    case 2: return launch_minmax_t<int32_t>(vol, n, scratch, out, st);
    case 3: return launch_minmax_t<int64_t>(vol, n, scratch, out, st);
    case 4: return launch_minmax_t<uint8_t>(vol, n, scratch, out, st);
    case 5: return launch_minmax_t<uint16_t>(vol, n, scratch, out, st);
    case 6: return launch_minmax_t<uint32_t>(vol, n, scratch, out, st);
    case 7: return launch_minmax_t<uint64_t>(vol, n, scratch, out, st);
    case 8: return launch_minmax_t<float>(vol, n, scratch, out, st);
    default: return launch_minmax_t<double>(vol, n, scratch, out, st);
    }
}

size_t minmax_scratch_bytes() { return kMinMaxParts * sizeof(Extreme<double>) + 64; }

template <typename T>
hipError_t launch_minmax_gather_t(const void* vol, const uint32_t* conn, uint64_t nconn, void* scratch, double* out,
                                  hipStream_t st)
{
    const uint32_t nparts = kMinMaxParts;
    Extreme<T>* parts = (Extreme<T>*)scratch;
    hipLaunchKernelGGL((k_minmax_gather_partial<T>), dim3(nparts), dim3(kBlock), 0, st, (const T*)vol, conn, nconn,
                       parts);
    hipLaunchKernelGGL((k_minmax_final<T>), dim3(1), dim3(kBlock), 0, st, (const T*)vol, parts, nparts, out);
    return hipGetLastError();
}

hipError_t launch_minmax_gather(int dtype, const void* vol, const uint32_t* conn, uint64_t nconn, void* scratch,
                                double* out, hipStream_t st)
{
    switch (dtype) {
    case 0: return launch_minmax_gather_t<int8_t>(vol, conn, nconn, scratch, out, st);
    case 1: return launch_minmax_gather_t<int16_t>(vol, conn, nconn, scratch, out, st);
    case 2: return launch_minmax_gather_t<int32_t>(vol, conn, nconn, scratch, out, st);
    case 3: return launch_minmax_gather_t<int64_t>(vol, conn, nconn, scratch, out, st);
    case 4: return launch_minmax_gather_t<uint8_t>(vol, conn, nconn, scratch, out, st);
    case 5: return launch_minmax_gather_t<uint16_t>(vol, conn, nconn, scratch, out, st);
    case 6: return launch_minmax_gather_t<uint32_t>(vol, conn, nconn, scratch, out, st);
    case 7: return launch_minmax_gather_t<uint64_t>(vol, conn, nconn, scratch, out, st);
    case 8: return launch_minmax_gather_t<float>(vol, conn, nconn, scratch, out, st);
    default: return launch_minmax_gather_t<double>(vol, conn, nconn, scratch, out, st);
    }
}

}  // namespace kmc
