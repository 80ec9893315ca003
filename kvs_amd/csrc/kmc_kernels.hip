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
 * Arithmetic follows Source/Core/Visualization/Mapper/MarchingCubes.cpp bit for
 * bit: the comparison double(v) > isolevel (:299-316), the FP64 clamp / ratio /
 * lerp of interpolate_vertex (:326-356), the float (p + min) * scale of the
 * callers (:169-173, :382-386) and the float cross product of Vector3::cross
 * (Source/Core/Matrix/Vector3.h:479-485).  The file is compiled with
 * -ffp-contract=off so that no multiply-add is fused.
 */
#include <hip/hip_runtime.h>
#include "kmc_job.h"
#include "mc_tables.h"

#pragma clang fp contract(off)

namespace kmc {

__constant__ uint64_t d_tri[256] = { KMC_TABLE_WORDS };

/* Owner node offset and axis of each cell edge (see kEdgeNode). */
__constant__ uint8_t d_edge_node[12][4] = {
    {0, 0, 0, 0}, {1, 0, 0, 1}, {0, 1, 0, 0}, {0, 0, 0, 1},
    {0, 0, 1, 0}, {1, 0, 1, 1}, {0, 1, 1, 0}, {0, 0, 1, 1},
    {0, 0, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}, {0, 1, 0, 2}
};
/* Table-orientation endpoints of each edge as corner offsets (x,y,z) packed
 * in 3 bits each: bit0 x, bit1 y, bit2 z. */
__constant__ uint8_t d_edge_ends[12][2] = {
    {0, 1}, {1, 3}, {3, 2}, {2, 0}, {4, 5}, {5, 7}, {7, 6}, {6, 4}, {0, 4}, {1, 5}, {3, 7}, {2, 6}
};

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
    double va = (double)vol[ia];
    double vb = (double)vol[ib];
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
/* LDS views                                                                   */
/* ------------------------------------------------------------------------- */

struct Views {
    uint64_t* bits;     // [3][nrb][ncp][nlev]
    uint32_t* pv;       // [2][R+1][NC][nlev]
    uint64_t* tm;       // [R][NC][nlev][3]
    uint32_t* pt;       // [R][NC][nlev]
    uint32_t* scan;     // [kBlock + 64]
    uint32_t* tot;      // [2][kMaxLevels]
    uint32_t* queue;    // [queue_cap]
    uint32_t nrb, ncp;
};

__device__ __forceinline__ Views make_views(const JobArgs& a, uint8_t* smem)
{
    const LdsLayout L = lds_layout(a.R, a.NC, a.nlev, a.queue_cap);
    Views v;
    v.bits = (uint64_t*)(smem + L.off_bits);
    v.pv = (uint32_t*)(smem + L.off_pv);
    v.tm = (uint64_t*)(smem + L.off_tm);
    v.pt = (uint32_t*)(smem + L.off_pt);
    v.scan = (uint32_t*)(smem + L.off_scan);
    v.tot = (uint32_t*)(smem + L.off_tot);
    v.queue = (uint32_t*)(smem + L.off_queue);
    v.nrb = L.nrb;
    v.ncp = L.ncp;
    return v;
}

/* Ring slot of node slice s (s >= -1). */
__device__ __forceinline__ uint32_t bslot(int s) { return (uint32_t)(s + 1) % 3u; }
__device__ __forceinline__ uint32_t pslot(int s) { return (uint32_t)(s + 1) & 1u; }

/* Chunk mask of slice-slot sl, LDS row r (row y = y0 - 1 + r), chunk c. */
__device__ __forceinline__ uint64_t B(const JobArgs& a, const Views& v, uint32_t sl, uint32_t r, uint32_t c, uint32_t l)
{
    return v.bits[((sl * v.nrb + r) * v.ncp + c) * a.nlev + l];
}

/* Mask of the nodes x < limit inside chunk c. */
__device__ __forceinline__ uint64_t valid_mask(uint32_t c, uint32_t limit)
{
    const uint32_t x0 = c * 64u;
    if (limit <= x0) return 0ull;
    const uint32_t n = limit - x0;
    return n >= 64u ? ~0ull : ((1ull << n) - 1ull);
}

/* Bits of chunk c shifted so that bit i holds node x+1. */
__device__ __forceinline__ uint64_t shift_next(const JobArgs& a, const Views& v, uint32_t sl, uint32_t r, uint32_t c, uint32_t l)
{
    return (B(a, v, sl, r, c, l) >> 1) | (B(a, v, sl, r, c + 1, l) << 63);
}

struct Crossings { uint64_t x, y, z; };

/* Crossing edges owned by the nodes of row y (LDS row r) of slice z in chunk c. */
__device__ __forceinline__ Crossings crossings(const JobArgs& a, const Views& v, int z, int y,
                                               uint32_t r, uint32_t c, uint32_t l)
{
    Crossings k;
    const uint32_t s0 = bslot(z);
    const uint64_t b = B(a, v, s0, r, c, l);
    k.x = (b ^ shift_next(a, v, s0, r, c, l)) & valid_mask(c, a.nx - 1);
    k.y = (y + 1 < (int)a.ny) ? ((b ^ B(a, v, s0, r + 1, c, l)) & valid_mask(c, a.nx)) : 0ull;
    k.z = (z + 1 < (int)a.nz) ? ((b ^ B(a, v, bslot(z + 1), r, c, l)) & valid_mask(c, a.nx)) : 0ull;
    return k;
}

/* Active-cell mask of the cells of row y (LDS row r) at cell slice zc, chunk c. */
__device__ __forceinline__ uint64_t active_cells(const JobArgs& a, const Views& v, int zc,
                                                 uint32_t r, uint32_t c, uint32_t l, uint64_t corner[8])
{
    const uint32_t s0 = bslot(zc), s1 = bslot(zc + 1);
    corner[0] = B(a, v, s0, r, c, l);
    corner[1] = shift_next(a, v, s0, r, c, l);
    corner[3] = B(a, v, s0, r + 1, c, l);
    corner[2] = shift_next(a, v, s0, r + 1, c, l);
    corner[4] = B(a, v, s1, r, c, l);
    corner[5] = shift_next(a, v, s1, r, c, l);
    corner[7] = B(a, v, s1, r + 1, c, l);
    corner[6] = shift_next(a, v, s1, r + 1, c, l);
    uint64_t o = 0, n = ~0ull;
#pragma unroll
    for (int k = 0; k < 8; ++k) { o |= corner[k]; n &= corner[k]; }
    return (o & ~n) & valid_mask(c, a.nx - 1);
}

__device__ __forceinline__ uint32_t cube_bits(const uint64_t corner[8], uint32_t i)
{
    uint32_t ci = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) ci |= (uint32_t)((corner[k] >> i) & 1ull) << k;
    return ci;
}

/* Cube index of cell (cx, cy, cz) from the bit ring (y0: first band row). */
__device__ __forceinline__ uint32_t cube_index_at(const JobArgs& a, const Views& v, int y0,
                                                  uint32_t cx, int cy, int cz, uint32_t l)
{
    uint32_t ci = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t dx = (k == 1 || k == 2 || k == 5 || k == 6) ? 1u : 0u;
        const uint32_t dy = (k == 2 || k == 3 || k == 6 || k == 7) ? 1u : 0u;
        const uint32_t dz = (k >= 4) ? 1u : 0u;
        const uint32_t x = cx + dx;
        const uint32_t r = (uint32_t)(cy + (int)dy - y0 + 1);
        const uint64_t m = B(a, v, bslot(cz + (int)dz), r, x >> 6, l);
        ci |= (uint32_t)((m >> (x & 63u)) & 1ull) << k;
    }
    return ci;
}

/* ------------------------------------------------------------------------- */
/* block scan of up to kMaxLevels u32 lanes                                    */
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

/* Exclusive scan across the workgroup of val[l] (l < nlev); returns the
 * exclusive prefix in val and the totals in total.  Uses sh[kMaxLevels*kWaves]. */
__device__ __forceinline__ void block_scan(uint32_t nlev, uint32_t val[kMaxLevels], uint32_t total[kMaxLevels], uint32_t* sh)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl[kMaxLevels];
#pragma unroll
    for (int l = 0; l < kMaxLevels; ++l) {
        incl[l] = 0;
        if (l < (int)nlev) {
            incl[l] = wave_incl_scan(val[l]);
            if (lane == 63) sh[l * kWaves + wave] = incl[l];
        }
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < kMaxLevels; ++l) {
        if (l < (int)nlev) {
            uint32_t before = 0, all = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t t = sh[l * kWaves + w];
                if (w < wave) before += t;
                all += t;
            }
            val[l] = before + incl[l] - val[l];
            total[l] = all;
        }
    }
    __syncthreads();
}

/* ------------------------------------------------------------------------- */
/* slice loading                                                               */
/* ------------------------------------------------------------------------- */

/* Classify node rows y0-1 .. y0+R+1 of slice s into chunk masks of its ring
 * slot.  Nodes outside the volume classify as all-zero. */
template <typename T>
__device__ void load_slice(const JobArgs& a, const Views& v, const T* vol, int s, int y0)
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t sl = bslot(s);
    const uint32_t total = v.nrb * a.NC;
    const bool slice_ok = s >= 0 && s < (int)a.nz;
    for (uint32_t j = wave; j < total; j += kWaves) {
        const uint32_t r = j / a.NC, c = j - r * a.NC;
        const int y = y0 - 1 + (int)r;
        const uint32_t x = c * 64u + lane;
        uint32_t bits = 0;
        if (slice_ok && y >= 0 && y < (int)a.ny && x < a.nx) {
            const T val = vol[((uint64_t)s * a.ny + (uint64_t)y) * a.nx + x];
            bits = Classify<T>::bits(val, a);
        }
        uint64_t* dst = v.bits + ((sl * v.nrb + r) * v.ncp + c) * a.nlev;
#pragma unroll
        for (int l = 0; l < kMaxLevels; ++l) {
            if (l < (int)a.nlev) {
                const uint64_t m = __ballot((bits >> l) & 1u);
                if (lane == 0) dst[l] = m;
            }
        }
    }
    for (uint32_t j = threadIdx.x; j < v.nrb * a.nlev; j += kBlock) {
        const uint32_t r = j / a.nlev, l = j - r * a.nlev;
        v.bits[((sl * v.nrb + r) * v.ncp + a.NC) * a.nlev + l] = 0ull;
    }
}

/* ------------------------------------------------------------------------- */
/* per-tile prefixes                                                           */
/* ------------------------------------------------------------------------- */

/* Vertex chunk prefixes of node slice z for band rows 0..R-1 (relative to
 * tile (z, b)) and for row R (relative to tile (z, b+1)); band totals go to
 * tot[0][l].  Writes pv slot of z. */
__device__ void vertex_prefixes(const JobArgs& a, const Views& v, int z, int y0)
{
    uint32_t* pv = v.pv + pslot(z) * (a.R + 1) * a.NC * a.nlev;
    const uint32_t nall = (a.R + 1) * a.NC;
    for (uint32_t j = threadIdx.x; j < nall; j += kBlock) {
        const uint32_t k = j / a.NC, c = j - k * a.NC;
        const int y = y0 + (int)k;
        for (uint32_t l = 0; l < a.nlev; ++l) {
            uint32_t cnt = 0;
            if (y < (int)a.ny) {
                const Crossings e = crossings(a, v, z, y, k + 1, c, l);
                cnt = __popcll(e.x) + __popcll(e.y) + __popcll(e.z);
            }
            pv[j * a.nlev + l] = cnt;
        }
    }
    __syncthreads();
    /* two segments: band rows [0, R*NC) and row R [R*NC, (R+1)*NC) */
    for (int seg = 0; seg < 2; ++seg) {
        const uint32_t j0 = seg == 0 ? 0u : a.R * a.NC;
        const uint32_t nj = seg == 0 ? a.R * a.NC : a.NC;
        const uint32_t q = (nj + kBlock - 1) / kBlock;
        const uint32_t jb = j0 + threadIdx.x * q;
        uint32_t val[kMaxLevels], total[kMaxLevels];
#pragma unroll
        for (int l = 0; l < kMaxLevels; ++l) { val[l] = 0; total[l] = 0; }
        for (uint32_t t = 0; t < q; ++t) {
            const uint32_t j = jb + t;
            if (j < j0 + nj) {
#pragma unroll
                for (int l = 0; l < kMaxLevels; ++l)
                    if (l < (int)a.nlev) val[l] += pv[j * a.nlev + l];
            }
        }
        block_scan(a.nlev, val, total, v.scan);
        for (uint32_t t = 0; t < q; ++t) {
            const uint32_t j = jb + t;
            if (j < j0 + nj) {
#pragma unroll
                for (int l = 0; l < kMaxLevels; ++l) {
                    if (l < (int)a.nlev) {
                        const uint32_t c = pv[j * a.nlev + l];
                        pv[j * a.nlev + l] = val[l];
                        val[l] += c;
                    }
                }
            }
        }
        if (seg == 0 && threadIdx.x == 0) {
#pragma unroll
            for (int l = 0; l < kMaxLevels; ++l)
                if (l < (int)a.nlev) v.tot[l] = total[l];
        }
        __syncthreads();
    }
}

/* Triangle counts of the cells of band rows at cell slice zc = z - 1: bit
 * planes of the per-cell counts into tm, chunk prefixes into pt, totals into
 * tot[1][l]. */
__device__ void triangle_prefixes(const JobArgs& a, const Views& v, int z, int y0)
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const int zc = z - 1;
    const uint32_t nj = a.R * a.NC;
    for (uint32_t j = wave; j < nj; j += kWaves) {
        const uint32_t k = j / a.NC, c = j - k * a.NC;
        const int y = y0 + (int)k;
        const bool row_ok = zc >= 0 && y + 1 < (int)a.ny;
        for (uint32_t l = 0; l < a.nlev; ++l) {
            uint32_t nt = 0;
            if (row_ok) {
                uint64_t corner[8];
                const uint64_t act = active_cells(a, v, zc, k + 1, c, l, corner);
                if ((act >> lane) & 1ull) nt = (uint32_t)(d_tri[cube_bits(corner, lane)] >> 60);
            }
            const uint64_t m0 = __ballot(nt & 1u), m1 = __ballot((nt >> 1) & 1u), m2 = __ballot((nt >> 2) & 1u);
            if (lane == 0) {
                uint64_t* t = v.tm + (j * a.nlev + l) * 3;
                t[0] = m0; t[1] = m1; t[2] = m2;
                v.pt[j * a.nlev + l] = __popcll(m0) + 2u * __popcll(m1) + 4u * __popcll(m2);
            }
        }
    }
    __syncthreads();
    const uint32_t q = (nj + kBlock - 1) / kBlock;
    const uint32_t jb = threadIdx.x * q;
    uint32_t val[kMaxLevels], total[kMaxLevels];
#pragma unroll
    for (int l = 0; l < kMaxLevels; ++l) { val[l] = 0; total[l] = 0; }
    for (uint32_t t = 0; t < q; ++t) {
        const uint32_t j = jb + t;
        if (j < nj) {
#pragma unroll
            for (int l = 0; l < kMaxLevels; ++l)
                if (l < (int)a.nlev) val[l] += v.pt[j * a.nlev + l];
        }
    }
    block_scan(a.nlev, val, total, v.scan);
    for (uint32_t t = 0; t < q; ++t) {
        const uint32_t j = jb + t;
        if (j < nj) {
#pragma unroll
            for (int l = 0; l < kMaxLevels; ++l) {
                if (l < (int)a.nlev) {
                    const uint32_t c = v.pt[j * a.nlev + l];
                    v.pt[j * a.nlev + l] = val[l];
                    val[l] += c;
                }
            }
        }
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int l = 0; l < kMaxLevels; ++l)
            if (l < (int)a.nlev) v.tot[kMaxLevels + l] = total[l];
    }
    __syncthreads();
}

/* ------------------------------------------------------------------------- */
/* emission                                                                    */
/* ------------------------------------------------------------------------- */

/* Rank of the vertex of (node x in band row k (0..R) of slice s, axis)
 * within its tile. */
__device__ __forceinline__ uint32_t vertex_rank(const JobArgs& a, const Views& v, int s, int y0,
                                                uint32_t x, uint32_t k, uint32_t axis, uint32_t l)
{
    const uint32_t c = x >> 6, i = x & 63u;
    const Crossings e = crossings(a, v, s, y0 + (int)k, k + 1, c, l);
    const uint64_t lo = (1ull << i) - 1ull;
    uint32_t r = v.pv[((pslot(s) * (a.R + 1) + k) * a.NC + c) * a.nlev + l];
    r += __popcll(e.x & lo) + __popcll(e.y & lo) + __popcll(e.z & lo);
    if (axis >= 1) r += (uint32_t)((e.x >> i) & 1ull);
    if (axis >= 2) r += (uint32_t)((e.y >> i) & 1ull);
    return r;
}

/* Coordinates of the shared vertex on the edge of node (x,y,z) along axis,
 * interpolated low -> high as calculate_isopoints does (:406,420,434). */
template <typename T>
__device__ __forceinline__ void node_edge_coords(const JobArgs& a, const T* vol, double iso,
                                                 uint32_t x, uint32_t y, uint32_t z, uint32_t axis, float out[3])
{
    interpolate<T>(a, vol, iso, x, y, z, x + (axis == 0), y + (axis == 1), z + (axis == 2), out);
}

/* Coordinates of local edge e of cell (cx,cy,cz) as a shared vertex. */
template <typename T>
__device__ __forceinline__ void cell_edge_coords(const JobArgs& a, const T* vol, double iso,
                                                 uint32_t cx, uint32_t cy, uint32_t cz, uint32_t e, float out[3])
{
    node_edge_coords<T>(a, vol, iso, cx + d_edge_node[e][0], cy + d_edge_node[e][1], cz + d_edge_node[e][2],
                        d_edge_node[e][3], out);
}

/* VertexNormal of the vertex on (x,y,z,axis): the float sum, in triangle order,
 * of the unnormalized face normals of the triangles that use it -- the order of
 * calculate_normals_on_vertex (:569-605).  Incident cells are visited in cell
 * scan order and triangles in table order. */
template <typename T>
__device__ void vertex_normal(const JobArgs& a, const Views& v, const T* vol, double iso, int y0,
                              uint32_t x, uint32_t y, uint32_t z, uint32_t axis, uint32_t l, float n[3])
{
    n[0] = 0.0f; n[1] = 0.0f; n[2] = 0.0f;
    /* (dx, dy, dz, local edge) of the four cells around the edge, scan order */
    int cd[4][4];
    if (axis == 0) {
        const int t[4][4] = { {0, -1, -1, 6}, {0, 0, -1, 4}, {0, -1, 0, 2}, {0, 0, 0, 0} };
        for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) cd[i][j] = t[i][j];
    } else if (axis == 1) {
        const int t[4][4] = { {-1, 0, -1, 5}, {0, 0, -1, 7}, {-1, 0, 0, 1}, {0, 0, 0, 3} };
        for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) cd[i][j] = t[i][j];
    } else {
        const int t[4][4] = { {-1, -1, 0, 10}, {0, -1, 0, 11}, {-1, 0, 0, 9}, {0, 0, 0, 8} };
        for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) cd[i][j] = t[i][j];
    }
    for (int q = 0; q < 4; ++q) {
        const int cx = (int)x + cd[q][0], cy = (int)y + cd[q][1], cz = (int)z + cd[q][2];
        if (cx < 0 || cy < 0 || cz < 0) continue;
        if (cx + 1 >= (int)a.nx || cy + 1 >= (int)a.ny || cz + 1 >= (int)a.nz) continue;
        const uint32_t ci = cube_index_at(a, v, y0, (uint32_t)cx, cy, cz, l);
        const uint64_t w = d_tri[ci];
        const int nt = (int)(w >> 60);
        const uint32_t me = (uint32_t)cd[q][3];
        for (int t = 0; t < nt; ++t) {
            const uint32_t e0 = (uint32_t)((w >> (12 * t)) & 0xF);
            const uint32_t e1 = (uint32_t)((w >> (12 * t + 8)) & 0xF);
            const uint32_t e2 = (uint32_t)((w >> (12 * t + 4)) & 0xF);
            if (e0 != me && e1 != me && e2 != me) continue;
            float p0[3], p1[3], p2[3], fn[3];
            cell_edge_coords<T>(a, vol, iso, cx, cy, cz, e0, p0);
            cell_edge_coords<T>(a, vol, iso, cx, cy, cz, e1, p1);
            cell_edge_coords<T>(a, vol, iso, cx, cy, cz, e2, p2);
            face_normal(p0, p1, p2, fn);
            n[0] = n[0] + fn[0];
            n[1] = n[1] + fn[1];
            n[2] = n[2] + fn[2];
        }
    }
}

/* Emit the vertices of tile (z, b) for level l. */
template <typename T>
__device__ void emit_vertices(const JobArgs& a, const Views& v, const T* vol, int z, int y0,
                              uint32_t l, uint64_t base)
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t nv = v.tot[l];
    const double iso = iso_of(a, (int)l);
    const uint32_t nj = a.R * a.NC;
    for (uint32_t w0 = 0; w0 < nv; w0 += a.queue_cap) {
        /* fill: queue[rank - w0] = x | k << 20 | axis << 28 */
        for (uint32_t j = wave; j < nj; j += kWaves) {
            const uint32_t k = j / a.NC, c = j - k * a.NC;
            const int y = y0 + (int)k;
            if (y >= (int)a.ny) continue;
            const Crossings e = crossings(a, v, z, y, k + 1, c, l);
            const uint64_t any = e.x | e.y | e.z;
            if (!((any >> lane) & 1ull)) continue;
            const uint64_t lo = (1ull << lane) - 1ull;
            uint32_t r = v.pv[((pslot(z) * (a.R + 1) + k) * a.NC + c) * a.nlev + l]
                         + __popcll(e.x & lo) + __popcll(e.y & lo) + __popcll(e.z & lo);
            const uint32_t x = c * 64u + lane;
            const uint64_t fx = (e.x >> lane) & 1ull, fy = (e.y >> lane) & 1ull, fz = (e.z >> lane) & 1ull;
            if (fx) { if (r >= w0 && r - w0 < a.queue_cap) v.queue[r - w0] = x | (k << 20) | (0u << 28); ++r; }
            if (fy) { if (r >= w0 && r - w0 < a.queue_cap) v.queue[r - w0] = x | (k << 20) | (1u << 28); ++r; }
            if (fz) { if (r >= w0 && r - w0 < a.queue_cap) v.queue[r - w0] = x | (k << 20) | (2u << 28); ++r; }
        }
        __syncthreads();
        const uint32_t nq = (nv - w0) < a.queue_cap ? (nv - w0) : a.queue_cap;
        for (uint32_t q = threadIdx.x; q < nq; q += kBlock) {
            const uint32_t e = v.queue[q];
            const uint32_t x = e & 0xFFFFFu, k = (e >> 20) & 0xFFu, axis = e >> 28;
            const uint32_t y = (uint32_t)y0 + k;
            const uint64_t vid = base + w0 + q;
            float p[3];
            node_edge_coords<T>(a, vol, iso, x, y, (uint32_t)z, axis, p);
            float* co = a.coords[0];
            float* no = a.normals[0];
#pragma unroll
            for (int m = 1; m < kMaxLevels; ++m)
                if (m == (int)l) { co = a.coords[m]; no = a.normals[m]; }
            uint64_t capv = a.cap_vertices[0];
#pragma unroll
            for (int m = 1; m < kMaxLevels; ++m)
                if (m == (int)l) capv = a.cap_vertices[m];
            if (vid < capv) {
                co[3 * vid + 0] = p[0];
                co[3 * vid + 1] = p[1];
                co[3 * vid + 2] = p[2];
                if (a.vnormal) {
                    float n[3];
                    vertex_normal<T>(a, v, vol, iso, y0, x, y, (uint32_t)z, axis, l, n);
                    no[3 * vid + 0] = n[0];
                    no[3 * vid + 1] = n[1];
                    no[3 * vid + 2] = n[2];
                }
            }
        }
        __syncthreads();
    }
}

/* Emit the triangles of the cells of band rows at cell slice z-1, level l. */
template <typename T>
__device__ void emit_triangles(const JobArgs& a, const Views& v, const T* vol, int z, int y0,
                               uint32_t l, uint64_t tbase, const uint64_t vbase[4])
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t nt = v.tot[kMaxLevels + l];
    const double iso = iso_of(a, (int)l);
    const uint32_t nj = a.R * a.NC;
    const int zc = z - 1;
    float* co = a.coords[0];
    float* no = a.normals[0];
    uint32_t* cn = a.conn[0];
    uint64_t capv = a.cap_vertices[0], capt = a.cap_triangles[0];
#pragma unroll
    for (int m = 1; m < kMaxLevels; ++m)
        if (m == (int)l) {
            co = a.coords[m]; no = a.normals[m]; cn = a.conn[m];
            capv = a.cap_vertices[m]; capt = a.cap_triangles[m];
        }
    for (uint32_t w0 = 0; w0 < nt; w0 += a.queue_cap) {
        /* fill: queue[rank - w0] = x | k << 20 | t << 26 */
        for (uint32_t j = wave; j < nj; j += kWaves) {
            const uint32_t k = j / a.NC, c = j - k * a.NC;
            const uint64_t* tmj = v.tm + (j * a.nlev + l) * 3;
            const uint64_t m0 = tmj[0], m1 = tmj[1], m2 = tmj[2];
            const uint32_t cnt = (uint32_t)((m0 >> lane) & 1ull) + 2u * (uint32_t)((m1 >> lane) & 1ull)
                               + 4u * (uint32_t)((m2 >> lane) & 1ull);
            if (cnt == 0) continue;
            const uint64_t lo = (1ull << lane) - 1ull;
            const uint32_t r0 = v.pt[j * a.nlev + l] + __popcll(m0 & lo) + 2u * __popcll(m1 & lo) + 4u * __popcll(m2 & lo);
            const uint32_t x = c * 64u + lane;
            for (uint32_t t = 0; t < cnt; ++t) {
                const uint32_t r = r0 + t;
                if (r >= w0 && r - w0 < a.queue_cap) v.queue[r - w0] = x | (k << 20) | (t << 26);
            }
        }
        __syncthreads();
        const uint32_t nq = (nt - w0) < a.queue_cap ? (nt - w0) : a.queue_cap;
        for (uint32_t q = threadIdx.x; q < nq; q += kBlock) {
            const uint32_t e = v.queue[q];
            const uint32_t x = e & 0xFFFFFu, k = (e >> 20) & 0x3Fu, t = e >> 26;
            const int y = y0 + (int)k;
            const uint64_t tri = tbase + w0 + q;
            if (tri >= capt) continue;
            const uint32_t ci = cube_index_at(a, v, y0, x, y, zc, l);
            const uint64_t w = d_tri[ci];
            const uint32_t ed[3] = { (uint32_t)((w >> (12 * t)) & 0xF), (uint32_t)((w >> (12 * t + 8)) & 0xF),
                                     (uint32_t)((w >> (12 * t + 4)) & 0xF) };
            if (a.dup) {
                /* extract_surfaces_with_duplication: table orientation (:206-228) */
                float p[3][3];
                for (int j = 0; j < 3; ++j) {
                    const uint32_t e0 = d_edge_ends[ed[j]][0], e1 = d_edge_ends[ed[j]][1];
                    interpolate<T>(a, vol, iso, x + (e0 & 1u), (uint32_t)y + ((e0 >> 1) & 1u), (uint32_t)zc + (e0 >> 2),
                                   x + (e1 & 1u), (uint32_t)y + ((e1 >> 1) & 1u), (uint32_t)zc + (e1 >> 2), p[j]);
                }
                float n[3];
                face_normal(p[0], p[1], p[2], n);
                if (3 * tri + 2 < capv) {
                    for (int j = 0; j < 3; ++j)
                        for (int m = 0; m < 3; ++m) co[9 * tri + 3 * j + m] = p[j][m];
                }
                no[3 * tri + 0] = n[0];
                no[3 * tri + 1] = n[1];
                no[3 * tri + 2] = n[2];
            } else {
                /* connect_isopoints: vertex ids of the three edges (:493-514) */
                uint32_t id[3];
                for (int j = 0; j < 3; ++j) {
                    const uint32_t dx = d_edge_node[ed[j]][0], dy = d_edge_node[ed[j]][1];
                    const uint32_t dz = d_edge_node[ed[j]][2], ax = d_edge_node[ed[j]][3];
                    const uint32_t kk = k + dy;
                    const int s = zc + (int)dz;
                    const uint32_t rank = vertex_rank(a, v, s, y0, x + dx, kk, ax, l);
                    const uint64_t b = vbase[(dz << 1) | (kk == a.R ? 1u : 0u)];
                    id[j] = (uint32_t)(b + rank);
                }
                cn[3 * tri + 0] = id[0];
                cn[3 * tri + 1] = id[1];
                cn[3 * tri + 2] = id[2];
                if (!a.vnormal) {
                    /* calculate_normals_on_polygon (:531-559) */
                    float p[3][3], n[3];
                    for (int j = 0; j < 3; ++j)
                        cell_edge_coords<T>(a, vol, iso, x, (uint32_t)y, (uint32_t)zc, ed[j], p[j]);
                    face_normal(p[0], p[1], p[2], n);
                    no[3 * tri + 0] = n[0];
                    no[3 * tri + 1] = n[1];
                    no[3 * tri + 2] = n[2];
                }
            }
        }
        __syncthreads();
    }
}

/* ------------------------------------------------------------------------- */
/* the z-marching kernel                                                       */
/* ------------------------------------------------------------------------- */

/* One workgroup per (slab, band).  kCount: write per-tile vertex / triangle
 * counts; otherwise emit the outputs at the prefixes of a.prefix. */
template <typename T, bool kCount>
__global__ __launch_bounds__(kBlock) void k_march(JobArgs a)
{
    extern __shared__ __align__(16) uint8_t smem[];
    const Views v = make_views(a, smem);
    const T* vol = (const T*)a.values;
    const uint32_t band = blockIdx.x % a.nb;
    const uint32_t slab = blockIdx.x / a.nb;
    const int zb = (int)(slab * a.slab);
    const int ze = (int)((slab + 1) * a.slab < a.nz ? (slab + 1) * a.slab : a.nz);
    const int y0 = (int)(band * a.R);

    load_slice<T>(a, v, vol, zb - 1, y0);
    load_slice<T>(a, v, vol, zb, y0);
    __syncthreads();
    if (!kCount && zb > 0) {
        vertex_prefixes(a, v, zb - 1, y0);
    }
    for (int z = zb; z < ze; ++z) {
        load_slice<T>(a, v, vol, z + 1, y0);
        __syncthreads();
        vertex_prefixes(a, v, z, y0);
        triangle_prefixes(a, v, z, y0);
        const uint64_t tile = (uint64_t)z * a.nb + band;
        if (kCount) {
            if (threadIdx.x < a.nlev) {
                const uint32_t l = threadIdx.x;
                a.counts[(uint64_t)(l * 2 + 0) * a.ntiles + tile] = v.tot[l];
                a.counts[(uint64_t)(l * 2 + 1) * a.ntiles + tile] = v.tot[kMaxLevels + l];
            }
        } else {
            for (uint32_t l = 0; l < a.nlev; ++l) {
                const uint64_t* pre_v = a.prefix + (uint64_t)(l * 2 + 0) * a.ntiles;
                const uint64_t* pre_t = a.prefix + (uint64_t)(l * 2 + 1) * a.ntiles;
                if (!a.dup) emit_vertices<T>(a, v, vol, z, y0, l, pre_v[tile]);
                uint64_t vb[4] = {0, 0, 0, 0};
                if (!a.dup && z > 0) {
                    const uint64_t tprev = (uint64_t)(z - 1) * a.nb + band;
                    vb[0] = pre_v[tprev];
                    vb[1] = (band + 1 < a.nb) ? pre_v[tprev + 1] : 0;
                    vb[2] = pre_v[tile];
                    vb[3] = (band + 1 < a.nb) ? pre_v[tile + 1] : 0;
                }
                if (z > 0) emit_triangles<T>(a, v, vol, z, y0, l, pre_t[tile], vb);
            }
        }
        __syncthreads();
    }
}

/* ------------------------------------------------------------------------- */
/* tile-count scan                                                             */
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

/* Exclusive prefix of each count array; the grand total of array arr goes
 * to totals[arr]. */
__global__ __launch_bounds__(kBlock) void k_scan_apply(const uint32_t* counts, uint64_t n, uint32_t nseg,
                                                       const uint64_t* partials, uint64_t* prefix, uint64_t* totals)
{
    __shared__ uint64_t sh[kWaves];
    __shared__ uint64_t base_sh;
    const uint32_t arr = blockIdx.x / nseg, seg = blockIdx.x - arr * nseg;
    const uint32_t* src = counts + (uint64_t)arr * n;
    uint64_t* dst = prefix + (uint64_t)arr * n;
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (uint32_t s = 0; s < seg; ++s) t += partials[(uint64_t)arr * nseg + s];
        base_sh = t;
        if (seg + 1 == nseg) totals[arr] = t + partials[(uint64_t)arr * nseg + seg];
    }
    const uint64_t b = (uint64_t)seg * kScanTile + (uint64_t)threadIdx.x * kScanPer;
    uint32_t vals[kScanPer];
    uint64_t s = 0;
#pragma unroll
    for (uint32_t i = 0; i < kScanPer; ++i) {
        const uint64_t g = b + i;
        vals[i] = g < n ? src[g] : 0u;
        s += vals[i];
    }
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
    uint64_t run = base_sh;
    for (int w = 0; w < wave; ++w) run += sh[w];
    run += incl - s;
#pragma unroll
    for (uint32_t i = 0; i < kScanPer; ++i) {
        const uint64_t g = b + i;
        if (g < n) dst[g] = run;
        run += vals[i];
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
__global__ __launch_bounds__(kBlock) void k_minmax_partial(const T* vol, uint64_t n, Extreme<T>* parts)
{
    __shared__ Extreme<T> sh[kBlock];
    Extreme<T> e;
    e.mn = T(0); e.mx = T(0); e.imn = 0; e.imx = 0; e.valid = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        const T v = vol[i];
        if (is_nan_val<T>(v)) continue;
        if (!e.valid) { e.mn = v; e.mx = v; e.imn = i; e.imx = i; e.valid = 1; }
        else {
            if (v < e.mn) { e.mn = v; e.imn = i; }
            if (v > e.mx) { e.mx = v; e.imx = i; }
        }
    }
    sh[threadIdx.x] = e;
    __syncthreads();
    for (uint32_t s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            Extreme<T> a = sh[threadIdx.x];
            merge_extreme(a, sh[threadIdx.x + s]);
            sh[threadIdx.x] = a;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = sh[0];
}

template <typename T>
__global__ void k_minmax_final(const T* vol, const Extreme<T>* parts, uint32_t nparts, double* out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Extreme<T> e = parts[0];
    for (uint32_t i = 1; i < nparts; ++i) merge_extreme(e, parts[i]);
    const T first = vol[0];
    if (is_nan_val<T>(first) || !e.valid) { e.mn = first; e.mx = first; }
    out[0] = (double)e.mn;
    out[1] = (double)e.mx;
}

/* ------------------------------------------------------------------------- */
/* host-side launch wrappers                                                   */
/* ------------------------------------------------------------------------- */

template <typename T>
hipError_t launch_march_t(const JobArgs& a, bool count, uint32_t lds_bytes, hipStream_t st)
{
    const dim3 grid(a.nb * a.nslabs), block(kBlock);
    if (count) hipLaunchKernelGGL((k_march<T, true>), grid, block, lds_bytes, st, a);
    else hipLaunchKernelGGL((k_march<T, false>), grid, block, lds_bytes, st, a);
    return hipGetLastError();
}

hipError_t launch_march(int dtype, const JobArgs& a, bool count, uint32_t lds_bytes, hipStream_t st)
{
    switch (dtype) {
    case 0: return launch_march_t<int8_t>(a, count, lds_bytes, st);
    case 1: return launch_march_t<int16_t>(a, count, lds_bytes, st);
    case 2: return launch_march_t<int32_t>(a, count, lds_bytes, st);
    case 3: return launch_march_t<int64_t>(a, count, lds_bytes, st);
    case 4: return launch_march_t<uint8_t>(a, count, lds_bytes, st);
    case 5: return launch_march_t<uint16_t>(a, count, lds_bytes, st);
    case 6: return launch_march_t<uint32_t>(a, count, lds_bytes, st);
    case 7: return launch_march_t<uint64_t>(a, count, lds_bytes, st);
    case 8: return launch_march_t<float>(a, count, lds_bytes, st);
    default: return launch_march_t<double>(a, count, lds_bytes, st);
    }
}

hipError_t launch_scan(const uint32_t* counts, uint64_t n, uint32_t narr, uint64_t* partials,
                       uint64_t* prefix, uint64_t* totals, hipStream_t st)
{
    const uint32_t nseg = (uint32_t)((n + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(k_scan_partials, dim3(nseg * narr), dim3(kBlock), 0, st, counts, n, nseg, partials);
    hipLaunchKernelGGL(k_scan_apply, dim3(nseg * narr), dim3(kBlock), 0, st, counts, n, nseg, partials, prefix, totals);
    return hipGetLastError();
}

uint32_t scan_partials_needed(uint64_t n, uint32_t narr)
{
    return (uint32_t)((n + kScanTile - 1) / kScanTile) * narr;
}

template <typename T>
hipError_t launch_minmax_t(const void* vol, uint64_t n, void* scratch, double* out, hipStream_t st)
{
    const uint32_t nparts = 1024;
    Extreme<T>* parts = (Extreme<T>*)scratch;
    hipLaunchKernelGGL((k_minmax_partial<T>), dim3(nparts), dim3(kBlock), 0, st, (const T*)vol, n, parts);
    hipLaunchKernelGGL((k_minmax_final<T>), dim3(1), dim3(64), 0, st, (const T*)vol, parts, nparts, out);
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

size_t minmax_scratch_bytes() { return 1024 * sizeof(Extreme<double>) + 64; }

}  // namespace kmc
