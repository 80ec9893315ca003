/*
 * kmc_job.h -- the argument block of one extraction job and the layout of
 * the per-workgroup LDS state, shared by host launch code and kernels.
 *
 * A job is one structured volume with up to kMaxLevels isolevels extracted in
 * the same pass (each level yields its own surface, identical to what a
 * separate kvs::MarchingCubes call at that level would return).
 *
 * Geometry of the work decomposition (all in reference scan order, x fastest):
 *   - a *band* is R consecutive node rows [y0, y0+R) spanning the full x range;
 *   - a *tile* (z, b) is band b at step z.  It emits the triangles of the
 *     cells of band b at cell slice z-1 and the vertices of the crossing edges
 *     whose low node lies in the band at node slice z (or z-1 with
 *     VertexNormal, whose normals need the face normals of the cells on both
 *     sides of a slice).  Tiles in (z, b) order enumerate vertices in node
 *     scan order x {x,y,z}-edge and triangles in cell scan order, which is the
 *     order of calculate_isopoints and connect_isopoints /
 *     extract_surfaces_with_duplication in
 *     Source/Core/Visualization/Mapper/MarchingCubes.cpp:155-243,365-521;
 *   - a *unit* is one level of one band over a slab of `slab` consecutive
 *     steps.  The emit pass runs one unit at a time per wave (a *team*): the
 *     wave marches z through its unit with its own LDS region, keeping a ring
 *     of node-slice classification masks (one new slice loaded per step) and,
 *     with VertexNormal, a ring of the face normals of the last two cell
 *     slices.  Teams never wait for each other.
 *
 * Passes of one job: classify (stream the values once, write the bit volume),
 * count (per-row vertex / triangle counts and in-row chunk prefixes from the
 * bit volume), scan (row prefixes), emit (outputs at the prefixes; raw values
 * are read again only at the endpoints of crossing edges).  The emit is flat
 * (compaction records, then one thread per output item) unless y or z exceeds
 * 2^20, when the tile teams below do it.
 *
 * Within a slice, the classification bits of 64 consecutive nodes of a row
 * are held as one 64-bit mask per level (a *chunk*), built with one wave
 * ballot; edge crossings and cell activity are then bitwise operations on
 * chunks, and ranks are popcounts.
 */
#pragma once
#include <stdint.h>

namespace kmc {

constexpr int kMaxLevels = 8;
constexpr int kBlock = 256;          // threads per workgroup of the dense passes (4 waves of 64)
constexpr int kWaves = kBlock / 64;
constexpr int kEmitTeams = 4;        // emit workgroup: 4 independent one-wave teams
constexpr int kEmitBlock = 64 * kEmitTeams;
constexpr uint32_t kTableBytes = 256 * 8;   // case table at the start of the emit workgroup's LDS

/* Crossing edges of one 64-node chunk with a vertex (flat emit): the
 * level-relative id of its first vertex and the x / y / z crossing-edge masks
 * of its nodes. */
struct alignas(32) ChunkCross {
    uint64_t base;
    uint64_t x, y, z;
};

struct JobArgs {
    const void* values;
    uint32_t nx, ny, nz;
    uint32_t R;        // node rows per band (emit tiles)
    uint32_t nb;       // number of bands, ceil(ny / R)
    uint32_t NC;       // 64-node chunks per row, ceil(nx / 64)
    uint32_t cshift;   // lanes are split into 64 >> cshift rows of (1 << cshift) chunk lanes
                       // (1 << cshift = the power of two >= min(NC, 64))
    uint32_t nlev;     // isolevels in this pass
    uint32_t slab;     // steps marched by one unit
    uint32_t nslabs;   // units per band, ceil(nz / slab)
    uint32_t dup;      // 1: duplicated vertices (per-triangle output)
    uint32_t vnormal;  // 1: VertexNormal (only without duplication)
    uint32_t queue_cap;
    uint32_t fn_cap;   // face normals held per cell slice (VertexNormal)
    uint32_t team_stride;  // bytes of LDS per team (>= the team layout's size)
    uint32_t teams;        // teams per emit workgroup (<= kEmitTeams; fewer when a team's LDS is large)
    uint32_t rec_levels;   // flat emit: bit l set when level l has records (levels without outputs get none)
    uint64_t nunits;   // nlev * nb * nslabs
    uint64_t nrows;    // nz * ny: counts and prefixes are kept per node row
    double minv, maxv; // clamp range of the interpolation (volume min/max)
    float minc[3];     // min object coordinate
    float sc[3];       // (max - min) / ncells, per axis (float ops, host side)
    double iso[kMaxLevels];
    float thr[kMaxLevels];    // float32 data: v > thr  <=>  double(v) > iso
    // counts[(l*2 + k) * nrows + z*ny + y], k = 0: vertices owned by node row
    // (y, z); k = 1: triangles of cell row (y, z-1)
    uint32_t* counts;
    // exclusive prefixes, same layout
    const uint64_t* prefix;
    // classification bit volume: bitvol[((l * nz + z) * ny + y) * NC + c] is the
    // 64-bit mask of nodes x in [64c, 64c+64) of row y, slice z, level l
    uint64_t* bitvol;
    // optional, same layout as bitvol: per chunk the exclusive prefix within its
    // row of the vertex counts (vpc) and of the triangle counts of cell row
    // (y, z-1) (tpc); written by the count pass when non-null
    uint32_t* vpc;
    uint32_t* tpc;
    // records of the compaction pass (flat emit): per vertex, in output order,
    // x | y << 20 | z << 40 | axis << 60; per triangle x | y << 20 | z << 40 |
    // t << 60 of its cell and the cell's cube index; per triangle its face normal
    // (VertexNormal); level l's items start at voff[l] / toff[l]
    uint64_t* vkey;
    uint64_t* trec;
    uint8_t* tci;
    float* fn;
    // per chunk (bit volume layout), written by the compaction for chunks with a vertex
    ChunkCross* cross;
    // VertexNormal: per vertex, for each of the four cells around its edge (in
    // cell scan order) the item index of that cell's first triangle record;
    // written by the triangles that use the edge (cells outside the volume:
    // not written, not read)
    uint32_t* vslot;
    uint64_t voff[kMaxLevels + 1];
    uint64_t toff[kMaxLevels + 1];
    // per-level outputs
    float* coords[kMaxLevels];
    float* normals[kMaxLevels];
    uint32_t* conn[kMaxLevels];
    uint64_t cap_vertices[kMaxLevels];   // capacity in coordinate triples
    uint64_t cap_triangles[kMaxLevels];
    // SlicePlane: nodes classify by the plane's sign instead of their values;
    // each vertex gets the colour of the value interpolated along its edge
    float plane[4];            // coefficients: plane(x,y,z) = plane[0]*x + plane[1]*y + plane[2]*z + plane[3]
    const uint8_t* ctable;     // colour map table, 3 bytes per entry (device)
    uint32_t cres;             // entries of ctable
    float crange[2];           // colour map range
    uint8_t* colors;           // one RGB per vertex (9 bytes per triangle)
};

/* LDS layout of one team (one level); sizes in bytes, offsets 16-byte
 * aligned, relative to the team's region.  Node slices live in a ring of
 * 3 mask slots (slices z-1 .. z+1), each holding node rows y0-1 .. y0+R+1.  Window vertex
 * prefixes cover node slices z-1 and z and node rows y0-1 .. y0+R; cell data
 * cover cell slice z-1, cell rows y0-1 .. y0+R-1 (row 0 is the halo row of
 * the band above, used for face normals only), in a ring of 2 cell slices
 * with VertexNormal. */
struct LdsLayout {
    uint32_t nrb;       // mask rows held per slice: y0-1 .. y0+R+1
    uint32_t nslots;    // node-slice mask slots
    uint32_t ncs;       // cell-slice slots
    uint32_t off_bits;  // uint64 [nslots][nrb][NC]         classification masks
    uint32_t off_wpv;   // uint32 [2][R+2][NC]              window vertex chunk prefixes
    uint32_t off_act;   // uint64 [ncs][R+1][NC]            active cells
    uint32_t off_pt;    // uint32 [ncs][R+1][NC]            triangle chunk prefixes
    uint32_t off_ap;    // uint32 [ncs][R+1][NC]            active-cell chunk prefixes (VertexNormal)
    uint32_t off_crec;  // uint32 [ncs][fn_cap]             per active cell: first triangle << 8 | cube index
    uint32_t off_fn;    // float  [ncs][fn_cap][3]          face normals (VertexNormal)
    uint32_t off_vq;    // uint32 [queue_cap]               vertex work queue
    uint32_t off_tq;    // uint32 [queue_cap]               triangle work queue
    uint32_t bytes;
};

#if defined(__HIPCC__)
__host__ __device__
#endif
inline uint32_t align16(uint32_t x) { return (x + 15u) & ~15u; }

#if defined(__HIPCC__)
__host__ __device__
#endif
inline LdsLayout lds_layout(const JobArgs& a)
{
    LdsLayout L;
    const uint32_t R = a.R, NC = a.NC;
    L.nrb = R + 3;
    L.nslots = 3u;
    L.ncs = a.vnormal ? 2u : 1u;
    uint32_t o = 0;
    L.off_bits = o;  o = align16(o + L.nslots * L.nrb * NC * 8u);
    L.off_wpv = o;   o = align16(o + 2u * (R + 2) * NC * 4u);
    L.off_act = o;   o = align16(o + L.ncs * (R + 1) * NC * 8u);
    L.off_pt = o;    o = align16(o + L.ncs * (R + 1) * NC * 4u);
    L.off_ap = o;    o = align16(o + (a.vnormal ? L.ncs * (R + 1) * NC * 4u : 0u));
    L.off_crec = o;  o = align16(o + (a.vnormal ? L.ncs * a.fn_cap * 4u : 0u));
    L.off_fn = o;    o = align16(o + (a.vnormal ? L.ncs * a.fn_cap * 12u : 0u));
    L.off_vq = o;    o = align16(o + a.queue_cap * 4u);
    L.off_tq = o;    o = align16(o + a.queue_cap * 4u);
    L.bytes = o;
    return L;
}

/* LDS of an emit workgroup of `teams` teams, team_stride bytes apart. */
#if defined(__HIPCC__)
__host__ __device__
#endif
inline uint32_t emit_lds_bytes(uint32_t team_stride, uint32_t teams) { return kTableBytes + teams * team_stride; }

}  // namespace kmc
