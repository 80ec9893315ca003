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
 *   - a *tile* (z, b) is band b of node slice z.  It owns the vertices of the
 *     crossing edges whose low node lies in the band at slice z, and the
 *     triangles of the cells of band b at cell slice z-1.  Tiles in (z, b)
 *     order enumerate vertices in node scan order x {x,y,z}-edge and triangles
 *     in cell scan order, which is exactly the order of calculate_isopoints
 *     and connect_isopoints / extract_surfaces_with_duplication in
 *     Source/Core/Visualization/Mapper/MarchingCubes.cpp:155-243,365-521;
 *   - a workgroup marches z through a *slab* of consecutive slices of one band.
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
constexpr int kBlock = 256;          // threads per workgroup (4 waves of 64)
constexpr int kWaves = kBlock / 64;

struct JobArgs {
    const void* values;
    uint32_t nx, ny, nz;
    uint32_t R;        // node rows per band
    uint32_t nb;       // number of bands, ceil(ny / R)
    uint32_t NC;       // 64-node chunks per row, ceil(nx / 64)
    uint32_t nlev;     // isolevels in this pass
    uint32_t slab;     // slices marched by one workgroup
    uint32_t nslabs;
    uint32_t dup;      // 1: duplicated vertices (per-triangle output)
    uint32_t vnormal;  // 1: VertexNormal (only without duplication)
    uint32_t queue_cap;
    uint64_t ntiles;   // nz * nb
    double minv, maxv; // clamp range of the interpolation (volume min/max)
    float minc[3];     // min object coordinate
    float sc[3];       // (max - min) / ncells, per axis (float ops, host side)
    double iso[kMaxLevels];
    float thr[kMaxLevels];    // float32 data: v > thr  <=>  double(v) > iso
    // counts[(l*2 + k) * ntiles + tile], k = 0 vertices, 1 triangles
    uint32_t* counts;
    // exclusive prefixes, same layout
    const uint64_t* prefix;
    // per-level outputs
    float* coords[kMaxLevels];
    float* normals[kMaxLevels];
    uint32_t* conn[kMaxLevels];
    uint64_t cap_vertices[kMaxLevels];   // capacity in coordinate triples
    uint64_t cap_triangles[kMaxLevels];
};

/* LDS layout of one workgroup; sizes in bytes, offsets 16-byte aligned. */
struct LdsLayout {
    uint32_t nrb;       // rows held per slice: y0-1 .. y0+R+1
    uint32_t ncp;       // chunks per row incl. one zero pad chunk
    uint32_t off_bits;  // uint64 [3][nrb][ncp][nlev]
    uint32_t off_pv;    // uint32 [2][R+1][NC][nlev]    vertex chunk prefixes
    uint32_t off_tm;    // uint64 [R][NC][nlev][3]       triangle-count bit planes
    uint32_t off_pt;    // uint32 [R][NC][nlev]          triangle chunk prefixes
    uint32_t off_scan;  // uint32 [kBlock + 64]
    uint32_t off_tot;   // uint32 [2][kMaxLevels]
    uint32_t off_queue; // uint32 [queue_cap]
    uint32_t bytes;
};

#if defined(__HIPCC__)
__host__ __device__
#endif
inline uint32_t align16(uint32_t x) { return (x + 15u) & ~15u; }

#if defined(__HIPCC__)
__host__ __device__
#endif
inline LdsLayout lds_layout(uint32_t R, uint32_t NC, uint32_t nlev, uint32_t queue_cap)
{
    LdsLayout L;
    L.nrb = R + 3;
    L.ncp = NC + 1;
    uint32_t o = 0;
    L.off_bits = o;  o = align16(o + 3u * L.nrb * L.ncp * nlev * 8u);
    L.off_pv = o;    o = align16(o + 2u * (R + 1) * NC * nlev * 4u);
    L.off_tm = o;    o = align16(o + R * NC * nlev * 3u * 8u);
    L.off_pt = o;    o = align16(o + R * NC * nlev * 4u);
    L.off_scan = o;  o = align16(o + (kBlock + 64) * 4u);
    L.off_tot = o;   o = align16(o + 2u * kMaxLevels * 4u);
    L.off_queue = o; o = align16(o + queue_cap * 4u);
    L.bytes = o;
    return L;
}

}  // namespace kmc
