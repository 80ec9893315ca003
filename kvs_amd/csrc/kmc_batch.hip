/*
 * kmc_batch.hip -- the batched form: many volumes of uneven sizes, each with
 * its own isolevels, extracted by one captured hipGraph.
 *
 * A batch is cut into passes (one volume, up to kMaxLevels isolevels each),
 * grouped by value type and emit mode.  Per group the graph holds one
 * classify, one count and one emit launch over the rows / units of all its
 * passes (a workgroup finds its pass by binary search of the pass start
 * table); the per-row counts of
 * all passes sit in one array and take one scan.  A layout kernel then turns
 * each surface's totals into offsets in a shared output arena, so that sizes
 * never go through the host inside the graph; when a run's surfaces do not fit
 * the arena, the emit writes nothing, the host grows the arena, captures the
 * graph again and reruns.  Each surface equals the plain call's output.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "kmc.h"
#include "kmc_guard.h"
#include "kmc_internal.h"
#include "kmc_job.h"

namespace kmc {

/* Where one surface of a run landed (element offsets into the arenas); the
 * entry after the last surface holds the totals, the overflow flag (nv) and
 * the record bytes the flat emit needs (rb). */
struct SegOut {
    uint64_t nv, nt;          // vertices (node-shared) and triangles counted
    uint64_t oc, on, ok;      // offsets of coords (floats), normals (floats), connections (uint32)
    uint64_t rb;
};

/* Flat-emit records of a pass: vertex keys, triangle records, cube indices,
 * face normals and vertex slots (VertexNormal), each 256-byte aligned. */
__host__ __device__ inline uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }
__host__ __device__ inline uint64_t record_bytes(uint64_t nv, uint64_t nt, bool vn)
{
    return align256(nv * 8) + align256(nt * 8) + align256(nt) + (vn ? align256(nt * 12) + align256(nv * 16) : 0);
}

/* Point pass a's records at base (or at nothing when base is null) and set its
 * per-level item offsets from the level totals nv[l] / nt[l]; a level given
 * zero totals gets no records, and the compaction skips it (rec_levels). */
__host__ __device__ inline void place_records(JobArgs& a, const uint64_t* nv, const uint64_t* nt, uint8_t* base)
{
    a.voff[0] = a.toff[0] = 0;
    a.rec_levels = 0;
    for (uint32_t l = 0; l < a.nlev; ++l) {
        a.voff[l + 1] = a.voff[l] + (base && !a.dup ? nv[l] : 0);
        a.toff[l + 1] = a.toff[l] + (base ? nt[l] : 0);
        if (base && ((!a.dup && nv[l]) || nt[l])) a.rec_levels |= 1u << l;
    }
    if (!base) {
        a.vkey = nullptr; a.trec = nullptr; a.tci = nullptr; a.fn = nullptr; a.vslot = nullptr;
        return;
    }
    const uint64_t V = a.voff[a.nlev], T = a.toff[a.nlev];
    uint8_t* p = base;
    a.vkey = (uint64_t*)p; p += align256(V * 8);
    a.trec = (uint64_t*)p; p += align256(T * 8);
    a.tci = p; p += align256(T);
    a.fn = a.vnormal ? (float*)p : nullptr;
    if (a.vnormal) p += align256(T * 12);
    a.vslot = a.vnormal ? (uint32_t*)p : nullptr;
}

constexpr uint32_t kLayoutBlock = 256;
constexpr uint32_t kLayoutPer = 8;

__device__ __forceinline__ uint64_t level_total(const JobArgs& a, uint32_t l, uint32_t k)
{
    const uint64_t i0 = (uint64_t)(2 * l + k) * a.nrows, last = i0 + a.nrows - 1;
    return a.prefix[last] + a.counts[last] - a.prefix[i0];
}

/* Block-wide exclusive scan of three 64-bit values; returns the totals. */
__device__ void block_scan3(uint64_t v[3], uint64_t tot[3])
{
    __shared__ uint64_t sh[3][kLayoutBlock / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (int q = 0; q < 3; ++q) {
        uint64_t incl = v[q];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t t = (uint64_t)__shfl_up((unsigned long long)incl, d);
            if ((int)lane >= d) incl += t;
        }
        if (lane == 63) sh[q][wave] = incl;
        v[q] = incl - v[q];
    }
    __syncthreads();
    for (int q = 0; q < 3; ++q) {
        uint64_t run = 0, all = 0;
        for (uint32_t w = 0; w < kLayoutBlock / 64; ++w) {
            if (w < wave) run += sh[q][w];
            all += sh[q][w];
        }
        v[q] += run;
        tot[q] = all;
    }
}

/* Sizes and arena offsets of every surface of the batch; points each pass's
 * level outputs into the arena, or at nothing (capacity 0) when the batch
 * does not fit.  out[nseg] receives the totals needed and the overflow flag. */
__global__ __launch_bounds__(kLayoutBlock) void k_batch_layout(JobArgs* jobs, const uint32_t* seg_map, uint32_t nseg,
                                                               float* arena_c, float* arena_n, uint32_t* arena_k,
                                                               uint64_t cap_c, uint64_t cap_n, uint64_t cap_k,
                                                               SegOut* out, uint32_t npass, uint8_t* arena_r,
                                                               uint64_t cap_r, uint64_t* istart)
{
    __shared__ uint32_t overflow;
    /* one sweep per kLayoutBlock * kLayoutPer surfaces, carrying the offsets */
    uint64_t carry[3] = { 0, 0, 0 };
    for (uint32_t b0 = 0; b0 < nseg; b0 += kLayoutBlock * kLayoutPer) {
        const uint32_t s0 = b0 + threadIdx.x * kLayoutPer;
        uint64_t sz[kLayoutPer][3];
        uint64_t nv[kLayoutPer], nt[kLayoutPer];
        uint64_t v[3] = { 0, 0, 0 };
        for (uint32_t i = 0; i < kLayoutPer; ++i) {
            nv[i] = nt[i] = 0;
            sz[i][0] = sz[i][1] = sz[i][2] = 0;
            if (s0 + i < nseg) {
                const uint32_t m = seg_map[s0 + i];
                const JobArgs& a = jobs[m >> 3];
                const uint32_t l = m & 7u;
                nv[i] = level_total(a, l, 0);
                nt[i] = level_total(a, l, 1);
                const uint64_t verts = a.dup ? 3 * nt[i] : nv[i];
                if (verts > 0) {
                    sz[i][0] = 3 * verts;
                    sz[i][1] = 3 * (a.vnormal ? nv[i] : nt[i]);
                    sz[i][2] = a.dup ? 0 : 3 * nt[i];
                }
            }
            v[0] += sz[i][0]; v[1] += sz[i][1]; v[2] += sz[i][2];
        }
        uint64_t tot[3];
        block_scan3(v, tot);
        for (uint32_t i = 0; i < kLayoutPer; ++i) {
            if (s0 + i < nseg) {
                SegOut o;
                o.nv = nv[i];
                o.nt = nt[i];
                o.oc = carry[0] + v[0];
                o.on = carry[1] + v[1];
                o.ok = carry[2] + v[2];
                out[s0 + i] = o;
            }
            v[0] += sz[i][0]; v[1] += sz[i][1]; v[2] += sz[i][2];
        }
        carry[0] += tot[0]; carry[1] += tot[1]; carry[2] += tot[2];
        __syncthreads();
    }
    /* flat-emit records of every pass (istart: null for the tile teams) */
    uint64_t rcarry[3] = { 0, 0, 0 };
    if (istart) {
        for (uint32_t b0 = 0; b0 < npass; b0 += kLayoutBlock * kLayoutPer) {
            const uint32_t p0 = b0 + threadIdx.x * kLayoutPer;
            uint64_t sz[kLayoutPer][3];
            uint64_t v[3] = { 0, 0, 0 };
            for (uint32_t i = 0; i < kLayoutPer; ++i) {
                sz[i][0] = sz[i][1] = sz[i][2] = 0;
                if (p0 + i < npass) {
                    const JobArgs& a = jobs[p0 + i];
                    for (uint32_t l = 0; l < a.nlev; ++l) {
                        sz[i][1] += a.dup ? 0 : level_total(a, l, 0);
                        sz[i][2] += level_total(a, l, 1);
                    }
                    sz[i][0] = record_bytes(sz[i][1], sz[i][2], a.vnormal != 0);
                }
                v[0] += sz[i][0]; v[1] += sz[i][1]; v[2] += sz[i][2];
            }
            uint64_t tot[3];
            block_scan3(v, tot);
            for (uint32_t i = 0; i < kLayoutPer; ++i) {
                if (p0 + i < npass) {
                    istart[p0 + i] = rcarry[1] + v[1];                 // vertex items
                    istart[npass + 1 + p0 + i] = rcarry[2] + v[2];     // triangle items
                    jobs[p0 + i].voff[0] = rcarry[0] + v[0];           // record byte offset (placed below)
                }
                v[0] += sz[i][0]; v[1] += sz[i][1]; v[2] += sz[i][2];
            }
            rcarry[0] += tot[0]; rcarry[1] += tot[1]; rcarry[2] += tot[2];
            __syncthreads();
        }
    }
    if (threadIdx.x == 0)
        overflow = (carry[0] > cap_c || carry[1] > cap_n || carry[2] > cap_k || (istart && rcarry[0] > cap_r)) ? 1u : 0u;
    __syncthreads();
    const bool fits = overflow == 0;
    if (istart) {
        for (uint32_t p = threadIdx.x; p < npass; p += kLayoutBlock) {
            JobArgs& a = jobs[p];
            uint64_t nv[kMaxLevels], nt[kMaxLevels];
            for (uint32_t l = 0; l < a.nlev; ++l) {
                nv[l] = level_total(a, l, 0);
                nt[l] = level_total(a, l, 1);
            }
            place_records(a, nv, nt, fits ? arena_r + a.voff[0] : nullptr);
            if (!fits) istart[p] = istart[npass + 1 + p] = 0;
        }
        if (threadIdx.x == 0) {
            istart[npass] = fits ? rcarry[1] : 0;
            istart[2 * npass + 1] = fits ? rcarry[2] : 0;
        }
    }
    for (uint32_t s = threadIdx.x; s < nseg; s += kLayoutBlock) {
        const uint32_t m = seg_map[s];
        JobArgs& a = jobs[m >> 3];
        const uint32_t l = m & 7u;
        const SegOut o = out[s];
        const uint64_t verts = a.dup ? 3 * o.nt : o.nv;
        a.coords[l] = arena_c + o.oc;
        a.normals[l] = arena_n + o.on;
        a.conn[l] = arena_k + o.ok;
        a.cap_vertices[l] = (fits && verts > 0) ? verts : 0;
        a.cap_triangles[l] = (fits && verts > 0) ? o.nt : 0;
    }
    if (threadIdx.x == 0) {
        SegOut t;
        t.nv = overflow;
        t.nt = 0;
        t.oc = carry[0];
        t.on = carry[1];
        t.ok = carry[2];
        t.rb = rcarry[0];
        out[nseg] = t;
    }
}

}  // namespace kmc

using namespace kmc;

/* ------------------------------------------------------------------------- */
/* host side                                                                   */
/* ------------------------------------------------------------------------- */

namespace {

struct UserJob {
    kmc_volume volume;           // values point at the device copy
    std::vector<double> levels;
    int normal_type;
    int dup;
    double mn, mx;
    float minc[3], maxc[3];
    bool empty;                  // Math::Equal(min, max)
    uint32_t first_surface;
};

struct Pass {
    uint32_t user;
    uint32_t level0;
    uint32_t nlev;
    int dtype;
    JobArgs a;
    uint32_t lds;
    uint64_t counts_off;         // first element of its counts in the shared arrays
    uint64_t bitvol_off;
    uint32_t seg0;               // first surface segment
};

struct Group {
    int dtype;
    int mode;                    // 0 duplicated, 1 PolygonNormal, 2 VertexNormal
    uint32_t p0, p1;             // passes [p0, p1)
    uint64_t rows, units;
    uint32_t max_levels;         // the most levels of one of its passes
    uint32_t emit_lds;
    uint32_t teams;                // emit teams per workgroup
    uint64_t* row_start_d;       // p1 - p0 + 1 entries
    uint64_t* unit_start_d;
    std::vector<uint64_t> row_start_h, unit_start_h;   // upload sources, kept while the copies may run
};

}  // namespace

struct kmc_batch {
    int device;
    uint32_t flags;
    std::vector<UserJob> users;
    std::vector<void*> owned_values;
    uint32_t n_surfaces;
    // plan
    uint32_t r_max, queue_cap;
    std::vector<Pass> passes;
    std::vector<Group> groups;
    std::vector<uint32_t> seg_pass_level;   // (pass << 3) | level, per segment
    std::vector<uint32_t> surface_seg;      // surface -> segment (or ~0u when the job is empty)
    std::vector<uint32_t> surface_user;     // surface -> job
    JobArgs* jobs_d;
    std::vector<JobArgs> jobs_h;            // host staging of the job table (two-phase use)
    uint32_t* seg_map_d;
    uint32_t* counts_d;
    uint64_t* prefix_d;
    uint64_t* scan_aux_d;
    uint64_t n_counts;
    uint32_t scan_parts;
    uint64_t* bitvol_d;
    uint64_t bitvol_words;
    bool flat;                              // flat emit (else the tile teams)
    bool flat_ok;                           // every pass's rows fit the flat records' 20-bit y / z
    uint32_t* cpre_d;                       // in-row chunk prefixes of the count pass (flat)
    ChunkCross* cross_d;                    // chunk crossing records of the compaction (flat)
    uint64_t* istart_d;                     // per-pass first vertex / triangle item (flat)
    std::vector<uint64_t> istart_h;
    uint8_t* arena_r;                       // flat-emit records
    uint64_t cap_r;
    SegOut* segout_d;
    SegOut* segout_h;                       // pinned
    float* arena_c;
    float* arena_n;
    uint32_t* arena_k;
    uint64_t cap_c, cap_n, cap_k;
    hipStream_t cap_stream;
    hipStream_t upload_stream;              // where table uploads are queued (the creating call's stream)
    hipGraph_t graph;
    hipGraphExec_t exec;
    hipEvent_t ev0, ev1;
    uint32_t graph_nodes;
    uint32_t recaptures;
    double last_ms;
    bool have_results;
};

namespace {

void free_plan_buffers(kmc_batch* b)
{
    for (Group& g : b->groups) {
        device_free(g.row_start_d);
        device_free(g.unit_start_d);
        g.row_start_d = g.unit_start_d = nullptr;
    }
}

void destroy_graph(kmc_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void destroy(kmc_batch* b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    /* blocks go back to the pool: nothing queued may still use them */
    (void)hipStreamSynchronize(b->upload_stream);
    if (b->cap_stream) (void)hipStreamSynchronize(b->cap_stream);
    destroy_graph(b);
    free_plan_buffers(b);
    device_free(b->jobs_d);
    device_free(b->seg_map_d);
    device_free(b->counts_d);
    device_free(b->prefix_d);
    device_free(b->scan_aux_d);
    device_free(b->bitvol_d);
    device_free(b->cpre_d);
    device_free(b->cross_d);
    device_free(b->istart_d);
    device_free(b->arena_r);
    device_free(b->segout_d);
    device_free(b->arena_c);
    device_free(b->arena_n);
    device_free(b->arena_k);
    host_pinned_free(b->segout_h);
    for (void* p : b->owned_values) device_free(p);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    delete b;
}

/* Tile geometry of every pass for (r_max, queue_cap); tile tables per group. */
kmc_status plan_tiles(kmc_batch* b, uint32_t r_max, uint32_t queue_cap)
{
    free_plan_buffers(b);
    b->r_max = r_max;
    b->queue_cap = queue_cap;
    /* the flat emit keeps no per-team LDS state: a pass whose team does not
     * fit only rules out the tile teams */
    for (Pass& p : b->passes) {
        if (!plan_geometry(p.a, &p.lds, r_max, queue_cap) && !b->flat) return KMC_ERR_INVALID_ARGUMENT;
        plan_rows(p.a);
    }
    for (Group& g : b->groups) {
        std::vector<uint64_t> rs, ts;
        uint64_t r = 0, t = 0;
        uint32_t stride = 0, teams = kEmitTeams;
        g.max_levels = 0;
        for (uint32_t i = g.p0; i < g.p1; ++i) {
            rs.push_back(r);
            ts.push_back(t);
            r += b->passes[i].a.nrows;
            t += b->passes[i].a.nunits;
            stride = std::max(stride, b->passes[i].a.team_stride);
            g.max_levels = std::max(g.max_levels, b->passes[i].a.nlev);
        }
        /* one emit launch per group: every team region of a workgroup is as
         * large as the group's largest, and as many teams as 64 KiB hold */
        teams = std::min<uint32_t>(kEmitTeams, (65536u - kTableBytes) / std::max(1u, stride));
        for (uint32_t i = g.p0; i < g.p1; ++i) {
            b->passes[i].a.team_stride = stride;
            b->passes[i].a.teams = teams;
        }
        g.teams = teams;
        g.emit_lds = emit_lds_bytes(stride, teams);
        rs.push_back(r);
        ts.push_back(t);
        g.rows = r;
        g.units = t;
        if (t > 0xFFFFFFFFull) return KMC_ERR_INVALID_ARGUMENT;
        g.row_start_d = (uint64_t*)device_alloc(rs.size() * 8);
        g.unit_start_d = (uint64_t*)device_alloc(ts.size() * 8);
        if (!g.row_start_d || !g.unit_start_d) return KMC_ERR_OUT_OF_MEMORY;
        g.row_start_h = std::move(rs);
        g.unit_start_h = std::move(ts);
        if (hipMemcpyAsync(g.row_start_d, g.row_start_h.data(), g.row_start_h.size() * 8, hipMemcpyHostToDevice,
                           b->upload_stream) != hipSuccess ||
            hipMemcpyAsync(g.unit_start_d, g.unit_start_h.data(), g.unit_start_h.size() * 8, hipMemcpyHostToDevice,
                           b->upload_stream) != hipSuccess)
            return KMC_ERR_DEVICE;
    }
    for (Pass& p : b->passes) {
        p.a.vpc = b->flat ? b->cpre_d + p.bitvol_off : nullptr;
        p.a.tpc = b->flat ? b->cpre_d + b->bitvol_words + p.bitvol_off : nullptr;
        p.a.cross = b->flat ? b->cross_d + p.bitvol_off : nullptr;
    }
    b->jobs_h.resize(b->passes.size());
    for (size_t i = 0; i < b->passes.size(); ++i) b->jobs_h[i] = b->passes[i].a;
    if (!b->jobs_h.empty() && hipMemcpyAsync(b->jobs_d, b->jobs_h.data(), b->jobs_h.size() * sizeof(JobArgs),
                                             hipMemcpyHostToDevice, b->upload_stream) != hipSuccess)
        return KMC_ERR_DEVICE;
    return KMC_OK;
}

/* Launches up to the output layout: classify, count, scan, layout (and, with
 * copy_sizes, the surface sizes to segout_h). */
hipError_t record_front(kmc_batch* b, hipStream_t st, bool copy_sizes)
{
    const uint32_t nseg = (uint32_t)b->seg_pass_level.size();
    if (nseg == 0) return hipSuccess;
    hipError_t e = hipSuccess;
    if (copy_sizes) {
        /* launched from the host (the serving form's two phases): each pass
         * as the plain call's classify and count, its arguments in the
         * launch -- for the large volumes served, faster than the batch
         * kernels' shared row ranges */
        for (const Group& g : b->groups)
            for (uint32_t i = g.p0; i < g.p1 && e == hipSuccess; ++i) {
                e = launch_classify(g.dtype, b->passes[i].a, st);
                if (e == hipSuccess) e = launch_count(b->passes[i].a, st);
            }
        if (e != hipSuccess) return e;
    } else {
        for (const Group& g : b->groups) {
            e = launch_classify_batch(g.dtype, b->jobs_d + g.p0, g.row_start_d, g.p1 - g.p0, g.rows, g.max_levels, st);
            if (e != hipSuccess) return e;
        }
        for (const Group& g : b->groups) {
            e = launch_count_batch(b->jobs_d + g.p0, g.row_start_d, g.p1 - g.p0, g.rows, st);
            if (e != hipSuccess) return e;
        }
    }
    uint64_t* partials = b->scan_aux_d;
    uint64_t* totals = b->scan_aux_d + b->scan_parts;
    e = launch_scan(b->counts_d, b->n_counts, 1, partials, b->prefix_d, totals, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_batch_layout, dim3(1), dim3(kLayoutBlock), 0, st, b->jobs_d, b->seg_map_d, nseg, b->arena_c,
                       b->arena_n, b->arena_k, b->cap_c, b->cap_n, b->cap_k, b->segout_d, (uint32_t)b->passes.size(),
                       b->arena_r, b->cap_r, b->flat ? b->istart_d : nullptr);
    e = hipGetLastError();
    if (e != hipSuccess || !copy_sizes) return e;
    return hipMemcpyAsync(b->segout_h, b->segout_d, (nseg + 1) * sizeof(SegOut), hipMemcpyDeviceToHost, st);
}

/* Launches of the emit.  host_sizes: the passes' record offsets are set on
 * the host (the serving form's two phases), so each pass runs as the plain
 * call's compaction and per-item kernels with its arguments in the launch; in
 * a captured graph they live on the device only. */
hipError_t record_emit(kmc_batch* b, hipStream_t st, bool host_sizes)
{
    hipError_t e = hipSuccess;
    if (b->flat) {
        const uint32_t np = (uint32_t)b->passes.size();
        for (const Group& g : b->groups) {
            if (host_sizes) {
                for (uint32_t i = g.p0; i < g.p1 && e == hipSuccess; ++i) e = launch_flat_emit(g.dtype, b->passes[i].a, st);
            } else {
                e = launch_compact_batch(b->jobs_d + g.p0, g.row_start_d, g.p1 - g.p0, g.rows, st);
                if (e == hipSuccess)
                    e = launch_flat_batch(g.dtype, g.mode, b->jobs_d + g.p0, b->istart_d + g.p0,
                                          b->istart_d + np + 1 + g.p0, g.p1 - g.p0, st);
            }
            if (e != hipSuccess) return e;
        }
        return e;
    }
    for (const Group& g : b->groups) {
        e = launch_emit_batch(g.dtype, g.mode, b->jobs_d + g.p0, g.unit_start_d, g.p1 - g.p0, g.units, g.emit_lds,
                              g.teams, st);
        if (e != hipSuccess) return e;
    }
    return e;
}

/* Record the batch's launches on st (the captured graph). */
hipError_t record(kmc_batch* b, hipStream_t st)
{
    const uint32_t nseg = (uint32_t)b->seg_pass_level.size();
    if (nseg == 0) return hipSuccess;
    hipError_t e = record_front(b, st, false);
    if (e == hipSuccess) e = record_emit(b, st, false);
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(b->segout_h, b->segout_d, (nseg + 1) * sizeof(SegOut), hipMemcpyDeviceToHost, st);
}

/* The capture stream and timing events, made on first use (graph form only). */
kmc_status graph_resources(kmc_batch* b)
{
    if (b->cap_stream) return KMC_OK;
    if (hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        (void)hipGetLastError();
        return KMC_ERR_DEVICE;
    }
    return KMC_OK;
}

kmc_status capture(kmc_batch* b)
{
    destroy_graph(b);
    if (graph_resources(b) != KMC_OK) return KMC_ERR_DEVICE;
    if (hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return KMC_ERR_DEVICE;
    const hipError_t rec = record(b, b->cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(b->cap_stream, &g);
    if (rec != hipSuccess || end != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return KMC_ERR_DEVICE;
    }
    b->graph = g;
    size_t n = 0;
    (void)hipGraphGetNodes(g, nullptr, &n);
    b->graph_nodes = (uint32_t)n;
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        return KMC_ERR_DEVICE;
    }
    ++b->recaptures;
    return KMC_OK;
}

kmc_status grow_arena(kmc_batch* b, uint64_t need_c, uint64_t need_n, uint64_t need_k, uint64_t need_r)
{
    auto grow = [](void*& p, uint64_t& cap, uint64_t need, size_t esz) -> bool {
        if (need <= cap) return true;
        device_free(p);
        const uint64_t n = need + need / 8 + 1024;   // headroom for runs over changed values
        p = device_alloc(n * esz);
        cap = p ? n : 0;
        return p != nullptr;
    };
    void* c = b->arena_c; void* n = b->arena_n; void* k = b->arena_k;
    void* r = b->arena_r;
    const bool ok = grow(c, b->cap_c, need_c, 4) && grow(n, b->cap_n, need_n, 4) && grow(k, b->cap_k, need_k, 4) &&
                    grow(r, b->cap_r, need_r, 1);
    b->arena_c = (float*)c; b->arena_n = (float*)n; b->arena_k = (uint32_t*)k; b->arena_r = (uint8_t*)r;
    return ok ? KMC_OK : KMC_ERR_OUT_OF_MEMORY;
}

/* Launch the graph on st and wait; grows the arena and reruns on overflow. */
kmc_status run(kmc_batch* b, hipStream_t st)
{
    const uint32_t nseg = (uint32_t)b->seg_pass_level.size();
    b->have_results = false;
    if (nseg == 0) { b->have_results = true; b->last_ms = 0; return KMC_OK; }
    for (int attempt = 0; attempt < 3; ++attempt) {
        if (!b->exec) {
            const kmc_status s = capture(b);
            if (s != KMC_OK) return s;
        }
        if (hipEventRecord(b->ev0, st) != hipSuccess || hipGraphLaunch(b->exec, st) != hipSuccess ||
            hipEventRecord(b->ev1, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
            (void)hipGetLastError();
            return KMC_ERR_DEVICE;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
        b->last_ms = ms;
        const SegOut& t = b->segout_h[nseg];
        if (t.nv == 0) { b->have_results = true; return KMC_OK; }
        const kmc_status s = grow_arena(b, t.oc, t.on, t.ok, t.rb);
        if (s != KMC_OK) return s;
        destroy_graph(b);
    }
    return KMC_ERR_DEVICE;
}

/* Time the candidate emit forms on this batch -- the flat emit and the tile
 * teams at several band heights and queue sizes; keep the fastest. */
kmc_status autotune(kmc_batch* b, hipStream_t st)
{
    struct Cand { uint32_t r, q; bool flat; };
    const Cand cands[] = { { 16, 128, true }, { 16, 128, false }, { 8, 128, false }, { 4, 128, false },
                           { 2, 128, false }, { 16, 256, false } };
    double best = 1e30;
    Cand best_c = { b->r_max, b->queue_cap, b->flat };
    for (const Cand& c : cands) {
        if (c.flat && !b->flat_ok) continue;
        b->flat = c.flat;
        kmc_status s = plan_tiles(b, c.r, c.q);
        if (s == KMC_OK) { destroy_graph(b); s = run(b, st); }
        if (s != KMC_OK) continue;          // a shape whose LDS does not fit is skipped
        double t = 1e30;
        for (int i = 0; i < 3 && s == KMC_OK; ++i) {
            s = run(b, st);
            t = std::min(t, b->last_ms);
        }
        if (s == KMC_OK && t < best) { best = t; best_c = c; }
    }
    b->flat = best_c.flat;
    kmc_status s = plan_tiles(b, best_c.r, best_c.q);
    if (s == KMC_OK) { destroy_graph(b); s = capture(b); }
    return s;
}

kmc_status create_impl(const kmc_job* jobs, uint32_t n_jobs, uint32_t flags, hipStream_t st, kmc_batch** out,
                       bool plan_only = false)
{
    if (!out || (n_jobs > 0 && !jobs)) return KMC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    for (uint32_t j = 0; j < n_jobs; ++j) {
        const kmc_status s = validate_volume(&jobs[j].volume);
        if (s != KMC_OK) return s;
        if (jobs[j].n_levels > 0 && !jobs[j].isolevels) return KMC_ERR_INVALID_ARGUMENT;
        if (jobs[j].normal_type != KMC_VERTEX_NORMAL && jobs[j].normal_type != KMC_POLYGON_NORMAL)
            return KMC_ERR_INVALID_ARGUMENT;
    }
    kmc_status s = ensure_device();
    if (s != KMC_OK) return s;

    kmc_batch* b = new kmc_batch();
    (void)hipGetDevice(&b->device);
    b->flags = flags;
    b->r_max = 16;
    b->queue_cap = 128;
    b->upload_stream = st;

    /* jobs: device values, min/max, object coordinates */
    b->users.resize(n_jobs);
    uint32_t nsurf = 0;
    for (uint32_t j = 0; j < n_jobs && s == KMC_OK; ++j) {
        UserJob& u = b->users[j];
        u.volume = jobs[j].volume;
        u.levels.assign(jobs[j].isolevels, jobs[j].isolevels + jobs[j].n_levels);
        u.normal_type = jobs[j].normal_type;
        u.dup = (u.normal_type == KMC_VERTEX_NORMAL) ? 0 : (jobs[j].duplication ? 1 : 0);   // MarchingCubes.cpp:69-73
        u.first_surface = nsurf;
        nsurf += jobs[j].n_levels;
        const void* dvals = nullptr;
        void* owned = nullptr;
        s = device_values(&u.volume, st, &dvals, &owned);
        if (s != KMC_OK) break;
        if (owned) b->owned_values.push_back(owned);
        u.volume.values = dvals;
        u.volume.memory = KMC_MEMORY_DEVICE;
        u.mn = u.volume.min_value;
        u.mx = u.volume.max_value;
        if (!u.volume.has_min_max_values) s = min_max_on_device(&u.volume, dvals, st, &u.mn, &u.mx);
        object_coords(&u.volume, u.minc, u.maxc);
        const double d = u.mn - u.mx;
        u.empty = (d < 0 ? -d : d) < 1.e-6;                    // MarchingCubes.cpp:108-110
    }
    if (s == KMC_OK && hipStreamSynchronize(st) != hipSuccess) s = KMC_ERR_DEVICE;
    if (s != KMC_OK) { destroy(b); return s; }
    b->n_surfaces = nsurf;
    b->surface_seg.assign(nsurf, ~0u);
    b->surface_user.resize(nsurf);
    for (uint32_t j = 0; j < n_jobs; ++j)
        for (uint32_t l = 0; l < b->users[j].levels.size(); ++l) b->surface_user[b->users[j].first_surface + l] = j;

    /* passes, grouped by value type and emit mode (one kernel instance each) */
    for (int gi = 0; gi < 3 * (KMC_REAL64 + 1); ++gi) {
        const int t = gi / 3, mode = gi % 3;
        Group g{};
        g.dtype = t;
        g.mode = mode;
        g.p0 = (uint32_t)b->passes.size();
        for (uint32_t j = 0; j < n_jobs; ++j) {
            const UserJob& u = b->users[j];
            const int um = u.dup ? 0 : (u.normal_type == KMC_VERTEX_NORMAL ? 2 : 1);
            if (u.volume.type != t || um != mode || u.empty) continue;
            const uint32_t per = levels_per_pass(&u.volume, !u.dup && u.normal_type == KMC_VERTEX_NORMAL);
            if (per == 0) { destroy(b); return KMC_ERR_INVALID_ARGUMENT; }
            for (uint32_t l0 = 0; l0 < u.levels.size(); l0 += per) {
                Pass p{};
                p.user = j;
                p.level0 = l0;
                p.nlev = std::min<uint32_t>((uint32_t)u.levels.size() - l0, per);
                p.dtype = t;
                init_job(p.a, &u.volume, u.volume.values, u.levels.data() + l0, p.nlev, u.normal_type, u.dup, u.mn,
                         u.mx, u.minc, u.maxc);
                b->passes.push_back(p);
            }
        }
        g.p1 = (uint32_t)b->passes.size();
        if (g.p1 > g.p0) b->groups.push_back(g);
    }

    /* shared counts / prefixes / masks, surface segments */
    uint64_t coff = 0, boff = 0;
    for (uint32_t i = 0; i < b->passes.size(); ++i) {
        Pass& p = b->passes[i];
        JobArgs& a = p.a;
        a.nrows = (uint64_t)a.nz * a.ny;
        a.NC = (a.nx + 63) / 64;
        p.counts_off = coff;
        p.bitvol_off = boff;
        coff += 2ull * p.nlev * a.nrows;
        boff += (uint64_t)p.nlev * a.nz * a.ny * a.NC;
        p.seg0 = (uint32_t)b->seg_pass_level.size();
        for (uint32_t l = 0; l < p.nlev; ++l) {
            b->surface_seg[b->users[p.user].first_surface + p.level0 + l] = (uint32_t)b->seg_pass_level.size();
            b->seg_pass_level.push_back((i << 3) | l);
        }
    }
    b->n_counts = coff;
    b->bitvol_words = boff;
    const uint32_t nseg = (uint32_t)b->seg_pass_level.size();
    if (!b->passes.empty()) {
        b->scan_parts = scan_partials_needed(coff, 1);
        b->counts_d = (uint32_t*)device_alloc(coff * 4);
        b->prefix_d = (uint64_t*)device_alloc(coff * 8);
        b->scan_aux_d = (uint64_t*)device_alloc((b->scan_parts + 1) * 8);
        b->bitvol_d = (uint64_t*)device_alloc(boff * 8);
        b->flat_ok = true;
        for (const Pass& p : b->passes) b->flat_ok = b->flat_ok && flat_emit_fits(p.a.ny, p.a.nz);
        b->flat = b->flat_ok;
        if (b->flat_ok) {
            b->cpre_d = (uint32_t*)device_alloc(2 * boff * 4);
            b->cross_d = (ChunkCross*)device_alloc(boff * sizeof(ChunkCross));
            b->istart_d = (uint64_t*)device_alloc((2 * b->passes.size() + 2) * 8);
            if (!b->cpre_d || !b->cross_d || !b->istart_d) {
                (void)hipGetLastError();
                destroy(b);
                return KMC_ERR_OUT_OF_MEMORY;
            }
        }
        b->jobs_d = (JobArgs*)device_alloc(b->passes.size() * sizeof(JobArgs));
        b->seg_map_d = (uint32_t*)device_alloc(nseg * 4);
        b->segout_d = (SegOut*)device_alloc((nseg + 1) * sizeof(SegOut));
        b->segout_h = (SegOut*)host_pinned_alloc((nseg + 1) * sizeof(SegOut));
        if (!b->counts_d || !b->prefix_d || !b->scan_aux_d || !b->bitvol_d || !b->jobs_d || !b->seg_map_d ||
            !b->segout_d || !b->segout_h) {
            (void)hipGetLastError();
            destroy(b);
            return KMC_ERR_OUT_OF_MEMORY;
        }
        for (Pass& p : b->passes) {
            p.a.counts = b->counts_d + p.counts_off;
            p.a.prefix = b->prefix_d + p.counts_off;
            p.a.bitvol = b->bitvol_d + p.bitvol_off;
        }
        if (hipMemcpyAsync(b->seg_map_d, b->seg_pass_level.data(), nseg * 4, hipMemcpyHostToDevice, st) !=
            hipSuccess) {
            destroy(b);
            return KMC_ERR_DEVICE;
        }
    }
    s = plan_tiles(b, 16, 0);
    if (s == KMC_OK && plan_only) { *out = b; return KMC_OK; }
    if (s == KMC_OK) s = run(b, st);            // sizes the arena (first capture writes nothing)
    if (s == KMC_OK && (flags & KMC_BATCH_AUTOTUNE)) s = autotune(b, st);
    if (s != KMC_OK) { destroy(b); return s; }
    *out = b;
    return KMC_OK;
}

void surface_view(const kmc_batch* b, uint32_t i, kmc_surface* o)
{
    const UserJob& u = b->users[b->surface_user[i]];
    const uint32_t l = i - u.first_surface;
    init_surface(o, u.levels[l], u.normal_type, u.dup, u.mn, u.mx, u.minc, u.maxc);
    o->memory = KMC_MEMORY_DEVICE;
    if (u.empty) { o->empty_range = 1; return; }
    const uint32_t sg = b->surface_seg[i];
    const SegOut& so = b->segout_h[sg];
    const uint64_t verts = u.dup ? 3 * so.nt : so.nv;
    if (verts == 0) return;
    o->n_vertices = verts;
    o->n_triangles = so.nt;
    o->n_connections = u.dup ? 0 : 3 * so.nt;
    o->n_normals = (!u.dup && u.normal_type == KMC_VERTEX_NORMAL) ? so.nv : so.nt;
    o->coords = b->arena_c + so.oc;
    o->normals = b->arena_n + so.on;
    o->connections = u.dup ? nullptr : b->arena_k + so.ok;
}

}  // namespace

/* ------------------------------------------------------------------------- */
/* two-phase use without a graph (the serving loop)                            */
/* ------------------------------------------------------------------------- */

namespace kmc {

kmc_status batch_plan(const kmc_job* jobs, uint32_t n_jobs, hipStream_t st, kmc_batch** out)
{
    return create_impl(jobs, n_jobs, 0, st, out, true);
}

hipError_t batch_launch_front(kmc_batch* b, hipStream_t st) { return record_front(b, st, true); }

uint32_t batch_surfaces(const kmc_batch* b) { return b->n_surfaces; }

/* Sizes of surface i after the front phase has completed (arrays unset). */
void batch_surface_sizes(const kmc_batch* b, uint32_t i, kmc_surface* o)
{
    surface_view(b, i, o);
    o->coords = nullptr;
    o->normals = nullptr;
    o->connections = nullptr;
}

/* Point surface i's outputs at caller memory sized as batch_surface_sizes
 * says (nullptr: write nothing for it). */
void batch_set_outputs(kmc_batch* b, uint32_t i, float* coords, float* normals, uint32_t* conn)
{
    const uint32_t sg = b->surface_seg[i];
    if (sg == ~0u) return;
    const uint32_t m = b->seg_pass_level[sg];
    Pass& p = b->passes[m >> 3];
    const uint32_t l = m & 7u;
    const SegOut& so = b->segout_h[sg];
    const uint64_t verts = p.a.dup ? 3 * so.nt : so.nv;
    const bool on = coords != nullptr && verts > 0;
    p.a.coords[l] = coords;
    p.a.normals[l] = normals;
    p.a.conn[l] = conn;
    p.a.cap_vertices[l] = on ? verts : 0;
    p.a.cap_triangles[l] = on ? so.nt : 0;
}

hipError_t batch_launch_emit(kmc_batch* b, hipStream_t st)
{
    if (b->passes.empty()) return hipSuccess;
    if (b->flat) {
        /* records of every pass from the sizes of the front phase */
        const uint32_t np = (uint32_t)b->passes.size();
        std::vector<uint64_t> off(np);
        uint64_t bytes = 0;
        b->istart_h.assign(2 * np + 2, 0);
        uint64_t vsum = 0, tsum = 0;
        /* only surfaces with outputs (admitted by the caller) get records */
        auto level_items = [&](const Pass& p, uint32_t l, uint64_t& nv, uint64_t& nt) {
            const bool on = p.a.coords[l] != nullptr;
            nv = on ? b->segout_h[p.seg0 + l].nv : 0;
            nt = on ? b->segout_h[p.seg0 + l].nt : 0;
        };
        for (uint32_t i = 0; i < np; ++i) {
            const Pass& p = b->passes[i];
            uint64_t V = 0, T = 0;
            for (uint32_t l = 0; l < p.nlev; ++l) {
                uint64_t nv, nt;
                level_items(p, l, nv, nt);
                V += p.a.dup ? 0 : nv;
                T += nt;
            }
            off[i] = bytes;
            bytes += record_bytes(V, T, p.a.vnormal != 0);
            b->istart_h[i] = vsum;
            b->istart_h[np + 1 + i] = tsum;
            vsum += V;
            tsum += T;
        }
        b->istart_h[np] = vsum;
        b->istart_h[2 * np + 1] = tsum;
        if (bytes > b->cap_r) {
            device_free(b->arena_r);
            b->arena_r = (uint8_t*)device_alloc(bytes);
            b->cap_r = b->arena_r ? bytes : 0;
            if (!b->arena_r) return hipErrorOutOfMemory;
        }
        for (uint32_t i = 0; i < np; ++i) {
            Pass& p = b->passes[i];
            uint64_t nv[kMaxLevels], nt[kMaxLevels];
            for (uint32_t l = 0; l < p.nlev; ++l) level_items(p, l, nv[l], nt[l]);
            place_records(p.a, nv, nt, b->arena_r + off[i]);
        }
        const hipError_t e = hipMemcpyAsync(b->istart_d, b->istart_h.data(), b->istart_h.size() * 8,
                                            hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return e;
    }
    b->jobs_h.resize(b->passes.size());
    for (size_t i = 0; i < b->passes.size(); ++i) b->jobs_h[i] = b->passes[i].a;
    const hipError_t e = hipMemcpyAsync(b->jobs_d, b->jobs_h.data(), b->jobs_h.size() * sizeof(JobArgs),
                                        hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    return record_emit(b, st, true);
}

/* Device bytes the plan holds besides the volumes (counts, prefixes, masks,
 * chunk prefixes, tables). */
uint64_t batch_work_bytes(const kmc_batch* b)
{
    return b->n_counts * 12 + b->bitvol_words * (b->flat_ok ? 48 : 8) + b->passes.size() * (sizeof(JobArgs) + 16) +
           b->seg_pass_level.size() * (4 + sizeof(SegOut)) + (uint64_t)(b->scan_parts + 1) * 8;
}

uint64_t batch_record_arena_bytes(const kmc_batch* b) { return b->flat ? b->cap_r : 0; }

/* Flat-emit record bytes surface i adds once its sizes are known. */
uint64_t batch_record_bytes(const kmc_batch* b, uint32_t i)
{
    const uint32_t sg = b->surface_seg[i];
    if (!b->flat || sg == ~0u) return 0;
    const Pass& p = b->passes[b->seg_pass_level[sg] >> 3];
    const SegOut& so = b->segout_h[sg];
    return record_bytes(p.a.dup ? 0 : so.nv, so.nt, p.a.vnormal != 0);
}

void batch_destroy(kmc_batch* b) { destroy(b); }

}  // namespace kmc

extern "C" {

kmc_status kmc_batch_create(const kmc_job* jobs, uint32_t n_jobs, uint32_t flags, void* stream, kmc_batch** out)
{
    return kmc::abi_guard([&]() -> kmc_status {
        const kmc_status s = create_impl(jobs, n_jobs, flags, (hipStream_t)stream, out);
        report(s);
        return s;
    });
}

kmc_status kmc_batch_run(kmc_batch* b, void* stream)
{
    return kmc::abi_guard([&]() -> kmc_status {
        if (!b) { report(KMC_ERR_INVALID_ARGUMENT); return KMC_ERR_INVALID_ARGUMENT; }
        if (hipSetDevice(b->device) != hipSuccess) { report(KMC_ERR_DEVICE); return KMC_ERR_DEVICE; }
        const kmc_status s = run(b, (hipStream_t)stream);
        report(s);
        return s;
    });
}

kmc_status kmc_batch_surfaces(const kmc_batch* b, kmc_surface* outs, uint32_t n_outs)
{
    return kmc::abi_guard([&]() -> kmc_status {
        if (!b || !outs || n_outs < b->n_surfaces) return KMC_ERR_INVALID_ARGUMENT;
        if (!b->have_results) return KMC_ERR_NOT_READY;
        for (uint32_t i = 0; i < b->n_surfaces; ++i) surface_view(b, i, &outs[i]);
        return KMC_OK;
    });
}

kmc_status kmc_batch_info_get(const kmc_batch* b, kmc_batch_info* info)
{
    return kmc::abi_guard([&]() -> kmc_status {
        if (!b || !info) return KMC_ERR_INVALID_ARGUMENT;
        std::memset(info, 0, sizeof(*info));
        info->n_jobs = (uint32_t)b->users.size();
        info->n_surfaces = b->n_surfaces;
        info->n_passes = (uint32_t)b->passes.size();
        info->n_type_groups = (uint32_t)b->groups.size();
        info->tile_rows = b->r_max;
        info->queue_capacity = b->queue_cap;
        info->graph_nodes = b->graph_nodes;
        info->recaptures = b->recaptures;
        for (const Pass& p : b->passes) info->nodes_classified += (uint64_t)p.a.nx * p.a.ny * p.a.nz * p.nlev;
        if (b->have_results)
            for (uint32_t s = 0; s < b->seg_pass_level.size(); ++s) info->n_triangles += b->segout_h[s].nt;
        info->arena_bytes = 4 * (b->cap_c + b->cap_n + b->cap_k);
        info->work_bytes = b->n_counts * 12 + b->bitvol_words * 8;
        info->last_run_ms = b->last_ms;
        info->flat_emit = b->flat ? 1u : 0u;
        return KMC_OK;
    });
}

void kmc_batch_destroy(kmc_batch* b) { destroy(b); }

kmc_status kmc_batch_extract(const kmc_job* jobs, uint32_t n_jobs, int32_t out_memory, void* stream,
                             kmc_surface* outs)
{
    return kmc::abi_guard([&]() -> kmc_status {
        if (out_memory != KMC_MEMORY_HOST && out_memory != KMC_MEMORY_DEVICE) {
            report(KMC_ERR_INVALID_ARGUMENT);
            return KMC_ERR_INVALID_ARGUMENT;
        }
        kmc_batch* b = nullptr;
        kmc_status s = create_impl(jobs, n_jobs, 0, (hipStream_t)stream, &b);
        if (s != KMC_OK) { report(s); return s; }
        if (!outs && b->n_surfaces > 0) { destroy(b); report(KMC_ERR_INVALID_ARGUMENT); return KMC_ERR_INVALID_ARGUMENT; }
        hipStream_t st = (hipStream_t)stream;
        const hipMemcpyKind kind = out_memory == KMC_MEMORY_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
        uint32_t done = 0;
        for (uint32_t i = 0; i < b->n_surfaces && s == KMC_OK; ++i, ++done) {
            kmc_surface v;
            surface_view(b, i, &v);
            SurfaceStorage* store = new SurfaceStorage();
            store->memory = out_memory;
            store->coords = store->connections = store->normals = nullptr;
            const size_t bc = v.n_vertices * 12, bn = v.n_normals * 12, bk = v.n_connections * 4;
            auto alloc = [&](void** p, size_t n) -> bool {
                if (n == 0) return true;
                if (out_memory == KMC_MEMORY_HOST) return (*p = host_block_alloc(n)) != nullptr;
                *p = device_alloc(n);
                return *p != nullptr;
            };
            bool ok = alloc(&store->coords, bc) && alloc(&store->normals, bn) && alloc(&store->connections, bk);
            if (ok && bc) ok = hipMemcpyAsync(store->coords, v.coords, bc, kind, st) == hipSuccess;
            if (ok && bn) ok = hipMemcpyAsync(store->normals, v.normals, bn, kind, st) == hipSuccess;
            if (ok && bk) ok = hipMemcpyAsync(store->connections, v.connections, bk, kind, st) == hipSuccess;
            v.memory = out_memory;
            v.coords = (float*)store->coords;
            v.normals = (float*)store->normals;
            v.connections = (uint32_t*)store->connections;
            v.internal = store;
            outs[i] = v;
            if (!ok) { (void)hipGetLastError(); s = KMC_ERR_OUT_OF_MEMORY; }
        }
        if (s == KMC_OK && hipStreamSynchronize(st) != hipSuccess) s = KMC_ERR_DEVICE;
        if (s != KMC_OK) {
            /* no copy still in flight may touch blocks that go back to the caches */
            (void)hipStreamSynchronize(st);
            (void)hipGetLastError();
            for (uint32_t i = 0; i < done; ++i) kmc_surface_release(&outs[i]);
        }
        destroy(b);
        report(s);
        return s;
    });
}

}  // extern "C"
