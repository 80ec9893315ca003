/*
 * kmc_internal.h -- helpers shared by the host-side modules of libkmc.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "kmc.h"
#include "kmc_job.h"

namespace kmc {

/* Caching device allocator: blocks return to a free list and are reused. */
void* device_alloc(size_t bytes);
void device_free(void* p);
void device_pool_trim();

/* Pooled pinned host memory for small staging buffers. */
void* host_pinned_alloc(size_t bytes);
void host_pinned_free(void* p);
void* host_block_alloc(size_t bytes);
void host_block_free(void* p);
hipError_t read_back(void* host, const void* dev, size_t n, hipStream_t st);
/* Pageable host memory to device memory (pinned staging for large copies). */
hipError_t upload_host(void* dst, const void* src, size_t bytes, hipStream_t st);
/* Several device ranges into host memory with one wait for the stream. */
struct ReadPart {
    void* host;
    const void* dev;
    size_t bytes;
};
hipError_t read_back_parts(const ReadPart* parts, int n, hipStream_t st);

/* Arrays owned by a kmc_surface (its `internal` handle). */
struct SurfaceStorage {
    int memory;
    void* coords;
    void* connections;
    void* normals;
    void* colors;
};
void free_storage(SurfaceStorage* s);
/* Move a device surface (arrays owned by its SurfaceStorage) to pinned host memory. */
kmc_status to_host(kmc_surface* s, hipStream_t st);

size_t type_size(int t);
float float_threshold(double iso);
kmc_status validate_volume(const kmc_volume* v);
void report(kmc_status s);
kmc_status ensure_device();
kmc_status device_values(const kmc_volume* v, hipStream_t st, const void** dvals, void** owned);
kmc_status min_max_on_device(const kmc_volume* v, const void* dvals, hipStream_t st, double* mn, double* mx);
void object_coords(const kmc_volume* v, float minc[3], float maxc[3]);
void init_surface(kmc_surface* s, double iso, int normal_type, int dup, double mn, double mx,
                  const float minc[3], const float maxc[3]);

/* Fill the per-job fields of a (geometry, levels, clamp range, coordinates). */
void init_job(JobArgs& a, const kmc_volume* v, const void* dvals, const double* iso, uint32_t nlev, int normal_type,
              int dup, double mn, double mx, const float minc[3], const float maxc[3]);

/* Row geometry of every emit (a.nx, a.ny, a.nz set): a.NC, a.cshift, a.nrows. */
void plan_rows(JobArgs& a);

/* True when the flat emit serves a volume of these y / z sizes (its records
 * pack y and z in 20 bits) and KMC_EMIT does not ask for the tile teams. */
bool flat_emit_fits(uint32_t ny, uint32_t nz);

/* Tile geometry and emit LDS size (a.nlev set): the tallest band R <= r_max
 * whose team LDS state fits the per-team budget (queue_cap 0: the default).
 * False if not even R = 1 fits a workgroup. */
bool plan_geometry(JobArgs& a, uint32_t* lds_bytes, uint32_t r_max = 16, uint32_t queue_cap = 0);

/* Isolevels per pass over this volume: kMaxLevels with the flat emit; with
 * the tile teams (one level per team) kMaxLevels whenever a team's LDS state
 * fits at all, 0 if it does not (for VertexNormal when vnormal). */
uint32_t levels_per_pass(const kmc_volume* v, int vnormal);

/* Kernel launchers (kmc_kernels.hip). */
hipError_t launch_emit(int dtype, const JobArgs& a, uint32_t lds_bytes, hipStream_t st);
/* Flat emit of a pass (a.vpc / a.tpc written by its count pass, a.vkey /
 * a.trec / a.tci / a.fn and a.voff / a.toff set): compaction records, then
 * one thread per vertex, per triangle and (VertexNormal) per vertex normal.
 * marks: three events recorded after the compaction, the vertices and the
 * triangles (timing), or null. */
hipError_t launch_flat_emit(int dtype, const JobArgs& a, hipStream_t st, hipEvent_t* marks = nullptr);
/* Flat emit of a batch group: compaction over the rows of its jobs, then the
 * per-item kernels over the items start[j] .. start[j + 1] of each job. */
hipError_t launch_compact_batch(const JobArgs* jobs, const uint64_t* row_start, uint32_t njobs, uint64_t rows,
                                hipStream_t st);
/* the per-item kernels of one pass whose record offsets are set (after its compaction) */
hipError_t launch_flat_items(int dtype, const JobArgs& a, hipStream_t st);
hipError_t launch_flat_batch(int dtype, int mode, const JobArgs* jobs, const uint64_t* vstart, const uint64_t* tstart,
                             uint32_t njobs, hipStream_t st);
hipError_t launch_classify(int dtype, const JobArgs& a, hipStream_t st);
/* SlicePlane: masks from the plane's sign (a.plane), and the duplicated emit
 * (compaction, then one thread per triangle with per-vertex colours). */
hipError_t launch_classify_plane(const JobArgs& a, hipStream_t st);
hipError_t launch_slice_emit(int dtype, const JobArgs& a, hipStream_t st);
/* SlicePlane by node-row thresholds: thresholds (ny * nz int32) and per cell
 * row triangle counts ((ny - 1) * (nz - 1) uint32); then, with their
 * exclusive prefix, one thread per triangle (a.toff[1] triangles) */
hipError_t launch_slice_runs_front(const JobArgs& a, int32_t* thr, uint32_t* counts, hipStream_t st);
hipError_t launch_slice_runs_emit(int dtype, const JobArgs& a, const int32_t* thr, const uint32_t* counts,
                                  const uint64_t* prefix, uint64_t* units, hipStream_t st);
hipError_t launch_count(const JobArgs& a, hipStream_t st);
hipError_t launch_scan(const uint32_t* counts, uint64_t n, uint32_t narr, uint64_t* partials,
                       uint64_t* prefix, uint64_t* totals, hipStream_t st);
uint32_t scan_partials_needed(uint64_t n, uint32_t narr);
hipError_t launch_minmax(int dtype, const void* vol, uint64_t n, void* scratch, double* out, hipStream_t st);
size_t minmax_scratch_bytes();
/* The unstructured form: vol[0], then vol[conn[i]] for every connection entry in order. */
hipError_t launch_minmax_gather(int dtype, const void* vol, const uint32_t* conn, uint64_t nconn, void* scratch,
                                double* out, hipStream_t st);
hipError_t launch_classify_batch(int dtype, const JobArgs* jobs, const uint64_t* row_start, uint32_t njobs,
                                 uint64_t rows, uint32_t max_levels, hipStream_t st);
hipError_t launch_count_batch(const JobArgs* jobs, const uint64_t* row_start, uint32_t njobs, uint64_t rows,
                              hipStream_t st);
/* mode: 0 duplicated vertices, 1 PolygonNormal, 2 VertexNormal (shared vertices) */
hipError_t launch_emit_batch(int dtype, int mode, const JobArgs* jobs, const uint64_t* unit_start, uint32_t njobs,
                             uint64_t units, uint32_t lds_bytes, uint32_t teams, hipStream_t st);

/* Two-phase use of a batch plan without a graph (kmc_batch.hip): plan, launch
 * the front (classify, count, scan, sizes), wait, read the sizes, point the
 * outputs, launch the emit. */
kmc_status batch_plan(const kmc_job* jobs, uint32_t n_jobs, hipStream_t st, kmc_batch** out);
hipError_t batch_launch_front(kmc_batch* b, hipStream_t st);
uint32_t batch_surfaces(const kmc_batch* b);
void batch_surface_sizes(const kmc_batch* b, uint32_t i, kmc_surface* o);
void batch_set_outputs(kmc_batch* b, uint32_t i, float* coords, float* normals, uint32_t* conn);
/* Flat-emit record bytes surface i adds once its sizes are known. */
uint64_t batch_record_bytes(const kmc_batch* b, uint32_t i);
hipError_t batch_launch_emit(kmc_batch* b, hipStream_t st);
uint64_t batch_work_bytes(const kmc_batch* b);
/* Device bytes of the flat-emit record arena the last batch_launch_emit used. */
uint64_t batch_record_arena_bytes(const kmc_batch* b);
void batch_destroy(kmc_batch* b);

}  // namespace kmc
