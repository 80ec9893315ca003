/*
 * kmc_noalloc.hip -- kmc_extract_noalloc: the plain one-level extraction with
 * no device allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * The caller passes the volume in device memory and the output arrays with
 * their capacities.  Every work array of the call -- classification masks,
 * per-row counts and prefixes, in-row chunk prefixes, chunk crossing records,
 * the emit's records, the min/max partials -- is carved from `scratch`, a
 * fixed device array of KMC_NOALLOC_SCRATCH_BYTES that each GPU the library
 * runs on holds until the process ends.  None of these arrays needs clearing:
 * every element a kernel reads was written earlier in the same call.
 *
 * Host code never passes `scratch` itself to a kernel or a copy: in host code
 * the name stands for a host-side shadow, not the device array.  Its device
 * address comes from hipGetSymbolAddress, once per device with that device
 * current, and is kept per device.
 */
#include <hip/hip_runtime.h>
#include <cstring>
#include <mutex>

#include "kmc.h"
#include "kmc_guard.h"
#include "kmc_internal.h"
#include "kmc_job.h"

namespace kmc {

namespace {

static __device__ __align__(256) uint8_t scratch[KMC_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;

struct DeviceScratch {
    std::once_flag once;
    hipError_t status = hipErrorNotInitialized;
    uint8_t* ptr = nullptr;     // device address of `scratch` on this device
    std::mutex mu;              // calls on one device share its array: one at a time
};
DeviceScratch g_scratch[kMaxDevices];

/* The device address of this device's `scratch` (the calling thread's
 * current device), looked up on the first call on the device. */
hipError_t scratch_of(int dev, uint8_t** out)
{
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    DeviceScratch& d = g_scratch[dev];
    std::call_once(d.once, [&d]() {
        void* p = nullptr;
        d.status = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (d.status == hipSuccess && !p) d.status = hipErrorInvalidSymbol;
        if (d.status != hipSuccess) (void)hipGetLastError();
        d.ptr = (uint8_t*)p;
    });
    *out = d.ptr;
    return d.status;
}

inline uint64_t up256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

/* Bump allocation inside the scratch array. */
struct Carve {
    uint8_t* base;
    uint64_t used = 0;
    template <typename T>
    T* take(uint64_t bytes)
    {
        T* p = (T*)(base + used);
        used += up256(bytes);
        return p;
    }
};

/* Scratch bytes of the arrays sized by the volume alone. */
uint64_t front_bytes(const JobArgs& a, bool minmax)
{
    const uint64_t R = a.nrows, C = a.nrows * a.NC;
    const uint64_t nparts = scan_partials_needed(R, 2);
    return up256(2 * R * 4) + up256(2 * R * 8) + up256((nparts + 2) * 8) + up256(C * 8) + up256(2 * C * 4) +
           (a.dup ? 0 : up256(C * sizeof(ChunkCross))) + (minmax ? up256(minmax_scratch_bytes() + 16) : 0);
}

/* Scratch bytes of the emit's records for nv vertices and nt triangles. */
uint64_t record_bytes(const JobArgs& a, uint64_t nv, uint64_t nt)
{
    return up256(nv * 8) + up256(nt * 8) + up256(nt) + (a.vnormal ? up256(nt * 12) + up256(nv * 16) : 0);
}

kmc_status extract_noalloc(const kmc_volume* v, double iso, int32_t normal_type, int32_t duplication,
                           hipStream_t st, float* coords, uint64_t coords_cap, uint32_t* conn, uint64_t conn_cap,
                           float* normals, uint64_t normals_cap, kmc_surface* out)
{
    kmc_status s = validate_volume(v);
    if (s != KMC_OK) return s;
    if (!out || v->memory != KMC_MEMORY_DEVICE) return KMC_ERR_INVALID_ARGUMENT;
    if (normal_type != KMC_VERTEX_NORMAL && normal_type != KMC_POLYGON_NORMAL) return KMC_ERR_INVALID_ARGUMENT;
    if (!flat_emit_fits(v->resolution[1], v->resolution[2])) return KMC_ERR_INVALID_ARGUMENT;
    const int dup = (normal_type == KMC_VERTEX_NORMAL) ? 0 : (duplication ? 1 : 0);   // MarchingCubes.cpp:69-73
    float minc[3], maxc[3];
    object_coords(v, minc, maxc);

    /* the work the volume alone fixes must fit before anything runs */
    JobArgs a;
    init_job(a, v, v->values, &iso, 1, normal_type, dup, v->min_value, v->max_value, minc, maxc);
    plan_rows(a);
    a.rec_levels = 1u;
    const uint64_t need_front = front_bytes(a, !v->has_min_max_values);
    if (need_front > KMC_NOALLOC_SCRATCH_BYTES) return KMC_ERR_SCRATCH;

    s = ensure_device();
    if (s != KMC_OK) return s;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { (void)hipGetLastError(); return KMC_ERR_DEVICE; }
    uint8_t* base = nullptr;
    if (scratch_of(dev, &base) != hipSuccess) return KMC_ERR_DEVICE;
    std::lock_guard<std::mutex> lk(g_scratch[dev].mu);

    Carve cv{base};
    const uint64_t R = a.nrows, C = a.nrows * a.NC;
    const uint32_t nparts = scan_partials_needed(R, 2);
    uint32_t* counts = cv.take<uint32_t>(2 * R * 4);
    uint64_t* prefix = cv.take<uint64_t>(2 * R * 8);
    uint64_t* aux = cv.take<uint64_t>((nparts + 2) * 8);
    a.bitvol = cv.take<uint64_t>(C * 8);
    a.vpc = cv.take<uint32_t>(2 * C * 4);
    a.tpc = a.vpc + C;
    a.cross = dup ? nullptr : cv.take<ChunkCross>(C * sizeof(ChunkCross));
    a.counts = counts;
    a.prefix = prefix;

    double mn = v->min_value, mx = v->max_value;
    if (!v->has_min_max_values) {
        /* GetMinMaxValues<T> (StructuredVolumeObject.cpp:66-111) in the scratch */
        uint8_t* mm = cv.take<uint8_t>(minmax_scratch_bytes() + 16);
        double* dout = (double*)(mm + minmax_scratch_bytes());
        double host[2] = { 0, 0 };
        const uint64_t n = (uint64_t)v->resolution[0] * v->resolution[1] * v->resolution[2];
        if (launch_minmax(v->type, v->values, n, mm, dout, st) != hipSuccess ||
            hipMemcpyAsync(host, dout, sizeof(host), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            (void)hipGetLastError();
            return KMC_ERR_DEVICE;
        }
        mn = host[0];
        mx = host[1];
        a.minv = mn;
        a.maxv = mx;
    }
    init_surface(out, iso, normal_type, dup, mn, mx, minc, maxc);
    const double d = mn - mx;
    if ((d < 0 ? -d : d) < 1.e-6) {
        out->empty_range = 1;   // Math::Equal(min, max): no surface (MarchingCubes.cpp:108-110)
        return KMC_OK;
    }

    uint64_t* partials = aux;
    uint64_t* totals = aux + nparts;
    if (launch_classify(v->type, a, st) != hipSuccess || launch_count(a, st) != hipSuccess ||
        launch_scan(counts, R, 2, partials, prefix, totals, st) != hipSuccess) {
        (void)hipGetLastError();
        return KMC_ERR_DEVICE;
    }
    uint64_t tot[2] = { 0, 0 };
    if (hipMemcpyAsync(tot, totals, sizeof(tot), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        (void)hipGetLastError();
        return KMC_ERR_DEVICE;
    }
    const uint64_t nv = tot[0], nt = tot[1];
    const uint64_t verts = dup ? 3 * nt : nv;
    if (verts == 0) return KMC_OK;   // coords.size() == 0: no arrays (:236-242, :282-289)
    out->n_vertices = verts;
    out->n_triangles = nt;
    out->n_connections = dup ? 0 : 3 * nt;
    out->n_normals = (!dup && normal_type == KMC_VERTEX_NORMAL) ? nv : nt;
    if (3 * out->n_vertices > coords_cap || 3 * out->n_normals > normals_cap || out->n_connections > conn_cap ||
        !coords || !normals || (out->n_connections && !conn))
        return KMC_ERR_CAPACITY;   // the sizes stay in *out; nothing was written
    if (cv.used + record_bytes(a, dup ? 0 : nv, nt) > KMC_NOALLOC_SCRATCH_BYTES) return KMC_ERR_SCRATCH;

    a.voff[0] = a.toff[0] = 0;
    a.voff[1] = dup ? 0 : nv;
    a.toff[1] = nt;
    a.vkey = cv.take<uint64_t>(a.voff[1] * 8);
    a.trec = cv.take<uint64_t>(nt * 8);
    a.tci = cv.take<uint8_t>(nt);
    a.fn = a.vnormal ? cv.take<float>(nt * 12) : nullptr;
    a.vslot = a.vnormal ? cv.take<uint32_t>(nv * 16) : nullptr;
    a.coords[0] = coords;
    a.normals[0] = normals;
    a.conn[0] = dup ? nullptr : conn;
    a.cap_vertices[0] = verts;
    a.cap_triangles[0] = nt;
    if (launch_flat_emit(v->type, a, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
        (void)hipGetLastError();
        return KMC_ERR_DEVICE;
    }
    out->coords = coords;
    out->normals = normals;
    out->connections = dup ? nullptr : conn;
    return KMC_OK;
}

}  // namespace

}  // namespace kmc

extern "C" kmc_status kmc_extract_noalloc(const kmc_volume* volume, double isolevel, int32_t normal_type,
                                          int32_t duplication, void* stream, float* coords, uint64_t coords_capacity,
                                          uint32_t* connections, uint64_t connections_capacity, float* normals,
                                          uint64_t normals_capacity, kmc_surface* out)
{
    return kmc::abi_guard([&]() -> kmc_status {
        const kmc_status s = kmc::extract_noalloc(volume, isolevel, normal_type, duplication, (hipStream_t)stream, coords,
                                                  coords_capacity, connections, connections_capacity, normals,
                                                  normals_capacity, out);
        kmc::report(s);
        return s;
    });
}
