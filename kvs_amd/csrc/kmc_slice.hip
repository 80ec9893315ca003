/*
 * kmc_slice.hip -- kmc_slice_plane: kvs::SlicePlane over a structured volume
 * (Source/Core/Visualization/Mapper/SlicePlane.cpp:97-280) on the device.
 *
 * The plane takes the place of the isosurface: nodes classify by the sign of
 * the plane equation (no value is read for that), the count / scan /
 * compaction passes of the marching-cubes path run unchanged in duplicated
 * mode, and one thread per triangle interpolates its three vertices along the
 * plane, colours each with the value interpolated at the same ratio, and
 * writes the negated face normal -- triangles in cell scan order, the order
 * of the reference's loop.
 */
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "kmc.h"
#include "kmc_guard.h"
#include "kmc_internal.h"
#include "kmc_job.h"

namespace kmc {

namespace {

inline size_t up256(size_t x) { return (x + 255) & ~size_t(255); }

/* MapperBase::setRange (MapperBase.cpp:54-83) for a transfer function
 * without a range of its own. */
void default_color_range(int type, double mn, double mx, float range[2])
{
    if (type == KMC_INT8) { range[0] = -128; range[1] = 127; return; }
    if (type == KMC_UINT8) { range[0] = 0; range[1] = 255; return; }
    range[0] = (float)mn;
    range[1] = (float)mx;
    const float d = range[0] - range[1];
    if ((double)(d < 0 ? -d : d) < 1.e-6) range[1] = range[1] + 1.0f;   // Math::Equal -> max + 1
}

/* SlicePlane::substitute_plane_equation in float, left to right (:936-971). */
float plane_value(const float c[4], float x, float y, float z)
{
    float s = c[0] * x;
    s = s + c[1] * y;
    s = s + c[2] * z;
    return s + c[3];
}

/* The threshold path (one threshold per node row, the plane being monotone
 * along x) when the plane value is finite at the eight corners of the node
 * box -- monotone in each coordinate, it is then finite everywhere inside --
 * unless KMC_SLICE=classify asks for the classification of every node. */
bool threshold_path(const kmc_volume* v, const float coef[4])
{
    const char* env = std::getenv("KMC_SLICE");
    if (env && std::strcmp(env, "classify") == 0) return false;
    for (int k = 0; k < 8; ++k) {
        const float x = (k & 1) ? (float)(v->resolution[0] - 1) : 0.0f;
        const float y = (k & 2) ? (float)(v->resolution[1] - 1) : 0.0f;
        const float z = (k & 4) ? (float)(v->resolution[2] - 1) : 0.0f;
        if (!std::isfinite(plane_value(coef, x, y, z))) return false;
    }
    return true;
}

/* Thresholds, per cell row triangle counts, their scan, then one thread per
 * triangle (no classification masks, no records). */
template <typename Alloc>
kmc_status slice_by_thresholds(const kmc_volume* v, JobArgs& a, hipStream_t st, Alloc& alloc, kmc_surface* out,
                               double mn, double mx, const float minc[3], const float maxc[3], const uint8_t* table,
                               uint32_t res)
{
    const uint64_t rows = (uint64_t)a.ny * a.nz, ncr = (uint64_t)(a.ny - 1) * (a.nz - 1);
    const uint32_t nparts = scan_partials_needed(ncr ? ncr : 1, 1);
    int32_t* thr = (int32_t*)alloc(rows * 4);
    uint32_t* counts = (uint32_t*)alloc((ncr ? ncr : 1) * 4);
    uint64_t* prefix = (uint64_t*)alloc((ncr ? ncr : 1) * 8);
    uint64_t* aux = (uint64_t*)alloc((nparts + 1) * 8);
    uint8_t* dtable = (uint8_t*)alloc(3 * (size_t)res);
    if (!thr || !counts || !prefix || !aux || !dtable) return KMC_ERR_OUT_OF_MEMORY;
    a.ctable = dtable;
    uint64_t nt = 0;
    if (hipMemcpyAsync(dtable, table, 3 * (size_t)res, hipMemcpyHostToDevice, st) != hipSuccess ||
        (ncr && (launch_slice_runs_front(a, thr, counts, st) != hipSuccess ||
                 launch_scan(counts, ncr, 1, aux, prefix, aux + nparts, st) != hipSuccess ||
                 read_back(&nt, aux + nparts, 8, st) != hipSuccess)) ||
        hipStreamSynchronize(st) != hipSuccess) {
        (void)hipGetLastError();
        return KMC_ERR_DEVICE;
    }
    init_surface(out, 0.0, KMC_POLYGON_NORMAL, 1, mn, mx, minc, maxc);
    out->color_type = 0;   // VertexColor
    SurfaceStorage* store = new SurfaceStorage();
    store->memory = KMC_MEMORY_DEVICE;
    out->internal = store;
    out->memory = KMC_MEMORY_DEVICE;
    if (nt == 0) return KMC_OK;
    out->n_triangles = nt;
    out->n_vertices = 3 * nt;
    out->n_normals = nt;
    out->n_colors = 3 * nt;
    store->coords = device_alloc(9 * nt * sizeof(float));
    store->normals = device_alloc(3 * nt * sizeof(float));
    store->colors = device_alloc(9 * nt);
    if (!store->coords || !store->normals || !store->colors) {
        kmc_surface_release(out);
        return KMC_ERR_OUT_OF_MEMORY;
    }
    out->coords = (float*)store->coords;
    out->normals = (float*)store->normals;
    out->colors = (uint8_t*)store->colors;
    a.toff[0] = 0;
    a.toff[1] = nt;
    a.coords[0] = out->coords;
    a.normals[0] = out->normals;
    a.colors = out->colors;
    a.cap_vertices[0] = 3 * nt;
    a.cap_triangles[0] = nt;
    uint64_t* units = (uint64_t*)alloc(2 * ((nt + 63) / 64) * 8);   // first and last cell row of each 64 triangles
    if (!units) {
        kmc_surface_release(out);
        return KMC_ERR_OUT_OF_MEMORY;
    }
    if (launch_slice_runs_emit(v->type, a, thr, counts, prefix, units, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        (void)hipGetLastError();
        kmc_surface_release(out);
        return KMC_ERR_DEVICE;
    }
    return KMC_OK;
}

kmc_status slice_plane_impl(const kmc_volume* v, const float coef[4], const uint8_t* table, uint32_t res,
                            const float* color_range, int32_t out_memory, hipStream_t st, kmc_surface* out)
{
    /* SlicePlane::exec / mapping (:65-150) */
    if (!v) return KMC_ERR_NULL_INPUT;
    if (v->kind != KMC_STRUCTURED) return KMC_ERR_NOT_STRUCTURED;
    if (v->veclen != 1) return KMC_ERR_NOT_SCALAR;
    kmc_status s = validate_volume(v);
    if (s != KMC_OK) return s;
    if (!coef || !table || res == 0 || !out) return KMC_ERR_INVALID_ARGUMENT;
    if (out_memory != KMC_MEMORY_HOST && out_memory != KMC_MEMORY_DEVICE) return KMC_ERR_INVALID_ARGUMENT;
    if (!flat_emit_fits(v->resolution[1], v->resolution[2])) return KMC_ERR_INVALID_ARGUMENT;
    s = ensure_device();
    if (s != KMC_OK) return s;

    const void* dvals = nullptr;
    void* owned = nullptr;
    s = device_values(v, st, &dvals, &owned);
    if (s != KMC_OK) return s;
    double mn = v->min_value, mx = v->max_value;
    /* setRange computes the volume's min/max when it has none cached; only
     * the colour range uses them */
    if (!v->has_min_max_values) s = min_max_on_device(v, dvals, st, &mn, &mx);
    float minc[3], maxc[3];
    object_coords(v, minc, maxc);

    std::vector<void*> work;
    auto alloc = [&](size_t bytes) -> void* {
        void* p = device_alloc(bytes ? bytes : 1);
        if (p) work.push_back(p);
        return p;
    };
    auto finish = [&](kmc_status status) {
        if (owned || !work.empty()) (void)hipStreamSynchronize(st);
        for (void* p : work) device_free(p);
        device_free(owned);
        return status;
    };
    if (s != KMC_OK) return finish(s);

    JobArgs a;
    const double iso = 0.0;
    init_job(a, v, dvals, &iso, 1, KMC_POLYGON_NORMAL, 1, mn, mx, minc, maxc);
    plan_rows(a);
    a.rec_levels = 1u;
    for (int k = 0; k < 4; ++k) a.plane[k] = coef[k];
    if (color_range) { a.crange[0] = color_range[0]; a.crange[1] = color_range[1]; }
    else default_color_range(v->type, mn, mx, a.crange);
    a.cres = res;

    if (threshold_path(v, coef)) {
        s = slice_by_thresholds(v, a, st, alloc, out, mn, mx, minc, maxc, table, res);
        if (s == KMC_OK && out_memory == KMC_MEMORY_HOST) {
            s = to_host(out, st);
            if (s != KMC_OK) kmc_surface_release(out);
        }
        return finish(s);
    }

    const uint64_t R = a.nrows, C = a.nrows * a.NC;
    const uint32_t nparts = scan_partials_needed(R, 2);
    uint32_t* counts = (uint32_t*)alloc(2 * R * 4);
    uint64_t* prefix = (uint64_t*)alloc(2 * R * 8);
    uint64_t* aux = (uint64_t*)alloc((nparts + 2) * 8);
    a.bitvol = (uint64_t*)alloc(C * 8);
    a.vpc = (uint32_t*)alloc(2 * C * 4);
    uint8_t* dtable = (uint8_t*)alloc(3 * (size_t)res);
    if (!counts || !prefix || !aux || !a.bitvol || !a.vpc || !dtable) return finish(KMC_ERR_OUT_OF_MEMORY);
    a.tpc = a.vpc + C;
    a.counts = counts;
    a.prefix = prefix;
    a.ctable = dtable;
    if (hipMemcpyAsync(dtable, table, 3 * (size_t)res, hipMemcpyHostToDevice, st) != hipSuccess)
        return finish(KMC_ERR_DEVICE);

    uint64_t* partials = aux;
    uint64_t* totals = aux + nparts;
    uint64_t tot[2] = { 0, 0 };
    if (launch_classify_plane(a, st) != hipSuccess || launch_count(a, st) != hipSuccess ||
        launch_scan(counts, R, 2, partials, prefix, totals, st) != hipSuccess ||
        read_back(tot, totals, sizeof(tot), st) != hipSuccess) {
        (void)hipGetLastError();
        return finish(KMC_ERR_DEVICE);
    }

    /* the reference's arrays: coords and colours per triangle corner, a
     * normal per triangle (:267-276) */
    init_surface(out, 0.0, KMC_POLYGON_NORMAL, 1, mn, mx, minc, maxc);
    out->color_type = 0;   // VertexColor
    SurfaceStorage* store = new SurfaceStorage();
    store->memory = KMC_MEMORY_DEVICE;
    out->internal = store;
    const uint64_t nt = tot[1];
    if (nt > 0) {
        out->n_triangles = nt;
        out->n_vertices = 3 * nt;
        out->n_normals = nt;
        out->n_colors = 3 * nt;
        store->coords = device_alloc(9 * nt * sizeof(float));
        store->normals = device_alloc(3 * nt * sizeof(float));
        store->colors = device_alloc(9 * nt);
        uint8_t* recs = (uint8_t*)alloc(up256(nt * 8) + nt);
        if (!store->coords || !store->normals || !store->colors || !recs) {
            kmc_surface_release(out);
            return finish(KMC_ERR_OUT_OF_MEMORY);
        }
        out->coords = (float*)store->coords;
        out->normals = (float*)store->normals;
        out->colors = (uint8_t*)store->colors;
        a.voff[0] = a.voff[1] = 0;
        a.toff[0] = 0;
        a.toff[1] = nt;
        a.trec = (uint64_t*)recs;
        a.tci = recs + up256(nt * 8);
        a.coords[0] = out->coords;
        a.normals[0] = out->normals;
        a.colors = out->colors;
        a.cap_vertices[0] = 3 * nt;
        a.cap_triangles[0] = nt;
        if (launch_slice_emit(v->type, a, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
            (void)hipGetLastError();
            kmc_surface_release(out);
            return finish(KMC_ERR_DEVICE);
        }
    }
    out->memory = KMC_MEMORY_DEVICE;
    if (out_memory == KMC_MEMORY_HOST) {
        s = to_host(out, st);
        if (s != KMC_OK) kmc_surface_release(out);
    }
    return finish(s);
}

}  // namespace

}  // namespace kmc

extern "C" kmc_status kmc_slice_plane(const kmc_volume* volume, const float coefficients[4],
                                      const uint8_t* color_table, uint32_t color_resolution,
                                      const float* color_range, int32_t out_memory, void* stream, kmc_surface* out)
{
    return kmc::abi_guard([&]() -> kmc_status {
        const kmc_status s = kmc::slice_plane_impl(volume, coefficients, color_table, color_resolution, color_range,
                                                   out_memory, (hipStream_t)stream, out);
        kmc::report(s);
        return s;
    });
}
