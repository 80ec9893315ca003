/*
 * kmc_export.hip -- the output side: a surface expanded per triangle on the
 * device (the triangle soup that PolygonExporter<kvs::Stl> builds,
 * Exporter/PolygonExporter.cpp:91-160, and that the GLSL polygon renderer
 * builds for polygon colours / normals, Renderer/PolygonRendererGLSL.cpp:105-186),
 * and the STL file the reference writes from it (FileFormat/STL/Stl.cpp:447-540).
 *
 * The expansion gathers each triangle's three vertex coordinates through its
 * connections and, for VertexNormal surfaces, averages the three vertex
 * normals as the exporter does: (n0 + n1 + n2) / 3.0f in float.  A duplicated
 * surface is already a triangle soup and is copied as it is.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "kmc.h"
#include "kmc_internal.h"
#include "kmc_job.h"

#pragma clang fp contract(off)

namespace kmc {

namespace {

/* One thread per triangle: its corners (through the connections) and normal. */
__global__ __launch_bounds__(kBlock) void k_expand(const float* coords, const uint32_t* conn, const float* normals,
                                                    uint64_t nt, int vertex_normals, float* out_coords,
                                                    float* out_normals)
{
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < nt; t += step) {
        const uint32_t i0 = conn[3 * t + 0], i1 = conn[3 * t + 1], i2 = conn[3 * t + 2];
        const uint32_t id[3] = { i0, i1, i2 };
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            out_coords[9 * t + 3 * j + 0] = coords[3 * (uint64_t)id[j] + 0];
            out_coords[9 * t + 3 * j + 1] = coords[3 * (uint64_t)id[j] + 1];
            out_coords[9 * t + 3 * j + 2] = coords[3 * (uint64_t)id[j] + 2];
        }
        if (vertex_normals) {
            /* ( normal0 + normal1 + normal2 ) / 3.0f (PolygonExporter.cpp:140-150) */
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float a = normals[3 * (uint64_t)i0 + k], b = normals[3 * (uint64_t)i1 + k];
                const float c = normals[3 * (uint64_t)i2 + k];
                const float s = a + b;
                out_normals[3 * t + k] = (s + c) / 3.0f;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) out_normals[3 * t + k] = normals[3 * t + k];
        }
    }
}

/* A device copy of a host surface's arrays (nothing is copied for a device one). */
struct DeviceView {
    const float* coords = nullptr;
    const uint32_t* conn = nullptr;
    const float* normals = nullptr;
    std::vector<void*> owned;
    ~DeviceView() { for (void* p : owned) device_free(p); }
};

kmc_status device_view(const kmc_surface* s, hipStream_t st, DeviceView& v)
{
    if (s->memory == KMC_MEMORY_DEVICE) {
        v.coords = s->coords;
        v.conn = s->connections;
        v.normals = s->normals;
        return KMC_OK;
    }
    auto up = [&](const void* src, size_t bytes) -> const void* {
        if (!bytes) return nullptr;
        void* p = device_alloc(bytes);
        if (!p) return nullptr;
        v.owned.push_back(p);
        return hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, st) == hipSuccess ? p : nullptr;
    };
    const size_t bc = s->n_vertices * 12, bk = s->n_connections * 4, bn = s->n_normals * 12;
    v.coords = (const float*)up(s->coords, bc);
    v.conn = (const uint32_t*)up(s->connections, bk);
    v.normals = (const float*)up(s->normals, bn);
    if ((bc && !v.coords) || (bk && !v.conn) || (bn && !v.normals)) return KMC_ERR_OUT_OF_MEMORY;
    return KMC_OK;
}

kmc_status expand_impl(const kmc_surface* in, int32_t out_memory, hipStream_t st, kmc_surface* out)
{
    if (!in || !out) return KMC_ERR_NULL_INPUT;
    if (in->polygon_type != 3) return KMC_ERR_NOT_TRIANGLES;   // "Input object is not triangle polygon."
    if (out_memory != KMC_MEMORY_HOST && out_memory != KMC_MEMORY_DEVICE) return KMC_ERR_INVALID_ARGUMENT;
    kmc_status s = ensure_device();
    if (s != KMC_OK) return s;
    const bool shared = in->n_connections > 0;
    const uint64_t nt = shared ? in->n_connections / 3 : in->n_vertices / 3;
    const bool vn = in->normal_type == KMC_VERTEX_NORMAL;
    *out = *in;
    out->internal = nullptr;
    out->coords = nullptr;
    out->normals = nullptr;
    out->connections = nullptr;
    out->colors = nullptr;
    out->n_colors = 0;
    out->n_connections = 0;
    out->n_triangles = nt;
    out->n_vertices = 3 * nt;
    out->n_normals = (vn || in->normal_type == KMC_POLYGON_NORMAL) && in->n_normals ? nt : 0;
    out->normal_type = KMC_POLYGON_NORMAL;
    SurfaceStorage* store = new SurfaceStorage();
    store->memory = KMC_MEMORY_DEVICE;
    out->internal = store;
    out->memory = KMC_MEMORY_DEVICE;
    if (nt == 0) {
        out->n_vertices = out->n_normals = 0;
        out->memory = out_memory;
        return KMC_OK;
    }
    DeviceView v;
    s = device_view(in, st, v);
    if (s == KMC_OK) {
        store->coords = device_alloc(9 * nt * 4);
        store->normals = device_alloc(3 * nt * 4);
        if (!store->coords || !store->normals) s = KMC_ERR_OUT_OF_MEMORY;
    }
    if (s == KMC_OK) {
        out->coords = (float*)store->coords;
        out->normals = (float*)store->normals;
        if (shared) {
            const uint64_t b = (nt + kBlock - 1) / kBlock;
            hipLaunchKernelGGL(k_expand, dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 256 * 8))),
                               dim3(kBlock), 0, st, v.coords, v.conn, v.normals, nt, vn ? 1 : 0, out->coords,
                               out->normals);
            if (hipGetLastError() != hipSuccess) s = KMC_ERR_DEVICE;
        } else if (hipMemcpyAsync(out->coords, v.coords, 9 * nt * 4, hipMemcpyDeviceToDevice, st) != hipSuccess ||
                   (in->n_normals && hipMemcpyAsync(out->normals, v.normals, 3 * nt * 4, hipMemcpyDeviceToDevice,
                                                    st) != hipSuccess)) {
            s = KMC_ERR_DEVICE;
        }
    }
    if (s == KMC_OK && hipStreamSynchronize(st) != hipSuccess) s = KMC_ERR_DEVICE;
    if (s == KMC_OK && out_memory == KMC_MEMORY_HOST) s = to_host(out, st);
    if (s != KMC_OK) { (void)hipGetLastError(); kmc_surface_release(out); }
    return s;
}

/* Stl::write_binary (:490-540) / write_ascii (:447-481). */
kmc_status write_stl(const kmc_surface* soup, const char* filename, int ascii)
{
    FILE* f = std::fopen(filename, "w");
    if (!f) return KMC_ERR_FILE;
    const uint64_t nt = soup->n_normals;
    const float* n = soup->normals;
    const float* c = soup->coords;
    bool ok = true;
    if (ascii) {
        std::fprintf(f, "solid %s\n", "Generated by KVS");
        for (uint64_t i = 0; i < nt && ok; ++i) {
            std::fprintf(f, "facet normal %f %f %f\n", n[3 * i], n[3 * i + 1], n[3 * i + 2]);
            std::fprintf(f, "outer loop\n");
            for (int j = 0; j < 3; ++j)
                std::fprintf(f, "vertex %f %f %f\n", c[9 * i + 3 * j], c[9 * i + 3 * j + 1], c[9 * i + 3 * j + 2]);
            std::fprintf(f, "endloop\n");
            std::fprintf(f, "endfacet\n");
        }
        std::fprintf(f, "endsolid\n");
    } else {
        /* the same bytes, assembled in blocks of facets and written with one
         * call per block: 80-byte header, 32-bit count, per facet its normal,
         * three corners and two zero bytes (50 bytes) */
        char header[80];
        std::memset(header, 0, 80);
        std::snprintf(header, 80, "Generated by KVS");
        const uint32_t count = (uint32_t)nt;
        ok = std::fwrite(header, 1, 80, f) == 80 && std::fwrite(&count, 4, 1, f) == 1;
        constexpr uint64_t kFacets = 1u << 16;
        std::vector<char> buf(std::min<uint64_t>(nt, kFacets) * 50);
        for (uint64_t i0 = 0; i0 < nt && ok; i0 += kFacets) {
            const uint64_t i1 = std::min<uint64_t>(nt, i0 + kFacets);
            char* o = buf.data();
            for (uint64_t i = i0; i < i1; ++i, o += 50) {
                std::memcpy(o, n + 3 * i, 12);
                std::memcpy(o + 12, c + 9 * i, 36);
                o[48] = o[49] = 0;
            }
            ok = std::fwrite(buf.data(), 50, i1 - i0, f) == i1 - i0;
        }
    }
    ok = std::fclose(f) == 0 && ok;
    return ok ? KMC_OK : KMC_ERR_FILE;
}

}  // namespace

}  // namespace kmc

extern "C" {

kmc_status kmc_surface_expand(const kmc_surface* surface, int32_t out_memory, void* stream, kmc_surface* out)
{
    const kmc_status s = kmc::expand_impl(surface, out_memory, (hipStream_t)stream, out);
    kmc::report(s);
    return s;
}

kmc_status kmc_export_stl(const kmc_surface* surface, const char* filename, int32_t ascii, void* stream)
{
    if (!filename) { kmc::report(KMC_ERR_INVALID_ARGUMENT); return KMC_ERR_INVALID_ARGUMENT; }
    kmc_surface soup;
    kmc_status s = kmc::expand_impl(surface, KMC_MEMORY_HOST, (hipStream_t)stream, &soup);
    if (s == KMC_OK) {
        s = kmc::write_stl(&soup, filename, ascii);
        kmc_surface_release(&soup);
    }
    kmc::report(s);
    return s;
}

}  // extern "C"
