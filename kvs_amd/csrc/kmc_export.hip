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
 *
 * The KVSML polygon file (FileFormat/KVSML/KVSMLPolygonObject.cpp:349-495) is
 * written from the surface's arrays as the reference's exporter and writer
 * lay it out: the same elements in the same order, printed by the rules of
 * the XML library the reference bundles, numbers formatted as its streams
 * format them; the PLY file (FileFormat/PLY/Ply.cpp:389-489) as its PLY
 * writer lays it out.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "kmc.h"
#include "kmc_guard.h"
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
        return upload_host(p, src, bytes, st) == hipSuccess ? p : nullptr;
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
    if (s != KMC_OK) {
        /* nothing queued may still read or write the blocks that go back to
         * the pool here (the outputs, and the device view's copies) */
        (void)hipStreamSynchronize(st);
        (void)hipGetLastError();
        kmc_surface_release(out);
    }
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

/* ------------------------------------------------------------------------- */
/* KVSML polygon files                                                         */
/* ------------------------------------------------------------------------- */

/* An XML element as the reference builds it (tag, attributes in insertion
 * order, element children or one text child). */
struct XmlNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<XmlNode> kids;
    bool has_text = false;
    std::string text;
    XmlNode& kid(const std::string& n)
    {
        kids.push_back(XmlNode());
        kids.back().name = n;
        return kids.back();
    }
};

/* The bundled TinyXML's printing (FileFormat/XML/TinyXML.cpp:1044-1080,
 * 1222-1335): four spaces of indent a level; an element without children
 * closes itself ("<a x=\"1\" />"); element children start on new lines, a text
 * child follows the tag directly, and the closing tag comes on a line of its
 * own; an attribute value holding '"' is quoted with '\''. */
void xml_print(std::string& out, const XmlNode& n, int depth)
{
    for (int i = 0; i < depth; ++i) out += "    ";
    out += "<" + n.name;
    for (const auto& a : n.attrs) {
        const char q = a.second.find('"') != std::string::npos ? '\'' : '"';
        out += " " + a.first + "=" + q + a.second + q;
    }
    if (n.kids.empty() && !n.has_text) {
        out += " />";
        return;
    }
    out += ">";
    if (n.has_text) out += n.text;
    for (const XmlNode& k : n.kids) {
        out += "\n";
        xml_print(out, k, depth + 1);
    }
    out += "\n";
    for (int i = 0; i < depth; ++i) out += "    ";
    out += "</" + n.name + ">";
}

/* Values as the reference's streams print them, each followed by `delim`:
 * floats as `ostream << float` (%g, six significant digits), bytes as
 * integers, uint32 as unsigned.  Large arrays are formatted by several
 * threads, in order. */
template <typename T>
void format_part(std::string& out, const T* v, size_t n, const char* delim)
{
    /* std::to_chars in the general format at precision 6 gives printf's
     * "%.6g" (the stream's conversion of the float's double value) */
    const size_t dl = std::strlen(delim);
    char buf[64];
    out.reserve(out.size() + n * (std::is_same<T, float>::value ? 12 : 8));
    for (size_t i = 0; i < n; ++i) {
        std::to_chars_result r;
        if constexpr (std::is_same<T, float>::value)
            r = std::to_chars(buf, buf + sizeof(buf), (double)v[i], std::chars_format::general, 6);
        else
            r = std::to_chars(buf, buf + sizeof(buf), (unsigned)v[i]);
        std::memcpy(r.ptr, delim, dl);
        out.append(buf, (size_t)(r.ptr - buf) + dl);
    }
}

template <typename T>
std::string format_values(const T* v, size_t n, const char* delim)
{
    constexpr size_t kPerThread = size_t(1) << 16;
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t parts = std::max<size_t>(1, std::min<size_t>({ (n + kPerThread - 1) / kPerThread, hw ? (size_t)hw : 1,
                                                               16 }));
    if (parts == 1) {
        std::string out;
        format_part(out, v, n, delim);
        return out;
    }
    std::vector<std::string> pieces(parts);
    const size_t per = (n + parts - 1) / parts;
    std::vector<std::thread> pool;
    for (size_t p = 0; p < parts; ++p)
        pool.emplace_back([&, p] {
            const size_t b = p * per, e = std::min(n, b + per);
            if (b < e) format_part(pieces[p], v + b, e - b, delim);
        });
    for (std::thread& t : pool) t.join();
    std::string out;
    size_t total = 0;
    for (const std::string& x : pieces) total += x.size();
    out.reserve(total);
    for (const std::string& x : pieces) out += x;
    return out;
}

/* kvs::File: the directory ("." without one) and the name without its last
 * extension (Utility/File.cpp:142-157, 244-268). */
void split_path(const std::string& path, std::string& dir, std::string& base)
{
    const size_t sep = path.find_last_of('/');
    std::string file;
    if (sep == std::string::npos) { dir = "."; file = path; }
    else if (sep == 0) { dir = "/"; file = path.substr(1); }
    else { dir = path.substr(0, sep); file = path.substr(sep + 1); }
    const size_t dot = file.find_last_of('.');
    base = dot == std::string::npos ? file : file.substr(0, dot);
}

bool write_file(const std::string& path, const void* data, size_t bytes)
{
    FILE* f = std::fopen(path.c_str(), bytes ? "wb" : "w");
    if (!f) return false;
    const bool ok = !bytes || std::fwrite(data, 1, bytes, f) == bytes;
    return std::fclose(f) == 0 && ok;
}

struct KvsmlWriter {
    int type;            // kmc_kvsml_writing
    std::string dir, base;
    bool ok = true;

    /* DataArrayTag::write (DataArrayTag.cpp:74-172) with the file names of
     * DataArray::GetDataFilename (DataArray.h:94-97) */
    template <typename T>
    void data_array(XmlNode& parent, const char* kvs_type, const T* v, size_t n, const char* what)
    {
        if (n == 0) return;
        XmlNode& e = parent.kid("DataArray");
        e.attrs.push_back({ "type", kvs_type });
        if (type == KMC_KVSML_ASCII) {
            e.has_text = true;
            e.text = format_values(v, n, " ");
            return;
        }
        const std::string file = base + "_" + what + ".dat";
        e.attrs.push_back({ "format", type == KMC_KVSML_EXTERNAL_ASCII ? "ascii" : "binary" });
        e.attrs.push_back({ "file", file });
        const std::string path = dir + "/" + file;
        if (type == KMC_KVSML_EXTERNAL_ASCII) {
            const std::string text = format_values(v, n, ", ");
            ok = write_file(path, text.data(), text.size()) && ok;
        } else {
            e.attrs.push_back({ "endian", "little" });
            ok = write_file(path, v, n * sizeof(T)) && ok;
        }
    }

    /* DataValueTag::write (DataValueTag.h:80-105) */
    template <typename T>
    void data_value(XmlNode& parent, const T* v, size_t n)
    {
        XmlNode& e = parent.kid("DataValue");
        e.has_text = true;
        e.text = format_values(v, n, " ");
    }

    /* WriteCoordData / WriteColorData / WriteOpacityData / WriteNormalData /
     * WriteConnectionData (KVSML/DataWriter.cpp): one value array becomes a
     * <DataValue> when it holds a single colour (3 bytes), opacity (1) or
     * normal (3). */
    void coords(XmlNode& p, const float* v, size_t n)
    {
        if (n) data_array(p.kid("Coord"), "float", v, n, "coord");
    }
    void colors(XmlNode& p, const uint8_t* v, size_t n)
    {
        if (!n) return;
        XmlNode& t = p.kid("Color");
        if (n == 3) data_value(t, v, n);
        else data_array(t, "uchar", v, n, "color");
    }
    void opacities(XmlNode& p, const uint8_t* v, size_t n)
    {
        if (!n) return;
        XmlNode& t = p.kid("Opacity");
        if (n == 1) data_value(t, v, n);
        else data_array(t, "uchar", v, n, "opacity");
    }
    void normals(XmlNode& p, const float* v, size_t n)
    {
        if (!n) return;
        XmlNode& t = p.kid("Normal");
        if (n == 3) data_value(t, v, n);
        else data_array(t, "float", v, n, "normal");
    }
    void connections(XmlNode& p, const uint32_t* v, size_t n)
    {
        if (n) data_array(p.kid("Connection"), "uint", v, n, "connect");
    }
};

/* Host copies of a surface's arrays (a host surface's arrays are used as they are). */
struct HostArrays {
    const float* coords = nullptr;
    const uint32_t* conn = nullptr;
    const float* normals = nullptr;
    const uint8_t* colors = nullptr;
    std::vector<float> c, n;
    std::vector<uint32_t> k;
    std::vector<uint8_t> cl;
};

kmc_status host_arrays(const kmc_surface* s, hipStream_t st, HostArrays& h)
{
    const size_t nc = 3 * s->n_vertices, nk = s->n_connections, nn = 3 * s->n_normals;
    const size_t ncl = s->colors ? 3 * s->n_colors : 0;
    if (s->memory == KMC_MEMORY_HOST) {
        h.coords = s->coords;
        h.conn = s->connections;
        h.normals = s->normals;
        h.colors = ncl ? s->colors : nullptr;
        return KMC_OK;
    }
    h.c.resize(nc);
    h.k.resize(nk);
    h.n.resize(nn);
    h.cl.resize(ncl);
    auto down = [&](void* dst, const void* src, size_t bytes) {
        return !bytes || hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st) == hipSuccess;
    };
    if (!down(h.c.data(), s->coords, nc * 4) || !down(h.k.data(), s->connections, nk * 4) ||
        !down(h.n.data(), s->normals, nn * 4) || !down(h.cl.data(), s->colors, ncl) ||
        hipStreamSynchronize(st) != hipSuccess) {
        (void)hipGetLastError();
        return KMC_ERR_DEVICE;
    }
    h.coords = h.c.data();
    h.conn = h.k.data();
    h.normals = h.n.data();
    h.colors = ncl ? h.cl.data() : nullptr;
    return KMC_OK;
}

/* PolygonExporter<KVSMLPolygonObject>::exec (Exporter/PolygonExporter.cpp:22-70)
 * then KVSMLPolygonObject::write (KVSMLPolygonObject.cpp:349-495). */
kmc_status export_kvsml(const kmc_surface* s, const uint8_t* color, int32_t opacity, const char* filename,
                        int32_t writing, hipStream_t st)
{
    if (!s || !filename) return KMC_ERR_INVALID_ARGUMENT;
    if (writing < KMC_KVSML_ASCII || writing > KMC_KVSML_EXTERNAL_BINARY || opacity > 255) return KMC_ERR_INVALID_ARGUMENT;
    if ((s->polygon_type != 3 && s->polygon_type != 4) || (s->color_type != 0 && s->color_type != 1) ||
        (s->normal_type != KMC_VERTEX_NORMAL && s->normal_type != KMC_POLYGON_NORMAL))
        return KMC_ERR_INVALID_ARGUMENT;
    HostArrays h;
    kmc_status status = host_arrays(s, st, h);
    if (status != KMC_OK) return status;

    const size_t ncoords = 3 * s->n_vertices, nconn = s->n_connections, nnormals = 3 * s->n_normals;
    const size_t ncolors = h.colors ? 3 * s->n_colors : (color ? 3 : 0);
    const uint8_t* colors = h.colors ? h.colors : color;
    const uint8_t op = (uint8_t)(opacity < 0 ? 0 : opacity);
    const size_t nopac = opacity < 0 ? 0 : 1;
    const bool vcolor = s->color_type == 0, vnormal = s->normal_type == KMC_VERTEX_NORMAL;

    KvsmlWriter w;
    w.type = writing;
    split_path(filename, w.dir, w.base);
    XmlNode root;
    root.name = "KVSML";
    XmlNode& object = root.kid("Object");
    object.attrs.push_back({ "type", "PolygonObject" });
    XmlNode& po = object.kid("PolygonObject");
    po.attrs.push_back({ "polygon_type", s->polygon_type == 3 ? "triangle" : "quadrangle" });
    po.attrs.push_back({ "color_type", vcolor ? "vertex" : "polygon" });
    po.attrs.push_back({ "normal_type", vnormal ? "vertex" : "polygon" });
    XmlNode& vertex = po.kid("Vertex");
    vertex.attrs.push_back({ "nvertices", std::to_string(s->n_vertices) });
    w.coords(vertex, h.coords, ncoords);
    if (vcolor) {
        w.colors(vertex, colors, ncolors);
        w.opacities(vertex, &op, nopac);
    }
    if (vnormal) w.normals(vertex, h.normals, nnormals);
    if (!vcolor || !vnormal || nopac || nconn) {
        const size_t per = s->polygon_type == 3 ? 3 : 4;
        XmlNode& polygon = po.kid("Polygon");
        polygon.attrs.push_back({ "npolygons", std::to_string(nconn ? nconn / per : s->n_vertices / per) });
        w.connections(polygon, h.conn, nconn);
        if (!vcolor) {
            w.colors(polygon, colors, ncolors);
            w.opacities(polygon, &op, nopac);
        }
        if (!vnormal) w.normals(polygon, h.normals, nnormals);
    }
    if (!w.ok) return KMC_ERR_FILE;

    std::string out = "<?xml version=\"1.0\" ?>\n<!-- Generated by kvs::KVSMLPolygonObject::write() -->\n";
    xml_print(out, root, 0);
    out += "\n";
    return write_file(filename, out.data(), out.size()) ? KMC_OK : KMC_ERR_FILE;
}

/* ------------------------------------------------------------------------- */
/* PLY files                                                                   */
/* ------------------------------------------------------------------------- */

/* The vertex arrays PolygonExporter<kvs::Ply>::exec hands the writer
 * (Exporter/PolygonExporter.cpp:178-363): a single colour repeated at every
 * vertex, per-vertex colours as they are; VertexNormal normals as they are,
 * PolygonNormal ones summed into each vertex of each triangle in triangle
 * order and divided by the vertex's triangle count (zero without
 * connections). */
struct PlyVertices {
    bool has_colors = false, has_normals = false;
    std::vector<uint8_t> colors;
    std::vector<float> normals;
};

bool ply_vertices(const kmc_surface* s, const HostArrays& h, const uint8_t* color, PlyVertices& pv)
{
    const uint64_t nv = s->n_vertices;
    if (h.colors && s->n_colors && s->n_colors != nv) return false;   // per-vertex colours, one a vertex
    if (h.colors && s->n_colors) {
        pv.has_colors = true;
        pv.colors.assign(h.colors, h.colors + 3 * s->n_colors);
    } else if (color) {
        pv.has_colors = true;
        pv.colors.resize(3 * nv);
        for (uint64_t i = 0; i < nv; ++i) std::memcpy(&pv.colors[3 * i], color, 3);
    }
    if (s->n_normals == 0) return true;
    pv.has_normals = true;
    if (s->normal_type == KMC_VERTEX_NORMAL) {
        if (s->n_normals != nv) return false;
        pv.normals.assign(h.normals, h.normals + 3 * s->n_normals);
        return true;
    }
    pv.normals.assign(3 * nv, 0.0f);
    std::vector<uint32_t> counter(nv, 0);
    const uint64_t nt = s->n_connections / 3;
    if (nt > s->n_normals) return false;   // PolygonNormal: one normal a triangle
    for (uint64_t t = 0; t < nt; ++t) {
        for (int j = 0; j < 3; ++j) {
            const uint32_t v = h.conn[3 * t + j];
            if (v >= nv) return false;
            for (int k = 0; k < 3; ++k) pv.normals[3 * (uint64_t)v + k] += h.normals[3 * t + k];
            counter[v] += 1;
        }
    }
    for (uint64_t i = 0; i < nv; ++i) {
        if (counter[i] == 0) continue;
        for (int k = 0; k < 3; ++k) pv.normals[3 * i + k] /= (float)counter[i];
    }
    return true;
}

/* Ply::write (FileFormat/PLY/Ply.cpp:389-489) through PlyFile's writer
 * (PlyFile.cpp:167-200 the name, 399-465 the header, 496-600 the records,
 * 1897-2000 the items): "%g " / "%u " / "%d " items and a newline per record
 * in ascii, packed native little-endian items in binary. */
kmc_status export_ply(const kmc_surface* s, const uint8_t* color, const char* filename, int32_t binary, hipStream_t st)
{
    if (!s || !filename) return KMC_ERR_INVALID_ARGUMENT;
    if (s->polygon_type != 3) return KMC_ERR_NOT_TRIANGLES;   // "Input object is not triangle polygon."
    HostArrays h;
    kmc_status status = host_arrays(s, st, h);
    if (status != KMC_OK) return status;
    PlyVertices pv;
    if (!ply_vertices(s, h, color, pv)) return KMC_ERR_INVALID_ARGUMENT;
    const uint64_t nv = s->n_vertices, nf = s->n_connections / 3;

    std::string name = filename;
    if (name.size() < 4 || name.compare(name.size() - 4, 4, ".ply") != 0) name += ".ply";
    std::string head = std::string("ply\nformat ") + (binary ? "binary_little_endian" : "ascii") + " 1.0\n";
    head += "element vertex " + std::to_string((int)nv) + "\n";
    head += "property float x\nproperty float y\nproperty float z\n";
    if (pv.has_colors) head += "property uchar red\nproperty uchar green\nproperty uchar blue\n";
    if (pv.has_normals) head += "property float nx\nproperty float ny\nproperty float nz\n";
    head += "element face " + std::to_string((int)nf) + "\n";
    if (nf) head += "property list uchar int vertex_indices\n";
    head += "end_header\n";

    FILE* f = std::fopen(name.c_str(), "wb");
    if (!f) return KMC_ERR_FILE;
    bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
    const size_t rec = 12 + (pv.has_colors ? 3 : 0) + (pv.has_normals ? 12 : 0);
    constexpr uint64_t kBlock = 1u << 16;
    std::string buf;
    for (uint64_t i0 = 0; i0 < nv && ok; i0 += kBlock) {
        const uint64_t i1 = std::min<uint64_t>(nv, i0 + kBlock);
        buf.clear();
        if (binary) {
            buf.resize((i1 - i0) * rec);
            char* o = &buf[0];
            for (uint64_t i = i0; i < i1; ++i) {
                std::memcpy(o, h.coords + 3 * i, 12);
                o += 12;
                if (pv.has_colors) { std::memcpy(o, &pv.colors[3 * i], 3); o += 3; }
                if (pv.has_normals) { std::memcpy(o, &pv.normals[3 * i], 12); o += 12; }
            }
        } else {
            for (uint64_t i = i0; i < i1; ++i) {
                format_part(buf, h.coords + 3 * i, 3, " ");
                if (pv.has_colors) format_part(buf, &pv.colors[3 * i], 3, " ");
                if (pv.has_normals) format_part(buf, &pv.normals[3 * i], 3, " ");
                buf += "\n";
            }
        }
        ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    }
    for (uint64_t t0 = 0; t0 < nf && ok; t0 += kBlock) {
        const uint64_t t1 = std::min<uint64_t>(nf, t0 + kBlock);
        buf.clear();
        for (uint64_t t = t0; t < t1; ++t) {
            if (binary) {
                const char three = 3;
                const int32_t idx[3] = { (int32_t)h.conn[3 * t], (int32_t)h.conn[3 * t + 1], (int32_t)h.conn[3 * t + 2] };
                buf.append(&three, 1);
                buf.append((const char*)idx, 12);
            } else {
                char line[48];
                const int n = std::snprintf(line, sizeof(line), "3 %d %d %d \n", (int32_t)h.conn[3 * t],
                                            (int32_t)h.conn[3 * t + 1], (int32_t)h.conn[3 * t + 2]);
                buf.append(line, (size_t)n);
            }
        }
        ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    }
    ok = std::fclose(f) == 0 && ok;
    return ok ? KMC_OK : KMC_ERR_FILE;
}

}  // namespace

}  // namespace kmc

extern "C" {

kmc_status kmc_surface_expand(const kmc_surface* surface, int32_t out_memory, void* stream, kmc_surface* out)
{
    return kmc::abi_guard([&]() -> kmc_status {
        const kmc_status s = kmc::expand_impl(surface, out_memory, (hipStream_t)stream, out);
        kmc::report(s);
        return s;
    });
}

kmc_status kmc_export_kvsml(const kmc_surface* surface, const uint8_t* color, int32_t opacity, const char* filename,
                            int32_t writing, void* stream)
{
    return kmc::abi_guard([&]() -> kmc_status {
        const kmc_status s = kmc::export_kvsml(surface, color, opacity, filename, writing, (hipStream_t)stream);
        kmc::report(s);
        return s;
    });
}

kmc_status kmc_export_ply(const kmc_surface* surface, const uint8_t* color, const char* filename, int32_t binary,
                          void* stream)
{
    return kmc::abi_guard([&]() -> kmc_status {
        const kmc_status s = kmc::export_ply(surface, color, filename, binary, (hipStream_t)stream);
        kmc::report(s);
        return s;
    });
}

kmc_status kmc_export_stl(const kmc_surface* surface, const char* filename, int32_t ascii, void* stream)
{
    return kmc::abi_guard([&]() -> kmc_status {
        if (!filename) { kmc::report(KMC_ERR_INVALID_ARGUMENT); return KMC_ERR_INVALID_ARGUMENT; }
        kmc_surface soup;
        kmc_status s = kmc::expand_impl(surface, KMC_MEMORY_HOST, (hipStream_t)stream, &soup);
        if (s == KMC_OK) {
            s = kmc::write_stl(&soup, filename, ascii);
            kmc_surface_release(&soup);
        }
        kmc::report(s);
        return s;
    });
}

}  // extern "C"
