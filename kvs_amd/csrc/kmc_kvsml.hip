/*
 * kmc_kvsml.hip -- volume ingestion straight to HBM, from KVSML and AVS field files:
 * kvs::StructuredVolumeObject::read (StructuredVolumeObject.cpp:171-229) over
 * KVSMLStructuredVolumeObject::read (KVSMLStructuredVolumeObject.cpp:159-300)
 * and DataArrayTag::read_data (DataArrayTag.cpp:217-420).
 *
 * External binary arrays -- the form large volumes take -- stream from the
 * file through two pinned staging buffers into device memory: while one
 * buffer's chunk is copied to the device, the next chunk is read into the
 * other, so disk reads and PCIe copies overlap and no host copy of the whole
 * volume exists.  Internal and ascii arrays are parsed on the host as the
 * reference parses them (atof / atoi and a static_cast).  Missing min/max
 * values are computed on the device with GetMinMaxValues<T>'s rules; missing
 * object coordinates follow updateMinMaxCoords (StructuredVolumeObject.cpp:348-424).
 */
#include <hip/hip_runtime.h>
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "kmc.h"
#include "kmc_guard.h"
#include "kmc_internal.h"

namespace kmc {

namespace {

/* ------------------------------------------------------------------------- */
/* a small XML reader: elements, attributes, text                              */
/* ------------------------------------------------------------------------- */

struct XmlNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::string text;
    std::vector<XmlNode> children;

    const XmlNode* child(const char* n) const
    {
        for (const XmlNode& c : children)
            if (c.name == n) return &c;
        return nullptr;
    }
    /* "" when the attribute is absent (kvs::XMLElement::AttributeValue) */
    std::string attr(const char* n) const
    {
        for (const auto& a : attrs)
            if (a.first == n) return a.second;
        return std::string();
    }
};

std::string decode_entities(const std::string& s)
{
    std::string out;
    out.reserve(s.size());
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] != '&') { out += s[i]; continue; }
        static const struct { const char* e; char c; } ents[] = {
            { "&amp;", '&' }, { "&lt;", '<' }, { "&gt;", '>' }, { "&quot;", '"' }, { "&apos;", '\'' } };
        bool done = false;
        for (const auto& en : ents) {
            const size_t n = std::strlen(en.e);
            if (s.compare(i, n, en.e) == 0) { out += en.c; i += n - 1; done = true; break; }
        }
        if (!done) out += '&';
    }
    return out;
}

struct XmlParser {
    const std::string& s;
    size_t p = 0;
    bool ok = true;

    explicit XmlParser(const std::string& src) : s(src) {}

    void skip_ws() { while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) ++p; }

    /* skip <?...?>, <!--...-->, <!...> */
    bool skip_misc()
    {
        skip_ws();
        if (s.compare(p, 2, "<?") == 0) { p = s.find("?>", p); if (p == std::string::npos) { ok = false; return false; } p += 2; return true; }
        if (s.compare(p, 4, "<!--") == 0) { p = s.find("-->", p); if (p == std::string::npos) { ok = false; return false; } p += 3; return true; }
        if (s.compare(p, 2, "<!") == 0) { p = s.find('>', p); if (p == std::string::npos) { ok = false; return false; } p += 1; return true; }
        return false;
    }

    std::string name()
    {
        const size_t b = p;
        while (p < s.size() && !strchr(" \t\r\n/>=", s[p])) ++p;
        return s.substr(b, p - b);
    }

    bool element(XmlNode& n)
    {
        while (skip_misc()) {}
        if (!ok || p >= s.size() || s[p] != '<') return false;
        ++p;
        n.name = name();
        for (;;) {
            skip_ws();
            if (p >= s.size()) { ok = false; return false; }
            if (s[p] == '/') {
                if (s.compare(p, 2, "/>") != 0) { ok = false; return false; }
                p += 2;
                return true;
            }
            if (s[p] == '>') { ++p; break; }
            const std::string an = name();
            skip_ws();
            if (p >= s.size() || s[p] != '=') { ok = false; return false; }
            ++p;
            skip_ws();
            if (p >= s.size() || (s[p] != '"' && s[p] != '\'')) { ok = false; return false; }
            const char q = s[p++];
            const size_t e = s.find(q, p);
            if (e == std::string::npos) { ok = false; return false; }
            n.attrs.emplace_back(an, decode_entities(s.substr(p, e - p)));
            p = e + 1;
        }
        /* content: text and child elements until </name> */
        for (;;) {
            const size_t lt = s.find('<', p);
            if (lt == std::string::npos) { ok = false; return false; }
            n.text += s.substr(p, lt - p);
            p = lt;
            if (s.compare(p, 2, "</") == 0) {
                const size_t gt = s.find('>', p);
                if (gt == std::string::npos) { ok = false; return false; }
                p = gt + 1;
                n.text = decode_entities(n.text);
                return true;
            }
            if (skip_misc()) continue;
            XmlNode c;
            if (!element(c)) { ok = false; return false; }
            n.children.push_back(std::move(c));
        }
    }
};

bool read_file(const std::string& path, std::string& out)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(n > 0 ? (size_t)n : 0);
    const bool ok = n <= 0 || std::fread(&out[0], 1, (size_t)n, f) == (size_t)n;
    std::fclose(f);
    return ok;
}

/* ------------------------------------------------------------------------- */
/* data arrays                                                                 */
/* ------------------------------------------------------------------------- */

/* DataArrayTag type names (DataArrayTag.cpp:225-300) -> kmc_type; -1 unknown */
int type_of(const std::string& t)
{
    if (t == "char") return KMC_INT8;
    if (t == "unsigned char" || t == "uchar") return KMC_UINT8;
    if (t == "short") return KMC_INT16;
    if (t == "unsigned short" || t == "ushort") return KMC_UINT16;
    if (t == "int") return KMC_INT32;
    if (t == "unsigned int" || t == "uint") return KMC_UINT32;
    if (t == "float") return KMC_REAL32;
    if (t == "double") return KMC_REAL64;
    return -1;
}

/* One text token converted as the reference converts it: kvsml::temporal::To<T>
 * (DataArray.h:27-43: atoi for char / uchar, else atof) for internal data,
 * atof for every type in external ascii files (DataArray.h:236-241); then a
 * static_cast to the value type. */
template <typename T>
void store_token(const char* tok, bool use_atoi, void* dst, size_t i)
{
    ((T*)dst)[i] = use_atoi ? (T)std::atoi(tok) : (T)std::atof(tok);
}

void store(int type, const char* tok, bool internal, void* dst, size_t i)
{
    const bool use_atoi = internal && (type == KMC_INT8 || type == KMC_UINT8);
    switch (type) {
    case KMC_INT8: store_token<int8_t>(tok, use_atoi, dst, i); break;
    case KMC_UINT8: store_token<uint8_t>(tok, use_atoi, dst, i); break;
    case KMC_INT16: store_token<int16_t>(tok, use_atoi, dst, i); break;
    case KMC_UINT16: store_token<uint16_t>(tok, use_atoi, dst, i); break;
    case KMC_INT32: store_token<int32_t>(tok, use_atoi, dst, i); break;
    case KMC_UINT32: store_token<uint32_t>(tok, use_atoi, dst, i); break;
    case KMC_REAL32: store_token<float>(tok, use_atoi, dst, i); break;
    default: store_token<double>(tok, use_atoi, dst, i); break;
    }
}

/* Tokens of text split at any of delim, as strtok / kvs::Tokenizer split them;
 * the first n tokens are stored (missing ones leave zeros). */
void parse_tokens(std::string text, const char* delim, int type, bool internal, size_t n, void* dst)
{
    char* save = nullptr;
    char* tok = strtok_r(&text[0], delim, &save);
    for (size_t i = 0; i < n && tok; ++i) {
        store(type, tok, internal, dst, i);
        tok = strtok_r(nullptr, delim, &save);
    }
}

void swap_bytes(uint8_t* p, size_t n, size_t esz)
{
    if (esz == 1) return;
    for (size_t i = 0; i < n; ++i)
        for (size_t a = 0, b = esz - 1; a < b; ++a, --b) std::swap(p[i * esz + a], p[i * esz + b]);
}

/* Where a data array goes: a host buffer, or device memory through pinned staging. */
struct Sink {
    int memory;
    void* dst;
    hipStream_t st;
};

constexpr size_t kStage = size_t(64) << 20;   // bytes per staging buffer
constexpr size_t kReaders = 8;                  // threads reading one range
constexpr size_t kMinPart = size_t(4) << 20;    // no thread reads less than this

/* Bytes [offset, offset + n) of file fd into dst, by up to kReaders threads
 * each reading (and, when asked, byte-swapping) a contiguous part of whole
 * elements: one thread copying out of the page cache reads about 20 GB/s,
 * below what the host-to-device copy takes. */
bool read_range(int fd, uint8_t* dst, size_t n, off_t offset, size_t esz, bool swap)
{
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t threads = std::max<size_t>(1, std::min<size_t>({ kReaders, hw ? (size_t)hw : 1, n / kMinPart + 1 }));
    const size_t part = ((n + threads - 1) / threads + esz - 1) / esz * esz;
    std::atomic<bool> ok{ true };
    auto work = [&](size_t t) {
        const size_t b = t * part;
        if (b >= n) return;
        const size_t e = std::min(n, b + part);
        for (size_t got = b; got < e;) {
            const ssize_t r = ::pread(fd, dst + got, e - got, offset + (off_t)got);
            if (r < 0 && errno == EINTR) continue;
            if (r <= 0) {
                ok = false;
                return;
            }
            got += (size_t)r;
        }
        if (swap) swap_bytes(dst + b, (e - b) / esz, esz);
    };
    std::vector<std::thread> pool;
    for (size_t t = 1; t < threads; ++t) pool.emplace_back(work, t);
    work(0);
    for (std::thread& th : pool) th.join();
    return ok;
}

/* External binary file -> sink.  To the device: chunks read into two pinned
 * buffers in turn, each copied asynchronously; a buffer is refilled only after
 * its previous copy has completed (its event). */
kmc_status read_binary(const std::string& path, size_t bytes, size_t esz, bool swap, const Sink& sink,
                       long offset = 0)
{
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) return KMC_ERR_FILE;
    kmc_status s = KMC_OK;
    if (sink.memory == KMC_MEMORY_HOST) {
        if (!read_range(fd, (uint8_t*)sink.dst, bytes, offset, esz, swap)) s = KMC_ERR_FILE;
        ::close(fd);
        return s;
    }
    uint8_t* stage[2] = { (uint8_t*)host_pinned_alloc(kStage), (uint8_t*)host_pinned_alloc(kStage) };
    hipEvent_t done[2] = { nullptr, nullptr };
    bool used[2] = { false, false };
    if (!stage[0] || !stage[1] || hipEventCreateWithFlags(&done[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&done[1], hipEventDisableTiming) != hipSuccess)
        s = KMC_ERR_OUT_OF_MEMORY;
    const size_t chunk = kStage / esz * esz;
    for (size_t off = 0, k = 0; s == KMC_OK && off < bytes; off += chunk, k ^= 1) {
        const size_t n = bytes - off < chunk ? bytes - off : chunk;
        if (used[k] && hipEventSynchronize(done[k]) != hipSuccess) { s = KMC_ERR_DEVICE; break; }
        if (!read_range(fd, stage[k], n, offset + (off_t)off, esz, swap)) { s = KMC_ERR_FILE; break; }
        if (hipMemcpyAsync((uint8_t*)sink.dst + off, stage[k], n, hipMemcpyHostToDevice, sink.st) != hipSuccess ||
            hipEventRecord(done[k], sink.st) != hipSuccess) {
            s = KMC_ERR_DEVICE;
            break;
        }
        used[k] = true;
    }
    if (hipStreamSynchronize(sink.st) != hipSuccess && s == KMC_OK) s = KMC_ERR_DEVICE;
    (void)hipGetLastError();
    for (int k = 0; k < 2; ++k) {
        if (done[k]) (void)hipEventDestroy(done[k]);
        host_pinned_free(stage[k]);
    }
    ::close(fd);
    return s;
}

/* DataArrayTag::read_data: internal text, external ascii or external binary. */
kmc_status read_array(const XmlNode& da, const std::string& dir, size_t n, int* type_out, const Sink& sink,
                      std::vector<uint8_t>* host_copy)
{
    const int type = type_of(da.attr("type"));
    if (type < 0) return KMC_ERR_FILE;
    *type_out = type;
    const size_t esz = type_size(type), bytes = n * esz;
    const std::string file = da.attr("file");
    std::vector<uint8_t> buf;
    if (file.empty() || da.attr("format") == "ascii") {
        buf.assign(bytes, 0);
        if (file.empty()) {
            parse_tokens(da.text, " ,\t\n", type, true, n, buf.data());
        } else {
            std::string text;
            if (!read_file(dir + "/" + file, text)) return KMC_ERR_FILE;
            parse_tokens(text, " ,\t\n", type, false, n, buf.data());
        }
    } else if (da.attr("format") == "binary") {
        const std::string endian = da.attr("endian");
        const bool swap = endian == "big";   // this library runs on little-endian hosts
        if (host_copy) {
            host_copy->assign(bytes, 0);
            Sink hs{ KMC_MEMORY_HOST, host_copy->data(), sink.st };
            const kmc_status s = read_binary(dir + "/" + file, bytes, esz, swap, hs);
            if (s != KMC_OK || !sink.dst) return s;
            if (sink.memory == KMC_MEMORY_HOST) { std::memcpy(sink.dst, host_copy->data(), bytes); return KMC_OK; }
            return hipMemcpyAsync(sink.dst, host_copy->data(), bytes, hipMemcpyHostToDevice, sink.st) == hipSuccess &&
                           hipStreamSynchronize(sink.st) == hipSuccess
                       ? KMC_OK
                       : KMC_ERR_DEVICE;
        }
        return read_binary(dir + "/" + file, bytes, esz, swap, sink);
    } else {
        return KMC_ERR_FILE;
    }
    if (host_copy) *host_copy = buf;
    if (!sink.dst) return KMC_OK;
    if (sink.memory == KMC_MEMORY_HOST) { std::memcpy(sink.dst, buf.data(), bytes); return KMC_OK; }
    return hipMemcpyAsync(sink.dst, buf.data(), bytes, hipMemcpyHostToDevice, sink.st) == hipSuccess &&
                   hipStreamSynchronize(sink.st) == hipSuccess
               ? KMC_OK
               : KMC_ERR_DEVICE;
}

bool parse_floats(const std::string& s, float* out, int n)
{
    std::string t = s;
    char* save = nullptr;
    char* tok = strtok_r(&t[0], " \n", &save);
    for (int i = 0; i < n; ++i) {
        if (!tok) return false;
        out[i] = (float)std::atof(tok);
        tok = strtok_r(nullptr, " \n", &save);
    }
    return true;
}

kmc_status read_volume(const char* filename, int32_t memory, hipStream_t st, kmc_volume* out)
{
    if (!filename || !out) return KMC_ERR_INVALID_ARGUMENT;
    if (memory != KMC_MEMORY_HOST && memory != KMC_MEMORY_DEVICE) return KMC_ERR_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    /* KVSMLStructuredVolumeObject::CheckExtension (:34-46) */
    const std::string fn(filename);
    const size_t dot = fn.find_last_of('.');
    const std::string ext = dot == std::string::npos ? "" : fn.substr(dot + 1);
    if (ext != "kvsml" && ext != "KVSML" && ext != "xml" && ext != "XML") return KMC_ERR_FILE;
    std::string doc;
    if (!read_file(fn, doc)) return KMC_ERR_FILE;
    const size_t slash = fn.find_last_of('/');
    const std::string dir = slash == std::string::npos ? "." : fn.substr(0, slash);

    XmlParser px(doc);
    XmlNode root;
    if (!px.element(root) || root.name != "KVSML") return KMC_ERR_FILE;
    const XmlNode* obj = root.child("Object");
    const XmlNode* vt = obj ? obj->child("StructuredVolumeObject") : nullptr;
    const XmlNode* node = vt ? vt->child("Node") : nullptr;
    const XmlNode* value = node ? node->child("Value") : nullptr;
    const XmlNode* da = value ? value->child("DataArray") : nullptr;
    if (!da || vt->attr("resolution").empty() || vt->attr("grid_type").empty() || value->attr("veclen").empty())
        return KMC_ERR_FILE;
    /* resolution (StructuredVolumeObjectTag.cpp:138-155), veclen (ValueTag.cpp:58-59) */
    uint32_t res[3];
    {
        std::string t = vt->attr("resolution");
        char* save = nullptr;
        char* tok = strtok_r(&t[0], " \n", &save);
        for (int k = 0; k < 3; ++k) {
            if (!tok) return KMC_ERR_FILE;
            res[k] = (uint32_t)std::atoi(tok);
            tok = strtok_r(nullptr, " \n", &save);
        }
    }
    const uint32_t veclen = (uint32_t)std::atoi(value->attr("veclen").c_str());
    const std::string grid = vt->attr("grid_type");
    const size_t nnodes = (size_t)res[0] * res[1] * res[2], n = nnodes * veclen;

    kmc_status s = ensure_device();   // device memory, or the device min / max scan
    if (s != KMC_OK) return s;
    const std::string tname = da->attr("type");
    const int type = type_of(tname);
    if (type < 0) return KMC_ERR_FILE;
    const size_t bytes = n * type_size(type);
    void* dst = nullptr;
    if (memory == KMC_MEMORY_DEVICE) dst = device_alloc(bytes ? bytes : 1);
    else dst = host_block_alloc(bytes ? bytes : 1);
    if (!dst) return KMC_ERR_OUT_OF_MEMORY;
    auto fail = [&](kmc_status e) {
        if (memory == KMC_MEMORY_DEVICE) device_free(dst);
        else host_block_free(dst);
        std::memset(out, 0, sizeof(*out));
        return e;
    };
    int got_type = -1;
    s = read_array(*da, dir, n, &got_type, Sink{ memory, dst, st }, nullptr);
    if (s != KMC_OK) return fail(s);

    out->values = dst;
    out->memory = memory;
    out->type = type;
    out->resolution[0] = res[0]; out->resolution[1] = res[1]; out->resolution[2] = res[2];
    out->veclen = veclen;
    out->kind = KMC_STRUCTURED;

    /* object coordinates: the Object tag's object_coord, else updateMinMaxCoords
     * (uniform: 0 .. resolution-1; rectilinear: first and last of each axis'
     * coordinates; curvilinear: a Math::Min / Math::Max scan) */
    float oc[6];
    if (!obj->attr("object_coord").empty()) {
        if (!parse_floats(obj->attr("object_coord"), oc, 6)) return fail(KMC_ERR_FILE);
    } else if (grid == "uniform") {
        for (int k = 0; k < 3; ++k) { oc[k] = 0.0f; oc[3 + k] = (float)res[k] - 1.0f; }
    } else if (grid == "rectilinear" || grid == "curvilinear") {
        const XmlNode* coord = node->child("Coord");
        const XmlNode* cda = coord ? coord->child("DataArray") : nullptr;
        if (!cda || type_of(cda->attr("type")) != KMC_REAL32) return fail(KMC_ERR_FILE);
        const size_t nc = grid == "rectilinear" ? (size_t)res[0] + res[1] + res[2] : nnodes * 3;
        std::vector<uint8_t> cbuf;
        int ct = -1;
        s = read_array(*cda, dir, nc, &ct, Sink{ KMC_MEMORY_HOST, nullptr, st }, &cbuf);
        if (s != KMC_OK || cbuf.size() < nc * 4) return fail(s != KMC_OK ? s : KMC_ERR_FILE);
        const float* c = (const float*)cbuf.data();
        if (grid == "rectilinear") {
            const size_t o[3] = { 0, res[0], (size_t)res[0] + res[1] };
            for (int k = 0; k < 3; ++k) { oc[k] = c[o[k]]; oc[3 + k] = c[o[k] + res[k] - 1]; }
        } else {
            for (int k = 0; k < 3; ++k) oc[k] = oc[3 + k] = c[k];
            for (size_t i = 3; i < nc; i += 3)
                for (int k = 0; k < 3; ++k) {
                    oc[k] = oc[k] < c[i + k] ? oc[k] : c[i + k];
                    oc[3 + k] = oc[3 + k] > c[i + k] ? oc[3 + k] : c[i + k];
                }
        }
    } else {
        for (int k = 0; k < 6; ++k) oc[k] = 0.0f;   // unknown grid type
    }
    out->has_min_max_coords = 1;
    for (int k = 0; k < 3; ++k) { out->min_coord[k] = oc[k]; out->max_coord[k] = oc[3 + k]; }

    /* min / max values (StructuredVolumeObject.cpp:214-226) */
    const std::string smin = value->attr("min_value"), smax = value->attr("max_value");
    double mn = smin.empty() ? 0.0 : std::atof(smin.c_str()), mx = smax.empty() ? 0.0 : std::atof(smax.c_str());
    if (smin.empty() || smax.empty()) {
        double cmn = 0, cmx = 0;
        if (veclen != 1) return fail(KMC_ERR_NOT_SCALAR);   // the magnitude scan of vector data is not supported
        if (memory == KMC_MEMORY_DEVICE) {
            s = min_max_on_device(out, dst, st, &cmn, &cmx);
        } else {
            void* tmp = device_alloc(bytes ? bytes : 1);
            if (!tmp) return fail(KMC_ERR_OUT_OF_MEMORY);
            s = hipMemcpyAsync(tmp, dst, bytes, hipMemcpyHostToDevice, st) == hipSuccess ? KMC_OK : KMC_ERR_DEVICE;
            if (s == KMC_OK) s = min_max_on_device(out, tmp, st, &cmn, &cmx);
            (void)hipStreamSynchronize(st);
            device_free(tmp);
        }
        if (s != KMC_OK) return fail(s);
        if (smin.empty()) mn = cmn;
        if (smax.empty()) mx = cmx;
    }
    out->has_min_max_values = 1;
    out->min_value = mn;
    out->max_value = mx;
    return KMC_OK;
}

/* ------------------------------------------------------------------------- */
/* AVS field                                                                   */
/* ------------------------------------------------------------------------- */

/* StructuredVolumeImporter( filename ) for an AVS field file
 * (Importer/StructuredVolumeImporter.cpp:37-60, 150-187) over
 * AVSField::read (FileFormat/AVSField/AVSField.cpp:174-419): the header's
 * "tag = value" lines up to the form feed, the values right after the first
 * "\f\f" of the file in the host's byte order, uniform fields only; object
 * coordinates 0 .. dim - 1; min / max by GetMinMaxValues on the device. */
kmc_status read_avs(const char* filename, int32_t memory, hipStream_t st, kmc_volume* out)
{
    if (!filename || !out) return KMC_ERR_INVALID_ARGUMENT;
    if (memory != KMC_MEMORY_HOST && memory != KMC_MEMORY_DEVICE) return KMC_ERR_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    /* AVSField::CheckExtension (:92-101) */
    const std::string fn(filename);
    const size_t dot = fn.find_last_of('.');
    const std::string ext = dot == std::string::npos ? "" : fn.substr(dot + 1);
    if (ext != "fld" && ext != "FLD") return KMC_ERR_FILE;
    FILE* f = std::fopen(filename, "rb");
    if (!f) return KMC_ERR_FILE;
    /* CheckHeader and read_header (:40-56, 216-300): lines of at most 255
     * characters; '#' lines skipped; a line starting with a form feed ends
     * the header; tag, '=', value split at " ,\t\n" */
    char buf[256];
    int veclen = 1, field = 1, dtype = -1;   // AVSField's defaults: uniform, no data type (AVSField.h:54-62)
    uint32_t dim[3] = { 1, 1, 1 };
    bool ok = std::fgets(buf, sizeof(buf), f) && std::strncmp(buf, "# AVS", 5) == 0;
    while (ok && std::fgets(buf, sizeof(buf), f)) {
        if (buf[0] == '#') continue;
        if (buf[0] == '\f') break;
        char* save = nullptr;
        const char* seps = " ,\t\n";
        char* tag = strtok_r(buf, seps, &save);
        char* equal = tag ? strtok_r(nullptr, seps, &save) : nullptr;
        char* value = equal ? strtok_r(nullptr, seps, &save) : nullptr;
        if (!equal || std::strcmp(equal, "=") || !value) continue;
        if (!std::strcmp(tag, "veclen")) veclen = std::atoi(value);
        if (!std::strcmp(tag, "dim1")) dim[0] = (uint32_t)std::atoi(value);
        if (!std::strcmp(tag, "dim2")) dim[1] = (uint32_t)std::atoi(value);
        if (!std::strcmp(tag, "dim3")) dim[2] = (uint32_t)std::atoi(value);
        if (!std::strcmp(tag, "field"))
            field = !std::strcmp(value, "uniform") ? 1 : !std::strcmp(value, "rectilinear") ? 2
                  : !std::strcmp(value, "irregular") ? 3 : 0;
        if (!std::strcmp(tag, "data"))
            dtype = !std::strcmp(value, "byte") ? KMC_UINT8 : !std::strcmp(value, "short") ? KMC_INT16
                  : !std::strcmp(value, "integer") ? KMC_INT32 : !std::strcmp(value, "float") ? KMC_REAL32
                  : !std::strcmp(value, "double") ? KMC_REAL64 : -1;
    }
    /* SkipHeader (:63-79): the values follow the first "\f\f" */
    long offset = -1;
    if (ok) {
        std::fseek(f, 0, SEEK_SET);
        int c;
        while ((c = std::fgetc(f)) != EOF)
            if (c == '\f' && std::fgetc(f) == '\f') { offset = std::ftell(f); break; }
    }
    std::fclose(f);
    if (!ok || offset < 0 || dtype < 0 || veclen < 1) return KMC_ERR_FILE;
    if (field != 1) return KMC_ERR_FILE;   // "'Uniform' type in kvs::AVSField format is only supported."

    kmc_status s = ensure_device();
    if (s != KMC_OK) return s;
    const size_t nnodes = (size_t)dim[0] * dim[1] * dim[2], n = nnodes * (size_t)veclen;
    const size_t bytes = n * type_size(dtype);
    void* dst = nullptr;
    if (memory == KMC_MEMORY_DEVICE) dst = device_alloc(bytes ? bytes : 1);
    else dst = host_block_alloc(bytes ? bytes : 1);
    if (!dst) return KMC_ERR_OUT_OF_MEMORY;
    auto fail = [&](kmc_status e) {
        if (memory == KMC_MEMORY_DEVICE) device_free(dst);
        else host_block_free(dst);
        std::memset(out, 0, sizeof(*out));
        return e;
    };
    s = read_binary(fn, bytes, type_size(dtype), false, Sink{ memory, dst, st }, offset);
    if (s != KMC_OK) return fail(s);
    out->values = dst;
    out->memory = memory;
    out->type = dtype;
    out->resolution[0] = dim[0]; out->resolution[1] = dim[1]; out->resolution[2] = dim[2];
    out->veclen = (uint32_t)veclen;
    out->kind = KMC_STRUCTURED;
    out->has_min_max_coords = 1;
    for (int k = 0; k < 3; ++k) { out->min_coord[k] = 0.0f; out->max_coord[k] = (float)dim[k] - 1.0f; }
    /* updateMinMaxValues */
    if (veclen != 1) return fail(KMC_ERR_NOT_SCALAR);   // the magnitude scan of vector data is not supported
    double mn = 0, mx = 0;
    if (memory == KMC_MEMORY_DEVICE) {
        s = min_max_on_device(out, dst, st, &mn, &mx);
    } else {
        void* tmp = device_alloc(bytes ? bytes : 1);
        if (!tmp) return fail(KMC_ERR_OUT_OF_MEMORY);
        s = hipMemcpyAsync(tmp, dst, bytes, hipMemcpyHostToDevice, st) == hipSuccess ? KMC_OK : KMC_ERR_DEVICE;
        if (s == KMC_OK) s = min_max_on_device(out, tmp, st, &mn, &mx);
        (void)hipStreamSynchronize(st);
        device_free(tmp);
    }
    if (s != KMC_OK) return fail(s);
    out->has_min_max_values = 1;
    out->min_value = mn;
    out->max_value = mx;
    return KMC_OK;
}

}  // namespace

}  // namespace kmc

extern "C" {

kmc_status kmc_avs_read_volume(const char* filename, int32_t memory, void* stream, kmc_volume* out)
{
    return kmc::abi_guard([&]() -> kmc_status {
        const kmc_status s = kmc::read_avs(filename, memory, (hipStream_t)stream, out);
        kmc::report(s);
        return s;
    });
}

kmc_status kmc_kvsml_read_volume(const char* filename, int32_t memory, void* stream, kmc_volume* out)
{
    return kmc::abi_guard([&]() -> kmc_status {
        const kmc_status s = kmc::read_volume(filename, memory, (hipStream_t)stream, out);
        kmc::report(s);
        return s;
    });
}

void kmc_kvsml_volume_release(kmc_volume* v)
{
    if (!v || !v->values) return;
    if (v->memory == KMC_MEMORY_DEVICE) kmc::device_free((void*)v->values);
    else kmc::host_block_free((void*)v->values);
    v->values = nullptr;
}

}  // extern "C"
