/*
 * kmc_api.hip -- host side of the C ABI (include/kmc.h): validation with the
 * reference's error behaviour, device memory, and the count -> scan -> emit
 * launch sequence of one extraction.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "kmc.h"
#include "kmc_guard.h"
#include "kmc_job.h"
#include "kmc_internal.h"

namespace kmc {

/* ------------------------------------------------------------------------- */
/* device memory pool                                                          */
/* ------------------------------------------------------------------------- */

namespace {

std::mutex g_pool_mutex;
std::multimap<size_t, void*> g_free;           // size -> block
std::unordered_map<void*, size_t> g_size;      // block -> size
size_t g_cached_bytes = 0;
/* Cached free bytes past which a free releases blocks: half of the device's
 * memory (64 GiB until the first allocation has asked the device).  Every
 * hipFree waits for the whole device, so a release while other streams are
 * busy stalls the releasing thread behind all of their kernels; below this
 * mark blocks are only released when an allocation fails. */
size_t g_keep_bytes = size_t(64) << 30;
bool g_keep_known = false;
const bool g_pool_trace = std::getenv("KMC_POOL_TRACE") != nullptr;

/* Size classes: 4 KiB steps up to 1 MiB, then four classes per power of two
 * (at most 25% slack), so that blocks are reused across similar sizes. */
size_t round_size(size_t n)
{
    if (n <= (size_t(1) << 20)) return (n + 4095) & ~size_t(4095);
    size_t p = size_t(1) << 20;
    while (p * 2 < n) p <<= 1;
    const size_t q = p / 4;
    return (n + q - 1) / q * q;
}

/* Take every cached block out of the pool (caller frees them unlocked). */
std::vector<void*> pool_take_all_locked()
{
    std::vector<void*> out;
    for (auto& kv : g_free) {
        out.push_back(kv.second);
        g_size.erase(kv.second);
    }
    g_free.clear();
    g_cached_bytes = 0;
    return out;
}

void free_blocks(const std::vector<void*>& blocks)
{
    for (void* p : blocks) (void)hipFree(p);
}

}  // namespace

void* device_alloc(size_t n)
{
    if (n == 0) n = 1;
    const size_t sz = round_size(n);
    {
        std::lock_guard<std::mutex> lk(g_pool_mutex);
        auto it = g_free.lower_bound(sz);
        if (it != g_free.end() && it->first <= sz * 2) {
            void* p = it->second;
            g_cached_bytes -= it->first;
            g_free.erase(it);
            return p;
        }
    }
    /* a miss: allocate without holding the pool lock */
    if (g_pool_trace) std::fprintf(stderr, "[kmc_pool] hipMalloc %zu bytes\n", sz);
    if (!g_keep_known) {
        size_t fr = 0, total = 0;
        if (hipMemGetInfo(&fr, &total) == hipSuccess && total > 0) {
            std::lock_guard<std::mutex> lk(g_pool_mutex);
            g_keep_bytes = std::max(g_keep_bytes, total / 2);
            g_keep_known = true;
        } else {
            (void)hipGetLastError();
        }
    }
    void* p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        std::vector<void*> all;
        {
            std::lock_guard<std::mutex> lk(g_pool_mutex);
            all = pool_take_all_locked();
        }
        if (g_pool_trace) std::fprintf(stderr, "[kmc_pool] out of memory: releasing %zu cached blocks\n", all.size());
        free_blocks(all);
        if (hipMalloc(&p, sz) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    std::lock_guard<std::mutex> lk(g_pool_mutex);
    g_size[p] = sz;
    return p;
}

void device_free(void* p)
{
    if (!p) return;
    std::vector<void*> release;
    {
        std::lock_guard<std::mutex> lk(g_pool_mutex);
        auto it = g_size.find(p);
        if (it == g_size.end()) return;
        g_free.emplace(it->second, p);
        g_cached_bytes += it->second;
        if (g_cached_bytes > g_keep_bytes) {
            /* release the largest cached blocks down to half the mark */
            while (g_cached_bytes > g_keep_bytes / 2 && !g_free.empty()) {
                auto last = std::prev(g_free.end());
                release.push_back(last->second);
                g_cached_bytes -= last->first;
                g_size.erase(last->second);
                g_free.erase(last);
            }
        }
    }
    if (!release.empty()) {
        if (g_pool_trace) std::fprintf(stderr, "[kmc_pool] releasing %zu cached blocks\n", release.size());
        free_blocks(release);
    }
}

void device_pool_trim()
{
    std::vector<void*> all;
    {
        std::lock_guard<std::mutex> lk(g_pool_mutex);
        all = pool_take_all_locked();
    }
    free_blocks(all);
}

/* Pinned host blocks for the surfaces returned in host memory (and host
 * volumes read from files), cached like the device blocks: hipHostMalloc pins
 * every page of a block and hipHostFree may wait for the device, each far
 * slower than the copy into the block.  Freed blocks stay cached up to
 * kHostKeep bytes (past that mark a free releases the largest ones), and all
 * are released when a pinned allocation fails. */
namespace {
std::mutex g_host_mutex;
std::multimap<size_t, void*> g_host_free;
std::unordered_map<void*, size_t> g_host_size;
size_t g_host_cached = 0;
constexpr size_t kHostKeep = size_t(4) << 30;
}

void* host_block_alloc(size_t n)
{
    if (n == 0) n = 1;
    const size_t sz = round_size(n);
    {
        std::lock_guard<std::mutex> lk(g_host_mutex);
        auto it = g_host_free.lower_bound(sz);
        if (it != g_host_free.end() && it->first <= sz * 2) {
            void* p = it->second;
            g_host_cached -= it->first;
            g_host_free.erase(it);
            return p;
        }
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, sz, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        std::vector<void*> all;
        {
            std::lock_guard<std::mutex> lk(g_host_mutex);
            for (auto& kv : g_host_free) {
                all.push_back(kv.second);
                g_host_size.erase(kv.second);
            }
            g_host_free.clear();
            g_host_cached = 0;
        }
        for (void* q : all) (void)hipHostFree(q);
        p = nullptr;
        if (hipHostMalloc(&p, sz, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    std::lock_guard<std::mutex> lk(g_host_mutex);
    g_host_size[p] = sz;
    return p;
}

void host_block_free(void* p)
{
    if (!p) return;
    std::vector<void*> release;
    {
        std::lock_guard<std::mutex> lk(g_host_mutex);
        auto it = g_host_size.find(p);
        if (it == g_host_size.end()) return;
        g_host_free.emplace(it->second, p);
        g_host_cached += it->second;
        while (g_host_cached > kHostKeep && !g_host_free.empty()) {
            auto last = std::prev(g_host_free.end());
            release.push_back(last->second);
            g_host_cached -= last->first;
            g_host_size.erase(last->second);
            g_host_free.erase(last);
        }
    }
    for (void* q : release) (void)hipHostFree(q);
}

/* Pinned host blocks for small staging buffers, by power-of-two size class;
 * reused for the life of the process (hipHostMalloc / hipHostFree are slow
 * and the latter may wait for the device). */
namespace {
std::mutex g_pin_mutex;
std::multimap<size_t, void*> g_pin_free;
std::unordered_map<void*, size_t> g_pin_size;
}

void* host_pinned_alloc(size_t n)
{
    size_t sz = 4096;
    while (sz < n) sz <<= 1;
    std::lock_guard<std::mutex> lk(g_pin_mutex);
    auto it = g_pin_free.find(sz);
    if (it != g_pin_free.end()) {
        void* p = it->second;
        g_pin_free.erase(it);
        return p;
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, sz, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_pin_size[p] = sz;
    return p;
}

void host_pinned_free(void* p)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_pin_mutex);
    auto it = g_pin_size.find(p);
    if (it != g_pin_size.end()) g_pin_free.emplace(it->second, p);
}

/* n bytes (totals, sizes, min/max) from device memory into host memory,
 * through a pinned staging block, then a wait for the stream: a copy into
 * pageable memory is staged by the runtime at a higher latency, and these
 * copies sit between a pass's front and its emit. */
hipError_t read_back_parts(const ReadPart* parts, int n, hipStream_t st)
{
    size_t total = 0;
    for (int i = 0; i < n; ++i) total += (parts[i].bytes + 15) & ~size_t(15);
    uint8_t* pin = (uint8_t*)host_pinned_alloc(total);
    if (!pin) return hipErrorOutOfMemory;
    hipError_t e = hipSuccess;
    for (int i = 0, off = 0; i < n && e == hipSuccess; off += (int)((parts[i].bytes + 15) & ~size_t(15)), ++i)
        e = hipMemcpyAsync(pin + off, parts[i].dev, parts[i].bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess)
        for (int i = 0, off = 0; i < n; off += (int)((parts[i].bytes + 15) & ~size_t(15)), ++i)
            std::memcpy(parts[i].host, pin + off, parts[i].bytes);
    host_pinned_free(pin);
    return e;
}

hipError_t read_back(void* host, const void* dev, size_t n, hipStream_t st)
{
    void* pin = host_pinned_alloc(n);
    if (!pin) return hipErrorOutOfMemory;
    hipError_t e = hipMemcpyAsync(pin, dev, n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) std::memcpy(host, pin, n);
    host_pinned_free(pin);
    return e;
}

/* ------------------------------------------------------------------------- */
/* surface ownership                                                           */
/* ------------------------------------------------------------------------- */

void free_storage(SurfaceStorage* s)
{
    if (!s) return;
    if (s->memory == KMC_MEMORY_DEVICE) {
        device_free(s->coords);
        device_free(s->connections);
        device_free(s->normals);
        device_free(s->colors);
    } else {
        host_block_free(s->coords);
        host_block_free(s->connections);
        host_block_free(s->normals);
        host_block_free(s->colors);
    }
    delete s;
}

/* ------------------------------------------------------------------------- */
/* helpers                                                                     */
/* ------------------------------------------------------------------------- */

size_t type_size(int t)
{
    switch (t) {
    case KMC_INT8: case KMC_UINT8: return 1;
    case KMC_INT16: case KMC_UINT16: return 2;
    case KMC_INT32: case KMC_UINT32: case KMC_REAL32: return 4;
    case KMC_INT64: case KMC_UINT64: case KMC_REAL64: return 8;
    default: return 0;
    }
}

/* The largest float t with double(t) <= iso, so that for float v:
 * double(v) > iso  <=>  v > t.  NaN stays NaN (every comparison false). */
float float_threshold(double iso)
{
    if (std::isnan(iso)) return NAN;
    float t = (float)iso;
    if ((double)t > iso) t = std::nextafter(t, -INFINITY);
    return t;
}

kmc_status validate_volume(const kmc_volume* v)
{
    if (!v) return KMC_ERR_NULL_INPUT;
    if (v->kind != KMC_STRUCTURED) return KMC_ERR_NOT_STRUCTURED;
    if (v->veclen != 1) return KMC_ERR_NOT_SCALAR;
    if (type_size(v->type) == 0) return KMC_ERR_UNSUPPORTED_TYPE;
    if (!v->values) return KMC_ERR_INVALID_ARGUMENT;
    if (v->resolution[0] == 0 || v->resolution[1] == 0 || v->resolution[2] == 0) return KMC_ERR_INVALID_ARGUMENT;
    if (v->resolution[0] > (1u << 20)) return KMC_ERR_INVALID_ARGUMENT;   // x packed in 20 bits
    if (v->memory != KMC_MEMORY_HOST && v->memory != KMC_MEMORY_DEVICE) return KMC_ERR_INVALID_ARGUMENT;
    return KMC_OK;
}

void report(kmc_status s)
{
    if (s == KMC_OK) return;
    if (std::getenv("KMC_QUIET")) return;
    std::fprintf(stderr, "KVS ERROR: %s\n", kmc_status_string(s));
}

namespace {
thread_local int t_device = 0;
thread_local bool t_device_set = false;
}

kmc_status ensure_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return KMC_ERR_NO_DEVICE;
    }
    if (!t_device_set) { t_device = 0; t_device_set = true; }
    if (t_device >= n) return KMC_ERR_NO_DEVICE;
    if (hipSetDevice(t_device) != hipSuccess) { (void)hipGetLastError(); return KMC_ERR_DEVICE; }
    return KMC_OK;
}

/* Device copy of the volume values (or the caller's device pointer). */
/* Pageable host memory to device memory.  The runtime stages a pageable
 * copy itself at about half the pinned rate; large copies instead go through
 * two pinned blocks in turn: a chunk is copied into one block by up to
 * kUploadThreads threads while the other block's transfer runs, and a block
 * is refilled only after its previous transfer has completed (its event). */
constexpr size_t kUploadChunk = size_t(64) << 20;
constexpr size_t kUploadPart = size_t(4) << 20;
constexpr size_t kUploadThreads = 8;

static void copy_parallel(uint8_t* dst, const uint8_t* src, size_t n)
{
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t threads =
        std::max<size_t>(1, std::min<size_t>({ kUploadThreads, hw ? (size_t)hw : 1, n / kUploadPart + 1 }));
    const size_t part = (n + threads - 1) / threads;
    auto work = [&](size_t t) {
        const size_t b = t * part;
        if (b < n) std::memcpy(dst + b, src + b, std::min(n, b + part) - b);
    };
    std::vector<std::thread> pool;
    for (size_t t = 1; t < threads; ++t) pool.emplace_back(work, t);
    work(0);
    for (std::thread& th : pool) th.join();
}

hipError_t upload_host(void* dst, const void* src, size_t bytes, hipStream_t st)
{
    if (bytes < 2 * kUploadPart) return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    /* at least four chunks where the volume allows, so that copies overlap */
    const size_t chunk = std::min(kUploadChunk, std::max(2 * kUploadPart, (bytes / 4 + 4095) & ~size_t(4095)));
    uint8_t* stage[2] = { (uint8_t*)host_pinned_alloc(kUploadChunk), (uint8_t*)host_pinned_alloc(kUploadChunk) };
    hipEvent_t done[2] = { nullptr, nullptr };
    bool used[2] = { false, false };
    hipError_t e = hipSuccess;
    if (!stage[0] || !stage[1] || hipEventCreateWithFlags(&done[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&done[1], hipEventDisableTiming) != hipSuccess)
        e = hipErrorOutOfMemory;
    for (size_t off = 0, k = 0; e == hipSuccess && off < bytes; off += chunk, k ^= 1) {
        const size_t n = std::min(chunk, bytes - off);
        if (used[k] && (e = hipEventSynchronize(done[k])) != hipSuccess) break;
        copy_parallel(stage[k], (const uint8_t*)src + off, n);
        if ((e = hipMemcpyAsync((uint8_t*)dst + off, stage[k], n, hipMemcpyHostToDevice, st)) != hipSuccess ||
            (e = hipEventRecord(done[k], st)) != hipSuccess)
            break;
        used[k] = true;
    }
    /* the blocks go back to the pool only when no transfer reads them */
    for (int k = 0; k < 2; ++k)
        if (used[k]) {
            const hipError_t w = hipEventSynchronize(done[k]);
            if (e == hipSuccess) e = w;
        }
    for (int k = 0; k < 2; ++k) {
        if (done[k]) (void)hipEventDestroy(done[k]);
        host_pinned_free(stage[k]);
    }
    if (e != hipSuccess) (void)hipGetLastError();
    return e;
}

kmc_status device_values(const kmc_volume* v, hipStream_t st, const void** dvals, void** owned)
{
    *owned = nullptr;
    if (v->memory == KMC_MEMORY_DEVICE) { *dvals = v->values; return KMC_OK; }
    const size_t bytes = (size_t)v->resolution[0] * v->resolution[1] * v->resolution[2] * type_size(v->type);
    void* p = device_alloc(bytes);
    if (!p) return KMC_ERR_OUT_OF_MEMORY;
    if (upload_host(p, v->values, bytes, st) != hipSuccess) {
        device_free(p);
        return KMC_ERR_DEVICE;
    }
    *owned = p;
    *dvals = p;
    return KMC_OK;
}

kmc_status min_max_on_device(const kmc_volume* v, const void* dvals, hipStream_t st, double* mn, double* mx)
{
    const uint64_t n = (uint64_t)v->resolution[0] * v->resolution[1] * v->resolution[2];
    void* scratch = device_alloc(minmax_scratch_bytes() + 16);
    if (!scratch) return KMC_ERR_OUT_OF_MEMORY;
    double* dout = (double*)((char*)scratch + minmax_scratch_bytes());
    kmc_status s = KMC_OK;
    double host[2] = {0, 0};
    if (launch_minmax(v->type, dvals, n, scratch, dout, st) != hipSuccess ||
        read_back(host, dout, sizeof(host), st) != hipSuccess) {
        s = KMC_ERR_DEVICE;
    }
    device_free(scratch);
    *mn = host[0];
    *mx = host[1];
    return s;
}

/* Row geometry every emit needs: 64-node chunks per row, the lane split of
 * the count and compaction passes, node rows. */
void plan_rows(JobArgs& a)
{
    a.NC = (a.nx + 63) / 64;
    a.cshift = 0;
    while ((1u << a.cshift) < a.NC && a.cshift < 6) ++a.cshift;
    a.nrows = (uint64_t)a.nz * a.ny;
}

bool flat_emit_fits(uint32_t ny, uint32_t nz)
{
    const char* env_emit = std::getenv("KMC_EMIT");
    return ny <= (1u << 20) && nz <= (1u << 20) && !(env_emit && std::strcmp(env_emit, "team") == 0);
}

bool plan_geometry(JobArgs& a, uint32_t* lds_bytes, uint32_t r_max, uint32_t queue_cap)
{
    /* tuning overrides: KMC_TILE_ROWS sets the largest R, KMC_QUEUE_CAP the work-queue
     * size, KMC_FN_CAP the face normals kept per cell slice (VertexNormal), KMC_TEAM_LDS
     * the LDS budget of one team, KMC_SLAB the steps per unit */
    static const char* env_r = std::getenv("KMC_TILE_ROWS");
    static const char* env_q = std::getenv("KMC_QUEUE_CAP");
    static const char* env_f = std::getenv("KMC_FN_CAP");
    static const char* env_b = std::getenv("KMC_TEAM_LDS");
    static const char* env_s = std::getenv("KMC_SLAB");
    if (env_r) r_max = (uint32_t)std::max(1, std::atoi(env_r));
    r_max = std::min<uint32_t>(r_max, 31);                 // cell rows are 5-bit queue fields
    if (queue_cap == 0) queue_cap = a.vnormal ? 112u : 128u;
    if (env_q) queue_cap = (uint32_t)std::max(64, std::atoi(env_q));
    plan_rows(a);
    a.queue_cap = queue_cap;
    a.fn_cap = a.vnormal ? (env_f ? (uint32_t)std::max(64, std::atoi(env_f)) : 112u) : 0u;
    /* the tallest band whose team fits the budget (VertexNormal: four 4-team
     * workgroups per CU, else five); for very long rows R = 1 with as many
     * teams as a 64 KiB workgroup holds */
    const uint32_t budget = env_b ? (uint32_t)std::max(1024, std::atoi(env_b))
                                  : (a.vnormal ? 40960u : 32256u) / kEmitTeams - kTableBytes / kEmitTeams;
    uint32_t pick = 0, pick_bytes = 0;
    for (uint32_t R = r_max; R >= 1 && !pick; --R) {
        a.R = R;
        const LdsLayout L = lds_layout(a);
        if (L.bytes <= budget) { pick = R; pick_bytes = L.bytes; }
    }
    a.teams = kEmitTeams;
    if (!pick) {
        a.R = 1;
        const LdsLayout L = lds_layout(a);
        a.teams = std::min<uint32_t>(kEmitTeams, (65536u - kTableBytes) / L.bytes);
        if (a.teams == 0) return false;
        pick = 1;
        pick_bytes = L.bytes;
    }
    a.R = pick;
    a.nb = (a.ny + pick - 1) / pick;
    a.team_stride = pick_bytes;
    *lds_bytes = emit_lds_bytes(pick_bytes, a.teams);
    /* steps per unit: about 65536 units when the volume allows (load balance over
     * some 4096 resident teams), at most 64 steps; VertexNormal needs 2 (a unit's
     * first step only prepares the face normals of the slice below) */
    const uint64_t want = ((uint64_t)a.nz * a.nb * std::max(1u, a.nlev)) / 65536;
    const uint64_t lo = a.vnormal ? 2 : 1;
    a.slab = env_s ? (uint32_t)std::max(1, std::atoi(env_s))
                   : (uint32_t)std::min<uint64_t>(64, std::max<uint64_t>(lo, want));
    a.nslabs = (a.nz + a.slab - 1) / a.slab;
    a.nunits = (uint64_t)a.nlev * a.nb * a.nslabs;
    return true;
}

uint32_t levels_per_pass(const kmc_volume* v, int vnormal)
{
    /* the flat emit keeps no per-team LDS state: every row length fits */
    if (flat_emit_fits(v->resolution[1], v->resolution[2])) return kMaxLevels;
    for (uint32_t nl = kMaxLevels; nl >= 1; --nl) {
        JobArgs a;
        std::memset(&a, 0, sizeof(a));
        a.nx = v->resolution[0];
        a.ny = v->resolution[1];
        a.nz = v->resolution[2];
        a.nlev = nl;
        a.vnormal = vnormal ? 1u : 0u;
        uint32_t lds = 0;
        if (plan_geometry(a, &lds)) return nl;
    }
    return 0;
}

void init_surface(kmc_surface* s, double iso, int normal_type, int dup, double mn, double mx,
                  const float minc[3], const float maxc[3])
{
    std::memset(s, 0, sizeof(*s));
    s->isolevel = iso;
    s->memory = KMC_MEMORY_DEVICE;
    s->polygon_type = 3;
    s->normal_type = dup ? KMC_POLYGON_NORMAL : normal_type;
    s->color_type = 1;
    s->min_value = mn;
    s->max_value = mx;
    for (int k = 0; k < 3; ++k) { s->min_coord[k] = minc[k]; s->max_coord[k] = maxc[k]; }
}

/* Object coordinates the mapper uses (MapperBase::setMinMaxCoords -> updateMinMaxCoords). */
void object_coords(const kmc_volume* v, float minc[3], float maxc[3])
{
    if (v->has_min_max_coords) {
        for (int k = 0; k < 3; ++k) { minc[k] = v->min_coord[k]; maxc[k] = v->max_coord[k]; }
    } else {
        for (int k = 0; k < 3; ++k) { minc[k] = 0.0f; maxc[k] = (float)v->resolution[k] - 1.0f; }
    }
}

static inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

/* Per-job fields of a pass of up to kMaxLevels levels. */
void init_job(JobArgs& a, const kmc_volume* v, const void* dvals, const double* iso, uint32_t nlev, int normal_type,
              int dup, double mn, double mx, const float minc[3], const float maxc[3])
{
    std::memset(&a, 0, sizeof(a));
    a.values = dvals;
    a.nx = v->resolution[0];
    a.ny = v->resolution[1];
    a.nz = v->resolution[2];
    a.nlev = nlev;
    a.dup = dup ? 1u : 0u;
    a.vnormal = (!dup && normal_type == KMC_VERTEX_NORMAL) ? 1u : 0u;
    a.minv = mn;
    a.maxv = mx;
    for (int k = 0; k < 3; ++k) {
        a.minc[k] = minc[k];
        const float span = maxc[k] - minc[k];          // (max_coord - min_coord), float
        a.sc[k] = span / (float)(v->resolution[k] - 1); // / Vec3(ncells), float (MarchingCubes.cpp:166-168)
    }
    for (uint32_t l = 0; l < nlev; ++l) {
        a.iso[l] = iso[l];
        a.thr[l] = float_threshold(iso[l]);
    }
}

/* Device time of each kernel of a pass, when this thread asked for it
 * (kmc_set_timing): HIP events on the pass's stream between the launches. */
namespace {
thread_local bool t_timing = false;
thread_local kmc_timing t_last;
thread_local hipEvent_t t_ev[9] = { nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr };
}

struct PassTimer {
    hipStream_t st;
    bool on;
    bool flat = false;
    explicit PassTimer(hipStream_t s) : st(s), on(t_timing)
    {
        for (int i = 0; i < 9 && on; ++i)
            if (!t_ev[i] && hipEventCreate(&t_ev[i]) != hipSuccess) { (void)hipGetLastError(); on = false; }
    }
    /* events between the flat emit's kernels (after compaction, vertices, triangles) */
    hipEvent_t* flat_marks() { flat = on; return on ? &t_ev[6] : nullptr; }
    bool mark(int i)
    {
        if (on && hipEventRecord(t_ev[i], st) != hipSuccess) { (void)hipGetLastError(); on = false; }
        return true;
    }
    void accumulate()
    {
        if (!on) return;
        float ms[5] = { 0, 0, 0, 0, 0 };
        for (int i = 0; i < 5; ++i) (void)hipEventElapsedTime(&ms[i], t_ev[i], t_ev[i + 1]);
        t_last.classify_ms += ms[0];
        t_last.count_ms += ms[1];
        t_last.scan_ms += ms[2];
        t_last.emit_ms += ms[4];
        t_last.passes += 1;
        if (flat) {
            float f[4] = { 0, 0, 0, 0 };
            (void)hipEventElapsedTime(&f[0], t_ev[4], t_ev[6]);
            (void)hipEventElapsedTime(&f[1], t_ev[6], t_ev[7]);
            (void)hipEventElapsedTime(&f[2], t_ev[7], t_ev[8]);
            (void)hipEventElapsedTime(&f[3], t_ev[8], t_ev[5]);
            t_last.compact_ms += f[0];
            t_last.vertices_ms += f[1];
            t_last.triangles_ms += f[2];
            t_last.vertex_normals_ms += f[3];
        }
    }
};

/* One pass of up to kMaxLevels levels over a device-resident volume.
 * On success outs[i] hold device arrays. */
kmc_status run_pass(const kmc_volume* v, const void* dvals, const double* iso, uint32_t nlev,
                    int normal_type, int dup, double mn, double mx, const float minc[3], const float maxc[3],
                    hipStream_t st, kmc_surface* outs)
{
    JobArgs a;
    init_job(a, v, dvals, iso, nlev, normal_type, dup, mn, mx, minc, maxc);
    /* flat emit (compaction records + one thread per item) unless the records'
     * 20-bit y / z fields do not fit or KMC_EMIT=team asks for the tile teams
     * (read per call); only the tile teams need their LDS state to fit */
    const bool flat = flat_emit_fits(a.ny, a.nz);
    uint32_t lds = 0;
    if (!plan_geometry(a, &lds) && !flat) return KMC_ERR_INVALID_ARGUMENT;
    plan_rows(a);
    a.rec_levels = (1u << nlev) - 1u;

    const uint32_t narr = 2 * nlev;
    const uint64_t n = a.nrows;
    const uint32_t nparts = scan_partials_needed(n, narr);
    const size_t counts_bytes = (size_t)narr * n * sizeof(uint32_t);
    const size_t prefix_bytes = (size_t)narr * n * sizeof(uint64_t);
    uint32_t* counts = (uint32_t*)device_alloc(counts_bytes);
    uint64_t* prefix = (uint64_t*)device_alloc(prefix_bytes);
    uint64_t* aux = (uint64_t*)device_alloc((size_t)(nparts + narr) * sizeof(uint64_t));
    const size_t chunks = (size_t)nlev * a.nz * a.ny * a.NC;
    uint64_t* bitvol = (uint64_t*)device_alloc(chunks * sizeof(uint64_t));
    uint32_t* cpre = flat ? (uint32_t*)device_alloc(2 * chunks * sizeof(uint32_t)) : nullptr;
    ChunkCross* cross = (flat && !dup) ? (ChunkCross*)device_alloc(chunks * sizeof(ChunkCross)) : nullptr;
    void* recs = nullptr;
    auto cleanup = [&]() {
        device_free(counts); device_free(prefix); device_free(aux); device_free(bitvol); device_free(cpre);
        device_free(cross); device_free(recs);
    };
    if (!counts || !prefix || !aux || !bitvol || (flat && !cpre) || (flat && !dup && !cross)) {
        cleanup();
        return KMC_ERR_OUT_OF_MEMORY;
    }
    a.bitvol = bitvol;
    if (flat) {
        a.vpc = cpre;
        a.tpc = cpre + chunks;
        a.cross = cross;
    }
    uint64_t* partials = aux;
    uint64_t* totals = aux + nparts;
    a.counts = counts;
    a.prefix = prefix;

    PassTimer timer(st);
    timer.mark(0);
    const bool front_ok = launch_classify(v->type, a, st) == hipSuccess && timer.mark(1) &&
                          launch_count(a, st) == hipSuccess && timer.mark(2) &&
                          launch_scan(counts, n, narr, partials, prefix, totals, st) == hipSuccess && timer.mark(3);
    if (!front_ok) {
        cleanup();
        return KMC_ERR_DEVICE;
    }
    std::vector<uint64_t> tot(narr, 0);
    if (read_back(tot.data(), totals, narr * sizeof(uint64_t), st) != hipSuccess) {
        cleanup();
        return KMC_ERR_DEVICE;
    }

    kmc_status status = KMC_OK;
    for (uint32_t l = 0; l < nlev; ++l) {
        kmc_surface* s = &outs[l];
        init_surface(s, iso[l], normal_type, dup, mn, mx, minc, maxc);
        const uint64_t nv = tot[l * 2 + 0], nt = tot[l * 2 + 1];
        SurfaceStorage* store = new SurfaceStorage();
        store->memory = KMC_MEMORY_DEVICE;
        store->coords = store->connections = store->normals = nullptr;
        s->internal = store;
        uint64_t verts = dup ? 3 * nt : nv;
        if (verts == 0) {
            /* coords.size() == 0: the reference sets no arrays (:236-242, :282-289) */
            s->n_triangles = 0;
            continue;
        }
        s->n_vertices = verts;
        s->n_triangles = nt;
        s->n_connections = dup ? 0 : 3 * nt;
        s->n_normals = (!dup && normal_type == KMC_VERTEX_NORMAL) ? nv : nt;
        store->coords = device_alloc(verts * 3 * sizeof(float));
        store->normals = s->n_normals ? device_alloc(s->n_normals * 3 * sizeof(float)) : nullptr;
        store->connections = s->n_connections ? device_alloc(s->n_connections * sizeof(uint32_t)) : nullptr;
        if (!store->coords || (s->n_normals && !store->normals) || (s->n_connections && !store->connections)) {
            status = KMC_ERR_OUT_OF_MEMORY;
            break;
        }
        s->coords = (float*)store->coords;
        s->normals = (float*)store->normals;
        s->connections = (uint32_t*)store->connections;
        a.coords[l] = s->coords;
        a.normals[l] = s->normals;
        a.conn[l] = s->connections;
        a.cap_vertices[l] = verts;
        a.cap_triangles[l] = nt;
    }
    if (status == KMC_OK && flat) {
        /* records: vertex keys, triangle records and cube indices, face normals */
        a.voff[0] = a.toff[0] = 0;
        for (uint32_t l = 0; l < nlev; ++l) {
            a.voff[l + 1] = a.voff[l] + (dup ? 0 : tot[l * 2 + 0]);
            a.toff[l + 1] = a.toff[l] + tot[l * 2 + 1];
        }
        const uint64_t nv = a.voff[nlev], nt = a.toff[nlev];
        const size_t bytes = align_up(nv * 8, 256) + align_up(nt * 8, 256) + align_up(nt, 256) +
                             (a.vnormal ? align_up(nt * 12, 256) + nv * 16 : 0);
        if (bytes) {
            recs = device_alloc(bytes);
            if (!recs) status = KMC_ERR_OUT_OF_MEMORY;
        }
        if (recs) {
            uint8_t* p = (uint8_t*)recs;
            a.vkey = (uint64_t*)p; p += align_up(nv * 8, 256);
            a.trec = (uint64_t*)p; p += align_up(nt * 8, 256);
            a.tci = p; p += align_up(nt, 256);
            a.fn = a.vnormal ? (float*)p : nullptr;
            if (a.vnormal) p += align_up(nt * 12, 256);
            a.vslot = a.vnormal ? (uint32_t*)p : nullptr;
        }
    }
    if (status == KMC_OK) {
        bool any = false;
        for (uint32_t l = 0; l < nlev; ++l) any = any || outs[l].n_vertices > 0;
        timer.mark(4);
        if (any && (flat ? launch_flat_emit(v->type, a, st, timer.flat_marks()) : launch_emit(v->type, a, lds, st)) !=
                       hipSuccess)
            status = KMC_ERR_DEVICE;
        timer.mark(5);
        if (status == KMC_OK && hipStreamSynchronize(st) != hipSuccess) status = KMC_ERR_DEVICE;
        if (status == KMC_OK) timer.accumulate();
    }
    cleanup();
    if (status != KMC_OK) {
        for (uint32_t l = 0; l < nlev; ++l) kmc_surface_release(&outs[l]);
    }
    return status;
}

kmc_status to_host(kmc_surface* s, hipStream_t st)
{
    SurfaceStorage* old = (SurfaceStorage*)s->internal;
    SurfaceStorage* hs = new SurfaceStorage();
    hs->memory = KMC_MEMORY_HOST;
    hs->coords = hs->connections = hs->normals = nullptr;
    const size_t bc = s->n_vertices * 3 * sizeof(float);
    const size_t bn = s->n_normals * 3 * sizeof(float);
    const size_t bk = s->n_connections * sizeof(uint32_t);
    const size_t bl = s->colors ? s->n_colors * 3 : 0;
    bool ok = true;
    if (bc) ok = ok && (hs->coords = host_block_alloc(bc)) != nullptr;
    if (bn) ok = ok && (hs->normals = host_block_alloc(bn)) != nullptr;
    if (bk) ok = ok && (hs->connections = host_block_alloc(bk)) != nullptr;
    if (bl) ok = ok && (hs->colors = host_block_alloc(bl)) != nullptr;
    if (ok && bc) ok = hipMemcpyAsync(hs->coords, s->coords, bc, hipMemcpyDeviceToHost, st) == hipSuccess;
    if (ok && bn) ok = hipMemcpyAsync(hs->normals, s->normals, bn, hipMemcpyDeviceToHost, st) == hipSuccess;
    if (ok && bk) ok = hipMemcpyAsync(hs->connections, s->connections, bk, hipMemcpyDeviceToHost, st) == hipSuccess;
    if (ok && bl) ok = hipMemcpyAsync(hs->colors, s->colors, bl, hipMemcpyDeviceToHost, st) == hipSuccess;
    if (ok) ok = hipStreamSynchronize(st) == hipSuccess;
    if (!ok) {
        /* no copy still in flight may write into blocks that go back to the cache */
        (void)hipStreamSynchronize(st);
        (void)hipGetLastError();
        free_storage(hs);
        return KMC_ERR_OUT_OF_MEMORY;
    }
    free_storage(old);
    s->internal = hs;
    s->memory = KMC_MEMORY_HOST;
    s->coords = (float*)hs->coords;
    s->normals = (float*)hs->normals;
    s->connections = (uint32_t*)hs->connections;
    s->colors = (uint8_t*)hs->colors;
    return KMC_OK;
}

kmc_status extract_levels_impl(const kmc_volume* v, const double* levels, uint32_t n_levels, int32_t normal_type,
                               int32_t duplication, int32_t out_memory, void* stream, kmc_surface* outs)
{
    std::memset(&t_last, 0, sizeof(t_last));
    kmc_status s = validate_volume(v);
    if (s != KMC_OK) return s;
    if (!outs || (n_levels > 0 && !levels)) return KMC_ERR_INVALID_ARGUMENT;
    if (normal_type != KMC_VERTEX_NORMAL && normal_type != KMC_POLYGON_NORMAL) return KMC_ERR_INVALID_ARGUMENT;
    if (out_memory != KMC_MEMORY_HOST && out_memory != KMC_MEMORY_DEVICE) return KMC_ERR_INVALID_ARGUMENT;
    const int dup = (normal_type == KMC_VERTEX_NORMAL) ? 0 : (duplication ? 1 : 0);   // :69-73
    s = ensure_device();
    if (s != KMC_OK) return s;
    hipStream_t st = (hipStream_t)stream;

    const void* dvals = nullptr;
    void* owned = nullptr;
    s = device_values(v, st, &dvals, &owned);
    if (s != KMC_OK) return s;

    double mn = v->min_value, mx = v->max_value;
    if (!v->has_min_max_values) {
        s = min_max_on_device(v, dvals, st, &mn, &mx);
        if (s != KMC_OK) { device_free(owned); return s; }
    }
    float minc[3], maxc[3];
    object_coords(v, minc, maxc);

    const double d = mn - mx;
    if ((d < 0 ? -d : d) < 1.e-6) {
        /* Math::Equal(min, max): no surface (MarchingCubes.cpp:108-110) */
        for (uint32_t l = 0; l < n_levels; ++l) {
            init_surface(&outs[l], levels[l], normal_type, dup, mn, mx, minc, maxc);
            outs[l].empty_range = 1;
            outs[l].memory = out_memory;
        }
        if (owned) { (void)hipStreamSynchronize(st); device_free(owned); }
        return KMC_OK;
    }

    const uint32_t per = levels_per_pass(v, !dup && normal_type == KMC_VERTEX_NORMAL);
    if (per == 0 && n_levels > 0) s = KMC_ERR_INVALID_ARGUMENT;   // rows too long for the tile teams' LDS
    for (uint32_t l0 = 0; l0 < n_levels && s == KMC_OK; l0 += per) {
        const uint32_t nl = (n_levels - l0) < per ? (n_levels - l0) : per;
        s = run_pass(v, dvals, levels + l0, nl, normal_type, dup, mn, mx, minc, maxc, st, outs + l0);
        if (s != KMC_OK) {
            for (uint32_t l = 0; l < l0; ++l) kmc_surface_release(&outs[l]);
        }
    }
    if (s == KMC_OK && out_memory == KMC_MEMORY_HOST) {
        for (uint32_t l = 0; l < n_levels && s == KMC_OK; ++l) s = to_host(&outs[l], st);
        if (s != KMC_OK)
            for (uint32_t l = 0; l < n_levels; ++l) kmc_surface_release(&outs[l]);
    }
    if (owned) { (void)hipStreamSynchronize(st); device_free(owned); }
    return s;
}

}  // namespace kmc

using namespace kmc;

extern "C" {

kmc_status kmc_init(int device)
{
    return kmc::abi_guard([&]() -> kmc_status {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) { (void)hipGetLastError(); return KMC_ERR_NO_DEVICE; }
        if (device < 0 || device >= n) return KMC_ERR_INVALID_ARGUMENT;
        t_device = device;
        t_device_set = true;
        return hipSetDevice(device) == hipSuccess ? KMC_OK : KMC_ERR_DEVICE;
    });
}

const char* kmc_status_string(kmc_status s)
{
    switch (s) {
    case KMC_OK: return "OK";
    case KMC_ERR_NULL_INPUT: return "Input object is NULL.";
    case KMC_ERR_NOT_STRUCTURED: return "Input object is not structured volume object.";
    case KMC_ERR_NOT_SCALAR: return "Input volume is not sclar field data.";
    case KMC_ERR_UNSUPPORTED_TYPE: return "Unsupported data type.";
    case KMC_ERR_INVALID_ARGUMENT: return "Invalid argument.";
    case KMC_ERR_OUT_OF_MEMORY: return "Cannot allocate memory.";
    case KMC_ERR_DEVICE: return "HIP runtime error.";
    case KMC_ERR_BUDGET: return "Request exceeds the memory budget.";
    case KMC_ERR_NOT_READY: return "Result not ready.";
    case KMC_ERR_NO_DEVICE: return "No HIP device available.";
    case KMC_ERR_SCRATCH: return "Work arrays exceed the fixed scratch of kmc_extract_noalloc.";
    case KMC_ERR_CAPACITY: return "Output arrays are too small.";
    case KMC_ERR_CELL_TYPE: return "Input volume is not tetrahedra, hexahedra, pyramid or prism-cell data.";
    case KMC_ERR_FILE: return "Cannot read or write the file.";
    case KMC_ERR_NOT_TRIANGLES: return "Input object is not triangle polygon.";
    }
    return "Unknown status.";
}

kmc_status kmc_extract(const kmc_volume* volume, double isolevel, int32_t normal_type, int32_t duplication,
                       int32_t out_memory, void* stream, kmc_surface* out)
{
    return kmc::abi_guard([&]() -> kmc_status {
        const kmc_status s = extract_levels_impl(volume, &isolevel, 1, normal_type, duplication, out_memory, stream, out);
        report(s);
        return s;
    });
}

kmc_status kmc_extract_levels(const kmc_volume* volume, const double* isolevels, uint32_t n_levels,
                              int32_t normal_type, int32_t duplication, int32_t out_memory, void* stream,
                              kmc_surface* outs)
{
    return kmc::abi_guard([&]() -> kmc_status {
        const kmc_status s = extract_levels_impl(volume, isolevels, n_levels, normal_type, duplication, out_memory,
                                                 stream, outs);
        report(s);
        return s;
    });
}

void kmc_surface_release(kmc_surface* s)
{
    if (!s) return;
    free_storage((SurfaceStorage*)s->internal);
    s->internal = nullptr;
    s->coords = nullptr;
    s->connections = nullptr;
    s->normals = nullptr;
    s->colors = nullptr;
    s->n_vertices = s->n_triangles = s->n_connections = s->n_normals = s->n_colors = 0;
}

void kmc_set_timing(int32_t enable) { t_timing = enable != 0; }

kmc_status kmc_last_timing(kmc_timing* out)
{
    return kmc::abi_guard([&]() -> kmc_status {
        if (!out) return KMC_ERR_INVALID_ARGUMENT;
        *out = t_last;
        return KMC_OK;
    });
}

kmc_status kmc_volume_min_max(const kmc_volume* v, void* stream, double* mn, double* mx)
{
    return kmc::abi_guard([&]() -> kmc_status {
        kmc_status s = validate_volume(v);
        if (s == KMC_OK && (!mn || !mx)) s = KMC_ERR_INVALID_ARGUMENT;
        if (s == KMC_OK) s = ensure_device();
        if (s != KMC_OK) { report(s); return s; }
        hipStream_t st = (hipStream_t)stream;
        const void* dvals = nullptr;
        void* owned = nullptr;
        s = device_values(v, st, &dvals, &owned);
        if (s == KMC_OK) s = min_max_on_device(v, dvals, st, mn, mx);
        if (owned) { (void)hipStreamSynchronize(st); device_free(owned); }
        report(s);
        return s;
    });
}

}  // extern "C"
