/*
 * kmc_server.hip -- the serving form: resident volumes, a request queue and
 * continuous batching over worker streams.
 *
 * A request is (resident volume, isolevel, normal type, duplication).  Worker
 * threads (each with a stream of its own, default priority) loop: take every
 * queued request that fits the free slots and the HBM budget, plan them as
 * one batch (the batched form's kernels, without a graph), run the front
 * passes (classify, count, scan, sizes), wait for their stream only, reserve
 * the exact output bytes, run the emit, wait, deliver.  While one worker waits
 * on the device the others admit and launch new batches, so requests that
 * arrive while others are on the GPU join the next batch and the batches
 * overlap on the device.
 *
 * Budget: masks, counts and prefixes of a request are reserved at admission;
 * its outputs once their sizes are known.  A request whose outputs do not fit
 * goes back to the front of the queue (deferred) and is redone when memory
 * frees up; a request that could not fit even alone fails with
 * KMC_ERR_BUDGET.  A batch never waits while holding reservations, so the
 * loop cannot deadlock.
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "kmc.h"
#include "kmc_guard.h"
#include "kmc_internal.h"

using namespace kmc;
using Clock = std::chrono::steady_clock;

namespace {

enum class State { Queued, Running, Done };

struct Request {
    uint32_t volume;
    double level;
    int32_t normal_type, duplication;
    uint64_t scratch;            // bytes reserved at admission
    Clock::time_point t_submit, t_done;
    State state = State::Queued;
    kmc_status status = KMC_OK;
    kmc_surface result;
};

struct Resident {
    kmc_volume v;                // device values, min/max set
    void* owned;
};

/* Device bytes of the masks, counts and prefixes of one single-level request. */
uint64_t request_scratch(const kmc_volume& v)
{
    /* per node row: counts and prefixes of vertices and triangles (2 x 12 B);
     * per 64-node chunk: its mask (8 B), its two in-row prefixes (8 B) and its
     * crossing record (32 B) */
    const uint64_t nx = v.resolution[0], ny = v.resolution[1], nz = v.resolution[2];
    const uint64_t nrows = ny * nz, NC = (nx + 63) / 64;
    return nrows * 2 * 12 + nrows * NC * 48 + 4096;
}

uint64_t output_bytes(const kmc_surface& s)
{
    return s.n_vertices * 12 + s.n_normals * 12 + s.n_connections * 4;
}

}  // namespace

struct kmc_server {
    kmc_server_config cfg;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<Resident> volumes;
    std::unordered_map<uint64_t, Request> reqs;
    std::deque<uint64_t> pending;
    uint64_t next_id = 1;
    uint64_t used = 0;           // budget bytes reserved
    uint32_t inflight = 0;       // requests admitted and not yet delivered
    uint32_t running = 0;        // batches between admission and delivery
    bool stop = false;
    bool trace = false;          // KMC_SERVER_TRACE: one line per batch on stderr
    std::vector<std::thread> workers;
    // statistics since the last reset
    kmc_server_stats st;
    std::vector<double> lat_ms;
    bool have_first = false;
    Clock::time_point first_submit, last_done;
};

namespace {

bool can_take(const kmc_server* S)
{
    if (S->pending.empty() || S->inflight >= S->cfg.max_inflight) return false;
    const Request& r = S->reqs.at(S->pending.front());
    return S->used + r.scratch <= S->cfg.budget_bytes;
}

void finish_locked(kmc_server* S, uint64_t id, Request& r, kmc_status status)
{
    r.status = status;
    r.state = State::Done;
    r.t_done = Clock::now();
    S->last_done = r.t_done;
    if (status == KMC_OK) {
        ++S->st.completed;
        S->st.triangles += r.result.n_triangles;
        S->lat_ms.push_back(std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count());
    } else {
        ++S->st.failed;
    }
    (void)id;
}

/* Copy a device surface to pinned host storage (and free the device arrays). */
kmc_status surface_to_host(kmc_surface* s, hipStream_t st)
{
    SurfaceStorage* d = (SurfaceStorage*)s->internal;
    SurfaceStorage* h = new SurfaceStorage();
    h->memory = KMC_MEMORY_HOST;
    h->coords = h->connections = h->normals = nullptr;
    const size_t bc = s->n_vertices * 12, bn = s->n_normals * 12, bk = s->n_connections * 4;
    bool ok = true;
    if (bc) ok = ok && (h->coords = host_block_alloc(bc)) != nullptr;
    if (bn) ok = ok && (h->normals = host_block_alloc(bn)) != nullptr;
    if (bk) ok = ok && (h->connections = host_block_alloc(bk)) != nullptr;
    if (ok && bc) ok = hipMemcpyAsync(h->coords, s->coords, bc, hipMemcpyDeviceToHost, st) == hipSuccess;
    if (ok && bn) ok = hipMemcpyAsync(h->normals, s->normals, bn, hipMemcpyDeviceToHost, st) == hipSuccess;
    if (ok && bk) ok = hipMemcpyAsync(h->connections, s->connections, bk, hipMemcpyDeviceToHost, st) == hipSuccess;
    if (ok) ok = hipStreamSynchronize(st) == hipSuccess;
    if (!ok) {
        /* no copy still in flight may write into blocks that go back to the cache */
        (void)hipStreamSynchronize(st);
        (void)hipGetLastError();
        free_storage(h);
        return KMC_ERR_OUT_OF_MEMORY;
    }
    free_storage(d);
    s->internal = h;
    s->memory = KMC_MEMORY_HOST;
    s->coords = (float*)h->coords;
    s->normals = (float*)h->normals;
    s->connections = (uint32_t*)h->connections;
    return KMC_OK;
}

void run_batch(kmc_server* S, hipStream_t st, const std::vector<uint64_t>& ids, bool& delivered)
{
    const size_t n = ids.size();
    /* requests on one volume with one emit mode become one job whose
     * isolevels are theirs in queue order: the volume is classified and
     * counted once for all of them (fused levels, as kmc_extract_levels);
     * surf[i] is request i's surface (job order, then level order) */
    std::vector<kmc_job> jobs;
    std::vector<std::vector<double>> levels;
    std::vector<size_t> job_of(n), pos_of(n), surf(n);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        std::vector<uint64_t> keys;
        for (size_t i = 0; i < n; ++i) {
            const Request& r = S->reqs.at(ids[i]);
            const int dup = r.normal_type == KMC_VERTEX_NORMAL ? 0 : (r.duplication ? 1 : 0);   // MarchingCubes.cpp:69-73
            const uint64_t key = (uint64_t)r.volume << 8 | (uint64_t)r.normal_type << 1 | (uint64_t)dup;
            size_t j = 0;
            while (j < keys.size() && keys[j] != key) ++j;
            if (j == keys.size()) {
                keys.push_back(key);
                kmc_job job{};
                job.volume = S->volumes[r.volume].v;
                job.normal_type = r.normal_type;
                job.duplication = dup;
                jobs.push_back(job);
                levels.emplace_back();
            }
            job_of[i] = j;
            pos_of[i] = levels[j].size();
            levels[j].push_back(r.level);
        }
    }
    std::vector<size_t> first(jobs.size(), 0);
    for (size_t j = 0, f = 0; j < jobs.size(); ++j) {
        jobs[j].isolevels = levels[j].data();
        jobs[j].n_levels = (uint32_t)levels[j].size();
        first[j] = f;
        f += levels[j].size();
    }
    for (size_t i = 0; i < n; ++i) surf[i] = first[job_of[i]] + pos_of[i];

    const Clock::time_point t0 = Clock::now();
    kmc_batch* b = nullptr;
    kmc_status s = batch_plan(jobs.data(), (uint32_t)jobs.size(), st, &b);
    const Clock::time_point t1 = Clock::now();
    if (s == KMC_OK && (batch_launch_front(b, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)) {
        (void)hipGetLastError();
        s = KMC_ERR_DEVICE;
    }
    const Clock::time_point t2 = Clock::now();
    Clock::time_point t3 = t2, t4 = t2;
    std::vector<kmc_surface> out(n);
    std::vector<uint64_t> obytes(n, 0);
    std::vector<char> admitted(n, 0);
    const bool front_ok = s == KMC_OK;
    if (s == KMC_OK) {
        for (size_t i = 0; i < n; ++i) {
            batch_surface_sizes(b, (uint32_t)surf[i], &out[i]);
            obytes[i] = output_bytes(out[i]) + batch_record_bytes(b, (uint32_t)surf[i]);
        }
        /* reserve outputs in queue order; the rest go back to the queue */
        std::lock_guard<std::mutex> lk(S->mu);
        bool defer_rest = false, any = false;
        for (size_t i = 0; i < n; ++i) {
            Request& r = S->reqs.at(ids[i]);
            if (r.scratch + obytes[i] > S->cfg.budget_bytes) {
                finish_locked(S, ids[i], r, KMC_ERR_BUDGET);      // cannot fit even alone
                S->used -= r.scratch;
                --S->inflight;
                continue;
            }
            /* the only batch on the device always makes progress with its first request */
            const bool alone = S->running == 1 && !any;
            if (!defer_rest && (S->used + obytes[i] <= S->cfg.budget_bytes || alone)) {
                admitted[i] = 1;
                any = true;
                S->used += obytes[i];
                S->st.peak_budget_bytes = std::max<uint64_t>(S->st.peak_budget_bytes, S->used);
            } else {
                defer_rest = true;
            }
        }
        /* deferred requests return to the front of the queue in their order */
        for (size_t k = n; k-- > 0;) {
            Request& r = S->reqs.at(ids[k]);
            if (admitted[k] || r.state == State::Done) continue;
            r.state = State::Queued;
            S->used -= r.scratch;
            --S->inflight;
            ++S->st.deferred;
            S->pending.push_front(ids[k]);
        }
    }
    if (s == KMC_OK) {
        for (size_t i = 0; i < n && s == KMC_OK; ++i) {
            if (!admitted[i]) { batch_set_outputs(b, (uint32_t)surf[i], nullptr, nullptr, nullptr); continue; }
            SurfaceStorage* store = new SurfaceStorage();
            store->memory = KMC_MEMORY_DEVICE;
            store->coords = store->connections = store->normals = nullptr;
            out[i].internal = store;
            out[i].memory = KMC_MEMORY_DEVICE;
            if (out[i].n_vertices) {
                store->coords = device_alloc(out[i].n_vertices * 12);
                store->normals = out[i].n_normals ? device_alloc(out[i].n_normals * 12) : nullptr;
                store->connections = out[i].n_connections ? device_alloc(out[i].n_connections * 4) : nullptr;
                if (!store->coords || (out[i].n_normals && !store->normals) ||
                    (out[i].n_connections && !store->connections))
                    s = KMC_ERR_OUT_OF_MEMORY;
            }
            out[i].coords = (float*)store->coords;
            out[i].normals = (float*)store->normals;
            out[i].connections = (uint32_t*)store->connections;
            batch_set_outputs(b, (uint32_t)surf[i], out[i].coords, out[i].normals, out[i].connections);
        }
        t3 = Clock::now();
        if (s == KMC_OK && (batch_launch_emit(b, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)) {
            (void)hipGetLastError();
            s = KMC_ERR_DEVICE;
        }
        if (s == KMC_OK) {
            /* what this batch holds on the device: its work arrays, the outputs
             * of its admitted requests and the records of their emit */
            uint64_t held = batch_work_bytes(b) + batch_record_arena_bytes(b);
            for (size_t i = 0; i < n; ++i) held += admitted[i] ? output_bytes(out[i]) : 0;
            std::lock_guard<std::mutex> lk(S->mu);
            S->st.peak_batch_device_bytes = std::max<uint64_t>(S->st.peak_batch_device_bytes, held);
        }
        if (s == KMC_OK && S->cfg.out_memory == KMC_MEMORY_HOST)
            for (size_t i = 0; i < n && s == KMC_OK; ++i)
                if (admitted[i]) s = surface_to_host(&out[i], st);
    }
    t4 = Clock::now();

    /* deliver before the batch's arrays go back to the pool: a free may
     * release cached blocks, and hipFree waits for the whole device */
    {
        std::lock_guard<std::mutex> lk(S->mu);
        for (size_t i = 0; i < n; ++i) {
            /* after a good front phase only admitted requests are still this
             * batch's (the others failed early or went back to the queue) */
            if (front_ok && !admitted[i]) continue;
            Request& r = S->reqs.at(ids[i]);
            if (s != KMC_OK) {
                if (admitted[i]) kmc_surface_release(&out[i]);
                finish_locked(S, ids[i], r, s);
            } else {
                r.result = out[i];
                finish_locked(S, ids[i], r, KMC_OK);
            }
            S->used -= r.scratch + (admitted[i] ? obytes[i] : 0);
            --S->inflight;
        }
        --S->running;
        ++S->st.batches;
        delivered = true;
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();

    const Clock::time_point t5 = Clock::now();
    if (b) batch_destroy(b);
    if (S->trace) {
        auto ms = [](Clock::time_point a, Clock::time_point c) {
            return std::chrono::duration<double, std::milli>(c - a).count();
        };
        size_t na = 0;
        uint64_t ob = 0;
        for (size_t i = 0; i < n; ++i) { na += admitted[i]; ob += admitted[i] ? obytes[i] : 0; }
        std::fprintf(stderr, "[kmc_server] batch %zu req %zu admitted %.1f MB out: plan %.3f front %.3f alloc %.3f "
                     "emit %.3f deliver %.3f destroy %.3f ms\n", n, na, ob / 1e6, ms(t0, t1), ms(t1, t2), ms(t2, t3),
                     ms(t3, t4), ms(t4, t5), ms(t5, Clock::now()));
    }
}

void worker_main(kmc_server* S)
{
    kmc_init(S->cfg.device);
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        st = nullptr;
    }
    for (;;) {
        std::vector<uint64_t> take;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return (S->stop && S->pending.empty()) || can_take(S); });
            if (S->stop && S->pending.empty()) break;
            while (!S->pending.empty() && take.size() < S->cfg.max_batch && S->inflight < S->cfg.max_inflight) {
                Request& r = S->reqs.at(S->pending.front());
                if (S->used + r.scratch > S->cfg.budget_bytes) break;
                S->used += r.scratch;
                ++S->inflight;
                r.state = State::Running;
                take.push_back(S->pending.front());
                S->pending.pop_front();
            }
            if (take.empty()) continue;
            ++S->running;
            S->st.peak_inflight = std::max(S->st.peak_inflight, S->inflight);
            S->st.peak_budget_bytes = std::max<uint64_t>(S->st.peak_budget_bytes, S->used);
        }
        bool delivered = false;
        try {
            run_batch(S, st, take, delivered);
        } catch (...) {
            /* host memory ran out inside the batch: before delivery its
             * requests that still run fail, so that no waiter waits for them
             * forever (their device blocks and output budget stay lost) */
            if (!delivered) {
                std::lock_guard<std::mutex> lk(S->mu);
                for (uint64_t id : take) {
                    auto it = S->reqs.find(id);
                    if (it == S->reqs.end() || it->second.state != State::Running) continue;
                    it->second.status = KMC_ERR_OUT_OF_MEMORY;
                    it->second.state = State::Done;
                    ++S->st.failed;
                    S->used -= std::min(S->used, it->second.scratch);
                    --S->inflight;
                }
                --S->running;
            }
            S->cv_done.notify_all();
            S->cv_work.notify_all();
        }
    }
    if (st) (void)hipStreamDestroy(st);
}

kmc_status take_result(kmc_server* S, uint64_t id, kmc_surface* out)
{
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return KMC_ERR_INVALID_ARGUMENT;
    Request& r = it->second;
    if (r.state != State::Done) return KMC_ERR_NOT_READY;
    const kmc_status s = r.status;
    if (s == KMC_OK) {
        if (out) *out = r.result;
        else kmc_surface_release(&r.result);
    }
    S->reqs.erase(it);
    return s;
}

}  // namespace

extern "C" {

kmc_status kmc_server_create(const kmc_server_config* config, kmc_server** out)
{
    if (!out) return KMC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    kmc_server_config cfg{};
    if (config) cfg = *config;
    if (cfg.workers == 0) cfg.workers = 3;
    if (cfg.max_inflight == 0) cfg.max_inflight = 256;
    if (cfg.max_batch == 0) cfg.max_batch = 64;
    if (cfg.budget_bytes == 0) cfg.budget_bytes = uint64_t(16) << 30;
    if (cfg.out_memory != KMC_MEMORY_HOST && cfg.out_memory != KMC_MEMORY_DEVICE) return KMC_ERR_INVALID_ARGUMENT;
    kmc_status s = kmc_init(cfg.device);
    if (s != KMC_OK) { report(s); return s; }
    kmc_server* S = new (std::nothrow) kmc_server();
    if (!S) return KMC_ERR_OUT_OF_MEMORY;
    S->cfg = cfg;
    S->trace = std::getenv("KMC_SERVER_TRACE") != nullptr;
    std::memset(&S->st, 0, sizeof(S->st));
    try {
        for (uint32_t i = 0; i < cfg.workers; ++i) S->workers.emplace_back(worker_main, S);
    } catch (...) {
        /* a worker could not be started: stop the ones that were */
        kmc_server_destroy(S);
        return KMC_ERR_OUT_OF_MEMORY;
    }
    *out = S;
    return KMC_OK;
}

kmc_status kmc_server_register_volume(kmc_server* S, const kmc_volume* v, uint32_t* volume_id)
{
    return kmc::abi_guard([&]() -> kmc_status {
        if (!S || !volume_id) return KMC_ERR_INVALID_ARGUMENT;
        kmc_status s = validate_volume(v);
        if (s == KMC_OK && levels_per_pass(v, 1) == 0) s = KMC_ERR_INVALID_ARGUMENT;   // rows too long for a tile
        if (s == KMC_OK) s = kmc_init(S->cfg.device);
        if (s != KMC_OK) { report(s); return s; }
        Resident r;
        r.v = *v;
        const void* dvals = nullptr;
        r.owned = nullptr;
        s = device_values(v, nullptr, &dvals, &r.owned);
        if (s == KMC_OK && !v->has_min_max_values) {
            double mn = 0, mx = 0;
            s = min_max_on_device(v, dvals, nullptr, &mn, &mx);
            r.v.min_value = mn;
            r.v.max_value = mx;
        }
        if (s == KMC_OK && hipStreamSynchronize(nullptr) != hipSuccess) s = KMC_ERR_DEVICE;
        if (s != KMC_OK) { device_free(r.owned); report(s); return s; }
        r.v.values = dvals;
        r.v.memory = KMC_MEMORY_DEVICE;
        r.v.has_min_max_values = 1;
        std::lock_guard<std::mutex> lk(S->mu);
        *volume_id = (uint32_t)S->volumes.size();
        S->volumes.push_back(r);
        return KMC_OK;
    });
}

kmc_status kmc_server_submit(kmc_server* S, uint32_t volume_id, double isolevel, int32_t normal_type,
                             int32_t duplication, uint64_t* request_id)
{
    return kmc::abi_guard([&]() -> kmc_status {
        if (!S || !request_id) return KMC_ERR_INVALID_ARGUMENT;
        if (normal_type != KMC_VERTEX_NORMAL && normal_type != KMC_POLYGON_NORMAL) return KMC_ERR_INVALID_ARGUMENT;
        std::lock_guard<std::mutex> lk(S->mu);
        if (volume_id >= S->volumes.size() || S->stop) return KMC_ERR_INVALID_ARGUMENT;
        Request r;
        r.volume = volume_id;
        r.level = isolevel;
        r.normal_type = normal_type;
        r.duplication = duplication ? 1 : 0;
        r.scratch = request_scratch(S->volumes[volume_id].v);
        r.t_submit = Clock::now();
        std::memset(&r.result, 0, sizeof(r.result));
        if (r.scratch > S->cfg.budget_bytes) return KMC_ERR_BUDGET;
        const uint64_t id = S->next_id++;
        if (!S->have_first) { S->have_first = true; S->first_submit = r.t_submit; }
        ++S->st.submitted;
        S->reqs.emplace(id, r);
        S->pending.push_back(id);
        *request_id = id;
        S->cv_work.notify_one();
        return KMC_OK;
    });
}

kmc_status kmc_server_poll(kmc_server* S, uint64_t id, kmc_surface* out)
{
    return kmc::abi_guard([&]() -> kmc_status {
        if (!S) return KMC_ERR_INVALID_ARGUMENT;
        std::lock_guard<std::mutex> lk(S->mu);
        return take_result(S, id, out);
    });
}

kmc_status kmc_server_wait(kmc_server* S, uint64_t id, int64_t timeout_ms, kmc_surface* out)
{
    return kmc::abi_guard([&]() -> kmc_status {
        if (!S) return KMC_ERR_INVALID_ARGUMENT;
        std::unique_lock<std::mutex> lk(S->mu);
        auto ready = [&] {
            auto it = S->reqs.find(id);
            return it == S->reqs.end() || it->second.state == State::Done;
        };
        if (timeout_ms < 0) S->cv_done.wait(lk, ready);
        else S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
        return take_result(S, id, out);
    });
}

kmc_status kmc_server_get_stats(kmc_server* S, kmc_server_stats* out)
{
    return kmc::abi_guard([&]() -> kmc_status {
        if (!S || !out) return KMC_ERR_INVALID_ARGUMENT;
        std::lock_guard<std::mutex> lk(S->mu);
        kmc_server_stats st = S->st;
        std::vector<double> l = S->lat_ms;
        if (!l.empty()) {
            std::sort(l.begin(), l.end());
            auto pct = [&](double p) {
                const size_t k = (size_t)std::min<double>((double)l.size() - 1, std::ceil(p * l.size()) - 1);
                return l[k];
            };
            st.p50_ms = pct(0.50);
            st.p99_ms = pct(0.99);
            st.max_ms = l.back();
            double sum = 0;
            for (double x : l) sum += x;
            st.mean_ms = sum / l.size();
        }
        if (S->have_first && st.completed > 0) {
            st.span_s = std::chrono::duration<double>(S->last_done - S->first_submit).count();
            st.triangles_per_s = st.span_s > 0 ? st.triangles / st.span_s : 0.0;
        }
        *out = st;
        return KMC_OK;
    });
}

kmc_status kmc_server_reset_stats(kmc_server* S)
{
    return kmc::abi_guard([&]() -> kmc_status {
        if (!S) return KMC_ERR_INVALID_ARGUMENT;
        std::lock_guard<std::mutex> lk(S->mu);
        std::memset(&S->st, 0, sizeof(S->st));
        S->lat_ms.clear();
        S->have_first = false;
        return KMC_OK;
    });
}

void kmc_server_destroy(kmc_server* S)
{
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (std::thread& t : S->workers) t.join();
    for (auto& kv : S->reqs)
        if (kv.second.state == State::Done && kv.second.status == KMC_OK) kmc_surface_release(&kv.second.result);
    (void)hipSetDevice(S->cfg.device);
    for (Resident& r : S->volumes) device_free(r.owned);
    delete S;
}

}  // extern "C"
