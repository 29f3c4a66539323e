// serve.cpp -- continuous-batching server over mdiscr_partition_batch.
//
// One scheduler thread owns nstreams slots (context + stream + events).  Each
// loop it (1) retires finished batches (their class counts come back in one
// pinned copy per batch), (2) admits queued requests in FIFO order while their
// output buffers fit the budget, and (3) hands every admitted request to the
// next free slot as one batched launch.  Output buffers come from a
// size-class cache carved from 1 GiB arena chunks, so steady-state admission
// calls neither hipMalloc nor hipFree (hipFree would wait for the device and
// stall the overlapping batches).  Kernels write class counts straight into
// mapped pinned memory, so a finished batch needs no copy back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/mdiscr.h"
#include "api_internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    const uint32_t* d_in = nullptr;
    uint64_t n = 0;
    mdiscr_discr d{};
    uint32_t* d_out = nullptr;
    size_t out_bytes = 0;      // size-class bytes held against the budget
    uint32_t h_counts[MDISCR_MAX_CLASSES] = {0};
    int state = 0;             // 0 queued, 1 admitted, 2 running, 3 done, 4 failed
    mdiscr_status err = MDISCR_OK;
    Clock::time_point t_submit, t_done;
};

struct Slot {
    mdiscr_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    uint32_t* h_counts = nullptr;  // mapped pinned, max_batch_requests * 64
    uint32_t* d_counts = nullptr;  // its device-side address (kernels write here)
    std::vector<uint64_t> tickets;
    bool busy = false;
};

size_t size_class(uint64_t bytes) {
    size_t c = 4096;
    while (c < bytes) c <<= 1;
    return c;
}

}  // namespace

struct mdiscr_server {
    mdiscr_server_config cfg{};
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::unordered_map<uint64_t, Request> reqs;
    std::deque<uint64_t> queued;     // waiting for budget
    std::deque<uint64_t> admitted;   // waiting for a slot
    std::map<size_t, std::vector<void*>> free_blocks;
    size_t held = 0, in_use = 0, peak = 0;
    uint64_t next_ticket = 1, batches = 0, batched_reqs = 0, completed = 0, elems = 0;
    uint64_t cancelled = 0;
    std::vector<double> latencies;
    bool stop = false;
    std::thread worker;

    // Output memory: size-class blocks carved from large arena chunks (one
    // hipMalloc per chunk), recycled through per-class free lists.  A block
    // is charged to the budget while it holds a request's output.
    static constexpr size_t kChunk = size_t(1) << 30;
    std::vector<std::pair<char*, size_t>> chunks;  // base, bytes used

    void* alloc_block(size_t cls) {
        if (in_use + cls > cfg.budget_bytes) return nullptr;
        auto& fl = free_blocks[cls];
        if (!fl.empty()) {
            void* p = fl.back();
            fl.pop_back();
            in_use += cls;
            peak = std::max(peak, in_use);
            return p;
        }
        // first chunk with room (reserved chunks fill in order)
        while (cur_chunk < chunks.size() && chunks[cur_chunk].second + cls > chunk_cap[cur_chunk])
            ++cur_chunk;
        if (cur_chunk == chunks.size()) {
            const size_t csize = std::max(kChunk, cls);
            if (held + csize > cfg.budget_bytes + kChunk) return nullptr;
            if (!map_chunk(csize)) return nullptr;
        }
        void* p = chunks[cur_chunk].first + chunks[cur_chunk].second;
        chunks[cur_chunk].second += cls;
        in_use += cls;
        peak = std::max(peak, in_use);
        return p;
    }
    size_t cur_chunk = 0;
    bool map_chunk(size_t csize) {
        void* p = nullptr;
        if (hipMalloc(&p, csize) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        chunks.emplace_back(static_cast<char*>(p), 0);
        chunk_cap.push_back(csize);
        held += csize;
        return true;
    }
    std::vector<size_t> chunk_cap;
    void free_block(void* p, size_t cls) {
        free_blocks[cls].push_back(p);
        in_use -= cls;
    }

    // Retire a slot's batch (mu held): completed, or failed when the device
    // reported an error for it, so that no waiter blocks on a batch that will
    // never signal.
    void retire(Slot& sl, mdiscr_status err) {
        const Clock::time_point now = Clock::now();
        for (size_t i = 0; i < sl.tickets.size(); ++i) {
            auto it = reqs.find(sl.tickets[i]);
            if (it == reqs.end()) continue;
            Request& r = it->second;
            r.t_done = now;
            if (err != MDISCR_OK) {
                r.state = 4;
                r.err = err;
                continue;
            }
            std::copy(sl.h_counts + i * MDISCR_MAX_CLASSES,
                      sl.h_counts + i * MDISCR_MAX_CLASSES + r.d.m, r.h_counts);
            r.state = 3;
            latencies.push_back(std::chrono::duration<double>(r.t_done - r.t_submit).count());
            ++completed;
            elems += r.n;
        }
        sl.tickets.clear();
        sl.busy = false;
    }

    void loop() {
        std::vector<mdiscr_request> batch;
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            bool progressed = false;
            // 1. retire finished batches
            for (auto& sl : slots) {
                if (!sl.busy) continue;
                const hipError_t q = hipEventQuery(sl.done);
                if (q == hipErrorNotReady) continue;
                if (q != hipSuccess) (void)hipGetLastError();
                retire(sl, q == hipSuccess ? MDISCR_OK : MDISCR_EDEVICE);
                progressed = true;
            }
            if (progressed) cv_done.notify_all();
            // 2. admit in FIFO order while outputs fit the budget
            while (!queued.empty()) {
                Request& r = reqs[queued.front()];
                const size_t cls = size_class(r.n * 4);
                void* p = alloc_block(cls);
                if (!p) break;
                r.d_out = static_cast<uint32_t*>(p);
                r.out_bytes = cls;
                r.state = 1;
                admitted.push_back(queued.front());
                queued.pop_front();
            }
            // 3. one batch per free slot, launched with mu released so that
            // submit, wait and release never block on a launch.  The slot is
            // claimed (busy) first; only this thread touches slots.
            for (auto& sl : slots) {
                if (sl.busy || admitted.empty()) continue;
                batch.clear();
                uint64_t belems = 0;
                while (!admitted.empty() && batch.size() < cfg.max_batch_requests) {
                    Request& r = reqs[admitted.front()];
                    if (!batch.empty() && belems + r.n > cfg.max_batch_elems) break;
                    const size_t i = batch.size();
                    batch.push_back(mdiscr_request{r.d_in, r.d_out, r.n, r.d,
                                                   sl.d_counts + i * MDISCR_MAX_CLASSES});
                    sl.tickets.push_back(admitted.front());
                    r.state = 2;
                    belems += r.n;
                    admitted.pop_front();
                }
                sl.busy = true;
                lk.unlock();
                mdiscr_status st = mdiscr_partition_batch(sl.ctx, batch.data(),
                                                          (uint32_t)batch.size(), sl.stream);
                if (st == MDISCR_OK)
                    st = hipEventRecord(sl.done, sl.stream) == hipSuccess ? MDISCR_OK
                                                                          : MDISCR_EDEVICE;
                lk.lock();
                if (st != MDISCR_OK) {
                    for (uint64_t t : sl.tickets) {
                        auto it = reqs.find(t);
                        if (it == reqs.end()) continue;
                        it->second.state = 4;
                        it->second.err = st;
                    }
                    sl.tickets.clear();
                    sl.busy = false;
                    cv_done.notify_all();
                    continue;
                }
                ++batches;
                batched_reqs += batch.size();
                progressed = true;
            }
            if (stop) {
                bool any_busy = false;
                for (auto& sl : slots) any_busy |= sl.busy;
                if (!any_busy && admitted.empty()) break;
            }
            if (!progressed) {
                bool any_busy = false;
                for (auto& sl : slots) any_busy |= sl.busy;
                // Poll running batches briefly; sleep until new work otherwise.
                if (any_busy)
                    cv_work.wait_for(lk, std::chrono::microseconds(2));
                else
                    cv_work.wait_for(lk, std::chrono::milliseconds(5));
            }
        }
    }
};

extern "C" {

mdiscr_status mdiscr_server_create(mdiscr_server** out, const mdiscr_server_config* cfg) {
    if (!out || !cfg || cfg->nstreams == 0 || cfg->nstreams > 16 || cfg->max_batch_requests == 0 ||
        cfg->max_batch_elems == 0 || cfg->budget_bytes == 0)
        return MDISCR_EINVAL;
    *out = nullptr;
    mdiscr_server* s = new (std::nothrow) mdiscr_server();
    if (!s) return MDISCR_ENOMEM;
    s->cfg = *cfg;
    if (s->cfg.max_pending == 0) s->cfg.max_pending = 1u << 20;
    if (hipSetDevice(cfg->device) != hipSuccess) {
        delete s;
        return MDISCR_EDEVICE;
    }
    s->slots.resize(cfg->nstreams);
    for (auto& sl : s->slots) {
        mdiscr_status st = mdiscr_ctx_create(&sl.ctx, cfg->device);
        const size_t cb = (size_t)cfg->max_batch_requests * MDISCR_MAX_CLASSES * 4;
        if (st != MDISCR_OK || hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void**)&sl.h_counts, cb, hipHostMallocMapped) != hipSuccess ||
            hipHostGetDevicePointer((void**)&sl.d_counts, sl.h_counts, 0) != hipSuccess) {
            (void)hipGetLastError();
            mdiscr_server_destroy(s);
            return MDISCR_EDEVICE;
        }
        // Every batch fits the slot's workspace and staging from the start, so
        // no launch grows a buffer while the other slots' batches run.
        const uint64_t max_elems = std::min<uint64_t>(cfg->max_batch_elems, MDISCR_MAX_ELEMS);
        if (ctx_reserve_batch(sl.ctx, max_elems, cfg->max_batch_requests) != MDISCR_OK) {
            mdiscr_server_destroy(s);
            return MDISCR_ENOMEM;
        }
    }
    for (uint64_t r = 0; r + mdiscr_server::kChunk <= std::min(cfg->reserve_bytes, cfg->budget_bytes);
         r += mdiscr_server::kChunk)
        if (!s->map_chunk(mdiscr_server::kChunk)) {
            mdiscr_server_destroy(s);
            return MDISCR_ENOMEM;
        }
    s->worker = std::thread([s] {
        (void)hipSetDevice(s->cfg.device);
        s->loop();
    });
    *out = s;
    return MDISCR_OK;
}

mdiscr_status mdiscr_server_destroy(mdiscr_server* s) {
    if (!s) return MDISCR_EINVAL;
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->stop = true;
            // queued requests that never got budget are abandoned
            for (uint64_t t : s->queued) s->reqs[t].state = 4;
            s->queued.clear();
        }
        s->cv_work.notify_all();
        s->worker.join();
    }
    (void)hipSetDevice(s->cfg.device);
    (void)hipDeviceSynchronize();
    for (auto& c : s->chunks) (void)hipFree(c.first);
    for (auto& sl : s->slots) {
        if (sl.ctx) mdiscr_ctx_destroy(sl.ctx);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.h_counts) (void)hipHostFree(sl.h_counts);
    }
    delete s;
    return MDISCR_OK;
}

mdiscr_status mdiscr_server_submit(mdiscr_server* s, const uint32_t* d_in, uint64_t n,
                                   mdiscr_discr d, uint64_t* ticket) {
    if (!s || !ticket) return MDISCR_EINVAL;
    if (d.m == 0 || d.m > MDISCR_MAX_CLASSES || (d.kind == MDISCR_DISCR_MASK && (d.m & (d.m - 1))) ||
        (d.kind != MDISCR_DISCR_MASK && d.kind != MDISCR_DISCR_MOD) || (n && !d_in) ||
        n > 0xFFFFFFFFull)
        return MDISCR_EINVAL;
    if (size_class(n * 4) > s->cfg.budget_bytes) return MDISCR_ENOMEM;  // could never fit
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->queued.size() + s->admitted.size() >= s->cfg.max_pending) return MDISCR_EBUSY;
        const uint64_t t = s->next_ticket++;
        Request& r = s->reqs[t];
        r.d_in = d_in;
        r.n = n;
        r.d = d;
        r.t_submit = Clock::now();
        s->queued.push_back(t);
        *ticket = t;
    }
    s->cv_work.notify_one();
    return MDISCR_OK;
}

mdiscr_status mdiscr_server_poll(mdiscr_server* s, uint64_t ticket, int* done) {
    if (!s || !done) return MDISCR_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return MDISCR_EINVAL;
    *done = it->second.state >= 3;
    return it->second.state == 4 ? it->second.err : MDISCR_OK;
}

mdiscr_status mdiscr_server_wait(mdiscr_server* s, uint64_t ticket, uint32_t** d_out,
                                 uint32_t* h_counts, double* latency_s) {
    if (!s) return MDISCR_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    // The entry is looked up again after every wake: other threads insert
    // (rehashing the map) and a cancel may remove it meanwhile.
    Request* rp = nullptr;
    s->cv_done.wait(lk, [&] {
        auto it = s->reqs.find(ticket);
        rp = it == s->reqs.end() ? nullptr : &it->second;
        return !rp || rp->state >= 3;
    });
    if (!rp) return MDISCR_EINVAL;  // unknown, or cancelled while waiting
    Request& r = *rp;
    if (r.state == 4) return r.err ? r.err : MDISCR_EDEVICE;
    if (d_out) *d_out = r.d_out;
    if (h_counts) std::copy(r.h_counts, r.h_counts + r.d.m, h_counts);
    if (latency_s) *latency_s = std::chrono::duration<double>(r.t_done - r.t_submit).count();
    return MDISCR_OK;
}

mdiscr_status mdiscr_server_cancel(mdiscr_server* s, uint64_t ticket) {
    if (!s) return MDISCR_EINVAL;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return MDISCR_EINVAL;
        Request& r = it->second;
        if (r.state == 0) {
            s->queued.erase(std::find(s->queued.begin(), s->queued.end(), ticket));
        } else if (r.state == 1) {
            s->admitted.erase(std::find(s->admitted.begin(), s->admitted.end(), ticket));
            s->free_block(r.d_out, r.out_bytes);
        } else {
            return MDISCR_EBUSY;  // on the GPU or finished: wait and release instead
        }
        s->reqs.erase(it);
        ++s->cancelled;
    }
    s->cv_done.notify_all();  // waiters on the ticket return
    s->cv_work.notify_one();  // freed budget may admit the next request
    return MDISCR_OK;
}

mdiscr_status mdiscr_server_release(mdiscr_server* s, uint64_t ticket) {
    if (!s) return MDISCR_EINVAL;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end() || (it->second.state != 3 && it->second.state != 4))
            return MDISCR_EINVAL;
        if (it->second.d_out) s->free_block(it->second.d_out, it->second.out_bytes);
        s->reqs.erase(it);
    }
    s->cv_work.notify_one();
    return MDISCR_OK;
}

mdiscr_status mdiscr_server_stats_get(mdiscr_server* s, mdiscr_server_stats* out) {
    if (!s || !out) return MDISCR_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    std::vector<double> v = s->latencies;
    std::sort(v.begin(), v.end());
    auto q = [&](double p) {
        if (v.empty()) return 0.0;
        size_t i = (size_t)(p * (double)(v.size() - 1) + 0.5);
        return v[std::min(i, v.size() - 1)];
    };
    double sum = 0;
    for (double x : v) sum += x;
    out->submitted = s->next_ticket - 1;
    out->completed = s->completed;
    out->batches = s->batches;
    out->elems = s->elems;
    out->p50_s = q(0.50);
    out->p99_s = q(0.99);
    out->mean_s = v.empty() ? 0.0 : sum / (double)v.size();
    out->max_s = v.empty() ? 0.0 : v.back();
    out->peak_bytes = s->peak;
    out->mean_batch_requests = s->batches ? (double)s->batched_reqs / (double)s->batches : 0.0;
    return MDISCR_OK;
}

}  // extern "C"
