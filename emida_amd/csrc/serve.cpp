// Request server: continuous batching of single-pattern requests over resident
// pattern-offset pipelines ("plans").  No counterpart in the reference, whose only
// driver (src/process_files.hpp) runs one fixed configuration over a file list.
//
// A plan is one emida_offset configuration (reference pattern, ROI set, slice size,
// policy) whose reference slices or spectra stay on the device.  Requests carry one
// uint16 pattern each and are copied into pinned staging buffers at submission.
// The scheduler thread forms a batch for a plan from whatever requests of that plan
// are queued (up to max_batch) as soon as the plan has a free slot, i.e. a device
// buffer set for one batch (patterns, working memory, results); new slots are created
// while the sum of slot sizes stays within the HBM budget, so admission is by
// memory; when other plans' idle slots fill the budget, the one idle longest is freed
// for a plan that has requests waiting.  Batches go round-robin over several streams and overlap on the device.
// Each (slot, batch size) pipeline - kernels plus the copies of the results into the
// slot's pinned buffers - is captured once in a hipGraph and replayed; batch sizes
// are rounded up to a power of two or 3/2 of one for this (graph_batch), the extra
// slices computing on stale slot contents whose results nobody reads.  One
// completion thread per stream waits for its batches in order, hands results to the
// requests and frees the slot and the staging buffers.
#include <chrono>
#include <deque>
#include <map>
#include <thread>
#include <unordered_map>

#include "offset.hpp"

namespace {

// Batch size a graph is captured for: the smallest of 1, 2, 3, 4, 6, 8, 12, 16, 24, ...
// (powers of two and 3/2 of them) that holds n, so a replay computes at most a third
// more slices than it needs (a power-of-two grid wastes up to all of them again) with
// two graphs per octave of sizes.
uint32_t graph_batch(uint32_t n) {
    uint32_t p = 1;
    while (p < n) {
        if (p >= 2 && p / 2 * 3 >= n) return p / 2 * 3;
        p *= 2;
    }
    return p;
}

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

struct emida_server_s {
    struct request {
        uint64_t id = 0;
        uint32_t plan = 0;
        void* staging = nullptr;
        double* out_offsets = nullptr;
        double* out_coefs = nullptr;
        uint64_t submit_ns = 0, launch_ns = 0, done_ns = 0;
        uint32_t batch = 0;
        bool done = false;
    };
    struct slot {
        uint32_t plan = 0;
        void* block = nullptr;  // device: patterns | scratch | maxes | neigh | offsets | coefs
        void* pics = nullptr;
        emida_offset_s::scratch_t scratch;
        void* maxes = nullptr;
        void* neigh = nullptr;
        double* offsets = nullptr;
        double* coefs = nullptr;
        double* h_offsets = nullptr;  // pinned
        double* h_coefs = nullptr;
        hipEvent_t done = nullptr;
        size_t bytes = 0;
        std::map<uint32_t, hipGraphExec_t> graphs;  // by captured batch size (graph_batch)
    };
    struct plan_rt {
        std::unique_ptr<emida_offset_s> off;
        uint32_t max_batch = 1;
        size_t pat_bytes = 0;
        size_t slot_bytes = 0;
        std::deque<request*> queue;
        std::vector<slot*> free_slots;
        uint32_t n_slots = 0;
    };
    struct launch_rec {
        slot* sl;
        std::vector<request*> reqs;
    };

    emida_server_config cfg{};
    int device = 0;
    std::vector<std::unique_ptr<plan_rt>> plans;
    std::vector<std::unique_ptr<slot>> slots;
    std::vector<hipStream_t> streams;
    hipStream_t capture_stream = nullptr;
    std::vector<std::deque<launch_rec>> inflight;
    std::vector<std::condition_variable> cv_inflight;
    std::unordered_map<uint64_t, request*> requests;
    std::map<size_t, std::vector<void*>> staging_free;  // by size
    std::vector<void*> staging_all;
    uint32_t staging_used = 0;
    uint64_t next_id = 1;
    uint32_t next_stream = 0;
    size_t slot_bytes_total = 0;
    size_t inflight_bytes = 0;
    emida_server_stats st{};
    bool stop = false;
    std::string worker_error;

    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_staging;
    std::thread scheduler;
    std::vector<std::thread> completers;

    // ------------------------------------------------------------ lifetime
    void start() {
        EMIDA_HIP(hipGetDevice(&device));
        streams.resize(cfg.streams);
        for (auto& s : streams) EMIDA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        EMIDA_HIP(hipStreamCreateWithFlags(&capture_stream, hipStreamNonBlocking));
        inflight.resize(cfg.streams);
        cv_inflight = std::vector<std::condition_variable>(cfg.streams);
        scheduler = std::thread([this] { guarded_loop([this] { schedule_loop(); }); });
        for (uint32_t k = 0; k < cfg.streams; ++k)
            completers.emplace_back([this, k] { guarded_loop([this, k] { complete_loop(k); }); });
    }

    // A worker that throws records the message, stops the server and wakes every waiter.
    template <class F>
    void guarded_loop(F&& f) {
        (void)hipSetDevice(device);
        try {
            f();
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> lock(mu);
            if (worker_error.empty()) worker_error = e.what();
            stop = true;
        }
        cv_sched.notify_all();
        cv_done.notify_all();
        cv_staging.notify_all();
        for (auto& c : cv_inflight) c.notify_all();
    }

    void shutdown() {
        {
            std::unique_lock<std::mutex> lock(mu);
            // drain: every submitted request completes first
            cv_done.wait(lock, [&] { return !worker_error.empty() || st.completed == st.submitted; });
            stop = true;
        }
        cv_sched.notify_all();
        for (auto& c : cv_inflight) c.notify_all();
        if (scheduler.joinable()) scheduler.join();
        for (auto& t : completers)
            if (t.joinable()) t.join();
        (void)hipSetDevice(device);
        for (auto& s : streams) (void)hipStreamSynchronize(s);
        for (auto& sl : slots) release_slot(*sl);
        for (void* p : staging_all) (void)hipHostFree(p);
        for (auto& kv : requests) delete kv.second;
        for (auto& s : streams) (void)hipStreamDestroy(s);
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
        plans.clear();
    }

    void check_running() {
        if (!worker_error.empty()) throw state_error("server stopped: " + worker_error);
        if (stop) throw state_error("server is shutting down");
    }

    // ------------------------------------------------------------ plans
    uint32_t add_plan(std::unique_ptr<emida_offset_s> off, const uint16_t* ref) {
        off->allocate(ref, true);
        auto p = std::make_unique<plan_rt>();
        p->pat_bytes = (size_t)off->src.area() * sizeof(uint16_t);
        p->max_batch = cfg.max_batch;
        auto slot_bytes = [&](uint32_t n) {
            const size_t tn = (size_t)off->nb * n, nk = (size_t)off->fit * off->fit;
            return emida_offset_s::al256(n * p->pat_bytes) + off->scratch_bytes(n) +
                   emida_offset_s::al256(tn * off->ti_bytes()) + emida_offset_s::al256(tn * nk * off->t_bytes()) +
                   emida_offset_s::al256(tn * 2 * sizeof(double)) + emida_offset_s::al256(tn * 6 * sizeof(double));
        };
        // a plan's batches stay within a quarter of the budget
        while (p->max_batch > 1 && slot_bytes(p->max_batch) > cfg.hbm_budget / 4) p->max_batch /= 2;
        p->slot_bytes = slot_bytes(p->max_batch);
        if (p->slot_bytes > cfg.hbm_budget)
            throw arg_error("one pattern of this plan needs more device memory than the server's budget");
        p->off = std::move(off);
        std::lock_guard<std::mutex> lock(mu);
        check_running();
        plans.push_back(std::move(p));
        st.device_bytes_resident += plans.back()->off->device_bytes;
        return (uint32_t)plans.size() - 1;
    }

    static void release_slot(slot& sl) {
        for (auto& g : sl.graphs) (void)hipGraphExecDestroy(g.second);
        sl.graphs.clear();
        if (sl.block) (void)hipFree(sl.block);
        if (sl.h_offsets) (void)hipHostFree(sl.h_offsets);
        if (sl.h_coefs) (void)hipHostFree(sl.h_coefs);
        if (sl.done) (void)hipEventDestroy(sl.done);
        sl.block = nullptr;
        sl.h_offsets = sl.h_coefs = nullptr;
        sl.done = nullptr;
    }

    slot* new_slot(uint32_t pi) {
        plan_rt& p = *plans[pi];
        emida_offset_s& o = *p.off;
        auto sl = std::make_unique<slot>();
        sl->plan = pi;
        sl->bytes = p.slot_bytes;
        const uint32_t n = p.max_batch;
        const size_t tn = (size_t)o.nb * n, nk = (size_t)o.fit * o.fit;
        sl->block = dmalloc(p.slot_bytes);
        char* b = (char*)sl->block;
        sl->pics = b;
        b += emida_offset_s::al256(n * p.pat_bytes);
        sl->scratch = o.carve(b, n);
        b += o.scratch_bytes(n);
        sl->maxes = b;
        b += emida_offset_s::al256(tn * o.ti_bytes());
        sl->neigh = b;
        b += emida_offset_s::al256(tn * nk * o.t_bytes());
        sl->offsets = (double*)b;
        b += emida_offset_s::al256(tn * 2 * sizeof(double));
        sl->coefs = (double*)b;
        EMIDA_HIP(hipHostMalloc((void**)&sl->h_offsets, tn * 2 * sizeof(double), hipHostMallocDefault));
        EMIDA_HIP(hipHostMalloc((void**)&sl->h_coefs, tn * 6 * sizeof(double), hipHostMallocDefault));
        EMIDA_HIP(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
        slots.push_back(std::move(sl));
        return slots.back().get();
    }

    // ------------------------------------------------------------ requests
    uint64_t submit(uint32_t pi, const uint16_t* pattern, double* offsets, double* coefs) {
        std::unique_lock<std::mutex> lock(mu);
        check_running();
        if (pi >= plans.size()) throw arg_error("unknown plan");
        const size_t bytes = plans[pi]->pat_bytes;
        cv_staging.wait(lock, [&] { return stop || staging_used < cfg.max_queued; });
        check_running();
        ++staging_used;
        void* buf = nullptr;
        auto& fl = staging_free[bytes];
        if (!fl.empty()) {
            buf = fl.back();
            fl.pop_back();
        }
        lock.unlock();
        try {
            if (!buf) {
                EMIDA_HIP(hipHostMalloc(&buf, bytes, hipHostMallocDefault));
                std::lock_guard<std::mutex> l2(mu);
                staging_all.push_back(buf);
            }
            std::memcpy(buf, pattern, bytes);
        } catch (...) {
            std::lock_guard<std::mutex> l2(mu);
            --staging_used;
            if (buf) staging_free[bytes].push_back(buf);
            cv_staging.notify_one();
            throw;
        }
        auto* r = new request;
        r->plan = pi;
        r->staging = buf;
        r->out_offsets = offsets;
        r->out_coefs = coefs;
        lock.lock();
        r->id = next_id++;
        r->submit_ns = now_ns();
        requests[r->id] = r;
        plans[pi]->queue.push_back(r);
        ++st.submitted;
        lock.unlock();
        cv_sched.notify_one();
        return r->id;
    }

    void wait(uint64_t id, emida_request_info* info) {
        std::unique_lock<std::mutex> lock(mu);
        auto it = requests.find(id);
        if (it == requests.end()) throw arg_error("unknown request (or already waited for)");
        request* r = it->second;
        cv_done.wait(lock, [&] { return r->done || !worker_error.empty(); });
        if (!r->done) throw state_error("server stopped: " + worker_error);
        if (info) {
            info->submit_ns = r->submit_ns;
            info->launch_ns = r->launch_ns;
            info->done_ns = r->done_ns;
            info->batch = r->batch;
            info->plan = r->plan;
        }
        requests.erase(it);
        delete r;
    }

    // ------------------------------------------------------------ scheduler
    // Plan with queued requests whose head is oldest and that can get a slot.
    int pick_plan() {
        int best = -1;
        uint64_t t = ~0ull;
        for (size_t i = 0; i < plans.size(); ++i) {
            plan_rt& p = *plans[i];
            if (p.queue.empty()) continue;
            const bool can = !p.free_slots.empty() ||
                             ((cfg.slots_per_plan == 0 || p.n_slots < cfg.slots_per_plan) &&
                              slot_bytes_total + p.slot_bytes <= cfg.hbm_budget);
            if (can && p.queue.front()->submit_ns < t) {
                t = p.queue.front()->submit_ns;
                best = (int)i;
            }
        }
        return best;
    }

    // A plan with queued requests that cannot get a slot although it is below its own
    // slot limit: the budget is taken (by other plans' slots).  Oldest head first.
    int starving_plan() const {
        int best = -1;
        uint64_t t = ~0ull;
        for (size_t i = 0; i < plans.size(); ++i) {
            const plan_rt& p = *plans[i];
            if (p.queue.empty() || !p.free_slots.empty()) continue;
            if (cfg.slots_per_plan != 0 && p.n_slots >= cfg.slots_per_plan) continue;
            if (slot_bytes_total + p.slot_bytes <= cfg.hbm_budget) continue;
            if (p.queue.front()->submit_ns < t) {
                t = p.queue.front()->submit_ns;
                best = (int)i;
            }
        }
        return best;
    }

    // Another plan holding an idle slot, preferably one with nothing queued; -1 when the
    // budget is held by slots in flight only (their completions free them).
    int victim_plan(int starving) const {
        int v = -1;
        for (size_t i = 0; i < plans.size(); ++i) {
            if ((int)i == starving || plans[i]->free_slots.empty()) continue;
            if (v < 0 || (plans[i]->queue.empty() && !plans[v]->queue.empty())) v = (int)i;
        }
        return v;
    }

    // Frees the idle slot of plan v that has been idle longest, giving its bytes back to
    // the budget: without this, plans whose idle slots fill the budget would keep the
    // requests of every other plan waiting forever.
    void evict_slot(int v, std::unique_lock<std::mutex>& lock) {
        plan_rt& p = *plans[v];
        slot* sl = p.free_slots.front();
        p.free_slots.erase(p.free_slots.begin());
        --p.n_slots;
        slot_bytes_total -= sl->bytes;
        std::unique_ptr<slot> owned;
        for (auto it = slots.begin(); it != slots.end(); ++it)
            if (it->get() == sl) {
                owned = std::move(*it);
                slots.erase(it);
                break;
            }
        st.slots = (uint32_t)slots.size();
        st.device_bytes_batches = slot_bytes_total;
        lock.unlock();
        release_slot(*owned);  // idle: no work of it is in flight
        lock.lock();
    }

    void schedule_loop() {
        std::unique_lock<std::mutex> lock(mu);
        while (true) {
            int pi = -1, victim = -1;
            cv_sched.wait(lock, [&] {
                if (stop) return true;
                if ((pi = pick_plan()) >= 0) return true;
                const int sp = starving_plan();
                victim = sp >= 0 ? victim_plan(sp) : -1;
                return victim >= 0;
            });
            if (stop) return;
            if (pi < 0) {
                evict_slot(victim, lock);
                continue;
            }
            plan_rt& p = *plans[pi];
            slot* sl;
            if (!p.free_slots.empty()) {
                sl = p.free_slots.back();
                p.free_slots.pop_back();
            } else {
                slot_bytes_total += p.slot_bytes;
                ++p.n_slots;
                lock.unlock();
                try {
                    sl = new_slot((uint32_t)pi);
                } catch (...) {
                    lock.lock();
                    slot_bytes_total -= p.slot_bytes;
                    --p.n_slots;
                    throw;
                }
                lock.lock();
                st.slots = (uint32_t)slots.size();
                st.device_bytes_batches = slot_bytes_total;
            }
            const uint32_t n = (uint32_t)std::min<size_t>(p.queue.size(), p.max_batch);
            launch_rec rec{sl, {}};
            for (uint32_t i = 0; i < n; ++i) {
                rec.reqs.push_back(p.queue.front());
                p.queue.pop_front();
            }
            const uint32_t k = next_stream++ % cfg.streams;
            inflight_bytes += sl->bytes;
            st.device_bytes_peak_inflight = std::max<uint64_t>(st.device_bytes_peak_inflight, inflight_bytes);
            lock.unlock();
            launch(p, *sl, rec.reqs, streams[k]);
            lock.lock();
            inflight[k].push_back(std::move(rec));
            cv_inflight[k].notify_one();
        }
    }

    void launch(plan_rt& p, slot& sl, const std::vector<request*>& reqs, hipStream_t s) {
        emida_offset_s& o = *p.off;
        const uint32_t n = (uint32_t)reqs.size();
        uint32_t nr = n;  // patterns the launch computes
        if (cfg.use_graphs) nr = std::min(graph_batch(n), p.max_batch);
        const size_t tn = (size_t)o.nb * nr;
        const uint64_t t = now_ns();
        for (uint32_t i = 0; i < n; ++i) {
            reqs[i]->launch_ns = t;
            reqs[i]->batch = n;
            EMIDA_HIP(hipMemcpyAsync((char*)sl.pics + i * p.pat_bytes, reqs[i]->staging, p.pat_bytes,
                                     hipMemcpyHostToDevice, s));
        }
        auto body = [&](hipStream_t q) {
            o.run(sl.pics, nr, sl.maxes, sl.neigh, sl.offsets, sl.coefs, sl.scratch, q);
            EMIDA_HIP(hipMemcpyAsync(sl.h_offsets, sl.offsets, tn * 2 * sizeof(double), hipMemcpyDeviceToHost, q));
            EMIDA_HIP(hipMemcpyAsync(sl.h_coefs, sl.coefs, tn * 6 * sizeof(double), hipMemcpyDeviceToHost, q));
        };
        if (cfg.use_graphs) {
            auto it = sl.graphs.find(nr);
            if (it == sl.graphs.end()) {
                hipGraph_t g = nullptr;
                hipGraphExec_t ge = nullptr;
                EMIDA_HIP(hipStreamBeginCapture(capture_stream, hipStreamCaptureModeRelaxed));
                try {
                    body(capture_stream);
                } catch (...) {
                    (void)hipStreamEndCapture(capture_stream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                EMIDA_HIP(hipStreamEndCapture(capture_stream, &g));
                const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                EMIDA_HIP(e);
                it = sl.graphs.emplace(nr, ge).first;
                std::lock_guard<std::mutex> lock(mu);
                ++st.graphs_captured;
            }
            EMIDA_HIP(hipGraphLaunch(it->second, s));
        } else {
            body(s);
        }
        EMIDA_HIP(hipEventRecord(sl.done, s));
        std::lock_guard<std::mutex> lock(mu);
        ++st.launches;
        if (cfg.use_graphs) ++st.graph_launches;
    }

    // ------------------------------------------------------------ completion
    void complete_loop(uint32_t k) {
        std::unique_lock<std::mutex> lock(mu);
        while (true) {
            cv_inflight[k].wait(lock, [&] { return stop || !inflight[k].empty(); });
            if (inflight[k].empty()) return;  // stop with nothing left on this stream
            slot* sl = inflight[k].front().sl;
            std::vector<request*> reqs = inflight[k].front().reqs;
            plan_rt& p = *plans[sl->plan];  // plan objects stay put; the vector may grow
            lock.unlock();
            EMIDA_HIP(hipEventSynchronize(sl->done));
            const uint64_t t = now_ns();
            const size_t nb = p.off->nb;
            for (size_t i = 0; i < reqs.size(); ++i) {
                if (reqs[i]->out_offsets)
                    std::memcpy(reqs[i]->out_offsets, sl->h_offsets + i * nb * 2, nb * 2 * sizeof(double));
                if (reqs[i]->out_coefs)
                    std::memcpy(reqs[i]->out_coefs, sl->h_coefs + i * nb * 6, nb * 6 * sizeof(double));
            }
            lock.lock();
            inflight[k].pop_front();
            for (request* r : reqs) {
                r->done_ns = t;
                r->done = true;
                staging_free[p.pat_bytes].push_back(r->staging);
                r->staging = nullptr;
                --staging_used;
            }
            st.completed += reqs.size();
            inflight_bytes -= sl->bytes;
            p.free_slots.push_back(sl);
            cv_sched.notify_one();
            cv_done.notify_all();
            cv_staging.notify_all();
        }
    }
};

extern "C" {

int emida_server_create(emida_server** out, const emida_server_config* cfg) {
    return guarded([&] {
        if (!out || !cfg) throw arg_error("null argument");
        *out = nullptr;
        if (cfg->hbm_budget == 0 || cfg->max_batch == 0 || cfg->streams == 0 || cfg->max_queued == 0)
            throw arg_error("hbm_budget, max_batch, streams and max_queued must be positive");
        if (cfg->streams > 16) throw arg_error("at most 16 streams");
        auto s = std::make_unique<emida_server_s>();
        s->cfg = *cfg;
        s->start();
        *out = s.release();
    });
}

void emida_server_destroy(emida_server* s) {
    if (!s) return;
    s->shutdown();
    delete s;
}

int emida_server_add_plan(emida_server* s, emida_type T, emida_size2 src_size, const uint16_t* ref,
                          const emida_size2* begins, uint32_t n_begins, emida_size2 slice_size,
                          emida_size2 cross_size, emida_cross_policy policy, emida_mfma_mode mfma_mode, int fitsize,
                          uint32_t* plan_id) {
    return guarded([&] {
        if (!s || !ref || !plan_id) throw arg_error("null argument");
        emida_offset* h = nullptr;
        const int rc = emida_offset_create(&h, T, EMIDA_U16, src_size, begins, n_begins, slice_size, cross_size,
                                           s->cfg.max_batch, policy, fitsize, 1);
        if (rc != EMIDA_OK) {
            if (rc == EMIDA_ERR_ARG) throw arg_error(g_last_error);
            if (rc == EMIDA_ERR_UNSUPPORTED) throw unsupported_error(g_last_error);
            throw hip_error(g_last_error);
        }
        std::unique_ptr<emida_offset_s> off(h);
        if (policy == EMIDA_CROSS_POLICY_BRUTE_MFMA) {
            if (mfma_mode != EMIDA_MFMA_AUTO) with_HT(mfma_mode, [](auto) {});
            off->mfma_mode = mfma_mode;
        }
        *plan_id = s->add_plan(std::move(off), ref);
    });
}

int emida_server_submit(emida_server* s, uint32_t plan_id, const uint16_t* pattern, double* offsets, double* coefs,
                        uint64_t* request_id) {
    return guarded([&] {
        if (!s || !pattern || !request_id) throw arg_error("null argument");
        *request_id = s->submit(plan_id, pattern, offsets, coefs);
    });
}

int emida_server_wait(emida_server* s, uint64_t request_id, emida_request_info* info) {
    return guarded([&] {
        if (!s) throw arg_error("null argument");
        s->wait(request_id, info);
    });
}

int emida_server_get_stats(emida_server* s, emida_server_stats* out) {
    return guarded([&] {
        if (!s || !out) throw arg_error("null argument");
        std::lock_guard<std::mutex> lock(s->mu);
        *out = s->st;
    });
}

int emida_server_plan_info(emida_server* s, uint32_t plan_id, emida_offset_info* info, uint32_t* max_batch) {
    return guarded([&] {
        if (!s || !info) throw arg_error("null argument");
        std::lock_guard<std::mutex> lock(s->mu);
        if (plan_id >= s->plans.size()) throw arg_error("unknown plan");
        const int rc = emida_offset_get_info(s->plans[plan_id]->off.get(), info);
        if (rc != EMIDA_OK) throw arg_error(g_last_error);
        if (max_batch) *max_batch = s->plans[plan_id]->max_batch;
    });
}

}  // extern "C"
