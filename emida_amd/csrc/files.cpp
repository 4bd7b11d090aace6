// Multi-threaded file pipeline: the reference's file_processor
// (src/process_files.hpp:98-241) behind one C entry point.
//
// Three stages joined by blocking queues, as the reference:
//  - loader threads (the caller plus load_workers - 1) take batches of batch_size
//    files with an atomic counter, read them into pinned host buffers
//    (tiff_read) and start the upload into the next free input buffer;
//  - one GPU thread runs get_offset_core on each uploaded batch;
//  - one finalize thread waits for each batch's results and hands them to the
//    caller's callback.
// Deliberate differences:
//  - a batch whose file cannot be read is reported to the callback with its error
//    instead of being dropped silently, and the pipeline always ends (the
//    reference's GPU thread waits forever when the last batch fails to load);
//  - each loader owns two pinned buffers and reuses one only after its upload
//    finished; the GPU thread takes result buffers from a free list that the
//    finalize thread refills, so no stage overwrites data another still reads;
//  - `ordered` delivers batches in file order (the reference prints them in the
//    order they finish).
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "io.hpp"
#include "offset.hpp"

namespace {

template <class V>
class blocking_queue {
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<V> q_;

public:
    void push(V v) {
        {
            std::lock_guard<std::mutex> lock(mu_);
            q_.push_back(std::move(v));
        }
        cv_.notify_one();
    }
    V pop() {
        std::unique_lock<std::mutex> lock(mu_);
        cv_.wait(lock, [&] { return !q_.empty(); });
        V v = std::move(q_.front());
        q_.pop_front();
        return v;
    }
};

// One batch on its way through the stages (offs_job, src/process_files.hpp:27-36).
struct batch_job {
    uint32_t first = 0;   // index of the batch's first file
    uint32_t count = 0;   // files in the batch (the last one may be short)
    uint32_t buffer = 0;  // input buffer holding the upload
    int status = EMIDA_OK;
    std::string error;
};

// Result buffers of one batch in flight (fin_job, src/process_files.hpp:39-50).
struct result_set {
    void* maxes = nullptr;
    void* neigh = nullptr;
};

struct finished {
    batch_job job;
    result_set* res = nullptr;  // null for a batch that was not computed
};

using clock_type = std::chrono::steady_clock;

double ms_since(clock_type::time_point t0) {
    return std::chrono::duration<double, std::milli>(clock_type::now() - t0).count();
}

struct pinned {
    void* p = nullptr;
    pinned() = default;
    void alloc(size_t bytes) { EMIDA_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault)); }
    pinned(const pinned&) = delete;
    pinned& operator=(const pinned&) = delete;
    ~pinned() {
        if (p) (void)hipHostFree(p);
    }
};

struct event {
    hipEvent_t e = nullptr;
    event() = default;
    void create() { EMIDA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }
    event(const event&) = delete;
    event& operator=(const event&) = delete;
    ~event() {
        if (e) (void)hipEventDestroy(e);
    }
};

int status_of(const std::exception& e) {
    if (dynamic_cast<const arg_error*>(&e)) return EMIDA_ERR_ARG;
    if (dynamic_cast<const unsupported_error*>(&e)) return EMIDA_ERR_UNSUPPORTED;
    if (dynamic_cast<const hip_error*>(&e)) return EMIDA_ERR_HIP;
    return EMIDA_ERR_STATE;
}

struct pipeline {
    emida_offset* h;
    const char* const* paths;
    uint32_t n_paths, n_batches;
    emida_files_callback cb;
    void* user;
    bool ordered;

    std::atomic<uint32_t> next_batch{0};
    std::atomic<bool> abort{false};
    std::mutex err_mu;
    int first_status = EMIDA_OK;
    std::string first_error;

    blocking_queue<batch_job> load_q;
    blocking_queue<finished> fin_q;
    blocking_queue<result_set*> free_q;

    std::mutex stat_mu;
    double load_ms = 0, wait_ms = 0;
    uint32_t failed = 0;

    void fail_pipeline(int status, const std::string& msg) {
        std::lock_guard<std::mutex> lock(err_mu);
        if (first_status == EMIDA_OK) {
            first_status = status;
            first_error = msg;
        }
        abort = true;
    }

    // load_rasters (src/process_files.hpp:98-133)
    void loader() {
        const size_t area = h->src.area();
        pinned buf[2];
        event staged[2];
        bool used[2] = {false, false};
        int k = 0;
        double my_load = 0, my_wait = 0;
        bool have_buffers = true;
        try {
            for (int i = 0; i < 2; ++i) {
                buf[i].alloc(area * h->batch * sizeof(uint16_t));
                staged[i].create();
            }
        } catch (const std::exception& e) {
            fail_pipeline(status_of(e), e.what());
            have_buffers = false;
        }
        uint32_t b;
        while ((b = next_batch.fetch_add(1)) < n_batches) {
            batch_job job;
            job.first = b * h->batch;
            job.count = std::min<uint32_t>(h->batch, n_paths - job.first);
            if (abort || !have_buffers) {
                job.status = EMIDA_ERR_STATE;
                job.error = "pipeline stopped after an earlier error";
                load_q.push(std::move(job));
                continue;
            }
            try {
                auto t0 = clock_type::now();
                if (used[k]) EMIDA_HIP(hipEventSynchronize(staged[k].e));
                uint16_t* dst = (uint16_t*)buf[k].p;
                for (uint32_t i = 0; i < job.count; ++i)
                    tiff_read(paths[job.first + i], dst + (size_t)i * area, h->src);
                my_load += ms_since(t0);
            } catch (const std::exception& e) {
                // the reference skips a batch with an unreadable file (load_tiff
                // returns false); report it and go on
                job.status = status_of(e);
                job.error = e.what();
                if (job.status == EMIDA_ERR_HIP) fail_pipeline(job.status, job.error);
                load_q.push(std::move(job));
                continue;
            }
            try {
                auto t0 = clock_type::now();
                job.buffer = h->upload(buf[k].p, staged[k].e);
                my_wait += ms_since(t0);
                used[k] = true;
                k ^= 1;
            } catch (const std::exception& e) {
                job.status = status_of(e);
                job.error = e.what();
                fail_pipeline(job.status, job.error);
            }
            load_q.push(std::move(job));
        }
        // the pinned buffers are freed on return: wait for their last uploads
        for (int i = 0; i < 2; ++i)
            if (used[i]) (void)hipEventSynchronize(staged[i].e);
        std::lock_guard<std::mutex> lock(stat_mu);
        load_ms += my_load;
        wait_ms += my_wait;
    }

    // compute_offsets_thread (src/process_files.hpp:170-207)
    void gpu() {
        for (uint32_t n = 0; n < n_batches; ++n) {
            batch_job job = load_q.pop();
            finished f;
            if (job.status == EMIDA_OK && abort) {
                h->release(job.buffer);
                job.status = EMIDA_ERR_STATE;
                job.error = "pipeline stopped after an earlier error";
            }
            if (job.status == EMIDA_OK) {
                result_set* r = free_q.pop();
                try {
                    h->core(r->maxes, r->neigh, job.buffer);
                    f.res = r;
                } catch (const std::exception& e) {
                    h->release(job.buffer);
                    free_q.push(r);
                    job.status = status_of(e);
                    job.error = e.what();
                    fail_pipeline(job.status, job.error);
                }
            }
            f.job = std::move(job);
            fin_q.push(std::move(f));
        }
    }

    void deliver(const batch_job& job, const double* offs, const double* coefs) {
        cb(user, job.first, job.count, job.status, job.status == EMIDA_OK ? nullptr : job.error.c_str(), offs, coefs);
    }

    // finalize_thread (src/process_files.hpp:210-241)
    void finalize() {
        const size_t per_file = h->total / h->batch;
        std::vector<double> offs(2 * (size_t)h->total), coefs(6 * (size_t)h->total);
        struct held {
            batch_job job;
            std::vector<double> offs, coefs;
        };
        std::map<uint32_t, held> waiting;  // ordered delivery: batches that finished early
        uint32_t next_first = 0;
        for (uint32_t n = 0; n < n_batches; ++n) {
            finished f = fin_q.pop();
            if (f.res) {
                try {
                    h->finalize(f.res->maxes, f.res->neigh, offs.data(), coefs.data());
                } catch (const std::exception& e) {
                    f.job.status = status_of(e);
                    f.job.error = e.what();
                    fail_pipeline(f.job.status, f.job.error);
                }
                free_q.push(f.res);
            }
            if (f.job.status != EMIDA_OK) {
                std::lock_guard<std::mutex> lock(stat_mu);
                ++failed;
            }
            const bool ok = f.job.status == EMIDA_OK;
            if (!ordered) {
                deliver(f.job, ok ? offs.data() : nullptr, ok ? coefs.data() : nullptr);
                continue;
            }
            held hd;
            hd.job = std::move(f.job);
            if (ok) {
                hd.offs.assign(offs.begin(), offs.begin() + 2 * per_file * hd.job.count);
                hd.coefs.assign(coefs.begin(), coefs.begin() + 6 * per_file * hd.job.count);
            }
            const uint32_t first = hd.job.first;
            waiting.emplace(first, std::move(hd));
            for (auto it = waiting.find(next_first); it != waiting.end(); it = waiting.find(next_first)) {
                const held& w = it->second;
                const bool wok = w.job.status == EMIDA_OK;
                deliver(w.job, wok ? w.offs.data() : nullptr, wok ? w.coefs.data() : nullptr);
                next_first += w.job.count;
                waiting.erase(it);
            }
        }
    }
};

}  // namespace

extern "C" int emida_process_files(emida_offset* h, const char* const* paths, uint32_t n_paths,
                                   uint32_t load_workers, int ordered, emida_files_callback cb, void* user,
                                   emida_files_stats* stats) {
    return guarded([&] {
        if (!h || !cb || (n_paths && !paths)) throw arg_error("null argument");
        h->require_allocated();
        if (h->IN != EMIDA_U16) throw arg_error("the file pipeline reads 16-bit TIFF patterns: IN must be U16");
        if (n_paths == 0) throw arg_error("The deformed list is empty.");
        if (load_workers == 0 || load_workers > 64) throw arg_error("load_workers must be in 1..64");
        for (uint32_t i = 0; i < n_paths; ++i)
            if (!paths[i]) throw arg_error("null path");
        EMIDA_HIP(hipSetDevice(h->device));
        const auto t0 = clock_type::now();

        pipeline p;
        p.h = h;
        p.paths = paths;
        p.n_paths = n_paths;
        p.n_batches = (n_paths + h->batch - 1) / h->batch;
        p.cb = cb;
        p.user = user;
        p.ordered = ordered != 0;

        // three result sets: one being filled, one being finalized, one spare
        const size_t nk = (size_t)h->fit * h->fit;
        pinned res_mem[3][2];
        result_set res[3];
        for (int i = 0; i < 3; ++i) {
            res_mem[i][0].alloc(h->total * h->ti_bytes());
            res_mem[i][1].alloc(h->total * nk * h->t_bytes());
            res[i] = {res_mem[i][0].p, res_mem[i][1].p};
            p.free_q.push(&res[i]);
        }

        const int device = h->device;
        std::thread gpu_thr([&] {
            (void)hipSetDevice(device);
            p.gpu();
        });
        std::thread fin_thr([&] {
            (void)hipSetDevice(device);
            p.finalize();
        });
        std::vector<std::thread> loaders;
        for (uint32_t i = 0; i + 1 < load_workers; ++i)
            loaders.emplace_back([&] {
                (void)hipSetDevice(device);
                p.loader();
            });
        p.loader();
        for (auto& t : loaders) t.join();
        gpu_thr.join();
        fin_thr.join();
        // results of the last batch were read by finalize; nothing stays queued
        EMIDA_HIP(hipStreamSynchronize(h->out_stream));
        for (int i = 0; i < 3; ++i) h->forget_slot(res[i].maxes, res[i].neigh);

        if (stats) {
            stats->batches = p.n_batches;
            stats->failed_batches = p.failed;
            stats->wall_ms = ms_since(t0);
            stats->load_ms = p.load_ms;
            stats->upload_wait_ms = p.wait_ms;
        }
        if (p.first_status != EMIDA_OK) {
            if (p.first_status == EMIDA_ERR_HIP) throw hip_error(p.first_error);
            throw state_error(p.first_error);
        }
    });
}
