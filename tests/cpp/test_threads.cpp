// The library's threaded host code under load: the request server (its scheduler
// thread, one completion thread per stream, four submitting threads with requests in
// flight and a thread reading the statistics) and the file pipeline (loader threads,
// the GPU thread, the finalize thread, results in completion order, one unreadable
// file).  Every served request must equal the same request served alone, bit for
// bit, and every batch of the file run must be reported once with the offsets of the
// others (all files are the same image).  The same source builds against a copy of
// the library whose host code is instrumented for ThreadSanitizer (tests/cpp/Makefile,
// target tsan).  Run from the repository root; exits non-zero on a failed check.
#include <emida_amd.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

static int g_failures = 0;
#define CHECK(cond, ...)                                                \
    do {                                                                \
        if (!(cond)) {                                                  \
            std::fprintf(stderr, "FAILED %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);                          \
            std::fprintf(stderr, "\n");                                 \
            ++g_failures;                                               \
        }                                                               \
    } while (0)

static void ok(int rc, const char* what) {
    if (rc != EMIDA_OK) {
        std::fprintf(stderr, "%s: error %d: %s\n", what, rc, emida_last_error());
        std::exit(2);
    }
}

// A smooth synthetic pattern shifted by (dx, dy) pixels.
static std::vector<uint16_t> pattern(uint32_t w, uint32_t h, double dx, double dy) {
    std::vector<uint16_t> p((size_t)w * h);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const double u = x - dx, v = y - dy;
            p[(size_t)y * w + x] = (uint16_t)(30000 + 9000 * std::sin(0.11 * u + 0.05 * v) +
                                              7000 * std::cos(0.07 * u - 0.13 * v) + 3000 * std::sin(0.021 * u * v / 16));
        }
    return p;
}

static void server_run() {
    const uint32_t W = 256;
    const emida_size2 src{W, W};
    const std::vector<uint16_t> ref = pattern(W, W, 0, 0);
    std::vector<std::vector<uint16_t>> pics;
    for (int i = 0; i < 6; ++i) pics.push_back(pattern(W, W, 0.7 * i - 1.5, 1.1 - 0.4 * i));
    std::vector<emida_size2> b_fft, b_brute;
    for (uint32_t y = 0; y < 4; ++y)
        for (uint32_t x = 0; x < 4; ++x) b_fft.push_back({16 + 56 * x, 16 + 56 * y});
    for (uint32_t i = 0; i < 3; ++i) b_brute.push_back({40 + 70 * i, 60 + 50 * i});

    emida_server_config cfg{};
    cfg.hbm_budget = 256ull << 20;
    cfg.max_batch = 8;
    cfg.streams = 4;
    cfg.max_queued = 24;
    cfg.use_graphs = 1;
    emida_server* srv = nullptr;
    ok(emida_server_create(&srv, &cfg), "server_create");
    uint32_t p_fft = 0, p_brute = 0;
    ok(emida_server_add_plan(srv, EMIDA_F32, src, ref.data(), b_fft.data(), (uint32_t)b_fft.size(), {32, 32},
                             {63, 63}, EMIDA_CROSS_POLICY_FFT, EMIDA_MFMA_F16_SPLIT, 5, &p_fft),
       "add_plan fft");
    ok(emida_server_add_plan(srv, EMIDA_F64, src, ref.data(), b_brute.data(), (uint32_t)b_brute.size(), {16, 16},
                             {31, 31}, EMIDA_CROSS_POLICY_BRUTE, EMIDA_MFMA_F16_SPLIT, 3, &p_brute),
       "add_plan brute");

    // lone results first: one request at a time
    const uint32_t plans[2] = {p_fft, p_brute};
    const size_t nroi[2] = {b_fft.size(), b_brute.size()};
    std::vector<double> lone[2];
    for (int k = 0; k < 2; ++k) {
        lone[k].resize(pics.size() * 2 * nroi[k]);
        for (size_t i = 0; i < pics.size(); ++i) {
            uint64_t id = 0;
            ok(emida_server_submit(srv, plans[k], pics[i].data(), lone[k].data() + i * 2 * nroi[k], nullptr, &id),
               "submit");
            ok(emida_server_wait(srv, id, nullptr), "wait");
        }
    }

    // then four submitting threads at once and a fifth reading the statistics
    const int kThreads = 4, kPerThread = 120;
    std::atomic<bool> stop{false};
    std::thread watcher([&] {
        while (!stop.load()) {
            emida_server_stats st{};
            ok(emida_server_get_stats(srv, &st), "get_stats");
            std::this_thread::yield();
        }
    });
    std::vector<std::thread> subs;
    std::mutex mu;
    for (int t = 0; t < kThreads; ++t)
        subs.emplace_back([&, t] {
            struct job {
                uint64_t id;
                int k;
                size_t i;
                std::vector<double> offs;
            };
            std::vector<job> mine;
            for (int j = 0; j < kPerThread; ++j) {
                const int k = (j + t) % 2;
                const size_t i = (size_t)(j * 7 + t) % pics.size();
                job jb{0, k, i, std::vector<double>(2 * nroi[k])};
                ok(emida_server_submit(srv, plans[k], pics[i].data(), jb.offs.data(), nullptr, &jb.id), "submit");
                mine.push_back(std::move(jb));
                if (mine.size() == 6) {  // a window of requests in flight per thread
                    for (auto& m : mine) {
                        ok(emida_server_wait(srv, m.id, nullptr), "wait");
                        std::lock_guard<std::mutex> lock(mu);
                        CHECK(std::memcmp(m.offs.data(), lone[m.k].data() + m.i * 2 * nroi[m.k],
                                          m.offs.size() * sizeof(double)) == 0,
                              "served result of pattern %zu, plan %d differs from the lone request", m.i, m.k);
                    }
                    mine.clear();
                }
            }
            for (auto& m : mine) {
                ok(emida_server_wait(srv, m.id, nullptr), "wait");
                std::lock_guard<std::mutex> lock(mu);
                CHECK(std::memcmp(m.offs.data(), lone[m.k].data() + m.i * 2 * nroi[m.k],
                                  m.offs.size() * sizeof(double)) == 0,
                      "served result of pattern %zu, plan %d differs from the lone request", m.i, m.k);
            }
        });
    for (auto& s : subs) s.join();
    stop.store(true);
    watcher.join();
    emida_server_stats st{};
    ok(emida_server_get_stats(srv, &st), "get_stats");
    CHECK(st.completed == st.submitted, "%llu of %llu requests completed", (unsigned long long)st.completed,
          (unsigned long long)st.submitted);
    std::printf("server: %llu requests, %llu launches, %llu graphs captured\n", (unsigned long long)st.completed,
                (unsigned long long)st.launches, (unsigned long long)st.graphs_captured);
    emida_server_destroy(srv);
}

struct files_result {
    std::mutex mu;
    std::vector<double> offs;
    std::vector<int> seen;
    uint32_t n_roi = 0;
    int failed = 0;
};

static void on_batch(void* user, uint32_t first, uint32_t count, int status, const char*, const double* offsets,
                     const double*) {
    auto* r = static_cast<files_result*>(user);
    std::lock_guard<std::mutex> lock(r->mu);
    if (status != EMIDA_OK) {
        ++r->failed;
        return;
    }
    for (uint32_t f = 0; f < count; ++f) {
        ++r->seen[first + f];
        std::memcpy(r->offs.data() + (size_t)(first + f) * 2 * r->n_roi, offsets + (size_t)f * 2 * r->n_roi,
                    2 * r->n_roi * sizeof(double));
    }
}

static void files_run() {
    const char* tif = "tests/golden/INITIAL_x0y0.tif";
    emida_size2 size{};
    ok(emida_tiff_size(tif, &size), "tiff_size");
    std::vector<uint16_t> ref((size_t)size.x * size.y);
    ok(emida_tiff_load(tif, ref.data(), size), "tiff_load");
    std::vector<emida_size2> begins;
    for (uint32_t i = 0; i < 5; ++i) begins.push_back({100 + 120 * i, 150 + 100 * i});
    const uint32_t batch = 3, n_files = 40;
    emida_offset* h = nullptr;
    ok(emida_offset_create(&h, EMIDA_F32, EMIDA_U16, size, begins.data(), (uint32_t)begins.size(), {64, 64},
                           {127, 127}, batch, EMIDA_CROSS_POLICY_FFT, 5, 8),
       "offset_create");
    ok(emida_offset_allocate_memory(h, ref.data()), "allocate_memory");
    std::vector<const char*> paths(n_files, tif);
    paths[17] = "tests/golden/no_such_file.tif";  // fails its batch only
    files_result res;
    res.n_roi = (uint32_t)begins.size();
    res.offs.assign((size_t)n_files * 2 * res.n_roi, 0.0);
    res.seen.assign(n_files, 0);
    emida_files_stats st{};
    ok(emida_process_files(h, paths.data(), n_files, 4, 0, on_batch, &res, &st), "process_files");
    CHECK(res.failed == 1, "%d failed batches, expected 1", res.failed);
    // every file of the other batches seen once, all with the same offsets (same file)
    int good = 0;
    for (uint32_t f = 0; f < n_files; ++f) {
        if (f / batch == 17 / batch) continue;
        CHECK(res.seen[f] == 1, "file %u reported %d times", f, res.seen[f]);
        CHECK(std::memcmp(res.offs.data() + (size_t)f * 2 * res.n_roi, res.offs.data(),
                          2 * res.n_roi * sizeof(double)) == 0,
              "file %u: offsets differ from file 0's", f);
        ++good;
    }
    std::printf("files: %u batches (%u failed), %d files checked\n", st.batches, st.failed_batches, good);
    emida_offset_destroy(h);
}

int main() {
    server_run();
    files_run();
    if (g_failures) {
        std::fprintf(stderr, "%d check(s) failed\n", g_failures);
        return 1;
    }
    std::printf("threaded host code: all checks passed\n");
    return 0;
}
