// C++ drop-in API test (include/emida/gpu_offset.hpp over libemida_amd.so).
//
// Restates the reference's end-to-end tests (test/get_offset_test.cpp:10-27 and
// :108-199) with the reference's own types and calls, and runs the reference's
// three-stage host pipeline pattern (src/process_files.hpp:98-233: loader threads
// calling transfer_pic_to_device_async concurrently, one GPU thread calling
// get_offset_core into alternating pinned result buffers, one finalize thread)
// and checks it against sequential get_offset calls.  Run from the repository
// root; exits non-zero on the first failed check.
#include <emida/gpu_offset.hpp>

#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

using namespace emida;

static int g_failures = 0;
#define CHECK(cond, ...)                                      \
    do {                                                      \
        if (!(cond)) {                                        \
            std::fprintf(stderr, "FAILED %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);                \
            std::fprintf(stderr, "\n");                       \
            ++g_failures;                                     \
        }                                                     \
    } while (0)

// Minimal reader of a C-order little-endian float64 .npy file.
static std::vector<double> load_npy_f64(const std::string& path, size_t expect) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    char magic[6];
    f.read(magic, 6);
    unsigned char ver[2];
    f.read((char*)ver, 2);
    uint32_t hlen = 0;
    if (ver[0] == 1) {
        uint16_t h16;
        f.read((char*)&h16, 2);
        hlen = h16;
    } else {
        f.read((char*)&hlen, 4);
    }
    std::string header(hlen, ' ');
    f.read(&header[0], hlen);
    if (header.find("<f8") == std::string::npos || header.find("'fortran_order': False") == std::string::npos)
        throw std::runtime_error(path + ": not a C-order float64 array");
    std::vector<double> v(expect);
    f.read((char*)v.data(), expect * sizeof(double));
    if (!f) throw std::runtime_error(path + ": short read");
    return v;
}

static std::vector<vec2<double>> load_goldens(const std::string& path) {
    std::ifstream f(path);
    std::vector<vec2<double>> out;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        vec2<double> v;
        is >> v.x >> v.y;
        out.push_back(v);
    }
    return out;
}

// get_slice_begins (src/slice_picture.hpp:56-66)
static std::vector<size2_t> slice_begins(size2_t src, size2_t size, size2_t step) {
    std::vector<size2_t> res;
    for (esize_t y = 0; y + size.y <= src.y; y += step.y)
        for (esize_t x = 0; x + size.x <= src.x; x += step.x) res.push_back({x, y});
    return res;
}

static std::vector<double> submatrix(const std::vector<double>& src, esize_t n, size2_t size) {
    std::vector<double> r(size.area());
    for (esize_t y = 0; y < size.y; ++y)
        for (esize_t x = 0; x < size.x; ++x) r[y * size.x + x] = src[y * n + x];
    return r;
}

static void test_bigger(const std::vector<double>& pic, const std::vector<double>& temp) {
    size2_t size{64, 64};
    auto a = submatrix(pic, 256, size);
    auto b = submatrix(temp, 256, size);
    auto offset = get_offset<double>(a.data(), b.data(), size, 1);
    CHECK(std::fabs(offset[0].x - 0.07583538046549165) <= 2e-14, "bigger x %.17g", offset[0].x);
    CHECK(std::fabs(offset[0].y - -0.0982055210473689) <= 2e-14, "bigger y %.17g", offset[0].y);
    std::printf("get_offset.bigger ok\n");
}

static void test_batched(const std::vector<double>& pic_file, const std::vector<double>& temp,
                         const std::vector<vec2<double>>& expected, cross_policy policy, esize_t batch) {
    size2_t src{256, 256};
    std::vector<double> pic;
    for (esize_t i = 0; i < batch; ++i) pic.insert(pic.end(), pic_file.begin(), pic_file.end());
    auto begins = slice_begins(src, {64, 64}, {32, 32});
    gpu_offset<double, double> offs(src, &begins, {64, 64}, {127, 127}, batch, policy);
    offs.allocate_memory(const_cast<double*>(temp.data()));
    auto [offsets, coefs] = offs.get_offset(pic.data());
    CHECK(offsets.size() == expected.size() * batch, "size %zu", offsets.size());
    double worst = 0;
    for (size_t i = 0; i < offsets.size(); ++i) {
        const auto& e = expected[i % expected.size()];
        worst = std::max(worst, std::max(std::fabs(offsets[i].x - e.x), std::fabs(offsets[i].y - e.y)));
    }
    CHECK(worst <= 7e-13, "batched policy %d batch %u: max error %.3g", (int)policy, batch, worst);
    (void)coefs;
    std::printf("get_offset.batched %s x%u ok (max error %.2e)\n", policy == CROSS_POLICY_FFT ? "FFT" : "BRUTE", batch,
                worst);
}

// A bounded blocking queue (the reference's concurrent_queue, src/concurrent_queue.hpp:14-77).
template <class T>
class queue {
    std::deque<T> q_;
    std::mutex m_;
    std::condition_variable cv_;

public:
    void push(T v) {
        {
            std::lock_guard<std::mutex> l(m_);
            q_.push_back(std::move(v));
        }
        cv_.notify_all();
    }
    T pop() {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return !q_.empty(); });
        T v = std::move(q_.front());
        q_.pop_front();
        return v;
    }
};

struct job {
    esize_t first;  // first batch index
    esize_t buffer;
    bool last;
};

// The reference's pipeline with float uint16 input: 3 loader threads, one GPU
// thread, one finalize thread; every batch's offsets must equal a sequential run.
static void test_threaded_pipeline(const std::vector<double>& pic_file, const std::vector<double>& temp_file,
                                   cross_policy policy) {
    const size2_t src{256, 256};
    const esize_t batch = 2, n_batches = 12, loaders = 3;
    auto begins = slice_begins(src, {32, 32}, {28, 28});
    std::vector<uint16_t> temp(src.area());
    for (size_t i = 0; i < temp.size(); ++i) temp[i] = (uint16_t)std::lround(temp_file[i] * 100.0);
    // batch b holds the fixture pattern shifted by b pixels (wrapping), as uint16
    std::vector<uint16_t*> host(n_batches);
    for (esize_t b = 0; b < n_batches; ++b) {
        host[b] = cuda_malloc_host<uint16_t>(src.area() * batch);
        for (esize_t k = 0; k < batch; ++k)
            for (esize_t y = 0; y < src.y; ++y)
                for (esize_t x = 0; x < src.x; ++x)
                    host[b][k * src.area() + y * src.x + x] =
                        (uint16_t)std::lround(pic_file[y * src.x + (x + b + k) % src.x] * 100.0);
    }
    gpu_offset<float, uint16_t> seq(src, &begins, {32, 32}, {63, 63}, batch, policy, 5, 1);
    seq.allocate_memory(temp.data());
    std::vector<offsets_t<double>> expected;
    for (esize_t b = 0; b < n_batches; ++b) expected.push_back(seq.get_offset(host[b]));

    gpu_offset<float, uint16_t> offs(src, &begins, {32, 32}, {63, 63}, batch, policy, 5, 2 * loaders);
    offs.allocate_memory(temp.data());
    const esize_t total = (esize_t)begins.size() * batch;
    data_index<float>* maxes[2] = {cuda_malloc_host<data_index<float>>(total), cuda_malloc_host<data_index<float>>(total)};
    float* neigh[2] = {cuda_malloc_host<float>(total * 25), cuda_malloc_host<float>(total * 25)};

    queue<job> load_q, fin_q;
    queue<int> free_results;  // result buffer pairs the GPU thread may fill
    free_results.push(0);
    free_results.push(1);
    std::atomic<esize_t> next{0};
    std::vector<offsets_t<double>> got(n_batches);

    std::vector<std::thread> threads;
    for (esize_t l = 0; l < loaders; ++l)
        threads.emplace_back([&] {
            esize_t b;
            while ((b = next.fetch_add(1)) < n_batches) {
                esize_t buf = offs.transfer_pic_to_device_async(host[b]);
                load_q.push({b, buf, false});
            }
        });
    std::thread gpu([&] {
        for (esize_t done = 0; done < n_batches; ++done) {
            job j = load_q.pop();
            int r = free_results.pop();
            offs.get_offset_core(maxes[r], neigh[r], j.buffer);
            fin_q.push({j.first, (esize_t)r, false});
        }
    });
    std::thread fin([&] {
        for (esize_t done = 0; done < n_batches; ++done) {
            job j = fin_q.pop();
            got[j.first] = offs.finalize(maxes[j.buffer], neigh[j.buffer]);
            free_results.push((int)j.buffer);
        }
    });
    for (auto& t : threads) t.join();
    gpu.join();
    fin.join();
    esize_t mismatches = 0;
    for (esize_t b = 0; b < n_batches; ++b)
        for (esize_t i = 0; i < total; ++i)
            if (!(got[b].offsets[i] == expected[b].offsets[i]) &&
                !(std::isnan(got[b].offsets[i].x) && std::isnan(expected[b].offsets[i].x)))
                ++mismatches;
    CHECK(mismatches == 0, "threaded pipeline (policy %d): %u offsets differ from sequential runs", (int)policy,
          mismatches);
    for (auto p : host) cuda_free_host(p);
    for (int r = 0; r < 2; ++r) {
        cuda_free_host(maxes[r]);
        cuda_free_host(neigh[r]);
    }
    std::printf("pipeline.threaded %s ok (%u batches, %u loaders)\n", policy == CROSS_POLICY_FFT ? "FFT" : "BRUTE",
                n_batches, loaders);
}

static void test_errors() {
    std::vector<size2_t> begins = {{200, 0}};
    bool threw = false;
    try {
        gpu_offset<double, uint16_t> bad({128, 128}, &begins, {64, 64}, {127, 127}, 1);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    CHECK(threw, "ROI outside the pattern must throw");
    std::printf("errors ok\n");
}

int main() {
    try {
        auto pic = load_npy_f64("tests/golden/data_pic.npy", 256 * 256);
        auto temp = load_npy_f64("tests/golden/data_temp.npy", 256 * 256);
        auto expected = load_goldens("tests/golden/get_offset_batched.txt");
        CHECK(expected.size() == 49, "goldens: %zu", expected.size());
        test_bigger(pic, temp);
        for (auto policy : {CROSS_POLICY_BRUTE, CROSS_POLICY_FFT})
            for (esize_t batch : {1u, 5u}) test_batched(pic, temp, expected, policy, batch);
        test_threaded_pipeline(pic, temp, CROSS_POLICY_FFT);
        test_threaded_pipeline(pic, temp, CROSS_POLICY_BRUTE);
        test_errors();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        return 2;
    }
    if (g_failures) {
        std::fprintf(stderr, "%d check(s) failed\n", g_failures);
        return 1;
    }
    std::printf("all C++ API checks passed\n");
    return 0;
}
