// The reference's cross_corr_fft test (test/cross_corr_fft_test.cpp:8-42) against the
// C++ drop-in, with the assertions the reference's test lacks, plus the drop-in's
// other test accessors (get_cu_pic, get_cu_temp, the public in_stream / out_stream).
//
// Expected values are direct DFTs in double written out here: R2C of the 2s x 2s
// buffer, product with the conjugate of the cu_temp half spectrum, and C2R as a
// column-then-row inverse (each row's C2R ignores the imaginary parts of its DC and
// Nyquist terms).  Run from the repository root; exits non-zero on a failed check.
#include <emida/gpu_offset.hpp>
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

using namespace emida;
using cd = std::complex<double>;

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

static const double kPi = 3.14159265358979323846;

// Unnormalised R2C of an n x m real array (rows of m): half spectrum [ky][kx], kx <= m/2.
static std::vector<cd> r2c(const std::vector<double>& a, int n, int m) {
    std::vector<cd> out((size_t)n * (m / 2 + 1));
    for (int ky = 0; ky < n; ++ky)
        for (int kx = 0; kx <= m / 2; ++kx) {
            cd s = 0;
            for (int y = 0; y < n; ++y)
                for (int x = 0; x < m; ++x)
                    s += a[(size_t)y * m + x] * std::polar(1.0, -2 * kPi * ((double)kx * x / m + (double)ky * y / n));
            out[(size_t)ky * (m / 2 + 1) + kx] = s;
        }
    return out;
}

// Unnormalised C2R of a half spectrum [ky][kx] (kx <= m/2) to n x m reals.
static std::vector<double> c2r(const std::vector<cd>& h, int n, int m) {
    const int nc = m / 2 + 1;
    std::vector<cd> g((size_t)n * nc);  // inverse along y, per column
    for (int y = 0; y < n; ++y)
        for (int kx = 0; kx < nc; ++kx) {
            cd s = 0;
            for (int ky = 0; ky < n; ++ky) s += h[(size_t)ky * nc + kx] * std::polar(1.0, 2 * kPi * (double)ky * y / n);
            g[(size_t)y * nc + kx] = s;
        }
    std::vector<double> out((size_t)n * m);
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < m; ++x) {
            double s = g[(size_t)y * nc].real() + g[(size_t)y * nc + m / 2].real() * ((x & 1) ? -1.0 : 1.0);
            for (int kx = 1; kx < m / 2; ++kx)
                s += 2 * (g[(size_t)y * nc + kx] * std::polar(1.0, 2 * kPi * (double)kx * x / m)).real();
            out[(size_t)y * m + x] = s;
        }
    return out;
}

static double max_rel_diff(const std::vector<double>& a, const std::vector<double>& b) {
    double scale = 0, d = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        scale = std::max(scale, std::fabs(b[i]));
        d = std::max(d, std::fabs(a[i] - b[i]));
    }
    return d / (scale > 0 ? scale : 1);
}

// test/cross_corr_fft_test.cpp:8-42, same inputs and calls
static void reference_test() {
    std::vector<double> in{1, 2, 3, 1, 0, 0, 0, 0,  //
                           4, 5, 6, 1, 0, 0, 0, 0,  //
                           7, 8, 9, 1, 0, 0, 0, 0,  //
                           7, 8, 9, 1, 0, 0, 0, 0,  //
                           0, 0, 0, 0, 0, 0, 0, 0,  //
                           0, 0, 0, 0, 0, 0, 0, 0,  //
                           0, 0, 0, 0, 0, 0, 0, 0,  //
                           0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<double> in_temp{6, 3, 2, 1, 0, 0, 0, 0, 0, 0,  //
                                12, 23, 3, 1, 0, 0, 0, 0, 0, 0,  //
                                1, 8, 9, 1, 0, 0, 0, 0, 0, 0,  //
                                7, 8, 9, 1, 0, 0, 0, 0, 0, 0,  //
                                0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  //
                                0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  //
                                0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  //
                                0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    size2_t slice_size{4, 4};
    std::vector<size2_t> begins{{0, 0}};
    gpu_offset<double, uint16_t> offs({10, 10}, &begins, slice_size, slice_size * 2 - 1, 1, CROSS_POLICY_FFT);
    offs.allocate_memory(nullptr);
    copy_to_device(in, offs.get_cu_pic());
    copy_to_device(in_temp, offs.get_cu_temp());

    offs.cross_corr_fft();

    // cu_temp's 80 doubles are the 8 x 5 complex half spectrum the product reads
    std::vector<cd> temp(40);
    for (int i = 0; i < 40; ++i) temp[i] = cd(in_temp[2 * i], in_temp[2 * i + 1]);
    std::vector<cd> p = r2c(in, 8, 8);
    for (size_t i = 0; i < p.size(); ++i) p[i] *= std::conj(temp[i]);
    const std::vector<double> exp = c2r(p, 8, 8);
    const std::vector<double> got = device_to_vector(offs.get_cu_cross_res(), 64);
    const double d = max_rel_diff(got, exp);
    CHECK(d < 1e-12, "cross_corr_fft on the reference's inputs: relative difference %.3g", d);
}

// allocate_memory with a reference: cu_temp holds the R2C of the zero-padded prepared
// reference slices (src/gpu_offset.hpp:185-210), and cross_corr_fft of a prepared
// pattern slice against it is the (2s)^2-scaled circular correlation
static void reference_spectrum_test() {
    const int W = 48, s = 16, m = 2 * s;
    std::vector<uint16_t> temp((size_t)W * W);
    for (int y = 0; y < W; ++y)
        for (int x = 0; x < W; ++x)
            temp[(size_t)y * W + x] = (uint16_t)(1000 + 400 * std::sin(0.3 * x + 0.11 * y * y) + 7 * ((x * 31 + y * 17) % 13));
    std::vector<size2_t> begins{{3, 5}, {20, 28}};
    gpu_offset<float, uint16_t> offs({(esize_t)W, (esize_t)W}, &begins, {(esize_t)s, (esize_t)s}, {31, 31}, 2,
                                     CROSS_POLICY_FFT, 5);
    offs.allocate_memory(temp.data());

    // the prepared reference slices: ((T)px - (T)sum / area) hx[x] hy[y] in float
    std::vector<double> hann(s);
    for (int i = 0; i < s; ++i) hann[i] = (float)(0.5 * (1 - std::cos(2 * kPi * i / (s - 1))));
    const std::vector<float> spec_dev = device_to_vector(offs.get_cu_temp(), (size_t)2 * m * (s + 1) * 2);
    for (int r = 0; r < 2; ++r) {
        uint32_t sum = 0;
        for (int y = 0; y < s; ++y)
            for (int x = 0; x < s; ++x) sum += temp[(size_t)(begins[r].y + y) * W + begins[r].x + x];
        const float mean = (float)sum / (float)(s * s);
        std::vector<double> pad((size_t)m * m, 0.0);
        for (int y = 0; y < s; ++y)
            for (int x = 0; x < s; ++x) {
                float v = (float)temp[(size_t)(begins[r].y + y) * W + begins[r].x + x];
                v -= mean;
                v = v * (float)hann[x] * (float)hann[y];
                pad[(size_t)y * m + x] = v;
            }
        const std::vector<cd> exp = r2c(pad, m, m);
        double scale = 0, d = 0;
        for (size_t i = 0; i < exp.size(); ++i) {
            const size_t k = ((size_t)r * exp.size() + i) * 2;
            const cd g(spec_dev[k], spec_dev[k + 1]);
            scale = std::max(scale, std::abs(exp[i]));
            d = std::max(d, std::abs(g - exp[i]));
        }
        CHECK(d / scale < 1e-5, "reference spectrum of ROI %d: relative difference %.3g", r, d / scale);
    }

    // a prepared pattern slice in every cu_pic buffer, correlated with its ROI's spectrum
    std::vector<float> pics((size_t)4 * m * m, 0.f);
    for (int b = 0; b < 4; ++b)
        for (int y = 0; y < s; ++y)
            for (int x = 0; x < s; ++x) pics[((size_t)b * m + y) * m + x] = (float)std::cos(0.2 * x * (b + 1) - 0.37 * y);
    copy_to_device(pics, offs.get_cu_pic());
    offs.cross_corr_fft();
    const std::vector<float> got = device_to_vector(offs.get_cu_cross_res(), (size_t)4 * m * m);
    for (int b = 0; b < 4; ++b) {
        std::vector<double> pd(pics.begin() + (size_t)b * m * m, pics.begin() + (size_t)(b + 1) * m * m);
        std::vector<cd> p = r2c(pd, m, m);
        const int r = b % 2;
        for (size_t i = 0; i < p.size(); ++i) {
            const size_t k = ((size_t)r * p.size() + i) * 2;
            p[i] *= std::conj(cd(spec_dev[k], spec_dev[k + 1]));
        }
        const std::vector<double> exp = c2r(p, m, m);
        const std::vector<double> gd(got.begin() + (size_t)b * m * m, got.begin() + (size_t)(b + 1) * m * m);
        const double d = max_rel_diff(gd, exp);
        CHECK(d < 1e-5, "cross_corr_fft slice %d: relative difference %.3g", b, d);
    }

    // the public streams: uploads run on in_stream (src/process_files.hpp:196)
    CHECK(offs.in_stream != nullptr && offs.out_stream != nullptr && offs.in_stream != offs.out_stream,
          "allocate_memory sets two streams");
    std::vector<uint16_t> two((size_t)2 * W * W);
    for (size_t i = 0; i < two.size(); ++i) two[i] = temp[i % temp.size()] + (uint16_t)(i % 5);
    const auto sync = offs.get_offset(two.data());
    const esize_t bi = offs.transfer_pic_to_device_async(two.data());
    CHECK(hipStreamSynchronize(offs.in_stream) == hipSuccess, "hipStreamSynchronize(in_stream)");
    std::vector<data_index<float>> maxes(offs.total_slices());
    std::vector<float> neigh(offs.neigh_size());
    offs.get_offset_core(maxes.data(), neigh.data(), bi);
    CHECK(hipStreamSynchronize(offs.out_stream) == hipSuccess, "hipStreamSynchronize(out_stream)");
    const auto fin = offs.finalize(maxes.data(), neigh.data());
    for (esize_t i = 0; i < offs.total_slices(); ++i)
        CHECK(fin.offsets[i] == sync.offsets[i], "offset %u differs between get_offset and the stream pipeline", i);
}

int main() {
    try {
        reference_test();
        reference_spectrum_test();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        return 2;
    }
    if (g_failures) {
        std::fprintf(stderr, "%d check(s) failed\n", g_failures);
        return 1;
    }
    std::printf("cross_corr_fft drop-in test passed\n");
    return 0;
}
