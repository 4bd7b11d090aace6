// emida/gpu_offset.hpp - C++17 drop-in for the reference's emida::gpu_offset<T, IN>
// and emida::get_offset<T> (src/gpu_offset.hpp:48-326, src/get_offset.hpp:8-20),
// implemented over the C ABI in emida_amd.h.  Link with -lemida_amd.
//
// Same constructor and member signatures; errors throw std::runtime_error after
// printing to stderr, like the reference's CUCH (src/host_helpers.hpp:14-25).
// Differences: `begins` is copied (the reference keeps the pointer); device
// memory is released by the destructor; in_stream / out_stream are HIP streams
// (hipStreamSynchronize(offs.in_stream) where the reference calls
// cudaStreamSynchronize), set by allocate_memory; wait_for_uploads() does the same.
#pragma once

#include <array>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "emida_amd.h"

// HIP's stream handle type (hipStream_t is `struct ihipStream_t*`), declared here so
// this header needs no HIP include; the two name the same type.
struct ihipStream_t;

namespace emida {

using esize_t = uint32_t;  // src/common.hpp:24

template <typename T>
struct data_index {  // src/common.hpp:26-31
    T data;
    esize_t index;
};

template <typename T>
inline bool operator==(const data_index<T>& a, const data_index<T>& b) {
    return a.data == b.data && a.index == b.index;
}

template <typename T>
struct vec2 {  // src/common.hpp:41-123 (the members used by callers)
    T x;
    T y;
    T area() const { return x * y; }
    vec2 operator+(const vec2& r) const { return {x + r.x, y + r.y}; }
    vec2 operator-(const vec2& r) const { return {x - r.x, y - r.y}; }
    template <typename U>
    vec2 operator*(const U& r) const { return {x * r, y * r}; }
    template <typename U>
    vec2 operator-(const U& r) const { return {x - r, y - r}; }
    esize_t pos(esize_t cols) const { return y * cols + x; }
};

template <typename T>
inline bool operator==(const vec2<T>& a, const vec2<T>& b) {
    return a.x == b.x && a.y == b.y;
}

using size2_t = vec2<esize_t>;

enum cross_policy {
    CROSS_POLICY_BRUTE = EMIDA_CROSS_POLICY_BRUTE,
    CROSS_POLICY_FFT = EMIDA_CROSS_POLICY_FFT,
    // Matrix-core brute correlation (no reference counterpart).
    CROSS_POLICY_BRUTE_MFMA = EMIDA_CROSS_POLICY_BRUTE_MFMA
};

template <typename T>
struct offsets_t {  // src/subpixel_max.hpp:69-74
    std::vector<vec2<T>> offsets;
    std::vector<std::array<T, 6>> coefs;
};

namespace detail {

inline void check(int rc, const char* what) {
    if (rc != EMIDA_OK) {
        std::string msg = std::string("emida error ") + std::to_string(rc) + " in " + what + ": " + emida_last_error();
        std::cerr << msg << "\n";
        throw std::runtime_error(msg);
    }
}

template <typename T>
constexpr emida_type type_code() {
    if constexpr (std::is_same<T, float>::value) return EMIDA_F32;
    else if constexpr (std::is_same<T, double>::value) return EMIDA_F64;
    else if constexpr (std::is_same<T, uint16_t>::value) return EMIDA_U16;
    else return EMIDA_I32;
}

inline emida_size2 c(size2_t s) { return {s.x, s.y}; }

}  // namespace detail

// gpu_offset<T, IN>: T in {float, double}; IN in {uint16_t, float, double}.
template <typename T, typename IN>
class gpu_offset {
    emida_offset* h_ = nullptr;
    esize_t total_ = 0, neigh_ = 0;

    offsets_t<double> unpack(const std::vector<double>& o, const std::vector<double>& cf) const {
        offsets_t<double> r;
        r.offsets.resize(total_);
        r.coefs.resize(total_);
        for (esize_t i = 0; i < total_; ++i) {
            r.offsets[i] = {o[2 * i], o[2 * i + 1]};
            for (int k = 0; k < 6; ++k) r.coefs[i][k] = cf[6 * i + k];
        }
        return r;
    }

public:
    // src/gpu_offset.hpp:96-97: the upload and download streams (set by allocate_memory)
    ihipStream_t* in_stream = nullptr;
    ihipStream_t* out_stream = nullptr;

    gpu_offset() = default;
    gpu_offset(size2_t src_size, const std::vector<size2_t>* begins, size2_t slice_size, size2_t cross_size,
               esize_t batch_size, cross_policy policy = CROSS_POLICY_BRUTE, int s = 3, size_t in_buffers_size = 1) {
        std::vector<emida_size2> b(begins->size());
        for (size_t i = 0; i < b.size(); ++i) b[i] = detail::c((*begins)[i]);
        detail::check(emida_offset_create(&h_, detail::type_code<T>(), detail::type_code<IN>(), detail::c(src_size),
                                          b.data(), (uint32_t)b.size(), detail::c(slice_size), detail::c(cross_size),
                                          batch_size, (emida_cross_policy)policy, s, (uint32_t)in_buffers_size),
                      "gpu_offset");
        emida_offset_info info;
        detail::check(emida_offset_get_info(h_, &info), "gpu_offset");
        total_ = info.total_slices;
        neigh_ = info.neigh_size;
    }
    gpu_offset(const gpu_offset&) = delete;
    gpu_offset& operator=(const gpu_offset&) = delete;
    gpu_offset(gpu_offset&& o) noexcept
        : h_(o.h_), total_(o.total_), neigh_(o.neigh_), in_stream(o.in_stream), out_stream(o.out_stream) {
        o.h_ = nullptr;
    }
    gpu_offset& operator=(gpu_offset&& o) noexcept {
        std::swap(h_, o.h_);
        total_ = o.total_;
        neigh_ = o.neigh_;
        in_stream = o.in_stream;
        out_stream = o.out_stream;
        return *this;
    }
    ~gpu_offset() {
        if (h_) emida_offset_destroy(h_);
    }

    esize_t total_slices() const { return total_; }
    esize_t neigh_size() const { return neigh_; }

    // Half type of CROSS_POLICY_BRUTE_MFMA (default EMIDA_MFMA_AUTO); call before
    // allocate_memory (no reference counterpart).
    void set_mfma_mode(emida_mfma_mode mode) {
        detail::check(emida_offset_set_mfma_mode(h_, mode), "set_mfma_mode");
    }
    // The half type in effect (what EMIDA_MFMA_AUTO resolves to).
    emida_mfma_mode mfma_mode() const {
        emida_mfma_mode m = EMIDA_MFMA_AUTO;
        detail::check(emida_offset_get_mfma_mode(h_, &m), "get_mfma_mode");
        return m;
    }
    // Exact peak equal to CROSS_POLICY_BRUTE's results (default); before allocate_memory.
    void set_mfma_refine(bool enable) {
        detail::check(emida_offset_set_mfma_refine(h_, enable ? 1 : 0), "set_mfma_refine");
    }

    // src/gpu_offset.hpp:127
    void allocate_memory(IN* temp) {
        detail::check(emida_offset_allocate_memory(h_, temp), "allocate_memory");
        void *a = nullptr, *b = nullptr;
        detail::check(emida_offset_get_streams(h_, &a, &b), "allocate_memory");
        in_stream = static_cast<ihipStream_t*>(a);
        out_stream = static_cast<ihipStream_t*>(b);
    }

    // src/gpu_offset.hpp:213
    esize_t transfer_pic_to_device_async(IN* pic) {
        uint32_t i = 0;
        detail::check(emida_offset_transfer_pic_to_device_async(h_, pic, &i), "transfer_pic_to_device_async");
        return i;
    }

    // cudaStreamSynchronize(in_stream), src/process_files.hpp:196
    void wait_for_uploads() { detail::check(emida_offset_wait_uploads(h_), "wait_for_uploads"); }

    // src/gpu_offset.hpp:220
    offsets_t<double> get_offset(IN* pic) {
        std::vector<double> o(2 * (size_t)total_), cf(6 * (size_t)total_);
        detail::check(emida_offset_get_offset(h_, pic, o.data(), cf.data()), "get_offset");
        return unpack(o, cf);
    }

    // src/gpu_offset.hpp:234
    void get_offset_core(data_index<T>* maxes_i, T* neighbors, size_t cu_pic_index) {
        detail::check(emida_offset_get_offset_core(h_, maxes_i, neighbors, (uint32_t)cu_pic_index), "get_offset_core");
    }

    // src/gpu_offset.hpp:280
    offsets_t<double> finalize(data_index<T>* maxes_i, T* neighbors) {
        std::vector<double> o(2 * (size_t)total_), cf(6 * (size_t)total_);
        detail::check(emida_offset_finalize(h_, maxes_i, neighbors, o.data(), cf.data()), "finalize");
        return unpack(o, cf);
    }

    // src/gpu_offset.hpp:303-316: R2C of the get_cu_pic buffers, product with the
    // conjugate get_cu_temp spectra, unnormalised C2R (FFT policy; synchronous)
    void cross_corr_fft() const { detail::check(emida_offset_cross_corr_fft(h_), "cross_corr_fft"); }

    // src/gpu_offset.hpp:318-325 (device pointers, for tests)
    T* get_cu_pic() {
        void* p = nullptr;
        detail::check(emida_offset_get_cu_pic(h_, &p), "get_cu_pic");
        return static_cast<T*>(p);
    }
    T* get_cu_temp() {
        void* p = nullptr;
        detail::check(emida_offset_get_cu_temp(h_, &p), "get_cu_temp");
        return static_cast<T*>(p);
    }
    // the buffers cross_corr_fft writes (the reference's private cu_cross_res_)
    T* get_cu_cross_res() {
        void* p = nullptr;
        detail::check(emida_offset_get_cu_cross_res(h_, &p), "get_cu_cross_res");
        return static_cast<T*>(p);
    }

    emida_offset* handle() { return h_; }
};

// copy_to_device / copy_from_device (src/host_helpers.hpp:49-72), synchronous
template <typename T>
inline void copy_to_device(const std::vector<T>& v, T* dst) {
    detail::check(emida_memcpy_h2d(dst, v.data(), v.size() * sizeof(T)), "copy_to_device");
}
template <typename T>
inline void copy_to_device(const T* src, size_t n, T* dst) {
    detail::check(emida_memcpy_h2d(dst, src, n * sizeof(T)), "copy_to_device");
}
template <typename T>
inline std::vector<T> device_to_vector(const T* src, size_t n) {
    std::vector<T> v(n);
    detail::check(emida_memcpy_d2h(v.data(), src, n * sizeof(T)), "device_to_vector");
    return v;
}

// get_offset<T>(pic, temp, size, cross_size, policy), src/get_offset.hpp:8-14: one ROI
// covering the whole size.x x size.y picture at {0, 0}, batch 1.
template <typename T>
inline std::vector<vec2<double>> get_offset(T* pic, T* temp, size2_t size, size2_t cross_size,
                                            cross_policy c_policy = CROSS_POLICY_BRUTE) {
    std::vector<size2_t> begins = {{0, 0}};
    gpu_offset<T, T> offs(size, &begins, size, cross_size, 1, c_policy);
    offs.allocate_memory(temp);
    return offs.get_offset(pic).offsets;
}

// src/get_offset.hpp:16-20
template <typename T>
inline std::vector<vec2<double>> get_offset(T* pic, T* temp, size2_t size, size_t) {
    return get_offset(pic, temp, size, size * 2 - 1);
}

// cuda_malloc_host (src/host_helpers.hpp:74-80)
template <typename T>
T* cuda_malloc_host(size_t num_elements) {
    void* p = nullptr;
    detail::check(emida_host_alloc(&p, num_elements * sizeof(T)), "host_alloc");
    return (T*)p;
}
inline void cuda_free_host(void* p) { detail::check(emida_host_free(p), "host_free"); }

}  // namespace emida
