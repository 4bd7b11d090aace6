// C ABI of the pattern-offset library (declared in include/emida_amd.h).
//
// emida_offset replaces the reference's gpu_offset<T, IN> (src/gpu_offset.hpp:48-326).
// Differences of the host side, all deliberate:
//  - three non-blocking streams (uploads, kernels, downloads) ordered by events;
//    the reference launches its kernels on the legacy default stream;
//  - an input buffer is not overwritten before the kernels that read it finished
//    (the reference relies on queue capacities);
//  - device results are double buffered so a download never races the next batch;
//  - the subpixel fit runs on the GPU inside the peak/FFT kernels;
//  - every buffer is released by emida_offset_destroy;
//  - ROI bounds and sizes are validated up front instead of reading out of bounds.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "emida_amd.h"
#include "kernels.hpp"

using namespace emida_amd;

namespace {

thread_local std::string g_last_error;

struct state_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return EMIDA_OK;
    } catch (const arg_error& e) {
        return fail(EMIDA_ERR_ARG, e.what());
    } catch (const unsupported_error& e) {
        return fail(EMIDA_ERR_UNSUPPORTED, e.what());
    } catch (const state_error& e) {
        return fail(EMIDA_ERR_STATE, e.what());
    } catch (const hip_error& e) {
        const std::string m = e.what();
        if (m.find("out of memory") != std::string::npos || m.find("OutOfMemory") != std::string::npos)
            return fail(EMIDA_ERR_NOMEM, m);
        return fail(EMIDA_ERR_HIP, m);
    } catch (const std::bad_alloc&) {
        return fail(EMIDA_ERR_NOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(EMIDA_ERR_HIP, e.what());
    }
}

size_t type_size(emida_type t) {
    switch (t) {
        case EMIDA_F32: return 4;
        case EMIDA_F64: return 8;
        case EMIDA_I32: return 4;
        case EMIDA_U16: return 2;
    }
    throw arg_error("unknown element type");
}

size2 S2(emida_size2 s) { return {s.x, s.y}; }

// generate_hanning (src/host_helpers.hpp:29-37): double expression, then cast.
template <class T>
std::vector<T> hanning(size_t n) {
    constexpr double PI = 3.14159265358979323846;
    std::vector<T> w(n);
    for (size_t i = 0; i < n; i++) w[i] = (T)(0.5 * (1 - cos(2 * PI * i / (n - 1))));
    return w;
}

template <class F>
void with_T(emida_type T, F&& f) {
    if (T == EMIDA_F32) f(float{});
    else if (T == EMIDA_F64) f(double{});
    else throw arg_error("T must be F32 or F64");
}

// IN element type and its sum type (sums_trait, src/cufft_helpers.hpp:157-179).
template <class F>
void with_IN(emida_type IN, F&& f) {
    if (IN == EMIDA_U16) f(uint16_t{}, uint32_t{});
    else if (IN == EMIDA_F32) f(float{}, float{});
    else if (IN == EMIDA_F64) f(double{}, double{});
    else throw arg_error("IN must be U16, F32 or F64");
}

template <class X>
struct tag {
    using type = X;
};

// Half type of a matrix-core mode.
template <class F>
void with_HT(emida_mfma_mode mode, F&& f) {
    if (mode == EMIDA_MFMA_BF16 || mode == EMIDA_MFMA_BF16_SPLIT) f(tag<__bf16>{});
    else if (mode == EMIDA_MFMA_F16 || mode == EMIDA_MFMA_F16_SPLIT) f(tag<_Float16>{});
    else throw arg_error("unknown matrix-core mode");
}
bool mfma_split(emida_mfma_mode mode) { return mode == EMIDA_MFMA_BF16_SPLIT || mode == EMIDA_MFMA_F16_SPLIT; }

void* dmalloc(size_t bytes) {
    void* p = nullptr;
    if (bytes == 0) bytes = 16;
    EMIDA_HIP(hipMalloc(&p, bytes));
    return p;
}

}  // namespace

struct emida_offset_s {
    emida_type T, IN;
    emida_cross_policy policy;
    int fit;
    size2 src, slice, cross, buf;
    uint32_t nb, batch, total, n_in;
    uint32_t side = 0;  // FFT slice side
    std::vector<size2> begins;
    size_t sum_bytes = 0;

    int device = 0;
    bool allocated = false;
    hipStream_t in_stream = nullptr, compute = nullptr, out_stream = nullptr;
    std::vector<void*> d_in;
    std::vector<hipEvent_t> ev_up, ev_done;
    std::atomic<uint32_t> rr{0};
    std::mutex upload_mu;
    std::condition_variable upload_cv;
    std::vector<char> pending;  // buffer holds an upload no get_offset_core consumed yet

    size2* d_begins = nullptr;
    void* d_hx = nullptr;
    void* d_hy = nullptr;
    void* d_sums = nullptr;
    void* d_ref = nullptr;     // brute: prepared reference slices; FFT: spectra
    void* d_slices = nullptr;  // brute: prepared pattern slices
    void* d_corr = nullptr;    // brute: correlation buffers + zero tail
    size_t corr_elems = 0;
    emida_mfma_mode mfma_mode = EMIDA_MFMA_BF16_SPLIT;
    void* d_ref_lo = nullptr;  // matrix-core policy: lo halves of the reference slices
    void* d_pic_hi = nullptr;  // matrix-core policy: prepared pattern slices, hi / lo halves
    void* d_pic_lo = nullptr;

    struct outset {
        void* maxes = nullptr;
        void* neigh = nullptr;
        double* offsets = nullptr;
        double* coefs = nullptr;
        hipEvent_t copied = nullptr;
        bool used = false;
    } out[2];
    uint32_t out_i = 0;

    struct slot {
        const void* maxes = nullptr;
        const void* neigh = nullptr;
        double* h_offsets = nullptr;
        double* h_coefs = nullptr;
        hipEvent_t done = nullptr;
        uint64_t stamp = 0;
    };
    std::vector<slot> slots;
    uint64_t stamp = 0;
    std::mutex slot_mu;

    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> tev;
    double t_ms = 0;
    uint64_t t_launches = 0;

    size_t device_bytes = 0;

    size_t ti_bytes() const { return T == EMIDA_F32 ? 8 : 16; }
    size_t t_bytes() const { return T == EMIDA_F32 ? 4 : 8; }
    size_t in_bytes() const { return type_size(IN); }

    void* alloc(size_t bytes) {
        void* p = dmalloc(bytes);
        device_bytes += bytes;
        return p;
    }

    ~emida_offset_s() {
        if (device >= 0) (void)hipSetDevice(device);
        if (compute) (void)hipStreamSynchronize(compute);
        if (in_stream) (void)hipStreamSynchronize(in_stream);
        if (out_stream) (void)hipStreamSynchronize(out_stream);
        for (void* p : d_in) (void)hipFree(p);
        for (auto e : ev_up) (void)hipEventDestroy(e);
        for (auto e : ev_done) (void)hipEventDestroy(e);
        void* ptrs[] = {d_begins, d_hx, d_hy, d_sums, d_ref, d_slices, d_corr, d_ref_lo, d_pic_hi, d_pic_lo};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        for (auto& o : out) {
            if (o.maxes) (void)hipFree(o.maxes);
            if (o.neigh) (void)hipFree(o.neigh);
            if (o.offsets) (void)hipFree(o.offsets);
            if (o.coefs) (void)hipFree(o.coefs);
            if (o.copied) (void)hipEventDestroy(o.copied);
        }
        for (auto& s : slots) {
            if (s.h_offsets) (void)hipHostFree(s.h_offsets);
            if (s.h_coefs) (void)hipHostFree(s.h_coefs);
            if (s.done) (void)hipEventDestroy(s.done);
        }
        for (auto& p : tev) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
        if (in_stream) (void)hipStreamDestroy(in_stream);
        if (compute) (void)hipStreamDestroy(compute);
        if (out_stream) (void)hipStreamDestroy(out_stream);
    }

    void require_allocated() const {
        if (!allocated) throw state_error("allocate_memory must be called first");
    }

    // ---- reference preparation (src/gpu_offset.hpp:185-210)
    void prepare_reference(const void* temp) {
        void* d_temp = dmalloc((size_t)src.area() * in_bytes());
        EMIDA_HIP(hipMemcpyAsync(d_temp, temp, (size_t)src.area() * in_bytes(), hipMemcpyHostToDevice, compute));
        with_T(T, [&](auto t) {
            using TT = decltype(t);
            with_IN(IN, [&](auto in, auto sm) {
                using IT = decltype(in);
                using ST = decltype(sm);
                const IT* pic = (const IT*)d_temp;
                ST* sums = (ST*)d_sums;  // first nb entries are free before any batch runs
                if (policy == EMIDA_CROSS_POLICY_BRUTE) {
                    launch_sum<IT, ST>(pic, sums, d_begins, src, slice, nb, nb, compute);
                    launch_prepare<IT, TT, ST>(pic, (TT*)d_ref, (const TT*)d_hx, (const TT*)d_hy, sums, d_begins,
                                               src, slice, slice, nb, nb, compute);
                } else if (policy == EMIDA_CROSS_POLICY_BRUTE_MFMA) {
                    launch_sum<IT, ST>(pic, sums, d_begins, src, slice, nb, nb, compute);
                    with_HT(mfma_mode, [&](auto ht) {
                        using HT = typename decltype(ht)::type;
                        launch_prepare_split<IT, ST, HT>(pic, (HT*)d_ref, (HT*)d_ref_lo, (const float*)d_hx,
                                                         (const float*)d_hy, sums, d_begins, src, slice, nb, nb,
                                                         compute);
                    });
                } else {
                    if (!std::is_same<IT, uint16_t>::value)
                        launch_sum<IT, ST>(pic, sums, d_begins, src, slice, nb, nb, compute);
                    launch_fft_ref<TT, IT, ST>(pic, sums, (const TT*)d_hx, (const TT*)d_hy, d_begins, src, side, nb,
                                               (cplx<TT>*)d_ref, compute);
                }
            });
        });
        EMIDA_HIP(hipStreamSynchronize(compute));
        EMIDA_HIP(hipFree(d_temp));
    }

    void allocate(const void* temp) {
        if (allocated) throw state_error("allocate_memory was already called");
        EMIDA_HIP(hipGetDevice(&device));
        EMIDA_HIP(hipStreamCreateWithFlags(&in_stream, hipStreamNonBlocking));
        EMIDA_HIP(hipStreamCreateWithFlags(&compute, hipStreamNonBlocking));
        EMIDA_HIP(hipStreamCreateWithFlags(&out_stream, hipStreamNonBlocking));
        const size_t pic_bytes = (size_t)src.area() * batch * in_bytes();
        for (uint32_t i = 0; i < n_in; ++i) {
            d_in.push_back(alloc(pic_bytes));
            hipEvent_t a, b;
            EMIDA_HIP(hipEventCreateWithFlags(&a, hipEventDisableTiming));
            EMIDA_HIP(hipEventCreateWithFlags(&b, hipEventDisableTiming));
            ev_up.push_back(a);
            ev_done.push_back(b);
        }
        pending.assign(n_in, 0);
        d_begins = (size2*)alloc(begins.size() * sizeof(size2));
        EMIDA_HIP(hipMemcpy(d_begins, begins.data(), begins.size() * sizeof(size2), hipMemcpyHostToDevice));
        with_T(T, [&](auto t) {
            using TT = decltype(t);
            auto hx = hanning<TT>(slice.x), hy = hanning<TT>(slice.y);
            d_hx = alloc(hx.size() * sizeof(TT));
            d_hy = alloc(hy.size() * sizeof(TT));
            EMIDA_HIP(hipMemcpy(d_hx, hx.data(), hx.size() * sizeof(TT), hipMemcpyHostToDevice));
            EMIDA_HIP(hipMemcpy(d_hy, hy.data(), hy.size() * sizeof(TT), hipMemcpyHostToDevice));
        });
        d_sums = alloc(std::max(total, nb) * sum_bytes);
        if (policy == EMIDA_CROSS_POLICY_BRUTE) {
            d_ref = alloc((size_t)nb * slice.area() * t_bytes());
            d_slices = alloc((size_t)total * slice.area() * t_bytes());
            corr_elems = (size_t)total * cross.area() + cross.x + 1;
            d_corr = alloc(corr_elems * t_bytes());
            EMIDA_HIP(hipMemset(d_corr, 0, corr_elems * t_bytes()));
        } else if (policy == EMIDA_CROSS_POLICY_BRUTE_MFMA) {
            const size_t half = 2;  // bf16 / fp16
            d_ref = alloc((size_t)nb * slice.area() * half);
            d_pic_hi = alloc((size_t)total * slice.area() * half);
            if (mfma_split(mfma_mode)) {
                d_ref_lo = alloc((size_t)nb * slice.area() * half);
                d_pic_lo = alloc((size_t)total * slice.area() * half);
            }
            corr_elems = (size_t)total * cross.area() + cross.x + 1;
            d_corr = alloc(corr_elems * sizeof(float));
            EMIDA_HIP(hipMemset(d_corr, 0, corr_elems * sizeof(float)));
        } else {
            const size_t spec = (size_t)(side + 1) * (2 * side) * 2 * t_bytes();
            d_ref = alloc(spec * nb);
        }
        const size_t nk = (size_t)fit * fit;
        for (auto& o : out) {
            o.maxes = alloc(total * ti_bytes());
            o.neigh = alloc(total * nk * t_bytes());
            o.offsets = (double*)alloc(total * 2 * sizeof(double));
            o.coefs = (double*)alloc(total * 6 * sizeof(double));
            EMIDA_HIP(hipEventCreateWithFlags(&o.copied, hipEventDisableTiming));
        }
        (void)lstsq_device_table();
        if (temp) prepare_reference(temp);
        allocated = true;
    }

    // ---- one batch on device data (src/gpu_offset.hpp:234-278)
    void run(const void* d_pics, uint32_t n_pat, void* maxes, void* neigh, double* offsets, double* coefs,
             hipStream_t st) {
        const uint32_t tn = nb * n_pat;
        hipEvent_t t0 = nullptr, t1 = nullptr;
        with_T(T, [&](auto t) {
            using TT = decltype(t);
            with_IN(IN, [&](auto in, auto sm) {
                using IT = decltype(in);
                using ST = decltype(sm);
                const IT* pics = (const IT*)d_pics;
                ST* sums = (ST*)d_sums;
                if (policy == EMIDA_CROSS_POLICY_BRUTE) {
                    launch_sum<IT, ST>(pics, sums, d_begins, src, slice, nb, tn, st);
                    launch_prepare<IT, TT, ST>(pics, (TT*)d_slices, (const TT*)d_hx, (const TT*)d_hy, sums, d_begins,
                                               src, slice, slice, nb, tn, st);
                    if (timing) t0 = record_start(st);
                    launch_cross_corr<TT>((const TT*)d_slices, (const TT*)d_ref, (TT*)d_corr, slice, cross, nb, n_pat,
                                          st);
                    if (timing) record_stop(st, t0);
                    launch_peak<TT>((const TT*)d_corr, (size_t)tn * cross.area(), cross.area(), cross, POLICY_BRUTE,
                                    fit, (data_index<TT>*)maxes, (TT*)neigh, offsets, coefs, tn, st);
                } else if (policy == EMIDA_CROSS_POLICY_BRUTE_MFMA) {
                    launch_sum<IT, ST>(pics, sums, d_begins, src, slice, nb, tn, st);
                    with_HT(mfma_mode, [&](auto ht) {
                        using HT = typename decltype(ht)::type;
                        launch_prepare_split<IT, ST, HT>(pics, (HT*)d_pic_hi, (HT*)d_pic_lo, (const float*)d_hx,
                                                         (const float*)d_hy, sums, d_begins, src, slice, nb, tn, st);
                        if (timing) t0 = record_start(st);
                        launch_mfma_corr<HT>((const HT*)d_pic_hi, (const HT*)d_pic_lo, (const HT*)d_ref,
                                             (const HT*)d_ref_lo, (float*)d_corr, slice, cross, nb, n_pat, st);
                        if (timing) record_stop(st, t0);
                    });
                    launch_peak<float>((const float*)d_corr, (size_t)tn * cross.area(), cross.area(), cross,
                                       POLICY_BRUTE, fit, (data_index<float>*)maxes, (float*)neigh, offsets, coefs,
                                       tn, st);
                } else {
                    if (!std::is_same<IT, uint16_t>::value)
                        launch_sum<IT, ST>(pics, sums, d_begins, src, slice, nb, tn, st);
                    if (timing) t0 = record_start(st);
                    launch_fft_pipe<TT, IT, ST>(pics, sums, (const TT*)d_hx, (const TT*)d_hy, d_begins, src, side,
                                                nb, tn, (const cplx<TT>*)d_ref, fit, (data_index<TT>*)maxes,
                                                (TT*)neigh, offsets, coefs, nullptr, st);
                    if (timing) record_stop(st, t0);
                }
            });
        });
        (void)t1;
    }

    hipEvent_t record_start(hipStream_t st) {
        hipEvent_t a, b;
        EMIDA_HIP(hipEventCreate(&a));
        EMIDA_HIP(hipEventCreate(&b));
        EMIDA_HIP(hipEventRecord(a, st));
        tev.emplace_back(a, b);
        return a;
    }
    void record_stop(hipStream_t st, hipEvent_t) { EMIDA_HIP(hipEventRecord(tev.back().second, st)); }

    slot& slot_for(const void* maxes, const void* neigh) {
        for (auto& s : slots)
            if (s.maxes == maxes && s.neigh == neigh) return s;
        if (slots.size() < 8) {
            slots.emplace_back();
            slot& s = slots.back();
            EMIDA_HIP(hipHostMalloc((void**)&s.h_offsets, total * 2 * sizeof(double), hipHostMallocDefault));
            EMIDA_HIP(hipHostMalloc((void**)&s.h_coefs, total * 6 * sizeof(double), hipHostMallocDefault));
            EMIDA_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
            s.maxes = maxes;
            s.neigh = neigh;
            return s;
        }
        slot* lru = &slots[0];
        for (auto& s : slots)
            if (s.stamp < lru->stamp) lru = &s;
        EMIDA_HIP(hipEventSynchronize(lru->done));
        lru->maxes = maxes;
        lru->neigh = neigh;
        return *lru;
    }

    void core(void* maxes_i, void* neighbors, uint32_t bi) {
        require_allocated();
        if (bi >= n_in) throw arg_error("buffer index out of range");
        outset& o = out[out_i];
        out_i ^= 1;
        EMIDA_HIP(hipStreamWaitEvent(compute, ev_up[bi], 0));
        if (o.used) EMIDA_HIP(hipStreamWaitEvent(compute, o.copied, 0));
        run(d_in[bi], batch, o.maxes, o.neigh, o.offsets, o.coefs, compute);
        EMIDA_HIP(hipEventRecord(ev_done[bi], compute));
        {
            std::lock_guard<std::mutex> lock(upload_mu);
            pending[bi] = 0;
        }
        upload_cv.notify_all();
        EMIDA_HIP(hipStreamWaitEvent(out_stream, ev_done[bi], 0));
        const size_t nk = (size_t)fit * fit;
        if (maxes_i) EMIDA_HIP(hipMemcpyAsync(maxes_i, o.maxes, total * ti_bytes(), hipMemcpyDeviceToHost, out_stream));
        if (neighbors)
            EMIDA_HIP(hipMemcpyAsync(neighbors, o.neigh, total * nk * t_bytes(), hipMemcpyDeviceToHost, out_stream));
        std::lock_guard<std::mutex> lock(slot_mu);
        slot& s = slot_for(maxes_i, neighbors);
        s.stamp = ++stamp;
        EMIDA_HIP(hipMemcpyAsync(s.h_offsets, o.offsets, total * 2 * sizeof(double), hipMemcpyDeviceToHost, out_stream));
        EMIDA_HIP(hipMemcpyAsync(s.h_coefs, o.coefs, total * 6 * sizeof(double), hipMemcpyDeviceToHost, out_stream));
        EMIDA_HIP(hipEventRecord(s.done, out_stream));
        EMIDA_HIP(hipEventRecord(o.copied, out_stream));
        o.used = true;
    }

    void finalize(const void* maxes_i, const void* neighbors, double* offsets, double* coefs) {
        require_allocated();
        const slot* found = nullptr;
        slot copy;
        {
            std::lock_guard<std::mutex> lock(slot_mu);
            for (auto& s : slots)
                if (s.maxes == maxes_i && s.neigh == neighbors && s.stamp != 0) {
                    copy = s;
                    found = &copy;
                    break;
                }
        }
        if (found) {
            // wait outside the lock so the next get_offset_core is not held up; the
            // caller reuses these result buffers only after this returns
            EMIDA_HIP(hipEventSynchronize(found->done));
            if (offsets) std::memcpy(offsets, found->h_offsets, total * 2 * sizeof(double));
            if (coefs) std::memcpy(coefs, found->h_coefs, total * 6 * sizeof(double));
            return;
        }
        if (!maxes_i || !neighbors) throw arg_error("finalize: maxes_i and neighbors are required");
        EMIDA_HIP(hipStreamSynchronize(out_stream));
        // foreign buffers: fit them on the GPU
        const size_t nk = (size_t)fit * fit;
        void* dm = dmalloc(total * ti_bytes());
        void* dn = dmalloc(total * nk * t_bytes());
        double* doff = (double*)dmalloc(total * 2 * sizeof(double));
        double* dco = (double*)dmalloc(total * 6 * sizeof(double));
        try {
            EMIDA_HIP(hipMemcpyAsync(dm, maxes_i, total * ti_bytes(), hipMemcpyHostToDevice, compute));
            EMIDA_HIP(hipMemcpyAsync(dn, neighbors, total * nk * t_bytes(), hipMemcpyHostToDevice, compute));
            const int pol = policy == EMIDA_CROSS_POLICY_FFT ? POLICY_FFT : POLICY_BRUTE;
            with_T(T, [&](auto t) {
                using TT = decltype(t);
                launch_subpixel<TT>((const TT*)dn, fit, total, (const data_index<TT>*)dm, cross, pol, doff, dco,
                                    compute);
            });
            if (offsets) EMIDA_HIP(hipMemcpyAsync(offsets, doff, total * 2 * sizeof(double), hipMemcpyDeviceToHost, compute));
            if (coefs) EMIDA_HIP(hipMemcpyAsync(coefs, dco, total * 6 * sizeof(double), hipMemcpyDeviceToHost, compute));
            EMIDA_HIP(hipStreamSynchronize(compute));
        } catch (...) {
            (void)hipStreamSynchronize(compute);
            (void)hipFree(dm);
            (void)hipFree(dn);
            (void)hipFree(doff);
            (void)hipFree(dco);
            throw;
        }
        EMIDA_HIP(hipFree(dm));
        EMIDA_HIP(hipFree(dn));
        EMIDA_HIP(hipFree(doff));
        EMIDA_HIP(hipFree(dco));
    }
};

// ====================================================================== C ABI
extern "C" {

const char* emida_last_error(void) { return g_last_error.c_str(); }

int emida_offset_create(emida_offset** out, emida_type T, emida_type IN, emida_size2 src_size,
                        const emida_size2* begins, uint32_t n_begins, emida_size2 slice_size,
                        emida_size2 cross_size, uint32_t batch_size, emida_cross_policy policy, int fitsize,
                        uint32_t in_buffers) {
    return guarded([&] {
        if (!out) throw arg_error("out is null");
        *out = nullptr;
        if (T != EMIDA_F32 && T != EMIDA_F64) throw arg_error("T must be F32 or F64");
        if (IN != EMIDA_U16 && IN != EMIDA_F32 && IN != EMIDA_F64) throw arg_error("IN must be U16, F32 or F64");
        if (policy != EMIDA_CROSS_POLICY_BRUTE && policy != EMIDA_CROSS_POLICY_FFT &&
            policy != EMIDA_CROSS_POLICY_BRUTE_MFMA)
            throw arg_error("unknown policy");
        if (policy == EMIDA_CROSS_POLICY_BRUTE_MFMA && T != EMIDA_F32)
            throw unsupported_error("matrix-core brute policy: results are float (T must be F32)");
        if (fitsize != 3 && fitsize != 5 && fitsize != 7 && fitsize != 9)
            throw arg_error("fitsize must be 3, 5, 7 or 9");
        if (!begins || n_begins == 0) throw arg_error("at least one ROI is required");
        if (batch_size == 0 || in_buffers == 0) throw arg_error("batch_size and in_buffers must be positive");
        if (slice_size.x < 2 || slice_size.y < 2) throw arg_error("slice size must be at least 2x2");
        for (uint32_t i = 0; i < n_begins; ++i)
            if ((uint64_t)begins[i].x + slice_size.x > src_size.x || (uint64_t)begins[i].y + slice_size.y > src_size.y)
                throw arg_error("ROI " + std::to_string(i) + " at (" + std::to_string(begins[i].x) + ", " +
                                std::to_string(begins[i].y) + ") does not fit in the pattern");
        if ((uint64_t)n_begins * batch_size > 0x7fffffffull) throw arg_error("too many slices");
        auto h = std::make_unique<emida_offset_s>();
        h->T = T;
        h->IN = IN;
        h->policy = policy;
        h->fit = fitsize;
        h->src = S2(src_size);
        h->slice = S2(slice_size);
        h->nb = n_begins;
        h->batch = batch_size;
        h->total = n_begins * batch_size;
        h->n_in = in_buffers;
        h->begins.resize(n_begins);
        for (uint32_t i = 0; i < n_begins; ++i) h->begins[i] = S2(begins[i]);
        h->sum_bytes = IN == EMIDA_U16 ? 4 : type_size(IN);
        if (policy == EMIDA_CROSS_POLICY_BRUTE_MFMA) check_mfma_shape(S2(slice_size), S2(cross_size));
        if (policy != EMIDA_CROSS_POLICY_FFT) {
            if (cross_size.x == 0 || cross_size.y == 0) throw arg_error("cross size must be positive");
            if (cross_size.x > 2 * slice_size.x - 1 || cross_size.y > 2 * slice_size.y - 1)
                throw arg_error("cross size must be at most 2*slice-1 per axis");
            h->cross = S2(cross_size);
            h->buf = h->cross;
        } else {
            // the FFT policy forces cross_size = 2s - 1 (src/gpu_offset.hpp:114)
            if (slice_size.x != slice_size.y)
                throw unsupported_error("FFT policy: square slices only");
            const uint32_t s = slice_size.x;
            if ((s & (s - 1)) != 0 || s > (uint32_t)fft_max_side(T == EMIDA_F64) / 2)
                throw unsupported_error("FFT policy: slice side must be a power of two <= " +
                                        std::to_string(fft_max_side(T == EMIDA_F64) / 2));
            h->side = s;
            h->cross = {2 * s - 1, 2 * s - 1};
            h->buf = {2 * s, 2 * s};
        }
        *out = h.release();
    });
}

void emida_offset_destroy(emida_offset* h) { delete h; }

int emida_offset_get_info(const emida_offset* h, emida_offset_info* info) {
    return guarded([&] {
        if (!h || !info) throw arg_error("null argument");
        info->total_slices = h->total;
        info->neigh_size = h->total * h->fit * h->fit;
        info->data_index_bytes = (uint32_t)h->ti_bytes();
        info->cross_size = {h->cross.x, h->cross.y};
        info->corr_buffer = {h->buf.x, h->buf.y};
        info->fft_size = h->side * 2;
        info->device_bytes = h->device_bytes;
    });
}

int emida_offset_allocate_memory(emida_offset* h, const void* temp) {
    return guarded([&] {
        if (!h) throw arg_error("null object");
        h->allocate(temp);
    });
}

int emida_offset_transfer_pic_to_device_async(emida_offset* h, const void* pic, uint32_t* buffer_index) {
    return guarded([&] {
        if (!h || !pic) throw arg_error("null argument");
        h->require_allocated();
        std::unique_lock<std::mutex> lock(h->upload_mu);
        const uint32_t i = h->rr.fetch_add(1) % h->n_in;
        h->upload_cv.wait(lock, [&] { return !h->pending[i]; });
        h->pending[i] = 1;
        EMIDA_HIP(hipStreamWaitEvent(h->in_stream, h->ev_done[i], 0));
        EMIDA_HIP(hipMemcpyAsync(h->d_in[i], pic, (size_t)h->src.area() * h->batch * h->in_bytes(),
                                 hipMemcpyHostToDevice, h->in_stream));
        EMIDA_HIP(hipEventRecord(h->ev_up[i], h->in_stream));
        if (buffer_index) *buffer_index = i;
    });
}

int emida_offset_wait_uploads(emida_offset* h) {
    return guarded([&] {
        if (!h) throw arg_error("null object");
        h->require_allocated();
        EMIDA_HIP(hipStreamSynchronize(h->in_stream));
    });
}

int emida_offset_get_offset_core(emida_offset* h, void* maxes_i, void* neighbors, uint32_t buffer_index) {
    return guarded([&] {
        if (!h) throw arg_error("null object");
        h->core(maxes_i, neighbors, buffer_index);
    });
}

int emida_offset_finalize(emida_offset* h, const void* maxes_i, const void* neighbors, double* offsets,
                          double* coefs) {
    return guarded([&] {
        if (!h) throw arg_error("null object");
        h->finalize(maxes_i, neighbors, offsets, coefs);
    });
}

int emida_offset_get_offset(emida_offset* h, const void* pic, double* offsets, double* coefs) {
    return guarded([&] {
        if (!h || !pic || !offsets) throw arg_error("null argument");
        h->require_allocated();
        auto& o = h->out[h->out_i];
        h->out_i ^= 1;
        EMIDA_HIP(hipStreamSynchronize(h->out_stream));
        EMIDA_HIP(hipStreamWaitEvent(h->compute, h->ev_done[0], 0));
        EMIDA_HIP(hipMemcpyAsync(h->d_in[0], pic, (size_t)h->src.area() * h->batch * h->in_bytes(),
                                 hipMemcpyHostToDevice, h->compute));
        h->run(h->d_in[0], h->batch, o.maxes, o.neigh, o.offsets, o.coefs, h->compute);
        EMIDA_HIP(hipMemcpyAsync(offsets, o.offsets, h->total * 2 * sizeof(double), hipMemcpyDeviceToHost, h->compute));
        if (coefs)
            EMIDA_HIP(hipMemcpyAsync(coefs, o.coefs, h->total * 6 * sizeof(double), hipMemcpyDeviceToHost, h->compute));
        EMIDA_HIP(hipEventRecord(h->ev_done[0], h->compute));
        EMIDA_HIP(hipStreamSynchronize(h->compute));
    });
}

int emida_offset_process_device(emida_offset* h, const void* d_pics, uint32_t n_patterns, void* d_maxes,
                                void* d_neighbors, double* d_offsets, double* d_coefs, void* stream) {
    return guarded([&] {
        if (!h || !d_pics) throw arg_error("null argument");
        h->require_allocated();
        if (n_patterns == 0 || n_patterns > h->batch) throw arg_error("n_patterns must be in 1..batch_size");
        hipStream_t st = stream ? (hipStream_t)stream : h->compute;
        h->run(d_pics, n_patterns, d_maxes, d_neighbors, d_offsets, d_coefs, st);
    });
}

int emida_offset_set_mfma_mode(emida_offset* h, emida_mfma_mode mode) {
    return guarded([&] {
        if (!h) throw arg_error("null object");
        if (h->allocated) throw state_error("set_mfma_mode must precede allocate_memory");
        with_HT(mode, [](auto) {});
        h->mfma_mode = mode;
    });
}

int emida_offset_set_timing(emida_offset* h, int enable) {
    return guarded([&] {
        if (!h) throw arg_error("null object");
        h->timing = enable != 0;
    });
}

int emida_offset_get_timing(emida_offset* h, const char** kernel_name, double* total_ms, uint64_t* launches) {
    return guarded([&] {
        if (!h) throw arg_error("null object");
        for (auto& p : h->tev) {
            EMIDA_HIP(hipEventSynchronize(p.second));
            float ms = 0;
            EMIDA_HIP(hipEventElapsedTime(&ms, p.first, p.second));
            h->t_ms += ms;
            h->t_launches += 1;
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
        h->tev.clear();
        if (kernel_name)
            *kernel_name = h->policy == EMIDA_CROSS_POLICY_FFT
                               ? "fft_kernel"
                               : (h->policy == EMIDA_CROSS_POLICY_BRUTE_MFMA ? "mfma_corr_kernel" : "cross_corr_kernel");
        if (total_ms) *total_ms = h->t_ms;
        if (launches) *launches = h->t_launches;
    });
}

// ------------------------------------------------------------------ memory helpers
int emida_host_alloc(void** ptr, size_t bytes) {
    return guarded([&] {
        if (!ptr) throw arg_error("null argument");
        EMIDA_HIP(hipHostMalloc(ptr, bytes ? bytes : 16, hipHostMallocDefault));
    });
}
int emida_host_free(void* ptr) {
    return guarded([&] { EMIDA_HIP(hipHostFree(ptr)); });
}
int emida_device_alloc(void** ptr, size_t bytes) {
    return guarded([&] {
        if (!ptr) throw arg_error("null argument");
        *ptr = dmalloc(bytes);
    });
}
int emida_device_free(void* ptr) {
    return guarded([&] { EMIDA_HIP(hipFree(ptr)); });
}
int emida_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    return guarded([&] { EMIDA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int emida_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    return guarded([&] { EMIDA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int emida_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    return guarded([&] { EMIDA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice)); });
}
int emida_memset_device(void* dst, int value, size_t bytes) {
    return guarded([&] { EMIDA_HIP(hipMemset(dst, value, bytes)); });
}
int emida_device_synchronize(void) {
    return guarded([&] { EMIDA_HIP(hipDeviceSynchronize()); });
}
int emida_device_query(char* name, size_t name_len, int* cu_count) {
    return guarded([&] {
        int dev = 0;
        EMIDA_HIP(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        EMIDA_HIP(hipGetDeviceProperties(&prop, dev));
        if (name && name_len) {
            std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
        }
        if (cu_count) *cu_count = prop.multiProcessorCount;
    });
}

// ------------------------------------------------------------------ kernel-level seam
int emida_run_sum(emida_type IN, const void* data, void* sums, const emida_size2* begins, emida_size2 src_size,
                  emida_size2 slice_size, uint32_t begins_size, uint32_t batch_size, void* stream) {
    return guarded([&] {
        if (!data || !sums || !begins) throw arg_error("null argument");
        with_IN(IN, [&](auto in, auto sm) {
            using IT = decltype(in);
            using ST = decltype(sm);
            launch_sum<IT, ST>((const IT*)data, (ST*)sums, (const size2*)begins, S2(src_size), S2(slice_size),
                               begins_size, begins_size * batch_size, (hipStream_t)stream);
        });
    });
}

int emida_run_prepare_pics(emida_type IN, emida_type T, emida_type S, const void* pic, void* slices,
                           const void* hanning_x, const void* hanning_y, const void* sums,
                           const emida_size2* begins, emida_size2 src_size, emida_size2 slice_size,
                           emida_size2 out_size, uint32_t begins_size, uint32_t batch_size, void* stream) {
    return guarded([&] {
        if (!pic || !slices || !hanning_x || !hanning_y || !sums || !begins) throw arg_error("null argument");
        if (out_size.x < slice_size.x || out_size.y < slice_size.y) throw arg_error("out_size smaller than slice");
        const size2 a = S2(src_size), b = S2(slice_size), o = S2(out_size);
        const uint32_t total = begins_size * batch_size;
        hipStream_t st = (hipStream_t)stream;
        const size2* bg = (const size2*)begins;
#define PREP(IT, TT, ST)                                                                                   \
    launch_prepare<IT, TT, ST>((const IT*)pic, (TT*)slices, (const TT*)hanning_x, (const TT*)hanning_y,   \
                               (const ST*)sums, bg, a, b, o, begins_size, total, st)
        if (T == EMIDA_F64 && IN == EMIDA_U16 && S == EMIDA_U16) PREP(uint16_t, double, uint32_t);
        else if (T == EMIDA_F64 && IN == EMIDA_U16 && S == EMIDA_F64) PREP(uint16_t, double, double);
        else if (T == EMIDA_F32 && IN == EMIDA_U16 && S == EMIDA_U16) PREP(uint16_t, float, uint32_t);
        else if (T == EMIDA_F32 && IN == EMIDA_U16 && S == EMIDA_F32) PREP(uint16_t, float, float);
        else if (T == EMIDA_F64 && IN == EMIDA_F64 && S == EMIDA_F64) PREP(double, double, double);
        else if (T == EMIDA_F64 && IN == EMIDA_F64 && S == EMIDA_U16) PREP(double, double, uint32_t);
        else if (T == EMIDA_F32 && IN == EMIDA_F64 && S == EMIDA_F64) PREP(double, float, double);
        else if (T == EMIDA_F32 && IN == EMIDA_F32 && S == EMIDA_F32) PREP(float, float, float);
        else if (T == EMIDA_F64 && IN == EMIDA_F32 && S == EMIDA_F32) PREP(float, double, float);
        else throw arg_error("unsupported (IN, T, S) combination");
#undef PREP
    });
}

int emida_run_cross_corr(emida_type T, const void* pics, const void* ref, void* res, emida_size2 size,
                         emida_size2 res_size, uint32_t ref_slices, uint32_t batch_size, void* stream) {
    return guarded([&] {
        if (!pics || !ref || !res) throw arg_error("null argument");
        hipStream_t st = (hipStream_t)stream;
        if (T == EMIDA_F32)
            launch_cross_corr<float>((const float*)pics, (const float*)ref, (float*)res, S2(size), S2(res_size),
                                     ref_slices, batch_size, st);
        else if (T == EMIDA_F64)
            launch_cross_corr<double>((const double*)pics, (const double*)ref, (double*)res, S2(size), S2(res_size),
                                      ref_slices, batch_size, st);
        else if (T == EMIDA_I32)
            launch_cross_corr<int32_t>((const int32_t*)pics, (const int32_t*)ref, (int32_t*)res, S2(size),
                                       S2(res_size), ref_slices, batch_size, st);
        else
            throw arg_error("T must be F32, F64 or I32");
    });
}

int emida_run_cross_corr_fft(emida_type T, const void* pics, const void* ref, void* res, emida_size2 size,
                             uint32_t ref_slices, uint32_t batch_size, void* stream) {
    return guarded([&] {
        if (!pics || !ref || !res) throw arg_error("null argument");
        if (size.x != size.y) throw unsupported_error("cross_corr_fft: square slices only");
        with_T(T, [&](auto t) {
            using TT = decltype(t);
            launch_fft_corr_prepared<TT>((const TT*)pics, (const TT*)ref, (TT*)res, size.x, ref_slices, batch_size,
                                         (hipStream_t)stream);
        });
    });
}

int emida_run_cross_corr_mfma(emida_mfma_mode mode, const float* pics, const float* ref, float* res,
                              emida_size2 size, emida_size2 res_size, uint32_t ref_slices, uint32_t batch_size,
                              void* stream) {
    return guarded([&] {
        if (!pics || !ref || !res) throw arg_error("null argument");
        check_mfma_shape(S2(size), S2(res_size));
        hipStream_t st = (hipStream_t)stream;
        const size_t area = (size_t)size.x * size.y, np = area * ref_slices * batch_size, nr = area * ref_slices;
        if (np == 0) return;
        const bool split = mfma_split(mode);
        with_HT(mode, [&](auto ht) {
            using HT = typename decltype(ht)::type;
            HT *ph = nullptr, *pl = nullptr, *rh = nullptr, *rl = nullptr;
            EMIDA_HIP(hipMallocAsync((void**)&ph, np * sizeof(HT), st));
            EMIDA_HIP(hipMallocAsync((void**)&rh, nr * sizeof(HT), st));
            if (split) {
                EMIDA_HIP(hipMallocAsync((void**)&pl, np * sizeof(HT), st));
                EMIDA_HIP(hipMallocAsync((void**)&rl, nr * sizeof(HT), st));
            }
            launch_split<HT>(pics, ph, pl, np, st);
            launch_split<HT>(ref, rh, rl, nr, st);
            launch_mfma_corr<HT>(ph, pl, rh, rl, res, S2(size), S2(res_size), ref_slices, batch_size, st);
            for (HT* p : {ph, pl, rh, rl})
                if (p) EMIDA_HIP(hipFreeAsync(p, st));
        });
    });
}

int emida_run_hadamard(emida_type T, void* pics, const void* ref, emida_size2 one_size, uint32_t ref_slices,
                       uint32_t batch_size, void* stream) {
    return guarded([&] {
        if (!pics || !ref) throw arg_error("null argument");
        with_T(T, [&](auto t) {
            using TT = decltype(t);
            launch_hadamard<TT>((cplx<TT>*)pics, (const cplx<TT>*)ref, S2(one_size), ref_slices, batch_size,
                                (hipStream_t)stream);
        });
    });
}

int emida_run_maxarg_reduce(emida_type T, const void* data, void* maxarg, emida_size2 size, uint32_t batch_size,
                            void* stream) {
    return guarded([&] {
        if (!data || !maxarg) throw arg_error("null argument");
        with_T(T, [&](auto t) {
            using TT = decltype(t);
            launch_maxarg<TT>((const TT*)data, size.x * size.y, (data_index<TT>*)maxarg, batch_size,
                              (hipStream_t)stream);
        });
    });
}

int emida_run_extract_neighbors(emida_type T, emida_cross_policy pos_policy, const void* data, size_t data_len,
                                const void* max_i, void* neighbors, int s, emida_size2 src_size,
                                uint32_t batch_size, void* stream) {
    return guarded([&] {
        if (!data || !max_i || !neighbors) throw arg_error("null argument");
        if (s < 1 || s > 9) throw arg_error("neighbourhood size must be 1..9");
        const int pol = pos_policy == EMIDA_CROSS_POLICY_FFT ? POLICY_FFT : POLICY_BRUTE;
        with_T(T, [&](auto t) {
            using TT = decltype(t);
            launch_extract<TT>((const TT*)data, data_len, (const data_index<TT>*)max_i, (TT*)neighbors, s,
                               S2(src_size), pol, batch_size, (hipStream_t)stream);
        });
    });
}

int emida_run_subpixel_max(emida_type T, const void* neighbors, int s, uint32_t n, double* offsets,
                           double* coefs, void* stream) {
    return guarded([&] {
        if (!neighbors) throw arg_error("null argument");
        with_T(T, [&](auto t) {
            using TT = decltype(t);
            launch_subpixel<TT>((const TT*)neighbors, s, n, nullptr, size2{0, 0}, POLICY_BRUTE, offsets, coefs,
                                (hipStream_t)stream);
        });
    });
}

int emida_lstsq_matrix(int s, double* out) {
    return guarded([&] {
        if (!out) throw arg_error("null argument");
        (void)lstsq_offset(s);
        lstsq_host(s, out);
    });
}

int emida_generate_hanning(emida_type T, size_t size, void* out) {
    return guarded([&] {
        if (!out) throw arg_error("null argument");
        with_T(T, [&](auto t) {
            using TT = decltype(t);
            auto w = hanning<TT>(size);
            std::memcpy(out, w.data(), size * sizeof(TT));
        });
    });
}

}  // extern "C"
