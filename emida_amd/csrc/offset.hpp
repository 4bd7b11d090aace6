// Host-side object behind the C ABI, shared by the C ABI files: emida_offset, one
// configuration of the pattern-offset pipeline with its resident reference data,
// plus the error plumbing.
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
#pragma once
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

namespace emida_amd {

inline thread_local std::string g_last_error;

struct state_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline int fail(int code, const std::string& msg) {
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

inline size_t type_size(emida_type t) {
    switch (t) {
        case EMIDA_F32: return 4;
        case EMIDA_F64: return 8;
        case EMIDA_I32: return 4;
        case EMIDA_U16: return 2;
    }
    throw arg_error("unknown element type");
}

inline size2 S2(emida_size2 s) { return {s.x, s.y}; }

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
inline bool mfma_split(emida_mfma_mode mode) { return mode == EMIDA_MFMA_BF16_SPLIT || mode == EMIDA_MFMA_F16_SPLIT; }

inline void* dmalloc(size_t bytes) {
    void* p = nullptr;
    if (bytes == 0) bytes = 16;
    EMIDA_HIP(hipMalloc(&p, bytes));
    return p;
}

}  // namespace emida_amd

using namespace emida_amd;

struct emida_offset_s {
    emida_type T, IN;
    emida_cross_policy policy;
    int fit;
    size2 src, slice, cross, buf;
    uint32_t nb, batch, total, n_in;
    uint32_t side = 0;  // FFT slice side (x)
    bool fftg = false;  // FFT policy through the three-kernel path (fftg.hip)
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
    void* d_corr = nullptr;    // brute: correlation buffers
    emida_mfma_mode mfma_mode = EMIDA_MFMA_AUTO;  // as set; mfma_mode_used() is the one in effect
    bool mfma_refine = true;   // matrix-core policy: exact peak from T-precision values
    void* d_ref_lo = nullptr;  // matrix-core policy: lo halves of the reference slices
    void* d_ref_t = nullptr;   // matrix-core policy with refinement: T-precision reference slices
    void* d_hxf = nullptr;     // matrix-core policy, T = double: float windows for the halves
    void* d_hyf = nullptr;
    void* d_pic_hi = nullptr;  // matrix-core policy: prepared pattern slices, hi / lo halves
    void* d_pic_lo = nullptr;
    void* d_smax = nullptr;    // matrix-core policy: screen maximum per slice row
    // matrix-core policy with refinement: the candidates published in place of the
    // screen (screen_lists) and the norms of the T-precision slices
    void* d_cnt = nullptr;
    void* d_ent = nullptr;
    void* d_pn2 = nullptr;
    void* d_rn2 = nullptr;
    // the reference's cross_corr_fft seam (get_cu_pic / get_cu_temp / cross_corr_fft,
    // src/gpu_offset.hpp:303-325), FFT policy: made on first use
    void* d_seam_pic = nullptr;   // total x (2s.y rows of 2s.x) T
    void* d_seam_temp = nullptr;  // nb x (2s.y rows of s.x + 1) complex T: R2C layout
    void* d_seam_res = nullptr;   // total x (2s.y rows of 2s.x) T
    void* d_seam_work = nullptr;
    size_t seam_chunk = 0;
    std::vector<unsigned char> h_ref;  // FFT policy: the reference pattern, for the seam's spectra

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

    // Per-batch working memory of run(): slice sums, prepared slices (or their
    // half-type planes) and correlation buffers, for a batch of up to n patterns.
    struct scratch_t {
        void* sums = nullptr;
        void* slices = nullptr;
        void* corr = nullptr;
        void* pic_hi = nullptr;
        void* pic_lo = nullptr;
        void* smax = nullptr;
        void* cnt = nullptr;
        void* ent = nullptr;
        void* pn2 = nullptr;
    };
    static constexpr int kParts = 9;
    static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
    size_t scratch_sizes(uint32_t n_pat, size_t* parts) const {
        const size_t tn = (size_t)nb * n_pat;
        parts[0] = al256(std::max<size_t>(tn, nb) * sum_bytes);
        for (int i = 1; i < kParts; ++i) parts[i] = 0;
        if (policy == EMIDA_CROSS_POLICY_BRUTE) {
            parts[1] = al256(tn * slice.area() * t_bytes());
            parts[2] = al256(tn * cross.area() * t_bytes());
        } else if (policy == EMIDA_CROSS_POLICY_BRUTE_MFMA) {
            if (mfma_refine) {
                parts[1] = al256(tn * slice.area() * t_bytes());
                parts[6] = al256(tn * cross.y * sizeof(uint32_t));
                parts[7] = al256(tn * cross.y * kListCap * sizeof(cand_entry));
                parts[8] = al256(tn * 2 * sizeof(double));  // norms, then zero-shift products
            } else {
                parts[2] = al256(tn * cross.area() * sizeof(float));  // the screen
            }
            if (IN != EMIDA_U16) {  // uint16 patterns are prepared by the matrix-core kernel itself
                parts[3] = al256(tn * slice.area() * 2);
                if (mfma_split(mfma_mode_used())) parts[4] = parts[3];
            }
            parts[5] = al256(tn * cross.y * sizeof(uint32_t));
        } else if (fftg) {
            parts[2] = al256(fftg_work_bytes(T == EMIDA_F64, slice, (uint32_t)tn, true));  // half spectra + corr
        }
        size_t sum = 0;
        for (int i = 0; i < kParts; ++i) sum += parts[i];
        return sum;
    }
    size_t scratch_bytes(uint32_t n_pat) const {
        size_t parts[kParts];
        return scratch_sizes(n_pat, parts);
    }
    scratch_t carve(void* base, uint32_t n_pat) const {
        size_t parts[kParts];
        scratch_sizes(n_pat, parts);
        char* b = (char*)base;
        void* ptr[kParts];
        for (int i = 0; i < kParts; ++i) {
            ptr[i] = parts[i] ? b : nullptr;
            b += parts[i];
        }
        return {ptr[0], ptr[1], ptr[2], ptr[3], ptr[4], ptr[5], ptr[6], ptr[7], ptr[8]};
    }
    const void* float_hx() const { return d_hxf ? d_hxf : d_hx; }
    const void* float_hy() const { return d_hyf ? d_hyf : d_hy; }
    scratch_t own_scratch() const {
        return {d_sums, d_slices, d_corr, d_pic_hi, d_pic_lo, d_smax, d_cnt, d_ent, d_pn2};
    }

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
        void* ptrs[] = {d_begins, d_hx,     d_hy,     d_sums,  d_ref, d_slices, d_corr,
                        d_ref_lo, d_pic_hi, d_pic_lo, d_ref_t, d_hxf, d_hyf,    d_smax,
                        d_cnt,    d_ent,    d_pn2,    d_rn2,
                        d_seam_pic, d_seam_temp, d_seam_res, d_seam_work};
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

    // Upload one batch of pictures into the next input buffer (round robin), after
    // the kernels that read its previous contents finished and get_offset_core
    // consumed it; `staged`, when given, is recorded once the host buffer was read.
    uint32_t upload(const void* pic, hipEvent_t staged = nullptr) {
        require_allocated();
        std::unique_lock<std::mutex> lock(upload_mu);
        const uint32_t i = rr.fetch_add(1) % n_in;
        upload_cv.wait(lock, [&] { return !pending[i]; });
        pending[i] = 1;
        try {
            EMIDA_HIP(hipStreamWaitEvent(in_stream, ev_done[i], 0));
            EMIDA_HIP(hipMemcpyAsync(d_in[i], pic, (size_t)src.area() * batch * in_bytes(), hipMemcpyHostToDevice,
                                     in_stream));
            EMIDA_HIP(hipEventRecord(ev_up[i], in_stream));
            if (staged) EMIDA_HIP(hipEventRecord(staged, in_stream));
        } catch (...) {
            pending[i] = 0;
            upload_cv.notify_all();
            throw;
        }
        return i;
    }

    // Give an uploaded buffer back without computing on it (error paths).
    void release(uint32_t i) {
        {
            std::lock_guard<std::mutex> lock(upload_mu);
            if (i < n_in) pending[i] = 0;
        }
        upload_cv.notify_all();
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
                    with_HT(mfma_mode_used(), [&](auto ht) {
                        using HT = typename decltype(ht)::type;
                        launch_prepare_split<IT, ST, HT>(pic, (HT*)d_ref, (HT*)d_ref_lo, (const float*)float_hx(),
                                                         (const float*)float_hy(), sums, d_begins, src, slice, nb, nb,
                                                         compute);
                    });
                    if (mfma_refine) {
                        launch_prepare<IT, TT, ST>(pic, (TT*)d_ref_t, (const TT*)d_hx, (const TT*)d_hy, sums,
                                                   d_begins, src, slice, slice, nb, nb, compute);
                        launch_slice_stats<TT>((const TT*)d_ref_t, nullptr, slice.area(), nb, nb, (double*)d_rn2,
                                               compute);
                    }
                } else {
                    if (!std::is_same<IT, uint16_t>::value)
                        launch_sum<IT, ST>(pic, sums, d_begins, src, slice, nb, nb, compute);
                    if (fftg) {
                        void* work = dmalloc(fftg_ref_work_bytes(T == EMIDA_F64, slice, nb));
                        try {
                            launch_fftg_ref<TT, IT, ST>(pic, sums, (const TT*)d_hx, (const TT*)d_hy, d_begins, src,
                                                        slice, nb, (cplx<TT>*)d_ref, work, compute);
                            EMIDA_HIP(hipStreamSynchronize(compute));
                        } catch (...) {
                            (void)hipStreamSynchronize(compute);
                            (void)hipFree(work);
                            throw;
                        }
                        EMIDA_HIP(hipFree(work));
                    } else {
                        launch_fft_ref<TT, IT, ST>(pic, sums, (const TT*)d_hx, (const TT*)d_hy, d_begins, src, side,
                                                   nb, (cplx<TT>*)d_ref, compute);
                    }
                }
            });
        });
        EMIDA_HIP(hipStreamSynchronize(compute));
        EMIDA_HIP(hipFree(d_temp));
    }

    // The matrix-core mode in effect: the one set, or for EMIDA_MFMA_AUTO the plain
    // half type when the peak is refined (exact results in every mode; one product in
    // three, and a third of the fp32 accumulations, which keeps the screen's bound
    // smaller than a split mode's) and the split one when the screen values are the
    // results; fp16 for uint16 patterns, bf16 (float's range) for float and double.
    emida_mfma_mode mfma_mode_used() const {
        if (mfma_mode != EMIDA_MFMA_AUTO) return mfma_mode;
        const bool u16 = IN == EMIDA_U16;
        if (mfma_refine) return u16 ? EMIDA_MFMA_F16 : EMIDA_MFMA_BF16;
        return u16 ? EMIDA_MFMA_F16_SPLIT : EMIDA_MFMA_BF16_SPLIT;
    }

    // resident_only: the reference data alone (for callers that bring their own
    // input, scratch and result buffers, such as the request server)
    void allocate(const void* temp, bool resident_only = false) {
        if (allocated) throw state_error("allocate_memory was already called");
        EMIDA_HIP(hipGetDevice(&device));
        EMIDA_HIP(hipStreamCreateWithFlags(&in_stream, hipStreamNonBlocking));
        EMIDA_HIP(hipStreamCreateWithFlags(&compute, hipStreamNonBlocking));
        EMIDA_HIP(hipStreamCreateWithFlags(&out_stream, hipStreamNonBlocking));
        const size_t pic_bytes = (size_t)src.area() * batch * in_bytes();
        for (uint32_t i = 0; i < (resident_only ? 0u : n_in); ++i) {
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
        if (policy == EMIDA_CROSS_POLICY_BRUTE_MFMA && T == EMIDA_F64) {
            auto hx = hanning<float>(slice.x), hy = hanning<float>(slice.y);
            d_hxf = alloc(hx.size() * sizeof(float));
            d_hyf = alloc(hy.size() * sizeof(float));
            EMIDA_HIP(hipMemcpy(d_hxf, hx.data(), hx.size() * sizeof(float), hipMemcpyHostToDevice));
            EMIDA_HIP(hipMemcpy(d_hyf, hy.data(), hy.size() * sizeof(float), hipMemcpyHostToDevice));
        }
        d_sums = alloc((resident_only ? nb : std::max(total, nb)) * sum_bytes);
        const size_t scratch_total = resident_only ? 0 : total;
        if (policy == EMIDA_CROSS_POLICY_BRUTE) {
            d_ref = alloc((size_t)nb * slice.area() * t_bytes());
            if (!resident_only) {
                d_slices = alloc(scratch_total * slice.area() * t_bytes());
                d_corr = alloc(scratch_total * cross.area() * t_bytes());
            }
        } else if (policy == EMIDA_CROSS_POLICY_BRUTE_MFMA) {
            const size_t half = 2;  // bf16 / fp16
            d_ref = alloc((size_t)nb * slice.area() * half);
            if (mfma_split(mfma_mode_used())) d_ref_lo = alloc((size_t)nb * slice.area() * half);
            if (mfma_refine) {
                d_ref_t = alloc((size_t)nb * slice.area() * t_bytes());
                d_rn2 = alloc((size_t)nb * sizeof(double));
            }
            if (!resident_only) {
                if (mfma_refine) {
                    d_slices = alloc(scratch_total * slice.area() * t_bytes());
                    d_cnt = alloc(scratch_total * cross.y * sizeof(uint32_t));
                    d_ent = alloc(scratch_total * cross.y * kListCap * sizeof(cand_entry));
                    d_pn2 = alloc(scratch_total * 2 * sizeof(double));  // norms, then zero-shift products
                } else {
                    d_corr = alloc(scratch_total * cross.area() * sizeof(float));
                }
                if (IN != EMIDA_U16) {
                    d_pic_hi = alloc(scratch_total * slice.area() * half);
                    if (mfma_split(mfma_mode_used())) d_pic_lo = alloc(scratch_total * slice.area() * half);
                }
                d_smax = alloc(scratch_total * cross.y * sizeof(uint32_t));
            }
        } else if (fftg) {
            d_ref = alloc(fftg_spectrum_bytes(T == EMIDA_F64, slice) * nb);
            if (!resident_only) d_corr = alloc(fftg_work_bytes(T == EMIDA_F64, slice, total, true));
        } else {
            const size_t m = fft_padded(side);
            const size_t spec = (m / 2 + 1) * m * 2 * t_bytes();
            d_ref = alloc(spec * nb);
        }
        const size_t nk = (size_t)fit * fit;
        for (auto& o : out) {
            if (resident_only) break;
            o.maxes = alloc(total * ti_bytes());
            o.neigh = alloc(total * nk * t_bytes());
            o.offsets = (double*)alloc(total * 2 * sizeof(double));
            o.coefs = (double*)alloc(total * 6 * sizeof(double));
            EMIDA_HIP(hipEventCreateWithFlags(&o.copied, hipEventDisableTiming));
        }
        (void)lstsq_device_table();  // the device tables, before any batch (or graph capture)
        (void)fft_twiddles<float>();
        if (temp) prepare_reference(temp);
        if (temp && policy == EMIDA_CROSS_POLICY_FFT && !resident_only)
            h_ref.assign((const unsigned char*)temp, (const unsigned char*)temp + (size_t)src.area() * in_bytes());
        allocated = true;
    }

    // ---- the reference's cross_corr_fft seam, FFT policy (src/gpu_offset.hpp:303-325)
    void seam_init() {
        require_allocated();
        if (policy != EMIDA_CROSS_POLICY_FFT) throw state_error("cross_corr_fft belongs to the FFT policy");
        if (d_seam_pic) return;
        const bool dbl = T == EMIDA_F64;
        if (fft_padded(slice.x) != 2 * slice.x || fft_padded(slice.y) != 2 * slice.y)
            throw unsupported_error("cross_corr_fft, get_cu_pic, get_cu_temp: 2s must be a power of two per axis");
        const size_t buf = (size_t)4 * slice.area() * t_bytes();
        const size_t spec = (size_t)(slice.x + 1) * 2 * slice.y * 2 * t_bytes();
        seam_chunk = fftg_chunk(dbl, slice, std::max(total, nb), 0);
        d_seam_pic = alloc(buf * total);
        d_seam_temp = alloc(spec * nb);
        d_seam_res = alloc(buf * std::max(total, nb));
        d_seam_work = alloc(seam_chunk * fftg_slice_work_bytes(dbl, slice));
        EMIDA_HIP(hipMemsetAsync(d_seam_pic, 0, buf * total, compute));
        EMIDA_HIP(hipMemsetAsync(d_seam_temp, 0, spec * nb, compute));
        if (!h_ref.empty()) {
            // R2C of the prepared reference, zero-padded to 2s.x x 2s.y (the reference's
            // in-place transform of cu_temp_, src/gpu_offset.hpp:185-210)
            void* d_temp = dmalloc(h_ref.size());
            try {
                EMIDA_HIP(hipMemcpyAsync(d_temp, h_ref.data(), h_ref.size(), hipMemcpyHostToDevice, compute));
                EMIDA_HIP(hipMemsetAsync(d_seam_res, 0, buf * nb, compute));
                with_T(T, [&](auto t) {
                    using TT = decltype(t);
                    with_IN(IN, [&](auto in, auto sm) {
                        using IT = decltype(in);
                        using ST = decltype(sm);
                        launch_sum<IT, ST>((const IT*)d_temp, (ST*)d_sums, d_begins, src, slice, nb, nb, compute);
                        launch_prepare<IT, TT, ST>((const IT*)d_temp, (TT*)d_seam_res, (const TT*)d_hx,
                                                   (const TT*)d_hy, (const ST*)d_sums, d_begins, src, slice,
                                                   size2{2 * slice.x, 2 * slice.y}, nb, nb, compute);
                        launch_fftg_std_r2c<TT>((const TT*)d_seam_res, 2 * slice.x, (cplx<TT>*)d_seam_temp, slice, nb,
                                                d_seam_work, seam_chunk, compute);
                    });
                });
                EMIDA_HIP(hipStreamSynchronize(compute));
            } catch (...) {
                (void)hipStreamSynchronize(compute);
                (void)hipFree(d_temp);
                throw;
            }
            EMIDA_HIP(hipFree(d_temp));
        }
        EMIDA_HIP(hipStreamSynchronize(compute));
    }

    void* cu_pic() {
        require_allocated();
        if (policy == EMIDA_CROSS_POLICY_FFT) {
            seam_init();
            return d_seam_pic;
        }
        if (!d_slices) throw state_error("get_cu_pic: this configuration keeps no prepared pattern slices");
        return d_slices;
    }
    void* cu_temp() {
        require_allocated();
        if (policy == EMIDA_CROSS_POLICY_FFT) {
            seam_init();
            return d_seam_temp;
        }
        void* p = policy == EMIDA_CROSS_POLICY_BRUTE ? d_ref : d_ref_t;
        if (!p) throw state_error("get_cu_temp: this configuration keeps no T-precision reference slices");
        return p;
    }
    void cross_corr_fft() {
        seam_init();
        with_T(T, [&](auto t) {
            using TT = decltype(t);
            launch_fftg_std_corr<TT>((const TT*)d_seam_pic, (const cplx<TT>*)d_seam_temp, (TT*)d_seam_res, slice, nb,
                                     total, d_seam_work, seam_chunk, compute);
        });
        EMIDA_HIP(hipStreamSynchronize(compute));
    }

    // Sums and preparation of a batch in one kernel for uint16 patterns (uint32 sums);
    // false when the input type or the slice size needs the separate kernels.
    template <class IT, class ST, class TT, class HT>
    bool fused_prepare(const IT* pics, ST* sums, TT* slices, HT* hi, HT* lo, double* stats, uint32_t total,
                       hipStream_t st) {
        if constexpr (std::is_same<IT, uint16_t>::value && std::is_same<ST, uint32_t>::value)
            return launch_sum_prepare<TT, HT>(pics, sums, slices, (const TT*)d_hx, (const TT*)d_hy, hi, lo,
                                              (const float*)float_hx(), (const float*)float_hy(),
                                              hi ? mfma_in_scale<HT>() : 1.0f, stats, (const TT*)d_ref_t, d_begins,
                                              src, slice, nb, total, st);
        else
            return false;
    }

    // ---- one batch on device data (src/gpu_offset.hpp:234-278)
    void run(const void* d_pics, uint32_t n_pat, void* maxes, void* neigh, double* offsets, double* coefs,
             hipStream_t st) {
        run(d_pics, n_pat, maxes, neigh, offsets, coefs, own_scratch(), st);
    }
    void run(const void* d_pics, uint32_t n_pat, void* maxes, void* neigh, double* offsets, double* coefs,
             const scratch_t& sc, hipStream_t st) {
        const uint32_t tn = nb * n_pat;
        hipEvent_t t0 = nullptr, t1 = nullptr;
        with_T(T, [&](auto t) {
            using TT = decltype(t);
            with_IN(IN, [&](auto in, auto sm) {
                using IT = decltype(in);
                using ST = decltype(sm);
                const IT* pics = (const IT*)d_pics;
                ST* sums = (ST*)sc.sums;
                if (policy == EMIDA_CROSS_POLICY_BRUTE) {
                    if (!fused_prepare<IT, ST, TT, _Float16>(pics, sums, (TT*)sc.slices, nullptr, nullptr, nullptr, tn,
                                                             st)) {
                        launch_sum<IT, ST>(pics, sums, d_begins, src, slice, nb, tn, st);
                        launch_prepare<IT, TT, ST>(pics, (TT*)sc.slices, (const TT*)d_hx, (const TT*)d_hy, sums,
                                                   d_begins, src, slice, slice, nb, tn, st);
                    }
                    if (timing) t0 = record_start(st);
                    launch_cross_corr<TT>((const TT*)sc.slices, (const TT*)d_ref, (TT*)sc.corr, slice, cross, nb, n_pat,
                                          st);
                    if (timing) record_stop(st, t0);
                    launch_peak<TT>((const TT*)sc.corr, (size_t)tn * cross.area(), cross.area(), cross, POLICY_BRUTE,
                                    fit, (data_index<TT>*)maxes, (TT*)neigh, offsets, coefs, tn, st);
                } else if (policy == EMIDA_CROSS_POLICY_BRUTE_MFMA) {
                    // with refinement the kernel publishes the candidates (screen_lists)
                    // instead of the screen
                    screen_lists lists{};
                    if (mfma_refine) {
                        lists.cnt = (uint32_t*)sc.cnt;
                        lists.ent = (cand_entry*)sc.ent;
                        lists.pn2 = (const double*)sc.pn2;
                        lists.rn2 = (const double*)d_rn2;
                        lists.z0 = (const double*)sc.pn2 + tn;
                        mfma_refine_bounds(mfma_mode_used(), slice, T == EMIDA_F64, lists.tol_rel, lists.tol_abs);
                    }
                    with_HT(mfma_mode_used(), [&](auto ht) {
                        using HT = typename decltype(ht)::type;
                        double* pn2 = mfma_refine ? (double*)sc.pn2 : nullptr;
                        // uint16 patterns: the matrix-core kernel prepares the pattern rows as
                        // it stages them, so only the sums (and, for the refinement, the T
                        // slices and their statistics) are made here
                        constexpr bool u16 = std::is_same<IT, uint16_t>::value && std::is_same<ST, uint32_t>::value;
                        HT* phi = u16 ? nullptr : (HT*)sc.pic_hi;
                        HT* plo = u16 ? nullptr : (HT*)sc.pic_lo;
                        if (!fused_prepare<IT, ST, TT, HT>(pics, sums, mfma_refine ? (TT*)sc.slices : nullptr, phi, plo,
                                                           pn2, tn, st)) {
                            launch_sum<IT, ST>(pics, sums, d_begins, src, slice, nb, tn, st);
                            if (!u16)
                                launch_prepare_split<IT, ST, HT>(pics, phi, plo, (const float*)float_hx(),
                                                                 (const float*)float_hy(), sums, d_begins, src, slice,
                                                                 nb, tn, st);
                            if (mfma_refine) {
                                launch_prepare<IT, TT, ST>(pics, (TT*)sc.slices, (const TT*)d_hx, (const TT*)d_hy,
                                                           sums, d_begins, src, slice, slice, nb, tn, st);
                                launch_slice_stats<TT>((const TT*)sc.slices, (const TT*)d_ref_t, slice.area(), nb,
                                                       tn, pn2, st);
                            }
                        }
                        const mfma_src msrc{(const uint16_t*)pics, (const uint32_t*)sums, (const float*)float_hx(),
                                            (const float*)float_hy(), d_begins, src};
                        if (timing) t0 = record_start(st);
                        launch_mfma_corr<HT>(phi, plo, (const HT*)d_ref, (const HT*)d_ref_lo,
                                             mfma_refine ? nullptr : (float*)sc.corr, slice, cross, nb, n_pat, st,
                                             (uint32_t*)sc.smax, mfma_refine ? &lists : nullptr, u16 ? &msrc : nullptr);
                        if (timing) record_stop(st, t0);
                    });
                    if (mfma_refine) {
                        launch_refine<TT>(nullptr, (size_t)tn * cross.area(), (const TT*)sc.slices,
                                          (const TT*)d_ref_t, slice, cross, nb, lists.tol_rel, lists.tol_abs, fit,
                                          (data_index<TT>*)maxes, (TT*)neigh, offsets, coefs, tn, st,
                                          (const uint32_t*)sc.smax, &lists);
                    } else {
                        launch_peak<float>((const float*)sc.corr, (size_t)tn * cross.area(), cross.area(), cross,
                                           POLICY_BRUTE, fit, (data_index<float>*)maxes, (float*)neigh, offsets,
                                           coefs, tn, st, (const uint32_t*)sc.smax);
                    }
                } else {
                    if (!std::is_same<IT, uint16_t>::value)
                        launch_sum<IT, ST>(pics, sums, d_begins, src, slice, nb, tn, st);
                    if (timing) t0 = record_start(st);
                    if (fftg)
                        launch_fftg_pipe<TT, IT, ST>(pics, sums, (const TT*)d_hx, (const TT*)d_hy, d_begins, src,
                                                     slice, nb, tn, (const cplx<TT>*)d_ref, fit,
                                                     (data_index<TT>*)maxes, (TT*)neigh, offsets, coefs, nullptr,
                                                     sc.corr, st);
                    else
                        launch_fft_pipe<TT, IT, ST>(pics, sums, (const TT*)d_hx, (const TT*)d_hy, d_begins, src,
                                                    side, nb, tn, (const cplx<TT>*)d_ref, fit,
                                                    (data_index<TT>*)maxes, (TT*)neigh, offsets, coefs, nullptr, st);
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

    // Drop the result slot kept for these host buffers before they are freed, so a
    // later buffer at the same address is not mistaken for them by finalize().
    void forget_slot(const void* maxes, const void* neigh) {
        std::lock_guard<std::mutex> lock(slot_mu);
        for (auto& s : slots)
            if (s.maxes == maxes && s.neigh == neigh && s.stamp != 0) {
                EMIDA_HIP(hipEventSynchronize(s.done));
                s.maxes = nullptr;
                s.neigh = nullptr;
                s.stamp = 0;
            }
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
