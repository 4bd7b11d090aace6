// C ABI of the pattern-offset library (declared in include/emida_amd.h).
#include "offset.hpp"

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
        if (policy == EMIDA_CROSS_POLICY_BRUTE_MFMA) {
            check_mfma_shape(S2(slice_size), S2(cross_size));
        }
        if (policy != EMIDA_CROSS_POLICY_FFT) {
            if (cross_size.x == 0 || cross_size.y == 0) throw arg_error("cross size must be positive");
            if (cross_size.x > 2 * slice_size.x - 1 || cross_size.y > 2 * slice_size.y - 1)
                throw arg_error("cross size must be at most 2*slice-1 per axis");
            h->cross = S2(cross_size);
            h->buf = h->cross;
        } else {
            // the FFT policy forces cross_size = 2s - 1 (src/gpu_offset.hpp:114); any
            // side per axis up to kFftMaxSlice, each zero-padded to the next power of
            // two >= 2s (same correlation values)
            if (slice_size.x > kFftMaxSlice || slice_size.y > kFftMaxSlice)
                throw unsupported_error("FFT policy: slice sides must be between 2 and " +
                                        std::to_string(kFftMaxSlice));
            h->side = slice_size.x;
            h->fftg = fftg_needed(T == EMIDA_F64, S2(slice_size));
            h->cross = {2 * slice_size.x - 1, 2 * slice_size.y - 1};
            h->buf = {2 * slice_size.x, 2 * slice_size.y};
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
        info->fft_size = h->side ? fft_padded(h->side) : 0;
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
        const uint32_t i = h->upload(pic);
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

int emida_offset_get_cu_pic(emida_offset* h, void** dptr) {
    return guarded([&] {
        if (!h || !dptr) throw arg_error("null argument");
        *dptr = h->cu_pic();
    });
}

int emida_offset_get_cu_temp(emida_offset* h, void** dptr) {
    return guarded([&] {
        if (!h || !dptr) throw arg_error("null argument");
        *dptr = h->cu_temp();
    });
}

int emida_offset_cross_corr_fft(emida_offset* h) {
    return guarded([&] {
        if (!h) throw arg_error("null object");
        h->cross_corr_fft();
    });
}

int emida_offset_get_cu_cross_res(emida_offset* h, void** dptr) {
    return guarded([&] {
        if (!h || !dptr) throw arg_error("null argument");
        h->seam_init();
        *dptr = h->d_seam_res;
    });
}

int emida_offset_get_streams(emida_offset* h, void** in_stream, void** out_stream) {
    return guarded([&] {
        if (!h) throw arg_error("null object");
        h->require_allocated();
        if (in_stream) *in_stream = (void*)h->in_stream;
        if (out_stream) *out_stream = (void*)h->out_stream;
    });
}

int emida_offset_set_mfma_mode(emida_offset* h, emida_mfma_mode mode) {
    return guarded([&] {
        if (!h) throw arg_error("null object");
        if (h->allocated) throw state_error("set_mfma_mode must precede allocate_memory");
        if (mode != EMIDA_MFMA_AUTO) with_HT(mode, [](auto) {});
        if ((mode == EMIDA_MFMA_F16 || mode == EMIDA_MFMA_F16_SPLIT) && h->IN != EMIDA_U16)
            throw unsupported_error(
                "fp16 matrix-core operands hold prepared values below 131008 in magnitude, which only uint16 "
                "patterns guarantee; use a bf16 mode for float or double patterns");
        h->mfma_mode = mode;
    });
}

int emida_offset_get_mfma_mode(emida_offset* h, emida_mfma_mode* mode) {
    return guarded([&] {
        if (!h || !mode) throw arg_error("null argument");
        *mode = h->mfma_mode_used();
    });
}

int emida_offset_set_mfma_refine(emida_offset* h, int enable) {
    return guarded([&] {
        if (!h) throw arg_error("null object");
        if (h->allocated) throw state_error("set_mfma_refine must precede allocate_memory");
        if (!enable && h->T != EMIDA_F32)
            throw arg_error("the unrefined matrix-core screen gives float results: T must be F32");
        h->mfma_refine = enable != 0;
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
                               ? (h->fftg ? "fftg_rows_kernel+fftg_cols_kernel+fftg_rows_inv_kernel+peak_kernel"
                                          : fft_padded(h->side) > 128 ? "fft256_kernel" : "fft_kernel")
                               : (h->policy == EMIDA_CROSS_POLICY_BRUTE_MFMA
                                      ? "mfma_corr_kernel"
                                      : cross_corr_kernel_name(h->T == EMIDA_F64, h->slice, h->cross, h->nb, h->batch));
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

int emida_device_mem_info(size_t* free_bytes, size_t* total_bytes) {
    return guarded([&] {
        size_t f = 0, t = 0;
        EMIDA_HIP(hipMemGetInfo(&f, &t));
        if (free_bytes) *free_bytes = f;
        if (total_bytes) *total_bytes = t;
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
        with_T(T, [&](auto t) {
            using TT = decltype(t);
            launch_fft_corr_prepared<TT>((const TT*)pics, (const TT*)ref, (TT*)res, S2(size), ref_slices, batch_size,
                                         (hipStream_t)stream);
        });
    });
}

int emida_run_cross_corr_fft_noalloc(emida_type T, const void* pics, const void* ref, void* res, emida_size2 size,
                                     uint32_t ref_slices, uint32_t batch_size, void* stream) {
    return guarded([&] {
        if (!pics || !ref || !res) throw arg_error("null argument");
        with_T(T, [&](auto t) {
            using TT = decltype(t);
            launch_fft_corr_prepared_noalloc<TT>((const TT*)pics, (const TT*)ref, (TT*)res, S2(size), ref_slices,
                                                 batch_size, (hipStream_t)stream);
        });
    });
}

int emida_fft_noalloc_scratch_bytes(size_t* bytes) {
    return guarded([&] {
        if (!bytes) throw arg_error("null argument");
        *bytes = kFftStaticScratchBytes;
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

int emida_fft_max_slice(emida_type T, uint32_t* side) {
    return guarded([&] {
        if (!side) throw arg_error("null argument");
        if (T != EMIDA_F32 && T != EMIDA_F64) throw arg_error("T must be F32 or F64");
        *side = kFftMaxSlice;
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
