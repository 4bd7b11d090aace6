/*
 * emida_amd.h - C ABI of the MI355X (gfx950) pattern-offset library.
 *
 * One `emida_offset` object plays the role of the reference's
 * `emida::gpu_offset<T, IN>` (src/gpu_offset.hpp:48-326): it holds the region of
 * interest (ROI) layout, the prepared reference pattern and the device buffers
 * of one batch shape, and runs the offset pipeline
 *   slice sums -> prepare (mean, Hann) -> cross-correlation (brute | FFT)
 *   -> arg-max -> F x F neighbourhood -> quadratic least-squares peak
 * on the GPU.
 *
 * Conventions
 *  - Every function returns EMIDA_OK (0) or an error code; emida_last_error()
 *    gives the message of the last failure on the calling thread (the reference
 *    prints and throws std::runtime_error, src/host_helpers.hpp:14-25; the C++
 *    wrapper include/emida/gpu_offset.hpp rethrows).
 *  - Sizes are {x = columns, y = rows}; images are row-major.
 *  - "host" pointers may be pageable or pinned (emida_host_alloc); "device"
 *    pointers come from emida_device_alloc or any HIP allocation on the current
 *    device.  `stream` arguments are hipStream_t values passed as void* (NULL
 *    means the object's own compute stream, or the null stream for the
 *    kernel-level functions).
 *  - No call allocates or frees device memory except create/allocate_memory/
 *    destroy and the explicit alloc/free helpers.
 */
#ifndef EMIDA_AMD_H
#define EMIDA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ types */

/* Element types. T (correlation precision) is F32 or F64; IN (pattern pixel
 * type) is U16, F32 or F64; I32 is accepted by emida_run_cross_corr only. */
typedef enum { EMIDA_F32 = 0, EMIDA_F64 = 1, EMIDA_I32 = 2, EMIDA_U16 = 3 } emida_type;

/* cross_policy, src/common.hpp:141-145.  BRUTE_MFMA (no reference counterpart) is
 * the brute-force correlation as an implicit GEMM on BF16/FP16 matrix cores with
 * FP32 accumulation (precision set by emida_offset_set_mfma_mode), refined to the
 * BRUTE policy's results by emida_offset_set_mfma_refine; it requires a slice width
 * that is a multiple of 8 up to 128. */
typedef enum {
    EMIDA_CROSS_POLICY_BRUTE = 0,
    EMIDA_CROSS_POLICY_FFT = 1,
    EMIDA_CROSS_POLICY_BRUTE_MFMA = 2
} emida_cross_policy;

/* Operand precision of EMIDA_CROSS_POLICY_BRUTE_MFMA.  *_SPLIT modes store every
 * prepared value as hi + lo halves and accumulate hi*hi + hi*lo + lo*hi (relative
 * representation error 2^-16 for bf16, 2^-22 for fp16).  fp16 operands are the
 * prepared values halved, so any |value| < 131008 fits: every uint16 input does, so
 * the fp16 modes are accepted for IN = U16 only (emida_offset_set_mfma_mode returns
 * EMIDA_ERR_UNSUPPORTED otherwise).  EMIDA_MFMA_AUTO (the default) takes the plain
 * mode when the peak is refined (the results are exact in every mode; one product
 * instead of three, and a third of the fp32 accumulations keeps the screen's bound,
 * and so the refinement's candidates, smaller than a split mode's) and the split mode
 * when the screen values are the results: F16 / F16_SPLIT for U16 patterns, BF16 /
 * BF16_SPLIT for F32 and F64 patterns. */
typedef enum {
    EMIDA_MFMA_AUTO = -1,
    EMIDA_MFMA_BF16 = 0,
    EMIDA_MFMA_BF16_SPLIT = 1,
    EMIDA_MFMA_F16 = 2,
    EMIDA_MFMA_F16_SPLIT = 3
} emida_mfma_mode;

enum {
    EMIDA_OK = 0,
    EMIDA_ERR_ARG = 1,         /* invalid argument */
    EMIDA_ERR_HIP = 2,         /* HIP runtime error */
    EMIDA_ERR_STATE = 3,       /* call order (e.g. compute before allocate_memory) */
    EMIDA_ERR_UNSUPPORTED = 4, /* valid in the reference but not implemented here */
    EMIDA_ERR_NOMEM = 5        /* device or pinned allocation failed */
};

/* size2_t, src/common.hpp:132 */
typedef struct {
    uint32_t x, y;
} emida_size2;

typedef struct emida_offset_s emida_offset;

/* Derived sizes of an object (src/gpu_offset.hpp:100-125). */
typedef struct {
    uint32_t total_slices;     /* ROIs x batch */
    uint32_t neigh_size;       /* fitsize^2 x total_slices */
    uint32_t data_index_bytes; /* sizeof(data_index<T>): 8 (F32) or 16 (F64) */
    emida_size2 cross_size;    /* effective correlation size (FFT forces 2s-1) */
    emida_size2 corr_buffer;   /* arg-max domain: cross_size (brute) or 2s (FFT) */
    uint32_t fft_size;         /* padded FFT side M of the x axis: the power of two >= 2 s.x (0 for brute) */
    uint64_t device_bytes;     /* device memory held by the object */
} emida_offset_info;

/* Message of the last failing call on this thread ("" if none). */
const char* emida_last_error(void);

/* ------------------------------------------------------------------ drop-in object */

/* gpu_offset(src_size, begins, slice_size, cross_size, batch_size, policy, s,
 * in_buffers_size), src/gpu_offset.hpp:100-125.  `begins` is copied (the
 * reference borrows it).  fitsize in {3,5,7,9}; in_buffers >= 1. */
int emida_offset_create(emida_offset** out, emida_type T, emida_type IN, emida_size2 src_size,
                        const emida_size2* begins, uint32_t n_begins, emida_size2 slice_size,
                        emida_size2 cross_size, uint32_t batch_size, emida_cross_policy policy, int fitsize,
                        uint32_t in_buffers);

/* Frees every device and pinned buffer of the object (the reference never frees). */
void emida_offset_destroy(emida_offset* h);

int emida_offset_get_info(const emida_offset* h, emida_offset_info* info);

/* allocate_memory(IN* temp), src/gpu_offset.hpp:127-211: allocates the buffers and
 * prepares the reference pattern `temp` (host, src_size pixels) once: slices for
 * BRUTE, spectra for FFT.  temp may be NULL (buffers only). */
int emida_offset_allocate_memory(emida_offset* h, const void* temp);

/* transfer_pic_to_device_async(IN* pic), src/gpu_offset.hpp:213-218: starts the
 * upload of batch_size patterns (host) into the next input buffer (round robin,
 * thread safe) and returns its index.  The upload is ordered after the previous
 * computation that read the same buffer; the call blocks while that buffer still
 * holds an upload that no get_offset_core call has consumed (a single-threaded
 * caller must not have more than in_buffers uploads outstanding).  `pic` must
 * stay unchanged until emida_offset_wait_uploads (or the consuming computation)
 * returns. */
int emida_offset_transfer_pic_to_device_async(emida_offset* h, const void* pic, uint32_t* buffer_index);

/* cudaStreamSynchronize(in_stream) as called by the reference's pipeline
 * (src/process_files.hpp:196). */
int emida_offset_wait_uploads(emida_offset* h);

/* get_offset_core(maxes_i, neighbors, buffer), src/gpu_offset.hpp:234-278: runs the
 * pipeline on input buffer `buffer_index` and copies the arg-max records
 * (data_index<T>[total_slices]) and neighbourhoods (T[neigh_size]) to the host
 * pointers, asynchronously on the output stream.  The subpixel fit also runs on
 * the GPU; emida_offset_finalize returns its result. */
int emida_offset_get_offset_core(emida_offset* h, void* maxes_i, void* neighbors, uint32_t buffer_index);

/* finalize(maxes_i, neighbors), src/gpu_offset.hpp:280-301: waits for the output
 * stream and writes offsets (double[2 x total_slices], x then y) and fit
 * coefficients (double[6 x total_slices], a..f of a + bx + cy + dx^2 + exy + fy^2).
 * For buffers filled by get_offset_core it returns the GPU fit of that call; for
 * any other buffers it uploads them and fits them on the GPU. */
int emida_offset_finalize(emida_offset* h, const void* maxes_i, const void* neighbors, double* offsets,
                          double* coefs);

/* get_offset(IN* pic), src/gpu_offset.hpp:220-231: synchronous upload, compute and
 * finalize of one batch.  coefs may be NULL. */
int emida_offset_get_offset(emida_offset* h, const void* pic, double* offsets, double* coefs);

/* Device-resident form of get_offset_core + finalize (no reference counterpart;
 * used by pipelines that keep patterns in HBM): `n_patterns` (1..batch_size)
 * patterns at device pointer d_pics; per-ROI results to device pointers (any may
 * be NULL), stream-ordered on `stream`. */
int emida_offset_process_device(emida_offset* h, const void* d_pics, uint32_t n_patterns, void* d_maxes,
                                void* d_neighbors, double* d_offsets, double* d_coefs, void* stream);

/* get_cu_pic(), src/gpu_offset.hpp:318-321: device pointer to the prepared pattern
 * slices.  FFT policy: the input of emida_offset_cross_corr_fft, total_slices buffers
 * of (2 s.x) x (2 s.y) T (rows of 2 s.x), zeroed when first requested; the pipeline
 * itself keeps its prepared slices on chip and does not write them.  BRUTE (and the
 * refined BRUTE_MFMA) policy: the s.x x s.y T slices get_offset_core prepares.
 * The FFT policy's seam needs 2s to be a power of two on both axes
 * (EMIDA_ERR_UNSUPPORTED otherwise). */
int emida_offset_get_cu_pic(emida_offset* h, void** dptr);
/* get_cu_temp(), src/gpu_offset.hpp:322-325.  FFT policy: the ROIs' reference spectra
 * in cuFFT's R2C layout, 2 s.y rows of s.x + 1 complex T per ROI (the bytes of the
 * reference's (2 s.x + 2) x 2 s.y in-place buffer), the unnormalised R2C of the
 * zero-padded prepared reference slices when allocate_memory had a reference.  BRUTE
 * (and refined BRUTE_MFMA): the prepared reference slices, s.x x s.y T each. */
int emida_offset_get_cu_temp(emida_offset* h, void** dptr);
/* cross_corr_fft(), src/gpu_offset.hpp:303-316, FFT policy: R2C of every cu_pic buffer,
 * product with the conjugate of its ROI's cu_temp spectrum (src/hadamard.cu:39-40),
 * unnormalised C2R (the column transform first, then each row's C2R, which ignores
 * the imaginary parts of its DC and Nyquist terms) into the buffers of
 * emida_offset_get_cu_cross_res; synchronous. */
int emida_offset_cross_corr_fft(emida_offset* h);
/* The correlation buffers cross_corr_fft writes (the reference's cu_cross_res_):
 * total_slices x (2 s.x) x (2 s.y) T.  No public accessor in the reference. */
int emida_offset_get_cu_cross_res(emida_offset* h, void** dptr);
/* The object's upload and download streams as hipStream_t values: the reference's
 * public in_stream / out_stream members (src/gpu_offset.hpp:96-97), made by
 * allocate_memory. */
int emida_offset_get_streams(emida_offset* h, void** in_stream, void** out_stream);

/* Operand precision of the BRUTE_MFMA policy (default EMIDA_MFMA_AUTO, see
 * emida_mfma_mode; the fp16 modes need IN = U16); call before
 * emida_offset_allocate_memory.  No reference counterpart. */
int emida_offset_set_mfma_mode(emida_offset* h, emida_mfma_mode mode);

/* The operand precision the BRUTE_MFMA policy runs with: the mode set, or what
 * EMIDA_MFMA_AUTO resolves to for the object's pixel type and refinement setting.
 * No reference counterpart. */
int emida_offset_get_mfma_mode(emida_offset* h, emida_mfma_mode* mode);

/* Exact peak for the BRUTE_MFMA policy (default on): the matrix-core correlation only
 * screens; every shift within its rounding bound of the screen maximum and the F x F
 * neighbourhood are evaluated the way the BRUTE policy evaluates them, in T, so the
 * results equal the BRUTE policy's (T = F32 or F64).  Off (T = F32 only): the float
 * screen values are used directly.  Call before emida_offset_allocate_memory.  No
 * reference counterpart. */
int emida_offset_set_mfma_refine(emida_offset* h, int enable);

/* Kernel timing with HIP events on the stream the kernels run on.  Returns the
 * summed time and launch count of the dominant kernel (the FFT pipeline kernel
 * for FFT, the correlation kernel for brute) since timing was enabled. */
int emida_offset_set_timing(emida_offset* h, int enable);
int emida_offset_get_timing(emida_offset* h, const char** kernel_name, double* total_ms, uint64_t* launches);

/* ------------------------------------------------------------------ memory helpers */

/* cuda_malloc_host (src/host_helpers.hpp:74-80): pinned host memory. */
int emida_host_alloc(void** ptr, size_t bytes);
int emida_host_free(void* ptr);
int emida_device_alloc(void** ptr, size_t bytes);
int emida_device_free(void* ptr);
int emida_memcpy_h2d(void* dst, const void* src, size_t bytes);
int emida_memcpy_d2h(void* dst, const void* src, size_t bytes);
int emida_memcpy_d2d(void* dst, const void* src, size_t bytes);
int emida_memset_device(void* dst, int value, size_t bytes);
int emida_device_synchronize(void);
/* Device name and compute-unit count of the current device. */
int emida_device_query(char* name, size_t name_len, int* cu_count);
/* Free and total device memory of the current device (hipMemGetInfo). */
int emida_device_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ------------------------------------------------------------------ kernel-level seam
 * Counterparts of the launchers in src/kernels.cuh:56-125 (the reference's tests call
 * them directly).  All pointers are device pointers. */

/* run_sum(data, sums, begins, src_size, slice_size, begins_size, batch_size),
 * src/kernels.cuh:117-119, src/sum.cu:160-167.  Sum type: uint32 for U16 input,
 * else IN. */
int emida_run_sum(emida_type IN, const void* data, void* sums, const emida_size2* begins, emida_size2 src_size,
                  emida_size2 slice_size, uint32_t begins_size, uint32_t batch_size, void* stream);

/* run_prepare_pics, src/kernels.cuh:93-104, src/prepare_pics.cu:46-63.  S is the sum
 * element type (U16 means uint32 sums). */
int emida_run_prepare_pics(emida_type IN, emida_type T, emida_type S, const void* pic, void* slices,
                           const void* hanning_x, const void* hanning_y, const void* sums,
                           const emida_size2* begins, emida_size2 src_size, emida_size2 slice_size,
                           emida_size2 out_size, uint32_t begins_size, uint32_t batch_size, void* stream);

/* run_cross_corr(pics, ref, res, size, res_size, ref_slices, batch_size),
 * src/kernels.cuh:55-62, src/cross_corr.cu:68-74.  T in {F32, F64, I32};
 * res_size <= 2*size+1 per axis. */
int emida_run_cross_corr(emida_type T, const void* pics, const void* ref, void* res, emida_size2 size,
                         emida_size2 res_size, uint32_t ref_slices, uint32_t batch_size, void* stream);

/* cross_corr_fft(), src/gpu_offset.hpp:303-316 (R2C, hadamard, C2R; unnormalised):
 * prepared size.x x size.y slices -> (2 size.x) x (2 size.y) circular correlation
 * buffers (rows of 2 size.x values), batch x ref_slices of them.  Sides 2..256 per
 * axis, square or not. */
int emida_run_cross_corr_fft(emida_type T, const void* pics, const void* ref, void* res, emida_size2 size,
                             uint32_t ref_slices, uint32_t batch_size, void* stream);

/* emida_run_cross_corr_fft without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as emida_run_cross_corr_fft (bit for bit); the reference spectra, and for
 * slice sides above 64 (float) the column-stage work buffer, live in a static device
 * array of the library, emida_fft_noalloc_scratch_bytes() = 32 MiB, one per GPU, held
 * until the process ends.  EMIDA_ERR_ARG when ref_slices spectra (66.5 KiB per slice
 * of side 64 in float, twice that in double; plus 264 KiB of work per slice for sides
 * 65..128) do not fit.  Every call on a GPU shares that GPU's array, and each call
 * leaves its contents there: calls that overlap on one GPU (other streams, other host
 * threads) race and must be serialised by the caller. */
int emida_run_cross_corr_fft_noalloc(emida_type T, const void* pics, const void* ref, void* res, emida_size2 size,
                                     uint32_t ref_slices, uint32_t batch_size, void* stream);
/* Size of the static scratch array emida_run_cross_corr_fft_noalloc works in. */
int emida_fft_noalloc_scratch_bytes(size_t* bytes);

/* Brute-force correlation of prepared float slices on matrix cores (the computation
 * of run_cross_corr, src/cross_corr.cu:68-74, in the given operand precision, FP32
 * accumulation, float results).  No reference counterpart.  size.x a multiple of 8
 * up to 128; res_size <= 2*size-1. */
int emida_run_cross_corr_mfma(emida_mfma_mode mode, const float* pics, const float* ref, float* res,
                              emida_size2 size, emida_size2 res_size, uint32_t ref_slices, uint32_t batch_size,
                              void* stream);

/* run_hadamard(pics, ref, one_size, ref_slices, batch_size), src/kernels.cuh:122-123,
 * src/hadamard.cu:43-49: pics[i] = pics[i] * conj(ref[i mod (one_size.area * ref_slices)])
 * on interleaved complex T. */
int emida_run_hadamard(emida_type T, void* pics, const void* ref, emida_size2 one_size, uint32_t ref_slices,
                       uint32_t batch_size, void* stream);

/* run_maxarg_reduce(data, maxes, maxarg, size, block_size, batch_size),
 * src/kernels.cuh:106-107, src/maxarg.cu:180-187: data_index<T> of the largest
 * value per slice of size.x*size.y values (lowest index on ties). */
int emida_run_maxarg_reduce(emida_type T, const void* data, void* maxarg, emida_size2 size, uint32_t batch_size,
                            void* stream);

/* run_extract_neighbors<T, pos_policy>(data, max_i, neighbors, s, src_size,
 * batch_size), src/kernels.cuh:109-110, src/extract_neighbors.cu:43-49.
 * pos_policy BRUTE = cross_res_pos_policy_id, FFT = cross_res_pos_policy_fft.
 * Reads at or past data_len elements (possible through the one-past-the-edge
 * clamp of the last slice) return 0. */
int emida_run_extract_neighbors(emida_type T, emida_cross_policy pos_policy, const void* data, size_t data_len,
                                const void* max_i, void* neighbors, int s, emida_size2 src_size,
                                uint32_t batch_size, void* stream);

/* subpixel_max_serial(pic, s, batch_size), src/subpixel_max.hpp:76-97, on the GPU:
 * offsets double[2n] (x, y within the s x s neighbourhood), coefs double[6n]. */
int emida_run_subpixel_max(emida_type T, const void* neighbors, int s, uint32_t n, double* offsets,
                           double* coefs, void* stream);

/* Largest slice side (per axis) the FFT policy accepts for element type T: 256 for
 * both (any side from 2 up to this, square or not; each axis' transform is padded to
 * the next power of two >= 2s, with the same correlation values).  No counterpart in
 * the reference, where cuFFT plans any size. */
int emida_fft_max_slice(emida_type T, uint32_t* side);

/* get_lstsq_matrix(s), src/subpixel_max.cpp:5-20: (A^T A)^-1 A^T as 6 x s^2 doubles. */
int emida_lstsq_matrix(int s, double* out);

/* generate_hanning<T>(size), src/host_helpers.hpp:29-37. */
int emida_generate_hanning(emida_type T, size_t size, void* out);

/* ======================================================================
 * Request server (serving form).  No counterpart in the reference: its only driver,
 * src/process_files.hpp:135-241, runs one fixed gpu_offset configuration over a file
 * list.  Here single-pattern requests for several resident configurations ("plans")
 * are batched continuously: a plan's queued requests launch together as soon as the
 * plan has a free slot (a device buffer set for one batch); slots are created while
 * their total size stays within hbm_budget; batches overlap on `streams` streams; a
 * (slot, batch size) pipeline is captured in a hipGraph once and replayed.  Each
 * request's results are the same as those of a lone request (the per-slice work does
 * not depend on the other slices of a batch).  Graphs exist per batch size on the grid
 * 1, 2, 3, 4, 6, 8, 12, 16, 24, ... (powers of two and 3/2 of them); a batch of n
 * requests runs the graph of the next size, the extra slices computing on stale slot
 * contents whose results are discarded.  Patterns are uint16.
 * ====================================================================== */
typedef struct emida_server_s emida_server;

typedef struct {
    uint64_t hbm_budget; /* bytes of device memory for batch slots (inputs, working memory, results) */
    uint32_t max_batch;  /* patterns per launch; lowered per plan so one slot fits in hbm_budget / 4 */
    uint32_t streams;    /* device streams batches are spread over (1..16) */
    uint32_t max_queued; /* pinned staging buffers: submitted requests not yet completed; submit blocks beyond */
    int use_graphs;      /* replay hipGraph-captured pipelines (1) or launch kernels directly (0) */
    uint32_t slots_per_plan; /* batch slots one plan may hold at once (0: as many as the budget allows) */
} emida_server_config;

typedef struct {
    uint64_t submit_ns, launch_ns, done_ns; /* steady clock of the library */
    uint32_t batch;                         /* requests in the launch that carried this one */
    uint32_t plan;
} emida_request_info;

typedef struct {
    uint64_t submitted, completed, launches, graph_launches, graphs_captured;
    uint64_t device_bytes_resident;       /* reference slices / spectra of all plans */
    uint64_t device_bytes_batches;        /* sum of created slots */
    uint64_t device_bytes_peak_inflight;  /* largest sum of slots in flight at once */
    uint32_t slots;
} emida_server_stats;

int emida_server_create(emida_server** out, const emida_server_config* cfg);
/* Completes every submitted request, then releases everything. */
void emida_server_destroy(emida_server* s);
/* A plan: the arguments of emida_offset_create (IN = U16) plus the reference pattern;
 * its reference slices or spectra are prepared now and stay resident.  mfma_mode is
 * used by EMIDA_CROSS_POLICY_BRUTE_MFMA only (EMIDA_MFMA_AUTO: fp16, refined). */
int emida_server_add_plan(emida_server* s, emida_type T, emida_size2 src_size, const uint16_t* ref,
                          const emida_size2* begins, uint32_t n_begins, emida_size2 slice_size,
                          emida_size2 cross_size, emida_cross_policy policy, emida_mfma_mode mfma_mode, int fitsize,
                          uint32_t* plan_id);
int emida_server_plan_info(emida_server* s, uint32_t plan_id, emida_offset_info* info, uint32_t* max_batch);
/* Copies the pattern (src_size.x * src_size.y uint16) and queues it; offsets
 * (double[2 * n_begins]) and coefs (double[6 * n_begins], may be null) are written
 * before the request completes.  Blocks while max_queued requests are pending. */
int emida_server_submit(emida_server* s, uint32_t plan_id, const uint16_t* pattern, double* offsets, double* coefs,
                        uint64_t* request_id);
/* Blocks until the request completed; info may be null.  A request is waited for once. */
int emida_server_wait(emida_server* s, uint64_t request_id, emida_request_info* info);
int emida_server_get_stats(emida_server* s, emida_server_stats* stats);

/* ======================================================================
 * Input stage: raw-strip TIFF patterns and the multi-threaded file pipeline.
 * ====================================================================== */

/* get_size(fname), src/load_tiff.hpp:57-69: width (x) and height (y) of the first
 * image.  EMIDA_ERR_ARG when the file cannot be opened or is not a classic TIFF. */
int emida_tiff_size(const char* path, emida_size2* size);

/* load_tiff(file_name, raster, size), src/load_tiff.hpp:14-53: the strips of an
 * uncompressed 16-bit single-sample TIFF copied in file order into raster
 * (size.x * size.y uint16, row-major, no byte swapping).  EMIDA_ERR_ARG for a size
 * other than `size`, an orientation other than top-left, or strips that are missing,
 * truncated, larger than the raster or do not fill it; EMIDA_ERR_UNSUPPORTED for compressed images
 * and other pixel formats (the reference copies their raw bytes regardless). */
int emida_tiff_load(const char* path, uint16_t* raster, emida_size2 size);

/* Results of one batch of files: files first .. first + count - 1 of the list;
 * offsets (double[2 * count * n_begins]) and coefs (double[6 * count * n_begins])
 * in file-major order, valid during the call.  status != EMIDA_OK (offsets and
 * coefs null, error names the file) when a file of the batch could not be read or
 * the pipeline stopped. */
typedef void (*emida_files_callback)(void* user, uint32_t first, uint32_t count, int status, const char* error,
                                     const double* offsets, const double* coefs);

typedef struct {
    uint32_t batches;        /* batches of batch_size files */
    uint32_t failed_batches; /* batches reported with status != EMIDA_OK */
    double wall_ms;          /* the whole call */
    double load_ms;          /* summed over loader threads: waiting for a free host buffer and reading files */
    double upload_wait_ms;   /* summed over loader threads: waiting for a free device input buffer */
} emida_files_stats;

/* file_processor::process_files, src/process_files.hpp:135-241, for an allocated
 * emida_offset with IN = U16: load_workers threads (the caller's included) read
 * batch_size files each into pinned buffers and upload them; one thread runs
 * get_offset_core per batch; one thread finalizes and calls cb once per batch (in
 * file order when `ordered`, else in completion order).  A file that cannot be read
 * fails its batch only.  Returns EMIDA_OK unless the pipeline itself failed (a HIP
 * error); EMIDA_ERR_ARG for an empty list.  The handle must not be used by other
 * threads during the call; the emida_offset's in_buffers_size sets how many uploads
 * may be in flight (the reference uses 2 * load_workers). */
int emida_process_files(emida_offset* h, const char* const* paths, uint32_t n_paths, uint32_t load_workers,
                        int ordered, emida_files_callback cb, void* user, emida_files_stats* stats);

/* load_slice_begins(file_name), src/args_parser.cpp:62-108: the slicepos file.  Its
 * first line holds the ROI "radius" (half the slice size; *radius_ok = 0 when it does
 * not parse), each further line "<mid_x> <mid_y>" (std::stoul rules: leading digits,
 * trailing characters ignored).  Lines without a space or with an unparsable number
 * are skipped with a warning on stderr (printed by calls that pass `mids`, so a
 * counting call with mids = NULL stays quiet).  Up to `capacity` middles are written;
 * *count receives the number of middles in the file. */
int emida_read_slicepos(const char* path, uint64_t* radius, int* radius_ok, emida_size2* mids, uint32_t capacity,
                        uint32_t* count);

/* file_processor::load_work, src/process_files.hpp:54-73: the deformed list, one
 * pattern per line, "<x> <y> <file name>" (the name is the rest of the line after the
 * character that follows y).  Up to `capacity` (x, y) pairs go to xy; the names, each
 * NUL-terminated, are packed into `names` while they fit in names_capacity bytes;
 * *count and *names_bytes receive the totals.  A line whose x or y does not parse
 * gives x = y = 0 and an empty name. */
int emida_read_deformed_list(const char* path, double* xy, uint32_t capacity, uint32_t* count, char* names,
                             size_t names_capacity, size_t* names_bytes);

/* draw_tiff(pic, size, file_name, vectors, vectors_pos), src/draw_tiff.hpp:26-59: writes
 * the pattern (size.x * size.y uint16) as an 8-bit RGBA TIFF, grey = pixel >> 8, with
 * each vector (double[2 * n_vectors], x then y) drawn in red as 20 points from its
 * position.  Vectors with NaN components are skipped (the reference would convert NaN
 * to int). */
int emida_draw_tiff(const uint16_t* pic, emida_size2 size, const char* path, const double* vectors,
                    const emida_size2* positions, uint32_t n_vectors);

/* ======================================================================
 * Deformation post-step: per pattern, the projective map between the ROI middles and
 * their displaced positions (an algebraic homography fit) and its strain and rotation.
 * ====================================================================== */

/* Pattern centre (x, y) and detector distance z, in pixels. */
typedef struct {
    double x, y, z;
} emida_calib;

/* The reference's post-processing of the offsets (python/process_offsets.py:31-97,
 * python/test.py:137-198, jove/analyze.py:6-95), one workgroup per pattern.  All
 * pointers are device pointers (any output may be NULL); stream-ordered on `stream`.
 *   mids     double[n_rois][2] ROI middles (x, y), shared by all patterns, or
 *            double[n_patterns][n_rois][2] when mids_per_pattern;
 *   offsets  double[n_patterns][n_rois][2], as the pipeline returns them;
 *   coefs    double[n_patterns][n_rois][6] fit coefficients, or NULL to use every ROI.
 * With coefs, a ROI is used when d = a c - b^2 > det_min (a = -coef[3], b = -coef[4]/2,
 * c = -coef[5]; the reference's threshold is 1e-6), quality = median of d over all ROIs,
 * and a pattern is fitted when more than min_rois ROIs are used (the reference: 10).
 * Without coefs every ROI is used (n_rois > 4) and quality is NaN.  Outputs per pattern:
 *   mask     int32: 1 fitted, 0 not (its matrices are NaN);
 *   H        double[9], the map divided by its [2][2] element (process_offsets.py);
 *   F        double[9], the map divided by the cube root of its determinant;
 *   epsilon  double[9], log of the right stretch U of F = R U (strain);
 *   omega    double[9], log of R by Rodrigues' formula (rotation; NaN for exactly no
 *            rotation, as the reference's 0/0).
 * Matrices are row-major.  The second row of each ROI pair is [0, 0, 0, -ox z, -oy z,
 * -z^2, ox ny, oy ny, z ny] as in process_offsets.py (jove/analyze.py:28 puts ox ny in
 * the oy column).  n_rois <= 8192. */
int emida_fit_deformation(const double* mids, int mids_per_pattern, uint32_t n_rois, const double* offsets,
                          const double* coefs, uint32_t n_patterns, emida_calib calib, double det_min,
                          uint32_t min_rois, int32_t* mask, double* quality, double* H, double* F,
                          double* epsilon, double* omega, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* EMIDA_AMD_H */
