// Host-side launchers of the gfx950 kernels (internal to the library).
#pragma once

#include <cmath>

#include "common.hpp"

namespace emida_amd {

// ---- prep.hip
template <class IN, class S>
void launch_sum(const IN* pics, S* sums, const size2* begins, size2 src, size2 slice, uint32_t nb, uint32_t total,
                hipStream_t st);
template <class IN, class T, class S>
void launch_prepare(const IN* pics, T* slices, const T* hx, const T* hy, const S* sums, const size2* begins,
                    size2 src, size2 slice, size2 out, uint32_t nb, uint32_t total, hipStream_t st);
// Sums and preparation of uint16 slices in one kernel (sum_kernel + prepare_kernel +
// prepare_split_kernel with identical results); slices (T, slice layout) and hi/lo
// (half planes scaled by in_scale) are optional.  stats (optional, with slices and the
// prepared T reference slices ref [roi][slice]): per slice, in double, the sum of
// squares of the prepared slice (stats[slice]) and its sum of products with the
// reference slice (stats[total + slice]).  Returns false, launching nothing, when a
// slice does not fit the LDS staging (kSumPrepareMaxLds bytes of pixels).
constexpr size_t kSumPrepareMaxLds = 32768;
template <class T, class HT>
bool launch_sum_prepare(const uint16_t* pics, uint32_t* sums, T* slices, const T* hx, const T* hy, HT* hi, HT* lo,
                        const float* hxf, const float* hyf, float in_scale, double* stats, const T* ref,
                        const size2* begins, size2 src, size2 slice, uint32_t nb, uint32_t total, hipStream_t st);
// The same statistics of n prepared slices of `area` T values: sums of squares into
// out[0..n) and, with ref (slice i against ref slice i % nb), sums of products into
// out[n..2n).
template <class T>
void launch_slice_stats(const T* slices, const T* ref, uint32_t area, uint32_t nb, uint32_t n, double* out,
                        hipStream_t st);

// ---- brute.hip
// res[(i*ref_slices + j)][...] = correlation of pics slice (i, j) with ref slice j.
template <class T>
void launch_cross_corr(const T* pics, const T* ref, T* res, size2 size, size2 res_size, uint32_t ref_slices,
                       uint32_t batch, hipStream_t st);
// Name of the kernel launch_cross_corr picks for this shape (float or double).
const char* cross_corr_kernel_name(bool is_double, size2 size, size2 res_size, uint32_t ref_slices, uint32_t batch);

// ---- peak.hip
// Offsets of the least-squares tables for fitsize 3, 5, 7, 9 in the device table.
int lstsq_offset(int fit);
const double* lstsq_device_table();  // uploads once per device
void lstsq_host(int s, double* out);  // exact (A^T A)^-1 A^T, 6 x s^2

// Per slice: arg-max over `area` values (lowest index on ties), F x F gather
// with the reference's clamp and position policy, GPU least-squares fit and the
// finalize formula.  Any output pointer may be null.
template <class T>
// row_max (optional, float screens): per-row maximum codes of corr ([slice][cross.y],
// screen_max_code); the arg-max then reads only the first row holding the maximum.
void launch_peak(const T* corr, size_t corr_len, uint32_t area, size2 cross, int policy, int fit,
                 data_index<T>* maxes, T* neigh, double* offsets, double* coefs, uint32_t total, hipStream_t st,
                 const uint32_t* row_max = nullptr);
template <class T>
void launch_maxarg(const T* data, uint32_t area, data_index<T>* out, uint32_t total, hipStream_t st);
template <class T>
void launch_extract(const T* data, size_t data_len, const data_index<T>* max_i, T* neigh, int s, size2 size,
                    int policy, uint32_t total, hipStream_t st);
// Fit only (subpixel_max_serial); with `maxes` non-null also the finalize formula
// (offsets become pattern offsets instead of neighbourhood positions).
template <class T>
void launch_subpixel(const T* neigh, int s, uint32_t n, const data_index<T>* maxes, size2 cross, int policy,
                     double* offsets, double* coefs, hipStream_t st);

// Candidates of the exact refinement as the matrix-core kernel publishes them in
// place of the screen: per (slice, row), every position whose screen value is within
// tau of a lower bound of the slice maximum (the largest row maximum the publishing
// workgroup holds, or the zero shift's dot product less tau), so each candidate of the
// refinement is among them.
constexpr int kListCap = 32;  // positions kept per row; a row with more makes its slice sweep
struct cand_entry {
    float v;     // screen value
    uint32_t x;  // column of the screen row
};
struct screen_lists {
    uint32_t* cnt;      // [slice][row]: number of such positions (may exceed kListCap)
    cand_entry* ent;    // [slice][row][kListCap]: the first kListCap of them, in no order
    const double* pn2;  // [slice]: sum of squares of the T-precision prepared pattern slice
    const double* rn2;  // [roi]: the same for the reference slice
    const double* z0;   // [slice]: sum of p * r over the slice (the zero shift), in double
    double tol_rel, tol_abs;
};
// tau = tol_rel ||p|| ||r|| + tol_abs (||p|| + ||r||) of one slice, evaluated the same
// way wherever it is used.
__host__ __device__ inline double refine_tau(const screen_lists& L, size_t slice, uint32_t roi) {
#pragma clang fp contract(off)
    const double np = sqrt(L.pn2[slice]), nr = sqrt(L.rn2[roi]);
    return L.tol_rel * np * nr + L.tol_abs * (np + nr);
}

// Precision form: exact peak from a float matrix-core screen of `total` slices and the
// T-precision prepared pattern (p, [slice][s]) and reference (r, [roi][s]) slices;
// outputs as launch_peak with the brute position policy.  The candidate bound is
// tau = tol_rel ||p|| ||r|| + tol_abs (||p|| + ||r||) per slice.  screen_max (optional):
// the screen maximum of each slice row as an ordered code (screen_max_code), from
// launch_mfma_corr; without it the kernel scans the screen for its maximum.  lists
// (optional, with screen_max): the published candidates; the screen is then not read
// (it may be null) and tau comes from the lists' norms.
template <class T>
void launch_refine(const float* screen, size_t corr_len, const T* p, const T* r, size2 s, size2 cross, uint32_t nb,
                   double tol_rel, double tol_abs, int fit, data_index<T>* maxes, T* neigh, double* offsets,
                   double* coefs, uint32_t total, hipStream_t st, const uint32_t* screen_max = nullptr,
                   const screen_lists* lists = nullptr);
// Screen error bound of launch_mfma_corr against the T-precision brute values, for
// launch_refine (both values' rounding, counted twice): mode as emida_mfma_mode.
void mfma_refine_bounds(int mode, size2 s, bool t_double, double& tol_rel, double& tol_abs);

// ---- fft.hip, fft256.hip
// W_M^t = exp(-2 pi i t / M) for every power of two 4 <= M <= kTwMaxM, the table of M at
// offset M - 4 (M entries), rounded from extended precision; a static device array per
// GPU, filled on first use (no allocation).
constexpr int kTwMaxM = 512;
template <class T>
const cplx<T>* fft_twiddles();
// Largest padded FFT side of the single-kernel pipelines: 128 for double, 256 for
// float (M = 256 in fft256.hip); fftg.hip takes every other shape up to kFftMaxSlice
// per axis.
int fft_max_side(bool is_double);
constexpr uint32_t kFftMaxSlice = 256;
// Transform side M of a slice side s: the power of two >= 2s (at least 4).
uint32_t fft_padded(uint32_t side);
// Reference spectra: prepared reference slices -> half spectra in the layout the
// pipeline kernel reads ((M/2+1) columns x M bit-reversed rows per ROI).
template <class T, class IN, class S>
void launch_fft_ref(const IN* pics, const S* sums, const T* hx, const T* hy, const size2* begins, size2 src,
                    uint32_t side, uint32_t nb, cplx<T>* spectra, hipStream_t st);
// Whole FFT pipeline of `total` slices (patterns x ROIs): prepare, R2C, multiply by
// conj(reference spectrum), C2R, arg-max, gather, fit.  corr_out (optional)
// receives the (2s)^2 circular correlation buffers.
template <class T, class IN, class S>
void launch_fft_pipe(const IN* pics, const S* sums, const T* hx, const T* hy, const size2* begins, size2 src,
                     uint32_t side, uint32_t nb, uint32_t total, const cplx<T>* spectra, int fit,
                     data_index<T>* maxes, T* neigh, double* offsets, double* coefs, T* corr_out, hipStream_t st);
// Prepared slices in, circular correlation buffers out (cross_corr_fft seam).
template <class T>
void launch_fft_corr_prepared(const T* pics, const T* ref, T* res, size2 side, uint32_t ref_slices, uint32_t batch,
                              hipStream_t st);
// The same without any allocation: the reference spectra (and the work buffer of the
// three-kernel path) live in the library's static device array of kFftStaticScratchBytes
// per GPU (shared by every call on that GPU: overlapping calls race).  Throws
// arg_error when the spectra (plus one slice's work buffer) exceed it.
constexpr size_t kFftStaticScratchBytes = (size_t)32 << 20;
template <class T>
void launch_fft_corr_prepared_noalloc(const T* pics, const T* ref, T* res, size2 side, uint32_t ref_slices,
                                      uint32_t batch, hipStream_t st);
template <class T>
void launch_hadamard(cplx<T>* pics, const cplx<T>* ref, size2 one, uint32_t ref_slices, uint32_t batch,
                     hipStream_t st);

// ---- fftg.hip: three-kernel FFT path (non-square slices, double sides above 64, sides
// above 128; transforms up to 512 per axis), work buffers of half spectra in HBM
bool fftg_needed(bool is_double, size2 side);               // the single-kernel paths do not take this shape
size_t fftg_spectrum_bytes(bool is_double, size2 side);      // reference spectrum of one ROI
size_t fftg_slice_work_bytes(bool is_double, size2 side);    // half spectra of one slice
size_t fftg_chunk(bool is_double, size2 side, uint32_t total, size_t extra_per_slice);  // slices per launch chunk
size_t fftg_ref_work_bytes(bool is_double, size2 side, uint32_t nb);  // work of launch_fftg_ref
// work of launch_fftg_pipe (with_corr: it also holds the chunk's correlation buffers)
size_t fftg_work_bytes(bool is_double, size2 side, uint32_t total, bool with_corr);
template <class T, class IN, class S>
void launch_fftg_ref(const IN* pics, const S* sums, const T* hx, const T* hy, const size2* begins, size2 src,
                     size2 side, uint32_t nb, cplx<T>* spectra, void* work, hipStream_t st);
// as launch_fft_pipe; work: fftg_work_bytes(.., total, true) bytes
template <class T, class IN, class S>
void launch_fftg_pipe(const IN* pics, const S* sums, const T* hx, const T* hy, const size2* begins, size2 src,
                      size2 side, uint32_t nb, uint32_t total, const cplx<T>* spectra, int fit, data_index<T>* maxes,
                      T* neigh, double* offsets, double* coefs, T* corr_out, void* work, hipStream_t st);
// prepared slices -> (2s.x x 2s.y) circular correlation buffers; spectra: nb ROIs'
// spectra (fftg_spectrum_bytes each), work: `chunk` slices of half spectra
template <class T>
void launch_fftg_corr_prepared(const T* pics, const T* ref, T* res, size2 side, uint32_t nb, uint32_t batch,
                               cplx<T>* spectra, void* work, size_t chunk, hipStream_t st);
// cuFFT-layout R2C (spectra [n][2s.y][s.x + 1]) and R2C -> product -> C2R of
// 2s.x x 2s.y buffers (the reference's cross_corr_fft); 2s a power of two per axis
template <class T>
void launch_fftg_std_r2c(const T* pic, uint32_t pitch, cplx<T>* spectra, size2 side, uint32_t n, void* work,
                         size_t chunk, hipStream_t st);
template <class T>
void launch_fftg_std_corr(const T* pic, const cplx<T>* spectra, T* res, size2 side, uint32_t nb, uint32_t total,
                          void* work, size_t chunk, hipStream_t st);

// ---- mfma.hip (matrix-core brute correlation, float results)
// Prepared slices as half-type hi (+ lo when lo != nullptr) planes [slice][y][x];
// corr (optional): the screen, [pattern * nb + roi][cross.y][cross.x].  screen_max
// (optional, [pattern * nb + roi][cross.y]): the largest value of each row of each
// slice (NaN ignored) as an ordered code, 0 for a row without one.  lists (optional,
// with screen_max): the refinement's candidates per row (screen_lists).
// src (optional): uint16 patterns prepared while they are staged, in place of phi/plo
// (which are then unused; the split mode follows rlo), with prepare_split_kernel's
// arithmetic, so the screen is the same.
struct mfma_src {
    const uint16_t* pics;   // [pattern][src.y][src.x]
    const uint32_t* sums;   // [pattern * nb + roi]: slice sums
    const float* hx;        // float Hann windows
    const float* hy;
    const size2* begins;    // ROI corners
    size2 src;              // pattern size
};
template <class HT>
void launch_mfma_corr(const HT* phi, const HT* plo, const HT* rhi, const HT* rlo, float* corr, size2 s, size2 cross,
                      uint32_t nb, uint32_t npat, hipStream_t st, uint32_t* screen_max = nullptr,
                      const screen_lists* lists = nullptr, const mfma_src* src = nullptr);
void check_mfma_shape(size2 s, size2 cross);
template <class IN, class S, class HT>
void launch_prepare_split(const IN* pics, HT* hi, HT* lo, const float* hx, const float* hy, const S* sums,
                          const size2* begins, size2 src, size2 slice, uint32_t nb, uint32_t total, hipStream_t st);

// Order-preserving uint32 code of a float (a < b  <=>  code(a) < code(b), -0 < +0);
// 0 is below every code and stands for "no value".
__host__ __device__ inline uint32_t screen_max_code(float v) {
    uint32_t u;
    __builtin_memcpy(&u, &v, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float screen_max_value(uint32_t c) {
    const uint32_t u = (c & 0x80000000u) ? (c & 0x7fffffffu) : ~c;
    float v;
    __builtin_memcpy(&v, &u, 4);
    return v;
}
template <class HT>
void launch_split(const float* v, HT* hi, HT* lo, size_t n, hipStream_t st);
// The factor prepare_split_kernel applies before the split (1 for bf16, 1/2 for fp16).
template <class HT>
float mfma_in_scale();

}  // namespace emida_amd
