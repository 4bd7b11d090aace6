// Building blocks of the LDS-resident FFT kernels (fft.hip, fft256.hip): complex
// helpers, the in-register radix-2 DFT, the two-level transform over a lane group
// and the reference's prepare expression.
#pragma once

#include <type_traits>

#include "kernels.hpp"

namespace emida_amd {

// ------------------------------------------------------------------ complex helpers
template <class T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
    return {a.re + b.re, a.im + b.im};
}
template <class T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
    return {a.re - b.re, a.im - b.im};
}
template <class T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> w) {
    return {a.re * w.re - a.im * w.im, a.re * w.im + a.im * w.re};
}
template <class T>
__device__ __forceinline__ cplx<T> cmul_conj(cplx<T> a, cplx<T> w) {  // a * conj(w)
    return {a.re * w.re + a.im * w.im, a.im * w.re - a.re * w.im};
}
// The reference's product p * conj(r) (src/hadamard.cu:39-40), same expression.
template <class T>
__device__ __forceinline__ cplx<T> hadamard(cplx<T> p, cplx<T> r) {
    return {p.re * r.re + p.im * r.im, -p.re * r.im + p.im * r.re};
}

// exp(-2 pi i j / 16), j < 8: correctly rounded cos/sin of multiples of pi/8.
constexpr double kC1 = 0.9238795325112867, kS1 = 0.3826834323650898, kR2 = 0.7071067811865476;

// v * W_16^j (conjugate twiddle when INV); j is a compile-time constant after unrolling.
template <class T, bool INV>
__device__ __forceinline__ cplx<T> twiddle16(cplx<T> v, int j) {
    const T sg = INV ? T(-1) : T(1);
    switch (j) {
        case 0: return v;
        case 4: return INV ? cplx<T>{-v.im, v.re} : cplx<T>{v.im, -v.re};
        case 2: return cmul(v, cplx<T>{T(kR2), sg * T(-kR2)});
        case 6: return cmul(v, cplx<T>{T(-kR2), sg * T(-kR2)});
        case 1: return cmul(v, cplx<T>{T(kC1), sg * T(-kS1)});
        case 3: return cmul(v, cplx<T>{T(kS1), sg * T(-kC1)});
        case 5: return cmul(v, cplx<T>{T(-kS1), sg * T(-kC1)});
        default: return cmul(v, cplx<T>{T(-kC1), sg * T(-kS1)});  // 7
    }
}

__host__ __device__ constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }
__host__ __device__ constexpr int brev_c(int v, int bits) {
    int r = 0;
    for (int b = 0; b < bits; ++b) r |= ((v >> b) & 1) << (bits - 1 - b);
    return r;
}

// v * W_16^j for any j (W_16^(j + 8) = -W_16^j).
template <class T, bool INV>
__device__ __forceinline__ cplx<T> twiddle16_any(cplx<T> v, int j) {
    j &= 15;
    const cplx<T> w = twiddle16<T, INV>(v, j & 7);
    return j >= 8 ? cplx<T>{-w.re, -w.im} : w;
}

// 4-point DFT of (a0, a1, a2, a3) in place (forward: W_4 = -i; INV: +i); HALF_ZERO:
// a2 = a3 = 0 (not read).
template <class T, bool INV, bool HALF_ZERO>
__device__ __forceinline__ void dft4(cplx<T>& a0, cplx<T>& a1, cplx<T>& a2, cplx<T>& a3) {
    if constexpr (HALF_ZERO) {
        const cplx<T> t0 = a0, t1 = a1;
        a0 = cadd(t0, t1);
        a2 = csub(t0, t1);
        // -i t1 forward, +i t1 inverse
        a1 = INV ? cplx<T>{t0.re - t1.im, t0.im + t1.re} : cplx<T>{t0.re + t1.im, t0.im - t1.re};
        a3 = INV ? cplx<T>{t0.re + t1.im, t0.im - t1.re} : cplx<T>{t0.re - t1.im, t0.im + t1.re};
    } else {
        const cplx<T> t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = csub(a1, a3);
        a0 = cadd(t0, t2);
        a2 = csub(t0, t2);
        a1 = INV ? cplx<T>{t1.re - t3.im, t1.im + t3.re} : cplx<T>{t1.re + t3.im, t1.im - t3.re};
        a3 = INV ? cplx<T>{t1.re + t3.im, t1.im - t3.re} : cplx<T>{t1.re - t3.im, t1.im + t3.re};
    }
}

// In-register DFT of N <= 16 points, natural order in and out.  N = 16: radix 4
// (n = n0 + 4 n1, k = k1 + 4 k0: 4-point DFTs over n1, twiddles W_16^(n0 k1), 4-point
// DFTs over n0; 8 non-trivial twiddle products instead of radix 2's 10); smaller N:
// radix-2 decimation in frequency.  The final digit reversal only renames registers.
template <class T, int N, bool INV, bool HALF_ZERO = false>
__device__ __forceinline__ void dft_reg(cplx<T> (&x)[N]) {
    static_assert(N >= 1 && N <= 16 && (N & (N - 1)) == 0, "N must be a power of two <= 16");
    if constexpr (N == 16) {
#pragma unroll
        for (int n0 = 0; n0 < 4; ++n0) {
            cplx<T> a0 = x[n0], a1 = x[n0 + 4], a2, a3;
            if constexpr (!HALF_ZERO) {
                a2 = x[n0 + 8];
                a3 = x[n0 + 12];
            }
            dft4<T, INV, HALF_ZERO>(a0, a1, a2, a3);
            x[n0] = a0;
            x[n0 + 4] = twiddle16_any<T, INV>(a1, n0);
            x[n0 + 8] = twiddle16_any<T, INV>(a2, 2 * n0);
            x[n0 + 12] = twiddle16_any<T, INV>(a3, 3 * n0);
        }
        cplx<T> y[16];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            cplx<T> b0 = x[4 * k1], b1 = x[4 * k1 + 1], b2 = x[4 * k1 + 2], b3 = x[4 * k1 + 3];
            dft4<T, INV, false>(b0, b1, b2, b3);
            y[k1] = b0;
            y[k1 + 4] = b1;
            y[k1 + 8] = b2;
            y[k1 + 12] = b3;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = y[i];
        return;
    }
    if constexpr (HALF_ZERO && N >= 2) {
        // first pass with x[N/2..N-1] = 0: x[k] unchanged, x[k + N/2] = x[k] W^k
#pragma unroll
        for (int k = 0; k < N / 2; ++k) x[k + N / 2] = twiddle16<T, INV>(x[k], k * (16 / N));
    }
#pragma unroll
    for (int len = HALF_ZERO ? N / 2 : N; len >= 2; len >>= 1) {
        const int half = len / 2;
#pragma unroll
        for (int s0 = 0; s0 < N; s0 += len) {
#pragma unroll
            for (int k = 0; k < half; ++k) {
                const cplx<T> a = x[s0 + k], b = x[s0 + k + half];
                x[s0 + k] = cadd(a, b);
                x[s0 + k + half] = twiddle16<T, INV>(csub(a, b), k * (16 / len));
            }
        }
    }
    cplx<T> y[N];
#pragma unroll
    for (int i = 0; i < N; ++i) y[i] = x[brev_c(i, ilog2(N))];
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = y[i];
}

// Intra-wave LDS hand-off between the lanes of one FFT group.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Twiddle W_M^(n k) of the two-level transforms (n < L, k < R): from a table of
// W_M^m (m < M), or, TW2D, from a table laid out [n][k] with rows of R + 1
// (load_twiddles_2d), where the address is a per-lane base plus a compile-time offset
// and the odd row stride keeps the lanes of a group on distinct banks.
template <int R, bool TW2D>
__device__ __forceinline__ int tw_index(int n, int k) {
    return TW2D ? n * (R + 1) + k : n * k;
}

// Forward transform: lane n2 holds x[n1] = input[L*n1 + n2]; on return lane j holds
// X[(j + L*i) + R*k2] in x[i*L + k2].  scr: L*(R+1) elements of this group.
// HALF_ZERO: the inputs x[R/2..R-1] are zero padding.
template <class T, int R, int L, int M, bool HALF_ZERO, bool TW2D = false>
__device__ __forceinline__ void fft_fwd(cplx<T> (&x)[R], int lane, cplx<T>* scr, const cplx<T>* tw) {
    constexpr int Q = R / L;
    dft_reg<T, R, false, HALF_ZERO>(x);
#pragma unroll
    for (int k1 = 1; k1 < R; ++k1) x[k1] = cmul(x[k1], tw[tw_index<R, TW2D>(lane, k1)]);
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) scr[lane * (R + 1) + k1] = x[k1];
    wave_sync();
    cplx<T> y[Q][L];
#pragma unroll
    for (int i = 0; i < Q; ++i)
#pragma unroll
        for (int m = 0; m < L; ++m) y[i][m] = scr[m * (R + 1) + lane + L * i];
    wave_sync();
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        dft_reg<T, L, false>(y[i]);
#pragma unroll
        for (int k2 = 0; k2 < L; ++k2) x[i * L + k2] = y[i][k2];
    }
}

// Inverse (unnormalised) transform from the frequency layout to natural order:
// on return lane n2 holds output[L*n1 + n2] in x[n1].
// TW_FIRST: load each L-point DFT's twiddles before it.  tw and scr share LDS, so a
// twiddle load written after a store to scr stays after it and each product waits for
// its own load: hidden at four waves per SIMD (fft.hip), exposed at two (fft256.hip,
// where loading first is 4% faster).
template <class T, int R, int L, int M, bool TW2D = false, bool TW_FIRST = false>
__device__ __forceinline__ void fft_inv(cplx<T> (&x)[R], int lane, cplx<T>* scr, const cplx<T>* tw) {
    constexpr int Q = R / L;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int k1 = lane + L * i;
        cplx<T> w[L];
        if constexpr (TW_FIRST) {
#pragma unroll
            for (int n2 = 1; n2 < L; ++n2) w[n2] = tw[tw_index<R, TW2D>(n2, k1)];
        }
        cplx<T> y[L];
#pragma unroll
        for (int k2 = 0; k2 < L; ++k2) y[k2] = x[i * L + k2];
        dft_reg<T, L, true>(y);
#pragma unroll
        for (int n2 = 0; n2 < L; ++n2)
            scr[n2 * (R + 1) + k1] =
                n2 == 0 ? y[0] : cmul_conj(y[n2], TW_FIRST ? w[n2] : tw[tw_index<R, TW2D>(n2, k1)]);
    }
    wave_sync();
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) x[k1] = scr[lane * (R + 1) + k1];
    wave_sync();
    dft_reg<T, R, true>(x);
}

// Prepared pixel (src/prepare_pics.cu:36-41), no contraction.
// Workgroup maximum / minimum for arrays used once per kernel: one barrier (the
// trailing one that would guard the array's reuse is not needed).  Every thread of
// the workgroup must call them (the wave steps read all 64 lanes).
template <class T, int NT>
__device__ __forceinline__ T block_max_once(T v, T* sv) {
    v = wave_reduce(v, -(T)__builtin_huge_val(), [](T x, T y) { return fmax(x, y); });
    if (threadIdx.x % kWave == 0) sv[threadIdx.x / kWave] = v;
    __syncthreads();
    v = sv[0];
#pragma unroll
    for (int k = 1; k < NT / kWave; ++k) v = fmax(v, sv[k]);
    return v;
}
template <int NT>
__device__ __forceinline__ uint32_t block_min_once_u32(uint32_t v, uint32_t* sv) {
    v = wave_reduce(v, 0xffffffffu, [](uint32_t x, uint32_t y) { return min(x, y); });
    if (threadIdx.x % kWave == 0) sv[threadIdx.x / kWave] = v;
    __syncthreads();
    v = sv[0];
#pragma unroll
    for (int k = 1; k < NT / kWave; ++k) v = min(v, sv[k]);
    return v;
}

template <class T, class IN>
__device__ __forceinline__ T prep_px(IN px, T mean, T wx, T wy) {
#pragma clang fp contract(off)
    T v = (T)px;
    v -= mean;
    v = v * wx * wy;
    return v;
}

template <class T>
__device__ __forceinline__ T mean_of(T sum, uint32_t area) {
#pragma clang fp contract(off)
    return sum / (T)area;
}

// Workgroup -> slice (pattern * nb + roi).  Workgroups are dispatched round robin
// over the 8 XCDs (b mod 8); XCD x takes a contiguous run of the ROI-major order
// (roi * npat + pattern), so the patterns of a ROI run back to back on one XCD and
// that ROI's reference spectrum stays in the XCD's L2.
__device__ __forceinline__ uint32_t xcd_slice(uint32_t b, uint32_t total, uint32_t nb) {
    const uint32_t o = xcd_remap(b, total);
    const uint32_t npat = total / nb;
    return (o % npat) * nb + o / npat;
}

template <class T, class IN, class SUM>
struct fft_args {
    const IN* pics;          // patterns (FROM_PATTERN)
    const SUM* sums;         // per-slice sums (FROM_PATTERN, IN != uint16)
    const T* hx;             // Hann windows (FROM_PATTERN)
    const T* hy;
    const size2* begins;     // ROI corners
    size2 src;               // pattern size
    uint32_t side;           // slice side s (at most half the transform size)
    uint32_t nb;             // ROIs per pattern
    uint32_t total;          // slices in this launch
    const T* prepared;       // prepared s x s slices (!FROM_PATTERN)
    const cplx<T>* spectra;  // reference spectra (PIPE)
    cplx<T>* spectra_out;    // reference spectra (!PIPE)
    int fit;
    const double* lsq;
    data_index<T>* maxes;
    T* neigh;
    double* offsets;
    double* coefs;
    T* corr_out;             // optional (2s)^2 circular buffers in the reference layout, row-major
    const cplx<T>* twt;      // twiddle tables (fft_twiddles)
};

// Copy of the M-entry table into LDS.
template <class T, int M>
__device__ __forceinline__ void load_twiddles(cplx<T>* tw, const cplx<T>* table, int tid, int nt) {
    for (int t = tid; t < M; t += nt) tw[t] = table[M - 4 + t];
}
// The same values laid out [n][k] = W_M^(n k), n < M / R, k < R, rows of R + 1
// (M + M / R elements).
template <class T, int M, int R>
__device__ __forceinline__ void load_twiddles_2d(cplx<T>* tw, const cplx<T>* table, int tid, int nt) {
    for (int t = tid; t < M + M / R; t += nt) {
        const int n = t / (R + 1), k = t % (R + 1);
        tw[t] = table[M - 4 + (k < R ? n * k % M : 0)];
    }
}

}  // namespace emida_amd
