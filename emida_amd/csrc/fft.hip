// Single-kernel FFT cross-correlation pipeline (LDS resident, one workgroup per
// ROI of one pattern).
//
// Reference: src/gpu_offset.hpp:303-316 (R2C of the zero-padded 2s x 2s slice,
// multiply by the conjugate reference spectrum, unnormalised C2R),
// src/hadamard.cu:27-41 (the product p * conj(r)), src/maxarg.cu and
// src/extract_neighbors.cu with cross_res_pos_policy_fft (src/kernels.cuh:36-53)
// for the circular result layout, src/subpixel_max.hpp for the fit.  The reference
// runs these as separate cuFFT plans and kernels with the correlation buffers in
// global memory; here one workgroup keeps the whole 2-D transform of one ROI in
// LDS and never writes the correlation to HBM.
//
// Transform layout (M = 2s, power of two, S = M/2 = slice side, NC = M/2 + 1):
//  1. rows: rows 2p and 2p+1 are packed as one complex row (re, im) and
//     transformed with an in-place radix-2 decimation-in-frequency FFT whose first
//     pass is pruned (the upper half of every row is zero padding).  The rows live
//     in the still unused upper halves of the spectrum columns.
//  2. split: the two real-row spectra are separated from the packed spectrum
//     (X[k] and conj(X[M-k])) into columns kx = 0..M/2 of the half spectrum.
//  3. columns: DIF FFT over y (first pass pruned), result in bit-reversed row
//     order; the reference spectrum is stored in the same order, so the product
//     needs no reordering.  The last DIF pass, the product and the first pass of
//     the inverse decimation-in-time FFT run on the same register pair.
//  4. inverse rows: rows 2q and 2q+1 of the column result become one complex
//     sequence Ga + i Gb (Hermitian extension of each half spectrum, imaginary
//     parts of the DC and Nyquist terms dropped as a C2R does); one inverse FFT
//     yields both real output rows.
//  5. arg-max over the (2s)^2 circular buffer, F x F gather through the FFT
//     position policy, double-precision least-squares fit and finalize.
#include <type_traits>

#include "kernels.hpp"

namespace emida_amd {

template <class T>
__device__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) {
    return {a.re + b.re, a.im + b.im};
}
template <class T>
__device__ inline cplx<T> csub(cplx<T> a, cplx<T> b) {
    return {a.re - b.re, a.im - b.im};
}
template <class T>
__device__ inline cplx<T> cmul(cplx<T> a, cplx<T> w) {
    return {a.re * w.re - a.im * w.im, a.re * w.im + a.im * w.re};
}
template <class T>
__device__ inline cplx<T> cmul_conj(cplx<T> a, cplx<T> w) {  // a * conj(w)
    return {a.re * w.re + a.im * w.im, a.im * w.re - a.re * w.im};
}

template <int M>
__device__ inline uint32_t bitrev(uint32_t k) {
    constexpr int L = __builtin_ctz(M);
    return __brev(k) >> (32 - L);
}

// Prepared pixel (src/prepare_pics.cu:36-41), no contraction.
template <class T, class IN>
__device__ inline T prep_px(IN px, T mean, T wx, T wy) {
#pragma clang fp contract(off)
    T v = (T)px;
    v -= mean;
    v = v * wx * wy;
    return v;
}

template <class T>
__device__ inline T mean_of(T sum, uint32_t area) {
#pragma clang fp contract(off)
    return sum / (T)area;
}

template <class T, class IN, class SUM>
struct fft_args {
    const IN* pics;        // patterns (FROM_PATTERN)
    const SUM* sums;       // per-slice sums (FROM_PATTERN, IN != uint16)
    const T* hx;           // Hann windows (FROM_PATTERN)
    const T* hy;
    const size2* begins;   // ROI corners
    size2 src;             // pattern size
    uint32_t nb;           // ROIs per pattern
    uint32_t total;        // slices in this launch
    const T* prepared;     // prepared slices (!FROM_PATTERN), s x s each
    const cplx<T>* spectra;  // reference spectra (PIPE)
    cplx<T>* spectra_out;    // reference spectra (!PIPE)
    int fit;
    const double* lsq;
    data_index<T>* maxes;
    T* neigh;
    double* offsets;
    double* coefs;
    T* corr_out;           // optional (2s)^2 circular buffers
};

template <class T, int M>
struct fft_geom {
    static constexpr int S = M / 2;
    static constexpr int NC = M / 2 + 1;
    static constexpr int CS = M + 1;  // column stride in complex elements
    static constexpr size_t lds_bytes = (size_t)(NC * CS + M / 2) * sizeof(cplx<T>);
    // row stage work area (packed row pair p, position x) in the upper column halves
    __device__ static inline int rw(int p, int x) {
        const int t = p * M + x;
        return (t / S) * CS + S + (t % S);
    }
    // inverse row work area (row pair q, position x)
    __device__ static inline int iw(int q, int x) { return (x >> 1) * CS + 2 * q + (x & 1); }
};

template <class T, class IN, class SUM, int M, int NT, bool FROM_PATTERN, bool PIPE>
__global__ __launch_bounds__(NT) void fft_kernel(fft_args<T, IN, SUM> a) {
    using G = fft_geom<T, M>;
    using C = cplx<T>;
    constexpr int S = G::S, NC = G::NC, CS = G::CS;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C* B = reinterpret_cast<C*>(lds_raw);
    C* tw = B + NC * CS;
    __shared__ uint32_t red_u[NT / kWave];
    __shared__ T sv[NT / kWave];
    __shared__ uint32_t si[NT / kWave];
    __shared__ T nbv[81];
    __shared__ double cf[6];

    const uint32_t sn = blockIdx.x;
    if (sn >= a.total) return;
    const uint32_t bn = sn % a.nb;
    const int tid = threadIdx.x;

    // twiddles W_M^j = exp(-2 pi i j / M), j < M/2, from double precision
    for (int j = tid; j < M / 2; j += NT) {
        double s, c;
        sincospi(2.0 * j / M, &s, &c);
        tw[j] = C{(T)c, (T)(-s)};
    }

    // ---- 0. input rows (prepared values), packed pairs, pruned first DIF pass
    if constexpr (FROM_PATTERN) {
        const size2 b0 = a.begins[bn];
        const IN* pic = a.pics + (size_t)(sn / a.nb) * a.src.area() + (size_t)b0.y * a.src.x + b0.x;
        T mean;
        if constexpr (std::is_same<IN, uint16_t>::value) {
            uint32_t su = 0;
            for (int e = tid; e < S * S; e += NT) su += pic[(size_t)(e / S) * a.src.x + (e % S)];
#pragma unroll
            for (int off = kWave / 2; off > 0; off >>= 1) su += __shfl_xor(su, off);
            if (tid % kWave == 0) red_u[tid / kWave] = su;
            __syncthreads();
            su = 0;
#pragma unroll
            for (int k = 0; k < NT / kWave; ++k) su += red_u[k];
            mean = mean_of<T>((T)su, S * S);
        } else {
            mean = mean_of<T>((T)a.sums[sn], S * S);
        }
        __syncthreads();  // twiddles visible
        for (int e = tid; e < (S / 2) * S; e += NT) {
            const int p = e / S, x = e % S;
            const T wx = a.hx[x];
            const T va = prep_px<T, IN>(pic[(size_t)(2 * p) * a.src.x + x], mean, wx, a.hy[2 * p]);
            const T vb = prep_px<T, IN>(pic[(size_t)(2 * p + 1) * a.src.x + x], mean, wx, a.hy[2 * p + 1]);
            const C z{va, vb};
            B[G::rw(p, x)] = z;
            B[G::rw(p, x + S)] = cmul(z, tw[x]);
        }
    } else {
        __syncthreads();
        const T* src = a.prepared + (size_t)sn * S * S;
        for (int e = tid; e < (S / 2) * S; e += NT) {
            const int p = e / S, x = e % S;
            const C z{src[(2 * p) * S + x], src[(2 * p + 1) * S + x]};
            B[G::rw(p, x)] = z;
            B[G::rw(p, x + S)] = cmul(z, tw[x]);
        }
    }
    __syncthreads();

    // ---- 1. row DIF passes, span S down to 2
#pragma unroll
    for (int len = S; len >= 2; len >>= 1) {
        const int half = len / 2;
        for (int b = tid; b < (S / 2) * (M / 2); b += NT) {
            const int p = b / (M / 2), j = b % (M / 2);
            const int k = j % half;
            const int i0 = (j / half) * len + k, i1 = i0 + half;
            const C w = tw[k * (M / len)];
            const C x0 = B[G::rw(p, i0)], x1 = B[G::rw(p, i1)];
            B[G::rw(p, i0)] = cadd(x0, x1);
            B[G::rw(p, i1)] = cmul(csub(x0, x1), w);
        }
        __syncthreads();
    }

    // ---- 2. split packed spectra into half-spectrum columns
    for (int e = tid; e < (S / 2) * NC; e += NT) {
        const int p = e / NC, k = e % NC;
        const C zk = B[G::rw(p, bitrev<M>(k))];
        const C zm = B[G::rw(p, bitrev<M>((M - k) & (M - 1)))];
        const C ra{(zk.re + zm.re) * (T)0.5, (zk.im - zm.im) * (T)0.5};
        const C rb{(zk.im + zm.im) * (T)0.5, (zm.re - zk.re) * (T)0.5};
        B[k * CS + 2 * p] = ra;
        B[k * CS + 2 * p + 1] = rb;
    }
    __syncthreads();

    // ---- 3. column DIF: pruned first pass, then spans S .. 4 (2 is fused below)
    for (int e = tid; e < NC * S; e += NT) {
        const int kx = e / S, y = e % S;
        B[kx * CS + y + S] = cmul(B[kx * CS + y], tw[y]);
    }
    __syncthreads();
#pragma unroll
    for (int len = S; len >= 4; len >>= 1) {
        const int half = len / 2;
        for (int b = tid; b < NC * (M / 2); b += NT) {
            const int kx = b / (M / 2), j = b % (M / 2);
            const int k = j % half;
            const int i0 = kx * CS + (j / half) * len + k, i1 = i0 + half;
            const C w = tw[k * (M / len)];
            const C x0 = B[i0], x1 = B[i1];
            B[i0] = cadd(x0, x1);
            B[i1] = cmul(csub(x0, x1), w);
        }
        __syncthreads();
    }

    if constexpr (!PIPE) {
        // last DIF pass and store of the reference spectrum (bit-reversed rows)
        C* out = a.spectra_out + (size_t)bn * NC * M;
        for (int b = tid; b < NC * (M / 2); b += NT) {
            const int kx = b / (M / 2), j = b % (M / 2);
            const C x0 = B[kx * CS + 2 * j], x1 = B[kx * CS + 2 * j + 1];
            out[kx * M + 2 * j] = cadd(x0, x1);
            out[kx * M + 2 * j + 1] = csub(x0, x1);
        }
        return;
    } else {
        // ---- 4. last DIF pass, product with conj(reference), first DIT pass
        const C* ref = a.spectra + (size_t)bn * NC * M;
        for (int b = tid; b < NC * (M / 2); b += NT) {
            const int kx = b / (M / 2), j = b % (M / 2);
            const int i0 = kx * CS + 2 * j;
            const C x0 = B[i0], x1 = B[i0 + 1];
            const C r0 = ref[kx * M + 2 * j], r1 = ref[kx * M + 2 * j + 1];
            const C h0 = cmul_conj(cadd(x0, x1), r0);
            const C h1 = cmul_conj(csub(x0, x1), r1);
            B[i0] = cadd(h0, h1);
            B[i0 + 1] = csub(h0, h1);
        }
        __syncthreads();

        // ---- 5. column inverse DIT, spans 4 .. M
#pragma unroll
        for (int len = 4; len <= M; len <<= 1) {
            const int half = len / 2;
            for (int b = tid; b < NC * (M / 2); b += NT) {
                const int kx = b / (M / 2), j = b % (M / 2);
                const int k = j % half;
                const int i0 = kx * CS + (j / half) * len + k, i1 = i0 + half;
                const C w = tw[k * (M / len)];
                const C x0 = B[i0];
                const C x1 = cmul_conj(B[i1], w);
                B[i0] = cadd(x0, x1);
                B[i1] = csub(x0, x1);
            }
            __syncthreads();
        }

        // ---- 6. inverse rows: Z = Ga + i Gb in bit-reversed order
        constexpr int E = (S * M + NT - 1) / NT;
        C zr[E];
#pragma unroll
        for (int it = 0; it < E; ++it) {
            const int e = tid + it * NT;
            if (e < S * M) {
                const int k = e / S, q = e % S;
                C ga, gb;
                if (k <= S) {
                    ga = B[k * CS + 2 * q];
                    gb = B[k * CS + 2 * q + 1];
                } else {
                    const C ca = B[(M - k) * CS + 2 * q], cb = B[(M - k) * CS + 2 * q + 1];
                    ga = C{ca.re, -ca.im};
                    gb = C{cb.re, -cb.im};
                }
                if (k == 0 || k == S) {
                    ga.im = 0;
                    gb.im = 0;
                }
                zr[it] = C{ga.re - gb.im, ga.im + gb.re};
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < E; ++it) {
            const int e = tid + it * NT;
            if (e < S * M) {
                const int k = e / S, q = e % S;
                B[G::iw(q, bitrev<M>(k))] = zr[it];
            }
        }
        __syncthreads();
#pragma unroll
        for (int len = 2; len <= M; len <<= 1) {
            const int half = len / 2;
            for (int b = tid; b < S * (M / 2); b += NT) {
                const int q = b % S, j = b / S;
                const int k = j % half;
                const int x0i = (j / half) * len + k, x1i = x0i + half;
                const C w = tw[k * (M / len)];
                const C x0 = B[G::iw(q, x0i)];
                const C x1 = cmul_conj(B[G::iw(q, x1i)], w);
                B[G::iw(q, x0i)] = cadd(x0, x1);
                B[G::iw(q, x1i)] = csub(x0, x1);
            }
            __syncthreads();
        }

        // ---- 7. arg-max over the circular (2s)^2 buffer (lowest index on ties)
        T best = -std::numeric_limits<T>::infinity();
        uint32_t bi = 0xffffffffu;
        T* cout = a.corr_out ? a.corr_out + (size_t)sn * M * M : nullptr;
        for (int idx = tid; idx < M * M; idx += NT) {
            const int y = idx / M, x = idx % M;
            const C z = B[G::iw(y >> 1, x)];
            const T v = (y & 1) ? z.im : z.re;
            if (cout) cout[idx] = v;
            if (v > best) {
                best = v;
                bi = (uint32_t)idx;
            }
        }
        block_argmax<T, NT>(best, bi, sv, si);

        // ---- 8. gather (FFT position policy, clamp to [0, 2s-1]) and fit
        const int fit = a.fit, kk = fit * fit, r = (fit - 1) / 2;
        const size2 cross{(uint32_t)M - 1, (uint32_t)M - 1};
        const size2 mp = shift_pos(POLICY_FFT, bi, cross);
        for (int t = tid; t < kk; t += NT) {
            const int i = t % fit, j = t / fit;
            const size2 c = clamp_pos((int)mp.x - r + i, (int)mp.y - r + j, cross);
            const int qx = (int)((c.x + S + 1) % M), qy = (int)((c.y + S + 1) % M);
            const C z = B[G::iw(qy >> 1, qx)];
            const T v = (qy & 1) ? z.im : z.re;
            nbv[t] = v;
            if (a.neigh) a.neigh[(size_t)sn * kk + t] = v;
        }
        __syncthreads();
        if (tid < 6) {
#pragma clang fp contract(off)
            const double* mrow = a.lsq + tid * kk;
            double sum = 0;
            for (int i = 0; i < kk; ++i) sum += mrow[i] * (double)nbv[i];
            cf[tid] = sum;
        }
        __syncthreads();
        if (tid == 0) {
#pragma clang fp contract(off)
            if (a.maxes) {
                a.maxes[sn].data = best;
                a.maxes[sn].index = bi;
            }
            const double b = cf[1], c = cf[2], d = cf[3], e = cf[4], f = cf[5];
            const double py = (b * e - 2 * c * d) / (4 * f * d - e * e);
            const double px = (-e * py - b) / (2 * d);
            const double ox = -((int)mp.x - ((int)cross.x / 2) - r + px);
            const double oy = -((int)mp.y - ((int)cross.y / 2) - r + py);
            if (a.offsets) {
                a.offsets[2 * (size_t)sn] = ox;
                a.offsets[2 * (size_t)sn + 1] = oy;
            }
            if (a.coefs)
                for (int k = 0; k < 6; ++k) a.coefs[6 * (size_t)sn + k] = cf[k];
        }
    }
}

// --------------------------------------------------------------- launchers
constexpr int kFftNT = 256;
constexpr size_t kMaxLds = 160 * 1024;

int fft_max_side(bool is_double) {
    (void)is_double;
    return 128;
}

template <class T, class IN, class SUM, int M, bool FP, bool PIPE>
void launch_fft_m(const fft_args<T, IN, SUM>& a, uint32_t grid, hipStream_t st) {
    using G = fft_geom<T, M>;
    static_assert(G::lds_bytes <= kMaxLds, "FFT working set exceeds LDS");
    auto kern = fft_kernel<T, IN, SUM, M, kFftNT, FP, PIPE>;
    EMIDA_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)G::lds_bytes));
    kern<<<grid, kFftNT, G::lds_bytes, st>>>(a);
    EMIDA_LAUNCHED("fft_kernel");
}

template <class T, class IN, class SUM, bool FP, bool PIPE>
void launch_fft_side(uint32_t side, const fft_args<T, IN, SUM>& a, uint32_t grid, hipStream_t st) {
    switch (2 * side) {
        case 4: return launch_fft_m<T, IN, SUM, 4, FP, PIPE>(a, grid, st);
        case 8: return launch_fft_m<T, IN, SUM, 8, FP, PIPE>(a, grid, st);
        case 16: return launch_fft_m<T, IN, SUM, 16, FP, PIPE>(a, grid, st);
        case 32: return launch_fft_m<T, IN, SUM, 32, FP, PIPE>(a, grid, st);
        case 64: return launch_fft_m<T, IN, SUM, 64, FP, PIPE>(a, grid, st);
        case 128: return launch_fft_m<T, IN, SUM, 128, FP, PIPE>(a, grid, st);
        default:
            throw unsupported_error("FFT policy: slice side must be a power of two between 2 and 64 (got " +
                                    std::to_string(side) + ")");
    }
}

template <class T, class IN, class S>
void launch_fft_ref(const IN* pics, const S* sums, const T* hx, const T* hy, const size2* begins, size2 src,
                    uint32_t side, uint32_t nb, cplx<T>* spectra, hipStream_t st) {
    fft_args<T, IN, S> a{};
    a.pics = pics;
    a.sums = sums;
    a.hx = hx;
    a.hy = hy;
    a.begins = begins;
    a.src = src;
    a.nb = nb;
    a.total = nb;
    a.spectra_out = spectra;
    if (nb == 0) return;
    launch_fft_side<T, IN, S, true, false>(side, a, nb, st);
}

template <class T, class IN, class S>
void launch_fft_pipe(const IN* pics, const S* sums, const T* hx, const T* hy, const size2* begins, size2 src,
                     uint32_t side, uint32_t nb, uint32_t total, const cplx<T>* spectra, int fit,
                     data_index<T>* maxes, T* neigh, double* offsets, double* coefs, T* corr_out, hipStream_t st) {
    fft_args<T, IN, S> a{};
    a.pics = pics;
    a.sums = sums;
    a.hx = hx;
    a.hy = hy;
    a.begins = begins;
    a.src = src;
    a.nb = nb;
    a.total = total;
    a.spectra = spectra;
    a.fit = fit;
    a.lsq = lstsq_device_table() + lstsq_offset(fit);
    a.maxes = maxes;
    a.neigh = neigh;
    a.offsets = offsets;
    a.coefs = coefs;
    a.corr_out = corr_out;
    if (total == 0) return;
    launch_fft_side<T, IN, S, true, true>(side, a, total, st);
}

template <class T>
void launch_fft_corr_prepared(const T* pics, const T* ref, T* res, uint32_t side, uint32_t ref_slices,
                              uint32_t batch, hipStream_t st) {
    if (ref_slices == 0 || batch == 0) return;
    const size_t nc = side + 1, m = 2 * (size_t)side;
    cplx<T>* spectra = nullptr;
    EMIDA_HIP(hipMallocAsync((void**)&spectra, nc * m * ref_slices * sizeof(cplx<T>), st));
    fft_args<T, T, T> a{};
    a.nb = ref_slices;
    a.total = ref_slices;
    a.prepared = ref;
    a.spectra_out = spectra;
    launch_fft_side<T, T, T, false, false>(side, a, ref_slices, st);
    fft_args<T, T, T> b{};
    b.nb = ref_slices;
    b.total = ref_slices * batch;
    b.prepared = pics;
    b.spectra = spectra;
    b.fit = 3;
    b.lsq = lstsq_device_table();
    b.corr_out = res;
    launch_fft_side<T, T, T, false, true>(side, b, b.total, st);
    EMIDA_HIP(hipFreeAsync(spectra, st));
}

// p = p * conj(r), the reference formula (src/hadamard.cu:39-40); ref repeats
// every one.area() * ref_slices elements.
template <class T>
__global__ void hadamard_kernel(cplx<T>* __restrict__ pics, const cplx<T>* __restrict__ ref, size_t ref_len,
                                size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const cplx<T> p = pics[i], r = ref[i % ref_len];
    pics[i] = cplx<T>{p.re * r.re + p.im * r.im, -p.re * r.im + p.im * r.re};
}

template <class T>
void launch_hadamard(cplx<T>* pics, const cplx<T>* ref, size2 one, uint32_t ref_slices, uint32_t batch,
                     hipStream_t st) {
    const size_t ref_len = (size_t)one.area() * ref_slices, n = ref_len * batch;
    if (n == 0) return;
    hadamard_kernel<T><<<(unsigned)div_up(n, (size_t)256), 256, 0, st>>>(pics, ref, ref_len, n);
    EMIDA_LAUNCHED("hadamard_kernel");
}

#define INST_FFT(T, IN, S)                                                                                    \
    template void launch_fft_ref<T, IN, S>(const IN*, const S*, const T*, const T*, const size2*, size2,      \
                                           uint32_t, uint32_t, cplx<T>*, hipStream_t);                        \
    template void launch_fft_pipe<T, IN, S>(const IN*, const S*, const T*, const T*, const size2*, size2,     \
                                            uint32_t, uint32_t, uint32_t, const cplx<T>*, int,                \
                                            data_index<T>*, T*, double*, double*, T*, hipStream_t);
INST_FFT(float, uint16_t, uint32_t)
INST_FFT(double, uint16_t, uint32_t)
INST_FFT(float, float, float)
INST_FFT(double, double, double)
INST_FFT(float, double, double)
INST_FFT(double, float, float)

template void launch_fft_corr_prepared<float>(const float*, const float*, float*, uint32_t, uint32_t, uint32_t,
                                              hipStream_t);
template void launch_fft_corr_prepared<double>(const double*, const double*, double*, uint32_t, uint32_t,
                                               uint32_t, hipStream_t);
template void launch_hadamard<float>(cplx<float>*, const cplx<float>*, size2, uint32_t, uint32_t, hipStream_t);
template void launch_hadamard<double>(cplx<double>*, const cplx<double>*, size2, uint32_t, uint32_t,
                                      hipStream_t);

}  // namespace emida_amd
