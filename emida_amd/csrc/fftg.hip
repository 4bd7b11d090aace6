// FFT cross-correlation for every slice shape the single-kernel pipelines (fft.hip,
// fft256.hip) do not take: non-square slices, double slices of side 65..256 and float
// slices of side 129..256 (transforms of up to 512 x 512).  cuFFT plans any
// 2s.x x 2s.y transform for the reference (src/gpu_offset.hpp:121-123,171-183); here
// the transform is Mx x My, each the power of two >= 2s of its axis, with the same
// correlation values (zero padding beyond 2s changes no shift |d| < s), read through
// the reference's 2s.x x 2s.y layout as in fft.hip.
//
// Three kernels per launch chunk, with a work buffer of half spectra in HBM
// ([slice][y][kx], kx <= Mx/2, the layout of an R2C result):
//  1. rows: two real rows packed as one complex row (prepared from the pattern or
//     read from prepared slices), an Mx-point transform, split into the two half
//     spectra (the split's factor 1/2 omitted: results come out doubled);
//  2. columns: a tile of adjacent half-spectrum columns staged through LDS, an
//     My-point transform each, the product with the conjugate reference spectrum
//     (src/hadamard.cu:39-40) and the inverse transform - or, for the reference, the
//     scaled spectrum stored ([roi][kx][ky], column-major, read by lanes in order);
//  3. rows, inverse: two output rows packed as Ga + i Gb (Hermitian extension, DC and
//     Nyquist imaginary parts dropped, as a C2R does), one inverse transform yields
//     both real rows, written as the reference's (2s.x x 2s.y) circular correlation
//     buffer; arg-max, gather (cross_res_pos_policy_fft, src/kernels.cuh:36-53), fit
//     and finalize follow in peak_kernel.
// Transforms are Stockham autosort FFTs (radix 4, one radix-2 pass for odd log2)
// in LDS, run by a group of lanes of one wave (Govindaraju et al., "High performance
// discrete Fourier transforms on graphics processors", SC 2008).
//
// Scale: the pattern's spectrum comes out doubled (the omitted split factor); the
// stored reference spectrum is 2 s.x s.y / (Mx My) times the true one (the same
// stages double it, launch_fftg_ref multiplies by s.x s.y / (Mx My)), so the
// unnormalised inverse carries the reference's 4 s.x s.y scale; all factors are
// exact powers of two when M = 2s.
//
// The same kernels give the reference's own cross_corr_fft (R2C, product with a
// spectrum in the R2C layout, C2R; src/gpu_offset.hpp:303-316) on 2s x 2s buffers
// (standard layouts, factor 1/2 applied to undo the doubling, exact).
#include "fft_common.hpp"

namespace emida_amd {

namespace {

constexpr int kGNT = 256;  // threads of the row kernels

// lanes per transform of N points: one radix-4 butterfly per lane and pass, at most a wave
__host__ __device__ constexpr int gsize(int n) { return n / 4 < 1 ? 1 : (n / 4 > kWave ? kWave : n / 4); }
// Stockham passes of N points (radix 4, plus one radix-2 pass for odd log2 N)
__host__ __device__ constexpr int npasses(int n) { return ilog2(n) / 2 + (ilog2(n) & 1); }

template <class T, bool INV>
__device__ __forceinline__ cplx<T> tw_at(const cplx<T>* tw, int t) {
    cplx<T> w = tw[t];
    if (INV) w.im = -w.im;
    return w;
}

// Radix-4 DFT of v (forward: W_4 = -i; inverse: +i).
template <class T, bool INV>
__device__ __forceinline__ void radix4(cplx<T> (&v)[4]) {
    const cplx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]), a2 = cadd(v[1], v[3]);
    const cplx<T> d = csub(v[1], v[3]);
    const cplx<T> a3 = INV ? cplx<T>{-d.im, d.re} : cplx<T>{d.im, -d.re};  // d * (+-i)
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
}

// Stockham autosort FFT of N points by the G lanes of one group (all in one wave):
// input in a (natural order), b scratch, both N complex in LDS; tw[t] = W_N^t =
// exp(-2 pi i t / N).  Returns the buffer holding the result in natural order (a
// when npasses(N) is even, b when odd).
template <class T, int N, bool INV, int G>
__device__ __forceinline__ cplx<T>* stockham(cplx<T>* a, cplx<T>* b, const cplx<T>* tw, int lane) {
    using C = cplx<T>;
    int ns = 1;
#pragma unroll
    for (int pass = 0; pass < ilog2(N) / 2; ++pass) {
        const int step = N / (4 * ns);
        for (int j = lane; j < N / 4; j += G) {
            const int k = j & (ns - 1);
            C v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = a[j + r * (N / 4)];
            if (ns > 1) {
#pragma unroll
                for (int r = 1; r < 4; ++r) v[r] = cmul(v[r], tw_at<T, INV>(tw, r * k * step));
            }
            radix4<T, INV>(v);
            const int d = (j - k) * 4 + k;
#pragma unroll
            for (int r = 0; r < 4; ++r) b[d + r * ns] = v[r];
        }
        wave_sync();
        C* t = a;
        a = b;
        b = t;
        ns *= 4;
    }
    if constexpr (ilog2(N) & 1) {  // ns = N / 2
        for (int j = lane; j < N / 2; j += G) {
            const C v0 = a[j], v1 = cmul(a[j + N / 2], tw_at<T, INV>(tw, j));
            b[j] = cadd(v0, v1);
            b[j + N / 2] = csub(v0, v1);
        }
        wave_sync();
        a = b;
    }
    return a;
}

// W_n^t for t < n from the twiddle tables (table of n at offset n - 4).
template <class T>
__device__ __forceinline__ void init_tw(cplx<T>* tw, const cplx<T>* table, int n, int tid, int nt) {
    for (int t = tid; t < n; t += nt) tw[t] = table[n - 4 + t];
}

}  // namespace

// Arguments of the three kernels (one launch chunk of slices first .. first + total - 1).
template <class T, class IN, class SUM>
struct fftg_args {
    const IN* pics;          // patterns (rows kernel, from pattern)
    const SUM* sums;         // per-slice sums (IN != uint16)
    const T* hx;             // Hann windows
    const T* hy;
    const size2* begins;     // ROI corners
    size2 src;               // pattern size
    const T* prepared;       // prepared slices (rows kernel, from prepared): [slice][rows][pitch]
    uint32_t pitch;          // prepared row pitch (elements)
    uint32_t in_rows, in_cols;  // rows / columns of input per slice (the rest is zero padding)
    uint32_t slice_elems;    // prepared elements per slice
    size2 side;              // slice size s
    uint32_t mx, my;         // transform sizes
    uint32_t nb;             // ROIs per pattern
    uint32_t first, total;   // slices of this chunk
    cplx<T>* work;           // [slice - first][my][mx/2 + 1]
    const cplx<T>* spectra;  // reference spectra (column kernel, product)
    cplx<T>* spectra_out;    // reference spectra (column kernel, spectrum out)
    T scale;                 // spectrum out: factor; product: factor on the product
    size2 out;               // output layout of the inverse rows kernel (2s.x x 2s.y)
    T* corr;                 // [slice - first][out.y][out.x]
    const cplx<T>* twt;      // twiddle tables (fft_twiddles)
};

// ---------------------------------------------------------------- kernel 1: rows
template <class T, class IN, class SUM, int MX, bool FROM_PATTERN>
__global__ __launch_bounds__(kGNT) void fftg_rows_kernel(fftg_args<T, IN, SUM> a) {
    using C = cplx<T>;
    constexpr int G = gsize(MX), NG = kGNT / G, NCX = MX / 2 + 1;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C* tw = reinterpret_cast<C*>(lds_raw);
    __shared__ uint32_t red_u[kGNT / kWave];
    const int tid = threadIdx.x, g = tid / G, lane = tid % G;
    C* A = tw + MX + g * 2 * MX;
    C* Bf = A + MX;
    const uint32_t sl = blockIdx.x, sn = a.first + sl;
    init_tw<T>(tw, a.twt, MX, tid, kGNT);

    const IN* pic = nullptr;
    const T* pre = nullptr;
    T mean = T(0);
    const int sx = (int)a.in_cols, sy = (int)a.in_rows;
    if constexpr (FROM_PATTERN) {
        const size2 b0 = a.begins[sn % a.nb];
        pic = a.pics + (size_t)(sn / a.nb) * a.src.area() + (size_t)b0.y * a.src.x + b0.x;
        if constexpr (std::is_same<IN, uint16_t>::value) {
            uint32_t su = 0;
            for (int y = tid / 64; y < sy; y += kGNT / 64)
                for (int x = tid % 64; x < sx; x += 64) su += pic[(size_t)y * a.src.x + x];
            su = wave_reduce(su, 0u, [](uint32_t x, uint32_t y) { return x + y; });
            if (tid % kWave == 0) red_u[tid / kWave] = su;
            __syncthreads();
            su = 0;
#pragma unroll
            for (int k = 0; k < kGNT / kWave; ++k) su += red_u[k];
            mean = mean_of<T>((T)su, sx * sy);
        } else {
            mean = mean_of<T>((T)a.sums[sn], sx * sy);
        }
    } else {
        pre = a.prepared + (size_t)sn * a.slice_elems;
    }
    __syncthreads();

    C* w = a.work + (size_t)sl * a.my * NCX;
    const int npairs = (sy + 1) / 2;
    for (int p0 = 0; p0 < npairs; p0 += NG) {
        const int p = p0 + g;
        const bool act = p < npairs, second = 2 * p + 1 < sy;
        for (int n = lane; n < MX; n += G) {
            T v0 = T(0), v1 = T(0);
            if (act && n < sx) {
                if constexpr (FROM_PATTERN) {
                    v0 = prep_px<T, IN>(pic[(size_t)(2 * p) * a.src.x + n], mean, a.hx[n], a.hy[2 * p]);
                    if (second) v1 = prep_px<T, IN>(pic[(size_t)(2 * p + 1) * a.src.x + n], mean, a.hx[n], a.hy[2 * p + 1]);
                } else {
                    v0 = pre[(size_t)(2 * p) * a.pitch + n];
                    if (second) v1 = pre[(size_t)(2 * p + 1) * a.pitch + n];
                }
            }
            A[n] = C{v0, v1};
        }
        wave_sync();
        const C* X = stockham<T, MX, false, G>(A, Bf, tw, lane);
        if (act) {
            C* r0 = w + (size_t)(2 * p) * NCX;
            for (int k = lane; k < NCX; k += G) {
                const C z = X[k], m = X[(MX - k) & (MX - 1)];
                r0[k] = C{z.re + m.re, z.im - m.im};
                if (second) r0[NCX + k] = C{z.im + m.im, m.re - z.re};
            }
        }
        wave_sync();  // X is the next row pair's input or scratch
    }
}

// ---------------------------------------------------------------- kernel 2: columns
enum { COLS_PIPE = 0, COLS_SPECTRUM = 1, COLS_STD_PIPE = 2, COLS_STD_SPECTRUM = 3 };

template <class T>
__host__ __device__ constexpr int col_tile(int my) {
    // columns per workgroup: at least 64 bytes of a row per load, at least one wave
    return (sizeof(T) == 4 ? 8 : 4) > kWave / gsize(my) ? (sizeof(T) == 4 ? 8 : 4) : kWave / gsize(my);
}

template <class T, int MY, int MODE>
__global__ __launch_bounds__(col_tile<T>(MY) * gsize(MY)) void fftg_cols_kernel(fftg_args<T, T, T> a) {
    using C = cplx<T>;
    constexpr int G = gsize(MY), CB = col_tile<T>(MY), NT = CB * G;
    constexpr bool RES_IN_B = npasses(MY) & 1;  // where stockham leaves a forward result
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C* tw = reinterpret_cast<C*>(lds_raw);
    C* tile = tw + MY;  // CB columns x 2 buffers x MY
    const int tid = threadIdx.x, g = tid / G, lane = tid % G;
    const uint32_t sl = blockIdx.y, sn = a.first + sl, roi = sn % a.nb;
    const int ncx = (int)a.mx / 2 + 1, kx0 = blockIdx.x * CB;
    init_tw<T>(tw, a.twt, MY, tid, NT);
    C* w = a.work + (size_t)sl * a.my * ncx;
    const int rows = MODE == COLS_SPECTRUM || MODE == COLS_PIPE ? (int)a.side.y : (int)a.my;
    // rows of the output the inverse rows kernel reads (the reference layout)
    const int sy = (int)a.side.y, pady = MY - 2 * sy;
    for (int e = tid; e < MY * CB; e += NT) {
        const int y = e / CB, c = e % CB;
        tile[c * 2 * MY + y] = (y < rows && kx0 + c < ncx) ? w[(size_t)y * ncx + kx0 + c] : C{T(0), T(0)};
    }
    __syncthreads();
    C* A = tile + g * 2 * MY;
    C* Bf = A + MY;
    C* X = stockham<T, MY, false, G>(A, Bf, tw, lane);
    const int kx = kx0 + g;
    const bool col_ok = kx < ncx;
    if constexpr (MODE == COLS_SPECTRUM || MODE == COLS_STD_SPECTRUM) {
        if (col_ok) {
            for (int k = lane; k < MY; k += G) {
                const C v = X[k];
                if constexpr (MODE == COLS_SPECTRUM)
                    a.spectra_out[((size_t)roi * ncx + kx) * MY + k] = C{v.re * a.scale, v.im * a.scale};
                else
                    a.spectra_out[((size_t)roi * MY + k) * ncx + kx] = C{v.re * a.scale, v.im * a.scale};
            }
        }
        return;
    } else {
        C* Y = RES_IN_B ? A : Bf;  // the other buffer
        for (int k = lane; k < MY; k += G) {
            C r = C{T(0), T(0)};
            if (col_ok) {
                if constexpr (MODE == COLS_PIPE) r = a.spectra[((size_t)roi * ncx + kx) * MY + k];
                else r = a.spectra[((size_t)roi * MY + k) * ncx + kx];
            }
            C h = hadamard(X[k], r);
            if constexpr (MODE == COLS_STD_PIPE) h = C{h.re * a.scale, h.im * a.scale};
            Y[k] = h;
        }
        wave_sync();
        C* Z = stockham<T, MY, true, G>(Y, X, tw, lane);
        __syncthreads();
        // every group's result sits in the same buffer of its column (same MY)
        const int zoff = (int)(Z - (tile + g * 2 * MY));
        for (int e = tid; e < MY * CB; e += NT) {
            const int y = e / CB, c = e % CB;
            const bool need = MODE == COLS_STD_PIPE || y <= sy || y > sy + pady;
            if (need && kx0 + c < ncx) w[(size_t)y * ncx + kx0 + c] = tile[c * 2 * MY + zoff + y];
        }
    }
}

// ---------------------------------------------------------------- kernel 3: rows, inverse
template <class T, int MX>
__global__ __launch_bounds__(kGNT) void fftg_rows_inv_kernel(fftg_args<T, T, T> a) {
    using C = cplx<T>;
    constexpr int G = gsize(MX), NG = kGNT / G, NCX = MX / 2 + 1, S = MX / 2;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C* tw = reinterpret_cast<C*>(lds_raw);
    const int tid = threadIdx.x, g = tid / G, lane = tid % G;
    C* A = tw + MX + g * 2 * MX;
    C* Bf = A + MX;
    const uint32_t sl = blockIdx.x;
    init_tw<T>(tw, a.twt, MX, tid, kGNT);
    __syncthreads();
    const C* w = a.work + (size_t)sl * a.my * NCX;
    T* out = a.corr + (size_t)sl * a.out.area();
    // reference-layout index k <= s is circular position k, k > s is k + M - 2s
    const int sx = (int)a.out.x / 2, sy = (int)a.out.y / 2;
    const int padx = MX - (int)a.out.x, pady = (int)a.my - (int)a.out.y;
    const int npairs = (int)a.out.y / 2;
    for (int q0 = 0; q0 < npairs; q0 += NG) {
        const int q = q0 + g;
        const bool act = q < npairs;
        const int ra = 2 * q, rb = 2 * q + 1;
        const int ya = act ? (ra <= sy ? ra : ra + pady) : 0, yb = act ? (rb <= sy ? rb : rb + pady) : 0;
        const C* ga_row = w + (size_t)ya * NCX;
        const C* gb_row = w + (size_t)yb * NCX;
        for (int k = lane; k < MX; k += G) {
            const bool mirror = k > S;
            const int col = mirror ? MX - k : k;
            C ga = ga_row[col], gb = gb_row[col];
            if (mirror) {
                ga.im = -ga.im;
                gb.im = -gb.im;
            }
            if (k == 0 || k == S) {
                ga.im = T(0);
                gb.im = T(0);
            }
            A[k] = C{ga.re - gb.im, ga.im + gb.re};
        }
        wave_sync();
        const C* X = stockham<T, MX, true, G>(A, Bf, tw, lane);
        if (act) {
            for (int rx = lane; rx < (int)a.out.x; rx += G) {
                const C v = X[rx <= sx ? rx : rx + padx];
                out[(size_t)ra * a.out.x + rx] = v.re;
                out[(size_t)rb * a.out.x + rx] = v.im;
            }
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------- launchers
namespace {

template <int N, class F>
void dispatch_n(uint32_t n, F&& f) {
    if constexpr (N <= 512) {
        if (n == (uint32_t)N) return f(std::integral_constant<int, N>{});
        return dispatch_n<N * 2>(n, std::forward<F>(f));
    } else {
        throw unsupported_error("FFT policy: transform side " + std::to_string(n) + " is beyond 512");
    }
}

template <class T>
size_t rows_lds(int mx) {
    return ((size_t)mx + (size_t)(kGNT / gsize(mx)) * 2 * mx) * sizeof(cplx<T>);
}
template <class T>
size_t cols_lds(int my) {
    return ((size_t)my + (size_t)col_tile<T>(my) * 2 * my) * sizeof(cplx<T>);
}

template <class K>
void set_lds(K kern, size_t bytes) {
    EMIDA_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

template <class T, class IN, class SUM, bool FP>
void run_rows(const fftg_args<T, IN, SUM>& a, hipStream_t st) {
    dispatch_n<4>(a.mx, [&](auto mc) {
        constexpr int MX = decltype(mc)::value;
        auto k = fftg_rows_kernel<T, IN, SUM, MX, FP>;
        static const bool once = (set_lds(k, rows_lds<T>(MX)), true);
        (void)once;
        k<<<a.total, kGNT, rows_lds<T>(MX), st>>>(a);
        EMIDA_LAUNCHED("fftg_rows_kernel");
    });
}

template <class T, int MODE>
void run_cols(const fftg_args<T, T, T>& a, hipStream_t st) {
    dispatch_n<4>(a.my, [&](auto mc) {
        constexpr int MY = decltype(mc)::value;
        auto k = fftg_cols_kernel<T, MY, MODE>;
        static const bool once = (set_lds(k, cols_lds<T>(MY)), true);
        (void)once;
        constexpr int CB = col_tile<T>(MY);
        const dim3 grid((unsigned)div_up((int)a.mx / 2 + 1, CB), a.total);
        k<<<grid, CB * gsize(MY), cols_lds<T>(MY), st>>>(a);
        EMIDA_LAUNCHED("fftg_cols_kernel");
    });
}

template <class T>
void run_rows_inv(const fftg_args<T, T, T>& a, hipStream_t st) {
    dispatch_n<4>(a.mx, [&](auto mc) {
        constexpr int MX = decltype(mc)::value;
        auto k = fftg_rows_inv_kernel<T, MX>;
        static const bool once = (set_lds(k, rows_lds<T>(MX)), true);
        (void)once;
        k<<<a.total, kGNT, rows_lds<T>(MX), st>>>(a);
        EMIDA_LAUNCHED("fftg_rows_inv_kernel");
    });
}

template <class T, class IN, class SUM>
fftg_args<T, T, T> as_t(const fftg_args<T, IN, SUM>& a) {
    fftg_args<T, T, T> b{};
    b.side = a.side;
    b.mx = a.mx;
    b.my = a.my;
    b.nb = a.nb;
    b.first = a.first;
    b.total = a.total;
    b.work = a.work;
    b.spectra = a.spectra;
    b.spectra_out = a.spectra_out;
    b.scale = a.scale;
    b.out = a.out;
    b.corr = a.corr;
    b.twt = a.twt;
    return b;
}

constexpr size_t kFftgWorkCap = (size_t)1 << 30;  // work buffer of one launch chunk

}  // namespace

bool fftg_needed(bool is_double, size2 side) {
    if (side.x != side.y) return true;
    const uint32_t m = fft_padded(side.x);
    return m > 256 || (m > 128 && is_double);
}

size_t fftg_spectrum_bytes(bool is_double, size2 side) {
    return (size_t)(fft_padded(side.x) / 2 + 1) * fft_padded(side.y) * (is_double ? 16 : 8);
}

size_t fftg_slice_work_bytes(bool is_double, size2 side) { return fftg_spectrum_bytes(is_double, side); }

size_t fftg_chunk(bool is_double, size2 side, uint32_t total, size_t extra_per_slice) {
    const size_t per = fftg_slice_work_bytes(is_double, side) + extra_per_slice;
    const size_t c = std::max<size_t>(1, kFftgWorkCap / per);
    return std::min<size_t>(c, total);
}

size_t fftg_ref_work_bytes(bool is_double, size2 side, uint32_t nb) {
    return fftg_chunk(is_double, side, nb, 0) * fftg_slice_work_bytes(is_double, side);
}

size_t fftg_work_bytes(bool is_double, size2 side, uint32_t total, bool with_corr) {
    const size_t corr = with_corr ? (size_t)4 * side.area() * (is_double ? 8 : 4) : 0;
    return fftg_chunk(is_double, side, total, corr) * (fftg_slice_work_bytes(is_double, side) + corr);
}

// Reference spectra ([roi][kx][ky], 2 s.x s.y / (Mx My) times the true spectrum) from
// the reference pattern; work holds fftg_ref_work_bytes bytes.
template <class T, class IN, class S>
void launch_fftg_ref(const IN* pics, const S* sums, const T* hx, const T* hy, const size2* begins, size2 src,
                     size2 side, uint32_t nb, cplx<T>* spectra, void* work, hipStream_t st) {
    if (nb == 0) return;
    fftg_args<T, IN, S> a{};
    a.twt = fft_twiddles<T>();
    a.pics = pics;
    a.sums = sums;
    a.hx = hx;
    a.hy = hy;
    a.begins = begins;
    a.src = src;
    a.in_rows = side.y;
    a.in_cols = side.x;
    a.side = side;
    a.mx = fft_padded(side.x);
    a.my = fft_padded(side.y);
    a.nb = nb;
    a.first = 0;
    a.total = nb;
    a.work = (cplx<T>*)work;
    a.spectra_out = spectra;
    // the row and column stages double it; s.x s.y / (Mx My) is exact
    a.scale = (T)((double)side.x * side.y / ((double)a.mx * a.my));
    const size_t chunk = fftg_chunk(sizeof(T) == 8, side, nb, 0);
    for (uint32_t f = 0; f < nb; f += (uint32_t)chunk) {
        a.first = f;
        a.total = (uint32_t)std::min<size_t>(chunk, nb - f);
        run_rows<T, IN, S, true>(a, st);
        run_cols<T, COLS_SPECTRUM>(as_t(a), st);
    }
}

// Pipeline of `total` slices: correlation buffers into corr (caller's, the whole
// launch) or the work buffer, then the peak kernel.
template <class T, class IN, class S>
void launch_fftg_pipe(const IN* pics, const S* sums, const T* hx, const T* hy, const size2* begins, size2 src,
                      size2 side, uint32_t nb, uint32_t total, const cplx<T>* spectra, int fit, data_index<T>* maxes,
                      T* neigh, double* offsets, double* coefs, T* corr_out, void* work, hipStream_t st) {
    if (total == 0) return;
    const bool dbl = sizeof(T) == 8;
    const size2 out{2 * side.x, 2 * side.y};
    const size_t corr_per = (size_t)out.area() * sizeof(T);
    // the chunk fftg_work_bytes(.., with_corr) sized the work buffer for
    const size_t chunk = fftg_chunk(dbl, side, total, corr_per);
    char* wb = (char*)work;
    T* corr_chunk = corr_out ? nullptr : (T*)(wb + chunk * fftg_slice_work_bytes(dbl, side));
    for (uint32_t f = 0; f < total; f += (uint32_t)chunk) {
        fftg_args<T, IN, S> a{};
        a.twt = fft_twiddles<T>();
        a.pics = pics;
        a.sums = sums;
        a.hx = hx;
        a.hy = hy;
        a.begins = begins;
        a.src = src;
        a.in_rows = side.y;
        a.in_cols = side.x;
        a.side = side;
        a.mx = fft_padded(side.x);
        a.my = fft_padded(side.y);
        a.nb = nb;
        a.first = f;
        a.total = (uint32_t)std::min<size_t>(chunk, total - f);
        a.work = (cplx<T>*)work;
        a.spectra = spectra;
        a.out = out;
        a.corr = corr_out ? corr_out + (size_t)f * out.area() : corr_chunk;
        run_rows<T, IN, S, true>(a, st);
        const fftg_args<T, T, T> b = as_t(a);
        run_cols<T, COLS_PIPE>(b, st);
        run_rows_inv<T>(b, st);
        launch_peak<T>(a.corr, (size_t)a.total * out.area(), out.area(), size2{out.x - 1, out.y - 1}, POLICY_FFT,
                       fit, maxes ? maxes + f : nullptr, neigh ? neigh + (size_t)f * fit * fit : nullptr,
                       offsets ? offsets + 2 * (size_t)f : nullptr, coefs ? coefs + 6 * (size_t)f : nullptr, a.total,
                       st);
    }
}

// Prepared s.x x s.y slices -> (2s.x) x (2s.y) circular correlation buffers
// (cross_corr_fft seam); spectra: nb spectra; work: chunk slices of half spectra.
template <class T>
void launch_fftg_corr_prepared(const T* pics, const T* ref, T* res, size2 side, uint32_t nb, uint32_t batch,
                               cplx<T>* spectra, void* work, size_t chunk, hipStream_t st) {
    fftg_args<T, T, T> a{};
    a.twt = fft_twiddles<T>();
    a.prepared = ref;
    a.pitch = side.x;
    a.slice_elems = side.area();
    a.in_rows = side.y;
    a.in_cols = side.x;
    a.side = side;
    a.mx = fft_padded(side.x);
    a.my = fft_padded(side.y);
    a.nb = nb;
    a.work = (cplx<T>*)work;
    a.spectra_out = spectra;
    a.scale = (T)((double)side.x * side.y / ((double)a.mx * a.my));
    a.out = size2{2 * side.x, 2 * side.y};
    for (uint32_t f = 0; f < nb; f += (uint32_t)chunk) {  // reference spectra, chunk by chunk
        a.first = f;
        a.total = (uint32_t)std::min<size_t>(chunk, nb - f);
        a.work = (cplx<T>*)work;
        run_rows<T, T, T, false>(a, st);
        fftg_args<T, T, T> b = a;
        b.work = (cplx<T>*)work;
        run_cols<T, COLS_SPECTRUM>(b, st);
    }
    // the spectrum kernel indexes ROIs from the slice index: shift it per chunk
    const uint32_t total = nb * batch;
    for (uint32_t f = 0; f < total; f += (uint32_t)chunk) {
        fftg_args<T, T, T> b = a;
        b.prepared = pics;
        b.first = f;
        b.total = (uint32_t)std::min<size_t>(chunk, total - f);
        b.spectra = spectra;
        b.spectra_out = nullptr;
        b.corr = res + (size_t)f * b.out.area();
        run_rows<T, T, T, false>(b, st);
        run_cols<T, COLS_PIPE>(b, st);
        run_rows_inv<T>(b, st);
    }
}

// The reference's own cross_corr_fft on its buffers (src/gpu_offset.hpp:303-316):
// R2C of every 2s.x x 2s.y real buffer of `pic`, product with the conjugate of the
// ROI's spectrum in `spectra` (R2C layout [ky][kx], kx <= s.x), unnormalised C2R into
// `res` (2s.x x 2s.y).  Both transform sides must be powers of two.
template <class T>
void launch_fftg_std_corr(const T* pic, const cplx<T>* spectra, T* res, size2 side, uint32_t nb, uint32_t total,
                          void* work, size_t chunk, hipStream_t st) {
    fftg_args<T, T, T> a{};
    a.twt = fft_twiddles<T>();
    a.prepared = pic;
    a.pitch = 2 * side.x;
    a.slice_elems = 4 * side.area();
    a.in_rows = 2 * side.y;
    a.in_cols = 2 * side.x;
    a.side = size2{side.x, side.y};
    a.mx = 2 * side.x;
    a.my = 2 * side.y;
    if (fft_padded(side.x) != a.mx || fft_padded(side.y) != a.my)
        throw unsupported_error("cross_corr_fft: 2s must be a power of two on both axes");
    a.nb = nb;
    a.spectra = spectra;
    a.scale = (T)0.5;  // undoes the doubling of the row and column stages (exact)
    a.out = size2{2 * side.x, 2 * side.y};
    a.work = (cplx<T>*)work;
    for (uint32_t f = 0; f < total; f += (uint32_t)chunk) {
        fftg_args<T, T, T> b = a;
        b.first = f;
        b.total = (uint32_t)std::min<size_t>(chunk, total - f);
        b.corr = res + (size_t)f * b.out.area();
        run_rows<T, T, T, false>(b, st);
        run_cols<T, COLS_STD_PIPE>(b, st);
        run_rows_inv<T>(b, st);
    }
}

// R2C of 2s.x x 2s.y real buffers (rows of `pitch` elements, 2s.y rows) into the R2C
// layout ([ky][kx], kx <= s.x) of `spectra`, unnormalised (cuFFT's R2C).
template <class T>
void launch_fftg_std_r2c(const T* pic, uint32_t pitch, cplx<T>* spectra, size2 side, uint32_t n, void* work,
                         size_t chunk, hipStream_t st) {
    fftg_args<T, T, T> a{};
    a.twt = fft_twiddles<T>();
    a.prepared = pic;
    a.pitch = pitch;
    a.slice_elems = pitch * 2 * side.y;
    a.in_rows = 2 * side.y;
    a.in_cols = 2 * side.x;
    a.side = side;
    a.mx = 2 * side.x;
    a.my = 2 * side.y;
    if (fft_padded(side.x) != a.mx || fft_padded(side.y) != a.my)
        throw unsupported_error("cross_corr_fft: 2s must be a power of two on both axes");
    a.nb = n;
    a.scale = (T)0.5;
    a.work = (cplx<T>*)work;
    a.spectra_out = spectra;
    for (uint32_t f = 0; f < n; f += (uint32_t)chunk) {
        fftg_args<T, T, T> b = a;
        b.first = f;
        b.total = (uint32_t)std::min<size_t>(chunk, n - f);
        run_rows<T, T, T, false>(b, st);
        run_cols<T, COLS_STD_SPECTRUM>(b, st);
    }
}

#define INST_FFTG(T, IN, S)                                                                                     \
    template void launch_fftg_ref<T, IN, S>(const IN*, const S*, const T*, const T*, const size2*, size2, size2, \
                                            uint32_t, cplx<T>*, void*, hipStream_t);                            \
    template void launch_fftg_pipe<T, IN, S>(const IN*, const S*, const T*, const T*, const size2*, size2, size2,\
                                             uint32_t, uint32_t, const cplx<T>*, int, data_index<T>*, T*, double*, \
                                             double*, T*, void*, hipStream_t);
INST_FFTG(float, uint16_t, uint32_t)
INST_FFTG(double, uint16_t, uint32_t)
INST_FFTG(float, float, float)
INST_FFTG(double, double, double)
INST_FFTG(float, double, double)
INST_FFTG(double, float, float)

template void launch_fftg_corr_prepared<float>(const float*, const float*, float*, size2, uint32_t, uint32_t,
                                               cplx<float>*, void*, size_t, hipStream_t);
template void launch_fftg_corr_prepared<double>(const double*, const double*, double*, size2, uint32_t, uint32_t,
                                                cplx<double>*, void*, size_t, hipStream_t);
template void launch_fftg_std_corr<float>(const float*, const cplx<float>*, float*, size2, uint32_t, uint32_t, void*,
                                          size_t, hipStream_t);
template void launch_fftg_std_corr<double>(const double*, const cplx<double>*, double*, size2, uint32_t, uint32_t,
                                           void*, size_t, hipStream_t);
template void launch_fftg_std_r2c<float>(const float*, uint32_t, cplx<float>*, size2, uint32_t, void*, size_t,
                                         hipStream_t);
template void launch_fftg_std_r2c<double>(const double*, uint32_t, cplx<double>*, size2, uint32_t, void*, size_t,
                                          hipStream_t);

}  // namespace emida_amd
