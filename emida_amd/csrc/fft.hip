// Single-kernel FFT cross-correlation pipeline (one workgroup per ROI of one
// pattern, everything resident in LDS and registers).
//
// Reference: src/gpu_offset.hpp:303-316 (R2C of the zero-padded 2s x 2s slice,
// product with the conjugate reference spectrum, unnormalised C2R),
// src/hadamard.cu:27-41 (p * conj(r)), src/maxarg.cu and src/extract_neighbors.cu with
// cross_res_pos_policy_fft (src/kernels.cuh:36-53) for the circular result layout,
// src/subpixel_max.hpp for the fit.  The reference runs separate cuFFT plans and
// kernels with every intermediate in global memory; here nothing but the inputs, the
// reference spectra and the per-ROI results touch HBM.
//
// Transform size.  The reference transforms the zero-padded 2s x 2s slice.  Here M is
// the power of two >= 2s; the padding beyond 2s changes no correlation value (shifts
// |d| < s are the linear correlation for any M >= 2s - 1), and the arg-max, the
// gather and the optional result buffer read the M x M circular result through the
// reference's 2s x 2s layout: its index k <= s is circular position k, k > s is
// position k + M - 2s (shift k - 2s).  For power-of-two slices M = 2s and both
// layouts coincide.
//
// FFT scheme.  A length-M transform is computed by a group of L lanes of
// one wave, each lane holding R = M/L values: an R-point DFT in registers over
// n = L*n1 + n2 (lane n2), a twiddle W_M^(n2*k1), one transpose through the group's
// LDS scratch, and L-point DFTs in registers.  Lane j then holds X[k] for
// k = (j + L*i) + R*k2 (i < R/L, k2 < L) - the "frequency layout".  The
// inverse starts from that layout and ends in natural order, so the product with the
// reference spectrum (stored in the same layout) needs no reordering.
//
// Stages for one ROI (s = slice side <= S = M/2, NC = S + 1 half-spectrum columns):
//  0. load the s x s ROI, integer sum (uint16) or given sum, prepare
//     ((px - mean) * hx * hy, the reference's order) into LDS;
//  A. rows 2p and 2p+1 packed as one complex row: ceil(s/2) forward transforms;
//  B. column kx of the half spectrum, split from the packed rows
//     (X[k] and conj(X[M-k])); columns 0 and S are real and packed into one
//     complex column; forward transform, product with conj(reference), inverse.
//     Both split steps omit their factor 1/2, so column spectra come out doubled;
//     the stored reference spectrum is scaled by 1/4 (half its true value), and the
//     product is exact in scale (powers of two only); for a padded transform the
//     stored spectrum also carries (2s/M)^2, which gives the unnormalised result the
//     reference's (2s)^2 scale (one rounding of the stored spectrum, none when M = 2s);
//  C. rows 2q and 2q+1 of the column result packed as Ga + i Gb (Hermitian
//     extension, DC and Nyquist imaginary parts dropped as a C2R does), inverse
//     transform yields both real correlation rows;
//  D. arg-max over the (2s)^2 circular buffer (reference layout), F x F gather through the FFT
//     position policy, double-precision least-squares fit, finalize.
#include <mutex>
#include <vector>

#include "fft_common.hpp"

namespace emida_amd {

// Bound of the kernel's static LDS (reductions, neighbourhood, least-squares rows).
constexpr size_t kStaticLds = 5120;

// Geometry of one transform size.
template <class T, int M>
struct fft_plan {
    static constexpr int S = M / 2;                  // largest slice side
    static constexpr int NC = S + 1;                 // half-spectrum columns
    static constexpr int L = M >= 64 ? 8 : (M >= 16 ? 4 : 2);  // lanes per transform
    static constexpr int R = M / L;                  // values per lane
    static constexpr int Q = R / L;
    static_assert(Q >= 1 && R * L == M, "bad transform split");
    static constexpr int NT = S * L < 64 ? 64 : S * L;
    // LDS slot stride (complex elements): holds a column of M values or one group's
    // transpose scratch; congruent to 8 mod 16 so the groups of a wave spread over
    // all banks.
    static constexpr int need = (L * (R + 1) > M ? L * (R + 1) : M);
    static constexpr int CS = ((need + 7) / 16) * 16 + 8;
    static constexpr int P_OFF = NC * CS - S * S / 2;  // prepared ROI (S*S reals) at the end
    static_assert(P_OFF >= (S / 2) * CS, "prepared ROI overlaps the row stage");
    static constexpr size_t lds_bytes = (size_t)(NC * CS + M + L) * sizeof(cplx<T>);  // + twiddles [L][R + 1]
    // waves per SIMD the LDS allows (160 KiB per CU, 4 SIMDs); registers are
    // budgeted for exactly that many
    static constexpr int WG_PER_CU = (int)((160 * 1024) / (lds_bytes + kStaticLds));
    static constexpr int WAVES_PER_SIMD_RAW = WG_PER_CU * (NT / 64) / 4;
    static constexpr int WAVES_PER_SIMD = WAVES_PER_SIMD_RAW < 1 ? 1 : (WAVES_PER_SIMD_RAW > 4 ? 4 : WAVES_PER_SIMD_RAW);
};

static __device__ cplx<float> g_tw_f[2 * kTwMaxM - 4];
static __device__ cplx<double> g_tw_d[2 * kTwMaxM - 4];

template <class T>
const cplx<T>* fft_twiddles() {
    static std::mutex mu;
    static std::vector<std::pair<const void*, const void*>> per_device;
    int dev = 0;
    EMIDA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)per_device.size() <= dev) per_device.resize(dev + 1, {nullptr, nullptr});
    if (!per_device[dev].first) {
        std::vector<cplx<float>> hf(2 * kTwMaxM - 4);
        std::vector<cplx<double>> hd(2 * kTwMaxM - 4);
        const long double pi = 3.141592653589793238462643383279502884L;
        for (int m = 4; m <= kTwMaxM; m *= 2)
            for (int t = 0; t < m; ++t) {
                const long double a = 2 * pi * t / m;
                long double c = cosl(a), sn = sinl(a);
                if (4 * t % m == 0) {  // multiples of a quarter turn: exact
                    c = (t == 0) ? 1 : (2 * t == m ? -1 : 0);
                    sn = (4 * t == m) ? 1 : (4 * t == 3 * m ? -1 : 0);
                }
                hd[m - 4 + t] = cplx<double>{(double)c, (double)-sn};
                hf[m - 4 + t] = cplx<float>{(float)c, (float)-sn};
            }
        EMIDA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_tw_f), hf.data(), hf.size() * sizeof(cplx<float>)));
        EMIDA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_tw_d), hd.data(), hd.size() * sizeof(cplx<double>)));
        void *pf = nullptr, *pd = nullptr;
        EMIDA_HIP(hipGetSymbolAddress(&pf, HIP_SYMBOL(g_tw_f)));
        EMIDA_HIP(hipGetSymbolAddress(&pd, HIP_SYMBOL(g_tw_d)));
        if (!pf || !pd) throw hip_error("hipGetSymbolAddress returned no address for the twiddle tables");
        per_device[dev] = {pf, pd};
    }
    return (const cplx<T>*)(sizeof(T) == 8 ? per_device[dev].second : per_device[dev].first);
}
template const cplx<float>* fft_twiddles<float>();
template const cplx<double>* fft_twiddles<double>();

// The larger of two transform outputs, a NaN operand ignored.  The outputs are
// arithmetic results, never signalling NaNs, so float takes v_max_f32 as is (IEEE
// maxNum) without the quieting of both operands fmax adds.
__device__ __forceinline__ float max_quiet(float a, float b) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double max_quiet(double a, double b) { return fmax(a, b); }

// The fit's least-squares rows (6 x kk doubles, kk = F^2 <= 81) in LDS: the loads are
// issued at the start (clamped addresses, no branch around them), stored once the
// first stage waits anyway, read by the fit at the end.
template <int NT, int N>
__device__ __forceinline__ void load_lsq(double (&v)[N], const double* lsq, int kk, int tid) {
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = kk > 0 ? lsq[min(tid + q * NT, 6 * kk - 1)] : 0.0;
}
template <int NT, int N>
__device__ __forceinline__ void store_lsq(double* dst, const double (&v)[N], int kk, int tid) {
#pragma unroll
    for (int q = 0; q < N; ++q)
        if (tid + q * NT < 6 * kk) dst[tid + q * NT] = v[q];
}

// PAD: the slice side is below M/2 (a runtime value); otherwise it is M/2 and every
// index expression below folds to constants.
template <class T, class IN, class SUM, int M, bool FROM_PATTERN, bool PIPE, bool PAD>
__global__ __launch_bounds__((fft_plan<T, M>::NT), (fft_plan<T, M>::WAVES_PER_SIMD)) void fft_kernel(fft_args<T, IN, SUM> a) {
    using P = fft_plan<T, M>;
    using C = cplx<T>;
    constexpr int S = P::S, L = P::L, R = P::R, Q = P::Q, CS = P::CS, NT = P::NT;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C* B = reinterpret_cast<C*>(lds_raw);
    C* tw = B + P::NC * CS;
    T* pre = reinterpret_cast<T*>(B + P::P_OFF);
    __shared__ uint32_t red_u[NT / kWave];
    __shared__ T sv[NT / kWave];
    __shared__ uint32_t si[NT / kWave];
    __shared__ T nbv[81];
    __shared__ double lsq_s[6 * 81];
    __shared__ double cf[6];
    constexpr int kLsqPer = (6 * 81 + NT - 1) / NT;  // least-squares elements per thread

    if (blockIdx.x >= a.total) return;
    // workgroup-uniform (a scalar register: addresses built from it stay scalar)
    const uint32_t sn = __builtin_amdgcn_readfirstlane(xcd_slice(blockIdx.x, a.total, a.nb));
    const uint32_t bn = sn % a.nb;
    const int tid = threadIdx.x;
    const int lane = tid % L, grp = tid / L;
    const int s = PAD ? (int)a.side : S;  // slice side, 2 <= s <= S
    const int pad = M - 2 * s;            // reference index k > s sits at circular position k + pad
    // stored reference spectra: x 1/4 (the omitted split factors) x (2s/M)^2, so an
    // M-point C2R carries the reference's (2s)^2 scale; s^2 / M^2 is exact in T
    const T sc = (T)(0.25 * ((double)(2 * s) / M) * ((double)(2 * s) / M));

    // ---- 0. prepared ROI into LDS; every global load of the stage is issued before
    // the first use (the window factors and the fit's least-squares rows too)
    const int kk = PIPE ? a.fit * a.fit : 0;
    if constexpr (FROM_PATTERN) {
        // the corner is workgroup-uniform: in scalar registers, every address below
        // is a scalar base plus a lane offset
        const size2 b0{(uint32_t)__builtin_amdgcn_readfirstlane(a.begins[bn].x),
                       (uint32_t)__builtin_amdgcn_readfirstlane(a.begins[bn].y)};
        const IN* pic = a.pics + (size_t)(sn / a.nb) * a.src.area() + (size_t)b0.y * a.src.x + b0.x;
        constexpr int E = (S * S + NT - 1) / NT;
        IN px[E];
        T wx[E], wy[E];
        if constexpr (!PAD && S == kWave) {
            // element tid + it NT is row w + it NT/64 (w = the wave), column = lane:
            // the row part of every address is wave-uniform (scalar registers)
            const uint32_t w = __builtin_amdgcn_readfirstlane((uint32_t)tid / kWave), col = (uint32_t)tid % kWave;
            const T hxv = a.hx[col];
#pragma unroll
            for (int it = 0; it < E; ++it) {
                const uint32_t row = w + it * (NT / kWave);
                const IN* prow = pic + (size_t)row * a.src.x;  // uniform
                px[it] = prow[col];
                wx[it] = hxv;
                wy[it] = a.hy[row];
            }
        } else {
#pragma unroll
            for (int it = 0; it < E; ++it) {
                const int e = PAD ? min(tid + it * NT, s * s - 1) : tid + it * NT;  // clamped: masked below
                px[it] = pic[(size_t)(e / s) * a.src.x + (e % s)];
                wx[it] = a.hx[e % s];
                wy[it] = a.hy[e / s];
            }
        }
        double lv[kLsqPer];
        load_lsq<NT>(lv, a.lsq, kk, tid);
        load_twiddles_2d<T, M, R>(tw, a.twt, tid, NT);
        store_lsq<NT>(lsq_s, lv, kk, tid);
        T mean;
        if constexpr (std::is_same<IN, uint16_t>::value) {
            uint32_t su = 0;
#pragma unroll
            for (int it = 0; it < E; ++it) su += tid + it * NT < s * s ? (uint32_t)px[it] : 0u;
            su = wave_reduce(su, 0u, [](uint32_t x, uint32_t y) { return x + y; });
            if (tid % kWave == 0) red_u[tid / kWave] = su;
            __syncthreads();
            su = 0;
#pragma unroll
            for (int k = 0; k < NT / kWave; ++k) su += red_u[k];
            mean = mean_of<T>((T)su, s * s);
        } else {
            mean = mean_of<T>((T)a.sums[sn], s * s);
        }
#pragma unroll
        for (int it = 0; it < E; ++it) {
            const int e = tid + it * NT;
            if (e < s * s) pre[e] = prep_px<T, IN>(px[it], mean, wx[it], wy[it]);
        }
    } else {
        const T* src = a.prepared + (size_t)sn * s * s;
        for (int e = tid; e < s * s; e += NT) pre[e] = src[e];
        double lv[kLsqPer];
        load_lsq<NT>(lv, a.lsq, kk, tid);
        load_twiddles_2d<T, M, R>(tw, a.twt, tid, NT);
        store_lsq<NT>(lsq_s, lv, kk, tid);
    }
    __syncthreads();

    // this group's column of the reference spectrum (column job c = grp of stage B),
    // loaded here so that the row stage and the column transform cover its latency
    // (half of it before the column transform and half after it: 4% slower; half here
    // and half before the column transform: 1.4% slower)
    C spec[R];
    if constexpr (PIPE) {
        if (grp < S) {
            const C* rc = a.spectra + ((size_t)bn * P::NC + grp) * M + lane * R;
#pragma unroll
            for (int r = 0; r < R; ++r) spec[r] = rc[r];
        }
    }
    // ---- A. packed row pairs, forward (an odd s leaves the last pair's second row zero)
    if (grp < (s + 1) / 2) {
        const int p = grp;
        const bool second = 2 * p + 1 < s;
        C x[R];
#pragma unroll
        for (int n1 = 0; n1 < R; ++n1) {
            const int n = L * n1 + lane;
            if (n1 < R / 2 && n < s) {
                x[n1] = C{pre[(2 * p) * s + n], second ? pre[(2 * p + 1) * s + n] : T(0)};
            } else {
                x[n1] = C{T(0), T(0)};
            }
        }
        C* slot = B + p * CS;
        fft_fwd<T, R, L, M, true, true>(x, lane, slot, tw);
#pragma unroll
        for (int i = 0; i < Q; ++i)
#pragma unroll
            for (int k2 = 0; k2 < L; ++k2) slot[(lane + L * i) + R * k2] = x[i * L + k2];
    }
    __syncthreads();

    // ---- B. half-spectrum columns
    C x[R];
    const int c = grp;  // column job (0 = packed columns 0 and S)
    if (c < S) {
        // loads and selects without branches (a branch per element would wait for
        // each element's loads in turn); the parity of row y is the lane's (L is even);
        // rows y >= s read in-bounds slots and are replaced by zero
        const bool odd = lane & 1;
        const int o2 = c == 0 ? S : M - c;
#pragma unroll
        for (int n1 = 0; n1 < R / 2; ++n1) {
            const int y = L * n1 + lane;
            const C* row = B + (y >> 1) * CS;
            const C zk = row[c], zm = row[o2];
            const C v = odd ? C{zk.im + zm.im, zm.re - zk.re} : C{zk.re + zm.re, zk.im - zm.im};
            x[n1] = y < s ? v : C{T(0), T(0)};
        }
#pragma unroll
        for (int n1 = R / 2; n1 < R; ++n1) x[n1] = C{T(0), T(0)};
        if (c == 0) {
            // packed real columns 0 and S (wave 0 only): Z[0] and Z[S] of the row pair
#pragma unroll
            for (int n1 = 0; n1 < R / 2; ++n1) {
                const int y = L * n1 + lane;
                const C* row = B + (y >> 1) * CS;
                const C z0 = row[0], zs = row[S];
                const C v = odd ? C{z0.im, zs.im} : C{z0.re, zs.re};
                x[n1] = y < s ? v : C{T(0), T(0)};
            }
        }
    }
    __syncthreads();  // row spectra consumed; slots become column scratch
    if (c < S) {
        C* slot = B + c * CS;
        fft_fwd<T, R, L, M, true, true>(x, lane, slot, tw);
        if (c == 0) {
            // split the packed real columns: P0[k] = (Z[k] + conj(Z[-k]))/2,
            // PS[k] = (Z[k] - conj(Z[-k]))/(2i), then H0 + i HS (PIPE) or the two
            // reference spectra, element by element
#pragma unroll
            for (int i = 0; i < Q; ++i)
#pragma unroll
                for (int k2 = 0; k2 < L; ++k2) slot[(lane + L * i) + R * k2] = x[i * L + k2];
            wave_sync();
            const size_t c0 = ((size_t)bn * P::NC + 0) * M + lane * R, cs = ((size_t)bn * P::NC + S) * M + lane * R;
#pragma unroll
            for (int i = 0; i < Q; ++i)
#pragma unroll
                for (int k2 = 0; k2 < L; ++k2) {
                    const int k = (lane + L * i) + R * k2, r = i * L + k2;
                    const C z = x[r], zm = slot[(M - k) & (M - 1)];
                    const C p0 = C{z.re + zm.re, z.im - zm.im}, ps = C{z.im + zm.im, zm.re - z.re};
                    if constexpr (PIPE) {
                        const C h0 = hadamard(p0, spec[r]), hs = hadamard(ps, a.spectra[cs + r]);
                        x[r] = C{h0.re - hs.im, h0.im + hs.re};  // H0 + i HS
                    } else {
                        a.spectra_out[c0 + r] = C{p0.re * sc, p0.im * sc};
                        a.spectra_out[cs + r] = C{ps.re * sc, ps.im * sc};
                    }
                }
            wave_sync();
        } else {
            if constexpr (PIPE) {
#pragma unroll
                for (int r = 0; r < R; ++r) x[r] = hadamard(x[r], spec[r]);
            } else {
                C* oc = a.spectra_out + ((size_t)bn * P::NC + c) * M + lane * R;
#pragma unroll
                for (int r = 0; r < R; ++r) oc[r] = C{x[r].re * sc, x[r].im * sc};
            }
        }
        if constexpr (PIPE) {
            fft_inv<T, R, L, M, true>(x, lane, slot, tw);
            if (c == 0) {
                C* slot_s = B + S * CS;
#pragma unroll
                for (int n1 = 0; n1 < R; ++n1) {
                    const int y = L * n1 + lane;
                    slot[y] = C{x[n1].re, T(0)};
                    slot_s[y] = C{x[n1].im, T(0)};
                }
            } else {
#pragma unroll
                for (int n1 = 0; n1 < R; ++n1) slot[L * n1 + lane] = x[n1];
            }
        }
    }
    if constexpr (!PIPE) {
        return;
    } else {
        __syncthreads();

        // ---- C. packed output row pairs, inverse
        const int q = grp;
        if (q < S) {
#pragma unroll
            for (int i = 0; i < Q; ++i)
#pragma unroll
                for (int k2 = 0; k2 < L; ++k2) {
                    const int k = (lane + L * i) + R * k2;
                    const bool mirror = k > S;
                    const C* col = B + (mirror ? M - k : k) * CS;
                    C ga = col[2 * q], gb = col[2 * q + 1];
                    if (mirror) {
                        ga.im = -ga.im;
                        gb.im = -gb.im;
                    }
                    if (k == 0 || k == S) {
                        ga.im = T(0);
                        gb.im = T(0);
                    }
                    x[i * L + k2] = C{ga.re - gb.im, ga.im + gb.re};
                }
        }
        __syncthreads();  // column results consumed; slots become row scratch
        // arg-max in two steps: the largest value (v_max, NaN ignored), then the
        // lowest index holding it, searched only by lanes that hold it
        // over the positions of the reference layout only (all of them when M = 2s),
        // with reference indices
        const uint32_t s2 = 2 * s;
        const int ya = 2 * q, yb = 2 * q + 1;
        const bool in_a = !PAD || ya <= s || ya > s + pad, in_b = !PAD || yb <= s || yb > s + pad;
        const uint32_t ra = (uint32_t)(ya <= s ? ya : ya - pad) * s2, rb = (uint32_t)(yb <= s ? yb : yb - pad) * s2;
        T vmax = -(T)__builtin_huge_val();
        if (q < S && (in_a || in_b)) {
            C* slot = B + q * CS;
            fft_inv<T, R, L, M, true>(x, lane, slot, tw);
            // the maximum through selects (no branch per element); positions outside
            // the reference layout count as -inf (there are none when M = 2s)
            const T lo = -(T)__builtin_huge_val();
#pragma unroll
            for (int n1 = 0; n1 < R; ++n1) {
                const int xx = L * n1 + lane;
                slot[xx] = x[n1];
                const bool in_x = !PAD || xx <= s || xx > s + pad;
                vmax = max_quiet(vmax, in_x && in_a ? x[n1].re : lo);
                vmax = max_quiet(vmax, in_x && in_b ? x[n1].im : lo);
            }
            if (a.corr_out) {
                T* cout = a.corr_out + (size_t)sn * s2 * s2;
#pragma unroll
                for (int n1 = 0; n1 < R; ++n1) {
                    const int xx = L * n1 + lane;
                    if (xx <= s || xx > s + pad) {
                        const uint32_t rx = (uint32_t)(xx <= s ? xx : xx - pad);
                        if (in_a) cout[ra + rx] = x[n1].re;
                        if (in_b) cout[rb + rx] = x[n1].im;
                    }
                }
            }
        }
        const T best = block_max_once<T, NT>(vmax, sv);
        uint32_t bi = 0xffffffffu;
        if (q < S && (in_a || in_b) && vmax == best) {
#pragma unroll
            for (int n1 = 0; n1 < R; ++n1) {
                const int xx = L * n1 + lane;
                if (xx <= s || xx > s + pad) {
                    const uint32_t rx = (uint32_t)(xx <= s ? xx : xx - pad);
                    if (in_a && x[n1].re == best) bi = min(bi, ra + rx);
                    if (in_b && x[n1].im == best) bi = min(bi, rb + rx);
                }
            }
        }
        bi = block_min_once_u32<NT>(bi, si);
        // the rest of the slice is wave 0's (its lanes exchange through LDS with
        // wave-level ordering only)
        if (tid >= kWave) return;

        // ---- D. gather (FFT position policy, clamp to [0, 2s-1]) and fit
        const int fit = a.fit, r = (fit - 1) / 2;
        const size2 cross{s2 - 1, s2 - 1};
        const size2 mp = shift_pos(POLICY_FFT, bi, cross);
        for (int t = tid; t < kk; t += kWave) {
            const int i = t % fit, j = t / fit;
            const size2 cc = clamp_pos((int)mp.x - r + i, (int)mp.y - r + j, cross);
            int qx = (int)((cc.x + s + 1) % s2), qy = (int)((cc.y + s + 1) % s2);
            if (qx > s) qx += pad;
            if (qy > s) qy += pad;
            const C z = B[(qy >> 1) * CS + qx];
            const T v = (qy & 1) ? z.im : z.re;
            nbv[t] = v;
            if (a.neigh) a.neigh[(size_t)sn * kk + t] = v;
        }
        wave_sync();
        if (tid < 6) {
#pragma clang fp contract(off)
            const double* mrow = lsq_s + tid * kk;
            double sum = 0;
            for (int i = 0; i < kk; ++i) sum += mrow[i] * (double)nbv[i];
            cf[tid] = sum;
        }
        wave_sync();
        if (tid == 0) {
#pragma clang fp contract(off)
            if (a.maxes) {
                a.maxes[sn].data = best;
                a.maxes[sn].index = bi;
            }
            const double b = cf[1], c2 = cf[2], d = cf[3], e = cf[4], f = cf[5];
            const double py = (b * e - 2 * c2 * d) / (4 * f * d - e * e);
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
constexpr size_t kMaxLds = 160 * 1024;

int fft_max_side(bool is_double) { return is_double ? 128 : 256; }

uint32_t fft_padded(uint32_t side) {
    uint32_t m = 4;
    while (m < 2 * side) m *= 2;
    return m;
}

template <class IN, class SUM, bool FP, bool PIPE>
void launch_fft256(fft_args<float, IN, SUM> a, hipStream_t st);

template <class T, class IN, class SUM, int M, bool FP, bool PIPE, bool PAD>
void launch_fft_mp(const fft_args<T, IN, SUM>& a, uint32_t grid, hipStream_t st) {
    using P = fft_plan<T, M>;
    static_assert(P::lds_bytes + kStaticLds <= kMaxLds, "FFT working set exceeds LDS");
    auto kern = fft_kernel<T, IN, SUM, M, FP, PIPE, PAD>;
    static const hipError_t attr =
        hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)P::lds_bytes);
    EMIDA_HIP(attr);
    kern<<<grid, P::NT, P::lds_bytes, st>>>(a);
    EMIDA_LAUNCHED("fft_kernel");
}

template <class T, class IN, class SUM, int M, bool FP, bool PIPE>
void launch_fft_m(const fft_args<T, IN, SUM>& a, uint32_t grid, hipStream_t st) {
    if (2 * a.side == (uint32_t)M) launch_fft_mp<T, IN, SUM, M, FP, PIPE, false>(a, grid, st);
    else launch_fft_mp<T, IN, SUM, M, FP, PIPE, true>(a, grid, st);
}

template <class T, class IN, class SUM, bool FP, bool PIPE>
void launch_fft_side(uint32_t side, fft_args<T, IN, SUM> a, uint32_t grid, hipStream_t st) {
    if (side < 2) throw unsupported_error("FFT policy: slice side must be at least 2");
    a.side = side;
    switch (fft_padded(side)) {
        case 4: return launch_fft_m<T, IN, SUM, 4, FP, PIPE>(a, grid, st);
        case 8: return launch_fft_m<T, IN, SUM, 8, FP, PIPE>(a, grid, st);
        case 16: return launch_fft_m<T, IN, SUM, 16, FP, PIPE>(a, grid, st);
        case 32: return launch_fft_m<T, IN, SUM, 32, FP, PIPE>(a, grid, st);
        case 64: return launch_fft_m<T, IN, SUM, 64, FP, PIPE>(a, grid, st);
        case 128: return launch_fft_m<T, IN, SUM, 128, FP, PIPE>(a, grid, st);
        case 256:
            if constexpr (std::is_same<T, float>::value) return launch_fft256<IN, SUM, FP, PIPE>(a, st);
            throw unsupported_error("FFT policy: slice sides above 64 need T = float");
        default:
            throw unsupported_error("FFT policy: slice side must be between 2 and 128 (got " +
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
    a.twt = fft_twiddles<T>();
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
    a.twt = fft_twiddles<T>();
    if (total == 0) return;
    launch_fft_side<T, IN, S, true, true>(side, a, total, st);
}

// Reference spectra from prepared reference slices, then the correlation of every
// prepared pattern slice.
template <class T>
static void fft_corr_prepared(const T* pics, const T* ref, T* res, uint32_t side, uint32_t ref_slices,
                              uint32_t batch, cplx<T>* spectra, hipStream_t st) {
    fft_args<T, T, T> a{};
    a.nb = ref_slices;
    a.total = ref_slices;
    a.prepared = ref;
    a.spectra_out = spectra;
    a.twt = fft_twiddles<T>();
    launch_fft_side<T, T, T, false, false>(side, a, ref_slices, st);
    fft_args<T, T, T> b{};
    b.nb = ref_slices;
    b.total = ref_slices * batch;
    b.prepared = pics;
    b.spectra = spectra;
    b.fit = 3;
    b.lsq = lstsq_device_table();
    b.corr_out = res;
    b.twt = a.twt;
    launch_fft_side<T, T, T, false, true>(side, b, b.total, st);
}

template <class T>
static size_t spectra_bytes(uint32_t side, uint32_t ref_slices) {
    const size_t m = fft_padded(side);
    return (m / 2 + 1) * m * ref_slices * sizeof(cplx<T>);
}

template <class T>
void launch_fft_corr_prepared(const T* pics, const T* ref, T* res, size2 side, uint32_t ref_slices, uint32_t batch,
                              hipStream_t st) {
    if (ref_slices == 0 || batch == 0) return;
    if (side.x < 2 || side.y < 2 || side.x > kFftMaxSlice || side.y > kFftMaxSlice)
        throw unsupported_error("cross_corr_fft: slice sides must be between 2 and " + std::to_string(kFftMaxSlice));
    const bool dbl = sizeof(T) == 8;
    cplx<T>* spectra = nullptr;
    void* work = nullptr;
    if (fftg_needed(dbl, side)) {
        const size_t chunk = fftg_chunk(dbl, side, std::max(ref_slices, ref_slices * batch), 0);
        EMIDA_HIP(hipMallocAsync((void**)&spectra, fftg_spectrum_bytes(dbl, side) * ref_slices, st));
        EMIDA_HIP(hipMallocAsync(&work, chunk * fftg_slice_work_bytes(dbl, side), st));
        launch_fftg_corr_prepared<T>(pics, ref, res, side, ref_slices, batch, spectra, work, chunk, st);
    } else {
        EMIDA_HIP(hipMallocAsync((void**)&spectra, spectra_bytes<T>(side.x, ref_slices), st));
        fft_corr_prepared<T>(pics, ref, res, side.x, ref_slices, batch, spectra, st);
    }
    if (work) EMIDA_HIP(hipFreeAsync(work, st));
    EMIDA_HIP(hipFreeAsync(spectra, st));
}

// ---- allocation-free variant: working memory in a static device array
// Each GPU holds its own copy of `scratch` for the life of the process; host code
// uses the address hipGetSymbolAddress gives for the current device, never the
// array's name.  Nothing in it needs clearing: every spectrum and work-buffer
// element is written before it is read.
static __device__ __align__(256) unsigned char scratch[kFftStaticScratchBytes];

static void* static_scratch() {
    static std::mutex mu;
    static std::vector<void*> per_device;
    int dev = 0;
    EMIDA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((size_t)dev >= per_device.size()) per_device.resize(dev + 1, nullptr);
    if (!per_device[dev]) {
        void* p = nullptr;
        EMIDA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw hip_error("hipGetSymbolAddress returned no address for the FFT scratch array");
        per_device[dev] = p;
    }
    return per_device[dev];
}

template <class T>
void launch_fft_corr_prepared_noalloc(const T* pics, const T* ref, T* res, size2 side, uint32_t ref_slices,
                                      uint32_t batch, hipStream_t st) {
    if (ref_slices == 0 || batch == 0) return;
    if (side.x < 2 || side.y < 2 || side.x > kFftMaxSlice || side.y > kFftMaxSlice)
        throw unsupported_error("cross_corr_fft: slice sides must be between 2 and " + std::to_string(kFftMaxSlice));
    const bool dbl = sizeof(T) == 8, general = fftg_needed(dbl, side);
    const size_t spec_one = general ? fftg_spectrum_bytes(dbl, side) : spectra_bytes<T>(side.x, 1);
    const size_t spec = (spec_one * ref_slices + 255) & ~(size_t)255;
    if (spec > kFftStaticScratchBytes)
        throw arg_error("cross_corr_fft without allocation: the reference spectra of " + std::to_string(ref_slices) +
                        " slices need " + std::to_string(spec) + " bytes, more than the " +
                        std::to_string(kFftStaticScratchBytes) + "-byte static scratch");
    // the rest holds the work buffer of the three-kernel path's launch chunks
    const size_t left = kFftStaticScratchBytes - spec;
    char* base = nullptr;
    if (general) {
        const size_t chunk = std::min<size_t>(left / fftg_slice_work_bytes(dbl, side), (size_t)ref_slices * batch);
        if (chunk == 0)
            throw arg_error("cross_corr_fft without allocation: no room for the work buffer beside " +
                            std::to_string(ref_slices) + " reference spectra in the " +
                            std::to_string(kFftStaticScratchBytes) + "-byte static scratch");
        base = (char*)static_scratch();
        launch_fftg_corr_prepared<T>(pics, ref, res, side, ref_slices, batch, (cplx<T>*)base, base + spec, chunk, st);
        return;
    }
    base = (char*)static_scratch();
    fft_corr_prepared<T>(pics, ref, res, side.x, ref_slices, batch, (cplx<T>*)base, st);
}

#define INST_CORR_PREPARED(T)                                                                                   \
    template void launch_fft_corr_prepared<T>(const T*, const T*, T*, size2, uint32_t, uint32_t, hipStream_t);   \
    template void launch_fft_corr_prepared_noalloc<T>(const T*, const T*, T*, size2, uint32_t, uint32_t,         \
                                                      hipStream_t);
INST_CORR_PREPARED(float)
INST_CORR_PREPARED(double)

// p = p * conj(r), the reference formula (src/hadamard.cu:39-40); ref repeats
// every one.area() * ref_slices elements.
template <class T>
__global__ void hadamard_kernel(cplx<T>* __restrict__ pics, const cplx<T>* __restrict__ ref, size_t ref_len,
                                size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pics[i] = hadamard(pics[i], ref[i % ref_len]);
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

template void launch_hadamard<float>(cplx<float>*, const cplx<float>*, size2, uint32_t, uint32_t, hipStream_t);
template void launch_hadamard<double>(cplx<double>*, const cplx<double>*, size2, uint32_t, uint32_t,
                                      hipStream_t);

}  // namespace emida_amd
