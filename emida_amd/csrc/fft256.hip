// FFT cross-correlation for slices of side 65..128 (256 x 256 transforms, float) in
// two kernels.  The half spectrum of one slice (129 x 256 complex) does not fit in LDS,
// but the column stage needs one column at a time: kernel 1 keeps the 64 packed
// row-pair spectra in LDS, runs every column's forward transform, product with the
// conjugate reference spectrum and inverse transform, and writes the column results
// (129 x 256 complex per slice) to a work buffer; kernel 2 reads them back in bands
// of 32 rows, runs the packed row-pair inverse transforms, finds the arg-max over the
// 256 x 256 circular buffer, recomputes the F rows of the neighbourhood, and fits.
// A slice side s < 128 is zero-padded to 256; the arg-max, the gather and the
// optional result buffer use the reference's 2s x 2s layout, as in fft.hip (index
// k <= s is circular position k, k > s is position k + 256 - 2s).
//
// Same semantics and scaling as the single-kernel pipeline (fft.hip): reference
// src/gpu_offset.hpp:303-316, src/hadamard.cu:27-41, src/maxarg.cu,
// src/extract_neighbors.cu with cross_res_pos_policy_fft (src/kernels.cuh:36-53),
// src/subpixel_max.hpp; the reference spectra use the same per-column layout and
// the 1/4 scale.  Launches run over chunks of slices so the work buffer stays
// bounded (kFft256Chunk slices).
#include "fft_common.hpp"

namespace emida_amd {

namespace {

constexpr int kM = 256, kS = 128, kNC = kS + 1, kL = 16, kR = 16;  // kS: largest slice side
constexpr int kCS = 280;          // LDS slot stride (complex): >= L (R + 1), 8 mod 16
constexpr int kNT1 = 1024;        // kernel 1: 64 lane groups, one per row pair / two column jobs
constexpr int kNT2 = 256;         // kernel 2: 16 lane groups, one per row pair of a band
constexpr int kBand = 32;         // kernel 2: rows per band
constexpr int kBS = kBand + 1;    // padded column stride of a band in LDS
constexpr size_t kLds1 = (size_t)(kS / 2) * kCS * sizeof(cplx<float>) + kM * sizeof(cplx<float>);
constexpr size_t kLds2 = ((size_t)kNC * kBS + (kNT2 / kL) * kCS + kM) * sizeof(cplx<float>);

__device__ __forceinline__ void init_twiddles(cplx<float>* tw, int tid, int nt) {
    for (int t = tid; t < kM; t += nt) {
        double s, c;
        sincospi(2.0 * t / kM, &s, &c);
        tw[t] = cplx<float>{(float)c, (float)(-s)};
    }
}

}  // namespace

// ---------------------------------------------------------------- kernel 1
template <class IN, class SUM, bool FROM_PATTERN, bool PIPE>
__global__ __launch_bounds__(kNT1, 1) void fft256_cols_kernel(fft_args<float, IN, SUM> a) {
    using C = cplx<float>;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C* B = reinterpret_cast<C*>(lds_raw);  // 64 slots of kCS
    C* tw = B + (kS / 2) * kCS;
    __shared__ uint32_t red_u[kNT1 / kWave];

    if (blockIdx.x >= a.total) return;
    const uint32_t sl = xcd_slice(blockIdx.x, a.total, a.nb);  // slice within the launch
    const uint32_t sn = a.first + sl;
    const uint32_t bn = sn % a.nb;
    const int tid = threadIdx.x, lane = tid % kL, grp = tid / kL;
    const int s = (int)a.side;  // slice side, s <= kS
    // stored reference spectra: x 1/4 x (2s/256)^2 (the reference's (2s)^2 scale)
    const float sc = (float)(0.25 * ((double)(2 * s) / kM) * ((double)(2 * s) / kM));
    init_twiddles(tw, tid, kNT1);

    // ---- 0. mean of the ROI (FROM_PATTERN), read again by the row stage
    float mean = 0.f;
    const IN* pic = nullptr;
    const float* pre = nullptr;
    if constexpr (FROM_PATTERN) {
        const size2 b0 = a.begins[bn];
        pic = a.pics + (size_t)(sn / a.nb) * a.src.area() + (size_t)b0.y * a.src.x + b0.x;
        if constexpr (std::is_same<IN, uint16_t>::value) {
            uint32_t su = 0;
            for (int e = tid; e < s * s; e += kNT1) su += pic[(size_t)(e / s) * a.src.x + (e % s)];
#pragma unroll
            for (int off = kWave / 2; off > 0; off >>= 1) su += __shfl_xor(su, off);
            if (tid % kWave == 0) red_u[tid / kWave] = su;
            __syncthreads();
            su = 0;
#pragma unroll
            for (int k = 0; k < kNT1 / kWave; ++k) su += red_u[k];
            mean = mean_of<float>((float)su, s * s);
        } else {
            mean = mean_of<float>((float)a.sums[sn], s * s);
        }
    } else {
        pre = a.prepared + (size_t)sn * s * s;
    }
    __syncthreads();

    // ---- A. packed row pairs (2g, 2g + 1), forward (an odd s leaves the last pair's
    // second row zero)
    if (grp < (s + 1) / 2) {
        const int g = grp;
        const bool second = 2 * g + 1 < s;
        C x[kR];
#pragma unroll
        for (int n1 = 0; n1 < kR; ++n1) {
            const int n = kL * n1 + lane;
            if (n1 < kR / 2 && n < s) {
                float v0, v1 = 0.f;
                if constexpr (FROM_PATTERN) {
                    v0 = prep_px<float, IN>(pic[(size_t)(2 * g) * a.src.x + n], mean, a.hx[n], a.hy[2 * g]);
                    if (second)
                        v1 = prep_px<float, IN>(pic[(size_t)(2 * g + 1) * a.src.x + n], mean, a.hx[n],
                                                a.hy[2 * g + 1]);
                } else {
                    v0 = pre[(2 * g) * s + n];
                    if (second) v1 = pre[(2 * g + 1) * s + n];
                }
                x[n1] = C{v0, v1};
            } else {
                x[n1] = C{0.f, 0.f};
            }
        }
        C* slot = B + g * kCS;
        fft_fwd<float, kR, kL, kM, true>(x, lane, slot, tw);
#pragma unroll
        for (int k2 = 0; k2 < kL; ++k2) slot[lane + kR * k2] = x[k2];
    }
    __syncthreads();

    // ---- B. column jobs c = g and c = g + 64 (c = 0: packed columns 0 and S)
    C xj[2][kR];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int c = grp + 64 * j;
#pragma unroll
        for (int n1 = 0; n1 < kR; ++n1) {
            const int y = kL * n1 + lane;
            if (n1 < kR / 2 && y < s) {
                const C* row = B + (y >> 1) * kCS;
                if (c == 0) {
                    const C z0 = row[0], zs = row[kS];
                    xj[j][n1] = (y & 1) ? C{z0.im, zs.im} : C{z0.re, zs.re};
                } else {
                    const C zk = row[c], zm = row[kM - c];
                    xj[j][n1] = (y & 1) ? C{zk.im + zm.im, zm.re - zk.re} : C{zk.re + zm.re, zk.im - zm.im};
                }
            } else {
                xj[j][n1] = C{0.f, 0.f};
            }
        }
    }
    __syncthreads();  // row spectra consumed; slots become per-group scratch
    C* slot = B + grp * kCS;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int c = grp + 64 * j;
        C x[kR];
#pragma unroll
        for (int r = 0; r < kR; ++r) x[r] = xj[j][r];
        fft_fwd<float, kR, kL, kM, true>(x, lane, slot, tw);
        if (c == 0) {
            // split the packed real columns: P0[k] = (Z[k] + conj(Z[-k]))/2,
            // PS[k] = (Z[k] - conj(Z[-k]))/(2i) (factors 1/2 omitted, as in fft.hip)
#pragma unroll
            for (int k2 = 0; k2 < kL; ++k2) slot[lane + kR * k2] = x[k2];
            wave_sync();
            C p0[kR], ps[kR];
#pragma unroll
            for (int k2 = 0; k2 < kL; ++k2) {
                const int k = lane + kR * k2;
                const C z = x[k2], zm = slot[(kM - k) & (kM - 1)];
                p0[k2] = C{z.re + zm.re, z.im - zm.im};
                ps[k2] = C{z.im + zm.im, zm.re - z.re};
            }
            wave_sync();
            if constexpr (PIPE) {
                const C* r0 = a.spectra + ((size_t)bn * kNC + 0) * kM + lane * kR;
                const C* rs = a.spectra + ((size_t)bn * kNC + kS) * kM + lane * kR;
#pragma unroll
                for (int r = 0; r < kR; ++r) {
                    const C h0 = hadamard(p0[r], r0[r]), hs = hadamard(ps[r], rs[r]);
                    x[r] = C{h0.re - hs.im, h0.im + hs.re};  // H0 + i HS
                }
            } else {
                C* o0 = a.spectra_out + ((size_t)bn * kNC + 0) * kM + lane * kR;
                C* os = a.spectra_out + ((size_t)bn * kNC + kS) * kM + lane * kR;
#pragma unroll
                for (int r = 0; r < kR; ++r) {
                    o0[r] = C{p0[r].re * sc, p0[r].im * sc};
                    os[r] = C{ps[r].re * sc, ps[r].im * sc};
                }
            }
        } else {
            if constexpr (PIPE) {
                const C* rc = a.spectra + ((size_t)bn * kNC + c) * kM + lane * kR;
#pragma unroll
                for (int r = 0; r < kR; ++r) x[r] = hadamard(x[r], rc[r]);
            } else {
                C* oc = a.spectra_out + ((size_t)bn * kNC + c) * kM + lane * kR;
#pragma unroll
                for (int r = 0; r < kR; ++r) oc[r] = C{x[r].re * sc, x[r].im * sc};
            }
        }
        if constexpr (PIPE) {
            fft_inv<float, kR, kL, kM>(x, lane, slot, tw);
            C* g = a.work + (size_t)sl * kNC * kM;
            if (c == 0) {
#pragma unroll
                for (int n1 = 0; n1 < kR; ++n1) {
                    const int y = kL * n1 + lane;
                    g[y] = C{x[n1].re, 0.f};
                    g[(size_t)kS * kM + y] = C{x[n1].im, 0.f};
                }
            } else {
#pragma unroll
                for (int n1 = 0; n1 < kR; ++n1) g[(size_t)c * kM + kL * n1 + lane] = x[n1];
            }
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------- kernel 2
// Packed inverse of output rows (2p, 2p + 1): frequency k of the row pair is
// Ga[k] + i Gb[k] with G[k] = col[k] for k <= S and conj(col[M - k]) above (DC and
// Nyquist imaginary parts dropped, as a C2R does).  Column c's values at the two rows
// are src[c * stride] and src[c * stride + 1].  On return lane n2 holds row 2p in
// x[n1].re and row 2p + 1 in x[n1].im at column L n1 + n2.
__device__ __forceinline__ void row_pair_inverse(cplx<float> (&x)[kR], int lane, cplx<float>* scr,
                                                 const cplx<float>* tw, const cplx<float>* src, size_t stride) {
    using C = cplx<float>;
#pragma unroll
    for (int k2 = 0; k2 < kL; ++k2) {
        const int k = lane + kR * k2;
        const bool mirror = k > kS;
        const int col = mirror ? kM - k : k;
        C ga = src[col * stride], gb = src[col * stride + 1];
        if (mirror) {
            ga.im = -ga.im;
            gb.im = -gb.im;
        }
        if (k == 0 || k == kS) {
            ga.im = 0.f;
            gb.im = 0.f;
        }
        x[k2] = C{ga.re - gb.im, ga.im + gb.re};
    }
    fft_inv<float, kR, kL, kM>(x, lane, scr, tw);
}

template <class IN, class SUM>
__global__ __launch_bounds__(kNT2, 2) void fft256_rows_kernel(fft_args<float, IN, SUM> a) {
    using C = cplx<float>;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C* Gb = reinterpret_cast<C*>(lds_raw);  // band: [kNC][kBS]
    C* scr = Gb + (size_t)kNC * kBS;        // per group kCS
    C* tw = scr + (kNT2 / kL) * kCS;
    float* rb = reinterpret_cast<float*>(Gb);  // neighbourhood rows after the band loop: [9][kM]
    __shared__ float sv[kNT2 / kWave];
    __shared__ uint32_t si[kNT2 / kWave];
    __shared__ float nbv[81];
    __shared__ double cf[6];

    if (blockIdx.x >= a.total) return;
    const uint32_t sl = xcd_slice(blockIdx.x, a.total, a.nb);
    const uint32_t sn = a.first + sl;
    const int tid = threadIdx.x, lane = tid % kL, grp = tid / kL;
    const C* g = a.work + (size_t)sl * kNC * kM;
    C* myscr = scr + grp * kCS;
    init_twiddles(tw, tid, kNT2);
    const int s = (int)a.side, pad = kM - 2 * s;
    const uint32_t s2 = 2 * s;
    float* cout = a.corr_out ? a.corr_out + (size_t)sn * s2 * s2 : nullptr;

    float vbest = -(float)__builtin_huge_val();
    uint32_t ibest = 0xffffffffu;
    C x[kR];  // one array for both inverse stages (keeps it in registers)
#pragma unroll 1
    for (int y0 = 0; y0 < kM; y0 += kBand) {
        if (y0 > s && y0 + kBand - 1 <= s + pad) continue;  // no row of the reference layout
        __syncthreads();  // previous band consumed
        for (int e = tid; e < kNC * kBand; e += kNT2) {
            const int c = e / kBand, r = e % kBand;
            Gb[c * kBS + r] = g[(size_t)c * kM + y0 + r];
        }
        __syncthreads();
        const int p = grp;  // row pair within the band
        row_pair_inverse(x, lane, myscr, tw, Gb + 2 * p, kBS);
        const int ya = y0 + 2 * p, yb = ya + 1;
        const bool in_a = ya <= s || ya > s + pad, in_b = yb <= s || yb > s + pad;
        const uint32_t ra = (uint32_t)(ya <= s ? ya : ya - pad) * s2, rb_ = (uint32_t)(yb <= s ? yb : yb - pad) * s2;
#pragma unroll
        for (int n1 = 0; n1 < kR; ++n1) {
            const int col = kL * n1 + lane;
            if (col > s && col <= s + pad) continue;
            const uint32_t rx = (uint32_t)(col <= s ? col : col - pad);
            // strictly larger wins; equal values keep the lower index (the scan order
            // per thread is increasing in neither, so compare indices)
            if (in_a && (x[n1].re > vbest || (x[n1].re == vbest && ra + rx < ibest))) {
                vbest = x[n1].re;
                ibest = ra + rx;
            }
            if (in_b && (x[n1].im > vbest || (x[n1].im == vbest && rb_ + rx < ibest))) {
                vbest = x[n1].im;
                ibest = rb_ + rx;
            }
            if (cout) {
                if (in_a) cout[ra + rx] = x[n1].re;
                if (in_b) cout[rb_ + rx] = x[n1].im;
            }
        }
    }
    // arg-max: the largest value (NaN ignored), then the lowest index holding it
    const float best = block_max<float, kNT2>(vbest, sv);
    uint32_t bi = (vbest == best) ? ibest : 0xffffffffu;
    bi = block_min_u32<kNT2>(bi, si);

    // ---- neighbourhood: recompute the F rows it touches (one lane group per row)
    const int fit = a.fit, kk = fit * fit, r = (fit - 1) / 2;
    const size2 cross{s2 - 1, s2 - 1};
    const size2 mp = shift_pos(POLICY_FFT, bi == 0xffffffffu ? 0u : bi, cross);
    __syncthreads();  // band buffer free: rb reuses it
    if (grp < fit) {
        const size2 cc = clamp_pos((int)mp.x, (int)mp.y - r + grp, cross);
        int qy = (int)((cc.y + s + 1) % s2);
        if (qy > s) qy += pad;
        const int p2 = qy >> 1;
        row_pair_inverse(x, lane, myscr, tw, g + 2 * p2, kM);
        // branch outside the loop: a select between the two members would index the
        // register array dynamically and move it to scratch memory
        if (qy & 1) {
#pragma unroll
            for (int n1 = 0; n1 < kR; ++n1) rb[grp * kM + kL * n1 + lane] = x[n1].im;
        } else {
#pragma unroll
            for (int n1 = 0; n1 < kR; ++n1) rb[grp * kM + kL * n1 + lane] = x[n1].re;
        }
    }
    __syncthreads();
    for (int t = tid; t < kk; t += kNT2) {
        const int i = t % fit, j = t / fit;
        const size2 cc = clamp_pos((int)mp.x - r + i, (int)mp.y - r + j, cross);
        int qx = (int)((cc.x + s + 1) % s2);
        if (qx > s) qx += pad;
        const float v = rb[j * kM + qx];
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

// ---------------------------------------------------------------- launchers
size_t fft256_chunk(uint32_t nb) {
    const size_t per = kFft256Chunk / nb;
    return (per < 1 ? 1 : per) * (size_t)nb;
}

size_t fft256_slice_work_bytes() { return (size_t)kNC * kM * sizeof(cplx<float>); }

size_t fft256_work_bytes(uint32_t nb, uint32_t total) {
    const size_t chunk = fft256_chunk(nb);
    return (total < chunk ? total : chunk) * (size_t)kNC * kM * sizeof(cplx<float>);
}

template <class IN, class SUM, bool FP, bool PIPE>
void launch_fft256(fft_args<float, IN, SUM> a, hipStream_t st) {
    auto k1 = fft256_cols_kernel<IN, SUM, FP, PIPE>;
    static const hipError_t attr1 =
        hipFuncSetAttribute((const void*)k1, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLds1);
    EMIDA_HIP(attr1);
    if constexpr (!PIPE) {
        a.first = 0;
        k1<<<a.total, kNT1, kLds1, st>>>(a);
        EMIDA_LAUNCHED("fft256_cols_kernel");
    } else {
        if (!a.work) throw arg_error("FFT 256: work buffer missing");
        auto k2 = fft256_rows_kernel<IN, SUM>;
        static const hipError_t attr2 =
            hipFuncSetAttribute((const void*)k2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLds2);
        EMIDA_HIP(attr2);
        const uint32_t total = a.total, chunk = a.chunk ? a.chunk : (uint32_t)fft256_chunk(a.nb);
        if (chunk % a.nb != 0) throw arg_error("FFT 256: launch chunk must be a multiple of the ROI count");
        for (uint32_t s0 = 0; s0 < total; s0 += chunk) {
            fft_args<float, IN, SUM> b = a;
            b.first = s0;
            b.total = std::min(chunk, total - s0);
            k1<<<b.total, kNT1, kLds1, st>>>(b);
            EMIDA_LAUNCHED("fft256_cols_kernel");
            k2<<<b.total, kNT2, kLds2, st>>>(b);
            EMIDA_LAUNCHED("fft256_rows_kernel");
        }
    }
}

#define INST_FFT256(IN, SUM)                                                                        \
    template void launch_fft256<IN, SUM, true, false>(fft_args<float, IN, SUM>, hipStream_t);      \
    template void launch_fft256<IN, SUM, true, true>(fft_args<float, IN, SUM>, hipStream_t);
INST_FFT256(uint16_t, uint32_t)
INST_FFT256(float, float)
INST_FFT256(double, double)
template void launch_fft256<float, float, false, false>(fft_args<float, float, float>, hipStream_t);
template void launch_fft256<float, float, false, true>(fft_args<float, float, float>, hipStream_t);

}  // namespace emida_amd
