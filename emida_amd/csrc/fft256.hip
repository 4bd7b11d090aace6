// FFT cross-correlation for float slices of side 65..128 (256 x 256 transforms) in one
// kernel per slice, with no intermediate in HBM.  The column results of one slice (129 x
// 256 complex, 264 KB) fit neither LDS nor, with the transforms' working registers, the
// register file alone: 512 threads in 32 lane groups of 16 keep three half-spectrum
// columns each in registers (48 complex values per lane) and the last 32 columns in LDS,
// and move them through LDS in bands of 64 rows for the inverse row transforms.
//  A. packed row pairs (2p, 2p + 1) -> forward 256-point transforms into LDS slots;
//  B. column jobs c = g + 32 j (j < 4; c = 0 packs the real columns 0 and 128): every
//     group reads its four columns out of the row spectra into registers, then per
//     column the forward transform, the product with the conjugate reference spectrum
//     and the inverse transform; jobs 0-2 stay in registers, job 3 goes to LDS;
//  C. four bands of 64 output rows: the band's rows of every column are gathered in LDS,
//     each group runs the packed inverse of one row pair (its transpose scratch over the
//     band once every group has read its inputs), arg-max over the reference layout
//     (lowest index on ties);
//  D. the F rows of the neighbourhood are transformed again (only the bands holding
//     them are gathered again), gather through the FFT position policy, fit and
//     finalize.
// A slice side s < 128 is zero-padded to 256; the arg-max, the gather and the optional
// result buffer use the reference's 2s x 2s layout, as in fft.hip (index k <= s is
// circular position k, k > s is position k + 256 - 2s).
//
// Same semantics and scaling as the single-kernel pipeline (fft.hip): reference
// src/gpu_offset.hpp:303-316, src/hadamard.cu:27-41, src/maxarg.cu,
// src/extract_neighbors.cu with cross_res_pos_policy_fft (src/kernels.cuh:36-53),
// src/subpixel_max.hpp; the reference spectra use the same per-column layout and the
// 1/4 scale.  For the reference (PIPE = false) the kernel stops after the forward
// column transforms and stores the scaled spectra.
#include "fft_common.hpp"

namespace emida_amd {

namespace {

constexpr int kM = 256, kS = 128, kNC = kS + 1, kL = 16, kR = 16;  // kS: largest slice side
constexpr int kNT = 512;          // threads: 32 lane groups of 16
constexpr int kG = kNT / kL;      // lane groups
constexpr int kJ = 4;             // column jobs per group (c = g + kG j covers 0..127)
constexpr int kJR = 3;            // jobs whose results stay in registers; job 3 goes to LDS
constexpr int kCS = 280;          // LDS slot stride (complex): >= L (R + 1), 8 mod 16
constexpr int kBand = 64;         // rows per band of the inverse row stage (kG row pairs)
constexpr int kBS = kBand + 1;    // padded column stride of a band in LDS
// LDS (complex elements): phase A: 64 slots of kCS (row spectra); phase B: group scratch
// in slots 0..31, job-3 columns from kCol; phases C-D: band [kNC][kBS] at 0 with the group
// scratch over it (kG slots of kCS), job-3 columns still from kCol
constexpr int kSlots = (kS / 2) * kCS;
constexpr int kLow = kG * kCS;         // band / scratch region
constexpr int kColStride = kM + 8;     // job-3 column stride
constexpr int kCol = kLow;
static_assert(kNC * kBS <= kLow, "band exceeds its region");
static_assert(kCol + kG * kColStride <= kSlots, "job-3 columns exceed the slots");
static_assert(6 * 81 <= kNT, "one least-squares element per thread");
static_assert(4 * 81 <= kNT, "four threads per neighbourhood value");
constexpr int kTw2 = kL * (kR + 1);  // the transforms' twiddles [n][k], rows of kR + 1
constexpr size_t kLds = (size_t)(kSlots + kM + kTw2) * sizeof(cplx<float>);

// Packed inverse of output rows (2p, 2p + 1): frequency k of the row pair is
// Ga[k] + i Gb[k] with G[k] = col[k] for k <= S and conj(col[M - k]) above (DC and
// Nyquist imaginary parts dropped, as a C2R does).  Column c's values at the two rows
// are src[c * stride] and src[c * stride + 1].  On return lane n2 holds row 2p in
// x[n1].re and row 2p + 1 in x[n1].im at column L n1 + n2.
__device__ __forceinline__ void row_pair_load(cplx<float> (&x)[kR], int lane, const cplx<float>* src, int stride) {
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
}

// Rows kBand b .. kBand b + 63 of every column into the band buffer (column c at c * kBS;
// the packed job 0 of group 0 gives columns 0 and S): jobs 0-2 from registers, job 3
// (columns 96..127) copied from its LDS store.
template <int B>
__device__ __forceinline__ void write_band(cplx<float>* band, const cplx<float> (&xs)[kJR][kR],
                                           const cplx<float>* cols, int g, int lane, int tid) {
    using C = cplx<float>;
    for (int e = tid; e < kG * kBand; e += kNT) {
        const int c = e / kBand, r = e % kBand;
        band[(kG * kJR + c) * kBS + r] = cols[c * kColStride + kBand * B + r];
    }
#pragma unroll
    for (int j = 0; j < kJR; ++j) {
        const int c = g + kG * j;
#pragma unroll
        for (int i = 0; i < kBand / kL; ++i) {
            const int n1 = (kBand / kL) * B + i, r = kL * i + lane;
            const C v = xs[j][n1];
            if (c == 0) {
                band[r] = C{v.re, 0.f};
                band[kS * kBS + r] = C{v.im, 0.f};
            } else {
                band[c * kBS + r] = v;
            }
        }
    }
}

// Row kL Q + lane of the register-held column results into a half-spectrum row gr
// (the packed job 0 of group 0 gives columns 0 and S).  Called from a switch on the
// row, so that every register index is static.
template <int Q>
__device__ __forceinline__ void grow_row(cplx<float>* gr, const cplx<float> (&xs)[kJR][kR], int g) {
    if (g == 0) {
        gr[0] = cplx<float>{xs[0][Q].re, 0.f};
        gr[kS] = cplx<float>{xs[0][Q].im, 0.f};
    } else {
        gr[g] = xs[0][Q];
    }
#pragma unroll
    for (int jj = 1; jj < kJR; ++jj) gr[g + kG * jj] = xs[jj][Q];
}

}  // namespace

// FULL: the slice side is kS (a compile-time constant, every bound below folds)
template <class IN, class SUM, bool FROM_PATTERN, bool PIPE, bool FULL>
__global__ __launch_bounds__(kNT, 1) void fft256_kernel(fft_args<float, IN, SUM> a) {
    using C = cplx<float>;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C* B = reinterpret_cast<C*>(lds_raw);  // 64 slots of kCS
    C* tw = B + kSlots;    // W_M^m, m < M (the neighbourhood's direct sums)
    C* tw2 = tw + kM;      // the same values [n][k] for the transforms
    __shared__ uint32_t red_u[kNT / kWave];
    __shared__ float sv[kNT / kWave];
    __shared__ uint32_t si[kNT / kWave];
    __shared__ float nbv[81];
    __shared__ double lsq_s[6 * 81];
    __shared__ double cf[6];

    if (blockIdx.x >= a.total) return;
    const uint32_t sn = xcd_slice(blockIdx.x, a.total, a.nb);
    const uint32_t bn = sn % a.nb;
    const int tid = threadIdx.x, lane = tid % kL, g = tid / kL;
    const int s = FULL ? kS : (int)a.side;  // slice side, s <= kS
    // stored reference spectra: x 1/4 x (2s/256)^2 (the reference's (2s)^2 scale)
    const float sc = (float)(0.25 * ((double)(2 * s) / kM) * ((double)(2 * s) / kM));

    // ---- 0. the slice's inputs: group g's row pairs g and g + 32, 16 values of each
    // row per lane, all loaded before the first use (addresses clamped into the slice,
    // values outside it masked), so the loads are in flight together
    const int npairs = (s + 1) / 2;
    using V = typename std::conditional<FROM_PATTERN, IN, float>::type;
    const V* base;
    size_t stride;
    if constexpr (FROM_PATTERN) {
        const size2 b0 = a.begins[bn];
        base = a.pics + (size_t)(sn / a.nb) * a.src.area() + (size_t)b0.y * a.src.x + b0.x;
        stride = a.src.x;
    } else {
        base = a.prepared + (size_t)sn * s * s;
        stride = (size_t)s;
    }
    V raw[2][kR / 2][2];
    float wx[kR / 2], wy[2][2];
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const int p = min(g + kG * it, npairs - 1);
#pragma unroll
        for (int n1 = 0; n1 < kR / 2; ++n1) {
            const int n = min(kL * n1 + lane, s - 1);
#pragma unroll
            for (int h = 0; h < 2; ++h) raw[it][n1][h] = base[(size_t)min(2 * p + h, s - 1) * stride + n];
        }
        if constexpr (FROM_PATTERN) {
            wy[it][0] = a.hy[2 * p];
            wy[it][1] = a.hy[min(2 * p + 1, s - 1)];
        }
    }
    if constexpr (FROM_PATTERN) {
#pragma unroll
        for (int n1 = 0; n1 < kR / 2; ++n1) wx[n1] = a.hx[min(kL * n1 + lane, s - 1)];
    }
    // the fit's least-squares rows (6 x F^2 <= 486 doubles), staged in LDS for the end
    const int kk = PIPE ? a.fit * a.fit : 0;
    const double lsq_v = tid < 6 * kk ? a.lsq[tid] : 0.0;
    load_twiddles<float, kM>(tw, a.twt, tid, kNT);  // while the loads are in flight
    load_twiddles_2d<float, kM, kR>(tw2, a.twt, tid, kNT);
    // mean of the ROI (FROM_PATTERN): every pixel is one (pair, row, column) of one lane
    float mean = 0.f;
    if constexpr (FROM_PATTERN) {
        if constexpr (std::is_same<IN, uint16_t>::value) {
            uint32_t su = 0;
#pragma unroll
            for (int it = 0; it < 2; ++it)
#pragma unroll
                for (int n1 = 0; n1 < kR / 2; ++n1)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int p = g + kG * it;
                        const bool in = p < npairs && 2 * p + h < s && kL * n1 + lane < s;
                        su += in ? (uint32_t)raw[it][n1][h] : 0u;
                    }
            su = wave_reduce(su, 0u, [](uint32_t x, uint32_t y) { return x + y; });
            if (tid % kWave == 0) red_u[tid / kWave] = su;
            __syncthreads();
            su = 0;
#pragma unroll
            for (int k = 0; k < kNT / kWave; ++k) su += red_u[k];
            mean = mean_of<float>((float)su, s * s);
        } else {
            mean = mean_of<float>((float)a.sums[sn], s * s);
        }
    }
    if (tid < 6 * kk) lsq_s[tid] = lsq_v;
    __syncthreads();  // twiddles and least-squares rows written

    // ---- A. packed row pairs (2p, 2p + 1), forward (an odd s leaves the last pair's
    // second row zero); group g takes pairs g and g + 32, slot p is its own scratch
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const int p = g + kG * it;
        if (p < npairs) {
            const bool second = 2 * p + 1 < s;
            C x[kR];
#pragma unroll
            for (int n1 = 0; n1 < kR; ++n1) {
                const int n = kL * n1 + lane;
                if (n1 < kR / 2 && n < s) {
                    float v0, v1;
                    if constexpr (FROM_PATTERN) {
                        v0 = prep_px<float, IN>(raw[it][n1][0], mean, wx[n1], wy[it][0]);
                        v1 = prep_px<float, IN>(raw[it][n1][1], mean, wx[n1], wy[it][1]);
                    } else {
                        v0 = raw[it][n1][0];
                        v1 = raw[it][n1][1];
                    }
                    x[n1] = C{v0, second ? v1 : 0.f};
                } else {
                    x[n1] = C{0.f, 0.f};
                }
            }
            C* slot = B + p * kCS;
            fft_fwd<float, kR, kL, kM, true, true>(x, lane, slot, tw2);
#pragma unroll
            for (int k2 = 0; k2 < kL; ++k2) slot[lane + kR * k2] = x[k2];
        }
    }
    __syncthreads();

    // ---- B. column jobs c = g + 32 j (c = 0: packed columns 0 and S), inputs split from
    // the row spectra into registers
    C xs[kJ][kR];  // inputs of the four jobs; results of jobs 0-2 (job 3's go to LDS)
#pragma unroll
    for (int j = 0; j < kJ; ++j) {
        const int c = g + kG * j;
#pragma unroll
        for (int n1 = 0; n1 < kR; ++n1) {
            const int y = kL * n1 + lane;
            if (n1 < kR / 2 && y < s) {
                const C* row = B + (y >> 1) * kCS;
                if (c == 0) {
                    const C z0 = row[0], zs = row[kS];
                    xs[j][n1] = (y & 1) ? C{z0.im, zs.im} : C{z0.re, zs.re};
                } else {
                    const C zk = row[c], zm = row[kM - c];
                    xs[j][n1] = (y & 1) ? C{zk.im + zm.im, zm.re - zk.re} : C{zk.re + zm.re, zk.im - zm.im};
                }
            } else {
                xs[j][n1] = C{0.f, 0.f};
            }
        }
    }
    __syncthreads();  // row spectra consumed; slot g becomes group g's scratch
    C* slot = B + g * kCS;
#pragma unroll
    for (int j = 0; j < kJ; ++j) {
        const int c = g + kG * j;
        // this column's reference spectrum, in flight during the forward transform
        C spec[kR];
        if constexpr (PIPE) {
            if (c != 0) {
                const C* rc = a.spectra + ((size_t)bn * kNC + c) * kM + lane * kR;
#pragma unroll
                for (int r = 0; r < kR; ++r) spec[r] = rc[r];
            }
        }
        // the twiddles and their addresses are reloaded per transform rather than kept
        // live across the column jobs (registers hold the four columns' results)
        asm volatile("" ::: "memory");
        fft_fwd<float, kR, kL, kM, true, true>(xs[j], lane, slot, tw2);
        if (c == 0) {
            // split the packed real columns: P0[k] = (Z[k] + conj(Z[-k]))/2,
            // PS[k] = (Z[k] - conj(Z[-k]))/(2i) (factors 1/2 omitted, as in fft.hip),
            // one frequency at a time, in place (no second register array)
#pragma unroll
            for (int k2 = 0; k2 < kL; ++k2) slot[lane + kR * k2] = xs[j][k2];
            wave_sync();
            const C* r0 = PIPE ? a.spectra + ((size_t)bn * kNC + 0) * kM + lane * kR : nullptr;
            const C* rs = PIPE ? a.spectra + ((size_t)bn * kNC + kS) * kM + lane * kR : nullptr;
            C* o0 = PIPE ? nullptr : a.spectra_out + ((size_t)bn * kNC + 0) * kM + lane * kR;
            C* os = PIPE ? nullptr : a.spectra_out + ((size_t)bn * kNC + kS) * kM + lane * kR;
#pragma unroll
            for (int k2 = 0; k2 < kL; ++k2) {
                const int k = lane + kR * k2;
                const C z = xs[j][k2], zm = slot[(kM - k) & (kM - 1)];
                const C p0 = C{z.re + zm.re, z.im - zm.im}, ps = C{z.im + zm.im, zm.re - z.re};
                if constexpr (PIPE) {
                    const C h0 = hadamard(p0, r0[k2]), hs = hadamard(ps, rs[k2]);
                    xs[j][k2] = C{h0.re - hs.im, h0.im + hs.re};  // H0 + i HS
                } else {
                    o0[k2] = C{p0.re * sc, p0.im * sc};
                    os[k2] = C{ps.re * sc, ps.im * sc};
                }
            }
            wave_sync();
        } else {
            if constexpr (PIPE) {
#pragma unroll
                for (int r = 0; r < kR; ++r) xs[j][r] = hadamard(xs[j][r], spec[r]);
            } else {
                C* oc = a.spectra_out + ((size_t)bn * kNC + c) * kM + lane * kR;
#pragma unroll
                for (int r = 0; r < kR; ++r) oc[r] = C{xs[j][r].re * sc, xs[j][r].im * sc};
            }
        }
        if constexpr (PIPE) {
            asm volatile("" ::: "memory");
            fft_inv<float, kR, kL, kM, true, true>(xs[j], lane, slot, tw2);
            if (j == kJ - 1) {
                // job 3: column c = 96 + g, rows kL n1 + lane, to the LDS column store
                C* col = B + kCol + g * kColStride;
#pragma unroll
                for (int n1 = 0; n1 < kR; ++n1) col[kL * n1 + lane] = xs[j][n1];
            }
        }
        wave_sync();
    }
    if constexpr (!PIPE) {
        return;
    } else {
        // ---- C. bands of 64 output rows: arg-max over the reference layout
        const int pad = kM - 2 * s;
        const uint32_t s2 = 2 * s;
        C* band = B;
        C* scr = B + g * kCS;  // over the band, once every group has read its inputs
        const C* cols = B + kCol;
        C (&xr)[kJR][kR] = *reinterpret_cast<C(*)[kJR][kR]>(&xs[0][0]);
        float* cout = a.corr_out ? a.corr_out + (size_t)sn * s2 * s2 : nullptr;
        float vbest = -(float)__builtin_huge_val();
        uint32_t ibest = 0xffffffffu;
        // columns kL n1 + lane of the reference layout (k <= s or k > s + pad) and their
        // position there
        bool ok_col[kR];
#pragma unroll
        for (int n1 = 0; n1 < kR; ++n1) ok_col[n1] = kL * n1 + lane <= s || kL * n1 + lane > s + pad;
        auto layout_col = [&](int k) { return (uint32_t)(k <= s ? k : k - pad); };
        // one band at a time (a loop, not unrolled: the bands' code shares registers);
        // the register rows a band needs are picked at compile time by the switch
        auto gather_band = [&](int b) {
            switch (b) {
                case 0: write_band<0>(band, xr, cols, g, lane, tid); break;
                case 1: write_band<1>(band, xr, cols, g, lane, tid); break;
                case 2: write_band<2>(band, xr, cols, g, lane, tid); break;
                default: write_band<3>(band, xr, cols, g, lane, tid); break;
            }
        };
#pragma unroll 1
        for (int b = 0; b < kM / kBand; ++b) {
            const int y0 = kBand * b;
            if (y0 > s && y0 + kBand - 1 <= s + pad) continue;  // no row of the reference layout
            __syncthreads();  // previous band (or the column stage's scratch) consumed
            gather_band(b);
            __syncthreads();
            C x[kR];
            row_pair_load(x, lane, band + 2 * g, kBS);
            __syncthreads();  // the band is read: its space becomes transpose scratch
            fft_inv<float, kR, kL, kM, true, true>(x, lane, scr, tw2);
            const int ya = y0 + 2 * g, yb = ya + 1;
            const bool in_a = ya <= s || ya > s + pad, in_b = yb <= s || yb > s + pad;
            const uint32_t ra = (uint32_t)(ya <= s ? ya : ya - pad) * s2;
            const uint32_t rb = (uint32_t)(yb <= s ? yb : yb - pad) * s2;
            // each row's largest in-layout value and its first column (strictly larger
            // wins in ascending column order), then folded in index order: row a, row b,
            // after every earlier band, so equal values keep the lower index
            float va = -(float)__builtin_huge_val(), vb = va;
            int ca = 0, cb = 0;
            if (pad == 0) {
#pragma unroll
                for (int n1 = 0; n1 < kR; ++n1) {
                    if (x[n1].re > va) va = x[n1].re, ca = n1;
                    if (x[n1].im > vb) vb = x[n1].im, cb = n1;
                }
            } else {
#pragma unroll
                for (int n1 = 0; n1 < kR; ++n1) {
                    const float ninf = -(float)__builtin_huge_val();
                    const float ua = ok_col[n1] ? x[n1].re : ninf, ub = ok_col[n1] ? x[n1].im : ninf;
                    if (ua > va) va = ua, ca = n1;
                    if (ub > vb) vb = ub, cb = n1;
                }
            }
            if (in_a && va > vbest) {
                vbest = va;
                ibest = ra + layout_col(kL * ca + lane);
            }
            if (in_b && vb > vbest) {
                vbest = vb;
                ibest = rb + layout_col(kL * cb + lane);
            }
            if (cout) {
#pragma unroll
                for (int n1 = 0; n1 < kR; ++n1) {
                    if (!ok_col[n1]) continue;
                    const uint32_t rx = layout_col(kL * n1 + lane);
                    if (in_a) cout[ra + rx] = x[n1].re;
                    if (in_b) cout[rb + rx] = x[n1].im;
                }
            }
        }
        // the largest value (NaN ignored), then the lowest index holding it
        const float best = block_max<float, kNT>(vbest, sv);
        uint32_t bi = (vbest == best) ? ibest : 0xffffffffu;
        bi = block_min_u32<kNT>(bi, si);

        // ---- D. neighbourhood values as direct sums over the half spectra of the F
        // circular rows it touches: row y of the C2R is
        //   out[x] = G[0] + (-1)^x G[S] + 2 sum_{k=1}^{S-1} Re(G[k] e^{2 pi i k x / M})
        // (G[k]: column k's value at row y; G[0], G[S] real), the same sum the row
        // transforms of the arg-max stage evaluate, rounded in another order
        const int fit = a.fit, r = (fit - 1) / 2;
        const size2 cross{s2 - 1, s2 - 1};
        const size2 mp = shift_pos(POLICY_FFT, bi == 0xffffffffu ? 0u : bi, cross);
        auto row_of = [&](int j) {  // circular row of neighbourhood row j
            const size2 cc = clamp_pos((int)mp.x, (int)mp.y - r + j, cross);
            int qy = (int)((cc.y + s + 1) % s2);
            if (qy > s) qy += pad;
            return qy;
        };
        // the rows' half spectra [j][k] over the band region (free since the block
        // reductions): register columns from the lane holding the row, the rest from
        // the column store
        C* grow = B;
#pragma unroll 1
        for (int j = 0; j < fit; ++j) {
            const int qy = row_of(j);  // the same on every thread
            if (lane == (qy & (kL - 1))) {
                C* gr = grow + j * kNC;
                switch (qy / kL) {
                    case 0: grow_row<0>(gr, xr, g); break;
                    case 1: grow_row<1>(gr, xr, g); break;
                    case 2: grow_row<2>(gr, xr, g); break;
                    case 3: grow_row<3>(gr, xr, g); break;
                    case 4: grow_row<4>(gr, xr, g); break;
                    case 5: grow_row<5>(gr, xr, g); break;
                    case 6: grow_row<6>(gr, xr, g); break;
                    case 7: grow_row<7>(gr, xr, g); break;
                    case 8: grow_row<8>(gr, xr, g); break;
                    case 9: grow_row<9>(gr, xr, g); break;
                    case 10: grow_row<10>(gr, xr, g); break;
                    case 11: grow_row<11>(gr, xr, g); break;
                    case 12: grow_row<12>(gr, xr, g); break;
                    case 13: grow_row<13>(gr, xr, g); break;
                    case 14: grow_row<14>(gr, xr, g); break;
                    default: grow_row<15>(gr, xr, g); break;
                }
            }
            if (tid < kG) grow[j * kNC + kG * kJR + tid] = cols[tid * kColStride + qy];
        }
        __syncthreads();
        {
            // four threads per value, frequencies k = 4 m + p; their partial sums are
            // added by lane shuffles
            const bool act = tid < 4 * kk;
            const int o = act ? tid >> 2 : 0, p = tid & 3;
            const int i = o % fit, j = o / fit;
            const size2 cc = clamp_pos((int)mp.x - r + i, (int)mp.y - r + j, cross);
            int qx = (int)((cc.x + s + 1) % s2);
            if (qx > s) qx += pad;
            const C* gr = grow + j * kNC;
            float acc = 0.f;
#pragma unroll 8
            for (int m = 0; m < kS / 4; ++m) {
                const int k = 4 * m + p;
                const C gk = gr[k], w = tw[(k * qx) & (kM - 1)];  // w = e^{-2 pi i k qx / M}
                acc = fmaf(gk.re, w.re, acc);
                acc = fmaf(gk.im, w.im, acc);
            }
            acc += __shfl_xor(acc, 1);
            acc += __shfl_xor(acc, 2);
            // the k = 0 term is in acc once, needed once: 2 acc - G[0]
            const float v = 2.f * acc - gr[0].re + ((qx & 1) ? -gr[kS].re : gr[kS].re);
            if (act && p == 0) {
                nbv[o] = v;
                if (a.neigh) a.neigh[(size_t)sn * kk + o] = v;
            }
        }
        __syncthreads();
        if (tid < 6) {
#pragma clang fp contract(off)
            const double* mrow = lsq_s + tid * kk;
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
}

// ---------------------------------------------------------------- launcher
template <class IN, class SUM, bool FP, bool PIPE, bool FULL>
static void launch_fft256_s(fft_args<float, IN, SUM> a, hipStream_t st) {
    auto k = fft256_kernel<IN, SUM, FP, PIPE, FULL>;
    static const hipError_t attr =
        hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLds);
    EMIDA_HIP(attr);
    k<<<a.total, kNT, kLds, st>>>(a);
    EMIDA_LAUNCHED("fft256_kernel");
}

template <class IN, class SUM, bool FP, bool PIPE>
void launch_fft256(fft_args<float, IN, SUM> a, hipStream_t st) {
    if (a.total == 0) return;
    if (a.side == (uint32_t)kS) launch_fft256_s<IN, SUM, FP, PIPE, true>(a, st);
    else launch_fft256_s<IN, SUM, FP, PIPE, false>(a, st);
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
