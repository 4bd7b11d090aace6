// Brute-force correlation as an implicit GEMM on CDNA4 matrix cores ("precision
// form"): BF16 or FP16 operands, FP32 accumulation, optionally split into hi + lo
// halves with three products (hi*hi + hi*lo + lo*hi) to win back precision.
//
// Reference semantics: src/cross_corr.cu:11-66, res[sy][sx] = sum over the overlap
// of pic[y + sy][x + sx] * ref[y][x].  For one ROI and a batch of patterns p this is
//   C[p][sy][sx] = sum_y sum_k P_p[y + sy][k] * B_y[k][sx],   B_y[k][sx] = R[y][k - sx]
// i.e. a GEMM with M = patterns, N = x-shifts and K = pixel columns, summed over the
// reference rows y, where B_y is the Toeplitz matrix of reference row y (zero outside
// the row).  All-zero Toeplitz tiles are skipped.
//
// Work decomposition: one workgroup = 32 patterns x 64 WX x-shifts (2 WX tiles of 32,
// wave wx of a row taking tiles wx and wx + WX) x 4 (4 / WX) consecutive y-shifts (4 per
// wave), WX = 4, 2 or 1 for pattern rows of up to 128, 64 or 32 columns, on a 1-D grid
// whose order keeps the y-shift tiles of one (ROI, pattern tile) on one XCD.  It walks
// the reference rows y; the pattern rows Y = y + sy of its y-shifts live in an LDS ring
// (one new row per step), and reference row y is staged as 8 shifted copies so that
// any 8-element window - one lane's B fragment - is a single aligned 16-byte LDS read.
// Per K tile a wave reads each B fragment once for its 4 y-shifts and each A fragment
// once for both of its x-shift tiles.  The next step's rows are loaded into registers
// before the matrix-core work of the current step.
//
// Matrix-core operand layout (v_mfma_f32_32x32x16_{bf16,f16}): lane l holds
// A[l % 32][8 (l / 32) + t] and B[8 (l / 32) + t][l % 32] for t < 8; the 16
// accumulators are D[8 (j / 4) + 4 (l / 32) + j % 4][l % 32], j < 16.
#include <cmath>
#include <type_traits>

#include "emida_amd.h"
#include "kernels.hpp"

namespace emida_amd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Operands enter the half type scaled by `in` (a power of two, exact) and the
// accumulators leave multiplied by `out` = 1 / in^2: fp16 keeps |v| < 65504 after
// halving, which every prepared uint16 value satisfies (|px - mean| <= 65535, window
// <= 1); bf16 has the fp32 exponent range and needs no scaling.
template <class HT>
struct mfma_vec;
template <>
struct mfma_vec<__bf16> {
    static constexpr float in = 1.0f, out = 1.0f;
    using type = bf16x8;
    __device__ static f32x16 mma(bf16x8 a, bf16x8 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct mfma_vec<_Float16> {
    static constexpr float in = 0.5f, out = 4.0f;
    using type = f16x8;
    __device__ static f32x16 mma(f16x8 a, f16x8 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};

// Prepared value in float, the reference's order (src/prepare_pics.cu:36-41), split
// into hi = round(v) and lo = round(v - hi) in the half type.
template <class IN, class S, class HT>
__global__ __launch_bounds__(256) void prepare_split_kernel(const IN* __restrict__ pics, HT* __restrict__ hi,
                                                            HT* __restrict__ lo, const float* __restrict__ hx,
                                                            const float* __restrict__ hy, const S* __restrict__ sums,
                                                            const size2* __restrict__ begins, size2 src, size2 slice,
                                                            uint32_t nb, uint32_t total) {
#pragma clang fp contract(off)
    // blocks_per_slice blocks per slice: wave-uniform 32-bit slice index
    const uint32_t area = slice.area(), bps = div_up(area, 256u);
    const uint32_t sn = blockIdx.x / bps;
    const uint32_t st = (blockIdx.x - sn * bps) * 256u + threadIdx.x;
    if (sn >= total || st >= area) return;
    const size_t tid = (size_t)sn * area + st;
    const uint32_t x = st % slice.x, y = st / slice.x;
    const size2 b0 = begins[sn % nb];
    float v = (float)pics[(size_t)(sn / nb) * src.area() + (size_t)(b0.y + y) * src.x + b0.x + x];
    v -= (float)sums[sn] / (float)area;
    v = v * hx[x] * hy[y];
    v *= mfma_vec<HT>::in;
    const HT h = (HT)v;
    hi[tid] = h;
    if (lo) lo[tid] = (HT)(v - (float)h);
}

constexpr int kMfmaWaves = 4;
constexpr int kMfmaTilesPerWave = 2;
constexpr int kMfmaM = 32;  // patterns per workgroup
constexpr int kMfmaNT = kMfmaWaves * 64;
constexpr int kMfmaTS = 4;  // y-shifts per wave

// Geometry for pattern rows of at most KMAX columns: WX waves side by side along the
// x-shifts (two 32-wide tiles each, 64 WX x-shifts per workgroup) and 4 / WX waves
// along the y-shifts (4 each).  KMAX = 128 pairs with WX = 4 (windows up to 256
// wide), 64 with 2 (up to 128), 32 with 1 (up to 64): every wave has work and the
// LDS ring stays at most 64 KiB per plane pair.
template <class HT, bool SPLIT, int KMAX, int WX>
struct mfma_plan {
    static constexpr int WY = kMfmaWaves / WX;
    static constexpr int ROWS = kMfmaTS * WY;        // y-shifts per workgroup = ring slots
    static constexpr int SXW = WX * kMfmaTilesPerWave * 32;  // x-shifts per workgroup
    static constexpr int CHMAX = KMAX / 8;           // 16-byte chunks per pattern row
    static constexpr int SWZ = 128 / KMAX;           // rows per 256-byte LDS sweep
    static constexpr int RP = KMAX + SXW + 8;        // length of one shifted reference copy
    static constexpr int NH = SPLIT ? 2 : 1;         // hi (+ lo) planes
    static constexpr size_t ring = (size_t)ROWS * kMfmaM * KMAX;  // elements of one plane of the row ring
    static constexpr size_t lds_a = NH * ring * sizeof(HT);
    static constexpr size_t lds_r = (size_t)NH * 8 * RP * sizeof(HT);
    static constexpr size_t lds_bytes = lds_a + lds_r;
    static constexpr int PF = (kMfmaM * CHMAX + kMfmaNT - 1) / kMfmaNT;  // prefetched row chunks per thread
    // 16-byte chunk c of pattern row m is stored at chunk swz(c, m): the 16 rows one
    // fragment read touches fall on 16 distinct 16-byte bank groups
    __device__ static int swz(int c, int m) { return c ^ ((m / SWZ) & (CHMAX - 1)); }
};

// hi/lo: prepared pattern slices [slice][y][x] (slice = pattern * nb + roi);
// rhi/rlo: prepared reference slices [roi][y][x]; corr: [slice][cross.y][cross.x].
// n_wg = nb * ceil(npat / 32) * ceil(cross.y / ROWS) workgroups (1-D grid).
// The 8 prepared values of a row chunk from its uint16 pixels, prepare_split_kernel's
// arithmetic: ((px - mean) * wx) * wy * in, then hi = round(v), lo = round(v - hi).
template <class HT, bool SPLIT, class V>
__device__ inline void prep_chunk(const uint32_t (&u)[4], float mean, const float (&wx)[8], float wy, V& h, V& l) {
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float v = (float)((u[k / 2] >> (16 * (k & 1))) & 0xffffu);
        v -= mean;
        v = v * wx[k];
        v = v * wy;
        v *= mfma_vec<HT>::in;
        const HT hh = (HT)v;
        h[k] = hh;
        if constexpr (SPLIT) l[k] = (HT)(v - (float)hh);
    }
}

// FP: the pattern rows are read as uint16 pixels and prepared while they are staged
// (phi/plo unused); otherwise phi/plo hold the prepared halves.
template <class HT, bool SPLIT, int KMAX, int WX, bool FP, bool FULL>
__global__ __launch_bounds__(kMfmaNT, 2) void mfma_corr_kernel(const HT* __restrict__ phi,
                                                                const HT* __restrict__ plo,
                                                                const HT* __restrict__ rhi,
                                                                const HT* __restrict__ rlo,
                                                                float* __restrict__ corr, size2 s, size2 cross,
                                                                uint32_t nb, uint32_t npat, uint32_t n_wg,
                                                                uint32_t* __restrict__ smax, screen_lists lists,
                                                                mfma_src src) {
    using P = mfma_plan<HT, SPLIT, KMAX, WX>;
    using V = typename mfma_vec<HT>::type;
    constexpr int NT = kMfmaNT, TS = kMfmaTS, ROWS = P::ROWS;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    // A: [NH][ROWS][32][KMAX] pattern-row ring, 16-byte chunk c of row m stored at
    //    chunk P::swz(c, m) (conflict-free fragment reads without padding);
    // Rs: [NH][8][RP] shifted reference copies.
    HT* A = reinterpret_cast<HT*>(lds_raw);
    HT* Rs = reinterpret_cast<HT*>(lds_raw + P::lds_a);

    // the wave index is wave-uniform: in a scalar register, the tile conditions below
    // are scalar branches
    const int tid = threadIdx.x, lane = tid % 64, wave = __builtin_amdgcn_readfirstlane(tid / 64);
    const uint32_t n_syt = div_up(cross.y, (uint32_t)ROWS);
    const uint32_t L = xcd_remap(blockIdx.x, n_wg);  // y-shift tiles of one (ROI, pattern tile) back to back
    const uint32_t syt = L % n_syt, slab = L / n_syt;
    const uint32_t roi = slab % nb, mt = slab / nb;
    const int rx = (int)(cross.x - 1) / 2, ry = (int)(cross.y - 1) / 2;
    const int sy0 = (int)syt * ROWS - ry;  // first y-shift of the workgroup
    const int off = P::SXW - rx;            // shifted-copy offset, see below
    // FULL: KMAX x KMAX slices (compile-time sizes, a fixed number of K steps)
    const int S = FULL ? KMAX : (int)s.x, SY = FULL ? KMAX : (int)s.y, CH = S / 8;
    const int K16 = (S + 15) / 16 * 16;
    const uint32_t p0 = mt * kMfmaM;
    const size_t slice_elems = (size_t)s.x * s.y;
    const HT* refh = rhi + (size_t)roi * slice_elems;
    const HT* refl = SPLIT ? rlo + (size_t)roi * slice_elems : nullptr;

    // this wave's two 32-wide x-shift tiles, wx and wx + WX, and its 4 y-shifts
    // sy0 + 4 wy + t.  The non-zero Toeplitz tiles per x-shift tile fall off linearly
    // from the centre (2, 4, .., 8, 8, .., 4, 2 K tiles for 128 columns), so tiles
    // wx and wx + WX always add up to the same work and no wave waits for another
    // at the per-row barriers.
    const int wx = wave % WX, wy = wave / WX;
    const int syw = sy0 + TS * wy;
    int sx0[kMfmaTilesPerWave];
    bool active[kMfmaTilesPerWave];
#pragma unroll
    for (int u = 0; u < kMfmaTilesPerWave; ++u) {
        sx0[u] = -rx + (wx + u * WX) * 32;
        active[u] = sx0[u] + rx < (int)cross.x;
    }

    f32x16 acc[TS][kMfmaTilesPerWave];
#pragma unroll
    for (int t = 0; t < TS; ++t)
#pragma unroll
        for (int u = 0; u < kMfmaTilesPerWave; ++u) acc[t][u] = f32x16{0};

    // zero LDS once: chunks past the slice width, rows of absent patterns and the
    // reference copies outside the row stay zero for the whole kernel
    for (int e = tid; e < (int)(P::lds_bytes / 16); e += NT) reinterpret_cast<uint4*>(lds_raw)[e] = uint4{0, 0, 0, 0};
    __syncthreads();

    // ---- staging through registers: global loads for step y + 1 are issued before the
    // matrix-core work of step y and written to LDS at the start of step y + 1
    V pfh[P::PF], pfl[P::PF];
    // FP: each thread's chunks (pattern m, columns 8c..8c+7) are the same in every row:
    // their pattern row pointer, mean and window values are set once; the pixels of
    // the next row (8 uint16 per chunk, packed in pairs) and its window value are
    // loaded ahead like the halves
    const uint16_t* pbase[P::PF];
    float pmean[P::PF], pwx[P::PF][8];
    bool pval[P::PF];
    uint32_t pfu[P::PF][4];
    float pwy = 0.f;
    if constexpr (FP) {
        const size2 b0 = src.begins[roi];
#pragma unroll
        for (int i = 0; i < P::PF; ++i) {
            const int e = tid + NT * i, m = e / CH, c = e % CH;
            pval[i] = e < kMfmaM * CH && p0 + m < npat;
            const uint32_t p = pval[i] ? p0 + m : p0;  // p0 < npat: a valid slice
            const int cc = pval[i] ? c : 0;
            pbase[i] = src.pics + (size_t)p * src.src.area() + (size_t)b0.y * src.src.x + b0.x + 8 * cc;
            pmean[i] = (float)src.sums[(size_t)p * nb + roi] / (float)(S * SY);
#pragma unroll
            for (int k = 0; k < 8; ++k) pwx[i][k] = src.hx[8 * cc + k];
        }
    }
    auto load_prow = [&](int Y) {
        if constexpr (FP) {
            if (Y >= 0 && Y < SY) {
                pwy = src.hy[Y];
#pragma unroll
                for (int i = 0; i < P::PF; ++i) {
                    const uint16_t* q = pbase[i] + (size_t)Y * src.src.x;
#pragma unroll
                    for (int k = 0; k < 4; ++k) pfu[i][k] = (uint32_t)q[2 * k] | ((uint32_t)q[2 * k + 1] << 16);
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < P::PF; ++i) {
                const int e = tid + NT * i;
                V vh{}, vl{};
                if (e < kMfmaM * CH && Y >= 0 && Y < SY) {
                    const int m = e / CH, c = e % CH;
                    if (p0 + m < npat) {
                        const size_t g = ((size_t)(p0 + m) * nb + roi) * slice_elems + (size_t)Y * S + 8 * c;
                        vh = *reinterpret_cast<const V*>(phi + g);
                        if constexpr (SPLIT) vl = *reinterpret_cast<const V*>(plo + g);
                    }
                }
                pfh[i] = vh;
                pfl[i] = vl;
            }
        }
    };
    // FP: the loaded pixels into halves (a separate step, placed after the matrix-core
    // work of a step rather than between the barrier and that work)
    auto cvt_prow = [&]() {
        if constexpr (FP) {
#pragma unroll
            for (int i = 0; i < P::PF; ++i) {
                V h{}, l{};
                if (pval[i]) prep_chunk<HT, SPLIT, V>(pfu[i], pmean[i], pwx[i], pwy, h, l);
                pfh[i] = h;
                pfl[i] = l;
            }
        }
    };
    // pattern row Y into its ring slot; rows past the slice are zero rows (read by the
    // last steps' y-shifts, so the matrix-core loop needs no per-row condition); rows
    // before it map to slots still holding the initial zeros
    auto store_prow = [&](int Y) {
        if (Y < 0) return;
        const bool zero = Y >= SY;
        const int slot = Y % ROWS;
#pragma unroll
        for (int i = 0; i < P::PF; ++i) {
            const int e = tid + NT * i;
            if (e < kMfmaM * CH) {
                const int m = e / CH, c = e % CH;
                const size_t o = ((size_t)slot * kMfmaM + m) * KMAX + 8 * P::swz(c, m);
                *reinterpret_cast<V*>(A + o) = zero ? V{} : pfh[i];
                if constexpr (SPLIT) *reinterpret_cast<V*>(A + P::ring + o) = zero ? V{} : pfl[i];
            }
        }
    };
    // reference row y as 8 shifted copies: Rs[rho][i] = R[y][i + rho - off] (0 outside
    // the row), so the window R[y][q .. q+7] is Rs[rho][jp - rho ..] with jp = q + off,
    // rho = jp & 7.  q spans [rx - SXW + 1, K16 - 8 + rx], hence jp in [1, K16 + SXW - 8].
    HT rfh = HT(0.0f), rfl = HT(0.0f);
    auto load_rrow = [&](int y) {
        if (tid < S) {
            rfh = refh[(size_t)y * S + tid];
            if constexpr (SPLIT) rfl = refl[(size_t)y * S + tid];
        }
    };
    auto store_rrow = [&]() {
        if (tid < S) {
#pragma unroll
            for (int rho = 0; rho < 8; ++rho) {
                const int i = tid - rho + off;
                Rs[rho * P::RP + i] = rfh;
                if constexpr (SPLIT) Rs[(8 + rho) * P::RP + i] = rfl;
            }
        }
    };

    // y range: some y-shift of the workgroup overlaps reference row y
    const int y_lo = max(0, -(sy0 + ROWS - 1)), y_hi = min(SY, SY - sy0);
    for (int Y = y_lo + sy0; Y < y_lo + sy0 + ROWS - 1; ++Y) {
        load_prow(Y);
        cvt_prow();
        store_prow(Y);
    }
    if (y_lo < y_hi) {
        load_prow(y_lo + sy0 + ROWS - 1);
        load_rrow(y_lo);
        cvt_prow();
    }
    const int m_a = lane % 32, half = lane / 32;
    for (int y = y_lo; y < y_hi; ++y) {
        store_prow(y + sy0 + ROWS - 1);
        store_rrow();
        __syncthreads();
        if (y + 1 < y_hi) {
            load_prow(y + sy0 + ROWS);
            load_rrow(y + 1);
        }
        for (int k0 = 0; k0 < K16; k0 += 16) {
            // a tile is all zero unless some 0 <= k - sx < S with k in [k0, k0 + 16) and
            // sx in [sx0, sx0 + 32)
            bool nz[kMfmaTilesPerWave];
            bool any = false;
#pragma unroll
            for (int u = 0; u < kMfmaTilesPerWave; ++u) {
                nz[u] = active[u] && !(k0 + 15 < sx0[u] || k0 - 31 - sx0[u] > S - 1);
                any = any || nz[u];
            }
            if (!any) continue;
            const int cs = 8 * P::swz(k0 / 8 + half, m_a);
            // the step for a compile-time set of non-zero tiles (U0, U1): straight-line
            // loads and products, the tile conditions tested once per step
            auto step = [&](auto u0, auto u1) {
                constexpr bool U[2] = {decltype(u0)::value, decltype(u1)::value};
                V bh[kMfmaTilesPerWave], bl[kMfmaTilesPerWave];
#pragma unroll
                for (int u = 0; u < kMfmaTilesPerWave; ++u) {
                    if (!U[u]) continue;
                    const int jp = k0 + 8 * half - sx0[u] - m_a + off;
                    const int rho = jp & 7;
                    bh[u] = *reinterpret_cast<const V*>(Rs + rho * P::RP + (jp - rho));
                    if constexpr (SPLIT) bl[u] = *reinterpret_cast<const V*>(Rs + (8 + rho) * P::RP + (jp - rho));
                }
#pragma unroll
                for (int t = 0; t < TS; ++t) {
                    const int Y = y + syw + t;  // >= -(ROWS - 1)
                    // rows outside the slice read zero rows; the split modes skip them
                    // unless the K loop is unrolled (three products per row: the skip pays
                    // there, ~4% the other way in the plain modes and 2% in the unrolled
                    // split loop, where the branch per row costs more than the products)
                    if constexpr (SPLIT && !FULL) {
                        if (Y < 0 || Y >= SY) continue;
                    }
                    const HT* arow = A + ((size_t)((Y + ROWS) % ROWS) * kMfmaM + m_a) * KMAX + cs;
                    const V ah = *reinterpret_cast<const V*>(arow);
                    V al{};
                    if constexpr (SPLIT) al = *reinterpret_cast<const V*>(arow + P::ring);
#pragma unroll
                    for (int u = 0; u < kMfmaTilesPerWave; ++u) {
                        if (!U[u]) continue;
                        acc[t][u] = mfma_vec<HT>::mma(ah, bh[u], acc[t][u]);
                        if constexpr (SPLIT) {
                            acc[t][u] = mfma_vec<HT>::mma(ah, bl[u], acc[t][u]);
                            acc[t][u] = mfma_vec<HT>::mma(al, bh[u], acc[t][u]);
                        }
                    }
                }
            };
            static_assert(kMfmaTilesPerWave == 2, "two tiles per wave");
            using T1 = std::true_type;
            using F1 = std::false_type;
            if (nz[0]) step(T1{}, F1{});
            if (nz[1]) step(F1{}, T1{});
        }
        // the next row's halves while this step's matrix-core work drains (converting
        // them halfway through the step costs 4 spilled registers at 128 columns)
        if (y + 1 < y_hi) cvt_prow();
        __syncthreads();
    }
    // ---- epilogue.  The workgroup holds complete screen rows (every x-shift of its
    // y-shifts and patterns), so a row's maximum is final here.  Per-row maxima (NaN
    // never wins a strict compare): for each y-shift, per accumulator row j the lane's
    // largest value over its two tiles, then over the 32 lanes of the half by a
    // reduce-scatter (xor 16, 8, 4, 2 halve the rows a lane keeps, xor 1 pairs the last
    // one: 16 lane exchanges instead of 80), so lane l ends with row j = 8 b4 + 4 b3 +
    // 2 b2 + b1 (bits of l) and the even lane of each pair folds pattern row
    // 8 (j/4) + 4 (l/32) + j%4 into the workgroup's LDS maximum of that screen row
    // (over the WX waves of the row).  Stored per (slice, row); with lists, every value
    // within tau of the largest row maximum of the workgroup's rows of its pattern is
    // published as well (a superset of the candidates: the slice maximum is at least that).
    if (smax) {
        __syncthreads();  // the row ring is free
        uint32_t* lcode = reinterpret_cast<uint32_t*>(lds_raw);  // [ROWS][32] row maximum codes
        uint32_t* lcnt = lcode + ROWS * kMfmaM;                  // [ROWS][32] published positions
        double* lthr = reinterpret_cast<double*>(lcnt + ROWS * kMfmaM);  // [ROWS][32] thresholds
        for (int i = tid; i < ROWS * kMfmaM; i += NT) lcode[i] = lcnt[i] = 0u;
        __syncthreads();
        auto keep_max = [](float& k, float o) {
            if (o > k) k = o;
        };
#pragma unroll
        for (int t = 0; t < TS; ++t) {
            const int row = syw + t + ry;
            if (row >= (int)cross.y) continue;
            float mx[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) mx[j] = -__builtin_huge_valf();
#pragma unroll
            for (int u = 0; u < kMfmaTilesPerWave; ++u) {
                const bool in = sx0[u] + m_a + rx < (int)cross.x;
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const float v = acc[t][u][j] * mfma_vec<HT>::out;
                    if (in && v > mx[j]) mx[j] = v;
                }
            }
#pragma unroll
            for (int w = 8; w >= 1; w >>= 1) {  // xor 2w: keep the upper w rows if that lane bit is set
                const bool up = lane & (2 * w);
#pragma unroll
                for (int i = 0; i < w; ++i) {
                    const float lo = mx[i], hi = mx[w + i];
                    float k = up ? hi : lo;
                    keep_max(k, __shfl_xor(up ? lo : hi, 2 * w));
                    mx[i] = k;
                }
            }
            keep_max(mx[0], __shfl_xor(mx[0], 1));
            const int j = 8 * ((lane >> 4) & 1) + 4 * ((lane >> 3) & 1) + 2 * ((lane >> 2) & 1) + ((lane >> 1) & 1);
            if (!(lane & 1) && mx[0] > -__builtin_huge_valf())
                atomicMax(&lcode[(TS * wy + t) * kMfmaM + 8 * (j / 4) + 4 * half + (j % 4)], screen_max_code(mx[0]));
        }
        __syncthreads();
        for (int i = tid; i < ROWS * kMfmaM; i += NT) {
            const int row = sy0 + i / kMfmaM + ry;
            const uint32_t p = p0 + i % kMfmaM;
            if (row >= (int)cross.y || p >= npat) continue;
            const size_t sl = (size_t)p * nb + roi;
            const uint32_t code = lcode[i];
            smax[sl * cross.y + row] = code;
            if (lists.cnt) {
                // the slice maximum is at least the largest row maximum of this
                // workgroup's rows of the pattern and at least z0 - tau (z0: the
                // pattern's dot product with the reference at the zero shift, whose
                // screen value is within tau/2 of it): positions more than tau below
                // that bound cannot be candidates
                uint32_t best = 0;
                for (int r = 0; r < ROWS; ++r) best = max(best, lcode[r * kMfmaM + i % kMfmaM]);
                const double tau = refine_tau(lists, sl, roi);
                const double lb = fmax(best ? (double)screen_max_value(best) : -__builtin_huge_val(),
                                       lists.z0[sl] - tau);
                lthr[i] = best ? lb - tau : __builtin_huge_val();
            }
        }
        if (lists.cnt) {
            __syncthreads();
#pragma unroll
            for (int t = 0; t < TS; ++t) {
                const int row = syw + t + ry;
                if (row >= (int)cross.y) continue;
#pragma unroll
                for (int u = 0; u < kMfmaTilesPerWave; ++u) {
                    const int x = sx0[u] + m_a + rx;
                    if (x >= (int)cross.x) continue;
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const int m = 8 * (j / 4) + 4 * half + (j % 4);
                        const uint32_t p = p0 + m;
                        if (p >= npat) continue;
                        const int li = (TS * wy + t) * kMfmaM + m;
                        const float v = acc[t][u][j] * mfma_vec<HT>::out;
                        if ((double)v >= lthr[li]) {
                            const uint32_t c = atomicAdd(&lcnt[li], 1u);
                            if (c < (uint32_t)kListCap)
                                lists.ent[(((size_t)p * nb + roi) * cross.y + row) * kListCap + c] =
                                    cand_entry{v, (uint32_t)x};
                        }
                    }
                }
            }
            __syncthreads();
            for (int i = tid; i < ROWS * kMfmaM; i += NT) {
                const int row = sy0 + i / kMfmaM + ry;
                const uint32_t p = p0 + i % kMfmaM;
                if (row >= (int)cross.y || p >= npat) continue;
                lists.cnt[((size_t)p * nb + roi) * cross.y + row] = lcnt[i];
            }
        }
    }
    if (!corr) return;
    // write the accumulators: D[8 (j/4) + 4 (l/32) + j%4][l%32]
#pragma unroll
    for (int t = 0; t < TS; ++t) {
        const int sy = syw + t;
        if (sy + ry >= (int)cross.y) continue;
#pragma unroll
        for (int u = 0; u < kMfmaTilesPerWave; ++u) {
            const int sx = sx0[u] + m_a;
            if (sx + rx >= (int)cross.x) continue;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const uint32_t m = 8 * (j / 4) + 4 * half + (j % 4);
                const uint32_t p = p0 + m;
                if (p >= npat) continue;
                corr[(((size_t)p * nb + roi) * cross.y + (sy + ry)) * cross.x + (sx + rx)] =
                    acc[t][u][j] * mfma_vec<HT>::out;
            }
        }
    }
}

// --------------------------------------------------------------- launchers
template <class HT, bool SPLIT, int KMAX, int WX, bool FP>
void launch_mfma_corr_g(const HT* phi, const HT* plo, const HT* rhi, const HT* rlo, float* corr, size2 s, size2 cross,
                        uint32_t nb, uint32_t npat, hipStream_t st, uint32_t* smax, const screen_lists& lists,
                        const mfma_src& src) {
    using P = mfma_plan<HT, SPLIT, KMAX, WX>;
    static_assert(P::lds_bytes <= 80 * 1024, "two workgroups per CU");
    static_assert((size_t)P::ROWS * kMfmaM * 16 <= P::lds_bytes, "epilogue row tables exceed the LDS");
    const bool full = s.x == (uint32_t)KMAX && s.y == (uint32_t)KMAX;
    auto kern = full ? mfma_corr_kernel<HT, SPLIT, KMAX, WX, FP, true> : mfma_corr_kernel<HT, SPLIT, KMAX, WX, FP, false>;
    static const hipError_t attr =
        hipFuncSetAttribute((const void*)mfma_corr_kernel<HT, SPLIT, KMAX, WX, FP, false>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)P::lds_bytes);
    static const hipError_t attr_full =
        hipFuncSetAttribute((const void*)mfma_corr_kernel<HT, SPLIT, KMAX, WX, FP, true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)P::lds_bytes);
    EMIDA_HIP(attr);
    EMIDA_HIP(attr_full);
    if (s.x > (uint32_t)KMAX || cross.x > (uint32_t)P::SXW) throw arg_error("matrix-core geometry mismatch");
    const uint64_t n_wg = (uint64_t)nb * div_up(npat, (uint32_t)kMfmaM) * div_up(cross.y, (uint32_t)P::ROWS);
    if (n_wg > (1u << 24)) throw unsupported_error("matrix-core brute policy: too many workgroups in one launch");
    kern<<<(uint32_t)n_wg, kMfmaNT, P::lds_bytes, st>>>(phi, plo, rhi, rlo, corr, s, cross, nb, npat,
                                                        (uint32_t)n_wg, smax, lists, src);
    EMIDA_LAUNCHED("mfma_corr_kernel");
}

// Row width class of a slice: 32, 64 or 128 columns (windows up to 64, 128, 256 wide).
template <class HT, bool SPLIT, bool FP>
void launch_mfma_corr_t(const HT* phi, const HT* plo, const HT* rhi, const HT* rlo, float* corr, size2 s, size2 cross,
                        uint32_t nb, uint32_t npat, hipStream_t st, uint32_t* smax, const screen_lists& lists,
                        const mfma_src& src) {
    if (s.x <= 32)
        launch_mfma_corr_g<HT, SPLIT, 32, 1, FP>(phi, plo, rhi, rlo, corr, s, cross, nb, npat, st, smax, lists, src);
    else if (s.x <= 64)
        launch_mfma_corr_g<HT, SPLIT, 64, 2, FP>(phi, plo, rhi, rlo, corr, s, cross, nb, npat, st, smax, lists, src);
    else
        launch_mfma_corr_g<HT, SPLIT, 128, 4, FP>(phi, plo, rhi, rlo, corr, s, cross, nb, npat, st, smax, lists, src);
}

void check_mfma_shape(size2 s, size2 cross) {
    if (s.x % 8 != 0 || s.x < 8 || s.x > 128 || s.y < 1)
        throw unsupported_error("matrix-core brute policy: slice width must be a multiple of 8 in [8, 128]");
    if (cross.x == 0 || cross.y == 0 || cross.x > 2 * s.x - 1 || cross.y > 2 * s.y - 1)
        throw arg_error("matrix-core brute policy: cross size must be in [1, 2*slice-1]");
}

template <class HT>
void launch_mfma_corr(const HT* phi, const HT* plo, const HT* rhi, const HT* rlo, float* corr, size2 s, size2 cross,
                      uint32_t nb, uint32_t npat, hipStream_t st, uint32_t* smax, const screen_lists* lists,
                      const mfma_src* src) {
    check_mfma_shape(s, cross);
    screen_lists none{};
    if (src && (!src->pics || !src->sums || !src->hx || !src->hy || !src->begins))
        throw arg_error("matrix-core brute policy: pattern staging needs the patterns, sums, windows and corners");
    if (!src && !phi) throw arg_error("matrix-core brute policy: no pattern slices");
    if (lists && (!smax || !lists->cnt || !lists->ent || !lists->pn2 || !lists->rn2 || !lists->z0))
        throw arg_error("matrix-core brute policy: candidate lists need the row maxima, counts, entries and norms");
    if (!corr && !smax) throw arg_error("matrix-core brute policy: no output");
    if (nb == 0 || npat == 0) return;
    const screen_lists& L = lists ? *lists : none;
    const bool split = rlo != nullptr;  // the reference's lo halves exist exactly in split modes
    if (src) {
        if (split)
            launch_mfma_corr_t<HT, true, true>(nullptr, nullptr, rhi, rlo, corr, s, cross, nb, npat, st, smax, L, *src);
        else
            launch_mfma_corr_t<HT, false, true>(nullptr, nullptr, rhi, nullptr, corr, s, cross, nb, npat, st, smax, L,
                                                *src);
    } else {
        if (split != (plo != nullptr)) throw arg_error("matrix-core brute policy: lo halves of one operand only");
        const mfma_src none_src{};
        if (split)
            launch_mfma_corr_t<HT, true, false>(phi, plo, rhi, rlo, corr, s, cross, nb, npat, st, smax, L, none_src);
        else
            launch_mfma_corr_t<HT, false, false>(phi, nullptr, rhi, nullptr, corr, s, cross, nb, npat, st, smax, L,
                                                 none_src);
    }
}

// |screen - T-precision brute value| <= tol_rel ||p|| ||r|| + tol_abs (||p|| + ||r||)
// for one shift, counted twice (the screen maximum's and the candidate's errors):
//  - each operand enters as a bf16/fp16 value (or hi + lo) of the float-prepared value:
//    relative error e_op = 2u + u^2 (one half) or 3.1 u^2 (hi + lo, lo*lo dropped), plus
//    2^-24 when T = double (the float rounding of the prepared value), plus for fp16 an
//    absolute error of the subnormal spacing (2^-24 after undoing the halving);
//  - products of halves are exact in fp32; their fp32 sum over n terms adds
//    gamma_n = n 2^-24 / (1 - n 2^-24) of the sum of magnitudes;
//  - the T-precision brute value itself is within gamma_n(u_T) of the exact sum;
//  - sums of magnitudes over an overlap are bounded by ||p|| ||r|| (Cauchy-Schwarz) and
//    by sqrt(n) ||p|| for the absolute terms.
void mfma_refine_bounds(int mode, size2 s, bool t_double, double& tol_rel, double& tol_abs) {
    const bool f16 = mode == EMIDA_MFMA_F16 || mode == EMIDA_MFMA_F16_SPLIT;
    const bool split = mode == EMIDA_MFMA_BF16_SPLIT || mode == EMIDA_MFMA_F16_SPLIT;
    const double u = f16 ? std::ldexp(1.0, -11) : std::ldexp(1.0, -8);
    const double n = (double)s.x * s.y;
    const double e_op = (split ? 3.1 * u * u : 2 * u + u * u) + (t_double ? 2.01 * std::ldexp(1.0, -24) : 0.0);
    const double nt = n * (split ? 3 : 1), u32 = std::ldexp(1.0, -24);
    const double g_acc = nt * u32 / (1 - nt * u32);
    const double ut = t_double ? std::ldexp(1.0, -53) : u32;
    const double g_t = n * ut / (1 - n * ut);
    const double h = f16 ? 2.1 * std::ldexp(1.0, -24) : 0.0;
    tol_rel = 2.02 * (e_op * (1 + g_acc) + g_acc * (1 + u) * (1 + u) + g_t);
    tol_abs = 2.02 * h * std::sqrt(n) * (1 + g_acc) * (split ? 2 : 1);
}

template <class IN, class S, class HT>
void launch_prepare_split(const IN* pics, HT* hi, HT* lo, const float* hx, const float* hy, const S* sums,
                          const size2* begins, size2 src, size2 slice, uint32_t nb, uint32_t total, hipStream_t st) {
    const size_t blocks = (size_t)div_up(slice.area(), 256u) * total;
    if (blocks == 0) return;
    if (blocks > 0xffffffffu) throw arg_error("prepare: too many slices for one launch");
    prepare_split_kernel<IN, S, HT><<<(unsigned)blocks, 256, 0, st>>>(pics, hi, lo, hx, hy, sums, begins, src, slice,
                                                                     nb, total);
    EMIDA_LAUNCHED("prepare_split_kernel");
}

template <class HT>
float mfma_in_scale() {
    return mfma_vec<HT>::in;
}
template float mfma_in_scale<__bf16>();
template float mfma_in_scale<_Float16>();

// Already prepared float slices -> hi (+ lo) halves.
template <class HT>
__global__ void split_kernel(const float* __restrict__ v, HT* __restrict__ hi, HT* __restrict__ lo, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float w = v[i] * mfma_vec<HT>::in;
    const HT h = (HT)w;
    hi[i] = h;
    if (lo) lo[i] = (HT)(w - (float)h);
}

template <class HT>
void launch_split(const float* v, HT* hi, HT* lo, size_t n, hipStream_t st) {
    if (n == 0) return;
    split_kernel<HT><<<(unsigned)div_up(n, (size_t)256), 256, 0, st>>>(v, hi, lo, n);
    EMIDA_LAUNCHED("split_kernel");
}

#define INST_MFMA(HT)                                                                                          \
    template void launch_mfma_corr<HT>(const HT*, const HT*, const HT*, const HT*, float*, size2, size2, uint32_t, \
                                       uint32_t, hipStream_t, uint32_t*, const screen_lists*, const mfma_src*); \
    template void launch_split<HT>(const float*, HT*, HT*, size_t, hipStream_t);                               \
    template void launch_prepare_split<uint16_t, uint32_t, HT>(const uint16_t*, HT*, HT*, const float*,        \
                                                               const float*, const uint32_t*, const size2*, size2, \
                                                               size2, uint32_t, uint32_t, hipStream_t);            \
    template void launch_prepare_split<float, float, HT>(const float*, HT*, HT*, const float*, const float*,    \
                                                         const float*, const size2*, size2, size2, uint32_t,    \
                                                         uint32_t, hipStream_t);                               \
    template void launch_prepare_split<double, double, HT>(const double*, HT*, HT*, const float*, const float*, \
                                                           const double*, const size2*, size2, size2, uint32_t, \
                                                           uint32_t, hipStream_t);
INST_MFMA(__bf16)
INST_MFMA(_Float16)

}  // namespace emida_amd
