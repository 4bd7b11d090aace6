// Definition-based ("brute") 2-D cross-correlation.
//
// Reference: src/cross_corr.cu:11-74.  For every shift (sx, sy) of the result
// (res index = shift + (res_size - 1) / 2) the sum over the overlap of
// pic[y + sy][x + sx] * ref[y][x] is accumulated in one register, y outer and x
// inner, as fused multiply-adds (the reference's nvcc build contracts
// `sum += a * b`), so double and float results are bit-identical to that order.
//
// Three kernels: a register-blocked one (below, "tiled") for slices up to 128 pixels
// wide, correlation windows up to 256 wide and launches of at least 256 workgroups;
// for other shapes whose slices fit in LDS, one with eight x-shifts per thread sliding
// a register window over zero-padded rows ("x8"); and a thread-per-shift one for the
// rest: one workgroup per (slice, 4-row band of shifts, 64-column band of shifts),
// staged in LDS when both slices fit in 64 KiB.
#include "kernels.hpp"

namespace emida_amd {

template <class T>
__device__ inline T mac(T a, T b, T acc) {
    return fma(a, b, acc);
}
template <>
__device__ inline int32_t mac<int32_t>(int32_t a, int32_t b, int32_t acc) {
    return (int32_t)((uint32_t)acc + (uint32_t)a * (uint32_t)b);
}

constexpr int kBruteTX = 64, kBruteTY = 4;
constexpr size_t kBruteLds = 64 * 1024;

template <class T, bool LDS>
__global__ __launch_bounds__(kBruteTX* kBruteTY) void cross_corr_kernel(const T* __restrict__ pics,
                                                                        const T* __restrict__ ref,
                                                                        T* __restrict__ res, size2 size,
                                                                        size2 res_size, uint32_t ref_slices) {
    extern __shared__ unsigned char smem_raw[];
    const uint32_t sn = blockIdx.z;
    const T* p = pics + (size_t)sn * size.area();
    const T* q = ref + (size_t)(sn % ref_slices) * size.area();
    if constexpr (LDS) {
        T* sp = reinterpret_cast<T*>(smem_raw);
        T* sq = sp + size.area();
        for (uint32_t i = threadIdx.y * kBruteTX + threadIdx.x; i < size.area(); i += kBruteTX * kBruteTY) {
            sp[i] = p[i];
            sq[i] = q[i];
        }
        __syncthreads();
        p = sp;
        q = sq;
    }
    const uint32_t rx = blockIdx.x * kBruteTX + threadIdx.x;
    const uint32_t ry = blockIdx.y * kBruteTY + threadIdx.y;
    if (rx >= res_size.x || ry >= res_size.y) return;
    const int sx = (int)rx - (int)((res_size.x - 1) / 2);
    const int sy = (int)ry - (int)((res_size.y - 1) / 2);
    const int x_end = min((int)size.x - sx, (int)size.x);
    const int y_end = min((int)size.y - sy, (int)size.y);
    const int x0 = max(-sx, 0);
    T sum = 0;
    for (int y = max(-sy, 0); y < y_end; ++y) {
        const T* pr = p + (size_t)(y + sy) * size.x + sx;
        const T* qr = q + (size_t)y * size.x;
        for (int x = x0; x < x_end; ++x) sum = mac<T>(pr[x], qr[x], sum);
    }
    res[(size_t)sn * res_size.area() + (size_t)ry * res_size.x + rx] = sum;
}

// The same sums, NX consecutive x-shifts per thread, for slices staged in LDS: an
// NX-value window slides along the pattern row (one new pattern value and one
// reference value per column feed NX fmas).  Each thread walks the union of its
// shifts' overlaps, rounded up to whole blocks of NX columns, so the window moves in
// registers by compile-time index within a block; the pattern rows are stored with
// NX - 1 zeros before and 2 NX - 2 after them and the reference rows with NX zeros
// after them, so a term outside a shift's own overlap multiplies an exact zero and
// leaves its sum unchanged (finite values), and every sum still sees its own terms
// y outer, x inner.  Indices: the window reads pattern column x + sx0 + t, t < NX,
// for x in [xu0, xu0 + NX * blocks) with xu0 + NX * blocks <= xu1 + NX - 1, which
// stays in [-(NX - 1), S + 2 NX - 2); reference columns stay below S + NX - 1.
constexpr int kXN = 8;  // x-shifts per thread (4: 0.146 ms, 16: 0.167 ms at 32 x 32 slices, 63 x 63 windows; 8: 0.130 ms)
constexpr int kXNTX = 64 / kXN, kXNTY = 256 / kXNTX;
constexpr int kXNPadL = kXN - 1, kXNPadR = 2 * kXN - 2;

template <class T>
__host__ __device__ inline size_t xn_lds_elems(size2 size) {
    return (size_t)size.y * (size.x + kXNPadL + kXNPadR) + (size_t)size.y * (size.x + kXN);
}

template <class T>
__global__ __launch_bounds__(kXNTX* kXNTY) void cross_corr_x8_kernel(const T* __restrict__ pics,
                                                                    const T* __restrict__ ref,
                                                                    T* __restrict__ res, size2 size,
                                                                    size2 res_size, uint32_t ref_slices) {
    constexpr int NX = kXN;
    extern __shared__ unsigned char smem_raw[];
    const uint32_t sn = blockIdx.z;
    const int S = (int)size.x, SY = (int)size.y, W = S + kXNPadL + kXNPadR, SQ = S + NX;
    T* sp = reinterpret_cast<T*>(smem_raw);  // SY rows of W: zeros, pattern row, zeros
    T* sq = sp + (size_t)SY * W;             // SY rows of SQ: reference row, zeros
    const T* p = pics + (size_t)sn * size.area();
    const T* q = ref + (size_t)(sn % ref_slices) * size.area();
    const int tid = threadIdx.y * kXNTX + threadIdx.x;
    for (int i = tid; i < SY * W; i += kXNTX * kXNTY) {
        const int y = i / W, c = i % W - kXNPadL;
        sp[i] = (c >= 0 && c < S) ? p[(size_t)y * S + c] : T(0);
    }
    for (int i = tid; i < SY * SQ; i += kXNTX * kXNTY) {
        const int y = i / SQ, c = i % SQ;
        sq[i] = c < S ? q[(size_t)y * S + c] : T(0);
    }
    __syncthreads();
    const uint32_t rx0 = (blockIdx.x * kXNTX + threadIdx.x) * NX;
    const uint32_t ry = blockIdx.y * kXNTY + threadIdx.y;
    if (rx0 >= res_size.x || ry >= res_size.y) return;
    const int nsh = min(NX, (int)(res_size.x - rx0));  // shifts of this thread in the window
    const int sx0 = (int)rx0 - (int)((res_size.x - 1) / 2);
    const int sy = (int)ry - (int)((res_size.y - 1) / 2);
    const int xu0 = max(0, -(sx0 + nsh - 1)), xu1 = min(S, S - sx0);  // union of the overlaps
    T acc[NX];
#pragma unroll
    for (int t = 0; t < NX; ++t) acc[t] = T(0);
    if (xu0 < xu1) {
        const int blocks = (xu1 - xu0 + NX - 1) / NX;
        for (int y = max(-sy, 0); y < min(SY, SY - sy); ++y) {
            const T* pr = sp + (size_t)(y + sy) * W + kXNPadL + sx0 + xu0;  // pr[j + t]: column xu0 + j + sx0 + t
            const T* qr = sq + (size_t)y * SQ + xu0;
            T w[2 * NX - 1];
#pragma unroll
            for (int t = 0; t < NX - 1; ++t) w[t] = pr[t];
            for (int b = 0; b < blocks; ++b, pr += NX, qr += NX) {
                T qv[NX];
#pragma unroll
                for (int u = 0; u < NX; ++u) {
                    w[NX - 1 + u] = pr[NX - 1 + u];
                    qv[u] = qr[u];
                }
#pragma unroll
                for (int u = 0; u < NX; ++u)
#pragma unroll
                    for (int t = 0; t < NX; ++t) acc[t] = mac<T>(w[u + t], qv[u], acc[t]);
#pragma unroll
                for (int t = 0; t < NX - 1; ++t) w[t] = w[NX + t];
            }
        }
    }
    T* o = res + (size_t)sn * res_size.area() + (size_t)ry * res_size.x + rx0;
#pragma unroll
    for (int t = 0; t < NX; ++t)
        if (t < nsh) o[t] = acc[t];
}

// ------------------------------------------------------------------ tiled kernel
// The same sums, register-blocked.  A workgroup takes 16 patterns of one ROI, a band
// of x-shifts (64 per wave) and a band of y-shifts (4 per wave); it walks the
// reference rows y and, per row, the pattern columns k in blocks of 8.  A thread
// holds 2 patterns x 8 consecutive x-shifts x 4 y-shifts of accumulators; per column
// k it reads 2 x 4 pattern values from LDS and slides an 8-wide window over the
// reference row, so each pattern value feeds 8 and each reference value 8 x 4
// multiply-adds.  The reference row is zero-padded in LDS and pattern rows are
// zero-padded to a multiple of 8, so the loop runs over the union of the thread's
// overlaps: a term outside a shift's overlap adds a product with an exact zero and
// leaves its sum unchanged (for finite values), and every sum still sees its own
// terms y outer, x inner, in the reference's order.
constexpr int kTiledP = 16;   // patterns per workgroup
constexpr int kTiledMI = 2;   // patterns per thread
constexpr int kTiledNI = 8;   // x-shifts per thread
constexpr int kTiledTT = 4;   // y-shifts per thread
constexpr int kTiledSXW = 64; // x-shifts per wave
constexpr size_t kTiledRingMax = 64 * 1024;

struct tiled_shape {
    int wx, wy;       // waves along x-shifts and y-shifts
    int rows;         // pattern rows per step (y-shifts per workgroup) = 4 wy
    int spad;         // slice width rounded up to 8
    int off;          // reference row offset in the padded LDS row
    int rlen;         // padded reference row length
    size_t lds;       // bytes: ring + reference row
};

template <class T>
inline bool tiled_plan(size2 size, size2 res, tiled_shape& t) {
    t.spad = (int)div_up(size.x, 8u) * 8;
    if (t.spad > 128 || res.x > 4u * kTiledSXW) return false;
    t.wx = (int)div_up(res.x, (uint32_t)kTiledSXW);
    t.wy = t.wx == 1 ? 4 : (t.wx == 2 ? 2 : 1);
    while (t.wy > 1 && (size_t)kTiledP * kTiledTT * t.wy * t.spad * sizeof(T) > kTiledRingMax) t.wy /= 2;
    t.rows = kTiledTT * t.wy;
    // a thread's window reads R[y][k - sx] for k in [k8 - 8, khi) and its x-shifts
    // sx in [-rx, 64 wx - rx); with off = 64 wx - rx + 16 every index k - sx + off lies
    // in [16, spad + rx + off)
    const int rx = (int)(res.x - 1) / 2;
    t.off = kTiledSXW * t.wx - rx + 16;
    t.rlen = t.spad + rx + t.off;
    t.lds = (size_t)kTiledP * t.rows * t.spad * sizeof(T) + (size_t)t.rlen * sizeof(T);
    return true;
}

template <class T>
struct pair_of {
    T v[2];
};

// Pattern ring: [rows][spad][16], column k of a row holding the 16 patterns' values in
// pattern order.  Staged with the pattern index fastest (consecutive threads write
// consecutive words); a lane reads its pattern pair of column k8 + j at a per-row base
// plus the compile-time offset 128 j bytes, so the inner loop needs no address
// arithmetic per read (the 8 pairs a wave reads are 128 contiguous bytes).
template <class T>
__global__ __launch_bounds__(256, 2) void cross_corr_tiled_kernel(const T* __restrict__ pics,
                                                                   const T* __restrict__ ref, T* __restrict__ res,
                                                                   size2 size, size2 res_size, uint32_t nb,
                                                                   uint32_t npat, tiled_shape sh, uint32_t n_wg) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* ring = reinterpret_cast<T*>(smem_raw);
    T* rrow = ring + (size_t)kTiledP * sh.rows * sh.spad;
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid % 64, wave = tid / 64;
    const int S = (int)size.x, SY = (int)size.y, SP = sh.spad, ROWS = sh.rows;
    const uint32_t n_syt = div_up(res_size.y, (uint32_t)ROWS);
    const uint32_t L = xcd_remap(blockIdx.x, n_wg);
    const uint32_t syt = L % n_syt, slab = L / n_syt;
    const uint32_t roi = slab % nb, p0 = (slab / nb) * kTiledP;
    const int rx = (int)(res_size.x - 1) / 2, ry = (int)(res_size.y - 1) / 2;
    const int sy0 = (int)syt * ROWS - ry;
    const size_t slice_elems = (size_t)size.x * size.y;
    const T* refs = ref + (size_t)roi * slice_elems;

    const int wx = (wave % sh.wx + (int)syt) % sh.wx, wy = wave / sh.wx;
    const int pg = lane % 8, sg = lane / 8;
    const int sxw0 = -rx + wx * kTiledSXW;          // first x-shift of the wave
    const int sxb = sxw0 + sg * kTiledNI;           // first x-shift of the thread
    const int syt0 = sy0 + wy * kTiledTT;           // first y-shift of the thread

    T acc[kTiledMI][kTiledNI][kTiledTT];
#pragma unroll
    for (int i = 0; i < kTiledMI; ++i)
#pragma unroll
        for (int n = 0; n < kTiledNI; ++n)
#pragma unroll
            for (int t = 0; t < kTiledTT; ++t) acc[i][n][t] = T(0);

    for (size_t e = tid; e * sizeof(T) < sh.lds; e += nt) ring[e] = T(0);
    __syncthreads();

    // pattern row Y into its ring slot; rows past the slice (read by the last
    // steps' y-shifts) are zeros, so the accumulation needs no per-row condition
    // (rows before it only ever map to slots still holding the initial zeros)
    auto stage_prow = [&](int Y) {
        if (Y < 0) return;
        T* dst = ring + (size_t)(Y % ROWS) * SP * kTiledP;
        if (Y >= SY) {
            for (int e = tid; e < kTiledP * SP; e += nt) dst[e] = T(0);
            return;
        }
        for (int e = tid; e < kTiledP * S; e += nt) {
            const int p = e % kTiledP, k = e / kTiledP;
            const uint32_t pp = p0 + p;
            const T v = pp < npat ? pics[((size_t)pp * nb + roi) * slice_elems + (size_t)Y * S + k] : T(0);
            dst[k * kTiledP + p] = v;
        }
    };
    auto stage_rrow = [&](int y) {
        for (int e = tid; e < S; e += nt) rrow[sh.off + e] = refs[(size_t)y * S + e];
    };

    // k range of the wave: union of the overlaps of its x-shifts, in blocks of 8
    const int klo = max(0, sxw0) & ~7;
    const int khi = min(SP, (min(S, S + sxw0 + kTiledSXW - 1) + 7) & ~7);

    const int y_lo = max(0, -(sy0 + ROWS - 1)), y_hi = min(SY, SY - sy0);
    for (int Y = y_lo + sy0; Y < y_lo + sy0 + ROWS - 1; ++Y) stage_prow(Y);
    for (int y = y_lo; y < y_hi; ++y) {
        stage_prow(y + sy0 + ROWS - 1);
        stage_rrow(y);
        __syncthreads();
        if (klo < khi) {
            // window values: of the block being accumulated, cur[j] = R[y][k8 + j - sxb],
            // and of the one before, prev[j] = R[y][k8 - 8 + j - sxb]; two register blocks
            // alternate in the two roles (no copies), two column blocks per iteration
            T wa[8], wb[8];
            const T* rp = rrow + sh.off - sxb;
#pragma unroll
            for (int j = 0; j < 8; ++j) wa[j] = rp[klo - 8 + j];
            // this lane's pattern pair in column klo of the rows of its 4 y-shifts
            const pair_of<T>* pr[kTiledTT];
#pragma unroll
            for (int t = 0; t < kTiledTT; ++t) {
                const int Y = y + syt0 + t;
                const int slot = (Y % ROWS + ROWS) % ROWS;
                pr[t] = reinterpret_cast<const pair_of<T>*>(ring + ((size_t)slot * SP + klo) * kTiledP) + pg;
            }
            // columns k8 .. k8 + 7: load the block into cur, accumulate
            auto block = [&](int k8, const T (&prev)[8], T (&cur)[8]) {
#pragma unroll
                for (int j = 0; j < 8; ++j) cur[j] = rp[k8 + j];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    pair_of<T> a[kTiledTT];
#pragma unroll
                    for (int t = 0; t < kTiledTT; ++t) a[t] = pr[t][j * (kTiledP / 2)];
#pragma unroll
                    for (int t = 0; t < kTiledTT; ++t) {
#pragma unroll
                        for (int n = 0; n < kTiledNI; ++n) {
                            const T b = (j - n >= 0) ? cur[(j - n) & 7] : prev[(8 + j - n) & 7];
#pragma unroll
                            for (int i = 0; i < kTiledMI; ++i) acc[i][n][t] = mac<T>(a[t].v[i], b, acc[i][n][t]);
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < kTiledTT; ++t) pr[t] += 8 * (kTiledP / 2);
            };
            int k8 = klo;
            for (; k8 + 8 < khi; k8 += 16) {
                block(k8, wa, wb);
                block(k8 + 8, wb, wa);
            }
            if (k8 < khi) block(k8, wa, wb);
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < kTiledTT; ++t) {
        const int sy = syt0 + t;
        if (sy + ry >= (int)res_size.y) continue;
#pragma unroll
        for (int i = 0; i < kTiledMI; ++i) {
            const uint32_t p = p0 + 2 * pg + i;
            if (p >= npat) continue;
            T* out = res + (((size_t)p * nb + roi) * res_size.y + (sy + ry)) * res_size.x;
#pragma unroll
            for (int n = 0; n < kTiledNI; ++n) {
                const int sx = sxb + n;
                if (sx + rx < (int)res_size.x) out[sx + rx] = acc[i][n][t];
            }
        }
    }
}

// Workgroups of the register-blocked kernel for this launch, or 0 when the
// thread-per-shift kernel runs instead (shape outside the tiled limits, or fewer
// workgroups than CUs, where the finer grid of the thread-per-shift kernel is faster).
template <class T>
uint64_t tiled_workgroups(size2 size, size2 res_size, uint32_t ref_slices, uint32_t batch, tiled_shape& sh) {
    if (!tiled_plan<T>(size, res_size, sh)) return 0;
    const uint64_t n_wg =
        (uint64_t)ref_slices * div_up(batch, (uint32_t)kTiledP) * div_up(res_size.y, (uint32_t)sh.rows);
    return (n_wg >= 256 && n_wg <= (1u << 24)) ? n_wg : 0;
}

const char* cross_corr_kernel_name(bool is_double, size2 size, size2 res_size, uint32_t ref_slices, uint32_t batch) {
    tiled_shape sh;
    const uint64_t n = is_double ? tiled_workgroups<double>(size, res_size, ref_slices, batch, sh)
                                 : tiled_workgroups<float>(size, res_size, ref_slices, batch, sh);
    if (n) return "cross_corr_tiled_kernel";
    const size_t xn = is_double ? xn_lds_elems<double>(size) * sizeof(double) : xn_lds_elems<float>(size) * sizeof(float);
    return xn <= kBruteLds ? "cross_corr_x8_kernel" : "cross_corr_kernel";
}

template <class T>
void launch_cross_corr(const T* pics, const T* ref, T* res, size2 size, size2 res_size, uint32_t ref_slices,
                       uint32_t batch, hipStream_t st) {
    if (res_size.x > 2 * size.x + 1 || res_size.y > 2 * size.y + 1)
        throw arg_error("cross_corr: res_size must be at most 2*size+1 per axis");
    const uint32_t total = ref_slices * batch;
    if (total == 0 || res_size.area() == 0) return;
    tiled_shape sh;
    if (const uint64_t n_wg = tiled_workgroups<T>(size, res_size, ref_slices, batch, sh)) {
        auto kern = cross_corr_tiled_kernel<T>;
        static const hipError_t attr =
            hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);
        EMIDA_HIP(attr);
        kern<<<(uint32_t)n_wg, 64 * sh.wx * sh.wy, sh.lds, st>>>(pics, ref, res, size, res_size, ref_slices, batch,
                                                                   sh, (uint32_t)n_wg);
        EMIDA_LAUNCHED("cross_corr_tiled_kernel");
        return;
    }
    if (total > 65535u) throw arg_error("cross_corr: more than 65535 slices in one launch");
    if (xn_lds_elems<T>(size) * sizeof(T) <= kBruteLds) {
        dim3 block(kXNTX, kXNTY);
        dim3 grid(div_up(res_size.x, (uint32_t)(kXNTX * kXN)), div_up(res_size.y, (uint32_t)kXNTY), total);
        cross_corr_x8_kernel<T><<<grid, block, xn_lds_elems<T>(size) * sizeof(T), st>>>(pics, ref, res, size,
                                                                                        res_size, ref_slices);
        EMIDA_LAUNCHED("cross_corr_x8_kernel");
        return;
    }
    dim3 block(kBruteTX, kBruteTY);
    dim3 grid(div_up(res_size.x, (uint32_t)kBruteTX), div_up(res_size.y, (uint32_t)kBruteTY), total);
    const size_t lds = 2 * (size_t)size.area() * sizeof(T);
    if (lds <= kBruteLds) {
        cross_corr_kernel<T, true><<<grid, block, lds, st>>>(pics, ref, res, size, res_size, ref_slices);
    } else {
        cross_corr_kernel<T, false><<<grid, block, 0, st>>>(pics, ref, res, size, res_size, ref_slices);
    }
    EMIDA_LAUNCHED("cross_corr_kernel");
}

template void launch_cross_corr<float>(const float*, const float*, float*, size2, size2, uint32_t, uint32_t,
                                       hipStream_t);
template void launch_cross_corr<double>(const double*, const double*, double*, size2, size2, uint32_t, uint32_t,
                                        hipStream_t);
template void launch_cross_corr<int32_t>(const int32_t*, const int32_t*, int32_t*, size2, size2, uint32_t,
                                         uint32_t, hipStream_t);

}  // namespace emida_amd
