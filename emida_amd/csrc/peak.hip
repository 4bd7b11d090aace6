// Arg-max, neighbourhood gather, least-squares subpixel peak and finalize.
//
// Reference: src/maxarg.cu:119-187 (arg-max per slice), src/extract_neighbors.cu:13-49
// (F x F gather with position policy and clamp), src/subpixel_max.hpp:14-97 and
// src/subpixel_max.cpp:5-58 (coefficients = (A^T A)^-1 A^T * neighbourhood, peak of
// the quadratic), src/gpu_offset.hpp:280-301 (offset formula).  The reference fits
// on the CPU; here one workgroup per slice does all four steps.  The fit runs in
// double without contraction in the reference's order, so it matches the CPU
// result bit for bit.
#include <mutex>
#include <vector>

#include "kernels.hpp"

namespace emida_amd {

// --------------------------------------------------------------- least-squares tables
// (A^T A)^-1 A^T for f(x,y) = a + bx + cy + dx^2 + exy + fy^2 on an s x s grid,
// by exact rational Gauss-Jordan elimination; every entry is rounded once.
namespace {
struct rational {
    __int128 n, d;
};
__int128 gcd_i128(__int128 a, __int128 b) {
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b != 0) {
        __int128 t = a % b;
        a = b;
        b = t;
    }
    return a;
}
rational make_r(__int128 n, __int128 d) {
    if (d < 0) {
        n = -n;
        d = -d;
    }
    __int128 g = gcd_i128(n, d);
    if (g == 0) g = 1;
    return {n / g, d / g};
}
}  // namespace

void lstsq_host(int s, double* out) {
    const int m = s * s;
    std::vector<std::vector<rational>> g(6, std::vector<rational>(6 + m));
    auto basis = [s](int k, int i) -> long long {
        const long long x = k % s, y = k / s;
        const long long v[6] = {1, x, y, x * x, x * y, y * y};
        return v[i];
    };
    for (int i = 0; i < 6; ++i) {
        for (int j = 0; j < 6; ++j) {
            long long acc = 0;
            for (int k = 0; k < m; ++k) acc += basis(k, i) * basis(k, j);
            g[i][j] = make_r(acc, 1);
        }
        for (int k = 0; k < m; ++k) g[i][6 + k] = make_r(basis(k, i), 1);
    }
    for (int c = 0; c < 6; ++c) {
        int piv = c;
        while (g[piv][c].n == 0) ++piv;
        std::swap(g[piv], g[c]);
        const rational p = g[c][c];
        for (auto& e : g[c]) e = make_r(e.n * p.d, e.d * p.n);
        for (int r = 0; r < 6; ++r) {
            if (r == c || g[r][c].n == 0) continue;
            const rational f = g[r][c];
            for (int k = 0; k < 6 + m; ++k) {
                const rational t = make_r(f.n * g[c][k].n, f.d * g[c][k].d);
                g[r][k] = make_r(g[r][k].n * t.d - t.n * g[r][k].d, g[r][k].d * t.d);
            }
        }
    }
    for (int i = 0; i < 6; ++i)
        for (int k = 0; k < m; ++k) out[i * m + k] = (double)(long long)g[i][6 + k].n / (double)(long long)g[i][6 + k].d;
}

int lstsq_offset(int fit) {
    switch (fit) {
        case 3: return 0;
        case 5: return 6 * 9;
        case 7: return 6 * (9 + 25);
        case 9: return 6 * (9 + 25 + 49);
        default: throw arg_error("fitsize must be 3, 5, 7 or 9");
    }
}

// The tables live in a static device array (each GPU holds its own copy): filled once
// per device with a copy, never allocated, so entry points that must not allocate can
// use them; host code uses the address hipGetSymbolAddress gives, never the name.
static __device__ double g_lstsq[6 * (9 + 25 + 49 + 81)];

const double* lstsq_device_table() {
    static std::mutex mu;
    static std::vector<double*> per_device;
    int dev = 0;
    EMIDA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)per_device.size() <= dev) per_device.resize(dev + 1, nullptr);
    if (!per_device[dev]) {
        std::vector<double> host(6 * (9 + 25 + 49 + 81));
        for (int s = 3; s <= 9; s += 2) lstsq_host(s, host.data() + lstsq_offset(s));
        EMIDA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_lstsq), host.data(), host.size() * sizeof(double)));
        void* d = nullptr;
        EMIDA_HIP(hipGetSymbolAddress(&d, HIP_SYMBOL(g_lstsq)));
        if (!d) throw hip_error("hipGetSymbolAddress returned no address for the least-squares table");
        per_device[dev] = (double*)d;
    }
    return per_device[dev];
}

// --------------------------------------------------------------- device pieces
// One coefficient: sum_i M[row][i] * nb[i] in double, i ascending
// (src/subpixel_max.hpp:14-29).
template <class T>
__device__ inline double fit_coef(const double* __restrict__ mrow, const T* nb, int k) {
#pragma clang fp contract(off)
    double sum = 0;
    for (int i = 0; i < k; ++i) sum += mrow[i] * (double)nb[i];
    return sum;
}

// Peak of the fitted quadratic (src/subpixel_max.hpp:91-92).
__device__ inline void fit_peak(const double* c, double& ox, double& oy) {
#pragma clang fp contract(off)
    const double b = c[1], cc = c[2], d = c[3], e = c[4], f = c[5];
    oy = (b * e - 2 * cc * d) / (4 * f * d - e * e);
    ox = (-e * oy - b) / (2 * d);
}

// Offset formula of finalize (src/gpu_offset.hpp:291-297).
__device__ inline void finalize_offset(uint32_t index, size2 cross, int policy, int fit, double sx, double sy,
                                       double& ox, double& oy) {
#pragma clang fp contract(off)
    const size2 mp = shift_pos(policy, index, cross);
    const int r = (fit - 1) / 2;
    ox = -((int)mp.x - ((int)cross.x / 2) - r + sx);
    oy = -((int)mp.y - ((int)cross.y / 2) - r + sy);
}

template <class T>
__device__ inline T lowest() {
    return -__builtin_huge_val();
}
template <>
__device__ inline float lowest<float>() {
    return -__builtin_huge_valf();
}

// --------------------------------------------------------------- fused peak kernel
template <class T, int NT>
__global__ __launch_bounds__(NT) void peak_kernel(const T* __restrict__ corr, size_t corr_len, uint32_t area,
                                                  size2 cross, int policy, int fit, const double* __restrict__ lsq,
                                                  data_index<T>* __restrict__ maxes, T* __restrict__ neigh,
                                                  double* __restrict__ offsets, double* __restrict__ coefs,
                                                  uint32_t total, const uint32_t* __restrict__ row_max) {
    __shared__ T sv[NT / kWave];
    __shared__ uint32_t si[NT / kWave];
    __shared__ T nbv[81];
    __shared__ double cf[6];
    const uint32_t sn = blockIdx.x;
    if (sn >= total) return;
    const T* d = corr + (size_t)sn * area;
    T best = lowest<T>();
    uint32_t bi = 0xffffffffu;
    if (row_max) {
        // the maximum from the row maxima; its lowest index lies in the first row whose
        // maximum equals it, so only that row is read
        const uint32_t* rm = row_max + (size_t)sn * cross.y;
        uint32_t mc = 0;
        for (uint32_t r = threadIdx.x; r < cross.y; r += NT) mc = max(mc, rm[r]);
        mc = ~block_min_u32<NT>(~mc, si);
        const T m = mc ? (T)screen_max_value(mc) : lowest<T>();
        if (m > lowest<T>()) {
            uint32_t row = 0xffffffffu, x = 0xffffffffu;
            for (uint32_t r = threadIdx.x; r < cross.y; r += NT)
                if (rm[r] == mc) row = min(row, r);
            row = block_min_u32<NT>(row, si);
            for (uint32_t i = threadIdx.x; i < cross.x; i += NT)
                if (d[(size_t)row * cross.x + i] == m) x = min(x, i);
            x = block_min_u32<NT>(x, si);
            if (x != 0xffffffffu) {
                best = m;
                bi = row * cross.x + x;
            }
        }
    }
    if (bi == 0xffffffffu) {
        for (uint32_t i = threadIdx.x; i < area; i += NT) {
            const T v = d[i];
            if (v > best) {
                best = v;
                bi = i;
            }
        }
        block_argmax<T, NT>(best, bi, sv, si);
    }
    const int k = fit * fit;
    const size2 mp = shift_pos(policy, bi, cross);
    const int r = (fit - 1) / 2;
    for (int t = threadIdx.x; t < k; t += NT) {
        const int i = t % fit, j = t / fit;
        const size2 c = clamp_pos((int)mp.x - r + i, (int)mp.y - r + j, cross);
        const size_t idx = pos_index(policy, sn, c, cross);
        const T v = idx < corr_len ? corr[idx] : (T)0;
        nbv[t] = v;
        if (neigh) neigh[(size_t)sn * k + t] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) cf[threadIdx.x] = fit_coef<T>(lsq + threadIdx.x * k, nbv, k);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (maxes) {
            maxes[sn].data = best;
            maxes[sn].index = bi;
        }
        double px, py, ox, oy;
        fit_peak(cf, px, py);
        finalize_offset(bi, cross, policy, fit, px, py, ox, oy);
        if (offsets) {
            offsets[2 * (size_t)sn] = ox;
            offsets[2 * (size_t)sn + 1] = oy;
        }
        if (coefs)
            for (int c = 0; c < 6; ++c) coefs[6 * (size_t)sn + c] = cf[c];
    }
}

template <class T, int NT>
__global__ __launch_bounds__(NT) void maxarg_kernel(const T* __restrict__ data, uint32_t area,
                                                    data_index<T>* __restrict__ out, uint32_t total) {
    __shared__ T sv[NT / kWave];
    __shared__ uint32_t si[NT / kWave];
    const uint32_t sn = blockIdx.x;
    if (sn >= total) return;
    const T* d = data + (size_t)sn * area;
    T best = lowest<T>();
    uint32_t bi = 0xffffffffu;
    for (uint32_t i = threadIdx.x; i < area; i += NT) {
        const T v = d[i];
        if (v > best) {
            best = v;
            bi = i;
        }
    }
    block_argmax<T, NT>(best, bi, sv, si);
    if (threadIdx.x == 0) {
        out[sn].data = best;
        out[sn].index = bi;
    }
}

template <class T>
__global__ void extract_kernel(const T* __restrict__ data, size_t data_len, const data_index<T>* __restrict__ max_i,
                               T* __restrict__ neigh, int s, size2 size, int policy, uint32_t total) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = s * s;
    const size_t sn = g / k;
    if (sn >= total) return;
    const int t = (int)(g % k), i = t % s, j = t / s, r = (s - 1) / 2;
    const size2 mp = shift_pos(policy, max_i[sn].index, size);
    const size2 c = clamp_pos((int)mp.x - r + i, (int)mp.y - r + j, size);
    const size_t idx = pos_index(policy, (uint32_t)sn, c, size);
    neigh[g] = idx < data_len ? data[idx] : (T)0;
}

template <class T>
__global__ void subpixel_kernel(const T* __restrict__ neigh, int s, uint32_t n, const double* __restrict__ lsq,
                                const data_index<T>* __restrict__ maxes, size2 cross, int policy,
                                double* __restrict__ offsets, double* __restrict__ coefs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = s * s;
    double c[6];
    for (int row = 0; row < 6; ++row) c[row] = fit_coef<T>(lsq + row * k, neigh + (size_t)i * k, k);
    double px, py;
    fit_peak(c, px, py);
    if (maxes) finalize_offset(maxes[i].index, cross, policy, s, px, py, px, py);
    if (offsets) {
        offsets[2 * (size_t)i] = px;
        offsets[2 * (size_t)i + 1] = py;
    }
    if (coefs)
        for (int row = 0; row < 6; ++row) coefs[6 * (size_t)i + row] = c[row];
}

// --------------------------------------------------------------- exact refinement
// Precision form: the matrix-core screen (float) locates the peak; the values the
// result depends on are then evaluated the way the brute kernels evaluate them (one
// accumulator, y outer, x inner, fma; src/cross_corr.cu:36-61) from the T-precision
// prepared slices, so they equal the brute policy's correlation values bit for bit.
//   candidates: every shift whose screen value is within tau of the screen maximum,
//     tau = tol_rel ||p|| ||r|| + tol_abs (||p|| + ||r||) bounding |screen - T value|
//     twice over (caller's bound), so the brute policy's arg-max is among them;
//   arg-max over the candidates' exact values, lowest index on ties;
//   the F x F neighbourhood (identity position policy and clamp, spilling into the next
//     slice as the brute path does) evaluated exactly; fit and finalize as peak_kernel.
// A slice with more than kRefineCap candidates (a flat, noise-dominated screen) or
// without a finite screen maximum is swept whole: every shift of the window is
// evaluated exactly, box by box, so its arg-max is the brute policy's as well.
constexpr int kRefineCap = 512;
constexpr int kSweepRows = 8;  // sweep boxes: kSweepRows x kBoxW shifts (one per thread)
constexpr int kRefineNT = 256;
constexpr int kRowList = 256;  // screens of at most this many rows use the row maxima

template <class T>
__device__ inline T exact_corr(const T* __restrict__ p, const T* __restrict__ r, size2 s, int sx, int sy) {
    const int x_end = min((int)s.x - sx, (int)s.x), y_end = min((int)s.y - sy, (int)s.y);
    const int x0 = max(-sx, 0);
    T sum = 0;
    for (int y = max(-sy, 0); y < y_end; ++y) {
        const T* pr = p + (size_t)(y + sy) * s.x + sx;
        const T* qr = r + (size_t)y * s.x;
        int x = x0;
        // the operands of 8 products are loaded before their 8 dependent fmas, so one
        // memory latency is paid per 8 terms; the summation order is unchanged
        for (; x + 8 <= x_end; x += 8) {
            T u[8], v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                u[i] = pr[x + i];
                v[i] = qr[x + i];
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) sum = fma(u[i], v[i], sum);
        }
        for (; x < x_end; ++x) sum = fma(pr[x], qr[x], sum);
    }
    return sum;
}

// Exact value of buffer element `g` of the launch (slice g / area, shift of g % area).
template <class T>
__device__ inline T exact_at(size_t g, size_t corr_len, const T* p, const T* r, size2 s, size2 cross, uint32_t nb) {
    if (g >= corr_len) return T(0);
    const uint32_t area = cross.area();
    const uint32_t sn = (uint32_t)(g / area), pos = (uint32_t)(g % area);
    const int sx = (int)(pos % cross.x) - (int)(cross.x - 1) / 2;
    const int sy = (int)(pos / cross.x) - (int)(cross.y - 1) / 2;
    const size_t a = (size_t)s.x * s.y;
    return exact_corr<T>(p + (size_t)sn * a, r + (size_t)(sn % nb) * a, s, sx, sy);
}

// Exact values of a box of shifts of one slice, all threads of the workgroup
// together: thread t < bw * bh takes position (x0 + t % bw, y0 + t / bw) of the cross
// buffer (shift = position - (cross - 1) / 2).  The lanes of a wave walk the union of the
// box's overlaps in lockstep (y outer, x inner): chunks of kBoxY rows of the pattern
// slice are staged in LDS with zeros wherever a shift leaves the slice (row stride
// below), every lane takes the same x at each step, and the reference value of that
// step is the same for every lane, read through the scalar unit; LDS serves one
// operand per fma.  The terms a lane adds outside its own overlap are exact zeros
// (fma(0, r, acc) == acc for finite values, and the accumulator is never -0), so each
// value equals exact_corr's bit for bit.  LDS: kBoxY + kBoxRows - 1 rows of
// box_stride(s.x + kBoxW - 1), which also holds the box values afterwards; boxes of at
// most 14 rows keep it at 23 KB for 128-wide double slices.
constexpr int kBoxRows = 14;
constexpr int kBoxW = 32;
constexpr int kBoxY = 4;
static_assert(kSweepRows * kBoxW <= kRefineNT && kSweepRows <= kBoxRows, "sweep box exceeds a box evaluation");

// Row stride of the allocation (elements): the smallest stride >= sx congruent to 9 mod
// 32.  A box whose conflict-free stride does not fit it falls back to this stride, on
// which consecutive box rows start 18 banks apart (at most 2-way conflicts).
__host__ __device__ constexpr int box_stride(int sx) { return sx + ((9 - sx % 32) + 32) % 32; }

template <class T, int NT>
__device__ inline void eval_box(const T* __restrict__ p, const T* __restrict__ r, size2 s, size2 cross, int x0,
                                int y0, int bw, int bh, T* prow, T* out) {
    const int tid = threadIdx.x, S = (int)s.x, SY = (int)s.y;
    const int rx = (int)(cross.x - 1) / 2, ry = (int)(cross.y - 1) / 2;
    const int sx_lo = x0 - rx, sy_lo = y0 - ry;  // shift range of the box
    const int sx_hi = sx_lo + bw - 1, sy_hi = sy_lo + bh - 1;
    const int xlo = max(0, -sx_hi), xhi = min(S, S - sx_lo);  // union of the overlaps
    const int ylo = max(0, -sy_hi), yhi = min(SY, SY - sy_lo);
    const int cw = xhi - xlo + bw - 1, c0 = xlo + sx_lo;  // staged columns, first one
    // a row stride congruent to bw modulo the 32 banks of 4 bytes makes the addresses of
    // a step, row * PS + column for position t = row * bw + column, consecutive modulo
    // the banks, so a step's reads are free of bank conflicts (refining 128x128 double
    // slices, the fallback stride cost 4.9 conflict cycles per LDS instruction, this 0.02)
    constexpr int kSweep = 128 / (int)sizeof(T);
    const int PSc = cw + ((bw - cw) % kSweep + kSweep) % kSweep;
    const int PS = PSc <= box_stride(S + kBoxW - 1) ? PSc : box_stride(cw);
    const int n = bw * bh;
    const bool wave_on = tid - tid % kWave < n;  // waves with no position skip the sums
    const int t = min(tid, n - 1), lx = t % bw, ly = t / bw;
    T acc = 0;
    for (int yc = ylo; yc < yhi; yc += kBoxY) {
        const int ny = min(kBoxY, yhi - yc);
        __syncthreads();
        for (int e = tid; e < (ny + bh - 1) * cw; e += NT) {
            const int j = e / cw, cc = e % cw, c = cc + c0, Y = yc + sy_lo + j;
            prow[j * PS + cc] = (Y >= 0 && Y < SY && c >= 0 && c < S) ? p[(size_t)Y * S + c] : T(0);
        }
        __syncthreads();
        if (!wave_on) continue;
        for (int k = 0; k < ny; ++k) {
            const T* pr = prow + (k + ly) * PS + lx - xlo;  // pr[x] = pattern (y + sy, x + sx)
            // the reference row is the same for every lane: read through the scalar unit
            // (before, a v_readlane pair per product broadcast it from the lane holding it)
            const T* rr = r + (size_t)(yc + k) * S;
            int x = xlo;
            // the operands of 8 products are loaded before their 8 dependent fmas
            for (; x + 8 <= xhi; x += 8) {
                T u[8], q[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    u[i] = pr[x + i];
                    q[i] = rr[x + i];
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) acc = fma(u[i], q[i], acc);
            }
            for (; x < xhi; ++x) acc = fma(pr[x], rr[x], acc);
        }
    }
    __syncthreads();  // out may alias the staged rows
    if (tid < n) out[tid] = acc;
    __syncthreads();
}

template <class T, int NT>
__global__ __launch_bounds__(NT) void refine_kernel(const float* __restrict__ screen, size_t corr_len,
                                                    const T* __restrict__ p, const T* __restrict__ r, size2 s,
                                                    size2 cross, uint32_t nb, double tol_rel, double tol_abs,
                                                    int fit, const double* __restrict__ lsq,
                                                    data_index<T>* __restrict__ maxes, T* __restrict__ neigh,
                                                    double* __restrict__ offsets, double* __restrict__ coefs,
                                                    uint32_t total, const uint32_t* __restrict__ smax,
                                                    screen_lists lists) {
    __shared__ float svf[NT / kWave];
    __shared__ uint32_t si[NT / kWave];
    __shared__ double sd[NT / kWave];
    __shared__ T sv[NT / kWave];
    __shared__ uint32_t cand[kRefineCap];
    __shared__ uint32_t ncand, nrow, overflow;
    __shared__ uint16_t rlist[kRowList];  // rows of the screen that can hold candidates
    __shared__ T nbv[81];
    __shared__ double cf[6];
    __shared__ int bb[4];
    extern __shared__ __align__(16) unsigned char refine_lds[];
    T* prow = reinterpret_cast<T*>(refine_lds);  // staged pattern rows of eval_box
    T* boxv = prow;  // the box values eval_box leaves (over the staged rows)
    const uint32_t sn = blockIdx.x;
    if (sn >= total) return;
    const int tid = threadIdx.x;
    const uint32_t area = cross.area();
    const size_t a = (size_t)s.x * s.y;
    const float* sc = screen ? screen + (size_t)sn * area : nullptr;  // none with candidate lists
    const T* ps = p + (size_t)sn * a;
    const T* rs = r + (size_t)(sn % nb) * a;

    // screen maximum (NaN ignored) and its lowest index; 8 loads in flight per thread.
    // Given the maximum from the matrix-core kernel, the index is found by the
    // candidate pass below instead (one read of the screen instead of two).
    constexpr int U = 8;
    float m = -__builtin_huge_valf();
    uint32_t mi = 0xffffffffu;
    uint32_t mcode = 0;  // largest row-maximum code (0: no value)
    if (smax) {
        for (uint32_t r = tid; r < cross.y; r += NT) mcode = max(mcode, smax[(size_t)sn * cross.y + r]);
        mcode = ~block_min_u32<NT>(~mcode, si);
        if (mcode) m = screen_max_value(mcode);
    } else {
        for (uint32_t i0 = tid; i0 < area; i0 += NT * U) {
            float v[U];
#pragma unroll
            for (int k = 0; k < U; ++k) v[k] = i0 + k * NT < area ? sc[i0 + k * NT] : -__builtin_huge_valf();
#pragma unroll
            for (int k = 0; k < U; ++k)
                if (v[k] > m) {
                    m = v[k];
                    mi = i0 + k * NT;
                }
        }
        block_argmax<float, NT>(m, mi, svf, si);
    }

    // ||p||, ||r|| in double (given with the candidate lists)
    double tau;
    if (lists.cnt) {
        tau = refine_tau(lists, sn, sn % nb);
    } else {
    double pp = 0, rr = 0;
    for (size_t i0 = tid; i0 < a; i0 += NT * U) {
        double u[U], w[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            u[k] = i0 + k * NT < a ? (double)ps[i0 + k * NT] : 0.0;
            w[k] = i0 + k * NT < a ? (double)rs[i0 + k * NT] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            pp += u[k] * u[k];
            rr += w[k] * w[k];
        }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        pp += __shfl_xor(pp, off);
        rr += __shfl_xor(rr, off);
    }
    if (tid % kWave == 0) sd[tid / kWave] = pp;
    __syncthreads();
    pp = 0;
    for (int k = 0; k < NT / kWave; ++k) pp += sd[k];
    __syncthreads();
    if (tid % kWave == 0) sd[tid / kWave] = rr;
    __syncthreads();
    rr = 0;
    for (int k = 0; k < NT / kWave; ++k) rr += sd[k];
    const double np = sqrt(pp), nr = sqrt(rr);
    tau = tol_rel * np * nr + tol_abs * (np + nr);
    }

    // candidates.  Given the row maxima, only rows whose maximum reaches the threshold
    // (or equals the slice maximum) can hold one, and only those rows are read.
    if (tid == 0) ncand = nrow = overflow = 0;
    __syncthreads();
    const double thr = (double)m - tau;
    uint32_t mil = 0xffffffffu;  // lowest index holding m (given maximum)
    auto take = [&](uint32_t idx, float v) {
        if (idx < area && (double)v >= thr) {
            const uint32_t c = atomicAdd(&ncand, 1u);
            if (c < kRefineCap) cand[c] = idx;
        }
        if (smax && idx < area && v == m && idx < mil) mil = idx;
    };
    if (lists.cnt) {
        // the rows that can hold a candidate, then the positions they published (a row
        // with more than kListCap of them makes the slice sweep)
        for (uint32_t r = tid; r < cross.y; r += NT) {
            const uint32_t code = smax[(size_t)sn * cross.y + r];
            if (code && ((double)screen_max_value(code) >= thr || code == mcode)) {
                const uint32_t k = atomicAdd(&nrow, 1u);
                if (k < (uint32_t)kRowList) rlist[k] = r;
                else overflow = 1;  // more rows than the list holds: sweep
            }
        }
        __syncthreads();
        const uint32_t n = min(nrow, (uint32_t)kRowList) * (uint32_t)kListCap;
        for (uint32_t e = tid; e < n; e += NT) {
            const uint32_t row = rlist[e / kListCap], k = e % kListCap;
            const size_t rr = (size_t)sn * cross.y + row;
            const uint32_t c = lists.cnt[rr];
            if (c > (uint32_t)kListCap) {
                if (k == 0) overflow = 1;
            } else if (k < c) {
                const cand_entry en = lists.ent[rr * kListCap + k];
                take(row * cross.x + en.x, en.v);
            }
        }
    } else if (smax && cross.y <= (uint32_t)kRowList) {
        for (uint32_t r = tid; r < cross.y; r += NT) {
            const uint32_t code = smax[(size_t)sn * cross.y + r];
            if (code && ((double)screen_max_value(code) >= thr || code == mcode)) rlist[atomicAdd(&nrow, 1u)] = r;
        }
        __syncthreads();
        const uint32_t n = nrow * cross.x;
        for (uint32_t i0 = tid; i0 < n; i0 += NT * U) {
            uint32_t idx[U];
            float v[U];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                const uint32_t e = i0 + k * NT;
                idx[k] = e < n ? (uint32_t)rlist[e / cross.x] * cross.x + e % cross.x : area;
                v[k] = e < n ? sc[idx[k]] : -__builtin_huge_valf();
            }
#pragma unroll
            for (int k = 0; k < U; ++k) take(idx[k], v[k]);
        }
    } else {
        for (uint32_t i0 = tid; i0 < area; i0 += NT * U) {
            float v[U];
#pragma unroll
            for (int k = 0; k < U; ++k) v[k] = i0 + k * NT < area ? sc[i0 + k * NT] : -__builtin_huge_valf();
#pragma unroll
            for (int k = 0; k < U; ++k) take(i0 + k * NT, v[k]);
        }
    }
    if (smax) mi = block_min_u32<NT>(mil, si);
    __syncthreads();
    const bool sweep = ncand > (uint32_t)kRefineCap || mi == 0xffffffffu || overflow;
    const uint32_t nc = sweep ? 0u : ncand;
    // exact values; arg-max with the lowest index on ties.  Candidates whose bounding
    // box fits one box evaluation take the staged path, others one chain each.
    if (tid == 0) {
        bb[0] = bb[1] = 1 << 30;
        bb[2] = bb[3] = -1;
    }
    __syncthreads();
    for (uint32_t c = tid; c < nc; c += NT) {
        const int cx = (int)(cand[c] % cross.x), cy = (int)(cand[c] / cross.x);
        atomicMin(&bb[0], cx);
        atomicMin(&bb[1], cy);
        atomicMax(&bb[2], cx);
        atomicMax(&bb[3], cy);
    }
    __syncthreads();
    const int k = fit * fit, rad = (fit - 1) / 2;
    // One box evaluation when the candidates' box grown by the fit radius fits the
    // workgroup: it holds every candidate and the in-window neighbourhood of any of
    // them (the chains of one box run side by side; a second box would run after).
    const int cbw = bb[2] - bb[0] + 1, cbh = bb[3] - bb[1] + 1;
    const int ex0 = max(0, bb[0] - rad), ey0 = max(0, bb[1] - rad);
    const int ebw = min((int)cross.x - 1, bb[2] + rad) - ex0 + 1, ebh = min((int)cross.y - 1, bb[3] + rad) - ey0 + 1;
    const bool merged = !sweep && ebh <= kBoxRows && ebw <= kBoxW && ebw * ebh <= NT;
    const bool boxed = merged || (!sweep && cbh <= kBoxRows && cbw <= kBoxW && cbw * cbh <= NT);
    const int bx0 = merged ? ex0 : bb[0], by0 = merged ? ey0 : bb[1], bw = merged ? ebw : cbw;
    if (boxed) eval_box<T, NT>(ps, rs, s, cross, bx0, by0, bw, merged ? ebh : cbh, prow, boxv);
    T best = lowest<T>();
    uint32_t bi = 0xffffffffu;
    for (uint32_t c = tid; c < nc; c += NT) {
        const uint32_t i = cand[c];
        const int cx = (int)(i % cross.x), cy = (int)(i / cross.x);
        const T v = boxed ? boxv[(cy - by0) * bw + (cx - bx0)]
                          : exact_at<T>((size_t)sn * area + i, corr_len, p, r, s, cross, nb);
        if (v > best || (v == best && i < bi)) {
            best = v;
            bi = i;
        }
    }
    if (sweep) {
        // the whole window, kSweepRows x kBoxW shifts at a time (one per thread)
        for (int y0 = 0; y0 < (int)cross.y; y0 += kSweepRows)
            for (int x0 = 0; x0 < (int)cross.x; x0 += kBoxW) {
                const int w = min(kBoxW, (int)cross.x - x0), h = min(kSweepRows, (int)cross.y - y0);
                eval_box<T, NT>(ps, rs, s, cross, x0, y0, w, h, prow, boxv);
                if (tid < w * h) {
                    const uint32_t i = (uint32_t)(y0 + tid / w) * cross.x + (uint32_t)(x0 + tid % w);
                    const T v = boxv[tid];
                    if (v > best || (v == best && i < bi)) {
                        best = v;
                        bi = i;
                    }
                }
                __syncthreads();  // the next box stages over these values
            }
    }
    block_argmax<T, NT>(best, bi, sv, si);

    // neighbourhood, exact: the in-window part from the merged box or as its own box,
    // positions the clamp puts one past the window (spilling into the next row or
    // slice) one chain each
    const size2 mp = shift_pos(POLICY_BRUTE, bi, cross);
    int nx0 = bx0, ny0 = by0, nbw = bw;
    if (!merged) {
        nx0 = max(0, (int)mp.x - rad);
        ny0 = max(0, (int)mp.y - rad);
        nbw = min((int)cross.x - 1, (int)mp.x + rad) - nx0 + 1;
        const int nbh = min((int)cross.y - 1, (int)mp.y + rad) - ny0 + 1;
        eval_box<T, NT>(ps, rs, s, cross, nx0, ny0, nbw, nbh, prow, boxv);
    }
    for (int t = tid; t < k; t += NT) {
        const int i = t % fit, j = t / fit;
        const size2 c = clamp_pos((int)mp.x - rad + i, (int)mp.y - rad + j, cross);
        const T v = (c.x < cross.x && c.y < cross.y)
                        ? boxv[((int)c.y - ny0) * nbw + ((int)c.x - nx0)]
                        : exact_at<T>(pos_index(POLICY_BRUTE, sn, c, cross), corr_len, p, r, s, cross, nb);
        nbv[t] = v;
        if (neigh) neigh[(size_t)sn * k + t] = v;
    }
    __syncthreads();
    if (tid < 6) cf[tid] = fit_coef<T>(lsq + tid * k, nbv, k);
    __syncthreads();
    if (tid == 0) {
        if (maxes) {
            maxes[sn].data = best;
            maxes[sn].index = bi;
        }
        double px, py, ox, oy;
        fit_peak(cf, px, py);
        finalize_offset(bi, cross, POLICY_BRUTE, fit, px, py, ox, oy);
        if (offsets) {
            offsets[2 * (size_t)sn] = ox;
            offsets[2 * (size_t)sn + 1] = oy;
        }
        if (coefs)
            for (int c = 0; c < 6; ++c) coefs[6 * (size_t)sn + c] = cf[c];
    }
}

// --------------------------------------------------------------- launchers
constexpr int kPeakNT = 256;

template <class T>
void launch_peak(const T* corr, size_t corr_len, uint32_t area, size2 cross, int policy, int fit,
                 data_index<T>* maxes, T* neigh, double* offsets, double* coefs, uint32_t total, hipStream_t st,
                 const uint32_t* row_max) {
    const double* lsq = lstsq_device_table() + lstsq_offset(fit);
    if (total == 0) return;
    peak_kernel<T, kPeakNT><<<total, kPeakNT, 0, st>>>(corr, corr_len, area, cross, policy, fit, lsq, maxes, neigh,
                                                      offsets, coefs, total, row_max);
    EMIDA_LAUNCHED("peak_kernel");
}

template <class T>
void launch_maxarg(const T* data, uint32_t area, data_index<T>* out, uint32_t total, hipStream_t st) {
    if (total == 0) return;
    maxarg_kernel<T, kPeakNT><<<total, kPeakNT, 0, st>>>(data, area, out, total);
    EMIDA_LAUNCHED("maxarg_kernel");
}

template <class T>
void launch_extract(const T* data, size_t data_len, const data_index<T>* max_i, T* neigh, int s, size2 size,
                    int policy, uint32_t total, hipStream_t st) {
    const size_t n = (size_t)total * s * s;
    if (n == 0) return;
    extract_kernel<T><<<(unsigned)div_up(n, (size_t)256), 256, 0, st>>>(data, data_len, max_i, neigh, s, size,
                                                                       policy, total);
    EMIDA_LAUNCHED("extract_kernel");
}

template <class T>
void launch_subpixel(const T* neigh, int s, uint32_t n, const data_index<T>* maxes, size2 cross, int policy,
                     double* offsets, double* coefs, hipStream_t st) {
    const double* lsq = lstsq_device_table() + lstsq_offset(s);
    if (n == 0) return;
    subpixel_kernel<T><<<div_up(n, 64u), 64, 0, st>>>(neigh, s, n, lsq, maxes, cross, policy, offsets, coefs);
    EMIDA_LAUNCHED("subpixel_kernel");
}

template <class T>
void launch_refine(const float* screen, size_t corr_len, const T* p, const T* r, size2 s, size2 cross, uint32_t nb,
                   double tol_rel, double tol_abs, int fit, data_index<T>* maxes, T* neigh, double* offsets,
                   double* coefs, uint32_t total, hipStream_t st, const uint32_t* screen_max,
                   const screen_lists* lists) {
    if (total == 0) return;
    if (s.x > 2 * (uint32_t)kWave) throw arg_error("exact refinement needs slices at most 128 wide");
    if (lists && (!screen_max || !lists->cnt || !lists->ent || !lists->pn2 || !lists->rn2))
        throw arg_error("exact refinement: candidate lists need the row maxima, counts, entries and norms");
    if (!lists && !screen) throw arg_error("exact refinement: no screen");
    screen_lists none{};
    const double* lsq = lstsq_device_table() + lstsq_offset(fit);
    const size_t lds = (size_t)(kBoxY + kBoxRows - 1) * box_stride((int)s.x + kBoxW - 1) * sizeof(T);
    refine_kernel<T, kRefineNT><<<total, kRefineNT, lds, st>>>(screen, corr_len, p, r, s, cross, nb, tol_rel, tol_abs,
                                                          fit, lsq, maxes, neigh, offsets, coefs, total,
                                                          screen_max, lists ? *lists : none);
    EMIDA_LAUNCHED("refine_kernel");
}
template void launch_refine<float>(const float*, size_t, const float*, const float*, size2, size2, uint32_t, double,
                                   double, int, data_index<float>*, float*, double*, double*, uint32_t,
                                   hipStream_t, const uint32_t*, const screen_lists*);
template void launch_refine<double>(const float*, size_t, const double*, const double*, size2, size2, uint32_t,
                                    double, double, int, data_index<double>*, double*, double*, double*, uint32_t,
                                    hipStream_t, const uint32_t*, const screen_lists*);

#define INST_PEAK(T)                                                                                          \
    template void launch_peak<T>(const T*, size_t, uint32_t, size2, int, int, data_index<T>*, T*, double*,     \
                                 double*, uint32_t, hipStream_t, const uint32_t*);                            \
    template void launch_maxarg<T>(const T*, uint32_t, data_index<T>*, uint32_t, hipStream_t);                \
    template void launch_extract<T>(const T*, size_t, const data_index<T>*, T*, int, size2, int, uint32_t,    \
                                    hipStream_t);                                                             \
    template void launch_subpixel<T>(const T*, int, uint32_t, const data_index<T>*, size2, int, double*,      \
                                     double*, hipStream_t);
INST_PEAK(float)
INST_PEAK(double)

}  // namespace emida_amd
