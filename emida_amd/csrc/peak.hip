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
        double* d = nullptr;
        EMIDA_HIP(hipMalloc(&d, host.size() * sizeof(double)));
        EMIDA_HIP(hipMemcpy(d, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice));
        per_device[dev] = d;
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
                                                  uint32_t total) {
    __shared__ T sv[NT / kWave];
    __shared__ uint32_t si[NT / kWave];
    __shared__ T nbv[81];
    __shared__ double cf[6];
    const uint32_t sn = blockIdx.x;
    if (sn >= total) return;
    const T* d = corr + (size_t)sn * area;
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

// --------------------------------------------------------------- launchers
constexpr int kPeakNT = 256;

template <class T>
void launch_peak(const T* corr, size_t corr_len, uint32_t area, size2 cross, int policy, int fit,
                 data_index<T>* maxes, T* neigh, double* offsets, double* coefs, uint32_t total, hipStream_t st) {
    const double* lsq = lstsq_device_table() + lstsq_offset(fit);
    if (total == 0) return;
    peak_kernel<T, kPeakNT><<<total, kPeakNT, 0, st>>>(corr, corr_len, area, cross, policy, fit, lsq, maxes, neigh,
                                                      offsets, coefs, total);
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

#define INST_PEAK(T)                                                                                          \
    template void launch_peak<T>(const T*, size_t, uint32_t, size2, int, int, data_index<T>*, T*, double*,     \
                                 double*, uint32_t, hipStream_t);                                             \
    template void launch_maxarg<T>(const T*, uint32_t, data_index<T>*, uint32_t, hipStream_t);                \
    template void launch_extract<T>(const T*, size_t, const data_index<T>*, T*, int, size2, int, uint32_t,    \
                                    hipStream_t);                                                             \
    template void launch_subpixel<T>(const T*, int, uint32_t, const data_index<T>*, size2, int, double*,      \
                                     double*, hipStream_t);
INST_PEAK(float)
INST_PEAK(double)

}  // namespace emida_amd
