// Slice sums and slice preparation (mean subtraction, 2-D Hann window).
//
// Reference: src/sum.cu:109-167 (per-ROI sums read straight from the pattern),
// src/prepare_pics.cu:11-63 (cut, subtract mean, window), sum types from
// src/cufft_helpers.hpp:157-179 (uint32 for uint16 pixels).
#include "kernels.hpp"

namespace emida_amd {

// Sum of one ROI in the reference's reduction order: 1024-thread groups, thread t
// of group b adds pixels i = nblk*1024*n + b*1024 + t for n = 0..9, then 32-lane
// shuffle-down trees over each 32-thread group and over the 32 group results
// (src/sum.cu:65-98, 109-158).  Groups are added in ascending order when a ROI
// needs more than one (the reference adds them atomically in no fixed order).
// For uint32 sums the order does not change the result.
template <class IN, class S>
__global__ __launch_bounds__(1024) void sum_kernel(const IN* __restrict__ pics, S* __restrict__ sums,
                                                   const size2* __restrict__ begins, size2 src, size2 slice,
                                                   uint32_t nb, uint32_t total) {
#pragma clang fp contract(off)
    constexpr uint32_t B = 1024, N = 10;
    __shared__ S part[32];
    __shared__ S acc;
    const uint32_t sn = blockIdx.x;
    if (sn >= total) return;
    const size2 b0 = begins[sn % nb];
    const IN* pic = pics + (size_t)(sn / nb) * src.area() + (size_t)b0.y * src.x + b0.x;
    const uint32_t area = slice.area();
    const uint32_t nblk = div_up(area, B * N);
    const uint32_t t = threadIdx.x;
    for (uint32_t b = 0; b < nblk; ++b) {
        S v = 0;
        for (uint32_t n = 0; n < N; ++n) {
            uint32_t i = nblk * B * n + b * B + t;
            uint32_t px = i % slice.x, py = i / slice.x;
            if (py < slice.y) v += (S)pic[(size_t)py * src.x + px];
        }
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) v += __shfl_down(v, off, 32);
        if (t % 32 == 0) part[t / 32] = v;
        __syncthreads();
        if (t < 32) {
            v = part[t];
#pragma unroll
            for (int off = 16; off > 0; off >>= 1) v += __shfl_down(v, off, 32);
            if (t == 0) acc = (nblk == 1) ? v : ((b == 0 ? (S)0 : acc) + v);
        }
        __syncthreads();
    }
    if (t == 0) sums[sn] = acc;
}

// One thread per ROI pixel (src/prepare_pics.cu:11-42):
//   out = ((T)px - (T)sum / (T)area) * hx[x] * hy[y]
// evaluated left to right without contraction; written at (x, y) of an out_size
// slice (padding untouched).
template <class IN, class T, class S>
__global__ __launch_bounds__(256) void prepare_kernel(const IN* __restrict__ pics, T* __restrict__ slices,
                                                      const T* __restrict__ hx, const T* __restrict__ hy,
                                                      const S* __restrict__ sums, const size2* __restrict__ begins,
                                                      size2 src, size2 slice, size2 out, uint32_t nb,
                                                      uint32_t total) {
#pragma clang fp contract(off)
    // blocks_per_slice blocks per slice: the slice index is wave-uniform 32-bit
    // arithmetic (no per-pixel 64-bit division)
    const uint32_t area = slice.area(), bps = div_up(area, 256u);
    const uint32_t sn = blockIdx.x / bps;
    const uint32_t st = (blockIdx.x - sn * bps) * 256u + threadIdx.x;
    if (sn >= total || st >= area) return;
    const uint32_t x = st % slice.x, y = st / slice.x;
    const size2 b0 = begins[sn % nb];
    T pixel = (T)pics[(size_t)(sn / nb) * src.area() + (size_t)(b0.y + y) * src.x + b0.x + x];
    pixel -= (T)sums[sn] / (T)area;
    pixel = pixel * hx[x] * hy[y];
    slices[(size_t)sn * out.area() + (size_t)y * out.x + x] = pixel;
}

// uint16 slices in one pass: one workgroup per slice stages the slice's pixels in LDS
// and sums them in uint32 (exact, so the order does not matter: the sum equals
// sum_kernel's), then writes the prepared T slice with prepare_kernel's arithmetic and
// the hi/lo halves of the float prepared value with prepare_split_kernel's (each
// output optional).  The pixels are read from HBM once instead of two or three times.
template <class T, class HT>
__global__ __launch_bounds__(256) void sum_prepare_kernel(const uint16_t* __restrict__ pics,
                                                          uint32_t* __restrict__ sums, T* __restrict__ slices,
                                                          const T* __restrict__ hx, const T* __restrict__ hy,
                                                          HT* __restrict__ hi, HT* __restrict__ lo,
                                                          const float* __restrict__ hxf, const float* __restrict__ hyf,
                                                          float in_scale, double* __restrict__ stats,
                                                          const T* __restrict__ ref,
                                                          const size2* __restrict__ begins, size2 src, size2 slice,
                                                          uint32_t nb, uint32_t total) {
#pragma clang fp contract(off)
    constexpr uint32_t NT = 256;
    extern __shared__ __align__(16) uint16_t spx[];  // the slice's pixels, row-major
    __shared__ uint32_t part[NT / kWave];
    __shared__ double partd[2][NT / kWave];
    if (blockIdx.x >= total) return;
    // ROI-major order: consecutive workgroups share a ROI (and its reference slice)
    const uint32_t npat = total / nb;
    const uint32_t sn = (blockIdx.x % npat) * nb + blockIdx.x / npat;
    const uint32_t area = slice.area(), sx = slice.x, tid = threadIdx.x;
    const size2 b0 = begins[sn % nb];
    const uint16_t* pic = pics + (size_t)(sn / nb) * src.area() + (size_t)b0.y * src.x + b0.x;
    uint32_t s = 0;
    constexpr uint32_t U = 16;  // loads in flight per thread
    for (uint32_t e0 = tid; e0 < area; e0 += NT * U) {
        uint16_t v[U];
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) {
            const uint32_t e = e0 + k * NT;
            v[k] = e < area ? pic[(size_t)(e / sx) * src.x + e % sx] : (uint16_t)0;
        }
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) {
            if (e0 + k * NT < area) spx[e0 + k * NT] = v[k];
            s += v[k];
        }
    }
    s = wave_reduce(s, 0u, [](uint32_t x, uint32_t y) { return x + y; });
    if (tid % kWave == 0) part[tid / kWave] = s;
    __syncthreads();
    uint32_t sum = 0;
    for (uint32_t k = 0; k < NT / kWave; ++k) sum += part[k];
    if (tid == 0 && sums) sums[sn] = sum;
    // rows of a multiple of 4 pixels: 4 consecutive pixels of one row per thread and
    // step (one 8-byte LDS read, vector table loads and stores); others one pixel
    const bool quad = sx % 4 == 0;
    const uint32_t step = quad ? 4 : 1;
    if (slices) {
        T* out = slices + (size_t)sn * area;
        const T mean = (T)sum / (T)area;
        double nn = 0, zz = 0;  // sums of squares and of products with the reference
        const T* rs = stats ? ref + (size_t)(sn % nb) * area : nullptr;
        for (uint32_t e = tid * step; e < area; e += NT * step) {
            const uint32_t x = e % sx, y = e / sx;
            if (quad) {
                typedef T t4 __attribute__((ext_vector_type(4)));
                typedef uint16_t u4 __attribute__((ext_vector_type(4)));
                const u4 q = *reinterpret_cast<const u4*>(spx + e);
                const t4 wx = *reinterpret_cast<const t4*>(hx + x);
                const T wy = hy[y];
                t4 v;
                v.x = ((T)q.x - mean) * wx.x * wy;
                v.y = ((T)q.y - mean) * wx.y * wy;
                v.z = ((T)q.z - mean) * wx.z * wy;
                v.w = ((T)q.w - mean) * wx.w * wy;
                *reinterpret_cast<t4*>(out + e) = v;
                if (rs) {
                    const t4 r = *reinterpret_cast<const t4*>(rs + e);
                    nn += (double)v.x * (double)v.x + (double)v.y * (double)v.y + (double)v.z * (double)v.z +
                          (double)v.w * (double)v.w;
                    zz += (double)v.x * (double)r.x + (double)v.y * (double)r.y + (double)v.z * (double)r.z +
                          (double)v.w * (double)r.w;
                }
            } else {
                T v = (T)spx[e];
                v -= mean;
                v = v * hx[x] * hy[y];
                out[e] = v;
                if (rs) {
                    nn += (double)v * (double)v;
                    zz += (double)v * (double)rs[e];
                }
            }
        }
        if (stats) {
            for (int off = kWave / 2; off > 0; off >>= 1) {
                nn += __shfl_xor(nn, off);
                zz += __shfl_xor(zz, off);
            }
            if (tid % kWave == 0) {
                partd[0][tid / kWave] = nn;
                partd[1][tid / kWave] = zz;
            }
            __syncthreads();
            if (tid < 2) {
                double t = 0;
                for (uint32_t k = 0; k < NT / kWave; ++k) t += partd[tid][k];
                stats[tid * (size_t)total + sn] = t;
            }
        }
    }
    if (hi) {
        const size_t o = (size_t)sn * area;
        const float mean = (float)sum / (float)area;
        for (uint32_t e = tid * step; e < area; e += NT * step) {
            const uint32_t x = e % sx, y = e / sx;
            if (quad) {
                typedef float f4 __attribute__((ext_vector_type(4)));
                typedef HT h4 __attribute__((ext_vector_type(4)));
                typedef uint16_t u4 __attribute__((ext_vector_type(4)));
                const u4 q = *reinterpret_cast<const u4*>(spx + e);
                const f4 wx = *reinterpret_cast<const f4*>(hxf + x);
                const float wy = hyf[y];
                f4 v;
                v.x = ((float)q.x - mean) * wx.x * wy * in_scale;
                v.y = ((float)q.y - mean) * wx.y * wy * in_scale;
                v.z = ((float)q.z - mean) * wx.z * wy * in_scale;
                v.w = ((float)q.w - mean) * wx.w * wy * in_scale;
                h4 h, l;
                h.x = (HT)v.x;
                h.y = (HT)v.y;
                h.z = (HT)v.z;
                h.w = (HT)v.w;
                *reinterpret_cast<h4*>(hi + o + e) = h;
                if (lo) {
                    l.x = (HT)(v.x - (float)h.x);
                    l.y = (HT)(v.y - (float)h.y);
                    l.z = (HT)(v.z - (float)h.z);
                    l.w = (HT)(v.w - (float)h.w);
                    *reinterpret_cast<h4*>(lo + o + e) = l;
                }
            } else {
                float v = (float)spx[e];
                v -= mean;
                v = v * hxf[x] * hyf[y];
                v *= in_scale;
                const HT h = (HT)v;
                hi[o + e] = h;
                if (lo) lo[o + e] = (HT)(v - (float)h);
            }
        }
    }
}

template <class T, class HT>
bool launch_sum_prepare(const uint16_t* pics, uint32_t* sums, T* slices, const T* hx, const T* hy, HT* hi, HT* lo,
                        const float* hxf, const float* hyf, float in_scale, double* stats, const T* ref,
                        const size2* begins, size2 src, size2 slice, uint32_t nb, uint32_t total, hipStream_t st) {
    if (stats && (!slices || !ref)) throw arg_error("sum_prepare: statistics need the prepared and reference slices");
    if (nb == 0 || total % nb) throw arg_error("sum_prepare: slices must be whole sets of ROIs");
    const size_t lds = (size_t)slice.area() * sizeof(uint16_t);
    if (lds > kSumPrepareMaxLds) return false;
    // the four-pixel steps read window tables and write outputs as vectors
    auto aligned = [](const void* q, size_t a) { return q == nullptr || (uintptr_t)q % a == 0; };
    if (!aligned(slices, 4 * sizeof(T)) || !aligned(hx, 4 * sizeof(T)) || !aligned(ref, 4 * sizeof(T)) ||
        !aligned(hi, 4 * sizeof(HT)) ||
        !aligned(lo, 4 * sizeof(HT)) || !aligned(hxf, 4 * sizeof(float)))
        return false;
    if (total == 0) return true;
    sum_prepare_kernel<T, HT><<<total, 256, lds, st>>>(pics, sums, slices, hx, hy, hi, lo, hxf, hyf, in_scale, stats,
                                                       ref, begins, src, slice, nb, total);
    EMIDA_LAUNCHED("sum_prepare_kernel");
    return true;
}

// Sums of squares and (with ref) of products with the ROI's reference slice, one
// slice per workgroup, in double.
template <class T>
__global__ __launch_bounds__(256) void slice_stats_kernel(const T* __restrict__ slices, const T* __restrict__ ref,
                                                          uint32_t area, uint32_t nb, uint32_t n,
                                                          double* __restrict__ out) {
    constexpr uint32_t NT = 256, U = 8;  // loads in flight per thread
    __shared__ double part[2][NT / kWave];
    const uint32_t sn = blockIdx.x, tid = threadIdx.x;
    if (sn >= n) return;
    const T* q = slices + (size_t)sn * area;
    const T* rs = ref ? ref + (size_t)(sn % nb) * area : nullptr;
    double nn = 0, zz = 0;
    for (uint32_t e0 = tid; e0 < area; e0 += NT * U) {
        double v[U], w[U];
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) {
            const bool in = e0 + k * NT < area;
            v[k] = in ? (double)q[e0 + k * NT] : 0.0;
            w[k] = in && rs ? (double)rs[e0 + k * NT] : 0.0;
        }
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) {
            nn += v[k] * v[k];
            zz += v[k] * w[k];
        }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        nn += __shfl_xor(nn, off);
        zz += __shfl_xor(zz, off);
    }
    if (tid % kWave == 0) {
        part[0][tid / kWave] = nn;
        part[1][tid / kWave] = zz;
    }
    __syncthreads();
    if (tid < (ref ? 2u : 1u)) {
        double t = 0;
        for (uint32_t k = 0; k < NT / kWave; ++k) t += part[tid][k];
        out[tid * (size_t)n + sn] = t;
    }
}

template <class T>
void launch_slice_stats(const T* slices, const T* ref, uint32_t area, uint32_t nb, uint32_t n, double* out,
                        hipStream_t st) {
    if (n == 0) return;
    slice_stats_kernel<T><<<n, 256, 0, st>>>(slices, ref, area, nb, n, out);
    EMIDA_LAUNCHED("slice_stats_kernel");
}
template void launch_slice_stats<float>(const float*, const float*, uint32_t, uint32_t, uint32_t, double*,
                                        hipStream_t);
template void launch_slice_stats<double>(const double*, const double*, uint32_t, uint32_t, uint32_t, double*,
                                         hipStream_t);

template <class IN, class S>
void launch_sum(const IN* pics, S* sums, const size2* begins, size2 src, size2 slice, uint32_t nb, uint32_t total,
                hipStream_t st) {
    if (total == 0) return;
    sum_kernel<IN, S><<<total, 1024, 0, st>>>(pics, sums, begins, src, slice, nb, total);
    EMIDA_LAUNCHED("sum_kernel");
}

template <class IN, class T, class S>
void launch_prepare(const IN* pics, T* slices, const T* hx, const T* hy, const S* sums, const size2* begins,
                    size2 src, size2 slice, size2 out, uint32_t nb, uint32_t total, hipStream_t st) {
    const size_t blocks = (size_t)div_up(slice.area(), 256u) * total;
    if (blocks == 0) return;
    if (blocks > 0xffffffffu) throw arg_error("prepare: too many slices for one launch");
    prepare_kernel<IN, T, S><<<(unsigned)blocks, 256, 0, st>>>(pics, slices, hx, hy, sums, begins, src, slice, out,
                                                               nb, total);
    EMIDA_LAUNCHED("prepare_kernel");
}

#define INST_SUM_PREP(T, HT)                                                                              \
    template bool launch_sum_prepare<T, HT>(const uint16_t*, uint32_t*, T*, const T*, const T*, HT*, HT*,   \
                                            const float*, const float*, float, double*, const T*,              \
                                            const size2*, size2, size2, uint32_t, uint32_t, hipStream_t);
INST_SUM_PREP(double, _Float16)
INST_SUM_PREP(double, __bf16)
INST_SUM_PREP(float, _Float16)
INST_SUM_PREP(float, __bf16)

#define INST_SUM(IN, S)                                                                                  \
    template void launch_sum<IN, S>(const IN*, S*, const size2*, size2, size2, uint32_t, uint32_t, hipStream_t);
INST_SUM(uint16_t, uint32_t)
INST_SUM(float, float)
INST_SUM(double, double)

#define INST_PREP(IN, T, S)                                                                              \
    template void launch_prepare<IN, T, S>(const IN*, T*, const T*, const T*, const S*, const size2*, size2, \
                                           size2, size2, uint32_t, uint32_t, hipStream_t);
INST_PREP(uint16_t, double, uint32_t)
INST_PREP(uint16_t, float, uint32_t)
INST_PREP(uint16_t, double, double)
INST_PREP(uint16_t, float, float)
INST_PREP(double, double, double)
INST_PREP(double, double, uint32_t)
INST_PREP(double, float, double)
INST_PREP(float, float, float)
INST_PREP(float, double, float)

}  // namespace emida_amd
