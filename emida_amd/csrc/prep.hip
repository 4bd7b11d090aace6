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
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t area = slice.area();
    const size_t sn = tid / area;
    if (sn >= total) return;
    const uint32_t st = (uint32_t)(tid % area);
    const uint32_t x = st % slice.x, y = st / slice.x;
    const size2 b0 = begins[sn % nb];
    T pixel = (T)pics[(sn / nb) * src.area() + (size_t)(b0.y + y) * src.x + b0.x + x];
    pixel -= (T)sums[sn] / (T)area;
    pixel = pixel * hx[x] * hy[y];
    slices[sn * out.area() + (size_t)y * out.x + x] = pixel;
}

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
    const size_t n = (size_t)slice.area() * total;
    if (n == 0) return;
    prepare_kernel<IN, T, S><<<(unsigned)div_up(n, (size_t)256), 256, 0, st>>>(pics, slices, hx, hy, sums, begins,
                                                                              src, slice, out, nb, total);
    EMIDA_LAUNCHED("prepare_kernel");
}

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
