// Definition-based ("brute") 2-D cross-correlation.
//
// Reference: src/cross_corr.cu:11-74.  For every shift (sx, sy) of the result
// (res index = shift + (res_size - 1) / 2) the sum over the overlap of
// pic[y + sy][x + sx] * ref[y][x] is accumulated in one register, y outer and x
// inner, as fused multiply-adds (the reference's nvcc build contracts
// `sum += a * b`), so double and float results are bit-identical to that order.
//
// Work decomposition: one workgroup per (slice, 4-row band of shifts, 64-column
// band of shifts), one thread per shift; the reference and pattern slices are
// staged in LDS when both fit in 64 KiB, and the overlap loop runs over LDS.
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

template <class T>
void launch_cross_corr(const T* pics, const T* ref, T* res, size2 size, size2 res_size, uint32_t ref_slices,
                       uint32_t batch, hipStream_t st) {
    if (res_size.x > 2 * size.x + 1 || res_size.y > 2 * size.y + 1)
        throw arg_error("cross_corr: res_size must be at most 2*size+1 per axis");
    const uint32_t total = ref_slices * batch;
    if (total == 0 || res_size.area() == 0) return;
    if (total > 65535u) throw arg_error("cross_corr: more than 65535 slices in one launch");
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
