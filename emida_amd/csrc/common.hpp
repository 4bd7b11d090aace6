// Shared types and device helpers of the gfx950 pattern-offset kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace emida_amd {

using esize_t = uint32_t;

// size2_t (reference src/common.hpp:41-132): {x = columns, y = rows}.
struct size2 {
    uint32_t x, y;
    __host__ __device__ uint32_t area() const { return x * y; }
};

// data_index<T> (reference src/common.hpp:26-31): value and index of a maximum.
template <class T>
struct data_index {
    T data;
    uint32_t index;
};
static_assert(sizeof(data_index<float>) == 8, "data_index<float> layout");
static_assert(sizeof(data_index<double>) == 16, "data_index<double> layout");

enum policy_t { POLICY_BRUTE = 0, POLICY_FFT = 1 };

template <class T>
struct cplx {
    T re, im;
};

constexpr int kWave = 64;

template <class A, class B>
__host__ __device__ inline A div_up(A a, B b) {
    return (a + b - 1) / b;
}

// Thrown by launchers on bad arguments; the C ABI turns it into EMIDA_ERR_ARG.
struct arg_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct unsupported_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct hip_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw hip_error(std::string(what) + ": " + hipGetErrorString(e));
}
#define EMIDA_HIP(call) ::emida_amd::hip_check((call), #call)
#define EMIDA_LAUNCHED(name) ::emida_amd::hip_check(hipGetLastError(), name)

// XCD-aware workgroup order.  Workgroup b of a 1-D grid of n runs on XCD b mod 8;
// this returns a logical index such that XCD x executes the contiguous run
// [x*(n/8) + min(x, n%8), ...) of logical indices, so that neighbouring logical
// indices (which share input data) share one XCD's L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
    constexpr uint32_t X = 8;
    const uint32_t x = b % X, l = b / X, per = n / X, rem = n % X;
    return x * per + (x < rem ? x : rem) + l;
}

// Position policies of the correlation buffer (reference src/kernels.cuh:22-53).
// Identity: buffer index = y * size.x + x.  FFT: the circular (2s) buffer is
// addressed through the shift (p + s + 1) mod 2s.
__host__ __device__ inline size2 shift_pos(int policy, uint32_t index, size2 size) {
    if (policy == POLICY_BRUTE) return {index % size.x, index / size.x};
    const size2 in{size.x + 1, size.y + 1};
    const size2 pos{index % in.x, index / in.x};
    return {(pos.x + (in.x - 1) / 2) % in.x, (pos.y + (in.y - 1) / 2) % in.y};
}

__host__ __device__ inline size_t pos_index(int policy, uint32_t pic, size2 p, size2 size) {
    if (policy == POLICY_BRUTE) return (size_t)pic * size.area() + (size_t)p.y * size.x + p.x;
    const size2 in{size.x + 1, size.y + 1};
    const size2 q{(p.x + (in.x / 2 + 1)) % in.x, (p.y + (in.y / 2 + 1)) % in.y};
    return (size_t)pic * in.area() + (size_t)q.y * in.x + q.x;
}

// Clamp of the neighbourhood gather (reference src/extract_neighbors.cu:31-35):
// positions are limited to [0, size], one past the last row/column.
__host__ __device__ inline size2 clamp_pos(int fx, int fy, size2 size) {
    uint32_t cx = fx < 0 ? 0u : (uint32_t)fx;
    uint32_t cy = fy < 0 ? 0u : (uint32_t)fy;
    cx = cx < size.x ? cx : size.x;
    cy = cy < size.y ? cy : size.y;
    return {cx, cy};
}

// Arg-max combine: larger value wins, lower index breaks ties.
template <class T>
__device__ inline void argmax_combine(T& v, uint32_t& i, T ov, uint32_t oi) {
    if (ov > v || (ov == v && oi < i)) {
        v = ov;
        i = oi;
    }
}

// ------------------------------------------------------------------ wave reductions
// DPP lane move of a 32- or 64-bit value: lane i receives the value of the lane CTRL
// selects; lanes in rows outside ROW_MASK receive `fill`.
template <int CTRL, int ROW_MASK, class T>
__device__ __forceinline__ T dpp_move(T fill, T v) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, fill), __builtin_bit_cast(int, v),
                                                                 CTRL, ROW_MASK, 0xF, false));
    } else {
        static_assert(sizeof(T) == 8, "32- or 64-bit values");
        const uint64_t f = __builtin_bit_cast(uint64_t, fill), x = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)f, (int)(uint32_t)x, CTRL, ROW_MASK, 0xF, false);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)(f >> 32), (int)(uint32_t)(x >> 32), CTRL,
                                                                  ROW_MASK, 0xF, false);
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    }
}
template <class T>
__device__ __forceinline__ T lane63(T v) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
    } else {
        const uint64_t x = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, 63);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), 63);
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    }
}
// Reduction of a whole wave (all 64 lanes active) with an associative, commutative op
// and its identity: quad swaps [1,0,3,2] and [2,3,0,1], half-row and row mirrors (every
// lane of a 16-lane row then holds the row's result), row broadcasts 15 (rows 1, 3) and
// 31 (rows 2, 3); lane 63 holds the wave's result, returned as a wave-uniform value.
// Six DPP steps instead of six LDS permutes with their index arithmetic.
template <class T, class OP>
__device__ __forceinline__ T wave_reduce(T v, T ident, OP op) {
    v = op(v, dpp_move<0xB1, 0xF>(ident, v));
    v = op(v, dpp_move<0x4E, 0xF>(ident, v));
    v = op(v, dpp_move<0x141, 0xF>(ident, v));
    v = op(v, dpp_move<0x140, 0xF>(ident, v));
    v = op(v, dpp_move<0x142, 0xA>(ident, v));
    v = op(v, dpp_move<0x143, 0xC>(ident, v));
    return lane63(v);
}

template <class T>
__device__ inline void wave_argmax(T& v, uint32_t& i) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        T ov = __shfl_xor(v, off);
        uint32_t oi = __shfl_xor(i, off);
        argmax_combine(v, i, ov, oi);
    }
}

// Workgroup arg-max; every thread receives the result.  `sv`/`si` hold one
// entry per wave.
template <class T, int NT>
__device__ inline void block_argmax(T& v, uint32_t& i, T* sv, uint32_t* si) {
    wave_argmax(v, i);
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) {
        sv[w] = v;
        si[w] = i;
    }
    __syncthreads();
    v = sv[0];
    i = si[0];
#pragma unroll
    for (int k = 1; k < NT / kWave; ++k) argmax_combine(v, i, sv[k], si[k]);
    __syncthreads();
}

// Workgroup maximum (IEEE maxNum: NaN operands are ignored); every thread receives it.
// Every thread of the workgroup calls it (the wave steps read all 64 lanes).
template <class T, int NT>
__device__ inline T block_max(T v, T* sv) {
    v = wave_reduce(v, -(T)__builtin_huge_val(), [](T x, T y) { return fmax(x, y); });
    if (threadIdx.x % kWave == 0) sv[threadIdx.x / kWave] = v;
    __syncthreads();
    v = sv[0];
#pragma unroll
    for (int k = 1; k < NT / kWave; ++k) v = fmax(v, sv[k]);
    __syncthreads();
    return v;
}

// Workgroup minimum of unsigned values; every thread receives it (and calls it).
template <int NT>
__device__ inline uint32_t block_min_u32(uint32_t v, uint32_t* sv) {
    v = wave_reduce(v, 0xffffffffu, [](uint32_t x, uint32_t y) { return min(x, y); });
    if (threadIdx.x % kWave == 0) sv[threadIdx.x / kWave] = v;
    __syncthreads();
    v = sv[0];
#pragma unroll
    for (int k = 1; k < NT / kWave; ++k) v = min(v, sv[k]);
    __syncthreads();
    return v;
}

}  // namespace emida_amd
