// CPU oracle for the EBSD pattern-offset path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, on the CPU, the arithmetic of the
// reference implementation's hot path so that the HIP library can be checked
// against it. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it; the product library never links or calls it.
//
// Every function cites the reference file:line it follows.  Floating-point
// operation order is reproduced exactly:
//   * the reference's GPU kernels were built by nvcc, which contracts
//     `sum += a * b` into fma(a, b, sum); those places call std::fma here;
//   * the reference's CPU finalize (subpixel fit) was built for plain x86-64
//     (no FMA); this file is compiled with -ffp-contract=off so no contraction
//     happens anywhere else.
//
// Parity notes (what is pinned by what) are listed in the header comment of
// oracle/oracle.py and in tests/test_oracle.py.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct sz2 {
    uint32_t x, y;
    uint32_t area() const { return x * y; }
};

template <class T>
struct data_index {
    T data;
    uint32_t index;
};

enum { T_F32 = 0, T_F64 = 1, T_I32 = 2, T_U16 = 3 };
enum { POLICY_BRUTE = 0, POLICY_FFT = 1 };

constexpr double PI = 3.14159265358979323846;

// ---------------------------------------------------------------- Hann window
// src/host_helpers.hpp:29-37 (generate_hanning): evaluated in double, then cast.
template <class T>
void hann(T* out, size_t n) {
    for (size_t i = 0; i < n; i++)
        out[i] = (T)(0.5 * (1 - cos(2 * PI * i / (n - 1))));
}

// ---------------------------------------------------------------- slice sums
// src/sum.cu:109-158. The reference reduces with 1024-thread blocks; thread t of
// block b adds N=10 pixels i = nblk*1024*n + b*1024 + t (n = 0..9) in order, then
// 32-lane shuffle-down trees (sum.cu:65-98) reduce the 32 warps and the 32 warp
// results.  For more than one block per slice the reference adds block results
// with atomicAdd in an unspecified order; blocks are added in ascending order here.
// For uint16 input the sum type is uint32 (src/cufft_helpers.hpp:157-179) and the
// order does not matter.
template <class RES>
RES warp_tree(RES* v) {
    RES nv[32];
    for (int off = 16; off > 0; off /= 2) {
        for (int l = 0; l < 32; ++l) nv[l] = v[l] + (l + off < 32 ? v[l + off] : v[l]);
        for (int l = 0; l < 32; ++l) v[l] = nv[l];
    }
    return v[0];
}

template <class IN, class RES>
RES slice_sum(const IN* pic, sz2 src, sz2 begin, sz2 slice) {
    const uint32_t B = 1024, N = 10;
    const uint32_t area = slice.area();
    const uint32_t nblk = (area + B * N - 1) / (B * N);
    RES total = 0;
    std::vector<RES> v(B);
    for (uint32_t b = 0; b < nblk; ++b) {
        for (uint32_t t = 0; t < B; ++t) {
            RES val = 0;
            for (uint32_t n = 0; n < N; ++n) {
                uint32_t i = nblk * B * n + b * B + t;
                uint32_t px = i % slice.x, py = i / slice.x;
                if (px < slice.x && py < slice.y)
                    val += (RES)pic[(size_t)(begin.y + py) * src.x + begin.x + px];
            }
            v[t] = val;
        }
        RES w[32];
        for (int wi = 0; wi < 32; ++wi) w[wi] = warp_tree(&v[wi * 32]);
        RES blk = warp_tree(w);
        total = (nblk == 1) ? blk : total + blk;
    }
    return total;
}

template <class IN, class RES>
void sums(const IN* pics, const uint32_t* begins, sz2 src, sz2 slice, uint32_t nb, uint32_t batch,
          RES* out) {
    for (uint32_t p = 0; p < batch; ++p)
        for (uint32_t j = 0; j < nb; ++j)
            out[p * nb + j] = slice_sum<IN, RES>(pics + (size_t)p * src.area(), src,
                                                 sz2{begins[2 * j], begins[2 * j + 1]}, slice);
}

// ---------------------------------------------------------------- prepare
// src/prepare_pics.cu:11-42: out = ((OUT)px - (OUT)sum / area) * hx[x] * hy[y],
// written into a slice of out_size (zero padding is left untouched).
template <class IN, class OUT, class S>
void prepare(const IN* pics, OUT* slices, const OUT* hx, const OUT* hy, const S* sumv,
             const uint32_t* begins, sz2 src, sz2 slice, sz2 out_size, uint32_t nb, uint32_t batch) {
    for (uint32_t sn = 0; sn < nb * batch; ++sn) {
        uint32_t bn = sn % nb, pn = sn / nb;
        for (uint32_t y = 0; y < slice.y; ++y)
            for (uint32_t x = 0; x < slice.x; ++x) {
                size_t pi = (size_t)pn * src.area() + (size_t)(begins[2 * bn + 1] + y) * src.x +
                            begins[2 * bn] + x;
                OUT pixel = (OUT)pics[pi];
                pixel -= (OUT)sumv[sn] / (OUT)slice.area();
                pixel = pixel * hx[x] * hy[y];
                slices[(size_t)sn * out_size.area() + (size_t)y * out_size.x + x] = pixel;
            }
    }
}

// ---------------------------------------------------------------- brute correlation
// src/cross_corr.cu:11-66: res[sy+r.y][sx+r.x] = sum over the overlap of
// pic[y+sy][x+sx] * ref[y][x], one accumulator, y outer, x inner, fma-contracted.
template <class T>
inline T mac(T a, T b, T acc) { return std::fma(a, b, acc); }
template <>
inline int32_t mac<int32_t>(int32_t a, int32_t b, int32_t acc) {
    return (int32_t)((uint32_t)acc + (uint32_t)a * (uint32_t)b);
}

template <class T>
T corr_at(const T* pic, const T* ref, sz2 size, int sx, int sy) {
    int x_end = std::min((int)size.x - sx, (int)size.x);
    int y_end = std::min((int)size.y - sy, (int)size.y);
    T sum = 0;
    for (int y = std::max(-sy, 0); y < y_end; ++y)
        for (int x = std::max(-sx, 0); x < x_end; ++x)
            sum = mac<T>(pic[(size_t)(y + sy) * size.x + (x + sx)], ref[(size_t)y * size.x + x], sum);
    return sum;
}

template <class T>
void cross_corr(const T* pics, const T* ref, T* res, sz2 size, sz2 res_size, uint32_t ref_slices,
                uint32_t batch) {
    sz2 r{(res_size.x - 1) / 2, (res_size.y - 1) / 2};
    const uint32_t total = ref_slices * batch;
#pragma omp parallel for schedule(dynamic) if (total > 1)
    for (int64_t sn = 0; sn < (int64_t)total; ++sn) {
        const T* p = pics + (size_t)sn * size.area();
        const T* q = ref + (size_t)(sn % ref_slices) * size.area();
        T* o = res + (size_t)sn * res_size.area();
        for (uint32_t y = 0; y < res_size.y; ++y)
            for (uint32_t x = 0; x < res_size.x; ++x)
                o[(size_t)y * res_size.x + x] = corr_at<T>(p, q, size, (int)x - (int)r.x, (int)y - (int)r.y);
    }
}

// ---------------------------------------------------------------- FFT-policy correlation
// src/gpu_offset.hpp:303-316 with src/hadamard.cu:39-40: unnormalised
// C2R(R2C(pic) * conj(R2C(ref))) over a (2s.x)x(2s.y) zero-padded grid equals
// (4 s.x s.y) times the circular correlation; with 2s padding the circular index q
// holds the linear shift q (q <= s) or q - 2s (q > s), and shift s has no overlap.
// The product is formed in double (fma, y outer, x inner) and rounded to T once.
template <class T>
void fft_corr(const T* pics, const T* ref, T* res, sz2 size, uint32_t ref_slices, uint32_t batch) {
    const sz2 m{2 * size.x, 2 * size.y};
    const double scale = (double)m.x * (double)m.y;
    const uint32_t total = ref_slices * batch;
    std::vector<double> pd, rd;
#pragma omp parallel for schedule(dynamic) if (total > 1) private(pd, rd)
    for (int64_t sn = 0; sn < (int64_t)total; ++sn) {
        pd.assign(pics + (size_t)sn * size.area(), pics + (size_t)(sn + 1) * size.area());
        const T* q = ref + (size_t)(sn % ref_slices) * size.area();
        rd.assign(q, q + size.area());
        T* o = res + (size_t)sn * m.area();
        for (uint32_t qy = 0; qy < m.y; ++qy) {
            int sy = qy <= size.y ? (int)qy : (int)qy - (int)m.y;
            for (uint32_t qx = 0; qx < m.x; ++qx) {
                int sx = qx <= size.x ? (int)qx : (int)qx - (int)m.x;
                double c = corr_at<double>(pd.data(), rd.data(), size, sx, sy);
                o[(size_t)qy * m.x + qx] = (T)(c * scale);
            }
        }
    }
}

// ---------------------------------------------------------------- arg-max
// src/maxarg.cu:119-177: per slice, the largest value with its index in the
// slice buffer.  Ties resolve to the lowest index (the reference's choice
// depends on its thread layout).
template <class T>
void maxarg(const T* data, uint32_t area, uint32_t batch, data_index<T>* out) {
    for (uint32_t b = 0; b < batch; ++b) {
        const T* d = data + (size_t)b * area;
        data_index<T> best{d[0], 0};
        for (uint32_t i = 1; i < area; ++i)
            if (d[i] > best.data) best = {d[i], i};
        out[b] = best;
    }
}

// ---------------------------------------------------------------- position policies
// src/kernels.cuh:22-53.
inline sz2 shift_pos(int policy, uint32_t index, sz2 size) {
    if (policy == POLICY_BRUTE) return {index % size.x, index / size.x};
    sz2 in{size.x + 1, size.y + 1};
    sz2 pos{index % in.x, index / in.x};
    return {(pos.x + (in.x - 1) / 2) % in.x, (pos.y + (in.y - 1) / 2) % in.y};
}

inline size_t pos_index(int policy, uint32_t pic, sz2 p, sz2 size) {
    if (policy == POLICY_BRUTE) return (size_t)pic * size.area() + (size_t)p.y * size.x + p.x;
    sz2 in{size.x + 1, size.y + 1};
    sz2 q{(p.x + (in.x / 2 + 1)) % in.x, (p.y + (in.y / 2 + 1)) % in.y};
    return (size_t)pic * in.area() + (size_t)q.y * in.x + q.x;
}

// ---------------------------------------------------------------- neighbourhood gather
// src/extract_neighbors.cu:13-41: positions are clamped to [0, size] (one past
// the last row/column), so for the identity policy a clamped read can land in
// the next row or the next slice.  `data_len` bounds the buffer; reads past it
// (possible only for the last slice) return 0.
template <class T>
void extract(const T* data, size_t data_len, const data_index<T>* max_i, T* neigh, int s, sz2 size,
             uint32_t batch, int policy) {
    int r = (s - 1) / 2;
    for (uint32_t idx = 0; idx < batch; ++idx) {
        sz2 mp = shift_pos(policy, max_i[idx].index, size);
        for (int i = 0; i < s; ++i)
            for (int j = 0; j < s; ++j) {
                int fx = (int)mp.x - r + i, fy = (int)mp.y - r + j;
                uint32_t cx = std::min<uint32_t>((uint32_t)std::max(fx, 0), size.x);
                uint32_t cy = std::min<uint32_t>((uint32_t)std::max(fy, 0), size.y);
                size_t k = pos_index(policy, idx, sz2{cx, cy}, size);
                neigh[(size_t)idx * s * s + j * s + i] = k < data_len ? data[k] : (T)0;
            }
    }
}

// ---------------------------------------------------------------- least-squares table
// src/subpixel_max.hpp:31-38 and src/subpixel_max.cpp:5-58 hold (A^T A)^-1 A^T for
// f(x,y) = a + bx + cy + dx^2 + exy + fy^2 sampled on an s x s grid, as exact
// rationals.  It is derived here by exact rational Gauss-Jordan elimination and
// rounded once per entry (numerator / denominator in double).
struct rat {
    __int128 n, d;
};
__int128 gcd128(__int128 a, __int128 b) {
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) {
        __int128 t = a % b;
        a = b;
        b = t;
    }
    return a;
}
rat mk(__int128 n, __int128 d) {
    if (d < 0) n = -n, d = -d;
    __int128 g = gcd128(n, d);
    if (g == 0) g = 1;
    return {n / g, d / g};
}
rat sub(rat a, rat b) { return mk(a.n * b.d - b.n * a.d, a.d * b.d); }
rat mul(rat a, rat b) { return mk(a.n * b.n, a.d * b.d); }
rat dvd(rat a, rat b) { return mk(a.n * b.d, a.d * b.n); }

void lstsq_matrix(int s, double* out) {
    const int m = s * s;
    std::vector<std::array<long long, 6>> A(m);
    for (int y = 0; y < s; ++y)
        for (int x = 0; x < s; ++x) A[y * s + x] = {1, x, y, (long long)x * x, (long long)x * y, (long long)y * y};
    // G = A^T A (6x6), augmented with A^T (6 x m)
    std::vector<std::vector<rat>> G(6, std::vector<rat>(6 + m));
    for (int i = 0; i < 6; ++i) {
        for (int j = 0; j < 6; ++j) {
            long long acc = 0;
            for (int k = 0; k < m; ++k) acc += A[k][i] * A[k][j];
            G[i][j] = mk(acc, 1);
        }
        for (int k = 0; k < m; ++k) G[i][6 + k] = mk(A[k][i], 1);
    }
    for (int c = 0; c < 6; ++c) {
        int piv = c;
        while (G[piv][c].n == 0) ++piv;
        std::swap(G[piv], G[c]);
        rat p = G[c][c];
        for (auto& e : G[c]) e = dvd(e, p);
        for (int r = 0; r < 6; ++r) {
            if (r == c || G[r][c].n == 0) continue;
            rat f = G[r][c];
            for (int k = 0; k < 6 + m; ++k) G[r][k] = sub(G[r][k], mul(f, G[c][k]));
        }
    }
    for (int i = 0; i < 6; ++i)
        for (int k = 0; k < m; ++k) out[i * m + k] = (double)(long long)G[i][6 + k].n / (double)(long long)G[i][6 + k].d;
}

// ---------------------------------------------------------------- subpixel fit
// src/subpixel_max.hpp:14-29 (multiply), :42-67 (coefs), :76-97 (peak): no FMA.
template <class T>
void subpixel(const T* neigh, int s, size_t n, double* offsets, double* coefs) {
    std::vector<double> M(6 * s * s);
    lstsq_matrix(s, M.data());
    const int k = s * s;
    for (size_t i = 0; i < n; ++i, neigh += k) {
        double c[6];
        for (int y = 0; y < 6; ++y) {
            double sum = 0;
            for (int j = 0; j < k; ++j) sum += (double)(M[y * k + j] * neigh[j]);
            c[y] = sum;
        }
        double a = c[0], b = c[1], cc = c[2], d = c[3], e = c[4], f = c[5];
        (void)a;
        double oy = (b * e - 2 * cc * d) / (4 * f * d - e * e);
        double ox = (-e * oy - b) / (2 * d);
        offsets[2 * i] = ox;
        offsets[2 * i + 1] = oy;
        for (int y = 0; y < 6; ++y) coefs[6 * i + y] = c[y];
    }
}

// ---------------------------------------------------------------- whole path
// src/gpu_offset.hpp:100-125 (sizes), :127-211 (reference preparation),
// :234-278 (one batch), :280-301 (finalize).
// peak_stats (optional, 3 per slice): the largest correlation value, the second
// largest (equal to it on a tie) and the largest magnitude - how well the integer
// peak is separated, for the tests' and the benchmark's conditioning criterion.
template <class T, class IN, class S>
int get_offset(int policy, const IN* pics, const IN* temp, sz2 src, const uint32_t* begins, uint32_t nb,
               sz2 slice, sz2 cross, uint32_t batch, int fit, double* offsets, double* coefs,
               data_index<T>* maxes_out, T* neigh_out, double* peak_stats) {
    if (fit != 3 && fit != 5 && fit != 7 && fit != 9) return -1;
    const uint32_t total = nb * batch;
    if (policy == POLICY_FFT) cross = sz2{2 * slice.x - 1, 2 * slice.y - 1};
    std::vector<T> hx(slice.x), hy(slice.y);
    hann<T>(hx.data(), slice.x);
    hann<T>(hy.data(), slice.y);
    std::vector<S> sref(nb), spic(total);
    sums<IN, S>(temp, begins, src, slice, nb, 1, sref.data());
    sums<IN, S>(pics, begins, src, slice, nb, batch, spic.data());
    std::vector<T> ref((size_t)nb * slice.area()), pic((size_t)total * slice.area());
    prepare<IN, T, S>(temp, ref.data(), hx.data(), hy.data(), sref.data(), begins, src, slice, slice, nb, 1);
    prepare<IN, T, S>(pics, pic.data(), hx.data(), hy.data(), spic.data(), begins, src, slice, slice, nb, batch);
    sz2 buf = policy == POLICY_FFT ? sz2{cross.x + 1, cross.y + 1} : cross;
    std::vector<T> corr((size_t)total * buf.area());
    if (policy == POLICY_FFT)
        fft_corr<T>(pic.data(), ref.data(), corr.data(), slice, nb, batch);
    else
        cross_corr<T>(pic.data(), ref.data(), corr.data(), slice, cross, nb, batch);
    std::vector<data_index<T>> mx(total);
    maxarg<T>(corr.data(), buf.area(), total, mx.data());
    if (peak_stats)
        for (uint32_t i = 0; i < total; ++i) {
            const T* d = corr.data() + (size_t)i * buf.area();
            double m1 = -HUGE_VAL, m2 = -HUGE_VAL, amax = 0;
            for (uint32_t k = 0; k < buf.area(); ++k) {
                const double v = (double)d[k];
                if (v > m1) {
                    m2 = m1;
                    m1 = v;
                } else if (v > m2) {
                    m2 = v;
                }
                amax = std::max(amax, std::fabs(v));
            }
            peak_stats[3 * (size_t)i] = m1;
            peak_stats[3 * (size_t)i + 1] = m2;
            peak_stats[3 * (size_t)i + 2] = amax;
        }
    std::vector<T> nbh((size_t)total * fit * fit);
    extract<T>(corr.data(), corr.size(), mx.data(), nbh.data(), fit, cross, total, policy);
    std::vector<double> sub(2 * total);
    subpixel<T>(nbh.data(), fit, total, sub.data(), coefs);
    const int r = (fit - 1) / 2;
    for (uint32_t i = 0; i < total; ++i) {
        sz2 mp = shift_pos(policy, mx[i].index, cross);
        offsets[2 * i] = -((int)mp.x - ((int)cross.x / 2) - r + sub[2 * i]);
        offsets[2 * i + 1] = -((int)mp.y - ((int)cross.y / 2) - r + sub[2 * i + 1]);
    }
    if (maxes_out) std::copy(mx.begin(), mx.end(), maxes_out);
    if (neigh_out) std::copy(nbh.begin(), nbh.end(), neigh_out);
    return 0;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

int or_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int or_hann(int T, size_t n, void* out) {
    if (T == T_F32) hann<float>((float*)out, n);
    else if (T == T_F64) hann<double>((double*)out, n);
    else return -1;
    return 0;
}

// IN: T_U16 -> uint32 sums, T_F32 -> float sums, T_F64 -> double sums.
int or_sums(int IN, const void* pics, const uint32_t* begins, uint32_t srcx, uint32_t srcy, uint32_t slx,
            uint32_t sly, uint32_t nb, uint32_t batch, void* out) {
    sz2 src{srcx, srcy}, sl{slx, sly};
    if (IN == T_U16) sums<uint16_t, uint32_t>((const uint16_t*)pics, begins, src, sl, nb, batch, (uint32_t*)out);
    else if (IN == T_F32) sums<float, float>((const float*)pics, begins, src, sl, nb, batch, (float*)out);
    else if (IN == T_F64) sums<double, double>((const double*)pics, begins, src, sl, nb, batch, (double*)out);
    else return -1;
    return 0;
}

// S: sum element type (T_U16 means uint32 sums).
int or_prepare(int IN, int T, int S, const void* pics, void* out, const void* hx, const void* hy,
               const void* sumv, const uint32_t* begins, uint32_t srcx, uint32_t srcy, uint32_t slx,
               uint32_t sly, uint32_t ox, uint32_t oy, uint32_t nb, uint32_t batch) {
    sz2 src{srcx, srcy}, sl{slx, sly}, os{ox, oy};
#define PREP(INT, OUTT, ST)                                                                          \
    prepare<INT, OUTT, ST>((const INT*)pics, (OUTT*)out, (const OUTT*)hx, (const OUTT*)hy,           \
                           (const ST*)sumv, begins, src, sl, os, nb, batch)
    if (T == T_F64) {
        if (IN == T_U16 && S == T_U16) PREP(uint16_t, double, uint32_t);
        else if (IN == T_U16 && S == T_F64) PREP(uint16_t, double, double);
        else if (IN == T_F64 && S == T_F64) PREP(double, double, double);
        else if (IN == T_F64 && S == T_U16) PREP(double, double, uint32_t);
        else if (IN == T_F32 && S == T_F32) PREP(float, double, float);
        else return -1;
    } else if (T == T_F32) {
        if (IN == T_U16 && S == T_U16) PREP(uint16_t, float, uint32_t);
        else if (IN == T_U16 && S == T_F32) PREP(uint16_t, float, float);
        else if (IN == T_F32 && S == T_F32) PREP(float, float, float);
        else if (IN == T_F64 && S == T_F64) PREP(double, float, double);
        else return -1;
    } else
        return -1;
#undef PREP
    return 0;
}

int or_cross_corr(int T, const void* pics, const void* ref, void* res, uint32_t sx, uint32_t sy, uint32_t rx,
                  uint32_t ry, uint32_t ref_slices, uint32_t batch) {
    sz2 s{sx, sy}, r{rx, ry};
    if (T == T_F64) cross_corr<double>((const double*)pics, (const double*)ref, (double*)res, s, r, ref_slices, batch);
    else if (T == T_F32) cross_corr<float>((const float*)pics, (const float*)ref, (float*)res, s, r, ref_slices, batch);
    else if (T == T_I32) cross_corr<int32_t>((const int32_t*)pics, (const int32_t*)ref, (int32_t*)res, s, r, ref_slices, batch);
    else return -1;
    return 0;
}

int or_fft_corr(int T, const void* pics, const void* ref, void* res, uint32_t sx, uint32_t sy, uint32_t ref_slices,
                uint32_t batch) {
    sz2 s{sx, sy};
    if (T == T_F64) fft_corr<double>((const double*)pics, (const double*)ref, (double*)res, s, ref_slices, batch);
    else if (T == T_F32) fft_corr<float>((const float*)pics, (const float*)ref, (float*)res, s, ref_slices, batch);
    else return -1;
    return 0;
}

int or_maxarg(int T, const void* data, uint32_t area, uint32_t batch, void* out) {
    if (T == T_F64) maxarg<double>((const double*)data, area, batch, (data_index<double>*)out);
    else if (T == T_F32) maxarg<float>((const float*)data, area, batch, (data_index<float>*)out);
    else return -1;
    return 0;
}

int or_extract(int T, int policy, const void* data, size_t data_len, const void* max_i, void* neigh, int s,
               uint32_t sx, uint32_t sy, uint32_t batch) {
    sz2 size{sx, sy};
    if (T == T_F64)
        extract<double>((const double*)data, data_len, (const data_index<double>*)max_i, (double*)neigh, s, size, batch, policy);
    else if (T == T_F32)
        extract<float>((const float*)data, data_len, (const data_index<float>*)max_i, (float*)neigh, s, size, batch, policy);
    else return -1;
    return 0;
}

int or_lstsq_matrix(int s, double* out) {
    if (s < 3 || s > 9 || s % 2 == 0) return -1;
    lstsq_matrix(s, out);
    return 0;
}

int or_subpixel(int T, const void* neigh, int s, size_t n, double* offsets, double* coefs) {
    if (s != 3 && s != 5 && s != 7 && s != 9) return -1;
    if (T == T_F64) subpixel<double>((const double*)neigh, s, n, offsets, coefs);
    else if (T == T_F32) subpixel<float>((const float*)neigh, s, n, offsets, coefs);
    else return -1;
    return 0;
}

int or_get_offset(int T, int IN, int policy, const void* pics, const void* temp, uint32_t srcx, uint32_t srcy,
                  const uint32_t* begins, uint32_t nb, uint32_t slx, uint32_t sly, uint32_t crx, uint32_t cry,
                  uint32_t batch, int fit, double* offsets, double* coefs, void* maxes_out, void* neigh_out,
                  double* peak_stats) {
    sz2 src{srcx, srcy}, sl{slx, sly}, cr{crx, cry};
#define GO(TT, INT, ST)                                                                                  \
    return get_offset<TT, INT, ST>(policy, (const INT*)pics, (const INT*)temp, src, begins, nb, sl, cr, \
                                   batch, fit, offsets, coefs, (data_index<TT>*)maxes_out, (TT*)neigh_out, \
                                   peak_stats)
    if (T == T_F64) {
        if (IN == T_U16) GO(double, uint16_t, uint32_t);
        if (IN == T_F64) GO(double, double, double);
        if (IN == T_F32) GO(double, float, float);
    } else if (T == T_F32) {
        if (IN == T_U16) GO(float, uint16_t, uint32_t);
        if (IN == T_F64) GO(float, double, double);
        if (IN == T_F32) GO(float, float, float);
    }
#undef GO
    return -1;
}

}  // extern "C"
