// Deformation post-step: per pattern, the projective map (homography) between the
// ROI middles and their displaced positions, fitted algebraically, and its strain and
// rotation.  One workgroup per pattern.
//
// Reference (Python post-processing of the offsets):
//  - python/process_offsets.py:31-97 (and python/test.py:137-198): relative to the
//    pattern centre (cx, cy) and detector distance z, old = mid - c, new = old + offset;
//    rows  [-ox z, -oy z, -z^2, 0, 0, 0, ox nx, oy nx, z nx]
//          [0, 0, 0, -ox z, -oy z, -z^2, ox ny, oy ny, z ny]
//    per ROI; the 3 x 3 map is the right singular vector of the smallest singular value
//    (up to scale); the script prints it divided by its [2][2] element;
//  - jove/analyze.py:37-95: ROIs are used where the fitted quadratic peak is a
//    maximum with determinant a c - b^2 > 1e-6 (a = -coef[3], b = -coef[4] / 2,
//    c = -coef[5]); quality = median of that determinant over all ROIs; a pattern needs
//    more than 10 such ROIs; the map is scaled to determinant 1 (divided by the cube
//    root of its determinant); polar decomposition F = R U with strain log(U) and
//    rotation log(R) from Rodrigues' formula.
// jove/analyze.py:28 fills the second row's oy column with ox ny; this follows
// process_offsets.py (oy ny), the form the algebra requires.
//
// Here the normal matrix G = A^T A (9 x 9) is accumulated in double and its eigenvector
// of the smallest eigenvalue found by cyclic Jacobi rotations; for the well-conditioned
// systems of ROI grids (largest / second-smallest singular value of A about 3) this
// equals the SVD's singular vector to ~1e-15.  The polar decomposition uses the
// eigen-decomposition of F^T F (3 x 3 Jacobi).  The quality determinant is evaluated
// without contraction, as numpy does, and the median is found by rank selection, so
// quality and the ROI selection equal numpy's bit for bit.
#include <cmath>

#include "offset.hpp"

namespace emida_amd {
namespace {

constexpr int kPostNT = 256;
constexpr int kPostChunk = 512;      // ROIs staged per pass of the normal-matrix sum
constexpr uint32_t kPostMaxRois = 8192;

// Cyclic Jacobi eigen-decomposition of the symmetric N x N matrix a (row-major, in
// place: eigenvalues end on the diagonal); v receives the eigenvectors as columns.
template <int N>
__device__ void jacobi_eigen(double* a, double* v) {
    for (int i = 0; i < N * N; ++i) v[i] = (i / N == i % N) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0, diag = 0;
        for (int i = 0; i < N; ++i) {
            diag += a[i * N + i] * a[i * N + i];
            for (int j = i + 1; j < N; ++j) off += a[i * N + j] * a[i * N + j];
        }
        if (!(off > 1e-30 * diag)) break;  // converged (also stops on NaN)
        for (int p = 0; p < N - 1; ++p)
            for (int q = p + 1; q < N; ++q) {
                const double apq = a[p * N + q];
                if (apq == 0.0) continue;
                const double theta = (a[q * N + q] - a[p * N + p]) / (2.0 * apq);
                const double t = fabs(theta) > 1e150 ? 0.5 / theta
                                                     : (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < N; ++k) {  // columns p, q
                    const double kp = a[k * N + p], kq = a[k * N + q];
                    a[k * N + p] = c * kp - s * kq;
                    a[k * N + q] = s * kp + c * kq;
                }
                for (int k = 0; k < N; ++k) {  // rows p, q
                    const double pk = a[p * N + k], qk = a[q * N + k];
                    a[p * N + k] = c * pk - s * qk;
                    a[q * N + k] = s * pk + c * qk;
                }
                for (int k = 0; k < N; ++k) {
                    const double kp = v[k * N + p], kq = v[k * N + q];
                    v[k * N + p] = c * kp - s * kq;
                    v[k * N + q] = s * kp + c * kq;
                }
            }
    }
}

// Entry k of a ROI's two rows (selects, no indexing into arrays):
//   first  [-ox z, -oy z, -z^2, 0, 0, 0, ox nx, oy nx, z nx]
//   second [0, 0, 0, -ox z, -oy z, -z^2, ox ny, oy ny, z ny]
__device__ __forceinline__ double row_entry(bool second, int k, double ox, double oy, double nv, double z) {
#pragma clang fp contract(off)
    const int j = k < 6 ? k - (second ? 3 : 0) : k;
    return j == 0 ? -ox * z : j == 1 ? -oy * z : j == 2 ? -z * z : j == 6 ? ox * nv : j == 7 ? oy * nv
           : j == 8 ? z * nv : 0.0;
}

struct post_args {
    const double* mids;    // [n_rois][2] or [n_patterns][n_rois][2]
    uint32_t mids_stride;  // doubles between patterns' mids (0: shared)
    uint32_t n;            // ROIs per pattern
    const double* offsets; // [n_patterns][n][2]
    const double* coefs;   // [n_patterns][n][6] or null
    uint32_t npat;
    double cx, cy, z;
    double det_min;
    uint32_t min_rois;
    int32_t* mask;
    double* quality;
    double* H;
    double* F;
    double* eps;
    double* omega;
};

__global__ __launch_bounds__(kPostNT) void deform_fit_kernel(post_args a) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char post_lds[];
    double* dq = reinterpret_cast<double*>(post_lds);          // [n] quality determinants (with coefs)
    double* st = dq + (a.coefs ? a.n : 0);                     // [4][kPostChunk] old x, old y, new x, new y
    int* sm = reinterpret_cast<int*>(st + 4 * kPostChunk);     // [kPostChunk] ROI used
    __shared__ double g[81], vec[81], f3[9], c3[9], v3[9], med[2];
    __shared__ double gpart[45];
    __shared__ int cnt_w[kPostNT / kWave], nan_w[kPostNT / kWave];

    const uint32_t p = blockIdx.x;
    if (p >= a.npat) return;
    const int tid = threadIdx.x;
    const uint32_t n = a.n;
    const double* mids = a.mids + (size_t)p * a.mids_stride;
    const double* offs = a.offsets + (size_t)p * n * 2;

    // ---- quality determinants, ROI selection count, NaN flag
    int cnt = 0, has_nan = 0;
    if (a.coefs) {
        const double* q = a.coefs + (size_t)p * n * 6;
        for (uint32_t i = tid; i < n; i += kPostNT) {
            const double qa = -q[6 * (size_t)i + 3], qb = -q[6 * (size_t)i + 4] / 2, qc = -q[6 * (size_t)i + 5];
            const double d = qa * qc - qb * qb;
            dq[i] = d;
            cnt += d > a.det_min;
            has_nan |= d != d;
        }
    } else {
        cnt = tid == 0 ? (int)n : 0;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        cnt += __shfl_xor(cnt, off);
        has_nan |= __shfl_xor(has_nan, off);
    }
    if (tid % kWave == 0) {
        cnt_w[tid / kWave] = cnt;
        nan_w[tid / kWave] = has_nan;
    }
    __syncthreads();
    cnt = 0;
    has_nan = 0;
    for (int w = 0; w < kPostNT / kWave; ++w) {
        cnt += cnt_w[w];
        has_nan |= nan_w[w];
    }

    // ---- median of the determinants (numpy: middle element, or the mean of the two)
    if (a.coefs) {
        const uint32_t k1 = (n - 1) / 2, k2 = n / 2;
        for (uint32_t i = tid; i < n; i += kPostNT) {
            const double di = dq[i];
            uint32_t less = 0, eq = 0;
            for (uint32_t j = 0; j < n; ++j) {
                const double dj = dq[j];
                less += dj < di;
                eq += dj == di;
            }
            if (less <= k1 && k1 < less + eq) med[0] = di;  // equal values: same bits
            if (less <= k2 && k2 < less + eq) med[1] = di;
        }
        __syncthreads();
        if (tid == 0 && a.quality)
            a.quality[p] = has_nan ? (double)NAN : (med[0] + med[1]) / 2;
    } else if (tid == 0 && a.quality) {
        a.quality[p] = NAN;
    }

    const bool ok = cnt > (int)a.min_rois;
    if (tid == 0 && a.mask) a.mask[p] = ok ? 1 : 0;
    if (!ok) {
        if (tid < 9) {
            if (a.H) a.H[9 * (size_t)p + tid] = NAN;
            if (a.F) a.F[9 * (size_t)p + tid] = NAN;
            if (a.eps) a.eps[9 * (size_t)p + tid] = NAN;
            if (a.omega) a.omega[9 * (size_t)p + tid] = NAN;
        }
        return;
    }

    // ---- normal matrix: thread t < 45 sums entry (i, j), i <= j, in ROI order
    int gi = 0, gj = 0;
    if (tid < 45) {
        int t = tid;
        while (t >= 9 - gi) {
            t -= 9 - gi;
            ++gi;
        }
        gj = gi + t;
    }
    double acc = 0;
    const double z = a.z;
    for (uint32_t c0 = 0; c0 < n; c0 += kPostChunk) {
        const uint32_t m = min((uint32_t)kPostChunk, n - c0);
        __syncthreads();
        for (uint32_t e = tid; e < m; e += kPostNT) {
            const uint32_t i = c0 + e;
            const double ox = mids[2 * (size_t)i] - a.cx, oy = mids[2 * (size_t)i + 1] - a.cy;
            const double nx = (mids[2 * (size_t)i] + offs[2 * (size_t)i]) - a.cx;
            const double ny = (mids[2 * (size_t)i + 1] + offs[2 * (size_t)i + 1]) - a.cy;
            st[e] = ox;
            st[kPostChunk + e] = oy;
            st[2 * kPostChunk + e] = nx;
            st[3 * kPostChunk + e] = ny;
            sm[e] = a.coefs ? (dq[i] > a.det_min) : 1;
        }
        __syncthreads();
        if (tid < 45) {
            for (uint32_t e = 0; e < m; ++e) {
                if (!sm[e]) continue;
                const double ox = st[e], oy = st[kPostChunk + e], nx = st[2 * kPostChunk + e],
                             ny = st[3 * kPostChunk + e];
                acc += row_entry(false, gi, ox, oy, nx, z) * row_entry(false, gj, ox, oy, nx, z) +
                       row_entry(true, gi, ox, oy, ny, z) * row_entry(true, gj, ox, oy, ny, z);
            }
            gpart[tid] = acc;
        }
    }
    __syncthreads();
    if (tid != 0) return;
    for (int t = 0, i = 0; i < 9; ++i)
        for (int j = i; j < 9; ++j, ++t) g[i * 9 + j] = g[j * 9 + i] = gpart[t];

    // ---- null vector: eigenvector of the smallest eigenvalue
    jacobi_eigen<9>(g, vec);
    int kmin = 0;
    for (int k = 1; k < 9; ++k)
        if (g[k * 9 + k] < g[kmin * 9 + kmin]) kmin = k;
    for (int k = 0; k < 9; ++k) f3[k] = vec[k * 9 + kmin];  // row-major 3 x 3

    if (a.H)
        for (int k = 0; k < 9; ++k) a.H[9 * (size_t)p + k] = f3[k] / f3[8];
    const double det = f3[0] * (f3[4] * f3[8] - f3[5] * f3[7]) - f3[1] * (f3[3] * f3[8] - f3[5] * f3[6]) +
                       f3[2] * (f3[3] * f3[7] - f3[4] * f3[6]);
    const double scale = cbrt(det);
    for (int k = 0; k < 9; ++k) f3[k] /= scale;
    if (a.F)
        for (int k = 0; k < 9; ++k) a.F[9 * (size_t)p + k] = f3[k];

    // ---- polar decomposition F = R U: U^2 = F^T F = V diag(l) V^T
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            c3[i * 3 + j] = f3[0 * 3 + i] * f3[0 * 3 + j] + f3[1 * 3 + i] * f3[1 * 3 + j] + f3[2 * 3 + i] * f3[2 * 3 + j];
    jacobi_eigen<3>(c3, v3);
    double ls[3], ui[9];
    for (int k = 0; k < 3; ++k) ls[k] = sqrt(c3[k * 3 + k]);  // singular values of F
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double e = 0, u = 0;
            for (int k = 0; k < 3; ++k) {
                e += log(ls[k]) * v3[i * 3 + k] * v3[j * 3 + k];
                u += v3[i * 3 + k] * v3[j * 3 + k] / ls[k];
            }
            if (a.eps) a.eps[9 * (size_t)p + i * 3 + j] = e;  // log(U)
            ui[i * 3 + j] = u;                                 // U^-1
        }
    double r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[i * 3 + j] = f3[i * 3 + 0] * ui[0 * 3 + j] + f3[i * 3 + 1] * ui[1 * 3 + j] + f3[i * 3 + 2] * ui[2 * 3 + j];
    double kk[9], ss = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            kk[i * 3 + j] = (r[i * 3 + j] - r[j * 3 + i]) / 2;
            ss += kk[i * 3 + j] * kk[i * 3 + j];
        }
    const double sin_t = sqrt(ss / 2);
    const double fac = asin(sin_t) / sin_t;  // 0 / 0 for no rotation, as the reference
    if (a.omega)
        for (int k = 0; k < 9; ++k) a.omega[9 * (size_t)p + k] = kk[k] * fac;
}

}  // namespace
}  // namespace emida_amd

extern "C" int emida_fit_deformation(const double* mids, int mids_per_pattern, uint32_t n_rois,
                                     const double* offsets, const double* coefs, uint32_t n_patterns,
                                     emida_calib calib, double det_min, uint32_t min_rois, int32_t* mask,
                                     double* quality, double* H, double* F, double* epsilon, double* omega,
                                     void* stream) {
    return guarded([&] {
        if (!mids || !offsets) throw arg_error("null argument");
        if (n_rois == 0 || n_rois > kPostMaxRois)
            throw arg_error("n_rois must be in 1.." + std::to_string(kPostMaxRois));
        if (!coefs && n_rois <= 4) throw arg_error("the algebraic fit needs more than 4 ROIs");
        if (n_patterns == 0) return;
        post_args a{};
        a.mids = mids;
        a.mids_stride = mids_per_pattern ? 2 * n_rois : 0;
        a.n = n_rois;
        a.offsets = offsets;
        a.coefs = coefs;
        a.npat = n_patterns;
        a.cx = calib.x;
        a.cy = calib.y;
        a.z = calib.z;
        a.det_min = det_min;
        a.min_rois = coefs ? min_rois : 4;
        a.mask = mask;
        a.quality = quality;
        a.H = H;
        a.F = F;
        a.eps = epsilon;
        a.omega = omega;
        const size_t lds = (coefs ? n_rois * sizeof(double) : 0) + 4 * kPostChunk * sizeof(double) +
                           kPostChunk * sizeof(int);
        static const hipError_t attr = hipFuncSetAttribute((const void*)deform_fit_kernel,
                                                           hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);
        EMIDA_HIP(attr);
        deform_fit_kernel<<<n_patterns, kPostNT, lds, (hipStream_t)stream>>>(a);
        EMIDA_LAUNCHED("deform_fit_kernel");
    });
}
