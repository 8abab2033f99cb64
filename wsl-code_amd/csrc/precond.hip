// Multigrid V-cycle for the pressure block of the MINRES preconditioner on
// gfx950.  Every level operator, prolongation and restriction is a sliced
// ELLPACK matrix with one 64-row slice per wavefront, so a slice's width is a
// scalar and each load instruction of the slice reads 64 consecutive entries.
// Smoothing is Chebyshev iteration on D^-1 A over [lmax / ratio, lmax]
// (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., Alg. 12.1)
// with the solution update lagged by one step so that no kernel writes a
// vector that other rows of the same kernel read.
#include <algorithm>
#include <cmath>

#include "precond.h"

namespace vemd {

template <typename T>
__device__ __forceinline__ T sell_dot(const Sell<T>& A, int64_t row, const T* __restrict__ x) {
    const int64_t s = row >> 6;
    const int w = A.width[s];
    const int64_t o = A.off[s] + (row & 63);
    T acc = T(0);
    for (int j = 0; j < w; ++j) acc += A.val[o + (int64_t)j * 64] * x[A.col[o + (int64_t)j * 64]];
    return acc;
}

template <typename T>
__device__ __forceinline__ T sell_dot2(const Sell<T>& A, int64_t row, const T* __restrict__ x,
                                       const T* __restrict__ d) {
    const int64_t s = row >> 6;
    const int w = A.width[s];
    const int64_t o = A.off[s] + (row & 63);
    T acc = T(0);
    for (int j = 0; j < w; ++j) {
        const int c = A.col[o + (int64_t)j * 64];
        acc += A.val[o + (int64_t)j * 64] * (x[c] + d[c]);
    }
    return acc;
}

#define AMG_ROW                                                  \
    if (st && st->done) return;                                  \
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;   \
    if (i >= n) return;

// First pre-smoothing step from x = 0: r = b, d = D^-1 b / centre, x = 0.
template <typename T>
__global__ __launch_bounds__(256) void k_pre0(int64_t n, const T* b, const T* dinv, T* r, T* d, T* x, T c,
                                            const MState* st) {
    AMG_ROW
    const T bi = b[i];
    r[i] = bi;
    d[i] = dinv[i] * bi * c;
    x[i] = T(0);
}

// Chebyshev step: r -= A d, d' = c1 d + c2 D^-1 r, x += d.
template <typename T>
__global__ __launch_bounds__(256) void k_step(int64_t n, Sell<T> A, T* r, const T* dinv, const T* din, T* dout, T* x,
                                            T c1, T c2, const MState* st) {
    AMG_ROW
    const T ri = r[i] - sell_dot(A, i, din);
    r[i] = ri;
    const T di = din[i];
    dout[i] = c1 * di + c2 * dinv[i] * ri;
    x[i] += di;
}

// Residual of the lagged iterate: r = b - A (x + d).
template <typename T>
__global__ __launch_bounds__(256) void k_resid(int64_t n, Sell<T> A, const T* b, const T* x, const T* d, T* r,
                                             const MState* st) {
    AMG_ROW
    r[i] = b[i] - sell_dot2(A, i, x, d);
}

// y = M v (restriction).
template <typename T>
__global__ __launch_bounds__(256) void k_spmv(int64_t n, Sell<T> M, const T* v, T* y, const MState* st) {
    AMG_ROW
    y[i] = sell_dot(M, i, v);
}

// Coarse correction of the lagged iterate: x += d + P xc.
template <typename T>
__global__ __launch_bounds__(256) void k_prol(int64_t n, Sell<T> P, const T* xc, T* x, const T* d,
                                            const MState* st) {
    AMG_ROW
    x[i] += d[i] + sell_dot(P, i, xc);
}

// First post-smoothing step: r = b - A x, d = D^-1 r / centre (x unchanged).
template <typename T>
__global__ __launch_bounds__(256) void k_post0(int64_t n, Sell<T> A, const T* b, const T* x, const T* dinv, T* r,
                                             T* d, T c, const MState* st) {
    AMG_ROW
    const T ri = b[i] - sell_dot(A, i, x);
    r[i] = ri;
    d[i] = dinv[i] * ri * c;
}

template <typename T>
__global__ __launch_bounds__(256) void k_addto(int64_t n, T* x, const T* d, const MState* st) {
    AMG_ROW
    x[i] += d[i];
}

// Coarsest level: x = C^-1 b with the dense inverse, one wavefront per row.
template <typename T>
__global__ __launch_bounds__(64) void k_dense(int64_t n, const T* cinv, const T* b, T* x, const MState* st) {
    if (st && st->done) return;
    const int64_t i = blockIdx.x;
    double acc = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += 64) acc += (double)cinv[i * n + j] * (double)b[j];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if (threadIdx.x == 0) x[i] = (T)acc;
}

// ---------------------------------------------------------------------------

template <typename T>
AmgDev<T>::~AmgDev() {
    for (void* p : allocs) (void)hipFree(p);
}

template <typename T>
static void* amg_alloc(AmgDev<T>& D, size_t bytes) {
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, std::max<size_t>(bytes, 256)));
    D.allocs.push_back(p);
    D.bytes += (int64_t)bytes;
    return p;
}

template <typename T, typename S>
static S* amg_copy(AmgDev<T>& D, const std::vector<S>& v) {
    S* p = (S*)amg_alloc(D, v.size() * sizeof(S));
    if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(S), hipMemcpyHostToDevice));
    return p;
}

template <typename T>
static Sell<T> to_sell(AmgDev<T>& D, const Csr& A) {
    Sell<T> s;
    s.nrows = A.nrows;
    s.nslices = (int)((A.nrows + 63) / 64);
    std::vector<int64_t> off(s.nslices + 1, 0);
    std::vector<int> width(s.nslices, 0);
    for (int sl = 0; sl < s.nslices; ++sl) {
        int w = 0;
        for (int64_t r = sl * 64LL; r < std::min<int64_t>(A.nrows, sl * 64LL + 64); ++r)
            w = std::max<int>(w, (int)(A.rp[r + 1] - A.rp[r]));
        width[sl] = w;
        off[sl + 1] = off[sl] + (int64_t)w * 64;
    }
    std::vector<int> col(off[s.nslices], 0);
    std::vector<T> val(off[s.nslices], T(0));
    for (int64_t r = 0; r < A.nrows; ++r) {
        const int64_t sl = r >> 6, l = r & 63;
        for (int64_t p = A.rp[r]; p < A.rp[r + 1]; ++p) {
            const int64_t j = p - A.rp[r];
            col[off[sl] + j * 64 + l] = A.ci[p];
            val[off[sl] + j * 64 + l] = (T)A.v[p];
        }
    }
    s.off = amg_copy(D, off);
    s.width = amg_copy(D, width);
    s.col = amg_copy(D, col);
    s.val = amg_copy(D, val);
    return s;
}

template <typename T>
AmgDev<T>* amg_upload(const Amg& H) {
    AmgDev<T>* D = new AmgDev<T>();
    try {
        D->lv.resize(H.lv.size());
        for (size_t l = 0; l < H.lv.size(); ++l) {
            const AmgLevel& h = H.lv[l];
            AmgLevelDev<T>& d = D->lv[l];
            d.n = h.A.nrows;
            d.A = to_sell(*D, h.A);
            if (l + 1 < H.lv.size()) {
                d.P = to_sell(*D, h.P);
                d.R = to_sell(*D, h.R);
            }
            std::vector<T> di(h.dinv.begin(), h.dinv.end());
            d.dinv = amg_copy(*D, di);
            const size_t vb = (size_t)d.n * sizeof(T);
            d.b = (T*)amg_alloc(*D, vb);
            d.x = (T*)amg_alloc(*D, vb);
            d.r = (T*)amg_alloc(*D, vb);
            d.d0 = (T*)amg_alloc(*D, vb);
            d.d1 = (T*)amg_alloc(*D, vb);
            for (T* p : {d.b, d.x, d.r, d.d0, d.d1}) HIPCHK(hipMemset(p, 0, vb));
            const double hi = h.lmax, lo = h.lmax / H.opt.cheb_ratio;
            d.centre = 0.5 * (hi + lo);
            d.half = 0.5 * (hi - lo);
            d.degree = l == 0 ? std::max(1, H.opt.degree0) : std::max(1, H.opt.degree);
        }
        std::vector<T> ci(H.coarse_inv.begin(), H.coarse_inv.end());
        D->cinv = amg_copy(*D, ci);
        D->nc = H.ncoarse;
    } catch (...) {
        delete D;
        throw;
    }
    return D;
}

template <typename T>
static void smooth_steps(const AmgLevelDev<T>& L, hipStream_t st, const MState* ms, T*& din) {
    const double sigma = L.centre / L.half;
    double rho = 1.0 / sigma;
    T* dout = din == L.d0 ? L.d1 : L.d0;
    const dim3 g((unsigned)((L.n + 255) / 256)), b(256);
    for (int k = 1; k < L.degree; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        hipLaunchKernelGGL((k_step<T>), g, b, 0, st, L.n, L.A, L.r, (const T*)L.dinv, (const T*)din, dout, L.x,
                           (T)(rn * rho), (T)(2.0 * rn / L.half), ms);
        std::swap(din, dout);
        rho = rn;
    }
}

template <typename T>
static void vcycle_level(const AmgDev<T>& D, size_t l, hipStream_t st, const MState* ms) {
    const AmgLevelDev<T>& L = D.lv[l];
    if (l + 1 == D.lv.size()) {
        hipLaunchKernelGGL((k_dense<T>), dim3((unsigned)D.nc), dim3(64), 0, st, D.nc, (const T*)D.cinv,
                           (const T*)L.b, L.x, ms);
        return;
    }
    const AmgLevelDev<T>& C = D.lv[l + 1];
    const dim3 g((unsigned)((L.n + 255) / 256)), b(256), gc((unsigned)((C.n + 255) / 256));
    const T c0 = (T)(1.0 / L.centre);
    T* din = L.d0;
    hipLaunchKernelGGL((k_pre0<T>), g, b, 0, st, L.n, (const T*)L.b, (const T*)L.dinv, L.r, L.d0, L.x, c0, ms);
    smooth_steps(L, st, ms, din);
    hipLaunchKernelGGL((k_resid<T>), g, b, 0, st, L.n, L.A, (const T*)L.b, (const T*)L.x, (const T*)din, L.r, ms);
    hipLaunchKernelGGL((k_spmv<T>), gc, b, 0, st, C.n, L.R, (const T*)L.r, C.b, ms);
    vcycle_level(D, l + 1, st, ms);
    hipLaunchKernelGGL((k_prol<T>), g, b, 0, st, L.n, L.P, (const T*)C.x, L.x, (const T*)din, ms);
    din = L.d0;
    hipLaunchKernelGGL((k_post0<T>), g, b, 0, st, L.n, L.A, (const T*)L.b, (const T*)L.x, (const T*)L.dinv, L.r,
                       L.d0, c0, ms);
    smooth_steps(L, st, ms, din);
    hipLaunchKernelGGL((k_addto<T>), g, b, 0, st, L.n, L.x, (const T*)din, ms);
}

template <typename T>
void amg_vcycle(const AmgDev<T>& D, hipStream_t st, const MState* ms) {
    vcycle_level(D, 0, st, ms);
    HIPCHK(hipGetLastError());
}

template struct AmgDev<double>;
template struct AmgDev<float>;
template AmgDev<double>* amg_upload<double>(const Amg&);
template AmgDev<float>* amg_upload<float>(const Amg&);
template void amg_vcycle<double>(const AmgDev<double>&, hipStream_t, const MState*);
template void amg_vcycle<float>(const AmgDev<float>&, hipStream_t, const MState*);

}  // namespace vemd
