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

// out = (A (x + y))_brow for one block row of bs = B rows (y may be null).
template <typename T, int B>
__device__ __forceinline__ void sell_row(const Sell<T>& A, int64_t brow, const T* __restrict__ x, T (&out)[B],
                                         const T* __restrict__ y = nullptr) {
    const int64_t s = brow >> 6;
    const int w = A.width[s];
    const int64_t o = A.off[s] + (brow & 63);
#pragma unroll
    for (int r = 0; r < B; ++r) out[r] = T(0);
    for (int j = 0; j < w; ++j) {
        const int64_t e = o + (int64_t)j * 64;
        const int64_t c = (int64_t)A.col[e] * B;
        const T* v = A.val + e * (B * B);
        T xc[B];
#pragma unroll
        for (int cc = 0; cc < B; ++cc) xc[cc] = y ? x[c + cc] + y[c + cc] : x[c + cc];
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int cc = 0; cc < B; ++cc) out[r] += v[r * B + cc] * xc[cc];
    }
}

template <typename T>
__device__ __forceinline__ T sell_dot(const Sell<T>& A, int64_t row, const T* __restrict__ x) {
    T o[1];
    sell_row<T, 1>(A, row, x, o);
    return o[0];
}

#define AMG_ROW                                                  \
    if (done && *done) return;                                   \
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;   \
    if (i >= n) return;

#define AMG_BROW                                                 \
    if (done && *done) return;                                   \
    const int64_t bi = (int64_t)blockIdx.x * 256 + threadIdx.x;  \
    if (bi >= A.nrows) return;

// First pre-smoothing step from x = 0: r = b, d = D^-1 b / centre, x = 0.
template <typename T>
__global__ __launch_bounds__(256) void k_pre0(int64_t n, const T* b, const T* dinv, T* r, T* d, T* x, T c,
                                            const int* done) {
    AMG_ROW
    const T bi = b[i];
    r[i] = bi;
    d[i] = dinv[i] * bi * c;
    x[i] = T(0);
}

// Chebyshev step per block row: r -= A d, d' = c1 d + c2 D^-1 r, x += d.
template <typename T, int B>
__global__ __launch_bounds__(256) void k_step(Sell<T> A, T* r, const T* dinv, const T* din, T* dout, T* x, T c1, T c2,
                                            const int* done) {
    AMG_BROW
    T ad[B];
    sell_row<T, B>(A, bi, din, ad);
#pragma unroll
    for (int k = 0; k < B; ++k) {
        const int64_t i = bi * B + k;
        const T ri = r[i] - ad[k];
        r[i] = ri;
        const T di = din[i];
        dout[i] = c1 * di + c2 * dinv[i] * ri;
        x[i] += di;
    }
}

// Residual of the lagged iterate: r = b - A x - A d.
template <typename T, int B>
__global__ __launch_bounds__(256) void k_resid(Sell<T> A, const T* b, const T* x, const T* d, T* r, const int* done) {
    AMG_BROW
    T ax[B];
    sell_row<T, B>(A, bi, x, ax, d);
#pragma unroll
    for (int k = 0; k < B; ++k) r[bi * B + k] = b[bi * B + k] - ax[k];
}

// First post-smoothing step: r = b - A x, d = D^-1 r / centre (x unchanged).
template <typename T, int B>
__global__ __launch_bounds__(256) void k_post0(Sell<T> A, const T* b, const T* x, const T* dinv, T* r, T* d, T c,
                                             const int* done) {
    AMG_BROW
    T ax[B];
    sell_row<T, B>(A, bi, x, ax);
#pragma unroll
    for (int k = 0; k < B; ++k) {
        const int64_t i = bi * B + k;
        const T ri = b[i] - ax[k];
        r[i] = ri;
        d[i] = dinv[i] * ri * c;
    }
}

// y = M v (restriction).
template <typename T>
__global__ __launch_bounds__(256) void k_spmv(int64_t n, Sell<T> M, const T* v, T* y, const int* done) {
    AMG_ROW
    y[i] = sell_dot(M, i, v);
}

// Coarse correction of the lagged iterate: x += d + P xc.
template <typename T>
__global__ __launch_bounds__(256) void k_prol(int64_t n, Sell<T> P, const T* xc, T* x, const T* d,
                                            const int* done) {
    AMG_ROW
    x[i] += d[i] + sell_dot(P, i, xc);
}

// First post-smoothing step: r = b - A x, d = D^-1 r / centre (x unchanged).
template <typename T>
__global__ __launch_bounds__(256) void k_post0(int64_t n, Sell<T> A, const T* b, const T* x, const T* dinv, T* r,
                                             T* d, T c, const int* done) {
    AMG_ROW
    const T ri = b[i] - sell_dot(A, i, x);
    r[i] = ri;
    d[i] = dinv[i] * ri * c;
}

template <typename T>
__global__ __launch_bounds__(256) void k_addto(int64_t n, T* x, const T* d, const int* done) {
    AMG_ROW
    x[i] += d[i];
}

// Coarsest level: x = C^-1 b with the dense inverse, one wavefront per row.
template <typename T>
__global__ __launch_bounds__(64) void k_dense(int64_t n, const T* cinv, const T* b, T* x, const int* done) {
    if (done && *done) return;
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
static Sell<T> to_sell(AmgDev<T>& D, const Csr& A, int bs) {
    Sell<T> s;
    s.bs = bs;
    s.nrows = A.nrows / bs;
    s.nslices = (int)((s.nrows + 63) / 64);
    // block columns of every block row
    std::vector<std::vector<int>> bc(s.nrows);
    for (int64_t br = 0; br < s.nrows; ++br) {
        std::vector<int>& v = bc[br];
        for (int k = 0; k < bs; ++k)
            for (int64_t p = A.rp[br * bs + k]; p < A.rp[br * bs + k + 1]; ++p) v.push_back(A.ci[p] / bs);
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    std::vector<int64_t> off(s.nslices + 1, 0);
    std::vector<int> width(s.nslices, 0);
    for (int sl = 0; sl < s.nslices; ++sl) {
        int w = 0;
        for (int64_t r = sl * 64LL; r < std::min<int64_t>(s.nrows, sl * 64LL + 64); ++r) w = std::max<int>(w, (int)bc[r].size());
        width[sl] = w;
        off[sl + 1] = off[sl] + (int64_t)w * 64;
    }
    std::vector<int> col(off[s.nslices], 0);
    std::vector<T> val(off[s.nslices] * bs * bs, T(0));
    for (int64_t br = 0; br < s.nrows; ++br) {
        const int64_t sl = br >> 6, l = br & 63;
        const std::vector<int>& v = bc[br];
        for (size_t j = 0; j < v.size(); ++j) col[off[sl] + (int64_t)j * 64 + l] = v[j];
        for (int k = 0; k < bs; ++k)
            for (int64_t p = A.rp[br * bs + k]; p < A.rp[br * bs + k + 1]; ++p) {
                const int64_t j = std::lower_bound(v.begin(), v.end(), A.ci[p] / bs) - v.begin();
                val[(off[sl] + j * 64 + l) * bs * bs + k * bs + A.ci[p] % bs] = (T)A.v[p];
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
            const int bs = l == 0 ? H.bs0 : 1;
            if (h.A.nrows % bs) throw Error(VEMD_ERR_INTERNAL, "finest multigrid level is not blocked");
            d.A = to_sell(*D, h.A, bs);
            if (l + 1 < H.lv.size()) {
                d.P = to_sell(*D, h.P, 1);
                d.R = to_sell(*D, h.R, 1);
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

template <typename T, int B>
static void smooth_steps(const AmgLevelDev<T>& L, hipStream_t st, const int* ms, T*& din) {
    const double sigma = L.centre / L.half;
    double rho = 1.0 / sigma;
    T* dout = din == L.d0 ? L.d1 : L.d0;
    const dim3 g((unsigned)((L.A.nrows + 255) / 256)), b(256);
    for (int k = 1; k < L.degree; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        hipLaunchKernelGGL((k_step<T, B>), g, b, 0, st, L.A, L.r, (const T*)L.dinv, (const T*)din, dout, L.x,
                           (T)(rn * rho), (T)(2.0 * rn / L.half), ms);
        std::swap(din, dout);
        rho = rn;
    }
}

template <typename T>
static void vcycle_level(const AmgDev<T>& D, size_t l, hipStream_t st, const int* ms);

// Pre-smoothing, residual, restriction, coarse cycle, prolongation and
// post-smoothing of one level whose operator has blocks of B rows.
template <typename T, int B>
static void level_cycle(const AmgDev<T>& D, size_t l, hipStream_t st, const int* ms) {
    const AmgLevelDev<T>& L = D.lv[l];
    const AmgLevelDev<T>& C = D.lv[l + 1];
    const dim3 g((unsigned)((L.n + 255) / 256)), gb((unsigned)((L.A.nrows + 255) / 256)), b(256),
        gc((unsigned)((C.n + 255) / 256));
    const T c0 = (T)(1.0 / L.centre);
    T* din = L.d0;
    hipLaunchKernelGGL((k_pre0<T>), g, b, 0, st, L.n, (const T*)L.b, (const T*)L.dinv, L.r, L.d0, L.x, c0, ms);
    smooth_steps<T, B>(L, st, ms, din);
    hipLaunchKernelGGL((k_resid<T, B>), gb, b, 0, st, L.A, (const T*)L.b, (const T*)L.x, (const T*)din, L.r, ms);
    hipLaunchKernelGGL((k_spmv<T>), gc, b, 0, st, C.n, L.R, (const T*)L.r, C.b, ms);
    vcycle_level(D, l + 1, st, ms);
    hipLaunchKernelGGL((k_prol<T>), g, b, 0, st, L.n, L.P, (const T*)C.x, L.x, (const T*)din, ms);
    din = L.d0;
    hipLaunchKernelGGL((k_post0<T, B>), gb, b, 0, st, L.A, (const T*)L.b, (const T*)L.x, (const T*)L.dinv, L.r, L.d0,
                       c0, ms);
    smooth_steps<T, B>(L, st, ms, din);
    hipLaunchKernelGGL((k_addto<T>), g, b, 0, st, L.n, L.x, (const T*)din, ms);
}

template <typename T>
static void vcycle_level(const AmgDev<T>& D, size_t l, hipStream_t st, const int* ms) {
    const AmgLevelDev<T>& L = D.lv[l];
    if (l + 1 == D.lv.size()) {
        hipLaunchKernelGGL((k_dense<T>), dim3((unsigned)D.nc), dim3(64), 0, st, D.nc, (const T*)D.cinv,
                           (const T*)L.b, L.x, ms);
        return;
    }
    switch (L.A.bs) {
        case 1: level_cycle<T, 1>(D, l, st, ms); break;
        case 3: level_cycle<T, 3>(D, l, st, ms); break;
        default: throw Error(VEMD_ERR_INTERNAL, "unsupported multigrid block size");
    }
}

template <typename T>
void amg_vcycle(const AmgDev<T>& D, hipStream_t st, const int* ms) {
    vcycle_level(D, 0, st, ms);
    HIPCHK(hipGetLastError());
}

template struct AmgDev<double>;
template struct AmgDev<float>;
template AmgDev<double>* amg_upload<double>(const Amg&);
template AmgDev<float>* amg_upload<float>(const Amg&);
template void amg_vcycle<double>(const AmgDev<double>&, hipStream_t, const int*);
template void amg_vcycle<float>(const AmgDev<float>&, hipStream_t, const int*);

}  // namespace vemd
