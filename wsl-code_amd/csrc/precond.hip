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
#include <cstdio>
#include <cstdlib>

#include "precond.h"

namespace vemd {

// Accumulates entries [0, w) of a slice row into out: U entries at a time,
// with their column indices, then their gathers, issued together (the
// gathers depend on the indices; batching keeps several in flight per lane).
// Gathered entries are x (+ y) times cs[column] when cs is non-null; the sum
// runs in entry order.
// HY, HC: the pending update y and the column scaling cs are present (known at
// compile time, so no gather waits behind a branch).
template <typename T, typename V, int B, int NC, int U, bool HY, bool HC>
__device__ __forceinline__ void sell_acc(const int* __restrict__ col, const V* __restrict__ vs, int64_t o, int w, int j,
                                         const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ cs,
                                         T (&out)[B]) {
    int k = 0;
    for (; k + U <= w; k += U) {
        int64_t c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = (int64_t)col[o + (int64_t)(k + u) * 64] * B;
        T xc[U][B];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int cc = 0; cc < B; ++cc) {
                const int64_t xi = (c[u] + cc) * NC + j;
                T xv = x[xi];
                if constexpr (HY) xv += y[xi];
                if constexpr (HC) xv *= cs[c[u] + cc];
                xc[u][cc] = xv;
            }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const V* v = vs + (int64_t)(k + u) * 64 * (B * B);
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
                for (int cc = 0; cc < B; ++cc) out[r] += (T)v[(r * B + cc) * 64] * xc[u][cc];
        }
    }
    if constexpr (B == 1) {
        // the tail as one more group, loads issued together: entries past w
        // re-read the row's last entry (in bounds) and contribute zero
        if (k < w) {
            int kk[U];
            int64_t c[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                kk[u] = k + u < w ? k + u : w - 1;
                c[u] = (int64_t)col[o + (int64_t)kk[u] * 64];
            }
            T xc[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t xi = c[u] * NC + j;
                T xv = x[xi];
                if constexpr (HY) xv += y[xi];
                if constexpr (HC) xv *= cs[c[u]];
                xc[u] = k + u < w ? xv : T(0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) out[0] += (T)vs[(int64_t)kk[u] * 64] * xc[u];
        }
    } else {
        for (; k < w; ++k) {
            const int64_t c = (int64_t)col[o + (int64_t)k * 64] * B;
            const V* v = vs + (int64_t)k * 64 * (B * B);
            T xc[B];
#pragma unroll
            for (int cc = 0; cc < B; ++cc) {
                const int64_t xi = (c + cc) * NC + j;
                T xv = x[xi];
                if constexpr (HY) xv += y[xi];
                if constexpr (HC) xv *= cs[c + cc];
                xc[cc] = xv;
            }
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
                for (int cc = 0; cc < B; ++cc) out[r] += (T)v[(r * B + cc) * 64] * xc[cc];
        }
    }
}

// out = (A (x + y))_brow, column j of NC interleaved columns (vector entry
// (row, j) at row * NC + j), for one block row of bs = B rows (y may be null);
// with ds non-null the gathered entries are scaled by ds[row] first.
// Lanes per row: small single-column scalar levels (below 32k rows by default,
// ~50 entries per row) spread each row over 8 consecutive lanes (entries
// k = sub, sub + 8, ...) summed with shuffles, which shortens each lane's chain
// of dependent loads; larger levels (scattered slice loads would multiply),
// block levels and the 32-column cycle keep one lane per row. The upload
// records the choice in the level's operator (Sell::lanes); the launch grids
// and the kernels read it there.
constexpr int LANES_SMALL = 8;
template <int B, int NC, typename T>
__host__ __device__ inline int lanes_per_row(const Sell<T>& A) {
    return B == 1 && NC == 1 ? A.lanes : 1;
}

// Row brow of a scalar SELL operator by the LP lanes of its group (lane sub):
// every lane's share as one group of U gathers (loads issued together; entries
// past w re-read entry w - 1 in bounds and contribute zero), then the sum over
// the group. Every lane of the group gets the row's value.
template <typename T, typename V, int U, bool HY, bool HC>
__device__ __forceinline__ T sell_row_lanes(const int* __restrict__ col, const V* __restrict__ vs, int64_t o, int w,
                                            int sub, const T* __restrict__ x, const T* __restrict__ y,
                                            const T* __restrict__ cs) {
    constexpr int LP = LANES_SMALL;
    T acc = T(0);
    for (int k0 = sub; k0 < w; k0 += LP * U) {
        int kk[U];
        int64_t c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u * LP;
            kk[u] = k < w ? k : w - 1;
            c[u] = (int64_t)col[o + (int64_t)kk[u] * 64];
        }
        T xc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            T xv = x[c[u]];
            if constexpr (HY) xv += y[c[u]];
            if constexpr (HC) xv *= cs[c[u]];
            xc[u] = k0 + u * LP < w ? xv : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc += (T)vs[(int64_t)kk[u] * 64] * xc[u];
    }
#pragma unroll
    for (int m = LP / 2; m > 0; m >>= 1) acc += __shfl_xor(acc, m, 64);
    return acc;
}

// The 32-column kernels give each block row one half-wave (its 32 columns),
// so every block value would be loaded by 32 lanes at once and one load
// instruction would fetch two values: the values cost far more load
// instructions (address work) than the gathers.  A 256-thread block of such a
// kernel covers 8 block rows of one slice; it stages their values (as T) and
// block columns in LDS with one load per value, and the lanes read them from
// there (broadcast reads).  Slices wider than SW entries are not staged.
template <typename T, int B>
struct Stage32 {
    static constexpr int SW = B == 1 ? 64 : 16;
    T v[SW][B * B][8];
    int c[SW][8];
};

template <typename T, int B, int NC, typename S>
__device__ __forceinline__ const S* stage_rows(const Sell<T>& A, S& st) {
    const int64_t br0 = (int64_t)blockIdx.x * (256 / NC);  // 8 block rows, within one slice of 64
    if (br0 >= A.nrows) return nullptr;                      // (uniform over the block)
    const int64_t sl = br0 >> 6;
    const int w = A.width[sl];
    if (w > S::SW) return nullptr;
    const int l0 = (int)(br0 & 63);
    const int64_t o = A.off[sl];
    // rows past nrows are slice padding: zero values, column 0, in bounds
    for (int e = threadIdx.x; e < w * B * B * 8; e += 256) {
        const int r8 = e & 7, rc = (e >> 3) % (B * B), k = (e >> 3) / (B * B);
        const int64_t at = (o + (int64_t)k * 64) * (B * B) + (int64_t)rc * 64 + l0 + r8;
        st.v[k][rc][r8] = A.hval ? (T)A.hval[at] : A.val[at];
    }
    for (int e = threadIdx.x; e < w * 8; e += 256) st.c[e >> 3][e & 7] = A.col[o + (int64_t)(e >> 3) * 64 + l0 + (e & 7)];
    __syncthreads();
    return &st;
}

// sell_acc over a staged row (r8 of the block's 8), in the same entry order.
template <typename T, int B, bool HY, bool HC>
__device__ __forceinline__ void sell_acc_staged(const Stage32<T, B>& S, int w, int r8, int j, const T* __restrict__ x,
                                                const T* __restrict__ y, const T* __restrict__ cs, T (&out)[B]) {
    constexpr int NC = MCOLS;
    constexpr int U = B == 1 ? 8 : 4;
#pragma unroll
    for (int r = 0; r < B; ++r) out[r] = T(0);
    int k = 0;
    for (; k + U <= w; k += U) {
        int64_t c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = (int64_t)S.c[k + u][r8] * B;
        T xc[U][B];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int cc = 0; cc < B; ++cc) {
                const int64_t xi = (c[u] + cc) * NC + j;
                T xv = x[xi];
                if constexpr (HY) xv += y[xi];
                if constexpr (HC) xv *= cs[c[u] + cc];
                xc[u][cc] = xv;
            }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
                for (int cc = 0; cc < B; ++cc) out[r] += S.v[k + u][r * B + cc][r8] * xc[u][cc];
    }
    for (; k < w; ++k) {
        const int64_t c = (int64_t)S.c[k][r8] * B;
        T xc[B];
#pragma unroll
        for (int cc = 0; cc < B; ++cc) {
            const int64_t xi = (c + cc) * NC + j;
            T xv = x[xi];
            if constexpr (HY) xv += y[xi];
            if constexpr (HC) xv *= cs[c + cc];
            xc[cc] = xv;
        }
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int cc = 0; cc < B; ++cc) out[r] += S.v[k][r * B + cc][r8] * xc[cc];
    }
}

// UG: gathers in flight per lane for scalar rows (0: the default of the level
// operators; the prolongation's short rows pass 4).
template <typename T, int B, int NC, int UG, bool HY, bool HC>
__device__ __forceinline__ void sell_row_impl(const Sell<T>& A, int64_t brow, int j, const T* __restrict__ x,
                                              T (&out)[B], const T* __restrict__ y, const T* __restrict__ ds,
                                              int sub, const Stage32<T, B>* stg) {
    if constexpr (NC == MCOLS) {
        if (stg) {
            sell_acc_staged<T, B, HY, HC>(*stg, A.width[brow >> 6], (int)(brow & 7), j, x, y, ds, out);
            return;
        }
    }
    if constexpr (UG == 0 && B == 1 && NC == 1) {
        if (lanes_per_row<B, NC>(A) > 1) {
            const int64_t s = brow >> 6;
            const int w = A.width[s];
            const int64_t o = A.off[s] + (brow & 63);
            if (A.hval)
                out[0] = sell_row_lanes<T, _Float16, 4, HY, HC>(A.col, A.hval + A.off[s] + (brow & 63), o, w, sub, x, y,
                                                                ds);
            else
                out[0] = sell_row_lanes<T, T, 4, HY, HC>(A.col, A.val + A.off[s] + (brow & 63), o, w, sub, x, y, ds);
            return;
        }
    }
    // scalar (coarse) levels: few rows, so each lane keeps many gathers in flight
    constexpr int U = UG > 0 ? UG : (B == 1 ? 16 : 2);
    const int64_t s = brow >> 6;
    const int w = A.width[s];
    const int64_t o = A.off[s] + (brow & 63);
#pragma unroll
    for (int r = 0; r < B; ++r) out[r] = T(0);
    if (A.hval)
        sell_acc<T, _Float16, B, NC, U, HY, HC>(A.col, A.hval + A.off[s] * (B * B) + (brow & 63), o, w, j, x, y, ds,
                                                out);
    else
        sell_acc<T, T, B, NC, U, HY, HC>(A.col, A.val + A.off[s] * (B * B) + (brow & 63), o, w, j, x, y, ds, out);
}

// Row brow of A times x (+ y) (scaled by ds per column): one uniform branch
// on which of y and ds are given, then loops without per-gather tests.
template <typename T, int B, int NC, int UG = 0>
__device__ __forceinline__ void sell_row(const Sell<T>& A, int64_t brow, int j, const T* __restrict__ x, T (&out)[B],
                                         const T* __restrict__ y = nullptr, const T* __restrict__ ds = nullptr,
                                         int sub = 0, const Stage32<T, B>* stg = nullptr) {
    if (y) {
        if (ds)
            sell_row_impl<T, B, NC, UG, true, true>(A, brow, j, x, out, y, ds, sub, stg);
        else
            sell_row_impl<T, B, NC, UG, true, false>(A, brow, j, x, out, y, ds, sub, stg);
    } else {
        if (ds)
            sell_row_impl<T, B, NC, UG, false, true>(A, brow, j, x, out, y, ds, sub, stg);
        else
            sell_row_impl<T, B, NC, UG, false, false>(A, brow, j, x, out, y, ds, sub, stg);
    }
}

// The staged block rows of a 32-column block-row kernel (null: not staged,
// or a single-column kernel).  Every thread of the block calls it before any
// returns (it holds a barrier).
template <typename T, int B, int NC>
__device__ __forceinline__ const Stage32<T, B>* stage_block(const Sell<T>& A) {
    if constexpr (NC == MCOLS) {
        __shared__ Stage32<T, B> st;
        return stage_rows<T, B, NC>(A, st);
    } else {
        return nullptr;
    }
}

// Thread t of an elementwise sweep over n rows x NC columns: entry t, row t / NC.
#define AMG_ROW                                                  \
    if (done && *done) return;                                   \
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;   \
    if (i >= n * NC) return;                                     \
    const int64_t row = i / NC;

// Thread t of a block-row sweep: block row t / NC, column t % NC.
#define AMG_BROW                                                          \
    if (done && *done) return;                                            \
    const int LPR = lanes_per_row<B, NC>(A);                              \
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;            \
    const int64_t bi = t / (NC * LPR);                                    \
    const int j = (int)((t / LPR) % NC);                                  \
    const int sub = (int)(t % LPR);                                       \
    const Stage32<T, B>* stg = stage_block<T, B, NC>(A);                  \
    if (bi >= A.nrows) return;

// Rows of finished instances of a batched solve, left alone by the finest
// level's block-row and prolongation kernels after the first cycle of a
// refinement pass (nothing reads them: the instances' Krylov output is not
// written either, and the hierarchy has no coupling between instances):
// instance of every 64-block-row tile and the instances' Krylov states; bs
// rows per block row.  ti null: every row is computed.
struct RowSkip {
    const int* ti = nullptr;
    const MState* st = nullptr;
    int bs = 1;
};
__device__ __forceinline__ bool skipped(const RowSkip& k, int64_t brow) {
    return k.ti && k.st[k.ti[brow >> 6]].done;
}

// Diagonal entry of D^-1 (null: unit diagonal) and the right-hand side of a
// scaled level (es non-null: es * b).
template <typename T>
__device__ __forceinline__ T dval(const T* dinv, int64_t row) {
    return dinv ? dinv[row] : T(1);
}
template <typename T>
__device__ __forceinline__ T bval(const T* b, const T* es, int64_t i, int64_t row) {
    return es ? b[i] * es[row] : b[i];
}

// First pre-smoothing step from x = 0: r = b, d = D^-1 b / centre, x = 0.
template <typename T, int NC>
__global__ __launch_bounds__(256) void k_pre0(int64_t n, const T* b, const T* es, const T* dinv, T* r, T* d, T* x,
                                            T c, const int* done) {
    AMG_ROW
    const T bi = bval(b, es, i, row);
    r[i] = bi;
    d[i] = dval(dinv, row) * bi * c;
    x[i] = T(0);
}

// Chebyshev step per block row: r -= A d, d' = c1 d + c2 D^-1 r, x += d.
// The last step of a sweep (last != 0) completes the iterate instead,
// x += d + d', and leaves r and d' unwritten (nothing reads them).
// With xs non-null the completed iterate leaves the scaled space: x = xs (x + d + d').
// With b0 non-null this is the first step of the sweep from x = 0 with r = b0
// (times es on a scaled level) and d = din formed by the caller.
template <typename T, int B, int NC>
__global__ __launch_bounds__(256) void k_step(Sell<T> A, T* r, const T* dinv, const T* din, T* dout, T* x, T c1, T c2,
                                            int last, const T* xs, const T* b0, const T* es, const int* done,
                                            RowSkip sk) {
    AMG_BROW
    if (skipped(sk, bi)) return;
    T ad[B];
    sell_row<T, B, NC>(A, bi, j, din, ad, nullptr, nullptr, sub, stg);
#pragma unroll
    for (int k = 0; k < B; ++k) {
        const int64_t i = (bi * B + k) * NC + j;
        const T ri = (b0 ? bval(b0, es, i, bi * B + k) : r[i]) - ad[k];
        const T di = din[i];
        const T dn = c1 * di + c2 * dval(dinv, bi * B + k) * ri;
        const T xo = b0 ? T(0) : x[i];
        if (last) {
            const T xn = xo + di + dn;
            x[i] = xs ? xn * xs[bi * B + k] : xn;
        } else {
            r[i] = ri;
            dout[i] = dn;
            x[i] = xo + di;
        }
    }
}

// First Chebyshev step from x = 0, with d0 = c0 D^-1 b formed on the fly:
// r = b - A d0, d' = c1 d0 + c2 D^-1 r, x = d0 (last: x = d0 + d' only).
template <typename T, int B, int NC>
__global__ __launch_bounds__(256) void k_step0(Sell<T> A, const T* b, const T* dinv, T* r, T* dout, T* x, T c0, T c1,
                                             T c2, int last, const int* done, RowSkip sk) {
    AMG_BROW
    if (skipped(sk, bi)) return;
    T ad[B];
    sell_row<T, B, NC>(A, bi, j, b, ad, nullptr, dinv, sub, stg);
#pragma unroll
    for (int k = 0; k < B; ++k) {
        const int64_t i = (bi * B + k) * NC + j;
        const T bv = b[i], dv = dval(dinv, bi * B + k);
        const T ri = bv - c0 * ad[k];
        const T di = c0 * dv * bv;
        const T dn = c1 * di + c2 * dv * ri;
        if (last) {
            x[i] = di + dn;
        } else {
            r[i] = ri;
            dout[i] = dn;
            x[i] = di;
        }
    }
}

// Residual of the lagged iterate: r = b - A (x + d) (d null: x is complete).
template <typename T, int B, int NC>
__global__ __launch_bounds__(256) void k_resid(Sell<T> A, const T* b, const T* es, const T* x, const T* d, T* r,
                                             const int* done) {
    AMG_BROW
    T ax[B];
    sell_row<T, B, NC>(A, bi, j, x, ax, d, nullptr, sub, stg);
#pragma unroll
    for (int k = 0; k < B; ++k) {
        const int64_t i = (bi * B + k) * NC + j;
        r[i] = bval(b, es, i, bi * B + k) - ax[k];
    }
}

// First post-smoothing step: r = b - A x, d = D^-1 r / centre (x unchanged).
template <typename T, int B, int NC>
__global__ __launch_bounds__(256) void k_post0(Sell<T> A, const T* b, const T* es, const T* x, const T* dinv, T* r,
                                             T* d, T c, const int* done, RowSkip sk) {
    AMG_BROW
    if (skipped(sk, bi)) return;
    T ax[B];
    sell_row<T, B, NC>(A, bi, j, x, ax, nullptr, nullptr, sub, stg);
#pragma unroll
    for (int k = 0; k < B; ++k) {
        const int64_t i = (bi * B + k) * NC + j;
        const T ri = bval(b, es, i, bi * B + k) - ax[k];
        r[i] = ri;
        d[i] = dval(dinv, bi * B + k) * ri * c;
    }
}

// y = R v (restriction).  One column: 32 lanes per row, partial sums
// combined in a fixed butterfly order; NC columns: one thread per (row, column).
template <typename T, int NC>
__global__ __launch_bounds__(256) void k_restrict(CsrDev<T> R, const T* v, T* y, const int* done) {
    if (done && *done) return;
    if constexpr (NC == 1) {
        const int64_t row = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 5;
        const int lane = threadIdx.x & 31;
        T acc = T(0);
        if (row < R.nrows) {
            if (R.hv) {
                for (int64_t p = R.rp[row] + lane; p < R.rp[row + 1]; p += 32) acc += (T)R.hv[p] * v[R.ci[p]];
                acc *= (T)R.hscale;
            } else
                for (int64_t p = R.rp[row] + lane; p < R.rp[row + 1]; p += 32) acc += R.v[p] * v[R.ci[p]];
        }
        for (int o = 16; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 32);
        if (row < R.nrows && lane == 0) y[row] = acc;
    } else {
        const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
        const int64_t row = t / NC;
        const int j = (int)(t % NC);
        if (row >= R.nrows) return;
        T acc = T(0);
        if (R.hv) {
            for (int64_t p = R.rp[row]; p < R.rp[row + 1]; ++p) acc += (T)R.hv[p] * v[(int64_t)R.ci[p] * NC + j];
            acc *= (T)R.hscale;
        } else
            for (int64_t p = R.rp[row]; p < R.rp[row + 1]; ++p) acc += R.v[p] * v[(int64_t)R.ci[p] * NC + j];
        y[t] = acc;
    }
}

// The next level's right-hand side straight from this level's iterate:
// y = Rb b - M (x + d) (d null: x complete), with M = R A, so the residual
// b - A (x + d) is never formed.  Both matrices are rows of 1 x B blocks.
// LPR lanes per row; each lane takes groups of U entries of both matrices,
// issuing the block columns of both, then the values and gathers of both,
// before adding (entries past a row re-read its last entry, in bounds, and
// add zero; rows are never empty); the lanes' sums combine in a fixed
// butterfly order.  RH, MH: values in half precision.
template <typename T, int B, int LPR, int U, bool RH, bool MH, bool HY>
__global__ __launch_bounds__(256) void k_rres(BcsrDev<T> R, BcsrDev<T> M, const T* __restrict__ b,
                                            const T* __restrict__ x, const T* __restrict__ d, T* __restrict__ y,
                                            const int* done, RowSkip sk) {
    if (done && *done) return;
    const int64_t row = ((int64_t)blockIdx.x * 256 + threadIdx.x) / LPR;
    const int lane = threadIdx.x % LPR;
    // a coarse row of a finished instance (an aggregate lies in one instance:
    // its first fine block row names it); the row's LPR lanes leave together
    if (sk.ti && row < M.nrows && skipped(sk, (int64_t)M.bc[M.rp[row]])) return;
    T acc = T(0);
    if (row < M.nrows) {
        T rb = T(0), mx = T(0);
        const int64_t r0 = R.rp[row], r1 = R.rp[row + 1], m0 = M.rp[row], m1 = M.rp[row + 1];
        for (int64_t p0 = r0 + lane, e0 = m0 + lane; p0 < r1 || e0 < m1; p0 += LPR * U, e0 += LPR * U) {
            int64_t pp[U], ee[U], cr[U], cm[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t p = p0 + LPR * u, e = e0 + LPR * u;
                pp[u] = p < r1 ? p : r1 - 1;
                ee[u] = e < m1 ? e : m1 - 1;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                cr[u] = (int64_t)R.bc[pp[u]] * B;
                cm[u] = (int64_t)M.bc[ee[u]] * B;
            }
            T bv[U][B], rv[U][B], xv[U][B], mv[U][B];
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int cc = 0; cc < B; ++cc) {
                    bv[u][cc] = b[cr[u] + cc];
                    if constexpr (RH)
                        rv[u][cc] = (T)R.hv[pp[u] * B + cc];
                    else
                        rv[u][cc] = R.v[pp[u] * B + cc];
                    T v = x[cm[u] + cc];
                    if constexpr (HY) v += d[cm[u] + cc];
                    xv[u][cc] = v;
                    if constexpr (MH)
                        mv[u][cc] = (T)M.hv[ee[u] * B + cc];
                    else
                        mv[u][cc] = M.v[ee[u] * B + cc];
                }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                T sr = T(0), sm = T(0);
#pragma unroll
                for (int cc = 0; cc < B; ++cc) {
                    sr += rv[u][cc] * bv[u][cc];
                    sm += mv[u][cc] * xv[u][cc];
                }
                rb += p0 + LPR * u < r1 ? sr : T(0);
                mx += e0 + LPR * u < m1 ? sm : T(0);
            }
        }
        if constexpr (RH) rb *= (T)R.hscale;
        if constexpr (MH) mx *= (T)M.hscale;
        acc = rb - mx;
    }
    for (int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
    if (row < M.nrows && lane == 0) y[row] = acc;
}

// The same residual-restriction on MCOLS interleaved columns: a thread per
// (coarse row, column), y = Rb b - M (x + d) over the row's entries in order
// (the residual b - A x is never formed, and M = R A gathers far fewer fine
// rows than A's rows followed by R's would).
template <typename T, int B, bool RH, bool MH, bool HY>
__global__ __launch_bounds__(256) void k_rres_mc(BcsrDev<T> R, BcsrDev<T> M, const T* __restrict__ b,
                                               const T* __restrict__ x, const T* __restrict__ d, T* __restrict__ y,
                                               const int* done) {
    if (done && *done) return;
    constexpr int NC = MCOLS;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t row = t / NC;
    const int j = (int)(t % NC);
    if (row >= M.nrows) return;
    T rb = T(0), mx = T(0);
    const int64_t r1 = R.rp[row + 1], m1 = M.rp[row + 1];
#pragma unroll 4
    for (int64_t p = R.rp[row]; p < r1; ++p) {
        const int64_t c = (int64_t)R.bc[p] * B;
#pragma unroll
        for (int cc = 0; cc < B; ++cc) {
            const T rv = RH ? (T)R.hv[p * B + cc] : R.v[p * B + cc];
            rb += rv * b[(c + cc) * NC + j];
        }
    }
#pragma unroll 4
    for (int64_t e = M.rp[row]; e < m1; ++e) {
        const int64_t c = (int64_t)M.bc[e] * B;
#pragma unroll
        for (int cc = 0; cc < B; ++cc) {
            T v = x[(c + cc) * NC + j];
            if constexpr (HY) v += d[(c + cc) * NC + j];
            const T mv = MH ? (T)M.hv[e * B + cc] : M.v[e * B + cc];
            mx += mv * v;
        }
    }
    if constexpr (RH) rb *= (T)R.hscale;
    if constexpr (MH) mx *= (T)M.hscale;
    y[row * NC + j] = rb - mx;
}

template <typename T, int B>
static void launch_rres_mc(const AmgLevelDev<T>& L, const T* b, const T* x, const T* d, T* y, hipStream_t st,
                           const int* ms) {
    const dim3 g((unsigned)((L.M.nrows * MCOLS + 255) / 256)), bl(256);
    const bool rh = L.Rb.hv != nullptr, mh = L.M.hv != nullptr;
#define VEMD_RRES_MC(RH, MH, HY) \
    hipLaunchKernelGGL((k_rres_mc<T, B, RH, MH, HY>), g, bl, 0, st, L.Rb, L.M, b, x, d, y, ms)
    if (d) {
        if (rh && mh) VEMD_RRES_MC(true, true, true);
        else if (rh) VEMD_RRES_MC(true, false, true);
        else if (mh) VEMD_RRES_MC(false, true, true);
        else VEMD_RRES_MC(false, false, true);
    } else {
        if (rh && mh) VEMD_RRES_MC(true, true, false);
        else if (rh) VEMD_RRES_MC(true, false, false);
        else if (mh) VEMD_RRES_MC(false, true, false);
        else VEMD_RRES_MC(false, false, false);
    }
#undef VEMD_RRES_MC
}

// Last Chebyshev step of the finest level (single column, one lane per block
// row, so a wavefront is one 64-position tile) with the Krylov output of
// KOut: x + d + d' (times xs on a scaled level) goes to the pressure block
// of z, with the tile's partials of b.z and z.r2.
template <typename T, int B>
__global__ __launch_bounds__(256) void k_step_z(Sell<T> A, const T* r, const T* dinv, const T* din, const T* x, T c1,
                                              T c2, const T* xs, const T* bcyc, KOut<T> ko, const int* done) {
    if (done && *done) return;
    const int64_t bi = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (bi >= A.nrows) return;  // whole wavefronts: the level has a multiple of 64 block rows
    const int64_t tile = bi >> 6;
    const MState* st = ko.st + (ko.tile_inst ? ko.tile_inst[tile] : 0);
    if (!ko.init && st->done) return;  // uniform over the tile
    const int it = ko.init ? 0 : st->itn;
    T* __restrict__ zo = ko.init ? ko.Z[0] : ko.Z[(it + 1) & 1];
    const T* __restrict__ r2 = ko.init ? nullptr : ko.R[it & 1];
    T ad[B];
    sell_row<T, B, 1>(A, bi, 0, din, ad);
    double d1 = 0.0, d2 = 0.0;
#pragma unroll
    for (int k = 0; k < B; ++k) {
        const int64_t i = bi * B + k;
        const T di = din[i];
        const T dn = c1 * di + c2 * dval(dinv, i) * (r[i] - ad[k]);
        const T xn = x[i] + di + dn;
        const T z = xs ? xn * xs[i] : xn;
        const int64_t kv = ko.pb + k * ko.cpad + bi;
        zo[kv] = z;
        d1 += (double)bcyc[i] * (double)z;
        if (r2) d2 += (double)z * (double)r2[kv];
    }
    for (int o = 32; o > 0; o >>= 1) {
        d1 += __shfl_xor(d1, o, 64);
        d2 += __shfl_xor(d2, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        ko.q1[tile] = d1;
        ko.q2[tile] = d2;
    }
}

// Coarse correction of the lagged iterate: x += d + P (xc + dc), with d this
// level's and dc the coarse level's pending Chebyshev update (null where the
// iterate is complete).
template <typename T, int NC>
__global__ __launch_bounds__(256) void k_prol(Sell<T> P, const T* xc, const T* dc, T* x, const T* d, const int* done,
                                            RowSkip sk) {
    const int64_t n = P.nrows;
    AMG_ROW
    if (skipped(sk, row / sk.bs)) return;
    T px[1];
    sell_row<T, 1, NC, 4>(P, row, (int)(i % NC), xc, px, dc);
    px[0] *= (T)P.hscale;  // 1 unless the values are stored scaled in half precision
    x[i] += d ? d[i] + px[0] : px[0];
}

template <typename T, int NC>
__global__ __launch_bounds__(256) void k_addto(int64_t n, T* x, const T* d, const T* xs, const int* done) {
    AMG_ROW
    const T xn = x[i] + d[i];
    x[i] = xs ? xn * xs[row] : xn;
}

// Coarsest level: x = C^-1 b with the dense inverse; one block per row, one
// wavefront reduction (NC = 1) or one thread per column.
template <typename T, int NC>
__global__ __launch_bounds__(64) void k_dense(int64_t n, const T* cinv, const T* b, T* x, const int* done) {
    if (done && *done) return;
    const int64_t i = blockIdx.x;
    if constexpr (NC == 1) {
        // groups of U loads issued together (the inverse is cold in the cache)
        constexpr int U = 8;
        double acc = 0.0;
        for (int64_t j0 = threadIdx.x; j0 < n; j0 += 64 * U) {
            T cv[U], bv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                // unconditional loads (a clamped index, in bounds), so none
                // waits for another; entries past n contribute zero
                const int64_t j = j0 + 64 * u < n ? j0 + 64 * u : n - 1;
                cv[u] = cinv[i * n + j];
                bv[u] = b[j];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) acc += j0 + 64 * u < n ? (double)cv[u] * (double)bv[u] : 0.0;
        }
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
        if (threadIdx.x == 0) x[i] = (T)acc;
    } else {
        for (int j = threadIdx.x; j < NC; j += 64) {
            double acc = 0.0;
            for (int64_t k = 0; k < n; ++k) acc += (double)cinv[i * n + k] * (double)b[k * NC + j];
            x[i * NC + j] = (T)acc;
        }
    }
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

// dsq non-null: half-precision values of D^-1/2 A D^-1/2 with dsq = D^1/2.
template <typename T>
// dsq: store D^-1/2 A D^-1/2 in half precision; half: store A's values
// divided by their largest magnitude in half precision (hscale keeps it).
static Sell<T> to_sell(AmgDev<T>& D, const Csr& A, int bs, const std::vector<double>* dsq = nullptr,
                       bool half = false, ValMap* vm = nullptr) {
    Sell<T> s;
    s.bs = bs;
    s.nrows = A.nrows / bs;
    s.nslices = (int)((s.nrows + 63) / 64);
    // rows below VEMD_MG_LANE_ROWS (default 32768) of a scalar level: 8 lanes each
    const char* lr = std::getenv("VEMD_MG_LANE_ROWS");
    const int64_t lane_rows = lr ? (int64_t)std::atoll(lr) : (int64_t)32768;
    s.lanes = bs == 1 && s.nrows < lane_rows ? LANES_SMALL : 1;
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
    const bool h16 = dsq || half;
    double amax = 0.0;
    if (half)
        for (double v : A.v) amax = std::max(amax, std::fabs(v));
    if (half && amax > 0.0) s.hscale = (float)amax;
    const double hinv = half && amax > 0.0 ? 1.0 / amax : 1.0;
    std::vector<T> val(h16 ? 0 : off[s.nslices] * bs * bs, T(0));
    std::vector<_Float16> hval(h16 ? off[s.nslices] * bs * bs : 0, (_Float16)0.0f);
    if (vm) {
        vm->n = (size_t)off[s.nslices] * bs * bs;
        vm->at.assign(A.nnz(), 0);
        if (vm->n > 0xffffffffull) throw Error(VEMD_ERR_MESH, "multigrid level too large for its value map");
    }
    for (int64_t br = 0; br < s.nrows; ++br) {
        const int64_t sl = br >> 6, l = br & 63;
        const std::vector<int>& v = bc[br];
        for (size_t j = 0; j < v.size(); ++j) col[off[sl] + (int64_t)j * 64 + l] = v[j];
        for (int k = 0; k < bs; ++k)
            for (int64_t p = A.rp[br * bs + k]; p < A.rp[br * bs + k + 1]; ++p) {
                const int64_t j = std::lower_bound(v.begin(), v.end(), A.ci[p] / bs) - v.begin();
                const int64_t at = (off[sl] + j * 64) * bs * bs + (k * bs + A.ci[p] % bs) * 64 + l;
                if (vm) vm->at[p] = (uint32_t)at;
                if (dsq)
                    hval[at] = (_Float16)(A.v[p] / ((*dsq)[br * bs + k] * (*dsq)[A.ci[p]]));
                else if (half)
                    hval[at] = (_Float16)(A.v[p] * hinv);
                else
                    val[at] = (T)A.v[p];
            }
    }
    s.off = amg_copy(D, off);
    s.width = amg_copy(D, width);
    s.col = amg_copy(D, col);
    if (h16)
        s.hval = amg_copy(D, hval);
    else
        s.val = amg_copy(D, val);
    return s;
}

template <typename T>
static CsrDev<T> to_csr(AmgDev<T>& D, const Csr& A, bool half = false) {
    CsrDev<T> c;
    c.nrows = A.nrows;
    c.rp = amg_copy(D, A.rp);
    c.ci = amg_copy(D, A.ci);
    if (half) {
        double amax = 0.0;
        for (double x : A.v) amax = std::max(amax, std::fabs(x));
        const double hinv = amax > 0.0 ? 1.0 / amax : 1.0;
        c.hscale = amax > 0.0 ? (float)amax : 1.0f;
        std::vector<_Float16> v(A.v.size());
        for (size_t i = 0; i < v.size(); ++i) v[i] = (_Float16)(A.v[i] * hinv);
        c.hv = amg_copy(D, v);
    } else {
        std::vector<T> v(A.v.begin(), A.v.end());
        c.v = amg_copy(D, v);
    }
    return c;
}

// Rows of 1 x bs blocks from compressed rows with sorted columns (missing
// components of a block are zero); with csc each column c is divided by
// csc[c] first; half: values divided by their largest magnitude in half
// precision.
template <typename T>
static BcsrDev<T> to_bcsr(AmgDev<T>& D, const Csr& A, int bs, const std::vector<double>* csc, bool half,
                          ValMap* vm = nullptr) {
    BcsrDev<T> m;
    m.nrows = A.nrows;
    m.bs = bs;
    std::vector<int64_t> rp(A.nrows + 1, 0);
    std::vector<int> bc;
    std::vector<double> v;
    bc.reserve(A.nnz() / bs + A.nrows);
    v.reserve(A.nnz() + A.nrows * bs);
    if (vm) vm->at.assign(A.nnz(), 0);
    for (int64_t i = 0; i < A.nrows; ++i) {
        int last = -1;
        for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            const int b = A.ci[p] / bs;
            if (b != last) {
                bc.push_back(b);
                v.resize(v.size() + bs, 0.0);
                last = b;
            }
            v[v.size() - bs + A.ci[p] % bs] = csc ? A.v[p] / (*csc)[A.ci[p]] : A.v[p];
            if (vm) vm->at[p] = (uint32_t)(v.size() - bs + A.ci[p] % bs);
        }
        rp[i + 1] = (int64_t)bc.size();
    }
    if (vm) {
        vm->n = v.size();
        if (vm->n > 0xffffffffull) throw Error(VEMD_ERR_MESH, "multigrid level too large for its value map");
    }
    m.rp = amg_copy(D, rp);
    m.bc = amg_copy(D, bc);
    if (half) {
        double amax = 0.0;
        for (double x : v) amax = std::max(amax, std::fabs(x));
        const double hinv = amax > 0.0 ? 1.0 / amax : 1.0;
        m.hscale = amax > 0.0 ? (float)amax : 1.0f;
        std::vector<_Float16> hv(v.size());
        for (size_t i = 0; i < v.size(); ++i) hv[i] = (_Float16)(v[i] * hinv);
        m.hv = amg_copy(D, hv);
    } else {
        std::vector<T> tv(v.begin(), v.end());
        m.v = amg_copy(D, tv);
    }
    return m;
}

// Half-precision storage (values divided by the largest magnitude) needs
// finite values; entries below 6e-5 of the largest lose relative precision,
// which changes the preconditioner by less than that.
static bool half_safe(const std::vector<const Csr*>& ms) {
    for (const Csr* m : ms)
        for (double v : m->v)
            if (!std::isfinite(v)) return false;
    return true;
}

// The residual restricted in one pass (k_rres) unless VEMD_MG_FUSED_RESTRICT=0.
static bool fused_restriction() {
    const char* e = std::getenv("VEMD_MG_FUSED_RESTRICT");
    return !(e && e[0] == '0');
}

// k_rres reads at least one entry of every row of R and M.
static bool rows_nonempty(const Csr& A) {
    for (int64_t i = 0; i < A.nrows; ++i)
        if (A.rp[i + 1] == A.rp[i]) return false;
    return true;
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
            const int bs = l == 0 ? H.bs0 : h.bs;
            if (h.A.nrows % bs) throw Error(VEMD_ERR_INTERNAL, "finest multigrid level is not blocked");
            // FP32 cycles run the finest level scaled to a unit diagonal with
            // its operator in half precision; VEMD_MG_HALF=0 keeps it in FP32
            const char* he = std::getenv("VEMD_MG_HALF");
            bool half = l == 0 && sizeof(T) == 4 && l + 1 < H.lv.size() && !(he && he[0] == '0');
            std::vector<double> dsq(half ? h.dinv.size() : 0);
            for (size_t i = 0; i < dsq.size(); ++i) {
                dsq[i] = 1.0 / std::sqrt(h.dinv[i]);
                if (!(h.dinv[i] > 0.0) || !std::isfinite(dsq[i])) half = false;
            }
            if (half) {
                d.A = to_sell(*D, h.A, bs, &dsq, false, &d.vA);
                Csr P = h.P, R = h.R;  // P <- D^1/2 P, R <- R D^1/2
                for (int64_t i = 0; i < P.nrows; ++i)
                    for (int64_t p = P.rp[i]; p < P.rp[i + 1]; ++p) P.v[p] *= dsq[i];
                for (int64_t p = 0; p < R.nnz(); ++p) R.v[p] *= dsq[R.ci[p]];
                // the transfers in half precision too where their values fit
                // (VEMD_MG_HALF_TRANSFER=0 keeps them in FP32)
                const char* ht = std::getenv("VEMD_MG_HALF_TRANSFER");
                const bool htr = !(ht && ht[0] == '0') && half_safe({&P, &R});
                d.P = to_sell(*D, P, 1, nullptr, htr, &d.vP);
                d.R = to_csr(*D, R, htr);
                d.vR.n = R.v.size();
                if (fused_restriction() && rows_nonempty(h.M) && rows_nonempty(h.R)) {
                    // M = R D^1/2 (D^-1/2 A D^-1/2) = (R A) D^-1/2 on the scaled
                    // iterate; the plain R on the unscaled right-hand side
                    d.M = to_bcsr(*D, h.M, bs, &dsq, htr && half_safe({&h.M}), &d.vM);
                    d.Rb = to_bcsr(*D, h.R, bs, nullptr, htr && half_safe({&h.R}), &d.vRb);
                }
                if (std::getenv("VEMD_VERBOSE"))
                    std::fprintf(stderr, "[vemd] finest-level transfers of the FP32 cycle in %s\n",
                                 htr ? "half precision" : "FP32 (values outside half precision's range)");
                std::vector<T> sc(dsq.size());
                for (size_t i = 0; i < dsq.size(); ++i) sc[i] = (T)(1.0 / dsq[i]);
                d.sc = amg_copy(*D, sc);
                d.dinv = nullptr;
            } else {
                d.A = to_sell(*D, h.A, bs, nullptr, false, &d.vA);
                if (l + 1 < H.lv.size()) {
                    d.P = to_sell(*D, h.P, 1, nullptr, false, &d.vP);
                    d.R = to_csr(*D, h.R);
                    d.vR.n = h.R.v.size();
                    if (fused_restriction() && rows_nonempty(h.M) && rows_nonempty(h.R)) {
                        d.M = to_bcsr(*D, h.M, bs, nullptr, false, &d.vM);
                        d.Rb = to_bcsr(*D, h.R, bs, nullptr, false, &d.vRb);
                    }
                }
                std::vector<T> di(h.dinv.begin(), h.dinv.end());
                d.dinv = amg_copy(*D, di);
            }
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

// New values of one device matrix: vals[p] for CSR entry p goes to index
// vm.at[p] (vm.at empty: p) of the device array dev, in T or, with h16, in
// half precision -- divided by the largest magnitude when normalise (hscale
// receives it), the conversions of to_sell / to_csr / to_bcsr.
template <typename T>
static void refill(const ValMap& vm, const std::vector<double>& vals, bool h16, bool normalise, const void* dev,
                   float* hscale) {
    double amax = 0.0;
    if (normalise)
        for (double v : vals) amax = std::max(amax, std::fabs(v));
    const double hinv = normalise && amax > 0.0 ? 1.0 / amax : 1.0;
    if (hscale) *hscale = normalise && amax > 0.0 ? (float)amax : 1.0f;
    const int64_t nnz = (int64_t)vals.size();
    auto at = [&](int64_t p) { return vm.at.empty() ? p : (int64_t)vm.at[p]; };
    if (h16) {
        std::vector<_Float16> out(vm.n, (_Float16)0.0f);
        parallel_for(nnz, [&](int64_t b, int64_t e) {
            for (int64_t p = b; p < e; ++p) out[at(p)] = (_Float16)(vals[p] * hinv);
        });
        HIPCHK(hipMemcpy(const_cast<void*>(dev), out.data(), out.size() * sizeof(_Float16), hipMemcpyHostToDevice));
    } else {
        std::vector<T> out(vm.n, T(0));
        parallel_for(nnz, [&](int64_t b, int64_t e) {
            for (int64_t p = b; p < e; ++p) out[at(p)] = (T)vals[p];
        });
        HIPCHK(hipMemcpy(const_cast<void*>(dev), out.data(), out.size() * sizeof(T), hipMemcpyHostToDevice));
    }
}

// Values of a CSR matrix with each entry scaled: by rs[row] and/or cs[col]
// (multiplied), or divided by both when div.
static std::vector<double> scaled_values(const Csr& A, const std::vector<double>* rs, const std::vector<double>* cs,
                                         bool div) {
    std::vector<double> v(A.v.size());
    parallel_for(A.nrows, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i)
            for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
                double x = A.v[p];
                if (div) {
                    double d = 1.0;
                    if (rs) d = (*rs)[i];
                    if (cs) d = rs ? d * (*cs)[A.ci[p]] : (*cs)[A.ci[p]];
                    x = x / d;
                } else {
                    if (rs) x *= (*rs)[i];
                    if (cs) x *= (*cs)[A.ci[p]];
                }
                v[p] = x;
            }
    });
    return v;
}

template <typename T>
bool amg_refresh(AmgDev<T>& D, const Amg& H) {
    if (D.lv.size() != H.lv.size() || D.nc != H.ncoarse) return false;
    // every shape and storage decision first, so that a mismatch changes nothing
    for (size_t l = 0; l < H.lv.size(); ++l) {
        const AmgLevel& h = H.lv[l];
        const AmgLevelDev<T>& d = D.lv[l];
        const bool last = l + 1 == H.lv.size();
        if (h.A.nrows != d.n || h.A.nnz() != (int64_t)d.vA.at.size()) return false;
        const char* he = std::getenv("VEMD_MG_HALF");
        bool half = l == 0 && sizeof(T) == 4 && !last && !(he && he[0] == '0');
        for (size_t i = 0; half && i < h.dinv.size(); ++i)
            if (!(h.dinv[i] > 0.0) || !std::isfinite(1.0 / std::sqrt(h.dinv[i]))) half = false;
        if (half != (d.sc != nullptr)) return false;
        if (last) continue;
        if (h.P.nnz() != (int64_t)d.vP.at.size() || h.R.nnz() != (int64_t)d.vR.n) return false;
        const bool hasM = d.M.nrows > 0;
        if (hasM && (h.M.nnz() != (int64_t)d.vM.at.size() || h.R.nnz() != (int64_t)d.vRb.at.size())) return false;
        if (half) {
            const bool htr = d.P.hval != nullptr;
            if (htr != (d.R.hv != nullptr) || (htr && !half_safe({&h.P, &h.R}))) return false;
            if (hasM && ((d.M.hv != nullptr) != (htr && half_safe({&h.M})) ||
                         (d.Rb.hv != nullptr) != (htr && half_safe({&h.R}))))
                return false;
        }
    }
    for (size_t l = 0; l < H.lv.size(); ++l) {
        const AmgLevel& h = H.lv[l];
        AmgLevelDev<T>& d = D.lv[l];
        const bool last = l + 1 == H.lv.size();
        if (d.sc) {
            std::vector<double> dsq(h.dinv.size());
            for (size_t i = 0; i < dsq.size(); ++i) dsq[i] = 1.0 / std::sqrt(h.dinv[i]);
            refill<T>(d.vA, scaled_values(h.A, &dsq, &dsq, true), true, false, d.A.hval, nullptr);
            const bool htr = d.P.hval != nullptr;
            refill<T>(d.vP, scaled_values(h.P, &dsq, nullptr, false), htr, htr,
                      htr ? (const void*)d.P.hval : (const void*)d.P.val, &d.P.hscale);
            refill<T>(d.vR, scaled_values(h.R, nullptr, &dsq, false), htr, htr,
                      htr ? (const void*)d.R.hv : (const void*)d.R.v, &d.R.hscale);
            if (d.M.nrows > 0) {
                const bool mh = d.M.hv != nullptr, rh = d.Rb.hv != nullptr;
                refill<T>(d.vM, scaled_values(h.M, nullptr, &dsq, true), mh, mh,
                          mh ? (const void*)d.M.hv : (const void*)d.M.v, &d.M.hscale);
                refill<T>(d.vRb, h.R.v, rh, rh, rh ? (const void*)d.Rb.hv : (const void*)d.Rb.v, &d.Rb.hscale);
            }
            std::vector<T> sc(dsq.size());
            for (size_t i = 0; i < dsq.size(); ++i) sc[i] = (T)(1.0 / dsq[i]);
            HIPCHK(hipMemcpy(d.sc, sc.data(), sc.size() * sizeof(T), hipMemcpyHostToDevice));
        } else {
            refill<T>(d.vA, h.A.v, false, false, d.A.val, nullptr);
            if (!last) {
                refill<T>(d.vP, h.P.v, false, false, d.P.val, nullptr);
                refill<T>(d.vR, h.R.v, false, false, d.R.v, nullptr);
                if (d.M.nrows > 0) {
                    refill<T>(d.vM, h.M.v, false, false, d.M.v, nullptr);
                    refill<T>(d.vRb, h.R.v, false, false, d.Rb.v, nullptr);
                }
            }
            std::vector<T> di(h.dinv.begin(), h.dinv.end());
            HIPCHK(hipMemcpy(d.dinv, di.data(), di.size() * sizeof(T), hipMemcpyHostToDevice));
        }
        const double hi = h.lmax, lo = h.lmax / H.opt.cheb_ratio;
        d.centre = 0.5 * (hi + lo);
        d.half = 0.5 * (hi - lo);
        d.degree = l == 0 ? std::max(1, H.opt.degree0) : std::max(1, H.opt.degree);
    }
    std::vector<T> ci(H.coarse_inv.begin(), H.coarse_inv.end());
    HIPCHK(hipMemcpy(D.cinv, ci.data(), ci.size() * sizeof(T), hipMemcpyHostToDevice));
    return true;
}

// Vectors of one level for NC columns.
template <typename T>
struct LevelVecs {
    T *b, *x, *r, *d0, *d1;
};
template <typename T, int NC>
static LevelVecs<T> vecs(const AmgLevelDev<T>& L) {
    if constexpr (NC == 1)
        return {L.b, L.x, L.r, L.d0, L.d1};
    else
        return {L.mb, L.mx, L.mr, L.md0, L.md1};
}

static dim3 grid_of(int64_t threads) { return dim3((unsigned)((threads + 255) / 256)); }

// The Chebyshev steps of one sweep after its first update: steps 2..degree
// (fused = false: the first update d = din is already formed) or 1..degree
// with the first update formed on the fly from x = 0 (fused = true).  The
// last step completes the iterate, so on return din is null; with a single
// step and no fused start the update stays pending in din.
template <typename T, int B, int NC>
static void smooth_steps(const AmgLevelDev<T>& L, const LevelVecs<T>& v, hipStream_t st, const int* ms, T*& din,
                         bool fused, const T* xs, bool from_b = false, const KOut<T>* ko = nullptr,
                         RowSkip sk = RowSkip{}) {
    const double sigma = L.centre / L.half;
    double rho = 1.0 / sigma;
    T* dout = din == v.d0 ? v.d1 : v.d0;
    for (int k = 1; k < L.degree; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        const int last = k + 1 == L.degree;
        if (last && ko && NC == 1) {
            hipLaunchKernelGGL((k_step_z<T, B>), grid_of(L.A.nrows), dim3(256), 0, st, L.A, (const T*)v.r,
                               (const T*)L.dinv, (const T*)din, (const T*)v.x, (T)(rn * rho), (T)(2.0 * rn / L.half),
                               xs, (const T*)v.b, *ko, ms);
        } else if (fused && k == 1) {
            hipLaunchKernelGGL((k_step0<T, B, NC>), grid_of(L.A.nrows * NC * lanes_per_row<B, NC>(L.A)), dim3(256), 0, st,
                               L.A, (const T*)v.b,
                               (const T*)L.dinv, v.r, din, v.x, (T)(1.0 / L.centre), (T)(rn * rho),
                               (T)(2.0 * rn / L.half), last, ms, sk);
        } else {
            const bool b0 = from_b && k == 1;
            hipLaunchKernelGGL((k_step<T, B, NC>), grid_of(L.A.nrows * NC * lanes_per_row<B, NC>(L.A)), dim3(256), 0, st,
                               L.A, v.r,
                               (const T*)L.dinv, (const T*)din, dout, v.x, (T)(rn * rho), (T)(2.0 * rn / L.half),
                               last, last ? xs : nullptr, b0 ? (const T*)v.b : nullptr, b0 ? (const T*)L.sc : nullptr,
                               ms, sk);
            std::swap(din, dout);
        }
        if (last) din = nullptr;
        rho = rn;
    }
}

template <typename T, int NC>
static const T* vcycle_level(const AmgDev<T>& D, size_t l, hipStream_t st, const int* ms, bool pre_done = false,
                             const KOut<T>* ko = nullptr);

// Pre-smoothing, residual, restriction, coarse cycle, prolongation and
// post-smoothing of one level whose operator has blocks of B rows.  The
// result is x + d with d the returned pending update: null unless the level
// smooths with a single step (then coarser levels leave the sum to the
// prolongation that reads them and the finest level adds it).
// y = Rb b - M (x + d) of the level (k_rres), specialised on which values
// are in half precision and on the pending update.
// Block levels (the finest, most rows): 16 lanes per row with 8 entries each in
// flight, so that fewer rounds of wavefronts pass the chain of dependent
// loads; scalar levels: 32 lanes, 4 entries each.
template <typename T, int B, int LPR, int U>
static void launch_rres_(const AmgLevelDev<T>& L, const T* b, const T* x, const T* d, T* y, hipStream_t st,
                         const int* ms, RowSkip sk) {
    const dim3 g = grid_of(L.M.nrows * LPR), bl(256);
    const bool rh = L.Rb.hv != nullptr, mh = L.M.hv != nullptr;
#define VEMD_RRES(RH, MH, HY) \
    hipLaunchKernelGGL((k_rres<T, B, LPR, U, RH, MH, HY>), g, bl, 0, st, L.Rb, L.M, b, x, d, y, ms, sk)
    if (d) {
        if (rh && mh) VEMD_RRES(true, true, true);
        else if (rh) VEMD_RRES(true, false, true);
        else if (mh) VEMD_RRES(false, true, true);
        else VEMD_RRES(false, false, true);
    } else {
        if (rh && mh) VEMD_RRES(true, true, false);
        else if (rh) VEMD_RRES(true, false, false);
        else if (mh) VEMD_RRES(false, true, false);
        else VEMD_RRES(false, false, false);
    }
#undef VEMD_RRES
}

// 32 lanes per row, 2 entries of each matrix in flight per lane (measured at
// 1M cells against 16 and 64 lanes and 1, 4 or 8 entries).
template <typename T, int B>
static void launch_rres(const AmgLevelDev<T>& L, const T* b, const T* x, const T* d, T* y, hipStream_t st,
                        const int* ms, RowSkip sk) {
    launch_rres_<T, B, 32, 2>(L, b, x, d, y, st, ms, sk);
}

template <typename T, int B, int NC>
static const T* level_cycle(const AmgDev<T>& D, size_t l, hipStream_t st, const int* ms, bool pre_done,
                            const KOut<T>* ko) {
    const AmgLevelDev<T>& L = D.lv[l];
    const AmgLevelDev<T>& C = D.lv[l + 1];
    const LevelVecs<T> v = vecs<T, NC>(L), vc = vecs<T, NC>(C);
    const dim3 g = grid_of(L.n * NC), gb = grid_of(L.A.nrows * NC * lanes_per_row<B, NC>(L.A)), b(256);
    const dim3 gr = NC == 1 ? grid_of(C.n * 32) : grid_of(C.n * NC);
    const T c0 = (T)(1.0 / L.centre);
    T* din = v.d0;
    // scalar levels form the first update inside the first step; block levels
    // (gathers of B consecutive entries) are cheaper with it materialised
    const bool fused = B == 1 && L.degree >= 2 && !L.sc;
    const T* sc = L.sc;
    // finished instances of a batch: their finest-level rows are left alone
    // after a pass's first cycle
    RowSkip sk;
    if (NC == 1 && l == 0 && ko && ko->tile_inst && !ko->init) sk = RowSkip{ko->tile_inst, ko->st, B};
    // pre_done: the caller has written d0 = b / centre (D^-1 or sc applied)
    if (!fused && !pre_done)
        hipLaunchKernelGGL((k_pre0<T, NC>), g, b, 0, st, L.n, (const T*)v.b, sc, (const T*)L.dinv, v.r, v.d0, v.x, c0,
                           ms);
    smooth_steps<T, B, NC>(L, v, st, ms, din, fused, (const T*)nullptr, pre_done, nullptr, sk);
    if (NC == 1 && L.M.nrows > 0) {
        launch_rres<T, B>(L, (const T*)v.b, (const T*)v.x, (const T*)din, vc.b, st, ms, sk);
    } else if (NC == MCOLS && L.M.nrows > 0) {
        launch_rres_mc<T, B>(L, (const T*)v.b, (const T*)v.x, (const T*)din, vc.b, st, ms);
    } else {
        hipLaunchKernelGGL((k_resid<T, B, NC>), gb, b, 0, st, L.A, (const T*)v.b, sc, (const T*)v.x, (const T*)din,
                           v.r, ms);
        hipLaunchKernelGGL((k_restrict<T, NC>), gr, b, 0, st, L.R, (const T*)v.r, vc.b, ms);
    }
    const T* dc = vcycle_level<T, NC>(D, l + 1, st, ms);
    hipLaunchKernelGGL((k_prol<T, NC>), g, b, 0, st, L.P, (const T*)vc.x, dc, v.x, (const T*)din, ms, sk);
    din = v.d0;
    hipLaunchKernelGGL((k_post0<T, B, NC>), gb, b, 0, st, L.A, (const T*)v.b, sc, (const T*)v.x, (const T*)L.dinv,
                       v.r, v.d0, c0, ms, sk);
    smooth_steps<T, B, NC>(L, v, st, ms, din, false, sc, false, ko, sk);
    if (l > 0 || !din) return din;
    hipLaunchKernelGGL((k_addto<T, NC>), g, b, 0, st, L.n, v.x, (const T*)din, sc, ms);
    return nullptr;
}

template <typename T, int NC>
static const T* vcycle_level(const AmgDev<T>& D, size_t l, hipStream_t st, const int* ms, bool pre_done,
                             const KOut<T>* ko) {
    const AmgLevelDev<T>& L = D.lv[l];
    if (l + 1 == D.lv.size()) {
        const LevelVecs<T> v = vecs<T, NC>(L);
        hipLaunchKernelGGL((k_dense<T, NC>), dim3((unsigned)D.nc), dim3(64), 0, st, D.nc, (const T*)D.cinv,
                           (const T*)v.b, v.x, ms);
        return nullptr;
    }
    switch (L.A.bs) {
        case 1: return level_cycle<T, 1, NC>(D, l, st, ms, pre_done, ko);
        case 3: return level_cycle<T, 3, NC>(D, l, st, ms, pre_done, ko);
        case 6:  // k = 2 (single column only: the multi-column path is k <= 1)
            if constexpr (NC == 1) return level_cycle<T, 6, NC>(D, l, st, ms, pre_done, ko);
            [[fallthrough]];
        default: throw Error(VEMD_ERR_INTERNAL, "unsupported multigrid block size");
    }
}

template <typename T>
bool amg_pre_fusable(const AmgDev<T>& D) {
    return D.lv.size() > 1 && D.lv[0].degree >= 2 && !(D.lv[0].A.bs == 1 && !D.lv[0].sc);
}

// The Krylov output needs the finest level's last step to be a k_step_z
// launch: a Chebyshev degree of at least 2, one lane per block row, whole
// 64-row tiles, and a coarser level (a single-level hierarchy ends in the
// dense solve).
template <typename T>
bool amg_kout_fusable(const AmgDev<T>& D) {
    const char* e = std::getenv("VEMD_MG_FUSED_DOTS");
    if (e && e[0] == '0') return false;
    const AmgLevelDev<T>& L = D.lv[0];
    return D.lv.size() > 1 && L.degree >= 2 && lanes_per_row<1, 1>(L.A) == 1 && L.A.nrows % 64 == 0;
}

template <typename T>
void amg_vcycle(const AmgDev<T>& D, hipStream_t st, const int* ms, bool pre_done, const KOut<T>* ko) {
    if (pre_done && !amg_pre_fusable(D)) throw Error(VEMD_ERR_INTERNAL, "multigrid cycle cannot take a formed start");
    if (ko && !amg_kout_fusable(D)) throw Error(VEMD_ERR_INTERNAL, "multigrid cycle cannot end in the Krylov output");
    vcycle_level<T, 1>(D, 0, st, ms, pre_done, ko);
    HIPCHK(hipGetLastError());
}

template <typename T>
void amg_multi_alloc(AmgDev<T>& D) {
    if (D.multi) return;
    for (AmgLevelDev<T>& L : D.lv) {
        const size_t vb = (size_t)L.n * MCOLS * sizeof(T);
        L.mb = (T*)amg_alloc(D, vb);
        L.mx = (T*)amg_alloc(D, vb);
        L.mr = (T*)amg_alloc(D, vb);
        L.md0 = (T*)amg_alloc(D, vb);
        L.md1 = (T*)amg_alloc(D, vb);
        for (T* p : {L.mb, L.mx, L.mr, L.md0, L.md1}) HIPCHK(hipMemset(p, 0, vb));
    }
    D.multi = true;
}

template <typename T>
void amg_vcycle_multi(const AmgDev<T>& D, hipStream_t st, const int* ms, bool pre_done) {
    if (!D.multi) throw Error(VEMD_ERR_STATE, "multi-column cycle buffers not allocated");
    if (pre_done && !amg_pre_fusable(D)) throw Error(VEMD_ERR_INTERNAL, "multigrid cycle cannot take a formed start");
    vcycle_level<T, MCOLS>(D, 0, st, ms, pre_done);
    HIPCHK(hipGetLastError());
}

template struct AmgDev<double>;
template struct AmgDev<float>;
template AmgDev<double>* amg_upload<double>(const Amg&);
template AmgDev<float>* amg_upload<float>(const Amg&);
template bool amg_refresh<double>(AmgDev<double>&, const Amg&);
template bool amg_refresh<float>(AmgDev<float>&, const Amg&);
template void amg_vcycle<double>(const AmgDev<double>&, hipStream_t, const int*, bool, const KOut<double>*);
template void amg_vcycle<float>(const AmgDev<float>&, hipStream_t, const int*, bool, const KOut<float>*);
template bool amg_pre_fusable<double>(const AmgDev<double>&);
template bool amg_pre_fusable<float>(const AmgDev<float>&);
template bool amg_kout_fusable<double>(const AmgDev<double>&);
template bool amg_kout_fusable<float>(const AmgDev<float>&);
template void amg_multi_alloc<double>(AmgDev<double>&);
template void amg_multi_alloc<float>(AmgDev<float>&);
template void amg_vcycle_multi<double>(const AmgDev<double>&, hipStream_t, const int*, bool);
template void amg_vcycle_multi<float>(const AmgDev<float>&, hipStream_t, const int*, bool);

}  // namespace vemd
