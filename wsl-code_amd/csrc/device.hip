// Device path of libvemd for gfx950: HBM operator store, one-launch matrix-free
// apply with two-sided edge slots, fused preconditioned MINRES with on-device
// scalar recurrences, and FP64 iterative refinement around an FP32 inner solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>

#include "apply.h"
#include "devstate.h"
#include "mgbuild.h"


namespace vemd {

// Entry i of the edge-dof block: the sum of the edge's two cell sides (a
// boundary edge's second slot is the zero slot).
// Edge of the edge-dof entry i (kk moments per edge).
__device__ __forceinline__ int64_t edge_of(int64_t i, int kk) {
    return kk == 2 ? i >> 1 : (kk == 1 ? i : (kk == 4 ? i >> 2 : i / kk));
}

template <typename T>
__device__ __forceinline__ T side_sum(const T* __restrict__ h, const int* __restrict__ hside, int kk, int64_t i) {
    const int64_t e = edge_of(i, kk);
    const int r = (int)(i - e * kk);
    return h[(int64_t)hside[2 * e] * kk + r] + h[(int64_t)hside[2 * e + 1] * kk + r];
}

// y[i] = side sum over the edge dofs (i < n); the padding lanes' slot of the
// vector (k + 1 entries at trash) is zeroed.
template <typename T>
__global__ __launch_bounds__(256) void k_combine(T* __restrict__ y, const T* __restrict__ h, const int* __restrict__ hside,
                                                 int kk, int64_t n, int64_t trash) {
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) y[i] = side_sum(h, hside, kk, i);
    if (blockIdx.x == 0 && (int)threadIdx.x < kk) y[trash + threadIdx.x] = T(0);
}

// Sums n partials in a fixed order with one block of 256 threads.
__device__ double block_sum_fixed(const double* p, int n) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += p[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// Multiplier rows of a plain apply, one block per instance: y_lambda of
// instance i = the sum of the per-tile partials of its tiles.
template <typename T>
__global__ void k_lam_row(const double* plam, const int* inst_tile0, T* y, int64_t lam_idx) {
    const int t0 = inst_tile0[blockIdx.x];
    double s = block_sum_fixed(plam + t0, inst_tile0[blockIdx.x + 1] - t0);
    if (threadIdx.x == 0) y[lam_idx + blockIdx.x] = (T)s;
}

// Sums of the partials src[k][b..e) (a null array sums to zero) in a fixed
// order by one block of FIN threads; every thread gets the results. The
// partials were written by other kernels, mostly on other XCDs, so each load
// misses this L2: every thread issues all its loads (32 / NA per array) before
// adding, and the block sum is two shuffle trees around one LDS exchange.
constexpr int FIN = 1024;
// Array k sums its range [b[k], e[k]).
template <int NA>
__device__ void block_sums_r(const double* const (&src)[NA], const int (&b)[NA], const int (&e)[NA],
                             double (&out)[NA]) {
    constexpr int U = 32 / NA, NW = FIN / 64;
    __shared__ double sh[NA][NW];
    double acc[NA];
    int len = 0;
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        acc[k] = 0.0;
        len = e[k] - b[k] > len ? e[k] - b[k] : len;
    }
    for (int i0 = (int)threadIdx.x; i0 < len; i0 += FIN * U) {
        double v[NA][U];
#pragma unroll
        for (int k = 0; k < NA; ++k)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + u * FIN;
                v[k][u] = src[k] && i < e[k] - b[k] ? src[k][b[k] + i] : 0.0;
            }
#pragma unroll
        for (int k = 0; k < NA; ++k)
#pragma unroll
            for (int u = 0; u < U; ++u) acc[k] += v[k][u];
    }
    const int w = (int)threadIdx.x >> 6, l = (int)threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        double x = acc[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (l == 0) sh[k][w] = x;
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
        for (int k = 0; k < NA; ++k) {
            double x = l < NW ? sh[k][l] : 0.0;
            for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
            if (l == 0) sh[k][0] = x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NA; ++k) out[k] = sh[k][0];
    __syncthreads();
}

template <int NA>
__device__ void block_sums_n(const double* const (&src)[NA], int b, int e, double (&out)[NA]) {
    int bb[NA], ee[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        bb[k] = b;
        ee[k] = e;
    }
    block_sums_r<NA>(src, bb, ee, out);
}

__device__ void block_sums(const double* a, const double* c, int b, int e, double& sa, double& sc) {
    const double* src[2] = {a, c};
    double out[2];
    block_sums_n<2>(src, b, e, out);
    sa = out[0];
    sc = out[1];
}

// The Krylov recurrences run per batch instance: every instance has its own
// MState; per-tile partials of the apply and per-chunk partials of the vector
// sweeps are summed over the instance's contiguous tile and chunk ranges by
// one block per instance.

// After the apply of iteration itn: alfa and the Lanczos coefficients.
template <typename T>
__global__ void k_finA(MState* sts, const double* pdot, const double* plam, const int* inst_tile0, T* Az,
                       const T* Z0, const T* Z1, int64_t lam_idx) {
    MState* st = sts + blockIdx.x;
    // the state, the stop flag and the partials are loaded together (one
    // round trip); a finished instance leaves after the sums (a uniform exit)
    MState sv;
    if (threadIdx.x == 0) sv = *st;
    const int done = st->done;
    double zaz, yl;
    block_sums(pdot, lam_idx >= 0 ? plam : nullptr, inst_tile0[blockIdx.x], inst_tile0[blockIdx.x + 1], zaz, yl);
    if (done) return;
    if (threadIdx.x == 0) {
        if (lam_idx >= 0) {  // the instance's multiplier row
            const T* z = (sv.itn & 1) ? Z1 : Z0;
            Az[lam_idx + blockIdx.x] = (T)yl;
            zaz += (double)z[lam_idx + blockIdx.x] * yl;
        }
        double beta = sv.beta;
        double alfa = zaz / (beta * beta) - (sv.itn > 0 ? sv.rho / sv.oldb : 0.0);
        st->alfa = alfa;
        st->c_az = 1.0 / beta;
        st->c_r1 = sv.itn > 0 ? beta / sv.oldb : 0.0;
        st->c_r2 = alfa / beta;
    }
}

template <typename T>
struct UpdArgs {
    T* R[2];
    T* Z[2];
    T* WB[2];
    T* x;
    const T* Az;         // cell-owned and multiplier rows of the apply
    const T* H;          // cell sides of the edge dofs
    const int* hside;
    int kk;
    const T* minv;
    int64_t edof;
    const VecChunk* chunks;
    double* p1;          // per-chunk partials
    double* p2;
    MState* st;          // per instance
    int amg;             // pressure chunks are preconditioned by the multigrid cycle
    int64_t pb;          // first entry of the pressure block
    T* bp;               // cycle input, cell-major: bp[gpos * q + r]
    int64_t cpad;
    int q;
    T* bd0;              // non-null: also the cycle's first update bd0 = c0 * bsc * bp
    const T* bsc;
    T c0;
};

// Cycle index of entry i of the pressure block (slot r = moment r).
template <typename T>
__device__ __forceinline__ int64_t cyc_index(const UpdArgs<T>& a, int64_t i, int r) {
    return (i - a.pb - (int64_t)r * a.cpad) * a.q + r;
}

// Block partials of two sums (256 threads), written by thread 0: a shuffle
// tree per wavefront, then the four wavefronts' sums in order.
__device__ __forceinline__ void block_pair(double d1, double d2, double* o1, double* o2) {
    __shared__ double sh[2][4];
    for (int o = 32; o > 0; o >>= 1) {
        d1 += __shfl_xor(d1, o, 64);
        d2 += __shfl_xor(d2, o, 64);
    }
    const int w = (int)threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][w] = d1;
        sh[1][w] = d2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *o1 = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
        if (o2) *o2 = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
    }
}

// Lanczos step, preconditioner, and the solution update of the previous
// iteration (delayed by one so that it needs no extra pass), in one sweep of
// one chunk per block.  Each thread takes the chunk's entries t, t + 256, ...
// (CHUNK / 256 of them) and issues every load of them -- the edge sides
// through the side table first -- before any arithmetic or store.
template <typename T>
__global__ __launch_bounds__(256) void k_update(UpdArgs<T> a) {
    constexpr int U = CHUNK / 256;
    const VecChunk ch = a.chunks[blockIdx.x];
    const MState* st = a.st + ch.inst;
    if (st->done) return;
    const int it = st->itn;
    T* __restrict__ r2 = a.R[it & 1];
    T* __restrict__ r1 = a.R[(it + 1) & 1];
    T* __restrict__ zo = a.Z[(it + 1) & 1];
    T* __restrict__ wn = a.WB[(it + 1) & 1];
    const T* __restrict__ wp = a.WB[it & 1];
    const T caz = (T)st->c_az, cr1 = (T)st->c_r1, cr2 = (T)st->c_r2;
    const bool pend = st->pending;
    const T uv = (T)st->u_v, uw1 = (T)st->u_w1, uw2 = (T)st->u_w2, ux = (T)st->u_x;
    const bool cyc = a.amg && ch.pslot;
    const int pr = ch.pslot - 1;
    const int64_t end = ch.start + ch.len;
    int64_t ii[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t i = ch.start + threadIdx.x + 256 * u;
        ok[u] = i < end;
        ii[u] = ok[u] ? i : ch.start;  // in bounds; the result is not stored
    }
    T az[U], r1v[U], r2v[U];
    if (ch.start < a.edof) {  // a chunk of edge dofs: the sum of the edge's two cell sides
        int2 sd[U];
        int rr[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = edge_of(ii[u], a.kk);
            rr[u] = (int)(ii[u] - e * a.kk);
            sd[u] = reinterpret_cast<const int2*>(a.hside)[e];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            az[u] = a.H[(int64_t)sd[u].x * a.kk + rr[u]] + a.H[(int64_t)sd[u].y * a.kk + rr[u]];
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) az[u] = a.Az[ii[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        r1v[u] = r1[ii[u]];
        r2v[u] = r2[ii[u]];
    }
    T zv[U], wnv[U], wpv[U], xv[U], mv[U], sv[U];
    if (pend) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            zv[u] = zo[ii[u]];
            wnv[u] = wn[ii[u]];
            wpv[u] = wp[ii[u]];
            xv[u] = a.x[ii[u]];
        }
    }
    if (!cyc) {
#pragma unroll
        for (int u = 0; u < U; ++u) mv[u] = a.minv[ii[u]];
    } else if (a.bd0) {
#pragma unroll
        for (int u = 0; u < U; ++u) sv[u] = a.bsc[cyc_index(a, ii[u], pr)];
    }
    double d1 = 0.0, d2 = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!ok[u]) continue;
        const int64_t i = ii[u];
        const T y = caz * az[u] - cr1 * r1v[u] - cr2 * r2v[u];
        if (pend) {
            const T w = uv * zv[u] - uw1 * wnv[u] - uw2 * wpv[u];
            wn[i] = w;
            a.x[i] = xv[u] + ux * w;
        }
        r1[i] = y;
        if (cyc) {
            const int64_t j = cyc_index(a, i, pr);
            a.bp[j] = y;  // preconditioned by the multigrid cycle, which also forms the dots
            if (a.bd0) a.bd0[j] = a.c0 * sv[u] * y;
        } else {
            const T zn = mv[u] * y;
            zo[i] = zn;
            d1 += (double)y * (double)zn;
            d2 += (double)zn * (double)r2v[u];
        }
    }
    block_pair(d1, d2, a.p1 + blockIdx.x, a.p2 + blockIdx.x);
}

// The same sweep in FP32 for two moments per edge (k = 1), by pairs of
// entries: thread t of a chunk takes entries 2p, 2p + 1 for p = t + 256 u,
// so every vector access is an 8-byte load or store per lane (the FP32 sweep
// with one entry per lane issues twice the instructions per byte of the FP64
// one), and in the edge chunks a pair is one edge's two moments: one side
// table entry and one 8-byte gather from each cell side.  Chunk starts are
// even except the one-entry chunk of an instance's multiplier; a chunk of odd
// length ends in a half pair.
__global__ __launch_bounds__(256) void k_update_pairs(UpdArgs<float> a) {
    constexpr int U = CHUNK / 512;
    const VecChunk ch = a.chunks[blockIdx.x];
    const MState* st = a.st + ch.inst;
    if (st->done) return;
    const int it = st->itn;
    float* __restrict__ r2 = a.R[it & 1];
    float* __restrict__ r1 = a.R[(it + 1) & 1];
    float* __restrict__ zo = a.Z[(it + 1) & 1];
    float* __restrict__ wn = a.WB[(it + 1) & 1];
    const float* __restrict__ wp = a.WB[it & 1];
    const float caz = (float)st->c_az, cr1 = (float)st->c_r1, cr2 = (float)st->c_r2;
    const bool pend = st->pending;
    const float uv = (float)st->u_v, uw1 = (float)st->u_w1, uw2 = (float)st->u_w2, ux = (float)st->u_x;
    const bool cyc = a.amg && ch.pslot;
    const int pr = ch.pslot - 1;
    const int64_t end = ch.start + ch.len;
    int64_t ii[U];
    bool ok0[U], ok1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t i = ch.start + 2 * ((int64_t)threadIdx.x + 256 * u);
        ok0[u] = i < end;
        ok1[u] = i + 1 < end;
        // lanes without a whole pair load the chunk's first pair (unused);
        // a one-entry chunk (an instance's multiplier: never an edge or
        // pressure chunk, whose indexing needs an entry of the chunk) loads
        // the vectors' first pair.  A half pair is done entry by entry below.
        ii[u] = ok1[u] ? i : (ch.len >= 2 ? ch.start : 0);
    }
    auto ld2 = [](const float* p, int64_t i) { return *reinterpret_cast<const float2*>(p + i); };
    float2 az[U], r1v[U], r2v[U];
    if (ch.start < a.edof) {  // a chunk of edge dofs: pair = the edge's two moments
        int2 sd[U];
#pragma unroll
        for (int u = 0; u < U; ++u) sd[u] = reinterpret_cast<const int2*>(a.hside)[ii[u] >> 1];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float2 h0 = ld2(a.H, (int64_t)sd[u].x * 2), h1 = ld2(a.H, (int64_t)sd[u].y * 2);
            az[u] = float2{h0.x + h1.x, h0.y + h1.y};
        }
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) az[u] = ld2(a.Az, ii[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        r1v[u] = ld2(r1, ii[u]);
        r2v[u] = ld2(r2, ii[u]);
    }
    float2 zv[U], wnv[U], wpv[U], xv[U], mv[U];
    if (pend) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            zv[u] = ld2(zo, ii[u]);
            wnv[u] = ld2(wn, ii[u]);
            wpv[u] = ld2(wp, ii[u]);
            xv[u] = ld2(a.x, ii[u]);
        }
    }
    float sv[U][2];
    if (!cyc) {
#pragma unroll
        for (int u = 0; u < U; ++u) mv[u] = ld2(a.minv, ii[u]);
    } else if (a.bd0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            sv[u][0] = a.bsc[cyc_index(a, ii[u], pr)];
            sv[u][1] = a.bsc[cyc_index(a, ii[u] + 1, pr)];
        }
    }
    double d1 = 0.0, d2 = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!ok0[u]) continue;
        if (!ok1[u]) {
            // the chunk's last entry without a partner (a multiplier chunk,
            // never an edge or pressure chunk)
            const int64_t i = ch.start + 2 * ((int64_t)threadIdx.x + 256 * u);
            const float y = caz * a.Az[i] - cr1 * r1[i] - cr2 * r2[i];
            if (pend) {
                const float w = uv * zo[i] - uw1 * wn[i] - uw2 * wp[i];
                wn[i] = w;
                a.x[i] += ux * w;
            }
            r1[i] = y;
            const float zn = a.minv[i] * y;
            zo[i] = zn;
            d1 += (double)y * (double)zn;
            d2 += (double)zn * (double)r2[i];
            continue;
        }
        const int64_t i = ii[u];
        const float y0 = caz * az[u].x - cr1 * r1v[u].x - cr2 * r2v[u].x;
        const float y1 = caz * az[u].y - cr1 * r1v[u].y - cr2 * r2v[u].y;
        if (pend) {
            const float w0 = uv * zv[u].x - uw1 * wnv[u].x - uw2 * wpv[u].x;
            const float w1 = uv * zv[u].y - uw1 * wnv[u].y - uw2 * wpv[u].y;
            *reinterpret_cast<float2*>(wn + i) = float2{w0, w1};
            *reinterpret_cast<float2*>(a.x + i) = float2{xv[u].x + ux * w0, xv[u].y + ux * w1};
        }
        *reinterpret_cast<float2*>(r1 + i) = float2{y0, y1};
        if (cyc) {
            const int64_t j0 = cyc_index(a, i, pr), j1 = cyc_index(a, i + 1, pr);
            a.bp[j0] = y0;  // preconditioned by the multigrid cycle, which also forms the dots
            a.bp[j1] = y1;
            if (a.bd0) {
                a.bd0[j0] = a.c0 * sv[u][0] * y0;
                a.bd0[j1] = a.c0 * sv[u][1] * y1;
            }
        } else {
            const float z0 = mv[u].x * y0, z1 = mv[u].y * y1;
            *reinterpret_cast<float2*>(zo + i) = float2{z0, z1};
            d1 += (double)y0 * (double)z0;
            d2 += (double)z0 * (double)r2v[u].x;
            d1 += (double)y1 * (double)z1;
            d2 += (double)z1 * (double)r2v[u].y;
        }
    }
    block_pair(d1, d2, a.p1 + blockIdx.x, a.p2 + blockIdx.x);
}

// Pressure chunks after the multigrid cycle: z_p = cycle output, with the
// partials of y.z and z.r2 (init: z of the right-hand side and b.z).
template <typename T>
__global__ __launch_bounds__(256) void k_pdot(UpdArgs<T> a, const int* pchunks, const T* xp, double* q1, double* q2,
                                            int init) {
    const int c = pchunks[blockIdx.x];
    const VecChunk ch = a.chunks[c];
    const MState* st = a.st + ch.inst;
    if (!init && st->done) return;
    const int it = init ? 0 : st->itn;
    T* __restrict__ zo = init ? a.Z[0] : a.Z[(it + 1) & 1];
    const T* __restrict__ r2 = a.R[it & 1];
    const int pr = ch.pslot - 1;
    double d1 = 0.0, d2 = 0.0;
    for (int64_t i = ch.start + threadIdx.x; i < ch.start + ch.len; i += 256) {
        const int64_t j = cyc_index(a, i, pr);
        const T z = xp[j];
        zo[i] = z;
        d1 += (double)a.bp[j] * (double)z;
        if (!init) d2 += (double)z * (double)r2[i];
    }
    block_pair(d1, d2, q1 + c, q2 + c);
}

// Givens recurrence of MINRES (Paige and Saunders) and the stopping test, one
// block per instance.
// With all_done non-null (a single instance) the block also publishes the
// instance's stop flag as the all-done flag.
// The pressure-block partials q1, q2 are per chunk (inst_q0 = inst_chunk0)
// or per tile (inst_q0 = inst_tile0) when the multigrid cycle wrote them.
__global__ void k_finB(MState* sts, const double* p1, const double* p2, const double* q1, const double* q2,
                       const int* inst_chunk0, int useq, int* all_done, const int* inst_q0) {
    MState* st = sts + blockIdx.x;
    // as in k_finA: state, stop flag and partials in one round trip
    MState sv;
    if (threadIdx.x == 0) sv = *st;
    const int done = st->done;
    const int c0 = inst_chunk0[blockIdx.x], c1 = inst_chunk0[blockIdx.x + 1];
    double b2, rho, qb2, qrho;
    {
        const int t0 = inst_q0[blockIdx.x], t1 = inst_q0[blockIdx.x + 1];
        const double* src[4] = {p1, p2, useq ? q1 : nullptr, useq ? q2 : nullptr};
        const int bb[4] = {c0, c0, t0, t0}, ee[4] = {c1, c1, t1, t1};
        double out[4];
        block_sums_r<4>(src, bb, ee, out);
        b2 = out[0];
        rho = out[1];
        qb2 = out[2];
        qrho = out[3];
    }
    if (done) return;
    b2 += qb2;
    rho += qrho;
    if (threadIdx.x == 0) {
        MState& s = sv;
        double beta_new = std::sqrt(b2 > 0.0 ? b2 : 0.0);
        s.rho = rho;
        s.oldb = s.beta;
        s.beta = beta_new;
        double oldeps = s.epsln;
        double delta = s.cs * s.dbar + s.sn * s.alfa;
        double gbar = s.sn * s.dbar - s.cs * s.alfa;
        s.epsln = s.sn * beta_new;
        s.dbar = -s.cs * beta_new;
        double gamma = std::sqrt(gbar * gbar + beta_new * beta_new);
        if (gamma < 1e-300) gamma = 1e-300;
        s.cs = gbar / gamma;
        s.sn = beta_new / gamma;
        double phi = s.cs * s.phibar;
        s.phibar = s.sn * s.phibar;
        s.u_v = 1.0 / (s.oldb * gamma);
        s.u_w1 = oldeps / gamma;
        s.u_w2 = delta / gamma;
        s.u_x = phi;
        s.pending = 1;
        s.itn += 1;
        if (s.phibar <= s.tol || s.itn >= s.maxit || !(beta_new > 0.0)) s.done = 1;
        *st = s;
        if (all_done) *all_done = s.done;
    }
}

// all_done = every instance done (lets the multigrid cycle skip its work).
__global__ void k_alldone(const MState* sts, int ninst, int* all_done) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < ninst; i += blockDim.x)
        if (!sts[i].done) any = 1;
    __syncthreads();
    if (threadIdx.x == 0) *all_done = any ? 0 : 1;
}

// r2 = b, r1 = 0, z = M^{-1} b, w = 0, x = 0 and partials of b.z, per chunk.
template <typename T>
__global__ __launch_bounds__(256) void k_init(UpdArgs<T> a, const T* b) {
    const VecChunk ch = a.chunks[blockIdx.x];
    const bool cyc = a.amg && ch.pslot;
    const int pr = ch.pslot - 1;
    double d = 0.0;
    for (int64_t i = ch.start + threadIdx.x; i < ch.start + ch.len; i += 256) {
        T bi = b[i];
        a.R[0][i] = bi;
        a.R[1][i] = T(0);
        a.Z[1][i] = T(0);
        a.WB[0][i] = T(0);
        a.WB[1][i] = T(0);
        a.x[i] = T(0);
        if (cyc) {
            a.bp[cyc_index(a, i, pr)] = bi;
        } else {
            T z = a.minv[i] * bi;
            a.Z[0][i] = z;
            d += (double)bi * (double)z;
        }
    }
    block_pair(d, 0.0, a.p1 + blockIdx.x, a.p2 + blockIdx.x);
}

// Initial state per instance: beta1 = sqrt(b.z), tolerance rtol_i * beta1;
// an instance with rtol_i >= 1 or a zero right-hand side starts done.
__global__ void k_finInit(MState* sts, const double* p1, const double* q1, const int* inst_chunk0, int useq,
                          const double* rtol, int maxit, const int* inst_q0) {
    const int c0 = inst_chunk0[blockIdx.x], c1 = inst_chunk0[blockIdx.x + 1];
    double b2, qb2 = 0.0, unused;
    block_sums(p1, nullptr, c0, c1, b2, unused);
    if (useq) block_sums(q1, nullptr, inst_q0[blockIdx.x], inst_q0[blockIdx.x + 1], qb2, unused);
    b2 += qb2;
    if (threadIdx.x == 0) {
        double beta1 = std::sqrt(b2 > 0.0 ? b2 : 0.0);
        MState s;
        std::memset(&s, 0, sizeof s);
        s.beta1 = beta1;
        s.beta = beta1;
        s.phibar = beta1;
        s.cs = -1.0;
        s.tol = rtol[blockIdx.x] * beta1;
        s.maxit = maxit;
        s.done = (beta1 > 0.0 && maxit > 0 && rtol[blockIdx.x] < 1.0) ? 0 : 1;
        sts[blockIdx.x] = s;
    }
}

// Pending solution update of the last iteration, per chunk.
template <typename T>
__global__ __launch_bounds__(256) void k_flush(UpdArgs<T> a) {
    const VecChunk ch = a.chunks[blockIdx.x];
    const MState* st = a.st + ch.inst;
    if (!st->pending) return;
    const int m = st->itn - 1;
    const T* z = a.Z[m & 1];
    const T* wpp = a.WB[m & 1];
    const T* wp = a.WB[(m + 1) & 1];
    const T uv = (T)st->u_v, uw1 = (T)st->u_w1, uw2 = (T)st->u_w2, ux = (T)st->u_x;
    for (int64_t i = ch.start + threadIdx.x; i < ch.start + ch.len; i += 256) {
        T w = uv * z[i] - uw1 * wpp[i] - uw2 * wp[i];
        a.x[i] += ux * w;
    }
}

// r = b - y over one chunk with the partial of |r|^2 (FP64).
__global__ __launch_bounds__(256) void k_residual(const VecChunk* chunks, const double* b, const double* y, double* r,
                                                  double* p) {
    const VecChunk ch = chunks[blockIdx.x];
    double d = 0.0;
    for (int64_t i = ch.start + threadIdx.x; i < ch.start + ch.len; i += 256) {
        double v = y ? b[i] - y[i] : b[i];  // y null: the residual of x = 0
        r[i] = v;
        d += v * v;
    }
    block_pair(d, 0.0, p + blockIdx.x, nullptr);
}

// Per-instance sums of chunk partials.
__global__ void k_inst_sum(const double* p, const int* inst_chunk0, double* out) {
    double s, unused;
    block_sums(p, nullptr, inst_chunk0[blockIdx.x], inst_chunk0[blockIdx.x + 1], s, unused);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

template <typename S, typename D>
__global__ __launch_bounds__(256) void k_cast(const S* a, D* b, int64_t n) {
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) b[i] = (D)a[i];
}

template <typename S>
__global__ __launch_bounds__(256) void k_add(double* x, const S* d, int64_t n) {
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) x[i] += (double)d[i];
}

// ---------------------------------------------------------------------------
// Host side of the device path.
// ---------------------------------------------------------------------------

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

template <typename T>
static void upload_ops(DeviceState& d, const Layout& L, DevOps<T>& o, const std::vector<T>& O, const std::vector<T>& Lv,
                       bool wide) {
    o.O = dupload(d, O);
    o.L = dupload(d, Lv);
    o.groups = dupload(d, wide ? L.groups32 : L.groups);
}

template <typename T>
static void alloc_krylov(DeviceState& d, DevKrylov<T>& k, int64_t n, int64_t nh) {
    for (int i = 0; i < 2; ++i) {
        k.R[i] = (T*)dalloc(d, n * sizeof(T));
        k.Z[i] = (T*)dalloc(d, n * sizeof(T));
        k.WB[i] = (T*)dalloc(d, n * sizeof(T));
        HIPCHK(hipMemset(k.R[i], 0, n * sizeof(T)));
        HIPCHK(hipMemset(k.Z[i], 0, n * sizeof(T)));
        HIPCHK(hipMemset(k.WB[i], 0, n * sizeof(T)));
    }
    k.x = (T*)dalloc(d, n * sizeof(T));
    k.Az = (T*)dalloc(d, n * sizeof(T));
    k.H = (T*)dalloc(d, nh * sizeof(T));
    k.b = (T*)dalloc(d, n * sizeof(T));
    // the zero slot (second side of boundary edges) is never written
    HIPCHK(hipMemset(k.H, 0, nh * sizeof(T)));
    HIPCHK(hipMemset(k.x, 0, n * sizeof(T)));
    HIPCHK(hipMemset(k.Az, 0, n * sizeof(T)));
    HIPCHK(hipMemset(k.b, 0, n * sizeof(T)));
}

template <typename T>
static void destroy_graph(DevKrylov<T>& k) {
    if (k.graph) (void)hipGraphExecDestroy(k.graph);
    if (k.graph_tail) (void)hipGraphExecDestroy(k.graph_tail);
    k.graph = nullptr;
    k.graph_tail = nullptr;
}

// Static scratch of the allocation-free apply (vemd_apply_device_static, below).
constexpr size_t STATIC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[MAX_DEVICES] = {};

// Records the device address of the static array on `device`, which must be
// current (each GPU holds its own copy at its own address).  In host code
// `scratch` names a host-side stand-in: only this address may reach kernels.
static void static_scratch_init(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw Error(VEMD_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
        throw Error(VEMD_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
}

static void* static_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < MAX_DEVICES ? g_scratch_ptr[device] : nullptr;
}

bool dev_mg_rebuild_possible(const Solver& s) {
    if (!s.dev || !s.amg || !s.splan || !s.dev->amg64 || !s.dev->amg32 || s.setup->sh.k > 2) return false;
    if (s.layout.lam && s.layout.ninst > 1) return false;  // per-instance rank-one terms: host rebuild
    if (s.amg->nvec0 > 1) return false;  // tentative prolongator rows with several entries: host rebuild
    for (const char* v : {"VEMD_MG_DEVICE_REBUILD", "VEMD_MG_REFILL"}) {
        const char* e = std::getenv(v);
        if (e && e[0] == '0') return false;
    }
    const char* ra = std::getenv("VEMD_MG_REAGGREGATE");
    return !(ra && ra[0] == '1');
}

// FP32 copy of an FP64 device array.
__global__ void k_narrow(int64_t n, const double* __restrict__ src, float* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

static void upload_theta(Solver& s, bool device_mg = false, bool rhs = true) {
    DeviceState& d = *s.dev;
    // the scales, diagonals and hierarchy below are read by queued applies,
    // the allocation-free ones included (those run on any stream, unordered)
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(d.o64.scale, s.scale.data(), s.scale.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.o64.minv, s.minv.data(), s.minv.size() * sizeof(double), hipMemcpyHostToDevice));
    const int64_t ns = (int64_t)s.scale.size(), nm = (int64_t)s.minv.size();
    hipLaunchKernelGGL(k_narrow, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, d.stream, ns,
                       (const double*)d.o64.scale, d.o32.scale);
    hipLaunchKernelGGL(k_narrow, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, d.stream, nm,
                       (const double*)d.o64.minv, d.o32.minv);
    HIPCHK(hipGetLastError());
    if (rhs) {
        std::vector<double> bi(s.layout.nvec, 0.0);
        s.to_internal(s.rhs_ref.data(), bi.data(), true);
        HIPCHK(hipMemcpy(d.bprob, bi.data(), bi.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPCHK(hipStreamSynchronize(d.stream));
    // the hierarchy depends on theta through the flux diagonal; with its
    // aggregates kept (the host rebuilt the values only) the uploaded
    // hierarchies are refilled in place, else uploaded afresh
    destroy_graph(d.k64);
    destroy_graph(d.k32);
    multi_invalidate(d);
    if (device_mg) {
        // the hierarchy's values rebuilt on the device from the flux diagonal
        // just uploaded; VEMD_MG_VERIFY=1 also rebuilds them on the host and
        // requires every value to be bitwise equal
        if (!d.mgb) d.mgb = mg_build_create(s, d.cell_of_gpos);
        if (mg_build_run(*d.mgb, s, d.o64.minv, *d.amg64, *d.amg32, d.stream)) {
            const char* vf = std::getenv("VEMD_MG_VERIFY");
            if (vf && vf[0] == '1') {
                s.build_amg();
                double mr = 0.0;
                const int64_t nd = mg_build_compare(*d.mgb, s, &mr);
                if (std::getenv("VEMD_VERBOSE"))
                    std::fprintf(stderr, "[vemd] device multigrid rebuild against the host build: %lld values differ\n",
                                 (long long)nd);
                if (nd != 0)
                    throw Error(VEMD_ERR_INTERNAL, "device multigrid rebuild differs from the host build in " +
                                                       std::to_string(nd) + " values (largest relative difference " +
                                                       std::to_string(mr) + ")");
            }
            return;
        }
        s.build_amg();  // values the device path cannot store as uploaded: host rebuild
    }
    // a host hierarchy whose values were last rebuilt on the device
    if (s.amg && s.amg_host_stale) s.build_amg();
    const char* rf = std::getenv("VEMD_MG_REFILL");  // 0: always upload afresh
    const bool refilled = !(rf && rf[0] == '0') && s.amg && s.amg_values_only && d.amg64 && d.amg32 &&
                          amg_refresh(*d.amg64, *s.amg) && amg_refresh(*d.amg32, *s.amg);
    if (!refilled) {
        delete d.amg64;
        delete d.amg32;
        d.amg64 = nullptr;
        d.amg32 = nullptr;
        if (s.amg) {
            d.amg64 = amg_upload<double>(*s.amg);
            d.amg32 = amg_upload<float>(*s.amg);
        }
    }
}

template <typename T>
static void autotune_apply(Solver& s);

void dev_upload(Solver& s, int device) {
    if (device_count() <= device || device < 0)
        throw Error(VEMD_ERR_NO_DEVICE, "no HIP device " + std::to_string(device) + " (the solve path has no CPU fallback)");
    dev_free(s);
    HIPCHK(hipSetDevice(device));
    s.dev = new DeviceState();
    DeviceState& d = *s.dev;
    d.device = device;
    s.device = device;
    const Layout& L = s.layout;
    HIPCHK(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&d.ev_order, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&d.ev_caller, hipEventDisableTiming));
    HIPCHK(hipEventRecord(d.ev_order, d.stream));
    static_scratch_init(device);
    {
        std::vector<double> O64, L64;
        std::vector<float> O32, L32;
        std::vector<uint32_t> wmask;
        pack_store(*s.mesh, *s.setup, L, O64, L64, O32, L32, wmask);
        upload_ops(d, L, d.o64, O64, L64, false);
        upload_ops(d, L, d.o32, O32, L32, true);
        d.wmask = dupload(d, wmask);
    }
    d.idx = dupload(d, L.idx);
    d.hside = dupload(d, L.hside);
    d.tile_group = dupload(d, L.tile_group);
    d.cell_of_gpos = dupload(d, L.cell_of_gpos);
    d.ntiles = (int)L.tile_group.size();
    d.o64.scale = (double*)dalloc(d, L.cpad * sizeof(double));
    d.o32.scale = (float*)dalloc(d, L.cpad * sizeof(float));
    d.o64.minv = (double*)dalloc(d, L.nvec * sizeof(double));
    d.o32.minv = (float*)dalloc(d, L.nvec * sizeof(float));
    d.pdot = (double*)dalloc(d, d.ntiles * sizeof(double));
    d.plam = (double*)dalloc(d, d.ntiles * sizeof(double));
    d.nblk = (int)std::min<int64_t>(2048, (L.nvec + 255) / 256);
    d.ninst = L.ninst;
    d.nchunks = (int)L.chunks.size();
    d.npchunks = (int)L.pchunks.size();
    d.chunks = dupload(d, L.chunks);
    d.pchunks = dupload(d, L.pchunks);
    d.inst_tile0 = dupload(d, L.inst_tile0);
    d.inst_chunk0 = dupload(d, L.inst_chunk0);
    d.p1 = (double*)dalloc(d, d.nchunks * sizeof(double));
    d.p2 = (double*)dalloc(d, d.nchunks * sizeof(double));
    // pressure-block partials: per chunk (k_pdot) or per tile (the cycle's last kernel)
    const int nq = std::max(d.nchunks, d.ntiles);
    d.q1 = (double*)dalloc(d, nq * sizeof(double));
    d.q2 = (double*)dalloc(d, nq * sizeof(double));
    for (double* p : {d.p1, d.p2}) HIPCHK(hipMemset(p, 0, d.nchunks * sizeof(double)));
    for (double* p : {d.q1, d.q2}) HIPCHK(hipMemset(p, 0, nq * sizeof(double)));
    {
        std::vector<int> ti(d.ntiles);
        for (int t = 0; t < d.ntiles; ++t) ti[t] = L.groups[L.tile_group[t]].inst;
        d.tile_inst = dupload(d, ti);
    }
    d.st = (MState*)dalloc(d, d.ninst * sizeof(MState));
    HIPCHK(hipMemset(d.st, 0, d.ninst * sizeof(MState)));
    HIPCHK(hipHostMalloc((void**)&d.st_host, d.ninst * sizeof(MState), hipHostMallocDefault));
    d.all_done = (int*)dalloc(d, sizeof(int));
    HIPCHK(hipMemset(d.all_done, 0, sizeof(int)));
    d.rtol = (double*)dalloc(d, d.ninst * sizeof(double));
    HIPCHK(hipHostMalloc((void**)&d.rtol_host, d.ninst * sizeof(double), hipHostMallocDefault));
    d.dsum = (double*)dalloc(d, d.ninst * sizeof(double));
    HIPCHK(hipHostMalloc((void**)&d.hsum, d.ninst * sizeof(double), hipHostMallocDefault));
    const int64_t nh = L.nside * (L.k + 1);
    alloc_krylov(d, d.k64, L.nvec, nh);
    alloc_krylov(d, d.k32, L.nvec, nh);
    d.h64 = (double*)dalloc(d, nh * sizeof(double));
    d.h32 = (float*)dalloc(d, nh * sizeof(float));
    HIPCHK(hipMemset(d.h64, 0, nh * sizeof(double)));
    HIPCHK(hipMemset(d.h32, 0, nh * sizeof(float)));
    d.x64 = (double*)dalloc(d, L.nvec * sizeof(double));
    d.r64 = (double*)dalloc(d, L.nvec * sizeof(double));
    d.y64 = (double*)dalloc(d, L.nvec * sizeof(double));
    d.b64 = (double*)dalloc(d, L.nvec * sizeof(double));
    d.r32 = (float*)dalloc(d, L.nvec * sizeof(float));
    d.bprob = (double*)dalloc(d, L.nvec * sizeof(double));
    for (double* p : {d.x64, d.r64, d.y64, d.b64, d.bprob}) HIPCHK(hipMemset(p, 0, L.nvec * sizeof(double)));
    HIPCHK(hipMemset(d.r32, 0, L.nvec * sizeof(float)));
    HIPCHK(hipEventCreate(&d.e0));
    HIPCHK(hipEventCreate(&d.e1));
    upload_theta(s);
    autotune_apply<double>(s);
    autotune_apply<float>(s);
}

void dev_refresh(Solver& s, bool device_mg) {
    if (!s.dev) return;
    HIPCHK(hipSetDevice(s.dev->device));
    auto t0 = std::chrono::steady_clock::now();
    upload_theta(s, device_mg, s.rhs_changed);
    if (std::getenv("VEMD_VERBOSE"))
        std::fprintf(stderr, "[vemd] theta-dependent device data %.3fs\n",
                     std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
}


void dev_errors(Solver& s, const void* x_dev, int nq, void* stream, double* out6) {
    if (!s.dev) throw Error(VEMD_ERR_STATE, "upload the solver before device error evaluation");
    DeviceState& d = *s.dev;
    HIPCHK(hipSetDevice(d.device));
    ScratchUse use(d, d.stream, true, (hipStream_t)stream);
    errors_device(s, (const double*)x_dev, nq, d.stream, out6);
}

void dev_free(Solver& s) {
    errdev_free(s.errdev);
    s.errdev = nullptr;
    if (!s.dev) return;
    DeviceState& d = *s.dev;
    (void)hipSetDevice(d.device);
    destroy_graph(d.k64);
    destroy_graph(d.k32);
    multi_free(d);
    delete d.amg64;
    delete d.amg32;
    delete d.mgb;
    for (void* p : d.allocs) (void)hipFree(p);
    if (d.st_host) (void)hipHostFree(d.st_host);
    if (d.hsum) (void)hipHostFree(d.hsum);
    if (d.rtol_host) (void)hipHostFree(d.rtol_host);
    if (d.e0) (void)hipEventDestroy(d.e0);
    if (d.e1) (void)hipEventDestroy(d.e1);
    if (d.ev_order) (void)hipEventDestroy(d.ev_order);
    if (d.ev_caller) (void)hipEventDestroy(d.ev_caller);
    if (d.stream) (void)hipStreamDestroy(d.stream);
    delete s.dev;
    s.dev = nullptr;
}

template <typename T>
static const DevOps<T>& ops(DeviceState& d);
template <>
const DevOps<double>& ops<double>(DeviceState& d) {
    return d.o64;
}
template <>
const DevOps<float>& ops<float>(DeviceState& d) {
    return d.o32;
}

template <typename T>
static T* side_scratch(DeviceState& d);
template <>
double* side_scratch<double>(DeviceState& d) {
    return d.h64;
}
template <>
float* side_scratch<float>(DeviceState& d) {
    return d.h32;
}

template <typename T>
static ApplyArgs<T> make_args(Solver& s, const T* x, T* y, T* h, const MState* st, const T* x0, const T* x1,
                              bool dot) {
    DeviceState& d = *s.dev;
    const DevOps<T>& o = ops<T>(d);
    const Layout& L = s.layout;
    ApplyArgs<T> a;
    a.groups = o.groups;
    a.tile_group = d.tile_group;
    a.O = o.O;
    a.L = o.L;
    a.scale = o.scale;
    a.idx = d.idx;
    a.wmask = d.wmask;
    a.x = x;
    a.xs[0] = x0;
    a.xs[1] = x1;
    a.y = y;
    a.h = h;
    a.cbase = L.cbase;
    a.cpad = L.cpad;
    a.lam_idx = L.lam ? L.lam_idx : -1;
    a.nvec = L.nvec;
    a.pdot = dot ? d.pdot : nullptr;
    a.plam = d.plam;
    a.st = st;
    a.ntiles = d.ntiles;
    a.per_xcd = (int)(((d.ntiles + 3) / 4 + 7) / 8);
    a.tile_base = 0;
    a.force_var = -1;
    return a;
}

// One apply on the solver stream: a single launch over all tiles.
template <typename T>
static void launch_apply(Solver& s, const ApplyArgs<T>& a) {
    dim3 grid(8 * a.per_xcd), block(256);
    if (s.layout.k == 0)
        hipLaunchKernelGGL((k_apply<T, 0>), grid, block, 0, s.dev->stream, a);
    else if (s.layout.k == 1)
        hipLaunchKernelGGL((k_apply<T, 1>), grid, block, 0, s.dev->stream, a);
    else if (s.layout.k == 2)
        hipLaunchKernelGGL((k_apply<T, 2>), grid, block, 0, s.dev->stream, a);
    else
        launch_apply_k3<T>(a, grid, s.dev->stream);
}

// Autotuner of the apply: for every valence, the variants of the cell body
// (operator-stream chunk shapes) are timed on the valence's largest group,
// launched alone, and the fastest is recorded in every group of that valence.
// VEMD_AUTOTUNE=0 keeps variant 0 everywhere.
template <typename T>
static void autotune_apply(Solver& s) {
    DeviceState& d = *s.dev;
    Layout& L = s.layout;
    std::vector<GroupDesc>& G = sizeof(T) == 8 ? L.groups : L.groups32;
    int* tuned = sizeof(T) == 8 ? d.tuned64 : d.tuned32;
    float* tuned_ms = sizeof(T) == 8 ? d.tuned_ms64 : d.tuned_ms32;
    for (int i = 0; i < 16; ++i) {
        tuned[i] = -1;
        tuned_ms[i] = 0.0f;
    }
    const char* env = std::getenv("VEMD_AUTOTUNE");
    const bool on = !(env && env[0] == '0') && L.k <= 1;  // k >= 2 builds one variant
    std::vector<int> rep(16, -1);  // largest group of each valence
    for (size_t gi = 0; gi < G.size(); ++gi) {
        const int ne = G[gi].ne;
        if (ne < 0 || ne >= 16) continue;
        if (rep[ne] < 0 || G[gi].npad > G[rep[ne]].npad) rep[ne] = (int)gi;
    }
    T* x = sizeof(T) == 8 ? (T*)d.x64 : (T*)d.r32;
    T* y = sizeof(T) == 8 ? (T*)d.y64 : (T*)d.k32.Az;
    T* h = side_scratch<T>(d);
    HIPCHK(hipMemsetAsync(x, 0, L.nvec * sizeof(T), d.stream));
    for (int ne = 0; ne < 16; ++ne) {
        if (rep[ne] < 0) continue;
        const GroupDesc& g = G[rep[ne]];
        const int nt = g.npad / TILE;
        int best = 0;
        float best_ms = 0.0f;
        if (on && nt >= 8) {
            for (int v = 0; v < NVARIANT; ++v) {
                ApplyArgs<T> a = make_args<T>(s, x, y, h, nullptr, nullptr, nullptr, false);
                a.force_var = v;
                a.tile_base = g.tile0;
                a.ntiles = nt;
                a.per_xcd = (nt + 31) / 32;
                launch_apply(s, a);  // warm
                HIPCHK(hipEventRecord(d.e0, d.stream));
                constexpr int R = 5;
                for (int r = 0; r < R; ++r) launch_apply(s, a);
                HIPCHK(hipEventRecord(d.e1, d.stream));
                HIPCHK(hipEventSynchronize(d.e1));
                float ms = 0.0f;
                HIPCHK(hipEventElapsedTime(&ms, d.e0, d.e1));
                ms /= R;
                if (v == 0 || ms < best_ms) {
                    best = v;
                    best_ms = ms;
                }
            }
        }
        tuned[ne] = best;
        tuned_ms[ne] = best_ms;
        for (GroupDesc& gg : G)
            if (gg.ne == ne) gg.var = best;
    }
    HIPCHK(hipGetLastError());
    GroupDesc* dg = sizeof(T) == 8 ? d.o64.groups : d.o32.groups;
    HIPCHK(hipMemcpy(dg, G.data(), G.size() * sizeof(GroupDesc), hipMemcpyHostToDevice));
    if (std::getenv("VEMD_VERBOSE")) {
        std::fprintf(stderr, "[vemd] apply autotune (%s): valence:variant(ms)", sizeof(T) == 8 ? "fp64" : "fp32");
        for (int ne = 0; ne < 16; ++ne)
            if (tuned[ne] >= 0) std::fprintf(stderr, " %d:%d(%.4f)", ne, tuned[ne], tuned_ms[ne]);
        std::fprintf(stderr, "\n");
    }
}

int dev_tuning(const Solver& s, int* out) {
    if (!s.dev) throw Error(VEMD_ERR_STATE, "solver not uploaded");
    for (int ne = 3; ne <= 10; ++ne) {
        out[2 * (ne - 3)] = s.dev->tuned64[ne];
        out[2 * (ne - 3) + 1] = s.dev->tuned32[ne];
    }
    return 16;
}

// y = K x in the internal layout: cell launch, the edge sums of the two cell
// sides, and the multiplier row.  With h_out the edge sums are left to the
// caller: h_out receives the cell sides and y's edge-dof block is untouched.
template <typename T>
static void apply_plain(Solver& s, const T* x, T* y, T* h_out = nullptr) {
    DeviceState& d = *s.dev;
    const Layout& L = s.layout;
    T* h = h_out ? h_out : side_scratch<T>(d);
    ApplyArgs<T> a = make_args<T>(s, x, y, h, nullptr, nullptr, nullptr, false);
    launch_apply(s, a);
    if (!h_out)
        hipLaunchKernelGGL((k_combine<T>), dim3(d.nblk), dim3(256), 0, d.stream, y, (const T*)h, (const int*)d.hside,
                           L.k + 1, L.cbase, L.trash_idx);
    if (s.layout.lam)
        hipLaunchKernelGGL((k_lam_row<T>), dim3(d.ninst), dim3(256), 0, d.stream, (const double*)d.plam,
                           (const int*)d.inst_tile0, y, s.layout.lam_idx);
    HIPCHK(hipGetLastError());
}

void dev_apply_int(Solver& s, int precision, const void* x, void* y, void* h, void* stream) {
    if (!s.dev) throw Error(VEMD_ERR_STATE, "solver not uploaded");
    HIPCHK(hipSetDevice(s.dev->device));
    // the caller's stream; NULL is the null (default) stream, as documented
    hipStream_t saved = s.dev->stream;
    s.dev->stream = (hipStream_t)stream;
    try {
        ScratchUse use(*s.dev, (hipStream_t)stream);
        if (precision == VEMD_FP64)
            apply_plain<double>(s, (const double*)x, (double*)y, (double*)h);
        else
            apply_plain<float>(s, (const float*)x, (float*)y, (float*)h);
    } catch (...) {
        s.dev->stream = saved;
        throw;
    }
    s.dev->stream = saved;
}

// ---------------------------------------------------------------------------
// Allocation-free apply.  The side slots and the multiplier-row partials of
// the apply live in one static device array per GPU instead of the solver's
// scratch, for callers that cannot allocate device memory (code under HIP
// graph capture, serving loops that forbid hipMalloc).  Each GPU the library
// runs on holds STATIC_SCRATCH_BYTES for it until the process ends; every call
// on a GPU shares that GPU's array and leaves its contents behind, so calls
// that overlap on one GPU race (the caller orders them).
// ---------------------------------------------------------------------------

template <typename T>
__global__ void k_fill(T* p, int n, T v) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

// The apply of apply_plain with h (side slots) and plam (multiplier-row
// partials) in the static array; kernel launches only.
template <typename T>
static void apply_static(Solver& s, const T* x, T* y, T* h, double* plam) {
    DeviceState& d = *s.dev;
    const Layout& L = s.layout;
    const int kk = L.k + 1;
    // the zero slot (second side of every boundary edge) must read zero, and
    // the array still holds what the last call on this GPU left there
    hipLaunchKernelGGL((k_fill<T>), dim3(1), dim3(64), 0, d.stream, h + (L.nside - 1) * kk, kk, T(0));
    ApplyArgs<T> a = make_args<T>(s, x, y, h, nullptr, nullptr, nullptr, false);
    a.plam = plam;
    launch_apply(s, a);
    hipLaunchKernelGGL((k_combine<T>), dim3(d.nblk), dim3(256), 0, d.stream, y, (const T*)h, (const int*)d.hside, kk,
                       L.cbase, L.trash_idx);
    if (L.lam)
        hipLaunchKernelGGL((k_lam_row<T>), dim3(d.ninst), dim3(256), 0, d.stream, (const double*)plam,
                           (const int*)d.inst_tile0, y,
                           L.lam_idx);
    HIPCHK(hipGetLastError());
}

size_t dev_static_scratch_need(const Solver& s, int precision) {
    const Layout& L = s.layout;
    const size_t w = precision == VEMD_FP64 ? sizeof(double) : sizeof(float);
    const size_t hbytes = ((size_t)L.nside * (L.k + 1) * w + 255) / 256 * 256;
    return hbytes + L.tile_group.size() * sizeof(double);
}

void dev_apply_static(Solver& s, int precision, const void* x, void* y, void* stream) {
    if (!s.dev) throw Error(VEMD_ERR_STATE, "solver not uploaded");
    DeviceState& d = *s.dev;
    HIPCHK(hipSetDevice(d.device));
    void* base = static_scratch(d.device);
    if (!base) throw Error(VEMD_ERR_STATE, "static scratch not set up on device " + std::to_string(d.device));
    const size_t need = dev_static_scratch_need(s, precision);
    size_t limit = STATIC_SCRATCH_BYTES;
    if (const char* e = std::getenv("VEMD_STATIC_SCRATCH_LIMIT"))  // lowers the limit (tests); cannot raise it
        limit = std::min(limit, (size_t)std::max(0LL, std::atoll(e)));
    if (need > limit)
        throw Error(VEMD_ERR_ARG, "the allocation-free apply needs " + std::to_string(need) +
                                      " bytes of static scratch for this mesh and precision, more than its limit of " +
                                      std::to_string(limit) + " bytes; use vemd_apply_device");
    const size_t hbytes = need - (size_t)d.ntiles * sizeof(double);
    double* plam = (double*)((char*)base + hbytes);
    hipStream_t saved = d.stream;
    d.stream = (hipStream_t)stream;
    try {
        if (precision == VEMD_FP64)
            apply_static<double>(s, (const double*)x, (double*)y, (double*)base, plam);
        else
            apply_static<float>(s, (const float*)x, (float*)y, (float*)base, plam);
    } catch (...) {
        d.stream = saved;
        throw;
    }
    d.stream = saved;
}

void dev_apply_ref(Solver& s, int precision, const double* x_ref, double* y_ref) {
    if (!s.dev) throw Error(VEMD_ERR_STATE, "solver not uploaded");
    DeviceState& d = *s.dev;
    HIPCHK(hipSetDevice(d.device));
    ScratchUse use(d, d.stream);
    const Layout& L = s.layout;
    std::vector<double> xi(L.nvec, 0.0), yi(L.nvec);
    s.to_internal(x_ref, xi.data(), false);
    if (precision == VEMD_FP64) {
        HIPCHK(hipMemcpyAsync(d.x64, xi.data(), L.nvec * sizeof(double), hipMemcpyHostToDevice, d.stream));
        apply_plain<double>(s, d.x64, d.y64);
        HIPCHK(hipMemcpyAsync(yi.data(), d.y64, L.nvec * sizeof(double), hipMemcpyDeviceToHost, d.stream));
    } else {
        std::vector<float> xf(xi.begin(), xi.end()), yf(L.nvec);
        HIPCHK(hipMemcpyAsync(d.r32, xf.data(), L.nvec * sizeof(float), hipMemcpyHostToDevice, d.stream));
        apply_plain<float>(s, d.r32, d.k32.Az);
        HIPCHK(hipMemcpyAsync(yf.data(), d.k32.Az, L.nvec * sizeof(float), hipMemcpyDeviceToHost, d.stream));
        HIPCHK(hipStreamSynchronize(d.stream));
        for (int64_t i = 0; i < L.nvec; ++i) yi[i] = yf[i];
    }
    HIPCHK(hipStreamSynchronize(d.stream));
    s.to_reference(yi.data(), y_ref, true);
}

template <typename T>
static AmgDev<T>* amg_of(DeviceState& d);
template <>
AmgDev<double>* amg_of<double>(DeviceState& d) {
    return d.amg64;
}
template <>
AmgDev<float>* amg_of<float>(DeviceState& d) {
    return d.amg32;
}

// Kernel arguments of the Krylov sweeps; with the multigrid cycle the
// pressure chunks of z come from the cycle instead of the diagonal.
template <typename T>
static UpdArgs<T> upd_args(Solver& s, DevKrylov<T>& k, bool amg) {
    DeviceState& d = *s.dev;
    const Layout& L = s.layout;
    UpdArgs<T> u{{k.R[0], k.R[1]}, {k.Z[0], k.Z[1]}, {k.WB[0], k.WB[1]}, k.x,  k.Az,      k.H, d.hside, L.k + 1,
                 ops<T>(d).minv,   L.cbase,          d.chunks,          d.p1,    d.p2, d.st, amg ? 1 : 0, 0, nullptr,
                 L.cpad,           L.q,               nullptr,           nullptr, T(0)};
    if (amg) {
        const AmgDev<T>& M = *amg_of<T>(d);
        u.pb = L.cbase + (int64_t)L.ni * L.cpad;
        u.bp = M.lv[0].b;
        if (amg_pre_fusable(M)) {  // the cycle's first update formed here
            u.bd0 = M.lv[0].d0;
            u.bsc = M.lv[0].sc ? M.lv[0].sc : M.lv[0].dinv;
            u.c0 = (T)(1.0 / M.lv[0].centre);
        }
    }
    return u;
}

// The Krylov output of the multigrid cycle's last kernel: the pressure block
// of z and the per-tile partials of the pressure-block dots.
template <typename T>
static KOut<T> kout(Solver& s, DevKrylov<T>& k, bool init) {
    DeviceState& d = *s.dev;
    const Layout& L = s.layout;
    KOut<T> o;
    o.Z[0] = k.Z[0];
    o.Z[1] = k.Z[1];
    o.R[0] = k.R[0];
    o.R[1] = k.R[1];
    o.st = d.st;
    o.tile_inst = d.ninst > 1 ? d.tile_inst : nullptr;
    o.pb = L.cbase + (int64_t)L.ni * L.cpad;
    o.cpad = L.cpad;
    o.q1 = d.q1;
    o.q2 = d.q2;
    o.init = init ? 1 : 0;
    return o;
}

// The FP32 sweep by entry pairs (k = 1) unless VEMD_PAIR_SWEEP=0.
static bool pair_sweep() {
    const char* e = std::getenv("VEMD_PAIR_SWEEP");
    return !(e && e[0] == '0');
}

// One MINRES iteration of every unfinished instance: apply, alfa,
// Lanczos/precondition/update, [pressure cycle], Givens.
template <typename T>
static void enqueue_iteration(Solver& s, DevKrylov<T>& k, bool amg) {
    DeviceState& d = *s.dev;
    const Layout& L = s.layout;
    ApplyArgs<T> a = make_args<T>(s, nullptr, k.Az, k.H, d.st, k.Z[0], k.Z[1], true);
    launch_apply(s, a);
    hipLaunchKernelGGL((k_finA<T>), dim3(d.ninst), dim3(FIN), 0, d.stream, d.st, d.pdot, d.plam, d.inst_tile0, k.Az,
                       k.Z[0], k.Z[1], L.lam ? L.lam_idx : (int64_t)-1);
    UpdArgs<T> u = upd_args(s, k, amg);
    if constexpr (std::is_same<T, float>::value) {
        if (L.k == 1 && pair_sweep())
            hipLaunchKernelGGL(k_update_pairs, dim3(d.nchunks), dim3(256), 0, d.stream, u);
        else
            hipLaunchKernelGGL((k_update<T>), dim3(d.nchunks), dim3(256), 0, d.stream, u);
    } else {
        hipLaunchKernelGGL((k_update<T>), dim3(d.nchunks), dim3(256), 0, d.stream, u);
    }
    const int* q0 = d.inst_chunk0;
    if (amg) {
        const AmgDev<T>& M = *amg_of<T>(d);
        if (amg_kout_fusable(M)) {  // z and its dots from the cycle's last kernel
            const KOut<T> ko = kout(s, k, false);
            amg_vcycle(M, d.stream, d.all_done, u.bd0 != nullptr, &ko);
            q0 = d.inst_tile0;
        } else {
            amg_vcycle(M, d.stream, d.all_done, u.bd0 != nullptr);
            hipLaunchKernelGGL((k_pdot<T>), dim3(d.npchunks), dim3(256), 0, d.stream, u, (const int*)d.pchunks,
                               (const T*)M.lv[0].x, d.q1, d.q2, 0);
        }
    }
    hipLaunchKernelGGL(k_finB, dim3(d.ninst), dim3(FIN), 0, d.stream, d.st, d.p1, d.p2, d.q1, d.q2, d.inst_chunk0,
                       amg ? 1 : 0, d.ninst == 1 ? d.all_done : nullptr, q0);
    if (d.ninst > 1) hipLaunchKernelGGL(k_alldone, dim3(1), dim3(256), 0, d.stream, d.st, d.ninst, d.all_done);
}

// Preconditioned MINRES on K x = b for every instance (internal order,
// precision T): b in k.b, solution in k.x.  Instance i stops when its
// M^{-1}-norm residual estimate falls to rtol[i] times its initial value
// (rtol[i] >= 1: left untouched, x_i = 0).  Fills its[i]; returns the largest.
template <typename T>
static int minres(Solver& s, DevKrylov<T>& k, const std::vector<double>& rtol, int maxit, int check_every,
                  bool use_graph, bool amg, std::vector<int>& its) {
    DeviceState& d = *s.dev;
    UpdArgs<T> u = upd_args(s, k, amg);
    check_every = std::max(1, check_every);
    // Near convergence a shorter graph is launched, so that fewer captured
    // iterations run (as early-returning kernels) after the stop test fires.
    const int tail = std::max(1, std::min(check_every, 4));
    if (use_graph && (k.graph_iters != check_every || k.graph_amg != amg || !k.graph)) {
        HIPCHK(hipStreamSynchronize(d.stream));
        destroy_graph(k);
        for (int pass = 0; pass < 2; ++pass) {
            hipGraph_t g;
            HIPCHK(hipStreamBeginCapture(d.stream, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < (pass == 0 ? check_every : tail); ++i) enqueue_iteration(s, k, amg);
            HIPCHK(hipStreamEndCapture(d.stream, &g));
            HIPCHK(hipGraphInstantiate(pass == 0 ? &k.graph : &k.graph_tail, g, nullptr, nullptr, 0));
            HIPCHK(hipGraphDestroy(g));
        }
        k.graph_iters = check_every;
        k.graph_amg = amg;
    }
    HIPCHK(hipStreamSynchronize(d.stream));  // the pinned tolerance buffer is reused
    for (int i = 0; i < d.ninst; ++i) d.rtol_host[i] = rtol[i];
    HIPCHK(hipMemcpyAsync(d.rtol, d.rtol_host, d.ninst * sizeof(double), hipMemcpyHostToDevice, d.stream));
    hipLaunchKernelGGL((k_init<T>), dim3(d.nchunks), dim3(256), 0, d.stream, u, (const T*)k.b);
    const int* q0 = d.inst_chunk0;
    if (amg) {
        const AmgDev<T>& M = *amg_of<T>(d);
        if (amg_kout_fusable(M)) {
            const KOut<T> ko = kout(s, k, true);
            amg_vcycle(M, d.stream, nullptr, false, &ko);
            q0 = d.inst_tile0;
        } else {
            amg_vcycle(M, d.stream, nullptr);
            hipLaunchKernelGGL((k_pdot<T>), dim3(d.npchunks), dim3(256), 0, d.stream, u, (const int*)d.pchunks,
                               (const T*)M.lv[0].x, d.q1, d.q2, 1);
        }
    }
    hipLaunchKernelGGL(k_finInit, dim3(d.ninst), dim3(FIN), 0, d.stream, d.st, d.p1, d.q1, d.inst_chunk0, amg ? 1 : 0,
                       (const double*)d.rtol, maxit, q0);
    hipLaunchKernelGGL(k_alldone, dim3(1), dim3(256), 0, d.stream, d.st, d.ninst, d.all_done);
    HIPCHK(hipGetLastError());
    static const bool verbose = std::getenv("VEMD_VERBOSE") != nullptr;
    auto last = std::chrono::steady_clock::now();
    // per instance: iteration count and log of the residual estimate at the
    // previous check (the convergence rate predicts the iterations left)
    std::vector<int> prev_it(d.ninst, -1);
    std::vector<double> prev_lg(d.ninst, 0.0);
    while (true) {
        HIPCHK(hipMemcpyAsync(d.st_host, d.st, d.ninst * sizeof(MState), hipMemcpyDeviceToHost, d.stream));
        HIPCHK(hipStreamSynchronize(d.stream));
        bool all = true;
        for (int i = 0; i < d.ninst && all; ++i) all = d.st_host[i].done != 0;
        if (all) break;
        double left = 0.0;  // predicted iterations until the slowest unfinished instance stops
        bool known = true;
        for (int i = 0; i < d.ninst; ++i) {
            const MState& m = d.st_host[i];
            if (m.done) continue;
            const double lg = std::log(std::max(m.phibar, 1e-300));
            if (prev_it[i] >= 0 && m.itn > prev_it[i] && lg < prev_lg[i]) {
                const double rate = (lg - prev_lg[i]) / (m.itn - prev_it[i]);
                left = std::max(left, (std::log(std::max(m.tol, 1e-300)) - lg) / rate);
            } else {
                known = false;
            }
            prev_it[i] = m.itn;
            prev_lg[i] = lg;
        }
        const bool near_end = known && left < check_every;
        if (verbose && std::chrono::steady_clock::now() - last > std::chrono::seconds(10)) {
            last = std::chrono::steady_clock::now();
            const MState& m = d.st_host[0];
            std::fprintf(stderr, "[vemd] minres %s: %d iterations, residual estimate %.3e of %.3e (instance 0 of %d)\n",
                         sizeof(T) == 8 ? "fp64" : "fp32", m.itn, m.phibar / m.beta1, rtol[0], d.ninst);
        }
        if (use_graph)
            HIPCHK(hipGraphLaunch(near_end ? k.graph_tail : k.graph, d.stream));
        else
            for (int i = 0; i < (near_end ? tail : check_every); ++i) enqueue_iteration(s, k, amg);
    }
    hipLaunchKernelGGL((k_flush<T>), dim3(d.nchunks), dim3(256), 0, d.stream, u);
    HIPCHK(hipGetLastError());
    its.assign(d.ninst, 0);
    int mx = 0;
    for (int i = 0; i < d.ninst; ++i) {
        its[i] = d.st_host[i].itn;
        mx = std::max(mx, its[i]);
    }
    return mx;
}

// r64 = b64 - K x64 (x64 = 0: r64 = b64, no apply); returns the 2-norm of
// every instance's block of r64.
static std::vector<double> residual64(Solver& s, bool x_zero = false) {
    DeviceState& d = *s.dev;
    if (!x_zero) apply_plain<double>(s, d.x64, d.y64);
    hipLaunchKernelGGL(k_residual, dim3(d.nchunks), dim3(256), 0, d.stream, (const VecChunk*)d.chunks,
                       (const double*)d.b64, x_zero ? (const double*)nullptr : (const double*)d.y64, d.r64, d.p1);
    hipLaunchKernelGGL(k_inst_sum, dim3(d.ninst), dim3(FIN), 0, d.stream, (const double*)d.p1,
                       (const int*)d.inst_chunk0, d.dsum);
    HIPCHK(hipMemcpyAsync(d.hsum, d.dsum, d.ninst * sizeof(double), hipMemcpyDeviceToHost, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    std::vector<double> r(d.ninst);
    for (int i = 0; i < d.ninst; ++i) r[i] = std::sqrt(d.hsum[i]);
    return r;
}

// Solve with b64 already on the device; solution in x64.  Iterative
// refinement per instance: FP64 residuals, inner MINRES (FP64, or FP32 for
// VEMD_MIXED) with per-instance tolerances, until every instance's true
// relative residual is below rtol.
static void solve_device(Solver& s, const vemd_solve_opts& o, vemd_solve_info* info) {
    DeviceState& d = *s.dev;
    const int64_t n = s.layout.nvec;
    const int ni = d.ninst;
    HIPCHK(hipMemsetAsync(d.x64, 0, n * sizeof(double), d.stream));
    std::vector<double> bn = residual64(s, true), rel(ni);  // r = b
    for (int i = 0; i < ni; ++i) rel[i] = bn[i] > 0 ? 1.0 : 0.0;
    std::vector<int> it_tot(ni, 0), its;
    int iters = 0, refin = 0;
    int maxit = o.max_iters > 0 ? o.max_iters : 1000000;
    bool graph = o.use_graph != 0;
    int chk = o.check_every > 0 ? o.check_every : 32;
    if (o.precond == VEMD_PC_MULTIGRID && !d.amg64)
        throw Error(VEMD_ERR_ARG, "no multigrid hierarchy on this solver");
    const bool amg = o.precond != VEMD_PC_JACOBI && d.amg64 != nullptr;
    if (s.layout.lam) {
        // the multiplier rows' preconditioner depends on the pressure block's
        const int nl = s.layout.ninst;
        std::vector<double> lm(nl);
        std::vector<float> lf(nl);
        for (int i = 0; i < nl; ++i) {
            lm[i] = amg ? s.amg_lam_minv[i] : s.minv[s.layout.lam_idx + i];
            lf[i] = (float)lm[i];
        }
        HIPCHK(hipMemcpyAsync(d.o64.minv + s.layout.lam_idx, lm.data(), nl * sizeof(double), hipMemcpyHostToDevice,
                              d.stream));
        HIPCHK(hipMemcpyAsync(d.o32.minv + s.layout.lam_idx, lf.data(), nl * sizeof(float), hipMemcpyHostToDevice,
                              d.stream));
        HIPCHK(hipStreamSynchronize(d.stream));  // the sources are on this stack frame
    }
    auto worst = [&]() { return *std::max_element(rel.begin(), rel.end()); };
    std::vector<double> target(ni);
    // Inner targets are relative to each pass's right-hand side and are met in
    // MINRES's preconditioned norm. On the first pass (smooth data) that norm
    // lags the true 2-norm residual by two to three orders of magnitude; on
    // later passes it tracks it while the target stays above ~1e-5, below
    // which an FP32 pass stagnates in the true residual. So every FP32 pass
    // stops at inner_rtol or at the outer tolerance if that is nearer.
    while (worst() > o.rtol && iters < maxit && refin < 100) {
        for (int i = 0; i < ni; ++i) {
            if (!(rel[i] > o.rtol))
                target[i] = 1.0;  // converged: no inner iterations
            else if (o.precision == VEMD_FP64)
                target[i] = std::min(0.5, 0.5 * o.rtol / rel[i]);
            else
                target[i] = std::max(o.inner_rtol > 0 ? o.inner_rtol : 1e-4, 0.5 * o.rtol / rel[i]);
        }
        if (o.precision == VEMD_FP64) {
            HIPCHK(hipMemcpyAsync(d.k64.b, d.r64, n * sizeof(double), hipMemcpyDeviceToDevice, d.stream));
            iters += minres<double>(s, d.k64, target, maxit - iters, chk, graph, amg, its);
            hipLaunchKernelGGL((k_add<double>), dim3(d.nblk), dim3(256), 0, d.stream, d.x64, d.k64.x, n);
        } else {
            hipLaunchKernelGGL((k_cast<double, float>), dim3(d.nblk), dim3(256), 0, d.stream, d.r64, d.k32.b, n);
            iters += minres<float>(s, d.k32, target, maxit - iters, chk, graph, amg, its);
            hipLaunchKernelGGL((k_add<float>), dim3(d.nblk), dim3(256), 0, d.stream, d.x64, d.k32.x, n);
        }
        for (int i = 0; i < ni; ++i) it_tot[i] += its[i];
        ++refin;
        std::vector<double> rn = residual64(s);
        for (int i = 0; i < ni; ++i) rel[i] = bn[i] > 0 ? rn[i] / bn[i] : 0.0;
        static const bool verbose = std::getenv("VEMD_VERBOSE") != nullptr;
        if (verbose)
            std::fprintf(stderr, "[vemd] refinement %d: inner target %.2e (instance 0), %d iterations, true relative "
                                 "residual %.3e (worst %.3e)\n",
                         refin, target[0], its[0], rel[0], worst());
    }
    s.inst_iterations = it_tot;
    s.inst_rel = rel;
    if (info) {
        info->iterations = iters;
        info->refinements = refin;
        info->rel_residual = worst();
    }
    if (worst() > o.rtol)
        throw Error(VEMD_ERR_NOT_CONVERGED, "not converged: relative residual " + std::to_string(worst()) + " after " +
                                                std::to_string(iters) + " iterations");
}

void dev_solve(Solver& s, const vemd_solve_opts& o, const double* rhs_ref, double* sol_ref, vemd_solve_info* info) {
    if (!s.dev) throw Error(VEMD_ERR_STATE, "solver not uploaded");
    DeviceState& d = *s.dev;
    HIPCHK(hipSetDevice(d.device));
    ScratchUse use(d, d.stream);
    const Layout& L = s.layout;
    std::vector<double> bi(L.nvec, 0.0);
    s.to_internal(rhs_ref ? rhs_ref : s.rhs_ref.data(), bi.data(), true);
    HIPCHK(hipMemcpyAsync(d.b64, bi.data(), L.nvec * sizeof(double), hipMemcpyHostToDevice, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    auto t0 = std::chrono::steady_clock::now();
    vemd_solve_info tmp{};
    std::string err;
    int code = 0;
    try {
        solve_device(s, o, &tmp);
    } catch (const Error& e) {
        if (e.code != VEMD_ERR_NOT_CONVERGED) throw;
        err = e.what();
        code = e.code;
    }
    auto t1 = std::chrono::steady_clock::now();
    tmp.seconds = std::chrono::duration<double>(t1 - t0).count();
    tmp.iter_ms = tmp.iterations > 0 ? 1e3 * tmp.seconds / tmp.iterations : 0.0;
    s.last_info = tmp;
    HIPCHK(hipMemcpyAsync(bi.data(), d.x64, L.nvec * sizeof(double), hipMemcpyDeviceToHost, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    if (sol_ref) s.to_reference(bi.data(), sol_ref, false);
    if (info) *info = tmp;
    if (code) throw Error(code, err);
}

void dev_solve_int(Solver& s, const vemd_solve_opts& o, const void* b_dev, void* x_dev, vemd_solve_info* info) {
    if (!s.dev) throw Error(VEMD_ERR_STATE, "solver not uploaded");
    DeviceState& d = *s.dev;
    HIPCHK(hipSetDevice(d.device));
    const int64_t n = s.layout.nvec;
    auto t0 = std::chrono::steady_clock::now();
    // b_dev may still be being written on the caller's stream
    ScratchUse use(d, d.stream, true, (hipStream_t)o.stream);
    HIPCHK(hipMemcpyAsync(d.b64, b_dev ? b_dev : d.bprob, n * sizeof(double), hipMemcpyDeviceToDevice, d.stream));
    vemd_solve_info tmp{};
    std::string err;
    int code = 0;
    try {
        solve_device(s, o, &tmp);
    } catch (const Error& e) {
        if (e.code != VEMD_ERR_NOT_CONVERGED) throw;
        err = e.what();
        code = e.code;
    }
    if (x_dev) HIPCHK(hipMemcpyAsync(x_dev, d.x64, n * sizeof(double), hipMemcpyDeviceToDevice, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    auto t1 = std::chrono::steady_clock::now();
    tmp.seconds = std::chrono::duration<double>(t1 - t0).count();
    tmp.iter_ms = tmp.iterations > 0 ? 1e3 * tmp.seconds / tmp.iterations : 0.0;
    s.last_info = tmp;
    if (info) *info = tmp;
    if (code) throw Error(code, err);
}

void dev_adjoint(Solver& s, const vemd_solve_opts& o, const double* sol_ref, const double* dldx_ref, double* grad,
                 vemd_solve_info* info) {
    if (!s.dev) throw Error(VEMD_ERR_STATE, "solver not uploaded");
    DeviceState& d = *s.dev;
    HIPCHK(hipSetDevice(d.device));
    ScratchUse use(d, d.stream);
    const Layout& L = s.layout;
    // adjoint rhs in internal form: the internal operator is S K~ with S the
    // row scaling, symmetric; K~^T lambda~ = g  <=>  (S K~) (S^{-1}... ) -- we
    // solve K_int mu = g_int with g_int = g (no row scaling: K_int symmetric,
    // lambda~ = S mu).
    std::vector<double> gi(L.nvec, 0.0), xi(L.nvec, 0.0);
    s.to_internal(dldx_ref, gi.data(), false);
    s.to_internal(sol_ref, xi.data(), false);
    HIPCHK(hipMemcpyAsync(d.b64, gi.data(), L.nvec * sizeof(double), hipMemcpyHostToDevice, d.stream));
    vemd_solve_info tmp{};
    solve_device(s, o, &tmp);
    // grad_E = -mu^T (dK_int/dtheta_E) x = mu_E^T (exp(-theta_E) A_E) x_E
    HIPCHK(hipMemcpyAsync(d.y64, xi.data(), L.nvec * sizeof(double), hipMemcpyHostToDevice, d.stream));
    double* g = (double*)dalloc(d, s.mesh->nc * sizeof(double));
    ApplyArgs<double> a = make_args<double>(s, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, false);
    if (L.k == 0)
        hipLaunchKernelGGL((k_adjoint<double, 0>), dim3(d.ntiles), dim3(64), 0, d.stream, a, d.x64, d.y64,
                           d.cell_of_gpos, g);
    else if (L.k == 1)
        hipLaunchKernelGGL((k_adjoint<double, 1>), dim3(d.ntiles), dim3(64), 0, d.stream, a, d.x64, d.y64,
                           d.cell_of_gpos, g);
    else if (L.k == 2)
        hipLaunchKernelGGL((k_adjoint<double, 2>), dim3(d.ntiles), dim3(64), 0, d.stream, a, d.x64, d.y64,
                           d.cell_of_gpos, g);
    else
        launch_adjoint_k3(a, d.ntiles, d.stream, d.x64, d.y64, d.cell_of_gpos, g);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(grad, g, s.mesh->nc * sizeof(double), hipMemcpyDeviceToHost, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    HIPCHK(hipFree(g));
    d.allocs.pop_back();
    if (info) *info = tmp;
}

}  // namespace vemd
