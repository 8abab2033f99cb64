// Device rebuild of the multigrid hierarchy's values after a theta change
// (the host algorithm of amg.cpp with its patterns kept; see mgbuild.h).
#include "mgbuild.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "solver.h"

// Every value here is computed with the host build's separately rounded
// products and sums: no multiply-add in this file is fused.
#pragma clang fp contract(off)

namespace vemd {

MgBuild::~MgBuild() {
    for (void* p : allocs) (void)hipFree(p);
}

static void* mg_alloc(MgBuild& B, size_t bytes) {
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, std::max<size_t>(bytes, 256)));
    B.allocs.push_back(p);
    B.bytes += (int64_t)bytes;
    return p;
}

template <typename S>
static S* mg_upload(MgBuild& B, const std::vector<S>& v) {
    S* p = (S*)mg_alloc(B, v.size() * sizeof(S));
    if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(S), hipMemcpyHostToDevice));
    return p;
}

static CsrPat mg_pattern(MgBuild& B, const Csr& A) {
    CsrPat c;
    c.nrows = A.nrows;
    c.ncols = A.ncols;
    c.nnz = A.nnz();
    c.rp = mg_upload(B, A.rp);
    c.ci = mg_upload(B, A.ci);
    return c;
}

constexpr int NB = 256;  // blocks of the level-0 diagonal sum (as on the host)

static dim3 grid1(int64_t n) { return dim3((unsigned)std::max<int64_t>(1, (n + 255) / 256)); }

// ---------------------------------------------------------------------------
// Kernels (one thread per row unless stated; rows are summed in the host's
// entry order)
// ---------------------------------------------------------------------------

// Level 0: rows g*Q .. g*Q+Q-1 of the Schur approximation, written by the
// thread of store position g (schur_values on the host): the cell's own block
// accumulated in registers, the other cells' blocks in place.
template <int Q>
__global__ __launch_bounds__(256) void k_schur(int64_t cpad, const int* __restrict__ cell_of_gpos,
                                               const int64_t* __restrict__ rp, double* __restrict__ v,
                                               const int* __restrict__ nbk, const int* __restrict__ self_slot,
                                               const int* __restrict__ n1, const int64_t* __restrict__ v0,
                                               const int64_t* __restrict__ w0, const int64_t* __restrict__ tflux,
                                               const int* __restrict__ oslot, const int64_t* __restrict__ ow,
                                               const int* __restrict__ on2, const double* __restrict__ W,
                                               const double* __restrict__ minv) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= cpad) return;
    const int c = cell_of_gpos[g];
    if (c < 0) {
        for (int r = 0; r < Q; ++r) v[rp[g * Q + r]] = 1.0;
        return;
    }
    const int nb = nbk[g];
    int64_t row[Q];
#pragma unroll
    for (int r = 0; r < Q; ++r) {
        row[r] = rp[g * Q + r];
        for (int e = 0; e < nb * Q; ++e) v[row[r] + e] = 0.0;
    }
    const int n = n1[c];
    const int64_t d0 = v0[c];
    const double* Wc = W + w0[c];
    const int ss = self_slot[c];
    double sb[Q * Q];
#pragma unroll
    for (int e = 0; e < Q * Q; ++e) sb[e] = 0.0;
    for (int i = 0; i < n; ++i) {
        const int64_t t = tflux[d0 + i];
        if (t < 0) continue;
        const double dinv = minv[t];
        double wc[Q];
#pragma unroll
        for (int r = 0; r < Q; ++r) wc[r] = Wc[r * n + i];
#pragma unroll
        for (int r = 0; r < Q; ++r)
#pragma unroll
            for (int s = 0; s < Q; ++s) sb[r * Q + s] += wc[r] * wc[s] * dinv;
        const int os = oslot[d0 + i];
        if (os < 0) continue;
        const double* Wo = W + ow[d0 + i];
        const int n2 = on2[d0 + i];
#pragma unroll
        for (int r = 0; r < Q; ++r)
#pragma unroll
            for (int s = 0; s < Q; ++s) v[row[r] + os * Q + s] += wc[r] * Wo[s * n2] * dinv;
    }
#pragma unroll
    for (int r = 0; r < Q; ++r)
#pragma unroll
        for (int s = 0; s < Q; ++s) v[row[r] + ss * Q + s] = sb[r * Q + s];
}

__global__ void k_dinv(int64_t n, const int64_t* __restrict__ diag, const double* __restrict__ v,
                       double* __restrict__ dinv) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double d = 1.0;
    const int64_t p = diag[i];
    if (p >= 0 && v[p] != 0.0) d = 1.0 / v[p];
    dinv[i] = d;
}

// Power iteration, one step, on the level's operator in the FP64 cycle's
// sliced-ELLPACK form (blocks of B rows; each row's entries in ascending
// column order, as in the host's compressed rows, then zero padding):
// y = D^-1 A x, and the sums of y^2 and x^2 over each group of 64 rows by the
// butterfly (lane l adds lane l ^ o for o = 32, ..., 1) in gy, gx.  Skipped
// once the iteration stopped.
template <int B>
__global__ __launch_bounds__(256) void k_pspmv(Sell<double> A, int64_t n, const double* __restrict__ dinv,
                                               const double* __restrict__ x, double* __restrict__ y,
                                               double* __restrict__ gy, double* __restrict__ gx,
                                               const double* __restrict__ sc) {
    __shared__ double sy[256 * B], sx[256 * B];
    if (sc[2] != 0.0) return;
    const int64_t br = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double out[B];
#pragma unroll
    for (int r = 0; r < B; ++r) out[r] = 0.0;
    if (br < A.nrows) {
        const int64_t sl = br >> 6, l = br & 63;
        const int w = A.width[sl];
        const int64_t o = A.off[sl];
        for (int j = 0; j < w; ++j) {
            const int64_t c = (int64_t)A.col[o + (int64_t)j * 64 + l] * B;
            const double* vb = A.val + (o + (int64_t)j * 64) * (B * B) + l;
            double xc[B];
#pragma unroll
            for (int cc = 0; cc < B; ++cc) xc[cc] = x[c + cc];
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
                for (int cc = 0; cc < B; ++cc) out[r] += vb[(r * B + cc) * 64] * xc[cc];
        }
    }
#pragma unroll
    for (int r = 0; r < B; ++r) {
        const int64_t i = br * B + r;
        double yy = 0.0, xx = 0.0;
        if (i < n) {
            const double yi = out[r] * dinv[i];
            y[i] = yi;
            yy = yi * yi;
            xx = x[i] * x[i];
        }
        sy[threadIdx.x * B + r] = yy;
        sx[threadIdx.x * B + r] = xx;
    }
    __syncthreads();
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    for (int gl = wv; gl < 4 * B; gl += 4) {
        double a = sy[gl * 64 + ln], b = sx[gl * 64 + ln];
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o, 64);
            b += __shfl_xor(b, o, 64);
        }
        const int64_t g = (int64_t)blockIdx.x * 4 * B + gl;
        if (ln == 0 && g * 64 < n) {
            gy[g] = a;
            gx[g] = b;
        }
    }
}

// The norms from the group sums (G groups): each run of 64 group sums by the
// butterfly, those added in order by thread 0 (chunks of FINCH runs through
// shared memory); then sc = [lam, 1/|y|, stop].
constexpr int FINCH = 1024;
__global__ __launch_bounds__(256) void k_pfin(int64_t G, const double* __restrict__ gy, const double* __restrict__ gx,
                                              double* __restrict__ sc) {
    __shared__ double hy[FINCH], hx[FINCH];
    if (sc[2] != 0.0) return;
    const int64_t H = (G + 63) / 64;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    double ny = 0.0, nx = 0.0;
    for (int64_t h0 = 0; h0 < H; h0 += FINCH) {
        for (int64_t h = h0 + w; h < std::min<int64_t>(H, h0 + FINCH); h += 4) {
            const int64_t g = h * 64 + l;
            double a = g < G ? gy[g] : 0.0, b = g < G ? gx[g] : 0.0;
            for (int o = 32; o > 0; o >>= 1) {
                a += __shfl_xor(a, o, 64);
                b += __shfl_xor(b, o, 64);
            }
            if (l == 0) {
                hy[h - h0] = a;
                hx[h - h0] = b;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int64_t h = h0; h < std::min<int64_t>(H, h0 + FINCH); ++h) {
                ny += hy[h - h0];
                nx += hx[h - h0];
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (!(ny > 0.0)) {
            sc[2] = 1.0;
        } else {
            sc[0] = std::sqrt(ny / nx);
            sc[1] = 1.0 / std::sqrt(ny);
        }
    }
}

__global__ void k_pscale(int64_t n, const double* __restrict__ y, double* __restrict__ x,
                         const double* __restrict__ sc) {
    if (sc[2] != 0.0) return;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    x[i] = y[i] * sc[1];
}

__global__ void k_preset(double* sc) {
    sc[0] = 0.0;
    sc[1] = 0.0;
    sc[2] = 0.0;
}

// Smoothed prolongator, row r: the products A(r,k) Pt(k,.) accumulated in
// P's slots (Pt has at most one entry per row), then
// P(r,c) = Pt(r,c) - omega dinv(r) (A Pt)(r,c), omega = 4 / (3 lmax).
__global__ void k_pvals(int64_t n, CsrPat A, const double* __restrict__ Av, CsrPat Pt, const double* __restrict__ Ptv,
                        CsrPat P, double* __restrict__ Pv, const double* __restrict__ dinv,
                        const double* __restrict__ sc) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    const int64_t p0 = P.rp[r], p1 = P.rp[r + 1];
    for (int64_t s = p0; s < p1; ++s) Pv[s] = 0.0;
    for (int64_t p = A.rp[r]; p < A.rp[r + 1]; ++p) {
        const int k = A.ci[p];
        const int64_t t = Pt.rp[k];
        if (t == Pt.rp[k + 1]) continue;
        const int c = Pt.ci[t];
        int64_t lo = p0, hi = p1;
        while (lo < hi) {
            const int64_t m = (lo + hi) >> 1;
            if (P.ci[m] < c) lo = m + 1;
            else hi = m;
        }
        Pv[lo] += Av[p] * Ptv[t];
    }
    const double lmax = 1.1 * sc[0];
    const double omega = 4.0 / (3.0 * lmax);
    const double f = omega * dinv[r];
    const int64_t t0 = Pt.rp[r], t1 = Pt.rp[r + 1];
    for (int64_t s = p0; s < p1; ++s) {
        double v = 0.0;
        if (t0 < t1 && Pt.ci[t0] == P.ci[s]) v += Ptv[t0];
        v -= f * Pv[s];
        Pv[s] = v;
    }
}

__global__ void k_permute(int64_t nnz, const int* __restrict__ perm, const double* __restrict__ src,
                          double* __restrict__ dst) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < nnz) dst[e] = src[perm[e]];
}

// C = A B on C's (known) pattern: one wavefront per row i, each lane owning
// entries of C's row; entry (i, j) adds A(i,k) B(k,j) over A's row in order
// (B(k,j) found by bisection of B's row k), the host product's order.
__global__ __launch_bounds__(256) void k_spgemm(int64_t n, CsrPat A, const double* __restrict__ Av, CsrPat Bm,
                                                const double* __restrict__ Bv, CsrPat C, double* __restrict__ Cv) {
    const int64_t i = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const int64_t c0 = C.rp[i], c1 = C.rp[i + 1], a0 = A.rp[i], a1 = A.rp[i + 1];
    for (int64_t s = c0 + lane; s < c1; s += 64) {
        const int j = C.ci[s];
        double acc = 0.0;
        for (int64_t p = a0; p < a1; ++p) {
            const int k = A.ci[p];
            const int64_t t1 = Bm.rp[k + 1];
            int64_t lo = Bm.rp[k], hi = t1;
            while (lo < hi) {
                const int64_t m = (lo + hi) >> 1;
                if (Bm.ci[m] < j) lo = m + 1;
                else hi = m;
            }
            if (lo < t1 && Bm.ci[lo] == j) acc += Av[p] * Bv[lo];
        }
        Cv[s] = acc;
    }
}

__global__ void k_spmv(int64_t n, const int64_t* __restrict__ rp, const int* __restrict__ ci,
                       const double* __restrict__ v, const double* __restrict__ x, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p) s += v[p] * x[ci[p]];
    y[i] = s;
}

// Sum of the level-0 diagonal over the rows of live store positions, in NB
// blocks of rows summed in order.
__global__ __launch_bounds__(NB) void k_sdiag(int64_t n, int q, const int* __restrict__ cell_of_gpos,
                                              const int64_t* __restrict__ diag, const double* __restrict__ v,
                                              double* __restrict__ out) {
    __shared__ double ps[NB];
    const int64_t bl = (n + NB - 1) / NB;
    const int k = threadIdx.x;
    double s = 0.0;
    const int64_t e = std::min<int64_t>(n, (k + 1) * bl);
    for (int64_t i = k * bl; i < e; ++i)
        if (cell_of_gpos[i / q] >= 0 && diag[i] >= 0) s += v[diag[i]];
    ps[k] = s;
    __syncthreads();
    if (k == 0) {
        double t = 0.0;
        for (int b = 0; b < NB; ++b) t += ps[b];
        *out = t;
    }
}

// ---------------------------------------------------------------------------
// Refill of an uploaded cycle from the rebuilt values
// ---------------------------------------------------------------------------

// Entry p of row i, as stored: x = v[p] (mode 0), x / (rs[i] cs[c]) (1),
// x rs[i] (2), x cs[c] (3), x / cs[c] (4), c the entry's column.
__device__ __forceinline__ double fill_value(int mode, double x, int64_t i, int c, const double* rs, const double* cs) {
    switch (mode) {
        case 1: return x / (rs[i] * cs[c]);
        case 2: return x * rs[i];
        case 3: return x * cs[c];
        case 4: return x / cs[c];
        default: return x;
    }
}

// Largest magnitude and a non-finite flag of the stored values, per block of
// 256 rows (the maximum does not depend on the order).
__global__ __launch_bounds__(256) void k_fstat(int64_t n, const int64_t* __restrict__ rp, const int* __restrict__ ci,
                                               const double* __restrict__ v, int mode, const double* __restrict__ rs,
                                               const double* __restrict__ cs, double* __restrict__ part) {
    __shared__ double sm[256], sb[256];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double m = 0.0, bad = 0.0;
    if (i < n)
        for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
            const double x = fill_value(mode, v[p], i, ci[p], rs, cs);
            if (!std::isfinite(x)) bad = 1.0;
            else m = std::max(m, std::fabs(x));
        }
    sm[threadIdx.x] = m;
    sb[threadIdx.x] = bad;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sm[threadIdx.x] = std::max(sm[threadIdx.x], sm[threadIdx.x + o]);
            sb[threadIdx.x] = std::max(sb[threadIdx.x], sb[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sm[0];
        part[2 * blockIdx.x + 1] = sb[0];
    }
}

__global__ __launch_bounds__(256) void k_fstat_fin(int nb, const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double sm[256], sb[256];
    double m = 0.0, bad = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        m = std::max(m, part[2 * b]);
        bad = std::max(bad, part[2 * b + 1]);
    }
    sm[threadIdx.x] = m;
    sb[threadIdx.x] = bad;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sm[threadIdx.x] = std::max(sm[threadIdx.x], sm[threadIdx.x + o]);
            sb[threadIdx.x] = std::max(sb[threadIdx.x], sb[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sm[0];
        out[1] = sb[0];
    }
}

// out[at[p]] (at null: out[p]) = the stored value of entry p times hinv, in O.
template <typename O>
__global__ void k_fill(int64_t n, const int64_t* __restrict__ rp, const int* __restrict__ ci,
                       const double* __restrict__ v, const uint32_t* __restrict__ at, int mode,
                       const double* __restrict__ rs, const double* __restrict__ cs, double hinv, O* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
        const double x = fill_value(mode, v[p], i, ci[p], rs, cs) * hinv;
        out[at ? (int64_t)at[p] : p] = (O)x;
    }
}

// dsq = D^-1/2 of a level (1 / sqrt(dinv)) and a flag for a diagonal that is
// not positive or gives a non-finite dsq.
__global__ void k_dsq(int64_t n, const double* __restrict__ dinv, double* __restrict__ dsq, double* __restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double d = 1.0 / std::sqrt(dinv[i]);
    dsq[i] = d;
    if (!(dinv[i] > 0.0) || !std::isfinite(d)) *bad = 1.0;
}

template <typename T>
__global__ void k_cast(int64_t n, const double* __restrict__ src, T* __restrict__ dst, int recip) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = (T)(recip ? 1.0 / src[i] : src[i]);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// Storage statistics (largest magnitude, non-finite flag) of the FP32 cycle's
// finest level: plain P, R, M; P, R, M as stored on the scaled level; D^-1/2.
enum { ST_P0, ST_R0, ST_M0, ST_PS, ST_RS, ST_MS, ST_DSQ, ST_N };

MgBuild* mg_build_create(const Solver& s, const int* cell_of_gpos_dev) {
    if (!s.amg || !s.splan) throw Error(VEMD_ERR_INTERNAL, "device multigrid rebuild without a host hierarchy");
    const Amg& H = *s.amg;
    const Mesh& m = *s.mesh;
    const Setup& S = *s.setup;
    const Layout& L = s.layout;
    const SchurPlan& SP = *s.splan;
    std::unique_ptr<MgBuild> B(new MgBuild());
    B->q = L.q;
    B->cpad = L.cpad;
    B->cell_of_gpos = cell_of_gpos_dev;
    B->meanzero = H.rho > 0.0;
    if (SP.pattern.nrows != H.lv[0].A.nrows || SP.pattern.nnz() != H.lv[0].A.nnz())
        throw Error(VEMD_ERR_INTERNAL, "device multigrid rebuild: level-0 pattern differs");
    // level-0 operator: per cell its flux dofs, W and the blocks of the other cells
    std::vector<int> n1(m.nc);
    std::vector<int64_t> v0(m.nc), w0(m.nc);
    const int64_t nv = S.ofs.v_off[m.nc];
    std::vector<int64_t> ow(nv, 0);
    std::vector<int> on2(nv, 0);
    for (int c = 0; c < m.nc; ++c) {
        n1[c] = S.sh.nflux(m.npc(c));
        v0[c] = S.ofs.v_off[c];
        w0[c] = S.ofs.w_off[c];
    }
    parallel_for(m.nc, [&](int64_t b, int64_t e) {
        for (int64_t c = b; c < e; ++c) {
            for (int i = 0; i < n1[c]; ++i) {
                const int64_t d = v0[c] + i;
                if (SP.other_slot[d] < 0) continue;
                const int ed = m.cell_edges[m.off[c] + i % m.npc((int)c)];
                const int o = m.edge_cells[2 * ed] == (int)c ? m.edge_cells[2 * ed + 1] : m.edge_cells[2 * ed];
                ow[d] = S.ofs.w_off[o] + SP.other_i[d];
                on2[d] = S.sh.nflux(m.npc(o));
            }
        }
    });
    B->nbk = mg_upload(*B, SP.nbk);
    B->self_slot = mg_upload(*B, SP.self_slot);
    B->n1 = mg_upload(*B, n1);
    B->v0 = mg_upload(*B, v0);
    B->w0 = mg_upload(*B, w0);
    B->tflux = mg_upload(*B, SP.tflux);
    B->oslot = mg_upload(*B, SP.other_slot);
    B->ow = mg_upload(*B, ow);
    B->on2 = mg_upload(*B, on2);
    B->W = mg_upload(*B, S.W);
    // the multiplier row and its sum over the first moments (theta-independent)
    std::vector<double> l0;
    if (B->meanzero) {
        l0.assign(H.lv[0].A.nrows, 0.0);
        for (int64_t g = 0; g < L.cpad; ++g) {
            const int c = L.cell_of_gpos[g];
            if (c < 0) continue;
            ++B->nlive;
            for (int r = 0; r < L.q; ++r) {
                l0[g * L.q + r] = S.L[(size_t)c * L.q + r];
                if (r == 0) B->ln0 += l0[g * L.q + r];
            }
        }
    } else {
        for (int64_t g = 0; g < L.cpad; ++g)
            if (L.cell_of_gpos[g] >= 0) ++B->nlive;
    }
    B->lv.resize(H.lv.size());
    for (size_t l = 0; l < H.lv.size(); ++l) {
        const AmgLevel& h = H.lv[l];
        MgLevelBuild& b = B->lv[l];
        const int64_t n = h.A.nrows;
        b.A = mg_pattern(*B, h.A);
        b.Av = (double*)mg_alloc(*B, h.A.nnz() * sizeof(double));
        std::vector<int64_t> diag(n, -1);
        for (int64_t i = 0; i < n; ++i)
            for (int64_t p = h.A.rp[i]; p < h.A.rp[i + 1]; ++p)
                if (h.A.ci[p] == i) diag[i] = p;
        b.diag = mg_upload(*B, diag);
        b.dinv = (double*)mg_alloc(*B, n * sizeof(double));
        if ((int64_t)h.ev.size() != n) throw Error(VEMD_ERR_INTERNAL, "device multigrid rebuild: no power-iteration vector");
        b.ev = mg_upload(*B, h.ev);
        b.y = (double*)mg_alloc(*B, n * sizeof(double));
        if (B->meanzero) b.l = (double*)mg_alloc(*B, n * sizeof(double));
        b.lmax = h.lmax;
        if (l + 1 == H.lv.size()) break;
        if (h.Pt.nrows != n || h.P.nrows != n) throw Error(VEMD_ERR_INTERNAL, "device multigrid rebuild: no aggregates");
        for (int64_t i = 0; i < n; ++i)
            if (h.Pt.rp[i + 1] - h.Pt.rp[i] > 1)
                throw Error(VEMD_ERR_INTERNAL, "device multigrid rebuild: tentative prolongator row with two entries");
        b.Pt = mg_pattern(*B, h.Pt);
        b.Ptv = mg_upload(*B, h.Pt.v);
        b.P = mg_pattern(*B, h.P);
        b.Pv = (double*)mg_alloc(*B, h.P.nnz() * sizeof(double));
        b.R = mg_pattern(*B, h.R);
        b.Rv = (double*)mg_alloc(*B, h.R.nnz() * sizeof(double));
        b.M = mg_pattern(*B, h.M);
        b.Mv = (double*)mg_alloc(*B, h.M.nnz() * sizeof(double));
        // R = P^T: entry e of R (row c, column r) is P's entry (r, c)
        std::vector<int> perm(h.R.nnz());
        {
            std::vector<int64_t> pos(h.R.rp.begin(), h.R.rp.end() - 1);
            for (int64_t r = 0; r < n; ++r)
                for (int64_t p = h.P.rp[r]; p < h.P.rp[r + 1]; ++p) perm[pos[h.P.ci[p]]++] = (int)p;
        }
        if (h.P.nnz() >= INT32_MAX) throw Error(VEMD_ERR_MESH, "multigrid prolongator too large");
        b.rperm = mg_upload(*B, perm);
    }
    if (B->meanzero) HIPCHK(hipMemcpy(B->lv[0].l, l0.data(), l0.size() * sizeof(double), hipMemcpyHostToDevice));
    // per level [lam, 1/|y|, stop], the diagonal sum, the storage statistics
    B->scal = (double*)mg_alloc(*B, (3 * B->lv.size() + 1 + 2 * ST_N) * sizeof(double));
    B->npart = (int)((H.lv[0].A.nrows + 255) / 256);
    B->part = (double*)mg_alloc(*B, 2 * (size_t)B->npart * sizeof(double));
    int64_t nmax = 0;
    for (const AmgLevel& h : H.lv) nmax = std::max(nmax, h.A.nrows);
    B->gpart = (double*)mg_alloc(*B, 2 * (size_t)((nmax + 63) / 64) * sizeof(double));
    return B.release();
}

// Values of the level's matrices and storage statistics read back at once.
struct FillStat {
    double amax = 0.0, bad = 0.0;
};

// Launches the statistics of one stored matrix into out (2 doubles).
static void fstat(MgBuild& B, int64_t n, const int64_t* rp, const int* ci, const double* v, int mode, const double* rs,
                  const double* cs, double* out, hipStream_t st) {
    const int nb = (int)((n + 255) / 256);
    if (nb > B.npart) throw Error(VEMD_ERR_INTERNAL, "device multigrid rebuild: statistics buffer too small");
    hipLaunchKernelGGL(k_fstat, grid1(n), dim3(256), 0, st, n, rp, ci, v, mode, rs, cs, B.part);
    hipLaunchKernelGGL(k_fstat_fin, dim3(1), dim3(256), 0, st, nb, (const double*)B.part, out);
}

template <typename O>
static void fill(int64_t n, const int64_t* rp, const int* ci, const double* v, const uint32_t* at, int mode,
                 const double* rs, const double* cs, double hinv, O* out, hipStream_t st) {
    hipLaunchKernelGGL((k_fill<O>), grid1(n), dim3(256), 0, st, n, rp, ci, v, at, mode, rs, cs, hinv, out);
}

// Device copy of a value map (made on first use, freed with the cycle).
template <typename T>
static const uint32_t* map_dev(AmgDev<T>& D, ValMap& vm) {
    if (vm.at.empty()) return nullptr;
    if (!vm.dev) {
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, vm.at.size() * sizeof(uint32_t)));
        D.allocs.push_back(p);
        HIPCHK(hipMemcpy(p, vm.at.data(), vm.at.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        vm.dev = (uint32_t*)p;
    }
    return vm.dev;
}

// The uploaded cycle's shapes against the rebuilt hierarchy (as amg_refresh).
template <typename T>
static bool shapes_match(const AmgDev<T>& D, const Amg& H) {
    if (D.lv.size() != H.lv.size() || D.nc != H.ncoarse) return false;
    for (size_t l = 0; l < H.lv.size(); ++l) {
        const AmgLevel& h = H.lv[l];
        const AmgLevelDev<T>& d = D.lv[l];
        if (h.A.nrows != d.n || h.A.nnz() != (int64_t)d.vA.at.size()) return false;
        if (d.sc && l != 0) return false;
        if (l + 1 == H.lv.size()) continue;
        if (h.P.nnz() != (int64_t)d.vP.at.size() || h.R.nnz() != (int64_t)d.vR.n) return false;
        if (d.M.nrows > 0 && (h.M.nnz() != (int64_t)d.vM.at.size() || h.R.nnz() != (int64_t)d.vRb.at.size()))
            return false;
    }
    return true;
}

template <typename T>
static void refill_level_values(AmgDev<T>& D, MgBuild& B, const Amg& H, const double* dsq, const double* stat,
                                hipStream_t st) {
    for (size_t l = 0; l < H.lv.size(); ++l) {
        AmgLevelDev<T>& d = D.lv[l];
        MgLevelBuild& b = B.lv[l];
        const bool last = l + 1 == H.lv.size();
        const int64_t n = b.A.nrows;
        auto hinv_of = [](double amax) { return amax > 0.0 ? 1.0 / amax : 1.0; };
        auto hscale_of = [](double amax) { return amax > 0.0 ? (float)amax : 1.0f; };
        if (d.sc) {
            fill(n, b.A.rp, b.A.ci, b.Av, map_dev(D, d.vA), 1, dsq, dsq, 1.0, (_Float16*)d.A.hval, st);
            const bool htr = d.P.hval != nullptr;
            const int64_t nc = b.R.nrows;
            if (htr) {
                fill(n, b.P.rp, b.P.ci, b.Pv, map_dev(D, d.vP), 2, dsq, nullptr, hinv_of(stat[ST_PS * 2]),
                     (_Float16*)d.P.hval, st);
                d.P.hscale = hscale_of(stat[ST_PS * 2]);
                fill(nc, b.R.rp, b.R.ci, b.Rv, (const uint32_t*)nullptr, 3, nullptr, dsq, hinv_of(stat[ST_RS * 2]),
                     (_Float16*)d.R.hv, st);
                d.R.hscale = hscale_of(stat[ST_RS * 2]);
            } else {
                fill(n, b.P.rp, b.P.ci, b.Pv, map_dev(D, d.vP), 2, dsq, nullptr, 1.0, (T*)d.P.val, st);
                fill(nc, b.R.rp, b.R.ci, b.Rv, (const uint32_t*)nullptr, 3, nullptr, dsq, 1.0, (T*)d.R.v, st);
            }
            if (d.M.nrows > 0) {
                if (d.M.hv) {
                    fill(nc, b.M.rp, b.M.ci, b.Mv, map_dev(D, d.vM), 4, nullptr, dsq, hinv_of(stat[ST_MS * 2]),
                         (_Float16*)d.M.hv, st);
                    d.M.hscale = hscale_of(stat[ST_MS * 2]);
                } else {
                    fill(nc, b.M.rp, b.M.ci, b.Mv, map_dev(D, d.vM), 4, nullptr, dsq, 1.0, (T*)d.M.v, st);
                }
                if (d.Rb.hv) {
                    fill(nc, b.R.rp, b.R.ci, b.Rv, map_dev(D, d.vRb), 0, nullptr, nullptr, hinv_of(stat[ST_R0 * 2]),
                         (_Float16*)d.Rb.hv, st);
                    d.Rb.hscale = hscale_of(stat[ST_R0 * 2]);
                } else {
                    fill(nc, b.R.rp, b.R.ci, b.Rv, map_dev(D, d.vRb), 0, nullptr, nullptr, 1.0, (T*)d.Rb.v, st);
                }
            }
            // sc = 1 / dsq
            hipLaunchKernelGGL((k_cast<T>), grid1(n), dim3(256), 0, st, n, dsq, d.sc, 1);
        } else {
            // (the FP64 cycle's operators were filled during the build)
            if (!std::is_same<T, double>::value)
                fill(n, b.A.rp, b.A.ci, b.Av, map_dev(D, d.vA), 0, nullptr, nullptr, 1.0, (T*)d.A.val, st);
            if (!last) {
                const int64_t nc = b.R.nrows;
                fill(n, b.P.rp, b.P.ci, b.Pv, map_dev(D, d.vP), 0, nullptr, nullptr, 1.0, (T*)d.P.val, st);
                fill(nc, b.R.rp, b.R.ci, b.Rv, (const uint32_t*)nullptr, 0, nullptr, nullptr, 1.0, (T*)d.R.v, st);
                if (d.M.nrows > 0) {
                    fill(nc, b.M.rp, b.M.ci, b.Mv, map_dev(D, d.vM), 0, nullptr, nullptr, 1.0, (T*)d.M.v, st);
                    fill(nc, b.R.rp, b.R.ci, b.Rv, map_dev(D, d.vRb), 0, nullptr, nullptr, 1.0, (T*)d.Rb.v, st);
                }
            }
            hipLaunchKernelGGL((k_cast<T>), grid1(n), dim3(256), 0, st, n, (const double*)b.dinv, d.dinv, 0);
        }
        const double hi = H.lv[l].lmax, lo = H.lv[l].lmax / H.opt.cheb_ratio;
        d.centre = 0.5 * (hi + lo);
        d.half = 0.5 * (hi - lo);
        d.degree = l == 0 ? std::max(1, H.opt.degree0) : std::max(1, H.opt.degree);
    }
    HIPCHK(hipGetLastError());
}

bool mg_build_run(MgBuild& B, Solver& s, const double* minv_dev, AmgDev<double>& D64, AmgDev<float>& D32,
                  hipStream_t st) {
    Amg& H = *s.amg;
    if (!shapes_match(D64, H) || !shapes_match(D32, H) || D64.lv[0].sc) return false;
    auto t0 = std::chrono::steady_clock::now();
    const size_t nl = B.lv.size();
    // level 0: the Schur approximation for the new flux diagonal
    {
        MgLevelBuild& b = B.lv[0];
        auto schur = B.q == 6 ? k_schur<6> : B.q == 3 ? k_schur<3> : k_schur<1>;
        if (B.q != 1 && B.q != 3 && B.q != 6)
            throw Error(VEMD_ERR_INTERNAL, "device multigrid rebuild: 1, 3 or 6 pressure moments");
        hipLaunchKernelGGL(schur, grid1(B.cpad), dim3(256), 0, st, B.cpad, B.cell_of_gpos, (const int64_t*)b.A.rp,
                           b.Av, (const int*)B.nbk, (const int*)B.self_slot, (const int*)B.n1, (const int64_t*)B.v0,
                           (const int64_t*)B.w0, (const int64_t*)B.tflux, (const int*)B.oslot, (const int64_t*)B.ow,
                           (const int*)B.on2, (const double*)B.W, minv_dev);
        if (B.meanzero)
            hipLaunchKernelGGL(k_sdiag, dim3(1), dim3(NB), 0, st, b.A.nrows, B.q, B.cell_of_gpos,
                               (const int64_t*)b.diag, (const double*)b.Av, B.scal + 3 * nl);
    }
    for (size_t l = 0; l < nl; ++l) {
        MgLevelBuild& b = B.lv[l];
        const int64_t n = b.A.nrows;
        double* sc = B.scal + 3 * l;
        hipLaunchKernelGGL(k_dinv, grid1(n), dim3(256), 0, st, n, (const int64_t*)b.diag, (const double*)b.Av, b.dinv);
        // the level's operator into the FP64 cycle, whose sliced form the
        // power iteration reads
        AmgLevelDev<double>& d64 = D64.lv[l];
        fill(n, b.A.rp, b.A.ci, b.Av, map_dev(D64, d64.vA), 0, nullptr, nullptr, 1.0, (double*)d64.A.val, st);
        const int bs = d64.A.bs;
        if (bs != 1 && bs != 3 && bs != 6) throw Error(VEMD_ERR_INTERNAL, "device multigrid rebuild: block size");
        auto pspmv = bs == 6 ? k_pspmv<6> : bs == 3 ? k_pspmv<3> : k_pspmv<1>;
        // spectral bound: 25 power-iteration steps from the previous iterate
        hipLaunchKernelGGL(k_preset, dim3(1), dim3(1), 0, st, sc);
        const int64_t G = (n + 63) / 64;
        for (int it = 0; it < 25; ++it) {
            hipLaunchKernelGGL(pspmv, grid1(d64.A.nrows), dim3(256), 0, st, d64.A, n, (const double*)b.dinv,
                               (const double*)b.ev, b.y, B.gpart, B.gpart + G, (const double*)sc);
            hipLaunchKernelGGL(k_pfin, dim3(1), dim3(256), 0, st, G, (const double*)B.gpart,
                               (const double*)(B.gpart + G), sc);
            hipLaunchKernelGGL(k_pscale, grid1(n), dim3(256), 0, st, n, (const double*)b.y, b.ev,
                               (const double*)sc);
        }
        if (l + 1 == nl) break;
        MgLevelBuild& c = B.lv[l + 1];
        hipLaunchKernelGGL(k_pvals, grid1(n), dim3(256), 0, st, n, b.A, (const double*)b.Av, b.Pt,
                           (const double*)b.Ptv, b.P, b.Pv, (const double*)b.dinv, (const double*)sc);
        hipLaunchKernelGGL(k_permute, grid1(b.R.nnz), dim3(256), 0, st, b.R.nnz, (const int*)b.rperm,
                           (const double*)b.Pv, b.Rv);
        hipLaunchKernelGGL(k_spgemm, grid1(b.R.nrows * 64), dim3(256), 0, st, b.R.nrows, b.R, (const double*)b.Rv,
                           b.A, (const double*)b.Av, b.M, b.Mv);
        hipLaunchKernelGGL(k_spgemm, grid1(b.M.nrows * 64), dim3(256), 0, st, b.M.nrows, b.M, (const double*)b.Mv,
                           b.P, (const double*)b.Pv, c.A, c.Av);
        if (B.meanzero)
            hipLaunchKernelGGL(k_spmv, grid1(b.R.nrows), dim3(256), 0, st, b.R.nrows, (const int64_t*)b.R.rp,
                               (const int*)b.R.ci, (const double*)b.Rv, (const double*)b.l, c.l);
    }
    HIPCHK(hipGetLastError());
    // storage statistics of the FP32 cycle's finest level (scaled, half precision)
    AmgLevelDev<float>& f0 = D32.lv[0];
    double* stat = B.scal + 3 * nl + 1;  // ST_N pairs follow the scalars
    double* dsq = nullptr;
    if (f0.sc) {
        MgLevelBuild& b = B.lv[0];
        const int64_t n = b.A.nrows, nc = b.R.nrows;
        dsq = b.y;  // the power iteration's scratch is free now
        HIPCHK(hipMemsetAsync(stat, 0, 2 * ST_N * sizeof(double), st));
        hipLaunchKernelGGL(k_dsq, grid1(n), dim3(256), 0, st, n, (const double*)b.dinv, dsq, stat + 2 * ST_DSQ + 1);
        fstat(B, n, b.P.rp, b.P.ci, b.Pv, 0, nullptr, nullptr, stat + 2 * ST_P0, st);
        fstat(B, nc, b.R.rp, b.R.ci, b.Rv, 0, nullptr, nullptr, stat + 2 * ST_R0, st);
        fstat(B, nc, b.M.rp, b.M.ci, b.Mv, 0, nullptr, nullptr, stat + 2 * ST_M0, st);
        fstat(B, n, b.P.rp, b.P.ci, b.Pv, 2, dsq, nullptr, stat + 2 * ST_PS, st);
        fstat(B, nc, b.R.rp, b.R.ci, b.Rv, 3, nullptr, dsq, stat + 2 * ST_RS, st);
        fstat(B, nc, b.M.rp, b.M.ci, b.Mv, 4, nullptr, dsq, stat + 2 * ST_MS, st);
    }
    // scalars, the coarsest operator and the restricted multiplier row to the host
    std::vector<double> scal(3 * nl + 1 + 2 * ST_N, 0.0);
    const MgLevelBuild& bl = B.lv.back();
    std::vector<double> cv(bl.A.nnz), lc(B.meanzero ? bl.A.nrows : 0);
    HIPCHK(hipMemcpyAsync(scal.data(), B.scal, scal.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(cv.data(), bl.Av, cv.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    if (B.meanzero)
        HIPCHK(hipMemcpyAsync(lc.data(), bl.l, lc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    const double* hst = scal.data() + 3 * nl + 1;
    if (f0.sc) {
        // the storage decisions of the upload must still hold (amg_refresh's checks)
        const bool htr = f0.P.hval != nullptr;
        if (hst[2 * ST_DSQ + 1] != 0.0) return false;
        if (htr && (hst[2 * ST_P0 + 1] != 0.0 || hst[2 * ST_R0 + 1] != 0.0)) return false;
        if (f0.M.nrows > 0 && ((f0.M.hv != nullptr) != (htr && hst[2 * ST_M0 + 1] == 0.0) ||
                               (f0.Rb.hv != nullptr) != (htr && hst[2 * ST_R0 + 1] == 0.0)))
            return false;
    }
    for (size_t l = 0; l < nl; ++l) {
        const double lam = scal[3 * l];
        if (!std::isfinite(lam)) return false;
        B.lv[l].lmax = 1.1 * lam;
    }
    // rank-one weight and the dense coarsest inverse (host, as amg_build)
    double rho = 0.0;
    if (B.meanzero) {
        const double sd = scal[3 * nl];
        if (!(sd > 0.0) || !(B.ln0 != 0.0)) return false;
        double c = 0.03;
        if (const char* e = std::getenv("VEMD_MG_RHO")) c = std::max(1e-12, std::atof(e));
        rho = c * (sd / ((double)B.nlive * B.q)) * (double)B.nlive / (B.ln0 * B.ln0);
    }
    const Csr& CA = H.lv.back().A;
    const int64_t nc = CA.nrows;
    std::vector<double> cinv(nc * nc, 0.0);
    for (int64_t i = 0; i < nc; ++i)
        for (int64_t p = CA.rp[i]; p < CA.rp[i + 1]; ++p) cinv[i * nc + CA.ci[p]] = cv[p];
    if (B.meanzero)
        for (int64_t i = 0; i < nc; ++i)
            for (int64_t j = 0; j < nc; ++j) cinv[i * nc + j] += rho * lc[i] * lc[j];
    dense_inverse(cinv, nc);
    // the host hierarchy keeps its patterns; its scalars follow the rebuild
    for (size_t l = 0; l < nl; ++l) H.lv[l].lmax = B.lv[l].lmax;
    H.rho = rho;
    H.coarse_inv = cinv;
    B.rho = rho;
    B.cinv = cinv;
    s.amg_lam_minv.assign(B.meanzero ? 1 : 0, B.meanzero ? 1.0 / rho : 0.0);
    s.amg_host_stale = true;
    // refill both cycles
    refill_level_values(D64, B, H, nullptr, hst, st);
    refill_level_values(D32, B, H, dsq, hst, st);
    {
        std::vector<float> cf(cinv.begin(), cinv.end());
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipMemcpy(D64.cinv, cinv.data(), cinv.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(D32.cinv, cf.data(), cf.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    if (std::getenv("VEMD_VERBOSE"))
        std::fprintf(stderr, "[vemd] multigrid values rebuilt on the device %.3fs (refill and coarsest inverse %.3fs)\n",
                     std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(),
                     std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count());
    return true;
}

int64_t mg_build_compare(const MgBuild& B, const Solver& s, double* maxrel) {
    const Amg& H = *s.amg;
    int64_t bad = 0;
    double mr = 0.0;
    auto cmp = [&](const double* dev, const std::vector<double>& host) {
        std::vector<double> v(host.size());
        if (!v.empty()) HIPCHK(hipMemcpy(v.data(), dev, v.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i] != host[i]) {
                ++bad;
                const double den = std::max(std::fabs(host[i]), 1e-300);
                mr = std::max(mr, std::fabs(v[i] - host[i]) / den);
            }
    };
    auto cmps = [&](double a, double b) {
        if (a != b) {
            ++bad;
            mr = std::max(mr, std::fabs(a - b) / std::max(std::fabs(b), 1e-300));
        }
    };
    if (B.lv.size() != H.lv.size()) return -1;
    for (size_t l = 0; l < H.lv.size(); ++l) {
        const AmgLevel& h = H.lv[l];
        const MgLevelBuild& b = B.lv[l];
        cmp(b.Av, h.A.v);
        cmp(b.dinv, h.dinv);
        cmp(b.ev, h.ev);
        cmps(b.lmax, h.lmax);
        if (l + 1 == H.lv.size()) break;
        cmp(b.Pv, h.P.v);
        cmp(b.Rv, h.R.v);
        cmp(b.Mv, h.M.v);
    }
    cmps(B.rho, H.rho);
    if (B.cinv.size() != H.coarse_inv.size()) return -1;
    for (size_t i = 0; i < B.cinv.size(); ++i) cmps(B.cinv[i], H.coarse_inv[i]);
    *maxrel = mr;
    return bad;
}

}  // namespace vemd
