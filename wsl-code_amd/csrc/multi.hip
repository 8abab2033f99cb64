// Multi-right-hand-side path on gfx950: up to 32 right-hand sides solved at
// once.  The element apply multiplies each cell's saddle-point block
// [[s A, -W^T], [-W, 0]] (m = n + q rows, m <= 26) by the cell's 32-column
// slice of the Krylov vectors with v_mfma_f32_32x32x2f32 (one wavefront per
// cell, FP32 products and accumulation); the Krylov vectors are interleaved
// ([entry][32] floats, 128 bytes per entry) so that every gathered row is one
// cache line.  The 32 MINRES recurrences run side by side with their own
// scalars and stopping tests, preconditioned by the same block-diagonal
// preconditioner (flux diagonal and a 32-column multigrid cycle), inside
// FP64 iterative refinement per column.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "devstate.h"

namespace vemd {

typedef float f16v __attribute__((ext_vector_type(16)));
constexpr int NR = MCOLS;     // right-hand-side columns
constexpr int OPMAX = 352;    // floats of the largest cell block (n = 23: A 276 + W 69, padded to 4)

// A valence group as the multi-RHS apply sees it: its real cells occupy the
// store positions [gpos0, gpos0 + count) (tile padding follows them).
struct MrGroup {
    int gpos0, count, ne, n;
    int64_t gop;    // first float of the group's cell blocks
    int64_t i_off;  // first entry of the group's edge-index block
};
constexpr int MR_MAXG = 32;

struct MultiState {
    float* op = nullptr;            // per cell: packed A (lower triangle) then W (q x n), padded to 4 floats
    MrGroup* mg = nullptr;          // group table of the apply
    int ngroups = 0;
    float *R[2] = {nullptr, nullptr}, *Z[2] = {nullptr, nullptr}, *WB[2] = {nullptr, nullptr};
    float *x = nullptr, *Az = nullptr, *Az1 = nullptr, *b = nullptr;  // [entry][NR]
    double *x64 = nullptr, *r64 = nullptr, *y64 = nullptr;  // [NR][nvec]
    MState* st = nullptr;
    MState* st_host = nullptr;
    double* rtol = nullptr;
    double* rtol_host = nullptr;
    double *pw = nullptr, *p1 = nullptr, *p2 = nullptr, *q1 = nullptr, *q2 = nullptr;
    double *part = nullptr, *dsum = nullptr, *hsum = nullptr;
    int* all_done = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_iters = 0;
    bool graph_amg = false;
    int pw_waves = 0;               // waves writing x.Kx partials (entries of pw per column)
    std::vector<MrGroup> mg_host;   // group table (per-valence launches)
    std::vector<int64_t> mr_cpw;   // per group: cells per wave, blocks, first wave
    std::vector<int> mr_blocks, mr_wave0;
    std::vector<void*> allocs;
};

static void* malloc_m(MultiState& m, size_t bytes) {
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, std::max<size_t>(bytes, 256)));
    m.allocs.push_back(p);
    return p;
}

void multi_invalidate(DeviceState& d) {
    if (!d.multi) return;
    if (d.multi->graph) (void)hipGraphExecDestroy(d.multi->graph);
    d.multi->graph = nullptr;
}

static void multi_release(MultiState* m) {
    if (!m) return;
    if (m->graph) (void)hipGraphExecDestroy(m->graph);
    for (void* p : m->allocs) (void)hipFree(p);
    if (m->st_host) (void)hipHostFree(m->st_host);
    if (m->rtol_host) (void)hipHostFree(m->rtol_host);
    if (m->hsum) (void)hipHostFree(m->hsum);
    delete m;
}

void multi_free(DeviceState& d) {
    multi_release(d.multi);
    d.multi = nullptr;
}

// ---------------------------------------------------------------------------
// MFMA element apply
// ---------------------------------------------------------------------------

struct MrArgs {
    const MrGroup* mg;
    int ngroups;
    const float* op;
    const uint32_t* idx;
    const float* scale;
    const float* X[2];        // input; with st, the buffer of each column's iteration parity
    float* Y;
    float* Y1;                // second sides of the edge dofs
    int64_t cbase, cpad;
    int64_t trash;            // the padding row (zeroed by each apply; no cell stores there)
    double* pw;               // per-wave, per-column partials of x.Kx (nullptr: not needed)
    const MState* st;         // per column
    const int* all_done;
    int off32;                // every row of X lies within 4 GiB of its buffer's start (32-bit gather offsets)
};

// The MFMA element apply on 32 interleaved columns: y = [[s A, -W^T], [-W,
// 0]] x per cell, one wavefront per cell, one launch per valence group with
// the valence a template argument (the MFMA step count, the row count and
// each lane's offsets into the staged operator are constants or loop
// invariants of a run of cells, and the kernel's registers are those of one
// valence body: 76-126 VGPRs, 4-5 waves per SIMD).
//
// v_mfma_f32_32x32x2f32 with the gathered rows as the A operand (lane l: row
// 2 kk + l / 32 of column l % 32) and the cell block as the B operand (lane
// l: entry (l % 32, 2 kk + l / 32)), so D = (block X)^T: lane l holds cell
// row l % 32 for the 16 columns 8 g + 4 (l / 32) + e (g, e < 4); the results
// pass through a per-wave LDS tile and leave as whole 128-byte rows.  A block
// of m rows fills m of the tile's 32 rows (a hexagon at k = 1: 18).
// ---------------------------------------------------------------------------

constexpr int MR_LDX = 36;  // row pitch of the per-wave K x tile (16-byte aligned rows)
#ifndef MR_WAVES
#define MR_WAVES 4  // occupancy bound of k_apply_mr (waves per SIMD)
#endif

template <int K, int NE>
struct MrC {
    static constexpr int KK = K + 1, Q = (K + 1) * (K + 2) / 2, NI = (Q - 1) + (Q - (K + 1));
    static constexpr int NED = KK * NE, N = NED + NI, M = N + Q, NA = N * (N + 1) / 2;
    static constexpr int NV4 = (NA + Q * N + 3) / 4, ZERO = NV4 * 4, KS = (M + 1) / 2;
    static_assert(ZERO < OPMAX, "staged operator and its zero entry must fit the LDS buffer");
    static_assert(M <= 32, "a cell block must fit the 32 rows of the MFMA tile");
};

// Loads of one cell (gp < 0: none): its packed operator (float4 per lane,
// two rounds), its scale and, for the lanes of edge rows, the raw edge index
// (the slot of other rows needs no load).  Every lane loads (indices clamped
// into the cell's data) and nothing loaded is used here: the slot is formed
// when the cell is staged (mr_slot), so that a wait for these loads does not
// also wait for the loads of the cells behind them.
template <int K, int NE>
__device__ __forceinline__ void mr_loads(const MrArgs& a, const MrGroup& g, int64_t gp, int l, float4& o0, float4& o1,
                                          uint32_t& raw, float& s) {
    using C = MrC<K, NE>;
    // no cell (gp < 0): the group's first cell is loaded instead and never
    // multiplied, so that every step issues the same loads (the waits for a
    // cell's loads then count exactly the loads issued after them)
    const int64_t cg = gp < 0 ? 0 : gp - g.gpos0;
    const float4* src = reinterpret_cast<const float4*>(a.op + g.gop) + cg * C::NV4;
    o0 = src[l < C::NV4 ? l : C::NV4 - 1];
    o1 = C::NV4 > 64 ? src[l + 64 < C::NV4 ? l + 64 : C::NV4 - 1] : float4{0.0f, 0.0f, 0.0f, 0.0f};
    s = a.scale[g.gpos0 + cg];
    const int le = l < C::NED ? l : 0;
    raw = a.idx[g.i_off + ((cg >> 6) * NE + le % NE) * 64 + (cg & 63)];
}

// The row slot of lane l (< M) of the cell at store position gp from the raw
// edge index of mr_loads.
template <int K, int NE>
__device__ __forceinline__ uint32_t mr_slot(const MrArgs& a, int64_t gp, int l, uint32_t raw) {
    using C = MrC<K, NE>;
    if (l < C::NED) return (raw & SIDE1) | ((raw & 0x7fffffffu) + (uint32_t)(l / NE));
    if (l < C::N) return (uint32_t)(a.cbase + (l - C::NED) * a.cpad + gp);
    if (l < C::M) return (uint32_t)(a.cbase + (C::NI + l - C::N) * a.cpad + gp);
    return 0;
}

// The cells of one group at store positions [p0, p1) (all real), pipelined as
// in k_apply_mr: while cell i is multiplied the rows of cell i + 1 are
// gathered and the operator of cell i + 2 is loaded.  U: Xj is the same
// buffer for the whole wave and every row lies within 4 GiB of it, so a
// gather is a wave-uniform base plus a 32-bit lane offset (no 64-bit address
// arithmetic and no 64-bit address registers per gather).
template <int K, int NE, bool U>
__device__ __forceinline__ void mr_run(const MrArgs& a, const MrGroup& g, int64_t p0, int64_t p1, int l,
                                        const float* __restrict__ Xj, float (*lop)[OPMAX], uint32_t (*lsl)[32],
                                        float* lkx, double& dot) {
    using C = MrC<K, NE>;
    const int hi = l >> 5, j = l & 31;
    float* const Y0 = a.Y;   // (the per-row choice between them then selects a value,
    float* const Y1 = a.Y1;  // not the address of an argument)
    // this lane's entry of the cell block at step kk: staged operator offset,
    // and whether it belongs to s A (else -W or zero)
    // two 16-bit offsets per register (the operator has at most 348 entries)
    uint32_t off2[(C::KS + 1) / 2] = {};
    uint32_t isA = 0;
#pragma unroll
    for (int kk = 0; kk < C::KS; ++kk) {
        const int c = 2 * kk + hi;
        int o = C::ZERO;
        if (j < C::M && c < C::M) {
            if (j < C::N && c < C::N) {
                const int r = j > c ? j : c, q = j > c ? c : j;
                o = r * (r + 1) / 2 + q;
                isA |= 1u << kk;
            } else if (j < C::N) {
                o = C::NA + (c - C::N) * C::N + j;
            } else if (c < C::N) {
                o = C::NA + (j - C::N) * C::N + c;
            }
        }
        off2[kk >> 1] |= (uint32_t)o << (16 * (kk & 1));
    }
    auto stage = [&](const float4& o0, const float4& o1, uint32_t raw, int64_t gp, int b, float (&B)[C::KS])
        __attribute__((always_inline)) {
        float4* d = reinterpret_cast<float4*>(lop[b]);
        if (l < C::NV4) d[l] = o0;
        if (C::NV4 > 64 && l + 64 < C::NV4) d[l + 64] = o1;
        if (l == 0) lop[b][C::ZERO] = 0.0f;
        if (l < C::M) lsl[b][l] = mr_slot<K, NE>(a, gp < 0 ? (int64_t)g.gpos0 : gp, l, raw);
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        __builtin_amdgcn_wave_barrier();
        // every lane gathers: a row the block does not have (the odd row
        // count's last step, upper half-wave) reads row 0 instead, which
        // meets only zero entries of the staged block (x is finite)
#pragma unroll
        for (int kk = 0; kk < C::KS; ++kk) {
            const int c = 2 * kk + hi;
            const uint32_t sl = lsl[b][c < C::M ? c : 0];
            if constexpr (U)
                B[kk] = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(Xj) +
                                                        (((sl & 0x7fffffffu) << 7) | (uint32_t)((l & 31) << 2)));
            else
                B[kk] = Xj[(int64_t)(sl & 0x7fffffffu) * NR + (l & 31)];
        }
    };
    auto multiply = [&](int b, float s, const float (&B)[C::KS]) __attribute__((always_inline)) {
        f16v acc = {};
#pragma unroll
        for (int kk = 0; kk < C::KS; ++kk) {
            const float v = lop[b][(off2[kk >> 1] >> (16 * (kk & 1))) & 0xffffu];
            const float av = (isA >> kk) & 1u ? s * v : -v;
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(B[kk], av, acc, 0, 0, 0);
        }
        // K x goes through LDS as a row-major tile (row j: its 32 columns).
        // Read back eight lanes to a row, each store instruction then writes
        // eight whole 128-byte rows (stored straight from the MFMA layout,
        // one instruction would touch 32 rows with 32 bytes each); and for
        // x.Kx the lanes that hold x read K x from the tile.
        if (j < C::M) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                *reinterpret_cast<float4*>(&lkx[j * MR_LDX + 8 * q + 4 * hi]) =
                    float4{acc[4 * q], acc[4 * q + 1], acc[4 * q + 2], acc[4 * q + 3]};
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int t = 0; t < (C::M + 7) / 8; ++t) {
            const int r = 8 * t + (l >> 3);
            if (r < C::M) {
                const uint32_t sl = lsl[b][r];
                float* __restrict__ dst = (sl & SIDE1) ? Y1 : Y0;
                *reinterpret_cast<float4*>(dst + (int64_t)(sl & 0x7fffffffu) * NR + 4 * (l & 7)) =
                    *reinterpret_cast<const float4*>(&lkx[r * MR_LDX + 4 * (l & 7)]);
            }
        }
        if (a.pw) {
#pragma unroll
            for (int kk = 0; kk < C::KS; ++kk) {
                const int c = 2 * kk + hi;
                if (c < C::M) dot += (double)B[kk] * (double)lkx[c * MR_LDX + (l & 31)];
            }
        }
        __builtin_amdgcn_wave_barrier();
    };
    // Three cells in flight: while cell c is multiplied, the rows of c + 1 and
    // c + 2 are being gathered and the operator of c + 3 is being loaded.
    // Cell j's registers, LDS buffer and rows are set j % 3; the loop body is
    // written three times with the roles rotated, so register sets are never
    // copied (a copy would wait for the loads still filling them).
    struct Set {
        float4 o0, o1;
        uint32_t sl;  // raw edge index (edge-row lanes)
        float s;
        int64_t gp;
    };
    Set S0, S1, S2;
    float B0[C::KS], B1[C::KS], B2[C::KS];
    auto load = [&](Set& t, int64_t gp) __attribute__((always_inline)) {
        t.gp = gp;
        mr_loads<K, NE>(a, g, gp, l, t.o0, t.o1, t.sl, t.s);
    };
    auto nxt = [&](int64_t gp) { return gp >= 0 && gp + 1 < p1 ? gp + 1 : (int64_t)-1; };
    load(S0, p0);
    load(S1, nxt(p0));
    stage(S0.o0, S0.o1, S0.sl, S0.gp, 0, B0);
    load(S2, nxt(S1.gp));
    stage(S1.o0, S1.o1, S1.sl, S1.gp, 1, B1);  // (no cell: staged and never multiplied)
    // step: multiply cell c (set c, buffer bc), stage cell n2 (two behind it)
    // into its buffer with its rows gathered, and load the cell after n2 into
    // c's registers
    auto step = [&](Set& c, Set& n2, float (&Bc)[C::KS], float (&Bn2)[C::KS], int bc, int bn2)
        __attribute__((always_inline)) {
        const float sc = c.s;
        load(c, nxt(n2.gp));
        stage(n2.o0, n2.o1, n2.sl, n2.gp, bn2, Bn2);
        multiply(bc, sc, Bc);
        __builtin_amdgcn_wave_barrier();  // buffer bc is free for the cell loaded now
    };
    while (S0.gp >= 0) {
        step(S0, S2, B0, B2, 0, 2);
        if (S1.gp < 0) break;
        step(S1, S0, B1, B0, 1, 0);
        if (S2.gp < 0) break;
        step(S2, S1, B2, B1, 2, 1);
    }
}

// One launch per valence group: the waves of the launch share the group's
// cells in contiguous runs; wave0 is the launch's first wave in the x.Kx
// partials (pw).  A single valence body per kernel keeps the register budget
// to that body's needs (about 85), so more waves fit on a SIMD and more row
// gathers are in flight.
template <int K, int NE>
__global__ __launch_bounds__(256, MR_WAVES) void k_apply_mr(MrArgs a, MrGroup g, int64_t cpw, int wave0, int zero_trash) {
    if (a.all_done && *a.all_done) return;
    __shared__ __attribute__((aligned(16))) float lop[4][3][OPMAX];
    __shared__ uint32_t lsl[4][3][32];
    __shared__ __attribute__((aligned(16))) float lkx[4][32 * MR_LDX];
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int wave = blockIdx.x * 4 + w;
    const int l = threadIdx.x & 63;
    const int col = l & 31;
    // Each column reads the buffer of its iteration parity.  The columns still
    // iterating share one parity (a refinement pass starts every column at
    // iteration 0 and each iteration advances all of them); a finished
    // column's products are never read, so it may read that buffer too.
    int par = 0;
    bool mixed = false;
    if (a.st) {
        const MState& sc = a.st[col];
        const bool act = !sc.done;
        const uint64_t odd = __ballot(act && (sc.itn & 1)), even = __ballot(act && !(sc.itn & 1));
        par = odd ? 1 : 0;
        mixed = odd && even;
    }
    // rows a cell does not have are not stored; the padding row reads zero
    if (zero_trash && wave == 0 && l < NR) a.Y[a.trash * NR + l] = 0.0f;
    double dot = 0.0;
    const int64_t end = (int64_t)g.gpos0 + g.count;
    const int64_t p0 = (int64_t)g.gpos0 + (int64_t)wave * cpw;
    const int64_t p1 = p0 + cpw < end ? p0 + cpw : end;
    if (p0 < p1) {
        if (a.off32 && !mixed)
            mr_run<K, NE, true>(a, g, p0, p1, l, a.X[__builtin_amdgcn_readfirstlane(par)], lop[w], lsl[w], lkx[w], dot);
        else
            mr_run<K, NE, false>(a, g, p0, p1, l, a.st ? a.X[a.st[col].itn & 1] : a.X[0], lop[w], lsl[w], lkx[w],
                                 dot);
    }
    if (a.pw) {
        dot += __shfl_xor(dot, 32, 64);
        if (l < 32) a.pw[(int64_t)(wave0 + wave) * NR + col] = dot;
    }
}

constexpr int MR_MIN_RUN = 1;  // fewest cells a wave of k_apply_mr runs through

template <int K>
static void mr_launch(int ne, dim3 grid, hipStream_t st, const MrArgs& a, const MrGroup& g, int64_t cpw, int wave0,
                       int zero_trash) {
    switch (ne) {
#define MR_CASE(NE_)                                                                                    \
    case NE_:                                                                                            \
        hipLaunchKernelGGL((k_apply_mr<K, NE_>), grid, dim3(256), 0, st, a, g, cpw, wave0, zero_trash); \
        break;
        MR_CASE(3)
        MR_CASE(4)
        MR_CASE(5)
        MR_CASE(6)
        MR_CASE(7)
        MR_CASE(8)
        MR_CASE(9)
        MR_CASE(10)
#undef MR_CASE
        default:
            throw Error(VEMD_ERR_MESH, "the multi-RHS apply handles polygons of 3 to 10 edges");
    }
}

template <int K>
static int mr_occupancy_k(int ne) {
    int bpc = 0;
    switch (ne) {
#define MR_OCC(NE_)                                                                         \
    case NE_:                                                                                \
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_apply_mr<K, NE_>, 256, 0)); \
        break;
        MR_OCC(3)
        MR_OCC(4)
        MR_OCC(5)
        MR_OCC(6)
        MR_OCC(7)
        MR_OCC(8)
        MR_OCC(9)
        MR_OCC(10)
#undef MR_OCC
        default:
            throw Error(VEMD_ERR_MESH, "the multi-RHS apply handles polygons of 3 to 10 edges");
    }
    return bpc;
}

static int mr_occupancy(int k, int ne) { return k == 0 ? mr_occupancy_k<0>(ne) : mr_occupancy_k<1>(ne); }

// ---------------------------------------------------------------------------
// Simultaneous MINRES over NR interleaved columns (one MState per column)
// ---------------------------------------------------------------------------

struct MrUpd {
    float* R[2];
    float* Z[2];
    float* WB[2];
    float* x;
    const float* Az;
    const float* Az1;
    const float* minv;
    int64_t edof;
    const VecChunk* chunks;
    double* p1;  // [chunk][NR]
    double* p2;
    MState* st;
    int amg;
    int64_t pb, cpad;
    int q;
    float* bp;   // multigrid input, [cell-major pressure row][NR]
    float* bd0 = nullptr;        // non-null: also the cycle's first update bd0 = c0 * bsc * bp
    const float* bsc = nullptr;  // per pressure row
    float c0 = 0.0f;
};

// Column partials of a block of 8 row groups x NR columns.
__device__ __forceinline__ void col_pair(double d1, double d2, double* o1, double* o2) {
    __shared__ double s1[8][NR], s2[8][NR];
    const int j = threadIdx.x & (NR - 1), r0 = threadIdx.x >> 5;
    s1[r0][j] = d1;
    s2[r0][j] = d2;
    __syncthreads();
    if (r0 == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < 8; ++k) {
            a += s1[k][j];
            b += s2[k][j];
        }
        o1[j] = a;
        if (o2) o2[j] = b;
    }
}

__global__ __launch_bounds__(256) void k_update_mr(MrUpd a) {
    const VecChunk ch = a.chunks[blockIdx.x];
    const int j = threadIdx.x & (NR - 1), r0 = threadIdx.x >> 5;
    const MState* st = a.st + j;
    double d1 = 0.0, d2 = 0.0;
    if (!st->done) {
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
        const bool edge = ch.start < a.edof;  // a chunk lies wholly inside or outside the edge dofs
        const int pr = ch.pslot - 1;
        const int64_t end = ch.start + ch.len;
        float* __restrict__ x = a.x;
        const float* __restrict__ Az = a.Az;
        const float* __restrict__ Az1 = a.Az1;
        // UR rows per thread and trip, every load issued before the first use
        // (rows past the chunk read its first row and store nothing)
        constexpr int UR = 4;
        for (int64_t row0 = ch.start + r0; row0 < end; row0 += 8 * UR) {
            int64_t e[UR];
            bool ok[UR];
#pragma unroll
            for (int u = 0; u < UR; ++u) {
                const int64_t row = row0 + 8 * u;
                ok[u] = row < end;
                e[u] = (ok[u] ? row : row0) * NR + j;
            }
            float r2v[UR], azv[UR], r1v[UR], zv[UR], wnv[UR], wpv[UR], xv[UR], mv[UR];
#pragma unroll
            for (int u = 0; u < UR; ++u) {
                r2v[u] = r2[e[u]];
                azv[u] = Az[e[u]];
                r1v[u] = r1[e[u]];
            }
            if (edge) {
#pragma unroll
                for (int u = 0; u < UR; ++u) azv[u] += Az1[e[u]];
            }
            if (pend) {
#pragma unroll
                for (int u = 0; u < UR; ++u) {
                    zv[u] = zo[e[u]];
                    wnv[u] = wn[e[u]];
                    wpv[u] = wp[e[u]];
                    xv[u] = x[e[u]];
                }
            }
            if (!cyc) {
#pragma unroll
                for (int u = 0; u < UR; ++u) mv[u] = a.minv[e[u] / NR];
            }
#pragma unroll
            for (int u = 0; u < UR; ++u) {
                if (!ok[u]) continue;
                const float y = caz * azv[u] - cr1 * r1v[u] - cr2 * r2v[u];
                if (pend) {
                    const float w = uv * zv[u] - uw1 * wnv[u] - uw2 * wpv[u];
                    wn[e[u]] = w;
                    x[e[u]] = xv[u] + ux * w;
                }
                r1[e[u]] = y;
                if (cyc) {
                    const int64_t row = row0 + 8 * u;
                    const int64_t prow = (row - a.pb - (int64_t)pr * a.cpad) * a.q + pr;
                    a.bp[prow * NR + j] = y;
                    if (a.bd0) a.bd0[prow * NR + j] = a.c0 * a.bsc[prow] * y;
                } else {
                    const float zn = mv[u] * y;
                    zo[e[u]] = zn;
                    d1 += (double)y * (double)zn;
                    d2 += (double)zn * (double)r2v[u];
                }
            }
        }
    }
    col_pair(d1, d2, a.p1 + (int64_t)blockIdx.x * NR, a.p2 + (int64_t)blockIdx.x * NR);
}

__global__ __launch_bounds__(256) void k_pdot_mr(MrUpd a, const int* pchunks, const float* xp, double* q1,
                                               double* q2, int init) {
    const int c = pchunks[blockIdx.x];
    const VecChunk ch = a.chunks[c];
    const int j = threadIdx.x & (NR - 1), r0 = threadIdx.x >> 5;
    const MState* st = a.st + j;
    double d1 = 0.0, d2 = 0.0;
    if (init || !st->done) {
        const int it = init ? 0 : st->itn;
        float* __restrict__ zo = init ? a.Z[0] : a.Z[(it + 1) & 1];
        const float* __restrict__ r2 = a.R[it & 1];
        const int pr = ch.pslot - 1;
        const int64_t end = ch.start + ch.len;
        const float* __restrict__ bp = a.bp;
        // UR rows per thread and trip, every load issued before the first use
        constexpr int UR = 4;
        for (int64_t row0 = ch.start + r0; row0 < end; row0 += 8 * UR) {
            int64_t e[UR], k[UR];
            bool ok[UR];
#pragma unroll
            for (int u = 0; u < UR; ++u) {
                const int64_t row0u = row0 + 8 * u;
                ok[u] = row0u < end;
                const int64_t row = ok[u] ? row0u : row0;
                e[u] = row * NR + j;
                k[u] = ((row - a.pb - (int64_t)pr * a.cpad) * a.q + pr) * NR + j;
            }
            float zv[UR], bv[UR], rv[UR];
#pragma unroll
            for (int u = 0; u < UR; ++u) {
                zv[u] = xp[k[u]];
                bv[u] = bp[k[u]];
                rv[u] = init ? 0.0f : r2[e[u]];
            }
#pragma unroll
            for (int u = 0; u < UR; ++u) {
                if (!ok[u]) continue;
                zo[e[u]] = zv[u];
                d1 += (double)bv[u] * (double)zv[u];
                if (!init) d2 += (double)zv[u] * (double)rv[u];
            }
        }
    }
    col_pair(d1, d2, q1 + (int64_t)c * NR, q2 + (int64_t)c * NR);
}

// Column sums of partials p[c][j] over c in [0, n) for the block's column j.
__device__ double col_sum(const double* p, int n, int j) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int c = threadIdx.x; c < n; c += 256) s += p[(int64_t)c * NR + j];
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

__global__ void k_finA_mr(MState* sts, const double* pw, int nwaves) {
    MState* st = sts + blockIdx.x;
    if (st->done) return;
    double zaz = col_sum(pw, nwaves, blockIdx.x);
    if (threadIdx.x == 0) {
        double beta = st->beta;
        double alfa = zaz / (beta * beta) - (st->itn > 0 ? st->rho / st->oldb : 0.0);
        st->alfa = alfa;
        st->c_az = 1.0 / beta;
        st->c_r1 = st->itn > 0 ? beta / st->oldb : 0.0;
        st->c_r2 = alfa / beta;
    }
}

__global__ void k_finB_mr(MState* sts, const double* p1, const double* p2, const double* q1, const double* q2,
                          int nchunks, int useq) {
    MState* st = sts + blockIdx.x;
    if (st->done) return;
    const int j = blockIdx.x;
    double b2 = col_sum(p1, nchunks, j) + (useq ? col_sum(q1, nchunks, j) : 0.0);
    double rho = col_sum(p2, nchunks, j) + (useq ? col_sum(q2, nchunks, j) : 0.0);
    if (threadIdx.x == 0) {
        double beta_new = std::sqrt(b2 > 0.0 ? b2 : 0.0);
        st->rho = rho;
        st->oldb = st->beta;
        st->beta = beta_new;
        double oldeps = st->epsln;
        double delta = st->cs * st->dbar + st->sn * st->alfa;
        double gbar = st->sn * st->dbar - st->cs * st->alfa;
        st->epsln = st->sn * beta_new;
        st->dbar = -st->cs * beta_new;
        double gamma = std::sqrt(gbar * gbar + beta_new * beta_new);
        if (gamma < 1e-300) gamma = 1e-300;
        st->cs = gbar / gamma;
        st->sn = beta_new / gamma;
        double phi = st->cs * st->phibar;
        st->phibar = st->sn * st->phibar;
        st->u_v = 1.0 / (st->oldb * gamma);
        st->u_w1 = oldeps / gamma;
        st->u_w2 = delta / gamma;
        st->u_x = phi;
        st->pending = 1;
        st->itn += 1;
        if (st->phibar <= st->tol || st->itn >= st->maxit || !(beta_new > 0.0)) st->done = 1;
    }
}

__global__ void k_alldone_mr(const MState* sts, int* all_done) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (threadIdx.x < NR && !sts[threadIdx.x].done) any = 1;
    __syncthreads();
    if (threadIdx.x == 0) *all_done = any ? 0 : 1;
}

__global__ __launch_bounds__(256) void k_init_mr(MrUpd a, const float* b) {
    const VecChunk ch = a.chunks[blockIdx.x];
    const int j = threadIdx.x & (NR - 1), r0 = threadIdx.x >> 5;
    const bool cyc = a.amg && ch.pslot;
    const int pr = ch.pslot - 1;
    double d = 0.0;
    for (int64_t row = ch.start + r0; row < ch.start + ch.len; row += 8) {
        const int64_t e = row * NR + j;
        float bi = b[e];
        a.R[0][e] = bi;
        a.R[1][e] = 0.0f;
        a.Z[1][e] = 0.0f;
        a.WB[0][e] = 0.0f;
        a.WB[1][e] = 0.0f;
        a.x[e] = 0.0f;
        if (cyc) {
            a.bp[((row - a.pb - (int64_t)pr * a.cpad) * a.q + pr) * NR + j] = bi;
        } else {
            float z = a.minv[row] * bi;
            a.Z[0][e] = z;
            d += (double)bi * (double)z;
        }
    }
    col_pair(d, 0.0, a.p1 + (int64_t)blockIdx.x * NR, nullptr);
}

__global__ void k_finInit_mr(MState* sts, const double* p1, const double* q1, int nchunks, int useq,
                             const double* rtol, int maxit) {
    const int j = blockIdx.x;
    double b2 = col_sum(p1, nchunks, j) + (useq ? col_sum(q1, nchunks, j) : 0.0);
    if (threadIdx.x == 0) {
        double beta1 = std::sqrt(b2 > 0.0 ? b2 : 0.0);
        MState s = {};
        s.beta1 = beta1;
        s.beta = beta1;
        s.phibar = beta1;
        s.cs = -1.0;
        s.tol = rtol[j] * beta1;
        s.maxit = maxit;
        s.done = (beta1 > 0.0 && maxit > 0 && rtol[j] < 1.0) ? 0 : 1;
        sts[j] = s;
    }
}

__global__ __launch_bounds__(256) void k_flush_mr(MrUpd a) {
    const VecChunk ch = a.chunks[blockIdx.x];
    const int j = threadIdx.x & (NR - 1), r0 = threadIdx.x >> 5;
    const MState* st = a.st + j;
    if (!st->pending) return;
    const int m = st->itn - 1;
    const float* z = a.Z[m & 1];
    const float* wpp = a.WB[m & 1];
    const float* wp = a.WB[(m + 1) & 1];
    const float uv = (float)st->u_v, uw1 = (float)st->u_w1, uw2 = (float)st->u_w2, ux = (float)st->u_x;
    for (int64_t row = ch.start + r0; row < ch.start + ch.len; row += 8) {
        const int64_t e = row * NR + j;
        float w = uv * z[e] - uw1 * wpp[e] - uw2 * wp[e];
        a.x[e] += ux * w;
    }
}

// FP64 outer loop: column-major [NR][n] <-> interleaved [n][NR].
__global__ __launch_bounds__(256) void k_cm_to_il(const double* r, float* b, int64_t n, int ncol) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n * NR) return;
    const int64_t row = t / NR;
    const int j = (int)(t % NR);
    b[t] = j < ncol ? (float)r[(int64_t)j * n + row] : 0.0f;
}

__global__ __launch_bounds__(256) void k_il_add_cm(double* x, const float* d, int64_t n, int ncol) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n * NR) return;
    const int64_t row = t / NR;
    const int j = (int)(t % NR);
    if (j < ncol) x[(int64_t)j * n + row] += (double)d[t];
}

// r = b - y (y's edge dofs: both sides) and block partials of |r|^2.
__global__ __launch_bounds__(256) void k_resid_col(const double* b, const double* y, int64_t n, double* r, double* part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        double v = y ? b[i] - y[i] : b[i];  // y null: the residual of x = 0
        r[i] = v;
        s += v * v;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_sum_to(const double* part, int n, double* out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static void multi_build(Solver& s, MultiState& m);

// The multi-right-hand-side state is built aside and published in d.multi only
// once every allocation and check has succeeded: a failure on the way (device
// out of memory, too many valence groups) frees what was allocated and leaves
// d.multi null, so a later call reports the error again or builds afresh
// instead of launching kernels on a half-built state.
static MultiState& multi_setup(Solver& s) {
    DeviceState& d = *s.dev;
    if (d.multi) return *d.multi;
    if (s.layout.lam || s.layout.ninst != 1)
        throw Error(VEMD_ERR_ARG, "multi-right-hand-side solves need one Dirichlet-pressure mesh");
    if (s.layout.k > 1) throw Error(VEMD_ERR_ARG, "multi-right-hand-side solves support k = 0 and k = 1");
    MultiState* m = new MultiState();
    try {
        multi_build(s, *m);
    } catch (...) {
        multi_release(m);
        throw;
    }
    d.multi = m;
    return *m;
}

static void multi_build(Solver& s, MultiState& m) {
    DeviceState& d = *s.dev;
    const Layout& L = s.layout;
    const Setup& S = *s.setup;
    // FP32 cell blocks: packed A then W, per group in store order
    std::vector<int64_t> gop(L.groups.size());
    int64_t tot = 0;
    for (size_t gi = 0; gi < L.groups.size(); ++gi) {
        const GroupDesc& g = L.groups[gi];
        const int n = g.n, sz = n * (n + 1) / 2 + L.q * n, sz4 = (sz + 3) / 4 * 4;
        gop[gi] = tot;
        tot += (int64_t)sz4 * g.npad;
    }
    std::vector<float> op(tot, 0.0f);
    parallel_for((int64_t)L.groups.size(), [&](int64_t b, int64_t e) {
        for (int64_t gi = b; gi < e; ++gi) {
            const GroupDesc& g = L.groups[gi];
            const int n = g.n, na = n * (n + 1) / 2, sz = na + L.q * n, sz4 = (sz + 3) / 4 * 4;
            for (int cg = 0; cg < g.npad; ++cg) {
                const int c = L.cell_of_gpos[g.gpos0 + cg];
                if (c < 0) continue;
                float* o = &op[gop[gi] + (int64_t)cg * sz4];
                const double* A = &S.A[S.ofs.a_off[c]];
                const double* W = &S.W[S.ofs.w_off[c]];
                for (int t = 0; t < na; ++t) o[t] = (float)A[t];
                for (int t = 0; t < L.q * n; ++t) o[na + t] = (float)W[t];
            }
        }
    });
    m.op = (float*)malloc_m(m, op.size() * sizeof(float));
    HIPCHK(hipMemcpy(m.op, op.data(), op.size() * sizeof(float), hipMemcpyHostToDevice));
    // the group table lives in LDS: at most MR_MAXG groups (VEMD_MR_MAX_GROUPS
    // lowers the cap; tests use it to reach this failure after allocations)
    int max_groups = MR_MAXG;
    if (const char* e = std::getenv("VEMD_MR_MAX_GROUPS")) max_groups = std::min(max_groups, std::max(0, std::atoi(e)));
    if ((int)L.groups.size() > max_groups)
        throw Error(VEMD_ERR_ARG, "too many valence groups for the multi-RHS apply: " + std::to_string(L.groups.size()) +
                                      " > " + std::to_string(max_groups));
    std::vector<MrGroup> mg(L.groups.size());
    for (size_t gi = 0; gi < L.groups.size(); ++gi) {
        const GroupDesc& g = L.groups[gi];
        int count = g.npad;
        while (count > 0 && L.cell_of_gpos[g.gpos0 + count - 1] < 0) --count;
        for (int cg = 0; cg < count; ++cg)
            if (L.cell_of_gpos[g.gpos0 + cg] < 0)
                throw Error(VEMD_ERR_INTERNAL, "padding inside a valence group");
        mg[gi] = MrGroup{g.gpos0, count, g.ne, g.n, gop[gi], g.i_off};
    }
    m.ngroups = (int)mg.size();
    m.mg = (MrGroup*)malloc_m(m, mg.size() * sizeof(MrGroup));
    HIPCHK(hipMemcpy(m.mg, mg.data(), mg.size() * sizeof(MrGroup), hipMemcpyHostToDevice));
    // one launch per valence group: as many waves as fit on the chip at once
    // (the occupancy of that valence's kernel), each over a contiguous run of
    // the group's cells
    int dev = 0, ncu = 256;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    int wave0 = 0;
    m.mg_host = mg;
    for (const MrGroup& g : mg) {
        int occ = mr_occupancy(s.layout.k, g.ne);
        if (const char* e = std::getenv("VEMD_MR_BLOCKS_PER_CU")) occ = std::max(1, std::atoi(e));
        const int64_t fit = (int64_t)ncu * std::max(occ, 1) * 4;
        const int64_t cpw = std::max<int64_t>(MR_MIN_RUN, (g.count + fit - 1) / fit);
        const int64_t waves = (g.count + cpw - 1) / cpw;
        const int blocks = (int)std::max<int64_t>(1, (waves + 3) / 4);
        m.mr_cpw.push_back(cpw);
        m.mr_blocks.push_back(blocks);
        m.mr_wave0.push_back(wave0);
        wave0 += 4 * blocks;
    }
    m.pw_waves = std::max(wave0, 1);
    const size_t fv = (size_t)L.nvec * NR * sizeof(float);
    for (int i = 0; i < 2; ++i) {
        m.R[i] = (float*)malloc_m(m, fv);
        m.Z[i] = (float*)malloc_m(m, fv);
        m.WB[i] = (float*)malloc_m(m, fv);
    }
    m.x = (float*)malloc_m(m, fv);
    m.Az = (float*)malloc_m(m, fv);
    m.b = (float*)malloc_m(m, fv);
    m.Az1 = (float*)malloc_m(m, (size_t)std::max<int64_t>(L.cbase, 1) * NR * sizeof(float));
    HIPCHK(hipMemset(m.Az1, 0, (size_t)std::max<int64_t>(L.cbase, 1) * NR * sizeof(float)));
    for (float* p : {m.R[0], m.R[1], m.Z[0], m.Z[1], m.WB[0], m.WB[1], m.x, m.Az, m.b}) HIPCHK(hipMemset(p, 0, fv));
    const size_t dv = (size_t)L.nvec * NR * sizeof(double);
    m.x64 = (double*)malloc_m(m, dv);
    m.r64 = (double*)malloc_m(m, dv);
    m.y64 = (double*)malloc_m(m, (size_t)L.nvec * sizeof(double));
    m.st = (MState*)malloc_m(m, NR * sizeof(MState));
    HIPCHK(hipMemset(m.st, 0, NR * sizeof(MState)));
    HIPCHK(hipHostMalloc((void**)&m.st_host, NR * sizeof(MState), hipHostMallocDefault));
    m.rtol = (double*)malloc_m(m, NR * sizeof(double));
    HIPCHK(hipHostMalloc((void**)&m.rtol_host, NR * sizeof(double), hipHostMallocDefault));
    const int nch = (int)L.chunks.size();
    m.pw = (double*)malloc_m(m, (size_t)m.pw_waves * NR * sizeof(double));
    for (double** p : {&m.p1, &m.p2, &m.q1, &m.q2}) {
        *p = (double*)malloc_m(m, (size_t)nch * NR * sizeof(double));
        HIPCHK(hipMemset(*p, 0, (size_t)nch * NR * sizeof(double)));
    }
    m.part = (double*)malloc_m(m, (size_t)d.nblk * sizeof(double));
    m.dsum = (double*)malloc_m(m, NR * sizeof(double));
    HIPCHK(hipHostMalloc((void**)&m.hsum, NR * sizeof(double), hipHostMallocDefault));
    m.all_done = (int*)malloc_m(m, sizeof(int));
    HIPCHK(hipMemset(m.all_done, 0, sizeof(int)));
    if (d.amg32) amg_multi_alloc(*d.amg32);
}

static MrArgs mr_args(Solver& s, MultiState& m, const float* x0, const float* x1, float* y, float* y1, bool solve) {
    DeviceState& d = *s.dev;
    MrArgs a;
    a.mg = m.mg;
    a.ngroups = m.ngroups;
    a.op = m.op;
    a.idx = d.idx;
    a.scale = d.o32.scale;
    a.X[0] = x0;
    a.X[1] = x1;
    a.Y = y;
    a.Y1 = y1;
    a.cbase = s.layout.cbase;
    a.trash = s.layout.trash_idx;
    a.cpad = s.layout.cpad;
    a.pw = solve ? m.pw : nullptr;
    a.st = solve ? m.st : nullptr;
    a.all_done = solve ? m.all_done : nullptr;
    a.off32 = (uint64_t)s.layout.nvec * NR * sizeof(float) <= (uint64_t(1) << 32) ? 1 : 0;
    return a;
}

static void launch_apply_mr(Solver& s, MultiState& m, const MrArgs& a, hipStream_t st) {
    int first = 1;  // the first launch zeroes the padding row
    for (size_t gi = 0; gi < m.mg_host.size(); ++gi) {
        const MrGroup& g = m.mg_host[gi];
        if (g.count == 0) continue;
        if (s.layout.k == 0)
            mr_launch<0>(g.ne, dim3(m.mr_blocks[gi]), st, a, g, m.mr_cpw[gi], m.mr_wave0[gi], first);
        else
            mr_launch<1>(g.ne, dim3(m.mr_blocks[gi]), st, a, g, m.mr_cpw[gi], m.mr_wave0[gi], first);
        first = 0;
    }
}

void dev_apply_multi(Solver& s, const void* x, void* y, void* y1, void* stream) {
    if (!s.dev) throw Error(VEMD_ERR_STATE, "solver not uploaded");
    HIPCHK(hipSetDevice(s.dev->device));
    MultiState& m = multi_setup(s);
    ScratchUse use(*s.dev, (hipStream_t)stream);
    MrArgs a = mr_args(s, m, (const float*)x, (const float*)x, (float*)y, (float*)y1, false);
    launch_apply_mr(s, m, a, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
}

static MrUpd mr_upd(Solver& s, MultiState& m, bool amg) {
    DeviceState& d = *s.dev;
    const Layout& L = s.layout;
    MrUpd u{{m.R[0], m.R[1]}, {m.Z[0], m.Z[1]}, {m.WB[0], m.WB[1]}, m.x, m.Az, m.Az1, d.o32.minv, L.cbase,
            d.chunks,         m.p1,           m.p2,           m.st, amg ? 1 : 0, 0,      L.cpad,        L.q,
            nullptr};
    if (amg) {
        const AmgDev<float>& M = *d.amg32;
        u.pb = L.cbase + (int64_t)L.ni * L.cpad;
        u.bp = M.lv[0].mb;
        if (amg_pre_fusable(M)) {  // the cycle's first update formed by the Krylov sweep
            u.bd0 = M.lv[0].md0;
            u.bsc = M.lv[0].sc ? M.lv[0].sc : M.lv[0].dinv;
            u.c0 = (float)(1.0 / M.lv[0].centre);
        }
    }
    return u;
}

static void enqueue_iteration_mr(Solver& s, MultiState& m, bool amg) {
    DeviceState& d = *s.dev;
    MrArgs a = mr_args(s, m, m.Z[0], m.Z[1], m.Az, m.Az1, true);
    launch_apply_mr(s, m, a, d.stream);
    hipLaunchKernelGGL(k_finA_mr, dim3(NR), dim3(256), 0, d.stream, m.st, (const double*)m.pw, m.pw_waves);
    MrUpd u = mr_upd(s, m, amg);
    hipLaunchKernelGGL(k_update_mr, dim3(d.nchunks), dim3(256), 0, d.stream, u);
    if (amg) {
        amg_vcycle_multi(*d.amg32, d.stream, m.all_done, u.bd0 != nullptr);
        hipLaunchKernelGGL(k_pdot_mr, dim3(d.npchunks), dim3(256), 0, d.stream, u, (const int*)d.pchunks,
                           (const float*)d.amg32->lv[0].mx, m.q1, m.q2, 0);
    }
    hipLaunchKernelGGL(k_finB_mr, dim3(NR), dim3(256), 0, d.stream, m.st, (const double*)m.p1, (const double*)m.p2,
                       (const double*)m.q1, (const double*)m.q2, d.nchunks, amg ? 1 : 0);
    hipLaunchKernelGGL(k_alldone_mr, dim3(1), dim3(64), 0, d.stream, (const MState*)m.st, m.all_done);
}

// Simultaneous preconditioned MINRES (FP32) on the NR columns of m.b; column
// j stops at rtol[j] (>= 1: untouched).  Returns the largest iteration count.
static int minres_mr(Solver& s, MultiState& m, const std::vector<double>& rtol, int maxit, int chk, bool graph,
                     bool amg, std::vector<int>& its) {
    DeviceState& d = *s.dev;
    MrUpd u = mr_upd(s, m, amg);
    if (graph && (m.graph_iters != chk || m.graph_amg != amg || !m.graph)) {
        HIPCHK(hipStreamSynchronize(d.stream));
        if (m.graph) (void)hipGraphExecDestroy(m.graph);
        m.graph = nullptr;
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(d.stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < chk; ++i) enqueue_iteration_mr(s, m, amg);
        HIPCHK(hipStreamEndCapture(d.stream, &g));
        HIPCHK(hipGraphInstantiate(&m.graph, g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
        m.graph_iters = chk;
        m.graph_amg = amg;
    }
    HIPCHK(hipStreamSynchronize(d.stream));
    for (int j = 0; j < NR; ++j) m.rtol_host[j] = rtol[j];
    HIPCHK(hipMemcpyAsync(m.rtol, m.rtol_host, NR * sizeof(double), hipMemcpyHostToDevice, d.stream));
    hipLaunchKernelGGL(k_init_mr, dim3(d.nchunks), dim3(256), 0, d.stream, u, (const float*)m.b);
    if (amg) {
        amg_vcycle_multi(*d.amg32, d.stream, nullptr);
        hipLaunchKernelGGL(k_pdot_mr, dim3(d.npchunks), dim3(256), 0, d.stream, u, (const int*)d.pchunks,
                           (const float*)d.amg32->lv[0].mx, m.q1, m.q2, 1);
    }
    hipLaunchKernelGGL(k_finInit_mr, dim3(NR), dim3(256), 0, d.stream, m.st, (const double*)m.p1, (const double*)m.q1,
                       d.nchunks, amg ? 1 : 0, (const double*)m.rtol, maxit);
    hipLaunchKernelGGL(k_alldone_mr, dim3(1), dim3(64), 0, d.stream, (const MState*)m.st, m.all_done);
    HIPCHK(hipGetLastError());
    while (true) {
        HIPCHK(hipMemcpyAsync(m.st_host, m.st, NR * sizeof(MState), hipMemcpyDeviceToHost, d.stream));
        HIPCHK(hipStreamSynchronize(d.stream));
        bool all = true;
        for (int j = 0; j < NR && all; ++j) all = m.st_host[j].done != 0;
        if (all) break;
        if (graph)
            HIPCHK(hipGraphLaunch(m.graph, d.stream));
        else
            for (int i = 0; i < chk; ++i) enqueue_iteration_mr(s, m, amg);
    }
    hipLaunchKernelGGL(k_flush_mr, dim3(d.nchunks), dim3(256), 0, d.stream, u);
    HIPCHK(hipGetLastError());
    its.assign(NR, 0);
    int mx = 0;
    for (int j = 0; j < NR; ++j) {
        its[j] = m.st_host[j].itn;
        mx = std::max(mx, its[j]);
    }
    return mx;
}

// r_j = b_j - K x_j (FP64, column-major) for j < ncol; returns |r_j|.
// Columns of r64 = b - K x64 (x_zero: x64 = 0, r64 = b, no applies) and their 2-norms.
static std::vector<double> residual_mr(Solver& s, MultiState& m, const double* b, int ncol, bool x_zero = false) {
    DeviceState& d = *s.dev;
    const int64_t n = s.layout.nvec;
    for (int j = 0; j < ncol; ++j) {
        if (!x_zero) dev_apply_int(s, VEMD_FP64, m.x64 + (int64_t)j * n, m.y64, nullptr, d.stream);
        hipLaunchKernelGGL(k_resid_col, dim3(d.nblk), dim3(256), 0, d.stream, b + (int64_t)j * n,
                           x_zero ? (const double*)nullptr : (const double*)m.y64, n, m.r64 + (int64_t)j * n, m.part);
        hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, d.stream, (const double*)m.part, d.nblk, m.dsum + j);
    }
    HIPCHK(hipMemcpyAsync(m.hsum, m.dsum, ncol * sizeof(double), hipMemcpyDeviceToHost, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    std::vector<double> r(ncol);
    for (int j = 0; j < ncol; ++j) r[j] = std::sqrt(m.hsum[j]);
    return r;
}

void dev_solve_multi(Solver& s, const vemd_solve_opts& o, int ncol, const double* b, double* x,
                     vemd_solve_info* info) {
    if (!s.dev) throw Error(VEMD_ERR_STATE, "solver not uploaded");
    DeviceState& d = *s.dev;
    HIPCHK(hipSetDevice(d.device));
    if (ncol < 1 || ncol > NR) throw Error(VEMD_ERR_ARG, "between 1 and 32 right-hand sides");
    if (o.precision != VEMD_MIXED)
        throw Error(VEMD_ERR_ARG, "multi-right-hand-side solves run FP32 MFMA iterations inside FP64 refinement "
                                  "(VEMD_MIXED)");
    if (o.precond == VEMD_PC_MULTIGRID && !d.amg32)
        throw Error(VEMD_ERR_ARG, "no multigrid hierarchy on this solver");
    MultiState& m = multi_setup(s);
    if (d.amg32) amg_multi_alloc(*d.amg32);  // the hierarchy is rebuilt when theta changes
    const bool amg = o.precond != VEMD_PC_JACOBI && d.amg32 != nullptr;
    const int64_t n = s.layout.nvec;
    auto t0 = std::chrono::steady_clock::now();
    // b may still be being written on the caller's stream
    ScratchUse use(d, d.stream, true, (hipStream_t)o.stream);
    HIPCHK(hipMemsetAsync(m.x64, 0, (size_t)ncol * n * sizeof(double), d.stream));
    std::vector<double> bn = residual_mr(s, m, b, ncol, true), rel(ncol);
    for (int j = 0; j < ncol; ++j) rel[j] = bn[j] > 0 ? 1.0 : 0.0;
    std::vector<int> tot(ncol, 0), its;
    int iters = 0, refin = 0;
    const int maxit = o.max_iters > 0 ? o.max_iters : 1000000;
    const int chk = o.check_every > 0 ? o.check_every : 32;
    auto worst = [&]() { return *std::max_element(rel.begin(), rel.end()); };
    std::vector<double> target(NR, 1.0);
    while (worst() > o.rtol && iters < maxit && refin < 100) {
        for (int j = 0; j < ncol; ++j)
            target[j] = rel[j] > o.rtol ? std::max(o.inner_rtol > 0 ? o.inner_rtol : 1e-4, 0.5 * o.rtol / rel[j]) : 1.0;
        hipLaunchKernelGGL(k_cm_to_il, dim3((unsigned)((n * NR + 255) / 256)), dim3(256), 0, d.stream,
                           (const double*)m.r64, m.b, n, ncol);
        iters += minres_mr(s, m, target, maxit - iters, chk, o.use_graph != 0, amg, its);
        hipLaunchKernelGGL(k_il_add_cm, dim3((unsigned)((n * NR + 255) / 256)), dim3(256), 0, d.stream, m.x64,
                           (const float*)m.x, n, ncol);
        for (int j = 0; j < ncol; ++j) tot[j] += its[j];
        ++refin;
        std::vector<double> rn = residual_mr(s, m, b, ncol);
        for (int j = 0; j < ncol; ++j) rel[j] = bn[j] > 0 ? rn[j] / bn[j] : 0.0;
        static const bool verbose = std::getenv("VEMD_VERBOSE") != nullptr;
        if (verbose)
            std::fprintf(stderr, "[vemd] multi-RHS refinement %d: %d iterations (column 0), worst true relative "
                                 "residual %.3e\n", refin, its[0], worst());
    }
    if (x) HIPCHK(hipMemcpyAsync(x, m.x64, (size_t)ncol * n * sizeof(double), hipMemcpyDeviceToDevice, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (info)
        for (int j = 0; j < ncol; ++j) {
            info[j].iterations = tot[j];
            info[j].refinements = refin;
            info[j].rel_residual = rel[j];
            info[j].seconds = secs;
            info[j].iter_ms = iters > 0 ? 1e3 * secs / iters : 0.0;
        }
    if (worst() > o.rtol)
        throw Error(VEMD_ERR_NOT_CONVERGED, "not converged: largest relative residual " + std::to_string(worst()));
}

}  // namespace vemd
