// The matrix-free cell apply of the device path, shared by the translation
// units that instantiate it (k = 0..2 in device.hip, k = 3 in device_k3.hip,
// which compile in parallel): operator-stream and buffer-resource helpers, the
// per-cell body (one cell per lane), the one-launch apply kernel and the
// adjoint contraction.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "devstate.h"

namespace vemd {

typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename T>
struct VT;
template <>
struct VT<double> {
    using V = d2v;
    static constexpr int W = 2;
};
template <>
struct VT<float> {
    using V = f4v;
    static constexpr int W = 4;
};

template <typename F, int... I>
__device__ __forceinline__ void static_for(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}

__host__ __device__ constexpr int tri_row(int t) {
    int i = 0;
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    return i;
}

template <typename T>
struct ApplyArgs {
    const GroupDesc* groups;
    const int* tile_group;
    const T* O;          // operator stream (packed A then W per cell, tile-major)
    const T* L;
    const T* scale;
    const uint32_t* idx;
    const uint32_t* wmask;  // row 0 of W per store position (bit j: -1, bit 16 + j: 0)
    const T* x;          // input when st == nullptr
    const T* xs[2];      // inputs selected by the Krylov iteration parity when st != nullptr
    T* y;                // cell-owned rows (internal dofs and pressure moments)
    T* h;                // edge results per cell side: k + 1 values per entry of idx
    int64_t cbase, cpad, lam_idx, nvec;
    int ntiles, per_xcd;
    int tile_base;       // first tile of the launch (the tuner launches one group's tiles)
    int force_var;       // >= 0: every group runs this variant (the tuner's trials)
    double* pdot;        // per-tile partial of x.Kx (nullptr: not needed)
    double* plam;        // per-tile partial of the multiplier row
    const MState* st;
};

// Buffer-resource access (raw buffer loads/stores): a scalar descriptor per
// array and one 32-bit per-lane byte offset, so streaming the operator costs no
// per-load 64-bit address registers; out-of-range accesses are dropped.
typedef unsigned int u2v __attribute__((ext_vector_type(2)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef __amdgpu_buffer_rsrc_t Rsrc;

__device__ __forceinline__ Rsrc make_rsrc(const void* p, int64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)(bytes < 0x7fffffff ? bytes : 0x7fffffff),
                                             0x00020000);
}

template <typename V>
__device__ __forceinline__ V bload_vec(Rsrc r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
__device__ __forceinline__ double bload1(Rsrc r, uint32_t voff, uint32_t soff, double*) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ float bload1(Rsrc r, uint32_t voff, uint32_t soff, float*) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ uint32_t bload_u32(Rsrc r, uint32_t voff, uint32_t soff) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}
__device__ __forceinline__ void bstore1(double v, Rsrc r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2v, v), r, voff, soff, 0);
}
__device__ __forceinline__ void bstore1(float v, Rsrc r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
}

typedef float f2v __attribute__((ext_vector_type(2)));

// The k + 1 adjacent moments of one edge: one 4-, 8- or 16-byte access.
template <typename T, int KK>
__device__ __forceinline__ void load_edge(Rsrc r, uint32_t off, T (&v)[KK]) {
    if constexpr (sizeof(T) == 8 && KK == 2) {
        d2v w = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0u, 0));
        v[0] = w[0];
        v[1] = w[1];
    } else if constexpr (sizeof(T) == 4 && KK == 4) {
        // (the whole vector is cast: a bit_cast of one subscripted element of
        // a vector reads element 0 whatever the subscript)
        const f4v w = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0u, 0));
        v[0] = w[0];
        v[1] = w[1];
        v[2] = w[2];
        v[3] = w[3];
    } else if constexpr (sizeof(T) * KK == 8 && KK == 2) {
        f2v w = __builtin_bit_cast(f2v, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0u, 0));
        v[0] = w[0];
        v[1] = w[1];
    } else {
#pragma unroll
        for (int i = 0; i < KK; ++i) v[i] = bload1(r, off + i * (uint32_t)sizeof(T), 0u, (T*)nullptr);
    }
}
template <typename T, int KK>
__device__ __forceinline__ void store_edge(Rsrc r, uint32_t off, const T (&v)[KK]) {
    if constexpr (sizeof(T) == 8 && KK == 2) {
        d2v w = {v[0], v[1]};
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4v, w), r, off, 0u, 0);
    } else if constexpr (sizeof(T) == 4 && KK == 4) {
        const f4v w = {v[0], v[1], v[2], v[3]};
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4v, w), r, off, 0u, 0);
    } else if constexpr (sizeof(T) * KK == 8 && KK == 2) {
        f2v w = {v[0], v[1]};
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2v, w), r, off, 0u, 0);
    } else {
#pragma unroll
        for (int i = 0; i < KK; ++i) bstore1(v[i], r, off + i * (uint32_t)sizeof(T), 0u);
    }
}

// Marks register values as used and redefined here (an empty asm statement),
// so the arithmetic that produces them cannot sink below this point.
template <typename T, int M>
__device__ __forceinline__ void pin(T (&v)[M]) {
#pragma unroll
    for (int i = 0; i < M; ++i) asm volatile("" : "+v"(v[i]));
}

// Shape constants of the cell operator of one polygon size.
template <typename T, int K, int NE>
struct CellShape {
    static constexpr int VW = VT<T>::W;
    static constexpr int KK = K + 1, Q = (K + 1) * (K + 2) / 2, NI = (Q - 1) + (Q - (K + 1)), NED = KK * NE,
                         N = NED + NI;
    // streamed part of W: rows 1..Q-1 at the NED edge columns
    static constexpr int NA = N * (N + 1) / 2, NAV = (NA + VW - 1) / VW, NW = (Q - 1) * NED,
                         NWV = (NW + VW - 1) / VW;
    static constexpr int NV = NAV + NWV;  // 16-byte vectors per cell: packed A, then W
};

// Apply variants: the operator stream of a cell is consumed in chunks of CH
// 16-byte vectors, with (PIPE) or without the next chunk in flight while one
// is used. The autotuner picks one per valence.
template <int V>
struct Variant;
template <>
struct Variant<0> {
    static constexpr int CH = 8, PIPE = 1;
};
template <>
struct Variant<1> {
    static constexpr int CH = 4, PIPE = 1;
};
template <>
struct Variant<2> {
    static constexpr int CH = 2, PIPE = 1;
};
template <>
struct Variant<3> {
    static constexpr int CH = 8, PIPE = 0;
};
constexpr int NVARIANT = 4;

// One cell per lane: gather signed-folded local dofs, y_E = K_E x_E, store the
// edge-dof results to the slot of the cell's side of each edge.  The cell's
// operator is one stream of 16-byte vectors (tile-major: vector v of lane l of
// tile tau at ((tau * NV + v) * 64 + l) * 16 bytes), read in chunks of CH
// vectors; with PIPE the next chunk is in flight while the current one is
// consumed.  y_u = s (A u - W^T (p / s)) so that A and W entries can be
// consumed in stream order; the exact part of W (row 0 from the cell's mask,
// the -1 entries of the internal columns) is applied after the stream.
template <typename T, int K, int NE, int CH, int PIPE>
__device__ __forceinline__ void cell_body(const ApplyArgs<T>& a, const T* __restrict__ x, const GroupDesc& g, int tau,
                                          int l, double& dot, double& lamp) {
    using SH = CellShape<T, K, NE>;
    using V = typename VT<T>::V;
    constexpr int VW = SH::VW, Q = SH::Q, NI = SH::NI, NED = SH::NED, N = SH::N, NA = SH::NA, NAV = SH::NAV,
                  NW = SH::NW, NV = SH::NV;
    constexpr int NC = (NV + CH - 1) / CH;
    constexpr uint32_t TS = sizeof(T);
    constexpr int KK = SH::KK;
    const int lane = tau * 64 + l;
    const uint32_t gpos = (uint32_t)(g.gpos0 + lane);
    const Rsrc rx = make_rsrc(x, a.nvec * (int64_t)TS);
    const Rsrc ry = make_rsrc(a.y, a.nvec * (int64_t)TS);
    // the tile's side slots are contiguous: edge j of lane l at (j * 64 + l)
    const Rsrc rh = make_rsrc(a.h + (g.i_off + (int64_t)tau * NE * 64) * KK, (int64_t)NE * 64 * KK * TS);
    const Rsrc ri = make_rsrc(a.idx + g.i_off + (int64_t)tau * NE * 64, (int64_t)NE * 64 * 4);
    const Rsrc rO = make_rsrc(a.O + g.o_off + (int64_t)tau * NV * 64 * VW, (int64_t)NV * 64 * 16);
    const uint32_t lv = (uint32_t)l * 16u;
    uint32_t id[NE];
    T u[N], ps[Q], acc[N], yp[Q];
#pragma unroll
    for (int j = 0; j < NE; ++j) id[j] = bload_u32(ri, (uint32_t)l * 4u, (uint32_t)j * 256u);
#pragma unroll
    for (int j = 0; j < NE; ++j) {
        T v[KK];
        load_edge<T, KK>(rx, (id[j] & 0x7fffffffu) * TS, v);
#pragma unroll
        for (int r = 0; r < KK; ++r) u[r * NE + j] = v[r];
    }
#pragma unroll
    for (int j = 0; j < NI; ++j) u[NED + j] = bload1(rx, gpos * TS, (uint32_t)((a.cbase + j * a.cpad) * TS), (T*)nullptr);
    T p[Q];
#pragma unroll
    for (int r = 0; r < Q; ++r) p[r] = bload1(rx, gpos * TS, (uint32_t)((a.cbase + (NI + r) * a.cpad) * TS), (T*)nullptr);
    const T s = a.scale[gpos];
    const uint32_t wm = a.wmask[gpos];
    const T rs = s > T(0) ? T(1) / s : T(0);
#pragma unroll
    for (int r = 0; r < Q; ++r) {
        ps[r] = p[r] * rs;
        yp[r] = T(0);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = T(0);
    auto load_chunk = [&](auto cc, V(&buf)[CH]) {
        constexpr int c0 = decltype(cc)::value * CH;
        static_for(
            [&](auto vv) {
                constexpr int v = c0 + decltype(vv)::value;
                if constexpr (v < NV) buf[decltype(vv)::value] = bload_vec<V>(rO, lv, (uint32_t)v * 1024u);
            },
            std::make_integer_sequence<int, CH>{});
    };
    auto use_chunk = [&](auto cc, const V(&buf)[CH]) {
        constexpr int c0 = decltype(cc)::value * CH;
        static_for(
            [&](auto vv) {
                constexpr int v = c0 + decltype(vv)::value;
                static_for(
                    [&](auto ll) {
                        constexpr int li = decltype(ll)::value;
                        if constexpr (v < NAV) {
                            constexpr int t = v * VW + li;
                            if constexpr (t < NA) {
                                constexpr int i = tri_row(t);
                                constexpr int j = t - i * (i + 1) / 2;
                                const T aij = buf[decltype(vv)::value][li];
                                acc[i] += aij * u[j];
                                if constexpr (i != j) acc[j] += aij * u[i];
                            }
                        } else if constexpr (v < NV) {
                            constexpr int t = (v - NAV) * VW + li;
                            if constexpr (t < NW) {
                                constexpr int r = 1 + t / NED, i = t % NED;
                                const T w = buf[decltype(vv)::value][li];
                                acc[i] -= w * ps[r];
                                yp[r] -= w * u[i];
                            }
                        }
                    },
                    std::make_integer_sequence<int, VW>{});
            },
            std::make_integer_sequence<int, CH>{});
    };
    V b0[CH], b1[CH];
    if constexpr (PIPE) {
    load_chunk(std::integral_constant<int, 0>{}, b0);
    static_for(
        [&](auto cc) {
            constexpr int c = decltype(cc)::value;
            if constexpr (c + 1 < NC) {
                if constexpr ((c & 1) == 0)
                    load_chunk(std::integral_constant<int, c + 1>{}, b1);
                else
                    load_chunk(std::integral_constant<int, c + 1>{}, b0);
            }
            __builtin_amdgcn_sched_barrier(0);
            if constexpr ((c & 1) == 0)
                use_chunk(cc, b0);
            else
                use_chunk(cc, b1);
            pin(acc);
            pin(yp);
            __builtin_amdgcn_sched_barrier(0);
        },
        std::make_integer_sequence<int, NC>{});
    } else {
    (void)b1;
    static_for(
        [&](auto cc) {
            load_chunk(cc, b0);
            use_chunk(cc, b0);
            pin(acc);
            pin(yp);
            __builtin_amdgcn_sched_barrier(0);
        },
        std::make_integer_sequence<int, NC>{});
    }
    // exact part of W: row 0 (+-1 or 0 at the first moment of each edge) and
    // W(r + 1, NED + r) = -1
#pragma unroll
    for (int j = 0; j < NE; ++j) {
        const T w0 = (wm >> (16 + j) & 1u) ? T(0) : ((wm >> j & 1u) ? T(-1) : T(1));
        acc[j] -= w0 * ps[0];
        yp[0] -= w0 * u[j];
    }
#pragma unroll
    for (int r = 0; r + 1 < Q; ++r) {
        acc[NED + r] += ps[r + 1];
        yp[r + 1] += u[NED + r];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] *= s;
    if (a.lam_idx >= 0) {
        const T xl = x[a.lam_idx + g.inst];  // the instance's multiplier
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            T lr = a.L[g.l_off + r * (int64_t)g.npad + lane];
            yp[r] -= lr * xl;
            lamp -= (double)lr * (double)p[r];
        }
    }
    if (a.pdot) {
#pragma unroll
        for (int i = 0; i < N; ++i) dot += (double)u[i] * (double)acc[i];
#pragma unroll
        for (int r = 0; r < Q; ++r) dot += (double)p[r] * (double)yp[r];
    }
    // edge results to the cell's own side slots: each store covers 64
    // consecutive slots, so every written line is written whole
#pragma unroll
    for (int j = 0; j < NE; ++j) {
        T v[KK];
#pragma unroll
        for (int r = 0; r < KK; ++r) v[r] = acc[r * NE + j];
        store_edge<T, KK>(rh, (uint32_t)((j * 64 + l) * KK) * TS, v);
    }
#pragma unroll
    for (int j = 0; j < NI; ++j) bstore1(acc[NED + j], ry, gpos * TS, (uint32_t)((a.cbase + j * a.cpad) * TS));
#pragma unroll
    for (int r = 0; r < Q; ++r) bstore1(yp[r], ry, gpos * TS, (uint32_t)((a.cbase + (NI + r) * a.cpad) * TS));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// The body of one valence in the variant var (wave-uniform).
template <typename T, int K, int NE>
__device__ __forceinline__ void cell_variant(int var, const ApplyArgs<T>& a, const T* __restrict__ x,
                                             const GroupDesc& g, int tau, int l, double& dot, double& lamp) {
    if constexpr (K >= 2) {  // one variant (the tuner's variants differ only in speed)
        (void)var;
        cell_body<T, K, NE, Variant<0>::CH, Variant<0>::PIPE>(a, x, g, tau, l, dot, lamp);
        return;
    }
    switch (var) {
        case 1: cell_body<T, K, NE, Variant<1>::CH, Variant<1>::PIPE>(a, x, g, tau, l, dot, lamp); break;
        case 2: cell_body<T, K, NE, Variant<2>::CH, Variant<2>::PIPE>(a, x, g, tau, l, dot, lamp); break;
        case 3: cell_body<T, K, NE, Variant<3>::CH, Variant<3>::PIPE>(a, x, g, tau, l, dot, lamp); break;
        default: cell_body<T, K, NE, Variant<0>::CH, Variant<0>::PIPE>(a, x, g, tau, l, dot, lamp); break;
    }
}

// One launch for all cells: each wave takes one tile of 64 cells of one
// valence group and runs the body specialised for that polygon size (a
// wave-uniform branch).  Consecutive tiles go to one XCD so that neighbouring
// cells share its L2.
template <typename T, int K>
__global__ __launch_bounds__(256) void k_apply(ApplyArgs<T> a) {
    const int b = blockIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int tile_r = ((b & 7) * a.per_xcd + (b >> 3)) * 4 + wave;
    if (tile_r >= a.ntiles) return;
    const int tile = a.tile_base + tile_r;
    const GroupDesc g = a.groups[__builtin_amdgcn_readfirstlane(a.tile_group[tile])];
    const MState* st = a.st ? a.st + g.inst : nullptr;  // the Krylov state of the tile's instance
    if (st && st->done) return;
    const int l = threadIdx.x & 63;
    const int tau = tile - g.tile0;
    const T* x = st ? a.xs[st->itn & 1] : a.x;
    double dot = 0.0, lamp = 0.0;
    const int var = a.force_var >= 0 ? a.force_var : g.var;
    switch (g.ne) {
        case 3: cell_variant<T, K, 3>(var, a, x, g, tau, l, dot, lamp); break;
        case 4: cell_variant<T, K, 4>(var, a, x, g, tau, l, dot, lamp); break;
        case 5: cell_variant<T, K, 5>(var, a, x, g, tau, l, dot, lamp); break;
        case 6: cell_variant<T, K, 6>(var, a, x, g, tau, l, dot, lamp); break;
        case 7: cell_variant<T, K, 7>(var, a, x, g, tau, l, dot, lamp); break;
        case 8: cell_variant<T, K, 8>(var, a, x, g, tau, l, dot, lamp); break;
        case 9: cell_variant<T, K, 9>(var, a, x, g, tau, l, dot, lamp); break;
        case 10: cell_variant<T, K, 10>(var, a, x, g, tau, l, dot, lamp); break;
        default: break;
    }
    if (a.pdot || a.lam_idx >= 0) {
        dot = wave_sum(dot);
        lamp = wave_sum(lamp);
        if (l == 0) {
            if (a.pdot) a.pdot[tile] = dot;
            if (a.plam) a.plam[tile] = lamp;
        }
    }
}

// Adjoint contraction per cell: g_E = lambda_E^T (exp(-theta_E) A_E) x_E.
template <typename T, int K, int NE>
__device__ __forceinline__ double cell_quad(const ApplyArgs<T>& a, const T* lam, const T* x, const GroupDesc& g,
                                            int tau, int l) {
    using SH = CellShape<T, K, NE>;
    using V = typename VT<T>::V;
    constexpr int VW = SH::VW, NI = SH::NI, NED = SH::NED, N = SH::N, NA = SH::NA, NAV = SH::NAV, NV = SH::NV;
    const int64_t gpos = g.gpos0 + tau * 64 + l;
    T u[N], m[N];
#pragma unroll
    for (int j = 0; j < NED; ++j) {
        const int e = j % NE, r = j / NE;
        uint32_t k = (a.idx[g.i_off + ((int64_t)tau * NE + e) * 64 + l] & 0x7fffffffu) + r;
        u[j] = x[k];
        m[j] = lam[k];
    }
#pragma unroll
    for (int j = 0; j < NI; ++j) {
        u[NED + j] = x[a.cbase + j * a.cpad + gpos];
        m[NED + j] = lam[a.cbase + j * a.cpad + gpos];
    }
    double s = 0.0;
    const V* Av = reinterpret_cast<const V*>(a.O + g.o_off) + (int64_t)tau * NV * 64 + l;
    static_for(
        [&](auto vv) {
            constexpr int v = decltype(vv)::value;
            V av = Av[v * 64];
            static_for(
                [&](auto ll) {
                    constexpr int li = decltype(ll)::value;
                    constexpr int t = v * VW + li;
                    if constexpr (t < NA) {
                        constexpr int i = tri_row(t);
                        constexpr int j = t - i * (i + 1) / 2;
                        double av_ = (double)av[li];
                        s += av_ * (double)m[i] * (double)u[j];
                        if constexpr (i != j) s += av_ * (double)m[j] * (double)u[i];
                    }
                },
                std::make_integer_sequence<int, VW>{});
        },
        std::make_integer_sequence<int, NAV>{});
    return s * (double)a.scale[gpos];
}

template <typename T, int K>
__global__ __launch_bounds__(64) void k_adjoint(ApplyArgs<T> a, const T* lam, const T* x, const int* cell_of_gpos,
                                              double* grad) {
    const int tile = blockIdx.x;
    const GroupDesc g = a.groups[a.tile_group[tile]];
    const int tau = tile - g.tile0, l = threadIdx.x;
    double v = 0.0;
    switch (g.ne) {
        case 3: v = cell_quad<T, K, 3>(a, lam, x, g, tau, l); break;
        case 4: v = cell_quad<T, K, 4>(a, lam, x, g, tau, l); break;
        case 5: v = cell_quad<T, K, 5>(a, lam, x, g, tau, l); break;
        case 6: v = cell_quad<T, K, 6>(a, lam, x, g, tau, l); break;
        case 7: v = cell_quad<T, K, 7>(a, lam, x, g, tau, l); break;
        case 8: v = cell_quad<T, K, 8>(a, lam, x, g, tau, l); break;
        case 9: v = cell_quad<T, K, 9>(a, lam, x, g, tau, l); break;
        case 10: v = cell_quad<T, K, 10>(a, lam, x, g, tau, l); break;
        default: break;
    }
    int c = cell_of_gpos[g.gpos0 + tau * 64 + l];
    if (c >= 0) grad[c] = v;
}

// Launches of the k = 3 kernels (device_k3.hip).
template <typename T>
void launch_apply_k3(const ApplyArgs<T>& a, dim3 grid, hipStream_t stream);
void launch_adjoint_k3(const ApplyArgs<double>& a, int ntiles, hipStream_t stream, const double* lam, const double* x,
                       const int* cell_of_gpos, double* grad);

}  // namespace vemd
