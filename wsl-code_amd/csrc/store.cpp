// Cell ordering, valence grouping, edge sides and packing of the HBM operator store.
#include "store.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <string>

namespace vemd {

// Hilbert index of (x, y) on a 2^16 x 2^16 grid.
static uint64_t hilbert(uint32_t x, uint32_t y) {
    uint64_t d = 0;
    for (uint32_t s = 1u << 15; s > 0; s >>= 1) {
        uint32_t rx = (x & s) > 0, ry = (y & s) > 0;
        d += (uint64_t)s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - x;
                y = s - 1 - y;
            }
            std::swap(x, y);
        }
    }
    return d;
}

static std::vector<uint64_t> hilbert_keys(const std::vector<double>& px, const std::vector<double>& py) {
    double x0 = *std::min_element(px.begin(), px.end()), x1 = *std::max_element(px.begin(), px.end());
    double y0 = *std::min_element(py.begin(), py.end()), y1 = *std::max_element(py.begin(), py.end());
    double sc = 65535.0 / std::max(std::max(x1 - x0, y1 - y0), 1e-300);
    std::vector<uint64_t> k(px.size());
    for (size_t i = 0; i < px.size(); ++i)
        k[i] = hilbert((uint32_t)((px[i] - x0) * sc), (uint32_t)((py[i] - y0) * sc));
    return k;
}

Layout layout_build(const Mesh& m, const Setup& S) {
    Layout L;
    const int nc = m.nc;
    L.k = S.sh.k;
    L.q = S.sh.q;
    L.ni = S.sh.ni;
    L.lam = S.desc.bc == VEMD_BC_FLUX_MEANZERO;
    L.nedge = m.ne;
    L.ncell = nc;
    for (int c = 0; c < nc; ++c) {
        L.max_ne = std::max(L.max_ne, m.npc(c));
        if (m.npc(c) > MAX_NE)
            throw Error(VEMD_ERR_MESH, "cell " + std::to_string(c) + " has more than 10 edges (device limit)");
    }
    if (L.k > 3) throw Error(VEMD_ERR_ARG, "the device path supports k = 0 to 3");
    // batch instance of every cell
    L.ninst = m.ninst();
    std::vector<int> inst(nc, 0);
    for (int i = 0; i < L.ninst && !m.inst_cell0.empty(); ++i)
        for (int c = m.inst_cell0[i]; c < m.inst_cell0[i + 1]; ++c) inst[c] = i;
    // spatial order of cells (instance-major)
    std::vector<uint64_t> hk = hilbert_keys(m.cx, m.cy);
    std::vector<int> order(nc);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        if (inst[a] != inst[b]) return inst[a] < inst[b];
        return hk[a] < hk[b] || (hk[a] == hk[b] && a < b);
    });
    // store order: instance, valence (largest first, so the longest tiles
    // start first), then spatial
    std::vector<int> rank(nc);
    for (int i = 0; i < nc; ++i) rank[order[i]] = i;
    std::vector<int> so(nc);
    std::iota(so.begin(), so.end(), 0);
    std::sort(so.begin(), so.end(), [&](int a, int b) {
        if (inst[a] != inst[b]) return inst[a] < inst[b];
        if (m.npc(a) != m.npc(b)) return m.npc(a) > m.npc(b);
        return rank[a] < rank[b];
    });
    // groups with padding to whole tiles
    int gpos = 0, tile = 0;
    L.cell_of_gpos.clear();
    L.gpos_of_cell.assign(nc, -1);
    L.inst_tile0.assign(L.ninst + 1, 0);
    L.inst_gpos0.assign(L.ninst + 1, 0);
    for (size_t i = 0; i < so.size();) {
        size_t j = i;
        while (j < so.size() && m.npc(so[j]) == m.npc(so[i]) && inst[so[j]] == inst[so[i]]) ++j;
        GroupDesc g{};
        g.ne = m.npc(so[i]);
        g.inst = inst[so[i]];
        g.n = S.sh.nflux(g.ne);
        int cnt = (int)(j - i);
        g.npad = (cnt + TILE - 1) / TILE * TILE;
        g.gpos0 = gpos;
        g.tile0 = tile;
        for (int t = 0; t < g.npad / TILE; ++t) L.tile_group.push_back((int)L.groups.size());
        for (int a = 0; a < g.npad; ++a) {
            int c = a < cnt ? so[i + a] : -1;
            L.cell_of_gpos.push_back(c);
            if (c >= 0) L.gpos_of_cell[c] = gpos + a;
        }
        gpos += g.npad;
        tile += g.npad / TILE;
        L.groups.push_back(g);
        i = j;
    }
    L.ngroups = (int)L.groups.size();
    for (const GroupDesc& g : L.groups) {
        L.inst_tile0[g.inst + 1] = g.tile0 + g.npad / TILE;
        L.inst_gpos0[g.inst + 1] = g.gpos0 + g.npad;
    }
    for (int i = 0; i < L.ninst; ++i) {
        L.inst_tile0[i + 1] = std::max(L.inst_tile0[i + 1], L.inst_tile0[i]);
        L.inst_gpos0[i + 1] = std::max(L.inst_gpos0[i + 1], L.inst_gpos0[i]);
    }
    L.spatial_gpos.resize(nc);
    for (int i = 0; i < nc; ++i) L.spatial_gpos[i] = L.gpos_of_cell[order[i]];
    L.cpad = gpos;
    // An interior edge's first side is the adjacent cell with the lower store
    // position.  Edges are numbered in the order of their first side's slot
    // in the apply's side array (group, tile, local edge, lane): for each
    // local edge j of a tile, the lanes that are the first side of their j-th
    // edge read consecutive edges of x in the apply, and the Krylov sweep's
    // edge sums read those side slots in order.  VEMD_EDGE_ORDER=hilbert
    // numbers edges by the spatial rank of their first owner instead.
    std::vector<int> side0(m.ne, INT32_MAX);
    for (int c = 0; c < nc; ++c)
        for (int j = m.off[c]; j < m.off[c + 1]; ++j)
            side0[m.cell_edges[j]] = std::min(side0[m.cell_edges[j]], L.gpos_of_cell[c]);
    L.edge_int.assign(m.ne, -1);
    const char* eord = std::getenv("VEMD_EDGE_ORDER");
    if (eord && std::string(eord) == "hilbert") {
        std::vector<int> efirst(m.ne, INT32_MAX);
        for (int c = 0; c < nc; ++c)
            for (int j = m.off[c]; j < m.off[c + 1]; ++j) {
                int e = m.cell_edges[j];
                efirst[e] = std::min(efirst[e], rank[c]);
            }
        std::vector<int> eo(m.ne);
        std::iota(eo.begin(), eo.end(), 0);
        std::sort(eo.begin(), eo.end(),
                  [&](int a, int b) { return efirst[a] < efirst[b] || (efirst[a] == efirst[b] && a < b); });
        for (int i = 0; i < m.ne; ++i) L.edge_int[eo[i]] = i;
    } else {
        int next = 0;
        for (const GroupDesc& g : L.groups)
            for (int t = 0; t < g.npad / TILE; ++t)
                for (int j = 0; j < g.ne; ++j)
                    for (int l = 0; l < TILE; ++l) {
                        const int gp = g.gpos0 + t * TILE + l;
                        const int c = L.cell_of_gpos[gp];
                        if (c < 0) continue;
                        const int e = m.cell_edges[m.off[c] + j];
                        if (side0[e] == gp) L.edge_int[e] = next++;
                    }
        if (next != m.ne) throw Error(VEMD_ERR_INTERNAL, "edge numbering missed edges");
    }
    // edges of an instance are contiguous (owners' ranks are instance-major)
    L.inst_edge0.assign(L.ninst + 1, 0);
    for (int e = 0; e < m.ne; ++e) L.inst_edge0[inst[m.edge_cells[2 * e]] + 1]++;
    for (int i = 0; i < L.ninst; ++i) L.inst_edge0[i + 1] += L.inst_edge0[i];
    // internal vector layout
    int kk = L.k + 1;
    L.cbase = (int64_t)kk * m.ne;
    int nslot = L.ni + L.q;
    // trash slots of padding lanes (k + 1 entries, aligned like an edge's
    // moments since cbase and cpad are even), then the multipliers (one per
    // instance: lam_idx + i)
    L.trash_idx = L.cbase + (int64_t)nslot * L.cpad;
    L.lam_idx = L.lam ? L.trash_idx + kk : -1;
    L.nvec = L.trash_idx + kk + (L.lam ? L.ninst : 0);
    // Krylov sweep chunks, instance-major (an instance's multiplier is a
    // chunk of its own, so that it follows that instance's recurrence)
    L.chunks.clear();
    L.pchunks.clear();
    L.inst_chunk0.assign(L.ninst + 1, 0);
    auto add = [&](int64_t b, int64_t e, int i, int pslot) {
        for (int64_t s = b; s < e; s += CHUNK) {
            if (pslot) L.pchunks.push_back((int)L.chunks.size());
            L.chunks.push_back(VecChunk{s, (int)std::min<int64_t>(CHUNK, e - s), i, pslot});
        }
    };
    for (int i = 0; i < L.ninst; ++i) {
        add((int64_t)kk * L.inst_edge0[i], (int64_t)kk * L.inst_edge0[i + 1], i, 0);
        for (int sl = 0; sl < nslot; ++sl)
            add(L.cbase + (int64_t)sl * L.cpad + L.inst_gpos0[i], L.cbase + (int64_t)sl * L.cpad + L.inst_gpos0[i + 1], i,
                sl >= L.ni ? 1 + sl - L.ni : 0);
        if (L.lam) add(L.lam_idx + i, L.lam_idx + i + 1, i, 0);
        if (i + 1 == L.ninst) add(L.trash_idx, L.trash_idx + kk, i, 0);  // trash slot
        L.inst_chunk0[i + 1] = (int)L.chunks.size();
    }
    // what the FP32 sweep by entry pairs (k = 1) assumes of the chunks: edge
    // and pressure chunks start at an even entry and hold whole pairs (it
    // reads an edge's two moments and a chunk's entries two at a time); only
    // a multiplier has a chunk of one entry
    for (const VecChunk& ch : L.chunks) {
        if (L.k != 1) break;
        const bool edge = ch.start < L.cbase, pres = ch.pslot != 0;
        const bool ok = ch.len >= 1 && (ch.len == 1 ? !edge && !pres : (ch.start % 2 == 0)) &&
                        (!(edge || pres) || ch.len % 2 == 0);
        if (!ok) throw Error(VEMD_ERR_INTERNAL, "Krylov sweep chunk of an unexpected shape");
    }
    // reference -> internal map
    L.int_of_ref.assign(S.ntot, -1);
    for (int e = 0; e < m.ne; ++e)
        for (int r = 0; r < kk; ++r) L.int_of_ref[(int64_t)r * m.ne + e] = (int64_t)kk * L.edge_int[e] + r;
    for (int c = 0; c < nc; ++c) {
        int64_t gp = L.gpos_of_cell[c];
        for (int i = 0; i < L.ni; ++i) L.int_of_ref[S.nflux_total - (int64_t)nc * L.ni + (int64_t)c * L.ni + i] = L.cbase + i * L.cpad + gp;
        for (int r = 0; r < L.q; ++r) L.int_of_ref[S.nflux_total + (int64_t)c * L.q + r] = L.cbase + (L.ni + r) * L.cpad + gp;
    }
    if (L.lam)
        for (int i = 0; i < L.ninst; ++i) L.int_of_ref[S.ntot - L.ninst + i] = L.lam_idx + i;
    // packed edge-dof indices (side0: the edges' first sides, above)
    // per-cell stream lengths in 16-byte vectors: packed A, then the rows
    // 1..q-1 of W at the edge columns, each padded
    auto nvec_of = [&](const GroupDesc& g, int vw) {
        int na = g.n * (g.n + 1) / 2, nw = (L.q - 1) * (L.k + 1) * g.ne;
        return (na + vw - 1) / vw + (nw + vw - 1) / vw;
    };
    int64_t ioff = 0, ooff2 = 0, ooff4 = 0, loff = 0;
    L.groups32 = L.groups;
    for (size_t gi = 0; gi < L.groups.size(); ++gi) {
        GroupDesc& g = L.groups[gi];
        g.i_off = ioff;
        g.o_off = ooff2;
        g.l_off = loff;
        L.groups32[gi] = g;
        L.groups32[gi].o_off = ooff4;
        ioff += (int64_t)g.ne * g.npad;
        ooff2 += (int64_t)nvec_of(g, 2) * 2 * g.npad;
        ooff4 += (int64_t)nvec_of(g, 4) * 4 * g.npad;
        loff += (int64_t)L.q * g.npad;
    }
    L.o_len2 = ooff2;
    L.o_len4 = ooff4;
    L.l_len = loff;
    L.idx.assign(ioff, (uint32_t)L.trash_idx);
    if (L.nvec >= (int64_t)SIDE1) throw Error(VEMD_ERR_MESH, "mesh too large for 31-bit dof indices");
    // The apply reads and writes the vectors through buffer resources with
    // 32-bit byte offsets (records beyond the range read as zero and writes are
    // dropped), so a vector of the solve must stay below 2^31 bytes in FP64:
    // at most 268,435,455 internal entries.  VEMD_MAX_VECTOR_BYTES lowers the
    // limit (tests force the error with it); it cannot raise it.
    int64_t vec_limit = (int64_t)INT32_MAX;
    if (const char* e = std::getenv("VEMD_MAX_VECTOR_BYTES")) vec_limit = std::min(vec_limit, (int64_t)std::max(0LL, std::atoll(e)));
    if (L.nvec * (int64_t)sizeof(double) > vec_limit)
        throw Error(VEMD_ERR_MESH, "mesh too large: " + std::to_string(L.nvec) + " internal vector entries (" +
                                       std::to_string(L.nvec * (int64_t)sizeof(double)) + " bytes in FP64) exceed the " +
                                       std::to_string(vec_limit) + "-byte range of the apply's 32-bit offsets");
    if (ioff + 1 >= (int64_t)INT32_MAX) throw Error(VEMD_ERR_MESH, "mesh too large for 31-bit side slots");
    L.nside = ioff + 1;
    L.hside.assign(2 * (size_t)m.ne, (int32_t)ioff);
    // one index per local edge: the edge's first moment (its k + 1 moments are
    // adjacent), flagged when the cell is the edge's second side
    for (auto& g : L.groups) {
        for (int a = 0; a < g.npad; ++a) {
            int c = L.cell_of_gpos[g.gpos0 + a];
            if (c < 0) continue;
            for (int j = 0; j < g.ne; ++j) {
                int e = m.cell_edges[m.off[c] + j];
                uint32_t id = (uint32_t)(kk * L.edge_int[e]);
                const bool second = L.gpos_of_cell[c] != side0[e];
                if (second) id |= SIDE1;
                const int64_t slot = g.i_off + ((int64_t)(a / TILE) * g.ne + j) * TILE + a % TILE;
                L.idx[slot] = id;
                L.hside[2 * (size_t)L.edge_int[e] + (second ? 1 : 0)] = (int32_t)slot;
            }
        }
    }
    for (int e = 0; e < m.ne; ++e)
        if (L.hside[2 * (size_t)e] == (int32_t)ioff) throw Error(VEMD_ERR_MESH, "edge without a first side");
    return L;
}

void pack_store(const Mesh& m, const Setup& S, const Layout& L, std::vector<double>& O64, std::vector<double>& L64,
                std::vector<float>& O32, std::vector<float>& L32, std::vector<uint32_t>& wmask) {
    O64.assign(L.o_len2, 0.0);
    O32.assign(L.o_len4, 0.0f);
    L64.assign(L.l_len, 0.0);
    L32.assign(L.l_len, 0.0f);
    wmask.assign(L.cpad, 0u);
    const int q = L.q;
    parallel_for((int64_t)L.groups.size(), [&](int64_t gb, int64_t ge) {
        std::vector<double> a;
        for (int64_t gi = gb; gi < ge; ++gi) {
            const GroupDesc& g = L.groups[gi];
            const int64_t off2 = g.o_off, off4 = L.groups32[gi].o_off;
            const int n = g.n, ned = (L.k + 1) * g.ne, na = n * (n + 1) / 2, nw = (q - 1) * ned;
            const int nav2 = (na + 1) / 2, nv2 = nav2 + (nw + 1) / 2;
            const int nav4 = (na + 3) / 4, nv4 = nav4 + (nw + 3) / 4;
            for (int lane = 0; lane < g.npad; ++lane) {
                int c = L.cell_of_gpos[g.gpos0 + lane];
                if (c < 0) continue;
                const int64_t tau = lane / TILE, l = lane % TILE;
                // element t of the stream (A entries first, W from the next whole vector)
                auto put = [&](int t_a, int t_w, double v) {
                    int t2 = t_a >= 0 ? t_a : nav2 * 2 + t_w;
                    int t4 = t_a >= 0 ? t_a : nav4 * 4 + t_w;
                    O64[off2 + ((tau * nv2 + t2 / 2) * TILE + l) * 2 + t2 % 2] = v;
                    O32[off4 + ((tau * nv4 + t4 / 4) * TILE + l) * 4 + t4 % 4] = (float)v;
                };
                const double* Ac = &S.A[S.ofs.a_off[c]];
                const double* Wc = &S.W[S.ofs.w_off[c]];
                const char* isb = &S.isb[S.ofs.v_off[c]];
                a.assign(Ac, Ac + na);
                // decouple essential boundary dofs: zero row/column, unit diagonal
                for (int i = 0; i < n; ++i)
                    for (int j = 0; j <= i; ++j)
                        if (isb[i] || isb[j]) a[i * (i + 1) / 2 + j] = (i == j) ? 1.0 : 0.0;
                for (int t = 0; t < na; ++t) put(t, -1, a[t]);
                for (int t = 0; t < nw; ++t) {
                    const int r = 1 + t / ned, i = t % ned;
                    put(-1, t, isb[i] ? 0.0 : Wc[r * n + i]);
                }
                // the exact part of W: row 0 and the internal columns
                uint32_t mk = 0;
                for (int i = 0; i < n; ++i)
                    for (int r = 0; r < q; ++r) {
                        double want;
                        if (i >= ned)
                            want = (i - ned < q - 1 && r == i - ned + 1) ? -1.0 : 0.0;
                        else if (r > 0)
                            continue;
                        else if (i >= g.ne)
                            want = 0.0;
                        else
                            want = std::fabs(Wc[i]) == 1.0 ? Wc[i] : 2.0;
                        if (Wc[r * n + i] != want)
                            throw Error(VEMD_ERR_INTERNAL, "unexpected divergence entry in cell " + std::to_string(c));
                    }
                for (int j = 0; j < g.ne; ++j) {
                    if (isb[j]) mk |= 1u << (16 + j);
                    else if (Wc[j] < 0.0) mk |= 1u << j;
                }
                wmask[g.gpos0 + lane] = mk;
                for (int r = 0; r < q; ++r) {
                    L64[g.l_off + (int64_t)r * g.npad + lane] = S.L[(size_t)c * q + r];
                    L32[g.l_off + (int64_t)r * g.npad + lane] = (float)S.L[(size_t)c * q + r];
                }
            }
        }
    });
    (void)m;
}

void pack_scale(const Setup& S, const Layout& L, std::vector<double>& s) {
    s.assign(L.cpad, 0.0);
    parallel_for(L.cpad, [&](int64_t b, int64_t e) {
        for (int64_t gp = b; gp < e; ++gp) {
            const int c = L.cell_of_gpos[gp];
            if (c >= 0) s[gp] = std::exp(-S.theta[c]);
        }
    });
}

DiagPlan diag_plan(const Mesh& m, const Setup& S, const Layout& L) {
    DiagPlan P;
    const int64_t nf = L.cbase + (int64_t)L.ni * L.cpad;
    P.c0.assign(nf, -1);
    P.c1.assign(nf, -1);
    for (int c = 0; c < m.nc; ++c) {
        const int n = S.sh.nflux(m.npc(c));
        for (int i = 0; i < n; ++i) {
            const int64_t v = S.ofs.v_off[c] + i;
            if (S.isb[v]) P.essential = true;
            const int64_t t = L.int_of_ref[S.gdof[v]];
            if (P.c0[t] < 0)
                P.c0[t] = v;
            else if (P.c1[t] < 0)
                P.c1[t] = v;
            else
                throw Error(VEMD_ERR_INTERNAL, "flux dof shared by more than two cells");
        }
    }
    return P;
}

void build_minv(const Mesh& m, const Setup& S, const Layout& L, const DiagPlan& plan, std::vector<double>& minv) {
    const int q = L.q;
    const int64_t nv = S.ofs.v_off[m.nc];
    // each local flux dof's diagonal, exp(-theta) A_ii (1 on an essential dof)
    std::vector<double> cv(nv);
    parallel_for(m.nc, [&](int64_t b, int64_t e) {
        for (int64_t c = b; c < e; ++c) {
            const int n = S.sh.nflux(m.npc((int)c));
            const double sc = std::exp(-S.theta[c]);
            const double* A = &S.A[S.ofs.a_off[c]];
            for (int i = 0; i < n; ++i) {
                const int64_t v = S.ofs.v_off[c] + i;
                cv[v] = S.isb[v] ? 1.0 : sc * A[i * (i + 1) / 2 + i];
            }
        }
    });
    // flux: the assembled diagonal, contributions added in cell order
    std::vector<double> d(L.nvec, 0.0);
    const int64_t nf = (int64_t)plan.c0.size();
    parallel_for(nf, [&](int64_t b, int64_t e) {
        for (int64_t t = b; t < e; ++t) {
            double v = 0.0;
            if (plan.c0[t] >= 0) v += cv[plan.c0[t]];
            if (plan.c1[t] >= 0) v += cv[plan.c1[t]];
            d[t] = v;
        }
    });
    // pressure: diag of W diag(A)^{-1} W^T per cell; multiplier: sum L^2 / S
    std::vector<double> lc(L.lam ? (size_t)m.nc : 0);
    parallel_for(m.nc, [&](int64_t b, int64_t e) {
        for (int64_t c = b; c < e; ++c) {
            const int n = S.sh.nflux(m.npc((int)c));
            const double* W = &S.W[S.ofs.w_off[c]];
            double lsum = 0.0;
            for (int r = 0; r < q; ++r) {
                double sr = 0.0;
                for (int i = 0; i < n; ++i) {
                    const int64_t v = S.ofs.v_off[c] + i;
                    if (S.isb[v]) continue;
                    const double w = W[r * n + i];
                    sr += w * w / d[L.int_of_ref[S.gdof[v]]];
                }
                if (!(sr > 0.0)) sr = 1.0;
                d[L.int_of_ref[S.nflux_total + c * q + r]] = sr;
                const double l = S.L[(size_t)c * q + r];
                lsum += l * l / sr;
            }
            if (L.lam) lc[c] = lsum;
        }
    });
    if (L.lam) {
        // per instance: its cells in order
        for (int i = 0; i < L.ninst; ++i) {
            const int c0 = m.inst_cell0.empty() ? 0 : m.inst_cell0[i];
            const int c1 = m.inst_cell0.empty() ? m.nc : m.inst_cell0[i + 1];
            double lam = 0.0;
            for (int c = c0; c < c1; ++c) lam += lc[c];
            d[L.lam_idx + i] = lam > 0 ? lam : 1.0;
        }
    }
    minv.assign(L.nvec, 1.0);
    parallel_for(L.nvec, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i)
            if (d[i] > 0.0) minv[i] = 1.0 / d[i];
    });
}

void build_rhs_ref(const Mesh& m, const Setup& S, std::vector<double>& b) {
    b.assign(S.ntot, 0.0);
    const int q = S.sh.q;
    for (int c = 0; c < m.nc; ++c) {
        int n = S.sh.nflux(m.npc(c));
        double sc = std::exp(-S.theta[c]);
        const double* A = &S.A[S.ofs.a_off[c]];
        const double* W = &S.W[S.ofs.w_off[c]];
        int64_t v0 = S.ofs.v_off[c];
        bool ess = false;  // a cell without essential dofs has no lift
        for (int i = 0; i < n && !ess; ++i) ess = S.isb[v0 + i] != 0;
        if (!ess) {
            for (int i = 0; i < n; ++i) b[S.gdof[v0 + i]] += S.ru[v0 + i];
            for (int r = 0; r < q; ++r) b[S.nflux_total + (int64_t)c * q + r] += S.f[(size_t)c * q + r];
            continue;
        }
        for (int i = 0; i < n; ++i) {
            int64_t g = S.gdof[v0 + i];
            if (S.isb[v0 + i]) {
                b[g] = S.ub[v0 + i];
                continue;
            }
            double lift = 0.0;
            for (int j = 0; j < n; ++j)
                if (S.isb[v0 + j]) lift += A[i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i] * S.ub[v0 + j];
            b[g] += S.ru[v0 + i] - sc * lift;
        }
        for (int r = 0; r < q; ++r) {
            double lift = 0.0;
            for (int j = 0; j < n; ++j)
                if (S.isb[v0 + j]) lift += W[r * n + j] * S.ub[v0 + j];
            b[S.nflux_total + (int64_t)c * q + r] += S.f[(size_t)c * q + r] - lift;
        }
    }
}

void build_row_scale(const Mesh& m, const Setup& S, const Layout& L, std::vector<double>& rs) {
    rs.assign(L.nvec, 1.0);
    // a cell writes its own pressure rows and its essential flux rows (a
    // boundary edge has one cell): no two cells write one entry
    parallel_for(m.nc, [&](int64_t b, int64_t e) {
        for (int64_t c = b; c < e; ++c) {
            const int n = S.sh.nflux(m.npc((int)c));
            for (int i = 0; i < n; ++i) {
                const int64_t v = S.ofs.v_off[c] + i;
                if (S.isb[v]) rs[L.int_of_ref[S.gdof[v]]] = std::exp(-S.theta[c]);
            }
            for (int r = 0; r < L.q; ++r) rs[L.int_of_ref[S.nflux_total + c * L.q + r]] = -1.0;
        }
    });
    if (L.lam)
        for (int i = 0; i < L.ninst; ++i) rs[L.lam_idx + i] = -1.0;
}

}  // namespace vemd
