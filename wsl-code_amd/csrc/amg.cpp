// Host construction of the smoothed-aggregation hierarchy (Vanek, Mandel and
// Brezina, "Algebraic multigrid by smoothed aggregation for second and fourth
// order elliptic problems", Computing 56, 1996): strength graph, greedy
// aggregation in spatial order, tentative prolongator from the near-null
// vector, one damped-Jacobi smoothing step, Galerkin coarse operators.
#include "amg.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace vemd {

void csr_spmv(const Csr& A, const double* x, double* y) {
    parallel_for(A.nrows, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            double s = 0.0;
            for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p) s += A.v[p] * x[A.ci[p]];
            y[i] = s;
        }
    });
}

Csr csr_transpose(const Csr& A) {
    Csr T;
    T.nrows = A.ncols;
    T.ncols = A.nrows;
    T.rp.assign(T.nrows + 1, 0);
    for (int c : A.ci) T.rp[c + 1]++;
    for (int64_t i = 0; i < T.nrows; ++i) T.rp[i + 1] += T.rp[i];
    T.ci.resize(A.nnz());
    T.v.resize(A.nnz());
    std::vector<int64_t> pos(T.rp.begin(), T.rp.end() - 1);
    for (int64_t i = 0; i < A.nrows; ++i)
        for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            int64_t q = pos[A.ci[p]]++;
            T.ci[q] = (int)i;
            T.v[q] = A.v[p];
        }
    return T;
}

// Gustavson's row-by-row product, two passes (structure, then values), rows
// in parallel; columns of each row sorted.
Csr csr_mult(const Csr& A, const Csr& B) {
    Csr C;
    C.nrows = A.nrows;
    C.ncols = B.ncols;
    C.rp.assign(C.nrows + 1, 0);
    parallel_for(A.nrows, [&](int64_t b, int64_t e) {
        std::vector<int64_t> mark(B.ncols, -1);
        for (int64_t i = b; i < e; ++i) {
            int64_t cnt = 0;
            for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
                int k = A.ci[p];
                for (int64_t t = B.rp[k]; t < B.rp[k + 1]; ++t)
                    if (mark[B.ci[t]] != i) {
                        mark[B.ci[t]] = i;
                        ++cnt;
                    }
            }
            C.rp[i + 1] = cnt;
        }
    });
    for (int64_t i = 0; i < C.nrows; ++i) C.rp[i + 1] += C.rp[i];
    C.ci.resize(C.rp[C.nrows]);
    C.v.resize(C.rp[C.nrows]);
    parallel_for(A.nrows, [&](int64_t b, int64_t e) {
        std::vector<double> acc(B.ncols, 0.0);
        std::vector<int64_t> mark(B.ncols, -1);
        std::vector<int> cols;
        for (int64_t i = b; i < e; ++i) {
            cols.clear();
            for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
                int k = A.ci[p];
                double a = A.v[p];
                for (int64_t t = B.rp[k]; t < B.rp[k + 1]; ++t) {
                    int j = B.ci[t];
                    if (mark[j] != i) {
                        mark[j] = i;
                        acc[j] = 0.0;
                        cols.push_back(j);
                    }
                    acc[j] += a * B.v[t];
                }
            }
            std::sort(cols.begin(), cols.end());
            int64_t o = C.rp[i];
            for (size_t t = 0; t < cols.size(); ++t) {
                C.ci[o + t] = cols[t];
                C.v[o + t] = acc[cols[t]];
            }
        }
    });
    return C;
}

SchurPlan schur_plan(const Mesh& m, const Setup& S, const Layout& L) {
    const int q = L.q;
    const int64_t cpad = L.cpad, n = (int64_t)q * cpad;
    SchurPlan P;
    // neighbour cells (self included) of every store position, sorted by position
    std::vector<std::vector<int>> nb(cpad);
    parallel_for(cpad, [&](int64_t b, int64_t e) {
        for (int64_t g = b; g < e; ++g) {
            int c = L.cell_of_gpos[g];
            std::vector<int>& v = nb[g];
            v.push_back((int)g);
            if (c < 0) continue;
            for (int j = m.off[c]; j < m.off[c + 1]; ++j) {
                int ed = m.cell_edges[j];
                for (int t = 0; t < 2; ++t) {
                    int o = m.edge_cells[2 * ed + t];
                    if (o >= 0 && o != c) v.push_back(L.gpos_of_cell[o]);
                }
            }
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
        }
    });
    Csr& A = P.pattern;
    A.nrows = A.ncols = n;
    A.rp.assign(n + 1, 0);
    for (int64_t g = 0; g < cpad; ++g)
        for (int r = 0; r < q; ++r) A.rp[g * q + r + 1] = L.cell_of_gpos[g] < 0 ? 1 : (int64_t)q * nb[g].size();
    for (int64_t i = 0; i < n; ++i) A.rp[i + 1] += A.rp[i];
    A.ci.resize(A.rp[n]);
    const int64_t nv = S.ofs.v_off[m.nc];
    P.self_slot.assign(m.nc, 0);
    P.nbk.assign(cpad, 1);
    P.other_slot.assign(nv, -1);
    P.other_i.assign(nv, -1);
    P.tflux.assign(nv, -1);
    parallel_for(cpad, [&](int64_t b, int64_t e) {
        for (int64_t g = b; g < e; ++g) {
            int c = L.cell_of_gpos[g];
            if (c < 0) {
                for (int r = 0; r < q; ++r) A.ci[A.rp[g * q + r]] = (int)(g * q + r);
                continue;
            }
            const std::vector<int>& v = nb[g];
            const int nbk = (int)v.size();
            P.nbk[g] = nbk;
            auto slot = [&](int gp) { return (int)(std::lower_bound(v.begin(), v.end(), gp) - v.begin()); };
            P.self_slot[c] = slot((int)g);
            const int n1 = S.sh.nflux(m.npc(c));
            const int64_t v0 = S.ofs.v_off[c];
            for (int i = 0; i < n1; ++i) {
                if (S.isb[v0 + i]) continue;
                const int64_t gd = S.gdof[v0 + i];
                P.tflux[v0 + i] = L.int_of_ref[gd];
                // the other cell of an edge dof
                if (i >= (S.sh.k + 1) * m.npc(c)) continue;
                const int ed = m.cell_edges[m.off[c] + i % m.npc(c)];
                int o = m.edge_cells[2 * ed] == c ? m.edge_cells[2 * ed + 1] : m.edge_cells[2 * ed];
                if (o < 0) continue;
                const int n2 = S.sh.nflux(m.npc(o));
                const int64_t w0 = S.ofs.v_off[o];
                int i2 = -1;
                for (int t = 0; t < n2; ++t)
                    if (S.gdof[w0 + t] == gd) {
                        i2 = t;
                        break;
                    }
                if (i2 < 0) throw Error(VEMD_ERR_INTERNAL, "edge dof not found in the neighbouring cell");
                P.other_slot[v0 + i] = slot(L.gpos_of_cell[o]);
                P.other_i[v0 + i] = i2;
            }
            for (int r = 0; r < q; ++r) {
                int64_t o = A.rp[g * q + r];
                for (int t = 0; t < nbk; ++t)
                    for (int s = 0; s < q; ++s) A.ci[o++] = (int)((int64_t)v[t] * q + s);
            }
        }
    });
    return P;
}

Csr schur_values(const SchurPlan& P, const Mesh& m, const Setup& S, const Layout& L, const std::vector<double>& minv) {
    const int q = L.q;
    const int64_t cpad = L.cpad;
    Csr A = P.pattern;
    A.v.assign(A.ci.size(), 0.0);
    parallel_for(cpad, [&](int64_t b, int64_t e) {
        std::vector<double> blk;
        for (int64_t g = b; g < e; ++g) {
            int c = L.cell_of_gpos[g];
            if (c < 0) {
                for (int r = 0; r < q; ++r) A.v[A.rp[g * q + r]] = 1.0;
                continue;
            }
            const int nbk = P.nbk[g];
            blk.assign((size_t)nbk * q * q, 0.0);  // block b: q x q, row-major
            const int n1 = S.sh.nflux(m.npc(c));
            const int64_t v0 = S.ofs.v_off[c];
            const double* Wc = &S.W[S.ofs.w_off[c]];
            double* bs = &blk[(size_t)P.self_slot[c] * q * q];
            for (int i = 0; i < n1; ++i) {
                if (P.tflux[v0 + i] < 0) continue;
                const double dinv = minv[P.tflux[v0 + i]];
                // this cell
                for (int r = 0; r < q; ++r)
                    for (int s = 0; s < q; ++s) bs[r * q + s] += Wc[r * n1 + i] * Wc[s * n1 + i] * dinv;
                // the other cell of an edge dof
                if (P.other_slot[v0 + i] < 0) continue;
                const int ed = m.cell_edges[m.off[c] + i % m.npc(c)];
                const int o = m.edge_cells[2 * ed] == c ? m.edge_cells[2 * ed + 1] : m.edge_cells[2 * ed];
                const int n2 = S.sh.nflux(m.npc(o));
                const int i2 = P.other_i[v0 + i];
                const double* Wo = &S.W[S.ofs.w_off[o]];
                double* bo = &blk[(size_t)P.other_slot[v0 + i] * q * q];
                for (int r = 0; r < q; ++r)
                    for (int s = 0; s < q; ++s) bo[r * q + s] += Wc[r * n1 + i] * Wo[s * n2 + i2] * dinv;
            }
            for (int r = 0; r < q; ++r) {
                int64_t o = A.rp[g * q + r];
                for (int t = 0; t < nbk; ++t)
                    for (int s = 0; s < q; ++s) A.v[o++] = blk[(size_t)t * q * q + r * q + s];
            }
        }
    });
    return A;
}

Csr schur_pressure(const Mesh& m, const Setup& S, const Layout& L, const std::vector<double>& minv) {
    return schur_values(schur_plan(m, S, L), m, S, L, minv);
}

// Sum over a wavefront of 64 values: lane l adds the value of lane l ^ o for
// o = 32, 16, ..., 1 (every lane ends with the same total).
static double butterfly64(double (&v)[64]) {
    for (int o = 32; o > 0; o >>= 1) {
        double t[64];
        for (int l = 0; l < 64; ++l) t[l] = v[l] + v[l ^ o];
        std::copy(t, t + 64, v);
    }
    return v[0];
}

// The sums of a[i]^2 and b[i]^2 over i < n in the order of the device
// rebuild: each group of 64 consecutive entries by the butterfly, each group
// of 64 consecutive group sums by the butterfly, then those in order.
static void tree_norms(int64_t n, const double* a, const double* b, double& na, double& nb) {
    const int64_t G = (n + 63) / 64, H = (G + 63) / 64;
    std::vector<double> ga(G), gb(G);
    parallel_for(G, [&](int64_t lo, int64_t hi) {
        for (int64_t g = lo; g < hi; ++g) {
            double va[64], vb[64];
            for (int l = 0; l < 64; ++l) {
                const int64_t i = g * 64 + l;
                va[l] = i < n ? a[i] * a[i] : 0.0;
                vb[l] = i < n ? b[i] * b[i] : 0.0;
            }
            ga[g] = butterfly64(va);
            gb[g] = butterfly64(vb);
        }
    });
    na = 0.0;
    nb = 0.0;
    for (int64_t h = 0; h < H; ++h) {
        double va[64], vb[64];
        for (int l = 0; l < 64; ++l) {
            const int64_t g = h * 64 + l;
            va[l] = g < G ? ga[g] : 0.0;
            vb[l] = g < G ? gb[g] : 0.0;
        }
        na += butterfly64(va);
        nb += butterfly64(vb);
    }
}

// Upper estimate of the spectral radius of D^-1 A by power iteration: iters
// steps from *ev when it holds a previous estimate's vector of the right
// length (a rebuild of the values after a theta change), else 25 from a
// fixed pseudo-random vector; the last iterate is left in *ev.  (Fewer steps
// from the previous vector measured up to 10% low after a change to an
// independent theta field, beyond the 1.1 margin.)  The norms are summed in
// the fixed order of tree_norms, whatever the thread count.
static double spectral_radius(const Csr& A, const std::vector<double>& dinv, std::vector<double>* ev, int iters) {
    const int64_t n = A.nrows;
    std::vector<double> x(n), y(n);
    if (ev && (int64_t)ev->size() == n) {
        x = *ev;
    } else {
        iters = 25;
        uint64_t s = 0x9e3779b97f4a7c15ull;
        for (int64_t i = 0; i < n; ++i) {
            s ^= s << 13;
            s ^= s >> 7;
            s ^= s << 17;
            x[i] = 0.5 + (double)(s >> 11) * (1.0 / 9007199254740992.0);
        }
    }
    double lam = 0.0;
    for (int it = 0; it < iters; ++it) {
        csr_spmv(A, x.data(), y.data());
        parallel_for(n, [&](int64_t b, int64_t e) {
            for (int64_t i = b; i < e; ++i) y[i] *= dinv[i];
        });
        double ny, nx;
        tree_norms(n, y.data(), x.data(), ny, nx);
        if (!(ny > 0.0)) break;
        lam = std::sqrt(ny / nx);
        const double sc = 1.0 / std::sqrt(ny);
        parallel_for(n, [&](int64_t b, int64_t e) {
            for (int64_t i = b; i < e; ++i) x[i] = y[i] * sc;
        });
    }
    if (ev) ev->swap(x);
    return 1.1 * lam;
}

// Greedy aggregation of the nodes on the strength graph (three passes: whole
// neighbourhoods, attachment to an adjacent aggregate, leftovers).  Nodes with
// no strong neighbour stay unaggregated (-1).
static int aggregate(const std::vector<std::vector<std::pair<int, double>>>& strong, const std::vector<int>& order,
                     std::vector<int>& agg) {
    const int nn = (int)strong.size();
    agg.assign(nn, -1);
    int na = 0;
    for (int i : order) {
        if (agg[i] >= 0 || strong[i].empty()) continue;
        bool free_nbhd = true;
        for (auto& p : strong[i])
            if (agg[p.first] >= 0) {
                free_nbhd = false;
                break;
            }
        if (!free_nbhd) continue;
        agg[i] = na;
        for (auto& p : strong[i]) agg[p.first] = na;
        ++na;
    }
    std::vector<int> first = agg;
    for (int i : order) {
        if (agg[i] >= 0 || strong[i].empty()) continue;
        double best = -1.0;
        int ba = -1;
        for (auto& p : strong[i])
            if (first[p.first] >= 0 && p.second > best) {
                best = p.second;
                ba = first[p.first];
            }
        if (ba >= 0) agg[i] = ba;
    }
    for (int i : order) {
        if (agg[i] >= 0 || strong[i].empty()) continue;
        agg[i] = na;
        for (auto& p : strong[i])
            if (agg[p.first] < 0) agg[p.first] = na;
        ++na;
    }
    return na;
}

void dense_inverse(std::vector<double>& a, int64_t n) {
    // Gauss-Jordan with partial pivoting (the coarsest operator has at most a
    // few thousand rows); the row eliminations of each pivot run in parallel.
    std::vector<double> inv(n * n, 0.0);
    for (int64_t i = 0; i < n; ++i) inv[i * n + i] = 1.0;
    for (int64_t c = 0; c < n; ++c) {
        int64_t p = c;
        for (int64_t r = c + 1; r < n; ++r)
            if (std::abs(a[r * n + c]) > std::abs(a[p * n + c])) p = r;
        if (!(std::abs(a[p * n + c]) > 0.0)) throw Error(VEMD_ERR_SINGULAR, "singular coarsest multigrid operator");
        if (p != c)
            for (int64_t j = 0; j < n; ++j) {
                std::swap(a[p * n + j], a[c * n + j]);
                std::swap(inv[p * n + j], inv[c * n + j]);
            }
        double d = 1.0 / a[c * n + c];
        for (int64_t j = 0; j < n; ++j) {
            a[c * n + j] *= d;
            inv[c * n + j] *= d;
        }
        auto eliminate = [&](int64_t rb, int64_t re) {
            for (int64_t r = rb; r < re; ++r) {
                if (r == c) continue;
                double f = a[r * n + c];
                if (f == 0.0) continue;
                for (int64_t j = 0; j < n; ++j) {
                    a[r * n + j] -= f * a[c * n + j];
                    inv[r * n + j] -= f * inv[c * n + j];
                }
            }
        };
        if (n >= 512)
            parallel_for(n, eliminate);
        else
            eliminate(0, n);
    }
    a.swap(inv);
}

Amg amg_build(Csr A, const std::vector<int>& node_of_row0, int nnodes0, const std::vector<double>& b00,
              const std::vector<int>& node_order0, const AmgOptions& o, const std::vector<double>* l0, double rho,
              const Amg* reuse, const std::vector<int>* row_inst, const std::vector<std::vector<double>>* nv) {
    Amg H;
    H.opt = o;
    const int K0 = reuse ? reuse->nvec0 : 1 + (nv ? (int)nv->size() : 0);  // vectors of the first coarsening
    H.nvec0 = K0;
    int cur_bs = 1;  // rows per node of the current level (level 0: the caller's)
    std::vector<double> lvec;
    std::vector<int> rinst;  // instance of each row of the current level (empty: one rank-one term)
    if (l0) {
        if ((int64_t)l0->size() != A.nrows) throw Error(VEMD_ERR_INTERNAL, "rank-one term of the wrong length");
        lvec = *l0;
        H.rho = rho;
        if (row_inst) rinst = *row_inst;
    }
    std::vector<int> nor = node_of_row0, order = node_order0;
    std::vector<double> b0 = b00;
    int nn = nnodes0;
    double eps = o.strength;
    for (int lev = 0;; ++lev) {
        AmgLevel L;
        L.bs = cur_bs;
        const int64_t n = A.nrows;
        L.dinv.assign(n, 1.0);
        for (int64_t i = 0; i < n; ++i)
            for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
                if (A.ci[p] == i && A.v[p] != 0.0) L.dinv[i] = 1.0 / A.v[p];
        // reuse: the previous hierarchy's aggregates (its tentative
        // prolongators) and depth; only the values are rebuilt
        const AmgLevel* old = reuse && lev < (int)reuse->lv.size() ? &reuse->lv[lev] : nullptr;
        if (old) L.ev = old->ev;
        L.lmax = spectral_radius(A, L.dinv, &L.ev, 25);
        bool last = n <= o.coarse_size || lev + 1 >= o.max_levels;
        int na = 0;
        std::vector<int> agg;
        if (reuse) {
            if (!old || old->A.nrows != n) throw Error(VEMD_ERR_INTERNAL, "multigrid reuse: level sizes differ");
            last = lev + 1 == (int)reuse->lv.size();
        }
        if (!last && !old) {
            // node-level strength matrix from the near-null vector
            std::vector<std::vector<int>> rows_of(nn);
            for (int64_t r = 0; r < n; ++r)
                if (nor[r] >= 0 && b0[r] != 0.0) rows_of[nor[r]].push_back((int)r);
            std::vector<std::vector<std::pair<int, double>>> nm(nn), strong(nn);
            parallel_for(nn, [&](int64_t b, int64_t e) {
                std::vector<std::pair<int, double>> t;
                for (int64_t i = b; i < e; ++i) {
                    t.clear();
                    for (int r : rows_of[i])
                        for (int64_t p = A.rp[r]; p < A.rp[r + 1]; ++p) {
                            int cj = A.ci[p];
                            if (nor[cj] < 0 || b0[cj] == 0.0) continue;
                            t.push_back({nor[cj], b0[r] * A.v[p] * b0[cj]});
                        }
                    std::sort(t.begin(), t.end(), [](auto& x, auto& y) { return x.first < y.first; });
                    auto& out = nm[i];
                    for (auto& pr : t)
                        if (!out.empty() && out.back().first == pr.first)
                            out.back().second += pr.second;
                        else
                            out.push_back(pr);
                }
            });
            std::vector<double> diag(nn, 0.0);
            for (int i = 0; i < nn; ++i)
                for (auto& pr : nm[i])
                    if (pr.first == i) diag[i] = pr.second;
            for (int i = 0; i < nn; ++i)
                for (auto& pr : nm[i]) {
                    if (pr.first == i) continue;
                    double s = std::abs(pr.second);
                    if (s >= eps * std::sqrt(std::abs(diag[i] * diag[pr.first])) && s > 0.0)
                        strong[i].push_back({pr.first, s / std::sqrt(std::abs(diag[i] * diag[pr.first]))});
                }
            na = aggregate(strong, order, agg);
            if (na == 0 || na >= nn) last = true;
        }
        if (last) {
            L.A = std::move(A);
            H.lv.push_back(std::move(L));
            break;
        }
        // tentative prolongator: the near-null vector restricted to each
        // aggregate, normalised; the norms are the coarse near-null vector
        std::vector<double> bc;
        Csr Pt;
        int K = 1;  // coarse rows per aggregate
        if (old) {
            Pt = old->Pt;
            K = lev == 0 ? K0 : 1;
            na = (int)Pt.ncols;
        } else if (lev == 0 && K0 > 1) {
            // several near-null vectors: each aggregate's rows times the
            // vectors, Gram-Schmidt in order (the vectors are independent on
            // any aggregate of whole cells); the first vector's coarse image
            // (its norm on the aggregate, in the aggregate's first row) is
            // the next level's near-null vector
            K = K0;
            std::vector<std::vector<int64_t>> arows(na);
            for (int64_t r = 0; r < n; ++r)
                if (nor[r] >= 0 && agg[nor[r]] >= 0) arows[agg[nor[r]]].push_back(r);
            std::vector<std::vector<std::pair<int, double>>> prow(n);
            bc.assign((size_t)na * K, 0.0);
            for (int a = 0; a < na; ++a) {
                const std::vector<int64_t>& rr = arows[a];
                const int m = (int)rr.size();
                std::vector<std::vector<double>> V(K, std::vector<double>(m));
                for (int i = 0; i < m; ++i) {
                    V[0][i] = b0[rr[i]];
                    for (int k = 1; k < K; ++k) V[k][i] = (*nv)[k - 1][rr[i]];
                }
                for (int k = 0; k < K; ++k) {
                    for (int j = 0; j < k; ++j) {
                        double d = 0.0;
                        for (int i = 0; i < m; ++i) d += V[k][i] * V[j][i];
                        for (int i = 0; i < m; ++i) V[k][i] -= d * V[j][i];
                    }
                    double nn = 0.0;
                    for (int i = 0; i < m; ++i) nn += V[k][i] * V[k][i];
                    nn = std::sqrt(nn);
                    if (!(nn > 0.0)) throw Error(VEMD_ERR_INTERNAL, "dependent near-null vectors on an aggregate");
                    for (int i = 0; i < m; ++i) V[k][i] /= nn;
                    if (k == 0) bc[(size_t)a * K] = nn;
                }
                for (int i = 0; i < m; ++i)
                    for (int k = 0; k < K; ++k)
                        if (V[k][i] != 0.0) prow[rr[i]].push_back({a * K + k, V[k][i]});
            }
            Pt.nrows = n;
            Pt.ncols = (int64_t)na * K;
            Pt.rp.assign(n + 1, 0);
            for (int64_t r = 0; r < n; ++r) {
                Pt.rp[r + 1] = Pt.rp[r] + (int64_t)prow[r].size();
                for (const auto& e : prow[r]) {
                    Pt.ci.push_back(e.first);
                    Pt.v.push_back(e.second);
                }
            }
            na = na * K;
        } else {
            bc.assign(na, 0.0);
            for (int64_t r = 0; r < n; ++r)
                if (nor[r] >= 0 && agg[nor[r]] >= 0) bc[agg[nor[r]]] += b0[r] * b0[r];
            for (double& v : bc) v = std::sqrt(v);
            Pt.nrows = n;
            Pt.ncols = na;
            Pt.rp.assign(n + 1, 0);
            for (int64_t r = 0; r < n; ++r) {
                bool on = nor[r] >= 0 && agg[nor[r]] >= 0 && b0[r] != 0.0;
                Pt.rp[r + 1] = Pt.rp[r] + (on ? 1 : 0);
                if (on) {
                    int a = agg[nor[r]];
                    Pt.ci.push_back(a);
                    Pt.v.push_back(b0[r] / bc[a]);
                }
            }
        }
        // smoothed prolongator P = (I - omega D^-1 A) Pt
        const double omega = 4.0 / (3.0 * L.lmax);
        Csr AP = csr_mult(A, Pt);
        Csr P;
        P.nrows = n;
        P.ncols = na;
        P.rp.assign(n + 1, 0);
        for (int64_t r = 0; r < n; ++r) {
            int64_t p1 = Pt.rp[r], e1 = Pt.rp[r + 1], p2 = AP.rp[r], e2 = AP.rp[r + 1];
            const double f = omega * L.dinv[r];
            while (p1 < e1 || p2 < e2) {
                int c1 = p1 < e1 ? Pt.ci[p1] : INT32_MAX, c2 = p2 < e2 ? AP.ci[p2] : INT32_MAX;
                double v = 0.0;
                int c = std::min(c1, c2);
                if (c1 == c) v += Pt.v[p1++];
                if (c2 == c) v -= f * AP.v[p2++];
                P.ci.push_back(c);
                P.v.push_back(v);
            }
            P.rp[r + 1] = (int64_t)P.ci.size();
        }
        AP = Csr();
        Csr R = csr_transpose(P);
        Csr M = csr_mult(R, A);
        Csr Ac = csr_mult(M, P);
        if (!lvec.empty()) {
            std::vector<double> lc(R.nrows);
            csr_spmv(R, lvec.data(), lc.data());
            lvec.swap(lc);
        }
        if (!rinst.empty()) {
            // an aggregate's instance: that of its rows (the tentative prolongator)
            std::vector<int> ci(Pt.ncols, -1);
            for (int64_t r = 0; r < Pt.nrows; ++r)
                for (int64_t p = Pt.rp[r]; p < Pt.rp[r + 1]; ++p) {
                    int& t = ci[Pt.ci[p]];
                    if (t >= 0 && t != rinst[r]) throw Error(VEMD_ERR_INTERNAL, "an aggregate joins two instances");
                    t = rinst[r];
                }
            rinst.swap(ci);
        }
        L.A = std::move(A);
        L.P = std::move(P);
        L.R = std::move(R);
        L.M = std::move(M);
        L.Pt = std::move(Pt);
        H.lv.push_back(std::move(L));
        A = std::move(Ac);
        cur_bs = K;
        if (old) continue;
        // the next level: K rows per aggregate (one node), ordered as created (spatially)
        nn = na / K;
        nor.resize(na);
        for (int c = 0; c < na; ++c) nor[c] = c / K;
        order.resize(nn);
        std::iota(order.begin(), order.end(), 0);
        b0 = bc;
        eps *= 0.5;
    }
    // dense inverse of the coarsest operator
    const AmgLevel& C = H.lv.back();
    H.ncoarse = C.A.nrows;
    if (H.ncoarse > 8 * (int64_t)std::max(o.coarse_size, 1000))
        throw Error(VEMD_ERR_INTERNAL, "multigrid coarsening stalled at " + std::to_string(H.ncoarse) + " unknowns");
    H.coarse_inv.assign(H.ncoarse * H.ncoarse, 0.0);
    for (int64_t i = 0; i < H.ncoarse; ++i)
        for (int64_t p = C.A.rp[i]; p < C.A.rp[i + 1]; ++p) H.coarse_inv[i * H.ncoarse + C.A.ci[p]] = C.A.v[p];
    if (!lvec.empty())
        for (int64_t i = 0; i < H.ncoarse; ++i)
            for (int64_t j = 0; j < H.ncoarse; ++j)
                if (rinst.empty() || (rinst[i] >= 0 && rinst[i] == rinst[j]))
                    H.coarse_inv[i * H.ncoarse + j] += rho * lvec[i] * lvec[j];
    dense_inverse(H.coarse_inv, H.ncoarse);
    return H;
}

}  // namespace vemd
