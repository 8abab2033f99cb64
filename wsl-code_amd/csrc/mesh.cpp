// Mesh readers and the reference's mesh conventions
// (CPU/mesh/straight_mesh.cpp:92-214 VTK, :215-451 Gmsh, :462-482 CCW,
// :504-576 cell properties, :584-660 edge numbering).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <memory>
#include <sstream>

#include "common.h"

namespace vemd {

// Twice the signed (shoelace) area of the polygon v.
static double shoelace2(const std::vector<int>& v, const std::vector<double>& xy) {
    double a = 0.0;
    int n = (int)v.size();
    for (int i = 0; i < n; ++i) {
        int j = (i + 1) % n;
        a += xy[2 * v[i]] * xy[2 * v[j] + 1] - xy[2 * v[j]] * xy[2 * v[i] + 1];
    }
    return a;
}

static void make_ccw(std::vector<int>& v, const std::vector<double>& xy) {
    if (shoelace2(v, xy) < 0) std::reverse(v.begin(), v.end());
}

void mesh_finalize(Mesh& m) {
    m.nc = (int)m.off.size() - 1;
    m.cx.assign(m.nc, 0.0);
    m.cy.assign(m.nc, 0.0);
    m.area.assign(m.nc, 0.0);
    m.diam.assign(m.nc, 0.0);
    parallel_for(m.nc, [&](int64_t b, int64_t e) {
        std::vector<double> x, y;
        for (int64_t c = b; c < e; ++c) {
            int s = m.off[c], n = m.npc((int)c);
            x.resize(n);
            y.resize(n);
            for (int i = 0; i < n; ++i) {
                x[i] = m.xy[2 * m.nodes[s + i]];
                y[i] = m.xy[2 * m.nodes[s + i] + 1];
            }
            if (n == 4) {  // quads: bounding-box centre and area
                double x0 = *std::min_element(x.begin(), x.end()), x1 = *std::max_element(x.begin(), x.end());
                double y0 = *std::min_element(y.begin(), y.end()), y1 = *std::max_element(y.begin(), y.end());
                m.cx[c] = (x0 + x1) / 2.0;
                m.cy[c] = (y0 + y1) / 2.0;
                m.area[c] = (x1 - x0) * (y1 - y0);
            } else {  // vertex average and shoelace area
                double sx = 0, sy = 0, a = 0;
                for (int i = 0; i < n; ++i) {
                    sx += x[i];
                    sy += y[i];
                    int j = (i + 1) % n;
                    a += x[i] * y[j] - x[j] * y[i];
                }
                m.cx[c] = sx / n;
                m.cy[c] = sy / n;
                m.area[c] = std::fabs(a) / 2.0;
            }
            // largest vertex distance, rounded as the reference does
            // (sqrt of the sum of squares, CPU/mesh/straight_mesh.cpp:562-571)
            double md = 0.0;
            for (int i = 0; i < n; ++i)
                for (int j = i + 1; j < n; ++j) {
                    const double dx = x[i] - x[j], dy = y[i] - y[j];
                    md = std::max(md, std::sqrt(dx * dx + dy * dy));
                }
            if (!(md > 0.0)) throw Error(VEMD_ERR_MESH, "cell " + std::to_string(c) + " is degenerate");
            m.diam[c] = md;
        }
    });
    // Edges: unique sorted (min,max) node pairs, numbered in sorted order.
    size_t nt = m.nodes.size();
    std::vector<std::pair<uint64_t, int>> all(nt);
    for (int c = 0; c < m.nc; ++c) {
        int s = m.off[c], n = m.npc(c);
        for (int j = 0; j < n; ++j) all[s + j] = {Mesh::key(m.nodes[s + j], m.nodes[s + (j + 1) % n]), s + j};
    }
    std::sort(all.begin(), all.end());
    m.cell_edges.assign(nt, -1);
    m.edge_nodes.clear();
    m.edge_occ.clear();
    int ne = 0;
    for (size_t i = 0; i < nt; ++i) {
        if (i == 0 || all[i].first != all[i - 1].first) {
            m.edge_nodes.push_back((int)(all[i].first >> 32));
            m.edge_nodes.push_back((int)(all[i].first & 0xffffffffu));
            m.edge_occ.push_back(0);
            ++ne;
        }
        m.cell_edges[all[i].second] = ne - 1;
        m.edge_occ[ne - 1]++;
    }
    m.ne = ne;
    m.edge_cells.assign(2 * (size_t)ne, -1);
    for (int c = 0; c < m.nc; ++c)
        for (int j = m.off[c]; j < m.off[c + 1]; ++j) {
            int e = m.cell_edges[j];
            if (m.edge_occ[e] > 2) throw Error(VEMD_ERR_MESH, "edge shared by more than two cells");
            if (m.edge_cells[2 * e] < 0)
                m.edge_cells[2 * e] = c;
            else
                m.edge_cells[2 * e + 1] = c;
        }
}

// Appends one cell (vertex indices already checked against the node count),
// counter-clockwise as the reference's readers store it
// (CPU/mesh/straight_mesh.cpp:462-482).  Rejects what the element setup
// cannot integrate: non-finite coordinates, a repeated vertex (a zero-length
// edge or a pinched polygon) and zero area.
static void push_cell(Mesh& m, std::vector<int> v) {
    const std::string cell = "cell " + std::to_string(m.off.size() - 1);
    for (int a : v)
        if (!std::isfinite(m.xy[2 * a]) || !std::isfinite(m.xy[2 * a + 1]))
            throw Error(VEMD_ERR_MESH, cell + " has a non-finite vertex coordinate");
    std::vector<int> u(v);
    std::sort(u.begin(), u.end());
    if (std::adjacent_find(u.begin(), u.end()) != u.end()) throw Error(VEMD_ERR_MESH, cell + " repeats a vertex");
    if (shoelace2(v, m.xy) == 0.0) throw Error(VEMD_ERR_MESH, cell + " has zero area");
    make_ccw(v, m.xy);
    m.nodes.insert(m.nodes.end(), v.begin(), v.end());
    m.off.push_back((int)m.nodes.size());
}

Mesh* mesh_read_msh(const char* path) {
    std::ifstream f(path);
    if (!f) throw Error(VEMD_ERR_IO, std::string("cannot open ") + path);
    std::unique_ptr<Mesh> m(new Mesh());
    std::string line;
    bool found = false;
    while (std::getline(f, line))
        if (line.find("$Nodes") != std::string::npos) {
            found = true;
            break;
        }
    if (!found) {
        throw Error(VEMD_ERR_IO, "$Nodes section not found");
    }
    int nblocks = 0, nn = 0;
    std::getline(f, line);
    std::istringstream(line) >> nblocks >> nn;
    if (nn <= 0) {
        throw Error(VEMD_ERR_IO, "invalid number of nodes");
    }
    m->nn = nn;
    m->xy.assign(2 * (size_t)nn, 0.0);
    for (int b = 0; b < nblocks; ++b) {
        std::getline(f, line);
        int dim, tag, par, cnt = 0;
        std::istringstream(line) >> dim >> tag >> par >> cnt;
        std::vector<int> tags;
        while ((int)tags.size() < cnt && std::getline(f, line)) {
            std::istringstream ts(line);
            int t;
            while ((int)tags.size() < cnt && ts >> t) tags.push_back(t);
        }
        for (int i = 0; i < cnt; ++i) {
            std::getline(f, line);
            std::istringstream cs(line);
            double x, y, z;
            if (!(cs >> x >> y >> z)) continue;
            int id = tags[i] - 1;
            if (id >= 0 && id < nn) {
                m->xy[2 * id] = x;
                m->xy[2 * id + 1] = y;
            }
        }
    }
    found = false;
    while (std::getline(f, line))
        if (line.find("$Elements") != std::string::npos) {
            found = true;
            break;
        }
    if (!found) {
        throw Error(VEMD_ERR_IO, "$Elements section not found");
    }
    std::getline(f, line);
    std::istringstream(line) >> nblocks;
    m->off.push_back(0);
    for (int b = 0; b < nblocks; ++b) {
        std::getline(f, line);
        int dim, tag, type, cnt = 0;
        std::istringstream(line) >> dim >> tag >> type >> cnt;
        // Gmsh element types 2/3/11/12 are read as 3/4/5/6-vertex polygons
        // (the reference's convention, CPU/mesh/straight_mesh.cpp:442-450)
        int nv = type == 2 ? 3 : type == 3 ? 4 : type == 11 ? 5 : type == 12 ? 6 : -1;
        for (int i = 0; i < cnt; ++i) {
            std::getline(f, line);
            if (nv < 3) continue;
            std::istringstream es(line);
            int etag;
            es >> etag;
            std::vector<int> v(nv);
            for (int j = 0; j < nv; ++j) {
                es >> v[j];
                v[j] -= 1;
                if (v[j] < 0 || v[j] >= nn) {
                    throw Error(VEMD_ERR_IO, "element references a missing node");
                }
            }
            push_cell(*m, v);
        }
    }
    if (m->off.size() < 2) {
        throw Error(VEMD_ERR_IO, "no polygon elements");
    }
    mesh_finalize(*m);
    return m.release();
}

Mesh* mesh_read_vtk(const char* path) {
    std::ifstream f(path);
    if (!f) throw Error(VEMD_ERR_IO, std::string("cannot open ") + path);
    auto seek = [&](const char* k) {
        std::string t;
        while (f >> t)
            if (t == k) return true;
        return false;
    };
    std::unique_ptr<Mesh> m(new Mesh());
    std::string dt;
    if (!seek("POINTS") || !(f >> m->nn >> dt) || m->nn <= 0) {
        throw Error(VEMD_ERR_IO, "invalid POINTS section");
    }
    m->xy.assign(2 * (size_t)m->nn, 0.0);
    for (int i = 0; i < m->nn; ++i) {
        double z;
        if (!(f >> m->xy[2 * i] >> m->xy[2 * i + 1] >> z)) {
            throw Error(VEMD_ERR_IO, "truncated POINTS section");
        }
    }
    int nc = 0, tot = 0;
    if (!seek("CELLS") || !(f >> nc >> tot) || nc <= 0) {
        throw Error(VEMD_ERR_IO, "invalid CELLS section");
    }
    m->off.push_back(0);
    for (int c = 0; c < nc; ++c) {
        int nv = 0;
        f >> nv;
        if (nv < 3) {
            throw Error(VEMD_ERR_IO, "cell with fewer than 3 vertices");
        }
        std::vector<int> v(nv);
        for (int i = 0; i < nv; ++i) {
            f >> v[i];
            if (v[i] < 0 || v[i] >= m->nn) {
                throw Error(VEMD_ERR_IO, "cell references a missing node");
            }
        }
        push_cell(*m, v);
    }
    int nt = 0;
    if (!seek("CELL_TYPES") || !(f >> nt) || nt != nc) {
        throw Error(VEMD_ERR_IO, "CELL_TYPES missing or count mismatch");
    }
    for (int c = 0; c < nc; ++c) {
        int t = 0;
        f >> t;
        if (t != 5 && t != 7 && t != 9) {
            throw Error(VEMD_ERR_IO, "unsupported VTK cell type");
        }
    }
    mesh_finalize(*m);
    return m.release();
}

Mesh* mesh_concat(const std::vector<const Mesh*>& parts) {
    if (parts.empty()) throw Error(VEMD_ERR_ARG, "no meshes to concatenate");
    std::vector<double> xy;
    std::vector<int> off{0}, nodes, arc_nodes, cell0{0};
    std::vector<double> arc_r;
    int nn = 0;
    for (const Mesh* p : parts) {
        if (!p->inst_cell0.empty()) throw Error(VEMD_ERR_ARG, "a part is already a batch");
        xy.insert(xy.end(), p->xy.begin(), p->xy.end());
        for (int c = 0; c < p->nc; ++c) {
            for (int j = p->off[c]; j < p->off[c + 1]; ++j) nodes.push_back(p->nodes[j] + nn);
            off.push_back((int)nodes.size());
        }
        // arcs in a fixed order (sorted keys) so that the union is reproducible
        std::vector<std::pair<uint64_t, double>> arcs(p->arcs.begin(), p->arcs.end());
        std::sort(arcs.begin(), arcs.end());
        for (auto& kv : arcs) {
            arc_nodes.push_back((int)(kv.first >> 32) + nn);
            arc_nodes.push_back((int)(kv.first & 0xffffffffu) + nn);
            arc_r.push_back(kv.second);
        }
        nn += p->nn;
        cell0.push_back(cell0.back() + p->nc);
    }
    Mesh* m = mesh_from_arrays(nn, xy.data(), (int)off.size() - 1, off.data(), nodes.data(), (int)arc_r.size(),
                               arc_nodes.data(), arc_r.data());
    if (parts.size() > 1) m->inst_cell0 = cell0;
    return m;
}

Mesh* mesh_from_arrays(int nn, const double* xy, int nc, const int* off, const int* nodes, int narc,
                       const int* arc_nodes, const double* arc_r) {
    if (nn <= 0 || nc <= 0 || !xy || !off || !nodes) throw Error(VEMD_ERR_ARG, "empty mesh arrays");
    if (off[0] != 0) throw Error(VEMD_ERR_ARG, "cell offsets must start at 0");
    if (narc < 0 || (narc > 0 && (!arc_nodes || !arc_r))) throw Error(VEMD_ERR_ARG, "arc arrays");
    std::unique_ptr<Mesh> m(new Mesh());
    m->nn = nn;
    m->xy.assign(xy, xy + 2 * (size_t)nn);
    m->off.push_back(0);
    for (int c = 0; c < nc; ++c) {
        int n = off[c + 1] - off[c];
        if (n < 3) throw Error(VEMD_ERR_MESH, "cell with fewer than 3 vertices");
        std::vector<int> v(nodes + off[c], nodes + off[c + 1]);
        for (int a : v)
            if (a < 0 || a >= nn) throw Error(VEMD_ERR_MESH, "cell references a missing node");
        push_cell(*m, v);
    }
    mesh_finalize(*m);
    for (int a = 0; a < narc; ++a) {
        const int i = arc_nodes[2 * a], j = arc_nodes[2 * a + 1];
        if (i < 0 || i >= nn || j < 0 || j >= nn) throw Error(VEMD_ERR_MESH, "arc references a missing node");
        if (!std::isfinite(arc_r[a])) throw Error(VEMD_ERR_MESH, "non-finite arc radius");
        m->arcs[Mesh::key(i, j)] = arc_r[a];
    }
    return m.release();
}

}  // namespace vemd
