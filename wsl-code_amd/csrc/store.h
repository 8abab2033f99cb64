// Host-side description of the HBM operator store: cell order, valence groups,
// internal dof numbering, edge sides and the packed per-precision arrays.
#pragma once

#include <cstdint>
#include <vector>

#include "common.h"

namespace vemd {

constexpr int TILE = 64;       // cells per tile: one wavefront, one lane per cell
constexpr int MAX_NE = 10;     // largest polygon the device kernels are built for
// Index flag of an edge dof seen from the second of its two cells: that cell's
// contribution goes to the side-1 array, the first cell's to the output vector,
// and the consumer adds the two (one writer per slot, no atomics, fixed order).
constexpr uint32_t SIDE1 = 0x80000000u;

// POD shared with the device code.
struct GroupDesc {
    int ne, n, npad;          // valence, local flux dofs, cells incl. padding
    int inst;                 // batch instance of the group's cells
    int gpos0, tile0;         // first store position, first global tile
    int var;                  // apply variant (stream chunk shape) chosen by the autotuner
    int64_t o_off, l_off, i_off;  // element offsets of the operator stream, L and index blocks
};

constexpr int CHUNK = 2048;    // entries per Krylov sweep chunk (256 threads x 8)

struct VecChunk {
    int64_t start;
    int len, inst;
    int pslot;  // 1 + pressure moment of the chunk's slot; 0 outside the pressure block
};

struct Layout {
    int k = 1, q = 0, ni = 0;
    int ngroups = 0, max_ne = 0;
    bool lam = false;
    int64_t nedge = 0, ncell = 0, cpad = 0;  // edges, real cells, padded cell slots
    int64_t cbase = 0;        // first cell-owned slot (= number of edge-dof entries)
    int64_t lam_idx = -1, trash_idx = 0, nvec = 0;
    std::vector<GroupDesc> groups;      // offsets of the FP64 store (vector width 2)
    std::vector<GroupDesc> groups32;    // the same groups with FP32 store offsets (width 4)
    std::vector<int> tile_group;        // group of each global tile
    std::vector<int> cell_of_gpos;      // mesh cell at each store position (-1 padding)
    std::vector<int> gpos_of_cell;
    std::vector<int> spatial_gpos;      // store positions of the cells in spatial (Hilbert) order
    std::vector<int> edge_int;          // internal number of each mesh edge
    std::vector<int64_t> int_of_ref;    // reference dof -> internal slot
    std::vector<uint32_t> idx;          // first moment of every local edge (tile-major per group), SIDE1 flag
    // Cell sides of the edges: the apply writes the k + 1 edge results of local
    // edge j of a cell to side slot p (the position of that edge's entry in
    // idx), values [p * (k + 1), (p + 1) * (k + 1)).  hside[2 e + s] is the
    // slot of side s of internal edge e; a boundary edge's side 1 is the zero
    // slot nside - 1, which no cell writes.
    std::vector<int32_t> hside;
    int64_t nside = 0;
    int64_t o_len2 = 0, o_len4 = 0, l_len = 0;  // packed sizes (elements) of the FP64 / FP32 streams
    // Batch instances (one for a plain mesh): each owns a contiguous range of
    // tiles, store positions and internal edges.
    int ninst = 1;
    std::vector<int> inst_tile0, inst_gpos0, inst_edge0;  // ninst + 1 each
    // Krylov vector sweeps: fixed chunks of at most CHUNK entries, each inside
    // one instance and one segment; an instance's chunks are contiguous
    // (inst_chunk0), so per-instance reductions sum a fixed range of partials.
    std::vector<VecChunk> chunks;
    std::vector<int> inst_chunk0;
    std::vector<int> pchunks;  // the chunks of the pressure block
};

// Builds ordering, grouping, numbering and edge sides from the mesh and setup.
Layout layout_build(const Mesh& m, const Setup& S);

// Packs each cell's operator stream -- A (packed symmetric lower triangle,
// essential boundary dofs decoupled) then the geometric part of W, rows
// 1..q-1 at the (k+1) ne edge columns, each padded to whole 16-byte vectors
// (2 FP64 / 4 FP32 values) -- tile-major: vector v of lane l of tile tau of a
// group at ((tau * nv + v) * 64 + l) * width; and L.  The rest of W is exact
// and not streamed: row 0 is +1 / -1 (direction sign) / 0 (decoupled boundary
// dof) at the first moment of each edge and 0 elsewhere -- recorded per store
// position in wmask (bit j: -1, bit 16 + j: 0, for local edge j) -- and the
// internal columns carry W(r + 1, edof + r) = -1 for r < q - 1 and zeros
// (the packing checks both).
void pack_store(const Mesh& m, const Setup& S, const Layout& L, std::vector<double>& O64, std::vector<double>& L64,
                std::vector<float>& O32, std::vector<float>& L32, std::vector<uint32_t>& wmask);

// Per-store-position cell scale exp(-theta) (padding 0).
void pack_scale(const Setup& S, const Layout& L, std::vector<double>& s);

// Theta-independent index data of the preconditioner diagonal, built once:
// for every internal flux slot the (at most two) local flux dofs of the setup
// (index into the per-cell dof arrays, cell order) whose diagonals sum there.
struct DiagPlan {
    std::vector<int64_t> c0, c1;  // -1: none
    bool essential = false;       // some flux dof is an essential boundary dof
};
DiagPlan diag_plan(const Mesh& m, const Setup& S, const Layout& L);

// Inverse of the block-diagonal preconditioner (internal order):
// flux: assembled diag of exp(-theta) A; pressure: sum_i W_ri^2 / d_i;
// multiplier: sum L_r^2 / S_r (the reference's choice, GPU/src/equation_gpu.cu:282-315).
// Cells in parallel; the sums run in cell order whatever the thread count.
void build_minv(const Mesh& m, const Setup& S, const Layout& L, const DiagPlan& plan, std::vector<double>& minv);

// Right-hand side in reference order and reference form: flux rows the natural
// boundary term minus the lift, pressure rows the source minus the lift,
// essential boundary rows their values, multiplier row 0.
void build_rhs_ref(const Mesh& m, const Setup& S, std::vector<double>& b);

// Diagonal scaling between the reference-form system (identity essential rows)
// and the internal symmetric one: pressure and multiplier rows -1, essential
// boundary rows exp(-theta_E), other flux rows 1.  Internal order.
void build_row_scale(const Mesh& m, const Setup& S, const Layout& L, std::vector<double>& rs);

}  // namespace vemd
