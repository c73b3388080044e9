// Host symbolic analysis for the multifrontal factorization of the
// 27-point x 3-component elastic operator on an nx x ny x nz grid.
//
// Replaces the ordering + symbolic phase that STRUMPACK runs in job 1
// (strumpack_finter.c:24-66: METIS nested dissection, STRUMPACK_reorder).
// The grid is regular, so the separators are chosen geometrically: each box
// is cut by a one-node-thick plane across its longest axis (a plane of
// nodes separates a 27-point stencil), recursively, down to small leaf
// boxes.  Every front's update set is the one-node halo of its box.
#pragma once
#include <cstdint>
#include <vector>

struct wt_front {
    int32_t level;      // depth in the elimination tree (0 = root)
    int32_t parent;     // -1 at the root
    int32_t child[2];   // -1 when absent
    int32_t slot;       // index of this front among its parent's children
    int64_t p, u;       // pivots, update variables
    int64_t pstart;     // first pivot in the elimination order
    int32_t box[6];     // x0, x1, y0, y1, z0, z1 (node coords, half-open)
    int64_t uoff;       // offset into upd_idx (u entries)
    int64_t moff;       // offset into relmap (u entries): position of each
                        // update variable in the parent's front index set
    int64_t Loff;       // (p+u) x p column block [L11\U11; L21], col-major
    int64_t Uoff;       // p x u block U12, col-major
    int64_t CBoff;      // u x u contribution block in the level-parity arena
    int64_t tileoff;    // ntile * nb * nb inverse-tile storage (x2: L, U)
    int64_t pivoff;     // p pivot entries
    int64_t Woff;       // u x nrhs solve update block (per nrhs unit: u)
};

struct wt_symbolic {
    int nx = 0, ny = 0, nz = 0;
    int nb = 64;                    // diagonal tile width
    int64_t n = 0;                  // unknowns
    int64_t nnz = 0;                // CSR nonzeros
    std::vector<wt_front> fronts;   // post-order (children before parents)
    std::vector<std::vector<int32_t>> levels;  // front ids per depth
    std::vector<int64_t> perm;      // perm[orig] = new position
    std::vector<int64_t> iperm;     // iperm[new] = orig
    std::vector<int64_t> upd_idx;   // concatenated update sets (new indices)
    std::vector<int32_t> relmap;    // concatenated child->parent positions
    std::vector<int64_t> dst;       // per CSR entry: offset in the factor arena
    int64_t factor_len = 0;         // factor arena entries (L parts + U12)
    int64_t cb_len[2] = {0, 0};     // CB arena entries for even / odd levels
    int64_t tile_len = 0;           // inverse-tile entries
    int64_t piv_len = 0;
    int64_t w_len[2] = {0, 0};      // solve update arena (per RHS column)
    int64_t max_p = 0, max_u = 0;
    double flops = 0;               // 8 (p^3/3 + p^2 u + p u^2) summed
};

// Builds the tree, the elimination order, the update sets, the
// child->parent maps and every arena offset.  rowptr/colidx: the CSR
// pattern (0-based) of the grid; dst is filled when they are given.
int wt_symbolic_build(int nx, int ny, int nz, int leaf_nodes, int nb, const int32_t* rowptr,
                      const int32_t* colidx, wt_symbolic& S);

// The CSR pattern of one grid on the host (same as the device pattern
// kernel): rows in y > x > z > component order, columns ascending, with
// row_base added to every row pointer and col_base to every column index.
// rowptr has 3 nx ny nz + 1 entries.
int wt_pattern_host(int nx, int ny, int nz, int32_t row_base, int32_t col_base, int32_t* rowptr,
                    int32_t* colidx);

// Symbolic data of the block-diagonal union of independent systems: the
// forest of their elimination trees, levels aligned by depth from each root,
// unknowns, CSR entries and factor storage concatenated in part order (part
// k's unknowns start at the sum of the previous parts' n).  Every part must
// have been built with its CSR pattern (dst filled) and the same nb.
int wt_symbolic_union(const std::vector<const wt_symbolic*>& parts, wt_symbolic& U);
