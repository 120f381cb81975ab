// gsg_setup.hpp -- FP64 host setup of the sparse-grid DG derivative operator:
// the 1-D multiwavelet basis, the hierarchical 1-D weak-derivative operator with
// central flux, and the sparse-grid layout tables the HIP kernels index with.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gsg {

constexpr int KMAX = 10;  // K_max, 1d_dg_functions.jl:7

// Legendre coefficients from Gram-Schmidt on monomials (dg_basis.jl:68-94):
// (KMAX+1) rows of 2(KMAX+1) coefficients.
const std::vector<double>& legendre_table();
// dg_basis(k) (dg_basis.jl:185-191): k rows of 2k coefficients.
std::vector<double> dg_basis(int k);

// The hierarchical 1-D operator periodic_DLF_matrix(k, n) (1d_derivative.jl:113-117,
// 136-148) restricted to the coupling pattern, stored as k x k blocks grouped by the
// output 1-D cell ("cell-row").  Cell-row r = base(l) + c (level l 0-based, cell c
// 0-based, base(0)=0, base(l)=2^(l-1)); its entries are sorted by input level then
// input cell, and the entries with input level <= m are [row_ptr[r*(n+2)],
// row_ptr[r*(n+2)+m+1]).
struct Op1D {
  int k = 0, n = 0;
  std::vector<int32_t> row_ptr;   // 2^n * (n+2)
  std::vector<int32_t> ent;       // packed (level << 24) | cell
  std::vector<double> val;        // nent * k * k, val[e*k*k + mo*k + mi]
  int64_t nent() const { return (int64_t)ent.size(); }
};
Op1D build_op1d(int k, int n);
// Dense (k 2^n)^2 row-major view of the same operator (zeros off the pattern).
std::vector<double> op1d_dense(const Op1D& op);

// Input cells (0-based, ascending, distinct) of level li coupled to output cell co of
// level lo by the weak derivative with central flux: a pair couples iff a jump point of
// one basis function lies in the closed support of the other.
int coupled_cells(int lo, int co, int li, int out[8]);

// Sparse-grid layout (dg_vmethods.jl:35-45,102-142; schemes.jl:21-27).
struct Layout {
  int D = 0, k = 0, n = 0, full = 0;
  int64_t N = 0;                 // get_size
  int nblk = 0;                  // admissible multi-levels
  std::vector<int32_t> blk_lev;  // nblk * D, 0-based levels
  std::vector<int64_t> blk_off;  // nblk + 1
  std::vector<int32_t> code2blk; // (n+1)^D mixed-radix (dim 1 fastest) -> block or -1
  // Offset of the block L' = L with level d replaced by l, for every block L,
  // direction d and level l (-1 when L' is cut): nblk * D * (n+1).
  std::vector<int64_t> blk_inoff;
};
Layout build_layout(int D, int k, int n, bool full);
int64_t layout_size(int D, int k, int n, bool full);
void layout_v2dref(const Layout& lt, int32_t* levels, int32_t* cells, int32_t* modes);

// Cell-level gather plan of the D directional operators on one sparse grid.  All k^D
// modes of an output cell share, per direction, the same coupled input cells and the
// same 1-D coefficient blocks, so the plan lists, per output cell g (vector order,
// g = DoF index / k^D) and direction d, the pairs (element offset of the coupled input
// cell's first mode, 1-D entry index of its k x k block), in the order of the 1-D row.
// Lists are padded to a multiple of PLAN_UNROLL with entries that point at a zero block
// (PLAN_UNROLL = 1: no padding; the kernels handle any list length).
constexpr int PLAN_UNROLL = 1;
struct CellPlan {
  std::vector<int32_t> ptr;  // ncells * D + 1, in entries
  std::vector<int32_t> ent;  // 2 per entry: (input offset, block index)
  int32_t zero_block = 0;    // index of the all-zero k x k block appended to the 1-D table
};
CellPlan build_cell_plan(const Layout& lt, const Op1D& op);

// Tile plan for the LDS kernel.  A tile is `cpt` consecutive output cells.  Its coupled
// input cells, keyed by (offset, direction), are deduplicated into "slots" that the
// workgroup loads into LDS as contiguous k^D blocks; when a tile needs more than `cap`
// slots, or more coefficient blocks than the LDS holds, its work is split into batches
// that reuse the LDS.  Per (batch, local cell, direction) the plan lists the entries in
// the order of the 1-D row, packed as (block index << 10) | slot, each list padded to a
// multiple of `pad` with (zero block, slot 0) entries (pad = 1: no padding).
struct TilePlan {
  int cpt = 0, cap = 0, ntiles = 0;
  int max_slots = 0;               // largest batch, in slots
  int max_ents = 0;                // largest batch, in entries (coefficient blocks staged)
  std::vector<int32_t> tile_bptr;  // ntiles + 1: batch ranges
  std::vector<int32_t> batch_sptr; // nbatch + 1: slot ranges
  std::vector<int32_t> slot_off;   // input cell offset (elements)
  std::vector<int32_t> slot_dir;   // direction whose input vector holds the slot
  std::vector<int32_t> batch_lptr; // nbatch * cpt * D + 1: entry ranges
  std::vector<int32_t> ent;        // (block << 10) | slot
};
// cpt: cells per tile; lds_elems: LDS elements for input cells per workgroup;
// ent_elems: LDS elements for the coefficient blocks of one batch (entries x k^2);
// max_ents: entries per batch.
TilePlan build_tile_plan(const Layout& lt, const Op1D& op, const CellPlan& cp, int cpt, int lds_elems, int ent_elems,
                         int max_ents, int pad);

// D_matrix(D, d, k, n; scheme) (multidim_derivative.jl:15-65) as a CSC matrix with
// 0-based indices, built from the block table.
struct Csc {
  int64_t m = 0, n = 0;
  std::vector<int64_t> colptr, rowval;
  std::vector<double> nzval;
};
Csc d_matrix_csc(const Layout& lt, const Op1D& op, int d /*0-based*/);

// ---- nodal basis and basis transforms of the Vlasov path (1d_nodal_basis.jl,
// multidim_nodal_basis.jl); k in {2, 3, 5}
enum : int { TB_MODAL = 0, TB_NODAL = 1, TB_POINTS = 2 };
bool nodal_k_ok(int k);
// transform_1D(k, n, from, to) (1d_nodal_basis.jl:383-414) as a dense (k 2^n)^2 row-major
// matrix, entries below 1e-12 dropped as the reference's thresholds do.  Supported:
// modal -> nodal, nodal -> modal, nodal -> points, points -> nodal.
std::vector<double> transform_1d(int k, int n, int from, int to);
// x coordinates of the 1-D collocation points in point-index order (eval_points_1D)
std::vector<double> points_1d(int k, int n);
// The same transform as a 1-D block table (rows: output cells, entries: input cells of
// nested support, ascending level and cell, k x k blocks), with the entries between cells
// that do not nest removed as the reference's D-dimensional assembly does
// (relevant_cell_1D, multidim_nodal_basis.jl:19-25).  The transforms are triangular in
// level, so applying this table along every direction in turn on the sparse grid gives
// exactly the reference's restricted Kronecker product (transform, :83-112) without its
// 1e-12 drop of small products.
Op1D transform_op1d(int k, int n, int from, int to);
// The D-dimensional transform that applying `op` (a transform_op1d table) along every
// direction in turn performs, as a CSC matrix with 0-based indices: the restricted
// Kronecker product of multidim_nodal_basis.jl:83-112 without its 1e-12 drop of small
// products.
Csc transform_csc(const Layout& lt, const Op1D& op);

// ---- projection of initial data (vcoeffs_DG, dg_vmethods.jl:149-179) by a tensor
// Gauss-Legendre rule of q nodes on each half of every 1-D cell (P = 2q nodes per cell,
// point p = s q + j on half s at node j).  nodes: the q Gauss-Legendre nodes on [-1, 1];
// wv: (n+1) x k x P, wv[(l k + m) P + p] = w_p v(k, l, 1, m+1, x_p) for the nodes x_p of
// the first cell of level l.  v of any other cell of the level is the same function
// shifted, so the table serves every cell.
void projection_tables(int k, int n, int q, std::vector<double>& nodes, std::vector<double>& wv);

}  // namespace gsg
