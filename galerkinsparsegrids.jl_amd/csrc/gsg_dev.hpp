// gsg_dev.hpp -- device-side descriptors shared by the HIP kernels and the host
// runtime.  A "solve" is one independent evolution (one operator on one state); a
// "tile" is a contiguous run of output DoFs of one solve.  One launch processes any
// list of tiles, so a single solve and a batch of uneven (D, k, n) solves use the
// same kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gsg {

constexpr int MAXD = 8;
constexpr int TILE = 256;  // DoFs per tile == threads per workgroup

// Geometry of the LDS fiber kernel.  A workgroup owns FIBER_CPT(D, K) consecutive output
// cells; one thread owns one task (direction d, cell c, fiber f): the K outputs of
// cell c whose mode digits other than digit d equal those of fiber f, f < K^(D-1).
constexpr int kpow(int b, int e) { return e <= 0 ? 1 : b * kpow(b, e - 1); }
// the shapes the fiber kernel serves: every thread at most one task, k^D <= 256
constexpr bool fiber_ok(int D, int K) {
  return D >= 1 && D <= 6 && K >= 1 && K <= 5 && kpow(K, D) <= TILE && D * kpow(K, D - 1) <= TILE;
}
// cells per workgroup: as many as keep every thread at one task, rounded down to a
// multiple of 64 / K^(D-1) when that divides 64, so no wavefront spans two directions
constexpr int fiber_cpt(int D, int K) {
  const int F = kpow(K, D - 1);
  const int c = TILE / (D * F);
  const int al = (F < 64 && 64 % F == 0) ? 64 / F : 1;
  return c >= al ? c / al * al : (c > 0 ? c : 1);
}
// doubles an LDS input slot of the FP64 fiber kernel may need for a cell of e doubles
// moved in 16-byte DMAs that start up to one double before it: e + 1 rounded up to even
constexpr int fiber_slot(int e) { return e % 2 ? e + 1 : e + 2; }
// every (cell, direction) pair fills whole wavefronts: its entry list and coefficient
// blocks are wave-uniform and read through the scalar cache
constexpr bool fiber_uniform(int D, int K) { return kpow(K, D - 1) % 64 == 0; }

// the direct fiber kernel: one cell per workgroup (k^D == 256), whole wavefronts per
// (cell, direction) pair, no LDS staging
constexpr bool direct_ok(int D, int K) {
  return fiber_ok(D, K) && kpow(K, D) == TILE && fiber_uniform(D, K);
}

constexpr int DIRECT_CPW = TILE / 64;  // direct kernel: cells (wavefronts) per workgroup
// the FP64 operator also runs the direct kernel at D = 4, k = 3 (one 81-value cell a
// wavefront, 27 fibers on its 64 lanes) and D = 3, k = 4 (64 values, 16 fibers)
// (measured and left out: D = 4, k = 2 runs 20-90% slower than the plan and fiber
// kernels; D = 3, k = 3 gains 9-16% for n >= 5 but loses 18% at n = 4)
constexpr bool direct64_ok(int D, int K) { return direct_ok(D, K) || (D == 4 && K == 3) || (D == 3 && K == 4); }

// The segmented kernel (float, k <= 4, K^(D-1) <= 64): a wavefront owns seg_cpw cells,
// one segment of seg_lanes lanes each (a lane per fiber, padded to a power of two >= 4
// so that every 4-lane MFMA block lies in one cell).
constexpr int pow2_ceil(int x) { return x <= 1 ? 1 : 2 * pow2_ceil((x + 1) / 2); }
constexpr bool seg_ok(int D, int K) { return fiber_ok(D, K) && D <= 4 && K <= 4 && kpow(K, D - 1) <= 64; }
constexpr int seg_lanes(int D, int K) { return kpow(K, D - 1) < 4 ? 4 : pow2_ceil(kpow(K, D - 1)); }
constexpr int seg_cpw(int D, int K) { return 64 / seg_lanes(D, K); }
constexpr int SEG_WAVES = TILE / 64;  // wavefronts per workgroup
// the FP64 segmented kernel (k_seg64): k = 2, D <= 4, several cells a wavefront
constexpr bool seg64_ok(int D, int K) { return seg_ok(D, K) && K == 2 && D >= 2 && seg_cpw(D, K) > 1; }

enum : int32_t {
  KERNEL_PLAN = 0,    // scalar-plan gather kernel (tiles of TILE DoFs)
  KERNEL_FIBER = 1,   // LDS fiber kernel (tile-plan tiles, lds_bytes > 0)
  KERNEL_DIRECT = 2,  // direct fiber kernel (single-cell tiles, DIRECT_CPW per workgroup)
  KERNEL_SEG = 3,     // segmented kernel (single-cell tiles, SEG_WAVES * seg_cpw per workgroup)
  KERNEL_SEG_ALT = 4, // the segmented kernel with the tuned alternative entries per round
};

enum : int32_t {
  SYS_SUM = 0,   // y = sum_d coef[d] D_d x_d   (transport, grad-sum, derivative)
  SYS_WAVE = 1,  // y = [v; sum_d D_d w_d], w_d = D_d u
};

// Operator tables of one (D, k, n, scheme) in device memory (shared by solves).
struct DevTables {
  const int32_t* blk_lev;    // nblk * D, 0-based levels
  const int64_t* blk_off;    // nblk + 1
  const int64_t* blk_inoff;  // nblk * D * (n+1)
  const int32_t* row_ptr;    // 2^n * (n+2)
  const int32_t* ent;        // packed (level << 24) | cell
  const double* val64;       // (nent + 1) * k * k, the last block is zero
  const float* val32;        // 2 (nent + 1) k^2: the hi table, then the lo table at val32 + val_elems
  const int32_t* plan_ptr;   // ncells * D + 1 (cell-level gather plan, entries)
  const int32_t* plan_ent;   // 2 per entry: input cell offset, block index
  // tile plan of the LDS block-gather kernel (k^D <= 256)
  const int32_t* tile_bptr;
  const int32_t* batch_sptr;
  const int32_t* slot_off;
  const int32_t* slot_dir;
  const int32_t* batch_lptr;
  const int32_t* tile_ent;
  int64_t val_elems;         // elements of val64 (and of each half of val32)
  const float* val32p;       // the hi and lo blocks padded to 4 x 4 (k <= 4), (nent + 1) * 16 each,
  int64_t val32p_elems;      // the lo half at val32p + val32p_elems
};

struct SolveDesc {
  DevTables t;
  int32_t D, K, n, full;
  int64_t N;        // DoFs of one component (state length N for SUM, 2N for WAVE)
  int32_t sys;      // SYS_*
  int32_t ndir;     // directions applied in SUM mode (D, or 1 for a single derivative)
  int32_t dir0;     // first direction in SUM mode
  int32_t pad;
  double coef[MAXD];
  double dt;
  // RK4 state (double) and stage buffers (element type = precision of the launch)
  double* y;
  double* acc;
  void* xs[3];  // xs[0]: stage-1 operator input copy (fp32 only); xs[1], xs[2]: ping-pong.
                // FP32 wave: [u; v; u_lo] (3N floats), u_lo = u - u the rounding remainder
  void* w;      // wave: D * N gradient buffer
  int32_t active;  // 0: skip (finished solve in a batch)
  int32_t pad2;
};

struct Tile {
  int32_t solve;
  int32_t blk;    // plan kernel: block containing `start`; fiber kernel: tile index;
                  // direct kernel: the cell's D list lengths, a byte per direction
  int64_t start;  // first DoF (component index, 0..N-1)
  int64_t end;    // one past the last DoF (end <= start: padding tile, does nothing)
  int64_t aux;    // direct kernel: device address of the cell's plan entries
};

enum : int32_t {
  OP_APPLY_SUM = 0,   // yout[i] = sum_d coef[d] (D_d x_d)[i],  x_d = xin + d*instride
  OP_APPLY_GRAD = 1,  // yout[d*N + i] = (D_d xin)[i]
  OP_RK_SUM = 2,      // transport stage `stage`
  OP_RK_WAVE_GRAD = 3,
  OP_RK_WAVE_SUM = 4,
  OP_APPLY_WAVE_SUM = 5,  // yout = [x_v; sum_d D_d w_d] with w = xin2 (D*N)
};

struct LaunchArgs {
  const SolveDesc* solves;
  const Tile* tiles;
  int32_t ntiles;
  int32_t op;
  int32_t stage;
  int32_t kernel;  // KERNEL_*
  const void* xin;   // APPLY ops
  const void* xin2;  // APPLY_WAVE_SUM: the gradient buffer
  void* yout;
  int64_t instride;
  int64_t split;       // float OP_APPLY_GRAD: element offset of the input's lo part (0: none)
  int32_t lds_bytes;   // > 0: tiles are tile-plan tiles, run the LDS kernel with this LDS
  int32_t lds_xelems;  // LDS elements reserved for input cells (then the per-direction
                       // results); the coefficient blocks follow
  int32_t lds_ments;   // entries per batch the coefficient and entry staging hold
  int32_t lds_mslots;  // slots per batch the offset staging holds
};

// Launch one operation over the given tiles for solves of shape (D, K); fp32 selects
// the float operator.  Returns hipSuccess or the launch error.
hipError_t launch_op(int D, int K, bool fp32, const LaunchArgs& a, hipStream_t s);

// ---- one launch for every class of a batch (FP32 RK stages) ----------------------
// The fused stage kernel runs, per workgroup item, the body of the kernel a class would
// launch on its own (the direct or segmented kernel of its (D, k), on the class's
// tiles), so a batch's stage (or wave pass) is one launch whatever its classes.  Bodies
// exist for the direct and segmented shapes (D, k in {2, 3, 4}); a class that runs
// another kernel (the plan kernel where it is tuned faster) keeps its own launch.
constexpr int FUSED_MAXSEG = 32;
// body id of a class: kernel (KERNEL_PLAN / KERNEL_SEG / KERNEL_DIRECT), D, k
constexpr int fused_body(int kernel, int D, int K) { return (kernel * 3 + (D - 2)) * 3 + (K - 2); }
inline bool fused_ok(int kernel, int D, int K) {
  if (D < 2 || D > 4 || K < 2 || K > 4) return false;
  if (kernel == KERNEL_DIRECT) return D == 4 && K == 4;
  return kernel == KERNEL_SEG && seg_ok(D, K) && seg_cpw(D, K) > 1;
}
struct FusedSeg {
  const Tile* tiles;  // the class's tiles (as its own launch would take them)
  int32_t ntiles;
  int32_t wg_end;     // one past the class's last item in the launch (items are workgroups)
  int16_t body;       // fused_body(...)
  int16_t op;         // OP_RK_SUM, OP_RK_WAVE_GRAD or OP_RK_WAVE_SUM
  int32_t pad;
};
struct FusedArgs {
  const SolveDesc* solves;
  const int32_t* order;  // item -> (segment << 24) | workgroup of the segment; null: the
                         // segments' items one after another (seg[s].wg_end)
  int32_t nseg;
  int32_t items;      // workgroup items of the launch (seg[nseg - 1].wg_end)
  FusedSeg seg[FUSED_MAXSEG];
};
// pass: FUSED_SUM (RK_SUM bodies only), FUSED_A (RK_SUM and RK_WAVE_GRAD), FUSED_B
// (RK_WAVE_SUM); grid workgroups run items grid, grid + wg, ... (grid <= items; grid ==
// items: one item a workgroup)
enum : int32_t { FUSED_SUM = 0, FUSED_A = 1, FUSED_B = 2 };
hipError_t launch_fused(int pass, int stage, const FusedArgs& f, int grid, hipStream_t s);

// Vlasov pointwise combination fused into the last n2p sweep (pdes.jl:182-191): vectors
// v < Dx are scaled by -v_point[v], vectors v >= Dx by F_point[v - Dx] (F may be null:
// no such vectors), summed as -(sum_d v_d P_d) + (sum_d F_d Q_d)
struct XComb {
  const double* v;
  const double* F;
  int32_t Dx;
  int32_t nv;
};
// The classical RK4 stage update fused into the last sweep of a right-hand side (one
// vector, no combination): the sweep result r at DoF i updates y, acc and the next stage
// input x (and its float copy x32 when not null); stage 0: none.
struct XRk {
  int32_t stage;
  double dt;
  double* y;
  double* acc;
  double* x;
  float* x32;
};
// entries a row of a sweep's fixed-width table (k = 2): the longest list a one-lane-per-DoF
// tile takes
constexpr int XELL = 8;
// one transform direction sweep over nvec vectors (k_xsweep; K in {2, 3, 5}); in_f32: the
// inputs are float; comb: write the combination of the nvec results (one vector); rk: the
// RK stage update instead of the result (one vector); ell (K = 2 only, or null): the sweep
// direction's fixed-width table, XELL (input cell offset, block) pairs a cell, the cell's
// list first and block -1 after it, for every cell whose list has at most XELL entries
hipError_t launch_xsweep(int K, int nvec, const LaunchArgs& a, hipStream_t s, bool in_f32 = false,
                         const XComb* comb = nullptr, const XRk* rk = nullptr, const int32_t* ell = nullptr);

// *dst = v, written by a one-thread kernel (stream-ordered; the value travels as a kernel
// argument, so the caller's copy may change as soon as this returns)
hipError_t launch_write_desc(SolveDesc* dst, const SolveDesc& v, hipStream_t s);

// y[i] = (float) x[i] for i < n, and the rounding remainders y[n + i] = (float)(x[i] -
// y[i]) for i < nlo (the lo part of the wave system's u, read by its gradient pass)
hipError_t launch_to_f32(const double* x, float* y, int64_t n, hipStream_t s, int64_t nlo = 0);
// y[i] = (double) x[i] for i < n
hipError_t launch_to_f64(const float* x, double* y, int64_t n, hipStream_t s);

// Projection of initial data onto the sparse grid (vcoeffs_DG): per-cell tables in
// device memory.  P = 2 q Gauss-Legendre points per cell and direction, ppc = P^D.
struct ProjArgs {
  const int32_t* cell_blk;   // per cell (vector order): its block
  const int32_t* blk_cell0;  // per block: its first cell
  const int32_t* blk_lev;    // nblk * D, 0-based levels
  int32_t D, K, q, P;
  int64_t ppc;               // points per cell (< 2^31)
};
// X[d ld + i]: coordinates of the tensor nodes of cells [cell0, cell0 + ncell), point
// index direction 1 fastest; t: the q Gauss-Legendre nodes on [-1, 1]
hipError_t launch_proj_nodes(const ProjArgs& a, const double* t, int64_t cell0, int64_t ncell, double* X, int64_t ld,
                             hipStream_t s);
// X1[(d ncell + c) P + p]: the direction-d coordinate of node p of cell cell0 + c
hipError_t launch_proj_nodes_1d(const ProjArgs& a, const double* t, int64_t cell0, int64_t ncell, double* X1,
                                hipStream_t s);
// tensor_construct on the device: out[i] = prod_d v[d][1-D index of DoF i in direction
// d] (the product in direction order from 1.0); v[d] has k 2^n entries
struct TensorVecs {
  const double* v[MAXD];
};
hipError_t launch_tensor(const ProjArgs& a, const TensorVecs& tv, int64_t N, double* out, hipStream_t s);
// contraction pass for direction d: per cell [P][R] -> [R][K]; wv: (n+1) x K x P
// weighted basis values
hipError_t launch_proj_pass(const ProjArgs& a, const double* wv, int d, int64_t cell0, int64_t ncell, int64_t R,
                            const double* in, double* out, hipStream_t s);

// CSR SpMV y = A x (rows ascending columns), value/vector type per precision.
struct DevCsr {
  int64_t m, n, nnz;
  const int64_t* rp;
  const int32_t* ci;
  const double* v64;
  const float* v32;  // 2 nnz: hi values, then lo values (v64 - hi)
};
hipError_t launch_csr(const DevCsr& A, bool fp32, const void* x, void* y, hipStream_t s);

}  // namespace gsg
