// gsg_kernels.hip -- gfx950 kernels of the sparse-grid DG derivative operator and
// its fused classical-RK4 stages.
//
// One thread owns one output DoF (output cell g, mode m') of one solve.  For every
// direction d the row of D_d restricted to the sparse grid is
//     (D_d x)[g, m'] = sum_{(off, e) in plan(g, d)} sum_{m_d} B_e[m'_d, m_d] x[off + mode(m' with digit d = m_d)]
// where plan(g, d) lists the coupled input cells of the output cell (the 1-D row of the
// output's coordinate along d, cut to the fiber's admissible levels) and B_e is the
// k x k block of the 1-D hierarchical operator.  The plan is shared by all k^D modes of
// a cell, so all index arithmetic happens once per cell on the host.
//
// When k^D is a multiple of 64 a wavefront covers 64 modes of one cell: the plan
// entries are wave-uniform, read with scalar loads into SGPRs, and the input values
// are buffer loads with the cell offset in the SGPR soffset and the lane's mode offset
// in a 32-bit VGPR.  Entries are consumed four at a time (lists are padded to a
// multiple of four with zero blocks) so each wave keeps 4k loads in flight.
//
// The RK4 stage update is fused into the epilogue: one transport stage is one launch
// that reads the stage input, writes the next stage input and the accumulator.
#include <hip/hip_runtime.h>

#include "gsg_dev.hpp"

namespace gsg {

namespace {

constexpr int ipow_c(int b, int e) { return e == 0 ? 1 : b * ipow_c(b, e - 1); }
__device__ __forceinline__ int ipow_rt(int b, int e) {
  int r = 1;
  for (int t = 0; t < e; ++t) r *= b;
  return r;
}

typedef __attribute__((address_space(4))) const int32_t c_i32;
typedef __attribute__((address_space(4))) const int64_t c_i64;

// scalar (SMEM) loads of read-only, wave-uniform data
__device__ __forceinline__ int32_t sld(const int32_t* p) { return *(c_i32*)(p); }
__device__ __forceinline__ int64_t sld(const int64_t* p) { return *(c_i64*)(p); }
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }

constexpr int BUF_DWORD3 = 0x00020000;  // raw buffer resource word 3, gfx9 family

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, BUF_DWORD3);
}

template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}

template <typename T>
__device__ __forceinline__ const T* vals_of(const DevTables& t);
template <>
__device__ __forceinline__ const double* vals_of<double>(const DevTables& t) {
  return t.val64;
}
template <>
__device__ __forceinline__ const float* vals_of<float>(const DevTables& t) {
  return t.val32;
}

template <typename T>
__device__ __forceinline__ const T* vals_of_csr(const DevCsr& A);
template <>
__device__ __forceinline__ const double* vals_of_csr<double>(const DevCsr& A) {
  return A.v64;
}
template <>
__device__ __forceinline__ const float* vals_of_csr<float>(const DevCsr& A) {
  return A.v32;
}

constexpr int U = 4;  // plan entries consumed per iteration of the main loop

// (D_d x)[g, m'] for one lane.  UNI: the plan entries are wave-uniform.  xr covers the
// input vector(s); xbase is the byte offset of direction d's input inside xr; md is the
// lane's mode digit d and mb its mode offset with digit d cleared; kd = K^d.  TX: the
// input's element type (converted to T before the product).  For T = float the entries
// come from the hi table into acc and from the lo table (at vals + val_elems) into acl.
template <int KT, typename T, bool UNI, int NU, typename TX = T>
__device__ __forceinline__ void plan_entries(const SolveDesc& S, __amdgpu_buffer_rsrc_t xr, uint32_t xbase, int32_t p,
                                             int K, int md, uint32_t mb, int kd, T& acc, T& acl) {
  constexpr int KM = KT ? KT : 16;
  constexpr bool HL = sizeof(T) == 4;
  const T* vals = vals_of<T>(S.t);
  constexpr uint32_t es = sizeof(TX);
  int32_t off[NU], eb[NU];
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    if constexpr (UNI) {
      off[u] = uni(sld(S.t.plan_ent + 2 * (p + u)));
      eb[u] = uni(sld(S.t.plan_ent + 2 * (p + u) + 1));
    } else {
      const int2 pe = *reinterpret_cast<const int2*>(S.t.plan_ent + 2 * (p + u));
      off[u] = pe.x;
      eb[u] = pe.y;
    }
  }
  T b[NU][KM], bl[NU][KM];
  TX xv[NU][KM];
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const T* vp = vals + (int64_t(eb[u]) * K + md) * K;
#pragma unroll
    for (int mm = 0; mm < KM; ++mm) {
      if (mm >= K) break;
      b[u][mm] = vp[mm];
      if constexpr (HL) bl[u][mm] = vp[S.t.val_elems + mm];
    }
  }
#pragma unroll
  for (int u = 0; u < NU; ++u) {
#pragma unroll
    for (int mm = 0; mm < KM; ++mm) {
      if (mm >= K) break;
      if constexpr (UNI)
        xv[u][mm] = buf_load<TX>(xr, (mb + uint32_t(mm * kd)) * es, uint32_t(off[u]) * es + xbase);
      else
        xv[u][mm] = buf_load<TX>(xr, (uint32_t(off[u]) + mb + uint32_t(mm * kd)) * es, xbase);
    }
  }
#pragma unroll
  for (int u = 0; u < NU; ++u) {
#pragma unroll
    for (int mm = 0; mm < KM; ++mm) {
      if (mm >= K) break;
      acc += b[u][mm] * T(xv[u][mm]);
      if constexpr (HL) acl += bl[u][mm] * T(xv[u][mm]);
    }
  }
}

// the direction-d sum of one lane; for T = float the hi and lo sums added in double
template <int KT, typename T, bool UNI, typename TX = T>
__device__ __forceinline__ double plan_dir(const SolveDesc& S, __amdgpu_buffer_rsrc_t xr, uint32_t xbase, int32_t g,
                                           int d, int D, int K, int md, uint32_t mb, int kd) {
  const int32_t* ptr = S.t.plan_ptr + int64_t(g) * D + d;
  int32_t p0, p1;
  if constexpr (UNI) {
    p0 = uni(sld(ptr));
    p1 = uni(sld(ptr + 1));
  } else {
    p0 = ptr[0];
    p1 = ptr[1];
  }
  T acc = T(0), acl = T(0);
  int32_t p = p0;
  for (; p + U <= p1; p += U) plan_entries<KT, T, UNI, U, TX>(S, xr, xbase, p, K, md, mb, kd, acc, acl);
  for (; p < p1; ++p) plan_entries<KT, T, UNI, 1, TX>(S, xr, xbase, p, K, md, mb, kd, acc, acl);
  if constexpr (sizeof(T) == 4) return double(acc) + double(acl);
  return double(acc);
}

template <typename T>
__device__ __forceinline__ const T* stage_in(const SolveDesc& S, int stage) {
  if (stage == 1) return sizeof(T) == 4 ? (const T*)S.xs[0] : (const T*)S.y;
  return (const T*)(stage == 3 ? S.xs[2] : S.xs[1]);
}

template <typename T>
__device__ __forceinline__ T* stage_out(const SolveDesc& S, int stage) {
  if (stage == 4) return sizeof(T) == 4 ? (T*)S.xs[0] : nullptr;
  return (T*)(stage == 2 ? S.xs[2] : S.xs[1]);
}

// Classical RK4 stage update at state index idx with stage derivative r:
//   acc = y + dt/6 k1; acc += dt/3 k2; acc += dt/3 k3; y = acc + dt/6 k4
//   stage inputs y + dt/2 k1, y + dt/2 k2, y + dt k3.
// rk_pre loads the state the update reads (so a kernel can issue it early), rk_post
// writes the update.
struct RkIn {
  double y, acc;
};
__device__ __forceinline__ RkIn rk_pre(const SolveDesc& S, int stage, int64_t idx) {
  RkIn v;
  v.y = stage <= 3 ? S.y[idx] : 0.0;
  v.acc = stage >= 2 ? S.acc[idx] : 0.0;
  return v;
}
// the next stage input x at out[idx] (and, with lo >= 0 in float, its rounding remainder
// at out[lo])
template <typename T>
__device__ __forceinline__ void put_stage(T* out, int64_t idx, int64_t lo, double x) {
  const T h = T(x);
  out[idx] = h;
  if (sizeof(T) == 4 && lo >= 0) out[lo] = T(x - double(h));
}
template <typename T>
__device__ __forceinline__ void rk_post(const SolveDesc& S, int stage, int64_t idx, const RkIn& v, double r,
                                        int64_t lo = -1) {
  const double dt = S.dt;
  T* out = stage_out<T>(S, stage);
  if (stage == 1) {
    S.acc[idx] = v.y + (dt / 6.0) * r;
    put_stage(out, idx, lo, v.y + (dt / 2.0) * r);
  } else if (stage == 2) {
    S.acc[idx] = v.acc + (dt / 3.0) * r;
    put_stage(out, idx, lo, v.y + (dt / 2.0) * r);
  } else if (stage == 3) {
    S.acc[idx] = v.acc + (dt / 3.0) * r;
    put_stage(out, idx, lo, v.y + dt * r);
  } else {
    const double yn = v.acc + (dt / 6.0) * r;
    S.y[idx] = yn;
    if (out) put_stage(out, idx, lo, yn);
  }
}
template <typename T>
__device__ __forceinline__ void rk_update(const SolveDesc& S, int stage, int64_t idx, double r, int64_t lo = -1) {
  rk_post<T>(S, stage, idx, rk_pre(S, stage, idx), r, lo);
}

// DT/KT = 0: D / k read at run time (shapes without a dedicated instantiation).
// wg: the workgroup's tile (blockIdx.x of k_plan, an item of the fused kernel).
template <int DT, int KT, typename T>
__device__ __forceinline__ void plan_body(const LaunchArgs& a, int wg) {
  constexpr bool UNI = DT > 0 && KT > 0 && ipow_c(KT, DT) % 64 == 0;
  const Tile* tp = a.tiles + wg;
  const int32_t solve = uni(sld(&tp->solve));
  const SolveDesc& S = a.solves[solve];
  if (!sld(&S.active)) return;
  const int D = DT ? DT : S.D;
  const int K = KT ? KT : S.K;
  const int KD = (DT && KT) ? ipow_c(KT, DT) : ipow_rt(K, D);
  const int64_t start = sld(&tp->start);
  const int64_t end = sld(&tp->end);
  int64_t i;
  int32_t g, ml;
  if constexpr (UNI) {
    const int wave = uni(int(threadIdx.x) >> 6);
    const int64_t i0 = start + wave * 64;
    if (i0 >= end) return;  // tiles hold whole cells, so a wave is all in or all out
    g = uni(int32_t(i0 / KD));
    ml = int32_t(i0 - int64_t(g) * KD) + (int(threadIdx.x) & 63);
    i = i0 + (int(threadIdx.x) & 63);
  } else {
    i = start + threadIdx.x;
    if (i >= end) return;
    g = int32_t(i / KD);
    ml = int32_t(i - int64_t(g) * KD);
  }
  const int64_t N = S.N;
  constexpr int DM = DT ? DT : MAXD;

  if (a.op == OP_APPLY_SUM || a.op == OP_RK_SUM) {
    const T* x0 = a.op == OP_RK_SUM ? stage_in<T>(S, a.stage) : (const T*)a.xin;
    const int64_t stride = a.op == OP_RK_SUM ? 0 : a.instride;
    const int64_t elems = stride ? stride * S.ndir : N;
    const __amdgpu_buffer_rsrc_t xr = make_rsrc(x0, uint32_t(elems * int64_t(sizeof(T))));
    double s = 0.0;
    int t = 0;
    int kd = 1;
#pragma unroll
    for (int d = 0; d < DM; ++d) {
      if (d >= D) break;
      if (d >= S.dir0 && d < S.dir0 + S.ndir) {
        const int md = (ml / kd) % K;
        const uint32_t mb = uint32_t(ml - md * kd);
        s += S.coef[d] * plan_dir<KT, T, UNI>(S, xr, uint32_t(t * stride * int64_t(sizeof(T))), g, d, D, K, md, mb, kd);
        ++t;
      }
      kd *= K;
    }
    if (a.op == OP_APPLY_SUM)
      ((double*)a.yout)[i] = s;
    else
      rk_update<T>(S, a.stage, i, s);
    return;
  }
  if (a.op == OP_APPLY_GRAD || a.op == OP_RK_WAVE_GRAD) {
    const T* x0 = a.op == OP_RK_WAVE_GRAD ? stage_in<T>(S, a.stage) : (const T*)a.xin;
    T* w = a.op == OP_RK_WAVE_GRAD ? (T*)S.w : (T*)a.yout;
    // float: the input's rounding remainder (u_lo at 2N of the wave stage input, at
    // a.split of an applied input) goes through the operator as well
    const int64_t lo = sizeof(T) == 4 ? (a.op == OP_RK_WAVE_GRAD ? 2 * N : a.split) : 0;
    const __amdgpu_buffer_rsrc_t xr = make_rsrc(x0, uint32_t((lo ? lo + N : N) * int64_t(sizeof(T))));
    int kd = 1;
#pragma unroll
    for (int d = 0; d < DM; ++d) {
      if (d >= D) break;
      const int md = (ml / kd) % K;
      const uint32_t mb = uint32_t(ml - md * kd);
      double r = plan_dir<KT, T, UNI>(S, xr, 0u, g, d, D, K, md, mb, kd);
      if (lo) r += plan_dir<KT, T, UNI>(S, xr, uint32_t(lo * int64_t(sizeof(T))), g, d, D, K, md, mb, kd);
      w[int64_t(d) * N + i] = T(r);
      kd *= K;
    }
    return;
  }
  // wave second pass: [u; v]' = [v; sum_d D_d w_d]
  const T* w = a.op == OP_RK_WAVE_SUM ? (const T*)S.w : (const T*)a.xin2;
  const __amdgpu_buffer_rsrc_t wr = make_rsrc(w, uint32_t(int64_t(D) * N * int64_t(sizeof(T))));
  double s = 0.0;
  int kd = 1;
#pragma unroll
  for (int d = 0; d < DM; ++d) {
    if (d >= D) break;
    const int md = (ml / kd) % K;
    const uint32_t mb = uint32_t(ml - md * kd);
    s += plan_dir<KT, T, UNI>(S, wr, uint32_t(int64_t(d) * N * int64_t(sizeof(T))), g, d, D, K, md, mb, kd);
    kd *= K;
  }
  if (a.op == OP_APPLY_WAVE_SUM) {
    const T* x = (const T*)a.xin;
    ((double*)a.yout)[i] = double(x[N + i]);
    ((double*)a.yout)[N + i] = s;
  } else {
    const T* xin = stage_in<T>(S, a.stage);
    rk_update<T>(S, a.stage, i, double(xin[N + i]), 2 * N + i);  // float: u's lo part too
    rk_update<T>(S, a.stage, N + i, s);
  }
}

template <int DT, int KT, typename T>
__global__ __launch_bounds__(TILE) void k_plan(LaunchArgs a) {
  plan_body<DT, KT, T>(a, int(blockIdx.x));
}


// ---------------------------------------------------------------------------------
// LDS fiber kernel (fiber_ok shapes).  A workgroup owns fiber_cpt(D, K) consecutive
// output cells.  For each batch of its tile plan it DMAs the batch's deduplicated input
// cells (contiguous k^D blocks) and, unless they are wave-uniform, the batch's k x k
// coefficient blocks straight into LDS (gfx950 buffer_load ... lds, 16 bytes a lane
// where the sizes allow), all in flight at once and without VGPRs.  Each thread owns
// one task (direction d, cell c, fiber f): the K outputs of cell c along mode digit d.
// Per entry of its (c, d) list it reads the K inputs of the fiber from LDS and does K^2
// FMAs with the entry's block, so every LDS read feeds K FMAs.  When the (c, d) pair
// fills whole wavefronts (K^(D-1) % 64 == 0) the list and the blocks are read through
// the scalar cache.  The per-direction results meet in LDS (over the dead input cells)
// and the fused epilogue runs one thread per output DoF.  The accumulation order is
// that of the 1-D row, entry after entry, mode after mode.
typedef __attribute__((address_space(3))) void* lds_ptr_t;

template <int VEC>
__device__ __forceinline__ void dma_to_lds(__amdgpu_buffer_rsrc_t r, uint32_t* lds_chunk, uint32_t voff) {
  if constexpr (VEC == 4)
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_ptr_t)lds_chunk, 16, (int)voff, 0, 0, 0);
  else
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_ptr_t)lds_chunk, 4, (int)voff, 0, 0, 0);
}

typedef __attribute__((address_space(4))) const float c_f32;
typedef __attribute__((address_space(4))) const double c_f64;
__device__ __forceinline__ float sld(const float* p) { return *(c_f32*)(p); }
__device__ __forceinline__ double sld(const double* p) { return *(c_f64*)(p); }

template <int DT, int KT, typename T>
__global__ __launch_bounds__(TILE) void k_fiber(LaunchArgs a) {
  constexpr int D = DT, K = KT;
  constexpr int KD = ipow_c(K, D);
  constexpr int F = KD / K;
  constexpr int KK = K * K;
  constexpr int CPT = fiber_cpt(D, K);
  constexpr int PD = CPT * F;  // tasks per direction
  constexpr bool UNIB = fiber_uniform(D, K);
  constexpr int DW = sizeof(T) / 4;                 // dwords per element
  static_assert(DW == 2, "the fiber kernel is FP64 only");
  // Input cells and coefficient blocks move 16 bytes (two doubles) a lane.  A cell of
  // odd k^D (a block of odd k^2) need not start on 16 bytes: its DMA starts at the
  // 16-byte boundary at or below it and covers KDS (KKS) doubles, so an LDS slot holds
  // KDS doubles and the reads skip the cell's shift there (0 or 1 double).
  constexpr int KKS = KK % 2 ? KK + 1 : KK;  // doubles per LDS coefficient block (the table is aligned)
  constexpr int BCH = KKS / 2;               // 16-byte chunks per block
  constexpr int NW = TILE / 64;
  static_assert(fiber_ok(D, K), "shape not served by the fiber kernel");
  extern __shared__ __align__(16) unsigned char smem[];
  // LDS: [input cells, later the per-direction results: lds_xelems]
  //      [coefficient blocks: lds_ments * k^2, absent when wave-uniform]
  //      [slot offsets: lds_mslots][entry words: lds_ments][list bounds: CPT * D + 1]
  T* X = reinterpret_cast<T*>(smem);
  T* Bl = X + a.lds_xelems;
  uint32_t* soff = reinterpret_cast<uint32_t*>(Bl + (UNIB ? 0 : a.lds_ments * KKS));
  int32_t* entl = reinterpret_cast<int32_t*>(soff + a.lds_mslots);
  int32_t* lptl = entl + a.lds_ments;
  const Tile* tp = a.tiles + blockIdx.x;
  const int64_t start = sld(&tp->start);
  const int64_t end = sld(&tp->end);
  if (end <= start) return;  // padding tile (whole workgroup)
  const int32_t solve = uni(sld(&tp->solve));
  const int32_t t = uni(sld(&tp->blk));
  const SolveDesc& S = a.solves[solve];
  if (!sld(&S.active)) return;
  const int64_t N = S.N;
  const int ncl = int((end - start) / KD);
  const int tid = threadIdx.x;
  const int wave = uni(tid >> 6);
  const int lane = tid & 63;
  // task of this thread
  int d = tid / PD;
  int c = (tid - d * PD) / F;
  const int f = tid - d * PD - c * F;
  if constexpr (UNIB) {
    d = uni(d);
    c = uni(c);
  }
  const bool act = d < D && c < ncl;
  int sd = 1;  // K^d: stride of mode digit d
#pragma unroll
  for (int e = 0; e + 1 < D; ++e)
    if (e < d) sd *= K;
  const int pos0 = f % sd + (f / sd) * sd * K;  // the fiber's element with digit d = 0

  const T* x0;
  int64_t dstride;
  if (a.op == OP_RK_SUM) {
    x0 = stage_in<T>(S, a.stage);
    dstride = 0;
  } else if (a.op == OP_APPLY_SUM) {
    x0 = (const T*)a.xin;
    dstride = a.instride;
  } else if (a.op == OP_APPLY_GRAD || a.op == OP_RK_WAVE_GRAD) {
    x0 = a.op == OP_RK_WAVE_GRAD ? stage_in<T>(S, a.stage) : (const T*)a.xin;
    dstride = 0;
  } else {
    x0 = a.op == OP_RK_WAVE_SUM ? (const T*)S.w : (const T*)a.xin2;
    dstride = N;
  }
  const __amdgpu_buffer_rsrc_t xr = make_rsrc(x0, uint32_t((dstride ? dstride * D : N) * int64_t(sizeof(T))));
  // parity of the input's base in doubles: a slot's cell starts on 16 bytes iff its
  // element offset plus this is even (the tables are device allocations: 256-byte aligned)
  const uint32_t xpar = uint32_t(reinterpret_cast<uintptr_t>(x0) >> 3) & 1u;
  // doubles per LDS input slot: k^D rounded up to even, and 2 more for an even k^D on an
  // input that does not start on 16 bytes (at most fiber_slot(k^D), what the host sized)
  const int KDS = KD % 2 ? KD + 1 : KD + 2 * int(xpar);
  const int XCH = KDS / 2;  // 16-byte chunks per slot
  const T* vals = vals_of<T>(S.t);
  const __amdgpu_buffer_rsrc_t vr = make_rsrc(vals, uint32_t(S.t.val_elems * int64_t(sizeof(T))));

  T acc[K];
#pragma unroll
  for (int i = 0; i < K; ++i) acc[i] = T(0);

  uint32_t* Xdw = reinterpret_cast<uint32_t*>(X);
  uint32_t* Bdw = reinterpret_cast<uint32_t*>(Bl);
  const int32_t b0 = uni(sld(S.t.tile_bptr + t));
  const int32_t b1 = uni(sld(S.t.tile_bptr + t + 1));
  for (int32_t b = b0; b < b1; ++b) {
    const int32_t s0 = uni(sld(S.t.batch_sptr + b));
    const int32_t ns = uni(sld(S.t.batch_sptr + b + 1)) - s0;
    const int32_t nxc = ns * XCH;  // 16-byte chunks of the batch's input cells
    const int32_t* lb = S.t.batch_lptr + int64_t(b) * CPT * D;
    const int32_t E0 = uni(sld(lb));
    const int32_t ne = uni(sld(lb + CPT * D)) - E0;
    for (int32_t q = tid; q < ns; q += TILE)
      soff[q] = uint32_t(S.t.slot_off[s0 + q]) + uint32_t(S.t.slot_dir[s0 + q] * dstride);
    for (int32_t q = tid; q < ne; q += TILE) entl[q] = S.t.tile_ent[E0 + q];
    for (int32_t q = tid; q <= CPT * D; q += TILE) lptl[q] = lb[q] - E0;
    __syncthreads();
    // input cells: chunk w of the batch -> chunk w % XCH of slot w / XCH, from the
    // 16-byte boundary at or below the slot's cell
    for (int32_t j = wave; j * 64 < nxc; j += NW) {
      const int32_t w = j * 64 + lane;
      if (w < nxc) {
        const int32_t sl = w / XCH;
        const uint32_t e0 = soff[sl] - ((soff[sl] + xpar) & 1u);
        dma_to_lds<4>(xr, Xdw + j * 64 * 4, (e0 + 2u * uint32_t(w - sl * XCH)) * 8u);
      }
    }
    if constexpr (!UNIB) {
      // coefficient blocks of the batch's entries, in entry order, KKS doubles apart
      const int32_t nbc = ne * BCH;
      for (int32_t j = wave; j * 64 < nbc; j += NW) {
        const int32_t w = j * 64 + lane;
        if (w < nbc) {
          const int32_t q = w / BCH;
          const uint32_t b0 = uint32_t(entl[q] >> 10) * KK;
          dma_to_lds<4>(vr, Bdw + j * 64 * 4, ((b0 & ~1u) + 2u * uint32_t(w - q * BCH)) * 8u);
        }
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (act) {
      int32_t p0 = lptl[c * D + d];
      int32_t p1 = lptl[c * D + d + 1];
      if constexpr (UNIB) {
        p0 = uni(p0);
        p1 = uni(p1);
      }
      for (int32_t p = p0; p < p1; ++p) {
        int32_t ew = entl[p];
        if constexpr (UNIB) ew = uni(ew);
        const int32_t sl = ew & 1023;
        const T* xs = X + sl * KDS + ((soff[sl] + xpar) & 1u) + pos0;
        T xv[K];
#pragma unroll
        for (int j = 0; j < K; ++j) xv[j] = xs[j * sd];
        if constexpr (UNIB) {
          const T* bp = vals + int64_t(ew >> 10) * KK;
          T bv[KK];
#pragma unroll
          for (int q = 0; q < KK; ++q) bv[q] = sld(bp + q);
#pragma unroll
          for (int i = 0; i < K; ++i)
#pragma unroll
            for (int j = 0; j < K; ++j) acc[i] += bv[i * K + j] * xv[j];
        } else {
          const T* bp = Bl + p * KKS + ((uint32_t(ew >> 10) * KK) & 1u);
#pragma unroll
          for (int i = 0; i < K; ++i)
#pragma unroll
            for (int j = 0; j < K; ++j) acc[i] += bp[i * K + j] * xv[j];
        }
      }
    }
    __syncthreads();
  }
  // per-direction results over the dead input cells: Y[(c * D + d) * KD + mode]
  T* Y = X;
  if (act) {
#pragma unroll
    for (int i = 0; i < K; ++i) Y[(c * D + d) * KD + pos0 + i * sd] = acc[i];
  }
  __syncthreads();

  for (int q = tid; q < ncl * KD; q += TILE) {
    const int cq = q / KD;
    const T* yq = Y + cq * D * KD + (q - cq * KD);
    const int64_t i = start + q;
    if (a.op == OP_APPLY_SUM || a.op == OP_RK_SUM) {
      T s = T(0);
#pragma unroll
      for (int e = 0; e < D; ++e) s += T(S.coef[e]) * yq[e * KD];
      if (a.op == OP_APPLY_SUM)
        ((double*)a.yout)[i] = double(s);
      else
        rk_update<T>(S, a.stage, i, double(s));
    } else if (a.op == OP_APPLY_GRAD || a.op == OP_RK_WAVE_GRAD) {
      T* w = a.op == OP_RK_WAVE_GRAD ? (T*)S.w : (T*)a.yout;
#pragma unroll
      for (int e = 0; e < D; ++e) w[int64_t(e) * N + i] = yq[e * KD];
    } else {
      T s = T(0);
#pragma unroll
      for (int e = 0; e < D; ++e) s += yq[e * KD];
      if (a.op == OP_APPLY_WAVE_SUM) {
        const T* x = (const T*)a.xin;
        ((double*)a.yout)[i] = double(x[N + i]);
        ((double*)a.yout)[N + i] = double(s);
      } else {
        const T* xin = stage_in<T>(S, a.stage);
        rk_update<T>(S, a.stage, i, double(xin[N + i]));
        rk_update<T>(S, a.stage, N + i, double(s));
      }
    }
  }
}

// ---------------------------------------------------------------------------------
// Direct kernel (direct_ok shapes: k^D == 256 and K^(D-1) a multiple of 64, i.e.
// D = 4, k = 4; and the FP64 operator at D = 4, k = 3, direct64_ok).  One wavefront per
// output cell, all D directions; the tile carries the address of the cell's plan
// entries and the D list lengths (details at k_direct).
#ifndef GSG_DIRECT_EB
#define GSG_DIRECT_EB 4
#endif
constexpr int DIRECT_EB = GSG_DIRECT_EB;  // entries per load round (fp32); the entry array has 16 spare words
static_assert(DIRECT_EB >= 2 && DIRECT_EB <= 8, "the entry prefetch reads at most 16 words past a list");

// v, or +0 when !keep (a bit mask: no branch, and NaN or Inf bits are cleared too)
__device__ __forceinline__ float keep_if(float v, bool keep) {
  return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v) & (keep ? 0xffffffffu : 0u));
}
__device__ __forceinline__ double keep_if(double v, bool keep) {
  return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, v) & (keep ? ~uint64_t(0) : uint64_t(0)));
}

// v_mfma_f32_4x4x1_16b_f32 with A broadcast: in every group of 2^CB blocks, each block
// takes A from the group's block `blk` (CBSZ CB, ABID blk; CB = 4: from lanes 4 blk ..
// 4 blk + 3 for the whole wavefront).  blk is a constant after unrolling.
typedef float v4f_t __attribute__((ext_vector_type(4)));
template <int CB>
__device__ __forceinline__ v4f_t mfma_abid(float a, float b, v4f_t c, int blk) {
#define GSG_ABID(X) \
  case X: return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, CB, X, 0);
  switch (blk) {
    GSG_ABID(0) GSG_ABID(1) GSG_ABID(2) GSG_ABID(3) GSG_ABID(4) GSG_ABID(5) GSG_ABID(6) GSG_ABID(7)
    GSG_ABID(8) GSG_ABID(9) GSG_ABID(10) GSG_ABID(11) GSG_ABID(12) GSG_ABID(13) GSG_ABID(14)
    default: return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, CB, 15, 0);
  }
#undef GSG_ABID
}

// The FP32 form's lo correction (the lo table times the input, or the hi table times the
// input's rounding remainder) is about 2^-24 of the hi product: one
// v_mfma_f32_4x4x4_16b_bf16 takes all K <= 4 terms of an entry at once (lane 4b + i
// gives row i of A as 4 bf16, lane 4b + n column n of B; checked on the GPU), and the
// bf16 rounding of its operands (2^-8 relative) adds about 2^-32 of the hi product.
typedef short v4s_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ v4s_t bf16x4(float a, float b, float c, float d) {
  typedef __bf16 v2b __attribute__((ext_vector_type(2)));
  typedef int v2i __attribute__((ext_vector_type(2)));
  const v2b p = {(__bf16)a, (__bf16)b}, q = {(__bf16)c, (__bf16)d};
  const v2i w = {__builtin_bit_cast(int, p), __builtin_bit_cast(int, q)};
  return __builtin_bit_cast(v4s_t, w);
}
template <int CB>
__device__ __forceinline__ v4f_t mfma_bf_abid(v4s_t a, v4s_t b, v4f_t c, int blk) {
#define GSG_ABID(X) \
  case X: return __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(a, b, c, CB, X, 0);
  switch (blk) {
    GSG_ABID(0) GSG_ABID(1) GSG_ABID(2) GSG_ABID(3) GSG_ABID(4) GSG_ABID(5) GSG_ABID(6) GSG_ABID(7)
    GSG_ABID(8) GSG_ABID(9) GSG_ABID(10) GSG_ABID(11) GSG_ABID(12) GSG_ABID(13) GSG_ABID(14)
    default: return __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(a, b, c, CB, 15, 0);
  }
#undef GSG_ABID
}

// wave-uniform element q of a block held one element per lane (lane q % k^2 holds q)
__device__ __forceinline__ float lane_bcast(float v, int q) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), q));
}
__device__ __forceinline__ double lane_bcast(double v, int q) {
  const int2 w = __builtin_bit_cast(int2, v);
  return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_readlane(w.x, q), __builtin_amdgcn_readlane(w.y, q)));
}

// loads and stores through the global address space (the pointers are device
// allocations): global_* instructions instead of flat_* ones
template <typename X>
__device__ __forceinline__ X gld(const X* p) {
  return *(const __attribute__((address_space(1))) X*)p;
}
template <typename X>
__device__ __forceinline__ void gst(X* p, X v) {
  *(__attribute__((address_space(1))) X*)p = v;
}

template <typename P>
__device__ __forceinline__ P sldp(P const* p) {  // scalar load of a pointer field
  return reinterpret_cast<P>(sld(reinterpret_cast<const int64_t*>(p)));
}

// The solve descriptor fields the direct kernel reads, fetched with straight-line
// scalar loads so they arrive in one round trip.
template <typename T, int D>
struct SolveView {
  int64_t N, val_elems;
  int32_t active;
  double dt;
  double coef[D];
  double *y, *acc;
  const T* vals;
  void *xs0, *xs1, *xs2, *w;
  __device__ __forceinline__ SolveView(const SolveDesc& S) {
    N = sld(&S.N);
    val_elems = sld(&S.t.val_elems);
    active = sld(&S.active);
    dt = sld(&S.dt);
#pragma unroll
    for (int e = 0; e < D; ++e) coef[e] = sld(&S.coef[e]);
    y = sldp(&S.y);
    acc = sldp(&S.acc);
    vals = sizeof(T) == 4 ? (const T*)sldp(&S.t.val32) : (const T*)sldp(&S.t.val64);
    xs0 = sldp(&S.xs[0]);
    xs1 = sldp(&S.xs[1]);
    xs2 = sldp(&S.xs[2]);
    w = sldp(&S.w);
  }
  // stage input / output buffers of an RK stage (see stage_in / stage_out)
  __device__ __forceinline__ const T* in(int stage) const {
    if (stage == 1) return sizeof(T) == 4 ? (const T*)xs0 : (const T*)y;
    return (const T*)(stage == 3 ? xs2 : xs1);
  }
  __device__ __forceinline__ T* out(int stage) const {
    if (stage == 4) return sizeof(T) == 4 ? (T*)xs0 : nullptr;
    return (T*)(stage == 2 ? xs2 : xs1);
  }
  // rk for the two consecutive state entries idx, idx + 1 (idx even: 16-byte aligned)
  // with their derivatives r0, r1, through 16-byte global accesses
  __device__ __forceinline__ void rk2(int stage, int64_t idx, double r0, double r1) const {
    typedef double d2 __attribute__((ext_vector_type(2)));
    typedef T t2 __attribute__((ext_vector_type(2)));
    typedef __attribute__((address_space(1))) d2 gd2;
    typedef __attribute__((address_space(1))) t2 gt2;
    const d2 r = {r0, r1};
    const d2 yi = stage <= 3 ? *(const gd2*)(y + idx) : d2{0.0, 0.0};
    const d2 ai = stage >= 2 ? *(const gd2*)(acc + idx) : d2{0.0, 0.0};
    T* o = out(stage);
    if (stage == 1) {
      *(gd2*)(acc + idx) = yi + (dt / 6.0) * r;
      const d2 x = yi + (dt / 2.0) * r;
      *(gt2*)(o + idx) = t2{T(x.x), T(x.y)};
    } else if (stage == 2) {
      *(gd2*)(acc + idx) = ai + (dt / 3.0) * r;
      const d2 x = yi + (dt / 2.0) * r;
      *(gt2*)(o + idx) = t2{T(x.x), T(x.y)};
    } else if (stage == 3) {
      *(gd2*)(acc + idx) = ai + (dt / 3.0) * r;
      const d2 x = yi + dt * r;
      *(gt2*)(o + idx) = t2{T(x.x), T(x.y)};
    } else {
      const d2 yn = ai + (dt / 6.0) * r;
      *(gd2*)(y + idx) = yn;
      if (o) *(gt2*)(o + idx) = t2{T(yn.x), T(yn.y)};
    }
  }
  // the next stage input at o[idx], and in float with lo >= 0 its rounding remainder at o[lo]
  __device__ __forceinline__ static void put(T* o, int64_t idx, int64_t lo, double x) {
    const T h = T(x);
    gst(o + idx, h);
    if (sizeof(T) == 4 && lo >= 0) gst(o + lo, T(x - double(h)));
  }
  // the RK4 update of rk_post with the loaded state (yi, ai); global (not flat) stores
  __device__ __forceinline__ void rk(int stage, int64_t idx, double yi, double ai, double r, int64_t lo = -1) const {
    T* o = out(stage);
    if (stage == 1) {
      gst(acc + idx, yi + (dt / 6.0) * r);
      put(o, idx, lo, yi + (dt / 2.0) * r);
    } else if (stage == 2) {
      gst(acc + idx, ai + (dt / 3.0) * r);
      put(o, idx, lo, yi + (dt / 2.0) * r);
    } else if (stage == 3) {
      gst(acc + idx, ai + (dt / 3.0) * r);
      put(o, idx, lo, yi + dt * r);
    } else {
      const double yn = ai + (dt / 6.0) * r;
      gst(y + idx, yn);
      if (o) put(o, idx, lo, yn);
    }
  }
  __device__ __forceinline__ double y_at(int64_t idx) const { return gld(y + idx); }
  __device__ __forceinline__ double acc_at(int64_t idx) const { return gld(acc + idx); }
};

// OPT / STG: the operation and RK stage, fixed at compile time.  One wavefront owns one
// output cell and all D directions; the workgroup's DIRECT_CPW wavefronts share nothing
// (no barrier), so a wavefront with short lists never waits for one with long lists.
// Lane f owns fiber f of every direction: for direction d its K outputs are the modes
// pos0_d(f) + r K^d.  The cell's entries (all directions, in direction order) are
// consumed in rounds of EB: the round's entry words come through the scalar cache one
// round ahead; each entry's input cell moves into the wavefront's own LDS slot with one
// 16-byte-per-lane buffer_load ... lds (one address per 16 bytes: the texture address
// unit, not the bandwidth, is what scattered 4-byte loads exhaust), the round's blocks
// into VGPRs (FP32: one 16-lane load for all EB blocks, broadcast by the MFMAs), and the
// lanes then read their strided fibers from LDS.  A padding tile
// or a finished solve runs with zero-sized buffer resources (loads return 0) and skips
// the stores.  FP64, k = 3: the 27 fibers take lanes 0..26 (the others read fiber 0 and
// store nothing), an 81-double cell moves with one DMA from the 16-byte boundary at or
// below it (41 lanes), and the round's blocks (wave-uniform) come through the scalar
// cache into SGPRs; the epilogue gives lane l outputs l and 64 + l.
// LDS of the direct kernel: per wavefront EB input slots and a row of direction sums
// FP64 entries per round, measured on the D = 4 classes (n = 8, >= 3 M DoFs, stage
// launch): 1: 0.086 / 0.080 ms (k = 4 / k = 3), 2: 0.090 / 0.105, 3: 0.110 / 0.118,
// 4: 0.108 / 0.126 (fewer LDS slots: more wavefronts a CU)
#ifndef GSG_DIRECT64_EB
#define GSG_DIRECT64_EB 1
#endif
template <typename T>
constexpr int direct_eb() { return sizeof(T) == 4 ? DIRECT_EB : GSG_DIRECT64_EB; }
// elements of an LDS input slot: a cell (D = 4, k = 4), or the 1 KiB window one
// 16-byte-a-lane DMA fills (FP64: an 81-double cell from the 16-byte boundary at or below
// it, or a 64-double cell at D = 3)
template <typename T, int D, int K>
constexpr int direct_slot() { return D == 4 && K == 4 ? 256 : 1024 / int(sizeof(T)); }
template <typename T, int D, int K>
constexpr int direct_lds_bytes() {
  return int(sizeof(T)) * DIRECT_CPW * (direct_eb<T>() * direct_slot<T, D, K>() + kpow(K, D));
}

// wg: the workgroup's group of DIRECT_CPW tiles (blockIdx.x of k_direct, an item of the
// fused kernel); lds: direct_lds_bytes<T, D, K>() bytes of LDS, 16-byte aligned.
template <int DT, int KT, typename T, int OPT, int STG>
__device__ __forceinline__ void direct_body(const LaunchArgs& a, int wg, unsigned char* lds) {
  constexpr int D = DT, K = KT;
  constexpr int KD = ipow_c(K, D);
  constexpr int F = KD / K;  // fibers per direction: lanes f >= F (k = 3) read fiber 0 and discard it
  constexpr int KK = K * K;
  constexpr int EB = direct_eb<T>();  // LDS slots per wavefront
  constexpr int SL = direct_slot<T, D, K>();
  constexpr uint32_t es = sizeof(T);
  constexpr bool FULL = D == 4 && K == 4;              // 256-value cells
  constexpr int CH = FULL ? KD * int(es) / 1024 : 1;  // 1 KiB DMA chunks per input cell
  // FP64 k = 3 and D = 3: the cell (81 doubles, or 64) moves with one DMA of the CHN
  // chunks from the 16-byte boundary at or below it (an 81-double cell need not start on
  // 16 bytes), and the blocks (wave-uniform) come through the scalar cache
  constexpr bool WIN = !FULL;
  // FP64 blocks come through the scalar cache (one entry a round: 16 doubles, 32 SGPRs;
  // at D = 4, k = 4 5-7% faster than a per-lane load and v_readlane broadcasts, which
  // won at 3 entries a round, where 48 doubles spilled SGPRs)
  static_assert(sizeof(T) == 4 || EB == 1 || !FULL, "FP64 scalar blocks at D = 4, k = 4 assume one entry a round");
  constexpr bool SB = !FULL || sizeof(T) == 8;
  constexpr uint32_t CHN = (uint32_t(KD) * es + 8u + 15u) / 16u;
  static_assert((FULL && direct_ok(D, K)) || (sizeof(T) == 8 && direct64_ok(D, K) && KD * 8 + 8 <= 1024),
                "shape not served by the direct kernel");
  T(*Xs)[EB][SL] = reinterpret_cast<T(*)[EB][SL]>(lds);
  T(*Yw)[KD] = reinterpret_cast<T(*)[KD]>(lds + sizeof(T) * DIRECT_CPW * EB * SL);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = uni(tid >> 6);
  T* Y = Yw[wave];
  // first scalar loads: this wavefront's tile (padding tiles repeat a real tile with an empty
  // range, so every address below is readable)
  const Tile* tp = a.tiles + int64_t(wg) * DIRECT_CPW + wave;
  const int64_t start = sld(&tp->start);
  const int64_t end = sld(&tp->end);
  const int32_t solve = uni(sld(&tp->solve));
  const uint32_t lens = uint32_t(uni(sld(&tp->blk)));
  const int32_t* ep = reinterpret_cast<const int32_t*>(sld(&tp->aux));
  // second scalar loads: the first entries and the solve descriptor
  int32_t ew[2 * EB];
#pragma unroll
  for (int q = 0; q < 2 * EB; ++q) ew[q] = uni(sld(ep + q));
  __builtin_amdgcn_sched_barrier(0);  // keep the entry loads ahead of the descriptor's
  const SolveView<T, D> V(a.solves[solve]);
  const bool valid = (end > start) & (V.active != 0);
  const uint32_t vmask = valid ? 0xffffffffu : 0u;
  const int64_t N = V.N;
  const T* x0;
  int64_t dstride = 0;
  if constexpr (OPT == OP_RK_SUM || OPT == OP_RK_WAVE_GRAD) {
    x0 = V.in(STG);
  } else if constexpr (OPT == OP_APPLY_SUM) {
    x0 = (const T*)a.xin;
    dstride = a.instride;
  } else if constexpr (OPT == OP_APPLY_GRAD) {
    x0 = (const T*)a.xin;
  } else if constexpr (OPT == OP_RK_WAVE_SUM) {
    x0 = (const T*)V.w;
    dstride = N;
  } else {
    x0 = (const T*)a.xin2;
    dstride = N;
  }
  constexpr bool MF = sizeof(T) == 4;  // fp32, k = 4: v_mfma_f32_4x4x1_16b_f32 (hi and lo tables)
  // float gradient: the input's rounding remainder (u_lo at 2N of the wave stage input, at
  // a.split of an applied input) runs through the directions' lists as a second part
  int64_t lo = 0;
  if constexpr (MF && OPT == OP_RK_WAVE_GRAD) lo = 2 * N;
  if constexpr (MF && OPT == OP_APPLY_GRAD) lo = a.split;
  const int npart = lo ? 2 : 1;
  const int64_t xelems = dstride ? dstride * D : (lo ? lo + N : N);
  const __amdgpu_buffer_rsrc_t xr = make_rsrc(x0, uint32_t(xelems * int64_t(es)) & vmask);
  const __amdgpu_buffer_rsrc_t vr = make_rsrc(V.vals, uint32_t(V.val_elems * (MF ? 2 : 1) * int64_t(es)) & vmask);

  typedef float v4f __attribute__((ext_vector_type(4)));
  const uint32_t brow = uint32_t(lane & 3) * uint32_t(K) * es;  // MF: this lane's block row
  const uint32_t bq = uint32_t(lane % KK) * es;                 // otherwise: element lane % k^2
  constexpr uint32_t OOR = 0x80000000u;  // added to the range-checked VGPR offset: loads return 0
  T* w = nullptr;
  if constexpr (OPT == OP_APPLY_GRAD) w = (T*)a.yout;
  if constexpr (OPT == OP_RK_WAVE_GRAD) w = (T*)V.w;
  int p0 = 0;  // first entry of direction d in the cell's list
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int sd = ipow_rt(K, d);
    const bool fib = lane < F;
    const int fc = fib ? lane : 0;
    const int vbd = fc % sd + (fc / sd) * sd * K;  // the fiber's element with digit d = 0
    const int len = int((lens >> (8 * d)) & 255u);
    const int p1 = p0 + len;
    const uint32_t xbase = uint32_t(int64_t(d) * dstride * int64_t(es));
    T acc[K];
    v4f accv = {0.f, 0.f, 0.f, 0.f}, accl = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < K; ++q) acc[q] = T(0);
    for (int part = 0; part < npart; ++part)
    for (int c0 = p0; c0 < p1; c0 += EB) {
      const int nb = p1 - c0;
      const uint32_t pbase = xbase + (part ? uint32_t(lo * int64_t(es)) : 0u);
      T bl[EB];
      T bs[SB ? EB : 1][SB ? KK : 1];
      int sh[EB];  // k = 3: the cell's offset past the 16-byte boundary, in elements
#pragma unroll
      for (int u = 0; u < EB; ++u) {
        const bool in = u < nb;
        const uint32_t cb = uint32_t(ew[2 * u]) * es + pbase;
        sh[u] = WIN ? int(cb & 15u) / int(es) : 0;
        if (in) {
          if constexpr (WIN) {
            __builtin_amdgcn_raw_ptr_buffer_load_lds(xr, (lds_ptr_t)Xs[wave][u], 16,
                                                     int(uint32_t(lane) < CHN ? uint32_t(lane) * 16u : OOR),
                                                     int(cb & ~15u), 0, 0);
          } else {
#pragma unroll
            for (int h = 0; h < CH; ++h)
              __builtin_amdgcn_raw_ptr_buffer_load_lds(xr, (lds_ptr_t)((unsigned char*)Xs[wave][u] + h * 1024), 16,
                                                       int(uint32_t(lane) * 16u + uint32_t(h) * 1024u), int(cb), 0, 0);
          }
        }
        if constexpr (SB) {
          const T* bp = V.vals + int64_t(in ? ew[2 * u + 1] : 0) * KK;
#pragma unroll
          for (int q = 0; q < KK; ++q) bs[u][q] = in ? sld(bp + q) : T(0);
        } else if constexpr (!MF) {
          const uint32_t bo = in ? uint32_t(ew[2 * u + 1] * KK) * es : 0u;
          const uint32_t vo = in ? 0u : OOR;
          bl[u] = buf_load<T>(vr, bq + vo, bo);
        }
      }
      // MF: the round's EB hi and EB lo blocks in one load: lane 4u + r (u < EB) takes row
      // r of slot u's hi block, lane 4 (EB + u) + r the same row of its lo block; each
      // slot's MFMAs broadcast A from that slot's 4-lane blocks (ABID u and EB + u)
      v4f bw = {0.f, 0.f, 0.f, 0.f};
      if constexpr (MF) {
        if (lane < 8 * EB) {
          const bool lo = lane >= 4 * EB;
          const int us = (lane >> 2) - (lo ? EB : 0);
          int eb = ew[1];
#pragma unroll
          for (int u = 1; u < EB; ++u) eb = us == u ? ew[2 * u + 1] : eb;
          const uint32_t lof = lo ? uint32_t(V.val_elems) * es : 0u;
          const uint32_t vo = us < nb ? brow + uint32_t(eb * KK) * es + lof : OOR;
          bw = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(vr, int(vo), 0, 0));
        }
      }
      // next round's entry words (this direction's next round, its first again for the
      // second part, or the next direction's first), in flight during this round's math
      const int nx = c0 + EB < p1 ? c0 + EB : (part + 1 < npart ? p0 : p1);
#pragma unroll
      for (int q = 0; q < 2 * EB; ++q) ew[q] = uni(sld(ep + 2 * nx + q));
      __builtin_amdgcn_s_waitcnt(0);  // the DMA into LDS (and everything else) has landed
      const v4s_t bwb = bf16x4(bw[0], bw[1], bw[2], bw[3]);  // MF: the block row as bf16 (lo correction)
#pragma unroll
      for (int u = 0; u < EB; ++u) {
        // unused slots hold stale (or never written) data: read them anyway, branch-free,
        // and clear the bits (a stale NaN times a zero block would still be NaN)
        const bool in = u < nb;
        const T* xs = Xs[wave][u] + sh[u] + vbd;
        T xv[K];
#pragma unroll
        for (int j = 0; j < K; ++j) xv[j] = keep_if(xs[j * sd], in);
        if constexpr (MF) {
          const v4s_t xb = bf16x4(float(xv[0]), float(xv[1]), float(xv[2]), float(xv[3]));
          if (part == 0) {  // hi table into accv, lo table into accl
#pragma unroll
            for (int j = 0; j < K; ++j) accv = mfma_abid<4>(bw[j], float(xv[j]), accv, u);
            accl = mfma_bf_abid<4>(bwb, xb, accl, EB + u);
          } else {  // the input's remainder times the hi table, into accl
            accl = mfma_bf_abid<4>(bwb, xb, accl, u);
          }
        } else {
#pragma unroll
          for (int r = 0; r < K; ++r)
#pragma unroll
            for (int j = 0; j < K; ++j) acc[r] += (SB ? bs[u][r * K + j] : lane_bcast(bl[u], r * K + j)) * xv[j];
        }
      }
    }
    p0 = p1;
    if constexpr (MF) {
      // the hi and lo sums and the coefficient in double; the direction sum is kept in
      // this wavefront's LDS row in float
      double r[K];
#pragma unroll
      for (int q = 0; q < K; ++q) r[q] = double(accv[q]) + double(accl[q]);
      if constexpr (OPT == OP_APPLY_GRAD || OPT == OP_RK_WAVE_GRAD) {
        if (valid) {
#pragma unroll
          for (int q = 0; q < K; ++q) w[int64_t(d) * N + start + vbd + q * sd] = T(r[q]);
        }
      } else {
        const double ce = (OPT == OP_APPLY_SUM || OPT == OP_RK_SUM) ? V.coef[d] : 1.0;
#pragma unroll
        for (int q = 0; q < K; ++q) {
          T* yp = Y + vbd + q * sd;
          *yp = d == 0 ? T(ce * r[q]) : T(double(*yp) + ce * r[q]);
        }
      }
    } else if constexpr (OPT == OP_APPLY_GRAD || OPT == OP_RK_WAVE_GRAD) {
      if (valid && fib) {
#pragma unroll
        for (int q = 0; q < K; ++q) w[int64_t(d) * N + start + vbd + q * sd] = acc[q];
      }
    } else if (fib) {
      // sum over directions in direction order (as the other kernels) in this
      // wavefront's LDS row: Y = c_0 a_0, then Y = fma(c_d, a_d, Y)
      const T ce = (OPT == OP_APPLY_SUM || OPT == OP_RK_SUM) ? T(V.coef[d]) : T(1);
#pragma unroll
      for (int q = 0; q < K; ++q) {
        T* yp = Y + vbd + q * sd;
        if (d == 0)
          *yp = T(0) + ce * acc[q];
        else
          *yp = *yp + ce * acc[q];
      }
    }
  }
  if (!valid) return;  // whole wavefront; no barrier follows
  if constexpr (OPT == OP_APPLY_GRAD || OPT == OP_RK_WAVE_GRAD) return;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if constexpr (OPT == OP_RK_SUM && FULL) {
    // RK epilogue: lane l owns outputs 2l, 2l + 1 and 128 + 2l, 129 + 2l, so each
    // 16-byte (2-double) access instruction covers 1 KiB contiguous across the wavefront
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int m = h * 128 + 2 * lane;
      V.rk2(STG, start + m, double(Y[m]), double(Y[m + 1]));
    }
    return;
  }
  // epilogue: lane l owns the 4 consecutive outputs 4l .. 4l + 3 (smaller cells: outputs
  // l and 64 + l of the k^D)
  constexpr int NT = FULL ? 4 : (KD + 63) / 64;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int m = FULL ? 4 * lane + t : 64 * t + lane;
    if (!FULL && m >= KD) break;
    const int64_t i = start + m;
    const T sum = Y[m];
    if constexpr (OPT == OP_APPLY_SUM) {
      ((double*)a.yout)[i] = double(sum);
    } else if constexpr (OPT == OP_RK_SUM) {
      V.rk(STG, i, STG <= 3 ? V.y_at(i) : 0.0, STG >= 2 ? V.acc_at(i) : 0.0, double(sum));
    } else if constexpr (OPT == OP_APPLY_WAVE_SUM) {
      const T* x = (const T*)a.xin;
      ((double*)a.yout)[i] = double(x[N + i]);
      ((double*)a.yout)[N + i] = double(sum);
    } else {
      const T vN = V.in(STG)[N + i];
      V.rk(STG, i, STG <= 3 ? V.y_at(i) : 0.0, STG >= 2 ? V.acc_at(i) : 0.0, double(vN), 2 * N + i);
      V.rk(STG, N + i, STG <= 3 ? V.y_at(N + i) : 0.0, STG >= 2 ? V.acc_at(N + i) : 0.0, double(sum));
    }
  }
}

template <int DT, int KT, typename T, int OPT, int STG>
__global__ __launch_bounds__(TILE) void k_direct(LaunchArgs a) {
  __shared__ __align__(16) unsigned char lds[direct_lds_bytes<T, DT, KT>()];
  direct_body<DT, KT, T, OPT, STG>(a, int(blockIdx.x), lds);
}

// ---------------------------------------------------------------------------------
// Segmented kernel (seg_ok shapes, float operator).  A wavefront owns seg_cpw(D, K)
// consecutive cells of one solve, one segment of S = seg_lanes(D, K) lanes each; lane
// fl of a segment owns fiber fl of every direction of its cell (lanes fl >= K^(D-1) are
// padding).  Direction by direction the wavefront walks the lists of all its cells in
// lockstep, EB entries a round: each lane loads its cell's entry word, one
// buffer_load ... lds moves every cell's input cell for slot u into the wavefront's LDS
// slot (16 bytes a lane, a segment of 16 S bytes per cell, starting at the 16-byte
// boundary at or below the cell), each lane loads row lane % 4
// of its cell's block from the 4 x 4-padded table, and v_mfma_f32_4x4x1_16b_f32 applies
// the blocks (a 4-lane MFMA block lies in one cell).  The per-direction results meet in
// the wavefront's LDS row and the fused epilogue runs over the wavefront's outputs.
// No barrier: the wavefronts of a workgroup are independent.
// entries per round (LDS slots per wavefront), measured per shape: 4, except D = 4,
// k = 3 (two 81-value cells a wavefront) where 2 slots (more wavefronts per CU) win by 9%
// (and 1 loses 6%)
constexpr int seg_eb(int D, int K) { return (D == 4 && K == 3) || seg_lanes(D, K) == 4 ? 2 : 4; }

// LDS of the segmented kernel: per wavefront EB input slots of 256 floats and a row of
// the wavefront's direction sums
// the tuned alternative (variant GSG_KERNEL_TILE_ALT): 4 entries a round where the rule
// says 2, 2 where it says 4
constexpr int seg_eb_alt(int D, int K) { return seg_eb(D, K) == 4 ? 2 : 4; }
constexpr int seg_lds_bytes(int D, int K, int EB) {
  return 4 * SEG_WAVES * (EB * 256 + seg_cpw(D, K) * kpow(K, D));
}
constexpr int seg_lds_bytes(int D, int K) { return seg_lds_bytes(D, K, seg_eb(D, K)); }

// wg: the workgroup's tiles (blockIdx.x of k_seg, an item of the fused kernel); dsel:
// the one direction an OP_APPLY_GRAD workgroup runs (-1: all); lds: seg_lds_bytes(D, K)
// bytes of LDS, 16-byte aligned.
template <int DT, int KT, int OPT, int STG, int EBX = 0>
__device__ __forceinline__ void seg_body(const LaunchArgs& a, int wg, int dsel, unsigned char* lds) {
  typedef float T;
  typedef float v4f __attribute__((ext_vector_type(4)));
  constexpr int D = DT, K = KT;
  constexpr int KD = ipow_c(K, D);
  constexpr int F = KD / K;
  constexpr int S = seg_lanes(D, K);
  constexpr int CPW = seg_cpw(D, K);
  constexpr int EB = EBX ? EBX : seg_eb(D, K);  // entries a round (EBX: the tuned alternative)
  static_assert(EB >= 1 && EB <= 8, "the block broadcast addresses at most 8 slots");
  constexpr uint32_t es = 4;
  constexpr int NOUT = CPW * KD;  // outputs of a wavefront
  // segments of 8 or more lanes load a round's blocks with NQ = EB / (S / 4) instructions:
  // in load q, lane 4v + r of a segment takes row r of slot q S / 4 + v's block of its
  // cell, and the MFMAs of slot u broadcast A over the segment's S / 4 blocks from its
  // block u % (S / 4) (ABID)
  constexpr int SPB = S / 4;  // slots whose blocks one load covers
  constexpr bool BW = S >= 8;
  // HL1: a load covers the hi and the lo blocks of every slot (2 EB <= SPB; the lo
  // blocks take the 4-lane blocks EB .. 2 EB - 1 of the segment); otherwise the lo
  // blocks come in loads of their own
  constexpr bool HL1 = BW && 2 * EB <= SPB;
  constexpr int NQ = BW ? (EB + SPB - 1) / SPB : 1;
  constexpr int CB = S >= 64 ? 4 : S >= 32 ? 3 : S >= 16 ? 2 : S >= 8 ? 1 : 0;
  // input cells move 16 bytes a lane: segment seg (S lanes, 16 S bytes) of a slot takes
  // its cell.  With k^D % 4 != 0 a cell need not start on 16 bytes: the DMA starts at
  // the 16-byte boundary below it and the fiber reads skip the cell's offset there
  // (0..3 floats).  The stage vectors start on 256 bytes (device allocations).
  constexpr bool ALN = KD % 4 == 0;  // every cell starts on 16 bytes
  constexpr int SLOT = 256;          // floats per LDS slot (all cells of the wavefront)
  static_assert(seg_ok(D, K) && KD * 4 + (ALN ? 0 : 12) <= 16 * S && S % 4 == 0,
                "shape not served by the segmented kernel");
  float(*Xs)[EB][SLOT] = reinterpret_cast<float(*)[EB][SLOT]>(lds);
  float(*Yw)[NOUT] = reinterpret_cast<float(*)[NOUT]>(lds + 4 * SEG_WAVES * EB * SLOT);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = uni(tid >> 6);
  const int seg = lane / S;
  const int fl = lane - seg * S;
  const bool fib = fl < F;  // a real fiber
  float* Y = Yw[wave];
  // this lane's cell (all tiles of a wavefront belong to one solve)
  const int64_t t0 = (int64_t(wg) * SEG_WAVES + wave) * CPW;
  const Tile* tp = a.tiles + t0 + seg;
  const int64_t start = tp->start;
  const int64_t end = tp->end;
  const uint32_t lens = uint32_t(tp->blk);
  const int32_t* ep = reinterpret_cast<const int32_t*>(tp->aux);
  const int32_t solve = uni(sld(&a.tiles[t0].solve));
  const SolveView<T, D> V(a.solves[solve]);
  const bool valid = (end > start) & (V.active != 0);  // per cell
  // the wavefront's cells are consecutive in the vector (padding tiles only at the end):
  // cell c starts at start0 + c k^D; bit c S of cmask says whether it is real
  const uint64_t cmask = __ballot(valid && fl == 0);
  const int64_t start0 = int64_t(__builtin_amdgcn_readfirstlane(uint32_t(start)));
  const uint32_t vmask = V.active != 0 ? 0xffffffffu : 0u;
  const int64_t N = V.N;
  const T* x0;
  int64_t dstride = 0;
  if constexpr (OPT == OP_RK_SUM || OPT == OP_RK_WAVE_GRAD) {
    x0 = V.in(STG);
  } else if constexpr (OPT == OP_APPLY_SUM) {
    x0 = (const T*)a.xin;
    dstride = a.instride;
  } else if constexpr (OPT == OP_APPLY_GRAD) {
    x0 = (const T*)a.xin;
  } else if constexpr (OPT == OP_RK_WAVE_SUM) {
    x0 = (const T*)V.w;
    dstride = N;
  } else {
    x0 = (const T*)a.xin2;
    dstride = N;
  }
  // gradient: the input's rounding remainder (u_lo at 2N of the wave stage input, at
  // a.split of an applied input) runs through the directions' lists as a second part
  int64_t lo = 0;
  if constexpr (OPT == OP_RK_WAVE_GRAD) lo = 2 * N;
  if constexpr (OPT == OP_APPLY_GRAD) lo = a.split;
  const int npart = lo ? 2 : 1;
  const int64_t xelems = dstride ? dstride * D : (lo ? lo + N : N);
  const __amdgpu_buffer_rsrc_t xr = make_rsrc(x0, uint32_t(xelems * 4) & vmask);
  // the padded hi blocks, then the padded lo blocks
  const uint32_t lofs = uint32_t(sld(&a.solves[solve].t.val32p_elems) * 4);
  const __amdgpu_buffer_rsrc_t br = make_rsrc(sldp(&a.solves[solve].t.val32p), (2u * lofs) & vmask);
  constexpr uint32_t OOR = 0x80000000u;  // added to a range-checked VGPR offset: loads return 0
  const uint32_t brow = uint32_t(lane & 3) * 16u;  // row lane % 4 of a padded block
  const uint32_t dmal = uint32_t(fl) * 16u;        // this lane's 16 bytes of its segment's cell
  const bool dmain = dmal < uint32_t(KD) * es + (ALN ? 0u : 12u);
  T* w = nullptr;
  if constexpr (OPT == OP_APPLY_GRAD) w = (T*)a.yout;
  if constexpr (OPT == OP_RK_WAVE_GRAD) w = (T*)V.w;
  // per direction: this cell's list length and first entry, and the wavefront's rounds
  int lraw[D], p0v[D], nrd[D];
  {
    int c = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      lraw[d] = int((lens >> (8 * d)) & 255u);
      p0v[d] = c;
      c += lraw[d];
      const int lv = valid ? lraw[d] : 0;
      int nr = 0;
#pragma unroll
      for (int q = 0; q < CPW; ++q) nr = max(nr, __builtin_amdgcn_readlane(lv, q * S));
      nrd[d] = nr;
    }
  }
  // entry words of a round (lanes past their list read their first entry, unused).
  // With at most GSG_SEW_CPW cells a wavefront (2: measured best; 4 slows the 4-cell
  // shapes by a third) they come through the scalar cache, cell by cell, and each lane
  // picks its cell's words (no vector-memory instruction); otherwise each lane loads
  // its own.
#ifndef GSG_SEW_CPW
#define GSG_SEW_CPW 2
#endif
  constexpr bool SEW = CPW <= GSG_SEW_CPW;
  // SEW: every cell's entry address, list lengths and first entries as wave-uniform
  // values, read from the cell's first lane once
  const int32_t* epq[SEW ? CPW : 1];
  int lqd[SEW ? CPW : 1][D], pqd[SEW ? CPW : 1][D];
  if constexpr (SEW) {
    const uint64_t epa = reinterpret_cast<uintptr_t>(ep);
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
      const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(epa)), q * S));
      const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(epa >> 32)), q * S));
      epq[q] = reinterpret_cast<const int32_t*>(uintptr_t((uint64_t(hi) << 32) | lo));
#pragma unroll
      for (int e = 0; e < D; ++e) {
        lqd[q][e] = __builtin_amdgcn_readlane(lraw[e], q * S);
        pqd[q][e] = __builtin_amdgcn_readlane(p0v[e], q * S);
      }
    }
  }
  int2 ewc[EB], ewn[EB];
  auto fetch = [&](int2(&dst)[EB], int dd, int c0) {
    if constexpr (SEW) {
#pragma unroll
      for (int q = 0; q < CPW; ++q) {
        const int lq = lqd[q][dd];
        const int pq = pqd[q][dd];
#pragma unroll
        for (int u = 0; u < EB; ++u) {
          const int idx = c0 + u < lq ? pq + c0 + u : 0;
          const int32_t wx = uni(sld(epq[q] + 2 * idx));
          const int32_t wy = uni(sld(epq[q] + 2 * idx + 1));
          if (q == 0 || seg == q) dst[u] = make_int2(wx, wy);
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < EB; ++u) {
        const bool in = c0 + u < lraw[dd];
        dst[u] = *reinterpret_cast<const int2*>(ep + 2 * (in ? p0v[dd] + c0 + u : 0));
      }
    }
  };
  fetch(ewc, 0, 0);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    if constexpr (OPT == OP_APPLY_GRAD) {
      if (dsel >= 0 && d != dsel) {  // wave-uniform
        if (d < dsel) fetch(ewc, d + 1 < D ? d + 1 : d, 0);  // the next direction's first entries
        continue;
      }
    }
    const int sdd = ipow_rt(K, d);
    const int flc = fib ? fl : 0;  // padding lanes read (and discard) fiber 0
    const int vbd = flc % sdd + (flc / sdd) * sdd * K;  // the fiber's element with digit d = 0
    const int len = lraw[d];
    const int nr = nrd[d];
    const uint32_t xbase = uint32_t(int64_t(d) * dstride * 4);
    v4f accv = {0.f, 0.f, 0.f, 0.f}, accl = {0.f, 0.f, 0.f, 0.f};
    for (int part = 0; part < npart; ++part)
    for (int c0 = 0; c0 < nr; c0 += EB) {
      const uint32_t pbase = xbase + (part ? uint32_t(lo * 4) : 0u);
      v4f bv[EB], bvl[EB];
      bool inu[EB];
      int shu[EB];  // the lane's cell's offset past a 16-byte boundary, in floats
#pragma unroll
      for (int u = 0; u < EB; ++u) {
        const bool in = c0 + u < len;
        inu[u] = in;
        const int2 ew = ewc[u];
        const uint32_t cb = uint32_t(ew.x) * es + pbase;  // the cell's byte offset
        shu[u] = ALN ? 0 : int(cb & 15u) >> 2;
        const uint32_t xo = in && dmain ? (ALN ? cb : cb & ~15u) + dmal : OOR;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(xr, (lds_ptr_t)Xs[wave][u], 16, int(xo), 0, 0, 0);
        if constexpr (!BW) {
          const uint32_t bo = in ? uint32_t(ew.y) * 64u + brow : OOR;
          bv[u] = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(br, int(bo), 0, 0));
          bvl[u] = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(br, int(in ? bo + lofs : OOR), 0, 0));
        }
      }
      v4f bw[NQ], bwl[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        bw[q] = v4f{0.f, 0.f, 0.f, 0.f};
        bwl[q] = v4f{0.f, 0.f, 0.f, 0.f};
        if constexpr (HL1) {
          const int vs = fl >> 2;            // 4-lane block of the segment: hi of slot vs, or
          const bool lob = vs >= EB;         // lo of slot vs - EB
          const int us = lob ? vs - EB : vs;
          if (vs < 2 * EB) {
            int eb = ewc[0].y;
#pragma unroll
            for (int u = 1; u < EB; ++u) eb = us == u ? ewc[u].y : eb;
            const uint32_t bo = c0 + us < len ? uint32_t(eb) * 64u + brow + (lob ? lofs : 0u) : OOR;
            bw[q] = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(br, int(bo), 0, 0));
          }
        } else if constexpr (BW) {
          const int us = q * SPB + (fl >> 2);  // the slot whose block row this lane loads
          if (us < EB) {
            int eb = ewc[0].y;
#pragma unroll
            for (int u = 1; u < EB; ++u) eb = us == u ? ewc[u].y : eb;
            const bool inb = c0 + us < len;
            const uint32_t bo = inb ? uint32_t(eb) * 64u + brow : OOR;
            bw[q] = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(br, int(bo), 0, 0));
            bwl[q] = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(br, int(inb ? bo + lofs : OOR), 0, 0));
          }
        }
      }
      // the next round's entry words (this direction's next round or the next
      // direction's first), in flight with this round's loads
      if (c0 + EB < nr)
        fetch(ewn, d, c0 + EB);
      else if (part + 1 < npart)
        fetch(ewn, d, 0);  // the direction's first entries again, for the second part
      else if (d + 1 < D)
        fetch(ewn, d + 1 < D ? d + 1 : d, 0);
      __builtin_amdgcn_s_waitcnt(0);  // the DMA into LDS (and everything else) has landed
#pragma unroll
      for (int u = 0; u < EB; ++u) {
        const bool use = inu[u] && fib;
#pragma unroll
        for (int j = 0; j < K; ++j) {
          const int el = vbd + j * sdd;  // element of the cell
          const float* xp = Xs[wave][u] + seg * (4 * S) + shu[u] + el;
          const float xv = keep_if(*xp, use);  // read unconditionally: no branch per element
          // part 0: hi table into accv, lo table into accl; part 1 (the input's
          // remainder): hi table into accl
          if constexpr (HL1) {
            if (part == 0) {
              accv = mfma_abid<CB>(bw[0][j], xv, accv, u);
              accl = mfma_abid<CB>(bw[0][j], xv, accl, EB + u);
            } else {
              accl = mfma_abid<CB>(bw[0][j], xv, accl, u);
            }
          } else if constexpr (BW) {
            if (part == 0) {
              accv = mfma_abid<CB>(bw[u / SPB][j], xv, accv, u % SPB);
              accl = mfma_abid<CB>(bwl[u / SPB][j], xv, accl, u % SPB);
            } else {
              accl = mfma_abid<CB>(bw[u / SPB][j], xv, accl, u % SPB);
            }
          } else {
            if (part == 0) {
              accv = __builtin_amdgcn_mfma_f32_4x4x1f32(bv[u][j], xv, accv, 0, 0, 0);
              accl = __builtin_amdgcn_mfma_f32_4x4x1f32(bvl[u][j], xv, accl, 0, 0, 0);
            } else {
              accl = __builtin_amdgcn_mfma_f32_4x4x1f32(bv[u][j], xv, accl, 0, 0, 0);
            }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < EB; ++u) ewc[u] = ewn[u];
      // the next round's DMA rewrites the slots only after these reads (the MFMAs above
      // consumed them)
    }
    if (nr == 0 && d + 1 < D) fetch(ewc, d + 1 < D ? d + 1 : d, 0);  // no round ran
    // the hi and lo sums and the coefficient in double
    double r[K];
#pragma unroll
    for (int q = 0; q < K; ++q) r[q] = double(accv[q]) + double(accl[q]);
    if constexpr (OPT == OP_APPLY_GRAD || OPT == OP_RK_WAVE_GRAD) {
      if (valid && fib) {
#pragma unroll
        for (int q = 0; q < K; ++q) w[int64_t(d) * N + start + vbd + q * sdd] = T(r[q]);
      }
    } else {
      // sum over directions in direction order in the wavefront's LDS row (float)
      const double ce = (OPT == OP_APPLY_SUM || OPT == OP_RK_SUM) ? V.coef[d] : 1.0;
      if (fib) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
          T* yp = Y + seg * KD + vbd + q * sdd;
          *yp = d == 0 ? T(ce * r[q]) : T(double(*yp) + ce * r[q]);
        }
      }
    }
  }
  if constexpr (OPT == OP_APPLY_GRAD || OPT == OP_RK_WAVE_GRAD) return;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if constexpr (OPT == OP_RK_SUM) {
    // RK epilogue on pairs of consecutive outputs with 16-byte accesses (rk2); pairs
    // start on an even state index, so an odd first output and an unpaired last one go
    // alone.  A pair may straddle two cells: each element keeps its cell's validity.
    auto real = [&](int m) { return ((cmask >> ((m / KD) * S)) & 1u) != 0; };
    auto one = [&](int m) {
      const int64_t i = start0 + m;
      V.rk(STG, i, STG <= 3 ? V.y_at(i) : 0.0, STG >= 2 ? V.acc_at(i) : 0.0, double(Y[m]));
    };
    const int sh = int(start0 & 1);
    const int np = (NOUT - sh) / 2;
#pragma unroll
    for (int p = lane; p < np; p += 64) {
      const int m = sh + 2 * p;
      const bool r0 = real(m), r1 = real(m + 1);
      if (r0 && r1) {
        V.rk2(STG, start0 + m, double(Y[m]), double(Y[m + 1]));
      } else {
        if (r0) one(m);
        if (r1) one(m + 1);
      }
    }
    if (lane == 0 && sh && real(0)) one(0);
    if (lane == 1 && ((NOUT - sh) & 1) && real(NOUT - 1)) one(NOUT - 1);
    return;
  }
#pragma unroll
  for (int m = lane; m < NOUT; m += 64) {
    const int c = m / KD;
    if (!((cmask >> (c * S)) & 1u)) continue;
    const int64_t i = start0 + m;
    const T sum = Y[m];
    if constexpr (OPT == OP_APPLY_SUM) {
      ((double*)a.yout)[i] = double(sum);
    } else if constexpr (OPT == OP_RK_SUM) {
      V.rk(STG, i, STG <= 3 ? V.y_at(i) : 0.0, STG >= 2 ? V.acc_at(i) : 0.0, double(sum));
    } else if constexpr (OPT == OP_APPLY_WAVE_SUM) {
      const T* x = (const T*)a.xin;
      ((double*)a.yout)[i] = double(x[N + i]);
      ((double*)a.yout)[N + i] = double(sum);
    } else {
      const T vN = V.in(STG)[N + i];
      V.rk(STG, i, STG <= 3 ? V.y_at(i) : 0.0, STG >= 2 ? V.acc_at(i) : 0.0, double(vN), 2 * N + i);
      V.rk(STG, N + i, STG <= 3 ? V.y_at(N + i) : 0.0, STG >= 2 ? V.acc_at(N + i) : 0.0, double(sum));
    }
  }
}

template <int DT, int KT, int OPT, int STG, int EBX = 0>
__global__ __launch_bounds__(TILE) void k_seg(LaunchArgs a) {
  __shared__ __align__(16) unsigned char lds[seg_lds_bytes(DT, KT, EBX ? EBX : seg_eb(DT, KT))];
  // OP_APPLY_GRAD on a small grid: grid row y runs direction y only (each direction
  // writes its own output), so the launch has D times the wavefronts
  const int dsel = (OPT == OP_APPLY_GRAD && gridDim.y > 1) ? int(blockIdx.y) : -1;
  seg_body<DT, KT, OPT, STG, EBX>(a, int(blockIdx.x), dsel, lds);
}

template <int DT, int KT, int OPT, int EBX = 0>
static void launch_seg_stage(const LaunchArgs& a, hipStream_t s) {
  const dim3 grid(unsigned(a.ntiles / (SEG_WAVES * seg_cpw(DT, KT))));
  switch (a.stage) {
    case 1: hipLaunchKernelGGL((k_seg<DT, KT, OPT, 1, EBX>), grid, dim3(TILE), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_seg<DT, KT, OPT, 2, EBX>), grid, dim3(TILE), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_seg<DT, KT, OPT, 3, EBX>), grid, dim3(TILE), 0, s, a); break;
    default: hipLaunchKernelGGL((k_seg<DT, KT, OPT, 4, EBX>), grid, dim3(TILE), 0, s, a); break;
  }
}

// the RK stages with the tuned alternative entries per round (the apply operations keep
// the rule's)
template <int DT, int KT>
static hipError_t launch_seg_alt(const LaunchArgs& a, hipStream_t s) {
  if (a.ntiles % (SEG_WAVES * seg_cpw(DT, KT))) return hipErrorInvalidValue;  // the host pads
  constexpr int E = seg_eb_alt(DT, KT);
  switch (a.op) {
    case OP_RK_SUM: launch_seg_stage<DT, KT, OP_RK_SUM, E>(a, s); break;
    case OP_RK_WAVE_GRAD: launch_seg_stage<DT, KT, OP_RK_WAVE_GRAD, E>(a, s); break;
    case OP_RK_WAVE_SUM: launch_seg_stage<DT, KT, OP_RK_WAVE_SUM, E>(a, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int DT, int KT>
static hipError_t launch_seg(const LaunchArgs& a, hipStream_t s) {
  if (a.ntiles % (SEG_WAVES * seg_cpw(DT, KT))) return hipErrorInvalidValue;  // the host pads
  const dim3 grid(unsigned(a.ntiles / (SEG_WAVES * seg_cpw(DT, KT))));
  // the gradient of a grid of fewer than 1024 workgroups (4 per CU) runs one direction
  // per grid row: measured on the 2 + 2 dimensional Vlasov derivative (k = 2, n = 8)
  const dim3 ggrid(grid.x, grid.x < 1024 ? unsigned(DT) : 1u);
  switch (a.op) {
    case OP_APPLY_SUM: hipLaunchKernelGGL((k_seg<DT, KT, OP_APPLY_SUM, 0>), grid, dim3(TILE), 0, s, a); break;
    case OP_APPLY_GRAD: hipLaunchKernelGGL((k_seg<DT, KT, OP_APPLY_GRAD, 0>), ggrid, dim3(TILE), 0, s, a); break;
    case OP_APPLY_WAVE_SUM: hipLaunchKernelGGL((k_seg<DT, KT, OP_APPLY_WAVE_SUM, 0>), grid, dim3(TILE), 0, s, a); break;
    case OP_RK_SUM: launch_seg_stage<DT, KT, OP_RK_SUM>(a, s); break;
    case OP_RK_WAVE_GRAD: launch_seg_stage<DT, KT, OP_RK_WAVE_GRAD>(a, s); break;
    case OP_RK_WAVE_SUM: launch_seg_stage<DT, KT, OP_RK_WAVE_SUM>(a, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// FP64 segmented kernel (seg64_ok shapes: FP64, k = 2, D <= 4).  A wavefront owns
// seg_cpw(D, 2) consecutive cells of one solve, one segment of S = seg_lanes(D, 2) lanes
// each; lane fl < k^(D-1) of a segment owns fiber fl of every direction of its cell.
// Direction by direction the wavefront walks its cells' lists in lockstep, one entry a
// round: each lane holds its cell's entry words (the next round's loaded during this
// round's math), one 16-byte-a-lane DMA moves every cell's input cell (8 k^D bytes, on
// 16 bytes) into the segment's part of the wavefront's 1 KiB LDS slot, each lane loads
// its cell's 2 x 2 block (two 16-byte loads) and does the k^2 FMAs of its fiber.  The
// per-direction results meet in the wavefront's LDS row (FP64) and the epilogue runs
// over the wavefront's outputs, as the direct kernel's.  No barrier.
template <int DT, int KT, int OPT, int STG>
__device__ __forceinline__ void seg64_body(const LaunchArgs& a, int wg, unsigned char* lds) {
  typedef double T;
  typedef double v2d __attribute__((ext_vector_type(2)));
  constexpr int D = DT, K = KT;
  constexpr int KD = ipow_c(K, D);
  constexpr int F = KD / K;
  constexpr int KK = K * K;
  constexpr int S = seg_lanes(D, K);
  constexpr int CPW = seg_cpw(D, K);
  constexpr int NOUT = CPW * KD;
  constexpr uint32_t es = 8;
  static_assert(seg64_ok(D, K) && KD * 8 <= 16 * S && KK == 4, "shape not served by the FP64 segmented kernel");
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = uni(tid >> 6);
  const int seg = lane / S;
  const int fl = lane - seg * S;
  const bool fib = fl < F;
  T* const Xw = reinterpret_cast<T*>(lds) + wave * 128;                       // 1 KiB slot
  T* const Y = reinterpret_cast<T*>(lds + SEG_WAVES * 1024) + wave * NOUT;  // direction sums
  const int64_t t0 = (int64_t(wg) * SEG_WAVES + wave) * CPW;
  const Tile* tp = a.tiles + t0 + seg;
  const int64_t start = tp->start;
  const int64_t end = tp->end;
  const uint32_t lens = uint32_t(tp->blk);
  const int32_t* ep = reinterpret_cast<const int32_t*>(tp->aux);
  const int32_t solve = uni(sld(&a.tiles[t0].solve));
  const SolveView<T, D> V(a.solves[solve]);
  const bool valid = (end > start) & (V.active != 0);  // per cell
  const uint64_t cmask = __ballot(valid && fl == 0);
  const int64_t start0 = int64_t(__builtin_amdgcn_readfirstlane(uint32_t(start)));
  const uint32_t vmask = V.active != 0 ? 0xffffffffu : 0u;
  const int64_t N = V.N;
  const T* x0;
  int64_t dstride = 0;
  if constexpr (OPT == OP_RK_SUM || OPT == OP_RK_WAVE_GRAD) {
    x0 = V.in(STG);
  } else if constexpr (OPT == OP_APPLY_SUM) {
    x0 = (const T*)a.xin;
    dstride = a.instride;
  } else if constexpr (OPT == OP_APPLY_GRAD) {
    x0 = (const T*)a.xin;
  } else if constexpr (OPT == OP_RK_WAVE_SUM) {
    x0 = (const T*)V.w;
    dstride = N;
  } else {
    x0 = (const T*)a.xin2;
    dstride = N;
  }
  const int64_t xelems = dstride ? dstride * D : N;
  const __amdgpu_buffer_rsrc_t xr = make_rsrc(x0, uint32_t(xelems * int64_t(es)) & vmask);
  const __amdgpu_buffer_rsrc_t vr = make_rsrc(V.vals, uint32_t(V.val_elems * int64_t(es)) & vmask);
  constexpr uint32_t OOR = 0x80000000u;
  const bool dmain = uint32_t(fl) * 16u < uint32_t(KD) * es;  // this lane's 16 bytes of its cell
  T* w = nullptr;
  if constexpr (OPT == OP_APPLY_GRAD) w = (T*)a.yout;
  if constexpr (OPT == OP_RK_WAVE_GRAD) w = (T*)V.w;
  int lraw[D], p0v[D], nrd[D];
  {
    int c = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      lraw[d] = int((lens >> (8 * d)) & 255u);
      p0v[d] = c;
      c += lraw[d];
      const int lv = valid ? lraw[d] : 0;
      int nr = 0;
#pragma unroll
      for (int q = 0; q < CPW; ++q) nr = max(nr, __builtin_amdgcn_readlane(lv, q * S));
      nrd[d] = nr;
    }
  }
  // this lane's cell's entry words of round c0 of direction dd (its first entry past the list)
  auto fetch = [&](int dd, int c0) {
    return *reinterpret_cast<const int2*>(ep + 2 * (c0 < lraw[dd] ? p0v[dd] + c0 : 0));
  };
  int2 ewc = fetch(0, 0);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int sd = ipow_rt(K, d);
    const int fc = fib ? fl : 0;
    const int vbd = fc % sd + (fc / sd) * sd * K;  // the fiber's element with digit d = 0
    const int len = lraw[d];
    const int nr = nrd[d];
    const uint32_t xbase = uint32_t(int64_t(d) * dstride * int64_t(es));
    T acc[K];
#pragma unroll
    for (int q = 0; q < K; ++q) acc[q] = T(0);
    for (int c0 = 0; c0 < nr; ++c0) {
      const bool in = c0 < len;
      const uint32_t cb = uint32_t(ewc.x) * es + xbase;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(xr, (lds_ptr_t)Xw, 16, int(in && dmain ? cb + uint32_t(fl) * 16u : OOR),
                                               0, 0, 0);
      const uint32_t bo = in ? uint32_t(ewc.y) * uint32_t(KK) * es : OOR;
      const v2d b01 = __builtin_bit_cast(v2d, __builtin_amdgcn_raw_buffer_load_b128(vr, int(bo), 0, 0));
      const v2d b23 = __builtin_bit_cast(v2d, __builtin_amdgcn_raw_buffer_load_b128(vr, int(in ? bo + 16u : OOR), 0, 0));
      // the next round's entry words (this direction's next round or the next direction's
      // first), in flight with this round's loads
      const int2 ewn = c0 + 1 < nr ? fetch(d, c0 + 1) : fetch(d + 1 < D ? d + 1 : d, 0);
      __builtin_amdgcn_s_waitcnt(0);  // the DMA into LDS (and everything else) has landed
      const T* xs = Xw + seg * (2 * S) + vbd;
      const T x0v = keep_if(xs[0], in), x1v = keep_if(xs[sd], in);
      acc[0] += b01.x * x0v;
      acc[0] += b01.y * x1v;
      acc[1] += b23.x * x0v;
      acc[1] += b23.y * x1v;
      ewc = ewn;
      // the next round's DMA rewrites the slot only after these reads (the FMAs consumed them)
    }
    if (nr == 0) ewc = fetch(d + 1 < D ? d + 1 : d, 0);
    if constexpr (OPT == OP_APPLY_GRAD || OPT == OP_RK_WAVE_GRAD) {
      if (valid && fib) {
#pragma unroll
        for (int q = 0; q < K; ++q) w[int64_t(d) * N + start + vbd + q * sd] = acc[q];
      }
    } else if (fib) {
      // sum over directions in direction order: Y = c_0 a_0, then Y = fma(c_d, a_d, Y)
      const T ce = (OPT == OP_APPLY_SUM || OPT == OP_RK_SUM) ? T(V.coef[d]) : T(1);
#pragma unroll
      for (int q = 0; q < K; ++q) {
        T* yp = Y + seg * KD + vbd + q * sd;
        if (d == 0)
          *yp = T(0) + ce * acc[q];
        else
          *yp = *yp + ce * acc[q];
      }
    }
  }
  if constexpr (OPT == OP_APPLY_GRAD || OPT == OP_RK_WAVE_GRAD) return;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int m = lane; m < NOUT; m += 64) {
    const int c = m / KD;
    if (!((cmask >> (c * S)) & 1u)) continue;
    const int64_t i = start0 + m;
    const T sum = Y[m];
    if constexpr (OPT == OP_APPLY_SUM) {
      ((double*)a.yout)[i] = double(sum);
    } else if constexpr (OPT == OP_RK_SUM) {
      V.rk(STG, i, STG <= 3 ? V.y_at(i) : 0.0, STG >= 2 ? V.acc_at(i) : 0.0, double(sum));
    } else if constexpr (OPT == OP_APPLY_WAVE_SUM) {
      const T* x = (const T*)a.xin;
      ((double*)a.yout)[i] = double(x[N + i]);
      ((double*)a.yout)[N + i] = double(sum);
    } else {
      const T vN = V.in(STG)[N + i];
      V.rk(STG, i, STG <= 3 ? V.y_at(i) : 0.0, STG >= 2 ? V.acc_at(i) : 0.0, double(vN));
      V.rk(STG, N + i, STG <= 3 ? V.y_at(N + i) : 0.0, STG >= 2 ? V.acc_at(N + i) : 0.0, double(sum));
    }
  }
}

template <int DT, int KT, int OPT, int STG>
__global__ __launch_bounds__(TILE) void k_seg64(LaunchArgs a) {
  __shared__ __align__(16) unsigned char lds[SEG_WAVES * (1024 + 8 * seg_cpw(DT, KT) * ipow_c(KT, DT))];
  seg64_body<DT, KT, OPT, STG>(a, int(blockIdx.x), lds);
}

template <int DT, int KT, int OPT>
static void launch_seg64_stage(const LaunchArgs& a, const dim3 grid, hipStream_t s) {
  switch (a.stage) {
    case 1: hipLaunchKernelGGL((k_seg64<DT, KT, OPT, 1>), grid, dim3(TILE), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_seg64<DT, KT, OPT, 2>), grid, dim3(TILE), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_seg64<DT, KT, OPT, 3>), grid, dim3(TILE), 0, s, a); break;
    default: hipLaunchKernelGGL((k_seg64<DT, KT, OPT, 4>), grid, dim3(TILE), 0, s, a); break;
  }
}

template <int DT, int KT>
static hipError_t launch_seg64(const LaunchArgs& a, hipStream_t s) {
  if (a.ntiles % (SEG_WAVES * seg_cpw(DT, KT))) return hipErrorInvalidValue;  // the host pads
  const dim3 grid(unsigned(a.ntiles / (SEG_WAVES * seg_cpw(DT, KT))));
  switch (a.op) {
    case OP_APPLY_SUM: hipLaunchKernelGGL((k_seg64<DT, KT, OP_APPLY_SUM, 0>), grid, dim3(TILE), 0, s, a); break;
    case OP_APPLY_GRAD: hipLaunchKernelGGL((k_seg64<DT, KT, OP_APPLY_GRAD, 0>), grid, dim3(TILE), 0, s, a); break;
    case OP_APPLY_WAVE_SUM: hipLaunchKernelGGL((k_seg64<DT, KT, OP_APPLY_WAVE_SUM, 0>), grid, dim3(TILE), 0, s, a); break;
    case OP_RK_SUM: launch_seg64_stage<DT, KT, OP_RK_SUM>(a, grid, s); break;
    case OP_RK_WAVE_GRAD: launch_seg64_stage<DT, KT, OP_RK_WAVE_GRAD>(a, grid, s); break;
    case OP_RK_WAVE_SUM: launch_seg64_stage<DT, KT, OP_RK_WAVE_SUM>(a, grid, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int DT, int KT, typename T, int OPT>
static void launch_direct_stage(const LaunchArgs& a, hipStream_t s) {
  switch (a.stage) {
    case 1: hipLaunchKernelGGL((k_direct<DT, KT, T, OPT, 1>), dim3(a.ntiles / DIRECT_CPW), dim3(TILE), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_direct<DT, KT, T, OPT, 2>), dim3(a.ntiles / DIRECT_CPW), dim3(TILE), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_direct<DT, KT, T, OPT, 3>), dim3(a.ntiles / DIRECT_CPW), dim3(TILE), 0, s, a); break;
    default: hipLaunchKernelGGL((k_direct<DT, KT, T, OPT, 4>), dim3(a.ntiles / DIRECT_CPW), dim3(TILE), 0, s, a); break;
  }
}

template <int DT, int KT, typename T>
static hipError_t launch_direct(const LaunchArgs& a, hipStream_t s) {
  if (a.ntiles % DIRECT_CPW) return hipErrorInvalidValue;  // the host pads to whole workgroups
  switch (a.op) {
    case OP_APPLY_SUM:
      hipLaunchKernelGGL((k_direct<DT, KT, T, OP_APPLY_SUM, 0>), dim3(a.ntiles / DIRECT_CPW), dim3(TILE), 0, s, a);
      break;
    case OP_APPLY_GRAD:
      hipLaunchKernelGGL((k_direct<DT, KT, T, OP_APPLY_GRAD, 0>), dim3(a.ntiles / DIRECT_CPW), dim3(TILE), 0, s, a);
      break;
    case OP_APPLY_WAVE_SUM:
      hipLaunchKernelGGL((k_direct<DT, KT, T, OP_APPLY_WAVE_SUM, 0>), dim3(a.ntiles / DIRECT_CPW), dim3(TILE), 0, s, a);
      break;
    case OP_RK_SUM: launch_direct_stage<DT, KT, T, OP_RK_SUM>(a, s); break;
    case OP_RK_WAVE_GRAD: launch_direct_stage<DT, KT, T, OP_RK_WAVE_GRAD>(a, s); break;
    case OP_RK_WAVE_SUM: launch_direct_stage<DT, KT, T, OP_RK_WAVE_SUM>(a, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Fused stage kernel: every class of a batch in one launch.  Items (workgroups) run the
// body the class's own launch would run on the same tiles, so the arithmetic is that of
// the per-class launches bit for bit.  The host interleaves the classes' items in runs
// of 8 (f.order), so the classes run side by side on every CU as on separate streams,
// and item 8 j + x of a class still runs on XCD x as in its own launch (the launch's
// item i runs on XCD i mod 8).
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int fused_lds_bytes() {
  int m = direct_lds_bytes<float, 4, 4>();
  for (int D = 2; D <= 4; ++D)
    for (int K = 2; K <= 4; ++K)
      if (seg_ok(D, K) && seg_cpw(D, K) > 1) m = cmax(m, seg_lds_bytes(D, K));
  return m;
}

template <int OPT, int STG>
__device__ __forceinline__ void fused_item(int body, const LaunchArgs& a, int wg, unsigned char* lds) {
  switch (body) {
#define GSG_FB_SEG(D, K) \
  case fused_body(KERNEL_SEG, D, K): seg_body<D, K, OPT, STG>(a, wg, -1, lds); return;
    case fused_body(KERNEL_DIRECT, 4, 4): direct_body<4, 4, float, OPT, STG>(a, wg, lds); return;
    GSG_FB_SEG(2, 2) GSG_FB_SEG(2, 3) GSG_FB_SEG(2, 4) GSG_FB_SEG(3, 2) GSG_FB_SEG(3, 3) GSG_FB_SEG(3, 4)
    GSG_FB_SEG(4, 2) GSG_FB_SEG(4, 3)
#undef GSG_FB_SEG
    default: return;
  }
}

template <int PASS, int STG>
__global__ __launch_bounds__(TILE) void k_fused(FusedArgs f) {
  __shared__ __align__(16) unsigned char lds[fused_lds_bytes()];
  {
    const int item = int(blockIdx.x);
    int s = 0, wg;
    if (f.order) {
      const int32_t rec = uni(sld(f.order + item));
      s = rec >> 24;
      wg = rec & 0xffffff;
    } else {
      while (s + 1 < f.nseg && item >= f.seg[s].wg_end) ++s;
      wg = item - (s ? f.seg[s - 1].wg_end : 0);
    }
    LaunchArgs a{};
    a.solves = f.solves;
    a.tiles = f.seg[s].tiles;
    a.ntiles = f.seg[s].ntiles;
    a.op = f.seg[s].op;
    a.stage = STG;
    const int body = f.seg[s].body;
    __syncthreads();
    if constexpr (PASS == FUSED_B) {
      fused_item<OP_RK_WAVE_SUM, STG>(body, a, wg, lds);
    } else if constexpr (PASS == FUSED_A) {
      if (a.op == OP_RK_WAVE_GRAD)
        fused_item<OP_RK_WAVE_GRAD, STG>(body, a, wg, lds);
      else
        fused_item<OP_RK_SUM, STG>(body, a, wg, lds);
    } else {
      fused_item<OP_RK_SUM, STG>(body, a, wg, lds);
    }
  }
}

template <int PASS>
static void launch_fused_pass(int stage, const FusedArgs& f, int grid, hipStream_t s) {
  switch (stage) {
    case 1: hipLaunchKernelGGL((k_fused<PASS, 1>), dim3(unsigned(grid)), dim3(TILE), 0, s, f); break;
    case 2: hipLaunchKernelGGL((k_fused<PASS, 2>), dim3(unsigned(grid)), dim3(TILE), 0, s, f); break;
    case 3: hipLaunchKernelGGL((k_fused<PASS, 3>), dim3(unsigned(grid)), dim3(TILE), 0, s, f); break;
    default: hipLaunchKernelGGL((k_fused<PASS, 4>), dim3(unsigned(grid)), dim3(TILE), 0, s, f); break;
  }
}

// A round of a transform sweep's plan entries: the first nvalid of NR entries, their
// operands loaded together so that a lane waits on one entry load and one operand load
// per round (the short lists of a sweep take one round).  The round is NR entries long for
// the whole wavefront, NR the largest nvalid of its lanes (xround_pick), so the wavefront
// issues no load instruction that none of its lanes needs; a lane with fewer entries
// loads its operands for the rest at offsets past the buffers' ends, which the hardware
// answers with 0 and no memory access.  No branch separates the loads (a branch makes the
// compiler wait on every load before the next).
template <int KT>
constexpr int xround_len() {
  return KT == 2 ? 8 : 4;
}
// the operands of entries (off[u], eb[u]), u < NR, applied to the NV vectors
template <int KT, typename TX, int NV, int NR>
__device__ __forceinline__ void xround_ops(const SolveDesc& S, __amdgpu_buffer_rsrc_t vr, const TX* xin, int64_t N,
                                           int v0, const uint32_t (&off)[xround_len<KT>()],
                                           const int32_t (&eb)[xround_len<KT>()], int nvalid, int md, uint32_t mb,
                                           int kd, double (&r)[NV]) {
  constexpr uint32_t es = sizeof(TX);
  const uint32_t vend = uint32_t(S.t.val_elems * int64_t(sizeof(double)));
  uint32_t xo[NR];
  double b[NR][KT];
#pragma unroll
  for (int u = 0; u < NR; ++u) {
    xo[u] = u < nvalid ? off[u] : uint32_t(N);  // N: past every vector's end
    const uint32_t vo = uint32_t(((int64_t(eb[u]) * KT + md) * KT) * int64_t(sizeof(double)));
#pragma unroll
    for (int mm = 0; mm < KT; ++mm)
      b[u][mm] = buf_load<double>(vr, u < nvalid ? vo + uint32_t(mm * sizeof(double)) : vend, 0u);
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const __amdgpu_buffer_rsrc_t xr = make_rsrc(xin + int64_t(v0 + v) * N, uint32_t(N * int64_t(es)));
    TX xv[NR][KT];
    if (KT == 2 && kd == 1) {
      // direction 0 at k = 2: a lane's two operands of an entry are adjacent, and aligned
      // to their pair (input cell offsets are multiples of 2^D, mb is even): one load
#pragma unroll
      for (int u = 0; u < NR; ++u) {
        const uint32_t o = (xo[u] + mb) * es;
        if constexpr (sizeof(TX) == 8) {
          typedef double d2 __attribute__((ext_vector_type(2)));
          const d2 q = __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(xr, int(o), 0, 0));
          xv[u][0] = q.x;
          xv[u][KT - 1] = q.y;
        } else {
          typedef float f2 __attribute__((ext_vector_type(2)));
          const f2 q = __builtin_bit_cast(f2, __builtin_amdgcn_raw_buffer_load_b64(xr, int(o), 0, 0));
          xv[u][0] = q.x;
          xv[u][KT - 1] = q.y;
        }
      }
    } else if (KT == 2 && kd < 64) {
      // k = 2, other directions with kd < 64: the lane whose mode digit d differs (lane ^ kd:
      // the same wavefront, the same cell, the same entries) needs the same two operands, so
      // each lane loads the one of its own mode and takes the other from that lane
#pragma unroll
      for (int u = 0; u < NR; ++u) {
        const TX own = buf_load<TX>(xr, (xo[u] + mb + uint32_t(md * kd)) * es, 0u);
        const TX other = __shfl_xor(own, kd);
        xv[u][0] = md == 0 ? own : other;
        xv[u][KT - 1] = md == 0 ? other : own;
      }
    } else {
#pragma unroll
      for (int u = 0; u < NR; ++u)
#pragma unroll
        for (int mm = 0; mm < KT; ++mm) xv[u][mm] = buf_load<TX>(xr, (xo[u] + mb + uint32_t(mm * kd)) * es, 0u);
    }
    // entries in list order, modes in order: the order of plan_dir's sums
    double acc = r[v];
#pragma unroll
    for (int u = 0; u < NR; ++u) {
#pragma unroll
      for (int mm = 0; mm < KT; ++mm) {
        const double t = acc + b[u][mm] * double(xv[u][mm]);
        acc = u < nvalid ? t : acc;
      }
    }
    r[v] = acc;
  }
}
// entries p, p + stride, ... of the plan list (p itself inside the list), then their operands
template <int KT, typename TX, int NV, int NR>
__device__ __forceinline__ void xround(const SolveDesc& S, __amdgpu_buffer_rsrc_t vr, const TX* xin, int64_t N, int v0,
                                       int32_t p, int32_t stride, int nvalid, int md, uint32_t mb, int kd,
                                       double (&r)[NV]) {
  constexpr int XR = xround_len<KT>();
  uint32_t off[XR];
  int32_t eb[XR];
#pragma unroll
  for (int u = 0; u < XR; ++u) {
    off[u] = 0;
    eb[u] = 0;
    if (u < NR) {  // compile time
      const int32_t q = u < nvalid ? p + u * stride : p;
      // the (offset, block) pair as one 64-bit global load (little endian: offset low)
      const uint64_t e = gld(reinterpret_cast<const uint64_t*>(S.t.plan_ent + 2 * q));
      off[u] = uint32_t(e);
      eb[u] = int32_t(uint32_t(e >> 32));
    }
  }
  xround_ops<KT, TX, NV, NR>(S, vr, xin, N, v0, off, eb, nvalid, md, mb, kd, r);
}
// the round for nw = the wavefront's largest nvalid (1 <= nw <= NR), one straight-line
// instantiation per length
template <int KT, typename TX, int NV, int NR>
__device__ __forceinline__ void xround_pick(int nw, const SolveDesc& S, __amdgpu_buffer_rsrc_t vr, const TX* xin,
                                            int64_t N, int v0, int32_t p, int32_t stride, int nvalid, int md,
                                            uint32_t mb, int kd, double (&r)[NV]) {
  if constexpr (NR > 1) {
    if (nw < NR) {
      xround_pick<KT, TX, NV, NR - 1>(nw, S, vr, xin, N, v0, p, stride, nvalid, md, mb, kd, r);
      return;
    }
  }
  xround<KT, TX, NV, NR>(S, vr, xin, N, v0, p, stride, nvalid, md, mb, kd, r);
}
// the same for entries already in registers (a row of the fixed-width table)
template <int KT, typename TX, int NV, int NR>
__device__ __forceinline__ void xround_ops_pick(int nw, const SolveDesc& S, __amdgpu_buffer_rsrc_t vr, const TX* xin,
                                                int64_t N, int v0, const uint32_t (&off)[xround_len<KT>()],
                                                const int32_t (&eb)[xround_len<KT>()], int nvalid, int md,
                                                uint32_t mb, int kd, double (&r)[NV]) {
  if constexpr (NR > 1) {
    if (nw < NR) {
      xround_ops_pick<KT, TX, NV, NR - 1>(nw, S, vr, xin, N, v0, off, eb, nvalid, md, mb, kd, r);
      return;
    }
  }
  xround_ops<KT, TX, NV, NR>(S, vr, xin, N, v0, off, eb, nvalid, md, mb, kd, r);
}
// the largest n over the wavefront's active lanes, for 1 <= n <= XR
template <int XR>
__device__ __forceinline__ int wave_max_count(int n) {
#pragma unroll
  for (int k = XR; k > 1; --k)
    if (__ballot(n >= k)) return k;
  return 1;
}

// The inputs of the classical RK4 stage update of DoF i (rk_stage_at), loaded before the
// sweep so the update waits on nothing but the sweep's sum.
__device__ __forceinline__ RkIn rk_stage_in(int stage, int64_t i, const double* y, const double* acc) {
  RkIn v{0.0, 0.0};
  if (stage <= 3) v.y = gld(y + i);
  if (stage >= 2) v.acc = gld(acc + i);
  return v;
}
// The classical RK4 stage update of DoF i from a right-hand-side value ri computed apart
// (the stage epilogue of the fused kernels, rk_update), applied by the last sweep of the
// Vlasov right-hand side; x receives the next stage input (stage < 4) or the new state's
// copy, and x32 (when not null) its float copy for an FP32 operator.
__device__ __forceinline__ void rk_stage_at(int stage, int64_t i, double dt, RkIn in, double* y, double* acc,
                                            double ri, double* x, float* x32) {
  double xn;
  if (stage == 1) {
    gst(acc + i, in.y + (dt / 6.0) * ri);
    xn = in.y + (dt / 2.0) * ri;
  } else if (stage == 2) {
    gst(acc + i, in.acc + (dt / 3.0) * ri);
    xn = in.y + (dt / 2.0) * ri;
  } else if (stage == 3) {
    gst(acc + i, in.acc + (dt / 3.0) * ri);
    xn = in.y + dt * ri;
  } else {
    xn = in.acc + (dt / 6.0) * ri;
    gst(y + i, xn);
  }
  gst(x + i, xn);
  if (x32) gst(x32 + i, float(xn));
}

// One direction sweep of a sparse-grid basis transform (the 1-D transform along
// direction S.dir0, each row cut to its fiber's admissible levels), FP64 arithmetic on
// TX inputs (float: the FP32 derivative output read directly), over vectors of S.N
// elements stored one after another.  NC = 0: vectors v0 .. v0 + NL - 1, v0 = NL blockIdx.y,
// go from xin + v N to yout + v N, each lane running the NL vectors on the entries it
// loaded once.  NC > 0 (the last n2p sweep of the Vlasov
// right-hand side, cb.nv = NC vectors): every lane runs the NC vectors for its DoF on the
// entries it loaded once, and writes only their combination
// -sum_{v<Dx} vpt_v y_v + sum_{v>=Dx} F_{v-Dx} y_v (the pointwise products of the
// steprule: -(sum_d v_d P_d) + (sum_d F_d Q_d), each sum in d order).  A tile with
// aux <= 1 is a DoF range, one lane per DoF.  A tile with aux = P > 1 is one cell with a
// long entry list (a coarse cell coupled to every finer nested cell): P lanes per DoF take
// every P-th entry and the partial sums are added in a fixed order through LDS.  ell (k = 2,
// when every short list of the sweep has at most XELL entries): the sweep direction's
// fixed-width table, XELL (offset, block) pairs a cell, the cell's list first and block -1
// after it, which the one-lane-per-DoF tiles read in place of the list bounds and entries.
template <int KT, typename TX, int NC, int NL = 1>
__global__ __launch_bounds__(TILE) void k_xsweep(LaunchArgs a, XComb cb, XRk rk, const int32_t* ell) {
  static_assert(NC == 0 || NL == 1, "a combination sweep runs its NC vectors");
  constexpr int NV = NC ? NC : NL;  // vectors a lane sums
  constexpr int XR = xround_len<KT>();
  __shared__ double part[NV * TILE];
  const Tile* tp = a.tiles + blockIdx.x;
  const SolveDesc& S = a.solves[0];
  const int D = S.D;
  const int d = S.dir0;
  int KD = 1, kd = 1;
  for (int e = 0; e < D; ++e) {
    if (e < d) kd *= KT;
    KD *= KT;
  }
  const int64_t N = S.N;
  const int64_t start = sld(&tp->start);
  const int64_t end = sld(&tp->end);
  const int parts = int(sld(&tp->aux));
  const int v0 = NC ? 0 : NL * int(blockIdx.y);
  const TX* xin = (const TX*)a.xin;
  const __amdgpu_buffer_rsrc_t vr = make_rsrc(S.t.val64, uint32_t(S.t.val_elems * int64_t(sizeof(double))));
  // the results at DoF i: the combination of the NC vectors' sums (each sum in vector
  // order, the pointwise factors loaded with the sums' operands), the RK stage update,
  // or the sum itself
  auto finish = [&](int64_t i, const double (&r)[NV], const double (&w)[NV], RkIn in) {
    if constexpr (NC > 0) {
      double c1 = 0.0, c2 = 0.0;
#pragma unroll
      for (int v = 0; v < NC; ++v) {
        const double t = w[v] * r[v];
        if (v < cb.Dx)
          c1 = v == 0 ? t : c1 + t;
        else
          c2 = v == cb.Dx ? t : c2 + t;
      }
      gst((double*)a.yout + i, cb.F ? -c1 + c2 : -c1);
    } else if (rk.stage) {
      rk_stage_at(rk.stage, i, rk.dt, in, rk.y, rk.acc, r[0], rk.x, rk.x32);
    } else {
#pragma unroll
      for (int v = 0; v < NV; ++v) gst((double*)a.yout + int64_t(v0 + v) * N + i, r[v]);
    }
  };
  auto pointwise = [&](int64_t i, double (&w)[NV]) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      if constexpr (NC > 0)
        w[v] = v < cb.Dx ? gld(cb.v + int64_t(v) * N + i) : (cb.F ? gld(cb.F + int64_t(v - cb.Dx) * N + i) : 0.0);
      else
        w[v] = 0.0;
    }
  };
  if (parts <= 1) {
    const int64_t i = start + threadIdx.x;
    if (i >= end) return;
    const int32_t g = int32_t(i / KD);
    const int ml = int(i - int64_t(g) * KD);
    const int md = (ml / kd) % KT;
    const uint32_t mb = uint32_t(ml - md * kd);
    double w[NV];
    pointwise(i, w);
    const RkIn in = rk.stage ? rk_stage_in(rk.stage, i, rk.y, rk.acc) : RkIn{0.0, 0.0};
    double r[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) r[v] = 0.0;
    if constexpr (XR == XELL) {
      if (ell) {
        // the cell's row of the direction's fixed-width table (its list, then block -1):
        // the entries load without waiting on the list bounds
        typedef unsigned int u4 __attribute__((ext_vector_type(4)));
        const __attribute__((address_space(1))) u4* row =
            (const __attribute__((address_space(1))) u4*)(ell + int64_t(g) * (2 * XELL));
        uint32_t off[XR];
        int32_t eb[XR];
        int nvalid = 0;
#pragma unroll
        for (int j = 0; j < XR / 2; ++j) {
          const u4 q = row[j];
          off[2 * j] = q.x;
          eb[2 * j] = int32_t(q.y);
          off[2 * j + 1] = q.z;
          eb[2 * j + 1] = int32_t(q.w);
        }
#pragma unroll
        for (int u = 0; u < XR; ++u) nvalid += eb[u] >= 0 ? 1 : 0;
        xround_ops_pick<KT, TX, NV, XR>(wave_max_count<XR>(nvalid), S, vr, xin, N, v0, off, eb, nvalid, md, mb, kd,
                                        r);
        finish(i, r, w, in);
        return;
      }
    }
    const int32_t* ptr = S.t.plan_ptr + int64_t(g) * D + d;
    const int32_t p0 = gld(ptr), p1 = gld(ptr + 1);
    for (int32_t pb = p0; pb < p1; pb += XR) {
      const int nvalid = p1 - pb < XR ? int(p1 - pb) : XR;
      xround_pick<KT, TX, NV, XR>(wave_max_count<XR>(nvalid), S, vr, xin, N, v0, pb, 1, nvalid, md, mb, kd, r);
    }
    finish(i, r, w, in);
    return;
  }
  if (end <= start) return;  // padding tile: the whole workgroup leaves before the barrier
  const int t = threadIdx.x;
  const int dof = t % KD;
  const int pp = t / KD;
  const int32_t g = int32_t(start / KD);
  const int md = (dof / kd) % KT;
  const uint32_t mb = uint32_t(dof - md * kd);
  const int32_t* ptr = S.t.plan_ptr + int64_t(g) * D + d;
  double r[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) r[v] = 0.0;
  double w[NV];
  RkIn in{0.0, 0.0};
  if (pp == 0) {
    pointwise(start + dof, w);
    if (rk.stage) in = rk_stage_in(rk.stage, start + dof, rk.y, rk.acc);
  }
  if (pp < parts) {
    const int32_t p0 = gld(ptr), p1 = gld(ptr + 1);
    for (int32_t pb = p0 + pp; pb < p1; pb += XR * parts) {
      const int32_t left = (p1 - pb + parts - 1) / parts;  // entries of this lane from pb on
      const int nvalid = left < XR ? int(left) : XR;
      xround_pick<KT, TX, NV, XR>(wave_max_count<XR>(nvalid), S, vr, xin, N, v0, pb, parts, nvalid, md, mb, kd, r);
    }
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) part[v * TILE + t] = r[v];
  __syncthreads();
  if (pp == 0) {
    double s[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      s[v] = part[v * TILE + dof];
      for (int q = 1; q < parts; ++q) s[v] += part[v * TILE + q * KD + dof];
    }
    finish(start + dof, s, w, in);
  }
}

__global__ void k_write_desc(SolveDesc* dst, SolveDesc v) {
  if (threadIdx.x == 0) *dst = v;
}

__global__ void k_to_f32(const double* __restrict__ x, float* __restrict__ y, int64_t n, int64_t nlo) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const float h = float(x[i]);
    y[i] = h;
    if (i < nlo) y[n + i] = float(x[i] - double(h));
  }
}
__global__ void k_to_f64(const float* __restrict__ x, double* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    y[i] = double(x[i]);
}

// float: the values as a hi and a lo table (A.v32, A.v32 + nnz), summed apart, added in double
template <typename T>
__global__ __launch_bounds__(256) void k_csr(DevCsr A, const T* __restrict__ x, double* __restrict__ y) {
  const int64_t row = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (row >= A.m) return;
  const T* v = vals_of_csr<T>(A);
  T acc = T(0), acl = T(0);
  for (int64_t p = A.rp[row]; p < A.rp[row + 1]; ++p) {
    const T xv = x[A.ci[p]];
    acc += v[p] * xv;
    if constexpr (sizeof(T) == 4) acl += v[A.nnz + p] * xv;
  }
  y[row] = sizeof(T) == 4 ? double(acc) + double(acl) : double(acc);
}

}  // namespace

template <int DT, int KT>
static hipError_t launch_tk(bool fp32, const LaunchArgs& a, hipStream_t s) {
  if (a.ntiles <= 0) return hipSuccess;
  if constexpr (DT > 0 && KT > 0 && direct_ok(DT, KT)) {
    if (a.kernel == KERNEL_DIRECT) return fp32 ? launch_direct<DT, KT, float>(a, s) : launch_direct<DT, KT, double>(a, s);
  }
  if constexpr (direct64_ok(DT, KT) && !direct_ok(DT, KT)) {
    if (a.kernel == KERNEL_DIRECT) return fp32 ? hipErrorInvalidValue : launch_direct<DT, KT, double>(a, s);
  }
  if (a.kernel == KERNEL_DIRECT) return hipErrorInvalidValue;  // no direct instantiation for this shape
  if constexpr (DT > 0 && KT > 0 && seg_ok(DT, KT)) {
    if (a.kernel == KERNEL_SEG) {
      if (fp32) return launch_seg<DT, KT>(a, s);
      if constexpr (seg64_ok(DT, KT)) return launch_seg64<DT, KT>(a, s);
      return hipErrorInvalidValue;
    }
    if (a.kernel == KERNEL_SEG_ALT) {
      if (a.op == OP_APPLY_SUM || a.op == OP_APPLY_GRAD || a.op == OP_APPLY_WAVE_SUM) {
        LaunchArgs b = a;  // the apply operations run the rule's entries per round
        b.kernel = KERNEL_SEG;
        return fp32 ? launch_seg<DT, KT>(b, s) : hipErrorInvalidValue;
      }
      return fp32 ? launch_seg_alt<DT, KT>(a, s) : hipErrorInvalidValue;
    }
  }
  if (a.kernel == KERNEL_SEG || a.kernel == KERNEL_SEG_ALT) return hipErrorInvalidValue;  // float, seg_ok shapes only
  if constexpr (DT > 0 && KT > 0 && fiber_ok(DT, KT)) {
    if (a.kernel == KERNEL_FIBER && a.lds_bytes > 0) {
      // the float operator (hi and lo tables) runs the plan, direct or segmented kernels
      if (fp32) return hipErrorInvalidValue;
      hipLaunchKernelGGL((k_fiber<DT, KT, double>), dim3(a.ntiles), dim3(TILE), a.lds_bytes, s, a);
      return hipGetLastError();
    }
  }
  if (a.kernel != KERNEL_PLAN) return hipErrorInvalidValue;  // tile-plan tiles need the fiber kernel
  if (fp32)
    hipLaunchKernelGGL((k_plan<DT, KT, float>), dim3(a.ntiles), dim3(TILE), 0, s, a);
  else
    hipLaunchKernelGGL((k_plan<DT, KT, double>), dim3(a.ntiles), dim3(TILE), 0, s, a);
  return hipGetLastError();
}

template <int DT>
static hipError_t launch_d(int K, bool fp32, const LaunchArgs& a, hipStream_t s) {
  switch (K) {
    case 1: return launch_tk<DT, 1>(fp32, a, s);
    case 2: return launch_tk<DT, 2>(fp32, a, s);
    case 3: return launch_tk<DT, 3>(fp32, a, s);
    case 4: return launch_tk<DT, 4>(fp32, a, s);
    case 5: return launch_tk<DT, 5>(fp32, a, s);
    default: return launch_tk<0, 0>(fp32, a, s);
  }
}

hipError_t launch_op(int D, int K, bool fp32, const LaunchArgs& a, hipStream_t s) {
  switch (D) {
    case 1: return launch_d<1>(K, fp32, a, s);
    case 2: return launch_d<2>(K, fp32, a, s);
    case 3: return launch_d<3>(K, fp32, a, s);
    case 4: return launch_d<4>(K, fp32, a, s);
    case 5: return launch_d<5>(K, fp32, a, s);
    case 6: return launch_d<6>(K, fp32, a, s);
    default: return launch_tk<0, 0>(fp32, a, s);
  }
}

static dim3 vec_grid(int64_t n) {
  const int64_t blocks = (n + 255) / 256;
  return dim3(unsigned(blocks < 65536 ? blocks : 65536));
}
template <typename TX, int NC, int NL = 1>
static hipError_t launch_xsweep_t(int K, int nvec, const LaunchArgs& a, const XComb& cb, const XRk& rk,
                                  const int32_t* ell, hipStream_t s) {
  const dim3 grid(unsigned(a.ntiles), unsigned(NC ? 1 : nvec / NL));
  switch (K) {
    case 2: hipLaunchKernelGGL((k_xsweep<2, TX, NC, NL>), grid, dim3(TILE), 0, s, a, cb, rk, ell); break;
    case 3: hipLaunchKernelGGL((k_xsweep<3, TX, NC, NL>), grid, dim3(TILE), 0, s, a, cb, rk, nullptr); break;
    case 5: hipLaunchKernelGGL((k_xsweep<5, TX, NC, NL>), grid, dim3(TILE), 0, s, a, cb, rk, nullptr); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_xsweep(int K, int nvec, const LaunchArgs& a, hipStream_t s, bool in_f32, const XComb* comb,
                         const XRk* rk, const int32_t* ell) {
  static_assert(xround_len<2>() == XELL, "a row of the fixed-width table is one round");
  if (a.ntiles <= 0 || nvec <= 0) return hipSuccess;
  if (ell && K != 2) return hipErrorInvalidValue;  // the table exists for k = 2 only
  const XComb none{nullptr, nullptr, 0, 1};
  const XRk nork{0, 0.0, nullptr, nullptr, nullptr, nullptr};
  if (rk && rk->stage && (comb || nvec != 1)) return hipErrorInvalidValue;  // one plain vector only
  const XRk& r = rk ? *rk : nork;
  if (comb) {
    // the combination reads FP64 sweep results of Dx or 2 Dx vectors (Dx <= 4)
    if (comb->nv != nvec || in_f32) return hipErrorInvalidValue;
    switch (nvec) {
      case 1: return launch_xsweep_t<double, 1>(K, nvec, a, *comb, r, ell, s);
      case 2: return launch_xsweep_t<double, 2>(K, nvec, a, *comb, r, ell, s);
      case 3: return launch_xsweep_t<double, 3>(K, nvec, a, *comb, r, ell, s);
      case 4: return launch_xsweep_t<double, 4>(K, nvec, a, *comb, r, ell, s);
      case 6: return launch_xsweep_t<double, 6>(K, nvec, a, *comb, r, ell, s);
      case 8: return launch_xsweep_t<double, 8>(K, nvec, a, *comb, r, ell, s);
      default: return hipErrorInvalidValue;
    }
  }
  if (nvec % 2 == 0) {  // vector pairs: a lane runs both on the entries it loaded once
    return in_f32 ? launch_xsweep_t<float, 0, 2>(K, nvec, a, none, r, ell, s)
                  : launch_xsweep_t<double, 0, 2>(K, nvec, a, none, r, ell, s);
  }
  return in_f32 ? launch_xsweep_t<float, 0>(K, nvec, a, none, r, ell, s)
                : launch_xsweep_t<double, 0>(K, nvec, a, none, r, ell, s);
}

hipError_t launch_fused(int pass, int stage, const FusedArgs& f, int grid, hipStream_t s) {
  if (f.items <= 0) return hipSuccess;
  if (grid <= 0 || grid > f.items || f.nseg <= 0 || f.nseg > FUSED_MAXSEG) return hipErrorInvalidValue;
  for (int i = 0; i < f.nseg; ++i) {  // every item must lie inside its class's tiles
    const FusedSeg& g = f.seg[i];
    const int wgs = g.wg_end - (i ? f.seg[i - 1].wg_end : 0);
    if (!fused_ok(g.body / 9, 2 + (g.body / 3) % 3, 2 + g.body % 3) || wgs < 0) return hipErrorInvalidValue;
    const int kernel = g.body / 9, D = 2 + (g.body / 3) % 3, K = 2 + g.body % 3;
    const int per = kernel == KERNEL_DIRECT ? DIRECT_CPW : kernel == KERNEL_SEG ? SEG_WAVES * seg_cpw(D, K) : 1;
    if (int64_t(wgs) * per != g.ntiles) return hipErrorInvalidValue;
  }
  if (f.seg[f.nseg - 1].wg_end != f.items) return hipErrorInvalidValue;
  if (grid != f.items) return hipErrorInvalidValue;  // one item a workgroup
  switch (pass) {
    case FUSED_SUM: launch_fused_pass<FUSED_SUM>(stage, f, grid, s); break;
    case FUSED_A: launch_fused_pass<FUSED_A>(stage, f, grid, s); break;
    case FUSED_B: launch_fused_pass<FUSED_B>(stage, f, grid, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_write_desc(SolveDesc* dst, const SolveDesc& v, hipStream_t s) {
  hipLaunchKernelGGL(k_write_desc, dim3(1), dim3(64), 0, s, dst, v);
  return hipGetLastError();
}

hipError_t launch_to_f32(const double* x, float* y, int64_t n, hipStream_t s, int64_t nlo) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_to_f32, dim3(unsigned(blocks < 65536 ? blocks : 65536)), dim3(256), 0, s, x, y, n, nlo);
  return hipGetLastError();
}
hipError_t launch_to_f64(const float* x, double* y, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_to_f64, dim3(unsigned(blocks < 65536 ? blocks : 65536)), dim3(256), 0, s, x, y, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Projection of initial data (vcoeffs_DG, dg_vmethods.jl:149-179): tensor Gauss-Legendre
// nodes per cell, then one contraction pass per direction.  The node values of a cell are
// stored with the direction-1 point index fastest; each pass contracts the slowest point
// axis of its input (reads coalesced across the remaining axes) and appends the mode axis
// fastest, so after the passes for directions D, D-1, ..., 1 a cell holds its k^D
// coefficients with the direction-1 mode fastest: the vhier order of a cell's modes.

// X[d ld + i] = direction-d coordinate of point i of cells [cell0, cell0 + npts / ppc)
__global__ void __launch_bounds__(256) k_proj_nodes(ProjArgs a, const double* __restrict__ t, int64_t cell0,
                                                    int64_t npts, double* __restrict__ X, int64_t ld) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < npts; i += stride) {
    const int64_t cl = i / a.ppc;
    uint32_t j = uint32_t(i - cl * a.ppc);
    const int64_t cell = cell0 + cl;
    const int32_t b = a.cell_blk[cell];
    uint32_t loc = uint32_t(cell - a.blk_cell0[b]);  // cell within its block, direction 1 fastest
    for (int d = 0; d < a.D; ++d) {
      const int l = a.blk_lev[b * a.D + d];
      const int sh = l > 0 ? l - 1 : 0;
      const uint32_t c = loc & ((1u << sh) - 1u);
      loc >>= sh;
      const uint32_t p = j % uint32_t(a.P);
      j /= uint32_t(a.P);
      const bool hi_half = p >= uint32_t(a.q);
      const double h = 1.0 / double(1u << sh);  // a power of two: the cell ends are exact
      const double lo = double(c) * h, hi = double(c + 1) * h, mid = 0.5 * (lo + hi);
      const double pa = hi_half ? mid : lo, pb = hi_half ? hi : mid;
      X[d * ld + i] = 0.5 * (pa + pb) + 0.5 * (pb - pa) * t[hi_half ? p - a.q : p];
    }
  }
}

// X1[(d ncell + c) P + p]: the direction-d coordinate of node p of cell cell0 + c (the
// per-direction factors of the tensor nodes, for callers that evaluate f by broadcasting)
__global__ void __launch_bounds__(256) k_proj_nodes_1d(ProjArgs a, const double* __restrict__ t, int64_t cell0,
                                                       int64_t ncell, double* __restrict__ X1) {
  const int64_t total = int64_t(a.D) * ncell * a.P;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < total; i += stride) {
    const uint32_t p = uint32_t(i % a.P);
    const int64_t dc = i / a.P;
    const int d = int(dc / ncell);
    const int64_t cell = cell0 + (dc - int64_t(d) * ncell);
    const int32_t b = a.cell_blk[cell];
    uint32_t loc = uint32_t(cell - a.blk_cell0[b]);
    uint32_t c = 0;
    int sh = 0;
    for (int e = 0; e <= d; ++e) {
      const int l = a.blk_lev[b * a.D + e];
      sh = l > 0 ? l - 1 : 0;
      c = loc & ((1u << sh) - 1u);
      loc >>= sh;
    }
    const bool hi_half = p >= uint32_t(a.q);
    const double h = 1.0 / double(1u << sh);
    const double lo = double(c) * h, hi = double(c + 1) * h, mid = 0.5 * (lo + hi);
    const double pa = hi_half ? mid : lo, pb = hi_half ? hi : mid;
    X1[i] = 0.5 * (pa + pb) + 0.5 * (pb - pa) * t[hi_half ? p - a.q : p];
  }
}

hipError_t launch_proj_nodes_1d(const ProjArgs& a, const double* t, int64_t cell0, int64_t ncell, double* X1,
                                hipStream_t s) {
  const int64_t n = int64_t(a.D) * ncell * a.P;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_proj_nodes_1d, vec_grid(n), dim3(256), 0, s, a, t, cell0, ncell, X1);
  return hipGetLastError();
}

template <int K>
__device__ __forceinline__ void proj_row(const double* __restrict__ w, int P, const double* __restrict__ src,
                                         int64_t R, double (&acc)[K]) {
  for (int p = 0; p < P; ++p) {
    const double v = src[p * R];
#pragma unroll
    for (int m = 0; m < K; ++m) acc[m] = fma(v, w[m * P + p], acc[m]);
  }
}

// One pass for direction d over cells [cell0, cell0 + nrows / R): per cell the input is
// [P][R] and the output [R][K], out[i K + m] = sum_p in[cell, p, r] wv[level_d(cell)][m][p]
// for row i = cell R + r.
template <int K>
__global__ void __launch_bounds__(256) k_proj_pass(ProjArgs a, const double* __restrict__ wv, int d, int64_t cell0,
                                                   int64_t nrows, int64_t R, const double* __restrict__ in,
                                                   double* __restrict__ out) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < nrows; i += stride) {
    const int64_t cl = i / R, r = i - cl * R;
    const int32_t b = a.cell_blk[cell0 + cl];
    const int l = a.blk_lev[b * a.D + d];
    const double* src = in + cl * a.P * R + r;
    double acc[K];
#pragma unroll
    for (int m = 0; m < K; ++m) acc[m] = 0.0;
    const int l0 = __builtin_amdgcn_readfirstlane(l);
    if (__all(l == l0))  // one level across the wavefront: the table row is read through the scalar cache
      proj_row<K>(wv + int64_t(l0) * K * a.P, a.P, src, R, acc);
    else
      proj_row<K>(wv + int64_t(l) * K * a.P, a.P, src, R, acc);
    double* dst = out + i * K;
#pragma unroll
    for (int m = 0; m < K; ++m) dst[m] = acc[m];
  }
}

// tensor_construct (tensor_construct.jl:57-63): out[i] = prod_d vecs[d][1-D index of DoF
// i in direction d], the product taken in direction order from 1.0
__global__ void __launch_bounds__(256) k_tensor(ProjArgs a, TensorVecs tv, int64_t N, double* __restrict__ out) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  int64_t KD = 1;
  for (int d = 0; d < a.D; ++d) KD *= a.K;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < N; i += stride) {
    const int64_t cell = i / KD;
    uint32_t m = uint32_t(i - cell * KD);
    const int32_t b = a.cell_blk[cell];
    uint32_t loc = uint32_t(cell - a.blk_cell0[b]);
    double v = 1.0;
    for (int d = 0; d < a.D; ++d) {
      const int l = a.blk_lev[b * a.D + d];
      const int sh = l > 0 ? l - 1 : 0;
      const uint32_t c = loc & ((1u << sh) - 1u);
      loc >>= sh;
      const uint32_t md = m % uint32_t(a.K);
      m /= uint32_t(a.K);
      const uint32_t base = l == 0 ? 0u : (1u << (l - 1));
      v = v * tv.v[d][int64_t(a.K) * (base + c) + md];
    }
    out[i] = v;
  }
}

hipError_t launch_tensor(const ProjArgs& a, const TensorVecs& tv, int64_t N, double* out, hipStream_t s) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tensor, vec_grid(N), dim3(256), 0, s, a, tv, N, out);
  return hipGetLastError();
}

hipError_t launch_proj_nodes(const ProjArgs& a, const double* t, int64_t cell0, int64_t ncell, double* X, int64_t ld,
                             hipStream_t s) {
  const int64_t n = ncell * a.ppc;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_proj_nodes, vec_grid(n), dim3(256), 0, s, a, t, cell0, n, X, ld);
  return hipGetLastError();
}

hipError_t launch_proj_pass(const ProjArgs& a, const double* wv, int d, int64_t cell0, int64_t ncell, int64_t R,
                            const double* in, double* out, hipStream_t s) {
  const int64_t n = ncell * R;
  if (n <= 0) return hipSuccess;
#define GSG_PROJ_CASE(KK)                                                                                  \
  case KK:                                                                                                 \
    hipLaunchKernelGGL((k_proj_pass<KK>), vec_grid(n), dim3(256), 0, s, a, wv, d, cell0, n, R, in, out); \
    break;
  switch (a.K) {
    GSG_PROJ_CASE(1) GSG_PROJ_CASE(2) GSG_PROJ_CASE(3) GSG_PROJ_CASE(4) GSG_PROJ_CASE(5)
    GSG_PROJ_CASE(6) GSG_PROJ_CASE(7) GSG_PROJ_CASE(8) GSG_PROJ_CASE(9) GSG_PROJ_CASE(10)
    default: return hipErrorInvalidValue;
  }
#undef GSG_PROJ_CASE
  return hipGetLastError();
}

hipError_t launch_csr(const DevCsr& A, bool fp32, const void* x, void* y, hipStream_t s) {
  if (A.m <= 0) return hipSuccess;
  const int64_t blocks = (A.m + 255) / 256;
  if (fp32)
    hipLaunchKernelGGL((k_csr<float>), dim3(unsigned(blocks)), dim3(256), 0, s, A, (const float*)x, (double*)y);
  else
    hipLaunchKernelGGL((k_csr<double>), dim3(unsigned(blocks)), dim3(256), 0, s, A, (const double*)x, (double*)y);
  return hipGetLastError();
}

}  // namespace gsg
