// gsg_kernels.hip -- gfx950 kernels of the sparse-grid DG derivative operator and
// its fused classical-RK4 stages.
//
// One thread owns one output DoF (L', C', m') of one solve and computes, for every
// direction d, the row of D_d restricted to the sparse grid: the 1-D operator row of
// the output's 1-D coordinate (l'_d, c'_d, m'_d) applied along the fiber through the
// DoF.  The fiber's admissible input levels are 0..n - sum_{e != d} l'_e, which is a
// prefix of the 1-D row (the row's entries are sorted by input level), so the sparse
// cutoff of multidim_derivative.jl:47 costs nothing.  Threads of a workgroup cover
// consecutive DoFs of the vhier vector, so for d >= 2 the k input values of one
// coupled cell are contiguous across lanes (coalesced) and for d = 1 groups of k lanes
// read the same addresses (broadcast).  The stage update of RK4 is fused into the
// epilogue, so one stage of transport is one launch that reads the stage input once
// from HBM and writes the next stage input and the accumulator.
#include <hip/hip_runtime.h>

#include "gsg_dev.hpp"

namespace gsg {

namespace {

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

// Decoded coordinates of one output DoF inside its level block.
struct Dof {
  int64_t cl;  // cell linear index (dimension 1 fastest)
  int64_t ml;  // mode linear index (dimension 1 fastest)
  int blk;
  int sumL;
  int lv[MAXD];
  int sh[MAXD];   // log2(cells) per dimension
  int cum[MAXD];  // bit offset of the cell coordinate in cl
};

template <int DT>
__device__ __forceinline__ void decode(const SolveDesc& S, int64_t i, int blk, int KD, Dof& q) {
  const int D = DT ? DT : S.D;
  while (i >= S.t.blk_off[blk + 1]) ++blk;
  const int64_t r = i - S.t.blk_off[blk];
  q.blk = blk;
  q.cl = r / KD;
  q.ml = r - q.cl * KD;
  int acc_sh = 0, sumL = 0;
#pragma unroll
  for (int e = 0; e < (DT ? DT : MAXD); ++e) {
    if (e >= D) break;
    const int l = S.t.blk_lev[blk * D + e];
    q.lv[e] = l;
    q.sh[e] = l > 0 ? l - 1 : 0;
    q.cum[e] = acc_sh;
    acc_sh += q.sh[e];
    sumL += l;
  }
  q.sumL = sumL;
}

// (D_d x)[DoF q], restricted to the sparse grid.  kd = K^d.
template <int KT, typename T>
__device__ __forceinline__ T apply_dir(const SolveDesc& S, const T* __restrict__ x, const Dof& q, int d, int D,
                                       int K, int KD, int kd) {
  const int n = S.n;
  const int ld = q.lv[d];
  const int mf = S.full ? n : n - (q.sumL - ld);
  const int64_t cmask = (int64_t(1) << q.sh[d]) - 1;
  const int cd = int((q.cl >> q.cum[d]) & cmask);
  const int64_t r1 = (ld == 0 ? 0 : (int64_t(1) << (ld - 1))) + cd;
  const int32_t* rp = S.t.row_ptr + r1 * (n + 2);
  const int e0 = rp[0], e1 = rp[mf + 1];
  const int64_t clo = q.cl & ((int64_t(1) << q.cum[d]) - 1);
  const int64_t chi = q.cl >> (q.cum[d] + q.sh[d]);
  const int64_t mq = q.ml / kd;
  const int64_t mlo = q.ml - mq * kd;
  const int md = int(mq % K);
  const int64_t mhi = mq / K;
  const int64_t mbase = mlo + int64_t(kd) * K * mhi;
  const int64_t* inoff = S.t.blk_inoff + (int64_t(q.blk) * D + d) * (n + 1);
  const T* vals = vals_of<T>(S.t);
  T acc = T(0);
  for (int e = e0; e < e1; ++e) {
    const int32_t pk = S.t.ent[e];
    const int li = pk >> 24;
    const int64_t ci = pk & 0xFFFFFF;
    const int sl = li > 0 ? li - 1 : 0;
    const int64_t cin = clo + ((ci + (chi << sl)) << q.cum[d]);
    const T* xp = x + inoff[li] + cin * KD + mbase;
    const T* vp = vals + (int64_t(e) * K + md) * K;
#pragma unroll
    for (int mm = 0; mm < (KT ? KT : 16); ++mm) {
      if (mm >= K) break;
      acc += vp[mm] * xp[int64_t(mm) * kd];
    }
  }
  return acc;
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
template <typename T>
__device__ __forceinline__ void rk_update(const SolveDesc& S, int stage, int64_t idx, double r) {
  const double dt = S.dt;
  T* out = stage_out<T>(S, stage);
  if (stage == 1) {
    const double yi = S.y[idx];
    S.acc[idx] = yi + (dt / 6.0) * r;
    out[idx] = T(yi + (dt / 2.0) * r);
  } else if (stage == 2) {
    const double yi = S.y[idx];
    S.acc[idx] = S.acc[idx] + (dt / 3.0) * r;
    out[idx] = T(yi + (dt / 2.0) * r);
  } else if (stage == 3) {
    const double yi = S.y[idx];
    S.acc[idx] = S.acc[idx] + (dt / 3.0) * r;
    out[idx] = T(yi + dt * r);
  } else {
    const double yn = S.acc[idx] + (dt / 6.0) * r;
    S.y[idx] = yn;
    if (out) out[idx] = T(yn);
  }
}

template <int DT, int KT, typename T>
__global__ __launch_bounds__(TILE) void k_op(LaunchArgs a) {
  const Tile tile = a.tiles[blockIdx.x];
  const SolveDesc& S = a.solves[tile.solve];
  if (!S.active) return;
  const int64_t i = tile.start + threadIdx.x;
  if (i >= tile.end) return;
  const int D = DT ? DT : S.D;
  const int K = KT ? KT : S.K;
  int kpow[MAXD + 1];
  kpow[0] = 1;
#pragma unroll
  for (int e = 0; e < (DT ? DT : MAXD); ++e) {
    if (e >= D) break;
    kpow[e + 1] = kpow[e] * K;
  }
  const int KD = kpow[D];
  Dof q;
  decode<DT>(S, i, tile.blk, KD, q);
  const int64_t N = S.N;

  if (a.op == OP_APPLY_SUM || a.op == OP_RK_SUM) {
    const T* x0 = a.op == OP_RK_SUM ? stage_in<T>(S, a.stage) : (const T*)a.xin;
    const int64_t stride = a.op == OP_RK_SUM ? 0 : a.instride;
    T s = T(0);
    int t = 0;
#pragma unroll
    for (int d = 0; d < (DT ? DT : MAXD); ++d) {
      if (d >= D) break;
      if (d < S.dir0 || d >= S.dir0 + S.ndir) continue;
      s += T(S.coef[d]) * apply_dir<KT, T>(S, x0 + t * stride, q, d, D, K, KD, kpow[d]);
      ++t;
    }
    if (a.op == OP_APPLY_SUM)
      ((double*)a.yout)[i] = double(s);
    else
      rk_update<T>(S, a.stage, i, double(s));
    return;
  }
  if (a.op == OP_APPLY_GRAD || a.op == OP_RK_WAVE_GRAD) {
    const T* x0 = a.op == OP_RK_WAVE_GRAD ? stage_in<T>(S, a.stage) : (const T*)a.xin;
    T* w = a.op == OP_RK_WAVE_GRAD ? (T*)S.w : (T*)a.yout;
#pragma unroll
    for (int d = 0; d < (DT ? DT : MAXD); ++d) {
      if (d >= D) break;
      w[int64_t(d) * N + i] = apply_dir<KT, T>(S, x0, q, d, D, K, KD, kpow[d]);
    }
    return;
  }
  // wave second pass: [u; v]' = [v; sum_d D_d w_d]
  const T* w = a.op == OP_RK_WAVE_SUM ? (const T*)S.w : (const T*)a.xin2;
  T s = T(0);
#pragma unroll
  for (int d = 0; d < (DT ? DT : MAXD); ++d) {
    if (d >= D) break;
    s += apply_dir<KT, T>(S, w + int64_t(d) * N, q, d, D, K, KD, kpow[d]);
  }
  if (a.op == OP_APPLY_WAVE_SUM) {
    const T* x = (const T*)a.xin;
    ((double*)a.yout)[i] = double(x[N + i]);
    ((double*)a.yout)[N + i] = double(s);
  } else {
    const T* xin = stage_in<T>(S, a.stage);
    const double rv = double(xin[N + i]);
    rk_update<T>(S, a.stage, i, rv);
    rk_update<T>(S, a.stage, N + i, double(s));
  }
}

__global__ void k_to_f32(const double* __restrict__ x, float* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    y[i] = float(x[i]);
}
__global__ void k_to_f64(const float* __restrict__ x, double* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    y[i] = double(x[i]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_csr(DevCsr A, const T* __restrict__ x, double* __restrict__ y) {
  const int64_t row = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (row >= A.m) return;
  const T* v = vals_of_csr<T>(A);
  T acc = T(0);
  for (int64_t p = A.rp[row]; p < A.rp[row + 1]; ++p) acc += v[p] * x[A.ci[p]];
  y[row] = double(acc);
}

}  // namespace

template <int DT, int KT>
static hipError_t launch_tk(bool fp32, const LaunchArgs& a, hipStream_t s) {
  if (a.ntiles <= 0) return hipSuccess;
  if (fp32)
    hipLaunchKernelGGL((k_op<DT, KT, float>), dim3(a.ntiles), dim3(TILE), 0, s, a);
  else
    hipLaunchKernelGGL((k_op<DT, KT, double>), dim3(a.ntiles), dim3(TILE), 0, s, a);
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

hipError_t launch_to_f32(const double* x, float* y, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_to_f32, dim3(unsigned(blocks < 65536 ? blocks : 65536)), dim3(256), 0, s, x, y, n);
  return hipGetLastError();
}
hipError_t launch_to_f64(const float* x, double* y, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_to_f64, dim3(unsigned(blocks < 65536 ? blocks : 65536)), dim3(256), 0, s, x, y, n);
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
