// gsg_setup.cpp -- FP64 host setup (basis, 1-D operator blocks, layout tables).
//
// Arithmetic follows the reference's order where it affects rounding: the basis is
// built by the same classical Gram-Schmidt sequence (dg_basis.jl:68-191), the flux
// traces are evaluated at the same +-5e-16 offsets (1d_derivative.jl:52-74), and
// the products Q'(AQ) accumulate in ascending inner index like SparseArrays'
// spmatmul.  No FMA contraction is allowed in this file.
#pragma clang fp contract(off)

#include "gsg_setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>

namespace gsg {

namespace {

// <x^i or sgn(x)x^i, x^j or sgn(x)x^j> on [-1,1]; coordinates 0..h-1 are x^i, h..2h-1
// are sgn(x) x^(i-h), with h = len/2 (dg_basis.jl:21-35).
double gram_entry(int i, int j, int len) {
  const int h = len / 2;
  auto par = [](int e) { return (e % 2 == 0) ? 1.0 : -1.0; };
  if (i < h && j >= h) std::swap(i, j);
  if (i < h) return (1.0 + par(i + j)) / double(1 + i + j);
  if (j < h) return (1.0 - par((i - h) + j)) / double(1 + (i - h) + j);
  return (1.0 + par((i - h) + (j - h))) / double(1 + (i - h) + (j - h));
}

using Vec = std::vector<double>;

double dot(const Vec& a, const Vec& b) {  // dg_basis.jl:37-50 (zero terms skipped)
  const int len = (int)a.size();
  double s = 0.0;
  for (int i = 0; i < len; ++i) {
    if (a[i] == 0.0) continue;
    for (int j = 0; j < len; ++j) {
      if (b[j] == 0.0) continue;
      s += gram_entry(i, j, len) * a[i] * b[j];
    }
  }
  return s;
}

double dot_unit(const Vec& a, int j1) {  // dg_basis.jl:52-60, j1 1-based
  const int len = (int)a.size();
  double s = 0.0;
  for (int i = 0; i < len; ++i) s += gram_entry(i, j1 - 1, len) * a[i];
  return s;
}

void axpy_neg(Vec& y, double a, const Vec& x) {
  for (size_t t = 0; t < y.size(); ++t) y[t] -= a * x[t];
}
void scale_inv(Vec& y, double s) {
  for (auto& v : y) v /= s;
}

// Classical Gram-Schmidt in the given order (dg_basis.jl:68-85).
std::vector<Vec> orthonormalise(const std::vector<Vec>& in) {
  std::vector<Vec> out = in;
  for (size_t i = 0; i < in.size(); ++i) {
    for (size_t j = 0; j < i; ++j) axpy_neg(out[i], dot(in[i], out[j]) / dot(out[j], out[j]), out[j]);
    scale_inv(out[i], std::sqrt(dot(out[i], out[i])));
  }
  return out;
}

std::vector<Vec> legendre_rows(int deg) {  // dg_basis.jl:88-94
  std::vector<Vec> e(deg + 1, Vec(2 * (deg + 1), 0.0));
  for (int j = 0; j <= deg; ++j) e[j][j] = 1.0;
  return orthonormalise(e);
}

std::vector<Vec> multiwavelets(int k) {  // dg_basis.jl:104-191
  const int len = 2 * k;
  std::vector<Vec> q(k, Vec(len, 0.0));
  for (int j = 0; j < k; ++j) q[j][k + j] = 1.0;  // sgn(x) x^j
  // remove the polynomial part (orthogonalize_1)
  const std::vector<Vec> leg = legendre_rows(k - 1);
  std::vector<Vec> q1 = q;
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) axpy_neg(q1[i], dot(q[i], leg[j]) / dot(leg[j], leg[j]), leg[j]);
  // extra vanishing moments (orthogonalize_2): uses the unmodified rows of q1
  std::vector<Vec> q2 = q1;
  for (int i = 1; i <= k - 1; ++i) {
    const Vec& fi = q1[i - 1];
    for (int j = i + 1; j <= k; ++j) axpy_neg(q2[j - 1], dot_unit(q2[j - 1], k + i) / dot_unit(fi, k + i), fi);
  }
  // reverse Gram-Schmidt (gram_schmidt_rev)
  std::vector<Vec> out(k, Vec(len, 0.0));
  for (int i = k - 1; i >= 0; --i) {
    const Vec& fi = q2[i];
    out[i] = fi;
    for (int j = i + 1; j < k; ++j) axpy_neg(out[i], dot(fi, out[j]) / dot(out[j], out[j]), out[j]);
    scale_inv(out[i], std::sqrt(dot(out[i], out[i])));
  }
  return out;
}

// array2poly (1d_dg_functions.jl:15-28): Horner over the two coefficient halves,
// zero outside [-1, 1].
double pair_eval(const double* v, int h, double x) {
  if (std::fabs(x) > 1.0) return 0.0;
  double s = 0.0;
  for (int i = h - 1; i >= 0; --i) {
    s *= x;
    const double odd = std::signbit(x) ? -v[i + h] : v[i + h];
    s += v[i] + odd;
  }
  return s;
}

struct Basis {
  int k;
  const double* leg;  // (KMAX+1) x 2(KMAX+1)
  std::vector<double> dg;  // k x 2k
  double legP(int deg, double x) const { return pair_eval(leg + deg * 2 * (KMAX + 1), KMAX + 1, x); }
  // v(k, level, cell, mode, x), cell/mode 1-based (dg_methods.jl:27-36)
  double hier(int level, int cell, int mode, double x) const {
    if (level == 0) return legP(mode - 1, 2 * x - 1) * std::sqrt(2.0);
    const double y = double(1 << level) * x - double(2 * cell - 1);
    return pair_eval(dg.data() + (mode - 1) * 2 * k, k, y) * std::sqrt(double(1 << level));
  }
  // basis(level, cell, mode, x) (1d_dg_functions.jl:83-93)
  double pos(int level, int cell, int mode, double x) const {
    const double y = double(1 << level) * x - double(cell - 1);
    return std::sqrt(2.0) * legP(mode - 1, 2 * y - 1) * std::pow(2.0, level / 2.0);
  }
};

// Gauss-Legendre rule on [-1,1] (Newton iteration on the three-term recurrence).
void gauss_rule(int m, std::vector<double>& x, std::vector<double>& w) {
  x.assign(m, 0.0);
  w.assign(m, 0.0);
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < m; ++i) {
    double z = std::cos(pi * (i + 0.75) / (m + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int j = 2; j <= m; ++j) {
        const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      dp = m * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = -z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

inline int64_t cell_base(int l) { return l == 0 ? 0 : (int64_t(1) << (l - 1)); }
inline int ncells(int l) { return 1 << (l > 0 ? l - 1 : 0); }

struct SparseCol {
  std::vector<int32_t> idx;
  std::vector<double> val;
};

}  // namespace

const std::vector<double>& legendre_table() {
  static std::vector<double> tab;
  static std::once_flag once;
  std::call_once(once, [] {
    const auto rows = legendre_rows(KMAX);
    tab.assign((KMAX + 1) * 2 * (KMAX + 1), 0.0);
    for (int r = 0; r <= KMAX; ++r) std::copy(rows[r].begin(), rows[r].end(), tab.begin() + r * 2 * (KMAX + 1));
  });
  return tab;
}

std::vector<double> dg_basis(int k) {
  if (k < 1 || k > KMAX) throw std::invalid_argument("k must be in 1..10");
  const auto rows = multiwavelets(k);
  std::vector<double> out(k * 2 * k);
  for (int r = 0; r < k; ++r) std::copy(rows[r].begin(), rows[r].end(), out.begin() + r * 2 * k);
  return out;
}

int coupled_cells(int lo, int co, int li, int out[8]) {
  const int nci = ncells(li);
  int tmp[8];
  int cnt = 0;
  auto wrap = [nci](int64_t c) { return int(((c % nci) + nci) % nci); };
  if (li == lo) {
    if (lo <= 1) {
      tmp[cnt++] = 0;
    } else {
      tmp[cnt++] = wrap(co - 1);
      tmp[cnt++] = co;
      tmp[cnt++] = wrap(co + 1);
    }
  } else if (li > lo) {
    if (lo == 0) {
      tmp[cnt++] = 0;
      tmp[cnt++] = nci - 1;
    } else {
      const int sh = li - 1 - lo;  // jump point q 2^-lo is fine-cell edge q << sh
      for (int64_t q = 2 * int64_t(co); q <= 2 * int64_t(co) + 2; ++q) {
        const int64_t p = q << sh;
        tmp[cnt++] = wrap(p - 1);
        tmp[cnt++] = wrap(p);
      }
    }
  } else {  // coarser input: its jump points must touch an endpoint of the output cell
    const int nco = ncells(lo);
    if (li == 0) {
      if (co == 0 || co == nco - 1) tmp[cnt++] = 0;
    } else {
      const int sh = lo - 1 - li;
      for (int64_t e = co; e <= int64_t(co) + 1; ++e) {
        if (e & ((int64_t(1) << sh) - 1)) continue;
        const int64_t q = e >> sh;  // jump point q 2^-li
        if (q & 1) {
          tmp[cnt++] = wrap((q - 1) / 2);
        } else {
          tmp[cnt++] = wrap(q / 2 - 1);
          tmp[cnt++] = wrap(q / 2);
        }
      }
    }
  }
  std::sort(tmp, tmp + cnt);
  int m = 0;
  for (int t = 0; t < cnt; ++t)
    if (m == 0 || out[m - 1] != tmp[t]) out[m++] = tmp[t];
  return m;
}

Op1D build_op1d(int k, int n) {
  if (k < 1 || k > KMAX) throw std::invalid_argument("k must be in 1..10");
  if (n < 0 || n > 20) throw std::invalid_argument("n must be in 0..20");
  Basis B{k, legendre_table().data(), dg_basis(k)};
  const int nc = 1 << n;
  const int64_t N1 = int64_t(k) << n;

  // ---- hier2pos Q (1d_dg_functions.jl:242-263), column by column, |q| > 1e-12 kept.
  std::vector<double> gx, gw;
  gauss_rule(12, gx, gw);
  std::vector<SparseCol> Q(N1);
  for (int l = 0; l <= n; ++l)
    for (int c = 1; c <= ncells(l); ++c)
      for (int m = 1; m <= k; ++m) {
        const int64_t j = k * (cell_base(l) + c - 1) + (m - 1);
        const int span = (l == 0) ? nc : (nc >> (l - 1));
        const int f0 = (l == 0) ? 0 : (c - 1) * span;
        SparseCol& col = Q[j];
        for (int fc = f0; fc < f0 + span; ++fc) {
          const double a = double(fc) / nc, b = double(fc + 1) / nc;
          for (int p = 1; p <= k; ++p) {
            double s = 0.0;
            for (size_t q = 0; q < gx.size(); ++q) {
              const double x = 0.5 * (a + b) + 0.5 * (b - a) * gx[q];
              s += gw[q] * B.pos(n, fc + 1, p, x) * B.hier(l, c, m, x);
            }
            s *= 0.5 * (b - a);
            if (std::fabs(s) > 1e-12) {
              col.idx.push_back(int32_t(fc * k + (p - 1)));
              col.val.push_back(s);
            }
          }
        }
      }

  // ---- position-basis operator A = (-Dint + LF)' (1d_derivative.jl:21-111), CSC.
  // legendreDlegendre (derivative_matrix_elements.jl:80-103) on the Legendre table.
  const int LL = 2 * (KMAX + 1);
  const double* leg = legendre_table().data();
  auto ldl = [&](int m1, int m2) {
    Vec a(leg + (m1 - 1) * LL, leg + m1 * LL), d(LL, 0.0);
    const double* v = leg + (m2 - 1) * LL;
    const int h = LL / 2;
    for (int i = 1; i <= LL; ++i) {
      if (i < h) d[i - 1] = i * v[i];
      else if (i > h && i < 2 * h) d[i - 1] = (i - h) * v[i];
    }
    return dot(a, d);
  };
  const double tiny = 5.0e-16;
  auto lfv = [&](int c1, int m1, int c2, int m2) {  // periodic_legvLFv, alpha = 0
    const double p1 = double(c2 - 1) / double(nc), p2 = double(c2) / double(nc);
    double left1 = B.pos(n, c1, m1, p1 - tiny), right1 = B.pos(n, c1, m1, p1 + tiny);
    double left2 = B.pos(n, c1, m1, p2 - tiny), right2 = B.pos(n, c1, m1, p2 + tiny);
    if (c2 == nc) right2 = B.pos(n, c1, m1, 0.0 + tiny);
    if (c2 == 1) left1 = B.pos(n, c1, m1, 1.0 - tiny);
    const double LF1 = 0.5 * (left1 + right1), LF2 = 0.5 * (left2 + right2);
    const double v1 = B.pos(n, c2, m2, p1 + tiny), v2 = B.pos(n, c2, m2, p2 - tiny);
    return LF2 * v2 - LF1 * v1;
  };
  std::vector<double> ldl_tab(k * k);
  for (int a = 1; a <= k; ++a)
    for (int b = 1; b <= k; ++b) ldl_tab[(a - 1) * k + (b - 1)] = ldl(a, b);
  // Apos[(cf,mf),(cg,mg)] = A[(cg,mg),(cf,mf)]; column g of Apos lists rows f.
  std::vector<SparseCol> Apos(N1);
  for (int cg = 1; cg <= nc; ++cg)
    for (int mg = 1; mg <= k; ++mg) {
      SparseCol& col = Apos[(cg - 1) * k + (mg - 1)];
      int cand[3] = {cg - 1, cg, cg + 1}, nf = 0, cells[3];
      for (int t = 0; t < 3; ++t) {
        int cf = ((cand[t] - 1) % nc + nc) % nc + 1;
        bool dup = false;
        for (int u = 0; u < nf; ++u) dup |= (cells[u] == cf);
        if (!dup) cells[nf++] = cf;
      }
      std::sort(cells, cells + nf);
      for (int t = 0; t < nf; ++t) {
        const int cf = cells[t];
        for (int mf = 1; mf <= k; ++mf) {
          double dv = 0.0;
          if (cf == cg) {
            dv = double(1 << (n + 1)) * ldl_tab[(mg - 1) * k + (mf - 1)];
            if (!(std::fabs(dv) > 1.0e-15)) dv = 0.0;
          }
          double lf = lfv(cg, mg, cf, mf);
          if (!(std::fabs(lf) > 1.0e-15)) lf = 0.0;
          const double a = (-dv) + lf;
          if (a != 0.0) {
            col.idx.push_back(int32_t((cf - 1) * k + (mf - 1)));
            col.val.push_back(a);
          }
        }
      }
    }

  // ---- block table skeleton on the coupling pattern
  Op1D op;
  op.k = k;
  op.n = n;
  op.row_ptr.assign(size_t(nc) * (n + 2), 0);
  for (int lo = 0; lo <= n; ++lo)
    for (int co = 0; co < ncells(lo); ++co) {
      const int64_t r = cell_base(lo) + co;
      for (int li = 0; li <= n; ++li) {
        op.row_ptr[r * (n + 2) + li] = int32_t(op.ent.size());
        int cells[8];
        const int cntc = coupled_cells(lo, co, li, cells);
        for (int t = 0; t < cntc; ++t) op.ent.push_back((li << 24) | cells[t]);
      }
      op.row_ptr[r * (n + 2) + n + 1] = int32_t(op.ent.size());
    }
  op.val.assign(op.ent.size() * k * k, 0.0);

  // ---- D = Q'(A Q) on the pattern (1d_derivative.jl:113-117)
  std::vector<double> aq(N1, 0.0);
  std::vector<int32_t> touched;
  std::vector<char> mark(N1, 0);
  for (int li = 0; li <= n; ++li)
    for (int ci = 0; ci < ncells(li); ++ci)
      for (int mi = 0; mi < k; ++mi) {
        const int64_t j = k * (cell_base(li) + ci) + mi;
        // column j of A*Q: scatter over ascending g (spmatmul order)
        for (size_t t = 0; t < Q[j].idx.size(); ++t) {
          const int32_t g = Q[j].idx[t];
          const double qg = Q[j].val[t];
          for (size_t u = 0; u < Apos[g].idx.size(); ++u) {
            const int32_t f = Apos[g].idx[u];
            if (!mark[f]) { mark[f] = 1; touched.push_back(f); }
            aq[f] += Apos[g].val[u] * qg;
          }
        }
        // rows coupled to (li, ci): by symmetry of the rule, coupled_cells(li, ci, lo)
        for (int lo = 0; lo <= n; ++lo) {
          int cells[8];
          const int cntc = coupled_cells(li, ci, lo, cells);
          for (int t = 0; t < cntc; ++t) {
            const int co = cells[t];
            const int64_t r = cell_base(lo) + co;
            int32_t e = op.row_ptr[r * (n + 2) + li], e1 = op.row_ptr[r * (n + 2) + li + 1];
            while (e < e1 && (op.ent[e] & 0xFFFFFF) != ci) ++e;
            if (e == e1) throw std::logic_error("coupling rule is not symmetric");
            for (int mo = 0; mo < k; ++mo) {
              const SparseCol& qi = Q[k * r + mo];
              double s = 0.0;
              for (size_t u = 0; u < qi.idx.size(); ++u) s += qi.val[u] * aq[qi.idx[u]];
              op.val[(size_t(e) * k + mo) * k + mi] = s;
            }
          }
        }
        for (int32_t f : touched) { aq[f] = 0.0; mark[f] = 0; }
        touched.clear();
      }
  return op;
}

std::vector<double> op1d_dense(const Op1D& op) {
  const int k = op.k, n = op.n;
  const int64_t N1 = int64_t(k) << n;
  std::vector<double> Dm(size_t(N1 * N1), 0.0);
  for (int64_t r = 0; r < (int64_t(1) << n); ++r)
    for (int32_t e = op.row_ptr[r * (n + 2)]; e < op.row_ptr[r * (n + 2) + n + 1]; ++e) {
      const int li = op.ent[e] >> 24, ci = op.ent[e] & 0xFFFFFF;
      const int64_t c0 = k * (cell_base(li) + ci);
      for (int mo = 0; mo < k; ++mo)
        for (int mi = 0; mi < k; ++mi) Dm[(k * r + mo) * N1 + c0 + mi] = op.val[(size_t(e) * k + mo) * k + mi];
    }
  return Dm;
}

// ---------------------------------------------------------------------------------
// Layout

namespace {
bool cut(const int* L, int D, int n, bool full) {
  if (full) return false;
  int s = 0;
  for (int d = 0; d < D; ++d) s += L[d];
  return s > n;  // 0-based levels: sum(L.I) > n + D on 1-based levels
}
bool next_levels(int* L, int D, int n) {
  for (int d = 0; d < D; ++d) {
    if (L[d] < n) { ++L[d]; return true; }
    L[d] = 0;
  }
  return false;
}
int64_t ipow(int64_t b, int e) {
  int64_t r = 1;
  while (e-- > 0) r *= b;
  return r;
}
}  // namespace

int64_t layout_size(int D, int k, int n, bool full) {
  if (D < 1 || D > 16) return -1;
  int L[16] = {0};
  int64_t size = 0;
  const int64_t kD = ipow(k, D);
  do {
    if (cut(L, D, n, full)) continue;
    int64_t p = 1;
    for (int d = 0; d < D; ++d) p *= ncells(L[d]);
    size += p * kD;
  } while (next_levels(L, D, n));
  return size;
}

Layout build_layout(int D, int k, int n, bool full) {
  if (D < 1 || D > 16) throw std::invalid_argument("D out of range");
  if (k < 1 || k > KMAX) throw std::invalid_argument("k must be in 1..10");
  if (n < 0 || n > 20) throw std::invalid_argument("n must be in 0..20");
  Layout lt;
  lt.D = D; lt.k = k; lt.n = n; lt.full = full ? 1 : 0;
  const int64_t ncode = ipow(n + 1, D);
  if (ncode > (int64_t(1) << 26)) throw std::invalid_argument("(n+1)^D too large");
  lt.code2blk.assign(size_t(ncode), -1);
  int L[16] = {0};
  int64_t off = 0;
  const int64_t kD = ipow(k, D);
  do {
    if (cut(L, D, n, full)) continue;
    int64_t code = 0;
    for (int d = D - 1; d >= 0; --d) code = code * (n + 1) + L[d];
    lt.code2blk[code] = lt.nblk++;
    for (int d = 0; d < D; ++d) lt.blk_lev.push_back(L[d]);
    lt.blk_off.push_back(off);
    int64_t p = 1;
    for (int d = 0; d < D; ++d) p *= ncells(L[d]);
    off += p * kD;
  } while (next_levels(L, D, n));
  lt.blk_off.push_back(off);
  lt.N = off;
  lt.blk_inoff.assign(size_t(lt.nblk) * D * (n + 1), -1);
  for (int b = 0; b < lt.nblk; ++b)
    for (int d = 0; d < D; ++d)
      for (int l = 0; l <= n; ++l) {
        int64_t code = 0;
        for (int e = D - 1; e >= 0; --e) code = code * (n + 1) + (e == d ? l : lt.blk_lev[b * D + e]);
        const int32_t bb = lt.code2blk[code];
        lt.blk_inoff[(size_t(b) * D + d) * (n + 1) + l] = bb < 0 ? -1 : lt.blk_off[bb];
      }
  return lt;
}

void layout_v2dref(const Layout& lt, int32_t* levels, int32_t* cells, int32_t* modes) {
  const int D = lt.D, k = lt.k;
  const int64_t kD = ipow(k, D);
  for (int b = 0; b < lt.nblk; ++b) {
    const int* L = &lt.blk_lev[b * D];
    for (int64_t r = 0; r < lt.blk_off[b + 1] - lt.blk_off[b]; ++r) {
      const int64_t j = lt.blk_off[b] + r;
      int64_t cl = r / kD, ml = r % kD;
      for (int d = 0; d < D; ++d) {
        const int nc = ncells(L[d]);
        levels[j * D + d] = L[d] + 1;
        cells[j * D + d] = int32_t(cl % nc) + 1;
        cl /= nc;
        modes[j * D + d] = int32_t(ml % k) + 1;
        ml /= k;
      }
    }
  }
}

CellPlan build_cell_plan(const Layout& lt, const Op1D& op) {
  const int D = lt.D, k = lt.k, n = lt.n;
  if (op.k != k || op.n != n) throw std::invalid_argument("operator/layout mismatch");
  if (lt.N >= (int64_t(1) << 31)) throw std::invalid_argument("grid too large for 32-bit offsets");
  const int64_t kD = ipow(k, D);
  const int64_t ncell = lt.N / kD;
  CellPlan P;
  P.zero_block = int32_t(op.ent.size());
  P.ptr.assign(size_t(ncell) * D + 1, 0);
  P.ent.reserve(size_t(ncell) * D * 8 * 2);
  int64_t g = 0;
  for (int b = 0; b < lt.nblk; ++b) {
    const int* L = &lt.blk_lev[b * D];
    int sh[16], cum[16], sumL = 0, acc = 0;
    for (int e = 0; e < D; ++e) {
      sh[e] = L[e] > 0 ? L[e] - 1 : 0;
      cum[e] = acc;
      acc += sh[e];
      sumL += L[e];
    }
    const int64_t cells = (lt.blk_off[b + 1] - lt.blk_off[b]) / kD;
    for (int64_t cl = 0; cl < cells; ++cl, ++g) {
      for (int d = 0; d < D; ++d) {
        P.ptr[g * D + d] = int32_t(P.ent.size() / 2);
        const int ld = L[d];
        const int mf = lt.full ? n : n - (sumL - ld);
        const int64_t cd = (cl >> cum[d]) & ((int64_t(1) << sh[d]) - 1);
        const int64_t r1 = cell_base(ld) + cd;
        const int32_t e0 = op.row_ptr[r1 * (n + 2)], e1 = op.row_ptr[r1 * (n + 2) + mf + 1];
        const int64_t clo = cl & ((int64_t(1) << cum[d]) - 1);
        const int64_t chi = cl >> (cum[d] + sh[d]);
        int32_t first = -1;
        for (int32_t e = e0; e < e1; ++e) {
          const int li = op.ent[e] >> 24;
          const int64_t ci = op.ent[e] & 0xFFFFFF;
          const int sl = li > 0 ? li - 1 : 0;
          const int64_t cin = clo + ((ci + (chi << sl)) << cum[d]);
          const int64_t base = lt.blk_inoff[(size_t(b) * D + d) * (n + 1) + li];
          if (base < 0) throw std::logic_error("coupled input block is cut");
          const int64_t off = base + cin * kD;
          if (first < 0) first = int32_t(off);
          P.ent.push_back(int32_t(off));
          P.ent.push_back(e);
        }
        if (first < 0) first = int32_t(lt.blk_off[b] + cl * kD);
        while ((P.ent.size() / 2 - P.ptr[g * D + d]) % PLAN_UNROLL) {
          P.ent.push_back(first);
          P.ent.push_back(P.zero_block);
        }
      }
    }
  }
  P.ptr[size_t(ncell) * D] = int32_t(P.ent.size() / 2);
  if (P.ent.size() / 2 >= (size_t(1) << 31)) throw std::invalid_argument("plan too large");
  return P;
}

TilePlan build_tile_plan(const Layout& lt, const Op1D& op, const CellPlan& cp, int cpt, int lds_elems, int ent_elems,
                         int max_ents, int pad) {
  const int D = lt.D, k = lt.k;
  const int64_t kD = ipow(k, D);
  if (kD > 256) throw std::invalid_argument("tile plan needs k^D <= 256");
  if (cpt < 1 || pad < 1) throw std::invalid_argument("tile plan: cpt and pad must be >= 1");
  if (op.ent.size() + 1 >= (size_t(1) << 21)) throw std::invalid_argument("1-D table too large for the tile plan");
  TilePlan P;
  P.cpt = cpt;
  P.cap = int(std::min<int64_t>(1023, lds_elems / kD));
  if (P.cap < 1) throw std::invalid_argument("LDS too small for one input cell");
  const int64_t kk = int64_t(k) * k;
  const int64_t ecap = std::min<int64_t>(max_ents, ent_elems / kk);  // entries per batch
  if (ecap < pad) throw std::invalid_argument("coefficient LDS too small");
  const int64_t ncell = lt.N / kD;
  P.ntiles = int((ncell + P.cpt - 1) / P.cpt);
  P.tile_bptr.push_back(0);
  P.batch_sptr.push_back(0);
  P.batch_lptr.push_back(0);
  std::vector<std::pair<int64_t, int32_t>> keys;  // (offset * 16 + dir, slot) of the open batch
  std::vector<std::vector<int32_t>> lists(size_t(P.cpt) * D);
  int64_t padded = 0;  // entries of the open batch once every list is padded
  auto close_batch = [&]() {
    const size_t e0 = P.ent.size();
    if (keys.empty()) {  // a batch always owns slot 0 (padding entries point at it)
      P.slot_off.push_back(0);
      P.slot_dir.push_back(0);
      keys.emplace_back(-1, 0);
    }
    for (auto& l : lists) {
      while (l.size() % pad) l.push_back((cp.zero_block << 10) | 0);
      P.ent.insert(P.ent.end(), l.begin(), l.end());
      P.batch_lptr.push_back(int32_t(P.ent.size()));
      l.clear();
    }
    P.max_slots = std::max(P.max_slots, int(keys.size()));
    P.max_ents = std::max(P.max_ents, int(P.ent.size() - e0));
    P.batch_sptr.push_back(int32_t(P.slot_off.size()));
    keys.clear();
    padded = 0;
  };
  for (int t = 0; t < P.ntiles; ++t) {
    const int64_t g0 = int64_t(t) * P.cpt;
    const int ncl = int(std::min<int64_t>(P.cpt, ncell - g0));
    for (int c = 0; c < ncl; ++c)
      for (int d = 0; d < D; ++d) {
        const int64_t g = g0 + c;
        std::vector<int32_t>& list = lists[size_t(c) * D + d];
        for (int32_t p = cp.ptr[g * D + d]; p < cp.ptr[g * D + d + 1]; ++p) {
          const int32_t off = cp.ent[2 * p], e = cp.ent[2 * p + 1];
          if (e == cp.zero_block) continue;
          const int64_t key = int64_t(off) * 16 + d;
          if (list.size() % pad == 0 && padded + pad > ecap) close_batch();
          int32_t slot = -1;
          for (auto& kv : keys)
            if (kv.first == key) { slot = kv.second; break; }
          if (slot < 0) {
            if (int(keys.size()) == P.cap) close_batch();
            slot = int32_t(keys.size());
            keys.emplace_back(key, slot);
            P.slot_off.push_back(off);
            P.slot_dir.push_back(d);
          }
          if (list.size() % pad == 0) padded += pad;
          list.push_back((e << 10) | slot);
        }
      }
    close_batch();  // every tile ends with its own batch
    P.tile_bptr.push_back(int32_t(P.batch_sptr.size() - 1));
  }
  return P;
}

Csc d_matrix_csc(const Layout& lt, const Op1D& op, int d) {
  const int D = lt.D, k = lt.k, n = lt.n;
  if (op.k != k || op.n != n) throw std::invalid_argument("operator/layout mismatch");
  const int64_t kD = ipow(k, D);
  Csc A;
  A.m = A.n = lt.N;
  A.colptr.assign(size_t(lt.N) + 1, 0);
  std::vector<std::pair<int64_t, double>> col;
  for (int b = 0; b < lt.nblk; ++b) {
    const int* L = &lt.blk_lev[b * D];
    // strides of the cell coordinate d inside a block
    int64_t cstride = 1;
    for (int e = 0; e < d; ++e) cstride *= ncells(L[e]);
    const int64_t mstride = ipow(k, d);
    for (int64_t r = 0; r < lt.blk_off[b + 1] - lt.blk_off[b]; ++r) {
      const int64_t j = lt.blk_off[b] + r;
      const int64_t cl = r / kD, ml = r % kD;
      const int li = L[d];
      const int ci = int((cl / cstride) % ncells(li));
      const int mi = int((ml / mstride) % k);
      const int64_t clo = cl % cstride, chi = cl / (cstride * ncells(li));
      const int64_t mlo = ml % mstride, mhi = ml / (mstride * k);
      col.clear();
      for (int lo = 0; lo <= n; ++lo) {
        const int64_t base = lt.blk_inoff[(size_t(b) * D + d) * (n + 1) + lo];
        if (base < 0) continue;
        int cells[8];
        const int cntc = coupled_cells(li, ci, lo, cells);
        for (int t = 0; t < cntc; ++t) {
          const int co = cells[t];
          const int64_t rr = cell_base(lo) + co;
          int32_t e = op.row_ptr[rr * (n + 2) + li], e1 = op.row_ptr[rr * (n + 2) + li + 1];
          while (e < e1 && (op.ent[e] & 0xFFFFFF) != ci) ++e;
          if (e == e1) throw std::logic_error("coupling rule is not symmetric");
          const int64_t clo_out = clo + cstride * (co + int64_t(ncells(lo)) * chi);
          for (int mo = 0; mo < k; ++mo) {
            const double v = op.val[(size_t(e) * k + mo) * k + mi];
            if (v == 0.0) continue;
            const int64_t row = base + clo_out * kD + mlo + mstride * (mo + int64_t(k) * mhi);
            col.emplace_back(row, v);
          }
        }
      }
      std::sort(col.begin(), col.end());
      for (auto& p : col) {
        A.rowval.push_back(p.first);
        A.nzval.push_back(p.second);
      }
      A.colptr[j + 1] = int64_t(A.rowval.size());
    }
  }
  return A;
}

// =================================================================================
// Nodal basis and the basis transforms of the Vlasov path (1d_nodal_basis.jl,
// multidim_nodal_basis.jl).  All 1-D matrices are dense (k 2^n)^2, row-major, indexed
// like the hierarchical vector: index k (cell_base(l) + c) + m for (level, 0-based cell,
// 0-based mode); "points" use the same index for (level, cell, node); "pos" (the
// position basis on the 2^n finest cells) uses fc k + m.
namespace {

// lag_nodal (1d_nodal_basis.jl:46-68): Lagrange polynomials of the coarsest level
double lag_nodal(int k, int mode, double x) {
  if (x < 0 || x > 1) return 0.0;
  switch (k) {
    case 2: return mode == 1 ? x : 1 - x;
    case 3:
      if (mode == 1) return 2 * (x - 0.5) * (x - 1);
      if (mode == 2) return -4 * x * (x - 1);
      return 2 * x * (x - 0.5);
    default:  // 5
      if (mode == 1) return 32.0 / 3 * (x - 0.25) * (x - 0.5) * (x - 0.75) * (x - 1);
      if (mode == 2) return -128.0 / 3 * x * (x - 0.5) * (x - 0.75) * (x - 1);
      if (mode == 3) return 64.0 * x * (x - 0.25) * (x - 0.75) * (x - 1);
      if (mode == 4) return -128.0 / 3 * x * (x - 0.25) * (x - 0.5) * (x - 1);
      return 32.0 / 3 * x * (x - 0.25) * (x - 0.5) * (x - 0.75);
  }
}

// h_nodal (1d_nodal_basis.jl:71-102): the multiresolution functions on [-1, 1]
double h_nodal(int k, int mode, double x) {
  switch (k) {
    case 2:
      if (mode == 1) return std::max(0.0, 1 - std::fabs(x));
      if (std::fabs(x) > 1 || x == 0) return 0.0;
      return x < 0 ? x + 1 : 0.0;
    case 3:
      if (mode == 1) return std::max(0.0, -4 * x * (x + 1));
      if (mode == 2) return std::max(0.0, -4 * x * (x - 1));
      if (std::fabs(x) > 1 || x == 0) return 0.0;
      return x < 0 ? 2 * (x + 1) * (x + 0.5) : 0.0;
    default:  // 5
      if (mode == 1) return (x < -1 || x > 0) ? 0.0 : -128.0 / 3 * x * (x + 1) * (x + 0.5) * (x + 0.25);
      if (mode == 2) return (x < -1 || x > 0) ? 0.0 : -128.0 / 3 * x * (x + 1) * (x + 0.5) * (x + 0.75);
      if (mode == 3) return (x > 1 || x < 0) ? 0.0 : -128.0 / 3 * x * (x - 1) * (x - 0.5) * (x - 0.75);
      if (mode == 4) return (x > 1 || x < 0) ? 0.0 : -128.0 / 3 * x * (x - 1) * (x - 0.5) * (x - 0.25);
      if (std::fabs(x) > 1 || x == 0) return 0.0;
      return x < 0 ? 32.0 / 3 * (x + 1) * (x + 0.75) * (x + 0.5) * (x + 0.25) : 0.0;
  }
}

// v_nodal (1d_nodal_basis.jl:105-111), cell and mode 1-based
double v_nodal(int k, int level, int cell, int mode, double x) {
  if (level == 0) return lag_nodal(k, mode, x);
  return h_nodal(k, mode, double(1 << level) * x - double(2 * cell - 1));
}

void threshold(std::vector<double>& a, double atol) {
  for (double& v : a)
    if (std::fabs(v) < atol) v = 0.0;
}

// the collocation points of eval_points_1D (1d_nodal_basis.jl:180-211): per cell the k-1
// nodes, evaluated from the right, then one more point evaluated from the left (x = 1
// on level 0, the cell midpoint on the finer levels)
struct Point1D {
  double x;
  bool left;
};
std::vector<Point1D> points_1d_list(int k, int n) {
  std::vector<Point1D> pts;
  for (int l = 0; l <= n; ++l)
    for (int c = 1; c <= ncells(l); ++c) {
      for (int q = 0; q < k - 1; ++q) {
        const double node = l == 0 ? double(q) / (k - 1) : (2.0 * q + 1) / (2.0 * (k - 1));
        pts.push_back({(c - 1 + node) / ncells(l), false});
      }
      pts.push_back({l == 0 ? 1.0 : (c - 0.5) / ncells(l), true});
    }
  return pts;
}

// nodal2points_1D (1d_nodal_basis.jl:215-236)
std::vector<double> nodal2points(int k, int n, double atol) {
  const std::vector<Point1D> pts = points_1d_list(k, n);
  const int64_t K = int64_t(k) << n;
  std::vector<double> A(size_t(K * K), 0.0);
  for (int l = 0; l <= n; ++l)
    for (int c = 1; c <= ncells(l); ++c)
      for (int m = 1; m <= k; ++m) {
        const int64_t j = k * (cell_base(l) + c - 1) + (m - 1);
        for (int64_t i = 0; i < K; ++i) {
          const double x = pts[i].left ? pts[i].x - 1e-16 : pts[i].x + 1e-16;
          A[i * K + j] = v_nodal(k, l, c, m, x);
        }
      }
  threshold(A, atol);
  return A;
}

// nodal2pos_1D (1d_nodal_basis.jl:300-360); the integrals of the piecewise polynomials
// are exact Gauss-Legendre sums on the two halves of each finest cell
std::vector<double> nodal2pos(const Basis& B, int n, double atol) {
  const int k = B.k;
  const int64_t K = int64_t(k) << n;
  const int nf = 1 << n;
  std::vector<double> gx, gw;
  gauss_rule(8, gx, gw);
  std::vector<double> A(size_t(K * K), 0.0);
  for (int l = 0; l <= n; ++l)
    for (int c = 1; c <= ncells(l); ++c) {
      const double nmin = double(c - 1) / ncells(l), nmax = double(c) / ncells(l);
      for (int m = 1; m <= k; ++m) {
        const int64_t j = k * (cell_base(l) + c - 1) + (m - 1);
        for (int fc = 1; fc <= nf; ++fc) {
          const double pmin = double(fc - 1) / nf, pmed = (fc - 0.5) / nf, pmax = double(fc) / nf;
          if (pmin > nmax || pmax < nmin) continue;
          for (int p = 1; p <= k; ++p) {
            double val = 0.0;
            const double ends[3] = {pmin, pmed, pmax};
            for (int h = 0; h < 2; ++h) {
              const double a = ends[h], b = ends[h + 1];
              double s = 0.0;
              for (size_t q = 0; q < gx.size(); ++q) {
                const double x = 0.5 * (a + b) + 0.5 * (b - a) * gx[q];
                s += gw[q] * v_nodal(k, l, c, m, x) * B.pos(n, fc, p, x);
              }
              val += 0.5 * (b - a) * s;
            }
            A[((fc - 1) * k + (p - 1)) * K + j] = val;
          }
        }
      }
    }
  threshold(A, atol);
  return A;
}

// hier2pos (1d_dg_functions.jl:242-263), dense, |q| > 1e-12 kept
std::vector<double> hier2pos_dense(const Basis& B, int n) {
  const int k = B.k;
  const int64_t K = int64_t(k) << n;
  const int nc = 1 << n;
  std::vector<double> gx, gw;
  gauss_rule(12, gx, gw);
  std::vector<double> Q(size_t(K * K), 0.0);
  for (int l = 0; l <= n; ++l)
    for (int c = 1; c <= ncells(l); ++c)
      for (int m = 1; m <= k; ++m) {
        const int64_t j = k * (cell_base(l) + c - 1) + (m - 1);
        const int span = (l == 0) ? nc : (nc >> (l - 1));
        const int f0 = (l == 0) ? 0 : (c - 1) * span;
        for (int fc = f0; fc < f0 + span; ++fc) {
          const double a = double(fc) / nc, b = double(fc + 1) / nc;
          for (int p = 1; p <= k; ++p) {
            double s = 0.0;
            for (size_t q = 0; q < gx.size(); ++q) {
              const double x = 0.5 * (a + b) + 0.5 * (b - a) * gx[q];
              s += gw[q] * B.pos(n, fc + 1, p, x) * B.hier(l, c, m, x);
            }
            s *= 0.5 * (b - a);
            if (std::fabs(s) > 1e-12) Q[(int64_t(fc) * k + (p - 1)) * K + j] = s;
          }
        }
      }
  return Q;
}

// dense inverse, Gauss-Jordan with partial pivoting
std::vector<double> invert(std::vector<double> A, int64_t K) {
  std::vector<double> X(size_t(K * K), 0.0);
  for (int64_t i = 0; i < K; ++i) X[i * K + i] = 1.0;
  for (int64_t c = 0; c < K; ++c) {
    int64_t p = c;
    for (int64_t r = c + 1; r < K; ++r)
      if (std::fabs(A[r * K + c]) > std::fabs(A[p * K + c])) p = r;
    if (A[p * K + c] == 0.0) throw std::runtime_error("singular basis transform");
    if (p != c)
      for (int64_t j = 0; j < K; ++j) {
        std::swap(A[p * K + j], A[c * K + j]);
        std::swap(X[p * K + j], X[c * K + j]);
      }
    const double inv = 1.0 / A[c * K + c];
    for (int64_t j = 0; j < K; ++j) {
      A[c * K + j] *= inv;
      X[c * K + j] *= inv;
    }
    for (int64_t r = 0; r < K; ++r) {
      if (r == c) continue;
      const double f = A[r * K + c];
      if (f == 0.0) continue;
      for (int64_t j = 0; j < K; ++j) {
        A[r * K + j] -= f * A[c * K + j];
        X[r * K + j] -= f * X[c * K + j];
      }
    }
  }
  return X;
}

// C = A B (dense, K x K)
std::vector<double> matmul(const std::vector<double>& A, const std::vector<double>& Bm, int64_t K) {
  std::vector<double> C(size_t(K * K), 0.0);
  for (int64_t i = 0; i < K; ++i)
    for (int64_t l = 0; l < K; ++l) {
      const double a = A[i * K + l];
      if (a == 0.0) continue;
      for (int64_t j = 0; j < K; ++j) C[i * K + j] += a * Bm[l * K + j];
    }
  return C;
}

}  // namespace

bool nodal_k_ok(int k) { return k == 2 || k == 3 || k == 5; }

std::vector<double> transform_1d(int k, int n, int from, int to) {
  if (!nodal_k_ok(k)) throw std::invalid_argument("the nodal basis exists for k in {2, 3, 5}");
  if (n < 0 || n > 12) throw std::invalid_argument("n must be in 0..12 for the basis transforms");
  const double atol = 1e-12;  // transform_1D's default
  const int64_t K = int64_t(k) << n;
  Basis B{k, legendre_table().data(), dg_basis(k)};
  if (from == TB_NODAL && to == TB_POINTS) return nodal2points(k, n, atol);
  if (from == TB_POINTS && to == TB_NODAL) {
    std::vector<double> X = invert(nodal2points(k, n, 1e-15), K);
    threshold(X, atol);
    return X;
  }
  if (from == TB_MODAL && to == TB_NODAL) {
    std::vector<double> p2n = invert(nodal2pos(B, n, atol), K);  // pos2nodal_1D(k, n)
    threshold(p2n, atol);
    std::vector<double> X = matmul(p2n, hier2pos_dense(B, n), K);
    threshold(X, atol);
    return X;
  }
  if (from == TB_NODAL && to == TB_MODAL) {
    const std::vector<double> Q = hier2pos_dense(B, n);
    std::vector<double> Qt(size_t(K * K));
    for (int64_t i = 0; i < K; ++i)
      for (int64_t j = 0; j < K; ++j) Qt[j * K + i] = Q[i * K + j];  // pos2hier = hier2pos'
    std::vector<double> X = matmul(Qt, nodal2pos(B, n, 1e-15), K);
    threshold(X, atol);
    return X;
  }
  throw std::invalid_argument("unsupported basis transform");
}

std::vector<double> points_1d(int k, int n) {
  std::vector<double> x;
  for (const Point1D& p : points_1d_list(k, n)) x.push_back(p.x);
  return x;
}

Op1D transform_op1d(int k, int n, int from, int to) {
  const std::vector<double> M = transform_1d(k, n, from, to);
  const int64_t K = int64_t(k) << n;
  Op1D op;
  op.k = k;
  op.n = n;
  const int64_t nrow = int64_t(1) << n;
  op.row_ptr.assign(size_t(nrow * (n + 2)), 0);
  auto nested = [](int l1, int c1, int l2, int c2) {  // relevant_cell_1D, 0-based cells
    const double a1 = double(c1) / ncells(l1), b1 = double(c1 + 1) / ncells(l1);
    const double a2 = double(c2) / ncells(l2), b2 = double(c2 + 1) / ncells(l2);
    return (a1 <= a2 && b1 >= b2) || (a1 >= a2 && b1 <= b2);
  };
  for (int lo = 0; lo <= n; ++lo)
    for (int co = 0; co < ncells(lo); ++co) {
      const int64_t r = cell_base(lo) + co;
      for (int li = 0; li <= n; ++li) {
        op.row_ptr[r * (n + 2) + li] = int32_t(op.ent.size());
        for (int ci = 0; ci < ncells(li); ++ci) {
          if (!nested(lo, co, li, ci)) continue;
          bool any = false;
          double blk[10 * 10];
          for (int mo = 0; mo < k; ++mo)
            for (int mi = 0; mi < k; ++mi) {
              const double v = M[(k * r + mo) * K + k * (cell_base(li) + ci) + mi];
              blk[mo * k + mi] = v;
              any |= v != 0.0;
            }
          if (!any) continue;
          op.ent.push_back(int32_t((li << 24) | ci));
          op.val.insert(op.val.end(), blk, blk + k * k);
        }
      }
      op.row_ptr[r * (n + 2) + n + 1] = int32_t(op.ent.size());
    }
  return op;
}

namespace {
// depth-first walk over one 1-D entry per direction for one input DoF of transform_csc;
// the per-direction lists are in ascending output level, so a sparse grid's level-sum
// cut ends a direction's loop
struct TcWalk {
  const Layout* lt;
  const Op1D* op;
  int D, k, n;
  int64_t kD;
  const std::vector<std::pair<int32_t, int32_t>>* lists[16];
  int min_[16], lo[16], co[16];
  int32_t ent[16];
  int64_t code_stride[17];
  std::vector<std::pair<int64_t, double>>* col;

  static int row_level(int64_t r) {
    int l = 0;
    while (cell_base(l + 1) <= r && l < 30) ++l;
    return l;
  }
  void leaf(int64_t code) {
    const int32_t bo = lt->code2blk[size_t(code)];
    if (bo < 0) return;
    int64_t cell = 0, cs = 1;
    for (int d = 0; d < D; ++d) {
      cell += co[d] * cs;
      cs *= ncells(lo[d]);
    }
    const int64_t base = lt->blk_off[bo] + cell * kD;
    for (int64_t mo = 0; mo < kD; ++mo) {
      double v = 1.0;
      int64_t t = mo;
      for (int d = 0; d < D && v != 0.0; ++d) {
        const int m = int(t % k);
        t /= k;
        v *= op->val[(size_t(ent[d]) * k + m) * k + min_[d]];
      }
      if (v != 0.0) col->emplace_back(base + mo, v);
    }
  }
  void walk(int d, int sum, int64_t code) {
    if (d == D) {
      leaf(code);
      return;
    }
    for (const auto& pr : *lists[d]) {
      const int l = row_level(pr.first);
      if (!lt->full && sum + l > n) break;
      lo[d] = l;
      co[d] = int(pr.first - cell_base(l));
      ent[d] = pr.second;
      walk(d + 1, sum + l, code + l * code_stride[d]);
    }
  }
};
}  // namespace

Csc transform_csc(const Layout& lt, const Op1D& op) {
  const int D = lt.D, k = lt.k, n = lt.n;
  if (op.k != k || op.n != n) throw std::invalid_argument("transform/layout mismatch");
  if (D < 1 || D > 16) throw std::invalid_argument("D must be in 1..16");
  // the 1-D table by input cell: (output cell-row, entry) pairs, output rows ascending
  const int64_t ncr = int64_t(1) << n;
  std::vector<std::vector<std::pair<int32_t, int32_t>>> bycol(static_cast<size_t>(ncr));
  for (int64_t r = 0; r < ncr; ++r)
    for (int32_t e = op.row_ptr[r * (n + 2)]; e < op.row_ptr[r * (n + 2) + n + 1]; ++e) {
      const int li = op.ent[e] >> 24, ci = op.ent[e] & 0xFFFFFF;
      bycol[size_t(cell_base(li) + ci)].emplace_back(int32_t(r), e);
    }
  Csc A;
  A.m = A.n = lt.N;
  A.colptr.assign(size_t(lt.N) + 1, 0);
  std::vector<std::pair<int64_t, double>> col;
  TcWalk w;
  w.lt = &lt;
  w.op = &op;
  w.D = D;
  w.k = k;
  w.n = n;
  w.kD = ipow(k, D);
  w.col = &col;
  w.code_stride[0] = 1;
  for (int d = 0; d < D; ++d) w.code_stride[d + 1] = w.code_stride[d] * (n + 1);
  for (int b = 0; b < lt.nblk; ++b) {
    const int* L = &lt.blk_lev[b * D];
    for (int64_t rloc = 0; rloc < lt.blk_off[b + 1] - lt.blk_off[b]; ++rloc) {
      const int64_t j = lt.blk_off[b] + rloc;
      int64_t cl = rloc / w.kD, ml = rloc % w.kD;
      for (int d = 0; d < D; ++d) {
        const int c = int(cl % ncells(L[d]));
        cl /= ncells(L[d]);
        w.min_[d] = int(ml % k);
        ml /= k;
        w.lists[d] = &bycol[size_t(cell_base(L[d]) + c)];
      }
      col.clear();
      w.walk(0, 0, 0);
      std::sort(col.begin(), col.end());
      for (auto& p : col) {
        A.rowval.push_back(p.first);
        A.nzval.push_back(p.second);
      }
      A.colptr[size_t(j) + 1] = int64_t(A.rowval.size());
    }
  }
  return A;
}

void projection_tables(int k, int n, int q, std::vector<double>& nodes, std::vector<double>& wv) {
  if (k < 1 || k > KMAX) throw std::invalid_argument("k must be in 1..10");
  if (n < 0 || n > 20) throw std::invalid_argument("n must be in 0..20");
  if (q < 1 || q > 32) throw std::invalid_argument("q must be in 1..32");
  Basis B{k, legendre_table().data(), dg_basis(k)};
  std::vector<double> gw;
  gauss_rule(q, nodes, gw);
  const int P = 2 * q;
  wv.assign(size_t(n + 1) * k * P, 0.0);
  for (int l = 0; l <= n; ++l) {
    // the first cell of the level: [0, 1/nc), halves [a, mid] and [mid, b]
    const double a = 0.0, b = 1.0 / ncells(l), mid = 0.5 * (a + b);
    for (int s = 0; s < 2; ++s) {
      const double pa = s ? mid : a, pb = s ? b : mid;
      for (int j = 0; j < q; ++j) {
        const double x = 0.5 * (pa + pb) + 0.5 * (pb - pa) * nodes[j];
        const double w = 0.5 * (pb - pa) * gw[j];
        for (int m = 1; m <= k; ++m) wv[(size_t(l) * k + (m - 1)) * P + s * q + j] = w * B.hier(l, 1, m, x);
      }
    }
  }
}

}  // namespace gsg
