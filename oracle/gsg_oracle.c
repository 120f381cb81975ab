/*
 * gsg_oracle.c -- FP64 CPU restatement of GalerkinSparseGrids.jl's operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (libgsg.so) never links, calls or falls back to it.
 *
 * It follows the reference algorithm literally (dense/CSC arithmetic in the
 * reference's order) so that a disagreement with the HIP path points at the HIP
 * path.  Every function cites the reference file:line it restates.  All indices in
 * the public API are 0-based; level/cell/mode tuples are 1-based like Julia's
 * CartesianIndex values.
 *
 * Pins (see tests/test_oracle_*.py): closed-form Legendre, orthonormality and
 * vanishing moments of the multiwavelets, orthogonality of hier2pos, skew symmetry
 * and D*const = 0 of the 1-D operator, brute-force Kronecker restriction for the
 * multi-D assembly, get_size closed form, layout round trip, energy conservation
 * and the derivative-accuracy bounds of the reference's own tests.
 * Parity unpinned: the fixed-step RK4 driver (the reference only calls ODE.jl's
 * adaptive ode45/ode78, pdes.jl:62-68) -- pinned only by energy conservation and
 * by the RK4 amplification polynomial on tiny systems.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define KMAX 10

typedef int64_t i64;

/* ------------------------------------------------------------------------- */
/* dg_basis.jl:21-35  product_matrix(i, j, n), 0-based i, j, n = vector length */
static double ipm1(int e) { return (e % 2 == 0) ? 1.0 : -1.0; }

static double product_matrix(int i, int j, int n) {
    int k = n / 2;
    if (i < k && j < k) return (1.0 + ipm1(i + j)) / (double)(1 + i + j);
    if (i >= k && j < k) return (1.0 - ipm1((i - k) + j)) / (double)(1 + (i - k) + j);
    if (i < k && j >= k) return product_matrix(j, i, n);
    return (1.0 + ipm1((i - k) + (j - k))) / (double)(1 + (i - k) + (j - k));
}

/* dg_basis.jl:37-50 */
static double inner_vv(const double *v1, const double *v2, int n) {
    double value = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            if (v1[i] == 0.0 || v2[j] == 0.0) continue;
            value += product_matrix(i, j, n) * v1[i] * v2[j];
        }
    return value;
}

/* dg_basis.jl:52-60  <v1, e_j> with j 1-based */
static double inner_vj(const double *v1, int j, int n) {
    double value = 0.0;
    for (int i = 0; i < n; i++) value += product_matrix(i, j - 1, n) * v1[i];
    return value;
}

/* dg_basis.jl:68-85  classical Gram-Schmidt; Q is nv x n row-major */
static void gram_schmidt(const double *Qi, double *Qf, int nv, int n) {
    memcpy(Qf, Qi, sizeof(double) * nv * n);
    for (int i = 0; i < nv; i++) {
        for (int j = 0; j < i; j++) {
            double proj = inner_vv(Qi + i * n, Qf + j * n, n) / inner_vv(Qf + j * n, Qf + j * n, n);
            for (int t = 0; t < n; t++) Qf[i * n + t] -= proj * Qf[j * n + t];
        }
        double nrm = sqrt(inner_vv(Qf + i * n, Qf + i * n, n));
        for (int t = 0; t < n; t++) Qf[i * n + t] /= nrm;
    }
}

/* dg_basis.jl:88-94  legendre(k): k+1 vectors of length 2(k+1) */
static void legendre(int k, double *out) {
    int nv = k + 1, n = 2 * (k + 1);
    double *Q = calloc((size_t)nv * n, sizeof(double));
    for (int j = 0; j < nv; j++) Q[j * n + j] = 1.0;
    gram_schmidt(Q, out, nv, n);
    free(Q);
}

/* dg_basis.jl:104-121 */
static void orthogonalize_1(const double *Qi, double *Qf, int k) {
    int n = 2 * k;
    double *leg = malloc(sizeof(double) * k * n);
    legendre(k - 1, leg);
    memcpy(Qf, Qi, sizeof(double) * k * n);
    for (int i = 0; i < k; i++)
        for (int j = 0; j < k; j++) {
            double proj = inner_vv(Qi + i * n, leg + j * n, n) / inner_vv(leg + j * n, leg + j * n, n);
            for (int t = 0; t < n; t++) Qf[i * n + t] -= proj * leg[j * n + t];
        }
    free(leg);
}

/* dg_basis.jl:130-146 */
static void orthogonalize_2(const double *Qi, double *Qf, int k) {
    int n = 2 * k;
    memcpy(Qf, Qi, sizeof(double) * k * n);
    for (int i = 1; i <= k - 1; i++) {
        const double *fi = Qi + (i - 1) * n;
        for (int j = i + 1; j <= k; j++) {
            double a = inner_vj(Qf + (j - 1) * n, k + i, n) / inner_vj(fi, k + i, n);
            for (int t = 0; t < n; t++) Qf[(j - 1) * n + t] -= a * fi[t];
        }
    }
}

/* dg_basis.jl:154-169 */
static void gram_schmidt_rev(const double *Qi, double *Qf, int k) {
    int n = 2 * k;
    memset(Qf, 0, sizeof(double) * k * n);
    for (int i = k - 1; i >= 0; i--) {
        const double *fi = Qi + i * n;
        memcpy(Qf + i * n, fi, sizeof(double) * n);
        for (int j = i + 1; j < k; j++) {
            double proj = inner_vv(fi, Qf + j * n, n) / inner_vv(Qf + j * n, Qf + j * n, n);
            for (int t = 0; t < n; t++) Qf[i * n + t] -= proj * Qf[j * n + t];
        }
        double nrm = sqrt(inner_vv(Qf + i * n, Qf + i * n, n));
        for (int t = 0; t < n; t++) Qf[i * n + t] /= nrm;
    }
}

/* dg_basis.jl:185-191  dg_basis(k): k vectors of length 2k (row = mode) */
int gsgo_dg_basis(int k, double *out) {
    if (k < 1 || k > KMAX) return -1;
    int n = 2 * k;
    double *Q0 = calloc((size_t)k * n, sizeof(double));
    double *Q1 = malloc(sizeof(double) * k * n);
    double *Q2 = malloc(sizeof(double) * k * n);
    for (int j = 1; j <= k; j++) Q0[(j - 1) * n + (k + j - 1)] = 1.0;
    orthogonalize_1(Q0, Q1, k);
    orthogonalize_2(Q1, Q2, k);
    gram_schmidt_rev(Q2, out, k);
    free(Q0); free(Q1); free(Q2);
    return 0;
}

/* 1d_dg_functions.jl:35  leg_coeffs = legendre(K_max): 11 x 22 */
int gsgo_legendre(int K, double *out) {
    if (K < 0 || K > 20) return -1;
    legendre(K, out);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Cached globals, as the module-load precomputation at 1d_dg_functions.jl:35,52-62 */
static double g_leg[(KMAX + 1) * 2 * (KMAX + 1)];
static double g_dg[KMAX + 1][KMAX * 2 * KMAX];
static int g_init = 0;

static void init_globals(void) {
    if (g_init) return;
    legendre(KMAX, g_leg);
    for (int k = 1; k <= KMAX; k++) gsgo_dg_basis(k, g_dg[k]);
    g_init = 1;
}

/* 1d_dg_functions.jl:15-28  array2poly (Horner over x^i and sgn(x)x^i parts) */
static double array2poly(const double *v, int n, double x) {
    if (fabs(x) > 1.0) return 0.0;
    int k = n / 2;
    double s = 0.0;
    for (int i = k; i >= 1; i--) {
        s *= x;
        double f = signbit(x) ? -v[i - 1 + k] : v[i - 1 + k];
        s += v[i - 1] + f;
    }
    return s;
}

/* 1d_dg_functions.jl:38-41 */
static double LegendreP(int k, double x) { return array2poly(g_leg + k * 2 * (KMAX + 1), 2 * (KMAX + 1), x); }

/* 1d_dg_functions.jl:66-69 */
static double hfun(int k, int mode, double x) { return array2poly(g_dg[k] + (mode - 1) * 2 * k, 2 * k, x); }

/* dg_methods.jl:27-36  v(k, level, cell, mode, x) */
static double vfun(int k, int level, int cell, int mode, double x) {
    if (level == 0) return LegendreP(mode - 1, 2 * x - 1) * sqrt(2.0);
    return hfun(k, mode, (double)(1 << level) * x - (double)(2 * cell - 1)) * sqrt((double)(1 << level));
}

/* 1d_dg_functions.jl:83-93 */
static double leg(int mode, double x) { return sqrt(2.0) * LegendreP(mode - 1, 2 * x - 1); }
static double basis_pos(int level, int cell, int mode, double x) {
    return leg(mode, (double)(1 << level) * x - (double)(cell - 1)) * pow(2.0, level / 2.0);
}

int gsgo_eval_v(int k, int level, int cell, int mode, const double *x, double *y, i64 npts) {
    init_globals();
    for (i64 i = 0; i < npts; i++) y[i] = vfun(k, level, cell, mode, x[i]);
    return 0;
}

int gsgo_eval_basis(int level, int cell, int mode, const double *x, double *y, i64 npts) {
    init_globals();
    for (i64 i = 0; i < npts; i++) y[i] = basis_pos(level, cell, mode, x[i]);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Gauss-Legendre nodes/weights on [-1,1] by Newton on P_n (textbook).  Replaces the
 * reference's hquadrature for polynomial integrands (exact there); see the header. */
static void gauss_legendre(int n, double *x, double *w) {
    for (int i = 0; i < n; i++) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            double z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) < 1e-16) break;
        }
        x[i] = -z;
        w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
}

#define NGL 12
static double g_glx[NGL], g_glw[NGL];
static int g_gl_init = 0;
static void init_gl(void) {
    if (g_gl_init) return;
    gauss_legendre(NGL, g_glx, g_glw);
    g_gl_init = 1;
}

/* 1-D hierarchical index helpers: additional_tools.jl:18-20 get_index_1D, 0-based */
static i64 idx1d(int k, int level, int cell, int mode) { /* level 0-based, cell/mode 1-based */
    i64 base = level == 0 ? 0 : ((i64)1 << (level - 1));
    return (i64)k * (base + cell - 1) + (mode - 1);
}
static void dec1d(int k, i64 i, int *level, int *cell, int *mode) {
    i64 c = i / k; /* global 1-D cell index */
    *mode = (int)(i % k) + 1;
    if (c == 0) { *level = 0; *cell = 1; return; }
    int l = 1;
    while (((i64)1 << l) <= c) l++;
    *level = l;
    *cell = (int)(c - ((i64)1 << (l - 1))) + 1;
}

/* 1d_dg_functions.jl:242-263  hier2pos(k, max_level): Q[pos, hier], dense row-major
 * (k 2^n)^2.  Each entry is pos_vcoeffs_DG's integral over one level-n cell
 * (1d_dg_functions.jl:103-117) of a polynomial of degree <= 2k-2 (the hierarchical
 * function is smooth on every level-n cell); 12-point Gauss-Legendre is exact there.
 * Entries with |val| <= ABS_TOL = 1e-12 are dropped as at :252. */
int gsgo_hier2pos(int k, int n, double *Q) {
    if (k < 1 || k > KMAX || n < 0 || n > 12) return -1;
    init_globals();
    init_gl();
    i64 N = (i64)k << n;
    memset(Q, 0, sizeof(double) * N * N);
    double h = 1.0 / (double)((i64)1 << n);
    for (i64 j = 0; j < N; j++) {
        int level, cell, mode;
        dec1d(k, j, &level, &cell, &mode);
        for (i64 fc = 1; fc <= ((i64)1 << n); fc++) {
            double a = (fc - 1) * h, b = fc * h;
            for (int p = 1; p <= k; p++) {
                double s = 0.0;
                for (int q = 0; q < NGL; q++) {
                    double x = 0.5 * (a + b) + 0.5 * (b - a) * g_glx[q];
                    s += g_glw[q] * basis_pos(n, (int)fc, p, x) * vfun(k, level, cell, mode, x);
                }
                s *= 0.5 * (b - a);
                if (fabs(s) > 1e-12) Q[((fc - 1) * k + (p - 1)) * N + j] = s;
            }
        }
    }
    return 0;
}

/* derivative_matrix_elements.jl:80-94 symbolic_diff */
static void symbolic_diff(const double *v, double *ans, int n) {
    int k = n / 2;
    for (int i = 1; i <= n; i++) {
        if (i < k) ans[i - 1] = i * v[i];
        else if (i > k && i < 2 * k) ans[i - 1] = (i - k) * v[i];
        else ans[i - 1] = 0.0;
    }
}

/* derivative_matrix_elements.jl:101-103 */
static double legendreDlegendre(int m1, int m2) {
    int n = 2 * (KMAX + 1);
    double d[2 * (KMAX + 1)];
    symbolic_diff(g_leg + (m2 - 1) * n, d, n);
    return inner_vv(g_leg + (m1 - 1) * n, d, n);
}

/* 1d_derivative.jl:52-74 periodic_legvLFv, alpha = 0.  The trace points carry the
 * reference's +-5e-16 offsets (point is an exact dyadic, tiny is rounded in). */
static double periodic_legvLFv(int level, int cell1, int mode1, int cell2, int mode2) {
    double point1 = (double)(cell2 - 1) / (double)(1 << level);
    double point2 = (double)(cell2) / (double)(1 << level);
    double tiny = 5.0e-16;
    double left1 = basis_pos(level, cell1, mode1, point1 - tiny);
    double right1 = basis_pos(level, cell1, mode1, point1 + tiny);
    double left2 = basis_pos(level, cell1, mode1, point2 - tiny);
    double right2 = basis_pos(level, cell1, mode1, point2 + tiny);
    if (cell2 == (1 << level)) right2 = basis_pos(level, cell1, mode1, 0.0 + tiny);
    if (cell2 == 1) left1 = basis_pos(level, cell1, mode1, 1.0 - tiny);
    double LF1 = 0.5 * (left1 + right1);
    double LF2 = 0.5 * (left2 + right2);
    double val1 = basis_pos(level, cell2, mode2, point1 + tiny);
    double val2 = basis_pos(level, cell2, mode2, point2 - tiny);
    return LF2 * val2 - LF1 * val1;
}

/* 1d_derivative.jl:108-111 periodic_pos_DLF_matrix = (-D_matrix(k,n) + LF)'.
 * Output dense row-major (k 2^n)^2 holding the transposed matrix, as returned. */
int gsgo_pos_dlf(int k, int n, double *Aout) {
    if (k < 1 || k > KMAX || n < 0 || n > 12) return -1;
    init_globals();
    i64 N = (i64)k << n;
    int nc = 1 << n;
    for (i64 t = 0; t < N * N; t++) Aout[t] = 0.0;
    for (int c1 = 1; c1 <= nc; c1++)
        for (int m1 = 1; m1 <= k; m1++) {
            i64 i = (i64)(c1 - 1) * k + (m1 - 1);
            for (int c2 = 1; c2 <= nc; c2++)
                for (int m2 = 1; m2 <= k; m2++) {
                    i64 j = (i64)(c2 - 1) * k + (m2 - 1);
                    /* 1d_derivative.jl:21-44 interior term, dropped below 1e-15 */
                    double dv = 0.0;
                    if (c1 == c2) {
                        dv = (double)(1 << (n + 1)) * legendreDlegendre(m1, m2);
                        if (!(fabs(dv) > 1.0e-15)) dv = 0.0;
                    }
                    /* 1d_derivative.jl:76-100 flux term, dropped below 1e-15 */
                    double lf = 0.0;
                    if (c1 == c2 || abs(c1 - c2) == 1 || abs(c1 - c2) == nc - 1) {
                        lf = periodic_legvLFv(n, c1, m1, c2, m2);
                        if (!(fabs(lf) > 1.0e-15)) lf = 0.0;
                    }
                    double a = (-dv) + lf;
                    Aout[j * N + i] = a; /* transpose */
                }
        }
    return 0;
}

/* 1d_derivative.jl:113-117 periodic_hier_DLF_matrix = Q' * (A * Q), dense row-major.
 * Products accumulate in ascending inner index, as SparseArrays' spmatmul does. */
int gsgo_hier_dlf(int k, int n, double *Dout) {
    if (k < 1 || k > KMAX || n < 0 || n > 11) return -1;
    i64 N = (i64)k << n;
    double *Q = malloc(sizeof(double) * N * N);
    double *A = malloc(sizeof(double) * N * N);
    double *AQ = calloc((size_t)(N * N), sizeof(double));
    if (!Q || !A || !AQ) { free(Q); free(A); free(AQ); return -2; }
    gsgo_hier2pos(k, n, Q);
    gsgo_pos_dlf(k, n, A);
    /* A is block-banded (periodic); only nonzeros matter to the sum. */
    for (i64 f = 0; f < N; f++)
        for (i64 g = 0; g < N; g++) {
            double a = A[f * N + g];
            if (a == 0.0) continue;
            for (i64 j = 0; j < N; j++) AQ[f * N + j] += a * Q[g * N + j];
        }
    for (i64 t = 0; t < N * N; t++) Dout[t] = 0.0;
    for (i64 f = 0; f < N; f++)
        for (i64 i = 0; i < N; i++) {
            double q = Q[f * N + i];
            if (q == 0.0) continue;
            for (i64 j = 0; j < N; j++) Dout[i * N + j] += q * AQ[f * N + j];
        }
    free(Q); free(A); free(AQ);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Sparse-grid layout: schemes.jl:21-27, dg_vmethods.jl:35-45,123-142 */
static int cells_of(int L1) { return 1 << (L1 - 2 > 0 ? L1 - 2 : 0); } /* L1 = 1-based level */
static i64 ipow_i(i64 b, int e) { i64 r = 1; while (e-- > 0) r *= b; return r; }

/* iterate CartesianIndices((n+1,...)) in Julia order (dimension 1 fastest) */
static int next_multi(int *L, int D, int maxv) { /* 1-based values 1..maxv */
    for (int d = 0; d < D; d++) {
        if (L[d] < maxv) { L[d]++; return 1; }
        L[d] = 1;
    }
    return 0;
}

static int cutoff(const int *L, int D, int n, int full) {
    if (full) return 0;
    int s = 0;
    for (int d = 0; d < D; d++) s += L[d];
    return s > n + D;
}

i64 gsgo_get_size(int D, int k, int n, int full) {
    if (D < 1 || D > 12 || k < 1 || n < 0) return -1;
    int L[12];
    for (int d = 0; d < D; d++) L[d] = 1;
    i64 size = 0, kD = ipow_i(k, D);
    do {
        if (cutoff(L, D, n, full)) continue;
        i64 p = 1;
        for (int d = 0; d < D; d++) p *= cells_of(L[d]);
        size += p * kD;
    } while (next_multi(L, D, n + 1));
    return size;
}

/* V2Dref: lev/cell/mode arrays of N x D (1-based values) in vhier order */
int gsgo_v2d(int D, int k, int n, int full, int32_t *lev, int32_t *cell, int32_t *mode) {
    int L[12], C[12], M[12], nc[12];
    for (int d = 0; d < D; d++) L[d] = 1;
    i64 j = 0;
    do {
        if (cutoff(L, D, n, full)) continue;
        for (int d = 0; d < D; d++) { nc[d] = cells_of(L[d]); C[d] = 1; }
        do {
            for (int d = 0; d < D; d++) M[d] = 1;
            do {
                for (int d = 0; d < D; d++) {
                    lev[j * D + d] = L[d]; cell[j * D + d] = C[d]; mode[j * D + d] = M[d];
                }
                j++;
            } while (next_multi(M, D, k));
            /* advance cells (mixed radix) */
            int d = 0;
            for (; d < D; d++) {
                if (C[d] < nc[d]) { C[d]++; break; }
                C[d] = 1;
            }
            if (d == D) break;
        } while (1);
    } while (next_multi(L, D, n + 1));
    return 0;
}

/* Dense map (level,cell,mode) -> index: levels are mixed radix (n+1)^D; within a
 * level block the index is cell_lin*k^D + mode_lin (dimension 1 fastest). */
typedef struct {
    int D, k, n, full;
    i64 N;
    i64 *blk_off; /* (n+1)^D entries, -1 when cut */
} layout_t;

static layout_t *layout_new(int D, int k, int n, int full) {
    layout_t *lt = calloc(1, sizeof(layout_t));
    lt->D = D; lt->k = k; lt->n = n; lt->full = full;
    i64 nb = ipow_i(n + 1, D);
    lt->blk_off = malloc(sizeof(i64) * nb);
    int L[12];
    for (int d = 0; d < D; d++) L[d] = 1;
    i64 off = 0, kD = ipow_i(k, D);
    do {
        i64 code = 0;
        for (int d = D - 1; d >= 0; d--) code = code * (n + 1) + (L[d] - 1);
        if (cutoff(L, D, n, full)) { lt->blk_off[code] = -1; continue; }
        lt->blk_off[code] = off;
        i64 p = 1;
        for (int d = 0; d < D; d++) p *= cells_of(L[d]);
        off += p * kD;
    } while (next_multi(L, D, n + 1));
    lt->N = off;
    return lt;
}
static void layout_free(layout_t *lt) { if (lt) { free(lt->blk_off); free(lt); } }

static i64 layout_index(const layout_t *lt, const int *L, const int *C, const int *M) {
    int D = lt->D, k = lt->k, n = lt->n;
    i64 code = 0;
    for (int d = D - 1; d >= 0; d--) code = code * (n + 1) + (L[d] - 1);
    i64 off = lt->blk_off[code];
    if (off < 0) return -1;
    i64 cl = 0, ml = 0;
    for (int d = D - 1; d >= 0; d--) { cl = cl * cells_of(L[d]) + (C[d] - 1); ml = ml * k + (M[d] - 1); }
    return off + cl * ipow_i(k, D) + ml;
}

int gsgo_d2v_index(int D, int k, int n, int full, const int32_t *lev, const int32_t *cell,
                   const int32_t *mode, i64 count, i64 *out) {
    layout_t *lt = layout_new(D, k, n, full);
    int L[12], C[12], M[12];
    for (i64 t = 0; t < count; t++) {
        for (int d = 0; d < D; d++) { L[d] = lev[t * D + d]; C[d] = cell[t * D + d]; M[d] = mode[t * D + d]; }
        out[t] = layout_index(lt, L, C, M);
    }
    layout_free(lt);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* CSC matrices (0-based storage), the SparseMatrixCSC of the reference */
typedef struct {
    i64 m, n, nnz;
    i64 *colptr, *rowval;
    double *nzval;
} gsgo_csc;

void gsgo_csc_free(gsgo_csc *A) {
    if (!A) return;
    free(A->colptr); free(A->rowval); free(A->nzval); free(A);
}
i64 gsgo_csc_nnz(const gsgo_csc *A) { return A->nnz; }
i64 gsgo_csc_rows(const gsgo_csc *A) { return A->m; }
int gsgo_csc_export(const gsgo_csc *A, i64 *colptr, i64 *rowval, double *nzval) {
    memcpy(colptr, A->colptr, sizeof(i64) * (A->n + 1));
    memcpy(rowval, A->rowval, sizeof(i64) * A->nnz);
    memcpy(nzval, A->nzval, sizeof(double) * A->nnz);
    return 0;
}

typedef struct { i64 r; double v; } rv_t;
static int cmp_rv(const void *a, const void *b) {
    i64 x = ((const rv_t *)a)->r, y = ((const rv_t *)b)->r;
    return (x > y) - (x < y);
}

/* multidim_derivative.jl:15-58  D_matrix(d, k, n, srefVD, srefDV; scheme): for each
 * column j take the 1-D column of periodic_DLF_matrix(k, n), substitute coordinate
 * d, drop cut rows, look the row up, then sparse(I, J, V, len, len, +).  d is
 * 1-based.  D1 is the dense hier 1-D operator (k 2^n)^2. */
static gsgo_csc *d_matrix_from(int D, int d, int k, int n, int full, const double *D1) {
    layout_t *lt = layout_new(D, k, n, full);
    i64 N = lt->N, N1 = (i64)k << n;
    int32_t *lev = malloc(sizeof(int32_t) * N * D), *cel = malloc(sizeof(int32_t) * N * D),
            *mod = malloc(sizeof(int32_t) * N * D);
    gsgo_v2d(D, k, n, full, lev, cel, mod);
    /* column lists of the 1-D operator (ascending rows, nonzeros only) */
    i64 *cp1 = calloc(N1 + 1, sizeof(i64));
    for (i64 j = 0; j < N1; j++) {
        i64 c = 0;
        for (i64 i = 0; i < N1; i++) c += (D1[i * N1 + j] != 0.0);
        cp1[j + 1] = cp1[j] + c;
    }
    i64 *ri1 = malloc(sizeof(i64) * (cp1[N1] + 1));
    for (i64 j = 0; j < N1; j++) {
        i64 p = cp1[j];
        for (i64 i = 0; i < N1; i++) if (D1[i * N1 + j] != 0.0) ri1[p++] = i;
    }
    gsgo_csc *A = calloc(1, sizeof(gsgo_csc));
    A->m = N; A->n = N;
    A->colptr = calloc(N + 1, sizeof(i64));
    i64 cap = 1024, nnz = 0;
    A->rowval = malloc(sizeof(i64) * cap);
    A->nzval = malloc(sizeof(double) * cap);
    rv_t *buf = malloc(sizeof(rv_t) * (N1 + 1));
    int L[12], C[12], M[12];
    for (i64 j = 0; j < N; j++) {
        for (int e = 0; e < D; e++) { L[e] = lev[j * D + e]; C[e] = cel[j * D + e]; M[e] = mod[j * D + e]; }
        int l = L[d - 1], c = C[d - 1], m = M[d - 1];
        i64 j1 = idx1d(k, l - 1, c, m);
        i64 cnt = 0;
        for (i64 p = cp1[j1]; p < cp1[j1 + 1]; p++) {
            i64 i1 = ri1[p];
            int l2, c2, m2;
            dec1d(k, i1, &l2, &c2, &m2);
            L[d - 1] = l2 + 1; C[d - 1] = c2; M[d - 1] = m2;
            if (cutoff(L, D, n, full)) continue;
            i64 i = layout_index(lt, L, C, M);
            buf[cnt].r = i; buf[cnt].v = D1[i1 * N1 + j1]; cnt++;
        }
        qsort(buf, cnt, sizeof(rv_t), cmp_rv);
        if (nnz + cnt > cap) {
            while (nnz + cnt > cap) cap *= 2;
            A->rowval = realloc(A->rowval, sizeof(i64) * cap);
            A->nzval = realloc(A->nzval, sizeof(double) * cap);
        }
        for (i64 t = 0; t < cnt; t++) { A->rowval[nnz] = buf[t].r; A->nzval[nnz] = buf[t].v; nnz++; }
        A->colptr[j + 1] = nnz;
    }
    A->nnz = nnz;
    free(buf); free(cp1); free(ri1); free(lev); free(cel); free(mod);
    layout_free(lt);
    return A;
}

gsgo_csc *gsgo_d_matrix(int D, int d, int k, int n, int full) {
    if (D < 1 || D > 12 || d < 1 || d > D || k < 1 || k > KMAX || n < 0 || n > 11) return NULL;
    i64 N1 = (i64)k << n;
    double *D1 = malloc(sizeof(double) * N1 * N1);
    if (gsgo_hier_dlf(k, n, D1)) { free(D1); return NULL; }
    gsgo_csc *A = d_matrix_from(D, d, k, n, full, D1);
    free(D1);
    return A;
}

/* SparseArrays spmatmul C = A*B: for each column j of B, for each nonzero B[p,j]
 * (ascending p), scatter A[:,p]*B[p,j]; rows of C are then sorted. */
static gsgo_csc *csc_mul(const gsgo_csc *A, const gsgo_csc *B) {
    i64 m = A->m, n = B->n;
    gsgo_csc *C = calloc(1, sizeof(gsgo_csc));
    C->m = m; C->n = n;
    C->colptr = calloc(n + 1, sizeof(i64));
    double *acc = calloc(m, sizeof(double));
    i64 *mark = malloc(sizeof(i64) * m);
    for (i64 i = 0; i < m; i++) mark[i] = -1;
    i64 *rows = malloc(sizeof(i64) * m);
    i64 cap = 1024, nnz = 0;
    C->rowval = malloc(sizeof(i64) * cap);
    C->nzval = malloc(sizeof(double) * cap);
    for (i64 j = 0; j < n; j++) {
        i64 cnt = 0;
        for (i64 pb = B->colptr[j]; pb < B->colptr[j + 1]; pb++) {
            i64 p = B->rowval[pb];
            double b = B->nzval[pb];
            for (i64 pa = A->colptr[p]; pa < A->colptr[p + 1]; pa++) {
                i64 r = A->rowval[pa];
                if (mark[r] != j) { mark[r] = j; acc[r] = 0.0; rows[cnt++] = r; }
                acc[r] += A->nzval[pa] * b;
            }
        }
        /* sort rows */
        rv_t *tmp = malloc(sizeof(rv_t) * (cnt + 1));
        for (i64 t = 0; t < cnt; t++) { tmp[t].r = rows[t]; tmp[t].v = acc[rows[t]]; }
        qsort(tmp, cnt, sizeof(rv_t), cmp_rv);
        if (nnz + cnt > cap) {
            while (nnz + cnt > cap) cap *= 2;
            C->rowval = realloc(C->rowval, sizeof(i64) * cap);
            C->nzval = realloc(C->nzval, sizeof(double) * cap);
        }
        for (i64 t = 0; t < cnt; t++) { C->rowval[nnz] = tmp[t].r; C->nzval[nnz] = tmp[t].v; nnz++; }
        free(tmp);
        C->colptr[j + 1] = nnz;
    }
    C->nnz = nnz;
    free(acc); free(mark); free(rows);
    return C;
}

/* C = A + B (same shape), union pattern, A's value first */
static gsgo_csc *csc_add(const gsgo_csc *A, const gsgo_csc *B) {
    gsgo_csc *C = calloc(1, sizeof(gsgo_csc));
    C->m = A->m; C->n = A->n;
    C->colptr = calloc(A->n + 1, sizeof(i64));
    i64 cap = A->nnz + B->nnz + 1, nnz = 0;
    C->rowval = malloc(sizeof(i64) * cap);
    C->nzval = malloc(sizeof(double) * cap);
    for (i64 j = 0; j < A->n; j++) {
        i64 pa = A->colptr[j], pb = B->colptr[j], ea = A->colptr[j + 1], eb = B->colptr[j + 1];
        while (pa < ea || pb < eb) {
            if (pb >= eb || (pa < ea && A->rowval[pa] < B->rowval[pb])) {
                C->rowval[nnz] = A->rowval[pa]; C->nzval[nnz++] = A->nzval[pa++];
            } else if (pa >= ea || B->rowval[pb] < A->rowval[pa]) {
                C->rowval[nnz] = B->rowval[pb]; C->nzval[nnz++] = B->nzval[pb++];
            } else {
                C->rowval[nnz] = A->rowval[pa]; C->nzval[nnz++] = A->nzval[pa++] + B->nzval[pb++];
            }
        }
        C->colptr[j + 1] = nnz;
    }
    C->nnz = nnz;
    return C;
}

/* multidim_derivative.jl:71-79 laplacian_matrix = sum_d D_d * D_d */
gsgo_csc *gsgo_laplacian(int D, int k, int n, int full) {
    i64 N1 = (i64)k << n;
    double *D1 = malloc(sizeof(double) * N1 * N1);
    if (gsgo_hier_dlf(k, n, D1)) { free(D1); return NULL; }
    gsgo_csc *lap = NULL;
    for (int d = 1; d <= D; d++) {
        gsgo_csc *Dd = d_matrix_from(D, d, k, n, full, D1);
        gsgo_csc *DD = csc_mul(Dd, Dd);
        gsgo_csc_free(Dd);
        if (!lap) lap = DD;
        else { gsgo_csc *t = csc_add(lap, DD); gsgo_csc_free(lap); gsgo_csc_free(DD); lap = t; }
    }
    free(D1);
    return lap;
}

/* pdes.jl:22-49 wave_data: RHS = [[0, I], [L, 0]] (2N x 2N) */
gsgo_csc *gsgo_wave_rhs(const gsgo_csc *lap) {
    i64 N = lap->n;
    gsgo_csc *R = calloc(1, sizeof(gsgo_csc));
    R->m = 2 * N; R->n = 2 * N;
    R->nnz = lap->nnz + N;
    R->colptr = calloc(2 * N + 1, sizeof(i64));
    R->rowval = malloc(sizeof(i64) * R->nnz);
    R->nzval = malloc(sizeof(double) * R->nnz);
    i64 p = 0;
    for (i64 j = 0; j < N; j++) {
        for (i64 q = lap->colptr[j]; q < lap->colptr[j + 1]; q++) {
            R->rowval[p] = lap->rowval[q] + N; R->nzval[p++] = lap->nzval[q];
        }
        R->colptr[j + 1] = p;
    }
    for (i64 i = 0; i < N; i++) {
        R->rowval[p] = i; R->nzval[p++] = 1.0;
        R->colptr[N + i + 1] = p;
    }
    return R;
}

/* SparseArrays A*x: column scatter in column order */
void gsgo_spmv(const gsgo_csc *A, const double *x, double *y) {
    for (i64 i = 0; i < A->m; i++) y[i] = 0.0;
    for (i64 j = 0; j < A->n; j++) {
        double xj = x[j];
        for (i64 p = A->colptr[j]; p < A->colptr[j + 1]; p++) y[A->rowval[p]] += A->nzval[p] * xj;
    }
}

/* ------------------------------------------------------------------------- */
/* Evolution systems: wave (pdes.jl:54-70 with the RHS closure (t,x)->RHS*x) and
 * transport F(u) = -sum_d a_d (D_d u) (the pattern of pdes.jl:179,186,191). */
typedef struct {
    int kind; /* 0 wave, 1 transport */
    int D;
    i64 len;
    gsgo_csc *rhs;     /* wave */
    gsgo_csc *Ds[12];  /* transport */
    double a[12];
    /* CSR copies for the multithreaded baseline (rows ascending columns) */
    i64 *rp, *ci;
    double *va;
    double *w;
} gsgo_sys;

static void csr_of(const gsgo_csc *A, i64 **rp, i64 **ci, double **va) {
    i64 m = A->m;
    *rp = calloc(m + 1, sizeof(i64));
    *ci = malloc(sizeof(i64) * (A->nnz + 1));
    *va = malloc(sizeof(double) * (A->nnz + 1));
    for (i64 p = 0; p < A->nnz; p++) (*rp)[A->rowval[p] + 1]++;
    for (i64 i = 0; i < m; i++) (*rp)[i + 1] += (*rp)[i];
    i64 *pos = malloc(sizeof(i64) * (m + 1));
    memcpy(pos, *rp, sizeof(i64) * (m + 1));
    for (i64 j = 0; j < A->n; j++)
        for (i64 p = A->colptr[j]; p < A->colptr[j + 1]; p++) {
            i64 q = pos[A->rowval[p]]++;
            (*ci)[q] = j; (*va)[q] = A->nzval[p];
        }
    free(pos);
}

gsgo_sys *gsgo_sys_wave(int D, int k, int n, int full) {
    gsgo_csc *lap = gsgo_laplacian(D, k, n, full);
    if (!lap) return NULL;
    gsgo_sys *s = calloc(1, sizeof(gsgo_sys));
    s->kind = 0; s->D = D;
    s->rhs = gsgo_wave_rhs(lap);
    gsgo_csc_free(lap);
    s->len = s->rhs->m;
    return s;
}

gsgo_sys *gsgo_sys_transport(int D, int k, int n, int full, const double *a) {
    i64 N1 = (i64)k << n;
    double *D1 = malloc(sizeof(double) * N1 * N1);
    if (D > 12 || gsgo_hier_dlf(k, n, D1)) { free(D1); return NULL; }
    gsgo_sys *s = calloc(1, sizeof(gsgo_sys));
    s->kind = 1; s->D = D;
    for (int d = 0; d < D; d++) { s->Ds[d] = d_matrix_from(D, d + 1, k, n, full, D1); s->a[d] = a[d]; }
    free(D1);
    s->len = s->Ds[0]->m;
    s->w = malloc(sizeof(double) * s->len * D);
    return s;
}

/* Build the transport RHS as one assembled CSC (-sum_d a_d D_d), used by the
 * multithreaded CSR baseline leg. */
static void sys_prepare_csr(gsgo_sys *s) {
    if (s->rp) return;
    if (s->kind == 0) { csr_of(s->rhs, &s->rp, &s->ci, &s->va); return; }
    gsgo_csc *acc = NULL;
    for (int d = 0; d < s->D; d++) {
        gsgo_csc *t = calloc(1, sizeof(gsgo_csc));
        t->m = s->Ds[d]->m; t->n = s->Ds[d]->n; t->nnz = s->Ds[d]->nnz;
        t->colptr = malloc(sizeof(i64) * (t->n + 1)); memcpy(t->colptr, s->Ds[d]->colptr, sizeof(i64) * (t->n + 1));
        t->rowval = malloc(sizeof(i64) * t->nnz); memcpy(t->rowval, s->Ds[d]->rowval, sizeof(i64) * t->nnz);
        t->nzval = malloc(sizeof(double) * t->nnz);
        for (i64 p = 0; p < t->nnz; p++) t->nzval[p] = -s->a[d] * s->Ds[d]->nzval[p];
        if (!acc) acc = t;
        else { gsgo_csc *u = csc_add(acc, t); gsgo_csc_free(acc); gsgo_csc_free(t); acc = u; }
    }
    csr_of(acc, &s->rp, &s->ci, &s->va);
    gsgo_csc_free(acc);
}

i64 gsgo_sys_len(const gsgo_sys *s) { return s->len; }

/* The reference's RHS evaluation, column-scatter SpMV, one thread */
void gsgo_sys_rhs(gsgo_sys *s, const double *x, double *y) {
    if (s->kind == 0) { gsgo_spmv(s->rhs, x, y); return; }
    i64 N = s->len;
    for (int d = 0; d < s->D; d++) gsgo_spmv(s->Ds[d], x, s->w + d * N);
    for (i64 i = 0; i < N; i++) {
        double acc = s->a[0] * s->w[i];
        for (int d = 1; d < s->D; d++) acc += s->a[d] * s->w[d * N + i];
        y[i] = -acc;
    }
}

/* Row-parallel CSR RHS over all cores (the analogue of the optional MKLSparse path,
 * GalerkinSparseGrids.jl:5-7) */
static void sys_rhs_csr(gsgo_sys *s, const double *x, double *y) {
    i64 m = s->len;
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < m; i++) {
        double acc = 0.0;
        for (i64 p = s->rp[i]; p < s->rp[i + 1]; p++) acc += s->va[p] * x[s->ci[p]];
        y[i] = acc;
    }
}

/* Classical fixed-step RK4 (Kutta).  The reference integrates with ODE.jl's adaptive
 * ode45/ode78 (pdes.jl:62-68); fixed-step RK4 is the build's integrator.
 * nthreads <= 1: the reference's single-threaded column-scatter SpMV;
 * nthreads > 1: row-parallel CSR. */
int gsgo_sys_rk4(gsgo_sys *s, double *y, double dt, i64 nsteps, int nthreads) {
    i64 N = s->len;
    double *k = malloc(sizeof(double) * N), *xs = malloc(sizeof(double) * N), *acc = malloc(sizeof(double) * N);
    if (nthreads > 1) {
        sys_prepare_csr(s);
#ifdef _OPENMP
        omp_set_num_threads(nthreads);
#endif
    }
    for (i64 st = 0; st < nsteps; st++) {
        /* stage 1 */
        if (nthreads > 1) sys_rhs_csr(s, y, k); else gsgo_sys_rhs(s, y, k);
#pragma omp parallel for if (nthreads > 1) schedule(static)
        for (i64 i = 0; i < N; i++) { acc[i] = y[i] + (dt / 6.0) * k[i]; xs[i] = y[i] + (dt / 2.0) * k[i]; }
        /* stage 2 */
        if (nthreads > 1) sys_rhs_csr(s, xs, k); else gsgo_sys_rhs(s, xs, k);
#pragma omp parallel for if (nthreads > 1) schedule(static)
        for (i64 i = 0; i < N; i++) { acc[i] += (dt / 3.0) * k[i]; xs[i] = y[i] + (dt / 2.0) * k[i]; }
        /* stage 3 */
        if (nthreads > 1) sys_rhs_csr(s, xs, k); else gsgo_sys_rhs(s, xs, k);
#pragma omp parallel for if (nthreads > 1) schedule(static)
        for (i64 i = 0; i < N; i++) { acc[i] += (dt / 3.0) * k[i]; xs[i] = y[i] + dt * k[i]; }
        /* stage 4 */
        if (nthreads > 1) sys_rhs_csr(s, xs, k); else gsgo_sys_rhs(s, xs, k);
#pragma omp parallel for if (nthreads > 1) schedule(static)
        for (i64 i = 0; i < N; i++) y[i] = acc[i] + (dt / 6.0) * k[i];
    }
    free(k); free(xs); free(acc);
    return 0;
}

void gsgo_sys_free(gsgo_sys *s) {
    if (!s) return;
    gsgo_csc_free(s->rhs);
    for (int d = 0; d < 12; d++) gsgo_csc_free(s->Ds[d]);
    free(s->rp); free(s->ci); free(s->va); free(s->w);
    free(s);
}

/* ------------------------------------------------------------------------- */
/* dg_vmethods.jl:149-179 vcoeffs_DG for D = 1 restated with composite Gauss-Legendre
 * on each smooth piece of the hierarchical function (the reference uses adaptive
 * hcubature, rtol 1e-8, dg_methods.jl:10-12,85-93).  f is given as samples: the
 * caller evaluates f at the nodes returned by gsgo_proj1d_nodes. */
i64 gsgo_proj1d_npts(int k, int n, int panels) {
    (void)k;
    /* every 1-D cell at every level is split at its midpoint into 2 pieces, each into
     * `panels` GL panels of NGL nodes; level 0 is one piece */
    i64 cells = 1; /* level 0 */
    for (int l = 1; l <= n; l++) cells += (i64)1 << (l - 1);
    return cells * 2 * panels * NGL;
}

int gsgo_proj1d_nodes(int k, int n, int panels, double *x) {
    init_gl();
    i64 p = 0;
    for (int l = 0; l <= n; l++) {
        int nc = 1 << (l > 0 ? l - 1 : 0);
        for (int c = 1; c <= nc; c++) {
            double a = (double)(c - 1) / nc, b = (double)c / nc, mid = 0.5 * (a + b);
            double seg[3] = {a, mid, b};
            for (int s = 0; s < 2; s++)
                for (int q = 0; q < panels; q++) {
                    double pa = seg[s] + (seg[s + 1] - seg[s]) * q / panels;
                    double pb = seg[s] + (seg[s + 1] - seg[s]) * (q + 1) / panels;
                    for (int g = 0; g < NGL; g++) x[p++] = 0.5 * (pa + pb) + 0.5 * (pb - pa) * g_glx[g];
                }
        }
    }
    (void)k;
    return 0;
}

int gsgo_proj1d(int k, int n, int panels, const double *fvals, double *coeffs) {
    init_globals();
    init_gl();
    i64 p = 0, j = 0;
    for (int l = 0; l <= n; l++) {
        int nc = 1 << (l > 0 ? l - 1 : 0);
        for (int c = 1; c <= nc; c++) {
            double a = (double)(c - 1) / nc, b = (double)c / nc, mid = 0.5 * (a + b);
            double seg[3] = {a, mid, b};
            i64 p0 = p;
            for (int m = 1; m <= k; m++) {
                double s = 0.0;
                i64 pp = p0;
                for (int sg = 0; sg < 2; sg++)
                    for (int q = 0; q < panels; q++) {
                        double pa = seg[sg] + (seg[sg + 1] - seg[sg]) * q / panels;
                        double pb = seg[sg] + (seg[sg + 1] - seg[sg]) * (q + 1) / panels;
                        double part = 0.0;
                        for (int g = 0; g < NGL; g++, pp++) {
                            double x = 0.5 * (pa + pb) + 0.5 * (pb - pa) * g_glx[g];
                            part += g_glw[g] * fvals[pp] * vfun(k, l, c, m, x);
                        }
                        s += part * 0.5 * (pb - pa);
                    }
                coeffs[j++] = s;
            }
            p = p0 + 2 * panels * NGL;
        }
    }
    return 0;
}

/* tensor_construct.jl:19-63 (vector form): value = prod_d coeff_d[(l_d, c_d, m_d)] */
int gsgo_tensor_construct(int D, int k, int n, int full, const double *const *c1d, double *out) {
    i64 N = gsgo_get_size(D, k, n, full);
    int32_t *lev = malloc(sizeof(int32_t) * N * D), *cel = malloc(sizeof(int32_t) * N * D),
            *mod = malloc(sizeof(int32_t) * N * D);
    gsgo_v2d(D, k, n, full, lev, cel, mod);
    for (i64 j = 0; j < N; j++) {
        double val = 1.0;
        for (int d = 0; d < D; d++)
            val *= c1d[d][idx1d(k, lev[j * D + d] - 1, cel[j * D + d], mod[j * D + d])];
        out[j] = val;
    }
    free(lev); free(cel); free(mod);
    return 0;
}

/* pdes.jl:258-273 energy_func for one state y = [u; v] */
double gsgo_energy(int D, int k, int n, int full, const double *y) {
    i64 N1 = (i64)k << n;
    double *D1 = malloc(sizeof(double) * N1 * N1);
    gsgo_hier_dlf(k, n, D1);
    i64 N = gsgo_get_size(D, k, n, full);
    double *w = malloc(sizeof(double) * N);
    double e = 0.0;
    for (int d = 1; d <= D; d++) {
        gsgo_csc *Dd = d_matrix_from(D, d, k, n, full, D1);
        gsgo_spmv(Dd, y, w);
        double s = 0.0;
        for (i64 i = 0; i < N; i++) s += w[i] * w[i];
        e += s;
        gsgo_csc_free(Dd);
    }
    double s = 0.0;
    for (i64 i = 0; i < N; i++) s += y[N + i] * y[N + i];
    free(D1); free(w);
    return e + s;
}
