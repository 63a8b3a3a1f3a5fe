/*
 * ro_dense.c -- blocked dense complex128 linear algebra of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see ro_dense.h).  Restates the semantics of the
 * LAPACK/BLAS routines the reference calls (zgetrf, zgetrs('N'), zgemm('N','N');
 * factorization_module.f90:71, 346-349; robin_tree_module.f90:550-605) with the
 * usual right-looking blocked algorithms:
 *   - zgemm: output tiles, operands packed into split real/imaginary panels,
 *     a 4 x 4 register micro-kernel (GCC vector extensions; an AVX2/FMA build of
 *     the same source is chosen at run time where the CPU has it);
 *   - zgetrf: 32-column panels factored column by column with partial pivoting
 *     (pivot = first entry of maximal |re|+|im|, izamax's rule), interchanges
 *     applied to the rest of the matrix, a unit-lower triangular solve for the
 *     U block row and a GEMM update of the trailing matrix;
 *   - zgetrs: row interchanges, then forward and backward block substitution
 *     with GEMM updates.
 * Complex division uses Smith's algorithm (as LAPACK's zladiv does, to avoid
 * overflow of |b|^2).
 */
#include "ro_dense.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double v4d __attribute__((vector_size(32)));

enum { MR = 4, NR = 4, KC = 256, TM = 192, TN = 96, NB = 32 };

static int g_inner = 0;

void ro_dense_set_inner(int nthreads) { g_inner = nthreads > 1 ? nthreads : 0; }

/* threads for a routine's own parallel loops (0: run on the calling thread) */
static int inner_par(void) {
#ifdef _OPENMP
    return (g_inner > 1 && omp_get_active_level() < omp_get_max_active_levels()) ? g_inner : 0;
#else
    return 0;
#endif
}

static void *xmalloc(size_t n) {
    void *p = NULL;
    if (posix_memalign(&p, 64, n ? n : 64)) p = NULL;
    if (!p) { fprintf(stderr, "ro_dense: out of memory\n"); abort(); }
    return p;
}

/* Smith's complex division a / b. */
static ro_cplx cdiv(ro_cplx a, ro_cplx b) {
    double ar = creal(a), ai = cimag(a), br = creal(b), bi = cimag(b);
    if (fabs(bi) <= fabs(br)) {
        double r = bi / br, den = br + bi * r;
        return __builtin_complex((ar + ai * r) / den, (ai - ar * r) / den);
    }
    double r = br / bi, den = bi + br * r;
    return __builtin_complex((ar * r + ai) / den, (ai * r - ar) / den);
}

static double cabs1(ro_cplx z) { return fabs(creal(z)) + fabs(cimag(z)); }

/* y -= x * s over n entries (complex, explicit real arithmetic) */
static void caxpy_neg(int n, ro_cplx s, const ro_cplx *restrict x, ro_cplx *restrict y) {
    const double sr = creal(s), si = cimag(s);
    const double *xd = (const double *)x;
    double *yd = (double *)y;
    for (int i = 0; i < n; ++i) {
        const double xr = xd[2 * i], xi = xd[2 * i + 1];
        yd[2 * i] -= xr * sr - xi * si;
        yd[2 * i + 1] -= xr * si + xi * sr;
    }
}

/* ----------------------------------------------------------------------------
 * GEMM
 * -------------------------------------------------------------------------- */

/* A rows [0, mc) x k [0, kc) into MR-row panels: per k, MR real parts then MR
 * imaginary parts (rows past mc are zero). */
static void pack_a(int mc, int kc, const ro_cplx *A, int lda, double *Ap) {
    for (int i0 = 0; i0 < mc; i0 += MR) {
        const int mr = mc - i0 < MR ? mc - i0 : MR;
        for (int k = 0; k < kc; ++k) {
            const double *a = (const double *)(A + i0 + (size_t)k * lda);
            double *o = Ap + (size_t)(i0 / MR) * kc * 2 * MR + (size_t)k * 2 * MR;
            int r = 0;
            for (; r < mr; ++r) { o[r] = a[2 * r]; o[MR + r] = a[2 * r + 1]; }
            for (; r < MR; ++r) { o[r] = 0.0; o[MR + r] = 0.0; }
        }
    }
}

/* B k [0, kc) x cols [0, nc) into NR-column panels: per k, NR real parts then NR
 * imaginary parts (columns past nc are zero). */
static void pack_b(int kc, int nc, const ro_cplx *B, int ldb, double *Bp) {
    for (int j0 = 0; j0 < nc; j0 += NR) {
        const int nr = nc - j0 < NR ? nc - j0 : NR;
        double *o = Bp + (size_t)(j0 / NR) * kc * 2 * NR;
        for (int k = 0; k < kc; ++k) {
            int c = 0;
            for (; c < nr; ++c) {
                const ro_cplx b = B[k + (size_t)(j0 + c) * ldb];
                o[(size_t)k * 2 * NR + c] = creal(b);
                o[(size_t)k * 2 * NR + NR + c] = cimag(b);
            }
            for (; c < NR; ++c) { o[(size_t)k * 2 * NR + c] = 0.0; o[(size_t)k * 2 * NR + NR + c] = 0.0; }
        }
    }
}

/* C tile (mc x nc) += alpha * (packed A panel) (packed B panel) over kc. */
static inline __attribute__((always_inline)) void macro_body(int mc, int nc, int kc, const double *Ap,
                                                             const double *Bp, ro_cplx alpha, ro_cplx *C,
                                                             int ldc) {
    const double alr = creal(alpha), ali = cimag(alpha);
    for (int j0 = 0; j0 < nc; j0 += NR) {
        const int nr = nc - j0 < NR ? nc - j0 : NR;
        const double *bp = Bp + (size_t)(j0 / NR) * kc * 2 * NR;
        for (int i0 = 0; i0 < mc; i0 += MR) {
            const int mr = mc - i0 < MR ? mc - i0 : MR;
            const double *ap = Ap + (size_t)(i0 / MR) * kc * 2 * MR;
            v4d r0 = {0, 0, 0, 0}, r1 = r0, r2 = r0, r3 = r0, q0 = r0, q1 = r0, q2 = r0, q3 = r0;
            for (int k = 0; k < kc; ++k) {
                const v4d ar = *(const v4d *)(ap + 8 * k);
                const v4d ai = *(const v4d *)(ap + 8 * k + 4);
                const double *b = bp + 8 * k;
                v4d br, bi;
                br = (v4d){b[0], b[0], b[0], b[0]}; bi = (v4d){b[4], b[4], b[4], b[4]};
                r0 += ar * br; r0 -= ai * bi; q0 += ar * bi; q0 += ai * br;
                br = (v4d){b[1], b[1], b[1], b[1]}; bi = (v4d){b[5], b[5], b[5], b[5]};
                r1 += ar * br; r1 -= ai * bi; q1 += ar * bi; q1 += ai * br;
                br = (v4d){b[2], b[2], b[2], b[2]}; bi = (v4d){b[6], b[6], b[6], b[6]};
                r2 += ar * br; r2 -= ai * bi; q2 += ar * bi; q2 += ai * br;
                br = (v4d){b[3], b[3], b[3], b[3]}; bi = (v4d){b[7], b[7], b[7], b[7]};
                r3 += ar * br; r3 -= ai * bi; q3 += ar * bi; q3 += ai * br;
            }
            const v4d rr[4] = {r0, r1, r2, r3}, qq[4] = {q0, q1, q2, q3};
            for (int c = 0; c < nr; ++c) {
                double *cd = (double *)(C + i0 + (size_t)(j0 + c) * ldc);
                for (int r = 0; r < mr; ++r) {
                    const double sr = rr[c][r], si = qq[c][r];
                    cd[2 * r] += alr * sr - ali * si;
                    cd[2 * r + 1] += alr * si + ali * sr;
                }
            }
        }
    }
}

__attribute__((target("avx2,fma"))) static void macro_avx2(int mc, int nc, int kc, const double *Ap,
                                                           const double *Bp, ro_cplx alpha, ro_cplx *C, int ldc) {
    macro_body(mc, nc, kc, Ap, Bp, alpha, C, ldc);
}

static void macro_base(int mc, int nc, int kc, const double *Ap, const double *Bp, ro_cplx alpha, ro_cplx *C,
                       int ldc) {
    macro_body(mc, nc, kc, Ap, Bp, alpha, C, ldc);
}

static int have_avx2(void) {
    static int v = -1;
    if (v < 0) {
        __builtin_cpu_init();
        v = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
    }
    return v;
}

/* one output tile: C[i0:i0+mc, j0:j0+nc] = alpha A B + beta C */
static void gemm_tile(int mc, int nc, int k, ro_cplx alpha, const ro_cplx *A, int lda, const ro_cplx *B, int ldb,
                      ro_cplx beta, ro_cplx *C, int ldc, double *Ap, double *Bp, int avx2) {
    for (int j = 0; j < nc; ++j) {
        ro_cplx *c = C + (size_t)j * ldc;
        if (beta == 0.0) memset(c, 0, sizeof(ro_cplx) * (size_t)mc);
        else if (beta != 1.0) for (int i = 0; i < mc; ++i) c[i] *= beta;
    }
    if (alpha == 0.0) return;
    for (int p0 = 0; p0 < k; p0 += KC) {
        const int kc = k - p0 < KC ? k - p0 : KC;
        pack_a(mc, kc, A + (size_t)p0 * lda, lda, Ap);
        pack_b(kc, nc, B + p0, ldb, Bp);
        if (avx2) macro_avx2(mc, nc, kc, Ap, Bp, alpha, C, ldc);
        else macro_base(mc, nc, kc, Ap, Bp, alpha, C, ldc);
    }
}

void ro_zgemm(int m, int n, int k, ro_cplx alpha, const ro_cplx *A, int lda, const ro_cplx *B, int ldb,
              ro_cplx beta, ro_cplx *C, int ldc) {
    if (m <= 0 || n <= 0) return;
    if ((double)m * n * (k > 0 ? k : 1) < 4096.0 || n < 2 || k < 1) {
        /* small products: plain column updates (zgemm's reference loop order) */
        for (int j = 0; j < n; ++j) {
            ro_cplx *c = C + (size_t)j * ldc;
            if (beta == 0.0) memset(c, 0, sizeof(ro_cplx) * (size_t)m);
            else if (beta != 1.0) for (int i = 0; i < m; ++i) c[i] *= beta;
            for (int l = 0; l < k; ++l) {
                const ro_cplx bv = alpha * B[l + (size_t)j * ldb];
                if (bv == 0.0) continue;
                caxpy_neg(m, -bv, A + (size_t)l * lda, c);
            }
        }
        return;
    }
    const int avx2 = have_avx2();
    const int tm = (m + TM - 1) / TM, tn = (n + TN - 1) / TN, ntask = tm * tn;
    const size_t abytes = (size_t)TM * KC * 2 * sizeof(double), bbytes = (size_t)KC * (TN + NR) * 2 * sizeof(double);
    const int nth = ntask > 1 ? inner_par() : 0;
    if (nth) {
#pragma omp parallel num_threads(nth)
        {
            double *Ap = xmalloc(abytes), *Bp = xmalloc(bbytes);
#pragma omp for schedule(dynamic, 1)
            for (int t = 0; t < ntask; ++t) {
                const int it = t % tm, jt = t / tm;
                const int i0 = it * TM, j0 = jt * TN;
                const int mc = m - i0 < TM ? m - i0 : TM, nc = n - j0 < TN ? n - j0 : TN;
                gemm_tile(mc, nc, k, alpha, A + i0, lda, B + (size_t)j0 * ldb, ldb, beta, C + i0 + (size_t)j0 * ldc,
                          ldc, Ap, Bp, avx2);
            }
            free(Ap);
            free(Bp);
        }
        return;
    }
    double *Ap = xmalloc(abytes), *Bp = xmalloc(bbytes);
    for (int t = 0; t < ntask; ++t) {
        const int it = t % tm, jt = t / tm;
        const int i0 = it * TM, j0 = jt * TN;
        const int mc = m - i0 < TM ? m - i0 : TM, nc = n - j0 < TN ? n - j0 : TN;
        gemm_tile(mc, nc, k, alpha, A + i0, lda, B + (size_t)j0 * ldb, ldb, beta, C + i0 + (size_t)j0 * ldc, ldc, Ap,
                  Bp, avx2);
    }
    free(Ap);
    free(Bp);
}

/* ----------------------------------------------------------------------------
 * LU
 * -------------------------------------------------------------------------- */

/* Row interchanges j in [j0, j1) (row j <-> row piv[j]) on columns [c0, c1). */
static void swap_rows(ro_cplx *A, int lda, const int *piv, int j0, int j1, int c0, int c1) {
    for (int c = c0; c < c1; ++c) {
        ro_cplx *col = A + (size_t)c * lda;
        for (int j = j0; j < j1; ++j) {
            const int p = piv[j];
            if (p != j) { const ro_cplx t = col[j]; col[j] = col[p]; col[p] = t; }
        }
    }
}

/* B[j0:j0+jb, cols] <- L11^-1 B[j0:j0+jb, cols], L11 unit lower (column-major in LU) */
static void trsm_lower_unit(const ro_cplx *LU, int ld, int j0, int jb, ro_cplx *B, int ldb, int ncols) {
    for (int c = 0; c < ncols; ++c) {
        ro_cplx *b = B + (size_t)c * ldb;
        for (int j = j0; j < j0 + jb; ++j) {
            const ro_cplx v = b[j];
            if (v == 0.0) continue;
            caxpy_neg(j0 + jb - j - 1, v, LU + (j + 1) + (size_t)j * ld, b + j + 1);
        }
    }
}

/* B[j0:j0+jb, cols] <- U11^-1 B[j0:j0+jb, cols], U11 upper */
static void trsm_upper(const ro_cplx *LU, int ld, int j0, int jb, ro_cplx *B, int ldb, int ncols) {
    for (int c = 0; c < ncols; ++c) {
        ro_cplx *b = B + (size_t)c * ldb;
        for (int j = j0 + jb - 1; j >= j0; --j) {
            b[j] = cdiv(b[j], LU[j + (size_t)j * ld]);
            const ro_cplx v = b[j];
            if (v == 0.0) continue;
            caxpy_neg(j - j0, v, LU + j0 + (size_t)j * ld, b + j0);
        }
    }
}

int ro_zgetrf(int n, ro_cplx *A, int *piv) {
    int info = 0;
    const int par = inner_par();
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int jb = n - j0 < NB ? n - j0 : NB;
        /* panel: columns [j0, j0+jb), rows [j0, n) */
        for (int j = j0; j < j0 + jb; ++j) {
            ro_cplx *cj = A + (size_t)j * n;
            int p = j;
            double best = cabs1(cj[j]);
            for (int i = j + 1; i < n; ++i) {
                const double v = cabs1(cj[i]);
                if (v > best) { best = v; p = i; }
            }
            piv[j] = p;
            if (best == 0.0) { if (!info) info = j + 1; continue; }
            if (p != j)
                for (int c = j0; c < j0 + jb; ++c) {
                    ro_cplx *col = A + (size_t)c * n;
                    const ro_cplx t = col[j]; col[j] = col[p]; col[p] = t;
                }
            const ro_cplx r = cdiv(1.0, cj[j]);
            for (int i = j + 1; i < n; ++i) cj[i] *= r;
            for (int c = j + 1; c < j0 + jb; ++c) {
                ro_cplx *col = A + (size_t)c * n;
                const ro_cplx u = col[j];
                if (u == 0.0) continue;
                caxpy_neg(n - j - 1, u, cj + j + 1, col + j + 1);
            }
        }
        const int rest = n - j0 - jb;
        /* interchanges on the columns left and right of the panel */
        if (par) {
#pragma omp parallel for schedule(static) num_threads(par)
            for (int c = 0; c < j0; ++c) swap_rows(A, n, piv, j0, j0 + jb, c, c + 1);
#pragma omp parallel for schedule(static) num_threads(par)
            for (int c = j0 + jb; c < n; ++c) swap_rows(A, n, piv, j0, j0 + jb, c, c + 1);
        } else {
            swap_rows(A, n, piv, j0, j0 + jb, 0, j0);
            swap_rows(A, n, piv, j0, j0 + jb, j0 + jb, n);
        }
        if (rest <= 0) continue;
        /* U12 = L11^-1 A12 */
        ro_cplx *A12 = A + (size_t)(j0 + jb) * n;
        if (par) {
#pragma omp parallel for schedule(static) num_threads(par)
            for (int c = 0; c < rest; ++c) trsm_lower_unit(A, n, j0, jb, A12 + (size_t)c * n, n, 1);
        } else {
            trsm_lower_unit(A, n, j0, jb, A12, n, rest);
        }
        /* A22 -= A21 U12 */
        ro_zgemm(rest, rest, jb, -1.0, A + (j0 + jb) + (size_t)j0 * n, n, A12 + j0, n, 1.0, A12 + j0 + jb, n);
    }
    return info;
}

void ro_zgetrs(int n, const ro_cplx *LU, const int *piv, int nrhs, ro_cplx *B, int ldb) {
    if (n <= 0 || nrhs <= 0) return;
    if (nrhs <= 2) {
        /* one or two columns: substitution column by column */
        for (int c = 0; c < nrhs; ++c) {
            ro_cplx *b = B + (size_t)c * ldb;
            for (int i = 0; i < n; ++i)
                if (piv[i] != i) { const ro_cplx t = b[i]; b[i] = b[piv[i]]; b[piv[i]] = t; }
            for (int j = 0; j < n; ++j) {
                const ro_cplx v = b[j];
                if (v == 0.0) continue;
                caxpy_neg(n - j - 1, v, LU + (j + 1) + (size_t)j * n, b + j + 1);
            }
            for (int j = n - 1; j >= 0; --j) {
                b[j] = cdiv(b[j], LU[j + (size_t)j * n]);
                const ro_cplx v = b[j];
                if (v == 0.0) continue;
                caxpy_neg(j, v, LU + (size_t)j * n, b);
            }
        }
        return;
    }
    const int par = inner_par();
    if (par) {
#pragma omp parallel for schedule(static) num_threads(par)
        for (int c = 0; c < nrhs; ++c) swap_rows(B + (size_t)c * ldb, ldb, piv, 0, n, 0, 1);
    } else {
        swap_rows(B, ldb, piv, 0, n, 0, nrhs);
    }
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int jb = n - j0 < NB ? n - j0 : NB;
        if (par) {
#pragma omp parallel for schedule(static) num_threads(par)
            for (int c = 0; c < nrhs; ++c) trsm_lower_unit(LU, n, j0, jb, B + (size_t)c * ldb, ldb, 1);
        } else {
            trsm_lower_unit(LU, n, j0, jb, B, ldb, nrhs);
        }
        if (j0 + jb < n)
            ro_zgemm(n - j0 - jb, nrhs, jb, -1.0, LU + (j0 + jb) + (size_t)j0 * n, n, B + j0, ldb, 1.0,
                     B + j0 + jb, ldb);
    }
    for (int j0 = ((n - 1) / NB) * NB; j0 >= 0; j0 -= NB) {
        const int jb = n - j0 < NB ? n - j0 : NB;
        if (par) {
#pragma omp parallel for schedule(static) num_threads(par)
            for (int c = 0; c < nrhs; ++c) trsm_upper(LU, n, j0, jb, B + (size_t)c * ldb, ldb, 1);
        } else {
            trsm_upper(LU, n, j0, jb, B, ldb, nrhs);
        }
        if (j0 > 0) ro_zgemm(j0, nrhs, jb, -1.0, LU + (size_t)j0 * n, n, B + j0, ldb, 1.0, B, ldb);
    }
}
