/*
 * ro_dense.h -- dense complex128 linear algebra of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY: part of the checker for the MI355X path (see
 * rs_oracle.h), never linked into the product library.
 *
 * The reference calls LAPACK/BLAS for every dense operation of its hot path
 * (factorization_module.f90:71, 129, 346-349; robin_tree_module.f90:550-605,
 * 898-959, 1094-1109).  These routines restate the semantics of the routines it
 * calls -- zgetrf (partial pivoting, pivot = first entry of maximal |re|+|im|,
 * as izamax chooses), zgetrs('N'), zgemm('N','N') -- as blocked algorithms, so
 * that the oracle can factor the 1024^2 and 512^2 trees in seconds.
 *
 * Results do not depend on the thread count: every output element is computed
 * by one thread with a fixed summation order (the work is split over output
 * tiles, never over the summation index).
 */
#ifndef RO_DENSE_H
#define RO_DENSE_H

#include <complex.h>

typedef double complex ro_cplx;

/* Inner parallelism: with nthreads > 1 the routines below spread their own work
 * over that many OpenMP threads (a nested team when called from a parallel loop
 * over the nodes of a level, if nesting is enabled); with nthreads <= 1 they run
 * on the calling thread. */
void ro_dense_set_inner(int nthreads);

/* C(m x n, ldc) = alpha * A(m x k, lda) * B(k x n, ldb) + beta * C. */
void ro_zgemm(int m, int n, int k, ro_cplx alpha, const ro_cplx *A, int lda, const ro_cplx *B, int ldb,
              ro_cplx beta, ro_cplx *C, int ldc);

/* zgetrf: LU with partial pivoting of the n x n matrix A (lda = n), piv 0-based
 * (row j was interchanged with row piv[j]).  Returns 0, or j+1 for the first
 * exactly zero pivot (the factorization is then still completed). */
int ro_zgetrf(int n, ro_cplx *A, int *piv);

/* zgetrs('N'): B(n x nrhs, ldb) <- A^-1 B with the factors of ro_zgetrf. */
void ro_zgetrs(int n, const ro_cplx *LU, const int *piv, int nrhs, ro_cplx *B, int ldb);

#endif
