/*
 * rs_oracle.c -- CPU restatement of robinsolver's HPS factor and solve.
 *
 * TEST INFRASTRUCTURE ONLY (see rs_oracle.h): the checker for the MI355X path,
 * never part of it.  Every function cites the reference file:line it follows.
 *
 * Parallelism: leaves, and the nodes of one level, are independent in the
 * reference recursion (robin_tree_module.f90:400-401, 861-862, 1138-1139), so
 * they are processed level by level with OpenMP.  Results do not depend on the
 * thread count: no value is accumulated across leaves or nodes.
 */
#include "rs_oracle.h"
#include "ro_dense.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

typedef struct ro_node ro_node;
struct ro_node {
    int id, lvl, isleaf;
    ro_node *parent, *c1, *c2;
    double box[2][3];   /* box[0][j] lower, box[1][j] upper (derived_type_module.f90:260) */
    int ptbox[3];
    int iface;          /* 1..6 */
    int isbndry[7];     /* 1..6 */
    int pf[7];          /* pt_face, 1..6 */
    cplx *rtr[7][7];    /* RtR(row,col) blocks, column-major pf(row) x pf(col) */
    cplx *rtr_full;     /* snapshot of the node's RtR in face order (test hook) */
    int P;
    /* merge data (derived_type_module.f90:262-263) */
    int isz;
    cplx *D_lu, *S_lu;
    int *D_piv, *S_piv;
    cplx nu;
    /* leaf data */
    int m;
    cplx *A_lu;
    int *A_piv;
    /* solve data */
    cplx *g[7], *h[7];
    cplx *sol;
};

struct ro_tree {
    int d, L;
    double domain[2][3];
    double h_tgt, kb;
    double h[3];
    cplx bp[2][2];
    ro_node *root;
    ro_node **nodes;    /* by id */
    ro_node **leaves;   /* DFS order */
    ro_node ***lvl_nodes;
    int *lvl_count;
    int n_leaves, n_nodes;
    int factored;
    char err[256];
};

static char g_create_err[256];
static int g_threads = 0;

void ro_set_threads(int n) { g_threads = n; }
const char *ro_create_error(void) { return g_create_err; }
const char *ro_last_error(const ro_tree *t) { return t ? t->err : g_create_err; }

static void set_err(ro_tree *t, const char *msg) {
    snprintf(t->err, sizeof t->err, "%s", msg);
}

static void *xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz);
    if (!p) { fprintf(stderr, "rs_oracle: out of memory\n"); abort(); }
    return p;
}

static void omp_setup(void) {
#ifdef _OPENMP
    if (g_threads > 0) omp_set_num_threads(g_threads);
#endif
}

/* ============================================================================
 * Dense complex linear algebra (restated LAPACK/BLAS semantics)
 * ========================================================================== */

/* zgetrf (factorization_module.f90:71): partial pivoting, pivot = first entry of
 * maximal |re|+|im| (izamax); blocked (ro_dense.c).  Returns info (0 ok, j+1 for
 * a zero pivot). */
static int lu_factor(int n, cplx *A, int *piv) { return ro_zgetrf(n, A, piv); }

/* zgetrs('n') (factorization_module.f90:346-349): B <- A^-1 B, nrhs columns. */
static void lu_solve(int n, const cplx *LU, const int *piv, int nrhs, cplx *B, int ldb) {
    ro_zgetrs(n, LU, piv, nrhs, B, ldb);
}

/* matrixdata%invvec (factorization_module.f90:291-366), LU mode:
 *   x present: x <- alpha*x + beta*A^-1 b   (b untouched)
 *   x absent : b <- beta*A^-1 b */
static void invvec(int n, const cplx *LU, const int *piv, int nrhs, cplx *b, cplx beta,
                   cplx *x, cplx alpha) {
    if (x) {
        cplx *tmp = xcalloc((size_t)n * nrhs, sizeof(cplx));
        memcpy(tmp, b, sizeof(cplx) * (size_t)n * nrhs);
        lu_solve(n, LU, piv, nrhs, tmp, n);
        /* alpha == 0 is defined as an overwrite: the reference multiplies
         * uninitialised memory by zero (robin_tree_module.f90:868-873). */
        for (size_t q = 0; q < (size_t)n * nrhs; ++q)
            x[q] = (alpha == 0.0 ? 0.0 : alpha * x[q]) + beta * tmp[q];
        free(tmp);
    } else {
        lu_solve(n, LU, piv, nrhs, b, n);
        for (size_t q = 0; q < (size_t)n * nrhs; ++q) b[q] *= beta;
    }
}

/* zgemm('n','n'): C(ldc) = alpha*A(m x k, lda)*B(k x n, ldb) + beta*C. */
static void gemm(int m, int n, int k, cplx alpha, const cplx *A, int lda, const cplx *B,
                 int ldb, cplx beta, cplx *C, int ldc) {
    ro_zgemm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

/* zgemv('n') with alpha = beta = 1: y += A x. */
static void gemv_acc(int m, int n, const cplx *A, const cplx *x, cplx *y) {
    gemm(m, 1, n, 1.0, A, m, x, n, 1.0, y, m);
}

/* ============================================================================
 * Index maps (fd_module.f90:669-1084)
 * ========================================================================== */

static int prod(const int *s, int n) { int p = 1; for (int q = 0; q < n; ++q) p *= s[q]; return p; }

/* get_pt_face (fd_module.f90:1033-1049); pf indexed 1..6. */
static void get_pt_face(int pf[7], const int ptbox[3], int d) {
    for (int j = 0; j < 7; ++j) pf[j] = 0;
    int pr = prod(ptbox, d);
    for (int j = 1; j <= d; ++j) { pf[2 * j - 1] = pr / ptbox[j - 1]; pf[2 * j] = pf[2 * j - 1]; }
}

/* get_h (fd_module.f90:1060-1084). */
static void get_h(double h[3], int shp[3], double h_tgt, int d, const double box[2][3]) {
    for (int j = 0; j < 3; ++j) { h[j] = 0.0; shp[j] = -1; }
    for (int j = 0; j < d; ++j) {
        int n = (int)ceil((box[1][j] - box[0][j]) / h_tgt);
        shp[j] = n > 3 ? n : 3;
        h[j] = (box[1][j] - box[0][j]) / shp[j];
    }
}

/* lin2igrid (fd_module.f90:950-975), 1-based. */
static void lin2igrid(int lin, int g[3], const int shp[3], int d) {
    g[0] = g[1] = g[2] = 1;
    for (int j = d; j >= 1; --j) {
        int p = prod(shp, j - 1);
        g[j - 1] = 1 + (lin - 1) / p;
        lin -= (g[j - 1] - 1) * p;
    }
}

/* igrid2lin, ordering 1 (fd_module.f90:813-854). */
static int igrid2lin(const int g[3], const int shp[3], int d) {
    int lin = 0;
    for (int j = d; j >= 2; --j) lin += prod(shp, j - 1) * (g[j - 1] - 1);
    return lin + g[0];
}

/* lin2fgrid (fd_module.f90:867-939): boundary points first, face by face. */
static void lin2fgrid(int lin, int g[3], const int shp[3], int d) {
    int pf[7];
    get_pt_face(pf, shp, d);
    int nb = 0;
    for (int f = 1; f <= 2 * d; ++f) nb += pf[f];
    if (lin > nb) { lin2igrid(lin - nb, g, shp, d); return; }
    int off = 0, face = 1;
    while (lin > off + pf[face]) { off += pf[face]; ++face; }
    int ax = (face - 1) / 2;
    int bshape[3], t[3], q = 0;
    for (int j = 0; j < 3; ++j) if (j != ax) bshape[q++] = shp[j];
    bshape[2] = -1;
    lin2igrid(lin - off, t, bshape, d - 1);
    q = 0;
    for (int j = 0; j < 3; ++j) {
        if (j == ax) g[j] = (face % 2 == 1) ? 0 : shp[j] + 1;
        else g[j] = t[q++];
    }
}

/* grid_classify (fd_module.f90:989-1022). */
static void grid_classify(int *dist, int fn[3], const int g[3], const int shp[3], int d) {
    fn[0] = fn[1] = fn[2] = 0;
    int dm = 1 << 30;
    for (int j = 0; j < d; ++j) {
        if (g[j] < dm) dm = g[j];
        if (shp[j] - g[j] + 1 < dm) dm = shp[j] - g[j] + 1;
    }
    *dist = dm;
    if (dm <= 1)
        for (int j = 0; j < d; ++j) {
            if (g[j] == 0 || g[j] == 1) fn[j] = 2 * (j + 1) - 1;
            else if (g[j] == shp[j] || g[j] == shp[j] + 1) fn[j] = 2 * (j + 1);
        }
}

/* fgrid2lin (fd_module.f90:717-800); returns -1 for a corner/edge point. */
static int fgrid2lin(const int g[3], const int shp[3], int d) {
    int pf[7];
    get_pt_face(pf, shp, d);
    int cnt = 0, bface = 0;
    for (int j = 0; j < d; ++j) {
        if (g[j] == 0) { bface = 2 * (j + 1) - 1; ++cnt; }
        else if (g[j] == shp[j] + 1) { bface = 2 * (j + 1); ++cnt; }
    }
    if (cnt > 1) return -1;
    if (cnt == 1) {
        int ax = (bface + 1) / 2 - 1, gg[3], ss[3], q = 0;
        for (int j = 0; j < 3; ++j) if (j != ax) { gg[q] = g[j]; ss[q] = shp[j]; ++q; }
        gg[2] = 0; ss[2] = 0;
        int off = 0;
        for (int f = 1; f < bface; ++f) off += pf[f];
        return off + igrid2lin(gg, ss, d - 1);
    }
    int nb = 0;
    for (int f = 1; f <= 2 * d; ++f) nb += pf[f];
    return igrid2lin(g, shp, d) + nb;
}

/* grid2real (fd_module.f90:669-700). */
static void grid2real(double x[3], const int g[3], const int shp[3], int d, const double h[3],
                      const double box[2][3]) {
    for (int j = 0; j < 3; ++j) x[j] = box[0][j] + g[j] * h[j] - h[j] / 2;
    for (int j = 0; j < d; ++j) {
        if (g[j] == 0) x[j] = box[0][j];
        else if (g[j] == shp[j] + 1) x[j] = box[1][j];
    }
}

/* ============================================================================
 * Geometry (robin_tree_module.f90:288-363, 1369-1412)
 * ========================================================================== */

/* node_divide: cut the longest axis (first on ties); child1 is the lower half. */
static void node_divide(int d, ro_node *nd) {
    double w[3] = {0, 0, 0};
    for (int j = 0; j < d; ++j) w[j] = nd->box[1][j] - nd->box[0][j];
    int e = 0;
    for (int j = 1; j < 3; ++j) if (w[j] > w[e]) e = j;
    memcpy(nd->c1->box, nd->box, sizeof nd->box);
    memcpy(nd->c2->box, nd->box, sizeof nd->box);
    nd->c1->box[1][e] = nd->box[0][e] + w[e] * 0.5;
    nd->c2->box[0][e] = nd->c1->box[1][e];
    nd->c1->iface = 2 * (e + 1);
    nd->c2->iface = 2 * (e + 1) - 1;
    for (int j = 1; j <= 2 * d; ++j)
        if (nd->isbndry[j]) {
            if (j != nd->c1->iface) nd->c1->isbndry[j] = 1;
            if (j != nd->c2->iface) nd->c2->isbndry[j] = 1;
        }
}

/* bin_tree: pairwise pre-order ids. */
static void bin_tree(ro_node *nd, int L, int d, int *counter) {
    if (nd->isleaf) return;
    nd->c1 = xcalloc(1, sizeof(ro_node));
    nd->c2 = xcalloc(1, sizeof(ro_node));
    node_divide(d, nd);
    nd->c1->id = ++*counter; nd->c1->lvl = nd->lvl + 1; nd->c1->isleaf = (nd->c1->lvl == L);
    nd->c1->parent = nd;
    nd->c2->id = ++*counter; nd->c2->lvl = nd->lvl + 1; nd->c2->isleaf = (nd->c2->lvl == L);
    nd->c2->parent = nd;
    bin_tree(nd->c1, L, d, counter);
    bin_tree(nd->c2, L, d, counter);
}

static void collect(ro_tree *t, ro_node *nd, int *leaf_counter) {
    t->nodes[nd->id] = nd;
    t->lvl_nodes[nd->lvl][t->lvl_count[nd->lvl]++] = nd;
    if (nd->isleaf) { t->leaves[(*leaf_counter)++] = nd; return; }
    collect(t, nd->c1, leaf_counter);
    collect(t, nd->c2, leaf_counter);
}

ro_tree *ro_create(int d, int L, const double *domain, double h_tgt, double kb,
                   const double *bp) {
    g_create_err[0] = 0;
    /* robin_tree_factor validation (robin_tree_module.f90:121-140) */
    for (int j = 0; j < 3; ++j)
        if (domain[2 * j + 1] - domain[2 * j] < 0) {
            snprintf(g_create_err, sizeof g_create_err,
                     "Error in ROBIN_TREE_FACTOR: Invalid operator domain format.");
            return NULL;
        }
    if (d != 2 && d != 3) {
        snprintf(g_create_err, sizeof g_create_err,
                 "Error in ROBIN_TREE_FACTOR: Dimension must be 2 or 3.");
        return NULL;
    }
    if (L < 0 || L > 24) {
        snprintf(g_create_err, sizeof g_create_err, "Error: invalid number of levels.");
        return NULL;
    }
    cplx b[2][2];
    for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) b[i][j] = bp[2 * (i + 2 * j)] + I * bp[2 * (i + 2 * j) + 1];
    if (cabs(b[0][0] * b[1][1] - b[0][1] * b[1][0]) < pow(10.0, -3.0)) {
        snprintf(g_create_err, sizeof g_create_err,
                 "Error: Boundary parameter matrix is nearly singular.");
        return NULL;
    }
    ro_tree *t = xcalloc(1, sizeof(ro_tree));
    t->d = d; t->L = L; t->h_tgt = h_tgt; t->kb = kb;
    for (int j = 0; j < 3; ++j) { t->domain[0][j] = domain[2 * j]; t->domain[1][j] = domain[2 * j + 1]; }
    memcpy(t->bp, b, sizeof b);
    t->root = xcalloc(1, sizeof(ro_node));
    t->root->lvl = 0; t->root->isleaf = (L == 0);
    memcpy(t->root->box, t->domain, sizeof t->domain);
    for (int j = 1; j <= 2 * d; ++j) t->root->isbndry[j] = 1;
    int counter = 0;
    bin_tree(t->root, L, d, &counter);
    t->n_nodes = (1 << (L + 1)) - 1;
    t->n_leaves = 1 << L;
    t->nodes = xcalloc(t->n_nodes, sizeof(ro_node *));
    t->leaves = xcalloc(t->n_leaves, sizeof(ro_node *));
    t->lvl_nodes = xcalloc(L + 1, sizeof(ro_node **));
    t->lvl_count = xcalloc(L + 1, sizeof(int));
    for (int l = 0; l <= L; ++l) t->lvl_nodes[l] = xcalloc((size_t)1 << l, sizeof(ro_node *));
    int lc = 0;
    collect(t, t->root, &lc);
    /* Per-leaf grid (compute_RTR leaf branch, robin_tree_module.f90:392).  The
     * reference keeps the last leaf's h in T%opts%h and reuses it in the solve;
     * leaves whose grids differ would be solved with a stencil they were not
     * factored with, so that case is refused here. */
    for (int q = 0; q < t->n_leaves; ++q) {
        ro_node *lf = t->leaves[q];
        double h[3];
        get_h(h, lf->ptbox, h_tgt, d, lf->box);
        if (q == 0) memcpy(t->h, h, sizeof h);
        else if (memcmp(h, t->h, sizeof h) != 0 || memcmp(lf->ptbox, t->leaves[0]->ptbox, sizeof lf->ptbox)) {
            set_err(t, "Error: leaves have non-uniform grids (unsupported).");
            snprintf(g_create_err, sizeof g_create_err, "%s", t->err);
            ro_destroy(t);
            return NULL;
        }
        get_pt_face(lf->pf, lf->ptbox, d);
        lf->m = prod(lf->ptbox, d);
        for (int f = 1; f <= 2 * d; ++f) lf->m += lf->pf[f];
    }
    /* Parent ptbox (robin_tree_module.f90:404-410). */
    for (int l = L - 1; l >= 0; --l)
        for (int q = 0; q < t->lvl_count[l]; ++q) {
            ro_node *nd = t->lvl_nodes[l][q];
            double diff[3];
            for (int j = 0; j < 3; ++j)
                diff[j] = fabs((nd->box[1][j] - nd->box[0][j]) - (nd->c1->box[1][j] - nd->c1->box[0][j]));
            int e = 0;
            for (int j = 1; j < d; ++j) if (diff[j] > diff[e]) e = j;
            memcpy(nd->ptbox, nd->c1->ptbox, sizeof nd->ptbox);
            nd->ptbox[e] *= 2;
            get_pt_face(nd->pf, nd->ptbox, d);
        }
    return t;
}

static void free_node_solve(ro_node *nd) {
    for (int f = 0; f < 7; ++f) { free(nd->g[f]); nd->g[f] = NULL; free(nd->h[f]); nd->h[f] = NULL; }
    free(nd->sol); nd->sol = NULL;
}

static void free_node(ro_node *nd) {
    if (!nd) return;
    for (int r = 0; r < 7; ++r) for (int c = 0; c < 7; ++c) free(nd->rtr[r][c]);
    free(nd->rtr_full); free(nd->D_lu); free(nd->S_lu); free(nd->D_piv); free(nd->S_piv);
    free(nd->A_lu); free(nd->A_piv);
    free_node_solve(nd);
    free(nd);
}

void ro_destroy(ro_tree *t) {
    if (!t) return;
    for (int q = 0; q < t->n_nodes; ++q) free_node(t->nodes[q]);
    if (t->lvl_nodes) for (int l = 0; l <= t->L; ++l) free(t->lvl_nodes[l]);
    free(t->lvl_nodes); free(t->lvl_count); free(t->nodes); free(t->leaves);
    free(t);
}

int ro_info(const ro_tree *t, int *n_leaves, int *m, int *n_int, int *nb, int *ptbox, double *h) {
    const ro_node *lf = t->leaves[0];
    int ni = prod(lf->ptbox, t->d);
    if (n_leaves) *n_leaves = t->n_leaves;
    if (m) *m = lf->m;
    if (n_int) *n_int = ni;
    if (nb) *nb = lf->m - ni;
    if (ptbox) memcpy(ptbox, lf->ptbox, 3 * sizeof(int));
    if (h) memcpy(h, t->h, 3 * sizeof(double));
    return 0;
}

int ro_leaf_box(const ro_tree *t, int leaf, double *box) {
    if (leaf < 0 || leaf >= t->n_leaves) return 1;
    const ro_node *lf = t->leaves[leaf];
    for (int j = 0; j < 3; ++j) { box[2 * j] = lf->box[0][j]; box[2 * j + 1] = lf->box[1][j]; }
    return 0;
}

int ro_leaf_points(const ro_tree *t, int leaf, double *xyz, int *face) {
    if (leaf < 0 || leaf >= t->n_leaves) return 1;
    const ro_node *lf = t->leaves[leaf];
    int nbp = 0;
    for (int f = 1; f <= 2 * t->d; ++f) nbp += lf->pf[f];
    for (int p = 1; p <= lf->m; ++p) {
        int g[3];
        lin2fgrid(p, g, lf->ptbox, t->d);
        grid2real(xyz + 3 * (p - 1), g, lf->ptbox, t->d, t->h, lf->box);
        if (face) {
            int fc = 0;
            if (p <= nbp) { int off = 0; fc = 1; while (p > off + lf->pf[fc]) { off += lf->pf[fc]; ++fc; } }
            face[p - 1] = fc;
        }
    }
    return 0;
}

/* ============================================================================
 * Leaf stage (fd_module.f90:49-278; robin_tree_module.f90:667-799)
 * ========================================================================== */

/* FD_matrix + FD_matrix_row: dense m x m, column-major. */
static void fd_matrix(const ro_tree *t, const ro_node *lf, const cplx *k_int, cplx *A) {
    int d = t->d, m = lf->m;
    const int *shp = lf->ptbox;
    const double *h = t->h;
    cplx alpha = I * t->kb;
    int nbp = m - prod(shp, d);
    memset(A, 0, sizeof(cplx) * (size_t)m * m);
    for (int p = 1; p <= m; ++p) {
        int g[3], fn[3], dist;
        lin2fgrid(p, g, shp, d);
        grid_classify(&dist, fn, g, shp, d);
        size_t row = (size_t)(p - 1);
        if (dist > 0) {
            A[row + row * m] += k_int[p - nbp - 1];
            for (int j = 1; j <= d; ++j) {
                double w[3];
                if (fn[j - 1] == 2 * j - 1) { w[0] = 8.0 / 3.0; w[1] = -4.0; w[2] = 4.0 / 3.0; }
                else if (fn[j - 1] == 2 * j) { w[0] = 4.0 / 3.0; w[1] = -4.0; w[2] = 8.0 / 3.0; }
                else { w[0] = 1.0; w[1] = -2.0; w[2] = 1.0; }
                double h2 = h[j - 1] * h[j - 1];
                A[row + row * m] += w[1] / h2;
                int s[3] = {g[0], g[1], g[2]};
                s[j - 1] += 1;
                A[row + (size_t)(fgrid2lin(s, shp, d) - 1) * m] += w[2] / h2;
                s[j - 1] -= 2;
                A[row + (size_t)(fgrid2lin(s, shp, d) - 1) * m] += w[0] / h2;
            }
        } else {
            int s1[3] = {g[0], g[1], g[2]}, s2[3] = {g[0], g[1], g[2]}, n = 0;
            for (int j = 0; j < d; ++j) {
                if (g[j] == 0) { n = j; s1[j] += 1; s2[j] += 2; }
                else if (g[j] == shp[j] + 1) { n = j; s1[j] -= 1; s2[j] -= 2; }
            }
            A[row + row * m] += alpha + (8.0 / 3.0) / h[n];
            A[row + (size_t)(fgrid2lin(s1, shp, d) - 1) * m] -= 3.0 / h[n];
            A[row + (size_t)(fgrid2lin(s2, shp, d) - 1) * m] += (1.0 / 3.0) / h[n];
        }
    }
}

int ro_leaf_matrix(const ro_tree *t, int leaf, const double *k_int_leaf, double *A) {
    if (leaf < 0 || leaf >= t->n_leaves) return 1;
    fd_matrix(t, t->leaves[leaf], (const cplx *)k_int_leaf, (cplx *)A);
    return 0;
}

/* Outgoing-data stencil on boundary row r of column vector x (RTR_update,
 * robin_tree_module.f90:667-718; fd_outgoing_bc, fd_module.f90:449-513):
 *   x_r <- sigma*x_r + tau*(8/3 x_r - 3 x_{r+s} + 1/3 x_{r+2s})/h_n. */
static cplx outgoing(const ro_tree *t, const ro_node *lf, int face, int j, const cplx *x, cplx sigma) {
    int d = t->d;
    int lin = j;
    for (int f = 1; f < face; ++f) lin += lf->pf[f];
    int g[3];
    lin2fgrid(lin, g, lf->ptbox, d);
    int n = (face + 1) / 2 - 1;
    int sh = -1 + 2 * (face % 2);
    int g1[3] = {g[0], g[1], g[2]}, g2[3] = {g[0], g[1], g[2]};
    g1[n] += sh; g2[n] += 2 * sh;
    double hn = t->h[n];
    cplx v = sigma * x[lin - 1] + (8.0 / 3.0) / hn * x[lin - 1];
    v += (-3.0) / hn * x[fgrid2lin(g1, lf->ptbox, d) - 1];
    v += (1.0 / 3.0) / hn * x[fgrid2lin(g2, lf->ptbox, d) - 1];
    return v;
}

/* RTR_solve (robin_tree_module.f90:728-799) with FD_solve (fd_module.f90:234-278). */
static int leaf_rtr(ro_tree *t, ro_node *lf, const cplx *k_int) {
    int d = t->d, m = lf->m;
    int nbp = 0;
    for (int f = 1; f <= 2 * d; ++f) nbp += lf->pf[f];
    lf->A_lu = xcalloc((size_t)m * m, sizeof(cplx));
    lf->A_piv = xcalloc(m, sizeof(int));
    fd_matrix(t, lf, k_int, lf->A_lu);
    if (lu_factor(m, lf->A_lu, lf->A_piv)) return 1;
    cplx *X = xcalloc((size_t)m * nbp, sizeof(cplx));
    for (int j = 0; j < nbp; ++j) X[j + (size_t)j * m] = 1.0;
    lu_solve(m, lf->A_lu, lf->A_piv, nbp, X, m);
    cplx sigma = -I * t->kb;
    cplx *col_new = xcalloc(nbp, sizeof(cplx));
    for (int c = 0; c < nbp; ++c) {
        cplx *x = X + (size_t)c * m;
        /* the reference updates in place; rows read (r+s, r+2s) are interior
         * rows that the update never writes, so the order does not matter */
        int q = 0;
        for (int f = 1; f <= 2 * d; ++f)
            for (int j = 1; j <= lf->pf[f]; ++j) col_new[q++] = outgoing(t, lf, f, j, x, sigma);
        memcpy(x, col_new, sizeof(cplx) * nbp);
    }
    free(col_new);
    int rs = 0;
    for (int r = 1; r <= 2 * d; ++r) {
        int cs = 0;
        for (int c = 1; c <= 2 * d; ++c) {
            cplx *B = xcalloc((size_t)lf->pf[r] * lf->pf[c], sizeof(cplx));
            for (int jc = 0; jc < lf->pf[c]; ++jc)
                for (int jr = 0; jr < lf->pf[r]; ++jr)
                    B[jr + (size_t)jc * lf->pf[r]] = X[(rs + jr) + (size_t)(cs + jc) * m];
            lf->rtr[r][c] = B;
            cs += lf->pf[c];
        }
        rs += lf->pf[r];
    }
    free(X);
    return 0;
}

static void snapshot_rtr(const ro_tree *t, ro_node *nd) {
    int d = t->d, P = 0;
    for (int f = 1; f <= 2 * d; ++f) P += nd->pf[f];
    nd->P = P;
    nd->rtr_full = xcalloc((size_t)P * P, sizeof(cplx));
    int rs = 0;
    for (int r = 1; r <= 2 * d; ++r) {
        int cs = 0;
        for (int c = 1; c <= 2 * d; ++c) {
            for (int jc = 0; jc < nd->pf[c]; ++jc)
                for (int jr = 0; jr < nd->pf[r]; ++jr)
                    nd->rtr_full[(rs + jr) + (size_t)(cs + jc) * P] =
                        nd->rtr[r][c][jr + (size_t)jc * nd->pf[r]];
            cs += nd->pf[c];
        }
        rs += nd->pf[r];
    }
}

/* ============================================================================
 * Merge (compute_RTR internal branch, robin_tree_module.f90:396-645)
 * ========================================================================== */

static int merge_node(ro_tree *t, ro_node *nd) {
    int d = t->d;
    ro_node *c1 = nd->c1, *c2 = nd->c2;
    int i1 = c1->iface, i2 = c2->iface;
    int isz = c1->pf[i1];
    if (c1->pf[i1] != c2->pf[i2]) {
        set_err(t, "Error: Interface sizes of node children do not match");
        return 1;
    }
    nd->isz = isz;
    cplx a11 = t->bp[0][0], a12 = t->bp[0][1], a21 = t->bp[1][0], a22 = t->bp[1][1];
    cplx nu = (-a11 * a22 + a12 * a21) / (2.0 * a11 * a12);
    cplx zeta = (a11 * a22 + a12 * a21) / (2.0 * a11 * a12);
    nd->nu = nu;
    size_t ii = (size_t)isz * isz;
    /* D^-1 = zeta I - T2_II; LU; explicit D overwrites T2_II (lines 440-455). */
    cplx *T2 = c2->rtr[i2][i2];
    for (size_t q = 0; q < ii; ++q) T2[q] = -T2[q];
    for (int j = 0; j < isz; ++j) T2[j + (size_t)j * isz] += zeta;
    nd->D_lu = xcalloc(ii, sizeof(cplx));
    nd->D_piv = xcalloc(isz, sizeof(int));
    memcpy(nd->D_lu, T2, ii * sizeof(cplx));
    if (lu_factor(isz, nd->D_lu, nd->D_piv)) { set_err(t, "Error: singular D block in merge"); return 1; }
    for (size_t q = 0; q < ii; ++q) T2[q] = 0.0;
    for (int j = 0; j < isz; ++j) T2[j + (size_t)j * isz] = 1.0;
    lu_solve(isz, nd->D_lu, nd->D_piv, isz, T2, isz);
    /* S^-1 = zeta I - T1_II - nu^2 D; LU (lines 457-467). */
    cplx *T1 = c1->rtr[i1][i1];
    for (size_t q = 0; q < ii; ++q) T1[q] = -T1[q] - (nu * nu) * T2[q];
    for (int j = 0; j < isz; ++j) T1[j + (size_t)j * isz] += zeta;
    nd->S_lu = xcalloc(ii, sizeof(cplx));
    nd->S_piv = xcalloc(isz, sizeof(int));
    memcpy(nd->S_lu, T1, ii * sizeof(cplx));
    if (lu_factor(isz, nd->S_lu, nd->S_piv)) { set_err(t, "Error: singular S block in merge"); return 1; }
    free(c1->rtr[i1][i1]); c1->rtr[i1][i1] = NULL;
    free(c2->rtr[i2][i2]); c2->rtr[i2][i2] = NULL;
    /* Blocks of the parent RtR (lines 478-615). */
    for (int col = 1; col <= 2 * d; ++col) {
        cplx *C11 = NULL, *C21 = NULL, *C12 = NULL, *C22 = NULL;
        int cs1 = 0, cs2 = 0;
        if (col != i1) {
            cs1 = c1->pf[col];
            C11 = xcalloc((size_t)isz * cs1, sizeof(cplx));
            C21 = xcalloc((size_t)isz * cs1, sizeof(cplx));
            invvec(isz, nd->S_lu, nd->S_piv, cs1, c1->rtr[i1][col], 1.0, C11, 0.0);
            invvec(isz, nd->D_lu, nd->D_piv, cs1, C11, nu, C21, 0.0);
        }
        if (col != i2) {
            cs2 = c1->pf[col];   /* size(child1%RtR(child2%iface,col),2), line 500 */
            C22 = xcalloc((size_t)isz * cs2, sizeof(cplx));
            C12 = xcalloc((size_t)isz * cs2, sizeof(cplx));
            invvec(isz, nd->D_lu, nd->D_piv, cs2, c2->rtr[i2][col], 1.0, C22, 0.0);
            invvec(isz, nd->S_lu, nd->S_piv, cs2, C22, nu, C12, 0.0);
            invvec(isz, nd->D_lu, nd->D_piv, cs2, C12, nu, C22, 1.0);
        }
        for (int row = 1; row <= 2 * d; ++row) {
            int pr = nd->pf[row], pc = nd->pf[col];
            cplx *B = xcalloc((size_t)pr * pc, sizeof(cplx));
            nd->rtr[row][col] = B;
            if (C11 && row != i1) {
                int rsz = c1->pf[row];
                gemm(rsz, cs1, isz, 1.0, c1->rtr[row][i1], rsz, C11, isz, 0.0, B, pr);
                const cplx *T = c1->rtr[row][col];
                for (int jc = 0; jc < cs1; ++jc)
                    for (int jr = 0; jr < rsz; ++jr) B[jr + (size_t)jc * pr] += T[jr + (size_t)jc * rsz];
            }
            if (C12 && row != i1) {
                int rsz = c1->pf[row];
                gemm(rsz, cs2, isz, -1.0, c1->rtr[row][i1], rsz, C12, isz, 0.0,
                     B + (size_t)(pc - cs2) * pr, pr);
            }
            if (C21 && row != i2) {
                int rsz = c2->pf[row];
                gemm(rsz, cs1, isz, -1.0, c2->rtr[row][i2], rsz, C21, isz, 0.0, B + (pr - rsz), pr);
            }
            if (C22 && row != i2) {
                int rsz = c2->pf[row];
                cplx *Bq = B + (pr - rsz) + (size_t)(pc - cs2) * pr;
                gemm(rsz, cs2, isz, 1.0, c2->rtr[row][i2], rsz, C22, isz, 0.0, Bq, pr);
                const cplx *T = c2->rtr[row][col];
                for (int jc = 0; jc < cs2; ++jc)
                    for (int jr = 0; jr < rsz; ++jr) Bq[jr + (size_t)jc * pr] += T[jr + (size_t)jc * rsz];
            }
        }
        free(C11); free(C21); free(C12); free(C22);
    }
    /* Free the children's T_NN (lines 618-645; here on every face, which the
     * reference leaks for faces d+1..2d).  T_NI and T_IN stay for the solve. */
    for (int r = 1; r <= 2 * d; ++r)
        for (int c = 1; c <= 2 * d; ++c) {
            if (r != i1 && c != i1) { free(c1->rtr[r][c]); c1->rtr[r][c] = NULL; }
            if (r != i2 && c != i2) { free(c2->rtr[r][c]); c2->rtr[r][c] = NULL; }
        }
    return 0;
}

/* one leaf (level L) or one merge (level < L) of ro_factor; 1 = singular leaf,
 * 2 = merge failure, 0 = ok */
static int factor_node(ro_tree *t, int l, int q, const cplx *k_int, int ni) {
    if (l == t->L) {
        ro_node *lf = t->leaves[q];
        if (leaf_rtr(t, lf, k_int + (size_t)q * ni)) return 1;
        snapshot_rtr(t, lf);
        return 0;
    }
    ro_node *nd = t->lvl_nodes[l][q];
    if (merge_node(t, nd)) return 2;
    snapshot_rtr(t, nd);
    return 0;
}

static int max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int ro_factor(ro_tree *t, const double *k_int_d) {
    omp_setup();
    const cplx *k_int = (const cplx *)k_int_d;
    int ni = prod(t->leaves[0]->ptbox, t->d);
    int bad = 0;
    /* Levels with at least as many nodes as threads run their nodes in parallel,
     * each node on one thread.  The nodes of a level with fewer nodes (the top
     * of the tree) run one after another, each with its dense operations spread
     * over all threads (ro_dense.c).  Either way a node's arithmetic is the
     * same. */
    const int nth = max_threads();
    for (int l = t->L; l >= 0; --l) {
        const int cnt = l == t->L ? t->n_leaves : t->lvl_count[l];
        const int outer = cnt >= nth;
        ro_dense_set_inner(outer ? 1 : nth);
        /* (the sequential loop is not an inactive parallel region: teams opened
         * inside one would be nested teams, which the OpenMP runtime does not
         * take from its thread pool) */
        if (outer) {
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
            for (int q = 0; q < cnt; ++q) bad |= factor_node(t, l, q, k_int, ni);
        } else {
            for (int q = 0; q < cnt; ++q) bad |= factor_node(t, l, q, k_int, ni);
        }
        if (bad) break;
    }
    ro_dense_set_inner(0);
    if (bad & 1) { set_err(t, "Error: singular leaf matrix"); return 1; }
    if (bad) return 1;
    t->factored = 1;
    return 0;
}

/* ============================================================================
 * Solve (robin_tree_module.f90:198-233, 813-1322; fd_module.f90:390-659)
 * ========================================================================== */

/* apply_invm (robin_tree_module.f90:1252-1322): y = alpha*y + inv(M) x. */
static void apply_invm(const ro_node *nd, cplx *y1, cplx *y2, const cplx *x1, const cplx *x2,
                       cplx alpha) {
    int n = nd->isz;
    cplx nu = nd->nu;
    cplx *interm = xcalloc(n, sizeof(cplx));
    cplx *xt = xcalloc(n, sizeof(cplx));
    memcpy(xt, x1, n * sizeof(cplx));
    invvec(n, nd->S_lu, nd->S_piv, 1, xt, 1.0, interm, 0.0);                 /* step 1 */
    for (int j = 0; j < n; ++j) y1[j] = (alpha == 0.0 ? 0.0 : alpha * y1[j]) + interm[j]; /* 2 */
    invvec(n, nd->D_lu, nd->D_piv, 1, interm, -nu, y2, alpha);                /* step 3 */
    memcpy(xt, x2, n * sizeof(cplx));
    invvec(n, nd->D_lu, nd->D_piv, 1, xt, 1.0, interm, 0.0);                 /* step 4 */
    for (int j = 0; j < n; ++j) y2[j] += interm[j];                           /* step 5 */
    invvec(n, nd->S_lu, nd->S_piv, 1, interm, 1.0, NULL, 0.0);               /* step 6 */
    for (int j = 0; j < n; ++j) y1[j] -= nu * interm[j];                      /* step 7 */
    invvec(n, nd->D_lu, nd->D_piv, 1, interm, nu * nu, y2, 1.0);              /* step 8 */
    free(interm); free(xt);
}

static void alloc_vecs(const ro_tree *t, ro_node *nd, cplx **v) {
    for (int f = 1; f <= 2 * t->d; ++f) if (!v[f]) v[f] = xcalloc(nd->pf[f], sizeof(cplx));
}

/* fd_leaf_solve + fd_outgoing_bc (fd_module.f90:390-513). */
static void leaf_up(const ro_tree *t, ro_node *lf, const cplx *f_leaf) {
    int m = lf->m, nbp = m - prod(lf->ptbox, t->d);
    lf->sol = xcalloc(m, sizeof(cplx));
    memcpy(lf->sol + nbp, f_leaf, sizeof(cplx) * (m - nbp));
    lu_solve(m, lf->A_lu, lf->A_piv, 1, lf->sol, m);
    cplx sigma = -I * t->kb;
    for (int f = 1; f <= 2 * t->d; ++f)
        for (int j = 1; j <= lf->pf[f]; ++j) lf->h[f][j - 1] = outgoing(t, lf, f, j, lf->sol, sigma);
}

/* forward_sub internal-node branch (robin_tree_module.f90:868-969). */
static void node_up(const ro_tree *t, ro_node *nd) {
    ro_node *c1 = nd->c1, *c2 = nd->c2;
    int i1 = c1->iface, i2 = c2->iface;
    apply_invm(nd, c1->g[i1], c2->g[i2], c1->h[i1], c2->h[i2], 0.0);
    if (nd->lvl > 0) {
        int ax = (i1 - 1) / 2;
        for (int j = 1; j <= 2 * t->d; ++j) {
            if ((j - 1) / 2 == ax) {
                ro_node *src = (j == i1) ? c2 : c1;
                int si = src->iface;
                memcpy(nd->h[j], src->h[j], sizeof(cplx) * nd->pf[j]);
                gemv_acc(src->pf[j], src->pf[si], src->rtr[j][si], src->g[si], nd->h[j]);
            } else {
                int half = nd->pf[j] / 2;
                memcpy(nd->h[j], c1->h[j], sizeof(cplx) * half);
                memcpy(nd->h[j] + half, c2->h[j], sizeof(cplx) * half);
                gemv_acc(c1->pf[j], c1->pf[i1], c1->rtr[j][i1], c1->g[i1], nd->h[j]);
                gemv_acc(c2->pf[j], c2->pf[i2], c2->rtr[j][i2], c2->g[i2], nd->h[j] + half);
            }
        }
    }
}

/* apply_bc (robin_tree_module.f90:1186-1237) + fd_make_bndry_vec
 * (fd_module.f90:611-659): physical-face g at every level. */
static void apply_bc(const ro_tree *t, ro_node *nd, const cplx *g_in) {
    int d = t->d;
    if (nd->isleaf) {
        int q = 0;
        while (t->leaves[q] != nd) ++q;
        const cplx *gl = g_in + (size_t)q * (nd->m - prod(nd->ptbox, d));
        int off = 0;
        for (int f = 1; f <= 2 * d; ++f) {
            if (nd->isbndry[f]) memcpy(nd->g[f], gl + off, sizeof(cplx) * nd->pf[f]);
            off += nd->pf[f];
        }
        return;
    }
    int any = 0;
    for (int f = 1; f <= 2 * d; ++f) any |= nd->isbndry[f];
    if (!any) return;
    apply_bc(t, nd->c1, g_in);
    apply_bc(t, nd->c2, g_in);
    for (int f = 1; f <= 2 * d; ++f) {
        if (!nd->isbndry[f]) continue;
        int mm = nd->pf[f];
        if (f == nd->c2->iface) memcpy(nd->g[f], nd->c1->g[f], sizeof(cplx) * mm);
        else if (f == nd->c1->iface) memcpy(nd->g[f], nd->c2->g[f], sizeof(cplx) * mm);
        else {
            memcpy(nd->g[f], nd->c1->g[f], sizeof(cplx) * (mm / 2));
            memcpy(nd->g[f] + mm / 2, nd->c2->g[f], sizeof(cplx) * (mm / 2));
        }
    }
}

/* backward_sub internal-node branch (robin_tree_module.f90:1039-1139). */
static void node_down(const ro_tree *t, ro_node *nd) {
    int d = t->d;
    ro_node *c1 = nd->c1, *c2 = nd->c2;
    int i1 = c1->iface, i2 = c2->iface, ax = (i1 - 1) / 2;
    for (int f = 1; f <= 2 * d; ++f) {
        if (nd->isbndry[f]) continue;
        if ((f - 1) / 2 != ax) {
            int half = nd->pf[f] / 2;
            memcpy(c1->g[f], nd->g[f], sizeof(cplx) * half);
            memcpy(c2->g[f], nd->g[f] + half, sizeof(cplx) * half);
        } else if (f == i1) {
            memcpy(c2->g[f], nd->g[f], sizeof(cplx) * nd->pf[f]);
        } else if (f == i2) {
            memcpy(c1->g[f], nd->g[f], sizeof(cplx) * nd->pf[f]);
        }
    }
    int n = nd->isz;
    cplx *Tg1 = xcalloc(n, sizeof(cplx)), *Tg2 = xcalloc(n, sizeof(cplx));
    for (int f = 1; f <= 2 * d; ++f) {
        if (f != i1) gemv_acc(c1->pf[i1], c1->pf[f], c1->rtr[i1][f], c1->g[f], Tg1);
        if (f != i2) gemv_acc(c2->pf[i2], c2->pf[f], c2->rtr[i2][f], c2->g[f], Tg2);
    }
    apply_invm(nd, c1->g[i1], c2->g[i2], Tg1, Tg2, 1.0);
    free(Tg1); free(Tg2);
}

/* fd_leaf_update (fd_module.f90:525-566). */
static void leaf_down(const ro_tree *t, ro_node *lf) {
    int m = lf->m;
    cplx *b = xcalloc(m, sizeof(cplx));
    int off = 0;
    for (int f = 1; f <= 2 * t->d; ++f) { memcpy(b + off, lf->g[f], sizeof(cplx) * lf->pf[f]); off += lf->pf[f]; }
    lu_solve(m, lf->A_lu, lf->A_piv, 1, b, m);
    for (int j = 0; j < m; ++j) lf->sol[j] += b[j];
    free(b);
}

int ro_solve(ro_tree *t, const double *f_d, const double *g_d, double *u_d) {
    if (!t->factored) { set_err(t, "Error: Robin Tree must be factored first."); return 1; }
    omp_setup();
    const cplx *f = (const cplx *)f_d, *g = (const cplx *)g_d;
    int d = t->d, ni = prod(t->leaves[0]->ptbox, d);
    for (int q = 0; q < t->n_nodes; ++q) {
        ro_node *nd = t->nodes[q];
        free_node_solve(nd);
        if (nd->lvl > 0 || nd->isleaf) alloc_vecs(t, nd, nd->h);
        alloc_vecs(t, nd, nd->g);
    }
#pragma omp parallel for schedule(dynamic)
    for (int q = 0; q < t->n_leaves; ++q) leaf_up(t, t->leaves[q], f + (size_t)q * ni);
    for (int l = t->L - 1; l >= 0; --l) {
#pragma omp parallel for schedule(dynamic)
        for (int q = 0; q < t->lvl_count[l]; ++q) node_up(t, t->lvl_nodes[l][q]);
    }
    apply_bc(t, t->root, g);
    for (int l = 0; l < t->L; ++l) {
#pragma omp parallel for schedule(dynamic)
        for (int q = 0; q < t->lvl_count[l]; ++q) node_down(t, t->lvl_nodes[l][q]);
    }
    cplx *u = (cplx *)u_d;
    int m = t->leaves[0]->m;
#pragma omp parallel for schedule(dynamic)
    for (int q = 0; q < t->n_leaves; ++q) {
        leaf_down(t, t->leaves[q]);
        memcpy(u + (size_t)q * m, t->leaves[q]->sol, sizeof(cplx) * m);
    }
    return 0;
}

/* ============================================================================
 * Test hooks
 * ========================================================================== */

int ro_num_nodes(const ro_tree *t) { return t->n_nodes; }

int ro_node_rtr(const ro_tree *t, int id, double *out, int *P) {
    if (id < 0 || id >= t->n_nodes) return 1;
    const ro_node *nd = t->nodes[id];
    if (!nd->rtr_full) return 2;
    if (P) *P = nd->P;
    if (out) memcpy(out, nd->rtr_full, sizeof(cplx) * (size_t)nd->P * nd->P);
    return 0;
}

int ro_apply_invm(const ro_tree *t, int id, const double *x, double *y) {
    if (id < 0 || id >= t->n_nodes) return 1;
    const ro_node *nd = t->nodes[id];
    if (nd->isleaf || !nd->D_lu) return 2;
    int n = nd->isz;
    apply_invm(nd, (cplx *)y, (cplx *)y + n, (const cplx *)x, (const cplx *)x + n, 0.0);
    return 0;
}

int ro_node_ptbox(const ro_tree *t, int id, int *ptbox) {
    if (id < 0 || id >= t->n_nodes) return 1;
    memcpy(ptbox, t->nodes[id]->ptbox, 3 * sizeof(int));
    return 0;
}

int ro_level_node_id(const ro_tree *t, int lvl, int idx) {
    if (lvl < 0 || lvl > t->L || idx < 0 || idx >= t->lvl_count[lvl]) return -1;
    return t->lvl_nodes[lvl][idx]->id;
}

int ro_node_iface(const ro_tree *t, int id, int *isz, int *c1i, int *c2i) {
    if (id < 0 || id >= t->n_nodes) return 1;
    const ro_node *nd = t->nodes[id];
    if (nd->isleaf) return 2;
    if (isz) *isz = nd->c1->pf[nd->c1->iface];
    if (c1i) *c1i = nd->c1->iface;
    if (c2i) *c2i = nd->c2->iface;
    return 0;
}
