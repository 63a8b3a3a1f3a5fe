/*
 * rs_oracle.h -- CPU restatement of robinsolver's HPS factor and solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X
 * implementation in robinsolver_amd/: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product path never
 * calls it.
 *
 * It restates, in plain C99 complex double, the reference algorithm:
 *   geometry            robin_tree_module.f90:288-363, 1369-1468; fd_module.f90:1033-1084
 *   leaf FD matrix      fd_module.f90:49-224 (index maps 669-1022)
 *   leaf LU + RtR       fd_module.f90:234-278; robin_tree_module.f90:667-799
 *   merge               robin_tree_module.f90:375-650
 *   solve passes        robin_tree_module.f90:813-1146, 1186-1322; fd_module.f90:390-566
 *   matrixdata          factorization_module.f90:24-141, 291-366
 *   output order        robin_output_module.f90:48-75, 282-322
 * Dense linear algebra (zgetrf/zgetrs/zgetri/zgemm/zgemv, which the reference
 * takes from an unpinned external LAPACK/BLAS) is written out here as plain
 * loops; zgetri is restated as zgetrs on the identity (equal to rounding).
 *
 * Complex arrays cross this interface as interleaved (re, im) doubles.
 */
#ifndef RS_ORACLE_H
#define RS_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ro_tree ro_tree;

/* Build the tree geometry (bin_tree + per-leaf get_h).  domain is the
 * reference's elliptic_operator%domain(1:2,1:3) in column-major order:
 * {x_lo, x_hi, y_lo, y_hi, z_lo, z_hi}.  bndry_params is robin_tree%bndry_params
 * (2,2) complex, column-major, interleaved.  Returns NULL on error (message via
 * ro_create_error()). */
ro_tree *ro_create(int d, int L, const double *domain, double h_tgt, double kb,
                   const double *bndry_params);
const char *ro_create_error(void);
void ro_destroy(ro_tree *t);
const char *ro_last_error(const ro_tree *t);
void ro_set_threads(int n);

/* sizes: n_leaves, m (unknowns per leaf), n_int, nb (boundary points per leaf),
 * ptbox[3], h[3]. */
int ro_info(const ro_tree *t, int *n_leaves, int *m, int *n_int, int *nb, int *ptbox,
            double *h);
/* Leaf box (2,3) column-major for leaf j in get_solution (DFS) order. */
int ro_leaf_box(const ro_tree *t, int leaf, double *box);
/* Coordinates (x,y,z) of every unknown of a leaf in fgrid order (boundary
 * points first, face by face, then interior x-fastest), grid2real
 * (fd_module.f90:669-700).  xyz has 3*m doubles; face[m] gets the face id 1..6
 * of boundary points and 0 for interior points (may be NULL). */
int ro_leaf_points(const ro_tree *t, int leaf, double *xyz, int *face);

/* Factor.  k_int: [n_leaves][n_int] complex, the potential sampled at the
 * interior points of each leaf in interior order. */
int ro_factor(ro_tree *t, const double *k_int);

/* Solve one RHS.  f: [n_leaves][n_int] complex (interior samples of f);
 * g: [n_leaves][nb] complex (samples of g_face at the leaf boundary points in
 * fgrid boundary order; only points on the physical boundary are read).
 * u: [n_leaves][m] complex output, the leaf solution vectors. */
int ro_solve(ro_tree *t, const double *f, const double *g, double *u);

/* Test hooks. */
/* Number of nodes (2^(L+1)-1); node n in pre-order id. */
int ro_num_nodes(const ro_tree *t);
/* Full RtR of a node in face order, P x P column-major, as computed at the
 * node (snapshot before any parent merge touches it).  Returns P via *P; pass
 * out=NULL to query P. */
int ro_node_rtr(const ro_tree *t, int node_id, double *out, int *P);
/* y = inv(M) x at internal node node_id via apply_invm with alpha = 0
 * (robin_tree_module.f90:1252-1322); x, y have 2*i complex entries. */
int ro_apply_invm(const ro_tree *t, int node_id, const double *x, double *y);
/* ptbox of a node (leaf grid, or doubled child grid for internal nodes). */
int ro_node_ptbox(const ro_tree *t, int node_id, int *ptbox);
/* Pre-order id of the idx-th node (left to right) of level lvl, -1 if none. */
int ro_level_node_id(const ro_tree *t, int lvl, int idx);
/* Interface size i of an internal node and the face ids of its children's
 * interfaces. */
int ro_node_iface(const ro_tree *t, int node_id, int *isz, int *c1_iface, int *c2_iface);
/* Dense leaf FD matrix (m x m, column-major) for leaf j, as FD_matrix builds it. */
int ro_leaf_matrix(const ro_tree *t, int leaf, const double *k_int_leaf, double *A);

#ifdef __cplusplus
}
#endif
#endif
