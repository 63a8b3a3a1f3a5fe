/*
 * robinsolver.h -- C ABI of the MI355X (gfx950) HPS factor and solve.
 *
 * Drop-in boundary for robinsolver's Fortran API.  Every entry point below is
 * what the Fortran shim (robinsolver_amd/fortran/robinsolver_f.f90) binds with
 * ISO_C_BINDING to keep the reference's module and procedure names; the
 * reference interface each one replaces is cited.  Plain pointers and sizes
 * only.  Complex arrays are interleaved (re, im) doubles, the memory layout of
 * Fortran complex(dp).  Every call returns 0 on success or a nonzero status;
 * rs_last_error() then holds the message (the reference's `stop` text where
 * the reference has one).
 *
 * Layouts:
 *   leaf order        DFS order of get_solution (robin_output_module.f90:282-322)
 *   point order       fgrid order of a leaf: boundary points face by face
 *                     (x1,x2,y1,y2,z1,z2), then interior points x-fastest
 *                     (fd_module.f90:867-939)
 *   k_int, f          [n_leaves][n_int]     samples at interior points
 *   g                 [n_leaves][n_bnd]     samples at leaf boundary points;
 *                                           only physical-boundary points are read
 *   u                 [n_leaves][m]         leaf solution vectors, m = n_bnd + n_int
 *   multi-RHS         an outer [nrhs] dimension in front of each of the above
 */
#ifndef ROBINSOLVER_H
#define ROBINSOLVER_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct rs_tree rs_tree;

enum {
    RS_FP64 = 0,   /* complex128 everywhere: the plain FP64 GPU path          */
    RS_MIXED = 1,  /* complex64 merge products and solve passes on the matrix
                      cores (FP32 accumulation), FP64 leaf factorization and
                      pivoted interface inverses, FP64 residual + iterative
                      refinement                                             */
    RS_MIXED_MERGE = 2  /* rs_check_gemm only: complex64 products of complex128
                           operands, complex128 result (the mixed merge)     */
};

typedef struct rs_opts {
    int precision;        /* RS_FP64 or RS_MIXED: the factorization's merge
                             products and the default solve precision.  A tree
                             factored in FP64 also takes RS_MIXED solves (FP64
                             factor data rounded to complex64 for the passes);
                             a tree with complex64 merges refuses RS_FP64
                             solves, whose accuracy it cannot give            */
    int skip_root_rtr;    /* 1: do not form the root RtR, which no solve reads  */
    int n_refine;         /* refinement steps of the mixed path                 */
    int verbose;          /* 1: progress lines as the reference prints them     */
    long long hbm_budget; /* bytes this tree may hold on the device (0 = no cap) */
    int autotune;         /* 1: time the matrix-core tile shapes of each solve
                             product once per workspace capacity (and of the merge
                             products on the first factorization) and keep the
                             fastest; the results do not depend on the choice.
                             A server tunes a tree when it is registered or after
                             its first factorization, not while dispatching     */
    int graphs;           /* 1: rs_solve / rs_solve_device capture the launch
                             sequence of each solve shape (precision, nrhs) once
                             as a hipGraph and replay it                          */
    int merge_fp64_levels;/* RS_MIXED trees: k >= 0 keeps complex128 merge
                             products at the top k levels (root first) and
                             runs the levels below in complex64; -1 (or k >= L)
                             keeps the whole factorization FP64               */
    int retain;           /* what a factored tree keeps beyond its solve data
                             (RS_RETAIN_* flags, 0 = the solve data only)     */
} rs_opts;

/* rs_opts.retain.  The solve data a tree keeps after rs_factor, as the
 * reference keeps D/S, nu, the children's T_NI and T_IN and the leaf factors
 * and frees T_II and T_NN (robin_tree_module.f90:472-475, 617-645): the leaf
 * block inverses, inv(M) of every merge and the children's T_NI / T_IN, in the
 * tree's precision (complex64 for RS_MIXED, complex128 for RS_FP64). */
enum {
    RS_RETAIN_MAPS = 1,     /* the full RtR map of every node (rs_node_rtr)            */
    RS_RETAIN_FP64 = 2,     /* RS_MIXED: also the complex128 solve data (FP64 solves) */
    RS_RETAIN_SCRATCH = 4   /* the factorization's working memory, so that a refactor
                               of the same plan allocates nothing                    */
};

/* Defaults: precision RS_FP64, skip_root_rtr 1, n_refine 2, verbose 0, no cap,
 * autotune 1, graphs 1, merge_fp64_levels 0 (an RS_MIXED tree runs every merge
 * product in complex64), retain 0. */
void rs_opts_default(rs_opts *o);

/* Create / destroy a tree handle (type robin_tree, robin_tree_module.f90:55-77;
 * destroy has no reference counterpart: the reference never frees a tree). */
int rs_tree_create(rs_tree **t, const rs_opts *o);
int rs_tree_destroy(rs_tree *t);

/* Geometry of robin_tree_factor (robin_tree_module.f90:96-156: validation,
 * bin_tree, node_divide, get_h).  domain = elliptic_operator%domain(1:2,1:3)
 * column-major {x_lo,x_hi,y_lo,y_hi,z_lo,z_hi}; kb = solve_opts%kb;
 * bndry_params = robin_tree%bndry_params(2,2) complex column-major.  Outputs the
 * sizes the caller needs to sample k, f and g. */
int rs_plan(rs_tree *t, int d, int L, const double *domain, double h_tgt, double kb,
            const double *bndry_params, int *n_leaves, int *m, int *n_int, int *n_bnd);

/* Coordinates (x,y,z) of every unknown of a leaf in point order (grid2real,
 * fd_module.f90:669-700) and the face id (1..6, 0 = interior) of each point.
 * face may be NULL. */
int rs_leaf_points(const rs_tree *t, int leaf, double *xyz, int *face);

/* rs_leaf_points of every leaf at once, written to device memory on `stream`
 * (a hipStream_t; NULL: the tree's stream): xyz [n_leaves][m][3], face
 * [n_leaves][m] (may be NULL); the same bits as rs_leaf_points.  So that
 * callers sample k, f and g on the GPU instead of in host callbacks (the
 * reference evaluates them on the host: fd_module.f90:136-145, 427-431,
 * 645-651). */
int rs_points_device(rs_tree *t, double *xyz, int *face, void *stream);

/* rs_factor with k at the interior points already in device memory
 * ([n_leaves][n_int] complex128, DFS leaf order), synchronous.  k is read
 * after the work queued on the default stream before the call. */
int rs_factor_device(rs_tree *t, const double *k_int);
/* Leaf grid: points per axis (ptbox) and spacing h, the values the reference
 * stores in node%ptbox and T%opts%h (fd_module.f90:1060-1084,
 * robin_tree_module.f90:392).  h(3) = 0 and ptbox(3) = -1 in 2D. */
int rs_leaf_grid(const rs_tree *t, int *ptbox, double *h);
/* Leaf box (2,3) column-major (leaf_solution%box, robin_output_module.f90:32-37). */
int rs_leaf_box(const rs_tree *t, int leaf, double *box);

/* robin_tree%factor (robin_tree_module.f90:96-185) with the potential sampled
 * at the interior points: k_int [n_leaves][n_int] complex, host memory. */
int rs_factor(rs_tree *t, const double *k_int);

/* robin_tree%solve (robin_tree_module.f90:198-233) for nrhs right-hand sides
 * at once, host memory: f [nrhs][n_leaves][n_int], g [nrhs][n_leaves][n_bnd].
 * Repeatable (the reference allows one solve per process).  precision < 0
 * uses rs_opts.precision. */
int rs_solve(rs_tree *t, int nrhs, const double *f, const double *g, int precision);

/* get_solution (robin_output_module.f90:48-75) for RHS number rhs of the last
 * rs_solve: u [n_leaves][m] complex, host memory. */
int rs_get_solution(const rs_tree *t, int rhs, double *u);

/* Device-resident variant of rs_solve (no reference counterpart: the caller's
 * inputs already live in HBM).  f_dev, g_dev, u_dev are device pointers with
 * the multi-RHS layouts above; stream is a hipStream_t (NULL = the tree's own
 * stream).  Asynchronous with respect to the host.
 * Ordering: with stream = NULL the solve runs on the tree's own stream after
 * the work the caller queued on the default stream (e.g. torch's default
 * stream, hipMemset, hipMemcpy) before the call, and work queued on the default
 * stream after the call waits for the solve.  A caller producing f/g on another
 * stream passes that stream and orders its own reads of u.  Consecutive solves
 * on one handle may use different streams: each waits for the previous one
 * (they share one workspace).  Calls on one handle must not overlap in host
 * threads. */
int rs_solve_device(rs_tree *t, int nrhs, const void *f_dev, const void *g_dev, void *u_dev,
                    int precision, void *stream);

/* Allocation-free variant of rs_solve_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, ordering and results (bit for bit) as rs_solve_device, but
 * the pass's whole workspace -- level vectors, K-split partial products,
 * refinement buffers, column tables and product descriptors -- lives in one
 * fixed device array of RS_NOALLOC_SCRATCH_MIB MiB that every GPU the library
 * runs on holds until the process ends.  A solve needing more fails with a
 * message that states the bytes needed (fewer right-hand sides per call fit:
 * about 60 MB per RHS at 1024^2 in RS_MIXED).  The tree must be factored, and
 * for RS_MIXED hold its complex64 data (factored with RS_MIXED).  Every call on
 * one GPU uses that GPU's one array: calls that overlap on the device (other
 * streams, trees or threads) race; calls queued on one stream are ordered.  No
 * hipMalloc, hipFree, graph or profiling event is used. */
#define RS_NOALLOC_SCRATCH_MIB 256
int rs_solve_device_noalloc(rs_tree *t, int nrhs, const void *f_dev, const void *g_dev, void *u_dev,
                            int precision, void *stream);

/* Number of FP64 refinement steps of the mixed path for later solves
 * (rs_opts.n_refine; no reference counterpart). */
int rs_set_refine(rs_tree *t, int n_refine);

/* Factored-tree persistence (no reference counterpart: the reference keeps no
 * factorization beyond its process -- write_variable is a stub,
 * robin_output_module.f90:157-221, robin_tree_memsize is empty,
 * robin_tree_module.f90:243-250 -- so every run re-factors).  rs_tree_save writes
 * a factored tree's plan parameters (the arguments of rs_plan) and its retained
 * solve data (the potential, the leaf block inverses, inv(M) of every merge
 * level and the children's T_NI / T_IN, in each precision the tree holds) to a
 * file.  rs_tree_load reads such a file into a handle: it plans the handle from
 * the file (dropping what the handle held), takes the saved tree's precision
 * (rs_opts.precision; the other options stay the handle's), uploads the data and
 * leaves the tree factored, so that its solves give the saved tree's bits; its
 * f_time is the load's wall time.  A file that is not a saved tree or of another
 * format version is refused before the handle changes; one whose sections do
 * not match its plan, or that ends early, is refused with a message and leaves
 * the handle planned but not factored.  Full node maps (RS_RETAIN_MAPS) and the
 * factorization's working memory are not saved. */
int rs_tree_save(rs_tree *t, const char *path);
int rs_tree_load(rs_tree *t, const char *path);

/* T%f_time and T%s_time (robin_tree_module.f90:63-64, time_module.f90:26-36):
 * wall seconds of the last factor and solve. */
int rs_timings(const rs_tree *t, double *f_time, double *s_time);

/* Message of the last failed call on this handle (or of the last failed
 * rs_tree_create when t is NULL). */
const char *rs_last_error(const rs_tree *t);

/* Inspection (robin_tree%inspect, robin_tree_module.f90:263-274): factored flag,
 * number of levels, device bytes held. */
int rs_inspect(const rs_tree *t, int *factored, int *nlvls, long long *device_bytes);

/* FD residual check (no reference counterpart; the reference's check_solution
 * compares with a known solution, robin_output_module.f90:87-138): for each of
 * nrhs solutions u of f, g (device arrays, the layouts of rs_solve_device),
 * rel[r] = ||r||_2 / ||(f, g)||_2 of the global FD system, r = the interior rows
 * (Delta_h + k) u - f (fd_module.f90:140-181), the physical Robin rows
 * (fd_module.f90:187-216) and the transmission rows zeta g_l + nu g_nb - h_l
 * between neighbouring leaves, all evaluated in FP64.  rel is host memory. */
int rs_residual_norms(rs_tree *t, int nrhs, const void *f_dev, const void *g_dev, const void *u_dev, double *rel);

/* Device memory (no reference counterpart; the reference's memsize is empty,
 * robin_tree_module.f90:243-250).  rs_memory_estimate: host-side, from the plan:
 * bytes the factorization keeps resident, its peak while it runs, and one solve
 * workspace of ncols right-hand sides at the tree's precision.
 * rs_memory_used: bytes the tree holds now and the most it has held. */
int rs_memory_estimate(const rs_tree *t, int ncols, long long *factor_resident, long long *factor_peak,
                       long long *solve_workspace);
int rs_memory_used(const rs_tree *t, long long *held, long long *peak);

/* Autotuning result (no reference counterpart): the GEMM tile chosen for each
 * product of the solve workspace of rs_solve / rs_solve_device at `precision`,
 * tiles[4 * lvl + k] for levels 0..L-1 and products k = 0 up (interface),
 * 1 up (parent data), 2 down (children's data), 3 down (interface); values as in
 * rs_check_gemm_tile, -1 for the shape rule or a product the level does not have.
 * Fails if that workspace does not exist yet (before the first solve). */
int rs_tile_choices(const rs_tree *t, int precision, int *tiles);

/* Per-level shape of the plan (no reference counterpart; host-side, needs no
 * GPU): nodes of level lvl, boundary points P of each, and for lvl < L the
 * interface size and the children's boundary size of its merge (0 at the
 * leaf level). */
int rs_level_info(const rs_tree *t, int lvl, int *nodes, int *P, int *isz, int *child_P);

/* Test hooks (no reference counterpart).  Node RtR of node idx (left to right)
 * of level lvl, P x P column-major, in face order (RtR(6,6) blocks of
 * robin_tree_node assembled, derived_type_module.f90:239-269).  P via *P; out
 * may be NULL to query P. */
int rs_node_rtr(const rs_tree *t, int lvl, int idx, double *out, int *P);

/* Kernel checks (test hooks, no reference counterpart), host memory in and out:
 *   rs_check_gemm: C = alpha A B + beta C for one column-major complex problem
 *     (M x K times K x N) through the batched MFMA GEMM, precision RS_FP64
 *     (complex128), RS_MIXED (the complex64 GEMM of the solve passes: data
 *     rounded to complex64 on the host, complex64 result) or RS_MIXED_MERGE
 *     (complex64 products of complex128 data, complex128 result);
 *     rows k >= ksplit of B are taken from B2 (ksplit >= K: unused); rmap, when
 *     not NULL, scatters output row r to row rmap[r] of C (ldc rows).
 *   rs_check_inverse: in-place inverse of `batch` n x n complex128 matrices,
 *     the interface inversion of the merge. */
int rs_check_gemm(int precision, int M, int N, int K, const double *A, int lda, const double *B,
                  const double *B2, int ldb, int ksplit, double *C, int ldc, const int *rmap,
                  const double *alpha, const double *beta);
int rs_check_inverse(int n, int batch, double *mats);
/* rs_check_gemm with the output tile shape forced (0: 64x64, 1: 32x64, 2: 16x64,
 * 3: 64x16, 4: 32x32, 5: 64x32, 6: 128x16, 7: skinny 64x16, 8: matrix-core-free
 * GEMV, 9: 128x64, 10: 64x128, 11..15: 64x64, 128x64, 64x128, 64x32, 128x128
 * with the operands copied straight into LDS, -1: by shape); a test hook for the
 * autotuner's premise that the tile never changes the results. */
int rs_check_gemm_tile(int precision, int M, int N, int K, const double *A, int lda, const double *B,
                       const double *B2, int ldb, int ksplit, double *C, int ldc, const int *rmap,
                       const double *alpha, const double *beta, int tile);
/* rs_check_gemm_tile with B's row k read at row bmap[k] of B (ldb rows; no B2),
 * as the down pass reads a node's incoming rows where the level above wrote
 * them. */
int rs_check_gemm_bmap(int precision, int M, int N, int K, const double *A, int lda, const double *B, int ldb,
                       const int *bmap, double *C, int ldc, const double *alpha, const double *beta, int tile);

/* Kernel timing (no reference counterpart; the reference's only timer is
 * time_module).  While enabled, every solve-pass launch is bracketed by HIP
 * events on the stream it runs on, and its algorithmic bytes and flops are
 * recorded.  rs_profile_read() synchronises and returns, for one category,
 * the summed milliseconds, launch count, bytes and flops since the last read
 * (and clears that category).  Categories are numbered 0..rs_profile_ncat()-1. */
int rs_profile(rs_tree *t, int enable);
int rs_profile_ncat(void);
const char *rs_profile_name(int cat);
int rs_profile_read(rs_tree *t, int cat, double *ms, int *launches, double *bytes, double *flops);

/* ---------------------------------------------------------------------------
 * Serving loop (no reference counterpart: the reference solves one right-hand
 * side per call and process, robin_tree_module.f90:198-233, 1014-1029).  A
 * server keeps factored trees resident, queues solve and factor requests per
 * tree in arrival order, and runs them on `lanes` streams: whenever a lane is
 * free it takes the tree whose oldest queued solve is oldest and coalesces that
 * tree's queued solves (up to max_cols right-hand sides) into one multi-RHS pass.
 * Requests that arrive while passes run join the next pass on a free lane.
 * Each lane owns one solve workspace arena, shared by every tree it runs and
 * grown to the largest tree's workspace at max_cols; a tree keeps per lane only
 * the descriptors of its products there (rs_server_lane_bytes gives both).
 * Passes and factorizations are admitted by free device memory (hipMemGetInfo
 * less hbm_reserve, or hbm_limit less what the server holds); idle lane arenas
 * and descriptors are evicted to make room.  A factor request waits for the tree's
 * earlier solves and holds back its later ones.  Results do not depend on how
 * requests are batched: every kernel computes each right-hand side the same
 * way whatever the batch, and the refinement count is fixed per tree.
 * ------------------------------------------------------------------------- */
typedef struct rs_server rs_server;

typedef struct rs_server_opts {
    int max_cols;          /* right-hand sides per pass (default 256)                     */
    int lanes;             /* passes in flight at once, one stream each (default 2)       */
    long long hbm_reserve; /* device bytes admission leaves free (default 2 GiB)          */
    long long hbm_limit;   /* > 0: admit against this many bytes held by the server's trees
                              and workspaces instead of the device's free memory          */
    int simulate;          /* 1: no device work; a pass takes sim_pass_s + sim_col_s per
                              right-hand side and a factorization sim_factor_s (tests of
                              the scheduling policy on a host without a GPU)              */
    double sim_pass_s, sim_col_s, sim_factor_s;
    int group;             /* trees per pass (default 1): above 1, a pass also takes the
                              queued solves of further trees of the same precision (oldest
                              first, within max_cols and the lane's arena) and runs every
                              level's products as one grouped launch across them; the
                              lane's arena then also holds the descriptors              */
} rs_server_opts;

typedef struct rs_server_stats {
    long long requests_done;   /* solve requests completed (done or failed)        */
    long long rhs_done;        /* right-hand sides solved                          */
    long long passes;          /* multi-RHS passes run                             */
    long long factors_done;    /* factorizations completed                         */
    long long admission_waits; /* times a request was held back for device memory */
    long long evictions;       /* lane arenas and descriptors freed to admit work  */
    long long failed;          /* requests that failed                             */
    double busy_s;             /* wall seconds with at least one pass in flight    */
    double lat_p50_s, lat_p99_s, lat_max_s;  /* solve latency, submission to completion */
    double mean_cols;          /* mean right-hand sides per pass                   */
    long long held_bytes;      /* device bytes the server accounts to its trees and
                                  lanes                                            */
    double mean_trees;         /* mean trees per pass (above 1 only with group > 1) */
} rs_server_stats;

enum { RS_REQ_QUEUED = 0, RS_REQ_RUNNING = 1, RS_REQ_DONE = 2, RS_REQ_FAILED = 3 };

void rs_server_opts_default(rs_server_opts *o);
int rs_server_create(rs_server **s, const rs_server_opts *o);
/* Completes every queued request, then stops the loop and frees the server's
 * streams and tables (the trees stay with their owner; destroy them after). */
int rs_server_destroy(rs_server *s);
/* Register a planned (and possibly factored) tree; its id is returned.  While
 * registered, the tree must be used only through the server.  The server sets
 * RS_RETAIN_SCRATCH in the tree's options: its factorization's working memory
 * stays allocated, so that a refactor allocates and frees nothing while passes
 * of other trees run. */
int rs_server_add_tree(rs_server *s, rs_tree *t, int *tree_id);
/* Queue a (re)factorization of tree tree_id with the potential k_dev: device
 * memory [n_leaves][n_int] complex128, valid until the request completes. */
int rs_server_factor(rs_server *s, int tree_id, const void *k_dev, long long *req);
/* Queue a solve of nrhs right-hand sides: device arrays f [nrhs][n_leaves][n_int],
 * g [nrhs][n_leaves][n_bnd] in, u [nrhs][n_leaves][m] out (complex128), valid
 * until the request completes.  nrhs <= max_cols. */
int rs_server_solve(rs_server *s, int tree_id, int nrhs, const void *f_dev, const void *g_dev, void *u_dev,
                    long long *req);
/* State of a request (RS_REQ_*), without blocking. */
int rs_server_test(rs_server *s, long long req, int *state);
/* Block until the request completes: 0 done, 1 failed (message by rs_server_error). */
int rs_server_wait(rs_server *s, long long req);
/* Block until every queued request has completed. */
int rs_server_drain(rs_server *s);
/* Times of a request (seconds on one monotonic clock) and the pass it ran in. */
int rs_server_request_info(rs_server *s, long long req, double *t_submit, double *t_start, double *t_done,
                           long long *pass);
/* Forget a completed request's record. */
int rs_server_release(rs_server *s, long long req);
int rs_server_get_stats(rs_server *s, rs_server_stats *st);
int rs_server_reset_stats(rs_server *s);
/* Message of a failed request, or of the last failed call when req < 0. */
const char *rs_server_error(rs_server *s, long long req);
/* What a server with this max_cols holds for tree t (planned): the lane arena
 * the tree needs (a lane's arena is the largest of these over its trees) and
 * the tree's descriptors on one lane (with group > 1 the arena holds them:
 * arena + descs per tree).  Host arithmetic, no device needed. */
int rs_server_lane_bytes(const rs_tree *t, int max_cols, long long *arena, long long *descs);

#ifdef __cplusplus
}
#endif
#endif
