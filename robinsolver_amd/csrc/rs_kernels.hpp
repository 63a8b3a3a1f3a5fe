// rs_kernels.hpp -- host-side launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

#include "rs_common.hpp"

namespace rs {

enum LeafMode { LEAF_FACTOR_T = 0, LEAF_UP = 1, LEAF_DOWN = 2 };

// One problem of a batched complex GEMM:
//   Out[rmap(r), cmap(c)] = alpha * sum_k A[r,k] B[k,c] + beta * Din[...]
// All matrices column-major; rows k >= ksplit of B come from B2 (row k-ksplit).
// Din is indexed (r, c) with ldd, or through the output maps when dmap != 0.
struct GemmDesc {
    const void* A;
    const void* B;
    const void* B2;
    void* Out;
    const void* Din;
    const int* rmap;
    const int* cmap;
    const int* bmap;     // B row k at B + bmap[k] (elements) instead of B + k (no B2 then)
    int lda, ldb, ldo, ldd;
    int ksplit, dmap;
    int bpair;           // bmap[2j + 1] = bmap[2j] + 1 and bmap[2j] even for every j
    int reserved;
    double ar, ai, br, bi;
};

// ws: workspace of gemm_ws_elems(...) complex elements of precision R for the
// K-split form (NULL: no split).  tile: one of GEMM_TILE_* (output tile rows x
// columns), or -1 for the shape rule gemm_tile_rule; results do not depend on it.
// GEMM_TILE_SKINNY: 64 x 16 tiles whose waves read A straight from memory
// (products of few columns).  GEMM_TILE_GEMV: products of at most
// GEMM_GEMV_MAX_N columns on the vector ALUs, one lane per output row and K
// slice (other products take the 64 x 16 tiles under this choice).
// GEMM_TILE_128x64 / 64x128: complex64 products only (complex128 ones take
// the 64 x 64 tiles under these choices).  GEMM_TILE_D*: the operands move
// from memory straight into LDS stages in the precision they are stored in
// (complex128 operands take D64x64 under D64x128 and D128x128).
enum { GEMM_TILE_64x64 = 0, GEMM_TILE_32x64 = 1, GEMM_TILE_16x64 = 2, GEMM_TILE_64x16 = 3, GEMM_TILE_32x32 = 4,
       GEMM_TILE_64x32 = 5, GEMM_TILE_128x16 = 6, GEMM_TILE_SKINNY = 7, GEMM_TILE_GEMV = 8, GEMM_TILE_128x64 = 9,
       GEMM_TILE_64x128 = 10, GEMM_TILE_D64x64 = 11, GEMM_TILE_D128x64 = 12, GEMM_TILE_D64x128 = 13,
       GEMM_TILE_D64x32 = 14, GEMM_TILE_D128x128 = 15, GEMM_TILES = 16 };
constexpr int GEMM_GEMV_MAX_N = 4;
// whether a tile choice is a kernel of its own for an N-column product with
// products in FP64 (fp64) or FP32, with operands and results in the precision
// of the products (same) or in FP64 (the tuners skip the ones that would
// repeat another)
inline bool gemm_tile_distinct(int tile, int N, bool fp64, bool same = true) {
    if (tile == GEMM_TILE_GEMV) return N <= GEMM_GEMV_MAX_N;
    if (tile == GEMM_TILE_128x64 || tile == GEMM_TILE_64x128) return !fp64 && N > 16;
    if (tile == GEMM_TILE_D64x128 || tile == GEMM_TILE_D128x128) return !fp64 && same && N > 16;
    if (tile == GEMM_TILE_D64x64 || tile == GEMM_TILE_D64x32 || tile == GEMM_TILE_D128x64) return N > 16;
    return true;
}
int gemm_tile_rule(int M, int N);
// ks_target: the K split aims at this many (row tile x slice) workgroups per
// launch; it never depends on N.
template <typename R>
void launch_gemm(const GemmDesc* d_descs, int nprob, int M, int N, int K, hipStream_t st, void* ws = nullptr,
                 int tile = -1, int ks_target = 256);
// The same products on complex128 operands and results with complex64 MFMA
// products and FP32 accumulation (the merge of the mixed-precision
// factorization).  ws as for launch_gemm<float>.
void launch_gemm_c64(const GemmDesc* d_descs, int nprob, int M, int N, int K, hipStream_t st, void* ws = nullptr,
                     int tile = -1, int ks_target = 256);
// Several groups of problems of different shapes in one launch (problems of
// several trees at one level): group q's problems are products of its own M, N
// and K with its own K split (gemm_kslices of its own nprob, M, K and
// ks_target, exactly as in a launch of that group alone) into its own
// workspace, so a group's results are the same bits in either form.
constexpr int GEMM_MAX_GROUPS = 8;
struct GemmGroup {
    const GemmDesc* descs = nullptr;
    int nprob = 0, M = 0, N = 0, K = 0;
    void* ws = nullptr;
    int ks_target = 256;
};
// the kernel-side form of a launch (kernel argument)
struct GemmLaunch {
    struct Grp {
        const GemmDesc* descs;
        void* ws;
        int M, N, K, tilesM, tilesN, ks, kchunk, wg0, nprob;
    } g[GEMM_MAX_GROUPS];
    int n;
};
template <typename R>
void launch_gemm_grouped(const GemmGroup* groups, int ngroups, hipStream_t st, int tile = -1);
int gemm_kslices(int nprob, int M, int K, int ks_target = 256);
size_t gemm_ws_elems(int nprob, int M, int N, int K, int ks_target = 256);
// K-split target of the solve passes: with few right-hand sides a long-K product
// (the top-level 3D maps: K >= 4096, or K >= 2048 over at least 4096 rows in
// all) is a matrix-vector stream that needs more workgroups than its row tiles
// give.  The factor's products keep the default.
inline int solve_ks_target(int nprob, int M, int K) {
    return (K >= 4096 || (K >= 2048 && (long long)nprob * M >= 4096)) ? 1024 : 256;
}
// K-split target of the factorization's merge products: the column tiles
// count toward filling the chip too (a merge product has one fixed N, so
// nothing asks for the same bits at every N; wide products then need no
// split, no partial sums and no reduction).
inline int merge_ks_target(int N) {
    const int tn = (N + 63) / 64;
    return tn >= 128 ? 2 : 256 / tn;
}

// leaf stage
size_t leaf_blocks_lds(const LeafGeom& G);
void launch_leaf_blocks(const LeafGeom& G, const double2* kint, double2* sinv, int* info, int n_leaves,
                        hipStream_t st);
template <typename R> int leaf_apply_chunk(const LeafGeom& G, int ncols, size_t lds_limit);
// Arguments of a leaf pass (see LeafIO in k_leaf.hip).
// Caller-side complex128 arrays (f, g, u) are reached through column tables:
// column c at tab[c], laid out [leaf][len]; the solve's own complex64 arrays are
// contiguous [col][leaf][len].
template <typename R> struct LeafArgs {
    int mode = LEAF_UP, ncols = 0, n_leaves = 0;
    const c_t<R>* sinv = nullptr;
    const double2* const* f64 = nullptr;
    const float2* f32 = nullptr;
    const float2* jump = nullptr;
    const c_t<R>* gin = nullptr;       // DOWN: incoming data, row b of leaf q in column c at
    const int* gsrc = nullptr;         //   gin[gsrc[q * nb + b] + c * gld]
    size_t gld = 0;
    c_t<R>* hout = nullptr;
    double2* const* u64 = nullptr;
    float2* u32 = nullptr;
    double2* T = nullptr;
    const float2* ubase32 = nullptr;   // u64 = ubase32 + x (complex64 base, contiguous)
    int uacc = 0;                      // u64 += x
};
template <typename R>
void launch_leaf_apply(const LeafGeom& G, const LeafArgs<R>& a, size_t lds_limit, hipStream_t st);
// whether the leaf maps (LEAF_FACTOR_T, complex64) of this geometry run on the
// matrix-core leaf kernel (square 2D leaves of 12, 14 or 16 points per axis)
bool leaf_mfma_factor_ok(const LeafGeom& G);
// u is complex128 (u64, column table) or complex64 (u32); f and g are column
// tables; rf, rg and rj are written in complex64.
void launch_residual(const LeafGeom& G, double2* const* u64, const float2* u32, const double2* kint,
                     const double2* const* f, const double2* const* g, const int* nbr, double2 nu, double2 zeta,
                     float2* rf, float2* rg, float2* rj, int n_leaves, int ncols, hipStream_t st);

// per-column squared norms of the residual (rf, rg, rj) and of the data (f, g on
// the physical boundary): out[2c], out[2c+1]
void launch_fd_norms(const LeafGeom& G, const float2* rf, const float2* rg, const float2* rj,
                     const double2* const* f, const double2* const* g, const int* nbr, int n_leaves, int ncols,
                     double* out, hipStream_t st);

// batched inverse
struct InvPlan {
    int n = 0, batch = 0, nb = 0;
    int ob = 0;   // columns per block of panels (0: every panel updates the whole matrix)
    size_t stride = 0, lds_limit = 0;
    bool small = true;
    void build(int n, int batch, size_t stride, size_t lds_limit);
    // false when no kernel of launch_inverse takes an n-point matrix (above ~5k
    // points the LDS panel of even one column exceeds the LDS)
    bool fits() const;
    size_t w_elems() const;      // complex elements of launch_inverse's W (all matrices)
    size_t desc_count() const;   // GEMM descriptors of launch_inverse
    size_t cmap_elems() const;   // column-map entries they point to
    // the descriptors (d, desc_count()) and column maps (cmaps, cmap_elems(), to
    // be copied to cmap_dev) of launch_inverse on mats with the W buffer W
    void make_descs(double2* mats, double2* W, const int* cmap_dev, GemmDesc* d, int* cmaps) const;
};
size_t inv_small_lds(int n);
void launch_inverse(const InvPlan& ip, double2* mats, int* pivs, double2* W, const GemmDesc* d_descs, int* info,
                    hipStream_t st);

// data movement
void launch_form_iface(double2* out, const double2* T, size_t tstride, int child_mul, int child_add, int b, int off,
                       int isz, int nodes, double2 zeta, double2 coef, const double2* D, hipStream_t st);
void launch_copy_blocks(double2* dst, size_t dstride, int ldd, const double2* src, size_t sstride, int lds, int rows,
                        int cols, int batch, double2 scale, hipStream_t st);
// M[bat] (n x n, leading dimension ld, batch matrices stride elements apart) += s I
void launch_add_diag(double2* M, size_t stride, int ld, int n, int batch, double2 s, hipStream_t st);
template <typename V>
void launch_root_gather(V* out, size_t ldo, const double2* const* gin, const int* leaf_of, const int* loc_of, int P,
                        int nbl, int ncols, hipStream_t st);
// the same from complex64 data, contiguous [col][leaf][nbl]
template <typename V>
void launch_root_gather32(V* out, size_t ldo, const float2* gin, const int* leaf_of, const int* loc_of, int P,
                          int n_leaves, int nbl, int ncols, hipStream_t st);
void launch_to_c64(float2* dst, const double2* src, size_t n, hipStream_t st);
// per node [T_NI | T_IN] (bn x i and i x bn, column-major) out of full b x b
// node maps of stride tstride, into d64 and/or d32 (either may be NULL)
void launch_extract_ni_in(const double2* T, size_t tstride, int b, int bn, int isz, int nodes, double2* d64,
                          float2* d32, hipStream_t st);
void launch_zero(double2* p, size_t n, hipStream_t st);
void launch_axpy(double2* y, const double2* x, size_t n, hipStream_t st);
// tab[c] = base + c * stride for c < ncols (column table of a contiguous array)
void launch_col_table(double2** tab, double2* base, size_t stride, int ncols, hipStream_t st);
void launch_leaf_points(const double* box, const int* pat, const double* h, const int* n, int d, int m, int nl,
                        double* xyz, int* face, hipStream_t st);

}  // namespace rs
