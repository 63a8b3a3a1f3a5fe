// k_misc.hip -- data-movement kernels of the merge and the solve passes: the
// interface blocks of the merge (robin_tree_module.f90:440-467), the face
// scatter/gather between a node and its children (robin_tree_module.f90:896-959,
// 1045-1075, 1205-1230), and precision conversion of the stored factors.
#include <algorithm>

#include "rs_common.hpp"
#include "rs_kernels.hpp"

namespace rs {

namespace {

// out[n] = zeta I - T_child(n)[off:off+i, off:off+i] + coef * D[n]
__global__ void form_iface_kernel(double2* __restrict__ out, const double2* __restrict__ T, size_t tstride,
                                  int child_mul, int child_add, int b, int off, int isz, double2 zeta,
                                  double2 coef, const double2* __restrict__ D) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ii = (size_t)isz * isz;
    if (e >= ii) return;
    const int node = blockIdx.y;
    const int r = (int)(e % isz), c = (int)(e / isz);
    const double2* Tc = T + (size_t)(child_mul * node + child_add) * tstride;
    double2 v = make_double2(-Tc[(size_t)(off + r) + (size_t)(off + c) * b].x, -Tc[(size_t)(off + r) + (size_t)(off + c) * b].y);
    if (r == c) v = cadd(v, zeta);
    if (D) v = cfma(coef, D[(size_t)node * ii + e], v);
    out[(size_t)node * ii + e] = v;
}

// dst[bat] (rows x cols, ldd) = scale * src[bat] (lds)
__global__ void copy_blocks_kernel(double2* __restrict__ dst, size_t dstride, int ldd, const double2* __restrict__ src,
                                   size_t sstride, int lds, int rows, int cols, double2 scale) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)rows * cols) return;
    const int r = (int)(e % rows), c = (int)(e / rows);
    const size_t bat = blockIdx.y;
    dst[bat * dstride + r + (size_t)c * ldd] = cmul(scale, src[bat * sstride + r + (size_t)c * lds]);
}

// M[bat] (n x n, ld) += s I
__global__ void add_diag_kernel(double2* __restrict__ M, size_t stride, int ld, int n, double2 s) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double2* p = M + blockIdx.y * stride + (size_t)r * (ld + 1);
    *p = cadd(*p, s);
}

// root boundary data from the physical data g (column table, [leaf][nbl] per column)
template <typename V>
__global__ void root_gather_kernel(V* __restrict__ out, size_t ldo, const double2* const* __restrict__ gin,
                                   const int* __restrict__ leaf_of, const int* __restrict__ loc_of, int P, int nbl) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = blockIdx.y;
    if (p >= P) return;
    out[(size_t)col * ldo + p] = ccast<V>(gld(gin[col] + ((size_t)leaf_of[p] * nbl + loc_of[p])));
}

template <typename V>
__global__ void root_gather32_kernel(V* __restrict__ out, size_t ldo, const float2* __restrict__ gin,
                                     const int* __restrict__ leaf_of, const int* __restrict__ loc_of, int P,
                                     int n_leaves, int nbl) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = blockIdx.y;
    if (p >= P) return;
    out[(size_t)col * ldo + p] = ccast<V>(gin[((size_t)col * n_leaves + leaf_of[p]) * nbl + loc_of[p]]);
}

__global__ void to_c64_kernel(float2* __restrict__ dst, const double2* __restrict__ src, size_t n) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) dst[e] = make_float2((float)src[e].x, (float)src[e].y);
}

// The blocks of the node maps that the solve reads, out of full b x b maps (node
// order: the non-interface faces first, the interface last): per node
// [T_NI (bn x i, leading dimension bn) | T_IN (i x bn, leading dimension i)],
// into complex128 and/or complex64.  One thread per compact element, reads and
// writes both running along columns.
__global__ void extract_ni_in_kernel(const double2* __restrict__ T, size_t tstride, int b, int bn, int isz,
                                     size_t total, double2* __restrict__ d64, float2* __restrict__ d32) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const size_t per = 2 * (size_t)bn * isz, ni = (size_t)bn * isz;
    const size_t n = e / per, r = e - n * per;
    const double2* Tn = T + n * tstride;
    double2 v;
    if (r < ni) {
        const size_t c = r / bn, q = r - c * bn;
        v = Tn[q + (bn + c) * (size_t)b];
    } else {
        const size_t r2 = r - ni, c = r2 / isz, q = r2 - c * isz;
        v = Tn[(bn + q) + c * (size_t)b];
    }
    if (d64) d64[e] = v;
    if (d32) d32[e] = make_float2((float)v.x, (float)v.y);
}

__global__ void zero_kernel(double2* __restrict__ p, size_t n) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[e] = make_double2(0.0, 0.0);
}

__global__ void axpy_kernel(double2* __restrict__ y, const double2* __restrict__ x, size_t n) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) y[e] = cadd(y[e], x[e]);
}

__global__ void col_table_kernel(double2** __restrict__ tab, double2* base, size_t stride, int ncols) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncols) tab[c] = base + (size_t)c * stride;
}

inline unsigned nblk(size_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void launch_form_iface(double2* out, const double2* T, size_t tstride, int child_mul, int child_add, int b, int off,
                       int isz, int nodes, double2 zeta, double2 coef, const double2* D, hipStream_t st) {
    hipLaunchKernelGGL(form_iface_kernel, dim3(nblk((size_t)isz * isz), nodes), dim3(256), 0, st, out, T, tstride,
                       child_mul, child_add, b, off, isz, zeta, coef, D);
}

void launch_copy_blocks(double2* dst, size_t dstride, int ldd, const double2* src, size_t sstride, int lds, int rows,
                        int cols, int batch, double2 scale, hipStream_t st) {
    hipLaunchKernelGGL(copy_blocks_kernel, dim3(nblk((size_t)rows * cols), batch), dim3(256), 0, st, dst, dstride,
                       ldd, src, sstride, lds, rows, cols, scale);
}

void launch_add_diag(double2* M, size_t stride, int ld, int n, int batch, double2 s, hipStream_t st) {
    if (n > 0 && batch > 0)
        hipLaunchKernelGGL(add_diag_kernel, dim3(nblk((size_t)n), batch), dim3(256), 0, st, M, stride, ld, n, s);
}

template <typename V>
void launch_root_gather(V* out, size_t ldo, const double2* const* gin, const int* leaf_of, const int* loc_of, int P,
                        int nbl, int ncols, hipStream_t st) {
    hipLaunchKernelGGL(root_gather_kernel<V>, dim3(nblk(P), ncols), dim3(256), 0, st, out, ldo, gin, leaf_of, loc_of,
                       P, nbl);
}

template <typename V>
void launch_root_gather32(V* out, size_t ldo, const float2* gin, const int* leaf_of, const int* loc_of, int P,
                          int n_leaves, int nbl, int ncols, hipStream_t st) {
    hipLaunchKernelGGL(root_gather32_kernel<V>, dim3(nblk(P), ncols), dim3(256), 0, st, out, ldo, gin, leaf_of,
                       loc_of, P, n_leaves, nbl);
}

void launch_to_c64(float2* dst, const double2* src, size_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(to_c64_kernel, dim3(nblk(n)), dim3(256), 0, st, dst, src, n);
}

void launch_extract_ni_in(const double2* T, size_t tstride, int b, int bn, int isz, int nodes, double2* d64,
                          float2* d32, hipStream_t st) {
    const size_t total = (size_t)nodes * 2 * bn * isz;
    if (total) hipLaunchKernelGGL(extract_ni_in_kernel, dim3(nblk(total)), dim3(256), 0, st, T, tstride, b, bn, isz,
                                  total, d64, d32);
}

void launch_zero(double2* p, size_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(zero_kernel, dim3(nblk(n)), dim3(256), 0, st, p, n);
}

void launch_axpy(double2* y, const double2* x, size_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(axpy_kernel, dim3(nblk(n)), dim3(256), 0, st, y, x, n);
}

// Coordinates of every point of every leaf (the host's Plan::leaf_points on
// the device): xyz [leaf][m][3], face [leaf][m].  box: [leaf][lo x, lo y, lo z,
// hi x, hi y, hi z]; pat: [m][g0, g1, g2, face].  The coordinate expressions are
// evaluated without contraction, so the bits are the host's.
__global__ void leaf_points_kernel(const double* __restrict__ box, const int* __restrict__ pat, double3 h, int3 n,
                                   int d, int m, int nl, double* __restrict__ xyz, int* __restrict__ face) {
#pragma clang fp contract(off)
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nl * m) return;
    const int leaf = (int)(e / m), p = (int)(e - (size_t)leaf * m);
    const double* b = box + 6 * (size_t)leaf;
    const int* g = pat + 4 * p;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const double hj = j == 0 ? h.x : (j == 1 ? h.y : h.z);
        const int nj = j == 0 ? n.x : (j == 1 ? n.y : n.z);
        const int gj = g[j];
        double v;
        if (j < d && gj == 0) v = b[j];
        else if (j < d && gj == nj + 1) v = b[3 + j];
        else v = b[j] + gj * hj - hj / 2;
        xyz[3 * e + j] = v;
    }
    if (face) face[e] = g[3];
}

void launch_leaf_points(const double* box, const int* pat, const double* h, const int* n, int d, int m, int nl,
                        double* xyz, int* face, hipStream_t st) {
    const size_t tot = (size_t)nl * m;
    if (tot == 0) return;
    hipLaunchKernelGGL(leaf_points_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, box, pat,
                       make_double3(h[0], h[1], h[2]), make_int3(n[0], n[1], n[2]), d, m, nl, xyz, face);
}

void launch_col_table(double2** tab, double2* base, size_t stride, int ncols, hipStream_t st) {
    if (ncols > 0) hipLaunchKernelGGL(col_table_kernel, dim3(nblk(ncols)), dim3(256), 0, st, tab, base, stride, ncols);
}

template void launch_root_gather<double2>(double2*, size_t, const double2* const*, const int*, const int*, int, int,
                                          int, hipStream_t);
template void launch_root_gather<float2>(float2*, size_t, const double2* const*, const int*, const int*, int, int, int,
                                         hipStream_t);
template void launch_root_gather32<double2>(double2*, size_t, const float2*, const int*, const int*, int, int, int,
                                            int, hipStream_t);
template void launch_root_gather32<float2>(float2*, size_t, const float2*, const int*, const int*, int, int, int, int,
                                           hipStream_t);

}  // namespace rs
