// k_inv.hip -- batched FP64 complex matrix inversion with partial pivoting
// for the interface systems of the merge.
//
// Reference: D = (zeta I - T2_II)^-1 by zgetrf + zgetri and the LU of
// S^-1 = zeta I - T1_II - nu^2 D by zgetrf (robin_tree_module.f90:440-467;
// factorization_module.f90:49-141).  The merge here inverts one matrix per
// node, W = nu^2 I - (zeta I - T1_II)(zeta I - T2_II), from which it forms
// inv(M) by products; the inverse is explicit (Gauss-Jordan, the pivot being
// the first entry of maximal |re|+|im| in the column, as zgetrf chooses it;
// the register panels compare the candidates on the top 52 bits of that
// value, so they can choose differently only between candidates equal to 12
// digits), so that the merge and the solve passes are plain matrix products
// on the matrix cores.
//
// Interfaces below 64 points use one workgroup per matrix with the whole matrix
// in LDS.  Larger ones use a blocked right-looking Gauss-Jordan: an n x nb
// panel is eliminated (in registers up to 4096 points, in LDS above), the row
// interchanges are applied to the other columns and the pivot rows moved out
// (W), and the rest of the matrix gets the rank-nb update A[:, R] += A[:, panel] W
// as a batched GEMM.  After the last panel the interchanges are undone as
// column interchanges.
#include <algorithm>
#include <cstring>
#include <utility>

#include "rs_common.hpp"
#include "rs_kernels.hpp"

namespace rs {

namespace {

// block-wide argmax of |re|+|im|; ties go to the smallest index
__device__ void block_argmax(double v, int i, double* rv, int* ri, double& best, int& bi) {
    wave_argmax(v, i);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) { rv[wave] = v; ri[wave] = i; }
    __syncthreads();
    best = rv[0]; bi = ri[0];
    for (int w = 1; w < nw; ++w)
        if (rv[w] > best || (rv[w] == best && ri[w] < bi)) { best = rv[w]; bi = ri[w]; }
    __syncthreads();
}

// Whole-matrix Gauss-Jordan in LDS, one workgroup per matrix.
__global__ void __launch_bounds__(256) inv_small_kernel(double2* __restrict__ mats, int n, size_t stride,
                                                       int* __restrict__ info) {
    extern __shared__ double2 sm[];
    double2* M = sm;
    double2* colj = sm + (size_t)n * n;
    int* piv = (int*)(colj + n);
    __shared__ double rv[4];
    __shared__ int ri[4];
    __shared__ int bad;
    const int tid = threadIdx.x, nt = blockDim.x;
    double2* G = mats + (size_t)blockIdx.x * stride;
    if (tid == 0) bad = 0;
    for (int e = tid; e < n * n; e += nt) M[e] = G[e];
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        double v = -1.0;
        int vi = n;
        for (int r = j + tid; r < n; r += nt) {
            double a = cabs1(M[r + j * n]);
            if (a > v) { v = a; vi = r; }
        }
        double best;
        int p;
        block_argmax(v, vi, rv, ri, best, p);
        if (!(best > 0.0) || p >= n) p = j;   // singular: no interchange, reported through info
        if (tid == 0) { piv[j] = p; if (!(best > 0.0)) bad = 1; }
        if (p != j)
            for (int c = tid; c < n; c += nt) {
                double2 t = M[j + c * n]; M[j + c * n] = M[p + c * n]; M[p + c * n] = t;
            }
        __syncthreads();
        for (int r = tid; r < n; r += nt) colj[r] = M[r + j * n];
        __syncthreads();
        const double2 inv = cinv(colj[j]);
        for (int c = tid; c < n; c += nt) M[j + c * n] = (c == j) ? inv : cmul(M[j + c * n], inv);
        __syncthreads();
        for (int e = tid; e < n * n; e += nt) {
            int r = e % n, c = e / n;
            if (r == j) continue;
            double2 base = (c == j) ? make_double2(0.0, 0.0) : M[e];
            M[e] = csub(base, cmul(colj[r], M[j + c * n]));
        }
        __syncthreads();
    }
    for (int r = tid; r < n; r += nt)
        for (int j = n - 1; j >= 0; --j) {
            int p = piv[j];
            if (p != j) { double2 t = M[r + j * n]; M[r + j * n] = M[r + p * n]; M[r + p * n] = t; }
        }
    __syncthreads();
    for (int e = tid; e < n * n; e += nt) G[e] = M[e];
    if (tid == 0 && bad) atomicOr(info, 2);
}

// Panel elimination of columns [j0, j0+nb) over all n rows, in LDS.
__global__ void __launch_bounds__(256) gj_panel_kernel(double2* __restrict__ mats, int n, size_t stride, int j0,
                                                      int nb, int* __restrict__ pivs, int* __restrict__ info) {
    extern __shared__ double2 sm[];
    double2* P = sm;              // [nb][n] column-major panel
    double2* colj = sm + (size_t)n * nb;
    __shared__ double rv[4];
    __shared__ int ri[4];
    __shared__ int bad;
    const int tid = threadIdx.x, nt = blockDim.x;
    double2* G = mats + (size_t)blockIdx.x * stride + (size_t)j0 * n;
    int* piv = pivs + (size_t)blockIdx.x * n;
    if (tid == 0) bad = 0;
    for (int e = tid; e < n * nb; e += nt) P[e] = G[e];
    __syncthreads();
    for (int jj = 0; jj < nb; ++jj) {
        const int j = j0 + jj;
        double v = -1.0;
        int vi = n;
        for (int r = j + tid; r < n; r += nt) {
            double a = cabs1(P[r + jj * n]);
            if (a > v) { v = a; vi = r; }
        }
        double best;
        int p;
        block_argmax(v, vi, rv, ri, best, p);
        if (!(best > 0.0) || p >= n) p = j;   // singular: no interchange, reported through info
        if (tid == 0) { piv[j] = p; if (!(best > 0.0)) bad = 1; }
        if (p != j)
            for (int c = tid; c < nb; c += nt) {
                double2 t = P[j + c * n]; P[j + c * n] = P[p + c * n]; P[p + c * n] = t;
            }
        __syncthreads();
        for (int r = tid; r < n; r += nt) colj[r] = P[r + jj * n];
        __syncthreads();
        const double2 inv = cinv(colj[j]);
        for (int c = tid; c < nb; c += nt) P[j + c * n] = (c == jj) ? inv : cmul(P[j + c * n], inv);
        __syncthreads();
        for (int e = tid; e < n * nb; e += nt) {
            int r = e % n, c = e / n;
            if (r == j) continue;
            double2 base = (c == jj) ? make_double2(0.0, 0.0) : P[e];
            P[e] = csub(base, cmul(colj[r], P[j + c * n]));
        }
        __syncthreads();
    }
    for (int e = tid; e < n * nb; e += nt) G[e] = P[e];
    if (tid == 0 && bad) atomicOr(info, 2);
}

// Panel elimination with the panel held in registers: thread t owns rows
// t, t + nt, ... (RPT of them) of the n x w panel (w <= NB columns, n <= 4096).
// One Gauss-Jordan step has one workgroup barrier: each wave finds its best
// row by a DPP argmax, the lane holding that row publishes it (with the old
// row j, from its owner), and the waves' results meet in one LDS atomic max;
// after the barrier every thread reads the pivot row, inverts its pivot and
// applies the rank-1 update to its own rows.  The pivot is the row of largest
// |re|+|im| compared on its top 52 bits (relative resolution 2^-41; the
// smallest row wins among equals), so it can differ from gj_panel_kernel's
// only between candidates equal to 12 digits.  The published rows are
// double-buffered by step parity: the writes of step s + 2 come after the
// barrier of step s + 1, which every reader of step s has passed.
//
// The step loop stays rolled.  The kernel runs once per panel, so a fully
// unrolled body is fetched from L2 instruction by instruction: unrolled over
// the 16 steps (10k-17k instructions) it ran at about 3.5 ns per instruction.
// To keep every register index static in a rolled loop, the columns rotate
// left by one each step: the current column is always a[.][0], and the column
// the step produces enters at a[.][NB-1]; after w steps original column c sits
// at position (c - w) mod NB.
template <int NB, int RPT, int LB = 1024>
__global__ void __launch_bounds__(LB) gj_panel_reg_kernel(double2* __restrict__ mats, int n, size_t stride,
                                                            int j0, int w, int* __restrict__ pivs,
                                                            int* __restrict__ info) {
    __shared__ unsigned long long best3[3];   // the step's pivot key (three slots: see below)
    __shared__ double2 cand[2][16][NB];   // each wave's best row (rotated like the registers)
    __shared__ double2 jrow[2][NB];       // old row j (becomes row p)
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    double2* G = mats + (size_t)blockIdx.x * stride + (size_t)j0 * n;
    int* piv = pivs + (size_t)blockIdx.x * n;
    double2 a[RPT][NB];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int r = tid + k * nt;
#pragma unroll
        for (int c = 0; c < NB; ++c) a[k][c] = (r < n && c < w) ? G[r + (size_t)c * n] : make_double2(0.0, 0.0);
    }
    if (tid < 3) best3[tid] = 0;
    __syncthreads();
    bool bad = false;
#pragma unroll 1
    for (int s = 0; s < w; ++s) {
        const int j = j0 + s, b = s & 1, b3 = s % 3;
        double v = -1.0;
        int vi = n;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = tid + k * nt;
            if (r >= j && r < n) {
                const double x = cabs1(a[k][0]);
                if (x > v) { v = x; vi = r; }
            }
        }
        wave_argmax(v, vi);
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = tid + k * nt;
            if (r == vi)
#pragma unroll
                for (int c = 0; c < NB; ++c) cand[b][wave][c] = a[k][c];
            if (r == j)
#pragma unroll
                for (int c = 0; c < NB; ++c) jrow[b][c] = a[k][c];
        }
        // the waves' candidates meet in one LDS atomic max of a key: the bits of
        // |re|+|im| (>= 0, so ordered like the values) above the lowest 11, then
        // 4095 - row (the smaller row wins among equal values)
        if (lane == 0 && v >= 0.0)
            atomicMax(&best3[b3], ((unsigned long long)__double_as_longlong(v) >> 11 << 12) |
                                      (unsigned long long)(4095 - vi));
        __syncthreads();
        const unsigned long long key = best3[b3];
        int p = 4095 - (int)(key & 4095);
        // slot (s + 2) % 3 was read in step s - 1, before this step's barrier, and is
        // next written in step s + 2, after the next one
        if (tid == 0) best3[(s + 2) % 3] = 0;
        if ((key >> 12) == 0 || p >= n) { bad = true; p = j; }   // singular: no interchange, reported through info
        const int qw = (p % nt) >> 6;   // the wave holding row p
        if (tid == 0) piv[j] = p;
        const double2* prow = cand[b][qw];   // the pivot row
        // 1 / pivot = conj(d) / |d|^2, with d scaled by a power of two (exact)
        // against overflow of |d|^2: one division on the step's critical path
        // (every thread: no divergent branch)
        double2 inv;
        {
            const double2 d = prow[0];
            const int ex = ilogb(fmax(fmax(fabs(d.x), fabs(d.y)), 1e-300));
            const double bx = scalbn(d.x, -ex), by = scalbn(d.y, -ex);
            const double rd = scalbn(1.0 / (bx * bx + by * by), -ex);
            inv = make_double2(bx * rd, -by * rd);
        }
        // the rank-1 update, rotated: a[c - 1] <- new column c, a[NB - 1] <- new column 0
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = tid + k * nt;
            if (r == j) {
#pragma unroll
                for (int c = 1; c < NB; ++c) a[k][c - 1] = cmul(prow[c], inv);
                a[k][NB - 1] = inv;
                continue;
            }
            if (r == p)
#pragma unroll
                for (int c = 0; c < NB; ++c) a[k][c] = jrow[b][c];
            const double2 m = cmul(a[k][0], inv);
#pragma unroll
            for (int c = 1; c < NB; ++c) a[k][c - 1] = cfms4(a[k][c], m, prow[c]);
            a[k][NB - 1] = make_double2(-m.x, -m.y);
        }
    }
    // (w == NB, the usual case, has the columns back in place: static addresses,
    // stores issued back to back)
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int r = tid + k * nt;
        if (r < n) {
            if (w == NB) {
#pragma unroll
                for (int c = 0; c < NB; ++c) G[r + (size_t)c * n] = a[k][c];
            } else {
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    const int oc = c + w < NB ? c + w : c + w - NB;   // original column at position c
                    if (oc < w) G[r + (size_t)oc * n] = a[k][c];
                }
            }
        }
    }
    if (bad && tid == 0) atomicOr(info, 2);
}

// Apply the row interchanges of pivots j0 .. j0+np-1 (in order) to the columns
// of [cb, ce) outside [j0, j0+np) (a subrange of it), move those columns'
// rows j0 .. j0+np-1 to W (np x (ce - cb - np), compact) and zero them.
// A group of np lanes per column (np <= 64 divides 256: a column's lanes are
// in one wave): the interchanges touch the contiguous rows j0 .. j0+np-1 and
// the rows piv[j], so lane i reads the rows whose values end at j0 + i and at
// piv[j0 + i] (from the composed permutation, built once per workgroup in
// LDS), moves the first to W, zeroes row j0 + i and stores the second.  Every
// load of a column comes before its stores in program order (the stores wait
// for the loaded values).
__global__ void __launch_bounds__(256) gj_swap_extract_grp_kernel(double2* __restrict__ mats, int n, size_t stride,
                                                                  int j0, int nb, int cb, int ce,
                                                                  const int* __restrict__ pivs,
                                                                  double2* __restrict__ W, size_t wstride) {
    extern __shared__ int rowat[];   // [n]: row whose value ends at each (touched) position
    const int tid = threadIdx.x;
    const int* piv = pivs + (size_t)blockIdx.y * n;
    if (tid < nb) {
        rowat[j0 + tid] = j0 + tid;
        const int p = piv[j0 + tid];
        rowat[p] = p;   // (a position in both sets gets the same value twice)
    }
    __syncthreads();
    if (tid == 0)
        for (int jj = 0; jj < nb; ++jj) {
            const int j = j0 + jj, p = piv[j];
            if (p != j) { const int t = rowat[j]; rowat[j] = rowat[p]; rowat[p] = t; }
        }
    __syncthreads();
    const int g = tid / nb, i = tid - g * nb;
    const int c2 = blockIdx.x * (256 / nb) + g;
    if (c2 >= ce - cb - nb) return;
    const int c = cb + c2 < j0 ? cb + c2 : cb + c2 + nb;
    double2* col = mats + (size_t)blockIdx.y * stride + (size_t)c * n;
    const int j = j0 + i, p = piv[j];
    const bool outside = p >= j0 + nb;   // (a pivot row inside the pivots' rows is some lane's j)
    const double2 v1 = col[rowat[j]];
    const double2 v2 = outside ? col[rowat[p]] : make_double2(0.0, 0.0);
    W[(size_t)blockIdx.y * wstride + (size_t)c2 * nb + i] = v1;
    if (outside) col[p] = v2;
    col[j] = make_double2(0.0, 0.0);
}

// (one thread per column, sequential interchanges: interfaces too large for the
// LDS position table of the kernel above, or pivot counts it does not take)
__global__ void gj_swap_extract_kernel(double2* __restrict__ mats, int n, size_t stride, int j0, int nb, int cb,
                                       int ce, const int* __restrict__ pivs, double2* __restrict__ W,
                                       size_t wstride) {
    const int c2 = blockIdx.x * blockDim.x + threadIdx.x;
    if (c2 >= ce - cb - nb) return;
    const int c = cb + c2 < j0 ? cb + c2 : cb + c2 + nb;
    double2* col = mats + (size_t)blockIdx.y * stride + (size_t)c * n;
    const int* piv = pivs + (size_t)blockIdx.y * n;
    double2* w = W + (size_t)blockIdx.y * wstride + (size_t)c2 * nb;
    for (int jj = 0; jj < nb; ++jj) {
        int j = j0 + jj, p = piv[j];
        if (p != j) { double2 t = col[j]; col[j] = col[p]; col[p] = t; }
    }
    for (int jj = 0; jj < nb; ++jj) { w[jj] = col[j0 + jj]; col[j0 + jj] = make_double2(0.0, 0.0); }
}

__global__ void gj_unpermute_kernel(double2* __restrict__ mats, int n, size_t stride, const int* __restrict__ pivs) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double2* M = mats + (size_t)blockIdx.y * stride;
    const int* piv = pivs + (size_t)blockIdx.y * n;
    for (int j = n - 1; j >= 0; --j) {
        int p = piv[j];
        if (p != j) { double2 t = M[r + (size_t)j * n]; M[r + (size_t)j * n] = M[r + (size_t)p * n]; M[r + (size_t)p * n] = t; }
    }
}

}  // namespace

size_t inv_small_lds(int n) { return (size_t)n * n * sizeof(double2) + (size_t)n * (sizeof(double2) + sizeof(int)); }

// rows per thread of the register panel kernel (0: use the LDS panel kernel).
// Interfaces of 513..1024 points take two rows per thread on half as many
// threads, still with 16-column panels (a third fewer waves to synchronise per
// step).  A step costs about 2.3 us at 512 points whatever the panel width (its
// latency chain: argmax, barrier, wave scan, pivot inverse, update), so
// 32-column panels were slower (103 us per panel against 2 x 37.6 us; factor at
// 1024^2 33.1 -> 34.7 ms) although they halve the launches.  Timed alone (one
// 16-column panel, events around the launch): at 1024 points about 15 us go to
// the launch and to one CU loading and storing the panel, 14 us to the pivot
// searches and barriers, 12 us to the updates; two rows per thread at 512 points
// (four waves instead of eight) measured the same factor time.
static int gj_reg_rows(int n) { return n <= 512 ? 1 : (n <= 2048 ? 2 : (n <= 4096 ? 4 : 0)); }

int gj_panel_width(int n, size_t lds_limit) {
    const int rpt = gj_reg_rows(n);
    int nb = rpt ? 16 / rpt : 16;
    if (rpt == 2 && n <= 1024) nb = 16;
    if (!rpt)
        while (nb > 1 && ((size_t)n * nb + n) * sizeof(double2) + 1024 > lds_limit) nb >>= 1;
    while (nb > 1 && n % nb) nb >>= 1;
    return nb;
}

void InvPlan::build(int n_, int batch_, size_t stride_, size_t lds_limit_) {
    n = n_; batch = batch_; stride = stride_; lds_limit = lds_limit_;
    // the whole-matrix LDS kernel (five barriers per column) below 64 points; from
    // 64 on the register panels are faster (64 points, 256 / 128 matrices:
    // 252 / 246 us with the LDS kernel)
    small = n < 64 && inv_small_lds(n) + 512 <= lds_limit;
    nb = small ? n : gj_panel_width(n, lds_limit);
    // Above 1024 points the register panels are 4 or 8 columns wide, and a
    // rank-nb update per panel would stream the whole matrix n / nb times:
    // the panels are grouped into blocks of OB columns instead, each panel's
    // update reaching only the rest of its block, and the columns outside a
    // block get one rank-OB update from the block's composed transform.
    ob = (!small && gj_reg_rows(n) >= 2 && nb < 16) ? 64 : 0;
}

bool InvPlan::fits() const {
    if (small || gj_reg_rows(n)) return true;   // register panels: rows up to 4096 (12-bit row in the pivot key)
    return nb >= 1 && ((size_t)n * nb + n) * sizeof(double2) + 1024 <= lds_limit;
}

size_t InvPlan::w_elems() const {
    if (small) return 0;
    if (!ob) return (size_t)nb * (n - nb) * batch;
    size_t w = 0;
    for (int J = 0; J < n; J += ob) {
        const int bw = std::min(ob, n - J);
        w = std::max(w, std::max((size_t)nb * (bw - nb), (size_t)bw * (n - bw)));
    }
    return w * batch;
}

// The update products of launch_inverse in launch order, each a batch of
// descriptors: unblocked, one rank-nb update per panel; blocked, per block its
// panels' updates inside the block, then the rank-bw update of the rest.
template <typename F> static void inv_steps(const InvPlan& ip, F&& f) {
    const int n = ip.n, nb = ip.nb;
    if (!ip.ob) {
        for (int j0 = 0; j0 < n; j0 += nb) f(j0, nb, 0, n, true);
        return;
    }
    for (int J = 0; J < n; J += ip.ob) {
        const int bw = std::min(ip.ob, n - J);
        for (int p = J; p < J + bw; p += nb) f(p, nb, J, J + bw, true);
        f(J, bw, 0, n, false);
    }
}

size_t InvPlan::desc_count() const {
    if (small) return 0;
    size_t c = 0;
    inv_steps(*this, [&](int, int np, int cb, int ce, bool) { if (ce - cb - np > 0) c += batch; });
    return c;
}

size_t InvPlan::cmap_elems() const {
    if (small) return 0;
    size_t c = 0;
    inv_steps(*this, [&](int, int np, int cb, int ce, bool) { c += ce - cb - np; });
    return c;
}

void InvPlan::make_descs(double2* mats, double2* W, const int* cmap_dev, GemmDesc* d, int* cmaps) const {
    if (small) return;
    size_t nd = 0, nc = 0;
    inv_steps(*this, [&](int j0, int np, int cb, int ce, bool) {
        const int m = ce - cb - np;
        for (int c2 = 0; c2 < m; ++c2) cmaps[nc + c2] = cb + c2 < j0 ? cb + c2 : cb + c2 + np;
        if (m > 0)
            for (int b = 0; b < batch; ++b) {
                GemmDesc g;
                std::memset(&g, 0, sizeof g);
                double2* Mb = mats + (size_t)b * stride;
                g.A = Mb + (size_t)j0 * n; g.lda = n;
                g.B = W + (size_t)b * (w_elems() / batch); g.ldb = np;
                g.Out = Mb; g.ldo = n;
                g.cmap = cmap_dev + nc;
                g.Din = Mb; g.ldd = n; g.dmap = 1;
                g.ar = 1.0; g.br = 1.0; g.ksplit = 1 << 30;
                d[nd++] = g;
            }
        nc += m;
    });
}

void launch_inverse(const InvPlan& ip, double2* mats, int* pivs, double2* W, const GemmDesc* d_descs,
                    int* info, hipStream_t st) {
    if (ip.batch <= 0) return;
    if (ip.small) {
        hipLaunchKernelGGL(inv_small_kernel, dim3(ip.batch), dim3(256), inv_small_lds(ip.n), st, mats, ip.n,
                           ip.stride, info);
        return;
    }
    const int n = ip.n, nb = ip.nb;
    const size_t plds = ((size_t)n * nb + n) * sizeof(double2);
    const size_t wstride = ip.w_elems() / ip.batch;
    const int rpt = gj_reg_rows(n);
    const int nt = rpt ? std::min(1024, ((n + rpt - 1) / rpt + 63) / 64 * 64) : 256;
    size_t nd = 0;
    inv_steps(ip, [&](int j0, int np, int cb, int ce, bool panel) {
        if (panel) {
            if (rpt == 1)
                hipLaunchKernelGGL((gj_panel_reg_kernel<16, 1>), dim3(ip.batch), dim3(nt), 0, st, mats, n, ip.stride,
                                   j0, nb, pivs, info);
            else if (rpt == 2 && nb == 16)
                hipLaunchKernelGGL((gj_panel_reg_kernel<16, 2, 512>), dim3(ip.batch), dim3(nt), 0, st, mats, n,
                                   ip.stride, j0, nb, pivs, info);
            else if (rpt == 2)
                hipLaunchKernelGGL((gj_panel_reg_kernel<8, 2>), dim3(ip.batch), dim3(nt), 0, st, mats, n, ip.stride,
                                   j0, nb, pivs, info);
            else if (rpt == 4)
                hipLaunchKernelGGL((gj_panel_reg_kernel<4, 4>), dim3(ip.batch), dim3(nt), 0, st, mats, n, ip.stride,
                                   j0, nb, pivs, info);
            else
                hipLaunchKernelGGL(gj_panel_kernel, dim3(ip.batch), dim3(256), plds, st, mats, n, ip.stride, j0, nb,
                                   pivs, info);
        }
        const int m = ce - cb - np;
        if (m <= 0) return;
        if ((size_t)n * sizeof(int) <= 32768 && np <= 64 && (np & (np - 1)) == 0) {
            const int cpb = 256 / np;
            hipLaunchKernelGGL(gj_swap_extract_grp_kernel, dim3((m + cpb - 1) / cpb, ip.batch), dim3(256),
                               (size_t)n * sizeof(int), st, mats, n, ip.stride, j0, np, cb, ce, pivs, W, wstride);
        } else {
            hipLaunchKernelGGL(gj_swap_extract_kernel, dim3((m + 255) / 256, ip.batch), dim3(256), 0, st, mats, n,
                               ip.stride, j0, np, cb, ce, pivs, W, wstride);
        }
        // a panel's rank-nb update (of the rest of the matrix, or of the rest of
        // its block) streams its columns once (at 1024^2 the top levels' launches
        // read and write 16.5 MB each in 8.5 us, 3.9 TB/s); 32x32 tiles (four
        // times the workgroups of 64x64) measured fastest (factor at 1024^2:
        // 64x64 38.4 ms, 32x64 38.0, 64x32 38.0, 16x64 38.0, 32x32 37.8); a
        // block's rank-64 update is a matrix-core product
        launch_gemm<double>(d_descs + nd, ip.batch, n, m, np, st, nullptr,
                            np >= 32 ? GEMM_TILE_D64x64 : GEMM_TILE_32x32);
        nd += ip.batch;
    });
    hipLaunchKernelGGL(gj_unpermute_kernel, dim3((n + 255) / 256, ip.batch), dim3(256), 0, st, mats, n, ip.stride,
                       pivs);
}

}  // namespace rs
