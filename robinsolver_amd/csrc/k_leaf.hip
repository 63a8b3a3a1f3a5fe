// k_leaf.hip -- leaf stage on gfx950: batched factorization of every leaf FD
// system and the leaf solves of the factor (RtR maps) and of both solve passes.
//
// Reference: the dense m x m leaf matrix of fd_module.f90:49-224, factored by
// zgetrf (fd_module.f90:273) and applied by zgetrs (fd_module.f90:275, 434, 562),
// with the outgoing-data stencil of robin_tree_module.f90:667-718 and
// fd_module.f90:449-513.
//
// MI355X design.  The leaf matrix A = [A_BB A_BI; A_IB A_II] has a diagonal A_BB
// (each boundary row couples its point to the two interior points along the
// normal), so the boundary unknowns are eliminated exactly:
//   u_I = S^-1 (f - A_IB A_BB^-1 g),  u_B = A_BB^-1 (g - A_BI u_I),
//   S   = A_II - A_IB A_BB^-1 A_BI    (same 5/7-point pattern as A_II).
// Ordered slow-axis last, S is block tridiagonal with dense s x s diagonal
// blocks (one x-row in 2D, one xy-layer in 3D) and off-diagonal blocks that are
// scalar multiples of the identity (the slow-axis stencil weight only depends
// on the layer index).  Block LU therefore needs only the K inverses
//   Shat_k^-1,  Shat_k = S_kk - lambda_k mu_{k-1} Shat_{k-1}^-1,
// each computed in LDS by Gauss-Jordan with partial pivoting.  Those K blocks
// (K s^2 entries, 4096 for a 16x16 leaf instead of 320^2 for the dense LU) are
// all that the solve streams per leaf.  The outgoing data needs no stencil
// pass: the boundary row gives (alpha + D_n) u = g, so
//   h = (sigma + D_n) u = g + (sigma - alpha) u_B.
#include <algorithm>
#include <type_traits>

#include "rs_common.hpp"
#include "rs_kernels.hpp"

namespace rs {

namespace {

// Per-axis constants of the eliminated system.
struct AxisCoef {
    double2 dd;     // A_BB entry: alpha + 8/(3h)
    double2 bndd;   // diagonal of S gained from eliminating the adjacent boundary point: 8/(h^3 dd)
    double2 eoff;   // S coupling of an edge point to its inward neighbour: (4/3)/h^2 - (8/9)/(h^3 dd)
    double cib;     // A_IB entry: (8/3)/h^2
    double mid;     // 1/h^2
    double w0e;     // centre weight at an edge point: -4/h^2
    double w0m;     // centre weight elsewhere: -2/h^2
    double a1, a2;  // A_BI entries: -3/h at the first interior point, 1/(3h) at the second
};

struct LeafCoef {
    AxisCoef ax[3];
};

__host__ LeafCoef make_coef(const LeafGeom& G) {
    LeafCoef C{};
    for (int j = 0; j < G.d; ++j) {
        double h = G.h[j];
        double2 dd = make_double2(8.0 / (3.0 * h), G.kb);
        double2 idd = cinv(dd);
        AxisCoef& a = C.ax[j];
        a.dd = dd;
        a.bndd = cscale(idd, 8.0 / (h * h * h));
        a.eoff = csub(make_double2(4.0 / 3.0 / (h * h), 0.0), cscale(idd, 8.0 / 9.0 / (h * h * h)));
        a.cib = 8.0 / 3.0 / (h * h);
        a.mid = 1.0 / (h * h);
        a.w0e = -4.0 / (h * h);
        a.w0m = -2.0 / (h * h);
        a.a1 = -3.0 / h;
        a.a2 = 1.0 / (3.0 * h);
    }
    return C;
}

__device__ inline void interior_coords(const LeafGeom& G, int q, int g[3]) {
    g[0] = q % G.n[0] + 1;
    int r = q / G.n[0];
    if (G.d == 2) { g[1] = r + 1; g[2] = 1; }
    else { g[1] = r % G.n[1] + 1; g[2] = r / G.n[1] + 1; }
}

__device__ inline int stride_of(const LeafGeom& G, int j) {
    return j == 0 ? 1 : (j == 1 ? G.n[0] : G.n[0] * G.n[1]);
}

// Local boundary index (fgrid order) of the face point of face f adjacent to
// interior point with grid coords g.
__device__ inline int face_point(const LeafGeom& G, int f, const int g[3]) {
    int j = (f - 1) >> 1;
    int t;
    if (G.d == 2) {
        t = g[1 - j] - 1;
    } else {
        int a0 = (j == 0) ? 1 : 0;
        int a1 = (j == 2) ? 1 : 2;
        t = (g[a0] - 1) + G.n[a0] * (g[a1] - 1);
    }
    return G.foff[f] + t;
}

// Decompose a local boundary index into (face, first interior point p1, axis j).
__device__ inline void boundary_info(const LeafGeom& G, int b, int& f, int& j, int& q1, int& step) {
    // scalar coordinates and selects only (a runtime index into a small array or
    // into G.n would put it in scratch memory)
    f = 1;
    while (f < 2 * G.d && b >= G.foff[f] + G.pf[f]) ++f;
    const int t = b - G.foff[f];
    j = (f - 1) >> 1;
    const int n0 = G.n[0], n1 = G.n[1], n2 = G.n[2];
    int g0 = 1, g1 = 1, g2 = 1;
    if (G.d == 2) {
        if (j == 0) g1 = t + 1;
        else g0 = t + 1;
    } else {
        // tangential axes a0 < a1 of face axis j; t runs fastest along a0
        const int na0 = (j == 0) ? n1 : n0;
        const int v0 = t % na0 + 1, v1 = t / na0 + 1;
        if (j == 0) { g1 = v0; g2 = v1; }
        else if (j == 1) { g0 = v0; g2 = v1; }
        else { g0 = v0; g1 = v1; }
    }
    const int nj = (j == 0) ? n0 : (j == 1 ? n1 : n2);
    const int gj = (f & 1) ? 1 : nj;
    if (j == 0) g0 = gj;
    else if (j == 1) g1 = gj;
    else g2 = gj;
    q1 = (g0 - 1) + n0 * ((g1 - 1) + (G.d == 3 ? n1 * (g2 - 1) : 0));
    step = ((f & 1) ? 1 : -1) * (j == 0 ? 1 : (j == 1 ? n0 : n0 * n1));
}

__device__ inline int local_to_node(const LeafGeom& G, int parity, int b) {
    int f = 1;
    while (f < 2 * G.d && b >= G.foff[f] + G.pf[f]) ++f;
    return G.noff[parity][f] + (b - G.foff[f]);
}

__device__ inline int node_to_local(const LeafGeom& G, int parity, int nc) {
    for (int f = 1; f <= 2 * G.d; ++f)
        if (nc >= G.noff[parity][f] && nc < G.noff[parity][f] + G.pf[f]) return G.foff[f] + (nc - G.noff[parity][f]);
    return -1;
}

// Slow-axis couplings of S: lambda_k (layer k to k-1), mu_k (layer k to k+1).
__device__ inline double2 slow_lambda(const LeafGeom& G, const LeafCoef& C, int k) {
    const AxisCoef& a = C.ax[G.d - 1];
    return (k + 1 == G.n[G.d - 1]) ? a.eoff : make_double2(a.mid, 0.0);
}
__device__ inline double2 slow_mu(const LeafGeom& G, const LeafCoef& C, int k) {
    const AxisCoef& a = C.ax[G.d - 1];
    return (k == 0) ? a.eoff : make_double2(a.mid, 0.0);
}

// Cooperative in-place Gauss-Jordan inversion with partial pivoting of an
// s x s column-major matrix in LDS.  Pivot = first entry of maximal |re|+|im|
// in the column (the zgetrf rule).  Sets *bad on an exact zero pivot.
__device__ void gj_invert_lds(double2* M, int s, int* piv, double2* colj, int* bad) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int j = 0; j < s; ++j) {
        if (tid < 64) {
            double best = -1.0;
            int bi = s;
            for (int i = j + tid; i < s; i += 64) {
                double v = cabs1(M[i + j * s]);
                if (v > best) { best = v; bi = i; }
            }
            wave_argmax(best, bi);
            if (tid == 0) {
                piv[j] = bi;
                if (!(best > 0.0)) *bad = 1;
            }
        }
        __syncthreads();
        const int p = piv[j];
        if (p != j)
            for (int c = tid; c < s; c += nt) {
                double2 t = M[j + c * s]; M[j + c * s] = M[p + c * s]; M[p + c * s] = t;
            }
        __syncthreads();
        for (int i = tid; i < s; i += nt) colj[i] = M[i + j * s];
        __syncthreads();
        const double2 inv = cinv(colj[j]);
        for (int c = tid; c < s; c += nt) M[j + c * s] = (c == j) ? inv : cmul(M[j + c * s], inv);
        __syncthreads();
        for (int e = tid; e < s * s; e += nt) {
            int i = e % s, c = e / s;
            if (i == j) continue;
            double2 base = (c == j) ? make_double2(0.0, 0.0) : M[e];
            M[e] = csub(base, cmul(colj[i], M[j + c * s]));
        }
        __syncthreads();
    }
    // undo the row interchanges as column interchanges, last first
    for (int i = tid; i < s; i += nt)
        for (int j = s - 1; j >= 0; --j) {
            int p = piv[j];
            if (p != j) { double2 t = M[i + j * s]; M[i + j * s] = M[i + p * s]; M[i + p * s] = t; }
        }
    __syncthreads();
}

// One workgroup per leaf: Shat_k^-1 for k = 0..K-1, FP64.
__global__ void __launch_bounds__(256) leaf_blocks_kernel(LeafGeom G, LeafCoef C, const double2* __restrict__ kint,
                                                         double2* __restrict__ sinv, int* __restrict__ info) {
    extern __shared__ double2 sm[];
    const int s = G.s, K = G.K, ss = s * s;
    double2* M = sm;
    double2* Pv = sm + ss;
    double2* colj = sm + 2 * ss;
    int* piv = (int*)(colj + s);
    __shared__ int bad;
    const int leaf = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) bad = 0;
    const double2* kl = kint + (size_t)leaf * G.n_int;
    for (int k = 0; k < K; ++k) {
        double2 lm = make_double2(0.0, 0.0);
        if (k > 0) lm = cmul(slow_lambda(G, C, k), slow_mu(G, C, k - 1));
        for (int e = tid; e < ss; e += nt) {
            int r = e % s, c = e / s;
            int q = k * s + r;
            int g[3];
            interior_coords(G, q, g);
            double2 v = make_double2(0.0, 0.0);
            if (r == c) {
                v = kl[q];
                for (int j = 0; j < G.d; ++j) {
                    const AxisCoef& a = C.ax[j];
                    bool edge = (g[j] == 1 || g[j] == G.n[j]);
                    v.x += edge ? a.w0e : a.w0m;
                    if (g[j] == 1) v = cadd(v, a.bndd);
                    if (g[j] == G.n[j]) v = cadd(v, a.bndd);
                }
            } else {
                for (int j = 0; j < G.d - 1; ++j) {
                    const AxisCoef& a = C.ax[j];
                    int st = stride_of(G, j);
                    if (c == r + st && g[j] < G.n[j]) v = (g[j] == 1) ? a.eoff : make_double2(a.mid, 0.0);
                    if (c == r - st && g[j] > 1) v = (g[j] == G.n[j]) ? a.eoff : make_double2(a.mid, 0.0);
                }
            }
            if (k > 0) v = csub(v, cmul(lm, Pv[e]));
            M[e] = v;
        }
        __syncthreads();
        gj_invert_lds(M, s, piv, colj, &bad);
        double2* out = sinv + ((size_t)leaf * K + k) * ss;
        for (int e = tid; e < ss; e += nt) { out[e] = M[e]; Pv[e] = M[e]; }
        __syncthreads();
    }
    if (tid == 0 && bad) atomicOr(info, 1);
}

template <typename R> struct LeafIO {
    int mode;                 // LEAF_FACTOR_T, LEAF_UP, LEAF_DOWN
    int ncols;                // columns: nb (factor) or nrhs
    int CH;                   // columns per workgroup (generic kernel)
    int n_leaves;
    const c_t<R>* sinv;       // [leaf][K][s*s]
    const double2* const* f64;  // interior data f, complex128: column c at f64[c], [leaf][n_int] ...
    const float2* f32;        //   ... or complex64 [col][leaf][n_int] (at most one set; none = 0)
    const float2* jump;       // UP: subtracted from the outgoing data h: [col][leaf][nb] local order
    const c_t<R>* gin;        // DOWN: incoming data g: row b (node order) of leaf q in column c at
    const int* gsrc;          //   gin[gsrc[q * nb + b] + c * gld]
    size_t gld;
    c_t<R>* hout;             // UP: outgoing data h: [col][leaf][nb] node order
    double2* const* u64;      // DOWN: solution, complex128: column c at u64[c], [leaf][m] ...
    float2* u32;              //   ... or complex64 [col][leaf][m] (one set); written, never read
    // complex128 solution written as a correction (refinement): u64 = base + x with
    // base = ubase32 [col][leaf][m] (complex64) when given, else the current u64 when uacc
    const float2* ubase32;
    int uacc;
    double2* T;               // FACTOR: [leaf][nb][nb] node order, column-major
};

// f at interior point q of (col, leaf); u at point p of (col, leaf)
template <typename V, typename R>
__device__ inline V load_f(const LeafIO<R>& io, int col, int leaf, int ni, int q) {
    if (io.f64) return ccast<V>(gld(io.f64[col] + ((size_t)leaf * ni + q)));
    if (io.f32) return ccast<V>(io.f32[((size_t)col * io.n_leaves + leaf) * ni + q]);
    return V{};
}

template <typename V, typename R>
__device__ inline void store_u(const LeafIO<R>& io, int col, int leaf, int m, int p, V x) {
    const size_t i = (size_t)leaf * m + p;
    if (io.u64) {
        double2* u = io.u64[col] + i;
        if (io.ubase32) gst(u, cadd(ccast<double2>(io.ubase32[(size_t)col * io.n_leaves * m + i]), ccast<double2>(x)));
        else if (io.uacc) gst(u, cadd(gld(u), ccast<double2>(x)));
        else gst(u, ccast<double2>(x));
    } else {
        io.u32[(size_t)col * io.n_leaves * m + i] = ccast<float2>(x);
    }
}


// (B Vt)[w][c] over w2 < s with four partial sums: the loads of four steps are
// issued together instead of each step waiting for its two LDS reads before
// extending one dependent chain (the FP64 leaf maps of the 1024^2 factor:
// 2.54 -> 2.27 ms).  Loading the next layer's block into registers during the
// current layer's products measured no faster, with or without this.
template <typename V>
__device__ inline V layer_dot(const V* __restrict__ B, const V* __restrict__ Vt, int s, int CH, int w, int c) {
    V a0 = V{}, a1 = V{}, a2 = V{}, a3 = V{};
    int w2 = 0;
    for (; w2 + 4 <= s; w2 += 4) {
        a0 = cfma(B[w + w2 * s], Vt[w2 * CH + c], a0);
        a1 = cfma(B[w + (w2 + 1) * s], Vt[(w2 + 1) * CH + c], a1);
        a2 = cfma(B[w + (w2 + 2) * s], Vt[(w2 + 2) * CH + c], a2);
        a3 = cfma(B[w + (w2 + 3) * s], Vt[(w2 + 3) * CH + c], a3);
    }
    for (; w2 < s; ++w2) a0 = cfma(B[w + w2 * s], Vt[w2 * CH + c], a0);
    return cadd(cadd(a0, a1), cadd(a2, a3));
}

// A workgroup barrier that waits for the wave's LDS accesses only: the
// global loads a thread has in flight into its registers stay in flight
// (__syncthreads also waits for those).
__device__ inline void lds_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(15 | (3 << 14) | (7 << 4));   // lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// EPT > 0: each thread holds up to EPT elements of the next layer's block in
// registers, loaded while the current layer is computed (s^2 <= EPT x 256);
// the layer barriers then wait for LDS accesses only.  EPT = 0: each layer's
// block is read when the layer starts.  The arithmetic is the same.
template <typename R, int EPT>
__global__ void __launch_bounds__(256) leaf_apply_kernel(LeafGeom G, LeafCoef C, LeafIO<R> io) {
    using V = c_t<R>;
    extern __shared__ unsigned char smraw[];
    const int s = G.s, K = G.K, ss = s * s, CH = io.CH, ni = G.n_int, nb = G.nb;
    V* Y = (V*)smraw;                // [ni][CH]
    V* B = Y + (size_t)ni * CH;      // [s*s]
    V* Vt = B + ss;                  // [s][CH]
    const int leaf = blockIdx.x, c0 = blockIdx.y * CH, tid = threadIdx.x, nt = blockDim.x;
    const int ncol = min(CH, io.ncols - c0);
    const int parity = leaf & 1;
    const V* sl = io.sinv + (size_t)leaf * K * ss;

    // incoming data g for boundary point b, column c (chunk-local)
    auto gval = [&](int b, int c) -> V {
        int col = c0 + c;
        if (io.mode == LEAF_FACTOR_T) {
            return (node_to_local(G, parity, col) == b) ? mk<R>(R(1), R(0)) : cz<R>();
        } else if (io.mode == LEAF_DOWN) {
            return io.gin[(size_t)io.gsrc[(size_t)leaf * nb + local_to_node(G, parity, b)] + (size_t)col * io.gld];
        }
        return cz<R>();
    };

    // 1. right-hand side r = f - A_IB A_BB^-1 g
    for (int e = tid; e < ni * CH; e += nt) {
        int q = e / CH, c = e % CH;
        V v = cz<R>();
        if (c < ncol) {
            if (io.mode != LEAF_FACTOR_T) v = load_f<V>(io, c0 + c, leaf, ni, q);
            if (io.mode != LEAF_UP) {
                int g[3];
                interior_coords(G, q, g);
                for (int j = 0; j < G.d; ++j) {
                    const AxisCoef& a = C.ax[j];
                    double2 w = cscale(cinv(a.dd), -a.cib);
                    V wr = ccast<V>(w);
                    if (g[j] == 1) v = cfma(wr, gval(face_point(G, 2 * j + 1, g), c), v);
                    if (g[j] == G.n[j]) v = cfma(wr, gval(face_point(G, 2 * j + 2, g), c), v);
                }
            }
        }
        Y[e] = v;
    }
    __syncthreads();

    if constexpr (EPT > 0) {
        // the forward sweep's blocks 0..K-1 and the backward sweep's K-2..0 as
        // one stream of 2K - 1 layer steps (the last step loads its own block
        // again; every load is unconditional, with the index clamped)
        V pre[EPT];
        int pidx[EPT];
#pragma unroll
        for (int i = 0; i < EPT; ++i) pidx[i] = min(tid + i * 256, ss - 1);
        const int nsteps = 2 * K - 1;
#pragma unroll
        for (int i = 0; i < EPT; ++i) pre[i] = gld(sl + pidx[i]);
        for (int j = 0; j < nsteps; ++j) {
            const int k = j < K ? j : 2 * K - 2 - j;
#pragma unroll
            for (int i = 0; i < EPT; ++i)
                if (tid + i * 256 < ss) B[tid + i * 256] = pre[i];
            const int j1 = min(j + 1, nsteps - 1);
            const V* src = sl + (size_t)(j1 < K ? j1 : 2 * K - 2 - j1) * ss;
#pragma unroll
            for (int i = 0; i < EPT; ++i) pre[i] = gld(src + pidx[i]);
            if (j < K) {   // 2. forward: y_k = Shat_k^-1 (r_k - lambda_k y_{k-1})
                V lam = ccast<V>(slow_lambda(G, C, k));
                for (int e = tid; e < s * CH; e += nt) {
                    V v = Y[(size_t)k * s * CH + e];
                    if (k > 0) v = csub(v, cmul(lam, Y[(size_t)(k - 1) * s * CH + e]));
                    Vt[e] = v;
                }
                lds_barrier();
                for (int e = tid; e < s * CH; e += nt) {
                    int w = e / CH, c = e % CH;
                    Y[(size_t)k * s * CH + e] = layer_dot(B, Vt, s, CH, w, c);
                }
            } else {       // 3. backward: x_k = y_k - mu_k Shat_k^-1 x_{k+1}
                V mu = ccast<V>(slow_mu(G, C, k));
                for (int e = tid; e < s * CH; e += nt) Vt[e] = cmul(mu, Y[(size_t)(k + 1) * s * CH + e]);
                lds_barrier();
                for (int e = tid; e < s * CH; e += nt) {
                    int w = e / CH, c = e % CH;
                    Y[(size_t)k * s * CH + e] = csub(Y[(size_t)k * s * CH + e], layer_dot(B, Vt, s, CH, w, c));
                }
            }
            lds_barrier();
        }
    } else {
    // 2. forward sweep: y_k = Shat_k^-1 (r_k - lambda_k y_{k-1})
    for (int k = 0; k < K; ++k) {
        for (int e = tid; e < ss; e += nt) B[e] = sl[(size_t)k * ss + e];
        V lam = ccast<V>(slow_lambda(G, C, k));
        for (int e = tid; e < s * CH; e += nt) {
            V v = Y[(size_t)k * s * CH + e];
            if (k > 0) v = csub(v, cmul(lam, Y[(size_t)(k - 1) * s * CH + e]));
            Vt[e] = v;
        }
        __syncthreads();
        for (int e = tid; e < s * CH; e += nt) {
            int w = e / CH, c = e % CH;
            Y[(size_t)k * s * CH + e] = layer_dot(B, Vt, s, CH, w, c);
        }
        __syncthreads();
    }
    // 3. backward sweep: x_k = y_k - mu_k Shat_k^-1 x_{k+1}
    for (int k = K - 2; k >= 0; --k) {
        for (int e = tid; e < ss; e += nt) B[e] = sl[(size_t)k * ss + e];
        V mu = ccast<V>(slow_mu(G, C, k));
        for (int e = tid; e < s * CH; e += nt) Vt[e] = cmul(mu, Y[(size_t)(k + 1) * s * CH + e]);
        __syncthreads();
        for (int e = tid; e < s * CH; e += nt) {
            int w = e / CH, c = e % CH;
            Y[(size_t)k * s * CH + e] = csub(Y[(size_t)k * s * CH + e], layer_dot(B, Vt, s, CH, w, c));
        }
        __syncthreads();
    }
    }
    // 4. boundary values and outputs
    const double2 sma = make_double2(0.0, -2.0 * G.kb);  // sigma - alpha
    for (int e = tid; e < nb * CH; e += nt) {
        int b = e / CH, c = e % CH;
        if (c >= ncol) continue;
        int f, j, q1, step;
        boundary_info(G, b, f, j, q1, step);
        const AxisCoef& a = C.ax[j];
        V gb = gval(b, c);
        V x1 = Y[(size_t)q1 * CH + c], x2 = Y[(size_t)(q1 + step) * CH + c];
        // u_B = (g - a1 x1 - a2 x2) / dd
        V num = csub(csub(gb, cscale(x1, (R)a.a1)), cscale(x2, (R)a.a2));
        V uB = cmul(num, ccast<V>(cinv(a.dd)));
        V hb = cfma(ccast<V>(sma), uB, gb);
        int col = c0 + c;
        if (io.mode == LEAF_FACTOR_T) {
            io.T[(size_t)leaf * nb * nb + (size_t)col * nb + local_to_node(G, parity, b)] = ccast<double2>(hb);
        } else if (io.mode == LEAF_UP) {
            if (io.jump) hb = csub(hb, ccast<V>(io.jump[((size_t)col * io.n_leaves + leaf) * nb + b]));
            io.hout[((size_t)col * io.n_leaves + leaf) * nb + local_to_node(G, parity, b)] = hb;
        } else {
            store_u(io, col, leaf, G.m, b, uB);
        }
    }
    if (io.mode == LEAF_DOWN) {
        for (int e = tid; e < ni * CH; e += nt) {
            int q = e / CH, c = e % CH;
            if (c >= ncol) continue;
            store_u(io, c0 + c, leaf, G.m, nb + q, Y[e]);
        }
    }
}

// Few right-hand sides (at most CW per workgroup, FP64, 2D layers of at most
// 16 points): one wave per leaf.  The block layer recursion of
// leaf_apply_kernel is latency-bound with one column -- a 256-thread
// workgroup per leaf waits for each layer's block (4 KB) before its 16 dot
// products -- so here each lane keeps its share of the next PF layers' blocks
// in flight in registers: lane (w, q) = (lane % 16, lane / 16) holds row w of
// the block at columns w2 = q, q + 4, ... (and the columns past the last
// multiple of four for q = 0), which is the partial sum a_q of layer_dot.  The
// four partial sums meet in LDS and are added in layer_dot's order, and every
// other step is leaf_apply_kernel's own arithmetic, so a column gets the same
// bits from either kernel.  The forward sweep's blocks 0..K-1 and the
// backward sweep's K-2..0 form one stream of 2K - 1 steps.  TAIL: s is not a
// multiple of four (the lanes q = 0 also hold the columns past the last one).
template <int CW, bool TAIL>
__global__ void __launch_bounds__(64) leaf_wave_kernel(LeafGeom G, LeafCoef C, LeafIO<double> io) {
    using V = double2;
    constexpr int PF = 4;   // steps of blocks in flight
    extern __shared__ unsigned char smraw[];
    const int s = G.s, K = G.K, ss = s * s, ni = G.n_int, nb = G.nb;
    V* Y = (V*)smraw;                // [ni][CW]
    V* Vt = Y + (size_t)ni * CW;     // [s][CW]
    V* Pp = Vt + 16 * CW;            // [4][16][CW] partial sums
    const int leaf = blockIdx.x, c0 = blockIdx.y * CW, lane = threadIdx.x;
    const int ncol = min(CW, io.ncols - c0);
    const int parity = leaf & 1;
    const V* sl = io.sinv + (size_t)leaf * K * ss;
    const int w = lane & 15, q = lane >> 4;
    const int s4 = s & ~3;
    const int nsteps = 2 * K - 1;

    auto gval = [&](int b, int c) -> V {
        const int col = c0 + c;
        if (io.mode == LEAF_DOWN)
            return io.gin[(size_t)io.gsrc[(size_t)leaf * nb + local_to_node(G, parity, b)] + (size_t)col * io.gld];
        return cz<double>();
    };
    // step j's block: the forward sweep's j, then the backward sweep's 2K - 2 - j
    auto blk = [&](int j) { return j < K ? j : 2 * K - 2 - j; };
    // An entry outside the block (w2 past the last multiple of four, or a
    // column past s) loads column 0 and is never read: step() skips it by the
    // same conditions.  Zeroing it here instead would make every load wait at
    // once for its data and leave no block in flight.
    V br[PF][4], bt[PF][3];
    auto load = [&](int j, V (&r)[4], V (&t)[3]) {
        const V* Bk = sl + (size_t)blk(min(j, nsteps - 1)) * ss;
        const int wl = w < s ? w : 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int w2 = q + 4 * u;
            r[u] = gld(Bk + wl + (size_t)(w2 < s4 ? w2 : 0) * s);
        }
        if constexpr (TAIL) {
#pragma unroll
            for (int u = 0; u < 3; ++u) {
                const int w2 = s4 + u;
                const bool ok = q == 0 && w2 < s;
                t[u] = gld(Bk + wl + (size_t)(ok ? w2 : 0) * s);
            }
        }
    };
#pragma unroll
    for (int p = 0; p < PF; ++p) load(p, br[p], bt[p]);

    // 1. right-hand side r = f - A_IB A_BB^-1 g (leaf_apply_kernel step 1)
    for (int e = lane; e < ni * CW; e += 64) {
        const int qp = e / CW, c = e % CW;
        V v = cz<double>();
        if (c < ncol) {
            v = load_f<V>(io, c0 + c, leaf, ni, qp);
            if (io.mode != LEAF_UP) {
                int g[3];
                interior_coords(G, qp, g);
                for (int j = 0; j < G.d; ++j) {
                    const AxisCoef& a = C.ax[j];
                    double2 wgt = cscale(cinv(a.dd), -a.cib);
                    V wr = ccast<V>(wgt);
                    if (g[j] == 1) v = cfma(wr, gval(face_point(G, 2 * j + 1, g), c), v);
                    if (g[j] == G.n[j]) v = cfma(wr, gval(face_point(G, 2 * j + 2, g), c), v);
                }
            }
        }
        Y[e] = v;
    }
    __syncthreads();
    // the first step's input: layer 0 of r
    for (int e = lane; e < s * CW; e += 64) Vt[e] = Y[e];
    __syncthreads();

    // 2./3. the 2K - 1 block steps
    auto step = [&](int j, V (&r)[4], V (&t)[3]) {
        // partial sums a_q of row w (layer_dot's chains)
        if (w < s) {
#pragma unroll
            for (int c = 0; c < CW; ++c) {
                if (c >= ncol) break;
                V a = cz<double>();
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (q + 4 * u < s4) a = cfma(r[u], Vt[(q + 4 * u) * CW + c], a);
                if (TAIL && q == 0) {
#pragma unroll
                    for (int u = 0; u < 3; ++u)
                        if (s4 + u < s) a = cfma(t[u], Vt[(s4 + u) * CW + c], a);
                }
                Pp[(q * 16 + w) * CW + c] = a;
            }
        }
        __syncthreads();
        // the layer's result, and the next step's input
        const int kb = blk(j);
        for (int e = lane; e < s * CW; e += 64) {
            const int wr = e / CW, c = e % CW;
            const V a0 = Pp[(0 * 16 + wr) * CW + c], a1 = Pp[(1 * 16 + wr) * CW + c];
            const V a2 = Pp[(2 * 16 + wr) * CW + c], a3 = Pp[(3 * 16 + wr) * CW + c];
            const V res = cadd(cadd(a0, a1), cadd(a2, a3));
            V* yk = Y + (size_t)kb * s * CW + e;
            const V y = j < K ? res : csub(*yk, res);
            *yk = y;
            if (j + 1 < K) {          // forward: r_{k+1} - lambda_{k+1} y_k
                const V lam = slow_lambda(G, C, j + 1);
                Vt[e] = csub(Y[(size_t)(j + 1) * s * CW + e], cmul(lam, y));
            } else if (j + 1 < nsteps) {   // backward: mu_k x_{k+1}, k = blk(j + 1) = kb - 1
                const V mu = slow_mu(G, C, kb - 1);
                Vt[e] = cmul(mu, y);
            }
        }
        __syncthreads();
    };
    for (int j0 = 0; j0 < nsteps; j0 += PF) {
#pragma unroll
        for (int p = 0; p < PF; ++p) {
            const int j = j0 + p;
            if (j < nsteps) step(j, br[p], bt[p]);
            load(j + PF, br[p], bt[p]);   // past the last step: the last block again (never read)
        }
    }

    // 4. boundary values and outputs (leaf_apply_kernel step 4)
    const double2 sma = make_double2(0.0, -2.0 * G.kb);  // sigma - alpha
    for (int e = lane; e < nb * CW; e += 64) {
        const int b = e / CW, c = e % CW;
        if (c >= ncol) continue;
        int f, j, q1, stp;
        boundary_info(G, b, f, j, q1, stp);
        const AxisCoef& a = C.ax[j];
        V gb = gval(b, c);
        V x1 = Y[(size_t)q1 * CW + c], x2 = Y[(size_t)(q1 + stp) * CW + c];
        V num = csub(csub(gb, cscale(x1, a.a1)), cscale(x2, a.a2));
        V uB = cmul(num, ccast<V>(cinv(a.dd)));
        V hb = cfma(ccast<V>(sma), uB, gb);
        const int col = c0 + c;
        if (io.mode == LEAF_UP) {
            if (io.jump) hb = csub(hb, ccast<V>(io.jump[((size_t)col * io.n_leaves + leaf) * nb + b]));
            io.hout[((size_t)col * io.n_leaves + leaf) * nb + local_to_node(G, parity, b)] = hb;
        } else {
            store_u(io, col, leaf, G.m, b, uB);
        }
    }
    if (io.mode == LEAF_DOWN) {
        for (int e = lane; e < ni * CW; e += 64) {
            const int qp = e / CW, c = e % CW;
            if (c >= ncol) continue;
            store_u(io, c0 + c, leaf, G.m, nb + qp, Y[e]);
        }
    }
}

// FP64 residual of the global FD system at the current leaf solutions u:
//   interior rows       rf = f - (Delta_h + k) u                     (fd_module.f90:140-181)
//   physical rows       rg = g - (alpha + D_n) u                     (fd_module.f90:187-216)
//   transmission rows   rj = zeta g_l + nu g_nb - h_l,  g = (alpha + D_n) u, h = (sigma + D_n) u
// The transmission condition is the pointwise form of the merge system
// M [g1; g2] = [h1; h2] (robin_tree_module.f90:425-467): it holds at every pair
// of coincident leaf boundary points once the tree has solved.  A correction
// solve takes (rf, rg) as (f, g) and subtracts rj from the leaf outgoing data.
// Evaluated in FP64; stored in complex64, the precision the correction pass
// computes in (rf [col][leaf][n_int], rg and rj [col][leaf][n_bnd]; rg is
// written on the physical boundary only).
template <typename U>
__device__ inline const U* ucolumn(double2* const* u64, const float2* u32, int col, int nl, int m);
template <>
__device__ inline const double2* ucolumn<double2>(double2* const* u64, const float2*, int col, int, int) {
    return u64[col];
}
template <>
__device__ inline const float2* ucolumn<float2>(double2* const*, const float2* u32, int col, int nl, int m) {
    return u32 + (size_t)col * nl * m;
}

// One thread per (leaf, point, RES_CPT columns); the loads of a thread are issued
// together.  More columns per thread share the geometry but lower occupancy:
// at 1024^2 with 256 right-hand sides 1, 2, 4, 8 measured 2.82, 2.86, 3.11, 3.65 ms.
#ifndef RS_RES_CPT
#define RS_RES_CPT 1
#endif
constexpr int RES_CPT = RS_RES_CPT;
// Column groups in the grid: each thread loops over its share of the columns
// with the point's geometry and k loaded once (k was otherwise fetched again
// for every column).  At 1024^2 with 256 right-hand sides: 256 groups (one
// column per thread) 2.85 ms, 32 groups 2.39, 16 groups 2.44, 8 groups 2.48,
// 4 groups 2.45, 1 group 2.99; two or four columns per loop step (8 or 16
// groups) 2.63-3.14.  (Both constants can be set when compiling, for such
// comparisons.)
#ifndef RS_RES_GROUPS
#define RS_RES_GROUPS 32
#endif
constexpr int RES_GROUPS = RS_RES_GROUPS;

template <typename U>
__global__ void __launch_bounds__(256) residual_kernel(LeafGeom G, LeafCoef C, double2* const* __restrict__ u64,
                                                       const float2* __restrict__ u32,
                                                       const double2* __restrict__ kint,
                                                       const double2* const* __restrict__ f,
                                                       const double2* const* __restrict__ g,
                                                       const int* __restrict__ nbr, double2 nu, double2 zeta,
                                                       float2* __restrict__ rf, float2* __restrict__ rg,
                                                       float2* __restrict__ rj, int n_leaves, int ncols) {
    const int m = G.m, nb = G.nb, ni = G.n_int;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n_leaves * m) return;
    const int leaf = (int)(e / m), p = (int)(e - (size_t)leaf * m);
    // the thread's columns: col0 = blockIdx.y * RES_CPT + k * gridDim.y * RES_CPT
    const int cstep = gridDim.y * RES_CPT;
    const U* ucols[RES_CPT];
    bool cok[RES_CPT];
    auto columns = [&](int col0) {
#pragma unroll
        for (int c = 0; c < RES_CPT; ++c) {
            cok[c] = col0 + c < ncols;
            ucols[c] = ucolumn<U>(u64, u32, cok[c] ? col0 + c : ncols - 1, n_leaves, m);
        }
    };
    const size_t ul0 = (size_t)leaf * m;
    if (p < nb) {
        int fc, j_, q1, st;
        boundary_info(G, p, fc, j_, q1, st);
        const int nl2 = nbr[(size_t)leaf * 7 + fc];
        const int f2 = (fc & 1) ? fc + 1 : fc - 1;
        const int b2 = G.foff[f2] + (p - G.foff[fc]);
        int ff, jj, q1n, stn;
        boundary_info(G, b2, ff, jj, q1n, stn);
        const size_t un0 = (size_t)(nl2 >= 0 ? nl2 : leaf) * m;
        const AxisCoef& A = C.ax[j_];
        const AxisCoef& B = C.ax[jj];
        const double2 sma = make_double2(0.0, -2.0 * G.kb);
        for (int col0 = blockIdx.y * RES_CPT; col0 < ncols; col0 += cstep) {
        columns(col0);
        U a0[RES_CPT], a1[RES_CPT], a2[RES_CPT], n0[RES_CPT], n1[RES_CPT], n2[RES_CPT];
        double2 gv[RES_CPT];
#pragma unroll
        for (int c = 0; c < RES_CPT; ++c) {
            const U* uc = ucols[c];
            a0[c] = gld(uc + (ul0 + p));
            a1[c] = gld(uc + (ul0 + nb + q1));
            a2[c] = gld(uc + (ul0 + nb + q1 + st));
            n0[c] = gld(uc + (un0 + b2));
            n1[c] = gld(uc + (un0 + nb + q1n));
            n2[c] = gld(uc + (un0 + nb + q1n + stn));
            // g only matters on the physical boundary; elsewhere every lane loads the
            // column's first element (one cache line) and the value is unused
            gv[c] = gld(g[cok[c] ? col0 + c : ncols - 1] + (nl2 < 0 ? (size_t)leaf * nb + p : 0));
        }
#pragma unroll
        for (int c = 0; c < RES_CPT; ++c) {
            if (!cok[c]) continue;
            const size_t o = ((size_t)(col0 + c) * n_leaves + leaf) * nb + p;
            const double2 up = ccast<double2>(a0[c]);
            double2 gA = cmul(A.dd, up);
            gA = cadd(gA, cscale(ccast<double2>(a1[c]), A.a1));
            gA = cadd(gA, cscale(ccast<double2>(a2[c]), A.a2));
            // (rg is read on the physical boundary only: the root gather's points)
            if (nl2 < 0) {
                rg[o] = ccast<float2>(csub(gv[c], gA));
                rj[o] = make_float2(0.f, 0.f);
            } else {
                double2 gN = cmul(B.dd, ccast<double2>(n0[c]));
                gN = cadd(gN, cscale(ccast<double2>(n1[c]), B.a1));
                gN = cadd(gN, cscale(ccast<double2>(n2[c]), B.a2));
                const double2 hA = cfma(sma, up, gA);
                rj[o] = ccast<float2>(csub(cadd(cmul(zeta, gA), cmul(nu, gN)), hA));
            }
        }
        }
        return;
    }
    // interior row: k u + sum over axes of the three-point second differences
    const int q = p - nb;
    int gq[3];
    interior_coords(G, q, gq);
    int im[3], ip[3];
    double wm[3], w0[3], wp[3];
#pragma unroll
    for (int jx = 0; jx < 3; ++jx) {
        im[jx] = ip[jx] = p;
        wm[jx] = w0[jx] = wp[jx] = 0.0;
        if (jx >= G.d) continue;
        const int stx = jx == 0 ? 1 : (jx == 1 ? G.n[0] : G.n[0] * G.n[1]);
        const int nj = jx == 0 ? G.n[0] : (jx == 1 ? G.n[1] : G.n[2]);
        const double ih2 = C.ax[jx].mid;
        if (gq[jx] == 1) { wm[jx] = 8.0 / 3.0; w0[jx] = -4.0; wp[jx] = 4.0 / 3.0; }
        else if (gq[jx] == nj) { wm[jx] = 4.0 / 3.0; w0[jx] = -4.0; wp[jx] = 8.0 / 3.0; }
        else { wm[jx] = 1.0; w0[jx] = -2.0; wp[jx] = 1.0; }
        wm[jx] *= ih2; w0[jx] *= ih2; wp[jx] *= ih2;
        im[jx] = (gq[jx] == 1) ? face_point(G, 2 * jx + 1, gq) : nb + q - stx;
        ip[jx] = (gq[jx] == nj) ? face_point(G, 2 * jx + 2, gq) : nb + q + stx;
    }
    const double2 kq = kint[(size_t)leaf * ni + q];
    for (int col0 = blockIdx.y * RES_CPT; col0 < ncols; col0 += cstep) {
    columns(col0);
    U uc0[RES_CPT], um[RES_CPT][3], upp[RES_CPT][3];
    double2 fv[RES_CPT];
#pragma unroll
    for (int c = 0; c < RES_CPT; ++c) {
        const U* uc = ucols[c] + ul0;
        uc0[c] = gld(uc + p);
#pragma unroll
        for (int jx = 0; jx < 3; ++jx) {
            um[c][jx] = gld(uc + im[jx]);
            upp[c][jx] = gld(uc + ip[jx]);
        }
        fv[c] = gld(f[cok[c] ? col0 + c : ncols - 1] + ((size_t)leaf * ni + q));
    }
#pragma unroll
    for (int c = 0; c < RES_CPT; ++c) {
        if (!cok[c]) continue;
        const double2 u0 = ccast<double2>(uc0[c]);
        double2 Au = cmul(kq, u0);
#pragma unroll
        for (int jx = 0; jx < 3; ++jx) {
            if (jx >= G.d) break;
            Au = cadd(Au, cscale(u0, w0[jx]));
            Au = cadd(Au, cscale(ccast<double2>(um[c][jx]), wm[jx]));
            Au = cadd(Au, cscale(ccast<double2>(upp[c][jx]), wp[jx]));
        }
        rf[((size_t)(col0 + c) * n_leaves + leaf) * ni + q] = ccast<float2>(csub(fv[c], Au));
    }
    }
}

// ---------------------------------------------------------------------------
// Multi-RHS leaf solve on the matrix cores (complex64, square 2D leaves of
// s = K = 12, 14 or 16 points per axis: K layers of s points).  Each wave owns 16 RHS
// columns; a workgroup covers 64 columns of one leaf.  The layer recursion
//   forward   y_k = Shat_k^-1 (r_k - lambda_k y_{k-1})
//   backward  x_k = y_k - mu_k Shat_k^-1 x_{k+1}
// is one 16x16x16 complex product per layer (v_mfma_f32_16x16x4f32, three real
// products per complex k-chunk).  The leaf's K blocks are staged once in LDS;
// y_k stays in registers in the MFMA output layout (lane l holds rows
// 4(l/16)..4(l/16)+3 of column l%16), and between layers the 16x16 operand
// goes through a small per-wave LDS tile to reach the B-operand layout.  The
// boundary values need x at the first two interior points along the normal:
// for x-faces those rows sit in one lane (s even), for y-faces they are
// layers 0,1 / K-2,K-1 of the same lane, so no further exchange is needed.
typedef float f4v __attribute__((ext_vector_type(4)));

// a[i] for a runtime i without indexing the register array (which would move it to scratch)
__device__ inline float2 pick4(float2 a0, float2 a1, float2 a2, float2 a3, int i) {
    return i == 0 ? a0 : (i == 1 ? a1 : (i == 2 ? a2 : a3));
}

// LDS exchange among the lanes of one wave.  A wave's LDS instructions are
// processed in order, so a write by one lane is seen by a later read of another
// lane of the same wave; only the compiler must not move LDS accesses across.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// f element type of a leaf pass: the caller's complex128 data through the
// column table, or the complex64 correction data (contiguous [col][leaf][n_int]).
template <typename FT> struct FCol;
template <> struct FCol<double2> {
    static __device__ const double2* base(const LeafIO<float>& io, int col, int leaf, int ni, int) {
        return io.f64[col] + (size_t)leaf * ni;
    }
};
template <> struct FCol<float2> {
    static __device__ const float2* base(const LeafIO<float>& io, int col, int leaf, int ni, int nl) {
        return io.f32 + ((size_t)col * nl + leaf) * ni;
    }
};

// Every global load in this kernel is unconditional (indices are clamped into
// range and out-of-range lanes select zero afterwards), so that loads issued
// ahead stay in flight: a load under a branch makes the compiler wait for
// all outstanding loads where the branches join.
// Output form of the down pass: the complex64 first iterate (contiguous), or
// the complex128 solution through the column table, written plainly, as
// u = base64 + x with a complex64 base, or as u += x.
enum { OUT_U32 = 0, OUT_U64 = 1, OUT_U64_BASE32 = 2, OUT_U64_ACC = 3 };

template <int K, int MODE, typename FT, int OUT>
__global__ void __launch_bounds__(256) leaf_mfma_kernel(LeafGeom G, LeafCoef C, LeafIO<float> io, int nchunks) {
    extern __shared__ float2 lds2[];
    // square leaves: K layers of s = K points (the launcher checks the geometry)
    constexpr int s = K, ni = K * K, nb = 4 * K, m = ni + nb;
    const int nl = io.n_leaves;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // LDS: the K blocks [K][16 x 16] column-major; per wave one 16 x 16 tile,
    // row-major with row stride LT (= 17: rows fall in different banks), and the
    // boundary data of its 16 columns, column stride BS = nb + 1.  The tile is
    // reused from one step to the next without a second copy: a wave's LDS
    // accesses are processed in order.
    constexpr int LT = 17, LTS = 16 * LT;
    constexpr int BS = nb + 1;
    float2* Bl = lds2;
    float2* scr = lds2 + K * 256 + wave * LTS;
    float2* bd = lds2 + K * 256 + 4 * LTS + wave * 16 * BS;
    const int leaf = blockIdx.x / nchunks, chunk = blockIdx.x - leaf * nchunks;
    const int ccol = lane & 15, grp = lane >> 4;
    const int mycol = chunk * 64 + wave * 16 + ccol;
    const bool colok = mycol < io.ncols;
    const int parity = leaf & 1;
    const size_t lc = (size_t)(colok ? mycol : 0) * nl + leaf;
    // The faces' offsets for this leaf's parity, read from the kernel arguments
    // once: a run-time index into them is a memory read, and such reads inside
    // the layer sweeps each came with a wait on the LDS counter as well.
    // Plain variables and arithmetic, not arrays or selection chains: a choice
    // among several such values can be compiled into an indexed load from a
    // table in scratch memory.
    // fe_f: end of face f in local order; fn_f: its start in node order.
    const int fe1 = G.foff[1] + G.pf[1], fe2 = G.foff[2] + G.pf[2], fe3 = G.foff[3] + G.pf[3];
    const int fn1 = G.noff[parity][1], fn2 = G.noff[parity][2], fn3 = G.noff[parity][3],
              fn4 = G.noff[parity][4];
    const int fs1 = fn1 - G.foff[1], fs2 = fn2 - G.foff[2], fs3 = fn3 - G.foff[3], fs4 = fn4 - G.foff[4];
    // local_to_node for the four faces of a 2D leaf (faces are consecutive in
    // local order, so fe1 <= fe2 <= fe3)
    auto l2n = [&](int b) {
        return b + fs1 + (b >= fe1 ? fs2 - fs1 : 0) + (b >= fe2 ? fs3 - fs2 : 0) + (b >= fe3 ? fs4 - fs3 : 0);
    };
    // x faces: rows 0, 1 (face 1) sit in the lanes of group 0 and rows s-1, s-2
    // (face 2) in those of group (s-1)/4, a different group (s >= 8): a lane
    // handles at most one x face
    const bool xf1 = grp == 0, xf2 = grp == (s - 1) / 4;
    const int fnx = xf1 ? fn1 : fn2;
    // Global loads and stores of interior data use a coalesced mapping: lane l
    // handles row rr = l % 16 of the wave's columns cc(i) = 4 i + l / 16, so 16
    // lanes cover 16 consecutive points of one column; tiles pass through the
    // wave's LDS tile to and from the MFMA layout.
    const int wcol0 = chunk * 64 + wave * 16;
    const int rr = lane & 15, cq = lane >> 4;
    const int rrc = rr < s ? rr : s - 1;
    // Lanes outside the leaf (rr >= s) or the batch load a valid point instead
    // (clamped indices) and their values are not masked: the rows of the layer
    // blocks beyond s are zero, so rows of the operand beyond s add nothing to
    // the product, and columns never mix, so a column beyond the batch only
    // computes values that are not stored.
    const FT* fcp[4];
    bool fok[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int col = wcol0 + 4 * i + cq;
        fok[i] = col < io.ncols && rr < s;
        fcp[i] = FCol<FT>::base(io, col < io.ncols ? col : 0, leaf, ni, nl);
    }
    // this lane's boundary point in the per-column loops over boundary data
    // (nb = 2 s + 2 K <= 64 for this kernel, so one point per lane)
    const bool bok = lane < nb;
    const int bi = bok ? lane : nb - 1;

    // the leaf's K block inverses into LDS (zero-padded to 16 x 16)
    {
        const float2* sl = io.sinv + (size_t)leaf * K * s * s;
        const int r = tid & 15, c = tid >> 4;
        const bool ok = r < s && c < s;
        const int off = (r < s ? r : 0) + (c < s ? c : 0) * s;
        float2 t[K];
#pragma unroll
        for (int k = 0; k < K; ++k) t[k] = sl[(size_t)k * s * s + off];
#pragma unroll
        for (int k = 0; k < K; ++k) Bl[k * 256 + tid] = ok ? t[k] : make_float2(0.f, 0.f);
    }
    // boundary data of the wave's 16 columns, read once with coalesced loads:
    // DOWN: incoming g (node order); UP: the transmission residual (local order);
    // FACTOR_T: column c is the unit vector of node-order boundary row c
    {
        if (MODE == LEAF_FACTOR_T) {
#pragma unroll
            for (int c = 0; c < 16; ++c)
                if (bok) bd[c * BS + bi] = (bi == wcol0 + c) ? make_float2(1.f, 0.f) : make_float2(0.f, 0.f);
        } else if (MODE == LEAF_DOWN || io.jump) {
            // stored in node order (g already is; the residual is in local order)
            float2 t[16];
            const size_t gs = (MODE == LEAF_DOWN) ? (size_t)io.gsrc[(size_t)leaf * nb + bi] : 0;
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const int col = (wcol0 + c) < io.ncols ? wcol0 + c : 0;
                if (MODE == LEAF_DOWN) t[c] = io.gin[gs + (size_t)col * io.gld];
                else t[c] = ccast<float2>(io.jump[((size_t)col * nl + leaf) * nb + bi]);
            }
            const int pos = (MODE == LEAF_DOWN) ? bi : l2n(bi);
#pragma unroll
            for (int c = 0; c < 16; ++c)
                if (bok) bd[c * BS + pos] = ((wcol0 + c) < io.ncols) ? t[c] : make_float2(0.f, 0.f);
        } else {
#pragma unroll
            for (int c = 0; c < 16; ++c)
                if (bok) bd[c * BS + bi] = make_float2(0.f, 0.f);
        }
    }
    __syncthreads();

    const AxisCoef& ax0 = C.ax[0];
    const AxisCoef& ax1 = C.ax[1];
    const float2 cbx = ccast<float2>(cscale(cinv(ax0.dd), -ax0.cib));
    const float2 cby = ccast<float2>(cscale(cinv(ax1.dd), -ax1.cib));
    // incoming g at point t of a face whose node offset is fn
    auto gval = [&](int fn, int t) -> float2 { return bd[ccol * BS + fn + t]; };
    // slow-axis couplings (axis 1, K layers): lambda_k couples layer k to k-1,
    // mu_k layer k to k+1; the end value applies at the last / first layer
    // (the inner layers' value is real: a scaling, not a complex product)
    const float cmid = (float)ax1.mid;
    const float2 cend = ccast<float2>(ax1.eoff);
    auto lam_times = [&](int k, float2 x) { return (k + 1 == K) ? cmul(cend, x) : cscale(x, cmid); };
    auto mu_times = [&](int k, float2 x) { return (k == 0) ? cmul(cend, x) : cscale(x, cmid); };

    float2 y[K][4];
    float2 prev[4];
    auto product = [&](int k, const float2 v[4], f4v& aR, f4v& aI) {
        float2* sb = scr;
        for (int vv = 0; vv < 4; ++vv) sb[(4 * grp + vv) * LT + ccol] = v[vv];
        wave_lds_sync();
        // three real products per complex k-chunk (Gauss): three independent
        // accumulation chains of four instead of two of eight
        f4v p1 = f4v{0.f, 0.f, 0.f, 0.f}, p2 = p1, p3 = p1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float2 a = Bl[k * 256 + ccol + 16 * (4 * q + grp)];
            const float2 b = sb[(4 * q + grp) * LT + ccol];
            p1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, p1, 0, 0, 0);
            p2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.y, p2, 0, 0, 0);
            p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x + a.y, b.x + b.y, p3, 0, 0, 0);
        }
        aR = p1 - p2;
        aI = (p3 - p1) - p2;
    };

    // raw[i] = f at row rr of layer k, column cc(i) (coalesced mapping)
    auto fetch = [&](int k, FT (&raw)[4]) {
#pragma unroll
        for (int i = 0; i < 4; ++i) raw[i] = (MODE == LEAF_FACTOR_T) ? FT{} : gld(fcp[i] + (k * s + rrc));
    };
    // the interior data runs PF layers ahead of the sweep (memory latency is
    // several layer steps long)
    constexpr int PF = 4;
    FT fbuf[PF][4];
#pragma unroll
    for (int k = 0; k < PF && k < K; ++k) fetch(k, fbuf[k]);
    // forward sweep
#pragma unroll
    for (int k = 0; k < K; ++k) {
        // the layer's f tile into the wave's LDS tile, then read back in the MFMA layout
        {
            float2* sb = scr;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                sb[rr * LT + 4 * i + cq] = ccast<float2>(fbuf[k % PF][i]);
            }
            wave_lds_sync();
        }
        if (k + PF < K) fetch(k + PF, fbuf[k % PF]);
        // incoming data of the boundary rows (down pass), added by selects:
        // the x face of the lane's group at row 0 / s-1 of every layer (one LDS
        // read), the y faces at every row of the first and last layer.  Columns
        // beyond the batch hold zeros here and are not stored anyway.
        constexpr bool GIN = MODE == LEAF_DOWN || MODE == LEAF_FACTOR_T;
        const float2 gx = GIN ? gval(fnx, k) : make_float2(0.f, 0.f);
        float2 v[4];
#pragma unroll
        for (int vv = 0; vv < 4; ++vv) {
            const int row = 4 * grp + vv;
            float2 r = scr[row * LT + ccol];
            if (GIN) {
                if (vv == 0) r = xf1 ? cfma(cbx, gx, r) : r;
                if (vv == ((s - 1) & 3)) r = xf2 ? cfma(cbx, gx, r) : r;
                if (k == 0 || k == K - 1) {
                    const float2 gy = gval(k == 0 ? fn3 : fn4, row < s ? row : s - 1);
                    r = row < s ? cfma(cby, gy, r) : r;
                }
            }
            if (k > 0) r = csub(r, lam_times(k, prev[vv]));
            v[vv] = r;
        }
        f4v aR, aI;
        product(k, v, aR, aI);
#pragma unroll
        for (int vv = 0; vv < 4; ++vv) { y[k][vv] = make_float2(aR[vv], aI[vv]); prev[vv] = y[k][vv]; }
    }
    // backward sweep
#pragma unroll
    for (int k = K - 2; k >= 0; --k) {
        float2 w[4];
#pragma unroll
        for (int vv = 0; vv < 4; ++vv) w[vv] = mu_times(k, y[k + 1][vv]);
        f4v aR, aI;
        product(k, w, aR, aI);
#pragma unroll
        for (int vv = 0; vv < 4; ++vv) y[k][vv] = csub(y[k][vv], make_float2(aR[vv], aI[vv]));
    }

    // store of solution point p of column col; base: the complex64 iterate for OUT_U64_BASE32
    auto put_col = [&](int col, double2* u64c, int p, float2 x, float2 base) {
        if (OUT == OUT_U32) io.u32[((size_t)col * nl + leaf) * m + p] = x;
        else if (OUT == OUT_U64) gst(u64c + p, ccast<double2>(x));
        else if (OUT == OUT_U64_BASE32) gst(u64c + p, cadd(ccast<double2>(base), ccast<double2>(x)));
        else gst(u64c + p, cadd(gld(u64c + p), ccast<double2>(x)));
    };
    if (MODE == LEAF_DOWN) {
        // interior solution: each layer's tile through LDS, stored 16 points per column at a time
        double2* ucp[4];
        const float2* bcp[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int col = wcol0 + 4 * i + cq;
            const int colc = col < io.ncols ? col : 0;
            ucp[i] = (OUT != OUT_U32) ? io.u64[colc] + (size_t)leaf * m : nullptr;
            bcp[i] = (OUT == OUT_U64_BASE32) ? io.ubase32 + ((size_t)colc * nl + leaf) * m : nullptr;
        }
        // the complex64 base runs two layers ahead of the stores
        float2 bbuf[2][4];
        auto bfetch = [&](int k, float2 (&bb)[4]) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                bb[i] = (OUT == OUT_U64_BASE32) ? bcp[i][nb + k * s + rrc] : make_float2(0.f, 0.f);
        };
        bfetch(0, bbuf[0]);
        if (K > 1) bfetch(1, bbuf[1]);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            float2 base[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) base[i] = bbuf[k & 1][i];
            if (k + 2 < K) bfetch(k + 2, bbuf[k & 1]);
            float2* sb = scr;
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) sb[(4 * grp + vv) * LT + ccol] = y[k][vv];
            wave_lds_sync();
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int col = wcol0 + 4 * i + cq;
                if (fok[i]) put_col(col, ucp[i], nb + k * s + rr, sb[rr * LT + 4 * i + cq], base[i]);
            }
            }
    }
    // boundary values: u_B = (g - a1 x1 - a2 x2) / dd (down) or h = g + (sigma - alpha) u_B
    // less the transmission residual (up), computed into the column's boundary slots of
    // the LDS area (node order: each slot is read and written by one lane only) and
    // stored afterwards with coalesced writes
    const float2 sma = make_float2(0.f, (float)(-2.0 * G.kb));
    // boundary-row coefficients of axes 0 and 1 (x1, x2 weights and 1/dd)
    const float a1x = (float)ax0.a1, a2x = (float)ax0.a2, a1y = (float)ax1.a1, a2y = (float)ax1.a2;
    const float2 iddx = ccast<float2>(cinv(ax0.dd)), iddy = ccast<float2>(cinv(ax1.dd));
    auto emit = [&](int fn, int t, float2 x1, float2 x2, float a1, float a2, float2 idd) {
        float2* slot = bd + ccol * BS + fn + t;
        const float2 gb = (MODE == LEAF_UP) ? make_float2(0.f, 0.f) : *slot;
        float2 num = csub(csub(gb, cscale(x1, a1)), cscale(x2, a2));
        float2 uB = cmul(num, idd);
        if (MODE == LEAF_UP) *slot = csub(cfma(sma, uB, gb), *slot);
        else if (MODE == LEAF_FACTOR_T) *slot = cfma(sma, uB, gb);
        else *slot = uB;
    };
    // one emit per layer serves both x faces (a lane handles at most one)
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (!colok) break;
        if (xf1 || xf2)
            emit(fnx, k, xf1 ? y[k][0] : pick4(y[k][0], y[k][1], y[k][2], y[k][3], (s - 1) & 3),
                 xf1 ? y[k][1] : pick4(y[k][0], y[k][1], y[k][2], y[k][3], (s - 2) & 3), a1x, a2x, iddx);
    }
#pragma unroll
    for (int vv = 0; vv < 4; ++vv) {
        const int row = 4 * grp + vv;
        if (row >= s || !colok) continue;
        emit(fn3, row, y[0][vv], y[1][vv], a1y, a2y, iddy);
        emit(fn4, row, y[K - 1][vv], y[K - 2][vv], a1y, a2y, iddy);
    }
    wave_lds_sync();
    // coalesced store of the wave's boundary values: up, h in node order; down, u_B
    // in local point order
    // (four columns at a time, their loads issued together)
    const int bnode = l2n(bi);
    for (int c0 = 0; c0 < 16; c0 += 4) {
        float2 base[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = (wcol0 + c0 + j) < io.ncols ? wcol0 + c0 + j : 0;
            base[j] = (MODE == LEAF_DOWN && OUT == OUT_U64_BASE32)
                          ? io.ubase32[((size_t)col * nl + leaf) * m + bi] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = c0 + j, col = wcol0 + c;
            if (!bok || col >= io.ncols) continue;
            if (MODE == LEAF_UP) {
                io.hout[((size_t)col * nl + leaf) * nb + bi] = bd[c * BS + bi];
            } else if (MODE == LEAF_FACTOR_T) {
                // column col of the leaf's map, node-order rows (complex128)
                io.T[((size_t)leaf * nb + col) * nb + bi] = ccast<double2>(bd[c * BS + bi]);
            } else {
                const float2 x = bd[c * BS + bnode];
                put_col(col, (OUT != OUT_U32) ? io.u64[col] + (size_t)leaf * m : nullptr, bi, x, base[j]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Multi-RHS 3D leaf solve on the matrix cores (complex64; layers of s <= 64
// points, K layers).  A workgroup takes 16 RHS columns of one leaf; wave w owns
// rows 16w..16w+15 of a layer (rows padded to S, a multiple of 16).  The same
// block recursion as the 2D kernel,
//   forward   y_k = Shat_k^-1 (r_k - lambda_k y_{k-1})
//   backward  x_k = y_k - mu_k Shat_k^-1 x_{k+1},
// where each layer product is an S x S by S x 16 product: the layer's operand
// goes through an LDS tile shared by the four waves, and each wave multiplies
// its 16 block rows, read from global memory one layer ahead, on
// v_mfma_f32_16x16x4f32 (three real products per complex k-chunk).  The
// right-hand sides of all layers, and then the solution, are held in LDS
// ([column][point], stride n_int + 1), so the interior loads and stores and the
// boundary values (from the first two interior points along each normal) are
// plain coalesced loops.
template <int K, int MODE, typename FT, int OUT>
__global__ void __launch_bounds__(256) leaf3_mfma_kernel(LeafGeom G, LeafCoef C, LeafIO<float> io) {
    extern __shared__ float2 lds3[];
    const int s = G.s, ni = G.n_int, nb = G.nb, m = G.m, nl = io.n_leaves;
    const int S = (s + 15) & ~15;
    const int XS = ni + 1;                 // column stride of X
    constexpr int VS = 17;                 // row stride of the operand tile
    float2* X = lds3;                      // [16][XS]
    float2* Vt = lds3 + 16 * XS;           // [S][VS]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nch = (io.ncols + 15) / 16;
    const int leaf = blockIdx.x / nch, c0 = (blockIdx.x - leaf * nch) * 16;
    const int parity = leaf & 1;
    const int ncol = min(16, io.ncols - c0);

    // 1. right-hand sides r = f - A_IB A_BB^-1 g of all layers into X
    for (int e = tid; e < ni * 16; e += 256) {
        const int q = e % ni, c = e / ni;
        float2 v = make_float2(0.f, 0.f);
        if (c < ncol) {
            v = ccast<float2>(gld(FCol<FT>::base(io, c0 + c, leaf, ni, nl) + q));
            if (MODE == LEAF_DOWN) {
                int g[3];
                interior_coords(G, q, g);
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const AxisCoef& a = C.ax[j];
                    const float2 w = ccast<float2>(cscale(cinv(a.dd), -a.cib));
                    const size_t cb = (size_t)(c0 + c) * io.gld;
                    const int* gs = io.gsrc + (size_t)leaf * nb;
                    if (g[j] == 1) v = cfma(w, io.gin[cb + gs[local_to_node(G, parity, face_point(G, 2 * j + 1, g))]], v);
                    if (g[j] == G.n[j]) v = cfma(w, io.gin[cb + gs[local_to_node(G, parity, face_point(G, 2 * j + 2, g))]], v);
                }
            }
        }
        X[c * XS + q] = v;
    }
    __syncthreads();

    // wave rows and MFMA lane roles
    const int r0 = 16 * wave;
    const bool active = r0 < S;
    const int ccol = lane & 15, grp = lane >> 4;
    const int arow = min(r0 + ccol, s - 1);             // A operand row of this lane (clamped)
    const bool arow_ok = r0 + ccol < s;
    const int nq = S / 4;                               // k-chunks of a layer product
    const float2* sl = io.sinv + (size_t)leaf * K * s * s;
    // A operand of layer k, chunk q: Shat_k^-1[arow][4q + grp] (zero outside s x s)
    float2 abuf[2][16];
    auto afetch = [&](int k, float2 (&a)[16]) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int col = 4 * q + grp;
            const int colc = min(col, s - 1);
            const float2 t = sl[(size_t)k * s * s + arow + (size_t)colc * s];
            a[q] = (q < nq && col < s && arow_ok) ? t : make_float2(0.f, 0.f);
        }
    };
    // y = Shat_k^-1 Vt (this wave's 16 rows, 16 columns)
    auto product = [&](const float2 (&a)[16], float2 (&out)[4]) {
        f4v p1 = f4v{0.f, 0.f, 0.f, 0.f}, p2 = p1, p3 = p1;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            if (q >= nq) break;
            const float2 b = Vt[(4 * q + grp) * VS + ccol];
            p1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q].x, b.x, p1, 0, 0, 0);
            p2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q].y, b.y, p2, 0, 0, 0);
            p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q].x + a[q].y, b.x + b.y, p3, 0, 0, 0);
        }
#pragma unroll
        for (int vv = 0; vv < 4; ++vv) out[vv] = make_float2(p1[vv] - p2[vv], (p3[vv] - p1[vv]) - p2[vv]);
    };

    // 2. forward sweep; y_k replaces r_k in X (each wave reads and writes only
    // its own rows of X, in its own order)
    auto xat = [&](int k, int R) -> float2& { return X[ccol * XS + k * s + R]; };
    // slow-axis couplings (axis 2, K layers), read once: lambda_k couples layer k
    // to k-1, mu_k layer k to k+1; the end value applies at the last / first layer
    const float2 cmid = ccast<float2>(make_double2(C.ax[2].mid, 0.0));
    const float2 cend = ccast<float2>(C.ax[2].eoff);
    afetch(0, abuf[0]);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (k + 1 < K) afetch(k + 1, abuf[(k + 1) & 1]);
        const float2 lam = (k + 1 == K) ? cend : cmid;
        if (active) {
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) {
                const int R = r0 + 4 * grp + vv;
                float2 v = make_float2(0.f, 0.f);
                if (R < s) {
                    v = xat(k, R);
                    if (k > 0) v = csub(v, cmul(lam, xat(k - 1, R)));
                }
                Vt[R * VS + ccol] = v;
            }
        }
        __syncthreads();
        if (active) {
            float2 yk[4];
            product(abuf[k & 1], yk);
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) {
                const int R = r0 + 4 * grp + vv;
                if (R < s) xat(k, R) = yk[vv];
            }
        }
        __syncthreads();
    }
    // 3. backward sweep: x_k replaces y_k in X
    if (K >= 2) afetch(K - 2, abuf[(K - 2) & 1]);
#pragma unroll
    for (int k = K - 2; k >= 0; --k) {
        if (k >= 1) afetch(k - 1, abuf[(k - 1) & 1]);
        const float2 mu = (k == 0) ? cend : cmid;
        if (active) {
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) {
                const int R = r0 + 4 * grp + vv;
                Vt[R * VS + ccol] = (R < s) ? cmul(mu, xat(k + 1, R)) : make_float2(0.f, 0.f);
            }
        }
        __syncthreads();
        if (active) {
            float2 pr[4];
            product(abuf[k & 1], pr);
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) {
                const int R = r0 + 4 * grp + vv;
                if (R < s) xat(k, R) = csub(xat(k, R), pr[vv]);
            }
        }
        __syncthreads();
    }

    // 4. outputs
    auto put = [&](int col, int p, float2 x) {
        const size_t i32 = ((size_t)col * nl + leaf) * m + p;
        if (OUT == OUT_U32) io.u32[i32] = x;
        else {
            double2* u = io.u64[col] + (size_t)leaf * m + p;
            if (OUT == OUT_U64) gst(u, ccast<double2>(x));
            else if (OUT == OUT_U64_BASE32) gst(u, cadd(ccast<double2>(io.ubase32[i32]), ccast<double2>(x)));
            else gst(u, cadd(gld(u), ccast<double2>(x)));
        }
    };
    if (MODE == LEAF_DOWN) {
        for (int e = tid; e < ni * 16; e += 256) {
            const int q = e % ni, c = e / ni;
            if (c < ncol) put(c0 + c, nb + q, X[c * XS + q]);
        }
    }
    const float2 sma = make_float2(0.f, (float)(-2.0 * G.kb));
    for (int e = tid; e < nb * 16; e += 256) {
        const int b = e % nb, c = e / nb;
        if (c >= ncol) continue;
        int f, j, q1, step;
        boundary_info(G, b, f, j, q1, step);
        // the axis constants by selects (a runtime index would go through scratch)
        const double2 dd = j == 0 ? C.ax[0].dd : (j == 1 ? C.ax[1].dd : C.ax[2].dd);
        const double a1 = j == 0 ? C.ax[0].a1 : (j == 1 ? C.ax[1].a1 : C.ax[2].a1);
        const double a2 = j == 0 ? C.ax[0].a2 : (j == 1 ? C.ax[1].a2 : C.ax[2].a2);
        const int col = c0 + c;
        const size_t gi = ((size_t)col * nl + leaf) * nb;   // (the jump and the outgoing data)
        const int pos = local_to_node(G, parity, b);
        const float2 gb = (MODE == LEAF_DOWN) ? io.gin[(size_t)io.gsrc[(size_t)leaf * nb + pos] + (size_t)col * io.gld]
                                              : make_float2(0.f, 0.f);
        const float2 x1 = X[c * XS + q1], x2 = X[c * XS + q1 + step];
        const float2 num = csub(csub(gb, cscale(x1, (float)a1)), cscale(x2, (float)a2));
        const float2 uB = cmul(num, ccast<float2>(cinv(dd)));
        if (MODE == LEAF_UP) {
            float2 h = cfma(sma, uB, gb);
            if (io.jump) h = csub(h, io.jump[gi + b]);
            io.hout[gi + pos] = h;
        } else {
            put(col, b, uB);
        }
    }
}

}  // namespace

void launch_residual(const LeafGeom& G, double2* const* u64, const float2* u32, const double2* kint,
                     const double2* const* f, const double2* const* g, const int* nbr, double2 nu, double2 zeta,
                     float2* rf, float2* rg, float2* rj, int n_leaves, int ncols, hipStream_t st) {
    LeafCoef C = make_coef(G);
    const int groups = std::min((ncols + RES_CPT - 1) / RES_CPT, RES_GROUPS);
    dim3 grid((unsigned)(((size_t)n_leaves * G.m + 255) / 256), (unsigned)groups);
    if (u64)
        hipLaunchKernelGGL(residual_kernel<double2>, grid, dim3(256), 0, st, G, C, u64, nullptr, kint, f, g, nbr, nu,
                           zeta, rf, rg, rj, n_leaves, ncols);
    else
        hipLaunchKernelGGL(residual_kernel<float2>, grid, dim3(256), 0, st, G, C, nullptr, u32, kint, f, g, nbr, nu,
                           zeta, rf, rg, rj, n_leaves, ncols);
}

// Per-column norms of the residual and of the data: one workgroup per column,
// a fixed element order and a fixed tree reduction (the same bits every call).
// out[2 col] = |rf|^2 + |rg|^2 + |rj|^2, out[2 col + 1] = |f|^2 + |g|^2 over
// the physical boundary points.
__global__ void __launch_bounds__(256) fd_norms_kernel(LeafGeom G, const float2* __restrict__ rf,
                                                       const float2* __restrict__ rg,
                                                       const float2* __restrict__ rj,
                                                       const double2* const* __restrict__ f,
                                                       const double2* const* __restrict__ g,
                                                       const int* __restrict__ nbr, int n_leaves,
                                                       double* __restrict__ out) {
    __shared__ double sr[256], sd[256];
    const int col = blockIdx.x, tid = threadIdx.x;
    const size_t ni = (size_t)n_leaves * G.n_int, nbt = (size_t)n_leaves * G.nb;
    double r = 0.0, d = 0.0;
    for (size_t e = tid; e < ni; e += 256) {
        const float2 a = rf[(size_t)col * ni + e];
        r += (double)a.x * a.x + (double)a.y * a.y;
        const double2 b = gld(f[col] + e);
        d += b.x * b.x + b.y * b.y;
    }
    for (size_t e = tid; e < nbt; e += 256) {
        const float2 c = rj[(size_t)col * nbt + e];
        r += (double)c.x * c.x + (double)c.y * c.y;
        const int leaf = (int)(e / G.nb), p = (int)(e % G.nb);
        int fc = 1;
        while (fc < 2 * G.d && p >= G.foff[fc] + G.pf[fc]) ++fc;
        if (nbr[(size_t)leaf * 7 + fc] < 0) {   // physical rows: rg is defined there only
            const float2 a = rg[(size_t)col * nbt + e];
            r += (double)a.x * a.x + (double)a.y * a.y;
            const double2 b = gld(g[col] + e);
            d += b.x * b.x + b.y * b.y;
        }
    }
    sr[tid] = r;
    sd[tid] = d;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) { sr[tid] += sr[tid + h]; sd[tid] += sd[tid + h]; }
        __syncthreads();
    }
    if (tid == 0) { out[2 * col] = sr[0]; out[2 * col + 1] = sd[0]; }
}

void launch_fd_norms(const LeafGeom& G, const float2* rf, const float2* rg, const float2* rj,
                     const double2* const* f, const double2* const* g, const int* nbr, int n_leaves, int ncols,
                     double* out, hipStream_t st) {
    hipLaunchKernelGGL(fd_norms_kernel, dim3(ncols), dim3(256), 0, st, G, rf, rg, rj, f, g, nbr, n_leaves, out);
}

size_t leaf_blocks_lds(const LeafGeom& G) {
    return (size_t)(2 * G.s * G.s + G.s) * sizeof(double2) + (size_t)G.s * sizeof(int) + 64;
}

void launch_leaf_blocks(const LeafGeom& G, const double2* kint, double2* sinv, int* info, int n_leaves,
                        hipStream_t st) {
    LeafCoef C = make_coef(G);
    size_t lds = leaf_blocks_lds(G);
    // one wave per leaf for blocks of at most 256 entries (2D leaves): the
    // Gauss-Jordan steps synchronise one wave instead of four, and four times as
    // many leaves fit a CU
    const int nt = G.s * G.s <= 256 ? 64 : 256;
    hipLaunchKernelGGL(leaf_blocks_kernel, dim3(n_leaves), dim3(nt), lds, st, G, C, kint, sinv, info);
}

template <typename R> int leaf_apply_chunk(const LeafGeom& G, int ncols, size_t lds_limit) {
    // largest CH (<= 16) whose LDS footprint fits the device limit (32 columns
    // per workgroup, one workgroup per CU: the FP64 leaf maps of the 1024^2
    // factor 2.27 -> 3.40 ms)
    size_t e = sizeof(c_t<R>);
    int ch = 16;
    while (ch > 1 && ((size_t)G.n_int * ch + (size_t)G.s * G.s + (size_t)G.s * ch) * e + 256 > lds_limit) ch >>= 1;
    if (ncols < ch) ch = ncols < 1 ? 1 : ncols;
    return ch;
}

// The 3D matrix-core path: layers of at most 64 points, 4, 6 or 8 layers, the
// same pass forms as the 2D kernel.
static bool mfma_leaf3_ok(const LeafGeom& G, const LeafArgs<float>& a) {
    const bool form = a.mode == LEAF_UP ? (a.f64 || a.f32)
                    : a.mode == LEAF_DOWN ? ((a.f64 && !a.ubase32 && !a.uacc) ||
                                             (a.f32 && a.u64 && (a.ubase32 || a.uacc)))
                    : false;
    return G.d == 3 && form && G.s <= 64 && (G.K == 4 || G.K == 6 || G.K == 8);
}

// The matrix-core path takes square 2D leaves of 12x12, 14x14 or 16x16 interior points.
static bool mfma_leaf_ok(const LeafGeom& G, const LeafArgs<float>& a) {
    // square leaves of 12, 14 or 16 points per axis (the kernel's layer size s = K)
    // forms with a matrix-core kernel: up with f in either precision; down with the
    // caller's f into u (complex64 or complex128), or with complex64 f as a correction
    const bool form = a.mode == LEAF_UP ? (a.f64 || a.f32)
                    : a.mode == LEAF_DOWN ? ((a.f64 && !a.ubase32 && !a.uacc) ||
                                             (a.f32 && a.u64 && (a.ubase32 || a.uacc)))
                    : (a.mode == LEAF_FACTOR_T && a.ncols == G.nb);
    return G.d == 2 && form && G.s == G.n[0] && G.s == G.K && G.nb == 4 * G.K && G.n_int == G.K * G.K &&
           (G.K == 12 || G.K == 14 || G.K == 16);
}

bool leaf_mfma_factor_ok(const LeafGeom& G) {
    LeafArgs<float> a;
    a.mode = LEAF_FACTOR_T; a.ncols = G.nb;
    return mfma_leaf_ok(G, a);
}

template <typename R>
void launch_leaf_apply(const LeafGeom& G, const LeafArgs<R>& a, size_t lds_limit, hipStream_t st) {
    LeafCoef C = make_coef(G);
    LeafIO<R> io;
    io.mode = a.mode; io.ncols = a.ncols; io.n_leaves = a.n_leaves;
    io.CH = leaf_apply_chunk<R>(G, a.ncols, lds_limit);
    io.sinv = a.sinv; io.f64 = a.f64; io.f32 = a.f32; io.jump = a.jump; io.gin = a.gin; io.hout = a.hout;
    io.gsrc = a.gsrc; io.gld = a.gld;
    io.u64 = a.u64; io.u32 = a.u32; io.T = a.T;
    io.ubase32 = a.ubase32; io.uacc = a.uacc;
    const int mode = a.mode, ncols = a.ncols, n_leaves = a.n_leaves;
    if constexpr (std::is_same<R, float>::value) {
        if (mfma_leaf_ok(G, a)) {
            const int nchunks = (ncols + 63) / 64;
            const size_t lds = ((size_t)G.K * 256 + 4 * 16 * 17 + 4 * 16 * (size_t)(G.nb + 1)) * sizeof(float2);
            dim3 grid((unsigned)n_leaves * nchunks);
#define RS_LEAF_MFMA(KK)                                                                                        \
    do {                                                                                                         \
        auto L = [&](auto kern) { hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, G, C, io, nchunks); };     \
        if (mode == LEAF_FACTOR_T) {                                                                             \
            L(leaf_mfma_kernel<KK, LEAF_FACTOR_T, float2, OUT_U32>);                                              \
        } else if (mode == LEAF_UP) {                                                                            \
            if (io.f32) L(leaf_mfma_kernel<KK, LEAF_UP, float2, OUT_U32>);                                       \
            else L(leaf_mfma_kernel<KK, LEAF_UP, double2, OUT_U32>);                                             \
        } else if (io.f32) {                                                                                     \
            if (io.ubase32) L(leaf_mfma_kernel<KK, LEAF_DOWN, float2, OUT_U64_BASE32>);                          \
            else L(leaf_mfma_kernel<KK, LEAF_DOWN, float2, OUT_U64_ACC>);                                        \
        } else {                                                                                                 \
            if (io.u64) L(leaf_mfma_kernel<KK, LEAF_DOWN, double2, OUT_U64>);                                    \
            else L(leaf_mfma_kernel<KK, LEAF_DOWN, double2, OUT_U32>);                                           \
        }                                                                                                        \
    } while (0)
            if (G.K == 16) RS_LEAF_MFMA(16);
            else if (G.K == 14) RS_LEAF_MFMA(14);
            else RS_LEAF_MFMA(12);
#undef RS_LEAF_MFMA
            return;
        }
    }
    if constexpr (std::is_same<R, float>::value) {
        if (mfma_leaf3_ok(G, a)) {
            const int S = (G.s + 15) & ~15;
            const size_t lds = ((size_t)16 * (G.n_int + 1) + (size_t)S * 17) * sizeof(float2);
            dim3 grid((unsigned)(n_leaves * ((ncols + 15) / 16)));
#define RS_LEAF3_MFMA(KK)                                                                                       \
    do {                                                                                                         \
        auto L = [&](auto kern) { hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, G, C, io); };              \
        if (mode == LEAF_UP) {                                                                                   \
            if (io.f32) L(leaf3_mfma_kernel<KK, LEAF_UP, float2, OUT_U32>);                                      \
            else L(leaf3_mfma_kernel<KK, LEAF_UP, double2, OUT_U32>);                                            \
        } else if (io.f32) {                                                                                     \
            if (io.ubase32) L(leaf3_mfma_kernel<KK, LEAF_DOWN, float2, OUT_U64_BASE32>);                         \
            else L(leaf3_mfma_kernel<KK, LEAF_DOWN, float2, OUT_U64_ACC>);                                       \
        } else {                                                                                                 \
            if (io.u64) L(leaf3_mfma_kernel<KK, LEAF_DOWN, double2, OUT_U64>);                                   \
            else L(leaf3_mfma_kernel<KK, LEAF_DOWN, double2, OUT_U32>);                                          \
        }                                                                                                        \
    } while (0)
            if (G.K == 8) RS_LEAF3_MFMA(8);
            else if (G.K == 6) RS_LEAF3_MFMA(6);
            else RS_LEAF3_MFMA(4);
#undef RS_LEAF3_MFMA
            return;
        }
    }
    if constexpr (std::is_same<R, double>::value) {
        // few columns of a 2D tree: one wave per leaf (same bits)
        if (G.d == 2 && G.s <= 16 && ncols <= 4 && (mode == LEAF_UP || mode == LEAF_DOWN)) {
            const int cw = ncols == 1 ? 1 : 4;
            io.CH = cw;
            const size_t lds = ((size_t)G.n_int * cw + 16 * cw + 64 * cw) * sizeof(double2);
            dim3 grid(n_leaves, (ncols + cw - 1) / cw);
            const bool tail = G.s % 4 != 0;
            if (cw == 1 && tail) hipLaunchKernelGGL((leaf_wave_kernel<1, true>), grid, dim3(64), lds, st, G, C, io);
            else if (cw == 1) hipLaunchKernelGGL((leaf_wave_kernel<1, false>), grid, dim3(64), lds, st, G, C, io);
            else if (tail) hipLaunchKernelGGL((leaf_wave_kernel<4, true>), grid, dim3(64), lds, st, G, C, io);
            else hipLaunchKernelGGL((leaf_wave_kernel<4, false>), grid, dim3(64), lds, st, G, C, io);
            return;
        }
    }
    size_t lds = ((size_t)G.n_int * io.CH + (size_t)G.s * G.s + (size_t)G.s * io.CH) * sizeof(c_t<R>);
    dim3 grid(n_leaves, (ncols + io.CH - 1) / io.CH);
    const int ss = G.s * G.s;
    if (ss <= 256) hipLaunchKernelGGL((leaf_apply_kernel<R, 1>), grid, dim3(256), lds, st, G, C, io);
    else if (ss <= 1024) hipLaunchKernelGGL((leaf_apply_kernel<R, 4>), grid, dim3(256), lds, st, G, C, io);
    else if (ss <= 4096) hipLaunchKernelGGL((leaf_apply_kernel<R, 16>), grid, dim3(256), lds, st, G, C, io);
    else hipLaunchKernelGGL((leaf_apply_kernel<R, 0>), grid, dim3(256), lds, st, G, C, io);
}

template int leaf_apply_chunk<double>(const LeafGeom&, int, size_t);
template int leaf_apply_chunk<float>(const LeafGeom&, int, size_t);
template void launch_leaf_apply<double>(const LeafGeom&, const LeafArgs<double>&, size_t, hipStream_t);
template void launch_leaf_apply<float>(const LeafGeom&, const LeafArgs<float>&, size_t, hipStream_t);

}  // namespace rs
