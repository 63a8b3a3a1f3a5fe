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
    f = 1;
    while (f < 2 * G.d && b >= G.foff[f] + G.pf[f]) ++f;
    int t = b - G.foff[f];
    j = (f - 1) >> 1;
    int g[3] = {1, 1, 1};
    if (G.d == 2) {
        g[1 - j] = t + 1;
    } else {
        int a0 = (j == 0) ? 1 : 0;
        int a1 = (j == 2) ? 1 : 2;
        g[a0] = t % G.n[a0] + 1;
        g[a1] = t / G.n[a0] + 1;
    }
    g[j] = (f & 1) ? 1 : G.n[j];
    q1 = (g[0] - 1) + G.n[0] * ((g[1] - 1) + (G.d == 3 ? G.n[1] * (g[2] - 1) : 0));
    step = ((f & 1) ? 1 : -1) * stride_of(G, j);
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
            for (int off = 32; off > 0; off >>= 1) {
                double ob = __shfl_xor(best, off);
                int oi = __shfl_xor(bi, off);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
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
    const double2* jump;      // UP: subtracted from the outgoing data h: [col][leaf][nb] local order
    const c_t<R>* gin;        // DOWN: incoming data g: [col][leaf][nb] node order
    c_t<R>* hout;             // UP: outgoing data h: [col][leaf][nb] node order
    double2* const* u64;      // DOWN: solution, complex128: column c at u64[c], [leaf][m] ...
    float2* u32;              //   ... or complex64 [col][leaf][m] (one set); written, never read
    double2* T;               // FACTOR: [leaf][nb][nb] node order, column-major
};

// f at interior point q of (col, leaf); u at point p of (col, leaf)
template <typename V, typename R>
__device__ inline V load_f(const LeafIO<R>& io, int col, int leaf, int ni, int q) {
    if (io.f64) return ccast<V>(io.f64[col][(size_t)leaf * ni + q]);
    if (io.f32) return ccast<V>(io.f32[((size_t)col * io.n_leaves + leaf) * ni + q]);
    return V{};
}

template <typename V, typename R>
__device__ inline void store_u(const LeafIO<R>& io, int col, int leaf, int m, int p, V x) {
    if (io.u64) io.u64[col][(size_t)leaf * m + p] = ccast<double2>(x);
    else io.u32[((size_t)col * io.n_leaves + leaf) * m + p] = ccast<float2>(x);
}


template <typename R>
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
            return io.gin[((size_t)col * io.n_leaves + leaf) * nb + local_to_node(G, parity, b)];
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
            V acc = cz<R>();
            for (int w2 = 0; w2 < s; ++w2) acc = cfma(B[w + w2 * s], Vt[w2 * CH + c], acc);
            Y[(size_t)k * s * CH + e] = acc;
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
            V acc = cz<R>();
            for (int w2 = 0; w2 < s; ++w2) acc = cfma(B[w + w2 * s], Vt[w2 * CH + c], acc);
            Y[(size_t)k * s * CH + e] = csub(Y[(size_t)k * s * CH + e], acc);
        }
        __syncthreads();
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

// FP64 residual of the global FD system at the current leaf solutions u:
//   interior rows       rf = f - (Delta_h + k) u                     (fd_module.f90:140-181)
//   physical rows       rg = g - (alpha + D_n) u                     (fd_module.f90:187-216)
//   transmission rows   rj = zeta g_l + nu g_nb - h_l,  g = (alpha + D_n) u, h = (sigma + D_n) u
// The transmission condition is the pointwise form of the merge system
// M [g1; g2] = [h1; h2] (robin_tree_module.f90:425-467): it holds at every pair
// of coincident leaf boundary points once the tree has solved.  A correction
// solve takes (rf, rg) as (f, g) and subtracts rj from the leaf outgoing data.
// column col of the current solution: complex128 through the column table, or
// the complex64 first iterate [col][leaf][m]
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

template <typename U>
__global__ void residual_kernel(LeafGeom G, LeafCoef C, double2* const* __restrict__ u64,
                                const float2* __restrict__ u32, const double2* __restrict__ kint,
                                const double2* const* __restrict__ f, const double2* const* __restrict__ g,
                                const int* __restrict__ nbr, double2 nu, double2 zeta, float2* __restrict__ rf,
                                double2* __restrict__ rg, double2* __restrict__ rj, int n_leaves, int ncols,
                                int cols_per_block) {
    // one block: one leaf, cols_per_block columns; threads run over (column, point)
    const int leaf = blockIdx.x % n_leaves;
    const int cb = blockIdx.x / n_leaves;
    const int nb = G.nb, ni = G.n_int;
    for (int e = threadIdx.x; e < cols_per_block * G.m; e += blockDim.x) {
    const int col = cb * cols_per_block + e / G.m;
    const int p = e % G.m;
    if (col >= ncols) break;
    const size_t lc = (size_t)col * n_leaves + leaf;
    const U* ucol = ucolumn<U>(u64, u32, col, n_leaves, G.m);
    const U* ul = ucol + (size_t)leaf * G.m;
    const double2 sma = make_double2(0.0, -2.0 * G.kb);
    auto uval = [&](const U* uu, int i) -> double2 { return ccast<double2>(uu[i]); };
    auto incoming = [&](const U* uu, int b) -> double2 {
        int f_, j_, q1, st;
        boundary_info(G, b, f_, j_, q1, st);
        const AxisCoef& a = C.ax[j_];
        double2 v = cmul(a.dd, uval(uu, b));
        v = cadd(v, cscale(uval(uu, nb + q1), a.a1));
        return cadd(v, cscale(uval(uu, nb + q1 + st), a.a2));
    };
    if (p < nb) {
        int fc = 1;
        while (fc < 2 * G.d && p >= G.foff[fc] + G.pf[fc]) ++fc;
        const double2 gA = incoming(ul, p);
        const int nl2 = nbr[(size_t)leaf * 7 + fc];
        if (nl2 < 0) {
            rg[lc * nb + p] = csub(g[col][(size_t)leaf * nb + p], gA);
            rj[lc * nb + p] = make_double2(0.0, 0.0);
        } else {
            const int f2 = (fc & 1) ? fc + 1 : fc - 1;
            const int b2 = G.foff[f2] + (p - G.foff[fc]);
            const double2 gN = incoming(ucol + (size_t)nl2 * G.m, b2);
            const double2 hA = cfma(sma, uval(ul, p), gA);
            rj[lc * nb + p] = csub(cadd(cmul(zeta, gA), cmul(nu, gN)), hA);
            rg[lc * nb + p] = make_double2(0.0, 0.0);
        }
        continue;
    }
    const int q = p - nb;
    int gq[3];
    interior_coords(G, q, gq);
    double2 Au = cmul(kint[(size_t)leaf * ni + q], uval(ul, p));
    for (int j = 0; j < G.d; ++j) {
        const AxisCoef& a = C.ax[j];
        const double ih2 = a.mid;
        const int st = stride_of(G, j);
        double wm, w0, wp;
        if (gq[j] == 1) { wm = 8.0 / 3.0; w0 = -4.0; wp = 4.0 / 3.0; }
        else if (gq[j] == G.n[j]) { wm = 4.0 / 3.0; w0 = -4.0; wp = 8.0 / 3.0; }
        else { wm = 1.0; w0 = -2.0; wp = 1.0; }
        const double2 um = (gq[j] == 1) ? uval(ul, face_point(G, 2 * j + 1, gq)) : uval(ul, nb + q - st);
        const double2 upp = (gq[j] == G.n[j]) ? uval(ul, face_point(G, 2 * j + 2, gq)) : uval(ul, nb + q + st);
        Au = cadd(Au, cscale(uval(ul, p), w0 * ih2));
        Au = cadd(Au, cscale(um, wm * ih2));
        Au = cadd(Au, cscale(upp, wp * ih2));
    }
    rf[lc * ni + q] = ccast<float2>(csub(f[col][(size_t)leaf * ni + q], Au));
    }
}

// ---------------------------------------------------------------------------
// Multi-RHS leaf solve on the matrix cores (complex64, 2D leaves with an even
// number s <= 16 of points per layer and K layers).  Each wave owns 16 RHS
// columns; a workgroup covers 64 columns of one leaf.  The layer recursion
//   forward   y_k = Shat_k^-1 (r_k - lambda_k y_{k-1})
//   backward  x_k = y_k - mu_k Shat_k^-1 x_{k+1}
// is one 16x16x16 complex product per layer (v_mfma_f32_16x16x4f32, four real
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

template <int K>
__global__ void __launch_bounds__(256) leaf_mfma_kernel(LeafGeom G, LeafCoef C, LeafIO<float> io, int nchunks) {
    extern __shared__ float2 lds2[];
    const int s = G.s, ni = G.n_int, nb = G.nb, m = G.m, nl = io.n_leaves;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float2* Bl = lds2;                                   // [K][16 x 16], column-major
    float2* scr = lds2 + K * 256 + wave * 512;           // two 16x16 tiles, row-major
    float2* bd = lds2 + K * 256 + 4 * 512 + wave * 16 * nb;  // per-column boundary data (g or jump)
    const int leaf = blockIdx.x / nchunks, chunk = blockIdx.x - leaf * nchunks;
    const int ccol = lane & 15, grp = lane >> 4;
    const int mycol = chunk * 64 + wave * 16 + ccol;
    const bool colok = mycol < io.ncols;
    const int parity = leaf & 1;
    const size_t lc = (size_t)(colok ? mycol : 0) * nl + leaf;
    // this lane's column of f: complex128 through the column table, or complex64 correction data
    const double2* f64p = (io.f64 && colok) ? io.f64[mycol] + (size_t)leaf * ni : nullptr;
    const float2* f32p = io.f32 ? io.f32 + lc * ni : nullptr;

    const float2* sl = io.sinv + (size_t)leaf * K * s * s;
    for (int e = tid; e < K * 256; e += 256) {
        int k = e >> 8, r = e & 15, c = (e >> 4) & 15;
        Bl[e] = (r < s && c < s) ? sl[(size_t)k * s * s + r + c * s] : make_float2(0.f, 0.f);
    }
    // boundary data of the wave's 16 columns, read once with coalesced loads:
    // DOWN: incoming g (node order); UP: the transmission residual (local order)
    {
        const int wcol0 = chunk * 64 + wave * 16;
        for (int e = lane; e < 16 * nb; e += 64) {
            const int c = e / nb, i = e - c * nb;
            const int col = wcol0 + c;
            float2 v = make_float2(0.f, 0.f);
            if (col < io.ncols) {
                const size_t gi = ((size_t)col * nl + leaf) * nb + i;
                if (io.mode == LEAF_DOWN) v = io.gin[gi];
                else if (io.jump) v = ccast<float2>(io.jump[gi]);
            }
            bd[e] = v;
        }
    }
    __syncthreads();

    const AxisCoef& ax0 = C.ax[0];
    const AxisCoef& ax1 = C.ax[1];
    const float2 cbx = ccast<float2>(cscale(cinv(ax0.dd), -ax0.cib));
    const float2 cby = ccast<float2>(cscale(cinv(ax1.dd), -ax1.cib));
    auto gval = [&](int face, int t) -> float2 { return bd[ccol * nb + G.noff[parity][face] + t]; };

    float2 y[K][4];
    float2 prev[4];
    int buf = 0;
    auto product = [&](int k, const float2 v[4], f4v& aR, f4v& aI) {
        float2* sb = scr + buf * 256;
        for (int vv = 0; vv < 4; ++vv) sb[(4 * grp + vv) * 16 + ccol] = v[vv];
        __syncthreads();
        aR = f4v{0.f, 0.f, 0.f, 0.f};
        aI = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float2 a = Bl[k * 256 + ccol + 16 * (4 * q + grp)];
            const float2 b = sb[(4 * q + grp) * 16 + ccol];
            aR = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, aR, 0, 0, 0);
            aR = __builtin_amdgcn_mfma_f32_16x16x4f32(-a.y, b.y, aR, 0, 0, 0);
            aI = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.y, aI, 0, 0, 0);
            aI = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.x, aI, 0, 0, 0);
        }
        buf ^= 1;
    };

    // forward sweep; the interior data of layer k+1 is fetched while layer k is
    // being processed
    auto fetch = [&](int k, double2 (&raw)[4]) {
#pragma unroll
        for (int vv = 0; vv < 4; ++vv) {
            const int row = 4 * grp + vv;
            double2 x = make_double2(0.0, 0.0);
            if (row < s && colok) {
                if (f64p) x = f64p[k * s + row];
                else if (f32p) x = ccast<double2>(f32p[k * s + row]);
            }
            raw[vv] = x;
        }
    };
    // the interior data runs PF layers ahead of the sweep (memory latency is
    // several layer steps long)
    constexpr int PF = 4;
    double2 fbuf[PF][4];
#pragma unroll
    for (int k = 0; k < PF && k < K; ++k) fetch(k, fbuf[k]);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        float2 v[4];
#pragma unroll
        for (int vv = 0; vv < 4; ++vv) {
            const int row = 4 * grp + vv;
            float2 r = ccast<float2>(fbuf[k % PF][vv]);
            if (row < s && colok && io.mode == LEAF_DOWN) {
                if (row == 0) r = cfma(cbx, gval(1, k), r);
                if (row == s - 1) r = cfma(cbx, gval(2, k), r);
                if (k == 0) r = cfma(cby, gval(3, row), r);
                if (k == K - 1) r = cfma(cby, gval(4, row), r);
            }
            if (k > 0) r = csub(r, cmul(ccast<float2>(slow_lambda(G, C, k)), prev[vv]));
            v[vv] = r;
        }
        if (k + PF < K) fetch(k + PF, fbuf[k % PF]);
        f4v aR, aI;
        product(k, v, aR, aI);
#pragma unroll
        for (int vv = 0; vv < 4; ++vv) { y[k][vv] = make_float2(aR[vv], aI[vv]); prev[vv] = y[k][vv]; }
    }
    // backward sweep
#pragma unroll
    for (int k = K - 2; k >= 0; --k) {
        float2 w[4];
        const float2 mu = ccast<float2>(slow_mu(G, C, k));
#pragma unroll
        for (int vv = 0; vv < 4; ++vv) w[vv] = cmul(mu, y[k + 1][vv]);
        f4v aR, aI;
        product(k, w, aR, aI);
#pragma unroll
        for (int vv = 0; vv < 4; ++vv) y[k][vv] = csub(y[k][vv], make_float2(aR[vv], aI[vv]));
    }
    if (!colok) return;

    double2* u64p = io.u64 ? io.u64[mycol] + (size_t)leaf * m : nullptr;
    float2* u32p = io.u32 ? io.u32 + lc * m : nullptr;
    auto put = [&](int p, float2 x) {
        if (u64p) u64p[p] = ccast<double2>(x);
        else u32p[p] = x;
    };
    if (io.mode == LEAF_DOWN) {
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) {
                const int row = 4 * grp + vv;
                if (row < s) put(nb + k * s + row, y[k][vv]);
            }
    }
    // boundary values: u_B = (g - a1 x1 - a2 x2) / dd, h = g + (sigma - alpha) u_B
    const float2 sma = make_float2(0.f, (float)(-2.0 * G.kb));
    auto emit = [&](int face, int t, float2 x1, float2 x2, const AxisCoef& a) {
        const int b = G.foff[face] + t;
        const float2 gb = (io.mode == LEAF_DOWN) ? gval(face, t) : make_float2(0.f, 0.f);
        float2 num = csub(csub(gb, cscale(x1, (float)a.a1)), cscale(x2, (float)a.a2));
        float2 uB = cmul(num, ccast<float2>(cinv(a.dd)));
        if (io.mode == LEAF_UP) {
            float2 hb = cfma(sma, uB, gb);
            hb = csub(hb, bd[ccol * nb + b]);
            io.hout[lc * nb + G.noff[parity][face] + t] = hb;
        } else {
            put(b, uB);
        }
    };
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (grp == 0) emit(1, k, y[k][0], y[k][1], ax0);
        if (grp == (s - 1) / 4) emit(2, k, pick4(y[k][0], y[k][1], y[k][2], y[k][3], (s - 1) & 3),
                                        pick4(y[k][0], y[k][1], y[k][2], y[k][3], (s - 2) & 3), ax0);
    }
#pragma unroll
    for (int vv = 0; vv < 4; ++vv) {
        const int row = 4 * grp + vv;
        if (row >= s) continue;
        emit(3, row, y[0][vv], y[1][vv], ax1);
        emit(4, row, y[K - 1][vv], y[K - 2][vv], ax1);
    }
}

}  // namespace

void launch_residual(const LeafGeom& G, double2* const* u64, const float2* u32, const double2* kint,
                     const double2* const* f, const double2* const* g, const int* nbr, double2 nu, double2 zeta,
                     float2* rf, double2* rg, double2* rj, int n_leaves, int ncols, hipStream_t st) {
    LeafCoef C = make_coef(G);
    const int cpb = 8;
    dim3 grid((unsigned)n_leaves * ((ncols + cpb - 1) / cpb));
    if (u64)
        hipLaunchKernelGGL(residual_kernel<double2>, grid, dim3(256), 0, st, G, C, u64, nullptr, kint, f, g, nbr,
                           nu, zeta, rf, rg, rj, n_leaves, ncols, cpb);
    else
        hipLaunchKernelGGL(residual_kernel<float2>, grid, dim3(256), 0, st, G, C, nullptr, u32, kint, f, g, nbr,
                           nu, zeta, rf, rg, rj, n_leaves, ncols, cpb);
}

size_t leaf_blocks_lds(const LeafGeom& G) {
    return (size_t)(2 * G.s * G.s + G.s) * sizeof(double2) + (size_t)G.s * sizeof(int) + 64;
}

void launch_leaf_blocks(const LeafGeom& G, const double2* kint, double2* sinv, int* info, int n_leaves,
                        hipStream_t st) {
    LeafCoef C = make_coef(G);
    size_t lds = leaf_blocks_lds(G);
    hipLaunchKernelGGL(leaf_blocks_kernel, dim3(n_leaves), dim3(256), lds, st, G, C, kint, sinv, info);
}

template <typename R> int leaf_apply_chunk(const LeafGeom& G, int ncols, size_t lds_limit) {
    // largest CH (<= 16) whose LDS footprint fits the device limit
    size_t e = sizeof(c_t<R>);
    int ch = 16;
    while (ch > 1 && ((size_t)G.n_int * ch + (size_t)G.s * G.s + (size_t)G.s * ch) * e + 256 > lds_limit) ch >>= 1;
    if (ncols < ch) ch = ncols < 1 ? 1 : ncols;
    return ch;
}

// The matrix-core path takes 2D leaves of 12x12, 14x14 or 16x16 interior points.
static bool mfma_leaf_ok(const LeafGeom& G, int mode) {
    return G.d == 2 && mode != LEAF_FACTOR_T && G.s == G.n[0] && G.s <= 16 && (G.s % 2) == 0 &&
           (G.K == 12 || G.K == 14 || G.K == 16);
}

template <typename R>
void launch_leaf_apply(const LeafGeom& G, const LeafArgs<R>& a, size_t lds_limit, hipStream_t st) {
    LeafCoef C = make_coef(G);
    LeafIO<R> io;
    io.mode = a.mode; io.ncols = a.ncols; io.n_leaves = a.n_leaves;
    io.CH = leaf_apply_chunk<R>(G, a.ncols, lds_limit);
    io.sinv = a.sinv; io.f64 = a.f64; io.f32 = a.f32; io.jump = a.jump; io.gin = a.gin; io.hout = a.hout;
    io.u64 = a.u64; io.u32 = a.u32; io.T = a.T;
    const int mode = a.mode, ncols = a.ncols, n_leaves = a.n_leaves;
    if constexpr (std::is_same<R, float>::value) {
        if (mfma_leaf_ok(G, mode)) {
            const int nchunks = (ncols + 63) / 64;
            const size_t lds = ((size_t)G.K * 256 + 4 * 512 + 4 * 16 * (size_t)G.nb) * sizeof(float2);
            dim3 grid((unsigned)n_leaves * nchunks);
            if (G.K == 16)
                hipLaunchKernelGGL(leaf_mfma_kernel<16>, grid, dim3(256), lds, st, G, C, io, nchunks);
            else if (G.K == 14)
                hipLaunchKernelGGL(leaf_mfma_kernel<14>, grid, dim3(256), lds, st, G, C, io, nchunks);
            else
                hipLaunchKernelGGL(leaf_mfma_kernel<12>, grid, dim3(256), lds, st, G, C, io, nchunks);
            return;
        }
    }
    size_t lds = ((size_t)G.n_int * io.CH + (size_t)G.s * G.s + (size_t)G.s * io.CH) * sizeof(c_t<R>);
    dim3 grid(n_leaves, (ncols + io.CH - 1) / io.CH);
    hipLaunchKernelGGL(leaf_apply_kernel<R>, grid, dim3(256), lds, st, G, C, io);
}

template int leaf_apply_chunk<double>(const LeafGeom&, int, size_t);
template int leaf_apply_chunk<float>(const LeafGeom&, int, size_t);
template void launch_leaf_apply<double>(const LeafGeom&, const LeafArgs<double>&, size_t, hipStream_t);
template void launch_leaf_apply<float>(const LeafGeom&, const LeafArgs<float>&, size_t, hipStream_t);

}  // namespace rs
