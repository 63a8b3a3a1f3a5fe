// tree.hip -- factor and solve orchestration of the HPS tree on one MI355X.
//
// Factor (robin_tree_factor / compute_RTR, robin_tree_module.f90:96-185, 375-650):
//   1. one launch factors every leaf (block-tridiagonal LU, k_leaf.hip) and one
//      launch builds every leaf RtR map;
//   2. per level, bottom-up, all nodes of the level at once, with
//      a = zeta I - T1_II, c = zeta I - T2_II and M = [a, nu I; nu I, c]:
//        Wi   = (nu^2 I - a c)^-1                                     (GEMM, FP64 pivoted inverse)
//        invM = [-c Wi, (I - (-c Wi) a) / nu; nu Wi, -Wi a]           (3 GEMMs + copy)
//        Y    = invM diag(T1_IN, T2_IN)                               (GEMM)
//        T    = diag(T1_NN, T2_NN) + diag(T1_NI, T2_NI) Y, scattered into the
//               parent's node order                                  (GEMM)
//      which is the reference's C11..C22 / zgemm block algebra written with
//      an explicit inv(M).  The reference inverts two matrices one after the
//      other, D = c^-1 and S = (a - nu^2 D)^-1 (robin_tree_module.f90:440-467);
//      the same inv(M) follows from one inverse of the i x i matrix
//      nu^2 I - a c (nu is nonzero whenever bndry_params passes its
//      determinant check), which halves the sequential pivoted eliminations
//      on the factor's critical path.  The root RtR, which no solve reads, is
//      skipped by default.
// Solve (forward_sub / backward_sub, robin_tree_module.f90:813-1146):
//   up:   leaf particular solves -> per level g_I = invM [h1_I; h2_I],
//         h_parent = P (h_N + T_NI g_I)
//   down: root g from the physical data -> per level children's g_N by gather,
//         g_I += invM [T1_IN g1_N; T2_IN g2_N] -> leaf updates u += A^-1 [g; 0]
//   every step batched over the nodes of a level and over all right-hand sides.
#include "tree.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace rs {

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) return fail(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)

GemmDesc desc0() {
    GemmDesc d;
    std::memset(&d, 0, sizeof d);
    d.ar = 1.0;
    d.ksplit = 1 << 30;
    return d;
}

}  // namespace

template <typename R> void SolveWork<R>::clear() {
    hV.clear(); gF.clear(); G.clear(); Z.clear();
    d_up_g.clear(); d_up_h.clear(); d_dn_t.clear(); d_dn_g.clear();
    ws = nullptr; rf = rg = rj = uA = nullptr;
    for (auto& t : tile) t.clear();
    arena.release();
    cap = 0;
    refine = false;
}

template <typename R> void SolveWork<R>::bind(char* base, const WorkLayout& lay) {
    using V_ = c_t<R>;
    auto at = [&](size_t off) -> char* { return off == WorkLayout::NONE ? nullptr : base + off; };
    const size_t L1 = lay.gF.size();
    hV.assign(L1, nullptr); gF.assign(L1, nullptr);
    for (size_t l = 0; l < L1; ++l) { hV[l] = (V_*)at(lay.hV[l]); gF[l] = (V_*)at(lay.gF[l]); }
    const size_t L = lay.G.size();
    G.assign(L, nullptr); Z.assign(L, nullptr);
    d_up_g.assign(L, nullptr); d_up_h.assign(L, nullptr); d_dn_t.assign(L, nullptr); d_dn_g.assign(L, nullptr);
    for (size_t l = 0; l < L; ++l) {
        G[l] = (V_*)at(lay.G[l]); Z[l] = (V_*)at(lay.Z[l]);
        d_up_g[l] = (GemmDesc*)at(lay.dug[l]); d_up_h[l] = (GemmDesc*)at(lay.duh[l]);
        d_dn_t[l] = (GemmDesc*)at(lay.ddt[l]); d_dn_g[l] = (GemmDesc*)at(lay.ddg[l]);
    }
    ws = at(lay.ws);
    rf = (float2*)at(lay.rf); rg = (float2*)at(lay.rg); rj = (float2*)at(lay.rj); uA = (float2*)at(lay.uA);
    cap = lay.cap;
    refine = lay.refine;
}

// The descriptors of node n's four solve products at one level (written by the
// host for a workspace of its own, and by solve_desc_kernel for one in the
// device scratch array):
//   up_g  g_I = invM [h1_I; h2_I]                          (B from the two children)
//   up_h  h_parent = P (h_N + T_NI g_I), per child c        (rows scattered by prow)
//   dn_t  Z_c = T_IN g_N of child c                         (g_N rows read through dsrc)
//   dn_g  g_I += invM Z, in place in U                      (the children's incoming interface rows)
__host__ __device__ inline void solve_descs(const SolveLevel& p, int n, GemmDesc* ug, GemmDesc* uh, GemmDesc* dt,
                                            GemmDesc* dg) {
    // children's maps: compact [T_NI (bn x i) | T_IN (i x bn)] per child
    const size_t e = p.esz, i = p.i, b = p.b, bn = p.bn, ii = i * i, tstride = 2 * bn * i;
    const int par = (p.lvl == 0) ? 0 : (n & 1);
    GemmDesc z = {};
    z.ar = 1.0;
    z.ksplit = 1 << 30;
    GemmDesc g = z;
    g.A = p.IM + (size_t)n * 4 * ii * e; g.lda = 2 * p.i;
    g.B = p.hC + ((size_t)(2 * n) * b + bn) * e; g.B2 = p.hC + ((size_t)(2 * n + 1) * b + bn) * e;
    g.ksplit = p.i; g.ldb = p.ldc;
    g.Out = p.Gl + (size_t)n * 2 * i * e; g.ldo = p.ldg;
    *ug = g;
    // (in place: the level's interface data becomes the children's incoming
    // interface rows, where the levels below and the leaves read it)
    GemmDesc q = z;
    q.A = p.IM + (size_t)n * 4 * ii * e; q.lda = 2 * p.i;
    q.B = p.Zl + (size_t)n * 2 * i * e; q.ldb = p.ldz;
    q.Din = p.Gl + (size_t)n * 2 * i * e; q.ldd = p.ldg; q.br = 1.0;
    q.Out = p.Gl + (size_t)n * 2 * i * e; q.ldo = p.ldg;
    *dg = q;
    for (int c = 0; c < 2; ++c) {
        const char* Tcn = p.Tc + (size_t)(2 * n + c) * tstride * e;
        if (p.lvl > 0) {
            GemmDesc u = z;
            u.A = Tcn; u.lda = p.bn;
            u.B = p.Gl + ((size_t)n * 2 * i + (size_t)c * i) * e; u.ldb = p.ldg;
            u.Out = p.hP + (size_t)n * p.Pp * e; u.ldo = p.ldp;
            u.rmap = p.prow + (size_t)(par * 2 + c) * bn;
            u.Din = p.hC + (size_t)(2 * n + c) * b * e; u.ldd = p.ldc; u.br = 1.0;
            uh[c] = u;
        }
        // B: the child's non-interface incoming rows, read where they were written
        GemmDesc t = z;
        t.A = Tcn + bn * i * e; t.lda = p.i;
        t.B = p.U; t.bmap = p.dsrc + (size_t)(2 * n + c) * bn; t.ldb = p.ldg; t.bpair = p.dpair;
        t.Out = p.Zl + ((size_t)n * 2 * i + (size_t)c * i) * e; t.ldo = p.ldz;
        dt[c] = t;
    }
}

__global__ void solve_desc_kernel(SolveLevel p, GemmDesc* ug, GemmDesc* uh, GemmDesc* dt, GemmDesc* dg) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= p.nodes) return;
    GemmDesc h[2];
    solve_descs(p, n, ug + n, h, dt + 2 * n, dg + n);
    if (p.lvl > 0) { uh[2 * n] = h[0]; uh[2 * n + 1] = h[1]; }
}

Tree::Tree() { rs_opts_default(&opts); }

Slot& Tree::slot(int i) {
    while ((int)slots.size() <= i) slots.push_back(new Slot);
    return *slots[i];
}

void Tree::clear_slots() {
    for (Slot* s : slots) delete s;
    slots.clear();
}

int Tree::ensure_device() {
    if (device_ready) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail("Error: no HIP device (the MI355X library needs a GPU).");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    lds_limit = prop.sharedMemPerBlock;
    // The tree's own stream does not synchronise implicitly with the default
    // stream; calls that run on it order themselves against the default stream
    // with two events instead (join_default_in / join_default_out), which also
    // holds for the launch of a captured graph.
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    if (!ev_in) HIPCHK(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
    if (!ev_out) HIPCHK(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
    device_ready = true;
    if (leaf_blocks_lds(G) > lds_limit) return fail("Error: leaf layer too large for LDS (reduce the leaf size).");
    const Plan& P = plan;
    for (DevBuf* b : d_prow) delete b;
    d_prow.assign(P.L, nullptr);
    for (int l = 0; l < P.L; ++l) {
        d_prow[l] = new DevBuf;
        if (upload(*d_prow[l], P.lv[l].prow)) return 1;
    }
    if (upload(d_root_leaf, P.root_leaf) || upload(d_root_loc, P.root_loc) || upload(d_leaf_nbr, P.leaf_nbr))
        return 1;
    for (DevBuf* b : d_src_dn) delete b;
    d_src_dn.assign(P.L, nullptr);
    for (int l = 0; l < P.L; ++l) {
        d_src_dn[l] = new DevBuf;
        if (upload(*d_src_dn[l], src_dn[l])) return 1;
    }
    if (upload(d_src_leaf, src_leaf)) return 1;
    return 0;
}

// Work the caller queued on the default stream (torch's default stream,
// hipMemset / hipMemcpy without a stream) before the call runs before the
// call's work on st.
int Tree::join_default_in(hipStream_t st) {
    HIPCHK(hipEventRecord(ev_in, nullptr));
    HIPCHK(hipStreamWaitEvent(st, ev_in, 0));
    return 0;
}

// Work queued on the default stream after the call runs after the call's work on st.
int Tree::join_default_out(hipStream_t st) {
    HIPCHK(hipEventRecord(ev_out, st));
    HIPCHK(hipStreamWaitEvent(nullptr, ev_out, 0));
    return 0;
}

Tree::~Tree() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    clear_slots();
    free_factor();
    for (DevBuf* b : d_prow) delete b;
    for (DevBuf* b : d_src_dn) delete b;
    if (stream) (void)hipStreamDestroy(stream);
}

int Tree::alloc(DevBuf& b, size_t bytes) {
    b.release();
    if (bytes == 0) bytes = 16;
    if (opts.hbm_budget > 0 && dev_bytes + (long long)bytes > opts.hbm_budget)
        return fail("Error: device memory budget of this tree exceeded.");
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) { b.p = nullptr; return fail(std::string("Error: hipMalloc failed: ") + hipGetErrorString(e)); }
    b.bytes = bytes;
    b.counter = &dev_bytes;
    dev_bytes += (long long)bytes;
    peak_bytes = std::max(peak_bytes, dev_bytes);
    return 0;
}

template <typename T_> int Tree::upload(DevBuf& b, const std::vector<T_>& v) {
    if (alloc(b, v.size() * sizeof(T_))) return 1;
    HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T_), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return 0;
}

void Tree::drop_factor_graph() {
    if (fgraph) (void)hipGraphExecDestroy(fgraph);
    fgraph = nullptr;
}

void Tree::free_factor() {
    auto drop = [&](std::vector<DevBuf*>& v) {
        for (DevBuf* b : v) delete b;
        v.clear();
    };
    drop_factor_graph();
    drop(T); drop(Tn); drop(Tn32); drop(invM); drop(invM32);
    ftile.clear();
    fdoff.clear();
    for (DevBuf* b : {&sinv, &sinv32, &info, &fscratch, &invM_t, &tmap[0], &tmap[1], &fdesc, &fcmap}) b->release();
    have32 = have64 = false;
    factored = false;
}

int Tree::plan_tree(int d, int L, const double* domain, double h_tgt, double kb, const double* bp) {
    std::string e = plan.build(d, L, domain, h_tgt, kb, bp);
    if (!e.empty()) return fail(e);
    const Plan& P = plan;
    const Level& LF = P.lv[P.L];
    G = LeafGeom{};
    G.d = d;
    for (int j = 0; j < 3; ++j) { G.n[j] = j < d ? LF.ptbox[j] : 1; G.h[j] = P.h[j]; }
    G.n_int = P.n_int; G.nb = P.nb; G.m = P.m;
    G.K = LF.ptbox[d - 1];
    G.s = P.n_int / G.K;
    for (int f = 0; f < 7; ++f) {
        G.pf[f] = LF.pf[f];
        G.foff[f] = P.leaf_foff[f];
        G.noff[0][f] = LF.noff[0][f];
        G.noff[1][f] = LF.noff[1][f];
    }
    G.kb = kb;
    if (device_ready) {
        // a re-plan on a live handle: drop everything sized for the old tree
        if (stream) (void)hipStreamSynchronize(stream);
        free_factor();
        clear_slots();
        tile_cache.clear();
        d_pts_box.release();
        d_pts_pat.release();
        device_ready = false;
        if (stream) { (void)hipStreamDestroy(stream); stream = nullptr; }
    }
    if (build_sources()) return 1;
    planned = true;
    return 0;
}

// The composition of the down pass's copies: a child's non-interface row j is
// its parent's row prow[parity][c][j], and its interface rows are the rows of
// the parent level's interface data that imap sends to them.  Followed from the
// root down, every node row ends at the root boundary data or at one interface
// value (the element that the level's dn_g product writes).
int Tree::build_sources() {
    const Plan& P = plan;
    const int L = P.L;
    iface_off.assign(L, 0);
    int off = P.lv[0].P;
    for (int l = 0; l < L; ++l) {
        iface_off[l] = off;
        off += P.lv[l].nodes * 2 * P.lv[l].isz;
    }
    iface_len = off;
    std::vector<int> S(P.lv[0].P);
    for (int r = 0; r < P.lv[0].P; ++r) S[r] = r;
    src_dn.assign(L, {});
    for (int l = 0; l < L; ++l) {
        const Level& V = P.lv[l];
        const int b = V.bc, bn = V.bn, i2 = 2 * V.isz;
        std::vector<int> Sn((size_t)2 * V.nodes * b, -1);
        std::vector<int>& dn = src_dn[l];
        dn.assign((size_t)2 * V.nodes * bn, 0);
        for (int n = 0; n < V.nodes; ++n) {
            const int par = (l == 0) ? 0 : (n & 1);
            for (int c = 0; c < 2; ++c)
                for (int j = 0; j < bn; ++j) {
                    const int v = S[(size_t)n * V.P + V.prow[(size_t)(par * 2 + c) * bn + j]];
                    Sn[(size_t)(2 * n + c) * b + j] = v;
                    dn[(size_t)(2 * n + c) * bn + j] = v;
                }
            for (int q = 0; q < i2; ++q) Sn[(size_t)2 * n * b + V.imap[q]] = iface_off[l] + n * i2 + q;
        }
        for (int v : Sn)
            if (v < 0) return fail("Error: internal: a node row without a source in the down pass.");
        S.swap(Sn);
    }
    src_leaf.swap(S);
    // rows read in 16-byte pairs by the direct-to-LDS GEMM tiles (with bn even,
    // the pairs of every node's map start at even offsets of the table)
    src_dn_pairs.assign(L, 0);
    for (int l = 0; l < L; ++l) {
        const std::vector<int>& dn = src_dn[l];
        bool ok = P.lv[l].bn % 2 == 0;
        for (size_t j = 0; ok && j + 1 < dn.size(); j += 2) ok = dn[j] % 2 == 0 && dn[j + 1] == dn[j] + 1;
        src_dn_pairs[l] = ok ? 1 : 0;
    }
    return 0;
}

int Tree::points_device(double* xyz, int* face, hipStream_t st) {
    if (!planned) return fail("Error: rs_plan must be called first.");
    if (!xyz) return fail("Error: null argument.");
    if (ensure_device()) return 1;
    const Plan& P = plan;
    if (!d_pts_box.p) {
        std::vector<double> box((size_t)P.n_leaves * 6);
        for (int q = 0; q < P.n_leaves; ++q)
            for (int j = 0; j < 3; ++j) {
                box[6 * (size_t)q + j] = P.nodes[P.L][q].box[0][j];
                box[6 * (size_t)q + 3 + j] = P.nodes[P.L][q].box[1][j];
            }
        if (upload(d_pts_box, box) || upload(d_pts_pat, P.point_pattern())) return 1;
    }
    const bool own = !st;
    if (own) st = stream;
    if (own && join_default_in(st)) return 1;
    launch_leaf_points(d_pts_box.as<double>(), d_pts_pat.as<int>(), P.h, P.lv[P.L].ptbox, P.d, P.m, P.n_leaves, xyz,
                       face, st);
    HIPCHK(hipGetLastError());
    return own ? join_default_out(st) : 0;
}

int Tree::factor(const double* k_host) {
    if (!planned) return fail("Error: rs_plan must be called before rs_factor.");
    if (ensure_device()) return 1;
    size_t n = (size_t)plan.n_leaves * plan.n_int;
    DevBuf tmp;
    if (alloc(tmp, n * sizeof(double2))) return 1;
    HIPCHK(hipMemcpy(tmp.p, k_host, n * sizeof(double2), hipMemcpyHostToDevice));
    return factor_device(tmp.as<double2>());
}

// Scratch of one level's merge, carved from the tree's merge arena (offsets in
// bytes, 256-aligned): a = zeta I - T1_II, c = zeta I - T2_II, W (inverted in
// place), Y, the K-split workspace of the level's products, and the
// Gauss-Jordan panel and pivots of large interfaces.
struct MergeLayout {
    size_t A = 0, C = 0, W = 0, Y = 0, ws = 0, GW = 0, piv = 0, total = 0;
    InvPlan ip;
};

// When the root's map is skipped, the root merge reads of each of its children's
// maps only the rows and columns on the root interface (T_II for a and c, T_IN
// and T_NI kept for the solve); the rest (T_NN, 44-56% of the map at 1024^2 and
// 64^3) is never read.  Level 1 then forms its nodes' maps in two kinds of
// blocks: a child's rows that land on the root interface (one run [a, a + m) of
// its bn rows) against all columns, and the child's other rows, in runs of m,
// against the columns on the root interface.  Returns m, with a per level-1
// node (the same for both children), or 0 where the rows are not one run or the
// runs do not tile bn.
static int root_child_split(const Plan& P, std::vector<int>& a) {
    if (P.L < 2) return 0;
    const Level& V0 = P.lv[0];
    const Level& V1 = P.lv[1];
    const int b = V0.bc, bn = V1.bn;
    std::vector<char> isI((size_t)2 * b, 0);
    for (int q = 0; q < 2 * V0.isz; ++q) isI[V0.imap[q]] = 1;
    int m = 0;
    a.assign(V1.nodes, 0);
    for (int n = 0; n < V1.nodes; ++n) {
        const int par = n & 1;
        for (int c = 0; c < 2; ++c) {
            int first = -1, cnt = 0, runs = 0;
            bool prev = false;
            for (int r = 0; r < bn; ++r) {
                const bool in = isI[(size_t)n * b + V1.prow[(size_t)(par * 2 + c) * bn + r]] != 0;
                if (in) { ++cnt; if (first < 0) first = r; if (!prev) ++runs; }
                prev = in;
            }
            if (runs != 1 || (m && cnt != m) || (c == 1 && first != a[n])) return 0;
            m = cnt;
            a[n] = first;
        }
        if (bn % m || a[n] % m) return 0;
    }
    return m;
}

// The split applies at level 1 of a tree whose root map is skipped and whose
// full maps are not kept; tn2: the problems of the second kind.
static int root_child_split(const Plan& P, const rs_opts& o, std::vector<int>& a, int& tn2) {
    tn2 = 0;
    if (!o.skip_root_rtr || (o.retain & RS_RETAIN_MAPS)) return 0;
    const int m = root_child_split(P, a);
    if (m) tn2 = 4 * P.lv[1].nodes * (P.lv[1].bn / m - 1);
    return m;
}

// with_rtr: the level builds its parents' maps (every level but the root, and
// the root too unless its map is skipped); Y exists only for them.
static MergeLayout merge_layout(const Level& V, size_t lds_limit, bool with_rtr, int tsplit = 0, int tn2 = 0) {
    MergeLayout m;
    const size_t nodes = V.nodes, i = V.isz, bn = V.bn, ii = i * i;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    m.A = take(nodes * ii * 16);
    m.C = take(nodes * ii * 16);
    m.W = take(nodes * ii * 16);
    m.Y = take(with_rtr ? nodes * 2 * i * 2 * bn * 16 : 0);
    size_t wse = 0;
    const int ni = (int)i, nbn = (int)bn;
    wse = std::max(wse, gemm_ws_elems((int)nodes, ni, ni, ni, merge_ks_target(ni)));
    wse = std::max(wse, gemm_ws_elems(2 * (int)nodes, ni, ni, ni, merge_ks_target(ni)));
    if (with_rtr) {
        wse = std::max(wse, gemm_ws_elems(2 * (int)nodes, 2 * ni, nbn, ni, merge_ks_target(nbn)));
        wse = std::max(wse, gemm_ws_elems(4 * (int)nodes, nbn, nbn, ni, merge_ks_target(nbn)));
        if (tsplit) {
            wse = std::max(wse, gemm_ws_elems(4 * (int)nodes, tsplit, nbn, ni, merge_ks_target(nbn)));
            wse = std::max(wse, gemm_ws_elems(tn2, tsplit, tsplit, ni, merge_ks_target(tsplit)));
        }
    }
    m.ws = take(wse * 16);
    m.ip.build((int)i, (int)nodes, ii, lds_limit);
    m.GW = take(m.ip.w_elems() * 16);
    m.piv = take(m.ip.small ? 0 : nodes * i * sizeof(int));
    m.total = off;
    return m;
}

// What factor_device allocates, and which of it the tree keeps (bytes).
struct FactorSizes {
    size_t sinv64 = 0, sinv32 = 0, maps_full = 0, tmap[2] = {0, 0}, invm_t = 0, arena = 0;
    std::vector<size_t> Tn64, Tn32, invm64, invm32, Tfull;   // per level
    long long fdesc = 0;
};

static FactorSizes factor_sizes(const Plan& P, const LeafGeom& G, const rs_opts& o, size_t lds) {
    FactorSizes z;
    const int L = P.L;
    const bool mixed = o.precision == RS_MIXED;
    const bool keep64 = !mixed || (o.retain & RS_RETAIN_FP64);
    const size_t nl = P.n_leaves;
    z.sinv64 = nl * (size_t)G.K * G.s * G.s * 16;
    z.sinv32 = mixed ? z.sinv64 / 2 : 0;
    z.Tn64.assign(L + 1, 0); z.Tn32.assign(L + 1, 0); z.Tfull.assign(L + 1, 0);
    z.invm64.assign(L, 0); z.invm32.assign(L, 0);
    z.Tfull[L] = nl * (size_t)P.nb * P.nb * 16;
    const long long gd = (long long)sizeof(GemmDesc);
    std::vector<int> ta;
    int tn2 = 0;
    const int tm = root_child_split(P, o, ta, tn2);
    for (int l = L - 1; l >= 0; --l) {
        const Level& V = P.lv[l];
        const size_t nodes = V.nodes, i = V.isz, bn = V.bn;
        const size_t compact = 2 * nodes * 2 * bn * i;   // the level's children
        if (keep64) z.Tn64[l + 1] = compact * 16;
        if (mixed) z.Tn32[l + 1] = compact * 8;
        const size_t im = nodes * 4 * i * i * 16;
        if (keep64) z.invm64[l] = im;
        else z.invm_t = std::max(z.invm_t, im);
        if (mixed) z.invm32[l] = im / 2;
        if (l > 0 || !o.skip_root_rtr) z.Tfull[l] = nodes * (size_t)V.P * V.P * 16;
        const MergeLayout m = merge_layout(V, lds, l > 0 || !o.skip_root_rtr, l == 1 ? tm : 0, l == 1 ? tn2 : 0);
        z.arena = std::max(z.arena, m.total);
        if (!m.ip.small) z.fdesc += (long long)m.ip.desc_count() * gd + (long long)m.ip.cmap_elems() * 4;
        z.fdesc += 4 * nodes * gd + ((l > 0 || !o.skip_root_rtr) ? 6 * nodes * gd : 0);
        if (l == 1 && tm) z.fdesc += (long long)tn2 * gd;
    }
    if (!z.arena) z.arena = 256;
    for (int l = 0; l <= L; ++l) {
        z.maps_full += z.Tfull[l];
        z.tmap[l & 1] = std::max(z.tmap[l & 1], z.Tfull[l]);
    }
    return z;
}

int Tree::factor_device(const double2* k) {
    if (!planned) return fail("Error: rs_plan must be called before rs_factor.");
    if (ensure_device()) return 1;
    const double t0 = now_s();
    // A refactor of the same plan overwrites the factor data in place.  When the
    // working memory was kept (RS_RETAIN_SCRATCH) the merge scratch and the
    // product descriptors of the first factorization are reused: no allocation,
    // and no host synchronisation until the end.  The solve workspaces' product
    // descriptors point into the retained solve data, which keeps its addresses.
    factored = false;
    have32 = have64 = false;
    const Plan& P = plan;
    const int L = P.L, nl = P.n_leaves;
    const bool mixed = opts.precision == RS_MIXED;
    const bool keep64 = !mixed || (opts.retain & RS_RETAIN_FP64);
    const bool keep_maps = (opts.retain & RS_RETAIN_MAPS) != 0;
    const bool keep_scratch = (opts.retain & RS_RETAIN_SCRATCH) != 0;
    for (int l = L - 1; l >= 0; --l)
        if (!merge_layout(P.lv[l], lds_limit, true).ip.fits())
            return fail("Error: interface of " + std::to_string(P.lv[l].isz) +
                        " points is larger than the interface inverses support (LDS panel).");
    const FactorSizes z = factor_sizes(P, G, opts, lds_limit);
    bool moved = false;
    auto need = [&](DevBuf& b, size_t bytes) -> int {
        if (!bytes) { if (b.p) { b.release(); moved = true; } return 0; }
        if (b.p && b.bytes == bytes) return 0;
        moved = true;
        return alloc(b, bytes);
    };
    auto needv = [&](std::vector<DevBuf*>& v, int l, size_t bytes) -> int {
        if (!v[l]) v[l] = new DevBuf;
        return need(*v[l], bytes);
    };
    T.resize(L + 1, nullptr); Tn.resize(L + 1, nullptr); Tn32.resize(L + 1, nullptr);
    invM.resize(L, nullptr); invM32.resize(L, nullptr);
    // addresses of the solve data before: the solve workspaces' descriptors and
    // captured graphs point into it
    auto solve_ptrs = [&]() {
        std::vector<const void*> v = {keep64 ? sinv.p : nullptr, sinv32.p};
        for (DevBuf* b : Tn) v.push_back(b ? b->p : nullptr);
        for (DevBuf* b : Tn32) v.push_back(b ? b->p : nullptr);
        for (DevBuf* b : invM) v.push_back(b ? b->p : nullptr);
        for (DevBuf* b : invM32) v.push_back(b ? b->p : nullptr);
        return v;
    };
    const std::vector<const void*> before = solve_ptrs();
    // the potential stays resident (the FP64 residual of the mixed path reads it).
    // An FP64 tree's complex64 copies (made for mixed solves) keep their buffers:
    // make_c64 refills them in place.
    if (need(kint, (size_t)nl * P.n_int * sizeof(double2)) || need(info, sizeof(int)) || need(sinv, z.sinv64) ||
        (mixed && need(sinv32, z.sinv32)) || need(invM_t, z.invm_t) || need(fscratch, z.arena))
        return 1;
    for (int l = 0; l <= L; ++l) {
        if (needv(Tn, l, z.Tn64[l]) || (mixed && needv(Tn32, l, z.Tn32[l])) ||
            needv(T, l, keep_maps ? z.Tfull[l] : 0))
            return 1;
        if (l < L && (needv(invM, l, z.invm64[l]) || (mixed && needv(invM32, l, z.invm32[l])))) return 1;
    }
    for (int q = 0; q < 2; ++q)
        if (need(tmap[q], keep_maps ? 0 : z.tmap[q])) return 1;
    if (solve_ptrs() != before) clear_slots();
    std::vector<MergeLayout> ml(L);
    std::vector<int> ta;
    int tn2 = 0;
    const int tm = root_child_split(P, opts, ta, tn2);
    for (int l = 0; l < L; ++l)
        ml[l] = merge_layout(P.lv[l], lds_limit, l > 0 || !opts.skip_root_rtr, l == 1 ? tm : 0, l == 1 ? tn2 : 0);
    char* ar = fscratch.as<char>();
    // full node maps of level l, and the FP64 inv(M) of level l
    auto Tfull = [&](int l) -> double2* { return keep_maps ? T[l]->as<double2>() : tmap[l & 1].as<double2>(); };
    auto IMf = [&](int l) -> double2* { return keep64 ? invM[l]->as<double2>() : invM_t.as<double2>(); };

    const double2 zeta = make_double2(P.zeta[0], P.zeta[1]);
    const double2 nu = make_double2(P.nu[0], P.nu[1]);
    const double2 nu2 = cmul(nu, nu);
    const double nn = nu.x * nu.x + nu.y * nu.y;
    const double2 rnu = make_double2(nu.x / nn, -nu.y / nn);
    // product descriptors of every level, in one array (kept for refactors while
    // the buffers they point to are)
    if (moved || fdoff.size() != (size_t)L || !fdesc.p) {
        fdoff.assign(L, MergeDescs{});
        size_t ncm = 0;
        for (int l = L - 1; l >= 0; --l)
            if (!ml[l].ip.small) {
                fdoff[l].cm = ncm;
                ncm += ml[l].ip.cmap_elems();
            }
        if (need(fcmap, ncm * sizeof(int))) return 1;
        std::vector<int> colmaps(ncm);
        std::vector<GemmDesc> v;
        for (int l = L - 1; l >= 0; --l) {
            const Level& V = P.lv[l];
            const MergeLayout& m = ml[l];
            MergeDescs& o = fdoff[l];
            const int nodes = V.nodes, i = V.isz, b = V.bc, bn = V.bn, Pp = V.P;
            const size_t ii = (size_t)i * i, tstride = (size_t)b * b;
            const double2* Tc = Tfull(l + 1);
            double2* Am = (double2*)(ar + m.A);
            double2* Cm = (double2*)(ar + m.C);
            double2* Wm = (double2*)(ar + m.W);
            double2* Y = (double2*)(ar + m.Y);
            double2* GW = (double2*)(ar + m.GW);
            double2* IM = IMf(l);
            if (!m.ip.small) {  // the Gauss-Jordan updates of W
                o.gj = v.size();
                v.resize(v.size() + m.ip.desc_count());
                m.ip.make_descs(Wm, GW, fcmap.as<int>() + o.cm, v.data() + o.gj, colmaps.data() + o.cm);
            }
            o.w = v.size();
            for (int n = 0; n < nodes; ++n) {  // W = -a c (nu^2 on the diagonal added after)
                GemmDesc g = desc0();
                g.A = Am + n * ii; g.lda = i;
                g.B = Cm + n * ii; g.ldb = i;
                g.Out = Wm + n * ii; g.ldo = i;
                g.ar = -1.0;
                v.push_back(g);
            }
            o.mb = v.size();
            for (int n = 0; n < nodes; ++n) {  // invM_22 = -Wi a
                GemmDesc g = desc0();
                g.A = Wm + n * ii; g.lda = i;
                g.B = Am + n * ii; g.ldb = i;
                g.Out = IM + n * 4 * ii + i + (size_t)i * 2 * i; g.ldo = 2 * i;
                g.ar = -1.0;
                v.push_back(g);
            }
            for (int n = 0; n < nodes; ++n) {  // invM_11 = -c Wi
                GemmDesc g = desc0();
                g.A = Cm + n * ii; g.lda = i;
                g.B = Wm + n * ii; g.ldb = i;
                g.Out = IM + n * 4 * ii; g.ldo = 2 * i;
                g.ar = -1.0;
                v.push_back(g);
            }
            o.m12 = v.size();
            for (int n = 0; n < nodes; ++n) {  // invM_12 = -(1 / nu) invM_11 a (I / nu added after)
                GemmDesc g = desc0();
                g.A = IM + n * 4 * ii; g.lda = 2 * i;
                g.B = Am + n * ii; g.ldb = i;
                g.Out = IM + n * 4 * ii + (size_t)i * 2 * i; g.ldo = 2 * i;
                g.ar = -rnu.x; g.ai = -rnu.y;
                v.push_back(g);
            }
            o.y = v.size();
            for (int n = 0; n < nodes && (l > 0 || !opts.skip_root_rtr); ++n)  // Y = invM diag(T1_IN, T2_IN)
                for (int c = 0; c < 2; ++c) {
                    GemmDesc g = desc0();
                    g.A = IM + n * 4 * ii + (size_t)c * i * 2 * i; g.lda = 2 * i;
                    g.B = Tc + (size_t)(2 * n + c) * tstride + bn; g.ldb = b;
                    g.Out = Y + (size_t)n * 2 * i * 2 * bn + (size_t)c * bn * 2 * i; g.ldo = 2 * i;
                    v.push_back(g);
                }
            o.t = v.size();
            o.tm = o.tn2 = 0;
            if (l == 1 && tm) {  // parent RtR where the root reads it (rows or columns on its interface)
                const int* prow = d_prow[l]->as<int>();
                o.tm = tm;
                for (int n = 0; n < nodes; ++n)
                    for (int c = 0; c < 2; ++c)
                        for (int c2 = 0; c2 < 2; ++c2) {   // the rows on the root interface, all columns
                            GemmDesc g = desc0();
                            const double2* Tcn = Tc + (size_t)(2 * n + c) * tstride;
                            const int a = ta[n], par = n & 1;
                            g.A = Tcn + (size_t)bn * b + a; g.lda = b;
                            g.B = Y + (size_t)n * 2 * i * 2 * bn + (size_t)c * i + (size_t)c2 * bn * 2 * i;
                            g.ldb = 2 * i;
                            g.Out = Tfull(l) + (size_t)n * Pp * Pp; g.ldo = Pp;
                            g.rmap = prow + (size_t)(par * 2 + c) * bn + a;
                            g.cmap = prow + (size_t)(par * 2 + c2) * bn;
                            if (c == c2) { g.Din = Tcn + a; g.ldd = b; g.br = 1.0; }
                            v.push_back(g);
                        }
                o.t2 = v.size();
                for (int n = 0; n < nodes; ++n)
                    for (int c = 0; c < 2; ++c)
                        for (int c2 = 0; c2 < 2; ++c2)
                            for (int r0 = 0; r0 < bn; r0 += tm) {   // the other rows, the columns on the root interface
                                const int a = ta[n], par = n & 1;
                                if (r0 == a) continue;
                                GemmDesc g = desc0();
                                const double2* Tcn = Tc + (size_t)(2 * n + c) * tstride;
                                g.A = Tcn + (size_t)bn * b + r0; g.lda = b;
                                g.B = Y + (size_t)n * 2 * i * 2 * bn + (size_t)c * i + (size_t)c2 * bn * 2 * i +
                                      (size_t)a * 2 * i;
                                g.ldb = 2 * i;
                                g.Out = Tfull(l) + (size_t)n * Pp * Pp; g.ldo = Pp;
                                g.rmap = prow + (size_t)(par * 2 + c) * bn + r0;
                                g.cmap = prow + (size_t)(par * 2 + c2) * bn + a;
                                if (c == c2) { g.Din = Tcn + r0 + (size_t)a * b; g.ldd = b; g.br = 1.0; }
                                v.push_back(g);
                            }
                o.tn2 = (int)(v.size() - o.t2);
            } else if (l > 0 || !opts.skip_root_rtr) {  // parent RtR, scattered into the parent's node order
                const int* prow = d_prow[l]->as<int>();
                for (int n = 0; n < nodes; ++n) {
                    const int par = (l == 0) ? 0 : (n & 1);
                    for (int c = 0; c < 2; ++c)
                        for (int c2 = 0; c2 < 2; ++c2) {
                            GemmDesc g = desc0();
                            const double2* Tcn = Tc + (size_t)(2 * n + c) * tstride;
                            g.A = Tcn + (size_t)bn * b; g.lda = b;
                            g.B = Y + (size_t)n * 2 * i * 2 * bn + (size_t)c * i + (size_t)c2 * bn * 2 * i;
                            g.ldb = 2 * i;
                            g.Out = Tfull(l) + (size_t)n * Pp * Pp; g.ldo = Pp;
                            g.rmap = prow + (size_t)(par * 2 + c) * bn;
                            g.cmap = prow + (size_t)(par * 2 + c2) * bn;
                            if (c == c2) { g.Din = Tcn; g.ldd = b; g.br = 1.0; }
                            v.push_back(g);
                        }
                }
            }
        }
        // the column maps go into the buffer the descriptors point to; upload()
        // waits for both copies
        if (ncm) HIPCHK(hipMemcpyAsync(fcmap.p, colmaps.data(), ncm * sizeof(int), hipMemcpyHostToDevice, stream));
        if (upload(fdesc, v)) return 1;
    }
    // leaves: block inverses and RtR maps (k may come from the default stream)
    if (join_default_in(stream)) return 1;
    HIPCHK(hipMemcpyAsync(kint.p, k, (size_t)nl * P.n_int * sizeof(double2), hipMemcpyDeviceToDevice, stream));
    HIPCHK(hipMemsetAsync(info.p, 0, sizeof(int), stream));
    const int c64_from = (mixed && opts.merge_fp64_levels >= 0) ? opts.merge_fp64_levels : L;
    // merges, bottom-up; every level of one launch sequence on the tree's stream.
    // The merge products are idempotent (each writes a buffer it does not read;
    // the diagonal terms of W and invM_12 are added by kernels of their own),
    // so the first factorization of a plan times their tile shapes and keeps the
    // fastest; later factorizations reuse the choice (results do not depend on it).
    const bool tune = opts.autotune && (int)ftile.size() != 6 * L;
    if (tune) ftile.assign(6 * L, -1);
    hipEvent_t te0 = nullptr, te1 = nullptr;
    if (tune) {
        HIPCHK(hipEventCreate(&te0));
        HIPCHK(hipEventCreate(&te1));
    }
    // In a mixed-precision tree with merge_fp64_levels = k >= 0 the merge
    // products of the levels below the top k run in complex64 on the matrix
    // cores (FP32 accumulation, complex128 operands and results); the interface
    // inverses around them stay FP64 with pivoting.
    merge_c64 = c64_from < L;
    auto gemm = [&](int lvl, const GemmDesc* d, int np, int M, int N, int K, void* gws, int t) {
        if (lvl >= c64_from)
            launch_gemm_c64(d, np, M, N, K, stream, gws, t, merge_ks_target(N));
        else
            launch_gemm<double>(d, np, M, N, K, stream, gws, t, merge_ks_target(N));
    };
    auto fgemm = [&](int site, const GemmDesc* d, int np, int M, int N, int K, void* gws) -> int {
        int tile = site < (int)ftile.size() ? ftile[site] : -1;
        if (tune) {
            float best = 0.f;
            for (int t = 0; t < GEMM_TILES; ++t) {
                if (!gemm_tile_distinct(t, N, site / 6 < c64_from, site / 6 < c64_from)) continue;
                (void)hipEventRecord(te0, stream);
                gemm(site / 6, d, np, M, N, K, gws, t);
                (void)hipEventRecord(te1, stream);
                if (hipEventSynchronize(te1) != hipSuccess) return fail("Error: GEMM autotuning failed.");
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, te0, te1);
                if (t == 0 || ms < best) { best = ms; tile = t; }
            }
            ftile[site] = tile;
        }
        gemm(site / 6, d, np, M, N, K, gws, tile);
        return 0;
    };
    // The launches from the leaves to the last merge, as one sequence.  A refactor
    // in place (RS_RETAIN_SCRATCH, nothing reallocated, the tiles tuned) replays
    // it from a graph captured on the first such refactor: about 650 launches,
    // most of them short (the Gauss-Jordan panels), whose gaps the graph removes.
    // The graph reads k through kint and clears nothing (info is cleared above).
    auto launch_seq = [&]() -> int {
        launch_leaf_blocks(G, kint.as<double2>(), sinv.as<double2>(), info.as<int>(), nl, stream);
        if (mixed) launch_to_c64(sinv32.as<float2>(), sinv.as<double2>(), z.sinv64 / 16, stream);
        // The leaf maps feed the bottom merge.  When that merge runs in complex64
        // (a mixed tree with complex64 merges down to the leaves) the maps come from
        // the complex64 block inverses on the matrix cores (one workgroup of 64
        // columns per leaf); otherwise they are FP64.
        if (L > 0 && c64_from <= L - 1 && leaf_mfma_factor_ok(G)) {
            LeafArgs<float> a;
            a.mode = LEAF_FACTOR_T; a.ncols = P.nb; a.n_leaves = nl;
            a.sinv = sinv32.as<float2>(); a.T = Tfull(L);
            launch_leaf_apply<float>(G, a, lds_limit, stream);
        } else {
            LeafArgs<double> a;
            a.mode = LEAF_FACTOR_T; a.ncols = P.nb; a.n_leaves = nl;
            a.sinv = sinv.as<double2>(); a.T = Tfull(L);
            launch_leaf_apply<double>(G, a, lds_limit, stream);
        }
        HIPCHK(hipGetLastError());
        for (int l = L - 1; l >= 0; --l) {
            const Level& V = P.lv[l];
            const MergeLayout& m = ml[l];
            const int nodes = V.nodes, i = V.isz, b = V.bc, bn = V.bn;
            const size_t ii = (size_t)i * i, tstride = (size_t)b * b;
            const double2* Tc = Tfull(l + 1);
            double2* Am = (double2*)(ar + m.A);
            double2* Cm = (double2*)(ar + m.C);
            double2* Wm = (double2*)(ar + m.W);
            double2* GW = (double2*)(ar + m.GW);
            int* piv = (int*)(ar + m.piv);
            void* gws = ar + m.ws;
            double2* IM = IMf(l);
            const MergeDescs& o = fdoff[l];
            const GemmDesc* fd = fdesc.as<GemmDesc>();
            // a = zeta I - T1_II, c = zeta I - T2_II
            launch_form_iface(Am, Tc, tstride, 2, 0, b, bn, i, nodes, zeta, make_double2(0, 0), nullptr, stream);
            launch_form_iface(Cm, Tc, tstride, 2, 1, b, bn, i, nodes, zeta, make_double2(0, 0), nullptr, stream);
            // Wi = (nu^2 I - a c)^-1
            if (fgemm(6 * l + 0, fd + o.w, nodes, i, i, i, gws)) return 1;
            launch_add_diag(Wm, ii, i, i, nodes, nu2, stream);
            launch_inverse(m.ip, Wm, piv, GW, fd + o.gj, info.as<int>(), stream);
            // invM = [-c Wi, I / nu - (1 / nu) (-c Wi) a; nu Wi, -Wi a]
            if (fgemm(6 * l + 1, fd + o.mb, 2 * nodes, i, i, i, gws) || fgemm(6 * l + 2, fd + o.m12, nodes, i, i, i, gws))
                return 1;
            launch_add_diag(IM + (size_t)i * 2 * i, 4 * ii, 2 * i, i, nodes, rnu, stream);
            launch_copy_blocks(IM + i, 4 * ii, 2 * i, Wm, ii, i, i, i, nodes, nu, stream);
            // Y = invM diag(T1_IN, T2_IN); parent T = diag(T_NN) + diag(T_NI) Y
            // (neither at a root whose map is skipped: nothing reads them)
            if ((l > 0 || !opts.skip_root_rtr) && fgemm(6 * l + 3, fd + o.y, 2 * nodes, 2 * i, bn, i, gws)) return 1;
            if (o.tm) {
                if (fgemm(6 * l + 4, fd + o.t, 4 * nodes, o.tm, bn, i, gws) ||
                    fgemm(6 * l + 5, fd + o.t2, o.tn2, o.tm, o.tm, i, gws))
                    return 1;
            } else if ((l > 0 || !opts.skip_root_rtr) && fgemm(6 * l + 4, fd + o.t, 4 * nodes, bn, bn, i, gws)) {
                return 1;
            }
            // the solve data of this merge: inv(M), and the children's T_NI / T_IN,
            // before the next level reuses the buffers they come from
            if (mixed) launch_to_c64(invM32[l]->as<float2>(), IM, (size_t)nodes * 4 * ii, stream);
            launch_extract_ni_in(Tc, tstride, b, bn, i, 2 * nodes, keep64 ? Tn[l + 1]->as<double2>() : nullptr,
                                 mixed ? Tn32[l + 1]->as<float2>() : nullptr, stream);
            HIPCHK(hipGetLastError());
        }
        return 0;
    };
    const FactorKey key = {mixed, keep64, keep_maps, opts.skip_root_rtr != 0, c64_from};
    const bool use_graph = opts.graphs && keep_scratch && !moved && !tune;
    if (!use_graph || !(fkey == key)) drop_factor_graph();
    if (!use_graph) {
        if (launch_seq()) return 1;
    } else {
        if (!fgraph) {
            HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            const int rc = launch_seq();
            hipGraph_t gr = nullptr;
            const hipError_t e = hipStreamEndCapture(stream, &gr);
            if (rc || e != hipSuccess) {
                if (gr) (void)hipGraphDestroy(gr);
                return rc ? 1 : fail(std::string("HIP error: stream capture: ") + hipGetErrorString(e));
            }
            const hipError_t e2 = hipGraphInstantiate(&fgraph, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e2 != hipSuccess) {
                fgraph = nullptr;
                return fail(std::string("HIP error: graph instantiation: ") + hipGetErrorString(e2));
            }
            fkey = key;
        }
        HIPCHK(hipGraphLaunch(fgraph, stream));
    }
    if (te0) (void)hipEventDestroy(te0);
    if (te1) (void)hipEventDestroy(te1);
    int hinfo = 0;
    HIPCHK(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (hinfo & 1) return fail("Error: singular leaf block in the leaf factorization.");
    if (hinfo & 2) return fail("Error: singular interface matrix in the merge.");
    // the working memory goes unless it is to be kept for refactors
    if (!keep_scratch) {
        drop_factor_graph();
        fscratch.release(); invM_t.release(); tmap[0].release(); tmap[1].release();
        if (!keep64) sinv.release();
        fdesc.release(); fcmap.release();
        fdoff.clear();
    }
    factored = true;
    have64 = keep64;
    have32 = mixed;
    f_time = now_s() - t0;
    return 0;
}

// The complex64 solve data of a tree factored in FP64 (made on its first mixed
// solve): the same blocks rounded, in place when the buffers exist.
int Tree::make_c64() {
    const Plan& P = plan;
    if (!have64) return fail("Error: the tree holds no FP64 solve data to round.");
    auto conv = [&](DevBuf& dst, const DevBuf& src) -> int {
        size_t n = src.bytes / sizeof(double2);
        if (ensure_buf(dst, n * sizeof(float2))) return 1;
        launch_to_c64(dst.as<float2>(), src.as<double2>(), n, stream);
        return 0;
    };
    Tn32.resize(P.L + 1, nullptr);
    invM32.resize(P.L, nullptr);
    if (conv(sinv32, sinv)) return 1;
    for (int l = 1; l <= P.L; ++l) { if (!Tn32[l]) Tn32[l] = new DevBuf; if (conv(*Tn32[l], *Tn[l])) return 1; }
    for (int l = 0; l < P.L; ++l) { if (!invM32[l]) invM32[l] = new DevBuf; if (conv(*invM32[l], *invM[l])) return 1; }
    HIPCHK(hipStreamSynchronize(stream));
    have32 = true;
    return 0;
}

WorkLayout Tree::work_layout(int cap, int esz, bool refine, bool tabs, bool descs) const {
    const Plan& P = plan;
    const int L = P.L;
    WorkLayout lay;
    lay.cap = cap < 1 ? 1 : cap;
    lay.esz = esz;
    lay.refine = refine;
    cap = lay.cap;
    size_t off = 0;
    auto take = [&](size_t bytes) { const size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    // U: the root boundary data and every level's interface data, one column
    // (iface_len elements) per right-hand side; G[l] is level l's slice of it
    lay.hV.assign(L + 1, WorkLayout::NONE); lay.gF.assign(L + 1, WorkLayout::NONE);
    lay.gF[0] = take((size_t)cap * iface_len * esz);
    for (int l = 0; l <= L; ++l)
        if (l > 0 || L == 0) lay.hV[l] = take((size_t)cap * P.lv[l].nodes * P.lv[l].P * esz);
    lay.G.assign(L, WorkLayout::NONE); lay.Z.assign(L, WorkLayout::NONE);
    lay.dug.assign(L, WorkLayout::NONE); lay.duh.assign(L, WorkLayout::NONE);
    lay.ddt.assign(L, WorkLayout::NONE); lay.ddg.assign(L, WorkLayout::NONE);
    size_t wse = 0;
    for (int l = 0; l < L; ++l) {
        const Level& Vl = P.lv[l];
        const size_t n = (size_t)cap * Vl.nodes * 2 * Vl.isz * esz;
        lay.G[l] = lay.gF[0] + (size_t)iface_off[l] * esz;
        lay.Z[l] = take(n);
        wse = std::max(wse, gemm_ws_elems(Vl.nodes, 2 * Vl.isz, cap, 2 * Vl.isz,
                                          solve_ks_target(Vl.nodes, 2 * Vl.isz, 2 * Vl.isz)));
        wse = std::max(wse, gemm_ws_elems(2 * Vl.nodes, Vl.bn, cap, Vl.isz, solve_ks_target(2 * Vl.nodes, Vl.bn, Vl.isz)));
        wse = std::max(wse, gemm_ws_elems(2 * Vl.nodes, Vl.isz, cap, Vl.bn, solve_ks_target(2 * Vl.nodes, Vl.isz, Vl.bn)));
    }
    if (wse) lay.ws = take(wse * esz);
    for (int l = 0; l < L && descs; ++l) {
        const size_t nd = P.lv[l].nodes, gd = sizeof(GemmDesc);
        lay.dug[l] = take(nd * gd);
        if (l > 0) lay.duh[l] = take(2 * nd * gd);
        lay.ddt[l] = take(2 * nd * gd);
        lay.ddg[l] = take(nd * gd);
    }
    if (refine) {
        const size_t nl = P.n_leaves;
        lay.rf = take((size_t)cap * nl * P.n_int * 8);
        lay.rg = take((size_t)cap * nl * P.nb * 8);
        lay.rj = take((size_t)cap * nl * P.nb * 8);
        lay.uA = take((size_t)cap * nl * P.m * 8);
    }
    if (tabs) lay.tabs = take((size_t)3 * cap * sizeof(void*));
    lay.total = off;
    return lay;
}

template <typename R> SolveLevel Tree::solve_level(const SolveWork<R>& w, int l) const {
    const Plan& P = plan;
    const Level& Vl = P.lv[l];
    const Level& C = P.lv[l + 1];
    const std::vector<DevBuf*>& Tt = std::is_same<R, double>::value ? Tn : Tn32;
    const std::vector<DevBuf*>& IMt = std::is_same<R, double>::value ? invM : invM32;
    SolveLevel p;
    p.IM = (const char*)IMt[l]->p;
    p.Tc = (const char*)Tt[l + 1]->p;
    p.hC = (char*)w.hV[l + 1];
    p.hP = l > 0 ? (char*)w.hV[l] : nullptr;
    p.Gl = (char*)w.G[l];
    p.Zl = (char*)w.Z[l];
    p.U = (char*)w.gF[0];
    p.prow = d_prow[l]->as<int>();
    p.dsrc = d_src_dn[l]->as<int>();
    p.dpair = src_dn_pairs[l];
    p.nodes = Vl.nodes; p.i = Vl.isz; p.b = Vl.bc; p.bn = Vl.bn; p.Pp = Vl.P;
    p.ldc = C.nodes * Vl.bc; p.ldg = iface_len; p.ldz = Vl.nodes * 2 * Vl.isz; p.ldp = Vl.nodes * Vl.P;
    p.lvl = l;
    p.esz = (int)sizeof(c_t<R>);
    return p;
}

long long Tree::lane_bytes(int cap) const {
    const bool mixed = opts.precision == RS_MIXED;
    return (long long)work_layout(cap, mixed ? 8 : 16, mixed && opts.n_refine > 0, false, false).total;
}

long long Tree::lane_desc_bytes() const {
    const WorkLayout d = work_layout(1, 8, false, false, true);
    const WorkLayout n = work_layout(1, 8, false, false, false);
    return d.total > n.total ? (long long)(d.total - n.total) : 256;   // (a tree without levels: one small buffer)
}

// Slot slot_id's workspace (at the tree's precision, cap columns) inside an
// arena the caller owns -- a server lane's, shared by every tree the lane runs.
// The product descriptors, which point into the arena and into this tree's
// solve data, live in the slot's own small buffer; they are written on stream
// st (before the pass that reads them) whenever the arena or its generation
// changed since the last binding.
int Tree::bind_slot(int slot_id, char* base, size_t bytes, unsigned long long gen, int cap, hipStream_t st) {
    if (!factored) return fail("Error: Robin Tree must be factored first.");
    const bool mixed = opts.precision == RS_MIXED;
    if (mixed && !have32 && make_c64()) return 1;
    const bool refine = mixed && opts.n_refine > 0;
    const WorkLayout lay = work_layout(cap, mixed ? 8 : 16, refine, false, false);
    if (lay.total > bytes) return fail("Error: the lane's workspace is smaller than this tree needs.");
    Slot& sl = slot(slot_id);
    auto bind = [&](auto& w) -> int {
        if (sl.ext_base == base && sl.ext_gen == gen && w.cap == lay.cap && w.refine == refine && !w.arena.p)
            return 0;
        w.clear();
        ++w.gen;
        sl.drop_graphs();
        w.bind(base, lay);
        // descriptor arrays of every level, consecutive in the slot's buffer
        // (256-byte aligned, as in a workspace of its own: lane_desc_bytes bytes)
        const Plan& P = plan;
        std::vector<size_t> off(4 * P.L, 0);
        size_t tot = 0;
        auto take = [&](size_t bytes) { const size_t o = tot; tot += (bytes + 255) & ~(size_t)255; return o; };
        for (int l = 0; l < P.L; ++l) {
            const size_t nd = P.lv[l].nodes * sizeof(GemmDesc);
            off[4 * l] = take(nd);
            off[4 * l + 1] = l > 0 ? take(2 * nd) : 0;
            off[4 * l + 2] = take(2 * nd);
            off[4 * l + 3] = take(nd);
        }
        if (ensure_buf(sl.descs, tot ? tot : 256)) return 1;
        char* db = sl.descs.as<char>();
        for (int l = 0; l < P.L; ++l) {
            w.d_up_g[l] = (GemmDesc*)(db + off[4 * l]);
            w.d_up_h[l] = l > 0 ? (GemmDesc*)(db + off[4 * l + 1]) : nullptr;
            w.d_dn_t[l] = (GemmDesc*)(db + off[4 * l + 2]);
            w.d_dn_g[l] = (GemmDesc*)(db + off[4 * l + 3]);
            const SolveLevel p = solve_level(w, l);
            hipLaunchKernelGGL(solve_desc_kernel, dim3((p.nodes + 255) / 256), dim3(256), 0, st, p, w.d_up_g[l],
                               w.d_up_h[l], w.d_dn_t[l], w.d_dn_g[l]);
        }
        HIPCHK(hipGetLastError());
        auto it = tile_cache.find(2LL * lay.cap + (mixed ? 0 : 1));
        for (int c = 0; c < 4; ++c) w.tile[c] = it != tile_cache.end() ? it->second[c] : std::vector<int>(P.L, -1);
        sl.ext_base = base;
        sl.ext_gen = gen;
        return 0;
    };
    return mixed ? bind(sl.w32) : bind(sl.w64);
}

template <typename R> int Tree::ensure_work(SolveWork<R>& w, int nrhs, bool refine) {
    if (w.cap >= nrhs && (w.refine || !refine)) return 0;
    refine = refine || w.refine;
    w.clear();
    ++w.gen;
    const Plan& P = plan;
    const int L = P.L;
    const WorkLayout lay = work_layout(nrhs, (int)sizeof(c_t<R>), refine, false);
    if (alloc(w.arena, lay.total)) return 1;
    w.bind((char*)w.arena.p, lay);
    // descriptors: built on the host, one upload (they are contiguous in the layout)
    if (L > 0) {
        const size_t lo = lay.dug[0];
        const size_t hi = lay.ddg[L - 1] + (size_t)P.lv[L - 1].nodes * sizeof(GemmDesc);
        std::vector<GemmDesc> hd((hi - lo + sizeof(GemmDesc) - 1) / sizeof(GemmDesc));
        char* hbase = (char*)hd.data() - lo;   // host image of the arena's descriptor range
        for (int l = 0; l < L; ++l) {
            const SolveLevel p = solve_level(w, l);
            GemmDesc* ug = (GemmDesc*)(hbase + lay.dug[l]);
            GemmDesc* dt = (GemmDesc*)(hbase + lay.ddt[l]);
            GemmDesc* dg = (GemmDesc*)(hbase + lay.ddg[l]);
            GemmDesc* uh = l > 0 ? (GemmDesc*)(hbase + lay.duh[l]) : nullptr;
            for (int n = 0; n < p.nodes; ++n) {
                GemmDesc h[2];
                solve_descs(p, n, ug + n, h, dt + 2 * n, dg + n);
                if (uh) { uh[2 * n] = h[0]; uh[2 * n + 1] = h[1]; }
            }
        }
        HIPCHK(hipMemcpyAsync((char*)w.arena.p + lo, hd.data(), hi - lo, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    const int cap = lay.cap;
    const long long key = 2LL * cap + (std::is_same<R, double>::value ? 1 : 0);
    auto it = tile_cache.find(key);
    if (it != tile_cache.end()) {
        for (int c = 0; c < 4; ++c) w.tile[c] = it->second[c];
        return 0;
    }
    for (auto& t : w.tile) t.assign(L, -1);
    if (opts.autotune && tune_work(w)) return 1;
    std::array<std::vector<int>, 4>& e = tile_cache[key];
    for (int c = 0; c < 4; ++c) e[c] = w.tile[c];
    return 0;
}

int Tree::prepare_tiles(int cap) {
    if (!factored || !opts.autotune) return 0;
    const bool mixed = opts.precision == RS_MIXED;
    const long long key = 2LL * cap + (mixed ? 0 : 1);
    if (tile_cache.count(key)) return 0;
    int rc;
    if (mixed) {
        SolveWork<float> w;
        rc = ensure_work(w, cap);
        w.clear();
    } else {
        SolveWork<double> w;
        rc = ensure_work(w, cap);
        w.clear();
    }
    if (rc) {
        // no room for a workspace now: the tile rule, and no later tuning
        std::array<std::vector<int>, 4>& e = tile_cache[key];
        for (auto& v : e) v.assign(plan.L, -1);
        err.clear();
    }
    return 0;
}

template <typename R> int Tree::tune_work(SolveWork<R>& w) {
    // The products are timed where they run: in the order of a pass (up_g and
    // up_h per level from the leaves up, then dn_t and dn_g per level from the
    // root down), every product of a sequence with one tile choice, events
    // between the launches.  Each product then meets the caches as in a solve:
    // its factor operand comes from HBM (the other levels' data went through
    // the caches since its last use) and the vectors it reads were just
    // written by the product before it.  (Timed alone and repeated, a product
    // finds its operands in the caches, which favours tiles that keep few
    // loads in flight.)  Two sequences per choice, the second timed; per
    // product the fastest choice is kept.  The products run on the workspace's
    // own buffers, whose contents every solve overwrites before reading them.
    const Plan& P = plan;
    const int L = P.L, N = w.cap;
    struct Site { int kind, l; const GemmDesc* d; int np, M, K; };
    std::vector<Site> seq;
    for (int l = L - 1; l >= 0; --l) {
        const Level& V = P.lv[l];
        seq.push_back({0, l, w.d_up_g[l], V.nodes, 2 * V.isz, 2 * V.isz});
        if (l > 0) seq.push_back({1, l, w.d_up_h[l], 2 * V.nodes, V.bn, V.isz});
    }
    for (int l = 0; l < L; ++l) {
        const Level& V = P.lv[l];
        seq.push_back({2, l, w.d_dn_t[l], 2 * V.nodes, V.isz, V.bn});
        seq.push_back({3, l, w.d_dn_g[l], V.nodes, 2 * V.isz, 2 * V.isz});
    }
    const size_t ns = seq.size();
    std::vector<hipEvent_t> ev(ns + 1, nullptr);
    int rc = 0;
    for (hipEvent_t& e : ev)
        if (hipEventCreate(&e) != hipSuccess) { e = nullptr; rc = fail("Error: hipEventCreate failed."); }
    std::vector<float> best(ns, 0.f);
    std::vector<int> bt(ns, -1);
    for (int t = 0; t < GEMM_TILES && !rc; ++t) {
        if (!gemm_tile_distinct(t, N, std::is_same<R, double>::value)) continue;
        for (int rep = 0; rep < 2; ++rep) {
            for (size_t j = 0; j < ns; ++j) {
                const Site& s = seq[j];
                if (rep) (void)hipEventRecord(ev[j], stream);
                launch_gemm<R>(s.d, s.np, s.M, N, s.K, stream, w.ws, t, solve_ks_target(s.np, s.M, s.K));
            }
            if (rep) (void)hipEventRecord(ev[ns], stream);
        }
        if (hipEventSynchronize(ev[ns]) != hipSuccess) { rc = fail("Error: GEMM autotuning failed."); break; }
        for (size_t j = 0; j < ns; ++j) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev[j], ev[j + 1]);
            if (bt[j] < 0 || ms < best[j]) { best[j] = ms; bt[j] = t; }
        }
    }
    if (!rc)
        for (size_t j = 0; j < ns; ++j) w.tile[seq[j].kind][seq[j].l] = bt[j];
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

long long Tree::work_bytes(int cap, int prec) const {
    // the workspace arena, plus the column tables of a contiguous-array call
    const bool refine = prec == RS_MIXED && opts.n_refine > 0;
    return (long long)work_layout(cap, prec == RS_MIXED ? 8 : 16, refine, false).total +
           3LL * cap * (long long)sizeof(void*);
}

void Tree::factor_bytes(long long* resident_out, long long* peak_out) const {
    // mirrors factor_device: everything is allocated up front; the working memory
    // is released at the end unless RS_RETAIN_SCRATCH keeps it
    const Plan& P = plan;
    const int L = P.L;
    const bool mixed = opts.precision == RS_MIXED;
    const bool keep64 = !mixed || (opts.retain & RS_RETAIN_FP64);
    const bool keep_maps = (opts.retain & RS_RETAIN_MAPS) != 0;
    const FactorSizes z = factor_sizes(P, G, opts, device_ready ? lds_limit : 163840);
    long long res = 0;
    for (int l = 0; l < L; ++l) res += (long long)(P.lv[l].prow.size() + src_dn[l].size()) * 4;
    res += (long long)src_leaf.size() * 4;
    res += (long long)(P.root_leaf.size() + P.root_loc.size() + P.leaf_nbr.size()) * 4;
    res += (long long)P.n_leaves * P.n_int * 16 + 4;   // k, info
    res += (long long)z.sinv32 + (keep64 ? (long long)z.sinv64 : 0);
    for (int l = 0; l <= L; ++l) res += (long long)(z.Tn64[l] + z.Tn32[l]);
    for (int l = 0; l < L; ++l) res += (long long)(z.invm64[l] + z.invm32[l]);
    if (keep_maps) res += (long long)z.maps_full;
    long long work = (long long)z.arena + (long long)z.invm_t + z.fdesc + (keep64 ? 0 : (long long)z.sinv64) +
                     (keep_maps ? 0 : (long long)(z.tmap[0] + z.tmap[1]));
    if (opts.retain & RS_RETAIN_SCRATCH) { res += work; work = 0; }
    if (resident_out) *resident_out = res;
    if (peak_out) *peak_out = res + work;
}

void Tree::pbeg(int cat, hipStream_t st) {
    if (!prof_on) return;
    ProfRec r;
    r.cat = cat;
    for (hipEvent_t* e : {&r.a, &r.b}) {
        if (ev_pool.empty()) { (void)hipEventCreate(e); }
        else { *e = ev_pool.back(); ev_pool.pop_back(); }
    }
    r.bytes = r.flops = 0;
    (void)hipEventRecord(r.a, st);
    prof.push_back(r);
}

void Tree::pend(double bytes, double flops, hipStream_t st) {
    if (!prof_on || prof.empty()) return;
    ProfRec& r = prof.back();
    r.bytes = bytes;
    r.flops = flops;
    (void)hipEventRecord(r.b, st);
}

int Tree::prof_collect() {
    for (ProfRec& r : prof) {
        HIPCHK(hipEventSynchronize(r.b));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
        prof_ms[r.cat] += ms;
        prof_n[r.cat] += 1;
        prof_bytes[r.cat] += r.bytes;
        prof_flops[r.cat] += r.flops;
        ev_pool.push_back(r.a);
        ev_pool.push_back(r.b);
    }
    prof.clear();
    return 0;
}

template <typename R>
int Tree::solve_pass(SolveWork<R>& w, int nrhs, const PassIO& io, hipStream_t st) {
    using V = c_t<R>;
    const Plan& P = plan;
    const int L = P.L, nl = P.n_leaves;
    const V* sv = std::is_same<R, double>::value ? (const V*)sinv.p : (const V*)sinv32.p;
    const double e = sizeof(V), nr = nrhs;
    const double ef = io.f64 ? 16.0 : (io.f32 ? 8.0 : 0.0);
    const double eu = (io.u64 ? 16.0 : 8.0) + (io.ubase32 ? 8.0 : (io.uacc ? 16.0 : 0.0));
    // algorithmic traffic of a leaf pass: the block inverses once, plus the
    // per-RHS vectors (f in; h out / g in, u out, u base in)
    const double sinv_b = (double)nl * G.K * G.s * G.s * e;
    const double leaf_fl = (double)nl * nr * (2.0 * G.K * G.s * G.s * 8.0);
    LeafArgs<R> up;
    up.mode = LEAF_UP; up.ncols = nrhs; up.n_leaves = nl; up.sinv = sv;
    up.f64 = io.f64; up.f32 = io.f32; up.jump = io.jump; up.hout = w.hV[L];
    pbeg(P_LEAF_UP, st);
    launch_leaf_apply<R>(G, up, lds_limit, st);
    pend(sinv_b + nr * nl * (ef * P.n_int + e * P.nb + (io.jump ? 8.0 * P.nb : 0.0)), leaf_fl, st);
    auto gemm = [&](int cat, int lvl, const GemmDesc* d, int np, int M, int N, int K, bool din) {
        const std::vector<int>& tiles = w.tile[cat - P_GEMM_UP_G];   // up_g, up_h, dn_t, dn_g
        pbeg(cat, st);
        launch_gemm<R>(d, np, M, N, K, st, w.ws, lvl < (int)tiles.size() ? tiles[lvl] : -1, solve_ks_target(np, M, K));
        pend((double)np * e * ((double)M * K + (double)K * N + (double)M * N * (din ? 2 : 1)),
             8.0 * np * (double)M * N * K, st);
    };
    for (int l = L - 1; l >= 0; --l) {
        const Level& Vl = P.lv[l];
        gemm(P_GEMM_UP_G, l, w.d_up_g[l], Vl.nodes, 2 * Vl.isz, nrhs, 2 * Vl.isz, false);
        if (l > 0) gemm(P_GEMM_UP_H, l, w.d_up_h[l], 2 * Vl.nodes, Vl.bn, nrhs, Vl.isz, true);
    }
    pbeg(P_GATHER, st);
    if (io.g32)
        launch_root_gather32<V>(w.gF[0], (size_t)iface_len, io.g32, d_root_leaf.as<int>(), d_root_loc.as<int>(),
                                P.lv[0].P, nl, P.nb, nrhs, st);
    else
        launch_root_gather<V>(w.gF[0], (size_t)iface_len, io.g, d_root_leaf.as<int>(), d_root_loc.as<int>(),
                              P.lv[0].P, P.nb, nrhs, st);
    pend(nr * P.lv[0].P * ((io.g32 ? 8.0 : 16.0) + e), 0, st);
    // (no copies down the tree: dn_t and the leaves read every incoming row
    // where the root gather or a level's dn_g wrote it)
    for (int l = 0; l < L; ++l) {
        const Level& Vl = P.lv[l];
        gemm(P_GEMM_DN_T, l, w.d_dn_t[l], 2 * Vl.nodes, Vl.isz, nrhs, Vl.bn, false);
        gemm(P_GEMM_DN_G, l, w.d_dn_g[l], Vl.nodes, 2 * Vl.isz, nrhs, 2 * Vl.isz, true);
    }
    LeafArgs<R> dn;
    dn.mode = LEAF_DOWN; dn.ncols = nrhs; dn.n_leaves = nl; dn.sinv = sv;
    dn.f64 = io.f64; dn.f32 = io.f32; dn.gin = w.gF[0];
    dn.gsrc = d_src_leaf.as<int>(); dn.gld = (size_t)iface_len;
    dn.u64 = io.u64; dn.u32 = io.u32; dn.ubase32 = io.ubase32; dn.uacc = io.uacc;
    pbeg(P_LEAF_DOWN, st);
    launch_leaf_apply<R>(G, dn, lds_limit, st);
    pend(sinv_b + nr * nl * (ef * P.n_int + e * P.nb + eu * P.m), leaf_fl, st);
    HIPCHK(hipGetLastError());
    return 0;
}

int Tree::slot_enter(Slot& sl, hipStream_t st) {
    // the slot's tables and workspaces may still be in use by a solve queued on
    // another stream: this stream waits for it
    if (sl.last_stream && sl.last_stream != st) HIPCHK(hipStreamWaitEvent(st, sl.done, 0));
    return 0;
}

int Tree::slot_leave(Slot& sl, hipStream_t st) {
    if (!sl.done) HIPCHK(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    HIPCHK(hipEventRecord(sl.done, st));
    sl.last_stream = st;
    return 0;
}

int Tree::solve_device(int nrhs, const double2* f, const double2* g, double2* u, int prec, hipStream_t st) {
    if (!factored) return fail("Error: Robin Tree must be factored first.");
    if (nrhs < 1) return fail("Error: nrhs must be positive.");
    const bool own = !st;
    if (own) st = stream;
    // column tables of the contiguous arrays, built on the stream
    Slot& sl = slot(0);
    if (own && join_default_in(st)) return 1;
    if (slot_enter(sl, st)) return 1;
    if (sl.tabs_cap < nrhs) {
        if (alloc(sl.tabs, (size_t)3 * nrhs * sizeof(void*))) return 1;
        sl.tabs_cap = nrhs;
    }
    double2** tab = sl.tabs.as<double2*>();
    const Plan& P = plan;
    const size_t nl = P.n_leaves;
    launch_col_table(tab, const_cast<double2*>(f), nl * P.n_int, nrhs, st);
    launch_col_table(tab + nrhs, const_cast<double2*>(g), nl * P.nb, nrhs, st);
    launch_col_table(tab + 2 * nrhs, u, nl * P.m, nrhs, st);
    RhsCols c;
    c.f = tab; c.g = tab + nrhs; c.u = tab + 2 * nrhs;
    if (solve_cols(nrhs, c, prec, st, 0, true)) return 1;
    if (slot_leave(sl, st)) return 1;
    return own ? join_default_out(st) : 0;
}

// rs_solve_device with the whole pass workspace inside a caller-side device
// arena (the library's fixed scratch array): it allocates nothing.  The
// workspace is cleared first (the arena keeps whatever the previous call left),
// the column tables and the products' descriptors are written on the device,
// and the launch sequence is the one rs_solve_device runs, so the results are
// the same bits.
int Tree::solve_device_noalloc(int nrhs, const double2* f, const double2* g, double2* u, int prec, hipStream_t st,
                               char* scratch, size_t scratch_bytes) {
    if (!factored) return fail("Error: Robin Tree must be factored first.");
    if (nrhs < 1) return fail("Error: nrhs must be positive.");
    if (!f || !g || !u || !scratch) return fail("Error: null argument.");
    if (prec < 0) prec = opts.precision;
    if (prec != RS_FP64 && prec != RS_MIXED) return fail("Error: precision must be RS_FP64 or RS_MIXED.");
    if (prec == RS_FP64 && merge_c64)
        return fail("Error: an FP64 solve needs a tree factored in FP64 (this tree's merge products are complex64).");
    if (prec == RS_FP64 && !have64)
        return fail("Error: this tree keeps complex64 solve data only (rs_opts.retain RS_RETAIN_FP64 keeps the "
                    "FP64 data for FP64 solves).");
    if (prec == RS_MIXED && !have32)
        return fail("Error: the allocation-free solve needs the tree's complex64 solve data: factor it with "
                    "RS_MIXED (or run one RS_MIXED rs_solve_device first).");
    const bool mixed = prec == RS_MIXED;
    const WorkLayout lay = work_layout(nrhs, mixed ? 8 : 16, mixed && opts.n_refine > 0, true);
    if (lay.total > scratch_bytes)
        return fail("Error: this solve needs " + std::to_string(lay.total) + " bytes of device scratch; the "
                    "allocation-free entry point has " + std::to_string(scratch_bytes) +
                    " (pass fewer right-hand sides per call, or use rs_solve_device).");
    const bool own = !st;
    if (own) st = stream;
    if (own && join_default_in(st)) return 1;
    launch_zero((double2*)scratch, lay.total / sizeof(double2), st);
    const Plan& P = plan;
    const size_t nl = P.n_leaves;
    double2** tab = (double2**)(scratch + lay.tabs);
    launch_col_table(tab, const_cast<double2*>(f), nl * P.n_int, nrhs, st);
    launch_col_table(tab + nrhs, const_cast<double2*>(g), nl * P.nb, nrhs, st);
    launch_col_table(tab + 2 * nrhs, u, nl * P.m, nrhs, st);
    RhsCols cols;
    cols.f = tab; cols.g = tab + nrhs; cols.u = tab + 2 * nrhs;
    SolveWork<double> w64;
    SolveWork<float> w32;
    auto prepare = [&](auto& w) {
        w.bind(scratch, lay);
        for (int l = 0; l < P.L; ++l) {
            const SolveLevel p = solve_level(w, l);
            hipLaunchKernelGGL(solve_desc_kernel, dim3((p.nodes + 255) / 256), dim3(256), 0, st, p, w.d_up_g[l],
                               w.d_up_h[l], w.d_dn_t[l], w.d_dn_g[l]);
        }
        // tile choices tuned for this width earlier, else the shape rule (the bits do not depend on them)
        auto it = tile_cache.find(2LL * lay.cap + (mixed ? 0 : 1));
        for (int c = 0; c < 4; ++c) w.tile[c] = it != tile_cache.end() ? it->second[c] : std::vector<int>(P.L, -1);
    };
    if (mixed) prepare(w32);
    else prepare(w64);
    HIPCHK(hipGetLastError());
    const bool was = prof_on;   // (the profiling events would be created on demand)
    prof_on = false;
    const int rc = solve_launches(nrhs, cols, prec, st, w64, w32);
    prof_on = was;
    if (rc) return 1;
    return own ? join_default_out(st) : 0;
}

int Tree::solve_cols(int nrhs, const RhsCols& cols, int prec, hipStream_t st, int slot_id, bool graph) {
    if (!factored) return fail("Error: Robin Tree must be factored first.");
    if (nrhs < 1) return fail("Error: nrhs must be positive.");
    if (prec < 0) prec = opts.precision;
    if (prec == RS_FP64 && merge_c64)
        return fail("Error: an FP64 solve needs a tree factored in FP64 (this tree's merge products are complex64).");
    if (prec == RS_FP64 && !have64)
        return fail("Error: this tree keeps complex64 solve data only (rs_opts.retain RS_RETAIN_FP64 keeps the "
                    "FP64 data for FP64 solves).");
    if (prec != RS_FP64 && prec != RS_MIXED) return fail("Error: precision must be RS_FP64 or RS_MIXED.");
    if (!st) st = stream;
    Slot& sl = slot(slot_id);
    // workspaces first: allocation and autotuning cannot be captured; a workspace
    // that moves invalidates the slot's captured solves
    // (a rebuilt workspace has a new layout even when the allocator returns the
    // same base address: the captured graphs go whenever it is rebuilt)
    if (prec == RS_MIXED) {
        if (!have32 && make_c64()) return 1;
        const unsigned long long gen = sl.w32.gen;
        if (ensure_work(sl.w32, nrhs, opts.n_refine > 0)) return 1;
        if (sl.w32.gen != gen) sl.drop_graphs();
    } else {
        const unsigned long long gen = sl.w64.gen;
        if (ensure_work(sl.w64, nrhs)) return 1;
        if (sl.w64.gen != gen) sl.drop_graphs();
    }
    if (!graph || !opts.graphs || prof_on) return solve_launches(nrhs, cols, prec, st, sl.w64, sl.w32);
    const int nref = prec == RS_MIXED ? opts.n_refine : 0;
    GraphEnt* hit = nullptr;
    for (GraphEnt& g : sl.graphs)
        if (g.prec == prec && g.nrhs == nrhs && g.n_refine == nref && g.tabs == (const void*)cols.f) hit = &g;
    if (!hit) {
        // capture this shape's launch sequence once
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = solve_launches(nrhs, cols, prec, st, sl.w64, sl.w32);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &gr);
        if (rc || e != hipSuccess) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc ? 1 : fail(std::string("HIP error: stream capture: ") + hipGetErrorString(e));
        }
        hipGraphExec_t ex = nullptr;
        const hipError_t e2 = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e2 != hipSuccess) return fail(std::string("HIP error: graph instantiation: ") + hipGetErrorString(e2));
        if ((int)sl.graphs.size() >= Slot::kMaxGraphs) {
            auto lru = std::min_element(sl.graphs.begin(), sl.graphs.end(),
                                        [](const GraphEnt& a, const GraphEnt& b) { return a.used < b.used; });
            (void)hipGraphExecDestroy(lru->exec);
            sl.graphs.erase(lru);
        }
        GraphEnt g;
        g.prec = prec; g.nrhs = nrhs; g.n_refine = nref; g.tabs = cols.f; g.exec = ex;
        sl.graphs.push_back(g);
        hit = &sl.graphs.back();
    }
    hit->used = ++sl.tick;
    HIPCHK(hipGraphLaunch(hit->exec, st));
    return 0;
}

// The launch sequence of one solve (no allocation, no synchronisation): the
// workspaces of `sl` are ready.
int Tree::solve_launches(int nrhs, const RhsCols& cols, int prec, hipStream_t st, SolveWork<double>& w64,
                         SolveWork<float>& w32) {
    if (prec == RS_MIXED) {
        PassIO io;
        io.f64 = cols.f; io.g = cols.g;
        if (opts.n_refine <= 0) {
            io.u64 = cols.u;
            return solve_pass(w32, nrhs, io, st);
        }
        // The first iterate is kept in complex64: its FP64 residual is exact for
        // the rounded iterate, and the correction restores full accuracy.  Then
        // u = u_{i} + K32^-1 (b - K u_{i}) for a fixed number of steps, so that a
        // right-hand side gets the same bits alone or batched.
        float2* uA = w32.uA;
        io.u32 = uA;
        if (solve_pass(w32, nrhs, io, st)) return 1;
        for (int it = 0; it < opts.n_refine; ++it) {
            if (residual(nrhs, it == 0 ? nullptr : cols.u, it == 0 ? uA : nullptr, cols.f, cols.g,
                         w32.rf, w32.rg, w32.rj, st))
                return 1;
            // the correction pass writes u = u_i + K32^-1 r directly (its leaf stage adds
            // the correction to the current iterate as it stores the solution)
            PassIO c;
            c.f32 = w32.rf; c.jump = w32.rj;
            c.g32 = w32.rg;
            c.u64 = cols.u;
            if (it == 0) c.ubase32 = uA;
            else c.uacc = 1;
            if (solve_pass(w32, nrhs, c, st)) return 1;
        }
        return 0;
    }
    PassIO io;
    io.f64 = cols.f; io.g = cols.g; io.u64 = cols.u;
    return solve_pass(w64, nrhs, io, st);
}

long long Tree::group_bytes(int nrhs) const {
    const bool mixed = opts.precision == RS_MIXED;
    return (long long)work_layout(nrhs, mixed ? 8 : 16, mixed && opts.n_refine > 0, false, true).total;
}

// The launches of solve_pass for several trees at once: the leaf stages and
// gathers per tree, each level's products as one grouped launch.  Up pass:
// level L_t - 1 - h of every tree t at height h; down pass: level d of every
// tree at depth d.  Within a tree the launch order is the one of solve_pass.
template <typename R>
int Tree::pass_grouped(const std::vector<Tree*>& ts, const std::vector<SolveWork<R>*>& ws,
                       const std::vector<int>& nrhs, const std::vector<PassIO>& io, hipStream_t st) {
    using V = c_t<R>;
    const int n = (int)ts.size();
    int maxL = 0;
    for (Tree* t : ts) maxL = std::max(maxL, t->plan.L);
    auto sinv_of = [&](Tree* t) {
        return std::is_same<R, double>::value ? (const V*)t->sinv.p : (const V*)t->sinv32.p;
    };
    for (int q = 0; q < n; ++q) {
        Tree* t = ts[q];
        LeafArgs<R> up;
        up.mode = LEAF_UP; up.ncols = nrhs[q]; up.n_leaves = t->plan.n_leaves; up.sinv = sinv_of(t);
        up.f64 = io[q].f64; up.f32 = io[q].f32; up.jump = io[q].jump; up.hout = ws[q]->hV[t->plan.L];
        launch_leaf_apply<R>(t->G, up, t->lds_limit, st);
    }
    std::vector<GemmGroup> g;
    auto add = [&](const GemmDesc* d, int np, int M, int N, int K, void* wsp) {
        GemmGroup e;
        e.descs = d; e.nprob = np; e.M = M; e.N = N; e.K = K; e.ws = wsp; e.ks_target = solve_ks_target(np, M, K);
        g.push_back(e);
    };
    for (int h = 0; h < maxL; ++h) {
        g.clear();
        for (int q = 0; q < n; ++q) {
            const int l = ts[q]->plan.L - 1 - h;
            if (l < 0) continue;
            const Level& Vl = ts[q]->plan.lv[l];
            add(ws[q]->d_up_g[l], Vl.nodes, 2 * Vl.isz, nrhs[q], 2 * Vl.isz, ws[q]->ws);
        }
        launch_gemm_grouped<R>(g.data(), (int)g.size(), st);
        g.clear();
        for (int q = 0; q < n; ++q) {
            const int l = ts[q]->plan.L - 1 - h;
            if (l <= 0) continue;
            const Level& Vl = ts[q]->plan.lv[l];
            add(ws[q]->d_up_h[l], 2 * Vl.nodes, Vl.bn, nrhs[q], Vl.isz, ws[q]->ws);
        }
        if (!g.empty()) launch_gemm_grouped<R>(g.data(), (int)g.size(), st);
    }
    for (int q = 0; q < n; ++q) {
        Tree* t = ts[q];
        const Plan& P = t->plan;
        if (io[q].g32)
            launch_root_gather32<V>(ws[q]->gF[0], (size_t)t->iface_len, io[q].g32, t->d_root_leaf.as<int>(),
                                    t->d_root_loc.as<int>(), P.lv[0].P, P.n_leaves, P.nb, nrhs[q], st);
        else
            launch_root_gather<V>(ws[q]->gF[0], (size_t)t->iface_len, io[q].g, t->d_root_leaf.as<int>(),
                                  t->d_root_loc.as<int>(), P.lv[0].P, P.nb, nrhs[q], st);
    }
    for (int d = 0; d < maxL; ++d) {
        g.clear();
        for (int q = 0; q < n; ++q) {
            if (d >= ts[q]->plan.L) continue;
            const Level& Vl = ts[q]->plan.lv[d];
            add(ws[q]->d_dn_t[d], 2 * Vl.nodes, Vl.isz, nrhs[q], Vl.bn, ws[q]->ws);
        }
        launch_gemm_grouped<R>(g.data(), (int)g.size(), st);
        g.clear();
        for (int q = 0; q < n; ++q) {
            if (d >= ts[q]->plan.L) continue;
            const Level& Vl = ts[q]->plan.lv[d];
            add(ws[q]->d_dn_g[d], Vl.nodes, 2 * Vl.isz, nrhs[q], 2 * Vl.isz, ws[q]->ws);
        }
        launch_gemm_grouped<R>(g.data(), (int)g.size(), st);
    }
    for (int q = 0; q < n; ++q) {
        Tree* t = ts[q];
        LeafArgs<R> dn;
        dn.mode = LEAF_DOWN; dn.ncols = nrhs[q]; dn.n_leaves = t->plan.n_leaves; dn.sinv = sinv_of(t);
        dn.f64 = io[q].f64; dn.f32 = io[q].f32; dn.gin = ws[q]->gF[0];
        dn.gsrc = t->d_src_leaf.as<int>(); dn.gld = (size_t)t->iface_len;
        dn.u64 = io[q].u64; dn.u32 = io[q].u32; dn.ubase32 = io[q].ubase32; dn.uacc = io[q].uacc;
        launch_leaf_apply<R>(t->G, dn, t->lds_limit, st);
    }
    return 0;
}

int Tree::solve_grouped(std::vector<GroupMember>& ms, hipStream_t st, std::string& err) {
    auto bad = [&](const std::string& m) { err = m; return 1; };
    if (ms.empty()) return 0;
    const int prec = ms[0].t->opts.precision;
    const bool mixed = prec == RS_MIXED;
    for (GroupMember& m : ms) {
        Tree* t = m.t;
        if (!t || !t->factored) return bad("Error: Robin Tree must be factored first.");
        if (t->opts.precision != prec) return bad("Error: a grouped pass needs trees of one precision.");
        if (m.nrhs < 1) return bad("Error: nrhs must be positive.");
        if (!mixed && t->merge_c64)
            return bad("Error: an FP64 solve needs a tree factored in FP64 (this tree's merge products are "
                       "complex64).");
        if (mixed && !t->have32 && t->make_c64()) return bad(t->err);
        if ((long long)m.bytes < t->group_bytes(m.nrhs)) return bad("Error: a grouped member's workspace is too small.");
    }
    const int n = (int)ms.size();
    std::vector<Tree*> ts(n);
    std::vector<int> nr(n);
    for (int q = 0; q < n; ++q) { ts[q] = ms[q].t; nr[q] = ms[q].nrhs; }
    // the members' workspaces and product descriptors, written on the stream
    auto prepare = [&](auto& w, GroupMember& m) {
        Tree* t = m.t;
        const WorkLayout lay =
            t->work_layout(m.nrhs, mixed ? 8 : 16, mixed && t->opts.n_refine > 0, false, true);
        w.bind(m.base, lay);
        for (int l = 0; l < t->plan.L; ++l) {
            const SolveLevel p = t->solve_level(w, l);
            hipLaunchKernelGGL(solve_desc_kernel, dim3((p.nodes + 255) / 256), dim3(256), 0, st, p, w.d_up_g[l],
                               w.d_up_h[l], w.d_dn_t[l], w.d_dn_g[l]);
        }
    };
    if (!mixed) {
        std::vector<SolveWork<double>> w(n);
        std::vector<SolveWork<double>*> wp(n);
        std::vector<PassIO> io(n);
        for (int q = 0; q < n; ++q) {
            prepare(w[q], ms[q]);
            wp[q] = &w[q];
            io[q].f64 = ms[q].cols.f; io[q].g = ms[q].cols.g; io[q].u64 = ms[q].cols.u;
        }
        if (hipGetLastError() != hipSuccess) return bad("HIP error: descriptor kernels of a grouped pass.");
        pass_grouped<double>(ts, wp, nr, io, st);
    } else {
        std::vector<SolveWork<float>> w(n);
        std::vector<SolveWork<float>*> wp(n);
        std::vector<PassIO> io(n);
        int maxref = 0;
        for (int q = 0; q < n; ++q) {
            prepare(w[q], ms[q]);
            wp[q] = &w[q];
            const int nref = ts[q]->opts.n_refine;
            maxref = std::max(maxref, nref);
            io[q].f64 = ms[q].cols.f; io[q].g = ms[q].cols.g;
            // the complex64 first iterate (refined below), or the answer itself
            if (nref <= 0) io[q].u64 = ms[q].cols.u;
            else io[q].u32 = w[q].uA;
        }
        if (hipGetLastError() != hipSuccess) return bad("HIP error: descriptor kernels of a grouped pass.");
        pass_grouped<float>(ts, wp, nr, io, st);
        // u = u_i + K32^-1 (b - K u_i) for each tree's fixed number of steps
        for (int it = 0; it < maxref; ++it) {
            std::vector<Tree*> st_t;
            std::vector<SolveWork<float>*> st_w;
            std::vector<int> st_n;
            std::vector<PassIO> st_io;
            for (int q = 0; q < n; ++q) {
                if (ts[q]->opts.n_refine <= it) continue;
                const RhsCols& c = ms[q].cols;
                if (ts[q]->residual(nr[q], it == 0 ? nullptr : c.u, it == 0 ? w[q].uA : nullptr, c.f, c.g, w[q].rf,
                                    w[q].rg, w[q].rj, st))
                    return bad(ts[q]->err);
                PassIO ci;
                ci.f32 = w[q].rf; ci.jump = w[q].rj; ci.g32 = w[q].rg; ci.u64 = c.u;
                if (it == 0) ci.ubase32 = w[q].uA;
                else ci.uacc = 1;
                st_t.push_back(ts[q]); st_w.push_back(wp[q]); st_n.push_back(nr[q]); st_io.push_back(ci);
            }
            pass_grouped<float>(st_t, st_w, st_n, st_io, st);
        }
    }
    if (hipGetLastError() != hipSuccess) return bad("HIP error: a grouped pass's launches.");
    return 0;
}

int Tree::reserve_slot(int slot_id, int cap) {
    if (!factored) return fail("Error: Robin Tree must be factored first.");
    Slot& sl = slot(slot_id);
    if (opts.precision == RS_MIXED) {
        if (!have32 && make_c64()) return 1;
        return ensure_work(sl.w32, cap, opts.n_refine > 0);
    }
    return ensure_work(sl.w64, cap);
}

int Tree::residual(int nrhs, double2* const* u64, const float2* u32, const double2* const* f,
                   const double2* const* g, float2* rf, float2* rg, float2* rj, hipStream_t st) {
    const Plan& P = plan;
    pbeg(P_RESIDUAL, st);
    launch_residual(G, u64, u32, kint.as<double2>(), f, g, d_leaf_nbr.as<int>(), make_double2(P.nu[0], P.nu[1]),
                    make_double2(P.zeta[0], P.zeta[1]), rf, rg, rj, P.n_leaves, nrhs, st);
    const double eu = u64 ? 16.0 : 8.0;
    pend((double)nrhs * P.n_leaves * (eu * P.m + 16.0 * P.n_int + 8.0 * P.n_int + 16.0 * P.nb + 8.0 * 2 * P.nb) +
             16.0 * P.n_leaves * P.n_int,
         (double)nrhs * P.n_leaves * P.m * 8.0 * (2 * G.d + 1), st);
    HIPCHK(hipGetLastError());
    return 0;
}

int Tree::residual_norms(int nrhs, const double2* f, const double2* g, const double2* u, double* out) {
    if (!factored) return fail("Error: Robin Tree must be factored first.");
    if (nrhs < 1) return fail("Error: nrhs must be positive.");
    const Plan& P = plan;
    const size_t nl = P.n_leaves;
    DevBuf rf, rg, rj, tabs, nrm;
    if (alloc(rf, nrhs * nl * P.n_int * 8) || alloc(rg, nrhs * nl * P.nb * 8) || alloc(rj, nrhs * nl * P.nb * 8) ||
        alloc(tabs, (size_t)3 * nrhs * sizeof(void*)) || alloc(nrm, (size_t)2 * nrhs * sizeof(double)))
        return 1;
    double2** tab = tabs.as<double2*>();
    if (join_default_in(stream)) return 1;
    launch_col_table(tab, const_cast<double2*>(f), nl * P.n_int, nrhs, stream);
    launch_col_table(tab + nrhs, const_cast<double2*>(g), nl * P.nb, nrhs, stream);
    launch_col_table(tab + 2 * nrhs, const_cast<double2*>(u), nl * P.m, nrhs, stream);
    const bool was = prof_on;
    prof_on = false;
    const int rc = residual(nrhs, tab + 2 * nrhs, nullptr, tab, tab + nrhs, rf.as<float2>(), rg.as<float2>(),
                            rj.as<float2>(), stream);
    prof_on = was;
    if (rc) return 1;
    launch_fd_norms(G, rf.as<float2>(), rg.as<float2>(), rj.as<float2>(), tab, tab + nrhs, d_leaf_nbr.as<int>(),
                    P.n_leaves, nrhs, nrm.as<double>(), stream);
    std::vector<double> h((size_t)2 * nrhs);
    HIPCHK(hipMemcpyAsync(h.data(), nrm.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int c = 0; c < nrhs; ++c) out[c] = h[2 * c + 1] > 0 ? std::sqrt(h[2 * c] / h[2 * c + 1]) : std::sqrt(h[2 * c]);
    return 0;
}

int Tree::solve_host(int nrhs, const double* f, const double* g, int prec) {
    if (!factored) return fail("Error: Robin Tree must be factored first.");
    if (nrhs < 1) return fail("Error: nrhs must be positive.");
    const double t0 = now_s();
    const Plan& P = plan;
    size_t nf = (size_t)nrhs * P.n_leaves * P.n_int, ng = (size_t)nrhs * P.n_leaves * P.nb,
           nu = (size_t)nrhs * P.n_leaves * P.m;
    if (st_f.bytes < nf * 16 && alloc(st_f, nf * 16)) return 1;
    if (st_g.bytes < ng * 16 && alloc(st_g, ng * 16)) return 1;
    if (st_u.bytes < nu * 16 && alloc(st_u, nu * 16)) return 1;
    HIPCHK(hipMemcpyAsync(st_f.p, f, nf * 16, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(st_g.p, g, ng * 16, hipMemcpyHostToDevice, stream));
    if (solve_device(nrhs, st_f.as<double2>(), st_g.as<double2>(), st_u.as<double2>(), prec, stream)) return 1;
    HIPCHK(hipStreamSynchronize(stream));
    last_nrhs = nrhs;
    s_time = now_s() - t0;
    return 0;
}

int Tree::get_solution(int rhs, double* u) const {
    if (!factored) return const_cast<Tree*>(this)->fail("Error in GET_SOLUTION: Robin Tree not factored.");
    if (rhs < 0 || rhs >= last_nrhs) return const_cast<Tree*>(this)->fail("Error: no such right-hand side.");
    size_t n = (size_t)plan.n_leaves * plan.m;
    hipError_t e = hipMemcpy(u, st_u.as<double2>() + (size_t)rhs * n, n * 16, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return const_cast<Tree*>(this)->fail(hipGetErrorString(e));
    return 0;
}

int Tree::node_rtr(int lvl, int idx, double* out, int* Pout) const {
    Tree* self = const_cast<Tree*>(this);
    if (!factored) return self->fail("Error: not factored.");
    if (lvl < 0 || lvl > plan.L || idx < 0 || idx >= plan.lv[lvl].nodes) return self->fail("Error: no such node.");
    if (lvl >= (int)T.size() || !T[lvl] || !T[lvl]->p)
        return self->fail("Error: RtR maps are kept only with rs_opts.retain RS_RETAIN_MAPS (and the root's "
                          "only with skip_root_rtr = 0).");
    const int Pp = plan.lv[lvl].P;
    if (Pout) *Pout = Pp;
    if (!out) return 0;
    std::vector<double2> h((size_t)Pp * Pp);
    hipError_t e = hipMemcpy(h.data(), T[lvl]->as<double2>() + (size_t)idx * Pp * Pp, h.size() * 16,
                             hipMemcpyDeviceToHost);
    if (e != hipSuccess) return self->fail(hipGetErrorString(e));
    std::vector<int> perm = plan.node_to_face_order(lvl, lvl == 0 ? 0 : (idx & 1));
    double2* o = (double2*)out;
    for (int c = 0; c < Pp; ++c)
        for (int r = 0; r < Pp; ++r) o[(size_t)perm[r] + (size_t)perm[c] * Pp] = h[(size_t)r + (size_t)c * Pp];
    return 0;
}

template int Tree::ensure_work<double>(SolveWork<double>&, int, bool);
template int Tree::tune_work<double>(SolveWork<double>&);
template int Tree::tune_work<float>(SolveWork<float>&);
template int Tree::ensure_work<float>(SolveWork<float>&, int, bool);

}  // namespace rs
