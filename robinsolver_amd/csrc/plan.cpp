// plan.cpp -- host geometry of the HPS tree (see plan.hpp).
#include "plan.hpp"

#include <cmath>
#include <complex>
#include <cstring>

namespace rs {

namespace {

void get_pt_face(int pf[7], const int pb[3], int d) {
    for (int j = 0; j < 7; ++j) pf[j] = 0;
    int pr = 1;
    for (int j = 0; j < d; ++j) pr *= pb[j];
    for (int j = 0; j < d; ++j) pf[2 * j + 1] = pf[2 * j + 2] = pr / pb[j];
}

// node_divide: cut the longest axis (first axis on ties), child1 = lower half
void divide(int d, const PlanNode& p, PlanNode& c1, PlanNode& c2, int& edge) {
    double w[3] = {0, 0, 0};
    for (int j = 0; j < d; ++j) w[j] = p.box[1][j] - p.box[0][j];
    int e = 0;
    for (int j = 1; j < 3; ++j)
        if (w[j] > w[e]) e = j;
    edge = e;
    std::memcpy(c1.box, p.box, sizeof p.box);
    std::memcpy(c2.box, p.box, sizeof p.box);
    c1.box[1][e] = p.box[0][e] + w[e] * 0.5;
    c2.box[0][e] = c1.box[1][e];
    c1.iface = 2 * (e + 1);
    c2.iface = 2 * (e + 1) - 1;
    for (int f = 0; f < 7; ++f) { c1.isbndry[f] = 0; c2.isbndry[f] = 0; }
    for (int f = 1; f <= 2 * d; ++f)
        if (p.isbndry[f]) {
            if (f != c1.iface) c1.isbndry[f] = 1;
            if (f != c2.iface) c2.isbndry[f] = 1;
        }
}

void node_order(int d, int own, int pf[7], int noff[7]) {
    int off = 0;
    for (int f = 1; f <= 2 * d; ++f)
        if (f != own) { noff[f] = off; off += pf[f]; }
    if (own) noff[own] = off;
    for (int f = 2 * d + 1; f < 7; ++f) noff[f] = 0;
    noff[0] = 0;
}

}  // namespace

std::string Plan::build(int d_, int L_, const double* domain_, double h_tgt_, double kb_, const double* bp_) {
    d = d_; L = L_; h_tgt = h_tgt_; kb = kb_;
    for (int j = 0; j < 3; ++j) { domain[0][j] = domain_[2 * j]; domain[1][j] = domain_[2 * j + 1]; }
    // robin_tree_factor validation (robin_tree_module.f90:121-140)
    for (int j = 0; j < 3; ++j)
        if (domain[1][j] - domain[0][j] < 0) return "Error in ROBIN_TREE_FACTOR: Invalid operator domain format.";
    if (d != 2 && d != 3) return "Error in ROBIN_TREE_FACTOR: Dimension must be 2 or 3.";
    if (L < 0 || L > 24) return "Error: number of levels must be in 0..24.";
    if (!(h_tgt > 0)) return "Error: h_tgt must be positive.";
    using cd = std::complex<double>;
    cd b[2][2];
    for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
            b[i][j] = cd(bp_[2 * (i + 2 * j)], bp_[2 * (i + 2 * j) + 1]);
            bp[i][j][0] = b[i][j].real(); bp[i][j][1] = b[i][j].imag();
        }
    if (std::abs(b[0][0] * b[1][1] - b[0][1] * b[1][0]) < std::pow(10.0, -3.0))
        return "Error: Boundary parameter matrix is nearly singular.";
    if (std::abs(b[0][0] * b[0][1]) == 0.0) return "Error: bndry_params(1,1)*bndry_params(1,2) must be nonzero.";
    // nu, zeta (robin_tree_module.f90:425-431)
    cd nu_ = (-b[0][0] * b[1][1] + b[0][1] * b[1][0]) / (2.0 * b[0][0] * b[0][1]);
    cd ze_ = (b[0][0] * b[1][1] + b[0][1] * b[1][0]) / (2.0 * b[0][0] * b[0][1]);
    nu[0] = nu_.real(); nu[1] = nu_.imag(); zeta[0] = ze_.real(); zeta[1] = ze_.imag();

    nodes.assign(L + 1, {});
    PlanNode root{};
    std::memcpy(root.box, domain, sizeof domain);
    root.iface = 0;
    for (int f = 1; f <= 2 * d; ++f) root.isbndry[f] = 1;
    root.cut = -1;
    nodes[0].push_back(root);
    for (int l = 0; l < L; ++l) {
        nodes[l + 1].resize((size_t)2 * nodes[l].size());
        for (size_t n = 0; n < nodes[l].size(); ++n) {
            int e;
            PlanNode& p = nodes[l][n];
            PlanNode& c1 = nodes[l + 1][2 * n];
            PlanNode& c2 = nodes[l + 1][2 * n + 1];
            divide(d, p, c1, c2, e);
            p.cut = e;
            c1.cut = c2.cut = -1;
        }
    }
    n_leaves = (int)nodes[L].size();
    // leaf grids (get_h); all leaves must agree, because the reference solves
    // every leaf with the last leaf's h (T%opts%h)
    for (int q = 0; q < n_leaves; ++q) {
        PlanNode& lf = nodes[L][q];
        double hh[3] = {0, 0, 0};
        int pb[3] = {-1, -1, -1};
        for (int j = 0; j < d; ++j) {
            int n = (int)std::ceil((lf.box[1][j] - lf.box[0][j]) / h_tgt);
            pb[j] = n > 3 ? n : 3;
            hh[j] = (lf.box[1][j] - lf.box[0][j]) / pb[j];
        }
        std::memcpy(lf.ptbox, pb, sizeof pb);
        if (q == 0) std::memcpy(h, hh, sizeof hh);
        else if (std::memcmp(h, hh, sizeof hh) || std::memcmp(pb, nodes[L][0].ptbox, sizeof pb))
            return "Error: leaves have non-uniform grids (unsupported).";
    }
    // parent ptbox (robin_tree_module.f90:404-410)
    for (int l = L - 1; l >= 0; --l)
        for (size_t n = 0; n < nodes[l].size(); ++n) {
            PlanNode& p = nodes[l][n];
            const PlanNode& c1 = nodes[l + 1][2 * n];
            double diff[3];
            for (int j = 0; j < 3; ++j) diff[j] = std::fabs((p.box[1][j] - p.box[0][j]) - (c1.box[1][j] - c1.box[0][j]));
            int e = 0;
            for (int j = 1; j < d; ++j)
                if (diff[j] > diff[e]) e = j;
            std::memcpy(p.ptbox, c1.ptbox, sizeof p.ptbox);
            p.ptbox[e] *= 2;
        }
    // levels: shapes must be uniform within a level
    lv.assign(L + 1, Level{});
    for (int l = 0; l <= L; ++l) {
        Level& V = lv[l];
        V.nodes = (int)nodes[l].size();
        std::memcpy(V.ptbox, nodes[l][0].ptbox, sizeof V.ptbox);
        V.cut = nodes[l][0].cut;
        for (const PlanNode& nd : nodes[l])
            if (std::memcmp(nd.ptbox, V.ptbox, sizeof V.ptbox) || nd.cut != V.cut)
                return "Error: nodes of one level differ in shape (unsupported).";
        get_pt_face(V.pf, V.ptbox, d);
        V.P = 0;
        for (int f = 1; f <= 2 * d; ++f) V.P += V.pf[f];
        if (l == 0) { V.own_iface[0] = V.own_iface[1] = 0; }
        else { V.own_iface[0] = nodes[l][0].iface; V.own_iface[1] = nodes[l].size() > 1 ? nodes[l][1].iface : 0; }
        for (int par = 0; par < 2; ++par) node_order(d, V.own_iface[par], V.pf, V.noff[par]);
    }
    for (int l = 0; l < L; ++l) {
        Level& V = lv[l];
        const Level& C = lv[l + 1];
        int i1 = C.own_iface[0], i2 = C.own_iface[1];
        if (C.pf[i1] != C.pf[i2]) return "Error: Interface sizes of node children do not match";
        V.isz = C.pf[i1];
        V.bc = C.P;
        V.bn = C.P - V.isz;
        V.prow.assign((size_t)4 * V.bn, 0);
        for (int par = 0; par < 2; ++par)
            for (int c = 0; c < 2; ++c) {
                int own = C.own_iface[c];
                for (int f = 1; f <= 2 * d; ++f) {
                    if (f == own) continue;
                    for (int t = 0; t < C.pf[f]; ++t) {
                        int r = C.noff[c][f] + t;
                        int pos = ((f - 1) / 2 == V.cut) ? V.noff[par][f] + t : V.noff[par][f] + c * C.pf[f] + t;
                        V.prow[(size_t)(par * 2 + c) * V.bn + r] = pos;
                    }
                }
            }
        V.imap.resize((size_t)2 * V.isz);
        for (int k = 0; k < 2 * V.isz; ++k) V.imap[k] = V.bn + k + (k >= V.isz ? V.bn : 0);
    }
    // leaf sizes
    const Level& LF = lv[L];
    n_int = 1;
    for (int j = 0; j < d; ++j) n_int *= LF.ptbox[j];
    nb = LF.P;
    m = nb + n_int;
    int off = 0;
    for (int f = 0; f < 7; ++f) leaf_foff[f] = 0;
    for (int f = 1; f <= 2 * d; ++f) { leaf_foff[f] = off; off += LF.pf[f]; }
    // root face concatenation (apply_bc, robin_tree_module.f90:1205-1230)
    root_leaf.clear(); root_loc.clear();
    for (int f = 1; f <= 2 * d; ++f) {
        // walk the tree: a face of a node is child1's then child2's part, or only
        // one child's part when the face is normal to the cut
        std::vector<std::pair<int, int>> stack{{0, 0}};  // (level, index)
        std::vector<std::pair<int, int>> order;
        // recursive in DFS order
        struct Rec {
            const Plan* P; int f;
            std::vector<int>* lf; std::vector<int>* lc;
            void go(int l, int n) {
                if (l == P->L) {
                    for (int t = 0; t < P->lv[l].pf[f]; ++t) { lf->push_back(n); lc->push_back(P->leaf_foff[f] + t); }
                    return;
                }
                int e = P->lv[l].cut;
                if ((f - 1) / 2 == e) go(l + 1, (f % 2 == 1) ? 2 * n : 2 * n + 1);
                else { go(l + 1, 2 * n); go(l + 1, 2 * n + 1); }
            }
        } rec{this, f, &root_leaf, &root_loc};
        rec.go(0, 0);
    }
    if ((int)root_leaf.size() != lv[0].P) return "Error: root boundary size mismatch.";
    // leaf neighbours across every face (leaves tile the domain on a regular grid)
    {
        double w[3];
        for (int j = 0; j < 3; ++j) w[j] = nodes[L][0].box[1][j] - nodes[L][0].box[0][j];
        int cnt[3] = {1, 1, 1};
        std::vector<int> pos((size_t)3 * n_leaves, 0);
        for (int q = 0; q < n_leaves; ++q)
            for (int j = 0; j < d; ++j) {
                int ix = (int)std::llround((nodes[L][q].box[0][j] - domain[0][j]) / w[j]);
                pos[(size_t)3 * q + j] = ix;
                if (ix + 1 > cnt[j]) cnt[j] = ix + 1;
            }
        std::vector<int> at((size_t)cnt[0] * cnt[1] * cnt[2], -1);
        auto key = [&](int x, int y, int z) { return ((size_t)z * cnt[1] + y) * cnt[0] + x; };
        for (int q = 0; q < n_leaves; ++q) {
            size_t kq = key(pos[3 * q], pos[3 * q + 1], pos[3 * q + 2]);
            if (at[kq] >= 0) return "Error: leaves overlap (internal geometry error).";
            at[kq] = q;
        }
        leaf_nbr.assign((size_t)7 * n_leaves, -1);
        for (int q = 0; q < n_leaves; ++q)
            for (int f = 1; f <= 2 * d; ++f) {
                int j = (f - 1) / 2, s = (f % 2 == 1) ? -1 : 1;
                int p3[3] = {pos[3 * q], pos[3 * q + 1], pos[3 * q + 2]};
                p3[j] += s;
                int nb_ = -1;
                if (p3[j] >= 0 && p3[j] < cnt[j]) nb_ = at[key(p3[0], p3[1], p3[2])];
                if ((nb_ < 0) != (nodes[L][q].isbndry[f] != 0)) return "Error: leaf adjacency inconsistent.";
                leaf_nbr[(size_t)7 * q + f] = nb_;
            }
    }
    return "";
}

void Plan::leaf_points(int leaf, double* xyz, int* face) const {
    const PlanNode& lf = nodes[L][leaf];
    const int* n = lf.ptbox;
    int p = 0;
    auto coord = [&](int j, int g) {
        if (g == 0) return lf.box[0][j];
        if (g == n[j] + 1) return lf.box[1][j];
        return lf.box[0][j] + g * h[j] - h[j] / 2;
    };
    // boundary points (lin2fgrid): face f, remaining axes first-fastest
    for (int f = 1; f <= 2 * d; ++f) {
        int ax = (f - 1) / 2;
        int rem[2], nr = 0;
        for (int j = 0; j < d; ++j)
            if (j != ax) rem[nr++] = j;
        int cnt = lv[L].pf[f];
        for (int t = 0; t < cnt; ++t, ++p) {
            int g[3] = {1, 1, 1};
            g[ax] = (f % 2 == 1) ? 0 : n[ax] + 1;
            if (nr >= 1) g[rem[0]] = (nr == 1) ? t + 1 : t % n[rem[0]] + 1;
            if (nr == 2) g[rem[1]] = t / n[rem[0]] + 1;
            for (int j = 0; j < 3; ++j) xyz[3 * p + j] = (j < d) ? coord(j, g[j]) : lf.box[0][j] + g[j] * h[j] - h[j] / 2;
            if (face) face[p] = f;
        }
    }
    for (int q = 0; q < n_int; ++q, ++p) {
        int g[3] = {1, 1, 1};
        g[0] = q % n[0] + 1;
        if (d == 2) g[1] = q / n[0] + 1;
        else { g[1] = (q / n[0]) % n[1] + 1; g[2] = q / (n[0] * n[1]) + 1; }
        for (int j = 0; j < 3; ++j) xyz[3 * p + j] = (j < d) ? coord(j, g[j]) : lf.box[0][j] + g[j] * h[j] - h[j] / 2;
        if (face) face[p] = 0;
    }
}

std::vector<int> Plan::point_pattern() const {
    const int* n = lv[L].ptbox;
    std::vector<int> pat;
    pat.reserve((size_t)m * 4);
    // boundary points (lin2fgrid): face f, remaining axes first-fastest
    for (int f = 1; f <= 2 * d; ++f) {
        int ax = (f - 1) / 2;
        int rem[2], nr = 0;
        for (int j = 0; j < d; ++j)
            if (j != ax) rem[nr++] = j;
        for (int t = 0; t < lv[L].pf[f]; ++t) {
            int g[3] = {1, 1, 1};
            g[ax] = (f % 2 == 1) ? 0 : n[ax] + 1;
            if (nr >= 1) g[rem[0]] = (nr == 1) ? t + 1 : t % n[rem[0]] + 1;
            if (nr == 2) g[rem[1]] = t / n[rem[0]] + 1;
            pat.insert(pat.end(), {g[0], g[1], g[2], f});
        }
    }
    for (int q = 0; q < n_int; ++q) {
        int g[3] = {1, 1, 1};
        g[0] = q % n[0] + 1;
        if (d == 2) g[1] = q / n[0] + 1;
        else { g[1] = (q / n[0]) % n[1] + 1; g[2] = q / (n[0] * n[1]) + 1; }
        pat.insert(pat.end(), {g[0], g[1], g[2], 0});
    }
    return pat;
}

std::vector<int> Plan::node_to_face_order(int lvl, int parity) const {
    const Level& V = lv[lvl];
    std::vector<int> perm(V.P);
    int off = 0;
    for (int f = 1; f <= 2 * d; ++f) {
        for (int t = 0; t < V.pf[f]; ++t) perm[V.noff[parity][f] + t] = off + t;
        off += V.pf[f];
    }
    return perm;
}

}  // namespace rs
