// capi.cpp -- the extern "C" boundary of include/robinsolver.h.
#include <cstring>
#include <new>
#include <string>

#include "tree.hpp"

struct rs_tree {
    rs::Tree t;
};

namespace {
std::string g_err;
thread_local std::string g_tls;
}  // namespace

extern "C" {

void rs_opts_default(rs_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->precision = RS_FP64;
    o->skip_root_rtr = 1;
    o->n_refine = 2;
    o->verbose = 0;
    o->hbm_budget = 0;
}

int rs_tree_create(rs_tree** t, const rs_opts* o) {
    if (!t) { g_err = "Error: null handle pointer."; return 1; }
    *t = nullptr;
    rs_tree* h = new (std::nothrow) rs_tree;
    if (!h) { g_err = "Error: out of host memory."; return 1; }
    if (o) h->t.opts = *o;
    if (h->t.opts.precision != RS_FP64 && h->t.opts.precision != RS_MIXED) {
        delete h;
        g_err = "Error: unknown precision mode.";
        return 1;
    }
    *t = h;
    return 0;
}

int rs_tree_destroy(rs_tree* t) {
    delete t;
    return 0;
}

int rs_plan(rs_tree* t, int d, int L, const double* domain, double h_tgt, double kb, const double* bp,
            int* n_leaves, int* m, int* n_int, int* n_bnd) {
    if (!t || !domain || !bp) { g_err = "Error: null argument."; return 1; }
    if (t->t.plan_tree(d, L, domain, h_tgt, kb, bp)) return 1;
    if (n_leaves) *n_leaves = t->t.plan.n_leaves;
    if (m) *m = t->t.plan.m;
    if (n_int) *n_int = t->t.plan.n_int;
    if (n_bnd) *n_bnd = t->t.plan.nb;
    return 0;
}

int rs_leaf_points(const rs_tree* t, int leaf, double* xyz, int* face) {
    if (!t || !t->t.planned || !xyz) return 1;
    if (leaf < 0 || leaf >= t->t.plan.n_leaves) return const_cast<rs::Tree&>(t->t).fail("Error: no such leaf.");
    t->t.plan.leaf_points(leaf, xyz, face);
    return 0;
}

int rs_leaf_grid(const rs_tree* t, int* ptbox, double* h) {
    if (!t || !t->t.planned) return 1;
    const rs::Plan& P = t->t.plan;
    for (int j = 0; j < 3; ++j) {
        if (ptbox) ptbox[j] = j < P.d ? P.lv[P.L].ptbox[j] : -1;
        if (h) h[j] = P.h[j];
    }
    return 0;
}

int rs_leaf_box(const rs_tree* t, int leaf, double* box) {
    if (!t || !t->t.planned || !box) return 1;
    if (leaf < 0 || leaf >= t->t.plan.n_leaves) return const_cast<rs::Tree&>(t->t).fail("Error: no such leaf.");
    const rs::PlanNode& n = t->t.plan.nodes[t->t.plan.L][leaf];
    for (int j = 0; j < 3; ++j) { box[2 * j] = n.box[0][j]; box[2 * j + 1] = n.box[1][j]; }
    return 0;
}

int rs_factor(rs_tree* t, const double* k_int) {
    if (!t || !k_int) { g_err = "Error: null argument."; return 1; }
    return t->t.factor(k_int);
}

int rs_solve(rs_tree* t, int nrhs, const double* f, const double* g, int precision) {
    if (!t || !f || !g) { g_err = "Error: null argument."; return 1; }
    return t->t.solve_host(nrhs, f, g, precision);
}

int rs_get_solution(const rs_tree* t, int rhs, double* u) {
    if (!t || !u) return 1;
    return t->t.get_solution(rhs, u);
}

int rs_solve_device(rs_tree* t, int nrhs, const void* f_dev, const void* g_dev, void* u_dev, int precision,
                    void* stream) {
    if (!t || !f_dev || !g_dev || !u_dev) { g_err = "Error: null argument."; return 1; }
    return t->t.solve_device(nrhs, (const double2*)f_dev, (const double2*)g_dev, (double2*)u_dev, precision,
                             (hipStream_t)stream);
}

int rs_timings(const rs_tree* t, double* f_time, double* s_time) {
    if (!t) return 1;
    if (f_time) *f_time = t->t.f_time;
    if (s_time) *s_time = t->t.s_time;
    return 0;
}

const char* rs_last_error(const rs_tree* t) {
    if (!t) return g_err.c_str();
    return t->t.err.c_str();
}

int rs_inspect(const rs_tree* t, int* factored, int* nlvls, long long* device_bytes) {
    if (!t) return 1;
    if (factored) *factored = t->t.factored ? 1 : 0;
    if (nlvls) *nlvls = t->t.plan.L;
    if (device_bytes) *device_bytes = t->t.dev_bytes;
    return 0;
}

int rs_level_info(const rs_tree* t, int lvl, int* nodes, int* P, int* isz, int* child_P) {
    if (!t || !t->t.planned || lvl < 0 || lvl > t->t.plan.L) return 1;
    const rs::Level& V = t->t.plan.lv[lvl];
    if (nodes) *nodes = V.nodes;
    if (P) *P = V.P;
    if (isz) *isz = V.isz;
    if (child_P) *child_P = V.bc;
    return 0;
}

int rs_node_rtr(const rs_tree* t, int lvl, int idx, double* out, int* P) {
    if (!t) return 1;
    return t->t.node_rtr(lvl, idx, out, P);
}

int rs_profile(rs_tree* t, int enable) {
    if (!t) return 1;
    t->t.prof_on = enable != 0;
    return 0;
}

int rs_profile_ncat(void) { return rs::P_NCAT; }

const char* rs_profile_name(int cat) {
    static const char* names[rs::P_NCAT] = {"leaf_up", "leaf_down", "gemm_up_g", "gemm_up_h",
                                            "gemm_dn_t", "gemm_dn_g", "gather", "residual"};
    return (cat >= 0 && cat < rs::P_NCAT) ? names[cat] : "";
}

int rs_profile_read(rs_tree* t, int cat, double* ms, int* launches, double* bytes, double* flops) {
    if (!t || cat < 0 || cat >= rs::P_NCAT) return 1;
    if (t->t.prof_collect()) return 1;
    rs::Tree& T = t->t;
    if (ms) *ms = T.prof_ms[cat];
    if (launches) *launches = T.prof_n[cat];
    if (bytes) *bytes = T.prof_bytes[cat];
    if (flops) *flops = T.prof_flops[cat];
    T.prof_ms[cat] = 0; T.prof_n[cat] = 0; T.prof_bytes[cat] = 0; T.prof_flops[cat] = 0;
    return 0;
}

}  // extern "C"
