// capi.cpp -- the extern "C" boundary of include/robinsolver.h.
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "tree.hpp"

namespace {
std::string g_err;
thread_local std::string g_tls;

// Device scratch of rs_solve_device_noalloc: one fixed array per GPU the
// library runs on, held until the process ends.  In host code `scratch` is only
// a stand-in for the symbol; its device address (one per device) comes from
// hipGetSymbolAddress and is the only form passed to kernels or copies.
constexpr size_t kNoallocScratchBytes = (size_t)RS_NOALLOC_SCRATCH_MIB << 20;
__device__ __align__(256) char scratch[kNoallocScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// the address of the current device's scratch array (looked up once per device)
char* noalloc_scratch(std::string& err) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) { err = "Error: no HIP device."; return nullptr; }
    if (dev >= kMaxDevices) { err = "Error: device index beyond the scratch table."; return nullptr; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("HIP error: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        g_scratch_ptr[dev] = (char*)p;
    }
    return g_scratch_ptr[dev];
}
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
    o->autotune = 1;
    o->graphs = 1;
    o->merge_fp64_levels = 0;
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

int rs_points_device(rs_tree* t, double* xyz, int* face, void* stream) {
    if (!t) { g_err = "Error: null handle."; return 1; }
    return t->t.points_device(xyz, face, (hipStream_t)stream);
}

int rs_factor_device(rs_tree* t, const double* k_int) {
    if (!t || !k_int) { g_err = "Error: null argument."; return 1; }
    if (!t->t.planned) return t->t.fail("Error: rs_plan must be called before rs_factor.");
    if (t->t.ensure_device()) return 1;
    return t->t.factor_device((const double2*)k_int);
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

int rs_solve_device_noalloc(rs_tree* t, int nrhs, const void* f_dev, const void* g_dev, void* u_dev, int precision,
                            void* stream) {
    if (!t || !f_dev || !g_dev || !u_dev) { g_err = "Error: null argument."; return 1; }
    char* sc = noalloc_scratch(t->t.err);
    if (!sc) return 1;
    return t->t.solve_device_noalloc(nrhs, (const double2*)f_dev, (const double2*)g_dev, (double2*)u_dev, precision,
                                     (hipStream_t)stream, sc, kNoallocScratchBytes);
}

int rs_set_refine(rs_tree* t, int n_refine) {
    if (!t || n_refine < 0 || n_refine > 16) return 1;
    t->t.opts.n_refine = n_refine;
    return 0;
}

int rs_tree_save(rs_tree* t, const char* path) {
    if (!t) { g_err = "Error: null argument."; return 1; }
    return t->t.save(path);
}

int rs_tree_load(rs_tree* t, const char* path) {
    if (!t) { g_err = "Error: null argument."; return 1; }
    return t->t.load(path);
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

int rs_residual_norms(rs_tree* t, int nrhs, const void* f_dev, const void* g_dev, const void* u_dev,
                      double* rel) {
    if (!t || !f_dev || !g_dev || !u_dev || !rel) return 1;
    return t->t.residual_norms(nrhs, (const double2*)f_dev, (const double2*)g_dev, (const double2*)u_dev, rel);
}

int rs_memory_estimate(const rs_tree* t, int ncols, long long* factor_resident, long long* factor_peak,
                       long long* solve_workspace) {
    if (!t) return 1;
    if (!t->t.planned) return const_cast<rs_tree*>(t)->t.fail("Error: rs_plan must be called first.");
    long long r = 0, p = 0;
    t->t.factor_bytes(&r, &p);
    if (factor_resident) *factor_resident = r;
    if (factor_peak) *factor_peak = p;
    if (solve_workspace) *solve_workspace = t->t.work_bytes(ncols < 1 ? 1 : ncols, t->t.opts.precision);
    return 0;
}

int rs_memory_used(const rs_tree* t, long long* held, long long* peak) {
    if (!t) return 1;
    if (held) *held = t->t.dev_bytes;
    if (peak) *peak = t->t.peak_bytes;
    return 0;
}

int rs_tile_choices(const rs_tree* t, int precision, int* tiles) {
    if (!t || !tiles) return 1;
    const rs::Tree& T = t->t;
    if (T.slots.empty()) return const_cast<rs_tree*>(t)->t.fail("Error: no solve workspace yet.");
    const rs::Slot& sl = *T.slots[0];
    const int L = T.plan.L;
    auto put = [&](const std::vector<int>(&tile)[4]) -> int {
        if (tile[0].size() != (size_t)L) return const_cast<rs_tree*>(t)->t.fail("Error: no solve workspace yet.");
        for (int l = 0; l < L; ++l)
            for (int k = 0; k < 4; ++k) tiles[4 * l + k] = (k == 1 && l == 0) ? -1 : tile[k][l];
        return 0;
    };
    return precision == RS_MIXED ? put(sl.w32.tile) : put(sl.w64.tile);
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

}  // extern "C"

namespace {

template <typename V>
std::vector<V> to_prec(const double* p, size_t n) {
    std::vector<V> v(n);
    for (size_t q = 0; q < n; ++q) { v[q].x = (decltype(v[q].x))p[2 * q]; v[q].y = (decltype(v[q].y))p[2 * q + 1]; }
    return v;
}

// S: precision of the operands and result, R: of the products
template <typename S, typename R>
int check_gemm_t(int tile, int M, int N, int K, const double* A, int lda, const double* B, const double* B2, int ldb,
                 int ksplit, double* C, int ldc, const int* rmap, const double* alpha, const double* beta,
                 const int* bmap = nullptr) {
    using V = rs::c_t<S>;
    size_t na = (size_t)lda * K, nb = (size_t)ldb * N, nc = (size_t)ldc * N;
    std::vector<V> hA = to_prec<V>(A, na), hB = to_prec<V>(B, nb), hC = to_prec<V>(C, nc);
    std::vector<V> hB2 = B2 ? to_prec<V>(B2, nb) : std::vector<V>();
    V *dA = nullptr, *dB = nullptr, *dB2 = nullptr, *dC = nullptr, *dD = nullptr;
    int *dmap = nullptr, *dbmap = nullptr;
    rs::GemmDesc* dd = nullptr;
    auto cleanup = [&]() {
        for (void* p : {(void*)dA, (void*)dB, (void*)dB2, (void*)dC, (void*)dD, (void*)dmap, (void*)dbmap, (void*)dd})
            if (p) (void)hipFree(p);
    };
    bool ok = hipMalloc(&dA, na * sizeof(V)) == hipSuccess && hipMalloc(&dB, nb * sizeof(V)) == hipSuccess &&
              hipMalloc(&dC, nc * sizeof(V)) == hipSuccess && hipMalloc(&dD, nc * sizeof(V)) == hipSuccess &&
              hipMalloc(&dd, sizeof(rs::GemmDesc)) == hipSuccess;
    if (ok && B2) ok = hipMalloc(&dB2, nb * sizeof(V)) == hipSuccess;
    if (ok && rmap) ok = hipMalloc(&dmap, (size_t)M * sizeof(int)) == hipSuccess;
    if (ok && bmap) ok = hipMalloc(&dbmap, (size_t)K * sizeof(int)) == hipSuccess;
    if (!ok) { cleanup(); g_err = "Error: hipMalloc failed in rs_check_gemm."; return 1; }
    (void)hipMemcpy(dA, hA.data(), na * sizeof(V), hipMemcpyHostToDevice);
    (void)hipMemcpy(dB, hB.data(), nb * sizeof(V), hipMemcpyHostToDevice);
    if (B2) (void)hipMemcpy(dB2, hB2.data(), nb * sizeof(V), hipMemcpyHostToDevice);
    (void)hipMemcpy(dC, hC.data(), nc * sizeof(V), hipMemcpyHostToDevice);
    (void)hipMemcpy(dD, hC.data(), nc * sizeof(V), hipMemcpyHostToDevice);
    if (rmap) (void)hipMemcpy(dmap, rmap, (size_t)M * sizeof(int), hipMemcpyHostToDevice);
    if (bmap) (void)hipMemcpy(dbmap, bmap, (size_t)K * sizeof(int), hipMemcpyHostToDevice);
    rs::GemmDesc d;
    std::memset(&d, 0, sizeof d);
    d.A = dA; d.lda = lda; d.B = dB; d.B2 = dB2; d.ldb = ldb; d.ksplit = B2 ? ksplit : (1 << 30);
    d.Out = dC; d.ldo = ldc; d.rmap = dmap; d.dmap = rmap ? 1 : 0;
    d.bmap = dbmap;
    if (bmap) {   // the row map's pair property, as the tree sets it for its down-pass maps
        bool pairs = K % 2 == 0;
        for (int k = 0; pairs && k + 1 < K; k += 2) pairs = bmap[k] % 2 == 0 && bmap[k + 1] == bmap[k] + 1;
        d.bpair = pairs ? 1 : 0;
    }
    d.Din = dD; d.ldd = ldc;
    d.ar = alpha[0]; d.ai = alpha[1]; d.br = beta[0]; d.bi = beta[1];
    (void)hipMemcpy(dd, &d, sizeof d, hipMemcpyHostToDevice);
    void* ws = nullptr;
    size_t wse = rs::gemm_ws_elems(1, M, N, K);
    // (without its workspace the product would run unsplit: other bits than the tree's)
    if (wse && hipMalloc(&ws, wse * sizeof(rs::c_t<R>)) != hipSuccess) {
        cleanup();
        g_err = "Error: hipMalloc failed in rs_check_gemm.";
        return 1;
    }
    if (sizeof(S) != sizeof(R))
        rs::launch_gemm_c64(dd, 1, M, N, K, nullptr, ws, tile);
    else
        rs::launch_gemm<R>(dd, 1, M, N, K, nullptr, ws, tile);
    hipError_t e = hipDeviceSynchronize();
    if (ws) (void)hipFree(ws);
    if (e == hipSuccess) e = hipMemcpy(hC.data(), dC, nc * sizeof(V), hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) { g_err = hipGetErrorString(e); return 1; }
    for (size_t q = 0; q < nc; ++q) { C[2 * q] = hC[q].x; C[2 * q + 1] = hC[q].y; }
    return 0;
}

}  // namespace

extern "C" {

int rs_check_gemm_tile(int precision, int M, int N, int K, const double* A, int lda, const double* B,
                       const double* B2, int ldb, int ksplit, double* C, int ldc, const int* rmap,
                       const double* alpha, const double* beta, int tile) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { g_err = "Error: no HIP device."; return 1; }
    if (precision == RS_MIXED)
        return check_gemm_t<float, float>(tile, M, N, K, A, lda, B, B2, ldb, ksplit, C, ldc, rmap, alpha, beta);
    if (precision == RS_MIXED_MERGE)
        return check_gemm_t<double, float>(tile, M, N, K, A, lda, B, B2, ldb, ksplit, C, ldc, rmap, alpha, beta);
    return check_gemm_t<double, double>(tile, M, N, K, A, lda, B, B2, ldb, ksplit, C, ldc, rmap, alpha, beta);
}

int rs_check_gemm_bmap(int precision, int M, int N, int K, const double* A, int lda, const double* B, int ldb,
                       const int* bmap, double* C, int ldc, const double* alpha, const double* beta, int tile) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { g_err = "Error: no HIP device."; return 1; }
    if (!bmap) { g_err = "Error: no row map."; return 1; }
    for (int k = 0; k < K; ++k)
        if (bmap[k] < 0 || bmap[k] >= ldb) { g_err = "Error: row map entry out of range."; return 1; }
    const int nok = 1 << 30;
    if (precision == RS_MIXED)
        return check_gemm_t<float, float>(tile, M, N, K, A, lda, B, nullptr, ldb, nok, C, ldc, nullptr, alpha, beta,
                                          bmap);
    if (precision == RS_MIXED_MERGE)
        return check_gemm_t<double, float>(tile, M, N, K, A, lda, B, nullptr, ldb, nok, C, ldc, nullptr, alpha, beta,
                                           bmap);
    return check_gemm_t<double, double>(tile, M, N, K, A, lda, B, nullptr, ldb, nok, C, ldc, nullptr, alpha, beta,
                                        bmap);
}

int rs_check_gemm(int precision, int M, int N, int K, const double* A, int lda, const double* B, const double* B2,
                  int ldb, int ksplit, double* C, int ldc, const int* rmap, const double* alpha, const double* beta) {
    return rs_check_gemm_tile(precision, M, N, K, A, lda, B, B2, ldb, ksplit, C, ldc, rmap, alpha, beta, -1);
}

int rs_check_inverse(int n, int batch, double* mats) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { g_err = "Error: no HIP device."; return 1; }
    if (n < 1 || batch < 1 || !mats) { g_err = "Error: bad arguments."; return 1; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) { g_err = "Error: hipGetDeviceProperties failed."; return 1; }
    rs::InvPlan ip;
    size_t nn = (size_t)n * n;
    ip.build(n, batch, nn, prop.sharedMemPerBlock);
    if (!ip.fits()) { g_err = "Error: matrix too large for the interface inverses."; return 1; }
    double2 *dm = nullptr, *W = nullptr;
    int *piv = nullptr, *info = nullptr, *cm = nullptr;
    rs::GemmDesc* dd = nullptr;
    std::vector<rs::GemmDesc> descs;
    std::vector<int> cmaps;
    auto release = [&]() {
        for (void* p : {(void*)dm, (void*)W, (void*)piv, (void*)info, (void*)cm, (void*)dd})
            if (p) (void)hipFree(p);
    };
    // every allocation and copy is checked before any launch
    hipError_t e = hipMalloc(&dm, nn * batch * 16);
    if (e == hipSuccess) e = hipMalloc(&info, sizeof(int));
    if (e == hipSuccess) e = hipMemset(info, 0, sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(dm, mats, nn * batch * 16, hipMemcpyHostToDevice);
    if (e == hipSuccess && !ip.small) {
        e = hipMalloc(&W, ip.w_elems() * 16 + 16);
        if (e == hipSuccess) e = hipMalloc(&piv, (size_t)n * batch * sizeof(int));
        cmaps.resize(ip.cmap_elems());
        if (e == hipSuccess) e = hipMalloc(&cm, cmaps.size() * sizeof(int) + 4);
        if (e == hipSuccess) {
            descs.resize(ip.desc_count());
            ip.make_descs(dm, W, cm, descs.data(), cmaps.data());
            e = hipMemcpy(cm, cmaps.data(), cmaps.size() * sizeof(int), hipMemcpyHostToDevice);
        }
        if (e == hipSuccess) {
            e = hipMalloc(&dd, descs.size() * sizeof(rs::GemmDesc));
            if (e == hipSuccess)
                e = hipMemcpy(dd, descs.data(), descs.size() * sizeof(rs::GemmDesc), hipMemcpyHostToDevice);
        }
    }
    if (e != hipSuccess) {
        release();
        g_err = std::string("HIP error: ") + hipGetErrorString(e);
        return 1;
    }
    rs::launch_inverse(ip, dm, piv, W, dd, info, nullptr);
    e = hipDeviceSynchronize();
    int hinfo = 0;
    if (e == hipSuccess) e = hipMemcpy(mats, dm, nn * batch * 16, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost);
    release();
    if (e != hipSuccess) { g_err = hipGetErrorString(e); return 1; }
    if (hinfo) { g_err = "Error: singular matrix."; return 2; }
    return 0;
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
