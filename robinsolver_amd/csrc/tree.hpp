// tree.hpp -- device-resident HPS tree: factor and solve orchestration.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <string>
#include <vector>

#include "../../include/robinsolver.h"
#include "plan.hpp"
#include "rs_kernels.hpp"

namespace rs {

// Device allocation owned by a tree; frees itself.
// The owning tree's byte counter is charged on allocation and credited on release.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    long long* counter = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        if (counter) *counter -= (long long)bytes;
        p = nullptr; bytes = 0; counter = nullptr;
    }
    template <typename T> T* as() const { return (T*)p; }
};

// Byte offsets of a solve workspace of `cap` columns inside one arena (256-byte
// aligned): per-level vectors, the K-split GEMM partials, the refinement buffers
// of the mixed path, optional column tables, and the products' descriptors.
struct WorkLayout {
    static constexpr size_t NONE = ~(size_t)0;
    int cap = 0, esz = 0;
    bool refine = false;
    std::vector<size_t> hV, gF, G, Z;          // per level (NONE where absent; gF[0] is U, G[l] slices of it)
    std::vector<size_t> dug, duh, ddt, ddg;    // descriptor arrays per level
    size_t ws = NONE, rf = NONE, rg = NONE, rj = NONE, uA = NONE, tabs = NONE;
    size_t total = 0;
};

// Solve workspace for one precision and a capacity of RHS columns.  Buffers are
// raw device pointers: into the workspace's own arena, or into a caller-side
// arena (the allocation-free entry point).
template <typename R> struct SolveWork {
    using V = c_t<R>;
    int cap = 0;
    bool refine = false;
    unsigned long long gen = 0;          // bumped on every rebuild (captured graphs depend on the layout)
    std::vector<V*> hV, gF, G, Z;        // per level
    void* ws = nullptr;                  // K-split GEMM workspace
    float2 *rf = nullptr, *rg = nullptr, *rj = nullptr, *uA = nullptr;   // refinement (complex64)
    std::vector<GemmDesc*> d_up_g, d_up_h, d_dn_t, d_dn_g;
    std::vector<int> tile[4];            // per product (up_g, up_h, dn_t, dn_g) and level: GEMM tile, -1 = by shape
    DevBuf arena;                        // the owned memory (empty for a caller-side arena)
    ~SolveWork() { clear(); }
    void clear();
    // point every buffer into base according to the layout
    void bind(char* base, const WorkLayout& lay);
};

// One level's solve products, for building their descriptors (pointers as
// bytes; esz = bytes per complex element of the pass).
struct SolveLevel {
    const char* IM;     // invM of the level's nodes, [nodes][2i][2i]
    const char* Tc;     // children's RtR maps, compact [2 nodes][T_NI | T_IN]
    char *hC, *hP;      // outgoing data of the children / of the nodes (NULL at the root)
    char *Gl, *Zl;      // the level's interface data (a slice of U), down-pass products
    char* U;            // every level's interface data and the root boundary data (column stride ldg)
    const int *prow, *dsrc;   // up-pass row scatter; the children's non-interface rows in U
    int nodes, i, b, bn, Pp, ldc, ldg, ldz, ldp, lvl, esz;
    int dpair;          // dsrc pairs contiguous and even (GemmDesc::bpair)
};

enum ProfCat { P_LEAF_UP = 0, P_LEAF_DOWN, P_GEMM_UP_G, P_GEMM_UP_H, P_GEMM_DN_T, P_GEMM_DN_G, P_GATHER,
               P_RESIDUAL, P_NCAT };

struct ProfRec {
    int cat;
    hipEvent_t a, b;
    double bytes, flops;
};

// Inputs and outputs of one tree solve pass (up + down) over all right-hand sides.
// Complex128 arrays are column tables (column c at tab[c], laid out [leaf][len]);
// complex64 arrays are contiguous [col][leaf][len].
struct PassIO {
    const double2* const* f64 = nullptr;  // interior data, complex128 ...
    const float2* f32 = nullptr;          //   ... or complex64
    const float2* jump = nullptr;         // transmission residual subtracted from the leaf outgoing data
    const double2* const* g = nullptr;    // physical boundary data (complex128) ...
    const float2* g32 = nullptr;          //   ... or complex64, contiguous
    double2* const* u64 = nullptr;        // solution out, complex128 ...
    float2* u32 = nullptr;                //   ... or complex64
    const float2* ubase32 = nullptr;      // complex128 out as a correction: u64 = ubase32 + x ...
    int uacc = 0;                         //   ... or u64 += x
};

// The right-hand sides of one solve as column tables in device memory:
// f[c] -> [leaf][n_int], g[c] -> [leaf][n_bnd], u[c] -> [leaf][m], complex128.
struct RhsCols {
    const double2* const* f = nullptr;
    const double2* const* g = nullptr;
    double2* const* u = nullptr;
};

// Solve workspaces of one concurrent caller (a stream): both precisions, and
// the column tables of the contiguous-array entry points.
// A captured solve: the launch sequence of one shape, replayed from a hipGraph.
struct GraphEnt {
    int prec = 0, nrhs = 0, n_refine = 0;
    const void* tabs = nullptr;    // the column tables the graph reads (fixed addresses)
    hipGraphExec_t exec = nullptr;
    unsigned long long used = 0;
};

struct Slot {
    SolveWork<double> w64;
    SolveWork<float> w32;
    DevBuf tabs;
    int tabs_cap = 0;
    std::vector<GraphEnt> graphs;  // at most kMaxGraphs, least recently used replaced
    unsigned long long tick = 0;
    // completion of the last solve queued on this slot, and its stream
    hipEvent_t done = nullptr;
    hipStream_t last_stream = nullptr;
    // a workspace in an arena the slot does not own (a server lane's): its base
    // and generation, and the slot's own product descriptors
    const char* ext_base = nullptr;
    unsigned long long ext_gen = 0;
    DevBuf descs;
    static constexpr int kMaxGraphs = 8;
    void drop_graphs() {
        for (GraphEnt& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }
    void clear() {
        drop_graphs(); w64.clear(); w32.clear(); tabs.release(); tabs_cap = 0;
        descs.release(); ext_base = nullptr; ext_gen = 0;
    }
    ~Slot() { drop_graphs(); if (done) (void)hipEventDestroy(done); }
};

struct Tree;
// A tree's share of a grouped pass: its right-hand sides (column tables) and
// its workspace inside the caller's memory.
struct GroupMember {
    Tree* t = nullptr;
    int nrhs = 0;
    RhsCols cols;
    char* base = nullptr;
    size_t bytes = 0;
};

struct Tree {
    rs_opts opts;
    Plan plan;
    bool planned = false, factored = false, device_ready = false;
    double f_time = 0, s_time = 0;
    std::string err;
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;   // ordering against the default stream
    size_t lds_limit = 65536;
    LeafGeom G{};
    long long dev_bytes = 0;

    // Retained solve data: the leaf block inverses, inv(M) of every merge level
    // and, per level l = 1..L, the children's [T_NI | T_IN] in a compact layout
    // (the only blocks of the node maps the solve reads), in complex128 (FP64
    // trees, or RS_RETAIN_FP64) and/or complex64 (mixed trees, or made on demand
    // for a mixed solve of an FP64 tree).
    DevBuf kint, info;
    DevBuf sinv, sinv32;
    std::vector<DevBuf*> Tn, Tn32;   // index l = 1..L
    std::vector<DevBuf*> invM, invM32;
    bool have64 = false;
    // Full node maps: per level when retained (RS_RETAIN_MAPS, for rs_node_rtr);
    // otherwise the factorization keeps two consecutive levels in tmap[l & 1].
    std::vector<DevBuf*> T;
    // the factorization's working memory (released after it unless RS_RETAIN_SCRATCH)
    DevBuf tmap[2];
    DevBuf invM_t;                // FP64 inv(M) of one level (mixed trees without FP64 solve data)
    DevBuf fscratch;              // merge scratch of the largest level
    std::vector<int> ftile;       // per level, 6 merge product sites: GEMM tile chosen by the first factor's timing
    // the launch sequence of an in-place refactor, captured, and the options it was captured under
    struct FactorKey {
        bool mixed = false, keep64 = false, keep_maps = false, skip_root = false;
        int c64_from = 0;
        bool operator==(const FactorKey& o) const {
            return mixed == o.mixed && keep64 == o.keep64 && keep_maps == o.keep_maps && skip_root == o.skip_root &&
                   c64_from == o.c64_from;
        }
    };
    hipGraphExec_t fgraph = nullptr;
    FactorKey fkey;
    void drop_factor_graph();
    // the merge's product descriptors of every level in one array (fdesc), the
    // Gauss-Jordan column maps in another (fcmap); per level the offsets of
    // each product's descriptors (elements)
    // tm > 0: the level's parent maps are formed only where the root reads them
    // (rows or columns on the root interface, tm of a child's rows; tn2 blocks of
    // the rest, from t2)
    struct MergeDescs { size_t gj = 0, w = 0, mb = 0, m12 = 0, y = 0, t = 0, cm = 0, t2 = 0; int tm = 0, tn2 = 0; };
    DevBuf fdesc, fcmap;
    std::vector<MergeDescs> fdoff;
    bool have32 = false;
    bool merge_c64 = false;      // some merge products of the factorization ran in complex64
    // maps
    std::vector<DevBuf*> d_prow;
    DevBuf d_root_leaf, d_root_loc, d_leaf_nbr;
    // Where the down pass finds every node's incoming data: all of it is held in
    // one array per right-hand side, [root boundary | interface data of level 0 |
    // ... | of level L-1] (iface_len elements; level l's slice at iface_off[l],
    // [nodes][2 isz]), and a node's incoming row is an element of it.  src_dn[l]:
    // per child of a level-l node, its bn non-interface rows ([2 nodes][bn]);
    // src_leaf: every leaf boundary row ([n_leaves][nb], node order).
    int iface_len = 0;
    std::vector<int> iface_off;
    std::vector<std::vector<int>> src_dn;
    std::vector<char> src_dn_pairs;   // per level: src_dn[l] in contiguous, even-aligned pairs
    std::vector<int> src_leaf;
    std::vector<DevBuf*> d_src_dn;
    DevBuf d_src_leaf;
    int build_sources();

    // kernel timing
    bool prof_on = false;
    std::vector<ProfRec> prof;
    std::vector<hipEvent_t> ev_pool;
    double prof_ms[P_NCAT] = {0}, prof_bytes[P_NCAT] = {0}, prof_flops[P_NCAT] = {0};
    int prof_n[P_NCAT] = {0};
    void pbeg(int cat, hipStream_t st);
    void pend(double bytes, double flops, hipStream_t st);
    int prof_collect();

    DevBuf st_f, st_g, st_u;     // staging of the host-memory API (double2)

    // solve workspaces; slot 0 serves rs_solve / rs_solve_device, a server uses
    // one slot per stream it runs this tree on
    std::vector<Slot*> slots;
    Slot& slot(int i);
    void clear_slots();
    int last_nrhs = 0;
    int last_prec = RS_FP64;
    long long peak_bytes = 0;    // high-water mark of dev_bytes

    Tree();
    ~Tree();
    int fail(const std::string& m) { err = m; return 1; }
    int alloc(DevBuf& b, size_t bytes);
    int ensure_buf(DevBuf& b, size_t bytes) { return (b.p && b.bytes == bytes) ? 0 : alloc(b, bytes); }
    template <typename T_> int upload(DevBuf& b, const std::vector<T_>& v);
    void free_factor();

    int ensure_device();
    int plan_tree(int d, int L, const double* domain, double h_tgt, double kb, const double* bp);
    int factor(const double* k_int_host);
    int factor_device(const double2* k_int_dev);
    int make_c64();
    // a factored tree to a file and back (persist.cpp)
    int save(const char* path);
    int load(const char* path);
    template <typename R> int ensure_work(SolveWork<R>& w, int nrhs, bool refine = false);
    // the arena layout of a workspace of cap columns (esz 8: complex64 passes,
    // 16: complex128), with refinement buffers and/or column tables
    WorkLayout work_layout(int cap, int esz, bool refine, bool tabs, bool descs = true) const;
    // a server lane's arena: its bytes for cap columns, the per-tree descriptor
    // bytes, and the binding of slot slot_id to it
    long long lane_bytes(int cap) const;
    long long lane_desc_bytes() const;
    int bind_slot(int slot_id, char* base, size_t bytes, unsigned long long gen, int cap, hipStream_t st);
    template <typename R> SolveLevel solve_level(const SolveWork<R>& w, int l) const;
    // time the tile shapes of every product of the workspace (at its capacity) and keep the fastest
    template <typename R> int tune_work(SolveWork<R>& w);
    // GEMM tile choices by (capacity, precision): they depend on the plan's
    // shapes only, so a workspace of a capacity already tuned copies them
    std::map<long long, std::array<std::vector<int>, 4>> tile_cache;
    // tune the solve products of capacity cap once (a workspace that is freed
    // again); used to keep the tuning out of a serving loop's dispatch path
    int prepare_tiles(int cap);
    // every point's coordinates and face id into device memory (leaf_points of all leaves)
    DevBuf d_pts_box, d_pts_pat;
    int points_device(double* xyz, int* face, hipStream_t st);
    bool tiles_pending(int cap) const {
        return opts.autotune && !tile_cache.count(2LL * cap + (opts.precision == RS_MIXED ? 0 : 1));
    }
    template <typename R> int solve_pass(SolveWork<R>& w, int nrhs, const PassIO& io, hipStream_t st);
    int residual(int nrhs, double2* const* u64, const float2* u32, const double2* const* f,
                 const double2* const* g, float2* rf, float2* rg, float2* rj, hipStream_t st);
    int solve_device(int nrhs, const double2* f, const double2* g, double2* u, int prec, hipStream_t st);
    // order a solve on stream st after the slot's previous solve (queued on any stream)
    int slot_enter(Slot& sl, hipStream_t st);
    int join_default_in(hipStream_t st);
    int join_default_out(hipStream_t st);
    int slot_leave(Slot& sl, hipStream_t st);
    int solve_cols(int nrhs, const RhsCols& c, int prec, hipStream_t st, int slot_id, bool graph = false);
    int solve_launches(int nrhs, const RhsCols& c, int prec, hipStream_t st, SolveWork<double>& w64,
                       SolveWork<float>& w32);
    // rs_solve_device with its workspace in the library's fixed device scratch
    // array: no device allocation (see capi)
    int solve_device_noalloc(int nrhs, const double2* f, const double2* g, double2* u, int prec, hipStream_t st,
                             char* scratch, size_t scratch_bytes);
    // allocate slot slot_id's workspace for cap columns at the tree's precision
    int reserve_slot(int slot_id, int cap);
    // One pass over several trees' right-hand sides (the same precision): each
    // tree's leaf stages and gathers launched per tree, every level's four
    // products launched once for all trees (grouped by height in the up pass,
    // by depth in the down pass).  A member's workspace (group_bytes of its
    // column count, product descriptors included) lies in the caller's memory.
    // Every right-hand side gets the bits of the tree's own solve.
    long long group_bytes(int nrhs) const;
    static int solve_grouped(std::vector<GroupMember>& ms, hipStream_t st, std::string& err);
    template <typename R>
    static int pass_grouped(const std::vector<Tree*>& ts, const std::vector<SolveWork<R>*>& ws,
                            const std::vector<int>& nrhs, const std::vector<PassIO>& io, hipStream_t st);
    // device bytes: a solve workspace of cap columns at precision prec (with the
    // refinement buffers of the mixed path), and the factor's resident and peak bytes
    long long work_bytes(int cap, int prec) const;
    void factor_bytes(long long* resident, long long* peak) const;
    int solve_host(int nrhs, const double* f, const double* g, int prec);
    // ||r||_2 / ||(f, g)||_2 of the global FD system per right-hand side (device arrays in)
    int residual_norms(int nrhs, const double2* f, const double2* g, const double2* u, double* out_host);
    int get_solution(int rhs, double* u) const;
    int node_rtr(int lvl, int idx, double* out, int* P) const;
};

}  // namespace rs

// the opaque handle of the C ABI
struct rs_tree {
    rs::Tree t;
};
