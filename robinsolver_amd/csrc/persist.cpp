// persist.cpp -- a factored tree written to a file and read back.
//
// The reference keeps no factorization beyond its process (write_variable is a
// stub, robin_output_module.f90:157-221; robin_tree_memsize is empty,
// robin_tree_module.f90:243-250): every run re-factors.  A file here holds the
// plan's parameters (the arguments of rs_plan) and the tree's retained solve
// data -- the potential, the leaf block inverses, inv(M) of every merge level
// and the children's T_NI / T_IN per level, in each precision the tree holds --
// so that a serving process restarts without factoring.  A loaded tree solves
// with the same bits as the tree that was saved (the same data, the same plan).
//
// Layout (little-endian): an 8-byte magic, the format version, the flags and
// plan parameters (Header), then one section per buffer: a 4-byte tag, the
// 8-byte length, the bytes.  The sections' lengths are checked against the
// plan before anything is uploaded.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "tree.hpp"

namespace rs {

namespace {

constexpr char MAGIC[8] = {'R', 'S', 'T', 'R', 'E', 'E', '\0', '\x01'};
constexpr unsigned VERSION = 1;

struct Header {   // 184 bytes, no padding
    char magic[8];
    unsigned version;
    int d, L;
    int precision;        // rs_opts.precision of the saved tree
    int have64, have32;   // precisions of the solve data in the file
    int merge_c64;        // the factorization ran complex64 merge products
    int n_leaves, m, n_int, nb;
    int reserved;         // zero (keeps the doubles 8-byte aligned without padding)
    double domain[6];     // rs_plan's arguments
    double h_tgt, kb;
    double bp[8];
};

// section tags: kind * 1000 + level (kind: 1 potential, 2 leaf blocks, 3 inv(M), 4 T_NI|T_IN;
// + 100 for the complex64 copies)
enum { TAG_K = 1000, TAG_SINV = 2000, TAG_INVM = 3000, TAG_TN = 4000, TAG_C64 = 100 };

struct File {
    FILE* f = nullptr;
    ~File() { if (f) std::fclose(f); }
};

constexpr size_t CHUNK = (size_t)64 << 20;   // host staging per copy

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) return fail(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)

}  // namespace

static_assert(sizeof(Header) == 184, "saved-tree header layout");

int Tree::save(const char* path) {
    if (!factored) return fail("Error: only a factored tree can be saved.");
    if (!path || !*path) return fail("Error: no file name.");
    const Plan& P = plan;
    HIPCHK(hipStreamSynchronize(stream));
    File fh;
    fh.f = std::fopen(path, "wb");
    if (!fh.f) return fail(std::string("Error: cannot open ") + path + " for writing.");
    Header h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.magic, MAGIC, 8);
    h.version = VERSION;
    h.d = P.d; h.L = P.L;
    h.precision = opts.precision;
    h.have64 = have64 ? 1 : 0;
    h.have32 = have32 ? 1 : 0;
    h.merge_c64 = merge_c64 ? 1 : 0;
    h.n_leaves = P.n_leaves; h.m = P.m; h.n_int = P.n_int; h.nb = P.nb;
    for (int j = 0; j < 3; ++j) { h.domain[2 * j] = P.domain[0][j]; h.domain[2 * j + 1] = P.domain[1][j]; }
    h.h_tgt = P.h_tgt; h.kb = P.kb;
    for (int c = 0; c < 2; ++c)
        for (int r = 0; r < 2; ++r) {
            h.bp[2 * (c * 2 + r)] = P.bp[r][c][0];
            h.bp[2 * (c * 2 + r) + 1] = P.bp[r][c][1];
        }
    if (std::fwrite(&h, sizeof h, 1, fh.f) != 1) return fail("Error: write failed.");
    std::vector<char> stage;
    auto section = [&](unsigned tag, const DevBuf& b) -> int {
        const unsigned long long n = b.p ? (unsigned long long)b.bytes : 0ULL;
        if (std::fwrite(&tag, 4, 1, fh.f) != 1 || std::fwrite(&n, 8, 1, fh.f) != 1) return fail("Error: write failed.");
        for (size_t off = 0; off < n; off += CHUNK) {
            const size_t len = std::min((size_t)n - off, CHUNK);
            stage.resize(len);
            HIPCHK(hipMemcpy(stage.data(), (const char*)b.p + off, len, hipMemcpyDeviceToHost));
            if (std::fwrite(stage.data(), 1, len, fh.f) != len) return fail("Error: write failed.");
        }
        return 0;
    };
    if (section(TAG_K, kint)) return 1;
    if (have64 && section(TAG_SINV, sinv)) return 1;
    if (have32 && section(TAG_SINV + TAG_C64, sinv32)) return 1;
    for (int l = 0; l < P.L; ++l) {
        if (have64 && section(TAG_INVM + l, *invM[l])) return 1;
        if (have32 && section(TAG_INVM + TAG_C64 + l, *invM32[l])) return 1;
    }
    for (int l = 1; l <= P.L; ++l) {
        if (have64 && section(TAG_TN + l, *Tn[l])) return 1;
        if (have32 && section(TAG_TN + TAG_C64 + l, *Tn32[l])) return 1;
    }
    if (std::fflush(fh.f) != 0) return fail("Error: write failed.");
    return 0;
}

int Tree::load(const char* path) {
    if (!path || !*path) return fail("Error: no file name.");
    const auto t0 = std::chrono::steady_clock::now();
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return fail(std::string("Error: cannot open ") + path + ".");
    Header h;
    if (std::fread(&h, sizeof h, 1, fh.f) != 1 || std::memcmp(h.magic, MAGIC, 8) != 0)
        return fail(std::string("Error: ") + path + " is not a saved tree.");
    if (h.version != VERSION) return fail("Error: saved tree format version " + std::to_string(h.version) +
                                          " (this library reads version " + std::to_string(VERSION) + ").");
    if (!h.have64 && !h.have32) return fail("Error: the saved tree holds no solve data.");
    if (h.precision != RS_FP64 && h.precision != RS_MIXED) return fail("Error: bad precision in the saved tree.");
    // plan from the file's parameters (this also drops whatever the handle held)
    factored = false;
    if (plan_tree(h.d, h.L, h.domain, h.h_tgt, h.kb, h.bp)) return 1;
    const Plan& P = plan;
    if (P.n_leaves != h.n_leaves || P.m != h.m || P.n_int != h.n_int || P.nb != h.nb)
        return fail("Error: the saved tree's sizes do not match its plan.");
    if (ensure_device()) return 1;
    opts.precision = h.precision;
    // expected bytes of every section, from the plan
    const size_t sinv_e = (size_t)P.n_leaves * G.K * G.s * G.s;
    auto invm_e = [&](int l) { return (size_t)P.lv[l].nodes * 4 * P.lv[l].isz * P.lv[l].isz; };
    auto tn_e = [&](int l) {   // level l = 1..L: the children of level l - 1
        const Level& V = P.lv[l - 1];
        return (size_t)2 * V.nodes * 2 * V.bn * V.isz;
    };
    std::vector<char> stage;
    auto section = [&](unsigned tag, DevBuf& b, size_t bytes) -> int {
        unsigned t = 0;
        unsigned long long n = 0;
        if (std::fread(&t, 4, 1, fh.f) != 1 || std::fread(&n, 8, 1, fh.f) != 1)
            return fail(std::string("Error: ") + path + " ends early.");
        if (t != tag || n != bytes)
            return fail(std::string("Error: ") + path + ": section " + std::to_string(t) + " of " +
                        std::to_string(n) + " bytes where " + std::to_string(tag) + " of " + std::to_string(bytes) +
                        " were expected.");
        if (ensure_buf(b, bytes)) return 1;
        for (size_t off = 0; off < bytes; off += CHUNK) {
            const size_t len = std::min(bytes - off, CHUNK);
            stage.resize(len);
            if (std::fread(stage.data(), 1, len, fh.f) != len) return fail(std::string("Error: ") + path + " ends early.");
            HIPCHK(hipMemcpy((char*)b.p + off, stage.data(), len, hipMemcpyHostToDevice));
        }
        return 0;
    };
    auto lvbuf = [](std::vector<DevBuf*>& v, int l) -> DevBuf& {
        if (!v[l]) v[l] = new DevBuf;
        return *v[l];
    };
    Tn.assign(P.L + 1, nullptr); Tn32.assign(P.L + 1, nullptr);
    invM.assign(P.L, nullptr); invM32.assign(P.L, nullptr);
    T.assign(P.L + 1, nullptr);
    if (section(TAG_K, kint, (size_t)P.n_leaves * P.n_int * sizeof(double2))) return 1;
    if (ensure_buf(info, sizeof(int))) return 1;
    if (h.have64 && section(TAG_SINV, sinv, sinv_e * 16)) return 1;
    if (h.have32 && section(TAG_SINV + TAG_C64, sinv32, sinv_e * 8)) return 1;
    for (int l = 0; l < P.L; ++l) {
        if (h.have64 && section(TAG_INVM + l, lvbuf(invM, l), invm_e(l) * 16)) return 1;
        if (h.have32 && section(TAG_INVM + TAG_C64 + l, lvbuf(invM32, l), invm_e(l) * 8)) return 1;
    }
    for (int l = 1; l <= P.L; ++l) {
        if (h.have64 && section(TAG_TN + l, lvbuf(Tn, l), tn_e(l) * 16)) return 1;
        if (h.have32 && section(TAG_TN + TAG_C64 + l, lvbuf(Tn32, l), tn_e(l) * 8)) return 1;
    }
    have64 = h.have64 != 0;
    have32 = h.have32 != 0;
    merge_c64 = h.merge_c64 != 0;
    clear_slots();
    factored = true;
    f_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();   // the load's wall time
    return 0;
}

}  // namespace rs
