// server.cpp -- the request loop above the tree solve (no reference counterpart:
// the reference solves one right-hand side per call, robin_tree_module.f90:198-233).
//
// One scheduler thread owns all device submissions.  Requests are queued per
// tree in arrival order.  A lane is one stream with its own column tables and
// one workspace arena, shared by every tree the lane runs and sized to the
// largest of them; a tree run on lane j keeps in its solve slot 1 + j only the
// descriptors of its products in that arena (slot 0 stays the tree's own entry
// points').  Passes of one lane are ordered by its stream; passes of different
// lanes never share memory and overlap on the device.  Whenever a lane is
// free, the tree whose oldest queued solve is oldest gets it, and that tree's
// queued solves up to max_cols right-hand sides form one pass: the callers'
// arrays are reached through the column tables, so nothing is copied.
// Factorizations run on a worker thread on the tree's own stream, once the
// tree's earlier passes have completed; its later solves wait for it.
// Admission: a pass needs its lane's arena to be large enough for the tree and
// the tree's descriptors on that lane, a factorization its resident and
// transient bytes; both are admitted against free device memory (or a byte
// limit), evicting idle lane arenas and descriptors first.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "tree.hpp"

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

enum ReqKind { RQ_SOLVE = 0, RQ_FACTOR = 1 };

struct Req {
    int kind = RQ_SOLVE, tree = -1, nrhs = 0, state = RS_REQ_QUEUED;
    const double2* f = nullptr;
    const double2* g = nullptr;
    double2* u = nullptr;
    const double2* k = nullptr;
    double t_submit = 0, t_start = 0, t_done = 0;
    long long pass = -1;
    std::string err;
};

struct TreeRec {
    rs_tree* h = nullptr;
    rs::Tree* t = nullptr;
    std::deque<long long> q;     // request ids, arrival order
    int inflight = 0;            // passes on the device
    bool factored = false, factoring = false;
    long long fbytes = 0;        // resident factor bytes accounted to the tree
    long long ftrans = 0;        // transient bytes of a running factorization
    std::vector<long long> ws;   // per lane: bytes of the tree's descriptors there (0 = none)
    std::vector<double> ws_used; // per lane: last use of them (eviction order)
};

struct Lane {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    void** tab_dev = nullptr;    // [3][max_cols]: f, g, u column pointers
    void** tab_host = nullptr;   // pinned staging of the same
    char* arena = nullptr;       // the lane's solve workspace (every tree's passes on this lane)
    long long arena_bytes = 0;   // its size (accounted also when simulated)
    unsigned long long arena_gen = 0;   // changes with every allocation (the trees rebind)
    double arena_used = 0;       // last use (eviction order)
    bool busy = false;
    std::vector<int> trees;      // the trees of the pass in flight (one, or the members of a grouped pass)
    std::vector<rs::GroupMember> members;
    int cols = 0;
    std::vector<long long> reqs;
    long long pass = -1;
    double t_sim_end = 0;
    bool failed = false;
    std::string err;
};

}  // namespace

struct rs_server {
    rs_server_opts o{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread loop_th;
    bool stop = false;
    std::vector<std::unique_ptr<TreeRec>> trees;
    std::unordered_map<long long, Req> reqs;
    long long next_id = 1, next_pass = 0;
    unsigned long long next_gen = 0;
    int queued = 0;               // requests not yet completed
    std::vector<Lane> lanes;
    std::vector<std::thread> fworkers;
    std::vector<char> fworker_done;   // set (under the lock) as a worker's last step
    int factors_running = 0;
    std::string err;
    // statistics
    long long st_done = 0, st_rhs = 0, st_passes = 0, st_factors = 0, st_waits = 0, st_evict = 0, st_failed = 0,
              st_members = 0;
    double st_busy = 0, st_busy_since = -1;
    std::vector<double> lat;

    int fail(const std::string& m) { err = m; return 1; }
    long long held() const {
        long long b = 0;
        for (const auto& tr : trees) {
            b += tr->fbytes + tr->ftrans;
            for (long long w : tr->ws) b += w;
        }
        for (const Lane& l : lanes) b += l.arena_bytes;
        return b;
    }
    long long avail() const {
        if (o.hbm_limit > 0) return o.hbm_limit - held();
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
        return (long long)fr - o.hbm_reserve;
    }
    bool any_busy() const {
        for (const Lane& l : lanes)
            if (l.busy) return true;
        return false;
    }
    void complete(long long id, bool ok, const std::string& msg, double t) {
        auto it = reqs.find(id);
        if (it == reqs.end()) return;
        Req& r = it->second;
        r.state = ok ? RS_REQ_DONE : RS_REQ_FAILED;
        r.t_done = t;
        if (!ok) { r.err = msg; ++st_failed; }
        if (r.kind == RQ_SOLVE) {
            ++st_done;
            if (ok) { st_rhs += r.nrhs; lat.push_back(r.t_done - r.t_submit); }
        } else if (ok) {
            ++st_factors;
        }
        --queued;
    }
    bool evict_for(long long need, const TreeRec* keep_tr, int keep_lane);
    void free_arena(Lane& l) {
        if (l.arena) (void)hipFree(l.arena);
        l.arena = nullptr;
        l.arena_bytes = 0;
    }
    bool pass_done(int j);
    void finish_pass(int j);
    bool dispatch_factors();
    bool dispatch_solves(std::unique_lock<std::mutex>& lk);
    void loop();
};

// Free idle lane arenas and descriptors, least recently used first, until
// `need` bytes are available.  A lane's arena is idle when the lane has no pass
// in flight; a tree's descriptors on a lane, when the lane is not running it.
// The lane keep_lane (about to run keep_tr) keeps both.
bool rs_server::evict_for(long long need, const TreeRec* keep_tr, int keep_lane) {
    if (avail() >= need) return true;
    struct Cand { TreeRec* tr; int lane; double used; };   // tr == nullptr: the lane's arena
    std::vector<Cand> c;
    for (int j = 0; j < (int)lanes.size(); ++j)
        if (lanes[j].arena_bytes && !lanes[j].busy && j != keep_lane) c.push_back({nullptr, j, lanes[j].arena_used});
    for (int ti = 0; ti < (int)trees.size(); ++ti) {
        TreeRec* tr = trees[ti].get();
        if (tr->factoring) continue;
        for (int j = 0; j < (int)lanes.size(); ++j) {
            if (!tr->ws[j] || (tr == keep_tr && j == keep_lane)) continue;
            const std::vector<int>& run = lanes[j].trees;
            if (lanes[j].busy && std::find(run.begin(), run.end(), ti) != run.end()) continue;
            c.push_back({tr, j, tr->ws_used[j]});
        }
    }
    std::sort(c.begin(), c.end(), [](const Cand& a, const Cand& b) { return a.used < b.used; });
    for (const Cand& e : c) {
        if (avail() >= need) break;
        if (!e.tr) {
            if (!o.simulate) free_arena(lanes[e.lane]);
            lanes[e.lane].arena_bytes = 0;
        } else {
            if (!o.simulate) e.tr->t->slot(1 + e.lane).clear();
            e.tr->ws[e.lane] = 0;
        }
        ++st_evict;
    }
    return avail() >= need;
}

bool rs_server::pass_done(int j) {
    Lane& ln = lanes[j];
    if (o.simulate) return now_s() >= ln.t_sim_end;
    hipError_t e = hipEventQuery(ln.done);
    if (e == hipErrorNotReady) return false;
    if (e != hipSuccess) { ln.failed = true; ln.err = std::string("HIP error: ") + hipGetErrorString(e); }
    return true;
}

void rs_server::finish_pass(int j) {
    Lane& ln = lanes[j];
    const double t = now_s();
    for (long long id : ln.reqs) complete(id, !ln.failed, ln.err, t);
    for (int ti : ln.trees) {
        TreeRec& tr = *trees[ti];
        --tr.inflight;
        tr.ws_used[j] = t;
    }
    ln.trees.clear();
    ln.arena_used = t;
    ln.busy = false;
    ln.reqs.clear();
    ln.failed = false;
    ln.err.clear();
    if (!any_busy() && st_busy_since >= 0) { st_busy += t - st_busy_since; st_busy_since = -1; }
    cv_done.notify_all();
}

// Start queued factorizations whose trees have no pass in flight.
bool rs_server::dispatch_factors() {
    bool any = false;
    for (size_t ti = 0; ti < trees.size(); ++ti) {
        TreeRec& tr = *trees[ti];
        if (tr.factoring || tr.inflight > 0 || tr.q.empty()) continue;
        Req& r = reqs[tr.q.front()];
        if (r.kind != RQ_FACTOR) continue;
        long long resident = 0, peak = 0;
        tr.t->factor_bytes(&resident, &peak);
        // the tile tuning after a first factorization holds a workspace of its own
        if (!o.simulate && tr.t->tiles_pending(o.max_cols))
            peak = std::max(peak, resident + tr.t->work_bytes(o.max_cols, tr.t->opts.precision));
        const long long need = std::max(0LL, resident - tr.fbytes) + (peak - resident);
        if (!evict_for(need, nullptr, -1)) {
            if (!any_busy() && factors_running == 0) {
                // nothing running will free memory: the request cannot be admitted
                const long long id = tr.q.front();
                tr.q.pop_front();
                complete(id, false, "Error: not enough device memory to factor this tree.", now_s());
                cv_done.notify_all();
                any = true;
            } else {
                ++st_waits;
            }
            break;   // arrival order: later factorizations wait behind this one
        }
        const long long id = tr.q.front();
        tr.q.pop_front();
        r.state = RS_REQ_RUNNING;
        r.t_start = now_s();
        tr.factoring = true;
        tr.fbytes = std::max(tr.fbytes, resident);
        tr.ftrans = peak - resident;
        ++factors_running;
        rs::Tree* t = tr.t;
        const double2* k = r.k;
        const bool sim = o.simulate != 0;
        const double sim_s = o.sim_factor_s;
        TreeRec* trp = &tr;
        const size_t wi = fworkers.size();
        fworker_done.push_back(0);
        const int mc = o.max_cols;
        fworkers.emplace_back([this, t, k, sim, sim_s, trp, id, wi, mc]() {
            int rc = 0;
            std::string msg;
            if (sim) {
                std::this_thread::sleep_for(std::chrono::duration<double>(sim_s));
            } else {
                rc = t->factor_device(k);
                if (rc) msg = t->err;
                // the solve products' tile tuning, here rather than in the dispatch path
                else (void)t->prepare_tiles(mc);
            }
            std::lock_guard<std::mutex> g(mu);
            trp->factoring = false;
            trp->ftrans = 0;
            trp->factored = (rc == 0);
            --factors_running;
            fworker_done[wi] = 1;
            complete(id, rc == 0, msg, now_s());
            cv_work.notify_all();
            cv_done.notify_all();
        });
        any = true;
    }
    return any;
}

// Fill free lanes with coalesced passes.
bool rs_server::dispatch_solves(std::unique_lock<std::mutex>& lk) {
    bool any = false;
    const bool group = o.group > 1;
    for (int j = 0; j < (int)lanes.size(); ++j) {
        if (lanes[j].busy) continue;
        // the tree whose oldest queued solve is oldest
        int best = -1;
        double best_t = 0;
        for (int ti = 0; ti < (int)trees.size(); ++ti) {
            TreeRec& tr = *trees[ti];
            if (tr.factoring || tr.q.empty()) continue;
            const Req& r = reqs[tr.q.front()];
            if (r.kind != RQ_SOLVE) continue;
            if (!tr.factored) {
                const long long id = tr.q.front();
                tr.q.pop_front();
                complete(id, false, "Error: Robin Tree must be factored first.", now_s());
                cv_done.notify_all();
                any = true;
                continue;
            }
            if (best < 0 || r.t_submit < best_t) { best = ti; best_t = r.t_submit; }
        }
        if (best < 0) break;
        TreeRec& tr = *trees[best];
        Lane& ln = lanes[j];
        // the lane's arena, grown to this tree's workspace when it is smaller
        // (the old one is freed first), and the tree's descriptors on the lane
        // (a grouped pass keeps them in the arena)
        // (grouped: the largest factored tree's, so that any tree's pass fits
        // and smaller trees' passes can join it)
        long long wneed = group ? 0 : tr.t->lane_bytes(o.max_cols);
        if (group)
            for (const auto& m : trees)
                if (m->factored && !m->factoring) wneed = std::max(wneed, m->t->group_bytes(o.max_cols));
        const long long grow = wneed > ln.arena_bytes ? wneed : 0;
        const long long dneed = (group || tr.ws[j]) ? 0 : tr.t->lane_desc_bytes();
        const long long need = (grow ? grow - ln.arena_bytes : 0) + dneed;
        if (need > 0) {
            if (!evict_for(need, &tr, j)) {
                if (!any_busy() && factors_running == 0) {
                    while (!tr.q.empty() && reqs[tr.q.front()].kind == RQ_SOLVE) {
                        complete(tr.q.front(), false, "Error: not enough device memory for a solve workspace.",
                                 now_s());
                        tr.q.pop_front();
                    }
                    cv_done.notify_all();
                    any = true;
                } else {
                    ++st_waits;
                }
                break;   // the oldest work waits for memory; later work does not overtake it
            }
            if (grow) {
                if (!o.simulate) {
                    free_arena(ln);
                    if (hipMalloc((void**)&ln.arena, (size_t)grow) != hipSuccess) {
                        (void)hipGetLastError();
                        ln.arena = nullptr;
                    }
                }
                ln.arena_bytes = (o.simulate || ln.arena) ? grow : 0;
                ln.arena_gen = ++next_gen;
            }
            if (!group) tr.ws[j] = tr.t->lane_desc_bytes();
        }
        if (!o.simulate && (!ln.arena_bytes ||
                            (!group && tr.t->bind_slot(1 + j, ln.arena, (size_t)ln.arena_bytes, ln.arena_gen,
                                                       o.max_cols, ln.st)))) {
            const std::string m = ln.arena_bytes ? tr.t->err : std::string("Error: could not allocate the lane's "
                                                                           "solve workspace.");
            tr.t->slot(1 + j).clear();
            tr.ws[j] = 0;
            while (!tr.q.empty() && reqs[tr.q.front()].kind == RQ_SOLVE) {
                complete(tr.q.front(), false, m, now_s());
                tr.q.pop_front();
            }
            cv_done.notify_all();
            any = true;
            continue;
        }
        // coalesce the tree's queued solves; a grouped pass then takes the
        // queued solves of further trees (oldest first, the same precision)
        // while the pass stays within max_cols and the lane's arena
        ln.reqs.clear();
        ln.trees.clear();
        ln.members.clear();
        int cols = 0;
        size_t used = 0;
        const double t0 = now_s();
        const long long pass = next_pass++;
        std::vector<std::vector<const Req*>> batch;
        auto take = [&](int ti, int room) {
            TreeRec& m = *trees[ti];
            std::vector<const Req*> b;
            int c = 0;
            while (!m.q.empty()) {
                Req& r = reqs[m.q.front()];
                if (r.kind != RQ_SOLVE || c + r.nrhs > room) break;
                r.state = RS_REQ_RUNNING;
                r.t_start = t0;
                r.pass = pass;
                c += r.nrhs;
                ln.reqs.push_back(m.q.front());
                b.push_back(&r);
                m.q.pop_front();
            }
            ln.trees.push_back(ti);
            ++m.inflight;
            rs::GroupMember gm;
            gm.t = m.t;
            gm.nrhs = c;
            if (group) {
                gm.bytes = (size_t)((m.t->group_bytes(c) + 255) / 256 * 256);
                gm.base = ln.arena ? ln.arena + used : nullptr;
                used += gm.bytes;
            }
            ln.members.push_back(gm);
            batch.push_back(b);
            cols += c;
        };
        take(best, o.max_cols);
        if (group) {
            std::vector<std::pair<double, int>> cand;
            for (int ti = 0; ti < (int)trees.size(); ++ti) {
                TreeRec& m = *trees[ti];
                if (ti == best || m.factoring || !m.factored || m.q.empty()) continue;
                if (m.t->opts.precision != tr.t->opts.precision) continue;
                const Req& r = reqs[m.q.front()];
                if (r.kind != RQ_SOLVE) continue;
                cand.push_back({r.t_submit, ti});
            }
            std::sort(cand.begin(), cand.end());
            for (const auto& cd : cand) {
                if ((int)ln.trees.size() >= o.group || cols >= o.max_cols) break;
                TreeRec& m = *trees[cd.second];
                const int first = reqs[m.q.front()].nrhs;
                if (cols + first > o.max_cols) continue;
                // the columns that fit, and the workspace they need
                int c = 0;
                for (long long id : m.q) {
                    const Req& r = reqs[id];
                    if (r.kind != RQ_SOLVE || cols + c + r.nrhs > o.max_cols) break;
                    c += r.nrhs;
                }
                const size_t bytes = (size_t)((m.t->group_bytes(c) + 255) / 256 * 256);
                if (used + bytes > (size_t)ln.arena_bytes) continue;
                take(cd.second, o.max_cols - cols);
            }
        }
        ln.busy = true;
        ln.cols = cols;
        ln.pass = pass;
        ++st_passes;
        st_members += (long long)ln.trees.size();
        if (st_busy_since < 0) st_busy_since = t0;
        any = true;
        if (o.simulate) {
            ln.t_sim_end = t0 + o.sim_pass_s + o.sim_col_s * cols;
            continue;
        }
        // column tables of the pass: the members' columns one after another
        const int mc = o.max_cols;
        int c = 0;
        for (size_t q = 0; q < ln.members.size(); ++q) {
            const rs::Plan& P = ln.members[q].t->plan;
            const size_t nf = (size_t)P.n_leaves * P.n_int, ng = (size_t)P.n_leaves * P.nb,
                         nu = (size_t)P.n_leaves * P.m;
            rs::RhsCols& rc_cols = ln.members[q].cols;
            rc_cols.f = (const double2* const*)(ln.tab_dev + c);
            rc_cols.g = (const double2* const*)(ln.tab_dev + mc + c);
            rc_cols.u = (double2* const*)(ln.tab_dev + 2 * mc + c);
            for (const Req* r : batch[q])
                for (int k = 0; k < r->nrhs; ++k, ++c) {
                    ln.tab_host[c] = (void*)(r->f + k * nf);
                    ln.tab_host[mc + c] = (void*)(r->g + k * ng);
                    ln.tab_host[2 * mc + c] = (void*)(r->u + k * nu);
                }
        }
        rs::Tree* t = tr.t;
        // launch without the lock: submitters may queue meanwhile (the lane, its
        // arena and the members' slots 1 + j belong to this thread until the pass
        // completes)
        lk.unlock();
        std::string msg;
        int rc = 0;
        if (hipMemcpyAsync(ln.tab_dev, ln.tab_host, (size_t)3 * mc * sizeof(void*), hipMemcpyHostToDevice,
                           ln.st) != hipSuccess) {
            rc = 1;
            msg = "HIP error: column table upload failed.";
        } else if (group) {
            rc = rs::Tree::solve_grouped(ln.members, ln.st, msg);
        } else {
            rs::Slot& sl = t->slot(1 + j);
            rc = t->slot_enter(sl, ln.st) || t->solve_cols(cols, ln.members[0].cols, -1, ln.st, 1 + j) ||
                 t->slot_leave(sl, ln.st);
            if (rc) msg = t->err;
        }
        if (hipEventRecord(ln.done, ln.st) != hipSuccess && !rc) {
            rc = 1;
            msg = "HIP error: event record failed.";
        }
        lk.lock();
        if (rc) {
            ln.failed = true;
            ln.err = msg;
        }
    }
    return any;
}

void rs_server::loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        bool progress = false;
        for (int j = 0; j < (int)lanes.size(); ++j)
            if (lanes[j].busy && pass_done(j)) { finish_pass(j); progress = true; }
        // join finished factor workers
        for (size_t w = 0; w < fworkers.size(); ++w)
            if (fworker_done[w] == 1 && fworkers[w].joinable()) { fworkers[w].join(); fworker_done[w] = 2; }
        if (stop && queued == 0 && !any_busy() && factors_running == 0) break;
        progress |= dispatch_factors();
        progress |= dispatch_solves(lk);
        if (!progress) {
            if (any_busy()) {
                // poll the passes in flight; new requests are picked up on the next turn
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            } else {
                cv_work.wait_for(lk, std::chrono::milliseconds(5));
            }
        }
    }
}

namespace {
thread_local std::string g_srv_msg;
std::string g_srv_err;
}  // namespace

extern "C" {

void rs_server_opts_default(rs_server_opts* o) {
    if (!o) return;
    *o = rs_server_opts{};
    o->max_cols = 256;
    o->lanes = 2;
    o->hbm_reserve = 2LL << 30;
    o->hbm_limit = 0;
    o->simulate = 0;
    o->sim_pass_s = 1e-3;
    o->sim_col_s = 1e-5;
    o->sim_factor_s = 5e-3;
    o->group = 1;
}

int rs_server_create(rs_server** out, const rs_server_opts* o) {
    if (!out) { g_srv_err = "Error: null handle pointer."; return 1; }
    *out = nullptr;
    rs_server_opts oo;
    rs_server_opts_default(&oo);
    if (o) oo = *o;
    if (oo.max_cols < 1 || oo.lanes < 1 || oo.lanes > 8) {
        g_srv_err = "Error: max_cols must be positive and lanes in 1..8.";
        return 1;
    }
    if (oo.group < 1 || oo.group > rs::GEMM_MAX_GROUPS) {
        g_srv_err = "Error: group must be in 1..8 (trees per pass).";
        return 1;
    }
    if (oo.simulate && oo.hbm_limit <= 0) {
        g_srv_err = "Error: a simulated server needs hbm_limit.";
        return 1;
    }
    std::unique_ptr<rs_server> s(new rs_server);
    s->o = oo;
    s->lanes.resize(oo.lanes);
    if (!oo.simulate) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
            g_srv_err = "Error: no HIP device (the MI355X library needs a GPU).";
            return 1;
        }
        for (Lane& l : s->lanes) {
            if (hipStreamCreateWithFlags(&l.st, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&l.done, hipEventDisableTiming) != hipSuccess ||
                hipMalloc((void**)&l.tab_dev, (size_t)3 * oo.max_cols * sizeof(void*)) != hipSuccess ||
                hipHostMalloc((void**)&l.tab_host, (size_t)3 * oo.max_cols * sizeof(void*), 0) != hipSuccess) {
                g_srv_err = "Error: could not create the server's streams and tables.";
                for (Lane& m : s->lanes) {
                    if (m.st) (void)hipStreamDestroy(m.st);
                    if (m.done) (void)hipEventDestroy(m.done);
                    if (m.tab_dev) (void)hipFree(m.tab_dev);
                    if (m.tab_host) (void)hipHostFree(m.tab_host);
                }
                return 1;
            }
        }
    }
    rs_server* sp = s.release();
    sp->loop_th = std::thread([sp]() { sp->loop(); });
    *out = sp;
    return 0;
}

int rs_server_destroy(rs_server* s) {
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->loop_th.joinable()) s->loop_th.join();
    for (std::thread& t : s->fworkers)
        if (t.joinable()) t.join();
    for (Lane& l : s->lanes) {
        if (l.st) (void)hipStreamSynchronize(l.st);
        if (l.st) (void)hipStreamDestroy(l.st);
        if (l.done) (void)hipEventDestroy(l.done);
        if (l.tab_dev) (void)hipFree(l.tab_dev);
        if (l.tab_host) (void)hipHostFree(l.tab_host);
        s->free_arena(l);
    }
    delete s;
    return 0;
}

int rs_server_add_tree(rs_server* s, rs_tree* t, int* tree_id) {
    if (!s || !t) return 1;
    // tile tuning of a factored tree's solve passes here (outside the lock: the
    // tree is not registered yet), not in the dispatch path
    if (!s->o.simulate && t->t.planned && t->t.factored) (void)t->t.prepare_tiles(s->o.max_cols);
    std::lock_guard<std::mutex> g(s->mu);
    if (!t->t.planned) return s->fail("Error: rs_plan must be called before rs_server_add_tree.");
    for (const auto& tr : s->trees)
        if (tr->h == t) return s->fail("Error: tree already registered with this server.");
    // A served tree keeps its factorization's working memory: a refactor then
    // allocates and frees nothing (a device free waits for the device to drain,
    // which stalls every lane's passes meanwhile).
    t->t.opts.retain |= RS_RETAIN_SCRATCH;
    std::unique_ptr<TreeRec> tr(new TreeRec);
    tr->h = t;
    tr->t = &t->t;
    tr->factored = !s->o.simulate && t->t.factored;
    if (tr->factored) {
        long long res = 0, peak = 0;
        t->t.factor_bytes(&res, &peak);
        tr->fbytes = res;
    }
    tr->ws.assign(s->lanes.size(), 0);
    tr->ws_used.assign(s->lanes.size(), 0.0);
    s->trees.push_back(std::move(tr));
    if (tree_id) *tree_id = (int)s->trees.size() - 1;
    return 0;
}

int rs_server_factor(rs_server* s, int tree_id, const void* k_dev, long long* req) {
    if (!s) return 1;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (tree_id < 0 || tree_id >= (int)s->trees.size()) return s->fail("Error: no such tree.");
        if (!k_dev && !s->o.simulate) return s->fail("Error: null potential.");
        const long long id = s->next_id++;
        Req& r = s->reqs[id];
        r.kind = RQ_FACTOR;
        r.tree = tree_id;
        r.k = (const double2*)k_dev;
        r.t_submit = now_s();
        s->trees[tree_id]->q.push_back(id);
        ++s->queued;
        if (req) *req = id;
    }
    s->cv_work.notify_all();
    return 0;
}

int rs_server_solve(rs_server* s, int tree_id, int nrhs, const void* f_dev, const void* g_dev, void* u_dev,
                    long long* req) {
    if (!s) return 1;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (tree_id < 0 || tree_id >= (int)s->trees.size()) return s->fail("Error: no such tree.");
        if (nrhs < 1 || nrhs > s->o.max_cols) return s->fail("Error: nrhs must be in 1..max_cols.");
        if (!s->o.simulate && (!f_dev || !g_dev || !u_dev)) return s->fail("Error: null array.");
        const long long id = s->next_id++;
        Req& r = s->reqs[id];
        r.kind = RQ_SOLVE;
        r.tree = tree_id;
        r.nrhs = nrhs;
        r.f = (const double2*)f_dev;
        r.g = (const double2*)g_dev;
        r.u = (double2*)u_dev;
        r.t_submit = now_s();
        s->trees[tree_id]->q.push_back(id);
        ++s->queued;
        if (req) *req = id;
    }
    s->cv_work.notify_all();
    return 0;
}

int rs_server_test(rs_server* s, long long req, int* state) {
    if (!s || !state) return 1;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end()) return s->fail("Error: no such request.");
    *state = it->second.state;
    return 0;
}

int rs_server_wait(rs_server* s, long long req) {
    if (!s) return 1;
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        auto it = s->reqs.find(req);
        if (it == s->reqs.end()) return s->fail("Error: no such request.");
        if (it->second.state == RS_REQ_DONE) return 0;
        if (it->second.state == RS_REQ_FAILED) return 1;
        s->cv_done.wait_for(lk, std::chrono::milliseconds(5));
    }
}

int rs_server_drain(rs_server* s) {
    if (!s) return 1;
    std::unique_lock<std::mutex> lk(s->mu);
    while (s->queued > 0) s->cv_done.wait_for(lk, std::chrono::milliseconds(5));
    return 0;
}

int rs_server_request_info(rs_server* s, long long req, double* t_submit, double* t_start, double* t_done,
                           long long* pass) {
    if (!s) return 1;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end()) return s->fail("Error: no such request.");
    const Req& r = it->second;
    if (t_submit) *t_submit = r.t_submit;
    if (t_start) *t_start = r.t_start;
    if (t_done) *t_done = r.t_done;
    if (pass) *pass = r.pass;
    return 0;
}

int rs_server_release(rs_server* s, long long req) {
    if (!s) return 1;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end()) return s->fail("Error: no such request.");
    if (it->second.state != RS_REQ_DONE && it->second.state != RS_REQ_FAILED)
        return s->fail("Error: request has not completed.");
    s->reqs.erase(it);
    return 0;
}

int rs_server_get_stats(rs_server* s, rs_server_stats* st) {
    if (!s || !st) return 1;
    std::lock_guard<std::mutex> g(s->mu);
    *st = rs_server_stats{};
    st->requests_done = s->st_done;
    st->rhs_done = s->st_rhs;
    st->passes = s->st_passes;
    st->factors_done = s->st_factors;
    st->admission_waits = s->st_waits;
    st->evictions = s->st_evict;
    st->failed = s->st_failed;
    st->busy_s = s->st_busy + (s->st_busy_since >= 0 ? now_s() - s->st_busy_since : 0.0);
    std::vector<double> l = s->lat;
    std::sort(l.begin(), l.end());
    auto q = [&](double p) -> double {
        if (l.empty()) return 0.0;
        size_t i = (size_t)std::min<double>((double)l.size() - 1, std::ceil(p * l.size()) - 1);
        return l[i];
    };
    st->lat_p50_s = q(0.50);
    st->lat_p99_s = q(0.99);
    st->lat_max_s = l.empty() ? 0.0 : l.back();
    st->mean_cols = s->st_passes ? (double)s->st_rhs / (double)s->st_passes : 0.0;
    st->mean_trees = s->st_passes ? (double)s->st_members / (double)s->st_passes : 0.0;
    st->held_bytes = s->held();
    return 0;
}

int rs_server_reset_stats(rs_server* s) {
    if (!s) return 1;
    std::lock_guard<std::mutex> g(s->mu);
    s->st_done = s->st_rhs = s->st_passes = s->st_factors = s->st_waits = s->st_evict = s->st_failed = 0;
    s->st_members = 0;
    s->st_busy = 0;
    s->st_busy_since = s->any_busy() ? now_s() : -1;
    s->lat.clear();
    return 0;
}

int rs_server_lane_bytes(const rs_tree* t, int max_cols, long long* arena, long long* descs) {
    if (!t || !t->t.planned || max_cols < 1) return 1;
    if (arena) *arena = t->t.lane_bytes(max_cols);
    if (descs) *descs = t->t.lane_desc_bytes();
    return 0;
}

const char* rs_server_error(rs_server* s, long long req) {
    if (!s) return g_srv_err.c_str();
    std::lock_guard<std::mutex> g(s->mu);
    if (req < 0) {
        g_srv_msg = s->err;
    } else {
        auto it = s->reqs.find(req);
        g_srv_msg = (it == s->reqs.end()) ? std::string("Error: no such request.") : it->second.err;
    }
    return g_srv_msg.c_str();
}

}  // extern "C"
