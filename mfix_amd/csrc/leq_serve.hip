// Serving: a request loop with continuous batching over the system tables.
//
// Requests (host A_M/B_M/VAR of any grid size) queue in arrival order.  Each
// scheduling round (mfix_leq_server_step) admits queued requests while a
// table slot is free and the HBM arena has room for the request's matrix,
// vectors and Krylov workspace, uploads them and runs their set-up, then
// advances every admitted system by one chunk of BiCGSTAB iterations - one
// launch per phase covering all of them - and retires the systems that
// finished: the solution goes back to the caller's VAR, the arena block and
// the slot are freed.  Systems admitted in later rounds join the running
// ones at the next iteration; each keeps its own state and reductions, so a
// result is bit-for-bit the same alone or in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "../../include/mfix_leq.h"
#include "leq_context.h"

namespace mfix {

// First-fit allocator over one device allocation of the HBM budget.
struct Arena {
    char *base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> length, coalesced
    void init(char *b, size_t n) {
        base = b;
        size = n;
        free_.clear();
        if (n) free_[0] = n;
    }
    bool alloc(size_t n, size_t *off) {
        n = (n + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const size_t rest = it->second - n;
            const size_t at = it->first + n;
            free_.erase(it);
            if (rest) free_[at] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = (n + 255) & ~(size_t)255;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Req {
    mfix_leq_request r;
    int status = 0;        // 0 queued, 1 running, 2 done, <0 rejected (error code)
    int table = 0, slot = -1;
    size_t off = 0, bytes = 0;
    mfix_leq_stats stats = {};
    std::chrono::steady_clock::time_point t_admit;
};

}  // namespace mfix

using namespace mfix;

struct mfix_leq_server {
    size_t budget = 0;
    int max_slots = 0;
    DevBuf pool;          // the arena
    DevBuf states;        // contiguous solver states: [table][slot]
    Arena arena;
    Group<double> g64;
    Group<float> g32;
    std::vector<int64_t> slot64, slot32;  // request id per slot, -1 free
    std::deque<int64_t> queue;
    std::map<int64_t, Req> reqs;
    int64_t next_id = 0;
    int chunk = 8;        // iterations per scheduling round
    MapSet setup64, setup32;
    std::vector<SolverState> hstate;
};

namespace {

int srv_bytes(const mfix_leq_request &r, DGrid *g_out, Part *p_out, size_t *bytes) {
    mfix_leq_grid hg = {};
    hg.imax = r.imax;
    hg.jmax = r.jmax;
    hg.kmax = r.kmax;
    hg.is_serial = 1;
    hg.icheck_bicgs = 1;
    const DGrid g = make_dgrid(hg);
    const Part P = make_part(g.n3, 256, 1);
    const int pc = parse_pc(r.pc);
    *bytes = ((9 * (size_t)g.n3 * sizeof(double) + 1023) & ~(size_t)255) + sys_bytes(g, P, pc, r.mixed != 0) + 1024;
    *g_out = g;
    *p_out = P;
    return 0;
}

// Place request `q` in slot `slot` of its table: upload A, B, VAR into its
// arena block and carve the workspace behind them.
template <typename T>
int admit(mfix_leq_server *S, Group<T> &G, Req &q, int slot, hipStream_t s) {
    DGrid g;
    Part P;
    size_t bytes;
    srv_bytes(q.r, &g, &P, &bytes);
    char *blk = S->arena.base + q.off;
    const size_t n3 = (size_t)g.n3;
    double *a = (double *)blk;
    double *b = a + 7 * n3;
    double *x = b + n3;
    char *ws = (char *)(((uintptr_t)(x + n3) + 255) & ~(uintptr_t)255);
    hipError_t e = hipMemcpyAsync(a, q.r.a_m, sizeof(double) * 7 * n3, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(b, q.r.b_m, sizeof(double) * n3, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(x, q.r.var, sizeof(double) * n3, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "request upload");
    SysDesc d = {};
    d.g = g;
    d.pc = parse_pc(q.r.pc);
    for (int c = 0; c < 4; ++c) d.sweep[c] = q.r.cmethod[c] ? q.r.cmethod[c] : ' ';
    d.tol = q.r.tol;
    d.itmax = q.r.itmax;
    d.minimize = 0;
    d.icheck = 1;
    d.delta = mixed_delta(g.n3);
    d.max_updates = 0;
    d.x = x;
    d.a = a;
    d.b = b;
    d.noise = nullptr;
    d.seed = q.r.noise_seed;
    sys_carve<T>(G.sys[slot], d, P, ws);
    // states live in one contiguous array so a round reads them back at once
    G.sys[slot].st = (SolverState *)S->states.ptr + (sizeof(T) == sizeof(float) ? S->max_slots : 0) + slot;
    G.sweep[slot] = d.sweep;
    q.slot = slot;
    q.status = 1;
    q.t_admit = std::chrono::steady_clock::now();
    return 0;
}

template <typename T>
std::vector<int> active_ids(const std::vector<int64_t> &slots) {
    std::vector<int> ids;
    for (size_t q = 0; q < slots.size(); ++q)
        if (slots[q] >= 0) ids.push_back((int)q);
    return ids;
}

}  // namespace

extern "C" {

int mfix_leq_server_create(size_t hbm_budget_bytes, int max_slots, mfix_leq_server **out) {
    if (!out || max_slots < 1 || max_slots > kHostStates / 2 || hbm_budget_bytes < ((size_t)1 << 20))
        return MFIX_LEQ_EINVAL;
    *out = nullptr;
    Context &C = ctx();
    int rc;
    if ((rc = C.ensure_init())) return rc;
    mfix_leq_server *S = new mfix_leq_server();
    S->budget = hbm_budget_bytes;
    S->max_slots = max_slots;
    if ((rc = S->pool.ensure(hbm_budget_bytes)) ||
        (rc = S->states.ensure(sizeof(SolverState) * 2 * (size_t)max_slots))) {
        S->pool.release();
        S->states.release();
        delete S;
        return rc;
    }
    S->arena.init((char *)S->pool.ptr, hbm_budget_bytes);
    S->g64.sys.assign(max_slots, Sys<double>{});
    S->g64.sweep.assign(max_slots, Sweep{'R', 'S', 'R', 'S'});
    S->g32.sys.assign(max_slots, Sys<float>{});
    S->g32.sweep.assign(max_slots, Sweep{'R', 'S', 'R', 'S'});
    S->g64.bs = S->g32.bs = 256;
    S->g64.mixed = false;
    S->g32.mixed = true;
    S->slot64.assign(max_slots, -1);
    S->slot32.assign(max_slots, -1);
    S->hstate.resize(2 * (size_t)max_slots);
    if (const char *v = std::getenv("MFIX_LEQ_SERVER_CHUNK")) S->chunk = std::max(1, std::atoi(v));
    *out = S;
    return 0;
}

int64_t mfix_leq_server_submit(mfix_leq_server *S, const mfix_leq_request *r) {
    if (!S || !r) return MFIX_LEQ_EINVAL;
    if (parse_pc(r->pc) < 0) return MFIX_LEQ_EBADPC;
    if (r->imax < 1 || r->jmax < 1 || r->kmax < 1 || !r->a_m || !r->b_m || !r->var) return MFIX_LEQ_EINVAL;
    if ((double)(r->imax + 2) * (r->jmax + 2) * (r->kmax + 2) >= 2147483647.0) return MFIX_LEQ_EINVAL;
    DGrid g;
    Part P;
    size_t bytes;
    srv_bytes(*r, &g, &P, &bytes);
    if (bytes > S->budget) return MFIX_LEQ_ENOMEM;
    if (S->queue.size() >= ((size_t)1 << 20)) return MFIX_LEQ_EBUSY;
    const int64_t id = S->next_id++;
    Req q;
    q.r = *r;
    q.bytes = bytes;
    S->reqs.emplace(id, q);
    S->queue.push_back(id);
    return id;
}

int mfix_leq_server_step(mfix_leq_server *S) {
    if (!S) return MFIX_LEQ_EINVAL;
    Context &C = ctx();
    const hipStream_t s = C.stream;
    int rc;
    // 1. admission in arrival order while a slot and arena room are free
    std::vector<int> new64, new32;
    while (!S->queue.empty()) {
        Req &q = S->reqs[S->queue.front()];
        std::vector<int64_t> &slots = q.r.mixed ? S->slot32 : S->slot64;
        auto free_slot = std::find(slots.begin(), slots.end(), (int64_t)-1);
        if (free_slot == slots.end()) break;
        size_t off;
        if (!S->arena.alloc(q.bytes, &off)) break;
        q.off = off;
        const int slot = (int)(free_slot - slots.begin());
        rc = q.r.mixed ? admit<float>(S, S->g32, q, slot, s) : admit<double>(S, S->g64, q, slot, s);
        if (rc) {
            S->arena.release(off, q.bytes);
            q.status = rc;
            S->queue.pop_front();
            continue;
        }
        *free_slot = S->queue.front();
        (q.r.mixed ? new32 : new64).push_back(slot);
        S->queue.pop_front();
    }
    // 2. set-up of the admitted systems, then one chunk over every system
    if (!new64.empty()) {
        if ((rc = upload_table<double>(S->g64, s)) || (rc = build_maps<double>(S->g64, new64, S->setup64, s)))
            return rc;
        enqueue_setup<double>(S->g64, S->setup64, s);
    }
    if (!new32.empty()) {
        if ((rc = upload_table<float>(S->g32, s)) || (rc = build_maps<float>(S->g32, new32, S->setup32, s)))
            return rc;
        enqueue_setup<float>(S->g32, S->setup32, s);
    }
    const std::vector<int> ids64 = active_ids<double>(S->slot64), ids32 = active_ids<float>(S->slot32);
    if (ids64.empty() && ids32.empty()) return 0;
    if ((rc = build_maps<double>(S->g64, ids64, S->g64.iter, s)) ||
        (rc = build_maps<float>(S->g32, ids32, S->g32.iter, s)))
        return rc;
    enqueue_iterations<double>(S->g64, S->g64.iter, S->chunk, s);
    enqueue_iterations<float>(S->g32, S->g32.iter, S->chunk, s);
    // 3. read every state back at once and retire the finished systems
    hipError_t e = hipMemcpyAsync(S->hstate.data(), S->states.ptr, sizeof(SolverState) * 2 * (size_t)S->max_slots,
                                  hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "server round");
    for (int t = 0; t < 2; ++t) {
        std::vector<int64_t> &slots = t ? S->slot32 : S->slot64;
        for (int q = 0; q < S->max_slots; ++q) {
            if (slots[q] < 0) continue;
            const SolverState &h = S->hstate[(size_t)t * S->max_slots + q];
            if (!h.done) continue;
            Req &R = S->reqs[slots[q]];
            const double *xsrc;
            int64_t n3;
            if (t) {
                const Sys<float> &Y = S->g32.sys[q];
                xsrc = h.x_cur == 1 ? Y.xb : Y.xa;
                n3 = Y.g.n3;
            } else {
                const Sys<double> &Y = S->g64.sys[q];
                xsrc = Y.xa;
                n3 = Y.g.n3;
            }
            e = hipMemcpyAsync(R.r.var, xsrc, sizeof(double) * n3, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) return hip_fail(e, "result download");
            fill_stats(h, &R.stats);
            R.stats.device_ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - R.t_admit).count();
            R.status = h.msolve_fatal ? MFIX_LEQ_EDGTSV : 2;
            S->arena.release(R.off, R.bytes);
            slots[q] = -1;
        }
    }
    return 0;
}

int mfix_leq_server_poll(mfix_leq_server *S, int64_t id, mfix_leq_stats *stats) {
    if (!S) return MFIX_LEQ_EINVAL;
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return MFIX_LEQ_EINVAL;
    const Req &R = it->second;
    if (R.status < 0) return R.status;
    if (R.status != 2) return 0;
    if (stats) *stats = R.stats;
    return 1;
}

int mfix_leq_server_pending(mfix_leq_server *S) {
    if (!S) return MFIX_LEQ_EINVAL;
    int n = (int)S->queue.size();
    for (int64_t v : S->slot64) n += v >= 0;
    for (int64_t v : S->slot32) n += v >= 0;
    return n;
}

int mfix_leq_server_destroy(mfix_leq_server *S) {
    if (!S) return 0;
    (void)hipStreamSynchronize(ctx().stream);
    S->g64.drop_graph();
    S->g32.drop_graph();
    for (DevBuf *b : {&S->pool, &S->states, &S->g64.dev, &S->g32.dev, &S->g64.iter.dev, &S->g32.iter.dev,
                      &S->setup64.dev, &S->setup32.dev})
        b->release();
    delete S;
    return 0;
}

}  // extern "C"
