// Request server: continuous batching of design evaluations on one GPU.
//
// Requests (a structured workload key + per-element design parameters) enter a
// queue; n_slots worker threads each drive one request at a time on their own
// HIP stream (one stream per context), so the device work of concurrent
// requests overlaps.  Contexts are cached per mesh key: the mesh, DofMap, CSR
// structure and element maps stay resident and are reused by every request on
// that mesh size.  A request is admitted only when its estimated device
// footprint fits the server's HBM budget; idle contexts of other meshes are
// evicted (least recently used first) to make room.  Launch shapes are fixed
// per context, and every reduction is deterministic, so a request's result
// does not depend on what else runs beside it.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "mastg/mastg.h"

namespace {

using Clock = std::chrono::steady_clock;
using Key = std::tuple<int, int, int, int>;

struct Req {
    int64_t id;
    mastg_request r;
    std::vector<double> params;
    mastg_request_result res{};
    Clock::time_point t_submit, t_start, t_end;
    int64_t footprint = 0;
};

struct CachedCtx {
    mastg_ctx* ctx = nullptr;
    mastg_workload* wl = nullptr;
    Key key;
    int64_t bytes = 0;
    bool busy = false;
    Clock::time_point last_use;
};

}  // namespace

struct mastg_server {
    mastg_server_opts o{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> queue;
    std::unordered_map<int64_t, std::unique_ptr<Req>> reqs;
    std::vector<std::unique_ptr<CachedCtx>> ctxs;
    std::vector<std::thread> workers;
    int64_t next_id = 1;
    int64_t bytes_in_use = 0;
    bool stop = false;
    std::string err;
};

namespace {

int64_t elems_of(const mastg_request& r) {
    if (r.kind == 1) return int64_t(r.nx) * r.ny * r.nz;
    return int64_t(r.nx) * (r.ny > 0 ? r.ny : r.nx);
}

// device bytes of a context for a structured mesh (element blocks, CSR values
// and columns, contribution maps, vectors), from the counts the structure will have
int64_t footprint(const mastg_request& r) {
    if (r.kind == 1) {
        const int64_t ne = elems_of(r);
        const int64_t nn = int64_t(r.nx + 1) * (r.ny + 1) * (r.nz + 1);
        const int64_t blocks = int64_t(3 * r.nx + 1) * (3 * r.ny + 1) * (3 * r.nz + 1) - 0;
        const int64_t nnz = blocks * 9;
        // + multigrid levels: coarse element blocks (1/8 + 1/64 + ...), their
        // node-block values and the lattice-ordered level vectors
        return ne * (64 * 9 * 8 + 64 * 4 + 8 * 4 * 2 + 8 * 3 * 8 + 24 * 8 + 8 * 4) + nnz * 12 +
               blocks * 8 + nn * 3 * 8 * 12 + ne * 700 + nn * 3 * 8 * 8 + (int64_t(64) << 20);
    }
    const int ny = r.ny > 0 ? r.ny : r.nx;
    const int64_t ne = elems_of(r);
    const int64_t nn = int64_t(r.nx + 1) * (ny + 1);
    const int64_t blocks = int64_t(3 * r.nx + 1) * (3 * ny + 1);
    const int64_t nnz = blocks * 25;
    return ne * (16 * 25 * 8 + 16 * 4 + 4 * 4 * 2 + 4 * 5 * 8) + nnz * 12 + blocks * 8 + nn * 5 * 8 * 12 +
           (int64_t(64) << 20);
}

Key key_of(const mastg_request& r) {
    return Key(r.kind, r.nx, r.kind == 1 ? r.ny : (r.ny > 0 ? r.ny : r.nx), r.kind == 1 ? r.nz : 0);
}

// takes cached context `c` out of the server's list (under the lock); the
// caller destroys it after releasing the lock
std::unique_ptr<CachedCtx> detach_locked(mastg_server* s, CachedCtx* c) {
    auto it = std::find_if(s->ctxs.begin(), s->ctxs.end(),
                           [&](const std::unique_ptr<CachedCtx>& p) { return p.get() == c; });
    std::unique_ptr<CachedCtx> out = std::move(*it);
    s->ctxs.erase(it);
    s->bytes_in_use -= out->bytes;
    return out;
}

void destroy_cached(std::vector<std::unique_ptr<CachedCtx>>& v) {
    for (auto& c : v) {
        if (c->ctx) mastg_ctx_destroy(c->ctx);
        if (c->wl) mastg_workload_destroy(c->wl);
    }
    v.clear();
}

// under the lock: an idle context for the key, or room to make one.  Idle
// contexts of other meshes are evicted (least recently used first) only when
// that makes the request fit; the evicted ones go to `victims`, destroyed by
// the caller once the lock is released.
CachedCtx* acquire_locked(mastg_server* s, Req* q, bool* need_create,
                          std::vector<std::unique_ptr<CachedCtx>>& victims) {
    const Key k = key_of(q->r);
    for (auto& c : s->ctxs)
        if (!c->busy && c->key == k) {
            c->busy = true;
            *need_create = false;
            return c.get();
        }
    const int64_t need = q->footprint;
    const int64_t short_by = s->bytes_in_use + need - s->o.hbm_budget;
    if (short_by > 0) {
        int64_t idle = 0;
        for (auto& c : s->ctxs)
            if (!c->busy) idle += c->bytes;
        if (idle < short_by) return nullptr;   // not admissible yet: keep the cache
        while (s->bytes_in_use + need > s->o.hbm_budget) {
            CachedCtx* victim = nullptr;
            for (auto& c : s->ctxs)
                if (!c->busy && (!victim || c->last_use < victim->last_use)) victim = c.get();
            victims.push_back(detach_locked(s, victim));
        }
    }
    auto c = std::make_unique<CachedCtx>();
    c->key = k;
    c->bytes = need;
    c->busy = true;
    s->bytes_in_use += need;
    CachedCtx* raw = c.get();
    s->ctxs.push_back(std::move(c));
    *need_create = true;
    return raw;
}

void run_request(mastg_server* s, Req* q, CachedCtx* cc, bool create) {
    int rc = MASTG_OK;
    if (create) {
        rc = mastg_workload_create(q->r.kind, q->r.nx, q->r.ny > 0 ? q->r.ny : q->r.nx, q->r.nz, 1, 0.0, 1.0, 0,
                                   &cc->wl);
        if (!rc) rc = mastg_ctx_create(mastg_workload_desc(cc->wl), &cc->ctx);
        if (!rc) {
            // fixed launch shapes: results independent of co-running work
            rc = mastg_set_tuning(cc->ctx, 8, 2048);
        }
    }
    double objective = 0;
    int iters = 0;
    if (!rc) {
        mastg_solve_opts so{};
        const bool plate = q->r.kind == 2;
        // plates take multigrid or CG + Jacobi (they refuse the hex8-only solvers)
        so.solver = (!plate || s->o.solver == MASTG_SOLVER_MG) ? s->o.solver : MASTG_SOLVER_CG_JACOBI;
        so.rtol = plate ? (s->o.rtol_plate > 0 ? s->o.rtol_plate : 1e-7) : (s->o.rtol > 0 ? s->o.rtol : 1e-10);
        mastg_design_result dr{};
        rc = mastg_design_eval(cc->ctx, q->params.data(), q->r.grad_out, MASTG_HOST, &so, &dr);
        objective = dr.objective;
        iters = dr.iters;
    }
    std::string msg;
    if (rc) msg = "request " + std::to_string(q->id) + ": " + mastg_last_error(cc->ctx);
    std::vector<std::unique_ptr<CachedCtx>> dead;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (rc) s->err = msg;
        q->res.objective = objective;
        q->res.iters = iters;
        q->t_end = Clock::now();
        q->res.status = rc ? rc : 2;
        q->res.t_queue = std::chrono::duration<double>(q->t_start - q->t_submit).count();
        q->res.t_run = std::chrono::duration<double>(q->t_end - q->t_start).count();
        if (rc) {
            // a context whose request failed (at set-up or in the evaluation)
            // is not reused: its device state may be partial
            dead.push_back(detach_locked(s, cc));
        } else {
            cc->busy = false;
            cc->last_use = Clock::now();
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
    destroy_cached(dead);
}

void worker(mastg_server* s) {
    for (;;) {
        Req* q = nullptr;
        CachedCtx* cc = nullptr;
        bool create = false;
        std::vector<std::unique_ptr<CachedCtx>> victims;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->stop && s->queue.empty()) return;
                // first admissible request in arrival order (bounded look-ahead)
                const size_t look = std::min<size_t>(s->queue.size(), 64);
                for (size_t i = 0; i < look && !q; i++) {
                    Req* cand = s->queue[i];
                    CachedCtx* c = acquire_locked(s, cand, &create, victims);
                    if (c) {
                        q = cand;
                        cc = c;
                        s->queue.erase(s->queue.begin() + long(i));
                    }
                }
                if (q) break;
                s->cv_work.wait(lk);
            }
            q->t_start = Clock::now();
            q->res.status = 1;
        }
        destroy_cached(victims);   // evicted contexts, outside the lock
        run_request(s, q, cc, create);
    }
}

}  // namespace

extern "C" {

int mastg_request_footprint(const mastg_request* r, int64_t* bytes) {
    if (!r || !bytes || (r->kind != 1 && r->kind != 2) || r->nx <= 0) return MASTG_ERR_INVALID;
    *bytes = footprint(*r);
    return MASTG_OK;
}

int mastg_server_create(const mastg_server_opts* o, mastg_server** out) {
    if (!o || !out || o->n_slots <= 0 || o->n_slots > 64) return MASTG_ERR_INVALID;
    auto* s = new mastg_server;
    s->o = *o;
    if (s->o.queue_capacity <= 0) s->o.queue_capacity = 4096;
    if (s->o.hbm_budget <= 0) s->o.hbm_budget = int64_t(64) << 30;
    for (int i = 0; i < s->o.n_slots; i++) s->workers.emplace_back(worker, s);
    *out = s;
    return MASTG_OK;
}

int mastg_server_submit(mastg_server* s, const mastg_request* r, int64_t* id) {
    if (!s || !r || !id || !r->params || (r->kind != 1 && r->kind != 2) || r->nx <= 0)
        return MASTG_ERR_INVALID;
    if (r->kind == 1 && (r->ny <= 0 || r->nz <= 0)) return MASTG_ERR_INVALID;
    const int64_t fp = footprint(*r);
    std::unique_ptr<Req> q(new Req);
    q->r = *r;
    q->footprint = fp;
    const int64_t ne = elems_of(*r);
    q->params.assign(r->params, r->params + ne);
    for (double p : q->params)
        if (!(p > 0.0) || !(p < 1e300)) return MASTG_ERR_NAN;
    std::lock_guard<std::mutex> lk(s->mu);
    if (fp > s->o.hbm_budget) return MASTG_ERR_ADMISSION;
    if (int(s->queue.size()) >= s->o.queue_capacity) return MASTG_ERR_BUSY;
    q->id = s->next_id++;
    q->t_submit = Clock::now();
    *id = q->id;
    s->queue.push_back(q.get());
    s->reqs[q->id] = std::move(q);
    s->cv_work.notify_one();
    return MASTG_OK;
}

int mastg_server_poll(mastg_server* s, int64_t id, mastg_request_result* res) {
    if (!s || !res) return MASTG_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return MASTG_ERR_INVALID;
    *res = it->second->res;
    if (res->status == 2 || res->status < 0) s->reqs.erase(it);   // retire on final read
    return MASTG_OK;
}

int mastg_server_wait(mastg_server* s, int64_t id, mastg_request_result* res) {
    if (!s || !res) return MASTG_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return MASTG_ERR_INVALID;
        if (it->second->res.status == 2 || it->second->res.status < 0) {
            *res = it->second->res;
            s->reqs.erase(it);
            return MASTG_OK;
        }
        s->cv_done.wait(lk);
    }
}

int mastg_server_stats(mastg_server* s, int64_t* bytes_in_use, int32_t* n_contexts, int32_t* queued) {
    if (!s) return MASTG_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (bytes_in_use) *bytes_in_use = s->bytes_in_use;
    if (n_contexts) *n_contexts = int32_t(s->ctxs.size());
    if (queued) *queued = int32_t(s->queue.size());
    return MASTG_OK;
}

const char* mastg_server_last_error(mastg_server* s) {
    if (!s) return "";
    // a copy per calling thread: the server's message may change under another worker
    thread_local std::string copy;
    std::lock_guard<std::mutex> lk(s->mu);
    copy = s->err;
    return copy.c_str();
}

int mastg_server_destroy(mastg_server* s) {
    if (!s) return MASTG_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (auto& c : s->ctxs) {
        mastg_ctx_destroy(c->ctx);
        mastg_workload_destroy(c->wl);
    }
    delete s;
    return MASTG_OK;
}

}  // extern "C"
