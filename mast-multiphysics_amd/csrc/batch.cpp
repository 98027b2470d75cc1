// Batched design evaluations: several designs of one mesh evaluated at once.
//
// MAST evaluates one design per FunctionEvaluation::evaluate call
// (src/optimization/function_evaluation.h:107; the SIMP cantilever's
// evaluate, examples/structural/example_6/example_6.cpp:450) on one libMesh
// system.  Here a batch holds n_lanes contexts over the same problem
// description, each with its own HIP stream (and so its own captured solver
// graphs) and its own host thread.  The designs of one mastg_batch_eval call
// are taken in turn by whichever lane is free; while one lane's solve sits in
// its latency-bound coarse multigrid levels or waits for a convergence check,
// the others' kernels fill the device.  Every lane is built like
// mastg_ctx_create's context and evaluates with mastg_design_eval, and the
// lanes share one set of launch shapes, so a design's result does not depend
// on its lane or on what runs beside it: it is bitwise what one context alone
// gives.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mastg/mastg.h"

struct mastg_batch {
    std::vector<mastg_ctx*> lanes;
    std::vector<std::thread> threads;
    int64_t n_elems = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    bool shapes_shared = false;
    // the current call
    uint64_t generation = 0;
    int busy = 0;                      // lanes still working on the call
    std::atomic<int> next{0};
    int n_designs = 0;
    const double* params = nullptr;
    double* dobj = nullptr;
    int where = MASTG_HOST;
    mastg_solve_opts opts{};
    bool has_opts = false;
    mastg_design_result* out = nullptr;
    double* seconds = nullptr;
    int first_bad = -1;                // lowest failing design index of the call
    int first_rc = MASTG_OK;
    std::string err;
};

namespace {

using Clock = std::chrono::steady_clock;

// one design on one lane; the first failure (lowest design index) is kept
void run_design(mastg_batch* b, int lane, int i) {
    const int64_t off = int64_t(i) * b->n_elems;
    const Clock::time_point t0 = Clock::now();
    const int rc = mastg_design_eval(b->lanes[size_t(lane)], b->params + off, b->dobj ? b->dobj + off : nullptr,
                                     b->where, b->has_opts ? &b->opts : nullptr, b->out ? b->out + i : nullptr);
    const double s = std::chrono::duration<double>(Clock::now() - t0).count();
    if (b->seconds) b->seconds[i] = s;
    if (rc) {
        std::lock_guard<std::mutex> lk(b->mu);
        if (b->first_bad < 0 || i < b->first_bad) {
            b->first_bad = i;
            b->first_rc = rc;
            try {
                b->err = "design " + std::to_string(i) + ": " + mastg_last_error(b->lanes[size_t(lane)]);
            } catch (...) {
                b->err.clear();   // the status still reports the failure
            }
        }
    }
}

// stop and join the lane threads started so far
void stop_lanes(mastg_batch* b) {
    {
        std::lock_guard<std::mutex> lk(b->mu);
        b->stop = true;
    }
    b->cv_work.notify_all();
    for (auto& t : b->threads) t.join();
    b->threads.clear();
}

void lane_loop(mastg_batch* b, int lane) {
    uint64_t seen = 0;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(b->mu);
            b->cv_work.wait(lk, [&] { return b->stop || b->generation != seen; });
            if (b->stop) return;
            seen = b->generation;
        }
        for (;;) {
            const int i = b->next.fetch_add(1);
            if (i >= b->n_designs) break;
            run_design(b, lane, i);
        }
        std::lock_guard<std::mutex> lk(b->mu);
        if (--b->busy == 0) b->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

int mastg_batch_create(const mastg_problem_desc* desc, int n_lanes, mastg_batch** out) {
    if (!desc || !out || n_lanes < 1 || n_lanes > 16) return MASTG_ERR_INVALID;
    *out = nullptr;
    mastg_batch* b = nullptr;
    try {
        b = new mastg_batch;
        b->n_elems = desc->n_elems;
        b->lanes.reserve(size_t(n_lanes));
        b->threads.reserve(size_t(n_lanes));
        // contexts are created one after another on the calling thread
        for (int l = 0; l < n_lanes; l++) {
            mastg_ctx* c = nullptr;
            const int rc = mastg_ctx_create(desc, &c);
            if (rc) {
                for (mastg_ctx* x : b->lanes) mastg_ctx_destroy(x);
                delete b;
                return rc;
            }
            b->lanes.push_back(c);
        }
        for (int l = 0; l < n_lanes; l++) b->threads.emplace_back(lane_loop, b, l);
    } catch (...) {   // host allocation or thread creation failed
        if (b) {
            stop_lanes(b);
            for (mastg_ctx* x : b->lanes) mastg_ctx_destroy(x);
            delete b;
        }
        return MASTG_ERR_OOM;
    }
    *out = b;
    return MASTG_OK;
}

int mastg_batch_lanes(const mastg_batch* b) { return b ? int(b->lanes.size()) : 0; }

int mastg_batch_lane(mastg_batch* b, int lane, mastg_ctx** ctx) {
    if (!b || !ctx || lane < 0 || lane >= int(b->lanes.size())) return MASTG_ERR_INVALID;
    *ctx = b->lanes[size_t(lane)];
    return MASTG_OK;
}

int mastg_batch_set_option(mastg_batch* b, int option, int value) {
    if (!b) return MASTG_ERR_INVALID;
    for (mastg_ctx* c : b->lanes) {
        const int rc = mastg_set_option(c, option, value);
        if (rc) return rc;
    }
    return MASTG_OK;
}

int mastg_batch_eval(mastg_batch* b, int n_designs, const double* params, double* dobj, int where,
                     const mastg_solve_opts* opts, mastg_design_result* out, double* seconds) {
    if (!b || n_designs < 0 || (n_designs > 0 && !params) || (where != MASTG_HOST && where != MASTG_DEVICE))
        return MASTG_ERR_INVALID;
    if (n_designs == 0) return MASTG_OK;
    b->params = params;
    b->dobj = dobj;
    b->where = where;
    b->has_opts = opts != nullptr;
    if (opts) b->opts = *opts;
    b->out = out;
    b->seconds = seconds;
    b->first_bad = -1;
    b->first_rc = MASTG_OK;
    b->err.clear();
    int start = 0;
    if (!b->shapes_shared) {
        // the first design runs on lane 0 alone (its solve picks the launch
        // shapes if the solver tunes them); every other lane then takes lane
        // 0's shapes, so results do not depend on the lane
        run_design(b, 0, 0);
        start = 1;
        int lanes_w = 0, grid = 0;
        if (mastg_get_tuning(b->lanes[0], &lanes_w, &grid) == MASTG_OK)
            for (mastg_ctx* c : b->lanes) (void)mastg_set_tuning(c, lanes_w, grid);   // fixed from here on
        b->shapes_shared = true;
    }
    {
        std::unique_lock<std::mutex> lk(b->mu);
        b->n_designs = n_designs;
        b->next.store(start);
        b->busy = int(b->lanes.size());
        b->generation++;
        b->cv_work.notify_all();
        b->cv_done.wait(lk, [&] { return b->busy == 0; });
    }
    return b->first_bad < 0 ? MASTG_OK : b->first_rc;
}

const char* mastg_batch_last_error(mastg_batch* b) {
    if (!b) return "";
    std::lock_guard<std::mutex> lk(b->mu);
    return b->err.c_str();
}

int mastg_batch_destroy(mastg_batch* b) {
    if (!b) return MASTG_OK;
    stop_lanes(b);
    for (mastg_ctx* c : b->lanes) mastg_ctx_destroy(c);
    delete b;
    return MASTG_OK;
}

}  // extern "C"
