// mgf_batch.cpp — many scenes on one device: the batched step and the served
// request loop (include/mgf.h).  Neither has a counterpart in the reference,
// which steps one CGridFluidSolver at a time (MPhysics/GridFluidSolver.cpp:
// 336-397).
//
// Every scene owns a HIP stream, so the device overlaps the kernels of
// different scenes.  A scene's step keeps its own adaptive control on the host
// (CG chunk lengths, per-scene convergence, the extrapolation's layer count),
// so host worker threads drive the scenes, each blocking only on its own
// scene's stream.  A scene computes exactly what it computes alone: the same
// kernels in the same order on its own buffers, with no state shared between
// scenes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "mgf_internal.h"

using namespace mgf;

namespace {

constexpr int kDefaultWorkers = 8;
constexpr int kMaxWorkers = 64;

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

mgf_status fail(mgf_status s, const std::string& msg) {
    set_error(msg);
    return s;
}

// Folds the statistics of one more mgf_step call into `acc`, as one mgf_step
// call over all the steps would have summed them.
void merge_stats(mgf_step_stats* acc, const mgf_step_stats& one, bool first) {
    if (first) {
        *acc = one;
        return;
    }
    acc->cg_iterations += one.cg_iterations;
    acc->refinement_sweeps += one.refinement_sweeps;
    acc->converged = acc->converged && one.converged;
    acc->residual_rel = one.residual_rel;
    acc->residual_recurrence = one.residual_recurrence;
    acc->b_norm = one.b_norm;
    acc->extrapolation_layers = one.extrapolation_layers;
    acc->steps += one.steps;
    acc->ms_advect += one.ms_advect;
    acc->ms_pressure += one.ms_pressure;
    acc->ms_extrapolate += one.ms_extrapolate;
    acc->ms_total += one.ms_total;
    acc->fluid_cells = one.fluid_cells;
    acc->true_converged = acc->true_converged && one.true_converged;
    if (!(one.div_residual <= acc->div_residual)) acc->div_residual = one.div_residual;
}

// Runs `nsteps` steps of one scene; with step_ms, one synchronised step at a
// time so that each step's latency is seen.
mgf_status run_scene(mgf_scene s, int32_t nsteps, mgf_step_stats* out, double* step_ms, std::string* err) {
    mgf_step_stats acc{};
    mgf_status result = MGF_OK;
    if (!step_ms || nsteps == 0) {
        result = mgf_step(s, nsteps, &acc);
    } else {
        for (int32_t k = 0; k < nsteps; ++k) {
            mgf_step_stats one{};
            const double t0 = now_ms();
            const mgf_status st = mgf_step(s, 1, &one);   // returns after the step's kernels finished
            step_ms[k] = now_ms() - t0;
            merge_stats(&acc, one, k == 0);
            if (st != MGF_OK && result == MGF_OK) result = st;
            if (st != MGF_OK && st != MGF_ERR_NOT_CONVERGED) break;
        }
    }
    if (result != MGF_OK && err) *err = mgf_last_error();
    if (out) *out = acc;
    return result;
}

}  // namespace

// ---------------------------------------------------------------------------
// Served loop state.
// A round of the served loop steps at least this many ready scenes in
// lockstep (one step each); below it, a request runs on its scene's own
// stream, which keeps the latency of a lightly loaded server low.
constexpr size_t kServeBatchMin = 8;

struct mgf_server_s {
    struct Request {
        int64_t ticket;
        int32_t nsteps;
        double t_submit;
        int32_t done = 0;          // steps already run (in lockstep rounds)
        double t_start = 0.0;
        mgf_step_stats acc{};
    };
    struct Resident {
        std::deque<Request> pending;
        bool running = false;
        bool closing = false;
        int64_t bytes = 0;
    };

    mgf_context ctx = nullptr;
    int64_t budget = 0;
    std::mutex mu;
    std::condition_variable cv_work;   // a scene became ready, or stop
    std::condition_variable cv_done;   // a request finished
    std::unordered_map<mgf_scene, Resident> scenes;
    std::deque<mgf_scene> ready;       // scenes with a pending request and none running
    std::deque<mgf_request_result> done;
    int64_t next_ticket = 1;
    int64_t resident = 0;
    int32_t queued = 0, running = 0;
    bool stop = false;
    bool in_round = false;             // a lockstep round is running
    std::vector<std::thread> workers;

    // Finishes a request (under the lock): its result, the scene back in the
    // ready queue when more requests wait.
    void finish_request(mgf_scene s, Resident& r, mgf_status st, double t_end) {
        const Request& q = r.pending.front();
        mgf_request_result res{};
        res.ticket = q.ticket;
        res.scene = s;
        res.steps = q.nsteps;
        res.t_submit_ms = q.t_submit;
        res.t_start_ms = q.t_start;
        res.t_end_ms = t_end;
        res.status = st;
        res.stats = q.acc;
        r.pending.pop_front();
        done.push_back(res);
    }

    // One lockstep round over the scenes in bs (taken from the ready queue,
    // marked running): one step each, then each request that has all its
    // steps, or failed, is finished.
    void round(std::vector<mgf_scene>& bs, std::unique_lock<std::mutex>& lk) {
        const double t0 = now_ms();
        for (mgf_scene s : bs) {
            Request& q = scenes.at(s).pending.front();
            if (q.done == 0) q.t_start = t0;
        }
        lk.unlock();
        std::vector<mgf_step_stats> st;
        std::vector<mgf_status> status;
        std::vector<std::string> errs;
        mgf_status r = MGF_OK;
        for (mgf_scene s : bs)
            if (r == MGF_OK) r = scene_prepare(s);
        if (r == MGF_OK) r = batch_lockstep(ctx, bs, 1, 0, st, status, errs, nullptr);
        std::vector<int64_t> bytes(bs.size(), 0);
        for (size_t i = 0; i < bs.size(); ++i) (void)mgf_scene_device_bytes(bs[i], &bytes[i]);
        const double t1 = now_ms();
        lk.lock();
        for (size_t i = 0; i < bs.size(); ++i) {
            mgf_scene s = bs[i];
            Resident& rr = scenes.at(s);
            Request& q = rr.pending.front();
            const mgf_status si = r != MGF_OK ? r : (i < status.size() ? status[i] : MGF_ERR_DEVICE);
            if (i < st.size()) merge_stats(&q.acc, st[i], q.done == 0);
            ++q.done;
            if (si != MGF_OK || q.done >= q.nsteps) finish_request(s, rr, si, t1);
            resident += bytes[i] - rr.bytes;
            rr.bytes = bytes[i];
            rr.running = false;
            --running;
            if (!rr.pending.empty()) ready.push_back(s);
        }
        cv_done.notify_all();
        if (!ready.empty()) cv_work.notify_all();
    }

    void work() {
        (void)hipSetDevice(ctx->device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_work.wait(lk, [&] { return stop || !ready.empty(); });
            if (ready.empty()) return;   // stop with nothing left
            // enough ready scenes that can step in lockstep, and no round in
            // flight: one round over all of them
            if (!in_round && ctx->lockstep) {
                std::vector<mgf_scene> bs;
                for (mgf_scene s : ready)
                    if (lockstep_eligible(s)) bs.push_back(s);
                if (bs.size() >= kServeBatchMin) {
                    std::unordered_set<mgf_scene> take(bs.begin(), bs.end());
                    std::deque<mgf_scene> rest;
                    for (mgf_scene s : ready)
                        if (!take.count(s)) rest.push_back(s);
                    ready.swap(rest);
                    for (mgf_scene s : bs) {
                        Resident& r = scenes.at(s);
                        if (r.pending.front().done == 0) --queued;
                        r.running = true;
                        ++running;
                    }
                    in_round = true;
                    round(bs, lk);
                    in_round = false;
                    continue;
                }
            }
            mgf_scene s = ready.front();
            ready.pop_front();
            Resident& r = scenes.at(s);
            Request& q = r.pending.front();
            if (q.done == 0) {
                --queued;
                q.t_start = now_ms();
            }
            const int32_t left = q.nsteps - q.done;
            r.running = true;
            ++running;
            lk.unlock();

            mgf_step_stats one{};
            const mgf_status st = run_scene(s, left, &one, nullptr, nullptr);
            const double t_end = now_ms();
            int64_t bytes = 0;
            (void)mgf_scene_device_bytes(s, &bytes);

            lk.lock();
            Resident& r2 = scenes.at(s);
            Request& q2 = r2.pending.front();
            merge_stats(&q2.acc, one, q2.done == 0);
            q2.done = q2.nsteps;
            finish_request(s, r2, st, t_end);
            resident += bytes - r2.bytes;
            r2.bytes = bytes;
            r2.running = false;
            --running;
            if (!r2.pending.empty()) ready.push_back(s);
            cv_done.notify_all();
            if (!ready.empty()) cv_work.notify_one();
        }
    }
};

extern "C" {

mgf_status mgf_scene_bytes_for(mgf_context ctx, const mgf_scene_desc* desc, int64_t* out) {
    return scene_bytes_for(ctx, desc, out);
}

mgf_status mgf_batch_step(mgf_context ctx, mgf_scene* scenes, int32_t n, int32_t nsteps, int32_t workers,
                          mgf_step_stats* stats, mgf_status* status, double* step_ms) {
    if (!ctx) return fail(MGF_ERR_INVALID, "null context");
    if (n < 0 || nsteps < 0) return fail(MGF_ERR_INVALID, "n and nsteps must be >= 0");
    if (n > 0 && !scenes) return fail(MGF_ERR_INVALID, "null scene array");
    std::unordered_set<mgf_scene> seen;
    for (int32_t i = 0; i < n; ++i) {
        if (!scenes[i]) return fail(MGF_ERR_INVALID, "null scene at index " + std::to_string(i));
        if (scenes[i]->ctx != ctx) return fail(MGF_ERR_INVALID, "scene " + std::to_string(i) + " belongs to another context");
        if (!seen.insert(scenes[i]).second) return fail(MGF_ERR_INVALID, "scene " + std::to_string(i) + " appears twice");
    }
    if (n == 0) return MGF_OK;
    // allocations a first step would make happen here, before any overlap
    for (int32_t i = 0; i < n; ++i) {
        const mgf_status st = scene_prepare(scenes[i]);
        if (st != MGF_OK) return st;
    }
    const size_t un = static_cast<size_t>(n);
    std::vector<mgf_status> st(un, MGF_OK);
    std::vector<std::string> err(un);
    std::vector<mgf_step_stats> local(un);
    // Scenes that can run in lockstep do (their pressure solves as
    // multi-scene launches); the others run on their own streams from the
    // worker threads.
    std::vector<int32_t> lock_idx, solo_idx;
    for (int32_t i = 0; i < n; ++i)
        ((ctx->lockstep && lockstep_eligible(scenes[i])) ? lock_idx : solo_idx).push_back(i);
    if (lock_idx.size() < 2) {
        solo_idx.insert(solo_idx.end(), lock_idx.begin(), lock_idx.end());
        lock_idx.clear();
    }
    if (!lock_idx.empty()) {
        std::vector<mgf_scene> ls;
        for (int32_t i : lock_idx) ls.push_back(scenes[i]);
        std::vector<mgf_step_stats> ls_stats;
        std::vector<mgf_status> ls_status;
        std::vector<std::string> ls_err;
        std::vector<double> ls_ms(step_ms ? ls.size() * size_t(nsteps) : 0);
        const mgf_status r = batch_lockstep(ctx, ls, nsteps, workers, ls_stats, ls_status, ls_err,
                                            step_ms ? ls_ms.data() : nullptr);
        const std::string batch_err = r != MGF_OK ? std::string(mgf_last_error()) : std::string();
        for (size_t k = 0; k < lock_idx.size(); ++k) {
            const size_t i = size_t(lock_idx[k]);
            local[i] = ls_stats[k];
            st[i] = ls_status[k] != MGF_OK ? ls_status[k] : r;
            err[i] = !ls_err[k].empty() ? ls_err[k] : batch_err;
            if (step_ms)
                for (int32_t q = 0; q < nsteps; ++q) step_ms[i * size_t(nsteps) + size_t(q)] = ls_ms[k * size_t(nsteps) + size_t(q)];
        }
    }
    // largest scenes first: the last scenes to start are the short ones
    std::vector<int32_t> order = solo_idx;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return scenes[a]->ncell > scenes[b]->ncell; });
    const int32_t nsolo = int32_t(order.size());
    const int W = std::max(1, std::min<int>(nsolo, workers <= 0 ? kDefaultWorkers : std::min(workers, kMaxWorkers)));
    std::atomic<int32_t> next{0};
    auto body = [&] {
        (void)hipSetDevice(ctx->device);
        for (;;) {
            const int32_t k = next.fetch_add(1);
            if (k >= nsolo) return;
            const int32_t i = order[size_t(k)];
            st[size_t(i)] = run_scene(scenes[i], nsteps, &local[size_t(i)],
                                      step_ms ? step_ms + size_t(i) * size_t(nsteps) : nullptr, &err[size_t(i)]);
        }
    };
    // Threads that cannot be started leave their share to the ones that run
    // (the calling thread always works), so a thread-creation failure only
    // narrows the overlap; nothing crosses the C boundary as an exception.
    std::vector<std::thread> pool;
    try {
        pool.reserve(size_t(W - 1));
        for (int w = 1; w < W; ++w) pool.emplace_back(body);
    } catch (...) {
    }
    body();
    for (auto& t : pool) t.join();
    mgf_status first = MGF_OK;
    for (int32_t i = 0; i < n; ++i) {
        if (stats) stats[i] = local[size_t(i)];
        if (status) status[i] = st[size_t(i)];
        if (st[size_t(i)] != MGF_OK && first == MGF_OK) {
            first = st[size_t(i)];
            set_error("scene " + std::to_string(i) + ": " + err[size_t(i)]);
        }
    }
    return first;
}

mgf_status mgf_batch_autotune(mgf_context ctx, mgf_scene* scenes, int32_t n, int32_t trial_steps, int32_t* chosen,
                              double* ms) {
    if (!ctx) return fail(MGF_ERR_INVALID, "null context");
    if (n < 0 || (n > 0 && !scenes)) return fail(MGF_ERR_INVALID, "bad scene array");
    if (trial_steps < 2) return fail(MGF_ERR_INVALID, "trial_steps must be >= 2");
    std::vector<mgf_scene> ls;
    std::unordered_set<mgf_scene> seen;
    for (int32_t i = 0; i < n; ++i) {
        if (!scenes[i] || scenes[i]->ctx != ctx) return fail(MGF_ERR_INVALID, "bad scene " + std::to_string(i));
        if (!seen.insert(scenes[i]).second) return fail(MGF_ERR_INVALID, "scene " + std::to_string(i) + " appears twice");
        if (lockstep_eligible(scenes[i])) ls.push_back(scenes[i]);
    }
    if (ls.size() < 2) return fail(MGF_ERR_INVALID, "fewer than two scenes can run in lockstep");
    for (mgf_scene s : ls) {
        const mgf_status st = scene_prepare(s);
        if (st != MGF_OK) return st;
    }
    // the V-cycle's kernels per level: fused descent / ascent on the large
    // levels only (AUTO) or on every 3-D level above the coarse cycle (FUSED)
    const int32_t cand[2] = {MGF_MG_AUTO, MGF_MG_FUSED};
    double best_ms = 0.0;
    int32_t best = MGF_MG_AUTO;
    for (int k = 0; k < 2; ++k) {
        for (mgf_scene s : ls) s->desc.mg_kernels = cand[k];
        std::vector<mgf_step_stats> st;
        std::vector<mgf_status> status;
        std::vector<std::string> err;
        std::vector<double> step_ms(ls.size() * size_t(trial_steps));
        const mgf_status r = batch_lockstep(ctx, ls, trial_steps, 0, st, status, err, step_ms.data());
        if (r != MGF_OK && r != MGF_ERR_NOT_CONVERGED) return r;
        // the fastest step after the first (which builds the plans)
        double t = step_ms[1];
        for (int32_t q = 2; q < trial_steps; ++q) t = std::min(t, step_ms[size_t(q)]);
        if (ms) ms[k] = t;
        if (k == 0 || t < best_ms) {
            best_ms = t;
            best = cand[k];
        }
    }
    for (mgf_scene s : ls) s->desc.mg_kernels = best;
    if (chosen) *chosen = best;
    return MGF_OK;
}

mgf_status mgf_server_create(mgf_context ctx, int64_t hbm_budget_bytes, int32_t slots, mgf_server* out) {
    if (!ctx || !out) return fail(MGF_ERR_INVALID, "null context or output");
    *out = nullptr;
    if (hbm_budget_bytes <= 0) return fail(MGF_ERR_INVALID, "budget must be > 0");
    const int W = slots <= 0 ? kDefaultWorkers : std::min<int>(slots, kMaxWorkers);
    mgf_server_s* srv = nullptr;
    try {
        srv = new mgf_server_s;
        srv->ctx = ctx;
        srv->budget = hbm_budget_bytes;
        for (int w = 0; w < W; ++w) srv->workers.emplace_back([srv] { srv->work(); });
    } catch (...) {
        // no exception crosses the C boundary: stop the workers that started
        if (srv) {
            {
                std::lock_guard<std::mutex> lk(srv->mu);
                srv->stop = true;
            }
            srv->cv_work.notify_all();
            for (auto& t : srv->workers) t.join();
            delete srv;
        }
        return fail(MGF_ERR_DEVICE, "could not start the server's worker threads");
    }
    *out = srv;
    return MGF_OK;
}

mgf_status mgf_server_destroy(mgf_server srv) {
    if (!srv) return fail(MGF_ERR_INVALID, "null server");
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        srv->cv_done.wait(lk, [&] { return srv->queued == 0 && srv->running == 0; });
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto& t : srv->workers) t.join();
    for (auto& kv : srv->scenes) (void)mgf_scene_destroy(kv.first);
    delete srv;
    return MGF_OK;
}

mgf_status mgf_server_open_scene(mgf_server srv, const mgf_scene_desc* desc, mgf_scene* out) {
    if (!srv || !out) return fail(MGF_ERR_INVALID, "null server or output");
    *out = nullptr;
    int64_t need = 0;
    mgf_status st = scene_bytes_for(srv->ctx, desc, &need);
    if (st != MGF_OK) return st;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->resident + need > srv->budget)
            return fail(MGF_ERR_BUDGET, "admission refused: " + std::to_string(srv->resident) + " + " +
                                            std::to_string(need) + " bytes exceed the budget of " +
                                            std::to_string(srv->budget));
        srv->resident += need;   // reserved while the scene is built
    }
    mgf_scene s = nullptr;
    st = mgf_scene_create(srv->ctx, desc, &s);
    if (st == MGF_OK) st = scene_prepare(s);
    int64_t bytes = 0;
    if (st == MGF_OK) st = mgf_scene_device_bytes(s, &bytes);
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->resident -= need;
    if (st != MGF_OK) {
        if (s) (void)mgf_scene_destroy(s);
        return st;
    }
    srv->resident += bytes;
    srv->scenes[s].bytes = bytes;
    *out = s;
    return MGF_OK;
}

mgf_status mgf_server_close_scene(mgf_server srv, mgf_scene scene) {
    if (!srv || !scene) return fail(MGF_ERR_INVALID, "null server or scene");
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        auto it = srv->scenes.find(scene);
        if (it == srv->scenes.end() || it->second.closing)
            return fail(MGF_ERR_INVALID, "scene is not open on this server");
        it->second.closing = true;
        srv->cv_done.wait(lk, [&] {
            const auto& r = srv->scenes.at(scene);
            return !r.running && r.pending.empty();
        });
        srv->resident -= srv->scenes.at(scene).bytes;
        srv->scenes.erase(scene);
    }
    return mgf_scene_destroy(scene);
}

mgf_status mgf_server_submit(mgf_server srv, mgf_scene scene, int32_t nsteps, int64_t* ticket) {
    if (!srv || !scene) return fail(MGF_ERR_INVALID, "null server or scene");
    if (nsteps < 0) return fail(MGF_ERR_INVALID, "nsteps < 0");
    std::lock_guard<std::mutex> lk(srv->mu);
    // the handle is only dereferenced once it is known to be open here
    auto it = srv->scenes.find(scene);
    if (it == srv->scenes.end() || it->second.closing) return fail(MGF_ERR_INVALID, "scene is not open on this server");
    auto& r = it->second;
    // solids added since admission count against the budget from now on;
    // while a worker steps the scene, the worker refreshes the count after
    // the request instead
    if (!r.running) {
        const int64_t bytes = scene->device_bytes;
        srv->resident += bytes - r.bytes;
        r.bytes = bytes;
    }
    const int64_t t = srv->next_ticket;
    try {
        r.pending.push_back({t, nsteps, now_ms()});
        if (!r.running && r.pending.size() == 1) srv->ready.push_back(scene);
    } catch (...) {
        if (!r.pending.empty() && r.pending.back().ticket == t) r.pending.pop_back();
        return fail(MGF_ERR_NOMEM, "request queue allocation failed");
    }
    ++srv->next_ticket;
    ++srv->queued;
    srv->cv_work.notify_one();
    if (ticket) *ticket = t;
    return MGF_OK;
}

mgf_status mgf_server_poll(mgf_server srv, mgf_request_result* out, int32_t cap, int32_t wait_ms, int32_t* count) {
    if (!srv || !count || (cap > 0 && !out)) return fail(MGF_ERR_INVALID, "null argument");
    if (cap < 0) return fail(MGF_ERR_INVALID, "cap < 0");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto ready = [&] { return !srv->done.empty() || (srv->queued == 0 && srv->running == 0); };
    if (srv->done.empty() && wait_ms != 0) {
        if (wait_ms < 0)
            srv->cv_done.wait(lk, ready);
        else
            srv->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
    }
    int32_t k = 0;
    while (k < cap && !srv->done.empty()) {
        out[k++] = srv->done.front();
        srv->done.pop_front();
    }
    *count = k;
    return MGF_OK;
}

mgf_status mgf_server_state(mgf_server srv, int64_t* resident_bytes, int32_t* queued, int32_t* running) {
    if (!srv) return fail(MGF_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(srv->mu);
    if (resident_bytes) *resident_bytes = srv->resident;
    if (queued) *queued = srv->queued;
    if (running) *running = srv->running;
    return MGF_OK;
}

}  // extern "C"
