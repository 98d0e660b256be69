// st_serve.hip -- continuous-batching server of independent flow solves.
//
// Requests (parameters, grid, initial field, number of RK3 steps) are
// queued by st_server_submit.  A host thread owned by the server admits them
// in arrival order while a slot is free and the HBM footprint of the solve
// (st_solver_footprint) fits in the remaining capacity, so requests that
// arrive while others run start as soon as memory frees up.  Every active
// solve runs on its own stream; its steps are replayed from a HIP graph that
// captures K steps (launch overhead paid once per K steps), and two chunks
// are kept in flight per solve so the device never waits for the host.  The
// work of all active solves therefore overlaps on the device.  A finished
// solve is copied out, its memory returned, and the next request admitted.
//
// Launch shapes (the z-chunk of the viscous march and the chunk of the y/z
// flux marches) come from an autotuner: for each grid-shape class (each
// extent rounded up to a power of two, plus the precision) the first
// admitted solve times candidate shapes on one step each and the fastest is
// cached for every later solve of that class.  Results do not depend on the
// launch shape.
#include <hip/hip_runtime.h>

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
#include <vector>

#include "st_solver.h"

namespace {

using Clock = std::chrono::steady_clock;

struct ServeReq {
    int64_t id = 0;
    st_params p{};
    std::vector<double> grid[13];
    std::vector<double> w0;
    int nsteps = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    int rc = ST_OK;  // status of a failed request (ST_ERR_NAN for a NaN residual)
    std::vector<double> w_out;
    st_scalars sc{};
    std::string err;
};

struct Slot {
    st_solver *s = nullptr;
    ServeReq *r = nullptr;
    int64_t bytes = 0;
    int launched = 0;       // steps enqueued
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int ev_steps[2] = {0, 0};  // steps completed when each event fires
    int nev = 0;               // events in flight (oldest first)
    int head = 0;
};

using TuneKey = std::tuple<int, int, int, int, int>;

int pow2_ceil(int n) {
    int v = 1;
    while (v < n) v <<= 1;
    return v;
}

}  // namespace

struct st_server {
    int64_t capacity = 0, used = 0;
    int max_slots = 1, K = 10;
    bool autotune = true;
    int device = 0;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServeReq *> pending;
    std::map<int64_t, std::unique_ptr<ServeReq>> reqs;
    std::vector<Slot> slots;
    std::map<TuneKey, std::pair<int, int>> tune;
    bool stop = false;
    int64_t next_id = 1;
    int64_t completed = 0, failed = 0;
    Clock::time_point t0;
    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

thread_local std::string g_serr;

int sfail(int code, const std::string &msg) {
    g_serr = msg;
    return code;
}

// Time one RK3 step of a freshly initialised solve (the second of two).
double time_step(st_solver *s, hipEvent_t a, hipEvent_t b) {
    if (st_advance(s, 1)) return 1e30;
    (void)hipEventRecord(a, s->stream);
    if (st_advance(s, 1)) return 1e30;
    (void)hipEventRecord(b, s->stream);
    if (hipEventSynchronize(b) != hipSuccess) return 1e30;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

// Pick (visc_chunk, march_chunk) for this solve's shape class: coordinate
// search over {8, 16, 32} for each, one timed step per candidate, the field
// reloaded afterwards.  Cached per class.
int autotune(st_server *sv, st_solver *s, ServeReq *r) {
    const TuneKey key{pow2_ceil(r->p.nx), pow2_ceil(r->p.ny), pow2_ceil(r->p.nz), r->p.precision,
                      r->p.conservative_viscous};
    {
        // st_server_stats reads the cache size from the client thread
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->tune.find(key);
        if (it != sv->tune.end()) return st_set_tuning(s, it->second.first, it->second.second);
    }
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return ST_ERR_HIP;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return ST_ERR_HIP;
    }
    const int cand[3] = {8, 16, 32};
    int best_m = 0, best_v = 0;
    double best = 1e30;
    int rc = ST_OK;
    for (int m : cand) {
        if ((rc = st_set_tuning(s, 0, m))) break;
        if ((rc = st_copy_from_field(s, r->w0.data())) || (rc = st_init_run(s))) break;
        const double t = time_step(s, a, b);
        if (t < best) { best = t; best_m = m; }
    }
    best = 1e30;
    for (int v : cand) {
        if (rc) break;
        if ((rc = st_set_tuning(s, v, best_m))) break;
        if ((rc = st_copy_from_field(s, r->w0.data())) || (rc = st_init_run(s))) break;
        const double t = time_step(s, a, b);
        if (t < best) { best = t; best_v = v; }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->tune[key] = {best_v, best_m};
    }
    return st_set_tuning(s, best_v, best_m);
}

void release(st_server *sv, Slot &sl) {
    if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
    if (sl.graph) (void)hipGraphDestroy(sl.graph);
    for (auto &e : sl.ev)
        if (e) (void)hipEventDestroy(e);
    if (sl.s) st_solver_destroy(sl.s);
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->used -= sl.bytes;
    }
    sl = Slot();
}

void finish(st_server *sv, Slot &sl, int rc, const std::string &err) {
    ServeReq *r = sl.r;
    std::lock_guard<std::mutex> lk(sv->mu);
    r->t_done = sv->now();
    r->state = rc ? 3 : 2;
    r->rc = rc;
    if (rc) {
        r->err = err;
        sv->failed++;
    } else {
        sv->completed++;
    }
    r->w0.clear();
    r->w0.shrink_to_fit();
    for (auto &gv : r->grid) {
        gv.clear();
        gv.shrink_to_fit();
    }
    sv->cv_done.notify_all();
}

// Set up one admitted request on its slot: solver, grid, tuning, field,
// start of run, and the graph of K steps.
int admit(st_server *sv, Slot &sl) {
    ServeReq *r = sl.r;
    int rc;
    if ((rc = st_solver_create(&r->p, &sl.s))) return rc;
    const std::vector<double> *gd = r->grid;
    if ((rc = st_set_grid(sl.s, gd[0].data(), gd[1].data(), gd[2].data(), gd[3].data(), gd[4].data(), gd[5].data(),
                          gd[6].data(), gd[7].data(), gd[8].data(), gd[9].data(), gd[10].data(), gd[11].data(),
                          gd[12].data())))
        return rc;
    if (sv->autotune && (rc = autotune(sv, sl.s, r))) return rc;
    // a fresh start of run: time, step counter and buffer order as after
    // st_solver_create, whatever the tuning steps left behind
    sl.s->cur = 0;
    sl.s->last_stage = -1;
    if ((rc = st_set_time(sl.s, 0.0, r->p.cfl < 0 ? -r->p.cfl : 0.0, 0))) return rc;
    // the Newton start of T(e) back at t0 (the tuning steps left the
    // temperature of an advanced state in w_aux(6))
    if (sv->autotune && st::launch_init_temperature(sl.s) != hipSuccess) return ST_ERR_HIP;
    if ((rc = st_copy_from_field(sl.s, r->w0.data())) || (rc = st_init_run(sl.s))) return rc;
    for (auto &e : sl.ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return ST_ERR_HIP;
    // K steps flip the right-hand-side buffers 3K times: K even keeps the
    // captured buffer order valid for every replay; K a multiple of the dt
    // period keeps the captured dt launches in phase.
    const int idt = r->p.iter_dt_recompute > 0 ? r->p.iter_dt_recompute : 1;
    st_solver *s = sl.s;
    if (sv->K >= 2 && sv->K % 2 == 0 && sv->K % idt == 0 && r->nsteps >= sv->K) {
        const int64_t icyc = s->icyc;
        const int cur = s->cur, last = s->last_stage;
        if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return ST_ERR_HIP;
        rc = st_advance(s, sv->K);
        hipError_t e = hipStreamEndCapture(s->stream, &sl.graph);
        s->icyc = icyc;
        s->cur = cur;
        s->last_stage = last;
        if (rc) return rc;
        if (e != hipSuccess) return ST_ERR_HIP;
        if (hipGraphInstantiate(&sl.exec, sl.graph, nullptr, nullptr, 0) != hipSuccess) return ST_ERR_HIP;
    }
    return ST_OK;
}

// Steps per event of a solve run without a graph (steps_per_graph 0 or 1).
constexpr int kPlainChunk = 10;

// Enqueue the next chunk of steps of a slot and record its event.  Every
// chunk holds at least one step, so each call advances sl.launched.
int launch_chunk(st_server *sv, Slot &sl) {
    st_solver *s = sl.s;
    const int left = sl.r->nsteps - sl.launched;
    int n;
    if (sl.exec && left >= sv->K) {
        if (hipGraphLaunch(sl.exec, s->stream) != hipSuccess) return ST_ERR_HIP;
        n = sv->K;
        s->icyc += n;
        s->last_stage = 2;
    } else {
        const int per = sv->K > 0 ? sv->K : kPlainChunk;
        n = left < per ? left : per;
        int rc = st_advance(s, n);
        if (rc) return rc;
    }
    sl.launched += n;
    const int slot = (sl.head + sl.nev) % 2;
    if (hipEventRecord(sl.ev[slot], s->stream) != hipSuccess) return ST_ERR_HIP;
    sl.ev_steps[slot] = sl.launched;
    sl.nev++;
    return ST_OK;
}

void serve_loop(st_server *sv) {
    (void)hipSetDevice(sv->device);
    for (;;) {
        bool idle = true;
        // admission, in arrival order, while a slot is free and memory fits
        for (;;) {
            ServeReq *r = nullptr;
            int64_t need = 0;
            {
                std::unique_lock<std::mutex> lk(sv->mu);
                if (sv->stop) return;
                if (sv->pending.empty()) break;
                bool free_slot = false;
                for (auto &sl : sv->slots) free_slot |= sl.r == nullptr;
                if (!free_slot) break;
                need = st_solver_footprint(&sv->pending.front()->p);
                if (sv->used + need > sv->capacity && sv->used > 0) break;  // wait for memory
                r = sv->pending.front();
                sv->pending.pop_front();
                r->state = 1;
                r->t_admit = sv->now();
                sv->used += need;
            }
            idle = false;
            Slot *slp = nullptr;
            for (auto &sl : sv->slots)
                if (sl.r == nullptr) { slp = &sl; break; }
            Slot &sl = *slp;
            sl.r = r;
            sl.bytes = need;
            int rc = admit(sv, sl);
            if (rc) {
                finish(sv, sl, rc, st_last_error());
                release(sv, sl);
            }
        }
        // progress of the active solves
        for (auto &sl : sv->slots) {
            if (!sl.r) continue;
            while (sl.nev > 0 && hipEventQuery(sl.ev[sl.head]) == hipSuccess) {
                sl.head = (sl.head + 1) % 2;
                sl.nev--;
                idle = false;
            }
            int rc = ST_OK;
            while (rc == ST_OK && sl.nev < 2 && sl.launched < sl.r->nsteps) {
                rc = launch_chunk(sv, sl);
                idle = false;
            }
            if (rc == ST_OK && sl.nev == 0 && sl.launched >= sl.r->nsteps) {
                ServeReq *r = sl.r;
                rc = st_compute_residual(sl.s, &r->sc);
                if (rc == ST_OK || rc == ST_ERR_NAN) {
                    r->w_out.resize((size_t)5 * sl.s->g.nvol);
                    const int rc2 = st_copy_to_field(sl.s, r->w_out.data());
                    if (rc == ST_OK) rc = rc2;
                }
                finish(sv, sl, rc, rc ? st_last_error() : "");
                release(sv, sl);
                idle = false;
            } else if (rc) {
                finish(sv, sl, rc, st_last_error());
                release(sv, sl);
            }
        }
        if (idle) {
            std::unique_lock<std::mutex> lk(sv->mu);
            if (sv->stop) return;
            bool active = false;
            for (auto &sl : sv->slots) active |= sl.r != nullptr;
            if (active) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            } else {
                sv->cv_work.wait_for(lk, std::chrono::milliseconds(5));
            }
        }
    }
}

}  // namespace

extern "C" {

const char *st_server_last_error(void) { return g_serr.c_str(); }

int st_server_create(int64_t hbm_capacity_bytes, int32_t max_slots, int32_t steps_per_graph, int32_t autotune,
                     st_server **out) {
    if (!out || hbm_capacity_bytes <= 0 || max_slots < 1 || max_slots > 64 || steps_per_graph < 0)
        return sfail(ST_ERR_ARG, "bad server arguments");
    st_server *sv = new (std::nothrow) st_server();
    if (!sv) return sfail(ST_ERR_NOMEM, "server allocation failed");
    if (hipGetDevice(&sv->device) != hipSuccess) {
        delete sv;
        return sfail(ST_ERR_HIP, "no device");
    }
    sv->capacity = hbm_capacity_bytes;
    sv->max_slots = max_slots;
    sv->K = steps_per_graph;
    sv->autotune = autotune != 0;
    sv->slots.resize(max_slots);
    sv->t0 = Clock::now();
    sv->th = std::thread(serve_loop, sv);
    *out = sv;
    return ST_OK;
}

int st_server_destroy(st_server *sv) {
    if (!sv) return ST_OK;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    if (sv->th.joinable()) sv->th.join();
    for (auto &sl : sv->slots) {
        if (sl.s) (void)hipStreamSynchronize(sl.s->stream);
        release(sv, sl);
    }
    delete sv;
    return ST_OK;
}

int st_server_submit(st_server *sv, const st_params *p, const double *x, const double *y, const double *z,
                     const double *yn, const double *dcsidx, const double *dcsidxs, const double *dcsidx2,
                     const double *detady, const double *detadys, const double *detady2, const double *dzitdz,
                     const double *dzitdzs, const double *dzitdz2, const double *w0, int32_t nsteps, int64_t *ticket) {
    if (!sv || !p || !w0 || !ticket || nsteps < 0) return sfail(ST_ERR_ARG, "bad submit arguments");
    const int64_t need = st_solver_footprint(p);
    if (need <= 0 || need > sv->capacity) return sfail(ST_ERR_ARG, "request larger than the server capacity");
    auto r = std::make_unique<ServeReq>();
    r->p = *p;
    r->nsteps = nsteps;
    const int nx = p->nx, ny = p->ny, nz = p->nz, ng = p->ng;
    const size_t n[13] = {(size_t)nx + 2 * ng, (size_t)ny + 2 * ng, (size_t)nz + 2 * ng, (size_t)ny + 1,
                          (size_t)nx, (size_t)nx, (size_t)nx, (size_t)ny, (size_t)ny, (size_t)ny,
                          (size_t)nz, (size_t)nz, (size_t)nz};
    const double *src[13] = {x, y, z, yn, dcsidx, dcsidxs, dcsidx2, detady, detadys, detady2, dzitdz, dzitdzs, dzitdz2};
    for (int q = 0; q < 13; q++) {
        if (!src[q]) return sfail(ST_ERR_ARG, "null grid array");
        r->grid[q].assign(src[q], src[q] + n[q]);
    }
    const size_t nvol = (size_t)(nx + 2 * ng) * (ny + 2 * ng) * (nz + 2 * ng);
    r->w0.assign(w0, w0 + 5 * nvol);
    std::lock_guard<std::mutex> lk(sv->mu);
    r->id = sv->next_id++;
    r->t_submit = sv->now();
    *ticket = r->id;
    sv->pending.push_back(r.get());
    sv->reqs[r->id] = std::move(r);
    sv->cv_work.notify_all();
    return ST_OK;
}

// Wait (up to timeout_s, < 0: forever) for a request; copies the final field
// (may be NULL), its scalars (may be NULL) and the times of submission,
// admission and completion in seconds since the server started (may be NULL),
// then forgets the request.  Returns ST_ERR_STATE on time-out.
int st_server_wait(st_server *sv, int64_t ticket, double timeout_s, double *w_out, st_scalars *sc, double times[3]) {
    if (!sv) return sfail(ST_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return sfail(ST_ERR_ARG, "unknown ticket");
    ServeReq *r = it->second.get();
    auto done = [&] { return r->state >= 2; };
    if (timeout_s < 0) sv->cv_done.wait(lk, done);
    else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
        return sfail(ST_ERR_STATE, "request not finished");
    int rc = ST_OK;
    if (r->state == 3) rc = sfail(r->rc, "request failed: " + r->err);
    if (w_out && !r->w_out.empty()) std::memcpy(w_out, r->w_out.data(), r->w_out.size() * sizeof(double));
    if (sc) *sc = r->sc;
    if (times) {
        times[0] = r->t_submit;
        times[1] = r->t_admit;
        times[2] = r->t_done;
    }
    sv->reqs.erase(it);
    return rc;
}

// Counters: [0] queued, [1] running, [2] completed, [3] failed, [4] HBM bytes
// reserved by running solves, [5] tuned shape classes.
int st_server_stats(st_server *sv, int64_t out[6]) {
    if (!sv || !out) return sfail(ST_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    int64_t running = 0;
    for (auto &kv : sv->reqs) running += kv.second->state == 1;
    out[0] = (int64_t)sv->pending.size();
    out[1] = running;
    out[2] = sv->completed;
    out[3] = sv->failed;
    out[4] = sv->used;
    out[5] = (int64_t)sv->tune.size();
    return ST_OK;
}

}  // extern "C"
