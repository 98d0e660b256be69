// st_serve.hip -- continuous-batching server of independent flow solves.
//
// Requests (parameters, grid, initial field, number of RK3 steps, optional
// start time and checkpoint period) are queued by st_server_submit2 and
// admitted in arrival order while a slot is free and their HBM footprint
// (st_solver_footprint) fits in the server's memory pool.
//
// Memory.  The server takes its whole HBM capacity as one device
// allocation when it is created and carves every admitted solve out of it
// (first fit on 256-byte blocks, neighbouring free blocks merged on
// release).  Admission by footprint is therefore an allocation in that
// pool, and no hipMalloc, hipFree or other device-wide synchronisation
// runs while solves are on the device.  Each slot owns, for the server's
// life, its stream, its pinned scalar mirror and table staging and its
// events.
//
// Threads.  The scheduler thread never waits on the device: it admits
// requests (pool block + slot), launches chunks of K steps on each active
// slot's stream (from a captured HIP graph once the class's launch shape
// is settled), keeps two chunks in flight per slot and polls their events.
// A worker thread does what blocks: the set-up of an admitted solve on its
// slot (tables, grid and field copies, start of run), the hand-back of a
// finished one (residual, field copy, pool block and slot released) and
// the writing of checkpoint files.  The work of all active solves overlaps
// on the device.
//
// Launch shapes are tuned online per grid-shape class (each extent rounded
// up to a power of two, precision, viscous form): the first chunks of the
// first solves of a class run with the candidate shapes in turn, timed by
// events on their own streams, and once every candidate has a time the
// fastest is kept for the class.  Results do not depend on the launch
// shape, so the tuning chunks are ordinary steps of the solves: no step is
// run twice and no solve waits for the tuner.
//
// Checkpoints (the restart files of a request, rst.bin + field_info0.dat
// in the formats of field.F90:1853-1966 and singleideal.F90:2546-2659):
// every checkpoint_every steps the slot's stream copies w and the scalars
// into pinned staging after the chunk that reaches the step, and the
// worker writes the files once the copy has landed.  A request resumed
// from them (st_server_submit2 with the stored time and step counter)
// continues as a plain solve restarted from the same files does.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <tuple>
#include <unordered_map>
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
    double time0 = 0.0;
    int64_t icyc0 = 0;
    int ckpt_every = 0;         // steps between checkpoints (0: none)
    std::string ckpt_dir;
    int64_t bytes = 0;          // footprint
    double t_submit = 0, t_admit = 0, t_done = 0;
    int state = 0;              // 0 queued, 1 running, 2 done, 3 failed
    int rc = ST_OK;             // status of a failed request (ST_ERR_NAN for a NaN residual)
    std::vector<double> w_out;
    st_scalars sc{};
    std::string err;
    int checkpoints = 0;        // checkpoint files written
};

// First-fit allocator over one device block.
struct Pool {
    char *base = nullptr;
    int64_t size = 0;
    std::map<int64_t, int64_t> free_blocks;  // offset -> bytes
    static int64_t round(int64_t n) { return (n + 255) / 256 * 256; }
    int64_t alloc(int64_t n) {
        n = round(n);
        for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
            if (it->second < n) continue;
            const int64_t off = it->first, left = it->second - n;
            free_blocks.erase(it);
            if (left > 0) free_blocks[off + n] = left;
            return off;
        }
        return -1;
    }
    void release(int64_t off, int64_t n) {
        n = round(n);
        auto it = free_blocks.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_blocks.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_blocks.erase(nx);
        }
        if (it != free_blocks.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_blocks.erase(it);
            }
        }
    }
};

enum SlotState { kFree, kSetup, kRunning, kDraining };

constexpr int kTabInts = 1 << 16;  // pinned table staging: nx + ny + nz + 3 ints at most

using TuneKey = std::tuple<int, int, int, int, int>;

struct Tune {
    std::vector<std::pair<int, int>> cand;  // (visc_chunk, march_chunk)
    std::vector<float> ms;                  // per step, < 0 while unmeasured
    std::vector<char> assigned;
    bool done = false;
    std::pair<int, int> best{0, 0};
};

struct Slot {
    // for the server's life
    hipStream_t stream = nullptr;
    st::DevScalars *hs = nullptr;
    int *tabs = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};          // end of each chunk in flight
    hipEvent_t tev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // timed bounds of a tuning chunk
    double *ckpt_w = nullptr;                       // pinned checkpoint staging (5 ghosted variables)
    st::DevScalars *ckpt_sc = nullptr;
    size_t ckpt_cap = 0;                            // doubles
    hipEvent_t ckpt_ev = nullptr;
    // per request
    SlotState state = kFree;
    int rc = ST_OK;            // failure found by the worker during set-up
    std::string err;
    st_solver *s = nullptr;
    ServeReq *r = nullptr;
    int64_t off = -1;
    int launched = 0;          // steps enqueued
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int ev_steps[2] = {0, 0};  // steps completed when each event fires
    int tcand[2] = {-1, -1};   // tuning candidate of each chunk in flight
    int tsteps[2] = {0, 0};
    int nev = 0;               // events in flight (oldest first)
    int head = 0;
    TuneKey key{};
    bool ckpt_pending = false;  // a checkpoint copy in flight or being written
    int ckpt_steps = 0;
    // batched launches
    bool batchable = false;                  // a chunk of its steps records in body form
    std::vector<st::BatchLaunchFn> sig;      // the launch sequence of one chunk (grouping key)
    st::Recorder rec;                        // the chunk being batched
    int inbatch = 0;                         // batched chunks in flight that hold this solve
    int lane = -1;                           // its batch lane (scheduler only)
    double lane_cells = 0.0;                 // what it added to its lane's cells
    bool own_work = true;                    // work on its own stream since it last joined a batch
    bool shape_set = false;                  // the tuned launch shape of its class is applied
};

// One launch of a batched chunk: the body instance, its workgroup shape,
// the batched grid and the byte offsets of its tables.
struct BatchOp {
    st::BatchLaunchFn fn;
    dim3 block;
    unsigned total;
    size_t ent, own, pack;
};

// A lane runs batched chunks on its own stream, at most two in flight,
// each with its own table buffer (device, and pinned staging for the
// upload).  Every batched solve is bound to one lane while it runs.
struct Lane {
    hipStream_t stream = nullptr;
    double cells = 0.0;                  // interior cells of the solves bound to the lane
    int nsolves = 0;                     // solves bound to the lane
    char *dtab[2] = {nullptr, nullptr};
    unsigned char *htab[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<int> members[2];
    int nev = 0, head = 0;
    // The lane's last layout (tables and launches).  A chunk of K steps
    // returns a solve to the same host state but for the step counter, and
    // the same arguments, so the layout is reused while the lane's solves
    // are those it was laid out for and each is in the state the last chunk
    // left it in.
    std::vector<int> cmem;
    std::vector<int64_t> cids;           // their requests (a slot is reused by later requests)
    std::vector<st::HostMark> cafter;    // each solve's state after the last chunk
    std::vector<unsigned char> chost;
    std::vector<BatchOp> cops;
    int64_t cver = 0, bver[2] = {-1, -1};  // layout version, and the version each device buffer holds
    // the launches of a reused layout, captured once per table buffer
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int64_t gver[2] = {-1, -1};            // layout version each graph was captured from
};

struct Batcher {
    // set by the client thread before the first submission, read by the
    // scheduler: the lanes are complete before `on` is seen true
    std::atomic<bool> on{false};
    int64_t max_cells = 0;               // solves with more interior cells run on their own streams (0: no limit)
    std::vector<Lane> lanes;
    size_t cap = 0;                      // bytes of one table buffer
    hipEvent_t join = nullptr;           // orders a joining solve's own-stream work before its lane
    std::atomic<int64_t> chunks{0}, launches{0}, solves{0};
    std::atomic<int64_t> layouts{0}, ns_layout{0}, ns_launch{0};  // host work of the scheduler
    std::atomic<int64_t> graph_chunks{0};  // batched chunks replayed from a captured graph
};

constexpr size_t kBatchTabBytes = (size_t)32 << 20;  // per table buffer
constexpr int kMaxLanes = 8;

constexpr int kPlainChunk = 10;  // steps per chunk of a server without graphs (steps_per_graph 0 or 1)

enum JobKind { kSetupJob, kFinishJob, kCheckpointJob };
struct Job {
    JobKind kind;
    int slot;
};

int pow2_ceil(int n) {
    int v = 1;
    while (v < n) v <<= 1;
    return v;
}

}  // namespace

struct st_server {
    int max_slots = 1, K = 10;
    bool autotune = true;
    int device = 0;
    Pool pool;
    int64_t used = 0;
    std::thread th, worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_job;
    std::deque<ServeReq *> pending;
    std::deque<Job> jobs;
    std::map<int64_t, std::unique_ptr<ServeReq>> reqs;
    std::vector<Slot> slots;
    std::map<TuneKey, Tune> tune;
    int tuned_classes = 0;
    Batcher bt;
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

bool graphable(const st_server *sv, const ServeReq *r) {
    // K steps flip the right-hand-side buffers 3K times: K even keeps the
    // captured buffer order valid for every replay; K a multiple of the dt
    // period keeps the captured dt launches in phase
    const int idt = r->p.iter_dt_recompute > 0 ? r->p.iter_dt_recompute : 1;
    return sv->K >= 2 && sv->K % 2 == 0 && sv->K % idt == 0;
}

// Steps per chunk: steps_per_graph, or kPlainChunk without graphs.
int chunk_steps(const st_server *sv) { return sv->K > 0 ? sv->K : kPlainChunk; }

// A batched chunk must return every solve to the same right-hand-side
// buffer order and dt phase, for its layout to serve the next chunk too:
// an even number of steps, a multiple of the dt period.
bool batch_phase_ok(const st_server *sv, const ServeReq *r) {
    const int idt = r->p.iter_dt_recompute > 0 ? r->p.iter_dt_recompute : 1;
    const int n = chunk_steps(sv);
    return n % 2 == 0 && n % idt == 0;
}

int capture(st_server *sv, Slot &sl) {
    st_solver *s = sl.s;
    const st::HostMark m = st::host_mark(s);
    if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return ST_ERR_HIP;
    int rc = st_advance(s, sv->K);
    hipError_t e = hipStreamEndCapture(s->stream, &sl.graph);
    st::host_restore(s, m);
    if (rc) return rc;
    if (e != hipSuccess) return ST_ERR_HIP;
    if (hipGraphInstantiate(&sl.exec, sl.graph, nullptr, nullptr, 0) != hipSuccess) return ST_ERR_HIP;
    return ST_OK;
}

// Worker: set up an admitted solve on its slot (no allocation; the copies
// block this thread only).
int setup(st_server *sv, Slot &sl) {
    ServeReq *r = sl.r;
    st::SlotMem m{sv->pool.base + sl.off, (size_t)r->bytes, sl.stream, sl.hs, sl.tabs, r->time0, r->icyc0};
    int rc = st::solver_create_in(&r->p, m, &sl.s);
    if (rc) return rc;
    const std::vector<double> *gd = r->grid;
    if ((rc = st_set_grid(sl.s, gd[0].data(), gd[1].data(), gd[2].data(), gd[3].data(), gd[4].data(), gd[5].data(),
                          gd[6].data(), gd[7].data(), gd[8].data(), gd[9].data(), gd[10].data(), gd[11].data(),
                          gd[12].data())))
        return rc;
    if ((rc = st_copy_from_field(sl.s, r->w0.data())) || (rc = st_init_run(sl.s))) return rc;
    // whether a chunk of its steps records in body form, and its launch
    // sequence (solves whose sequences agree are batched together)
    sl.batchable = false;
    sl.sig.clear();
    const int64_t cells = (int64_t)r->p.nx * r->p.ny * r->p.nz;
    if (sv->bt.on && batch_phase_ok(sv, r) && (sv->bt.max_cells <= 0 || cells <= sv->bt.max_cells)) {
        const st::HostMark m = st::host_mark(sl.s);
        sl.rec.recs.clear();
        sl.rec.bad = false;
        sl.s->rec = &sl.rec;
        rc = st_advance(sl.s, chunk_steps(sv));
        sl.s->rec = nullptr;
        st::host_restore(sl.s, m);
        if (rc) return rc;
        sl.batchable = !sl.rec.bad && !sl.rec.recs.empty();
        for (const auto &q : sl.rec.recs) sl.sig.push_back(q.fn);
        sl.rec.recs.clear();
    }
    if (r->ckpt_every > 0) {
        const size_t need = (size_t)5 * sl.s->g.nvol;
        if (sl.ckpt_cap < need) {
            if (sl.ckpt_w) (void)hipHostFree(sl.ckpt_w);
            sl.ckpt_w = nullptr;
            sl.ckpt_cap = 0;
            if (hipHostMalloc((void **)&sl.ckpt_w, need * sizeof(double), hipHostMallocDefault) != hipSuccess)
                return st::set_error(ST_ERR_NOMEM, "checkpoint staging allocation failed");
            sl.ckpt_cap = need;
        }
    }
    return ST_OK;
}

bool write_checkpoint(const ServeReq *r, const st_solver *s, const double *w, const st::DevScalars &sc) {
    (void)mkdir(r->ckpt_dir.c_str(), 0755);
    const st::Geo &g = s->g;
    std::string path = r->ckpt_dir + "/rst.bin";
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    // interior cells of each variable, x fastest (write_field, field.F90:1853-1900)
    bool ok = true;
    for (int m = 0; m < 5 && ok; m++)
        for (int k = 1; k <= g.nz && ok; k++)
            for (int j = 1; j <= g.ny && ok; j++) {
                const double *row = w + (size_t)m * g.nvol + st::gidx(g, 1, j, k);
                ok = std::fwrite(row, sizeof(double), (size_t)g.nx, f) == (size_t)g.nx;
            }
    ok = std::fclose(f) == 0 && ok;
    if (!ok) return false;
    // field_info0.dat, the keys of write_field_info (singleideal.F90:2620-2659)
    path = r->ckpt_dir + "/field_info0.dat";
    f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "[field_info]\nicyc0 = %lld\ntime0 = %.17g\nitav = 0\nitslice_p3d = 0\n", (long long)sc.icyc,
                 sc.time);
    const char *zero_keys[] = {"time_from_last_rst",  "time_from_last_write", "time_from_last_stat",
                               "time_from_last_slice", "time_from_last_probe", "time_from_last_slice_p3d"};
    for (const char *k : zero_keys) std::fprintf(f, "%s = 0.0\n", k);
    std::fprintf(f, "istore = 0\n");
    return std::fclose(f) == 0;
}

void record_done(st_server *sv, ServeReq *r, int rc, const std::string &err) {
    // caller holds sv->mu
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

// Worker: hand a finished (or failed) solve back: scalars and field, then
// the graph, the solver, the pool block and the slot.
void finish(st_server *sv, Slot &sl) {
    ServeReq *r = sl.r;
    int rc = sl.rc;
    std::string err = sl.err;
    if (sl.s && rc == ST_OK) {
        rc = st_compute_residual(sl.s, &r->sc);
        if (rc == ST_OK || rc == ST_ERR_NAN) {
            err = rc ? st_last_error() : "";
            r->w_out.resize((size_t)5 * sl.s->g.nref);  // the reference layout
            const int rc2 = st_copy_to_field(sl.s, r->w_out.data());
            if (rc == ST_OK && rc2 != ST_OK) {
                rc = rc2;
                err = st_last_error();
            }
        } else {
            err = st_last_error();
        }
    }
    if (sl.s) (void)hipStreamSynchronize(sl.stream);
    if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
    if (sl.graph) (void)hipGraphDestroy(sl.graph);
    if (sl.s) st_solver_destroy(sl.s);
    std::lock_guard<std::mutex> lk(sv->mu);
    record_done(sv, r, rc, err);
    sv->pool.release(sl.off, r->bytes);
    sv->used -= r->bytes;
    sl.s = nullptr;
    sl.r = nullptr;
    sl.off = -1;
    sl.exec = nullptr;
    sl.graph = nullptr;
    sl.launched = 0;
    sl.nev = 0;
    sl.head = 0;
    sl.rc = ST_OK;
    sl.err.clear();
    sl.ckpt_pending = false;
    sl.ckpt_steps = 0;
    sl.batchable = false;
    sl.sig.clear();
    sl.rec.recs.clear();
    sl.inbatch = 0;
    sl.own_work = true;
    sl.shape_set = false;
    sl.state = kFree;
    sv->cv_work.notify_all();
}

void worker_loop(st_server *sv) {
    (void)hipSetDevice(sv->device);
    for (;;) {
        Job j;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_job.wait(lk, [&] { return sv->stop || !sv->jobs.empty(); });
            if (sv->jobs.empty()) return;  // stop requested and nothing left
            j = sv->jobs.front();
            sv->jobs.pop_front();
        }
        Slot &sl = sv->slots[j.slot];
        if (j.kind == kSetupJob) {
            int rc = setup(sv, sl);
            std::string err = rc ? std::string(st_last_error()) : "";
            std::lock_guard<std::mutex> lk(sv->mu);
            sl.rc = rc;
            sl.err = err;
            sl.state = kRunning;  // the scheduler launches it, or hands a failed set-up back
            sv->cv_work.notify_all();
        } else if (j.kind == kCheckpointJob) {
            const bool ok = write_checkpoint(sl.r, sl.s, sl.ckpt_w, *sl.ckpt_sc);
            std::lock_guard<std::mutex> lk(sv->mu);
            if (ok) sl.r->checkpoints++;
            else {
                sl.rc = ST_ERR_ARG;
                sl.err = "checkpoint files could not be written under " + sl.r->ckpt_dir;
            }
            sl.ckpt_pending = false;
            sv->cv_work.notify_all();
        } else {
            finish(sv, sl);
        }
    }
}

void push_job(st_server *sv, JobKind k, int slot) {
    // caller holds sv->mu
    sv->jobs.push_back(Job{k, slot});
    sv->cv_job.notify_all();
}

// Scheduler: the candidate shapes of a class.
Tune make_tune() {
    Tune t;
    t.cand = {{0, 0}, {8, 8}, {8, 16}, {16, 16}, {16, 32}, {32, 16}, {32, 32}, {8, 32}, {32, 8}};
    t.ms.assign(t.cand.size(), -1.f);
    t.assigned.assign(t.cand.size(), 0);
    return t;
}

// Scheduler: enqueue the next chunk of steps of a slot and record its
// event.  Every chunk holds at least one step.
int launch_chunk(st_server *sv, Slot &sl) {
    st_solver *s = sl.s;
    ServeReq *r = sl.r;
    const int left = r->nsteps - sl.launched;
    const int per = chunk_steps(sv);
    int n = left < per ? left : per;
    const int slot = (sl.head + sl.nev) % 2;
    int cand = -1;
    Tune *t = nullptr;
    if (sv->autotune) {
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->tune.find(sl.key);
        if (it == sv->tune.end()) it = sv->tune.emplace(sl.key, make_tune()).first;
        t = &it->second;
        if (!t->done && n == per) {
            for (size_t q = 0; q < t->cand.size(); q++)
                if (!t->assigned[q]) {
                    t->assigned[q] = 1;
                    cand = (int)q;
                    break;
                }
        }
    }
    int rc;
    if (cand >= 0) {
        if ((rc = st_set_tuning(s, t->cand[cand].first, t->cand[cand].second))) return rc;
        if (hipEventRecord(sl.tev[slot][0], s->stream) != hipSuccess) return ST_ERR_HIP;
        if ((rc = st_advance(s, n))) return rc;
        if (hipEventRecord(sl.tev[slot][1], s->stream) != hipSuccess) return ST_ERR_HIP;
    } else {
        bool use_graph = false;
        if (n == sv->K && graphable(sv, r)) {
            if (!sl.exec && (!t || t->done)) {
                if (t && (rc = st_set_tuning(s, t->best.first, t->best.second))) return rc;
                if ((rc = capture(sv, sl))) return rc;
            }
            use_graph = sl.exec != nullptr;
        }
        if (use_graph) {
            if (hipGraphLaunch(sl.exec, s->stream) != hipSuccess) return ST_ERR_HIP;
            s->icyc += n;
            s->last_stage = 2;
        } else {
            if (t && !t->done && (rc = st_set_tuning(s, 0, 0))) return rc;
            if ((rc = st_advance(s, n))) return rc;
        }
    }
    sl.launched += n;
    // a checkpoint after the chunk that reaches the period (or the last step)
    if (r->ckpt_every > 0 && !sl.ckpt_pending &&
        (sl.launched % r->ckpt_every == 0 || sl.launched == r->nsteps)) {
        if (hipMemcpyAsync(sl.ckpt_w, s->w, sizeof(double) * 5 * s->g.nvol, hipMemcpyDeviceToHost, s->stream) !=
                hipSuccess ||
            hipMemcpyAsync(sl.ckpt_sc, s->ds, sizeof(st::DevScalars), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
            hipEventRecord(sl.ckpt_ev, s->stream) != hipSuccess)
            return ST_ERR_HIP;
        std::lock_guard<std::mutex> lk(sv->mu);
        sl.ckpt_pending = true;
        sl.ckpt_steps = sl.launched;
    }
    if (hipEventRecord(sl.ev[slot], s->stream) != hipSuccess) return ST_ERR_HIP;
    sl.ev_steps[slot] = sl.launched;
    sl.tcand[slot] = cand;
    sl.tsteps[slot] = n;
    sl.nev++;
    return ST_OK;
}

// Scheduler: retire the chunks whose events have fired; a tuning chunk's
// time goes to its class.
void retire(st_server *sv, Slot &sl, bool *progress) {
    while (sl.nev > 0 && hipEventQuery(sl.ev[sl.head]) == hipSuccess) {
        const int h = sl.head;
        if (sl.tcand[h] >= 0) {
            float ms = 0.f;
            const bool ok = hipEventElapsedTime(&ms, sl.tev[h][0], sl.tev[h][1]) == hipSuccess;
            std::lock_guard<std::mutex> lk(sv->mu);
            Tune &t = sv->tune[sl.key];
            if (ok) t.ms[sl.tcand[h]] = ms / (float)sl.tsteps[h];
            else t.assigned[sl.tcand[h]] = 0;
            bool all = true;
            for (float v : t.ms) all &= v >= 0.f;
            if (all && !t.done) {
                size_t b = 0;
                for (size_t q = 1; q < t.ms.size(); q++)
                    if (t.ms[q] < t.ms[b]) b = q;
                t.best = t.cand[b];
                t.done = true;
                sv->tuned_classes++;
            }
            sl.tcand[h] = -1;
        }
        sl.head = (sl.head + 1) % 2;
        sl.nev--;
        *progress = true;
    }
}

// Scheduler: a slot that stops early gives back the tuning candidates of
// its chunks in flight, so that another solve of the class measures them.
void drop_candidates(st_server *sv, Slot &sl) {
    std::lock_guard<std::mutex> lk(sv->mu);
    for (int q = 0; q < 2; q++)
        if (sl.tcand[q] >= 0) {
            Tune &t = sv->tune[sl.key];
            if (t.ms[sl.tcand[q]] < 0.f) t.assigned[sl.tcand[q]] = 0;
            sl.tcand[q] = -1;
        }
}

// Scheduler: whether the launch shape of a solve's class is settled.
bool class_done(st_server *sv, const Slot &sl, Tune *best = nullptr) {
    if (!sv->autotune) return true;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->tune.find(sl.key);
    if (it == sv->tune.end() || !it->second.done) return false;
    if (best) *best = it->second;
    return true;
}

// Scheduler: a solve whose next chunk belongs in a batch (it shares its
// lane with another solve, its class is tuned and a whole chunk is left);
// it launches nothing on its own stream.  A solve alone in its lane runs on
// its own stream, as without batching.
bool batch_bound(st_server *sv, const Slot &sl) {
    return sv->bt.on && sl.batchable && sl.lane >= 0 && sv->bt.lanes[sl.lane].nsolves >= 2 &&
           sl.r->nsteps - sl.launched >= chunk_steps(sv) && class_done(sv, sl);
}

// Scheduler: retire the batched chunks whose events have fired.
void retire_batch(st_server *sv, Lane &ln, bool *progress) {
    while (ln.nev > 0 && hipEventQuery(ln.ev[ln.head]) == hipSuccess) {
        for (int q : ln.members[ln.head]) sv->slots[q].inbatch--;
        ln.members[ln.head].clear();
        ln.head = (ln.head + 1) % 2;
        ln.nev--;
        *progress = true;
    }
}

// Lay out the tables of one batched chunk of the solves `mem` (their
// recorders hold the same launch sequence) in bt.host: per launch, the
// solves' entries, the owner table (one int per 8 workgroups) and the
// packed arguments, each 256-byte aligned.  Identical tables (a launch
// repeated with the same arguments, entries shared by launches of the same
// grids) are stored once.  false if the layout exceeds a table buffer.
bool lay_out(st_server *sv, const std::vector<int> &mem, std::vector<unsigned char> &host,
             std::vector<BatchOp> &ops) {
    host.clear();
    ops.clear();
    std::unordered_map<std::string, size_t> seen;
    auto put = [&](const void *p, size_t n) -> size_t {
        std::string key((const char *)p, n);
        auto it = seen.find(key);
        if (it != seen.end()) return it->second;
        const size_t off = (host.size() + 255) / 256 * 256;
        host.resize(off + n);
        std::memcpy(host.data() + off, p, n);
        seen.emplace(std::move(key), off);
        return off;
    };
    const size_t nops = sv->slots[mem[0]].rec.recs.size();
    std::vector<st::BatchEntry> ent(mem.size());
    std::vector<int> own;
    std::vector<unsigned char> packs;
    for (size_t i = 0; i < nops; i++) {
        const st::BatchRec &r0 = sv->slots[mem[0]].rec.recs[i];
        const size_t psz = r0.pack.size();
        unsigned total = 0;
        own.clear();
        packs.resize(psz * mem.size());
        for (size_t q = 0; q < mem.size(); q++) {
            const st::BatchRec &r = sv->slots[mem[q]].rec.recs[i];
            const unsigned nb = r.grid.x * r.grid.y * r.grid.z;
            const unsigned nbp = (nb + st::kBatchGroup - 1) / st::kBatchGroup * st::kBatchGroup;
            ent[q] = st::BatchEntry{(int)total, (int)nb, (int)r.grid.x, (int)r.grid.y, (int)r.grid.z, {0, 0, 0}};
            own.insert(own.end(), nbp / st::kBatchGroup, (int)q);
            std::memcpy(packs.data() + q * psz, r.pack.data(), psz);
            total += nbp;
        }
        BatchOp op;
        op.fn = r0.fn;
        op.block = r0.block;
        op.total = total;
        op.ent = put(ent.data(), sizeof(st::BatchEntry) * mem.size());
        op.own = put(own.data(), sizeof(int) * own.size());
        op.pack = put(packs.data(), packs.size());
        ops.push_back(op);
        if (host.size() > sv->bt.cap) return false;
    }
    return true;
}

void batch_error(st_server *sv, Lane &ln, const std::vector<int> &mem, const char *what) {
    // launches already enqueued still use the solves' memory
    (void)hipStreamSynchronize(ln.stream);
    std::lock_guard<std::mutex> lk(sv->mu);
    for (int q : mem) {
        sv->slots[q].rc = ST_ERR_HIP;
        sv->slots[q].err = what;
    }
}

// Scheduler: record one chunk of each solve of `mem` and lay the lane's
// tables out anew.  A solve whose recording does not follow its launch
// sequence leaves the batch path; solves that do not fit in a table buffer
// wait for a later chunk.  The layout is kept for reuse when every solve
// came back to its state but for the step counter.  false if no solve is
// left.
bool record_chunk(st_server *sv, Lane &ln, std::vector<int> &mem) {
    std::vector<st::HostMark> marks;
    std::vector<int> rec_ok;
    for (int q : mem) {
        Slot &sl = sv->slots[q];
        if (!sl.shape_set) {
            Tune t;
            if (sv->autotune && class_done(sv, sl, &t)) (void)st_set_tuning(sl.s, t.best.first, t.best.second);
            sl.shape_set = true;
        }
        const st::HostMark m = st::host_mark(sl.s);
        sl.rec.recs.clear();
        sl.rec.bad = false;
        sl.s->rec = &sl.rec;
        const int rc = st_advance(sl.s, chunk_steps(sv));
        sl.s->rec = nullptr;
        bool same = rc == ST_OK && !sl.rec.bad && sl.rec.recs.size() == sl.sig.size();
        for (size_t i = 0; same && i < sl.sig.size(); i++) same = sl.rec.recs[i].fn == sl.sig[i];
        if (!same) {
            // not batchable after all: it runs on its own stream
            st::host_restore(sl.s, m);
            sl.batchable = false;
            sl.rec.recs.clear();
            continue;
        }
        marks.push_back(m);
        rec_ok.push_back(q);
    }
    mem.swap(rec_ok);
    while (!mem.empty() && !lay_out(sv, mem, ln.chost, ln.cops)) {
        // too large for a table buffer: the last solve waits for the next chunk
        Slot &sl = sv->slots[mem.back()];
        st::host_restore(sl.s, marks.back());
        sl.rec.recs.clear();
        if (mem.size() == 1) sl.batchable = false;
        mem.pop_back();
        marks.pop_back();
    }
    ln.cver++;
    ln.cmem.clear();
    ln.cids.clear();
    ln.cafter.clear();
    if (mem.empty()) return false;
    bool periodic = true;
    std::vector<st::HostMark> after;
    for (size_t i = 0; i < mem.size(); i++) {
        Slot &sl = sv->slots[mem[i]];
        after.push_back(st::host_mark(sl.s));
        periodic &= st::same_state(marks[i], after.back(), false);
        sl.rec.recs.clear();
    }
    if (periodic) {
        ln.cmem = mem;
        ln.cafter = after;
        for (int q : mem) ln.cids.push_back(sv->slots[q].r->id);
    }
    sv->bt.layouts++;
    return true;
}

// Scheduler: bind a solve that is about to join a batch to the lane that
// holds the fewest cells (the lanes keep their own counts: the scheduler
// never reads another slot's solver, which the worker may be creating or
// destroying).
void bind_lane(st_server *sv, Slot &sl) {
    std::vector<Lane> &lanes = sv->bt.lanes;
    int best = 0;
    for (int l = 1; l < (int)lanes.size(); l++)
        if (lanes[l].cells < lanes[best].cells) best = l;
    sl.lane = best;
    sl.lane_cells = (double)sl.s->g.nint;
    lanes[best].cells += sl.lane_cells;
    lanes[best].nsolves++;
}

// Scheduler: a solve that leaves (hand-back) gives its cells back to its lane.
void unbind_lane(st_server *sv, Slot &sl) {
    if (sl.lane < 0) return;
    sv->bt.lanes[sl.lane].cells -= sl.lane_cells;
    sv->bt.lanes[sl.lane].nsolves--;
    sl.lane = -1;
    sl.lane_cells = 0.0;
}

// Scheduler: the next batched chunk of lane `li`.  Its solves are the
// ready ones of the lane (set up, no chunk in flight on their own stream,
// no checkpoint copy pending) whose launch sequence is that of the least
// advanced of them.  Each is recorded for K steps (which advances its host state as
// enqueuing does), the tables are laid out and uploaded (skipped when the
// buffer already holds them), and one launch per recorded launch follows
// on the lane's stream.
void launch_batch(st_server *sv, int li, bool *progress) {
    Batcher &bt = sv->bt;
    Lane &ln = bt.lanes[li];
    if (ln.nev >= 2) return;
    std::vector<int> ready;
    int lead = -1;
    for (int q = 0; q < (int)sv->slots.size(); q++) {
        Slot &sl = sv->slots[q];
        bool ok;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            ok = sl.state == kRunning && sl.s && sl.rc == ST_OK && !sl.ckpt_pending;
        }
        if (!ok || sl.nev > 0 || sl.lane != li || !batch_bound(sv, sl)) continue;
        ready.push_back(q);
        // the group of the least advanced ready solve (the oldest among
        // equals): groups of different launch sequences in one lane take
        // turns
        const Slot &ld = sv->slots[lead < 0 ? q : lead];
        if (lead < 0 || sl.launched < ld.launched || (sl.launched == ld.launched && sl.r->id < ld.r->id)) lead = q;
    }
    if (lead < 0) return;
    std::vector<int> mem;
    for (int q : ready)
        if (sv->slots[q].sig == sv->slots[lead].sig && (int)mem.size() < st::kBatchMax) mem.push_back(q);
    const auto t0 = Clock::now();
    bool reuse = mem == ln.cmem;
    for (size_t i = 0; reuse && i < mem.size(); i++)
        reuse = sv->slots[mem[i]].r->id == ln.cids[i] &&
                st::same_state(st::host_mark(sv->slots[mem[i]].s), ln.cafter[i], true);
    if (reuse) {
        // the same chunk again: advance each solve's host state as its
        // recording would
        for (size_t i = 0; i < mem.size(); i++) {
            ln.cafter[i].icyc += chunk_steps(sv);
            st::host_restore(sv->slots[mem[i]].s, ln.cafter[i]);
        }
    } else if (!record_chunk(sv, ln, mem)) {
        return;
    }
    const auto t1 = Clock::now();
    if (mem.empty()) return;
    const int b = (ln.head + ln.nev) % 2;
    for (int q : mem) {
        Slot &sl = sv->slots[q];
        if (!sl.own_work) continue;
        if (hipEventRecord(bt.join, sl.stream) != hipSuccess || hipStreamWaitEvent(ln.stream, bt.join, 0) != hipSuccess)
            return batch_error(sv, ln, mem, "batched chunk: stream ordering failed");
        sl.own_work = false;
    }
    if (ln.bver[b] != ln.cver) {
        std::memcpy(ln.htab[b], ln.chost.data(), ln.chost.size());
        if (hipMemcpyAsync(ln.dtab[b], ln.htab[b], ln.chost.size(), hipMemcpyHostToDevice, ln.stream) != hipSuccess)
            return batch_error(sv, ln, mem, "batched chunk: table upload failed");
        ln.bver[b] = ln.cver;
    }
    char *d = ln.dtab[b];
    auto enqueue = [&]() -> bool {
        for (const BatchOp &op : ln.cops)
            if (op.fn(op.total, op.block, ln.stream, d + op.ent, d + op.pack, (const int *)(d + op.own)) != hipSuccess)
                return false;
        return true;
    };
    // with graphs on, a layout used a second time is captured (once per
    // buffer) and replayed from then on
    const bool graphs = sv->K >= 2;
    if (reuse && graphs && ln.gver[b] != ln.cver) {
        if (ln.exec[b]) (void)hipGraphExecDestroy(ln.exec[b]);
        if (ln.graph[b]) (void)hipGraphDestroy(ln.graph[b]);
        ln.exec[b] = nullptr;
        ln.graph[b] = nullptr;
        ln.gver[b] = -1;
        if (hipStreamBeginCapture(ln.stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return batch_error(sv, ln, mem, "batched chunk: capture failed");
        const bool ok = enqueue();
        const bool ended = hipStreamEndCapture(ln.stream, &ln.graph[b]) == hipSuccess;
        if (!ok || !ended || hipGraphInstantiate(&ln.exec[b], ln.graph[b], nullptr, nullptr, 0) != hipSuccess)
            return batch_error(sv, ln, mem, "batched chunk: capture failed");
        ln.gver[b] = ln.cver;
    }
    if (reuse && graphs) {
        if (hipGraphLaunch(ln.exec[b], ln.stream) != hipSuccess) return batch_error(sv, ln, mem, "batched graph launch failed");
        bt.graph_chunks++;
    } else if (!enqueue()) {
        return batch_error(sv, ln, mem, "batched launch failed");
    }
    if (hipEventRecord(ln.ev[b], ln.stream) != hipSuccess)
        return batch_error(sv, ln, mem, "batched chunk: event failed");
    ln.members[b] = mem;
    ln.nev++;
    bt.chunks++;
    bt.launches += (int64_t)ln.cops.size();
    bt.solves += (int64_t)mem.size();
    bt.ns_layout += std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
    bt.ns_launch += std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - t1).count();
    for (int q : mem) {
        Slot &sl = sv->slots[q];
        sl.rec.recs.clear();
        sl.launched += chunk_steps(sv);
        sl.inbatch++;
        ServeReq *r = sl.r;
        // a checkpoint after the chunk that reaches the period (or the last
        // step): copied on the solve's own stream once the batch has passed
        if (r->ckpt_every > 0 && (sl.launched % r->ckpt_every == 0 || sl.launched == r->nsteps)) {
            st_solver *s = sl.s;
            if (hipStreamWaitEvent(sl.stream, ln.ev[b], 0) != hipSuccess ||
                hipMemcpyAsync(sl.ckpt_w, s->w, sizeof(double) * 5 * s->g.nvol, hipMemcpyDeviceToHost, sl.stream) !=
                    hipSuccess ||
                hipMemcpyAsync(sl.ckpt_sc, s->ds, sizeof(st::DevScalars), hipMemcpyDeviceToHost, sl.stream) !=
                    hipSuccess ||
                hipEventRecord(sl.ckpt_ev, sl.stream) != hipSuccess)
                return batch_error(sv, ln, {q}, "checkpoint copy failed");
            sl.own_work = true;
            std::lock_guard<std::mutex> lk(sv->mu);
            sl.ckpt_pending = true;
            sl.ckpt_steps = sl.launched;
        }
    }
    *progress = true;
}

void serve_loop(st_server *sv) {
    (void)hipSetDevice(sv->device);
    for (;;) {
        bool progress = false;
        {
            // admission, in arrival order, while a slot is free and the pool
            // holds the footprint
            std::unique_lock<std::mutex> lk(sv->mu);
            if (sv->stop) return;
            while (!sv->pending.empty()) {
                int fs = -1;
                for (int q = 0; q < (int)sv->slots.size(); q++)
                    if (sv->slots[q].state == kFree) {
                        fs = q;
                        break;
                    }
                if (fs < 0) break;
                ServeReq *r = sv->pending.front();
                const int64_t off = sv->pool.alloc(r->bytes);
                if (off < 0) break;  // wait for memory
                sv->pending.pop_front();
                Slot &sl = sv->slots[fs];
                sl.state = kSetup;
                sl.r = r;
                sl.off = off;
                sl.rc = ST_OK;
                sl.key = TuneKey{pow2_ceil(r->p.nx), pow2_ceil(r->p.ny), pow2_ceil(r->p.nz), r->p.precision,
                                 r->p.conservative_viscous};
                sl.tcand[0] = sl.tcand[1] = -1;
                sv->used += r->bytes;
                r->state = 1;
                r->t_admit = sv->now();
                push_job(sv, kSetupJob, fs);
                progress = true;
            }
        }
        bool active = false;
        for (int q = 0; q < (int)sv->slots.size(); q++) {
            Slot &sl = sv->slots[q];
            SlotState st;
            bool ckpt_pending;
            {
                std::lock_guard<std::mutex> lk(sv->mu);
                st = sl.state;
                ckpt_pending = sl.ckpt_pending;
            }
            if (st != kRunning) {
                active |= st != kFree;
                continue;
            }
            active = true;
            // a batchable solve is bound to a lane once it runs
            if (sv->bt.on && sl.batchable && sl.lane < 0 && sl.s) bind_lane(sv, sl);
            retire(sv, sl, &progress);
            // a landed checkpoint copy goes to the worker
            if (ckpt_pending && sl.s && hipEventQuery(sl.ckpt_ev) == hipSuccess) {
                std::lock_guard<std::mutex> lk(sv->mu);
                if (sl.ckpt_steps > 0) {
                    sl.ckpt_steps = 0;
                    push_job(sv, kCheckpointJob, q);
                    progress = true;
                }
            }
            int rc;
            {
                std::lock_guard<std::mutex> lk(sv->mu);
                rc = sl.rc;  // a failed set-up, or a checkpoint that could not be written
            }
            while (rc == ST_OK && sl.nev < 2 && sl.launched < sl.r->nsteps && !ckpt_pending && sl.inbatch == 0 &&
                   !batch_bound(sv, sl)) {
                rc = launch_chunk(sv, sl);
                sl.own_work = true;
                if (rc) {
                    std::lock_guard<std::mutex> lk(sv->mu);
                    sl.err = st_last_error();
                }
                {
                    std::lock_guard<std::mutex> lk(sv->mu);
                    ckpt_pending = sl.ckpt_pending;
                }
                progress = true;
            }
            bool ckpt_busy;
            {
                std::lock_guard<std::mutex> lk(sv->mu);
                ckpt_busy = sl.ckpt_pending;
            }
            // a solve held by a batched chunk in flight stays until the chunk has passed
            if (sl.inbatch == 0 && (rc != ST_OK || (sl.nev == 0 && sl.launched >= sl.r->nsteps && !ckpt_busy))) {
                if (rc != ST_OK) drop_candidates(sv, sl);
                unbind_lane(sv, sl);
                std::lock_guard<std::mutex> lk(sv->mu);
                sl.rc = rc;
                sl.state = kDraining;
                push_job(sv, kFinishJob, q);
                progress = true;
            }
        }
        if (sv->bt.on) {
            for (int l = 0; l < (int)sv->bt.lanes.size(); l++) {
                Lane &ln = sv->bt.lanes[l];
                retire_batch(sv, ln, &progress);
                launch_batch(sv, l, &progress);
                active |= ln.nev > 0;
            }
        }
        if (!progress) {
            std::unique_lock<std::mutex> lk(sv->mu);
            if (sv->stop) return;
            if (active) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            } else {
                sv->cv_work.wait_for(lk, std::chrono::milliseconds(5));
            }
        }
    }
}

// The lanes' streams, events, graphs and table buffers.
void free_lanes(Batcher &bt) {
    for (Lane &ln : bt.lanes) {
        if (ln.stream) (void)hipStreamSynchronize(ln.stream);
        for (int b = 0; b < 2; b++) {
            if (ln.exec[b]) (void)hipGraphExecDestroy(ln.exec[b]);
            if (ln.graph[b]) (void)hipGraphDestroy(ln.graph[b]);
            if (ln.ev[b]) (void)hipEventDestroy(ln.ev[b]);
            if (ln.dtab[b]) (void)hipFree(ln.dtab[b]);
            if (ln.htab[b]) (void)hipHostFree(ln.htab[b]);
        }
        if (ln.stream) (void)hipStreamDestroy(ln.stream);
    }
    bt.lanes.clear();
    if (bt.join) (void)hipEventDestroy(bt.join);
    bt.join = nullptr;
}

void destroy_slots(st_server *sv) {
    // batched chunks still in flight use the solves' pool memory: the lanes
    // drain first
    sv->bt.on = false;
    free_lanes(sv->bt);
    for (auto &sl : sv->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
        if (sl.graph) (void)hipGraphDestroy(sl.graph);
        if (sl.s) st_solver_destroy(sl.s);
        for (auto &e : sl.ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &pr : sl.tev)
            for (auto &e : pr)
                if (e) (void)hipEventDestroy(e);
        if (sl.ckpt_ev) (void)hipEventDestroy(sl.ckpt_ev);
        if (sl.ckpt_w) (void)hipHostFree(sl.ckpt_w);
        if (sl.ckpt_sc) (void)hipHostFree(sl.ckpt_sc);
        if (sl.hs) (void)hipHostFree(sl.hs);
        if (sl.tabs) (void)hipHostFree(sl.tabs);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        sl = Slot();
    }
    if (sv->pool.base) (void)hipFree(sv->pool.base);
    sv->pool.base = nullptr;
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
    sv->max_slots = max_slots;
    sv->K = steps_per_graph;
    sv->autotune = autotune != 0;
    sv->slots.resize(max_slots);
    // the whole capacity, once: every admitted solve is carved out of it
    sv->pool.size = Pool::round(hbm_capacity_bytes);
    hipError_t e = hipMalloc((void **)&sv->pool.base, (size_t)sv->pool.size);
    if (e == hipSuccess) sv->pool.free_blocks[0] = sv->pool.size;
    for (auto &sl : sv->slots) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipHostMalloc((void **)&sl.hs, sizeof(st::DevScalars), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void **)&sl.tabs, sizeof(int) * kTabInts, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void **)&sl.ckpt_sc, sizeof(st::DevScalars), hipHostMallocDefault);
        for (auto &ev : sl.ev)
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        for (auto &pr : sl.tev)
            for (auto &ev : pr)
                if (e == hipSuccess) e = hipEventCreate(&ev);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.ckpt_ev, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        destroy_slots(sv);
        delete sv;
        return sfail(ST_ERR_NOMEM, std::string("server memory: ") + hipGetErrorString(e));
    }
    sv->t0 = Clock::now();
    sv->th = std::thread(serve_loop, sv);
    sv->worker = std::thread(worker_loop, sv);
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
    // the worker finishes the jobs already queued, then stops
    sv->cv_job.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    destroy_slots(sv);
    delete sv;
    return ST_OK;
}

int st_server_submit2(st_server *sv, const st_params *p, const double *x, const double *y, const double *z,
                      const double *yn, const double *dcsidx, const double *dcsidxs, const double *dcsidx2,
                      const double *detady, const double *detadys, const double *detady2, const double *dzitdz,
                      const double *dzitdzs, const double *dzitdz2, const double *w0, int32_t nsteps, double time0,
                      int64_t icyc0, int32_t checkpoint_every, const char *checkpoint_dir, int64_t *ticket) {
    if (!sv || !p || !w0 || !ticket || nsteps < 0 || checkpoint_every < 0 || icyc0 < 0)
        return sfail(ST_ERR_ARG, "bad submit arguments");
    if (checkpoint_every > 0 && (!checkpoint_dir || !*checkpoint_dir))
        return sfail(ST_ERR_ARG, "a checkpoint period needs a checkpoint directory");
    // checkpoints fall on chunk boundaries: K steps, or kPlainChunk without graphs
    const int per = chunk_steps(sv);
    if (checkpoint_every > 0 && checkpoint_every % per != 0)
        return sfail(ST_ERR_ARG, "the checkpoint period must be a multiple of the steps per chunk");
    int rc = st::validate_params(p);
    if (rc) return sfail(rc, st_last_error());
    if (p->ng != 3) return sfail(ST_ERR_ARG, "the server takes solves with three ghost layers");
    if (p->nx + p->ny + p->nz + 3 > kTabInts) return sfail(ST_ERR_ARG, "grid too large for the server");
    if (p->bctags[0] == 10 || p->bctags[0] == 9)
        return sfail(ST_ERR_ARG, "the server takes no inflow planes (recycling or laminar inflow)");
    const int64_t need = st_solver_footprint(p);
    if (need <= 0 || Pool::round(need) > sv->pool.size)
        return sfail(ST_ERR_ARG, "request larger than the server capacity");
    auto r = std::make_unique<ServeReq>();
    r->p = *p;
    r->nsteps = nsteps;
    r->time0 = time0;
    r->icyc0 = icyc0;
    r->ckpt_every = checkpoint_every;
    if (checkpoint_every > 0) r->ckpt_dir = checkpoint_dir;
    r->bytes = need;
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

int st_server_submit(st_server *sv, const st_params *p, const double *x, const double *y, const double *z,
                     const double *yn, const double *dcsidx, const double *dcsidxs, const double *dcsidx2,
                     const double *detady, const double *detadys, const double *detady2, const double *dzitdz,
                     const double *dzitdzs, const double *dzitdz2, const double *w0, int32_t nsteps, int64_t *ticket) {
    return st_server_submit2(sv, p, x, y, z, yn, dcsidx, dcsidxs, dcsidx2, detady, detadys, detady2, dzitdz, dzitdzs,
                             dzitdz2, w0, nsteps, 0.0, 0, 0, nullptr, ticket);
}

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

int st_server_set_batching(st_server *sv, int32_t lanes, int64_t max_cells) {
    if (!sv) return sfail(ST_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->next_id != 1) return sfail(ST_ERR_STATE, "batching is set before the first submission");
    if (chunk_steps(sv) % 2 != 0)
        return sfail(ST_ERR_ARG, "batched chunks need an even number of steps (steps_per_graph, or 10 without graphs)");
    Batcher &bt = sv->bt;
    if (lanes < 0 || lanes > kMaxLanes || max_cells < 0) return sfail(ST_ERR_ARG, "batch lanes out of range");
    bt.max_cells = max_cells;
    if (!bt.lanes.empty()) {
        if (lanes != 0 && lanes != (int)bt.lanes.size())
            return sfail(ST_ERR_STATE, "the number of batch lanes is set once");
        bt.on = lanes != 0;
        return ST_OK;
    }
    if (lanes == 0) return ST_OK;
    // each lane's stream, its events and two table buffers, once
    bt.lanes.resize(lanes);
    hipError_t e = hipEventCreateWithFlags(&bt.join, hipEventDisableTiming);
    for (Lane &ln : bt.lanes) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking);
        for (int b = 0; b < 2; b++) {
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ln.ev[b], hipEventDisableTiming);
            if (e == hipSuccess) e = hipMalloc((void **)&ln.dtab[b], kBatchTabBytes);
            if (e == hipSuccess) e = hipHostMalloc((void **)&ln.htab[b], kBatchTabBytes, hipHostMallocDefault);
        }
    }
    if (e != hipSuccess) {
        free_lanes(bt);  // a later call starts afresh
        return sfail(ST_ERR_NOMEM, std::string("batch tables: ") + hipGetErrorString(e));
    }
    bt.cap = kBatchTabBytes;
    bt.on = true;
    return ST_OK;
}

int st_server_batch_stats(st_server *sv, int64_t out[7]) {
    if (!sv || !out) return sfail(ST_ERR_ARG, "null argument");
    out[0] = sv->bt.chunks;
    out[1] = sv->bt.launches;
    out[2] = sv->bt.solves;
    out[3] = sv->bt.layouts;
    out[4] = sv->bt.ns_layout / 1000;
    out[5] = sv->bt.ns_launch / 1000;
    out[6] = sv->bt.graph_chunks;
    return ST_OK;
}

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
    out[5] = sv->tuned_classes;
    return ST_OK;
}

}  // extern "C"
