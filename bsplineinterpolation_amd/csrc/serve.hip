// serve.hip -- the serving loop (intp_server_* / intp_serve_*): builds and
// evaluations submitted by any number of client threads, admitted by a
// dispatcher thread against an HBM budget, run on several HIP streams,
// completed through HIP events.
//
// Dispatcher turn (continuous batching: a turn never waits for the work of
// an earlier turn):
//   1. take every queued request, oldest first;
//      eval  -> launched on the next stream, after a stream wait on its
//               spline's build event; waits while its build is queued;
//               fails with INTP_E_EVICTED if the spline is gone;
//      build -> admitted if its control points fit the budget beside the
//               resident splines, after evicting least-recently-used idle
//               splines if needed; otherwise it stays queued for a later
//               turn (evaluations of other splines are not held up);
//   2. poll the events of launched requests, oldest first on each stream
//      (events on one stream complete in order, so polling stops at a
//      stream's first unfinished request); a finished request gets its
//      completion time, unpins its spline, and wakes waiting clients;
//   3. sleep until a submission, or briefly while work is in flight.
// Builds and evaluations go through intp_build / intp_build_refined /
// intp_eval, so their results equal those calls' on the same inputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/intp_amd.h"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace intp_amd;

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kMaxServeStreams = 16;

struct Request {
    int kind = INTP_REQ_BUILD;
    int state = INTP_REQ_QUEUED;
    int status = INTP_OK;
    double t_submit = 0, t_done = 0;  // server clock, s
    double bytes = 0;                 // algorithmic bytes
    int64_t spline = -1;              // build: its own id; eval: the target
    // build
    intp_grid_desc desc{};
    std::vector<std::vector<double>> coords_copy;  // non-uniform node coordinates
    const double* values = nullptr;
    int precision = INTP_PREC_F64;
    int sweeps = 0;
    int64_t footprint = 0;
    bool waited = false;  // held back for memory at least once
    // eval
    int64_t npts = 0;
    const double* coords = nullptr;
    int64_t cs_dim = 0, cs_pt = 0;
    std::vector<int32_t> orders;
    int32_t nout = 0;
    double* out = nullptr;
    int64_t os_out = 0, os_pt = 0;
    uint32_t flags = 0;
    // running
    hipEvent_t done = nullptr;
};

struct Resident {
    intp_spline h = nullptr;
    int64_t bytes = 0;
    int pins = 0;         // launched, unfinished evaluations
    bool ready = false;   // build finished
    bool released = false;
    hipEvent_t built = nullptr;
    std::list<int64_t>::iterator lru;
};

// Table bytes a build keeps resident (refined: FP32 control points and
// the gradient table), plus what a refined build needs while it runs
// (FP32 correction and two FP64 residual buffers).
int64_t resident_bytes_of(const intp_grid_desc& d, int precision) {
    int64_t n = 1;
    for (int i = 0; i < d.dim; ++i) n *= d.n[i];
    return n * (precision == INTP_PREC_F32_REFINED ? 4 + gradient_table_bytes(d.dim) : 8);
}

int64_t build_footprint(const intp_grid_desc& d, int precision) {
    int64_t n = 1;
    for (int i = 0; i < d.dim; ++i) n *= d.n[i];
    return resident_bytes_of(d, precision) + (precision == INTP_PREC_F32_REFINED ? n * (4 + 16) : 0);
}

// Algorithmic HBM bytes of a build (2 D N s per solve; refined: S+1 FP32
// solves, S FP64 residuals with the subtraction of f, the FP32 updates
// added on store, and with a gradient table (up to 3 dimensions) the
// table from c and the last correction, and the last update).
double build_bytes(const intp_grid_desc& d, int precision, int sweeps) {
    double n = 1;
    for (int i = 0; i < d.dim; ++i) n *= static_cast<double>(d.n[i]);
    const double D = d.dim;
    if (precision != INTP_PREC_F32_REFINED) return 2 * D * n * 8;
    const double S = sweeps;
    const double base = (S + 1) * 2 * D * n * 4 + S * (2 * D * n * 8 + 8 * n);
    const int cb = gradient_table_bytes(d.dim);
    if (cb == 0) return base + S * n * 4;
    return base + (S > 0 ? S - 1 : 0) * n * 4 + n * ((S > 0 ? 8 : 4) + cb) + (S > 0 ? 12 * n : 0);
}

}  // namespace

struct intp_server_s {
    int device = 0;
    intp_ctx ctx = nullptr;
    int64_t budget = 0;
    int max_inflight = 256;
    std::vector<hipStream_t> streams;
    size_t next_stream = 0;
    Clock::time_point t0 = Clock::now();

    // Guarded by mu: the tables below and every Request's state, status,
    // t_done and done event.  A Request's arguments are written before it
    // is queued and only read afterwards, so the dispatcher reads them
    // without the lock; the lock is not held across HIP launches, so
    // submissions never wait for one.
    std::mutex mu;
    std::condition_variable wake;  // dispatcher: new work
    std::condition_variable done;  // clients: a request finished
    bool stop = false;
    int64_t next_id = 1;
    std::deque<int64_t> queue;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    std::unordered_map<int64_t, Resident> splines;
    std::list<int64_t> lru;  // front: least recently used
    int64_t resident = 0;
    // launched requests per stream, in launch order (= completion order)
    std::vector<std::deque<int64_t>> inflight;
    int64_t inflight_n = 0;
    int64_t unfinished = 0;  // submitted, not yet done or failed

    // statistics of the current period (mu)
    std::vector<double> lat[2];
    double bytes[2] = {0, 0};
    int64_t failed[2] = {0, 0};
    int64_t evictions = 0, admission_waits = 0;
    double first_submit = -1, last_done = 0;

    // dispatcher thread only
    std::vector<hipEvent_t> spare_events;
    std::vector<intp_spline> to_release;

    std::thread dispatcher;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    void run();
    void finish(Request& r, int status);
    int dispatch(int64_t id, std::unique_lock<std::mutex>& lk);
    void poll_inflight();
    void drop(std::unordered_map<int64_t, Resident>::iterator it);
    hipEvent_t take_event();
};

hipEvent_t intp_server_s::take_event() {
    if (!spare_events.empty()) {
        hipEvent_t e = spare_events.back();
        spare_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

// Marks a request finished (dispatcher, holding mu).
void intp_server_s::finish(Request& r, int status) {
    r.status = status;
    r.state = status == INTP_OK ? INTP_REQ_DONE : INTP_REQ_FAILED;
    r.t_done = now();
    if (status == INTP_OK) {
        lat[r.kind].push_back((r.t_done - r.t_submit) * 1e3);
        bytes[r.kind] += r.bytes;
        last_done = std::max(last_done, r.t_done);
    } else {
        ++failed[r.kind];
    }
    if (r.done) {
        spare_events.push_back(r.done);
        r.done = nullptr;
    }
    --unfinished;
    done.notify_all();
}

// Removes a resident spline; its handle is released after the lock drops
// (dispatcher, holding mu).
void intp_server_s::drop(std::unordered_map<int64_t, Resident>::iterator it) {
    Resident& s = it->second;
    resident -= s.bytes;
    lru.erase(s.lru);
    if (s.built) spare_events.push_back(s.built);
    if (s.h) to_release.push_back(s.h);
    splines.erase(it);
}

// Admits and launches one request.  Returns 1 if it ran or failed, 0 if it
// must wait (a build for memory, an evaluation for its queued build), -1
// if the in-flight limit is reached.
// Called and returns with lk held; drops it around the HIP launches.
int intp_server_s::dispatch(int64_t id, std::unique_lock<std::mutex>& lk) {
    if (inflight_n >= max_inflight) return -1;
    Request& r = *reqs[id];
    const size_t si = next_stream++ % streams.size();
    hipStream_t st = streams[si];
    if (r.kind == INTP_REQ_BUILD) {
        if (r.footprint > budget) {
            finish(r, fail(INTP_E_NOMEM, "build larger than the server's HBM budget"));
            return 1;
        }
        // evict least-recently-used idle splines until it fits
        for (auto it = lru.begin(); resident + r.footprint > budget && it != lru.end();) {
            auto sp = splines.find(*it++);
            if (sp->second.ready && sp->second.pins == 0) {
                drop(sp);
                ++evictions;
            }
        }
        if (resident + r.footprint > budget) {
            if (!r.waited) ++admission_waits;
            r.waited = true;
            return 0;  // wait for running work to finish
        }
        // reserve the memory under the request's id, then launch unlocked
        Resident res;
        res.bytes = resident_bytes_of(r.desc, r.precision);
        lru.push_back(id);
        res.lru = std::prev(lru.end());
        resident += res.bytes;
        splines.emplace(id, res);
        lk.unlock();
        for (intp_spline h : to_release) (void)intp_spline_release(h);
        to_release.clear();
        intp_spline h = nullptr;
        int rc = r.precision == INTP_PREC_F32_REFINED
                     ? intp_build_refined(ctx, &r.desc, r.values, INTP_MEM_DEVICE, r.sweeps, &h, st)
                     : intp_build(ctx, &r.desc, r.values, INTP_MEM_DEVICE, &h, st);
        hipEvent_t ev = nullptr, built = nullptr;
        if (rc == INTP_OK) {
            ev = take_event();
            built = take_event();
            if (!ev || !built || hipEventRecord(ev, st) != hipSuccess || hipEventRecord(built, st) != hipSuccess) {
                (void)hipStreamSynchronize(st);  // the build is in flight; let it finish before dropping it
                rc = fail(INTP_E_HIP, "hipEventRecord");
            }
        }
        lk.lock();
        auto it = splines.find(id);
        if (rc != INTP_OK) {
            if (ev) spare_events.push_back(ev);
            if (built) spare_events.push_back(built);
            it->second.h = h;  // released with the entry
            drop(it);
            finish(r, rc);
            return 1;
        }
        it->second.h = h;
        it->second.built = built;
        r.done = ev;
    } else {
        auto it = splines.find(r.spline);
        if (it == splines.end()) {
            // its build may still be queued, waiting for memory: wait with it
            auto bq = reqs.find(r.spline);
            if (bq != reqs.end() && bq->second->kind == INTP_REQ_BUILD && bq->second->state == INTP_REQ_QUEUED)
                return 0;
        }
        if (it == splines.end() || it->second.released) {
            finish(r, fail(INTP_E_EVICTED, "spline " + std::to_string(r.spline) + " is not resident"));
            return 1;
        }
        Resident& s = it->second;
        ++s.pins;  // not evictable while the launch is prepared and runs
        lru.splice(lru.end(), lru, s.lru);
        const intp_spline h = s.h;
        const hipEvent_t wait_for = s.ready ? nullptr : s.built;
        lk.unlock();
        for (intp_spline v : to_release) (void)intp_spline_release(v);
        to_release.clear();
        int rc = INTP_OK;
        if (wait_for && hipStreamWaitEvent(st, wait_for, 0) != hipSuccess) rc = fail(INTP_E_HIP, "hipStreamWaitEvent");
        if (rc == INTP_OK)
            rc = intp_eval(h, r.npts, r.coords, r.cs_dim, r.cs_pt, r.nout, r.orders.data(), r.out, r.os_out, r.os_pt,
                           r.flags, st);
        hipEvent_t ev = nullptr;
        if (rc == INTP_OK) {
            ev = take_event();
            if (!ev || hipEventRecord(ev, st) != hipSuccess) {
                (void)hipStreamSynchronize(st);  // the launch is in flight; finish it before reporting
                if (ev) spare_events.push_back(ev);
                ev = nullptr;
                rc = fail(INTP_E_HIP, "hipEventRecord");
            }
        }
        lk.lock();
        if (rc != INTP_OK) {
            --splines.find(r.spline)->second.pins;
            finish(r, rc);
            return 1;
        }
        r.done = ev;
    }
    r.state = INTP_REQ_RUNNING;
    inflight[si].push_back(id);
    ++inflight_n;
    return 1;
}

// Completes launched requests whose events have fired, and frees released
// splines nothing uses any more (dispatcher, holding mu).
void intp_server_s::poll_inflight() {
    for (std::deque<int64_t>& q_stream : inflight) {
        while (!q_stream.empty()) {
            Request& r = *reqs[q_stream.front()];
            const hipError_t q = hipEventQuery(r.done);
            if (q == hipErrorNotReady) break;
            q_stream.pop_front();
            --inflight_n;
            auto it = splines.find(r.spline);
            if (it != splines.end()) {
                if (r.kind == INTP_REQ_EVAL)
                    --it->second.pins;
                else
                    it->second.ready = true;
            }
            finish(r, q == hipSuccess ? INTP_OK : hip_fail(q, "request"));
        }
    }
    for (auto it = splines.begin(); it != splines.end();) {
        auto cur = it++;
        if (cur->second.released && cur->second.ready && cur->second.pins == 0) drop(cur);
    }
}

void intp_server_s::run() {
    DeviceScope scope(device);
    std::unique_lock<std::mutex> lk(mu);
    std::deque<int64_t> waiting;  // requests held back (memory, their build), oldest first
    for (;;) {
        bool progressed = false;
        // requests that waited go first, in arrival order
        std::deque<int64_t> turn;
        turn.swap(waiting);
        while (!queue.empty()) {
            turn.push_back(queue.front());
            queue.pop_front();
        }
        for (size_t i = 0; i < turn.size(); ++i) {
            const int rc = dispatch(turn[i], lk);
            if (rc < 0) {
                waiting.insert(waiting.end(), turn.begin() + static_cast<std::ptrdiff_t>(i), turn.end());
                break;
            }
            if (rc == 0)
                waiting.push_back(turn[i]);
            else
                progressed = true;
        }
        const int64_t before = inflight_n;
        poll_inflight();
        progressed = progressed || inflight_n != before;
        if (!to_release.empty()) {
            lk.unlock();
            for (intp_spline h : to_release) (void)intp_spline_release(h);
            to_release.clear();
            lk.lock();
        }
        if (stop && queue.empty() && waiting.empty() && inflight_n == 0) break;
        if (!progressed) {
            if (inflight_n == 0 && queue.empty() && waiting.empty())
                wake.wait(lk, [&] { return stop || !queue.empty(); });
            else
                wake.wait_for(lk, std::chrono::microseconds(20));
        }
    }
}

extern "C" {

int intp_server_create(intp_ctx ctx, const intp_server_config* cfg, intp_server* out) {
    if (!ctx || !out) return fail(INTP_E_INVALID, "null argument");
    intp_server_config c{};
    if (cfg) c = *cfg;
    if (c.streams == 0) c.streams = 4;
    if (c.max_inflight == 0) c.max_inflight = 256;
    if (c.streams < 1 || c.streams > kMaxServeStreams || c.max_inflight < 1 || c.hbm_budget < 0)
        return fail(INTP_E_INVALID, "server config out of range");
    std::unique_ptr<intp_server_s> s(new intp_server_s());
    s->ctx = ctx;
    s->device = ctx->device;
    DeviceScope scope(s->device);
    if (!scope.ok()) return fail(INTP_E_HIP, "cannot select device");
    if (c.hbm_budget == 0) {
        size_t free_b = 0, total_b = 0;
        HIP_TRY(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        c.hbm_budget = std::max<int64_t>(int64_t{1} << 30, static_cast<int64_t>(free_b) - (int64_t{8} << 30));
    }
    s->budget = c.hbm_budget;
    s->max_inflight = c.max_inflight;
    for (int i = 0; i < c.streams; ++i) {
        hipStream_t st = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        s->streams.push_back(st);
    }
    s->inflight.resize(s->streams.size());
    intp_server_s* raw = s.get();
    s->dispatcher = std::thread([raw] { raw->run(); });
    *out = s.release();
    return INTP_OK;
}

int intp_server_destroy(intp_server s) {
    if (!s) return fail(INTP_E_INVALID, "null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->wake.notify_all();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    DeviceScope scope(s->device);
    for (auto& kv : s->splines) {
        (void)intp_spline_release(kv.second.h);
        if (kv.second.built) (void)hipEventDestroy(kv.second.built);
    }
    for (auto& kv : s->reqs)
        if (kv.second->done) (void)hipEventDestroy(kv.second->done);
    for (hipEvent_t e : s->spare_events) (void)hipEventDestroy(e);
    for (hipStream_t st : s->streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    delete s;
    return INTP_OK;
}

int intp_serve_build(intp_server s, const intp_grid_desc* desc, const double* values, int32_t precision,
                     int32_t sweeps, int64_t* request) {
    if (!s || !desc || !values || !request) return fail(INTP_E_INVALID, "null argument");
    if (precision != INTP_PREC_F64 && precision != INTP_PREC_F32_REFINED)
        return fail(INTP_E_INVALID, "precision must be INTP_PREC_F64 or INTP_PREC_F32_REFINED");
    if (desc->dim < 1 || desc->dim > INTP_MAX_DIM) return fail(INTP_E_UNSUPPORTED, "dimension without kernels");
    for (int d = 0; d < desc->dim; ++d)
        if (desc->n[d] < 1) return fail(INTP_E_INVALID, "empty dimension");
    Request r;
    r.kind = INTP_REQ_BUILD;
    r.desc = *desc;
    for (int d = 0; d < desc->dim; ++d)
        if (desc->coords[d]) {
            r.coords_copy.emplace_back(desc->coords[d], desc->coords[d] + desc->n[d] + (desc->periodic[d] ? 1 : 0));
            r.desc.coords[d] = r.coords_copy.back().data();
        }
    r.values = values;
    r.precision = precision;
    r.sweeps = sweeps;
    r.footprint = build_footprint(*desc, precision);
    r.bytes = build_bytes(*desc, precision, sweeps);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        const int64_t id = s->next_id++;
        r.spline = id;
        r.t_submit = s->now();
        if (s->first_submit < 0) s->first_submit = r.t_submit;
        s->reqs.emplace(id, std::make_unique<Request>(std::move(r)));
        s->queue.push_back(id);
        ++s->unfinished;
        *request = id;
    }
    s->wake.notify_one();
    return INTP_OK;
}

int intp_serve_eval(intp_server s, int64_t spline, int64_t npts, const double* coords, int64_t cs_dim,
                    int64_t cs_pt, int32_t nout, const int32_t* orders, double* out, int64_t os_out, int64_t os_pt,
                    uint32_t flags, int64_t* request) {
    if (!s || !request || (npts > 0 && (!coords || !out)) || nout < 1 || !orders)
        return fail(INTP_E_INVALID, "null argument");
    if (flags & (INTP_EVAL_BOUNDS_CHECK | INTP_EVAL_HOST_MEM | INTP_EVAL_TIMED))
        return fail(INTP_E_INVALID, "served evaluations take device buffers without bounds checks or timing");
    Request r;
    r.kind = INTP_REQ_EVAL;
    r.spline = spline;
    r.npts = npts;
    r.coords = coords;
    r.cs_dim = cs_dim;
    r.cs_pt = cs_pt;
    r.nout = nout;
    r.out = out;
    r.os_out = os_out;
    r.os_pt = os_pt;
    r.flags = flags;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto b = s->reqs.find(spline);
        if (b == s->reqs.end() && spline > 0 && spline < s->next_id)
            return fail(INTP_E_EVICTED, "spline " + std::to_string(spline) + " is not resident");
        if (b == s->reqs.end() || b->second->kind != INTP_REQ_BUILD)
            return fail(INTP_E_INVALID, "no build request " + std::to_string(spline));
        const intp_grid_desc& d = b->second->desc;
        r.orders.assign(orders, orders + static_cast<int64_t>(nout) * d.dim);
        int64_t n = 1;
        for (int i = 0; i < d.dim; ++i) n *= d.n[i];
        const double s_out = (flags & INTP_EVAL_OUT_F32) ? 4 : 8;
        const double s_tab = b->second->precision == INTP_PREC_F32_REFINED ? 4 : 8;
        r.bytes = static_cast<double>(npts) * (d.dim * 8 + nout * s_out) + static_cast<double>(n) * s_tab;
        const int64_t id = s->next_id++;
        r.t_submit = s->now();
        if (s->first_submit < 0) s->first_submit = r.t_submit;
        s->reqs.emplace(id, std::make_unique<Request>(std::move(r)));
        s->queue.push_back(id);
        ++s->unfinished;
        *request = id;
    }
    s->wake.notify_one();
    return INTP_OK;
}

int intp_serve_poll(intp_server s, int64_t request, int32_t* state, int32_t* status) {
    if (!s || !state) return fail(INTP_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request);
    if (it == s->reqs.end()) return fail(INTP_E_INVALID, "no such request");
    *state = it->second->state;
    if (status) *status = it->second->status;
    return INTP_OK;
}

int intp_serve_wait(intp_server s, int64_t request) {
    if (!s) return fail(INTP_E_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->reqs.find(request) == s->reqs.end()) return fail(INTP_E_INVALID, "no such request");
    int st = INTP_OK;
    // looked up on every wake-up: the table may grow (or a statistics reset
    // may forget the finished request) while this thread sleeps
    s->done.wait(lk, [&] {
        auto it = s->reqs.find(request);
        if (it == s->reqs.end()) return true;
        st = it->second->status;
        return it->second->state >= INTP_REQ_DONE;
    });
    if (st != INTP_OK) return fail(st, "request " + std::to_string(request) + " failed: " + intp_status_string(st));
    return INTP_OK;
}

int intp_serve_drain(intp_server s) {
    if (!s) return fail(INTP_E_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->done.wait(lk, [&] { return s->unfinished == 0; });
    return INTP_OK;
}

int intp_serve_release(intp_server s, int64_t spline) {
    if (!s) return fail(INTP_E_INVALID, "null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->splines.find(spline);
        if (it == s->splines.end()) return INTP_OK;  // already evicted
        it->second.released = true;
    }
    s->wake.notify_one();
    return INTP_OK;
}

int intp_serve_spline(intp_server s, int64_t spline, intp_spline* out) {
    if (!s || !out) return fail(INTP_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->splines.find(spline);
    if (it == s->splines.end() || it->second.released)
        return fail(INTP_E_EVICTED, "spline " + std::to_string(spline) + " is not resident");
    if (!it->second.ready) return fail(INTP_E_INVALID, "spline " + std::to_string(spline) + " is still building");
    const int rc = intp_spline_retain(it->second.h);
    if (rc != INTP_OK) return rc;
    *out = it->second.h;
    return INTP_OK;
}

int intp_serve_stats(intp_server s, int32_t reset, intp_serve_counters* out) {
    if (!s || !out) return fail(INTP_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof(*out));
    for (int k = 0; k < 2; ++k) {
        intp_serve_type_counters& t = out->type[k];
        std::vector<double> v = s->lat[k];
        t.completed = static_cast<int64_t>(v.size());
        t.failed = s->failed[k];
        t.bytes = s->bytes[k];
        if (!v.empty()) {
            std::sort(v.begin(), v.end());
            auto pct = [&](double p) {
                const size_t i = std::min(v.size() - 1, static_cast<size_t>(p * static_cast<double>(v.size() - 1) + 0.5));
                return v[i];
            };
            t.latency_p50_ms = pct(0.50);
            t.latency_p99_ms = pct(0.99);
            double sum = 0;
            for (double x : v) sum += x;
            t.latency_mean_ms = sum / static_cast<double>(v.size());
            t.latency_max_ms = v.back();
        }
    }
    out->resident_splines = static_cast<int64_t>(s->splines.size());
    out->resident_bytes = s->resident;
    out->evictions = s->evictions;
    out->admission_waits = s->admission_waits;
    out->first_submit_s = s->first_submit;
    out->last_done_s = s->last_done;
    if (reset) {
        for (int k = 0; k < 2; ++k) {
            s->lat[k].clear();
            s->bytes[k] = 0;
            s->failed[k] = 0;
        }
        s->evictions = s->admission_waits = 0;
        s->first_submit = -1;
        s->last_done = 0;
        // forget finished requests, keeping the build records of resident
        // splines (evaluations look their grids up)
        for (auto it = s->reqs.begin(); it != s->reqs.end();) {
            const Request& r = *it->second;
            const bool finished = r.state >= INTP_REQ_DONE;
            const bool keep = r.kind == INTP_REQ_BUILD && s->splines.count(it->first);
            if (finished && !keep)
                it = s->reqs.erase(it);
            else
                ++it;
        }
    }
    return INTP_OK;
}

}  // extern "C"
