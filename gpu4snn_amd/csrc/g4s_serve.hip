// g4s_serve.hip -- serving loop over the batched form (include/gpu4snn.h,
// g4s_server_*).  Host code only.
//
// Requests (a network spec and a step count) are submitted from any thread.
// One worker thread owns a g4s_batch and loops:
//   1. admit pending requests in arrival order while their device footprint
//      (g4s_batch_footprint) fits the budget -- the head of the queue waits
//      for memory rather than being overtaken;
//   2. run one chunk: every admitted request advances, each up to `chunk`
//      steps, in the same launches (continuous batching: requests that
//      arrived during the previous chunk join here);
//   3. retire requests that reached their step count: copy their counts and
//      raster to the request record, free their device memory, stamp the
//      completion time.
// Network state stays resident on the device between chunks.  Times are
// seconds on the steady clock since the server was created.
#include "g4s_internal.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

using namespace g4s;

namespace {

enum { kPending = 0, kActive = 1, kDone = 2, kFailed = 3 };

struct Request {
    g4s_spec spec{};
    int64_t steps = 0;
    bool keep_raster = false;
    int state = kPending;
    int error = 0;
    int32_t slot = -1;
    int64_t bytes = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int64_t t = 0, spikes = 0;
    std::vector<int32_t> counts;
    std::vector<uint32_t> raster;
};

}  // namespace

struct g4s_server {
    g4s_server_opts opts{};
    g4s_batch *batch = nullptr;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::map<int64_t, Request> reqs;
    std::deque<int64_t> pending;
    std::vector<int64_t> active;
    int64_t admitting = 0;  // taken from `pending`, not yet active or failed
    bool fixed_slice = false;  // the caller chose the slice length
    int64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    g4s_server_stats stats{};

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// Slice length of a request admitted now.  Unless the caller fixed it:
// shorter slices while few requests are resident (a chunk's steps are then
// latency-bound, and more slices per instance run them sooner), kBusySlice
// once they pile up (the device is throughput-bound, and longer slices look
// up each spike fewer times).  Measured: at 50 requests per second (about 10
// resident) p50 / p99 0.035 / 0.105 s with 256-neuron slices, 0.039 / 0.105 s
// with 384; at 400 per second (about 85 resident) 0.126 / 0.296 s with 256,
// 0.121 / 0.270 s with 384, 0.19 / 0.38 s with 768 beyond 48 resident; at
// 1000 per second (offered about 2540 simulated seconds per second) 2198
// served with 384, 2363 with 512, 2417 with 768.
constexpr int32_t kIdleSlice = 256, kLightSlice = 384, kBusySlice = 768;
constexpr size_t kLightActive = 32, kBusyActive = 192;
// Called by the worker, the only thread that changes `active` and
// `admitting`, so it reads them without the lock.
int32_t slice_for(const g4s_server *s, const g4s_spec &spec)
{
    if (s->fixed_slice) return s->opts.slice_len;
    const size_t resident = s->active.size() + (size_t)s->admitting;
    const int32_t want = resident < kLightActive ? kIdleSlice : resident < kBusyActive ? kLightSlice : kBusySlice;
    return (int32_t)std::min<int64_t>(want, (spec.n + 63) / 64 * 64);
}

// Worker: admit, run a chunk, retire; sleeps while there is nothing to do.
void serve_loop(g4s_server *s)
{
    for (;;) {
        std::vector<int64_t> admit;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty() || !s->active.empty(); });
            if (s->stop) return;
            int64_t room = s->opts.budget > 0 ? s->opts.budget - s->stats.bytes_in_use : INT64_MAX;
            while (!s->pending.empty()) {
                Request &r = s->reqs[s->pending.front()];
                if (r.bytes > room) break;
                room -= r.bytes;
                s->stats.bytes_in_use += r.bytes;  // reserved before the add
                admit.push_back(s->pending.front());
                s->pending.pop_front();
                s->admitting += 1;
            }
        }
        // build and upload outside the lock (submitters keep going)
        for (int64_t id : admit) {
            Request *r;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                r = &s->reqs[id];
            }
            int32_t slot = -1;
            const int rc =
                batch_add_sliced(s->batch, &r->spec, r->steps, r->keep_raster ? 1 : 0, slice_for(s, r->spec), &slot);
            std::lock_guard<std::mutex> lk(s->mu);
            r->t_admit = s->now();
            s->admitting -= 1;
            if (rc) {
                r->state = kFailed;
                r->error = rc;
                r->t_done = r->t_admit;
                s->stats.bytes_in_use -= r->bytes;
                s->stats.failed += 1;
                s->cv_done.notify_all();
                continue;
            }
            r->slot = slot;
            r->state = kActive;
            s->active.push_back(id);
            s->stats.max_bytes = std::max(s->stats.max_bytes, s->stats.bytes_in_use);
        }
        int32_t n_adv = 0;
        const auto c0 = std::chrono::steady_clock::now();
        int rc = g4s_batch_step(s->batch, &n_adv);
        const double busy = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
        g4s_timing tm{};
        g4s_batch_timing(s->batch, &tm);
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.chunks += n_adv > 0 ? 1 : 0;
        s->stats.busy_ms += busy;
        s->stats.device_ms += tm.device_ms;
        s->stats.steps += tm.steps;
        s->stats.max_active = std::max<int64_t>(s->stats.max_active, (int64_t)s->active.size());
        // retire (or fail everything on a device error)
        std::vector<int64_t> still;
        for (int64_t id : s->active) {
            Request &r = s->reqs[id];
            g4s_inst_info info{};
            const int irc = rc ? rc : g4s_batch_info(s->batch, r.slot, &info);
            if (irc == G4S_OK && !info.done) {
                still.push_back(id);
                continue;
            }
            if (irc == G4S_OK) {
                r.t = info.t;
                r.spikes = info.spikes;
                r.counts.resize((size_t)info.t);
                g4s_batch_counts(s->batch, r.slot, r.counts.data());
                if (r.keep_raster) {
                    r.raster.resize((size_t)info.t * (size_t)((info.n + 31) / 32));
                    g4s_batch_raster(s->batch, r.slot, r.raster.data());
                }
                r.state = kDone;
                s->stats.completed += 1;
                s->stats.sim_steps_done += info.t;
            } else {
                r.state = kFailed;
                r.error = irc;
                s->stats.failed += 1;
            }
            g4s_batch_remove(s->batch, r.slot);
            r.slot = -1;
            r.t_done = s->now();
            s->stats.bytes_in_use -= r.bytes;
        }
        s->active.swap(still);
        s->cv_done.notify_all();
    }
}

Request *find(g4s_server *s, int64_t id)
{
    auto it = s->reqs.find(id);
    return it == s->reqs.end() ? nullptr : &it->second;
}

}  // namespace

extern "C" {

int g4s_server_create(const g4s_server_opts *opts, g4s_server **out)
{
    if (!out) return fail(G4S_ERR_ARG, "out is NULL");
    *out = nullptr;
    g4s_server_opts o{};
    if (opts) o = *opts;
    if (o.budget < 0 || o.chunk < 0) return fail(G4S_ERR_ARG, "bad server options");
    g4s_batch_opts bo{};
    bo.chunk = o.chunk > 0 ? o.chunk : 100;
    // Serving uses slices shorter than the batch default while the load is
    // light (an earlier build at 50 requests per second, p50 / p99: 0.043 /
    // 0.122 s at 256, 0.048 / 0.120 s at 384, 0.051-0.053 / 0.131-0.139 s
    // at 512, 0.062 / 0.156 s at 768, the same aggregate rate).
    // slice_for() picks each request's slice length by load; the footprints
    // reserved for requests are those of the shortest (kIdleSlice), the
    // largest.
    bo.slice_len = o.slice_len > 0 ? o.slice_len : kIdleSlice;
    int rc;
    g4s_batch *b = nullptr;
    if ((rc = g4s_batch_create(o.accum, &bo, &b))) return rc;
    g4s_server *s = new g4s_server();
    s->opts = o;
    s->opts.chunk = bo.chunk;
    s->opts.slice_len = bo.slice_len;
    s->fixed_slice = o.slice_len > 0;
    s->batch = b;
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread(serve_loop, s);
    *out = s;
    return G4S_OK;
}

int g4s_server_submit(g4s_server *s, const g4s_spec *spec, int64_t steps, int keep_raster, int64_t *id)
{
    if (!s || !id) return fail(G4S_ERR_ARG, "NULL argument");
    int rc = validate_spec(spec);
    if (rc) return rc;
    if (steps < 1 || steps > 0x7FFFFFFFLL) return fail(G4S_ERR_ARG, "steps out of range");
    const int64_t fp = g4s_batch_footprint(spec, s->opts.chunk, s->opts.slice_len);
    if (s->opts.budget > 0 && fp > s->opts.budget)
        return fail(G4S_ERR_BUDGET, "request needs %lld bytes, budget %lld", (long long)fp, (long long)s->opts.budget);
    std::lock_guard<std::mutex> lk(s->mu);
    const int64_t k = s->next_id++;
    Request &r = s->reqs[k];
    r.spec = *spec;
    r.steps = steps;
    r.keep_raster = keep_raster != 0;
    r.bytes = fp;
    r.t_submit = s->now();
    s->pending.push_back(k);
    s->stats.submitted += 1;
    s->cv_work.notify_one();
    *id = k;
    return G4S_OK;
}

int g4s_server_info(g4s_server *s, int64_t id, g4s_request_info *info)
{
    if (!s || !info) return fail(G4S_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    const Request *r = find(s, id);
    if (!r) return fail(G4S_ERR_ARG, "no request %lld", (long long)id);
    info->state = r->state;
    info->error = r->error;
    info->n = r->spec.n;
    info->steps = r->steps;
    info->t = r->t;
    info->spikes = r->spikes;
    info->submit_s = r->t_submit;
    info->admit_s = r->t_admit;
    info->done_s = r->t_done;
    return G4S_OK;
}

int g4s_server_wait(g4s_server *s, int64_t id, double timeout_s)
{
    if (!s) return fail(G4S_ERR_ARG, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    Request *r = find(s, id);
    if (!r) return fail(G4S_ERR_ARG, "no request %lld", (long long)id);
    const bool ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 1e9),
                                        [&] { return r->state == kDone || r->state == kFailed; });
    if (!ok) return fail(G4S_ERR_TIMEOUT, "request %lld not finished within %.1f s", (long long)id, timeout_s);
    return r->state == kDone ? G4S_OK : r->error;
}

int g4s_server_drain(g4s_server *s, double timeout_s)
{
    if (!s) return fail(G4S_ERR_ARG, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    const bool ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 1e9), [&] {
        return s->pending.empty() && s->admitting == 0 && s->active.empty();
    });
    return ok ? G4S_OK : fail(G4S_ERR_TIMEOUT, "server not drained within %.1f s", timeout_s);
}

int g4s_server_counts(g4s_server *s, int64_t id, int32_t *counts)
{
    if (!s) return fail(G4S_ERR_ARG, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    const Request *r = find(s, id);
    if (!r) return fail(G4S_ERR_ARG, "no request %lld", (long long)id);
    if (r->state != kDone) return fail(G4S_ERR_STATE, "request %lld not done", (long long)id);
    if (counts && !r->counts.empty()) std::copy(r->counts.begin(), r->counts.end(), counts);
    return G4S_OK;
}

int g4s_server_raster(g4s_server *s, int64_t id, uint32_t *words)
{
    if (!s) return fail(G4S_ERR_ARG, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    const Request *r = find(s, id);
    if (!r) return fail(G4S_ERR_ARG, "no request %lld", (long long)id);
    if (r->state != kDone) return fail(G4S_ERR_STATE, "request %lld not done", (long long)id);
    if (!r->keep_raster) return fail(G4S_ERR_STATE, "request %lld keeps no raster", (long long)id);
    if (words && !r->raster.empty()) std::copy(r->raster.begin(), r->raster.end(), words);
    return G4S_OK;
}

int g4s_server_release(g4s_server *s, int64_t id)
{
    if (!s) return fail(G4S_ERR_ARG, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(G4S_ERR_ARG, "no request %lld", (long long)id);
    if (it->second.state != kDone && it->second.state != kFailed)
        return fail(G4S_ERR_STATE, "request %lld still running", (long long)id);
    s->reqs.erase(it);
    return G4S_OK;
}

int g4s_server_get_stats(g4s_server *s, g4s_server_stats *out)
{
    if (!s || !out) return fail(G4S_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->pending = (int64_t)s->pending.size();
    out->active = (int64_t)s->active.size();
    out->now_s = s->now();
    return G4S_OK;
}

int g4s_server_destroy(g4s_server *s)
{
    if (!s) return G4S_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    g4s_batch_destroy(s->batch);
    delete s;
    return G4S_OK;
}

}  // extern "C"
