// g4s_batch.hip -- batched form: many independent networks resident on one
// device, advanced together by one launch per step (include/gpu4snn.h,
// g4s_batch_*).
//
// Each instance is a network of its own (state, connectivity, slice offset
// table) in one device allocation.  A chunk of `cs` steps is one launch of
// k_batch_step per step over the work list {slot, slice} of every unfinished
// instance, replayed from a captured HIP graph.  The kernels read the
// instance descriptors and the work list from device memory, so the graph
// stays valid while instances come and go; it is captured again only when
// the chunk length, the padded grid, the LDS size or a pool address changes.
// Raster rows (instances that keep them) and per-step counts are copied to
// the host once per chunk.
#include "g4s_build.h"
#include "g4s_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

using namespace g4s;

#define HIP_OK G4S_HIP_OK

namespace {

constexpr int kBlock = 256;          // threads per workgroup of the batched kernels
constexpr int kDefaultChunk = 100;
constexpr int kMaxGroups = 4;         // schedule 0: instance groups on their own streams
// Default group count.  At 768-neuron slices, 256 networks, in a process
// whose only streams are the batch's: 4 groups run 3237-3249 (atomic) and
// 2089-2092 (det) simulated seconds per second, 3 groups 3182-3206 and
// 2012-2021, 2 groups 2942 and 1811 (one box).  With one more stream in the
// process (a single network's, as in the benchmark that runs both), 4 groups
// fell to 2890 while 3 groups kept 3172-3227, most likely because the
// process has 4 hardware queues and a fifth stream then shares one with a
// group.  3 is the default.
constexpr int kDefaultGroups = 3;
constexpr int kMaxSliceLen = 4096;
// Queue schedule: a work unit packs its chunk step and slice in 16 bits each.
constexpr int kMaxQueueField = 65535;

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct Layout {  // sub-allocation of one instance's device block
    size_t v, u, a, b, c, d, lastI, nospks, bm0, bm1, syn, soff, total;
};

Layout layout_of(int64_t n, int32_t nsyn, int64_t P)
{
    Layout l{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += align256(bytes);
        return at;
    };
    const size_t words = (size_t)((n + 31) / 32);
    l.v = take(n * 4); l.u = take(n * 4); l.a = take(n * 4); l.b = take(n * 4);
    l.c = take(n * 4); l.d = take(n * 4); l.lastI = take(n * 4); l.nospks = take(n * 4);
    l.bm0 = take(words * 4); l.bm1 = take(words * 4);
    l.syn = take((size_t)n * (size_t)nsyn * 8);
    l.soff = take((size_t)n * (size_t)(P + 1) * 2);
    l.total = off;
    return l;
}

// Default slice length: 768 neurons (fewer when N is smaller).  Measured on
// 256 networks with N log-uniform on [1e3, 5e4] and Nsyn in {100, 500,
// 1000}, 256-thread workgroups: atomic mode 3205-3212 simulated seconds per
// second at 768, against 2396 (384), 2788 (512), 3027 (640), 3045 (704),
// 3094 (832), 3131 (896) and 3117 (1024); det mode 2025 at 768, 2011 at 512,
// 1962 at 640, 1752 at 1024.  Whole multiples of the workgroup (3 neurons per
// thread) win; fewer, longer slices pay the fixed part of a workgroup (bitmap
// scan, offsets, update) less often.  A Nsyn-scaled choice (about one
// 2048-entry window per slice and step) is slower.
int default_slice_len(int64_t n, int32_t nsyn)
{
    (void)nsyn;
    return (int)std::min<int64_t>(768, (n + 63) / 64 * 64);
}

struct Inst {
    bool used = false;
    g4s_spec spec{};
    DevNet d{};
    void *block = nullptr;
    size_t block_bytes = 0;  // allocated size of block (a reused block may be larger)
    int64_t bytes = 0;      // device block + this instance's share of the pools
    int64_t t = 0, steps = 0, spikes = 0;
    bool keep_raster = false;
    std::vector<int32_t> counts;
    std::vector<uint32_t> raster;
    size_t ring_off = 0, bcnt_off = 0;  // in the pools, this chunk
};

}  // namespace

struct g4s_batch {
    int accum = G4S_ACCUM_ATOMIC;
    g4s_batch_opts opts{};
    int chunk = kDefaultChunk;
    int block = kBlock;  // threads per workgroup of the per-step kernel
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // schedule 0: instance groups; group 0 runs on `stream`, group g >= 1 on
    // side[g-1], forked after the chunk's uploads and joined before its copies
    int groups = 3;
    hipStream_t side[kMaxGroups - 1] = {};
    hipEvent_t fork = nullptr, join[kMaxGroups - 1] = {};
    std::vector<Inst> inst;
    int64_t bytes_in_use = 0;
    // descriptor table and its pinned mirror
    DevNet *d_desc = nullptr, *h_desc = nullptr;
    int desc_cap = 0;
    // work list, padded to the grid
    uint2 *d_work = nullptr, *h_work = nullptr;
    int work_cap = 0;
    // chunk pools: raster rows, per-slice counts, per-instance counts
    uint32_t *d_ring = nullptr, *h_ring = nullptr;
    size_t ring_cap = 0;
    int32_t *d_bcnt = nullptr;
    size_t bcnt_cap = 0;
    int32_t *d_slots = nullptr, *h_slots = nullptr;
    int32_t *d_counts = nullptr, *h_counts = nullptr;
    size_t counts_cap = 0;
    int slots_cap = 0;
    // persistent schedule: unit queue, {claimed, reserved, error}, finished
    // slices per slot
    unsigned long long *d_units = nullptr, *h_units = nullptr;
    size_t units_cap = 0;
    int32_t *d_ctr = nullptr, *h_ctr = nullptr;
    int32_t *d_done = nullptr;
    size_t done_cap = 0;
    int chunk_grid = 0;
    size_t chunk_grid_lds = 0;
    // graph of one chunk, one per group
    hipGraph_t graph[kMaxGroups] = {};
    hipGraphExec_t exec[kMaxGroups] = {};
    int g_cs = 0, g_groups = 0, g_grid[kMaxGroups] = {};
    size_t g_lds = 0;
    int64_t gen = 0, g_gen = -1;
    g4s_timing timing{};
    // Device blocks of removed instances, kept for the next instances that
    // fit them: hipFree takes about 0.25 ms and waits for the device, which
    // a server admitting and retiring hundreds of requests per second pays
    // on every request.  At most kSpareBytes in all.
    std::vector<std::pair<void *, size_t>> spare;
    size_t spare_bytes = 0;
};

namespace {

constexpr size_t kSpareBytes = (size_t)16 << 30;

void drop_spare(g4s_batch *b)
{
    for (auto &e : b->spare) (void)hipFree(e.first);
    b->spare.clear();
    b->spare_bytes = 0;
}

// A device block of at least `need` bytes: the smallest spare block no more
// than an eighth larger, else a new allocation (after releasing the spare
// blocks if the device is out of memory).  *got = its size; null on failure.
void *take_block(g4s_batch *b, size_t need, size_t *got)
{
    size_t best = 0;
    int k_best = -1;
    for (size_t k = 0; k < b->spare.size(); ++k) {
        const size_t have = b->spare[k].second;
        if (have >= need && have - need <= need / 8 && (k_best < 0 || have < best)) {
            best = have;
            k_best = (int)k;
        }
    }
    if (k_best >= 0) {
        void *p = b->spare[(size_t)k_best].first;
        b->spare.erase(b->spare.begin() + k_best);
        b->spare_bytes -= best;
        *got = best;
        return p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, need) != hipSuccess) {
        (void)hipGetLastError();  // the failed allocation's status is not a later launch's
        if (b->spare.empty()) return nullptr;
        drop_spare(b);
        p = nullptr;
        if (hipMalloc(&p, need) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    *got = need;
    return p;
}

// Keep a block no longer used by any queued work for reuse, releasing the
// oldest spare blocks beyond kSpareBytes.
void put_block(g4s_batch *b, void *p, size_t bytes)
{
    if (!p) return;
    if (bytes > kSpareBytes / 4) {
        (void)hipFree(p);
        return;
    }
    size_t drop = 0;
    while (drop < b->spare.size() && b->spare_bytes + bytes > kSpareBytes) {
        (void)hipFree(b->spare[drop].first);
        b->spare_bytes -= b->spare[drop].second;
        ++drop;
    }
    b->spare.erase(b->spare.begin(), b->spare.begin() + (ptrdiff_t)drop);
    b->spare.emplace_back(p, bytes);
    b->spare_bytes += bytes;
}

template <class T>
int grow_dev(g4s_batch *b, T **d, T **h, size_t *cap, size_t need)
{
    if (need <= *cap) return G4S_OK;
    size_t want = std::max<size_t>(need, *cap + *cap / 2);
    want = std::max<size_t>(want, 64);
    if (*d) (void)hipFree(*d);
    if (h && *h) (void)hipHostFree(*h);
    *d = nullptr;
    if (h) *h = nullptr;
    if (hipMalloc((void **)d, want * sizeof(T)) != hipSuccess) {
        *cap = 0;
        return fail(G4S_ERR_MEM, "batch pool of %zu bytes", want * sizeof(T));
    }
    if (h && hipHostMalloc((void **)h, want * sizeof(T)) != hipSuccess) {
        *cap = 0;
        return fail(G4S_ERR_MEM, "pinned staging of %zu bytes", want * sizeof(T));
    }
    *cap = want;
    b->gen += 1;
    return G4S_OK;
}

int grow_int(g4s_batch *b, int32_t **d, int32_t **h, int *cap, int need)
{
    size_t c = (size_t)*cap;
    int rc = grow_dev(b, d, h, &c, (size_t)need);
    *cap = (int)c;
    return rc;
}

void drop_graph(g4s_batch *b)
{
    for (int g = 0; g < kMaxGroups; ++g) {
        if (b->exec[g]) (void)hipGraphExecDestroy(b->exec[g]);
        if (b->graph[g]) (void)hipGraphDestroy(b->graph[g]);
        b->exec[g] = nullptr;
        b->graph[g] = nullptr;
    }
    b->g_groups = 0;
}

int check_slot(const g4s_batch *b, int32_t slot)
{
    if (!b) return fail(G4S_ERR_ARG, "batch is NULL");
    if (slot < 0 || slot >= (int32_t)b->inst.size() || !b->inst[(size_t)slot].used)
        return fail(G4S_ERR_ARG, "no instance in slot %d", slot);
    return G4S_OK;
}

// Tuned slice lengths per bucket (floor(log2 N), Nsyn, drive class, accum),
// filled by g4s_batch_tune / g4s_batch_tune_mix for the life of the
// process.  The drive class -- regime and thalamic scale, which set the
// firing rate -- stands for the rate bucket: it is known when an instance is
// added, the rate only after it has run.
struct TunedLen {
    int slice_len;
    float us_per_inst_step;
    float rate_hz;
};
using BucketKey = std::tuple<int, int, int, int>;
std::mutex g_tune_mu;
std::map<BucketKey, TunedLen> g_tuned;

int ilog2i(int64_t x)
{
    int r = -1;
    while (x > 0) {
        x >>= 1;
        ++r;
    }
    return r;
}

// Drive class: regime and thalamic scale in quarters (scale 0.8-1.5 of the
// balanced regime falls in classes 3, 4 and 5).
int drive_class(const g4s_spec *s)
{
    const float q = std::min(std::max(s->thal_scale, 0.0f), 64.0f) * 4.0f;
    return s->regime * 1024 + (int)q;
}

BucketKey bucket_of(const g4s_spec *s, int accum)
{
    return BucketKey(ilog2i(s->n), (int)s->nsyn, drive_class(s), accum);
}

int slice_len_for(const g4s_spec *s, int32_t requested, int accum)
{
    const int64_t n = s->n;
    if (requested > 0) return requested;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(bucket_of(s, accum));
        if (it != g_tuned.end()) return (int)std::min<int64_t>(it->second.slice_len, (n + 63) / 64 * 64);
    }
    return default_slice_len(n, s->nsyn);
}

}  // namespace

extern "C" {

int64_t g4s_batch_footprint(const g4s_spec *spec, int32_t chunk, int32_t slice_len)
{
    if (validate_spec(spec)) return -1;
    const int64_t n = spec->n;
    const int L = slice_len_for(spec, slice_len, -1);
    const int64_t P = (n + L - 1) / L;
    const int64_t words = (n + 31) / 32;
    const int cs = chunk > 0 ? chunk : kDefaultChunk;
    return (int64_t)layout_of(n, spec->nsyn, P).total + (int64_t)cs * (words + P) * 4 +
           (int64_t)sizeof(DevNet) + 64;
}

int g4s_batch_create(int accum, const g4s_batch_opts *opts, g4s_batch **out)
{
    if (!out) return fail(G4S_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (accum != G4S_ACCUM_ATOMIC && accum != G4S_ACCUM_DET) return fail(G4S_ERR_ARG, "unknown accum %d", accum);
    g4s_batch_opts o{};
    if (opts) o = *opts;
    if (o.chunk < 0 || o.chunk > 100000) return fail(G4S_ERR_ARG, "chunk %d out of range", o.chunk);
    if (o.slice_len < 0 || o.slice_len % 64 != 0 || o.slice_len > kMaxSliceLen)
        return fail(G4S_ERR_ARG, "slice_len must be a multiple of 64 <= %d", kMaxSliceLen);
    if (o.schedule < 0 || o.schedule > 1 || o.budget < 0) return fail(G4S_ERR_ARG, "bad options");
    if (o.block != 0 && o.block != 256 && o.block != 512 && o.block != 1024)
        return fail(G4S_ERR_ARG, "block must be 0, 256, 512 or 1024");
    if (o.schedule == 1 && o.block > 256) return fail(G4S_ERR_ARG, "the queue schedule runs 256-thread workgroups");
    // the queue schedule packs a work unit as rel << 48 | p << 32 | slot
    if (o.schedule == 1 && o.chunk > kMaxQueueField)
        return fail(G4S_ERR_UNSUPPORTED, "the queue schedule needs chunk <= %d", kMaxQueueField);
    if (o.streams < 0 || o.streams > kMaxGroups) return fail(G4S_ERR_ARG, "streams must be 0..%d", kMaxGroups);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(G4S_ERR_HIP, "no HIP device");
    g4s_batch *b = new g4s_batch();
    b->accum = accum;
    b->opts = o;
    b->chunk = o.chunk > 0 ? o.chunk : kDefaultChunk;
    b->block = o.block > 0 ? o.block : kBlock;
    b->groups = o.schedule == 1 ? 1 : (o.streams > 0 ? o.streams : kDefaultGroups);
    bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&b->ev0) == hipSuccess && hipEventCreate(&b->ev1) == hipSuccess;
    if (ok && b->groups > 1) ok = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
    for (int g = 1; ok && g < b->groups; ++g)  // default priority, like `stream`
        ok = hipStreamCreateWithFlags(&b->side[g - 1], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&b->join[g - 1], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        g4s_batch_destroy(b);
        return fail(G4S_ERR_HIP, "stream/event creation failed");
    }
    *out = b;
    return G4S_OK;
}

}  // extern "C"

namespace {
// g4s_batch_add with the slice length given (0: the batch's rule).
int batch_add(g4s_batch *b, const g4s_spec *spec, int64_t steps, int keep_raster, int32_t slice_len, int32_t *slot)
{
    if (!b || !slot) return fail(G4S_ERR_ARG, "NULL argument");
    int rc = validate_spec(spec);
    if (rc) return rc;
    if (steps < 0 || steps > 0x7FFFFFFFLL) return fail(G4S_ERR_ARG, "steps out of range");
    if (spec->nsyn > 65535) return fail(G4S_ERR_UNSUPPORTED, "batched instances need Nsyn < 65536");
    const int64_t n = spec->n;
    const int L = slice_len > 0 ? slice_len : slice_len_for(spec, b->opts.slice_len, b->accum);
    const bool det = b->accum == G4S_ACCUM_DET;
    if (batch_lds(L, det, b->block) > 64 * 1024)
        return fail(G4S_ERR_UNSUPPORTED, "slice length %d needs %zu bytes of LDS", L, batch_lds(L, det, b->block));
    const int64_t P = (n + L - 1) / L;
    if (b->opts.schedule == 1 && P > kMaxQueueField)
        return fail(G4S_ERR_UNSUPPORTED, "the queue schedule needs at most %d slices per instance (N=%lld gives %lld)",
                    kMaxQueueField, (long long)n, (long long)P);
    const int64_t fp = g4s_batch_footprint(spec, b->chunk, L);
    if (b->opts.budget > 0 && b->bytes_in_use + fp > b->opts.budget)
        return fail(G4S_ERR_BUDGET, "instance needs %lld bytes, %lld of %lld in use", (long long)fp,
                    (long long)b->bytes_in_use, (long long)b->opts.budget);

    const Layout lay = layout_of(n, spec->nsyn, P);
    size_t blk_bytes = 0;
    void *blk = take_block(b, lay.total, &blk_bytes);
    if (!blk) return fail(G4S_ERR_MEM, "hipMalloc(%zu) for an instance of N=%lld", lay.total, (long long)n);
    char *base = (char *)blk;
    DevNet d{};
    d.n = (int32_t)n;
    d.n_exc = (int32_t)n_exc_of(spec);
    d.nsyn = spec->nsyn;
    d.words = (int32_t)((n + 31) / 32);
    d.seed = spec->seed;
    regime_gains(spec->regime, spec->thal_scale, &d.ge, &d.gi);
    d.v = (float *)(base + lay.v); d.u = (float *)(base + lay.u);
    d.a = (float *)(base + lay.a); d.b = (float *)(base + lay.b);
    d.c = (float *)(base + lay.c); d.d = (float *)(base + lay.d);
    d.lastI = (float *)(base + lay.lastI);
    d.nospks = (int32_t *)(base + lay.nospks);
    d.bm0 = (uint32_t *)(base + lay.bm0); d.bm1 = (uint32_t *)(base + lay.bm1);
    uint2 *syn = (uint2 *)(base + lay.syn);
    d.syn = syn;
    d.P = (int32_t)P;
    d.L = L;
    if ((rc = build_synapses(spec, syn, b->stream))) {
        (void)hipFree(blk);
        return rc;
    }
    uint16_t *soff = (uint16_t *)(base + lay.soff);
    hipError_t e = launch_init(d, b->stream);
    if (e == hipSuccess) e = launch_slice_offsets(d, soff, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        (void)hipFree(blk);  // not reused: the device may be in an error state
        return fail(G4S_ERR_HIP, "instance initialisation: %s", hipGetErrorString(e));
    }
    d.soff = soff;

    int32_t s = -1;
    for (size_t k = 0; k < b->inst.size(); ++k)
        if (!b->inst[k].used) {
            s = (int32_t)k;
            break;
        }
    if (s < 0) {
        s = (int32_t)b->inst.size();
        b->inst.emplace_back();
    }
    Inst &in = b->inst[(size_t)s];
    in = Inst();
    in.used = true;
    in.spec = *spec;
    in.d = d;
    in.block = blk;
    in.block_bytes = blk_bytes;
    in.bytes = fp;
    in.steps = steps;
    in.keep_raster = keep_raster != 0;
    in.counts.reserve((size_t)steps);
    if (in.keep_raster) in.raster.reserve((size_t)steps * (size_t)d.words);
    b->bytes_in_use += fp;
    *slot = s;
    return G4S_OK;
}
}  // namespace

int g4s::batch_add_sliced(g4s_batch *b, const g4s_spec *spec, int64_t steps, int keep_raster, int32_t slice_len,
                          int32_t *slot)
{
    return batch_add(b, spec, steps, keep_raster, slice_len, slot);
}

extern "C" {

int g4s_batch_add(g4s_batch *b, const g4s_spec *spec, int64_t steps, int keep_raster, int32_t *slot)
{
    return batch_add(b, spec, steps, keep_raster, 0, slot);
}

int g4s_batch_step(g4s_batch *b, int32_t *active)
{
    if (!b) return fail(G4S_ERR_ARG, "batch is NULL");
    const auto wall0 = std::chrono::steady_clock::now();
    std::vector<int32_t> act;
    int64_t most = 0;
    for (size_t k = 0; k < b->inst.size(); ++k) {
        const Inst &in = b->inst[k];
        if (in.used && in.t < in.steps) {
            act.push_back((int32_t)k);
            most = std::max<int64_t>(most, in.steps - in.t);
        }
    }
    if (active) *active = (int32_t)act.size();
    b->timing = g4s_timing{};
    if (act.empty()) return G4S_OK;
    const int cs = (int)std::min<int64_t>(b->chunk, most);
    const bool det = b->accum == G4S_ACCUM_DET;

    // pools: instances that keep rasters first, so one copy covers them
    std::stable_sort(act.begin(), act.end(), [&](int32_t x, int32_t y) {
        return b->inst[(size_t)x].keep_raster > b->inst[(size_t)y].keep_raster;
    });
    size_t ring = 0, ring_keep = 0, bcnt = 0;

    int Lmax = 64;
    for (int32_t s : act) {
        Inst &in = b->inst[(size_t)s];
        in.ring_off = ring;
        if (in.keep_raster) {
            ring += (size_t)cs * (size_t)in.d.words;
            ring_keep = ring;
        }
        in.bcnt_off = bcnt;
        bcnt += (size_t)cs * (size_t)in.d.P;

        Lmax = std::max(Lmax, (int)in.d.L);
    }
    int rc;
    if ((rc = grow_dev(b, &b->d_ring, &b->h_ring, &b->ring_cap, std::max<size_t>(ring, 1)))) return rc;
    if ((rc = grow_dev<int32_t>(b, &b->d_bcnt, nullptr, &b->bcnt_cap, bcnt))) return rc;
    size_t dc = (size_t)b->desc_cap;
    if ((rc = grow_dev(b, &b->d_desc, &b->h_desc, &dc, b->inst.size()))) return rc;
    b->desc_cap = (int)dc;
    if ((rc = grow_int(b, &b->d_slots, &b->h_slots, &b->slots_cap, (int)act.size()))) return rc;
    if ((rc = grow_dev(b, &b->d_counts, &b->h_counts, &b->counts_cap, act.size() * (size_t)cs))) return rc;

    // descriptors of this chunk
    for (int32_t s : act) {
        Inst &in = b->inst[(size_t)s];
        DevNet d = in.d;
        d.raster = in.keep_raster ? b->d_ring + in.ring_off : nullptr;
        d.bcnt = b->d_bcnt + in.bcnt_off;
        d.t_base = (int32_t)in.t;
        d.rel_end = (int32_t)std::min<int64_t>(cs, in.steps - in.t);
        d.last_t = (int32_t)(in.steps - 1);
        b->h_desc[s] = d;
    }
    // work lists.  Schedule 0 splits the instances into groups of about
    // equal cost (longest first, each to the cheapest group); each group is
    // stepped by its own chain of launches on its own stream, so while one
    // group's step drains its last workgroups another group's step fills the
    // freed CUs.  Within a group, workgroup `blk` runs on XCD blk % 8.  Every
    // instance goes whole to one XCD (its slices share that XCD's L2: the
    // bitmap and the neighbouring segments of each row), instances are dealt
    // to XCDs longest-first to the least loaded one (cost ~ slices * L *
    // Nsyn), and each XCD runs its costlier slices first.
    std::vector<int32_t> order(act);
    auto cost = [&](int32_t s) {
        const DevNet &d = b->inst[(size_t)s].d;
        return (double)d.P * ((double)d.L * 8.0 + (double)d.L * (double)d.nsyn * 0.02);
    };
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return cost(x) > cost(y); });
    const int G = b->opts.schedule == 1 ? 1 : std::max(1, std::min<int>(b->groups, (int)act.size()));
    std::vector<int32_t> grp[kMaxGroups];
    double gload[kMaxGroups] = {};
    for (int32_t s : order) {
        const int g = (int)(std::min_element(gload, gload + G) - gload);
        gload[g] += cost(s);
        grp[g].push_back(s);
    }
    int ggrid[kMaxGroups] = {}, goff[kMaxGroups] = {};
    std::vector<uint2> wl;
    for (int g = 0; g < G; ++g) {
        std::vector<uint2> xcd[8];
        double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int32_t s : grp[g]) {
            const int x = (int)(std::min_element(load, load + 8) - load);
            load[x] += cost(s);
            for (int p = 0; p < b->inst[(size_t)s].d.P; ++p) xcd[x].push_back(make_uint2((uint32_t)s, (uint32_t)p));
        }
        size_t per = 0;
        for (auto &v : xcd) per = std::max(per, v.size());
        ggrid[g] = (int)std::max<size_t>(64, (per * 8 + 63) / 64 * 64);
        goff[g] = (int)wl.size();
        for (int blk = 0; blk < ggrid[g]; ++blk) {
            const std::vector<uint2> &v = xcd[blk % 8];
            const size_t q = (size_t)(blk / 8);
            wl.push_back(q < v.size() ? v[q] : make_uint2(kNoWork, 0u));
        }
    }
    const int grid = (int)wl.size();
    size_t wc = (size_t)b->work_cap;
    if ((rc = grow_dev(b, &b->d_work, &b->h_work, &wc, (size_t)grid))) return rc;
    b->work_cap = (int)wc;
    std::copy(wl.begin(), wl.end(), b->h_work);
    for (size_t k = 0; k < act.size(); ++k) b->h_slots[k] = act[k];
    HIP_OK(hipMemcpyAsync(b->d_desc, b->h_desc, b->inst.size() * sizeof(DevNet), hipMemcpyHostToDevice, b->stream));
    HIP_OK(hipMemcpyAsync(b->d_work, b->h_work, (size_t)grid * sizeof(uint2), hipMemcpyHostToDevice, b->stream));
    HIP_OK(hipMemcpyAsync(b->d_slots, b->h_slots, act.size() * sizeof(int32_t), hipMemcpyHostToDevice, b->stream));

    const size_t lds = batch_lds(Lmax, det, b->block);
    int launch_grid = grid;
    if (b->opts.schedule == 1) {
        // queue: the first step of every instance, costlier instances first
        size_t total = 0;
        for (int32_t s : act) {
            const Inst &in = b->inst[(size_t)s];
            total += (size_t)in.d.P * (size_t)std::min<int64_t>(cs, in.steps - in.t);
        }
        if (total > (size_t)0x7FFFFFFF)
            return fail(G4S_ERR_UNSUPPORTED, "the queue schedule holds < 2^31 work units per chunk (%zu)", total);
        if ((rc = grow_dev(b, &b->d_units, &b->h_units, &b->units_cap, total))) return rc;
        size_t dn = b->done_cap;
        if ((rc = grow_dev<int32_t>(b, &b->d_done, nullptr, &dn, b->inst.size()))) return rc;
        b->done_cap = dn;
        if (!b->d_ctr) {
            HIP_OK(hipMalloc((void **)&b->d_ctr, 4 * sizeof(int32_t)));
            HIP_OK(hipHostMalloc((void **)&b->h_ctr, 4 * sizeof(int32_t)));
        }
        size_t n0 = 0;
        for (int32_t s : order)
            for (int p = 0; p < b->inst[(size_t)s].d.P; ++p)
                b->h_units[n0++] = ((unsigned long long)p << 32) | (uint32_t)s;  // rel 0
        HIP_OK(launch_fill_u32((uint32_t *)b->d_units, total * 2, 0xFFFFFFFFu, b->stream));
        HIP_OK(launch_fill_u32((uint32_t *)b->d_done, b->inst.size(), 0u, b->stream));
        HIP_OK(hipMemcpyAsync(b->d_units, b->h_units, n0 * sizeof(unsigned long long), hipMemcpyHostToDevice,
                              b->stream));
        b->h_ctr[0] = 0;
        b->h_ctr[1] = (int32_t)n0;
        b->h_ctr[2] = 0;
        b->h_ctr[3] = 0;
        HIP_OK(hipMemcpyAsync(b->d_ctr, b->h_ctr, 4 * sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
        if (b->chunk_grid == 0 || b->chunk_grid_lds != lds) {
            const int per_cu = batch_chunk_blocks_per_cu(det, lds);
            if (per_cu <= 0) return fail(G4S_ERR_HIP, "no occupancy for the batched chunk kernel");
            b->chunk_grid = per_cu * device_cus();
            b->chunk_grid_lds = lds;
        }
        launch_grid = (int)std::min<size_t>((size_t)b->chunk_grid, total);
        HIP_OK(hipEventRecord(b->ev0, b->stream));
        HIP_OK(launch_batch_chunk(b->d_desc, b->d_units, b->d_ctr, b->d_done, (int)total, det, lds,
                                  Launch{launch_grid, kBlock, b->stream}));
    } else {
        auto st = [&](int g) { return g == 0 ? b->stream : b->side[g - 1]; };
        HIP_OK(hipEventRecord(b->ev0, b->stream));
        if (G > 1) {  // the other groups start after this chunk's uploads
            HIP_OK(hipEventRecord(b->fork, b->stream));
            for (int g = 1; g < G; ++g) HIP_OK(hipStreamWaitEvent(st(g), b->fork, 0));
        }
        if (b->opts.graph < 0) {
            for (int rel = 0; rel < cs; ++rel)
                for (int g = 0; g < G; ++g)
                    HIP_OK(launch_batch_step(b->d_desc, b->d_work + goff[g], ggrid[g], rel, det, lds,
                                             Launch{ggrid[g], b->block, st(g)}));
        } else {
            bool same = b->g_groups == G && b->g_cs == cs && b->g_lds == lds && b->g_gen == b->gen;
            for (int g = 0; same && g < G; ++g) same = b->g_grid[g] == ggrid[g];
            if (!same) {
                drop_graph(b);
                for (int g = 0; g < G; ++g) {
                    HIP_OK(hipStreamBeginCapture(st(g), hipStreamCaptureModeThreadLocal));
                    hipError_t e = hipSuccess;
                    for (int rel = 0; rel < cs && e == hipSuccess; ++rel)
                        e = launch_batch_step(b->d_desc, b->d_work + goff[g], ggrid[g], rel, det, lds,
                                              Launch{ggrid[g], b->block, st(g)});
                    hipGraph_t gr = nullptr;
                    hipError_t e2 = hipStreamEndCapture(st(g), &gr);
                    if (e != hipSuccess || e2 != hipSuccess) {
                        if (gr) (void)hipGraphDestroy(gr);
                        drop_graph(b);
                        return fail(G4S_ERR_HIP, "batch graph capture: %s",
                                    hipGetErrorString(e != hipSuccess ? e : e2));
                    }
                    b->graph[g] = gr;
                    if (hipGraphInstantiate(&b->exec[g], gr, nullptr, nullptr, 0) != hipSuccess) {
                        drop_graph(b);
                        return fail(G4S_ERR_HIP, "batch graph instantiation failed");
                    }
                    b->g_grid[g] = ggrid[g];
                }
                b->g_groups = G;
                b->g_cs = cs;
                b->g_lds = lds;
                b->g_gen = b->gen;
            }
            for (int g = 0; g < G; ++g) HIP_OK(hipGraphLaunch(b->exec[g], st(g)));
        }
        for (int g = 1; g < G; ++g) {  // join before the chunk's copies on `stream`
            HIP_OK(hipEventRecord(b->join[g - 1], st(g)));
            HIP_OK(hipStreamWaitEvent(b->stream, b->join[g - 1], 0));
        }
    }
    HIP_OK(hipEventRecord(b->ev1, b->stream));
    if (b->opts.schedule == 1)
        HIP_OK(hipMemcpyAsync(b->h_ctr, b->d_ctr, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, b->stream));
    HIP_OK(launch_batch_counts(b->d_desc, b->d_slots, (int)act.size(), cs, b->d_counts, b->stream));
    HIP_OK(hipMemcpyAsync(b->h_counts, b->d_counts, act.size() * (size_t)cs * sizeof(int32_t), hipMemcpyDeviceToHost,
                          b->stream));
    if (ring_keep)
        HIP_OK(hipMemcpyAsync(b->h_ring, b->d_ring, ring_keep * sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    if (b->opts.schedule == 1 && b->h_ctr[2] != 0)
        return fail(G4S_ERR_TIMEOUT, "a bounded wait of the batched chunk expired");
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, b->ev0, b->ev1));

    int64_t spikes = 0, steps = 0;
    for (size_t k = 0; k < act.size(); ++k) {
        Inst &in = b->inst[(size_t)act[k]];
        const int adv = (int)std::min<int64_t>(cs, in.steps - in.t);
        for (int r = 0; r < adv; ++r) {
            const int32_t c = b->h_counts[k * (size_t)cs + (size_t)r];
            in.counts.push_back(c);
            in.spikes += c;
            spikes += c;
        }
        if (in.keep_raster) {
            const uint32_t *src = b->h_ring + in.ring_off;
            in.raster.insert(in.raster.end(), src, src + (size_t)adv * (size_t)in.d.words);
        }
        in.t += adv;
        steps += adv;
    }
    b->timing.device_ms = ms;
    b->timing.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    b->timing.steps = steps;
    b->timing.spikes = spikes;
    b->timing.launches = b->opts.schedule == 1 ? 1 : cs * G;
    b->timing.grid = launch_grid;
    return G4S_OK;
}

int g4s_batch_run(g4s_batch *b)
{
    if (!b) return fail(G4S_ERR_ARG, "batch is NULL");
    g4s_timing total{};
    for (;;) {
        int32_t active = 0;
        int rc = g4s_batch_step(b, &active);
        if (rc) return rc;
        if (active == 0) break;
        total.device_ms += b->timing.device_ms;
        total.wall_ms += b->timing.wall_ms;
        total.steps += b->timing.steps;
        total.spikes += b->timing.spikes;
        total.launches += b->timing.launches;
        total.grid = std::max(total.grid, b->timing.grid);
    }
    b->timing = total;
    return G4S_OK;
}

int g4s_batch_info(const g4s_batch *b, int32_t slot, g4s_inst_info *info)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    if (!info) return fail(G4S_ERR_ARG, "info is NULL");
    const Inst &in = b->inst[(size_t)slot];
    info->n = in.d.n;
    info->nsyn = in.d.nsyn;
    info->slice_len = in.d.L;
    info->t = in.t;
    info->steps = in.steps;
    info->spikes = in.spikes;
    info->bytes = in.bytes;
    info->done = in.t >= in.steps;
    info->slices = in.d.P;
    return G4S_OK;
}

int g4s_batch_counts(const g4s_batch *b, int32_t slot, int32_t *counts)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    const Inst &in = b->inst[(size_t)slot];
    if (counts && !in.counts.empty()) memcpy(counts, in.counts.data(), in.counts.size() * sizeof(int32_t));
    return G4S_OK;
}

int g4s_batch_raster(const g4s_batch *b, int32_t slot, uint32_t *words)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    const Inst &in = b->inst[(size_t)slot];
    if (!in.keep_raster) return fail(G4S_ERR_STATE, "instance %d keeps no raster", slot);
    if (words && !in.raster.empty()) memcpy(words, in.raster.data(), in.raster.size() * sizeof(uint32_t));
    return G4S_OK;
}

int g4s_batch_state(g4s_batch *b, int32_t slot, float *v, float *u, float *I, int32_t *nospks)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    const DevNet &d = b->inst[(size_t)slot].d;
    const size_t bytes = (size_t)d.n * 4;
    HIP_OK(hipStreamSynchronize(b->stream));
    if (v) HIP_OK(hipMemcpy(v, d.v, bytes, hipMemcpyDeviceToHost));
    if (u) HIP_OK(hipMemcpy(u, d.u, bytes, hipMemcpyDeviceToHost));
    if (I) HIP_OK(hipMemcpy(I, d.lastI, bytes, hipMemcpyDeviceToHost));
    if (nospks) HIP_OK(hipMemcpy(nospks, d.nospks, bytes, hipMemcpyDeviceToHost));
    return G4S_OK;
}

int g4s_batch_remove(g4s_batch *b, int32_t slot)
{
    int rc = check_slot(b, slot);
    if (rc) return rc;
    Inst &in = b->inst[(size_t)slot];
    HIP_OK(hipStreamSynchronize(b->stream));  // no queued work still uses the block
    put_block(b, in.block, in.block_bytes);
    b->bytes_in_use -= in.bytes;
    in = Inst();
    return G4S_OK;
}

int g4s_batch_timing(const g4s_batch *b, g4s_timing *out)
{
    if (!b || !out) return fail(G4S_ERR_ARG, "NULL argument");
    *out = b->timing;
    return G4S_OK;
}

int g4s_batch_tune(int accum, const g4s_spec *rep, int32_t trial_steps, int force, g4s_batch_plan *out)
{
    int rc = validate_spec(rep);
    if (rc) return rc;
    if (!out) return fail(G4S_ERR_ARG, "out is NULL");
    if (accum != G4S_ACCUM_ATOMIC && accum != G4S_ACCUM_DET) return fail(G4S_ERR_ARG, "unknown accum %d", accum);
    if (trial_steps < 10) trial_steps = 100;
    const BucketKey key = bucket_of(rep, accum);
    if (!force) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            out->slice_len = it->second.slice_len;
            out->us_per_inst_step = it->second.us_per_inst_step;
            out->rate_hz = it->second.rate_hz;
            out->copies = 0;
            return G4S_OK;
        }
    }
    // Each candidate runs enough copies (different seeds) to keep the device
    // full -- about 2000 workgroups per step, within 4 GiB -- so that
    // throughput, not the latency of one instance, decides.
    const int64_t n = rep->n;
    const int warm = std::max(20, trial_steps / 4);
    TunedLen best{0, INFINITY, 0.0f};
    int copies = 0;
    for (int L : {128, 256, 512, 768, 1024, 2048}) {
        if (L > (n + 63) / 64 * 64 * 2 && L > 128) continue;  // one slice already covers N
        if (batch_lds(L, accum == G4S_ACCUM_DET) > 64 * 1024) continue;
        const int64_t P = (n + L - 1) / L;
        const int64_t fp = g4s_batch_footprint(rep, trial_steps + warm, L);
        copies = (int)std::max<int64_t>(4, std::min<int64_t>({(2000 + P - 1) / P, 1024, ((int64_t)4 << 30) / fp}));
        g4s_batch_opts o{};
        o.chunk = trial_steps + warm;
        o.slice_len = L;
        g4s_batch *b = nullptr;
        if ((rc = g4s_batch_create(accum, &o, &b))) return rc;
        for (int k = 0; k < copies && rc == G4S_OK; ++k) {
            g4s_spec sp = *rep;
            sp.seed = rep->seed + 0x9E3779B97F4A7C15ull * (uint64_t)(k + 1);
            int32_t slot;
            rc = g4s_batch_add(b, &sp, warm + trial_steps, 0, &slot);
        }
        // one chunk of warm-up, one timed chunk
        if (rc == G4S_OK) {
            b->chunk = warm;
            int32_t act = 0;
            rc = g4s_batch_step(b, &act);
        }
        g4s_timing tm{};
        if (rc == G4S_OK) {
            b->chunk = trial_steps;
            int32_t act = 0;
            rc = g4s_batch_step(b, &act);
            tm = b->timing;
        }
        g4s_batch_destroy(b);
        if (rc == G4S_ERR_UNSUPPORTED || rc == G4S_ERR_MEM) {
            rc = G4S_OK;
            continue;
        }
        if (rc) return rc;
        const float us = (float)(tm.device_ms * 1000.0 / ((double)trial_steps * copies));
        if (us < best.us_per_inst_step) {
            best.slice_len = L;
            best.us_per_inst_step = us;
            best.rate_hz = (float)((double)tm.spikes / ((double)tm.steps * (double)n) * 1000.0);
        }
    }
    if (best.slice_len == 0) return fail(G4S_ERR_UNSUPPORTED, "no slice length ran for N=%lld", (long long)n);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[key] = best;
    }
    out->slice_len = best.slice_len;
    out->us_per_inst_step = best.us_per_inst_step;
    out->rate_hz = best.rate_hz;
    out->copies = copies;
    return G4S_OK;
}

}  // extern "C"

namespace {
// Device time per step of the whole mix `specs`, each instance at the slice
// length of its bucket in `len`: a batch built from the same seeds, `warm`
// steps, then `steps` timed steps.
int time_mix(int accum, const g4s_batch_opts &o, const std::vector<g4s_spec> &specs,
             const std::vector<BucketKey> &keys, const std::map<BucketKey, int> &len, int warm, int steps, double *us,
             std::vector<int64_t> *spikes)
{
    g4s_batch_opts bo = o;
    bo.chunk = std::max(warm, steps);
    g4s_batch *b = nullptr;
    int rc = g4s_batch_create(accum, &bo, &b);
    if (rc) return rc;
    std::vector<int32_t> slots(specs.size());
    for (size_t k = 0; k < specs.size() && rc == G4S_OK; ++k)
        rc = batch_add(b, &specs[k], warm + steps, 0, len.at(keys[k]), &slots[k]);
    int32_t act = 0;
    if (rc == G4S_OK) {
        b->chunk = warm;
        rc = g4s_batch_step(b, &act);
    }
    if (rc == G4S_OK) {
        b->chunk = steps;
        rc = g4s_batch_step(b, &act);
    }
    if (rc == G4S_OK) {
        *us = b->timing.device_ms * 1000.0 / steps;
        if (spikes) {
            spikes->assign(specs.size(), 0);
            for (size_t k = 0; k < specs.size(); ++k) (*spikes)[k] = b->inst[(size_t)slots[k]].spikes;
        }
    }
    g4s_batch_destroy(b);
    return rc;
}
}  // namespace

extern "C" int g4s_batch_tune_mix(int accum, const g4s_spec *specs, int32_t count, int32_t trial_steps,
                                  const g4s_batch_opts *opts, int32_t max_trials, g4s_mix_plan *out)
{
    if (!specs || count <= 0 || !out) return fail(G4S_ERR_ARG, "bad arguments");
    if (accum != G4S_ACCUM_ATOMIC && accum != G4S_ACCUM_DET) return fail(G4S_ERR_ARG, "unknown accum %d", accum);
    g4s_batch_opts o{};
    if (opts) o = *opts;
    if (o.schedule != 0 || o.slice_len != 0) return fail(G4S_ERR_ARG, "tuning needs schedule 0 and slice_len 0");
    if (trial_steps < 10) trial_steps = 100;
    if (max_trials <= 0) max_trials = 40;
    std::vector<g4s_spec> sp(specs, specs + count);
    std::vector<BucketKey> keys((size_t)count);
    std::map<BucketKey, int> len;        // current plan
    std::map<BucketKey, double> weight;  // synapses of the bucket: its share of the work
    for (int32_t k = 0; k < count; ++k) {
        int rc = validate_spec(&sp[(size_t)k]);
        if (rc) return rc;
        keys[(size_t)k] = bucket_of(&sp[(size_t)k], accum);
        len[keys[(size_t)k]] = default_slice_len(sp[(size_t)k].n, sp[(size_t)k].nsyn);
        weight[keys[(size_t)k]] += (double)sp[(size_t)k].n * (double)sp[(size_t)k].nsyn;
    }
    // the largest bucket's lengths cover all: a bucket's candidates may not
    // exceed its largest N (rounded up to 64)
    std::map<BucketKey, int64_t> nmax;
    for (int32_t k = 0; k < count; ++k)
        nmax[keys[(size_t)k]] = std::max<int64_t>(nmax[keys[(size_t)k]], sp[(size_t)k].n);
    const int warm = std::max(20, trial_steps);
    int trials = 0;
    double base = 0.0;
    std::vector<int64_t> spk;
    int rc = time_mix(accum, o, sp, keys, len, warm, trial_steps, &base, &spk);
    if (rc) return rc;
    ++trials;
    const double start = base;
    // buckets by weight, heaviest first
    std::vector<std::pair<double, BucketKey>> order;
    for (auto &kv : weight) order.emplace_back(-kv.second, kv.first);
    std::sort(order.begin(), order.end());
    constexpr int kLens[] = {256, 384, 512, 640, 768, 1024, 1280, 1536, 2048};
    constexpr int kNL = (int)(sizeof(kLens) / sizeof(kLens[0]));
    const bool det = accum == G4S_ACCUM_DET;
    int changed = 0;
    for (auto &ob : order) {
        const BucketKey key = ob.second;
        const int64_t cap = (nmax[key] + 63) / 64 * 64;
        int idx = 0;
        while (idx + 1 < kNL && kLens[idx] < len[key]) ++idx;
        // step down, then up; follow an improving direction while it improves
        for (int dir : {-1, 1}) {
            bool moved = false;
            for (int j = idx + dir; j >= 0 && j < kNL && trials < max_trials; j += dir) {
                const int L = kLens[j];
                if (L > cap && kLens[j - dir] >= cap) break;  // no longer changes the slicing
                if (batch_lds(L, det) > 64 * 1024) break;
                const int old = len[key];
                len[key] = L;
                double us = 0.0;
                rc = time_mix(accum, o, sp, keys, len, warm, trial_steps, &us, nullptr);
                ++trials;
                if (rc == G4S_ERR_UNSUPPORTED || rc == G4S_ERR_MEM) rc = G4S_OK, us = INFINITY;
                if (rc) return rc;
                if (us < base * 0.985) {  // a change must beat timing noise
                    base = us;
                    idx = j;
                    moved = true;
                    ++changed;
                } else {
                    len[key] = old;
                    break;
                }
            }
            if (moved) break;
            if (trials >= max_trials) break;
        }
        if (trials >= max_trials) break;
    }
    // confirm against the untuned plan once more; keep the tuned plan only
    // if it still wins
    std::map<BucketKey, int> dflt;
    for (int32_t k = 0; k < count; ++k)
        dflt[keys[(size_t)k]] = default_slice_len(sp[(size_t)k].n, sp[(size_t)k].nsyn);
    double us_t = 0.0, us_d = 0.0;
    for (int rep = 0; rep < 2; ++rep) {  // alternated, summed: drift hits both alike
        double a = 0.0, c = 0.0;
        if ((rc = time_mix(accum, o, sp, keys, len, warm, trial_steps, &a, nullptr))) return rc;
        if ((rc = time_mix(accum, o, sp, keys, dflt, warm, trial_steps, &c, nullptr))) return rc;
        us_t += a / 2;
        us_d += c / 2;
        trials += 2;
    }
    const bool keep = us_t < us_d;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (auto &kv : (keep ? len : dflt)) {
            double spikes = 0.0, neurons = 0.0;
            for (int32_t k = 0; k < count; ++k)
                if (keys[(size_t)k] == kv.first) {
                    spikes += (double)spk[(size_t)k];
                    neurons += (double)sp[(size_t)k].n;
                }
            g_tuned[kv.first] = TunedLen{kv.second, (float)(keep ? us_t : us_d),
                                         (float)(spikes / std::max(1.0, neurons) / (warm + trial_steps) * 1000.0)};
        }
    }
    out->buckets = (int32_t)len.size();
    out->changed = keep ? changed : 0;
    out->trials = trials;
    out->us_per_step_default = (float)us_d;
    out->us_per_step_tuned = (float)(keep ? us_t : us_d);
    out->us_per_step_first = (float)start;
    return G4S_OK;
}

extern "C" {

int g4s_batch_destroy(g4s_batch *b)
{
    if (!b) return G4S_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (hipStream_t st : b->side)
        if (st) (void)hipStreamSynchronize(st);
    drop_graph(b);
    for (Inst &in : b->inst)
        if (in.block) (void)hipFree(in.block);
    drop_spare(b);
    for (void *p : {(void *)b->d_desc, (void *)b->d_work, (void *)b->d_ring, (void *)b->d_bcnt, (void *)b->d_slots,
                    (void *)b->d_counts, (void *)b->d_units, (void *)b->d_ctr, (void *)b->d_done})
        if (p) (void)hipFree(p);
    for (void *p : {(void *)b->h_desc, (void *)b->h_work, (void *)b->h_ring, (void *)b->h_slots, (void *)b->h_counts,
                    (void *)b->h_units, (void *)b->h_ctr})
        if (p) (void)hipHostFree(p);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (int g = 0; g < kMaxGroups - 1; ++g) {
        if (b->join[g]) (void)hipEventDestroy(b->join[g]);
        if (b->side[g]) (void)hipStreamDestroy(b->side[g]);
    }
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return G4S_OK;
}

}  // extern "C"
