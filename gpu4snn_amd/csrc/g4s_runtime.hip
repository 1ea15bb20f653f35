// g4s_runtime.hip -- host runtime and C ABI (include/gpu4snn.h).
//
// Replaces the reference's main() (gpu4snn.cu:735-1332): allocation, the
// connectivity build, initNeuron, the per-step launch loop with its
// device->host spike copy, and the SKL single launch.  Differences by design:
// steps run in device-resident chunks (no per-step host round trip); the
// raster is a per-step bitmap copied once per chunk; timing uses HIP events.
#include "g4s_build.h"
#include "g4s_internal.h"

#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

using namespace g4s;

namespace {
thread_local std::string g_err;
}  // namespace

// Device scratch of g4s_run_noalloc: the per-step control rows, per-slice
// spike counts, grid-barrier counters, error flag and step info that
// g4s_run keeps in buffers of the handle.  One array per device for the
// life of the process; calls on one device share it (see
// g4s_run_noalloc).  Host code uses the address returned by
// hipGetSymbolAddress for the current device, never the array's name.
static __device__ __align__(256) unsigned char scratch[G4S_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device.
int scratch_of_current_device(unsigned char **out)
{
    int dev = 0;
    G4S_HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(G4S_ERR_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(G4S_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch[dev] = (unsigned char *)p;
    }
    *out = g_scratch[dev];
    return G4S_OK;
}

// Layout of the scratch array for a run of chunks of `cs` steps over P
// slices: barrier counters, the control block (error flag, then cs + 1
// control rows), step info, per-slice counts (cs x P).
struct ScratchLayout {
    size_t bar, err, tinfo, ctl, bcnt, end;
};
ScratchLayout scratch_layout(int cs, int P)
{
    ScratchLayout l{};
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    l.bar = 0;
    l.err = up(l.bar + (size_t)kBarWords * 4);
    l.ctl = l.err + (size_t)kCtlHead * 4;  // the control block: error flag, then rows
    l.tinfo = up(l.ctl + (size_t)(cs + 1) * kCtlStride * 4);
    l.bcnt = up(l.tinfo + 8);
    l.end = l.bcnt + (size_t)cs * (size_t)std::max(P, 0) * 4;
    return l;
}
}  // namespace

namespace g4s {

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int validate_spec(const g4s_spec *s)
{
    if (!s) return fail(G4S_ERR_ARG, "spec is NULL");
    if (s->n < 2 || s->n > 0x7FFFFFFFLL) return fail(G4S_ERR_ARG, "N=%lld outside [2, 2^31)", (long long)s->n);
    if (s->nsyn < 1 || (int64_t)s->nsyn > s->n - 1)
        return fail(G4S_ERR_ARG, "Nsyn=%d must be in [1, N-1=%lld]", s->nsyn, (long long)(s->n - 1));
    if (s->regime < 0 || s->regime > 2) return fail(G4S_ERR_ARG, "regime=%d not in {0,1,2}", s->regime);
    if (!(s->thal_scale >= 0.0f) || s->thal_scale > 1e6f) return fail(G4S_ERR_ARG, "thal_scale out of range");
    if (s->n_exc > s->n) return fail(G4S_ERR_ARG, "n_exc > N");
    return G4S_OK;
}

int64_t n_exc_of(const g4s_spec *s) { return s->n_exc >= 0 ? s->n_exc : (4 * s->n) / 5; }

int build_synapses(const g4s_spec *spec, uint2 *syn, hipStream_t stream)
{
    const int64_t n = spec->n;
    const int32_t nsyn = spec->nsyn;
    const int64_t n_exc = n_exc_of(spec);
    if (nsyn <= kDeviceBuildMaxNsyn) {
        const float scale = 1000.0f / (float)nsyn;  // the host builder's gains
        // queued on `stream`; the callers wait for the stream after the
        // instance's initialisation, which follows the build there
        const hipError_t e = launch_build_rows(spec->seed, n, n_exc, nsyn, scale * 0.5f, -scale, syn, 0, n, stream);
        if (e != hipSuccess) return fail(G4S_ERR_HIP, "device connectivity build: %s", hipGetErrorString(e));
        return G4S_OK;
    }
    const size_t ns = (size_t)n * (size_t)nsyn;
    std::vector<uint2> tab(ns);
    build_table(spec->seed, n, n_exc, nsyn, tab.data(), 0);
    if (hipMemcpy(syn, tab.data(), ns * sizeof(uint2), hipMemcpyHostToDevice) != hipSuccess)
        return fail(G4S_ERR_HIP, "upload of the synapse table failed");
    return G4S_OK;
}

}  // namespace g4s

#define HIP_OK G4S_HIP_OK

namespace {

template <class T>
hipError_t dalloc(T **p, size_t count)
{
    return hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
}

}  // namespace

struct g4s_net {
    g4s_spec spec{};
    DevNet d{};
    hipStream_t stream = nullptr;
    // Per staging buffer k (chunks alternate between two so that the host
    // can consume chunk c while chunk c+1 runs): events around the chunk's
    // step launches and at the end of its copies, pinned copies of its raster
    // rows, control words, error flag and step info.
    hipEvent_t ev0[2] = {nullptr, nullptr}, ev1[2] = {nullptr, nullptr}, done_ev[2] = {nullptr, nullptr};
    int64_t t = 0;
    int chunk_cap = 0;             // raster rows / ctl entries allocated
    // The control rows (chunk_cap + 1 of them) and the barrier words are
    // zero: cleared after the last chunk's results were queued for copy, so
    // that the next chunk starts without a clearing launch.
    bool ctl_clean = false;
    uint32_t *h_raster[2] = {nullptr, nullptr};
    int32_t *h_ctl[2] = {nullptr, nullptr};  // control block copies: error flag, then the rows
    int32_t *h_ctl_dev[2] = {nullptr, nullptr};  // their device addresses (mapped pinned memory)
    bool slices_ready = false;
    // Per-step AB launches captured once per chunk length and device layout.
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_steps = 0, graph_block = 0;
    int64_t gen = 0, graph_gen = -1;  // gen: bumped when DevNet changes
    int32_t *h_tinfo = nullptr;       // pinned staging of d.tinfo, [2][2]
    // isyn0/isyn1 hold the pending input only after the global-atomic schemes
    // (N, S, AP atomic); the slice and gather schemes derive the input from
    // the previous step's queue or bitmap instead.
    bool isyn_valid = true;
    int32_t last_count = 0;     // spikes of step t-1
    int64_t est_count = 0;      // recent max spikes per step (grid sizing)
    int persist_cache[1][17] = {};  // resident workgroups per CU per (kernel, block/64)
    int cus = 0;
    g4s_timing timing{};
    // what the last run launched; g4s_get_launch_info builds its report
    struct LastLaunch {
        int alg = -1, block = 0, grid = 0;
        bool slice = false, det = false;
    } last_launch;
    // resident workgroups per CU of SKL's instantiation per (L, block, det)
    std::map<std::tuple<int, int, bool>, int> skl_res;
    // Per-step device time of the last run (ms), from the steps' start
    // stamps on the device wall clock (clock_khz).
    std::vector<float> step_ms;
    int clock_khz = 0;
    std::vector<void *> allocs;
    // Event pairs bracketing each launch of the dominant kernel of a chunk
    // (g4s_run_opts.kernel_events); grown on demand, destroyed with the net.
    std::vector<hipEvent_t> kev;
    bool kev_on = false;
    int kev_used = 0;

    template <class T>
    int alloc(T **p, size_t count)
    {
        hipError_t e = dalloc(p, count);
        if (e != hipSuccess) return fail(G4S_ERR_MEM, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
        allocs.push_back((void *)*p);
        return G4S_OK;
    }
};

namespace {

// Grow the event pool so that two more events are available.
hipError_t kev_reserve(g4s_net *net)
{
    while ((size_t)net->kev_used + 2 > net->kev.size()) {
        hipEvent_t e = nullptr;
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return r;
        net->kev.push_back(e);
    }
    return hipSuccess;
}

// Ordinary launches: hand the next event pair to the dispatch itself.
hipError_t kev_stamp(g4s_net *net, Launch *L)
{
    L->ev_start = L->ev_stop = nullptr;
    if (!net->kev_on) return hipSuccess;
    hipError_t r = kev_reserve(net);
    if (r != hipSuccess) return r;
    L->ev_start = net->kev[net->kev_used++];
    L->ev_stop = net->kev[net->kev_used++];
    return hipSuccess;
}

// Sum of the bracketed kernel durations of the chunk (after a stream sync).
hipError_t kev_collect(g4s_net *net, double *ms)
{
    for (int k = 0; k + 1 < net->kev_used; k += 2) {
        float x = 0.0f;
        hipError_t r = hipEventElapsedTime(&x, net->kev[k], net->kev[k + 1]);
        if (r != hipSuccess) return r;
        *ms += x;
    }
    net->kev_used = 0;
    return hipSuccess;
}

int ensure_chunk(g4s_net *net, int chunk)
{
    if (chunk <= net->chunk_cap) return G4S_OK;
    // from here on the buffers are being replaced: a failed allocation below
    // must leave a handle whose next call allocates again
    net->chunk_cap = 0;
    net->gen += 1;
    if (net->d.raster) { (void)hipFree(net->d.raster); net->allocs.erase(std::find(net->allocs.begin(), net->allocs.end(), (void *)net->d.raster)); }
    // the control block: error flag (d.err) and control rows (d.ctl) in one allocation
    if (net->d.err) { (void)hipFree(net->d.err); net->allocs.erase(std::find(net->allocs.begin(), net->allocs.end(), (void *)net->d.err)); }
    for (int k = 0; k < 2; ++k) {
        if (net->h_raster[k]) (void)hipHostFree(net->h_raster[k]);
        if (net->h_ctl[k]) (void)hipHostFree(net->h_ctl[k]);
        net->h_raster[k] = nullptr;
        net->h_ctl[k] = nullptr;
        net->h_ctl_dev[k] = nullptr;
    }
    net->d.raster = nullptr; net->d.ctl = nullptr; net->d.err = nullptr;
    net->ctl_clean = false;
    int rc;
    if ((rc = net->alloc(&net->d.raster, (size_t)chunk * net->d.words))) return rc;
    int32_t *blk = nullptr;
    if ((rc = net->alloc(&blk, (size_t)kCtlHead + (size_t)(chunk + 1) * kCtlStride))) return rc;
    net->d.err = blk;
    net->d.ctl = blk + kCtlHead;
    HIP_OK(launch_fill_u32((uint32_t *)blk, kCtlHead, 0u, net->stream));
    if (net->d.bcnt) {
        (void)hipFree(net->d.bcnt);
        net->allocs.erase(std::find(net->allocs.begin(), net->allocs.end(), (void *)net->d.bcnt));
        net->d.bcnt = nullptr;
    }
    if (net->d.P > 0 && (rc = net->alloc(&net->d.bcnt, (size_t)chunk * net->d.P))) return rc;
    for (int k = 0; k < 2; ++k) {
        HIP_OK(hipHostMalloc((void **)&net->h_raster[k], (size_t)chunk * net->d.words * sizeof(uint32_t)));
        HIP_OK(hipHostMalloc((void **)&net->h_ctl[k], ((size_t)kCtlHead + (size_t)(chunk + 1) * kCtlStride) * sizeof(int32_t)));
        HIP_OK(hipHostGetDevicePointer((void **)&net->h_ctl_dev[k], net->h_ctl[k], 0));
    }
    net->chunk_cap = chunk;
    net->gen += 1;
    return G4S_OK;
}

Launch plain(g4s_net *net, int64_t threads, int block)
{
    int64_t g = (threads + block - 1) / block;
    g = std::max<int64_t>(1, std::min<int64_t>(g, 1 << 20));
    return Launch{(int)g, block, net->stream};
}

// Grid of a persistent kernel, whose workgroups wait on each other: at most
// the workgroups the device holds at once (resident_per_cu: the compiler's
// occupancy of the kernel, its LDS and the occupancy API, whichever is
// least), so every workgroup is resident while the others wait for it.
int persistent_launch(g4s_net *net, PersistentKernel k, int block, int64_t useful_threads, Launch *L)
{
    int &cached = net->persist_cache[(int)k][block / 64];
    if (cached == 0) cached = persistent_blocks_per_cu(k, block);
    if (net->cus == 0) net->cus = device_cus();
    if (cached <= 0 || net->cus <= 0)
        return fail(G4S_ERR_HIP, "no resident capacity known for persistent kernel %d at block %d", (int)k, block);
    const int per_cu = cached;
    const int maxb = per_cu * net->cus;
    int64_t want = (useful_threads + block - 1) / block;
    int g = (int)std::max<int64_t>(1, std::min<int64_t>(maxb, want));
    *L = Launch{g, block, net->stream};
    return G4S_OK;
}

// Slice partition for the LDS-accumulating schemes: `want` slices (0 = one
// per CU), slice length L a multiple of 32 (whole bitmap words) and at most
// 16*1024 neurons; offsets stored as uint16 (Nsyn < 65536).  Rebuilt when the
// slice count changes.
int64_t pick_slice_len(int64_t n, int64_t want)
{
    int64_t P = std::max<int64_t>(1, std::min<int64_t>(want, (n + 63) / 64));
    int64_t L = ((n + P - 1) / P + 31) / 32 * 32;
    return std::min<int64_t>(L, 16 * 1024);
}

int build_slices(g4s_net *net, int want)
{
    if (net->d.nsyn > 65535) return fail(G4S_ERR_UNSUPPORTED, "slice schemes need Nsyn < 65536");
    if (net->cus == 0) net->cus = device_cus();
    const int64_t n = net->d.n;
    if (want <= 0) want = std::max(net->cus, 1);
    const int64_t L = pick_slice_len(n, want);
    const int64_t P = (n + L - 1) / L;
    if (net->slices_ready && net->d.L == L) return G4S_OK;
    // LDS holds acc[L] plus the pass workspace (checked per run against
    // the LDS of one workgroup); a thread updates at most 16 neurons
    if (L > 16 * 1024)
        return fail(G4S_ERR_UNSUPPORTED, "N=%lld needs %lld slices of %lld: beyond the slice schemes' LDS", (long long)n,
                    (long long)P, (long long)L);
    net->slices_ready = false;  // a failed allocation below leaves the handle to rebuild
    net->gen += 1;
    for (void **pp : {(void **)&net->d.soff, (void **)&net->d.bcnt}) {
        if (*pp) {
            (void)hipFree(*pp);
            net->allocs.erase(std::find(net->allocs.begin(), net->allocs.end(), *pp));
            *pp = nullptr;
        }
    }
    net->d.P = (int32_t)P;
    net->d.L = (int32_t)L;
    uint16_t *soff = nullptr;
    int rc;
    if ((rc = net->alloc(&soff, (size_t)n * (size_t)(P + 1)))) return rc;
    if ((rc = net->alloc(&net->d.bcnt, (size_t)std::max(net->chunk_cap, 1) * (size_t)P))) return rc;
    HIP_OK(launch_slice_offsets(net->d, soff, net->stream));
    HIP_OK(hipStreamSynchronize(net->stream));
    net->d.soff = soff;
    net->slices_ready = true;
    net->gen += 1;
    return G4S_OK;
}

// Workgroup size of the slice kernels: requested, or by default 1024
// threads in det mode and one thread per neuron of the slice (in whole
// waves, at most 1024) in atomic mode.  More threads per slice run faster
// (the autotuner picks 1024 at N=1e5; det at N=1e5 takes 13.1 us per step
// with 1024 threads and 21.6 with one thread per neuron, whose waves then
// hold more entries than the wave-local accumulation keeps in registers).
// Det results do not depend on the thread count; in atomic mode more threads
// let the adds of a target arrive further from source order, and the first
// steps then drift from the ordered sum sooner.
int slice_block(const g4s_net *net, int requested, bool det)
{
    if (requested > 0) return requested;
    if (det) return 1024;
    return (int)std::min<int64_t>(1024, std::max<int64_t>(64, ((int64_t)net->d.L + 63) / 64 * 64));
}

void drop_graph(g4s_net *net)
{
    if (net->graph_exec) (void)hipGraphExecDestroy(net->graph_exec);
    if (net->graph) (void)hipGraphDestroy(net->graph);
    net->graph_exec = nullptr;
    net->graph = nullptr;
    net->graph_steps = 0;
}

// Graph of an AB chunk of `cs` steps: captured again when the chunk length,
// tiling, mode or any device pointer of the network changed.  Called before
// the chunk's opening event, so capture and instantiation (host work) stay
// out of the chunk's device time.
int prepare_ab_graph(g4s_net *net, int cs, int block, bool det)
{
    const Launch L{net->d.P, block, net->stream};
    const int key = block * 2 + (det ? 1 : 0);
    if (!net->graph_exec || net->graph_steps != cs || net->graph_gen != net->gen || net->graph_block != key) {
        drop_graph(net);
        HIP_OK(hipStreamBeginCapture(net->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (int rel = 0; rel < cs && e == hipSuccess; ++rel) e = launch_slice_step(net->d, rel, det, L);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(net->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(G4S_ERR_HIP, "graph capture of %d steps: %s", cs, hipGetErrorString(e != hipSuccess ? e : e2));
        }
        net->graph = g;
        HIP_OK(hipGraphInstantiate(&net->graph_exec, g, nullptr, nullptr, 0));
        net->graph_steps = cs;
        net->graph_gen = net->gen;
        net->graph_block = key;
    }
    return G4S_OK;
}

// AB chunk of `cs` steps: replayed from the graph prepare_ab_graph made, or
// launched directly.
int run_ab_chunk(g4s_net *net, int cs, bool use_graph, int block, bool det)
{
    if (use_graph) {
        HIP_OK(hipGraphLaunch(net->graph_exec, net->stream));
        return G4S_OK;
    }
    const Launch L{net->d.P, block, net->stream};
    for (int rel = 0; rel < cs; ++rel) {
        Launch Lk = L;
        HIP_OK(kev_stamp(net, &Lk));
        HIP_OK(launch_slice_step(net->d, rel, det, Lk));
    }
    return G4S_OK;
}

}  // namespace

extern "C" {

const char *g4s_last_error(void) { return g_err.c_str(); }

int g4s_version(void) { return 100; }

int g4s_build_synapses_host(const g4s_spec *spec, uint32_t *post, float *w, int32_t threads)
{
    int rc = validate_spec(spec);
    if (rc) return rc;
    if (!post || !w) return fail(G4S_ERR_ARG, "NULL output");
    const size_t ns = (size_t)spec->n * (size_t)spec->nsyn;
    std::vector<uint2> tab(ns);
    build_table(spec->seed, spec->n, n_exc_of(spec), spec->nsyn, tab.data(), threads);
    for (size_t k = 0; k < ns; ++k) {
        post[k] = tab[k].x;
        w[k] = u2f(tab[k].y);
    }
    return G4S_OK;
}

int g4s_create(const g4s_spec *spec, g4s_net **out)
{
    if (!out) return fail(G4S_ERR_ARG, "out is NULL");
    *out = nullptr;
    int rc = validate_spec(spec);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(G4S_ERR_HIP, "no HIP device");

    g4s_net *net = new g4s_net();
    net->spec = *spec;
    DevNet &d = net->d;
    d.n = (int32_t)spec->n;
    d.n_exc = (int32_t)n_exc_of(spec);
    d.nsyn = spec->nsyn;
    d.words = (int32_t)((spec->n + 31) / 32);
    d.seed = spec->seed;
    regime_gains(spec->regime, spec->thal_scale, &d.ge, &d.gi);

    auto bail = [&](int code) { g4s_destroy(net); return code; };
    if (hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(G4S_ERR_HIP, "hipStreamCreate failed"));
    {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&net->clock_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess)
            net->clock_khz = 0;
    }
    for (int k = 0; k < 2; ++k)
        if (hipEventCreate(&net->ev0[k]) != hipSuccess || hipEventCreate(&net->ev1[k]) != hipSuccess ||
            hipEventCreateWithFlags(&net->done_ev[k], hipEventDisableTiming) != hipSuccess)
            return bail(fail(G4S_ERR_HIP, "hipEventCreate failed"));
    const size_t n = (size_t)spec->n;
    float **fields[] = {&d.v, &d.u, &d.a, &d.b, &d.c, &d.d, &d.lastI, &d.isyn0, &d.isyn1};
    for (float **f : fields)
        if ((rc = net->alloc(f, n))) return bail(rc);
    if ((rc = net->alloc(&d.nospks, n))) return bail(rc);
    // bitmaps padded to whole groups of four words: the slice kernels read
    // them 16 bytes at a time
    const size_t bm_words = ((size_t)d.words + 3) & ~(size_t)3;
    if ((rc = net->alloc(&d.bm0, bm_words))) return bail(rc);
    if ((rc = net->alloc(&d.bm1, bm_words))) return bail(rc);
    if ((rc = net->alloc(&d.q0, n))) return bail(rc);
    if ((rc = net->alloc(&d.q1, n))) return bail(rc);
    if ((rc = net->alloc(&d.bar, kBarWords))) return bail(rc);
    int32_t *tinfo = nullptr;
    if ((rc = net->alloc(&tinfo, 2))) return bail(rc);
    d.tinfo = tinfo;
    if (hipHostMalloc((void **)&net->h_tinfo, 4 * sizeof(int32_t)) != hipSuccess)
        return bail(fail(G4S_ERR_MEM, "hipHostMalloc failed"));
    uint2 *syn = nullptr;
    const size_t ns = n * (size_t)spec->nsyn;
    if ((rc = net->alloc(&syn, ns))) return bail(rc);
    d.syn = syn;
    if ((rc = build_synapses(spec, syn, net->stream))) return bail(rc);
    if ((rc = ensure_chunk(net, 1000))) return bail(rc);
    if ((rc = g4s_reset(net))) return bail(rc);
    *out = net;
    return G4S_OK;
}

int g4s_reset(g4s_net *net)
{
    if (!net) return fail(G4S_ERR_ARG, "net is NULL");
    if (!net->d.err) return fail(G4S_ERR_STATE, "handle has no control block (a reallocation failed)");
    HIP_OK(launch_init(net->d, net->stream));
    HIP_OK(launch_fill_u32((uint32_t *)net->d.err, 1, 0u, net->stream));
    HIP_OK(hipStreamSynchronize(net->stream));
    net->t = 0;
    net->isyn_valid = true;
    net->last_count = 0;
    net->est_count = std::max<int64_t>(64, net->spec.n / 16);
    return G4S_OK;
}

int64_t g4s_current_step(const g4s_net *net) { return net ? net->t : -1; }

namespace {

bool uses_isyn(int alg, int accum)
{
    return accum == G4S_ACCUM_ATOMIC && (alg == G4S_ALG_AP || alg == G4S_ALG_N || alg == G4S_ALG_S);
}

// Launches of one step of a per-step atomic scheme (AP, N, S); returns the
// HIP status.  Det mode runs every scheme through the slice kernels.
hipError_t launch_step(g4s_net *net, int alg, int block, int t, int rel, int last, int *launches, Launch *L, int *rc)
{
    DevNet &d = net->d;
    *rc = G4S_OK;
    hipError_t e;
    if (alg == G4S_ALG_AP) {
        *launches += 1;
        if ((*rc = persistent_launch(net, kPersistAP, block, d.n, L))) return hipSuccess;
        if ((e = kev_stamp(net, L)) != hipSuccess) return e;
        return launch_ap(d, t, rel, last, net->last_count, *L);
    }
    *launches += 2;
    e = launch_update(d, t, rel, last, plain(net, d.n, block));
    if (e != hipSuccess) return e;
    if (alg == G4S_ALG_N) {
        *L = plain(net, net->est_count * 64, block);
        if ((e = kev_stamp(net, L)) != hipSuccess) return e;
        return launch_push_rows(d, t, rel, *L);
    }
    *L = plain(net, net->est_count * (int64_t)d.nsyn, block);
    if ((e = kev_stamp(net, L)) != hipSuccess) return e;
    return launch_push_flat(d, t, rel, *L);
}

}  // namespace

namespace {

// SKL's resident workgroups per CU at slice length L (cached per handle:
// the occupancy query is host work on every run otherwise).
int skl_resident(g4s_net *net, int L, int block, bool det)
{
    const auto key = std::make_tuple(L, block, det);
    auto it = net->skl_res.find(key);
    if (it != net->skl_res.end()) return it->second;
    const int r = skl_slice_blocks_per_cu(L, block, det);
    if (r > 0) net->skl_res[key] = r;
    return r;
}

// What a run launched, for g4s_get_launch_info.
void record_launch(g4s_net *net, int alg, bool slice, int block, int grid, bool det)
{
    net->last_launch.alg = alg;
    net->last_launch.slice = slice;
    net->last_launch.block = block;
    net->last_launch.grid = grid;
    net->last_launch.det = det;
}

// Launch report of the dominant kernel of the last run.
g4s_launch_info launch_report(g4s_net *net)
{
    g4s_launch_info li{};
    const int alg = net->last_launch.alg;
    if (alg < 0) return li;
    const bool slice = net->last_launch.slice, det = net->last_launch.det;
    const int block = net->last_launch.block;
    if (net->cus == 0) net->cus = device_cus();
    li.cus = net->cus;
    li.grid = net->last_launch.grid;
    li.block = block;
    char targs[32] = "";
    const char *name;
    if (slice) {
        name = alg == G4S_ALG_SKL ? "k_skl_slice" : "k_slice_step";
        slice_kernel_targs(net->d.L, block, det, targs, sizeof(targs));
        li.lds_bytes = (int32_t)slice_lds(net->d.L, block, det);
        li.resident_per_cu = alg == G4S_ALG_SKL ? skl_resident(net, net->d.L, block, det)
                                                : slice_step_blocks_per_cu(net->d.L, block, det);
        li.persistent = alg == G4S_ALG_SKL;
    } else {
        name = alg == G4S_ALG_AP ? "k_ap" : alg == G4S_ALG_N ? "k_push_rows" : "k_push_flat";
        li.persistent = alg == G4S_ALG_AP;
        if (alg == G4S_ALG_AP) li.resident_per_cu = persistent_blocks_per_cu(kPersistAP, block);
    }
    snprintf(li.kernel, sizeof(li.kernel), targs[0] ? "%s<%s>" : "%s", name, targs);
    if (const KernelRes *r = kernel_res(name, targs)) {
        li.vgprs = r->vgprs;
        li.sgprs = r->sgprs;
        li.waves_per_simd = r->occupancy;
        li.lds_bytes += r->lds;
    }
    return li;
}

// Capture range of a profiler run with --selected-regions, closed on every
// return path.
struct ProfileRange {
    ProfileRange() { roctxProfilerResume(0); }
    ~ProfileRange() { roctxProfilerPause(0); }
};

// Outputs of g4s_run_noalloc: device buffers of the caller.
struct NoAllocOut {
    uint32_t *raster;  // steps x words, or NULL
    int32_t *counts;   // steps, or NULL
};

// Restores the handle's own buffers after a g4s_run_noalloc call, whatever
// path the call leaves by.
struct ScratchSwap {
    g4s_net *net = nullptr;
    int32_t *ctl = nullptr, *bcnt = nullptr, *err = nullptr;
    uint32_t *bar = nullptr, *raster = nullptr;
    const int32_t *tinfo = nullptr;
    void take(g4s_net *n)
    {
        net = n;
        ctl = n->d.ctl; bcnt = n->d.bcnt; err = n->d.err; bar = n->d.bar; raster = n->d.raster; tinfo = n->d.tinfo;
    }
    ~ScratchSwap()
    {
        if (!net) return;
        net->d.ctl = ctl; net->d.bcnt = bcnt; net->d.err = err; net->d.bar = bar; net->d.raster = raster;
        net->d.tinfo = tinfo;
    }
};

int run_impl(g4s_net *net, int alg, int accum, int32_t steps, uint32_t *raster, int32_t *counts,
             const g4s_run_opts *opts, g4s_chunk_fn fn, void *user, const NoAllocOut *na = nullptr)
{
    if (!net) return fail(G4S_ERR_ARG, "net is NULL");
    if (alg < G4S_ALG_AP || alg > G4S_ALG_SKL) return fail(G4S_ERR_ARG, "unknown alg %d", alg);
    if (accum != G4S_ACCUM_ATOMIC && accum != G4S_ACCUM_DET) return fail(G4S_ERR_ARG, "unknown accum %d", accum);
    if (steps < 0) return fail(G4S_ERR_ARG, "steps < 0");
    if (net->t + (int64_t)steps > 0x7FFFFFFFLL) return fail(G4S_ERR_UNSUPPORTED, "step index overflow");
    g4s_run_opts o{};
    if (opts) o = *opts;
    if (o.block < 0 || o.block % 64 != 0 || o.block > 1024) return fail(G4S_ERR_ARG, "block must be a multiple of 64 <= 1024");
    const int block = o.block > 0 ? std::min(o.block, 256) : 256;  // launch bound of the other kernels
    int chunk = o.chunk > 0 ? o.chunk : 1000;
    int rc;
    if (na) {
        // nothing is allocated: the handle's staging (chunk_cap rows) and
        // slice partition must already exist
        if (o.kernel_events) return fail(G4S_ERR_ARG, "kernel_events needs event allocation: not in g4s_run_noalloc");
        if (net->chunk_cap < 1) return fail(G4S_ERR_STATE, "handle has no staging buffers");
        chunk = std::min(chunk, net->chunk_cap);
    } else if ((rc = ensure_chunk(net, chunk))) {
        return rc;
    }
    // AB and SKL run the slice kernels in both modes.  Det mode runs every
    // scheme through them: the ascending-source order of each input fixes a
    // gather by target, which the slices' ordered accumulation is; AP, N and
    // S then launch the slice step once per step like AB.
    const bool det = accum == G4S_ACCUM_DET;
    const bool slice = alg == G4S_ALG_AB || alg == G4S_ALG_SKL || det;
    const bool slice_steps = slice && alg != G4S_ALG_SKL;  // per-step slice launches
    int sblock = 0;
    if (slice && na) {
        // the partition the handle holds (that of its last g4s_run of a
        // slice scheme); building one would allocate
        if (!net->slices_ready)
            return fail(G4S_ERR_STATE, "no slice partition: run g4s_run once with this scheme and options first");
        sblock = slice_block(net, o.block, det);
    } else if (slice && o.slices > 0) {
        if ((rc = build_slices(net, o.slices))) return rc;
        sblock = slice_block(net, o.block, det);
    } else if (slice) {
        // default tiling: about one slice per CU; more slices while a
        // slice's workspace would not fit the LDS of one workgroup
        if (net->cus == 0) net->cus = device_cus();
        for (int64_t want = std::max(net->cus, 1);; want *= 2) {
            if ((rc = build_slices(net, (int)want))) return rc;
            sblock = slice_block(net, o.block, det);
            if (slice_lds(net->d.L, sblock, det) <= kSliceLdsMax || net->d.L <= 64) break;
        }
    }
    if (slice && (int64_t)net->d.L > 16 * (int64_t)sblock)
        return fail(G4S_ERR_ARG, "block %d too small for slices of %d neurons (max 16 per thread)", sblock, net->d.L);
    if (slice && slice_lds(net->d.L, sblock, det) > kSliceLdsMax)
        return fail(G4S_ERR_UNSUPPORTED, "slices of %d neurons at block %d need %zu bytes of LDS (max %zu)", net->d.L,
                    sblock, slice_lds(net->d.L, sblock, det), kSliceLdsMax);

    DevNet &d = net->d;
    Launch L{};
    if (alg == G4S_ALG_SKL) {
        // every workgroup must be resident: P slices within what the device
        // holds at once of the kernel instantiation that will run
        const int per_cu = skl_resident(net, d.L, sblock, det);
        const int cap = per_cu * std::max(net->cus, 1);
        if (per_cu <= 0 || cap < d.P)
            return fail(G4S_ERR_UNSUPPORTED, "SKL needs %d resident workgroups of %d threads, device holds %d", d.P,
                        sblock, cap);
        L = Launch{d.P, sblock, net->stream};
    }
    // The global-atomic schemes read the pending input from isyn[t&1]; after
    // a scheme that keeps none, rebuild it from the queue of step t-1.  This
    // must precede the AP sentinel fill below, which overwrites that queue.
    if (uses_isyn(alg, accum) && !net->isyn_valid && steps > 0) {
        const int t = (int)net->t;
        float *cur = (t & 1) ? d.isyn1 : d.isyn0;
        float *nxt = (t & 1) ? d.isyn0 : d.isyn1;
        HIP_OK(launch_fill_u32((uint32_t *)cur, (size_t)d.n, 0u, net->stream));
        HIP_OK(launch_fill_u32((uint32_t *)nxt, (size_t)d.n, 0u, net->stream));
        if (t > 0) HIP_OK(launch_push_bitmap(d, t - 1, net->stream));
    }
    if (steps > 0) net->isyn_valid = uses_isyn(alg, accum);
    // AP consumers wait on kSentinel slots: restore the invariant when AP
    // follows another scheme (those store queue slots without clearing).
    if (alg == G4S_ALG_AP && accum == G4S_ACCUM_ATOMIC && steps > 0) {
        HIP_OK(launch_fill_u32(d.q0, (size_t)d.n, kSentinel, net->stream));
        HIP_OK(launch_fill_u32(d.q1, (size_t)d.n, kSentinel, net->stream));
    }
    // Per-launch event brackets need separate launches: they replace the
    // graph replay of AB steps for this run.
    net->kev_on = o.kernel_events != 0;
    net->kev_used = 0;
    if (na)
        net->step_ms.clear();  // keeps its capacity: no host allocation either
    else
        net->step_ms.assign((size_t)steps, 0.0f);
    const bool use_graph = o.graph >= 0 && !net->kev_on && !na;
    // g4s_run_noalloc: the run's scratch buffers from the device's static
    // array, chunks no longer than it holds; the handle's own buffers are
    // restored on return
    ScratchSwap swap;
    if (na) {
        unsigned char *sc = nullptr;
        if ((rc = scratch_of_current_device(&sc))) return rc;
        const int P = slice ? d.P : 0;
        int cs_max = chunk;
        while (cs_max > 0 && scratch_layout(cs_max, P).end > (size_t)G4S_NOALLOC_SCRATCH_BYTES) cs_max /= 2;
        if (cs_max < 1)
            return fail(G4S_ERR_UNSUPPORTED, "%d slices need %zu bytes of scratch per step; the static scratch holds %d",
                        P, scratch_layout(1, P).end, G4S_NOALLOC_SCRATCH_BYTES);
        chunk = cs_max;
        const ScratchLayout l = scratch_layout(chunk, P);
        swap.take(net);
        d.bar = (uint32_t *)(sc + l.bar);
        d.err = (int32_t *)(sc + l.err);
        d.tinfo = (const int32_t *)(sc + l.tinfo);
        d.ctl = (int32_t *)(sc + l.ctl);
        d.bcnt = slice ? (int32_t *)(sc + l.bcnt) : nullptr;
        d.raster = nullptr;
        // the array keeps the last call's contents: the error flag starts clear
        HIP_OK(launch_fill_u32((uint32_t *)d.err, 1, 0u, net->stream));
    }

    // The step loop is the profiler's capture range (rocprofv3
    // --selected-regions), as cudaProfilerStart/Stop bracket the reference's
    // simulation (cu:1209, cu:1302): setup such as the connectivity build
    // stays out of the trace.
    ProfileRange prange;
    const auto wall0 = std::chrono::steady_clock::now();
    double dev_ms = 0.0, kern_ms = 0.0;
    int64_t spikes = 0;
    int launches = 0, dom_grid = 0, kern_launches = 0;
    const int64_t t_end = net->t + steps;
    const int64_t words = d.words;
    // A streaming run hands chunk c to the callback while chunk c+1 runs.
    // AP sizes a chunk's first step from the previous chunk's last count,
    // and event-bracketed runs collect per chunk: those run one at a time.
    const bool overlap = fn && !net->kev_on && !(alg == G4S_ALG_AP && accum == G4S_ACCUM_ATOMIC);
    struct Pending { int64_t c0; int cs; int t0; int buf; };
    Pending pend[2];
    int npend = 0;

    // A chunk's device time: for SKL, one launch whose workgroup 0 stamps
    // each step's start, the stamps (first step to the chunk's end stamp)
    // instead of event markers around the launch, which cost each run a few
    // microseconds; other schemes and runs with kernel events use the markers.
    const bool stamp_time = alg == G4S_ALG_SKL && !net->kev_on && net->clock_khz > 0;
    auto ctl_stamp = [](const int32_t *rows, int r) {
        return (uint64_t)(uint32_t)rows[r * kCtlStride + kCtlStamp] |
               ((uint64_t)(uint32_t)rows[r * kCtlStride + kCtlStamp + 1] << 32);
    };
    // consume the oldest enqueued chunk
    auto finish = [&](const Pending &pc) -> int {
        HIP_OK(hipEventSynchronize(net->done_ev[pc.buf]));
        if (net->h_ctl[pc.buf][0]) return fail(G4S_ERR_TIMEOUT, "a bounded device wait expired at chunk starting step %d", pc.t0);
        float ms = 0.0f;
        if (stamp_time) {  // the chunk's first step start to its end stamp
            ms = (float)((double)(int64_t)(ctl_stamp(net->h_ctl[pc.buf] + kCtlHead, pc.cs) -
                                           ctl_stamp(net->h_ctl[pc.buf] + kCtlHead, 0)) /
                         net->clock_khz);
        } else {
            HIP_OK(hipEventElapsedTime(&ms, net->ev0[pc.buf], net->ev1[pc.buf]));
        }
        dev_ms += ms;
        if (net->kev_on) {
            kern_launches += net->kev_used / 2;
            HIP_OK(kev_collect(net, &kern_ms));
        }
        const int32_t *hc = net->h_ctl[pc.buf] + kCtlHead;
        int64_t mx = 0;
        auto stamp = [&](int r) { return ctl_stamp(hc, r); };
        for (int r = 0; r < pc.cs; ++r) {
            const int32_t cnt = hc[r * kCtlStride + kCtlCount];
            spikes += cnt;
            mx = std::max<int64_t>(mx, cnt);
            if (counts) counts[pc.c0 + r] = cnt;
            if ((size_t)(pc.c0 + r) < net->step_ms.size() && net->clock_khz > 0)
                net->step_ms[(size_t)(pc.c0 + r)] = (float)((double)(int64_t)(stamp(r + 1) - stamp(r)) / net->clock_khz);
        }
        net->last_count = hc[(pc.cs - 1) * kCtlStride + kCtlCount];
        net->est_count = std::max<int64_t>(64, mx + mx / 4);
        if (raster) memcpy(raster + (size_t)pc.c0 * words, net->h_raster[pc.buf], (size_t)pc.cs * words * sizeof(uint32_t));
        if (fn) {
            std::vector<int32_t> cc((size_t)pc.cs);
            for (int r = 0; r < pc.cs; ++r) cc[(size_t)r] = hc[r * kCtlStride + kCtlCount];
            const int urc = fn(user, pc.t0, pc.cs, net->h_raster[pc.buf], cc.data());
            if (urc) return fail(G4S_ERR_STATE, "stream callback returned %d at step %d", urc, pc.t0);
        }
        return G4S_OK;
    };
    // drain what is still in flight when the run stops early
    auto abandon = [&](int code) {
        for (int k = 0; k < npend; ++k) (void)hipEventSynchronize(net->done_ev[pend[k].buf]);
        return code;
    };

    int chunk_no = 0;
    for (int64_t c0 = 0; c0 < steps; c0 += chunk, ++chunk_no) {
        const int cs = (int)std::min<int64_t>(chunk, steps - c0);
        const int t0 = (int)(net->t);
        const int buf = chunk_no & 1;
        if (slice_steps && use_graph && (rc = prepare_ab_graph(net, cs, sblock, det))) return abandon(rc);
        // control rows and barrier words start at zero: cleared here only if
        // the previous chunk did not leave them so (and always for the
        // caller-independent scratch of g4s_run_noalloc)
        if (na || !net->ctl_clean)
            HIP_OK(launch_fill2_u32((uint32_t *)d.ctl, (size_t)(cs + 1) * kCtlStride, d.bar, kBarWords, 0u, net->stream));
        if (!na) net->ctl_clean = false;
        if (na) d.raster = na->raster ? na->raster + (size_t)c0 * (size_t)words : nullptr;
        if (slice_steps) {
            int32_t *ti = net->h_tinfo + 2 * buf;
            ti[0] = t0;
            ti[1] = (int32_t)(t_end - 1);
            HIP_OK(hipMemcpyAsync((void *)d.tinfo, ti, 2 * sizeof(int32_t), hipMemcpyHostToDevice, net->stream));
        }
        if (!stamp_time) HIP_OK(hipEventRecord(net->ev0[buf], net->stream));
        if (slice_steps) {
            if ((rc = run_ab_chunk(net, cs, use_graph, sblock, det))) return abandon(rc);
            launches += cs;
            dom_grid = d.P;
        } else if (alg == G4S_ALG_SKL) {
            Launch Lk = L;
            HIP_OK(kev_stamp(net, &Lk));
            const hipError_t e = launch_skl_slice(d, t0, cs, (int)(t_end - 1), det, Lk);
            if (e != hipSuccess) return abandon(fail(G4S_ERR_HIP, "SKL launch (grid %d): %s", L.grid, hipGetErrorString(e)));
            launches += 1;
            dom_grid = L.grid;
        } else {
            for (int rel = 0; rel < cs; ++rel) {
                const int t = t0 + rel;
                hipError_t e = launch_step(net, alg, block, t, rel, t == t_end - 1, &launches, &L, &rc);
                if (rc) return abandon(rc);
                if (e != hipSuccess) return abandon(fail(G4S_ERR_HIP, "step %d launch (alg %d, grid %d): %s", t, alg, L.grid, hipGetErrorString(e)));
                dom_grid = L.grid;
            }
        }
        if (!stamp_time) HIP_OK(hipEventRecord(net->ev1[buf], net->stream));
        // one kernel: the chunk's end stamp, the step counts (summed over the
        // slices for slice schemes), the error flag and control rows written
        // straight to the host buffer, and the caller's counts.  Without a
        // raster copy after it, its dispatch signals the chunk's done event.
        const bool copy_raster = raster || fn;
        HIP_OK(launch_finish_chunk(d, cs, slice, net->h_ctl_dev[buf], na && na->counts ? na->counts + c0 : nullptr,
                                   copy_raster ? nullptr : net->done_ev[buf], net->stream));
        if (copy_raster) {
            HIP_OK(hipMemcpyAsync(net->h_raster[buf], d.raster, (size_t)cs * words * sizeof(uint32_t), hipMemcpyDeviceToHost, net->stream));
            HIP_OK(hipEventRecord(net->done_ev[buf], net->stream));
        }
        if (!na) {
            // clear for the next chunk, after this one's copies in stream
            // order: the host's wait for done_ev does not include it
            HIP_OK(launch_fill2_u32((uint32_t *)d.ctl, (size_t)(net->chunk_cap + 1) * kCtlStride, d.bar, kBarWords, 0u,
                                    net->stream));
            net->ctl_clean = true;
        }
        net->t += cs;
        pend[npend++] = Pending{c0, cs, t0, buf};
        if (!overlap || npend == 2) {
            rc = finish(pend[0]);
            pend[0] = pend[1];
            npend -= 1;
            if (rc) return abandon(rc);
        }
    }
    while (npend > 0) {
        rc = finish(pend[0]);
        pend[0] = pend[1];
        npend -= 1;
        if (rc) return abandon(rc);
    }
    net->timing.device_ms = dev_ms;
    net->timing.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    net->timing.steps = steps;
    net->timing.spikes = spikes;
    net->timing.launches = launches;
    net->timing.grid = dom_grid;
    net->timing.kernel_ms = kern_ms;
    net->timing.kernel_launches = kern_launches;
    net->kev_on = false;
    record_launch(net, alg, slice, slice ? sblock : block, dom_grid, det);
    return G4S_OK;
}

}  // namespace

int g4s_run(g4s_net *net, int alg, int accum, int32_t steps, uint32_t *raster, int32_t *counts,
            const g4s_run_opts *opts)
{
    return run_impl(net, alg, accum, steps, raster, counts, opts, nullptr, nullptr);
}

int g4s_run_noalloc(g4s_net *net, int alg, int accum, int32_t steps, uint32_t *raster_dev, int32_t *counts_dev,
                    const g4s_run_opts *opts)
{
    const NoAllocOut na{raster_dev, counts_dev};
    return run_impl(net, alg, accum, steps, nullptr, nullptr, opts, nullptr, nullptr, &na);
}

int g4s_run_stream(g4s_net *net, int alg, int accum, int32_t steps, g4s_chunk_fn fn, void *user,
                   const g4s_run_opts *opts)
{
    if (!fn) return fail(G4S_ERR_ARG, "callback is NULL");
    return run_impl(net, alg, accum, steps, nullptr, nullptr, opts, fn, user);
}

int g4s_set_regime(g4s_net *net, int regime, float thal_scale)
{
    if (!net) return fail(G4S_ERR_ARG, "net is NULL");
    g4s_spec s = net->spec;
    s.regime = regime;
    s.thal_scale = thal_scale;
    int rc = validate_spec(&s);
    if (rc) return rc;
    net->spec = s;
    regime_gains(regime, thal_scale, &net->d.ge, &net->d.gi);
    net->gen += 1;
    return G4S_OK;
}

int g4s_get_timing(const g4s_net *net, g4s_timing *out)
{
    if (!net || !out) return fail(G4S_ERR_ARG, "NULL argument");
    *out = net->timing;
    return G4S_OK;
}

int g4s_get_step_times(const g4s_net *net, float *ms, int32_t n)
{
    if (!net || (!ms && n > 0) || n < 0) return fail(G4S_ERR_ARG, "bad arguments");
    const size_t k = std::min<size_t>((size_t)n, net->step_ms.size());
    if (k) memcpy(ms, net->step_ms.data(), k * sizeof(float));
    return (int)k;
}

int g4s_get_launch_info(g4s_net *net, g4s_launch_info *out)
{
    if (!net || !out) return fail(G4S_ERR_ARG, "NULL argument");
    *out = launch_report(net);
    return G4S_OK;
}

int g4s_kernel_resources(const char *name, const char *targs, int32_t out[6])
{
    if (!name || !out) return fail(G4S_ERR_ARG, "NULL argument");
    const KernelRes *r = kernel_res(name, targs ? targs : "");
    if (!r) return fail(G4S_ERR_ARG, "no kernel %s<%s> in this build", name, targs ? targs : "");
    out[0] = r->occupancy;
    out[1] = r->vgprs;
    out[2] = r->agprs;
    out[3] = r->sgprs;
    out[4] = r->lds;
    out[5] = r->scratch;
    return G4S_OK;
}

int g4s_get_state(g4s_net *net, float *v, float *u, float *I, int32_t *nospks)
{
    if (!net) return fail(G4S_ERR_ARG, "net is NULL");
    const size_t n = (size_t)net->d.n;
    if (v) HIP_OK(hipMemcpy(v, net->d.v, n * sizeof(float), hipMemcpyDeviceToHost));
    if (u) HIP_OK(hipMemcpy(u, net->d.u, n * sizeof(float), hipMemcpyDeviceToHost));
    if (I) HIP_OK(hipMemcpy(I, net->d.lastI, n * sizeof(float), hipMemcpyDeviceToHost));
    if (nospks) HIP_OK(hipMemcpy(nospks, net->d.nospks, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    return G4S_OK;
}

int g4s_get_params(g4s_net *net, float *a, float *b, float *c, float *d)
{
    if (!net) return fail(G4S_ERR_ARG, "net is NULL");
    const size_t n = (size_t)net->d.n;
    if (a) HIP_OK(hipMemcpy(a, net->d.a, n * sizeof(float), hipMemcpyDeviceToHost));
    if (b) HIP_OK(hipMemcpy(b, net->d.b, n * sizeof(float), hipMemcpyDeviceToHost));
    if (c) HIP_OK(hipMemcpy(c, net->d.c, n * sizeof(float), hipMemcpyDeviceToHost));
    if (d) HIP_OK(hipMemcpy(d, net->d.d, n * sizeof(float), hipMemcpyDeviceToHost));
    return G4S_OK;
}

int g4s_set_state(g4s_net *net, const float *v, const float *u)
{
    if (!net) return fail(G4S_ERR_ARG, "net is NULL");
    const size_t n = (size_t)net->d.n;
    if (v) HIP_OK(hipMemcpy(net->d.v, v, n * sizeof(float), hipMemcpyHostToDevice));
    if (u) HIP_OK(hipMemcpy(net->d.u, u, n * sizeof(float), hipMemcpyHostToDevice));
    return G4S_OK;
}

int g4s_get_synapses(g4s_net *net, uint32_t *post, float *w)
{
    if (!net) return fail(G4S_ERR_ARG, "net is NULL");
    const size_t ns = (size_t)net->d.n * (size_t)net->d.nsyn;
    std::vector<uint2> tab(ns);
    HIP_OK(hipMemcpy(tab.data(), net->d.syn, ns * sizeof(uint2), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < ns; ++k) {
        if (post) post[k] = tab[k].x;
        if (w) w[k] = u2f(tab[k].y);
    }
    return G4S_OK;
}

int g4s_destroy(g4s_net *net)
{
    if (!net) return G4S_OK;
    if (net->stream) (void)hipStreamSynchronize(net->stream);
    drop_graph(net);
    for (void *p : net->allocs) (void)hipFree(p);
    if (net->h_tinfo) (void)hipHostFree(net->h_tinfo);
    for (int k = 0; k < 2; ++k) {
        if (net->h_raster[k]) (void)hipHostFree(net->h_raster[k]);
        if (net->h_ctl[k]) (void)hipHostFree(net->h_ctl[k]);
    }
    for (hipEvent_t e : net->kev) (void)hipEventDestroy(e);
    for (int k = 0; k < 2; ++k) {
        if (net->ev0[k]) (void)hipEventDestroy(net->ev0[k]);
        if (net->ev1[k]) (void)hipEventDestroy(net->ev1[k]);
        if (net->done_ev[k]) (void)hipEventDestroy(net->done_ev[k]);
    }
    if (net->stream) (void)hipStreamDestroy(net->stream);
    delete net;
    return G4S_OK;
}

int g4s_build_synapses_device(const g4s_spec *spec, uint32_t *post, float *w)
{
    int rc = validate_spec(spec);
    if (rc) return rc;
    if (!post || !w) return fail(G4S_ERR_ARG, "NULL output");
    if (spec->nsyn > kDeviceBuildMaxNsyn) return fail(G4S_ERR_UNSUPPORTED, "device build needs Nsyn <= %d", kDeviceBuildMaxNsyn);
    const size_t ns = (size_t)spec->n * (size_t)spec->nsyn;
    uint2 *d = nullptr;
    if (hipMalloc((void **)&d, ns * sizeof(uint2)) != hipSuccess) return fail(G4S_ERR_MEM, "hipMalloc(%zu)", ns * 8);
    rc = build_synapses(spec, d, nullptr);
    std::vector<uint2> tab;
    if (rc == G4S_OK) {
        tab.resize(ns);
        if (hipMemcpy(tab.data(), d, ns * sizeof(uint2), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(G4S_ERR_HIP, "copy of the device table failed");
    }
    (void)hipFree(d);
    if (rc) return rc;
    for (size_t k = 0; k < ns; ++k) {
        post[k] = tab[k].x;
        w[k] = u2f(tab[k].y);
    }
    return G4S_OK;
}

int g4s_device_thalamic(uint64_t seed, uint32_t t, int64_t n, float *out)
{
    if (n <= 0 || !out) return fail(G4S_ERR_ARG, "bad arguments");
    float *dout = nullptr;
    HIP_OK(hipMalloc((void **)&dout, (size_t)n * sizeof(float)));
    hipError_t e = launch_thalamic(seed, t, n, dout, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, dout, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(dout);
    if (e != hipSuccess) return fail(G4S_ERR_HIP, "device thalamic: %s", hipGetErrorString(e));
    return G4S_OK;
}

int g4s_device_normal_words(const uint32_t *x0, const uint32_t *x1, int64_t n, float *out)
{
    if (n <= 0 || !x0 || !x1 || !out) return fail(G4S_ERR_ARG, "bad arguments");
    uint32_t *d0 = nullptr, *d1 = nullptr;
    float *dout = nullptr;
    const size_t b = (size_t)n * 4;
    hipError_t e = hipMalloc((void **)&d0, b);
    if (e == hipSuccess) e = hipMalloc((void **)&d1, b);
    if (e == hipSuccess) e = hipMalloc((void **)&dout, b);
    if (e == hipSuccess) e = hipMemcpy(d0, x0, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d1, x1, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_normal_words(d0, d1, n, dout, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, dout, b, hipMemcpyDeviceToHost);
    for (void *p : {(void *)d0, (void *)d1, (void *)dout})
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return fail(G4S_ERR_HIP, "device normals: %s", hipGetErrorString(e));
    return G4S_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Autotuner: scheme x tiling candidates timed on the live network, cached
// per (log2 N, Nsyn, log2 rate, accum) bucket.
// ---------------------------------------------------------------------------
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

namespace {

using Bucket = std::tuple<int, int, int, int>;
std::mutex g_tune_mu;
std::map<Bucket, g4s_plan> g_tune_cache;

int ilog2(double x) { return x <= 0 ? -100 : (int)std::floor(std::log2(x)); }

// Time `plan` from a reset network: warm-up then `steps` steps.
int time_plan(g4s_net *net, g4s_plan *plan, int steps)
{
    int rc = g4s_reset(net);
    if (rc) return rc;
    const int warm = std::max(10, steps / 4);
    if ((rc = g4s_run(net, plan->alg, plan->accum, warm, nullptr, nullptr, &plan->opts))) return rc;
    if ((rc = g4s_run(net, plan->alg, plan->accum, steps, nullptr, nullptr, &plan->opts))) return rc;
    plan->us_per_step = (float)(net->timing.device_ms * 1000.0 / steps);
    plan->rate_hz = (float)((double)net->timing.spikes / steps / (double)net->d.n * 1000.0);
    return G4S_OK;
}

}  // namespace

extern "C" int g4s_autotune(g4s_net *net, int accum, int32_t trial_steps, int force, g4s_plan *best)
{
    if (!net || !best) return fail(G4S_ERR_ARG, "NULL argument");
    if (accum != G4S_ACCUM_ATOMIC && accum != G4S_ACCUM_DET) return fail(G4S_ERR_ARG, "unknown accum %d", accum);
    if (trial_steps < 1) trial_steps = 100;
    // probe the firing rate with a default plan to find the bucket
    g4s_plan probe{};
    probe.alg = G4S_ALG_AB;
    probe.accum = accum;
    int rc = time_plan(net, &probe, std::max(20, trial_steps / 2));
    if (rc) return rc;
    const Bucket key{ilog2((double)net->d.n), net->d.nsyn, ilog2(probe.rate_hz), accum};
    if (!force) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            *best = it->second;
            return g4s_reset(net);
        }
    }
    std::vector<g4s_plan> cands;
    for (int alg = G4S_ALG_AP; alg <= G4S_ALG_SKL; ++alg) {
        const bool slice = alg == G4S_ALG_AB || alg == G4S_ALG_SKL;
        if (slice) {
            for (int P : {64, 128, 192, 256, 512})
                for (int B : {256, 512, 1024}) {
                    g4s_plan c{};
                    c.alg = alg; c.accum = accum; c.opts.slices = P; c.opts.block = B;
                    cands.push_back(c);
                }
        } else {
            for (int B : {128, 256}) {
                g4s_plan c{};
                c.alg = alg; c.accum = accum; c.opts.block = B;
                cands.push_back(c);
            }
        }
    }
    g4s_plan win{};
    win.us_per_step = INFINITY;
    for (g4s_plan &c : cands) {
        rc = time_plan(net, &c, trial_steps);
        if (rc == G4S_ERR_UNSUPPORTED || rc == G4S_ERR_ARG) continue;  // tiling not valid here
        if (rc) return rc;
        if (c.us_per_step < win.us_per_step) win = c;
    }
    if (!(win.us_per_step < INFINITY)) return fail(G4S_ERR_UNSUPPORTED, "no candidate plan ran");
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = win;
    }
    *best = win;
    return g4s_reset(net);
}
