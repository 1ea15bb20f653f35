// g4s_runtime.hip -- host runtime and C ABI (include/gpu4snn.h).
//
// Replaces the reference's main() (gpu4snn.cu:735-1332): allocation, the
// connectivity build, initNeuron, the per-step launch loop with its
// device->host spike copy, and the SKL single launch.  Differences by design:
// steps run in device-resident chunks (no per-step host round trip); the
// raster is a per-step bitmap copied once per chunk; timing uses HIP events.
#include "../../include/gpu4snn.h"
#include "g4s_build.h"
#include "g4s_internal.h"

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace g4s;

namespace {

thread_local std::string g_err;

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

#define HIP_OK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return fail(G4S_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

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
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int64_t t = 0;
    int chunk_cap = 0;          // raster rows / ctl entries allocated
    uint32_t *h_raster = nullptr;  // pinned staging
    int32_t *h_ctl = nullptr;      // pinned staging
    bool det_ready = false;
    int last_accum = G4S_ACCUM_ATOMIC;  // mode of the previous step
    int64_t est_count = 0;      // recent max spikes per step (grid sizing)
    g4s_timing timing{};
    std::vector<void *> allocs;

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

int ensure_chunk(g4s_net *net, int chunk)
{
    if (chunk <= net->chunk_cap) return G4S_OK;
    if (net->d.raster) { hipFree(net->d.raster); net->allocs.erase(std::find(net->allocs.begin(), net->allocs.end(), (void *)net->d.raster)); }
    if (net->d.ctl) { hipFree(net->d.ctl); net->allocs.erase(std::find(net->allocs.begin(), net->allocs.end(), (void *)net->d.ctl)); }
    if (net->h_raster) hipHostFree(net->h_raster);
    if (net->h_ctl) hipHostFree(net->h_ctl);
    net->d.raster = nullptr; net->d.ctl = nullptr; net->h_raster = nullptr; net->h_ctl = nullptr;
    int rc;
    if ((rc = net->alloc(&net->d.raster, (size_t)chunk * net->d.words))) return rc;
    if ((rc = net->alloc(&net->d.ctl, (size_t)chunk * kCtlStride))) return rc;
    HIP_OK(hipHostMalloc((void **)&net->h_raster, (size_t)chunk * net->d.words * sizeof(uint32_t)));
    HIP_OK(hipHostMalloc((void **)&net->h_ctl, (size_t)chunk * kCtlStride * sizeof(int32_t)));
    net->chunk_cap = chunk;
    return G4S_OK;
}

int build_det_tables(g4s_net *net)
{
    if (net->det_ready) return G4S_OK;
    const int64_t n = net->d.n;
    const size_t ns = (size_t)n * (size_t)net->d.nsyn;
    std::vector<uint2> tab(ns), in_tab(ns);
    std::vector<int64_t> in_ptr((size_t)n + 1);
    HIP_OK(hipMemcpy(tab.data(), net->d.syn, ns * sizeof(uint2), hipMemcpyDeviceToHost));
    build_incoming(n, net->d.nsyn, tab.data(), in_ptr.data(), in_tab.data(), 0);
    int64_t *dp = nullptr;
    uint2 *dt = nullptr;
    int rc;
    if ((rc = net->alloc(&dp, (size_t)n + 1))) return rc;
    if ((rc = net->alloc(&dt, ns))) return rc;
    HIP_OK(hipMemcpy(dp, in_ptr.data(), ((size_t)n + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dt, in_tab.data(), ns * sizeof(uint2), hipMemcpyHostToDevice));
    net->d.in_ptr = dp;
    net->d.in_syn = dt;
    net->det_ready = true;
    return G4S_OK;
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

Launch plain(g4s_net *net, int64_t threads, int block)
{
    int64_t g = (threads + block - 1) / block;
    g = std::max<int64_t>(1, std::min<int64_t>(g, 1 << 20));
    return Launch{(int)g, block, net->stream};
}

int coop_launch(g4s_net *net, CoopKernel k, int block, int64_t useful_threads, Launch *L)
{
    int maxb = coop_max_blocks(k, block);
    if (maxb <= 0) return fail(G4S_ERR_HIP, "no co-resident capacity for cooperative kernel %d", (int)k);
    int64_t want = (useful_threads + block - 1) / block;
    int g = (int)std::max<int64_t>(1, std::min<int64_t>(maxb, want));
    *L = Launch{g, block, net->stream};
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
    if (hipEventCreate(&net->ev0) != hipSuccess || hipEventCreate(&net->ev1) != hipSuccess)
        return bail(fail(G4S_ERR_HIP, "hipEventCreate failed"));
    const size_t n = (size_t)spec->n;
    float **fields[] = {&d.v, &d.u, &d.a, &d.b, &d.c, &d.d, &d.lastI, &d.isyn0, &d.isyn1};
    for (float **f : fields)
        if ((rc = net->alloc(f, n))) return bail(rc);
    if ((rc = net->alloc(&d.nospks, n))) return bail(rc);
    if ((rc = net->alloc(&d.bm0, (size_t)d.words))) return bail(rc);
    if ((rc = net->alloc(&d.bm1, (size_t)d.words))) return bail(rc);
    if ((rc = net->alloc(&d.q0, n))) return bail(rc);
    if ((rc = net->alloc(&d.q1, n))) return bail(rc);
    if ((rc = net->alloc(&d.bar, 1))) return bail(rc);
    if ((rc = net->alloc(&d.err, 1))) return bail(rc);
    uint2 *syn = nullptr;
    const size_t ns = n * (size_t)spec->nsyn;
    if ((rc = net->alloc(&syn, ns))) return bail(rc);
    d.syn = syn;
    {
        std::vector<uint2> tab(ns);
        build_table(spec->seed, spec->n, d.n_exc, spec->nsyn, tab.data(), 0);
        if (hipMemcpy(syn, tab.data(), ns * sizeof(uint2), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(G4S_ERR_HIP, "upload of the synapse table failed"));
    }
    if ((rc = ensure_chunk(net, 1000))) return bail(rc);
    if ((rc = g4s_reset(net))) return bail(rc);
    *out = net;
    return G4S_OK;
}

int g4s_reset(g4s_net *net)
{
    if (!net) return fail(G4S_ERR_ARG, "net is NULL");
    HIP_OK(launch_init(net->d, net->stream));
    HIP_OK(launch_fill_u32((uint32_t *)net->d.err, 1, 0u, net->stream));
    HIP_OK(hipStreamSynchronize(net->stream));
    net->t = 0;
    net->last_accum = G4S_ACCUM_ATOMIC;
    net->est_count = std::max<int64_t>(64, net->spec.n / 16);
    return G4S_OK;
}

int64_t g4s_current_step(const g4s_net *net) { return net ? net->t : -1; }

int g4s_run(g4s_net *net, int alg, int accum, int32_t steps, uint32_t *raster, int32_t *counts,
            const g4s_run_opts *opts)
{
    if (!net) return fail(G4S_ERR_ARG, "net is NULL");
    if (alg < G4S_ALG_AP || alg > G4S_ALG_SKL) return fail(G4S_ERR_ARG, "unknown alg %d", alg);
    if (accum != G4S_ACCUM_ATOMIC && accum != G4S_ACCUM_DET) return fail(G4S_ERR_ARG, "unknown accum %d", accum);
    if (steps < 0) return fail(G4S_ERR_ARG, "steps < 0");
    if (net->t + (int64_t)steps > 0x7FFFFFFFLL) return fail(G4S_ERR_UNSUPPORTED, "step index overflow");
    g4s_run_opts o{};
    if (opts) o = *opts;
    const int block = o.block > 0 ? o.block : 256;
    if (block % 64 != 0 || block > 256) return fail(G4S_ERR_ARG, "block must be a multiple of 64 <= 256");
    const int chunk = o.chunk > 0 ? o.chunk : 1000;
    int rc;
    if ((rc = ensure_chunk(net, chunk))) return rc;
    if (accum == G4S_ACCUM_DET && (rc = build_det_tables(net))) return rc;

    DevNet &d = net->d;
    // AP consumers wait on kSentinel slots: restore the invariant when AP
    // follows another scheme (those store queue slots without clearing).
    if (alg == G4S_ALG_AP && accum == G4S_ACCUM_ATOMIC) {
        HIP_OK(launch_fill_u32(d.q0, (size_t)d.n, kSentinel, net->stream));
        HIP_OK(launch_fill_u32(d.q1, (size_t)d.n, kSentinel, net->stream));
    }
    // det mode keeps no pending-input buffers; rebuild them when switching
    // back to atomic mode in the middle of a simulation.
    if (accum == G4S_ACCUM_ATOMIC && net->last_accum == G4S_ACCUM_DET && net->t > 0)
        HIP_OK(launch_gather_isyn(d, (int)net->t, net->stream));
    if (steps > 0) net->last_accum = accum;
    const auto wall0 = std::chrono::steady_clock::now();
    double dev_ms = 0.0;
    int64_t spikes = 0;
    int launches = 0, dom_grid = 0;
    const int64_t t_end = net->t + steps;
    const int64_t words = d.words;

    for (int64_t c0 = 0; c0 < steps; c0 += chunk) {
        const int cs = (int)std::min<int64_t>(chunk, steps - c0);
        const int t0 = (int)(net->t);
        HIP_OK(launch_fill_u32((uint32_t *)d.ctl, (size_t)cs * kCtlStride, 0u, net->stream));
        HIP_OK(launch_fill_u32(d.bar, 1, 0u, net->stream));
        HIP_OK(hipEventRecord(net->ev0, net->stream));
        Launch L{};
        if (alg == G4S_ALG_SKL) {
            const CoopKernel k = accum == G4S_ACCUM_DET ? kCoopSKLDet : kCoopSKL;
            if ((rc = coop_launch(net, k, block, std::max<int64_t>(d.n, (int64_t)net->est_count * d.nsyn), &L))) return rc;
            hipError_t e = accum == G4S_ACCUM_DET ? launch_skl_det(d, t0, cs, (int)(t_end - 1), L)
                                                  : launch_skl(d, t0, cs, (int)(t_end - 1), L);
            if (e != hipSuccess) return fail(G4S_ERR_HIP, "SKL cooperative launch (grid %d): %s", L.grid, hipGetErrorString(e));
            launches += 1;
            dom_grid = L.grid;
        } else {
            for (int rel = 0; rel < cs; ++rel) {
                const int t = t0 + rel;
                const int last = (t == t_end - 1);
                hipError_t e = hipSuccess;
                if (accum == G4S_ACCUM_DET) {
                    switch (alg) {
                    case G4S_ALG_AP:
                        if ((rc = coop_launch(net, kCoopAPDet, block, d.n, &L))) return rc;
                        e = launch_ap_det(d, t, rel, last, L);
                        break;
                    case G4S_ALG_N:
                        L = plain(net, d.n, block);
                        e = launch_pull_thread(d, t, rel, last, L, false);
                        break;
                    case G4S_ALG_S:
                        L = plain(net, ((int64_t)d.n + 63) / 64 * 64, block);
                        e = launch_pull_wave(d, t, rel, last, L);
                        break;
                    case G4S_ALG_AB:
                        if ((rc = coop_launch(net, kCoopPull, block, d.n, &L))) return rc;
                        e = launch_pull_thread(d, t, rel, last, L, true);
                        break;
                    }
                    launches += 1;
                } else {
                    if (alg == G4S_ALG_AP) {
                        if ((rc = coop_launch(net, kCoopAP, block, d.n, &L))) return rc;
                        e = launch_ap(d, t, rel, last, L);
                        launches += 1;
                    } else {
                        e = launch_update(d, t, rel, last, plain(net, d.n, block));
                        if (e == hipSuccess) {
                            if (alg == G4S_ALG_N) {
                                L = plain(net, net->est_count * 64, block);
                                e = launch_push_rows(d, t, rel, L);
                            } else if (alg == G4S_ALG_S) {
                                L = plain(net, net->est_count * (int64_t)d.nsyn, block);
                                e = launch_push_flat(d, t, rel, L, false);
                            } else {
                                if ((rc = coop_launch(net, kCoopPushFlat, block, (int64_t)d.n * d.nsyn, &L))) return rc;
                                e = launch_push_flat(d, t, rel, L, true);
                            }
                        }
                        launches += 2;
                    }
                }
                if (e != hipSuccess) return fail(G4S_ERR_HIP, "step %d launch (alg %d, grid %d): %s", t, alg, L.grid, hipGetErrorString(e));
                dom_grid = L.grid;
            }
        }
        HIP_OK(hipEventRecord(net->ev1, net->stream));
        HIP_OK(hipMemcpyAsync(net->h_ctl, d.ctl, (size_t)cs * kCtlStride * sizeof(int32_t), hipMemcpyDeviceToHost, net->stream));
        if (raster)
            HIP_OK(hipMemcpyAsync(net->h_raster, d.raster, (size_t)cs * words * sizeof(uint32_t), hipMemcpyDeviceToHost, net->stream));
        int32_t err = 0;
        HIP_OK(hipMemcpyAsync(&err, d.err, sizeof(int32_t), hipMemcpyDeviceToHost, net->stream));
        HIP_OK(hipStreamSynchronize(net->stream));
        if (err) return fail(G4S_ERR_TIMEOUT, "a bounded device wait expired at chunk starting step %d", t0);
        float ms = 0.0f;
        HIP_OK(hipEventElapsedTime(&ms, net->ev0, net->ev1));
        dev_ms += ms;
        int64_t mx = 0;
        for (int r = 0; r < cs; ++r) {
            const int32_t cnt = net->h_ctl[r * kCtlStride + kCtlCount];
            spikes += cnt;
            mx = std::max<int64_t>(mx, cnt);
            if (counts) counts[c0 + r] = cnt;
        }
        net->est_count = std::max<int64_t>(64, mx + mx / 4);
        if (raster) memcpy(raster + (size_t)c0 * words, net->h_raster, (size_t)cs * words * sizeof(uint32_t));
        net->t += cs;
    }
    net->timing.device_ms = dev_ms;
    net->timing.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    net->timing.steps = steps;
    net->timing.spikes = spikes;
    net->timing.launches = launches;
    net->timing.grid = dom_grid;
    return G4S_OK;
}

int g4s_get_timing(const g4s_net *net, g4s_timing *out)
{
    if (!net || !out) return fail(G4S_ERR_ARG, "NULL argument");
    *out = net->timing;
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
    if (net->stream) hipStreamSynchronize(net->stream);
    for (void *p : net->allocs) hipFree(p);
    if (net->h_raster) hipHostFree(net->h_raster);
    if (net->h_ctl) hipHostFree(net->h_ctl);
    if (net->ev0) hipEventDestroy(net->ev0);
    if (net->ev1) hipEventDestroy(net->ev1);
    if (net->stream) hipStreamDestroy(net->stream);
    delete net;
    return G4S_OK;
}

int g4s_device_thalamic(uint64_t seed, uint32_t t, int64_t n, float *out)
{
    if (n <= 0 || !out) return fail(G4S_ERR_ARG, "bad arguments");
    float *dout = nullptr;
    HIP_OK(hipMalloc((void **)&dout, (size_t)n * sizeof(float)));
    hipError_t e = launch_thalamic(seed, t, n, dout, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, dout, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
    hipFree(dout);
    if (e != hipSuccess) return fail(G4S_ERR_HIP, "device thalamic: %s", hipGetErrorString(e));
    return G4S_OK;
}

int g4s_device_normal_words(const uint32_t *x0, const uint32_t *x1, int64_t n, float *out)
{
    if (n <= 0 || !x0 || !x1 || !out) return fail(G4S_ERR_ARG, "bad arguments");
    uint32_t *d0 = nullptr, *d1 = nullptr;
    float *dout = nullptr;
    const size_t b = (size_t)n * 4;
    HIP_OK(hipMalloc((void **)&d0, b));
    HIP_OK(hipMalloc((void **)&d1, b));
    HIP_OK(hipMalloc((void **)&dout, b));
    hipError_t e = hipMemcpy(d0, x0, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d1, x1, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_normal_words(d0, d1, n, dout, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, dout, b, hipMemcpyDeviceToHost);
    hipFree(d0);
    hipFree(d1);
    hipFree(dout);
    if (e != hipSuccess) return fail(G4S_ERR_HIP, "device normals: %s", hipGetErrorString(e));
    return G4S_OK;
}

}  // extern "C"
