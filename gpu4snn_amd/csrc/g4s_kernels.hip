// g4s_kernels.hip -- gfx950 kernels of the spike-propagation hot path.
//
// One simulation step = (a) Izhikevich update of every neuron with its
// thalamic draw, (b) spike detection compacted by wave64 ballot into the
// step's bitmap, raster row and spike queue, (c) propagation of the step's
// spikes into the next step's synaptic input.  Reference: gpu4snn.cu
// stateupdate (cu:334-417), propagatespikes/deliverspks1/deliverspks2/
// AB_deliverspks/SKL_deliverspks (cu:323-331, 420-467, 554-584, 669-716).
//
// Two accumulation modes:
//   atomic: (c) scatters fp32 atomics into isyn[(t+1)&1] (reference order:
//           none);
//   det:    (c) is fused into the next step's (a): each neuron gathers its
//           incoming synapses in ascending source order from the previous
//           step's bitmap, summing from 0.0f.  The result equals a sequential
//           push in ascending spike order, bit for bit.
// Compiled with -ffp-contract=off: the neuron arithmetic matches the host.
#include "g4s_internal.h"

#include <hip/hip_cooperative_groups.h>

namespace g4s {

namespace {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ int bits_below(uint64_t mask)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ int ld_agent(const int32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool error_set(const DevNet &d) { return ld_agent(d.err) != 0; }

__device__ __forceinline__ void raise_error(const DevNet &d)
{
    __hip_atomic_store(d.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One Izhikevich step of neuron j at step t with synaptic input `isyn`
// (cu:358-407).  Writes v, u, nospks and, on the run's last step, lastI.
__device__ __forceinline__ bool neuron_step(const DevNet &d, int t, int j, float isyn, bool last)
{
    const float z = thalamic(d.seed, (uint32_t)t, (uint32_t)j);
    const float g = (j < d.n_exc) ? d.ge : d.gi;
    float I = g * z;
    I = I + isyn;
    float v = d.v[j], u = d.u[j];
    const bool spk = izh_step(v, u, d.a[j], d.b[j], d.c[j], d.d[j], I);
    d.v[j] = v;
    d.u[j] = u;
    if (spk) d.nospks[j] += 1;
    if (last) d.lastI[j] = I;
    return spk;
}

// Spike output of one wave covering neurons [wave_base, wave_base+64):
// two bitmap words (always written, so stale bits are cleared), the raster
// row of the chunk, and the queue entries appended at one atomic per wave.
// kPublish: queue slots are stored with agent-scope atomics because
// consumers of the same launch read them (AP scheme).
template <bool kPublish>
__device__ __forceinline__ void emit(const DevNet &d, int t, int rel, int wave_base, bool spk, int j)
{
    const uint64_t mask = __ballot(spk);
    const int lane = lane_id();
    uint32_t *bm = (t & 1) ? d.bm1 : d.bm0;
    const int w = (wave_base >> 5) + lane;
    if (lane < 2 && w < d.words) {
        const uint32_t bits = lane ? (uint32_t)(mask >> 32) : (uint32_t)mask;
        bm[w] = bits;
        if (d.raster) d.raster[(size_t)rel * (size_t)d.words + (size_t)w] = bits;
    }
    if (mask) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&d.ctl[rel * kCtlStride + kCtlCount], __popcll(mask));
        base = __shfl(base, 0);
        if (spk) {
            uint32_t *q = (t & 1) ? d.q1 : d.q0;
            const int pos = base + bits_below(mask);
            if (kPublish)
                __hip_atomic_store(&q[pos], (uint32_t)j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else
                q[pos] = (uint32_t)j;
        }
    }
}

// Synaptic input of neuron j at step t in atomic mode: read and clear.
__device__ __forceinline__ float take_input(const DevNet &d, int t, int j)
{
    float *in = (t & 1) ? d.isyn1 : d.isyn0;
    const float x = in[j];
    in[j] = 0.0f;
    return x;
}

__device__ __forceinline__ bool spiked(const uint32_t *bm, uint32_t i)
{
    return (bm[i >> 5] >> (i & 31u)) & 1u;
}

// Ordered gather of neuron j's input from the previous step's spikes:
// incoming synapses in ascending source id, summed from 0.0f.  Loads are
// issued eight at a time; the additions stay in order.
__device__ __forceinline__ float gather_input(const DevNet &d, int t, int j)
{
    if (t == 0) return 0.0f;
    const uint32_t *bm = ((t - 1) & 1) ? d.bm1 : d.bm0;
    const int64_t e0 = d.in_ptr[j], e1 = d.in_ptr[j + 1];
    float acc = 0.0f;
    int64_t e = e0;
    for (; e + 8 <= e1; e += 8) {
        uint2 s[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] = d.in_syn[e + k];
        bool h[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) h[k] = spiked(bm, s[k].x);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (h[k]) acc = acc + u2f(s[k].y);
    }
    for (; e < e1; ++e) {
        const uint2 s = d.in_syn[e];
        if (spiked(bm, s.x)) acc = acc + u2f(s.y);
    }
    return acc;
}

// One wave scatters row s of the outgoing table into `out`.
__device__ __forceinline__ void push_row(const DevNet &d, uint32_t s, float *out, int lane)
{
    const uint2 *__restrict__ row = d.syn + (size_t)s * (size_t)d.nsyn;
    const int nsyn = d.nsyn;
    int k = lane;
    for (; k + 3 * kWave < nsyn; k += 4 * kWave) {
        const uint2 e0 = row[k], e1 = row[k + kWave], e2 = row[k + 2 * kWave], e3 = row[k + 3 * kWave];
        unsafeAtomicAdd(&out[e0.x], u2f(e0.y));
        unsafeAtomicAdd(&out[e1.x], u2f(e1.y));
        unsafeAtomicAdd(&out[e2.x], u2f(e2.y));
        unsafeAtomicAdd(&out[e3.x], u2f(e3.y));
    }
    for (; k < nsyn; k += kWave) {
        const uint2 e = row[k];
        unsafeAtomicAdd(&out[e.x], u2f(e.y));
    }
}

// Flattened synapse-parallel scatter of the step's queue: index g covers
// synapse g % Nsyn of queued row g / Nsyn.  Grid-stride.
__device__ __forceinline__ void push_flat(const DevNet &d, int t, int count, int64_t gid, int64_t stride)
{
    const uint32_t *q = (t & 1) ? d.q1 : d.q0;
    float *out = ((t + 1) & 1) ? d.isyn1 : d.isyn0;
    const int64_t total = (int64_t)count * d.nsyn;
    const double inv = 1.0 / (double)d.nsyn;
    for (int64_t g = gid; g < total; g += stride) {
        int64_t qi = (int64_t)((double)g * inv);
        int64_t k = g - qi * d.nsyn;
        if (k < 0) { qi -= 1; k += d.nsyn; }
        else if (k >= d.nsyn) { qi += 1; k -= d.nsyn; }
        const uint2 e = d.syn[(size_t)q[qi] * (size_t)d.nsyn + (size_t)k];
        unsafeAtomicAdd(&out[e.x], u2f(e.y));
    }
}

// Grid-wide barrier for cooperative launches: monotonically increasing
// arrival counter, target = epoch * gridDim.x.  Bounded: returns false (and
// raises the error flag) if the wait exceeds kSpinLimit or another
// workgroup raised the flag.
__device__ bool grid_sync(const DevNet &d, uint32_t &epoch)
{
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        epoch += 1;
        const uint32_t target = epoch * gridDim.x;
        __hip_atomic_fetch_add(d.bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        uint32_t spins = 0;
        while (__hip_atomic_load(d.bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (error_set(d)) { ok = 0; break; }
            if (++spins > kSpinLimit) { raise_error(d); ok = 0; break; }
            __builtin_amdgcn_s_sleep(1);
        }
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_init(DevNet d)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < d.n; j += stride) {
        const Params p = neuron_params(d.seed, (uint32_t)j, (uint32_t)d.n_exc);
        d.a[j] = p.a; d.b[j] = p.b; d.c[j] = p.c; d.d[j] = p.d;
        d.v[j] = -65.0f;
        d.u[j] = -65.0f * p.b;
        d.nospks[j] = 0;
        d.lastI[j] = 0.0f;
        d.isyn0[j] = 0.0f;
        d.isyn1[j] = 0.0f;
        d.q0[j] = kSentinel;
        d.q1[j] = kSentinel;
    }
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < d.words; w += stride) {
        d.bm0[w] = 0u;
        d.bm1[w] = 0u;
    }
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t *p, size_t n, uint32_t val)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = val;
}

// (a)+(b) for the atomic-mode per-step schemes N, S, AB.
__global__ void __launch_bounds__(256) k_update(DevNet d, int t, int rel, int last)
{
    const int stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x; base < d.n; base += stride) {
        const int j = base + (int)threadIdx.x;
        bool spk = false;
        if (j < d.n) spk = neuron_step(d, t, j, take_input(d, t, j), last != 0);
        emit<false>(d, t, rel, base + ((int)threadIdx.x & ~(kWave - 1)), spk, j);
    }
}

// (c) N scheme: one wave per queued spiking row.
__global__ void __launch_bounds__(256) k_push_rows(DevNet d, int t, int rel)
{
    const int count = d.ctl[rel * kCtlStride + kCtlCount];
    const uint32_t *q = (t & 1) ? d.q1 : d.q0;
    float *out = ((t + 1) & 1) ? d.isyn1 : d.isyn0;
    const int waves = gridDim.x * (blockDim.x / kWave);
    const int lane = lane_id();
    for (int r = blockIdx.x * (blockDim.x / kWave) + (int)(threadIdx.x / kWave); r < count; r += waves)
        push_row(d, q[r], out, lane);
}

// (c) S scheme (launched plainly) and AB scheme (launched cooperatively
// with a persistent, co-resident grid): thread per synapse of queued rows.
__global__ void __launch_bounds__(256) k_push_flat(DevNet d, int t, int rel)
{
    const int count = d.ctl[rel * kCtlStride + kCtlCount];
    push_flat(d, t, count, (int64_t)blockIdx.x * blockDim.x + threadIdx.x,
              (int64_t)gridDim.x * blockDim.x);
}

// AP scheme: one co-resident launch per step.  Every workgroup updates its
// share of neurons and publishes spikes into the queue, then becomes a
// consumer that claims queued rows one at a time (atomic head) and
// scatters them.  A row is claimed as soon as its slot is published, so
// propagation overlaps the update.  The queue slot holds kSentinel until
// published and is reset to kSentinel once consumed.
__global__ void __launch_bounds__(256) k_ap(DevNet d, int t, int rel, int last)
{
    const int stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x; base < d.n; base += stride) {
        const int j = base + (int)threadIdx.x;
        bool spk = false;
        if (j < d.n) spk = neuron_step(d, t, j, take_input(d, t, j), last != 0);
        emit<true>(d, t, rel, base + ((int)threadIdx.x & ~(kWave - 1)), spk, j);
    }
    __syncthreads();
    int32_t *ctl = d.ctl + rel * kCtlStride;
    if (threadIdx.x == 0)
        __hip_atomic_fetch_add(&ctl[kCtlDone], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);

    uint32_t *q = (t & 1) ? d.q1 : d.q0;
    float *out = ((t + 1) & 1) ? d.isyn1 : d.isyn0;
    const int lane = lane_id();
    for (;;) {
        int item = 0, go = 0;
        uint32_t s = 0;
        if (lane == 0) {
            item = atomicAdd(&ctl[kCtlHead], 1);
            uint32_t spins = 0;
            for (;;) {
                if (item < ld_agent(&ctl[kCtlCount])) { go = 1; break; }
                if (__hip_atomic_load(&ctl[kCtlDone], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) ==
                    (int)gridDim.x) {
                    go = item < ld_agent(&ctl[kCtlCount]);
                    break;
                }
                if (error_set(d)) break;
                if (++spins > kSpinLimit) { raise_error(d); break; }
                __builtin_amdgcn_s_sleep(1);
            }
            if (go) {
                spins = 0;
                for (;;) {
                    s = __hip_atomic_load(&q[item], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (s != kSentinel) break;
                    if (++spins > kSpinLimit) { raise_error(d); go = 0; break; }
                    __builtin_amdgcn_s_sleep(1);
                }
                if (go) __hip_atomic_store(&q[item], kSentinel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        go = __shfl(go, 0);
        if (!go) break;
        s = __shfl(s, 0);
        push_row(d, s, out, lane);
    }
}

// SKL scheme: the whole chunk in one co-resident launch; per step an
// update phase, a grid barrier, a synapse-parallel scatter and a barrier.
__global__ void __launch_bounds__(256) k_skl(DevNet d, int t0, int steps, int last_t)
{
    uint32_t epoch = 0;
    const int stride = gridDim.x * blockDim.x;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int rel = 0; rel < steps; ++rel) {
        const int t = t0 + rel;
        for (int base = blockIdx.x * blockDim.x; base < d.n; base += stride) {
            const int j = base + (int)threadIdx.x;
            bool spk = false;
            if (j < d.n) spk = neuron_step(d, t, j, take_input(d, t, j), t == last_t);
            emit<false>(d, t, rel, base + ((int)threadIdx.x & ~(kWave - 1)), spk, j);
        }
        if (!grid_sync(d, epoch)) return;
        const int count = d.ctl[rel * kCtlStride + kCtlCount];
        push_flat(d, t, count, gid, (int64_t)stride);
        if (!grid_sync(d, epoch)) return;
    }
}

// det mode, N and AB schemes: thread per neuron gathers its input and
// updates (AB: launched cooperatively with the persistent grid).
__global__ void __launch_bounds__(256) k_pull_thread(DevNet d, int t, int rel, int last)
{
    const int stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x; base < d.n; base += stride) {
        const int j = base + (int)threadIdx.x;
        bool spk = false;
        if (j < d.n) spk = neuron_step(d, t, j, gather_input(d, t, j), last != 0);
        emit<false>(d, t, rel, base + ((int)threadIdx.x & ~(kWave - 1)), spk, j);
    }
}

// det mode, S scheme: synapse-parallel gather.  A wave owns 64 neurons; for
// each of them the 64 lanes load 64 incoming synapses at a time, test the
// source bits and the hits are added in lane (= ascending source) order.
__global__ void __launch_bounds__(256) k_pull_wave(DevNet d, int t, int rel, int last)
{
    const int lane = lane_id();
    const uint32_t *bm = ((t - 1) & 1) ? d.bm1 : d.bm0;
    const int waves = gridDim.x * (blockDim.x / kWave);
    for (int wb = (blockIdx.x * (blockDim.x / kWave) + (int)(threadIdx.x / kWave)) * kWave; wb < d.n;
         wb += waves * kWave) {
        float mine = 0.0f;
        if (t > 0) {
            const int nj = min(kWave, d.n - wb);
            for (int p = 0; p < nj; ++p) {
                const int j = wb + p;
                const int64_t e0 = d.in_ptr[j], e1 = d.in_ptr[j + 1];
                float acc = 0.0f;
                for (int64_t e = e0; e < e1; e += kWave) {
                    const int64_t my = e + lane;
                    uint2 s = make_uint2(0u, 0u);
                    bool hit = false;
                    if (my < e1) {
                        s = d.in_syn[my];
                        hit = spiked(bm, s.x);
                    }
                    uint64_t m = __ballot(hit);
                    while (m) {
                        const int k = __builtin_ctzll(m);
                        acc = acc + __builtin_bit_cast(float, __builtin_amdgcn_readlane((int)s.y, k));
                        m &= m - 1;
                    }
                }
                if (lane == p) mine = acc;
            }
        }
        const int j = wb + lane;
        bool spk = false;
        if (j < d.n) spk = neuron_step(d, t, j, mine, last != 0);
        emit<false>(d, t, rel, wb, spk, j);
    }
}

// det mode, AP scheme: one co-resident launch per step; waves claim
// 64-neuron tiles from an atomic work counter (dynamic load balance).
__global__ void __launch_bounds__(256) k_ap_det(DevNet d, int t, int rel, int last)
{
    int32_t *ctl = d.ctl + rel * kCtlStride;
    const int lane = lane_id();
    const int tiles = (d.n + kWave - 1) / kWave;
    for (;;) {
        int tile = 0;
        if (lane == 0) tile = atomicAdd(&ctl[kCtlHead], 1);
        tile = __shfl(tile, 0);
        if (tile >= tiles) break;
        const int j = tile * kWave + lane;
        bool spk = false;
        if (j < d.n) spk = neuron_step(d, t, j, gather_input(d, t, j), last != 0);
        emit<false>(d, t, rel, tile * kWave, spk, j);
    }
}

// det mode, SKL scheme: the whole chunk in one co-resident launch; one
// gather+update phase and one grid barrier per step.
__global__ void __launch_bounds__(256) k_skl_det(DevNet d, int t0, int steps, int last_t)
{
    uint32_t epoch = 0;
    const int stride = gridDim.x * blockDim.x;
    for (int rel = 0; rel < steps; ++rel) {
        const int t = t0 + rel;
        for (int base = blockIdx.x * blockDim.x; base < d.n; base += stride) {
            const int j = base + (int)threadIdx.x;
            bool spk = false;
            if (j < d.n) spk = neuron_step(d, t, j, gather_input(d, t, j), t == last_t);
            emit<false>(d, t, rel, base + ((int)threadIdx.x & ~(kWave - 1)), spk, j);
        }
        if (!grid_sync(d, epoch)) return;
    }
}

// Switching from det to atomic mode at step t: the pending input of step t
// is gathered from step t-1's bitmap into isyn[t&1]; isyn[(t+1)&1] cleared.
__global__ void __launch_bounds__(256) k_gather_isyn(DevNet d, int t)
{
    float *in = (t & 1) ? d.isyn1 : d.isyn0;
    float *nx = (t & 1) ? d.isyn0 : d.isyn1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < d.n; j += stride) {
        in[j] = gather_input(d, t, (int)j);
        nx[j] = 0.0f;
    }
}

__global__ void __launch_bounds__(256) k_thalamic(uint64_t seed, uint32_t t, int64_t n, float *out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = thalamic(seed, t, (uint32_t)i);
}

__global__ void __launch_bounds__(256) k_normal_words(const uint32_t *x0, const uint32_t *x1, int64_t n,
                                                      float *out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = normal(x0[i], x1[i]);
}

int grid_for(int64_t n, int block)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 1048576) g = 1048576;
    return (int)g;
}

hipError_t coop(const void *fn, const Launch &L, void **args)
{
    return hipLaunchCooperativeKernel(fn, dim3(L.grid), dim3(L.block), args, 0, L.stream);
}

}  // namespace

hipError_t launch_init(const DevNet &d, hipStream_t s)
{
    const int64_t m = d.n > d.words ? d.n : d.words;
    hipLaunchKernelGGL(k_init, dim3(grid_for(m, 256)), dim3(256), 0, s, d);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t val, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for((int64_t)n, 256)), dim3(256), 0, s, p, n, val);
    return hipGetLastError();
}

hipError_t launch_update(const DevNet &d, int t, int rel, int last, const Launch &L)
{
    hipLaunchKernelGGL(k_update, dim3(L.grid), dim3(L.block), 0, L.stream, d, t, rel, last);
    return hipGetLastError();
}

hipError_t launch_push_rows(const DevNet &d, int t, int rel, const Launch &L)
{
    hipLaunchKernelGGL(k_push_rows, dim3(L.grid), dim3(L.block), 0, L.stream, d, t, rel);
    return hipGetLastError();
}

hipError_t launch_push_flat(const DevNet &d, int t, int rel, const Launch &L, bool cooperative)
{
    if (!cooperative) {
        hipLaunchKernelGGL(k_push_flat, dim3(L.grid), dim3(L.block), 0, L.stream, d, t, rel);
        return hipGetLastError();
    }
    DevNet dd = d;
    void *args[] = {&dd, &t, &rel};
    return coop((const void *)k_push_flat, L, args);
}

hipError_t launch_ap(const DevNet &d, int t, int rel, int last, const Launch &L)
{
    DevNet dd = d;
    void *args[] = {&dd, &t, &rel, &last};
    return coop((const void *)k_ap, L, args);
}

hipError_t launch_skl(const DevNet &d, int t0, int steps, int last_t, const Launch &L)
{
    DevNet dd = d;
    void *args[] = {&dd, &t0, &steps, &last_t};
    return coop((const void *)k_skl, L, args);
}

hipError_t launch_pull_thread(const DevNet &d, int t, int rel, int last, const Launch &L, bool cooperative)
{
    if (!cooperative) {
        hipLaunchKernelGGL(k_pull_thread, dim3(L.grid), dim3(L.block), 0, L.stream, d, t, rel, last);
        return hipGetLastError();
    }
    DevNet dd = d;
    void *args[] = {&dd, &t, &rel, &last};
    return coop((const void *)k_pull_thread, L, args);
}

hipError_t launch_pull_wave(const DevNet &d, int t, int rel, int last, const Launch &L)
{
    hipLaunchKernelGGL(k_pull_wave, dim3(L.grid), dim3(L.block), 0, L.stream, d, t, rel, last);
    return hipGetLastError();
}

hipError_t launch_ap_det(const DevNet &d, int t, int rel, int last, const Launch &L)
{
    DevNet dd = d;
    void *args[] = {&dd, &t, &rel, &last};
    return coop((const void *)k_ap_det, L, args);
}

hipError_t launch_skl_det(const DevNet &d, int t0, int steps, int last_t, const Launch &L)
{
    DevNet dd = d;
    void *args[] = {&dd, &t0, &steps, &last_t};
    return coop((const void *)k_skl_det, L, args);
}

hipError_t launch_gather_isyn(const DevNet &d, int t, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_isyn, dim3(grid_for(d.n, 256)), dim3(256), 0, s, d, t);
    return hipGetLastError();
}

hipError_t launch_thalamic(uint64_t seed, uint32_t t, int64_t n, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_thalamic, dim3(grid_for(n, 256)), dim3(256), 0, s, seed, t, n, out);
    return hipGetLastError();
}

hipError_t launch_normal_words(const uint32_t *x0, const uint32_t *x1, int64_t n, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_normal_words, dim3(grid_for(n, 256)), dim3(256), 0, s, x0, x1, n, out);
    return hipGetLastError();
}

int coop_max_blocks(CoopKernel k, int block)
{
    const void *fn = nullptr;
    switch (k) {
    case kCoopPushFlat: fn = (const void *)k_push_flat; break;
    case kCoopAP: fn = (const void *)k_ap; break;
    case kCoopSKL: fn = (const void *)k_skl; break;
    case kCoopPull: fn = (const void *)k_pull_thread; break;
    case kCoopAPDet: fn = (const void *)k_ap_det; break;
    case kCoopSKLDet: fn = (const void *)k_skl_det; break;
    }
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0) != hipSuccess) return 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return per_cu * cus;
}

}  // namespace g4s
