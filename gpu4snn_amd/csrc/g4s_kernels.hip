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
//   det:    (c) is fused into the next step's (a): each slice of neurons
//           gathers the segments of the previous step's spiking rows that
//           target it and sums every neuron's input in ascending source
//           order from 0.0f (every scheme; see slice_accumulate).  The result
//           equals a sequential push in ascending spike order, bit for bit.
// Compiled with -ffp-contract=off: the neuron arithmetic matches the host.
#include "g4s_internal.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

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

// Start of step `rel` of a chunk on the device's constant-rate wall clock,
// in words 2-3 of the step's control row; row `cs` (one past the chunk's
// last step) receives the chunk's end.  Workgroup 0 of the step's first
// kernel stamps it, so consecutive stamps bracket one step.
__device__ __forceinline__ void stamp_step(int32_t *ctl, int rel)
{
    const uint64_t c = (uint64_t)wall_clock64();
    ctl[rel * kCtlStride + kCtlStamp] = (int32_t)(uint32_t)c;
    ctl[rel * kCtlStride + kCtlStamp + 1] = (int32_t)(uint32_t)(c >> 32);
}

__device__ __forceinline__ void raise_error(const DevNet &d)
{
    __hip_atomic_store(d.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One Izhikevich step of neuron j at step t with synaptic input `isyn`
// (cu:358-407).  Writes v, u, nospks and, on the run's last step, lastI.
template <class Net>
__device__ __forceinline__ bool neuron_step(const Net &d, int t, int j, float isyn, bool last)
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

// Spike output of one workgroup tile [base, base + blockDim.x): each wave's
// bitmap and raster words, and the tile's queue entries appended at one
// device-scope atomic per workgroup.  Such atomics on one address are
// serialised (about 30 ns each on MI355X), so one per wave would cost more.
// Every thread of the workgroup calls it (it has barriers); at most 1024
// threads.  kPublish (AP): the word ctl[kCtlCount..+1] is one 64-bit value,
// spikes queued so far in the low half and finished workgroups in the high
// half (`finished` adds one), and the slots are stored with agent-scope
// atomics, because consumers of the same launch read them.
template <bool kPublish>
__device__ __forceinline__ void emit_tile(const DevNet &d, int t, int rel, int base, bool spk, int j, bool finished)
{
    __shared__ int s_wcnt[1024 / kWave];
    __shared__ int s_base;
    const int lane = lane_id();
    const int wave = (int)threadIdx.x / kWave;
    const int nwave = (int)blockDim.x / kWave;
    const int wave_base = base + wave * kWave;
    const uint64_t mask = __ballot(spk);
    if (wave_base < d.n) {
        uint32_t *bm = (t & 1) ? d.bm1 : d.bm0;
        const int w = (wave_base >> 5) + lane;
        if (lane < 2 && w < d.words) {
            const uint32_t bits = lane ? (uint32_t)(mask >> 32) : (uint32_t)mask;
            bm[w] = bits;
            if (d.raster) d.raster[(size_t)rel * (size_t)d.words + (size_t)w] = bits;
        }
    }
    if (lane == 0) s_wcnt[wave] = __popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int k = 0; k < nwave; ++k) tot += s_wcnt[k];
        int32_t *ctl = d.ctl + rel * kCtlStride + kCtlCount;
        if (kPublish) {
            const uint64_t inc = (uint64_t)tot + (finished ? (1ull << 32) : 0ull);
            s_base = (int)(uint32_t)__hip_atomic_fetch_add(reinterpret_cast<uint64_t *>(ctl), inc, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT);
        } else {
            s_base = tot ? atomicAdd(ctl, tot) : 0;
        }
    }
    __syncthreads();
    if (spk) {
        int pos = s_base + bits_below(mask);
        for (int k = 0; k < wave; ++k) pos += s_wcnt[k];
        uint32_t *q = (t & 1) ? d.q1 : d.q0;
        if (kPublish)
            __hip_atomic_store(&q[pos], (uint32_t)j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
            q[pos] = (uint32_t)j;
    }
    __syncthreads();  // s_wcnt and s_base are rewritten by the next tile
}

// Synaptic input of neuron j at step t in atomic mode: read and clear.
__device__ __forceinline__ float take_input(const DevNet &d, int t, int j)
{
    float *in = (t & 1) ? d.isyn1 : d.isyn0;
    const float x = in[j];
    in[j] = 0.0f;
    return x;
}

// One wave scatters row s of the outgoing table into `out`.  Queue entries
// outside [0, n) (an unpublished kSentinel slot) are ignored.
__device__ __forceinline__ void push_row(const DevNet &d, uint32_t s, float *out, int lane)
{
    if (s >= (uint32_t)d.n) return;
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
        const uint32_t s = q[qi];
        if (s >= (uint32_t)d.n) continue;
        const uint2 e = d.syn[(size_t)s * (size_t)d.nsyn + (size_t)k];
        unsafeAtomicAdd(&out[e.x], u2f(e.y));
    }
}

// Grid-wide barrier of a persistent launch.  Workgroup b arrives on the
// counter of its group b % kBarGroups (two groups per XCD) with a returning
// atomic add; the group's last arriver of the epoch (old value e*m - 1, m
// workgroups in the group) stores e into the group's release flag.  Lanes
// 0..kBarGroups-1 of the first wave poll the flags together, and lane
// kBarGroups the error flag in the same round trip, until every flag holds
// the epoch.  The polled lines change once per barrier, not at every
// arrival, which measured 0.6-0.9 us faster than polling the counters
// themselves (241 workgroups).  Bounded: returns false (and raises the
// error flag) if the wait exceeds kSpinLimit, or returns false once another
// workgroup has raised the flag.
// It orders only data written and read with agent-scope accesses (the spike
// bitmap of the SKL kernel): every wave waits for its own stores to complete
// before the workgroup arrives, a group's flag is stored only after all of
// its members' arrivals, and the loads after the barrier are issued only
// once every flag has been seen.  It pays no release (an L2 write-back of
// the XCD) and no acquire (an L2 invalidate, which would also evict the
// static synapse rows that the next step reads).  Between its two workgroup
// barriers a non-polling wave stores the step's spike count.
static_assert(kBarGroups < kWave, "the barrier's polling wave also reads the error flag");
__device__ bool grid_sync(const DevNet &d, uint32_t &epoch, const int *s_wcnt, int32_t *cnt_dst)
{
    __shared__ int s_ok;
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores have completed
    __syncthreads();
    // the workgroup's spike count of the step (s_wcnt: one total per wave,
    // written before the call), summed by a wave that does not poll
    const int nw = (int)blockDim.x / kWave;
    if ((int)threadIdx.x == (nw > 1 ? kWave : 0)) {
        int c = 0;
        for (int w = 0; w < nw; ++w) c += s_wcnt[w];
        *cnt_dst = c;
    }
    epoch += 1;
    if (threadIdx.x < kWave) {
        const int lane = (int)threadIdx.x;
        const uint32_t G = gridDim.x;
        uint32_t *flags = d.bar + kBarGroups * kBarStride;
        if (lane == 0) {
            const uint32_t grp = blockIdx.x % kBarGroups;
            const uint32_t members = G / kBarGroups + (grp < G % kBarGroups ? 1u : 0u);
            const uint32_t old = __hip_atomic_fetch_add(&d.bar[grp * kBarStride], 1u, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
            if (old + 1u == epoch * members)
                __hip_atomic_store(&flags[grp * kBarStride], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // lanes 0..kBarGroups-1 read the flags and lane kBarGroups the
        // error flag, in one round trip per poll; groups without members
        // (grids smaller than kBarGroups) are never waited on
        const bool waits = lane < kBarGroups && (uint32_t)lane < G;
        const uint32_t *src = lane < kBarGroups ? &flags[lane * kBarStride]
                                                : (lane == kBarGroups ? (const uint32_t *)d.err : &flags[0]);
        int ok = 1;
        uint32_t spins = 0;
        for (;;) {
            const uint32_t x = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__all(!waits || x >= epoch)) break;
            if (__any(lane == kBarGroups && x != 0u)) {  // another workgroup gave up
                ok = 0;
                break;
            }
            if (++spins > kSpinLimit) {
                if (lane == 0) raise_error(d);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (lane == 0) s_ok = ok;
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
        if (d.isyn0) {  // single-network handles; batched instances have none
            d.isyn0[j] = 0.0f;
            d.isyn1[j] = 0.0f;
            d.q0[j] = kSentinel;
            d.q1[j] = kSentinel;
        }
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

__global__ void __launch_bounds__(256) k_fill2_u32(uint32_t *p, size_t n, uint32_t *q, size_t m, uint32_t val)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n + m; i += stride) {
        if (i < n)
            p[i] = val;
        else
            q[i - n] = val;
    }
}

// (a)+(b) for the atomic-mode per-step schemes N, S, AB.
__global__ void __launch_bounds__(256) k_update(DevNet d, int t, int rel, int last)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) stamp_step(d.ctl, rel);
    const int stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x; base < d.n; base += stride) {
        const int j = base + (int)threadIdx.x;
        bool spk = false;
        if (j < d.n) spk = neuron_step(d, t, j, take_input(d, t, j), last != 0);
        emit_tile<false>(d, t, rel, base, spk, j, false);
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

// Scatter every row whose bit is set in step t's bitmap into isyn[(t+1)&1]:
// a wave per 64 neurons walks its two bitmap words.
__global__ void __launch_bounds__(256) k_push_bitmap(DevNet d, int t)
{
    const uint32_t *bm = (t & 1) ? d.bm1 : d.bm0;
    float *out = ((t + 1) & 1) ? d.isyn1 : d.isyn0;
    const int lane = lane_id();
    const int waves = gridDim.x * (blockDim.x / kWave);
    const int groups = (d.words + 1) / 2;
    for (int gi = blockIdx.x * (blockDim.x / kWave) + (int)(threadIdx.x / kWave); gi < groups; gi += waves) {
        uint64_t m = bm[2 * gi];
        if (2 * gi + 1 < d.words) m |= (uint64_t)bm[2 * gi + 1] << 32;
        while (m) {
            const int k = __builtin_ctzll(m);
            m &= m - 1;
            push_row(d, (uint32_t)(gi * 64 + k), out, lane);
        }
    }
}

// (c) S scheme: thread per synapse of queued rows.
__global__ void __launch_bounds__(256) k_push_flat(DevNet d, int t, int rel)
{
    const int count = d.ctl[rel * kCtlStride + kCtlCount];
    push_flat(d, t, count, (int64_t)blockIdx.x * blockDim.x + threadIdx.x,
              (int64_t)gridDim.x * blockDim.x);
}

// AP scheme: one co-resident launch per step.  Every workgroup updates its
// share of neurons and publishes their spikes into the queue, then its waves
// become consumers that scatter queued rows.  A row is scattered as soon as
// its slot is published, so propagation overlaps the update.
//
// Control: ctl[kCtlCount] (low half) and ctl[kCtlCount + 1] (high half) form
// one 64-bit word: spikes queued so far, and workgroups finished.  Each
// workgroup adds its tile's spike total (and, with its last tile, one
// finished workgroup) in a single atomic.  Device-scope atomics on one
// address are serialised (about 30 ns each on MI355X), so the launch makes
// one per tile and no per-row claims: consumer wave w takes queue items w,
// w + W, w + 2W, ... (W waves in the grid).  It waits until its item is below
// the count, or until the word shows every workgroup finished, which makes
// the count final.  A slot holds kSentinel until its id is stored (the count
// is raised first).  Everything consumers read from the launch is written
// with agent-scope atomics, so neither a release (an L2 write-back) nor an
// acquire (an L2 invalidate) is needed.
// Step t re-arms the next step's buffer q[(t+1)&1] (only its first
// count(t-1) slots can be non-sentinel); nothing reads that buffer during
// step t, and consumed slots keep their ids so the last step's queue stays
// valid for whichever scheme runs next.  Workgroups of at most 256 threads.
__global__ void __launch_bounds__(256) k_ap(DevNet d, int t, int rel, int last, int prev_count)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) stamp_step(d.ctl, rel);
    const int stride = gridDim.x * blockDim.x;
    {
        uint32_t *qn = ((t + 1) & 1) ? d.q1 : d.q0;
        const int used = t == 0 ? 0 : (rel ? d.ctl[(rel - 1) * kCtlStride + kCtlCount] : prev_count);
        for (int i = blockIdx.x * blockDim.x + (int)threadIdx.x; i < used; i += stride) qn[i] = kSentinel;
    }
    const uint64_t *word = reinterpret_cast<const uint64_t *>(d.ctl + rel * kCtlStride + kCtlCount);
    const uint32_t *q = (t & 1) ? d.q1 : d.q0;
    const int lane = lane_id();
    const int wave = (int)threadIdx.x / kWave;
    const int nwave = (int)blockDim.x / kWave;
    for (int base = blockIdx.x * blockDim.x; base < d.n; base += stride) {
        const int j = base + (int)threadIdx.x;
        bool spk = false;
        if (j < d.n) spk = neuron_step(d, t, j, take_input(d, t, j), last != 0);
        emit_tile<true>(d, t, rel, base, spk, j, base + stride >= d.n);
    }

    float *out = ((t + 1) & 1) ? d.isyn1 : d.isyn0;
    const int waves = (int)gridDim.x * nwave;
    int final_count = -1;  // known once every workgroup has finished
    for (int item = (int)blockIdx.x * nwave + wave;; item += waves) {
        int go = 0;
        uint32_t s = 0;
        if (lane == 0) {
            if (final_count >= 0) {
                go = item < final_count;
            } else {
                uint32_t spins = 0;
                for (;;) {
                    const uint64_t v = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (item < (int)(uint32_t)v) { go = 1; break; }
                    if ((int)(v >> 32) == (int)gridDim.x) {
                        final_count = (int)(uint32_t)v;
                        break;
                    }
                    if (error_set(d)) break;
                    if (++spins > kSpinLimit) { raise_error(d); break; }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            if (go) {
                uint32_t spins = 0;
                for (;;) {
                    s = __hip_atomic_load(&q[item], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (s != kSentinel) break;
                    if (++spins > kSpinLimit) { raise_error(d); go = 0; break; }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        go = __shfl(go, 0);
        if (!go) break;
        s = __shfl(s, 0);
        push_row(d, s, out, lane);
    }
}

// ---------------------------------------------------------------------------
// Slice-owned propagation (AB, SKL and the batched form, both modes).
// Workgroup p owns the neurons of slice p.  Its input for step t is the part
// of every row that spiked at step t-1 that targets slice p: a contiguous
// segment of the sorted row, located by the offset table.  Segments are
// accumulated in LDS, so there are no global float atomics and no
// pending-input array in HBM; the owner then updates its neurons directly.
// ---------------------------------------------------------------------------

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b to XCD b mod
// 8); consecutive slices go to the same XCD so rows' neighbouring segments
// and offsets share its L2.  XCD x holds g/8 workgroups, one more if
// x < g mod 8, and takes the next run of slices: a bijection for any g.
__host__ __device__ inline int slice_of_block(int b, int g)
{
    const int x = b % 8, q = g / 8, r = g % 8;
    return x * q + (x < r ? x : r) + b / 8;
}

// Bitmaps are read in groups of four words (16 bytes, one load) through a
// raw buffer resource: dword 3 of a gfx9-family buffer descriptor with a
// 32-bit data format and no index stride, and the buffer's extent rounded
// up to a whole group (bitmap allocations are padded to it).  Loads past
// the extent read 0.  kSc1 is the load's agent-scope coherence bit (sc1),
// the bit an agent-scope relaxed atomic load carries on gfx950.
constexpr int kBufWord3 = 0x00020000;
constexpr int kSc1 = 16;

template <class P>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t bm_rsrc(P bm, int W)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)bm, (short)0, ((W + 3) & ~3) * 4, kBufWord3);
}

// Group g of a bitmap of W words: words 4g .. 4g+3, those at or past W read
// as 0.  kCoh: the words were stored by other workgroups of the same launch
// with agent-scope atomics, and the load is coherent at agent scope.
template <bool kCoh>
__device__ __forceinline__ uint4 bm_group(__amdgpu_buffer_rsrc_t r, int g, int W)
{
    const auto x = __builtin_amdgcn_raw_buffer_load_b128(r, g * 16, 0, kCoh ? kSc1 : 0);
    uint4 y = make_uint4(x[0], x[1], x[2], x[3]);
    const int over = 4 * g + 4 - W;  // words of the group at or past W
    if (over > 0) {
        y.w = 0u;
        if (over > 1) y.z = 0u;
        if (over > 2) y.y = 0u;
        if (over > 3) y.x = 0u;
    }
    return y;
}

__device__ __forceinline__ int popc4(uint4 x) { return __popc(x.x) + __popc(x.y) + __popc(x.z) + __popc(x.w); }

// Spike ids of bitmap group g, ascending, numbered idx, idx+1, ... in the
// pass; those numbered in [lo, hi) go to ids[number - lo] (kAll: all of
// them are known to be, with lo = 0).
template <bool kAll>
__device__ __forceinline__ void put_group(uint4 x, int g, int idx, uint32_t *ids, int lo, int hi)
{
    const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        uint32_t bits = w[e];
        while (bits) {
            if (kAll || (idx >= lo && idx < hi)) ids[idx - lo] = (uint32_t)(4 * g + e) * 32u + (uint32_t)__builtin_ctz(bits);
            bits &= bits - 1u;
            ++idx;
        }
    }
}

// Bitmap and raster words of one wave covering [wave_base, wave_base+64)
// of a slice kernel; returns the wave's spike mask.  A word is written only
// if it starts below `limit` (the end of the slice): slices are multiples of
// 32 neurons, so the second word of a slice's last wave may be the next
// slice's.  kCoh: bitmap words are stored with agent-scope atomics (read by
// other workgroups of the same launch).
template <bool kCoh = false, class Net>
__device__ __forceinline__ uint64_t emit_bits(const Net &d, int t, int rel, int wave_base, bool spk, int limit)
{
    const uint64_t mask = __ballot(spk);
    if (wave_base >= limit) return mask;  // uniform over the wave
    const int lane = lane_id();
    auto bm = (t & 1) ? d.bm1 : d.bm0;
    const int w = (wave_base >> 5) + lane;
    if (lane < 2 && w < d.words && (w << 5) < limit) {
        const uint32_t bits = lane ? (uint32_t)(mask >> 32) : (uint32_t)mask;
        if (kCoh)
            __hip_atomic_store(&bm[w], bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
            bm[w] = bits;
        if (d.raster) d.raster[(size_t)rel * (size_t)d.words + (size_t)w] = bits;
    }
    return mask;
}

// Spike count of slice p at step t (waves' ballots of each tile) written to
// bcnt[rel][p] for the host's per-step counts.  No global atomics.  *s_cnt
// is zero on entry (zeroed before a workgroup barrier that precedes the
// call) and again on return for the next step's call.
template <int R>
__device__ __forceinline__ void slice_count(const DevNet &d, int rel, int p, int ntile, const uint64_t (&masks)[R],
                                            int *s_cnt)
{
    int c = 0;
#pragma unroll
    for (int k = 0; k < R; ++k)
        if (k < ntile) c += __popcll(masks[k]);
    if (lane_id() == 0 && c) atomicAdd(s_cnt, c);
    __syncthreads();
    if (threadIdx.x == 0) {
        d.bcnt[(size_t)rel * (size_t)d.P + (size_t)p] = *s_cnt;
        *s_cnt = 0;
    }
}

// Exclusive prefix sum over the workgroup of one value per thread; returns
// the thread's prefix and sets *total.  s_tmp holds blockDim/64 ints.  One
// barrier: every thread reads all wave totals.  The caller guarantees a
// workgroup barrier between the reads of the previous use of s_tmp and
// this call.
// Inclusive scans over the 64 lanes of a wave (lane 0 first) by data-parallel
// primitives: row shifts by 1, 2, 4 and 8 scan each row of 16 lanes (a lane
// with no source lane in its row adds the identity 0), then the last lane
// of row 0 is added to row 1 and of row 2 to row 3 (row broadcast 15), and
// the last lane of row 1 to rows 2 and 3 (row broadcast 31).  A few cycles
// per step, where a lane shuffle is an LDS round trip.
__device__ __forceinline__ int wave_inclusive_sum(int x)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1 and 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2 and 3
    return x;
}
__device__ __forceinline__ uint32_t wave_inclusive_max(uint32_t x)
{
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false));
    return x;
}
// The value of the wave's last lane, in every lane.
__device__ __forceinline__ int wave_last(int x) { return __builtin_amdgcn_readlane(x, kWave - 1); }

// This thread's wave index, as a wave-uniform value.
__device__ __forceinline__ int wave_index() { return __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave); }

// Spike ids of two bitmap groups per lane, group ga of every lane of the
// wave before group ga + 64 of any (the lanes' groups ga are consecutive):
// the wave's spikes of the two are numbered from `run` (wave-uniform) in
// ascending id and placed by put_group.  One wave scan over both counts
// packed in 16-bit halves (at most 128 per group, 8192 per wave: no carry).
// Returns run advanced past both.
template <bool kAll>
__device__ __forceinline__ int put_pair(uint4 xa, uint4 xb, int ga, int run, uint32_t *ids, int lo, int hi)
{
    const int pk = popc4(xa) | (popc4(xb) << 16);
    const int incl = wave_inclusive_sum(pk);
    const int tot = wave_last(incl);
    const int ex = incl - pk;
    put_group<kAll>(xa, ga, run + (ex & 0xFFFF), ids, lo, hi);
    put_group<kAll>(xb, ga + kWave, run + (tot & 0xFFFF) + (ex >> 16), ids, lo, hi);
    return run + (tot & 0xFFFF) + (tot >> 16);
}

// Per-wave values vals[0..nw) in LDS (nw = waves of the workgroup, at most
// 64): lane l reads vals[l] and one wave scan gives, in every lane, the sum
// over the waves before this one and (*total) over all of them.  A loop
// over the waves instead costs each wave a few scalar instructions per
// wave, and the CU's waves share one scalar unit.
__device__ __forceinline__ int waves_prefix_sum(const int *vals, int *total)
{
    const int lane = lane_id();
    const int nw = (int)blockDim.x / kWave;
    const int wave = wave_index();
    const int incl = wave_inclusive_sum(lane < nw ? vals[lane] : 0);
    *total = __builtin_amdgcn_readlane(incl, nw - 1);
    const int before = __builtin_amdgcn_readlane(incl, wave > 0 ? wave - 1 : 0);
    return wave > 0 ? before : 0;
}
// The same for a per-wave value already held by lane `wave` (0 elsewhere).
__device__ __forceinline__ int waves_prefix_sum_of(int lane_value, int *total)
{
    const int nw = (int)blockDim.x / kWave;
    const int wave = wave_index();
    const int incl = wave_inclusive_sum(lane_value);
    *total = __builtin_amdgcn_readlane(incl, nw - 1);
    const int before = __builtin_amdgcn_readlane(incl, wave > 0 ? wave - 1 : 0);
    return wave > 0 ? before : 0;
}
// The same for the maximum (values >= 0).
__device__ __forceinline__ uint32_t waves_prefix_max(const int *vals, uint32_t *total)
{
    const int lane = lane_id();
    const int nw = (int)blockDim.x / kWave;
    const int wave = wave_index();
    const uint32_t incl = wave_inclusive_max(lane < nw ? (uint32_t)vals[lane] : 0u);
    *total = (uint32_t)__builtin_amdgcn_readlane((int)incl, nw - 1);
    const uint32_t before = (uint32_t)__builtin_amdgcn_readlane((int)incl, wave > 0 ? wave - 1 : 0);
    return wave > 0 ? before : 0u;
}

__device__ __forceinline__ int block_exclusive_scan(int v, int *s_tmp, int *total)
{
    const int lane = lane_id();
    const int wave = (int)threadIdx.x / kWave;
    const int incl = wave_inclusive_sum(v);
    if (lane == kWave - 1) s_tmp[wave] = incl;
    __syncthreads();
    return waves_prefix_sum(s_tmp, total) + incl - v;
}

// Spikes per pass per thread and entry loads in flight per thread:
// single-network slices (up to 1024 threads) and batched slices (256).
// Large single-network slices (kLargeSlice neurons or more: networks of
// about 10^6 neurons) take more spikes per pass and twice the entry loads in
// flight, cutting the dependent levels of a step whose spikes need several
// passes: 8 spikes per thread in both modes (det's workspace, which also
// holds the window's counting sort, then fits slices up to about 3970
// neurons at 1024 threads).  Det finds each entry's spike by the owner map
// (measured 2.5% faster at N=10^6 than the search; the same in atomic mode).
constexpr int kSpikesSingle = 2, kEntriesSingle = 4;
constexpr int kLargeSlice = 2048, kEntriesLarge = 8;
constexpr int kSpikesLargeAtomic = 8, kSpikesLargeDet = 8;
constexpr bool kOwnerLargeAtomic = false, kOwnerLargeDet = true;
// Batched slices: 6 entry loads in flight per thread, and 6 spikes per pass
// per thread in atomic mode, 2 in det mode.  At the default 768-neuron
// slices 7 workgroups' LDS then fits a CU in both modes (atomic 21.3 KB,
// det 22.6 KB, whose window also holds its counting sort).  Measured on 256
// networks: atomic 3282-3294 simulated seconds per second against 3200-3219
// with 8 entries and 4 spikes (the register savings remove its spill); det
// 2051-2055 against 2016-2022 with 8 entries (25.6 KB, 6 workgroups).
template <bool kDet>
constexpr int entries_batch() { return 6; }
template <bool kDet>
constexpr int spikes_batch() { return kDet ? 2 : 6; }
// How an entry of a window finds its spike: an owner map (batched slices,
// whose workgroups are issue-bound: a few instructions per entry) or a
// binary search over the segment offsets, all of a thread's entries in
// lockstep (single-network slices, one latency-bound workgroup per CU: no
// extra barriers).
constexpr bool kOwnerSingle = false, kOwnerBatch = true;

// LDS workspace of slice_accumulate.  acc[L] (floats, padded to 16 bytes),
// then, for a pass of cap = kSPT*B spikes and windows of Wd = kEPT*B
// entries:
//   base   u64 x cap   per spike: table index of its segment's first entry
//                      minus the segment's offset in the pass (the pass's
//                      spike ids are gathered in its first half)
//   scan   int x (B/64 + 1), padded to 16 bytes
//   search lookup: segment offsets int x cap, padded to 16 bytes
//   atomic, owner lookup: own u16 x Wd, twice (maps of even and odd windows)
//   det:    w f32 x Wd | keys u16 x Wd (the owner map too) | cursor, start int x L
//   atomic, wave-local lookup (search plans with at most 64*kSPT < 256
//           spikes per wave): a map u8 x kWaveMap per wave after the search
//           lookup's offsets; the offsets region holds each wave's spike ids
//           and the base region each wave's segment bases
//   det, wave-local (the same plans, slices below kDetLocalSlice): after
//           the cursors, 16-byte aligned, a map u8 x kDetMap per wave, then
//           a byte and a u16 count per (wave, target), then weights f32 for
//           kDetMap entries per wave, and 64 bytes that the sum's 16-wide
//           reads may pass over
constexpr int kWaveMap = 384;  // entries per wave-local chunk: 6 per lane
constexpr int kDetSlots = 6;   // det, wave-local: entries per lane, all in registers
constexpr int kDetMap = kDetSlots * 64;
__host__ __device__ constexpr bool wave_local_plan(bool det, int spt, bool owner)
{
    return !det && !owner && 64 * spt < 256;
}
__host__ __device__ constexpr bool det_local_plan(bool det, int spt, bool owner)
{
    return det && !owner && 64 * spt < 256;
}
__host__ __device__ inline size_t det_local_offset(size_t det_end) { return (det_end + 15) & ~(size_t)15; }
// Det, wave-local: slices shorter than this (the per-(wave, target) bytes
// and counts grow with the slice; longer slices take the windows).
constexpr int kDetLocalSlice = 1024;
// The stable placement's 48 bytes per target follow the maps, then the weights.
__host__ __device__ inline size_t det_weights_offset(int L, int nw) { return (size_t)nw * kDetMap + (size_t)L * 48; }
__host__ __device__ inline int acc_floats(int L) { return (L + 3) & ~3; }
__host__ __device__ inline size_t scan_bytes(int block) { return ((size_t)(block / 64 + 1) * 4 + 15) & ~(size_t)15; }
__host__ __device__ inline size_t slice_lds_bytes(int L, int block, bool det, int ept, int spt, bool owner)
{
    const size_t cap = (size_t)spt * (size_t)block, wd = (size_t)ept * (size_t)block;
    size_t b = (size_t)acc_floats(L) * 4 + cap * 8 + scan_bytes(block);
    if (!owner) b += (cap * 4 + 15) & ~(size_t)15;
    if (det && det_local_plan(det, spt, owner) && L < kDetLocalSlice)
        return det_local_offset(b + wd * 6 + (size_t)L * 8) + det_weights_offset(L, block / 64) +
               (size_t)(block / 64) * kDetMap * 4 + 64;
    if (det) return b + wd * 6 + (size_t)L * 8;
    if (wave_local_plan(det, spt, owner)) b += (size_t)(block / 64) * kWaveMap;
    return b + (owner ? wd * 4 : 0);
}

// Single-network slice plans: large (slices of kLargeSlice neurons or more
// whose workspace fits the LDS) or the default.
template <bool kDet, bool kLarge>
struct SlicePlan {
    static constexpr int kSPT = kLarge ? (kDet ? kSpikesLargeDet : kSpikesLargeAtomic) : kSpikesSingle;
    static constexpr int kEPT = kLarge ? kEntriesLarge : kEntriesSingle;
    static constexpr bool kOwner = kLarge ? (kDet ? kOwnerLargeDet : kOwnerLargeAtomic) : kOwnerSingle;
    static size_t lds(int L, int block) { return slice_lds_bytes(L, block, kDet, kEPT, kSPT, kOwner); }
};
__host__ inline bool large_slice_plan(int L, int block, bool det)
{
    const size_t lds = det ? SlicePlan<true, true>::lds(L, block) : SlicePlan<false, true>::lds(L, block);
    return L >= kLargeSlice && lds <= kSliceLdsMax;
}
__host__ inline size_t single_slice_lds(int L, int block, bool det)
{
    const bool large = large_slice_plan(L, block, det);
    if (det) return large ? SlicePlan<true, true>::lds(L, block) : SlicePlan<true, false>::lds(L, block);
    return large ? SlicePlan<false, true>::lds(L, block) : SlicePlan<false, false>::lds(L, block);
}

// One {target, weight bits} table entry (generic or global pointer).
__device__ __forceinline__ uint2 entry_at(const uint2 *p, size_t i) { return p[i]; }
__device__ __forceinline__ uint2 entry_at(__attribute__((address_space(1))) const unsigned long long *p, size_t i)
{
    const unsigned long long x = p[i];
    return make_uint2((uint32_t)x, (uint32_t)(x >> 32));
}

// Zero this thread's kEPT positions of an owner map.
template <int kEPT>
__device__ __forceinline__ void clear_owner(uint16_t *own)
{
    // The zero passes through an empty asm operand so that it is made here:
    // hoisted out of the window loop, the zero vector of the 16-byte store
    // was kept in scratch memory.
    uint32_t zero = 0u;
    asm volatile("" : "+v"(zero));
    uint32_t *z = (uint32_t *)(own + (size_t)threadIdx.x * kEPT);
#pragma unroll
    for (int q = 0; q < kEPT / 2; ++q) z[q] = zero;
}

// Owner map of a window: on entry own[] holds, at the window offset of each
// non-empty segment's first entry, that segment's spike index, and zero
// elsewhere (non-empty segments start at distinct offsets, and spike
// indices grow with the offset).  A max-scan over the window, started from
// *carry (the owner of the previous window's last entry), leaves in own[e]
// the spike whose segment holds entry f0+e, and sets *carry for the next
// window.  Thread i scans positions [i*kEPT, (i+1)*kEPT).  Begins with the
// owner-map reads and ends with a workgroup barrier.
template <int kEPT>
__device__ __forceinline__ void owner_scan(uint16_t *own, int *s_wave, uint32_t *carry)
{
    const int lane = lane_id();
    const int wave = (int)threadIdx.x / kWave;
    const int nw = (int)blockDim.x / kWave;
    uint32_t *ow = (uint32_t *)(own + (size_t)threadIdx.x * kEPT);
    uint32_t raw[kEPT / 2];
    uint32_t run = 0;
#pragma unroll
    for (int q = 0; q < kEPT / 2; ++q) {
        raw[q] = ow[q];
        run = max(run, max(raw[q] & 0xFFFFu, raw[q] >> 16));
    }
    const uint32_t incl = wave_inclusive_max(run);
    uint32_t pre = (uint32_t)__shfl_up((int)incl, 1);
    if (lane == 0) pre = 0;
    if (lane == kWave - 1) s_wave[wave] = (int)incl;
    __syncthreads();
    uint32_t tot = 0;
    pre = max(pre, max(*carry, waves_prefix_max(s_wave, &tot)));
    tot = max(tot, *carry);
#pragma unroll
    for (int q = 0; q < kEPT / 2; ++q) {
        const uint32_t lo16 = max(pre, raw[q] & 0xFFFFu);
        pre = max(lo16, raw[q] >> 16);
        ow[q] = lo16 | (pre << 16);
    }
    *carry = tot;
    __syncthreads();
}

// acc[e.x - lo] += weight of e for each entry e of x (target 0xFFFFFFFF:
// none): binary32 adds applied atomically, in any order, as compare-and-swap
// loops on the bit patterns.  Every first attempt of the thread is issued
// before any result is looked at, and a lost race retries from the value it
// saw.  On MI355X this runs the batched step 1.5x faster than the LDS float
// atomic add (ds_add_f32).
template <int K>
__device__ __forceinline__ void lds_add_cas(float *acc, const uint2 (&x)[K], int lo)
{
    uint32_t *ai = (uint32_t *)acc;
    uint32_t seen[K], got[K];
#pragma unroll
    for (int k = 0; k < K; ++k) seen[k] = x[k].x != 0xFFFFFFFFu ? ai[(int)x[k].x - lo] : 0u;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        got[k] = seen[k];
        if (x[k].x != 0xFFFFFFFFu)
            got[k] = atomicCAS(&ai[(int)x[k].x - lo], seen[k], __float_as_uint(__uint_as_float(seen[k]) + u2f(x[k].y)));
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        uint32_t cur = got[k];
        while (cur != seen[k]) {  // lost a race: at least one other add landed
            seen[k] = cur;
            cur = atomicCAS(&ai[(int)x[k].x - lo], seen[k], __float_as_uint(__uint_as_float(seen[k]) + u2f(x[k].y)));
        }
    }
}

// LDS accesses of the other lanes of this wave that precede the call in
// program order are visible to the accesses that follow it.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Entries c0 + 64i + lane (i < K, below wlen) of a wave-local chunk: all map
// reads, then all base reads, then all table loads are issued before the
// first wait (the reads are unconditional; a slot past wlen reads spike 0's
// base and loads nothing), then the adds.
template <int K, class Net>
__device__ __forceinline__ void wave_chunk(const Net &d, int lo, float *acc, int c0, int wlen, const uint64_t *wbs,
                                           const uint8_t *wmap)
{
    const int lane = lane_id();
    uint32_t jm[K];
#pragma unroll
    for (int i = 0; i < K; ++i) jm[i] = wmap[kWave * i + lane];
    uint64_t idx[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const int f = c0 + kWave * i + lane;
        idx[i] = wbs[f < wlen ? jm[i] : 0u] + (uint64_t)f;
    }
    uint2 x[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        x[i] = make_uint2(0xFFFFFFFFu, 0u);
        if (c0 + kWave * i + lane < wlen) x[i] = entry_at(d.syn, (size_t)idx[i]);
    }
    lds_add_cas<K>(acc, x, lo);
}

// Atomic-mode accumulation of the spikes of one wave's bitmap groups (its
// lane's groups g0 and g0 + 64 hold kept0 and kept1), with no exchange
// between the waves of the workgroup: the adds may land in any order, so
// each wave takes the entries of its own spikes.  Per batch of up to
// 64*kSPT of the wave's spikes (one batch unless the wave holds more):
//  1. the spike ids, ascending, into the wave's id array (wids);
//  2. lane per spike: the row's segment bounds for slice p, a wave scan of
//     the lengths numbering the concatenation of the batch's segments, and
//     per spike the table index of its entry 0 of that numbering (wbs);
//  3. per chunk of kWaveMap entries: every spike writes its index into the
//     map at the positions of its segment, then lane l takes entries
//     c0 + 64i + l: one map read, one base read, one table load each.
// Only wave-level ordering is needed (wave_sync); the caller brackets the
// call with workgroup barriers (acc cleared before, every add done after).
template <bool kCoh, int kSPT, class Net>
__device__ __forceinline__ void wave_accumulate(const Net &d, int p, int lo, float *acc, uint4 kept0, uint4 kept1,
                                                int g0, uint32_t *wids, uint64_t *wbs, uint8_t *wmap)
{
    constexpr int kWaveCap = kWave * kSPT;
    constexpr int kEPTw = kWaveMap / kWave;
    static_assert(kWaveCap <= 256, "wave-local spike indices are bytes");
    const int lane = lane_id();
    const uint32_t stride = (uint32_t)d.P + 1u;
    const int pk = popc4(kept0) | (popc4(kept1) << 16);
    const int incl = wave_inclusive_sum(pk);
    const int tot = wave_last(incl);
    const int T0 = tot & 0xFFFF;
    const int wtot = T0 + (tot >> 16);
    const int ex = incl - pk;
    const int i0 = ex & 0xFFFF, i1 = T0 + (ex >> 16);
    for (int b0 = 0; b0 < wtot; b0 += kWaveCap) {  // uniform over the wave
        put_group<false>(kept0, g0, i0, wids, b0, b0 + kWaveCap);
        put_group<false>(kept1, g0 + kWave, i1, wids, b0, b0 + kWaveCap);
        wave_sync();
        const int nb = min(kWaveCap, wtot - b0);
        // a lane without a spike loads row 0's bounds (in range) and keeps
        // length 0, so that every bound load issues before the first wait
        uint32_t sid[kSPT], a0[kSPT], a1[kSPT];
#pragma unroll
        for (int q = 0; q < kSPT; ++q) sid[q] = lane + kWave * q < nb ? wids[lane + kWave * q] : 0u;
#pragma unroll
        for (int q = 0; q < kSPT; ++q) {
            const auto o = d.soff + ((uint64_t)sid[q] * stride + (uint32_t)p);
            a0[q] = o[0];
            a1[q] = o[1];
        }
        int len[kSPT], run[kSPT];
        int wlen = 0;
#pragma unroll
        for (int q = 0; q < kSPT; ++q) {
            len[q] = lane + kWave * q < nb ? (int)(a1[q] - a0[q]) : 0;
            const int inc = wave_inclusive_sum(len[q]);
            run[q] = wlen + inc - len[q];
            wlen += wave_last(inc);
        }
#pragma unroll
        for (int q = 0; q < kSPT; ++q)
            if (len[q] > 0)
                wbs[lane + kWave * q] = (uint64_t)sid[q] * (uint64_t)d.nsyn + (uint64_t)a0[q] - (uint64_t)run[q];
        for (int c0 = 0; c0 < wlen; c0 += kWaveMap) {  // uniform over the wave
#pragma unroll
            for (int q = 0; q < kSPT; ++q) {
                const int s = max(run[q], c0), e = min(run[q] + len[q], c0 + kWaveMap);
                for (int f = s; f < e; ++f) wmap[f - c0] = (uint8_t)(lane + kWave * q);
            }
            wave_sync();
            if (wlen - c0 <= kWaveMap / 2)  // uniform: half the slots hold every entry
                wave_chunk<kEPTw / 2>(d, lo, acc, c0, wlen, wbs, wmap);
            else
                wave_chunk<kEPTw>(d, lo, acc, c0, wlen, wbs, wmap);
            wave_sync();  // the map, ids and bases are rewritten next
        }
    }
}

// Det mode, wave-local: one wave's spikes and entries, all in registers.
// The same steps as wave_accumulate (ids, segment bounds, map) for a single
// batch and a single chunk of at most kDetSlots*64 entries; entry
// c0 + 64i + lane lands in x[i] (target 0xFFFFFFFF past the wave's last).
// Returns, uniform over the wave: 0 if the wave holds more than 64*kSPT
// spikes or more entries than that (nothing is loaded then), 1 if it holds
// no entry, 2 if its entries' loads are in flight.
template <bool kCoh, int kSPT, class Net>
__device__ __forceinline__ int wave_collect(const Net &d, int p, uint4 kept0, uint4 kept1, int g0, uint32_t *wids,
                                             uint64_t *wbs, uint8_t *wmap, uint2 (&x)[kDetSlots])
{
    constexpr int kWaveCap = kWave * kSPT;
    static_assert(kWaveCap <= 256, "wave-local spike indices are bytes");
    const int lane = lane_id();
    const uint32_t stride = (uint32_t)d.P + 1u;
#pragma unroll
    for (int i = 0; i < kDetSlots; ++i) x[i] = make_uint2(0xFFFFFFFFu, 0u);
    const int pk = popc4(kept0) | (popc4(kept1) << 16);
    const int incl = wave_inclusive_sum(pk);
    const int tot = wave_last(incl);
    const int T0 = tot & 0xFFFF;
    const int wtot = T0 + (tot >> 16);
    if (wtot > kWaveCap) return 0;  // uniform
    if (wtot == 0) return 1;
    const int ex = incl - pk;
    put_group<true>(kept0, g0, ex & 0xFFFF, wids, 0, 0);
    put_group<true>(kept1, g0 + kWave, T0 + (ex >> 16), wids, 0, 0);
    wave_sync();
    uint32_t sid[kSPT], a0[kSPT], a1[kSPT];
#pragma unroll
    for (int q = 0; q < kSPT; ++q) sid[q] = lane + kWave * q < wtot ? wids[lane + kWave * q] : 0u;
#pragma unroll
    for (int q = 0; q < kSPT; ++q) {
        const auto o = d.soff + ((uint64_t)sid[q] * stride + (uint32_t)p);
        a0[q] = o[0];
        a1[q] = o[1];
    }
    int len[kSPT], run[kSPT];
    int wlen = 0;
#pragma unroll
    for (int q = 0; q < kSPT; ++q) {
        len[q] = lane + kWave * q < wtot ? (int)(a1[q] - a0[q]) : 0;
        const int inc = wave_inclusive_sum(len[q]);
        run[q] = wlen + inc - len[q];
        wlen += wave_last(inc);
    }
    if (wlen > kDetMap) return 0;  // uniform
    if (wlen == 0) return 1;
#pragma unroll
    for (int q = 0; q < kSPT; ++q)
        if (len[q] > 0) {
            wbs[lane + kWave * q] = (uint64_t)sid[q] * (uint64_t)d.nsyn + (uint64_t)a0[q] - (uint64_t)run[q];
            for (int f = run[q]; f < run[q] + len[q]; ++f) wmap[f] = (uint8_t)(lane + kWave * q);
        }
    wave_sync();
    uint32_t jm[kDetSlots];
#pragma unroll
    for (int i = 0; i < kDetSlots; ++i) jm[i] = kWave * i + lane < wlen ? (uint32_t)wmap[kWave * i + lane] : 0u;
    uint64_t idx[kDetSlots];
#pragma unroll
    for (int i = 0; i < kDetSlots; ++i) idx[i] = wbs[jm[i]] + (uint64_t)(kWave * i + lane);
#pragma unroll
    for (int i = 0; i < kDetSlots; ++i)
        if (kWave * i + lane < wlen) x[i] = entry_at(d.syn, (size_t)idx[i]);
    return 2;
}

// Det mode, last phase: the bucket dw[b0 .. b1) added in order to a, the
// weights read 16 at a time (reads may pass the bucket's ends, within the
// workspace).
__device__ __forceinline__ float bucket_sum(const float *dw, int b0, int b1, float a)
{
    for (int q0 = b0 & ~3; q0 < b1; q0 += 16) {
        float4 wv[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) wv[c] = *(const float4 *)(dw + q0 + 4 * c);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float h[4] = {wv[c].x, wv[c].y, wv[c].z, wv[c].w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int q = q0 + 4 * c + e;
                if (q >= b0 && q < b1) a = a + h[e];
            }
        }
    }
    return a;
}
// Each target j in [tb, te) adds its bucket dw[dstart[j] .. dend[j]); dend
// is left zero.  Ends with a workgroup barrier.
__device__ __forceinline__ void bucket_sums(float *acc, const float *dw, const int *dstart, int *dend, int tb, int te)
{
    for (int j = tb; j < te; ++j) {
        acc[j] = bucket_sum(dw, dstart[j], dend[j], acc[j]);
        dend[j] = 0;
    }
    __syncthreads();
}

// Det mode, stable placement: acc[j] <- acc[j] plus the entries of every
// wave (x[i] is entry 64i + lane of this wave's concatenation, target j +
// lo) added in the order (wave, i, lane), which is ascending source order:
// each wave's spikes are ascending and precede the next wave's, and a
// wave's entries follow its spikes.  No sort by key: an entry's place in
// its target's bucket is the number of that target's entries in earlier
// waves plus those before it in this wave (earlier slots, then lower lanes
// of its own slot).  Per slot, the lanes that share a target find each other
// through a byte per (wave, target) (each stores its lane id, all read back
// the same one of them) and ballots; a count per (wave, target) carries the
// place across slots.
// reg (48*L bytes, 16-byte aligned): per target a row of 16 bytes in each of
// two byte arrays -- lane ids at 0, counts (zero on entry) at 32*L -- wave w
// at byte (w + j) & 15 of row j, so that a wave's accesses to different
// targets spread over the LDS banks and a target's counts are one 16-byte
// load.  After the scan, per target a row of 16 u16 places (the same
// rotation) at 0, over the lane ids and the 16*L bytes after them.  dstart
// and dend get the buckets' bounds.  Ends with a workgroup barrier if the
// slice is longer than the workgroup; otherwise thread j has summed target j,
// the neuron its update reads, and the caller's next barrier orders the rest.
template <int K>
__device__ __forceinline__ void det_stable_add(float *acc, int lo, const uint2 (&x)[K], uint8_t *reg, int L,
                                               int *dstart, int *dend, float *dw, int *scratch, int tb, int te)
{
    const int lane = lane_id(), wave = wave_index();
    uint8_t *rep8 = reg;
    uint8_t *cnt8 = reg + 32 * L;
    uint16_t *off16 = (uint16_t *)reg;
    int rk[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        __builtin_amdgcn_sched_barrier(0);  // one slot at a time
        const bool valid = x[i].x != 0xFFFFFFFFu;
        const int j = valid ? (int)x[i].x - lo : 0;
        const int a = 16 * j + ((wave + j) & 15);
        if (valid) rep8[a] = (uint8_t)lane;
        // the byte read back may be another lane's store (the wave-level
        // sync also keeps the compiler from reusing this lane's own value)
        wave_sync();
        const int rep = valid ? (int)rep8[a] : -1;
        const int prev = valid ? (int)cnt8[a] : 0;
        uint64_t mine = valid ? (1ull << lane) : 0ull;
        uint64_t pend = __ballot(valid && rep != lane);  // lanes whose target another lane shares
        while (pend) {  // uniform: one pass per shared target
            const int r = __builtin_amdgcn_readlane(rep, (int)__builtin_ctzll(pend));
            const uint64_t grp = __ballot(valid && rep == r);
            if (valid && rep == r) mine = grp;
            pend &= ~grp;
        }
        const int below = bits_below(mine), n = __popcll(mine);
        rk[i] = prev + below;
        if (valid && below == n - 1) cnt8[a] = (uint8_t)(prev + n);  // a wave holds at most 128 spikes
    }
    __syncthreads();
    // buckets: each target's total, a scan over the targets, then each
    // wave's first place in the bucket
    auto bytes4 = [](uint32_t v) {
        const uint32_t t = (v & 0x00FF00FFu) + ((v >> 8) & 0x00FF00FFu);
        return (int)((t & 0xFFFFu) + (t >> 16));
    };
    int csum = 0;
    for (int j = tb; j < te; ++j) {
        const uint4 row = *(const uint4 *)(cnt8 + 16 * j);
        csum += (bytes4(row.x) + bytes4(row.y)) + (bytes4(row.z) + bytes4(row.w));
    }
    int all = 0;
    int st = block_exclusive_scan(csum, scratch, &all);
    // a slice no longer than the workgroup: thread j's one target is j, its
    // bucket bounds stay in registers and its input starts from 0.0f (acc
    // is cleared on entry to the accumulation)
    const bool own = L <= (int)blockDim.x;  // uniform
    int b0own = 0, b1own = 0;
    for (int j = tb; j < te; ++j) {
        const uint4 row = *(const uint4 *)(cnt8 + 16 * j);
        // rotate the row right by (j & 15) bytes: byte w is then wave w's count
        const int q = (j >> 2) & 3, sh = j & 3;
        const uint32_t r4[4] = {row.x, row.y, row.z, row.w};
        uint32_t dq[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            dq[k] = q == 0 ? r4[k] : q == 1 ? r4[(k + 1) & 3] : q == 2 ? r4[(k + 2) & 3] : r4[(k + 3) & 3];
        uint32_t c[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) c[k] = __builtin_amdgcn_alignbyte(dq[(k + 1) & 3], dq[k], (uint32_t)sh);
        if (own) b0own = st;
        else dstart[j] = st;
        // the places in wave order, two per register, then rotated left by
        // (j & 15) u16 back to the row's order and stored as two 16-byte words
        uint32_t pl[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t a = (uint32_t)st;
            st += (int)((c[k >> 1] >> (16 * (k & 1))) & 0xFFu);
            pl[k] = a | ((uint32_t)st << 16);
            st += (int)((c[k >> 1] >> (16 * (k & 1) + 8)) & 0xFFu);
        }
        const int h = (j >> 1) & 7;
#pragma unroll
        for (int b = 1; b < 8; b <<= 1)
            if (h & b) {  // rotate the eight words up by b
                uint32_t t[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) t[k] = pl[(k - b) & 7];
#pragma unroll
                for (int k = 0; k < 8; ++k) pl[k] = t[k];
            }
        if (j & 1) {
            uint32_t t[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) t[k] = __builtin_amdgcn_alignbit(pl[k], pl[(k + 7) & 7], 16u);
#pragma unroll
            for (int k = 0; k < 8; ++k) pl[k] = t[k];
        }
        uint4 *o = (uint4 *)(off16 + 16 * j);
        o[0] = make_uint4(pl[0], pl[1], pl[2], pl[3]);
        o[1] = make_uint4(pl[4], pl[5], pl[6], pl[7]);
        if (own) b1own = st;
        else dend[j] = st;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < K; ++i)
        if (x[i].x != 0xFFFFFFFFu) {
            const int j = (int)x[i].x - lo;
            dw[(int)off16[16 * j + ((wave + j) & 15)] + rk[i]] = u2f(x[i].y);
        }
    __syncthreads();
    if (own) {
        if (tb < te) acc[tb] = bucket_sum(dw, b0own, b1own, 0.0f);
    } else {
        bucket_sums(acc, dw, dstart, dend, tb, te);
    }
}

// Det mode, one window of a four-wave workgroup whose wave w holds window
// entries (w*K + i)*64 + lane in x[i] (targets j + lo, 0xFFFFFFFF for none):
// acc[j] <- acc[j] plus the window's entries for j in window order, which is
// (wave, i, lane).  Placement as in det_stable_add, with one byte per (wave,
// target): the lanes' ids in dw (free until the weights are placed; 4*len
// bytes), the counts in dcur[j] (zero on entry, byte w for wave w).  The scan
// then leaves in dstart[j] and dcur[j] the first places of waves 0, 1 and 2,
// 3 (u16 each), and target j's bucket ends where target j+1's starts.
// Leaves dcur zero; the caller ends the window with a workgroup barrier.
// A count beyond a byte (256 of a wave's entries for one target: segments
// of about one entry, all to one neuron) is flagged instead: the call then
// clears the counts and returns false, and the caller sorts the window.
// ovf: an LDS flag, zero on entry and on return.
template <int K>
__device__ __forceinline__ bool det_window_stable(float *acc, int lo, int len, const uint2 (&x)[K], float *dw,
                                                  int *dstart, int *dcur, int *scratch, int tb, int te, int *ovf)
{
    const int lane = lane_id(), wave = wave_index();
    uint8_t *rep8 = (uint8_t *)dw;
    uint8_t *cnt8 = (uint8_t *)dcur;
    int rk[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        __builtin_amdgcn_sched_barrier(0);  // one slot at a time
        const bool valid = x[i].x != 0xFFFFFFFFu;
        const int a = valid ? 4 * ((int)x[i].x - lo) + wave : 0;
        if (valid) rep8[a] = (uint8_t)lane;
        wave_sync();  // another lane's id may be read back (and this lane's own is not reused)
        const int rep = valid ? (int)rep8[a] : -1;
        const int prev = valid ? (int)cnt8[a] : 0;
        uint64_t mine = valid ? (1ull << lane) : 0ull;
        uint64_t pend = __ballot(valid && rep != lane);
        while (pend) {  // uniform: one pass per shared target
            const int r = __builtin_amdgcn_readlane(rep, (int)__builtin_ctzll(pend));
            const uint64_t grp = __ballot(valid && rep == r);
            if (valid && rep == r) mine = grp;
            pend &= ~grp;
        }
        const int below = bits_below(mine), n = __popcll(mine);
        rk[i] = prev + below;
        if (valid && below == n - 1) {
            cnt8[a] = (uint8_t)(prev + n);
            if (prev + n > 255) *ovf = 1;
        }
    }
    __syncthreads();
    if (*ovf) {  // uniform
        for (int j = tb; j < te; ++j) dcur[j] = 0;
        __syncthreads();  // every thread has read the flag
        if (threadIdx.x == 0) *ovf = 0;
        __syncthreads();
        return false;
    }
    int csum = 0;
    for (int j = tb; j < te; ++j) {
        const uint32_t c = (uint32_t)dcur[j];
        csum += (int)((c & 0xFFu) + ((c >> 8) & 0xFFu) + ((c >> 16) & 0xFFu) + (c >> 24));
    }
    int nwin = 0;
    int st = block_exclusive_scan(csum, scratch, &nwin);
    for (int j = tb; j < te; ++j) {
        const uint32_t c = (uint32_t)dcur[j];
        const uint32_t p0 = (uint32_t)st, p1 = p0 + (c & 0xFFu), p2 = p1 + ((c >> 8) & 0xFFu), p3 = p2 + ((c >> 16) & 0xFFu);
        st = (int)(p3 + (c >> 24));
        dstart[j] = (int)(p0 | (p1 << 16));
        dcur[j] = (int)(p2 | (p3 << 16));
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < K; ++i)
        if (x[i].x != 0xFFFFFFFFu) {
            const int j = (int)x[i].x - lo;
            const uint32_t pw = (uint32_t)(wave < 2 ? dstart[j] : dcur[j]);
            dw[(int)((pw >> (16 * (wave & 1))) & 0xFFFFu) + rk[i]] = u2f(x[i].y);
        }
    __syncthreads();
    for (int j = tb; j < te; ++j) {
        const int s0 = dstart[j] & 0xFFFF;
        const int s1 = j + 1 < len ? (dstart[j + 1] & 0xFFFF) : nwin;
        float a = acc[j];
        for (int q = s0; q < s1; ++q) a = a + dw[q];
        acc[j] = a;
        dcur[j] = 0;
    }
    return true;
}

// acc[0..len) <- input of slice p from the rows that spiked at step t-1.
// Three dependent global levels, each with all its loads in flight:
//  1. the step's bitmap, compacted in LDS into the ascending spike list
//     (the lanes of a wave own consecutive groups of four words);
//  2. thread per spike: the row's segment bounds for slice p (offset table),
//     and a scan of the segment lengths gives each segment's offset in the
//     concatenation of the pass's segments;
//  3. thread per synapse entry of that concatenation, by windows of Wd
//     entries: an owner map (segment heads, then a max-scan) or a binary
//     search over the segment offsets names each entry's spike, and one
//     table load per entry follows.
// Spikes beyond kSPT*B are handled in further passes.  dyn is the LDS
// workspace (acc first).  Ends with a workgroup barrier, except after the
// det wave-local sums of a slice no longer than the workgroup, where each
// thread's acc[tid] is its own sum (the update reads only that).
// kDet = false: binary32 adds applied atomically (compare-and-swap), any order.
// kDet = true: every acc[j] is summed in ascending presynaptic id from 0.0f,
// bit-equal to a sequential push.  The entries of a pass are in ascending
// (source, target) order; each window of them is counting-sorted by target
// in LDS (count, scan, place) and each target then adds its entries in key
// order, so the sum never depends on the order of the LDS atomics.
// kCoh: bitmap words are read with agent-scope coherent loads (written by
// other workgroups of the same launch).  kZeroed: acc[0..len) is zero on
// entry, ordered before the call by a workgroup barrier (the caller clears
// it as it reads it).  pre() is called once, by every thread,
// while the first bitmap loads are in flight.
struct NoPre {
    __device__ void operator()() const {}
};
// backoff (a persistent caller's register, or null): after a failed fast
// lookup the next kFastBackoff steps take the general path directly.
constexpr int kFastBackoff = 32;
template <bool kDet, int kEPT, int kSPT, bool kOwner, bool kCoh = false, bool kZeroed = false, class Net,
          class Pre = NoPre>
__device__ __forceinline__ void slice_accumulate(const Net &d, int t, int p, int lo, int len, float *dyn,
                                                 Pre pre = Pre(), int *backoff = nullptr)
{
    const int B = (int)blockDim.x;
    const int tid = (int)threadIdx.x;
    const int cap = kSPT * B;
    const int Wd = kEPT * B;
    const int L = d.L;
    float *acc = dyn;
    char *w8 = (char *)(dyn + acc_floats(L));
    uint64_t *base = (uint64_t *)w8;
    uint32_t *ids = (uint32_t *)w8;
    int *scratch = (int *)(w8 + (size_t)cap * 8);
    char *w1 = w8 + (size_t)cap * 8 + scan_bytes(B);
    int *segp = (int *)w1;  // search lookup
    char *w2 = kOwner ? w1 : w1 + (((size_t)cap * 4 + 15) & ~(size_t)15);
    float *dw = (float *)w2;
    uint16_t *own0 = (uint16_t *)(kDet ? w2 + (size_t)Wd * 4 : w2);
    uint16_t *own1 = kDet ? own0 : own0 + Wd;
    uint16_t *dkey = (uint16_t *)(w2 + (size_t)Wd * 4);  // det: also the owner map
    int *dcur = (int *)(w2 + (size_t)Wd * 6);
    int *dstart = dcur + L;
    __shared__ int s_ovf;  // det windows: a wave's count beyond a byte (det_window_stable)
    if (kDet && tid == 0) s_ovf = 0;
    if (!kZeroed)
        for (int i = tid; i < len; i += B) {
            acc[i] = 0.0f;
            if (kDet) dcur[i] = 0;
        }
    if (t == 0) {
        pre();
        __syncthreads();
        return;
    }
    const auto bm = ((t - 1) & 1) ? d.bm1 : d.bm0;
    const int W = d.words;
    // Bitmap ownership: groups of four consecutive words, one 16-byte load
    // each, and the lanes of a wave take consecutive groups, so that one
    // load instruction of a wave reads up to 1 KB of consecutive words
    // (whole 128-byte lines; a run of words per thread would touch one line
    // per lane, and coherent loads are not merged in the caches).  The
    // groups are spread evenly over the waves: wave v owns Gw = ceil(groups
    // / waves) of them from v*Gw, and its lane l takes groups v*Gw + 64k +
    // l for 64k + l < Gw, k < G = ceil(Gw / 64).  The ascending word order
    // is (wave, k, lane, word of the group).  The first two groups of a
    // lane stay in registers for the id pass below; longer bitmaps (large
    // networks) are read again two groups at a time.
    const int lane = lane_id(), wave = tid / kWave;
    const int nwaves = B / kWave;
    const int Gw = ((W + 3) / 4 + nwaves - 1) / nwaves;
    const int G = (Gw + kWave - 1) / kWave;
    const int g0 = Gw * wave + lane;
    const auto rs = bm_rsrc(bm, W);
    const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
    auto group = [&](int k) { return kWave * k + lane < Gw ? bm_group<kCoh>(rs, g0 + kWave * k, W) : zero4; };
    const uint4 kept0 = group(0);
    const uint4 kept1 = group(1);
    // the caller's work that needs no spikes (the neurons' thalamic draws)
    // runs while the bitmap loads are in flight
    pre();
    // targets owned by this thread in the ordered phases: a contiguous run
    const int tper = (len + B - 1) / B;
    const int tb = min(tid * tper, len), te = min(tb + tper, len);
    constexpr bool kWaveLocal = wave_local_plan(kDet, kSPT, kOwner);
    if constexpr (kWaveLocal) {
        if (G <= 2) {  // uniform: every group of the wave is in registers
            if (!kZeroed) __syncthreads();  // acc is cleared before any wave adds into it
            uint8_t *wmap = (uint8_t *)(w1 + (((size_t)cap * 4 + 15) & ~(size_t)15)) + (size_t)wave * kWaveMap;
            wave_accumulate<kCoh, kSPT>(d, p, lo, acc, kept0, kept1, g0, (uint32_t *)segp + wave * kWave * kSPT,
                                        base + wave * kWave * kSPT, wmap);
            __syncthreads();
            return;
        }
    }
    constexpr bool kDetLocal = det_local_plan(kDet, kSPT, kOwner);
    if constexpr (kDetLocal) {
        // Every wave collects its own entries, and each entry's place in its
        // target's bucket follows from the wave, slot and lane that hold it
        // (det_stable_add).  If a wave holds more than fits in registers,
        // the windows below do the step instead.
        if (G <= 2 && L < kDetLocalSlice) {  // uniform
            __shared__ int s_fit[1024 / kWave];
            char *wl = (char *)dyn + det_local_offset((size_t)(w2 - (char *)dyn) + (size_t)Wd * 6 + (size_t)L * 8);
            uint8_t *reg = (uint8_t *)(wl + (size_t)nwaves * kDetMap);
            float *dw32 = (float *)(wl + det_weights_offset(L, nwaves));
            for (int i = tid; i < 4 * L; i += B) ((uint32_t *)(reg + 32 * L))[i] = 0u;  // the counts
            uint2 x[kDetSlots];
            // vote: 0 = does not fit, 1 = no entries, 2 = entries (their
            // loads stay in flight across the barrier)
            const int fit = wave_collect<kCoh, kSPT>(d, p, kept0, kept1, g0, (uint32_t *)segp + wave * kWave * kSPT,
                                                     base + wave * kWave * kSPT, (uint8_t *)wl + (size_t)wave * kDetMap,
                                                     x);
            if (lane == 0) s_fit[wave] = fit;
            __syncthreads();  // also orders the clearing of cursors and counts before their use
            const bool all_fit = __ballot(lane < nwaves && s_fit[lane] == 0) == 0;  // uniform
            if (all_fit) {
                if (__ballot(lane < nwaves && s_fit[lane] == 2) != 0)  // uniform: some input this step
                    det_stable_add<kDetSlots>(acc, lo, x, reg, L, dstart, dcur, dw32, scratch, tb, te);
                return;
            }
        }
    }
    int mine = popc4(kept0) + popc4(kept1);
    for (int k = 2; k < G; k += 2) {
        const uint4 xa = group(k), xb = group(k + 1);
        mine += popc4(xa) + popc4(xb);
    }
    const uint32_t stride = (uint32_t)d.P + 1u;
    int total = 0;
    int wrun = 0, wsum = 0;  // general path: spikes before this wave, in it
    uint32_t sid[kSPT], e0[kSPT];
    int lens[kSPT], head[kSPT];
    int m = 0, E = 0;
    // Fast lookup (search plans whose every bitmap word sits in registers):
    // each wave compacts its own spikes into a private part of the ids array
    // and its lanes load their segment bounds at once; one workgroup pass
    // over the waves' (spikes, entries) totals then places every spike.  Two
    // workgroup barriers before the entries instead of four, and the offset
    // loads go out before the first of them.  A wave with more than 64*kSPT
    // spikes sends the whole workgroup down the general path below (at the
    // cost of the two barriers); a persistent caller then skips the attempt
    // for kFastBackoff steps, since busy steps come in runs.
    constexpr bool kFast = !kOwner && kWave * kSPT < 256 && !kWaveLocal;
    bool fast = false;
    if constexpr (kFast) {
        const bool attempt = G <= 2 && (backoff == nullptr || *backoff == 0);  // uniform
        if (backoff != nullptr && *backoff > 0) --*backoff;
        if (attempt) {
            const int nw = nwaves;
            const int pk = popc4(kept0) | (popc4(kept1) << 16);
            const int incl = wave_inclusive_sum(pk);
            const int tot = wave_last(incl);
            const int wtot = (tot & 0xFFFF) + (tot >> 16);
            constexpr int kWaveCap = kWave * kSPT;
            const bool wfit = wtot <= kWaveCap;  // uniform over the wave
            uint32_t *wids = ids + wave * kWaveCap;
            if (wfit && mine > 0) {  // both groups' ids, as put_pair places them
                const int ex = incl - pk;
                put_group<true>(kept0, g0, ex & 0xFFFF, wids, 0, 0);
                put_group<true>(kept1, g0 + kWave, (tot & 0xFFFF) + (ex >> 16), wids, 0, 0);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // lane l takes the wave's spikes l, l + 64, ...; in a wave with
            // spikes that fits, a lane without a spike loads row 0's bounds
            // (in range) and keeps length 0, so that every load of the wave
            // issues before the first wait; other waves load nothing
            uint32_t e1[kSPT];
            bool has[kSPT];
#pragma unroll
            for (int q = 0; q < kSPT; ++q) {
                has[q] = wfit && lane + kWave * q < wtot;
                sid[q] = has[q] ? wids[lane + kWave * q] : 0u;
                e0[q] = e1[q] = 0u;
            }
            if (wfit && wtot > 0) {  // uniform over the wave
#pragma unroll
                for (int q = 0; q < kSPT; ++q) {
                    const auto o = d.soff + ((uint64_t)sid[q] * stride + (uint32_t)p);
                    e0[q] = o[0];
                    e1[q] = o[1];
                }
            }
            int runq[kSPT];
            int wlen = 0;
#pragma unroll
            for (int q = 0; q < kSPT; ++q) {
                lens[q] = has[q] ? (int)(e1[q] - e0[q]) : 0;
                const int inc = wave_inclusive_sum(lens[q]);
                runq[q] = wlen + inc - lens[q];  // entries before this spike within the wave
                wlen += wave_last(inc);
            }
            // per wave: entries (23 bits) and spikes (8 bits), or the top
            // bit and the spikes (31 bits) when the wave does not fit
            const bool packs = wfit && wlen < (1 << 23);
            if (lane == 0) scratch[wave] = packs ? (int)(((uint32_t)wlen << 8) | (uint32_t)wtot) : -1;
            __syncthreads();
            // lane l decodes wave l's totals; wave scans place this wave
            const uint32_t xl = lane < nw ? (uint32_t)scratch[lane] : 0u;
            const bool over = xl == 0xFFFFFFFFu;
            const bool all_fit = __ballot(over) == 0;  // uniform
            int sp_all = 0, en_all = 0;
            const int sp_before = waves_prefix_sum_of(over ? 0 : (int)(xl & 0xFFu), &sp_all);
            const int en_before = waves_prefix_sum_of(over ? 0 : (int)(xl >> 8), &en_all);
            if (!all_fit && backoff != nullptr) *backoff = kFastBackoff;
            if (all_fit) {
                fast = true;
                total = m = sp_all;
                E = en_all;
                // base[i] overwrites ids[2i..2i+1]: every id was read before
                // the barrier above
#pragma unroll
                for (int q = 0; q < kSPT; ++q) {
                    head[q] = -1;
                    const int j = lane + kWave * q;
                    if (j < wtot) {
                        const int i = sp_before + j, run = en_before + runq[q];
                        base[i] = (uint64_t)sid[q] * (uint64_t)d.nsyn + (uint64_t)e0[q] - (uint64_t)run;
                        segp[i] = run;
                    }
                }
            }
            // placed spikes are visible to every thread; or, on the general
            // path, every thread has read the totals before its scan reuses
            // the scratch ints.  With no spike at all, the closing barrier
            // below is the only one needed.
            if (!(all_fit && sp_all == 0)) __syncthreads();  // uniform
        }
    }
    if (!fast) {
        // spikes before this wave: a scan over the waves' totals
        wsum = wave_last(wave_inclusive_sum(mine));
        if (lane == kWave - 1) scratch[wave] = wsum;
        __syncthreads();
        wrun = waves_prefix_sum(scratch, &total);
    }
    // 1. spike ids of the first pass, ascending, from the kept groups; later
    // groups and later passes read the bitmap again.  The conditions are
    // uniform over the wave (put_pair scans it).
    if (!fast && wsum > 0 && wrun < cap) {
        int run = put_pair<false>(kept0, kept1, g0, wrun, ids, 0, cap);
        for (int k = 2; k < G && run < cap; k += 2) {
            run = put_pair<false>(group(k), group(k + 1), g0 + kWave * k, run, ids, 0, cap);
        }
    }
    for (int pass = 0; pass < total; pass += cap) {
        if (!fast) {  // the general path places the pass's spikes
        if (pass > 0 && wsum > 0 && wrun < pass + cap && wrun + wsum > pass) {  // uniform over the wave
            int run = wrun;
            for (int k = 0; k < G && run < pass + cap; k += 2) {
                run = put_pair<false>(group(k), group(k + 1), g0 + kWave * k, run, ids, pass, pass + cap);
            }
        }
        __syncthreads();
        m = min(cap, total - pass);
        // 2. segment bounds, thread per spike
        int lsum = 0;
        if (tid * kSPT >= m) {
#pragma unroll
            for (int k = 0; k < kSPT; ++k) {
                sid[k] = e0[k] = 0;
                lens[k] = 0;
            }
        } else {
            // rows clamped to the pass's last spike: every load issues
            // before the first wait
            uint32_t e1[kSPT];
#pragma unroll
            for (int k = 0; k < kSPT; ++k) sid[k] = ids[min(tid * kSPT + k, m - 1)];
#pragma unroll
            for (int k = 0; k < kSPT; ++k) {
                const auto o = d.soff + ((uint64_t)sid[k] * stride + (uint32_t)p);
                e0[k] = o[0];
                e1[k] = o[1];
            }
#pragma unroll
            for (int k = 0; k < kSPT; ++k) lens[k] = tid * kSPT + k < m ? (int)(e1[k] - e0[k]) : 0;
        }
#pragma unroll
        for (int k = 0; k < kSPT; ++k) lsum += lens[k];
        if (kOwner) clear_owner<kEPT>(own0);  // the first window's map, before the scan's barriers
        E = 0;
        int run = block_exclusive_scan(lsum, scratch, &E);
        // base[i] overwrites ids[2i..2i+1]: every id was read before the scan
#pragma unroll
        for (int k = 0; k < kSPT; ++k) {
            const int i = tid * kSPT + k;
            // offset of the segment's first entry in the pass, -1 if empty
            head[k] = lens[k] > 0 ? run : -1;
            if (i < m) {
                base[i] = (uint64_t)sid[k] * (uint64_t)d.nsyn + (uint64_t)e0[k] - (uint64_t)run;
                if (!kOwner) segp[i] = run;
            }
            run += lens[k];
        }
        if (!kOwner) __syncthreads();
        }  // general path
        // 3. entries, thread per synapse, windows of Wd entries with all of a
        // window's loads in flight.  Lanes take consecutive entries
        // (f = f0 + k*B + tid): the loads coalesce, and a target's
        // contributions reach the LDS atomics nearly in ascending source
        // order, which keeps atomic mode on the ordered sum for many steps.
        uint32_t carry = 0;
        // det with four waves and slices of at most Wd targets: each wave
        // takes a contiguous quarter of the window (entry f0 + (w*kEPT + k)*64
        // + lane in its slot k), so that the order (wave, slot, lane) is the
        // window's order and the entries are placed without a sort
        // (det_window_stable); otherwise entry f0 + k*B + tid
        const bool stable_win = kDet && B == 4 * kWave && L <= Wd;  // uniform
        const int wv = wave_index(), ln = lane_id();
        auto off_of = [&](int k) { return stable_win ? (wv * kEPT + k) * kWave + ln : k * B + tid; };
        for (int f0 = 0, win = 0; f0 < E; f0 += Wd, ++win) {
            int sa[kEPT];
            if (kOwner) {
                uint16_t *own = (win & 1) ? own1 : own0;
#pragma unroll
                for (int k = 0; k < kSPT; ++k)
                    if (head[k] >= f0 && head[k] < f0 + Wd) own[head[k] - f0] = (uint16_t)(tid * kSPT + k);
                __syncthreads();
                // atomic mode: the other map is the next window's; its last
                // readers (the previous window's entries) passed the barrier above
                if (!kDet) clear_owner<kEPT>((win & 1) ? own0 : own1);
                owner_scan<kEPT>(own, scratch, &carry);
#pragma unroll
                for (int k = 0; k < kEPT; ++k) sa[k] = own[off_of(k)];
            } else {
                // the last spike a < m with segp[a] <= f, which holds entry f
                // (segp ascending from segp[0] = 0; empty segments repeat
                // their successor's offset): branchless, kEPT searches in
                // lockstep, log2(m) uniform steps.  A probe past m-1 reads
                // m-1 instead, which is the answer whenever it passes.
#pragma unroll
                for (int k = 0; k < kEPT; ++k) sa[k] = 0;
                for (int step = m > 1 ? 1 << (31 - __builtin_clz((unsigned)(m - 1))) : 0; step > 0; step >>= 1) {
                    int c[kEPT], v[kEPT];
#pragma unroll
                    for (int k = 0; k < kEPT; ++k) {
                        c[k] = min(sa[k] + step, m - 1);
                        v[k] = segp[c[k]];
                    }
#pragma unroll
                    for (int k = 0; k < kEPT; ++k) sa[k] = v[k] <= f0 + off_of(k) ? c[k] : sa[k];
                }
            }
            // every lookup (sa < m) before the first load
            uint64_t idx[kEPT];
#pragma unroll
            for (int k = 0; k < kEPT; ++k) idx[k] = base[sa[k]] + (uint64_t)(f0 + off_of(k));
            uint2 x[kEPT];
#pragma unroll
            for (int k = 0; k < kEPT; ++k) {
                x[k] = make_uint2(0xFFFFFFFFu, 0u);
                if (f0 + off_of(k) < E) x[k] = entry_at(d.syn, (size_t)idx[k]);
            }
            if (!kDet) {
                lds_add_cas<kEPT>(acc, x, lo);
            } else if (stable_win &&
                       det_window_stable<kEPT>(acc, lo, len, x, dw, dstart, dcur, scratch, tb, te, &s_ovf)) {
                if (kOwner) clear_owner<kEPT>(own0);
                __syncthreads();
            } else {
                // ordered: counting sort of the window by target, then each
                // target adds its entries in window-offset (= source) order
#pragma unroll
                for (int k = 0; k < kEPT; ++k)
                    if (x[k].x != 0xFFFFFFFFu) atomicAdd(&dcur[(int)x[k].x - lo], 1);
                __syncthreads();
                // starts of the targets' buckets; the cursor restarts there
                int csum = 0;
                for (int j = tb; j < te; ++j) csum += dcur[j];
                int nwin = 0;
                int st = block_exclusive_scan(csum, scratch, &nwin);
                for (int j = tb; j < te; ++j) {
                    const int c = dcur[j];
                    dstart[j] = st;
                    dcur[j] = st;
                    st += c;
                }
                __syncthreads();
                // keys (window offset = ascending source) into the buckets
#pragma unroll
                for (int k = 0; k < kEPT; ++k)
                    if (x[k].x != 0xFFFFFFFFu) dkey[atomicAdd(&dcur[(int)x[k].x - lo], 1)] = (uint16_t)(k * B + tid);
                __syncthreads();
                // each entry's rank in its bucket puts its weight in key order;
                // the bucket's keys are read eight at a time (16-byte aligned
                // LDS reads; dkey is 16-byte aligned and Wd a multiple of 8)
#pragma unroll
                for (int k = 0; k < kEPT; ++k)
                    if (x[k].x != 0xFFFFFFFFu) {
                        const int j = (int)x[k].x - lo;
                        const int s0 = dstart[j], s1 = dcur[j];
                        const int key = k * B + tid;
                        int r = 0;
                        for (int q = s0 & ~7; q < s1; q += 8) {
                            const uint4 kv = *(const uint4 *)(dkey + q);
                            const uint32_t h[4] = {kv.x, kv.y, kv.z, kv.w};
#pragma unroll
                            for (int e = 0; e < 8; ++e) {
                                const int kq = (int)((h[e >> 1] >> ((e & 1) * 16)) & 0xFFFFu);
                                // position q+e in [s0, s1) and key below ours
                                const bool in = (uint32_t)(q + e - s0) < (uint32_t)(s1 - s0);
                                r += (in && kq < key) ? 1 : 0;
                            }
                        }
                        dw[s0 + r] = u2f(x[k].y);
                    }
                __syncthreads();
                // each target adds its bucket in order; the keys are read,
                // so the map is cleared for the next window
                for (int j = tb; j < te; ++j) {
                    const int s0 = dstart[j], s1 = dcur[j];
                    float a = acc[j];
                    for (int q = s0; q < s1; ++q) a = a + dw[q];
                    acc[j] = a;
                    dcur[j] = 0;
                }
                if (kOwner) clear_owner<kEPT>(own0);
                __syncthreads();
            }
        }
        __syncthreads();
    }
    if (total == 0) __syncthreads();  // otherwise the last pass ended with one
}

// Offset table: soff[i*(P+1) + p] = first index of row i whose target is
// >= min(p*L, N).  One workgroup per row (grid-stride over rows), one pass
// over the row: entry e (targets ascending, x_{-1} = -1) is the answer for
// every boundary p*L in (x_{e-1}, x_e], and index Nsyn for the boundaries
// beyond the last target, up to p = P.  Each p is written once.
__global__ void __launch_bounds__(256) k_slice_offsets(DevNet d, uint16_t *soff)
{
    const int L = d.L, P = d.P, ns = d.nsyn;
    for (int64_t i = blockIdx.x; i < (int64_t)d.n; i += gridDim.x) {
        const uint2 *row = d.syn + (size_t)i * (size_t)ns;
        uint16_t *out = soff + (size_t)i * (size_t)(P + 1);
        for (int e = (int)threadIdx.x; e <= ns; e += (int)blockDim.x) {
            const int p0 = e > 0 ? (int)(row[e - 1].x / (uint32_t)L) + 1 : 0;
            const int p1 = e < ns ? min((int)(row[e].x / (uint32_t)L), P) : P;
            for (int q = p0; q <= p1; ++q) out[q] = (uint16_t)e;
        }
    }
}

// The end of a chunk of `steps` steps, right after its step kernels (its
// start is the chunk's end stamp, control row `steps`).  With `reduce`
// (slice schemes: per-slice counts in bcnt) ctl[rel].count = sum over slices
// of bcnt[rel][p], one workgroup per step.  Then the control block -- the
// kCtlHead words at d.err, then rows 0..steps -- is written to `host`, the
// device address of mapped pinned host memory that the host reads after the
// chunk's completion event (no copy follows), and, if `counts` is set, each
// step's count to counts[rel] (a caller's device buffer).
__device__ __forceinline__ void host_store(int32_t *p, int32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void __launch_bounds__(256) k_finish_chunk(DevNet d, int steps, int reduce, int32_t *host, int32_t *counts)
{
    __shared__ int part[256];
    const int tid = (int)threadIdx.x;
    if (blockIdx.x == 0) {
        if (tid == 0) stamp_step(d.ctl, steps);
        __syncthreads();
        // the head words and the end row, one per thread
        if (tid < kCtlHead + kCtlStride) {
            const int32_t v = tid < kCtlHead ? d.err[tid] : d.ctl[steps * kCtlStride + tid - kCtlHead];
            host_store(&host[tid < kCtlHead ? tid : kCtlHead + steps * kCtlStride + tid - kCtlHead], v);
        }
    }
    for (int rel = blockIdx.x; rel < steps; rel += gridDim.x) {
        int32_t *row = d.ctl + (size_t)rel * kCtlStride;
        if (reduce) {
            int s = 0;
            const int32_t *b = d.bcnt + (size_t)rel * (size_t)d.P;
            for (int p = tid; p < d.P; p += blockDim.x) s += b[p];
            part[tid] = s;
            __syncthreads();
            for (int w = (int)blockDim.x / 2; w > 0; w >>= 1) {
                if (tid < w) part[tid] += part[tid + w];
                __syncthreads();
            }
            if (tid == 0) row[kCtlCount] = part[0];
            __syncthreads();
        }
        if (tid < kCtlStride) {
            const int32_t v = row[tid];
            host_store(&host[kCtlHead + rel * kCtlStride + tid], v);
            if (tid == kCtlCount && counts) counts[rel] = v;
        }
    }
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores have completed
}

// AB scheme: one launch per step (rel-th of the chunk; t = tinfo[0] + rel),
// grid = P slices.  Thread state is loaded from and stored to HBM.
template <int R, bool kDet, bool kLarge>
__global__ void __launch_bounds__(1024) k_slice_step(DevNet d, int rel)
{
    extern __shared__ __align__(16) float dyn[];  // slice_accumulate needs 16-byte alignment
    __shared__ int s_cnt;
    float *acc = dyn;
    const int t = d.tinfo[0] + rel;
    const bool last = t == d.tinfo[1];
    if (threadIdx.x == 0) s_cnt = 0;  // slice_accumulate's barriers precede slice_count
    if (blockIdx.x == 0 && threadIdx.x == 0) stamp_step(d.ctl, rel);
    const int p = slice_of_block((int)blockIdx.x, (int)gridDim.x);
    const int lo = p * d.L;
    const int len = min(d.L, d.n - lo);
    const int B = (int)blockDim.x;
    const int ntile = (len + B - 1) / B;
    // The neurons' state and thalamic draws do not depend on the spikes:
    // load and compute them first so that their latency overlaps the
    // accumulation's dependent loads.  (Computing the draws while the
    // accumulation's first bitmap loads are in flight instead measured the
    // same.)
    float v[R], u[R], a[R], b[R], c[R], dd[R], I0[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int off = k * B + (int)threadIdx.x;
        v[k] = u[k] = a[k] = b[k] = c[k] = dd[k] = I0[k] = 0.0f;
        if (k < ntile && off < len) {
            const int j = lo + off;
            v[k] = d.v[j]; u[k] = d.u[j];
            a[k] = d.a[j]; b[k] = d.b[j]; c[k] = d.c[j]; dd[k] = d.d[j];
            I0[k] = ((j < d.n_exc) ? d.ge : d.gi) * thalamic(d.seed, (uint32_t)t, (uint32_t)j);
        }
    }
    using Plan = SlicePlan<kDet, kLarge>;
    slice_accumulate<kDet, Plan::kEPT, Plan::kSPT, Plan::kOwner>(d, t, p, lo, len, dyn);
    uint64_t masks[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        masks[k] = 0;
        if (k >= ntile) continue;  // uniform
        const int off = k * B + (int)threadIdx.x;
        const int j = lo + off;
        bool spk = false;
        if (off < len) {
            const float I = I0[k] + acc[off];  // (g * z) + input, as neuron_step
            spk = izh_step(v[k], u[k], a[k], b[k], c[k], dd[k], I);
            d.v[j] = v[k];
            d.u[j] = u[k];
            if (spk) d.nospks[j] += 1;
            if (last) d.lastI[j] = I;
        }
        masks[k] = emit_bits(d, t, rel, lo + k * B + ((int)threadIdx.x & ~(kWave - 1)), spk, lo + len);
    }
    slice_count<R>(d, rel, p, ntile, masks, &s_cnt);
}

// Descriptor of a batched instance as the batched kernels use it: the
// fields they read, with pointers in the global address space.  The table
// in device memory holds DevNets; pointers loaded from memory are generic,
// and generic (flat) accesses would also wait on LDS traffic.
#define G4S_GLOBAL __attribute__((address_space(1)))
struct BNet {
    int32_t n, n_exc, nsyn, words;
    uint64_t seed;
    float ge, gi;
    int32_t P, L, t_base, rel_end, last_t;
    G4S_GLOBAL float *v, *u, *a, *b, *c, *d, *lastI;
    G4S_GLOBAL int32_t *nospks, *bcnt;
    G4S_GLOBAL const unsigned long long *syn;  // {target, weight bits} as one word
    G4S_GLOBAL const uint16_t *soff;
    G4S_GLOBAL uint32_t *bm0, *bm1, *raster;
};

__device__ __forceinline__ BNet bnet_of(const DevNet &x)
{
    BNet r;
    r.n = x.n; r.n_exc = x.n_exc; r.nsyn = x.nsyn; r.words = x.words;
    r.seed = x.seed; r.ge = x.ge; r.gi = x.gi;
    r.P = x.P; r.L = x.L; r.t_base = x.t_base; r.rel_end = x.rel_end; r.last_t = x.last_t;
    r.v = (G4S_GLOBAL float *)x.v; r.u = (G4S_GLOBAL float *)x.u;
    r.a = (G4S_GLOBAL float *)x.a; r.b = (G4S_GLOBAL float *)x.b;
    r.c = (G4S_GLOBAL float *)x.c; r.d = (G4S_GLOBAL float *)x.d;
    r.lastI = (G4S_GLOBAL float *)x.lastI;
    r.nospks = (G4S_GLOBAL int32_t *)x.nospks;
    r.bcnt = (G4S_GLOBAL int32_t *)x.bcnt;
    r.syn = (G4S_GLOBAL const unsigned long long *)x.syn;
    r.soff = (G4S_GLOBAL const uint16_t *)x.soff;
    r.bm0 = (G4S_GLOBAL uint32_t *)x.bm0; r.bm1 = (G4S_GLOBAL uint32_t *)x.bm1;
    r.raster = (G4S_GLOBAL uint32_t *)x.raster;
    return r;
}

// Batched form.  A work unit is slice p of one instance at chunk step rel:
// it accumulates that slice's input from the instance's previous step,
// updates the slice's neurons and records its spike count.
template <bool kDet>
__device__ __forceinline__ void batch_unit(const BNet &d, int rel, int p, float *dyn, int *s_cnt)
{
    const int t = d.t_base + rel;
    const int lo = p * d.L;
    const int len = min(d.L, d.n - lo);
    float *acc = dyn;
    const int B = (int)blockDim.x;
    const int tid = (int)threadIdx.x;
    if (tid == 0) *s_cnt = 0;  // slice_accumulate's barriers precede the count atomics
    slice_accumulate<kDet, entries_batch<kDet>(), spikes_batch<kDet>(), kOwnerBatch>(d, t, p, lo, len, dyn);
    const bool last = t == d.last_t;
    int c = 0;
    for (int base = 0; base < len; base += B) {  // uniform
        const int off = base + tid;
        bool spk = false;
        if (off < len) spk = neuron_step(d, t, lo + off, acc[off], last);
        const uint64_t mask = emit_bits(d, t, rel, lo + base + (tid & ~(kWave - 1)), spk, lo + len);
        c += __popcll(mask);
    }
    if (lane_id() == 0 && c) atomicAdd(s_cnt, c);
    __syncthreads();
    if (tid == 0) d.bcnt[(size_t)rel * (size_t)d.P + (size_t)p] = *s_cnt;
}

// One launch per step: workgroup b takes work unit work[b] = {slot, p} at
// step rel; the host orders the units so that each XCD holds whole
// instances.  Instances whose chunk ends before rel do nothing.
// 256-thread workgroups: 7 waves per SIMD (66 VGPRs, no scratch, in both
// modes), and 7 workgroups' LDS fits a CU at the default 768-neuron slices.
template <bool kDet, int kB>
__global__ void __launch_bounds__(kB, kB == 256 ? 7 : 1) k_batch_step(const DevNet *__restrict__ descs,
                                                                     const uint2 *__restrict__ work, int rel)
{
    extern __shared__ __align__(16) float dyn[];  // slice_accumulate needs 16-byte alignment
    __shared__ int s_cnt;
    const uint2 wu = work[blockIdx.x];
    if (wu.x == kNoWork) return;
    // a private copy: the kernel's stores could otherwise alias the table,
    // forcing a reload of every field after each store
    const BNet d = bnet_of(descs[wu.x]);
    if (rel >= d.rel_end) return;
    batch_unit<kDet>(d, rel, (int)wu.y, dyn, &s_cnt);
}

// One persistent launch per chunk, scheduled by data flow.  units[] holds
// every work unit of the chunk in the order it becomes ready: {slot,
// rel << 16 | p}; the host writes the units of each instance's first step,
// the rest of the array is kNoWork.  A workgroup claims the next position
// (ctr[0]), waits until that unit is published, and runs it.  The workgroup
// that completes the last slice of an instance's step (done[slot] counts
// finished slices) publishes the instance's next step (positions reserved
// at ctr[1]).  Instances thus advance independently; a position is only
// waited on after every earlier position was claimed by a running
// workgroup, and those publish all that is left, so every wait ends.  Waits
// are bounded (ctr[2] = error flag).  Grid: any size.
template <bool kDet>
__global__ void __launch_bounds__(256, 6) k_batch_chunk(const DevNet *__restrict__ descs, unsigned long long *units,
                                                     int32_t *ctr, int32_t *done, int total)
{
    extern __shared__ __align__(16) float dyn[];  // slice_accumulate needs 16-byte alignment
    __shared__ int s_cnt;
    __shared__ int s_pos, s_pub, s_base;
    __shared__ unsigned long long s_unit;
    const int tid = (int)threadIdx.x;
    for (;;) {
        if (tid == 0) {
            int pos = atomicAdd(&ctr[0], 1);
            unsigned long long u = 0xFFFFFFFFFFFFFFFFull;
            if (pos < total) {
                uint32_t spins = 0;
                for (;;) {
                    u = __hip_atomic_load(&units[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if ((uint32_t)u != kNoWork) break;
                    if (__hip_atomic_load(&ctr[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                        ++spins > kSpinLimit) {
                        __hip_atomic_store(&ctr[2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        pos = total;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            s_pos = pos;
            s_unit = u;
        }
        __syncthreads();
        if (s_pos >= total) return;  // uniform
        const unsigned long long u = s_unit;
        const uint32_t slot = (uint32_t)u;
        const int rel = (int)(u >> 48);
        const int p = (int)((u >> 32) & 0xFFFFu);
        const BNet d = bnet_of(descs[slot]);
        batch_unit<kDet>(d, rel, p, dyn, &s_cnt);
        // release this slice; the last one of the step publishes the next
        __syncthreads();
        if (tid == 0) {
            const int old = __hip_atomic_fetch_add(&done[slot], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            const bool pub = old == (rel + 1) * d.P - 1 && rel + 1 < d.rel_end;
            s_pub = pub ? 1 : 0;
            if (pub) s_base = atomicAdd(&ctr[1], d.P);
        }
        __syncthreads();
        if (s_pub) {
            const unsigned long long hi = (unsigned long long)(rel + 1) << 48;
            for (int q = tid; q < d.P; q += (int)blockDim.x)
                __hip_atomic_store(&units[s_base + q], hi | ((unsigned long long)q << 32) | slot, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    }
}

// Per-step spike counts of the chunk's instances: workgroup k sums the
// slices' counts of instance slots[k] for every step of the chunk.
__global__ void __launch_bounds__(256) k_batch_counts(const DevNet *descs, const int32_t *slots, int cs,
                                                      int32_t *counts)
{
    const DevNet &d = descs[slots[blockIdx.x]];
    for (int rel = (int)threadIdx.x; rel < cs; rel += (int)blockDim.x) {
        int s = 0;
        if (rel < d.rel_end) {
            const int32_t *row = d.bcnt + (size_t)rel * (size_t)d.P;
            for (int p = 0; p < d.P; ++p) s += row[p];
        }
        counts[(size_t)blockIdx.x * (size_t)cs + (size_t)rel] = s;
    }
}

// SKL scheme: the whole chunk in one co-resident launch of P workgroups.
// Each thread keeps the state of its R neurons (j = lo + k*B + tid) in
// registers for the whole chunk; per step: slice accumulation from the
// previous step's region queue, update, spike output, one grid barrier.
template <int R, bool kDet, bool kLarge>
__global__ void __launch_bounds__(1024) k_skl_slice(DevNet d, int t0, int steps, int last_t)
{
    extern __shared__ __align__(16) float dyn[];  // slice_accumulate needs 16-byte alignment
    __shared__ int s_wcnt[1024 / kWave];  // per-wave spike totals of the step
    // the first step's start stamp is the launch's start (its state loads
    // belong to the chunk's time); later steps stamp their own start
    if (blockIdx.x == 0 && threadIdx.x == 0) stamp_step(d.ctl, 0);
    float *acc = dyn;
    const int p = slice_of_block((int)blockIdx.x, (int)gridDim.x);
    const int lo = p * d.L;
    const int len = min(d.L, d.n - lo);
    const int B = (int)blockDim.x;
    const int tid = (int)threadIdx.x;
    const int ntile = (len + B - 1) / B;
    float v[R], u[R], a[R], b[R], c[R], dd[R], g[R];
    int nos[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int off = k * B + tid;
        nos[k] = 0;
        v[k] = u[k] = a[k] = b[k] = c[k] = dd[k] = g[k] = 0.0f;
        if (off < len) {
            const int j = lo + off;
            v[k] = d.v[j]; u[k] = d.u[j];
            a[k] = d.a[j]; b[k] = d.b[j]; c[k] = d.c[j]; dd[k] = d.d[j];
            g[k] = (j < d.n_exc) ? d.ge : d.gi;
        }
    }
    uint32_t epoch = 0;
    int backoff = 0;  // slice_accumulate's fast-lookup backoff, uniform over the workgroup
    // atomic mode: each neuron's input is cleared as the update reads it, so
    // the accumulation needs no clearing barrier of its own
    if constexpr (!kDet) {
        for (int i = tid; i < len; i += B) acc[i] = 0.0f;
        __syncthreads();
    }
    for (int rel = 0; rel < steps; ++rel) {
        const int t = t0 + rel;
        if (blockIdx.x == 0 && tid == 0 && rel > 0) stamp_step(d.ctl, rel);
        // det mode computes the step's thalamic draws while the accumulation's
        // first bitmap loads are in flight (18.25 -> 18.0 us per step at
        // N=100k); in atomic mode the draws follow the accumulation, which
        // measured faster than overlapping them with the bitmap loads (13.35
        // -> 13.8) or with the segment-bound loads (9.23 -> 9.96).  Atomic
        // mode clears each neuron's input as the update reads it.
        float I0[R];
        using Plan = SlicePlan<kDet, kLarge>;
        if constexpr (kDet) {
            auto draws = [&]() {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int off = k * B + tid;
                    I0[k] = 0.0f;
                    if (k < ntile && off < len) I0[k] = g[k] * thalamic(d.seed, (uint32_t)t, (uint32_t)(lo + off));
                }
            };
            slice_accumulate<kDet, Plan::kEPT, Plan::kSPT, Plan::kOwner, true>(d, t, p, lo, len, dyn, draws, &backoff);
        } else {
            slice_accumulate<kDet, Plan::kEPT, Plan::kSPT, Plan::kOwner, true, true>(d, t, p, lo, len, dyn, NoPre(),
                                                                                    &backoff);
        }
        uint64_t masks[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            masks[k] = 0;
            if (k >= ntile) continue;  // uniform
            const int off = k * B + tid;
            const int j = lo + off;
            bool spk = false;
            if (off < len) {
                if constexpr (!kDet) I0[k] = g[k] * thalamic(d.seed, (uint32_t)t, (uint32_t)j);
                float I = I0[k];  // g * z, as neuron_step
                I = I + acc[off];
                if constexpr (!kDet) acc[off] = 0.0f;
                spk = izh_step(v[k], u[k], a[k], b[k], c[k], dd[k], I);
                nos[k] += spk ? 1 : 0;
                if (t == last_t) d.lastI[j] = I;
            }
            masks[k] = emit_bits<true>(d, t, rel, lo + k * B + (tid & ~(kWave - 1)), spk, lo + len);
        }
        int c = 0;
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (k < ntile) c += __popcll(masks[k]);
        if (lane_id() == 0) s_wcnt[tid / kWave] = c;
        if (!grid_sync(d, epoch, s_wcnt, &d.bcnt[(size_t)rel * (size_t)d.P + (size_t)p])) break;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int off = k * B + tid;
        if (off < len) {
            const int j = lo + off;
            d.v[j] = v[k];
            d.u[j] = u[k];
            d.nospks[j] += nos[k];
        }
    }
}

// ---------------------------------------------------------------------------
// Connectivity on the device, bit-equal to the host builder (build_table):
// row i holds the values of Floyd's algorithm over the m = N-1 non-self
// candidates -- for d = 0..Nsyn-1 with jj_d = m-Nsyn+d: t_d uniform on
// [0, jj_d] from Philox (i, d, CONN); take t_d unless already taken, else
// jj_d -- mapped v -> v + (v >= i), sorted, with weights g*U(Philox(i, j,
// WEIGHT)).  Restated for a workgroup per row:
//  * a value below J0 = jj_0 can only be taken as a draw, so draw d takes
//    t_d < J0 exactly when d is the first draw of that value: the least d
//    per value, from an LDS hash table (open addressing, at most half full);
//  * a value jj_b >= J0 is taken either as draw b's fallback (bit b) or as
//    a draw t_k = jj_b at k >= b; only draws t_k >= J0 read that state, so
//    one thread resolves those draws in ascending d over an Nsyn-bit set --
//    in sparse rows a handful per row (about Nsyn^2 / 2N);
//  * the row's values, distinct, are put in order by buckets of equal value
//    ranges (a count per bucket, a scan, a place per value) and a value's
//    rank is its bucket's start plus the values below it in its bucket.
// kBuildThreads threads per row; Nsyn <= kDevRowMax.
// ---------------------------------------------------------------------------
constexpr int kDevRowMax = 1024;
constexpr int kBuildThreads = 256;
constexpr int kBuildPer = kDevRowMax / kBuildThreads;  // draws per thread
constexpr uint32_t kHashEmpty = 0xFFFFFFFFu;           // no value (values are below 2^31)

__global__ void __launch_bounds__(kBuildThreads) k_build_rows(uint64_t seed, int64_t n, int64_t n_exc, int32_t nsyn,
                                                              float g_exc, float g_inh, uint2 *tab, int64_t row0,
                                                              int64_t rows)
{
    __shared__ uint32_t tv[kDevRowMax];
    __shared__ uint32_t val[kDevRowMax];
    // the hash table (keys, least draw) until the draws are settled, then the
    // buckets: counts and starts (kDevRowMax + 1), placed values, row order
    __shared__ uint32_t pool[4 * kDevRowMax];
    __shared__ uint32_t taken[kDevRowMax / 32], hib[kDevRowMax / 32];
    __shared__ int s_scan[kBuildThreads / kWave];
    uint32_t *hkey = pool, *hmin = pool + 2 * kDevRowMax;
    uint32_t *bstart = pool, *bkt = pool + kDevRowMax + 32, *ord = pool + 2 * kDevRowMax + 64;
    const int tid = (int)threadIdx.x;
    const uint64_t m = (uint64_t)(n - 1);
    const uint64_t J0 = m - (uint64_t)nsyn;
    int S = 64;
    while (S < nsyn) S <<= 1;
    const int H = 2 * S;  // hash slots
    const int hshift = 32 - (31 - __builtin_clz((unsigned)H));
    // buckets: S of them, value v in bucket (v * mul) >> 32, non-decreasing in
    // v and below S for v < n (mul = 2^32 when N <= S: one value per bucket)
    const uint64_t mul = (uint64_t)n <= (uint64_t)S ? (1ull << 32) : (((uint64_t)S << 32) / (uint64_t)n);
    const int words = (nsyn + 31) / 32;
    for (int64_t i = row0 + blockIdx.x; i < row0 + rows; i += gridDim.x) {
        // 1. draws; the table and the bit sets cleared
        uint32_t t[kBuildPer];
#pragma unroll
        for (int k = 0; k < kBuildPer; ++k) {
            const int d = tid + kBuildThreads * k;
            t[k] = kHashEmpty;
            if (d < nsyn) {
                const uint64_t jj = J0 + (uint64_t)d;
                const Word4 o = draw(seed, (uint32_t)i, (uint32_t)d, kTagConn);
                const uint64_t r = ((uint64_t)o.x1 << 32) | o.x0;
                t[k] = (uint32_t)__umul64hi(r, jj + 1u);
                tv[d] = t[k];
            }
        }
        for (int h = tid; h < H; h += kBuildThreads) {
            hkey[h] = kHashEmpty;
            hmin[h] = kHashEmpty;
        }
        for (int w = tid; w < words; w += kBuildThreads) {
            taken[w] = 0u;
            hib[w] = 0u;
        }
        __syncthreads();
        // 2. the least draw of each value below J0
        int slot[kBuildPer];
#pragma unroll
        for (int k = 0; k < kBuildPer; ++k) {
            const int d = tid + kBuildThreads * k;
            slot[k] = -1;
            if (d < nsyn && (uint64_t)t[k] < J0) {
                uint32_t h = (t[k] * 0x9E3779B1u) >> hshift;
                for (;;) {  // ends: the table is at most half full
                    const uint32_t old = atomicCAS(&hkey[h], kHashEmpty, t[k]);
                    if (old == kHashEmpty || old == t[k]) break;
                    h = (h + 1u) & (uint32_t)(H - 1);
                }
                atomicMin(&hmin[h], (uint32_t)d);
                slot[k] = (int)h;
            }
        }
        __syncthreads();
        // 3. draws below J0 settle at once (fallbacks of repeated values take
        //    their jj); the others are marked for step 4
#pragma unroll
        for (int k = 0; k < kBuildPer; ++k) {
            const int d = tid + kBuildThreads * k;
            if (d >= nsyn) continue;
            if (slot[k] >= 0) {
                if (hmin[slot[k]] == (uint32_t)d) {
                    val[d] = t[k];
                } else {
                    val[d] = (uint32_t)(J0 + (uint64_t)d);
                    atomicOr(&taken[d >> 5], 1u << (d & 31));
                }
            } else {
                atomicOr(&hib[d >> 5], 1u << (d & 31));
            }
        }
        __syncthreads();
        // 4. draws at or above J0, in ascending d, by one thread; the others
        //    clear the bucket counts (over the table, no longer read)
        if (tid == 0) {
            for (int w = 0; w < words; ++w)
                for (uint32_t bits = hib[w]; bits; bits &= bits - 1u) {
                    const int d = 32 * w + __builtin_ctz(bits);
                    const uint32_t b = (uint32_t)((uint64_t)tv[d] - J0);
                    const bool is_taken = (int)b == d || ((taken[b >> 5] >> (b & 31)) & 1u);
                    if (is_taken) {
                        val[d] = (uint32_t)(J0 + (uint64_t)d);
                        taken[d >> 5] |= 1u << (d & 31);
                    } else {
                        val[d] = tv[d];
                        taken[b >> 5] |= 1u << (b & 31);
                    }
                }
        }
        for (int q = tid; q <= S; q += kBuildThreads) bstart[q] = 0u;
        __syncthreads();
        // 5. skip the presynaptic neuron itself; count the buckets
        uint32_t v[kBuildPer], bk[kBuildPer], pos[kBuildPer];
#pragma unroll
        for (int k = 0; k < kBuildPer; ++k) {
            const int d = tid + kBuildThreads * k;
            v[k] = bk[k] = pos[k] = 0u;
            if (d < nsyn) {
                v[k] = val[d] + ((int64_t)val[d] >= i ? 1u : 0u);
                bk[k] = (uint32_t)(((uint64_t)v[k] * mul) >> 32);
                pos[k] = atomicAdd(&bstart[bk[k]], 1u);
            }
        }
        __syncthreads();
        // 6. bucket starts: thread q scans buckets [q*per, (q+1)*per)
        {
            const int per = (S + kBuildThreads - 1) / kBuildThreads;
            const int q0 = min(tid * per, S), q1 = min(q0 + per, S);
            int c = 0;
            for (int q = q0; q < q1; ++q) c += (int)bstart[q];
            int tot = 0;
            const int lane = lane_id(), wave = tid / kWave;
            const int incl = wave_inclusive_sum(c);
            if (lane == kWave - 1) s_scan[wave] = incl;
            __syncthreads();  // every count was read above
            int run = waves_prefix_sum(s_scan, &tot) + incl - c;
            for (int q = q0; q < q1; ++q) {
                const int cq = (int)bstart[q];
                bstart[q] = (uint32_t)run;
                run += cq;
            }
            if (tid == 0) bstart[S] = (uint32_t)nsyn;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kBuildPer; ++k)
            if (tid + kBuildThreads * k < nsyn) bkt[bstart[bk[k]] + pos[k]] = v[k];
        __syncthreads();
        // 7. rank = bucket start + the bucket's values below this one
#pragma unroll
        for (int k = 0; k < kBuildPer; ++k)
            if (tid + kBuildThreads * k < nsyn) {
                const uint32_t b0 = bstart[bk[k]], b1 = bstart[bk[k] + 1];
                uint32_t r = b0;
                for (uint32_t q = b0; q < b1; ++q) r += bkt[q] < v[k] ? 1u : 0u;
                ord[r] = v[k];
            }
        __syncthreads();
        // 8. the row in order, with its weights
        const float g = i < n_exc ? g_exc : g_inh;
        uint2 *out = tab + (size_t)(i - row0) * (size_t)nsyn;
        for (int d = tid; d < nsyn; d += kBuildThreads) {
            const uint32_t j = ord[d];
            const Word4 o = draw(seed, (uint32_t)i, j, kTagWeight);
            out[d] = make_uint2(j, f2u(g * u01(o.x0)));
        }
        __syncthreads();  // the next row clears what this one still reads
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

// Ordinary launch of `kernel`; with L's events set, the dispatch packet
// stamps them at the kernel's start and end.
template <typename... Args, typename F = void (*)(Args...)>
hipError_t plain_launch(F kernel, const Launch &L, uint32_t shmem, Args... args)
{
    if (L.ev_start)
        hipExtLaunchKernelGGL(kernel, dim3(L.grid), dim3(L.block), shmem, L.stream, L.ev_start, L.ev_stop, 0, args...);
    else
        hipLaunchKernelGGL(kernel, dim3(L.grid), dim3(L.block), shmem, L.stream, args...);
    return hipGetLastError();
}

// The large plan exists for R >= 2 (slices of kLargeSlice neurons or more
// at up to 1024 threads).
template <int R, bool kDet>
const void *skl_slice_fn_r(bool large)
{
    if constexpr (R >= 2)
        if (large) return (const void *)k_skl_slice<R, kDet, true>;
    return (const void *)k_skl_slice<R, kDet, false>;
}

// The SKL grid waits on itself (one grid barrier per step), so every
// workgroup must be resident at once: the runtime sizes it from
// skl_slice_blocks_per_cu.  It is an ordinary launch: the barrier is the
// library's own, and its waits are bounded.
template <int R, bool kDet>
hipError_t launch_skl_slice_r(const DevNet &d, int t0, int steps, int last_t, const Launch &L)
{
    const uint32_t lds = (uint32_t)single_slice_lds(d.L, L.block, kDet);
    if constexpr (R >= 2)
        if (large_slice_plan(d.L, L.block, kDet))
            return plain_launch(k_skl_slice<R, kDet, true>, L, lds, d, t0, steps, last_t);
    return plain_launch(k_skl_slice<R, kDet, false>, L, lds, d, t0, steps, last_t);
}

template <int R, bool kDet>
hipError_t launch_slice_step_r(const DevNet &d, int rel, const Launch &L)
{
    const uint32_t lds = (uint32_t)single_slice_lds(d.L, L.block, kDet);
    if constexpr (R >= 2)
        if (large_slice_plan(d.L, L.block, kDet)) return plain_launch(k_slice_step<R, kDet, true>, L, lds, d, rel);
    return plain_launch(k_slice_step<R, kDet, false>, L, lds, d, rel);
}

template <bool kDet>
const void *skl_slice_fn(int per_thread, bool large)
{
    if (per_thread <= 1) return skl_slice_fn_r<1, kDet>(large);
    if (per_thread <= 2) return skl_slice_fn_r<2, kDet>(large);
    if (per_thread <= 4) return skl_slice_fn_r<4, kDet>(large);
    if (per_thread <= 8) return skl_slice_fn_r<8, kDet>(large);
    return skl_slice_fn_r<16, kDet>(large);
}

}  // namespace

hipError_t launch_init(const DevNet &d, hipStream_t s)
{
    const int64_t m = d.n > d.words ? d.n : d.words;
    hipLaunchKernelGGL(k_init, dim3(grid_for(m, 256)), dim3(256), 0, s, d);
    return hipGetLastError();
}


hipError_t launch_fill2_u32(uint32_t *p, size_t n, uint32_t *q, size_t m, uint32_t val, hipStream_t s)
{
    if (n + m == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill2_u32, dim3(grid_for((int64_t)(n + m), 256)), dim3(256), 0, s, p, n, q, m, val);
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
    return plain_launch(k_update, L, 0, d, t, rel, last);
}

hipError_t launch_push_rows(const DevNet &d, int t, int rel, const Launch &L)
{
    return plain_launch(k_push_rows, L, 0, d, t, rel);
}

hipError_t launch_push_bitmap(const DevNet &d, int t, hipStream_t s)
{
    const int64_t groups = ((int64_t)d.words + 1) / 2;
    hipLaunchKernelGGL(k_push_bitmap, dim3(grid_for(groups * 64, 256)), dim3(256), 0, s, d, t);
    return hipGetLastError();
}

hipError_t launch_finish_chunk(const DevNet &d, int steps, bool reduce, int32_t *host, int32_t *counts, hipEvent_t done,
                               hipStream_t s)
{
    // one workgroup per step; `done` (if set) is signalled by this dispatch's end
    const dim3 grid(steps < 4096 ? (steps > 0 ? steps : 1) : 4096);
    if (done)
        hipExtLaunchKernelGGL(k_finish_chunk, grid, dim3(256), 0, s, nullptr, done, 0, d, steps, reduce ? 1 : 0, host,
                              counts);
    else
        hipLaunchKernelGGL(k_finish_chunk, grid, dim3(256), 0, s, d, steps, reduce ? 1 : 0, host, counts);
    return hipGetLastError();
}

hipError_t launch_push_flat(const DevNet &d, int t, int rel, const Launch &L)
{
    return plain_launch(k_push_flat, L, 0, d, t, rel);
}

hipError_t launch_ap(const DevNet &d, int t, int rel, int last, int prev_count, const Launch &L)
{
    // consumers wait on producers of the same grid: the runtime sizes it
    // from persistent_blocks_per_cu, so every workgroup is resident at once
    return plain_launch(k_ap, L, 0, d, t, rel, last, prev_count);
}

hipError_t launch_slice_offsets(const DevNet &d, uint16_t *soff, hipStream_t s)
{
    const int grid = (int)std::min<int64_t>(d.n, 65536);
    hipLaunchKernelGGL(k_slice_offsets, dim3(grid), dim3(256), 0, s, d, soff);
    return hipGetLastError();
}

template <bool kDet>
hipError_t launch_slice_step_t(const DevNet &d, int rel, const Launch &L)
{
    const int per_thread = (d.L + L.block - 1) / L.block;
    if (per_thread <= 1) return launch_slice_step_r<1, kDet>(d, rel, L);
    if (per_thread <= 2) return launch_slice_step_r<2, kDet>(d, rel, L);
    if (per_thread <= 4) return launch_slice_step_r<4, kDet>(d, rel, L);
    if (per_thread <= 8) return launch_slice_step_r<8, kDet>(d, rel, L);
    if (per_thread <= 16) return launch_slice_step_r<16, kDet>(d, rel, L);
    return hipErrorInvalidValue;
}

template <bool kDet>
hipError_t launch_skl_slice_t(const DevNet &d, int t0, int steps, int last_t, const Launch &L)
{
    const int per_thread = (d.L + L.block - 1) / L.block;
    if (per_thread <= 1) return launch_skl_slice_r<1, kDet>(d, t0, steps, last_t, L);
    if (per_thread <= 2) return launch_skl_slice_r<2, kDet>(d, t0, steps, last_t, L);
    if (per_thread <= 4) return launch_skl_slice_r<4, kDet>(d, t0, steps, last_t, L);
    if (per_thread <= 8) return launch_skl_slice_r<8, kDet>(d, t0, steps, last_t, L);
    if (per_thread <= 16) return launch_skl_slice_r<16, kDet>(d, t0, steps, last_t, L);
    return hipErrorInvalidValue;
}

hipError_t launch_slice_step(const DevNet &d, int rel, bool det, const Launch &L)
{
    return det ? launch_slice_step_t<true>(d, rel, L) : launch_slice_step_t<false>(d, rel, L);
}

hipError_t launch_skl_slice(const DevNet &d, int t0, int steps, int last_t, bool det, const Launch &L)
{
    return det ? launch_skl_slice_t<true>(d, t0, steps, last_t, L) : launch_skl_slice_t<false>(d, t0, steps, last_t, L);
}

void slice_kernel_targs(int L, int block, bool det, char *out, size_t cap)
{
    const int r = (L + block - 1) / block;
    const int R = r <= 1 ? 1 : r <= 2 ? 2 : r <= 4 ? 4 : r <= 8 ? 8 : 16;
    const bool large = R >= 2 && large_slice_plan(L, block, det);
    snprintf(out, cap, "%d,%s,%s", R, det ? "true" : "false", large ? "true" : "false");
}

int skl_slice_blocks_per_cu(int L, int block, bool det)
{
    const int r = (L + block - 1) / block;
    const bool large = r >= 2 && large_slice_plan(L, block, det);
    const void *fn = det ? skl_slice_fn<true>(r, large) : skl_slice_fn<false>(r, large);
    char targs[32];
    slice_kernel_targs(L, block, det, targs, sizeof(targs));
    return resident_per_cu("k_skl_slice", targs, fn, block, single_slice_lds(L, block, det));
}

template <int R, bool kDet>
const void *slice_step_fn_r(bool large)
{
    if constexpr (R >= 2)
        if (large) return (const void *)k_slice_step<R, kDet, true>;
    return (const void *)k_slice_step<R, kDet, false>;
}

int slice_step_blocks_per_cu(int L, int block, bool det)
{
    const int r = (L + block - 1) / block;
    const bool large = r >= 2 && large_slice_plan(L, block, det);
    const int R = r <= 1 ? 1 : r <= 2 ? 2 : r <= 4 ? 4 : r <= 8 ? 8 : 16;
    const void *fn = nullptr;
    switch (R * 2 + (det ? 1 : 0)) {
    case 2: fn = slice_step_fn_r<1, false>(large); break;
    case 3: fn = slice_step_fn_r<1, true>(large); break;
    case 4: fn = slice_step_fn_r<2, false>(large); break;
    case 5: fn = slice_step_fn_r<2, true>(large); break;
    case 8: fn = slice_step_fn_r<4, false>(large); break;
    case 9: fn = slice_step_fn_r<4, true>(large); break;
    case 16: fn = slice_step_fn_r<8, false>(large); break;
    case 17: fn = slice_step_fn_r<8, true>(large); break;
    case 32: fn = slice_step_fn_r<16, false>(large); break;
    default: fn = slice_step_fn_r<16, true>(large); break;
    }
    char targs[32];
    slice_kernel_targs(L, block, det, targs, sizeof(targs));
    return resident_per_cu("k_slice_step", targs, fn, block, single_slice_lds(L, block, det));
}

size_t slice_lds(int L, int block, bool det) { return single_slice_lds(L, block, det); }
size_t batch_lds(int L, bool det, int block)
{
    return slice_lds_bytes(L, block, det, det ? entries_batch<true>() : entries_batch<false>(),
                           det ? spikes_batch<true>() : spikes_batch<false>(), kOwnerBatch);
}

hipError_t launch_batch_chunk(const DevNet *descs, unsigned long long *units, int32_t *ctr, int32_t *done, int total,
                              bool det, size_t lds, const Launch &L)
{
    if (det) return plain_launch(k_batch_chunk<true>, L, (uint32_t)lds, descs, units, ctr, done, total);
    return plain_launch(k_batch_chunk<false>, L, (uint32_t)lds, descs, units, ctr, done, total);
}

int batch_chunk_blocks_per_cu(bool det, size_t lds)
{
    int per_cu = 0;
    const void *fn = det ? (const void *)k_batch_chunk<true> : (const void *)k_batch_chunk<false>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, lds) != hipSuccess) return 0;
    return per_cu;
}

hipError_t launch_batch_step(const DevNet *descs, const uint2 *work, int grid, int rel, bool det, size_t lds,
                             const Launch &L)
{
    Launch Lb = L;
    Lb.grid = grid;
    switch (L.block * 2 + (det ? 1 : 0)) {
    case 512: return plain_launch(k_batch_step<false, 256>, Lb, (uint32_t)lds, descs, work, rel);
    case 513: return plain_launch(k_batch_step<true, 256>, Lb, (uint32_t)lds, descs, work, rel);
    case 1024: return plain_launch(k_batch_step<false, 512>, Lb, (uint32_t)lds, descs, work, rel);
    case 1025: return plain_launch(k_batch_step<true, 512>, Lb, (uint32_t)lds, descs, work, rel);
    case 2048: return plain_launch(k_batch_step<false, 1024>, Lb, (uint32_t)lds, descs, work, rel);
    case 2049: return plain_launch(k_batch_step<true, 1024>, Lb, (uint32_t)lds, descs, work, rel);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batch_counts(const DevNet *descs, const int32_t *slots, int nslots, int cs, int32_t *counts,
                               hipStream_t s)
{
    if (nslots <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_counts, dim3(nslots), dim3(256), 0, s, descs, slots, cs, counts);
    return hipGetLastError();
}

hipError_t launch_build_rows(uint64_t seed, int64_t n, int64_t n_exc, int32_t nsyn, float g_exc, float g_inh, uint2 *tab,
                             int64_t row0, int64_t rows, hipStream_t s)
{
    if (nsyn < 1 || nsyn > kDevRowMax || rows <= 0) return hipErrorInvalidValue;
    const int grid = (int)std::min<int64_t>(rows, 65536);
    hipLaunchKernelGGL(k_build_rows, dim3(grid), dim3(kBuildThreads), 0, s, seed, n, n_exc, nsyn, g_exc, g_inh, tab, row0,
                       rows);
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

const KernelRes *kernel_res(const char *name, const char *targs)
{
    for (int k = 0; k < kKernelResCount; ++k)
        if (strcmp(kKernelRes[k].name, name) == 0 && strcmp(kKernelRes[k].targs, targs) == 0) return &kKernelRes[k];
    return nullptr;
}

int resident_per_cu(const char *name, const char *targs, const void *fn, int block, size_t dyn_lds)
{
    constexpr int64_t kLdsPerCu = 160 * 1024, kLdsGranule = 1024, kSimdsPerCu = 4;
    const KernelRes *r = kernel_res(name, targs);
    if (!r || r->occupancy <= 0 || block <= 0) return 0;
    const int waves = (block + kWave - 1) / kWave;
    const int per_simd = (int)((waves + kSimdsPerCu - 1) / kSimdsPerCu);  // a workgroup's waves on each SIMD
    int64_t n = r->occupancy / per_simd;
    const int64_t lds = (int64_t)dyn_lds + r->lds;
    if (lds > 0) n = std::min<int64_t>(n, kLdsPerCu / ((lds + kLdsGranule - 1) / kLdsGranule * kLdsGranule));
    int api = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&api, fn, block, dyn_lds) != hipSuccess) return 0;
    return (int)std::max<int64_t>(0, std::min<int64_t>(n, api));
}

int persistent_blocks_per_cu(PersistentKernel k, int block)
{
    switch (k) {
    case kPersistAP: return resident_per_cu("k_ap", "", (const void *)k_ap, block, 0);
    }
    return 0;
}

int device_cus()
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return cus;
}

}  // namespace g4s
