// g4s_internal.h -- device-side view of one network and the launch entry
// points shared between g4s_kernels.hip and g4s_runtime.hip.
#pragma once
#include "../../include/gpu4snn.h"
#include "g4s_math.h"

namespace g4s {

// Host helpers shared by the runtime files.  fail() sets the calling
// thread's message (g4s_last_error) and returns `code`.
int fail(int code, const char *fmt, ...);
int validate_spec(const g4s_spec *s);
int64_t n_exc_of(const g4s_spec *s);
// Outgoing table of `spec` into device memory `syn`: built on the device
// (Nsyn <= kDeviceBuildMaxNsyn; queued on `stream`, not waited for) or on
// the host and uploaded; same table.
int build_synapses(const g4s_spec *spec, uint2 *syn, hipStream_t stream);
// g4s_batch_add with the instance's slice length given (0: the batch's rule).
int batch_add_sliced(g4s_batch *b, const g4s_spec *spec, int64_t steps, int keep_raster, int32_t slice_len,
                     int32_t *slot);

#define G4S_HIP_OK(expr)                                                                           \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return ::g4s::fail(G4S_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

constexpr uint32_t kSentinel = 0xFFFFFFFFu;   // empty spike-queue slot
constexpr int kCtlStride = 4;                 // per-step control words
// Control row of a step: [0] spikes of the step (k_ap keeps finished
// workgroups in word 1), [2..3] the step's start on the device wall clock.
enum { kCtlCount = 0, kCtlStamp = 2 };
// A run's control block: kCtlHead words (word 0 the error flag of bounded
// waits) followed by the control rows; the chunk-end kernel writes both to
// the host in one pass.
constexpr int kCtlHead = 4;
constexpr uint32_t kSpinLimit = 1u << 24;     // bounded device waits (~10 s)
// Grid barrier: one arrival counter and one release flag per group of
// workgroups (blockIdx % 16, two groups per XCD), each in its own 256-byte
// line; the counters first, then the flags.  Measured on MI355X, 241
// workgroups of 1024 threads, barrier alone: 2.02-2.05 us with 16 groups,
// 2.14 with 8, 2.09 with 32; counters polled directly (no flags) 2.67-2.99.
constexpr int kBarGroups = 16;
constexpr int kBarStride = 64;
constexpr int kBarWords = 2 * kBarGroups * kBarStride;

// Everything a kernel needs about one network, passed by value.
//  - neuron state/parameters are structure-of-arrays fp32 (the reference's
//    36-byte AoS Neuron, cu:50-62, split so each field is one coalesced
//    stream);
//  - syn is the outgoing table: row i = Nsyn {target, weight-bits} pairs,
//    ascending target, row offset i*Nsyn (64-bit);
//  - isyn0/isyn1, bm0/bm1, q0/q1 are double-buffered by step parity: step t
//    reads input buffer t&1 and scatters into (t+1)&1, so propagation of
//    step t can never reach a neuron's step-t update;
//  - ctl holds a control row per step of the current chunk (kCtl*), plus
//    one more row for the chunk's end stamp.
struct DevNet {
    int32_t n, n_exc, nsyn, words;
    uint64_t seed;
    float ge, gi;
    float *v, *u, *a, *b, *c, *d, *lastI;
    float *isyn0, *isyn1;
    int32_t *nospks;
    const uint2 *syn;
    uint32_t *bm0, *bm1;
    uint32_t *q0, *q1;
    int32_t *ctl;
    uint32_t *raster;
    uint32_t *bar;
    int32_t *err;
    // Slice partition of the targets for the LDS-accumulating schemes: slice
    // p owns neurons [p*L, min((p+1)*L, n)), L a multiple of 32.  soff[i*(P+1)
    // + p] is the index in row i of its first target >= p*L (rows ascending).
    const uint16_t *soff;
    int32_t P, L;
    // Per-slice spike counts of each step of the chunk ([chunk][P]), written
    // without atomics and summed into ctl by k_finish_chunk.
    int32_t *bcnt;
    // Per-step kernels replayed from a HIP graph read the chunk's first
    // step and the run's last step here: tinfo[0] = t0, tinfo[1] = last.
    const int32_t *tinfo;
    // Batched instances (descriptor table in device memory): step of the
    // chunk's first launch, launches of the chunk that advance this
    // instance, and the instance's last step (lastI is written there).
    int32_t t_base, rel_end, last_t, pad_;
};

// Work unit of a batched launch: slice `p` of the instance in slot `slot`;
// slot == kNoWork pads the grid.
constexpr uint32_t kNoWork = 0xFFFFFFFFu;

struct Launch {
    int grid, block;
    hipStream_t stream;
    // Optional: events stamped by the dispatch itself at the kernel's start
    // and end (hipExtLaunchKernelGGL) -- the dominant kernel's timing.
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
};

// Kernel launchers (g4s_kernels.hip).  Each returns the HIP status.
hipError_t launch_init(const DevNet &d, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t val, hipStream_t s);
// Two ranges filled by one launch.
hipError_t launch_fill2_u32(uint32_t *p, size_t n, uint32_t *q, size_t m, uint32_t val, hipStream_t s);
hipError_t launch_update(const DevNet &d, int t, int rel, int last, const Launch &L);
hipError_t launch_push_rows(const DevNet &d, int t, int rel, const Launch &L);
// Scatter every row set in step t's bitmap into isyn[(t+1)&1] (rebuild of
// the pending input buffer when a global-atomic scheme follows one that
// keeps none).
hipError_t launch_push_bitmap(const DevNet &d, int t, hipStream_t s);
// The end of a chunk: its end stamp into control row `steps`; with `reduce`,
// each step's count summed from the per-slice counts (bcnt); then the
// control block (the kCtlHead words at d.err, rows 0..steps) written to
// `host` (device address of mapped pinned host memory) and, if `counts` is
// not null, each step's count to counts[r].  `done`, if not null, is an
// event that the dispatch's own end signals.
hipError_t launch_finish_chunk(const DevNet &d, int steps, bool reduce, int32_t *host, int32_t *counts, hipEvent_t done,
                               hipStream_t s);
hipError_t launch_push_flat(const DevNet &d, int t, int rel, const Launch &L);
hipError_t launch_ap(const DevNet &d, int t, int rel, int last, int prev_count, const Launch &L);
hipError_t launch_slice_offsets(const DevNet &d, uint16_t *soff, hipStream_t s);
// AB step `rel` of the chunk; the absolute step is tinfo[0] + rel.  det:
// ordered accumulation (bit-equal to a sequential ascending push).
hipError_t launch_slice_step(const DevNet &d, int rel, bool det, const Launch &L);
hipError_t launch_skl_slice(const DevNet &d, int t0, int steps, int last_t, bool det, const Launch &L);
// Co-resident workgroups per CU of the SKL slice kernel used for slice
// length L at `block` threads (resident_per_cu of that instantiation).
int skl_slice_blocks_per_cu(int L, int block, bool det);
// Template arguments ("R,det,large") of the slice kernels (k_slice_step,
// k_skl_slice) launched for slice length L at `block` threads, and the
// workgroups of k_slice_step one CU holds at once.
void slice_kernel_targs(int L, int block, bool det, char *out, size_t cap);
int slice_step_blocks_per_cu(int L, int block, bool det);
// Dynamic LDS bytes of the slice kernels (single network at `block`
// threads; batched at 256).
// LDS of one workgroup on MI355X (gfx950): single-network slice kernels run
// one workgroup per CU and may use all of it.
constexpr size_t kSliceLdsMax = 160 * 1024;
size_t slice_lds(int L, int block, bool det);
size_t batch_lds(int L, bool det, int block = 256);
// Batched step `rel` of a chunk: one workgroup per work unit, grid = units
// (padded entries exit).  L.block = 256, 512 or 1024.
hipError_t launch_batch_step(const DevNet *descs, const uint2 *work, int grid, int rel, bool det, size_t lds,
                             const Launch &L);
// One persistent launch running a whole chunk from the device work queue
// `units` (total positions; ctr = {claimed, reserved, error}; done[slot] =
// finished slices of the instance in this chunk).
hipError_t launch_batch_chunk(const DevNet *descs, unsigned long long *units, int32_t *ctr, int32_t *done, int total,
                              bool det, size_t lds, const Launch &L);
int batch_chunk_blocks_per_cu(bool det, size_t lds);
// counts[k * cs + rel] = spikes of instance slots[k] at chunk step rel.
hipError_t launch_batch_counts(const DevNet *descs, const int32_t *slots, int nslots, int cs, int32_t *counts,
                               hipStream_t s);
// Rows [row0, row0+rows) of the outgoing table built on the device, equal
// to build_table's; tab points at row row0.  Nsyn <= 1024.
constexpr int kDeviceBuildMaxNsyn = 1024;
hipError_t launch_build_rows(uint64_t seed, int64_t n, int64_t n_exc, int32_t nsyn, float g_exc, float g_inh, uint2 *tab,
                             int64_t row0, int64_t rows, hipStream_t s);
hipError_t launch_thalamic(uint64_t seed, uint32_t t, int64_t n, float *out, hipStream_t s);
hipError_t launch_normal_words(const uint32_t *x0, const uint32_t *x1, int64_t n, float *out,
                               hipStream_t s);

// Resource usage of one kernel instantiation as the compiler reported it
// for gfx950 (generated at build time into g4s_resources.cpp): name, template
// arguments ("1,false,false"), occupancy in waves per SIMD (limited by the
// kernel's VGPRs, AGPRs and SGPRs), registers, static LDS bytes per
// workgroup and scratch bytes per lane.
struct KernelRes {
    const char *name, *targs;
    int occupancy, vgprs, agprs, sgprs, lds, scratch;
};
extern const KernelRes kKernelRes[];
extern const int kKernelResCount;
const KernelRes *kernel_res(const char *name, const char *targs);
// Workgroups of `block` threads and `dyn_lds` bytes of dynamic LDS that one
// CU holds at once: the compiler's occupancy over the workgroup's waves per
// SIMD, the CU's 160 KB of LDS, and the occupancy API, whichever is least
// (the API alone overstates kernels whose SGPRs limit them, e.g. 106 SGPRs
// hold 7 waves per SIMD, not 8).  0 if the kernel is not in the table.
int resident_per_cu(const char *name, const char *targs, const void *fn, int block, size_t dyn_lds);

// Co-resident workgroups of each persistent kernel at `block` threads.
enum PersistentKernel { kPersistAP };
int persistent_blocks_per_cu(PersistentKernel k, int block);
int device_cus();

}  // namespace g4s
