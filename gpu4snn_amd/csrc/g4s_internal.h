// g4s_internal.h -- device-side view of one network and the launch entry
// points shared between g4s_kernels.hip and g4s_runtime.hip.
#pragma once
#include "g4s_math.h"

namespace g4s {

constexpr uint32_t kSentinel = 0xFFFFFFFFu;   // empty spike-queue slot
constexpr int kCtlStride = 4;                 // per-step control words
enum { kCtlCount = 0, kCtlHead = 1, kCtlDone = 2 };
constexpr uint32_t kSpinLimit = 1u << 26;     // bounded device waits (seconds)

// Everything a kernel needs about one network, passed by value.
//  - neuron state/parameters are structure-of-arrays fp32 (the reference's
//    36-byte AoS Neuron, cu:50-62, split so each field is one coalesced
//    stream);
//  - syn is the outgoing table: row i = Nsyn {target, weight-bits} pairs,
//    ascending target, row offset i*Nsyn (64-bit);
//  - in_ptr/in_syn is the incoming table (det mode): entries {source,
//    weight-bits} of target j in ascending source order;
//  - isyn0/isyn1, bm0/bm1, q0/q1 are double-buffered by step parity: step t
//    reads input buffer t&1 and scatters into (t+1)&1, so propagation of
//    step t can never reach a neuron's step-t update;
//  - ctl holds [count, head, done, -] per step of the current chunk.
struct DevNet {
    int32_t n, n_exc, nsyn, words;
    uint64_t seed;
    float ge, gi;
    float *v, *u, *a, *b, *c, *d, *lastI;
    float *isyn0, *isyn1;
    int32_t *nospks;
    const uint2 *syn;
    const int64_t *in_ptr;
    const uint2 *in_syn;
    uint32_t *bm0, *bm1;
    uint32_t *q0, *q1;
    int32_t *ctl;
    uint32_t *raster;
    uint32_t *bar;
    int32_t *err;
};

struct Launch {
    int grid, block;
    hipStream_t stream;
};

// Kernel launchers (g4s_kernels.hip).  Each returns the HIP status.
hipError_t launch_init(const DevNet &d, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, size_t n, uint32_t val, hipStream_t s);
hipError_t launch_update(const DevNet &d, int t, int rel, int last, const Launch &L);
hipError_t launch_push_rows(const DevNet &d, int t, int rel, const Launch &L);
hipError_t launch_push_flat(const DevNet &d, int t, int rel, const Launch &L, bool cooperative);
hipError_t launch_ap(const DevNet &d, int t, int rel, int last, const Launch &L);
hipError_t launch_skl(const DevNet &d, int t0, int steps, int last_t, const Launch &L);
hipError_t launch_pull_thread(const DevNet &d, int t, int rel, int last, const Launch &L, bool cooperative);
hipError_t launch_pull_wave(const DevNet &d, int t, int rel, int last, const Launch &L);
hipError_t launch_ap_det(const DevNet &d, int t, int rel, int last, const Launch &L);
hipError_t launch_skl_det(const DevNet &d, int t0, int steps, int last_t, const Launch &L);
hipError_t launch_gather_isyn(const DevNet &d, int t, hipStream_t s);
hipError_t launch_thalamic(uint64_t seed, uint32_t t, int64_t n, float *out, hipStream_t s);
hipError_t launch_normal_words(const uint32_t *x0, const uint32_t *x1, int64_t n, float *out,
                               hipStream_t s);

// Co-resident workgroups of each cooperative kernel at `block` threads.
enum CoopKernel { kCoopPushFlat, kCoopAP, kCoopSKL, kCoopPull, kCoopAPDet, kCoopSKLDet };
int coop_max_blocks(CoopKernel k, int block);

}  // namespace g4s
