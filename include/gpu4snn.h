/*
 * gpu4snn.h -- C ABI of the MI355X (gfx950) GPU4SNN spike-propagation library
 * (gpu4snn_amd/libgpu4snn.so).
 *
 * The reference (Crypto-TII/GPU4SNN, gpu4snn.cu; "cu:L" below) has no library
 * API: its hot path is a set of kernels called from main().  Each entry point
 * here names the reference code it replaces.  Conventions:
 *   - every function returns G4S_OK (0) or a negative G4S_ERR_* code and
 *     sets a thread-local message readable with g4s_last_error();
 *   - the library owns all device memory; callers own every host buffer;
 *   - a handle must not be used from two threads at once;
 *   - all device work runs on the library's own HIP stream of the handle.
 */
#ifndef GPU4SNN_H
#define GPU4SNN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define G4S_OK               0
#define G4S_ERR_ARG         -1  /* invalid argument (N <= 1, Nsyn >= N, unknown alg ...) */
#define G4S_ERR_HIP         -2  /* HIP runtime error (no device, launch failure ...)     */
#define G4S_ERR_MEM         -3  /* device or host allocation failed                      */
#define G4S_ERR_STATE       -4  /* call not valid in the handle's state                  */
#define G4S_ERR_UNSUPPORTED -5  /* size outside what the selected scheme supports        */
#define G4S_ERR_TIMEOUT     -6  /* a bounded device-side wait expired; run aborted       */
#define G4S_ERR_BUDGET      -7  /* serving: request exceeds the device-memory budget      */

/* Propagation schemes: the reference's `dynamic` selector (cu:1083, cu:1230). */
#define G4S_ALG_AP  0  /* per-spike work items (reference: child grid per spike, cu:408-409) */
#define G4S_ALG_N   1  /* neuron-parallel (reference deliverspks1, cu:420-443)            */
#define G4S_ALG_S   2  /* synapse-parallel (reference deliverspks2, cu:446-467)           */
#define G4S_ALG_AB  3  /* slice-owned propagation, one launch per step (AB_deliverspks, cu:554-584) */
#define G4S_ALG_SKL 4  /* a whole chunk of steps in one persistent launch (SKL_deliverspks, cu:669-716) */

/* Accumulation modes of the synaptic input. */
#define G4S_ACCUM_ATOMIC 0 /* fp32 atomic scatter, any order: the reference's semantics   */
#define G4S_ACCUM_DET    1 /* ordered gather: each input summed in ascending presynaptic id */

/* Network specification.  Replaces main()'s argv + hard-coded constants
 * (cu:738-745, cu:362-374, cu:876-899). */
typedef struct g4s_spec {
    int64_t  n;          /* neurons N (cu:738)                                     */
    int32_t  nsyn;       /* outgoing synapses per neuron N_syn (cu:739), < N        */
    int32_t  regime;     /* 0 quiet, 1 balanced, 2 irregular (cu:362-374)           */
    uint64_t seed;       /* key of every Philox stream of this network              */
    float    thal_scale; /* multiplies the regime gains; 1.0 reproduces cu:362-374  */
    int32_t  reserved;   /* must be 0                                               */
    int64_t  n_exc;      /* excitatory neurons; < 0 selects floor(4N/5) (cu:744)    */
} g4s_spec;

typedef struct g4s_net g4s_net;

/* Tuning of one run; zero-initialised fields select the defaults. */
typedef struct g4s_run_opts {
    int32_t block;       /* threads per workgroup (multiple of 64; <= 256, or <= 1024
                            for AB/SKL), 0 = chosen per scheme (AB/SKL: one thread
                            per neuron of a slice, at most 1024)                   */
    int32_t slices;      /* AB/SKL: target slices P (workgroups), 0 = about one per
                            CU, more if a slice's LDS workspace would not fit      */
    int32_t chunk;       /* steps per device chunk between host copies, 0 = 1000       */
    int32_t graph;       /* AB: >= 0 replays the chunk's per-step launches from a
                            captured HIP graph, < 0 launches them directly          */
    int32_t kernel_events; /* != 0: bracket every launch of the dominant kernel with
                            HIP events (g4s_timing.kernel_ms); AB steps are then
                            launched directly instead of from a graph              */
} g4s_run_opts;

/* Timing of the last g4s_run call. */
typedef struct g4s_timing {
    double  device_ms;   /* device time of the step loop, raster copies excluded: HIP  *
                          * events around each chunk's launches; SKL without kernel    *
                          * events: its launch's start to the chunk's end stamp (device *
                          * wall clock)                                                */
    double  wall_ms;     /* host steady-clock time of the whole call                   */
    int64_t steps;       /* steps simulated                                            */
    int64_t spikes;      /* spikes in those steps                                      */
    int32_t launches;    /* kernel launches issued                                     */
    int32_t grid;        /* workgroups of the dominant kernel                          */
    double  kernel_ms;   /* with kernel_events: summed event-timed duration of the
                            dominant kernel's launches, else 0                         */
    int32_t kernel_launches; /* launches that kernel_ms covers                         */
    int32_t reserved;
} g4s_timing;

/* Allocate device memory, build the connectivity table, initialise neurons
 * and zero the synaptic input.  Replaces cu:753-777 (allocation), cu:862-912
 * (host connectivity sampler + upload) and cu:1094 (initNeuron). */
int g4s_create(const g4s_spec *spec, g4s_net **out);

/* Re-initialise neuron parameters/state (initNeuron, cu:286-320), zero the
 * synaptic input and spike flags (fixes cu:776-777), and set step t = 0. */
int g4s_reset(g4s_net *net);

/* Simulate `steps` steps with scheme `alg` and accumulation mode `accum`,
 * continuing from the handle's current step.  Replaces the step loop
 * cu:1215-1271 (AP/N/S/AB) and the SKL launch cu:1282-1284.
 *   raster : NULL or steps * ceil(N/32) words; bit (i & 31) of word
 *            [s * ceil(N/32) + (i >> 5)] is set when neuron i spiked at step
 *            t0 + s (the reference's spiketimes rows, cu:1256-1261)
 *   counts : NULL or `steps` spike counts (SpksPerStep, cu:1267)
 *   opts   : NULL for defaults */
int g4s_run(g4s_net *net, int alg, int accum, int32_t steps,
            uint32_t *raster, int32_t *counts, const g4s_run_opts *opts);

/* Streaming form of g4s_run: after each chunk (opts->chunk steps) the
 * callback receives its raster rows (rows x ceil(N/32) words, layout as
 * above) and per-step counts, starting at step t0; both buffers are valid
 * only during the call.  The next chunk already runs on the device while the
 * callback works (except for AP in atomic mode and kernel_events runs), so
 * writing the reference's spiketimes.csv rows (cu:1256-1261) overlaps the
 * simulation.  A nonzero return stops the run (G4S_ERR_STATE). */
typedef int (*g4s_chunk_fn)(void *user, int64_t t0, int32_t rows, const uint32_t *raster, const int32_t *counts);
int g4s_run_stream(g4s_net *net, int alg, int accum, int32_t steps, g4s_chunk_fn fn, void *user,
                   const g4s_run_opts *opts);

int g4s_get_timing(const g4s_net *net, g4s_timing *out);

/* No counterpart in the reference.  For callers that cannot allocate
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  g4s_run on a handle that a g4s_run of the same scheme
 * has prepared (its slice partition and staging exist; a zero-step g4s_run
 * with the same options prepares it), writing the raster rows (layout as
 * g4s_run's, steps x ceil(N/32) words) and the per-step spike counts into
 * the caller's DEVICE buffers raster_dev and counts_dev (either may be
 * NULL).  It calls no hipMalloc, hipFree, hipHostMalloc, event or graph
 * creation: every buffer the run needs beyond the network's own state (the
 * per-step control rows, per-slice spike counts, grid-barrier counters, the
 * error flag and step info) comes from one static device array of
 * G4S_NOALLOC_SCRATCH_BYTES per device, held until the process ends.  Runs
 * are cut into chunks that fit it; a network whose slices need more than it
 * holds for one step (P > about 260000 slices) gets G4S_ERR_UNSUPPORTED.
 * It uses the slice partition the handle holds (opts->slices is not
 * applied) and launches AB steps directly (no graph); kernel_events is
 * refused (G4S_ERR_ARG).  Every call on one device shares that device's
 * array, which keeps the last call's contents: calls that overlap on one
 * device (two threads, two handles) race and must be serialised by the
 * caller.  Returns when the run is complete, as g4s_run. */
#define G4S_NOALLOC_SCRATCH_BYTES (1 << 20)
int g4s_run_noalloc(g4s_net *net, int alg, int accum, int32_t steps, uint32_t *raster_dev, int32_t *counts_dev,
                    const g4s_run_opts *opts);

/* Device time of each step of the last g4s_run / g4s_run_stream, in ms:
 * the interval between the starts of consecutive steps on the device's
 * constant-rate wall clock (the last step of a chunk ends at the chunk's end
 * stamp), i.e. kernel time plus the launch gaps inside the chunk.  Copies at
 * most n values and returns how many (>= 0), or a G4S_ERR_* code.  The
 * reference's per-step TimeKernelUpdate/TimeSpent rows of computetime.csv
 * (cu:1219-1267). */
int g4s_get_step_times(const g4s_net *net, float *ms, int32_t n);

/* Launch report of the dominant kernel of the last g4s_run: the counterpart
 * of the grid, block and occupancy figures the reference prints for each
 * kernel it launches (cu:800-838, cu:1089-1201). */
typedef struct g4s_launch_info {
    char    kernel[64];       /* kernel and template arguments, e.g. "k_skl_slice<1,false,false>" */
    int32_t grid, block;      /* workgroups and threads per workgroup                      */
    int32_t lds_bytes;        /* dynamic + static LDS per workgroup                        */
    int32_t vgprs, sgprs;     /* compiler's register allocation: per lane, per wave        */
    int32_t waves_per_simd;   /* compiler's occupancy for those registers                  */
    int32_t resident_per_cu;  /* workgroups of this launch one CU holds at once            */
    int32_t cus;              /* compute units of the device                               */
    int32_t persistent;       /* 1: its workgroups wait on each other, so the grid is at
                                 most resident_per_cu * cus                               */
} g4s_launch_info;
int g4s_get_launch_info(g4s_net *net, g4s_launch_info *out);

/* Host-only: the compiler's resource report of one kernel instantiation of
 * this build (name "k_skl_slice", template arguments "1,false,false"; "" for
 * a kernel that is not a template).  out[0..5] = waves per SIMD, VGPRs,
 * AGPRs, SGPRs, static LDS bytes per workgroup, scratch bytes per lane.
 * G4S_ERR_ARG if the build has no such kernel.  No reference counterpart:
 * the reference asks the occupancy API (cu:800-838). */
int g4s_kernel_resources(const char *name, const char *targs, int32_t out[6]);

/* Change the firing regime / thalamic scale of an existing network without
 * rebuilding its connectivity: the reference's per-run `mode` argument of
 * stateupdate (cu:344-347, cu:362-374, loop cu:1082). */
int g4s_set_regime(g4s_net *net, int regime, float thal_scale);

/* Current step index (number of steps simulated since the last reset). */
int64_t g4s_current_step(const g4s_net *net);

/* A propagation plan: scheme, accumulation mode and tiling. */
typedef struct g4s_plan {
    int32_t alg;          /* G4S_ALG_*                                        */
    int32_t accum;        /* G4S_ACCUM_*                                      */
    g4s_run_opts opts;    /* tiling (block, slices), chunk, graph             */
    float   us_per_step;  /* measured device time of the plan                 */
    float   rate_hz;      /* firing rate seen while measuring                 */
} g4s_plan;

/* Autotuner (no reference counterpart; the reference fixes its launch shapes
 * by occupancy queries, cu:780-838, cu:957-1079).  Times every scheme and
 * tiling candidate for `accum` on this network (trial_steps steps each after
 * a short warm-up) and returns the fastest.  Plans are cached per bucket of
 * (log2 N, Nsyn, log2 firing rate, accum) for the life of the process; a
 * bucket hit returns the cached plan without timing unless `force` is set.
 * The network is reset (step 0) on return. */
int g4s_autotune(g4s_net *net, int accum, int32_t trial_steps, int force, g4s_plan *best);

/* Copy the structure-of-arrays state out (any pointer may be NULL).  I is the
 * total input of the last simulated step; replaces the AoS Neuron copy-back
 * behind neuroninfo(after).csv (cu:1099, cu:1318-1327). */
int g4s_get_state(g4s_net *net, float *v, float *u, float *I, int32_t *nospks);
int g4s_get_params(g4s_net *net, float *a, float *b, float *c, float *d);
/* Overwrite v and u (NULL leaves one unchanged), e.g. to inject a NaN. */
int g4s_set_state(g4s_net *net, const float *v, const float *u);

/* Copy the connectivity table: row i holds the Nsyn targets of neuron i in
 * ascending order (cu:895-896 layout [N][N_syn]) and their weights. */
int g4s_get_synapses(g4s_net *net, uint32_t *post, float *w);

/* Free every device allocation of the handle (cu:1331 cudaDeviceReset). */
int g4s_destroy(g4s_net *net);

const char *g4s_last_error(void);
int g4s_version(void);

/* Host-only: the connectivity builder without a device (same table as
 * g4s_create uploads).  `threads` <= 0 uses the hardware concurrency. */
int g4s_build_synapses_host(const g4s_spec *spec, uint32_t *post, float *w, int32_t threads);

/* The same table built by the device builder (used by g4s_create and the
 * batched form whenever Nsyn <= 1024) and copied back; Nsyn > 1024 returns
 * G4S_ERR_UNSUPPORTED.  No reference counterpart: the reference's device
 * generator is never launched (cu:220-268, cu:915-953). */
int g4s_build_synapses_device(const g4s_spec *spec, uint32_t *post, float *w);

/* Device diagnostics: thalamic draws of neurons [0, n) at step t under
 * `seed` computed by a gfx950 kernel, and the Box-Muller transform of given
 * word pairs on the device. */
int g4s_device_thalamic(uint64_t seed, uint32_t t, int64_t n, float *out);
int g4s_device_normal_words(const uint32_t *x0, const uint32_t *x1, int64_t n, float *out);

/* ------------------------------------------------------------------------
 * Batched form (no reference counterpart: the reference simulates one
 * network per process, cu:735-1332).  A batch holds many independent
 * networks of uneven N/Nsyn resident on one device.  Each chunk advances
 * every unfinished instance by up to `chunk` steps: by default ONE persistent
 * launch per chunk whose workgroups take work units (one slice of one
 * instance at one step, the AB slice scheme) from a device queue, each
 * instance's next step queued as soon as its previous step completes; or one
 * launch per step over all instances, replayed from a captured HIP graph.
 * Instances may be added and removed between chunks (continuous batching);
 * an instance's results do not depend on which other instances share its
 * launches.
 * ---------------------------------------------------------------------- */
typedef struct g4s_batch g4s_batch;

typedef struct g4s_batch_opts {
    int32_t chunk;       /* steps per chunk, 0 = 100                                */
    int32_t slice_len;   /* neurons per slice (multiple of 64), 0 = the bucket's
                            g4s_batch_tune choice, else min(768, N)              */
    int32_t graph;       /* schedule 0: >= 0 replays the chunk's launches from a HIP
                            graph, < 0 launches them directly                       */
    int32_t schedule;    /* 0: one launch per step over all instances;
                            1: one persistent launch per chunk, work units queued on
                            the device as each instance's previous step completes */
    int64_t budget;      /* device bytes the batch may hold, 0 = no limit            */
    int32_t block;       /* schedule 0: threads per workgroup, 256 (0), 512 or 1024   */
    int32_t streams;     /* schedule 0: the instances are split into this many groups
                            of about equal cost, each stepped on its own stream, so
                            one group's launches fill another's tail; 0 = 3, 1..4    */
} g4s_batch_opts;

typedef struct g4s_inst_info {
    int64_t n;           /* neurons                                                  */
    int32_t nsyn;        /* synapses per neuron                                      */
    int32_t slice_len;   /* neurons per slice in use                                 */
    int64_t t;           /* steps simulated                                          */
    int64_t steps;       /* steps requested                                          */
    int64_t spikes;      /* spikes so far                                            */
    int64_t bytes;       /* device bytes held                                        */
    int32_t done;        /* t == steps                                               */
    int32_t slices;      /* slices (workgroups per launch)                           */
} g4s_inst_info;

/* Device bytes one instance of `spec` holds in a batch with this chunk
 * length and slice length (0 = the batch default): state, connectivity,
 * slice offsets, raster ring and counters.  The admission rule of serving. */
int64_t g4s_batch_footprint(const g4s_spec *spec, int32_t chunk, int32_t slice_len);

int g4s_batch_create(int accum, const g4s_batch_opts *opts, g4s_batch **out);

/* Build an instance (connectivity on the host, uploaded) and schedule
 * `steps` steps.  keep_raster != 0 keeps its raster rows on the host.
 * Returns G4S_ERR_BUDGET when it does not fit the remaining budget. */
int g4s_batch_add(g4s_batch *b, const g4s_spec *spec, int64_t steps, int keep_raster, int32_t *slot);

/* Run one chunk; *active = instances advanced (0: nothing left to do). */
int g4s_batch_step(g4s_batch *b, int32_t *active);

/* Run chunks until every instance is done. */
int g4s_batch_run(g4s_batch *b);

int g4s_batch_info(const g4s_batch *b, int32_t slot, g4s_inst_info *info);
/* Per-step spike counts (info.t entries) and raster rows (info.t x
 * ceil(N/32) words, keep_raster instances only) of an instance. */
int g4s_batch_counts(const g4s_batch *b, int32_t slot, int32_t *counts);
int g4s_batch_raster(const g4s_batch *b, int32_t slot, uint32_t *words);
int g4s_batch_state(g4s_batch *b, int32_t slot, float *v, float *u, float *I, int32_t *nospks);
/* Release an instance; its slot may be reused.  Its device block is kept
 * for later instances that fit it (up to 16 GB of such blocks per batch,
 * the oldest freed beyond), and freed by g4s_batch_destroy. */
int g4s_batch_remove(g4s_batch *b, int32_t slot);
/* Timing of the last g4s_batch_step / g4s_batch_run call (launches = kernel
 * launches of the batched step kernel, grid = its largest grid). */
int g4s_batch_timing(const g4s_batch *b, g4s_timing *out);
int g4s_batch_destroy(g4s_batch *b);

/* Batch autotuner on one network shape (no reference counterpart).  Times
 * the slice lengths 128 .. 2048 on a batch of copies of `rep` (different
 * seeds; enough copies to fill the device, so throughput decides) for
 * trial_steps steps after a warm-up, and records the fastest for the bucket
 * (floor(log2 N), Nsyn, drive class, accum; see g4s_batch_tune_mix) for the
 * life of the process: g4s_batch_add then uses it for every instance in that
 * bucket when the batch's slice_len is 0.  A recorded bucket is returned
 * without timing unless `force`.  On a mix of shapes g4s_batch_tune_mix
 * judges better: a homogeneous trial does not see which instances set the
 * mix's step time. */
typedef struct g4s_batch_plan {
    int32_t slice_len;         /* chosen neurons per slice                  */
    int32_t copies;            /* instances timed (0: cached plan)           */
    float   us_per_inst_step;  /* device time per instance and step          */
    float   rate_hz;           /* firing rate seen while timing              */
} g4s_batch_plan;
int g4s_batch_tune(int accum, const g4s_spec *rep, int32_t trial_steps, int force, g4s_batch_plan *out);

/* Batch autotuner on the mix itself (no reference counterpart).  Times the
 * whole batch of `specs` -- built from the same seeds for every trial, with
 * the batch options `opts` (schedule 0, slice_len 0) -- for trial_steps
 * steps after as many warm-up steps, and by coordinate descent over the
 * buckets (floor(log2 N), Nsyn, drive class = regime and thalamic scale in
 * quarters, which set the firing rate), heaviest bucket (most synapses)
 * first, moves one bucket's slice length at a time through 256 .. 2048 while
 * the mix's device time per step drops by more than 1.5%, within max_trials
 * trials (<= 0: 40).  The plan is then timed against the untuned one
 * (every instance at min(768, N)), twice each, alternated, and kept only if
 * faster: it is recorded
 * per bucket for the life of the process, and g4s_batch_add with slice_len 0
 * uses it.  The single-network autotuner (g4s_autotune) also picks the
 * scheme; a batch runs every instance with the slice scheme, the only one
 * that serves many networks in one launch. */
typedef struct g4s_mix_plan {
    int32_t buckets;             /* buckets in the mix                      */
    int32_t changed;             /* buckets whose length the tuner moved    */
    int32_t trials;              /* timed trials                            */
    float   us_per_step_default; /* device us per step, every instance at min(768, N) */
    float   us_per_step_tuned;   /* device us per step with the kept plan   */
    float   us_per_step_first;   /* first trial (the starting plan)          */
} g4s_mix_plan;
int g4s_batch_tune_mix(int accum, const g4s_spec *specs, int32_t count, int32_t trial_steps,
                       const g4s_batch_opts *opts, int32_t max_trials, g4s_mix_plan *out);

/* ------------------------------------------------------------------------
 * Serving (no reference counterpart).  A server owns one batch and a worker
 * thread that admits submitted requests in arrival order while their
 * footprint (g4s_batch_footprint) fits the budget, runs one chunk of every
 * admitted request per iteration (requests arriving meanwhile join at the
 * next chunk: continuous batching), and retires finished requests, copying
 * their counts (and raster) to host memory and releasing their instances.
 * All functions are thread-safe.  Times are seconds since server creation.
 * ---------------------------------------------------------------------- */
typedef struct g4s_server g4s_server;

typedef struct g4s_server_opts {
    int32_t accum;       /* G4S_ACCUM_* of every request                        */
    int32_t chunk;       /* steps per chunk, 0 = 100                            */
    int32_t slice_len;   /* as g4s_batch_opts.slice_len, 0 = by load: 256 while
                            fewer than 32 requests are resident, 384 while
                            fewer than 192, else 768                          */
    int32_t reserved;    /* must be 0                                           */
    int64_t budget;      /* device bytes for resident requests, 0 = no limit   */
} g4s_server_opts;

#define G4S_REQ_PENDING 0
#define G4S_REQ_ACTIVE  1
#define G4S_REQ_DONE    2
#define G4S_REQ_FAILED  3

typedef struct g4s_request_info {
    int32_t state;       /* G4S_REQ_*                                           */
    int32_t error;       /* G4S_ERR_* of a failed request                       */
    int64_t n;           /* neurons                                             */
    int64_t steps;       /* steps requested                                     */
    int64_t t;           /* steps simulated (set when done)                     */
    int64_t spikes;      /* spikes (set when done)                              */
    double  submit_s, admit_s, done_s;
} g4s_request_info;

typedef struct g4s_server_stats {
    int64_t submitted, completed, failed, pending, active;
    int64_t chunks;          /* chunks run                                      */
    int64_t steps;           /* instance-steps simulated                        */
    int64_t sim_steps_done;  /* steps of completed requests                     */
    int64_t bytes_in_use, max_bytes, max_active;
    double  busy_ms;         /* host time inside chunk calls                    */
    double  device_ms;       /* device time of the chunks                       */
    double  now_s;
} g4s_server_stats;

int g4s_server_create(const g4s_server_opts *opts, g4s_server **out);
/* Queue a request; G4S_ERR_BUDGET if it can never fit the budget. */
int g4s_server_submit(g4s_server *s, const g4s_spec *spec, int64_t steps, int keep_raster, int64_t *id);
int g4s_server_info(g4s_server *s, int64_t id, g4s_request_info *info);
/* Block until the request is done (G4S_OK) or failed (its error code);
 * G4S_ERR_TIMEOUT after timeout_s (<= 0: no limit). */
int g4s_server_wait(g4s_server *s, int64_t id, double timeout_s);
/* Block until nothing is pending or active. */
int g4s_server_drain(g4s_server *s, double timeout_s);
int g4s_server_counts(g4s_server *s, int64_t id, int32_t *counts);
int g4s_server_raster(g4s_server *s, int64_t id, uint32_t *words);
/* Forget a finished request's results. */
int g4s_server_release(g4s_server *s, int64_t id);
int g4s_server_get_stats(g4s_server *s, g4s_server_stats *out);
/* Stop the worker (requests still queued are dropped) and free everything. */
int g4s_server_destroy(g4s_server *s);

#ifdef __cplusplus
}
#endif
#endif
