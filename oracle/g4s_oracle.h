/*
 * g4s_oracle.h -- CPU oracle for the GPU4SNN spike-propagation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (gpu4snn_amd/) never links, loads or calls it.
 *
 * It restates, in plain C compiled with -ffp-contract=off, the reference's
 * method for one Izhikevich network (Crypto-TII/GPU4SNN gpu4snn.cu):
 *   - neuron parameters            gpu4snn.cu:286-320   (initNeuron)
 *   - thalamic input + state update gpu4snn.cu:352-416  (stateupdate)
 *   - regime gains                 gpu4snn.cu:362-374
 *   - propagation semantics        gpu4snn.cu:323-331, 420-467, 554-584, 692-704
 *   - connectivity / weight rule   gpu4snn.cu:876-899   (with the fixes below)
 *   - E/I split                    gpu4snn.cu:744
 *   - step order and raster        gpu4snn.cu:1215-1271
 * with these deliberate changes (the reference is time-seeded and not
 * reproducible): every random draw comes from a counter-based Philox4x32-10
 * stream keyed by a 64-bit seed; targets exclude the presynaptic neuron
 * itself; synaptic input starts at zero in every run; deterministic
 * accumulation adds each postsynaptic input in ascending presynaptic order.
 *
 * Pinning status of every function is listed next to its declaration.
 */
#ifndef G4S_ORACLE_H
#define G4S_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox stream tags: counter word 2 of every draw. */
enum { ORC_TAG_PARAM = 1, ORC_TAG_CONN = 2, ORC_TAG_WEIGHT = 3, ORC_TAG_THAL = 4 };

/* Philox4x32-10 (Salmon et al., SC'11).  Pinned: Random123 known-answer
 * vectors and rocRAND's philox4x32_10_engine (tests/test_oracle_rng.py). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* (2*(x>>9)+1) * 2^-24, exact, in (0,1).  Pinned: closed form. */
float orc_u01(uint32_t x);
/* natural log for positive normal floats.  Pinned: libm log within 3 ulp. */
float orc_logf(float x);
/* square root for positive normal floats.  Pinned: sqrtf within 1 ulp. */
float orc_sqrtf(float x);
/* cos(2*pi*((x>>8)+0.5)/2^24).  Pinned: libm cos within 2e-7 absolute. */
float orc_cos2pi(uint32_t x);
/* Box-Muller normal from two words.  Pinned: moments + KS vs N(0,1). */
float orc_normal(uint32_t x0, uint32_t x1);

/* thalamic N(0,1) draw of neuron i at step t under 64-bit seed. */
float orc_thalamic(uint64_t seed, uint32_t t, uint32_t i);

/* Array forms of the transforms above (for tests over millions of inputs). */
void orc_logf_many(const float *x, float *out, int64_t n);
void orc_sqrtf_many(const float *x, float *out, int64_t n);
void orc_cos2pi_many(const uint32_t *x, float *out, int64_t n);
void orc_normal_many(const uint32_t *x0, const uint32_t *x1, float *out, int64_t n);
/* out[i] = orc_thalamic(seed, t, i0 + i) for i < n. */
void orc_thalamic_many(uint64_t seed, uint32_t t, uint32_t i0, float *out, int64_t n);
/* out[4*k..4*k+3] = philox(ctr = (a0 + k, b, tag, 0), key = seed) */
void orc_philox_many(uint64_t seed, uint32_t a0, uint32_t b, uint32_t tag, uint32_t *out, int64_t n);

/* Regime gains (gpu4snn.cu:362-374): mode 0 quiet 2.5/1, 1 balanced 5/2,
 * 2 irregular 7.5/3; each multiplied by thal_scale in fp32. */
void orc_gains(int regime, float thal_scale, float *ge, float *gi);

/* Neuron parameters (gpu4snn.cu:286-320).  Pinned: closed form per id. */
void orc_init_params(uint64_t seed, int64_t n, int64_t n_exc,
                     float *a, float *b, float *c, float *d, float *v, float *u);

/* Fixed-degree connectivity: nsyn distinct targets != i per row, Floyd's
 * sampling over Philox, rows sorted ascending, weights keyed by (i, j).
 * Pinned: invariants (distinct, no self, sorted, weight range), chi-square
 * uniformity of the subset law on tiny N, in-degree mean. */
int orc_build_synapses(uint64_t seed, int64_t n, int64_t n_exc, int32_t nsyn,
                       uint32_t *post, float *w);

/* Deterministic fp32 simulation of `steps` steps starting at step t0.
 * State arrays are in/out.  isyn holds the input accumulated for the next
 * step (zero at the start of a run).  raster (steps*n bytes, may be NULL)
 * receives 1 for each spike; counts (steps, may be NULL) the spikes per
 * step; lastI (may be NULL) the total input of the last step.
 * Pinned: hand-evaluated single steps, RS/FS textbook firing, brute-force
 * dense propagation on tiny N, and the firing rates of an independent numpy
 * model of the same network at N=2500, Nsyn=1000 in all three regimes
 * (tests/golden/model_rates.json).  The reference notebook's spike totals
 * (nb:297-614) are not a pin: they come from another revision of the
 * reference source and fire about 5x faster than the current one. */
typedef struct {
    int64_t n, n_exc;
    int32_t nsyn;
    uint64_t seed;
    float ge, gi;
    const uint32_t *post;
    const float *w;
    const float *a, *b, *c, *d;
} orc_net;

int orc_simulate(const orc_net *net, int32_t t0, int32_t steps,
                 float *v, float *u, float *isyn, int32_t *nospks,
                 uint8_t *raster, int32_t *counts, float *lastI);

/* orc_simulate with the spiking rows of each step added in descending
 * presynaptic order when `descending` is nonzero: the same terms summed in
 * another order, i.e. one of the orders an atomic scatter may take.  Used
 * to show that reassociation alone moves spike times.  Pinned: descending
 * = 0 is orc_simulate bit for bit; with descending != 0 the per-step sums
 * differ from the ascending ones only by rounding (same spikes for steps
 * whose near-threshold margins exceed the rounding difference). */
int orc_simulate_order(const orc_net *net, int32_t t0, int32_t steps,
                       float *v, float *u, float *isyn, int32_t *nospks,
                       uint8_t *raster, int32_t *counts, float *lastI, int32_t descending);

/* orc_simulate on `threads` host threads (<= 0: one per core), bit-identical
 * to it: each thread updates a range of neurons, then sums into its own
 * range of targets the rows that spiked, in ascending presynaptic order
 * (each row's part for the range found by binary search in the sorted
 * row).  The all-cores CPU baseline.
 * Pinned: equal to orc_simulate bit for bit (tests/test_oracle_model.py). */
int orc_simulate_mt(const orc_net *net, int32_t t0, int32_t steps,
                    float *v, float *u, float *isyn, int32_t *nospks,
                    uint8_t *raster, int32_t *counts, float *lastI, int32_t threads);

/* fp64 twin of orc_simulate with teacher forcing: presynaptic spikes are
 * taken from `forced_raster` (steps*n bytes, e.g. a GPU raster) instead of
 * being generated; only neurons listed in `sub` (nsub ids) are integrated.
 * Records v after every step for each listed neuron in v_trace (steps*nsub)
 * and its own spikes in own_raster (steps*nsub, may be NULL).
 * Pinned: equals orc_simulate's voltages within fp32 rounding when forced
 * with orc_simulate's own raster over a short horizon. */
int orc_simulate_f64_forced(const orc_net *net, int32_t t0, int32_t steps,
                            const uint8_t *forced_raster,
                            const int64_t *sub, int64_t nsub,
                            double *v, double *u, double *v_trace, uint8_t *own_raster);

/* Same integration for a subsample given as ascending disjoint id ranges
 * [lo[r], hi[r]), with the forcing raster bit-packed (steps rows of
 * words_per_row words; bit i & 31 of word i >> 5 = neuron i spiked).  Each
 * spiking row contributes through a binary search of its sorted targets per
 * range, so N = 1e6 networks are cheap.  Inputs are added in ascending
 * presynaptic order, as in orc_simulate_f64_forced.
 * Pinned: bit-identical to orc_simulate_f64_forced on the same subsample
 * (tests/test_oracle_model.py). */
int orc_simulate_f64_forced_ranges(const orc_net *net, int32_t t0, int32_t steps,
                                   const uint32_t *forced_words, int64_t words_per_row,
                                   const int64_t *lo, const int64_t *hi, int32_t nranges,
                                   double *v, double *u, double *v_trace, uint8_t *own_raster);

#ifdef __cplusplus
}
#endif
#endif
