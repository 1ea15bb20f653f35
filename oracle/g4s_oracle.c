/*
 * g4s_oracle.c -- CPU oracle for the GPU4SNN hot path (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this file, and only as the checker.  Built with -O2 -ffp-contract=off
 * -fno-fast-math on x86-64 (SSE float arithmetic, no flush-to-zero), so each
 * float operation written here is one IEEE-754 binary32 operation.
 *
 * Reference citations are to Crypto-TII/GPU4SNN gpu4snn.cu (cu:L).
 */
#include "g4s_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <unistd.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers:  */
/* as easy as 1, 2, 3", SC'11.  Round: two 32x32->64 products of words 0  */
/* and 2 with the Philox multipliers, xor with words 1/3 and the key;     */
/* the key is bumped by the Weyl constants between rounds.                */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void draw(uint64_t seed, uint32_t a, uint32_t b, uint32_t tag, uint32_t out[4])
{
    uint32_t ctr[4] = {a, b, tag, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* ------------------------------------------------------------------ */
/* Transforms.  Only +, -, *, fmaf and integer operations are used, so  */
/* any IEEE-754 implementation gives the same bits.                     */
/* ------------------------------------------------------------------ */
float orc_u01(uint32_t x)
{
    /* 2*(x>>9)+1 < 2^24 is exact in binary32; scaling by 2^-24 is exact. */
    return (float)(((x >> 9) << 1) | 1u) * 0x1p-24f;
}

/* 1/y for y in [1,4): magic-constant seed, three Newton steps
 * r <- r + r*(1 - y*r) written with fmaf. */
static float rcp_nr(float y)
{
    float r = u2f(0x7EF311C3u - f2u(y));
    for (int k = 0; k < 3; ++k) {
        float e = fmaf(-y, r, 1.0f);
        r = fmaf(r, e, r);
    }
    return r;
}

/* log(x) = e*ln2 + 2*atanh(s), s = (m-1)/(m+1), m in [sqrt(1/2), sqrt(2)).
 * |s| <= 0.1716, the atanh series is cut after s^11 (error < 1e-11). */
float orc_logf(float x)
{
    uint32_t bits = f2u(x);
    int32_t e = (int32_t)(bits >> 23) - 127;
    float m = u2f((bits & 0x007FFFFFu) | 0x3F800000u);
    if (m > 0x1.6a09e6p+0f) {   /* 0x3FB504F3 */
        m = m * 0.5f;
        e += 1;
    }
    float f = m - 1.0f;
    float s = f * rcp_nr(m + 1.0f);
    float z = s * s;
    float p = fmaf(z, 0x1.745d18p-4f, 0x1.c71c72p-4f);
    p = fmaf(z, p, 0x1.24924ap-3f);
    p = fmaf(z, p, 0x1.99999ap-3f);
    p = fmaf(z, p, 0x1.555556p-2f);
    p = p * z;
    float s2 = s + s;
    float lm = fmaf(s2, p, s2);
    float ef = (float)e;
    /* ln2 split: hi has 16 trailing zero bits so ef*hi is exact. */
    return fmaf(ef, 0x1.62e4p-1f, fmaf(ef, 0x1.7f7d1cp-20f, lm));
}

/* sqrt(x): inverse square root by magic seed + three Newton steps, then one
 * residual correction s <- s + (x - s*s) * y/2. */
float orc_sqrtf(float x)
{
    float y = u2f(0x5F3759DFu - (f2u(x) >> 1));
    float h = 0.5f * x;
    for (int k = 0; k < 3; ++k) {
        float hy = h * y;
        y = y * fmaf(-hy, y, 1.5f);
    }
    float s = x * y;
    float r = fmaf(-s, s, x);
    return fmaf(r, 0.5f * y, s);
}

/* Taylor polynomials on [0, pi/4]; truncation error < 2e-9. */
static float sin_poly(float x)
{
    float x2 = x * x;
    float p = fmaf(x2, 0x1.71de3ap-19f, -0x1.a01a02p-13f);
    p = fmaf(x2, p, 0x1.111112p-7f);
    p = fmaf(x2, p, -0x1.555556p-3f);
    return fmaf(x * x2, p, x);
}

static float cos_poly(float x)
{
    float x2 = x * x;
    float p = fmaf(x2, -0x1.27e4fcp-22f, 0x1.a01a02p-16f);
    p = fmaf(x2, p, -0x1.6c16c2p-10f);
    p = fmaf(x2, p, 0x1.555556p-5f);
    p = fmaf(x2, p, -0.5f);
    return fmaf(x2, p, 1.0f);
}

/* cos(2*pi*(k+0.5)/2^24), k = x>>8.  The reduction is exact in integers:
 * quadrant q = k>>22, offset r in [0,2^22); offsets in the upper half of a
 * quadrant are reflected (r -> 2^22-1-r) and sin/cos swapped, so the
 * polynomial argument (2r+1)*pi/2^24 lies in (0, pi/4]. */
float orc_cos2pi(uint32_t x)
{
    uint32_t k = x >> 8;
    uint32_t q = k >> 22;
    uint32_t r = k & 0x3FFFFFu;
    int use_sin = (int)(q & 1u);
    if (r >= 0x200000u) {
        r = 0x3FFFFFu - r;
        use_sin ^= 1;
    }
    float a = (float)(2u * r + 1u) * 0x1.921fb6p-23f;   /* pi / 2^24 */
    float v = use_sin ? sin_poly(a) : cos_poly(a);
    return (q == 1u || q == 2u) ? -v : v;
}

float orc_normal(uint32_t x0, uint32_t x1)
{
    float u1 = orc_u01(x0);
    float l = orc_logf(u1);
    float w = -2.0f * l;
    float rad = orc_sqrtf(w);
    return rad * orc_cos2pi(x1);
}

float orc_thalamic(uint64_t seed, uint32_t t, uint32_t i)
{
    uint32_t o[4];
    draw(seed, t, i, ORC_TAG_THAL, o);
    return orc_normal(o[0], o[1]);
}

void orc_logf_many(const float *x, float *out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) out[i] = orc_logf(x[i]);
}

void orc_sqrtf_many(const float *x, float *out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) out[i] = orc_sqrtf(x[i]);
}

void orc_cos2pi_many(const uint32_t *x, float *out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) out[i] = orc_cos2pi(x[i]);
}

void orc_normal_many(const uint32_t *x0, const uint32_t *x1, float *out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) out[i] = orc_normal(x0[i], x1[i]);
}

void orc_thalamic_many(uint64_t seed, uint32_t t, uint32_t i0, float *out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) out[i] = orc_thalamic(seed, t, i0 + (uint32_t)i);
}

void orc_philox_many(uint64_t seed, uint32_t a0, uint32_t b, uint32_t tag, uint32_t *out, int64_t n)
{
    for (int64_t k = 0; k < n; ++k) draw(seed, a0 + (uint32_t)k, b, tag, out + 4 * k);
}

/* cu:362-374 */
void orc_gains(int regime, float thal_scale, float *ge, float *gi)
{
    float e, g;
    switch (regime) {
    case 0: e = 2.5f; g = 1.0f; break;
    case 2: e = 7.5f; g = 3.0f; break;
    default: e = 5.0f; g = 2.0f; break;
    }
    *ge = e * thal_scale;
    *gi = g * thal_scale;
}

/* cu:286-320.  r*r replaces powf(r,2) (identical for a square). */
void orc_init_params(uint64_t seed, int64_t n, int64_t n_exc,
                     float *a, float *b, float *c, float *d, float *v, float *u)
{
    for (int64_t i = 0; i < n; ++i) {
        uint32_t o[4];
        draw(seed, (uint32_t)i, 0u, ORC_TAG_PARAM, o);
        float r = orc_u01(o[0]);
        if (i < n_exc) {
            float r2 = r * r;
            a[i] = 0.02f;
            b[i] = 0.2f;
            c[i] = -65.0f + 15.0f * r2;
            d[i] = 8.0f - 6.0f * r2;
        } else {
            a[i] = 0.02f + 0.08f * r;
            b[i] = 0.25f - 0.05f * r;
            c[i] = -65.0f;
            d[i] = 2.0f;
        }
        v[i] = -65.0f;
        u[i] = v[i] * b[i];
    }
}

static int cmp_u32(const void *x, const void *y)
{
    uint32_t p = *(const uint32_t *)x, q = *(const uint32_t *)y;
    return (p > q) - (p < q);
}

/* cu:876-899 restated.  Row i draws nsyn distinct values from the m = n-1
 * non-self candidates with Floyd's algorithm: for jj = m-nsyn .. m-1 draw
 * t uniform on [0, jj] (64-bit multiply-shift of Philox words 0,1 under
 * counter (i, d, TAG_CONN)); insert t, or jj when t is already present.
 * A value v maps to neuron v + (v >= i).  The row is sorted ascending.
 * The weight of synapse (i -> j) is g * U(Philox(i, j, TAG_WEIGHT).x0) with
 * g = 0.5*1000/nsyn for excitatory rows and -1000/nsyn for inhibitory rows
 * (cu:890, cu:893). */
int orc_build_synapses(uint64_t seed, int64_t n, int64_t n_exc, int32_t nsyn,
                       uint32_t *post, float *w)
{
    if (n < 2 || nsyn < 1 || (int64_t)nsyn > n - 1 || n > 0x7FFFFFFFLL) return -1;
    const uint64_t m = (uint64_t)(n - 1);
    uint8_t *mark = (uint8_t *)calloc((size_t)m, 1);
    if (!mark) return -2;
    const float scale = 1000.0f / (float)nsyn;
    const float g_exc = scale * 0.5f, g_inh = -scale;
    for (int64_t i = 0; i < n; ++i) {
        uint32_t *row = post + (size_t)i * (size_t)nsyn;
        for (int32_t d = 0; d < nsyn; ++d) {
            uint64_t jj = m - (uint64_t)nsyn + (uint64_t)d;
            uint32_t o[4];
            draw(seed, (uint32_t)i, (uint32_t)d, ORC_TAG_CONN, o);
            uint64_t r = ((uint64_t)o[1] << 32) | o[0];
            uint64_t t = (uint64_t)(((unsigned __int128)r * (jj + 1u)) >> 64);
            uint64_t val = mark[t] ? jj : t;
            mark[val] = 1;
            row[d] = (uint32_t)val;
        }
        for (int32_t d = 0; d < nsyn; ++d) {
            mark[row[d]] = 0;
            if ((int64_t)row[d] >= i) row[d] += 1u;
        }
        qsort(row, (size_t)nsyn, sizeof(uint32_t), cmp_u32);
        const float g = (i < n_exc) ? g_exc : g_inh;
        for (int32_t d = 0; d < nsyn; ++d) {
            uint32_t o[4];
            draw(seed, (uint32_t)i, row[d], ORC_TAG_WEIGHT, o);
            w[(size_t)i * (size_t)nsyn + (size_t)d] = g * orc_u01(o[0]);
        }
    }
    free(mark);
    return 0;
}

/* One step of cu:352-416 for neuron i.  The expression order is the
 * reference's left-to-right C order: ((((0.04*v)*v + 5*v) + 140) - u) + I,
 * times 0.5, added to v; twice; then u += a*(b*v - u).  Returns 1 on spike
 * (v >= 30 or v not finite, cu:400) after the reset v=c, u+=d. */
static int izh_step(float *vp, float *up, float a, float b, float c, float d, float I)
{
    float v = *vp, u = *up;
    for (int h = 0; h < 2; ++h) {
        float q = 0.04f * v;
        q = q * v;
        float l = 5.0f * v;
        float s = q + l;
        s = s + 140.0f;
        s = s - u;
        s = s + I;
        v = v + 0.5f * s;
    }
    float bv = b * v;
    u = u + a * (bv - u);
    int spk = 0;
    if (v >= 30.0f || !isfinite(v)) {
        v = c;
        u = u + d;
        spk = 1;
    }
    *vp = v;
    *up = u;
    return spk;
}

int orc_simulate(const orc_net *net, int32_t t0, int32_t steps,
                 float *v, float *u, float *isyn, int32_t *nospks,
                 uint8_t *raster, int32_t *counts, float *lastI)
{
    return orc_simulate_order(net, t0, steps, v, u, isyn, nospks, raster, counts, lastI, 0);
}

int orc_simulate_order(const orc_net *net, int32_t t0, int32_t steps,
                       float *v, float *u, float *isyn, int32_t *nospks,
                       uint8_t *raster, int32_t *counts, float *lastI, int32_t descending)
{
    const int64_t n = net->n;
    uint8_t *spk = (uint8_t *)malloc((size_t)n);
    if (!spk) return -2;
    for (int32_t s = 0; s < steps; ++s) {
        const uint32_t t = (uint32_t)(t0 + s);
        int32_t cnt = 0;
        for (int64_t i = 0; i < n; ++i) {
            float z = orc_thalamic(net->seed, t, (uint32_t)i);
            float g = (i < net->n_exc) ? net->ge : net->gi;
            float I = g * z;                      /* cu:378-383 */
            I = I + isyn[i];                      /* cu:386 */
            isyn[i] = 0.0f;                       /* cu:396 */
            if (lastI && s == steps - 1) lastI[i] = I;
            spk[i] = (uint8_t)izh_step(&v[i], &u[i], net->a[i], net->b[i], net->c[i],
                                       net->d[i], I);
            if (spk[i]) { nospks[i] += 1; cnt += 1; }
        }
        /* Propagation of step t into the input of step t+1, in ascending
         * presynaptic order (cu:323-331 semantics without the atomics), or
         * descending for the reassociation study. */
        for (int64_t k = 0; k < n; ++k) {
            const int64_t i = descending ? n - 1 - k : k;
            if (!spk[i]) continue;
            const uint32_t *row = net->post + (size_t)i * (size_t)net->nsyn;
            const float *wr = net->w + (size_t)i * (size_t)net->nsyn;
            for (int32_t k = 0; k < net->nsyn; ++k) isyn[row[k]] = isyn[row[k]] + wr[k];
        }
        if (raster) memcpy(raster + (size_t)s * (size_t)n, spk, (size_t)n);
        if (counts) counts[s] = cnt;
    }
    free(spk);
    return 0;
}

int orc_simulate_f64_forced(const orc_net *net, int32_t t0, int32_t steps,
                            const uint8_t *forced_raster,
                            const int64_t *sub, int64_t nsub,
                            double *v, double *u, double *v_trace, uint8_t *own_raster)
{
    const int64_t n = net->n;
    /* Input to each listed neuron from the forced presynaptic spikes: build
     * a map neuron -> index in sub for the listed neurons. */
    int64_t *slot = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    double *isyn = (double *)calloc((size_t)nsub, sizeof(double));
    if (!slot || !isyn) { free(slot); free(isyn); return -2; }
    for (int64_t i = 0; i < n; ++i) slot[i] = -1;
    for (int64_t k = 0; k < nsub; ++k) slot[sub[k]] = k;
    for (int32_t s = 0; s < steps; ++s) {
        const uint32_t t = (uint32_t)(t0 + s);
        for (int64_t k = 0; k < nsub; ++k) {
            const int64_t i = sub[k];
            double z = (double)orc_thalamic(net->seed, t, (uint32_t)i);
            double g = (i < net->n_exc) ? (double)net->ge : (double)net->gi;
            double I = g * z + isyn[k];
            isyn[k] = 0.0;
            double vv = v[k], uu = u[k];
            for (int h = 0; h < 2; ++h)
                vv += 0.5 * (0.04 * vv * vv + 5.0 * vv + 140.0 - uu + I);
            uu += (double)net->a[i] * ((double)net->b[i] * vv - uu);
            int sp = 0;
            if (vv >= 30.0 || !isfinite(vv)) { vv = (double)net->c[i]; uu += (double)net->d[i]; sp = 1; }
            v[k] = vv; u[k] = uu;
            if (v_trace) v_trace[(size_t)s * (size_t)nsub + (size_t)k] = vv;
            if (own_raster) own_raster[(size_t)s * (size_t)nsub + (size_t)k] = (uint8_t)sp;
        }
        const uint8_t *fr = forced_raster + (size_t)s * (size_t)n;
        for (int64_t i = 0; i < n; ++i) {
            if (!fr[i]) continue;
            const uint32_t *row = net->post + (size_t)i * (size_t)net->nsyn;
            const float *wr = net->w + (size_t)i * (size_t)net->nsyn;
            for (int32_t k = 0; k < net->nsyn; ++k) {
                int64_t sl = slot[row[k]];
                if (sl >= 0) isyn[sl] += (double)wr[k];
            }
        }
    }
    free(slot);
    free(isyn);
    return 0;
}

/* One fp64 Izhikevich step of listed neuron i (cu:352-407 in double). */
static int orc_f64_step(const orc_net *net, uint32_t t, int64_t i, double in, double *v, double *u)
{
    double z = (double)orc_thalamic(net->seed, t, (uint32_t)i);
    double g = (i < net->n_exc) ? (double)net->ge : (double)net->gi;
    double I = g * z + in;
    double vv = *v, uu = *u;
    for (int h = 0; h < 2; ++h)
        vv += 0.5 * (0.04 * vv * vv + 5.0 * vv + 140.0 - uu + I);
    uu += (double)net->a[i] * ((double)net->b[i] * vv - uu);
    int sp = 0;
    if (vv >= 30.0 || !isfinite(vv)) { vv = (double)net->c[i]; uu += (double)net->d[i]; sp = 1; }
    *v = vv;
    *u = uu;
    return sp;
}

/* first index of the ascending row with target >= key */
static int32_t orc_lower_bound(const uint32_t *row, int32_t len, uint32_t key)
{
    int32_t a = 0, b = len;
    while (a < b) {
        const int32_t m = (a + b) / 2;
        if (row[m] < key) a = m + 1;
        else b = m;
    }
    return a;
}

int orc_simulate_f64_forced_ranges(const orc_net *net, int32_t t0, int32_t steps,
                                   const uint32_t *forced_words, int64_t words_per_row,
                                   const int64_t *lo, const int64_t *hi, int32_t nranges,
                                   double *v, double *u, double *v_trace, uint8_t *own_raster)
{
    const int64_t n = net->n;
    int64_t nsub = 0;
    int64_t base[64];
    if (nranges < 1 || nranges > 64) return -1;
    for (int32_t r = 0; r < nranges; ++r) {
        if (lo[r] < 0 || hi[r] > n || lo[r] >= hi[r] || (r && lo[r] < hi[r - 1])) return -1;
        base[r] = nsub;
        nsub += hi[r] - lo[r];
    }
    double *isyn = (double *)calloc((size_t)nsub, sizeof(double));
    if (!isyn) return -2;
    for (int32_t s = 0; s < steps; ++s) {
        const uint32_t t = (uint32_t)(t0 + s);
        for (int32_t r = 0; r < nranges; ++r)
            for (int64_t i = lo[r]; i < hi[r]; ++i) {
                const int64_t k = base[r] + (i - lo[r]);
                const int sp = orc_f64_step(net, t, i, isyn[k], &v[k], &u[k]);
                isyn[k] = 0.0;
                if (v_trace) v_trace[(size_t)s * (size_t)nsub + (size_t)k] = v[k];
                if (own_raster) own_raster[(size_t)s * (size_t)nsub + (size_t)k] = (uint8_t)sp;
            }
        const uint32_t *fr = forced_words + (size_t)s * (size_t)words_per_row;
        for (int64_t w = 0; w < words_per_row; ++w) {
            uint32_t bits = fr[w];
            while (bits) {
                const int64_t i = w * 32 + __builtin_ctz(bits);
                bits &= bits - 1u;
                if (i >= n) break;
                const uint32_t *row = net->post + (size_t)i * (size_t)net->nsyn;
                const float *wr = net->w + (size_t)i * (size_t)net->nsyn;
                for (int32_t r = 0; r < nranges; ++r) {
                    int32_t e = orc_lower_bound(row, net->nsyn, (uint32_t)lo[r]);
                    for (; e < net->nsyn && (int64_t)row[e] < hi[r]; ++e)
                        isyn[base[r] + ((int64_t)row[e] - lo[r])] += (double)wr[e];
                }
            }
        }
    }
    free(isyn);
    return 0;
}

/* ------------------------------------------------------------------------
 * Multithreaded orc_simulate (all-cores CPU baseline).
 * ---------------------------------------------------------------------- */
typedef struct {
    const orc_net *net;
    int32_t t0, steps, nthreads, k;
    float *v, *u, *isyn, *lastI;
    int32_t *nospks, *counts, *part;
    uint8_t *spk, *raster;
    int64_t *list;   /* spiking ids of the step, ascending */
    int64_t *nlist;
    pthread_barrier_t *bar;
} orc_mt_arg;

static void *orc_mt_worker(void *p)
{
    orc_mt_arg *a = (orc_mt_arg *)p;
    const orc_net *net = a->net;
    const int64_t n = net->n;
    const int64_t per = (n + a->nthreads - 1) / a->nthreads;
    const int64_t lo = (int64_t)a->k * per, hi = lo + per < n ? lo + per : n;
    for (int32_t s = 0; s < a->steps; ++s) {
        const uint32_t t = (uint32_t)(a->t0 + s);
        int32_t cnt = 0;
        for (int64_t i = lo; i < hi; ++i) {
            float z = orc_thalamic(net->seed, t, (uint32_t)i);
            float g = (i < net->n_exc) ? net->ge : net->gi;
            float I = g * z;
            I = I + a->isyn[i];
            a->isyn[i] = 0.0f;
            if (a->lastI && s == a->steps - 1) a->lastI[i] = I;
            a->spk[i] = (uint8_t)izh_step(&a->v[i], &a->u[i], net->a[i], net->b[i], net->c[i], net->d[i], I);
            if (a->spk[i]) { a->nospks[i] += 1; cnt += 1; }
        }
        a->part[a->k] = cnt;
        pthread_barrier_wait(a->bar);
        if (a->k == 0) {  /* ascending spike list and the step's outputs */
            int64_t m = 0;
            int32_t total = 0;
            for (int64_t i = 0; i < n; ++i)
                if (a->spk[i]) a->list[m++] = i;
            for (int32_t q = 0; q < a->nthreads; ++q) total += a->part[q];
            *a->nlist = m;
            if (a->raster) memcpy(a->raster + (size_t)s * (size_t)n, a->spk, (size_t)n);
            if (a->counts) a->counts[s] = total;
        }
        pthread_barrier_wait(a->bar);
        /* rows that spiked, ascending, into this thread's targets */
        const int64_t m = *a->nlist;
        for (int64_t q = 0; q < m; ++q) {
            const int64_t i = a->list[q];
            const uint32_t *row = net->post + (size_t)i * (size_t)net->nsyn;
            const float *wr = net->w + (size_t)i * (size_t)net->nsyn;
            int32_t e = orc_lower_bound(row, net->nsyn, (uint32_t)lo);
            for (; e < net->nsyn && (int64_t)row[e] < hi; ++e) a->isyn[row[e]] = a->isyn[row[e]] + wr[e];
        }
        pthread_barrier_wait(a->bar);
    }
    return NULL;
}

int orc_simulate_mt(const orc_net *net, int32_t t0, int32_t steps,
                    float *v, float *u, float *isyn, int32_t *nospks,
                    uint8_t *raster, int32_t *counts, float *lastI, int32_t threads)
{
    if (threads <= 0) threads = (int32_t)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if ((int64_t)threads > net->n) threads = (int32_t)net->n;
    uint8_t *spk = (uint8_t *)malloc((size_t)net->n);
    int64_t *list = (int64_t *)malloc(sizeof(int64_t) * (size_t)net->n);
    int32_t *part = (int32_t *)calloc((size_t)threads, sizeof(int32_t));
    orc_mt_arg *args = (orc_mt_arg *)calloc((size_t)threads, sizeof(orc_mt_arg));
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    int64_t nlist = 0;
    pthread_barrier_t bar;
    int rc = 0;
    if (!spk || !list || !part || !args || !th || pthread_barrier_init(&bar, NULL, (unsigned)threads) != 0) {
        rc = -2;
    } else {
        for (int32_t k = 0; k < threads; ++k) {
            orc_mt_arg a = {net, t0, steps, threads, k, v, u, isyn, lastI, nospks, counts, part, spk, raster,
                            list, &nlist, &bar};
            args[k] = a;
        }
        for (int32_t k = 1; k < threads; ++k) pthread_create(&th[k], NULL, orc_mt_worker, &args[k]);
        orc_mt_worker(&args[0]);
        for (int32_t k = 1; k < threads; ++k) pthread_join(th[k], NULL);
        pthread_barrier_destroy(&bar);
    }
    free(spk); free(list); free(part); free(args); free(th);
    return rc;
}
