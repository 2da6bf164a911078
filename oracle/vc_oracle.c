/*
 * vc_oracle.c -- CPU restatement of the patchV1 simulation step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is compiled
 * into oracle/libvc_oracle.so and called only by tests/, by the smoke check
 * and by the CPU-baseline leg of bench.py.  The product (libvc_hip.so and the
 * Python driver) never links, loads or calls it.
 *
 * It restates, in plain C and binary32 arithmetic with no FP contraction,
 * the reference's equations in the reference's own data layouts:
 *   - LGN precompute     src/discrete_input_convol.cu:8-33 (kernels),
 *                        :230-305 (temporal weights), :307-425 (spatial
 *                        weights, sample coordinates), :746-881 (maxConvol)
 *   - LGN frame response src/discrete_input_convol.cu:35-76, 1003-1069,
 *                        1143-1211 (mean over surround samples, contrast,
 *                        spatial reduction)
 *   - LGN temporal sum,  src/discrete_input_convol.cu:1212-1340,
 *     rate and spikes    :1637-1821, :1587-1635; src/LGN_props.cuh:94-100
 *   - V1 update          src/coredynamics.cu:1235-1629 (non-learning path),
 *                        src/coredynamics.cuh:27-118 (AdEx), src/condShape.h:13-41;
 *                        several spikes per step (tRef < dt):
 *                        src/coredynamics.cu:530-836
 *   - spike compaction   src/coredynamics.cu:1994-2046 (lists in ascending
 *                        order: the reference's atomicAdd_block order is
 *                        unspecified, ascending is one of its outcomes)
 *   - recurrent gather   src/coredynamics.cu:2609-2922, dense conMat
 *                        [nblock][maxNear][bs pre][bs post]; the dense-scan
 *                        form recal_G_mat_nd, :2364-2607
 *   - delayed recurrent  src/coredynamics.cu:2062-2309 (recal_G_mat, axonal
 *     gather             delays from a ring of past step outputs)
 *   - far connections    src/coredynamics.cu:181-372, 2314-2349 (recal_G_vec
 *                        with delays, sum_G)
 *   - plasticity,        src/discrete_input_convol.cu:1736-1821 (the LGN
 *     learning = 3       trace), src/coredynamics.cu:1631-1992 (the FF_E
 *                        triplet rule, homeostasis, clip), pinned by float64
 *                        sums over spike times (tests/test_learn_oracle.py)
 * Reductions follow the reference's block_reduce tree (32-lane shuffle-down
 * halving per warp, then halving over the warp partials,
 * src/util/cuda_util.cuh:22-127).
 * exp/log on the per-step path use include/vc_math.h (the same functions
 * the GPU uses); the one-time precompute uses the C library's expf, logf,
 * lgammaf, cosf, sinf as the reference uses CUDA's.
 *
 * Parity pins (tests/test_oracle.py): closed forms of the conductance ODE,
 * Heun amplification for constant conductances, spike-time interpolation,
 * depC steady state, brute-force tiny networks, logistic identities
 * f(0)=spont f(1)=1, Poisson time rescaling, TW against gamma pdfs, SW sum,
 * uniform-frame zero contrast.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "../include/vc_math.h"
#include "../include/vc_abi.h"

#define ORC_VE 0.0f     /* src/CONST.h:7-9 */
#define ORC_VI (-75.0f)
#define ORC_VL (-65.0f)
#define ORC_SQRT2 1.4142135623730951   /* src/CONST.h:32, a double */

/* ---------------------------------------------------------------------- */
/* reference block_reduce over nThreads (a power of two, >= 32); x[i] for  */
/* i >= n contribute 0.                                                    */
static float orc_block_reduce(const float *x, int n, int nThreads) {
    float part[32];
    int W = nThreads / 32;
    for (int w = 0; w < W; ++w) {
        float a[32];
        for (int l = 0; l < 32; ++l) {
            int i = w * 32 + l;
            a[l] = (i < n) ? x[i] : 0.0f;
        }
        for (int off = 16; off > 0; off >>= 1)
            for (int l = 0; l < off; ++l) a[l] = a[l] + a[l + off];
        part[w] = a[0];
    }
    for (int off = W / 2; off > 0; off >>= 1)
        for (int l = 0; l < off; ++l) part[l] = part[l] + part[l + off];
    return part[0];
}

float orc_block_reduce_export(const float *x, int n, int nThreads) {
    return orc_block_reduce(x, n, nThreads);
}

/* ---------------------------------------------------------------------- */
/* ConductanceShape, src/condShape.h:28-41                                 */
static inline void orc_single_input(float *g, float *h, float f, float ddt,
                                    float rise, float decay, float dod) {
    float etr = vc_expf(-ddt / rise);
    *g = *g + f * dod * (vc_expf(-ddt / decay) - etr);
    *h = *h + f * etr;
}
static inline void orc_decay(float *g, float *h, float ddt, float rise,
                             float decay, float dod) {
    float etr = vc_expf(-ddt / rise);
    float etd = vc_expf(-ddt / decay);
    *g = *g * etd + *h * dod * (etd - etr);
    *h = *h * etr;
}

/* exported for the closed-form tests */
void orc_cond_single(float *g, float *h, float f, float ddt, float rise,
                     float decay) {
    float dod = rise / (decay - rise);
    orc_single_input(g, h, f, ddt, rise, decay, dod);
}
void orc_cond_decay(float *g, float *h, float ddt, float rise, float decay) {
    float dod = rise / (decay - rise);
    orc_decay(g, h, ddt, rise, decay, dod);
}

float orc_expf(float x) { return vc_expf(x); }
float orc_logf(float x) { return vc_logf(x); }
void orc_cos2pi_vec(uint32_t n, const float *x, float *out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = vc_cos2pif(x[i]);
}
void orc_normal_vec(uint32_t n, const uint32_t *a, const uint32_t *b, float *out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = vc_normal(a[i], b[i]);
}
/* The normal of counter (ent, step, stream | pack << 8, trial) under key
 * (klo, khi): the draw the noise and failure terms use. */
float orc_draw_normal(uint32_t klo, uint32_t khi, uint32_t ent, uint32_t step,
                      uint32_t stream, uint32_t pack, uint32_t trial) {
    vc_u32x4 r = vc_draw4(klo, khi, ent, step, stream, pack, trial);
    return vc_normal(r.v[0], r.v[1]);
}
void orc_philox(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
    vc_u32x4 c;
    for (int i = 0; i < 4; ++i) c.v[i] = ctr[i];
    c = vc_philox4x32_10(c, k0, k1);
    for (int i = 0; i < 4; ++i) out[i] = c.v[i];
}

/* ====================================================================== */
/* V1                                                                      */
/* ====================================================================== */
typedef struct {
    float v, v0, a0, b0, a1, b1, vR, vThres, tRef, tBack, tsp, C, gL, depC;
    float w0, w, tau_w, a, b, vT, deltaT;
    int spiked;
} orc_adex;

/* AdEx constructor, src/coredynamics.cuh:40-49.  deltaT == 0 is the LIF
 * extension: the exponential current is absent. */
static void orc_adex_init(orc_adex *m, const vc_v1_params *p, int t, float w0,
                          float v0, float tBack, float dep) {
    m->w0 = w0; m->tau_w = p->tau_w[t]; m->a = p->a[t]; m->b = p->b[t];
    m->v0 = v0; m->tBack = tBack; m->vR = p->vR[t]; m->vThres = p->vThres[t];
    m->gL = p->gL[t]; m->C = p->C[t]; m->tRef = p->tRef[t]; m->vT = p->vT[t];
    m->deltaT = p->deltaT[t];
    m->spiked = 0;
    m->v = v0; m->w = w0; m->tsp = 0.0f;
    float targetV = ORC_VL + (m->vT - ORC_VL) * dep;
    if (targetV > m->vT) {
        m->depC = (m->gL + m->a) * (targetV - ORC_VL) - m->gL * m->deltaT;
    } else if (m->deltaT > 0.0f) {
        m->depC = (m->gL + m->a) * (targetV - ORC_VL) -
                  m->gL * m->deltaT * vc_expf((targetV - m->vT) / m->deltaT);
    } else {
        m->depC = (m->gL + m->a) * (targetV - ORC_VL) - 0.0f;
    }
}

/* src/coredynamics.cuh:75-87: noise0 joins the first slope, noise1 the
 * second */
static void orc_rk2(orc_adex *m, float dt, float noise0, float noise1) {
    float dTgL = m->deltaT * m->gL;
    float fk1, fk2;
    if (m->deltaT > 0.0f)
        fk1 = (-m->a0 * m->v0 + m->b0 + dTgL * vc_expf((m->v0 - m->vT) / m->deltaT) - m->w0) / m->C;
    else
        fk1 = (-m->a0 * m->v0 + m->b0 - m->w0) / m->C;
    fk1 += noise0;
    float gk1 = (m->a * (m->v0 - ORC_VL) - m->w0) / m->tau_w;
    float v1 = m->v0 + dt * fk1;
    float w1 = m->w0 + dt * gk1;
    if (m->deltaT > 0.0f)
        fk2 = (-m->a1 * v1 + m->b1 + dTgL * vc_expf((v1 - m->vT) / m->deltaT) - w1) / m->C;
    else
        fk2 = (-m->a1 * v1 + m->b1 - w1) / m->C;
    fk2 += noise1;
    float gk2 = (m->a * (v1 - ORC_VL) - w1) / m->tau_w;
    m->w = m->w0 + dt * (gk1 + gk2) / 2.0f;
    m->v = m->v0 + dt * (fk1 + fk2) / 2.0f;
}

static inline int orc_type_of(const vc_v1_params *p, uint32_t local) {
    for (uint32_t j = 0; j < p->nType; ++j)
        if (local < p->typeAcc[j]) return (int)j;
    return (int)p->nType - 1;
}

/* The Philox key of neuron n's draws: the run's seed and n, or with batched
 * trials its trial's seed and its id within the trial (off = the trial's
 * first neuron). */
typedef struct { uint32_t lo, hi, ent, off; } orc_key;
static inline orc_key orc_key_of(const vc_v1_params *p, uint32_t n) {
    orc_key k = {p->seed_lo, p->seed_hi, n, 0u};
    if (p->blockTrial) {
        uint32_t tr = p->blockTrial[n / p->blockSize];
        k.lo = p->trialSeed[2 * tr];
        k.hi = p->trialSeed[2 * tr + 1];
        k.off = p->trialV1Off[tr];
        k.ent = n - k.off;
    }
    return k;
}

/* Synaptic failure (src/coredynamics.cu:1476-1480 for LGN synapses,
 * :2766-2769 for V1 ones): strength f times 1 - p + sqrt(p (1 - p) / nSyn)
 * N(0, 1); the normal is drawn for (post neuron, step, stream, pack). */
static inline float orc_fail(float f, float pf, float nSyn, orc_key k, uint32_t step,
                             uint32_t stream, uint32_t pack, uint32_t trial) {
    float sd = sqrtf(pf * (1.0f - pf) / nSyn);
    vc_u32x4 r = vc_draw4(k.lo, k.hi, k.ent, step, stream, pack, trial);
    float rnd = vc_normal(r.v[0], r.v[1]);
    return f * (1.0f - pf + sd * rnd);
}

/* Background drive (extension): one Philox draw per neuron-step gives the
 * Poisson count (inversion) and the event time. */
static inline uint32_t orc_bg_draw(const vc_v1_params *p, int t, uint32_t n,
                                   uint32_t step, float *frac) {
    orc_key k = orc_key_of(p, n);
    vc_u32x4 r = vc_draw4(k.lo, k.hi, k.ent, step, VC_STREAM_V1_BG, 0, p->trial);
    float u = vc_u01_open0(r.v[0]);
    *frac = vc_u01_open1(r.v[1]);
    return vc_poisson_inv(u, p->bgLambda[t], p->bgP0[t]);
}

/* One V1 neuron, src/coredynamics.cu:1296-1629.  Returns sInfo. */
static float orc_update_neuron(const vc_v1_params *p, uint32_t n, float *v,
                               float *w, float *tBack, float *gFF, float *hFF,
                               float *gE, float *hE, float *gI, float *hI,
                               const float *tonicDep, const uint32_t *nLGN,
                               const float *sLGN, const uint32_t *idLGN,
                               const float *lgnSInfo, uint32_t step,
                               float *out_dep, const float *gapS, const float *gap,
                               float *lastNoise) {
    const uint32_t nV1 = p->nV1;
    const uint32_t local = n % p->blockSize;
    const int t = orc_type_of(p, local);
    const float dt = p->dt;
    orc_adex m;
    orc_adex_init(&m, p, t, w[n], v[n], tBack[n], tonicDep[n]);

    float ge[VC_MAX_NG], he[VC_MAX_NG], gi[VC_MAX_NG], hi[VC_MAX_NG];
    float gE_t1 = 0.0f, gI_t1 = 0.0f;
    for (uint32_t ig = 0; ig < p->ngE; ++ig) {
        float g = gE[ig * nV1 + n], h = hE[ig * nV1 + n];
        ge[ig] = g; he[ig] = h;
        orc_decay(&g, &h, dt, p->riseE[ig], p->decayE[ig], p->dodE[ig]);
        gE[ig * nV1 + n] = g; hE[ig * nV1 + n] = h;
        gE_t1 += g;
    }
    for (uint32_t ig = 0; ig < p->ngI; ++ig) {
        float g = gI[ig * nV1 + n], h = hI[ig * nV1 + n];
        gi[ig] = g; hi[ig] = h;
        orc_decay(&g, &h, dt, p->riseI[ig], p->decayI[ig], p->dodI[ig]);
        gI[ig * nV1 + n] = g; hI[ig * nV1 + n] = h;
        gI_t1 += g;
    }
    float g0[VC_MAX_NG], h0[VC_MAX_NG], g1[VC_MAX_NG], h1[VC_MAX_NG];
    for (uint32_t ig = 0; ig < p->ngFF; ++ig) {
        g1[ig] = gFF[ig * nV1 + n]; h1[ig] = hFF[ig * nV1 + n];
        g0[ig] = g1[ig]; h0[ig] = h1[ig];
        orc_decay(&g1[ig], &h1[ig], dt, p->riseFF[ig], p->decayFF[ig], p->dodFF[ig]);
    }
    /* voltage noise, src/coredynamics.cu:1424-1445: white (tauNoise = 0)
     * enters both slopes as noise1/dt, Ornstein-Uhlenbeck (tauNoise > 0) as
     * the previous value (noise0) and the new one (noise1) */
    const orc_key key = orc_key_of(p, n);
    const float tau = p->tauNoise;
    float noise0 = 0.0f;
    if (tau > 0.0f) noise0 = lastNoise ? lastNoise[n] : 0.0f;
    float ratio = p->noiseOnTonic ? tonicDep[n] : 1.0f;
    float noise1 = sqrtf(2.0f * p->noisyDep[t] * (p->vT[t] - p->vR[t]) * ratio * dt);
    if (noise1 > 0.0f) {
        vc_u32x4 r = vc_draw4(key.lo, key.hi, key.ent, step, VC_STREAM_V1_NOISE, 0, p->trial);
        noise1 *= vc_normal(r.v[0], r.v[1]);
        if (tau > 0.0f) {
            noise1 /= tau;
            float hk1 = -noise0 / tau;
            float hk2 = -(noise0 + dt * hk1 + noise1) / tau;
            noise1 += noise0 + dt * (hk1 + hk2) / 2.0f;
            if (lastNoise) lastNoise[n] = noise1;
        }
    }
    const int backing = (m.tBack < dt) && (m.tBack > 0.0f);
    if (out_dep) out_dep[n] = m.depC + noise1;
    if (backing)
        for (uint32_t ig = 0; ig < p->ngFF; ++ig)
            orc_decay(&g0[ig], &h0[ig], m.tBack, p->riseFF[ig], p->decayFF[ig], p->dodFF[ig]);

    const uint32_t mLGN = nLGN[n];
    for (uint32_t i = 0; i < mLGN; ++i) {
        uint32_t lid = i * nV1 + n;
        float f = sLGN[lid];
        float s = lgnSInfo[idLGN[lid]];
        uint32_t nsp = (uint32_t)floorf(s);
        if (nsp > 0) {
            float tsp = (s - (float)nsp) * dt;
            if (p->synFailFF[t] > 0.0f)
                f = orc_fail(f, p->synFailFF[t], p->synPerConFF[t], key, step,
                             VC_STREAM_V1_FAILFF, i, p->trial);
            if (f > 0.0f) {
                float ddt = 0.0f;
                if (backing) ddt = m.tBack - tsp;
                for (uint32_t ig = 0; ig < p->ngFF; ++ig) {
                    float str = f * p->pFF[t * p->ngFF + ig];
                    if (backing && ddt > 0.0f)
                        orc_single_input(&g0[ig], &h0[ig], str * (float)nsp, ddt,
                                         p->riseFF[ig], p->decayFF[ig], p->dodFF[ig]);
                    orc_single_input(&g1[ig], &h1[ig], str * (float)nsp, dt - tsp,
                                     p->riseFF[ig], p->decayFF[ig], p->dodFF[ig]);
                }
            }
        }
    }
    if (p->bgStrength[t] > 0.0f) {
        float frac;
        uint32_t nsp = orc_bg_draw(p, t, n, step, &frac);
        if (nsp > 0) {
            float tsp = frac * dt;
            float f = p->bgStrength[t];
            float ddt = 0.0f;
            if (backing) ddt = m.tBack - tsp;
            for (uint32_t ig = 0; ig < p->ngFF; ++ig) {
                float str = f * p->pBG[t * p->ngFF + ig];
                if (backing && ddt > 0.0f)
                    orc_single_input(&g0[ig], &h0[ig], str * (float)nsp, ddt,
                                     p->riseFF[ig], p->decayFF[ig], p->dodFF[ig]);
                orc_single_input(&g1[ig], &h1[ig], str * (float)nsp, dt - tsp,
                                 p->riseFF[ig], p->decayFF[ig], p->dodFF[ig]);
            }
        }
    }
    float gE_t0 = 0.0f;
    for (uint32_t ig = 0; ig < p->ngFF; ++ig) {
        gE_t0 += g0[ig];
        gE_t1 += g1[ig];
        gFF[ig * nV1 + n] = g1[ig];
        hFF[ig * nV1 + n] = h1[ig];
    }
    for (uint32_t ig = 0; ig < p->ngE; ++ig) {
        if (backing) orc_decay(&ge[ig], &he[ig], m.tBack, p->riseE[ig], p->decayE[ig], p->dodE[ig]);
        gE_t0 += ge[ig];
    }
    float gI_t0 = 0.0f;
    for (uint32_t ig = 0; ig < p->ngI; ++ig) {
        if (backing) orc_decay(&gi[ig], &hi[ig], m.tBack, p->riseI[ig], p->decayI[ig], p->dodI[ig]);
        gI_t0 += gi[ig];
    }

    /* gap junctions of inhibitory neurons: gapS joins a, gap joins b
     * (src/coredynamics.cuh:107-117; compute_V_collect_spike_learnFF_fast
     * takes gapS for every inhibitory neuron, src/coredynamics.cu:1344) */
    float lgapS = 0.0f, lgap = 0.0f;
    if (local >= p->nE) {
        uint32_t gi = (n / p->blockSize) * (p->blockSize - p->nE) + (local - p->nE);
        if (gapS) lgapS = gapS[gi];
        if (gap) lgap = gap[gi];
    }
    m.tsp = 0.0f;
    if (m.tBack < dt) {
        m.a0 = gE_t0 + gI_t0 + m.gL + lgapS;
        m.b0 = gE_t0 * ORC_VE + gI_t0 * ORC_VI + m.gL * ORC_VL + m.depC + lgap;
        m.a1 = gE_t1 + gI_t1 + m.gL + lgapS;
        m.b1 = gE_t1 * ORC_VE + gI_t1 * ORC_VI + m.gL * ORC_VL + m.depC + lgap;
        float new_dt = dt - m.tBack;
        if (backing) {
            float A = m.a * (m.v0 - ORC_VL);
            m.w0 = (m.w0 - A) * vc_expf(-m.tBack / m.tau_w) + A;
        }
        if (tau > 0.0f)
            orc_rk2(&m, new_dt, noise0, noise1);
        else
            orc_rk2(&m, new_dt, noise1 / dt, noise1 / dt);
        if (m.v > m.vThres) {
            float ratio = (m.vThres - m.v0) / (m.v - m.v0);
            m.tsp = m.tBack + ratio * new_dt;
            m.w0 += ratio * (m.w - m.w0) + m.b;
            m.spiked = 1;
            m.tBack = m.tsp + m.tRef;
            m.v = m.vR;
        } else {
            if (m.tBack > 0.0f) m.tBack = 0.0f;
        }
    }
    if (m.tBack >= dt) {
        m.v = m.vR;
        float A = m.a * (m.v - ORC_VL);
        m.w = (m.w0 - A) * vc_expf(-(dt - m.tsp) / m.tau_w) + A;
        m.tBack -= dt;
    }
    float sInfo = m.spiked ? 1.0f + m.tsp / dt : m.v;
    v[n] = m.v;
    if (p->iModel == 1) w[n] = m.w;
    tBack[n] = m.tBack;
    return sInfo;
}

/* One V1 neuron that may fire several times in a step,
 * compute_V_collect_spike_learnFF, src/coredynamics.cu:530-836 (the
 * reference runs it for every neuron when dt > min tRef, :6297).  The state
 * at the start of the step goes into (g0, ge, gi); g1 collects every input
 * of the step once (first pass).  The do-while re-integrates the rest of the
 * step from vR after each spike whose refractory period ends inside it:
 * g0/ge/gi decay from the previous restart new_t0 to tBack, and the LGN
 * inputs (list order, then the background one) that arrived in [new_t0,
 * tBack) are added at tBack.  A synapse's failure factor is drawn in the
 * first pass and kept (f[i]).  White noise: `noise1 /= dt` on every pass
 * (:720-723).  sInfo = nsp + mean(tsp)/dt (:815-821). */
static float orc_update_neuron_multi(const vc_v1_params *p, uint32_t n, float *v,
                                     float *w, float *tBack, float *gFF, float *hFF,
                                     float *gE, float *hE, float *gI, float *hI,
                                     const float *tonicDep, const uint32_t *nLGN,
                                     const float *sLGN, const uint32_t *idLGN,
                                     const float *lgnSInfo, uint32_t step,
                                     float *out_dep, const float *gapS, const float *gap,
                                     float *lastNoise) {
    const uint32_t nV1 = p->nV1;
    const uint32_t local = n % p->blockSize;
    const int t = orc_type_of(p, local);
    const float dt = p->dt;
    orc_adex m;
    orc_adex_init(&m, p, t, w[n], v[n], tBack[n], tonicDep[n]);
    uint32_t spikeCount = 0;

    /* :549-580: cortical g decayed to the end of the step; ge/gi keep t = 0 */
    float ge[VC_MAX_NG], he[VC_MAX_NG], gi[VC_MAX_NG], hi[VC_MAX_NG];
    float gE_t1 = 0.0f, gI_t1 = 0.0f;
    for (uint32_t ig = 0; ig < p->ngE; ++ig) {
        float g = gE[ig * nV1 + n], h = hE[ig * nV1 + n];
        ge[ig] = g; he[ig] = h;
        orc_decay(&g, &h, dt, p->riseE[ig], p->decayE[ig], p->dodE[ig]);
        gE[ig * nV1 + n] = g; hE[ig * nV1 + n] = h;
        gE_t1 += g;
    }
    for (uint32_t ig = 0; ig < p->ngI; ++ig) {
        float g = gI[ig * nV1 + n], h = hI[ig * nV1 + n];
        gi[ig] = g; hi[ig] = h;
        orc_decay(&g, &h, dt, p->riseI[ig], p->decayI[ig], p->dodI[ig]);
        gI[ig * nV1 + n] = g; hI[ig * nV1 + n] = h;
        gI_t1 += g;
    }
    /* :582-596 */
    float g0[VC_MAX_NG], h0[VC_MAX_NG], g1[VC_MAX_NG], h1[VC_MAX_NG];
    for (uint32_t ig = 0; ig < p->ngFF; ++ig) {
        g1[ig] = gFF[ig * nV1 + n]; h1[ig] = hFF[ig * nV1 + n];
        g0[ig] = g1[ig]; h0[ig] = h1[ig];
        orc_decay(&g1[ig], &h1[ig], dt, p->riseFF[ig], p->decayFF[ig], p->dodFF[ig]);
    }
    /* :606-622 noise, as in the one-spike update */
    const orc_key key = orc_key_of(p, n);
    const float tau = p->tauNoise;
    float noise0 = 0.0f;
    if (tau > 0.0f) noise0 = lastNoise ? lastNoise[n] : 0.0f;
    float ratio0 = p->noiseOnTonic ? tonicDep[n] : 1.0f;
    float noise1 = sqrtf(2.0f * p->noisyDep[t] * (p->vT[t] - p->vR[t]) * ratio0 * dt);
    if (noise1 > 0.0f) {
        vc_u32x4 r = vc_draw4(key.lo, key.hi, key.ent, step, VC_STREAM_V1_NOISE, 0, p->trial);
        noise1 *= vc_normal(r.v[0], r.v[1]);
        if (tau > 0.0f) {
            noise1 /= tau;
            float hk1 = -noise0 / tau;
            float hk2 = -(noise0 + dt * hk1 + noise1) / tau;
            noise1 += noise0 + dt * (hk1 + hk2) / 2.0f;
            if (lastNoise) lastNoise[n] = noise1;
        }
    }
    int backingUpFromRef = (m.tBack < dt) && (m.tBack > 0.0f);
    float new_t0 = 0.0f, sInfo = 0.0f;
    if (out_dep) out_dep[n] = m.depC + noise1;
    const uint32_t mLGN = nLGN[n];
    uint32_t bgNsp = 0;
    float bgTsp = 0.0f;
    if (p->bgStrength[t] > 0.0f) {
        float frac;
        bgNsp = orc_bg_draw(p, t, n, step, &frac);
        bgTsp = frac * dt;
    }
    float lgapS = 0.0f, lgap = 0.0f;
    if (local >= p->nE) {
        uint32_t gid = (n / p->blockSize) * (p->blockSize - p->nE) + (local - p->nE);
        if (gapS) lgapS = gapS[gid];
        if (gap) lgap = gap[gid];
    }
    do {
        float dtBack = 0.0f;
        if (backingUpFromRef) dtBack = m.tBack - new_t0;
        /* :645-651 */
        for (uint32_t ig = 0; ig < p->ngFF; ++ig)
            if (backingUpFromRef)
                orc_decay(&g0[ig], &h0[ig], dtBack, p->riseFF[ig], p->decayFF[ig], p->dodFF[ig]);
        /* :653-693 the LGN inputs, then the background one (extension) */
        for (uint32_t i = 0; i < mLGN; ++i) {
            uint32_t lid = i * nV1 + n;
            float s = lgnSInfo[idLGN[lid]];
            uint32_t nsp = (uint32_t)floorf(s);
            float tsp = (s - (float)nsp) * dt;
            if (nsp > 0 && tsp >= new_t0) {
                /* f[i]: the reference draws the failure factor in the first
                 * pass and keeps it; the draw depends on (neuron, step, i)
                 * only, so drawing it again in a later pass gives it back */
                float fi = sLGN[lid];
                if (p->synFailFF[t] > 0.0f)
                    fi = orc_fail(fi, p->synFailFF[t], p->synPerConFF[t], key, step,
                                  VC_STREAM_V1_FAILFF, i, p->trial);
                if (fi > 0.0f) {
                    float ddt = 0.0f;
                    if (backingUpFromRef) ddt = m.tBack - tsp;
                    for (uint32_t ig = 0; ig < p->ngFF; ++ig) {
                        float str = fi * p->pFF[t * p->ngFF + ig];
                        if (backingUpFromRef && ddt > 0.0f)
                            orc_single_input(&g0[ig], &h0[ig], str * (float)nsp, ddt,
                                             p->riseFF[ig], p->decayFF[ig], p->dodFF[ig]);
                        if (spikeCount == 0)
                            orc_single_input(&g1[ig], &h1[ig], str * (float)nsp, dt - tsp,
                                             p->riseFF[ig], p->decayFF[ig], p->dodFF[ig]);
                    }
                }
            }
        }
        if (bgNsp > 0 && bgTsp >= new_t0) {
            float f = p->bgStrength[t];
            float ddt = 0.0f;
            if (backingUpFromRef) ddt = m.tBack - bgTsp;
            for (uint32_t ig = 0; ig < p->ngFF; ++ig) {
                float str = f * p->pBG[t * p->ngFF + ig];
                if (backingUpFromRef && ddt > 0.0f)
                    orc_single_input(&g0[ig], &h0[ig], str * (float)bgNsp, ddt,
                                     p->riseFF[ig], p->decayFF[ig], p->dodFF[ig]);
                if (spikeCount == 0)
                    orc_single_input(&g1[ig], &h1[ig], str * (float)bgNsp, dt - bgTsp,
                                     p->riseFF[ig], p->decayFF[ig], p->dodFF[ig]);
            }
        }
        /* :697-727 collect */
        float gE_t0 = 0.0f;
        for (uint32_t ig = 0; ig < p->ngFF; ++ig) {
            gE_t0 += g0[ig];
            if (spikeCount == 0) {
                gE_t1 += g1[ig];
                gFF[ig * nV1 + n] = g1[ig];
                hFF[ig * nV1 + n] = h1[ig];
            }
        }
        for (uint32_t ig = 0; ig < p->ngE; ++ig) {
            if (backingUpFromRef)
                orc_decay(&ge[ig], &he[ig], dtBack, p->riseE[ig], p->decayE[ig], p->dodE[ig]);
            gE_t0 += ge[ig];
        }
        float gI_t0 = 0.0f;
        for (uint32_t ig = 0; ig < p->ngI; ++ig) {
            if (backingUpFromRef)
                orc_decay(&gi[ig], &hi[ig], dtBack, p->riseI[ig], p->decayI[ig], p->dodI[ig]);
            gI_t0 += gi[ig];
        }
        /* :730-800 stepping */
        m.tsp = 0.0f;
        if (m.tBack < dt) {
            m.a0 = gE_t0 + gI_t0 + m.gL + lgapS;
            m.b0 = gE_t0 * ORC_VE + gI_t0 * ORC_VI + m.gL * ORC_VL + m.depC + lgap;
            if (spikeCount == 0) {
                m.a1 = gE_t1 + gI_t1 + m.gL + lgapS;
                m.b1 = gE_t1 * ORC_VE + gI_t1 * ORC_VI + m.gL * ORC_VL + m.depC + lgap;
            }
            float new_dt = dt - m.tBack;
            if (backingUpFromRef) {                      /* rk2_vFixedBefore(dtBack) */
                float A = m.a * (m.v0 - ORC_VL);
                m.w0 = (m.w0 - A) * vc_expf(-dtBack / m.tau_w) + A;
            }
            if (tau > 0.0f) {
                orc_rk2(&m, new_dt, noise0, noise1);
            } else {
                noise1 /= dt;
                orc_rk2(&m, new_dt, noise1, noise1);
            }
            if (m.v > m.vThres) {
                new_t0 = m.tBack;                        /* compute_spike_time */
                float ratio = (m.vThres - m.v0) / (m.v - m.v0);
                m.tsp = new_t0 + ratio * new_dt;
                m.w0 += ratio * (m.w - m.w0) + m.b;
                sInfo += m.tsp;
                spikeCount++;
                m.tBack = m.tsp + m.tRef;
                backingUpFromRef = m.tBack < dt;
                if (backingUpFromRef) m.v0 = m.vR;       /* reset(true) */
                else m.v = m.vR;                          /* reset(false) */
            } else {
                if (m.tBack > 0.0f) m.tBack = 0.0f;
                backingUpFromRef = 0;
            }
        }
        if (m.tBack >= dt) {                             /* rk2_vFixedAfter */
            m.v = m.vR;
            float A = m.a * (m.v - ORC_VL);
            m.w = (m.w0 - A) * vc_expf(-(dt - m.tsp) / m.tau_w) + A;
            m.tBack -= dt;
        }
    } while (backingUpFromRef);
    if (spikeCount > 0) {
        sInfo /= (float)spikeCount * dt;
        sInfo += (float)spikeCount;
    } else {
        sInfo = m.v;
    }
    v[n] = m.v;
    if (p->iModel == 1) w[n] = m.w;
    tBack[n] = m.tBack;
    return sInfo;
}

/* 1: every neuron takes the multi-spike update, as the reference does for
 * the whole network when dt > min tRef; 0: only types with tRef < dt do
 * (the two give the same results, which tests/ check). */
static int orc_multi_all = 0;
void orc_set_multi_all(int on) { orc_multi_all = on; }

/*
 * rand_spInit, src/coredynamics.cu:5-146 (initial spikes and the spike-train
 * ring).  Per neuron: type t, chance = sp0[t*2 + 0] when it has more than
 * SCsplit LGN inputs, else sp0[t*2 + 1]; rand < chance spikes at tsp (ring
 * slot 0 = 1 + tsp; v = vR; tBack = tRef - (1 - tsp) dt if positive; AdEx w
 * advanced as :75-79 writes it, A = a (vR - vL) tau_w); then reference
 * slots i = depthRef-1 down to 1: outside the refractory period (ref < dt)
 * a spike with chance * (dt - ref)/dt at (dt - ref)/dt of a uniform, else
 * the neuron's v; inside it vR; ref -= dt after each slot.  Uniforms from
 * Philox counter (neuron, reference slot, VC_STREAM_V1_INIT, trial), word 0
 * for the test, word 1 for the time, in (0, 1] as curand_uniform.
 * Reference slot j goes to ring slot (cur0 + depth - depthRef + j) % depth
 * of train [depth][nV1]; ring slots of steps before -depthRef get v.
 */
int orc_v1_initial_spikes(const vc_v1_params *p, const uint32_t *nLGN, const vc_v1_spinit *q,
                          float *v, float *w, float *tBack, float *train, uint32_t depth,
                          uint32_t cur0) {
    const uint32_t nV1 = p->nV1, D = q->depthRef;
    if (D == 0 || depth < D || cur0 >= depth) return VC_E_INVALID;
    for (uint32_t n = 0; n < nV1; ++n) {
        const orc_key key = orc_key_of(p, n);
        const int t = orc_type_of(p, n % p->blockSize);
        const uint32_t iLGN = nLGN ? nLGN[n] : 0;
        const float chance = iLGN > q->SCsplit ? q->sp0[t * 2 + 0] : q->sp0[t * 2 + 1];
        const float dt = p->dt, tRef = p->tRef[t];
        vc_u32x4 r = vc_draw4(key.lo, key.hi, key.ent, 0, VC_STREAM_V1_INIT, 0, p->trial);
        float rnd = vc_u01_open0(r.v[0]);
        float ref = 0.0f;
        float sInfo = v[n];
        if (chance > 0.0f) {
            if (rnd < chance) {
                float tsp = vc_u01_open0(r.v[1]);
                sInfo = 1.0f + tsp;
                float tb = tRef - (1.0f - tsp) * dt;
                if (tb > 0.0f) tBack[n] = tb;
                float v0 = p->vR[t];
                v[n] = v0;
                if (p->iModel == 1) {
                    float A = p->a[t] * (v0 - ORC_VL) * p->tau_w[t];
                    float w0 = w[n] + p->b[t];
                    w[n] = (w0 - A) * vc_expf(-dt * (1.0f - tsp) / p->tau_w[t]) + A;
                }
                ref = (1.0f - tsp) * dt + tRef - dt;
            }
        }
        train[(size_t)((cur0 + depth - D) % depth) * nV1 + n] = sInfo;
        for (uint32_t i = D - 1; i > 0; --i) {
            float val;
            if (ref < dt) {
                if (ref < 0.0f) ref = 0.0f;
                r = vc_draw4(key.lo, key.hi, key.ent, i, VC_STREAM_V1_INIT, 0, p->trial);
                if (vc_u01_open0(r.v[0]) < chance * (dt - ref) / dt) {
                    float tsp = vc_u01_open0(r.v[1]) * (dt - ref) / dt;
                    val = 1.0f + tsp;
                    ref = tRef + (1.0f - tsp) * dt;
                } else {
                    val = v[n];
                }
            } else {
                val = p->vR[t];
            }
            ref -= dt;
            train[(size_t)((cur0 + depth - D + i) % depth) * nV1 + n] = val;
        }
        for (uint32_t k = D; k < depth; ++k)
            train[(size_t)((cur0 + depth - 1 - k) % depth) * nV1 + n] = v[n];
    }
    return VC_OK;
}

/*
 * recal_G_mat_nd_fast (src/coredynamics.cu:2609-2922, the gather at
 * :2721-2916): every post neuron sums, over its neighbour slots ib and the
 * E list then the I list of block nbr[b][ib] (ipre/npre, og/oh in the
 * reference's compacted layout, E list first), strength * p[type][ig] *
 * og/oh.  `step` is the step the spikes were fired in (the failure draws'
 * key).
 */
int orc_v1_rec_fast(const vc_v1_params *p, float *gE, float *hE, float *gI, float *hI,
                    uint32_t maxNear, const uint32_t *nNear, const uint32_t *nbr,
                    const float *conMat, const uint32_t *ipre, const uint32_t *npre,
                    const float *og, const float *oh, uint32_t step) {
    const uint32_t nV1 = p->nV1, bs = p->blockSize, nblock = p->nblock;
    const uint32_t nE = p->nE, nI = bs - nE;
    const uint32_t block_ng = p->ngE * nE + p->ngI * nI;
#pragma omp parallel for schedule(static)
    for (uint32_t n = 0; n < nV1; ++n) {
        uint32_t b = n / bs, l = n % bs;
        int t = orc_type_of(p, l);
        float lgE[VC_MAX_NG] = {0}, lhE[VC_MAX_NG] = {0};
        float lgI[VC_MAX_NG] = {0}, lhI[VC_MAX_NG] = {0};
        /* synaptic failure by presynaptic type (E: the first, I: the
         * last) and this neuron's type, drawn per (this neuron, step,
         * presynaptic neuron) */
        const orc_key key = orc_key_of(p, n);
        const float pfE = p->synFail[0 * p->nType + t];
        const float pfI = p->synFail[(p->nType - 1) * p->nType + t];
        const float nsE = p->synPerCon[0 * p->nType + t];
        const float nsI = p->synPerCon[(p->nType - 1) * p->nType + t];
        for (uint32_t ib = 0; ib < nNear[b]; ++ib) {
            uint32_t bid = nbr[b * maxNear + ib];
            uint32_t cE = npre[bid], cI = npre[nblock + bid];
            uint64_t local_bid = (uint64_t)b * maxNear + ib;
            for (uint32_t i = 0; i < cE; ++i) {
                uint32_t pre = ipre[bid * bs + i];
                float strength = conMat[(local_bid * bs + pre) * bs + l];
                if (strength > 0.0f && pfE > 0.0f)
                    strength = orc_fail(strength, pfE, nsE, key, step, VC_STREAM_V1_FAIL,
                                        bid * bs + pre - key.off, p->trial);
                if (strength > 0.0f) {
                    for (uint32_t ig = 0; ig < p->ngE; ++ig) {
                        float str = strength * p->pE[t * p->ngE + ig];
                        uint32_t id = block_ng * bid + i;
                        lgE[ig] += str * og[ig * cE + id];
                        lhE[ig] += str * oh[ig * cE + id];
                    }
                }
            }
            for (uint32_t i = 0; i < cI; ++i) {
                uint32_t pre = ipre[bid * bs + cE + i];
                float strength = conMat[(local_bid * bs + pre) * bs + l];
                if (strength > 0.0f && pfI > 0.0f)
                    strength = orc_fail(strength, pfI, nsI, key, step, VC_STREAM_V1_FAIL,
                                        bid * bs + pre - key.off, p->trial);
                if (strength > 0.0f) {
                    for (uint32_t ig = 0; ig < p->ngI; ++ig) {
                        float str = strength * p->pI[t * p->ngI + ig];
                        uint32_t id = block_ng * bid + cE * p->ngE + i;
                        lgI[ig] += str * og[ig * cI + id];
                        lhI[ig] += str * oh[ig * cI + id];
                    }
                }
            }
        }
        for (uint32_t ig = 0; ig < p->ngE; ++ig) {
            gE[ig * nV1 + n] += lgE[ig];
            hE[ig * nV1 + n] += lhE[ig];
        }
        for (uint32_t ig = 0; ig < p->ngI; ++ig) {
            gI[ig * nV1 + n] += lgI[ig];
            hI[ig * nV1 + n] += lhI[ig];
        }
    }
    return VC_OK;
}

/*
 * One V1 step in the reference's order: the update kernel (c-FF, d, e) for
 * every neuron, then recal_G for the spikes of this step.
 *   ipre [nblock*bs], npre [2*nblock], og/oh [nblock*(ngE*nE + ngI*nI)]:
 *   the reference's spike lists and unit increments (E list first).
 *   conMat: dense [nblock][maxNear][bs][bs].
 * If do_rec is 0 the recurrent increments are not applied (the lists are
 * still produced).  gapS/gap: the inhibitory neurons' gap-junction terms
 * (NULL: none); lastNoise: the Ornstein-Uhlenbeck noise state (NULL
 * without it).
 */
int orc_v1_step(const vc_v1_params *p, float *v, float *w, float *tBack,
                float *gFF, float *hFF, float *gE, float *hE, float *gI,
                float *hI, const float *tonicDep, const uint32_t *nLGN,
                const float *sLGN, const uint32_t *idLGN,
                const float *lgnSInfo, uint32_t maxNear, const uint32_t *nNear,
                const uint32_t *nbr, const float *conMat, uint32_t *ipre,
                uint32_t *npre, float *og, float *oh, uint32_t step,
                float *out_sInfo, float *out_dep, int do_rec,
                const float *gapS, const float *gap, float *lastNoise) {
    const uint32_t nV1 = p->nV1, bs = p->blockSize, nblock = p->nblock;
    const uint32_t nE = p->nE, nI = bs - nE;
    if (nV1 != nblock * bs) return VC_E_INVALID;
    float *sInfo = out_sInfo;
    int own = 0;
    if (!sInfo) { sInfo = (float *)malloc(sizeof(float) * nV1); own = 1; }

#pragma omp parallel for schedule(static)
    for (uint32_t n = 0; n < nV1; ++n) {
        const int t = orc_type_of(p, n % bs);
        if (orc_multi_all || p->tRef[t] < p->dt)
            sInfo[n] = orc_update_neuron_multi(p, n, v, w, tBack, gFF, hFF, gE, hE, gI,
                                               hI, tonicDep, nLGN, sLGN, idLGN, lgnSInfo,
                                               step, out_dep, gapS, gap, lastNoise);
        else
            sInfo[n] = orc_update_neuron(p, n, v, w, tBack, gFF, hFF, gE, hE, gI,
                                         hI, tonicDep, nLGN, sLGN, idLGN, lgnSInfo,
                                         step, out_dep, gapS, gap, lastNoise);
    }
    /* compaction, src/coredynamics.cu:1994-2046, ascending order */
    const uint32_t block_ng = p->ngE * nE + p->ngI * nI;
#pragma omp parallel for schedule(static)
    for (uint32_t b = 0; b < nblock; ++b) {
        uint32_t cE = 0, cI = 0;
        for (uint32_t l = 0; l < bs; ++l) {
            if (sInfo[b * bs + l] >= 1.0f) { if (l < nE) cE++; else cI++; }
        }
        npre[b] = cE; npre[nblock + b] = cI;
        uint32_t iE = 0, iI = 0;
        for (uint32_t l = 0; l < bs; ++l) {
            float s = sInfo[b * bs + l];
            if (!(s >= 1.0f)) continue;
            float nsp = floorf(s);
            float tsp = (s - nsp) * p->dt;
            uint32_t i = block_ng * b;
            if (l < nE) {
                i += iE;
                for (uint32_t ig = 0; ig < p->ngE; ++ig) {
                    float lg = 0.0f, lh = 0.0f;
                    orc_single_input(&lg, &lh, nsp, p->dt - tsp, p->riseE[ig], p->decayE[ig], p->dodE[ig]);
                    og[ig * cE + i] = lg; oh[ig * cE + i] = lh;
                }
                ipre[b * bs + iE] = l;
                iE++;
            } else {
                i += cE * p->ngE + iI;
                for (uint32_t ig = 0; ig < p->ngI; ++ig) {
                    float lg = 0.0f, lh = 0.0f;
                    orc_single_input(&lg, &lh, nsp, p->dt - tsp, p->riseI[ig], p->decayI[ig], p->dodI[ig]);
                    og[ig * cI + i] = lg; oh[ig * cI + i] = lh;
                }
                ipre[b * bs + cE + iI] = l;
                iI++;
            }
        }
    }
    if (do_rec)
        orc_v1_rec_fast(p, gE, hE, gI, hI, maxNear, nNear, nbr, conMat, ipre, npre, og, oh,
                        step);
    if (own) free(sInfo);
    return VC_OK;
}

/*
 * The dense-scan recurrent gather, recal_G_mat_nd (src/coredynamics.cu
 * :2364-2607, without learning and gap junctions): every post neuron n
 * visits, for ib < nNear[b] and every presynaptic index i < bs of block
 * nbr[b][ib] in order, the pairs with conMat != 0 whose step output
 * sInfo[pre] > 0 (a spike), draws the failure factor, and for strength > 0
 * adds strength * p[type][ig] * og/oh of the presynaptic neuron.  og/oh are
 * read per neuron, as the multi-spike update writes them
 * (src/coredynamics.cu:1205-1230: single_input(nsp, dt - tsp) with nsp and
 * tsp from the neuron's sInfo).  `step` is the step the spikes were fired
 * in (the failure draws' key).
 */
int orc_v1_rec_dense(const vc_v1_params *p, float *gE, float *hE, float *gI, float *hI,
                     uint32_t maxNear, const uint32_t *nNear, const uint32_t *nbr,
                     const float *conMat, const float *sInfo, uint32_t step) {
    const uint32_t nV1 = p->nV1, bs = p->blockSize, nE = p->nE;
#pragma omp parallel for schedule(static)
    for (uint32_t n = 0; n < nV1; ++n) {
        const uint32_t b = n / bs, l = n % bs;
        const int t = orc_type_of(p, l);
        const orc_key key = orc_key_of(p, n);
        float lgE[VC_MAX_NG] = {0}, lhE[VC_MAX_NG] = {0};
        float lgI[VC_MAX_NG] = {0}, lhI[VC_MAX_NG] = {0};
        for (uint32_t ib = 0; ib < nNear[b]; ++ib) {
            const uint32_t bid = nbr[b * maxNear + ib];
            const uint64_t local_bid = (uint64_t)b * maxNear + ib;
            for (uint32_t i = 0; i < bs; ++i) {
                const uint32_t pre = bid * bs + i;
                float strength = conMat[(local_bid * bs + i) * bs + l];
                const float s = sInfo[pre];
                if (!(strength != 0.0f && s > 0.0f)) continue;
                const int jtype = orc_type_of(p, i);
                const float pf = p->synFail[jtype * p->nType + t];
                if (pf > 0.0f)
                    strength = orc_fail(strength, pf, p->synPerCon[jtype * p->nType + t], key,
                                        step, VC_STREAM_V1_FAIL, pre - key.off, p->trial);
                if (!(strength > 0.0f)) continue;
                const float nsp = floorf(s);
                const float tsp = (s - nsp) * p->dt;
                if (i < nE) {
                    for (uint32_t ig = 0; ig < p->ngE; ++ig) {
                        float g = 0.0f, h = 0.0f;
                        orc_single_input(&g, &h, nsp, p->dt - tsp, p->riseE[ig], p->decayE[ig],
                                         p->dodE[ig]);
                        const float str = strength * p->pE[t * p->ngE + ig];
                        lgE[ig] += str * g;
                        lhE[ig] += str * h;
                    }
                } else {
                    for (uint32_t ig = 0; ig < p->ngI; ++ig) {
                        float g = 0.0f, h = 0.0f;
                        orc_single_input(&g, &h, nsp, p->dt - tsp, p->riseI[ig], p->decayI[ig],
                                         p->dodI[ig]);
                        const float str = strength * p->pI[t * p->ngI + ig];
                        lgI[ig] += str * g;
                        lhI[ig] += str * h;
                    }
                }
            }
        }
        for (uint32_t ig = 0; ig < p->ngE; ++ig) {
            gE[ig * nV1 + n] += lgE[ig];
            hE[ig * nV1 + n] += lhE[ig];
        }
        for (uint32_t ig = 0; ig < p->ngI; ++ig) {
            gI[ig * nV1 + n] += lgI[ig];
            hI[ig * nV1 + n] += lhI[ig];
        }
    }
    return VC_OK;
}

/*
 * Gap sums of the inhibitory neurons after a step, for the next one
 * (recal_G_mat_nd, src/coredynamics.cu:2524-2540, 2599-2600; with delays
 * recal_G_mat, :2237-2260, 2305-2306): for inhibitory neuron m of block b,
 * neighbour slots ib < nNear[b], partners l < nI (block-local index nE + l)
 * with gapMat[b][ib][l][m] != 0, gap += strength * v, v = the partner's
 * step output in ring slot cur (without delays) or cur - it2post with
 * it2post = ceil((delayMat[b][ib][nE+l][nE+m] / speed) / dt) (delays),
 * replaced by vThres[partner's type] when > 0 (a spike).
 *   gapMat [nblock][maxNear][nI][nI]; gapDist [nblock][maxNear][nI][nI] =
 *   the inhibitory-to-inhibitory part of V1_delayMat (distances in mm), or
 *   NULL (no delays); train [depth][nV1]; gap [nblock*nI] out.
 * The reference's recal_G_mat_nd_fast indexes gapMat with (i - nE) for a
 * partner index i < nI, which wraps below zero; this follows the other two
 * kernels, where i is the block-local index (>= nE).
 */
int orc_v1_gap(const vc_v1_params *p, uint32_t maxNear, const uint32_t *nNear,
               const uint32_t *nbr, const float *gapMat, const float *gapDist,
               float speed, const float *train, uint32_t depth, uint32_t cur,
               float *gap) {
    const uint32_t nV1 = p->nV1, bs = p->blockSize, nE = p->nE, nI = bs - nE;
    if (nV1 != p->nblock * bs || depth == 0 || cur >= depth) return VC_E_INVALID;
#pragma omp parallel for schedule(static)
    for (uint32_t gi = 0; gi < p->nblock * nI; ++gi) {
        uint32_t b = gi / nI, m = gi % nI;
        float s = 0.0f;
        for (uint32_t ib = 0; ib < nNear[b]; ++ib) {
            uint32_t bid = nbr[b * maxNear + ib];
            uint64_t local_bid = (uint64_t)b * maxNear + ib;
            for (uint32_t l = 0; l < nI; ++l) {
                float strength = gapMat[(local_bid * nI + l) * nI + m];
                if (strength == 0.0f) continue;
                uint32_t slot = cur;
                if (gapDist) {
                    float d = gapDist[(local_bid * nI + l) * nI + m];
                    float time2post = d / speed;
                    uint32_t it2post = (uint32_t)ceilf(time2post / p->dt);
                    int64_t j0 = (int64_t)cur - (int64_t)it2post;
                    while (j0 < 0) j0 += depth;
                    slot = (uint32_t)(j0 % depth);
                }
                float v = train[(uint64_t)slot * nV1 + (uint64_t)bid * bs + nE + l];
                if (v > 0.0f) v = p->vThres[orc_type_of(p, nE + l)];
                s += strength * v;
            }
        }
        gap[gi] = s;
    }
    return VC_OK;
}

/*
 * Recurrent input with axonal delays: recal_G_mat, src/coredynamics.cu
 * :2062-2309 (noDelay = 0; synFail 0, no learning, no gap junctions), in its
 * own loop order: for every post neuron n, neighbour blocks ib, every
 * presynaptic neuron i with conMat != 0, ring slots j = 0, 1 of the delay,
 * receptors ig; the sums are added to gE/hE/gI/hI at the end (:2286-2303).
 *   conMat, delayMat: dense [nblock][maxNear][bs pre][bs post] (strength,
 *   distance in mm); train: [depth][nV1] step outputs (sInfo), the step
 *   whose spikes are newest in slot cur.
 * Returns VC_E_INVALID if a connected pair's delay does not fit the ring.
 */
int orc_v1_rec_delay(const vc_v1_params *p, float *gE, float *hE, float *gI,
                     float *hI, uint32_t maxNear, const uint32_t *nNear,
                     const uint32_t *nbr, const float *conMat,
                     const float *delayMat, float speed, const float *train,
                     uint32_t depth, uint32_t cur) {
    const uint32_t nV1 = p->nV1, bs = p->blockSize;
    const float dt = p->dt;
    int bad = 0;
    if (nV1 != p->nblock * bs || depth == 0 || cur >= depth) return VC_E_INVALID;
#pragma omp parallel for schedule(static) reduction(|:bad)
    for (uint32_t n = 0; n < nV1; ++n) {
        uint32_t b = n / bs, l = n % bs;
        int t = orc_type_of(p, l);
        float lgE[VC_MAX_NG] = {0}, lhE[VC_MAX_NG] = {0};
        float lgI[VC_MAX_NG] = {0}, lhI[VC_MAX_NG] = {0};
        for (uint32_t ib = 0; ib < nNear[b]; ++ib) {
            uint32_t bid = nbr[b * maxNear + ib];
            uint64_t local_bid = (uint64_t)b * maxNear + ib;
            for (uint32_t i = 0; i < bs; ++i) {
                uint32_t ipre = bid * bs + i;
                uint64_t mid = (local_bid * bs + i) * bs + l;
                float strength = conMat[mid];
                if (strength == 0.0f) continue;
                int inh = i >= p->nE;
                float time2post = delayMat[mid] / speed;
                uint32_t it2post = (uint32_t)ceilf(time2post / dt);
                time2post = (float)it2post * dt - time2post;
                if (it2post >= depth) { bad = 1; continue; }
                int j0 = (int)cur - (int)it2post;
                if (j0 < 0) j0 += (int)depth;
                for (uint32_t j = 0; j < 2; ++j) {
                    uint64_t isp = (uint64_t)nV1 * ((j0 + j) % depth) + ipre;
                    float dtsp = train[isp];
                    if (dtsp >= 1.0f) {
                        uint32_t nsp = (uint32_t)floorf(dtsp);
                        dtsp = (dtsp - (float)nsp + (float)j) * dt - time2post;
                        if (dtsp < dt && dtsp >= 0.0f) {
                            float f = strength;
                            if (f > 0.0f) {
                                dtsp = dt - dtsp;
                                if (!inh) {
                                    for (uint32_t ig = 0; ig < p->ngE; ++ig)
                                        orc_single_input(&lgE[ig], &lhE[ig],
                                                         (float)nsp * f * p->pE[t * p->ngE + ig],
                                                         dtsp, p->riseE[ig], p->decayE[ig],
                                                         p->dodE[ig]);
                                } else {
                                    for (uint32_t ig = 0; ig < p->ngI; ++ig)
                                        orc_single_input(&lgI[ig], &lhI[ig],
                                                         (float)nsp * f * p->pI[t * p->ngI + ig],
                                                         dtsp, p->riseI[ig], p->decayI[ig],
                                                         p->dodI[ig]);
                                }
                            }
                        }
                    }
                }
            }
        }
        for (uint32_t ig = 0; ig < p->ngE; ++ig) {
            gE[ig * nV1 + n] += lgE[ig];
            hE[ig * nV1 + n] += lhE[ig];
        }
        for (uint32_t ig = 0; ig < p->ngI; ++ig) {
            gI[ig * nV1 + n] += lgI[ig];
            hI[ig * nV1 + n] += lhI[ig];
        }
    }
    return bad ? VC_E_INVALID : VC_OK;
}

/*
 * Far (long-range) connections with delays: recal_G_vec, src/coredynamics.cu
 * :181-372 (the noFarDelay = 0 branch; synFail 0), then sum_G (:2314-2349):
 * for every post neuron n with far inputs, its list in file order, ring
 * slots k = 0, 1 of the delay, receptors; the sum is added to gE/hE/gI/hI.
 *   fOff[nV1+1], fPre[], fW[] (strength, already * sRatioV1), fDist[] (mm);
 *   train: [depth][nV1] step outputs, the newest in slot cur.
 * The reference keeps one ring per connection, fTrainDepth = ceil(d/speed/dt)
 * + 1 slots long, written every step (src/patch.cu:3838-3851, 5808); the
 * slots it reads hold the same steps' outputs as the shared ring here.
 */
int orc_v1_rec_far(const vc_v1_params *p, float *gE, float *hE, float *gI,
                   float *hI, const uint32_t *fOff, const uint32_t *fPre,
                   const float *fW, const float *fDist, float speed,
                   const float *train, uint32_t depth, uint32_t cur) {
    const uint32_t nV1 = p->nV1, bs = p->blockSize;
    const float dt = p->dt;
    int bad = 0, anySpiked = 0;
    if (depth == 0 || cur >= depth) return VC_E_INVALID;
    /* fill_fSpikeTrain's test (src/stats.cu:114-126): recal_G_vec and sum_G
     * run only in a step in which some far presynaptic neuron's output is
     * above 1.0 (src/patch.cu:6597-6598); batched trials (p->blockTrial) are
     * separate runs of the reference, so the test is per trial */
    uint32_t nTrial = 1;
    if (p->blockTrial)
        for (uint32_t b = 0; b < p->nblock; ++b)
            if (p->blockTrial[b] + 1 > nTrial) nTrial = p->blockTrial[b] + 1;
    unsigned char *farSpiked = (unsigned char *)calloc(nTrial, 1);
    if (!farSpiked) return VC_E_INVALID;
    for (uint32_t j = 0; j < fOff[nV1]; ++j)
        if (fPre[j] < nV1 && train[(uint64_t)nV1 * cur + fPre[j]] > 1.0f) {
            farSpiked[p->blockTrial ? p->blockTrial[fPre[j] / bs] : 0] = 1;
            anySpiked = 1;
        }
    if (!anySpiked) { free(farSpiked); return VC_OK; }
#pragma omp parallel for schedule(static) reduction(|:bad)
    for (uint32_t n = 0; n < nV1; ++n) {
        if (fOff[n + 1] == fOff[n]) continue;          /* sum_G: nVec[id] > 0 */
        if (!farSpiked[p->blockTrial ? p->blockTrial[n / bs] : 0]) continue;
        int t = orc_type_of(p, n % bs);
        float lgE[VC_MAX_NG] = {0}, lhE[VC_MAX_NG] = {0};
        float lgI[VC_MAX_NG] = {0}, lhI[VC_MAX_NG] = {0};
        for (uint32_t j = fOff[n]; j < fOff[n + 1]; ++j) {
            uint32_t ipre = fPre[j], tid = ipre % bs;
            float strength = fW[j];
            float time2post = fDist[j] / speed;
            uint32_t it2post = (uint32_t)ceilf(time2post / dt);
            time2post = (float)it2post * dt - time2post;
            if (ipre >= nV1 || it2post >= depth) { bad = 1; continue; }
            int k0 = (int)cur - (int)it2post;
            if (k0 < 0) k0 += (int)depth;
            for (uint32_t k = 0; k < 2; ++k) {
                float dtsp = train[(uint64_t)nV1 * ((k0 + k) % depth) + ipre];
                if (dtsp >= 1.0f) {
                    uint32_t nsp = (uint32_t)floorf(dtsp);
                    dtsp = (dtsp - (float)nsp + (float)k) * dt - time2post;
                    if (dtsp < dt && dtsp >= 0.0f) {
                        float f = strength;
                        if (f > 0.0f) {
                            dtsp = dt - dtsp;
                            if (tid < p->nE) {
                                for (uint32_t ig = 0; ig < p->ngE; ++ig)
                                    orc_single_input(&lgE[ig], &lhE[ig],
                                                     (float)nsp * f * p->pE[t * p->ngE + ig], dtsp,
                                                     p->riseE[ig], p->decayE[ig], p->dodE[ig]);
                            } else {
                                for (uint32_t ig = 0; ig < p->ngI; ++ig)
                                    orc_single_input(&lgI[ig], &lhI[ig],
                                                     (float)nsp * f * p->pI[t * p->ngI + ig], dtsp,
                                                     p->riseI[ig], p->decayI[ig], p->dodI[ig]);
                            }
                        }
                    }
                }
            }
        }
        for (uint32_t ig = 0; ig < p->ngE; ++ig) {
            gE[ig * nV1 + n] += lgE[ig];
            hE[ig * nV1 + n] += lhE[ig];
        }
        for (uint32_t ig = 0; ig < p->ngI; ++ig) {
            gI[ig * nV1 + n] += lgI[ig];
            hI[ig * nV1 + n] += lhI[ig];
        }
    }
    free(farSpiked);
    return bad ? VC_E_INVALID : VC_OK;
}

/* ====================================================================== */
/* LGN                                                                     */
/* ====================================================================== */

/* Exact-float bilinear filtering of one channel with border addressing
 * (0 outside) and normalised coordinates, the semantics of
 * tex2DLayered with cudaFilterModeLinear / cudaAddressModeBorder
 * (src/patch.h:95-109) restated in binary32 arithmetic. */
static inline float orc_texel(const float *img, int W, int H, int i, int j) {
    if (i < 0 || j < 0 || i >= W || j >= H) return 0.0f;
    return img[(size_t)j * W + i];
}
static float orc_bilinear(const float *img, int W, int H, float x, float y) {
    float xt = x * (float)W - 0.5f;
    float yt = y * (float)H - 0.5f;
    if (!(xt > -2.0f && xt < (float)W + 1.0f && yt > -2.0f && yt < (float)H + 1.0f))
        return 0.0f;
    float fx = floorf(xt), fy = floorf(yt);
    float ax = xt - fx, ay = yt - fy;
    int i0 = (int)fx, j0 = (int)fy;
    float t00 = orc_texel(img, W, H, i0, j0);
    float t10 = orc_texel(img, W, H, i0 + 1, j0);
    float t01 = orc_texel(img, W, H, i0, j0 + 1);
    float t11 = orc_texel(img, W, H, i0 + 1, j0 + 1);
    float bx = 1.0f - ax, by = 1.0f - ay;
    return by * (bx * t00 + ax * t10) + ay * (bx * t01 + ax * t11);
}
float orc_bilinear_export(const float *img, int W, int H, float x, float y) {
    return orc_bilinear(img, W, H, x, y);
}

/* get_intensity, src/discrete_input_convol.cu:35-76 */
static float orc_intensity(uint32_t cone, const float *frame, int W, int H,
                           float x, float y) {
    size_t np = (size_t)W * H;
    switch (cone) {
    case 0: return orc_bilinear(frame, W, H, x, y);
    case 1: return orc_bilinear(frame + np, W, H, x, y);
    case 2: return orc_bilinear(frame + 2 * np, W, H, x, y);
    case 3: return (orc_bilinear(frame, W, H, x, y) + orc_bilinear(frame + np, W, H, x, y)) / 2.0f;
    case 4: return orc_bilinear(frame, W, H, x, y) * 0.45436904f +
                   orc_bilinear(frame + np, W, H, x, y) * 0.47353315f +
                   orc_bilinear(frame + 2 * np, W, H, x, y) * 0.0720978f;
    default: return 0.0f;
    }
}

static inline float orc_clip_contrast(float I, float mean) {
    float c = (mean > 0.0f) ? I / mean - 1.0f : I;
    if (fabsf(c) > 1.0f) c = copysignf(1.0f, c);
    return c;
}

void orc_lgn_frame_response(const float *frame0, uint32_t W, uint32_t H,
                            uint32_t nLGN, uint32_t nSample1D, const float *SCx,
                            const float *SCy, const float *SW,
                            const uint32_t *coneType, const uint32_t *cellTrial,
                            uint64_t trialFrameStride, float *R) {
    const int nS = (int)(nSample1D * nSample1D);
#pragma omp parallel for schedule(static)
    for (uint32_t c = 0; c < nLGN; ++c) {
        const float *frame = cellTrial ? frame0 + (size_t)cellTrial[c] * trialFrameStride : frame0;
        float *a = (float *)malloc(sizeof(float) * nS * 4);
        float *bS = a + nS, *fC = a + 2 * nS, *fS = a + 3 * nS;
        uint32_t tC = coneType[c], tS = coneType[nLGN + c];
        const float *xC = SCx + ((size_t)c * 2 + 0) * nS, *yC = SCy + ((size_t)c * 2 + 0) * nS;
        const float *xS = SCx + ((size_t)c * 2 + 1) * nS, *yS = SCy + ((size_t)c * 2 + 1) * nS;
        for (int s = 0; s < nS; ++s) {
            a[s] = orc_intensity(tC, frame, (int)W, (int)H, xS[s], yS[s]);
            bS[s] = orc_intensity(tS, frame, (int)W, (int)H, xS[s], yS[s]);
        }
        float meanC = orc_block_reduce(a, nS, nS) / (float)nS;
        float meanS = orc_block_reduce(bS, nS, nS) / (float)nS;
        for (int s = 0; s < nS; ++s) {
            float I = orc_intensity(tC, frame, (int)W, (int)H, xC[s], yC[s]);
            fC[s] = SW[s] * orc_clip_contrast(I, meanC);
            fS[s] = SW[s] * orc_clip_contrast(bS[s], meanS);
        }
        R[c * 2 + 0] = orc_block_reduce(fC, nS, nS);
        R[c * 2 + 1] = orc_block_reduce(fS, nS, nS);
        free(a);
    }
}

static inline int64_t orc_floor_div(int64_t a, int64_t b) {
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
    return q;
}

/* src/LGN_props.cuh:94-100; 1.0 is a double literal in the reference */
static inline float orc_transform(float C50, float K, float A, float B,
                                  float input) {
    float X = (float)(1.0 / (double)(1.0f + vc_expf(K * (C50 - input))));
    return A * X + B;
}

/* one LGN cell, one step: returns sInfo */
static float orc_lgn_cell_step(const vc_lgn_params *p, const float *TW,
                               const float *kC, const float *kS,
                               const float *maxConvol, const float *c50,
                               const float *K, const float *A, const float *B,
                               float *leftTimeRate, float *lastNegLogRand,
                               const float *Rring, uint32_t c, uint32_t it,
                               float *fr_out, float *convol_out, uint32_t *nsp_out,
                               float *tsp_out) {
    const uint32_t nKS = p->nKernelSample;
    float prod[2][1024];
    for (int r = 0; r < 2; ++r) {
        for (uint32_t i = 0; i < nKS; ++i) {
            /* the tap's time in units of dt/denorm (src/patch.cu:4999-5030):
             * frame f covers [f ntPerFrame, (f+1) ntPerFrame) of them */
            int64_t tau = (int64_t)it + 1 - (int64_t)p->nRetrace +
                          (int64_t)p->iKernelSampleT0 + (int64_t)i * p->kernelSampleInterval;
            int64_t den = p->denorm ? (int64_t)p->denorm : 1;
            int64_t f = orc_floor_div(tau * den, (int64_t)p->ntPerFrame);
            int64_t slot = f % (int64_t)p->maxFrame;
            if (slot < 0) slot += p->maxFrame;
            float Rv = Rring[((size_t)slot * p->nLGN + c) * 2 + r];
            prod[r][i] = Rv * TW[((size_t)c * 2 + r) * nKS + i];
        }
    }
    /* the reference sums the taps with block_reduce2 over a patch of nSample
     * threads; nSample >= 64 >= nKS here, so the patch is one tree. */
    float convolC = 0.0f + orc_block_reduce(prod[0], (int)nKS, 256);
    float convolS = 0.0f + orc_block_reduce(prod[1], (int)nKS, 256);
    convolC *= kC[c];
    convolS *= kS[c];
    float convol = convolC + convolS;
    if (convol_out) *convol_out = convol;

    float current = convol * p->convolRatio;
    float mx = maxConvol[c];
    if (current < 0.0f) current = 0.0f;
    float fr = p->frRatio * mx * orc_transform(c50[c], K[c], A[c], B[c], current / mx);
    if (fr_out) *fr_out = fr;

    /* get_spike, src/discrete_input_convol.cu:1587-1622 */
    float rate = fr / 1000.0f;
    float lTR = leftTimeRate[c], lNL = lastNegLogRand[c];
    float rT = p->dt * rate;
    float next_rT = lNL - lTR;
    uint32_t nsp = 0;
    float tsp = 0.0f;
    if (next_rT > rT) {
        lTR += rT;
    } else {
        float old_rT = 0.0f;
        uint32_t k = 0;
        vc_u32x4 pack;
        float tr = p->tRef * rate;
        uint32_t klo = p->seed_lo, khi = p->seed_hi, ent = c;
        if (p->cellTrial) {
            uint32_t t = p->cellTrial[c];
            klo = p->trialSeed[2 * t];
            khi = p->trialSeed[2 * t + 1];
            ent = c - p->trialLgnOff[t];
        }
        do {
            if ((k & 3u) == 0u)
                pack = vc_draw4(klo, khi, ent, it, VC_STREAM_LGN_SPIKE, k >> 2, p->trial);
            float u = vc_u01_open0(pack.v[k & 3u]);
            k++;
            float x = -vc_logf(u);
            lNL = (x > tr) ? x : tr;
            old_rT = next_rT;
            next_rT += lNL;
            nsp++;
        } while (next_rT <= rT);
        lTR = rT - old_rT;
        tsp = old_rT / (rate * (float)nsp);
    }
    float sInfo = (float)nsp + tsp / p->dt;
    if (sInfo > 0.0f) lastNegLogRand[c] = lNL;
    leftTimeRate[c] = lTR;
    if (nsp_out) *nsp_out = nsp;
    if (tsp_out) *tsp_out = tsp;
    return sInfo;
}

/* The LTP trace of an LGN cell over one step (LGN_nonlinear's learning part,
 * src/discrete_input_convol.cu:1736-1743, 1775-1821): x is the value at the
 * step start (the surface slot varSlot); with a spike the step's nsp spikes
 * are added and then x decays to their mean time tsp (the value kept in the
 * third slot), then to the step end (the other slot); without one x decays
 * over dt. */
static void orc_lgn_trace_step(float *x, uint32_t nsp, float tsp, float sInfo, float dt,
                               float tau, float *xStart, float *xSpike) {
    float v = *x;
    *xStart = v;
    float rest = dt;
    if (sInfo > 0.0f) {
        v = v + (float)nsp;
        v = v * vc_expf(-tsp / tau);
        rest = dt - tsp;
    }
    *xSpike = v;
    v = v * vc_expf(-rest / tau);
    *x = v;
}

/* orc_lgn_steps with the cells' LTP trace (xVar [nLGN] in/out; xStart,
 * xSpike [nsteps][nLGN] out; xVar NULL = no trace). */
void orc_lgn_steps_learn(const vc_lgn_params *p, const float *TW, const float *kC,
                         const float *kS, const float *maxConvol, const float *c50,
                         const float *K, const float *A, const float *B,
                         float *leftTimeRate, float *lastNegLogRand,
                         const float *Rring, uint32_t it0, uint32_t nsteps,
                         float *sInfo, float *fr, float *convol, float tauLTP,
                         float *xVar, float *xStart, float *xSpike) {
    for (uint32_t s = 0; s < nsteps; ++s) {
        uint32_t it = it0 + s;
#pragma omp parallel for schedule(static)
        for (uint32_t c = 0; c < p->nLGN; ++c) {
            float frv, cv, tsp;
            uint32_t nsp;
            float si = orc_lgn_cell_step(p, TW, kC, kS, maxConvol, c50, K, A, B,
                                         leftTimeRate, lastNegLogRand, Rring, c,
                                         it, &frv, &cv, &nsp, &tsp);
            size_t o = (size_t)s * p->nLGN + c;
            sInfo[o] = si;
            if (fr) fr[o] = frv;
            if (convol) convol[o] = cv;
            if (xVar)
                orc_lgn_trace_step(&xVar[c], nsp, tsp, si, p->dt, tauLTP, &xStart[o],
                                   &xSpike[o]);
        }
    }
}

void orc_lgn_steps(const vc_lgn_params *p, const float *TW, const float *kC,
                   const float *kS, const float *maxConvol, const float *c50,
                   const float *K, const float *A, const float *B,
                   float *leftTimeRate, float *lastNegLogRand,
                   const float *Rring, uint32_t it0, uint32_t nsteps,
                   float *sInfo, float *fr, float *convol) {
    orc_lgn_steps_learn(p, TW, kC, kS, maxConvol, c50, K, A, B, leftTimeRate,
                        lastNegLogRand, Rring, it0, nsteps, sInfo, fr, convol, 0.0f, NULL,
                        NULL, NULL);
}

/* One step of the feed-forward plasticity with learning = 3 (only the FF_E
 * rule), compute_V_collect_spike_learnFF_fast, src/coredynamics.cu:1631-1992:
 * for each excitatory neuron (local id < nE), row 0 of its post variables is
 * the value at the step start and row 1 the value at its spike time; the
 * homeostatic change of its total (:1765-1800); per synapse with a strength
 * (:1803-1927) the LTD term at the LGN cell's spike (whether or not the
 * neuron spiked) and the LTP term at the neuron's spike (whether or not the
 * cell spiked), the clip; then the post variables take the step's spikes and
 * decay to the step end (:1930-1960).  sLGN is updated in place; the trace
 * arrays are [nV1] (inhibitory entries untouched).  lgnS, xStart, xSpike are
 * the step's rows over the LGN cells. */
int orc_v1_learn_ff(const vc_v1_params *p, const vc_learn_ff_params *lp,
                    const uint32_t *nLGN, float *sLGN, const uint32_t *idLGN,
                    float *vLTD, float *vTrip, float *vAvg, float *totalFF,
                    const float *totalFFinf, const float *sInfo, const float *lgnS,
                    const float *xStart, const float *xSpike) {
#pragma omp parallel for schedule(static)
    for (uint32_t n = 0; n < p->nV1; ++n) {
        if (n % p->blockSize >= p->nE) continue;
        const float dt = p->dt;
        float nsp = 0.0f, tsp = dt;
        if (sInfo[n] > 0.0f) {
            nsp = floorf(sInfo[n]);
            tsp = (sInfo[n] - nsp) * dt;
        }
        /* lFF[row][0] = LTD, lFF[row][1] = triplet; lAvg[row] */
        float lFF[2][2] = {{vLTD[n], vTrip[n]}, {0.0f, 0.0f}};
        float lAvg[2] = {vAvg[n], 0.0f};
        if (nsp > 0.0f) {
            lFF[1][0] = lFF[0][0] * vc_expf(-tsp / lp->tauLTD);
            lFF[1][1] = lFF[0][1] * vc_expf(-tsp / lp->tauTrip);
            lAvg[1] = lAvg[0] * vc_expf(-tsp / lp->tauAvg);
        }
        uint32_t m = nLGN[n];
        float f0tot = 0.0f, fnew = 0.0f, hc = 0.0f, df_tot = 0.0f;
        if (lp->applyHomeo) {
            f0tot = totalFF[n];
            float finf = totalFFinf[n];
            float d = f0tot - finf;
            if (d > 0.0f || lp->symmetricHomeo) {
                fnew = d * lp->expHomeo + finf;
                hc = lp->applyHomeo == 1 ? fnew / f0tot : (fnew - f0tot) / (float)m;
            } else {
                fnew = f0tot;
                hc = lp->applyHomeo == 1 ? 1.0f : 0.0f;
            }
        }
        for (uint32_t i = 0; i < m; ++i) {
            size_t lid = (size_t)i * p->nV1 + n;
            float f = sLGN[lid];
            if (f == 0.0f && !lp->rebound) continue;
            if (lp->applyHomeo == 1) f *= hc;
            if (lp->applyHomeo == 2) f += hc;
            uint32_t c = idLGN[lid];
            float sF = lgnS[c];
            float nspF = floorf(sF);
            float tspF = (sF > 0.0f ? sF - nspF : 1.0f) * dt;
            if (nspF > 0.0f) {                 /* LTD */
                int row;
                float del;
                if (tspF < tsp) { row = 0; del = tspF; }
                else            { row = 1; del = tspF - tsp; }
                float A_LTD = lp->A_ratio;
                if (lp->targetFR > 0.0f) A_LTD *= lp->tauLTP * lAvg[row] * lAvg[row];
                float df = lFF[row][0] * vc_expf(-del / lp->tauLTD) * A_LTD;
                f -= df;
                if (lp->applyHomeo) df_tot -= df;
            }
            if (nsp > 0.0f) {                  /* LTP */
                float pre, del;
                if (tspF < tsp) { pre = xSpike[c]; del = tsp - tspF; }
                else            { pre = xStart[c]; del = tsp; }
                float df = pre * vc_expf(-del / lp->tauLTP) * lFF[1][1] * lp->A_LTP;
                if (lp->applyHomeo) df_tot += df;
                f += df;
            }
            if (f < lp->gmin) {
                if (lp->applyHomeo) df_tot += lp->gmin - f;
                f = lp->gmin;
            }
            if (f > lp->gmax) {
                if (lp->applyHomeo) df_tot -= f - lp->gmax;
                f = lp->gmax;
            }
            sLGN[lid] = f;
        }
        if (lp->applyHomeo) totalFF[n] = fnew + df_tot;
        int row = nsp > 0.0f ? 1 : 0;
        float del = dt;
        if (nsp > 0.0f) del -= tsp;
        lFF[row][0] += nsp;
        lFF[row][1] += nsp;
        lFF[row][0] = lFF[row][0] * vc_expf(-del / lp->tauLTD);
        lFF[row][1] = lFF[row][1] * vc_expf(-del / lp->tauTrip);
        lAvg[row] += nsp;
        lAvg[row] = lAvg[row] * vc_expf(-del / lp->tauAvg);
        vLTD[n] = lFF[row][0];
        vTrip[n] = lFF[row][1];
        vAvg[n] = lAvg[row];
    }
    return 0;
}

void orc_lgn_init(const vc_lgn_params *p, float *leftTimeRate,
                  float *lastNegLogRand) {
    for (uint32_t c = 0; c < p->nLGN; ++c) {
        uint32_t klo = p->seed_lo, khi = p->seed_hi, ent = c;
        if (p->cellTrial) {
            uint32_t t = p->cellTrial[c];
            klo = p->trialSeed[2 * t];
            khi = p->trialSeed[2 * t + 1];
            ent = c - p->trialLgnOff[t];
        }
        vc_u32x4 r = vc_draw4(klo, khi, ent, 0xffffffffu, VC_STREAM_LGN_INIT, 0, p->trial);
        leftTimeRate[c] = 0.0f;
        lastNegLogRand[c] = -vc_logf(vc_u01_open0(r.v[0]));
    }
}

/* ---------------------------------------------------------------------- */
/* LGN precompute (setup; C library transcendentals)                       */

/* src/discrete_input_convol.cu:14-33 */
static float orc_temporal_kernel(float tau, float tauR, float tauD, float ratio,
                                 float nR, float nD, float lfac1, float lfac2) {
    float tau1 = tau / tauR;
    float tau2 = tau / tauD;
    float A1 = (nR - 1) * logf(tau1);
    float A2 = (nD - 1) * logf(tau2);
    float exponent = A1 - tau1 - lfac1;
    float tpR = expf(exponent) / tauR;
    exponent = A2 - tau2 - lfac2;
    float tpD = expf(exponent) / tauD;
    return tpR - ratio * tpD;
}

static float orc_spatial_kernel(float x, float y, float rx, float ry) {
    return expf(-x * x / (rx * rx) - y * y / (ry * ry));
}

void orc_lgn_precompute(const vc_lgn_setup *q, float *TW, float *SW,
                        float *SCx, float *SCy, float *maxConvol) {
    const uint32_t nLGN = q->nLGN, n1D = q->nSample1D, nS = n1D * n1D;
    const uint32_t nKS = q->nKernelSample;
    const float kernelSampleDt = (float)q->kernelSampleInterval * q->dt;
    const float kernelSampleT0 = (float)q->iKernelSampleT0 * q->dt;
    /* temporal weights, src/discrete_input_convol.cu:230-305 */
#pragma omp parallel for schedule(static)
    for (uint32_t c = 0; c < nLGN; ++c) {
        float *absw = (float *)malloc(sizeof(float) * nS);
        for (uint32_t r = 0; r < 2; ++r) {
            uint32_t lid = r * nLGN + c;
            float lfac1 = lgammaf(q->nR[lid]);
            float lfac2 = lgammaf(q->nD[lid]);
            float T0 = kernelSampleT0 - q->delay[lid];
            float total = 0.0f;
            float *tw = TW + ((size_t)c * 2 + r) * nKS;
            uint32_t nPatch = (nKS + nS - 1) / nS;
            for (uint32_t ip = 0; ip < nPatch; ++ip) {
                uint32_t nAct = (ip + 1 < nPatch) ? nS : nKS - ip * nS;
                for (uint32_t tid = 0; tid < nS; ++tid) {
                    if (tid < nAct) {
                        uint32_t j = ip * nS + tid;
                        uint32_t twid = nKS - 1 - j;
                        float t = (float)twid * kernelSampleDt + T0;
                        float v = 0.0f;
                        if (!(t < 0.0f))
                            v = orc_temporal_kernel(t, q->tauR[lid], q->tauD[lid], q->ratio[lid],
                                                    q->nR[lid], q->nD[lid], lfac1, lfac2);
                        tw[j] = v;
                        absw[tid] = fabsf(v);
                    } else {
                        absw[tid] = 0.0f;
                    }
                }
                total += orc_block_reduce(absw, (int)nS, (int)nS);
            }
            for (uint32_t j = 0; j < nKS; ++j) tw[j] /= total;
        }
        free(absw);
    }
    /* shared spatial weights from cell 0 centre, :332-336 */
    {
        float rx = q->rw[0], ry = q->rh[0];
        float wSpan = (float)((double)(q->nsig * rx) / ORC_SQRT2);
        float dw = 2 * wSpan / (float)n1D;
        float hSpan = (float)((double)(q->nsig * ry) / ORC_SQRT2);
        float dh = 2 * hSpan / (float)n1D;
        float *sw = (float *)malloc(sizeof(float) * nS);
        for (uint32_t ty = 0; ty < n1D; ++ty)
            for (uint32_t tx = 0; tx < n1D; ++tx) {
                float w = (float)((tx + 0.5) * (double)dw - (double)wSpan);
                float h = (float)((ty + 0.5) * (double)dh - (double)hSpan);
                sw[ty * n1D + tx] = orc_spatial_kernel(w, h, rx, ry) * 1.0f;
            }
        float tot = orc_block_reduce(sw, (int)nS, (int)nS);
        for (uint32_t s = 0; s < nS; ++s) SW[s] = sw[s] / tot;
        free(sw);
    }
    /* sample coordinates, flat retina, :398-417 */
#pragma omp parallel for schedule(static)
    for (uint32_t c = 0; c < nLGN; ++c) {
        for (uint32_t r = 0; r < 2; ++r) {
            uint32_t lid = r * nLGN + c;
            float rx = q->rw[lid], ry = q->rh[lid];
            float wSpan = (float)((double)(q->nsig * rx) / ORC_SQRT2);
            float dw = 2 * wSpan / (float)n1D;
            float hSpan = (float)((double)(q->nsig * ry) / ORC_SQRT2);
            float dh = 2 * hSpan / (float)n1D;
            float polar = q->polar[lid], ecc = q->ecc[lid];
            float coso = cosf(q->orient[lid]), sino = sinf(q->orient[lid]);
            for (uint32_t ty = 0; ty < n1D; ++ty)
                for (uint32_t tx = 0; tx < n1D; ++tx) {
                    float w = (float)((tx + 0.5) * (double)dw - (double)wSpan);
                    float h = (float)((ty + 0.5) * (double)dh - (double)hSpan);
                    float x = q->x0 + (ecc * cosf(polar) + w * coso - h * sino) * q->normViewDistance;
                    float y = q->y0 + (ecc * sinf(polar) + w * sino + h * coso) * q->normViewDistance;
                    size_t id = ((size_t)c * 2 + r) * nS + ty * n1D + tx;
                    SCx[id] = x;
                    SCy[id] = y;
                }
        }
    }
    /* maxConvol, src/discrete_input_convol.cu:746-881, 1024 threads */
    const uint32_t m1D = q->maxSample1D, mS = m1D * m1D;
    const uint32_t nPatch = mS / 1024;
#pragma omp parallel for schedule(dynamic, 4)
    for (uint32_t c = 0; c < nLGN; ++c) {
        float *swC = (float *)malloc(sizeof(float) * mS * 2);
        float *swS = swC + mS;
        float buf[1024];
        float rxs = q->rw[nLGN + c], rys = q->rh[nLGN + c];
        float xs = q->polar[nLGN + c], ys = q->ecc[nLGN + c];
        float wSpan = (float)((double)(q->nsig * rxs) / ORC_SQRT2);
        float hSpan = (float)((double)(q->nsig * rys) / ORC_SQRT2);
        float rxc = q->rw[c], ryc = q->rh[c];
        float R = wSpan > hSpan ? wSpan : hSpan;
        float xc = q->polar[c], yc = q->ecc[c];
        wSpan = (float)((double)(q->nsig * rxc) / ORC_SQRT2);
        hSpan = (float)((double)(q->nsig * ryc) / ORC_SQRT2);
        float dx = yc * cosf(xc) - ys * cosf(xs);
        float dy = yc * sinf(xc) - ys * sinf(xs);
        float r = sqrtf(dx * dx + dy * dy);
        float cov = q->covariant[c];
        float orients = q->orient[nLGN + c], orientc = q->orient[c];
        r += wSpan > hSpan ? wSpan : hSpan;
        float coss = cosf(orients);
        R = r > R ? r : R;
        float sins = sinf(orients);
        float cosc = cosf(orientc), sinc = sinf(orientc);
        float ks = q->k[nLGN + c], kc = q->k[c];
        float ds = 2 * R / (float)m1D;
        float sumC = 0.0f, sumS = 0.0f;
        for (uint32_t ip = 0; ip < nPatch; ++ip) {
            for (uint32_t t = 0; t < 1024; ++t) {
                uint32_t pid = ip * 1024 + t;
                uint32_t wi = pid % m1D, hi = pid / m1D;
                float lx = (float)((wi + 0.5) * (double)ds - (double)R);
                float ly = (float)((hi + 0.5) * (double)ds - (double)R);
                float x = cosc * (lx - dx) + sinc * (ly - dy);
                float y = -sinc * (lx - dx) + cosc * (ly - dy);
                swC[pid] = orc_spatial_kernel(x, y, rxc, ryc);
            }
            sumC += orc_block_reduce(swC + ip * 1024, 1024, 1024);
            for (uint32_t t = 0; t < 1024; ++t) {
                uint32_t pid = ip * 1024 + t;
                uint32_t wi = pid % m1D, hi = pid / m1D;
                float lx = (float)((wi + 0.5) * (double)ds - (double)R);
                float ly = (float)((hi + 0.5) * (double)ds - (double)R);
                float x = coss * lx + sins * ly;
                float y = -sins * lx + coss * ly;
                swS[pid] = orc_spatial_kernel(x, y, rxs, rys);
            }
            sumS += orc_block_reduce(swS + ip * 1024, 1024, 1024);
        }
        for (uint32_t pid = 0; pid < mS; ++pid) { swC[pid] /= sumC; swS[pid] /= sumS; }
        float convol = 0.0f;
        for (uint32_t it = 0; it < nKS; ++it) {
            float tc = TW[(size_t)c * 2 * nKS + it];
            float ts = TW[((size_t)c * 2 + 1) * nKS + it];
            for (uint32_t ip = 0; ip < nPatch; ++ip) {
                for (uint32_t t = 0; t < 1024; ++t) {
                    uint32_t pid = ip * 1024 + t;
                    float lc = swC[pid] * tc * kc;
                    float ls = swS[pid] * ts * ks;
                    float d;
                    if (fabsf(lc) > fabsf(ls)) d = fabsf(lc) + ls * copysignf(cov, lc);
                    else d = fabsf(ls) + lc * copysignf(cov, ls);
                    buf[t] = d;
                }
                convol += orc_block_reduce(buf, 1024, 1024);
            }
        }
        maxConvol[c] = convol;
        free(swC);
    }
}

int orc_num_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
