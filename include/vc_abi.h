/*
 * vc_abi.h -- C ABI of libvc_hip.so, the MI355X (gfx950) implementation of
 * the patchV1 simulation step.
 *
 * Every entry point:
 *   - takes device pointers and sizes only (no framework types), plus a
 *     `stream` argument that is a hipStream_t passed as void* (NULL = the
 *     null stream);
 *   - allocates nothing: the caller owns every buffer;
 *   - returns 0 on success or a negative VC_E_* code; vc_last_error() returns
 *     a message describing the most recent failure on the calling thread.
 *     Nothing calls exit().
 *
 * Reference interfaces replaced (paths relative to the reference repo):
 *   vc_lgn_precompute      store_PM + parvo_maxConvol
 *                            src/discrete_input_convol.cu:427-744, 746-881;
 *                            launched at src/patch.cu:5243-5320
 *   vc_lgn_frame_response  the per-frame part of LGN_convol_parvo
 *                            src/discrete_input_convol.cu:1003-1223
 *   vc_lgn_steps           the temporal part of LGN_convol_parvo (taps x
 *                            frames, :1143-1340) fused with LGN_nonlinear and
 *                            get_spike (:1587-1821), for a run of steps
 *   vc_v1_step             compute_V_collect_spike_learnFF_fast
 *                            src/coredynamics.cu:1235-2059 (stages c-FF, d, e)
 *                            fused with recal_G_mat_nd_fast for the previous
 *                            step's spikes, src/coredynamics.cu:2609-2922;
 *                            for a type with tRef < dt the multi-spike
 *                            compute_V_collect_spike_learnFF, :427-1233, and
 *                            its dense gather recal_G_mat_nd, :2364-2607
 *                            (the same sums over the same spikes)
 *   vc_v1_rec_apply        recal_G_mat_nd_fast on its own (the pending
 *                            recurrent increments of the last step)
 *   vc_v1_rec_delay        recal_G_mat (axonal delays, noDelay = 0)
 *                            src/coredynamics.cu:2062-2309
 *   vc_v1_rec_delay_table  its per-synapse delay arithmetic, :2173-2186
 *   vc_lgn_steps_learn     vc_lgn_steps plus the LGN cells' LTP trace of the
 *                            feed-forward plasticity, LGN_nonlinear's
 *                            learning part, src/discrete_input_convol.cu
 *                            :1736-1743, 1775-1821
 *   vc_v1_learn_ff         the feed-forward plasticity of
 *                            compute_V_collect_spike_learnFF(_fast) with
 *                            learning = 3, src/coredynamics.cu:1631-1992
 *                            (and :843-1233, the same code)
 *   vc_v1_rec_distances    the V1_delayMat distances of genCon
 *                            src/preprocess/connect.cu:424-466
 *   vc_v1_rec_far          recal_G_vec (far connections with delays) +
 *                            sum_G, src/coredynamics.cu:181-372, 2314-2349
 *   vc_v1_delay_run        the delayed main loop (step, recal_G_mat,
 *                            recal_G_vec), src/patch.cu:5966-6700
 * The data-structure layouts are documented field by field below.
 */
#ifndef VC_ABI_H
#define VC_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VC_MAX_NTYPE 2      /* src/CONST.h:14  MAX_NTYPE   */
#define VC_MAX_NG    2      /* src/CONST.h:27-30 MAX_NGTYPE_* */

#define VC_OK             0
#define VC_E_INVALID    (-1)   /* bad argument or size           */
#define VC_E_UNSUPPORTED (-2)  /* a reference feature not built  */
#define VC_E_HIP        (-3)   /* a HIP runtime call failed      */

/* ------------------------------------------------------------------ V1 -- */

/* Neuron model and synapse parameters; one copy per simulation (or trial).
 * Types: neuron i of a block has type j = first j with (i % blockSize) <
 * typeAcc[j] (src/coredynamics.cu:1335-1342); E neurons are local ids < nE. */
typedef struct {
    uint32_t nV1, nblock, blockSize, nE;     /* nE = typeAcc[nTypeE-1]          */
    uint32_t nType;
    uint32_t typeAcc[VC_MAX_NTYPE];
    uint32_t ngFF, ngE, ngI;                  /* receptor counts (<= VC_MAX_NG)  */
    int32_t  iModel;                          /* 1: AdEx, w stored; 0: w not kept */
    float dt;                                 /* ms                              */
    /* per type (src/patch.cu:180-374 cfg keys of the same names) */
    float vR[VC_MAX_NTYPE], vThres[VC_MAX_NTYPE], gL[VC_MAX_NTYPE], C[VC_MAX_NTYPE];
    float tRef[VC_MAX_NTYPE], vT[VC_MAX_NTYPE], deltaT[VC_MAX_NTYPE];
    float tau_w[VC_MAX_NTYPE], a[VC_MAX_NTYPE], b[VC_MAX_NTYPE];
    float pFF[VC_MAX_NTYPE * VC_MAX_NG];      /* [type][ig] receptor split       */
    float pE[VC_MAX_NTYPE * VC_MAX_NG];
    float pI[VC_MAX_NTYPE * VC_MAX_NG];
    /* ConductanceShape (src/condShape.h:13-41): dod = rise/(decay-rise) */
    float riseFF[VC_MAX_NG], decayFF[VC_MAX_NG], dodFF[VC_MAX_NG];
    float riseE[VC_MAX_NG],  decayE[VC_MAX_NG],  dodE[VC_MAX_NG];
    float riseI[VC_MAX_NG],  decayI[VC_MAX_NG],  dodI[VC_MAX_NG];
    /* background Poisson drive (extension; not in the reference): per type a
     * rate lambda = rate*dt events per step, p0 = exp(-lambda) (precomputed
     * on the host), strength and receptor split.  bgStrength 0 = off. */
    float bgStrength[VC_MAX_NTYPE], bgLambda[VC_MAX_NTYPE], bgP0[VC_MAX_NTYPE];
    float pBG[VC_MAX_NTYPE * VC_MAX_NG];
    uint32_t seed_lo, seed_hi, trial;         /* Philox key and trial word       */
    /* Batched trials (NULL = one trial): blocks of several independent
     * trials are concatenated; blockTrial[b] is the trial of block b,
     * trialV1Off[t] its first neuron, trialSeed[2t..2t+1] its Philox key.
     * A trial's draws use its own key and local neuron ids, so a trial gives
     * the same result batched as alone. */
    const uint32_t *blockTrial;
    const uint32_t *trialV1Off;
    const uint32_t *trialSeed;
    /* Launch shape of the one-step kernels (vc_v1_step, vc_v1_run,
     * vc_v1_step_snapshot); 0 = default.  stepWG: threads per workgroup (64,
     * 128 or 256, dividing blockSize); gatherTile: masked-CSR strengths in
     * flight per lane in the recurrent gather (4, 8 or 16).  Results do not
     * depend on them (visual-cortex_amd/tune.py picks them by timing). */
    uint32_t stepWG, gatherTile;
    /* Synaptic failure and voltage noise (src/patch.cu:273-280, cfg keys of
     * the same names; 0 = off).  A failing synapse's strength f becomes
     * f * (1 - p + sqrt(p (1 - p) / nSyn) * N(0,1)) each time it transmits
     * (src/coredynamics.cu:1476-1480, 2766-2769): synFailFF/synPerConFF per
     * post type for LGN inputs, synFail/synPerCon [pre type][post type] for
     * V1 inputs.  noisyDep per type: sigma = sqrt(2 noisyDep (vT - vR) r dt)
     * N(0,1), r = the neuron's tonicDep when noiseOnTonic, else 1; white
     * (tauNoise 0: sigma/dt added to both RK2 slopes) or Ornstein-Uhlenbeck
     * (tauNoise > 0, state in vc_v1_state.lastNoise), src/coredynamics.cu
     * :1424-1441, 1549-1554, src/coredynamics.cuh:79-87. */
    float synFailFF[VC_MAX_NTYPE], synPerConFF[VC_MAX_NTYPE];
    float synFail[VC_MAX_NTYPE * VC_MAX_NTYPE], synPerCon[VC_MAX_NTYPE * VC_MAX_NTYPE];
    float noisyDep[VC_MAX_NTYPE];
    float tauNoise;
    uint32_t noiseOnTonic;
} vc_v1_params;

/* Persistent V1 state, all [..][nV1] float32, receptor-major.
 * gapS and gap (may be NULL = zero) are the reference's d_gapS and d_gap
 * over the mI = nblock*nI inhibitory neurons, inhibitory neuron l >= nE of
 * block b at [b*nI + l - nE]: gapS = the sum of its gap-junction strengths
 * (src/patch.cu:3360-3396, 3686-3723), gap = sum over its junctions of
 * strength x the partner's voltage (vThres when it spiked) in the step
 * before (vc_v1_gap).  The step adds gapS to the membrane's a0/a1 and gap
 * to b0/b1 of inhibitory neurons (src/coredynamics.cuh:107-117,
 * src/coredynamics.cu:1344, 1536-1543). */
typedef struct {
    float *v, *w, *tBack;
    float *gFF, *hFF;          /* [ngFF][nV1] */
    float *gE, *hE;            /* [ngE][nV1]  */
    float *gI, *hI;            /* [ngI][nV1]  */
    const float *tonicDep;     /* [nV1]       */
    const float *gapS;         /* [mI] or NULL */
    float *gap;                /* [mI] or NULL */
    float *lastNoise;          /* [nV1] OU noise of the last step (tauNoise > 0),
                                  the reference's last_noise; else NULL */
} vc_v1_state;

/* LGN->V1 synapses, ELL layout transposed like the reference's sLGN
 * (src/patch.h:30-77): entry i of neuron n at [i*nV1 + n]. */
typedef struct {
    const uint32_t *nLGN;      /* [nV1] inputs of each neuron (<= maxLGN)       */
    const float *sLGN;         /* [maxLGN][nV1] strength, already * sRatioLGN   */
    const uint32_t *idLGN;     /* [maxLGN][nV1] LGN cell index                  */
    uint32_t maxLGN;
} vc_ff_conn;

/* Event-driven LGN->V1 delivery (for maxLGN <= 32): the reverse of
 * vc_ff_conn, revTgt[revOff[c] .. revOff[c+1]) = (n << 5) | i for every
 * synapse i of neuron n fed by LGN cell c.  When cell c spikes in step s of
 * a frame period, vc_lgn_steps sets bit i of mask[s*nV1 + n]; vc_v1_step then
 * visits only the set bits of its neuron's word, in synapse order (the
 * reference visits every synapse and skips those with no spike,
 * src/coredynamics.cu:1466-1473), and clears the word for the next period. */
typedef struct {
    const uint32_t *revOff;    /* [nLGN + 1] */
    const uint32_t *revTgt;    /* [revOff[nLGN]] */
    uint32_t *mask;            /* [ntPerFrame][nV1], zero between uses */
    uint32_t nV1;
} vc_ff_events;

/* V1->V1 near-neighbour synapses in "masked CSR": the reference's dense
 * V1_conMat[nblock][maxNear][blockSize pre][blockSize post]
 * (src/patch.cu:3441-3461) with only its strictly positive entries kept.
 * Row r = (b*maxNear + ib)*blockSize + pre, split into blockSize/64 segments
 * of 64 posts (one wavefront each):
 *   rowMask[r*nSeg + s] bit l  <=> conMat[...][pre][s*64 + l] > 0
 *   rowOff [r*nSeg + s]        =   index in w[] of the first kept entry of
 *                                  segment s (entries in post order)           */
typedef struct {
    uint32_t maxNear;           /* nearNeighborBlock                            */
    const uint32_t *nNear;      /* [nblock] near neighbours of each block       */
    const uint32_t *nbr;        /* [nblock][maxNear] neighbour block ids        */
    const uint64_t *rowMask;    /* [nblock*maxNear*blockSize][blockSize/64]     */
    const uint32_t *rowOff;     /* [nblock*maxNear*blockSize][blockSize/64]     */
    const float *w;             /* kept strengths                                */
    /* Optional (NULL = not used): the same strengths in the dense V1_conMat
     * layout [nblock][maxNear][blockSize][blockSize], zeros included.  When
     * given, the forward kernels read a spike's 64 strengths for a wavefront
     * straight from its row (one round trip fewer than rowMask/rowOff/w);
     * the results are identical.  The adjoint always uses w. */
    const float *wDense;
    /* Optional gap junctions between inhibitory neurons (InhGap; gapPre NULL
     * = none): the strictly positive entries of the reference's
     * V1_gapMat[nblock][maxNear][nI pre][nI post] (already * gapRatio,
     * src/patch.cu:3688-3723) as an ELL over the mI inhibitory neurons,
     * junction j of inhibitory neuron i at [j*mI + i], ordered by neighbour
     * slot, then presynaptic index (the reference's summation order,
     * src/coredynamics.cu:2524-2540). */
    uint32_t maxGap;
    const uint32_t *nGap;      /* [mI] junctions of each inhibitory neuron      */
    const uint32_t *gapPre;    /* [maxGap][mI] global id of the partner neuron  */
    const float *gapG;         /* [maxGap][mI] strength                          */
} vc_rec_conn;

/* Spikes of one step ("per-block spike bookkeeping"): one 64-bit spike mask
 * per 64-neuron tile, and the unit conductance increments of each spiking
 * neuron stored at slot tile*64 + rank, rank = number of lower spiking lanes
 * of the tile (so a tile's slots list its spikes in ascending neuron order,
 * E before I).  og/oh are [VC_MAX_NG][nV1]. */
typedef struct {
    uint64_t *mask;             /* [nV1/64] */
    float *og, *oh;             /* [VC_MAX_NG][nV1] */
} vc_spikes;

/* One simulation step on all V1 neurons: apply the recurrent increments of
 * `prev` (the previous step's spikes), decay, LGN + background input, RK2
 * update, spike time, refractory handling, compaction into `next`.  A
 * neuron whose type has tRef < dt may fire several times in the step: after
 * each spike whose refractory period ends inside it the rest of the step is
 * integrated again from vR, and out_sInfo holds nsp + mean(tsp)/dt, its
 * increments in `next` being those of nsp spikes at the mean time
 * (compute_V_collect_spike_learnFF, src/coredynamics.cu:617-836, 1205-1230).
 * tRef must be at least dt/512.
 * ffMask (may be NULL) is this step's row of vc_ff_events.mask: when given,
 * only the synapses whose bit is set are visited and the row is cleared;
 * when NULL every synapse's LGN cell is read.
 * out_sInfo[nV1] receives 1 + tsp/dt for a spike and v otherwise;
 * out_dep[nV1] (may be NULL) receives depC; out_g_prev (may be NULL)
 * receives [2*(ngE+ngI)][nV1] = gE[ngE], gI[ngI], hE[ngE], hI[ngI]: the
 * cortical conductances at the end of the previous step after its recurrent
 * increments, i.e. the reference's gE/gI/hE/hI output of that step. */
int vc_v1_step(void *stream, const vc_v1_params *p, const vc_v1_state *s,
               const vc_ff_conn *ff, const vc_rec_conn *rec,
               const vc_spikes *prev, const vc_spikes *next,
               const float *lgnSInfo, uint32_t *ffMask, uint32_t step,
               float *out_sInfo, float *out_dep, float *out_g_prev);

/* `nsteps` consecutive vc_v1_step launches from the native side: step i
 * reads the spikes of spk2[cur ^ (i&1)], writes spk2[cur ^ ((i+1)&1)], takes
 * its LGN input from lgnSInfo + i*lgnStride (and ffMask + i*maskStride
 * when ffMask is not NULL) and writes out_sInfo +
 * i*outStride (strides in floats; 0 reuses one row).  After the call the
 * last step's spikes are in spk2[cur ^ (nsteps&1)].  (The reference's host
 * loop, src/patch.cu:5966-6913, launched one step at a time.)  With gap
 * junctions (rec->gapPre) each step is followed by vc_v1_gap on its output
 * row, so s->gap holds the last step's gap on return. */
int vc_v1_run(void *stream, const vc_v1_params *p, const vc_v1_state *s,
              const vc_ff_conn *ff, const vc_rec_conn *rec, const vc_spikes *spk2,
              uint32_t cur, const float *lgnSInfo, uint64_t lgnStride,
              uint32_t *ffMask, uint64_t maskStride, uint32_t it0,
              uint32_t nsteps, float *out_sInfo, uint64_t outStride);

/* vc_v1_run in ONE launch: the neurons' state stays in registers
 * across the steps and each workgroup waits only for its neighbour blocks'
 * previous step.  `scratch` (64-byte aligned, vc_v1_persistent_scratch_bytes
 * long) holds the tagged per-tile spike records the workgroups exchange;
 * the records carry `epoch`, which must differ between calls that share a
 * scratch buffer: pass 1, 2, ..., 65535, 1, ... (the call with epoch 1
 * clears the buffer).  Results are identical to
 * vc_v1_run; the last step's spikes are left in spk2 as vc_v1_run leaves them.  Every wait is
 * bounded: on a time-out *status (caller-zeroed, sticky) becomes nonzero
 * and the results of the call are invalid.  Returns VC_E_UNSUPPORTED when
 * the grid (nV1/256 workgroups) cannot be co-resident on the device, or
 * with gap junctions (s->gapS, s->gap or rec->gapPre set); use vc_v1_run
 * then.  A call that returns an error, or has nsteps = 0, launches nothing
 * (the buffer is not cleared): the next call may pass the same epoch.  No counterpart in the reference (its host loop launches
 * each step, src/patch.cu:5966-6913). */
int vc_v1_run_persistent(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                         const vc_ff_conn *ff, const vc_rec_conn *rec, const vc_spikes *spk2,
                         uint32_t cur, const float *lgnSInfo, uint64_t lgnStride,
                         uint32_t *ffMask, uint64_t maskStride, uint32_t it0,
                         uint32_t nsteps, float *out_sInfo, uint64_t outStride,
                         void *scratch, uint32_t epoch, uint32_t *status);

/* Bytes of the scratch buffer of vc_v1_run_persistent. */
uint64_t vc_v1_persistent_scratch_bytes(const vc_v1_params *p);

/* vc_v1_run_persistent without a scratch argument: the call allocates
 * nothing and takes no workspace.  No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * tagged spike records live in a static device array inside the library,
 * one per GPU (8.5 MiB, held until the process ends), which limits nV1 to
 * vc_v1_run_persistent_noalloc_max_v1() = 131072 neurons; larger networks
 * get VC_E_INVALID.  The library cycles the epoch itself, per GPU (the first
 * launching call on a GPU and every 65535th after it clear the array with a
 * kernel; refused calls and calls of 0 steps use up no epoch).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (other streams, other host threads) race and their results are invalid;
 * calls one after the other are safe.  `stream` must belong to the current
 * device.  Inputs, outputs, *status and the return codes are those of
 * vc_v1_run_persistent, whose results it reproduces bit for bit. */
int vc_v1_run_persistent_noalloc(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                                 const vc_ff_conn *ff, const vc_rec_conn *rec,
                                 const vc_spikes *spk2, uint32_t cur, const float *lgnSInfo,
                                 uint64_t lgnStride, uint32_t *ffMask, uint64_t maskStride,
                                 uint32_t it0, uint32_t nsteps, float *out_sInfo,
                                 uint64_t outStride, uint32_t *status);

/* The largest nV1 vc_v1_run_persistent_noalloc accepts. */
uint32_t vc_v1_run_persistent_noalloc_max_v1(void);

/* Apply the recurrent increments of `prev` to gE/hE/gI/hI without stepping
 * (used after the last step, before reading or saving the state).  With
 * synaptic failure (p->synFail) it returns VC_E_UNSUPPORTED: use
 * vc_v1_rec_apply_at. */
int vc_v1_rec_apply(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                    const vc_rec_conn *rec, const vc_spikes *prev);

/* vc_v1_rec_apply for the spikes fired in step `step`: the synaptic-failure
 * draws of a spike are keyed by the step it was fired in, as vc_v1_step
 * keys those of the spikes it applies (its step - 1). */
int vc_v1_rec_apply_at(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                       const vc_rec_conn *rec, const vc_spikes *prev, uint32_t step);

/* vc_v1_gap: s->gap[mI] from rec's gap junctions (rec->gapPre must be set):
 * for inhibitory neuron i, gap = sum over its junctions j, in the ELL's
 * order, of gapG[j] * v_pre, with v_pre = train[slot*nV1 + gapPre[j]]
 * replaced by vThres[partner's type] when it is > 0 (a spike), slot = (cur +
 * depth - lag[j*mI + i]) % depth (lag may be NULL = 0).  train is a ring of
 * `depth` step outputs (out_sInfo), the newest in slot cur; without axonal
 * delays pass the step's out_sInfo, depth 1, cur 0, lag NULL
 * (recal_G_mat_nd, src/coredynamics.cu:2524-2540, 2600); with delays the
 * lag of each junction is ceil((distance/speedOfThought)/dt) of the pair
 * (recal_G_mat, src/coredynamics.cu:2237-2260, 2306).  The reference
 * computes gap after each step's update, for the next step; vc_v1_run does
 * that itself, callers of vc_v1_step call vc_v1_gap after it (and once
 * before the first step, on the initial voltages, as the reference's
 * pre-loop recal_G does, src/patch.cu:5687-5770). */
int vc_v1_gap(void *stream, const vc_v1_params *p, const vc_v1_state *s,
              const vc_rec_conn *rec, const float *train, uint32_t depth, uint32_t cur,
              const uint32_t *lag);

/* depC[nV1] of every neuron from s->tonicDep, as the step kernels compute it
 * (the AdEx constructor's tonic drive, src/coredynamics.cuh:40-49); the
 * reference keeps it as the first array of its V1 state block
 * (src/patch.cu:4001-4019), which its snapshots store (:6897-6898). */
int vc_v1_dep_current(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                      float *out_dep);

/* Initial V1 state: v, w, tBack, g/h as given; clears a spike buffer. */
int vc_spikes_clear(void *stream, uint32_t nV1, const vc_spikes *sp);

/* ---- initial state --------------------------------------------------- */
/* The draws of the reference's initial V1 state, per type (cfg keys of the
 * same names, src/patch.cu:225-229, 311): w0 [type][mean, std] (AdEx),
 * gFF0 [type][ig][g mean, g std/mean, h mean, h std/mean], gE0, gI0 alike;
 * useV0 (use_v0). */
typedef struct {
    uint32_t nV1, nblock, blockSize, nType;
    uint32_t typeAcc[VC_MAX_NTYPE];
    uint32_t ngFF, ngE, ngI;
    int32_t iModel;
    uint32_t useV0;
    float vR[VC_MAX_NTYPE], vT[VC_MAX_NTYPE];
    float w0[VC_MAX_NTYPE * 2];
    float gFF0[VC_MAX_NTYPE * VC_MAX_NG * 4];
    float gE0[VC_MAX_NTYPE * VC_MAX_NG * 4];
    float gI0[VC_MAX_NTYPE * VC_MAX_NG * 4];
} vc_v1_init;

/* The reference's initial V1 state, drawn on the HOST (all pointers are host
 * memory; no device work): src/patch.cu:4405-4570 with the same generator
 * (std::default_random_engine seeded with `seed`, one
 * std::normal_distribution<float> for every truncated-normal draw,
 * std::uniform_real_distribution<float>) in the same order.  w (iModel 1)
 * ~ N(w0); v = vR + (vT - vR) * tonicDep, or uniform in [vR, vT) when
 * useV0 or max tonicDep > 1; gFF/hFF, gE/hE, gI/hI ~ N(mean, mean*relstd)
 * redrawn while negative, gFF = hFF = 0 for a neuron without LGN input.
 * Arrays are [ng][nV1].  The reference seeds this engine with its `seed`
 * after the increments that precede it (seed + 1 + nV1 when the LGN
 * parameters are read from fLGN, src/patch.cu:2686, 4224-4228). */
int vc_v1_initial_state(uint64_t seed, const vc_v1_init *q, const float *tonicDep,
                        const uint32_t *nLGN, float *v, float *w, float *gFF, float *hFF,
                        float *gE, float *hE, float *gI, float *hI);

/* Initial spikes of rand_spInit (src/coredynamics.cu:5-146): per type a
 * chance per step, sp0[type*2 + 0] for a neuron with more than SCsplit LGN
 * inputs and sp0[type*2 + 1] otherwise (the cfg's spE0/spI0 rates times
 * dt/1000, src/patch.cu:811-837).  depthRef is the reference's trainDepth,
 * the number of steps of spikes it generates. */
typedef struct {
    float sp0[VC_MAX_NTYPE * 2];
    uint32_t SCsplit;
    uint32_t depthRef;
} vc_v1_spinit;

/* rand_spInit on the device: a neuron spikes with its chance at a uniform
 * time (reference spikeTrain slot 0), which puts v at vR, w (AdEx) and
 * tBack past the spike; then slots depthRef-1, ..., 1 in that order each
 * hold a spike with chance * (dt - ref)/dt outside the refractory period
 * (the neuron's v otherwise, vR inside it).  Reference slot j holds step
 * j - depthRef of the run (the reference's pre-loop recal_G reads slot
 * depthRef - 1 as the step before the first), written to ring slot (cur0 +
 * depth - depthRef + j) % depth of train [depth][nV1] (cur0: the ring slot
 * of step 0; depth >= depthRef; older slots get v) with its spike mask in
 * trainMask [depth][nV1/64] (may be NULL).  Updates s->v, s->w, s->tBack.
 * Draws: Philox counter (neuron, reference slot, VC_STREAM_V1_INIT, trial)
 * in place of cuRAND's per-neuron stream. */
int vc_v1_initial_spikes(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                         const uint32_t *nLGN, const vc_v1_spinit *q, float *train,
                         uint64_t *trainMask, uint32_t depth, uint32_t cur0);

/* One step's outputs sInfo[nV1] as a spike buffer: each tile's spike mask
 * and every spike's unit increments at slot tile*64 + rank, as the step
 * kernels write them; used for the spikes the first step gathers (the
 * reference's pre-loop recal_G, src/patch.cu:5687-5770). */
int vc_spikes_compact(void *stream, const vc_v1_params *p, const float *sInfo,
                      const vc_spikes *sp);

/* ---- axonal delays (noDelay = 0) ----------------------------------------
 * The reference's recal_G_mat, src/coredynamics.cu:2062-2309 (launched at
 * src/patch.cu:6518-6527 in place of recal_G_mat_nd_fast): each spike reaches
 * its post neuron distance/speedOfThought ms after it was fired.
 *
 * vc_v1_rec_distances: out_dist[i] = the pre-post distance (mm) of masked-CSR
 * strength w[i], from the double positions pos[nV1][2]: x = (float)(x_pre -
 * x_post), d = sqrtf(x*x + y*y) (src/preprocess/connect.cu:424-428, 466: the
 * reference's V1_delayMat, kept here for the connected pairs only). */
int vc_v1_rec_distances(void *stream, const vc_v1_params *p, const vc_rec_conn *rec,
                        const double *pos, float *out_dist);

/* vc_v1_rec_delay_table: one 16-byte record per masked-CSR strength (in w's
 * order), out_syn[4i..4i+3] = {w[i] bits, rem bits, it2post, 0} with
 * time2post = dist[i]/speed, it2post = ceil(time2post/dt), rem = it2post*dt -
 * time2post: the reference's per-step arithmetic of src/coredynamics.cu
 * :2173-2186, done once (speed in mm/ms, dt in ms).  out_syn is 16-byte
 * aligned. */
int vc_v1_rec_delay_table(void *stream, uint64_t n, const float *w, const float *dist,
                          float speed, float dt, uint32_t *out_syn);

/* vc_v1_rec_delay: adds to s->gE/hE/gI/hI the recurrent input that arrives
 * during the step whose outputs are in slot `cur` of the spike-train ring:
 *   train[depth][nV1]        each step's out_sInfo (1 + tsp/dt for a spike, v
 *                            otherwise), step t in slot t % depth (the
 *                            reference's spikeTrain / currentTimeSlot);
 *   trainMask[depth][nV1/64] each step's spike masks (vc_spikes.mask);
 *   syn                      vc_v1_rec_delay_table's records (rec->w itself
 *                            is not read).
 * Only the last `window` slots (ages 0 .. window-1; 0 = depth, at most
 * depth) are scanned for spikes: window must exceed the largest it2post, i.e.
 * be at least the reference's trainDepth = ceil(maxDistance/speed/dt) + 1
 * (src/patch.cu:3813); the ring itself may be longer (the far connections'
 * delays, vc_v1_rec_far).  A synapse whose delay does not fit is skipped and
 * sets *status (may be NULL) to 1.  Use it with vc_v1_step given an empty
 * `prev` (no undelayed increments): step t = vc_v1_step writing slot t %
 * depth, then vc_v1_rec_delay(cur = t % depth), as the reference runs
 * compute_V then recal_G_mat.  Masked CSR only (rec->wDense is ignored);
 * maxNear * blockSize/64 <= 512. */
int vc_v1_rec_delay(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                    const vc_rec_conn *rec, const uint32_t *syn,
                    const float *train, const uint64_t *trainMask, uint32_t depth,
                    uint32_t cur, uint32_t window, uint32_t *status);

/* Far (long-range) connections, the reference's V1_vec lists (src/patch.cu
 * :3308-3352: per post neuron, presynaptic ids, strengths * sRatioV1 and
 * distances in mm), as a transposed ELL: connection j < nFar[n] of post
 * neuron n at [j][nV1], in the file's order. */
typedef struct {
    uint32_t maxFar;            /* max nFar; 0 = no far connections             */
    const uint32_t *nFar;       /* [nV1] */
    const uint32_t *pre;        /* [maxFar][nV1] global presynaptic id          */
    const float *w;             /* [maxFar][nV1] strength, already * sRatioV1   */
    const float *dist;          /* [maxFar][nV1] distance, mm                   */
    const uint64_t *preMask;    /* [nV1/64] bit n%64 of word n/64 set when n is
                                   the presynaptic neuron of a far connection  */
    /* the same connections grouped by presynaptic neuron: connection k of
     * presynaptic neuron q (outOff[q] <= k < outOff[q+1]) is connection
     * outJ[k] of post neuron outPost[k]; every pre[j][n] < nV1 */
    const uint32_t *outOff;     /* [nV1 + 1] */
    const uint32_t *outPost;    /* [outOff[nV1]] */
    const uint32_t *outJ;       /* [outOff[nV1]] */
    uint64_t *hit;              /* [nV1][(maxFar + 63)/64] work bits: zero before
                                   the first call; every call leaves it zero   */
} vc_far_conn;

/* vc_v1_rec_far: adds to s->gE/hE/gI/hI the far connections' input arriving
 * during the step in slot `cur` of the ring vc_v1_rec_delay reads (same
 * train, depth and cur): the reference's recal_G_vec with distance-dependent
 * delays (noFarDelay 0, src/coredynamics.cu:181-372) followed by sum_G
 * (:2314-2349).  With delayed near connections (noDelay 0) call it after
 * vc_v1_rec_delay of the same step, as the reference adds the far sums after
 * recal_G_mat.  With undelayed near connections (noDelay 1, noFarDelay 0) the
 * ring serves the far input alone: step t is vc_v1_step writing ring slot
 * cur (nothing pending in `prev`), then vc_v1_rec_apply_at(step t) of the
 * step's own spikes (recal_G_mat_nd), then this call.  Per connection time2post =
 * dist/speed, it2post = ceil(time2post/dt); depth must exceed every it2post
 * (the reference's fTrainDepth = it2post + 1, src/patch.cu:3847).  As the
 * reference, it adds nothing in a step in which no far presynaptic neuron
 * (far->preMask) has train[cur][n] > 1.0 (src/patch.cu:6597-6598, src/stats.cu
 * :114-126); with batched trials (p->blockTrial) that gate is per trial,
 * over the trial's own far presynaptic neurons, as each trial is a run of
 * its own.  trainMask[depth][nV1/64] holds each step's spike masks, as for
 * vc_v1_rec_delay.  Only the connections of presynaptic neurons that fired
 * inside the ring are visited (far->out*), and far->hit is the calls' work
 * memory (calls on one far->hit must not overlap).  *status (may be NULL)
 * |= 1 when a spiking presynaptic neuron has a connection whose delay does
 * not fit the ring, |= 2 for an out* entry out of range (post >= nV1 or
 * j >= maxFar); such connections are skipped. */
int vc_v1_rec_far(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                  const vc_far_conn *far, const float *train, const uint64_t *trainMask,
                  uint32_t depth, uint32_t cur, float speed, uint32_t *status);

/* vc_v1_delay_run: nsteps steps of the delayed path in one call (the
 * reference's main loop with noDelay = 0, src/patch.cu:5966-6700: compute_V,
 * recal_G_mat, then recal_G_vec + sum_G): step i = it0 + i is vc_v1_step
 * (prev = {zeroMask, og, oh}: nothing pending; next = ring slot (slot0 + i)
 * % depth, whose row of train receives the step's outputs), then
 * vc_v1_rec_delay (depth, window), then vc_v1_rec_far when far is not NULL
 * and has connections.  lgnSInfo + i*lgnStride is step i's LGN output row,
 * ffMask + i*maskStride its event mask (or NULL); out_sInfo + i*outStride
 * receives a copy of step i's outputs when out_sInfo is not NULL.  zeroMask
 * is nV1/64 zero words; og, oh [VC_MAX_NG][nV1] are work rows.  The same
 * launches as the per-step calls, issued from one host call.  syn NULL: the
 * near connections are undelayed (noDelay 1 with delayed far connections,
 * noFarDelay 0) and vc_v1_rec_apply_at of the step's own spikes
 * (recal_G_mat_nd) takes the place of vc_v1_rec_delay (window unused). */
int vc_v1_delay_run(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                    const vc_ff_conn *ff, const vc_rec_conn *rec, const uint32_t *syn,
                    const vc_far_conn *far, float *train, uint64_t *trainMask,
                    uint32_t depth, uint32_t window, uint32_t slot0,
                    const float *lgnSInfo, uint64_t lgnStride, uint32_t *ffMask,
                    uint64_t maskStride, uint32_t it0, uint32_t nsteps, float speed,
                    const uint64_t *zeroMask, float *og, float *oh,
                    float *out_sInfo, uint64_t outStride, uint32_t *status);

/* ---- reverse mode of the V1 step (no counterpart in the reference; the
 * forward it differentiates is vc_v1_step, i.e. compute_V_collect_spike_
 * learnFF_fast + recal_G_mat_nd_fast, src/coredynamics.cu:1235-1629,
 * 2609-2922, for the conductance LIF model). */

/* Rows of a step snapshot, each [nV1]: the state at the start of a step with
 * the cortical conductances after the step's recurrent increments. */
#define VC_SNAP_ROWS 14
#define VC_SNAP_V     0
#define VC_SNAP_TBACK 1
#define VC_SNAP_GFF   2     /* + ig, ig < VC_MAX_NG */
#define VC_SNAP_HFF   4
#define VC_SNAP_GE    6
#define VC_SNAP_HE    8
#define VC_SNAP_GI   10
#define VC_SNAP_HI   12

/* vc_v1_step that also writes the step's snapshot to out_snap
 * [VC_SNAP_ROWS][nV1] (rows past the receptor counts are left untouched). */
int vc_v1_step_snapshot(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                        const vc_ff_conn *ff, const vc_rec_conn *rec,
                        const vc_spikes *prev, const vc_spikes *next,
                        const float *lgnSInfo, uint32_t *ffMask, uint32_t step,
                        float *out_sInfo, float *out_snap);

/* Inputs and accumulators of the reverse sweep over steps it0 .. it0+n-1. */
typedef struct {
    const float *snap;          /* [n][VC_SNAP_ROWS][nV1] snapshots of the steps    */
    const vc_spikes *spk;       /* [n + 1]: spk[k] = spikes of step it0+k-1 (the
                                   input of step it0+k), as vc_v1_step wrote them */
    const float *sInfo;         /* [n][nV1] the forward's output of each step       */
    uint64_t lgnStride;         /* floats between the LGN rows of two steps        */
    const float *lgnSInfo;      /* [n][lgnStride] LGN output of each step          */
    const float *gOut;          /* [n][nV1] dL/dsInfo of each step                 */
    float *lam;                 /* [VC_SNAP_ROWS][nV1] dL/d(state): in, after step
                                   it0+n-1 (zero at the end of the run); out, at the
                                   start of step it0 (before its increments)        */
    float *lamTsp;              /* [2][nV1] dL/d(spike time) accumulators, slot
                                   step & 1; zero before the first call of a run    */
    float *dsLGN;               /* [maxLGN][nV1] += dL/d(sLGN)                      */
    float *dw;                  /* [nnz] += dL/d(w), the order of vc_rec_conn.w     */
} vc_v1_adjoint;

/* Reverse sweep over steps it0+n-1 down to it0 for the conductance LIF
 * model (iModel 0, a = 0, deltaT = 0 for every type; VC_E_UNSUPPORTED
 * otherwise), at the fixed spike pattern of the forward: spike times are
 * differentiated through their interpolation, the refractory back-up and the
 * conductance increments they cause.  s supplies tonicDep and w. */
int vc_v1_adjoint_run(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                      const vc_ff_conn *ff, const vc_rec_conn *rec,
                      const vc_v1_adjoint *a, uint32_t it0, uint32_t n);

/* ----------------------------------------------------------------- LGN -- */

/* Per-cell LGN parameters in the reference's SoA order [region][nLGN]
 * (src/LGN_props.cuh:14-100), region 0 = centre, 1 = surround. Angles in
 * radians; the stimulus is flat (flat_retina = true). */
typedef struct {
    uint32_t nLGN, nSample1D, nKernelSample, kernelSampleInterval,
             iKernelSampleT0, maxSample1D;
    float dt, nsig, normViewDistance, x0, y0;
    const float *polar, *ecc, *rw, *rh, *orient, *k;        /* [2][nLGN] */
    const float *tauR, *tauD, *delay, *ratio, *nR, *nD;     /* [2][nLGN] */
    const float *covariant;                                 /* [nLGN]    */
} vc_lgn_setup;

/* Temporal weights TW[nLGN][2][nKernelSample], shared spatial weights
 * SW[nSample1D^2], sample coordinates SCx/SCy[nLGN][2][nSample1D^2] (texture
 * coordinates normalised to [0,1]), and maxConvol[nLGN]. */
int vc_lgn_precompute(void *stream, const vc_lgn_setup *q, float *TW, float *SW,
                      float *SCx, float *SCy, float *maxConvol);

/* Spatial response of every cell to one frame: R[nLGN][2] = sum over samples
 * of SW * clip(I/mean - 1, +-1), with the mean taken over the surround
 * samples (src/discrete_input_convol.cu:1055-1069, 1143-1211).  frame is
 * [3][height][width] (L, M, S) float32; coneType is [2][nLGN].  With batched
 * trials (cellTrial != NULL) cell c reads frame + cellTrial[c] *
 * trialFrameStride (floats), its trial's copy of the frame. */
int vc_lgn_frame_response(void *stream, const float *frame, uint32_t width,
                          uint32_t height, uint32_t nLGN, uint32_t nSample1D,
                          const float *SCx, const float *SCy, const float *SW,
                          const uint32_t *coneType, const uint32_t *cellTrial,
                          uint64_t trialFrameStride, float *R);

/* The same for cells c0 .. c1-1 only (R of the other cells is not written),
 * so that a caller can spread one frame's spatial work over several calls,
 * e.g. over the steps of the frame period before the frame enters. */
int vc_lgn_frame_response_cells(void *stream, const float *frame, uint32_t width,
                                uint32_t height, uint32_t nLGN, uint32_t c0, uint32_t c1,
                                uint32_t nSample1D, const float *SCx, const float *SCy,
                                const float *SW, const uint32_t *coneType,
                                const uint32_t *cellTrial, uint64_t trialFrameStride, float *R);

/* The per-region sample geometry of every cell, geom[nLGN][2][VC_LGN_GEOM]:
 * ecc cos(polar), ecc sin(polar), cos and sin of the orientation, the
 * sample grid's step and half span in x and in y -- everything the flat-
 * retina texture coordinates of a region's samples need besides the
 * sample's grid position (src/discrete_input_convol.cu:398-417).  64 bytes
 * per cell instead of the 16 bytes per sample of the SCx/SCy tables. */
#define VC_LGN_GEOM 8
int vc_lgn_sample_geometry(void *stream, const vc_lgn_setup *q, float *geom);

/* vc_lgn_frame_response_cells with the coordinates regenerated from geom
 * (and the setup's x0, y0, normViewDistance) by the same operations that fill
 * SCx/SCy: the results are bit-identical, and the coordinate tables are not
 * read.  nSample1D must be 8, 16 or 32. */
int vc_lgn_frame_response_geom(void *stream, const float *frame, uint32_t width,
                               uint32_t height, uint32_t nLGN, uint32_t c0, uint32_t c1,
                               uint32_t nSample1D, const float *geom, float x0, float y0,
                               float normViewDistance, const float *SW,
                               const uint32_t *coneType, const uint32_t *cellTrial,
                               uint64_t trialFrameStride, float *R);

/* The same for nFrames frames [nFrames][3][height][width] in one launch:
 * R[nFrames][nLGN][2]. */
int vc_lgn_frame_responses(void *stream, const float *frames, uint32_t nFrames,
                           uint32_t width, uint32_t height, uint32_t nLGN,
                           uint32_t nSample1D, const float *SCx, const float *SCy,
                           const float *SW, const uint32_t *coneType, float *R);

/* Temporal weights alone (the TW part of vc_lgn_precompute), for callers that
 * change the temporal parameters without re-deriving maxConvol. */
int vc_lgn_temporal_weights(void *stream, const vc_lgn_setup *q, float *TW);

typedef struct {
    uint32_t nLGN, nKernelSample, kernelSampleInterval, iKernelSampleT0;
    uint32_t nRetrace, ntPerFrame, maxFrame;
    float dt, convolRatio, frRatio, tRef;
    uint32_t seed_lo, seed_hi, trial;
    /* Batched trials (NULL = one trial): cellTrial[c], trialLgnOff[t] first
     * cell of trial t, trialSeed[2t..2t+1] its Philox key. */
    const uint32_t *cellTrial;
    const uint32_t *trialLgnOff;
    const uint32_t *trialSeed;
    /* Frames per ntPerFrame steps (find_denorm, src/util/util.h:250-278;
     * 0 is taken as 1): time is counted in units of dt/denorm and frame f
     * covers [f ntPerFrame, (f+1) ntPerFrame) of them, i.e. steps
     * [f ntPerFrame/denorm, (f+1) ntPerFrame/denorm) (src/patch.cu:4999-5030). */
    uint32_t denorm;
} vc_lgn_params;

typedef struct {
    const float *TW;                 /* [nLGN][2][nKernelSample]   */
    const float *kC, *kS;            /* [nLGN] centre / surround k */
    const float *maxConvol;          /* [nLGN]                     */
    const float *c50, *sharpness, *logA, *logB;   /* [nLGN]        */
    float *leftTimeRate, *lastNegLogRand;         /* [nLGN] state  */
} vc_lgn_state;

/* Steps it0 .. it0+nsteps-1 of stages a (temporal part) and b for every LGN
 * cell; the steps must lie within one period of ntPerFrame steps (denorm
 * frames).  Rring is [maxFrame][nLGN][2]: the spatial responses of frame f
 * are in slot f mod maxFrame (pre-stimulus frames f < 0 included); tap i of
 * step it samples frame floor((it + 1 - nRetrace + iKernelSampleT0 + i
 * kernelSampleInterval) denorm / ntPerFrame), which must be in the ring.  Writes
 * sInfo[nsteps][nLGN] (nsp + mean tsp/dt) and fr[nsteps][nLGN] (Hz; required,
 * it is also the hand-over buffer between the rate and spike kernels).  If ev
 * is not NULL, every spike of step it0+s is also delivered as bits of
 * ev->mask[s*nV1 + n] (see vc_ff_events). */
int vc_lgn_steps(void *stream, const vc_lgn_params *p, const vc_lgn_state *s,
                 const float *Rring, uint32_t it0, uint32_t nsteps,
                 float *sInfo, float *fr, const vc_ff_events *ev);

/* Plasticity of the LGN->V1 strengths: the reference's learning = 3, the
 * triplet rule on the feed-forward synapses onto excitatory neurons, with
 * optional homeostasis of each neuron's total feed-forward strength
 * (src/coredynamics.cu:1631-1992, src/discrete_input_convol.cu:1736-1821,
 * the constants of learnFF_pre/learnFF_post, src/condShape.h:94-155, set up
 * at src/patch.cu:1092-1330, 4332-4352).  The traces decay as
 * x *= exp(-t/tau) (decay/if_decay, src/condShape.h:44-51). */
typedef struct {
    float tauLTP;            /* LGN (pre) trace, ms: lFF_E_pre.tauLTP[0]        */
    float tauLTD, tauTrip;   /* V1 (post) traces, ms: lFF_E_post.tau[0], [1]    */
    float tauAvg;            /* filtered post spike count, ms: lFF_E_post.tau[2] */
    float A_LTP;             /* A_LGN[0] * sRatioLGN[0]                          */
    float A_ratio;           /* the LTD amplitude (learnFF_post's A_ratio)       */
    float targetFR;          /* Hz; > 0: the LTD amplitude is scaled by
                                tauLTP * avg * avg at the pre spike              */
    float gmax, gmin;        /* strength bounds, already * sRatioLGN[0]          */
    uint32_t applyHomeo;     /* 0 none, 1 multiplicative, 2 additive             */
    uint32_t symmetricHomeo; /* relax towards the target from below too          */
    float expHomeo;          /* exp(-dt/decayTimeHomeoFF)                        */
    uint32_t rebound;        /* 0: a strength of exactly 0 is left alone         */
} vc_learn_ff_params;

/* The plastic state, all [nV1] float32 (the reference's vLTD_FF_E,
 * vTrip_FF_E and vAvgE over [nblock][nE], here indexed by global neuron id:
 * inhibitory entries are neither read nor written). */
typedef struct {
    float *sLGN;             /* [maxLGN][nV1]: the strengths vc_ff_conn points
                                to, updated in place                            */
    float *vLTD, *vTrip;     /* post traces, zero at the start of a run          */
    float *vAvg;             /* filtered post spike count, zero at the start     */
    float *totalFF;          /* applyHomeo: total strength of each neuron (the
                                reference's initial sum over its list, in list
                                order, src/patch.cu:4339-4344); else NULL        */
    const float *totalFFinf; /* applyHomeo: its target, FF_InfRatio * the initial
                                total; else NULL                                */
} vc_learn_ff_state;

/* vc_lgn_steps, and the LTP trace x of each LGN cell (the reference's
 * surface layers 1..3 of LGNspikeSurface, src/discrete_input_convol.cu
 * :1736-1743, 1775-1821): xVar[nLGN] holds x at the start of step it0 on
 * entry (0 before the first step, logRand_init) and at the end of the last
 * step on exit; for step it0+s, xStart[s*nLGN + c] receives x at the start of
 * the step and xSpike[s*nLGN + c] x at the cell's spike time: the reference
 * adds the step's nsp spikes first, then decays to the mean spike time tsp
 * (x_spike = (x + nsp) exp(-tsp/tauLTP)), and the step ends at x_spike
 * exp(-(dt - tsp)/tauLTP); without a spike xSpike = xStart and the step ends
 * at x exp(-dt/tauLTP).  tsp is the draw's own spike time, not the one
 * recovered from sInfo.  spikeBits[s*nWords + c/64] bit c%64 (nWords =
 * (nLGN + 63)/64) is set when cell c spiked in step it0+s, the bitmap
 * vc_v1_learn_ff tests before it reads a cell's row.  Batched trials learn
 * as each alone (every quantity is per cell). */
int vc_lgn_steps_learn(void *stream, const vc_lgn_params *p, const vc_lgn_state *s,
                       const float *Rring, uint32_t it0, uint32_t nsteps, float *sInfo,
                       float *fr, const vc_ff_events *ev, float tauLTP, float *xVar,
                       float *xStart, float *xSpike, uint64_t *spikeBits);

/* One step of the feed-forward plasticity, after that step's vc_v1_step (the
 * reference learns at the end of the same kernel, after the strengths were
 * used): for each excitatory neuron n, with nsp = floor(sInfo[n]) and tsp =
 * (sInfo[n] - nsp) dt its spikes of the step (tsp = dt without), every
 * synapse i < nLGN[n] whose strength f is not 0 (or every one with rebound):
 * f is first scaled (applyHomeo 1) or shifted (2) by the homeostatic change
 * of the neuron's total; if its LGN cell c spiked (lgnSInfo[c] >= 1, at
 * tspF) f -= the LTD term ltd(t) * exp(-|tspF - t|/tauLTD) * A (t = the step
 * start when tspF < tsp, else tsp); if the neuron spiked, f += the LTP term
 * x_pre * exp(-dtp/tauLTP) * trip(tsp) * A_LTP (x_pre = xSpike[c], dtp = tsp
 * - tspF when tspF < tsp, else xStart[c], dtp = tsp); f is clipped to
 * [gmin, gmax].  Then the post traces take the step's spikes and decay to the
 * step end, in the reference's operation order.  sInfo is the step's V1
 * output row, lgnSInfo/xStart/xSpike/lgnSpikeBits the step's rows of
 * vc_lgn_steps_learn over nLGN cells.  Batched trials (p->blockTrial) learn
 * as each alone: a neuron's synapses and LGN cells lie in its trial, and the
 * rule's constants are shared. */
int vc_v1_learn_ff(void *stream, const vc_v1_params *p, const vc_learn_ff_params *lp,
                   const vc_ff_conn *ff, const vc_learn_ff_state *ls, const float *sInfo,
                   const float *lgnSInfo, const float *xStart, const float *xSpike,
                   const uint64_t *lgnSpikeBits, uint32_t nLGN);

/* The rate part of vc_lgn_steps alone (no spike draw), for any run of steps
 * and any maxFrame: fr[nsteps][nLGN] (Hz).  R is read at slot f mod maxFrame
 * for frame f; a caller holding every frame of a run passes maxFrame = the
 * number of frames held. */
int vc_lgn_rates(void *stream, const vc_lgn_params *p, const vc_lgn_state *s,
                 const float *R, uint32_t it0, uint32_t nsteps, float *fr);

/* ---- reverse mode of the LGN rates (no counterpart in the reference; the
 * forward it differentiates is vc_lgn_frame_responses + vc_lgn_rates, i.e.
 * LGN_convol_parvo and the rate of LGN_nonlinear,
 * src/discrete_input_convol.cu:1003-1340, 1587-1622).  maxConvol is held
 * constant.  Every output is ACCUMULATED into (the caller zeroes it first). */

/* Given gFr[nsteps][nLGN] = dL/dfr of vc_lgn_rates(p, s, R, it0, nsteps):
 *   dR[maxFrame][nLGN][2]   += dL/dR     (same slots as R; maxFrame must hold
 *                                         every frame the run touches)
 *   dTW[nLGN][2][nKS]       += dL/dTW
 *   dK[nLGN][2]             += dL/d(kC, kS) */
int vc_lgn_rates_adjoint(void *stream, const vc_lgn_params *p, const vc_lgn_state *s,
                         const float *R, uint32_t it0, uint32_t nsteps, const float *gFr,
                         float *dR, float *dTW, float *dK);

/* Given dR[nFrames][nLGN][2] for the frames of vc_lgn_frame_responses:
 * dFrames[nFrames][3][height][width] += dL/dframes (float atomics: the order
 * of the additions, hence the last bits, varies from run to run). */
int vc_lgn_frame_responses_adjoint(void *stream, const float *frames, uint32_t nFrames,
                                   uint32_t width, uint32_t height, uint32_t nLGN,
                                   uint32_t nSample1D, const float *SCx, const float *SCy,
                                   const float *SW, const uint32_t *coneType,
                                   const float *dR, float *dFrames);

/* Given dTW[nLGN][2][nKS] for the weights of vc_lgn_temporal_weights(q):
 * dTauR[2][nLGN], dTauD[2][nLGN] = dL/d(tauR, tauD) (written, not
 * accumulated), through the normalisation by sum |TW|. */
int vc_lgn_temporal_adjoint(void *stream, const vc_lgn_setup *q, const float *dTW,
                            float *dTauR, float *dTauD);

/* LGN Poisson state initialisation (logRand_init, src/coredynamics.cu:148-177):
 * leftTimeRate = 0, lastNegLogRand = -log(u) with u a Philox draw. */
int vc_lgn_init(void *stream, const vc_lgn_params *p, float *leftTimeRate,
                float *lastNegLogRand);

const char *vc_last_error(void);
int vc_version(void);

#ifdef __cplusplus
}
#endif
#endif /* VC_ABI_H */
