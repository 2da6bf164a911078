// vc_v1_learn.hip -- plasticity of the LGN->V1 strengths (the reference's
// learning = 3: the triplet rule on the feed-forward synapses onto excitatory
// neurons, with optional homeostasis of each neuron's total strength).
//
// k_v1_learn_ff runs after a step's k_v1_step, one lane per V1 neuron
// (inhibitory lanes exit), and restates the learning part of
// compute_V_collect_spike_learnFF_fast (src/coredynamics.cu:1631-1992) with
// learning = 3, where only the FF_E rule is active (src/patch.cu:1166-1171):
// the post traces and the filtered rate of the neuron decayed to its spike
// time, the homeostatic change of its total, then for each of its synapses
// LTD at the LGN cell's spike and LTP at its own spike, the clip to
// [gmin, gmax], and the post traces carried to the end of the step.  The
// operations and their order are the reference's, in float32.
//
// Traffic per excitatory neuron-step: its sInfo, three traces (read and
// written) and, with homeostasis, its total and target (about 36 B), and per
// synapse the strength and LGN id (8 B read) plus the strength written back
// when it changed.  Lanes are neurons, so every [i][nV1] access of a
// wavefront is one contiguous 256-byte segment.  Whether a synapse's LGN
// cell spiked is a bit test in the step's spike bitmap, staged in LDS (2 KB
// for 16,384 cells): a per-lane gather of the cell's sInfo would touch one
// cache line per lane per synapse, which measured 20-27 us per step at the
// full patch; the cell's sInfo and trace rows are read only for a spiking
// cell or a spiking neuron.
#include "vc_common.h"

namespace {

// decay / if_decay, src/condShape.h:44-51: x * exp(-t/tau)
__device__ __forceinline__ float decayed(float x, float tau, float t) {
    return x * vc_expf(-t / tau);
}

constexpr uint32_t kMaxLdsWords = 4096;    // bitmaps of up to 262,144 cells in LDS

template <bool kLds>
__global__ __launch_bounds__(256) void k_v1_learn_ff(
    vc_v1_params p, vc_learn_ff_params lp, vc_ff_conn ff, vc_learn_ff_state ls,
    const float *__restrict__ sInfo, const float *__restrict__ lgnS,
    const float *__restrict__ xStart, const float *__restrict__ xSpike,
    const uint64_t *__restrict__ bitsG, uint32_t nWords) {
    __shared__ uint64_t bitsL[kLds ? kMaxLdsWords : 1];
    if (kLds) {
        for (uint32_t w = threadIdx.x; w < nWords; w += blockDim.x) bitsL[w] = bitsG[w];
        __syncthreads();
    }
    const uint64_t *bits = kLds ? bitsL : bitsG;
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= p.nV1 || n % p.blockSize >= p.nE) return;
    const float dt = p.dt;
    const float si = sInfo[n];
    float nsp, tsp;
    if (si > 0.0f) {                                        // :1632-1639
        nsp = floorf(si);
        tsp = (si - nsp) * dt;
    } else {
        nsp = 0.0f;
        tsp = dt;
    }
    // post traces at the step start (row 0) and at the spike (row 1)
    const float ltd0 = ls.vLTD[n], trip0 = ls.vTrip[n], avg0 = ls.vAvg[n];
    float ltd1 = 0.0f, trip1 = 0.0f, avg1 = 0.0f;
    if (nsp > 0.0f) {                                       // :1694-1727
        ltd1 = decayed(ltd0, lp.tauLTD, tsp);
        trip1 = decayed(trip0, lp.tauTrip, tsp);
        avg1 = decayed(avg0, lp.tauAvg, tsp);
    }
    const uint32_t m = ff.nLGN[n];
    // homeostasis of the total strength, :1765-1800
    float total0 = 0.0f, newTotal = 0.0f, change = 0.0f, deltaF = 0.0f;
    if (lp.applyHomeo) {
        total0 = ls.totalFF[n];
        const float inf = ls.totalFFinf[n];
        const float d = total0 - inf;
        if (d > 0.0f || lp.symmetricHomeo) {
            newTotal = d * lp.expHomeo + inf;
            change = (lp.applyHomeo == 1u) ? newTotal / total0 : (newTotal - total0) / (float)m;
        } else {
            newTotal = total0;
            change = (lp.applyHomeo == 1u) ? 1.0f : 0.0f;
        }
    }
    // kChunk synapses at a time: their strengths and LGN ids are loaded
    // together, then processed in list order
    constexpr uint32_t kChunk = 8;
    for (uint32_t i0 = 0; i0 < m; i0 += kChunk) {          // :1803-1927
        float fs[kChunk], sFs[kChunk], xp[kChunk];
        uint32_t cs[kChunk];
#pragma unroll
        for (uint32_t k = 0; k < kChunk; ++k) {
            const bool in = i0 + k < m;
            const size_t lid = (size_t)(in ? i0 + k : 0u) * p.nV1 + n;
            fs[k] = in ? ls.sLGN[lid] : 0.0f;
            cs[k] = in ? ff.idLGN[lid] : 0u;
        }
        // a cell that did not spike has sInfo exactly 0 (vc_lgn_steps)
#pragma unroll
        for (uint32_t k = 0; k < kChunk; ++k) {
            const uint32_t c = cs[k];
            const bool spk = (i0 + k < m) && ((bits[c >> 6] >> (c & 63u)) & 1ull);
            sFs[k] = spk ? lgnS[c] : 0.0f;
        }
        if (nsp > 0.0f) {                                   // the pre traces LTP reads
#pragma unroll
            for (uint32_t k = 0; k < kChunk; ++k) {
                const float sF = sFs[k];
                const float tspF = (sF > 0.0f ? sF - floorf(sF) : 1.0f) * dt;
                xp[k] = (i0 + k < m) ? (tspF < tsp ? xSpike[cs[k]] : xStart[cs[k]]) : 0.0f;
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kChunk; ++k) {
            if (i0 + k >= m) break;
            const float f0 = fs[k];
            if (f0 == 0.0f && !lp.rebound) continue;        // pruned
            float f = f0;
            if (lp.applyHomeo == 1u) f *= change;
            else if (lp.applyHomeo == 2u) f += change;
            const float sF = sFs[k];
            const float nspF = floorf(sF);
            const float tspF = (sF > 0.0f ? sF - nspF : 1.0f) * dt;
            const bool early = tspF < tsp;                  // pre spike before the post one
            if (nspF > 0.0f) {                              // LTD, with or without a post spike
                const float av = early ? avg0 : avg1;
                float A = lp.A_ratio;
                if (lp.targetFR > 0.0f) A *= lp.tauLTP * av * av;
                const float df =
                    decayed(early ? ltd0 : ltd1, lp.tauLTD, early ? tspF : tspF - tsp) * A;
                f -= df;
                if (lp.applyHomeo) deltaF -= df;
            }
            if (nsp > 0.0f) {                               // LTP, with or without a pre spike
                const float df =
                    decayed(xp[k], lp.tauLTP, early ? tsp - tspF : tsp) * trip1 * lp.A_LTP;
                if (lp.applyHomeo) deltaF += df;
                f += df;
            }
            if (f < lp.gmin) {
                if (lp.applyHomeo) deltaF += lp.gmin - f;
                f = lp.gmin;
            }
            if (f > lp.gmax) {
                if (lp.applyHomeo) deltaF -= f - lp.gmax;
                f = lp.gmax;
            }
            if (__float_as_uint(f) != __float_as_uint(f0))
                ls.sLGN[(size_t)(i0 + k) * p.nV1 + n] = f;
        }
    }
    if (lp.applyHomeo) ls.totalFF[n] = newTotal + deltaF;
    // the post traces take the step's spikes and decay to its end, :1930-1960
    const float rest = (nsp > 0.0f) ? dt - tsp : dt;
    float ltd = (nsp > 0.0f) ? ltd1 : ltd0;
    float trip = (nsp > 0.0f) ? trip1 : trip0;
    float avg = (nsp > 0.0f) ? avg1 : avg0;
    ltd += nsp;
    trip += nsp;
    ltd = decayed(ltd, lp.tauLTD, rest);
    trip = decayed(trip, lp.tauTrip, rest);
    avg += nsp;
    avg = decayed(avg, lp.tauAvg, rest);
    ls.vLTD[n] = ltd;
    ls.vTrip[n] = trip;
    ls.vAvg[n] = avg;
}

// Four wavefronts share each 64 neurons: wavefront w takes synapses w, w+4,
// w+8, ... (at most 8 for m <= 32, loaded together).  Four times the
// wavefronts of k_v1_learn_ff hide the load latency that one wavefront per
// SIMD cannot.  Wavefront 0 stores the post traces after a barrier (every
// wavefront has read them by then).  With homeostasis (kHomeo, maxLGN <= 32)
// each synapse's changes to the running total (-LTD, +LTP, the clip) go to
// LDS, and wavefront 0 adds them in synapse order, as the reference's
// sequential loop does, so the total is the same float.
constexpr uint32_t kHomeoSyn = 32;

template <bool kLds, bool kHomeo>
__global__ __launch_bounds__(256) void k_v1_learn_ff_split(
    vc_v1_params p, vc_learn_ff_params lp, vc_ff_conn ff, vc_learn_ff_state ls,
    const float *__restrict__ sInfo, const float *__restrict__ lgnS,
    const float *__restrict__ xStart, const float *__restrict__ xSpike,
    const uint64_t *__restrict__ bitsG, uint32_t nWords) {
    __shared__ uint64_t bitsL[kLds ? kMaxLdsWords : 1];
    __shared__ float terms[kHomeo ? kHomeoSyn * 3 * 64 : 1];   // [i][3][lane]
    if (kLds) {
        for (uint32_t w = threadIdx.x; w < nWords; w += blockDim.x) bitsL[w] = bitsG[w];
        __syncthreads();
    }
    const uint64_t *bits = kLds ? bitsL : bitsG;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t n = blockIdx.x * 64u + (threadIdx.x & 63u);
    const bool mine = n < p.nV1 && n % p.blockSize < p.nE;
    const float dt = p.dt;
    float nsp = 0.0f, tsp = dt, ltd0 = 0.0f, trip0 = 0.0f, avg0 = 0.0f;
    float ltd1 = 0.0f, trip1 = 0.0f, avg1 = 0.0f;
    uint32_t m = 0;
    if (mine) {
        const float si = sInfo[n];
        if (si > 0.0f) {
            nsp = floorf(si);
            tsp = (si - nsp) * dt;
        }
        ltd0 = ls.vLTD[n];
        trip0 = ls.vTrip[n];
        avg0 = ls.vAvg[n];
        if (nsp > 0.0f) {
            ltd1 = decayed(ltd0, lp.tauLTD, tsp);
            trip1 = decayed(trip0, lp.tauTrip, tsp);
            avg1 = decayed(avg0, lp.tauAvg, tsp);
        }
        m = ff.nLGN[n];
    }
    float newTotal = 0.0f, change = 0.0f;
    if (kHomeo && mine) {                                   // as k_v1_learn_ff
        const float total0 = ls.totalFF[n];
        const float inf = ls.totalFFinf[n];
        const float d = total0 - inf;
        if (d > 0.0f || lp.symmetricHomeo) {
            newTotal = d * lp.expHomeo + inf;
            change = (lp.applyHomeo == 1u) ? newTotal / total0 : (newTotal - total0) / (float)m;
        } else {
            newTotal = total0;
            change = (lp.applyHomeo == 1u) ? 1.0f : 0.0f;
        }
    }
    constexpr uint32_t kPer = 8;
    for (uint32_t j0 = 0; wave + 4u * j0 < m; j0 += kPer) {
        float fs[kPer], sFs[kPer], xp[kPer];
        uint32_t cs[kPer];
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t i = wave + 4u * (j0 + k);
            const bool in = i < m;
            const size_t lid = (size_t)(in ? i : 0u) * p.nV1 + n;
            fs[k] = in ? ls.sLGN[lid] : 0.0f;
            cs[k] = in ? ff.idLGN[lid] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t c = cs[k];
            const bool spk = (wave + 4u * (j0 + k) < m) && ((bits[c >> 6] >> (c & 63u)) & 1ull);
            sFs[k] = spk ? lgnS[c] : 0.0f;
        }
        if (nsp > 0.0f) {
#pragma unroll
            for (uint32_t k = 0; k < kPer; ++k) {
                const float sF = sFs[k];
                const float tspF = (sF > 0.0f ? sF - floorf(sF) : 1.0f) * dt;
                xp[k] = (wave + 4u * (j0 + k) < m)
                            ? (tspF < tsp ? xSpike[cs[k]] : xStart[cs[k]]) : 0.0f;
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t i = wave + 4u * (j0 + k);
            if (i >= m) break;
            const float f0 = fs[k];
            float tL = 0.0f, tP = 0.0f, tC = 0.0f;          // changes to the running total
            if (f0 == 0.0f && !lp.rebound) {
                if (kHomeo) {
                    terms[(i * 3 + 0) * 64 + lane] = 0.0f;
                    terms[(i * 3 + 1) * 64 + lane] = 0.0f;
                    terms[(i * 3 + 2) * 64 + lane] = 0.0f;
                }
                continue;
            }
            float f = f0;
            if (kHomeo) {
                if (lp.applyHomeo == 1u) f *= change;
                else f += change;
            }
            const float sF = sFs[k];
            const float nspF = floorf(sF);
            const float tspF = (sF > 0.0f ? sF - nspF : 1.0f) * dt;
            const bool early = tspF < tsp;
            if (nspF > 0.0f) {
                const float av = early ? avg0 : avg1;
                float A = lp.A_ratio;
                if (lp.targetFR > 0.0f) A *= lp.tauLTP * av * av;
                const float df = decayed(early ? ltd0 : ltd1, lp.tauLTD, early ? tspF : tspF - tsp) * A;
                f -= df;
                tL = -df;
            }
            if (nsp > 0.0f) {
                const float df = decayed(xp[k], lp.tauLTP, early ? tsp - tspF : tsp) * trip1 * lp.A_LTP;
                f += df;
                tP = df;
            }
            if (f < lp.gmin) {
                tC = lp.gmin - f;
                f = lp.gmin;
            }
            if (f > lp.gmax) {
                tC = -(f - lp.gmax);
                f = lp.gmax;
            }
            if (kHomeo) {
                terms[(i * 3 + 0) * 64 + lane] = tL;
                terms[(i * 3 + 1) * 64 + lane] = tP;
                terms[(i * 3 + 2) * 64 + lane] = tC;
            }
            if (__float_as_uint(f) != __float_as_uint(f0)) ls.sLGN[(size_t)i * p.nV1 + n] = f;
        }
    }
    __syncthreads();
    if (wave != 0 || !mine) return;
    if (kHomeo) {                          // the reference's order: synapse by synapse
        float deltaF = 0.0f;
        for (uint32_t i = 0; i < m; ++i) {
            deltaF += terms[(i * 3 + 0) * 64 + lane];
            deltaF += terms[(i * 3 + 1) * 64 + lane];
            deltaF += terms[(i * 3 + 2) * 64 + lane];
        }
        ls.totalFF[n] = newTotal + deltaF;
    }
    const float rest = (nsp > 0.0f) ? dt - tsp : dt;
    float ltd = (nsp > 0.0f) ? ltd1 : ltd0;
    float trip = (nsp > 0.0f) ? trip1 : trip0;
    float avg = (nsp > 0.0f) ? avg1 : avg0;
    ltd += nsp;
    trip += nsp;
    ltd = decayed(ltd, lp.tauLTD, rest);
    trip = decayed(trip, lp.tauTrip, rest);
    avg += nsp;
    avg = decayed(avg, lp.tauAvg, rest);
    ls.vLTD[n] = ltd;
    ls.vTrip[n] = trip;
    ls.vAvg[n] = avg;
}

}  // namespace

extern "C" int vc_v1_learn_ff(void *stream, const vc_v1_params *p, const vc_learn_ff_params *lp,
                              const vc_ff_conn *ff, const vc_learn_ff_state *ls,
                              const float *sInfo, const float *lgnSInfo, const float *xStart,
                              const float *xSpike, const uint64_t *lgnSpikeBits,
                              uint32_t nLGN) {
    if (!p || !lp || !ff || !ls || !sInfo || !lgnSInfo || !xStart || !xSpike || !lgnSpikeBits) {
        vc_set_error("vc_v1_learn_ff: null argument");
        return VC_E_INVALID;
    }
    if (!ff->nLGN || !ff->idLGN || !ls->sLGN || !ls->vLTD || !ls->vTrip || !ls->vAvg) {
        vc_set_error("vc_v1_learn_ff: null connection or trace array");
        return VC_E_INVALID;
    }
    if (ls->sLGN != ff->sLGN) {
        vc_set_error("vc_v1_learn_ff: ls->sLGN must be the strengths ff->sLGN points to");
        return VC_E_INVALID;
    }
    if (lp->applyHomeo > 2u) {
        vc_set_error("vc_v1_learn_ff: applyHomeo %u (0, 1 or 2)", lp->applyHomeo);
        return VC_E_INVALID;
    }
    if (lp->applyHomeo && (!ls->totalFF || !ls->totalFFinf)) {
        vc_set_error("vc_v1_learn_ff: homeostasis needs totalFF and totalFFinf");
        return VC_E_INVALID;
    }
    if (!(lp->tauLTP > 0.0f && lp->tauLTD > 0.0f && lp->tauTrip > 0.0f && lp->tauAvg > 0.0f)) {
        vc_set_error("vc_v1_learn_ff: the trace time constants must be positive");
        return VC_E_INVALID;
    }
    if (p->nV1 == 0 || p->blockSize == 0 || p->nV1 != p->nblock * p->blockSize) {
        vc_set_error("vc_v1_learn_ff: nV1 %u != nblock %u x blockSize %u", p->nV1, p->nblock,
                     p->blockSize);
        return VC_E_INVALID;
    }
    const uint32_t nWords = (nLGN + 63u) / 64u;
    const dim3 grid((p->nV1 + 63u) / 64u);
    const bool lds = nWords <= kMaxLdsWords;
    if (lp->applyHomeo == 0u) {           // 4 wavefronts per 64 neurons
        if (lds)
            hipLaunchKernelGGL((k_v1_learn_ff_split<true, false>), grid, dim3(256), 0,
                               (hipStream_t)stream, *p, *lp, *ff, *ls, sInfo, lgnSInfo, xStart,
                               xSpike, lgnSpikeBits, nWords);
        else
            hipLaunchKernelGGL((k_v1_learn_ff_split<false, false>), grid, dim3(256), 0,
                               (hipStream_t)stream, *p, *lp, *ff, *ls, sInfo, lgnSInfo, xStart,
                               xSpike, lgnSpikeBits, nWords);
    } else if (ff->maxLGN <= kHomeoSyn) {  // the same, the total's terms added in order
        if (lds)
            hipLaunchKernelGGL((k_v1_learn_ff_split<true, true>), grid, dim3(256), 0,
                               (hipStream_t)stream, *p, *lp, *ff, *ls, sInfo, lgnSInfo, xStart,
                               xSpike, lgnSpikeBits, nWords);
        else
            hipLaunchKernelGGL((k_v1_learn_ff_split<false, true>), grid, dim3(256), 0,
                               (hipStream_t)stream, *p, *lp, *ff, *ls, sInfo, lgnSInfo, xStart,
                               xSpike, lgnSpikeBits, nWords);
    } else if (lds)
        hipLaunchKernelGGL(k_v1_learn_ff<true>, dim3((p->nV1 + 255u) / 256u), dim3(256), 0,
                           (hipStream_t)stream, *p, *lp, *ff, *ls, sInfo, lgnSInfo, xStart,
                           xSpike, lgnSpikeBits, nWords);
    else
        hipLaunchKernelGGL(k_v1_learn_ff<false>, dim3((p->nV1 + 255u) / 256u), dim3(256), 0,
                           (hipStream_t)stream, *p, *lp, *ff, *ls, sInfo, lgnSInfo, xStart,
                           xSpike, lgnSpikeBits, nWords);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}
