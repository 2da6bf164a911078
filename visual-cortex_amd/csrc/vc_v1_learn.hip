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
// when it changed; the LGN cell's sInfo and trace values are gathers from
// arrays of nLGN floats that stay in L2.  Lanes are neurons, so every
// [i][nV1] access of a wavefront is one contiguous 256-byte segment.
#include "vc_common.h"

namespace {

// decay / if_decay, src/condShape.h:44-51: x * exp(-t/tau)
__device__ __forceinline__ float decayed(float x, float tau, float t) {
    return x * vc_expf(-t / tau);
}

__global__ __launch_bounds__(256) void k_v1_learn_ff(
    vc_v1_params p, vc_learn_ff_params lp, vc_ff_conn ff, vc_learn_ff_state ls,
    const float *__restrict__ sInfo, const float *__restrict__ lgnS,
    const float *__restrict__ xStart, const float *__restrict__ xSpike) {
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
    for (uint32_t i = 0; i < m; ++i) {                      // :1803-1927
        const size_t lid = (size_t)i * p.nV1 + n;
        const float f0 = ls.sLGN[lid];
        if (f0 == 0.0f && !lp.rebound) continue;            // pruned
        float f = f0;
        if (lp.applyHomeo == 1u) f *= change;
        else if (lp.applyHomeo == 2u) f += change;
        const uint32_t c = ff.idLGN[lid];
        const float sF = lgnS[c];
        const float nspF = floorf(sF);
        const float tspF = (sF > 0.0f ? sF - nspF : 1.0f) * dt;
        const bool early = tspF < tsp;                      // pre spike before the post one
        if (nspF > 0.0f) {                                  // LTD, with or without a post spike
            const float av = early ? avg0 : avg1;
            float A = lp.A_ratio;
            if (lp.targetFR > 0.0f) A *= lp.tauLTP * av * av;
            const float df = decayed(early ? ltd0 : ltd1, lp.tauLTD, early ? tspF : tspF - tsp) * A;
            f -= df;
            if (lp.applyHomeo) deltaF -= df;
        }
        if (nsp > 0.0f) {                                   // LTP, with or without a pre spike
            const float xpre = early ? xSpike[c] : xStart[c];
            const float df = decayed(xpre, lp.tauLTP, early ? tsp - tspF : tsp) * trip1 * lp.A_LTP;
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
        if (__float_as_uint(f) != __float_as_uint(f0)) ls.sLGN[lid] = f;
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

}  // namespace

extern "C" int vc_v1_learn_ff(void *stream, const vc_v1_params *p, const vc_learn_ff_params *lp,
                              const vc_ff_conn *ff, const vc_learn_ff_state *ls,
                              const float *sInfo, const float *lgnSInfo, const float *xStart,
                              const float *xSpike) {
    if (!p || !lp || !ff || !ls || !sInfo || !lgnSInfo || !xStart || !xSpike) {
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
    if (p->blockTrial) {
        vc_set_error("vc_v1_learn_ff: plasticity with batched trials is not built");
        return VC_E_UNSUPPORTED;
    }
    if (p->nV1 == 0 || p->blockSize == 0 || p->nV1 != p->nblock * p->blockSize) {
        vc_set_error("vc_v1_learn_ff: nV1 %u != nblock %u x blockSize %u", p->nV1, p->nblock,
                     p->blockSize);
        return VC_E_INVALID;
    }
    hipLaunchKernelGGL(k_v1_learn_ff, dim3((p->nV1 + 255u) / 256u), dim3(256), 0,
                       (hipStream_t)stream, *p, *lp, *ff, *ls, sInfo, lgnSInfo, xStart, xSpike);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}
