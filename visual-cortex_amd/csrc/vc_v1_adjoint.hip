// vc_v1_adjoint.hip -- reverse sweep of the V1 step (vc_v1.hip) for the
// conductance LIF model (iModel 0, a = 0, deltaT = 0), at a fixed spike
// pattern.
//
// The forward step of neuron n at step `it` (src/coredynamics.cu:1367-1589,
// 2609-2922; src/coredynamics.cuh:27-118; src/condShape.h:13-41):
//   ge' = ge + sum_pre w p og_pre, he' likewise     (increments of step it-1's spikes)
//   (geD, heD) = decay(ge', he', dt)                 -> next state; gE_t1 += geD
//   (g1, h1)   = decay(gFF, hFF, dt) + LGN/background inputs at dt - tsp
//   (g0, h0)   = (gFF, hFF), decayed to tBack and given the inputs before tBack
//                when backing up from refractoriness (0 < tBack < dt)
//   gE_t0 = sum g0 + sum ge'(tBack);  gE_t1 = sum g1 + sum geD  (I likewise)
//   Heun on [tBack, dt]: v1 = v0 + h fk1, v = v0 + h (fk1 + fk2)/2, h = dt - tBack
//   spike: ratio = (vThres - v0)/(v - v0), tsp = tBack + ratio h,
//          tBack' = tsp + tRef - dt, v' = vR;  else v' = v, tBack' = 0;
//   refractory (tBack >= dt): v' = vR, tBack' = tBack - dt.
//   A spike at tsp sends og = dod (e^{-(dt-tsp)/decay} - e^{-(dt-tsp)/rise}),
//   oh = e^{-(dt-tsp)/rise} to its targets at step it+1.
// The kernel below runs this map backwards for one step with the branches
// (spike or not, backing, refractory) taken from the forward's record:
//   phase 1, one thread per neuron: recompute the step from its snapshot,
//     then propagate lam (the adjoint of the state after the step) and the
//     adjoint of the neuron's own spike time to the state before the step,
//     to its LGN synapse strengths, and to the incremented ge', he', gi', hi';
//   phase 2, the recurrent synapses of step it-1's spikes, lane = post as in
//     the forward gather: dw += p (lam_g' og + lam_h' oh), and the adjoint of
//     each presynaptic spike time, summed over the wavefront's posts, is added
//     to lamTsp[(it-1) & 1] of the presynaptic neuron (float atomics).
#include "vc_v1_device.h"

namespace {

constexpr int kWG = 256;
constexpr int kMaxEnt = 512;
constexpr int kBatch = 64;
constexpr int kW = 32;             // strengths (and gradients) loaded per group
constexpr int kFF = 32;            // LGN synapses of a neuron held in registers

// Contribution of an input of unit size at lag ddt (single_input) and the
// derivatives with respect to ddt: A = dod (Ed - Er), B = Er.
struct Shape {
    float A, B, dA, dB;
};
__device__ __forceinline__ Shape shape_at(float ddt, float rise, float decay, float dod) {
    Shape s;
    const float Er = vc_expf(-ddt / rise);
    const float Ed = vc_expf(-ddt / decay);
    s.A = dod * (Ed - Er);
    s.B = Er;
    s.dA = dod * (-Ed / decay + Er / rise);
    s.dB = -Er / rise;
    return s;
}

struct Adj {
    float ge[VC_MAX_NG], he[VC_MAX_NG], gi[VC_MAX_NG], hi[VC_MAX_NG];
};

// FF / background input of one synapse: forward contribution into (g0, g1,
// h1) and, in reverse, its strength gradient and its tBack gradient.
struct FFIn {
    float nsp, tsp;
};

__global__ __launch_bounds__(kWG) void k_v1_adj(
    vc_v1_params p, V1Const k, vc_v1_state s, vc_ff_conn ff, vc_rec_conn rec, vc_spikes prev,
    const float *__restrict__ snap, const float *__restrict__ sInfoRec,
    const float *__restrict__ lgnSInfo, const float *__restrict__ gOut, float *__restrict__ lam,
    float *__restrict__ lamTspOwn, float *__restrict__ lamTspPrev, float *__restrict__ dsLGN,
    float *__restrict__ dw, uint32_t step) {
    __shared__ uint32_t sbrow[kWG / 64][kBatch];
    __shared__ uint32_t sbslot[kWG / 64][kBatch];
    __shared__ uint32_t sbpre[kWG / 64][kBatch];
    __shared__ uint64_t smk[kMaxEnt];
    __shared__ uint32_t sbid[kMaxEnt];
    // per wavefront, each lane's spike-time contribution to the kW spikes of
    // a group, transposed for the per-spike sums (row stride 65: no bank
    // conflicts when lane j reads row j)
    __shared__ float sred[kWG / 64][kW][65];
    const uint32_t nV1 = p.nV1;
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t b = n / p.blockSize;
    const uint32_t local = n - b * p.blockSize;
    const uint32_t lane = vc_lane();
    const uint32_t wv = threadIdx.x >> 6;
    const int t = type_of(p, local);
    const float dt = p.dt;
    const uint32_t nSeg = k.nSeg;
    const uint32_t myseg = local >> 6;

    // neighbour spike masks of step it-1 (shared by the workgroup)
    const uint32_t nb = rec.nNear[b];
    const uint32_t nEnt = nb * nSeg;
    for (uint32_t e = threadIdx.x; e < nEnt; e += blockDim.x) {
        const uint32_t ib = e / nSeg, tt = e - ib * nSeg;
        const uint32_t bid = rec.nbr[b * rec.maxNear + ib];
        smk[e] = prev.mask[bid * nSeg + tt];
        sbid[e] = bid;
    }

    // ------------------------------------------------------------ phase 1
    const float v0 = snap[VC_SNAP_V * nV1 + n];
    const float tB = snap[VC_SNAP_TBACK * nV1 + n];
    float gF[VC_MAX_NG], hF[VC_MAX_NG], ge[VC_MAX_NG], he[VC_MAX_NG], gi[VC_MAX_NG],
        hi[VC_MAX_NG];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        gF[ig] = ig < p.ngFF ? snap[(VC_SNAP_GFF + ig) * nV1 + n] : 0.0f;
        hF[ig] = ig < p.ngFF ? snap[(VC_SNAP_HFF + ig) * nV1 + n] : 0.0f;
        ge[ig] = ig < p.ngE ? snap[(VC_SNAP_GE + ig) * nV1 + n] : 0.0f;
        he[ig] = ig < p.ngE ? snap[(VC_SNAP_HE + ig) * nV1 + n] : 0.0f;
        gi[ig] = ig < p.ngI ? snap[(VC_SNAP_GI + ig) * nV1 + n] : 0.0f;
        hi[ig] = ig < p.ngI ? snap[(VC_SNAP_HI + ig) * nV1 + n] : 0.0f;
    }
    const float sInfo = sInfoRec[n];
    const bool spiked = sInfo >= 1.0f;
    const float gout = gOut[n];
    float lamTsp = lamTspOwn[n];
    lamTspOwn[n] = 0.0f;                       // the slot serves step it-2 next
    // adjoint of the state after the step
    const float lv1 = lam[VC_SNAP_V * nV1 + n], ltB1 = lam[VC_SNAP_TBACK * nV1 + n];
    float lg1[VC_MAX_NG], lh1[VC_MAX_NG], lgeD[VC_MAX_NG], lheD[VC_MAX_NG], lgiD[VC_MAX_NG],
        lhiD[VC_MAX_NG];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        lg1[ig] = lam[(VC_SNAP_GFF + ig) * nV1 + n];
        lh1[ig] = lam[(VC_SNAP_HFF + ig) * nV1 + n];
        lgeD[ig] = lam[(VC_SNAP_GE + ig) * nV1 + n];
        lheD[ig] = lam[(VC_SNAP_HE + ig) * nV1 + n];
        lgiD[ig] = lam[(VC_SNAP_GI + ig) * nV1 + n];
        lhiD[ig] = lam[(VC_SNAP_HI + ig) * nV1 + n];
    }

    // ---- recompute the forward step
    const float vThres = p.vThres[t], gL = p.gL[t], Cm = p.C[t], vT = p.vT[t];
    const float tau_w = p.tau_w[t], a = p.a[t];
    const float dep = s.tonicDep[n];
    float depC;
    {
        const float targetV = VC_VL + (vT - VC_VL) * dep;
        depC = (gL + a) * (targetV - VC_VL) - 0.0f;      // deltaT = 0
    }
    const float w0s = s.w[n];
    float geD[VC_MAX_NG], giD[VC_MAX_NG];
    float gE_t1 = 0.0f, gI_t1 = 0.0f;
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        float g = ge[ig], h = he[ig];
        decay_const(g, h, k.etrE[ig], k.etdE[ig], p.dodE[ig]);
        geD[ig] = g;
        if (ig < p.ngE) gE_t1 += g;
        g = gi[ig];
        h = hi[ig];
        decay_const(g, h, k.etrI[ig], k.etdI[ig], p.dodI[ig]);
        giD[ig] = g;
        if (ig < p.ngI) gI_t1 += g;
    }
    const bool backing = (tB < dt) && (tB > 0.0f);
    float g1[VC_MAX_NG], g0[VC_MAX_NG];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        float g = gF[ig], h = hF[ig];
        decay_const(g, h, k.etrFF[ig], k.etdFF[ig], p.dodFF[ig]);
        g1[ig] = g;
        g = gF[ig];
        h = hF[ig];
        if (backing) decay_var(g, h, tB, p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
        g0[ig] = g;
    }
    const uint32_t mLGN = ff.nLGN[n];
    // The neuron's first kFF LGN synapses: strengths, cells and the cells'
    // spike words, loaded together (two round trips).  Only synapses whose
    // cell spiked contribute; they are visited in list order through a
    // wave-uniform slot index (the slots that spiked for any lane), so the
    // input arithmetic has one copy.  Synapses past kFF take a per-lane loop.
    float fv[kFF], sv[kFF];
    {
        uint32_t fid[kFF];
#pragma unroll
        for (int j = 0; j < kFF; ++j) {
            fv[j] = 0.0f;
            fid[j] = 0u;
            if ((uint32_t)j < mLGN) {
                fv[j] = ff.sLGN[(uint32_t)j * nV1 + n];
                fid[j] = ff.idLGN[(uint32_t)j * nV1 + n];
            }
        }
#pragma unroll
        for (int j = 0; j < kFF; ++j) sv[j] = (uint32_t)j < mLGN ? lgnSInfo[fid[j]] : 0.0f;
    }
    uint32_t spkAny = 0;
#pragma unroll
    for (int j = 0; j < kFF; ++j) spkAny |= (sv[j] >= 1.0f ? 1u : 0u) << j;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) spkAny |= (uint32_t)__shfl_xor((int)spkAny, off);
    spkAny = __builtin_amdgcn_readfirstlane(spkAny);
    // the value of slot j (wave-uniform) of a register array
    auto slot = [](const float (&a)[kFF], uint32_t j) {
        float x = 0.0f;
#pragma unroll
        for (int q = 0; q < kFF; ++q)
            if ((uint32_t)q == j) x = a[q];
        return x;
    };
    // LGN input (f, sI) of the neuron into g1 / g0
    auto ff_fwd = [&](float f, float sI) {
        const float nsp = (float)(uint32_t)floorf(sI), tsp = (sI - nsp) * dt;
        #pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) if (ig < p.ngFF) {
            const float str = f * p.pFF[t * p.ngFF + ig] * nsp;
            const Shape s1 = shape_at(dt - tsp, p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
            g1[ig] = g1[ig] + str * s1.A;
            if (backing && tB - tsp > 0.0f) {
                const Shape s0 = shape_at(tB - tsp, p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
                g0[ig] = g0[ig] + str * s0.A;
            }
        }
    };
    // LGN inputs in list order, then the background drive
    for (uint32_t any = spkAny; any != 0u;) {
        const uint32_t j = (uint32_t)__builtin_ctz(any);
        any &= any - 1u;
        const float sI = slot(sv, j), f = slot(fv, j);
        if (sI >= 1.0f) ff_fwd(f, sI);
    }
    for (uint32_t j = kFF; j < mLGN; ++j) {
        const uint32_t lid = j * nV1 + n;
        const float sI = lgnSInfo[ff.idLGN[lid]];
        if (sI >= 1.0f) ff_fwd(ff.sLGN[lid], sI);
    }
    FFIn bg{0.0f, 0.0f};
    if (p.bgStrength[t] > 0.0f) {
        uint32_t klo = p.seed_lo, khi = p.seed_hi, ent = n;
        if (p.blockTrial != nullptr) {
            const uint32_t tr = p.blockTrial[b];
            klo = p.trialSeed[2 * tr];
            khi = p.trialSeed[2 * tr + 1];
            ent = n - p.trialV1Off[tr];
        }
        const vc_u32x4 r = vc_draw4(klo, khi, ent, step, VC_STREAM_V1_BG, 0u, p.trial);
        const float u = vc_u01_open0(r.v[0]);
        const float frac = vc_u01_open1(r.v[1]);
        bg.nsp = (float)vc_poisson_inv(u, p.bgLambda[t], p.bgP0[t]);
        bg.tsp = frac * dt;
        if (bg.nsp > 0.0f) {
            #pragma unroll
            for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) if (ig < p.ngFF) {
                const float str = p.bgStrength[t] * p.pBG[t * p.ngFF + ig] * bg.nsp;
                const Shape s1 = shape_at(dt - bg.tsp, p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
                g1[ig] = g1[ig] + str * s1.A;
                if (backing && tB - bg.tsp > 0.0f) {
                    const Shape s0 = shape_at(tB - bg.tsp, p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
                    g0[ig] = g0[ig] + str * s0.A;
                }
            }
        }
    }
    float gE_t0 = 0.0f, gI_t0 = 0.0f;
    float EdE[VC_MAX_NG], ErE[VC_MAX_NG], EdI[VC_MAX_NG], ErI[VC_MAX_NG];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        EdE[ig] = ErE[ig] = EdI[ig] = ErI[ig] = 1.0f;
        if (ig < p.ngFF) {
            gE_t0 += g0[ig];
            gE_t1 += g1[ig];
        }
        if (backing) {
            ErE[ig] = vc_expf(-tB / p.riseE[ig]);
            EdE[ig] = vc_expf(-tB / p.decayE[ig]);
            ErI[ig] = vc_expf(-tB / p.riseI[ig]);
            EdI[ig] = vc_expf(-tB / p.decayI[ig]);
        }
        if (ig < p.ngE)
            gE_t0 += backing ? ge[ig] * EdE[ig] + he[ig] * p.dodE[ig] * (EdE[ig] - ErE[ig]) : ge[ig];
        if (ig < p.ngI)
            gI_t0 += backing ? gi[ig] * EdI[ig] + hi[ig] * p.dodI[ig] * (EdI[ig] - ErI[ig]) : gi[ig];
    }

    // ---- reverse of the membrane update
    float lv0 = 0.0f, ltB = 0.0f;
    float lgE_t0 = 0.0f, lgI_t0 = 0.0f, lgE_t1 = 0.0f, lgI_t1 = 0.0f;
    if (tB < dt) {
        const float a0 = gE_t0 + gI_t0 + gL + 0.0f;
        const float b0 = gE_t0 * VC_VE + gI_t0 * VC_VI + gL * VC_VL + depC + 0.0f;
        const float a1 = gE_t1 + gI_t1 + gL + 0.0f;
        const float b1 = gE_t1 * VC_VE + gI_t1 * VC_VI + gL * VC_VL + depC + 0.0f;
        const float hstep = dt - tB;
        float w0 = w0s;
        if (backing) w0 = (w0 - 0.0f) * vc_expf(-tB / tau_w) + 0.0f;
        const float fk1 = (-a0 * v0 + b0 - w0) / Cm;
        const float gk1 = (a * (v0 - VC_VL) - w0) / tau_w;
        const float v1 = v0 + hstep * fk1;
        const float w1 = w0 + hstep * gk1;
        const float fk2 = (-a1 * v1 + b1 - w1) / Cm;
        const float v = v0 + hstep * (fk1 + fk2) / 2.0f;
        float lv = 0.0f, lh = 0.0f;
        if (spiked) {
            const float ratio = (vThres - v0) / (v - v0);
            const float lt = lamTsp + gout / dt + ltB1;       // tsp -> sInfo, tBack'
            lh += lt * ratio;
            ltB += lt;
            const float lr = lt * hstep;
            lv = lr * (-ratio / (v - v0));
            lv0 += lr * (ratio - 1.0f) / (v - v0);
        } else {
            lv = lv1 + gout;                                  // v' = v, sInfo = v
        }
        lv0 += lv;
        float lfk1 = lv * hstep / 2.0f, lfk2 = lv * hstep / 2.0f;
        lh += lv * (fk1 + fk2) / 2.0f;
        const float la1 = -lfk2 * v1 / Cm, lv1s = -lfk2 * a1 / Cm, lb1 = lfk2 / Cm;
        const float lw1 = -lfk2 / Cm;
        float lw0 = lw1 * (1.0f - hstep / tau_w);
        lh += lw1 * (-w0 / tau_w);
        lv0 += lv1s;
        lh += lv1s * fk1;
        lfk1 += lv1s * hstep;
        const float la0 = -lfk1 * v0 / Cm, lb0 = lfk1 / Cm;
        lv0 += -lfk1 * a0 / Cm;
        lw0 += -lfk1 / Cm;
        if (backing) ltB += lw0 * (-w0 / tau_w);
        ltB -= lh;
        lgE_t0 = la0 + lb0 * VC_VE;
        lgI_t0 = la0 + lb0 * VC_VI;
        lgE_t1 = la1 + lb1 * VC_VE;
        lgI_t1 = la1 + lb1 * VC_VI;
    } else {
        ltB += ltB1;                                          // tBack' = tBack - dt
    }

    // ---- reverse of the conductances
    Adj L;
    float lgF[VC_MAX_NG], lhF[VC_MAX_NG], lg0[VC_MAX_NG], lg1t[VC_MAX_NG];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        const float lgeDt = lgeD[ig] + lgE_t1, lgiDt = lgiD[ig] + lgI_t1;
        const float lgB = lgE_t0, liB = lgI_t0;
        // geD = ge etd + he dod (etd - etr), heD = he etr
        float lge = lgeDt * k.etdE[ig], lhe = lgeDt * p.dodE[ig] * (k.etdE[ig] - k.etrE[ig]) +
                                              lheD[ig] * k.etrE[ig];
        float lgi = lgiDt * k.etdI[ig], lhi = lgiDt * p.dodI[ig] * (k.etdI[ig] - k.etrI[ig]) +
                                              lhiD[ig] * k.etrI[ig];
        if (backing) {
            lge += lgB * EdE[ig];
            lhe += lgB * p.dodE[ig] * (EdE[ig] - ErE[ig]);
            lgi += liB * EdI[ig];
            lhi += liB * p.dodI[ig] * (EdI[ig] - ErI[ig]);
            if (ig < p.ngE)
                ltB += lgB * (ge[ig] * (-EdE[ig] / p.decayE[ig]) +
                              he[ig] * p.dodE[ig] * (-EdE[ig] / p.decayE[ig] + ErE[ig] / p.riseE[ig]));
            if (ig < p.ngI)
                ltB += liB * (gi[ig] * (-EdI[ig] / p.decayI[ig]) +
                              hi[ig] * p.dodI[ig] * (-EdI[ig] / p.decayI[ig] + ErI[ig] / p.riseI[ig]));
        } else {
            lge += lgB;
            lgi += liB;
        }
        L.ge[ig] = ig < p.ngE ? lge : 0.0f;
        L.he[ig] = ig < p.ngE ? lhe : 0.0f;
        L.gi[ig] = ig < p.ngI ? lgi : 0.0f;
        L.hi[ig] = ig < p.ngI ? lhi : 0.0f;
        // FF: g1 = gF etd + hF dod (etd - etr) + inputs, h1 = hF etr + inputs;
        // g0 = gF (or decayed to tB) + inputs before tB
        lg1t[ig] = lg1[ig] + lgE_t1;
        lg0[ig] = lgE_t0;
        float lgf = lg1t[ig] * k.etdFF[ig];
        float lhf = lg1t[ig] * p.dodFF[ig] * (k.etdFF[ig] - k.etrFF[ig]) + lh1[ig] * k.etrFF[ig];
        if (backing) {
            const float Er = vc_expf(-tB / p.riseFF[ig]), Ed = vc_expf(-tB / p.decayFF[ig]);
            lgf += lg0[ig] * Ed;
            lhf += lg0[ig] * p.dodFF[ig] * (Ed - Er);
            if (ig < p.ngFF)
                ltB += lg0[ig] * (gF[ig] * (-Ed / p.decayFF[ig]) +
                                  hF[ig] * p.dodFF[ig] * (-Ed / p.decayFF[ig] + Er / p.riseFF[ig]));
        } else {
            lgf += lg0[ig];
        }
        lgF[ig] = ig < p.ngFF ? lgf : 0.0f;
        lhF[ig] = ig < p.ngFF ? lhf : 0.0f;
    }
    // inputs: strength gradients and the tBack dependence of the backed-up part
    auto ff_rev = [&](float f, float sI) {
        const float nsp = (float)(uint32_t)floorf(sI), tsp = (sI - nsp) * dt;
        float df = 0.0f;
        #pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) if (ig < p.ngFF) {
            const float pn = p.pFF[t * p.ngFF + ig] * nsp;
            const Shape s1 = shape_at(dt - tsp, p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
            df += pn * (lg1t[ig] * s1.A + lh1[ig] * s1.B);
            if (backing && tB - tsp > 0.0f) {
                const Shape s0 = shape_at(tB - tsp, p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
                df += pn * lg0[ig] * s0.A;
                ltB += f * pn * lg0[ig] * s0.dA;
            }
        }
        return df;
    };
    for (uint32_t any = spkAny; any != 0u;) {
        const uint32_t j = (uint32_t)__builtin_ctz(any);
        any &= any - 1u;
        const float sI = slot(sv, j), f = slot(fv, j);
        if (sI >= 1.0f) dsLGN[j * nV1 + n] += ff_rev(f, sI);
    }
    for (uint32_t j = kFF; j < mLGN; ++j) {
        const uint32_t lid = j * nV1 + n;
        const float sI = lgnSInfo[ff.idLGN[lid]];
        if (sI >= 1.0f) dsLGN[lid] += ff_rev(ff.sLGN[lid], sI);
    }
    if (bg.nsp > 0.0f && backing && tB - bg.tsp > 0.0f) {
        #pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) if (ig < p.ngFF) {
            const float str = p.bgStrength[t] * p.pBG[t * p.ngFF + ig] * bg.nsp;
            const Shape s0 = shape_at(tB - bg.tsp, p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
            ltB += str * lg0[ig] * s0.dA;
        }
    }
    // adjoint of the state before the step (the increments are additive, so
    // the cortical rows are the adjoints of ge', he', gi', hi')
    lam[VC_SNAP_V * nV1 + n] = lv0;
    lam[VC_SNAP_TBACK * nV1 + n] = ltB;
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        lam[(VC_SNAP_GFF + ig) * nV1 + n] = lgF[ig];
        lam[(VC_SNAP_HFF + ig) * nV1 + n] = lhF[ig];
        lam[(VC_SNAP_GE + ig) * nV1 + n] = L.ge[ig];
        lam[(VC_SNAP_HE + ig) * nV1 + n] = L.he[ig];
        lam[(VC_SNAP_GI + ig) * nV1 + n] = L.gi[ig];
        lam[(VC_SNAP_HI + ig) * nV1 + n] = L.hi[ig];
    }

    // ------------------------------------------------------------ phase 2
    __syncthreads();                           // smk / sbid staged
    const uint32_t bs = p.blockSize;
    uint32_t *brow = sbrow[wv], *bslot = sbslot[wv], *bpre = sbpre[wv];
    const uint32_t nChunk = (nEnt + 63u) / 64u;
    uint32_t c = 0;
    uint64_t nz = 0;
    // the increments of an E (I) spike, per receptor: p * (lam_g og + lam_h oh)
    // and their spike-time derivatives og' = (og + dod oh)/decay - dod oh/rise,
    // oh' = oh/rise
    float pE[VC_MAX_NG], pI[VC_MAX_NG], irE[VC_MAX_NG], idE[VC_MAX_NG], irI[VC_MAX_NG],
        idI[VC_MAX_NG];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        pE[ig] = ig < p.ngE ? p.pE[t * p.ngE + ig] : 0.0f;
        pI[ig] = ig < p.ngI ? p.pI[t * p.ngI + ig] : 0.0f;
        irE[ig] = 1.0f / p.riseE[ig];
        idE[ig] = 1.0f / p.decayE[ig];
        irI[ig] = 1.0f / p.riseI[ig];
        idI[ig] = 1.0f / p.decayI[ig];
    }
    float (*red)[65] = sred[wv];
    while (true) {
        uint32_t cnt = 0;
        while (cnt < (uint32_t)kBatch) {
            while (nz == 0ull && c < nChunk) {
                const uint32_t e = c * 64u + lane;
                nz = __ballot(e < nEnt && smk[e] != 0ull);
                ++c;
            }
            if (nz == 0ull) break;
            const uint32_t j = (uint32_t)__builtin_ctzll(nz);
            const uint32_t e = (c - 1u) * 64u + j;
            const uint64_t m = smk[e];
            const uint32_t np = (uint32_t)__builtin_popcountll(m);
            if (cnt + np > (uint32_t)kBatch) break;        // next batch takes the tile
            nz &= nz - 1ull;
            const uint32_t bid = sbid[e];
            const uint32_t ib = e / nSeg, tt = e - ib * nSeg;
            if ((m >> lane) & 1ull) {
                const uint32_t rank = vc_mbcnt(m);
                const uint32_t pos = cnt + rank;
                const uint32_t pre = tt * 64u + lane;
                brow[pos] = (b * rec.maxNear + ib) * bs + pre;
                bslot[pos] = (bid * bs + tt * 64u + rank) | ((pre >= p.nE ? 1u : 0u) << 31);
                bpre[pos] = bid * bs + pre;
            }
            cnt += np;
        }
        if (cnt == 0u) break;
        __builtin_amdgcn_wave_barrier();
        // lane kk holds the header of spike kk
        uint64_t rm = 0;
        uint32_t off = 0, isI = 0;
        float og0 = 0.f, oh0 = 0.f, og1 = 0.f, oh1 = 0.f;
        if (lane < cnt) {
            const uint32_t row = brow[lane], sw = bslot[lane];
            isI = sw >> 31;
            const uint32_t slot = sw & 0x7fffffffu;
            const size_t hid = (size_t)row * nSeg + myseg;
            rm = rec.rowMask[hid];
            off = rec.rowOff[hid];
            og0 = prev.og[slot];
            oh0 = prev.oh[slot];
            og1 = prev.og[nV1 + slot];
            oh1 = prev.oh[nV1 + slot];
        }
        __builtin_amdgcn_wave_barrier();
        for (uint32_t k0 = 0; k0 < cnt; k0 += kW) {
            float wv_[kW], dv_[kW];
            uint32_t ix[kW];
#pragma unroll
            for (int jj = 0; jj < kW; ++jj) {
                const uint32_t kk = k0 + jj;
                wv_[jj] = 0.0f;
                dv_[jj] = 0.0f;
                ix[jj] = 0u;
                if (kk < cnt) {
                    const uint64_t rmk = vc_readlane_u64(rm, kk);
                    ix[jj] = vc_readlane_u32(off, kk) + vc_mbcnt(rmk);
                    if ((rmk >> lane) & 1ull) {
                        wv_[jj] = rec.w[ix[jj]];
                        dv_[jj] = dw[ix[jj]];
                    }
                }
            }
#pragma unroll
            for (int jj = 0; jj < kW; ++jj) {
                const uint32_t kk = k0 + jj;
                float dts = 0.0f;
                if (kk < cnt) {
                    const uint64_t rmk = vc_readlane_u64(rm, kk);
                    const float g[2] = {vc_readlane_f32(og0, kk), vc_readlane_f32(og1, kk)};
                    const float h[2] = {vc_readlane_f32(oh0, kk), vc_readlane_f32(oh1, kk)};
                    const bool e = vc_readlane_u32(isI, kk) == 0u;
                    float contrib = 0.0f;
                    if ((rmk >> lane) & 1ull) {
#pragma unroll
                        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
                            const float pp = e ? pE[ig] : pI[ig];
                            const float lg = e ? L.ge[ig] : L.gi[ig];
                            const float lhh = e ? L.he[ig] : L.hi[ig];
                            const float ir = e ? irE[ig] : irI[ig];
                            const float id = e ? idE[ig] : idI[ig];
                            const float dod = e ? p.dodE[ig] : p.dodI[ig];
                            if (pp != 0.0f) {
                                contrib += pp * (lg * g[ig] + lhh * h[ig]);
                                const float dg = (g[ig] + dod * h[ig]) * id - dod * h[ig] * ir;
                                const float dh = h[ig] * ir;
                                dts += pp * (lg * dg + lhh * dh);
                            }
                        }
                        dw[ix[jj]] = dv_[jj] + contrib;
                    }
                }
                red[jj][lane] = wv_[jj] * dts;
            }
            // per spike, the sum over the wavefront's posts: lane j sums row j
            __builtin_amdgcn_wave_barrier();
            if (lane < (uint32_t)kW && k0 + lane < cnt) {
                float tot = 0.0f;
                for (int l = 0; l < 64; ++l) tot += red[lane][l];
                if (tot != 0.0f) atomicAdd(lamTspPrev + bpre[k0 + lane], tot);
            }
            __builtin_amdgcn_wave_barrier();
        }
        __builtin_amdgcn_wave_barrier();
    }
}

}  // namespace

extern "C" int vc_v1_adjoint_run(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                                 const vc_ff_conn *ff, const vc_rec_conn *rec,
                                 const vc_v1_adjoint *a, uint32_t it0, uint32_t n) {
    if (!p || !s || !ff || !rec || !a || !a->snap || !a->spk || !a->sInfo || !a->lgnSInfo ||
        !a->gOut || !a->lam || !a->lamTsp || !a->dsLGN || !a->dw || !s->tonicDep || !s->w) {
        vc_set_error("vc_v1_adjoint_run: null argument");
        return VC_E_INVALID;
    }
    if (p->blockSize == 0 || p->blockSize % 64 != 0 || p->nV1 != p->nblock * p->blockSize ||
        p->nType < 1 || p->nType > VC_MAX_NTYPE) {
        vc_set_error("vc_v1_adjoint_run: bad sizes");
        return VC_E_INVALID;
    }
    if (p->iModel != 0) {
        vc_set_error("vc_v1_adjoint_run: only the LIF model (iModel 0) is differentiated");
        return VC_E_UNSUPPORTED;
    }
    if (s->gapS || s->gap || rec->gapPre) {
        vc_set_error("vc_v1_adjoint_run: gap junctions are not differentiated");
        return VC_E_UNSUPPORTED;
    }
    for (uint32_t t = 0; t < p->nType; ++t) {
        bool off = p->noisyDep[t] == 0.0f && p->synFailFF[t] == 0.0f;
        for (uint32_t u = 0; u < p->nType; ++u) off = off && p->synFail[t * p->nType + u] == 0.0f;
        if (!off) {
            vc_set_error("vc_v1_adjoint_run: voltage noise and synaptic failure are not "
                         "differentiated");
            return VC_E_UNSUPPORTED;
        }
    }
    for (uint32_t t = 0; t < p->nType; ++t) {
        if (p->a[t] != 0.0f || p->deltaT[t] != 0.0f) {
            vc_set_error("vc_v1_adjoint_run: type %u has a = %g, deltaT = %g (conductance LIF "
                         "needs both 0)", t, p->a[t], p->deltaT[t]);
            return VC_E_UNSUPPORTED;
        }
        if (!(p->tRef[t] >= p->dt)) {
            vc_set_error("vc_v1_adjoint_run: tRef < dt is not built");
            return VC_E_UNSUPPORTED;
        }
    }
    if (rec->maxNear * (p->blockSize / 64) > (uint32_t)kMaxEnt) {
        vc_set_error("vc_v1_adjoint_run: too many neighbour tiles");
        return VC_E_UNSUPPORTED;
    }
    const V1Const k = make_const(p);
    const uint32_t wg = wg_size(p->blockSize);
    const size_t nV1 = p->nV1;
    for (uint32_t q = n; q-- > 0;) {
        const uint32_t it = it0 + q;
        hipLaunchKernelGGL(k_v1_adj, dim3(p->nV1 / wg), dim3(wg), 0, (hipStream_t)stream, *p, k,
                           *s, *ff, *rec, a->spk[q], a->snap + (size_t)q * VC_SNAP_ROWS * nV1,
                           a->sInfo + (size_t)q * nV1, a->lgnSInfo + (size_t)q * a->lgnStride,
                           a->gOut + (size_t)q * nV1, a->lam, a->lamTsp + (size_t)(it & 1u) * nV1,
                           a->lamTsp + (size_t)((it + 1u) & 1u) * nV1, a->dsLGN, a->dw, it);
        VC_HIP_CHECK(hipGetLastError());
    }
    return VC_OK;
}
