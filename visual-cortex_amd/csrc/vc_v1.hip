// vc_v1.hip -- stages c, d, e of the V1 simulation step on gfx950.
//
// One thread per V1 neuron, workgroups of up to 256 threads inside one V1
// block, so every wavefront is one 64-neuron tile.  The kernel is latency
// bound at the full-patch size (one wavefront per SIMD), so it is organised
// around few dependent memory round trips:
//   round 1: state, the LGN synapse indices/strengths of the neuron, and the
//            spike masks of the 3x3 neighbour blocks (one 64-bit mask per
//            tile, held lane-parallel: lane l <- tile entry l);
//   round 2: the LGN spike words, and for the spikes of the neighbourhood
//            (listed in neighbour order, ascending neuron id, by a ballot
//            scan over the non-empty tiles) one masked-CSR row-segment header
//            per spike (lane k <- spike k);
//   round 3: the synaptic weights each lane owns (rank via mbcnt).
// Sums are accumulated in the reference order: recurrent E and I sums over
// neighbour blocks then ascending presynaptic id (src/coredynamics.cu
// :2721-2838), LGN synapses in list order (src/coredynamics.cu:1466-1502).
//   stage d: decay, Heun RK2 AdEx/LIF, spike time, refractory back-up
//            (src/coredynamics.cu:1367-1589, src/coredynamics.cuh:27-118);
//   stage e: ballot compaction -- one spike mask per tile and the unit
//            increments at slot tile*64 + rank (src/coredynamics.cu:1994-2046).
#include "vc_v1_device.h"

namespace {

constexpr int kWG = 256;           // max workgroup size
constexpr int kMaxEnt = 512;       // neighbour-tile entries staged in LDS per wavefront
constexpr int kBatch = 64;
constexpr int kFF = 32;            // LGN synapses loaded per round
constexpr int kW = 8;              // weight loads in flight per lane
constexpr int kEv = 4;             // LGN spike events prefetched per neuron



// LGN (or background) conductance input of one synapse, src/coredynamics.cu
// :1474-1500: nsp spikes at tsp into (g1, h1) over dt - tsp and, when the
// neuron is backing up from refractoriness, into (g0, h0) over tBack - tsp.
// Not inlined: it runs only for synapses whose cell spiked, and one copy of
// its code serves every synapse slot.
struct FFAcc {
    float g0[VC_MAX_NG], h0[VC_MAX_NG], g1[VC_MAX_NG], h1[VC_MAX_NG];
};
struct FFShape {
    float p[VC_MAX_NG], rise[VC_MAX_NG], decay[VC_MAX_NG], dod[VC_MAX_NG];
    uint32_t ng;
};
__device__ __noinline__ FFAcc ff_add(FFAcc a, const FFShape sh, float f, float nsp, float tsp,
                                     float dt, float tBack, bool backing) {
    if (f > 0.0f) {
        float ddt = 0.0f;
        if (backing) ddt = tBack - tsp;
#pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
            if (ig < sh.ng) {
                const float str = f * sh.p[ig];
                if (backing && ddt > 0.0f)
                    single_input(a.g0[ig], a.h0[ig], str * nsp, ddt, sh.rise[ig], sh.decay[ig],
                                 sh.dod[ig]);
                single_input(a.g1[ig], a.h1[ig], str * nsp, dt - tsp, sh.rise[ig], sh.decay[ig],
                             sh.dod[ig]);
            }
        }
    }
    return a;
}


// Adds the recurrent increments of a batch of K <= 64 spikes listed in
// brow/bslot (this wavefront's LDS rows) to the calling lane's sums.
__device__ __forceinline__ void flush_batch(const vc_v1_params &p, const vc_rec_conn &rec,
                                            const vc_spikes &prev, uint32_t K, uint32_t lane,
                                            uint32_t nSeg, uint32_t myseg, int t,
                                            const uint32_t *brow, const uint32_t *bslot,
                                            float lgE[VC_MAX_NG], float lhE[VC_MAX_NG],
                                            float lgI[VC_MAX_NG], float lhI[VC_MAX_NG]) {
    // lane kk fetches the row-segment header and the increments of spike kk
    uint64_t rm = 0;
    uint32_t off = 0, isI = 0;
    float og0 = 0.f, oh0 = 0.f, og1 = 0.f, oh1 = 0.f;
    if (lane < K) {
        const uint32_t row = brow[lane];
        const uint32_t sw = bslot[lane];
        isI = sw >> 31;
        const uint32_t slot = sw & 0x7fffffffu;
        const size_t hid = (size_t)row * nSeg + myseg;
        rm = rec.rowMask[hid];
        off = rec.rowOff[hid];
        og0 = prev.og[slot];
        oh0 = prev.oh[slot];
        og1 = prev.og[p.nV1 + slot];
        oh1 = prev.oh[p.nV1 + slot];
    }
    const float pE0 = p.pE[t * p.ngE + 0], pE1 = p.pE[t * p.ngE + (p.ngE > 1 ? 1 : 0)];
    const float pI0 = p.pI[t * p.ngI + 0], pI1 = p.pI[t * p.ngI + (p.ngI > 1 ? 1 : 0)];
    for (uint32_t k0 = 0; k0 < K; k0 += kW) {
        float wv[kW];
#pragma unroll
        for (int j = 0; j < kW; ++j) {
            const uint32_t kk = k0 + j;
            wv[j] = 0.0f;
            if (kk < K) {
                const uint64_t rmk = vc_readlane_u64(rm, kk);
                const uint32_t idx = vc_readlane_u32(off, kk) + vc_mbcnt(rmk);
                if ((rmk >> lane) & 1ull) wv[j] = rec.w[idx];
            }
        }
#pragma unroll
        for (int j = 0; j < kW; ++j) {
            const uint32_t kk = k0 + j;
            if (kk < K) {
                const uint64_t rmk = vc_readlane_u64(rm, kk);
                const float g0 = vc_readlane_f32(og0, kk), h0 = vc_readlane_f32(oh0, kk);
                const float g1 = vc_readlane_f32(og1, kk), h1 = vc_readlane_f32(oh1, kk);
                const uint32_t inh = vc_readlane_u32(isI, kk);
                if ((rmk >> lane) & 1ull) {
                    const float strength = wv[j];
                    if (inh == 0u) {
                        float str = strength * pE0;
                        lgE[0] += str * g0;
                        lhE[0] += str * h0;
                        if (p.ngE > 1) {
                            str = strength * pE1;
                            lgE[1] += str * g1;
                            lhE[1] += str * h1;
                        }
                    } else {
                        float str = strength * pI0;
                        lgI[0] += str * g0;
                        lhI[0] += str * h0;
                        if (p.ngI > 1) {
                            str = strength * pI1;
                            lgI[1] += str * g1;
                            lhI[1] += str * h1;
                        }
                    }
                }
            }
        }
    }
}

// Recurrent increments of the spikes of `prev` for the calling neuron.
// smk/sbid (this wavefront's LDS rows) hold the spike mask and block id of
// neighbour-tile entry e = ib*nSeg + tt, e < nEnt.  Entries are scanned in
// order with a ballot over 64 entries at a time; the spikes of non-empty
// tiles are appended to the batch, which is flushed from one place.
__device__ __forceinline__ void rec_gather(const vc_v1_params &p, const vc_rec_conn &rec,
                                        const vc_spikes &prev, uint32_t b, uint32_t myseg,
                                        uint32_t nSeg, uint32_t nEnt, int t,
                                        const uint64_t *smk, const uint32_t *sbid,
                                        uint32_t *brow, uint32_t *bslot,
                                        float lgE[VC_MAX_NG], float lhE[VC_MAX_NG],
                                        float lgI[VC_MAX_NG], float lhI[VC_MAX_NG]) {
    const uint32_t lane = vc_lane();
    const uint32_t bs = p.blockSize;
    const uint32_t nChunk = (nEnt + 63u) / 64u;
    uint32_t cnt = 0;                     // entries in the batch (wave-uniform)
    uint32_t c = 0;                       // next chunk of 64 entries to scan
    uint64_t nz = 0;                      // non-empty entries left in chunk c-1
    while (true) {
        // fill the batch up to 64 spikes or until the entries are exhausted
        while (cnt < (uint32_t)kBatch) {
            while (nz == 0ull && c < nChunk) {
                const uint32_t e = c * 64u + lane;
                nz = __ballot(e < nEnt && smk[e] != 0ull);
                ++c;
            }
            if (nz == 0ull) break;
            const uint32_t j = (uint32_t)__builtin_ctzll(nz);
            nz &= nz - 1ull;
            const uint32_t e = (c - 1u) * 64u + j;
            const uint64_t m = smk[e];
            const uint32_t bid = sbid[e];
            const uint32_t ib = e / nSeg, tt = e - ib * nSeg;
            if ((m >> lane) & 1ull) {
                const uint32_t rank = vc_mbcnt(m);
                const uint32_t pos = cnt + rank;
                const uint32_t pre = tt * 64u + lane;
                brow[pos] = (b * rec.maxNear + ib) * bs + pre;
                bslot[pos] = (bid * bs + tt * 64u + rank) | ((pre >= p.nE ? 1u : 0u) << 31);
            }
            cnt += (uint32_t)__builtin_popcountll(m);
        }
        if (cnt == 0u) break;
        const uint32_t K = cnt < (uint32_t)kBatch ? cnt : (uint32_t)kBatch;
        __builtin_amdgcn_wave_barrier();
        flush_batch(p, rec, prev, K, lane, nSeg, myseg, t, brow, bslot, lgE, lhE, lgI, lhI);
        __builtin_amdgcn_wave_barrier();
        const uint32_t rest = cnt - K;
        uint32_t r0 = 0, r1 = 0;
        if (lane < rest) {
            r0 = brow[K + lane];
            r1 = bslot[K + lane];
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < rest) {
            brow[lane] = r0;
            bslot[lane] = r1;
        }
        __builtin_amdgcn_wave_barrier();
        cnt = rest;
    }
}

template <bool kUpdate, bool kEvents>
__global__ __launch_bounds__(kWG) void k_v1_step(vc_v1_params p, V1Const k, vc_v1_state s,
                                                 vc_ff_conn ff, vc_rec_conn rec,
                                                 vc_spikes prev, vc_spikes next,
                                                 const float *__restrict__ lgnSInfo,
                                                 uint32_t *__restrict__ ffMask,
                                                 uint32_t step, float *__restrict__ out_sInfo,
                                                 float *__restrict__ out_dep,
                                                 float *__restrict__ out_g_prev,
                                                 float *__restrict__ out_snap) {
    __shared__ uint32_t sbrow[kWG / 64][2 * kBatch];
    __shared__ uint32_t sbslot[kWG / 64][2 * kBatch];
    __shared__ uint64_t smk[kMaxEnt];          // neighbour-tile masks of this block (all
    __shared__ uint32_t sbid[kMaxEnt];         // wavefronts of a workgroup share the block)
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

    // ---- round 1: neighbour spike masks (lane-parallel), LGN synapses, state
    const uint32_t nb = rec.nNear[b];
    const uint32_t nEnt = nb * nSeg;
    for (uint32_t e = threadIdx.x; e < nEnt; e += blockDim.x) {
        const uint32_t ib = e / nSeg, tt = e - ib * nSeg;
        const uint32_t bid = rec.nbr[b * rec.maxNear + ib];
        smk[e] = prev.mask[bid * nSeg + tt];
        sbid[e] = bid;
    }
    constexpr bool events = kUpdate && kEvents;
    constexpr int nFF = (kUpdate && !kEvents) ? kFF : 1;    // ELL synapses held in registers
    const uint32_t mLGN = (kUpdate && !events) ? ff.nLGN[n] : 0u;
    const uint32_t ffBits = events ? ffMask[n] : 0u;
    float fv[nFF];
    uint32_t fid[nFF];
#pragma unroll
    for (int j = 0; j < nFF; ++j) {
        fv[j] = 0.0f;
        fid[j] = 0u;
        if ((uint32_t)j < mLGN) {
            const uint32_t lid = (uint32_t)j * nV1 + n;
            fv[j] = ff.sLGN[lid];
            fid[j] = ff.idLGN[lid];
        }
    }
    float ge[VC_MAX_NG], he[VC_MAX_NG], gi[VC_MAX_NG], hi[VC_MAX_NG];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        ge[ig] = he[ig] = gi[ig] = hi[ig] = 0.0f;
        if (ig < p.ngE) {
            ge[ig] = s.gE[ig * nV1 + n];
            he[ig] = s.hE[ig * nV1 + n];
        }
        if (ig < p.ngI) {
            gi[ig] = s.gI[ig * nV1 + n];
            hi[ig] = s.hI[ig * nV1 + n];
        }
    }

    // ---- round 2: LGN spike words, then the recurrent gather
    float sv[nFF];
#pragma unroll
    for (int j = 0; j < nFF; ++j) sv[j] = ((uint32_t)j < mLGN) ? lgnSInfo[fid[j]] : 0.0f;
    // event path: the first kEv synapses that received a spike this step
    constexpr int nEv = events ? kEv : 1;
    uint32_t evj[nEv];
    float evf[nEv];
    uint32_t evid[nEv];
    {
        uint32_t bits = ffBits;
#pragma unroll
        for (int k = 0; k < nEv; ++k) {
            evj[k] = 0xffffffffu;
            evf[k] = 0.0f;
            evid[k] = 0u;
            if (bits) {
                const uint32_t j = (uint32_t)__builtin_ctz(bits);
                bits &= bits - 1u;
                const uint32_t lid = j * nV1 + n;
                evj[k] = j;
                evf[k] = ff.sLGN[lid];
                evid[k] = ff.idLGN[lid];
            }
        }
    }

    float lgE[VC_MAX_NG] = {0.f, 0.f}, lhE[VC_MAX_NG] = {0.f, 0.f};
    float lgI[VC_MAX_NG] = {0.f, 0.f}, lhI[VC_MAX_NG] = {0.f, 0.f};
    __syncthreads();
    rec_gather(p, rec, prev, b, myseg, nSeg, nEnt, t, smk, sbid, sbrow[wv], sbslot[wv],
               lgE, lhE, lgI, lhI);

    // conductances at the end of the previous step (reference: after recal_G)
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        ge[ig] = ge[ig] + lgE[ig];
        he[ig] = he[ig] + lhE[ig];
        gi[ig] = gi[ig] + lgI[ig];
        hi[ig] = hi[ig] + lhI[ig];
    }
    if (out_snap) {                 // state at the start of the step, cortical g/h incremented
        out_snap[VC_SNAP_V * nV1 + n] = s.v[n];
        out_snap[VC_SNAP_TBACK * nV1 + n] = s.tBack[n];
        for (uint32_t ig = 0; ig < p.ngFF; ++ig) {
            out_snap[(VC_SNAP_GFF + ig) * nV1 + n] = s.gFF[ig * nV1 + n];
            out_snap[(VC_SNAP_HFF + ig) * nV1 + n] = s.hFF[ig * nV1 + n];
        }
        for (uint32_t ig = 0; ig < p.ngE; ++ig) {
            out_snap[(VC_SNAP_GE + ig) * nV1 + n] = ge[ig];
            out_snap[(VC_SNAP_HE + ig) * nV1 + n] = he[ig];
        }
        for (uint32_t ig = 0; ig < p.ngI; ++ig) {
            out_snap[(VC_SNAP_GI + ig) * nV1 + n] = gi[ig];
            out_snap[(VC_SNAP_HI + ig) * nV1 + n] = hi[ig];
        }
    }
    if (out_g_prev) {
        const uint32_t ng = p.ngE + p.ngI;
        for (uint32_t ig = 0; ig < p.ngE; ++ig) {
            out_g_prev[ig * nV1 + n] = ge[ig];
            out_g_prev[(ng + ig) * nV1 + n] = he[ig];
        }
        for (uint32_t ig = 0; ig < p.ngI; ++ig) {
            out_g_prev[(p.ngE + ig) * nV1 + n] = gi[ig];
            out_g_prev[(ng + p.ngE + ig) * nV1 + n] = hi[ig];
        }
    }
    if (!kUpdate) {
        for (uint32_t ig = 0; ig < p.ngE; ++ig) {
            s.gE[ig * nV1 + n] = ge[ig];
            s.hE[ig * nV1 + n] = he[ig];
        }
        for (uint32_t ig = 0; ig < p.ngI; ++ig) {
            s.gI[ig * nV1 + n] = gi[ig];
            s.hI[ig * nV1 + n] = hi[ig];
        }
        return;
    }

    // ---- stage d: model set-up (AdEx constructor, src/coredynamics.cuh:40-49)
    const float vR = p.vR[t], vThres = p.vThres[t], gL = p.gL[t], C = p.C[t];
    const float tRef = p.tRef[t], vT = p.vT[t], deltaT = p.deltaT[t];
    const float tau_w = p.tau_w[t], a = p.a[t], bb = p.b[t];
    float v0 = s.v[n];
    float w0 = s.w[n];
    float tBack = s.tBack[n];
    const float dep = s.tonicDep[n];
    float depC;
    {
        float targetV = VC_VL + (vT - VC_VL) * dep;
        if (targetV > vT) {
            depC = (gL + a) * (targetV - VC_VL) - gL * deltaT;
        } else if (deltaT > 0.0f) {
            depC = (gL + a) * (targetV - VC_VL) - gL * deltaT * vc_expf((targetV - vT) / deltaT);
        } else {
            depC = (gL + a) * (targetV - VC_VL) - 0.0f;
        }
    }

    float gE_t1 = 0.0f, gI_t1 = 0.0f;
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        if (ig < p.ngE) {
            float g = ge[ig], h = he[ig];
            decay_const(g, h, k.etrE[ig], k.etdE[ig], p.dodE[ig]);
            s.gE[ig * nV1 + n] = g;
            s.hE[ig * nV1 + n] = h;
            gE_t1 += g;
        }
    }
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        if (ig < p.ngI) {
            float g = gi[ig], h = hi[ig];
            decay_const(g, h, k.etrI[ig], k.etdI[ig], p.dodI[ig]);
            s.gI[ig * nV1 + n] = g;
            s.hI[ig * nV1 + n] = h;
            gI_t1 += g;
        }
    }
    float g0[VC_MAX_NG], h0[VC_MAX_NG], g1[VC_MAX_NG], h1[VC_MAX_NG];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        g0[ig] = h0[ig] = g1[ig] = h1[ig] = 0.0f;
        if (ig < p.ngFF) {
            g1[ig] = s.gFF[ig * nV1 + n];
            h1[ig] = s.hFF[ig * nV1 + n];
            g0[ig] = g1[ig];
            h0[ig] = h1[ig];
            decay_const(g1[ig], h1[ig], k.etrFF[ig], k.etdFF[ig], p.dodFF[ig]);
        }
    }
    const bool backing = (tBack < dt) && (tBack > 0.0f);
    if (out_dep) out_dep[n] = depC;
    if (backing) {
#pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig)
            if (ig < p.ngFF) decay_var(g0[ig], h0[ig], tBack, p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
    }

    // ---- stage c-FF: LGN synapses in list order; only synapses whose LGN
    // cell spiked (sInfo >= 1) do work (reference: nsp_FF > 0).
    FFAcc acc;
    FFShape shFF, shBG;
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        acc.g0[ig] = g0[ig]; acc.h0[ig] = h0[ig]; acc.g1[ig] = g1[ig]; acc.h1[ig] = h1[ig];
        const uint32_t pi = t * p.ngFF + (ig < p.ngFF ? ig : 0u);
        shFF.p[ig] = p.pFF[pi];
        shBG.p[ig] = p.pBG[pi];
        shFF.rise[ig] = shBG.rise[ig] = p.riseFF[ig];
        shFF.decay[ig] = shBG.decay[ig] = p.decayFF[ig];
        shFF.dod[ig] = shBG.dod[ig] = p.dodFF[ig];
    }
    shFF.ng = shBG.ng = p.ngFF;
    if (events && __ballot(ffBits != 0u) != 0ull) {
        // synapses whose LGN cell spiked, in synapse order
        float evs[nEv];
#pragma unroll
        for (int k = 0; k < nEv; ++k) evs[k] = (evj[k] != 0xffffffffu) ? lgnSInfo[evid[k]] : 0.0f;
#pragma unroll
        for (int k = 0; k < nEv; ++k) {
            if (evj[k] != 0xffffffffu) {
                const float sI = evs[k];
                const float nsp = (float)(uint32_t)floorf(sI);
                acc = ff_add(acc, shFF, evf[k], nsp, (sI - nsp) * dt, dt, tBack, backing);
            }
        }
        uint32_t rest = ffBits;
#pragma unroll
        for (int k = 0; k < nEv; ++k) rest &= rest - 1u;    // drop the prefetched bits
        while (rest) {
            const uint32_t j = (uint32_t)__builtin_ctz(rest);
            rest &= rest - 1u;
            const uint32_t lid = j * nV1 + n;
            const float sI = lgnSInfo[ff.idLGN[lid]];
            const float nsp = (float)(uint32_t)floorf(sI);
            acc = ff_add(acc, shFF, ff.sLGN[lid], nsp, (sI - nsp) * dt, dt, tBack, backing);
        }
        if (ffBits) ffMask[n] = 0u;
    }
    if (!events) {
        uint32_t spk = 0;
#pragma unroll
        for (int j = 0; j < nFF; ++j) spk |= (sv[j] >= 1.0f ? 1u : 0u) << j;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) spk |= (uint32_t)__shfl_xor((int)spk, off);
        spk = __builtin_amdgcn_readfirstlane(spk);     // slots that spiked for any lane
#pragma unroll
        for (int j = 0; j < nFF; ++j) {
            if ((spk >> j) & 1u) {
                const float sI = sv[j];
                if (sI >= 1.0f) {
                    const float nsp = (float)(uint32_t)floorf(sI);
                    acc = ff_add(acc, shFF, fv[j], nsp, (sI - nsp) * dt, dt, tBack, backing);
                }
            }
        }
    }
    for (uint32_t i = nFF; i < mLGN; ++i) {            // lists longer than kFF
        const uint32_t lid = i * nV1 + n;
        const float sI = lgnSInfo[ff.idLGN[lid]];
        if (sI >= 1.0f) {
            const float nsp = (float)(uint32_t)floorf(sI);
            acc = ff_add(acc, shFF, ff.sLGN[lid], nsp, (sI - nsp) * dt, dt, tBack, backing);
        }
    }
    // background Poisson drive (extension): one more input after the LGN list
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
        const uint32_t nsp = vc_poisson_inv(u, p.bgLambda[t], p.bgP0[t]);
        if (nsp > 0u)
            acc = ff_add(acc, shBG, p.bgStrength[t], (float)nsp, frac * dt, dt, tBack, backing);
    }
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        g0[ig] = acc.g0[ig]; h0[ig] = acc.h0[ig]; g1[ig] = acc.g1[ig]; h1[ig] = acc.h1[ig];
    }
    float gE_t0 = 0.0f;
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        if (ig < p.ngFF) {
            gE_t0 += g0[ig];
            gE_t1 += g1[ig];
            s.gFF[ig * nV1 + n] = g1[ig];
            s.hFF[ig * nV1 + n] = h1[ig];
        }
    }
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        if (ig < p.ngE) {
            if (backing) decay_var(ge[ig], he[ig], tBack, p.riseE[ig], p.decayE[ig], p.dodE[ig]);
            gE_t0 += ge[ig];
        }
    }
    float gI_t0 = 0.0f;
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        if (ig < p.ngI) {
            if (backing) decay_var(gi[ig], hi[ig], tBack, p.riseI[ig], p.decayI[ig], p.dodI[ig]);
            gI_t0 += gi[ig];
        }
    }

    // ---- stage d: Heun RK2 (src/coredynamics.cu:1534-1589)
    float v = v0, w = w0, tsp = 0.0f;
    bool spiked = false;
    if (tBack < dt) {
        const float a0 = gE_t0 + gI_t0 + gL + 0.0f;
        const float b0 = gE_t0 * VC_VE + gI_t0 * VC_VI + gL * VC_VL + depC + 0.0f;
        const float a1 = gE_t1 + gI_t1 + gL + 0.0f;
        const float b1 = gE_t1 * VC_VE + gI_t1 * VC_VI + gL * VC_VL + depC + 0.0f;
        const float new_dt = dt - tBack;
        if (backing) {
            const float A = a * (v0 - VC_VL);
            w0 = (w0 - A) * vc_expf(-tBack / tau_w) + A;
        }
        const float dTgL = deltaT * gL;
        float fk1, fk2;
        if (deltaT > 0.0f)
            fk1 = (-a0 * v0 + b0 + dTgL * vc_expf((v0 - vT) / deltaT) - w0) / C;
        else
            fk1 = (-a0 * v0 + b0 - w0) / C;
        const float gk1 = (a * (v0 - VC_VL) - w0) / tau_w;
        const float v1 = v0 + new_dt * fk1;
        const float w1 = w0 + new_dt * gk1;
        if (deltaT > 0.0f)
            fk2 = (-a1 * v1 + b1 + dTgL * vc_expf((v1 - vT) / deltaT) - w1) / C;
        else
            fk2 = (-a1 * v1 + b1 - w1) / C;
        const float gk2 = (a * (v1 - VC_VL) - w1) / tau_w;
        w = w0 + new_dt * (gk1 + gk2) / 2.0f;
        v = v0 + new_dt * (fk1 + fk2) / 2.0f;
        if (v > vThres) {
            const float ratio = (vThres - v0) / (v - v0);
            tsp = tBack + ratio * new_dt;
            w0 += ratio * (w - w0) + bb;
            spiked = true;
            tBack = tsp + tRef;
            v = vR;
        } else {
            if (tBack > 0.0f) tBack = 0.0f;
        }
    }
    if (tBack >= dt) {
        v = vR;
        const float A = a * (v - VC_VL);
        w = (w0 - A) * vc_expf(-(dt - tsp) / tau_w) + A;
        tBack -= dt;
    }
    const float sInfo = spiked ? 1.0f + tsp / dt : v;
    s.v[n] = v;
    if (p.iModel == 1) s.w[n] = w;
    s.tBack[n] = tBack;
    out_sInfo[n] = sInfo;

    // ---- stage e: ballot compaction (src/coredynamics.cu:1994-2046)
    const bool exist = sInfo >= 1.0f;
    const uint64_t m = __ballot(exist);
    const uint32_t tile = n >> 6;
    if (lane == 0) next.mask[tile] = m;
    if (exist) {
        const float nsp = floorf(sInfo);
        const float tsp2 = (sInfo - nsp) * dt;
        const uint32_t slot = tile * 64u + vc_mbcnt(m);
        if (local < p.nE) {
#pragma unroll
            for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
                if (ig < p.ngE) {
                    float lg = 0.0f, lh = 0.0f;
                    single_input(lg, lh, nsp, dt - tsp2, p.riseE[ig], p.decayE[ig], p.dodE[ig]);
                    next.og[ig * nV1 + slot] = lg;
                    next.oh[ig * nV1 + slot] = lh;
                }
            }
        } else {
#pragma unroll
            for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
                if (ig < p.ngI) {
                    float lg = 0.0f, lh = 0.0f;
                    single_input(lg, lh, nsp, dt - tsp2, p.riseI[ig], p.decayI[ig], p.dodI[ig]);
                    next.og[ig * nV1 + slot] = lg;
                    next.oh[ig * nV1 + slot] = lh;
                }
            }
        }
    }
}

__global__ void k_clear_masks(uint64_t *mask, uint32_t nTiles) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nTiles) mask[i] = 0ull;
}

int check_params(const vc_v1_params *p, const vc_rec_conn *rec) {
    if (!p) { vc_set_error("null params"); return VC_E_INVALID; }
    if (p->blockSize == 0 || p->blockSize % 64 != 0 || p->blockSize > 4096) {
        vc_set_error("blockSize %u must be a positive multiple of 64 (<= 4096)", p->blockSize);
        return VC_E_INVALID;
    }
    if (p->nV1 != p->nblock * p->blockSize) {
        vc_set_error("nV1 %u != nblock %u * blockSize %u", p->nV1, p->nblock, p->blockSize);
        return VC_E_INVALID;
    }
    if (p->nType < 1 || p->nType > VC_MAX_NTYPE || p->typeAcc[p->nType - 1] != p->blockSize) {
        vc_set_error("nType %u / typeAcc invalid", p->nType);
        return VC_E_INVALID;
    }
    if (p->ngFF < 1 || p->ngFF > VC_MAX_NG || p->ngE < 1 || p->ngE > VC_MAX_NG ||
        p->ngI < 1 || p->ngI > VC_MAX_NG) {
        vc_set_error("receptor counts out of range (1..%d)", VC_MAX_NG);
        return VC_E_INVALID;
    }
    if (p->nE > p->blockSize) { vc_set_error("nE > blockSize"); return VC_E_INVALID; }
    for (uint32_t t = 0; t < p->nType; ++t) {
        if (!(p->tRef[t] >= p->dt)) {
            vc_set_error("tRef[%u]=%g < dt=%g: the multi-spike-per-step variant "
                         "(compute_V_collect_spike_learnFF) is not built", t, p->tRef[t], p->dt);
            return VC_E_UNSUPPORTED;
        }
        if (p->deltaT[t] < 0.0f) { vc_set_error("deltaT < 0"); return VC_E_INVALID; }
    }
    if ((uint64_t)p->nblock * (rec ? rec->maxNear : 1) * p->blockSize >= (1ull << 32)) {
        vc_set_error("nblock * maxNear * blockSize must be < 2^32");
        return VC_E_UNSUPPORTED;
    }
    if (rec) {
        if (rec->maxNear * (p->blockSize / 64) > (uint32_t)kMaxEnt) {
            vc_set_error("maxNear %u x tiles %u exceeds the %d neighbour tiles staged in LDS",
                         rec->maxNear, p->blockSize / 64, kMaxEnt);
            return VC_E_UNSUPPORTED;
        }
    }
    return VC_OK;
}



}  // namespace

extern "C" int vc_v1_step(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                          const vc_ff_conn *ff, const vc_rec_conn *rec,
                          const vc_spikes *prev, const vc_spikes *next,
                          const float *lgnSInfo, uint32_t *ffMask, uint32_t step,
                          float *out_sInfo, float *out_dep, float *out_g_prev) {
    int rc = check_params(p, rec);
    if (rc) return rc;
    if (!s || !ff || !rec || !prev || !next || !lgnSInfo || !out_sInfo) {
        vc_set_error("vc_v1_step: null argument");
        return VC_E_INVALID;
    }
    if (prev->mask == next->mask) {
        vc_set_error("vc_v1_step: prev and next spike buffers must differ");
        return VC_E_INVALID;
    }
    if (ffMask && ff->maxLGN > 32u) {
        vc_set_error("vc_v1_step: event-driven LGN input needs maxLGN <= 32 (got %u)", ff->maxLGN);
        return VC_E_INVALID;
    }
    V1Const k = make_const(p);
    uint32_t wg = wg_size(p->blockSize);
    if (ffMask)
        hipLaunchKernelGGL((k_v1_step<true, true>), dim3(p->nV1 / wg), dim3(wg), 0,
                           (hipStream_t)stream, *p, k, *s, *ff, *rec, *prev, *next, lgnSInfo, ffMask,
                           step, out_sInfo, out_dep, out_g_prev, nullptr);
    else
        hipLaunchKernelGGL((k_v1_step<true, false>), dim3(p->nV1 / wg), dim3(wg), 0,
                           (hipStream_t)stream, *p, k, *s, *ff, *rec, *prev, *next, lgnSInfo, ffMask,
                           step, out_sInfo, out_dep, out_g_prev, nullptr);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_v1_step_snapshot(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                                   const vc_ff_conn *ff, const vc_rec_conn *rec,
                                   const vc_spikes *prev, const vc_spikes *next,
                                   const float *lgnSInfo, uint32_t *ffMask, uint32_t step,
                                   float *out_sInfo, float *out_snap) {
    int rc = check_params(p, rec);
    if (rc) return rc;
    if (!s || !ff || !rec || !prev || !next || !lgnSInfo || !out_sInfo || !out_snap) {
        vc_set_error("vc_v1_step_snapshot: null argument");
        return VC_E_INVALID;
    }
    if (prev->mask == next->mask) {
        vc_set_error("vc_v1_step_snapshot: prev and next spike buffers must differ");
        return VC_E_INVALID;
    }
    if (ffMask && ff->maxLGN > 32u) {
        vc_set_error("vc_v1_step_snapshot: event-driven LGN input needs maxLGN <= 32");
        return VC_E_INVALID;
    }
    V1Const k = make_const(p);
    uint32_t wg = wg_size(p->blockSize);
    if (ffMask)
        hipLaunchKernelGGL((k_v1_step<true, true>), dim3(p->nV1 / wg), dim3(wg), 0,
                           (hipStream_t)stream, *p, k, *s, *ff, *rec, *prev, *next, lgnSInfo, ffMask,
                           step, out_sInfo, nullptr, nullptr, out_snap);
    else
        hipLaunchKernelGGL((k_v1_step<true, false>), dim3(p->nV1 / wg), dim3(wg), 0,
                           (hipStream_t)stream, *p, k, *s, *ff, *rec, *prev, *next, lgnSInfo, ffMask,
                           step, out_sInfo, nullptr, nullptr, out_snap);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_v1_run(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                         const vc_ff_conn *ff, const vc_rec_conn *rec, const vc_spikes *spk2,
                         uint32_t cur, const float *lgnSInfo, uint64_t lgnStride,
                         uint32_t *ffMask, uint64_t maskStride, uint32_t it0,
                         uint32_t nsteps, float *out_sInfo, uint64_t outStride) {
    int rc = check_params(p, rec);
    if (rc) return rc;
    if (!s || !ff || !rec || !spk2 || !lgnSInfo || !out_sInfo || cur > 1) {
        vc_set_error("vc_v1_run: bad argument");
        return VC_E_INVALID;
    }
    if (spk2[0].mask == spk2[1].mask) {
        vc_set_error("vc_v1_run: the two spike buffers must differ");
        return VC_E_INVALID;
    }
    if (ffMask && ff->maxLGN > 32u) {
        vc_set_error("vc_v1_run: event-driven LGN input needs maxLGN <= 32 (got %u)", ff->maxLGN);
        return VC_E_INVALID;
    }
    V1Const k = make_const(p);
    const uint32_t wg = wg_size(p->blockSize);
    for (uint32_t i = 0; i < nsteps; ++i) {
        const uint32_t nx = 1u - cur;
        if (ffMask)
            hipLaunchKernelGGL((k_v1_step<true, true>), dim3(p->nV1 / wg), dim3(wg), 0,
                               (hipStream_t)stream, *p, k, *s, *ff, *rec, spk2[cur], spk2[nx],
                               lgnSInfo + i * lgnStride, ffMask + i * maskStride, it0 + i,
                               out_sInfo + i * outStride, nullptr, nullptr, nullptr);
        else
            hipLaunchKernelGGL((k_v1_step<true, false>), dim3(p->nV1 / wg), dim3(wg), 0,
                               (hipStream_t)stream, *p, k, *s, *ff, *rec, spk2[cur], spk2[nx],
                               lgnSInfo + i * lgnStride, nullptr, it0 + i,
                               out_sInfo + i * outStride, nullptr, nullptr, nullptr);
        cur = nx;
    }
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_v1_rec_apply(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                               const vc_rec_conn *rec, const vc_spikes *prev) {
    int rc = check_params(p, rec);
    if (rc) return rc;
    if (!s || !rec || !prev) {
        vc_set_error("vc_v1_rec_apply: null argument");
        return VC_E_INVALID;
    }
    V1Const k = make_const(p);
    uint32_t wg = wg_size(p->blockSize);
    vc_ff_conn ff{};
    vc_spikes nx{};
    hipLaunchKernelGGL((k_v1_step<false, false>), dim3(p->nV1 / wg), dim3(wg), 0, (hipStream_t)stream,
                       *p, k, *s, ff, *rec, *prev, nx, nullptr, nullptr, 0u, nullptr, nullptr,
                       nullptr, nullptr);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_spikes_clear(void *stream, uint32_t nV1, const vc_spikes *sp) {
    if (!sp || !sp->mask || nV1 % 64 != 0) {
        vc_set_error("vc_spikes_clear: bad arguments");
        return VC_E_INVALID;
    }
    uint32_t nT = nV1 / 64;
    hipLaunchKernelGGL(k_clear_masks, dim3((nT + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       sp->mask, nT);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}
