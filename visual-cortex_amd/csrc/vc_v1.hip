// vc_v1.hip -- stages c, d, e of the V1 simulation step on gfx950.
//
// One thread per V1 neuron, workgroups of up to 256 threads (4 wavefronts)
// inside one V1 block, so every wavefront is one 64-neuron tile.
//
//   c-rec  recurrent gather of the previous step's spikes: the workgroup
//          stages the spike masks of its neighbour blocks in LDS; each
//          wavefront lists the spikes (neighbour order, ascending neuron id)
//          into an LDS batch, fetches one masked-CSR row segment per spike
//          with one vector load (lane k <- spike k), then walks the batch in
//          order, each lane loading only the weights whose mask bit it owns
//          (rank via mbcnt).  Sums are accumulated in the reference order
//          (src/coredynamics.cu:2721-2838).
//   c-FF   LGN input, ELL/transposed like the reference (src/coredynamics.cu
//          :1457-1513), loads issued 8 synapses at a time.
//   d      decay, Heun RK2 AdEx/LIF, spike time, refractory back-up
//          (src/coredynamics.cu:1367-1589, src/coredynamics.cuh:27-118).
//   e      ballot compaction: one spike mask per tile and the unit increments
//          at slot tile*64 + rank (src/coredynamics.cu:1994-2046).
#include "vc_common.h"

namespace {

constexpr int kWG = 256;          // max workgroup size
constexpr int kMaxNearTiles = 512; // neighbour blocks x tiles per block staged in LDS
constexpr int kBatch = 64;

// per-launch constants derived on the host from vc_v1_params
struct V1Const {
    float etrFF[VC_MAX_NG], etdFF[VC_MAX_NG];
    float etrE[VC_MAX_NG], etdE[VC_MAX_NG];
    float etrI[VC_MAX_NG], etdI[VC_MAX_NG];
    uint32_t nSeg;                // blockSize / 64
};

__device__ __forceinline__ void single_input(float &g, float &h, float f, float ddt,
                                             float rise, float decay, float dod) {
    float etr = vc_expf(-ddt / rise);
    g = g + f * dod * (vc_expf(-ddt / decay) - etr);
    h = h + f * etr;
}
__device__ __forceinline__ void decay_var(float &g, float &h, float ddt, float rise,
                                          float decay, float dod) {
    float etr = vc_expf(-ddt / rise);
    float etd = vc_expf(-ddt / decay);
    g = g * etd + h * dod * (etd - etr);
    h = h * etr;
}
__device__ __forceinline__ void decay_const(float &g, float &h, float etr, float etd,
                                            float dod) {
    g = g * etd + h * dod * (etd - etr);
    h = h * etr;
}

__device__ __forceinline__ int type_of(const vc_v1_params &p, uint32_t local) {
    int t = (int)p.nType - 1;
    for (int j = (int)p.nType - 1; j >= 0; --j)
        if (local < p.typeAcc[j]) t = j;
    return t;
}

__device__ __forceinline__ uint64_t emask_of(uint32_t tile, uint32_t nE) {
    uint32_t lo = tile * 64u;
    if (nE >= lo + 64u) return ~0ull;
    if (nE <= lo) return 0ull;
    return (1ull << (nE - lo)) - 1ull;
}

// Recurrent increments of the spikes in `prev` for the calling neuron.
// Must be called by all threads of the workgroup (barriers inside).
__device__ __forceinline__ void rec_gather(const vc_v1_params &p, const V1Const &k,
                                           const vc_rec_conn &rec, const vc_spikes &prev,
                                           uint32_t b, uint32_t local, int t,
                                           float lgE[VC_MAX_NG], float lhE[VC_MAX_NG],
                                           float lgI[VC_MAX_NG], float lhI[VC_MAX_NG]) {
    __shared__ uint64_t smask[kMaxNearTiles];
    __shared__ uint32_t sbid[64];
    __shared__ uint32_t brow[kWG / 64][2 * kBatch];
    __shared__ uint32_t bslot[kWG / 64][2 * kBatch];

    const uint32_t nTile = k.nSeg;
    const uint32_t nb = rec.nNear[b];
    const uint32_t lane = vc_lane();
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t myseg = local >> 6;
    const uint32_t bs = p.blockSize;

    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
        sbid[i] = rec.nbr[b * rec.maxNear + i];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb * nTile; i += blockDim.x) {
        uint32_t ib = i / nTile, tt = i % nTile;
        smask[i] = prev.mask[sbid[ib] * nTile + tt];
    }
    __syncthreads();

    uint32_t cnt = 0;   // entries in this wavefront's batch (uniform)

    auto flush = [&](uint32_t K) {
        // lane kk loads row segment header and increments of spike kk
        uint64_t rm = 0;
        uint32_t off = 0;
        float og0 = 0.f, oh0 = 0.f, og1 = 0.f, oh1 = 0.f;
        uint32_t isI = 0;
        if (lane < K) {
            uint32_t row = brow[wv][lane];
            uint32_t sw = bslot[wv][lane];
            isI = sw >> 31;
            uint32_t slot = sw & 0x7fffffffu;
            size_t hid = (size_t)row * nTile + myseg;
            rm = rec.rowMask[hid];
            off = rec.rowOff[hid];
            og0 = prev.og[slot];
            oh0 = prev.oh[slot];
            og1 = prev.og[p.nV1 + slot];
            oh1 = prev.oh[p.nV1 + slot];
        }
        for (uint32_t k0 = 0; k0 < K; k0 += 8) {
            float wv8[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                uint32_t kk = k0 + j;
                wv8[j] = 0.0f;
                if (kk < K) {
                    const uint64_t rmk = vc_readlane_u64(rm, kk);
                    const uint32_t offk = vc_readlane_u32(off, kk);
                    const uint32_t idx = offk + vc_mbcnt(rmk);
                    if ((rmk >> lane) & 1ull) wv8[j] = rec.w[idx];
                }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                uint32_t kk = k0 + j;
                if (kk < K) {
                    // wave-uniform values of spike kk, read before any divergence
                    const uint64_t rmk = vc_readlane_u64(rm, kk);
                    const float g0 = vc_readlane_f32(og0, kk), h0 = vc_readlane_f32(oh0, kk);
                    const float g1 = vc_readlane_f32(og1, kk), h1 = vc_readlane_f32(oh1, kk);
                    const uint32_t inh = vc_readlane_u32(isI, kk);
                    if ((rmk >> lane) & 1ull) {
                        float strength = wv8[j];
                        if (inh == 0u) {
                            float str = strength * p.pE[t * p.ngE + 0];
                            lgE[0] += str * g0;
                            lhE[0] += str * h0;
                            if (p.ngE > 1) {
                                str = strength * p.pE[t * p.ngE + 1];
                                lgE[1] += str * g1;
                                lhE[1] += str * h1;
                            }
                        } else {
                            float str = strength * p.pI[t * p.ngI + 0];
                            lgI[0] += str * g0;
                            lhI[0] += str * h0;
                            if (p.ngI > 1) {
                                str = strength * p.pI[t * p.ngI + 1];
                                lgI[1] += str * g1;
                                lhI[1] += str * h1;
                            }
                        }
                    }
                }
            }
        }
    };

    for (uint32_t ib = 0; ib < nb; ++ib) {
        const uint32_t bid = sbid[ib];
        const uint32_t rowbase = (b * rec.maxNear + ib) * bs;
        for (uint32_t part = 0; part < 2; ++part) {          // E spikes, then I
            for (uint32_t tt = 0; tt < nTile; ++tt) {
                const uint64_t m = smask[ib * nTile + tt];
                if (m == 0ull) continue;
                const uint64_t em = emask_of(tt, p.nE);
                const uint64_t mm = part == 0 ? (m & em) : (m & ~em);
                if (mm == 0ull) continue;
                const uint32_t c = (uint32_t)__builtin_popcountll(mm);
                if ((mm >> lane) & 1ull) {
                    uint32_t pos = cnt + vc_mbcnt(mm);
                    uint32_t rank = vc_mbcnt(m);
                    brow[wv][pos] = rowbase + tt * 64u + lane;
                    bslot[wv][pos] = (bid * bs + tt * 64u + rank) | (part << 31);
                }
                cnt += c;
                if (cnt >= (uint32_t)kBatch) {
                    __builtin_amdgcn_wave_barrier();
                    flush(kBatch);
                    __builtin_amdgcn_wave_barrier();
                    uint32_t rest = cnt - kBatch;
                    if (lane < rest) {
                        brow[wv][lane] = brow[wv][kBatch + lane];
                        bslot[wv][lane] = bslot[wv][kBatch + lane];
                    }
                    __builtin_amdgcn_wave_barrier();
                    cnt = rest;
                }
            }
        }
    }
    if (cnt > 0) {
        __builtin_amdgcn_wave_barrier();
        flush(cnt);
    }
}

template <bool kUpdate>
__global__ __launch_bounds__(kWG) void k_v1_step(vc_v1_params p, V1Const k, vc_v1_state s,
                                                 vc_ff_conn ff, vc_rec_conn rec,
                                                 vc_spikes prev, vc_spikes next,
                                                 const float *__restrict__ lgnSInfo,
                                                 uint32_t step, float *__restrict__ out_sInfo,
                                                 float *__restrict__ out_dep,
                                                 float *__restrict__ out_g_prev) {
    const uint32_t nV1 = p.nV1;
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t b = n / p.blockSize;
    const uint32_t local = n - b * p.blockSize;
    const int t = type_of(p, local);
    const float dt = p.dt;

    float lgE[VC_MAX_NG] = {0.f, 0.f}, lhE[VC_MAX_NG] = {0.f, 0.f};
    float lgI[VC_MAX_NG] = {0.f, 0.f}, lhI[VC_MAX_NG] = {0.f, 0.f};
    rec_gather(p, k, rec, prev, b, local, t, lgE, lhE, lgI, lhI);

    // conductances at the end of the previous step (reference: after recal_G)
    float ge[VC_MAX_NG], he[VC_MAX_NG], gi[VC_MAX_NG], hi[VC_MAX_NG];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        if (ig < p.ngE) {
            ge[ig] = s.gE[ig * nV1 + n] + lgE[ig];
            he[ig] = s.hE[ig * nV1 + n] + lhE[ig];
        }
        if (ig < p.ngI) {
            gi[ig] = s.gI[ig * nV1 + n] + lgI[ig];
            hi[ig] = s.hI[ig * nV1 + n] + lhI[ig];
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

    // ---- stage c-FF: LGN synapses (transposed ELL), 8 loads in flight
    const uint32_t mLGN = ff.nLGN[n];
    for (uint32_t i0 = 0; i0 < mLGN; i0 += 8) {
        float fv[8], sv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            uint32_t i = i0 + j;
            fv[j] = 0.0f;
            sv[j] = 0.0f;
            if (i < mLGN) {
                uint32_t lid = i * nV1 + n;
                fv[j] = ff.sLGN[lid];
                sv[j] = lgnSInfo[ff.idLGN[lid]];
            }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (i0 + j < mLGN) {
                float sI = sv[j];
                uint32_t nsp = (uint32_t)floorf(sI);
                if (nsp > 0u) {
                    float tsp = (sI - (float)nsp) * dt;
                    float f = fv[j];
                    if (f > 0.0f) {
                        float ddt = 0.0f;
                        if (backing) ddt = tBack - tsp;
#pragma unroll
                        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
                            if (ig < p.ngFF) {
                                float str = f * p.pFF[t * p.ngFF + ig];
                                if (backing && ddt > 0.0f)
                                    single_input(g0[ig], h0[ig], str * (float)nsp, ddt,
                                                 p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
                                single_input(g1[ig], h1[ig], str * (float)nsp, dt - tsp,
                                             p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
                            }
                        }
                    }
                }
            }
        }
    }
    // background Poisson drive (extension)
    if (p.bgStrength[t] > 0.0f) {
        vc_u32x4 r = vc_draw4(p.seed_lo, p.seed_hi, n, step, VC_STREAM_V1_BG, 0u, p.trial);
        float u = vc_u01_open0(r.v[0]);
        float frac = vc_u01_open1(r.v[1]);
        uint32_t nsp = vc_poisson_inv(u, p.bgLambda[t], p.bgP0[t]);
        if (nsp > 0u) {
            float tsp = frac * dt;
            float f = p.bgStrength[t];
            float ddt = 0.0f;
            if (backing) ddt = tBack - tsp;
#pragma unroll
            for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
                if (ig < p.ngFF) {
                    float str = f * p.pBG[t * p.ngFF + ig];
                    if (backing && ddt > 0.0f)
                        single_input(g0[ig], h0[ig], str * (float)nsp, ddt,
                                     p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
                    single_input(g1[ig], h1[ig], str * (float)nsp, dt - tsp,
                                 p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
                }
            }
        }
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
            float A = a * (v0 - VC_VL);
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
            float ratio = (vThres - v0) / (v - v0);
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
        float A = a * (v - VC_VL);
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
    if (vc_lane() == 0) next.mask[tile] = m;
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
    if (rec) {
        if (rec->maxNear * (p->blockSize / 64) > (uint32_t)kMaxNearTiles || rec->maxNear > 64) {
            vc_set_error("maxNear %u x tiles %u exceeds the LDS staging limit %d",
                         rec->maxNear, p->blockSize / 64, kMaxNearTiles);
            return VC_E_UNSUPPORTED;
        }
    }
    return VC_OK;
}

V1Const make_const(const vc_v1_params *p) {
    V1Const k;
    for (int ig = 0; ig < VC_MAX_NG; ++ig) {
        k.etrFF[ig] = vc_expf(-p->dt / p->riseFF[ig]);
        k.etdFF[ig] = vc_expf(-p->dt / p->decayFF[ig]);
        k.etrE[ig] = vc_expf(-p->dt / p->riseE[ig]);
        k.etdE[ig] = vc_expf(-p->dt / p->decayE[ig]);
        k.etrI[ig] = vc_expf(-p->dt / p->riseI[ig]);
        k.etdI[ig] = vc_expf(-p->dt / p->decayI[ig]);
    }
    k.nSeg = p->blockSize / 64;
    return k;
}

uint32_t wg_size(uint32_t blockSize) {
    if (blockSize % 256 == 0) return 256;
    if (blockSize % 128 == 0) return 128;
    return 64;
}

}  // namespace

extern "C" int vc_v1_step(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                          const vc_ff_conn *ff, const vc_rec_conn *rec,
                          const vc_spikes *prev, const vc_spikes *next,
                          const float *lgnSInfo, uint32_t step, float *out_sInfo,
                          float *out_dep, float *out_g_prev) {
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
    V1Const k = make_const(p);
    uint32_t wg = wg_size(p->blockSize);
    hipLaunchKernelGGL(k_v1_step<true>, dim3(p->nV1 / wg), dim3(wg), 0, (hipStream_t)stream,
                       *p, k, *s, *ff, *rec, *prev, *next, lgnSInfo, step, out_sInfo, out_dep,
                       out_g_prev);
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
    hipLaunchKernelGGL(k_v1_step<false>, dim3(p->nV1 / wg), dim3(wg), 0, (hipStream_t)stream,
                       *p, k, *s, ff, *rec, *prev, nx, nullptr, 0u, nullptr, nullptr, nullptr);
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
