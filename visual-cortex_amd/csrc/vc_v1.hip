// vc_v1.hip -- stages c, d, e of the V1 simulation step on gfx950.
//
// One thread per V1 neuron, workgroups of up to 256 threads inside one V1
// block, so every wavefront is one 64-neuron tile.  The kernel is latency
// bound at the full-patch size (one wavefront per SIMD), so it is organised
// around few dependent memory round trips:
//   trip 1: state, the LGN synapse indices/strengths of the neuron, and the
//            spike masks of the 3x3 neighbour blocks (one 64-bit mask per
//            tile, held lane-parallel: lane l <- tile entry l);
//   trip 2: the LGN spike words, and for the spikes of the neighbourhood
//            (listed in neighbour order, ascending neuron id, by a ballot
//            scan over the non-empty tiles) one masked-CSR row-segment header
//            per spike (lane k <- spike k);
//   trip 3: the synaptic weights each lane owns (rank via mbcnt).
// Sums are accumulated in the reference order: recurrent E and I sums over
// neighbour blocks then ascending presynaptic id (src/coredynamics.cu
// :2721-2838), LGN synapses in list order (src/coredynamics.cu:1466-1502).
//   stage d: decay, Heun RK2 AdEx/LIF, spike time, refractory back-up
//            (src/coredynamics.cu:1367-1589, src/coredynamics.cuh:27-118);
//   stage e: ballot compaction -- one spike mask per tile and the unit
//            increments at slot tile*64 + rank (src/coredynamics.cu:1994-2046).
#include "vc_v1_device.h"

#include <mutex>

namespace {

constexpr int kWG = 256;           // max workgroup size
constexpr int kMaxEnt = 512;       // neighbour-tile entries staged in LDS per wavefront
constexpr int kBatch = 64;
constexpr int kFF = 32;            // LGN synapses loaded per round
// The recurrent gather's weight layout and tile, a template parameter kG of
// the kernels: 0 = dense rows (kWD strengths in flight per lane), else the
// masked CSR with kG strengths in flight per lane (4, 8 or 16).  A compile-
// time choice: the dense path's 32 loads in flight would otherwise set the
// register count of every kernel.
constexpr int kWD = 32;
constexpr uint32_t kTileStep = 4;  // default kG of the one-step kernel
constexpr uint32_t kTileMulti = 8; // kG of the persistent kernel
constexpr int kEv = 4;             // LGN spike events prefetched per neuron



// LGN (or background) conductance input of one synapse, src/coredynamics.cu
// :1474-1500: nsp spikes at tsp into (g1, h1) over dt - tsp and, when the
// neuron is backing up from refractoriness, into (g0, h0) over tBack - tsp.
// pm[ig] is the receptor split of the input's kind (LGN or background).
struct FFAcc {
    float g0[VC_MAX_NG], h0[VC_MAX_NG], g1[VC_MAX_NG], h1[VC_MAX_NG];
};
struct FFShape {
    float p[VC_MAX_NG], rise[VC_MAX_NG], decay[VC_MAX_NG], dod[VC_MAX_NG];
    uint32_t ng;
};
__device__ __forceinline__ void ff_apply(FFAcc &a, const FFShape &sh, const float pm[VC_MAX_NG],
                                         float f, float nsp, float tsp, float dt, float tBack,
                                         bool backing) {
    if (f > 0.0f) {
        float ddt = 0.0f;
        if (backing) ddt = tBack - tsp;
#pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
            if (ig < sh.ng) {
                const float str = f * pm[ig];
                if (backing && ddt > 0.0f)
                    single_input(a.g0[ig], a.h0[ig], str * nsp, ddt, sh.rise[ig], sh.decay[ig],
                                 sh.dod[ig]);
                single_input(a.g1[ig], a.h1[ig], str * nsp, dt - tsp, sh.rise[ig], sh.decay[ig],
                             sh.dod[ig]);
            }
        }
    }
}

// Loads and stores of the spike buffers.  Within one launch of the
// persistent kernel they are exchanged between workgroups that may sit on
// different XCDs (each with its own L2): kCoh selects agent-coherent
// accesses (relaxed agent-scope atomics, which bypass or write through the
// XCD's L2), otherwise ordinary ones (the data was written by an earlier
// launch).
template <bool kCoh, typename T>
__device__ __forceinline__ T ld_spk(const T *p) {
    if (kCoh) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return *p;
}
template <bool kCoh, typename T>
__device__ __forceinline__ void st_spk(T *p, T v) {
    if (kCoh) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

// Spike records of the persistent kernel.  Every 64-bit word carries a
// 32-bit value and, in its high half, the tag of the step that wrote it, so a
// reader needs no flag: it polls the words it needs until their tag is the
// step it expects (each word is written and read whole).  Per 64-neuron tile
// and step: a record of kRecW words, [0]/[1] the low/high half of the tile's
// spike mask and [2 + 4r + j] (r < kRecSpk) the increments og0, oh0, og1, oh1
// (j = 0..3) of its spike of rank r; spikes of rank >= kRecSpk use the
// overflow words [(tile*64 + rank)*4 + j].
constexpr int kRecSpk = 3;
constexpr int kRecW = 16;                    // 64-bit words per record
constexpr uint32_t kSpinMax = 1u << 20;      // bound of every wait

__device__ __forceinline__ uint64_t tagged(uint32_t v, uint32_t tag) {
    return ((uint64_t)tag << 32) | v;
}
// The values of N tagged words p[0..N-1] once each carries `tag`: all N
// loads are issued together, then only the words still carrying another tag
// are loaded again (bounded wait; on a time-out *status is set and the stale
// values returned).
template <int N>
__device__ __forceinline__ void ld_tagged(const uint64_t *p, uint32_t tag, uint32_t *status,
                                          uint32_t out[N]) {
    uint64_t v[N];
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = ld_spk<true>(p + j);
    for (uint32_t spin = 0;; ++spin) {
        bool ok = true;
#pragma unroll
        for (int j = 0; j < N; ++j) ok = ok && (uint32_t)(v[j] >> 32) == tag;
        if (ok) break;
        if (spin >= kSpinMax || ((spin & 255u) == 255u && ld_spk<true>(status) != 0u)) {
            atomicOr(status, 1u);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int j = 0; j < N; ++j)
            if ((uint32_t)(v[j] >> 32) != tag) v[j] = ld_spk<true>(p + j);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] = (uint32_t)v[j];
}

// Philox key of neuron n (block b) and its entity id: with batched trials the
// trial's own key and the neuron's id within its trial (off = the trial's
// first neuron), so a trial draws the same numbers batched as alone.
struct V1Key {
    uint32_t lo, hi, ent, off;
};
__device__ __forceinline__ V1Key v1_key(const vc_v1_params &p, uint32_t b, uint32_t n) {
    V1Key k{p.seed_lo, p.seed_hi, n, 0u};
    if (p.blockTrial != nullptr) {
        const uint32_t tr = p.blockTrial[b];
        k.lo = p.trialSeed[2 * tr];
        k.hi = p.trialSeed[2 * tr + 1];
        k.off = p.trialV1Off[tr];
        k.ent = n - k.off;
    }
    return k;
}

// Synaptic failure: the strength f of a transmitting synapse times
// 1 - p + sqrt(p (1 - p) / nSyn) * N(0,1) (src/coredynamics.cu:1476-1480,
// 2766-2769), the normal drawn from (key, step, stream, pack).
__device__ __forceinline__ float fail_factor(float f, float pf, float nSyn, const V1Key &k,
                                             uint32_t step, uint32_t stream, uint32_t pack,
                                             uint32_t trial) {
    const float sd = __builtin_sqrtf(pf * (1.0f - pf) / nSyn);
    const vc_u32x4 r = vc_draw4(k.lo, k.hi, k.ent, step, stream, pack, trial);
    return f * (1.0f - pf + sd * vc_normal(r.v[0], r.v[1]));
}

// Adds one spike's increments (strength w) to the calling lane's sums in the
// reference's order.
__device__ __forceinline__ void add_spike(const vc_v1_params &p, float strength, uint32_t inh,
                                          float pE0, float pE1, float pI0, float pI1, float g0,
                                          float h0, float g1, float h1, float lgE[VC_MAX_NG],
                                          float lhE[VC_MAX_NG], float lgI[VC_MAX_NG],
                                          float lhI[VC_MAX_NG]) {
    // Selects on values, not on which array is updated: a branch between the
    // E and I sums is otherwise merged into a select between the arrays'
    // addresses, which puts them in scratch memory.
    const bool e = inh == 0u;
    const float str0 = strength * (e ? pE0 : pI0);
    const float a0 = str0 * g0, b0 = str0 * h0;
    lgE[0] = e ? lgE[0] + a0 : lgE[0];
    lhE[0] = e ? lhE[0] + b0 : lhE[0];
    lgI[0] = e ? lgI[0] : lgI[0] + a0;
    lhI[0] = e ? lhI[0] : lhI[0] + b0;
    const bool e1 = e && p.ngE > 1, i1 = !e && p.ngI > 1;
    if (e1 || i1) {
        const float str1 = strength * (e ? pE1 : pI1);
        const float a1 = str1 * g1, b1 = str1 * h1;
        lgE[1] = e1 ? lgE[1] + a1 : lgE[1];
        lhE[1] = e1 ? lhE[1] + b1 : lhE[1];
        lgI[1] = i1 ? lgI[1] + a1 : lgI[1];
        lhI[1] = i1 ? lhI[1] + b1 : lhI[1];
    }
}

// Spike records being read (persistent kernel): the tagged words and the tag.
struct RecIn {
    const uint64_t *words;        // scratch base
    uint32_t rec, ovf;            // word offsets of the record and overflow buffers read
    uint32_t tag;
    uint32_t *status;
};

// One wavefront's copy of a batch's row-segment headers and increments,
// read by the lanes the spikes reach.
struct WaveStage {
    uint32_t rmLo[kBatch], rmHi[kBatch], off[kBatch];
    float4 inc[kBatch];
};

// Transpose of the 64 x 64 bit matrix whose row k is lane k's word: on
// return lane l holds column l (bit k set when bit l of lane k's word was).
// Six butterfly stages exchange blocks of 32, 16, ..., 1 bits between lanes
// k and k ^ j.
__device__ __forceinline__ uint64_t transpose64(uint64_t a, uint32_t lane) {
    constexpr uint64_t kMask[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull,
                                   0x00FF00FF00FF00FFull, 0x0F0F0F0F0F0F0F0Full,
                                   0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const uint32_t j = 32u >> s;
        const uint32_t plo = (uint32_t)__shfl_xor((int)(uint32_t)a, (int)j);
        const uint32_t phi = (uint32_t)__shfl_xor((int)(uint32_t)(a >> 32), (int)j);
        const uint64_t b = ((uint64_t)phi << 32) | plo;
        const bool low = (lane & j) == 0u;
        const uint64_t lo = low ? a : b, hi = low ? b : a;
        const uint64_t t = ((lo >> j) ^ hi) & kMask[s];
        a = low ? (lo ^ (t << j)) : (hi ^ t);
    }
    return a;
}

// Adds the recurrent increments of a batch of K <= 64 spikes listed in
// brow/bslot (this wavefront's LDS rows) to the calling lane's sums.  With
// kRec the increments of spike k are the four tagged words at ri.words +
// bsrc[k]; otherwise they are read from prev.og/oh.  kSt: synaptic failure
// may be on (p.synFail); without it the failure draws are compiled out.
template <bool kCoh, bool kRec, int kG, bool kSt>
__device__ __forceinline__ void flush_batch(const vc_v1_params &p, const vc_rec_conn &rec,
                                            const vc_spikes &prev, uint32_t K, uint32_t lane,
                                            uint32_t nSeg, uint32_t myseg, int t,
                                            const uint32_t *brow, const uint32_t *bslot,
                                            const uint32_t *bsrc, const RecIn &ri,
                                            const V1Key &key, uint32_t step, WaveStage &ws,
                                            float lgE[VC_MAX_NG], float lhE[VC_MAX_NG],
                                            float lgI[VC_MAX_NG], float lhI[VC_MAX_NG]) {
    // lane kk fetches the row-segment header and the increments of spike kk
    constexpr bool dense = kG == 0;
    constexpr int kW = kG > 0 ? kG : 1;
    uint64_t rm = 0;
    uint32_t off = 0, isI = 0, row = 0, slot = 0;
    float og0 = 0.f, oh0 = 0.f, og1 = 0.f, oh1 = 0.f;
    if (lane < K) {
        row = brow[lane];
        const uint32_t sw = bslot[lane];
        isI = sw >> 31;
        slot = sw & 0x7fffffffu;
        if (!dense) {
            const size_t hid = (size_t)row * nSeg + myseg;
            rm = rec.rowMask[hid];
            off = rec.rowOff[hid];
        }
    }
    // the increments of spike `lane` (after the dense weights are in flight:
    // with kRec they may have to wait for their tags)
    auto load_increments = [&]() {
        if (lane >= K) return;
        if (kRec) {
            uint32_t u[4];
            ld_tagged<4>(ri.words + bsrc[lane], ri.tag, ri.status, u);
            og0 = __uint_as_float(u[0]);
            oh0 = __uint_as_float(u[1]);
            og1 = __uint_as_float(u[2]);
            oh1 = __uint_as_float(u[3]);
        } else {
            og0 = ld_spk<kCoh>(prev.og + slot);
            oh0 = ld_spk<kCoh>(prev.oh + slot);
            og1 = ld_spk<kCoh>(prev.og + p.nV1 + slot);
            oh1 = ld_spk<kCoh>(prev.oh + p.nV1 + slot);
        }
    };
    const float pE0 = p.pE[t * p.ngE + 0], pE1 = p.pE[t * p.ngE + (p.ngE > 1 ? 1 : 0)];
    const float pI0 = p.pI[t * p.ngI + 0], pI1 = p.pI[t * p.ngI + (p.ngI > 1 ? 1 : 0)];
    // synaptic failure of the spike kk's synapse onto this lane's neuron
    // (presynaptic type: E or I by the spike's flag), 1 when off
    const float pfE = p.synFail[0 * p.nType + t], pfI = p.synFail[(p.nType - 1u) * p.nType + t];
    const float nsE = p.synPerCon[0 * p.nType + t], nsI = p.synPerCon[(p.nType - 1u) * p.nType + t];
    const uint32_t bs = nSeg * 64u;
    auto failed = [&](float w, uint32_t kk) -> float {
        if constexpr (!kSt) return w;
        const uint32_t inh = vc_readlane_u32(isI, kk);
        const float pf = inh ? pfI : pfE;
        if (!(pf > 0.0f)) return w;
        const uint32_t pre = (vc_readlane_u32(slot, kk) / bs) * bs + vc_readlane_u32(row, kk) % bs;
        return fail_factor(w, pf, inh ? nsI : nsE, key, step, VC_STREAM_V1_FAIL, pre - key.off,
                           p.trial);
    };
    if constexpr (dense) {
        // a spike's strengths for this wavefront are one coalesced 256-byte
        // row segment; absent synapses hold +0 and add +0 to the sums
        const float *wrow = rec.wDense + (size_t)myseg * 64u + lane;
        const size_t bs = (size_t)nSeg * 64u;
        for (uint32_t k0 = 0; k0 < K; k0 += kWD) {
            float wv[kWD];
#pragma unroll
            for (int j = 0; j < kWD; ++j) {
                const uint32_t kk = k0 + j;
                wv[j] = kk < K ? wrow[(size_t)vc_readlane_u32(row, kk) * bs] : 0.0f;
            }
            if (k0 == 0) load_increments();
#pragma unroll
            for (int j = 0; j < kWD; ++j) {
                const uint32_t kk = k0 + j;
                if (kk < K && wv[j] > 0.0f) wv[j] = failed(wv[j], kk);
                if (kk < K && wv[j] > 0.0f)
                    add_spike(p, wv[j], vc_readlane_u32(isI, kk), pE0, pE1, pI0, pI1,
                              vc_readlane_f32(og0, kk), vc_readlane_f32(oh0, kk),
                              vc_readlane_f32(og1, kk), vc_readlane_f32(oh1, kk), lgE, lhE, lgI,
                              lhI);
            }
        }
        return;
    }
    load_increments();
    // Masked CSR: each lane visits only the spikes whose row segment has a
    // strength for its own post.  Lane k stages spike k's header and
    // increments; the transpose of the batch's K x 64 segment masks gives
    // every lane the set of spikes that reach it, walked in ascending order
    // (the order of the sums), kW strengths in flight at a time.
    if (lane < K) {
        ws.rmLo[lane] = (uint32_t)rm;
        ws.rmHi[lane] = (uint32_t)(rm >> 32);
        ws.off[lane] = off;
        ws.inc[lane] = make_float4(og0, oh0, og1, oh1);
    }
    uint64_t col = transpose64(lane < K ? rm : 0ull, lane);
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    while (__ballot(col != 0ull) != 0ull) {
        uint32_t kk[kW];
        float wv[kW];
#pragma unroll
        for (int j = 0; j < kW; ++j) {
            kk[j] = 64u;
            wv[j] = 0.0f;
            if (col != 0ull) {
                const uint32_t k = (uint32_t)__builtin_ctzll(col);
                col &= col - 1ull;
                kk[j] = k;
                const uint64_t rmk = ((uint64_t)ws.rmHi[k] << 32) | ws.rmLo[k];
                wv[j] = rec.w[ws.off[k] + vc_mbcnt(rmk)];
            }
        }
#pragma unroll
        for (int j = 0; j < kW; ++j) {
            const uint32_t k = kk[j];
            if (k < 64u) {
                const uint32_t sw = bslot[k];
                const uint32_t inh = sw >> 31;
                float w = wv[j];
                if constexpr (kSt) {
                    const float pf = inh ? pfI : pfE;
                    if (pf > 0.0f) {
                        const uint32_t pre = ((sw & 0x7fffffffu) / bs) * bs + brow[k] % bs;
                        w = fail_factor(w, pf, inh ? nsI : nsE, key, step, VC_STREAM_V1_FAIL,
                                        pre - key.off, p.trial);
                    }
                }
                if (w > 0.0f) {
                    const float4 in = ws.inc[k];
                    add_spike(p, w, inh, pE0, pE1, pI0, pI1, in.x, in.y, in.z, in.w, lgE, lhE,
                              lgI, lhI);
                }
            }
        }
    }
    __builtin_amdgcn_wave_barrier();        // the stage is rewritten by the next batch
}

// LDS of the workgroup's spike list (see rec_gather).
constexpr int kList = 512;                   // spikes listed per pass
struct GatherLds {
    uint32_t off[kMaxEnt];                   // list position of entry e's first spike
    uint32_t wsum[kWG / 64];
    uint32_t row[kList], slot[kList], src[kList];
    WaveStage ws[kWG / 64];                  // one per wavefront (flush_batch)
};

// Recurrent increments of the spikes of `prev` for the calling neuron.
// smk/sbid (LDS) hold the spike mask and block id of neighbour-tile entry
// e = ib*nSeg + tt, e < nEnt.  The workgroup lists the spikes once, in entry
// order and then ascending neuron id (popcounts, a block-wide prefix sum,
// then every thread writes its entries' spikes), and every wavefront adds
// the listed spikes' increments for its own 64 posts, 64 spikes per batch.
// Lists longer than kList are done in several passes.  Every thread of the
// workgroup must call it.
template <bool kCoh, bool kRec, int kG, bool kSt>
__device__ __forceinline__ void rec_gather(const vc_v1_params &p, const vc_rec_conn &rec,
                                        const vc_spikes &prev, uint32_t b, uint32_t myseg,
                                        uint32_t nSeg, uint32_t nEnt, int t,
                                        const uint64_t *smk, const uint32_t *sbid,
                                        const RecIn &ri, GatherLds &g, const V1Key &key,
                                        uint32_t step,
                                        float lgE[VC_MAX_NG], float lhE[VC_MAX_NG],
                                        float lgI[VC_MAX_NG], float lhI[VC_MAX_NG]) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6;
    const uint32_t nw = blockDim.x >> 6;
    const uint32_t bs = p.blockSize;
    // each thread owns a run of `per` consecutive entries
    const uint32_t per = (nEnt + blockDim.x - 1u) / blockDim.x;
    const uint32_t eb = tid * per;
    uint32_t mine = 0;
    for (uint32_t i = 0; i < per; ++i)
        if (eb + i < nEnt) mine += (uint32_t)__builtin_popcountll(smk[eb + i]);
    uint32_t incl = mine;                                   // inclusive scan in the wavefront
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == 63u) g.wsum[wv] = incl;
    __syncthreads();
    uint32_t base = 0, total = 0;
    for (uint32_t w = 0; w < nw; ++w) {
        base += w < wv ? g.wsum[w] : 0u;
        total += g.wsum[w];
    }
    {
        uint32_t o = base + incl - mine;
        for (uint32_t i = 0; i < per; ++i)
            if (eb + i < nEnt) {
                g.off[eb + i] = o;
                o += (uint32_t)__builtin_popcountll(smk[eb + i]);
            }
    }
    for (uint32_t L0 = 0; L0 < total; L0 += (uint32_t)kList) {
        // list the spikes at positions [L0, L0 + kList)
        for (uint32_t i = 0; i < per; ++i) {
            const uint32_t e = eb + i;
            if (e >= nEnt) break;
            uint64_t m = smk[e];
            const uint32_t o = g.off[e];
            if (m == 0ull || o >= L0 + (uint32_t)kList || o + (uint32_t)__builtin_popcountll(m) <= L0)
                continue;
            const uint32_t bid = sbid[e];
            const uint32_t ib = e / nSeg, tt = e - ib * nSeg;
            for (uint32_t rank = 0; m != 0ull; ++rank) {
                const uint32_t l = (uint32_t)__builtin_ctzll(m);
                m &= m - 1ull;
                const uint32_t pos = o + rank;
                if (pos < L0 || pos >= L0 + (uint32_t)kList) continue;
                const uint32_t pre = tt * 64u + l;
                g.row[pos - L0] = (b * rec.maxNear + ib) * bs + pre;
                g.slot[pos - L0] = (bid * bs + tt * 64u + rank) | ((pre >= p.nE ? 1u : 0u) << 31);
                if (kRec) {
                    const uint32_t tile = bid * nSeg + tt;
                    g.src[pos - L0] = rank < (uint32_t)kRecSpk
                                          ? ri.rec + tile * kRecW + 2u + 4u * rank
                                          : ri.ovf + (tile * 64u + rank) * 4u;
                }
            }
        }
        __syncthreads();
        const uint32_t K = (total - L0) < (uint32_t)kList ? (total - L0) : (uint32_t)kList;
        for (uint32_t k0 = 0; k0 < K; k0 += (uint32_t)kBatch) {
            const uint32_t kb = (K - k0) < (uint32_t)kBatch ? (K - k0) : (uint32_t)kBatch;
            flush_batch<kCoh, kRec, kG, kSt>(p, rec, prev, kb, lane, nSeg, myseg, t, g.row + k0,
                                    g.slot + k0, g.src + k0, ri, key, step, g.ws[wv], lgE, lhE,
                                    lgI, lhI);
        }
        __syncthreads();                    // the list is rewritten by the next pass
    }
}

// ---------------------------------------------------- per-neuron pieces --
// Shared by the one-step kernel and the persistent multi-step kernel.

struct NState {                      // one neuron's state between steps
    float v, w, tBack;
    float gFF[VC_MAX_NG], hFF[VC_MAX_NG];
    float gE[VC_MAX_NG], hE[VC_MAX_NG], gI[VC_MAX_NG], hI[VC_MAX_NG];
};

__device__ __forceinline__ void load_cortical(const vc_v1_params &p, const vc_v1_state &s,
                                              uint32_t n, NState &x) {
    const uint32_t nV1 = p.nV1;
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        x.gE[ig] = x.hE[ig] = x.gI[ig] = x.hI[ig] = 0.0f;
        if (ig < p.ngE) {
            x.gE[ig] = s.gE[ig * nV1 + n];
            x.hE[ig] = s.hE[ig * nV1 + n];
        }
        if (ig < p.ngI) {
            x.gI[ig] = s.gI[ig * nV1 + n];
            x.hI[ig] = s.hI[ig * nV1 + n];
        }
    }
}

__device__ __forceinline__ void load_rest(const vc_v1_params &p, const vc_v1_state &s,
                                          uint32_t n, NState &x) {
    const uint32_t nV1 = p.nV1;
    x.v = s.v[n];
    x.w = s.w[n];
    x.tBack = s.tBack[n];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        x.gFF[ig] = ig < p.ngFF ? s.gFF[ig * nV1 + n] : 0.0f;
        x.hFF[ig] = ig < p.ngFF ? s.hFF[ig * nV1 + n] : 0.0f;
    }
}

__device__ __forceinline__ void store_state(const vc_v1_params &p, const vc_v1_state &s,
                                            uint32_t n, const NState &x) {
    const uint32_t nV1 = p.nV1;
    s.v[n] = x.v;
    if (p.iModel == 1) s.w[n] = x.w;
    s.tBack[n] = x.tBack;
    #pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) if (ig < p.ngFF) {
        s.gFF[ig * nV1 + n] = x.gFF[ig];
        s.hFF[ig * nV1 + n] = x.hFF[ig];
    }
    #pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) if (ig < p.ngE) {
        s.gE[ig * nV1 + n] = x.gE[ig];
        s.hE[ig * nV1 + n] = x.hE[ig];
    }
    #pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) if (ig < p.ngI) {
        s.gI[ig * nV1 + n] = x.gI[ig];
        s.hI[ig * nV1 + n] = x.hI[ig];
    }
}

// tonic drive of the AdEx constructor, src/coredynamics.cuh:40-49
__device__ __forceinline__ float dep_current(const vc_v1_params &p, int t, float dep) {
    const float gL = p.gL[t], vT = p.vT[t], deltaT = p.deltaT[t], a = p.a[t];
    const float targetV = VC_VL + (vT - VC_VL) * dep;
    if (targetV > vT) return (gL + a) * (targetV - VC_VL) - gL * deltaT;
    if (deltaT > 0.0f)
        return (gL + a) * (targetV - VC_VL) - gL * deltaT * vc_expf((targetV - vT) / deltaT);
    return (gL + a) * (targetV - VC_VL) - 0.0f;
}

// LGN input of one neuron, loaded in two rounds so that other loads overlap.
// Event path: the neuron's word of synapse bits whose cell spiked; ELL path:
// the first kFF synapses (strength, cell) and their cells' spike words.
template <bool kEvents> struct FFPre;
template <> struct FFPre<true> {
    uint32_t bits;
    uint32_t evj[kEv], evid[kEv];
    float evf[kEv];
};
template <> struct FFPre<false> {
    uint32_t m;
    float fv[kFF], sv[kFF];
    uint32_t fid[kFF];
};

__device__ __forceinline__ void ff_trip1(FFPre<true> &f, const vc_ff_conn &, const uint32_t *ffMask,
                                          uint32_t n, uint32_t) {
    f.bits = ffMask[n];
}
__device__ __forceinline__ void ff_trip1(FFPre<false> &f, const vc_ff_conn &ff, const uint32_t *,
                                          uint32_t n, uint32_t nV1) {
    f.m = ff.nLGN[n];
#pragma unroll
    for (int j = 0; j < kFF; ++j) {
        f.fv[j] = 0.0f;
        f.fid[j] = 0u;
        if ((uint32_t)j < f.m) {
            const uint32_t lid = (uint32_t)j * nV1 + n;
            f.fv[j] = ff.sLGN[lid];
            f.fid[j] = ff.idLGN[lid];
        }
    }
}
// events: the first kEv synapses that received a spike this step
__device__ __forceinline__ void ff_trip2(FFPre<true> &f, const vc_ff_conn &ff, const float *,
                                          uint32_t n, uint32_t nV1) {
    uint32_t bits = f.bits;
#pragma unroll
    for (int k = 0; k < kEv; ++k) {
        f.evj[k] = 0xffffffffu;
        f.evf[k] = 0.0f;
        f.evid[k] = 0u;
        if (bits) {
            const uint32_t j = (uint32_t)__builtin_ctz(bits);
            bits &= bits - 1u;
            const uint32_t lid = j * nV1 + n;
            f.evj[k] = j;
            f.evf[k] = ff.sLGN[lid];
            f.evid[k] = ff.idLGN[lid];
        }
    }
}
__device__ __forceinline__ void ff_trip2(FFPre<false> &f, const vc_ff_conn &, const float *lgnSInfo,
                                          uint32_t, uint32_t) {
#pragma unroll
    for (int j = 0; j < kFF; ++j) f.sv[j] = ((uint32_t)j < f.m) ? lgnSInfo[f.fid[j]] : 0.0f;
}

// Background input of one neuron for a step (extension: 2 kHz Poisson drive
// applied after the LGN list): nsp events at tsp, nsp = 0 for none.
struct BGIn {
    float f, nsp, tsp;
};
// One Philox draw per neuron-step gives the Poisson count (by inversion) and
// the event time.
__device__ __forceinline__ BGIn bg_input(const vc_v1_params &p, int t, const V1Key &key,
                                         uint32_t step) {
    BGIn bg{0.0f, 0.0f, 0.0f};
    if (p.bgStrength[t] > 0.0f) {
        const vc_u32x4 r = vc_draw4(key.lo, key.hi, key.ent, step, VC_STREAM_V1_BG, 0u, p.trial);
        const float u = vc_u01_open0(r.v[0]);
        const float frac = vc_u01_open1(r.v[1]);
        const uint32_t nsp = vc_poisson_inv(u, p.bgLambda[t], p.bgP0[t]);
        if (nsp > 0u) bg = BGIn{p.bgStrength[t], (float)nsp, frac * p.dt};
    }
    return bg;
}

// LGN synapses in list order, then the background input; only synapses
// whose cell spiked (sInfo >= 1) do work (reference: nsp_FF > 0).  Every
// input of a lane goes through the one inlined copy of ff_apply in the loop
// (no call, so no stack frame).
// LGN synaptic failure of one neuron (pf = 0: off) and its draws' key.
struct FailFF {
    float pf, nSyn;
    V1Key key;
    uint32_t step, trial;
};

__device__ __forceinline__ void ff_inputs(FFAcc &acc, const FFShape &shFF, const FFShape &shBG,
                                          FFPre<true> &f, const vc_ff_conn &ff,
                                          const float *lgnSInfo, uint32_t *ffMask, uint32_t n,
                                          uint32_t nV1, float dt, float tBack, bool backing,
                                          const BGIn &bg, const FailFF &fl) {
    float evs[kEv];
    if (__ballot(f.bits != 0u) != 0ull) {
#pragma unroll
        for (int k = 0; k < kEv; ++k)
            evs[k] = (f.evj[k] != 0xffffffffu) ? lgnSInfo[f.evid[k]] : 0.0f;
    }
    uint32_t rest = f.bits;
    uint32_t nPre = 0;                       // prefetched events consumed
    bool bgLeft = bg.nsp > 0.0f;
    while (rest != 0u || bgLeft) {
        float fv, nsp, tsp, pm[VC_MAX_NG];
        if (rest != 0u) {
            const uint32_t j = (uint32_t)__builtin_ctz(rest);
            rest &= rest - 1u;
            float sI;
            if (nPre < (uint32_t)kEv) {      // the next prefetched event (ascending j)
                fv = f.evf[0];
                sI = evs[0];
#pragma unroll
                for (int k = 0; k + 1 < kEv; ++k) {
                    f.evf[k] = f.evf[k + 1];
                    evs[k] = evs[k + 1];
                }
                ++nPre;
            } else {
                const uint32_t lid = j * nV1 + n;
                fv = ff.sLGN[lid];
                sI = lgnSInfo[ff.idLGN[lid]];
            }
            nsp = (float)(uint32_t)floorf(sI);
            tsp = (sI - nsp) * dt;
            if (fl.pf > 0.0f && nsp > 0.0f)
                fv = fail_factor(fv, fl.pf, fl.nSyn, fl.key, fl.step, VC_STREAM_V1_FAILFF, j,
                                 fl.trial);
#pragma unroll
            for (int ig = 0; ig < VC_MAX_NG; ++ig) pm[ig] = shFF.p[ig];
        } else {
            fv = bg.f;
            nsp = bg.nsp;
            tsp = bg.tsp;
            bgLeft = false;
#pragma unroll
            for (int ig = 0; ig < VC_MAX_NG; ++ig) pm[ig] = shBG.p[ig];
        }
        ff_apply(acc, shFF, pm, fv, nsp, tsp, dt, tBack, backing);
    }
    if (f.bits) ffMask[n] = 0u;
}
__device__ __forceinline__ void ff_inputs(FFAcc &acc, const FFShape &shFF, const FFShape &shBG,
                                          FFPre<false> &f, const vc_ff_conn &ff,
                                          const float *lgnSInfo, uint32_t *, uint32_t n,
                                          uint32_t nV1, float dt, float tBack, bool backing,
                                          const BGIn &bg, const FailFF &fl) {
    uint32_t spk = 0;
#pragma unroll
    for (int j = 0; j < kFF; ++j) spk |= (f.sv[j] >= 1.0f ? 1u : 0u) << j;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) spk |= (uint32_t)__shfl_xor((int)spk, off);
    spk = __builtin_amdgcn_readfirstlane(spk);           // slots that spiked for any lane
    // one inlined input body: the slots that spiked for some lane (in list
    // order, a wave-uniform index into the registers), then the synapses
    // past kFF, then the background input
    uint32_t j = 0;
    uint32_t i = kFF;
    bool bgLeft = bg.nsp > 0.0f;
    for (;;) {
        float fv = 0.0f, sI = 0.0f, pm[VC_MAX_NG];
        bool isBG = false;
        uint32_t slotj = 0;                              // the synapse's slot in the list
        if (spk != 0u) {
            j = (uint32_t)__builtin_ctz(spk);
            slotj = j;
            spk &= spk - 1u;
#pragma unroll
            for (int k = 0; k < kFF; ++k) {      // j is wave-uniform: scalar compares
                if ((uint32_t)k == j) {
                    fv = f.fv[k];
                    sI = f.sv[k];
                }
            }
        } else if (i < f.m) {                            // lists longer than kFF
            const uint32_t lid = i * nV1 + n;
            slotj = i;
            ++i;
            sI = lgnSInfo[ff.idLGN[lid]];
            fv = ff.sLGN[lid];
        } else if (bgLeft) {
            bgLeft = false;
            isBG = true;
        } else {
            break;                                       // this lane is done
        }
        float nsp, tsp;
        if (isBG) {
            fv = bg.f;
            nsp = bg.nsp;
            tsp = bg.tsp;
        } else {
            if (!(sI >= 1.0f)) fv = 0.0f;                // this lane's cell did not spike
            nsp = (float)(uint32_t)floorf(sI);
            tsp = (sI - nsp) * dt;
            if (fl.pf > 0.0f && fv > 0.0f)
                fv = fail_factor(fv, fl.pf, fl.nSyn, fl.key, fl.step, VC_STREAM_V1_FAILFF, slotj,
                                 fl.trial);
        }
#pragma unroll
        for (int ig = 0; ig < VC_MAX_NG; ++ig) pm[ig] = isBG ? shBG.p[ig] : shFF.p[ig];
        ff_apply(acc, shFF, pm, fv, nsp, tsp, dt, tBack, backing);
    }
}

// Stage c-FF of one neuron: the LGN and background inputs into the FF
// receptors at the end of the step (g1, h1) and, when backing up from
// refractoriness (0 < tBack < dt), at tBack (g0, h0).  Depends on the
// neuron's own state only, not on the other neurons' spikes.
template <bool kEvents, bool kSt>
__device__ __forceinline__ FFAcc ff_stage(const vc_v1_params &p, const V1Const &k,
                                          const vc_ff_conn &ff, int t, uint32_t n, uint32_t b,
                                          uint32_t step, const NState &x, FFPre<kEvents> &f,
                                          const float *lgnSInfo, uint32_t *ffMask) {
    const float dt = p.dt;
    const float tBack = x.tBack;
    FFAcc acc;
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        acc.g0[ig] = acc.h0[ig] = acc.g1[ig] = acc.h1[ig] = 0.0f;
        if (ig < p.ngFF) {
            acc.g1[ig] = x.gFF[ig];
            acc.h1[ig] = x.hFF[ig];
            acc.g0[ig] = acc.g1[ig];
            acc.h0[ig] = acc.h1[ig];
            decay_const(acc.g1[ig], acc.h1[ig], k.etrFF[ig], k.etdFF[ig], p.dodFF[ig]);
        }
    }
    const bool backing = (tBack < dt) && (tBack > 0.0f);
    if (backing) {
#pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig)
            if (ig < p.ngFF)
                decay_var(acc.g0[ig], acc.h0[ig], tBack, p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
    }
    FFShape shFF, shBG;
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        const uint32_t pi = t * p.ngFF + (ig < p.ngFF ? ig : 0u);
        shFF.p[ig] = p.pFF[pi];
        shBG.p[ig] = p.pBG[pi];
        shFF.rise[ig] = shBG.rise[ig] = p.riseFF[ig];
        shFF.decay[ig] = shBG.decay[ig] = p.decayFF[ig];
        shFF.dod[ig] = shBG.dod[ig] = p.dodFF[ig];
    }
    shFF.ng = shBG.ng = p.ngFF;
    const V1Key key = v1_key(p, b, n);
    const BGIn bg = bg_input(p, t, key, step);
    const FailFF fl{kSt ? p.synFailFF[t] : 0.0f, p.synPerConFF[t], key, step, p.trial};
    ff_inputs(acc, shFF, shBG, f, ff, lgnSInfo, ffMask, n, p.nV1, dt, tBack, backing, bg, fl);
    return acc;
}

// Voltage noise of one neuron for a step (src/coredynamics.cu:1424-1441):
// sigma = sqrt(2 noisyDep (vT - vR) r dt) N(0,1), r = tonicDep when
// noiseOnTonic; Ornstein-Uhlenbeck when tauNoise > 0 (`last` is the previous
// step's value and becomes this step's).  Returns the terms the RK2 slopes
// take, n0 and n1 (white: sigma/dt both; OU: last, new), and in *dep the
// value the reference adds to depC in its dep output (sigma, or the new OU
// value).  All 0 without noise.
__device__ __forceinline__ void noise_terms(const vc_v1_params &p, int t, const V1Key &key,
                                            uint32_t step, float tonic, float &last, float &n0,
                                            float &n1, float &dep) {
    const float dt = p.dt, tau = p.tauNoise;
    const float ratio = p.noiseOnTonic ? tonic : 1.0f;
    float noise1 = __builtin_sqrtf(2.0f * p.noisyDep[t] * (p.vT[t] - p.vR[t]) * ratio * dt);
    const float noise0 = tau > 0.0f ? last : 0.0f;
    if (noise1 > 0.0f) {
        const vc_u32x4 r = vc_draw4(key.lo, key.hi, key.ent, step, VC_STREAM_V1_NOISE, 0u,
                                    p.trial);
        noise1 *= vc_normal(r.v[0], r.v[1]);
        if (tau > 0.0f) {
            noise1 /= tau;
            const float hk1 = -noise0 / tau;
            const float hk2 = -(noise0 + dt * hk1 + noise1) / tau;
            noise1 += noise0 + dt * (hk1 + hk2) / 2.0f;
            last = noise1;
        }
    }
    dep = noise1;
    if (tau > 0.0f) {
        n0 = noise0;
        n1 = noise1;
    } else {
        n0 = n1 = noise1 / dt;
    }
}

// Stage d of one neuron: x holds the state at the start of the step with the
// cortical g/h already incremented by the previous step's spikes, acc the FF
// receptors of ff_stage; on return x holds the state after the step.  gapS
// and gap are the neuron's gap-junction terms (0 for excitatory neurons and
// without gap junctions).  Returns sInfo.
__device__ __forceinline__ float membrane_stage(const vc_v1_params &p, const V1Const &k, int t,
                                                float depC, NState &x, const FFAcc &acc,
                                                float gapS, float gap, float n0, float n1) {
    const float dt = p.dt;
    const float vR = p.vR[t], vThres = p.vThres[t], gL = p.gL[t], C = p.C[t];
    const float tRef = p.tRef[t], vT = p.vT[t], deltaT = p.deltaT[t];
    const float tau_w = p.tau_w[t], a = p.a[t], bb = p.b[t];
    const float v0 = x.v;
    float w0 = x.w;
    float tBack = x.tBack;
    const bool backing = (tBack < dt) && (tBack > 0.0f);
    // cortical conductances decayed to the end of the step (the next state)
    float geN[VC_MAX_NG], heN[VC_MAX_NG], giN[VC_MAX_NG], hiN[VC_MAX_NG];
    float gE_t1 = 0.0f, gI_t1 = 0.0f;
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        float g = x.gE[ig], h = x.hE[ig];
        if (ig < p.ngE) {
            decay_const(g, h, k.etrE[ig], k.etdE[ig], p.dodE[ig]);
            gE_t1 += g;
        }
        geN[ig] = g;
        heN[ig] = h;
    }
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        float g = x.gI[ig], h = x.hI[ig];
        if (ig < p.ngI) {
            decay_const(g, h, k.etrI[ig], k.etdI[ig], p.dodI[ig]);
            gI_t1 += g;
        }
        giN[ig] = g;
        hiN[ig] = h;
    }
    float gE_t0 = 0.0f;
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        if (ig < p.ngFF) {
            gE_t0 += acc.g0[ig];
            gE_t1 += acc.g1[ig];
            x.gFF[ig] = acc.g1[ig];
            x.hFF[ig] = acc.h1[ig];
        }
    }
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        if (ig < p.ngE) {
            if (backing) decay_var(x.gE[ig], x.hE[ig], tBack, p.riseE[ig], p.decayE[ig], p.dodE[ig]);
            gE_t0 += x.gE[ig];
        }
    }
    float gI_t0 = 0.0f;
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        if (ig < p.ngI) {
            if (backing) decay_var(x.gI[ig], x.hI[ig], tBack, p.riseI[ig], p.decayI[ig], p.dodI[ig]);
            gI_t0 += x.gI[ig];
        }
    }
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        x.gE[ig] = geN[ig];
        x.hE[ig] = heN[ig];
        x.gI[ig] = giN[ig];
        x.hI[ig] = hiN[ig];
    }

    // ---- Heun RK2 (src/coredynamics.cu:1534-1589)
    float v = v0, w = w0, tsp = 0.0f;
    bool spiked = false;
    if (tBack < dt) {
        const float a0 = gE_t0 + gI_t0 + gL + gapS;
        const float b0 = gE_t0 * VC_VE + gI_t0 * VC_VI + gL * VC_VL + depC + gap;
        const float a1 = gE_t1 + gI_t1 + gL + gapS;
        const float b1 = gE_t1 * VC_VE + gI_t1 * VC_VI + gL * VC_VL + depC + gap;
        const float new_dt = dt - tBack;
        if (backing) {
            const float A = a * (v0 - VC_VL);
            w0 = (w0 - A) * vc_expf(-tBack / tau_w) + A;
        }
        const float dTgL = deltaT * gL;
        float fk1, fk2;
        if (deltaT > 0.0f)
            fk1 = (-a0 * v0 + b0 + dTgL * vc_expf((v0 - vT) / deltaT) - w0) / C + n0;
        else
            fk1 = (-a0 * v0 + b0 - w0) / C + n0;
        const float gk1 = (a * (v0 - VC_VL) - w0) / tau_w;
        const float v1 = v0 + new_dt * fk1;
        const float w1 = w0 + new_dt * gk1;
        if (deltaT > 0.0f)
            fk2 = (-a1 * v1 + b1 + dTgL * vc_expf((v1 - vT) / deltaT) - w1) / C + n1;
        else
            fk2 = (-a1 * v1 + b1 - w1) / C + n1;
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
    x.v = v;
    if (p.iModel == 1) x.w = w;
    x.tBack = tBack;
    return spiked ? 1.0f + tsp / dt : v;
}

// Stage d when the neuron's refractory period is shorter than dt, so that it
// may fire several times in one step (compute_V_collect_spike_learnFF,
// src/coredynamics.cu:617-836; the reference runs it for every neuron when
// dt > min tRef, and for a type with tRef >= dt it gives what membrane_stage
// gives).  After each spike whose refractory period ends inside the step
// (tBack < dt) the rest of the step is integrated again from vR: the
// conductances at the new tBack are those at the previous restart decayed
// over dtBack = tBack - new_t0 plus the LGN and background inputs that
// arrived in [new_t0, tBack) (re-read from the synapse lists; the failure
// draws are keyed by the synapse, so a re-read finds the same factor);
// w is carried as the reference carries it (rk2_vFixedBefore over dtBack
// after compute_spike_time has moved w0 to the spike), and white noise is
// divided by dt once more in every further pass, as the reference's
// `noise1 /= dt` inside its loop does.  sInfo = nsp + mean(tsp)/dt.
// Arguments as membrane_stage, plus what the re-read needs.
__device__ __forceinline__ float membrane_multi(const vc_v1_params &p, const V1Const &k, int t,
                                                float depC, NState &x, const FFAcc &acc,
                                                float gapS, float gap, float n0, float n1,
                                                const vc_ff_conn &ff, const float *lgnSInfo,
                                                uint32_t n, uint32_t step, const V1Key &key) {
    const float dt = p.dt;
    const uint32_t nV1 = p.nV1;
    const float vR = p.vR[t], vThres = p.vThres[t], gL = p.gL[t], C = p.C[t];
    const float tRef = p.tRef[t], vT = p.vT[t], deltaT = p.deltaT[t];
    const float tau_w = p.tau_w[t], a = p.a[t], bb = p.b[t];
    const bool white = !(p.tauNoise > 0.0f);
    float tBack = x.tBack;
    bool backing = (tBack < dt) && (tBack > 0.0f);
    // cortical conductances decayed to the end of the step (the next state)
    float geN[VC_MAX_NG], heN[VC_MAX_NG], giN[VC_MAX_NG], hiN[VC_MAX_NG];
    float gE_t1 = 0.0f, gI_t1 = 0.0f;
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        float g = x.gE[ig], h = x.hE[ig];
        if (ig < p.ngE) {
            decay_const(g, h, k.etrE[ig], k.etdE[ig], p.dodE[ig]);
            gE_t1 += g;
        }
        geN[ig] = g;
        heN[ig] = h;
    }
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        float g = x.gI[ig], h = x.hI[ig];
        if (ig < p.ngI) {
            decay_const(g, h, k.etrI[ig], k.etdI[ig], p.dodI[ig]);
            gI_t1 += g;
        }
        giN[ig] = g;
        hiN[ig] = h;
    }
    float g0f[VC_MAX_NG], h0f[VC_MAX_NG];    // FF receptors at the current restart
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        g0f[ig] = acc.g0[ig];
        h0f[ig] = acc.h0[ig];
        if (ig < p.ngFF) {
            gE_t1 += acc.g1[ig];
            x.gFF[ig] = acc.g1[ig];
            x.hFF[ig] = acc.h1[ig];
        }
    }
    float pFF[VC_MAX_NG], pBG[VC_MAX_NG];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        const uint32_t pi = t * p.ngFF + (ig < p.ngFF ? ig : 0u);
        pFF[ig] = p.pFF[pi];
        pBG[ig] = p.pBG[pi];
    }
    const BGIn bg = bg_input(p, t, key, step);
    const float pf = p.synFailFF[t];
    const uint32_t m = ff.nLGN[n];

    float v0 = x.v, w0 = x.w, v = v0, w = w0, tsp = 0.0f;
    float new_t0 = 0.0f, sumTsp = 0.0f, a1 = 0.0f, b1 = 0.0f;
    uint32_t count = 0;
    bool first = true;
    do {
        const float dtBack = tBack - new_t0;
        if (!first) {                        // a restart after a spike (backing is set)
#pragma unroll
            for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig)
                if (ig < p.ngFF)
                    decay_var(g0f[ig], h0f[ig], dtBack, p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
            for (uint32_t i = 0; i < m; ++i) {          // LGN inputs in [new_t0, tBack)
                const uint32_t lid = i * nV1 + n;
                const float sI = lgnSInfo[ff.idLGN[lid]];
                const float nsp = (float)(uint32_t)floorf(sI);
                if (!(nsp > 0.0f)) continue;
                const float ts = (sI - nsp) * dt;
                if (!(ts >= new_t0)) continue;
                float f = ff.sLGN[lid];
                if (pf > 0.0f)
                    f = fail_factor(f, pf, p.synPerConFF[t], key, step, VC_STREAM_V1_FAILFF, i,
                                    p.trial);
                const float ddt = tBack - ts;
                if (f > 0.0f && ddt > 0.0f) {
#pragma unroll
                    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig)
                        if (ig < p.ngFF)
                            single_input(g0f[ig], h0f[ig], f * pFF[ig] * nsp, ddt, p.riseFF[ig],
                                         p.decayFF[ig], p.dodFF[ig]);
                }
            }
            if (bg.nsp > 0.0f && bg.tsp >= new_t0) {    // the background input
                const float ddt = tBack - bg.tsp;
                if (ddt > 0.0f) {
#pragma unroll
                    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig)
                        if (ig < p.ngFF)
                            single_input(g0f[ig], h0f[ig], bg.f * pBG[ig] * bg.nsp, ddt,
                                         p.riseFF[ig], p.decayFF[ig], p.dodFF[ig]);
                }
            }
        }
        float gE_t0 = 0.0f;
#pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig)
            if (ig < p.ngFF) gE_t0 += g0f[ig];
#pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
            if (ig < p.ngE) {
                if (backing)
                    decay_var(x.gE[ig], x.hE[ig], dtBack, p.riseE[ig], p.decayE[ig], p.dodE[ig]);
                gE_t0 += x.gE[ig];
            }
        }
        float gI_t0 = 0.0f;
#pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
            if (ig < p.ngI) {
                if (backing)
                    decay_var(x.gI[ig], x.hI[ig], dtBack, p.riseI[ig], p.decayI[ig], p.dodI[ig]);
                gI_t0 += x.gI[ig];
            }
        }
        tsp = 0.0f;
        if (tBack < dt) {
            const float a0 = gE_t0 + gI_t0 + gL + gapS;
            const float b0 = gE_t0 * VC_VE + gI_t0 * VC_VI + gL * VC_VL + depC + gap;
            if (first) {
                a1 = gE_t1 + gI_t1 + gL + gapS;
                b1 = gE_t1 * VC_VE + gI_t1 * VC_VI + gL * VC_VL + depC + gap;
            }
            const float new_dt = dt - tBack;
            if (backing) {
                const float A = a * (v0 - VC_VL);
                w0 = (w0 - A) * vc_expf(-dtBack / tau_w) + A;
            }
            if (!first && white) n0 = n1 = n1 / dt;
            const float dTgL = deltaT * gL;
            float fk1, fk2;
            if (deltaT > 0.0f)
                fk1 = (-a0 * v0 + b0 + dTgL * vc_expf((v0 - vT) / deltaT) - w0) / C + n0;
            else
                fk1 = (-a0 * v0 + b0 - w0) / C + n0;
            const float gk1 = (a * (v0 - VC_VL) - w0) / tau_w;
            const float v1 = v0 + new_dt * fk1;
            const float w1 = w0 + new_dt * gk1;
            if (deltaT > 0.0f)
                fk2 = (-a1 * v1 + b1 + dTgL * vc_expf((v1 - vT) / deltaT) - w1) / C + n1;
            else
                fk2 = (-a1 * v1 + b1 - w1) / C + n1;
            const float gk2 = (a * (v1 - VC_VL) - w1) / tau_w;
            w = w0 + new_dt * (gk1 + gk2) / 2.0f;
            v = v0 + new_dt * (fk1 + fk2) / 2.0f;
            if (v > vThres) {
                new_t0 = tBack;
                const float ratio = (vThres - v0) / (v - v0);
                tsp = new_t0 + ratio * new_dt;
                w0 += ratio * (w - w0) + bb;
                sumTsp += tsp;
                ++count;
                tBack = tsp + tRef;
                backing = tBack < dt;
                if (backing)
                    v0 = vR;
                else
                    v = vR;
            } else {
                if (tBack > 0.0f) tBack = 0.0f;
                backing = false;
            }
        }
        if (tBack >= dt) {
            v = vR;
            const float A = a * (v - VC_VL);
            w = (w0 - A) * vc_expf(-(dt - tsp) / tau_w) + A;
            tBack -= dt;
        }
        first = false;
        // tRef >= dt/1024 (checked on the host) ends the loop within 1025 passes
    } while (backing && count <= 1024u);
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        x.gE[ig] = geN[ig];
        x.hE[ig] = heN[ig];
        x.gI[ig] = giN[ig];
        x.hI[ig] = hiN[ig];
    }
    x.v = v;
    if (p.iModel == 1) x.w = w;
    x.tBack = tBack;
    if (count > 0u) return sumTsp / ((float)count * dt) + (float)count;
    return v;
}

// Stage d: the multi-spike form for types whose tRef < dt (compiled into
// the kSt instances only), else the one-spike form.
template <bool kSt>
__device__ __forceinline__ float membrane(const vc_v1_params &p, const V1Const &k, int t,
                                          float depC, NState &x, const FFAcc &acc, float gapS,
                                          float gap, float n0, float n1, const vc_ff_conn &ff,
                                          const float *lgnSInfo, uint32_t n, uint32_t step,
                                          const V1Key &key) {
    if constexpr (kSt) {
        if (p.tRef[t] < p.dt)
            return membrane_multi(p, k, t, depC, x, acc, gapS, gap, n0, n1, ff, lgnSInfo, n,
                                  step, key);
    }
    return membrane_stage(p, k, t, depC, x, acc, gapS, gap, n0, n1);
}

// Spike records being written (persistent kernel).
struct RecOut {
    uint64_t *words;              // scratch base
    uint32_t rec, ovf;            // word offsets of the record and overflow buffers written
    uint32_t tag;
};

// Stage e: ballot compaction -- one spike mask per tile and the unit
// increments at slot tile*64 + rank (src/coredynamics.cu:1994-2046).  With
// kRec (persistent kernel) the mask and the increments also go to the tile's
// tagged record words.
template <bool kRec>
__device__ __forceinline__ void compact(const vc_v1_params &p, const vc_spikes &next,
                                        const RecOut &ro, uint32_t n, uint32_t local,
                                        uint32_t lane, float sInfo) {
    const uint32_t nV1 = p.nV1;
    const float dt = p.dt;
    const bool exist = sInfo >= 1.0f;
    const uint64_t m = __ballot(exist);
    const uint32_t tile = n >> 6;
    if (lane == 0) {
        next.mask[tile] = m;
        if (kRec) {
            uint64_t *r = ro.words + ro.rec + (size_t)tile * kRecW;
            st_spk<true>(r, tagged((uint32_t)m, ro.tag));
            st_spk<true>(r + 1, tagged((uint32_t)(m >> 32), ro.tag));
        }
    }
    if (exist) {
        const float nsp = floorf(sInfo);
        const float tsp2 = (sInfo - nsp) * dt;
        const uint32_t rank = vc_mbcnt(m);
        const uint32_t slot = tile * 64u + rank;
        const bool isE = local < p.nE;
        const uint32_t ng = isE ? p.ngE : p.ngI;
        float lg[VC_MAX_NG], lh[VC_MAX_NG];
#pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
            lg[ig] = lh[ig] = 0.0f;
            if (ig < ng) {
                if (isE)
                    single_input(lg[ig], lh[ig], nsp, dt - tsp2, p.riseE[ig], p.decayE[ig],
                                 p.dodE[ig]);
                else
                    single_input(lg[ig], lh[ig], nsp, dt - tsp2, p.riseI[ig], p.decayI[ig],
                                 p.dodI[ig]);
                next.og[ig * nV1 + slot] = lg[ig];
                next.oh[ig * nV1 + slot] = lh[ig];
            }
        }
        if (kRec) {
            uint64_t *w = ro.words + (rank < (uint32_t)kRecSpk
                                          ? ro.rec + (size_t)tile * kRecW + 2u + 4u * rank
                                          : ro.ovf + ((size_t)tile * 64u + rank) * 4u);
            st_spk<true>(w + 0, tagged(__float_as_uint(lg[0]), ro.tag));
            st_spk<true>(w + 1, tagged(__float_as_uint(lh[0]), ro.tag));
            st_spk<true>(w + 2, tagged(__float_as_uint(lg[1]), ro.tag));
            st_spk<true>(w + 3, tagged(__float_as_uint(lh[1]), ro.tag));
        }
    }
}

// ------------------------------------------------------ one-step kernel --
// kSt: noise or synaptic failure may be on (their draws are compiled out of
// the other instances).
template <bool kUpdate, bool kEvents, int kG, bool kSt>
__global__ __launch_bounds__(kWG) void k_v1_step(vc_v1_params p, V1Const k, vc_v1_state s,
                                                 vc_ff_conn ff, vc_rec_conn rec,
                                                 vc_spikes prev, vc_spikes next,
                                                 const float *__restrict__ lgnSInfo,
                                                 uint32_t *__restrict__ ffMask,
                                                 uint32_t step, float *__restrict__ out_sInfo,
                                                 float *__restrict__ out_dep,
                                                 float *__restrict__ out_g_prev,
                                                 float *__restrict__ out_snap) {
    __shared__ GatherLds glds;
    __shared__ uint64_t smk[kMaxEnt];          // neighbour-tile masks of this block (all
    __shared__ uint32_t sbid[kMaxEnt];         // wavefronts of a workgroup share the block)
    const uint32_t nV1 = p.nV1;
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t b = n / p.blockSize;
    const uint32_t local = n - b * p.blockSize;
    const uint32_t lane = vc_lane();
    const uint32_t wv = threadIdx.x >> 6;
    const int t = type_of(p, local);
    const uint32_t nSeg = k.nSeg;
    const uint32_t myseg = local >> 6;

    // ---- memory trip 1: neighbour spike masks (lane-parallel), LGN synapses, state
    const uint32_t nb = rec.nNear[b];
    const uint32_t nEnt = nb * nSeg;
    for (uint32_t e = threadIdx.x; e < nEnt; e += blockDim.x) {
        const uint32_t ib = e / nSeg, tt = e - ib * nSeg;
        const uint32_t bid = rec.nbr[b * rec.maxNear + ib];
        smk[e] = prev.mask[bid * nSeg + tt];
        sbid[e] = bid;
    }
    FFPre<kEvents> f;
    if (kUpdate) ff_trip1(f, ff, ffMask, n, nV1);
    NState x;
    load_cortical(p, s, n, x);
    // ---- memory trip 2: LGN spike words, then the recurrent gather
    if (kUpdate) ff_trip2(f, ff, lgnSInfo, n, nV1);
    float lgE[VC_MAX_NG] = {0.f, 0.f}, lhE[VC_MAX_NG] = {0.f, 0.f};
    float lgI[VC_MAX_NG] = {0.f, 0.f}, lhI[VC_MAX_NG] = {0.f, 0.f};
    __syncthreads();
    const RecIn noRec{nullptr, 0u, 0u, 0u, nullptr};
    // (the failure draws of the spikes of step - 1 are keyed by that step)
    rec_gather<false, false, kG, kSt>(p, rec, prev, b, myseg, nSeg, nEnt, t, smk, sbid, noRec, glds,
                                 v1_key(p, b, n), step - 1u, lgE, lhE, lgI, lhI);
    // conductances at the end of the previous step (reference: after recal_G)
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        x.gE[ig] = x.gE[ig] + lgE[ig];
        x.hE[ig] = x.hE[ig] + lhE[ig];
        x.gI[ig] = x.gI[ig] + lgI[ig];
        x.hI[ig] = x.hI[ig] + lhI[ig];
    }
    if (out_snap) {                 // state at the start of the step, cortical g/h incremented
        out_snap[VC_SNAP_V * nV1 + n] = s.v[n];
        out_snap[VC_SNAP_TBACK * nV1 + n] = s.tBack[n];
        #pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) if (ig < p.ngFF) {
            out_snap[(VC_SNAP_GFF + ig) * nV1 + n] = s.gFF[ig * nV1 + n];
            out_snap[(VC_SNAP_HFF + ig) * nV1 + n] = s.hFF[ig * nV1 + n];
        }
        #pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) if (ig < p.ngE) {
            out_snap[(VC_SNAP_GE + ig) * nV1 + n] = x.gE[ig];
            out_snap[(VC_SNAP_HE + ig) * nV1 + n] = x.hE[ig];
        }
        #pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) if (ig < p.ngI) {
            out_snap[(VC_SNAP_GI + ig) * nV1 + n] = x.gI[ig];
            out_snap[(VC_SNAP_HI + ig) * nV1 + n] = x.hI[ig];
        }
    }
    if (out_g_prev) {
        const uint32_t ng = p.ngE + p.ngI;
        #pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) if (ig < p.ngE) {
            out_g_prev[ig * nV1 + n] = x.gE[ig];
            out_g_prev[(ng + ig) * nV1 + n] = x.hE[ig];
        }
        #pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) if (ig < p.ngI) {
            out_g_prev[(p.ngE + ig) * nV1 + n] = x.gI[ig];
            out_g_prev[(ng + p.ngE + ig) * nV1 + n] = x.hI[ig];
        }
    }
    if (!kUpdate) {
        #pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) if (ig < p.ngE) {
            s.gE[ig * nV1 + n] = x.gE[ig];
            s.hE[ig * nV1 + n] = x.hE[ig];
        }
        #pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) if (ig < p.ngI) {
            s.gI[ig * nV1 + n] = x.gI[ig];
            s.hI[ig * nV1 + n] = x.hI[ig];
        }
        return;
    }
    load_rest(p, s, n, x);
    const float tonic = s.tonicDep[n];
    const float depC = dep_current(p, t, tonic);
    float n0 = 0.0f, n1 = 0.0f, dn = 0.0f;
    if constexpr (kSt) {
        float last = s.lastNoise ? s.lastNoise[n] : 0.0f;
        noise_terms(p, t, v1_key(p, b, n), step, tonic, last, n0, n1, dn);
        if (s.lastNoise) s.lastNoise[n] = last;
    }
    if (out_dep) out_dep[n] = depC + dn;
    const FFAcc acc = ff_stage<kEvents, kSt>(p, k, ff, t, n, b, step, x, f, lgnSInfo, ffMask);
    float gapS = 0.0f, gap = 0.0f;           // inhibitory neurons' gap-junction terms
    if (local >= p.nE) {
        const uint32_t gi = b * (p.blockSize - p.nE) + (local - p.nE);
        if (s.gapS) gapS = s.gapS[gi];
        if (s.gap) gap = s.gap[gi];
    }
    const float sInfo = membrane<kSt>(p, k, t, depC, x, acc, gapS, gap, n0, n1, ff, lgnSInfo, n,
                                      step, v1_key(p, b, n));
    store_state(p, s, n, x);
    out_sInfo[n] = sInfo;
    const RecOut noOut{nullptr, 0u, 0u, 0u};
    compact<false>(p, next, noOut, n, local, lane, sInfo);
}

// ------------------------------------------------------ gap junctions --
// The inhibitory neurons' gap sums of recal_G_mat_nd / recal_G_mat
// (src/coredynamics.cu:2524-2540, 2237-2260): one lane per inhibitory
// neuron, its junctions in the ELL's order (neighbour slot, then partner
// index: the reference's order), a partner's step output replaced by its
// type's vThres when it spiked.  lag[j] < depth (checked by the caller).
__global__ __launch_bounds__(256) void k_v1_gap(vc_v1_params p, vc_rec_conn rec,
                                                const float *__restrict__ train, uint32_t depth,
                                                uint32_t cur, const uint32_t *__restrict__ lag,
                                                uint32_t mI, float *__restrict__ gap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mI) return;
    const uint32_t cnt = rec.nGap[i];
    float sum = 0.0f;
    for (uint32_t j = 0; j < cnt; ++j) {
        const uint64_t e = (uint64_t)j * mI + i;
        const uint32_t pre = rec.gapPre[e];
        uint32_t slot = cur;
        if (lag) slot = (cur + depth - lag[e] % depth) % depth;
        float v = train[(uint64_t)slot * p.nV1 + pre];
        if (v > 0.0f) v = p.vThres[type_of(p, pre % p.blockSize)];
        sum = sum + rec.gapG[e] * v;
    }
    gap[i] = sum;
}

// ----------------------------------------------- persistent multi-step --
// Runs nsteps steps in one launch; every workgroup (a quarter, half or whole
// V1 block) keeps its neurons' state in registers.  Spikes travel between
// workgroups as tagged records (see kRecSpk) in `words`: the records of step
// s go to buffer (s+1)&1 with tag s+2, a prologue writing those of the step
// before it0 (from sa) as s = -1.  A step reads its neighbours' records of
// the previous step, waiting on each word's tag; it never waits on anything
// else.  A buffer is rewritten (step s+2) only after every neighbour has
// published step s+1, which it does after reading the buffer in step s+1.
// Tags are epoch*65536 + s+2 with a per-launch epoch, so words left by
// earlier launches never match.  The vc_spikes buffers sa/sb receive every
// step's spikes as vc_v1_run leaves them.  Every wait is bounded: on a
// time-out *status is set and the kernel runs to its end (results invalid,
// no hang).  The grid must fit on the device at once.  kSt as in k_v1_step.
template <bool kEvents, int kG, bool kSt>
__global__ __launch_bounds__(kWG) void k_v1_multi(vc_v1_params p, V1Const k, vc_v1_state s,
                                                  vc_ff_conn ff, vc_rec_conn rec,
                                                  vc_spikes sa, vc_spikes sb,
                                                  const float *__restrict__ lgnSInfo,
                                                  uint64_t lgnStride, uint32_t *__restrict__ ffMask,
                                                  uint64_t maskStride, uint32_t it0, uint32_t nsteps,
                                                  float *__restrict__ out_sInfo, uint64_t outStride,
                                                  uint64_t *words, uint32_t epoch,
                                                  uint32_t *status) {
    // the step chain is the critical path; the LGN stage's waves that share
    // the SIMDs run a frame period ahead, so they yield issue slots to these
    __builtin_amdgcn_s_setprio(3);
    __shared__ GatherLds glds;
    __shared__ uint64_t smk[kMaxEnt];
    __shared__ uint32_t sbid[kMaxEnt];
    const uint32_t nV1 = p.nV1;
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t b = n / p.blockSize;
    const uint32_t local = n - b * p.blockSize;
    const uint32_t lane = vc_lane();
    const uint32_t wv = threadIdx.x >> 6;
    const int t = type_of(p, local);
    const uint32_t nSeg = k.nSeg;
    const uint32_t myseg = local >> 6;
    const uint32_t nb = rec.nNear[b];
    const uint32_t nEnt = nb * nSeg;
    const uint32_t nTiles = nV1 / 64u;
    const uint32_t recW = nTiles * kRecW;                     // words of one record buffer
    const uint32_t ovfW = nTiles * 64u * 4u;                  // words of one overflow buffer
    auto rec_of = [&](uint32_t buf) { return buf * recW; };
    auto ovf_of = [&](uint32_t buf) { return 2u * recW + buf * ovfW; };
    for (uint32_t e = threadIdx.x; e < nEnt; e += blockDim.x)
        sbid[e] = rec.nbr[b * rec.maxNear + e / nSeg];
    NState x;
    load_cortical(p, s, n, x);
    load_rest(p, s, n, x);
    const float tonic = s.tonicDep[n];
    const float depC = dep_current(p, t, tonic);
    const V1Key key = v1_key(p, b, n);
    float last = kSt && s.lastNoise ? s.lastNoise[n] : 0.0f;      // OU noise state
    // prologue: the spikes of the step before it0 (sa) as records of step -1
    {
        const uint32_t tile = n >> 6;
        const uint64_t m = sa.mask[tile];
        const RecOut ro{words, rec_of(0), ovf_of(0), (epoch << 16) + 1u};
        uint64_t *r = words + ro.rec + (size_t)tile * kRecW;
        if (lane == 0) {
            st_spk<true>(r, tagged((uint32_t)m, ro.tag));
            st_spk<true>(r + 1, tagged((uint32_t)(m >> 32), ro.tag));
        }
        if ((m >> lane) & 1ull) {
            const uint32_t rank = vc_mbcnt(m);
            const uint32_t slot = tile * 64u + rank;
            uint64_t *w = words + (rank < (uint32_t)kRecSpk
                                       ? ro.rec + (size_t)tile * kRecW + 2u + 4u * rank
                                       : ro.ovf + ((size_t)tile * 64u + rank) * 4u);
            st_spk<true>(w + 0, tagged(__float_as_uint(sa.og[slot]), ro.tag));
            st_spk<true>(w + 1, tagged(__float_as_uint(sa.oh[slot]), ro.tag));
            st_spk<true>(w + 2, tagged(__float_as_uint(sa.og[nV1 + slot]), ro.tag));
            st_spk<true>(w + 3, tagged(__float_as_uint(sa.oh[nV1 + slot]), ro.tag));
        }
    }
    for (uint32_t q = 0; q < nsteps; ++q) {
        const vc_spikes &prev = (q & 1u) ? sb : sa;
        const vc_spikes &next = (q & 1u) ? sa : sb;
        const float *lgn = lgnSInfo + q * lgnStride;
        uint32_t *fm = kEvents ? ffMask + q * maskStride : nullptr;
        // the LGN / background part does not depend on the neighbours: it
        // runs before their spikes are awaited
        FFPre<kEvents> f;
        ff_trip1(f, ff, fm, n, nV1);
        ff_trip2(f, ff, lgn, n, nV1);
        const FFAcc acc = ff_stage<kEvents, kSt>(p, k, ff, t, n, b, it0 + q, x, f, lgn, fm);
        // the neighbours' spike masks of the previous step (tag q+1)
        const RecIn ri{words, rec_of(q & 1u), ovf_of(q & 1u), (epoch << 16) + q + 1u, status};
        __syncthreads();                      // every wavefront is done with smk
        // one entry (both mask words) per thread, issued together
        for (uint32_t e = threadIdx.x; e < nEnt; e += blockDim.x) {
            const uint32_t tile = sbid[e] * nSeg + (e - (e / nSeg) * nSeg);
            uint32_t u[2];
            ld_tagged<2>(words + ri.rec + (size_t)tile * kRecW, ri.tag, status, u);
            smk[e] = ((uint64_t)u[1] << 32) | u[0];
        }
        float lgE[VC_MAX_NG] = {0.f, 0.f}, lhE[VC_MAX_NG] = {0.f, 0.f};
        float lgI[VC_MAX_NG] = {0.f, 0.f}, lhI[VC_MAX_NG] = {0.f, 0.f};
        __syncthreads();
        rec_gather<true, true, kG, kSt>(p, rec, prev, b, myseg, nSeg, nEnt, t, smk, sbid, ri, glds,
                                   key, it0 + q - 1u, lgE, lhE, lgI, lhI);
#pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
            x.gE[ig] = x.gE[ig] + lgE[ig];
            x.hE[ig] = x.hE[ig] + lhE[ig];
            x.gI[ig] = x.gI[ig] + lgI[ig];
            x.hI[ig] = x.hI[ig] + lhI[ig];
        }
        float n0 = 0.0f, n1 = 0.0f, dn;
        if constexpr (kSt) noise_terms(p, t, key, it0 + q, tonic, last, n0, n1, dn);
        const float sInfo = membrane<kSt>(p, k, t, depC, x, acc, 0.0f, 0.0f, n0, n1, ff, lgn, n,
                                          it0 + q, key);
        out_sInfo[q * outStride + n] = sInfo;
        const RecOut ro{words, rec_of((q + 1u) & 1u), ovf_of((q + 1u) & 1u),
                        (epoch << 16) + q + 2u};
        compact<true>(p, next, ro, n, local, lane, sInfo);
    }
    store_state(p, s, n, x);
    if (kSt && s.lastNoise) s.lastNoise[n] = last;
}

__global__ void k_fill_u32(uint32_t *a, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_dep_current(vc_v1_params p, const float *__restrict__ tonicDep,
                              float *__restrict__ out) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= p.nV1) return;
    out[n] = dep_current(p, type_of(p, n % p.blockSize), tonicDep[n]);
}

__global__ void k_clear_masks(uint64_t *mask, uint32_t nTiles) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nTiles) mask[i] = 0ull;
}

// Voltage noise, synaptic failure or a refractory period shorter than dt
// (several spikes per step): the kernels' instances with the draws and the
// multi-spike update.
bool has_noise_or_failure(const vc_v1_params *p) {
    for (uint32_t t = 0; t < p->nType; ++t) {
        if (p->noisyDep[t] > 0.0f || p->synFailFF[t] > 0.0f || p->tRef[t] < p->dt) return true;
        for (uint32_t u = 0; u < p->nType; ++u)
            if (p->synFail[t * p->nType + u] > 0.0f) return true;
    }
    return false;
}

// Launch of the one-step kernel with gather variant g (0 = dense rows, else
// the masked-CSR tile) and wg threads per workgroup.
template <bool kUpdate, bool kEvents>
void launch_step(uint32_t g, uint32_t wg, hipStream_t st, const vc_v1_params &p,
                 const V1Const &k, const vc_v1_state &s, const vc_ff_conn &ff,
                 const vc_rec_conn &rec, const vc_spikes &prev, const vc_spikes &next,
                 const float *lgn, uint32_t *fm, uint32_t step, float *o_sInfo, float *o_dep,
                 float *o_g, float *o_snap) {
    const dim3 grid(p.nV1 / wg), block(wg);
    const bool sto = has_noise_or_failure(&p);
#define VC_STEP(G, S)                                                                           \
    hipLaunchKernelGGL((k_v1_step<kUpdate, kEvents, G, S>), grid, block, 0, st, p, k, s, ff,    \
                       rec, prev, next, lgn, fm, step, o_sInfo, o_dep, o_g, o_snap)
    if (sto) {
        switch (g) {
        case 0: VC_STEP(0, true); break;
        case 4: VC_STEP(4, true); break;
        case 16: VC_STEP(16, true); break;
        default: VC_STEP(8, true); break;
        }
    } else {
        switch (g) {
        case 0: VC_STEP(0, false); break;
        case 4: VC_STEP(4, false); break;
        case 16: VC_STEP(16, false); break;
        default: VC_STEP(8, false); break;
        }
    }
#undef VC_STEP
}

// The launch shape of the one-step kernel: the gather variant (0 when the
// dense rows are given, else p->gatherTile or the default tile) and the
// workgroup size (p->stepWG or the largest of 256/128/64 dividing blockSize).
int step_shape(const vc_v1_params *p, const vc_rec_conn *rec, uint32_t *g, uint32_t *wg) {
    const uint32_t t = p->gatherTile ? p->gatherTile : kTileStep;
    if (t != 4 && t != 8 && t != 16) {
        vc_set_error("gatherTile %u must be 0 (default), 4, 8 or 16", p->gatherTile);
        return VC_E_INVALID;
    }
    *g = rec->wDense ? 0u : t;
    *wg = p->stepWG ? p->stepWG : wg_size(p->blockSize);
    if (*wg % 64 != 0 || *wg > (uint32_t)kWG || p->blockSize % *wg != 0) {
        vc_set_error("stepWG %u must be 64, 128 or 256 and divide blockSize %u", p->stepWG,
                     p->blockSize);
        return VC_E_INVALID;
    }
    return VC_OK;
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
        // tRef < dt selects the multi-spike update; a lower bound keeps its
        // per-step loop short (at most 513 spikes in a step)
        if (!(p->tRef[t] >= p->dt / 512.0f)) {
            vc_set_error("tRef[%u]=%g must be at least dt/512 (dt=%g)", t, p->tRef[t], p->dt);
            return VC_E_INVALID;
        }
        if (p->deltaT[t] < 0.0f) { vc_set_error("deltaT < 0"); return VC_E_INVALID; }
    }
    for (uint32_t t = 0; t < p->nType; ++t) {
        if (!(p->synFailFF[t] >= 0.0f && p->synFailFF[t] <= 1.0f) ||
            (p->synFailFF[t] > 0.0f && !(p->synPerConFF[t] > 0.0f)) || !(p->noisyDep[t] >= 0.0f)) {
            vc_set_error("type %u: synFailFF in [0, 1] with synPerConFF > 0, noisyDep >= 0", t);
            return VC_E_INVALID;
        }
        for (uint32_t u = 0; u < p->nType; ++u) {
            const float f = p->synFail[t * p->nType + u];
            if (!(f >= 0.0f && f <= 1.0f) || (f > 0.0f && !(p->synPerCon[t * p->nType + u] > 0.0f))) {
                vc_set_error("synFail[%u][%u] must be in [0, 1] with synPerCon > 0", t, u);
                return VC_E_INVALID;
            }
        }
    }
    if (!(p->tauNoise >= 0.0f)) { vc_set_error("tauNoise < 0"); return VC_E_INVALID; }
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



// The OU noise needs its state array.
int noise_ok(const vc_v1_params *p, const vc_v1_state *s, const char *who) {
    bool noisy = false;
    for (uint32_t t = 0; t < p->nType; ++t) noisy = noisy || p->noisyDep[t] > 0.0f;
    if (noisy && p->tauNoise > 0.0f && !s->lastNoise) {
        vc_set_error("%s: Ornstein-Uhlenbeck noise (tauNoise > 0) needs s->lastNoise", who);
        return VC_E_INVALID;
    }
    return VC_OK;
}

// Gap junctions need s->gap to write to; the ELL arrays come together.
int gap_ok(const vc_v1_params *p, const vc_v1_state *s, const vc_rec_conn *rec,
           const char *who) {
    if (rec->gapPre && (!rec->nGap || !rec->gapG || !s->gap || rec->maxGap == 0)) {
        vc_set_error("%s: gap junctions need nGap, gapPre, gapG, maxGap > 0 and s->gap", who);
        return VC_E_INVALID;
    }
    if (rec->gapPre && p->nE >= p->blockSize) {
        vc_set_error("%s: gap junctions without inhibitory neurons", who);
        return VC_E_INVALID;
    }
    return VC_OK;
}

bool has_gap(const vc_v1_state *s, const vc_rec_conn *rec) {
    return s->gapS || s->gap || rec->gapPre;
}

}  // namespace

extern "C" int vc_v1_gap(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                         const vc_rec_conn *rec, const float *train, uint32_t depth,
                         uint32_t cur, const uint32_t *lag) {
    int rc = check_params(p, rec);
    if (rc) return rc;
    if (!s || !rec || !train || !rec->gapPre || depth == 0 || cur >= depth) {
        vc_set_error("vc_v1_gap: bad argument (gap junctions, train, depth > cur needed)");
        return VC_E_INVALID;
    }
    if ((rc = gap_ok(p, s, rec, "vc_v1_gap"))) return rc;
    const uint32_t mI = p->nblock * (p->blockSize - p->nE);
    hipLaunchKernelGGL(k_v1_gap, dim3((mI + 255) / 256), dim3(256), 0, (hipStream_t)stream, *p,
                       *rec, train, depth, cur, lag, mI, s->gap);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

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
    if ((rc = noise_ok(p, s, "vc_v1_step"))) return rc;
    if (prev->mask == next->mask) {
        vc_set_error("vc_v1_step: prev and next spike buffers must differ");
        return VC_E_INVALID;
    }
    if (ffMask && ff->maxLGN > 32u) {
        vc_set_error("vc_v1_step: event-driven LGN input needs maxLGN <= 32 (got %u)", ff->maxLGN);
        return VC_E_INVALID;
    }
    V1Const k = make_const(p);
    uint32_t g, wg;
    if ((rc = step_shape(p, rec, &g, &wg))) return rc;
    if (ffMask)
        launch_step<true, true>(g, wg, (hipStream_t)stream, *p, k, *s, *ff, *rec, *prev, *next,
                                lgnSInfo, ffMask, step, out_sInfo, out_dep, out_g_prev, nullptr);
    else
        launch_step<true, false>(g, wg, (hipStream_t)stream, *p, k, *s, *ff, *rec, *prev, *next,
                                 lgnSInfo, ffMask, step, out_sInfo, out_dep, out_g_prev, nullptr);
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
    if (has_gap(s, rec) || has_noise_or_failure(p)) {
        vc_set_error("vc_v1_step_snapshot: the adjoint's forward has no gap junctions, noise, "
                     "synaptic failure or tRef < dt");
        return VC_E_UNSUPPORTED;
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
    uint32_t g, wg;
    if ((rc = step_shape(p, rec, &g, &wg))) return rc;
    if (ffMask)
        launch_step<true, true>(g, wg, (hipStream_t)stream, *p, k, *s, *ff, *rec, *prev, *next,
                                lgnSInfo, ffMask, step, out_sInfo, nullptr, nullptr, out_snap);
    else
        launch_step<true, false>(g, wg, (hipStream_t)stream, *p, k, *s, *ff, *rec, *prev, *next,
                                 lgnSInfo, ffMask, step, out_sInfo, nullptr, nullptr, out_snap);
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
    if ((rc = gap_ok(p, s, rec, "vc_v1_run"))) return rc;
    if ((rc = noise_ok(p, s, "vc_v1_run"))) return rc;
    V1Const k = make_const(p);
    uint32_t g, wg;
    if ((rc = step_shape(p, rec, &g, &wg))) return rc;
    const uint32_t mI = p->nblock * (p->blockSize - p->nE);
    for (uint32_t i = 0; i < nsteps; ++i) {
        const uint32_t nx = 1u - cur;
        if (ffMask)
            launch_step<true, true>(g, wg, (hipStream_t)stream, *p, k, *s, *ff, *rec, spk2[cur],
                                    spk2[nx], lgnSInfo + i * lgnStride, ffMask + i * maskStride,
                                    it0 + i, out_sInfo + i * outStride, nullptr, nullptr, nullptr);
        else
            launch_step<true, false>(g, wg, (hipStream_t)stream, *p, k, *s, *ff, *rec, spk2[cur],
                                     spk2[nx], lgnSInfo + i * lgnStride, nullptr, it0 + i,
                                     out_sInfo + i * outStride, nullptr, nullptr, nullptr);
        if (rec->gapPre && mI)          // recal_G's gap sums of this step, for the next
            hipLaunchKernelGGL(k_v1_gap, dim3((mI + 255) / 256), dim3(256), 0,
                               (hipStream_t)stream, *p, *rec, out_sInfo + i * outStride, 1u, 0u,
                               nullptr, mI, s->gap);
        cur = nx;
    }
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" uint64_t vc_v1_persistent_scratch_bytes(const vc_v1_params *p) {
    if (!p) return 0;
    const uint64_t nTiles = p->nV1 / 64u;
    return 2ull * nTiles * (kRecW + 64u * 4u) * sizeof(uint64_t);
}

namespace {

// The persistent launch behind vc_v1_run_persistent and its allocation-free
// variant: `words` is the record buffer, `clearWords32` the number of 32-bit
// words cleared (by a fill kernel, so that the call may be captured) when
// epoch is 1.  *launched (may be null) tells whether kernels were enqueued:
// a refused call, or one of 0 steps, uses up no epoch.
int run_persistent(const char *who, void *stream, const vc_v1_params *p, const vc_v1_state *s,
                   const vc_ff_conn *ff, const vc_rec_conn *rec, const vc_spikes *spk2,
                   uint32_t cur, const float *lgnSInfo, uint64_t lgnStride, uint32_t *ffMask,
                   uint64_t maskStride, uint32_t it0, uint32_t nsteps, float *out_sInfo,
                   uint64_t outStride, uint64_t *words, uint32_t epoch, uint64_t clearWords32,
                   uint32_t *status, bool *launched = nullptr) {
    if (launched) *launched = false;
    int rc = check_params(p, rec);
    if (rc) return rc;
    if (!s || !ff || !rec || !spk2 || !lgnSInfo || !out_sInfo || !words || !status || cur > 1) {
        vc_set_error("%s: bad argument", who);
        return VC_E_INVALID;
    }
    if (spk2[0].mask == spk2[1].mask) {
        vc_set_error("%s: the two spike buffers must differ", who);
        return VC_E_INVALID;
    }
    if (ffMask && ff->maxLGN > 32u) {
        vc_set_error("%s: event-driven LGN input needs maxLGN <= 32", who);
        return VC_E_INVALID;
    }
    if (rec->maxNear > 64u) {
        vc_set_error("%s: more than 64 neighbour blocks", who);
        return VC_E_UNSUPPORTED;
    }
    if (has_gap(s, rec)) {
        vc_set_error("%s: gap junctions need the partners' voltages of every step (vc_v1_run)",
                     who);
        return VC_E_UNSUPPORTED;
    }
    if ((rc = noise_ok(p, s, who))) return rc;
    if (epoch == 0 || epoch > 0xffffu || nsteps > 0xfffdu) {
        vc_set_error("%s: epoch %u (1..65535) or %u steps (<= 65533) out of range", who, epoch,
                     nsteps);
        return VC_E_INVALID;
    }
    if (nsteps == 0) return VC_OK;
    const uint32_t wg = wg_size(p->blockSize);
    const uint32_t grid = p->nV1 / wg;
    constexpr int kT = (int)kTileMulti;
    const bool dense = rec->wDense != nullptr;
    const bool sto = has_noise_or_failure(p);
    const void *fn = sto ? (ffMask ? (dense ? (const void *)k_v1_multi<true, 0, true>
                                            : (const void *)k_v1_multi<true, kT, true>)
                                   : (dense ? (const void *)k_v1_multi<false, 0, true>
                                            : (const void *)k_v1_multi<false, kT, true>))
                         : (ffMask ? (dense ? (const void *)k_v1_multi<true, 0, false>
                                            : (const void *)k_v1_multi<true, kT, false>)
                                   : (dense ? (const void *)k_v1_multi<false, 0, false>
                                            : (const void *)k_v1_multi<false, kT, false>));
    int dev = 0, ncu = 0, per = 0;
    VC_HIP_CHECK(hipGetDevice(&dev));
    VC_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    VC_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, (int)wg, 0));
    if ((uint64_t)per * (uint64_t)ncu < grid) {
        vc_set_error("%s: %u workgroups cannot be co-resident (%d per CU x %d CUs); use "
                     "vc_v1_run", who, grid, per, ncu);
        return VC_E_UNSUPPORTED;
    }
    if (((uintptr_t)words & 63u) != 0) {
        vc_set_error("%s: scratch must be 64-byte aligned", who);
        return VC_E_INVALID;
    }
    const uint64_t nWords32 = vc_v1_persistent_scratch_bytes(p) / 4u;
    if (nWords32 >= (1ull << 32) || clearWords32 >= (1ull << 32) ||
        2ull * (p->nV1 / 64u) * (kRecW + 256u) >= (1ull << 32)) {
        vc_set_error("%s: nV1 %u too large for 32-bit record offsets", who, p->nV1);
        return VC_E_UNSUPPORTED;
    }
    hipStream_t st = (hipStream_t)stream;
    if (launched) *launched = true;
    if (epoch == 1) {               // a new cycle of epochs: no tag of an earlier one may match
        hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)((clearWords32 + 255) / 256)), dim3(256),
                           0, st, (uint32_t *)words, (uint32_t)clearWords32, 0u);
        VC_HIP_CHECK(hipGetLastError());
    }
    // An ordinary launch: every workgroup fits on the device at once (checked
    // above), the only waits are of these workgroups on one another and are
    // bounded, so kernels of other streams can at most delay some of them.
    // (A cooperative launch would also serialise the host with the device.)
    const V1Const k = make_const(p);
#define VC_MULTI(E, G, S)                                                                       \
    hipLaunchKernelGGL((k_v1_multi<E, G, S>), dim3(grid), dim3(wg), 0, st, *p, k, *s, *ff, *rec, \
                       spk2[cur], spk2[1u - cur], lgnSInfo, lgnStride, ffMask, maskStride, it0,  \
                       nsteps, out_sInfo, outStride, words, epoch, status)
    if (sto) {
        if (ffMask) {
            if (dense) VC_MULTI(true, 0, true); else VC_MULTI(true, kT, true);
        } else {
            if (dense) VC_MULTI(false, 0, true); else VC_MULTI(false, kT, true);
        }
    } else {
        if (ffMask) {
            if (dense) VC_MULTI(true, 0, false); else VC_MULTI(true, kT, false);
        } else {
            if (dense) VC_MULTI(false, 0, false); else VC_MULTI(false, kT, false);
        }
    }
#undef VC_MULTI
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

// ---- records of the allocation-free persistent launch ----------------------
// A file-scope device array: every GPU the library runs on holds its own copy
// until the process ends.  Host code never passes its name to a kernel (on
// the host it is only a stand-in symbol); it looks up the device address per
// GPU once with hipGetSymbolAddress and uses that.
constexpr uint32_t kNoAllocMaxV1 = 131072u;            // 128 blocks of 1024 neurons
constexpr uint64_t kNoAllocWords = 2ull * (kNoAllocMaxV1 / 64u) * (kRecW + 64u * 4u);
static __device__ __align__(256) uint64_t noalloc_scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
struct NoAllocDevice {
    uint64_t *words = nullptr;      // device address of noalloc_scratch on this GPU
    uint32_t epoch = 0;             // of the last launch on this GPU (1..65535, cycling)
};
std::mutex g_noalloc_mu;
NoAllocDevice g_noalloc[kMaxDevices];

}  // namespace

extern "C" int vc_v1_run_persistent(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                                    const vc_ff_conn *ff, const vc_rec_conn *rec,
                                    const vc_spikes *spk2, uint32_t cur, const float *lgnSInfo,
                                    uint64_t lgnStride, uint32_t *ffMask, uint64_t maskStride,
                                    uint32_t it0, uint32_t nsteps, float *out_sInfo,
                                    uint64_t outStride, void *scratch, uint32_t epoch,
                                    uint32_t *status) {
    if (!p) { vc_set_error("null params"); return VC_E_INVALID; }
    return run_persistent("vc_v1_run_persistent", stream, p, s, ff, rec, spk2, cur, lgnSInfo,
                          lgnStride, ffMask, maskStride, it0, nsteps, out_sInfo, outStride,
                          (uint64_t *)scratch, epoch, vc_v1_persistent_scratch_bytes(p) / 4u,
                          status);
}

extern "C" uint32_t vc_v1_run_persistent_noalloc_max_v1(void) { return kNoAllocMaxV1; }

extern "C" int vc_v1_run_persistent_noalloc(void *stream, const vc_v1_params *p,
                                            const vc_v1_state *s, const vc_ff_conn *ff,
                                            const vc_rec_conn *rec, const vc_spikes *spk2,
                                            uint32_t cur, const float *lgnSInfo,
                                            uint64_t lgnStride, uint32_t *ffMask,
                                            uint64_t maskStride, uint32_t it0, uint32_t nsteps,
                                            float *out_sInfo, uint64_t outStride,
                                            uint32_t *status) {
    const char *who = "vc_v1_run_persistent_noalloc";
    if (!p) { vc_set_error("%s: null params", who); return VC_E_INVALID; }
    if (p->nV1 > kNoAllocMaxV1) {
        vc_set_error("%s: nV1 %u exceeds the %u neurons its static record buffer holds; use "
                     "vc_v1_run_persistent with a caller buffer", who, p->nV1, kNoAllocMaxV1);
        return VC_E_INVALID;
    }
    int dev = 0;
    VC_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        vc_set_error("%s: device ordinal %d beyond %d", who, dev, kMaxDevices);
        return VC_E_UNSUPPORTED;
    }
    // The lock covers the whole call (it only enqueues work): the epoch is
    // used up only by a call that launches, so that epoch 1, which clears the
    // whole array, is never skipped.  The clear matters because a later call
    // with more neurons must not meet tags that an earlier cycle of epochs
    // left beyond this call's records.
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    NoAllocDevice &d = g_noalloc[dev];
    if (!d.words) {
        void *addr = nullptr;
        VC_HIP_CHECK(hipGetSymbolAddress(&addr, HIP_SYMBOL(noalloc_scratch)));
        d.words = (uint64_t *)addr;
    }
    const uint32_t epoch = d.epoch % 0xffffu + 1u;
    bool launched = false;
    const int rc = run_persistent(who, stream, p, s, ff, rec, spk2, cur, lgnSInfo, lgnStride,
                                  ffMask, maskStride, it0, nsteps, out_sInfo, outStride, d.words,
                                  epoch, kNoAllocWords * 2u, status, &launched);
    if (launched) d.epoch = epoch;
    return rc;
}

static bool any_syn_fail(const vc_v1_params *p) {
    for (uint32_t i = 0; i < p->nType * p->nType; ++i)
        if (p->synFail[i] > 0.0f) return true;
    return false;
}

extern "C" int vc_v1_rec_apply(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                               const vc_rec_conn *rec, const vc_spikes *prev) {
    if (p && any_syn_fail(p)) {
        vc_set_error("vc_v1_rec_apply: with synaptic failure use vc_v1_rec_apply_at (the "
                     "failure draws are keyed by the spikes' step)");
        return VC_E_UNSUPPORTED;
    }
    return vc_v1_rec_apply_at(stream, p, s, rec, prev, 0xffffffffu);
}

extern "C" int vc_v1_rec_apply_at(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                                  const vc_rec_conn *rec, const vc_spikes *prev, uint32_t step) {
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
    // (the gather of a k_v1_step launched for step s + 1 keys its failure
    // draws by s, the step of the spikes)
    launch_step<false, false>(rec->wDense ? 0u : kTileMulti, wg, (hipStream_t)stream, *p, k, *s,
                              ff, *rec, *prev, nx, nullptr, nullptr, step + 1u, nullptr, nullptr,
                              nullptr, nullptr);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_v1_dep_current(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                                 float *out_dep) {
    if (!p || !s || !s->tonicDep || !out_dep || p->blockSize == 0) {
        vc_set_error("vc_v1_dep_current: null argument");
        return VC_E_INVALID;
    }
    hipLaunchKernelGGL(k_dep_current, dim3((p->nV1 + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, *p, s->tonicDep, out_dep);
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
