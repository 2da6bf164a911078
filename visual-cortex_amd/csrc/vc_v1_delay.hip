// vc_v1_delay.hip -- recurrent input with axonal delays on gfx950: the
// reference's recal_G_mat (src/coredynamics.cu:2062-2309, used when noDelay
// is 0), which delivers every spike distance / speedOfThought after it was
// fired, from a ring of the last trainDepth steps' outputs.
//
// The reference scans, for every post neuron, every presynaptic neuron of its
// neighbour blocks and reads two ring slots of each connected one.  Here only
// the presynaptic neurons that fired inside the ring's window are visited
// (the "spikers"), found from the per-step spike masks; the work per post
// neuron is a sequential sum in the reference's order (neighbour blocks,
// ascending presynaptic id, slot j = 0 then 1, receptor), so the kernel is
// organised to keep many loads in flight per wavefront rather than to split
// a post neuron's sum.  The sums are added to gE/hE/gI/hI at the end, as the
// reference does.  The far connections' delayed input (recal_G_vec) reads
// the same ring.
#include "vc_v1_device.h"

namespace {

constexpr int kMaxEntD = 512;      // neighbour tiles of one block (maxNear * blockSize/64)
// spikers whose synapse records are loaded together: 16 when the LDS allows
// (staging the records and their pairs takes kGD*3.5 KiB per workgroup), else 8
constexpr uint32_t kLdsMax = 160u * 1024u;
constexpr int kMaxDepthLds = 32;   // rings up to this long are staged in LDS
constexpr int kStv = kMaxDepthLds + 1;  // LDS row stride of one spiker's ring values

// Distances of the masked-CSR strengths: one wavefront per row segment
// (row r = (b*maxNear + ib)*blockSize + pre, segment s of 64 posts).
__global__ __launch_bounds__(256) void k_rec_distances(uint32_t blockSize, uint32_t maxNear,
                                                       uint64_t nRowSeg,
                                                       const uint32_t *__restrict__ nNear,
                                                       const uint32_t *__restrict__ nbr,
                                                       const uint64_t *__restrict__ rowMask,
                                                       const uint32_t *__restrict__ rowOff,
                                                       const double *__restrict__ pos,
                                                       float *__restrict__ dist) {
    const uint64_t rs = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (rs >= nRowSeg) return;
    const uint32_t lane = vc_lane();
    const uint32_t nSeg = blockSize >> 6;
    const uint64_t r = rs / nSeg;
    const uint32_t s = (uint32_t)(rs - r * nSeg);
    const uint64_t bib = r / blockSize;                 // b*maxNear + ib
    const uint32_t pre = (uint32_t)(r - bib * blockSize);
    const uint32_t b = (uint32_t)(bib / maxNear), ib = (uint32_t)(bib - (uint64_t)b * maxNear);
    if (ib >= nNear[b]) return;
    const uint64_t m = rowMask[rs];
    if (!((m >> lane) & 1ull)) return;
    const uint64_t idx = (uint64_t)rowOff[rs] + vc_mbcnt(m);
    const uint64_t ipre = (uint64_t)nbr[bib] * blockSize + pre;
    const uint64_t ipost = (uint64_t)b * blockSize + s * 64u + lane;
    // src/preprocess/connect.cu:427-428: difference in double, then float
    const float x = (float)(pos[2 * ipre] - pos[2 * ipost]);
    const float y = (float)(pos[2 * ipre + 1] - pos[2 * ipost + 1]);
    dist[idx] = sqrtf(x * x + y * y);
}

// One 16-byte record per synapse: the strength, and the delay in steps and
// remainder the reference derives from the distance every step
// (src/coredynamics.cu:2173-2186): time2post = d / speed, it2post =
// ceil(time2post / dt), rem = it2post*dt - time2post; {w, rem, it2post, 0}.
__global__ __launch_bounds__(256) void k_rec_delay_table(uint64_t n, const float *__restrict__ w,
                                                         const float *__restrict__ dist,
                                                         float speed, float dt,
                                                         uint4 *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float time2post = dist[i] / speed;
    const uint32_t it2post = (uint32_t)ceilf(time2post / dt);
    time2post = (float)it2post * dt - time2post;
    out[i] = make_uint4(__float_as_uint(w[i]), __float_as_uint(time2post), it2post, 0u);
}

// One post neuron per thread, workgroups (up to 256 threads) inside one V1
// block, so every wavefront of a workgroup has the same neighbour tiles.
//   1. the spike masks of the ring's last `win` slots (the window the near
//      delays reach; the ring can be longer for the far connections) of every
//      neighbour tile are OR-ed into su[] (the window's spikers);
//   2. the window's spikers are listed in visiting order (neighbour block,
//      ascending presynaptic id) in items[] (a scan of su[]'s popcounts);
//   3. each wavefront walks the list 64 spikers at a time: lane k loads
//      spiker k's row-segment header and, for windows of up to kMaxDepthLds
//      slots, its `win` ring values into LDS, by age, together with a bit mask
//      of the ages holding a spike (one round trip); then the strengths and
//      delays of kGD spikers at a time are loaded (one round trip) and
//      accumulated in order, a lane reading ring values only when one of
//      its two slots holds a spike.
template <int kGD_>
__global__ __launch_bounds__(256) void k_rec_delay(vc_v1_params p, vc_v1_state s,
                                                   vc_rec_conn rec,
                                                   const uint4 *__restrict__ syn,
                                                   const float *__restrict__ train,
                                                   const uint64_t *__restrict__ tmask,
                                                   uint32_t depth, uint32_t cur, uint32_t win,
                                                   uint32_t *__restrict__ status,
                                                   float *__restrict__ copyOut) {
    __shared__ uint64_t su[kMaxEntD];
    __shared__ uint32_t soff[kMaxEntD + 1];
    __shared__ uint32_t sbid[kMaxEntD];
    extern __shared__ uint16_t items[];                    // [maxNear*blockSize] (dynamic)
    __shared__ float stv[4][64 * kStv];
    __shared__ float gw[4][kGD_][64], gr[4][kGD_][64];     // per wavefront: strengths,
    __shared__ uint32_t gi[4][kGD_][64];                   // remainders, steps of kGD spikers
    __shared__ uint16_t pl[4][kGD_ * 64];                  // (spiker, lane) pairs to evaluate
    __shared__ float4 tq[4][64][2];                        // their increments, per slot j
    __shared__ uint32_t tf[4][64];                         // accepted slots, presynaptic type
    __shared__ uint32_t slp[4][64], sgp[4][64];            // the batch's spikers: local, global id
    const uint32_t nV1 = p.nV1, bs = p.blockSize, nSeg = bs >> 6;
    const uint32_t nTiles = nV1 >> 6;
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t b = n / bs;
    const uint32_t local = n - b * bs;
    const uint32_t lane = vc_lane();
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t myseg = local >> 6;
    const int t = type_of(p, local);
    const float dt = p.dt;
    const uint32_t nb = rec.nNear[b];
    const uint32_t nEnt = nb * nSeg;
    // ---- 1. window spike masks per neighbour tile
    for (uint32_t e = threadIdx.x; e < nEnt; e += blockDim.x) su[e] = 0ull;
    for (uint32_t ib = threadIdx.x; ib < nb; ib += blockDim.x)
        sbid[ib] = rec.nbr[b * rec.maxNear + ib];
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < nEnt * win; q += blockDim.x) {
        const uint32_t a = q / nEnt, e = q - a * nEnt;
        const uint32_t k = cur >= a ? cur - a : cur + depth - a;
        const uint32_t ib = e / nSeg, tt = e - ib * nSeg;
        const uint64_t m = tmask[(size_t)k * nTiles + sbid[ib] * nSeg + tt];
        if (m) atomicOr((unsigned long long *)&su[e], (unsigned long long)m);
    }
    __syncthreads();
    // ---- 2. list of the window's spikers, e*64 + lane, in visiting order
    if (wv == 0) {
        uint32_t carry = 0;
        for (uint32_t c = 0; c < nEnt; c += 64) {
            const uint32_t e = c + lane;
            const uint32_t v = e < nEnt ? (uint32_t)__popcll(su[e]) : 0u;
            uint32_t incl = v;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(incl, d);
                if (lane >= (uint32_t)d) incl += y;
            }
            if (e < nEnt) soff[e] = carry + incl - v;
            carry += vc_readlane_u32(incl, 63);
        }
        if (lane == 0) soff[nEnt] = carry;
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < nEnt; e += blockDim.x) {
        uint64_t u = su[e];
        uint32_t o = soff[e];
        while (u) {
            items[o++] = (uint16_t)(e * 64u + (uint32_t)__builtin_ctzll(u));
            u &= u - 1ull;
        }
    }
    __syncthreads();
    const uint32_t total = __builtin_amdgcn_readfirstlane(soff[nEnt]);

    float ipE[VC_MAX_NG], ipI[VC_MAX_NG];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        ipE[ig] = ig < p.ngE ? p.pE[t * p.ngE + ig] : 0.0f;
        ipI[ig] = ig < p.ngI ? p.pI[t * p.ngI + ig] : 0.0f;
    }
    float lgE[VC_MAX_NG] = {0.f, 0.f}, lhE[VC_MAX_NG] = {0.f, 0.f};
    float lgI[VC_MAX_NG] = {0.f, 0.f}, lhI[VC_MAX_NG] = {0.f, 0.f};
    bool late = false;
    const bool staged = win <= (uint32_t)kMaxDepthLds;
    float *sv = stv[wv];

    // ---- 3. the spikers, 64 at a time (every wavefront of the workgroup
    // takes the same trips: the list is the workgroup's)
    for (uint32_t base = 0; base < total; base += 64u) {
        const uint32_t nIt = min(64u, total - base);
        uint64_t hm = 0ull;
        uint32_t ho = 0u, lpre = 0u, gpre = 0u;
        uint32_t am = ~0u;                             // bit a: a spike a steps before cur
        if (lane < nIt) {
            const uint32_t it = items[base + lane];
            const uint32_t e = it >> 6, ib = e / nSeg, tt = e - ib * nSeg;
            lpre = tt * 64u + (it & 63u);
            const size_t rs = (((size_t)b * rec.maxNear + ib) * bs + lpre) * nSeg + myseg;
            hm = rec.rowMask[rs];
            ho = rec.rowOff[rs];
            gpre = sbid[ib] * bs + lpre;
            slp[wv][lane] = lpre;
            sgp[wv][lane] = gpre;
            if (staged) {                              // by age a = 0 .. win-1
                am = 0u;
                for (uint32_t a0 = 0; a0 < win; a0 += 8u) {
                    float x[8];
#pragma unroll
                    for (uint32_t i = 0; i < 8u; ++i) {
                        const uint32_t a = a0 + i;
                        const uint32_t k = cur >= a ? cur - a : cur + depth - a;
                        x[i] = a < win ? train[(size_t)k * nV1 + gpre] : 0.0f;
                    }
#pragma unroll
                    for (uint32_t i = 0; i < 8u; ++i) {
                        const uint32_t a = a0 + i;
                        if (a < win) {
                            sv[lane * kStv + a] = x[i];
                            am |= (x[i] >= 1.0f ? 1u : 0u) << a;
                        }
                    }
                }
            }
        }
        __syncthreads();
        // (a) the records of all the batch's spikers in flight together: their
        // random 16-byte reads are the kernel's longest round trip
        uint64_t mine = 0ull;                          // bit k: lane has a synapse from spiker k
        float wv_[64], rem[64];
        uint32_t i2p[64];
#pragma unroll
        for (int k = 0; k < 64; ++k) {
            wv_[k] = rem[k] = 0.0f;
            i2p[k] = 0u;
            if ((uint32_t)k < nIt) {
                const uint64_t m = vc_readlane_u64(hm, k);
                if ((m >> lane) & 1ull) {
                    const uint32_t idx = vc_readlane_u32(ho, k) + vc_mbcnt(m);
                    const uint4 r = syn[idx];
                    wv_[k] = __uint_as_float(r.x);
                    rem[k] = __uint_as_float(r.y);
                    i2p[k] = r.z;
                    mine |= 1ull << k;
                }
            }
        }
        // then kGD spikers at a time
#pragma unroll
        for (int c = 0; c < 64 / kGD_; ++c) {
            const uint32_t g0 = (uint32_t)c * kGD_;
            if (g0 >= nIt) break;
            // (b) the synapses one of whose two slots holds a spike, staged in LDS
            uint32_t act = 0u;
#pragma unroll
            for (int g = 0; g < kGD_; ++g) {
                if (!((mine >> (g0 + g)) & 1ull)) continue;
                const uint32_t it2post = i2p[g0 + g];
                if (it2post >= win) {                 // window too short for this delay
                    late = true;
                    continue;
                }
                if (staged) {   // slot j = 0 is it2post steps old, j = 1 one step newer
                    // (for it2post = 0, slot j = 1 is never accepted: its dtsp is >= dt)
                    const uint32_t a = vc_readlane_u32(am, g0 + g);
                    const uint32_t need = it2post ? (a >> (it2post - 1u)) & 3u : a & 1u;
                    if (!need) continue;
                }
                act |= 1u << g;
                gw[wv][g][lane] = wv_[g0 + g];
                gr[wv][g][lane] = rem[g0 + g];
                gi[wv][g][lane] = it2post;
            }
            // (c) the (spiker, lane) pairs to evaluate, listed spiker-major, so
            // each lane's pairs are in its summation order
            uint32_t mypos[kGD_];
            uint32_t P = 0u;
#pragma unroll
            for (int g = 0; g < kGD_; ++g) {
                const uint64_t m = __ballot((act >> g) & 1u);
                mypos[g] = P + vc_mbcnt(m);
                if ((act >> g) & 1u) pl[wv][mypos[g]] = (uint16_t)(lane | ((uint32_t)g << 6));
                P += (uint32_t)__popcll(m);
            }
            // (d) 64 pairs at a time, one per lane: the increments of src/coredynamics.cu
            // :2173-2235 (single_input's two terms per receptor, per accepted slot)
            for (uint32_t pb = 0; pb < P; pb += 64u) {
                const uint32_t pi = pb + lane;
                uint32_t fl = 0u;
                float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0;
                if (pi < P) {
                    const uint32_t code = pl[wv][pi];
                    const uint32_t l = code & 63u, g = code >> 6, kk = g0 + g;
                    const float f = gw[wv][g][l];
                    const float time2post = gr[wv][g][l];
                    const uint32_t it2post = gi[wv][g][l];
                    const uint32_t j0 = cur >= it2post ? cur - it2post : cur + depth - it2post;
                    const uint32_t j1 = j0 + 1u == depth ? 0u : j0 + 1u;
                    float tv0, tv1;
                    if (staged) {                      // by age: it2post, it2post - 1
                        tv0 = sv[kk * kStv + it2post];
                        tv1 = it2post ? sv[kk * kStv + it2post - 1u] : 0.0f;
                    } else {
                        tv0 = train[(size_t)j0 * nV1 + sgp[wv][kk]];
                        tv1 = train[(size_t)j1 * nV1 + sgp[wv][kk]];
                    }
                    // shapes by the presynaptic type, split by the post neuron's type
                    const bool inh = slp[wv][kk] >= p.nE;
                    const int to = type_of(p, myseg * 64u + l);
                    const uint32_t ng = inh ? p.ngI : p.ngE;
                    float ip[VC_MAX_NG], rise[VC_MAX_NG], decay[VC_MAX_NG], dod[VC_MAX_NG];
#pragma unroll
                    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
                        ip[ig] = ig < ng ? (inh ? p.pI[to * p.ngI + ig] : p.pE[to * p.ngE + ig])
                                         : 0.0f;
                        rise[ig] = inh ? p.riseI[ig] : p.riseE[ig];
                        decay[ig] = inh ? p.decayI[ig] : p.decayE[ig];
                        dod[ig] = inh ? p.dodI[ig] : p.dodE[ig];
                    }
                    fl = inh ? 4u : 0u;
#pragma unroll
                    for (int jj = 0; jj < 2; ++jj) {
                        float dtsp = jj ? tv1 : tv0;
                        if (dtsp >= 1.0f) {
                            const float nsp = floorf(dtsp);
                            dtsp = (dtsp - nsp + (float)jj) * dt - time2post;
                            if (dtsp < dt && dtsp >= 0.0f && f > 0.0f) {
                                const float ddt = dt - dtsp;
                                const float a = nsp * f;
                                float tg[VC_MAX_NG], th[VC_MAX_NG];
#pragma unroll
                                for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
                                    const float fi = a * ip[ig];
                                    const float etr = vc_expf(-ddt / rise[ig]);
                                    tg[ig] = fi * dod[ig] * (vc_expf(-ddt / decay[ig]) - etr);
                                    th[ig] = fi * etr;
                                }
                                const float4 q = make_float4(tg[0], th[0], tg[1], th[1]);
                                if (jj) q1 = q; else q0 = q;
                                fl |= 1u << jj;
                            }
                        }
                    }
                }
                tq[wv][lane][0] = q0;
                tq[wv][lane][1] = q1;
                tf[wv][lane] = fl;
                // (e) every lane adds its pairs of this pass in its order: spiker,
                // slot j, receptor (the I or E sums by value: +0 leaves the other
                // sum unchanged)
#pragma unroll
                for (int g = 0; g < kGD_; ++g) {
                    if (!((act >> g) & 1u)) continue;
                    const uint32_t q = mypos[g] - pb;
                    if (q >= 64u) continue;
                    const uint32_t f2 = tf[wv][q];
                    const bool inh = (f2 >> 2) & 1u;
                    const uint32_t ng = inh ? p.ngI : p.ngE;
#pragma unroll
                    for (int jj = 0; jj < 2; ++jj) {
                        if (!((f2 >> jj) & 1u)) continue;
                        const float4 v = tq[wv][q][jj];
                        const float tg[VC_MAX_NG] = {v.x, v.z}, th[VC_MAX_NG] = {v.y, v.w};
#pragma unroll
                        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
                            if (ig < ng) {
                                lgE[ig] = lgE[ig] + (inh ? 0.0f : tg[ig]);
                                lhE[ig] = lhE[ig] + (inh ? 0.0f : th[ig]);
                                lgI[ig] = lgI[ig] + (inh ? tg[ig] : 0.0f);
                                lhI[ig] = lhI[ig] + (inh ? th[ig] : 0.0f);
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();                               // before sv is rewritten
    }
    if (late && status) atomicOr(status, 1u);
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        if (ig < p.ngE) {
            s.gE[ig * nV1 + n] = s.gE[ig * nV1 + n] + lgE[ig];
            s.hE[ig * nV1 + n] = s.hE[ig * nV1 + n] + lhE[ig];
        }
        if (ig < p.ngI) {
            s.gI[ig * nV1 + n] = s.gI[ig * nV1 + n] + lgI[ig];
            s.hI[ig * nV1 + n] = s.hI[ig * nV1 + n] + lhI[ig];
        }
    }
    if (copyOut) copyOut[n] = train[(size_t)cur * nV1 + n];   // the step's output row
}

// Far (long-range) connections with delays: the reference's recal_G_vec,
// src/coredynamics.cu:181-372 (noFarDelay 0, synFail 0), whose sums sum_G
// (:2314-2349) adds to gE/hE/gI/hI after recal_G_mat.  Far spikes are rare,
// so the work is driven by them rather than by the 2-million-entry lists:
//   k_far_mark: one lane per (age a, 64-neuron tile) of the ring finds the
//     far presynaptic neurons that fired a steps ago; for each, the
//     wavefront's lanes go through its connections together, and those whose
//     delay puts that spike in one of their two slots (it2post = a for slot
//     j = 0, a + 1 for slot j = 1) get their bit set in the post neuron's
//     hit words (an OR, so the order of the marks does not matter);
//   k_far_eval: one thread per post neuron takes its hit words (and clears
//     them) and evaluates the marked connections in list order, reading their
//     strengths, distances and ring values, so the per-post sum is the
//     reference's.  The ring holds the same steps' outputs as the
//     reference's per-connection rings, and a slot's value is >= 1 exactly
//     when its mask bit is set.
// As in the reference (src/patch.cu:6597-6598, src/stats.cu:114-126) nothing
// is added in a step in which no far presynaptic neuron has an output above
// 1.0: a workgroup with anything to add tests that from the step's spike
// masks (fc.preMask & trainMask[cur], then the values of the bits set).
__global__ __launch_bounds__(256) void k_far_mark(uint32_t nV1, float dt, vc_far_conn fc,
                                                  const uint64_t *__restrict__ trainMask,
                                                  uint32_t depth, uint32_t cur, float speed,
                                                  uint32_t *__restrict__ status) {
    const uint32_t nTiles = nV1 >> 6;
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = vc_lane();
    uint32_t a = 0u, t = 0u;
    uint64_t m = 0ull;
    if (q < (uint64_t)depth * nTiles) {
        a = (uint32_t)(q / nTiles);
        t = (uint32_t)(q - (uint64_t)a * nTiles);
        const uint32_t k = cur >= a ? cur - a : cur + depth - a;  // the slot of age a
        m = trainMask[(size_t)k * nTiles + t] & fc.preMask[t];
    }
    const uint32_t hw = (fc.maxFar + 63u) >> 6;
    uint32_t bad = 0u;
    // the wavefront takes its lanes' spikes one at a time, the 64 lanes
    // sharing the spiking neuron's connections
    uint64_t live = __ballot(m != 0ull);
    while (live) {
        const uint32_t src = (uint32_t)__builtin_ctzll(live);
        const uint64_t ms = vc_readlane_u64(m, src);
        const uint32_t aa = vc_readlane_u32(a, src);
        const uint32_t pre = vc_readlane_u32(t, src) * 64u + (uint32_t)__builtin_ctzll(ms);
        if (lane == src) m &= m - 1ull;
        const uint32_t r1 = fc.outOff[pre + 1];
        for (uint32_t r = fc.outOff[pre] + lane; r < r1; r += 64u) {
            const uint32_t post = fc.outPost[r], j = fc.outJ[r];
            if (post >= nV1 || j >= fc.maxFar) { bad |= 2u; continue; }
            const float d = fc.dist[(size_t)j * nV1 + post];
            const uint32_t it2post = (uint32_t)ceilf((d / speed) / dt);
            if (it2post >= depth) { bad |= 1u; continue; }
            if (it2post == aa || it2post == aa + 1u)
                atomicOr((unsigned long long *)&fc.hit[(size_t)post * hw + (j >> 6)],
                         1ull << (j & 63u));
        }
        live = __ballot(m != 0ull);
    }
    if (bad && status) atomicOr(status, bad);
}

__global__ __launch_bounds__(256) void k_far_eval(vc_v1_params p, vc_v1_state s,
                                                  vc_far_conn fc,
                                                  const float *__restrict__ train,
                                                  const uint64_t *__restrict__ trainMask,
                                                  uint32_t depth, uint32_t cur, float speed) {
    const uint32_t nV1 = p.nV1, nTiles = nV1 >> 6;
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t cnt = n < nV1 ? fc.nFar[n] : 0u;
    const uint32_t hw = (fc.maxFar + 63u) >> 6;
    const float dt = p.dt;
    const int to = type_of(p, n % p.blockSize);
    float lgE[VC_MAX_NG], lhE[VC_MAX_NG], lgI[VC_MAX_NG], lhI[VC_MAX_NG];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) lgE[ig] = lhE[ig] = lgI[ig] = lhI[ig] = 0.0f;
    uint32_t any = 0u;
    for (uint32_t c = 0; n < nV1 && c < hw; ++c) {
        uint64_t hh = fc.hit[(size_t)n * hw + c];
        if (!hh) continue;
        fc.hit[(size_t)n * hw + c] = 0ull;
        while (hh) {                                  // marked connections in list order
            const uint32_t j = c * 64u + (uint32_t)__builtin_ctzll(hh);
            hh &= hh - 1ull;
            if (j >= cnt) continue;
            any = 1u;
            const size_t e = (size_t)j * nV1 + n;
            const uint32_t ipre = fc.pre[e];
            const float strength = fc.w[e];
            float time2post = fc.dist[e] / speed;
            const uint32_t it2post = (uint32_t)ceilf(time2post / dt);
            time2post = (float)it2post * dt - time2post;
            const uint32_t j0 = cur >= it2post ? cur - it2post : cur + depth - it2post;
            const uint32_t j1 = j0 + 1u == depth ? 0u : j0 + 1u;
            const float tv[2] = {train[(size_t)j0 * nV1 + ipre], train[(size_t)j1 * nV1 + ipre]};
            const bool inh = (ipre % p.blockSize) >= p.nE;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                float dtsp = tv[k];
                if (dtsp >= 1.0f) {
                    const float nsp = floorf(dtsp);
                    dtsp = (dtsp - nsp + (float)k) * dt - time2post;
                    if (dtsp < dt && dtsp >= 0.0f && strength > 0.0f) {
                        const float ddt = dt - dtsp;
                        const float a = nsp * strength;
                        if (inh) {
#pragma unroll
                            for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig)
                                if (ig < p.ngI)
                                    single_input(lgI[ig], lhI[ig], a * p.pI[to * p.ngI + ig],
                                                 ddt, p.riseI[ig], p.decayI[ig], p.dodI[ig]);
                        } else {
#pragma unroll
                            for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig)
                                if (ig < p.ngE)
                                    single_input(lgE[ig], lhE[ig], a * p.pE[to * p.ngE + ig],
                                                 ddt, p.riseE[ig], p.decayE[ig], p.dodE[ig]);
                        }
                    }
                }
            }
        }
    }
    // the reference's gate: some far presynaptic output of this step above 1.0
    // (of this trial's neurons when trials are batched: a workgroup lies in one
    // V1 block, and far connections stay inside a trial)
    if (!__syncthreads_or((int)any)) return;
    const uint32_t myTrial = p.blockTrial ? p.blockTrial[(blockIdx.x * blockDim.x) / p.blockSize]
                                          : 0u;
    int fired = 0;
    for (uint32_t t = threadIdx.x; t < nTiles && !fired; t += blockDim.x) {
        if (p.blockTrial && p.blockTrial[(t * 64u) / p.blockSize] != myTrial) continue;
        uint64_t m = trainMask[(size_t)cur * nTiles + t] & fc.preMask[t];
        while (m) {
            const uint32_t b = (uint32_t)__builtin_ctzll(m);
            m &= m - 1ull;
            if (train[(size_t)cur * nV1 + t * 64u + b] > 1.0f) { fired = 1; break; }
        }
    }
    if (!__syncthreads_or(fired)) return;
    if (n >= nV1 || cnt == 0u) return;                // sum_G: nVec[id] > 0 only
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        if (ig < p.ngE) {
            s.gE[ig * nV1 + n] = s.gE[ig * nV1 + n] + lgE[ig];
            s.hE[ig * nV1 + n] = s.hE[ig * nV1 + n] + lhE[ig];
        }
        if (ig < p.ngI) {
            s.gI[ig * nV1 + n] = s.gI[ig * nV1 + n] + lgI[ig];
            s.hI[ig * nV1 + n] = s.hI[ig * nV1 + n] + lhI[ig];
        }
    }
}

// Static LDS of k_rec_delay<kGD_> (its declarations above), bytes.
template <int kGD_>
constexpr size_t lds_static() {
    return sizeof(uint64_t) * kMaxEntD + sizeof(uint32_t) * (2 * kMaxEntD + 1) +
           sizeof(float) * 4 * 64 * kStv + 3 * sizeof(float) * 4 * kGD_ * 64 +
           sizeof(uint16_t) * 4 * kGD_ * 64 + sizeof(float) * 4 * 64 * 8 +
           sizeof(uint32_t) * 4 * 64 * 3;
}

bool rec_ok(const vc_v1_params *p, const vc_rec_conn *rec, const char *who) {
    if (!p || !rec || !rec->nNear || !rec->nbr || !rec->rowMask || !rec->rowOff || !rec->w) {
        vc_set_error("%s: null argument (the masked CSR is required)", who);
        return false;
    }
    if (p->blockSize == 0 || p->blockSize % 64 != 0 || p->nV1 != p->nblock * p->blockSize) {
        vc_set_error("%s: blockSize %u must be a multiple of 64 and nV1 = nblock*blockSize",
                     who, p->blockSize);
        return false;
    }
    if ((uint64_t)rec->maxNear * (p->blockSize / 64) > (uint64_t)kMaxEntD) {
        vc_set_error("%s: maxNear*blockSize/64 = %u exceeds %d", who,
                     rec->maxNear * (p->blockSize / 64), kMaxEntD);
        return false;
    }
    return true;
}

// The delayed recurrent kernels (k_rec_delay, k_far_eval) deliver every
// spike whole: they have no synaptic-failure draw.
bool rec_failure_refused(const vc_v1_params *p, const char *who) {
    for (uint32_t i = 0; i < VC_MAX_NTYPE * VC_MAX_NTYPE; ++i)
        if (p->synFail[i] > 0.0f) {
            vc_set_error("%s: synaptic failure between V1 neurons (synFail) is not "
                         "implemented with conduction delays", who);
            return true;
        }
    return false;
}

}  // namespace

extern "C" int vc_v1_rec_distances(void *stream, const vc_v1_params *p, const vc_rec_conn *rec,
                                   const double *pos, float *out_dist) {
    if (!rec_ok(p, rec, "vc_v1_rec_distances")) return VC_E_INVALID;
    if (!pos || !out_dist) {
        vc_set_error("vc_v1_rec_distances: null argument");
        return VC_E_INVALID;
    }
    const uint64_t nRowSeg = (uint64_t)p->nblock * rec->maxNear * p->blockSize *
                             (p->blockSize / 64);
    const uint64_t grid = (nRowSeg + 3) / 4;
    if (grid == 0) return VC_OK;
    if (grid > 0x7fffffffull) {
        vc_set_error("vc_v1_rec_distances: %llu row segments is too many for one launch",
                     (unsigned long long)nRowSeg);
        return VC_E_INVALID;
    }
    hipLaunchKernelGGL(k_rec_distances, dim3((uint32_t)grid), dim3(256), 0, (hipStream_t)stream,
                       p->blockSize, rec->maxNear, nRowSeg, rec->nNear, rec->nbr, rec->rowMask,
                       rec->rowOff, pos, out_dist);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_v1_rec_delay_table(void *stream, uint64_t n, const float *w, const float *dist,
                                     float speed, float dt, uint32_t *out_syn) {
    if (n == 0) return VC_OK;
    if (!w || !dist || !out_syn || !(speed > 0.0f) || !(dt > 0.0f)) {
        vc_set_error("vc_v1_rec_delay_table: need w, dist, out_syn, speed > 0 and dt > 0");
        return VC_E_INVALID;
    }
    if (((uintptr_t)out_syn & 15u) != 0) {
        vc_set_error("vc_v1_rec_delay_table: out_syn must be 16-byte aligned");
        return VC_E_INVALID;
    }
    const uint64_t grid = (n + 255) / 256;
    if (grid > 0x7fffffffull) {
        vc_set_error("vc_v1_rec_delay_table: %llu synapses is too many for one launch",
                     (unsigned long long)n);
        return VC_E_INVALID;
    }
    hipLaunchKernelGGL(k_rec_delay_table, dim3((uint32_t)grid), dim3(256), 0,
                       (hipStream_t)stream, n, w, dist, speed, dt, (uint4 *)out_syn);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

// vc_v1_rec_delay, optionally also copying ring row `cur` to copyOut[nV1]
static int rec_delay_launch(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                            const vc_rec_conn *rec, const uint32_t *syn, const float *train,
                            const uint64_t *trainMask, uint32_t depth, uint32_t cur,
                            uint32_t window, uint32_t *status, float *copyOut) {
    if (!rec_ok(p, rec, "vc_v1_rec_delay")) return VC_E_INVALID;
    if (rec_failure_refused(p, "vc_v1_rec_delay")) return VC_E_UNSUPPORTED;
    if (!s || !s->gE || !s->hE || !s->gI || !s->hI || !syn || !train || !trainMask) {
        vc_set_error("vc_v1_rec_delay: null argument");
        return VC_E_INVALID;
    }
    if (((uintptr_t)syn & 15u) != 0) {
        vc_set_error("vc_v1_rec_delay: syn must be 16-byte aligned");
        return VC_E_INVALID;
    }
    if (depth == 0 || cur >= depth || window > depth) {
        vc_set_error("vc_v1_rec_delay: need cur (%u) < depth (%u) and window (%u) <= depth",
                     cur, depth, window);
        return VC_E_INVALID;
    }
    const uint32_t win = window ? window : depth;
    if (p->ngE > VC_MAX_NG || p->ngI > VC_MAX_NG || p->nType > VC_MAX_NTYPE) {
        vc_set_error("vc_v1_rec_delay: receptor or type count out of range");
        return VC_E_INVALID;
    }
    const uint32_t wg = wg_size(p->blockSize);
    const size_t dyn = (size_t)rec->maxNear * p->blockSize * sizeof(uint16_t);
    const uint4 *sy = (const uint4 *)syn;
    if (dyn + lds_static<16>() + 1024u <= kLdsMax)     // margin for alignment padding
        hipLaunchKernelGGL(k_rec_delay<16>, dim3(p->nV1 / wg), dim3(wg), dyn, (hipStream_t)stream,
                           *p, *s, *rec, sy, train, trainMask, depth, cur, win, status,
                           copyOut);
    else
        hipLaunchKernelGGL(k_rec_delay<8>, dim3(p->nV1 / wg), dim3(wg), dyn, (hipStream_t)stream,
                           *p, *s, *rec, sy, train, trainMask, depth, cur, win, status,
                           copyOut);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_v1_rec_delay(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                               const vc_rec_conn *rec, const uint32_t *syn,
                               const float *train, const uint64_t *trainMask, uint32_t depth,
                               uint32_t cur, uint32_t window, uint32_t *status) {
    return rec_delay_launch(stream, p, s, rec, syn, train, trainMask, depth, cur, window,
                            status, nullptr);
}

extern "C" int vc_v1_rec_far(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                             const vc_far_conn *far, const float *train,
                             const uint64_t *trainMask, uint32_t depth, uint32_t cur,
                             float speed, uint32_t *status) {
    if (!p || !s || !far || !s->gE || !s->hE || !s->gI || !s->hI) {
        vc_set_error("vc_v1_rec_far: null argument");
        return VC_E_INVALID;
    }
    if (far->maxFar == 0 || p->nV1 == 0) return VC_OK;
    if (rec_failure_refused(p, "vc_v1_rec_far")) return VC_E_UNSUPPORTED;
    if (!far->nFar || !far->pre || !far->w || !far->dist || !far->preMask || !far->outOff ||
        !far->outPost || !far->outJ || !far->hit || !train || !trainMask) {
        vc_set_error("vc_v1_rec_far: null argument (every vc_far_conn array, train and "
                     "trainMask are required)");
        return VC_E_INVALID;
    }
    if (p->blockSize == 0 || p->blockSize % 64 != 0 || p->nV1 != p->nblock * p->blockSize) {
        vc_set_error("vc_v1_rec_far: need blockSize a multiple of 64 and nV1 = nblock*blockSize");
        return VC_E_INVALID;
    }
    if (depth == 0 || cur >= depth || !(speed > 0.0f) || !(p->dt > 0.0f)) {
        vc_set_error("vc_v1_rec_far: need cur (%u) < depth (%u), speed > 0 and dt > 0", cur,
                     depth);
        return VC_E_INVALID;
    }
    if (p->ngE > VC_MAX_NG || p->ngI > VC_MAX_NG || p->nType > VC_MAX_NTYPE) {
        vc_set_error("vc_v1_rec_far: receptor or type count out of range");
        return VC_E_INVALID;
    }
    const uint64_t nq = (uint64_t)depth * (p->nV1 / 64u);
    if ((nq + 255u) / 256u > 0x7fffffffull) {
        vc_set_error("vc_v1_rec_far: ring of %u steps is too long for one launch", depth);
        return VC_E_INVALID;
    }
    hipLaunchKernelGGL(k_far_mark, dim3((uint32_t)((nq + 255u) / 256u)), dim3(256), 0,
                       (hipStream_t)stream, p->nV1, p->dt, *far, trainMask, depth, cur, speed,
                       status);
    VC_HIP_CHECK(hipGetLastError());
    const uint32_t wg = wg_size(p->blockSize);          // workgroups inside one V1 block
    hipLaunchKernelGGL(k_far_eval, dim3(p->nV1 / wg), dim3(wg), 0,
                       (hipStream_t)stream, *p, *s, *far, train, trainMask, depth, cur, speed);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_v1_delay_run(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                               const vc_ff_conn *ff, const vc_rec_conn *rec, const uint32_t *syn,
                               const vc_far_conn *far, float *train, uint64_t *trainMask,
                               uint32_t depth, uint32_t window, uint32_t slot0,
                               const float *lgnSInfo, uint64_t lgnStride, uint32_t *ffMask,
                               uint64_t maskStride, uint32_t it0, uint32_t nsteps, float speed,
                               const uint64_t *zeroMask, float *og, float *oh,
                               float *out_sInfo, uint64_t outStride, uint32_t *status) {
    if (!p || !train || !trainMask || !lgnSInfo || !zeroMask || !og || !oh) {
        vc_set_error("vc_v1_delay_run: null argument");
        return VC_E_INVALID;
    }
    if (depth == 0 || slot0 >= depth) {
        vc_set_error("vc_v1_delay_run: need slot0 (%u) < depth (%u)", slot0, depth);
        return VC_E_INVALID;
    }
    if (rec_failure_refused(p, "vc_v1_delay_run")) return VC_E_UNSUPPORTED;
    const uint32_t nTiles = p->nV1 / 64u;
    const vc_spikes none = {(uint64_t *)zeroMask, og, oh};
    for (uint32_t i = 0; i < nsteps; ++i) {
        const uint32_t slot = (uint32_t)(((uint64_t)slot0 + i) % depth);
        const vc_spikes ring = {trainMask + (size_t)slot * nTiles, og, oh};
        float *row = train + (size_t)slot * p->nV1;
        int rc = vc_v1_step(stream, p, s, ff, rec, &none, &ring, lgnSInfo + i * lgnStride,
                            ffMask ? ffMask + i * maskStride : nullptr, it0 + i, row, nullptr,
                            nullptr);
        if (rc != VC_OK) return rc;
        if (syn) {
            // the delayed gather also copies the step's outputs to out_sInfo
            rc = rec_delay_launch(stream, p, s, rec, syn, train, trainMask, depth, slot, window,
                                  status, out_sInfo ? out_sInfo + i * outStride : nullptr);
        } else {
            // undelayed near connections: the step's own spikes (recal_G_mat_nd)
            rc = vc_v1_rec_apply_at(stream, p, s, rec, &ring, it0 + i);
            if (rc == VC_OK && out_sInfo)
                VC_HIP_CHECK(hipMemcpyAsync(out_sInfo + i * outStride, row,
                                            (size_t)p->nV1 * sizeof(float),
                                            hipMemcpyDeviceToDevice, (hipStream_t)stream));
        }
        if (rc != VC_OK) return rc;
        if (far && far->maxFar) {
            rc = vc_v1_rec_far(stream, p, s, far, train, trainMask, depth, slot, speed, status);
            if (rc != VC_OK) return rc;
        }
    }
    return VC_OK;
}
