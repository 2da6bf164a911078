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
// reference does.
#include "vc_v1_device.h"

namespace {

constexpr int kMaxEntD = 512;      // neighbour tiles of one block (maxNear * blockSize/64)
constexpr int kGD = 32;            // spikers whose strengths are loaded together
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

// Per-synapse delay in steps and the remainder the reference derives from
// the distance every step (src/coredynamics.cu:2173-2186):
// time2post = d / speed, it2post = ceil(time2post / dt), rem = it2post*dt -
// time2post; packed as rem's bits | it2post << 32.
__global__ __launch_bounds__(256) void k_rec_delay_table(uint64_t n, const float *__restrict__ dist,
                                                         float speed, float dt,
                                                         uint64_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float time2post = dist[i] / speed;
    const uint32_t it2post = (uint32_t)ceilf(time2post / dt);
    time2post = (float)it2post * dt - time2post;
    out[i] = ((uint64_t)it2post << 32) | __float_as_uint(time2post);
}

// One post neuron per thread, workgroups (up to 256 threads) inside one V1
// block, so every wavefront of a workgroup has the same neighbour tiles.
//   1. the spike masks of all `depth` ring slots of every neighbour tile are
//      OR-ed into su[] (the window's spikers);
//   2. the window's spikers are listed in visiting order (neighbour block,
//      ascending presynaptic id) in items[] (a scan of su[]'s popcounts);
//   3. each wavefront walks the list 64 spikers at a time: lane k loads
//      spiker k's row-segment header and, for rings of up to kMaxDepthLds
//      slots, its `depth` ring values into LDS together with a bit mask of
//      the slots holding a spike (one round trip); then the strengths and
//      delays of kGD spikers at a time are loaded (one round trip) and
//      accumulated in order, a lane reading ring values only when one of
//      its two slots holds a spike.
template <int kGD_>
__global__ __launch_bounds__(256) void k_rec_delay(vc_v1_params p, vc_v1_state s,
                                                   vc_rec_conn rec,
                                                   const uint64_t *__restrict__ delays,
                                                   const float *__restrict__ train,
                                                   const uint64_t *__restrict__ tmask,
                                                   uint32_t depth, uint32_t cur,
                                                   uint32_t *__restrict__ status) {
    __shared__ uint64_t su[kMaxEntD];
    __shared__ uint32_t soff[kMaxEntD + 1];
    __shared__ uint32_t sbid[kMaxEntD];
    __shared__ uint16_t items[kMaxEntD * 64];
    __shared__ float stv[4][64 * kStv];
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
    for (uint32_t q = threadIdx.x; q < nEnt * depth; q += blockDim.x) {
        const uint32_t k = q / nEnt, e = q - k * nEnt;
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
    const uint32_t total = soff[nEnt];

    float ipE[VC_MAX_NG], ipI[VC_MAX_NG];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        ipE[ig] = ig < p.ngE ? p.pE[t * p.ngE + ig] : 0.0f;
        ipI[ig] = ig < p.ngI ? p.pI[t * p.ngI + ig] : 0.0f;
    }
    float lgE[VC_MAX_NG] = {0.f, 0.f}, lhE[VC_MAX_NG] = {0.f, 0.f};
    float lgI[VC_MAX_NG] = {0.f, 0.f}, lhI[VC_MAX_NG] = {0.f, 0.f};
    bool late = false;
    const bool staged = depth <= (uint32_t)kMaxDepthLds;
    const uint2 *delays2 = reinterpret_cast<const uint2 *>(delays);   // {rem bits, it2post}
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
            if (staged) {
                am = 0u;
                for (uint32_t k0 = 0; k0 < depth; k0 += 8u) {
                    float x[8];
#pragma unroll
                    for (uint32_t i = 0; i < 8u; ++i)
                        x[i] = k0 + i < depth ? train[(size_t)(k0 + i) * nV1 + gpre] : 0.0f;
#pragma unroll
                    for (uint32_t i = 0; i < 8u; ++i) {
                        const uint32_t k = k0 + i;
                        if (k < depth) {
                            sv[lane * kStv + k] = x[i];
                            const uint32_t age = cur >= k ? cur - k : cur + depth - k;
                            am |= (x[i] >= 1.0f ? 1u : 0u) << age;
                        }
                    }
                }
            }
        }
        __syncthreads();
        for (uint32_t g0 = 0; g0 < nIt; g0 += kGD_) {
            uint64_t mine = 0ull;                      // bit g: lane has a synapse from spiker g
            float wv_[kGD_], rem[kGD_];
            uint32_t i2p[kGD_];
#pragma clang loop unroll(full)
            for (int g = 0; g < kGD_; ++g) {
                wv_[g] = rem[g] = 0.0f;
                i2p[g] = 0u;
                const uint32_t kk = g0 + g;
                if (kk < nIt) {
                    const uint64_t m = vc_readlane_u64(hm, kk);
                    if ((m >> lane) & 1ull) {
                        const uint32_t idx = vc_readlane_u32(ho, kk) + vc_mbcnt(m);
                        wv_[g] = rec.w[idx];
                        const uint2 dl = delays2[idx];
                        rem[g] = __uint_as_float(dl.x);
                        i2p[g] = dl.y;
                        mine |= 1ull << g;
                    }
                }
            }
#pragma clang loop unroll(full)
            for (int g = 0; g < kGD_; ++g) {
                if (!((mine >> g) & 1ull)) continue;
                const uint32_t kk = g0 + g;
                // src/coredynamics.cu:2173-2201 (it2post and the remainder precomputed)
                const uint32_t it2post = i2p[g];
                const float time2post = rem[g];
                if (it2post >= depth) {               // ring too short for this delay
                    late = true;
                    continue;
                }
                if (staged) {   // slot j = 0 is it2post steps old, j = 1 one step newer
                    const uint32_t a = vc_readlane_u32(am, kk);
                    const uint32_t need = it2post ? (a >> (it2post - 1u)) & 3u
                                                  : (a & 1u) | ((a >> (depth - 1u)) & 1u);
                    if (!need) continue;                  // neither slot holds a spike
                }
                const uint32_t j0 = cur >= it2post ? cur - it2post : cur + depth - it2post;
                const uint32_t j1 = j0 + 1u == depth ? 0u : j0 + 1u;
                float tv0, tv1;
                if (staged) {
                    tv0 = sv[kk * kStv + j0];
                    tv1 = sv[kk * kStv + j1];
                } else {
                    const uint32_t gp = vc_readlane_u32(gpre, kk);
                    tv0 = train[(size_t)j0 * nV1 + gp];
                    tv1 = train[(size_t)j1 * nV1 + gp];
                }
                const bool inh = vc_readlane_u32(lpre, kk) >= p.nE;
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    float dtsp = j ? tv1 : tv0;
                    if (dtsp >= 1.0f) {                // src/coredynamics.cu:2213-2235
                        const float nsp = floorf(dtsp);
                        dtsp = (dtsp - nsp + (float)j) * dt - time2post;
                        if (dtsp < dt && dtsp >= 0.0f) {
                            const float f = wv_[g];
                            if (f > 0.0f) {
                                dtsp = dt - dtsp;
                                const float a = nsp * f;
                                if (!inh) {
#pragma unroll
                                    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig)
                                        if (ig < p.ngE)
                                            single_input(lgE[ig], lhE[ig], a * ipE[ig], dtsp,
                                                         p.riseE[ig], p.decayE[ig], p.dodE[ig]);
                                } else {
#pragma unroll
                                    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig)
                                        if (ig < p.ngI)
                                            single_input(lgI[ig], lhI[ig], a * ipI[ig], dtsp,
                                                         p.riseI[ig], p.decayI[ig], p.dodI[ig]);
                                }
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

extern "C" int vc_v1_rec_delay_table(void *stream, uint64_t n, const float *dist, float speed,
                                     float dt, uint64_t *out_delays) {
    if (n == 0) return VC_OK;
    if (!dist || !out_delays || !(speed > 0.0f) || !(dt > 0.0f)) {
        vc_set_error("vc_v1_rec_delay_table: need dist, out_delays, speed > 0 and dt > 0");
        return VC_E_INVALID;
    }
    const uint64_t grid = (n + 255) / 256;
    if (grid > 0x7fffffffull) {
        vc_set_error("vc_v1_rec_delay_table: %llu synapses is too many for one launch",
                     (unsigned long long)n);
        return VC_E_INVALID;
    }
    hipLaunchKernelGGL(k_rec_delay_table, dim3((uint32_t)grid), dim3(256), 0,
                       (hipStream_t)stream, n, dist, speed, dt, out_delays);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_v1_rec_delay(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                               const vc_rec_conn *rec, const uint64_t *delays,
                               const float *train, const uint64_t *trainMask, uint32_t depth,
                               uint32_t cur, uint32_t *status) {
    if (!rec_ok(p, rec, "vc_v1_rec_delay")) return VC_E_INVALID;
    if (!s || !s->gE || !s->hE || !s->gI || !s->hI || !delays || !train || !trainMask) {
        vc_set_error("vc_v1_rec_delay: null argument");
        return VC_E_INVALID;
    }
    if (depth == 0 || cur >= depth) {
        vc_set_error("vc_v1_rec_delay: need cur (%u) < depth (%u)", cur, depth);
        return VC_E_INVALID;
    }
    if (p->ngE > VC_MAX_NG || p->ngI > VC_MAX_NG || p->nType > VC_MAX_NTYPE) {
        vc_set_error("vc_v1_rec_delay: receptor or type count out of range");
        return VC_E_INVALID;
    }
    const uint32_t wg = wg_size(p->blockSize);
    hipLaunchKernelGGL(k_rec_delay<kGD>, dim3(p->nV1 / wg), dim3(wg), 0, (hipStream_t)stream, *p,
                       *s, *rec, delays, train, trainMask, depth, cur, status);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}
