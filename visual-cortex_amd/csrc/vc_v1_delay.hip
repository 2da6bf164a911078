// vc_v1_delay.hip -- recurrent input with axonal delays on gfx950: the
// reference's recal_G_mat (src/coredynamics.cu:2062-2309, used when noDelay
// is 0), which delivers every spike distance / speedOfThought after it was
// fired, from a ring of the last trainDepth steps' outputs.
//
// The reference scans, for every post neuron, every presynaptic neuron of its
// neighbour blocks and reads two ring slots of each connected one.  Here a
// workgroup (one V1 block, or part of one) first ORs the spike masks of all
// trainDepth slots of every neighbour tile into LDS, so only presynaptic
// neurons that fired inside the window are visited.  For each of those the
// wavefront (64 post neurons) reads the masked-CSR row-segment header once,
// each connected lane its strength and distance, then the two slots its delay
// selects.  Visiting order per post neuron is the reference's: neighbour
// blocks, ascending presynaptic id, slot j = 0 then 1, receptor; the sums are
// added to gE/hE/gI/hI at the end, as the reference does.
#include "vc_v1_device.h"

namespace {

constexpr int kMaxEntD = 512;      // neighbour tiles of one block (maxNear * blockSize/64)
constexpr int kGD = 8;             // presynaptic neurons per group of loads in flight

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

// One post neuron per thread, workgroups inside one V1 block.
__global__ __launch_bounds__(256) void k_rec_delay(vc_v1_params p, vc_v1_state s,
                                                   vc_rec_conn rec,
                                                   const float *__restrict__ dist, float speed,
                                                   const float *__restrict__ train,
                                                   const uint64_t *__restrict__ tmask,
                                                   uint32_t depth, uint32_t cur,
                                                   uint32_t *__restrict__ status) {
    __shared__ uint64_t su[kMaxEntD];
    const uint32_t nV1 = p.nV1, bs = p.blockSize, nSeg = bs >> 6;
    const uint32_t nTiles = nV1 >> 6;
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t b = n / bs;
    const uint32_t local = n - b * bs;
    const uint32_t lane = vc_lane();
    const uint32_t myseg = local >> 6;
    const int t = type_of(p, local);
    const float dt = p.dt;
    const uint32_t nb = rec.nNear[b];
    const uint32_t nEnt = nb * nSeg;
    // spike masks of the window, OR-ed over the ring, per neighbour tile
    for (uint32_t e = threadIdx.x; e < nEnt; e += blockDim.x) su[e] = 0ull;
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < nEnt * depth; q += blockDim.x) {
        const uint32_t k = q / nEnt, e = q - k * nEnt;
        const uint32_t ib = e / nSeg, tt = e - ib * nSeg;
        const uint32_t bid = rec.nbr[b * rec.maxNear + ib];
        const uint64_t m = tmask[(size_t)k * nTiles + bid * nSeg + tt];
        if (m) atomicOr((unsigned long long *)&su[e], (unsigned long long)m);
    }
    __syncthreads();

    float ipE[VC_MAX_NG], ipI[VC_MAX_NG];
#pragma unroll
    for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
        ipE[ig] = ig < p.ngE ? p.pE[t * p.ngE + ig] : 0.0f;
        ipI[ig] = ig < p.ngI ? p.pI[t * p.ngI + ig] : 0.0f;
    }
    float lgE[VC_MAX_NG] = {0.f, 0.f}, lhE[VC_MAX_NG] = {0.f, 0.f};
    float lgI[VC_MAX_NG] = {0.f, 0.f}, lhI[VC_MAX_NG] = {0.f, 0.f};
    bool late = false;

    for (uint32_t e = 0; e < nEnt; ++e) {
        uint64_t u = vc_readlane_u64(su[e], 0);      // wave-uniform
        if (!u) continue;
        const uint32_t ib = e / nSeg, tt = e - ib * nSeg;
        const uint32_t bid = rec.nbr[b * rec.maxNear + ib];
        const size_t rowBase = ((size_t)b * rec.maxNear + ib) * bs + tt * 64u;
        // row-segment headers of the tile's window spikes, lane l <- pre tt*64+l
        uint64_t hm = 0ull;
        uint32_t ho = 0u;
        if ((u >> lane) & 1ull) {
            const size_t rs = (rowBase + lane) * nSeg + myseg;
            hm = rec.rowMask[rs];
            ho = rec.rowOff[rs];
        }
        const float *trow = train + (size_t)bid * bs + tt * 64u;
        while (u) {
            uint32_t pl[kGD];
            bool mine[kGD];
            float wv[kGD], dv[kGD];
#pragma unroll
            for (int g = 0; g < kGD; ++g) {
                pl[g] = 64u;
                mine[g] = false;
                wv[g] = dv[g] = 0.0f;
                if (u) {
                    pl[g] = (uint32_t)__builtin_ctzll(u);
                    u &= u - 1ull;
                    const uint64_t m = vc_readlane_u64(hm, pl[g]);
                    if ((m >> lane) & 1ull) {
                        const uint32_t idx = vc_readlane_u32(ho, pl[g]) + vc_mbcnt(m);
                        wv[g] = rec.w[idx];
                        dv[g] = dist[idx];
                        mine[g] = true;
                    }
                }
            }
            float tv[kGD][2], rem[kGD];
#pragma unroll
            for (int g = 0; g < kGD; ++g) {
                tv[g][0] = tv[g][1] = 0.0f;
                rem[g] = 0.0f;
                if (mine[g]) {
                    // src/coredynamics.cu:2173-2201
                    float time2post = dv[g] / speed;
                    const uint32_t it2post = (uint32_t)ceilf(time2post / dt);
                    time2post = (float)it2post * dt - time2post;
                    rem[g] = time2post;
                    if (it2post >= depth) {           // ring too short for this delay
                        late = true;
                        mine[g] = false;
                    } else {
                        const uint32_t j0 = cur >= it2post ? cur - it2post : cur + depth - it2post;
                        const uint32_t j1 = j0 + 1u == depth ? 0u : j0 + 1u;
                        tv[g][0] = trow[(size_t)j0 * nV1 + pl[g]];
                        tv[g][1] = trow[(size_t)j1 * nV1 + pl[g]];
                    }
                }
            }
#pragma unroll
            for (int g = 0; g < kGD; ++g) {
                if (!mine[g]) continue;
                const bool inh = tt * 64u + pl[g] >= p.nE;
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    float dtsp = tv[g][j];
                    if (dtsp >= 1.0f) {                // src/coredynamics.cu:2213-2235
                        const float nsp = floorf(dtsp);
                        dtsp = (dtsp - nsp + (float)j) * dt - rem[g];
                        if (dtsp < dt && dtsp >= 0.0f) {
                            const float f = wv[g];
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

extern "C" int vc_v1_rec_delay(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                               const vc_rec_conn *rec, const float *dist, float speed,
                               const float *train, const uint64_t *trainMask, uint32_t depth,
                               uint32_t cur, uint32_t *status) {
    if (!rec_ok(p, rec, "vc_v1_rec_delay")) return VC_E_INVALID;
    if (!s || !s->gE || !s->hE || !s->gI || !s->hI || !dist || !train || !trainMask) {
        vc_set_error("vc_v1_rec_delay: null argument");
        return VC_E_INVALID;
    }
    if (!(speed > 0.0f) || depth == 0 || cur >= depth) {
        vc_set_error("vc_v1_rec_delay: need speed > 0 and cur (%u) < depth (%u)", cur, depth);
        return VC_E_INVALID;
    }
    if (p->ngE > VC_MAX_NG || p->ngI > VC_MAX_NG || p->nType > VC_MAX_NTYPE) {
        vc_set_error("vc_v1_rec_delay: receptor or type count out of range");
        return VC_E_INVALID;
    }
    const uint32_t wg = wg_size(p->blockSize);
    hipLaunchKernelGGL(k_rec_delay, dim3(p->nV1 / wg), dim3(wg), 0, (hipStream_t)stream, *p, *s,
                       *rec, dist, speed, train, trainMask, depth, cur, status);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}
