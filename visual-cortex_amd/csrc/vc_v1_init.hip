// vc_v1_init.hip -- the initial V1 state of a run.
//
// vc_v1_initial_state: the reference's host-side draws of the initial
// membrane potentials, adaptation currents and conductances
// (src/patch.cu:4404-4570), with the generator the reference uses there
// (std::default_random_engine seeded once, std::normal_distribution<float>
// and std::uniform_real_distribution<float>), in its order, so that a run
// starts from the reference's own initial state.
//
// vc_v1_initial_spikes: rand_spInit (src/coredynamics.cu:5-146), the
// initial spikes drawn with a per-type chance per step, and the spike-train
// ring it fills; vc_spikes_compact turns one step's outputs into a spike
// buffer (the spikes the first step gathers).
#include "vc_common.h"
#include "vc_v1_device.h"

#include <functional>
#include <random>

namespace {

__device__ __forceinline__ int init_type(const vc_v1_params &p, uint32_t local) {
    for (uint32_t j = 0; j + 1 < p.nType; ++j)
        if (local < p.typeAcc[j]) return (int)j;
    return (int)p.nType - 1;
}

// One thread per neuron, 64-thread workgroups (one spike-mask tile each).
// Chain slot 0 is the reference's spikeTrain slot 0 (its "initial" spike,
// which also resets v/w/tBack), chain slots depthRef-1, ..., 1 follow it in
// the reference's loop order; chain slot j holds step j - depthRef of this
// run, i.e. ring slot (cur0 + depth - depthRef + j) % depth; ring slots of
// older steps (depth > depthRef) hold v.  Draws: counter (neuron, chain
// slot, VC_STREAM_V1_INIT, trial), word 0 the spike test, word 1 the time.
__global__ __launch_bounds__(64) void k_v1_sp_init(vc_v1_params p, vc_v1_spinit q,
                                                    const uint32_t *__restrict__ nLGN,
                                                    float *__restrict__ v, float *__restrict__ w,
                                                    float *__restrict__ tBack,
                                                    float *__restrict__ train,
                                                    uint64_t *__restrict__ trainMask,
                                                    uint32_t depth, uint32_t cur0) {
    const uint32_t nV1 = p.nV1;
    const uint32_t n = blockIdx.x * 64u + threadIdx.x;      // nV1 % 64 == 0: always < nV1
    const uint32_t b = n / p.blockSize;
    const uint32_t local = n - b * p.blockSize;
    const int t = init_type(p, local);
    uint32_t klo = p.seed_lo, khi = p.seed_hi, ent = n;
    if (p.blockTrial != nullptr) {
        const uint32_t tr = p.blockTrial[b];
        klo = p.trialSeed[2 * tr];
        khi = p.trialSeed[2 * tr + 1];
        ent = n - p.trialV1Off[tr];
    }
    const float dt = p.dt, tRef = p.tRef[t];
    const uint32_t iLGN = nLGN ? nLGN[n] : 0u;
    const float chance = q.sp0[t * 2 + (iLGN > q.SCsplit ? 0 : 1)];
    const uint32_t D = q.depthRef;
    auto slot_of = [&](uint32_t j) { return (cur0 + depth - D + j) % depth; };
    auto put = [&](uint32_t j, float val) {
        const uint32_t sl = slot_of(j);
        train[(size_t)sl * nV1 + n] = val;
        const uint64_t m = __ballot(val >= 1.0f);
        if (trainMask != nullptr && threadIdx.x == 0) trainMask[(size_t)sl * (nV1 / 64u) + n / 64u] = m;
    };
    const vc_u32x4 r0 = vc_draw4(klo, khi, ent, 0u, VC_STREAM_V1_INIT, 0u, p.trial);
    float vv = v[n];
    float sInfo = vv;
    float ref = 0.0f;
    if (chance > 0.0f && vc_u01_open0(r0.v[0]) < chance) {
        const float tsp = vc_u01_open0(r0.v[1]);
        sInfo = 1.0f + tsp;
        const float tb = tRef - (1.0f - tsp) * dt;
        if (tb > 0.0f) tBack[n] = tb;
        const float v0 = p.vR[t];
        v[n] = v0;
        vv = v0;
        if (p.iModel == 1) {
            const float A = p.a[t] * (v0 - VC_VL) * p.tau_w[t];
            const float w0 = w[n] + p.b[t];
            w[n] = (w0 - A) * vc_expf(-dt * (1.0f - tsp) / p.tau_w[t]) + A;
        }
        ref = (1.0f - tsp) * dt + tRef - dt;
    }
    put(0u, sInfo);
    for (uint32_t i = D - 1u; i > 0u; --i) {
        float val;
        if (ref < dt) {
            if (ref < 0.0f) ref = 0.0f;
            const vc_u32x4 r = vc_draw4(klo, khi, ent, i, VC_STREAM_V1_INIT, 0u, p.trial);
            if (vc_u01_open0(r.v[0]) < chance * (dt - ref) / dt) {
                const float tsp = vc_u01_open0(r.v[1]) * (dt - ref) / dt;
                val = 1.0f + tsp;
                ref = tRef + (1.0f - tsp) * dt;
            } else {
                val = vv;
            }
        } else {
            val = p.vR[t];
        }
        ref -= dt;
        put(i, val);
    }
    for (uint32_t k = D; k < depth; ++k) {                 // steps before -depthRef
        const uint32_t sl = (cur0 + depth - 1u - k) % depth;
        train[(size_t)sl * nV1 + n] = vv;
        if (trainMask != nullptr && threadIdx.x == 0) trainMask[(size_t)sl * (nV1 / 64u) + n / 64u] = 0ull;
    }
}

// Ballot compaction of one step's outputs into a spike buffer: the tile's
// spike mask and each spike's unit increments at slot tile*64 + rank, as
// the step kernels write them (src/coredynamics.cu:1994-2046, 1205-1230).
__global__ __launch_bounds__(64) void k_spikes_compact(vc_v1_params p,
                                                        const float *__restrict__ sInfo,
                                                        vc_spikes sp) {
    const uint32_t n = blockIdx.x * 64u + threadIdx.x;
    const uint32_t nV1 = p.nV1;
    const uint32_t local = n % p.blockSize;
    const float s = sInfo[n];
    const bool exist = s >= 1.0f;
    const uint64_t m = __ballot(exist);
    const uint32_t tile = n >> 6;
    if (threadIdx.x == 0) sp.mask[tile] = m;
    if (exist) {
        const float dt = p.dt;
        const float nsp = floorf(s);
        const float tsp = (s - nsp) * dt;
        const uint32_t slot = tile * 64u + vc_mbcnt(m);
        const bool isE = local < p.nE;
        const uint32_t ng = isE ? p.ngE : p.ngI;
#pragma unroll
        for (uint32_t ig = 0; ig < VC_MAX_NG; ++ig) {
            if (ig < ng) {
                float g = 0.0f, h = 0.0f;
                if (isE)
                    single_input(g, h, nsp, dt - tsp, p.riseE[ig], p.decayE[ig], p.dodE[ig]);
                else
                    single_input(g, h, nsp, dt - tsp, p.riseI[ig], p.decayI[ig], p.dodI[ig]);
                sp.og[ig * nV1 + slot] = g;
                sp.oh[ig * nV1 + slot] = h;
            }
        }
    }
}

int init_params_ok(const vc_v1_params *p, const char *who) {
    if (!p || p->nV1 == 0 || p->nV1 % 64u != 0 || p->blockSize == 0 || p->blockSize % 64u != 0 ||
        p->nV1 != p->nblock * p->blockSize || p->nType < 1 || p->nType > VC_MAX_NTYPE ||
        p->ngE < 1 || p->ngE > VC_MAX_NG || p->ngI < 1 || p->ngI > VC_MAX_NG) {
        vc_set_error("%s: bad V1 parameters", who);
        return VC_E_INVALID;
    }
    return VC_OK;
}

}  // namespace

extern "C" int vc_v1_initial_spikes(void *stream, const vc_v1_params *p, const vc_v1_state *s,
                                    const uint32_t *nLGN, const vc_v1_spinit *q, float *train,
                                    uint64_t *trainMask, uint32_t depth, uint32_t cur0) {
    int rc = init_params_ok(p, "vc_v1_initial_spikes");
    if (rc) return rc;
    if (!s || !s->v || !s->w || !s->tBack || !q || !train || q->depthRef == 0 ||
        depth < q->depthRef || cur0 >= depth) {
        vc_set_error("vc_v1_initial_spikes: bad argument (state, train, 1 <= depthRef <= depth, "
                     "cur0 < depth)");
        return VC_E_INVALID;
    }
    for (uint32_t i = 0; i < 2u * p->nType; ++i) {
        if (!(q->sp0[i] >= 0.0f && q->sp0[i] <= 1.0f)) {
            vc_set_error("vc_v1_initial_spikes: sp0[%u] = %g is not a chance per step in [0, 1]",
                         i, q->sp0[i]);
            return VC_E_INVALID;
        }
    }
    hipLaunchKernelGGL(k_v1_sp_init, dim3(p->nV1 / 64u), dim3(64), 0, (hipStream_t)stream, *p,
                       *q, nLGN, s->v, s->w, s->tBack, train, trainMask, depth, cur0);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_spikes_compact(void *stream, const vc_v1_params *p, const float *sInfo,
                                 const vc_spikes *sp) {
    int rc = init_params_ok(p, "vc_spikes_compact");
    if (rc) return rc;
    if (!sInfo || !sp || !sp->mask || !sp->og || !sp->oh) {
        vc_set_error("vc_spikes_compact: null argument");
        return VC_E_INVALID;
    }
    hipLaunchKernelGGL(k_spikes_compact, dim3(p->nV1 / 64u), dim3(64), 0, (hipStream_t)stream,
                       *p, sInfo, *sp);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

// ---------------------------------------------------------------- host --
// src/patch.cu:4405-4570, in the reference's order: one engine seeded with
// `seed`; one normal distribution object shared by every truncated-normal
// draw (the reference's static `norm` inside its draw lambda, so the second
// value of each polar pair carries over from one array to the next); per
// type: w (AdEx) for every block, then v; then per type and FF receptor g
// then h for every block, a neuron without LGN input getting g = h = 0; then
// per type the E receptors' g and h for every block, then the I ones'.
extern "C" int vc_v1_initial_state(uint64_t seed, const vc_v1_init *q, const float *tonicDep,
                                   const uint32_t *nLGN, float *v, float *w, float *gFF,
                                   float *hFF, float *gE, float *hE, float *gI, float *hI) {
    if (!q || !v || !gFF || !hFF || !gE || !hE || !gI || !hI || !tonicDep || !nLGN ||
        (q->iModel == 1 && !w)) {
        vc_set_error("vc_v1_initial_state: null argument");
        return VC_E_INVALID;
    }
    const uint32_t nV1 = q->nV1, bs = q->blockSize, nT = q->nType;
    if (nT < 1 || nT > VC_MAX_NTYPE || bs == 0 || nV1 != q->nblock * bs || q->ngFF < 1 ||
        q->ngFF > VC_MAX_NG || q->ngE < 1 || q->ngE > VC_MAX_NG || q->ngI < 1 ||
        q->ngI > VC_MAX_NG || q->typeAcc[nT - 1] != bs) {
        vc_set_error("vc_v1_initial_state: bad sizes");
        return VC_E_INVALID;
    }
    uint32_t acc0[VC_MAX_NTYPE + 1] = {0};
    for (uint32_t i = 0; i < nT; ++i) acc0[i + 1] = q->typeAcc[i];
    std::default_random_engine rGen(seed);
    std::normal_distribution<float> norm(0.0f, 1.0f);
    auto truncated = [&](float mean, float sd, bool nonNeg) {
        float r;
        do {
            r = norm(rGen) * sd + mean;
        } while (nonNeg && r < 0.0f);
        return r;
    };
    float tonicMax = tonicDep[0];
    for (uint32_t i = 1; i < nV1; ++i) tonicMax = tonicDep[i] > tonicMax ? tonicDep[i] : tonicMax;
    for (uint32_t i = 0; i < nT; ++i) {
        if (q->iModel == 1) {
            const float mean = q->w0[i * 2 + 0], sd = q->w0[i * 2 + 1];
            for (uint32_t b = 0; b < q->nblock; ++b)
                for (uint32_t j = acc0[i]; j < acc0[i + 1]; ++j)
                    w[b * bs + j] = truncated(mean, sd, false);
        }
        if (q->useV0 || tonicMax > 1.0f) {
            std::uniform_real_distribution<float> uni(q->vR[i], q->vT[i]);
            for (uint32_t b = 0; b < q->nblock; ++b)
                for (uint32_t j = acc0[i]; j < acc0[i + 1]; ++j) v[b * bs + j] = uni(rGen);
        } else {
            for (uint32_t b = 0; b < q->nblock; ++b)
                for (uint32_t j = acc0[i]; j < acc0[i + 1]; ++j)
                    v[b * bs + j] = q->vR[i] + (q->vT[i] - q->vR[i]) * tonicDep[b * bs + j];
        }
    }
    const uint32_t ngF = q->ngFF;
    for (uint32_t i = 0; i < nT; ++i) {
        for (uint32_t k = 0; k < ngF; ++k) {
            const float *d = q->gFF0 + (i * ngF + k) * 4;
            for (uint32_t b = 0; b < q->nblock; ++b) {
                for (uint32_t j = acc0[i]; j < acc0[i + 1]; ++j)
                    gFF[(size_t)k * nV1 + b * bs + j] = truncated(d[0], d[0] * d[1], true);
                for (uint32_t j = acc0[i]; j < acc0[i + 1]; ++j)
                    hFF[(size_t)k * nV1 + b * bs + j] = truncated(d[2], d[2] * d[3], true);
                for (uint32_t j = acc0[i]; j < acc0[i + 1]; ++j) {
                    if (nLGN[b * bs + j] == 0) {
                        gFF[(size_t)k * nV1 + b * bs + j] = 0.0f;
                        hFF[(size_t)k * nV1 + b * bs + j] = 0.0f;
                    }
                }
            }
        }
    }
    for (uint32_t i = 0; i < nT; ++i) {
        for (uint32_t k = 0; k < q->ngE; ++k) {
            const float *d = q->gE0 + (i * q->ngE + k) * 4;
            for (uint32_t b = 0; b < q->nblock; ++b) {
                for (uint32_t j = acc0[i]; j < acc0[i + 1]; ++j)
                    gE[(size_t)k * nV1 + b * bs + j] = truncated(d[0], d[0] * d[1], true);
                for (uint32_t j = acc0[i]; j < acc0[i + 1]; ++j)
                    hE[(size_t)k * nV1 + b * bs + j] = truncated(d[2], d[2] * d[3], true);
            }
        }
        for (uint32_t k = 0; k < q->ngI; ++k) {
            const float *d = q->gI0 + (i * q->ngI + k) * 4;
            for (uint32_t b = 0; b < q->nblock; ++b) {
                for (uint32_t j = acc0[i]; j < acc0[i + 1]; ++j)
                    gI[(size_t)k * nV1 + b * bs + j] = truncated(d[0], d[0] * d[1], true);
                for (uint32_t j = acc0[i]; j < acc0[i + 1]; ++j)
                    hI[(size_t)k * nV1 + b * bs + j] = truncated(d[2], d[2] * d[3], true);
            }
        }
    }
    return VC_OK;
}
