// vc_v1_device.h -- pieces of the V1 step shared by the forward kernel
// (vc_v1.hip) and its adjoint (vc_v1_adjoint.hip).
#pragma once
#include "vc_common.h"

namespace {

// per-launch constants derived on the host from vc_v1_params
struct V1Const {
    float etrFF[VC_MAX_NG], etdFF[VC_MAX_NG];
    float etrE[VC_MAX_NG], etdE[VC_MAX_NG];
    float etrI[VC_MAX_NG], etdI[VC_MAX_NG];
    uint32_t nSeg;                 // blockSize / 64
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
