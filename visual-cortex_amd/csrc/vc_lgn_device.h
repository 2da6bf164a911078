// vc_lgn_device.h -- device helpers of the LGN stage shared by the forward
// kernels (vc_lgn.hip) and their adjoints (vc_lgn_adjoint.hip).
#pragma once
#include "vc_common.h"
#include <math.h>

namespace {

constexpr double kSqrt2 = 1.4142135623730951;  // src/CONST.h:32 (double)
constexpr int kMaxTaps = 64;                    // temporal kernel samples per region

// What a region's sample texture coordinates need besides the sample's grid
// position (flat retina, src/discrete_input_convol.cu:398-417): ex = ecc
// cos(polar), ey = ecc sin(polar), the orientation's cos/sin and the grid's
// step and half span in each direction (VC_LGN_GEOM floats per region).
struct SampleGeom {
    float ex, ey, coso, sino, dw, wSpan, dh, hSpan;
};
static_assert(sizeof(SampleGeom) == VC_LGN_GEOM * sizeof(float), "geometry record");

__device__ __forceinline__ SampleGeom sample_geom(const vc_lgn_setup &q, uint32_t lid) {
    SampleGeom g;
    const uint32_t n1D = q.nSample1D;
    const float rx = q.rw[lid], ry = q.rh[lid];
    g.wSpan = (float)((double)(q.nsig * rx) / kSqrt2);
    g.dw = 2 * g.wSpan / (float)n1D;
    g.hSpan = (float)((double)(q.nsig * ry) / kSqrt2);
    g.dh = 2 * g.hSpan / (float)n1D;
    const float polar = q.polar[lid], ecc = q.ecc[lid];
    g.coso = cosf(q.orient[lid]);
    g.sino = sinf(q.orient[lid]);
    g.ex = ecc * cosf(polar);
    g.ey = ecc * sinf(polar);
    return g;
}

// Normalised texture coordinates of grid sample (tx, ty) of a region: the
// reference's expression, evaluated in the same order and precision.
__device__ __forceinline__ void sample_xy(const SampleGeom &g, float x0, float y0, float nvd,
                                          uint32_t tx, uint32_t ty, float &x, float &y) {
    const float w = (float)((tx + 0.5) * (double)g.dw - (double)g.wSpan);
    const float h = (float)((ty + 0.5) * (double)g.dh - (double)g.hSpan);
    x = x0 + (g.ex + w * g.coso - h * g.sino) * nvd;
    y = y0 + (g.ey + w * g.sino + h * g.coso) * nvd;
}

// block_reduce of src/util/cuda_util.cuh:89-101 for blockDim.x threads (a
// power of two >= 32).  Every thread gets the result.
__device__ float block_reduce_ref(float x, float *part /* >= 32 floats */) {
    const uint32_t tid = threadIdx.x;
    const uint32_t W = blockDim.x / 32u;
    x = vc_tree32(x);
    __syncthreads();
    if ((tid & 31u) == 0u) part[tid >> 5] = x;
    __syncthreads();
    if (tid == 0) {
        for (uint32_t off = W / 2; off > 0; off >>= 1)
            for (uint32_t l = 0; l < off; ++l) part[l] = part[l] + part[l + off];
    }
    __syncthreads();
    float r = part[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ float texel(const float *img, int W, int H, int i, int j) {
    if (i < 0 || j < 0 || i >= W || j >= H) return 0.0f;
    return img[(size_t)j * W + i];
}

// exact-float bilinear with border addressing and normalised coordinates
__device__ __forceinline__ float bilinear(const float *img, int W, int H, float x, float y) {
    float xt = x * (float)W - 0.5f;
    float yt = y * (float)H - 0.5f;
    if (!(xt > -2.0f && xt < (float)W + 1.0f && yt > -2.0f && yt < (float)H + 1.0f))
        return 0.0f;
    float fx = floorf(xt), fy = floorf(yt);
    float ax = xt - fx, ay = yt - fy;
    int i0 = (int)fx, j0 = (int)fy;
    float t00 = texel(img, W, H, i0, j0);
    float t10 = texel(img, W, H, i0 + 1, j0);
    float t01 = texel(img, W, H, i0, j0 + 1);
    float t11 = texel(img, W, H, i0 + 1, j0 + 1);
    float bx = 1.0f - ax, by = 1.0f - ay;
    return by * (bx * t00 + ax * t10) + ay * (bx * t01 + ax * t11);
}

// get_intensity, src/discrete_input_convol.cu:35-76
__device__ float intensity(uint32_t cone, const float *frame, int W, int H, float x, float y) {
    const size_t np = (size_t)W * H;
    switch (cone) {
    case 0: return bilinear(frame, W, H, x, y);
    case 1: return bilinear(frame + np, W, H, x, y);
    case 2: return bilinear(frame + 2 * np, W, H, x, y);
    case 3: return (bilinear(frame, W, H, x, y) + bilinear(frame + np, W, H, x, y)) / 2.0f;
    case 4: return bilinear(frame, W, H, x, y) * 0.45436904f +
                   bilinear(frame + np, W, H, x, y) * 0.47353315f +
                   bilinear(frame + 2 * np, W, H, x, y) * 0.0720978f;
    default: return 0.0f;
    }
}

__device__ __forceinline__ float clip_contrast(float I, float mean) {
    float c = (mean > 0.0f) ? I / mean - 1.0f : I;
    if (fabsf(c) > 1.0f) c = copysignf(1.0f, c);
    return c;
}

__device__ __forceinline__ int64_t floor_div(int64_t a, int64_t b) {
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
    return q;
}

__device__ __forceinline__ float transform(float C50, float K, float A, float B, float input) {
    float X = (float)(1.0 / (double)(1.0f + vc_expf(K * (C50 - input))));
    return A * X + B;
}

// src/discrete_input_convol.cu:14-33
__device__ float temporal_kernel(float tau, float tauR, float tauD, float ratio, float nR,
                                 float nD, float lfac1, float lfac2) {
    float tau1 = tau / tauR;
    float tau2 = tau / tauD;
    float A1 = (nR - 1) * logf(tau1);
    float A2 = (nD - 1) * logf(tau2);
    float exponent = A1 - tau1 - lfac1;
    float tpR = expf(exponent) / tauR;
    exponent = A2 - tau2 - lfac2;
    float tpD = expf(exponent) / tauD;
    return tpR - ratio * tpD;
}

__device__ __forceinline__ float spatial_kernel(float x, float y, float rx, float ry) {
    return expf(-x * x / (rx * rx) - y * y / (ry * ry));
}

}  // namespace
