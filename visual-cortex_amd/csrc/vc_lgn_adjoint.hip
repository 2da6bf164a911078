// vc_lgn_adjoint.hip -- reverse mode of the LGN stage (stage a and the rate
// of stage b) on gfx950.
//
// Forward being differentiated (vc_lgn.hip):
//   R_r[F]   = sum_s SW_s * clip(I_r(x_s) / mean_r - 1, +-1)      per frame F
//   conv_r   = sum_i TW_r[i] * R_r[F(it, i)]                        per step it
//   fr       = frRatio * max * (A * sigma(K (x - C50)) + B),  x = max(0, convolRatio *
//              (kC conv_C + kS conv_S)) / max
// with F(it, i) = floor((it + iT0 - nRetrace + 1 + i * interval) / ntPerFrame)
// (src/discrete_input_convol.cu:1143-1340, 1587-1622; LGN_props.cuh:94-100).
// max (maxConvol) is a set-up constant and is not differentiated.
//
//  k_lgn_rate_adj   one workgroup per cell and window of steps.  The step
//      gradient g(it) = dL/dconvol is formed in LDS, then summed over the
//      runs of steps that share a (tap, frame) pair: the frame of tap i
//      changes every ntPerFrame steps at offsets that are multiples of
//      L = gcd(interval, ntPerFrame), so segment sums of length L give every
//      G[i][F] = sum_{it : F(it,i) = F} g(it) with ntPerFrame/L additions.
//      Then dR_r[F] = k_r sum_i TW_r[i] G[i][F], dTW_r[i] = k_r sum_F R_r[F]
//      G[i][F], dk_r = sum_it g(it) conv_r(it).
//  k_lgn_frame_adj  64 consecutive cells x 4 frames per workgroup, one
//      wavefront per cell: recomputes the samples and pushes dR through the
//      clipped contrast, the surround mean and the bilinear taps into an LDS
//      tile over the group's texel box, then into dFrames (float atomics).
//  k_lgn_tw_adj     one thread per (cell, region): dTW through the
//      normalised difference of gamma densities to (tauR, tauD).
#include "vc_lgn_device.h"

namespace {

constexpr int kAdjWG = 256;
constexpr int kAdjT = 2048;      // steps per window
constexpr int kAdjF = 40;        // frames per window

__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    return x;
}

__device__ __forceinline__ float block_sum(float x, float *part /* >= 16 */) {
    x = wave_sum(x);
    const uint32_t w = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63u) == 0u) part[w] = x;
    __syncthreads();
    float r = 0.0f;
    for (uint32_t k = 0; k < nw; ++k) r += part[k];
    __syncthreads();
    return r;
}

__device__ __forceinline__ uint32_t slot_of(int64_t F, uint32_t maxFrame) {
    int64_t s = F % (int64_t)maxFrame;
    if (s < 0) s += maxFrame;
    return (uint32_t)s;
}

__global__ __launch_bounds__(kAdjWG) void k_lgn_rate_adj(
    vc_lgn_params p, vc_lgn_state st, const float *__restrict__ R, uint32_t it0, uint32_t T,
    int64_t u0, int64_t Fmin, uint32_t nF, uint32_t L, int64_t jmin, uint32_t nSeg,
    const float *__restrict__ gFr, float *__restrict__ dR, float *__restrict__ dTW,
    float *__restrict__ dK) {
    __shared__ float sTW[2][kMaxTaps];
    __shared__ float sR[2][kAdjF];
    __shared__ float sg[kAdjT];
    __shared__ float sS[kAdjT + 1];
    __shared__ float sG[kMaxTaps][kAdjF + 1];
    __shared__ float part[16];
    const uint32_t c = blockIdx.x, tid = threadIdx.x, nKS = p.nKernelSample;
    const int64_t npf = p.ntPerFrame, itv = p.kernelSampleInterval;
    for (uint32_t j = tid; j < 2 * nKS; j += blockDim.x)
        sTW[j / nKS][j % nKS] = st.TW[(size_t)c * 2 * nKS + j];
    for (uint32_t j = tid; j < 2 * nF; j += blockDim.x) {
        const uint32_t r = j / nF, f = j % nF;
        sR[r][f] = R[((size_t)slot_of(Fmin + f, p.maxFrame) * p.nLGN + c) * 2 + r];
    }
    __syncthreads();
    const float kC = st.kC[c], kS = st.kS[c], mx = st.maxConvol[c];
    const float C50 = st.c50[c], K = st.sharpness[c], A = st.logA[c];
    const float dfr = p.frRatio * A * K * p.convolRatio;
    float dkC = 0.0f, dkS = 0.0f;
    const uint32_t dq = (uint32_t)(itv / npf), dr = (uint32_t)(itv % npf);
    for (uint32_t s = tid; s < T; s += blockDim.x) {
        const int64_t u = u0 + s;
        // frame of tap i, floor((u + i itv) / npf) - Fmin, walked incrementally
        const int64_t q0 = floor_div(u, npf);
        uint32_t rr = (uint32_t)(u - q0 * npf);
        uint32_t f = (uint32_t)(q0 - Fmin);
        float convC = 0.0f, convS = 0.0f;
        for (uint32_t i = 0; i < nKS; ++i) {
            convC += sR[0][f] * sTW[0][i];
            convS += sR[1][f] * sTW[1][i];
            rr += dr;
            f += dq;
            if (rr >= (uint32_t)npf) {
                rr -= (uint32_t)npf;
                f += 1;
            }
        }
        const float cur = (convC * kC + convS * kS) * p.convolRatio;
        float g = 0.0f;
        if (cur > 0.0f) {
            const float X = 1.0f / (1.0f + vc_expf(K * (C50 - cur / mx)));
            g = gFr[(size_t)s * p.nLGN + c] * dfr * X * (1.0f - X);
        }
        sg[s] = g;
        dkC += g * convC;
        dkS += g * convS;
    }
    __syncthreads();
    // segment sums: segment j holds the steps with u in [jL, (j+1)L)
    for (uint32_t js = tid; js < nSeg; js += blockDim.x) {
        const int64_t ulo = (jmin + js) * (int64_t)L;
        int64_t a = ulo - u0, b = ulo + L - u0;
        if (a < 0) a = 0;
        if (b > (int64_t)T) b = T;
        float sum = 0.0f;
        for (int64_t s = a; s < b; ++s) sum += sg[s];
        sS[js] = sum;
    }
    __syncthreads();
    // G[i][F]: tap i sees frame F for u in [F npf - i itv, (F+1) npf - i itv)
    const uint32_t per = (uint32_t)(npf / L);
    for (uint32_t e = tid; e < nKS * nF; e += blockDim.x) {
        const uint32_t i = e / nF, f = e % nF;
        const int64_t jlo = ((Fmin + f) * npf - (int64_t)i * itv) / (int64_t)L - jmin;
        float sum = 0.0f;
        for (uint32_t k = 0; k < per; ++k) {
            const int64_t js = jlo + k;
            if (js >= 0 && js < (int64_t)nSeg) sum += sS[js];
        }
        sG[i][f] = sum;
    }
    __syncthreads();
    for (uint32_t e = tid; e < 2 * nF; e += blockDim.x) {
        const uint32_t r = e / nF, f = e % nF;
        float acc = 0.0f;
        for (uint32_t i = 0; i < nKS; ++i) acc += sTW[r][i] * sG[i][f];
        float *o = dR + ((size_t)slot_of(Fmin + f, p.maxFrame) * p.nLGN + c) * 2 + r;
        *o += (r == 0 ? kC : kS) * acc;
    }
    for (uint32_t e = tid; e < 2 * nKS; e += blockDim.x) {
        const uint32_t r = e / nKS, i = e % nKS;
        float acc = 0.0f;
        for (uint32_t f = 0; f < nF; ++f) acc += sR[r][f] * sG[i][f];
        dTW[((size_t)c * 2 + r) * nKS + i] += (r == 0 ? kC : kS) * acc;
    }
    dkC = block_sum(dkC, part);
    dkS = block_sum(dkS, part);
    if (tid == 0) {
        dK[(size_t)c * 2 + 0] += dkC;
        dK[(size_t)c * 2 + 1] += dkS;
    }
}

// d clip(I/mean - 1) with respect to I and to mean (zero where clipped)
__device__ __forceinline__ void contrast_adj(float I, float mean, float &dI, float &dmean) {
    if (mean > 0.0f) {
        const float c = I / mean - 1.0f;
        const bool pass = !(fabsf(c) > 1.0f);
        dI = pass ? 1.0f / mean : 0.0f;
        dmean = pass ? -I / (mean * mean) : 0.0f;
    } else {
        dI = (fabsf(I) > 1.0f) ? 0.0f : 1.0f;
        dmean = 0.0f;
    }
}

// bilinear footprint of one sample (the addressing of bilinear())
struct Tap {
    int i0, j0;
    float ax, ay;
    bool ok;
};
__device__ __forceinline__ Tap tap_of(int W, int H, float x, float y) {
    Tap t;
    const float xt = x * (float)W - 0.5f;
    const float yt = y * (float)H - 0.5f;
    t.ok = xt > -2.0f && xt < (float)W + 1.0f && yt > -2.0f && yt < (float)H + 1.0f;
    const float fx = floorf(t.ok ? xt : 0.0f), fy = floorf(t.ok ? yt : 0.0f);
    t.ax = (t.ok ? xt : 0.0f) - fx;
    t.ay = (t.ok ? yt : 0.0f) - fy;
    t.i0 = (int)fx;
    t.j0 = (int)fy;
    return t;
}

__device__ __forceinline__ void cone_weights(uint32_t cone, float w[3]) {
    w[0] = w[1] = w[2] = 0.0f;
    switch (cone) {
    case 0: w[0] = 1.0f; break;
    case 1: w[1] = 1.0f; break;
    case 2: w[2] = 1.0f; break;
    case 3: w[0] = 0.5f; w[1] = 0.5f; break;
    case 4: w[0] = 0.45436904f; w[1] = 0.47353315f; w[2] = 0.0720978f; break;
    default: break;
    }
}

constexpr int kGrpCells = 64;     // consecutive cells per workgroup (one wavefront per cell)
constexpr int kGrpFrames = 4;     // frames per workgroup
constexpr int kTile = 12288;      // LDS floats of the footprint tile (48 KB)
constexpr int kMaxSPL = 16;       // samples per lane (nS <= 1024)

struct FrameAdjCtx {
    float *tile;                  // [planes][bh][bw] in LDS, or null (global atomics)
    int imin, jmin, bw, area;
    uint32_t coneMask;            // cone types present (one tile plane each)
    float *dframe;                // [3][H][W]
    int W, H;
};

__device__ __forceinline__ void scatter_sample(const FrameAdjCtx &x, uint32_t cone, float px,
                                               float py, float d) {
    if (d == 0.0f || cone > 4u) return;
    const Tap t = tap_of(x.W, x.H, px, py);
    if (!t.ok) return;
    const float bx = 1.0f - t.ax, by = 1.0f - t.ay;
    const int ii[4] = {t.i0, t.i0 + 1, t.i0, t.i0 + 1};
    const int jj[4] = {t.j0, t.j0, t.j0 + 1, t.j0 + 1};
    const float ww[4] = {by * bx, by * t.ax, t.ay * bx, t.ay * t.ax};
    if (x.tile) {
        const int plane = __builtin_popcount(x.coneMask & ((1u << cone) - 1u));
        float *tp = x.tile + plane * x.area;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (ii[k] < 0 || jj[k] < 0 || ii[k] >= x.W || jj[k] >= x.H) continue;
            atomicAdd(tp + (jj[k] - x.jmin) * x.bw + (ii[k] - x.imin), ww[k] * d);
        }
    } else {
        float cw[3];
        cone_weights(cone, cw);
        const size_t np = (size_t)x.W * x.H;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (ii[k] < 0 || jj[k] < 0 || ii[k] >= x.W || jj[k] >= x.H) continue;
            const size_t o = (size_t)jj[k] * x.W + ii[k];
            for (int ch = 0; ch < 3; ++ch)
                if (cw[ch] != 0.0f) unsafeAtomicAdd(x.dframe + ch * np + o, cw[ch] * ww[k] * d);
        }
    }
}

// Frame adjoint: workgroup = kGrpCells consecutive cells x kGrpFrames frames.
// Each wavefront takes one cell at a time (lane = sample, nS/64 samples per
// lane); the cells' contributions are summed in an LDS tile over the texel
// box of the group's samples (one plane per cone type present), which is
// then added to dFrames with one atomic per texel and channel.  Groups whose
// box does not fit the tile add their samples to dFrames directly.
__global__ __launch_bounds__(256) void k_lgn_frame_adj(
    const float *__restrict__ frames, uint64_t frameStride, int W, int H, uint32_t nLGN,
    uint32_t nS, uint32_t nF, const float *__restrict__ SCx, const float *__restrict__ SCy,
    const float *__restrict__ SW, const uint32_t *__restrict__ coneType,
    const float *__restrict__ dR, float *__restrict__ dFrames) {
    __shared__ float tile[kTile];
    __shared__ int sbox[4];
    __shared__ uint32_t sMask;
    const uint32_t g0 = blockIdx.x * kGrpCells;
    const uint32_t nc = (nLGN - g0) < (uint32_t)kGrpCells ? (nLGN - g0) : (uint32_t)kGrpCells;
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t spl = (nS + 63u) / 64u;
    if (threadIdx.x == 0) {
        sbox[0] = sbox[1] = 0x7fffffff;
        sbox[2] = sbox[3] = -0x7fffffff;
        sMask = 0u;
    }
    __syncthreads();
    // texel box of every sample of the group and the cone types in use
    {
        int imin = 0x7fffffff, jmin = 0x7fffffff, imax = -0x7fffffff, jmax = -0x7fffffff;
        uint32_t cm = 0;
        for (uint32_t cc = wv; cc < nc; cc += nw) {
            const uint32_t c = g0 + cc;
            const uint32_t t0 = coneType[c], t1 = coneType[nLGN + c];
            if (t0 <= 4u) cm |= 1u << t0;
            if (t1 <= 4u) cm |= 1u << t1;
            for (uint32_t r = 0; r < 2; ++r)
                for (uint32_t k = 0; k < spl; ++k) {
                    const uint32_t s = lane + 64u * k;
                    if (s >= nS) break;
                    const size_t id = ((size_t)c * 2 + r) * nS + s;
                    const Tap t = tap_of(W, H, SCx[id], SCy[id]);
                    if (!t.ok) continue;
                    imin = min(imin, t.i0);
                    jmin = min(jmin, t.j0);
                    imax = max(imax, t.i0 + 1);
                    jmax = max(jmax, t.j0 + 1);
                }
        }
        atomicMin(&sbox[0], imin);
        atomicMin(&sbox[1], jmin);
        atomicMax(&sbox[2], imax);
        atomicMax(&sbox[3], jmax);
        atomicOr(&sMask, cm);
    }
    __syncthreads();
    const int imin = max(sbox[0], 0), jmin = max(sbox[1], 0);
    const int imax = min(sbox[2], W - 1), jmax = min(sbox[3], H - 1);
    if (imin > imax || jmin > jmax) return;                  // no sample reads the frame
    FrameAdjCtx x;
    x.imin = imin;
    x.jmin = jmin;
    x.bw = imax - imin + 1;
    x.area = x.bw * (jmax - jmin + 1);
    x.coneMask = sMask;
    x.W = W;
    x.H = H;
    const int nPl = __builtin_popcount(x.coneMask);
    x.tile = ((int64_t)x.area * nPl <= kTile) ? tile : nullptr;
    const float inv_nS = 1.0f / (float)nS;
    for (uint32_t q = 0; q < (uint32_t)kGrpFrames; ++q) {
        const uint32_t f = blockIdx.y * kGrpFrames + q;
        if (f >= nF) break;
        const float *frame = frames + (size_t)f * frameStride;
        x.dframe = dFrames + (size_t)f * frameStride;
        if (x.tile) {
            for (int e = threadIdx.x; e < x.area * nPl; e += blockDim.x) tile[e] = 0.0f;
            __syncthreads();
        }
        for (uint32_t cc = wv; cc < nc; cc += nw) {
            const uint32_t c = g0 + cc;
            const float dRC = dR[((size_t)f * nLGN + c) * 2 + 0];
            const float dRS = dR[((size_t)f * nLGN + c) * 2 + 1];
            if (dRC == 0.0f && dRS == 0.0f) continue;        // uniform over the wavefront
            const uint32_t tC = coneType[c], tS = coneType[nLGN + c];
            const float *xC = SCx + (size_t)c * 2 * nS, *yC = SCy + (size_t)c * 2 * nS;
            const float *xS = xC + nS, *yS = yC + nS;
            // surround intensities of both cone types (one evaluation when the
            // types agree), kept for the second pass
            float sumC = 0.0f, sumS = 0.0f;
            float icc[kMaxSPL], iss[kMaxSPL];
#pragma unroll
            for (int k = 0; k < kMaxSPL; ++k) {
                icc[k] = iss[k] = 0.0f;
                const uint32_t s = lane + 64u * k;
                if ((uint32_t)k < spl && s < nS) {
                    iss[k] = intensity(tS, frame, W, H, xS[s], yS[s]);
                    sumC += tC == tS ? iss[k] : intensity(tC, frame, W, H, xS[s], yS[s]);
                    sumS += iss[k];
                }
            }
            const float meanC = wave_sum(sumC) * inv_nS, meanS = wave_sum(sumS) * inv_nS;
            float tmC = 0.0f, tmS = 0.0f;
#pragma unroll
            for (int k = 0; k < kMaxSPL; ++k) {
                const uint32_t s = lane + 64u * k;
                if ((uint32_t)k < spl && s < nS) {
                    icc[k] = intensity(tC, frame, W, H, xC[s], yC[s]);
                    float dI, dm;
                    contrast_adj(icc[k], meanC, dI, dm);
                    tmC += dRC * SW[s] * dm;
                    contrast_adj(iss[k], meanS, dI, dm);
                    tmS += dRS * SW[s] * dm;
                }
            }
            const float dMeanC = wave_sum(tmC) * inv_nS, dMeanS = wave_sum(tmS) * inv_nS;
#pragma unroll
            for (int k = 0; k < kMaxSPL; ++k) {
                const uint32_t s = lane + 64u * k;
                if ((uint32_t)k < spl && s < nS) {
                    float dIC, dIS, dm;
                    contrast_adj(icc[k], meanC, dIC, dm);
                    contrast_adj(iss[k], meanS, dIS, dm);
                    const float sw = SW[s];
                    scatter_sample(x, tC, xC[s], yC[s], dRC * sw * dIC);
                    const float dSs = dRS * sw * dIS + dMeanS;
                    if (tC == tS) {
                        scatter_sample(x, tC, xS[s], yS[s], dMeanC + dSs);
                    } else {
                        scatter_sample(x, tC, xS[s], yS[s], dMeanC);
                        scatter_sample(x, tS, xS[s], yS[s], dSs);
                    }
                }
            }
        }
        if (x.tile) {
            __syncthreads();
            const size_t np = (size_t)W * H;
            for (int e = threadIdx.x; e < x.area * nPl; e += blockDim.x) {
                const float v = tile[e];
                if (v == 0.0f) continue;
                const int pl = e / x.area, o = e - pl * x.area;
                // cone type of plane pl: the pl-th set bit of the mask
                uint32_t m = x.coneMask;
                for (int k = 0; k < pl; ++k) m &= m - 1u;
                float cw[3];
                cone_weights((uint32_t)__builtin_ctz(m), cw);
                const int jy = o / x.bw, ix = o - jy * x.bw;
                const size_t po = (size_t)(jmin + jy) * W + (imin + ix);
                for (int ch = 0; ch < 3; ++ch)
                    if (cw[ch] != 0.0f) unsafeAtomicAdd(x.dframe + ch * np + po, cw[ch] * v);
            }
            __syncthreads();
        }
    }
}

// dL/d(tauR, tauD) of one (cell, region) from dL/dTW (normalised weights)
__global__ void k_lgn_tw_adj(vc_lgn_setup q, const float *__restrict__ dTW,
                             float *__restrict__ dTauR, float *__restrict__ dTauD) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= 2 * q.nLGN) return;
    const uint32_t r = id / q.nLGN, c = id - r * q.nLGN, lid = id;
    const uint32_t nKS = q.nKernelSample;
    const float kernelSampleDt = (float)q.kernelSampleInterval * q.dt;
    const float T0 = (float)q.iKernelSampleT0 * q.dt - q.delay[lid];
    const double tauR = q.tauR[lid], tauD = q.tauD[lid], ratio = q.ratio[lid];
    const double nR = q.nR[lid], nD = q.nD[lid];
    const double lfR = lgamma(nR), lfD = lgamma(nD);
    const float *g = dTW + ((size_t)c * 2 + r) * nKS;
    double S = 0.0, Av = 0.0, bR = 0.0, bD = 0.0, cR = 0.0, cD = 0.0;
    for (uint32_t j = 0; j < nKS; ++j) {
        const float t = (float)(nKS - 1 - j) * kernelSampleDt + T0;
        if (t < 0.0f) continue;
        const double tt = t;
        const double tpR = exp((nR - 1.0) * log(tt / tauR) - tt / tauR - lfR) / tauR;
        const double tpD = exp((nD - 1.0) * log(tt / tauD) - tt / tauD - lfD) / tauD;
        const double v = tpR - ratio * tpD;
        const double dvR = tpR * (tt / tauR - nR) / tauR;
        const double dvD = -ratio * tpD * (tt / tauD - nD) / tauD;
        const double sg = v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : 0.0);
        S += fabs(v);
        Av += (double)g[j] * v;
        bR += sg * dvR;
        bD += sg * dvD;
        cR += (double)g[j] * dvR;
        cD += (double)g[j] * dvD;
    }
    // TW_j = v_j / S:  dL/dtau = (sum_j g_j dv_j - (sum_j g_j v_j / S) sum_k sign(v_k) dv_k) / S
    dTauR[lid] = (float)((cR - Av / S * bR) / S);
    dTauD[lid] = (float)((cD - Av / S * bD) / S);
}

uint32_t gcd_u32(uint32_t a, uint32_t b) {
    while (b) {
        uint32_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

int64_t floor_div_h(int64_t a, int64_t b) {
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
    return q;
}

}  // namespace

extern "C" int vc_lgn_rates_adjoint(void *stream, const vc_lgn_params *p, const vc_lgn_state *s,
                                    const float *R, uint32_t it0, uint32_t nsteps,
                                    const float *gFr, float *dR, float *dTW, float *dK) {
    if (!p || !s || !R || !gFr || !dR || !dTW || !dK) {
        vc_set_error("vc_lgn_rates_adjoint: null argument");
        return VC_E_INVALID;
    }
    if (p->nKernelSample == 0 || p->nKernelSample > (uint32_t)kMaxTaps || p->ntPerFrame == 0 ||
        p->kernelSampleInterval == 0 || p->maxFrame == 0) {
        vc_set_error("vc_lgn_rates_adjoint: unsupported kernel sampling (taps %u, interval %u, "
                     "ntPerFrame %u)", p->nKernelSample, p->kernelSampleInterval, p->ntPerFrame);
        return VC_E_UNSUPPORTED;
    }
    if (p->denorm > 1u) {
        vc_set_error("vc_lgn_rates_adjoint: frames that are not a whole number of steps "
                     "(denorm %u) are not built", p->denorm);
        return VC_E_UNSUPPORTED;
    }
    if (nsteps == 0 || p->nLGN == 0) return VC_OK;
    const int64_t npf = p->ntPerFrame, itv = p->kernelSampleInterval;
    const int64_t span = (int64_t)(p->nKernelSample - 1) * itv;
    // window length so that every window's taps cover at most kAdjF frames
    int64_t Tw = (int64_t)(kAdjF - 2) * npf - span;
    if (Tw > kAdjT) Tw = kAdjT;
    if (Tw < 1) {
        vc_set_error("vc_lgn_rates_adjoint: a %lld-step tap span over %lld-step frames is too "
                     "long", (long long)span, (long long)npf);
        return VC_E_UNSUPPORTED;
    }
    const uint32_t L = gcd_u32((uint32_t)itv, (uint32_t)npf);
    const int64_t c0 = (int64_t)p->iKernelSampleT0 - (int64_t)p->nRetrace + 1;
    for (uint32_t w0 = 0; w0 < nsteps; w0 += (uint32_t)Tw) {
        const uint32_t T = (uint32_t)((nsteps - w0) < (uint32_t)Tw ? (nsteps - w0) : Tw);
        const int64_t u0 = (int64_t)it0 + w0 + c0;
        const int64_t Fmin = floor_div_h(u0, npf);
        const int64_t Fmax = floor_div_h(u0 + T - 1 + span, npf);
        const uint32_t nF = (uint32_t)(Fmax - Fmin + 1);
        if (nF > (uint32_t)kAdjF || nF > p->maxFrame) {
            vc_set_error("vc_lgn_rates_adjoint: a window spans %u frames but R holds %u "
                         "(maxFrame must cover every frame of the run)", nF, p->maxFrame);
            return VC_E_INVALID;
        }
        const int64_t jmin = floor_div_h(u0, L), jmax = floor_div_h(u0 + T - 1, L);
        const uint32_t nSeg = (uint32_t)(jmax - jmin + 1);
        hipLaunchKernelGGL(k_lgn_rate_adj, dim3(p->nLGN), dim3(kAdjWG), 0, (hipStream_t)stream,
                           *p, *s, R, it0 + w0, T, u0, Fmin, nF, L, jmin, nSeg,
                           gFr + (size_t)w0 * p->nLGN, dR, dTW, dK);
        VC_HIP_CHECK(hipGetLastError());
    }
    return VC_OK;
}

extern "C" int vc_lgn_frame_responses_adjoint(void *stream, const float *frames,
                                              uint32_t nFrames, uint32_t width, uint32_t height,
                                              uint32_t nLGN, uint32_t nSample1D,
                                              const float *SCx, const float *SCy,
                                              const float *SW, const uint32_t *coneType,
                                              const float *dR, float *dFrames) {
    const uint32_t nS = nSample1D * nSample1D;
    if (!frames || !SCx || !SCy || !SW || !coneType || !dR || !dFrames) {
        vc_set_error("vc_lgn_frame_responses_adjoint: null argument");
        return VC_E_INVALID;
    }
    if (nS < 32 || nS > 1024 || (nS & (nS - 1)) != 0) {
        vc_set_error("nSpatialSample1D^2 = %u must be a power of two in [32, 1024]", nS);
        return VC_E_UNSUPPORTED;
    }
    if (width == 0 || height == 0 || nLGN == 0 || nFrames > 65535u) {
        vc_set_error("vc_lgn_frame_responses_adjoint: empty frame or LGN, or > 65535 frames");
        return VC_E_INVALID;
    }
    if (nFrames == 0) return VC_OK;
    if (nS > 64u * kMaxSPL) {
        vc_set_error("vc_lgn_frame_responses_adjoint: %u samples per region > %d", nS,
                     64 * kMaxSPL);
        return VC_E_UNSUPPORTED;
    }
    const dim3 grid((nLGN + kGrpCells - 1) / kGrpCells, (nFrames + kGrpFrames - 1) / kGrpFrames);
    hipLaunchKernelGGL(k_lgn_frame_adj, grid, dim3(256), 0, (hipStream_t)stream, frames,
                       3ull * width * height, (int)width, (int)height, nLGN, nS, nFrames, SCx,
                       SCy, SW, coneType, dR, dFrames);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_lgn_temporal_adjoint(void *stream, const vc_lgn_setup *q, const float *dTW,
                                       float *dTauR, float *dTauD) {
    if (!q || !dTW || !dTauR || !dTauD) {
        vc_set_error("vc_lgn_temporal_adjoint: null argument");
        return VC_E_INVALID;
    }
    if (q->nLGN == 0 || q->nKernelSample == 0) return VC_OK;
    const uint32_t n = 2 * q->nLGN;
    hipLaunchKernelGGL(k_lgn_tw_adj, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       *q, dTW, dTauR, dTauD);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}
