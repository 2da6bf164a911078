// vc_lgn.hip -- stages a and b of the simulation step (LGN) on gfx950.
//
//  k_lgn_frame_response  once per stimulus frame: one workgroup per LGN cell,
//      one thread per spatial sample.  Bilinear samples of the frame at the
//      cell's centre/surround sample coordinates, mean over the surround
//      samples, clipped contrast, weighted spatial sums R_C, R_S.  The sums
//      use the reference's block_reduce tree (32-lane halving, then halving
//      over the warp partials), so R is bitwise the reference formulation's
//      per-frame value (src/discrete_input_convol.cu:1003-1069, 1143-1211).
//  k_lgn_rate  for a run of steps within one frame period: one wavefront
//      per cell, lane = step, forms tap x frame products from the cached
//      per-frame responses (the separable form: R depends only on the frame,
//      so it is not recomputed per step as the reference does), sums them
//      with the same tree and applies the logistic rate.
//  k_lgn_spike  the time-rescaled Poisson spike draw, one thread per cell,
//      sequential over the steps (src/discrete_input_convol.cu:1587-1821).
//  k_lgn_tw / k_lgn_sw / k_lgn_sc / k_lgn_maxconvol  one-time tables
//      (store_PM, parvo_maxConvol).
//
// MFMA is not used: every contraction here is a per-cell dot product over
// that cell's own irregular sample set (no operand is shared across cells),
// i.e. a batched GEMV, not a GEMM.
#include "vc_lgn_device.h"

namespace {

// block_reduce_ref of two values at once: the same tree for each, one set
// of barriers.
__device__ void block_reduce_ref2(float &x, float &y, float *part /* >= 64 floats */) {
    const uint32_t tid = threadIdx.x;
    const uint32_t W = blockDim.x / 32u;
    x = vc_tree32(x);
    y = vc_tree32(y);
    __syncthreads();
    if ((tid & 31u) == 0u) {
        part[tid >> 5] = x;
        part[32 + (tid >> 5)] = y;
    }
    __syncthreads();
    if (tid == 0) {
        for (uint32_t off = W / 2; off > 0; off >>= 1)
            for (uint32_t l = 0; l < off; ++l) {
                part[l] = part[l] + part[l + off];
                part[32 + l] = part[32 + l] + part[32 + l + off];
            }
    }
    __syncthreads();
    x = part[0];
    y = part[32];
    __syncthreads();
}

__global__ __launch_bounds__(1024) void k_lgn_frame_response(
    const float *__restrict__ frame0, int W, int H, uint32_t nLGN, uint32_t c0,
    const float *__restrict__ SCx, const float *__restrict__ SCy,
    const float *__restrict__ SW, const uint32_t *__restrict__ coneType,
    const uint32_t *__restrict__ cellTrial, uint64_t trialFrameStride,
    uint64_t frameStride, float *__restrict__ R) {
    __shared__ float part[64];
    const uint32_t c = c0 + blockIdx.x;
    // blockIdx.y: frame of a batch of frames frameStride floats apart
    frame0 += (size_t)blockIdx.y * frameStride;
    R += (size_t)blockIdx.y * nLGN * 2;
    const float *frame = cellTrial ? frame0 + (size_t)cellTrial[c] * trialFrameStride : frame0;
    const uint32_t s = threadIdx.x;
    const uint32_t nS = blockDim.x;
    const uint32_t tC = coneType[c], tS = coneType[nLGN + c];
    const size_t iC = ((size_t)c * 2 + 0) * nS + s;
    const size_t iS = ((size_t)c * 2 + 1) * nS + s;
    const float xC = SCx[iC], yC = SCy[iC], xS = SCx[iS], yS = SCy[iS];
    const float sw = SW[s];
    const float IC_S = intensity(tC, frame, W, H, xS, yS);
    const float IS_S = (tS == tC) ? IC_S : intensity(tS, frame, W, H, xS, yS);
    const float IC_C = intensity(tC, frame, W, H, xC, yC);
    float sumC = IC_S, sumS = IS_S;
    block_reduce_ref2(sumC, sumS, part);
    const float meanC = sumC / (float)nS;
    const float meanS = sumS / (float)nS;
    float rC = sw * clip_contrast(IC_C, meanC);
    float rS = sw * clip_contrast(IS_S, meanS);
    block_reduce_ref2(rC, rS, part);
    if (s == 0) {
        R[(size_t)c * 2 + 0] = rC;
        R[(size_t)c * 2 + 1] = rS;
    }
}

// The same response with one wavefront per cell (nS a multiple of 64): lane
// l holds samples l + 64k, k < kPer.  The reference's block reduction
// (src/discrete_input_convol.cu:1055-1069, :1143-1211; emulated by
// block_reduce_ref2) sums each warp of 32 consecutive samples with a 32-lane
// tree -- here the half-wave trees of register k give warps 2k and 2k+1 --
// and then halves over the nS/32 warp partials, here in registers.  The
// additions are the same, so the results are identical; there is no barrier.
// With a sample-geometry table (geom, VC_LGN_GEOM floats per region; SCx,
// SCy unused) the coordinates are regenerated from the sample's grid position
// with the table kernel's own operations (sample_xy), so they equal the SC
// tables bit for bit and the 16 bytes per sample are not read.
struct GeomArgs {
    const float *geom;
    float x0, y0, nvd;
    uint32_t n1D;
};
template <int kPer, bool kGeom>
__global__ __launch_bounds__(256) void k_lgn_frame_response_w(
    const float *__restrict__ frame0, int W, int H, uint32_t nLGN, uint32_t c0, uint32_t c1,
    const float *__restrict__ SCx, const float *__restrict__ SCy,
    const float *__restrict__ SW, const uint32_t *__restrict__ coneType,
    const uint32_t *__restrict__ cellTrial, uint64_t trialFrameStride,
    uint64_t frameStride, float *__restrict__ R, GeomArgs ga) {
    constexpr uint32_t nS = 64u * kPer, nW = 2u * kPer;
    const uint32_t c = c0 + __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) +
                                                           (threadIdx.x >> 6));
    if (c >= c1) return;                      // whole wavefronts
    const uint32_t lane = threadIdx.x & 63u;
    frame0 += (size_t)blockIdx.y * frameStride;
    R += (size_t)blockIdx.y * nLGN * 2;
    const float *frame = cellTrial ? frame0 + (size_t)cellTrial[c] * trialFrameStride : frame0;
    const uint32_t tC = coneType[c], tS = coneType[nLGN + c];
    SampleGeom gC{}, gS{};
    if (kGeom) {                               // wave-uniform: scalar loads
        const float *g = ga.geom + (size_t)c * 2 * VC_LGN_GEOM;
        gC = SampleGeom{g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7]};
        gS = SampleGeom{g[8], g[9], g[10], g[11], g[12], g[13], g[14], g[15]};
    }
    float icS[kPer], isS[kPer], icC[kPer], sw[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint32_t smp = lane + 64u * k;
        float xC, yC, xS, yS;
        if (kGeom) {
            const uint32_t tx = smp % ga.n1D, ty = smp / ga.n1D;
            sample_xy(gC, ga.x0, ga.y0, ga.nvd, tx, ty, xC, yC);
            sample_xy(gS, ga.x0, ga.y0, ga.nvd, tx, ty, xS, yS);
        } else {
            const size_t iC = ((size_t)c * 2 + 0) * nS + smp;
            const size_t iS = ((size_t)c * 2 + 1) * nS + smp;
            xC = SCx[iC]; yC = SCy[iC]; xS = SCx[iS]; yS = SCy[iS];
        }
        sw[k] = SW[smp];
        icS[k] = intensity(tC, frame, W, H, xS, yS);
        // centre and surround of one cone type see the same surround intensity
        isS[k] = (tS == tC) ? icS[k] : intensity(tS, frame, W, H, xS, yS);
        icC[k] = intensity(tC, frame, W, H, xC, yC);
    }
    // warp partials, then the halving over them (as block_reduce_ref2)
    auto reduce = [&](const float (&x)[kPer], const float (&y)[kPer], float &sx, float &sy) {
        float px[nW], py[nW];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const float a = vc_tree32(x[k]), b = vc_tree32(y[k]);
            px[2 * k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a), 0));
            px[2 * k + 1] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a), 32));
            py[2 * k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(b), 0));
            py[2 * k + 1] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(b), 32));
        }
#pragma unroll
        for (uint32_t off = nW / 2; off > 0; off >>= 1) {
#pragma unroll
            for (uint32_t l = 0; l < off; ++l) {
                px[l] = px[l] + px[l + off];
                py[l] = py[l] + py[l + off];
            }
        }
        sx = px[0];
        sy = py[0];
    };
    float sumC, sumS;
    reduce(icS, isS, sumC, sumS);
    const float meanC = sumC / (float)nS;
    const float meanS = sumS / (float)nS;
    float rC[kPer], rS[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        rC[k] = sw[k] * clip_contrast(icC[k], meanC);
        rS[k] = sw[k] * clip_contrast(isS[k], meanS);
    }
    float outC, outS;
    reduce(rC, rS, outC, outS);
    if (lane == 0) {
        R[(size_t)c * 2 + 0] = outC;
        R[(size_t)c * 2 + 1] = outS;
    }
}

// Launch of the frame response of cells c0 .. c1-1 for nFrames frames
// frameStride floats apart: one wavefront per cell when nS is a multiple of
// 64, else one workgroup.
hipError_t launch_frame_response(hipStream_t st, uint32_t nS, uint32_t nFrames,
                                 const float *frame, uint32_t width, uint32_t height,
                                 uint32_t nLGN, uint32_t c0, uint32_t c1, const float *SCx,
                                 const float *SCy, const float *SW, const uint32_t *coneType,
                                 const uint32_t *cellTrial, uint64_t trialFrameStride,
                                 uint64_t frameStride, float *R, const GeomArgs *ga = nullptr) {
    if (c1 <= c0) return hipSuccess;
    const uint32_t nc = c1 - c0;
    const dim3 grid((nc + 3u) / 4u, nFrames), block(256);
    const GeomArgs g0{nullptr, 0.0f, 0.0f, 0.0f, 0u};
    if (ga) {                         // square grids of 64, 256 or 1024 samples
        switch (nS) {
#define VC_FRG(K)                                                                                \
    hipLaunchKernelGGL((k_lgn_frame_response_w<K, true>), grid, block, 0, st, frame, (int)width, \
                       (int)height, nLGN, c0, c1, SCx, SCy, SW, coneType, cellTrial,            \
                       trialFrameStride, frameStride, R, *ga)
        case 64: VC_FRG(1); break;
        case 256: VC_FRG(4); break;
        case 1024: VC_FRG(16); break;
        default: return hipErrorInvalidValue;
#undef VC_FRG
        }
        return hipGetLastError();
    }
#define VC_FR(K)                                                                                  \
    hipLaunchKernelGGL((k_lgn_frame_response_w<K, false>), grid, block, 0, st, frame, (int)width, \
                       (int)height, nLGN, c0, c1, SCx, SCy, SW, coneType, cellTrial,             \
                       trialFrameStride, frameStride, R, g0)
    switch (nS) {
    case 64: VC_FR(1); break;
    case 128: VC_FR(2); break;
    case 256: VC_FR(4); break;
    case 512: VC_FR(8); break;
    case 1024: VC_FR(16); break;
    default:
        hipLaunchKernelGGL(k_lgn_frame_response, dim3(nc, nFrames), dim3(nS), 0, st, frame,
                           (int)width, (int)height, nLGN, c0, SCx, SCy, SW, coneType, cellTrial,
                           trialFrameStride, frameStride, R);
    }
#undef VC_FR
    return hipGetLastError();
}

// in-register 32-element halving tree (the same additions as vc_tree32)
__device__ __forceinline__ float tree32_regs(float a[32]) {
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) {
#pragma unroll
        for (int l = 0; l < off; ++l) a[l] = a[l] + a[l + off];
    }
    return a[0];
}

constexpr int kMaxFrame = 16;              // frames under one wavefront's tap windows
constexpr int kWin = 16;                    // rate window of k_lgn_spike


// Stage a (temporal part) and the rate of stage b for every (cell, step):
// one wavefront per cell, lane = step, so the cell's temporal weights and
// frame responses are wave-uniform (scalar loads); the tap products and the
// reference's tap-sum tree are evaluated in registers.  kNT > 0 fixes the
// tap count at compile time (the configurations' 40), 0 takes it from p.
template <int kNT>
__global__ __launch_bounds__(256) void k_lgn_rate(vc_lgn_params p, vc_lgn_state st,
                                                  const float *__restrict__ Rring,
                                                  uint32_t it0, uint32_t nsteps,
                                                  float *__restrict__ frOut) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t wpc = (nsteps + 63u) / 64u;               // wavefronts per cell
    const uint32_t wave = __builtin_amdgcn_readfirstlane(g >> 6);
    const uint32_t c = wave / wpc;
    if (c >= p.nLGN) return;
    const uint32_t s = (wave - c * wpc) * 64u + (g & 63u);
    const uint32_t nKS = p.nKernelSample;
    const uint32_t it = it0 + (s < nsteps ? s : 0u);
    // time in units of dt/denorm: frame f covers [f ntPerFrame, (f+1) ntPerFrame)
    const int64_t den = p.denorm ? (int64_t)p.denorm : 1;
    const int64_t base = ((int64_t)it + (int64_t)p.iKernelSampleT0 - (int64_t)p.nRetrace + 1) * den;
    // responses of the frames under the first step's window, wave-uniform
    const int64_t fw = floor_div(((int64_t)(it0 + (wave - c * wpc) * 64u) + (int64_t)p.iKernelSampleT0 -
                                  (int64_t)p.nRetrace + 1) * den, (int64_t)p.ntPerFrame);
    // frame responses of this cell under the wavefront's windows, in LDS so
    // that each lane can pick its tap's frame with one read
    __shared__ float sR[256 / 64][2][kMaxFrame + 1];
    const uint32_t wv = threadIdx.x >> 6, ln = threadIdx.x & 63u;
    if (ln < 2u * (kMaxFrame + 1)) {
        const uint32_t r = ln / (kMaxFrame + 1), d = ln % (kMaxFrame + 1);
        int64_t slot = (fw + d) % (int64_t)p.maxFrame;
        if (slot < 0) slot += p.maxFrame;
        sR[wv][r][d] = Rring[((size_t)slot * p.nLGN + c) * 2 + r];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // frame of tap i: floor((base + i*interval) / ntPerFrame), walked
    // incrementally from tap 0 (quotient dd relative to fw, remainder rr)
    const int64_t q0 = floor_div(base, (int64_t)p.ntPerFrame);
    const uint32_t rr0 = (uint32_t)(base - q0 * (int64_t)p.ntPerFrame);
    const int dd0 = (int)(q0 - fw);
    const uint32_t npf = p.ntPerFrame;
    const uint32_t tstep = p.kernelSampleInterval * (uint32_t)den;   // one tap, in dt/denorm
    const uint32_t dq = tstep / npf, dr = tstep % npf;
    float conv[2];
#pragma unroll 1
    for (int r = 0; r < 2; ++r) {              // one region at a time: half the registers
        float a0[32], a1[32];
        uint32_t rr = rr0;
        int dd = dd0;
        // the cell's temporal weights of region r: lane i holds tap i, read
        // back per tap as a scalar operand
        const int tw = __float_as_int(ln < nKS ? st.TW[((size_t)c * 2 + r) * nKS + ln] : 0.0f);
        constexpr int kT = kNT > 0 ? kNT : kMaxTaps;   // taps walked (the rest are zeros)
#pragma unroll
        for (int i = 0; i < kMaxTaps; ++i) {
            float v = 0.0f;
            if (i < kT) {
                if (kNT > 0 || i < (int)nKS)
                    v = sR[wv][r][dd] * __int_as_float(__builtin_amdgcn_readlane(tw, i));
                rr += dr;
                dd += (int)dq;
                if (rr >= npf) {
                    rr -= npf;
                    dd += 1;
                }
            }
            if (i < 32) a0[i] = v; else a1[i - 32] = v;
        }
        const float P0 = tree32_regs(a0);
        const float P1 = tree32_regs(a1);
        // halving over the 8 warp partials of a 256-thread block, 6 of them zero
        const float sum = ((P0 + 0.0f) + (0.0f + 0.0f)) + ((P1 + 0.0f) + (0.0f + 0.0f));
        conv[r] = 0.0f + sum;
    }
    const float convolC = conv[0] * st.kC[c];
    const float convolS = conv[1] * st.kS[c];
    const float convol = convolC + convolS;
    float current = convol * p.convolRatio;
    if (current < 0.0f) current = 0.0f;
    const float mx = st.maxConvol[c];
    const float fr = p.frRatio * mx * transform(st.c50[c], st.sharpness[c], st.logA[c], st.logB[c],
                                                current / mx);
    if (s < nsteps) frOut[(size_t)s * p.nLGN + c] = fr;
}

// k_lgn_rate for a short run of steps (at most kShortRun, e.g. a chunk that
// ends a V1 launch early): lane = cell, each wavefront a slice of kStepsPerWave
// steps in sequence (blockIdx.y), so that every lane works (lane = step would
// leave most of a wavefront idle).  The cell's temporal weights are
// registers, its frame responses a per-lane LDS row; the tap -> frame walk is
// wave-uniform.  Per (cell, step) the products, the tap-sum tree and the rate
// are those of k_lgn_rate, bit for bit.
constexpr uint32_t kShortRun = 32;
constexpr uint32_t kStepsPerWave = 4;
template <int kNT>
__global__ __launch_bounds__(64) void k_lgn_rate_cells(vc_lgn_params p, vc_lgn_state st,
                                                       const float *__restrict__ Rring,
                                                       uint32_t it0, uint32_t nsteps,
                                                       float *__restrict__ frOut) {
    constexpr int kF = kMaxFrame + 1;
    constexpr int kRow = 2 * kF + 1;                   // odd stride: no bank conflicts
    __shared__ float sR[64 * kRow];
    const uint32_t ln = threadIdx.x;
    const uint32_t c = blockIdx.x * 64u + ln;
    const bool valid = c < p.nLGN;
    const uint32_t cc = valid ? c : 0u;
    const uint32_t npf = p.ntPerFrame;
    const uint32_t den = p.denorm ? p.denorm : 1u;           // time in units of dt/denorm
    // frames under the run's windows: from step 0's first tap on
    const int64_t base0 = (int64_t)it0 + (int64_t)p.iKernelSampleT0 - (int64_t)p.nRetrace + 1;
    const int64_t fw = floor_div(base0 * (int64_t)den, (int64_t)npf);
    // the run's taps span at most (nsteps - 1 + nRetrace) denorm / npf + 2 frames
    const uint32_t nF = ((nsteps - 1u + p.nRetrace) * den) / npf + 2u;
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int d = 0; d < kF; ++d) {
            if ((uint32_t)d >= nF) break;
            int64_t slot = (fw + d) % (int64_t)p.maxFrame;
            if (slot < 0) slot += p.maxFrame;
            sR[ln * kRow + r * kF + d] = Rring[((size_t)slot * p.nLGN + cc) * 2 + r];
        }
    float tw[2][kNT];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int i = 0; i < kNT; ++i) tw[r][i] = st.TW[((size_t)cc * 2 + r) * kNT + i];
    const float kC = st.kC[cc], kS = st.kS[cc], mx = st.maxConvol[cc];
    const float c50 = st.c50[cc], sh = st.sharpness[cc], lA = st.logA[cc], lB = st.logB[cc];
    const uint32_t tstep = p.kernelSampleInterval * den;
    const uint32_t dq = tstep / npf, dr = tstep % npf;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    const uint32_t s1 = (blockIdx.y + 1u) * kStepsPerWave < nsteps ? (blockIdx.y + 1u) * kStepsPerWave
                                                                    : nsteps;
    for (uint32_t s = blockIdx.y * kStepsPerWave; s < s1; ++s) {
        const int64_t base = (base0 + (int64_t)s) * (int64_t)den;
        const int64_t q0 = floor_div(base, (int64_t)npf);
        const uint32_t rr0 = (uint32_t)(base - q0 * (int64_t)npf);
        const int dd0 = (int)(q0 - fw);
        float conv[2];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            float a0[32], a1[32];
            uint32_t rr = rr0;
            int dd = dd0;
#pragma unroll
            for (int i = 0; i < kMaxTaps; ++i) {
                float v = 0.0f;
                if (i < kNT) {
                    v = sR[ln * kRow + r * kF + dd] * tw[r][i];
                    rr += dr;
                    dd += (int)dq;
                    if (rr >= npf) {
                        rr -= npf;
                        dd += 1;
                    }
                }
                if (i < 32) a0[i] = v; else a1[i - 32] = v;
            }
            const float P0 = tree32_regs(a0);
            const float P1 = tree32_regs(a1);
            const float sum = ((P0 + 0.0f) + (0.0f + 0.0f)) + ((P1 + 0.0f) + (0.0f + 0.0f));
            conv[r] = 0.0f + sum;
        }
        const float convolC = conv[0] * kC;
        const float convolS = conv[1] * kS;
        const float convol = convolC + convolS;
        float current = convol * p.convolRatio;
        if (current < 0.0f) current = 0.0f;
        const float fr = p.frRatio * mx * transform(c50, sh, lA, lB, current / mx);
        if (valid) frOut[(size_t)s * p.nLGN + c] = fr;
    }
}

// k_lgn_rate with the tap count fixed when it is the configurations' 40.
void launch_rate(hipStream_t st, uint64_t waves, const vc_lgn_params &p, const vc_lgn_state &s,
                 const float *R, uint32_t it0, uint32_t nsteps, float *fr) {
    const dim3 grid((uint32_t)((waves * 64 + 255) / 256)), block(256);
    if (p.nKernelSample == 40)
        hipLaunchKernelGGL(k_lgn_rate<40>, grid, block, 0, st, p, s, R, it0, nsteps, fr);
    else
        hipLaunchKernelGGL(k_lgn_rate<0>, grid, block, 0, st, p, s, R, it0, nsteps, fr);
}

// Stage b, the time-rescaled Poisson draw (get_spike,
// src/discrete_input_convol.cu:1587-1622): sequential in time per cell.
// The steps are taken 64 at a time: first every step's draw (no memory
// round trip inside the loop: the rates are prefetched, the outputs stored
// without waiting), then the deliveries of the window's spikes to the
// synapse bits of the V1 neurons they feed, in groups of kDelGroup (spike,
// cell) pairs whose target lists are loaded together.
constexpr int kDelGroup = 4;
// The LGN cells' LTP trace of the feed-forward plasticity (NULL var = off).
struct LgnTrace {
    float tauLTP;
    float *var;        // [nLGN] trace at the start of the next step
    float *xStart;     // [nsteps][nLGN] trace at the start of each step
    float *xSpike;     // [nsteps][nLGN] trace at the step's spike time
    uint64_t *bits;    // [nsteps][(nLGN + 63)/64] bit c%64 of word c/64: cell c spiked
};

__global__ __launch_bounds__(256) void k_lgn_spike(vc_lgn_params p, vc_lgn_state st,
                                                   uint32_t it0, uint32_t nsteps,
                                                   const float *__restrict__ frIn,
                                                   float *__restrict__ sInfoOut,
                                                   vc_ff_events ev, LgnTrace lt) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = c < p.nLGN;            // every lane runs the step loop (ballots)
    const uint32_t lane = vc_lane();
    float lTR = valid ? st.leftTimeRate[c] : 0.0f;
    float lNL = valid ? st.lastNegLogRand[c] : 1.0f;
    uint32_t klo = p.seed_lo, khi = p.seed_hi, ent = c;     // Philox key, entity id
    if (valid && p.cellTrial != nullptr) {
        const uint32_t tr = p.cellTrial[c];
        klo = p.trialSeed[2 * tr];
        khi = p.trialSeed[2 * tr + 1];
        ent = c - p.trialLgnOff[tr];
    }
    const bool trace = lt.var != nullptr;
    float x = (trace && valid) ? lt.var[c] : 0.0f;
    // this cell's target list, read once
    uint32_t myk0 = 0, myk1 = 0;
    if (ev.mask != nullptr && valid) {
        myk0 = ev.revOff[c];
        myk1 = ev.revOff[c + 1];
    }
    // one step of one cell: rate = fr/1000 and rT = dt*rate come precomputed
    // (the same operations, taken out of the time-sequential chain, which
    // without a spike is one compare and one add per step)
    // nsp and tsp: the step's spike count and mean spike time
    auto draw = [&](uint32_t s, float rate, float rT, uint32_t &nsp, float &tsp) -> float {
        const uint32_t it = it0 + s;
        float next_rT = lNL - lTR;
        nsp = 0;
        tsp = 0.0f;
        float lNLn = lNL;
        if (next_rT > rT) {
            lTR += rT;
            return 0.0f;                       // nsp + tsp/dt with nsp = tsp = 0
        } else {
            float old_rT = 0.0f;
            uint32_t kdraw = 0;
            vc_u32x4 pack;
            const float tr = p.tRef * rate;
            do {
                if ((kdraw & 3u) == 0u)
                    pack = vc_draw4(klo, khi, ent, it, VC_STREAM_LGN_SPIKE, kdraw >> 2, p.trial);
                const float u = vc_u01_open0(pack.v[kdraw & 3u]);
                kdraw++;
                const float x = -vc_logf(u);
                lNLn = (x > tr) ? x : tr;
                old_rT = next_rT;
                next_rT += lNLn;
                nsp++;
            } while (next_rT <= rT);
            lTR = rT - old_rT;
            tsp = old_rT / (rate * (float)nsp);
        }
        const float sInfo = (float)nsp + tsp / p.dt;
        if (sInfo > 0.0f) lNL = lNLn;
        return sInfo;
    };
    for (uint32_t w0 = 0; w0 < nsteps; w0 += 64u) {
        const uint32_t wn = (nsteps - w0) < 64u ? (nsteps - w0) : 64u;
        // ---- the draws of the window; the rates arrive kWin steps at a time
        uint64_t spiked = 0;                   // bit s: this cell spiked in step w0 + s
        float cur[kWin], nxt[kWin];
#pragma unroll
        for (uint32_t j = 0; j < (uint32_t)kWin; ++j)
            cur[j] = (valid && j < wn) ? frIn[(size_t)(w0 + j) * p.nLGN + c] : 0.0f;
        for (uint32_t s0 = 0; s0 < wn; s0 += (uint32_t)kWin) {
#pragma unroll
            for (uint32_t j = 0; j < (uint32_t)kWin; ++j) {
                const uint32_t sn = s0 + (uint32_t)kWin + j;
                nxt[j] = (valid && sn < wn) ? frIn[(size_t)(w0 + sn) * p.nLGN + c] : 0.0f;
            }
            float rate[kWin], rT[kWin];
#pragma unroll
            for (uint32_t j = 0; j < (uint32_t)kWin; ++j) {
                rate[j] = cur[j] / 1000.0f;
                rT[j] = p.dt * rate[j];
            }
#pragma unroll
            for (uint32_t j = 0; j < (uint32_t)kWin; ++j) {
                const uint32_t sl = s0 + j;
                if (valid && sl < wn) {
                    uint32_t nsp;
                    float tsp;
                    const float sInfo = draw(w0 + sl, rate[j], rT[j], nsp, tsp);
                    const size_t o = (size_t)(w0 + sl) * p.nLGN + c;
                    sInfoOut[o] = sInfo;
                    if (sInfo >= 1.0f) spiked |= 1ull << sl;
                    if (trace) {
                        // LGN_nonlinear's learning part: the spikes are added
                        // before the decay to their mean time
                        lt.xStart[o] = x;
                        float rest = p.dt;
                        if (sInfo > 0.0f) {
                            x += (float)nsp;
                            x *= vc_expf(-tsp / lt.tauLTP);
                            rest = p.dt - tsp;
                        }
                        lt.xSpike[o] = x;
                        x *= vc_expf(-rest / lt.tauLTP);
                    }
                }
            }
#pragma unroll
            for (uint32_t j = 0; j < (uint32_t)kWin; ++j) cur[j] = nxt[j];
        }
        if (trace) {
            // the window's spike bitmaps: lane s stores the wavefront's word of
            // step w0 + s (a wavefront holds 64 consecutive cells)
            uint64_t mine = 0;
            for (uint32_t s = 0; s < wn; ++s) {
                const uint64_t b = __ballot((spiked >> s) & 1ull);
                if (lane == s) mine = b;
            }
            const uint32_t nWords = (p.nLGN + 63u) / 64u;
            if (lane < wn && (c >> 6) < nWords) lt.bits[(size_t)(w0 + lane) * nWords + (c >> 6)] = mine;
        }
        if (ev.mask == nullptr) continue;
        // ---- deliveries, in (step, cell) order: the wavefront walks the
        // spikes kDelGroup at a time, lanes over each spike's targets; the
        // group's target loads are issued together before its atomics
        uint32_t sstep = 0;
        uint64_t sm = __ballot(spiked & 1ull);
        auto next_pair = [&](uint32_t &l, uint32_t &sp) -> bool {
            while (sm == 0ull) {
                if (++sstep >= wn) return false;
                sm = __ballot((spiked >> sstep) & 1ull);
            }
            l = (uint32_t)__builtin_ctzll(sm);
            sm &= sm - 1ull;
            sp = sstep;
            return true;
        };
        for (;;) {
            uint32_t gl[kDelGroup], gs[kDelGroup], gk0[kDelGroup], gk1[kDelGroup];
            uint32_t ng = 0;
#pragma unroll
            for (int q = 0; q < kDelGroup; ++q) {
                gl[q] = gs[q] = gk0[q] = gk1[q] = 0u;
                if ((uint32_t)q == ng && next_pair(gl[q], gs[q])) {
                    gk0[q] = vc_readlane_u32(myk0, gl[q]);
                    gk1[q] = vc_readlane_u32(myk1, gl[q]);
                    ++ng;
                }
            }
            if (ng == 0) break;
            // the first two 64-target chunks of every pair in the group
            uint32_t tg[kDelGroup][2];
#pragma unroll
            for (int q = 0; q < kDelGroup; ++q)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const uint32_t k = gk0[q] + 64u * (uint32_t)h + lane;
                    tg[q][h] = ((uint32_t)q < ng && k < gk1[q]) ? ev.revTgt[k] : 0xffffffffu;
                }
#pragma unroll
            for (int q = 0; q < kDelGroup; ++q) {
                if ((uint32_t)q >= ng) break;
                uint32_t *row = ev.mask + (size_t)(w0 + gs[q]) * ev.nV1;
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    if (tg[q][h] != 0xffffffffu)
                        atomicOr(row + (tg[q][h] >> 5), 1u << (tg[q][h] & 31u));
                for (uint32_t k = gk0[q] + 128u + lane; k < gk1[q]; k += 64u) {
                    const uint32_t t2 = ev.revTgt[k];
                    atomicOr(row + (t2 >> 5), 1u << (t2 & 31u));
                }
            }
            if (ng < (uint32_t)kDelGroup) break;
        }
    }
    if (valid) {
        st.leftTimeRate[c] = lTR;
        st.lastNegLogRand[c] = lNL;
        if (trace) lt.var[c] = x;
    }
}

__global__ void k_lgn_init(vc_lgn_params p, float *lTR, float *lNL) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= p.nLGN) return;
    uint32_t klo = p.seed_lo, khi = p.seed_hi, ent = c;
    if (p.cellTrial != nullptr) {
        const uint32_t tr = p.cellTrial[c];
        klo = p.trialSeed[2 * tr];
        khi = p.trialSeed[2 * tr + 1];
        ent = c - p.trialLgnOff[tr];
    }
    vc_u32x4 r = vc_draw4(klo, khi, ent, 0xffffffffu, VC_STREAM_LGN_INIT, 0u, p.trial);
    lTR[c] = 0.0f;
    lNL[c] = -vc_logf(vc_u01_open0(r.v[0]));
}

// ---------------------------------------------------------------- setup --


// store_temporalWeight, one workgroup of nS threads per (cell, region)
__global__ __launch_bounds__(1024) void k_lgn_tw(vc_lgn_setup q, float *__restrict__ TW) {
    __shared__ float part[32];
    const uint32_t c = blockIdx.x, r = blockIdx.y, tid = threadIdx.x, nS = blockDim.x;
    const uint32_t lid = r * q.nLGN + c, nKS = q.nKernelSample;
    const float kernelSampleDt = (float)q.kernelSampleInterval * q.dt;
    const float kernelSampleT0 = (float)q.iKernelSampleT0 * q.dt;
    const float lfac1 = lgammaf(q.nR[lid]);
    const float lfac2 = lgammaf(q.nD[lid]);
    const float T0 = kernelSampleT0 - q.delay[lid];
    float *tw = TW + ((size_t)c * 2 + r) * nKS;
    const uint32_t nPatch = (nKS + nS - 1) / nS;
    float total = 0.0f;
    for (uint32_t ip = 0; ip < nPatch; ++ip) {
        const uint32_t nAct = (ip + 1 < nPatch) ? nS : nKS - ip * nS;
        float a = 0.0f;
        if (tid < nAct) {
            uint32_t j = ip * nS + tid;
            uint32_t twid = nKS - 1 - j;
            float t = (float)twid * kernelSampleDt + T0;
            float v = 0.0f;
            if (!(t < 0.0f))
                v = temporal_kernel(t, q.tauR[lid], q.tauD[lid], q.ratio[lid], q.nR[lid],
                                    q.nD[lid], lfac1, lfac2);
            tw[j] = v;
            a = fabsf(v);
        }
        total += block_reduce_ref(a, part);
    }
    __syncthreads();
    for (uint32_t j = tid; j < nKS; j += nS) tw[j] /= total;
}

// shared spatial weights from cell 0's centre (src/discrete_input_convol.cu:332-336)
__global__ __launch_bounds__(1024) void k_lgn_sw(vc_lgn_setup q, float *__restrict__ SW) {
    __shared__ float part[32];
    const uint32_t n1D = q.nSample1D;
    const uint32_t tx = threadIdx.x % n1D, ty = threadIdx.x / n1D;
    const float rx = q.rw[0], ry = q.rh[0];
    const float wSpan = (float)((double)(q.nsig * rx) / kSqrt2);
    const float dw = 2 * wSpan / (float)n1D;
    const float hSpan = (float)((double)(q.nsig * ry) / kSqrt2);
    const float dh = 2 * hSpan / (float)n1D;
    const float w = (float)((tx + 0.5) * (double)dw - (double)wSpan);
    const float h = (float)((ty + 0.5) * (double)dh - (double)hSpan);
    const float sw = spatial_kernel(w, h, rx, ry) * 1.0f;
    const float tot = block_reduce_ref(sw, part);
    SW[threadIdx.x] = sw / tot;
}

// the per-region sample geometry of every cell: geom[(c*2 + r)*VC_LGN_GEOM ...]
__global__ __launch_bounds__(256) void k_lgn_geom(vc_lgn_setup q, float *__restrict__ geom) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2u * q.nLGN) return;
    const uint32_t c = i >> 1, r = i & 1u;
    const SampleGeom g = sample_geom(q, r * q.nLGN + c);
    float *o = geom + (size_t)i * VC_LGN_GEOM;
    o[0] = g.ex; o[1] = g.ey; o[2] = g.coso; o[3] = g.sino;
    o[4] = g.dw; o[5] = g.wSpan; o[6] = g.dh; o[7] = g.hSpan;
}

// sample texture coordinates, flat retina (src/discrete_input_convol.cu:398-417)
__global__ __launch_bounds__(1024) void k_lgn_sc(vc_lgn_setup q, float *__restrict__ SCx,
                                                 float *__restrict__ SCy) {
    const uint32_t c = blockIdx.x, r = blockIdx.y;
    const uint32_t n1D = q.nSample1D, nS = n1D * n1D;
    const uint32_t tx = threadIdx.x % n1D, ty = threadIdx.x / n1D;
    const SampleGeom g = sample_geom(q, r * q.nLGN + c);
    float x, y;
    sample_xy(g, q.x0, q.y0, q.normViewDistance, tx, ty, x, y);
    const size_t id = ((size_t)c * 2 + r) * nS + threadIdx.x;
    SCx[id] = x;
    SCy[id] = y;
}

// parvo_maxConvol, src/discrete_input_convol.cu:746-881 (1024 threads per cell)
__global__ __launch_bounds__(1024) void k_lgn_maxconvol(vc_lgn_setup q, const float *__restrict__ TW,
                                                        float *__restrict__ maxConvol) {
    extern __shared__ float sw[];          // [2][mS]
    __shared__ float part[32];
    const uint32_t c = blockIdx.x, tid = threadIdx.x, nLGN = q.nLGN, nKS = q.nKernelSample;
    const uint32_t m1D = q.maxSample1D, mS = m1D * m1D, nPatch = mS / 1024;
    float *swC = sw, *swS = sw + mS;
    const float rxs = q.rw[nLGN + c], rys = q.rh[nLGN + c];
    const float xs = q.polar[nLGN + c], ys = q.ecc[nLGN + c];
    float wSpan = (float)((double)(q.nsig * rxs) / kSqrt2);
    float hSpan = (float)((double)(q.nsig * rys) / kSqrt2);
    const float rxc = q.rw[c], ryc = q.rh[c];
    float R = wSpan > hSpan ? wSpan : hSpan;
    const float xc = q.polar[c], yc = q.ecc[c];
    wSpan = (float)((double)(q.nsig * rxc) / kSqrt2);
    hSpan = (float)((double)(q.nsig * ryc) / kSqrt2);
    const float dx = yc * cosf(xc) - ys * cosf(xs);
    const float dy = yc * sinf(xc) - ys * sinf(xs);
    float r = sqrtf(dx * dx + dy * dy);
    const float cov = q.covariant[c];
    const float orients = q.orient[nLGN + c], orientc = q.orient[c];
    r += wSpan > hSpan ? wSpan : hSpan;
    const float coss = cosf(orients);
    R = r > R ? r : R;
    const float sins = sinf(orients);
    const float cosc = cosf(orientc), sinc = sinf(orientc);
    const float ks = q.k[nLGN + c], kc = q.k[c];
    const float ds = 2 * R / (float)m1D;
    float sumC = 0.0f, sumS = 0.0f;
    for (uint32_t ip = 0; ip < nPatch; ++ip) {
        const uint32_t pid = ip * 1024 + tid;
        const uint32_t wi = pid % m1D, hi = pid / m1D;
        const float lx = (float)((wi + 0.5) * (double)ds - (double)R);
        const float ly = (float)((hi + 0.5) * (double)ds - (double)R);
        float x = cosc * (lx - dx) + sinc * (ly - dy);
        float y = -sinc * (lx - dx) + cosc * (ly - dy);
        const float lc = spatial_kernel(x, y, rxc, ryc);
        swC[pid] = lc;
        sumC += block_reduce_ref(lc, part);
        x = coss * lx + sins * ly;
        y = -sins * lx + coss * ly;
        const float ls = spatial_kernel(x, y, rxs, rys);
        swS[pid] = ls;
        sumS += block_reduce_ref(ls, part);
    }
    __syncthreads();
    for (uint32_t ip = 0; ip < nPatch; ++ip) {
        const uint32_t pid = ip * 1024 + tid;
        swC[pid] /= sumC;
        swS[pid] /= sumS;
    }
    __syncthreads();
    float convol = 0.0f;
    for (uint32_t it = 0; it < nKS; ++it) {
        const float tc = TW[(size_t)c * 2 * nKS + it];
        const float ts = TW[((size_t)c * 2 + 1) * nKS + it];
        for (uint32_t ip = 0; ip < nPatch; ++ip) {
            const uint32_t pid = ip * 1024 + tid;
            const float lc = swC[pid] * tc * kc;
            const float ls = swS[pid] * ts * ks;
            float d;
            if (fabsf(lc) > fabsf(ls)) d = fabsf(lc) + ls * copysignf(cov, lc);
            else d = fabsf(ls) + lc * copysignf(cov, ls);
            convol += block_reduce_ref(d, part);
        }
    }
    if (tid == 0) maxConvol[c] = convol;
}

}  // namespace

extern "C" int vc_lgn_precompute(void *stream, const vc_lgn_setup *q, float *TW, float *SW,
                                 float *SCx, float *SCy, float *maxConvol) {
    if (!q || !TW || !SW || !SCx || !SCy || !maxConvol) {
        vc_set_error("vc_lgn_precompute: null argument");
        return VC_E_INVALID;
    }
    const uint32_t nS = q->nSample1D * q->nSample1D;
    if (nS < 32 || nS > 1024 || (nS & (nS - 1)) != 0) {
        vc_set_error("nSpatialSample1D^2 = %u must be a power of two in [32, 1024]", nS);
        return VC_E_UNSUPPORTED;
    }
    const uint32_t mS = q->maxSample1D * q->maxSample1D;
    if (mS < 1024 || mS % 1024 != 0 || 2 * mS * sizeof(float) > 64 * 1024) {
        vc_set_error("maxSample1D^2 = %u must be a multiple of 1024 and fit in LDS", mS);
        return VC_E_UNSUPPORTED;
    }
    if (q->nLGN == 0 || q->nKernelSample == 0) {
        vc_set_error("vc_lgn_precompute: empty LGN");
        return VC_E_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_lgn_tw, dim3(q->nLGN, 2), dim3(nS), 0, st, *q, TW);
    VC_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_lgn_sw, dim3(1), dim3(nS), 0, st, *q, SW);
    VC_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_lgn_sc, dim3(q->nLGN, 2), dim3(nS), 0, st, *q, SCx, SCy);
    VC_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_lgn_maxconvol, dim3(q->nLGN), dim3(1024), 2 * mS * sizeof(float), st,
                       *q, TW, maxConvol);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_lgn_frame_response_cells(void *stream, const float *frame, uint32_t width,
                                           uint32_t height, uint32_t nLGN, uint32_t c0,
                                           uint32_t c1, uint32_t nSample1D, const float *SCx,
                                           const float *SCy, const float *SW,
                                           const uint32_t *coneType, const uint32_t *cellTrial,
                                           uint64_t trialFrameStride, float *R) {
    const uint32_t nS = nSample1D * nSample1D;
    if (!frame || !SCx || !SCy || !SW || !coneType || !R) {
        vc_set_error("vc_lgn_frame_response: null argument");
        return VC_E_INVALID;
    }
    if (nS < 32 || nS > 1024 || (nS & (nS - 1)) != 0) {
        vc_set_error("nSpatialSample1D^2 = %u must be a power of two in [32, 1024]", nS);
        return VC_E_UNSUPPORTED;
    }
    if (width == 0 || height == 0 || nLGN == 0) {
        vc_set_error("vc_lgn_frame_response: empty frame or LGN");
        return VC_E_INVALID;
    }
    if (c0 > c1 || c1 > nLGN) {
        vc_set_error("vc_lgn_frame_response: cells %u..%u outside 0..%u", c0, c1, nLGN);
        return VC_E_INVALID;
    }
    VC_HIP_CHECK(launch_frame_response((hipStream_t)stream, nS, 1u, frame, width, height, nLGN,
                                       c0, c1, SCx, SCy, SW, coneType, cellTrial,
                                       trialFrameStride, (uint64_t)0, R));
    return VC_OK;
}

extern "C" int vc_lgn_frame_response(void *stream, const float *frame, uint32_t width,
                                     uint32_t height, uint32_t nLGN, uint32_t nSample1D,
                                     const float *SCx, const float *SCy, const float *SW,
                                     const uint32_t *coneType, const uint32_t *cellTrial,
                                     uint64_t trialFrameStride, float *R) {
    return vc_lgn_frame_response_cells(stream, frame, width, height, nLGN, 0u, nLGN, nSample1D,
                                       SCx, SCy, SW, coneType, cellTrial, trialFrameStride, R);
}

extern "C" int vc_lgn_sample_geometry(void *stream, const vc_lgn_setup *q, float *geom) {
    if (!q || !geom || !q->rw || !q->rh || !q->polar || !q->ecc || !q->orient) {
        vc_set_error("vc_lgn_sample_geometry: null argument");
        return VC_E_INVALID;
    }
    if (q->nLGN == 0 || q->nSample1D == 0) {
        vc_set_error("vc_lgn_sample_geometry: empty LGN or sample grid");
        return VC_E_INVALID;
    }
    hipLaunchKernelGGL(k_lgn_geom, dim3((2u * q->nLGN + 255u) / 256u), dim3(256), 0,
                       (hipStream_t)stream, *q, geom);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_lgn_frame_response_geom(void *stream, const float *frame, uint32_t width,
                                          uint32_t height, uint32_t nLGN, uint32_t c0,
                                          uint32_t c1, uint32_t nSample1D, const float *geom,
                                          float x0, float y0, float normViewDistance,
                                          const float *SW, const uint32_t *coneType,
                                          const uint32_t *cellTrial, uint64_t trialFrameStride,
                                          float *R) {
    const uint32_t nS = nSample1D * nSample1D;
    if (!frame || !geom || !SW || !coneType || !R) {
        vc_set_error("vc_lgn_frame_response_geom: null argument");
        return VC_E_INVALID;
    }
    if (nS != 64 && nS != 256 && nS != 1024) {
        vc_set_error("vc_lgn_frame_response_geom: nSpatialSample1D %u must be 8, 16 or 32",
                     nSample1D);
        return VC_E_UNSUPPORTED;
    }
    if (width == 0 || height == 0 || nLGN == 0 || c0 > c1 || c1 > nLGN) {
        vc_set_error("vc_lgn_frame_response_geom: empty frame or LGN, or cells %u..%u outside "
                     "0..%u", c0, c1, nLGN);
        return VC_E_INVALID;
    }
    const GeomArgs ga{geom, x0, y0, normViewDistance, nSample1D};
    VC_HIP_CHECK(launch_frame_response((hipStream_t)stream, nS, 1u, frame, width, height, nLGN,
                                       c0, c1, nullptr, nullptr, SW, coneType, cellTrial,
                                       trialFrameStride, (uint64_t)0, R, &ga));
    return VC_OK;
}

extern "C" int vc_lgn_frame_responses(void *stream, const float *frames, uint32_t nFrames,
                                      uint32_t width, uint32_t height, uint32_t nLGN,
                                      uint32_t nSample1D, const float *SCx, const float *SCy,
                                      const float *SW, const uint32_t *coneType, float *R) {
    const uint32_t nS = nSample1D * nSample1D;
    if (!frames || !SCx || !SCy || !SW || !coneType || !R) {
        vc_set_error("vc_lgn_frame_responses: null argument");
        return VC_E_INVALID;
    }
    if (nS < 32 || nS > 1024 || (nS & (nS - 1)) != 0) {
        vc_set_error("nSpatialSample1D^2 = %u must be a power of two in [32, 1024]", nS);
        return VC_E_UNSUPPORTED;
    }
    if (width == 0 || height == 0 || nLGN == 0 || nFrames > 65535u) {
        vc_set_error("vc_lgn_frame_responses: empty frame or LGN, or more than 65535 frames");
        return VC_E_INVALID;
    }
    if (nFrames == 0) return VC_OK;
    const uint64_t fs = 3ull * width * height;
    VC_HIP_CHECK(launch_frame_response((hipStream_t)stream, nS, nFrames, frames, width, height,
                                       nLGN, 0u, nLGN, SCx, SCy, SW, coneType, nullptr,
                                       (uint64_t)0, fs, R));
    return VC_OK;
}

namespace {
int lgn_steps(void *stream, const vc_lgn_params *p, const vc_lgn_state *s, const float *Rring,
              uint32_t it0, uint32_t nsteps, float *sInfo, float *fr, const vc_ff_events *ev,
              LgnTrace lt) {
    if (!p || !s || !Rring || !sInfo) {
        vc_set_error("vc_lgn_steps: null argument");
        return VC_E_INVALID;
    }
    if (p->nKernelSample == 0 || p->nKernelSample > (uint32_t)kMaxTaps) {
        vc_set_error("nKernelSample %u unsupported (1..%d)", p->nKernelSample, kMaxTaps);
        return VC_E_UNSUPPORTED;
    }
    if (p->maxFrame == 0 || p->ntPerFrame == 0) {
        vc_set_error("maxFrame %u / ntPerFrame %u must be positive", p->maxFrame, p->ntPerFrame);
        return VC_E_INVALID;
    }
    const uint32_t den = p->denorm ? p->denorm : 1u;
    // one wavefront covers 64 steps: its taps span at most kMaxFrame + 1 frames
    if ((63u + p->nRetrace) * den / p->ntPerFrame + 1u > (uint32_t)kMaxFrame) {
        vc_set_error("a %u-step window over frames of %u/%u steps spans too many frames",
                     p->nRetrace, p->ntPerFrame, den);
        return VC_E_UNSUPPORTED;
    }
    // the window spans nRetrace steps; every tap's frame must be within maxFrame of the first
    if ((p->nRetrace * den + p->ntPerFrame - 1) / p->ntPerFrame + 1 > p->maxFrame) {
        vc_set_error("maxFrame %u too small for a %u-step window", p->maxFrame, p->nRetrace);
        return VC_E_INVALID;
    }
    if (!fr) {
        vc_set_error("vc_lgn_steps: fr (the [nsteps][nLGN] rate buffer) is required");
        return VC_E_INVALID;
    }
    if (nsteps > p->ntPerFrame - it0 % p->ntPerFrame) {
        vc_set_error("vc_lgn_steps: steps %u..%u cross a period boundary (%u steps)", it0,
                     it0 + nsteps - 1, p->ntPerFrame);
        return VC_E_INVALID;
    }
    if (nsteps == 0 || p->nLGN == 0) return VC_OK;
    if (nsteps <= kShortRun && p->nKernelSample == 40) {
        hipLaunchKernelGGL(k_lgn_rate_cells<40>,
                           dim3((p->nLGN + 63u) / 64u, (nsteps + kStepsPerWave - 1u) / kStepsPerWave),
                           dim3(64), 0, (hipStream_t)stream, *p, *s, Rring, it0, nsteps, fr);
    } else {
        const uint64_t waves = (uint64_t)p->nLGN * ((nsteps + 63u) / 64u);
        launch_rate((hipStream_t)stream, waves, *p, *s, Rring, it0, nsteps, fr);
    }
    VC_HIP_CHECK(hipGetLastError());
    vc_ff_events e{};
    if (ev) {
        if (!ev->revOff || !ev->revTgt || !ev->mask || ev->nV1 == 0) {
            vc_set_error("vc_lgn_steps: incomplete vc_ff_events");
            return VC_E_INVALID;
        }
        e = *ev;
    }
    hipLaunchKernelGGL(k_lgn_spike, dim3((p->nLGN + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, *p, *s, it0, nsteps, fr, sInfo, e, lt);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}
}  // namespace

extern "C" int vc_lgn_steps(void *stream, const vc_lgn_params *p, const vc_lgn_state *s,
                            const float *Rring, uint32_t it0, uint32_t nsteps, float *sInfo,
                            float *fr, const vc_ff_events *ev) {
    return lgn_steps(stream, p, s, Rring, it0, nsteps, sInfo, fr, ev, LgnTrace{});
}

extern "C" int vc_lgn_steps_learn(void *stream, const vc_lgn_params *p, const vc_lgn_state *s,
                                  const float *Rring, uint32_t it0, uint32_t nsteps,
                                  float *sInfo, float *fr, const vc_ff_events *ev, float tauLTP,
                                  float *xVar, float *xStart, float *xSpike,
                                  uint64_t *spikeBits) {
    if (!xVar || !xStart || !xSpike || !spikeBits) {
        vc_set_error("vc_lgn_steps_learn: null trace buffer");
        return VC_E_INVALID;
    }
    if (!(tauLTP > 0.0f)) {
        vc_set_error("vc_lgn_steps_learn: tauLTP %g must be positive", (double)tauLTP);
        return VC_E_INVALID;
    }
    return lgn_steps(stream, p, s, Rring, it0, nsteps, sInfo, fr, ev,
                     LgnTrace{tauLTP, xVar, xStart, xSpike, spikeBits});
}

extern "C" int vc_lgn_rates(void *stream, const vc_lgn_params *p, const vc_lgn_state *s,
                            const float *R, uint32_t it0, uint32_t nsteps, float *fr) {
    if (!p || !s || !R || !fr) {
        vc_set_error("vc_lgn_rates: null argument");
        return VC_E_INVALID;
    }
    if (p->nKernelSample == 0 || p->nKernelSample > (uint32_t)kMaxTaps) {
        vc_set_error("nKernelSample %u unsupported (1..%d)", p->nKernelSample, kMaxTaps);
        return VC_E_UNSUPPORTED;
    }
    if (p->maxFrame == 0 || p->ntPerFrame == 0) {
        vc_set_error("vc_lgn_rates: maxFrame and ntPerFrame must be positive");
        return VC_E_INVALID;
    }
    if ((63u + p->nRetrace) * (p->denorm ? p->denorm : 1u) / p->ntPerFrame + 1u >
        (uint32_t)kMaxFrame) {
        vc_set_error("a %u-step window over frames of %u/%u steps spans too many frames",
                     p->nRetrace, p->ntPerFrame, p->denorm);
        return VC_E_UNSUPPORTED;
    }
    if (nsteps == 0 || p->nLGN == 0) return VC_OK;
    const uint64_t waves = (uint64_t)p->nLGN * ((nsteps + 63u) / 64u);
    if (waves * 64 / 256 + 1 > 0x7fffffffull) {
        vc_set_error("vc_lgn_rates: %u steps x %u cells is too large for one launch", nsteps,
                     p->nLGN);
        return VC_E_INVALID;
    }
    launch_rate((hipStream_t)stream, waves, *p, *s, R, it0, nsteps, fr);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_lgn_temporal_weights(void *stream, const vc_lgn_setup *q, float *TW) {
    if (!q || !TW) {
        vc_set_error("vc_lgn_temporal_weights: null argument");
        return VC_E_INVALID;
    }
    const uint32_t nS = q->nSample1D * q->nSample1D;
    if (nS < 32 || nS > 1024 || (nS & (nS - 1)) != 0) {
        vc_set_error("nSpatialSample1D^2 = %u must be a power of two in [32, 1024]", nS);
        return VC_E_UNSUPPORTED;
    }
    if (q->nLGN == 0 || q->nKernelSample == 0) {
        vc_set_error("vc_lgn_temporal_weights: empty LGN");
        return VC_E_INVALID;
    }
    hipLaunchKernelGGL(k_lgn_tw, dim3(q->nLGN, 2), dim3(nS), 0, (hipStream_t)stream, *q, TW);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}

extern "C" int vc_lgn_init(void *stream, const vc_lgn_params *p, float *leftTimeRate,
                           float *lastNegLogRand) {
    if (!p || !leftTimeRate || !lastNegLogRand) {
        vc_set_error("vc_lgn_init: null argument");
        return VC_E_INVALID;
    }
    if (p->nLGN == 0) return VC_OK;
    hipLaunchKernelGGL(k_lgn_init, dim3((p->nLGN + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       *p, leftTimeRate, lastNegLogRand);
    VC_HIP_CHECK(hipGetLastError());
    return VC_OK;
}
