/*
 * vc_oracle_grad.c -- float64 CPU reference of the LGN rate loss and its
 * gradient.
 *
 * TEST INFRASTRUCTURE ONLY (part of oracle/libvc_oracle.so, called only by
 * tests/ and bench.py's checks; the product never links or calls it).
 *
 * Forward (float64 throughout, from float32 inputs), restating the
 * reference's parvo LGN convolution and rate (src/discrete_input_convol.cu
 * :14-33 temporal kernel, :230-305 temporal weights, :35-76 intensity,
 * :1003-1069 and :1143-1211 surround mean / contrast / spatial sum,
 * :1212-1340 tap sum, :1637-1680 rate; src/LGN_props.cuh:94-100):
 *   TW_r[j]  = v_j / sum_k |v_k|,  v_j = G(t_j; nR, tauR) - ratio G(t_j; nD, tauD)
 *   R_r[F]   = sum_s SW_s clip(I_r(x_s)/mean_r - 1, +-1),  mean over surround samples
 *   conv_r   = sum_i TW_r[i] R_r[F(it, i)],  F(it,i) = floor((it + iT0 - nRetrace + 1
 *              + i interval) / ntPerFrame); frames F < 0 are the pre-stimulus frame
 *   fr       = frRatio max (A / (1 + exp(K (C50 - x))) + B),
 *              x = max(0, convolRatio (kC conv_C + kS conv_S)) / max
 *   loss     = sum_{it,c} (fr - target)^2 / (T nLGN)
 * Parameters theta[c][6] = (kC, kS, tauR_C, tauD_C, tauR_S, tauD_S).
 *
 * Gradient: a direct reverse sweep over every (cell, step, tap) and every
 * (cell, frame, sample) -- no reorganisation of the sums -- so that it is an
 * independent check of the GPU adjoint's segment-sum formulation.  It is
 * pinned by central finite differences of orc_lgn64_loss
 * (tests/test_oracle_grad.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint32_t nLGN, nS, nKS, interval, iT0, nRetrace, npf, nF, T, W, H;
    float dt, convolRatio, frRatio;
    const float *SCx, *SCy, *SW;                 /* [nLGN][2][nS], [nS]   */
    const uint32_t *coneType;                    /* [2][nLGN]             */
    const float *delay, *ratio, *nR, *nD;        /* [2][nLGN]             */
    const float *maxConvol, *c50, *sharpness, *logA, *logB;   /* [nLGN]   */
    const double *initFrame;                     /* [3][H][W]             */
} orc_lgn64;

static inline double tex64(const double *img, int W, int H, int i, int j) {
    if (i < 0 || j < 0 || i >= W || j >= H) return 0.0;
    return img[(size_t)j * W + i];
}

/* bilinear sample with border addressing (the texture fetch of
 * src/discrete_input_convol.cu:35-76); w[4], idx[4] receive the taps */
static double bil64(const double *img, int W, int H, float x, float y, double w[4],
                    long idx[4]) {
    for (int k = 0; k < 4; ++k) { w[k] = 0.0; idx[k] = -1; }
    double xt = (double)x * W - 0.5, yt = (double)y * H - 0.5;
    if (!(xt > -2.0 && xt < W + 1.0 && yt > -2.0 && yt < H + 1.0)) return 0.0;
    double fx = floor(xt), fy = floor(yt);
    double ax = xt - fx, ay = yt - fy;
    int i0 = (int)fx, j0 = (int)fy;
    int ii[4] = {i0, i0 + 1, i0, i0 + 1}, jj[4] = {j0, j0, j0 + 1, j0 + 1};
    double ww[4] = {(1 - ax) * (1 - ay), ax * (1 - ay), (1 - ax) * ay, ax * ay};
    double v = 0.0;
    for (int k = 0; k < 4; ++k) {
        if (ii[k] < 0 || jj[k] < 0 || ii[k] >= W || jj[k] >= H) continue;
        w[k] = ww[k];
        idx[k] = (long)jj[k] * W + ii[k];
        v += ww[k] * img[idx[k]];
    }
    return v;
}

static const double kCone4[3] = {0.45436904, 0.47353315, 0.0720978};

/* intensity of cone type `cone` and, in (chw[3]), its weight on channel ch */
static double inten64(uint32_t cone, const double *frame, int W, int H, float x, float y,
                      double chw[3], double w[4], long idx[4]) {
    const size_t np = (size_t)W * H;
    chw[0] = chw[1] = chw[2] = 0.0;
    switch (cone) {
    case 0: chw[0] = 1; break;
    case 1: chw[1] = 1; break;
    case 2: chw[2] = 1; break;
    case 3: chw[0] = 0.5; chw[1] = 0.5; break;
    case 4: chw[0] = kCone4[0]; chw[1] = kCone4[1]; chw[2] = kCone4[2]; break;
    default: return 0.0;
    }
    double v = 0.0;
    for (int ch = 0; ch < 3; ++ch)
        if (chw[ch] != 0.0) v += chw[ch] * bil64(frame + ch * np, W, H, x, y, w, idx);
    return v;
}

/* normalised temporal weights of one (cell, region) */
static void tw64(const orc_lgn64 *m, uint32_t c, int r, double tauR, double tauD,
                 double *tw) {
    const uint32_t lid = r * m->nLGN + c;
    const float ksDt = (float)m->interval * m->dt;
    const float T0 = (float)m->iT0 * m->dt - m->delay[lid];
    const double nR = m->nR[lid], nD = m->nD[lid], ratio = m->ratio[lid];
    double S = 0.0;
    for (uint32_t j = 0; j < m->nKS; ++j) {
        const float t = (float)(m->nKS - 1 - j) * ksDt + T0;
        double v = 0.0;
        if (!(t < 0.0f)) {
            const double tt = t;
            const double gR = exp((nR - 1) * log(tt / tauR) - tt / tauR - lgamma(nR)) / tauR;
            const double gD = exp((nD - 1) * log(tt / tauD) - tt / tauD - lgamma(nD)) / tauD;
            v = gR - ratio * gD;
        }
        tw[j] = v;
        S += fabs(v);
    }
    for (uint32_t j = 0; j < m->nKS; ++j) tw[j] /= S;
}

/* spatial responses R[2] of cell c to `frame`; *clipped counts clipped samples */
static void resp64(const orc_lgn64 *m, const double *frame, uint32_t c, double R[2],
                   long *clipped) {
    const uint32_t nS = m->nS;
    const uint32_t tC = m->coneType[c], tS = m->coneType[m->nLGN + c];
    const float *xC = m->SCx + ((size_t)c * 2 + 0) * nS, *yC = m->SCy + ((size_t)c * 2 + 0) * nS;
    const float *xS = m->SCx + ((size_t)c * 2 + 1) * nS, *yS = m->SCy + ((size_t)c * 2 + 1) * nS;
    double chw[3], w[4];
    long idx[4];
    double mC = 0.0, mS = 0.0;
    for (uint32_t s = 0; s < nS; ++s) {
        mC += inten64(tC, frame, m->W, m->H, xS[s], yS[s], chw, w, idx);
        mS += inten64(tS, frame, m->W, m->H, xS[s], yS[s], chw, w, idx);
    }
    mC /= nS;
    mS /= nS;
    double rC = 0.0, rS = 0.0;
    for (uint32_t s = 0; s < nS; ++s) {
        const double IC = inten64(tC, frame, m->W, m->H, xC[s], yC[s], chw, w, idx);
        const double IS = inten64(tS, frame, m->W, m->H, xS[s], yS[s], chw, w, idx);
        double cc = mC > 0 ? IC / mC - 1.0 : IC;
        double cs = mS > 0 ? IS / mS - 1.0 : IS;
        if (fabs(cc) > 1.0) { cc = copysign(1.0, cc); if (clipped) ++*clipped; }
        if (fabs(cs) > 1.0) { cs = copysign(1.0, cs); if (clipped) ++*clipped; }
        rC += m->SW[s] * cc;
        rS += m->SW[s] * cs;
    }
    R[0] = rC;
    R[1] = rS;
}

static inline int64_t fdiv64(int64_t a, int64_t b) {
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
    return q;
}

/* per-cell rates: R64 [nF+1][2] (index 0 = pre-stimulus), tw [2][nKS] */
static double cell_rate(const orc_lgn64 *m, uint32_t c, const double *theta,
                        const double *R64, const double *tw, uint32_t it, double *convC,
                        double *convS, double *X, int *active) {
    const int64_t c0 = (int64_t)m->iT0 - (int64_t)m->nRetrace + 1;
    double a = 0.0, b = 0.0;
    for (uint32_t i = 0; i < m->nKS; ++i) {
        int64_t F = fdiv64((int64_t)it + c0 + (int64_t)i * m->interval, m->npf);
        const double *Rf = R64 + 2 * (F < 0 ? 0 : F + 1);
        a += tw[i] * Rf[0];
        b += tw[m->nKS + i] * Rf[1];
    }
    *convC = a;
    *convS = b;
    const double mx = m->maxConvol[c];
    double cur = (theta[0] * a + theta[1] * b) * m->convolRatio;
    *active = cur > 0.0;
    if (cur < 0.0) cur = 0.0;
    const double x = cur / mx;
    const double Xv = 1.0 / (1.0 + exp(m->sharpness[c] * (m->c50[c] - x)));
    *X = Xv;
    return m->frRatio * mx * (m->logA[c] * Xv + m->logB[c]);
}

/* The loss restricted to the cells in `cells` (all if NULL); *pattern gets
 * a signature of the activation pattern for finite-difference checks: per
 * cell an FNV-1a hash of the number of clipped samples of every frame, the
 * signs of the temporal weights (the normalisation by sum |TW| has a kink
 * where a weight changes sign) and the rectifier state of every step, summed
 * over the cells.  Two runs with equal signatures have, with overwhelming
 * probability, crossed no kink between them (a count alone would miss kinks
 * crossed in opposite directions). */
double orc_lgn64_loss(const orc_lgn64 *m, const double *frames, const double *theta,
                      const float *target, const int32_t *cells, int ncells, long *pattern) {
    const int nc = cells ? ncells : (int)m->nLGN;
    const size_t fsz = 3ull * m->W * m->H;
    double total = 0.0;
    unsigned long pat = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total, pat)
    for (int k = 0; k < nc; ++k) {
        const uint32_t c = cells ? (uint32_t)cells[k] : (uint32_t)k;
        const double *th = theta + (size_t)c * 6;
        double *R64 = malloc(sizeof(double) * 2 * (m->nF + 1));
        double *tw = malloc(sizeof(double) * 2 * m->nKS);
        uint64_t h = 1469598103934665603ull;
#define ORC_MIX(x) (h = (h ^ (uint64_t)(x)) * 1099511628211ull)
        long clipped = 0;
        resp64(m, m->initFrame, c, R64, &clipped);   /* the pre-stimulus frame is no variable */
        for (uint32_t f = 0; f < m->nF; ++f) {
            clipped = 0;
            resp64(m, frames + f * fsz, c, R64 + 2 * (f + 1), &clipped);
            ORC_MIX(clipped);
        }
        tw64(m, c, 0, th[2], th[3], tw);
        tw64(m, c, 1, th[4], th[5], tw + m->nKS);
        for (uint32_t j = 0; j < 2 * m->nKS; ++j) ORC_MIX(tw[j] < 0.0);
        double acc = 0.0;
        for (uint32_t it = 0; it < m->T; ++it) {
            double a, b, X;
            int on;
            const double fr = cell_rate(m, c, th, R64, tw, it, &a, &b, &X, &on);
            const double d = fr - (double)target[(size_t)it * m->nLGN + c];
            acc += d * d;
            ORC_MIX(on);
        }
#undef ORC_MIX
        total += acc;
        pat += h;
        free(R64);
        free(tw);
    }
    if (pattern) *pattern = (long)pat;
    return total / ((double)m->T * m->nLGN);
}

/* Rates fr[T][nLGN] (float64). */
void orc_lgn64_rates(const orc_lgn64 *m, const double *frames, const double *theta,
                     double *fr) {
    const size_t fsz = 3ull * m->W * m->H;
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < (int)m->nLGN; ++k) {
        const uint32_t c = (uint32_t)k;
        const double *th = theta + (size_t)c * 6;
        double *R64 = malloc(sizeof(double) * 2 * (m->nF + 1));
        double *tw = malloc(sizeof(double) * 2 * m->nKS);
        resp64(m, m->initFrame, c, R64, NULL);
        for (uint32_t f = 0; f < m->nF; ++f) resp64(m, frames + f * fsz, c, R64 + 2 * (f + 1), NULL);
        tw64(m, c, 0, th[2], th[3], tw);
        tw64(m, c, 1, th[4], th[5], tw + m->nKS);
        for (uint32_t it = 0; it < m->T; ++it) {
            double a, b, X;
            int on;
            fr[(size_t)it * m->nLGN + c] = cell_rate(m, c, th, R64, tw, it, &a, &b, &X, &on);
        }
        free(R64);
        free(tw);
    }
}

/* d(normalised TW_j)/d(tauR), d/d(tauD) of one (cell, region) */
static void tw64_grad(const orc_lgn64 *m, uint32_t c, int r, double tauR, double tauD,
                      double *dR, double *dD) {
    const uint32_t lid = r * m->nLGN + c;
    const float ksDt = (float)m->interval * m->dt;
    const float T0 = (float)m->iT0 * m->dt - m->delay[lid];
    const double nR = m->nR[lid], nD = m->nD[lid], ratio = m->ratio[lid];
    const uint32_t n = m->nKS;
    double *v = malloc(sizeof(double) * 3 * n);
    double *vR = v + n, *vD = v + 2 * n;
    double S = 0.0, SR = 0.0, SD = 0.0;
    for (uint32_t j = 0; j < n; ++j) {
        const float t = (float)(n - 1 - j) * ksDt + T0;
        v[j] = vR[j] = vD[j] = 0.0;
        if (!(t < 0.0f)) {
            const double tt = t;
            const double gR = exp((nR - 1) * log(tt / tauR) - tt / tauR - lgamma(nR)) / tauR;
            const double gD = exp((nD - 1) * log(tt / tauD) - tt / tauD - lgamma(nD)) / tauD;
            v[j] = gR - ratio * gD;
            /* d/dtau of the gamma density G(t; n, tau) = G (t/tau - n) / tau */
            vR[j] = gR * (tt / tauR - nR) / tauR;
            vD[j] = -ratio * gD * (tt / tauD - nD) / tauD;
        }
        const double sg = v[j] > 0 ? 1.0 : (v[j] < 0 ? -1.0 : 0.0);
        S += fabs(v[j]);
        SR += sg * vR[j];
        SD += sg * vD[j];
    }
    for (uint32_t j = 0; j < n; ++j) {
        dR[j] = vR[j] / S - v[j] * SR / (S * S);
        dD[j] = vD[j] / S - v[j] * SD / (S * S);
    }
    free(v);
}

/* Full gradient of orc_lgn64_loss (all cells): dframes [nF][3][H][W] and
 * dtheta [nLGN][6], both written. */
void orc_lgn64_grad(const orc_lgn64 *m, const double *frames, const double *theta,
                    const float *target, double *dframes, double *dtheta) {
    const size_t fsz = 3ull * m->W * m->H;
    const uint32_t nF = m->nF, nKS = m->nKS, nS = m->nS;
    const double N = (double)m->T * m->nLGN;
    const int64_t c0 = (int64_t)m->iT0 - (int64_t)m->nRetrace + 1;
    double *dRall = calloc((size_t)m->nLGN * 2 * nF, sizeof(double));   /* [nLGN][nF][2] */
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < (int)m->nLGN; ++k) {
        const uint32_t c = (uint32_t)k;
        const double *th = theta + (size_t)c * 6;
        double *R64 = malloc(sizeof(double) * 2 * (nF + 1));
        double *tw = malloc(sizeof(double) * 2 * nKS);
        double *dtw = calloc(2 * nKS, sizeof(double));
        double *dR = dRall + (size_t)c * 2 * nF;
        resp64(m, m->initFrame, c, R64, NULL);
        for (uint32_t f = 0; f < nF; ++f) resp64(m, frames + f * fsz, c, R64 + 2 * (f + 1), NULL);
        tw64(m, c, 0, th[2], th[3], tw);
        tw64(m, c, 1, th[4], th[5], tw + nKS);
        double dkC = 0.0, dkS = 0.0;
        for (uint32_t it = 0; it < m->T; ++it) {
            double a, b, X;
            int on;
            const double fr = cell_rate(m, c, th, R64, tw, it, &a, &b, &X, &on);
            if (!on) continue;
            const double gfr = 2.0 * (fr - (double)target[(size_t)it * m->nLGN + c]) / N;
            /* fr = frRatio mx (A X + B), X = sigma(K (x - C50)), x = cur / mx */
            const double gcur = gfr * m->frRatio * m->logA[c] * X * (1.0 - X) * m->sharpness[c];
            const double gconv = gcur * m->convolRatio;
            dkC += gconv * a;
            dkS += gconv * b;
            for (uint32_t i = 0; i < nKS; ++i) {
                int64_t F = fdiv64((int64_t)it + c0 + (int64_t)i * m->interval, m->npf);
                const double *Rf = R64 + 2 * (F < 0 ? 0 : F + 1);
                dtw[i] += gconv * th[0] * Rf[0];
                dtw[nKS + i] += gconv * th[1] * Rf[1];
                if (F >= 0) {
                    dR[2 * F + 0] += gconv * th[0] * tw[i];
                    dR[2 * F + 1] += gconv * th[1] * tw[nKS + i];
                }
            }
        }
        double *gR = malloc(sizeof(double) * 2 * nKS);
        double *gD = malloc(sizeof(double) * 2 * nKS);
        double out[6] = {dkC, dkS, 0, 0, 0, 0};
        for (int r = 0; r < 2; ++r) {
            tw64_grad(m, c, r, th[2 + 2 * r], th[3 + 2 * r], gR + r * nKS, gD + r * nKS);
            for (uint32_t j = 0; j < nKS; ++j) {
                out[2 + 2 * r] += dtw[r * nKS + j] * gR[r * nKS + j];
                out[3 + 2 * r] += dtw[r * nKS + j] * gD[r * nKS + j];
            }
        }
        memcpy(dtheta + (size_t)c * 6, out, sizeof(out));
        free(gR); free(gD); free(R64); free(tw); free(dtw);
    }
    /* frames: one frame per task, so no two tasks touch the same pixels */
    memset(dframes, 0, sizeof(double) * fsz * nF);
#pragma omp parallel for schedule(dynamic, 1)
    for (int f = 0; f < (int)nF; ++f) {
        const double *frame = frames + f * fsz;
        double *dframe = dframes + f * fsz;
        const size_t np = (size_t)m->W * m->H;
        double *IC = malloc(sizeof(double) * 3 * nS);
        double *ISs = IC + nS, *ICs = IC + 2 * nS;
        for (uint32_t c = 0; c < m->nLGN; ++c) {
            const double dRC = dRall[((size_t)c * nF + f) * 2 + 0];
            const double dRS = dRall[((size_t)c * nF + f) * 2 + 1];
            if (dRC == 0.0 && dRS == 0.0) continue;
            const uint32_t tC = m->coneType[c], tS = m->coneType[m->nLGN + c];
            const float *xC = m->SCx + ((size_t)c * 2) * nS, *yC = m->SCy + ((size_t)c * 2) * nS;
            const float *xS = m->SCx + ((size_t)c * 2 + 1) * nS, *yS = m->SCy + ((size_t)c * 2 + 1) * nS;
            double chw[3], w[4];
            long idx[4];
            double mC = 0.0, mS = 0.0;
            for (uint32_t s = 0; s < nS; ++s) {
                ICs[s] = inten64(tC, frame, m->W, m->H, xS[s], yS[s], chw, w, idx);
                ISs[s] = inten64(tS, frame, m->W, m->H, xS[s], yS[s], chw, w, idx);
                IC[s] = inten64(tC, frame, m->W, m->H, xC[s], yC[s], chw, w, idx);
                mC += ICs[s];
                mS += ISs[s];
            }
            mC /= nS;
            mS /= nS;
            /* R_C = sum_s SW clip(IC/mC - 1): d/dIC and d/dmC */
            double gmC = 0.0, gmS = 0.0;
            double *dIC = malloc(sizeof(double) * 2 * nS), *dIS = dIC + nS;
            for (uint32_t s = 0; s < nS; ++s) {
                double cc = mC > 0 ? IC[s] / mC - 1.0 : IC[s];
                double cs = mS > 0 ? ISs[s] / mS - 1.0 : ISs[s];
                const double pc = fabs(cc) > 1.0 ? 0.0 : 1.0, ps = fabs(cs) > 1.0 ? 0.0 : 1.0;
                dIC[s] = dRC * m->SW[s] * pc * (mC > 0 ? 1.0 / mC : 1.0);
                dIS[s] = dRS * m->SW[s] * ps * (mS > 0 ? 1.0 / mS : 1.0);
                if (mC > 0) gmC += dRC * m->SW[s] * pc * (-IC[s] / (mC * mC));
                if (mS > 0) gmS += dRS * m->SW[s] * ps * (-ISs[s] / (mS * mS));
            }
            for (uint32_t s = 0; s < nS; ++s) {
                /* centre sample, surround sample (centre cone: mean), surround sample */
                const float px[3] = {xC[s], xS[s], xS[s]}, py[3] = {yC[s], yS[s], yS[s]};
                const uint32_t cone[3] = {tC, tC, tS};
                const double d[3] = {dIC[s], gmC / nS, dIS[s] + gmS / nS};
                for (int q = 0; q < 3; ++q) {
                    if (d[q] == 0.0) continue;
                    inten64(cone[q], frame, m->W, m->H, px[q], py[q], chw, w, idx);
                    for (int ch = 0; ch < 3; ++ch) {
                        if (chw[ch] == 0.0) continue;
                        for (int t = 0; t < 4; ++t)
                            if (idx[t] >= 0) dframe[ch * np + idx[t]] += d[q] * chw[ch] * w[t];
                    }
                }
            }
            free(dIC);
        }
        free(IC);
    }
    free(dRall);
}
