/*
 * vc_oracle_v1grad.c -- float64 CPU reference of the V1 step of the
 * conductance LIF model and its gradient with respect to every synaptic
 * weight, at a fixed spike pattern.
 *
 * TEST INFRASTRUCTURE ONLY (part of oracle/libvc_oracle.so; called by
 * tests/ only; the product never links or calls it).
 *
 * The forward restates, in float64, the per-step equations of
 * src/coredynamics.cu:1367-1589 (decay, LGN and background inputs, refractory
 * back-up, Heun RK2 of src/coredynamics.cuh:27-118 with deltaT = 0, spike
 * interpolation), :1994-2046 (unit increments og/oh of src/condShape.h:13-41)
 * and :2609-2922 (recurrent sums over a dense conMat[b][ib][pre][post]).
 * The spike decision of step it for neuron n is taken from `pattern` when it
 * is given (a fixed spike pattern), else v > vThres.
 *
 * The gradient is NOT a hand-derived adjoint: every arithmetic operation of
 * the forward is recorded on a tape (each node keeps its two parents and the
 * partial derivatives with respect to them) and the tape is swept backwards
 * once (reverse-mode automatic differentiation).  It is therefore an
 * independent check of the GPU's hand-written reverse sweep; the tape itself
 * is pinned by central finite differences (tests/test_oracle_v1grad.py).
 *
 * Every value also carries a tangent (forward-mode automatic differentiation
 * with dual numbers): given a direction in weight space, the same forward
 * returns the directional derivative of every output.  With an anchor (the
 * float32 trajectory's state at the start of every step), the primal state
 * is reset to it at each step, so the tangent is the derivative of the step
 * map at exactly that trajectory -- the limit of finite differences at the
 * fixed spike pattern without the drift of a float64 trajectory forced onto
 * a float32 spike pattern over a long run.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "../include/vc_math.h"
#include "../include/vc_abi.h"

/* ------------------------------------------------------------- the tape -- */
typedef struct {
    double v;
    int64_t id;      /* node index, or -1 for a constant */
    double d;        /* tangent: directional derivative */
} num;

typedef struct {
    int64_t a, b;
    double da, db;
} node;

static _Thread_local struct {     /* one tape per calling thread */
    node *nd;
    int64_t n, cap;
    int on;
    int failed;
} tp;

static int64_t push(int64_t a, double da, int64_t b, double db) {
    if (!tp.on || (a < 0 && b < 0 && a != -2)) return -1;
    if (tp.n == tp.cap) {
        int64_t nc = tp.cap ? tp.cap * 2 : (1 << 20);
        node *p = realloc(tp.nd, sizeof(node) * nc);
        if (!p) { tp.failed = 1; return -1; }
        tp.nd = p;
        tp.cap = nc;
    }
    tp.nd[tp.n] = (node){a, b, da, db};
    return tp.n++;
}
static inline num K(double v) { num r = {v, -1, 0.0}; return r; }
static inline num leaf(double v) {
    num r = {v, -1, 0.0};
    if (tp.on) r.id = push(-2, 0.0, -1, 0.0);   /* a = -2 marks a leaf */
    return r;
}
static inline num add(num x, num y) {
    num r = {x.v + y.v, push(x.id, 1.0, y.id, 1.0), x.d + y.d};
    return r;
}
static inline num sub(num x, num y) {
    num r = {x.v - y.v, push(x.id, 1.0, y.id, -1.0), x.d - y.d};
    return r;
}
static inline num mul(num x, num y) {
    num r = {x.v * y.v, push(x.id, y.v, y.id, x.v), x.d * y.v + x.v * y.d};
    return r;
}
static inline num dvd(num x, num y) {
    const double q = x.v / y.v;
    num r = {q, push(x.id, 1.0 / y.v, y.id, -x.v / (y.v * y.v)), (x.d - q * y.d) / y.v};
    return r;
}
static inline num ex(num x) {
    double e = exp(x.v);
    num r = {e, push(x.id, e, -1, 0.0), e * x.d};
    return r;
}
static inline num scl(num x, double c) { num r = {x.v * c, push(x.id, c, -1, 0.0), x.d * c}; return r; }

/* ------------------------------------------------------------ the model -- */
#define NG VC_MAX_NG

typedef struct {
    num v, tB, gF[NG], hF[NG], ge[NG], he[NG], gi[NG], hi[NG];
} st64;

/* g, h after an input f at lag ddt (single_input) */
static void input(num *g, num *h, num f, num ddt, double rise, double decay, double dod) {
    num er = ex(scl(ddt, -1.0 / rise));
    num ed = ex(scl(ddt, -1.0 / decay));
    *g = add(*g, mul(scl(f, dod), sub(ed, er)));
    *h = add(*h, mul(f, er));
}
/* decay of (g, h) over ddt (decay_var) */
static void decay(num *g, num *h, num ddt, double rise, double decay_, double dod) {
    num er = ex(scl(ddt, -1.0 / rise));
    num ed = ex(scl(ddt, -1.0 / decay_));
    num ng = add(mul(*g, ed), mul(scl(*h, dod), sub(ed, er)));
    *h = mul(*h, er);
    *g = ng;
}

static int type_of64(const vc_v1_params *p, uint32_t local) {
    int t = (int)p->nType - 1;
    for (int j = (int)p->nType - 1; j >= 0; --j)
        if (local < p->typeAcc[j]) t = j;
    return t;
}

/*
 * Run T steps.  init [VC_SNAP_ROWS][nV1] (v, tBack, gFF, hFF, gE, hE, gI, hI
 * rows as in a snapshot), tonicDep [nV1], wstate [nV1] (the constant w of the
 * LIF model), sLGN [maxLGN][nV1] and conMat [nblock][maxNear][bs][bs]
 * (float64 weights), lgnSInfo [T][nLGN].  Writes sInfo [T][nV1].  If gOut
 * ([T][nV1] dL/dsInfo) is given, the gradient is written to dsLGN and dConMat
 * (same shapes as the weights).  Tangent mode (no tape): uSLGN and uConMat
 * (NULL = zero) are a direction in weight space and dOut [T][nV1] receives
 * the directional derivative of every output; anchor [T][VC_SNAP_ROWS][nV1]
 * (may be NULL) resets the primal state of every step to a given trajectory
 * (rows as in a snapshot, the cortical conductances after the step's
 * recurrent increments).  Returns 0, or -1 if the tape ran out of memory,
 * -2 for an unsupported model.
 */
int orc_v1_lif64(const vc_v1_params *p, uint32_t T, const float *init, const float *tonicDep,
                 const float *wstate, uint32_t maxLGN, const uint32_t *nLGN_of,
                 const uint32_t *idLGN, const double *sLGN, const float *lgnSInfo,
                 uint32_t nLGN, uint32_t maxNear, const uint32_t *nNear, const uint32_t *nbr,
                 const double *conMat, const uint8_t *pattern, double *sInfo,
                 const double *gOut, double *dsLGN, double *dConMat,
                 const double *uSLGN, const double *uConMat, const float *anchor,
                 double *dOut) {
    const uint32_t nV1 = p->nV1, bs = p->blockSize, nb = p->nblock;
    const double dt = p->dt;
    if (p->iModel != 0) return -2;
    for (uint32_t t = 0; t < p->nType; ++t)
        if (p->a[t] != 0.0f || p->deltaT[t] != 0.0f) return -2;
    tp.n = 0;
    tp.on = gOut != NULL;
    tp.failed = 0;
    /* weights: leaves */
    const size_t nW = (size_t)nb * maxNear * bs * bs;
    /* tape leaves of the weights (only with a tape: without one they are
     * plain constants read straight from conMat and sLGN) */
    num *W = NULL, *F = NULL;
    if (tp.on) {
        W = malloc(sizeof(num) * nW);
        F = malloc(sizeof(num) * (size_t)maxLGN * nV1);
        for (size_t i = 0; i < nW; ++i) W[i] = conMat[i] != 0.0 ? leaf(conMat[i]) : K(0.0);
        for (size_t i = 0; i < (size_t)maxLGN * nV1; ++i) F[i] = leaf(sLGN[i]);
    }
    st64 *S = malloc(sizeof(st64) * nV1);
    for (uint32_t n = 0; n < nV1; ++n) {
        S[n].v = K(init[VC_SNAP_V * nV1 + n]);
        S[n].tB = K(init[VC_SNAP_TBACK * nV1 + n]);
        for (int ig = 0; ig < NG; ++ig) {
            S[n].gF[ig] = K(init[(VC_SNAP_GFF + ig) * nV1 + n]);
            S[n].hF[ig] = K(init[(VC_SNAP_HFF + ig) * nV1 + n]);
            S[n].ge[ig] = K(init[(VC_SNAP_GE + ig) * nV1 + n]);
            S[n].he[ig] = K(init[(VC_SNAP_HE + ig) * nV1 + n]);
            S[n].gi[ig] = K(init[(VC_SNAP_GI + ig) * nV1 + n]);
            S[n].hi[ig] = K(init[(VC_SNAP_HI + ig) * nV1 + n]);
        }
    }
    /* spikes of the previous step: og/oh per receptor, and the sInfo nodes */
    num *og = malloc(sizeof(num) * NG * nV1), *oh = malloc(sizeof(num) * NG * nV1);
    uint8_t *spk = calloc(nV1, 1), *spkNext = calloc(nV1, 1);
    int64_t *outId = malloc(sizeof(int64_t) * (size_t)T * nV1);
    num *ogN = malloc(sizeof(num) * NG * nV1), *ohN = malloc(sizeof(num) * NG * nV1);
    /* the previous step's spiking neurons of each block, ascending */
    uint32_t *spkList = malloc(sizeof(uint32_t) * nV1), *spkOff = malloc(sizeof(uint32_t) * (nb + 1));
    const int tape = tp.on;
    for (uint32_t it = 0; it < T; ++it) {
        const float *lgn = lgnSInfo + (size_t)it * nLGN;
        spkOff[0] = 0;
        for (uint32_t b = 0, k = 0; b < nb; ++b) {
            for (uint32_t pre = 0; pre < bs; ++pre)
                if (spk[b * bs + pre]) spkList[k++] = pre;
            spkOff[b + 1] = k;
        }
        /* without a tape every neuron's step is independent of the others' */
#pragma omp parallel for schedule(static) if (!tape)
        for (uint32_t n = 0; n < nV1; ++n) {
            const uint32_t b = n / bs, post = n % bs;
            const int t = type_of64(p, post);
            st64 s = S[n];
            /* recurrent increments of the previous step's spikes */
            for (uint32_t ib = 0; ib < nNear[b]; ++ib) {
                const uint32_t bid = nbr[b * maxNear + ib];
                for (uint32_t k = spkOff[bid]; k < spkOff[bid + 1]; ++k) {
                    const uint32_t pre = spkList[k];
                    const uint32_t pg = bid * bs + pre;
                    const size_t wi = (((size_t)b * maxNear + ib) * bs + pre) * bs + post;
                    num w = W ? W[wi] : K(conMat[wi]);
                    if (uConMat) w.d = uConMat[wi];
                    if (w.id < 0 && w.v == 0.0) continue;
                    if (pre < p->nE) {
                        for (uint32_t ig = 0; ig < p->ngE; ++ig) {
                            num sw = scl(w, p->pE[t * p->ngE + ig]);
                            s.ge[ig] = add(s.ge[ig], mul(sw, og[ig * nV1 + pg]));
                            s.he[ig] = add(s.he[ig], mul(sw, oh[ig * nV1 + pg]));
                        }
                    } else {
                        for (uint32_t ig = 0; ig < p->ngI; ++ig) {
                            num sw = scl(w, p->pI[t * p->ngI + ig]);
                            s.gi[ig] = add(s.gi[ig], mul(sw, og[ig * nV1 + pg]));
                            s.hi[ig] = add(s.hi[ig], mul(sw, oh[ig * nV1 + pg]));
                        }
                    }
                }
            }
            if (anchor) {          /* the primal state of the given trajectory */
                const float *a = anchor + (size_t)it * VC_SNAP_ROWS * nV1 + n;
                s.v.v = a[(size_t)VC_SNAP_V * nV1];
                s.tB.v = a[(size_t)VC_SNAP_TBACK * nV1];
                for (int ig = 0; ig < NG; ++ig) {
                    s.gF[ig].v = a[(size_t)(VC_SNAP_GFF + ig) * nV1];
                    s.hF[ig].v = a[(size_t)(VC_SNAP_HFF + ig) * nV1];
                    s.ge[ig].v = a[(size_t)(VC_SNAP_GE + ig) * nV1];
                    s.he[ig].v = a[(size_t)(VC_SNAP_HE + ig) * nV1];
                    s.gi[ig].v = a[(size_t)(VC_SNAP_GI + ig) * nV1];
                    s.hi[ig].v = a[(size_t)(VC_SNAP_HI + ig) * nV1];
                }
            }
            const double vThres = p->vThres[t], gL = p->gL[t], Cm = p->C[t], vT = p->vT[t];
            const double tau_w = p->tau_w[t];
            const double targetV = -65.0 + (vT - (-65.0)) * tonicDep[n];
            const double depC = gL * (targetV - (-65.0));
            const num dtK = K(dt);
            /* cortical conductances at t + dt */
            num gE1 = K(0.0), gI1 = K(0.0), gE0 = K(0.0), gI0 = K(0.0);
            st64 o = s;
            for (uint32_t ig = 0; ig < p->ngE; ++ig) {
                decay(&o.ge[ig], &o.he[ig], dtK, p->riseE[ig], p->decayE[ig], p->dodE[ig]);
                gE1 = add(gE1, o.ge[ig]);
            }
            for (uint32_t ig = 0; ig < p->ngI; ++ig) {
                decay(&o.gi[ig], &o.hi[ig], dtK, p->riseI[ig], p->decayI[ig], p->dodI[ig]);
                gI1 = add(gI1, o.gi[ig]);
            }
            const int backing = s.tB.v < dt && s.tB.v > 0.0;
            num g0[NG], h0[NG];
            for (uint32_t ig = 0; ig < p->ngFF; ++ig) {
                decay(&o.gF[ig], &o.hF[ig], dtK, p->riseFF[ig], p->decayFF[ig], p->dodFF[ig]);
                g0[ig] = s.gF[ig];
                h0[ig] = s.hF[ig];
                if (backing)
                    decay(&g0[ig], &h0[ig], s.tB, p->riseFF[ig], p->decayFF[ig], p->dodFF[ig]);
            }
            /* LGN inputs in list order, then background */
            for (uint32_t j = 0; j < nLGN_of[n]; ++j) {
                const size_t lid = (size_t)j * nV1 + n;
                const float sI = lgn[idLGN[lid]];
                if (!(sI >= 1.0f)) continue;
                const double nsp = floor((double)sI), tsp = ((double)sI - nsp) * dt;
                for (uint32_t ig = 0; ig < p->ngFF; ++ig) {
                    num fw = F ? F[lid] : K(sLGN[lid]);
                    if (uSLGN) fw.d = uSLGN[lid];
                    num f = scl(fw, p->pFF[t * p->ngFF + ig] * nsp);
                    input(&o.gF[ig], &o.hF[ig], f, K(dt - tsp), p->riseFF[ig], p->decayFF[ig],
                          p->dodFF[ig]);
                    if (backing && s.tB.v - tsp > 0.0)
                        input(&g0[ig], &h0[ig], f, sub(s.tB, K(tsp)), p->riseFF[ig],
                              p->decayFF[ig], p->dodFF[ig]);
                }
            }
            if (p->bgStrength[t] > 0.0f) {
                const vc_u32x4 r = vc_draw4(p->seed_lo, p->seed_hi, n, it, VC_STREAM_V1_BG, 0u,
                                            p->trial);
                const float u = vc_u01_open0(r.v[0]);
                const float frac = vc_u01_open1(r.v[1]);
                const uint32_t nsp = vc_poisson_inv(u, p->bgLambda[t], p->bgP0[t]);
                if (nsp > 0) {
                    const double tsp = (double)frac * dt;
                    for (uint32_t ig = 0; ig < p->ngFF; ++ig) {
                        num f = K((double)p->bgStrength[t] * p->pBG[t * p->ngFF + ig] * nsp);
                        input(&o.gF[ig], &o.hF[ig], f, K(dt - tsp), p->riseFF[ig],
                              p->decayFF[ig], p->dodFF[ig]);
                        if (backing && s.tB.v - tsp > 0.0)
                            input(&g0[ig], &h0[ig], f, sub(s.tB, K(tsp)), p->riseFF[ig],
                                  p->decayFF[ig], p->dodFF[ig]);
                    }
                }
            }
            for (uint32_t ig = 0; ig < p->ngFF; ++ig) {
                gE0 = add(gE0, g0[ig]);
                gE1 = add(gE1, o.gF[ig]);
            }
            for (uint32_t ig = 0; ig < p->ngE; ++ig) {
                num g = s.ge[ig], h = s.he[ig];
                if (backing) decay(&g, &h, s.tB, p->riseE[ig], p->decayE[ig], p->dodE[ig]);
                gE0 = add(gE0, g);
            }
            for (uint32_t ig = 0; ig < p->ngI; ++ig) {
                num g = s.gi[ig], h = s.hi[ig];
                if (backing) decay(&g, &h, s.tB, p->riseI[ig], p->decayI[ig], p->dodI[ig]);
                gI0 = add(gI0, g);
            }
            /* membrane */
            num vNew = K(p->vR[t]), tBNew = K(0.0), out;
            int fired = 0;
            if (s.tB.v < dt) {
                num a0 = add(add(gE0, gI0), K(gL));
                num b0 = add(add(scl(gE0, 0.0), scl(gI0, -75.0)), K(gL * -65.0 + depC));
                num a1 = add(add(gE1, gI1), K(gL));
                num b1 = add(add(scl(gE1, 0.0), scl(gI1, -75.0)), K(gL * -65.0 + depC));
                num h = sub(dtK, s.tB);
                num w0 = K(wstate[n]);
                if (backing) w0 = mul(w0, ex(scl(s.tB, -1.0 / tau_w)));
                num fk1 = scl(sub(add(scl(mul(a0, s.v), -1.0), b0), w0), 1.0 / Cm);
                num gk1 = scl(w0, -1.0 / tau_w);
                num v1 = add(s.v, mul(h, fk1));
                num w1 = add(w0, mul(h, gk1));
                num fk2 = scl(sub(add(scl(mul(a1, v1), -1.0), b1), w1), 1.0 / Cm);
                num v = add(s.v, scl(mul(h, add(fk1, fk2)), 0.5));
                fired = pattern ? pattern[(size_t)it * nV1 + n] != 0 : (v.v > vThres);
                if (fired) {
                    num ratio = dvd(sub(K(vThres), s.v), sub(v, s.v));
                    num tsp = add(s.tB, mul(ratio, h));
                    tBNew = sub(add(tsp, K(p->tRef[t])), dtK);
                    out = add(K(1.0), scl(tsp, 1.0 / dt));
                    /* unit increments sent to the targets at the next step */
                    num lag = sub(dtK, tsp);
                    const int isE = post < p->nE;
                    const uint32_t ng = isE ? p->ngE : p->ngI;
                    for (uint32_t ig = 0; ig < ng; ++ig) {
                        num g = K(0.0), hh = K(0.0);
                        input(&g, &hh, K(1.0), lag, isE ? p->riseE[ig] : p->riseI[ig],
                              isE ? p->decayE[ig] : p->decayI[ig], isE ? p->dodE[ig] : p->dodI[ig]);
                        ogN[ig * nV1 + n] = g;
                        ohN[ig * nV1 + n] = hh;
                    }
                } else {
                    vNew = v;
                    out = v;
                }
            } else {
                tBNew = sub(s.tB, dtK);
                out = vNew;
            }
            spkNext[n] = (uint8_t)fired;
            sInfo[(size_t)it * nV1 + n] = out.v;
            if (dOut) dOut[(size_t)it * nV1 + n] = out.d;
            outId[(size_t)it * nV1 + n] = out.id;
            o.v = vNew;
            o.tB = tBNew;
            S[n] = o;
        }
        /* this step's spikes feed the next step */
        memcpy(spk, spkNext, nV1);
        num *t1 = og; og = ogN; ogN = t1;
        num *t2 = oh; oh = ohN; ohN = t2;
    }
    int rc = tp.failed ? -1 : 0;
    if (gOut && !tp.failed) {
        double *adj = calloc(tp.n, sizeof(double));
        for (size_t i = 0; i < (size_t)T * nV1; ++i)
            if (outId[i] >= 0) adj[outId[i]] += gOut[i];
        for (int64_t i = tp.n - 1; i >= 0; --i) {
            const node *q = &tp.nd[i];
            const double g = adj[i];
            if (g == 0.0 || q->a == -2) continue;
            if (q->a >= 0) adj[q->a] += g * q->da;
            if (q->b >= 0) adj[q->b] += g * q->db;
        }
        for (size_t i = 0; i < nW; ++i) dConMat[i] = W[i].id >= 0 ? adj[W[i].id] : 0.0;
        for (size_t i = 0; i < (size_t)maxLGN * nV1; ++i) dsLGN[i] = F[i].id >= 0 ? adj[F[i].id] : 0.0;
        free(adj);
    }
    free(W); free(F); free(S); free(og); free(oh); free(ogN); free(ohN);
    free(spk); free(spkNext); free(outId); free(spkList); free(spkOff);
    free(tp.nd);
    tp.nd = NULL;
    tp.n = tp.cap = 0;
    return rc;
}
