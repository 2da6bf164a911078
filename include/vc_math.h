/*
 * vc_math.h -- deterministic scalar math shared by the HIP kernels and the
 * CPU oracle of the V1 simulation step.
 *
 * Why this exists: the simulation step is chaotic (a one-ulp change in a
 * membrane potential can move a spike by a step after a few hundred ms), so
 * the GPU path and the CPU restatement must evaluate the hot-path
 * transcendentals with *identical* operation sequences.  Everything here is
 * written only in IEEE-754 binary32 add/sub/mul/div, exact bit manipulation
 * and round-to-nearest-even conversions, and is compiled without FP
 * contraction on both sides (-ffp-contract=off), so the host compiler and the
 * gfx950 compiler produce bit-identical results.
 *
 *   vc_expf  : Cody-Waite argument reduction x = k ln2 + r, |r| <= ln2/2,
 *              degree-7 Taylor polynomial of e^r in Horner form, exact
 *              scaling by 2^k (two-step for subnormal results).
 *   vc_logf  : x = 2^e m with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
 *              log m = 2 atanh(s) = 2s (1 + s^2/3 + s^4/5 + ... + s^10/11).
 *   Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
 *              "Parallel random numbers: as easy as 1, 2, 3", SC'11).
 *
 * The reference calls the CUDA libm expf/logf (src/DIRECTIVE.h:73-77 maps
 * `exponential`/`logarithm` to expf/logf); these routines replace them with
 * our own functions whose error against the correctly rounded result is
 * bounded by a few ulp (tested in tests/test_vc_math.py).  The reference's
 * cuRAND MRG32k3a streams are replaced by Philox draws keyed by
 * (seed, entity, step, stream) -- see vc_uniform01().
 */
#ifndef VC_MATH_H
#define VC_MATH_H

#include <stdint.h>

#ifdef __HIPCC__
#define VC_HD __host__ __device__ __forceinline__
#else
#define VC_HD static inline
#endif

/* ---- bit casts -------------------------------------------------------- */
VC_HD uint32_t vc_f2u(float f) {
    union { float f; uint32_t u; } c; c.f = f; return c.u;
}
VC_HD float vc_u2f(uint32_t u) {
    union { float f; uint32_t u; } c; c.u = u; return c.f;
}

/* ---- exp ---------------------------------------------------------------- */
#define VC_LOG2E   1.44269504088896341f
#define VC_LN2_HI  0.693145751953125f          /* 16 significant bits: k*hi exact for |k|<256 */
#define VC_LN2_LO  1.42860682030941723212e-6f  /* ln2 - VC_LN2_HI */
#define VC_EXP_MAX 88.7228317f                 /* above: +inf */
#define VC_EXP_MIN (-103.972084f)              /* below: +0   */

VC_HD float vc_expf(float x) {
    if (x != x) return x;                          /* NaN */
    if (x > VC_EXP_MAX) return vc_u2f(0x7f800000u);
    if (x < VC_EXP_MIN) return 0.0f;
    float kf = __builtin_rintf(x * VC_LOG2E);
    int k = (int)kf;
    float r = (x - kf * VC_LN2_HI) - kf * VC_LN2_LO;
    /* e^r = 1 + r + r^2/2! + ... + r^7/7!  (Horner) */
    float p = 1.98412698412698413e-4f;             /* 1/7! */
    p = p * r + 1.38888888888888889e-3f;           /* 1/6! */
    p = p * r + 8.33333333333333333e-3f;           /* 1/5! */
    p = p * r + 4.16666666666666667e-2f;           /* 1/4! */
    p = p * r + 1.66666666666666667e-1f;           /* 1/3! */
    p = p * r + 0.5f;                              /* 1/2! */
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    if (k > 127) {                                 /* x close to VC_EXP_MAX */
        p = p * 2.0f;
        k -= 1;
    }
    if (k >= -126) {
        return p * vc_u2f((uint32_t)(k + 127) << 23);
    }
    /* subnormal result: exact scale by 2^(k+64) then one rounding multiply */
    return (p * vc_u2f((uint32_t)(k + 64 + 127) << 23)) * 5.42101086242752217e-20f; /* 2^-64 */
}

/* ---- log ---------------------------------------------------------------- */
VC_HD float vc_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return vc_u2f(0x7fc00000u);      /* NaN */
    if (x == 0.0f) return vc_u2f(0xff800000u);     /* -inf */
    if (x == vc_u2f(0x7f800000u)) return x;        /* +inf */
    int e = 0;
    if (x < 1.17549435e-38f) {                     /* subnormal input */
        x = x * 8388608.0f;                        /* 2^23, exact */
        e = -23;
    }
    uint32_t u = vc_f2u(x);
    e += (int)((u >> 23) & 0xffu) - 127;
    float m = vc_u2f((u & 0x007fffffu) | 0x3f800000u);   /* [1, 2) */
    if (m > 1.41421356f) {
        m = m * 0.5f;                              /* exact */
        e += 1;
    }
    float f = m - 1.0f;                            /* exact (Sterbenz) */
    float s = f / (2.0f + f);
    float z = s * s;
    float q = 0.181818181818181818f;               /* 2/11 */
    q = q * z + 0.222222222222222222f;             /* 2/9  */
    q = q * z + 0.285714285714285714f;             /* 2/7  */
    q = q * z + 0.4f;                              /* 2/5  */
    q = q * z + 0.666666666666666667f;             /* 2/3  */
    float logm = (s + s) + s * (q * z);
    float ef = (float)e;
    return ef * VC_LN2_HI + (ef * VC_LN2_LO + logm);
}

/* ---- Philox4x32-10 ------------------------------------------------------ */
typedef struct { uint32_t v[4]; } vc_u32x4;

VC_HD void vc_mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

VC_HD vc_u32x4 vc_philox4x32_10(vc_u32x4 ctr, uint32_t k0, uint32_t k1) {
    for (int i = 0; i < 10; ++i) {
        uint32_t hi0, lo0, hi1, lo1;
        vc_mulhilo32(0xD2511F53u, ctr.v[0], &hi0, &lo0);
        vc_mulhilo32(0xCD9E8D57u, ctr.v[2], &hi1, &lo1);
        vc_u32x4 n;
        n.v[0] = hi1 ^ ctr.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ ctr.v[3] ^ k1;
        n.v[3] = lo0;
        ctr = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return ctr;
}

/* Uniform in (0, 1]: the convention of curand_uniform, which the reference's
 * Poisson draw relies on (-log(u) must be finite). Exact in binary32. */
VC_HD float vc_u01_open0(uint32_t x) {
    return (float)((x >> 8) + 1u) * 5.9604644775390625e-8f;   /* 2^-24 */
}
/* Uniform in [0, 1). */
VC_HD float vc_u01_open1(uint32_t x) {
    return (float)(x >> 8) * 5.9604644775390625e-8f;
}

/* Random streams (the third counter word). */
#define VC_STREAM_LGN_SPIKE  1u
#define VC_STREAM_LGN_INIT   2u
#define VC_STREAM_V1_BG      3u
#define VC_STREAM_V1_INIT    4u
#define VC_STREAM_V1_FAILFF  5u   /* LGN->V1 synaptic failure, pack = synapse slot     */
#define VC_STREAM_V1_FAIL    6u   /* V1->V1 synaptic failure, pack = presynaptic id    */
#define VC_STREAM_V1_NOISE   7u   /* voltage noise (noisyDep)                          */

/* ---- cos(2 pi x) and a standard normal ---------------------------------- */
/* cos(2 pi x) for x in [0, 1).  The reduction to z in [0, 1/8] is exact in
 * binary32 (1 - x, 1/2 - y and 1/4 - z are exact for the ranges they are
 * taken on); then the Taylor series of cos (degree 8) or sin (degree 9) of
 * theta = 2 pi z, |theta| <= pi/4, in Horner form in theta^2 (truncation
 * below 3e-8). */
VC_HD float vc_cos2pif(float x) {
    float y = x > 0.5f ? 1.0f - x : x;             /* cos(2 pi x) = cos(2 pi (1 - x))  */
    float sgn = 1.0f;
    if (y > 0.25f) { y = 0.5f - y; sgn = -1.0f; }  /* cos(pi - t) = -cos t             */
    float r;
    if (y <= 0.125f) {
        float t = y * 6.28318548f;
        float t2 = t * t;
        float q = 2.48015873e-5f;                  /* 1/8!  */
        q = q * t2 - 1.38888889e-3f;               /* 1/6!  */
        q = q * t2 + 4.16666667e-2f;               /* 1/4!  */
        q = q * t2 - 0.5f;                         /* 1/2!  */
        r = q * t2 + 1.0f;
    } else {                                       /* cos(2 pi y) = sin(2 pi (1/4 - y)) */
        float t = (0.25f - y) * 6.28318548f;
        float t2 = t * t;
        float q = 2.75573192e-6f;                  /* 1/9!  */
        q = q * t2 - 1.98412698e-4f;               /* 1/7!  */
        q = q * t2 + 8.33333333e-3f;               /* 1/5!  */
        q = q * t2 - 0.166666667f;                 /* 1/3!  */
        r = t + t * (q * t2);
    }
    return sgn * r;
}

/* A standard normal from two 32-bit draws by the Box-Muller transform
 * (Box and Muller, "A note on the generation of random normal deviates",
 * Ann. Math. Stat. 29, 1958): sqrt(-2 log u1) cos(2 pi u2), u1 in (0, 1],
 * u2 in [0, 1).  It replaces cuRAND's curand_normal (src/coredynamics.cu
 * :1430, 1479, 2769); the square root is the correctly rounded one on both
 * sides. */
VC_HD float vc_normal(uint32_t a, uint32_t b) {
    const float u1 = vc_u01_open0(a);
    const float u2 = vc_u01_open1(b);
    const float r = __builtin_sqrtf(-2.0f * vc_logf(u1));
    return r * vc_cos2pif(u2);
}

/* Counter layout: (entity id, step, stream | (pack << 8), trial). */
VC_HD vc_u32x4 vc_draw4(uint32_t seed_lo, uint32_t seed_hi, uint32_t entity,
                        uint32_t step, uint32_t stream, uint32_t pack,
                        uint32_t trial) {
    vc_u32x4 c;
    c.v[0] = entity;
    c.v[1] = step;
    c.v[2] = stream | (pack << 8);
    c.v[3] = trial;
    return vc_philox4x32_10(c, seed_lo, seed_hi);
}

/* Poisson count by sequential inversion given p0 = exp(-lambda) and u in
 * (0,1]; the same float sequence on both sides. Capped at 64 events. */
VC_HD uint32_t vc_poisson_inv(float u, float lambda, float p0) {
    uint32_t k = 0;
    float p = p0;
    float F = p0;
    while (u > F && k < 64u) {
        k += 1u;
        p = p * lambda / (float)k;
        F = F + p;
    }
    return k;
}

#endif /* VC_MATH_H */
