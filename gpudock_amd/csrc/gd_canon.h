/*
 * gd_canon.h -- the canonical FP32 arithmetic of the REMC docking step.
 *
 * Every floating-point operation that decides a Monte Carlo trajectory is
 * written here once, in terms of IEEE-754 operations that are correctly
 * rounded on both the x86 host and gfx950 (add, sub, mul, fma, div, sqrt,
 * rint, compares) and integer bit manipulation.  Nothing calls a vendor
 * math library: exp, log, sin and cos are built from those operations.
 * Compiled with contraction off (-ffp-contract=off) on both sides, the same
 * inputs therefore give the same bits on the CPU and on the GPU, which is
 * what lets a CPU replay reproduce the GPU's accept/exchange decisions.
 *
 * Methods (written from the published algorithms, constants derived here):
 *   exp  : Cody-Waite reduction x = n*ln2 + r with ln2 split into a float
 *          and its float remainder, Taylor polynomial of degree 7 on
 *          |r| <= ln2/2 (truncation error < 6e-9 relative), scale by 2^n
 *          assembled from exponent bits; a shorter branch-free variant
 *          (one-constant reduction, fitted degree-6 polynomial) serves the
 *          energy loops.
 *   log  : x = m*2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
 *          log m = 2 atanh(s) summed to s^11.
 *   sin/cos : reduction by n = rint(x*2/pi) with pi/2 split in three
 *          floats, Taylor polynomials on |r| <= pi/4, quadrant selection.
 *   Philox4x32-10 : Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 *          as easy as 1, 2, 3", SC'11 (multipliers 0xD2511F53, 0xCD9E8D57,
 *          Weyl key increments 0x9E3779B9, 0xBB67AE85).
 *
 *   1/sqrt, 1/x : integer-bit seeds (0x5f3759df - i/2, 0x7ef311c3 - i) and
 *          three Newton iterations each, all in fma.
 *
 * Energy forms follow the reference kernel gpudocksm-rem-1.2/src/
 * calcenergy_d.cu:69-135 (pair terms), :142-149 (hydrophobic restraint),
 * :186-202 (KDE), :240-268 (position restraints), :276-309 (distance,
 * averaging, normalisation); the algebra is the same, the rounding order is
 * the one written below.
 */
#ifndef GD_CANON_H
#define GD_CANON_H

#include <stdint.h>

#if defined(__HIPCC__)
#define GD_HD __host__ __device__ __forceinline__
#else
#define GD_HD static inline
#endif

/* Inside a branch whose condition is wavefront-uniform: keeps the compiler
 * from evaluating the branch body unconditionally (it has no effect on
 * the values computed). */
#if defined(__HIP_DEVICE_COMPILE__)
#define GD_KEEP_BRANCH() asm volatile("")
/* Has x computed at this point (no effect on its value): lets a register
 * whose last use follows be reused instead of copied. */
#define GD_COMPUTED(x) asm volatile("" : : "v"(x))
#else
#define GD_KEEP_BRANCH() ((void) 0)
#define GD_COMPUTED(x) ((void) (x))
#endif

/* ------------------------------------------------------------------ */
/* bit casts                                                           */
/* ------------------------------------------------------------------ */
GD_HD float gd_as_float(uint32_t u)
{
  union { uint32_t u; float f; } v;
  v.u = u;
  return v.f;
}

GD_HD uint32_t gd_as_uint(float f)
{
  union { uint32_t u; float f; } v;
  v.f = f;
  return v.u;
}

GD_HD float gd_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
GD_HD float gd_rint(float x) { return __builtin_rintf(x); }
GD_HD float gd_sqrt(float x) { return __builtin_sqrtf(x); }

#define GD_INF_BITS 0x7f800000u

/* 1/sqrt(x), x positive normal: seed from the exponent/mantissa bits
 * (0x5f3759df - bits/2), then one step y <- y (1 + e p(e)), e = 1 - x y^2,
 * |e| <= 0.07.  p is a cubic fitted here for minimax relative error of
 * the result over that range of e (the series of (1 - e)^-1/2 = 1 + e/2 +
 * 3e^2/8 + 5e^3/16 + 35e^4/128 + ..., re-weighted: 2.3e-8 in exact
 * arithmetic); with the FP32 roundings at most 1.19 ulp over every float in
 * [1, 4), which repeats for every even exponent step.  9 ops. */
GD_HD float gd_rsqrt(float x)
{
  const float y = gd_as_float(0x5f3759dfu - (gd_as_uint(x) >> 1));
  const float e = gd_fma(-x, y * y, 1.0f);
  float p = gd_fma(0x1.18ed8ap-2f, e, 0x1.41793ep-2f);
  p = gd_fma(p, e, 0x1.7fffd4p-2f);
  p = gd_fma(p, e, 0x1.ffff90p-2f);
  return gd_fma(y, e * p, y);
}

/* 1/b, b positive normal below 2^126: seed 0x7ef311c3 - bits (relative
 * error d <= 5.1e-2); one third-order step y <- y (1 + e + e^2),
 * e = 1 - b y (error d^3 <= 1.3e-4), then one Newton step y <- y + y e
 * (error d^2 ~ 2e-8): below 1 ulp, 6 ops. */
GD_HD float gd_rcp(float b)
{
  float y = gd_as_float(0x7ef311c3u - gd_as_uint(b));
  const float e = gd_fma(-b, y, 1.0f);
  y = gd_fma(y, gd_fma(e, e, e), y);
  y = gd_fma(y, gd_fma(-b, y, 1.0f), y);
  return y;
}

/* ------------------------------------------------------------------ */
/* exp                                                                  */
/* ------------------------------------------------------------------ */
/* Results for x > 88 are +inf and for x < -86 exactly 0, so no
 * subnormal result is ever produced (both sides agree on those too). */
GD_HD float gd_exp(float x)
{
  if (x != x)
    return x;
  if (x > 88.0f)
    return gd_as_float(GD_INF_BITS);
  if (x < -86.0f)
    return 0.0f;
  const float n = gd_rint(x * 0x1.715476p+0f);         /* x / ln2 */
  float r = gd_fma(-n, 0x1.62e430p-1f, x);              /* ln2 (float) */
  r = gd_fma(-n, -0x1.05c610p-29f, r);                   /* ln2 - float(ln2) */
  float p = 0x1.a01a02p-13f;                             /* 1/7! */
  p = gd_fma(p, r, 0x1.6c16c2p-10f);                     /* 1/6! */
  p = gd_fma(p, r, 0x1.111112p-7f);                      /* 1/5! */
  p = gd_fma(p, r, 0x1.555556p-5f);                      /* 1/4! */
  p = gd_fma(p, r, 0x1.555556p-3f);                      /* 1/3! */
  p = gd_fma(p, r, 0.5f);
  p = gd_fma(p, r, 1.0f);
  p = gd_fma(p, r, 1.0f);
  const int ni = (int) n;                                /* in [-125, 127] */
  return p * gd_as_float((uint32_t) (ni + 127) << 23);
}

/* exp for the energy loops, where the argument is never NaN and never
 * above 88 (every PMF, H-bond and KDE exponent is <= 18): branch-free.
 * The argument is clamped at -86 (below it the result is e^-86 ~ 4e-38,
 * which no energy sum can tell from 0).  n = round(x / ln2) comes from
 * adding 1.5 * 2^23 in one fma (the integer lands in the low mantissa bits,
 * so 2^n is one shift-and-add of those bits); then a one-constant
 * Cody-Waite step (|r| error <= 2.4e-7 at |x| = 86) and a degree-5
 * polynomial 1 + r + r^2 q(r) fitted here for minimax relative error on
 * |r| <= ln2/2 (1.05e-7 before FP32 rounding); total error below 3 ulp on
 * [-20, 20] (mean 0.75 ulp) and below 5 ulp down to -86, a relative 1.8e-7
 * at most: two orders under the energies' 1e-5 bar, for one fma per
 * exponential fewer than a degree-6 fit.  Returns the polynomial p and
 * sets *scale = 2^n; e^x = p * 2^n. */
GD_HD float gd_exp_hot_split(float x, float *scale)
{
  const float xc = __builtin_fmaxf(x, -86.0f);
  const float t = gd_fma(xc, 0x1.715476p+0f, 0x1.8p+23f);
  const float n = t - 0x1.8p+23f;
  const float r = gd_fma(-n, 0x1.62e430p-1f, xc);
  float p = 0x1.106288p-7f;
  p = gd_fma(p, r, 0x1.5729f2p-5f);
  p = gd_fma(p, r, 0x1.5557aep-3f);
  p = gd_fma(p, r, 0x1.fffdfcp-2f);
  p = gd_fma(p, r, 1.0f);
  p = gd_fma(p, r, 1.0f);
  /* bits(t) = 0x4b400000 + n: (bits << 23) + (127 << 23) = bits of 2^n */
  *scale = gd_as_float((gd_as_uint(t) << 23) + 0x3f800000u);
  return p;
}

/* e^x = p * 2^n of gd_exp_hot_split.  The product is exact (n >= -124 keeps
 * it normal), so c + e^x rounds like gd_fma(p, 2^n, c): callers that add
 * the exponential to something use that form, one op fewer. */
GD_HD float gd_exp_hot(float x)
{
  float s;
  const float p = gd_exp_hot_split(x, &s);
  return p * s;
}

/* ------------------------------------------------------------------ */
/* log                                                                  */
/* ------------------------------------------------------------------ */
GD_HD float gd_log(float x)
{
  if (x != x)
    return x;
  if (x < 0.0f)
    return gd_as_float(0x7fc00000u);
  if (x == 0.0f)
    return -gd_as_float(GD_INF_BITS);
  if (x == gd_as_float(GD_INF_BITS))
    return x;
  int e = 0;
  if (x < 0x1p-126f) {            /* subnormal: scale into the normal range */
    x = x * 0x1p+24f;
    e = -24;
  }
  const uint32_t b = gd_as_uint(x);
  e += (int) ((b >> 23) & 0xffu) - 127;
  float m = gd_as_float((b & 0x007fffffu) | 0x3f800000u);   /* [1, 2) */
  if (m > 0x1.6a09e6p+0f) {                                  /* sqrt(2) */
    m = m * 0.5f;
    e += 1;
  }
  const float f = m - 1.0f;
  const float s = f / (m + 1.0f);
  const float s2 = s * s;
  float p = 0x1.745d18p-4f;              /* 1/11 */
  p = gd_fma(p, s2, 0x1.c71c72p-4f);     /* 1/9 */
  p = gd_fma(p, s2, 0x1.24924ap-3f);     /* 1/7 */
  p = gd_fma(p, s2, 0x1.99999ap-3f);     /* 1/5 */
  p = gd_fma(p, s2, 0x1.555556p-2f);     /* 1/3 */
  p = p * s2;
  const float lm = gd_fma(2.0f * s, p, 2.0f * s);            /* 2 atanh(s) */
  const float ef = (float) e;
  return gd_fma(ef, 0x1.62e430p-1f, gd_fma(ef, -0x1.05c610p-29f, lm));
}

/* ------------------------------------------------------------------ */
/* sin / cos                                                            */
/* ------------------------------------------------------------------ */
GD_HD void gd_sincos(float x, float *sp, float *cp)
{
  const float n = gd_rint(x * 0x1.45f306p-1f);          /* x * 2/pi */
  float r = gd_fma(-n, 0x1.921fb6p+0f, x);
  r = gd_fma(-n, -0x1.777a5cp-25f, r);
  r = gd_fma(-n, -0x1.000000p-49f, r);
  const float r2 = r * r;
  /* sin r = r - r^3/3! + r^5/5! - r^7/7! + r^9/9! */
  float ps = 0x1.71de3ap-19f;            /* 1/9! */
  ps = gd_fma(ps, r2, -0x1.a01a02p-13f); /* -1/7! */
  ps = gd_fma(ps, r2, 0x1.111112p-7f);   /* 1/5! */
  ps = gd_fma(ps, r2, -0x1.555556p-3f);  /* -1/3! */
  ps = ps * r2;
  const float s = gd_fma(ps, r, r);
  /* cos r = 1 - r^2/2! + r^4/4! - r^6/6! + r^8/8! - r^10/10! */
  float pc = -0x1.27e4fcp-22f;           /* -1/10! */
  pc = gd_fma(pc, r2, 0x1.a01a02p-16f);  /* 1/8! */
  pc = gd_fma(pc, r2, -0x1.6c16c2p-10f); /* -1/6! */
  pc = gd_fma(pc, r2, 0x1.555556p-5f);   /* 1/4! */
  pc = gd_fma(pc, r2, -0.5f);
  const float c = gd_fma(pc, r2, 1.0f);
  const int q = ((int) n) & 3;
  float so, co;
  if (q == 0) { so = s;  co = c;  }
  else if (q == 1) { so = c;  co = -s; }
  else if (q == 2) { so = -s; co = -c; }
  else { so = -c; co = s;  }
  *sp = so;
  *cp = co;
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 and the (0,1] mapping                                  */
/* ------------------------------------------------------------------ */
typedef struct { uint32_t v[4]; } gd_u32x4;

GD_HD gd_u32x4 gd_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                uint32_t k0, uint32_t k1)
{
  for (int i = 0; i < 10; ++i) {
    const uint64_t p0 = (uint64_t) 0xD2511F53u * (uint64_t) c0;
    const uint64_t p1 = (uint64_t) 0xCD9E8D57u * (uint64_t) c2;
    const uint32_t hi0 = (uint32_t) (p0 >> 32), lo0 = (uint32_t) p0;
    const uint32_t hi1 = (uint32_t) (p1 >> 32), lo1 = (uint32_t) p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  gd_u32x4 r;
  r.v[0] = c0; r.v[1] = c1; r.v[2] = c2; r.v[3] = c3;
  return r;
}

/* Top 24 bits plus one, times 2^-24: a value in (0, 1], exact in float.
 * Same half-open side as the reference's curand_uniform. */
GD_HD float gd_u01(uint32_t w) { return (float) ((w >> 8) + 1u) * 0x1p-24f; }

/* Counter layout.  key = (seed, complex id).
 *   MC step s of replica r: ctr (r, s, 0, 0) gives u0..u3, ctr (r, s, 1, 0)
 *   gives u4..u7; u0..u5 perturb the six degrees of freedom, u6 is the
 *   Metropolis draw, u7 the second draw of the reversed test.
 *   Exchange number x of temperature chain c, temperature pair j:
 *   ctr (c, x, 2, j >> 2), element j & 3. */
#define GD_CTR_MOVE_A 0u
#define GD_CTR_MOVE_B 1u
#define GD_CTR_EXCHANGE 2u

/* The reference's draw toggles (toggle.h:9,24).  fixed_draws = IS_RANDOM 0:
 * every draw is 20 (util_d.cu:106-108), so each move adds 39 x move_scale
 * (move_d.cu:81-82), every MC step is accepted (accept_d.cu:19-20) and a
 * temperature pair swaps only when its probability exceeds 20
 * (exchangereplicas_d.cu:61).  reverse = IS_REVERSE 1 (accept_d.cu:24-27):
 * the Metropolis test is taken again on a second draw with the exponent's
 * sign flipped, and that second test decides. */
#define GD_FIXED_DRAW 20.0f

/* The perturbation of one degree of freedom in [-1, 1] (move_d.cu:81). */
GD_HD float gd_perturbation(int fixed_draws, uint32_t w)
{
  return 2.0f * (fixed_draws ? GD_FIXED_DRAW : gd_u01(w)) - 1.0f;
}

/* Accept_d's decision (accept_d.cu:17-27): u and u2 the step's two draws,
 * minus_beta = -1 / T. */
GD_HD int gd_metropolis(float e_new, float e_old, float minus_beta, float u, float u2,
                        int fixed_draws, int reverse)
{
  if (fixed_draws)
    return 1;
  const float delta_energy = e_new - e_old;
  if (reverse)
    return u2 < gd_exp(-delta_energy * minus_beta);
  return u < gd_exp(delta_energy * minus_beta);
}

/* The draw a temperature pair's swap probability is compared with. */
GD_HD float gd_exchange_draw(int fixed_draws, uint32_t w)
{
  return fixed_draws ? GD_FIXED_DRAW : gd_u01(w);
}

/* ------------------------------------------------------------------ */
/* Rigid move (reference move_d.cu:79-132)                              */
/* ------------------------------------------------------------------ */
/* R = Rz(a) Ry(b) Rx(g) for (a, b, g) = (m[3], m[4], m[5]). */
GD_HD void gd_rotation(const float m[6], float rot[9])
{
  float s1, c1, s2, c2, s3, c3;
  gd_sincos(m[3], &s1, &c1);
  gd_sincos(m[4], &s2, &c2);
  gd_sincos(m[5], &s3, &c3);
  rot[0] = c1 * c2;
  rot[1] = c1 * s2 * s3 - c3 * s1;
  rot[2] = s1 * s3 + c1 * c3 * s2;
  rot[3] = c2 * s1;
  rot[4] = c1 * c3 + s1 * s2 * s3;
  rot[5] = c3 * s1 * s2 - c1 * s3;
  rot[6] = -s2;
  rot[7] = c2 * s3;
  rot[8] = c2 * c3;
}

/* x' = R x + t + c, summed left to right as in the reference. */
GD_HD void gd_transform(const float rot[9], const float m[6], const float c[3],
                        float x, float y, float z, float *ox, float *oy, float *oz)
{
  *ox = rot[0] * x + rot[1] * y + rot[2] * z + m[0] + c[0];
  *oy = rot[3] * x + rot[4] * y + rot[5] * z + m[1] + c[1];
  *oz = rot[6] * x + rot[7] * y + rot[8] * z + m[2] + c[2];
}

/* ------------------------------------------------------------------ */
/* Pair terms (reference calcenergy_d.cu:69-135)                        */
/* ------------------------------------------------------------------ */
/* Per (ligand type, point type) table entry.  A table row holds
 * GD_ROW_STRIDE entries: the 30 point types and two all-zero entries used
 * by padding points. */
typedef struct __attribute__((aligned(16))) {
  float p1a, p2a, pmf0, pmf1, hdb0, hdb1, pad0, pad1;
} gd_pair_param;

#define GD_ROW_STRIDE 32
#define GD_PAD_TYPE 30   /* point type of padding points: all-zero entry */

/* A pocket point's fourth word holds its type as the byte offset of its
 * entry in a table row: type << GD_PT_TYPE_SHIFT (sizeof(gd_pair_param)). */
#define GD_PT_TYPE_SHIFT 5

/* Scalars shared by every pair.  The distance-dependent electrostatics
 * g(s), s = el1 d (calcenergy_d.cu:99-105), is 1 / s for s >= 1 and
 * el0 + a1 s^2 + b1 s^3 below; the pose sums ele_pt q g.  The pair loop
 * takes el1 g instead -- 1 / d far out, el1 el0 + d^2 (ea1 + eb1 d) for
 * d^2 < dcut2 = 1 / el1^2, with ea1 = a1 el1^3, eb1 = b1 el1^4 -- and the
 * pose's electrostatic sum is multiplied by inv_el1 = 1 / el1 once.  The
 * soft L-J denominator 1 + lj0 p1 (1 + lj1 d^2) is 1 + p1 (lj0 + lj01 d^2),
 * lj01 = lj0 lj1. */
typedef struct {
  float lj0, lj1, el0, el1, a1, b1, inv_el1, ea, eb, dcut2, lj01, ea1, eb1, el01, pad0, pad1;
} gd_pair_scalars;

GD_HD void gd_pair_scalars_init(gd_pair_scalars *sc, float lj0, float lj1, float el0,
                                float el1, float a1, float b1)
{
  sc->lj0 = lj0;
  sc->lj1 = lj1;
  sc->el0 = el0;
  sc->el1 = el1;
  sc->a1 = a1;
  sc->b1 = b1;
  sc->inv_el1 = 1.0f / el1;
  sc->ea = (a1 * el1) * el1;
  sc->eb = ((b1 * el1) * el1) * el1;
  sc->dcut2 = 1.0f / (el1 * el1);
  sc->lj01 = lj0 * lj1;
  sc->ea1 = sc->ea * el1;
  sc->eb1 = sc->eb * el1;
  sc->el01 = el0 * el1;
  sc->pad0 = 0.0f;
  sc->pad1 = 0.0f;
}


/* Accumulators of one lane. */
typedef struct {
  float vdw, ele, pmf, psp, hdb;
} gd_pair_acc;

GD_HD float gd_dist2(float dx, float dy, float dz)
{
  return gd_fma(dz, dz, gd_fma(dy, dy, dx * dx));
}

/* Hydrophobic contact kernel of one pair (calcenergy_d.cu:75-81):
 * hpp (1 - (3.5 u - 4.5 u^2 + 2.5 u^3 - 0.5 u^4)) for u = d^2 / 81 <= 1, else
 * 0.  The polynomial is (1 - u)^3 (1 - u / 2) = w^3 (1 + w) / 2, w = 1 - u:
 * with w clamped at 0 the value is an exact 0 beyond the cut, so no test is
 * needed.  The pair adds twice the kernel, hacc + w^3 hpp (1 + w); the
 * caller halves the atom's sum once (an exact scaling). */
GD_HD float gd_hpc_acc2(float d2, float hpp, float hacc)
{
  const float w = __builtin_fmaxf(gd_fma(d2, -0x1.948b10p-7f, 1.0f), 0.0f);   /* 1/81 */
  const float w3 = (w * w) * w;
  return gd_fma(w3, gd_fma(w, hpp, hpp), hacc);
}

/* The five directly summed pair terms (ele: times el1, see
 * gd_pair_scalars).  psp_val is the PSP table value
 * psp[ligand type][residue] of the point (any value for points that are
 * not class 2); it is added for class-2 points inside the PMF radius
 * (calcenergy_d.cu:123-126).  acc->ele collects ele_point * g for the
 * current ligand atom; the caller multiplies that sum by the atom's charge
 * once per atom (q sum(ele g) = sum(q ele g), calcenergy_d.cu:99-105).
 * hdb_row = 0 promises that no entry of this ligand type's row has
 * hdb0 > 0.1, so the H-bond test can be left out; near_any = 0 promises
 * d2 >= dcut2 (no short-range electrostatics). */
GD_HD int gd_pair_terms(float d2, float ele_pt, const gd_pair_param *pp,
                        const gd_pair_scalars *sc, int is_c2, float psp_val,
                        int hdb_row, int near_any, gd_pair_acc *acc)
{
  /* all inverse powers from one reciprocal square root */
  const float r1 = gd_rsqrt(d2);
  const float d = d2 * r1;
  const float r2 = r1 * r1;
  const float r3 = r2 * r1;
  const float r6 = r3 * r3;
  const float r9 = r6 * r3;

  /* soft L-J numerator and denominator */
  const float p1 = pp->p1a * r9;
  const float p2 = pp->p2a * r6;
  const float p4 = gd_fma(p1, gd_fma(sc->lj01, d2, sc->lj0), 1.0f);

  /* distance-dependent electrostatics times el1 (the pose's sum is scaled
   * by 1 / el1 once); the short-range form only where some pair of the
   * call's group (near_any) is inside it (r1's other uses come first, so
   * the far form takes r1 in place) */
  GD_COMPUTED(r9);
  GD_COMPUTED(d);
  if (near_any) {
    GD_KEEP_BRANCH();
    const float gn = gd_fma(d2, gd_fma(d, sc->eb1, sc->ea1), sc->el01);
    acc->ele = gd_fma(ele_pt, d2 < sc->dcut2 ? gn : r1, acc->ele);
  } else {
    acc->ele = gd_fma(ele_pt, r1, acc->ele);
  }

  /* PMF contact denominator */
  const float dm = d - pp->pmf0;
  float es;
  const float ep = gd_exp_hot_split(gd_fma(-0.5f, d, 6.0f) * dm, &es);
  const float q1 = gd_fma(ep, es, 1.0f);   /* 1 + e^z, one rounding as before */

  /* one reciprocal serves both quotients: 1/(p4 q1) */
  const float rr = gd_rcp(p4 * q1);
  acc->vdw = gd_fma((p1 - p2) * q1, rr, acc->vdw);
  acc->pmf = gd_fma(pp->pmf1 * p4, rr, acc->pmf);

  /* pocket-specific; is_c2 = 0 promises psp_val = 0 for every point of the
   * call (a sum that starts at +0 never becomes -0, so adding 0 leaves its
   * bits unchanged and may be left out) */
  if (is_c2)
    acc->psp += (dm <= 0.0f) ? psp_val : 0.0f;

  /* hydrogen bond */
  if (hdb_row && pp->hdb0 > 0.1f) {
    const float h = (d - pp->hdb0) * pp->hdb1;
    acc->hdb = gd_fma(pp->hdb1, gd_exp_hot((-0.5f * h) * h), acc->hdb);
  }
  /* contact (d <= pmf0): the PSP gate above and the MCC contact test
   * (calcmcc_d.cu:72-73, 182-189) */
  return dm <= 0.0f;
}

/* Matthews correlation of the contact confusion counts, in the reference's
 * FP32 expression order (calcmcc_d.cu:196-207); 10000 when undefined. */
#define GD_CMCC_INVALID 10000.0f
GD_HD float gd_mcc(float tp, float fn, float fp, float tn)
{
  const float dividend = gd_sqrt((tp + fp) * (tp + fn) * (tn + fp) * (tn + fn));
  return dividend != 0.0f ? (tp * tn - fp * fn) / dividend : GD_CMCC_INVALID;
}

/* Point types that form an H-bond with at least one ligand type
 * (hdb0 > 0.1 somewhere in their column of the [n_lt][n_pt] table). */
GD_HD uint32_t gd_hdb_point_mask(const float *hdb0, int n_lt, int n_pt)
{
  uint32_t m = 0u;
  for (int lt = 0; lt < n_lt; ++lt)
    for (int pt = 0; pt < n_pt; ++pt)
      if (hdb0[lt * n_pt + pt] > 0.1f)
        m |= 1u << pt;
  return m;
}

/* Order of the pocket points inside the canonical sums: stable by
 * (hydrophobic class first, then H-bond-capable type first, point type,
 * class-2 flag), so that the hydrophobic points fill the leading
 * wavefront-wide chunks (the rest skip that term), the lanes of one
 * wavefront mostly see one point type (coherent branches, broadcast table
 * reads) and the H-bond branch is skipped by wavefronts outside the
 * H-bond-capable groups. */
GD_HD int gd_point_sort_key(int type, int is_c2, int hpc, uint32_t hdb_mask)
{
  return (hpc ? 0 : 256) + (((hdb_mask >> type) & 1u) ? 0 : 128) + type * 4 +
         (is_c2 ? 2 : 0) + (hpc ? 1 : 0);
}

/* Summation blocks.  The ligand atoms are cut into at most GD_MAX_BLOCKS
 * blocks of gd_block_size(lna) consecutive atoms, and the 64-pair rounds of
 * the KDE pair list into at most GD_MAX_BLOCKS blocks of
 * gd_block_size(rounds) consecutive rounds.  Each lane sums each term within
 * a block from +0 in pair order; a lane's total is its block partials added
 * in block order from +0; the per-atom hydrophobic-restraint terms sum the
 * same way.  The blocks depend on the complex only, so a pose's energy has
 * the same bits whether one wavefront evaluates every block in order or
 * several wavefronts evaluate blocks in parallel and one adds the partials. */
#define GD_MAX_BLOCKS 16

GD_HD int gd_block_size(int n)
{
  return n > GD_MAX_BLOCKS ? (n + GD_MAX_BLOCKS - 1) / GD_MAX_BLOCKS : 1;
}

GD_HD int gd_block_count(int n)
{
  const int bs = gd_block_size(n);
  return (n + bs - 1) / bs;
}

/* Raw (averaged, not normalised) terms in the parameter-file order used by
 * the weight and normalisation files: vdw ele pmf hpc hdb dst psp kde lhm. */
enum {
  GD_F_VDW = 0, GD_F_ELE, GD_F_PMF, GD_F_HPC, GD_F_HDB, GD_F_DST, GD_F_PSP,
  GD_F_KDE, GD_F_LHM, GD_F_NTERM
};

/* Output slots, same order as the reference Energy.e[] (dock.h:100-115). */
enum {
  GD_E_VDW = 0, GD_E_ELE, GD_E_PMF, GD_E_PSP, GD_E_HDB, GD_E_HPC, GD_E_KDE,
  GD_E_LHM, GD_E_DST, GD_E_TOTAL, GD_E_N
};

/* slot index of each file-order term */
#define GD_FILE_TO_SLOT_INIT {GD_E_VDW, GD_E_ELE, GD_E_PMF, GD_E_HPC, GD_E_HDB, \
                              GD_E_DST, GD_E_PSP, GD_E_KDE, GD_E_LHM}

/* term_order 0: parameter k applies to file-order term k (the documented
 * order, cpu_mic calcenergy_cpu.C:220-239); 1: parameter k applies to
 * slot k (rem-1.2 calcenergy_d.cu:301-309,349-358). */
/* ------------------------------------------------------------------ */
/* Bayesian total (gpudocksm-bayes-1.0 calcenergy_d.cu:315-367 with the   */
/* pdfs of util_d.cu:220-303 and the fitted parameters of               */
/* distribution.h:21-77)                                                 */
/* ------------------------------------------------------------------ */
#define GD_PDF_FLOOR 1e-8f   /* the reference's floor for a zero pdf */

/* log10 x = ln x / ln 10 */
GD_HD float gd_log10(float x) { return gd_log(x) * 0x1.bcb7b2p-2f; }

/* a pdf that is 0, infinite or NaN counts as the floor.  The reference
 * floors exact zeros only (its infinite and NaN cases poison the total);
 * with this exp a pdf whose exponential factor is below e^-86 is 0. */
GD_HD float gd_pdf_floor(float v)
{
  return (v > 0.0f && v < gd_as_float(GD_INF_BITS)) ? v : GD_PDF_FLOOR;
}

GD_HD float gd_norm_pdf(float x, float loc, float scale)
{
  const float norm = 1.0f / (scale * 0x1.40d932p+1f);     /* sqrt(2 pi) */
  const float d = x - loc;
  return gd_pdf_floor(norm * gd_exp(0.0f - d * d / (2.0f * scale * scale)));
}

GD_HD float gd_cauchy_pdf(float x, float loc, float scale)
{
  const float norm = 1.0f / (0x1.921fb6p+1f * scale);      /* pi */
  const float t = (x - loc) / scale;
  return gd_pdf_floor(norm * (1.0f / (1.0f + t * t)));
}

/* e^-z / (1 + e^-z)^2 is even in z: evaluated at |z| so that e^-|z| <= 1
 * never overflows (the reference's e^-z does for z < -88) */
GD_HD float gd_logistic_pdf(float x, float loc, float scale)
{
  const float z = (x - loc) / scale;
  const float e = gd_exp(-__builtin_fabsf(z));
  const float q = 1.0f + e;
  return gd_pdf_floor((1.0f / scale) * (e / (q * q)));
}

GD_HD float gd_wald_pdf(float x, float loc, float scale)
{
  const float nx = (x - loc) / scale;
  if (!(nx > 0.0f))
    return GD_PDF_FLOOR;
  const float norm = 1.0f / (gd_sqrt(0x1.921fb6p+2f * (nx * nx * nx)) * scale);   /* 2 pi */
  const float t = nx - 1.0f;
  return gd_pdf_floor(norm * gd_exp(-(t * t) / (2.0f * nx)));
}

/* Fitted (location, scale) of each normalised term for high- and
 * low-quality decoys, file order vdw ele pmf hpc hdb dst psp kde lhm
 * (distribution.h:21-77). */
#define GD_BAYES_HIGH_INIT {                                                   \
    {0.05286683f, 0.29916642f}, {-0.3653806f, 0.00681102f},                   \
    {-0.0221777f, 0.11847532f}, {-1.0041231f, 0.1151713f},                    \
    {0.60986808f, 0.30974999f}, {-0.8246522f, 0.05218529f},                   \
    {-0.2346999f, 0.09871713f}, {-1.0112055f, 0.09882933f},                   \
    {-0.3377100f, 0.09282947f}}
#define GD_BAYES_LOW_INIT {                                                    \
    {0.15818486f, 0.29032471f}, {-0.3647824f, 0.00534225f},                   \
    {0.04257942f, 0.08801551f}, {-1.0019512f, 0.10312611f},                   \
    {0.76967708f, 0.11224029f}, {-0.7509618f, 0.08757299f},                   \
    {-0.2022064f, 0.07854212f}, {-1.004605f, 0.02967548f},                    \
    {-0.2387023f, 0.12134983f}}

/* Family of each term's pdf, file order (calcenergy_d.cu:319-342 of
 * bayes-1.0): 0 normal, 1 Cauchy, 2 logistic, 3 Wald.  Low-decoy psp and
 * hdb use the logistic and normal forms with the parameters fitted for a
 * Laplace and a logistic, as the reference does. */
#define GD_BAYES_FAMILY_INIT {0, 1, 2, 3, 0, 2, 2, 3, 2}

GD_HD float gd_bayes_pdf(int family, float x, float loc, float scale)
{
  switch (family) {
  case 0: return gd_norm_pdf(x, loc, scale);
  case 1: return gd_cauchy_pdf(x, loc, scale);
  case 2: return gd_logistic_pdf(x, loc, scale);
  default: return gd_wald_pdf(x, loc, scale);
  }
}

/* sum_k log10 p_low(e_k) - sum_k log10 p_high(e_k), k in file order */
GD_HD float gd_bayes_total(const float norm_file[GD_F_NTERM])
{
  const float hi[GD_F_NTERM][2] = GD_BAYES_HIGH_INIT;
  const float lo[GD_F_NTERM][2] = GD_BAYES_LOW_INIT;
  const int fam[GD_F_NTERM] = GD_BAYES_FAMILY_INIT;
  float ph = 0.0f, pl = 0.0f;
  for (int k = 0; k < GD_F_NTERM; ++k) {
    ph += gd_log10(gd_bayes_pdf(fam[k], norm_file[k], hi[k][0], hi[k][1]));
    pl += gd_log10(gd_bayes_pdf(fam[k], norm_file[k], lo[k][0], lo[k][1]));
  }
  return pl - ph;
}

/* combine: 0 = weighted sum (calcenergy_d.cu:349-358), 1 = Bayesian total */
GD_HD void gd_finalize(const float sums[GD_F_NTERM], float lna_f,
                       const float a[10], const float b[10], const float w[10],
                       int term_order, int combine, float out[GD_E_N])
{
  float raw[GD_F_NTERM];
  raw[GD_F_VDW] = sums[GD_F_VDW] / lna_f;
  raw[GD_F_ELE] = sums[GD_F_ELE] / lna_f;
  raw[GD_F_PMF] = sums[GD_F_PMF] / lna_f;
  raw[GD_F_HPC] = sums[GD_F_HPC] / lna_f;
  raw[GD_F_HDB] = sums[GD_F_HDB] / lna_f / 0x1.40d932p+1f * -1.0f;   /* sqrt(2 pi) */
  raw[GD_F_DST] = sums[GD_F_DST];
  raw[GD_F_PSP] = sums[GD_F_PSP] / lna_f;
  raw[GD_F_KDE] = sums[GD_F_KDE] / lna_f;
  raw[GD_F_LHM] = sums[GD_F_LHM];
  const int f2s[GD_F_NTERM] = GD_FILE_TO_SLOT_INIT;
  float slot[GD_E_N];
  for (int k = 0; k < GD_F_NTERM; ++k)
    slot[f2s[k]] = raw[k];
  float total = 0.0f;
  if (term_order == 0) {
    for (int k = 0; k < GD_F_NTERM; ++k) {
      const float v = gd_fma(a[k], raw[k], b[k]);
      slot[f2s[k]] = v;
      total = gd_fma(w[k], v, total);
    }
  } else {
    for (int i = 0; i < GD_F_NTERM; ++i) {
      const float v = gd_fma(a[i], slot[i], b[i]);
      slot[i] = v;
      total = gd_fma(w[i], v, total);
    }
  }
  if (combine == 1) {
    float nf[GD_F_NTERM];
    for (int k = 0; k < GD_F_NTERM; ++k)
      nf[k] = slot[f2s[k]];
    total = gd_bayes_total(nf);
  }
  slot[GD_E_TOTAL] = total;
  for (int i = 0; i < GD_E_N; ++i)
    out[i] = slot[i];
}

#endif /* GD_CANON_H */
