/*
 * gd_oracle.c -- CPU oracle of the REMC docking core.  TEST INFRASTRUCTURE:
 * only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * load this library, to check results against; the product never links it.
 *
 * Three independent pieces:
 *
 * 1. Reference-order restatement (orc_energy_ref32 / orc_energy_ref64 /
 *    orc_move_ref32 / orc_ref_run).  The formulas are those of the
 *    reference kernel gpudocksm-rem-1.2/src/calcenergy_d.cu:69-309 written
 *    in the loop order of the reference's CPU restatement
 *    SimilarityMatirxGen/src/unused/cpu_mic/calcenergy_cpu.C:42-243
 *    (ligand atom outer, pocket point inner, plain sequential sums, libm
 *    expf/sqrtf/logf), term order of calcenergy_cpu.C:220-239.  The move
 *    follows move_d.cu:79-132, accept accept_d.cu:18-51, exchange
 *    exchangereplicas_d.cu:28-78, step order montecarlo_d.cu:81-102.  The
 *    FP64 variant uses the same forms in double (the serial original,
 *    gpudocksm-1.1-untested/src/complex.C:1058-1203, has the same algebra).
 *    This is what per-pose energies are compared against.
 *
 * 2. Canonical replay (orc_replay_energy / orc_replay_run).  The same
 *    algorithm evaluated with the canonical FP32 arithmetic of
 *    gpudock_amd/csrc/gd_canon.h and the wave-of-64 summation order the
 *    GPU kernel uses (lane j takes pocket points j, j+64, ...; per-lane
 *    sums reduced by the pairing tree of an xor butterfly 1,2,4,...,32).  Its purpose is
 *    decision parity: given the same Philox stream it must reproduce the
 *    GPU's Monte Carlo trajectory bit for bit.  The arithmetic spec is
 *    shared on purpose; the formulas themselves are pinned by part 1, the
 *    published golden vector and closed forms in the tests.
 *
 * 3. Philox4x32-10 and the (0,1] mapping are taken from gd_canon.h and
 *    pinned by known-answer vectors generated from Random123's philox.h.
 *
 * Deliberate deviations shared with the product (both restate the same
 * choice): an MCS restraint with no valid field is skipped (the reference
 * computes tcc*sqrt(0/0) = NaN, calcenergy_d.cu:255-258); pos == 0 gives
 * lhm 0 (the serial original guards it, complex.C:1195-1196, the GPU
 * kernel takes log(0/0), calcenergy_d.cu:267-268).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "gd_remc.h"
#include "gd_canon.h"

#define ORC_EXPORT __attribute__((visibility("default")))

/* ====================================================================== */
/* 1. reference-order restatement                                          */
/* ====================================================================== */

/* move_d.cu:86-132 */
ORC_EXPORT void orc_move_ref32(const float m[6], const float *ox, const float *oy,
                               const float *oz, int lna, const float c[3], float *nx,
                               float *ny, float *nz, float cen[3])
{
  const float s1 = sinf(m[3]), c1 = cosf(m[3]);
  const float s2 = sinf(m[4]), c2 = cosf(m[4]);
  const float s3 = sinf(m[5]), c3 = cosf(m[5]);
  float r[3][3];
  r[0][0] = c1 * c2;
  r[0][1] = c1 * s2 * s3 - c3 * s1;
  r[0][2] = s1 * s3 + c1 * c3 * s2;
  r[1][0] = c2 * s1;
  r[1][1] = c1 * c3 + s1 * s2 * s3;
  r[1][2] = c3 * s1 * s2 - c1 * s3;
  r[2][0] = -1 * s2;
  r[2][1] = c2 * s3;
  r[2][2] = c2 * c3;
  for (int l = 0; l < lna; ++l) {
    const float x = ox[l], y = oy[l], z = oz[l];
    nx[l] = r[0][0] * x + r[0][1] * y + r[0][2] * z + m[0] + c[0];
    ny[l] = r[1][0] * x + r[1][1] * y + r[1][2] * z + m[1] + c[1];
    nz[l] = r[2][0] * x + r[2][1] * y + r[2][2] * z + m[2] + c[2];
  }
  for (int i = 0; i < 3; ++i)
    cen[i] = c[i] + m[i];
}

static const int kFileToSlot[9] = {GD_E_VDW, GD_E_ELE, GD_E_PMF, GD_E_HPC, GD_E_HDB,
                                   GD_E_DST, GD_E_PSP, GD_E_KDE, GD_E_LHM};

/* calcenergy_cpu.C:24-243 with the rem-1.2 table layout */
ORC_EXPORT void orc_energy_ref32(const gd_ligand *lig, const gd_protein *prt,
                                 const gd_psp *psp, const gd_kde *kde,
                                 const gd_mcs *mcs, const gd_enepara *ep,
                                 gd_complex_size sz, int prt_conf, const float *lx,
                                 const float *ly, const float *lz, const float cen[3],
                                 int term_order, float out[10])
{
  const int lna = sz.lna, pnp = sz.pnp;
  const size_t po = (size_t) prt_conf * pnp;
  float evdw = 0.0f, eele = 0.0f, epmf = 0.0f, epsp = 0.0f, ehdb = 0.0f, ehpc = 0.0f,
        ekde = 0.0f, elhm = 0.0f;

  for (int l = 0; l < lna; ++l) {
    float hpc1 = 0.0f;
    const int lig_t = lig->t[l];
    for (int j = 0; j < pnp; ++j) {
      const size_t p = po + j;
      const int prt_t = prt->t[p];
      const float dx = lx[l] - prt->x[p];
      const float dy = ly[l] - prt->y[p];
      const float dz = lz[l] - prt->z[p];
      const float dst_pow2 = dx * dx + dy * dy + dz * dz;
      const float dst_pow4 = dst_pow2 * dst_pow2;
      const float dst = sqrtf(dst_pow2);

      if (prt->c0_and_d12_or_c2[p] == 1 && dst_pow2 <= 81.0f) {
        hpc1 += prt->hpp[p] *
                (1.0f - (3.5f / 81.0f * dst_pow2 - 4.5f / 81.0f / 81.0f * dst_pow4 +
                         2.5f / 81.0f / 81.0f / 81.0f * dst_pow4 * dst_pow2 -
                         0.5f / 81.0f / 81.0f / 81.0f / 81.0f * dst_pow4 * dst_pow4));
      }

      const float p1 = ep->p1a[lig_t][prt_t] / (dst_pow4 * dst_pow4 * dst);
      const float p2 = ep->p2a[lig_t][prt_t] / (dst_pow4 * dst_pow2);
      const float p4 = p1 * ep->lj0 * (1.0f + ep->lj1 * dst_pow2) + 1.0f;
      evdw += (p1 - p2) / p4;

      const float s1 = ep->el1 * dst;
      float g1;
      if (s1 < 1)
        g1 = ep->el0 + ep->a1 * s1 * s1 + ep->b1 * s1 * s1 * s1;
      else
        g1 = 1.0f / s1;
      eele += lig->c[l] * prt->ele[p] * g1;

      const float dst_minus_pmf0 = dst - ep->pmf0[lig_t][prt_t];
      epmf += ep->pmf1[lig_t][prt_t] /
              (1.0f + expf((-0.5f * dst + 6.0f) * dst_minus_pmf0));

      if (prt->c[p] == 2 && dst_minus_pmf0 <= 0)
        epsp += psp->psp[(size_t) lig_t * psp->n_res + prt->seq3r[p]];

      const float hdb0 = ep->hdb0[lig_t][prt_t];
      if (hdb0 > 0.1f) {
        const float hdb1 = ep->hdb1[lig_t][prt_t];
        const float hdb3 = (dst - hdb0) * hdb1;
        ehdb += hdb1 * expf(-0.5f * hdb3 * hdb3);
      }
    }
    const float hpc2 = (hpc1 - ep->hpl0[lig_t]) / ep->hpl1[lig_t];
    ehpc += 0.5f * hpc2 * hpc2 - ep->hpl2[lig_t];
  }

  for (int l = 0; l < lna; ++l) {
    float kde_val = 0.0f, kde_sz = 0.0f;
    const int lig_t = lig->t[l];
    for (int k = 0; k < sz.pnk; ++k) {
      if (lig_t == kde->t[k]) {
        const float dx = lx[l] - kde->x[k];
        const float dy = ly[l] - kde->y[k];
        const float dz = lz[l] - kde->z[k];
        const float d2 = dx * dx + dy * dy + dz * dz;
        kde_val += expf(ep->kde2 * d2);
        kde_sz += 1.0f;
      }
    }
    if (kde_sz != 0.0f)
      ekde += (kde_val / kde_sz);
  }
  ekde = ekde / ep->kde3;

  for (int m = 0; m < sz.pos; ++m) {
    float lhm_val = 0.0f, lhm_sz = 0.0f;
    for (int l = 0; l < lna; ++l) {
      const size_t f = (size_t) m * mcs->n_field + lig->n[l] + 1;
      if (mcs->x[f] != GD_MCS_INVALID_COORD) {
        const float dx = lx[l] - mcs->x[f];
        const float dy = ly[l] - mcs->y[f];
        const float dz = lz[l] - mcs->z[f];
        lhm_val += dx * dx + dy * dy + dz * dz;
        lhm_sz += 1.0f;
      }
    }
    if (lhm_sz != 0.0f)
      elhm += mcs->tcc[m] * sqrtf(lhm_val / lhm_sz);
  }

  const float dx = cen[0] - prt->pocket_center[0];
  const float dy = cen[1] - prt->pocket_center[1];
  const float dz = cen[2] - prt->pocket_center[2];
  const float edst = sqrtf(dx * dx + dy * dy + dz * dz);

  float e[9];
  e[0] = evdw / lna;
  e[1] = eele / lna;
  e[2] = epmf / lna;
  e[3] = ehpc / lna;
  e[4] = ehdb / lna / sqrtf(2.0f * 3.1415926535f) * -1.0f;
  e[5] = edst;
  e[6] = epsp / lna;
  e[7] = ekde / lna;
  e[8] = sz.pos > 0 ? logf(elhm / sz.pos) : 0.0f;

  float slot[10];
  for (int k = 0; k < 9; ++k)
    slot[kFileToSlot[k]] = e[k];
  float etotal = 0.0f;
  if (term_order == 0) {
    for (int k = 0; k < 9; ++k) {
      e[k] = ep->a_para[k] * e[k] + ep->b_para[k];
      slot[kFileToSlot[k]] = e[k];
    }
    for (int k = 0; k < 9; ++k)
      etotal += ep->w[k] * e[k];
  } else {
    for (int i = 0; i < 9; ++i)
      slot[i] = ep->a_para[i] * slot[i] + ep->b_para[i];
    for (int i = 0; i < 9; ++i)
      etotal += ep->w[i] * slot[i];
  }
  slot[9] = etotal;
  memcpy(out, slot, sizeof slot);
}

/* The same in double; coordinates are moved in double from the float
 * conformer and movematrix. */
ORC_EXPORT void orc_energy_ref64(const gd_ligand *lig, const gd_protein *prt,
                                 const gd_psp *psp, const gd_kde *kde,
                                 const gd_mcs *mcs, const gd_enepara *ep,
                                 gd_complex_size sz, int lig_conf, int prt_conf,
                                 const float mv[6], int term_order, double out[10])
{
  const int lna = sz.lna, pnp = sz.pnp;
  const size_t po = (size_t) prt_conf * pnp, lo = (size_t) lig_conf * lna;
  double m[6];
  for (int k = 0; k < 6; ++k)
    m[k] = mv[k];
  const double s1 = sin(m[3]), c1 = cos(m[3]), s2 = sin(m[4]), c2 = cos(m[4]),
               s3 = sin(m[5]), c3 = cos(m[5]);
  const double r[9] = {c1 * c2, c1 * s2 * s3 - c3 * s1, s1 * s3 + c1 * c3 * s2,
                       c2 * s1, c1 * c3 + s1 * s2 * s3, c3 * s1 * s2 - c1 * s3,
                       -s2,     c2 * s3,                c2 * c3};
  double cen[3];
  for (int i = 0; i < 3; ++i)
    cen[i] = (double) lig->center[i] + m[i];
  double *lx = (double *) malloc(sizeof(double) * 3 * lna);
  double *ly = lx + lna, *lz = ly + lna;
  for (int l = 0; l < lna; ++l) {
    const double x = lig->x[lo + l], y = lig->y[lo + l], z = lig->z[lo + l];
    lx[l] = r[0] * x + r[1] * y + r[2] * z + cen[0];
    ly[l] = r[3] * x + r[4] * y + r[5] * z + cen[1];
    lz[l] = r[6] * x + r[7] * y + r[8] * z + cen[2];
  }
  const double el0 = ep->el0, el1 = ep->el1, a1 = 4.0 - 3.0 * el0, b1 = 2.0 * el0 - 3.0;
  double evdw = 0, eele = 0, epmf = 0, epsp = 0, ehdb = 0, ehpc = 0, ekde = 0, elhm = 0;
  for (int l = 0; l < lna; ++l) {
    double hpc1 = 0.0;
    const int lt = lig->t[l];
    for (int j = 0; j < pnp; ++j) {
      const size_t p = po + j;
      const int pt = prt->t[p];
      const double dx = lx[l] - prt->x[p], dy = ly[l] - prt->y[p], dz = lz[l] - prt->z[p];
      const double d2 = dx * dx + dy * dy + dz * dz;
      const double d = sqrt(d2);
      if (prt->c0_and_d12_or_c2[p] == 1 && d2 <= 81.0) {
        const double u = d2 / 81.0;
        hpc1 += (double) prt->hpp[p] *
                (1.0 - 0.5 * (7.0 * u - 9.0 * u * u + 5.0 * u * u * u - u * u * u * u));
      }
      const double p1 = (double) ep->p1a[lt][pt] / pow(d, 9);
      const double p2 = (double) ep->p2a[lt][pt] / pow(d, 6);
      const double p4 = p1 * (double) ep->lj0 * (1.0 + (double) ep->lj1 * d2) + 1.0;
      evdw += (p1 - p2) / p4;
      const double s = el1 * d;
      const double g = s < 1.0 ? el0 + a1 * s * s + b1 * s * s * s : 1.0 / s;
      eele += (double) lig->c[l] * (double) prt->ele[p] * g;
      const double dm = d - (double) ep->pmf0[lt][pt];
      epmf += (double) ep->pmf1[lt][pt] / (1.0 + exp((-0.5 * d + 6.0) * dm));
      if (prt->c[p] == 2 && dm <= 0)
        epsp += psp->psp[(size_t) lt * psp->n_res + prt->seq3r[p]];
      if (ep->hdb0[lt][pt] > 0.1f) {
        const double h = (d - (double) ep->hdb0[lt][pt]) * (double) ep->hdb1[lt][pt];
        ehdb += (double) ep->hdb1[lt][pt] * exp(-0.5 * h * h);
      }
    }
    const double z = (hpc1 - (double) ep->hpl0[lt]) / (double) ep->hpl1[lt];
    ehpc += 0.5 * z * z - (double) ep->hpl2[lt];
  }
  for (int l = 0; l < lna; ++l) {
    double v = 0.0;
    int n = 0;
    for (int k = 0; k < sz.pnk; ++k)
      if (lig->t[l] == kde->t[k]) {
        const double dx = lx[l] - kde->x[k], dy = ly[l] - kde->y[k], dz = lz[l] - kde->z[k];
        v += exp((double) ep->kde2 * (dx * dx + dy * dy + dz * dz));
        ++n;
      }
    if (n)
      ekde += v / n;
  }
  ekde /= (double) ep->kde3;
  for (int mm = 0; mm < sz.pos; ++mm) {
    double s = 0.0;
    int c = 0;
    for (int l = 0; l < lna; ++l) {
      const size_t f = (size_t) mm * mcs->n_field + lig->n[l] + 1;
      if (mcs->x[f] != GD_MCS_INVALID_COORD) {
        const double dx = lx[l] - mcs->x[f], dy = ly[l] - mcs->y[f], dz = lz[l] - mcs->z[f];
        s += dx * dx + dy * dy + dz * dz;
        ++c;
      }
    }
    if (c)
      elhm += (double) mcs->tcc[mm] * sqrt(s / c);
  }
  const double ddx = cen[0] - prt->pocket_center[0], ddy = cen[1] - prt->pocket_center[1],
               ddz = cen[2] - prt->pocket_center[2];
  double e[9];
  e[0] = evdw / lna;
  e[1] = eele / lna;
  e[2] = epmf / lna;
  e[3] = ehpc / lna;
  e[4] = -ehdb / lna / sqrt(2.0 * 3.14159265358979323846);
  e[5] = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
  e[6] = epsp / lna;
  e[7] = ekde / lna;
  e[8] = sz.pos > 0 ? log(elhm / sz.pos) : 0.0;
  double slot[10];
  for (int k = 0; k < 9; ++k)
    slot[kFileToSlot[k]] = e[k];
  double etotal = 0.0;
  if (term_order == 0) {
    for (int k = 0; k < 9; ++k) {
      e[k] = (double) ep->a_para[k] * e[k] + (double) ep->b_para[k];
      slot[kFileToSlot[k]] = e[k];
      etotal += (double) ep->w[k] * e[k];
    }
  } else {
    for (int i = 0; i < 9; ++i) {
      slot[i] = (double) ep->a_para[i] * slot[i] + (double) ep->b_para[i];
      etotal += (double) ep->w[i] * slot[i];
    }
  }
  slot[9] = etotal;
  memcpy(out, slot, sizeof slot);
  free(lx);
}

/* Convenience: reference-order FP32 energy of a pose given by conformer
 * indices and a movematrix. */
ORC_EXPORT void orc_pose_ref32(const gd_ligand *lig, const gd_protein *prt,
                               const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                               const gd_enepara *ep, gd_complex_size sz, int lig_conf,
                               int prt_conf, const float mv[6], int term_order,
                               float out[10])
{
  const int lna = sz.lna;
  const size_t lo = (size_t) lig_conf * lna;
  float *buf = (float *) malloc(sizeof(float) * 3 * lna);
  float cen[3];
  orc_move_ref32(mv, lig->x + lo, lig->y + lo, lig->z + lo, lna, lig->center, buf,
                 buf + lna, buf + 2 * lna, cen);
  orc_energy_ref32(lig, prt, psp, kde, mcs, ep, sz, prt_conf, buf, buf + lna,
                   buf + 2 * lna, cen, term_order, out);
  free(buf);
}

/* ---------------------------------------------------------------------- */
/* Decision rules and the launch schedule, shared by the reference-order    */
/* run and the canonical replay.  canonical = 0 evaluates the exponentials  */
/* with libm expf as the reference does; 1 with gd_canon.h's gd_exp as the   */
/* GPU does.                                                                */
/* ---------------------------------------------------------------------- */
static float orc_expf(int canonical, float x) { return canonical ? gd_exp(x) : expf(x); }

/* Accept_d (accept_d.cu:22-23): delta_energy = E_new - E_old (the total),
 * accepted when the draw u in (0, 1] is below expf(delta * minus_beta);
 * minus_beta = -1 / T is below zero. */
ORC_EXPORT int orc_accept(int canonical, float e_new, float e_old, float minus_beta, float u)
{
  const float delta_energy = e_new - e_old;
  return u < orc_expf(canonical, delta_energy * minus_beta);
}

/* Accept_d with the reference's draw toggles in the canonical arithmetic
 * (gd_metropolis; accept_d.cu:17-27): u and u2 the step's two draws. */
ORC_EXPORT int orc_metropolis(float e_new, float e_old, float minus_beta, float u, float u2,
                              int fixed_draws, int reverse)
{
  return gd_metropolis(e_new, e_old, minus_beta, u, u2, fixed_draws, reverse);
}

/* ExchangeReplicas_d for one temperature chain (exchangereplicas_d.cu:28-78).
 * temps[t]: temperature index held by chain slot t (updated in place);
 * ener[t]: its total energy; minus_beta[i]: of temperature i; mode_t 0 pairs
 * temperatures (0,1), (2,3), ...; 1 pairs (1,2), (3,4), ...; u[j]: the draw
 * of pair j; acs[t] (or NULL) counts the swaps of slot t.  Returns the number
 * of pairs visited. */
ORC_EXPORT int orc_exchange_chain(int canonical, int n_tmp, const float *minus_beta,
                                  int *temps, const float *ener, int mode_t, const float *u,
                                  int *acs)
{
  int temp_orders[64];
  for (int t = 0; t < n_tmp; ++t)
    temp_orders[temps[t]] = t;
  const int maxpair = n_tmp / 2 - (n_tmp % 2 == 0) * (mode_t == 1);
  for (int pair = 0; pair < maxpair; ++pair) {
    const int i1 = (pair << 1) + mode_t;
    const int i2 = i1 + 1;
    const int o1 = temp_orders[i1];
    const int o2 = temp_orders[i2];
    const int tt = temps[o1];
    const float etot1 = ener[o1];
    const float etot2 = ener[o2];
    const float minus_beta1 = minus_beta[temps[o1]];
    const float minus_beta2 = minus_beta[temps[o2]];
    const float delta = (minus_beta1 - minus_beta2) * (etot2 - etot1);
    if (orc_expf(canonical, delta) > u[pair]) {
      temps[o1] = temps[o2];
      temps[o2] = tt;
      if (acs) {
        acs[o1] += 1;
        acs[o2] += 1;
      }
    }
  }
  return maxpair;
}

/* The launch sequence of gpusingle.cu:8-20:
 *   for (s1 = 0; s1 < steps_total; s1 += steps_per_dump)
 *     for (s2 = 0; s2 < steps_per_dump; s2 += steps_per_exchange)
 *       MonteCarlo_d (steps_per_exchange steps), then
 *       ExchangeReplicas_d in mode_t = (s2 / steps_per_exchange) % 2.
 * Launch x runs MC steps [x spe, (x + 1) spe) of the run (the Philox step
 * counter), and is followed by exchange number x.  mode[x] for x < max
 * (mode may be NULL); returns the number of launches, -1 for a step size
 * below 1. */
ORC_EXPORT int orc_schedule(int steps_total, int steps_per_dump, int steps_per_exchange,
                            int *mode, int max)
{
  if (steps_per_dump < 1 || steps_per_exchange < 1)
    return -1;
  int x = 0;
  for (int s1 = 0; s1 < steps_total; s1 += steps_per_dump)
    for (int s2 = 0; s2 < steps_per_dump; s2 += steps_per_exchange, ++x)
      if (mode && x < max)
        mode[x] = (s2 / steps_per_exchange) % 2;
  return x;
}

/* The exchange draws of chain c after launch x: Philox counter (c, x, 2,
 * j >> 2), element j & 3, for pair j (20 for every pair with fixed_draws). */
static void exchange_draws(uint32_t c, uint32_t x, int n_pair, uint32_t seed, uint32_t cid,
                           int fixed_draws, float *u)
{
  gd_u32x4 u4;
  for (int j = 0; j < n_pair; ++j) {
    if ((j & 3) == 0)
      u4 = gd_philox4x32_10(c, x, GD_CTR_EXCHANGE, (uint32_t) (j >> 2), seed, cid);
    u[j] = gd_exchange_draw(fixed_draws, u4.v[j & 3]);
  }
}

/* Reference-order REMC on the CPU (the CPU baseline): the Metropolis and
 * exchange rules of the reference on the Philox stream, OpenMP over
 * replicas as montecarlo_cpu.C:34-55 does.  Energies come from
 * orc_energy_ref32, so trajectories are statistically, not bitwise, equal
 * to the GPU's.  Runs the launches of orc_schedule(n_steps, steps_per_dump,
 * steps_per_exchange) from the loaded pose (IS_AWAY 0), each followed by an
 * exchange when n_tmp > 1 and exchange != 0.  acs: [n_rep] accepted moves;
 * e_final: [n_rep * 10]. */
static int ref_run(const gd_ligand *lig, const gd_protein *prt, const gd_psp *psp,
                   const gd_kde *kde, const gd_mcs *mcs, const gd_enepara *ep,
                   const gd_temp *temp, gd_complex_size sz, const float move_scale[6],
                   int n_steps, int steps_per_dump, int steps_per_exchange, int exchange,
                   uint32_t seed,
                   uint32_t cid, int term_order, int n_threads, int *acs, int *acs_x,
                   int *idx_tmp_out, float *e_final, float *mv_final, float *e_samp,
                   int *t_samp, int samp_stride, int n_samp)
{
  const int n = sz.n_rep, lna = sz.lna;
  float *mv = (float *) calloc((size_t) n * 6, sizeof(float));
  float *e = (float *) calloc((size_t) n * 10, sizeof(float));
  int *it = (int *) malloc(sizeof(int) * n);
  for (int r = 0; r < n; ++r) {
    it[r] = (r / sz.n_lig) % sz.n_tmp;
    acs[r] = 0;
    if (acs_x)
      acs_x[r] = 0;
  }
#ifdef _OPENMP
  if (n_threads > 0)
    omp_set_num_threads(n_threads);
#else
  (void) n_threads;
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int r = 0; r < n; ++r) {
    const int l = r % sz.n_lig, p = r / (sz.n_tmp * sz.n_lig);
    orc_pose_ref32(lig, prt, psp, kde, mcs, ep, sz, l, p, mv + (size_t) r * 6,
                   term_order, e + (size_t) r * 10);
  }
  const int n_launch = orc_schedule(n_steps, steps_per_dump, steps_per_exchange, NULL, 0);
  if (n_launch < 0) {
    free(mv); free(e); free(it);
    return -1;
  }
  int *modes = (int *) malloc(sizeof(int) * (n_launch + 1));
  orc_schedule(n_steps, steps_per_dump, steps_per_exchange, modes, n_launch);
  for (int x = 0; x < n_launch; ++x) {
    const int s0 = x * steps_per_exchange, s1 = s0 + steps_per_exchange;
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 0; r < n; ++r) {
      const int l = r % sz.n_lig, p = r / (sz.n_tmp * sz.n_lig);
      const size_t lo = (size_t) l * lna;
      float *buf = (float *) malloc(sizeof(float) * 3 * lna);
      float cen[3], mn[6], en[10];
      float *mo = mv + (size_t) r * 6, *eo = e + (size_t) r * 10;
      const float mb = temp[it[r]].minus_beta;
      for (int s = s0; s < s1; ++s) {
        if (e_samp && s % samp_stride == 0 && s / samp_stride < n_samp) {
          /* the state before step s, as the per-step records hold it */
          e_samp[(size_t) r * n_samp + s / samp_stride] = eo[9];
          t_samp[(size_t) r * n_samp + s / samp_stride] = it[r];
        }
        const gd_u32x4 ra = gd_philox4x32_10((uint32_t) r, (uint32_t) s, GD_CTR_MOVE_A, 0u, seed, cid);
        const gd_u32x4 rb = gd_philox4x32_10((uint32_t) r, (uint32_t) s, GD_CTR_MOVE_B, 0u, seed, cid);
        const uint32_t w[7] = {ra.v[0], ra.v[1], ra.v[2], ra.v[3], rb.v[0], rb.v[1], rb.v[2]};
        for (int k = 0; k < 6; ++k)
          mn[k] = move_scale[k] * (2.0f * gd_u01(w[k]) - 1.0f) + mo[k];
        orc_move_ref32(mn, lig->x + lo, lig->y + lo, lig->z + lo, lna, lig->center, buf,
                       buf + lna, buf + 2 * lna, cen);
        orc_energy_ref32(lig, prt, psp, kde, mcs, ep, sz, p, buf, buf + lna,
                         buf + 2 * lna, cen, term_order, en);
        if (orc_accept(0, en[9], eo[9], mb, gd_u01(w[6]))) {
          memcpy(mo, mn, sizeof mn);
          memcpy(eo, en, sizeof en);
          acs[r]++;
        }
      }
      free(buf);
    }
    if (exchange && sz.n_tmp > 1) {
      const int n_chain = sz.n_prt * sz.n_lig;
      float mbs[64], u[32];
      for (int t = 0; t < sz.n_tmp; ++t)
        mbs[t] = temp[t].minus_beta;
      for (int c = 0; c < n_chain; ++c) {
        const int p = c / sz.n_lig, l = c % sz.n_lig;
        const int base = sz.n_tmp * sz.n_lig * p + l;
        int temps[64], acx[64];
        float ener[64];
        for (int t = 0; t < sz.n_tmp; ++t) {
          temps[t] = it[base + sz.n_lig * t];
          ener[t] = e[(size_t) (base + sz.n_lig * t) * 10 + 9];
          acx[t] = 0;
        }
        exchange_draws((uint32_t) c, (uint32_t) x, sz.n_tmp / 2, seed, cid, 0, u);
        orc_exchange_chain(0, sz.n_tmp, mbs, temps, ener, modes[x], u, acx);
        for (int t = 0; t < sz.n_tmp; ++t) {
          it[base + sz.n_lig * t] = temps[t];
          if (acs_x)
            acs_x[base + sz.n_lig * t] += acx[t];
        }
      }
    }
  }
  free(modes);
  if (e_final)
    memcpy(e_final, e, sizeof(float) * 10 * n);
  if (mv_final)
    memcpy(mv_final, mv, sizeof(float) * 6 * n);
  if (idx_tmp_out)
    memcpy(idx_tmp_out, it, sizeof(int) * n);
  free(mv);
  free(e);
  free(it);
  return 0;
}

ORC_EXPORT int orc_ref_run(const gd_ligand *lig, const gd_protein *prt,
                           const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                           const gd_enepara *ep, const gd_temp *temp,
                           gd_complex_size sz, const float move_scale[6], int n_steps,
                           int steps_per_dump, int steps_per_exchange, int exchange,
                           uint32_t seed, uint32_t cid, int term_order, int n_threads,
                           int *acs, int *acs_x, int *idx_tmp_out, float *e_final,
                           float *mv_final)
{
  return ref_run(lig, prt, psp, kde, mcs, ep, temp, sz, move_scale, n_steps, steps_per_dump,
                 steps_per_exchange, exchange, seed, cid, term_order, n_threads, acs, acs_x,
                 idx_tmp_out, e_final, mv_final, NULL, NULL, 1, 0);
}

/* orc_ref_run plus samples of the trajectory: for replica r, the total
 * energy and temperature index of the state before every samp_stride-th
 * step, e_samp/t_samp[r * n_samp + k] for step k * samp_stride. */
ORC_EXPORT int orc_ref_run_samples(const gd_ligand *lig, const gd_protein *prt,
                                   const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                                   const gd_enepara *ep, const gd_temp *temp,
                                   gd_complex_size sz, const float move_scale[6],
                                   int n_steps, int steps_per_dump, int steps_per_exchange,
                                   int exchange, uint32_t seed, uint32_t cid, int term_order,
                                   int n_threads, int *acs, int *acs_x, float *e_samp,
                                   int *t_samp, int samp_stride, int n_samp)
{
  if (samp_stride < 1)
    return -1;
  return ref_run(lig, prt, psp, kde, mcs, ep, temp, sz, move_scale, n_steps, steps_per_dump,
                 steps_per_exchange, exchange, seed, cid, term_order, n_threads, acs, acs_x,
                 NULL, NULL, NULL, e_samp, t_samp, samp_stride, n_samp);
}

/* ====================================================================== */
/* 2. canonical replay                                                     */
/* ====================================================================== */
#define W 64

/* pairing tree of the GPU wave sum: lanes j and j^1 first, then j^2, j^4,
 * j^8, j^16, j^32 */
static float bfly(float *v)
{
  float t[W];
  for (int off = 1; off < W; off <<= 1) {
    for (int j = 0; j < W; ++j)
      t[j] = v[j] + v[j ^ off];
    memcpy(v, t, sizeof t);
  }
  return v[0];
}

typedef struct {
  const gd_ligand *lig;
  const gd_protein *prt;
  const gd_psp *psp;
  const gd_kde *kde;
  const gd_mcs *mcs;
  const gd_enepara *ep;
  gd_complex_size sz;
  int term_order;
  gd_pair_scalars sc;
  /* KDE points per ligand atom type in original order */
  int kseg[GD_MAXTP2 + 1];
  float *kx, *ky, *kz;
  /* canonical point order per protein conformer (stable by sort key) */
  int *perm;
  int combine;
} rp_ctx;

static const gd_protein *g_sort_prt;
static size_t g_sort_base;
static uint32_t g_sort_hdb_mask;

static int point_key(int i)
{
  const gd_protein *p = g_sort_prt;
  const size_t q = g_sort_base + i;
  return gd_point_sort_key(p->t[q], p->c[q] == 2, p->c0_and_d12_or_c2[q] != 0,
                           g_sort_hdb_mask);
}

/* stable insertion-merge sort of indices by point_key (n is small) */
static void stable_sort_points(int *idx, int n)
{
  int *tmp = (int *) malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int w = 1; w < n; w *= 2) {
    for (int lo = 0; lo < n; lo += 2 * w) {
      int mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
      int a = lo, b = mid, k = lo;
      while (a < mid && b < hi)
        tmp[k++] = point_key(idx[b]) < point_key(idx[a]) ? idx[b++] : idx[a++];
      while (a < mid) tmp[k++] = idx[a++];
      while (b < hi) tmp[k++] = idx[b++];
    }
    memcpy(idx, tmp, sizeof(int) * n);
  }
  free(tmp);
}

static void rp_init(rp_ctx *c, const gd_ligand *lig, const gd_protein *prt,
                    const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                    const gd_enepara *ep, gd_complex_size sz, int term_order)
{
  c->lig = lig; c->prt = prt; c->psp = psp; c->kde = kde; c->mcs = mcs; c->ep = ep;
  c->sz = sz; c->term_order = term_order; c->combine = 0;
  gd_pair_scalars_init(&c->sc, ep->lj0, ep->lj1, ep->el0, ep->el1, ep->a1, ep->b1);
  const int n = sz.pnk > 0 ? sz.pnk : 1;
  c->kx = (float *) malloc(sizeof(float) * 3 * n);
  c->ky = c->kx + n;
  c->kz = c->ky + n;
  int cnt = 0;
  for (int t = 0; t < GD_MAXTP2; ++t) {
    c->kseg[t] = cnt;
    for (int k = 0; k < sz.pnk; ++k)
      if (kde->t[k] == t) {
        c->kx[cnt] = kde->x[k];
        c->ky[cnt] = kde->y[k];
        c->kz[cnt] = kde->z[k];
        ++cnt;
      }
  }
  c->kseg[GD_MAXTP2] = cnt;
  c->perm = (int *) malloc(sizeof(int) * (size_t) sz.n_prt * sz.pnp);
  for (int cf = 0; cf < sz.n_prt; ++cf) {
    int *pp = c->perm + (size_t) cf * sz.pnp;
    for (int j = 0; j < sz.pnp; ++j)
      pp[j] = j;
    g_sort_prt = prt;
    g_sort_base = (size_t) cf * sz.pnp;
    g_sort_hdb_mask = gd_hdb_point_mask(&ep->hdb0[0][0], GD_MAXTP2, GD_MAXTP1);
    stable_sort_points(pp, sz.pnp);
  }
}

static void rp_free(rp_ctx *c)
{
  free(c->kx);
  free(c->perm);
}

/* canonical energy of atoms (ax, ay, az)[lna] placed with centre cen */
/* ref: MCC reference contacts [lna][pnp] (original point order) or NULL;
 * *cmcc gets the Matthews correlation of this pose's contacts (0 without
 * ref), calcmcc_d.cu:148-213 */
static void rp_energy(const rp_ctx *c, int prt_conf, const float *ax, const float *ay,
                      const float *az, const float cen[3], float out[10], const char *ref,
                      float *cmcc)
{
  int tp = 0, fp = 0, fn = 0, tn = 0;
  const gd_complex_size sz = c->sz;
  const int lna = sz.lna, pnp = sz.pnp;
  const size_t po = (size_t) prt_conf * pnp;
  const gd_enepara *ep = c->ep;
  const gd_protein *prt = c->prt;
  /* lane totals: block partials added in block order (gd_block_size) */
  gd_pair_acc tot[W];
  memset(tot, 0, sizeof tot);
  float eele_tot[W];
  for (int j = 0; j < W; ++j)
    eele_tot[j] = 0.0f;
  float ehpc = 0.0f;
  const int bs = gd_block_size(lna);
  for (int b0 = 0; b0 < lna; b0 += bs) {
  const int b1 = b0 + bs < lna ? b0 + bs : lna;
  gd_pair_acc acc[W];
  memset(acc, 0, sizeof acc);
  float eele[W];
  for (int j = 0; j < W; ++j)
    eele[j] = 0.0f;
  float ehpc_b = 0.0f;
  for (int l = b0; l < b1; ++l) {
    const float lx = ax[l], ly = ay[l], lz = az[l];
    const int lt = c->lig->t[l];
    const float lq = c->lig->c[l];
    const float *psp_row = c->psp->psp + (size_t) lt * c->psp->n_res;
    float hacc[W];
    for (int j = 0; j < W; ++j) {
      hacc[j] = 0.0f;
      acc[j].ele = 0.0f;
      for (int p = j; p < pnp; p += W) {
        const size_t q = po + c->perm[po + p];
        const float d2 = gd_dist2(lx - prt->x[q], ly - prt->y[q], lz - prt->z[q]);
        if (prt->c0_and_d12_or_c2[q])
          hacc[j] = gd_hpc_acc2(d2, prt->hpp[q], hacc[j]);
        const int pt = prt->t[q];
        gd_pair_param pp;
        pp.p1a = ep->p1a[lt][pt]; pp.p2a = ep->p2a[lt][pt];
        pp.pmf0 = ep->pmf0[lt][pt]; pp.pmf1 = ep->pmf1[lt][pt];
        pp.hdb0 = ep->hdb0[lt][pt]; pp.hdb1 = ep->hdb1[lt][pt];
        pp.pad0 = 0.0f; pp.pad1 = 0.0f;
        const int is_c2 = prt->c[q] == 2;
        const float pv = is_c2 ? psp_row[prt->seq3r[q]] : 0.0f;
        const int now = gd_pair_terms(d2, prt->ele[q], &pp, &c->sc, is_c2, pv, 1, 1, &acc[j]);
        if (ref) {
          const int was = ref[(size_t) l * pnp + c->perm[po + p]];
          tp += was && now;
          fn += was && !now;
          fp += !was && now;
          tn += !was && !now;
        }
      }
    }
    for (int j = 0; j < W; ++j)
      eele[j] = gd_fma(lq, acc[j].ele, eele[j]);
    const float hs = 0.5f * bfly(hacc);   /* the pairs added twice the kernel */
    const float z = (hs - ep->hpl0[lt]) * (1.0f / ep->hpl1[lt]);
    ehpc_b += 0.5f * z * z - ep->hpl2[lt];
  }
  for (int j = 0; j < W; ++j) {
    tot[j].vdw += acc[j].vdw;
    tot[j].pmf += acc[j].pmf;
    tot[j].psp += acc[j].psp;
    tot[j].hdb += acc[j].hdb;
    eele_tot[j] += eele[j];
  }
  ehpc += ehpc_b;
  }

  /* KDE as one list of (atom, point) pairs: atoms < W, each atom's points
   * in segment order, then atoms >= W the same way, each group padded to
   * whole rounds of W pairs; pair q of a group goes to lane q % W, weighted
   * by 1 / (segment length); lane partials per block of rounds */
  int n_pair[2] = {0, 0};
  for (int g = 0; g < 2; ++g) {
    const int la = g ? W : 0, lb = g ? lna : (lna < W ? lna : W);
    for (int l = la; l < lb; ++l)
      n_pair[g] += c->kseg[c->lig->t[l] + 1] - c->kseg[c->lig->t[l]];
  }
  const int rounds_lo = (n_pair[0] + W - 1) / W;
  const int n_rounds = rounds_lo + (n_pair[1] + W - 1) / W;
  const int kbs = gd_block_size(n_rounds > 0 ? n_rounds : 1);
  float ekde = 0.0f, ekde_lane[W], ekde_blk[GD_MAX_BLOCKS][W];
  memset(ekde_blk, 0, sizeof ekde_blk);
  for (int g = 0; g < 2; ++g) {
    const int la = g ? W : 0, lb = g ? lna : (lna < W ? lna : W);
    int q = 0;
    for (int l = la; l < lb; ++l) {
      const int lt = c->lig->t[l];
      const int s0 = c->kseg[lt], s1 = c->kseg[lt + 1];
      const int n = s1 - s0;
      const float w = n > 0 ? 1.0f / (float) n : 0.0f;
      for (int i = s0; i < s1; ++i, ++q) {
        float es;
        const float e = gd_exp_hot_split(ep->kde2 * gd_dist2(ax[l] - c->kx[i], ay[l] - c->ky[i],
                                                             az[l] - c->kz[i]), &es);
        const int kb = ((g ? rounds_lo : 0) + q / W) / kbs;
        ekde_blk[kb][q % W] = gd_fma(e * es, w, ekde_blk[kb][q % W]);
      }
    }
  }
  for (int j = 0; j < W; ++j) {
    ekde_lane[j] = 0.0f;
    for (int kb = 0; kb * kbs < n_rounds; ++kb)
      ekde_lane[j] += ekde_blk[kb][j];
  }
  ekde = bfly(ekde_lane) / ep->kde3;

  float elhm = 0.0f;
  for (int m = 0; m < sz.pos; ++m) {
    float s[W], cn[W];
    for (int j = 0; j < W; ++j) {
      s[j] = 0.0f;
      cn[j] = 0.0f;
      for (int a = j; a < lna && a < j + 2 * W; a += W) {
        const size_t f = (size_t) m * c->mcs->n_field + c->lig->n[a] + 1;
        if (c->mcs->x[f] != GD_MCS_INVALID_COORD) {
          s[j] += gd_dist2(ax[a] - c->mcs->x[f], ay[a] - c->mcs->y[f], az[a] - c->mcs->z[f]);
          cn[j] += 1.0f;
        }
      }
    }
    const float ss = bfly(s), cs = bfly(cn);
    if (cs > 0.0f)
      elhm += c->mcs->tcc[m] * gd_sqrt(ss / cs);
  }
  elhm = sz.pos > 0 ? gd_log(elhm / (float) sz.pos) : 0.0f;

  const float edst = gd_sqrt(gd_dist2(cen[0] - prt->pocket_center[0],
                                      cen[1] - prt->pocket_center[1],
                                      cen[2] - prt->pocket_center[2]));
  float t[W];
  float sums[GD_F_NTERM];
  for (int j = 0; j < W; ++j) t[j] = tot[j].vdw;
  sums[GD_F_VDW] = bfly(t);
  sums[GD_F_ELE] = bfly(eele_tot) * c->sc.inv_el1;   /* the loop sums el1 g */
  for (int j = 0; j < W; ++j) t[j] = tot[j].pmf;
  sums[GD_F_PMF] = bfly(t);
  sums[GD_F_HPC] = ehpc;
  for (int j = 0; j < W; ++j) t[j] = tot[j].hdb;
  sums[GD_F_HDB] = bfly(t);
  sums[GD_F_DST] = edst;
  for (int j = 0; j < W; ++j) t[j] = tot[j].psp;
  sums[GD_F_PSP] = bfly(t);
  sums[GD_F_KDE] = ekde;
  sums[GD_F_LHM] = elhm;
  gd_finalize(sums, (float) lna, ep->a_para, ep->b_para, ep->w, c->term_order, c->combine,
              out);
  if (cmcc)
    *cmcc = ref ? gd_mcc((float) tp, (float) fn, (float) fp, (float) tn) : 0.0f;
}

static void rp_place(const rp_ctx *c, int lig_conf, const float m[6], float *ax, float *ay,
                     float *az, float cen[3])
{
  const int lna = c->sz.lna;
  const size_t lo = (size_t) lig_conf * lna;
  float rot[9];
  gd_rotation(m, rot);
  for (int a = 0; a < lna; ++a)
    gd_transform(rot, m, c->lig->center, c->lig->x[lo + a], c->lig->y[lo + a],
                 c->lig->z[lo + a], &ax[a], &ay[a], &az[a]);
  for (int i = 0; i < 3; ++i)
    cen[i] = c->lig->center[i] + m[i];
}

/* Canonical FP32 energies of given poses, slot order, [n_pose * 10]. */
ORC_EXPORT void orc_replay_energy(const gd_ligand *lig, const gd_protein *prt,
                                  const gd_psp *psp, const gd_kde *kde,
                                  const gd_mcs *mcs, const gd_enepara *ep,
                                  gd_complex_size sz, int n_pose, const int *lig_conf,
                                  const int *prt_conf, const float *move, int term_order,
                                  float *out)
{
  rp_ctx c;
  rp_init(&c, lig, prt, psp, kde, mcs, ep, sz, term_order);
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < n_pose; ++i) {
    float ax[GD_MAX_LNA], ay[GD_MAX_LNA], az[GD_MAX_LNA], cen[3];
    rp_place(&c, lig_conf[i], move + (size_t) i * 6, ax, ay, az, cen);
    rp_energy(&c, prt_conf[i], ax, ay, az, cen, out + (size_t) i * 10, NULL, NULL);
  }
  rp_free(&c);
}

/* Canonical replay of gd_run: the same initialisation, MC steps, exchange
 * cadence and counter keys.  Fills results[n_rep] (energy_best_fp64 is
 * left zero).  The launches are those of orc_schedule (whole steps_per_dump
 * windows of whole launches, gpusingle.cu:8-20). */
/* The replay with RecordLigand_d (util_d.cu:72-94): before every MC step s
 * with s % record_stride == 0 the old state goes to
 * rec[r * n_rec + s / record_stride] (record_stride 0: no records). */
ORC_EXPORT int orc_replay_run_rec(const gd_ligand *lig, const gd_protein *prt,
                                  const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                                  const gd_enepara *ep, const gd_temp *temp,
                                  const gd_replica *replica, const gd_mcpara *mc,
                                  gd_complex_size sz, const gd_options *opt,
                                  gd_replica_result *res, int record_stride,
                                  gd_step_record *rec, int n_rec)
{
  if (orc_schedule(mc->steps_total, mc->steps_per_dump, mc->steps_per_exchange, NULL, 0) < 0)
    return -1;
  rp_ctx c;
  rp_init(&c, lig, prt, psp, kde, mcs, ep, sz, opt->term_order);
  c.combine = opt->combine;
  const int n = sz.n_rep;
  int *lc = (int *) malloc(sizeof(int) * n), *pc = (int *) malloc(sizeof(int) * n);
  int *it = (int *) malloc(sizeof(int) * n);
  for (int r = 0; r < n; ++r) {
    if (replica) {
      pc[r] = replica[r].idx_prt; it[r] = replica[r].idx_tmp; lc[r] = replica[r].idx_lig;
    } else {
      pc[r] = r / (sz.n_tmp * sz.n_lig);
      it[r] = (r / sz.n_lig) % sz.n_tmp;
      lc[r] = r % sz.n_lig;
    }
    memset(&res[r], 0, sizeof res[r]);
  }
  /* MCC reference contacts (InitRefMatrix_d, calcmcc_d.cu:57-83): replica
   * 0's ligand at movematrix 0 against replica 0's protein conformer */
  char *ref = NULL;
  float *cm = (float *) calloc((size_t) n, sizeof(float));
  if (opt->mcc) {
    ref = (char *) malloc((size_t) sz.lna * sz.pnp);
    const float m0[6] = {0, 0, 0, 0, 0, 0};
    float ax[GD_MAX_LNA], ay[GD_MAX_LNA], az[GD_MAX_LNA], cen[3];
    rp_place(&c, lc[0], m0, ax, ay, az, cen);
    const size_t po = (size_t) pc[0] * sz.pnp;
    for (int a = 0; a < sz.lna; ++a)
      for (int q = 0; q < sz.pnp; ++q) {
        const size_t i = po + q;
        const float d2 = gd_dist2(ax[a] - prt->x[i], ay[a] - prt->y[i], az[a] - prt->z[i]);
        const float d = d2 * gd_rsqrt(d2);
        ref[(size_t) a * sz.pnp + q] = (d - ep->pmf0[lig->t[a]][prt->t[i]]) <= 0.0f;
      }
  }
  /* initial state (MonteCarlo_Init_d) */
#pragma omp parallel for schedule(dynamic, 1)
  for (int r = 0; r < n; ++r) {
    float m[6] = {0, 0, 0, 0, 0, 0};
    if (opt->away)
      for (int k = 0; k < 6; ++k)
        m[k] = mc->move_scale[k] * 44.5f + m[k];
    float ax[GD_MAX_LNA], ay[GD_MAX_LNA], az[GD_MAX_LNA], cen[3], e[10];
    rp_place(&c, lc[r], m, ax, ay, az, cen);
    rp_energy(&c, pc[r], ax, ay, az, cen, e, ref, &cm[r]);
    res[r].cmcc_final = cm[r];
    res[r].cmcc_best = cm[r];
    memcpy(res[r].move_final, m, sizeof m);
    memcpy(res[r].move_best, m, sizeof m);
    memcpy(res[r].energy_final, e, sizeof e);
    memcpy(res[r].energy_best, e, sizeof e);
    res[r].step_best = -1;
  }
  const int spe = mc->steps_per_exchange;
  const int n_launch = orc_schedule(mc->steps_total, mc->steps_per_dump, spe, NULL, 0);
  int *modes = (int *) malloc(sizeof(int) * (n_launch > 0 ? n_launch : 1));
  if (n_launch > 0)
    orc_schedule(mc->steps_total, mc->steps_per_dump, spe, modes, n_launch);
  for (int x = 0; x < n_launch; ++x) {
    const int s0 = x * spe, s1 = s0 + spe;
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 0; r < n; ++r) {
      gd_replica_result *R = &res[r];
      const float mb = temp[it[r]].minus_beta;
      float ax[GD_MAX_LNA], ay[GD_MAX_LNA], az[GD_MAX_LNA], cen[3], en[10], mn[6];
      for (int s = s0; s < s1; ++s) {
        if (record_stride > 0 && s % record_stride == 0 && s / record_stride < n_rec) {
          gd_step_record *q = &rec[(size_t) r * n_rec + s / record_stride];
          q->replica.idx_rep = r;
          q->replica.idx_prt = pc[r];
          q->replica.idx_tmp = it[r];
          q->replica.idx_lig = lc[r];
          memcpy(q->energy, R->energy_final, sizeof(float) * 10);
          q->energy[10] = R->cmcc_final;
          memcpy(q->movematrix, R->move_final, sizeof(float) * 6);
          q->step = s;
        }
        const gd_u32x4 ra = gd_philox4x32_10((uint32_t) r, (uint32_t) s, GD_CTR_MOVE_A, 0u,
                                             opt->seed, opt->complex_id);
        const gd_u32x4 rb = gd_philox4x32_10((uint32_t) r, (uint32_t) s, GD_CTR_MOVE_B, 0u,
                                             opt->seed, opt->complex_id);
        const uint32_t w[8] = {ra.v[0], ra.v[1], ra.v[2], ra.v[3],
                               rb.v[0], rb.v[1], rb.v[2], rb.v[3]};
        for (int k = 0; k < 6; ++k)
          mn[k] = mc->move_scale[k] * gd_perturbation(opt->fixed_draws, w[k]) +
                  R->move_final[k];
        rp_place(&c, lc[r], mn, ax, ay, az, cen);
        float cmn = 0.0f;
        rp_energy(&c, pc[r], ax, ay, az, cen, en, ref, &cmn);
        if (gd_metropolis(en[GD_E_TOTAL], R->energy_final[GD_E_TOTAL], mb, gd_u01(w[6]),
                          gd_u01(w[7]), opt->fixed_draws, opt->reverse)) {
          R->acs_mc++;
          memcpy(R->move_final, mn, sizeof mn);
          memcpy(R->energy_final, en, sizeof en);
          R->cmcc_final = cmn;
          if (en[GD_E_TOTAL] < R->energy_best[GD_E_TOTAL]) {
            memcpy(R->move_best, mn, sizeof mn);
            memcpy(R->energy_best, en, sizeof en);
            R->cmcc_best = cmn;
            R->step_best = s;
          }
        }
      }
    }
    if (opt->exchange && sz.n_tmp > 1) {
      const int n_chain = sz.n_prt * sz.n_lig;
      float mbs[64], u[32];
      for (int t = 0; t < sz.n_tmp; ++t)
        mbs[t] = temp[t].minus_beta;
      for (int ch = 0; ch < n_chain; ++ch) {
        const int p = ch / sz.n_lig, l = ch % sz.n_lig;
        const int base = sz.n_tmp * sz.n_lig * p + l;
        int temps[64], acx[64];
        float ener[64];
        for (int t = 0; t < sz.n_tmp; ++t) {
          temps[t] = it[base + sz.n_lig * t];
          ener[t] = res[base + sz.n_lig * t].energy_final[GD_E_TOTAL];
          acx[t] = 0;
        }
        exchange_draws((uint32_t) ch, (uint32_t) x, sz.n_tmp / 2, opt->seed, opt->complex_id,
                       opt->fixed_draws, u);
        orc_exchange_chain(1, sz.n_tmp, mbs, temps, ener, modes[x], u, acx);
        for (int t = 0; t < sz.n_tmp; ++t) {
          it[base + sz.n_lig * t] = temps[t];
          res[base + sz.n_lig * t].acs_exchg += acx[t];
        }
      }
    }
  }
  free(modes);
  for (int r = 0; r < n; ++r)
    res[r].idx_tmp_final = it[r];
  free(ref);
  free(cm);
  free(lc);
  free(pc);
  free(it);
  rp_free(&c);
  return 0;
}

ORC_EXPORT int orc_replay_run(const gd_ligand *lig, const gd_protein *prt,
                              const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                              const gd_enepara *ep, const gd_temp *temp,
                              const gd_replica *replica, const gd_mcpara *mc,
                              gd_complex_size sz, const gd_options *opt,
                              gd_replica_result *res)
{
  return orc_replay_run_rec(lig, prt, psp, kde, mcs, ep, temp, replica, mc, sz, opt, res, 0,
                            NULL, 0);
}

/* ====================================================================== */
/* math/RNG probes for the known-answer and accuracy tests                 */
/* ====================================================================== */
ORC_EXPORT void orc_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                           uint32_t k1, uint32_t out[4])
{
  const gd_u32x4 r = gd_philox4x32_10(c0, c1, c2, c3, k0, k1);
  memcpy(out, r.v, sizeof r.v);
}

/* canonical Bayesian pdfs and total, for the closed-form tests */
ORC_EXPORT void orc_bayes_pdf(int family, int n, const float *x, float loc, float scale,
                              float *y)
{
  for (int i = 0; i < n; ++i)
    y[i] = gd_bayes_pdf(family, x[i], loc, scale);
}

ORC_EXPORT float orc_bayes_total(const float norm_file[9])
{
  return gd_bayes_total(norm_file);
}

/* The canonical move (gd_canon.h gd_rotation + gd_transform, what the GPU
 * runs) of n points (x, y, z) about centre c by movematrix m. */
ORC_EXPORT void orc_canon_move(const float m[6], const float *x, const float *y,
                               const float *z, int n, const float c[3], float *ox, float *oy,
                               float *oz, float cen[3])
{
  float rot[9];
  gd_rotation(m, rot);
  for (int a = 0; a < n; ++a)
    gd_transform(rot, m, c, x[a], y[a], z[a], &ox[a], &oy[a], &oz[a]);
  for (int i = 0; i < 3; ++i)
    cen[i] = c[i] + m[i];
}

ORC_EXPORT void orc_canon_math(int fn, int n, const float *x, float *y, float *y2)
{
  for (int i = 0; i < n; ++i) {
    if (fn == 0)
      y[i] = gd_exp(x[i]);
    else if (fn == 1)
      y[i] = gd_log(x[i]);
    else if (fn == 2)
      gd_sincos(x[i], &y[i], &y2[i]);
    else if (fn == 3)
      y[i] = gd_u01(gd_as_uint(x[i]));
    else if (fn == 4)
      y[i] = gd_rsqrt(x[i]);
    else if (fn == 5)
      y[i] = gd_rcp(x[i]);
    else if (fn == 6)
      y[i] = gd_exp_hot(x[i]);
  }
}
