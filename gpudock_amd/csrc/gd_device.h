/*
 * gd_device.h -- device-side data layout shared by the HIP kernels and the
 * host code that builds it.
 *
 * One "complex pack" holds everything the Monte Carlo kernels read for one
 * complex; it is built once per gd_run() call (or once per admitted job in
 * the batched and serving paths) and stays resident in HBM.
 */
#ifndef GD_DEVICE_H
#define GD_DEVICE_H

#include <stdint.h>
#include "gd_canon.h"

#define GD_WAVE 64

/* Per-complex constants and device pointers (kernel argument). */
typedef struct gd_pack {
  /* sizes */
  int n_lig, n_prt, n_tmp, n_rep;
  int lna, pnp, pnk, pos;
  int n_rows;            /* distinct ligand atom types = table rows */
  int n_res;             /* PSP residues */
  int term_order;
  int pnp_pad;           /* pnp rounded up to a multiple of 64 */

  /* pocket, per protein conformer: [n_prt][pnp_pad]; points pnp..pnp_pad-1
   * are padding that contributes exactly zero to every term */
  const float4 *pt_pos;  /* x, y, z, bits(type | flags | seq << 8) */
  const float2 *pt_eh;   /* ele, hpp */
  /* pair table, [n_rows][GD_ROW_STRIDE] */
  const gd_pair_param *tab;
  /* ligand: conformer coordinates [n_lig][lna] (x, y, z, 0) */
  const float4 *lig_xyz;
  const int *lig_row;    /* [lna] table row of each atom */
  const int *lig_type;   /* [lna] atom type */
  const float *lig_q;    /* [lna] charge */
  /* KDE points sorted by type, and per-row segments [n_rows + 1] */
  const float4 *kde;
  const int *kde_seg;    /* start of row r's points at kde_seg[2r], end kde_seg[2r+1] */
  const float *kde_inv_n;/* [n_rows] 1.0f / (float) segment length (0 if empty) */
  /* position restraints: targets per (m, atom) [pos][lna] (x, y, z, valid) */
  const float4 *mcs;
  const float *mcs_tcc;  /* [pos] */
  /* PSP table [GD_MAXTP2][n_res] */
  const float *psp;
  /* per type (24 entries) */
  const float *hpl0, *hpl1, *hpl2;
  const float *inv_hpl1;    /* 1.0f / hpl1 */
  /* temperatures */
  const float *minus_beta;  /* [n_tmp] */

  gd_pair_scalars sc;
  float kde2, kde3;
  float center[3];          /* ligand coord_orig.center */
  float pocket_center[3];   /* same for every protein conformer */
  float move_scale[6];
  float a_para[10], b_para[10], w[10];
} gd_pack;

#define GD_MAXTP1_DEV 30

/* Replica state in HBM, structure of arrays over [n_rep]. */
typedef struct gd_state {
  float *mv;        /* [n_rep][8]  movematrix_old (6 used) */
  float *e;         /* [n_rep][10] energy_old */
  float *mv_best;   /* [n_rep][8] */
  float *e_best;    /* [n_rep][10] */
  int *step_best;   /* [n_rep] */
  int *idx_tmp;     /* [n_rep] temperature index held by the slot */
  int *acs_mc;      /* [n_rep] */
  int *acs_x;       /* [n_rep] */
} gd_state;

/* One workgroup task: up to waves-per-group replicas of one protein
 * conformer. */
typedef struct gd_task {
  int prt;
  int r_begin;
  int r_count;
  int pad;
} gd_task;

#endif /* GD_DEVICE_H */
