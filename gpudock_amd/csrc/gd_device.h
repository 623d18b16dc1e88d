/*
 * gd_device.h -- device-side data layout shared by the HIP kernels and the
 * host code that builds it.
 *
 * A complex (one docking job: ligand ensemble, protein ensemble, pocket
 * tables, KDE, MCS, PSP, temperature ladder and its replica state) is
 * described by one gd_pack that lives in HBM.  The MC, exchange and
 * advance kernels take an array of packs, so one launch serves one complex
 * (the plain Run() path) or many (batched and served jobs) alike.
 */
#ifndef GD_DEVICE_H
#define GD_DEVICE_H

#include <stdint.h>
#include "gd_canon.h"

#define GD_WAVE 64
/* bytes of one staged 64-point pocket chunk: three 8-byte planes */
#define GD_LDS_CHUNK_BYTES (3 * 8 * GD_WAVE)

/* Replica state of one complex in HBM, structure of arrays over its n_rep
 * replicas (local replica index r). */
typedef struct gd_state {
  float *mv;        /* [n_rep][8]  movematrix_old (6 used) */
  float *e;         /* [n_rep][10] energy_old */
  float *mv_best;   /* [n_rep][8] */
  float *e_best;    /* [n_rep][10] */
  int *step_best;   /* [n_rep] */
  int *idx_tmp;     /* [n_rep] temperature index held by the slot */
  int *acs_mc;      /* [n_rep] */
  int *acs_x;       /* [n_rep] */
  float *cmcc;      /* [n_rep] energy_old.cmcc (0 without MCC) */
  float *cmcc_best; /* [n_rep] cmcc of the best pose */
} gd_state;

/* Per ligand atom, what the pair loop reads once per atom: one 32-byte
 * record, one scalar load. */
typedef struct gd_atom_rec {
  int row;                       /* table row of the atom's type */
  float q;                       /* charge */
  int hdb;                       /* 1 if that row has an H-bond pair */
  float hpl0, inv_hpl1, hpl2;    /* hydrophobic restraint of the atom's type */
  int pad0, pad1;
} gd_atom_rec;

/* One complex. */
typedef struct gd_pack {
  /* sizes */
  int n_lig, n_prt, n_tmp, n_rep;
  int lna, pnp, pnk, pos;
  int n_rows;            /* distinct ligand atom types = table rows */
  int n_res;             /* PSP residues */
  int pad_psp;
  int term_order;
  int combine;           /* total: 0 weighted sum, 1 Bayesian */
  int pnp_pad;           /* pnp rounded up to a multiple of 64 */
  int kde_n_lo, kde_n_hi;/* (atom, KDE point) pairs of atoms < 64 and >= 64 */

  /* run control */
  int steps_run;         /* MC steps the job runs: whole windows of whole
                            launches (gpusingle.cu:8-20) */
  int spe;               /* steps per exchange = MC steps per cycle */
  int lpw;               /* cycles (launches) per steps_per_dump window: the
                            exchange mode restarts at 0 in every window */
  int exchange;          /* temperature swaps on/off */
  int away;              /* start 44.5 * move_scale away (IS_AWAY) */
  int fixed_draws;       /* IS_RANDOM 0: every draw is 20 (gd_perturbation) */
  int reverse;           /* IS_REVERSE 1: the reversed Metropolis test */
  uint32_t seed, cid;    /* Philox key */
  int *step;             /* [2] device counters: next MC step, exchanges done */

  /* LDS carve-up for this complex (bytes); the pocket, when staged, starts
   * at 0 in 64-point chunks of GD_LDS_CHUNK_BYTES (gd_stage) */
  int off_tab, off_kde, lds_bytes;
  int kde_in_lds;
  int pocket_in_lds;     /* 0: pocket points read through the caches */
  /* launches that split a replica over several wavefronts (one replica per
   * workgroup) also stage, after the KDE points, the (atom, KDE point) pair
   * list, the position-restraint targets and the PSP rows of the staged
   * protein conformer, when split_extra = 1 (they fit); split_lds_bytes is
   * then the staging size of such launches (else lds_bytes) */
  int split_extra;
  int off_pairs, off_mcs, off_psp, split_lds_bytes;

  /* pocket, per protein conformer: [n_prt][pnp_pad]; points pnp..pnp_pad-1
   * are padding that contributes exactly zero to every term */
  const float4 *pt_pos;  /* x, y, z, bits(type << GD_PT_TYPE_SHIFT) */
  const float2 *pt_eh;   /* ele, hpp (0 outside the hydrophobic class) */
  /* pair table, [n_rows][GD_ROW_STRIDE] */
  const gd_pair_param *tab;
  /* ligand: conformer coordinates [n_lig][lna] (x, y, z, 0) */
  const float4 *lig_xyz;
  const int *lig_row;    /* [lna] table row of each atom */
  const int *lig_type;   /* [lna] atom type */
  const float *lig_q;    /* [lna] charge */
  /* KDE points grouped by table row; row r spans [kde_seg[2r], kde_seg[2r+1]);
   * (x, y, z, 1 / that segment's length) */
  const float4 *kde;
  /* (atom, KDE point) pairs of a pose: atoms 0..63, each atom's segment in
   * order, padded with zeros to a multiple of 64; then atoms 64.. the same
   * way.  Entry = point << 16 | (atom & 63) << 2 */
  const uint32_t *kde_pairs;
  const int *kde_seg;
  const float *kde_inv_n;/* [n_rows] 1.0f / (float) segment length (0 if empty) */
  /* position restraints: targets per (m, atom) [pos][lna] (x, y, z, valid) */
  const float4 *mcs;
  const float *mcs_tcc;  /* [pos] */
  /* PSP value per (protein conformer, table row, point):
   * [n_prt][n_rows][pnp_pad], 0 outside class 2 and for padding */
  const float *psp_pt;
  /* [n_prt][2]: kh = chunks of 64 points, from the first, that hold every
   * point of the hydrophobic class (the canonical order puts them first);
   * kp >= kh = chunks from the first that hold every class-2 point, the
   * only points with a PSP value */
  const int *chunk_bounds;
  /* [n_rows] 1 if some point type pairs with the row's ligand type through
   * an H-bond term (hdb0 > 0.1) */
  const int *row_hdb;
  const gd_atom_rec *atom; /* [lna] */
  /* per ligand type (24 entries each) */
  const float *hpl0, *hpl1, *hpl2;
  const float *inv_hpl1; /* 1.0f / hpl1 */
  /* temperatures */
  const float *minus_beta;  /* [n_tmp] */
  /* replicas */
  const int *rep_lig;    /* [n_rep] ligand conformer */
  const int *rep_prt;    /* [n_rep] protein conformer */
  gd_state S;
  /* per-step records (RecordLigand_d): NULL, or [n_rep][rec_per_window]
   * records of GD_REC_WORDS words for the steps of the current
   * steps_per_dump window with step % rec_stride == 0 */
  uint32_t *rec;
  int rec_stride, rec_window, rec_per_window, rec_pad;
  /* MCC (IS_CALCU_MCC): reference contacts of replica 0's ligand conformer
   * at its start pose, per protein conformer in its canonical point order:
   * bit j of mcc_ref[(c * lna + l) * nchunk + k] = contact of atom l with
   * point k * 64 + j; NULL = no MCC */
  const unsigned long long *mcc_ref;

  gd_pair_scalars sc;
  float kde2, kde3;
  float center[3];          /* ligand coord_orig.center */
  float pocket_center[3];
  float move_scale[6];
  float a_para[10], b_para[10], w[10];
} gd_pack;

/* words of one step record: replica (4), energy (10 + MCC), movematrix (6),
 * step -- the layout of gd_step_record */
#define GD_REC_WORDS 22

/* One workgroup task: up to waves-per-group replicas of one protein
 * conformer of one complex. */
typedef struct gd_task {
  int cx;
  int prt;
  int r_begin;
  int r_count;
} gd_task;

/* One temperature chain: (complex, protein conformer * n_lig + ligand conformer). */
typedef struct gd_chain {
  int cx;
  int c;
} gd_chain;

/* Per-wave LDS slot holding the replica's uniform state between steps:
 * m_old[6] e_old[10] m_new[6] e_new[10] cmcc_old */
#define GD_WSLOT_FLOATS 36
#define GD_WS_MOLD 0
#define GD_WS_EOLD 6
#define GD_WS_MNEW 16
#define GD_WS_ENEW 22
#define GD_WS_CMCC 32     /* energy_old.cmcc */

/* A replica split over several wavefronts keeps the partial sums of its
 * summation blocks in LDS beside its slot, with three placed poses and
 * the conformer: this many bytes (atom and KDE blocks, restraint values,
 * poses, conformer, sums);
 * complexes with more position restraints than GD_SPLIT_MAX_POS run one
 * wavefront per replica. */
#define GD_SPLIT_MAX_POS 64
#define GD_SPLIT_AREA_BYTES \
  ((GD_MAX_BLOCKS * 6 * 64 + GD_MAX_BLOCKS * 4 + GD_SPLIT_MAX_POS + 3 * (6 * 64 + 16) + \
    6 * 64 + 16) * 4)

#endif /* GD_DEVICE_H */
