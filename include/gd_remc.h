/*
 * gd_remc.h -- C ABI of the MI355X replica-exchange Monte Carlo docking core.
 *
 * The plain per-complex entry point gd_run() replaces the reference's
 *
 *   void Run (const Ligand *, const Protein *, const Psp *, const Kde *,
 *             const Mcs *, const EnePara *, const Temp *, const Replica *,
 *             const McPara *, McLog *, const ComplexSize);
 *
 * (gpudocksm-rem-1.2/src/run.h:6-17, body run.cu:40-352).  The structs below
 * carry the same fields as the reference's prepared host structs
 * (gpudocksm-rem-1.2/src/dock.h), with pointer + count in place of the
 * fixed capacities of size.h, so that every size the reference can hold
 * and the larger configurations beyond MAXLIG/MAXREP fit.
 *
 * Conventions
 *   - Plain pointers to host memory; nothing here takes ownership.
 *   - Every function returns GD_OK (0) or a negative gd_status; the text of
 *     the last error of the calling thread is gd_last_error().  Nothing
 *     calls exit(), unlike the reference (cutil_yeah.h:26-42).
 *   - Energies are float[GD_E_N] in the reference's Energy.e[] slot order
 *     (dock.h:100-115): vdw ele pmf psp hdb hpc kde lhm dst total.
 *   - Weights, normalisation a/b and the parameter files are in file order
 *     vdw ele pmf hpc hdb dst psp kde lhm (see gd_options.term_order).
 */
#ifndef GD_REMC_H
#define GD_REMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* capacities of the type tables (size.h:52-70) */
#define GD_MAXTP1 30   /* protein effective-point types */
#define GD_MAXTP2 24   /* ligand atom types */
#define GD_MAXTP3 50   /* electrostatic point classes */
#define GD_MAXTP4 20   /* residue types */
#define GD_MAXWEI 10   /* energy terms + total */
#define GD_MAX_LNA 128 /* ligand atoms per conformer held by one wavefront */
#define GD_MCS_INVALID_COORD 728492.0f  /* size.h:97 */

enum gd_status {
  GD_OK = 0,
  GD_ERR_ARG = -1,       /* invalid argument or size */
  GD_ERR_HIP = -2,       /* HIP runtime error */
  GD_ERR_NOMEM = -3,     /* device allocation failed / admission refused */
  GD_ERR_NODEV = -4,     /* no usable GPU */
  GD_ERR_IO = -5,        /* file could not be read or parsed */
  GD_ERR_CANCELLED = -6, /* job cancelled (serving) */
  GD_ERR_TIMEOUT = -7    /* wait timed out; the job is still queued or running */
};

enum gd_energy_slot {
  GD_E_VDW_SLOT = 0, GD_E_ELE_SLOT, GD_E_PMF_SLOT, GD_E_PSP_SLOT, GD_E_HDB_SLOT,
  GD_E_HPC_SLOT, GD_E_KDE_SLOT, GD_E_LHM_SLOT, GD_E_DST_SLOT, GD_E_TOTAL_SLOT
};

enum gd_combine {
  GD_COMBINE_LINEAR = 0,  /* sum of w_k e_k (calcenergy_d.cu:349-358) */
  GD_COMBINE_BAYES = 1    /* IS_BAYE: sum log10 p_low(e_k) - sum log10 p_high(e_k)
                             over fitted per-term pdfs (gpudocksm-bayes-1.0
                             calcenergy_d.cu:315-367, distribution.h:21-77) */
};

enum gd_term_order {
  GD_TERM_ORDER_FILE = 0, /* parameter k applies to file-order term k */
  GD_TERM_ORDER_SLOT = 1  /* rem-1.2 behaviour: parameter k applies to slot k
                             (calcenergy_d.cu:301-309,349-358) */
};

/* ComplexSize, dock.h:77-90 */
typedef struct gd_complex_size {
  int n_lig, n_prt, n_tmp, n_rep;   /* n_rep = n_lig * n_prt * n_tmp */
  int lna, pnp, pnk, pos;
} gd_complex_size;

/* Ligand ensemble: the used fields of Ligand (dock.h:191-216) after
 * OptimizeLigand (util.C:199-226).  Coordinates are centroid-relative
 * (coord_orig), conformer-major: atom a of conformer i at [i * lna + a].
 * t, c and n are shared by all conformers, as the loader makes them
 * (load.C:296-330). */
typedef struct gd_ligand {
  const float *x, *y, *z;  /* [n_lig * lna] */
  const int *t;            /* [lna] atom type, 0 .. GD_MAXTP2-1 */
  const float *c;          /* [lna] partial charge */
  const int *n;            /* [lna] atom number; MCS field index is n + 1 */
  float center[3];         /* coord_orig.center = pocket centre */
} gd_ligand;

/* Protein ensemble: Protein (dock.h:254-274) after OptimizeProtein
 * (util.C:252-329).  Every array is [n_prt * pnp], conformer-major. */
typedef struct gd_protein {
  const float *x, *y, *z;
  const int *t;                 /* point type, 0 .. GD_MAXTP1-1 */
  const int *c;                 /* point class */
  const float *ele;             /* ele_tab[t == 0 ? d + 30 : t] */
  const int *seq3r;             /* residue index into the PSP table */
  const int *c0_and_d12_or_c2;  /* (c == 0 && d == 12) || c == 2 */
  const float *hpp;             /* hpp_tab[d] */
  float pocket_center[3];
} gd_protein;

/* Psp (dock.h:289-295): psp[t * n_res + seq3r] for ligand type t. */
typedef struct gd_psp {
  const float *psp;
  int n_res;
} gd_psp;

/* Kde (dock.h:317-325), [pnk] each */
typedef struct gd_kde {
  const float *x, *y, *z;
  const int *t;
} gd_kde;

/* Mcs[pos] (dock.h:347-354): field f of restraint m at [m * n_field + f],
 * GD_MCS_INVALID_COORD where absent (load.C:1014-1020). */
typedef struct gd_mcs {
  const float *x, *y, *z;
  const float *tcc;   /* [pos] */
  int n_field;
} gd_mcs;

/* EnePara (dock.h:386-421), the tables OptimizeEnepara builds
 * (util.C:404-458), indexed [ligand type][point type]. */
typedef struct gd_enepara {
  float p1a[GD_MAXTP2][GD_MAXTP1];
  float p2a[GD_MAXTP2][GD_MAXTP1];
  float lj0, lj1;
  float el0, el1;
  float ele[GD_MAXTP3];
  float a1, b1;
  float pmf0[GD_MAXTP2][GD_MAXTP1];
  float pmf1[GD_MAXTP2][GD_MAXTP1];
  float hdb0[GD_MAXTP2][GD_MAXTP1];
  float hdb1[GD_MAXTP2][GD_MAXTP1];
  float hdb2[GD_MAXTP2][GD_MAXTP1];
  float hpp[GD_MAXTP4];
  float hpl0[GD_MAXTP2], hpl1[GD_MAXTP2], hpl2[GD_MAXTP2];
  float kde2, kde3;
  float w[GD_MAXWEI];
  float a_para[GD_MAXWEI];
  float b_para[GD_MAXWEI];
} gd_enepara;

/* EnePara0 (dock.h:363-380): raw parameter-file values, [point][ligand]. */
typedef struct gd_enepara0 {
  float vdw[GD_MAXTP1][GD_MAXTP2][2];
  float ele[GD_MAXTP3];
  float pmf[GD_MAXTP1][GD_MAXTP2][2];
  float hpp[GD_MAXTP4];
  float hpl[GD_MAXTP2][2];
  float hdb[GD_MAXTP1][GD_MAXTP2][2];
  float lj[3];
  float el[2];
  float kde;
  float w[GD_MAXWEI];
  float a_para[GD_MAXWEI];
  float b_para[GD_MAXWEI];
} gd_enepara0;

/* Temp (dock.h:429-434) */
typedef struct gd_temp {
  float t;
  float minus_beta;
  int order;
} gd_temp;

/* Replica (dock.h:443-450); flat index n_tmp*n_lig*p + n_lig*t + l */
typedef struct gd_replica {
  int idx_rep, idx_prt, idx_tmp, idx_lig;
} gd_replica;

/* McPara (dock.h:460-470) without the output paths.  The run follows the
 * reference's launch loop (gpusingle.cu:8-20): windows of steps_per_dump
 * steps, each run as whole launches of steps_per_exchange MC steps with an
 * exchange after each, the exchange mode (0: temperature pairs (0,1),
 * (2,3), ...; 1: (1,2), ...) alternating from 0 in every window; the last,
 * partial window runs whole.  Steps run = ceil(steps_total / steps_per_dump)
 * * ceil(steps_per_dump / steps_per_exchange) * steps_per_exchange.  Both
 * step sizes must be >= 1 (GD_ERR_ARG otherwise). */
typedef struct gd_mcpara {
  int steps_total;
  int steps_per_dump;
  int steps_per_exchange;
  float move_scale[6];   /* translation x y z (A), rotation x y z (rad):
                            uniform half-widths (move_d.cu:79-84) */
} gd_mcpara;

/* McLog (dock.h:472-482).  acs_mc / acs_temp_exchg are [n_rep] arrays the
 * caller provides (or NULL); they are accumulated into, as run.cu:284-296
 * does.  As the reference takes them (gpusingle.cu:10,31,41): t0 = seconds
 * of the MC loop (the cycle launches, after setup, upload and the initial
 * energies), t1 = t0 plus the trajectory dumps; t2 (unused there) = the
 * whole gd_run call, setup to results on the host (FP64 re-score, copies). */
typedef struct gd_mclog {
  double t0, t1, t2;
  long long ac_mc;
  int *acs_mc;
  long long ac_temp_exchg;
  int *acs_temp_exchg;
} gd_mclog;

/* Settings the reference keeps in toggle.h or time-seeds; no counterpart
 * argument in Run(). */
typedef struct gd_options {
  uint32_t seed;        /* Philox key word 0 (the reference seeds from time) */
  uint32_t complex_id;  /* Philox key word 1 */
  int exchange;         /* IS_EXCHANGE (toggle.h:12): swap temperatures */
  int away;             /* IS_AWAY (toggle.h:15): start 44.5*move_scale away */
  int term_order;       /* enum gd_term_order */
  int device;           /* HIP device ordinal */
  int record_stride;    /* IS_OUTPUT (toggle.h): record the state before every
                           record_stride-th MC step; 0 = no records */
  const char *record_prefix; /* gd_run writes <prefix>_%04d.gdtraj per
                                steps_per_dump window (gpusingle.cu:26-34);
                                NULL = keep records on the device only */
  int mcc;              /* IS_CALCU_MCC (toggle.h:18): Matthews correlation of
                           each pose's contacts against the contacts of
                           replica 0's ligand at its start pose */
  int combine;          /* enum gd_combine: the total energy */
  int fixed_draws;      /* IS_RANDOM == 0 (toggle.h:9) when 1: the reference's
                           deterministic validation mode -- every draw is 20,
                           so each move adds 39 x move_scale, every MC step is
                           accepted and a temperature pair swaps only when its
                           probability exceeds 20 (util_d.cu:106-108) */
  int reverse;          /* IS_REVERSE (toggle.h:24) when 1: the Metropolis test
                           is taken on a second draw with the exponent's sign
                           flipped (accept_d.cu:24-27) */
} gd_options;

/* LigRecordSingleStep (dock.h:487-494): the state of one replica slot
 * before MC step `step` (RecordLigand_d, util_d.cu:72-94).  88 bytes. */
typedef struct gd_step_record {
  gd_replica replica;          /* idx_rep, idx_prt, idx_tmp held then, idx_lig */
  float energy[GD_MAXWEI + 1]; /* energy_old e[10] (slot order) + cmcc (0
                                  without gd_options.mcc) */
  float movematrix[6];
  int step;
} gd_step_record;

/* .gdtraj file: this header, then n_rep x n_step records, replica-major
 * (the reference's LigRecord[n_rep].step[] order); the reference writes
 * the same records as HDF5 (hdf5io.C:10-30), which this library does not
 * link. */
typedef struct gd_traj_header {
  char magic[8];               /* "GDTRAJ01" */
  int32_t n_rep;
  int32_t n_step;              /* records per replica in this file */
  int32_t record_bytes;        /* sizeof(gd_step_record) = 88 */
  int32_t first_step;          /* MC step of each replica's first record */
  int32_t stride;              /* MC steps between records */
  int32_t reserved[3];
} gd_traj_header;

/* Per-replica outcome (no counterpart: the reference leaves the final
 * state on the device, montecarlo_d.cu:98-102, and never extracts a best
 * pose, README.org:492-505). */
typedef struct gd_replica_result {
  float move_final[6];
  float energy_final[GD_MAXWEI];
  float move_best[6];
  float energy_best[GD_MAXWEI];
  double energy_best_fp64[GD_MAXWEI]; /* FP64 re-score of move_best */
  int step_best;        /* MC step that produced move_best, -1 = start */
  int idx_tmp_final;    /* temperature index held at the end */
  int acs_mc;           /* accepted MC moves */
  int acs_exchg;        /* accepted temperature swaps involving this slot */
  float cmcc_final;     /* MCC of the final pose (gd_options.mcc; else 0) */
  float cmcc_best;      /* MCC of the best pose */
} gd_replica_result;

/* ---------------------------------------------------------------------- */
/* Plain per-complex run: drop-in for Run() (run.h:6-17).                   */
/* results: [n_rep] or NULL.  opt: NULL = defaults (seed 1, complex 0,      */
/* exchange on, away off, file term order, device 0).                       */
/* ---------------------------------------------------------------------- */
int gd_run(const gd_ligand *lig, const gd_protein *prt, const gd_psp *psp,
           const gd_kde *kde, const gd_mcs *mcs, const gd_enepara *enepara,
           const gd_temp *temp, const gd_replica *replica,
           const gd_mcpara *mcpara, gd_mclog *mclog, gd_complex_size size,
           const gd_options *opt, gd_replica_result *results);

/* Energy of given poses on the GPU (CalcEnergy_d, calcenergy_d.cu:22-374,
 * evaluated for caller-chosen movematrices).  pose i uses ligand conformer
 * lig_conf[i], protein conformer prt_conf[i] and movematrix move[6*i..].
 * energy: [n_pose * GD_MAXWEI] FP32; energy_fp64: same in FP64, or NULL. */
int gd_score_poses(const gd_ligand *lig, const gd_protein *prt,
                   const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                   const gd_enepara *enepara, gd_complex_size size,
                   int n_pose, const int *lig_conf, const int *prt_conf,
                   const float *move, int term_order, int device,
                   float *energy, double *energy_fp64);

/* gd_score_poses without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  The complex's device image, the pose
 * arrays and the energies are staged in a static workspace of 64 MiB that
 * each GPU the library runs on holds until the process ends; a call whose
 * complex and poses need more returns GD_ERR_ARG (split the poses, or use
 * gd_score_poses).  The workspace keeps the last call's contents; every
 * byte a call reads it has written first.  Calls on one GPU share that
 * workspace: calls that overlap on one GPU (from several threads) race.
 * Work is ordered on the default stream; the call returns when it is done.
 * Results are those of gd_score_poses, bit for bit. */
int gd_score_poses_noalloc(const gd_ligand *lig, const gd_protein *prt,
                           const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                           const gd_enepara *enepara, gd_complex_size size,
                           int n_pose, const int *lig_conf, const int *prt_conf,
                           const float *move, int term_order, int device,
                           float *energy, double *energy_fp64);

/* ---------------------------------------------------------------------- */
/* Session: the same run split into steps, with the complex and the replica */
/* state resident on the GPU between calls (gd_run is create + init +       */
/* cycles + results + destroy).  No counterpart in the reference, whose     */
/* Run() does all of it in one call (run.cu:40-352).                        */
/* ---------------------------------------------------------------------- */
typedef struct gd_session gd_session;

int gd_session_create(const gd_ligand *lig, const gd_protein *prt,
                      const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                      const gd_enepara *enepara, const gd_temp *temp,
                      const gd_replica *replica, const gd_mcpara *mcpara,
                      gd_complex_size size, const gd_options *opt,
                      gd_session **out);
/* initial energies (MonteCarlo_Init_d, montecarlo_d.cu:12-55) */
int gd_session_init(gd_session *s);
/* n_cycles x (steps_per_exchange MC steps, then one exchange); queued on
 * the session's stream, returns before the GPU finishes */
int gd_session_cycles(gd_session *s, int n_cycles);
int gd_session_sync(gd_session *s);
/* MC steps done so far */
long long gd_session_steps_done(const gd_session *s);
/* copies the state back; mclog and results as in gd_run (either NULL) */
int gd_session_results(gd_session *s, gd_mclog *mclog,
                       gd_replica_result *results);
/* kernel timing with HIP events on the session stream (off by default) */
int gd_session_set_timing(gd_session *s, int on);
int gd_session_timing(gd_session *s, double *mc_ms, int *mc_launches,
                      double *x_ms, int *x_launches);
/* per-step records (gd_options.record_stride > 0) of the current
 * steps_per_dump window, the device buffer the reference copies out per dump
 * (gpusingle.cu:26-34): *first_step and *n_step records per replica, then,
 * if dst is not NULL, [n_rep][n_step] records into dst */
int gd_session_records(gd_session *s, int *first_step, int *n_step,
                       gd_step_record *dst);
/* LDS bytes per workgroup and whether KDE points are staged in LDS */
int gd_session_layout(const gd_session *s, int *lds_bytes, int *kde_in_lds,
                      int *n_tasks, int *waves_per_group);
void gd_session_destroy(gd_session *s);

/* ---------------------------------------------------------------------- */
/* Engine: many complexes resident on one GPU, advanced together.  One      */
/* cycle = one MC launch over every replica of every complex (each complex  */
/* takes min(steps_per_exchange, steps left) steps), one exchange launch    */
/* over every temperature chain, one launch that advances the per-complex   */
/* step counters held in device memory; a complex whose steps are done is   */
/* left untouched.  A complex's results are identical whether it runs alone */
/* or with others, and identical to gd_run on the same inputs.  No          */
/* counterpart in the reference, which runs one complex per process         */
/* (main.C:11-119 -> Run, run.cu:40-352).                                   */
/* ---------------------------------------------------------------------- */
typedef struct gd_engine gd_engine;

/* A job = the arguments of one gd_run call (pointers are read at admission
 * only).  opt NULL = the gd_run defaults. */
typedef struct gd_job {
  const gd_ligand *lig;
  const gd_protein *prt;
  const gd_psp *psp;
  const gd_kde *kde;
  const gd_mcs *mcs;
  const gd_enepara *enepara;
  const gd_temp *temp;
  const gd_replica *replica;   /* NULL = the SetReplica layout */
  const gd_mcpara *mcpara;
  gd_complex_size size;
  const gd_options *opt;
} gd_job;

enum gd_kde_policy {
  GD_KDE_STAGE_AUTO = 0,   /* KDE points in LDS when they fit */
  GD_KDE_STAGE_NEVER = 1   /* KDE points read from HBM (L2) */
};

typedef struct gd_engine_config {
  int device;
  size_t arena_bytes;  /* device memory reserved up front; 0 = 1 GiB */
  int wpg;             /* waves (replicas) per workgroup: 4, 8 (default), 16 */
  int persistent;      /* 1: resident workgroups pull tasks from a queue */
  int use_graph;       /* 1: cycles replay one captured hipGraph */
  int kde_policy;      /* enum gd_kde_policy */
  size_t staging_bytes;/* pinned host memory for uploads; 0 = 64 MiB */
  int occ;             /* MC kernel register budget: 0 (compiler's choice),
                          6 or 8 waves per SIMD */
  int split;           /* wavefronts per replica in the MC cycles: 0 = chosen
                          from the resident replica count (several when they
                          are fewer than the device's SIMDs), 1 = one, or 4, 8,
                          16 (one replica per workgroup); the energies are the
                          same bits for every choice.  One wavefront per
                          replica is used whatever is asked when a resident
                          complex has more than 64 position restraints, its
                          pair list, restraint targets and PSP rows do not fit
                          in LDS beside the split area, or the KDE points are
                          kept in HBM */
  int prio;            /* 0-3: issue priority of the MC launches' wavefronts in
                          their SIMD against other streams' launches on the
                          device (s_setprio); 0 = the default */
} gd_engine_config;

#define GD_TUNE_MAX 32
typedef struct gd_tune_report {
  int n;               /* launch shapes tried */
  int wpg[GD_TUNE_MAX];
  int persistent[GD_TUNE_MAX];
  int kde_global[GD_TUNE_MAX];
  int occ[GD_TUNE_MAX];
  double ms[GD_TUNE_MAX];   /* the MC launches of one cycle */
  int split[GD_TUNE_MAX];   /* wavefronts per replica */
  int chosen_wpg, chosen_persistent, chosen_kde_global, chosen_occ, chosen_split;
} gd_tune_report;

typedef struct gd_engine_layout_info {
  int stage_bytes;     /* LDS staged per workgroup */
  int wpg, persistent;
  int n_tasks, n_chains, n_complexes;
  int grid;            /* workgroups per cycle (all MC launches) */
  int resident_groups; /* workgroups the device holds at once */
  int kde_in_lds;      /* 1 if every complex stages its KDE points */
  size_t arena_used, arena_size;
  int n_classes;       /* MC launches per cycle (occupancy classes) */
  int kde_global;      /* 1: KDE points read from HBM for every complex */
  int occ;             /* MC kernel register budget (waves per SIMD, 0 = default) */
  int split;           /* wavefronts per replica in the MC cycles */
} gd_engine_layout_info;

int gd_engine_create(const gd_engine_config *cfg, gd_engine **out);
void gd_engine_destroy(gd_engine *e);
/* device bytes a job occupies once admitted (HBM admission control) */
int gd_job_footprint(const gd_job *job, int device, size_t *bytes);
/* admit: upload + initial energies; GD_ERR_NOMEM if the arena is full */
int gd_engine_add(gd_engine *e, const gd_job *job, int *handle);
/* retire (waits for queued cycles first) */
int gd_engine_remove(gd_engine *e, int handle);
/* queue n cycles; returns before the GPU finishes */
int gd_engine_cycles(gd_engine *e, int n);
int gd_engine_sync(gd_engine *e);
/* cycles still needed by the job after the queued ones */
int gd_engine_cycles_left(gd_engine *e, int handle);
/* results of one job, as gd_run (waits for queued cycles) */
int gd_engine_results(gd_engine *e, int handle, gd_mclog *mclog,
                      gd_replica_result *results);
/* time each launch shape on the resident jobs, keep the fastest */
int gd_engine_autotune(gd_engine *e, gd_tune_report *rep);
int gd_engine_set_timing(gd_engine *e, int on);
int gd_engine_timing(gd_engine *e, double *mc_ms, int *mc_launches,
                     double *x_ms, int *x_launches);
int gd_engine_layout(gd_engine *e, gd_engine_layout_info *info);

/* Batch: admit every job, autotune (optional), run until every job has
 * done its steps, collect.  results[i]: [jobs[i].size.n_rep] or NULL;
 * results and mclogs may be NULL. */
typedef struct gd_batch_config {
  int device;
  int wpg;             /* 0 = 8 */
  int persistent;
  int no_graph;        /* 1: launch the cycle kernels one by one */
  int autotune;
  int kde_policy;
  int occ;             /* MC kernel register budget: 0, 6 or 8 waves per SIMD */
  int split;           /* wavefronts per replica: as gd_engine_config.split */
} gd_batch_config;

typedef struct gd_batch_report {
  double admit_s, tune_s, run_s, results_s;
  int cycles;
  gd_tune_report tune;
  gd_engine_layout_info layout;
} gd_batch_report;

int gd_batch_run(int n_jobs, const gd_job *jobs, const gd_batch_config *cfg,
                 gd_replica_result **results, gd_mclog *mclogs,
                 gd_batch_report *report);

/* ---------------------------------------------------------------------- */
/* Serving: a request loop over one engine.  Jobs are submitted from any    */
/* thread at any time (the inputs are copied at submission), wait in FIFO  */
/* order until their HBM footprint fits, then run in the same cycles as     */
/* every other resident job (continuous batching); a finished job is read   */
/* back asynchronously and retired, freeing its memory for the next.  Each  */
/* job's results are identical to gd_run of it alone.  No counterpart in    */
/* the reference (one complex per process, main.C:11-119).                  */
/* ---------------------------------------------------------------------- */
typedef struct gd_server gd_server;

#define GD_SERVER_MAX_LANES 4

/* Launch shape: on a mixed stream of small and medium complexes, wpg 16,
 * persistent 0, occ 8 and kde_policy GD_KDE_STAGE_NEVER served a 1,000-job burst
 * 3 % faster than wpg 8, persistent 1, occ 6, and at a load below capacity
 * the jobs' p50 latency was 2.4 times lower (fewer, larger workgroup tasks
 * per job). */
typedef struct gd_server_config {
  int device;
  size_t arena_bytes;          /* HBM for resident jobs; 0 = 4 GiB */
  int wpg;                     /* waves per workgroup: 4, 8 (default), 16 */
  int persistent;              /* 1: persistent MC launches with a work queue */
  int use_graph;               /* 1: replay captured cycles between changes */
  int kde_policy;              /* enum gd_kde_policy */
  int max_resident_replicas;   /* admission cap besides memory; 0 = none */
  int cycles_per_tick;         /* cycles queued per admission round; 0 = 1 */
  int max_ticks_in_flight;     /* ticks queued ahead of the GPU; 0 = 2 */
  int keep_results;            /* 1: results kept until gd_server_fetch;
                                  0: only latency statistics are kept */
  size_t readback_bytes;       /* pinned host memory for read-back; 0 = 64 MiB */
  int occ;                     /* MC kernel register budget: 0, 6 or 8 */
  int split;                   /* wavefronts per replica: as gd_engine_config.split */
  /* Lanes: every job advances steps_per_exchange steps per cycle, and a
   * cycle lasts as long as the slowest workgroup task in it, which is set by
   * the largest complex resident (one wavefront walks lna x pnp pairs per
   * step).  With lanes > 1 the server runs that many engines, each with its
   * own stream, worker thread, 1/lanes of arena_bytes and readback_bytes,
   * and its own cycles, side by side on the device; a job goes to the first
   * lane k < lanes - 1 with lna x pnp <= lane_max_pairs[k], else to the
   * last.  Admission stays FIFO within a lane; max_resident_replicas counts
   * every lane.  0 or 1 = one engine.  The lanes' streams share the
   * process's hardware queues (GPU_MAX_HW_QUEUES, 4 by default): with four
   * lanes two of them shared a queue and their launches ran in turn; with 8
   * queues they did not, but four lanes served no better than three.
   * lane_split[k] > 1 runs lane k with that many wavefronts per replica (4,
   * 8 or 16) and KDE points staged in LDS: a few small jobs then fill the
   * device, so they are done in a few cycles of their own instead of waiting
   * in a large resident set.  A job that lane cannot split (more than 64
   * restraints, or its tables do not fit in LDS beside the split area) goes
   * to the next lane.  0 = split as above.
   * lane_max_jobs[k] > 0: lane k takes a job only while fewer jobs than
   * that are queued or resident on it; the others go to the next lane.  A
   * split lane holds a few jobs at a time at low latency and hands the rest
   * of a heavy load to the one-wavefront lanes, which carry more work per
   * second.  0 = no limit; the last lane takes every job routed to it. */
  int lanes;
  int lane_max_pairs[GD_SERVER_MAX_LANES - 1];
  int lane_split[GD_SERVER_MAX_LANES];
  int lane_max_jobs[GD_SERVER_MAX_LANES];
  /* lane_prio[k]: gd_engine_config.prio of lane k (a lane of small jobs
   * above the others: its wavefronts issue first) */
  int lane_prio[GD_SERVER_MAX_LANES];
  /* lane_max_total[k] > 0: lane k takes a job only while fewer jobs than
   * that are queued or resident in the whole server, so a split lane serves
   * light load and steps aside under heavy load.  0 = no limit. */
  int lane_max_total[GD_SERVER_MAX_LANES];
} gd_server_config;

enum gd_job_state {
  GD_JOB_QUEUED = 0, GD_JOB_RUNNING = 1, GD_JOB_DONE = 2, GD_JOB_FAILED = 3,
  GD_JOB_CANCELLED = 4
};

typedef struct gd_server_job_info {
  long long ticket;
  int state;                   /* enum gd_job_state */
  int status;                  /* gd_status of a failed job */
  int n_rep, steps_total;
  double t_submit, t_admit, t_done;   /* seconds since gd_server_create */
  double latency;              /* t_done - t_submit */
  size_t footprint;            /* device bytes while resident */
  int lane;                    /* the lane the job was routed to */
} gd_server_job_info;

typedef struct gd_server_stats {
  long long submitted, completed, failed, cancelled;
  int queued, running;
  double p50_latency, p99_latency, mean_latency, max_latency;  /* s */
  double p50_wait, p99_wait;   /* submission to admission, s */
  double pose_evals;           /* MC pose evaluations of completed jobs */
  double first_submit, last_done;
  double pose_evals_per_s;     /* pose_evals / (last_done - first_submit) */
  size_t arena_size, arena_peak;   /* summed over lanes */
  int running_peak;
  long long cycles, rebuilds;      /* summed over lanes */
  int lanes;
  long long lane_cycles[GD_SERVER_MAX_LANES];
  long long lane_completed[GD_SERVER_MAX_LANES];
  int lane_split[GD_SERVER_MAX_LANES];   /* wavefronts per replica of each
                                            lane's last cycle */
} gd_server_stats;

int gd_server_create(const gd_server_config *cfg, gd_server **out);
/* copies the job's inputs; GD_ERR_NOMEM if it can never fit the arena */
int gd_server_submit(gd_server *s, const gd_job *job, long long *ticket);
/* until the job is done, failed or cancelled (GD_OK) or timeout_s passes
 * (GD_ERR_TIMEOUT); timeout_s < 0 waits without limit */
int gd_server_wait(gd_server *s, long long ticket, double timeout_s,
                   gd_server_job_info *info);
/* results of a finished job (keep_results = 1), then forgets the job;
 * results: [n_rep] or NULL */
int gd_server_fetch(gd_server *s, long long ticket, gd_mclog *mclog,
                    gd_replica_result *results, gd_server_job_info *info);
int gd_server_cancel(gd_server *s, long long ticket);
/* until nothing is queued or running */
int gd_server_drain(gd_server *s);
int gd_server_get_stats(gd_server *s, gd_server_stats *st);
/* cancels what is queued or running and stops the loop */
void gd_server_destroy(gd_server *s);

/* ---------------------------------------------------------------------- */
/* Host preparation (util.C:199-558)                                        */
/* ---------------------------------------------------------------------- */
/* OptimizeEnepara, util.C:404-458 */
int gd_optimize_enepara(const gd_enepara0 *src, gd_enepara *dst);
/* SetTemperature, util.C:490-532 */
int gd_set_temperature(gd_temp *temp, int num_temp, float floor_temp,
                       float ceiling_temp);
/* SetReplica, util.C:537-558 (the replica table only) */
int gd_set_replica(gd_replica *replica, gd_complex_size size);

/* ---------------------------------------------------------------------- */
/* Input files of one complex (load.C:99-1126) and their preparation        */
/* (util.C:199-402)                                                         */
/* ---------------------------------------------------------------------- */
/* The per-complex files of a docking run: InputFiles (dock.h:62-75), set by
 * the -l, -p, -s and -id options (util.C:28-195). */
typedef struct gd_input_files {
  const char *ligand_sdf;   /* SD file: atom block, OB_ATOM_TYPES,
                               OB_ATOMIC_CHARGES, ENSEMBLE_COORDS */
  const char *molid;        /* tag of the ligand-name property; NULL = "MOLID" */
  const char *protein_pdb;  /* PDB file, one MODEL ... ENDMDL per conformer */
  const char *lhm_ff;       /* LHM file: KDE, PSP and MCS lines */
  const char *ligand_id;    /* the MCS lines of this id are read (-id) */
  int flags;                /* GD_LOAD_* */
} gd_input_files;

enum {
  /* Peptide-plate points of every PDB model average the N of their residue
   * and the C and O of the previous one.  Without it, models from the third
   * on take C and O of residue serial 0 instead, as the reference's
   * loadProtein does (load.C:886 reads an unfilled conformer). */
  GD_LOAD_PLATE_ALL_MODELS = 1
};

typedef struct gd_inputs gd_inputs;   /* host-memory prepared complex */

/* What Run() takes about one complex, in memory owned by the gd_inputs
 * object (valid until gd_inputs_free). */
typedef struct gd_inputs_view {
  gd_complex_size size;     /* n_lig n_prt lna pnp pnk pos; n_tmp = n_rep = 0 */
  gd_ligand lig;
  gd_protein prt;
  gd_psp psp;
  gd_kde kde;
  gd_mcs mcs;               /* n_field = lna + 2 */
  const float *lens_rmsd;   /* [n_lig] RMSD to the native pose (Ligand0.lens_rmsd) */
  const int *prt_d;         /* [n_prt * pnp] residue code of each point (Protein0.d) */
  const int *prt_r;         /* [n_prt * pnp] residue index of each point (Protein0.r) */
  const char *ligand_name;  /* the line after the molid tag; "" if absent */
} gd_inputs_view;

/* loadLigand + loadProtein + loadLHM (load.C:99-419, 595-983, 1010-1126),
 * then OptimizeLigand/OptimizeProtein/OptimizePsp/OptimizeKde/OptimizeMcs
 * (util.C:199-402), the sequence of dock.C:40-42,70-74.  enepara0 supplies
 * the ELE and HPP tables of the protein points (util.C:263-273). */
int gd_load_inputs(const gd_input_files *files, const gd_enepara0 *enepara0,
                   gd_inputs **out);
int gd_inputs_get(const gd_inputs *in, gd_inputs_view *view);
void gd_inputs_free(gd_inputs *in);

/* ---------------------------------------------------------------------- */
/* Parameter files (load.C:1138-1419)                                       */
/* ---------------------------------------------------------------------- */
/* loadEnePara (load.C:1138-1324): gpudocksm.ff; zero-fills first. */
int gd_load_enepara(const char *path, gd_enepara0 *enepara0);
/* loadWeight (load.C:1389-1419) */
int gd_load_weight(const char *path, gd_enepara0 *enepara0);
/* loadNorPara (load.C:1326-1387) */
int gd_load_norpara(const char *path_a, const char *path_b,
                    gd_enepara0 *enepara0);

/* ---------------------------------------------------------------------- */
/* Measurement (no counterpart in the reference)                             */
/* ---------------------------------------------------------------------- */
/* FP32 VALU issue rate of the device: waves_per_simd wavefronts on every
 * SIMD run independent FMA chains with all operands in vector registers;
 * *ns_per_inst = time per wavefront FMA instruction per SIMD, *tflops =
 * the FMA rate over the whole device, *n_simd = SIMDs of the device.  The
 * ceiling the MC kernel's instruction rate is read against. */
int gd_valu_probe(int device, int waves_per_simd, double *ns_per_inst, double *tflops,
                  int *n_simd);

/* Issue rate per instruction kind: the same probe for each of the
 * GD_VALU_KINDS instruction kinds the energy loops are made of (FMA with
 * vector operands, FMA with a scalar-register operand, mul, max, cndmask
 * with a scalar mask, integer shift-add, compare, DPP add, readlane, and
 * FMA/max alternating); ns_per_inst[k] = time per wavefront instruction per
 * SIMD of kind k, named by gd_valu_kind_name(k); n_kinds must be
 * GD_VALU_KINDS. */
#define GD_VALU_KINDS 10
int gd_valu_probe_kinds(int device, int waves_per_simd, double *ns_per_inst, int n_kinds);
const char *gd_valu_kind_name(int kind);

/* ---------------------------------------------------------------------- */
const char *gd_last_error(void);
int gd_device_count(void);
const char *gd_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GD_REMC_H */
