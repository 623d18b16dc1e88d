/*
 * gd_engine.h -- the resident-complex engine (host side).
 */
#ifndef GD_ENGINE_H
#define GD_ENGINE_H

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "gd_internal.h"

enum { GD_KDE_AUTO = GD_KDE_STAGE_AUTO, GD_KDE_GLOBAL = GD_KDE_STAGE_NEVER };

/* Host image of one complex: every device array in its final layout, plus
 * the pack whose pointers are filled in at upload. */
struct GdImage {
  gd_pack P;
  std::vector<float4> pt_pos;
  std::vector<float2> pt_eh;
  std::vector<gd_pair_param> tab;
  std::vector<float4> lig_xyz;
  std::vector<int> lig_row, lig_type;
  std::vector<float> lig_q;
  std::vector<float4> kde;
  std::vector<uint32_t> kde_pairs;
  std::vector<int> kde_seg;
  std::vector<float> kde_inv;
  std::vector<float4> mcs;
  std::vector<float> tcc;
  std::vector<float> psp_pt;     /* [n_prt][n_rows][pnp_pad] */
  std::vector<int> chunk_bounds; /* [n_prt][2]: kh, kp */
  std::vector<int> row_hdb;      /* [n_rows] */
  std::vector<gd_atom_rec> atom; /* [lna] */
  std::vector<float> hpl;
  std::vector<float> mb;
  std::vector<int> rep_lig, rep_prt, idx_tmp;
  std::vector<unsigned long long> mcc_ref;   /* [n_prt][lna][nchunk] */
};

/* Byte offsets of a complex's arrays inside its arena block.  The replica
 * state [mv, step) is contiguous, so it is saved, restored and read back
 * as one copy. */
struct GdBlock {
  size_t pt_pos, pt_eh, tab, lig_xyz, lig_row, lig_type, lig_q, kde, kde_pairs, kde_seg, kde_inv, mcs,
      tcc, psp_pt, chunk_bounds, row_hdb, atom, hpl, mb, rep_lig, rep_prt, mv, e, mv_best, e_best, step_best, idx_tmp, acs_mc,
      acs_x, cmcc, cmcc_best, step, pack, rec, rec_bytes, mcc_ref, total;
};

/* temp, replica and mc may be NULL (pose scoring) */
int gd_build_image(GdImage &I, const gd_ligand *lig, const gd_protein *prt,
                   const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                   const gd_enepara *ep, const gd_temp *temp, const gd_replica *replica,
                   const gd_mcpara *mc, const gd_complex_size &sz, const gd_options &o,
                   int max_lds, int kde_policy);
GdBlock gd_layout_block(const GdImage &I);
/* device bytes of the image once uploaded */
size_t gd_image_footprint(const GdImage &I);
/* the arena block of the image, laid out in host memory for the device
 * address dev (fills I.P's pointers; the pack itself sits at b.pack) */
void gd_pack_image(GdImage &I, char *host, char *dev, GdBlock &b);
/* workgroup tasks of one complex: runs of replicas on one protein
 * conformer, at most wpg each */
void gd_complex_tasks(const GdImage &I, int cx, int wpg, std::vector<gd_task> &out);
/* fill per-replica results from a host copy of the state region */
void gd_parse_state(const GdImage &I, const GdBlock &b, const char *state, const double *e64,
                    gd_mclog *mclog, gd_replica_result *results);

/* One resident complex. */
struct GdSlot {
  std::unique_ptr<GdImage> img;
  GdBlock blk;
  size_t off = 0, bytes = 0;        /* arena block */
  gd_pack *pack_dev = nullptr;      /* the complex's own pack copy */
  int handle = -1;
  int cx_index = -1;                /* index in the engine's pack array */
  int total_steps = 0, spe = 1;
  int steps_done = 0, cycles_done = 0;  /* host mirror of the device counters */
  size_t state_bytes() const { return blk.step - blk.mv; }
};

/* The read-back of one complex's results: its state region, then the FP64
 * energies of its best poses, copied into host memory at `host`. */
struct GdCollect {
  int handle = -1;
  char *host = nullptr;
  size_t host_off = 0, host_bytes = 0;
};

struct gd_engine {
  gd_engine_config conf{};
  gd_launch_cfg lc{8, 16, 0, 0};
  int max_lds = 0, n_cu = 0;
  GdArena arena;
  GdArena pinned;                   /* pinned host staging for uploads */
  hipStream_t stream = nullptr;
  size_t ctl_off = 0;               /* queue counter */
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<GdSlot *> slots;
  std::vector<int> pending_init;    /* admitted, initial energies not yet queued */
  int next_handle = 1;
  bool dirty = true;
  size_t packs_off = 0, packs_bytes = 0, tasks_off = 0, tasks_bytes = 0, chains_off = 0,
         chains_bytes = 0;
  int n_cx = 0, n_tasks = 0, n_chain = 0;
  /* MC launches of a cycle: complexes whose LDS need gives the same
   * workgroups per CU share one launch (tasks [off, off + n) of the task
   * list, queue counter = class index), so a large pocket does not lower
   * the occupancy of the small ones */
  struct Class {
    int stage, off, n, grid, per_cu;
  };
  std::vector<Class> classes;
  int kde_global = 0;               /* 1: every complex reads KDE points from HBM */
  int split_eff = 1;                /* wavefronts per replica of the current launches
                                       (lc.split 0 = chosen at each rebuild) */
  int timing = 0;
  std::vector<hipEvent_t> ev_mc, ev_x;
  double mc_ms = 0.0, x_ms = 0.0;
  int mc_launches = 0, x_launches = 0;
  long long n_cycles = 0, n_rebuilds = 0, n_captures = 0;

  /* arena blocks and graphs still referenced by queued work: released once
   * the event recorded after that work has completed */
  struct Deferred {
    hipEvent_t ev;
    size_t off, bytes;
    hipGraph_t graph;
    hipGraphExec_t gexec;
    bool host;                      /* block of the pinned staging arena */
  };
  std::vector<Deferred> deferred;

  gd_engine() = default;
  gd_engine(const gd_engine &) = delete;
  gd_engine &operator=(const gd_engine &) = delete;
  ~gd_engine();

  int init(const gd_engine_config *cfg);
  /* device task lists current with the members (rebuilt after a change) */
  int ensure_built();
  /* build + admit (admission never waits for queued cycles) */
  int add(const gd_ligand *lig, const gd_protein *prt, const gd_psp *psp, const gd_kde *kde,
          const gd_mcs *mcs, const gd_enepara *ep, const gd_temp *temp,
          const gd_replica *replica, const gd_mcpara *mc, const gd_complex_size &sz,
          const gd_options &o, int *handle);
  /* admit a built image; it is moved into the engine only on success */
  int add_image(std::unique_ptr<GdImage> &img, int *handle);
  int build_image(GdImage &I, const gd_job &job) const;
  /* a complex this engine could run split over several wavefronts */
  bool can_split(const GdImage &I) const;
  /* retire; its memory is released once queued work no longer needs it */
  int remove(int handle);
  GdSlot *find(int handle);
  int cycles(int n);
  int sync();
  /* synchronous results of one complex */
  int results(int handle, gd_mclog *mclog, gd_replica_result *results);
  /* queue, for complexes cs[0..n), one FP64 re-score launch of all their
   * replicas' best poses and the read-back of each into cs[i].host
   * (cs[i].host_bytes >= collect_bytes()); *done (created here) is recorded
   * after the copies */
  size_t collect_bytes(int handle);
  int collect_async(GdCollect *cs, int n, hipEvent_t *done);
  int autotune(gd_tune_report *rep);
  void drain_timing();
  /* release deferred memory whose work has completed */
  void reap();
  int grid() const;                 /* workgroups per cycle, all classes */
  int resident() const;             /* workgroups resident at once, all classes */
  int *queue_ptr();
  int *done_ptr();
  size_t free_bytes() const { return arena.size() - arena.used(); }

private:
  int rebuild();
  void drop_graph();
  void defer(size_t off, size_t bytes, hipGraph_t g = nullptr, hipGraphExec_t x = nullptr,
             bool host = false);
  /* stream-ordered copy of host bytes to device memory, staged through
   * pinned memory so that it does not wait for queued work */
  int upload(char *dev, const void *src, size_t bytes);
  int upload_image(GdImage &I, size_t off, GdBlock &b);
  int enqueue_cycle();
  int trial_cycle();
  /* wavefronts per replica for the resident complexes under lc.split */
  int choose_split() const;
};
typedef gd_engine gd_engine_t;

#endif /* GD_ENGINE_H */
