/*
 * gd_launch.h -- host-side launch interface of the kernels in gd_kernels.hip.
 */
#ifndef GD_LAUNCH_H
#define GD_LAUNCH_H

#include <hip/hip_runtime.h>
#include "gd_device.h"

/* MC launch shape. */
typedef struct gd_launch_cfg {
  int wpg;          /* wavefronts (replicas) per workgroup: 4, 8 or 16 */
  int stage_bytes;  /* dynamic LDS for staging: max over resident complexes */
  int persistent;   /* 1: grid of resident workgroups pulling from a queue */
  int occ;          /* register budget for waves per SIMD: 0 (compiler), 6 or 8 */
  int split;        /* wavefronts per replica in MC cycles: 1, or 4, 8, 16 with one
                       replica per workgroup (tasks of one replica) */
  int prio;         /* issue priority of the MC launches' wavefronts in their SIMD
                       against other launches' (s_setprio 0-3) */
} gd_launch_cfg;

hipError_t gd_launch_mc(const gd_launch_cfg &C, const gd_pack *packs, const gd_task *tasks,
                        int n_tasks, int *queue, int grid, int mode, hipStream_t st);
/* exchange over chains[0 .. n_chain), then the advance of
 * gd_launch_advance: by one workgroup for few chains, else by the last
 * workgroup to finish (*done: a zeroed counter the launch leaves zeroed).
 * With no chains, the advance alone. */
hipError_t gd_launch_exchange(const gd_pack *packs, const gd_chain *chains, int n_chain,
                              int n_cx, int *queue, int n_queue, int *done, hipStream_t st);
/* advances the step counters of packs[0 .. n_cx) and zeroes queue[0 .. n_queue) */
hipError_t gd_launch_advance(const gd_pack *packs, int n_cx, int *queue, int n_queue,
                             hipStream_t st);
hipError_t gd_launch_score(const gd_pack *packs, const gd_task *tasks, int n_tasks,
                           const int *pose_ids, const int *lig_conf, const float *move,
                           int stage_bytes, float *energy, hipStream_t st);
/* pose i: movematrix at move[i * move_stride .. +6] */
/* blocks x 256 lanes, each 128 x iters FMAs; out holds blocks x 256 floats */
hipError_t gd_launch_valu_probe(float *out, int blocks, int iters, hipStream_t st);
/* one kind of gd_valu_probe_kinds (0 .. GD_VALU_KINDS - 1) */
hipError_t gd_launch_valu_kind(int kind, float *out, int blocks, int iters, hipStream_t st);
hipError_t gd_launch_rescore64(const gd_pack *packs, int cx_index, const int *lig_conf,
                               const int *prt_conf, const float *move, int move_stride,
                               int n_pose, double *energy, hipStream_t st);
/* items[i] = (complex index in packs, replica): FP64 energy of its best pose */
hipError_t gd_launch_rescore64_best(const gd_pack *packs, const int2 *items, int n,
                                    double *energy, hipStream_t st);
hipError_t gd_set_kernel_attrs(int stage_bytes);

/* Bytes of the static per-GPU workspace of gd_score_poses_noalloc. */
#define GD_SCRATCH_BYTES ((size_t) 64 << 20)
/* Device address of that workspace on `device`, which must be the current
 * device; resolved once per device (hipGetSymbolAddress) and kept. */
hipError_t gd_scratch_region(int device, char **ptr);
hipError_t gd_mc_occupancy(const gd_launch_cfg &C, int stage_bytes, int *blocks_per_cu);

#endif /* GD_LAUNCH_H */
