/*
 * gd_launch.h -- host-side launch interface of the kernels in gd_kernels.hip.
 */
#ifndef GD_LAUNCH_H
#define GD_LAUNCH_H

#include <hip/hip_runtime.h>
#include "gd_device.h"

/* LDS carve-up chosen by the host for one pack. */
typedef struct gd_launch_cfg {
  int off_eh;      /* bytes: pocket float2 array */
  int off_tab;     /* bytes: pair-parameter rows */
  int off_kde;     /* bytes: KDE points (when kde_in_lds) */
  int lds_bytes;   /* total dynamic LDS per workgroup */
  int kde_in_lds;
} gd_launch_cfg;

hipError_t gd_launch_mc(const gd_pack &P, const gd_state &S, const gd_task *tasks,
                        int n_tasks, const int *rep_lig, const gd_launch_cfg &C,
                        int s_begin, int n_steps, uint32_t seed, uint32_t cid,
                        int mode, hipStream_t st);
hipError_t gd_launch_exchange(const gd_pack &P, const gd_state &S, int x,
                              uint32_t seed, uint32_t cid, hipStream_t st);
hipError_t gd_launch_score(const gd_pack &P, const gd_task *tasks, int n_tasks,
                           const int *pose_ids, const int *lig_conf,
                           const float *move, const gd_launch_cfg &C,
                           float *energy, hipStream_t st);
hipError_t gd_launch_rescore64(const gd_pack &P, const int *lig_conf,
                               const int *prt_conf, const float *move, int n_pose,
                               double *energy, hipStream_t st);
hipError_t gd_set_kernel_attrs(int lds_bytes);
int gd_waves_per_group(void);

#endif /* GD_LAUNCH_H */
