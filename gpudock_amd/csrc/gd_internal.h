/*
 * gd_internal.h -- host-side internals shared by the C ABI sources.
 */
#ifndef GD_INTERNAL_H
#define GD_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <map>
#include <string>
#include <vector>

#include "../../include/gd_remc.h"
#include "gd_device.h"
#include "gd_launch.h"

/* error reporting: sets the calling thread's gd_last_error() text */
int gd_fail(int code, const char *fmt, ...);

#define GD_HIP_TRY(expr)                                                       \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return gd_fail(GD_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                     __FILE__, __LINE__);                                      \
  } while (0)

int gd_check_inputs(const gd_ligand *lig, const gd_protein *prt, const gd_psp *psp,
                    const gd_kde *kde, const gd_mcs *mcs, const gd_enepara *ep,
                    const gd_complex_size &sz);

/* The reference's launch schedule (gpusingle.cu:8-20): windows of
 * steps_per_dump steps start at s1 = 0, spd, ... while s1 < steps_total; each
 * window runs whole launches of steps_per_exchange MC steps at s2 = 0, spe,
 * ... while s2 < spd, each followed by one exchange in mode (s2 / spe) % 2.
 * *launches_per_window = ceil(spd / spe) and *steps_run = windows x launches
 * per window x spe, the MC steps that actually run (above steps_total when
 * spd does not divide steps_total or spe does not divide spd). */
int gd_schedule(const gd_mcpara *mc, int *steps_run, int *launches_per_window);

/* First-fit sub-allocator over one device allocation (or one pinned host
 * allocation, for staging copies).  Jobs are admitted and retired without
 * hipMalloc/hipFree (which synchronise the device). */
class GdArena {
public:
  GdArena() = default;
  GdArena(const GdArena &) = delete;
  GdArena &operator=(const GdArena &) = delete;
  ~GdArena();
  hipError_t init(size_t bytes, bool pinned_host = false);
  /* offset of a block of `bytes` (256-B aligned) or SIZE_MAX; high: taken
   * from the top of the highest range that fits (for small blocks) */
  size_t alloc(size_t bytes, bool high = false);
  void free(size_t off, size_t bytes);
  char *base() const { return base_; }
  size_t size() const { return size_; }
  size_t used() const { return used_; }
  static size_t round(size_t b) { return (b + 255) & ~(size_t) 255; }

private:
  char *base_ = nullptr;
  size_t size_ = 0, used_ = 0;
  bool pinned_ = false;
  std::map<size_t, size_t> free_;  /* offset -> length */
};

#endif /* GD_INTERNAL_H */
