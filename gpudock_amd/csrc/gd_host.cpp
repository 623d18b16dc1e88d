/*
 * gd_host.cpp -- error reporting, device queries, argument checks and the
 * device-memory arena shared by the C ABI.
 */
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "gd_internal.h"

static thread_local std::string g_err;

int gd_fail(int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

/* error hook for the preparation/parsing code in gd_prep.cpp */
int gd_prep_fail(int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

extern "C" const char *gd_last_error(void) { return g_err.c_str(); }
extern "C" const char *gd_version(void) { return "gd-remc-mi355x 0.2 (gfx950)"; }

extern "C" int gd_device_count(void)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess)
    return 0;
  return n;
}

/* ------------------------------------------------------------------------ */
/* argument checks                                                           */
/* ------------------------------------------------------------------------ */
int gd_schedule(const gd_mcpara *mc, int *steps_run, int *launches_per_window)
{
  if (!mc)
    return gd_fail(GD_ERR_ARG, "null mcpara");
  const long long total = mc->steps_total, spd = mc->steps_per_dump,
                  spe = mc->steps_per_exchange;
  if (total < 0 || spe < 1 || spd < 1)
    return gd_fail(GD_ERR_ARG,
                   "steps_total=%lld steps_per_dump=%lld steps_per_exchange=%lld (need "
                   "steps_total >= 0, steps_per_dump >= 1, steps_per_exchange >= 1)",
                   total, spd, spe);
  const long long lpw = (spd + spe - 1) / spe;
  const long long windows = (total + spd - 1) / spd;
  const long long run = windows * lpw * spe;
  if (run > (1LL << 30))
    return gd_fail(GD_ERR_ARG, "schedule runs %lld MC steps (limit 2^30)", run);
  *steps_run = (int) run;
  *launches_per_window = (int) lpw;
  return GD_OK;
}

int gd_check_inputs(const gd_ligand *lig, const gd_protein *prt, const gd_psp *psp,
                        const gd_kde *kde, const gd_mcs *mcs, const gd_enepara *ep,
                    const gd_complex_size &sz)
{
  if (!lig || !prt || !psp || !kde || !mcs || !ep)
    return gd_fail(GD_ERR_ARG, "null input struct");
  if (sz.lna < 1 || sz.lna > GD_MAX_LNA)
    return gd_fail(GD_ERR_ARG, "lna=%d outside [1, %d]", sz.lna, GD_MAX_LNA);
  if (sz.pnp < 1 || sz.n_lig < 1 || sz.n_prt < 1)
    return gd_fail(GD_ERR_ARG, "empty ensemble (pnp=%d n_lig=%d n_prt=%d)", sz.pnp,
                sz.n_lig, sz.n_prt);
  if (sz.pnk < 0 || sz.pos < 0)
    return gd_fail(GD_ERR_ARG, "negative pnk/pos");
  if (!lig->x || !lig->y || !lig->z || !lig->t || !lig->c || !lig->n)
    return gd_fail(GD_ERR_ARG, "null ligand array");
  if (!prt->x || !prt->y || !prt->z || !prt->t || !prt->c || !prt->ele ||
      !prt->seq3r || !prt->c0_and_d12_or_c2 || !prt->hpp)
    return gd_fail(GD_ERR_ARG, "null protein array");
  for (int a = 0; a < sz.lna; ++a)
    if (lig->t[a] < 0 || lig->t[a] >= GD_MAXTP2)
      return gd_fail(GD_ERR_ARG, "ligand atom %d has type %d outside [0, %d)", a,
                  lig->t[a], GD_MAXTP2);
  const size_t np = (size_t) sz.n_prt * sz.pnp;
  for (size_t i = 0; i < np; ++i) {
    if (prt->t[i] < 0 || prt->t[i] >= GD_MAXTP1)
      return gd_fail(GD_ERR_ARG, "protein point %zu has type %d outside [0, %d)", i,
                  prt->t[i], GD_MAXTP1);
    if (prt->c[i] == 2 && (prt->seq3r[i] < 0 || prt->seq3r[i] >= psp->n_res))
      return gd_fail(GD_ERR_ARG, "protein point %zu: seq3r %d outside PSP table (%d)",
                  i, prt->seq3r[i], psp->n_res);
    if (prt->seq3r[i] >= (1 << 23))
      return gd_fail(GD_ERR_ARG, "seq3r %d too large", prt->seq3r[i]);
  }
  if (psp->n_res < 1 || !psp->psp)
    return gd_fail(GD_ERR_ARG, "empty PSP table");
  if (sz.pnk > 0 && (!kde->x || !kde->y || !kde->z || !kde->t))
    return gd_fail(GD_ERR_ARG, "null KDE array");
  if (sz.pos > 0 && (!mcs->x || !mcs->y || !mcs->z || !mcs->tcc))
    return gd_fail(GD_ERR_ARG, "null MCS array");
  if (sz.pos > 0)
    for (int a = 0; a < sz.lna; ++a)
      if (lig->n[a] + 1 < 0 || lig->n[a] + 1 >= mcs->n_field)
        return gd_fail(GD_ERR_ARG, "ligand atom number %d outside MCS fields (%d)",
                    lig->n[a], mcs->n_field);
  return GD_OK;
}

/* ------------------------------------------------------------------------ */
/* arena                                                                     */
/* ------------------------------------------------------------------------ */
GdArena::~GdArena()
{
  if (base_) {
    if (pinned_)
      (void) hipHostFree(base_);
    else
      (void) hipFree(base_);
  }
}

hipError_t GdArena::init(size_t bytes, bool pinned_host)
{
  size_ = round(bytes);
  pinned_ = pinned_host;
  hipError_t e = pinned_host ? hipHostMalloc((void **) &base_, size_ ? size_ : 256, 0)
                             : hipMalloc((void **) &base_, size_ ? size_ : 256);
  if (e != hipSuccess) {
    base_ = nullptr;
    size_ = 0;
    return e;
  }
  free_.clear();
  free_[0] = size_;
  used_ = 0;
  return hipSuccess;
}

size_t GdArena::alloc(size_t bytes, bool high)
{
  bytes = round(bytes ? bytes : 1);
  if (high) {
    /* from the end of the highest free range that fits: small, short-lived
     * blocks gather at the top and leave the low ranges whole */
    for (auto it = free_.rbegin(); it != free_.rend(); ++it) {
      if (it->second >= bytes) {
        const size_t off = it->first + it->second - bytes;
        if (it->second == bytes)
          free_.erase(std::next(it).base());
        else
          it->second -= bytes;
        used_ += bytes;
        return off;
      }
    }
    return (size_t) -1;
  }
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second >= bytes) {
      const size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > bytes)
        free_[off + bytes] = len - bytes;
      used_ += bytes;
      return off;
    }
  }
  return (size_t) -1;
}

void GdArena::free(size_t off, size_t bytes)
{
  bytes = round(bytes ? bytes : 1);
  used_ -= bytes;
  auto it = free_.emplace(off, bytes).first;
  auto nx = std::next(it);
  if (nx != free_.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    free_.erase(nx);
  }
  if (it != free_.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      free_.erase(it);
    }
  }
}
