/*
 * gd_host.cpp -- host side of the C ABI (include/gd_remc.h).
 *
 * Builds the device-resident complex pack, owns the replica state, and
 * drives the kernel cadence of the reference driver
 * (gpusingle.cu:3-38: MonteCarlo_d launches of steps_per_exchange steps,
 * each followed by ExchangeReplicas_d) on one HIP stream, with no host
 * synchronisation between launches.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gd_remc.h"
#include "gd_device.h"
#include "gd_launch.h"

/* ------------------------------------------------------------------------ */
/* errors                                                                    */
/* ------------------------------------------------------------------------ */
static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(GD_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                       \
  } while (0)

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
extern "C" const char *gd_version(void) { return "gd-remc-mi355x 0.1 (gfx950)"; }

extern "C" int gd_device_count(void)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess)
    return 0;
  return n;
}

/* ------------------------------------------------------------------------ */
/* device buffers                                                            */
/* ------------------------------------------------------------------------ */
struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release()
  {
    if (p)
      (void) hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t bytes)
  {
    release();
    n = bytes;
    return hipMalloc(&p, bytes ? bytes : 16);
  }
  template <class T> hipError_t upload(const std::vector<T> &v)
  {
    hipError_t e = alloc(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty())
      return e;
    return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  }
  template <class T> T *as() const { return (T *) p; }
};

/* ------------------------------------------------------------------------ */
/* argument checks                                                           */
/* ------------------------------------------------------------------------ */
static int check_inputs(const gd_ligand *lig, const gd_protein *prt, const gd_psp *psp,
                        const gd_kde *kde, const gd_mcs *mcs, const gd_enepara *ep,
                        const gd_complex_size &sz)
{
  if (!lig || !prt || !psp || !kde || !mcs || !ep)
    return fail(GD_ERR_ARG, "null input struct");
  if (sz.lna < 1 || sz.lna > GD_MAX_LNA)
    return fail(GD_ERR_ARG, "lna=%d outside [1, %d]", sz.lna, GD_MAX_LNA);
  if (sz.pnp < 1 || sz.n_lig < 1 || sz.n_prt < 1)
    return fail(GD_ERR_ARG, "empty ensemble (pnp=%d n_lig=%d n_prt=%d)", sz.pnp,
                sz.n_lig, sz.n_prt);
  if (sz.pnk < 0 || sz.pos < 0)
    return fail(GD_ERR_ARG, "negative pnk/pos");
  if (!lig->x || !lig->y || !lig->z || !lig->t || !lig->c || !lig->n)
    return fail(GD_ERR_ARG, "null ligand array");
  if (!prt->x || !prt->y || !prt->z || !prt->t || !prt->c || !prt->ele ||
      !prt->seq3r || !prt->c0_and_d12_or_c2 || !prt->hpp)
    return fail(GD_ERR_ARG, "null protein array");
  for (int a = 0; a < sz.lna; ++a)
    if (lig->t[a] < 0 || lig->t[a] >= GD_MAXTP2)
      return fail(GD_ERR_ARG, "ligand atom %d has type %d outside [0, %d)", a,
                  lig->t[a], GD_MAXTP2);
  const size_t np = (size_t) sz.n_prt * sz.pnp;
  for (size_t i = 0; i < np; ++i) {
    if (prt->t[i] < 0 || prt->t[i] >= GD_MAXTP1)
      return fail(GD_ERR_ARG, "protein point %zu has type %d outside [0, %d)", i,
                  prt->t[i], GD_MAXTP1);
    if (prt->c[i] == 2 && (prt->seq3r[i] < 0 || prt->seq3r[i] >= psp->n_res))
      return fail(GD_ERR_ARG, "protein point %zu: seq3r %d outside PSP table (%d)",
                  i, prt->seq3r[i], psp->n_res);
    if (prt->seq3r[i] >= (1 << 23))
      return fail(GD_ERR_ARG, "seq3r %d too large", prt->seq3r[i]);
  }
  if (psp->n_res < 1 || !psp->psp)
    return fail(GD_ERR_ARG, "empty PSP table");
  if (sz.pnk > 0 && (!kde->x || !kde->y || !kde->z || !kde->t))
    return fail(GD_ERR_ARG, "null KDE array");
  if (sz.pos > 0 && (!mcs->x || !mcs->y || !mcs->z || !mcs->tcc))
    return fail(GD_ERR_ARG, "null MCS array");
  if (sz.pos > 0)
    for (int a = 0; a < sz.lna; ++a)
      if (lig->n[a] + 1 < 0 || lig->n[a] + 1 >= mcs->n_field)
        return fail(GD_ERR_ARG, "ligand atom number %d outside MCS fields (%d)",
                    lig->n[a], mcs->n_field);
  return GD_OK;
}

/* ------------------------------------------------------------------------ */
/* the complex pack                                                          */
/* ------------------------------------------------------------------------ */
struct Pack {
  gd_pack P;
  gd_launch_cfg C;
  DevBuf pt_pos, pt_eh, tab, lig_xyz, lig_row, lig_type, lig_q, kde, kde_seg, kde_inv,
      mcs, mcs_tcc, psp, hpl, minus_beta;
};

static int build_pack(Pack &K, const gd_ligand *lig, const gd_protein *prt,
                      const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                      const gd_enepara *ep, const gd_temp *temp,
                      const gd_complex_size &sz, int term_order,
                      const float move_scale[6], int device)
{
  gd_pack &P = K.P;
  memset(&P, 0, sizeof P);
  P.n_lig = sz.n_lig; P.n_prt = sz.n_prt; P.n_tmp = sz.n_tmp; P.n_rep = sz.n_rep;
  P.lna = sz.lna; P.pnp = sz.pnp; P.pnk = sz.pnk; P.pos = sz.pos;
  P.n_res = psp->n_res;
  P.term_order = term_order;

  /* table rows: distinct ligand atom types in order of first appearance */
  int row_of_type[GD_MAXTP2];
  for (int t = 0; t < GD_MAXTP2; ++t)
    row_of_type[t] = -1;
  std::vector<int> row_type;
  std::vector<int> lig_row(sz.lna), lig_type(sz.lna);
  std::vector<float> lig_q(sz.lna);
  for (int a = 0; a < sz.lna; ++a) {
    const int t = lig->t[a];
    if (row_of_type[t] < 0) {
      row_of_type[t] = (int) row_type.size();
      row_type.push_back(t);
    }
    lig_row[a] = row_of_type[t];
    lig_type[a] = t;
    lig_q[a] = lig->c[a];
  }
  P.n_rows = (int) row_type.size();

  std::vector<gd_pair_param> tab((size_t) P.n_rows * GD_ROW_STRIDE);
  memset(tab.data(), 0, tab.size() * sizeof(gd_pair_param));
  for (int r = 0; r < P.n_rows; ++r) {
    const int lt = row_type[r];
    for (int pt = 0; pt < GD_MAXTP1; ++pt) {
      gd_pair_param &q = tab[(size_t) r * GD_ROW_STRIDE + pt];
      q.p1a = ep->p1a[lt][pt];
      q.p2a = ep->p2a[lt][pt];
      q.pmf0 = ep->pmf0[lt][pt];
      q.pmf1 = ep->pmf1[lt][pt];
      q.hdb0 = ep->hdb0[lt][pt];
      q.hdb1 = ep->hdb1[lt][pt];
      q.pad0 = 0.0f;
      q.pad1 = 0.0f;
    }
  }

  /* pocket points, per conformer, in the canonical order (stable by
   * gd_point_sort_key) */
  const int pnp_pad = (sz.pnp + GD_WAVE - 1) / GD_WAVE * GD_WAVE;
  P.pnp_pad = pnp_pad;
  const size_t np = (size_t) sz.n_prt * pnp_pad;
  std::vector<float4> pt_pos(np);
  std::vector<float2> pt_eh(np);
  {
    /* padding: a far point of the all-zero table entry, no flags, ele 0 */
    const uint32_t pbits = GD_PAD_TYPE;
    float pw;
    memcpy(&pw, &pbits, 4);
    for (size_t i = 0; i < np; ++i) {
      pt_pos[i] = make_float4(1.0e3f, 1.0e3f, 1.0e3f, pw);
      pt_eh[i] = make_float2(0.0f, 0.0f);
    }
  }
  std::vector<int> perm(sz.pnp);
  for (int cf = 0; cf < sz.n_prt; ++cf) {
    const size_t base = (size_t) cf * sz.pnp;
    const size_t obase = (size_t) cf * pnp_pad;
    for (int j = 0; j < sz.pnp; ++j)
      perm[j] = j;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) {
      return gd_point_sort_key(prt->t[base + a], prt->c[base + a] == 2,
                               prt->c0_and_d12_or_c2[base + a] != 0) <
             gd_point_sort_key(prt->t[base + b], prt->c[base + b] == 2,
                               prt->c0_and_d12_or_c2[base + b] != 0);
    });
    for (int j = 0; j < sz.pnp; ++j) {
    const size_t i = base + perm[j];
    const size_t o = obase + j;
    uint32_t bits = (uint32_t) prt->t[i] & GD_PT_TYPE_MASK;
    if (prt->c[i] == 2)
      bits |= GD_PT_C2;
    if (prt->c0_and_d12_or_c2[i])
      bits |= GD_PT_HPC;
    bits |= (uint32_t) (prt->c[i] == 2 ? prt->seq3r[i] : 0) << GD_PT_SEQ_SHIFT;
    float w;
    memcpy(&w, &bits, 4);
    pt_pos[o] = make_float4(prt->x[i], prt->y[i], prt->z[i], w);
    pt_eh[o] = make_float2(prt->ele[i], prt->hpp[i]);
    }
  }

  /* ligand conformers */
  std::vector<float4> lig_xyz((size_t) sz.n_lig * sz.lna);
  for (size_t i = 0; i < lig_xyz.size(); ++i)
    lig_xyz[i] = make_float4(lig->x[i], lig->y[i], lig->z[i], 0.0f);

  /* KDE points grouped by table row (stable within a type) */
  std::vector<float4> kde_pts;
  std::vector<int> kde_seg(2 * (size_t) P.n_rows, 0);
  for (int r = 0; r < P.n_rows; ++r) {
    kde_seg[2 * r] = (int) kde_pts.size();
    for (int k = 0; k < sz.pnk; ++k)
      if (kde->t[k] == row_type[r])
        kde_pts.push_back(make_float4(kde->x[k], kde->y[k], kde->z[k], 0.0f));
    kde_seg[2 * r + 1] = (int) kde_pts.size();
  }
  const int pnk_used = (int) kde_pts.size();
  std::vector<float> kde_inv(P.n_rows);
  for (int r = 0; r < P.n_rows; ++r) {
    const int n = kde_seg[2 * r + 1] - kde_seg[2 * r];
    kde_inv[r] = n > 0 ? 1.0f / (float) n : 0.0f;
  }

  /* MCS targets per (restraint, ligand atom) */
  std::vector<float4> mcs_t((size_t) sz.pos * sz.lna);
  std::vector<float> tcc(sz.pos);
  for (int m = 0; m < sz.pos; ++m) {
    tcc[m] = mcs->tcc[m];
    for (int a = 0; a < sz.lna; ++a) {
      const size_t f = (size_t) m * mcs->n_field + lig->n[a] + 1;
      const float x = mcs->x[f];
      const bool valid = x != GD_MCS_INVALID_COORD;
      mcs_t[(size_t) m * sz.lna + a] =
          valid ? make_float4(x, mcs->y[f], mcs->z[f], 1.0f) : make_float4(0, 0, 0, 0);
    }
  }

  std::vector<float> psp_v((size_t) GD_MAXTP2 * psp->n_res);
  memcpy(psp_v.data(), psp->psp, psp_v.size() * sizeof(float));
  std::vector<float> hpl(4 * GD_MAXTP2);
  for (int t = 0; t < GD_MAXTP2; ++t) {
    hpl[t] = ep->hpl0[t];
    hpl[GD_MAXTP2 + t] = ep->hpl1[t];
    hpl[2 * GD_MAXTP2 + t] = ep->hpl2[t];
    hpl[3 * GD_MAXTP2 + t] = 1.0f / ep->hpl1[t];
  }
  std::vector<float> mb(std::max(sz.n_tmp, 1), 0.0f);
  if (temp)
    for (int t = 0; t < sz.n_tmp; ++t)
      mb[t] = temp[t].minus_beta;

  HIP_TRY(hipSetDevice(device));
  HIP_TRY(K.pt_pos.upload(pt_pos));
  HIP_TRY(K.pt_eh.upload(pt_eh));
  HIP_TRY(K.tab.upload(tab));
  HIP_TRY(K.lig_xyz.upload(lig_xyz));
  HIP_TRY(K.lig_row.upload(lig_row));
  HIP_TRY(K.lig_type.upload(lig_type));
  HIP_TRY(K.lig_q.upload(lig_q));
  HIP_TRY(K.kde.upload(kde_pts));
  HIP_TRY(K.kde_seg.upload(kde_seg));
  HIP_TRY(K.kde_inv.upload(kde_inv));
  HIP_TRY(K.mcs.upload(mcs_t));
  HIP_TRY(K.mcs_tcc.upload(tcc));
  HIP_TRY(K.psp.upload(psp_v));
  HIP_TRY(K.hpl.upload(hpl));
  HIP_TRY(K.minus_beta.upload(mb));

  P.pt_pos = K.pt_pos.as<float4>();
  P.pt_eh = K.pt_eh.as<float2>();
  P.tab = K.tab.as<gd_pair_param>();
  P.lig_xyz = K.lig_xyz.as<float4>();
  P.lig_row = K.lig_row.as<int>();
  P.lig_type = K.lig_type.as<int>();
  P.lig_q = K.lig_q.as<float>();
  P.kde = K.kde.as<float4>();
  P.kde_seg = K.kde_seg.as<int>();
  P.kde_inv_n = K.kde_inv.as<float>();
  P.mcs = K.mcs.as<float4>();
  P.mcs_tcc = K.mcs_tcc.as<float>();
  P.psp = K.psp.as<float>();
  P.hpl0 = K.hpl.as<float>();
  P.hpl1 = P.hpl0 + GD_MAXTP2;
  P.hpl2 = P.hpl0 + 2 * GD_MAXTP2;
  P.inv_hpl1 = P.hpl0 + 3 * GD_MAXTP2;
  P.minus_beta = K.minus_beta.as<float>();

  P.sc.lj0 = ep->lj0;
  P.sc.lj1 = ep->lj1;
  P.sc.el0 = ep->el0;
  P.sc.el1 = ep->el1;
  P.sc.a1 = ep->a1;
  P.sc.b1 = ep->b1;
  P.sc.inv_el1 = 1.0f / ep->el1;
  P.sc.pad = 0.0f;
  P.kde2 = ep->kde2;
  P.kde3 = ep->kde3;
  for (int i = 0; i < 3; ++i) {
    P.center[i] = lig->center[i];
    P.pocket_center[i] = prt->pocket_center[i];
  }
  for (int k = 0; k < 6; ++k)
    P.move_scale[k] = move_scale ? move_scale[k] : 0.0f;
  for (int k = 0; k < 10; ++k) {
    P.a_para[k] = ep->a_para[k];
    P.b_para[k] = ep->b_para[k];
    P.w[k] = ep->w[k];
  }
  P.pnk = pnk_used;

  /* LDS carve-up */
  int max_lds = 0;
  HIP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  gd_launch_cfg &C = K.C;
  auto al16 = [](int v) { return (v + 15) & ~15; };
  C.off_eh = al16(pnp_pad * 16);
  C.off_tab = al16(C.off_eh + pnp_pad * 8);
  C.off_kde = al16(C.off_tab + P.n_rows * GD_ROW_STRIDE * (int) sizeof(gd_pair_param));
  const int with_kde = C.off_kde + pnk_used * 16;
  if (with_kde <= max_lds) {
    C.kde_in_lds = 1;
    C.lds_bytes = with_kde;
  } else {
    C.kde_in_lds = 0;
    C.lds_bytes = C.off_kde;
  }
  if (C.lds_bytes > max_lds)
    return fail(GD_ERR_ARG, "pocket of %d points needs %d B of LDS (device has %d)",
                sz.pnp, C.lds_bytes, max_lds);
  HIP_TRY(gd_set_kernel_attrs(std::max(C.lds_bytes, 16)));
  return GD_OK;
}

/* Split index runs with equal protein conformer into tasks of <= wpg. */
static std::vector<gd_task> make_tasks(const std::vector<int> &prt_of, int wpg)
{
  std::vector<gd_task> tasks;
  const int n = (int) prt_of.size();
  int i = 0;
  while (i < n) {
    int j = i;
    while (j < n && prt_of[j] == prt_of[i] && j - i < wpg)
      ++j;
    gd_task t;
    t.prt = prt_of[i];
    t.r_begin = i;
    t.r_count = j - i;
    t.pad = 0;
    tasks.push_back(t);
    i = j;
  }
  return tasks;
}

/* ------------------------------------------------------------------------ */
/* session                                                                   */
/* ------------------------------------------------------------------------ */
struct gd_session {
  Pack K;
  gd_state S;
  DevBuf mv, e, mv_best, e_best, step_best, idx_tmp, acs_mc, acs_x;
  DevBuf tasks, rep_lig;
  std::vector<int> rep_lig_h, rep_prt_h;
  int n_tasks = 0;
  gd_complex_size sz;
  gd_mcpara mc;
  gd_options opt;
  hipStream_t stream = nullptr;
  long long steps_done = 0;
  int cycles_done = 0;
  int initialised = 0;
  double compute_s = 0.0;
  /* timing */
  int timing = 0;
  std::vector<hipEvent_t> ev_mc, ev_x;
  double mc_ms = 0.0, x_ms = 0.0;
  int mc_launches = 0, x_launches = 0;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  int pending_interval = 0;
};

static void drain_timing(gd_session *s)
{
  for (size_t i = 0; i + 1 < s->ev_mc.size(); i += 2) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, s->ev_mc[i], s->ev_mc[i + 1]) == hipSuccess) {
      s->mc_ms += ms;
      s->mc_launches++;
    }
    (void) hipEventDestroy(s->ev_mc[i]);
    (void) hipEventDestroy(s->ev_mc[i + 1]);
  }
  s->ev_mc.clear();
  for (size_t i = 0; i + 1 < s->ev_x.size(); i += 2) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, s->ev_x[i], s->ev_x[i + 1]) == hipSuccess) {
      s->x_ms += ms;
      s->x_launches++;
    }
    (void) hipEventDestroy(s->ev_x[i]);
    (void) hipEventDestroy(s->ev_x[i + 1]);
  }
  s->ev_x.clear();
}

extern "C" int gd_session_create(const gd_ligand *lig, const gd_protein *prt,
                                 const gd_psp *psp, const gd_kde *kde,
                                 const gd_mcs *mcs, const gd_enepara *enepara,
                                 const gd_temp *temp, const gd_replica *replica,
                                 const gd_mcpara *mcpara, gd_complex_size size,
                                 const gd_options *opt, gd_session **out)
{
  if (!out)
    return fail(GD_ERR_ARG, "null session out pointer");
  *out = nullptr;
  int rc = check_inputs(lig, prt, psp, kde, mcs, enepara, size);
  if (rc)
    return rc;
  if (!temp || !mcpara)
    return fail(GD_ERR_ARG, "null temp/mcpara");
  if (size.n_tmp < 1 || size.n_tmp > 64)
    return fail(GD_ERR_ARG, "n_tmp=%d outside [1, 64]", size.n_tmp);
  if ((long long) size.n_rep != (long long) size.n_lig * size.n_prt * size.n_tmp)
    return fail(GD_ERR_ARG, "n_rep=%d != n_lig*n_prt*n_tmp", size.n_rep);
  if (mcpara->steps_total < 0 || mcpara->steps_per_exchange < 1)
    return fail(GD_ERR_ARG, "steps_total=%d steps_per_exchange=%d",
                mcpara->steps_total, mcpara->steps_per_exchange);
  int ndev = gd_device_count();
  gd_options o;
  if (opt)
    o = *opt;
  else {
    o.seed = 1; o.complex_id = 0; o.exchange = 1; o.away = 0;
    o.term_order = GD_TERM_ORDER_FILE; o.device = 0;
  }
  if (ndev < 1)
    return fail(GD_ERR_NODEV, "no HIP device visible");
  if (o.device < 0 || o.device >= ndev)
    return fail(GD_ERR_ARG, "device %d of %d", o.device, ndev);

  gd_session *s = new gd_session();
  s->sz = size;
  s->mc = *mcpara;
  s->opt = o;
  rc = build_pack(s->K, lig, prt, psp, kde, mcs, enepara, temp, size, o.term_order,
                  mcpara->move_scale, o.device);
  if (rc) {
    delete s;
    return rc;
  }
  const int n = size.n_rep;
  s->rep_lig_h.resize(n);
  s->rep_prt_h.resize(n);
  std::vector<int> idx_tmp(n);
  for (int r = 0; r < n; ++r) {
    int p, t, l;
    if (replica) {
      p = replica[r].idx_prt; t = replica[r].idx_tmp; l = replica[r].idx_lig;
    } else {
      p = r / (size.n_tmp * size.n_lig);
      t = (r / size.n_lig) % size.n_tmp;
      l = r % size.n_lig;
    }
    if (p < 0 || p >= size.n_prt || t < 0 || t >= size.n_tmp || l < 0 || l >= size.n_lig) {
      delete s;
      return fail(GD_ERR_ARG, "replica %d has indices (%d,%d,%d) out of range", r, p, t, l);
    }
    s->rep_prt_h[r] = p;
    s->rep_lig_h[r] = l;
    idx_tmp[r] = t;
  }
  std::vector<gd_task> tasks = make_tasks(s->rep_prt_h, gd_waves_per_group());
  s->n_tasks = (int) tasks.size();
  std::vector<float> zeros8((size_t) n * 8, 0.0f), zeros10((size_t) n * 10, 0.0f);
  std::vector<int> zi(n, 0);
#define S_TRY(expr)                                                          \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      int c_ = fail(GD_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));      \
      delete s;                                                              \
      return c_;                                                             \
    }                                                                        \
  } while (0)
  S_TRY(s->mv.upload(zeros8));
  S_TRY(s->mv_best.upload(zeros8));
  S_TRY(s->e.upload(zeros10));
  S_TRY(s->e_best.upload(zeros10));
  S_TRY(s->step_best.upload(zi));
  S_TRY(s->idx_tmp.upload(idx_tmp));
  S_TRY(s->acs_mc.upload(zi));
  S_TRY(s->acs_x.upload(zi));
  S_TRY(s->tasks.upload(tasks));
  S_TRY(s->rep_lig.upload(s->rep_lig_h));
  S_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  S_TRY(hipEventCreate(&s->ev_begin));
  S_TRY(hipEventCreate(&s->ev_end));
#undef S_TRY
  s->S.mv = s->mv.as<float>();
  s->S.e = s->e.as<float>();
  s->S.mv_best = s->mv_best.as<float>();
  s->S.e_best = s->e_best.as<float>();
  s->S.step_best = s->step_best.as<int>();
  s->S.idx_tmp = s->idx_tmp.as<int>();
  s->S.acs_mc = s->acs_mc.as<int>();
  s->S.acs_x = s->acs_x.as<int>();
  *out = s;
  return GD_OK;
}

extern "C" void gd_session_destroy(gd_session *s)
{
  if (!s)
    return;
  (void) hipSetDevice(s->opt.device);
  if (s->stream)
    (void) hipStreamSynchronize(s->stream);
  drain_timing(s);
  if (s->ev_begin) (void) hipEventDestroy(s->ev_begin);
  if (s->ev_end) (void) hipEventDestroy(s->ev_end);
  if (s->stream)
    (void) hipStreamDestroy(s->stream);
  delete s;
}

extern "C" int gd_session_init(gd_session *s)
{
  if (!s)
    return fail(GD_ERR_ARG, "null session");
  HIP_TRY(hipSetDevice(s->opt.device));
  HIP_TRY(gd_launch_mc(s->K.P, s->S, s->tasks.as<gd_task>(), s->n_tasks,
                       s->rep_lig.as<int>(), s->K.C, 0, 0, s->opt.seed,
                       s->opt.complex_id, s->opt.away ? 2 : 1, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->initialised = 1;
  return GD_OK;
}

extern "C" int gd_session_cycles(gd_session *s, int n_cycles)
{
  if (!s)
    return fail(GD_ERR_ARG, "null session");
  if (!s->initialised)
    return fail(GD_ERR_ARG, "session not initialised");
  if (n_cycles < 0)
    return fail(GD_ERR_ARG, "negative cycle count");
  HIP_TRY(hipSetDevice(s->opt.device));
  const int spe = s->mc.steps_per_exchange;
  for (int c = 0; c < n_cycles; ++c) {
    const int s0 = (int) s->steps_done;
    if (s->timing) {
      hipEvent_t a, b;
      HIP_TRY(hipEventCreate(&a));
      HIP_TRY(hipEventCreate(&b));
      HIP_TRY(hipEventRecord(a, s->stream));
      HIP_TRY(gd_launch_mc(s->K.P, s->S, s->tasks.as<gd_task>(), s->n_tasks,
                           s->rep_lig.as<int>(), s->K.C, s0, spe, s->opt.seed,
                           s->opt.complex_id, 0, s->stream));
      HIP_TRY(hipEventRecord(b, s->stream));
      s->ev_mc.push_back(a);
      s->ev_mc.push_back(b);
    } else {
      HIP_TRY(gd_launch_mc(s->K.P, s->S, s->tasks.as<gd_task>(), s->n_tasks,
                           s->rep_lig.as<int>(), s->K.C, s0, spe, s->opt.seed,
                           s->opt.complex_id, 0, s->stream));
    }
    s->steps_done += spe;
    if (s->opt.exchange && s->sz.n_tmp > 1) {
      if (s->timing) {
        hipEvent_t a, b;
        HIP_TRY(hipEventCreate(&a));
        HIP_TRY(hipEventCreate(&b));
        HIP_TRY(hipEventRecord(a, s->stream));
        HIP_TRY(gd_launch_exchange(s->K.P, s->S, s->cycles_done, s->opt.seed,
                                   s->opt.complex_id, s->stream));
        HIP_TRY(hipEventRecord(b, s->stream));
        s->ev_x.push_back(a);
        s->ev_x.push_back(b);
      } else {
        HIP_TRY(gd_launch_exchange(s->K.P, s->S, s->cycles_done, s->opt.seed,
                                   s->opt.complex_id, s->stream));
      }
    }
    s->cycles_done++;
  }
  return GD_OK;
}

/* a shorter final cycle when steps_total is not a multiple of steps_per_exchange */
static int session_tail(gd_session *s, int n_steps)
{
  HIP_TRY(gd_launch_mc(s->K.P, s->S, s->tasks.as<gd_task>(), s->n_tasks,
                       s->rep_lig.as<int>(), s->K.C, (int) s->steps_done, n_steps,
                       s->opt.seed, s->opt.complex_id, 0, s->stream));
  s->steps_done += n_steps;
  if (s->opt.exchange && s->sz.n_tmp > 1)
    HIP_TRY(gd_launch_exchange(s->K.P, s->S, s->cycles_done, s->opt.seed,
                               s->opt.complex_id, s->stream));
  s->cycles_done++;
  return GD_OK;
}

extern "C" int gd_session_sync(gd_session *s)
{
  if (!s)
    return fail(GD_ERR_ARG, "null session");
  HIP_TRY(hipSetDevice(s->opt.device));
  HIP_TRY(hipStreamSynchronize(s->stream));
  drain_timing(s);
  return GD_OK;
}

extern "C" long long gd_session_steps_done(const gd_session *s)
{
  return s ? s->steps_done : -1;
}

extern "C" int gd_session_set_timing(gd_session *s, int on)
{
  if (!s)
    return fail(GD_ERR_ARG, "null session");
  s->timing = on ? 1 : 0;
  return GD_OK;
}

extern "C" int gd_session_timing(gd_session *s, double *mc_ms, int *mc_launches,
                                 double *x_ms, int *x_launches)
{
  if (!s)
    return fail(GD_ERR_ARG, "null session");
  drain_timing(s);
  if (mc_ms) *mc_ms = s->mc_ms;
  if (mc_launches) *mc_launches = s->mc_launches;
  if (x_ms) *x_ms = s->x_ms;
  if (x_launches) *x_launches = s->x_launches;
  return GD_OK;
}

extern "C" int gd_session_layout(const gd_session *s, int *lds_bytes, int *kde_in_lds,
                                 int *n_tasks, int *waves_per_group)
{
  if (!s)
    return fail(GD_ERR_ARG, "null session");
  if (lds_bytes) *lds_bytes = s->K.C.lds_bytes;
  if (kde_in_lds) *kde_in_lds = s->K.C.kde_in_lds;
  if (n_tasks) *n_tasks = s->n_tasks;
  if (waves_per_group) *waves_per_group = gd_waves_per_group();
  return GD_OK;
}

extern "C" int gd_session_results(gd_session *s, gd_mclog *mclog,
                                  gd_replica_result *results)
{
  if (!s)
    return fail(GD_ERR_ARG, "null session");
  HIP_TRY(hipSetDevice(s->opt.device));
  HIP_TRY(hipStreamSynchronize(s->stream));
  const int n = s->sz.n_rep;
  std::vector<float> mv((size_t) n * 8), e((size_t) n * 10), mvb((size_t) n * 8),
      eb((size_t) n * 10);
  std::vector<int> stepb(n), it(n), acs(n), acx(n);
  HIP_TRY(hipMemcpy(mv.data(), s->mv.p, mv.size() * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(e.data(), s->e.p, e.size() * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(mvb.data(), s->mv_best.p, mvb.size() * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(eb.data(), s->e_best.p, eb.size() * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(stepb.data(), s->step_best.p, n * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(it.data(), s->idx_tmp.p, n * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(acs.data(), s->acs_mc.p, n * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(acx.data(), s->acs_x.p, n * 4, hipMemcpyDeviceToHost));

  std::vector<double> e64;
  if (results) {
    /* FP64 re-score of every replica's best pose */
    std::vector<float> moves((size_t) n * 6);
    for (int r = 0; r < n; ++r)
      for (int k = 0; k < 6; ++k)
        moves[(size_t) r * 6 + k] = mvb[(size_t) r * 8 + k];
    DevBuf d_mv, d_lc, d_pc, d_e;
    HIP_TRY(d_mv.upload(moves));
    HIP_TRY(d_lc.upload(s->rep_lig_h));
    HIP_TRY(d_pc.upload(s->rep_prt_h));
    HIP_TRY(d_e.alloc((size_t) n * 10 * sizeof(double)));
    HIP_TRY(gd_launch_rescore64(s->K.P, d_lc.as<int>(), d_pc.as<int>(),
                                d_mv.as<float>(), n, d_e.as<double>(), s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    e64.resize((size_t) n * 10);
    HIP_TRY(hipMemcpy(e64.data(), d_e.p, e64.size() * 8, hipMemcpyDeviceToHost));
    for (int r = 0; r < n; ++r) {
      gd_replica_result &R = results[r];
      for (int k = 0; k < 6; ++k) {
        R.move_final[k] = mv[(size_t) r * 8 + k];
        R.move_best[k] = mvb[(size_t) r * 8 + k];
      }
      for (int i = 0; i < 10; ++i) {
        R.energy_final[i] = e[(size_t) r * 10 + i];
        R.energy_best[i] = eb[(size_t) r * 10 + i];
        R.energy_best_fp64[i] = e64[(size_t) r * 10 + i];
      }
      R.step_best = stepb[r];
      R.idx_tmp_final = it[r];
      R.acs_mc = acs[r];
      R.acs_exchg = acx[r];
    }
  }
  if (mclog) {
    for (int r = 0; r < n; ++r) {
      mclog->ac_mc += acs[r];
      mclog->ac_temp_exchg += acx[r];
      if (mclog->acs_mc)
        mclog->acs_mc[r] += acs[r];
      if (mclog->acs_temp_exchg)
        mclog->acs_temp_exchg[r] += acx[r];
    }
  }
  return GD_OK;
}

/* ------------------------------------------------------------------------ */
/* gd_run: drop-in for Run() (run.cu:40-352)                                 */
/* ------------------------------------------------------------------------ */
extern "C" int gd_run(const gd_ligand *lig, const gd_protein *prt, const gd_psp *psp,
                      const gd_kde *kde, const gd_mcs *mcs, const gd_enepara *enepara,
                      const gd_temp *temp, const gd_replica *replica,
                      const gd_mcpara *mcpara, gd_mclog *mclog, gd_complex_size size,
                      const gd_options *opt, gd_replica_result *results)
{
  gd_session *s = nullptr;
  int rc = gd_session_create(lig, prt, psp, kde, mcs, enepara, temp, replica, mcpara,
                             size, opt, &s);
  if (rc)
    return rc;
  auto t_begin = std::chrono::steady_clock::now();
  rc = gd_session_init(s);
  if (!rc) {
    const int spe = mcpara->steps_per_exchange;
    const int full = mcpara->steps_total / spe;
    const int tail = mcpara->steps_total - full * spe;
    rc = gd_session_cycles(s, full);
    if (!rc && tail > 0)
      rc = session_tail(s, tail);
    if (!rc)
      rc = gd_session_sync(s);
  }
  auto t_comp = std::chrono::steady_clock::now();
  if (!rc)
    rc = gd_session_results(s, mclog, results);
  auto t_end = std::chrono::steady_clock::now();
  if (!rc && mclog) {
    mclog->t0 += std::chrono::duration<double>(t_comp - t_begin).count();
    mclog->t1 += std::chrono::duration<double>(t_end - t_begin).count();
  }
  gd_session_destroy(s);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* gd_score_poses                                                            */
/* ------------------------------------------------------------------------ */
extern "C" int gd_score_poses(const gd_ligand *lig, const gd_protein *prt,
                              const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                              const gd_enepara *enepara, gd_complex_size size,
                              int n_pose, const int *lig_conf, const int *prt_conf,
                              const float *move, int term_order, int device,
                              float *energy, double *energy_fp64)
{
  int rc = check_inputs(lig, prt, psp, kde, mcs, enepara, size);
  if (rc)
    return rc;
  if (n_pose < 0 || (n_pose > 0 && (!lig_conf || !prt_conf || !move)))
    return fail(GD_ERR_ARG, "bad pose arrays");
  if (!energy && !energy_fp64)
    return fail(GD_ERR_ARG, "no output requested");
  for (int i = 0; i < n_pose; ++i)
    if (lig_conf[i] < 0 || lig_conf[i] >= size.n_lig || prt_conf[i] < 0 ||
        prt_conf[i] >= size.n_prt)
      return fail(GD_ERR_ARG, "pose %d: conformer index out of range", i);
  const int ndev = gd_device_count();
  if (ndev < 1)
    return fail(GD_ERR_NODEV, "no HIP device visible");
  if (device < 0 || device >= ndev)
    return fail(GD_ERR_ARG, "device %d of %d", device, ndev);
  if (n_pose == 0)
    return GD_OK;
  Pack K;
  rc = build_pack(K, lig, prt, psp, kde, mcs, enepara, nullptr, size, term_order,
                  nullptr, device);
  if (rc)
    return rc;
  /* poses grouped by protein conformer */
  std::vector<int> order(n_pose);
  for (int i = 0; i < n_pose; ++i)
    order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return prt_conf[a] < prt_conf[b]; });
  std::vector<int> prt_sorted(n_pose);
  for (int i = 0; i < n_pose; ++i)
    prt_sorted[i] = prt_conf[order[i]];
  std::vector<gd_task> tasks = make_tasks(prt_sorted, gd_waves_per_group());
  std::vector<int> lc(lig_conf, lig_conf + n_pose), pc(prt_conf, prt_conf + n_pose);
  std::vector<float> mv(move, move + (size_t) n_pose * 6);
  DevBuf d_tasks, d_order, d_lc, d_pc, d_mv, d_e, d_e64;
  HIP_TRY(d_tasks.upload(tasks));
  HIP_TRY(d_order.upload(order));
  HIP_TRY(d_lc.upload(lc));
  HIP_TRY(d_pc.upload(pc));
  HIP_TRY(d_mv.upload(mv));
  hipStream_t st;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  int err = GD_OK;
  hipError_t he = hipSuccess;
  if (energy) {
    he = d_e.alloc((size_t) n_pose * 10 * sizeof(float));
    if (he == hipSuccess)
      he = gd_launch_score(K.P, d_tasks.as<gd_task>(), (int) tasks.size(),
                           d_order.as<int>(), d_lc.as<int>(), d_mv.as<float>(), K.C,
                           d_e.as<float>(), st);
  }
  if (he == hipSuccess && energy_fp64) {
    he = d_e64.alloc((size_t) n_pose * 10 * sizeof(double));
    if (he == hipSuccess)
      he = gd_launch_rescore64(K.P, d_lc.as<int>(), d_pc.as<int>(), d_mv.as<float>(),
                               n_pose, d_e64.as<double>(), st);
  }
  if (he == hipSuccess)
    he = hipStreamSynchronize(st);
  if (he == hipSuccess && energy)
    he = hipMemcpy(energy, d_e.p, (size_t) n_pose * 10 * sizeof(float),
                   hipMemcpyDeviceToHost);
  if (he == hipSuccess && energy_fp64)
    he = hipMemcpy(energy_fp64, d_e64.p, (size_t) n_pose * 10 * sizeof(double),
                   hipMemcpyDeviceToHost);
  if (he != hipSuccess)
    err = fail(GD_ERR_HIP, "score: %s", hipGetErrorString(he));
  (void) hipStreamDestroy(st);
  return err;
}
