/*
 * gd_image.cpp -- host image of one complex and its upload into HBM.
 *
 * The image holds every device array of a complex in its final layout:
 * pocket points per protein conformer in the canonical order (stable sort
 * by gd_point_sort_key, padded to whole wavefronts with points that add
 * exact zeros), pair-parameter rows for the ligand's atom types, KDE points
 * grouped by those rows, position restraints resolved per (restraint,
 * atom), PSP, temperatures and the replica table (SetReplica, util.C:537-558).
 * The upload places all of it, the replica state and the step counters in
 * one contiguous arena block.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "gd_internal.h"
#include "gd_engine.h"

/* ------------------------------------------------------------------------ */
/* host image of a complex                                                   */
/* ------------------------------------------------------------------------ */
namespace {

template <class T> size_t vbytes(const std::vector<T> &v) { return v.size() * sizeof(T); }

}  // namespace

int gd_build_image(GdImage &I, const gd_ligand *lig, const gd_protein *prt,
                   const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                   const gd_enepara *ep, const gd_temp *temp, const gd_replica *replica,
                   const gd_mcpara *mc, const gd_complex_size &sz, const gd_options &o,
                   int max_lds, int kde_policy)
{
  gd_pack &P = I.P;
  memset(&P, 0, sizeof P);
  P.n_lig = sz.n_lig; P.n_prt = sz.n_prt; P.n_tmp = sz.n_tmp; P.n_rep = sz.n_rep;
  P.lna = sz.lna; P.pnp = sz.pnp; P.pnk = sz.pnk; P.pos = sz.pos;
  P.n_res = psp->n_res;
  P.term_order = o.term_order;
  P.combine = o.combine;
  P.steps_run = 0;
  P.spe = mc ? mc->steps_per_exchange : 1;
  P.lpw = 1;
  if (mc) {
    const int rc = gd_schedule(mc, &P.steps_run, &P.lpw);
    if (rc)
      return rc;
  }
  P.exchange = o.exchange;
  P.away = o.away;
  P.fixed_draws = o.fixed_draws != 0;
  P.reverse = o.reverse != 0;
  P.seed = o.seed;
  P.cid = o.complex_id;
  if (o.record_stride > 0 && mc) {
    /* records of one steps_per_dump window stay on the device until read */
    const int w = mc->steps_per_dump;
    if (w < 1 || w % mc->steps_per_exchange != 0 || w % o.record_stride != 0)
      return gd_fail(GD_ERR_ARG,
                     "records need steps_per_dump (%d) to be a multiple of "
                     "steps_per_exchange (%d) and of record_stride (%d)",
                     w, mc->steps_per_exchange, o.record_stride);
    P.rec_stride = o.record_stride;
    P.rec_window = w;
    P.rec_per_window = w / o.record_stride;
  }

  /* table rows: distinct ligand atom types in order of first appearance */
  int row_of_type[GD_MAXTP2];
  for (int t = 0; t < GD_MAXTP2; ++t)
    row_of_type[t] = -1;
  std::vector<int> row_type;
  I.lig_row.resize(sz.lna);
  I.lig_type.resize(sz.lna);
  I.lig_q.resize(sz.lna);
  for (int a = 0; a < sz.lna; ++a) {
    const int t = lig->t[a];
    if (row_of_type[t] < 0) {
      row_of_type[t] = (int) row_type.size();
      row_type.push_back(t);
    }
    I.lig_row[a] = row_of_type[t];
    I.lig_type[a] = t;
    I.lig_q[a] = lig->c[a];
  }
  P.n_rows = (int) row_type.size();

  I.tab.assign((size_t) P.n_rows * GD_ROW_STRIDE, gd_pair_param());
  memset(I.tab.data(), 0, vbytes(I.tab));
  for (int r = 0; r < P.n_rows; ++r) {
    const int lt = row_type[r];
    for (int pt = 0; pt < GD_MAXTP1; ++pt) {
      gd_pair_param &q = I.tab[(size_t) r * GD_ROW_STRIDE + pt];
      q.p1a = ep->p1a[lt][pt];
      q.p2a = ep->p2a[lt][pt];
      q.pmf0 = ep->pmf0[lt][pt];
      q.pmf1 = ep->pmf1[lt][pt];
      q.hdb0 = ep->hdb0[lt][pt];
      q.hdb1 = ep->hdb1[lt][pt];
    }
  }
  I.row_hdb.assign(P.n_rows, 0);
  for (int r = 0; r < P.n_rows; ++r)
    for (int pt = 0; pt < GD_MAXTP1; ++pt)
      if (ep->hdb0[row_type[r]][pt] > 0.1f)
        I.row_hdb[r] = 1;

  /* pocket points per conformer, canonical order (stable by
   * gd_point_sort_key), padded to a multiple of 64 with zero points */
  const int pnp_pad = (sz.pnp + GD_WAVE - 1) / GD_WAVE * GD_WAVE;
  P.pnp_pad = pnp_pad;
  const size_t np = (size_t) sz.n_prt * pnp_pad;
  {
    const uint32_t pbits = (uint32_t) GD_PAD_TYPE << GD_PT_TYPE_SHIFT;
    float pw;
    memcpy(&pw, &pbits, 4);
    I.pt_pos.assign(np, make_float4(1.0e3f, 1.0e3f, 1.0e3f, pw));
    I.pt_eh.assign(np, make_float2(0.0f, 0.0f));
  }
  const uint32_t hdb_mask = gd_hdb_point_mask(&ep->hdb0[0][0], GD_MAXTP2, GD_MAXTP1);
  std::vector<int> perm(sz.pnp);
  std::vector<int> perms((size_t) sz.n_prt * sz.pnp);   /* sorted slot -> point */
  I.chunk_bounds.assign((size_t) 2 * sz.n_prt, 0);
  for (int cf = 0; cf < sz.n_prt; ++cf) {
    const size_t base = (size_t) cf * sz.pnp;
    const size_t obase = (size_t) cf * pnp_pad;
    for (int j = 0; j < sz.pnp; ++j)
      perm[j] = j;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) {
      return gd_point_sort_key(prt->t[base + a], prt->c[base + a] == 2,
                               prt->c0_and_d12_or_c2[base + a] != 0, hdb_mask) <
             gd_point_sort_key(prt->t[base + b], prt->c[base + b] == 2,
                               prt->c0_and_d12_or_c2[base + b] != 0, hdb_mask);
    });
    for (int j = 0; j < sz.pnp; ++j)
      perms[(size_t) cf * sz.pnp + j] = perm[j];
    int n_hpc = 0, end_c2 = 0;
    for (int j = 0; j < sz.pnp; ++j) {
      n_hpc += prt->c0_and_d12_or_c2[base + j] != 0;
      if (prt->c[base + perm[j]] == 2)
        end_c2 = j + 1;
    }
    /* [0, kh) holds every hydrophobic-class point; [0, kp) every class-2
     * point (the only ones with a PSP value); class 2 implies the
     * hydrophobic class in every input the loaders build, so kp = kh there */
    const int kh = (n_hpc + GD_WAVE - 1) / GD_WAVE;
    const int kc = (end_c2 + GD_WAVE - 1) / GD_WAVE;
    I.chunk_bounds[2 * cf] = kh;
    I.chunk_bounds[2 * cf + 1] = kc > kh ? kc : kh;
    for (int j = 0; j < sz.pnp; ++j) {
      const size_t i = base + perm[j];
      const uint32_t bits = (uint32_t) prt->t[i] << GD_PT_TYPE_SHIFT;
      float w;
      memcpy(&w, &bits, 4);
      I.pt_pos[obase + j] = make_float4(prt->x[i], prt->y[i], prt->z[i], w);
      /* hpp of points outside the hydrophobic class is stored as 0: their
       * contribution is then an exact +0 and the loop needs no flag test */
      I.pt_eh[obase + j] =
          make_float2(prt->ele[i], prt->c0_and_d12_or_c2[i] ? prt->hpp[i] : 0.0f);
    }
  }

  I.lig_xyz.resize((size_t) sz.n_lig * sz.lna);
  for (size_t i = 0; i < I.lig_xyz.size(); ++i)
    I.lig_xyz[i] = make_float4(lig->x[i], lig->y[i], lig->z[i], 0.0f);

  /* KDE points grouped by table row, original order within a type */
  I.kde.clear();
  I.kde_seg.assign(2 * (size_t) P.n_rows, 0);
  for (int r = 0; r < P.n_rows; ++r) {
    I.kde_seg[2 * r] = (int) I.kde.size();
    for (int k = 0; k < sz.pnk; ++k)
      if (kde->t[k] == row_type[r])
        I.kde.push_back(make_float4(kde->x[k], kde->y[k], kde->z[k], 0.0f));
    I.kde_seg[2 * r + 1] = (int) I.kde.size();
  }
  P.pnk = (int) I.kde.size();
  if (P.pnk > 0xfffe)
    return gd_fail(GD_ERR_ARG, "%d KDE points of the ligand's atom types (at most 65534)",
                   P.pnk);
  I.kde_inv.resize(P.n_rows);
  for (int r = 0; r < P.n_rows; ++r) {
    const int n = I.kde_seg[2 * r + 1] - I.kde_seg[2 * r];
    I.kde_inv[r] = n > 0 ? 1.0f / (float) n : 0.0f;
    /* the weight of a pair rides with its KDE point: every point of the
     * segment pairs with atoms of this row only */
    for (int k = I.kde_seg[2 * r]; k < I.kde_seg[2 * r + 1]; ++k)
      I.kde[k].w = I.kde_inv[r];
  }
  /* one more point after the segments, of weight 0: the padding pairs name
   * it, so they add exact zeros and need no mask */
  I.kde.push_back(make_float4(0.0f, 0.0f, 0.0f, 0.0f));
  const uint32_t kde_pad = (uint32_t) P.pnk << 16;
  /* the pose's (atom, KDE point) pairs, atoms < 64 then atoms >= 64 (lane
   * a holds atoms a and a + 64), each group padded to whole wavefronts */
  I.kde_pairs.clear();
  for (int g = 0; g < 2; ++g) {
    const int la = g ? GD_WAVE : 0;
    const int lb = g ? sz.lna : std::min(sz.lna, GD_WAVE);
    const size_t start = I.kde_pairs.size();
    for (int a = la; a < lb; ++a) {
      const int r = I.lig_row[a];
      for (int k = I.kde_seg[2 * r]; k < I.kde_seg[2 * r + 1]; ++k)
        I.kde_pairs.push_back((uint32_t) k << 16 | (uint32_t) (a & (GD_WAVE - 1)) << 2);
    }
    const int n = (int) (I.kde_pairs.size() - start);
    (g ? P.kde_n_hi : P.kde_n_lo) = n;
    I.kde_pairs.resize(start + (size_t) (n + GD_WAVE - 1) / GD_WAVE * GD_WAVE, kde_pad);
  }
  if (I.kde_pairs.empty())
    I.kde_pairs.push_back(kde_pad);

  I.mcs.resize((size_t) sz.pos * sz.lna);
  I.tcc.resize(sz.pos);
  for (int m = 0; m < sz.pos; ++m) {
    I.tcc[m] = mcs->tcc[m];
    for (int a = 0; a < sz.lna; ++a) {
      const size_t f = (size_t) m * mcs->n_field + lig->n[a] + 1;
      const float x = mcs->x[f];
      I.mcs[(size_t) m * sz.lna + a] = x != GD_MCS_INVALID_COORD
                                           ? make_float4(x, mcs->y[f], mcs->z[f], 1.0f)
                                           : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  }

  /* PSP value of every (protein conformer, table row, sorted point):
   * psp[row type][seq3r] for class-2 points, 0 for the others and for
   * padding, so the energy loop reads one contiguous row per ligand atom
   * and adds it under the PMF-radius test alone */
  I.psp_pt.assign((size_t) sz.n_prt * P.n_rows * pnp_pad, 0.0f);
  for (int cf = 0; cf < sz.n_prt; ++cf)
    for (int r = 0; r < P.n_rows; ++r) {
      float *dst = &I.psp_pt[((size_t) cf * P.n_rows + r) * pnp_pad];
      const float *src = psp->psp + (size_t) row_type[r] * psp->n_res;
      for (int j = 0; j < sz.pnp; ++j) {
        const size_t i = (size_t) cf * sz.pnp + perms[(size_t) cf * sz.pnp + j];
        if (prt->c[i] == 2)
          dst[j] = src[prt->seq3r[i]];
      }
    }
  I.hpl.resize(4 * GD_MAXTP2);
  for (int t = 0; t < GD_MAXTP2; ++t) {
    I.hpl[t] = ep->hpl0[t];
    I.hpl[GD_MAXTP2 + t] = ep->hpl1[t];
    I.hpl[2 * GD_MAXTP2 + t] = ep->hpl2[t];
    I.hpl[3 * GD_MAXTP2 + t] = 1.0f / ep->hpl1[t];
  }
  I.atom.assign(sz.lna, gd_atom_rec());
  for (int a = 0; a < sz.lna; ++a) {
    gd_atom_rec &ar = I.atom[a];
    const int t = I.lig_type[a];
    ar.row = I.lig_row[a];
    ar.q = I.lig_q[a];
    ar.hdb = I.row_hdb[ar.row];
    ar.hpl0 = I.hpl[t];
    ar.inv_hpl1 = I.hpl[3 * GD_MAXTP2 + t];
    ar.hpl2 = I.hpl[2 * GD_MAXTP2 + t];
    ar.pad0 = ar.pad1 = 0;
  }
  I.mb.assign(std::max(sz.n_tmp, 1), 0.0f);
  if (temp)
    for (int t = 0; t < sz.n_tmp; ++t)
      I.mb[t] = temp[t].minus_beta;

  /* replicas */
  const int n = sz.n_rep;
  I.rep_lig.resize(n);
  I.rep_prt.resize(n);
  I.idx_tmp.resize(n);
  for (int r = 0; r < n; ++r) {
    int p, t, l;
    if (replica) {
      p = replica[r].idx_prt; t = replica[r].idx_tmp; l = replica[r].idx_lig;
    } else {
      p = r / (sz.n_tmp * sz.n_lig);
      t = (r / sz.n_lig) % sz.n_tmp;
      l = r % sz.n_lig;
    }
    if (p < 0 || p >= sz.n_prt || t < 0 || t >= sz.n_tmp || l < 0 || l >= sz.n_lig)
      return gd_fail(GD_ERR_ARG, "replica %d has indices (%d,%d,%d) out of range", r, p, t, l);
    I.rep_prt[r] = p;
    I.rep_lig[r] = l;
    I.idx_tmp[r] = t;
  }
  /* The exchange treats slots n_tmp*n_lig*p + n_lig*t + l, t = 0..n_tmp-1, as
   * the temperature chain of (p, l) (exchangereplicas_d.cu:28-38) and inverts
   * their temperature indices: every slot of the chain must hold protein
   * conformer p and ligand conformer l, as SetReplica lays them out
   * (util.C:537-558), and the chain's temperatures must be a permutation of
   * 0..n_tmp-1. */
  if (replica) {
    std::vector<int> seen(sz.n_tmp);
    for (int p = 0; p < sz.n_prt; ++p)
      for (int l = 0; l < sz.n_lig; ++l) {
        std::fill(seen.begin(), seen.end(), 0);
        for (int t = 0; t < sz.n_tmp; ++t) {
          const int r = sz.n_tmp * sz.n_lig * p + sz.n_lig * t + l;
          if (I.rep_prt[r] != p || I.rep_lig[r] != l)
            return gd_fail(GD_ERR_ARG,
                           "replica %d holds conformers (prt %d, lig %d); its slot belongs to "
                           "the chain of (prt %d, lig %d)",
                           r, I.rep_prt[r], I.rep_lig[r], p, l);
          if (seen[I.idx_tmp[r]]++)
            return gd_fail(GD_ERR_ARG,
                           "temperature %d appears twice in the chain of (prt %d, lig %d)",
                           I.idx_tmp[r], p, l);
        }
      }
  }

  gd_pair_scalars_init(&P.sc, ep->lj0, ep->lj1, ep->el0, ep->el1, ep->a1, ep->b1);

  /* MCC reference contacts (InitRefMatrix_d, calcmcc_d.cu:57-83): replica 0's
   * ligand conformer at its start pose (movematrix 0, before IS_AWAY) against
   * replica 0's protein conformer, contact = d <= pmf0 with the pair loop's
   * own d; stored per protein conformer in that conformer's point order */
  I.mcc_ref.clear();
  if (o.mcc && mc) {
    const int nchunk = pnp_pad / GD_WAVE;
    const int l0 = I.rep_lig[0], p0 = I.rep_prt[0];
    const float m0[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float rot[9];
    gd_rotation(m0, rot);
    std::vector<char> ref((size_t) sz.lna * sz.pnp);
    for (int a = 0; a < sz.lna; ++a) {
      const size_t li = (size_t) l0 * sz.lna + a;
      float lx, ly, lz;
      gd_transform(rot, m0, lig->center, lig->x[li], lig->y[li], lig->z[li], &lx, &ly, &lz);
      for (int q = 0; q < sz.pnp; ++q) {
        const size_t pi = (size_t) p0 * sz.pnp + q;
        const float d2 = gd_dist2(lx - prt->x[pi], ly - prt->y[pi], lz - prt->z[pi]);
        const float d = d2 * gd_rsqrt(d2);
        ref[(size_t) a * sz.pnp + q] = (d - ep->pmf0[lig->t[a]][prt->t[pi]]) <= 0.0f;
      }
    }
    I.mcc_ref.assign((size_t) sz.n_prt * sz.lna * nchunk, 0ull);
    for (int cf = 0; cf < sz.n_prt; ++cf)
      for (int a = 0; a < sz.lna; ++a)
        for (int j = 0; j < sz.pnp; ++j)
          if (ref[(size_t) a * sz.pnp + perms[(size_t) cf * sz.pnp + j]])
            I.mcc_ref[((size_t) cf * sz.lna + a) * nchunk + j / GD_WAVE] |=
                1ull << (j % GD_WAVE);
  }
  P.kde2 = ep->kde2;
  P.kde3 = ep->kde3;
  for (int i = 0; i < 3; ++i) {
    P.center[i] = lig->center[i];
    P.pocket_center[i] = prt->pocket_center[i];
  }
  for (int k = 0; k < 6; ++k)
    P.move_scale[k] = mc ? mc->move_scale[k] : 0.0f;
  for (int k = 0; k < 10; ++k) {
    P.a_para[k] = ep->a_para[k];
    P.b_para[k] = ep->b_para[k];
    P.w[k] = ep->w[k];
  }

  /* LDS carve-up */
  auto al16 = [](int v) { return (v + 15) & ~15; };
  /* pocket points staged in 64-point chunks of three 8-byte planes,
   * {x, y}, {z, type}, {ele, hpp}: contiguous per lane, so free of bank
   * conflicts, and one address per chunk; a pocket too large for LDS
   * beside the table rows is read through the caches instead
   * (whole-protein pockets) */
  const int reserve = 16 * GD_WSLOT_FLOATS * 4 + 64;   /* wave slots of up to 16 waves */
  const int tab_bytes = P.n_rows * GD_ROW_STRIDE * (int) sizeof(gd_pair_param);
  P.pocket_in_lds = 1;
  P.off_tab = (pnp_pad / GD_WAVE) * GD_LDS_CHUNK_BYTES;
  if (P.off_tab + tab_bytes + reserve > max_lds) {
    P.pocket_in_lds = 0;
    P.off_tab = 0;
  }
  P.off_kde = al16(P.off_tab + tab_bytes);
  const int with_kde = al16(P.off_kde + (P.pnk + 1) * 16);   /* with the padding point */
  if (kde_policy != GD_KDE_GLOBAL && with_kde + reserve <= max_lds) {
    P.kde_in_lds = 1;
    P.lds_bytes = with_kde;
  } else {
    P.kde_in_lds = 0;
    P.lds_bytes = P.off_kde;
  }
  if (P.lds_bytes + reserve > max_lds)
    return gd_fail(GD_ERR_ARG, "%d ligand atom types need %d B of LDS (device has %d)",
                   P.n_rows, P.lds_bytes + reserve, max_lds);
  /* split launches: pair list, restraint targets and PSP rows after the KDE
   * points, beside the wave slot and the partial-sum area of one replica */
  const int split_reserve = GD_WSLOT_FLOATS * 4 + (int) GD_SPLIT_AREA_BYTES + 64;
  P.off_pairs = with_kde;
  P.off_mcs = al16(P.off_pairs + (int) I.kde_pairs.size() * 4);
  P.off_psp = al16(P.off_mcs + P.pos * P.lna * 16);
  const int split_bytes = al16(P.off_psp + P.n_rows * pnp_pad * 4);
  P.split_extra = P.pocket_in_lds && P.kde_in_lds && split_bytes + split_reserve <= max_lds;
  P.split_lds_bytes = P.split_extra ? split_bytes : P.lds_bytes;
  return GD_OK;
}

GdBlock gd_layout_block(const GdImage &I)
{
  GdBlock b;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += GdArena::round(bytes ? bytes : 1);
    return o;
  };
  const size_t n = I.P.n_rep;
  b.pt_pos = take(vbytes(I.pt_pos));
  b.pt_eh = take(vbytes(I.pt_eh));
  b.tab = take(vbytes(I.tab));
  b.lig_xyz = take(vbytes(I.lig_xyz));
  b.lig_row = take(vbytes(I.lig_row));
  b.lig_type = take(vbytes(I.lig_type));
  b.lig_q = take(vbytes(I.lig_q));
  b.kde = take(vbytes(I.kde));
  b.kde_pairs = take(vbytes(I.kde_pairs));
  b.kde_seg = take(vbytes(I.kde_seg));
  b.kde_inv = take(vbytes(I.kde_inv));
  b.mcs = take(vbytes(I.mcs));
  b.tcc = take(vbytes(I.tcc));
  b.psp_pt = take(vbytes(I.psp_pt));
  b.chunk_bounds = take(vbytes(I.chunk_bounds));
  b.row_hdb = take(vbytes(I.row_hdb));
  b.atom = take(vbytes(I.atom));
  b.hpl = take(vbytes(I.hpl));
  b.mb = take(vbytes(I.mb));
  b.rep_lig = take(n * 4);
  b.rep_prt = take(n * 4);
  b.mv = take(n * 8 * 4);
  b.e = take(n * 10 * 4);
  b.mv_best = take(n * 8 * 4);
  b.e_best = take(n * 10 * 4);
  b.step_best = take(n * 4);
  b.idx_tmp = take(n * 4);
  b.acs_mc = take(n * 4);
  b.acs_x = take(n * 4);
  b.cmcc = take(n * 4);
  b.cmcc_best = take(n * 4);
  b.step = take(2 * 4);
  b.pack = take(sizeof(gd_pack));
  b.rec_bytes = (size_t) n * I.P.rec_per_window * GD_REC_WORDS * 4;
  b.rec = I.P.rec_stride > 0 ? take(b.rec_bytes) : 0;
  b.mcc_ref = I.mcc_ref.empty() ? 0 : take(vbytes(I.mcc_ref));
  b.total = off;
  return b;
}

size_t gd_image_footprint(const GdImage &I) { return gd_layout_block(I).total; }

/* Lay an image out as its arena block: the arrays at their block offsets in
 * host memory `host` (gd_image_footprint bytes), the replica state and step
 * counters zeroed, the pack's pointers aimed at the block's device address
 * `dev`, and the pack itself at b.pack.  One copy of `host` to `dev` then
 * makes the complex resident. */
void gd_pack_image(GdImage &I, char *host, char *dev, GdBlock &b)
{
  b = gd_layout_block(I);
  memset(host, 0, b.total);
  gd_pack &P = I.P;
  auto put = [&](size_t off, const void *src, size_t bytes) {
    if (bytes)
      memcpy(host + off, src, bytes);
  };
  put(b.pt_pos, I.pt_pos.data(), vbytes(I.pt_pos));
  put(b.pt_eh, I.pt_eh.data(), vbytes(I.pt_eh));
  put(b.tab, I.tab.data(), vbytes(I.tab));
  put(b.lig_xyz, I.lig_xyz.data(), vbytes(I.lig_xyz));
  put(b.lig_row, I.lig_row.data(), vbytes(I.lig_row));
  put(b.lig_type, I.lig_type.data(), vbytes(I.lig_type));
  put(b.lig_q, I.lig_q.data(), vbytes(I.lig_q));
  put(b.kde, I.kde.data(), vbytes(I.kde));
  put(b.kde_pairs, I.kde_pairs.data(), vbytes(I.kde_pairs));
  put(b.kde_seg, I.kde_seg.data(), vbytes(I.kde_seg));
  put(b.kde_inv, I.kde_inv.data(), vbytes(I.kde_inv));
  put(b.mcs, I.mcs.data(), vbytes(I.mcs));
  put(b.tcc, I.tcc.data(), vbytes(I.tcc));
  put(b.psp_pt, I.psp_pt.data(), vbytes(I.psp_pt));
  put(b.chunk_bounds, I.chunk_bounds.data(), vbytes(I.chunk_bounds));
  put(b.row_hdb, I.row_hdb.data(), vbytes(I.row_hdb));
  put(b.atom, I.atom.data(), vbytes(I.atom));
  put(b.hpl, I.hpl.data(), vbytes(I.hpl));
  put(b.mb, I.mb.data(), vbytes(I.mb));
  put(b.rep_lig, I.rep_lig.data(), vbytes(I.rep_lig));
  put(b.rep_prt, I.rep_prt.data(), vbytes(I.rep_prt));
  put(b.idx_tmp, I.idx_tmp.data(), vbytes(I.idx_tmp));
  put(b.mcc_ref, I.mcc_ref.data(), vbytes(I.mcc_ref));

  P.pt_pos = (const float4 *) (dev + b.pt_pos);
  P.pt_eh = (const float2 *) (dev + b.pt_eh);
  P.tab = (const gd_pair_param *) (dev + b.tab);
  P.lig_xyz = (const float4 *) (dev + b.lig_xyz);
  P.lig_row = (const int *) (dev + b.lig_row);
  P.lig_type = (const int *) (dev + b.lig_type);
  P.lig_q = (const float *) (dev + b.lig_q);
  P.kde = (const float4 *) (dev + b.kde);
  P.kde_pairs = (const uint32_t *) (dev + b.kde_pairs);
  P.kde_seg = (const int *) (dev + b.kde_seg);
  P.kde_inv_n = (const float *) (dev + b.kde_inv);
  P.mcs = (const float4 *) (dev + b.mcs);
  P.mcs_tcc = (const float *) (dev + b.tcc);
  P.psp_pt = (const float *) (dev + b.psp_pt);
  P.chunk_bounds = (const int *) (dev + b.chunk_bounds);
  P.row_hdb = (const int *) (dev + b.row_hdb);
  P.atom = (const gd_atom_rec *) (dev + b.atom);
  P.hpl0 = (const float *) (dev + b.hpl);
  P.hpl1 = P.hpl0 + GD_MAXTP2;
  P.hpl2 = P.hpl0 + 2 * GD_MAXTP2;
  P.inv_hpl1 = P.hpl0 + 3 * GD_MAXTP2;
  P.minus_beta = (const float *) (dev + b.mb);
  P.rep_lig = (const int *) (dev + b.rep_lig);
  P.rep_prt = (const int *) (dev + b.rep_prt);
  P.S.mv = (float *) (dev + b.mv);
  P.S.e = (float *) (dev + b.e);
  P.S.mv_best = (float *) (dev + b.mv_best);
  P.S.e_best = (float *) (dev + b.e_best);
  P.S.step_best = (int *) (dev + b.step_best);
  P.S.idx_tmp = (int *) (dev + b.idx_tmp);
  P.S.acs_mc = (int *) (dev + b.acs_mc);
  P.S.acs_x = (int *) (dev + b.acs_x);
  P.S.cmcc = (float *) (dev + b.cmcc);
  P.S.cmcc_best = (float *) (dev + b.cmcc_best);
  P.mcc_ref = I.mcc_ref.empty() ? nullptr : (const unsigned long long *) (dev + b.mcc_ref);
  P.step = (int *) (dev + b.step);
  P.rec = P.rec_stride > 0 ? (uint32_t *) (dev + b.rec) : nullptr;
  memcpy(host + b.pack, &P, sizeof P);
}

/* Tasks of one complex: runs of replicas with one protein conformer,
 * chunked by wpg. */
void gd_complex_tasks(const GdImage &I, int cx, int wpg, std::vector<gd_task> &out)
{
  const int n = I.P.n_rep;
  int i = 0;
  while (i < n) {
    int j = i;
    while (j < n && I.rep_prt[j] == I.rep_prt[i] && j - i < wpg)
      ++j;
    gd_task t;
    t.cx = cx;
    t.prt = I.rep_prt[i];
    t.r_begin = i;
    t.r_count = j - i;
    out.push_back(t);
    i = j;
  }
}

void gd_parse_state(const GdImage &I, const GdBlock &b, const char *state, const double *e64,
                    gd_mclog *mclog, gd_replica_result *results)
{
  const int n = I.P.n_rep;
  auto at = [&](size_t off) { return state + (off - b.mv); };
  const float *mv = (const float *) at(b.mv);
  const float *e = (const float *) at(b.e);
  const float *mvb = (const float *) at(b.mv_best);
  const float *eb = (const float *) at(b.e_best);
  const int *stepb = (const int *) at(b.step_best);
  const int *it = (const int *) at(b.idx_tmp);
  const int *acs = (const int *) at(b.acs_mc);
  const int *acx = (const int *) at(b.acs_x);
  const float *cm = (const float *) at(b.cmcc);
  const float *cmb = (const float *) at(b.cmcc_best);
  if (results) {
    for (int r = 0; r < n; ++r) {
      gd_replica_result &R = results[r];
      for (int k = 0; k < 6; ++k) {
        R.move_final[k] = mv[(size_t) r * 8 + k];
        R.move_best[k] = mvb[(size_t) r * 8 + k];
      }
      for (int i = 0; i < 10; ++i) {
        R.energy_final[i] = e[(size_t) r * 10 + i];
        R.energy_best[i] = eb[(size_t) r * 10 + i];
        R.energy_best_fp64[i] = e64 ? e64[(size_t) r * 10 + i] : 0.0;
      }
      R.step_best = stepb[r];
      R.idx_tmp_final = it[r];
      R.acs_mc = acs[r];
      R.acs_exchg = acx[r];
      R.cmcc_final = cm[r];
      R.cmcc_best = cmb[r];
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
}
