/*
 * gd_prep.cpp -- host preparation and parameter-file parsing.
 *
 *   gd_optimize_enepara  OptimizeEnepara   util.C:404-458
 *   gd_set_temperature   SetTemperature    util.C:490-532
 *   gd_set_replica       SetReplica        util.C:537-558
 *   gd_load_enepara      loadEnePara       load.C:1138-1324
 *   gd_load_weight       loadWeight        load.C:1389-1419
 *   gd_load_norpara      loadNorPara       load.C:1326-1387
 *
 * Differences from the reference, on purpose: the EnePara0 it fills is
 * zeroed first (the reference reads never-written heap for the HDB/ELE
 * entries a file omits, dock.C:38), a WEI line with fewer than ten values
 * leaves the rest zero (load.C:1269-1274 reads one past its array), and
 * errors are returned instead of exit().
 */
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gd_remc.h"

int gd_prep_fail(int code, const char *fmt, ...);

/* The type-name tables are the file format of gpudocksm.ff (data.C:239-587):
 * ligand SYBYL types, coarse-grained protein point names, residue letters. */
static const char *const kLigTypes[GD_MAXTP2] = {
    "Br", "C.1", "C.2", "C.3", "C.ar", "C.cat", "Cl", "F", "I", "N.1", "N.2", "N.3",
    "N.4", "N.am", "N.ar", "N.pl3", "O.2", "O.3", "O.co2", "P.3", "S.2", "S.3", "S.O",
    "S.O2"};
static const char *const kPointTypes[GD_MAXTP1] = {
    "CA", "PP", "A1", "C1", "D1", "E1", "E2", "F1", "F2", "H1", "H2", "I1", "K1", "K2",
    "L1", "M1", "M2", "N1", "P1", "Q1", "Q2", "R1", "R2", "S1", "T1", "V1", "W1", "W2",
    "Y1", "Y2"};
static const char kResOne[GD_MAXTP4 + 1] = "ARCPLIEHWTMDGFKQSNYV";

extern "C" int gd_lig_type_code(const char *name)
{
  for (int i = 0; i < GD_MAXTP2; ++i)
    if (!strcmp(name, kLigTypes[i]))
      return i;
  return -1;
}

extern "C" int gd_point_type_code(const char *name)
{
  for (int i = 0; i < GD_MAXTP1; ++i)
    if (!strcmp(name, kPointTypes[i]))
      return i;
  return -1;
}

extern "C" int gd_residue_code_one(char c)
{
  for (int i = 0; i < GD_MAXTP4; ++i)
    if (kResOne[i] == c)
      return i;
  return -1;
}

/* ------------------------------------------------------------------------ */
extern "C" int gd_optimize_enepara(const gd_enepara0 *src, gd_enepara *dst)
{
  if (!src || !dst)
    return gd_prep_fail(GD_ERR_ARG, "null enepara");
  memset(dst, 0, sizeof *dst);
  const float sqrt_2_pi = sqrtf(2.0f * 3.1415926535f);
  for (int i = 0; i < GD_MAXTP2; ++i)
    for (int j = 0; j < GD_MAXTP1; ++j) {
      const float r0 = src->vdw[j][i][0] * src->lj[2];
      dst->p1a[i][j] = 2.0f * src->vdw[j][i][1] * powf(r0, 9.0f);
      dst->p2a[i][j] = 3.0f * src->vdw[j][i][1] * powf(r0, 6.0f);
    }
  dst->lj0 = src->lj[0];
  dst->lj1 = src->lj[1];
  dst->el0 = src->el[0];
  dst->a1 = 4.0f - 3.0f * src->el[0];
  dst->b1 = 2.0f * src->el[0] - 3.0f;
  dst->el1 = src->el[1];
  for (int i = 0; i < GD_MAXTP3; ++i)
    dst->ele[i] = src->ele[i];
  for (int i = 0; i < GD_MAXTP2; ++i)
    for (int j = 0; j < GD_MAXTP1; ++j) {
      dst->pmf0[i][j] = src->pmf[j][i][0];
      dst->pmf1[i][j] = src->pmf[j][i][1];
      dst->hdb0[i][j] = src->hdb[j][i][0];
      dst->hdb1[i][j] = 1.0f / src->hdb[j][i][1];
      dst->hdb2[i][j] = -1.0f / (src->hdb[j][i][1] * sqrt_2_pi);
    }
  for (int i = 0; i < GD_MAXTP4; ++i)
    dst->hpp[i] = src->hpp[i];
  for (int i = 0; i < GD_MAXTP2; ++i) {
    dst->hpl0[i] = src->hpl[i][0];
    dst->hpl1[i] = src->hpl[i][1];
    dst->hpl2[i] = logf(1.0f / (dst->hpl1[i] * sqrt_2_pi));
  }
  dst->kde2 = -0.5f / (src->kde * src->kde);
  dst->kde3 = powf(src->kde * sqrt_2_pi, 3.0f);
  for (int i = 0; i < GD_MAXWEI; ++i) {
    dst->w[i] = src->w[i];
    dst->a_para[i] = src->a_para[i];
    dst->b_para[i] = src->b_para[i];
  }
  return GD_OK;
}

extern "C" int gd_set_temperature(gd_temp *temp, int num_temp, float floor_temp,
                                  float ceiling_temp)
{
  if (!temp || num_temp < 1)
    return gd_prep_fail(GD_ERR_ARG, "num_temp=%d", num_temp);
  if (!(floor_temp > 0.0f) || !(ceiling_temp > 0.0f))
    return gd_prep_fail(GD_ERR_ARG, "temperatures must be positive");
  if (num_temp == 1) {
    const float beta = 1.0f / (1.0f * floor_temp);
    temp[0].t = floor_temp;
    temp[0].order = 0;
    temp[0].minus_beta = 0.0f - beta;
    return GD_OK;
  }
  /* float throughout, as util.C:511-523 (std::log/std::exp of float) */
  const float ratio = expf(logf(ceiling_temp / floor_temp) / (float) (num_temp - 1));
  float a = floor_temp;
  for (int i = 0; i < num_temp; ++i) {
    const float beta = 1.0f / (1.0f * a);
    temp[i].t = a;
    temp[i].order = i;
    temp[i].minus_beta = 0.0f - beta;
    a *= ratio;
  }
  return GD_OK;
}

extern "C" int gd_set_replica(gd_replica *replica, gd_complex_size size)
{
  if (!replica || size.n_lig < 1 || size.n_prt < 1 || size.n_tmp < 1)
    return gd_prep_fail(GD_ERR_ARG, "bad replica layout");
  for (int i = 0; i < size.n_prt; ++i)
    for (int j = 0; j < size.n_tmp; ++j)
      for (int k = 0; k < size.n_lig; ++k) {
        const int f = size.n_tmp * size.n_lig * i + size.n_lig * j + k;
        replica[f].idx_rep = f;
        replica[f].idx_prt = i;
        replica[f].idx_tmp = j;
        replica[f].idx_lig = k;
      }
  return GD_OK;
}

/* ------------------------------------------------------------------------ */
/* parameter files                                                           */
/* ------------------------------------------------------------------------ */
static std::vector<std::string> split_ws(const std::string &line)
{
  std::vector<std::string> v;
  std::istringstream is(line);
  std::string w;
  while (is >> w)
    v.push_back(w);
  return v;
}

extern "C" int gd_load_enepara(const char *path, gd_enepara0 *ep)
{
  if (!path || !ep)
    return gd_prep_fail(GD_ERR_ARG, "null argument");
  std::ifstream f(path);
  if (!f.is_open())
    return gd_prep_fail(GD_ERR_IO, "cannot open %s", path);
  memset(ep, 0, sizeof *ep);
  std::string line;
  int lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    if (line.size() <= 3)
      continue;
    const std::string key = line.substr(0, 3);
    const std::vector<std::string> d = split_ws(line);
    auto need = [&](size_t n) { return d.size() >= n; };
    if (key == "VDW" || key == "PMF" || key == "HDB") {
      if (!need(5))
        return gd_prep_fail(GD_ERR_IO, "%s:%d: short %s line", path, lineno, key.c_str());
      const int pt = gd_point_type_code(d[1].c_str());
      const int lt = gd_lig_type_code(d[2].c_str());
      if (pt < 0 || lt < 0)
        return gd_prep_fail(GD_ERR_IO, "%s:%d: unknown type %s/%s", path, lineno,
                            d[1].c_str(), d[2].c_str());
      float(*tab)[GD_MAXTP2][2] = key == "VDW" ? ep->vdw : key == "PMF" ? ep->pmf : ep->hdb;
      tab[pt][lt][0] = (float) atof(d[3].c_str());
      tab[pt][lt][1] = (float) atof(d[4].c_str());
    } else if (key == "PLJ") {
      if (!need(4))
        return gd_prep_fail(GD_ERR_IO, "%s:%d: short PLJ", path, lineno);
      for (int i = 0; i < 3; ++i)
        ep->lj[i] = (float) atof(d[i + 1].c_str());
    } else if (key == "ELE") {
      if (!need(3))
        return gd_prep_fail(GD_ERR_IO, "%s:%d: short ELE", path, lineno);
      int slot;
      if (d[1].size() >= 2 && d[1][1] != 'C')
        slot = gd_point_type_code(d[1].c_str());
      else
        slot = d[1].empty() ? -1 : gd_residue_code_one(d[1][0]) + 30;
      if (slot < 0 || slot >= GD_MAXTP3)
        return gd_prep_fail(GD_ERR_IO, "%s:%d: unknown ELE class %s", path, lineno,
                            d[1].c_str());
      ep->ele[slot] = (float) atof(d[2].c_str());
    } else if (key == "PEL") {
      if (!need(3))
        return gd_prep_fail(GD_ERR_IO, "%s:%d: short PEL", path, lineno);
      ep->el[0] = (float) atof(d[1].c_str());
      ep->el[1] = (float) atof(d[2].c_str());
    } else if (key == "HPP") {
      if (!need(3) || d[1].size() != 1 || gd_residue_code_one(d[1][0]) < 0)
        return gd_prep_fail(GD_ERR_IO, "%s:%d: bad HPP", path, lineno);
      ep->hpp[gd_residue_code_one(d[1][0])] = (float) atof(d[2].c_str());
    } else if (key == "HPL") {
      if (!need(4))
        return gd_prep_fail(GD_ERR_IO, "%s:%d: short HPL", path, lineno);
      const int lt = gd_lig_type_code(d[1].c_str());
      if (lt < 0)
        return gd_prep_fail(GD_ERR_IO, "%s:%d: unknown ligand type %s", path, lineno,
                            d[1].c_str());
      ep->hpl[lt][0] = (float) atof(d[2].c_str());
      ep->hpl[lt][1] = (float) atof(d[3].c_str());
    } else if (key == "WEI") {
      for (int i = 0; i < GD_MAXWEI; ++i)
        ep->w[i] = (size_t) (i + 1) < d.size() ? (float) atof(d[i + 1].c_str()) : 0.0f;
    } else if (key == "KDE") {
      if (!need(2))
        return gd_prep_fail(GD_ERR_IO, "%s:%d: short KDE", path, lineno);
      ep->kde = (float) atof(d[1].c_str());
    }
  }
  /* hydrophobicity scale mapped onto [-1, 1] (load.C:1293-1322) */
  float lo = 1e6f, hi = -1e6f;
  for (int i = 0; i < GD_MAXTP4; ++i) {
    if (ep->hpp[i] < lo) lo = ep->hpp[i];
    if (ep->hpp[i] > hi) hi = ep->hpp[i];
  }
  for (int i = 0; i < GD_MAXTP4; ++i) {
    float v = (float) (((ep->hpp[i] - lo) / (hi - lo)) * 2.0 - 1.0);
    if (v < -1.0f) v = -1.0f;
    if (v > 1.0f) v = 1.0f;
    ep->hpp[i] = v;
  }
  return GD_OK;
}

static int load_column(const char *path, float *dst, int maxn)
{
  std::ifstream f(path);
  if (!f.is_open())
    return gd_prep_fail(GD_ERR_IO, "cannot open %s", path);
  std::string line;
  int i = 0;
  while (std::getline(f, line) && i < maxn) {
    std::istringstream is(line.substr(0, 30));
    float v = 0.0f;
    is >> v;
    dst[i++] = v;
  }
  return GD_OK;
}

extern "C" int gd_load_weight(const char *path, gd_enepara0 *ep)
{
  if (!path || !ep)
    return gd_prep_fail(GD_ERR_ARG, "null argument");
  return load_column(path, ep->w, GD_MAXWEI);
}

extern "C" int gd_load_norpara(const char *path_a, const char *path_b, gd_enepara0 *ep)
{
  if (!path_a || !path_b || !ep)
    return gd_prep_fail(GD_ERR_ARG, "null argument");
  int rc = load_column(path_a, ep->a_para, GD_MAXWEI);
  if (rc)
    return rc;
  return load_column(path_b, ep->b_para, GD_MAXWEI);
}
