/*
 * gd_load.cpp -- the input files of one complex and their host preparation.
 *
 *   SD file: native pose, atom types and charges,   loadLigand  load.C:99-419
 *            conformer ensemble + RMSD filter
 *   PDB file: coarse-grained effective points,      loadProtein load.C:595-983
 *            one conformer per MODEL/ENDMDL
 *   LHM file: KDE points, PSP table, MCS restraints loadLHM     load.C:1010-1126
 *   preparation for Run()                           OptimizeLigand util.C:199-226,
 *            CopyProteinResidue/OptimizeProtein util.C:252-329, OptimizePsp/Kde/Mcs
 *            util.C:335-402
 *
 * Everything lives in host memory owned by a gd_inputs object; gd_inputs_get
 * hands out the gd_ligand/gd_protein/gd_psp/gd_kde/gd_mcs views that Run()
 * and the engine take.  Numbers are parsed with atof/atoi on the same
 * fixed columns and tokens as the reference and rounded to float where it
 * rounds, and every float sum (centroids, side-chain centroids) is taken in
 * the reference's order, so the prepared arrays are bitwise the reference's.
 *
 * Differences, on purpose:
 *   - unopenable or malformed files (a short atom line, a missing property
 *     block, a conformer line with the wrong number of values, an unknown
 *     residue or ligand atom type, a PDB without ENDMDL) are errors, not
 *     exit() or reads of never-written memory;
 *   - the RMSD of an ensemble conformer to the native pose (kept when above
 *     0.1 A, load.C:330-412) is computed in double precision by Horn's
 *     quaternion method (B. K. P. Horn, J. Opt. Soc. Am. A 4:629, 1987)
 *     instead of the single-precision Kabsch routine rmsd.f:16; it agrees
 *     to float rounding, so only a conformer within ~1e-6 A of the 0.1 A
 *     cut could be kept differently;
 *   - capacities follow the data: no MAXPRO/MAXKDE/MAXPOS/MAXMCS limits
 *     (size.h); the 70-conformer ensemble limit MAXEN2 (load.C:205,409)
 *     is kept because it changes which conformers are read;
 *   - PSP entries of an unknown ligand type and MCS fields above lna + 1
 *     are dropped: no ligand atom can read them.
 * One reference quirk is kept by default, because it changes coordinates:
 * the peptide-plate points of the third and later PDB models (see
 * load_pdb); GD_LOAD_PLATE_ALL_MODELS turns it off.
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/gd_remc.h"

int gd_prep_fail(int code, const char *fmt, ...);
extern "C" int gd_lig_type_code(const char *name);
extern "C" int gd_point_type_code(const char *name);

namespace {

/* ligand atom type of an unknown SYBYL name (BADKDE, size.h:91) */
const int kBadType = 50;
/* ensemble conformers read from the SD file and kept (MAXEN2, size.h:31) */
const int kMaxConf = 70;
/* the conformer filter keeps RMSD > 0.1 A (load.C:388) */
const double kMinRmsd = 0.1;

/* residue names in residue-code order (getResCode, data.C:239-288); the
 * one-letter codes are "ARCPLIEHWTMDGFKQSNYV" in the same order */
const char *const kRes3[GD_MAXTP4] = {"ALA", "ARG", "CYS", "PRO", "LEU", "ILE", "GLU",
                                       "HIS", "TRP", "THR", "MET", "ASP", "GLY", "PHE",
                                       "LYS", "GLN", "SER", "ASN", "TYR", "VAL"};
const char kRes1[GD_MAXTP4 + 1] = "ARCPLIEHWTMDGFKQSNYV";

int residue_code(const std::string &name)
{
  for (int i = 0; i < GD_MAXTP4; ++i)
    if (name == kRes3[i])
      return i;
  return -1;
}

int lig_code(const std::string &name)
{
  const int c = gd_lig_type_code(name.c_str());
  return c < 0 ? kBadType : c;
}

/* Side-chain points of a residue (load.C:502-574): point "<letter>1" of
 * class 2 and, where the point table has it, "<letter>2" of class 3; GLY
 * has none.  Both lists are in the order the reference pushes them. */
struct SideChain {
  int n, type[2], cls[2];
};

SideChain side_chain(int res)
{
  SideChain s{0, {0, 0}, {0, 0}};
  for (int k = 1; k <= 2; ++k) {
    const char name[3] = {kRes1[res], (char) ('0' + k), 0};
    const int t = gd_point_type_code(name);
    if (t < 0)
      break;
    s.type[s.n] = t;
    s.cls[s.n] = k + 1;
    ++s.n;
  }
  return s;
}

/* Which atoms of a residue make up each point type (load.C:813-883):
 *   CA: the CA atom;
 *   the only side-chain point of a residue: every atom but N, CA, C, O;
 *   point 1 of a two-point side chain: every atom but N, CA, C, O, CB, CG;
 *   point 2: CB and CG. */
enum AtomSet { SET_CA, SET_SIDE, SET_DISTAL, SET_PROXIMAL, SET_PLATE };

bool atom_in(AtomSet s, const std::string &a)
{
  const bool bb = a == " N  " || a == " CA " || a == " C  " || a == " O  ";
  const bool cbg = a == " CB " || a == " CG ";
  switch (s) {
  case SET_CA: return a == " CA ";
  case SET_SIDE: return !bb;
  case SET_DISTAL: return !bb && !cbg;
  case SET_PROXIMAL: return cbg;
  default: return false;
  }
}

/* std::getline over the whole file: '\n' separated, a final '\n' ends the
 * last line rather than starting an empty one */
bool read_lines(const char *path, std::vector<std::string> &out)
{
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open())
    return false;
  out.clear();
  std::string line;
  while (std::getline(f, line))
    out.push_back(line);
  return true;
}

std::vector<std::string> tokens(const std::string &line)
{
  std::vector<std::string> v;
  std::istringstream is(line);
  std::string w;
  while (is >> w)
    v.push_back(w);
  return v;
}

bool starts(const std::string &s, const char *p)
{
  const size_t n = strlen(p);
  return s.size() >= n && s.compare(0, n, p) == 0;
}

/* substr as std::string::substr, but "" past the end instead of throwing */
std::string col(const std::string &s, size_t pos, size_t n)
{
  return pos <= s.size() ? s.substr(pos, n) : std::string();
}

float fcol(const std::string &s, size_t pos, size_t n)
{
  return (float) atof(col(s, pos, n).c_str());
}

/* Smallest sum of squared distances sum_i |R x_i + t - y_i|^2 over proper
 * rotations R and translations t: Horn's closed form, E = |x - xc|^2 +
 * |y - yc|^2 - 2 lambda_max(N), N the 4x4 symmetric matrix of the
 * cross-covariance, lambda_max by cyclic Jacobi rotations. */
double superpose_residual(const float *x, const float *y, int n)
{
  double xc[3] = {0, 0, 0}, yc[3] = {0, 0, 0};
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      xc[k] += x[3 * i + k];
      yc[k] += y[3 * i + k];
    }
  for (int k = 0; k < 3; ++k) {
    xc[k] /= n;
    yc[k] /= n;
  }
  double S[3][3] = {{0}}, gx = 0.0, gy = 0.0;
  for (int i = 0; i < n; ++i) {
    double a[3], b[3];
    for (int k = 0; k < 3; ++k) {
      a[k] = x[3 * i + k] - xc[k];
      b[k] = y[3 * i + k] - yc[k];
      gx += a[k] * a[k];
      gy += b[k] * b[k];
    }
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q)
        S[p][q] += a[p] * b[q];
  }
  const double xx = S[0][0], xy = S[0][1], xz = S[0][2], yx = S[1][0], yy = S[1][1],
               yz = S[1][2], zx = S[2][0], zy = S[2][1], zz = S[2][2];
  double N[4][4] = {{xx + yy + zz, yz - zy, zx - xz, xy - yx},
                    {yz - zy, xx - yy - zz, xy + yx, zx + xz},
                    {zx - xz, xy + yx, -xx + yy - zz, yz + zy},
                    {xy - yx, zx + xz, yz + zy, -xx - yy + zz}};
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 4; ++p)
      for (int q = p + 1; q < 4; ++q)
        off += N[p][q] * N[p][q];
    if (off < 1e-30)
      break;
    for (int p = 0; p < 4; ++p)
      for (int q = p + 1; q < 4; ++q) {
        if (N[p][q] == 0.0)
          continue;
        const double theta = (N[q][q] - N[p][p]) / (2.0 * N[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) /
                         (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 4; ++k) {
          const double a = N[k][p], b = N[k][q];
          N[k][p] = c * a - s * b;
          N[k][q] = s * a + c * b;
        }
        for (int k = 0; k < 4; ++k) {
          const double a = N[p][k], b = N[q][k];
          N[p][k] = c * a - s * b;
          N[q][k] = s * a + c * b;
        }
      }
  }
  double lmax = N[0][0];
  for (int k = 1; k < 4; ++k)
    lmax = std::max(lmax, N[k][k]);
  return std::max(0.0, gx + gy - 2.0 * lmax);
}

} // namespace

struct gd_inputs {
  gd_complex_size size{};
  std::string ligand_name;
  /* ligand: conformer-major coordinates, shared per-atom properties */
  std::vector<float> lx, ly, lz, lc, lens;
  std::vector<int> lt, ln;
  float center[3] = {0, 0, 0};
  /* protein: [n_prt * pnp] */
  std::vector<float> px, py, pz, pele, phpp;
  std::vector<int> pt, pc, pseq3r, phpc, pd, pr;
  std::vector<float> psp;
  int n_res = 1;
  std::vector<float> kx, ky, kz;
  std::vector<int> kt;
  std::vector<float> mx, my, mz, tcc;
  int n_field = 0;
};

/* ------------------------------------------------------------------------ */
/* SD file (load.C:99-419)                                                    */
/* ------------------------------------------------------------------------ */
static int load_sdf(const char *path, const char *molid, gd_inputs &in)
{
  std::vector<std::string> L;
  if (!read_lines(path, L))
    return gd_prep_fail(GD_ERR_IO, "cannot open %s", path);
  const int nl = (int) L.size();

  /* the counts line is read once a "$$$$" line closes more than ten lines
   * (load.C:118-131) */
  int lna = -1, lnb = 0;
  for (int i = 0; i < nl; ++i)
    if (L[i] == "$$$$" && i + 1 > 10) {
      lna = atoi(col(L[3], 0, 3).c_str());
      lnb = atoi(col(L[3], 3, 3).c_str());
    }
  if (lna < 0)
    return gd_prep_fail(GD_ERR_IO, "%s: no \"$$$$\" record terminator after the counts line",
                        path);
  if (lna < 1 || lna > GD_MAX_LNA)
    return gd_prep_fail(GD_ERR_IO, "%s: %d atoms, outside [1, %d]", path, lna, GD_MAX_LNA);
  if (4 + lna > nl)
    return gd_prep_fail(GD_ERR_IO, "%s: atom block shorter than %d lines", path, lna);

  /* property blocks: the value is the line after the tag line; the tests
   * run in the reference's order, first match wins (load.C:149-230) */
  std::vector<int> types;
  std::vector<float> charges;
  std::vector<std::vector<float>> raw;
  bool have_types = false, have_charges = false;
  const std::string mol = molid ? molid : "MOLID";
  static const char *const skip[] = {"SMILES_CANONICAL", "OB_MW", "OB_logP", "OB_PSA",
                                     "OB_MR", "MCT_HBD", "MCT_HBA"};
  for (int i = 4 + lna + lnb; i < nl - 1; ++i) {
    const std::string &s = L[i];
    if (s.find(mol) != std::string::npos) {
      in.ligand_name = L[i + 1];
      continue;
    }
    bool skipped = false;
    for (const char *k : skip)
      if (s.find(k) != std::string::npos) {
        skipped = true;
        break;
      }
    if (skipped)
      continue;
    if (s.find("OB_ATOM_TYPES") != std::string::npos) {
      types.clear();
      for (const std::string &w : tokens(L[i + 1]))
        types.push_back(lig_code(w));
      have_types = true;
    } else if (s.find("OB_ATOMIC_CHARGES") != std::string::npos) {
      charges.clear();
      for (const std::string &w : tokens(L[i + 1]))
        charges.push_back((float) atof(w.c_str()));
      have_charges = true;
    } else if (s.find("ENSEMBLE_COORDS") != std::string::npos) {
      /* one conformer per line until an empty line (load.C:204-229) */
      while (true) {
        ++i;
        if (i >= nl || L[i].empty() || (int) raw.size() >= kMaxConf)
          break;
        std::vector<float> v;
        for (const std::string &w : tokens(L[i]))
          v.push_back((float) atof(w.c_str()));
        if ((int) v.size() != 3 * lna)
          return gd_prep_fail(GD_ERR_IO, "%s:%d: ensemble conformer has %zu values, expected %d",
                              path, i + 1, v.size(), 3 * lna);
        raw.push_back(std::move(v));
      }
    }
  }
  if (!have_types || (int) types.size() < lna)
    return gd_prep_fail(GD_ERR_IO, "%s: OB_ATOM_TYPES missing or shorter than %d atoms", path,
                        lna);
  if (!have_charges || (int) charges.size() < lna)
    return gd_prep_fail(GD_ERR_IO, "%s: OB_ATOMIC_CHARGES missing or shorter than %d atoms",
                        path, lna);
  for (int a = 0; a < lna; ++a)
    if (types[a] == kBadType)
      return gd_prep_fail(GD_ERR_IO, "%s: atom %d has a SYBYL type outside the %d-type table",
                          path, a + 1, GD_MAXTP2);

  /* native pose from the atom block, relative to its float centroid
   * (load.C:234-279) */
  std::vector<float> ref(3 * (size_t) lna);
  float cen[3] = {0.0f, 0.0f, 0.0f};
  for (int a = 0; a < lna; ++a) {
    const std::string &s = L[4 + a];
    if (s.size() < 31)
      return gd_prep_fail(GD_ERR_IO, "%s:%d: short atom line", path, 5 + a);
    ref[3 * a + 0] = fcol(s, 0, 10);
    ref[3 * a + 1] = fcol(s, 10, 10);
    ref[3 * a + 2] = fcol(s, 20, 10);
    for (int k = 0; k < 3; ++k)
      cen[k] += ref[3 * a + k];
  }
  for (int k = 0; k < 3; ++k)
    cen[k] /= (float) lna;
  for (int a = 0; a < lna; ++a)
    for (int k = 0; k < 3; ++k)
      ref[3 * a + k] -= cen[k];

  /* conformer 0 is the native pose; each ensemble conformer is centred on
   * its own centroid and kept when its RMSD to the native pose exceeds
   * 0.1 A (load.C:330-412) */
  std::vector<std::vector<float>> conf{ref};
  std::vector<float> lens{0.0f};
  for (size_t j = 0; j < raw.size(); ++j) {
    std::vector<float> mob = raw[j];
    float mc[3] = {0.0f, 0.0f, 0.0f};
    for (int a = 0; a < lna; ++a)
      for (int k = 0; k < 3; ++k)
        mc[k] += mob[3 * a + k];
    for (int k = 0; k < 3; ++k)
      mc[k] /= (float) lna;
    for (int a = 0; a < lna; ++a)
      for (int k = 0; k < 3; ++k)
        mob[3 * a + k] -= mc[k];
    const float e = (float) superpose_residual(mob.data(), ref.data(), lna);
    const float rms = std::sqrt(e / (float) lna);
    if ((double) rms > kMinRmsd) {
      conf.push_back(std::move(mob));
      lens.push_back(rms);
    }
    if ((int) conf.size() >= kMaxConf)
      break;
  }

  const int n_lig = (int) conf.size();
  in.size.n_lig = n_lig;
  in.size.lna = lna;
  in.lx.resize((size_t) n_lig * lna);
  in.ly.resize(in.lx.size());
  in.lz.resize(in.lx.size());
  for (int i = 0; i < n_lig; ++i)
    for (int a = 0; a < lna; ++a) {
      in.lx[(size_t) i * lna + a] = conf[i][3 * a + 0];
      in.ly[(size_t) i * lna + a] = conf[i][3 * a + 1];
      in.lz[(size_t) i * lna + a] = conf[i][3 * a + 2];
    }
  in.lens = lens;
  in.lt.assign(types.begin(), types.begin() + lna);
  in.lc.assign(charges.begin(), charges.begin() + lna);
  in.ln.resize(lna);
  for (int a = 0; a < lna; ++a)
    in.ln[a] = a;
  /* coord_orig.center = pocket_center = the native centroid (util.C:218-221) */
  for (int k = 0; k < 3; ++k)
    in.center[k] = cen[k];
  return GD_OK;
}

/* ------------------------------------------------------------------------ */
/* PDB file (load.C:595-983) and OptimizeProtein (util.C:252-329)             */
/* ------------------------------------------------------------------------ */
static bool is_atom(const std::string &s) { return s.size() > 53 && starts(s, "ATOM  "); }
/* ENDMDL is recognised only on lines of at most 53 characters (load.C:733,913) */
static bool is_endmdl(const std::string &s) { return s.size() <= 53 && starts(s, "ENDMDL"); }

static int load_pdb(const char *path, const gd_enepara0 *ep0, int flags, gd_inputs &in)
{
  std::vector<std::string> L;
  if (!read_lines(path, L))
    return gd_prep_fail(GD_ERR_IO, "cannot open %s", path);

  /* effective points from the CA atoms of the first model (load.C:626-737):
   * per residue its CA point, then the plate point of the peptide bond to
   * the previous residue (from the third residue on), then its side chain */
  struct Pt {
    int r, t, d, c;
    AtomSet set;
  };
  std::vector<Pt> pts;
  std::vector<int> seq3;
  for (size_t i = 0; i < L.size(); ++i) {
    const std::string &s = L[i];
    if (is_atom(s)) {
      const int res = residue_code(col(s, 17, 3));
      if (res < 0)
        return gd_prep_fail(GD_ERR_IO, "%s:%zu: unknown residue %s", path, i + 1,
                            col(s, 17, 3).c_str());
      if (col(s, 12, 4) != " CA ")
        continue;
      const int pnr = (int) seq3.size();
      seq3.push_back(atoi(col(s, 22, 4).c_str()));
      pts.push_back({pnr, 0, res, 0, SET_CA});
      if (pnr > 1)
        pts.push_back({pnr - 1, 1, res, 1, SET_PLATE});
      const SideChain sc = side_chain(res);
      for (int k = 0; k < sc.n; ++k)
        pts.push_back({pnr, sc.type[k], res, sc.cls[k],
                       sc.n == 1 ? SET_SIDE : (k == 0 ? SET_DISTAL : SET_PROXIMAL)});
    } else if (is_endmdl(s)) {
      break;
    }
  }
  const int pnp = (int) pts.size(), pnr = (int) seq3.size();
  if (pnp == 0)
    return gd_prep_fail(GD_ERR_IO, "%s: no CA atom before the first ENDMDL", path);

  /* Coordinates (load.C:743-934): each point is the float mean of its atoms
   * over the lines of one model, summed in line order, and a model ends at
   * its ENDMDL line.  Atoms are matched by residue serial number alone.
   * Lines are walked once; each adds to the points it belongs to, which
   * keeps every point's summation order. */
  std::unordered_map<int, std::vector<int>> own, plate_n, plate_co, plate_co_late;
  for (int p = 0; p < pnp; ++p) {
    const Pt &q = pts[p];
    if (q.set != SET_PLATE) {
      own[seq3[q.r]].push_back(p);
    } else if (q.r > 0 && q.r < pnr - 1) {
      plate_n[seq3[q.r]].push_back(p);
      plate_co[seq3[q.r - 1]].push_back(p);
      plate_co_late[0].push_back(p);
    }
  }
  /* From the third model on, the reference looks the previous residue's
   * serial number up in a conformer whose residue tables are not filled
   * yet (load.C:886 reads myprt, which load.C:906 has moved to
   * prt[model - 1]); the zeroed tables give serial 0, so a plate point
   * averages the N atom of its residue and the C and O atoms of residue 0.
   * GD_LOAD_PLATE_ALL_MODELS takes the previous residue in every model. */
  const bool all_models = (flags & GD_LOAD_PLATE_ALL_MODELS) != 0;
  int model = 0;
  std::vector<float> sx(pnp, 0.0f), sy(pnp, 0.0f), sz(pnp, 0.0f), sn(pnp, 0.0f);
  std::vector<std::vector<float>> mx, my, mz;
  auto add = [&](int p, float x, float y, float z) {
    sx[p] += x;
    sy[p] += y;
    sz[p] += z;
    sn[p] += 1.0f;
  };
  for (const std::string &s : L) {
    if (is_atom(s)) {
      const int serial = atoi(col(s, 22, 4).c_str());
      const std::string atom = col(s, 12, 4);
      const float x = fcol(s, 30, 8), y = fcol(s, 38, 8), z = fcol(s, 46, 8);
      auto it = own.find(serial);
      if (it != own.end())
        for (int p : it->second)
          if (atom_in(pts[p].set, atom))
            add(p, x, y, z);
      if (atom == " N  ") {
        auto jt = plate_n.find(serial);
        if (jt != plate_n.end())
          for (int p : jt->second)
            add(p, x, y, z);
      } else if (atom == " C  " || atom == " O  ") {
        const auto &co = (model < 2 || all_models) ? plate_co : plate_co_late;
        auto jt = co.find(serial);
        if (jt != co.end())
          for (int p : jt->second)
            add(p, x, y, z);
      }
    } else if (is_endmdl(s)) {
      /* a point with no atoms in this model keeps the zero of a fresh
       * conformer (the reference leaves whatever the array held) */
      std::vector<float> cx(pnp, 0.0f), cy(pnp, 0.0f), cz(pnp, 0.0f);
      for (int p = 0; p < pnp; ++p) {
        if (sn[p] > 0.0f) {
          cx[p] = sx[p] / sn[p];
          cy[p] = sy[p] / sn[p];
          cz[p] = sz[p] / sn[p];
        }
        sx[p] = sy[p] = sz[p] = sn[p] = 0.0f;
      }
      mx.push_back(std::move(cx));
      my.push_back(std::move(cy));
      mz.push_back(std::move(cz));
      ++model;
    }
  }
  const int n_prt = (int) mx.size();
  if (n_prt == 0)
    return gd_prep_fail(GD_ERR_IO, "%s: no ENDMDL record (coordinates are taken per model)",
                        path);

  /* CopyProteinResidue for every conformer, identity order (util.C:252-329) */
  in.size.n_prt = n_prt;
  in.size.pnp = pnp;
  const size_t np = (size_t) n_prt * pnp;
  in.px.resize(np);
  in.py.resize(np);
  in.pz.resize(np);
  in.pele.resize(np);
  in.phpp.resize(np);
  in.pt.resize(np);
  in.pc.resize(np);
  in.pseq3r.resize(np);
  in.phpc.resize(np);
  in.pd.resize(np);
  in.pr.resize(np);
  for (int m = 0; m < n_prt; ++m)
    for (int p = 0; p < pnp; ++p) {
      const size_t i = (size_t) m * pnp + p;
      const Pt &q = pts[p];
      in.px[i] = mx[m][p];
      in.py[i] = my[m][p];
      in.pz[i] = mz[m][p];
      in.pt[i] = q.t;
      in.pc[i] = q.c;
      in.pele[i] = ep0->ele[q.t == 0 ? q.d + 30 : q.t];
      in.pseq3r[i] = seq3[q.r];
      in.phpc[i] = (q.c == 0 && q.d == 12) || q.c == 2;
      in.phpp[i] = ep0->hpp[q.d];
      in.pd[i] = q.d;
      in.pr[i] = q.r;
    }
  return GD_OK;
}

/* ------------------------------------------------------------------------ */
/* LHM file (load.C:1010-1126), OptimizePsp/Kde/Mcs (util.C:335-402)          */
/* ------------------------------------------------------------------------ */
static int load_lhm(const char *path, const char *ligand_id, gd_inputs &in)
{
  std::vector<std::string> L;
  if (!read_lines(path, L))
    return gd_prep_fail(GD_ERR_IO, "cannot open %s", path);
  const std::string id = ligand_id ? ligand_id : "";
  const int lna = in.size.lna;
  /* the energy reads MCS field n + 1 of ligand atom n < lna */
  in.n_field = lna + 2;
  std::map<std::pair<int, int>, float> psp;   /* (type, residue) -> value, last wins */
  int max_res = -1;
  for (size_t i = 0; i < L.size(); ++i) {
    const std::string &s = L[i];
    if (s.size() <= 3)
      continue;
    const std::string key = s.substr(0, 3);
    if (key == "KDE") {
      const std::vector<std::string> d = tokens(s);
      if (d.size() < 5)
        return gd_prep_fail(GD_ERR_IO, "%s:%zu: short KDE line", path, i + 1);
      in.kt.push_back(lig_code(d[1]));
      in.kx.push_back((float) atof(d[2].c_str()));
      in.ky.push_back((float) atof(d[3].c_str()));
      in.kz.push_back((float) atof(d[4].c_str()));
    } else if (key == "PSP") {
      const std::vector<std::string> d = tokens(s);
      if (d.size() < 4)
        return gd_prep_fail(GD_ERR_IO, "%s:%zu: short PSP line", path, i + 1);
      const int r = atoi(d[1].c_str());
      if (r < 0 || r >= (1 << 23))
        return gd_prep_fail(GD_ERR_IO, "%s:%zu: PSP residue %d out of range", path, i + 1, r);
      const int t = lig_code(d[2]);
      if (t == kBadType)
        continue;
      psp[{t, r}] = (float) atof(d[3].c_str());
      max_res = std::max(max_res, r);
    } else if (key == "MCS") {
      /* every MCS line is one restraint slot; only those of ligand_id get
       * a weight and fields (load.C:1091-1118) */
      const std::vector<std::string> d = tokens(s);
      const size_t nf = (size_t) in.n_field;
      in.tcc.push_back(0.0f);
      in.mx.resize(in.mx.size() + nf, GD_MCS_INVALID_COORD);
      in.my.resize(in.my.size() + nf, GD_MCS_INVALID_COORD);
      in.mz.resize(in.mz.size() + nf, GD_MCS_INVALID_COORD);
      if (d.size() < 2 || d[1] != id)
        continue;
      if (d.size() < 4)
        return gd_prep_fail(GD_ERR_IO, "%s:%zu: short MCS line", path, i + 1);
      const size_t row = in.tcc.size() - 1;
      in.tcc[row] = (float) atof(d[2].c_str());
      const int total = atoi(d[3].c_str());
      if (total < 0 || d.size() < 4 + 4 * (size_t) total)
        return gd_prep_fail(GD_ERR_IO, "%s:%zu: MCS line shorter than its %d fields", path,
                            i + 1, total);
      for (int a = 0; a < total; ++a) {
        const int f = atoi(d[4 + 4 * a].c_str());
        if (f < 0)
          return gd_prep_fail(GD_ERR_IO, "%s:%zu: negative MCS field %d", path, i + 1, f);
        if (f >= in.n_field)
          continue;
        in.mx[row * nf + f] = (float) atof(d[5 + 4 * a].c_str());
        in.my[row * nf + f] = (float) atof(d[6 + 4 * a].c_str());
        in.mz[row * nf + f] = (float) atof(d[7 + 4 * a].c_str());
      }
    }
  }
  in.size.pnk = (int) in.kt.size();
  in.size.pos = (int) in.tcc.size();

  /* PSP as [ligand type][residue], wide enough for every residue a class-2
   * point indexes (OptimizePsp, util.C:335-342) */
  const size_t np = in.pc.size();
  for (size_t i = 0; i < np; ++i)
    if (in.pc[i] == 2)
      max_res = std::max(max_res, in.pseq3r[i]);
  in.n_res = max_res + 1 > 0 ? max_res + 1 : 1;
  in.psp.assign((size_t) GD_MAXTP2 * in.n_res, 0.0f);
  for (const auto &kv : psp)
    in.psp[(size_t) kv.first.first * in.n_res + kv.first.second] = kv.second;
  return GD_OK;
}

/* ------------------------------------------------------------------------ */
extern "C" int gd_load_inputs(const gd_input_files *files, const gd_enepara0 *enepara0,
                              gd_inputs **out)
{
  if (!out)
    return gd_prep_fail(GD_ERR_ARG, "null output");
  *out = nullptr;
  if (!files || !enepara0 || !files->ligand_sdf || !files->protein_pdb || !files->lhm_ff)
    return gd_prep_fail(GD_ERR_ARG, "null input file path or parameters");
  gd_inputs *in = new gd_inputs();
  int rc = load_sdf(files->ligand_sdf, files->molid, *in);
  if (!rc)
    rc = load_pdb(files->protein_pdb, enepara0, files->flags, *in);
  if (!rc)
    rc = load_lhm(files->lhm_ff, files->ligand_id, *in);
  if (rc) {
    delete in;
    return rc;
  }
  *out = in;
  return GD_OK;
}

extern "C" int gd_inputs_get(const gd_inputs *in, gd_inputs_view *v)
{
  if (!in || !v)
    return gd_prep_fail(GD_ERR_ARG, "null argument");
  memset(v, 0, sizeof *v);
  v->size = in->size;
  v->lig.x = in->lx.data();
  v->lig.y = in->ly.data();
  v->lig.z = in->lz.data();
  v->lig.t = in->lt.data();
  v->lig.c = in->lc.data();
  v->lig.n = in->ln.data();
  for (int k = 0; k < 3; ++k) {
    v->lig.center[k] = in->center[k];
    v->prt.pocket_center[k] = in->center[k];
  }
  v->prt.x = in->px.data();
  v->prt.y = in->py.data();
  v->prt.z = in->pz.data();
  v->prt.t = in->pt.data();
  v->prt.c = in->pc.data();
  v->prt.ele = in->pele.data();
  v->prt.seq3r = in->pseq3r.data();
  v->prt.c0_and_d12_or_c2 = in->phpc.data();
  v->prt.hpp = in->phpp.data();
  v->psp.psp = in->psp.data();
  v->psp.n_res = in->n_res;
  v->kde.x = in->kx.data();
  v->kde.y = in->ky.data();
  v->kde.z = in->kz.data();
  v->kde.t = in->kt.data();
  v->mcs.x = in->mx.data();
  v->mcs.y = in->my.data();
  v->mcs.z = in->mz.data();
  v->mcs.tcc = in->tcc.data();
  v->mcs.n_field = in->n_field;
  v->lens_rmsd = in->lens.data();
  v->prt_d = in->pd.data();
  v->prt_r = in->pr.data();
  v->ligand_name = in->ligand_name.c_str();
  return GD_OK;
}

extern "C" void gd_inputs_free(gd_inputs *in) { delete in; }
