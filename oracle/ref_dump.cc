// Runs the reference's own loaders and host preparation on one complex and
// writes the prepared structures to a flat binary file.  It is compiled
// together with the reference sources gpudocksm-rem-1.2/src/{load,data,util,
// quicksort}.C and rmsd.f, where they lie, by oracle/Makefile (target
// _ref/ref_dump); nothing of it ships.  The call sequence is that of
// gpudocksm-rem-1.2/src/dock.C:26-90, with every input struct
// value-initialised (the reference leaves Kde0.pnk, Psp0 and EnePara0
// uninitialised, dock.C:33-38).
//
// usage: ref_dump SDF PDB LHM_FF PARA_FF WEIGHT NOR_A NOR_B LIGAND_ID
//                 N_TMP FLOOR CEIL OUT
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <string>

#include "dock.h"
#include "load.h"
#include "size.h"
#include "util.h"

static FILE *g_out;

static void put(const char *name, char type, const void *data, uint64_t count)
{
  const size_t esz = (type == 'd') ? 8 : 4;
  std::fwrite(name, 1, std::strlen(name) + 1, g_out);
  std::fwrite(&type, 1, 1, g_out);
  std::fwrite(&count, 8, 1, g_out);
  std::fwrite(data, esz, count, g_out);
}

int main(int argc, char **argv)
{
  if (argc != 13) {
    std::fprintf(stderr, "usage: %s SDF PDB LHM_FF PARA_FF WEIGHT NOR_A NOR_B ID "
                         "N_TMP FLOOR CEIL OUT\n", argv[0]);
    return 2;
  }
  InputFiles *in = new InputFiles[1];
  in->lig_file.path = argv[1];
  in->lig_file.molid = "MOLID";
  in->prt_file.path = argv[2];
  in->lhm_file.path = argv[3];
  in->lhm_file.ligand_id = argv[8];
  in->enepara_file.path = argv[4];
  in->weight_file.path = argv[5];
  in->norpara_file.path_a = argv[6];
  in->norpara_file.path_b = argv[7];
  ExchgPara ex;
  ex.num_temp = std::atoi(argv[9]);
  ex.floor_temp = (float) std::atof(argv[10]);
  ex.ceiling_temp = (float) std::atof(argv[11]);

  Ligand0 *lig0 = new Ligand0[MAXEN2]();
  Protein0 *prt0 = new Protein0[MAXEN1]();
  Psp0 *psp0 = new Psp0();
  Kde0 *kde0 = new Kde0();
  Mcs0 *mcs0 = new Mcs0[MAXPOS]();
  EnePara0 *ep0 = new EnePara0();

  loadLigand(&in->lig_file, lig0);
  loadProtein(&in->prt_file, prt0);
  loadLHM(&in->lhm_file, psp0, kde0, mcs0);
  loadEnePara(&in->enepara_file, ep0);
  loadWeight(&in->weight_file, ep0);
  loadNorPara(&in->norpara_file, ep0);

  ComplexSize cs;
  cs.n_prt = in->prt_file.conf_total;
  cs.n_tmp = ex.num_temp;
  cs.n_lig = in->lig_file.conf_total;
  cs.n_rep = cs.n_lig * cs.n_prt * cs.n_tmp;
  cs.lna = in->lig_file.lna;
  cs.pnp = in->prt_file.pnp;
  cs.pnk = kde0->pnk;
  cs.pos = in->lhm_file.pos;

  Ligand *lig = (Ligand *) std::calloc(cs.n_lig, sizeof(Ligand));
  Protein *prt = new Protein[cs.n_prt]();
  Psp *psp = new Psp();
  Kde *kde = new Kde();
  Mcs *mcs = new Mcs[cs.pos > 0 ? cs.pos : 1]();
  EnePara *ep = new EnePara();
  Temp *temp = new Temp[cs.n_tmp]();

  OptimizeLigand(lig0, lig, cs);
  OptimizeProtein(prt0, prt, ep0, lig0, cs);
  OptimizePsp(psp0, psp, lig, prt);
  OptimizeKde(kde0, kde);
  OptimizeMcs(mcs0, mcs, cs);
  OptimizeEnepara(ep0, ep);
  InitLigCoord(lig, cs);
  SetTemperature(temp, &ex);

  g_out = std::fopen(argv[12], "wb");
  if (!g_out)
    return 3;
  std::fwrite("GDREF1\n", 1, 7, g_out);
  const int sizes[8] = {cs.n_lig, cs.n_prt, cs.n_tmp, cs.n_rep, cs.lna, cs.pnp, cs.pnk, cs.pos};
  put("size", 'i', sizes, 8);

  const int lna = cs.lna, pnp = cs.pnp;
  float *buf = new float[(size_t) cs.n_lig * lna * 3];
  for (int i = 0; i < cs.n_lig; ++i)
    for (int a = 0; a < lna; ++a) {
      buf[((size_t) i * lna + a) * 3 + 0] = lig[i].coord_orig.x[a];
      buf[((size_t) i * lna + a) * 3 + 1] = lig[i].coord_orig.y[a];
      buf[((size_t) i * lna + a) * 3 + 2] = lig[i].coord_orig.z[a];
    }
  put("lig_xyz", 'f', buf, (uint64_t) cs.n_lig * lna * 3);
  for (int a = 0; a < lna; ++a) {
    buf[a * 3 + 0] = lig[0].coord_new.x[a];
    buf[a * 3 + 1] = lig[0].coord_new.y[a];
    buf[a * 3 + 2] = lig[0].coord_new.z[a];
  }
  put("lig0_new_xyz", 'f', buf, (uint64_t) lna * 3);
  delete[] buf;
  put("lig_center", 'f', lig[0].coord_orig.center, 3);
  put("lig_t", 'i', lig[0].t, lna);
  put("lig_c", 'f', lig[0].c, lna);
  put("lig_n", 'i', lig[0].n, lna);
  float lens[MAXEN2];
  for (int i = 0; i < cs.n_lig; ++i)
    lens[i] = lig0[i].lens_rmsd;
  put("lig_lens_rmsd", 'f', lens, cs.n_lig);

  float *fx = new float[(size_t) cs.n_prt * pnp];
  int *ix = new int[(size_t) cs.n_prt * pnp];
#define PRT_F(field)                                                    \
  for (int i = 0; i < cs.n_prt; ++i)                                     \
    for (int j = 0; j < pnp; ++j)                                        \
      fx[(size_t) i * pnp + j] = prt[i].field[j];                        \
  put("prt_" #field, 'f', fx, (uint64_t) cs.n_prt * pnp);
#define PRT_I(field)                                                    \
  for (int i = 0; i < cs.n_prt; ++i)                                     \
    for (int j = 0; j < pnp; ++j)                                        \
      ix[(size_t) i * pnp + j] = prt[i].field[j];                        \
  put("prt_" #field, 'i', ix, (uint64_t) cs.n_prt * pnp);
  PRT_F(x) PRT_F(y) PRT_F(z) PRT_I(t) PRT_I(c) PRT_F(ele) PRT_I(seq3r)
  PRT_I(c0_and_d12_or_c2) PRT_F(hpp)
  /* raw loader fields of Protein0, for checking a loader */
#define PRT0_I(field)                                                   \
  for (int i = 0; i < cs.n_prt; ++i)                                     \
    for (int j = 0; j < pnp; ++j)                                        \
      ix[(size_t) i * pnp + j] = prt0[i].field[j];                       \
  put("prt0_" #field, 'i', ix, (uint64_t) cs.n_prt * pnp);
  PRT0_I(d) PRT0_I(r)
  delete[] fx;
  delete[] ix;
  put("pocket_center", 'f', prt[0].pocket_center, 3);

  float *ps = new float[(size_t) MAXTP2 * MAXPRO];
  for (int t = 0; t < MAXTP2; ++t)
    for (int r = 0; r < MAXPRO; ++r)
      ps[(size_t) t * MAXPRO + r] = psp->psp[t][r];
  put("psp", 'f', ps, (uint64_t) MAXTP2 * MAXPRO);
  delete[] ps;

  put("kde_x", 'f', kde->x, cs.pnk);
  put("kde_y", 'f', kde->y, cs.pnk);
  put("kde_z", 'f', kde->z, cs.pnk);
  put("kde_t", 'i', kde->t, cs.pnk);

  float *mb = new float[(size_t) (cs.pos > 0 ? cs.pos : 1) * MAXMCS * 3];
  float *tc = new float[cs.pos > 0 ? cs.pos : 1];
  for (int m = 0; m < cs.pos; ++m) {
    tc[m] = mcs[m].tcc;
    for (int f = 0; f < MAXMCS; ++f) {
      mb[((size_t) m * MAXMCS + f) * 3 + 0] = mcs[m].x[f];
      mb[((size_t) m * MAXMCS + f) * 3 + 1] = mcs[m].y[f];
      mb[((size_t) m * MAXMCS + f) * 3 + 2] = mcs[m].z[f];
    }
  }
  put("mcs_xyz", 'f', mb, (uint64_t) cs.pos * MAXMCS * 3);
  put("mcs_tcc", 'f', tc, cs.pos);
  delete[] mb;
  delete[] tc;

  put("ep_p1a", 'f', ep->p1a, MAXTP2 * MAXTP1);
  put("ep_p2a", 'f', ep->p2a, MAXTP2 * MAXTP1);
  put("ep_pmf0", 'f', ep->pmf0, MAXTP2 * MAXTP1);
  put("ep_pmf1", 'f', ep->pmf1, MAXTP2 * MAXTP1);
  put("ep_hdb0", 'f', ep->hdb0, MAXTP2 * MAXTP1);
  put("ep_hdb1", 'f', ep->hdb1, MAXTP2 * MAXTP1);
  put("ep_hdb2", 'f', ep->hdb2, MAXTP2 * MAXTP1);
  const float sc[8] = {ep->lj0, ep->lj1, ep->el0, ep->el1, ep->a1, ep->b1, ep->kde2, ep->kde3};
  put("ep_scalars", 'f', sc, 8);
  put("ep_ele", 'f', ep->ele, MAXTP3);
  put("ep_hpp", 'f', ep->hpp, MAXTP4);
  put("ep_hpl0", 'f', ep->hpl0, MAXTP2);
  put("ep_hpl1", 'f', ep->hpl1, MAXTP2);
  put("ep_hpl2", 'f', ep->hpl2, MAXTP2);
  put("ep_w", 'f', ep->w, MAXWEI);
  put("ep_a", 'f', ep->a_para, MAXWEI);
  put("ep_b", 'f', ep->b_para, MAXWEI);
  float tb[2 * MAXTMP];
  for (int t = 0; t < cs.n_tmp; ++t)
    tb[t] = temp[t].minus_beta;
  put("temp_minus_beta", 'f', tb, cs.n_tmp);
  std::fclose(g_out);
  std::printf("ref_dump: n_lig=%d n_prt=%d lna=%d pnp=%d pnk=%d pos=%d\n", cs.n_lig,
              cs.n_prt, cs.lna, cs.pnp, cs.pnk, cs.pos);
  return 0;
}
