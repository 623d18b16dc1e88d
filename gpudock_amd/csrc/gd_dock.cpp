/*
 * gd_dock -- command-line docking run: the reference's `dock` program
 * (gpudocksm-rem-1.2/src/dock.C:14-130) over libgdremc.
 *
 *   options           ParseArguments, util.C:28-195 (same names and defaults)
 *   loading + prep    dock.C:40-90 -> gd_load_enepara/weight/norpara,
 *                     gd_load_inputs, gd_optimize_enepara, gd_set_temperature
 *   run               Run(), dock.C:103 -> gd_run
 *   summary           PrintSummary, util.C:1017-1131
 *
 * The toggles the reference compiles in (toggle.h) are options here, with
 * rem-1.2's values as defaults: IS_EXCHANGE 0 (-exchange), IS_AWAY 1
 * (-away), IS_CALCU_MCC 1 (-mcc), IS_BAYE 0 (-bayes).  IS_OUTPUT writes
 * per-step records only when -o names a prefix (the reference always
 * writes under output_<time>).  Paths the reference hard-codes in the
 * working directory (gpudocksm.ff, 08_nor_a, 08_nor_b) are options with
 * those defaults.  The random streams are keyed by -seed instead of the
 * clock, so a run is reproducible.  Errors print a message and exit with
 * status 1, like the reference's exit(EXIT_FAILURE).
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gd_remc.h"

namespace {

struct Args {
  std::string lig, prt, lhm, id = "1a07C1", weight, para = "gpudocksm.ff",
                                  nor_a = "08_nor_a", nor_b = "08_nor_b", molid = "MOLID",
                                  out, csv;
  int ns = 3000, nc = 5, nt = 1, spd = 1000, device = 0;
  float floor_temp = 0.3f, ceiling_temp = 0.3f, ts = 0.001f, rs = 3.1415f;
  unsigned seed = 1;
  int exchange = 0, away = 1, mcc = 1, bayes = 0, slot_order = 0, plate_all = 0;
};

void usage(const char *bin)
{
  std::printf(
      "usage: %s -p PROTEIN.pdb -l LIGAND.sdf -s LHM.ff [options]\n"
      "  -id ID            ligand id whose MCS lines are read (1a07C1)\n"
      "  -opt_fn FILE      term weights (default: the WEI line of -para)\n"
      "  -ns N             MC steps (3000)\n"
      "  -nc N             steps per exchange (5)\n"
      "  -nt N             temperatures (1)\n"
      "  -floor_temp T     lowest temperature (0.3)\n"
      "  -ceiling_temp T   highest temperature (0.3)\n"
      "  -t S              translational move scale, A (0.001)\n"
      "  -r S              rotational move scale, rad (3.1415)\n"
      "  -para FILE        force-field parameters (gpudocksm.ff)\n"
      "  -nor_a FILE, -nor_b FILE  normalisation (08_nor_a, 08_nor_b)\n"
      "  -molid TAG        SD property naming the ligand (MOLID)\n"
      "  -exchange 0|1     temperature exchange (0)\n"
      "  -away 0|1         start 44.5 move scales away (1)\n"
      "  -mcc 0|1          contact Matthews correlation per pose (1)\n"
      "  -bayes            Bayesian total energy\n"
      "  -slot_order       weights by energy slot, as rem-1.2's kernel applies them\n"
      "  -plate_all_models peptide plates from N, C, O in every PDB model\n"
      "  -seed N           random stream key (1)\n"
      "  -device N         GPU ordinal (0)\n"
      "  -o PREFIX         write per-step records to PREFIX_%%04d.gdtraj\n"
      "  -spd N            steps per record file (1000)\n"
      "  -csv FILE         per-replica results\n",
      bin);
}

bool parse(int argc, char **argv, Args &a)
{
  for (int i = 1; i < argc; ++i) {
    const std::string k = argv[i];
    auto val = [&](void) -> const char * {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "%s needs a value\n", k.c_str());
        std::exit(1);
      }
      return argv[++i];
    };
    if (k == "-h" || k == "--help") { usage(argv[0]); std::exit(0); }
    else if (k == "-p") a.prt = val();
    else if (k == "-l") a.lig = val();
    else if (k == "-s") a.lhm = val();
    else if (k == "-id") a.id = val();
    else if (k == "-opt_fn") a.weight = val();
    else if (k == "-ns") a.ns = std::atoi(val());
    else if (k == "-nc") a.nc = std::atoi(val());
    else if (k == "-nt") a.nt = std::atoi(val());
    else if (k == "-floor_temp") a.floor_temp = (float) std::atof(val());
    else if (k == "-ceiling_temp") a.ceiling_temp = (float) std::atof(val());
    else if (k == "-t") a.ts = (float) std::atof(val());
    else if (k == "-r") a.rs = (float) std::atof(val());
    else if (k == "-para") a.para = val();
    else if (k == "-nor_a") a.nor_a = val();
    else if (k == "-nor_b") a.nor_b = val();
    else if (k == "-molid") a.molid = val();
    else if (k == "-exchange") a.exchange = std::atoi(val());
    else if (k == "-away") a.away = std::atoi(val());
    else if (k == "-mcc") a.mcc = std::atoi(val());
    else if (k == "-bayes") a.bayes = 1;
    else if (k == "-slot_order") a.slot_order = 1;
    else if (k == "-plate_all_models") a.plate_all = 1;
    else if (k == "-seed") a.seed = (unsigned) std::strtoul(val(), nullptr, 10);
    else if (k == "-device") a.device = std::atoi(val());
    else if (k == "-o") a.out = val();
    else if (k == "-spd") a.spd = std::atoi(val());
    else if (k == "-csv") a.csv = val();
    else {
      std::fprintf(stderr, "unknown option %s\n", k.c_str());
      return false;
    }
  }
  /* util.C:152-166 */
  if (a.prt.empty()) {
    std::printf("Provide target protein structure\n");
    return false;
  }
  if (a.lig.empty()) {
    std::printf("Provide compound library in SD format\n");
    return false;
  }
  if (a.lhm.empty()) {
    std::printf("Provide LHM potentials\n");
    return false;
  }
  if (a.nt < 1) {
    std::printf("number of temperature cannot set to be zero\n");
    return false;
  }
  return true;
}

int fail(const char *what)
{
  std::fprintf(stderr, "%s: %s\n", what, gd_last_error());
  return 1;
}

const char *const kSlotNames[GD_MAXWEI] = {"vdw", "ele", "pmf", "psp", "hdb",
                                           "hpc", "kde", "lhm", "dst", "total"};

} // namespace

int main(int argc, char **argv)
{
  Args a;
  if (!parse(argc, argv, a)) {
    usage(argv[0]);
    return 1;
  }

  gd_enepara0 ep0;
  if (gd_load_enepara(a.para.c_str(), &ep0))
    return fail("parameters");
  if (!a.weight.empty() && gd_load_weight(a.weight.c_str(), &ep0))
    return fail("weights");
  if (gd_load_norpara(a.nor_a.c_str(), a.nor_b.c_str(), &ep0))
    return fail("normalisation");
  gd_enepara ep;
  if (gd_optimize_enepara(&ep0, &ep))
    return fail("parameters");

  gd_input_files f = {};
  f.ligand_sdf = a.lig.c_str();
  f.molid = a.molid.c_str();
  f.protein_pdb = a.prt.c_str();
  f.lhm_ff = a.lhm.c_str();
  f.ligand_id = a.id.c_str();
  f.flags = a.plate_all ? GD_LOAD_PLATE_ALL_MODELS : 0;
  gd_inputs *in = nullptr;
  if (gd_load_inputs(&f, &ep0, &in))
    return fail("inputs");
  gd_inputs_view v;
  gd_inputs_get(in, &v);

  gd_complex_size cs = v.size;
  cs.n_tmp = a.nt;
  cs.n_rep = cs.n_lig * cs.n_prt * cs.n_tmp;
  std::vector<gd_temp> temp(a.nt);
  if (gd_set_temperature(temp.data(), a.nt, a.floor_temp, a.ceiling_temp)) {
    gd_inputs_free(in);
    return fail("temperatures");
  }

  gd_mcpara mc = {};
  mc.steps_total = a.ns;
  mc.steps_per_dump = a.spd;
  mc.steps_per_exchange = a.nc;
  for (int k = 0; k < 3; ++k) {
    mc.move_scale[k] = a.ts;
    mc.move_scale[k + 3] = a.rs;
  }
  gd_options o = {};
  o.seed = a.seed;
  o.exchange = a.exchange;
  o.away = a.away;
  o.term_order = a.slot_order ? GD_TERM_ORDER_SLOT : GD_TERM_ORDER_FILE;
  o.device = a.device;
  o.mcc = a.mcc;
  o.combine = a.bayes ? GD_COMBINE_BAYES : GD_COMBINE_LINEAR;
  if (!a.out.empty()) {
    o.record_stride = 1;
    o.record_prefix = a.out.c_str();
  }

  std::vector<int> acs(cs.n_rep, 0), acx(cs.n_rep, 0);
  gd_mclog log = {};
  log.acs_mc = acs.data();
  log.acs_temp_exchg = acx.data();
  std::vector<gd_replica_result> res(cs.n_rep);
  const int rc = gd_run(&v.lig, &v.prt, &v.psp, &v.kde, &v.mcs, &ep, temp.data(), nullptr,
                        &mc, &log, cs, &o, res.data());
  if (rc) {
    gd_inputs_free(in);
    return fail("run");
  }

  /* PrintSummary (util.C:1017-1131) */
  const char *bar =
      "================================================================================\n";
  std::printf("\n%sInputs and Outputs\n%s", bar, bar);
  std::printf("ligand file\t\t\t%s\n", a.lig.c_str());
  std::printf("ligand name\t\t\t%s\n", v.ligand_name);
  std::printf("protein file\t\t\t%s\n", a.prt.c_str());
  std::printf("lhm file\t\t\t%s\n", a.lhm.c_str());
  std::printf("enepara file\t\t\t%s\n", a.para.c_str());
  std::printf("weight file\t\t\t%s\n", a.weight.empty() ? "(WEI line)" : a.weight.c_str());
  std::printf("out files\t\t\t%s\n", a.out.empty() ? "(none)" : (a.out + "_XXXX.gdtraj").c_str());
  std::printf("steps_per_dump\t\t\t%d\n", mc.steps_per_dump);
  std::printf("%sReplica Exchange Monte Carlo parameters\n%s", bar, bar);
  std::printf("steps_total\t\t\t%d\n", mc.steps_total);
  std::printf("steps_per_dump\t\t\t%d\n", mc.steps_per_dump);
  std::printf("steps_per_exchange\t\t%d\n", mc.steps_per_exchange);
  std::printf("translational scale\t\t");
  for (int i = 0; i < 3; ++i)
    std::printf("%.8f ", mc.move_scale[i]);
  std::printf("\nrotational scale\t\t");
  for (int i = 3; i < 6; ++i)
    std::printf("%.8f ", mc.move_scale[i]);
  std::printf("\nligand conformations\t\t%d\n", cs.n_lig);
  std::printf("prt conformations\t\t%d\n", cs.n_prt);
  std::printf("temperatures\t\t\t%d\n", cs.n_tmp);
  std::printf("replica ensembles\t\t%d\n", cs.n_rep);
  std::printf("size_lig\t\t\t%d\n", cs.lna);
  std::printf("size_prt\t\t\t%d\n", cs.pnp);
  std::printf("size_pnk\t\t\t%d\n", cs.pnk);
  std::printf("size_mcs\t\t\t%d\n", cs.pos);
  const long long moves = (long long) mc.steps_total * cs.n_rep;
  std::printf("AR of MC \t\t\t%lld / %lld \t%f\n", log.ac_mc, moves,
              (float) log.ac_mc / (float) moves);
  const long long swaps = (long long) (mc.steps_total / mc.steps_per_exchange) * cs.n_rep;
  std::printf("AR of temp exchange \t\t%lld / %lld \t%f\n", log.ac_temp_exchg, swaps,
              swaps ? (float) log.ac_temp_exchg / (float) swaps : 0.0f);
  int best = 0;
  for (int r = 1; r < cs.n_rep; ++r)
    if (res[r].energy_best[GD_MAXWEI - 1] < res[best].energy_best[GD_MAXWEI - 1])
      best = r;
  std::printf("%sBest pose\n%s", bar, bar);
  std::printf("replica\t\t\t\t%d (step %d)\n", best, res[best].step_best);
  std::printf("movematrix\t\t\t");
  for (int k = 0; k < 6; ++k)
    std::printf("%.6f ", res[best].move_best[k]);
  std::printf("\n");
  for (int k = 0; k < GD_MAXWEI; ++k)
    std::printf("%s\t\t\t\t%.6f\t(FP64 %.6f)\n", kSlotNames[k], res[best].energy_best[k],
                res[best].energy_best_fp64[k]);
  std::printf("%sPerformance\n%s", bar, bar);
  const double mcs0 = (double) moves / log.t0;
  std::printf("compute time\t\t\t%.3f seconds\n", log.t0);
  std::printf("time per MC sweep per replica\t%.3f * 1e-6 seconds\n", 1e6 / mcs0);
  std::printf("MC sweeps per second\t\t%.3f\n", mcs0);
  const double mcs1 = (double) moves / log.t1;
  std::printf("wall time\t\t\t%.3f seconds\n", log.t1);
  std::printf("MC sweeps per second\t\t%.3f\n", mcs1);
  /* no counterpart in the reference's report: the whole call, with setup,
   * upload, initial energies, FP64 re-score and read-back */
  std::printf("call time\t\t\t%.3f seconds\n", log.t2);
  std::printf("MC sweeps per second\t\t%.3f\n%s", (double) moves / log.t2, bar);

  if (!a.csv.empty()) {
    FILE *out = std::fopen(a.csv.c_str(), "w");
    if (!out) {
      std::fprintf(stderr, "cannot open %s\n", a.csv.c_str());
      gd_inputs_free(in);
      return 1;
    }
    std::fprintf(out, "replica,step_best,acs_mc,acs_exchg,idx_tmp_final,cmcc_best");
    for (int k = 0; k < 6; ++k)
      std::fprintf(out, ",move_best%d", k);
    for (int k = 0; k < GD_MAXWEI; ++k)
      std::fprintf(out, ",best_%s", kSlotNames[k]);
    for (int k = 0; k < GD_MAXWEI; ++k)
      std::fprintf(out, ",best64_%s", kSlotNames[k]);
    std::fprintf(out, "\n");
    for (int r = 0; r < cs.n_rep; ++r) {
      const gd_replica_result &q = res[r];
      std::fprintf(out, "%d,%d,%d,%d,%d,%.9g", r, q.step_best, q.acs_mc, q.acs_exchg,
                   q.idx_tmp_final, q.cmcc_best);
      for (int k = 0; k < 6; ++k)
        std::fprintf(out, ",%.9g", q.move_best[k]);
      for (int k = 0; k < GD_MAXWEI; ++k)
        std::fprintf(out, ",%.9g", q.energy_best[k]);
      for (int k = 0; k < GD_MAXWEI; ++k)
        std::fprintf(out, ",%.17g", q.energy_best_fp64[k]);
      std::fprintf(out, "\n");
    }
    std::fclose(out);
  }
  gd_inputs_free(in);
  return 0;
}
