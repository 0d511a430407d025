// nnpp_main.cpp — C++ step driver over the C ABI; the counterpart of the reference's main()
// (main.cpp:153-1028) with the same command line and the same input/output files:
//
//   nnpp_main <prefix> <jobname> <startfileNo> [--species S] [--chem reference|generic]
//             [--precision fp64|mixed] [--steps N] [--rtol R] [--device D] [--volumes]
//             [--implicit [--newton-tol T] [--lin-rtol L]] [--bunny]
//
// Reads <job>.data1/.data2/.datac*/.nodes/.bctypes/.bcvalues*/.icvalues* (main.cpp:210-289),
// builds the horizons with r = delta (1 + 1e-5) and 10 dx cells (main.cpp:304-305), sets the
// icvalues initial state (main.cpp:400-421), optionally restarts from
// <job><prefix>-<startfileNo>.output.csv (main.cpp:506-542), then marches to stepsNo writing
// <job><prefix>-<s>.output.csv, <job><prefix>.volume and <job><prefix>.energy every sub_steps
// steps (main.cpp:935-999). The time step is the explicit update of nnpp_step, or with
// --implicit the reference's own modified-Newton step (nnpp_step_implicit, main.cpp:547-1025)
// with its tolerances (EPS = 1e-5 unless overridden); the state then starts from the raw
// initial values and the first Newton iterate gets the Dirichlet values, as main.cpp:556-640.
// --bunny is the reference's BunnyConditions = 1 (main.cpp:40): the initial values come from
// the implant points <job>.bunnyNodes (bunnyNodesNo of .data1, main.cpp:236-240), classified on
// the GPU (nnpp_implant_classify, main.cpp:316-320, 422-457) instead of the icvalues files.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "nnpp.h"

namespace {

void die(const char* what, int st, nnpp_ctx* c = nullptr) {
  std::fprintf(stderr, "nnpp_main: %s failed (status %d)%s%s\n", what, st, c ? ": " : "",
               c ? nnpp_last_error(c) : "");
  std::exit(1);
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 4) {  // main.cpp:172-177 (the reference also needs the third argument, :181)
    std::printf("Usage: %s <prefix> <jobname> <startfileNo> [--species S] [--chem reference|generic]"
                " [--precision fp64|mixed] [--steps N] [--rtol R] [--device D] [--volumes]"
                " [--implicit [--newton-tol T] [--lin-rtol L]] [--bunny]\n",
                argv[0]);
    return 0;
  }
  const std::string prefix = argv[1], job = argv[2];
  const long start_no = std::atol(argv[3]);
  int S = 10, chem = -1, prec = NNPP_FP64, device = 0, use_vol = 0, implicit = 0, bunny = 0;
  long steps_override = -1;
  double rtol = 1e-10, newton_tol = -1.0, lin_rtol = -1.0;
  for (int a = 4; a < argc; a++) {
    const std::string k = argv[a];
    auto next = [&](void) -> const char* {
      if (a + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", k.c_str());
        std::exit(2);
      }
      return argv[++a];
    };
    if (k == "--species") S = std::atoi(next());
    else if (k == "--chem") chem = std::string(next()) == "generic" ? NNPP_CHEM_GENERIC : NNPP_CHEM_REFERENCE;
    else if (k == "--precision") prec = std::string(next()) == "mixed" ? NNPP_MIXED : NNPP_FP64;
    else if (k == "--steps") steps_override = std::atol(next());
    else if (k == "--rtol") rtol = std::atof(next());
    else if (k == "--device") device = std::atoi(next());
    else if (k == "--volumes") use_vol = 1;
    else if (k == "--implicit") implicit = 1;
    else if (k == "--bunny") bunny = 1;
    else if (k == "--newton-tol") newton_tol = std::atof(next());
    else if (k == "--lin-rtol") lin_rtol = std::atof(next());
    else {
      std::fprintf(stderr, "unknown option %s\n", k.c_str());
      return 2;
    }
  }
  if (chem < 0) chem = (S == 10) ? NNPP_CHEM_REFERENCE : NNPP_CHEM_GENERIC;

  nnpp_params p;
  int st = nnpp_read_params(job.c_str(), S, chem, &p);
  if (st) die("reading parameters", st);
  if (steps_override >= 0) p.n_steps = (int32_t)steps_override;
  const int32_t n = p.n_nodes;
  const int F = S + 1;
  std::vector<double> x(n), y(n), z(n), vol;
  std::printf("Reading data... ");
  if ((st = nnpp_read_nodes((job + ".nodes").c_str(), n, x.data(), y.data(), z.data())))
    die("reading .nodes", st);
  std::vector<int32_t> bct(n), phase(n);
  std::vector<double> bcv((size_t)F * n), ic((size_t)F * n), U((size_t)F * n);
  if ((st = nnpp_read_ivector((job + ".bctypes").c_str(), n, bct.data()))) die("reading .bctypes", st);
  for (int f = 0; f < F; f++) {
    const std::string sfx = std::to_string(f + 1);
    if ((st = nnpp_read_vector((job + ".bcvalues" + sfx).c_str(), n, bcv.data() + (size_t)f * n)))
      die(("reading .bcvalues" + sfx).c_str(), st);
    if ((st = nnpp_read_vector((job + ".icvalues" + sfx).c_str(), n, ic.data() + (size_t)f * n)))
      die(("reading .icvalues" + sfx).c_str(), st);
  }
  if (use_vol) {
    vol.resize(n);
    if ((st = nnpp_read_vector((job + ".volumes").c_str(), n, vol.data()))) die("reading .volumes", st);
  }
  std::printf("finished.\n");
  double energy_bound = -1e300;  // main.cpp:240: nodes.maxCoeff()
  for (int32_t i = 0; i < n; i++)
    energy_bound = std::fmax(energy_bound, std::fmax(x[i], std::fmax(y[i], z[i])));

  nnpp_ctx* c = nullptr;
  if ((st = nnpp_create(&c, device))) die("nnpp_create", st);
  if ((st = nnpp_set_params(c, &p))) die("nnpp_set_params", st, c);
  if ((st = nnpp_set_nodes(c, n, x.data(), y.data(), z.data(), use_vol ? vol.data() : nullptr)))
    die("nnpp_set_nodes", st, c);
  double t0 = now();
  int64_t nnz = 0;
  int32_t max_len = 0;
  if ((st = nnpp_build_neighbors(c, p.delta * (1.00 + 1e-5), p.dx * 10., p.dy * 10., p.dz * 10.,
                                 1024, &nnz, &max_len)))
    die("nnpp_build_neighbors", st, c);
  std::printf("Building horizons took %g seconds. Max. horizon length: %d\n", now() - t0, max_len);
  std::printf("%d nodes; horizon radius: %g\n", n, p.delta);
  if (bunny) {  // main.cpp:236-240, 316-325, 422-457
    double d1[4];
    if ((st = nnpp_read_vector((job + ".data1").c_str(), 4, d1))) die("reading .data1", st);
    const int64_t nb = (int64_t)d1[3];
    std::vector<double> bx(nb), by(nb), bz(nb);
    if ((st = nnpp_read_nodes((job + ".bunnyNodes").c_str(), nb, bx.data(), by.data(), bz.data())))
      die("reading .bunnyNodes", st);
    std::vector<int32_t> counts(n);
    t0 = now();
    st = nnpp_implant_classify(c, nb, bx.data(), by.data(), bz.data(),
                               3. * 0.4 * p.dx * (1.00 + 1e-5), 1024, counts.data(), ic.data());
    if (st) die("nnpp_implant_classify", st, c);
    int32_t mx = 0;
    double mean = 0.0;
    for (int32_t i = 0; i < n; i++) {
      mx = counts[i] > mx ? counts[i] : mx;
      mean += counts[i];
    }
    std::printf("Building bunnyNodes took %g seconds. Max. horizonsBunny length: %d \n"
                "Mean horizonBunny length: %g\n", now() - t0, mx, n ? mean / n : 0.0);
  }
  if ((st = nnpp_set_bc(c, bct.data(), bcv.data()))) die("nnpp_set_bc", st, c);
  if ((st = nnpp_init_state(c, ic.data()))) die("nnpp_init_state", st, c);
  if ((st = nnpp_set_solver(c, rtol, 20000, 1e-3))) die("nnpp_set_solver", st, c);
  if (implicit && (st = nnpp_set_state(c, NNPP_SLOT_N, ic.data())))  // raw U^n, main.cpp:400-411
    die("nnpp_set_state", st, c);
  if ((st = nnpp_get_state(c, NNPP_SLOT_N, U.data()))) die("nnpp_get_state", st, c);
  if ((st = nnpp_get_phase(c, phase.data()))) die("nnpp_get_phase", st, c);
  nnpp_newton_opts nopt;
  nnpp_newton_defaults(&nopt);
  if (newton_tol >= 0) nopt.newton_tol = newton_tol;
  if (lin_rtol >= 0) nopt.lin_rtol = lin_rtol;
  const double V0 = p.dx * p.dy * p.dz;
  int64_t solid = 0;
  double vol0rel = 0.0;
  for (int32_t i = 0; i < n; i++)  // main.cpp:472-479, from the initial values
    if (phase[i] == 1) {
      solid++;
      vol0rel += (ic[i] - p.CSat / p.CSolid);
    }
  std::printf("Initial volume in mm^3: %g\n", solid * V0 * 1e9);

  const std::string base = job + prefix;
  std::FILE* fe = std::fopen((base + ".energy").c_str(), "w");  // main.cpp:495-503
  std::FILE* fv = std::fopen((base + ".volume").c_str(), "w");
  if (fe) std::fclose(fe);
  if (fv) std::fclose(fv);

  long s0 = 1;
  if (start_no > 0) {  // restart branch (main.cpp:506-542)
    const std::string f = base + "-" + std::to_string(start_no) + ".output.csv";
    if ((st = nnpp_read_start_csv(f.c_str(), n, S, U.data(), phase.data()))) die(f.c_str(), st);
    if ((st = nnpp_set_state(c, NNPP_SLOT_N, U.data()))) die("nnpp_set_state", st, c);
    if ((st = nnpp_set_phase(c, phase.data()))) die("nnpp_set_phase", st, c);
    double v = 0.0;
    for (int32_t i = 0; i < n; i++)
      if (phase[i] == 1) v += (U[i] - p.CSat / p.CSolid);
    s0 = start_no;
    nnpp_set_time(c, s0, vol0rel != 0.0 ? v / vol0rel : 0.0, vol0rel);
    std::printf("Volume rel: %g\n", vol0rel != 0.0 ? v / vol0rel : 0.0);
  }

  std::printf("Time marching started:\n");
  const double tm0 = now();
  const int sub = p.sub_steps > 0 ? p.sub_steps : 8;
  long s = s0;
  while (s <= p.n_steps) {
    // march to the next output step (a multiple of sub_steps) or the end
    long next_out = ((s + sub - 1) / sub) * sub;
    if (next_out > p.n_steps) next_out = p.n_steps;
    const int32_t k = (int32_t)(next_out - s + 1);
    nnpp_step_stats ss;
    std::memset(&ss, 0, sizeof(ss));
    const double ts = now();
    if (implicit) {
      nnpp_newton_stats ns;
      st = nnpp_step_implicit(c, k, prec, 1, &nopt, &ns);
      if (st && st != NNPP_ERR_NOCONV) die("nnpp_step_implicit", st, c);
      if (st == NNPP_ERR_NOCONV) std::printf("warning: a linear solve hit its iteration limit\n");
      ss.steps = ns.steps;
      ss.solid_count = ns.solid_count;
      ss.voln1rel = ns.voln1rel;
      std::printf("Time for a step: %g seconds. (%d steps, %d Newton iterations, %d unconverged,"
                  " %d BiCGSTAB + %d CG iterations, norm of residual %.3e)\n",
                  (now() - ts) / (ns.steps ? ns.steps : 1), ns.steps, ns.newton_iterations,
                  ns.unconverged_steps, ns.lin_iterations, ns.cg_iterations, ns.last_resmax);
    } else {
      st = nnpp_step(c, k, prec, 1, &ss);
      if (st && st != NNPP_ERR_NOCONV) die("nnpp_step", st, c);
      if (st == NNPP_ERR_NOCONV) std::printf("warning: potential solve hit its iteration limit\n");
      std::printf("Time for a step: %g seconds. (%lld steps, %lld CG iterations, relres %.2e)\n",
                  (now() - ts) / (ss.steps ? ss.steps : 1), (long long)ss.steps,
                  (long long)ss.cg_iterations, ss.last_relres);
    }
    s += ss.steps;
    const long done = s - 1;
    const double t = done * p.dt;
    if (ss.solid_count <= 0) {  // main.cpp:924-932
      std::printf("Solid phase corroded completely at t: %g seconds.\n", t);
      break;
    }
    if (done % sub == 0) {  // outputs (main.cpp:935-999)
      if ((st = nnpp_get_state(c, NNPP_SLOT_N, U.data()))) die("nnpp_get_state", st, c);
      if ((st = nnpp_get_phase(c, phase.data()))) die("nnpp_get_phase", st, c);
      std::printf("\nt: %g seconds\n", t);
      std::printf("Volume in mm^2: %g\n", ss.solid_count * V0 * 1e9);
      std::printf("Volume rel: %g\n", ss.voln1rel);
      if ((fv = std::fopen((base + ".volume").c_str(), "a"))) {
        std::fprintf(fv, "%.17g,\n", ss.voln1rel);
        std::fclose(fv);
      }
      double energy = 0.0;
      for (int32_t i = 0; i < n; i++)
        if (std::fabs(x[i]) < energy_bound || std::fabs(y[i]) < energy_bound)
          energy += U[i] * U[i] * p.dx * p.dy * p.dz;
      std::printf("Energy: %g\n", energy);
      const long num = (start_no > 0) ? done + start_no : done;
      const std::string out = base + "-" + std::to_string(num) + ".output.csv";
      if ((st = nnpp_write_output_csv(out.c_str(), n, S, x.data(), y.data(), z.data(), U.data(),
                                      phase.data())))
        die(out.c_str(), st);
      if ((fe = std::fopen((base + ".energy").c_str(), "a"))) {
        std::fprintf(fe, "%g\n", energy);  // default stream precision, main.cpp:996
        std::fclose(fe);
      }
    }
  }
  std::printf("Time marching took %g seconds.\n", now() - tm0);
  nnpp_destroy(c);
  return 0;
}
