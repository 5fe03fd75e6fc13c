// prc_d2z -- host driver of the delta2zreion path on one MI355X.
//
// Keeps the reference d2z program's interface (marcelo-alvarez/pr-calc,
// c++/src/main_d2z.C:15-109): the same parameter file (io_d2z.C:16-162), the
// same command line (commandline_d2z.C:6-127), a dense N^3 float32 ICs file in
// (io_d2z.C:164-207) and a dense N^3 float32 .zreion file out (io_d2z.C:209-246),
// the per-radius table of delta2zreion_d2z.C:46-52,102-103 and the optical
// depth of GetHistory/tau (bubblehistory_d2z.C:5-41, cosmology.C:52-83).  The
// FFTW-MPI transforms and delta2zreion() are replaced by calls into
// libprc_d2z (include/prc_d2z.h); there is no MPI: one process drives one GPU.
//
//   prc_d2z <paramfile> [-f name] [-r Rmax] [-m Mmin] [-z zeta] [-l] [-v] [-e]
//                       [-M plain|fused|mixed|half] [-b radii_per_launch] [-d device]
//                       [-T fcoll table]
//
// Differences from the reference, on purpose:
//   * every failure prints a message and exits with status 1 (the reference
//     prints and continues, or exits with status 0);
//   * -l (lean mode) is accepted and ignored: the device holds the spectrum and
//     the intermediates, so the extra forward transforms buy nothing;
//   * the history correction of main_d2z.C:78-90 (GetBubbleMFP, WriteMFP,
//     WriteMFPPDF, GetCHistory, CorrectHistory) runs as the reference's does and
//     the .zreion written is the corrected field; the skewers draw Philox
//     numbers (seed 1) instead of rand().  fcoll_table.tab (or -T <file>) comes
//     from the colossus-based table builder; without it the program says so and
//     writes the uncorrected field (the reference crashes on the missing file).
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <string>
#include <vector>

#include "prc_d2z.h"

namespace {

constexpr int kNHistory = PRC_D2Z_NHISTORY;
constexpr double kHistoryInitial = 49.9;  // HISTORY_INITIAL, allvars_d2z.h:2
constexpr double kHistoryFinal = 0.01;    // HISTORY_FINAL, allvars_d2z.h:3

struct Options {
  std::string paramfile, basename = "output/output";
  float Rmax = 256, zeta = 10, M_min = 1e8;  // commandline_d2z.C:47-49
  int verbose = 0, everbose = 0, lean = 0;
  int mode = PRC_D2Z_FUSED_FP32, radii_per_launch = 0, device = 0;
  std::string fcoll_table = "fcoll_table.tab";  // Float2FloatTable file of GetCHistory
};

struct Param {
  float Omegam = 0, Omegab = 0, Omegal = 0, h = 0, ns = 0, sigma8 = 0, w = 0;
  float BoxSize = 0, zInit = 0, Rmin = 0;
  int N = 0, Nscales = 0;
  std::string DeltaFile;
  int RecordBubbleSizes = 0;
};

[[noreturn]] void die(const char* fmt, const char* a = "", const char* b = "") {
  fprintf(stderr, "prc_d2z: ");
  fprintf(stderr, fmt, a, b);
  fprintf(stderr, "\n");
  exit(1);
}

void usage() {
  printf("\n usage: prc_d2z <parameterfile> [options]\n");
  printf("\n OPTIONS:\n");
  printf("\n   -h show this message");
  printf("\n   -f <File base name> [default = 'output'; output files are output/<name>.*]");
  printf("\n   -r <Rmfp> [default = 256; the mean free path in units of Mpc/h]");
  printf("\n   -m <Mmin> [default = 1e8; the minimum halo mass hosting ionizing sources, in Msun]");
  printf("\n   -z <zeta> [default =  10; the ionizing efficiency of halos]");
  printf("\n   -l lean mode (accepted; no effect on the GPU)");
  printf("\n   -v print stdout to the terminal instead of the .stdout file");
  printf("\n   -e print stderr to the terminal instead of the .stderr file");
  printf("\n   -M plain|fused|mixed|half  GPU execution path [default = fused]");
  printf("\n   -b <n> radii per fused batch, 1..32 [default = 32]");
  printf("\n   -d <n> HIP device [default = 0]");
  printf("\n   -T <file> collapsed-fraction table of the history correction [default = fcoll_table.tab]");
  printf("\n\n");
  exit(0);
}

// getopt "hvelf:r:m:z:" (commandline_d2z.C:54-99) plus the execution options.
Options command_line(int argc, char** argv) {
  Options o;
  if (argc < 2) usage();
  o.paramfile = argv[1];
  opterr = 0;
  int c;
  while ((c = getopt(argc - 1, &argv[1], "hvelf:r:m:z:M:b:d:T:")) != -1) {
    switch (c) {
      case 'h': usage();
      case 'f': o.basename = std::string("output/") + optarg; break;
      case 'r': o.Rmax = (float)atof(optarg); break;
      case 'm': o.M_min = (float)atof(optarg); break;
      case 'z': o.zeta = (float)atof(optarg); break;
      case 'v': o.verbose = 1; break;
      case 'e': o.everbose = 1; break;
      case 'l': o.lean = 1; break;
      case 'M':
        if (!strcmp(optarg, "plain")) o.mode = PRC_D2Z_PLAIN_FP32;
        else if (!strcmp(optarg, "fused")) o.mode = PRC_D2Z_FUSED_FP32;
        else if (!strcmp(optarg, "mixed")) o.mode = PRC_D2Z_FUSED_MIXED;
        else if (!strcmp(optarg, "half")) o.mode = PRC_D2Z_FUSED_HALF;
        else die("unknown -M value '%s'", optarg);
        break;
      case 'b': o.radii_per_launch = atoi(optarg); break;
      case 'd': o.device = atoi(optarg); break;
      case 'T': o.fcoll_table = optarg; break;
      case '?':
        fprintf(stderr, "\n Unknown option or missing argument `-%c'.\n", optopt);
        usage();
      default: usage();
    }
  }
  return o;
}

// Tag/value parameter file (io_d2z.C:16-162): lines with at least two words,
// '%' or '#' comments, each of the 14 tags at most once; writes
// <paramfile>-usedvalues.  Unknown or repeated tags are fatal here.
Param read_parameter_file(const std::string& fname) {
  Param p;
  enum Kind { F, I, S };
  struct Tag {
    const char* name;
    Kind kind;
    void* addr;
    bool seen;
  };
  std::vector<Tag> tags = {
      {"Omegam", F, &p.Omegam, false}, {"Omegab", F, &p.Omegab, false}, {"Omegal", F, &p.Omegal, false},
      {"h", F, &p.h, false},           {"ns", F, &p.ns, false},         {"sigma8", F, &p.sigma8, false},
      {"w", F, &p.w, false},           {"BoxSize", F, &p.BoxSize, false}, {"zInit", F, &p.zInit, false},
      {"Rmin", F, &p.Rmin, false},     {"N", I, &p.N, false},           {"Nscales", I, &p.Nscales, false},
      {"DeltaFile", S, &p.DeltaFile, false}, {"RecordBubbleSizes", I, &p.RecordBubbleSizes, false},
  };
  FILE* fd = fopen(fname.c_str(), "r");
  if (!fd) die("Parameter file %s not found.", fname.c_str());
  FILE* fout = fopen((fname + "-usedvalues").c_str(), "w");
  char buf[400], b1[200], b2[200], b3[400];
  int nerr = 0;
  while (fgets(buf, sizeof buf, fd)) {
    if (sscanf(buf, "%199s%199s%399s", b1, b2, b3) < 2) continue;
    if (b1[0] == '%' || b1[0] == '#') continue;
    Tag* t = nullptr;
    for (auto& x : tags)
      if (!strcmp(b1, x.name) && !x.seen) t = &x;
    if (!t) {
      printf("Error in %s: Tag '%s' not allowed or multiple defined.\n", fname.c_str(), b1);
      ++nerr;
      continue;
    }
    t->seen = true;
    switch (t->kind) {
      case F:
        *(float*)t->addr = (float)atof(b2);
        if (fout) fprintf(fout, "%-35s%f\n", b1, *(float*)t->addr);
        break;
      case I:
        *(int*)t->addr = atoi(b2);
        if (fout) fprintf(fout, "%-35s%d\n", b1, *(int*)t->addr);
        break;
      case S:
        *(std::string*)t->addr = b2;
        if (fout) fprintf(fout, "%-35s%s\n", b1, b2);
        break;
    }
  }
  fclose(fd);
  if (fout) fclose(fout);
  if (nerr) die("%s: invalid parameter file", fname.c_str());
  for (auto& x : tags)
    if (!x.seen && strcmp(x.name, "RecordBubbleSizes") != 0) die("%s: missing tag %s", fname.c_str(), x.name);
  return p;
}

// Dense N^3 float32 (x slowest, no header); logs mean and sigma in double
// (io_d2z.C:164-207).
std::vector<float> read_delta_file(const std::string& fname, int N) {
  const size_t n = (size_t)N * N * N;
  std::vector<float> d(n);
  FILE* f = fopen(fname.c_str(), "rb");
  if (!f) die("cannot open density file %s", fname.c_str());
  const size_t got = fread(d.data(), sizeof(float), n, f);
  fclose(f);
  if (got != n) die("density file %s is shorter than N^3 floats", fname.c_str());
  double mean = 0, sigma = 0;
  for (size_t i = 0; i < n; ++i) {
    mean += d[i];
    sigma += d[i] * d[i];
  }
  mean /= pow((float)N, 3);
  sigma = sqrt(sigma / pow((float)N, 3));
  printf("\n Input data read, mean = %le sigma = %le", mean, sigma);
  return d;
}

// tau of cosmology.C:52-83 with the reference's float accumulator
float tau_es(const std::vector<float>& za, const std::vector<float>& xa, float omegam, float omegab, float h) {
  const float thompson = 6.65e-25f, YHe = 0.25f;
  const float hubble0 = (float)(100. / 3.086e19 * h);
  const float ne0 = (float)((double)(omegab * h * h) * 1.88e-29 / 1.67e-24);
  const float c = 3e10f;
  const float dtau0 = thompson * ne0 * c / hubble0;
  float tau = 0;
  const int nz = (int)za.size();
  for (int i = 1; i < nz - 1; ++i) {
    const float z = za[i], x = xa[i];
    const float dz = fabsf(za[i + 1] - za[i]);
    const float NHe = (z < 3) ? 2.f : 1.f;
    const double dtau = (double)(x * dtau0) * pow((double)(1 + z), 2) /
                        sqrt(omegam * pow((double)(1 + z), 3) + 1 - omegam) *
                        (double)(1 - YHe + NHe * YHe / 4) * (double)dz;
    tau = (float)((double)tau + dtau);
  }
  return tau;
}

void check(int rc, prc_d2z_ctx* ctx, const char* what) {
  if (rc != PRC_D2Z_OK) die("%s failed: %s", what, prc_d2z_last_error(ctx));
}

// GetHistory (bubblehistory_d2z.C:5-41): integer counts per bin from the GPU,
// history[i] = count / N^3 in float, then the float running sum;
// history_z[i] = 49.9 + i dzhistory.  The reference never sets history_z[0]
// (its loop starts at 1); here it is 49.9.  Returns tau_es of the history.
float get_history(prc_d2z_ctx* ctx, const Param& P, std::vector<float>& history, std::vector<float>& history_z) {
  std::vector<uint32_t> counts(kNHistory);
  check(prc_d2z_history(ctx, counts.data()), ctx, "history");
  const float dzhistory = (float)((kHistoryFinal - kHistoryInitial) / (kNHistory - 1));
  history.assign(kNHistory, 0.0f);
  history_z.assign(kNHistory, 0.0f);
  const double n3 = pow((double)P.N, 3);
  for (int i = 0; i < kNHistory; ++i) history[i] = (float)((double)counts[i] / n3);
  for (int i = 1; i < kNHistory; ++i) history[i] += history[i - 1];
  history_z[0] = (float)kHistoryInitial;
  for (int i = 1; i < kNHistory; ++i) history_z[i] = (float)(kHistoryInitial + i * dzhistory);
  return tau_es(history_z, history, P.Omegam, P.Omegab, P.h);
}

// growth (cosmology.C:174-212) with the reference's float/double mixture:
// float variables, the 2.5 and 4./7. and Ma et al. coefficients in double
float growth_ref(float z, float omegam, float omegal, float w) {
  const float x = 1 + z, x2 = x * x, x3 = x * x * x;
  const float x3w = powf(x3, w);
  const float omega = omegam * x3 / (omegam * x3 + (1 - omegam - omegal) * x2 + omegal);
  const float lambda = omegal * x3 * x3w / (omegam * x3 + (1 - omegam - omegal) * x2 + omegal * x3 * x3w);
  float g = (float)(2.5 * omega / (pow((double)omega, 4. / 7.) - lambda + (1 + omega / 2) * (1 + lambda / 70)));
  float g0 = (float)(2.5 * omegam / (pow((double)omegam, 4. / 7.) - omegal + (1 + omegam / 2) * (1 + omegal / 70)));
  const float t = (float)(-(0.255 + 0.305 * w + 0.0027 / w) * (1 - omega) -
                          (0.366 + 0.266 * w - 0.0700 / w) * logf(omega));
  const float t0 = (float)(-(0.255 + 0.305 * w + 0.0027 / w) * (1 - omegam) -
                           (0.366 + 0.266 * w - 0.0700 / w) * logf(omegam));
  g *= powf(-w, t);
  g0 *= powf(-w, t0);
  return (g / x) / g0;
}

// deltac (cosmology.C:141-172): the collapse threshold over D(z); the open
// (Omega_Lambda = 0) and the flat fitting forms, float variables
float deltac_ref(float z, float omegam, float omegal, float w) {
  const float pi = 3.14159265358979f;
  const float x = 1 + z, x2 = x * x, x3 = x * x * x;
  const float omega = omegam * x3 / (omegam * x3 + (1 - omegam - omegal) * x2 + omegal);
  const float eta = (float)log((2 / omega - 1) + sqrt(pow((double)(2 / omega - 1), 2) - 1));
  float dc;
  if ((omegam < 1) && (omegal == 0)) {
    dc = (float)(1.5 * (3 * sinhf(eta) * (sinhf(eta) - eta) / pow((double)(coshf(eta) - 1), 2) - 2) *
                 (1 + pow((double)(2 * pi / (sinhf(eta) - eta)), 2. / 3.)));
  } else {
    dc = (float)(3 * pow((double)(12 * pi), 2. / 3.) / 20 * (1 + 0.0123 * log10f(omega)));
  }
  dc /= growth_ref(z, omegam, omegal, w);
  return dc;
}

// fcoll (cosmology.C:311-324): erfc(deltac / sqrt(2) / sigma_min), float erfc
float fcoll_ref(float z, float omegam, float omegal, float w, float sigma_min) {
  const float arg = (float)((double)deltac_ref(z, omegam, omegal, w) / sqrt(2.) / (double)sigma_min);
  return erfcf(arg);
}

// Float2FloatTable::Float2Float (tables.C:27-38)
float table_value(const std::vector<float>& t, float tmin, float tmax, float v) {
  const int n = (int)t.size();
  if (v <= tmin) return t[0];
  if (v >= tmax) return t[n - 1];
  const float delta = (tmax - tmin) / (float)n;
  const int bin = (int)((v - tmin) / delta);
  if (bin >= n - 1) return t[n - 1];  // the reference reads t[n] here
  const float f = (v - (tmin + (float)bin * delta)) / delta;
  return (1 - f) * t[bin] + f * t[bin + 1];
}

}  // namespace

int main(int argc, char** argv) {
  Options o = command_line(argc, argv);
  mkdir("output", 0755);
  if (!o.verbose && !freopen((o.basename + ".stdout").c_str(), "w", stdout)) die("cannot open .stdout");
  if (!o.everbose && !freopen((o.basename + ".stderr").c_str(), "w", stderr)) die("cannot open .stderr");
  setvbuf(stdout, NULL, _IONBF, 0);
  setvbuf(stderr, NULL, _IONBF, 0);
  printf("\n Command line: %s %s", argv[0], argv[argc - 1]);
  for (int i = 1; i < argc - 1; ++i) printf(" %s", argv[i]);
  printf("\n");

  const Param P = read_parameter_file(o.paramfile);
  printf("\n Parameter file read...");

  prc_d2z_params lp;
  prc_d2z_default_params(&lp);
  lp.Omegam = P.Omegam;
  lp.Omegab = P.Omegab;
  lp.Omegal = P.Omegal;
  lp.h = P.h;
  lp.ns = P.ns;
  lp.sigma8 = P.sigma8;
  lp.w = P.w;
  lp.BoxSize = P.BoxSize;
  lp.zInit = P.zInit;
  lp.Rmin = P.Rmin;
  lp.N = P.N;
  lp.Nscales = P.Nscales;
  lp.RecordBubbleSizes = P.RecordBubbleSizes;
  lp.Rmax = o.Rmax;
  lp.zeta = o.zeta;
  lp.M_min = o.M_min;
  lp.mode = o.mode;
  lp.radii_per_launch = o.radii_per_launch;
  lp.device = o.device;

  prc_d2z_ctx* ctx = nullptr;
  check(prc_d2z_create(&lp, &ctx), nullptr, "prc_d2z_create");
  printf("\n Arrays allocated... (%.2f GB on the device)", prc_d2z_device_bytes(ctx) / 1e9);

  {
    std::vector<float> delta = read_delta_file(P.DeltaFile, P.N);
    check(prc_d2z_set_delta(ctx, delta.data(), PRC_D2Z_DENSE), ctx, "prc_d2z_set_delta");
  }
  check(prc_d2z_run(ctx), ctx, "prc_d2z_run");

  // per-radius table (delta2zreion_d2z.C:46-52,102-103)
  std::vector<prc_d2z_radius_row> rows(P.Nscales);
  check(prc_d2z_get_radius_table(ctx, rows.data(), P.Nscales), ctx, "radius table");
  printf("\n");
  printf("\n Nscales = %2d    Rmin=%5.3f", P.Nscales, rows[0].R_mpch);
  printf("\n                 Rmax=%5.3f\n", rows[P.Nscales - 1].R_mpch);
  printf("\n      Scale       R [Mpc/h]        sigma(R)      sigma_grid");
  printf("\n ==========================================================\n");
  for (int r = 0; r < P.Nscales; ++r)
    printf("\n %10d %15.3f %15.3f %15.3f", r, rows[r].R_mpch, rows[r].sigma, rows[r].sigma_grid);
  printf("\n");

  // GetHistory of the uncorrected field (main_d2z.C:76): the input of the
  // bubble statistics and of the correction
  std::vector<float> history, history_z;
  (void)get_history(ctx, P, history, history_z);

  // GetBubbleMFP (bubblehistory_d2z.C:82-242), WriteMFP and WriteMFPPDF
  // (io_d2z.C:343-400)
  std::vector<float> mfp(PRC_D2Z_NX_MFP), mfp_x(PRC_D2Z_NX_MFP), dfdmfp((size_t)PRC_D2Z_NX_MFP * PRC_D2Z_NMFP);
  check(prc_d2z_bubble_mfp(ctx, history.data(), 0, 1, mfp.data(), mfp_x.data(), dfdmfp.data()), ctx, "bubble mfp");
  printf("\n Writing bubble mfp...");
  {
    FILE* f = fopen((o.basename + ".mfp").c_str(), "w");
    if (!f) die("cannot write %s.mfp", o.basename.c_str());
    for (int i = 0; i < PRC_D2Z_NX_MFP; ++i) fprintf(f, "%f %f\n", mfp_x[i], mfp[i]);
    fclose(f);
  }
  printf("\n Writing bubble mfp...");
  {
    FILE* f = fopen((o.basename + ".mfppdf").c_str(), "wb");
    if (!f) die("cannot write %s.mfppdf", o.basename.c_str());
    // BoxSize in Mpc, as the reference holds it after main_d2z.C:29
    const float box = P.BoxSize / P.h;
    const float cell = box / (float)P.N;
    const int header1[2] = {PRC_D2Z_NX_MFP, PRC_D2Z_NMFP};
    const float header2[4] = {0.01f, 0.999f, log10f(cell), (float)log10(1.0 * (double)box)};
    if (fwrite(header1, sizeof(int), 2, f) != 2 || fwrite(header2, sizeof(float), 4, f) != 4 ||
        fwrite(dfdmfp.data(), sizeof(float), dfdmfp.size(), f) != dfdmfp.size())
      die("cannot write %s.mfppdf", o.basename.c_str());
    fclose(f);
  }

  // GetCHistory and CorrectHistory (bubblehistory_d2z.C:44-80,301-343) with the
  // Float2FloatTable file (tables.C:9-25: float min, float max, int n, float[n])
  std::vector<float> table;
  float tmin = 0, tmax = 0;
  if (FILE* f = fopen(o.fcoll_table.c_str(), "rb")) {
    int n = 0;
    if (fread(&tmin, 4, 1, f) != 1 || fread(&tmax, 4, 1, f) != 1 || fread(&n, 4, 1, f) != 1 || n < 2 ||
        n > (1 << 28))
      die("malformed table %s", o.fcoll_table.c_str());
    table.resize(n);
    if (fread(table.data(), 4, (size_t)n, f) != (size_t)n) die("short table %s", o.fcoll_table.c_str());
    fclose(f);
    std::vector<float> chistory(PRC_D2Z_NCHISTORY);
    check(prc_d2z_corrected_history(ctx, table.data(), n, tmin, tmax, mfp.data(), chistory.data()), ctx,
          "corrected history");
    check(prc_d2z_correct_zreion(ctx, history.data(), chistory.data()), ctx, "history correction");
  } else {
    printf("\n %s not found: the history correction is skipped and the uncorrected zreion is written",
           o.fcoll_table.c_str());
  }

  // WriteZreionFile (io_d2z.C:209-246)
  printf("\n Writing reionization redshifts...");
  {
    const size_t n = (size_t)P.N * P.N * P.N;
    std::vector<float> z(n);
    check(prc_d2z_get_zreion(ctx, z.data(), PRC_D2Z_DENSE), ctx, "prc_d2z_get_zreion");
    FILE* f = fopen((o.basename + ".zreion").c_str(), "wb");
    if (!f || fwrite(z.data(), sizeof(float), n, f) != n) die("cannot write %s.zreion", o.basename.c_str());
    fclose(f);
  }
  printf("\n Reionization redshifts written...");

  // GetHistory of the written (corrected) field and its tau (main_d2z.C:97-98)
  const float tau = get_history(ctx, P, history, history_z);
  printf("\n\n tau = %f", tau);

  // WriteHistory (io_d2z.C:284-314): z, x_HII(z), the fcoll table at z and
  // fcoll(z) of the barrier's sigma_min.  The reference reads fcoll_table.tab
  // here unconditionally; without it the third column is nan.
  {
    prc_d2z_scalars sc;
    check(prc_d2z_get_scalars(ctx, &sc), ctx, "scalars");
    printf("\n Writing ionization history...");
    FILE* f = fopen((o.basename + ".history").c_str(), "w");
    if (!f) die("cannot write %s.history", o.basename.c_str());
    for (int i = 0; i < kNHistory; ++i) {
      const float zi = history_z[i];
      const float ft = table.empty() ? NAN : table_value(table, tmin, tmax, zi);
      fprintf(f, "%e %e %e %e\n", zi, history[i], ft, fcoll_ref(zi, P.Omegam, P.Omegal, P.w, sc.sigma_min));
    }
    fclose(f);
  }
  printf("\n\n");
  prc_d2z_destroy(ctx);
  return 0;
}
