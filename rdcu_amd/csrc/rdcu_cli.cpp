// rdcu_cli.cpp -- drop-in replacement for the rdcu executable's non-stiff path.
//
// Same flags, defaults, help text, factor-index listing, coupling syntax,
// input files and binary output files as the reference driver
// (rdcu.cu:599-1273, step_eval rdcu.cu:252-341).  The integrator behind it is
// whatever library provides include/rdcu_dp45.h: the HIP library (binary
// `rdcu`) or the CPU oracle (binary `rdcu_cpu`, used as the checker and the
// timed CPU baseline).
//
// Differences from the reference, all deliberate:
//   * -S (BDF) and -D4 (Jacobian output, which needs -S) are rejected with a
//     message and exit status 2: the stiff path is not part of this build.
//   * An out-of-range field or factor index in a term is an error (exit 2)
//     instead of an out-of-bounds read.
//   * The output-time loop checks its bound before reading t_eval
//     (rdcu.cu:268 reads one past the end).
//   * Integrator failures are reported and give exit status 1.
//   * FILEBASEY.dat imaginary parts hold the Nyquist leakage of the first
//     and mixed derivative blocks, as the reference's complex transform
//     does; block 0 (uninitialised memory in the reference) and the pure
//     second derivatives (roundoff) hold 0.
//   * FILEBASEfs.dat is replaced atomically (temporary file, then rename).

#include <getopt.h>
#include <sys/time.h>
#include <unistd.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "rdcu_dp45.h"

namespace {

// strtok-based integer list, as parse_list (rdcu.cu:343-357)
void parse_list(const char *arg, const char *delim, int *lst, int *len, int max_len) {
  if (!arg) return;
  char *copy = strdup(arg);
  char *tok = strtok(copy, delim);
  while (tok != NULL) {
    lst[(*len)++] = atoi(tok);
    tok = strtok(NULL, delim);
    if (*len >= max_len) break;
  }
  free(copy);
}

// strtok-based double list returning the characters consumed, as fparse_list
// (rdcu.cu:359-374)
int fparse_list(const char *arg, const char *delim, double *lst, int *len, int max_len) {
  int chars = 0;
  if (!arg) return chars;
  char *copy = strdup(arg);
  char *tok = strtok(copy, delim);
  while (tok != NULL) {
    chars += (int)strlen(tok) + 1;
    lst[(*len)++] = atof(tok);
    tok = strtok(NULL, delim);
    if (*len >= max_len) break;
  }
  free(copy);
  return chars;
}

struct TermSpec {
  double C;
  std::vector<int> ints;  // field, then (index, power) pairs
};

// "C,field,idx,eta,..." -> term (rdcu.cu:654-661, 831-835)
TermSpec parse_term(const char *s) {
  TermSpec t;
  double C[1];
  int nC = 0;
  int chars = fparse_list(s, ",", C, &nC, 1);
  t.C = nC ? C[0] : 0.0;
  std::vector<int> buf(1024, 0);
  int cnt = 0;
  if ((size_t)chars <= strlen(s)) parse_list(s + chars, ",", buf.data(), &cnt, 1024);
  t.ints.assign(buf.begin(), buf.begin() + cnt);
  return t;
}

double elapsed(const timeval &a) {
  timeval b;
  gettimeofday(&b, NULL);
  return b.tv_sec - a.tv_sec + 1e-6 * (b.tv_usec - a.tv_usec);
}

struct Driver {
  rdcu_ctx *ctx = nullptr;
  int n = 1, ndim = 1, N = 1, nblk = 3;
  int dense = 1, verbose = 0, steps = 0;
  double t0 = 0, t1 = 100;
  std::vector<double> t_eval;
  size_t eval_i = 0;
  std::string filebase;
  timeval start;
  FILE *out = nullptr, *outsteps = nullptr, *outtimes = nullptr, *outstates = nullptr,
       *outf = nullptr, *outY = nullptr;
  int failed = 0;
};

// Host buffers of one output frame (state, RHS, factor blocks), shared by
// every instance of the process: frames are written one at a time.  They are
// page-locked (rdcu_host_alloc), so the device copies into them run as DMA,
// and grow on demand.  release_frames() frees them before exit.
double *g_frame[3] = {nullptr, nullptr, nullptr};
size_t g_frame_cap[3] = {0, 0, 0};

double *frame_buf(int i, size_t count) {
  if (count > g_frame_cap[i]) {
    rdcu_host_free(g_frame[i]);
    g_frame[i] = nullptr;
    g_frame_cap[i] = 0;
    void *q = nullptr;
    if (rdcu_host_alloc(count * sizeof(double), &q)) return nullptr;
    g_frame[i] = static_cast<double *>(q);
    g_frame_cap[i] = count;
  }
  return g_frame[i];
}

void release_frames() {
  for (int i = 0; i < 3; ++i) {
    rdcu_host_free(g_frame[i]);
    g_frame[i] = nullptr;
    g_frame_cap[i] = 0;
  }
}

// The frame buffers one output of `d` needs (NULL where its -D level writes
// nothing); false if an allocation failed.
bool frame_bufs(const Driver &d, double **y, double **f, double **Y) {
  const size_t nN = (size_t)d.n * d.N;
  *y = frame_buf(0, nN);
  *f = d.dense >= 2 ? frame_buf(1, nN) : nullptr;
  *Y = d.dense >= 3 ? frame_buf(2, 2 * nN * d.nblk) : nullptr;
  if (*y && (d.dense < 2 || *f) && (d.dense < 3 || *Y)) return true;
  fprintf(stderr, "frame buffers: %s\n", rdcu_last_error());
  return false;
}

void put(FILE *f, const void *p, size_t sz, size_t cnt) {
  if (f && cnt) {
    fwrite(p, sz, cnt, f);
    fflush(f);
  }
}

// FILEBASEfs.dat = state, t, h (the layout -R reads back, rdcu.cu:322-328).
// Written to a temporary name and renamed over the old file, so a run killed
// mid-write leaves the previous restart point intact.
void write_fs(const std::string &filebase, const double *y, size_t nN, double t, double h) {
  const std::string fn = filebase + "fs.dat", tmp = fn + ".tmp";
  FILE *fs = fopen(tmp.c_str(), "wb");
  if (!fs) return;
  bool ok = fwrite(y, sizeof(double), nN, fs) == nN;
  ok = ok && fwrite(&t, sizeof(double), 1, fs) == 1;
  ok = ok && fwrite(&h, sizeof(double), 1, fs) == 1;
  ok = (fclose(fs) == 0) && ok;
  if (ok && rename(tmp.c_str(), fn.c_str()) == 0) return;
  fprintf(stderr, "could not write %s\n", fn.c_str());
  remove(tmp.c_str());
}

// Per-attempt output, as step_eval (rdcu.cu:252-341).
int step_eval(rdcu_ctx *ctx, double t, double h, int success, int current, double *stop,
              void *user) {
  Driver &d = *static_cast<Driver *>(user);
  const size_t nN = (size_t)d.n * d.N;
  if (d.dense >= 1) put(d.outsteps, &h, sizeof(double), 1);
  if (!success) return 0;
  d.steps++;
  int eval = 0;
  while (d.eval_i < d.t_eval.size() && t >= d.t_eval[d.eval_i]) {
    if (!current) {
      fprintf(stderr, "internal error: output time crossed by a non-current attempt\n");
      d.failed = 1;
      return 1;
    }
    const double te = d.t_eval[d.eval_i];
    double *yb, *fb, *Yb;
    if (!frame_bufs(d, &yb, &fb, &Yb)) {
      d.failed = 1;
      return 1;
    }
    if (rdcu_frame(ctx, te, yb, fb, Yb)) {
      fprintf(stderr, "rdcu_frame: %s\n", rdcu_last_error());
      d.failed = 1;
      return 1;
    }
    if (d.dense >= 1) put(d.outstates, yb, sizeof(double), nN);
    if (d.dense >= 2) put(d.outf, fb, sizeof(double), nN);
    if (d.dense >= 3) put(d.outY, Yb, sizeof(double), 2 * nN * d.nblk);
    put(d.outtimes, &te, sizeof(double), 1);
    d.eval_i++;
    eval = 1;
  }
  *stop = d.eval_i < d.t_eval.size() ? d.t_eval[d.eval_i] : INFINITY;
  if (eval) {
    double el = elapsed(d.start);
    double frac = (t - d.t0) / (d.t1 - d.t0);
    double eta = el / ((t - d.t0 + h) / (d.t1 - d.t0)) * (1 - frac);
    if (d.verbose) {
      printf("%.3f\t%1.3e\t%1.3e\t%f\t%i\t\r", frac, el, eta, h, d.steps);
      fflush(stdout);
    }
    fprintf(d.out, "%.3f\t%1.3e\t%1.3e\t%f\t%i\t\n", frac, el, eta, h, d.steps);
    fflush(d.out);
    double *yb = frame_buf(0, nN);
    if (!yb || rdcu_get_state(ctx, yb)) {
      fprintf(stderr, "rdcu_get_state: %s\n", rdcu_last_error());
      d.failed = 1;
      return 1;
    }
    write_fs(d.filebase, yb, nN, t, h);
  }
  return 0;
}

void print_help(int n, int ndim) {
  printf("usage:\trdcu [-hvSFR] [-n NFIELDS] [-N NUMS] [-L LENGTHS]\n");
  printf("\t[-c COUPLING] [-A AMPLITUDE] [-t TIME] [-d DT] [-s SEED] \n");
  printf("\t[-D DENSITY] [-g GPU] [-r RTOL] [-a ATOL]  FILEBASE \n\n");
  printf("-h for help. Call with -n and -N to see Y indices. \n");
  printf("-v for verbose output, including progress \n");
  printf("-S for stiff system. Use the BDF method with finite differences instead of RK45 "
         "pseudospectral method. (Not available in this build.)\n");
  printf("-F for fixed timestep \n");
  printf("-R to reload final state from FILEBASEfs.dat if possible\n");
  printf("NFIELDS is the number of fields. Default 1\n");
  printf("NUMS is number of grid points in each dimension (up to three), separated by commas "
         "(no spaces). Default 128\n");
  printf("LENGTHS is domain length in each dimension (up to three), separated by commas (no "
         "spaces). Default 1.0\n");
  printf("COUPLING is a list of coupling terms, separated by commas (no spaces). Default empty\n");
  printf("\tThe first value is the term coefficent. \n");
  printf("\tThe second value is an integer specifying the field that the term appears in. \n");
  printf("\tThe following 2N values are pairs of integers specifing the indices and powers for "
         "each factor that appear in the term. \n");
  printf("\tFactor indices for the given -n and -N values appear below. \n");
  printf("\tYou may specify additional coupling terms on separate lines in this format in the "
         "input file FILEBASEcoupling.dat. \n");
  printf("AMPLITUDE is uniform random initial condition amplitude. Default 1.0 \n");
  printf("\tYou may also provide a binary input file FILEBASEic.dat with the initial condition\n");
  printf("TIME is total integration time. Default 1e2 \n");
  printf("DT is the time between outputs. Default 1e0 \n");
  printf("\t You may include firststep, minstep, maxstep as additional comma-separated values \n");
  printf("\t If not included, the first step is 1/100th the value of dt \n");
  printf("SEED is random seed. Default 1 \n");
  printf("GPU is the index of the gpu. Default 0\n");
  printf("DENSITY is the output density. Default 1\n");
  printf("\t1 for the timesteps (FILEBASEtimes.dat) and evaluated state values "
         "(FILEBASEstates.dat), \n");
  printf("\t2 to include time derivatives (FILEBASEf.dat), 3 to include factors (FILEBASEY.dat), "
         "\n");
  printf("\t4 to include CSR Jacobian elements (FILEBASErows.dat, FILEBASEcols.dat, and "
         "FILEBASEvals.dat). \n");
  printf("RTOL is relative error tolerance. Default 1E-8\n");
  printf("ATOL is absolute error tolerance. Default 1E-8\n");
  printf("FILEBASE is base file name for output. \n");
  printf("\n");
  printf("Example: ./rdcu -N 128,128 -L 100.0,100.0 -n 2 -c 1.0,0,0,1 -c 1.0,0,6,1 -c "
         "1.0,0,12,1 -c -2.0,0,7,1 -c -2.0,0,13,1 -c -1.0,0,0,3 -c -1.0,0,0,1,1,2 -c "
         "-0.8,0,0,2,1,1 -c -0.8,0,1,3 -c 1.0,1,1,1 -c 1.0,1,7,1 -c 1.0,1,13,1 -c 2.0,1,6,1 -c "
         "2.0,1,12,1 -c -1.0,1,0,2,1,1 -c -1.0,1,1,3 -c 0.8,1,0,3 -c 0.8,1,0,1,1,2 -v -D3 "
         "2dcgle\n");
  printf("\n");
  printf("Indices for %i field(s) in %i dimension(s):\n", n, ndim);
  int l = 0;
  for (int i = 0; i < n; i++) printf("%i: u%i\n", l++, i);
  for (int j = 0; j < ndim; j++)
    for (int i = 0; i < n; i++) printf("%i: u%i_%i\n", l++, i, j);
  for (int j = 0; j < ndim; j++)
    for (int k = 0; k < ndim; k++)
      for (int i = 0; i < n; i++) printf("%i: u%i_%i%i\n", l++, i, j, k);
}

// One instance of the command line: everything main() sets up before the
// integrator exists (rdcu.cu:599-1059, non-stiff parts).
struct Instance {
  Driver d;
  double t = 0, h = 0, t1 = 1e2, dt = 1e0, atl = 1e-8, rtl = 1e-8, hmin = 0, hmax = INFINITY;
  int gpu = 0, fixed = 0, reload = 0;
  std::vector<double> coef;
  std::vector<int> tfield, nfac, fac;
  rdcu_problem prob;
  std::vector<double> yloc;
};

enum { PREP_OK = 0, PREP_FAIL = 1, PREP_USAGE = 2, PREP_HELP = 3 };

// Parses one command line and prepares its instance: flags, coupling terms,
// .out header, initial state (file, -R reload or glibc rand), output files
// and output times.  argv[0] is the program name.
int prepare(int argc, char *argv[], Instance &I) {
  Driver &d = I.d;
  gettimeofday(&d.start, NULL);

  // defaults (rdcu.cu:603-607)
  double A = 1E-3, A0 = 0.0;
  int seed = 1, n = 1, ndim = 1, ndim2 = 0, verbose = 0, help = 1, dense = 1, stiff = 0;
  int Nsloc[3] = {128, 128, 128};
  double Lsloc[3] = {1.0, 1.0, 1.0};
  std::vector<TermSpec> terms;
  const char *filebase = nullptr;

  // getopt loop (rdcu.cu:612-731); optind = 0 restarts glibc's scan for a
  // new argument vector
  optind = 0;
  int ch;
  while (optind < argc) {
    if ((ch = getopt(argc, argv, "hvFRSn:N:L:c:A:t:d:s:D:g:r:a:")) != -1) {
      switch (ch) {
        case 'h': help = 1; break;
        case 'v': verbose = 1; break;
        case 'F': I.fixed = 1; break;
        case 'R': I.reload = 1; break;
        case 'S': stiff = 1; break;
        case 'n': if (optarg) n = atoi(optarg); break;
        case 'N':
          if (optarg) {
            ndim = 0;
            parse_list(optarg, ",", Nsloc, &ndim, 3);
          }
          break;
        case 'L':
          if (optarg && ndim2 < 3) fparse_list(optarg, ",", Lsloc, &ndim2, 3);
          break;
        case 'c':
          if (optarg) terms.push_back(parse_term(optarg));
          break;
        case 'A':
          if (optarg) {
            double As[4] = {0, 0, 0, 0};
            int nAs = 0;
            fparse_list(optarg, ",", As, &nAs, 4);
            A = As[0];
            if (nAs > 1) A0 = As[1];
          }
          break;
        case 't': if (optarg) I.t1 = atof(optarg); break;
        case 'd':
          if (optarg) {
            double dts[4];
            int ndts = 0;
            fparse_list(optarg, ",", dts, &ndts, 4);
            I.dt = dts[0];
            if (ndts > 1) I.h = dts[1];
            if (ndts > 2) I.hmin = dts[2];
            if (ndts > 3) I.hmax = dts[3];
          }
          break;
        case 's': if (optarg) seed = atoi(optarg); break;
        case 'D': if (optarg) dense = atoi(optarg); break;
        case 'g': if (optarg) I.gpu = atoi(optarg); break;
        case 'r': if (optarg) I.rtl = atof(optarg); break;
        case 'a': if (optarg) I.atl = atof(optarg); break;
      }
    } else {
      filebase = argv[optind];
      optind++;
      help = 0;
    }
  }

  if (help) {
    print_help(n, ndim);
    return PREP_HELP;
  }
  if (stiff) {
    fprintf(stderr, "rdcu: the stiff BDF path (-S) is not part of this build\n");
    return PREP_USAGE;
  }
  if (dense >= 4) {
    fprintf(stderr, "rdcu: -D4 (CSR Jacobian output) needs the stiff path (-S), which is not "
                    "part of this build\n");
    return PREP_USAGE;
  }
  const double t1 = I.t1, dt = I.dt;
  const int fixed = I.fixed, reload = I.reload;
  double &t = I.t, &h = I.h;

  // factor names (rdcu.cu:789-805)
  std::vector<std::string> sterms;
  char buf[64];
  for (int i = 0; i < n; i++) {
    snprintf(buf, sizeof buf, "u%i", i);
    sterms.push_back(buf);
  }
  for (int j = 0; j < ndim; j++)
    for (int i = 0; i < n; i++) {
      snprintf(buf, sizeof buf, "u%i_%i", i, j);
      sterms.push_back(buf);
    }
  for (int j = 0; j < ndim; j++)
    for (int k = 0; k < ndim; k++)
      for (int i = 0; i < n; i++) {
        snprintf(buf, sizeof buf, "u%i_%i%i", i, j, k);
        sterms.push_back(buf);
      }

  if (ndim2 < ndim)
    printf("Warning: Number of length scales %i smaller than number of dimensions %i. Using "
           "defaults.\n",
           ndim2, ndim);

  d.filebase = filebase;
  std::string file = d.filebase + ".out";
  d.out = fopen(file.c_str(), "w");
  if (!d.out) {
    fprintf(stderr, "rdcu: cannot open %s\n", file.c_str());
    return PREP_FAIL;
  }
  for (int i = 0; i < argc; i++) fprintf(d.out, "%s ", argv[i]);
  fprintf(d.out, "\n");
  fflush(d.out);

  // coupling file, one term per whitespace-separated token (rdcu.cu:822-837)
  file = d.filebase + "coupling.dat";
  if (FILE *in = fopen(file.c_str(), "r")) {
    if (verbose) {
      printf("Using coupling terms from file\n");
      fprintf(d.out, "Using coupling terms from file\n");
    }
    char line[1024];
    while (fscanf(in, "%1023s", line) == 1) terms.push_back(parse_term(line));
    fclose(in);
  }

  const int nblk = 1 + ndim + ndim * ndim;
  const int nidx = n * nblk;
  // terms -> dictionary arrays; a trailing index without a power gets power 0
  std::vector<double> &coef = I.coef;
  std::vector<int> &tfield = I.tfield, &nfac = I.nfac, &fac = I.fac;
  for (const TermSpec &t : terms) {
    int fld = t.ints.empty() ? 0 : t.ints[0];
    int np = (int)t.ints.size();
    int nf = np / 2;
    if (fld < 0 || fld >= n) {
      fprintf(stderr, "rdcu: term field %d out of range (n=%d)\n", fld, n);
      return PREP_USAGE;
    }
    coef.push_back(t.C);
    tfield.push_back(fld);
    nfac.push_back(nf);
    for (int j = 0; j < nf; ++j) {
      int ix = t.ints[1 + 2 * j];
      int pw = 2 + 2 * j < np ? t.ints[2 + 2 * j] : 0;
      if (ix < 0 || ix >= nidx) {
        fprintf(stderr, "rdcu: factor index %d out of range (%d indices)\n", ix, nidx);
        return PREP_USAGE;
      }
      fac.push_back(ix);
      fac.push_back(pw);
    }
  }

  if (verbose) {  // equations (rdcu.cu:839-859)
    size_t off = 0;
    std::vector<size_t> offs;
    for (size_t i = 0; i < terms.size(); ++i) {
      offs.push_back(off);
      off += 2 * nfac[i];
    }
    for (int k = 0; k < n; k++) {
      printf("%s'=", sterms[k].c_str());
      fprintf(d.out, "%s'=", sterms[k].c_str());
      for (size_t i = 0; i < terms.size(); i++) {
        if (tfield[i] != k) continue;
        printf("%.3f", coef[i]);
        fprintf(d.out, "%.3f", coef[i]);
        for (int j = 0; j < nfac[i]; j++) {
          printf("(%s)^%i", sterms[fac[offs[i] + 2 * j]].c_str(), fac[offs[i] + 2 * j + 1]);
          fprintf(d.out, "(%s)^%i", sterms[fac[offs[i] + 2 * j]].c_str(),
                  fac[offs[i] + 2 * j + 1]);
        }
        printf(" + ");
        fprintf(d.out, " + ");
      }
      printf("0\n");
      fprintf(d.out, "0\n");
    }
  }
  fflush(stdout);
  fflush(d.out);

  t = 0;
  long N = Nsloc[0];
  for (int i = 1; i < ndim; i++) N *= Nsloc[i];
  const size_t nN = (size_t)n * N;

  srand(seed);
  if (h <= 0.0) h = fixed ? dt : dt / 100;  // rdcu.cu:917-924

  std::vector<double> &yloc = I.yloc;
  yloc.assign(nN, 0.0);
  // initial conditions (rdcu.cu:926-1008)
  file = d.filebase + "ic.dat";
  int loaded = 0;
  if (FILE *in = fopen(file.c_str(), "r")) {
    loaded = 1;
    printf("Using initial conditions from file\n");
    fprintf(d.out, "Using initial conditions from file\n");
    size_t rd = fread(yloc.data(), sizeof(double), nN, in);
    if (rd != nN) {
      loaded = 0;
      printf("initial conditions file not compatible with N!\n");
      fprintf(d.out, "initial conditions file not compatible with N!\n");
    }
    fclose(in);
  }
  file = d.filebase + "fs.dat";
  int reloaded = 0;
  FILE *in = nullptr;
  if (reload && (in = fopen(file.c_str(), "r"))) {
    reloaded = 1;
    loaded = 1;
    printf("Reloading final conditions from file\n");
    fprintf(d.out, "Reloading final conditions from file\n");
    size_t rd = fread(yloc.data(), sizeof(double), nN, in);
    if (rd != nN) {
      printf("final conditions file not compatible with N!\n");
      fprintf(d.out, "final conditions file not compatible with N!\n");
      reloaded = 0;
      loaded = 0;
    }
    if (reloaded) {
      rd = fread(&t, sizeof(double), 1, in);
      rd = fread(&h, sizeof(double), 1, in);
      if (rd != 1) {
        printf("Couldn't read start time and step!\n");
        fprintf(d.out, "Couldn't read start time and step!\n");
        reloaded = 0;
      }
    }
    fclose(in);
    printf("Restarting at t=%f with h=%f\n", t, h);
    fprintf(d.out, "Restarting at t=%f with h=%f\n", t, h);
  }
  if (!reloaded && !loaded) {
    printf("Using random initial conditions\n");
    fprintf(d.out, "Using random initial conditions\n");
    for (int i = 0; i < n; i++)
      for (long j = 0; j < N; j++) yloc[N * i + j] = A0 + A / 2 * (2.0 / RAND_MAX * rand() - 1);
    write_fs(d.filebase, yloc.data(), nN, t, h);
  }
  fflush(d.out);
  fflush(stdout);

  const char *wt = loaded ? "ab" : "wb";
  auto openf = [&](const char *suffix) { return fopen((d.filebase + suffix).c_str(), wt); };
  if (dense >= 1) {
    d.outsteps = openf("steps.dat");
    d.outtimes = openf("times.dat");
    d.outstates = openf("states.dat");
  }
  if (dense >= 2) d.outf = openf("f.dat");
  if (dense >= 3) d.outY = openf("Y.dat");

  // output times as products dt*nn (rdcu.cu:1052-1059)
  int n0 = int(t / dt) + 1;
  int n1 = int(t1 / dt) + 1;
  for (int nn = n0; nn < n1; nn++) d.t_eval.push_back(dt * nn);

  rdcu_problem &prob = I.prob;
  memset(&prob, 0, sizeof prob);
  prob.n = n;
  prob.ndim = ndim;
  for (int a = 0; a < 3; ++a) {
    prob.Ns[a] = Nsloc[a];
    prob.Ls[a] = Lsloc[a];
  }
  prob.nterms = (int)coef.size();
  prob.coef = coef.data();
  prob.field = tfield.data();
  prob.nfac = nfac.data();
  prob.fac = fac.data();

  d.n = n;
  d.ndim = ndim;
  d.N = (int)N;
  d.nblk = nblk;
  d.dense = dense;
  d.verbose = verbose;
  d.t0 = t;
  d.t1 = t1;
  return PREP_OK;
}

// Initial frame (rdcu.cu:1155-1191), once the instance has a handle; keyed on
// the -R flag, as there.
int first_frame(Instance &I) {
  Driver &d = I.d;
  const size_t nN = I.yloc.size();
  if (d.dense >= 1 && !I.reload) {
    put(d.outstates, I.yloc.data(), sizeof(double), nN);
    put(d.outtimes, &I.t, sizeof(double), 1);
  }
  if (d.dense >= 2 && !I.reload) {
    double *yb, *fb, *Yb;
    if (!frame_bufs(d, &yb, &fb, &Yb)) return 1;
    if (rdcu_rhs(d.ctx, I.yloc.data(), fb, Yb)) {
      fprintf(stderr, "rdcu_rhs: %s\n", rdcu_last_error());
      return 1;
    }
    put(d.outf, fb, sizeof(double), nN);
    if (d.dense >= 3) put(d.outY, Yb, sizeof(double), 2 * nN * d.nblk);
  }
  return 0;
}

// Closing lines of the .out file (rdcu.cu:1265-1273) and the output files.
void finish(Instance &I, bool failed) {
  Driver &d = I.d;
  if (failed) fprintf(d.out, "integration failed: %s\n", rdcu_last_error());
  double rt = elapsed(d.start);
  printf("\nruntime: %f\n", rt);
  fprintf(d.out, "\nsteps: %i\n", d.steps);
  fprintf(d.out, "runtime: %f\n", rt);
  fflush(d.out);
  fclose(d.out);
  d.out = nullptr;
  for (FILE **f : {&d.outtimes, &d.outsteps, &d.outstates, &d.outf, &d.outY})
    if (*f) {
      fclose(*f);
      *f = nullptr;
    }
}

int run_single(Instance &I) {
  Driver &d = I.d;
  int st = rdcu_create(&I.prob, I.atl, I.rtl, I.fixed, I.yloc.data(), I.gpu, &d.ctx);
  if (st) {
    printf("integrator initialization failed: %s\n", rdcu_last_error());
    fprintf(d.out, "integrator initialization failed: %s\n", rdcu_last_error());
    fclose(d.out);
    d.out = nullptr;
    return 1;
  }
  if (first_frame(I)) return 1;
  double stop0 = d.t_eval.empty() ? INFINITY : d.t_eval[0];
  st = rdcu_run(d.ctx, &I.t, &I.h, I.hmin, I.hmax, I.t1, stop0, &step_eval, &d);
  const bool failed = st || d.failed;
  if (failed) fprintf(stderr, "rdcu_run failed: %s\n", rdcu_last_error());
  finish(I, failed);
  rdcu_destroy(d.ctx);
  d.ctx = nullptr;
  return failed ? 1 : 0;
}

}  // namespace

#ifndef RDCU_SWEEP
int main(int argc, char *argv[]) {
  Instance I;
  const int p = prepare(argc, argv, I);
  if (p == PREP_HELP) return 0;
  if (p != PREP_OK) return p;
  const int rc = run_single(I);
  release_frames();
  return rc;
}
#endif

#ifdef RDCU_SWEEP
// rdcu_sweep MANIFEST -- a parameter sweep as batches instead of one process
// per instance (the job array of sweep.sh:21-41, whose every task runs
// makelines.py:260-268 -> ./rdcu).  Every non-empty line of MANIFEST ('#'
// starts a comment) is the argument list of one rdcu run: the same flags and
// FILEBASE, split at whitespace.  Each instance writes exactly the files that
// rdcu would write for its line.  Instances that agree on -F, -r, -a, -g and
// the step bounds of -d share one batch (rdcu_batch_*); a batch runs one
// attempt per call, so every attempt's step size (steps.dat) and every
// output time are seen as rdcu sees them.  A group the batch cannot hold
// (3-D grids, lengths other than 2^k and 3*2^k) runs one instance at a time.
namespace {

std::vector<std::string> split_ws(const std::string &line) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && isspace((unsigned char)line[i])) ++i;
    size_t j = i;
    while (j < line.size() && !isspace((unsigned char)line[j])) ++j;
    if (j > i) out.push_back(line.substr(i, j - i));
    i = j;
  }
  return out;
}

struct GroupKey {
  int fixed, gpu;
  double atl, rtl, hmin, hmax;
  bool operator<(const GroupKey &o) const {
    if (fixed != o.fixed) return fixed < o.fixed;
    if (gpu != o.gpu) return gpu < o.gpu;
    if (atl != o.atl) return atl < o.atl;
    if (rtl != o.rtl) return rtl < o.rtl;
    if (hmin != o.hmin) return hmin < o.hmin;
    return hmax < o.hmax;
  }
};

// Runs the instances of one group as a batch; 0 on success.
int run_batch(std::vector<Instance *> &g) {
  const int cnt = (int)g.size();
  const Instance &g0 = *g[0];
  std::vector<rdcu_problem> probs(cnt);
  std::vector<const double *> y0s(cnt);
  for (int k = 0; k < cnt; ++k) {
    probs[k] = g[k]->prob;
    y0s[k] = g[k]->yloc.data();
  }
  rdcu_batch *b = nullptr;
  if (rdcu_batch_create(probs.data(), cnt, y0s.data(), g0.atl, g0.rtl, g0.fixed, g0.gpu, &b)) {
    fprintf(stderr, "rdcu_sweep: %d instance(s) not batched (%s): running them one at a time\n",
            cnt, rdcu_last_error());
    int rc = 0;
    for (Instance *I : g) rc |= run_single(*I);
    return rc;
  }
  int rc = 0;
  for (int k = 0; k < cnt && !rc; ++k) {
    if (rdcu_batch_member(b, k, &g[k]->d.ctx) || first_frame(*g[k])) {
      fprintf(stderr, "rdcu_sweep: %s\n", rdcu_last_error());
      rc = 1;
    }
  }
  std::vector<double> t(cnt), h(cnt), t1(cnt);
  std::vector<long long> att(cnt), acc0(cnt);
  for (int k = 0; k < cnt; ++k) {
    t[k] = g[k]->t;
    h[k] = g[k]->h;
    t1[k] = g[k]->t1;
  }
  long long stats[4];
  while (!rc) {
    for (int k = 0; k < cnt; ++k) {
      rdcu_get_stats(g[k]->d.ctx, stats);
      acc0[k] = stats[1];
    }
    if (rdcu_batch_run(b, t.data(), h.data(), g0.hmin, g0.hmax, t1.data(), 1, att.data())) {
      fprintf(stderr, "rdcu_batch_run failed: %s\n", rdcu_last_error());
      rc = 1;
      break;
    }
    bool any = false;
    for (int k = 0; k < cnt; ++k) {
      if (att[k] == 0) continue;
      any = true;
      rdcu_get_stats(g[k]->d.ctx, stats);
      double stop;
      step_eval(g[k]->d.ctx, t[k], h[k], stats[1] > acc0[k], 1, &stop, &g[k]->d);
      if (g[k]->d.failed) rc = 1;
    }
    if (!any) break;
  }
  for (int k = 0; k < cnt; ++k) {
    g[k]->t = t[k];
    g[k]->h = h[k];
    if (g[k]->d.out) finish(*g[k], rc != 0);
  }
  rdcu_batch_destroy(b);
  return rc;
}

}  // namespace

int main(int argc, char *argv[]) {
  if (argc != 2) {
    fprintf(stderr, "usage: rdcu_sweep MANIFEST\n"
                    "  every non-empty line of MANIFEST ('#' starts a comment) holds the\n"
                    "  arguments of one rdcu run (flags and FILEBASE); the runs are\n"
                    "  integrated side by side and each writes rdcu's files.\n");
    return 2;
  }
  FILE *mf = fopen(argv[1], "r");
  if (!mf) {
    fprintf(stderr, "rdcu_sweep: cannot open %s\n", argv[1]);
    return 1;
  }
  std::vector<std::unique_ptr<Instance>> inst;
  std::string line;
  int lineno = 0, c;
  bool eof = false;
  while (!eof) {
    line.clear();
    while ((c = fgetc(mf)) != EOF && c != '\n') line.push_back((char)c);
    eof = c == EOF;
    ++lineno;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::vector<std::string> tok = split_ws(line);
    if (tok.empty()) continue;
    tok.insert(tok.begin(), "rdcu");
    std::vector<char *> av;
    for (auto &s : tok) av.push_back(&s[0]);
    av.push_back(nullptr);
    inst.push_back(std::make_unique<Instance>());
    const int p = prepare((int)tok.size(), av.data(), *inst.back());
    if (p != PREP_OK) {
      fprintf(stderr, "rdcu_sweep: %s line %d: %s\n", argv[1], lineno,
              p == PREP_HELP ? "no FILEBASE" : "not a runnable rdcu command line");
      fclose(mf);
      return p == PREP_HELP ? 2 : p;
    }
  }
  fclose(mf);
  std::map<GroupKey, std::vector<Instance *>> groups;
  for (auto &I : inst)
    groups[GroupKey{I->fixed, I->gpu, I->atl, I->rtl, I->hmin, I->hmax}].push_back(I.get());
  int rc = 0;
  for (auto &kv : groups) rc |= run_batch(kv.second);
  release_frames();
  return rc;
}
#endif
