// rd_runtime.hip -- host runtime and C ABI (include/rdcu_dp45.h) of the
// MI355X DP45 pseudospectral integrator.
//
// One context = one problem on one GPU, one HIP stream (default priority).
// All DP45 decisions (accept/reject, step size, buffer rotation, FSAL) are
// made on the device by k_control, so a whole attempt is a fixed kernel
// sequence captured once as a hipGraph.  rdcu_run replays it in chunks that
// halt on the device at the next output time; the host reads the per-attempt
// log once per chunk and calls the step callback for each attempt, which is
// the role of dp45_run/step_eval (dp45.cu:154-169, rdcu.cu:252-341) without a
// host round trip per attempt (the reference blocks on cublasDnrm2 every
// attempt, dp45.cu:118).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "rd_kernels.h"
#include "rdcu_dp45.h"

using rd::cx;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHECK(expr)                                                                    \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(RDCU_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

// exp(-2 pi i m / L), exact at quarter turns (same table recipe as the CPU
// reference implementation, so both sides multiply by identical twiddles).
cx unit_root(long m, long L) {
  m %= L;
  if (m < 0) m += L;
  if ((4 * m) % L == 0) {
    switch ((4 * m) / L) {
      case 0: return {1.0, 0.0};
      case 1: return {0.0, -1.0};
      case 2: return {-1.0, 0.0};
      default: return {0.0, 1.0};
    }
  }
  long double x = 2.0L * M_PIl * (long double)m / (long double)L;
  return {(double)cosl(x), (double)(-sinl(x))};
}

bool make_plan(int L, std::vector<int> &rad) {
  rad.clear();
  if (L < 1) return false;
  int a = 0, b = 0, c = 0, m = L;
  while (m % 5 == 0) { m /= 5; ++c; }
  while (m % 3 == 0) { m /= 3; ++b; }
  while (m % 2 == 0) { m /= 2; ++a; }
  if (m != 1) return false;
  for (int i = 0; i < c; ++i) rad.push_back(5);
  for (int i = 0; i < b; ++i) rad.push_back(3);
  if (a % 3 == 1) rad.push_back(2);
  if (a % 3 == 2) rad.push_back(4);
  for (int i = 0; i < a / 3; ++i) rad.push_back(8);
  return true;
}

// DP45 coefficients (dp45.cu:6-21).
const double A2[1] = {1.0 / 5};
const double A3[2] = {3.0 / 40, 9.0 / 40};
const double A4[3] = {44.0 / 45, -56.0 / 15, 32.0 / 9};
const double A5[4] = {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729};
const double A6[5] = {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656};
const double B7[6] = {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84};
const double EE[7] = {-71.0 / 57600, 0, 71.0 / 16695, -71.0 / 1920, 17253.0 / 339200, -22.0 / 525,
                      1.0 / 40};
const double PP[7][4] = {
    {1.0, -8048581381.0 / 2820520608, 8663915743.0 / 2820520608, -12715105075.0 / 11282082432},
    {0, 0, 0, 0},
    {0, 131558114200.0 / 32700410799, -68118460800.0 / 10900136933, 87487479700.0 / 32700410799},
    {0, -1754552775.0 / 470086768, 14199869525.0 / 1410260304, -10690763975.0 / 1880347072},
    {0, 127303824393.0 / 49829197408, -318862633887.0 / 49829197408,
     701980252875.0 / 199316789632},
    {0, -282668133.0 / 205662961, 2019193451.0 / 616988883, -1453857185.0 / 822651844},
    {0, 40617522.0 / 29380423, -110615467.0 / 29380423, 69997945.0 / 29380423}};
const double *const STAGE_A[6] = {A2, A3, A4, A5, A6, B7};

const double kPI = 3.14159265358979323846;  // the reference's PI (rdcu.cu:14)
constexpr size_t kLdsMax = 64 * 1024;
constexpr int kChunkMax = 64;

}  // namespace

struct rdcu_ctx {
  int dev = 0;
  hipStream_t stream = nullptr;
  // problem
  int n = 0, ndim = 0, Ns[3] = {1, 1, 1};
  double Ls[3] = {1, 1, 1};
  long N = 0, rows = 0, Nh = 0;
  int Nx = 0, Nxh = 0, L = 0, nblk = 0;
  int nf_used = 0, nf_all = 0;
  rd::Grid grid{};
  rd::FftPlan px{}, p1{}, p0{};  // x, axis 1 (3D only), axis 0 (ndim >= 2)
  rd::cx *twN = nullptr;
  rd::Bufs bufs{};
  rd::Dict dict{};
  rd::FacList fl_used{}, fl_all{};
  rd::DevState *st = nullptr;
  rd::DevState *sh = nullptr;  // pinned host shadow
  double *rhs_in = nullptr, *rhs_out = nullptr, *Yreal = nullptr, *yeval = nullptr;
  std::vector<void *> allocs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // launch shapes
  int r2c_rpb = 1, r2c_thr = 64;
  int col_tc = 1, col_thr = 64;
  int c2r_grp = 1, c2r_grp_all = 1, c2r_thr = 64;
  size_t lds_r2c = 0, lds_col = 0, lds_c2r = 0, lds_c2r_all = 0;
  long nlines = 0, nred = 0;
  // host-side flags and counters
  int fixed = 0;
  bool k1_valid = false;
  long long nrhs = 0, nsync = 0;

  ~rdcu_ctx() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void *p : allocs) (void)hipFree(p);
    if (sh) (void)hipHostFree(sh);
    if (stream) (void)hipStreamDestroy(stream);
  }

  template <class T>
  int dalloc(T **p, size_t count) {
    void *q = nullptr;
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(RDCU_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(q);
    *p = static_cast<T *>(q);
    return RDCU_OK;
  }
  template <class T>
  int upload(T **p, const std::vector<T> &h) {
    int s = dalloc(p, h.size());
    if (s) return s;
    if (!h.empty()) HIPCHECK(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return RDCU_OK;
  }
  int make_fft(rd::FftPlan &p, int len) {
    std::vector<int> rad;
    if (!make_plan(len, rad) || (int)rad.size() > rd::kMaxPass)
      return fail(RDCU_EUNSUPPORTED, "transform length " + std::to_string(len) +
                                         " must have only the prime factors 2, 3, 5");
    p.L = len;
    p.np = (int)rad.size();
    for (int i = 0; i < p.np; ++i) p.rad[i] = rad[i];
    std::vector<cx> tw(len);
    for (int m = 0; m < len; ++m) tw[m] = unit_root(m, len);
    cx *d = nullptr;
    int s = upload(&d, tw);
    p.tw = d;
    return s;
  }
};

namespace {

// Enqueues one RHS evaluation.  stage 2..7: input = DP45 stage sum, output =
// stage slot; stage 0: input/output given by kind/pointer.
int enqueue_rhs(rdcu_ctx *c, int stage, int src_kind, const double *src_ptr, int dst_kind,
                int dst_slot, double *dst_ptr, int check, int err) {
  rd::StageCoef sc{};
  int rsrc = src_kind;
  if (stage >= 2) {
    sc.ns = stage - 1;
    for (int j = 0; j < sc.ns; ++j) sc.a[j] = STAGE_A[stage - 2][j];
    rsrc = rd::SRC_STAGE;
  }
  const long nrows = (long)c->n * c->rows;
  dim3 g1((unsigned)((nrows + c->r2c_rpb - 1) / c->r2c_rpb));
  hipLaunchKernelGGL(rd::k_row_r2c, g1, dim3(c->r2c_thr), c->lds_r2c, c->stream, c->grid, c->px,
                     c->twN, c->st, check, c->bufs, rsrc, src_ptr, sc, c->r2c_rpb);
  const int nd = c->ndim;
  const int ntc = (c->Nxh + c->col_tc - 1) / c->col_tc;
  if (nd == 3) {
    rd::ColGeom cg{c->Ns[1], (long)c->Nxh, c->Ns[0], c->Nh, (long)c->Ns[1] * c->Nxh, ntc,
                   c->col_tc, 1};
    dim3 gg((unsigned)((long)c->n * c->Ns[0] * ntc));
    hipLaunchKernelGGL(rd::k_col, gg, dim3(c->col_thr), c->lds_col, c->stream, c->grid, c->p1, cg,
                       c->st, check, c->bufs, c->fl_used, 0);
  }
  if (nd >= 2) {
    rd::ColGeom cg{};
    cg.L = c->Ns[0];
    cg.ntc = ntc;
    cg.tc = c->col_tc;
    cg.axis = 0;
    cg.gs1 = c->Nh;
    if (nd == 2) {
      cg.es = c->Nxh;
      cg.gdiv = 1;
      cg.gs2 = 0;
    } else {
      cg.es = (long)c->Ns[1] * c->Nxh;
      cg.gdiv = c->Ns[1];
      cg.gs2 = c->Nxh;
    }
    dim3 gg((unsigned)((long)c->n * cg.gdiv * ntc));
    hipLaunchKernelGGL(rd::k_col, gg, dim3(c->col_thr), c->lds_col, c->stream, c->grid, c->p0, cg,
                       c->st, check, c->bufs, c->fl_used, 1);
  }
  if (nd == 3 && c->nf_used > 0) {
    rd::ColGeom cg{c->Ns[1], (long)c->Nxh, c->Ns[0], c->Nh, (long)c->Ns[1] * c->Nxh, ntc,
                   c->col_tc, 1};
    dim3 gg((unsigned)((long)c->nf_used * c->Ns[0] * ntc));
    hipLaunchKernelGGL(rd::k_col, gg, dim3(c->col_thr), c->lds_col, c->stream, c->grid, c->p1, cg,
                       c->st, check, c->bufs, c->fl_used, 2);
  }
  rd::ErrCoef ec{};
  for (int j = 0; j < 7; ++j) ec.e[j] = EE[j];
  int csrc = stage >= 2 ? rd::SRC_YT : src_kind;
  hipLaunchKernelGGL(rd::k_row_c2r, dim3((unsigned)c->rows), dim3(c->c2r_thr), c->lds_c2r,
                     c->stream, c->grid, c->px, c->twN, c->st, check, c->bufs, c->fl_used, c->dict,
                     0, csrc, src_ptr, dst_kind, dst_slot, dst_ptr, err, ec, c->c2r_grp,
                     (double *)nullptr);
  HIPCHECK(hipGetLastError());
  c->nrhs++;
  return RDCU_OK;
}

// Every derivative block of every field, real space, into c->Yreal
// (S must hold the current forward spectra).
int enqueue_all_factors(rdcu_ctx *c) {
  if (c->nf_all == 0) return RDCU_OK;
  const int nd = c->ndim;
  const int ntc = (c->Nxh + c->col_tc - 1) / c->col_tc;
  if (nd >= 2) {
    rd::ColGeom cg{};
    cg.L = c->Ns[0];
    cg.ntc = ntc;
    cg.tc = c->col_tc;
    cg.gs1 = c->Nh;
    if (nd == 2) {
      cg.es = c->Nxh;
      cg.gdiv = 1;
    } else {
      cg.es = (long)c->Ns[1] * c->Nxh;
      cg.gdiv = c->Ns[1];
      cg.gs2 = c->Nxh;
    }
    dim3 gg((unsigned)((long)c->n * cg.gdiv * ntc));
    hipLaunchKernelGGL(rd::k_col, gg, dim3(c->col_thr), c->lds_col, c->stream, c->grid, c->p0, cg,
                       c->st, 0, c->bufs, c->fl_all, 1);
  }
  if (nd == 3) {
    rd::ColGeom cg{c->Ns[1], (long)c->Nxh, c->Ns[0], c->Nh, (long)c->Ns[1] * c->Nxh, ntc,
                   c->col_tc, 1};
    dim3 gg((unsigned)((long)c->nf_all * c->Ns[0] * ntc));
    hipLaunchKernelGGL(rd::k_col, gg, dim3(c->col_thr), c->lds_col, c->stream, c->grid, c->p1, cg,
                       c->st, 0, c->bufs, c->fl_all, 2);
  }
  rd::ErrCoef ec{};
  hipLaunchKernelGGL(rd::k_row_c2r, dim3((unsigned)c->rows), dim3(c->c2r_thr), c->lds_c2r_all,
                     c->stream, c->grid, c->px, c->twN, c->st, 0, c->bufs, c->fl_all, c->dict, 1,
                     (int)rd::SRC_PTR, (const double *)nullptr, (int)rd::DST_PTR, 0,
                     (double *)nullptr, 0, ec, c->c2r_grp_all, c->Yreal);
  HIPCHECK(hipGetLastError());
  return RDCU_OK;
}

int enqueue_attempt(rdcu_ctx *c) {
  for (int s = 2; s <= 7; ++s) {
    int dk = s < 7 ? rd::DST_KS : rd::DST_K7;
    int st = enqueue_rhs(c, s, rd::SRC_STAGE, nullptr, dk, s - 1, nullptr, 1, s == 7 && !c->fixed);
    if (st) return st;
  }
  const double *part = c->bufs.part;
  long cnt = c->nlines;
  if (!c->fixed && c->nlines > rd::kRedCap) {
    hipLaunchKernelGGL(rd::k_reduce, dim3((unsigned)c->nred), dim3(256), 0, c->stream,
                       c->bufs.part, c->nlines, c->bufs.part2, c->st, 1);
    part = c->bufs.part2;
    cnt = c->nred;
  }
  hipLaunchKernelGGL(rd::k_control, dim3(1), dim3(256), 0, c->stream, c->st, part, cnt,
                     (long)c->n * c->N);
  HIPCHECK(hipGetLastError());
  return RDCU_OK;
}

int read_state(rdcu_ctx *c) {
  HIPCHECK(hipMemcpyAsync(c->sh, c->st, sizeof(rd::DevState), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->nsync++;
  return RDCU_OK;
}

int begin(rdcu_ctx *c, double t_stop, int set_th, double t, double h, double hmin, double hmax,
          double t1) {
  hipLaunchKernelGGL(rd::k_begin, dim3(1), dim3(1), 0, c->stream, c->st, t_stop, set_th, t, h,
                     hmin, hmax, t1);
  HIPCHECK(hipGetLastError());
  return RDCU_OK;
}

int ensure_k1(rdcu_ctx *c) {
  if (c->k1_valid) return RDCU_OK;
  int s = enqueue_rhs(c, 0, rd::SRC_Y, nullptr, rd::DST_K1, 0, nullptr, 0, 0);
  if (s) return s;
  c->k1_valid = true;
  return RDCU_OK;
}

// Launch shapes: rows per block and threads for the x passes, column tile
// width for the strided passes, factor group size for the inverse x pass.
int choose_shapes(rdcu_ctx *c) {
  const int L = c->L, Nx = c->Nx, n = c->n;
  auto round64 = [](long v) { return (int)std::max<long>(64, std::min<long>(256, (v + 63) / 64 * 64)); };
  c->r2c_rpb = std::max(1, std::min(8, 256 / std::max(1, L)));
  c->r2c_thr = round64((long)c->r2c_rpb * L / 4);
  c->lds_r2c = (size_t)2 * c->r2c_rpb * L * sizeof(cx);
  if (c->lds_r2c > kLdsMax) return fail(RDCU_EUNSUPPORTED, "x axis too long");
  if (c->ndim >= 2) {
    int Lmax = c->Ns[0];
    if (c->ndim == 3) Lmax = std::max(Lmax, c->Ns[1]);
    int tc = 1;
    while (tc * 2 <= 16 && (size_t)3 * (tc * 2) * Lmax * sizeof(cx) <= 48 * 1024) tc *= 2;
    if ((size_t)3 * tc * Lmax * sizeof(cx) > kLdsMax)
      return fail(RDCU_EUNSUPPORTED, "strided axis too long");
    c->col_tc = tc;
    c->col_thr = round64((long)tc * Lmax / 4);
    c->lds_col = (size_t)3 * tc * Lmax * sizeof(cx);
  }
  auto c2r_lds = [&](int nf, int grp) {
    size_t fft = (size_t)2 * grp * L * sizeof(cx);
    size_t sq = (size_t)n * Nx * sizeof(double);
    return (size_t)(n + nf) * Nx * sizeof(double) + std::max(fft, sq);
  };
  int grp = std::max(1, c->nf_used);
  while (grp > 1 && c2r_lds(c->nf_used, grp) > kLdsMax) --grp;
  if (c2r_lds(c->nf_used, grp) > kLdsMax)
    return fail(RDCU_EUNSUPPORTED, "x axis too long for the number of factors");
  c->c2r_grp = grp;
  c->lds_c2r = c2r_lds(c->nf_used, grp);
  c->c2r_thr = round64((long)Nx / 2);
  // store mode keeps no value rows: only the FFT scratch of its groups
  int ga = std::max(1, c->nf_all);
  while (ga > 1 && (size_t)2 * ga * L * sizeof(cx) > kLdsMax) --ga;
  c->c2r_grp_all = ga;
  c->lds_c2r_all = (size_t)2 * ga * L * sizeof(cx);
  return RDCU_OK;
}

}  // namespace

extern "C" {

const char *rdcu_last_error(void) { return g_err.c_str(); }
const char *rdcu_backend(void) { return "hip-gfx950"; }

int rdcu_create(const rdcu_problem *p, double atol, double rtol, int fixedstep, const double *y0,
                int device, rdcu_ctx **out) {
  if (!out) return fail(RDCU_EINVAL, "null output handle");
  *out = nullptr;
  if (!p) return fail(RDCU_EINVAL, "null problem");
  if (p->n < 1) return fail(RDCU_EINVAL, "n must be >= 1");
  if (p->n > rd::kNMax) return fail(RDCU_EUNSUPPORTED, "at most 8 fields are supported");
  if (p->ndim < 1 || p->ndim > 3) return fail(RDCU_EINVAL, "ndim must be 1, 2 or 3");
  std::unique_ptr<rdcu_ctx> c(new rdcu_ctx());
  c->n = p->n;
  c->ndim = p->ndim;
  c->N = 1;
  for (int a = 0; a < 3; ++a) {
    c->Ns[a] = a < p->ndim ? p->Ns[a] : 1;
    c->Ls[a] = a < p->ndim ? p->Ls[a] : 1.0;
  }
  for (int a = 0; a < p->ndim; ++a) {
    if (c->Ns[a] < 2 || c->Ns[a] % 2) return fail(RDCU_EUNSUPPORTED, "grid sizes must be even");
    c->N *= c->Ns[a];
  }
  c->Nx = c->Ns[p->ndim - 1];
  c->L = c->Nx / 2;
  c->Nxh = c->L + 1;
  c->rows = c->N / c->Nx;
  c->Nh = c->rows * c->Nxh;
  c->nblk = 1 + c->ndim + c->ndim * c->ndim;
  c->fixed = fixedstep ? 1 : 0;

  // dictionary and factor lists
  if (p->nterms < 0 || (p->nterms > 0 && (!p->coef || !p->field || !p->nfac || !p->fac)))
    return fail(RDCU_EINVAL, "bad dictionary arrays");
  const int nidx = c->n * c->nblk;
  std::vector<char> seen(nidx, 0);
  {
    long off = 0;
    for (int i = 0; i < p->nterms; ++i) {
      if (p->field[i] < 0 || p->field[i] >= c->n) return fail(RDCU_EINVAL, "term field out of range");
      if (p->nfac[i] < 0) return fail(RDCU_EINVAL, "negative factor count");
      for (int j = 0; j < p->nfac[i]; ++j) {
        int ix = p->fac[off + 2 * j];
        if (ix < 0 || ix >= nidx) return fail(RDCU_EINVAL, "factor index out of range");
        seen[ix] = 1;
      }
      off += 2 * p->nfac[i];
    }
  }
  std::vector<int> ublk, ufld, ustart(c->n + 1, 0), ablk, afld, astart(c->n + 1, 0);
  std::vector<int> slot_of(nidx, -1);
  for (int f = 0; f < c->n; ++f) {
    ustart[f] = (int)ublk.size();
    astart[f] = (int)ablk.size();
    for (int b = 1; b < c->nblk; ++b) {
      ablk.push_back(b);
      afld.push_back(f);
      if (seen[b * c->n + f]) {
        slot_of[b * c->n + f] = c->n + (int)ublk.size();
        ublk.push_back(b);
        ufld.push_back(f);
      }
    }
  }
  ustart[c->n] = (int)ublk.size();
  astart[c->n] = (int)ablk.size();
  for (int f = 0; f < c->n; ++f) slot_of[f] = f;
  c->nf_used = (int)ublk.size();
  c->nf_all = (int)ablk.size();
  std::vector<int> tfield, tstart, fslot, fpow;
  std::vector<double> tcoef;
  {
    long off = 0;
    for (int i = 0; i < p->nterms; ++i) {
      tfield.push_back(p->field[i]);
      tcoef.push_back(p->coef[i]);
      tstart.push_back((int)fslot.size());
      for (int j = 0; j < p->nfac[i]; ++j) {
        fslot.push_back(slot_of[p->fac[off + 2 * j]]);
        fpow.push_back(p->fac[off + 2 * j + 1]);
      }
      off += 2 * p->nfac[i];
    }
    tstart.push_back((int)fslot.size());
  }

  int st = RDCU_OK;
  HIPCHECK(hipSetDevice(device));
  c->dev = device;
  HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));

  // transform plans and tables
  if ((st = c->make_fft(c->px, c->L))) return st;
  if (c->ndim >= 2 && (st = c->make_fft(c->p0, c->Ns[0]))) return st;
  if (c->ndim == 3 && (st = c->make_fft(c->p1, c->Ns[1]))) return st;
  {
    std::vector<cx> twN(c->Nx);
    for (int m = 0; m < c->Nx; ++m) twN[m] = unit_root(m, c->Nx);
    if ((st = c->upload(&c->twN, twN))) return st;
  }
  rd::Grid &g = c->grid;
  g.n = c->n;
  g.ndim = c->ndim;
  for (int a = 0; a < 3; ++a) g.Ns[a] = c->Ns[a];
  g.Nx = c->Nx;
  g.Nxh = c->Nxh;
  g.L = c->L;
  g.N = c->N;
  g.rows = c->rows;
  g.Nh = c->Nh;
  g.invN = 1.0 / (double)c->N;
  for (int a = 0; a < c->ndim; ++a) {
    const int Na = c->Ns[a];
    std::vector<double> kt(Na);
    for (int na = 0; na < Na; ++na) {
      int nn = na < Na / 2 ? na : na - Na;
      kt[na] = ((double)(2 * nn) * kPI) / c->Ls[a];
    }
    double *d = nullptr;
    if ((st = c->upload(&d, kt))) return st;
    g.kt[a] = d;
  }
  // dictionary and factor lists on the device
  {
    int *d_tf, *d_ts, *d_fs, *d_fp;
    double *d_tc;
    if ((st = c->upload(&d_tf, tfield)) || (st = c->upload(&d_tc, tcoef)) ||
        (st = c->upload(&d_ts, tstart)) || (st = c->upload(&d_fs, fslot)) ||
        (st = c->upload(&d_fp, fpow)))
      return st;
    c->dict = rd::Dict{p->nterms, d_tf, d_tc, d_ts, d_fs, d_fp};
    int *ub, *uf, *us, *ab, *af, *as;
    if ((st = c->upload(&ub, ublk)) || (st = c->upload(&uf, ufld)) || (st = c->upload(&us, ustart)) ||
        (st = c->upload(&ab, ablk)) || (st = c->upload(&af, afld)) || (st = c->upload(&as, astart)))
      return st;
    c->fl_used = rd::FacList{c->nf_used, ub, uf, us};
    c->fl_all = rd::FacList{c->nf_all, ab, af, as};
  }
  if ((st = choose_shapes(c.get()))) return st;

  // buffers
  const size_t nN = (size_t)c->n * c->N;
  for (int i = 0; i < 3; ++i)
    if ((st = c->dalloc(&c->bufs.y[i], nN))) return st;
  for (int i = 0; i < 7; ++i)
    if ((st = c->dalloc(&c->bufs.k[i], nN))) return st;
  if ((st = c->dalloc(&c->bufs.S, (size_t)c->n * c->Nh))) return st;
  if ((st = c->dalloc(&c->bufs.F, c->ndim >= 2 ? (size_t)c->nf_all * c->Nh : 1))) return st;
  c->nlines = (long)c->n * c->rows;
  c->nred = (c->nlines + rd::kRedCap - 1) / rd::kRedCap;
  if ((st = c->dalloc(&c->bufs.part, (size_t)c->nlines))) return st;
  if ((st = c->dalloc(&c->bufs.part2, (size_t)c->nred))) return st;
  if ((st = c->dalloc(&c->rhs_in, nN)) || (st = c->dalloc(&c->rhs_out, nN)) ||
      (st = c->dalloc(&c->yeval, nN)))
    return st;
  if (y0) HIPCHECK(hipMemcpy(c->bufs.y[0], y0, nN * sizeof(double), hipMemcpyHostToDevice));
  else HIPCHECK(hipMemset(c->bufs.y[0], 0, nN * sizeof(double)));

  // device state
  {
    void *q = nullptr;
    HIPCHECK(hipHostMalloc(&q, sizeof(rd::DevState), hipHostMallocDefault));
    c->sh = static_cast<rd::DevState *>(q);
    std::memset(c->sh, 0, sizeof(rd::DevState));
    c->sh->iy = 0;
    c->sh->iyt = 1;
    c->sh->iyl = 2;
    c->sh->k1s = 0;
    c->sh->k7s = 6;
    c->sh->atol = atol;
    c->sh->rtol = rtol;
    c->sh->fixed = c->fixed;
    c->sh->hmax = INFINITY;
    c->sh->t1 = INFINITY;
    c->sh->hmin = 0;
    c->sh->t_stop = INFINITY;
    c->sh->halted = 1;
    if ((st = c->dalloc(&c->st, 1))) return st;
    HIPCHECK(hipMemcpy(c->st, c->sh, sizeof(rd::DevState), hipMemcpyHostToDevice));
  }

  // capture one attempt
  HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  st = enqueue_attempt(c.get());
  hipGraph_t graph = nullptr;
  hipError_t ce = hipStreamEndCapture(c->stream, &graph);
  if (st) return st;
  if (ce != hipSuccess) return fail(RDCU_EDEVICE, std::string("capture: ") + hipGetErrorString(ce));
  c->graph = graph;
  HIPCHECK(hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0));
  c->nrhs = 0;
  *out = c.release();
  return RDCU_OK;
}

int rdcu_destroy(rdcu_ctx *c) {
  if (c) {
    (void)hipSetDevice(c->dev);
    (void)hipStreamSynchronize(c->stream);
  }
  delete c;
  return RDCU_OK;
}

int rdcu_step(rdcu_ctx *c, double *t, double *h, double hmin, double hmax, int *accepted) {
  (void)hmin;  // dp45_step does not test hmin; only the run loop does
  if (!c || !t || !h || !accepted) return fail(RDCU_EINVAL, "null argument");
  HIPCHECK(hipSetDevice(c->dev));
  int st;
  if ((st = begin(c, INFINITY, 1, *t, *h, -INFINITY, hmax, INFINITY))) return st;
  if ((st = ensure_k1(c))) return st;
  HIPCHECK(hipGraphLaunch(c->gexec, c->stream));
  c->nrhs += 6;
  if ((st = read_state(c))) return st;
  if (c->sh->nlog != 1) return fail(RDCU_ESTATE, "attempt did not run");
  *accepted = c->sh->log_acc[0];
  *t = c->sh->t;
  *h = c->sh->h;
  return RDCU_OK;
}

int rdcu_run(rdcu_ctx *c, double *t, double *h, double hmin, double hmax, double t1, double stop,
             rdcu_step_fn fn, void *user) {
  if (!c || !t || !h) return fail(RDCU_EINVAL, "null argument");
  HIPCHECK(hipSetDevice(c->dev));
  int st;
  if ((st = begin(c, stop, 1, *t, *h, hmin, hmax, t1))) return st;
  c->k1_valid = false;  // dp45_run recomputes k1 = f(y) (dp45.cu:157)
  if ((st = ensure_k1(c))) return st;
  double tc = *t, hc = *h;
  while (tc < t1 && hc > hmin) {
    double te = std::min(stop, t1);
    long K = 1;
    if (te > tc && hc > 0) {
      double est = std::ceil((te - tc) / hc) + 1.0;
      K = est > kChunkMax ? kChunkMax : (long)est;
      if (K < 1) K = 1;
    }
    if ((st = begin(c, stop, 0, 0, 0, hmin, hmax, t1))) return st;
    for (long i = 0; i < K; ++i) HIPCHECK(hipGraphLaunch(c->gexec, c->stream));
    if ((st = read_state(c))) return st;
    const int nl = c->sh->nlog;
    c->nrhs += 6L * nl;
    if (nl == 0) break;
    for (int i = 0; i < nl; ++i) {
      tc = c->sh->log_t[i];
      hc = c->sh->log_h[i];
      *t = tc;
      *h = hc;
      if (fn && fn(c, tc, hc, c->sh->log_acc[i], i == nl - 1, &stop, user))
        return fail(RDCU_ECALLBACK, "callback stop");
    }
    tc = c->sh->t;
    hc = c->sh->h;
  }
  *t = c->sh->t;
  *h = c->sh->h;
  return RDCU_OK;
}

int rdcu_eval(rdcu_ctx *c, double t_eval, double *y_out) {
  if (!c || !y_out) return fail(RDCU_EINVAL, "null argument");
  if (!c->sh->eval_valid) return fail(RDCU_ESTATE, "no accepted step to interpolate");
  HIPCHECK(hipSetDevice(c->dev));
  rd::DenseCoef dc;
  for (int j = 0; j < 7; ++j)
    for (int q = 0; q < 4; ++q) dc.p[j][q] = PP[j][q];
  const long nN = (long)c->n * c->N;
  hipLaunchKernelGGL(rd::k_dense, dim3((unsigned)((nN + 255) / 256)), dim3(256), 0, c->stream,
                     c->st, c->bufs, dc, t_eval, c->yeval, nN);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(y_out, c->yeval, nN * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return RDCU_OK;
}

int rdcu_get_state(rdcu_ctx *c, double *y_out) {
  if (!c || !y_out) return fail(RDCU_EINVAL, "null argument");
  HIPCHECK(hipSetDevice(c->dev));
  const long nN = (long)c->n * c->N;
  HIPCHECK(hipMemcpyAsync(y_out, c->bufs.y[c->sh->iy], nN * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return RDCU_OK;
}

int rdcu_set_state(rdcu_ctx *c, const double *y) {
  if (!c || !y) return fail(RDCU_EINVAL, "null argument");
  HIPCHECK(hipSetDevice(c->dev));
  const long nN = (long)c->n * c->N;
  HIPCHECK(hipMemcpyAsync(c->bufs.y[c->sh->iy], y, nN * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->k1_valid = false;
  c->sh->eval_valid = 0;
  return RDCU_OK;
}

int rdcu_rhs(rdcu_ctx *c, const double *y, double *f_out, double *Y_out) {
  if (!c || !y) return fail(RDCU_EINVAL, "null argument");
  HIPCHECK(hipSetDevice(c->dev));
  const long N = c->N, nN = (long)c->n * N;
  HIPCHECK(hipMemcpyAsync(c->rhs_in, y, nN * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int st = enqueue_rhs(c, 0, rd::SRC_PTR, c->rhs_in, rd::DST_PTR, 0, c->rhs_out, 0, 0);
  if (st) return st;
  if (Y_out) {
    if (!c->Yreal && (st = c->dalloc(&c->Yreal, (size_t)std::max(1, c->nf_all) * N))) return st;
    if ((st = enqueue_all_factors(c))) return st;
  }
  if (f_out)
    HIPCHECK(hipMemcpyAsync(f_out, c->rhs_out, nN * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  std::vector<double> yr;
  if (Y_out && c->nf_all) {
    yr.resize((size_t)c->nf_all * N);
    HIPCHECK(hipMemcpyAsync(yr.data(), c->Yreal, yr.size() * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (Y_out) {
    for (int blk = 0; blk < c->nblk; ++blk)
      for (int f = 0; f < c->n; ++f) {
        double *dst = Y_out + 2 * ((size_t)(blk * c->n + f) * N);
        // factor list "all" is field-major: factor of (f, blk) is f*(nblk-1) + blk-1
        const double *src = blk == 0 ? y + (size_t)f * N
                                     : yr.data() + (size_t)(f * (c->nblk - 1) + blk - 1) * N;
        for (long x = 0; x < N; ++x) {
          dst[2 * x] = src[x];
          dst[2 * x + 1] = 0.0;
        }
      }
  }
  return RDCU_OK;
}

int rdcu_get_stats(rdcu_ctx *c, long long out[4]) {
  if (!c || !out) return fail(RDCU_EINVAL, "null argument");
  out[0] = c->sh->attempts;
  out[1] = c->sh->accepted;
  out[2] = c->nrhs;
  out[3] = c->nsync;
  return RDCU_OK;
}

}  // extern "C"
