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
#include <hip/hiprtc.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "rd_kernels.h"
#include "rd_kernels_reg.h"
#include "rd_adjoint.h"
#include "rdcu_dp45.h"

#include "rd_jit_src.inc"

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
constexpr int kGraphAttempts = 8;  // attempts per multi-attempt graph; divides kChunkMax
static_assert(kGraphAttempts % 2 == 0, "reverse-step groups alternate the two adjoint buffers");

// Device scratch of rdcu_rhs_noalloc: one array per GPU for the life of the
// process, so that the call allocates nothing.  Host code never uses the
// array's name as an address (on the host it is only a stand-in); the
// device address comes from hipGetSymbolAddress, once per device.
constexpr size_t kScratchBytes = size_t(128) << 20;
__device__ __align__(256) char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_call_mu[kMaxDevices];

size_t round_up256(size_t b) { return (b + 255) / 256 * 256; }

// Device address of this GPU's scratch array (the device must be current).
int scratch_base(int dev, char **out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(RDCU_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(RDCU_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<char *>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return RDCU_OK;
}

// Calls on one GPU share its array: they are serialised within the process.
std::mutex &scratch_mutex(int dev) { return g_scratch_call_mu[dev]; }

// 3-D: a factor list split into the factors without k_z (zf: multiplied in
// the y pass) and the others (zd: through the z pass and the inverse y pass),
// each with the slots of its factors in the full list.
struct ZSplit {
  rd::FacList zf{}, zd{};
  int zf_max = 0, zd_max = 0;  // most factors of one field
};

}  // namespace

struct rdcu_ctx {
  int dev = 0;
  hipStream_t stream = nullptr;
  // problem
  int n = 0, ndim = 0, Ns[3] = {1, 1, 1};
  double Ls[3] = {1, 1, 1};
  long N = 0, rows = 0, Nh = 0;
  int Nx = 0, Nxh = 0, Nxp = 0, L = 0, nblk = 0;
  int nf_used = 0, nf_all = 0;
  rd::Grid grid{};
  rd::FftPlan px{}, p1{}, p0{};  // x, axis 1 (3D only), axis 0 (ndim >= 2)
  rd::cx *twN = nullptr;
  rd::Bufs bufs{};
  rd::Dict dict{};
  rd::DictP dictp{};
  bool dictk_ok = false;
  // x-inverse kernel with the dictionary compiled in (hiprtc); null if not built
  hipModule_t jit_mod = nullptr;
  hipFunction_t jit_rowi = nullptr;
  hipFunction_t jit_rowi_tw = nullptr;  // the same with the twiddles in registers
  // both compiled lean (fewer values in flight) for two wavefronts per SIMD
  hipFunction_t jit_rowi_lean = nullptr, jit_rowi_lean_tw = nullptr;
  int ri_tw = 0;  // the tuner's choice: 0 plain, 1 tw, 2 lean, 3 lean tw
  // 1-D: a whole attempt in one workgroup (null if not built or not eligible)
  hipFunction_t jit_att = nullptr;
  size_t lds_att = 0;
  // gradients: the vector-Jacobian row kernel with the dictionary compiled in
  hipModule_t jit_vjp_mod = nullptr;
  hipFunction_t jit_vjp = nullptr;
  std::string jit_log;
  std::vector<int> h_tfield, h_tstart, h_fslot, h_fpow;
  std::vector<double> h_tcoef;
  int col_lds = 0;
  rd::FacList fl_used{}, fl_all{};
  rd::FacList fl_imag{};  // fl_all with the imaginary-part codes (Y output)
  // 3-D splits of fl_used, fl_all and fl_imag, and of the adjoint's list for
  // seven stage inputs side by side (zs7: built at its first use)
  ZSplit zs_used, zs_all, zs_imag, zs7;
  rd::DevState *st = nullptr;
  rd::DevState *sh = nullptr;  // pinned host shadow
  double *rhs_in = nullptr, *rhs_out = nullptr, *Yreal = nullptr, *yeval = nullptr;
  double *Ycx = nullptr;  // factor blocks as complex values (FILEBASEY.dat layout)
  std::vector<void *> allocs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // kGraphAttempts attempts in one graph: fewer graph launches, and no gap
  // between the graphs of consecutive attempts inside it
  hipGraph_t graph_m = nullptr;
  hipGraphExec_t gexec_m = nullptr;
  // launch shapes
  int r2c_rpb = 1, r2c_thr = 64;
  int col_tc = 1, col_thr = 64;
  int c2r_grp = 1, c2r_grp_all = 1, c2r_thr = 64;
  size_t lds_r2c = 0, lds_col = 0, lds_c2r = 0, lds_c2r_all = 0;
  long nlines = 0, nred = 0;
  // register-resident path: elements per thread for the x axis and the
  // strided axes (0 = not eligible); shapes picked by the autotuner
  int reg = 0, Ex = 0, E0 = 0, E1 = 0;
  int rf_rpb = 1, cr_tc = 1, ri_G = 1, ri_R = 1, cr_fpw = 1, max_fpf = 1;
  int cr_ntc = 0;  // strided tiles of even widths per column group (0: tiles of cr_tc)
  size_t lds_rf = 0, lds_cr0 = 0, lds_cr1 = 0, lds_ri = 0;
  bool lds_rhs = true, lds_y = true;  // the LDS kernels fit (see choose_shapes)
  std::string tune_log;
  // host-side flags and counters
  int fixed = 0;
  bool k1_valid = false;
  long long nrhs = 0, nsync = 0;
  // member of a batch: stream and device state belong to the batch, and
  // attempts run only through the batch
  bool member = false;
  bool borrowed_stream = false;  // stream owned by another context
  std::vector<int> h_ublk, h_ufld, h_prog;  // referenced factors; dictionary program
  // adjoint: companion problem applying the derivative blocks to slot weights
  std::unique_ptr<rdcu_ctx> adj_pc;
  // right-hand side given as a callback (rdcu_create_fn): no dictionary
  rdcu_dydt_fn fn = nullptr;
  void *fn_user = nullptr;
  int fn_nchunk = 1;

  ~rdcu_ctx() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (gexec_m) (void)hipGraphExecDestroy(gexec_m);
    if (graph_m) (void)hipGraphDestroy(graph_m);
    if (jit_mod) (void)hipModuleUnload(jit_mod);
    if (jit_vjp_mod) (void)hipModuleUnload(jit_vjp_mod);
    if (graph) (void)hipGraphDestroy(graph);
    for (void *p : allocs) (void)hipFree(p);
    adj_pc.reset();
    if (sh && !member) (void)hipHostFree(sh);
    if (stream && !member && !borrowed_stream) (void)hipStreamDestroy(stream);
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
    p.code = 0;
    for (int i = 0; i < p.np; ++i) p.code |= (unsigned long long)rad[i] << (4 * i);
    std::vector<cx> tw(len);
    for (int m = 0; m < len; ++m) tw[m] = unit_root(m, len);
    cx *d = nullptr;
    int s = upload(&d, tw);
    p.tw = d;
    return s;
  }
};

namespace {

// Splits the factor list (blocks blk, field f's factors [start[f],
// start[f+1])) into zf (the factors without k_z) and zd (the others) and
// uploads both halves.
int build_zsplit(rdcu_ctx *c, int nfields, const std::vector<int> &blk,
                 const std::vector<int> &start, ZSplit &z) {
  std::vector<int> b[2], f[2], sl[2], st[2];
  for (int k = 0; k < 2; ++k) st[k].assign(nfields + 1, 0);
  for (int fd = 0; fd < nfields; ++fd) {
    for (int k = 0; k < 2; ++k) st[k][fd] = (int)b[k].size();
    for (int m = start[fd]; m < start[fd + 1]; ++m) {
      const int k = rd::zfree_blk(blk[m], c->ndim) ? 0 : 1;
      b[k].push_back(blk[m]);
      f[k].push_back(fd);
      sl[k].push_back(m);
    }
  }
  rd::FacList *out[2] = {&z.zf, &z.zd};
  int *mx[2] = {&z.zf_max, &z.zd_max};
  for (int k = 0; k < 2; ++k) {
    st[k][nfields] = (int)b[k].size();
    int *db = nullptr, *df = nullptr, *ds = nullptr, *dst = nullptr, s;
    if ((s = c->upload(&db, b[k])) || (s = c->upload(&df, f[k])) || (s = c->upload(&ds, sl[k])) ||
        (s = c->upload(&dst, st[k])))
      return s;
    *out[k] = rd::FacList{(int)b[k].size(), db, df, dst, ds};
    *mx[k] = 0;
    for (int fd = 0; fd < nfields; ++fd) *mx[k] = std::max(*mx[k], st[k][fd + 1] - st[k][fd]);
  }
  return RDCU_OK;
}

// Kernel launches of one RHS evaluation, in order.  stage 2..7: input =
// DP45 stage sum, output = stage slot; stage 0: input/output given by
// kind/pointer.
struct Launch {
  const char *name;
  std::function<void()> go;
};

std::vector<Launch> rhs_launches(rdcu_ctx *c, int stage, int src_kind, const double *src_ptr,
                                 int dst_kind, int dst_slot, double *dst_ptr, int check, int err) {
  std::vector<Launch> v;
  rd::StageCoef sc{};
  int rsrc = src_kind;
  if (stage >= 2) {
    sc.ns = stage - 1;
    for (int j = 0; j < sc.ns; ++j) sc.a[j] = STAGE_A[stage - 2][j];
    rsrc = rd::SRC_STAGE;
  }
  const long nrows = (long)c->n * c->rows;
  dim3 g1((unsigned)((nrows + c->r2c_rpb - 1) / c->r2c_rpb));
  v.push_back({"row_r2c", [=]() {
                 hipLaunchKernelGGL(rd::k_row_r2c, g1, dim3(c->r2c_thr), c->lds_r2c, c->stream,
                                    c->grid, c->px, c->twN, c->st, check, c->bufs, rsrc, src_ptr,
                                    sc, c->r2c_rpb);
               }});
  const int nd = c->ndim;
  const int ntc = (c->Nxh + c->col_tc - 1) / c->col_tc;
  // 3-D: the factors without k_z in the y forward pass (mode 3), the others
  // through the z pass and the inverse y pass
  const ZSplit &zs = c->zs_used;
  if (nd == 3) {
    rd::ColGeom cg{c->Ns[1], (long)c->Nxp, c->Ns[0], c->Nh, (long)c->Ns[1] * c->Nxp, ntc,
                   c->col_tc, 1};
    dim3 gg((unsigned)((long)c->n * c->Ns[0] * ntc));
    const bool zf = zs.zf.nf > 0;
    v.push_back({"col_fwd", [=]() {
                   hipLaunchKernelGGL(rd::k_col, gg, dim3(c->col_thr), c->lds_col, c->stream,
                                      c->grid, c->p1, cg, c->st, check, c->bufs,
                                      zf ? zs.zf : c->fl_used, zf ? 3 : 0);
                 }});
  }
  if (nd == 2 || (nd == 3 && zs.zd.nf > 0)) {
    rd::ColGeom cg{};
    cg.L = c->Ns[0];
    cg.ntc = ntc;
    cg.tc = c->col_tc;
    cg.axis = 0;
    cg.gs1 = c->Nh;
    if (nd == 2) {
      cg.es = c->Nxp;
      cg.gdiv = 1;
      cg.gs2 = 0;
    } else {
      cg.es = (long)c->Ns[1] * c->Nxp;
      cg.gdiv = c->Ns[1];
      cg.gs2 = c->Nxp;
    }
    dim3 gg((unsigned)((long)c->n * cg.gdiv * ntc));
    const rd::FacList fl = nd == 3 ? zs.zd : c->fl_used;
    v.push_back({"col_mult", [=]() {
                   hipLaunchKernelGGL(rd::k_col, gg, dim3(c->col_thr), c->lds_col, c->stream,
                                      c->grid, c->p0, cg, c->st, check, c->bufs, fl, 1);
                 }});
  }
  if (nd == 3 && zs.zd.nf > 0) {
    rd::ColGeom cg{c->Ns[1], (long)c->Nxp, c->Ns[0], c->Nh, (long)c->Ns[1] * c->Nxp, ntc,
                   c->col_tc, 1};
    dim3 gg((unsigned)((long)zs.zd.nf * c->Ns[0] * ntc));
    v.push_back({"col_inv", [=]() {
                   hipLaunchKernelGGL(rd::k_col, gg, dim3(c->col_thr), c->lds_col, c->stream,
                                      c->grid, c->p1, cg, c->st, check, c->bufs, zs.zd, 2);
                 }});
  }
  rd::ErrCoef ec{};
  for (int j = 0; j < 7; ++j) ec.e[j] = EE[j];
  int csrc = stage >= 2 ? rd::SRC_YT : src_kind;
  v.push_back({"row_c2r", [=]() {
                 hipLaunchKernelGGL(rd::k_row_c2r, dim3((unsigned)c->rows), dim3(c->c2r_thr),
                                    c->lds_c2r, c->stream, c->grid, c->px, c->twN, c->st, check,
                                    c->bufs, c->fl_used, c->dict, 0, csrc, src_ptr, dst_kind,
                                    dst_slot, dst_ptr, err, ec, c->c2r_grp, (double *)nullptr);
               }});
  return v;
}

std::vector<Launch> rhs_launches_reg(rdcu_ctx *c, int src_kind, const double *src_ptr, int dst_kind,
                                     int dst_slot, double *dst_ptr, int check);

int enqueue_rhs(rdcu_ctx *c, int stage, int src_kind, const double *src_ptr, int dst_kind,
                int dst_slot, double *dst_ptr, int check, int err) {
  if (c->reg && stage == 0 && !err) {
    for (auto &l : rhs_launches_reg(c, src_kind, src_ptr, dst_kind, dst_slot, dst_ptr, check))
      l.go();
    HIPCHECK(hipGetLastError());
    c->nrhs++;
    return RDCU_OK;
  }
  for (auto &l : rhs_launches(c, stage, src_kind, src_ptr, dst_kind, dst_slot, dst_ptr, check, err))
    l.go();
  HIPCHECK(hipGetLastError());
  c->nrhs++;
  return RDCU_OK;
}

// ---- register-resident path ------------------------------------------------

// Elements per thread of the register Stockham for a length (0 = use the
// LDS path): 8 for powers of two >= 8 and for 3*2^k >= 24 (whose leading
// radix-3 pass runs through the line in LDS, lds_pass).
int reg_elems(int L) {
  auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
  if (L >= 8 && pow2(L)) return 8;
  if (L >= 24 && L % 3 == 0 && pow2(L / 3)) return 8;
  return 0;
}

// Transform lengths with kernels compiled for that length (constant radix
// plan and index arithmetic); other lengths run the run-time-plan kernels.
// x lines: the half lengths of 256- and 512-point rows; strided lines: 256
// and 512 points.
constexpr int kRowLs[] = {128, 256};
constexpr int kColLs[] = {256, 512};
// with 8 values per thread also the 3*2^k lengths of 96- and 192-point grids
constexpr int kRowLs3[] = {48, 96};
constexpr int kColLs3[] = {96, 192};

template <int E, bool NT, int LS>
void go_rowf(rdcu_ctx *c, dim3 g, dim3 b, size_t lds, int check, int src_kind,
             const double *src_ptr, const rd::StageCoef &sc, int rpb) {
  hipLaunchKernelGGL((rd::k_rowf_reg<E, NT, LS>), g, b, lds, c->stream, c->grid, c->px, c->twN,
                     c->st, check, c->bufs, src_kind, src_ptr, sc, rpb);
}

// Lengths 3*2^k run the run-time-plan kernels instantiated with LS = -1
// (their radix-3 pass through LDS); powers of two those with LS = 0.
bool odd_len(int L) { return L % 3 == 0; }

template <int E, bool NT>
void go_rowf_l(rdcu_ctx *c, dim3 g, dim3 b, size_t lds, int check, int src_kind,
               const double *src_ptr, const rd::StageCoef &sc, int rpb) {
  if constexpr (E == 8) {
    if (c->L == kRowLs[0]) return go_rowf<E, NT, kRowLs[0]>(c, g, b, lds, check, src_kind, src_ptr, sc, rpb);
    if (c->L == kRowLs[1]) return go_rowf<E, NT, kRowLs[1]>(c, g, b, lds, check, src_kind, src_ptr, sc, rpb);
    if (c->L == kRowLs3[0]) return go_rowf<E, NT, kRowLs3[0]>(c, g, b, lds, check, src_kind, src_ptr, sc, rpb);
    if (c->L == kRowLs3[1]) return go_rowf<E, NT, kRowLs3[1]>(c, g, b, lds, check, src_kind, src_ptr, sc, rpb);
  }
  if (odd_len(c->L)) return go_rowf<E, NT, -1>(c, g, b, lds, check, src_kind, src_ptr, sc, rpb);
  go_rowf<E, NT, 0>(c, g, b, lds, check, src_kind, src_ptr, sc, rpb);
}

template <int E>
void launch_rowf(rdcu_ctx *c, int check, int src_kind, const double *src_ptr, rd::StageCoef sc,
                 int rpb) {
  const long nrows = (long)c->n * c->rows;
  const int T = c->L / E;
  dim3 g((unsigned)((nrows + rpb - 1) / rpb));
  size_t lds = (size_t)(c->L + rpb * rd::padded_len(c->L)) * sizeof(cx);
  if (c->grid.nt_store)
    go_rowf_l<E, true>(c, g, dim3(rpb * T), lds, check, src_kind, src_ptr, sc, rpb);
  else
    go_rowf_l<E, false>(c, g, dim3(rpb * T), lds, check, src_kind, src_ptr, sc, rpb);
}

template <int E, bool NT, int LS>
void go_col(rdcu_ctx *c, dim3 g, dim3 b, size_t lds, const rd::FftPlan &p, const rd::ColGeom &cg,
            int check, const rd::FacList &fl, int mode, int fpw) {
  if constexpr (E == 8) {
    // the 3-D y forward pass: the build without the multiply path (100
    // instead of 180 registers, more blocks per CU: 0.103 -> 0.085 ms at
    // 256^3).  The y inverse pass keeps the general build: with more blocks
    // in flight its six factor spectra's scattered stores took longer
    // (0.342 -> 0.359 ms).
    if (mode == 0) {
      hipLaunchKernelGGL((rd::k_col_reg<E, NT, LS, false, true>), g, b, lds, c->stream, c->grid, p,
                         cg, c->st, check, c->bufs, fl, mode, fpw);
      return;
    }
    // several factors per block (the 3-D y pass with the factors without
    // k_z, or the multiply pass with fpw > 1): their forward spectrum stays
    // in registers (col_kr)
    if (mode == 3 || (mode == 1 && fpw > 1)) {
      hipLaunchKernelGGL((rd::k_col_reg<E, NT, LS, false, false, true>), g, b, lds, c->stream,
                         c->grid, p, cg, c->st, check, c->bufs, fl, mode, fpw);
      return;
    }
  }
  hipLaunchKernelGGL((rd::k_col_reg<E, NT, LS>), g, b, lds, c->stream, c->grid, p, cg, c->st,
                     check, c->bufs, fl, mode, fpw);
}

// Strided lines of at least this length with 8 values per thread read their
// twiddles and wavenumbers from global memory (k_col_reg<8, NT, 0, true>):
// the LDS copies would leave room for a single column per block.
constexpr int kColTgMin = 1024;
bool col_tg(int L, int E) { return E == 8 && L >= kColTgMin && !odd_len(L); }

// LDS of a strided-pass block of tc columns of length L (mode 1: with the
// wavenumbers, and with the kept spectrum when a block takes several
// factors, fpw > 1).
// The strided kernel keeps a multi-factor forward spectrum in registers
// (k_col_reg<..., KR>) with 8 values per thread, except on lines long enough
// to read their twiddles from global memory.
bool col_kr(int L, int E) { return E == 8 && !col_tg(L, E); }

size_t col_lds(int L, int tc, int E, bool mode1, int fpw) {
  const bool tg = col_tg(L, E), keep = mode1 && fpw > 1;
  return (size_t)((tg ? 0 : L) + tc * rd::padded_len(L) + (keep ? tc * L : 0)) * sizeof(cx) +
         (mode1 && !tg ? (size_t)L * sizeof(double) : 0);
}

template <int E, bool NT>
void go_col_l(rdcu_ctx *c, dim3 g, dim3 b, size_t lds, const rd::FftPlan &p, const rd::ColGeom &cg,
              int check, const rd::FacList &fl, int mode, int fpw) {
  if constexpr (E == 8 || E == 16) {
    if (cg.L == kColLs[0]) return go_col<E, NT, kColLs[0]>(c, g, b, lds, p, cg, check, fl, mode, fpw);
    if (cg.L == kColLs[1]) return go_col<E, NT, kColLs[1]>(c, g, b, lds, p, cg, check, fl, mode, fpw);
  }
  if constexpr (E == 8) {
    if (col_tg(cg.L, E)) {
      hipLaunchKernelGGL((rd::k_col_reg<8, NT, 0, true>), g, b, lds, c->stream, c->grid, p, cg,
                         c->st, check, c->bufs, fl, mode, fpw);
      return;
    }
    if (cg.L == kColLs3[0]) return go_col<E, NT, kColLs3[0]>(c, g, b, lds, p, cg, check, fl, mode, fpw);
    if (cg.L == kColLs3[1]) return go_col<E, NT, kColLs3[1]>(c, g, b, lds, p, cg, check, fl, mode, fpw);
    if (odd_len(cg.L)) return go_col<E, NT, -1>(c, g, b, lds, p, cg, check, fl, mode, fpw);
  }
  go_col<E, NT, 0>(c, g, b, lds, p, cg, check, fl, mode, fpw);
}

template <int E>
void launch_colr(rdcu_ctx *c, const rd::FftPlan &p, rd::ColGeom cg, long ngroups, int check,
                 const rd::FacList &fl, int mode) {
  const int T = cg.L / E;
  const int fpw = std::max(1, c->cr_fpw);
  cg.nparts = mode == 1 ? (c->max_fpf + fpw - 1) / fpw : 1;
  if (cg.nparts < 1) cg.nparts = 1;
  cg.nblk = (int)(ngroups * cg.ntc * cg.nparts);
  dim3 g((unsigned)((cg.nblk + 7) / 8 * 8));
  // mode 3 (3-D y pass with the factors without k_z): all of a field's such
  // factors in one block
  const size_t lds =
      mode == 3 ? col_lds(cg.L, cg.tc, E, true, col_kr(cg.L, E) ? 1 : std::max(1, c->zs_used.zf_max))
                : col_lds(cg.L, cg.tc, E, mode == 1, col_kr(cg.L, E) ? 1 : fpw);
  if (c->grid.nt_store)
    go_col_l<E, true>(c, g, dim3(cg.tc * T), lds, p, cg, check, fl, mode, fpw);
  else
    go_col_l<E, false>(c, g, dim3(cg.tc * T), lds, p, cg, check, fl, mode, fpw);
}

// (a generated kernel keeps no program in LDS)
size_t rowi_lds(const rdcu_ctx *c, int G, int R) {
  return (size_t)R *
         rd::rowi_region(c->L, c->Nx, c->n, c->nf_used, c->dictp.nwords, G, !c->jit_rowi) *
         sizeof(cx);
}

// The generated x-inverse kernel of variant v (0 plain, 1 twiddles in
// registers, 2 lean, 3 lean with the twiddles in registers); null if that
// variant was not built.
hipFunction_t jit_rowi_variant(const rdcu_ctx *c, int v) {
  switch (v) {
    case 1: return c->jit_rowi_tw;
    case 2: return c->jit_rowi_lean;
    case 3: return c->jit_rowi_lean_tw;
    default: return c->jit_rowi;
  }
}

template <int E>
void launch_rowi(rdcu_ctx *c, int check, int src_kind, const double *src_ptr, int dst_kind,
                 int dst_slot, double *dst_ptr, int err, rd::StageCoef nx, int G, int R) {
  const int T = c->L / E;
  rd::ErrCoef ec{};
  for (int j = 0; j < 7; ++j) ec.e[j] = EE[j];
  const size_t lds = rowi_lds(c, G, R);
  const unsigned grid = (unsigned)(c->rows / R), block = (unsigned)(R * G * T);
  if (c->jit_rowi) {
    rd::Grid a_g = c->grid;
    rd::FftPlan a_px = c->px;
    const rd::cx *a_tw = c->twN;
    rd::DevState *a_st = c->st;
    int a_check = check;
    rd::Bufs a_b = c->bufs;
    rd::FacList a_fl = c->fl_used;
    rd::DictP a_d = c->dictp;
    int a_src = src_kind;
    const double *a_sp = src_ptr;
    int a_dk = dst_kind, a_ds = dst_slot;
    double *a_dp = dst_ptr;
    int a_err = err;
    rd::ErrCoef a_ec = ec;
    rd::StageCoef a_nx = nx;
    int a_G = G;
    void *args[] = {&a_g, &a_px, &a_tw, &a_st, &a_check, &a_b, &a_fl, &a_d, &a_src,
                    &a_sp, &a_dk, &a_ds, &a_dp, &a_err, &a_ec, &a_nx, &a_G};
    hipFunction_t fn = jit_rowi_variant(c, c->ri_tw);
    // a failed launch is reported by the hipGetLastError check after each
    // group of launches (and inside a graph capture it fails the capture)
    (void)hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, (unsigned)lds, c->stream, args,
                                nullptr);
    return;
  }
  if (odd_len(c->L))
    hipLaunchKernelGGL((rd::k_rowi_reg<E, -1>), dim3(grid), dim3(block), lds, c->stream, c->grid,
                       c->px, c->twN, c->st, check, c->bufs, c->fl_used, c->dictp, src_kind,
                       src_ptr, dst_kind, dst_slot, dst_ptr, err, ec, nx, G);
  else
    hipLaunchKernelGGL(rd::k_rowi_reg<E>, dim3(grid), dim3(block), lds, c->stream, c->grid, c->px,
                       c->twN, c->st, check, c->bufs, c->fl_used, c->dictp, src_kind, src_ptr,
                       dst_kind, dst_slot, dst_ptr, err, ec, nx, G);
}

void rowf(rdcu_ctx *c, int check, int src_kind, const double *src_ptr, rd::StageCoef sc, int rpb) {
  launch_rowf<8>(c, check, src_kind, src_ptr, sc, rpb);
}
void colr(rdcu_ctx *c, const rd::FftPlan &p, int E, rd::ColGeom cg, long ngroups, int check,
          const rd::FacList &fl, int mode) {
  if (E == 16) launch_colr<16>(c, p, cg, ngroups, check, fl, mode);
  else launch_colr<8>(c, p, cg, ngroups, check, fl, mode);
}
void rowi(rdcu_ctx *c, int check, int src_kind, const double *src_ptr, int dst_kind, int dst_slot,
          double *dst_ptr, int err, rd::StageCoef nx, int G, int R) {
  launch_rowi<8>(c, check, src_kind, src_ptr, dst_kind, dst_slot, dst_ptr, err, nx, G, R);
}

rd::StageCoef stage_coef(int stage) {
  rd::StageCoef sc{};
  if (stage >= 2) {
    sc.ns = stage - 1;
    for (int j = 0; j < sc.ns; ++j) sc.a[j] = STAGE_A[stage - 2][j];
  }
  return sc;
}

// Column tiles of the strided passes: with ntc > 0, ntc tiles of even
// widths per group, each block as wide as the widest (bal = 1); otherwise
// tiles of tc columns from column 0 (bal = 0, ntc = as many as needed).
int col_tiles(const rdcu_ctx *c, int &tc, int &ntc) {
  if (ntc <= 0 || ntc > c->Nxh) {
    ntc = (c->Nxh + tc - 1) / tc;
    return 0;
  }
  tc = (c->Nxh + ntc - 1) / ntc;
  return 1;
}

// Strided passes of the register path for one RHS (after the x forward pass).
std::vector<Launch> col_launches_reg(rdcu_ctx *c, int check, int tc, int ntc) {
  std::vector<Launch> v;
  if (c->col_lds) {  // LDS column kernels inside the register path (comparison switch)
    for (auto &l : rhs_launches(c, 0, rd::SRC_Y, nullptr, rd::DST_PTR, 0, nullptr, check, 0))
      if (std::string(l.name).rfind("col", 0) == 0) v.push_back(l);
    return v;
  }
  const int nd = c->ndim;
  const int bal = col_tiles(c, tc, ntc);
  // 3-D: the y forward pass also makes the factors without k_z (mode 3);
  // the z pass and the inverse y pass take the others
  const ZSplit &zs = c->zs_used;
  if (nd == 3) {
    rd::ColGeom cg{c->Ns[1], (long)c->Nxp, c->Ns[0], c->Nh, (long)c->Ns[1] * c->Nxp, ntc, tc, 1};
    cg.bal = bal;
    long ng = (long)c->n * c->Ns[0];
    if (zs.zf.nf > 0)
      v.push_back({"col_fwd", [=]() { colr(c, c->p1, c->E1, cg, ng, check, zs.zf, 3); }});
    else
      v.push_back({"col_fwd", [=]() { colr(c, c->p1, c->E1, cg, ng, check, c->fl_used, 0); }});
  }
  if (nd == 2 || (nd == 3 && zs.zd.nf > 0)) {
    rd::ColGeom cg{};
    cg.L = c->Ns[0];
    cg.ntc = ntc;
    cg.tc = tc;
    cg.bal = bal;
    cg.gs1 = c->Nh;
    if (nd == 2) {
      cg.es = c->Nxp;
      cg.gdiv = 1;
    } else {
      cg.es = (long)c->Ns[1] * c->Nxp;
      cg.gdiv = c->Ns[1];
      cg.gs2 = c->Nxp;
    }
    long ng = (long)c->n * cg.gdiv;
    const rd::FacList fl = nd == 3 ? zs.zd : c->fl_used;
    v.push_back({"col_mult", [=]() { colr(c, c->p0, c->E0, cg, ng, check, fl, 1); }});
  }
  if (nd == 3 && zs.zd.nf > 0) {
    rd::ColGeom cg{c->Ns[1], (long)c->Nxp, c->Ns[0], c->Nh, (long)c->Ns[1] * c->Nxp, ntc, tc, 1};
    cg.bal = bal;
    long ng = (long)zs.zd.nf * c->Ns[0];
    v.push_back({"col_inv", [=]() { colr(c, c->p1, c->E1, cg, ng, check, zs.zd, 2); }});
  }
  return v;
}

// One plain RHS on the register path (no stage fusion).
std::vector<Launch> rhs_launches_reg(rdcu_ctx *c, int src_kind, const double *src_ptr, int dst_kind,
                                     int dst_slot, double *dst_ptr, int check) {
  std::vector<Launch> v;
  rd::StageCoef none{};
  v.push_back({"row_fwd", [=]() { rowf(c, check, src_kind, src_ptr, none, c->rf_rpb); }});
  for (auto &l : col_launches_reg(c, check, c->cr_tc, c->cr_ntc)) v.push_back(l);
  v.push_back({"row_inv", [=]() {
                 rowi(c, check, src_kind, src_ptr, dst_kind, dst_slot, dst_ptr, 0, none, c->ri_G, c->ri_R);
               }});
  return v;
}

// One DP45 attempt on the register path: x forward of stage 2, then per
// stage the strided passes and the x inverse pass fused with the next
// stage's sum and x forward pass; stage 7 adds the error tree.
std::vector<Launch> attempt_launches_reg(rdcu_ctx *c) {
  std::vector<Launch> v;
  rd::StageCoef s2 = stage_coef(2);
  v.push_back({"row_fwd", [=]() { rowf(c, 1, rd::SRC_STAGE, nullptr, s2, c->rf_rpb); }});
  for (int s = 2; s <= 7; ++s) {
    for (auto &l : col_launches_reg(c, 1, c->cr_tc, c->cr_ntc)) v.push_back(l);
    rd::StageCoef nx = s < 7 ? stage_coef(s + 1) : rd::StageCoef{};
    int dk = s < 7 ? rd::DST_KS : rd::DST_K7;
    int err = s == 7 && !c->fixed;
    v.push_back({"row_inv", [=]() {
                   rowi(c, 1, rd::SRC_YT, nullptr, dk, s - 1, nullptr, err, nx, c->ri_G, c->ri_R);
                 }});
  }
  return v;
}

// Straight-line evaluator of a dictionary structure: per field, terms in
// input order, term = 1 * v^p * ..., acc = fma(C, term, acc) from 0 -- the
// same operations as dict_field, with slots and powers as compile-time
// constants.  The coefficients are read from the dictionary program at fixed
// word offsets (uniform loads), so every dictionary with the same structure
// shares one compiled kernel.
std::string jit_dict_source(int n, int nslots, const std::vector<int> &tfield,
                            const std::vector<int> &tstart, const std::vector<int> &fslot,
                            const std::vector<int> &fpow, const std::string &name = "JitDict") {
  auto powexpr = [](const std::string &v, int p) {
    if (p == 0) return std::string("1.0");
    int a = p < 0 ? -p : p;
    std::string e = v;
    for (int j = 1; j < a; ++j) e = "(" + e + " * " + v + ")";
    return p < 0 ? "(1.0 / " + e + ")" : e;
  };
  const int nt = (int)tfield.size();
  // word offset of each term's coefficient in the program (layout of DictP)
  std::vector<int> coff(nt, 0);
  {
    int pos = n + 1;
    for (int f = 0; f < n; ++f)
      for (int i = 0; i < nt; ++i)
        if (tfield[i] == f) {
          coff[i] = pos + 1;
          pos += 3 + 2 * (tstart[i + 1] - tstart[i]);
        }
  }
  std::string s;
  s += "namespace rd {\nstruct " + name + " {\n  static constexpr bool kProg = false;\n";
  s += "  template <int XM>\n  __device__ __forceinline__ static void eval(const int *__restrict__ P, "
       "int f, const double *vals, int Nx, int x0, int stride, double (&acc)[XM]) {\n";
  s += "    switch (f) {\n";
  char buf[256];
  for (int f = 0; f < n; ++f) {
    snprintf(buf, sizeof buf, "    case %d: {\n", f);
    s += buf;
    for (int i = 0; i < nt; ++i)
      if (tfield[i] == f) {
        snprintf(buf, sizeof buf, "      const double c%d = __hiloint2double(P[%d], P[%d]);\n", i,
                 coff[i] + 1, coff[i]);
        s += buf;
      }
    s += "#pragma unroll\n      for (int i = 0; i < XM; ++i) {\n";
    s += "        const int x = x0 + stride * i;\n        const bool in = x < Nx;\n";
    s += "        double a = 0.0;\n";
    std::vector<char> used(nslots, 0);
    for (int i = 0; i < nt; ++i)
      if (tfield[i] == f)
        for (int j = tstart[i]; j < tstart[i + 1]; ++j) used[fslot[j]] = 1;
    for (int sl = 0; sl < nslots; ++sl)
      if (used[sl]) {
        snprintf(buf, sizeof buf, "        const double v%d = in ? vals[%d * Nx + x] : 0.0;\n", sl,
                 sl);
        s += buf;
      }
    for (int i = 0; i < nt; ++i) {
      if (tfield[i] != f) continue;
      s += "        { double t = 1.0;";
      for (int j = tstart[i]; j < tstart[i + 1]; ++j)
        s += " t = t * " + powexpr("v" + std::to_string(fslot[j]), fpow[j]) + ";";
      snprintf(buf, sizeof buf, " a = __builtin_fma(c%d, t, a); }\n", i);
      s += buf;
    }
    s += "        acc[i] = a;\n      }\n    } break;\n";
  }
  s += "    default:\n#pragma unroll\n      for (int i = 0; i < XM; ++i) acc[i] = 0.0;\n";
  s += "    }\n  }\n};\n}  // namespace rd\n";
  return s;
}

std::string jit_arch(int dev) {
  hipDeviceProp_t prop;
  std::string arch = "--offload-arch=gfx950";
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
    std::string a = prop.gcnArchName;
    arch = "--offload-arch=" + a.substr(0, a.find(':'));
  }
  return arch;
}

// hiprtc compile of `src` for `arch`; code objects are cached per process,
// keyed by source and target.  Returns false with a log on failure.
bool jit_compile(const std::string &src, const std::string &arch, std::vector<char> &code,
                 std::string &log) {
  static std::mutex cache_mu;
  static std::map<std::string, std::vector<char>> cache;
  const std::string key = arch + "\n" + src;
  {
    std::lock_guard<std::mutex> lk(cache_mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      code = it->second;
      return true;
    }
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "rdcu_rowi_jit.hip", kJitNumHeaders, kJitHeaders,
                          kJitHeaderNames) != HIPRTC_SUCCESS) {
    log = "hiprtcCreateProgram failed";
    return false;
  }
  const char *opts[] = {arch.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  hiprtcResult r = hiprtcCompileProgram(prog, 9, opts);
  if (r != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string lg(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &lg[0]);
    log = "hiprtc: " + lg;
    hiprtcDestroyProgram(&prog);
    return false;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  code.resize(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  std::lock_guard<std::mutex> lk(cache_mu);
  cache[key] = code;
  return true;
}

bool jit_enabled() {
  const char *env = getenv("RDCU_JIT");
  return !(env && !strcmp(env, "0"));
}

// Compiles the x-inverse kernel with the dictionary structure built in.  On
// any failure the interpreter kernel (k_rowi_reg) stays in use; RDCU_JIT=0
// disables it.
// True if a generated kernel keeps values in scratch (private) memory.  Such
// kernels are not used: every kernel here is meant to run from registers and
// LDS, and a spilling strided kernel once gave wrong values on the GPU.
bool jit_spills(hipFunction_t f) {
  int local = 0;
  if (hipFuncGetAttribute(&local, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, f) != hipSuccess) {
    (void)hipGetLastError();
    return true;
  }
  return local > 0;
}

// LDS of the single-workgroup 1-D attempt (256 threads: G = 256/T line
// groups), or 0 where it does not apply (not 1-D, not the register path, fewer
// line groups than fields, or more LDS than a CU has).
size_t attempt1d_lds(const rdcu_ctx *c) {
  if (c->ndim != 1 || !c->reg || c->Ex != 8) return 0;
  const int T = c->L / c->Ex;
  if (T < 1 || 256 % T) return 0;
  const int G = 256 / T;
  if (G < c->n) return 0;
  size_t lds = (size_t)(c->L + (long)G * rd::padded_len(c->L)) * sizeof(cx);
  lds = std::max(lds, (size_t)rd::rowi_region(c->L, c->Nx, c->n, c->nf_used, c->dictp.nwords, G) *
                          sizeof(cx));
  lds = std::max(lds, (size_t)c->n * sizeof(double));
  return lds <= 160 * 1024 ? lds : 0;
}

void build_jit(rdcu_ctx *c) {
  if (!jit_enabled()) return;
  std::string src = "#include \"rd_kernels_reg.h\"\n" +
                    jit_dict_source(c->n, c->n + c->nf_used, c->h_tfield, c->h_tstart,
                                    c->h_fslot, c->h_fpow);
  src += "extern \"C\" __global__ __launch_bounds__(256) void rdcu_rowi_jit(rd::Grid g, rd::FftPlan px, "
         "const rd::cx *__restrict__ twN, rd::DevState *st, int check_active, rd::Bufs b, "
         "rd::FacList fl, rd::DictP dict, int src_kind, const double *src_ptr, int dst_kind, "
         "int dst_slot, double *dst_ptr, int err, rd::ErrCoef ec, rd::StageCoef nx, int G) {\n";
  const std::string E = std::to_string(c->Ex);
  // The grid sizes and the factor count are constants of the generated
  // kernel (index arithmetic folds, and the 1-D-only table reads disappear
  // in 2-D and 3-D); the launch passes the same values in g and fl.
  const rd::Grid &gd = c->grid;
  char buf[512];
  snprintf(buf, sizeof buf,
           "namespace rd {\nstruct JitGeom {\n  static constexpr int n = %d, ndim = %d, Nx = %d, "
           "Nxp = %d, nf = %d;\n  static constexpr long N = %ldL, rows = %ldL;\n};\n}\n",
           gd.n, gd.ndim, gd.Nx, gd.Nxp, c->fl_used.nf, gd.N, gd.rows);
  src = buf + src;
  const std::string sig = src.substr(src.rfind("extern \"C\""));
  src += "  rd::rowi_rows<" + E + ", rd::JitDict, " + std::to_string(c->L) +
         ", rd::JitGeom>(g, px, twN, st, b, fl, dict, src_kind, src_ptr, dst_kind, dst_slot, "
         "dst_ptr, err, ec, nx, G, check_active ? st : nullptr);\n}\n";
  // the same kernel with the transforms' twiddles in registers, and both
  // compiled lean for two wavefronts per SIMD: at 256^3 the full kernel needs
  // more than 256 registers and a SIMD holds one of its wavefronts, four rows
  // per CU where the LDS has room for seven (the tuner times all four)
  if (c->Ex == 8) {
    const struct { const char *name; const char *bounds; const char *args; } vs[] = {
        {"rdcu_rowi_jit_tw", "256", "true, false"},
        {"rdcu_rowi_jit_lean", "256, 2", "false, true"},
        {"rdcu_rowi_jit_lean_tw", "256, 2", "true, true"}};
    for (const auto &v : vs) {
      std::string sv = sig;
      sv.replace(sv.find("__launch_bounds__(256)"), 22, std::string("__launch_bounds__(") + v.bounds + ")");
      sv.replace(sv.find("rdcu_rowi_jit"), 13, v.name);
      src += sv + "  rd::rowi_rows<" + E + ", rd::JitDict, " + std::to_string(c->L) +
             ", rd::JitGeom, " + v.args + ">(g, px, twN, st, b, fl, dict, src_kind, src_ptr, "
             "dst_kind, dst_slot, dst_ptr, err, ec, nx, G, check_active ? st : nullptr);\n}\n";
    }
  }
  const size_t att_lds = attempt1d_lds(c);
  if (att_lds)
    src += "extern \"C\" __global__ __launch_bounds__(256) void rdcu_attempt1d_jit(rd::Grid g, "
           "rd::FftPlan px, const rd::cx *__restrict__ twN, rd::DevState *st, rd::Bufs b, "
           "rd::FacList fl, rd::DictP dict, rd::AttemptCoef ac, int fixed) {\n"
           "  rd::attempt1d_body<" + E + ", rd::JitDict, " + std::to_string(c->L) +
           ", rd::JitGeom>(g, px, twN, st, b, fl, dict, ac, fixed);\n}\n";
  std::vector<char> code;
  if (!jit_compile(src, jit_arch(c->dev), code, c->jit_log)) return;
  if (hipModuleLoadData(&c->jit_mod, code.data()) != hipSuccess ||
      hipModuleGetFunction(&c->jit_rowi, c->jit_mod, "rdcu_rowi_jit") != hipSuccess ||
      jit_spills(c->jit_rowi)) {
    c->jit_log = c->jit_rowi && jit_spills(c->jit_rowi) ? "spills to scratch" : "module load failed";
    if (c->jit_mod) (void)hipModuleUnload(c->jit_mod);
    c->jit_mod = nullptr;
    c->jit_rowi = nullptr;
    return;
  }
  // a variant that does not load or would spill is left out
  auto variant = [&](hipFunction_t *f, const char *name) {
    if (c->Ex != 8 || hipModuleGetFunction(f, c->jit_mod, name) != hipSuccess || jit_spills(*f))
      *f = nullptr;
  };
  variant(&c->jit_rowi_tw, "rdcu_rowi_jit_tw");
  variant(&c->jit_rowi_lean, "rdcu_rowi_jit_lean");
  variant(&c->jit_rowi_lean_tw, "rdcu_rowi_jit_lean_tw");
  if (att_lds && hipModuleGetFunction(&c->jit_att, c->jit_mod, "rdcu_attempt1d_jit") == hipSuccess &&
      !jit_spills(c->jit_att))
    c->lds_att = att_lds;
  else
    c->jit_att = nullptr;
  c->jit_log = "ok";
}

// Autotuner: every candidate shape of each kernel computes the same values
// (the arithmetic does not depend on the tiling), so the fastest is kept.
// Candidates are timed alone with HIP events on the plain RHS of the
// current state.
// Shapes given by the caller instead of the autotuner (profiling runs that
// must see only the launches of one shape), checked against the same limits.

// Whether every strided-pass launch with tc columns per block fits the LDS
// of a CU (axis 0 keeps the forward spectrum; in 3-D axis 1 does not).
bool col_fits(const rdcu_ctx *c, int tc, int fpw) {
  size_t lds = col_lds(c->Ns[0], tc, c->E0, true, col_kr(c->Ns[0], c->E0) ? 1 : fpw);
  if (c->ndim == 3) {
    const int zf = c->zs_used.zf_max;
    lds = std::max(lds, col_lds(c->Ns[1], tc, c->E1, zf > 0,
                                col_kr(c->Ns[1], c->E1) ? 1 : std::max(1, zf)));
  }
  return lds <= 160 * 1024;
}

// Largest strided-pass block for the values per thread in use (the
// kernels' launch bounds).
int col_threads(const rdcu_ctx *c) {
  auto lim = [](int L, int E) {
    return E == 8 && !odd_len(L) ? rd::col_max_threads<8>() : rd::col_max_threads<16>();
  };
  int m = lim(c->Ns[0], c->E0);
  if (c->ndim == 3) m = std::min(m, lim(c->Ns[1], c->E1));
  return m;
}

int pow2_above(int x) {
  int p = 1;
  while (p <= x) p *= 2;
  return p;
}

int fixed_shapes(rdcu_ctx *c, const char *spec) {
  int rpb = 0, tc = 0, fpw = 0, G = 0, R = 1, Ec = 0, ntc = 0, tw = 0;
  const int got = sscanf(spec, "%d,%d,%d,%d,%d,%d,%d,%d", &rpb, &tc, &fpw, &G, &R, &Ec, &ntc, &tw);
  if (got < 4) return fail(RDCU_EINVAL, "RDCU_SHAPES must be rpb,tc,fpw,G[,R[,Ecol[,ntc[,tw]]]]");
  if (c->ndim >= 2 && tc >= 1 && col_tiles(c, tc, ntc) == 0) ntc = 0;
  if (Ec == 16) {  // 16 values per thread on the power-of-two strided axes
    if (c->E0 == 8 && c->Ns[0] >= 16 && !odd_len(c->Ns[0])) c->E0 = 16;
    if (c->ndim == 3 && c->E1 == 8 && c->Ns[1] >= 16 && !odd_len(c->Ns[1])) c->E1 = 16;
  } else if (Ec != 0 && Ec != 8) {
    return fail(RDCU_EINVAL, "RDCU_SHAPES column values per thread must be 8 or 16");
  }
  const int T = c->L / c->Ex;
  int Lmax = c->Ns[0], Emin = c->E0;
  if (c->ndim == 3) {
    Lmax = std::max(Lmax, c->Ns[1]);
    Emin = std::min(c->E0, c->E1);
  }
  const bool ok =
      rpb >= 1 && rpb * T <= 256 &&
      (size_t)(c->L + rpb * rd::padded_len(c->L)) * sizeof(cx) <= 160 * 1024 &&
      (c->ndim < 2 || (tc >= 1 && tc * (Lmax / Emin) <= col_threads(c) &&
                       col_fits(c, tc, fpw) &&
                       fpw >= 1 && fpw <= std::max(1, c->max_fpf))) &&
      G >= c->n && R >= 1 && (R == 1 || c->Ex <= 8) &&
      R * G * T <= 256 &&
      c->rows % R == 0 && rowi_lds(c, G, R) <= 160 * 1024;
  if (!ok) return fail(RDCU_EINVAL, std::string("RDCU_SHAPES out of range: ") + spec);
  c->rf_rpb = rpb;
  c->cr_tc = tc;
  c->cr_ntc = ntc;
  c->cr_fpw = fpw;
  c->ri_G = G;
  c->ri_R = R;
  c->ri_tw = tw >= 0 && tw <= 3 && jit_rowi_variant(c, tw) ? tw : 0;
  c->tune_log = std::string("fixed ") + spec;
  return RDCU_OK;
}

int autotune(rdcu_ctx *c) {
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  auto time_of = [&](const std::function<void()> &f) -> double {
    for (int i = 0; i < 2; ++i) f();
    (void)hipEventRecord(e0, c->stream);
    const int reps = 10;
    for (int i = 0; i < reps; ++i) f();
    (void)hipEventRecord(e1, c->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (hipGetLastError() != hipSuccess) return 1e30;
    return ms / reps;
  };
  const int T = c->L / c->Ex;
  rd::StageCoef none{};
  char line[160];
  // x forward: rows per block
  double best = 1e30;
  for (int rpb = std::max(1, 64 / T); rpb <= 128; rpb *= 2) {  // from one wavefront up
    if (rpb * T > 256 || (size_t)(c->L + rpb * rd::padded_len(c->L)) * sizeof(cx) > 160 * 1024) break;
    double tm = time_of([&]() { rowf(c, 0, rd::SRC_Y, nullptr, none, rpb); });
    snprintf(line, sizeof line, "row_fwd rpb=%d %.2fus; ", rpb, tm * 1e3);
    c->tune_log += line;
    if (tm < best) {
      best = tm;
      c->rf_rpb = rpb;
    }
  }
  // strided passes: values per thread (8 or 16 for power-of-two lengths:
  // the same butterflies either way), column tile width, factors per block
  if (c->ndim >= 2) {
    best = 1e30;
    const int E0b = c->E0, E1b = c->E1;
    int bestE0 = E0b, bestE1 = E1b;
    int best_fpw = c->max_fpf;
    for (int ev = 0; ev < 2; ++ev) {
      if (ev == 1) {
        bool any = false;
        if (E0b == 8 && c->Ns[0] >= 16 && !odd_len(c->Ns[0])) c->E0 = 16, any = true;
        if (c->ndim == 3 && E1b == 8 && c->Ns[1] >= 16 && !odd_len(c->Ns[1]))
          c->E1 = 16, any = true;
        if (!any) break;
      }
      int Lmax = c->Ns[0], Tmax = c->Ns[0] / c->E0;
      if (c->ndim == 3) {
        Lmax = std::max(Lmax, c->Ns[1]);
        Tmax = std::max(Tmax, c->Ns[1] / c->E1);
      }
      for (int tc0 = std::max(1, 64 / Tmax); tc0 <= 64; tc0 *= 2) {
        if (tc0 * Tmax > col_threads(c) || !col_fits(c, tc0, 1)) break;
        // tiles of tc0 columns (the last one holding the remainder), and the
        // same columns in one tile fewer, spread evenly (blocks one column
        // wider), so that the block count can land on a multiple of the CUs
        const int full = (c->Nxh + tc0 - 1) / tc0;
        for (int k = 0; k < 2; ++k) {
          int tc = tc0, ntc = k == 0 ? 0 : c->Nxh / tc0;
          if (k == 1 && (ntc < 1 || ntc == full)) continue;
          col_tiles(c, tc, ntc);
          if (k == 0) ntc = 0;  // tiles of tc0 from column 0
          if (tc * Tmax > col_threads(c) || !col_fits(c, tc, 1)) continue;
          for (int fpw = c->max_fpf; fpw >= 1; fpw = fpw > 1 ? (fpw + 1) / 2 : 0) {
            if (!col_fits(c, tc, fpw)) continue;
            c->cr_fpw = fpw;
            auto ls = col_launches_reg(c, 0, tc, ntc);
            double tm = time_of([&]() { for (auto &l : ls) l.go(); });
            snprintf(line, sizeof line, "col E=%d tc=%d ntc=%d fpw=%d %.2fus; ", c->E0, tc, ntc,
                     fpw, tm * 1e3);
            c->tune_log += line;
            if (tm < best) {
              best = tm;
              c->cr_tc = tc;
              c->cr_ntc = ntc;
              best_fpw = fpw;
              bestE0 = c->E0;
              bestE1 = c->E1;
            }
            if (fpw == 1) break;
          }
        }
      }
    }
    c->E0 = bestE0;
    c->E1 = bestE1;
    c->cr_fpw = best_fpw;
  }
  int best_tw = 0;
  // x inverse: factor lines per group, rows per block (dividing the rows).
  // Groups wider than the factor count (powers of two above it) leave lines
  // idle but spread the dictionary and the stage sums over more threads,
  // which pays when there are few rows (1-D: one row).
  best = 1e30;
  const int Gf = std::max(c->n, c->nf_used);
  // (with a generated kernel, also its variant with the twiddles in
  // registers, timed on the fused pass of an attempt's stage, where it runs)
  rd::StageCoef nx3 = stage_coef(3);
  for (int tw = 0; tw <= 3; ++tw) {
    if (tw > 0 && !jit_rowi_variant(c, tw)) continue;
    c->ri_tw = tw;
    for (int G = c->n; G * T <= 256; G = G < Gf ? G + 1 : (G == Gf ? pow2_above(Gf) : 2 * G)) {
      for (int R = 1; R * G * T <= 256 && (R == 1 || c->Ex <= 8); R *= 2) {
        if (c->rows % R) break;
        if (rowi_lds(c, G, R) > 160 * 1024) break;
        double tm = time_of([&]() {
          rowi(c, 0, rd::SRC_Y, nullptr, rd::DST_KS, 1, nullptr, 0, nx3, G, R);
        });
        snprintf(line, sizeof line, "row_inv tw=%d G=%d R=%d %.2fus; ", tw, G, R, tm * 1e3);
        c->tune_log += line;
        if (tm < best) {
          best = tm;
          c->ri_G = G;
          c->ri_R = R;
          best_tw = tw;
        }
      }
    }
  }
  c->ri_tw = best_tw;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipGetLastError());
  return RDCU_OK;
}

// Raise the dynamic-LDS limit of every register-path kernel to the 160 KB of
// a gfx950 CU.
int allow_big_lds() {
  const int lim = 160 * 1024;
  const void *fs[] = {
      (const void *)rd::k_rowf_reg<8, false>,     (const void *)rd::k_rowf_reg<8, true>,
      (const void *)rd::k_rowf_reg<8, false, -1>, (const void *)rd::k_rowf_reg<8, true, -1>,
      (const void *)rd::k_col_reg<8, false>,      (const void *)rd::k_col_reg<8, true>,
      (const void *)rd::k_col_reg<16, false>,     (const void *)rd::k_col_reg<16, true>,
      (const void *)rd::k_col_reg<8, false, -1>,  (const void *)rd::k_col_reg<8, true, -1>,
      (const void *)rd::k_rowi_reg<8>,            (const void *)rd::k_rowi_reg<8, -1>,
      (const void *)rd::k_rowi_vjp<8>,            (const void *)rd::k_rowi_vjp<8, -1>,
      (const void *)rd::k_rowf_reg<8, false, kRowLs[0]>,
      (const void *)rd::k_rowf_reg<8, true, kRowLs[0]>,
      (const void *)rd::k_rowf_reg<8, false, kRowLs[1]>,
      (const void *)rd::k_rowf_reg<8, true, kRowLs[1]>,
      (const void *)rd::k_col_reg<8, false, kColLs[0]>,
      (const void *)rd::k_col_reg<8, true, kColLs[0]>,
      (const void *)rd::k_col_reg<8, false, kColLs[1]>,
      (const void *)rd::k_col_reg<8, true, kColLs[1]>,
      (const void *)rd::k_col_reg<16, false, kColLs[0]>,
      (const void *)rd::k_col_reg<16, true, kColLs[0]>,
      (const void *)rd::k_col_reg<16, false, kColLs[1]>,
      (const void *)rd::k_col_reg<16, true, kColLs[1]>,
      (const void *)rd::k_rowf_reg<8, false, kRowLs3[0]>,
      (const void *)rd::k_rowf_reg<8, true, kRowLs3[0]>,
      (const void *)rd::k_rowf_reg<8, false, kRowLs3[1]>,
      (const void *)rd::k_rowf_reg<8, true, kRowLs3[1]>,
      (const void *)rd::k_col_reg<8, false, kColLs3[0]>,
      (const void *)rd::k_col_reg<8, true, kColLs3[0]>,
      (const void *)rd::k_col_reg<8, false, kColLs3[1]>,
      (const void *)rd::k_col_reg<8, true, kColLs3[1]>,
      (const void *)rd::k_col_reg<8, false, 0, false, true>,
      (const void *)rd::k_col_reg<8, true, 0, false, true>,
      (const void *)rd::k_col_reg<8, false, -1, false, true>,
      (const void *)rd::k_col_reg<8, true, -1, false, true>,
      (const void *)rd::k_col_reg<8, false, kColLs[0], false, true>,
      (const void *)rd::k_col_reg<8, true, kColLs[0], false, true>,
      (const void *)rd::k_col_reg<8, false, kColLs[1], false, true>,
      (const void *)rd::k_col_reg<8, true, kColLs[1], false, true>,
      (const void *)rd::k_col_reg<8, false, kColLs3[0], false, true>,
      (const void *)rd::k_col_reg<8, true, kColLs3[0], false, true>,
      (const void *)rd::k_col_reg<8, false, kColLs3[1], false, true>,
      (const void *)rd::k_col_reg<8, true, kColLs3[1], false, true>,
      (const void *)rd::k_col_reg<8, false, 0, true>,
      (const void *)rd::k_col_reg<8, true, 0, true>,
      (const void *)rd::k_col_reg<8, false, 0, false, false, true>,
      (const void *)rd::k_col_reg<8, true, 0, false, false, true>,
      (const void *)rd::k_col_reg<8, false, -1, false, false, true>,
      (const void *)rd::k_col_reg<8, true, -1, false, false, true>,
      (const void *)rd::k_col_reg<8, false, kColLs[0], false, false, true>,
      (const void *)rd::k_col_reg<8, true, kColLs[0], false, false, true>,
      (const void *)rd::k_col_reg<8, false, kColLs[1], false, false, true>,
      (const void *)rd::k_col_reg<8, true, kColLs[1], false, false, true>,
      (const void *)rd::k_col_reg<8, false, kColLs3[0], false, false, true>,
      (const void *)rd::k_col_reg<8, true, kColLs3[0], false, false, true>,
      (const void *)rd::k_col_reg<8, false, kColLs3[1], false, false, true>,
      (const void *)rd::k_col_reg<8, true, kColLs3[1], false, false, true>};
  for (const void *f : fs)
    HIPCHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  return RDCU_OK;
}

// Runs fn (a set-up such as raising kernels' dynamic-LDS limit, which holds
// for the device current when it runs) once per device and slot, whichever
// thread gets there first; calls that arrive meanwhile wait for it.  A failure
// is returned to the caller and the set-up is tried again next time.
int once_per_device(int dev, int slot, int (*fn)()) {
  static std::mutex mu;
  static bool done[kMaxDevices][2] = {};
  if (dev < 0 || dev >= kMaxDevices || slot < 0 || slot > 1)
    return fail(RDCU_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (done[dev][slot]) return RDCU_OK;
  const int st = fn();
  if (st == RDCU_OK) done[dev][slot] = true;
  return st;
}

// Every factor of fl (every derivative block of every field: fl_all, or
// their imaginary parts: fl_imag; zs its 3-D split), real space, into out (S
// must hold the current forward spectra, in 3-D after the y forward pass; it
// is only read).
int enqueue_all_factors(rdcu_ctx *c, const rd::FacList &fl, const ZSplit &zs, double *out) {
  if (c->nf_all == 0) return RDCU_OK;
  const int nd = c->ndim;
  const int ntc = (c->Nxh + c->col_tc - 1) / c->col_tc;
  if (nd == 3 && zs.zf.nf > 0) {  // the factors without k_z: multiply, inverse y
    rd::ColGeom cg{c->Ns[1], (long)c->Nxp, c->Ns[0], c->Nh, (long)c->Ns[1] * c->Nxp, ntc,
                   c->col_tc, 1};
    dim3 gg((unsigned)((long)c->n * c->Ns[0] * ntc));
    hipLaunchKernelGGL(rd::k_col, gg, dim3(c->col_thr), c->lds_col, c->stream, c->grid, c->p1, cg,
                       c->st, 0, c->bufs, zs.zf, 4);
  }
  if (nd == 2 || (nd == 3 && zs.zd.nf > 0)) {
    rd::ColGeom cg{};
    cg.L = c->Ns[0];
    cg.ntc = ntc;
    cg.tc = c->col_tc;
    cg.gs1 = c->Nh;
    if (nd == 2) {
      cg.es = c->Nxp;
      cg.gdiv = 1;
    } else {
      cg.es = (long)c->Ns[1] * c->Nxp;
      cg.gdiv = c->Ns[1];
      cg.gs2 = c->Nxp;
    }
    dim3 gg((unsigned)((long)c->n * cg.gdiv * ntc));
    hipLaunchKernelGGL(rd::k_col, gg, dim3(c->col_thr), c->lds_col, c->stream, c->grid, c->p0, cg,
                       c->st, 0, c->bufs, nd == 3 ? zs.zd : fl, 1);
  }
  if (nd == 3 && zs.zd.nf > 0) {
    rd::ColGeom cg{c->Ns[1], (long)c->Nxp, c->Ns[0], c->Nh, (long)c->Ns[1] * c->Nxp, ntc,
                   c->col_tc, 1};
    dim3 gg((unsigned)((long)zs.zd.nf * c->Ns[0] * ntc));
    hipLaunchKernelGGL(rd::k_col, gg, dim3(c->col_thr), c->lds_col, c->stream, c->grid, c->p1, cg,
                       c->st, 0, c->bufs, zs.zd, 2);
  }
  rd::ErrCoef ec{};
  hipLaunchKernelGGL(rd::k_row_c2r, dim3((unsigned)c->rows), dim3(c->c2r_thr), c->lds_c2r_all,
                     c->stream, c->grid, c->px, c->twN, c->st, 0, c->bufs, fl, c->dict, 1,
                     (int)rd::SRC_PTR, (const double *)nullptr, (int)rd::DST_PTR, 0,
                     (double *)nullptr, 0, ec, c->c2r_grp_all, out);
  HIPCHECK(hipGetLastError());
  return RDCU_OK;
}

std::vector<Launch> attempt_launches_reg(rdcu_ctx *c);

int enqueue_attempt(rdcu_ctx *c) {
  if (c->reg && c->jit_att) {  // 1-D: the whole attempt, control included, in one workgroup
    rd::AttemptCoef ac{};
    ac.sc2 = stage_coef(2);
    for (int s = 2; s <= 7; ++s) ac.nx[s - 2] = s < 7 ? stage_coef(s + 1) : rd::StageCoef{};
    for (int j = 0; j < 7; ++j) ac.ec.e[j] = EE[j];
    rd::Grid a_g = c->grid;
    rd::FftPlan a_px = c->px;
    const rd::cx *a_tw = c->twN;
    rd::DevState *a_st = c->st;
    rd::Bufs a_b = c->bufs;
    rd::FacList a_fl = c->fl_used;
    rd::DictP a_d = c->dictp;
    int a_fixed = c->fixed;
    void *args[] = {&a_g, &a_px, &a_tw, &a_st, &a_b, &a_fl, &a_d, &ac, &a_fixed};
    HIPCHECK(hipModuleLaunchKernel(c->jit_att, 1, 1, 1, 256, 1, 1, (unsigned)c->lds_att,
                                   c->stream, args, nullptr));
    return RDCU_OK;
  }
  if (c->reg) {
    for (auto &l : attempt_launches_reg(c)) l.go();
  }
  for (int s = 2; !c->reg && s <= 7; ++s) {
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

// Enqueues K attempts: whole groups through the multi-attempt graph, the rest
// one graph each. Every attempt checks on the device whether it still runs, so
// the grouping does not change which attempts take effect.
int launch_attempts(rdcu_ctx *c, long long K) {
  long long i = 0;
  if (c->gexec_m)
    for (; i + kGraphAttempts <= K; i += kGraphAttempts)
      HIPCHECK(hipGraphLaunch(c->gexec_m, c->stream));
  for (; i < K; ++i) HIPCHECK(hipGraphLaunch(c->gexec, c->stream));
  return RDCU_OK;
}

int no_dict() {
  return fail(RDCU_EUNSUPPORTED, "this call needs a dictionary problem (the handle's RHS is a callback)");
}

int fn_step(rdcu_ctx *c, double *t, double *h, double hmax, int *accepted);
int fn_run(rdcu_ctx *c, double *t, double *h, double hmin, double hmax, double t1, double stop,
           rdcu_step_fn cb, void *user, long long max_attempts, long long *done);

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
// Shapes of the LDS-resident kernels.  lds_rhs: they can evaluate the RHS
// (needed when the register path does not apply); lds_y: the all-factor
// output of rdcu_rhs (strided pass + C2R store mode) fits.
int choose_shapes(rdcu_ctx *c) {
  const int L = c->L, Nx = c->Nx, n = c->n;
  auto round64 = [](long v) { return (int)std::max<long>(64, std::min<long>(256, (v + 63) / 64 * 64)); };
  c->lds_rhs = true;
  c->lds_y = true;
  c->r2c_rpb = std::max(1, std::min(8, 256 / std::max(1, L)));
  c->r2c_thr = round64((long)c->r2c_rpb * L / 4);
  c->lds_r2c = (size_t)2 * c->r2c_rpb * L * sizeof(cx);
  if (c->lds_r2c > kLdsMax) c->lds_rhs = false;
  if (c->ndim >= 2) {
    int Lmax = c->Ns[0];
    if (c->ndim == 3) Lmax = std::max(Lmax, c->Ns[1]);
    int tc = 1;
    while (tc * 2 <= 16 && (size_t)3 * (tc * 2) * Lmax * sizeof(cx) <= 48 * 1024) tc *= 2;
    if ((size_t)3 * tc * Lmax * sizeof(cx) > kLdsMax) c->lds_rhs = c->lds_y = false;
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
  if (c2r_lds(c->nf_used, grp) > kLdsMax) c->lds_rhs = false;
  c->c2r_grp = grp;
  c->lds_c2r = c2r_lds(c->nf_used, grp);
  c->c2r_thr = round64((long)Nx / 2);
  // store mode keeps no value rows: only the FFT scratch of its groups
  int ga = std::max(1, c->nf_all);
  while (ga > 1 && (size_t)2 * ga * L * sizeof(cx) > kLdsMax) --ga;
  if ((size_t)2 * ga * L * sizeof(cx) > kLdsMax) c->lds_y = false;
  c->c2r_grp_all = ga;
  c->lds_c2r_all = (size_t)2 * ga * L * sizeof(cx);
  return RDCU_OK;
}

}  // namespace

extern "C" {

const char *rdcu_last_error(void) { return g_err.c_str(); }
const char *rdcu_backend(void) { return "hip-gfx950"; }

// Where a batch member's stream and device state live (owned by the batch).
struct MemberOpts {
  hipStream_t stream;
  rd::DevState *st;  // device
  rd::DevState *sh;  // pinned host shadow
};

int create_ctx(const rdcu_problem *p, double atol, double rtol, int fixedstep, const double *y0,
               int device, const MemberOpts *mo, std::unique_ptr<rdcu_ctx> &c);

int rdcu_create(const rdcu_problem *p, double atol, double rtol, int fixedstep, const double *y0,
                int device, rdcu_ctx **out) {
  if (!out) return fail(RDCU_EINVAL, "null output handle");
  *out = nullptr;
  std::unique_ptr<rdcu_ctx> c;
  int st = create_ctx(p, atol, rtol, fixedstep, y0, device, nullptr, c);
  if (st) return st;
  *out = c.release();
  return RDCU_OK;
}

int create_ctx(const rdcu_problem *p, double atol, double rtol, int fixedstep, const double *y0,
               int device, const MemberOpts *mo, std::unique_ptr<rdcu_ctx> &c) {
  if (!p) return fail(RDCU_EINVAL, "null problem");
  if (p->n < 1) return fail(RDCU_EINVAL, "n must be >= 1");
  if (p->n > rd::kNMax) return fail(RDCU_EUNSUPPORTED, "at most 8 fields are supported");
  if (p->ndim < 1 || p->ndim > 3) return fail(RDCU_EINVAL, "ndim must be 1, 2 or 3");
  c.reset(new rdcu_ctx());
  c->member = mo != nullptr;
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
  c->Nxp = (c->Nxh + 7) / 8 * 8;
  c->rows = c->N / c->Nx;
  c->Nh = c->rows * c->Nxp;
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
  c->max_fpf = 1;
  for (int f = 0; f < c->n; ++f) c->max_fpf = std::max(c->max_fpf, ustart[f + 1] - ustart[f]);
  c->cr_fpw = c->max_fpf;
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
  {  // set-up of this GPU's rdcu_rhs_noalloc scratch address (once per device)
    char *base = nullptr;
    if ((st = scratch_base(device, &base))) return st;
  }
  if (mo) c->stream = mo->stream;
  else HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));

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
  g.Nxp = c->Nxp;
  // streaming stores once a state vector no longer fits the 32 MB of L2
  g.nt_store = (double)c->n * c->N * sizeof(double) >= 32.0 * (1 << 20) ? 1 : 0;
  g.L = c->L;
  g.N = c->N;
  g.rows = c->rows;
  g.Nh = c->Nh;
  g.invN = 1.0 / (double)c->N;
  g.invNxy = c->ndim == 3 ? 1.0 / (double)((long)c->Ns[1] * c->Ns[2]) : g.invN;
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
    // the same dictionary as a program for the register path: per-field term
    // lists (input order kept), records nfac, coef (2 words), (slot, pow)...
    std::vector<int> prog(c->n + 1, 0);
    for (int f = 0; f < c->n; ++f) {
      prog[f] = (int)prog.size();
      for (int i = 0; i < p->nterms; ++i) {
        if (tfield[i] != f) continue;
        const int nfac_i = tstart[i + 1] - tstart[i];
        int lo, hi;
        std::memcpy(&lo, reinterpret_cast<const char *>(&tcoef[i]), 4);
        std::memcpy(&hi, reinterpret_cast<const char *>(&tcoef[i]) + 4, 4);
        prog.push_back(nfac_i);
        prog.push_back(lo);
        prog.push_back(hi);
        for (int j = tstart[i]; j < tstart[i + 1]; ++j) {
          prog.push_back(fslot[j]);
          prog.push_back(fpow[j]);
        }
      }
    }
    prog[c->n] = (int)prog.size();
    bool ok = prog.size() <= 4096;
    int *d_prog = nullptr;
    if ((st = c->upload(&d_prog, prog))) return st;
    c->dictp = rd::DictP{d_prog, (int)prog.size()};
    c->h_tfield = tfield;
    c->h_tstart = tstart;
    c->h_fslot = fslot;
    c->h_fpow = fpow;
    c->h_tcoef = tcoef;
    c->dictk_ok = ok;
    int *ub, *uf, *us, *ab, *af, *as;
    if ((st = c->upload(&ub, ublk)) || (st = c->upload(&uf, ufld)) || (st = c->upload(&us, ustart)) ||
        (st = c->upload(&ab, ablk)) || (st = c->upload(&af, afld)) || (st = c->upload(&as, astart)))
      return st;
    c->fl_used = rd::FacList{c->nf_used, ub, uf, us};
    c->h_ublk = ublk;
    c->h_ufld = ufld;
    c->h_prog = prog;
    c->fl_all = rd::FacList{c->nf_all, ab, af, as};
    std::vector<int> iblk(ablk);
    for (int &v : iblk) v += rd::kImagBlk;
    int *ib;
    if ((st = c->upload(&ib, iblk))) return st;
    c->fl_imag = rd::FacList{c->nf_all, ib, af, as};
    if (c->ndim == 3) {
      if ((st = build_zsplit(c.get(), c->n, ublk, ustart, c->zs_used)) ||
          (st = build_zsplit(c.get(), c->n, ablk, astart, c->zs_all)) ||
          (st = build_zsplit(c.get(), c->n, iblk, astart, c->zs_imag)))
        return st;
      // the z pass splits only the factors that go through it
      c->max_fpf = std::max(1, c->zs_used.zd_max);
      c->cr_fpw = c->max_fpf;
    }
  }
  if ((st = choose_shapes(c.get()))) return st;
  // register-resident path when every transform length allows it
  c->Ex = reg_elems(c->L);
  c->E0 = c->ndim >= 2 ? reg_elems(c->Ns[0]) : 0;
  c->E1 = c->ndim == 3 ? reg_elems(c->Ns[1]) : 0;
  {
    bool ok = c->Ex > 0 && (c->ndim < 2 || c->E0 > 0) && (c->ndim < 3 || c->E1 > 0);
    const int T = c->Ex ? c->L / c->Ex : 0;
    ok = ok && c->n * T <= 256;  // one block of <= 256 threads per row position
    ok = ok && (size_t)rd::rowi_region(c->L, c->Nx, c->n, c->nf_used, c->dictp.nwords, c->n) *
                       sizeof(cx) <=
                   160 * 1024;
    ok = ok && c->dictk_ok;
    if (c->ndim >= 2) ok = ok && c->Ns[0] / c->E0 <= 256 && (size_t)3 * c->Ns[0] * sizeof(cx) <= 150 * 1024;
    if (c->ndim == 3) ok = ok && c->Ns[1] / c->E1 <= 256 && (size_t)2 * c->Ns[1] * sizeof(cx) <= 160 * 1024;
    c->reg = ok ? 1 : 0;
    // RDCU_PATH=lds selects the LDS-resident kernels for every size (the two
    // paths compute identical values; this exists to compare them)
    const char *path = getenv("RDCU_PATH");
    if (path && !strcmp(path, "lds")) c->reg = 0;
    const char *colp = getenv("RDCU_COL");
    if (colp && !strcmp(colp, "lds")) c->col_lds = 1;
    c->ri_G = c->n;
    if (!c->reg && !c->lds_rhs)
      return fail(RDCU_EUNSUPPORTED, "grid lines too long for the kernels of this build");
    if (c->col_lds && !c->lds_y) return fail(RDCU_EUNSUPPORTED, "strided axis too long");
  }

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
    if (mo) {
      c->sh = mo->sh;
    } else {
      void *q = nullptr;
      HIPCHECK(hipHostMalloc(&q, sizeof(rd::DevState), hipHostMallocDefault));
      c->sh = static_cast<rd::DevState *>(q);
    }
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
    if (mo) c->st = mo->st;
    else if ((st = c->dalloc(&c->st, 1))) return st;
    HIPCHECK(hipMemcpy(c->st, c->sh, sizeof(rd::DevState), hipMemcpyHostToDevice));
  }

  if (mo) {  // batch member: the batch tunes, compiles and captures for its classes
    if (!c->reg)
      return fail(RDCU_EUNSUPPORTED,
                  "batched instances need every transform length of the form 2^k or 3*2^k");
    return RDCU_OK;
  }

  if (c->reg) {
    if ((st = once_per_device(device, 0, allow_big_lds))) return st;
    build_jit(c.get());
    const char *shp = getenv("RDCU_SHAPES");  // "rpb,tc,fpw,G": fixed shapes, no tuning
    if (shp && *shp) {
      if ((st = fixed_shapes(c.get(), shp))) return st;
    } else if ((st = autotune(c.get()))) {
      return st;
    }
    const char *vb = getenv("RDCU_VERBOSE");
    if (vb && *vb && *vb != '0')
      fprintf(stderr, "rdcu jit: %s\nrdcu autotune: %s\nrdcu shapes: %d,%d,%d,%d,%d,%d,%d,%d\n",
              c->jit_log.c_str(), c->tune_log.c_str(), c->rf_rpb, c->cr_tc, c->cr_fpw, c->ri_G,
              c->ri_R, c->E0 == 16 ? 16 : 8, c->cr_ntc, c->ri_tw);
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
  if (!c->member) {  // a batch member's attempts run through its batch's graph
    HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    st = RDCU_OK;
    for (int i = 0; i < kGraphAttempts && !st; ++i) st = enqueue_attempt(c.get());
    hipGraph_t gm = nullptr;
    ce = hipStreamEndCapture(c->stream, &gm);
    if (st) return st;
    if (ce != hipSuccess) return fail(RDCU_EDEVICE, std::string("capture: ") + hipGetErrorString(ce));
    c->graph_m = gm;
    HIPCHECK(hipGraphInstantiate(&c->gexec_m, gm, nullptr, nullptr, 0));
  }
  c->nrhs = 0;
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
  if (c->member) return fail(RDCU_ESTATE, "a batch member steps only through its batch");
  HIPCHECK(hipSetDevice(c->dev));
  if (c->fn) return fn_step(c, t, h, hmax, accepted);
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
  if (c->member) return fail(RDCU_ESTATE, "a batch member steps only through its batch");
  HIPCHECK(hipSetDevice(c->dev));
  if (c->fn) return fn_run(c, t, h, hmin, hmax, t1, stop, fn, user, -1, nullptr);
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
    if ((st = launch_attempts(c, K))) return st;
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

int rdcu_advance(rdcu_ctx *c, double *t, double *h, double hmin, double hmax, double t1,
                 long long max_attempts, long long *done) {
  if (!c || !t || !h) return fail(RDCU_EINVAL, "null argument");
  if (c->member) return fail(RDCU_ESTATE, "a batch member steps only through its batch");
  HIPCHECK(hipSetDevice(c->dev));
  if (c->fn) return fn_run(c, t, h, hmin, hmax, t1, INFINITY, nullptr, nullptr, max_attempts, done);
  int st;
  if ((st = begin(c, INFINITY, 1, *t, *h, hmin, hmax, t1))) return st;
  c->k1_valid = false;
  if ((st = ensure_k1(c))) return st;
  long long k = 0;
  double tc = *t, hc = *h;
  while (k < max_attempts && tc < t1 && hc > hmin) {
    long long K = std::min<long long>(kChunkMax, max_attempts - k);
    if ((st = begin(c, INFINITY, 0, 0, 0, hmin, hmax, t1))) return st;
    if ((st = launch_attempts(c, K))) return st;
    if ((st = read_state(c))) return st;
    const int nl = c->sh->nlog;
    c->nrhs += 6L * nl;
    k += nl;
    tc = c->sh->t;
    hc = c->sh->h;
    if (nl < K) break;
  }
  *t = c->sh->t;
  *h = c->sh->h;
  if (done) *done = k;
  return RDCU_OK;
}

int rdcu_time_rhs(rdcu_ctx *c, int reps, double *ms, int max, int *count, char *names,
                  int names_len) {
  if (!c || !ms || max < 1 || reps < 1) return fail(RDCU_EINVAL, "bad argument");
  if (c->fn) return no_dict();
  HIPCHECK(hipSetDevice(c->dev));
  auto ls = c->reg ? rhs_launches_reg(c, rd::SRC_Y, nullptr, rd::DST_PTR, 0, c->rhs_out, 0)
                   : rhs_launches(c, 0, rd::SRC_Y, nullptr, rd::DST_PTR, 0, c->rhs_out, 0, 0);
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  auto timed = [&](const std::function<void()> &body, double *out) -> int {
    for (int r = 0; r < 3; ++r) body();  // warm
    HIPCHECK(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) body();
    HIPCHECK(hipEventRecord(e1, c->stream));
    HIPCHECK(hipEventSynchronize(e1));
    float f = 0;
    HIPCHECK(hipEventElapsedTime(&f, e0, e1));
    *out = (double)f / reps;
    return RDCU_OK;
  };
  std::string nm = "rhs";
  int cnt = 0, st = RDCU_OK;
  st = timed([&]() { for (auto &l : ls) l.go(); }, &ms[cnt++]);
  for (size_t i = 0; !st && i < ls.size() && cnt < max; ++i) {
    st = timed(ls[i].go, &ms[cnt++]);
    nm += std::string(",") + ls[i].name;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) return st;
  HIPCHECK(hipGetLastError());
  if (count) *count = cnt;
  if (names && names_len > 0) snprintf(names, names_len, "%s", nm.c_str());
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

// RHS of the device state y_dev into c->rhs_out and, if Y_out is wanted,
// every factor block packed as complex values into c->Ycx; then the copies
// to the host (asynchronous on the library's stream; the caller syncs).
int enqueue_rhs_out(rdcu_ctx *c, const double *y_dev, double *f_out, double *Y_out) {
  const long N = c->N, nN = (long)c->n * N;
  int st = enqueue_rhs(c, 0, rd::SRC_PTR, y_dev, rd::DST_PTR, 0, c->rhs_out, 0, 0);
  if (st) return st;
  if (f_out)
    HIPCHECK(hipMemcpyAsync(f_out, c->rhs_out, nN * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (Y_out) {
    if (!c->Yreal && (st = c->dalloc(&c->Yreal, (size_t)std::max(1, c->nf_all) * N))) return st;
    const size_t ny = 2 * (size_t)c->nblk * nN;
    if (!c->Ycx && (st = c->dalloc(&c->Ycx, ny))) return st;
    const dim3 pg((unsigned)((N + 255) / 256), (unsigned)(c->nblk * c->n));
    if ((st = enqueue_all_factors(c, c->fl_all, c->zs_all, c->Yreal))) return st;
    hipLaunchKernelGGL(rd::k_pack_y, pg, dim3(256), 0, c->stream, y_dev, c->Yreal,
                       (double2 *)c->Ycx, c->n, c->nblk, N);
    // imaginary parts (Nyquist leakage), through the same buffer
    if (c->nf_all > 0) {
      if ((st = enqueue_all_factors(c, c->fl_imag, c->zs_imag, c->Yreal))) return st;
      hipLaunchKernelGGL(rd::k_pack_y_imag, pg, dim3(256), 0, c->stream, c->Yreal,
                         (double2 *)c->Ycx, c->n, c->nblk, c->ndim, N);
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(Y_out, c->Ycx, ny * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  return RDCU_OK;
}

int rdcu_rhs(rdcu_ctx *c, const double *y, double *f_out, double *Y_out) {
  if (!c || !y) return fail(RDCU_EINVAL, "null argument");
  if (c->fn) return no_dict();
  if (Y_out && !c->lds_y) return fail(RDCU_EUNSUPPORTED, "factor output not available for lines this long");
  HIPCHECK(hipSetDevice(c->dev));
  const long nN = (long)c->n * c->N;
  HIPCHECK(hipMemcpyAsync(c->rhs_in, y, nN * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int st = enqueue_rhs_out(c, c->rhs_in, f_out, Y_out);
  if (st) return st;
  HIPCHECK(hipStreamSynchronize(c->stream));
  return RDCU_OK;
}

int rdcu_rhs_noalloc(rdcu_ctx *c, const double *y, double *f_out, double *Y_out) {
  if (!c || !y) return fail(RDCU_EINVAL, "null argument");
  if (c->fn) return no_dict();
  if (Y_out && !c->lds_y) return fail(RDCU_EUNSUPPORTED, "factor output not available for lines this long");
  HIPCHECK(hipSetDevice(c->dev));
  char *base = nullptr;
  int st = scratch_base(c->dev, &base);
  if (st) return st;
  // every intermediate buffer of rdcu_rhs, carved out of the scratch array
  const long nN = (long)c->n * c->N;
  const int nfac = Y_out ? c->nf_all : c->nf_used;
  const size_t b_in = round_up256(nN * sizeof(double));
  const size_t b_S = round_up256((size_t)c->n * c->Nh * sizeof(cx));
  const size_t b_F = round_up256(c->ndim >= 2 ? (size_t)std::max(1, nfac) * c->Nh * sizeof(cx) : 16);
  const size_t b_out = b_in;
  const size_t b_Yr = Y_out ? round_up256((size_t)std::max(1, c->nf_all) * c->N * sizeof(double)) : 0;
  const size_t b_Yc = Y_out ? round_up256(2 * (size_t)c->nblk * nN * sizeof(double)) : 0;
  const size_t need = b_in + b_S + b_F + b_out + b_Yr + b_Yc;
  if (need > kScratchBytes)
    return fail(RDCU_EINVAL, "rdcu_rhs_noalloc: this problem needs " + std::to_string(need) +
                                 " bytes of device scratch, more than the " +
                                 std::to_string(kScratchBytes) + " the library holds per GPU");
  std::lock_guard<std::mutex> lk(scratch_mutex(c->dev));
  // the context's own buffers are swapped out for the call, then restored;
  // on every return, an error one included, the stream is drained first, so
  // that no work reading or writing the array is in flight once the lock is
  // released (the guard is destroyed after this)
  struct Swap {
    rdcu_ctx *c;
    rd::Bufs b;
    double *in, *out, *Yr, *Yc;
    ~Swap() {
      (void)hipStreamSynchronize(c->stream);
      c->bufs = b;
      c->rhs_in = in;
      c->rhs_out = out;
      c->Yreal = Yr;
      c->Ycx = Yc;
    }
  } keep{c, c->bufs, c->rhs_in, c->rhs_out, c->Yreal, c->Ycx};
  char *p = base;
  c->rhs_in = reinterpret_cast<double *>(p);
  p += b_in;
  c->bufs.S = reinterpret_cast<cx *>(p);
  p += b_S;
  c->bufs.F = reinterpret_cast<cx *>(p);
  p += b_F;
  c->rhs_out = reinterpret_cast<double *>(p);
  p += b_out;
  c->Yreal = Y_out ? reinterpret_cast<double *>(p) : nullptr;
  p += b_Yr;
  c->Ycx = Y_out ? reinterpret_cast<double *>(p) : nullptr;
  HIPCHECK(hipMemcpyAsync(c->rhs_in, y, nN * sizeof(double), hipMemcpyHostToDevice, c->stream));
  st = enqueue_rhs_out(c, c->rhs_in, f_out, Y_out);
  if (st) return st;
  HIPCHECK(hipStreamSynchronize(c->stream));
  return RDCU_OK;
}

int rdcu_frame(rdcu_ctx *c, double t_eval, double *y_out, double *f_out, double *Y_out) {
  if (!c || !y_out) return fail(RDCU_EINVAL, "null argument");
  if (c->fn && (f_out || Y_out)) return no_dict();
  if (!c->sh->eval_valid) return fail(RDCU_ESTATE, "no accepted step to interpolate");
  if (Y_out && !c->lds_y) return fail(RDCU_EUNSUPPORTED, "factor output not available for lines this long");
  HIPCHECK(hipSetDevice(c->dev));
  rd::DenseCoef dc;
  for (int j = 0; j < 7; ++j)
    for (int q = 0; q < 4; ++q) dc.p[j][q] = PP[j][q];
  const long nN = (long)c->n * c->N;
  hipLaunchKernelGGL(rd::k_dense, dim3((unsigned)((nN + 255) / 256)), dim3(256), 0, c->stream,
                     c->st, c->bufs, dc, t_eval, c->yeval, nN);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(y_out, c->yeval, nN * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (f_out || Y_out) {
    int st = enqueue_rhs_out(c, c->yeval, f_out, Y_out);
    if (st) return st;
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  return RDCU_OK;
}

int rdcu_host_alloc(size_t bytes, void **out) {
  if (!out) return fail(RDCU_EINVAL, "null argument");
  *out = nullptr;
  hipError_t e = hipHostMalloc(out, bytes ? bytes : 16, hipHostMallocDefault);
  if (e != hipSuccess) {
    *out = nullptr;
    return fail(RDCU_ENOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  }
  return RDCU_OK;
}

int rdcu_host_free(void *p) {
  if (p) HIPCHECK(hipHostFree(p));
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

#include "rd_batch.hip"
#include "rd_adjoint.hip"
#include "rd_callback.hip"
