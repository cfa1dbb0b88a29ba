// rd_adjoint.hip -- coefficient updates, step schedules and the discrete
// adjoint of DP45 on a frozen schedule (included at the end of
// rd_runtime.hip: same translation unit).
//
// Forward: the attempt graph with the device step control switched to fixed
// steps whose sizes come from the schedule; the start state of every step is
// saved (checkpoints), and each output is the dense output of the step that
// reaches it.  Reverse, step by step from the last: the seven stages are
// recomputed from the checkpoint, then the stage adjoints are accumulated in
// reverse stage order through vector-Jacobian products of the RHS
// (k_rowi_vjp + the derivative blocks applied by a companion problem).

namespace {

bool adj_store_enabled() {
  const char *e = getenv("RDCU_ADJ_STORE");
  return !(e && !strcmp(e, "0"));
}

// RAII list of device allocations of one call, ordered on the call's stream
// (stream-ordered allocation: no device-wide synchronisation, so calls on
// other streams keep running; plain hipMalloc where pools are unsupported).
struct DevScratch {
  hipStream_t s;
  std::vector<std::pair<void *, bool>> p;  // (pointer, stream-ordered)
  explicit DevScratch(hipStream_t st) : s(st) {}
  ~DevScratch() {
    for (auto &q : p) (void)(q.second ? hipFreeAsync(q.first, s) : hipFree(q.first));
  }
  template <class T>
  int get(T **out, size_t count) {
    void *q = nullptr;
    const size_t bytes = std::max<size_t>(1, count) * sizeof(T);
    bool async = hipMallocAsync(&q, bytes, s) == hipSuccess;
    if (!async) {
      (void)hipGetLastError();
      hipError_t e = hipMalloc(&q, bytes);
      if (e != hipSuccess)
        return fail(RDCU_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    p.push_back({q, async});
    *out = static_cast<T *>(q);
    return RDCU_OK;
  }
};

unsigned blocks_of(long n) { return (unsigned)((n + 255) / 256); }

int place_outputs(double t0, long long ns, const double *hs, int nout, const double *t_out,
                  std::vector<long long> &ostep, std::vector<double> &tn) {
  tn.assign(ns + 1, t0);
  for (long long k = 0; k < ns; ++k) tn[k + 1] = tn[k] + hs[k];
  ostep.assign(nout, -1);
  long long k = 0;
  for (int o = 0; o < nout; ++o) {
    if (o > 0 && t_out[o] < t_out[o - 1]) return fail(RDCU_EINVAL, "output times must ascend");
    if (!(t_out[o] > t0)) return fail(RDCU_EINVAL, "output time not after t0");
    while (k < ns && tn[k + 1] < t_out[o]) ++k;
    if (k >= ns) return fail(RDCU_EINVAL, "output time beyond the schedule");
    ostep[o] = k;
  }
  return RDCU_OK;
}

// Straight-line per-term work of the vector-Jacobian row kernel for one
// dictionary structure: the operations of the interpreter in rowi_vjp_body
// (per point: fields ascending, terms in input order, factors in order;
// powers as repeated products; term partials reduced per wave), with the
// slot values, weights and per-term partials in registers.
std::string jit_vjp_source(int n, int nslots, const std::vector<int> &tfield,
                           const std::vector<int> &tstart, const std::vector<int> &fslot,
                           const std::vector<int> &fpow) {
  auto powexpr = [](int sl, int p) {
    const std::string v = "v" + std::to_string(sl);
    if (p == 0) return std::string("1.0");
    const int a = p < 0 ? -p : p;
    std::string e = v;
    for (int j = 1; j < a; ++j) e = "(" + e + " * " + v + ")";
    return p < 0 ? "(1.0 / " + e + ")" : e;
  };
  const int nt = (int)tfield.size();
  std::vector<int> order, coff;  // program order of the terms; coefficient word offsets
  {
    int pos = n + 1;
    for (int f = 0; f < n; ++f)
      for (int i = 0; i < nt; ++i)
        if (tfield[i] == f) {
          order.push_back(i);
          coff.push_back(pos + 1);
          pos += 3 + 2 * (tstart[i + 1] - tstart[i]);
        }
  }
  std::vector<char> used(nslots, 0), grad(nslots, 0);
  for (int i : order)
    for (int j = tstart[i]; j < tstart[i + 1]; ++j) {
      used[fslot[j]] = 1;
      if (fpow[j] != 0) grad[fslot[j]] = 1;
    }
  char buf[256];
  std::string s = "namespace rd {\nstruct JitVjp {\n  static constexpr bool kGen = true;\n"
                  "  __device__ __forceinline__ static void run(const int *__restrict__ P, "
                  "const double *vals, const double *wrow, double *gacc, double *red, int Nx, "
                  "int tid, int nt, int nwv) {\n";
  for (size_t q = 0; q < order.size(); ++q) {
    snprintf(buf, sizeof buf, "    const double c%zu = __hiloint2double(P[%d], P[%d]);\n    double a%zu = 0.0;\n",
             q, coff[q] + 1, coff[q], q);
    s += buf;
  }
  s += "    for (int x = tid; x < Nx; x += nt) {\n";
  for (int sl = 0; sl < nslots; ++sl) {
    if (used[sl]) {
      snprintf(buf, sizeof buf, "      const double v%d = vals[%d * Nx + x];\n", sl, sl);
      s += buf;
    }
    if (grad[sl]) {
      snprintf(buf, sizeof buf, "      double g%d = 0.0;\n", sl);
      s += buf;
    }
  }
  for (int f = 0, q = 0; f < n; ++f) {
    snprintf(buf, sizeof buf, "      {\n        const double w = wrow[%d * Nx + x];\n", f);
    s += buf;
    for (; q < (int)order.size() && tfield[order[q]] == f; ++q) {
      const int i = order[q];
      s += "        { double t = 1.0;";
      for (int j = tstart[i]; j < tstart[i + 1]; ++j) s += " t = t * " + powexpr(fslot[j], fpow[j]) + ";";
      snprintf(buf, sizeof buf, " a%d = a%d + w * t;\n", q, q);
      s += buf;
      for (int j = tstart[i]; j < tstart[i + 1]; ++j) {
        if (fpow[j] == 0) continue;
        snprintf(buf, sizeof buf, "          { double d = %d.0 * %s;", fpow[j],
                 powexpr(fslot[j], fpow[j] - 1).c_str());
        s += buf;
        for (int k = tstart[i]; k < tstart[i + 1]; ++k)
          if (k != j) s += " d = d * " + powexpr(fslot[k], fpow[k]) + ";";
        snprintf(buf, sizeof buf, " g%d = g%d + (w * c%d) * d; }\n", fslot[j], fslot[j], q);
        s += buf;
      }
      s += "        }\n";
    }
    s += "      }\n";
  }
  for (int sl = 0; sl < nslots; ++sl)
    if (grad[sl]) {
      snprintf(buf, sizeof buf, "      gacc[%d * Nx + x] = g%d;\n", sl, sl);
      s += buf;
    }
  s += "    }\n";
  for (size_t q = 0; q < order.size(); ++q) {
    snprintf(buf, sizeof buf,
             "    { const double r_ = wave_sum(a%zu); if ((tid & 63) == 0) red[%zu * nwv + (tid >> 6)] = r_; }\n",
             q, q);
    s += buf;
  }
  s += "  }\n};\n}  // namespace rd\n";
  return s;
}

// Compiles the vector-Jacobian row kernel for c's dictionary structure
// (once per handle, on the first gradient call).  Without it (hiprtc
// disabled or failing, or a kernel that would spill) the interpreter runs.
void build_jit_vjp(rdcu_ctx *c) {
  if (c->jit_vjp || !jit_enabled()) return;
  std::string src = "#include \"rd_adjoint.h\"\n" +
                    jit_vjp_source(c->n, c->n + c->nf_used, c->h_tfield, c->h_tstart, c->h_fslot,
                                   c->h_fpow);
  src += "extern \"C\" __global__ __launch_bounds__(256) void rdcu_rowi_vjp_jit(rd::Grid g, "
         "rd::FftPlan px, const rd::cx *__restrict__ twN, rd::Bufs b, rd::FacList fl, rd::DictP dict, "
         "const int *__restrict__ tmap, const double *__restrict__ u, const double *__restrict__ W, "
         "double *out_id, double *Gk, double *P, int G, rd::cx *Spc) {\n"
         "  rd::rowi_vjp_body<" + std::to_string(c->Ex) + ", rd::JitVjp, " +
         (odd_len(c->L) ? "-1" : "0") + ">(g, px, twN, b, fl, dict, "
         "tmap, u, W, out_id, Gk, P, G, Spc);\n}\n";
  std::vector<char> code;
  std::string log;
  const char *vb = getenv("RDCU_VERBOSE");
  if (!jit_compile(src, jit_arch(c->dev), code, log)) {
    if (vb && *vb) fprintf(stderr, "rdcu vjp jit: %s\n", log.c_str());
    return;
  }
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  if (hipModuleLoadData(&mod, code.data()) != hipSuccess ||
      hipModuleGetFunction(&fn, mod, "rdcu_rowi_vjp_jit") != hipSuccess || jit_spills(fn)) {
    (void)hipGetLastError();
    if (mod) (void)hipModuleUnload(mod);
    if (vb && *vb) fprintf(stderr, "rdcu vjp jit: not used (load failed or spills)\n");
    return;
  }
  c->jit_vjp_mod = mod;
  c->jit_vjp = fn;
  if (vb && *vb) fprintf(stderr, "rdcu vjp jit: ok\n");
}

// Companion problem: one field per referenced factor m of c, whose RHS is
// sign_m * D_m applied to it (sign -1 for first derivatives: D^T = -D).
int adj_companion(rdcu_ctx *c) {
  if (c->adj_pc) return RDCU_OK;
  build_jit_vjp(c);
  const int nf = c->nf_used;
  if (nf > rd::kNMax)
    return fail(RDCU_EUNSUPPORTED, "gradients need at most 8 referenced derivative factors");
  std::vector<double> coef(nf);
  std::vector<int> field(nf), nfac(nf, 1), fac(2 * nf);
  for (int m = 0; m < nf; ++m) {
    const int blk = c->h_ublk[m];
    coef[m] = blk <= c->ndim ? -1.0 : 1.0;
    field[m] = m;
    fac[2 * m] = blk * nf + m;
    fac[2 * m + 1] = 1;
  }
  rdcu_problem pp{};
  pp.n = nf;
  pp.ndim = c->ndim;
  for (int a = 0; a < 3; ++a) {
    pp.Ns[a] = c->Ns[a];
    pp.Ls[a] = c->Ls[a];
  }
  pp.nterms = nf;
  pp.coef = coef.data();
  pp.field = field.data();
  pp.nfac = nfac.data();
  pp.fac = fac.data();
  std::unique_ptr<rdcu_ctx> pc;
  int st = create_ctx(&pp, 1e-8, 1e-8, 1, nullptr, c->dev, nullptr, pc);
  if (st) return st;
  HIPCHECK(hipStreamSynchronize(pc->stream));
  HIPCHECK(hipStreamDestroy(pc->stream));
  pc->stream = c->stream;
  pc->borrowed_stream = true;
  c->adj_pc = std::move(pc);
  return RDCU_OK;
}

// J(u)^T w -> out_id (identity slots) and Dk (derivative blocks, signed);
// coefficient partials added into P.
// pre: u's x-forward half spectra and factor spectra already computed
// (S, F of *pre); null: computed here into c's buffers.
int vjp(rdcu_ctx *c, const double *u, const double *w, double *out_id, double *Gk, double *Dk,
        double *P, const int *tmap, const rd::Bufs *pre = nullptr) {
  rd::StageCoef none{};
  if (!pre) {
    rowf(c, 0, rd::SRC_PTR, u, none, c->rf_rpb);
    for (auto &l : col_launches_reg(c, 0, c->cr_tc, c->cr_ntc)) l.go();
  }
  const rd::Bufs bv = pre ? *pre : c->bufs;
  const int nf = c->nf_used, T = c->L / c->Ex;
  // all 256 threads when the line groups divide them (idle lines, but the
  // per-term loop over the row's points gets the most threads)
  const int G = (256 % T == 0 && 256 / T >= nf) ? 256 / T : std::max(1, std::min(nf, 256 / T));
  const int nt = G * T;
  const int nterms = (int)c->h_tcoef.size();
  const size_t nred = std::max<size_t>(nt, (nt % 64 == 0) ? (size_t)nterms * (nt / 64) : 0);
  const size_t lds = (size_t)G * rd::padded_len(c->L) * sizeof(cx) +
                     (size_t)(2 * (c->n + nf) + c->n) * c->Nx * sizeof(double) +
                     (size_t)((c->dictp.nwords + 1) & ~1) * sizeof(int) + nred * sizeof(double);
  if (lds > 160 * 1024) return fail(RDCU_EUNSUPPORTED, "row too long for the adjoint kernel");
  // the companion's x forward pass runs inside the row kernel (its S)
  rdcu_ctx *pc = nf > 0 ? c->adj_pc.get() : nullptr;
  cx *spc = pc && pc->reg && c->Ex == 8 ? pc->bufs.S : nullptr;
  if (c->jit_vjp && nt % 64 == 0) {
    rd::Grid a_g = c->grid;
    rd::FftPlan a_px = c->px;
    const rd::cx *a_tw = c->twN;
    rd::Bufs a_b = bv;
    rd::FacList a_fl = c->fl_used;
    rd::DictP a_d = c->dictp;
    const int *a_tm = tmap;
    const double *a_u = u, *a_w = w;
    double *a_oid = out_id, *a_gk = Gk, *a_p = P;
    int a_G = G;
    cx *a_s = spc;
    void *args[] = {&a_g, &a_px, &a_tw, &a_b, &a_fl, &a_d, &a_tm, &a_u, &a_w, &a_oid, &a_gk, &a_p, &a_G, &a_s};
    HIPCHECK(hipModuleLaunchKernel(c->jit_vjp, (unsigned)c->rows, 1, 1, (unsigned)nt, 1, 1,
                                   (unsigned)lds, c->stream, args, nullptr));
  } else if (odd_len(c->L))
    hipLaunchKernelGGL((rd::k_rowi_vjp<8, -1>), dim3((unsigned)c->rows), dim3(nt), lds, c->stream,
                       c->grid, c->px, c->twN, bv, c->fl_used, c->dictp, tmap, u, w, out_id,
                       Gk, P, G, spc);
  else
    hipLaunchKernelGGL(rd::k_rowi_vjp<8>, dim3((unsigned)c->rows), dim3(nt), lds, c->stream,
                       c->grid, c->px, c->twN, bv, c->fl_used, c->dictp, tmap, u, w, out_id,
                       Gk, P, G, spc);
  HIPCHECK(hipGetLastError());
  if (pc && !spc) {
    int st = enqueue_rhs(pc, 0, rd::SRC_PTR, Gk, rd::DST_PTR, 0, Dk, 0, 0);
    if (st) return st;
  } else if (pc) {  // the companion's strided and x-inverse passes
    auto ls = rhs_launches_reg(pc, rd::SRC_PTR, Gk, rd::DST_PTR, 0, Dk, 0);
    if (ls.empty() || std::string(ls[0].name) != "row_fwd")
      return fail(RDCU_EDEVICE, "companion pipeline does not start with its x forward pass");
    for (size_t i = 1; i < ls.size(); ++i) ls[i].go();
    HIPCHECK(hipGetLastError());
    pc->nrhs++;
  }
  return RDCU_OK;
}

}  // namespace

extern "C" {

int rdcu_set_coef(rdcu_ctx *c, const double *coef) {
  if (!c || (!coef && !c->h_tcoef.empty())) return fail(RDCU_EINVAL, "null argument");
  if (c->fn) return no_dict();
  HIPCHECK(hipSetDevice(c->dev));
  const int nt = (int)c->h_tcoef.size();
  for (int i = 0; i < nt; ++i) c->h_tcoef[i] = coef[i];
  // coefficient words of the program (layout of DictP: fields' term lists)
  std::vector<int> &prog = c->h_prog;
  int pos = c->n + 1;
  for (int f = 0; f < c->n; ++f)
    for (int i = 0; i < nt; ++i)
      if (c->h_tfield[i] == f) {
        std::memcpy(&prog[pos + 1], &c->h_tcoef[i], sizeof(double));
        pos += 3 + 2 * (c->h_tstart[i + 1] - c->h_tstart[i]);
      }
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipMemcpy(const_cast<int *>(c->dictp.prog), prog.data(), prog.size() * sizeof(int),
                     hipMemcpyHostToDevice));
  if (nt)
    HIPCHECK(hipMemcpy(const_cast<double *>(c->dict.tcoef), c->h_tcoef.data(), nt * sizeof(double),
                       hipMemcpyHostToDevice));
  c->k1_valid = false;
  return RDCU_OK;
}

int rdcu_schedule(rdcu_ctx *c, double t0, double h0, double hmin, double hmax, double t1,
                  long long max_steps, double *hs, long long *nsteps) {
  if (!c || !hs || !nsteps) return fail(RDCU_EINVAL, "null argument");
  if (c->fn) return no_dict();
  if (c->member) return fail(RDCU_ESTATE, "a batch member steps only through its batch");
  HIPCHECK(hipSetDevice(c->dev));
  int st;
  if ((st = begin(c, INFINITY, 1, t0, h0, hmin, hmax, t1))) return st;
  c->k1_valid = false;
  if ((st = ensure_k1(c))) return st;
  long long k = 0;
  double tc = t0, hc = h0;
  while (tc < t1 && hc > hmin) {
    const long long K = kChunkMax;
    if ((st = begin(c, INFINITY, 0, 0, 0, hmin, hmax, t1))) return st;
    if ((st = launch_attempts(c, K))) return st;
    if ((st = read_state(c))) return st;
    const int nl = c->sh->nlog;
    c->nrhs += 6L * nl;
    for (int i = 0; i < nl; ++i) {
      const double hu = i == 0 ? hc : c->sh->log_h[i - 1];  // size the attempt used
      if (!c->sh->log_acc[i]) continue;
      if (k >= max_steps) return fail(RDCU_EINVAL, "more accepted steps than max_steps");
      hs[k++] = hu;
    }
    tc = c->sh->t;
    hc = c->sh->h;
    if (nl < K) break;
  }
  *nsteps = k;
  return RDCU_OK;
}

}  // extern "C"

namespace {

// rdcu_loss_grad with y0, target and grad_y0 in host memory (dev = false) or
// device memory (dev = true: target is read in place, y0 and the gradient
// are copied device to device).
int loss_grad(rdcu_ctx *c, const double *y0, double t0, long long ns, const double *hs, int nout,
              const double *t_out, const double *target, double *loss, double *grad_coef,
              double *grad_y0, bool dev) {
  if (!c || !y0 || !hs || !t_out || !target || !loss) return fail(RDCU_EINVAL, "null argument");
  if (c->fn) return no_dict();
  if (c->member) return fail(RDCU_ESTATE, "a batch member steps only through its batch");
  if (ns < 1 || nout < 1) return fail(RDCU_EINVAL, "need at least one step and one output");
  if (!c->reg) return fail(RDCU_EUNSUPPORTED, "gradients need transform lengths 2^k or 3*2^k");
  HIPCHECK(hipSetDevice(c->dev));
  std::vector<long long> ostep;
  std::vector<double> tn;
  int st = place_outputs(t0, ns, hs, nout, t_out, ostep, tn);
  if (st) return st;
  const bool grad = grad_coef || grad_y0;
  if (grad && (st = adj_companion(c))) return st;
  const long N = c->N, nN = (long)c->n * N;
  const double cnt = (double)nout * (double)nN;
  const unsigned nbl = blocks_of(nN);
  DevScratch ds(c->stream);
  double *d_hs, *d_tg, *d_part, *ych = nullptr, *gbar = nullptr;
  if ((st = ds.get(&d_hs, ns)) || (st = ds.get(&d_part, (size_t)nout * nbl))) return st;
  if (dev) d_tg = const_cast<double *>(target);
  else if ((st = ds.get(&d_tg, (size_t)nout * nN))) return st;
  if (grad && ((st = ds.get(&ych, (size_t)ns * nN)) || (st = ds.get(&gbar, (size_t)nout * nN))))
    return st;
  // The stage values of every step are kept from the forward pass when they
  // fit in a quarter of the free device memory (RDCU_ADJ_STORE=0: never);
  // otherwise the reverse sweep recomputes them from the checkpoints.  Both
  // give the same bits.
  double *kst = nullptr;
  if (grad && adj_store_enabled()) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess &&
        (double)7 * ns * nN * sizeof(double) <= 0.25 * (double)fr &&
        (st = ds.get(&kst, (size_t)7 * ns * nN)))
      return st;
  }
  HIPCHECK(hipMemcpyAsync(d_hs, hs, ns * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (!dev)
    HIPCHECK(hipMemcpyAsync(d_tg, target, (size_t)nout * nN * sizeof(double),
                            hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->bufs.y[c->sh->iy], y0, nN * sizeof(double),
                          dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
  c->k1_valid = false;
  c->sh->eval_valid = 0;

  // ---- forward on the frozen schedule
  hipLaunchKernelGGL(rd::k_set_sched, dim3(1), dim3(1), 0, c->stream, c->st, 1, d_hs, ns, 1);
  if ((st = begin(c, INFINITY, 1, t0, hs[0], -INFINITY, INFINITY, INFINITY))) return st;
  hipLaunchKernelGGL(rd::k_set_sched, dim3(1), dim3(1), 0, c->stream, c->st, 1, d_hs, ns, 1);
  c->k1_valid = false;
  if ((st = ensure_k1(c))) return st;
  rd::DenseCoef dc;
  for (int j = 0; j < 7; ++j)
    for (int q = 0; q < 4; ++q) dc.p[j][q] = PP[j][q];
  // with RDCU_VERBOSE: HIP-event times of the forward pass and the reverse
  // sweep, and the kernels of one reverse step (stderr)
  const char *vb = getenv("RDCU_VERBOSE");
  struct PhaseEvents {
    hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
    ~PhaseEvents() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } pev;
  const bool timing = vb && *vb && *vb != '0';
  if (timing)
    for (hipEvent_t &x : pev.e) HIPCHECK(hipEventCreate(&x));
  size_t step_nodes = 0;
  if (timing) HIPCHECK(hipEventRecord(pev.e[0], c->stream));
  int o = 0;
  for (long long k = 0; k < ns; ++k) {
    if (grad)
      hipLaunchKernelGGL(rd::k_save_state, dim3(blocks_of(nN)), dim3(256), 0, c->stream, c->st,
                         c->bufs, ych + (size_t)k * nN, nN);
    HIPCHECK(hipGraphLaunch(c->gexec, c->stream));
    if (kst)
      hipLaunchKernelGGL(rd::k_save_stages, dim3(blocks_of(nN)), dim3(256), 0, c->stream, c->st,
                         c->bufs, kst + (size_t)k * 7 * nN, nN);
    for (; o < nout && ostep[o] == k; ++o) {
      hipLaunchKernelGGL(rd::k_dense, dim3(blocks_of(nN)), dim3(256), 0, c->stream, c->st, c->bufs,
                         dc, t_out[o], c->yeval, nN);
      hipLaunchKernelGGL(rd::k_loss, dim3(nbl), dim3(256), 0, c->stream, c->yeval,
                         d_tg + (size_t)o * nN, cnt, grad ? gbar + (size_t)o * nN : nullptr,
                         d_part + (size_t)o * nbl, nN);
    }
  }
  HIPCHECK(hipGetLastError());
  if (timing) HIPCHECK(hipEventRecord(pev.e[1], c->stream));
  c->nrhs += 6 * ns;
  hipLaunchKernelGGL(rd::k_set_sched, dim3(1), dim3(1), 0, c->stream, c->st, c->fixed, nullptr,
                     0LL, 0);
  std::vector<double> part((size_t)nout * nbl);
  HIPCHECK(hipMemcpyAsync(part.data(), d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  if ((st = read_state(c))) return st;
  c->k1_valid = false;
  double L = 0.0;
  for (double v : part) L += v;
  *loss = L / cnt;
  if (!grad) return RDCU_OK;

  // ---- reverse sweep
  const int nf = c->nf_used, nt = (int)c->h_tcoef.size();
  double *K[7], *Ys[7], *kbar[7], *ybar, *lam, *out_id, *Gk, *Dk, *P;
  int *tmap, *fmap;
  // the seven stage inputs side by side (stage s at Ys[s-1]; Ys[0] is the
  // checkpoint's copy): their spectra are computed together, as 7n fields
  double *Yall;
  if ((st = ds.get(&Yall, (size_t)7 * nN))) return st;
  for (int j = 0; j < 7; ++j) {
    Ys[j] = Yall + (size_t)j * nN;
    if ((!kst && (st = ds.get(&K[j], nN))) || (st = ds.get(&kbar[j], nN))) return st;
  }
  if ((st = ds.get(&ybar, nN)) || (st = ds.get(&lam, nN)) || (st = ds.get(&out_id, nN)) ||
      (st = ds.get(&Gk, (size_t)nf * N)) || (st = ds.get(&Dk, (size_t)nf * N)) ||
      (st = ds.get(&P, (size_t)std::max(1, nt) * c->rows)) || (st = ds.get(&tmap, nt)) ||
      (st = ds.get(&fmap, nf)))
    return st;
  {
    std::vector<int> tm;
    for (int f = 0; f < c->n; ++f)
      for (int i = 0; i < nt; ++i)
        if (c->h_tfield[i] == f) tm.push_back(i);
    if (nt) HIPCHECK(hipMemcpyAsync(tmap, tm.data(), nt * sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (nf)
      HIPCHECK(hipMemcpyAsync(fmap, c->h_ufld.data(), nf * sizeof(int), hipMemcpyHostToDevice,
                              c->stream));
  }
  HIPCHECK(hipMemsetAsync(lam, 0, nN * sizeof(double), c->stream));
  HIPCHECK(hipMemsetAsync(P, 0, (size_t)std::max(1, nt) * c->rows * sizeof(double), c->stream));
  // Outputs of each step (contiguous in o) and their dense-output weights
  // c_oj = hh * sum_p P_jp x^p (dp45.cu:70-76), computed here as before.
  std::vector<int> obeg(ns, 0), oend(ns, 0);
  std::vector<double> oc((size_t)7 * nout, 0.0);
  for (int q = 0; q < nout; ++q) {
    const long long k = ostep[q];
    if (oend[k] == 0) obeg[k] = q;
    oend[k] = q + 1;
    const double hh = tn[k + 1] - tn[k];
    const double x = (t_out[q] - tn[k]) / hh;
    const double xp[4] = {x, x * x, x * x * x, x * x * x * x};
    for (int j = 0; j < 7; ++j) {
      double cj = 0.0;
      for (int pw = 0; pw < 4; ++pw) cj += PP[j][pw] * xp[pw];
      oc[(size_t)7 * q + j] = hh * cj;
    }
  }
  int *d_obeg, *d_oend;
  double *d_oc, *ycur;
  long long *kdev;
  if ((st = ds.get(&d_obeg, ns)) || (st = ds.get(&d_oend, ns)) ||
      (st = ds.get(&d_oc, (size_t)7 * nout)) || (st = ds.get(&kdev, 1)))
    return st;
  ycur = Ys[0];
  // Spectra of the 7 stage inputs in one x-forward and one strided pipeline:
  // the problem seen as 7n fields (field s*n+f = field f of stage s+1) with
  // the factor list repeated per stage; stage s's spectra at offset s-1.
  cx *S7, *F7;
  int *fl7b, *fl7f, *fl7s;
  const int nf7 = 7 * nf;
  std::vector<int> h7b(std::max(1, nf7)), h7f(std::max(1, nf7)), h7s(7 * c->n + 1);
  {
    std::vector<int> ust(c->n + 1, 0);
    for (int m = 0; m < nf; ++m) ust[c->h_ufld[m] + 1]++;
    for (int f = 0; f < c->n; ++f) ust[f + 1] += ust[f];
    for (int q = 0; q < 7; ++q) {
      for (int m = 0; m < nf; ++m) {
        h7b[q * nf + m] = c->h_ublk[m];
        h7f[q * nf + m] = q * c->n + c->h_ufld[m];
      }
      for (int f = 0; f < c->n; ++f) h7s[q * c->n + f] = q * nf + ust[f];
    }
    h7s[7 * c->n] = nf7;
  }
  if ((st = ds.get(&S7, (size_t)7 * c->n * c->Nh)) ||
      (st = ds.get(&F7, (size_t)std::max(1, nf7) * c->Nh)) || (st = ds.get(&fl7b, h7b.size())) ||
      (st = ds.get(&fl7f, h7f.size())) || (st = ds.get(&fl7s, h7s.size())))
    return st;
  HIPCHECK(hipMemcpyAsync(fl7b, h7b.data(), h7b.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(fl7f, h7f.data(), h7f.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(fl7s, h7s.data(), h7s.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  const bool batched_spectra = c->ndim >= 2;  // in 1-D the row kernel reads S only
  // 3-D: the repeated list's split (factors without k_z), kept by the handle
  if (c->ndim == 3 && !c->zs7.zd.fstart &&
      (st = build_zsplit(c, 7 * c->n, h7b, h7s, c->zs7)))
    return st;
  auto enqueue_spectra7 = [&]() {
    const rd::Grid g0 = c->grid;
    const rd::Bufs b0 = c->bufs;
    const rd::FacList f0 = c->fl_used;
    const ZSplit z0 = c->zs_used;
    const int n0 = c->n, nf0 = c->nf_used;
    c->grid.n = 7 * n0;
    c->n = 7 * n0;
    c->nf_used = nf7;
    c->fl_used = rd::FacList{nf7, fl7b, fl7f, fl7s};
    c->zs_used = c->zs7;
    c->bufs.S = S7;
    c->bufs.F = F7;
    rd::StageCoef none{};
    rowf(c, 0, rd::SRC_PTR, Yall, none, c->rf_rpb);
    for (auto &l : col_launches_reg(c, 0, c->cr_tc, c->cr_ntc)) l.go();
    c->grid = g0;
    c->bufs = b0;
    c->fl_used = f0;
    c->zs_used = z0;
    c->n = n0;
    c->nf_used = nf0;
  };
  HIPCHECK(hipMemcpyAsync(d_obeg, obeg.data(), ns * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(d_oend, oend.data(), ns * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(d_oc, oc.data(), oc.size() * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
  // The rest of a reverse step: the stage adjoints in reverse stage order
  // (vector-Jacobian products and their accumulation), then the step index.
  auto enqueue_step_tail = [&](double *lam_in, double *ybar_out) -> int {
    const unsigned nb = blocks_of(nN);
    int s2;
    for (int s = 7; s >= 1; --s) {
      const double *u = s == 1 ? ycur : Ys[s - 1];
      rd::Bufs pre = c->bufs;
      pre.S = S7 + (size_t)(s - 1) * c->n * c->Nh;
      pre.F = F7 + (size_t)(s - 1) * nf * c->Nh;
      if ((s2 = vjp(c, u, kbar[s - 1], out_id, Gk, Dk, P, tmap, batched_spectra ? &pre : nullptr)))
        return s2;
      rd::AdjK kb{};
      for (int j = 0; j < 7; ++j) kb.k[j] = kbar[j];
      if (s >= 2)
        for (int j = 0; j < s - 1; ++j) kb.c[j] = STAGE_A[s - 2][j];
      hipLaunchKernelGGL(rd::k_adj_stage, dim3(nb), dim3(256), 0, c->stream,
                         (const double *)out_id, (const double *)Dk, (const int *)fmap, nf,
                         (const double *)(s == 7 ? lam_in : nullptr), ybar_out, kb,
                         (const double *)d_hs, (const long long *)kdev, N, nN);
    }
    hipLaunchKernelGGL(rd::k_adj_next, dim3(1), dim3(64), 0, c->stream, kdev);
    HIPCHECK(hipGetLastError());
    return RDCU_OK;
  };
  // One reverse step (the step index from kdev), with lam_in the adjoint of
  // the step's end state and ybar_out that of its start.
  auto enqueue_step = [&](double *lam_in, double *ybar_out) -> int {
    const unsigned nb = blocks_of(nN);
    if (kst) {  // stored stages: the step's point-wise start in one kernel
      rd::StageIn6 si6{};
      for (int s = 2; s <= 7; ++s) {
        for (int j = 0; j < s - 1; ++j) si6.a[s - 2][j] = STAGE_A[s - 2][j];
        si6.out[s - 2] = Ys[s - 1];
      }
      rd::AdjK kz{};
      for (int j = 0; j < 7; ++j) kz.k[j] = kbar[j];
      hipLaunchKernelGGL(rd::k_adj_begin, dim3(nb), dim3(256), 0, c->stream, ych,
                         (const long long *)kdev, ycur, (const double *)kst, (const double *)d_hs,
                         si6, ybar_out, kz, (const double *)gbar, (const int *)d_obeg,
                         (const int *)d_oend, (const double *)d_oc, nN);
      if (batched_spectra) enqueue_spectra7();
      return enqueue_step_tail(lam_in, ybar_out);
    }
    hipLaunchKernelGGL(rd::k_adj_fetch, dim3(nb), dim3(256), 0, c->stream, ych, kdev, ycur, nN);
    // stage inputs from the checkpoint (the forward's operations); the stage
    // values stored by the forward pass, or recomputed
    int s2;
    if (!kst && (s2 = enqueue_rhs(c, 0, rd::SRC_PTR, ycur, rd::DST_PTR, 0, K[0], 0, 0))) return s2;
    for (int s = 2; s <= 7; ++s) {
      rd::StageIn si{};
      si.ns = s - 1;
      for (int j = 0; j < s - 1; ++j) {
        si.k[j] = kst ? kst + (size_t)j * nN : K[j];
        si.a[j] = STAGE_A[s - 2][j];
      }
      hipLaunchKernelGGL(rd::k_stage_in, dim3(nb), dim3(256), 0, c->stream, ycur, si,
                         (const double *)d_hs, (const long long *)kdev, kst ? 7 * nN : 0L,
                         Ys[s - 1], nN);
      if (!kst && (s2 = enqueue_rhs(c, 0, rd::SRC_PTR, Ys[s - 1], rd::DST_PTR, 0, K[s - 1], 0, 0)))
        return s2;
    }
    if (batched_spectra) enqueue_spectra7();
    rd::AdjK kz{};
    for (int j = 0; j < 7; ++j) kz.k[j] = kbar[j];
    hipLaunchKernelGGL(rd::k_adj_zero, dim3(nb), dim3(256), 0, c->stream, kz, ybar_out, nN);
    hipLaunchKernelGGL(rd::k_adj_dense, dim3(nb), dim3(256), 0, c->stream, ybar_out, kz,
                       (const double *)gbar, (const int *)d_obeg, (const int *)d_oend,
                       (const double *)d_oc, (const long long *)kdev, nN);
    return enqueue_step_tail(lam_in, ybar_out);
  };
  // The step is captured twice, once per role of the two adjoint buffers
  // (they swap every step), and replayed ns times; x[2] holds kGraphAttempts
  // steps (both roles in turn) for the whole groups, which start at even steps.
  struct Execs {
    hipStream_t s;
    hipGraphExec_t x[3] = {nullptr, nullptr, nullptr};
    ~Execs() {
      (void)hipStreamSynchronize(s);
      for (auto &e : x)
        if (e) (void)hipGraphExecDestroy(e);
    }
  } ex{c->stream};
  double *buf[2] = {lam, ybar};
  for (int par = 0; par < 3 && (par < 2 ? par < ns : ns >= kGraphAttempts); ++par) {
    HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int s2 = RDCU_OK;
    if (par < 2) s2 = enqueue_step(buf[par], buf[1 - par]);
    for (int i = 0; par == 2 && i < kGraphAttempts && !s2; ++i)
      s2 = enqueue_step(buf[i & 1], buf[1 - (i & 1)]);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(c->stream, &g);
    if (s2 || ce != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return s2 ? s2 : fail(RDCU_EDEVICE, std::string("reverse step capture: ") + hipGetErrorString(ce));
    }
    if (par == 0 && timing) (void)hipGraphGetNodes(g, nullptr, &step_nodes);
    const hipError_t ie = hipGraphInstantiate(&ex.x[par], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess)
      return fail(RDCU_EDEVICE, std::string("reverse step graph: ") + hipGetErrorString(ie));
  }
  hipLaunchKernelGGL(rd::k_adj_setk, dim3(1), dim3(64), 0, c->stream, kdev, ns - 1);
  if (timing) HIPCHECK(hipEventRecord(pev.e[2], c->stream));
  long long i = 0;
  if (ex.x[2])
    for (; i + kGraphAttempts <= ns; i += kGraphAttempts) HIPCHECK(hipGraphLaunch(ex.x[2], c->stream));
  for (; i < ns; ++i) HIPCHECK(hipGraphLaunch(ex.x[i & 1], c->stream));
  if (timing) HIPCHECK(hipEventRecord(pev.e[3], c->stream));
  if (!kst) c->nrhs += 7 * ns;
  lam = buf[ns & 1];
  std::vector<double> hp((size_t)std::max(1, nt) * c->rows);
  HIPCHECK(hipMemcpyAsync(hp.data(), P, hp.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  if (grad_y0)
    HIPCHECK(hipMemcpyAsync(grad_y0, lam, nN * sizeof(double),
                            dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (timing) {
    float fwd = 0, rev = 0;
    HIPCHECK(hipEventElapsedTime(&fwd, pev.e[0], pev.e[1]));
    HIPCHECK(hipEventElapsedTime(&rev, pev.e[2], pev.e[3]));
    fprintf(stderr,
            "rdcu loss_grad: %lld steps; forward with checkpoints %.3f ms (%.1f us per step); "
            "reverse sweep %.3f ms (%.1f us per step, %zu graph nodes per step)\n",
            ns, fwd, 1e3 * fwd / (double)ns, rev, 1e3 * rev / (double)ns, step_nodes);
  }
  if (grad_coef)
    for (int t = 0; t < nt; ++t) {
      double sacc = 0.0;
      for (long r = 0; r < c->rows; ++r) sacc += hp[(size_t)t * c->rows + r];
      grad_coef[t] = sacc;
    }
  return RDCU_OK;
}

}  // namespace

extern "C" {

int rdcu_loss_grad(rdcu_ctx *c, const double *y0, double t0, long long ns, const double *hs,
                   int nout, const double *t_out, const double *target, double *loss,
                   double *grad_coef, double *grad_y0) {
  return loss_grad(c, y0, t0, ns, hs, nout, t_out, target, loss, grad_coef, grad_y0, false);
}

int rdcu_loss_grad_device(rdcu_ctx *c, const double *y0_dev, double t0, long long ns,
                          const double *hs, int nout, const double *t_out,
                          const double *target_dev, double *loss, double *grad_coef,
                          double *grad_y0_dev) {
  return loss_grad(c, y0_dev, t0, ns, hs, nout, t_out, target_dev, loss, grad_coef, grad_y0_dev,
                   true);
}

int rdcu_loss_grad_batch(rdcu_grad_job *jobs, int count, int concurrency) {
  if (count < 0 || (count > 0 && !jobs)) return fail(RDCU_EINVAL, "bad argument");
  std::set<rdcu_ctx *> seen;
  for (int i = 0; i < count; ++i) {
    jobs[i].status = RDCU_OK;
    if (!jobs[i].ctx || !seen.insert(jobs[i].ctx).second)
      return fail(RDCU_EINVAL, "every job needs its own handle");
  }
  // companion problems (they compile and tune) are built here, one at a time
  for (int i = 0; i < count; ++i) {
    rdcu_grad_job &j = jobs[i];
    if (!(j.grad_coef || j.grad_y0) || !j.ctx->reg || j.ctx->member) continue;
    HIPCHECK(hipSetDevice(j.ctx->dev));
    if (int st = adj_companion(j.ctx)) return j.status = st;
  }
  // Each worker thread runs whole jobs, each on its job's own stream, so up
  // to `concurrency` instances' kernels share the GPU.
  const int W = std::max(1, std::min(count, concurrency > 0 ? concurrency : 4));
  std::atomic<int> next{0};
  std::mutex mu;
  int first = -1;
  std::string first_msg;
  auto work = [&]() {
    for (;;) {
      const int i = next.fetch_add(1);
      if (i >= count) break;
      rdcu_grad_job &j = jobs[i];
      j.status = rdcu_loss_grad(j.ctx, j.y0, j.t0, j.nsteps, j.hs, j.nout, j.t_out, j.target,
                                j.loss, j.grad_coef, j.grad_y0);
      if (j.status) {
        std::lock_guard<std::mutex> lk(mu);
        if (first < 0 || i < first) {
          first = i;
          first_msg = rdcu_last_error();
        }
      }
    }
  };
  std::vector<std::thread> th;
  for (int w = 0; w < W; ++w) th.emplace_back(work);
  for (auto &t : th) t.join();
  if (first >= 0) return fail(jobs[first].status, first_msg);
  return RDCU_OK;
}

}  // extern "C"
