// rd_callback.hip -- integrators whose right-hand side is a host callback
// (included at the end of rd_runtime.hip: same translation unit).
//
// dp45_init takes any dydt (dp45.h:8, dp45.cu:171-177); rdcu_create takes a
// dictionary, whose RHS the kernels evaluate themselves.  rdcu_create_fn keeps
// the reference's generality: the stage sums, the error norm and the step
// control run on the device as for a dictionary problem, and the callback is
// handed device pointers for dydt(t + c_s h, ytemp, k_s) at every stage
// (dp45.cu:90-106).  The callback sees t and h, so an attempt is enqueued by
// the host and read back once (the reference blocks once per attempt too, in
// cublasDnrm2, dp45.cu:118); the attempts cannot be captured in a graph.

namespace {

// Stage times of dp45.cu:90-106: t + h*c_s, and t + h for the last stage.
const double STAGE_C[6] = {1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0, 1.0};

int fn_call(rdcu_ctx *c, double t, const double *y, double *f) {
  if (c->fn(t, y, f, (void *)c->stream, c->fn_user))
    return fail(RDCU_ECALLBACK, "the right-hand side callback returned an error");
  HIPCHECK(hipGetLastError());
  return RDCU_OK;
}

// k1 = f(t, y) (dp45.cu:157), with the buffer roles of the host shadow.
int fn_ensure_k1(rdcu_ctx *c, double t) {
  if (c->k1_valid) return RDCU_OK;
  int st = fn_call(c, t, c->bufs.y[c->sh->iy], c->bufs.k[c->sh->k1s]);
  if (st) return st;
  c->k1_valid = true;
  return RDCU_OK;
}

// One attempt from (t, h), which the device state already holds; the host
// shadow holds the buffer roles.  Ends with the state read back.
int fn_attempt(rdcu_ctx *c, double t, double h) {
  const long nN = (long)c->n * c->N;
  const dim3 g((unsigned)((nN + 255) / 256)), b(256);
  int st;
  for (int s = 2; s <= 7; ++s) {
    hipLaunchKernelGGL(rd::k_fn_stage, g, b, 0, c->stream, c->st, c->bufs, stage_coef(s), nN);
    HIPCHECK(hipGetLastError());
    const double ts = s < 7 ? t + h * STAGE_C[s - 2] : t + h;
    double *ks = s < 7 ? c->bufs.k[s - 1] : c->bufs.k[c->sh->k7s];
    if ((st = fn_call(c, ts, c->bufs.y[c->sh->iyt], ks))) return st;
  }
  const double *part = c->bufs.part;
  if (!c->fixed) {
    rd::ErrCoef ec{};
    for (int j = 0; j < 7; ++j) ec.e[j] = EE[j];
    hipLaunchKernelGGL(rd::k_fn_err, dim3((unsigned)((long)c->n * c->fn_nchunk)), dim3(256), 0,
                       c->stream, c->st, c->bufs, ec, c->N, c->fn_nchunk, c->bufs.part);
    if (c->fn_nchunk > 1) {
      hipLaunchKernelGGL(rd::k_fn_lines, dim3((unsigned)c->n), dim3(256), 0, c->stream,
                         c->bufs.part, c->fn_nchunk, c->bufs.part2);
      part = c->bufs.part2;
    }
  }
  hipLaunchKernelGGL(rd::k_control, dim3(1), dim3(256), 0, c->stream, c->st, part, (long)c->n, nN);
  HIPCHECK(hipGetLastError());
  c->nrhs += 6;
  return read_state(c);
}

int fn_step(rdcu_ctx *c, double *t, double *h, double hmax, int *accepted) {
  int st;
  if ((st = begin(c, INFINITY, 1, *t, *h, -INFINITY, hmax, INFINITY))) return st;
  if ((st = fn_ensure_k1(c, *t))) return st;
  if ((st = fn_attempt(c, *t, *h))) return st;
  if (c->sh->nlog != 1) return fail(RDCU_ESTATE, "attempt did not run");
  *accepted = c->sh->log_acc[0];
  *t = c->sh->t;
  *h = c->sh->h;
  return RDCU_OK;
}

// The run loop of dp45_run (dp45.cu:154-169) one attempt at a time: with cb,
// rdcu_run's (every attempt is current); without, rdcu_advance's (at most
// max_attempts attempts, their count in *done).
int fn_run(rdcu_ctx *c, double *t, double *h, double hmin, double hmax, double t1, double stop,
           rdcu_step_fn cb, void *user, long long max_attempts, long long *done) {
  int st;
  if ((st = begin(c, INFINITY, 1, *t, *h, hmin, hmax, t1))) return st;
  c->k1_valid = false;  // dp45_run recomputes k1 = f(y) (dp45.cu:157)
  if ((st = fn_ensure_k1(c, *t))) return st;
  long long k = 0;
  double tc = *t, hc = *h;
  while (tc < t1 && hc > hmin && (max_attempts < 0 || k < max_attempts)) {
    if ((st = begin(c, INFINITY, 0, 0, 0, hmin, hmax, t1))) return st;
    if ((st = fn_attempt(c, tc, hc))) return st;
    if (c->sh->nlog != 1) break;
    ++k;
    tc = c->sh->t;
    hc = c->sh->h;
    *t = tc;
    *h = hc;
    if (cb && cb(c, tc, hc, c->sh->log_acc[0], 1, &stop, user))
      return fail(RDCU_ECALLBACK, "callback stop");
  }
  *t = c->sh->t;
  *h = c->sh->h;
  if (done) *done = k;
  return RDCU_OK;
}

}  // namespace

extern "C" {

int rdcu_create_fn(int n, long long N, double atol, double rtol, int fixedstep, const double *y0,
                   rdcu_dydt_fn fn, void *user, int device, rdcu_ctx **out) {
  if (!out) return fail(RDCU_EINVAL, "null output handle");
  *out = nullptr;
  if (!fn) return fail(RDCU_EINVAL, "null right-hand side callback");
  if (n < 1 || N < 1) return fail(RDCU_EINVAL, "n and N must be >= 1");
  const long long nchunk = (N + rd::kFnChunk - 1) / rd::kFnChunk;
  if (n > rd::kRedCap || nchunk > rd::kRedCap)
    return fail(RDCU_EUNSUPPORTED, "at most 4096 fields of at most 2^24 values");
  std::unique_ptr<rdcu_ctx> c(new rdcu_ctx());
  c->fn = fn;
  c->fn_user = user;
  c->fn_nchunk = (int)nchunk;
  c->n = n;
  c->ndim = 1;
  c->N = (long)N;
  c->rows = 1;
  c->Ns[0] = (int)std::min<long long>(N, 1 << 30);
  c->nblk = 1;
  c->fixed = fixedstep ? 1 : 0;
  c->dev = device;
  HIPCHECK(hipSetDevice(device));
  int st;
  {
    char *base = nullptr;
    if ((st = scratch_base(device, &base))) return st;
  }
  HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  const size_t nN = (size_t)n * N;
  for (int i = 0; i < 3; ++i)
    if ((st = c->dalloc(&c->bufs.y[i], nN))) return st;
  for (int i = 0; i < 7; ++i)
    if ((st = c->dalloc(&c->bufs.k[i], nN))) return st;
  if ((st = c->dalloc(&c->bufs.part, (size_t)n * nchunk)) || (st = c->dalloc(&c->bufs.part2, n)) ||
      (st = c->dalloc(&c->yeval, nN)))
    return st;
  c->nlines = n;
  if (y0) HIPCHECK(hipMemcpy(c->bufs.y[0], y0, nN * sizeof(double), hipMemcpyHostToDevice));
  else HIPCHECK(hipMemset(c->bufs.y[0], 0, nN * sizeof(double)));
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
  c->sh->t_stop = INFINITY;
  c->sh->halted = 1;
  if ((st = c->dalloc(&c->st, 1))) return st;
  HIPCHECK(hipMemcpy(c->st, c->sh, sizeof(rd::DevState), hipMemcpyHostToDevice));
  *out = c.release();
  return RDCU_OK;
}

int rdcu_rhs_device(rdcu_ctx *c, const double *y_dev, double *f_dev) {
  if (!c || !y_dev || !f_dev) return fail(RDCU_EINVAL, "null argument");
  if (c->fn) return no_dict();
  HIPCHECK(hipSetDevice(c->dev));
  int st = enqueue_rhs(c, 0, rd::SRC_PTR, y_dev, rd::DST_PTR, 0, f_dev, 0, 0);
  if (st) return st;
  HIPCHECK(hipStreamSynchronize(c->stream));
  return RDCU_OK;
}

}  // extern "C"
