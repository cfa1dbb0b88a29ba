// rd_batch.hip -- batched integration of many independent problems on one GPU
// (included at the end of rd_runtime.hip: same translation unit).
//
// The reference runs a parameter sweep as one process per instance
// (sweep.sh:21-41 -> makelines.py:260-268 -> ./rdcu).  Here the instances of
// a sweep live side by side in HBM and every DP45 attempt of all of them is
// one hipGraph: per shape class (same grid, same transform lengths) one
// launch per pass walks the (instance, block) items of the whole class, with
// a device table of per-instance buffers, state, factor lists and
// dictionaries; step control then runs for every instance at once.  Each
// instance keeps its own t, h and accept/reject history: an instance whose
// attempts have ended (t >= t1, h <= hmin) is skipped by every launch, and the
// arithmetic per instance is the single-instance sequence, so its steps are
// bit-identical to a run of that problem alone.

struct BatchGroup {  // instances of one class sharing a dictionary structure
  std::vector<int> all;  // its members
  int i0 = 0, i1 = 0;    // range in the class's current table
  hipFunction_t fn = nullptr;  // specialised x-inverse kernel (null: interpreter)
  // its variants: 1 twiddles in registers, 2 lean, 3 lean with the twiddles
  // in registers (null: not built)
  hipFunction_t fnv[4] = {};
};

struct BatchClass {
  std::vector<int> members;  // batch member indices (creation order)
  std::vector<BatchGroup> groups;
  int off = 0;               // offset of this class in the batch's current instance table
  int cnt = 0;               // instances of this class in the current table
  rdcu_ctx *c0 = nullptr;    // a representative member (geometry, plans, tables)
  int Ex = 0, E0 = 0, rows = 0, nf_max = 0, nwords_max = 0, max_fpf = 1;
  int rpb = 1, tc = 1, fpw = 1, G = 1, R = 1;
  int rv = 0;        // x-inverse variant (BatchGroup::fnv), the tuner's choice
  long grid_rf = 0;  // persistent grid of the x-forward pass (0: one block per item)
  std::string tune_log;
};

struct rdcu_batch {
  int dev = 0;
  hipStream_t stream = nullptr;
  hipStream_t ls = nullptr;  // stream the launch helpers enqueue on (stream, or a class's)
  std::vector<std::unique_ptr<rdcu_ctx>> members;
  std::vector<int> pos;  // member -> device state slot (= position in the full table)
  std::vector<int> tab;  // current table position -> member
  bool table_full = true;
  std::vector<BatchClass> classes;
  rd::Inst *d_inst = nullptr;
  rd::DevState *d_st = nullptr, *h_st = nullptr;  // instance table order
  int *d_done = nullptr, *h_done = nullptr;
  double *d_t = nullptr, *d_h = nullptr, *d_t1 = nullptr, *h_buf = nullptr;
  // One attempt of each class's instances is one graph, replayed on the
  // class's own stream, so that the classes' launches overlap instead of
  // adding up; fork and join events order them after and before the work on
  // `stream` (state set-up, done flags).
  std::vector<hipStream_t> cstream;
  std::vector<hipGraphExec_t> cexec;  // null for a class with no instance in the table
  std::vector<hipEvent_t> cjoin;
  hipEvent_t fork = nullptr;
  bool captured = false;
  hipModule_t jit_mod = nullptr;
  std::string jit_log = "off";
  bool k1_valid = false;
  int fixed = 0;
  long long nsync = 0;

  ~rdcu_batch() {
    for (hipStream_t s : cstream) (void)hipStreamSynchronize(s);
    if (stream) (void)hipStreamSynchronize(stream);
    for (hipGraphExec_t e : cexec)
      if (e) (void)hipGraphExecDestroy(e);
    for (hipEvent_t e : cjoin) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    for (hipStream_t s : cstream) (void)hipStreamDestroy(s);
    if (jit_mod) (void)hipModuleUnload(jit_mod);
    members.clear();  // before the stream and state they borrow
    for (void *p : {(void *)d_inst, (void *)d_st, (void *)d_done, (void *)d_t, (void *)d_h,
                    (void *)d_t1})
      if (p) (void)hipFree(p);
    if (h_st) (void)hipHostFree(h_st);
    if (h_done) (void)hipHostFree(h_done);
    if (h_buf) (void)hipHostFree(h_buf);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int ninst(const rdcu_batch *b) { return (int)b->members.size(); }


// LDS row region of the class (largest factor count and program).
long rowi_region_b(const BatchClass &cl, int G) {
  const rdcu_ctx *c = cl.c0;
  return rd::rowi_region(c->L, c->Nx, c->n, cl.nf_max, cl.nwords_max, G);
}
size_t rowi_lds(const BatchClass &cl, int G, int R) {
  return (size_t)R * rowi_region_b(cl, G) * sizeof(cx);
}

// Batch classes of these lengths run kernels compiled for the length (x lines:
// half lengths of 64-, 128- and 256-point rows; strided lines of 64, 128 and
// 256 points); other lengths run the run-time-plan kernels.
constexpr int kRowLsB[] = {32, 64, 128};
constexpr int kColLsB[] = {64, 128, 256};
// and the 3*2^k lengths of the 96^2 and 192^2 grids (x lines 48 and 96,
// strided lines 96 and 192)
constexpr int kRowLsB3[] = {48, 96};
constexpr int kColLsB3[] = {96, 192};

// Raises the dynamic-LDS limit of every batched kernel to the 160 KB of a
// gfx950 CU (for the device current when it runs).
int allow_big_lds_b() {
  const int lim = 160 * 1024;
  for (const void *f :
       {(const void *)rd::k_rowf_b<8>, (const void *)rd::k_rowf_b<8, -1>,
        (const void *)rd::k_col_b<8>, (const void *)rd::k_col_b<8, -1>,
        (const void *)rd::k_rowi_b<8>, (const void *)rd::k_rowi_b<8, -1>,
        (const void *)rd::k_rowf_b<8, kRowLsB[0]>, (const void *)rd::k_rowf_b<8, kRowLsB[1]>,
        (const void *)rd::k_rowf_b<8, kRowLsB[2]>, (const void *)rd::k_col_b<8, kColLsB[0]>,
        (const void *)rd::k_col_b<8, kColLsB[1]>, (const void *)rd::k_col_b<8, kColLsB[2]>,
        (const void *)rd::k_rowf_b<8, kRowLsB3[0]>, (const void *)rd::k_rowf_b<8, kRowLsB3[1]>,
        (const void *)rd::k_col_b<8, kColLsB3[0]>, (const void *)rd::k_col_b<8, kColLsB3[1]>,
        (const void *)rd::k_col_b<8, 0, true>, (const void *)rd::k_col_b<8, -1, true>,
        (const void *)rd::k_col_b<8, kColLsB[0], true>, (const void *)rd::k_col_b<8, kColLsB[1], true>,
        (const void *)rd::k_col_b<8, kColLsB[2], true>, (const void *)rd::k_col_b<8, kColLsB3[0], true>,
        (const void *)rd::k_col_b<8, kColLsB3[1], true>})
    HIPCHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
  return RDCU_OK;
}

template <int E, int LS>
void go_rowf_b(rdcu_batch *b, const BatchClass &cl, dim3 g, dim3 bl, size_t lds, int per,
               int src_kind, const rd::StageCoef &sc, int rpb, int check) {
  const rdcu_ctx *c = cl.c0;
  hipLaunchKernelGGL((rd::k_rowf_b<E, LS>), g, bl, lds, b->ls, b->d_inst + cl.off, cl.cnt, per,
                     c->px, c->twN, src_kind, sc, rpb, check);
}

template <int E, int LS>
void go_col_b(rdcu_batch *b, const BatchClass &cl, dim3 g, dim3 bl, size_t lds, int per,
              const rd::ColGeom &cg, int fpw, int check) {
  if constexpr (E == 8) {
    if (fpw > 1) {  // several factors per block: their forward spectrum in registers
      hipLaunchKernelGGL((rd::k_col_b<E, LS, true>), g, bl, lds, b->ls, b->d_inst + cl.off, cl.cnt,
                         per, cl.c0->p0, cg, 1, fpw, check);
      return;
    }
  }
  hipLaunchKernelGGL((rd::k_col_b<E, LS>), g, bl, lds, b->ls, b->d_inst + cl.off, cl.cnt, per,
                     cl.c0->p0, cg, 1, fpw, check);
}

template <int E>
void launch_rowf_b(rdcu_batch *b, const BatchClass &cl, int src_kind, rd::StageCoef sc, int rpb,
                   long grid, int check) {
  const rdcu_ctx *c = cl.c0;
  const int T = c->L / E;
  const int per = (int)(((long)c->n * c->rows + rpb - 1) / rpb);
  const long total = (long)cl.cnt * per;
  if (total == 0) return;
  const long g = grid > 0 ? std::min(grid, total) : total;
  const size_t lds = (size_t)(c->L + rpb * rd::padded_len(c->L)) * sizeof(cx);
  const dim3 gd((unsigned)g), bl(rpb * T);
  if constexpr (E == 8) {
    if (c->L == kRowLsB[0]) return go_rowf_b<E, kRowLsB[0]>(b, cl, gd, bl, lds, per, src_kind, sc, rpb, check);
    if (c->L == kRowLsB[1]) return go_rowf_b<E, kRowLsB[1]>(b, cl, gd, bl, lds, per, src_kind, sc, rpb, check);
    if (c->L == kRowLsB[2]) return go_rowf_b<E, kRowLsB[2]>(b, cl, gd, bl, lds, per, src_kind, sc, rpb, check);
    if (c->L == kRowLsB3[0]) return go_rowf_b<E, kRowLsB3[0]>(b, cl, gd, bl, lds, per, src_kind, sc, rpb, check);
    if (c->L == kRowLsB3[1]) return go_rowf_b<E, kRowLsB3[1]>(b, cl, gd, bl, lds, per, src_kind, sc, rpb, check);
    if (odd_len(c->L)) return go_rowf_b<E, -1>(b, cl, gd, bl, lds, per, src_kind, sc, rpb, check);
  }
  go_rowf_b<E, 0>(b, cl, gd, bl, lds, per, src_kind, sc, rpb, check);
}

template <int E>
void launch_col_b(rdcu_batch *b, const BatchClass &cl, int tc, int fpw, int check) {
  const rdcu_ctx *c = cl.c0;
  rd::ColGeom cg{};
  cg.L = c->Ns[0];
  cg.ntc = (c->Nxh + tc - 1) / tc;
  cg.tc = tc;
  cg.gs1 = c->Nh;
  cg.es = c->Nxp;
  cg.gdiv = 1;
  cg.nparts = (cl.max_fpf + fpw - 1) / fpw;
  const int per = c->n * cg.ntc * cg.nparts;
  cg.nblk = per;
  const int T = cg.L / E;
  const long total = (long)cl.cnt * per;
  if (total == 0) return;
  const long g = (total + 7) / 8 * 8;  // one item per block, a multiple of the 8 XCDs
  // the kept forward spectrum (tc*L) only for several factors per block with
  // 16 values per thread; 8 per thread keep it in registers (go_col_b)
  const bool keep = fpw > 1 && E != 8;
  const size_t lds = (size_t)(cg.L + tc * rd::padded_len(cg.L) + (keep ? tc * cg.L : 0)) * sizeof(cx) +
                     (size_t)cg.L * sizeof(double);
  const dim3 gd((unsigned)g), bl(tc * T);
  if constexpr (E == 8) {
    if (cg.L == kColLsB[0]) return go_col_b<E, kColLsB[0]>(b, cl, gd, bl, lds, per, cg, fpw, check);
    if (cg.L == kColLsB[1]) return go_col_b<E, kColLsB[1]>(b, cl, gd, bl, lds, per, cg, fpw, check);
    if (cg.L == kColLsB[2]) return go_col_b<E, kColLsB[2]>(b, cl, gd, bl, lds, per, cg, fpw, check);
    if (cg.L == kColLsB3[0]) return go_col_b<E, kColLsB3[0]>(b, cl, gd, bl, lds, per, cg, fpw, check);
    if (cg.L == kColLsB3[1]) return go_col_b<E, kColLsB3[1]>(b, cl, gd, bl, lds, per, cg, fpw, check);
    if (odd_len(cg.L)) return go_col_b<E, -1>(b, cl, gd, bl, lds, per, cg, fpw, check);
  }
  go_col_b<E, 0>(b, cl, gd, bl, lds, per, cg, fpw, check);
}

template <int E>
void launch_rowi_b(rdcu_batch *b, const BatchClass &cl, int src_kind, int dst_kind, int dst_slot,
                   int err, rd::StageCoef nx, int G, int R, int check) {
  const rdcu_ctx *c = cl.c0;
  const int T = c->L / E;
  rd::ErrCoef ec{};
  for (int j = 0; j < 7; ++j) ec.e[j] = EE[j];
  const size_t lds = rowi_lds(cl, G, R);
  long region = rowi_region_b(cl, G);
  for (const BatchGroup &gr : cl.groups) {
    const rd::Inst *in = b->d_inst + cl.off + gr.i0;
    int cnt = gr.i1 - gr.i0, rows = (int)c->rows;
    if (cnt == 0) continue;
    const unsigned grid = (unsigned)((long)cnt * rows / R);  // R rows per block, R | rows
    const unsigned block = (unsigned)(R * G * T);
    hipFunction_t fn = cl.rv > 0 && gr.fnv[cl.rv] ? gr.fnv[cl.rv] : gr.fn;
    if (fn) {
      rd::FftPlan px = c->px;
      const rd::cx *tw = c->twN;
      int a_src = src_kind, a_dk = dst_kind, a_ds = dst_slot, a_err = err, a_G = G, a_chk = check;
      void *args[] = {&in, &cnt, &rows, &px, &tw, &a_src, &a_dk, &a_ds, &a_err, &ec, &nx, &a_G,
                      &a_chk, &region};
      // a failed launch is reported by the hipGetLastError check after each
      // group of launches (and inside a graph capture it fails the capture)
      (void)hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, (unsigned)lds, b->ls, args,
                                  nullptr);
    } else if (odd_len(c->L)) {
      hipLaunchKernelGGL((rd::k_rowi_b<E, -1>), dim3(grid), dim3(block), lds, b->ls, in, cnt,
                         rows, c->px, c->twN, src_kind, dst_kind, dst_slot, err, ec, nx, G, check,
                         region);
    } else {
      hipLaunchKernelGGL(rd::k_rowi_b<E>, dim3(grid), dim3(block), lds, b->ls, in, cnt, rows,
                         c->px, c->twN, src_kind, dst_kind, dst_slot, err, ec, nx, G, check,
                         region);
    }
  }
}

void rowf_b(rdcu_batch *b, const BatchClass &cl, int src_kind, rd::StageCoef sc, int rpb,
            long grid, int check) {
  launch_rowf_b<8>(b, cl, src_kind, sc, rpb, grid, check);
}
void col_b(rdcu_batch *b, const BatchClass &cl, int tc, int fpw, int check) {
  if (cl.c0->ndim < 2) return;
  launch_col_b<8>(b, cl, tc, fpw, check);
}
void rowi_b(rdcu_batch *b, const BatchClass &cl, int src_kind, int dst_kind, int dst_slot, int err,
            rd::StageCoef nx, int G, int R, int check) {
  launch_rowi_b<8>(b, cl, src_kind, dst_kind, dst_slot, err, nx, G, R, check);
}

// One plain RHS of every instance of a class (input y, output k slot `slot`
// or the FSAL slot with DST_K1).
void rhs_b(rdcu_batch *b, const BatchClass &cl, int dst_kind, int slot, int check) {
  rd::StageCoef none{};
  rowf_b(b, cl, rd::SRC_Y, none, cl.rpb, cl.grid_rf, check);
  col_b(b, cl, cl.tc, cl.fpw, check);
  rowi_b(b, cl, rd::SRC_Y, dst_kind, slot, 0, none, cl.G, cl.R, check);
}

// One DP45 attempt of every active instance of a class (the single-instance
// sequence of attempt_launches_reg), then their step control, on b->ls.
int enqueue_attempt_b(rdcu_batch *b, const BatchClass &cl) {
  rowf_b(b, cl, rd::SRC_STAGE, stage_coef(2), cl.rpb, cl.grid_rf, 1);
  for (int s = 2; s <= 7; ++s) {
    col_b(b, cl, cl.tc, cl.fpw, 1);
    rd::StageCoef nx = s < 7 ? stage_coef(s + 1) : rd::StageCoef{};
    const int dk = s < 7 ? rd::DST_KS : rd::DST_K7;
    rowi_b(b, cl, rd::SRC_YT, dk, s - 1, s == 7 && !b->fixed, nx, cl.G, cl.R, 1);
  }
  hipLaunchKernelGGL(rd::k_control_b, dim3((unsigned)cl.cnt), dim3(256), 0, b->ls,
                     b->d_inst + cl.off, b->d_done + cl.off);
  HIPCHECK(hipGetLastError());
  return RDCU_OK;
}

// k attempts of every instance in the table: each class's graph k times on
// its own stream, after the work already on b->stream and before what
// follows there.
int replay_b(rdcu_batch *b, long long k) {
  HIPCHECK(hipEventRecord(b->fork, b->stream));
  const int nc = (int)b->classes.size();
  for (int c = 0; c < nc; ++c)
    if (b->cexec[c]) HIPCHECK(hipStreamWaitEvent(b->cstream[c], b->fork, 0));
  for (long long r = 0; r < k; ++r)
    for (int c = 0; c < nc; ++c)
      if (b->cexec[c]) HIPCHECK(hipGraphLaunch(b->cexec[c], b->cstream[c]));
  for (int c = 0; c < nc; ++c)
    if (b->cexec[c]) {
      HIPCHECK(hipEventRecord(b->cjoin[c], b->cstream[c]));
      HIPCHECK(hipStreamWaitEvent(b->stream, b->cjoin[c], 0));
    }
  return RDCU_OK;
}

// Launch shapes per class, timed on the plain RHS of all its instances
// (every candidate computes the same values).
int autotune_b(rdcu_batch *b, BatchClass &cl) {
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  auto time_of = [&](const std::function<void()> &f) -> double {
    f();
    (void)hipEventRecord(e0, b->stream);
    const int reps = 5;
    for (int i = 0; i < reps; ++i) f();
    (void)hipEventRecord(e1, b->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (hipGetLastError() != hipSuccess) return 1e30;
    return ms / reps;
  };
  const rdcu_ctx *c = cl.c0;
  const int T = c->L / cl.Ex;
  rd::StageCoef none{};
  char line[160];
  const long caps[] = {0, 4096};
  double best = 1e30;
  for (int rpb = std::max(1, 64 / T); rpb <= 128; rpb *= 2) {  // from one wavefront up
    if (rpb * T > 256 || (size_t)(c->L + rpb * rd::padded_len(c->L)) * sizeof(cx) > 160 * 1024)
      break;
    for (long cap : caps) {
      double tm = time_of([&]() { rowf_b(b, cl, rd::SRC_Y, none, rpb, cap, 0); });
      snprintf(line, sizeof line, "row_fwd rpb=%d grid=%ld %.1fus; ", rpb, cap, tm * 1e3);
      cl.tune_log += line;
      if (tm < best) {
        best = tm;
        cl.rpb = rpb;
        cl.grid_rf = cap;
      }
    }
  }
  if (c->ndim >= 2) {
    best = 1e30;
    const int Tc = c->Ns[0] / cl.E0;
    for (int tc = std::max(1, 64 / Tc); tc <= 64; tc *= 2) {
      if (tc * Tc > 256 ||
          (size_t)(3 * c->Ns[0] + tc * rd::padded_len(c->Ns[0]) + tc * c->Ns[0]) * sizeof(cx) >
              160 * 1024)
        break;
      for (int fpw = cl.max_fpf; fpw >= 1; fpw = fpw > 1 ? (fpw + 1) / 2 : 0) {
        double tm = time_of([&]() { col_b(b, cl, tc, fpw, 0); });
        snprintf(line, sizeof line, "col tc=%d fpw=%d %.1fus; ", tc, fpw, tm * 1e3);
        cl.tune_log += line;
        if (tm < best) {
          best = tm;
          cl.tc = tc;
          cl.fpw = fpw;
        }
        if (fpw == 1) break;
      }
    }
  }
  best = 1e30;
  // factor lines per group, also wider than the factor count (powers of two;
  // idle lines, but more threads per row: at short rows G*T is a fraction of
  // a wave)
  const int Gf = std::max(c->n, cl.nf_max);
  // and the generated kernel's variants (those built for every group;
  // RDCU_BATCH_XINV=v times only variant v, for tests)
  int best_rv = 0;
  const char *xv = getenv("RDCU_BATCH_XINV");
  const int only = xv && *xv ? atoi(xv) : -1;
  for (int rv = 0; rv < 4; ++rv) {
    bool all = true;
    for (const BatchGroup &gr : cl.groups) all = all && gr.fn && (rv == 0 || gr.fnv[rv]);
    if ((rv > 0 && !all) || (only >= 0 && rv != only)) continue;
    cl.rv = rv;
    for (int G = c->n; G * T <= 256; G = G < Gf ? G + 1 : (G == Gf ? pow2_above(Gf) : 2 * G)) {
      for (int R = 1; R * G * T <= 256 && c->rows % R == 0 && (R == 1 || cl.Ex <= 8); R *= 2) {
        if (rowi_lds(cl, G, R) > 160 * 1024) break;
        double tm = time_of([&]() { rowi_b(b, cl, rd::SRC_Y, rd::DST_KS, 1, 0, none, G, R, 0); });
        snprintf(line, sizeof line, "row_inv v=%d G=%d R=%d %.1fus; ", rv, G, R, tm * 1e3);
        cl.tune_log += line;
        if (tm < best) {
          best = tm;
          cl.G = G;
          cl.R = R;
          best_rv = rv;
        }
      }
    }
  }
  cl.rv = best_rv;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHECK(hipGetLastError());
  return RDCU_OK;
}

// Specialised x-inverse kernels of every (class, dictionary structure) pair,
// compiled as one hiprtc program: kernel rdcu_rowi_b_<class>_<group>.
void build_jit_b(rdcu_batch *b) {
  if (!jit_enabled()) return;
  std::string src = "#include \"rd_kernels_reg.h\"\n";
  std::vector<std::string> names;
  for (size_t ci = 0; ci < b->classes.size(); ++ci) {
    BatchClass &cl = b->classes[ci];
    for (size_t gi = 0; gi < cl.groups.size(); ++gi) {
      const rdcu_ctx *m = b->members[cl.groups[gi].all[0]].get();
      const std::string tag = std::to_string(ci) + "_" + std::to_string(gi);
      src += jit_dict_source(m->n, m->n + m->nf_used, m->h_tfield, m->h_tstart, m->h_fslot,
                             m->h_fpow, "JitDict_" + tag);
      // sizes shared by the group (one grid, one dictionary structure) as
      // compile-time constants, as for a single problem's kernel
      const rd::Grid &gd = m->grid;
      char gbuf[512];
      snprintf(gbuf, sizeof gbuf,
               "namespace rd {\nstruct JitGeom_%s {\n  static constexpr int n = %d, ndim = %d, "
               "Nx = %d, Nxp = %d, nf = %d;\n  static constexpr long N = %ldL, rows = %ldL;\n};\n}\n",
               tag.c_str(), gd.n, gd.ndim, gd.Nx, gd.Nxp, m->fl_used.nf, gd.N, gd.rows);
      src += gbuf;
      // the kernel, and (8 values per thread) its variants with the twiddles
      // in registers and compiled lean for two wavefronts per SIMD
      static const struct { const char *sfx, *bounds, *args; } vs[] = {
          {"", "256", ""}, {"_v1", "256", ", true, false"},
          {"_v2", "256, 2", ", false, true"}, {"_v3", "256, 2", ", true, true"}};
      for (int v = 0; v < (cl.Ex == 8 ? 4 : 1); ++v) {
        src += std::string("extern \"C\" __global__ __launch_bounds__(") + vs[v].bounds +
               ") void rdcu_rowi_b_" + tag + vs[v].sfx +
               "(const rd::Inst *__restrict__ in, int ninst, int rows, rd::FftPlan px, const "
               "rd::cx *__restrict__ twN, int src_kind, int dst_kind, int dst_slot, int err, "
               "rd::ErrCoef ec, rd::StageCoef nx, int G, int check, long region) {\n";
        src += "  rd::rowi_batch<" + std::to_string(cl.Ex) + ", rd::JitDict_" + tag + ", " +
               std::to_string(m->L) + ", rd::JitGeom_" + tag + vs[v].args +
               ">(in, ninst, rows, px, twN, src_kind, dst_kind, dst_slot, err, ec, nx, G, check, "
               "region);\n}\n";
      }
    }
  }
  std::vector<char> code;
  std::string log;
  if (!jit_compile(src, jit_arch(b->dev), code, log)) {
    b->jit_log = log;
    return;
  }
  hipModule_t mod = nullptr;
  if (hipModuleLoadData(&mod, code.data()) != hipSuccess) {
    b->jit_log = "module load failed";
    return;
  }
  b->jit_mod = mod;
  // kernels that fail to load or spill are left out above; their lookup
  // errors are not kept as the thread's last error
  struct ClearLast {
    ~ClearLast() { (void)hipGetLastError(); }
  } clear_last;
  for (size_t ci = 0; ci < b->classes.size(); ++ci) {
    BatchClass &cl = b->classes[ci];
    for (size_t gi = 0; gi < cl.groups.size(); ++gi) {
      const std::string nm = "rdcu_rowi_b_" + std::to_string(ci) + "_" + std::to_string(gi);
      hipFunction_t f = nullptr;
      if (hipModuleGetFunction(&f, mod, nm.c_str()) == hipSuccess && !jit_spills(f))
        cl.groups[gi].fn = f;
      // a variant that does not load or would spill is left out
      for (int v = 1; cl.groups[gi].fn && cl.Ex == 8 && v < 4; ++v) {
        const std::string nv = nm + "_v" + std::to_string(v);
        f = nullptr;
        if (hipModuleGetFunction(&f, mod, nv.c_str()) == hipSuccess && !jit_spills(f))
          cl.groups[gi].fnv[v] = f;
      }
    }
  }
  b->jit_log = "ok";
}

std::string structure_key(const rdcu_ctx *m) {
  std::string k;
  auto add = [&](const std::vector<int> &v) {
    for (int x : v) k += std::to_string(x) + ",";
    k += "|";
  };
  add(m->h_tfield);
  add(m->h_tstart);
  add(m->h_fslot);
  add(m->h_fpow);
  return k;
}

// Instance table of the members with keep[i] set, class by class and group
// by group (creation order within a group), uploaded to the device.
int build_table(rdcu_batch *b, const std::vector<char> &keep) {
  std::vector<rd::Inst> tab;
  b->tab.clear();
  bool full = true;
  for (BatchClass &cl : b->classes) {
    cl.off = (int)tab.size();
    for (BatchGroup &gr : cl.groups) {
      gr.i0 = (int)tab.size() - cl.off;
      for (int i : gr.all) {
        if (!keep[i]) {
          full = false;
          continue;
        }
        rdcu_ctx *m = b->members[i].get();
        tab.push_back(rd::Inst{m->grid, m->bufs, m->st, m->fl_used, m->dictp});
        b->tab.push_back(i);
      }
      gr.i1 = (int)tab.size() - cl.off;
    }
    cl.cnt = (int)tab.size() - cl.off;
  }
  b->table_full = full;
  if (!tab.empty())
    HIPCHECK(hipMemcpyAsync(b->d_inst, tab.data(), sizeof(rd::Inst) * tab.size(),
                            hipMemcpyHostToDevice, b->stream));
  HIPCHECK(hipStreamSynchronize(b->stream));
  return RDCU_OK;
}

// Captures one attempt of each class's instances in the current table (the
// caller has synchronised the class streams: their previous graphs are idle).
int capture_b(rdcu_batch *b) {
  b->captured = false;
  for (size_t c = 0; c < b->classes.size(); ++c) {
    if (b->cexec[c]) (void)hipGraphExecDestroy(b->cexec[c]);
    b->cexec[c] = nullptr;
    const BatchClass &cl = b->classes[c];
    if (cl.cnt == 0) continue;
    HIPCHECK(hipStreamBeginCapture(b->cstream[c], hipStreamCaptureModeThreadLocal));
    b->ls = b->cstream[c];
    int st = enqueue_attempt_b(b, cl);
    b->ls = b->stream;
    hipGraph_t graph = nullptr;
    hipError_t ce = hipStreamEndCapture(b->cstream[c], &graph);
    if (st) {
      if (graph) (void)hipGraphDestroy(graph);
      return st;
    }
    if (ce != hipSuccess) return fail(RDCU_EDEVICE, std::string("capture: ") + hipGetErrorString(ce));
    const hipError_t ie = hipGraphInstantiate(&b->cexec[c], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess)
      return fail(RDCU_EDEVICE, std::string("instantiate: ") + hipGetErrorString(ie));
  }
  b->captured = true;
  return RDCU_OK;
}

int full_table(rdcu_batch *b) {
  if (b->table_full && b->captured) return RDCU_OK;
  int st = build_table(b, std::vector<char>(b->members.size(), 1));
  if (st) return st;
  return capture_b(b);
}

int read_states_b(rdcu_batch *b) {
  HIPCHECK(hipMemcpyAsync(b->h_st, b->d_st, sizeof(rd::DevState) * ninst(b),
                          hipMemcpyDeviceToHost, b->stream));
  HIPCHECK(hipStreamSynchronize(b->stream));
  b->nsync++;
  return RDCU_OK;
}

int begin_b(rdcu_batch *b, const double *t, const double *h, const double *t1, double hmin,
            double hmax, int nolog) {
  const int n = ninst(b);
  double *ht = b->h_buf, *hh = b->h_buf + n, *ht1 = b->h_buf + 2 * n;
  for (int i = 0; i < n; ++i) {
    const int p = b->pos[i];
    ht[p] = t[i];
    hh[p] = h[i];
    ht1[p] = t1 ? t1[i] : INFINITY;
  }
  HIPCHECK(hipMemcpyAsync(b->d_t, ht, 3 * sizeof(double) * n, hipMemcpyHostToDevice, b->stream));
  hipLaunchKernelGGL(rd::k_begin_b, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream,
                     b->d_inst, n, b->d_t, b->d_h, b->d_t1, hmin, hmax, nolog);
  HIPCHECK(hipGetLastError());
  return RDCU_OK;
}

int ensure_k1_b(rdcu_batch *b, bool force) {
  bool ok = b->k1_valid && !force;
  for (auto &m : b->members) ok = ok && m->k1_valid;
  if (ok) return RDCU_OK;
  for (const BatchClass &cl : b->classes) rhs_b(b, cl, rd::DST_K1, 0, 0);
  HIPCHECK(hipGetLastError());
  for (auto &m : b->members) m->k1_valid = true;
  b->k1_valid = true;
  return RDCU_OK;
}

}  // namespace

extern "C" {

int rdcu_batch_create(const rdcu_problem *probs, int count, const double *const *y0, double atol,
                      double rtol, int fixedstep, int device, rdcu_batch **out) {
  if (!out) return fail(RDCU_EINVAL, "null output handle");
  *out = nullptr;
  if (!probs || count < 1) return fail(RDCU_EINVAL, "a batch needs at least one problem");
  for (int i = 0; i < count; ++i) {
    if (probs[i].ndim > 2) return fail(RDCU_EUNSUPPORTED, "batched instances are 1-D or 2-D");
    if (probs[i].n < 1 || probs[i].n > rd::kNMax) return fail(RDCU_EINVAL, "bad field count");
    long rows = 1;
    for (int a = 0; a + 1 < probs[i].ndim; ++a) rows *= probs[i].Ns[a];
    if ((long)probs[i].n * rows > rd::kRedCap)
      return fail(RDCU_EUNSUPPORTED, "batched instances have at most 4096 grid lines");
  }
  std::unique_ptr<rdcu_batch> b(new rdcu_batch());
  b->dev = device;
  b->fixed = fixedstep ? 1 : 0;
  HIPCHECK(hipSetDevice(device));
  HIPCHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  b->ls = b->stream;
  HIPCHECK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  {
    void *q = nullptr;
    HIPCHECK(hipMalloc(&q, sizeof(rd::DevState) * count));
    b->d_st = static_cast<rd::DevState *>(q);
    HIPCHECK(hipHostMalloc(&q, sizeof(rd::DevState) * count, hipHostMallocDefault));
    b->h_st = static_cast<rd::DevState *>(q);
    HIPCHECK(hipMalloc(&q, sizeof(int) * count));
    b->d_done = static_cast<int *>(q);
    HIPCHECK(hipHostMalloc(&q, sizeof(int) * count, hipHostMallocDefault));
    b->h_done = static_cast<int *>(q);
    HIPCHECK(hipMalloc(&q, 3 * sizeof(double) * count));
    b->d_t = static_cast<double *>(q);
    b->d_h = b->d_t + count;
    b->d_t1 = b->d_t + 2 * count;
    HIPCHECK(hipHostMalloc(&q, 3 * sizeof(double) * count, hipHostMallocDefault));
    b->h_buf = static_cast<double *>(q);
  }
  // classes: same grid and transform lengths; within a class, instances
  // with the same dictionary structure are contiguous
  struct Key {
    int n, ndim, Ns[3];
    bool operator<(const Key &o) const {
      if (n != o.n) return n < o.n;
      if (ndim != o.ndim) return ndim < o.ndim;
      for (int a = 0; a < 3; ++a)
        if (Ns[a] != o.Ns[a]) return Ns[a] < o.Ns[a];
      return false;
    }
  };
  std::map<Key, std::vector<int>> byclass;
  for (int i = 0; i < count; ++i) {
    Key k{probs[i].n, probs[i].ndim, {1, 1, 1}};
    for (int a = 0; a < probs[i].ndim; ++a) k.Ns[a] = probs[i].Ns[a];
    byclass[k].push_back(i);
  }
  // table order: class by class; members are created in that order so that
  // their device states are contiguous in table order
  std::vector<int> order;
  b->pos.assign(count, 0);
  b->members.resize(count);
  int st = RDCU_OK;
  for (auto &kv : byclass) {
    // structure groups need the members' dictionaries: create first, sort after
    for (int i : kv.second) order.push_back(i);
  }
  for (int p = 0; p < count; ++p) {
    const int i = order[p];
    MemberOpts mo{b->stream, b->d_st + p, b->h_st + p};
    std::unique_ptr<rdcu_ctx> m;
    if ((st = create_ctx(&probs[i], atol, rtol, fixedstep, y0 ? y0[i] : nullptr, device, &mo, m)))
      return fail(st, "instance " + std::to_string(i) + ": " + g_err);
    b->members[i] = std::move(m);
  }
  // within each class, order by dictionary structure; device states follow
  // the members, so rebuild the table order and move the states with it
  std::vector<int> final_order;
  int off = 0;
  for (auto &kv : byclass) {
    BatchClass cl;
    std::map<std::string, std::vector<int>> bys;
    for (int i : kv.second) bys[structure_key(b->members[i].get())].push_back(i);
    cl.off = off;
    for (auto &sv : bys) {
      BatchGroup gr;
      gr.i0 = (int)cl.members.size();
      for (int i : sv.second) cl.members.push_back(i);
      gr.i1 = (int)cl.members.size();
      gr.all = sv.second;
      cl.groups.push_back(gr);
    }
    cl.cnt = (int)cl.members.size();
    for (int i : cl.members) final_order.push_back(i);
    off += (int)cl.members.size();
    b->classes.push_back(std::move(cl));
  }
  // device state slot of member i must be its table position: reassign
  {
    std::vector<rd::DevState> hs(count);
    for (int p = 0; p < count; ++p) {
      const int i = final_order[p];
      hs[p] = *b->members[i]->sh;
    }
    for (int p = 0; p < count; ++p) {
      const int i = final_order[p];
      b->pos[i] = p;
      b->members[i]->st = b->d_st + p;
      b->members[i]->sh = b->h_st + p;
      b->h_st[p] = hs[p];
    }
    HIPCHECK(hipMemcpy(b->d_st, b->h_st, sizeof(rd::DevState) * count, hipMemcpyHostToDevice));
  }
  // instance table (the full table's order is final_order: position = state slot)
  {
    void *q = nullptr;
    HIPCHECK(hipMalloc(&q, sizeof(rd::Inst) * count));
    b->d_inst = static_cast<rd::Inst *>(q);
    if ((st = build_table(b.get(), std::vector<char>(count, 1)))) return st;
  }
  if ((st = once_per_device(b->dev, 1, allow_big_lds_b)) ||
      (st = once_per_device(b->dev, 0, allow_big_lds)))
    return st;
  for (BatchClass &cl : b->classes) {
    cl.c0 = b->members[cl.members[0]].get();
    cl.Ex = cl.c0->Ex;
    cl.E0 = cl.c0->E0;
    cl.rows = (int)cl.c0->rows;
    for (int i : cl.members) {
      rdcu_ctx *m = b->members[i].get();
      cl.nf_max = std::max(cl.nf_max, m->nf_used);
      cl.nwords_max = std::max(cl.nwords_max, m->dictp.nwords);
      cl.max_fpf = std::max(cl.max_fpf, m->max_fpf);
    }
    cl.G = cl.c0->n;
    cl.fpw = cl.max_fpf;
  }
  build_jit_b(b.get());
  for (BatchClass &cl : b->classes) {
    if (rowi_lds(cl, cl.c0->n, 1) > 160 * 1024)
      return fail(RDCU_EUNSUPPORTED, "batched instance too large for one block per row");
    const char *shp = getenv("RDCU_SHAPES");
    if (!(shp && *shp)) {
      if ((st = autotune_b(b.get(), cl))) return st;
    }
    // members use the class shapes for their own plain-RHS calls
    for (int i : cl.members) {
      rdcu_ctx *m = b->members[i].get();
      m->rf_rpb = cl.rpb;
      m->cr_tc = cl.tc;
      m->cr_fpw = cl.fpw;
      m->ri_G = cl.G;
      m->ri_R = cl.R;
    }
    const char *vb = getenv("RDCU_VERBOSE");
    if (vb && *vb && *vb != '0')
      fprintf(stderr, "rdcu batch class %dx%d (%zu instances, %zu structures, jit %s): %s -> "
                      "%d,%d,%d,%d,%d x-inverse variant %d x-forward grid %ld\n",
              cl.c0->Ns[0], cl.c0->Ns[1], cl.members.size(), cl.groups.size(),
              b->jit_log.c_str(), cl.tune_log.c_str(), cl.rpb, cl.tc, cl.fpw, cl.G, cl.R, cl.rv,
              cl.grid_rf);
  }
  for (size_t c = 0; c < b->classes.size(); ++c) {  // one stream per class (default priority)
    hipStream_t cs = nullptr;
    hipEvent_t ce = nullptr;
    HIPCHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    b->cstream.push_back(cs);
    HIPCHECK(hipEventCreateWithFlags(&ce, hipEventDisableTiming));
    b->cjoin.push_back(ce);
    b->cexec.push_back(nullptr);
  }
  if ((st = capture_b(b.get()))) return st;
  HIPCHECK(hipStreamSynchronize(b->stream));
  *out = b.release();
  return RDCU_OK;
}

int rdcu_batch_destroy(rdcu_batch *b) {
  if (b) (void)hipSetDevice(b->dev);
  delete b;
  return RDCU_OK;
}

int rdcu_batch_member(rdcu_batch *b, int i, rdcu_ctx **out) {
  if (!b || !out) return fail(RDCU_EINVAL, "null argument");
  if (i < 0 || i >= ninst(b)) return fail(RDCU_EINVAL, "instance index out of range");
  *out = b->members[i].get();
  return RDCU_OK;
}

int rdcu_batch_time(rdcu_batch *b, int reps, double *ms, double *bytes, int *shape,
                    int max_classes, int *nclasses) {
  if (!b || !ms || reps < 1 || max_classes < 0) return fail(RDCU_EINVAL, "bad argument");
  HIPCHECK(hipSetDevice(b->dev));
  int st;
  if ((st = full_table(b))) return st;  // every instance in the table
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  auto timed = [&](const std::function<void()> &body, double *out) -> int {
    for (int r = 0; r < 2; ++r) body();  // warm
    HIPCHECK(hipEventRecord(e0, b->stream));
    for (int r = 0; r < reps; ++r) body();
    HIPCHECK(hipEventRecord(e1, b->stream));
    HIPCHECK(hipEventSynchronize(e1));
    float f = 0;
    HIPCHECK(hipEventElapsedTime(&f, e0, e1));
    *out = (double)f / reps;
    return RDCU_OK;
  };
  rd::StageCoef none{};
  // the plain RHS into the stage-2 slot (k2), as the tuner does
  st = timed([&]() {
    for (const BatchClass &cl : b->classes) rhs_b(b, cl, rd::DST_KS, 1, 0);
  }, &ms[0]);
  const int nc = (int)b->classes.size();
  for (int k = 0; !st && k < nc && k < max_classes; ++k) {
    const BatchClass &cl = b->classes[k];
    const rdcu_ctx *c0 = cl.c0;
    double *o = ms + 1 + 3 * k;
    st = timed([&]() { rowf_b(b, cl, rd::SRC_Y, none, cl.rpb, cl.grid_rf, 0); }, &o[0]);
    if (!st && c0->ndim >= 2) st = timed([&]() { col_b(b, cl, cl.tc, cl.fpw, 0); }, &o[1]);
    else o[1] = 0.0;
    if (!st)
      st = timed([&]() { rowi_b(b, cl, rd::SRC_Y, rd::DST_KS, 1, 0, none, cl.G, cl.R, 0); },
                 &o[2]);
    if (bytes) {
      double bf = 0, bc = 0, bi = 0;
      for (int i : cl.members) {
        const rdcu_ctx *m = b->members[i].get();
        const double nN = (double)m->n * m->N, MN = (double)m->nf_used * m->N;
        bf += 16 * nN;
        if (m->ndim >= 2) {
          bc += 8 * nN + 8 * MN;
          bi += 8 * MN + 16 * nN;
        } else {
          bi += 24 * nN;
        }
      }
      bytes[3 * k] = bf;
      bytes[3 * k + 1] = bc;
      bytes[3 * k + 2] = bi;
    }
    if (shape) {
      shape[4 * k] = c0->ndim;
      shape[4 * k + 1] = c0->Ns[0];
      shape[4 * k + 2] = c0->ndim >= 2 ? c0->Ns[1] : 1;
      shape[4 * k + 3] = (int)cl.members.size();
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) return st;
  HIPCHECK(hipGetLastError());
  if (nclasses) *nclasses = nc;
  return RDCU_OK;
}

int rdcu_batch_step(rdcu_batch *b, double *t, double *h, double hmin, double hmax, int *accepted) {
  (void)hmin;  // as rdcu_step: one attempt regardless of hmin
  if (!b || !t || !h || !accepted) return fail(RDCU_EINVAL, "null argument");
  HIPCHECK(hipSetDevice(b->dev));
  int st;
  if ((st = full_table(b))) return st;
  if ((st = begin_b(b, t, h, nullptr, -INFINITY, hmax, 0))) return st;
  if ((st = ensure_k1_b(b, false))) return st;
  if ((st = replay_b(b, 1))) return st;
  if ((st = read_states_b(b))) return st;
  for (int i = 0; i < ninst(b); ++i) {
    const rd::DevState &s = b->h_st[b->pos[i]];
    if (s.nlog != 1) return fail(RDCU_ESTATE, "attempt of instance " + std::to_string(i) + " did not run");
    accepted[i] = s.log_acc[0];
    t[i] = s.t;
    h[i] = s.h;
    b->members[i]->nrhs += 6;
  }
  return RDCU_OK;
}

int rdcu_batch_run(rdcu_batch *b, double *t, double *h, double hmin, double hmax, const double *t1,
                   long long max_attempts, long long *attempts) {
  if (!b || !t || !h || !t1) return fail(RDCU_EINVAL, "null argument");
  if (max_attempts < 0) return fail(RDCU_EINVAL, "max_attempts must be >= 0");
  HIPCHECK(hipSetDevice(b->dev));
  const int n = ninst(b);
  std::vector<long long> a0(n);
  for (int i = 0; i < n; ++i) a0[i] = b->h_st[b->pos[i]].attempts;
  int st;
  if ((st = full_table(b))) return st;
  if ((st = begin_b(b, t, h, t1, hmin, hmax, 1))) return st;
  // dp45_run computes k1 = f(y) at the start of a run (dp45.cu:157)
  if ((st = ensure_k1_b(b, true))) return st;
  long long done = 0;
  bool all = false;
  // with RDCU_VERBOSE, one line per table: the graph replays it ran, its
  // instances, and the wall time it took (where a sweep's time goes)
  const char *vb = getenv("RDCU_VERBOSE");
  const bool verbose = vb && *vb && *vb != '0';
  auto now = []() { return std::chrono::steady_clock::now(); };
  auto t_tab = now();
  long long tab_replays = 0;
  while (!all && done < max_attempts) {
    const long long k = std::min<long long>(kChunkMax, max_attempts - done);
    if ((st = replay_b(b, k))) return st;
    done += k;
    tab_replays += k;
    const int tn = (int)b->tab.size();
    HIPCHECK(hipMemcpyAsync(b->h_done, b->d_done, sizeof(int) * tn, hipMemcpyDeviceToHost,
                            b->stream));
    HIPCHECK(hipStreamSynchronize(b->stream));
    b->nsync++;
    int active = 0;
    std::vector<char> keep(n, 0);
    for (int p = 0; p < tn; ++p)
      if (!b->h_done[p]) {
        ++active;
        keep[b->tab[p]] = 1;
      }
    all = active == 0;
    if (verbose && (all || done >= max_attempts || 2 * active <= tn)) {
      const double sec = std::chrono::duration<double>(now() - t_tab).count();
      fprintf(stderr, "rdcu batch table: %d instances, %lld replays, %.3f s (%.1f us per replay)\n",
              tn, tab_replays, sec, 1e6 * sec / (double)tab_replays);
      t_tab = now();
      tab_replays = 0;
    }
    // once half the table has finished, later attempts launch only the
    // instances still integrating (a new table and graph)
    if (!all && done < max_attempts && 2 * active <= tn) {
      if ((st = build_table(b, keep)) || (st = capture_b(b))) return st;
    }
  }
  if ((st = read_states_b(b))) return st;
  for (int i = 0; i < n; ++i) {
    const rd::DevState &s = b->h_st[b->pos[i]];
    t[i] = s.t;
    h[i] = s.h;
    const long long na = s.attempts - a0[i];
    if (attempts) attempts[i] = na;
    b->members[i]->nrhs += 6 * na;
  }
  return RDCU_OK;
}

}  // extern "C"
