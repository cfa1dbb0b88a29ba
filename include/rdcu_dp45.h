/*
 * rdcu_dp45.h -- C ABI of the explicit (Dormand-Prince 4/5, pseudospectral) integrator.
 *
 * This is the drop-in boundary for rdcu's non-stiff path.  It replaces the five
 * module-static calls of the reference integrator
 *     dp45_init / dp45_step / dp45_run / dp45_eval / dp45_destroy   (dp45.h:7-11)
 * together with the RHS callback the reference driver plugs into them
 *     dydt / makeY                                                  (rdcu.cu:117-184)
 * by handle-based calls with status codes.  The RHS callback becomes a
 * dictionary descriptor (rdcu_problem) because the RHS is evaluated on the device.
 *
 * Two libraries export exactly these symbols:
 *   rdcu_amd/lib/librdcu_hip.so   -- the MI355X (gfx950) implementation, HIP kernels
 *   oracle/_bin/liboracle.so      -- the FP64 CPU restatement used as the checker
 * Plain pointers and sizes only; every array argument is host memory unless a
 * function name ends in _device.
 */
#ifndef RDCU_DP45_H
#define RDCU_DP45_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference returns nothing and checks no CUDA status
 * (dp45.cu:87-254); every call here returns one of these. */
enum {
  RDCU_OK = 0,
  RDCU_EINVAL = 1,       /* bad argument (sizes, indices, null pointers)            */
  RDCU_EUNSUPPORTED = 2, /* valid for the reference but not built here (odd N ...)  */
  RDCU_EDEVICE = 3,      /* a HIP runtime call failed                               */
  RDCU_ENOMEM = 4,       /* allocation failed                                       */
  RDCU_ESTATE = 5,       /* call not valid in the current state (eval before a step) */
  RDCU_ECALLBACK = 6     /* the step callback asked to stop                         */
};

/* Problem description: grid, domain and monomial dictionary.
 * Replaces the fields of the reference's `parameters` struct that the RHS reads
 * (rdcu.cu:17-76: n, ndim, Ns, Ls, c, nprods, C, nterms).
 * Term i contributes  coef[i] * prod_j Y[fac[2j]]^fac[2j+1]  to field[i]
 * (rdcu.cu:175-183); its pairs start at offset 2*sum_{i'<i} nfac[i'] in fac.
 * Factor index k means Y block k/n of field k%n: block 0 = u, 1+a = d/dx_a,
 * 1+ndim+a*ndim+b = d/dx_b d/dx_a (rdcu.cu:155-161, 769-785).  Axis 0 is the
 * slowest-varying axis of the row-major grid. */
typedef struct rdcu_problem {
  int n;              /* number of fields (-n)                 */
  int ndim;           /* 1, 2 or 3 (number of values in -N)    */
  int Ns[3];          /* grid points per axis (-N)             */
  double Ls[3];       /* domain length per axis (-L)           */
  int nterms;         /* number of dictionary terms            */
  const double *coef; /* [nterms]                              */
  const int *field;   /* [nterms]                              */
  const int *nfac;    /* [nterms]                              */
  const int *fac;     /* [2 * sum(nfac)] (index, power) pairs  */
} rdcu_problem;

typedef struct rdcu_ctx rdcu_ctx;

/* Creates an integrator for one problem on device `device` (ignored by the CPU
 * oracle) with initial state y0[n*N] (field-major, row-major grid).
 * Replaces dp45_init (dp45.h:8, dp45.cu:171-225) and the FFT plan creation
 * (rdcu.cu:899-900).  fixedstep != 0 selects -F (every attempt accepted). */
int rdcu_create(const rdcu_problem *prob, double atol, double rtol, int fixedstep,
                const double *y0, int device, rdcu_ctx **out);

/* A right-hand side given as code: the dydt callback of dp45_init (dp45.h:8,
 * dp45.cu:171-177), called as dydt(t + c_s*h, ytemp, k_s) at every stage
 * (dp45.cu:90-106) and dydt(t, y, k1) first (dp45.cu:157).  fn must make
 * f_dev[0 .. n*N) = dydt(t, y_dev) with work ordered after everything already
 * on `stream` (the integrator's hipStream_t) and complete, or ordered before
 * anything enqueued on `stream` later, when it returns: the simplest correct
 * callback synchronises `stream`, computes, and synchronises its own work.
 * y_dev and f_dev are device memory (host memory in the CPU oracle, where
 * stream is NULL).  A nonzero return aborts the call with RDCU_ECALLBACK. */
typedef int (*rdcu_dydt_fn)(double t, const double *y_dev, double *f_dev, void *stream,
                            void *user);

/* Creates an integrator for a state of n*N values (n fields of N values; the
 * error norm sums squares per field by the pairwise tree, then over the
 * fields) whose right-hand side is fn.  Replaces dp45_init with an arbitrary
 * dydt (dp45.h:8).  The same step control, run loop and dense output as a
 * dictionary problem, on the device: rdcu_step, rdcu_run, rdcu_advance,
 * rdcu_eval, rdcu_get_state, rdcu_set_state, rdcu_get_stats and rdcu_destroy
 * take the handle; calls that need the dictionary (rdcu_rhs, rdcu_frame with
 * f or Y, rdcu_time_rhs, coefficients and gradients) return
 * RDCU_EUNSUPPORTED.  Attempts run one at a time, each with one host
 * synchronisation (the callback sees t and h). */
int rdcu_create_fn(int n, long long N, double atol, double rtol, int fixedstep, const double *y0,
                   rdcu_dydt_fn fn, void *user, int device, rdcu_ctx **out);

/* Frees everything.  Replaces dp45_destroy (dp45.h:10, dp45.cu:227-254). */
int rdcu_destroy(rdcu_ctx *ctx);

/* One DP45 attempt from (*t, *h).  *accepted = 1 on acceptance (t and h
 * advanced, hmax applied), 0 on rejection (h shrunk).  Replaces dp45_step
 * (dp45.h:7, dp45.cu:87-147); the FSAL copy k1 <- k7 the reference does in
 * dp45_run (dp45.cu:165-167) is done here after an accepted attempt. */
int rdcu_step(rdcu_ctx *ctx, double *t, double *h, double hmin, double hmax, int *accepted);

/* Per-attempt callback of rdcu_run, the role of step_eval (rdcu.cu:252-341).
 * t, h: values after the attempt; success: 1 accepted / 0 rejected.
 * current: 1 if the integrator still holds this attempt's stages, so that
 * rdcu_eval and rdcu_get_state refer to it.  *stop: the stop time in force;
 * the callback may lower or raise it.  Every accepted attempt that reaches
 * t >= *stop is reported with current == 1.  A nonzero return aborts the run. */
typedef int (*rdcu_step_fn)(rdcu_ctx *ctx, double t, double h, int success, int current,
                            double *stop, void *user);

/* Integrates while (*t < t1 && *h > hmin), calling fn after every attempt.
 * Computes k1 = f(y) first, as dp45_run does (dp45.cu:157).  `stop` is the
 * first stop time (normally the first output time).  Replaces dp45_run
 * (dp45.h:9, dp45.cu:154-169). */
int rdcu_run(rdcu_ctx *ctx, double *t, double *h, double hmin, double hmax, double t1,
             double stop, rdcu_step_fn fn, void *user);

/* Runs attempts while (*t < t1 && *h > hmin), at most max_attempts of them,
 * with no per-attempt callback (a run without outputs, e.g. a burn-in or a
 * benchmark).  *done receives the number of attempts run.  Same loop and step
 * control as rdcu_run (dp45.cu:154-169). */
int rdcu_advance(rdcu_ctx *ctx, double *t, double *h, double hmin, double hmax, double t1,
                 long long max_attempts, long long *done);

/* Timing of the RHS pipeline on the library's stream with HIP events.
 * ms[0]: one RHS evaluation of the current state (all its kernels), averaged
 * over `reps` back-to-back evaluations; ms[1..count-1]: each kernel of that
 * pipeline launched alone `reps` times, averaged.  `names` receives the
 * comma-separated kernel names ("rhs,<k1>,<k2>,...").  The CPU oracle fills
 * ms[0] only (wall clock). */
int rdcu_time_rhs(rdcu_ctx *ctx, int reps, double *ms, int max, int *count, char *names,
                  int names_len);

/* Dense output of the last accepted step at t_eval into y_out[n*N].
 * Replaces dp45_eval (dp45.h:11, dp45.cu:149-152), which returned a borrowed
 * device pointer; this copies out. */
int rdcu_eval(rdcu_ctx *ctx, double t_eval, double *y_out);

/* Copies the current state y into y_out[n*N] (the D2H before fs.dat, rdcu.cu:320). */
int rdcu_get_state(rdcu_ctx *ctx, double *y_out);

/* Replaces the current state (invalidates the FSAL stage). */
int rdcu_set_state(rdcu_ctx *ctx, const double *y);

/* Evaluates the dictionary RHS at y[n*N]: f_out[n*N] (may be NULL) and, if
 * Y_out is not NULL, all n*(1+ndim+ndim^2) factor blocks as complex values
 * Y_out[2*((blk*n+field)*N + x) + {0,1}] in the layout of FILEBASEY.dat.
 * Replaces dydt/makeY (rdcu.cu:117-184).  The imaginary parts are those of
 * the reference's complex transform (rdcu.cu:151-164): the Nyquist leakage of
 * the first and mixed derivative blocks; 0 in block 0 (uninitialised memory
 * in the reference) and in the pure second derivatives (roundoff there). */
int rdcu_rhs(rdcu_ctx *ctx, const double *y, double *f_out, double *Y_out);

/* The dictionary RHS on device memory: f_dev[n*N] = dydt(y_dev[n*N]), the
 * values of rdcu_rhs, complete when the call returns (the work runs on the
 * handle's own stream).  For a callback of rdcu_create_fn that evaluates a
 * dictionary, and for callers whose state already lives on the device.  The
 * CPU oracle takes host memory. */
int rdcu_rhs_device(rdcu_ctx *ctx, const double *y_dev, double *f_dev);

/* rdcu_rhs that allocates nothing: the same values, bit for bit, with every
 * intermediate device buffer of the call (input staging, x-forward spectra,
 * factor spectra, output staging, factor blocks) taken from one static device
 * array of 128 MiB per GPU, which each GPU the library runs on holds until the
 * process ends.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory during their loop (tight serving or fitting loops,
 * harnesses that forbid hipMalloc); the context itself is created beforehand.
 * A problem needing more scratch than the array holds is rejected with
 * RDCU_EINVAL.  Calls on one GPU share its array: within a process they are
 * serialised (concurrent callers wait); nothing in it needs to start at zero,
 * and it keeps the last call's contents. */
int rdcu_rhs_noalloc(rdcu_ctx *ctx, const double *y, double *f_out, double *Y_out);

/* One output frame of step_eval (rdcu.cu:252-341): the dense output at t_eval
 * into y_out[n*N] (as rdcu_eval) and, when not NULL, the RHS at that state
 * into f_out and its factor blocks into Y_out (as rdcu_rhs on y_out).  Same
 * values, bit for bit, as those two calls; the state stays on the device in
 * between, and Y's complex layout is assembled there.  Replaces the dp45_eval
 * / cudaMemcpy / dydt / makeY sequence of rdcu.cu:270-291. */
int rdcu_frame(rdcu_ctx *ctx, double t_eval, double *y_out, double *f_out, double *Y_out);

/* Page-locked host memory (*out = NULL on failure) for the buffers of
 * rdcu_frame, rdcu_eval, rdcu_rhs and rdcu_get_state: the copies into it run
 * as DMA at full link speed.  The reference copies into pageable memory
 * (rdcu.cu:277, 283, 288, 320).  The CPU build returns ordinary memory. */
int rdcu_host_alloc(size_t bytes, void **out);
int rdcu_host_free(void *p);

/* ---- Batches: many independent problems integrated side by side ----------
 * The reference runs a parameter sweep as one process per instance (a SLURM
 * job array, sweep.sh:21-41, each running makelines.py:260-268 -> ./rdcu).
 * A batch holds `count` problems (any mix of 1-D/2-D grids and dictionaries;
 * transform lengths 2^k or 3*2^k) with initial states y0[i][n_i*N_i].  Every
 * instance has its own t, h and step history, computed with exactly the
 * arithmetic of a single rdcu_ctx on the same problem. */
typedef struct rdcu_batch rdcu_batch;

int rdcu_batch_create(const rdcu_problem *probs, int count, const double *const *y0,
                      double atol, double rtol, int fixedstep, int device, rdcu_batch **out);

int rdcu_batch_destroy(rdcu_batch *batch);

/* One attempt of every instance from (t[i], h[i]), as rdcu_step (dp45.cu:87-147)
 * for each; accepted[i] = 1 or 0. */
int rdcu_batch_step(rdcu_batch *batch, double *t, double *h, double hmin, double hmax,
                    int *accepted);

/* Runs every instance while (t[i] < t1[i] && h[i] > hmin), at most max_attempts
 * attempts each, as rdcu_advance does for one (dp45.cu:154-169, k1 = f(y) first).
 * attempts[i] (may be NULL) receives the attempts run by instance i. */
int rdcu_batch_run(rdcu_batch *batch, double *t, double *h, double hmin, double hmax,
                   const double *t1, long long max_attempts, long long *attempts);

/* Instance i as an ordinary handle for rdcu_get_state, rdcu_set_state,
 * rdcu_eval, rdcu_rhs, rdcu_get_stats.  It belongs to the batch (do not pass it
 * to rdcu_destroy) and steps only through the batch (rdcu_step, rdcu_run and
 * rdcu_advance return RDCU_ESTATE for it). */
int rdcu_batch_member(rdcu_batch *batch, int i, rdcu_ctx **out);

/* Timing of one plain RHS of every instance, with HIP events on the batch's
 * stream (a diagnostic, like rdcu_time_rhs; no counterpart in the reference,
 * whose sweep runs one process per instance).  All instances, on their current
 * states, with the batch's tuned launch shapes:
 *   ms[0]: every shape class's x-forward, strided and x-inverse launches back to
 *          back, averaged over `reps` evaluations;
 *   ms[1+3k .. 3+3k]: class k's three launches, each alone, averaged over `reps`
 *          (classes ordered by field count, dimension, then grid sizes).
 * bytes[3k .. 3k+2] receive the algorithmic HBM bytes of class k's launches,
 * summed over its instances (half spectra, one read and one write per array per
 * pass: x forward 16nN; strided 8nN + 8MN; x inverse 8MN + 16nN, with M the
 * instance's referenced factors; in 1-D no strided pass and 24nN for the x
 * inverse).  shape[4k .. 4k+3] receives ndim, N0, N1 (1 in 1-D) and the
 * instance count.  Up to max_classes classes are filled (ms needs
 * 1 + 3*max_classes entries); *nclasses receives the number of classes.  Writes
 * every instance's stage-2 buffer, which every attempt overwrites.  The CPU
 * oracle fills ms[0] only (wall clock of the instances' RHS one after another)
 * and sets *nclasses to 0. */
int rdcu_batch_time(rdcu_batch *batch, int reps, double *ms, double *bytes, int *shape,
                    int max_classes, int *nclasses);

/* ---- Differentiation: gradients of an output-matching loss ---------------
 * No counterpart in the reference, whose only derivative code is the Jacobian
 * chain rule of its stiff path (rdcu.cu:204-216).  These serve fitting
 * dictionary coefficients to trajectories. */

/* Replaces the dictionary coefficients (coef[nterms], in the problem's term
 * order); fields, factors and powers stay.  Invalidates the FSAL stage. */
int rdcu_set_coef(rdcu_ctx *ctx, const double *coef);

/* Step schedule of an adaptive run: from the current state at (t0, h0), runs
 * attempts as rdcu_advance does while (t < t1 && h > hmin) and writes the size
 * of every accepted step to hs[0 .. *nsteps).  More than max_steps accepted
 * steps is an error (RDCU_EINVAL).  The state advances. */
int rdcu_schedule(rdcu_ctx *ctx, double t0, double h0, double hmin, double hmax, double t1,
                  long long max_steps, double *hs, long long *nsteps);

/* Loss on a frozen step schedule and its gradient by the discrete adjoint.
 * From y0[n*N] at t0, takes nsteps DP45 steps of sizes hs[] (every step
 * accepted, no error control), evaluates the dense output at t_out[0..nout)
 * (ascending, in (t0, t0 + sum hs]: each in the first step whose end reaches
 * it, as rdcu_run's stops), and returns
 *     loss = sum_o sum_x (y(t_o)[x] - target[o*n*N + x])^2 / (nout*n*N).
 * grad_coef[nterms] and grad_y0[n*N] (either may be NULL; both NULL gives the
 * loss only) receive dloss/dcoef and dloss/dy0 of exactly these steps: a
 * reverse sweep over the steps, each recomputing its stages from a checkpoint
 * of its start state, through the spectral factors and the monomial Jacobian.
 * The state afterwards is undefined (set it before stepping again). */
int rdcu_loss_grad(rdcu_ctx *ctx, const double *y0, double t0, long long nsteps,
                   const double *hs, int nout, const double *t_out, const double *target,
                   double *loss, double *grad_coef, double *grad_y0);

/* rdcu_loss_grad with y0_dev[n*N], target_dev[nout*n*N] and grad_y0_dev[n*N]
 * in device memory (target read in place; the small arrays hs, t_out, loss
 * and grad_coef stay in host memory): the same values, for callers whose
 * trajectories live on the device (torch tensors).  Complete on return.  The
 * CPU oracle takes host memory. */
int rdcu_loss_grad_device(rdcu_ctx *ctx, const double *y0_dev, double t0, long long nsteps,
                          const double *hs, int nout, const double *t_out,
                          const double *target_dev, double *loss, double *grad_coef,
                          double *grad_y0_dev);

/* One instance of a gradient batch: the arguments of rdcu_loss_grad, and its
 * status (an RDCU_* code) on return. */
typedef struct rdcu_grad_job {
  rdcu_ctx *ctx;
  const double *y0;
  double t0;
  long long nsteps;
  const double *hs;
  int nout;
  const double *t_out;
  const double *target;
  double *loss;
  double *grad_coef;
  double *grad_y0;
  int status;
} rdcu_grad_job;

/* rdcu_loss_grad for jobs[0 .. count), each on its own handle (all distinct),
 * up to `concurrency` of them at a time (<= 0: a default).  Every job gives
 * exactly what rdcu_loss_grad gives for it alone.  Returns RDCU_OK, or the
 * first failing job's status (its message is then the last error). */
int rdcu_loss_grad_batch(rdcu_grad_job *jobs, int count, int concurrency);

/* Counters: out[0] attempts, out[1] accepted, out[2] RHS evaluations,
 * out[3] host synchronisations (chunks). */
int rdcu_get_stats(rdcu_ctx *ctx, long long out[4]);

/* Human-readable message for the last failing call on this thread. */
const char *rdcu_last_error(void);

/* Name of the backend: "hip-gfx950" or "cpu-oracle". */
const char *rdcu_backend(void);

#ifdef __cplusplus
}
#endif
#endif /* RDCU_DP45_H */
