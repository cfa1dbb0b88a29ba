"""ctypes binding of the integrator C ABI (include/rdcu_dp45.h).

The same binding drives two shared libraries that export identical symbols:

* ``rdcu_amd/lib/librdcu_hip.so`` -- the MI355X implementation (the product);
* ``oracle/_bin/liboracle.so``   -- the FP64 CPU restatement (test checker only).

``Integrator`` mirrors the reference's integrator module (dp45.h:7-11:
dp45_init / dp45_step / dp45_run / dp45_eval / dp45_destroy) plus the RHS
callback the reference driver plugs into it (dydt/makeY, rdcu.cu:117-184),
with handles instead of module statics and exceptions instead of silent
failures.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIP_LIB = os.path.join(ROOT, "rdcu_amd", "lib", "librdcu_hip.so")

# status codes of rdcu_dp45.h
OK, EINVAL, EUNSUPPORTED, EDEVICE, ENOMEM, ESTATE, ECALLBACK = range(7)
_ERRNAMES = {1: "EINVAL", 2: "EUNSUPPORTED", 3: "EDEVICE", 4: "ENOMEM", 5: "ESTATE", 6: "ECALLBACK"}

SYMBOLS = (
    "rdcu_create", "rdcu_destroy", "rdcu_step", "rdcu_run", "rdcu_advance", "rdcu_time_rhs",
    "rdcu_eval", "rdcu_get_state", "rdcu_set_state", "rdcu_rhs", "rdcu_get_stats",
    "rdcu_last_error", "rdcu_backend",
    "rdcu_batch_create", "rdcu_batch_destroy", "rdcu_batch_step", "rdcu_batch_run",
    "rdcu_batch_member", "rdcu_set_coef", "rdcu_schedule", "rdcu_loss_grad",
    "rdcu_loss_grad_batch", "rdcu_frame", "rdcu_host_alloc", "rdcu_host_free",
    "rdcu_rhs_noalloc", "rdcu_create_fn", "rdcu_rhs_device", "rdcu_loss_grad_device",
    "rdcu_batch_time",
)


class RdcuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class _Problem(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("ndim", ctypes.c_int), ("Ns", ctypes.c_int * 3),
        ("Ls", ctypes.c_double * 3), ("nterms", ctypes.c_int),
        ("coef", ctypes.POINTER(ctypes.c_double)), ("field", ctypes.POINTER(ctypes.c_int)),
        ("nfac", ctypes.POINTER(ctypes.c_int)), ("fac", ctypes.POINTER(ctypes.c_int)),
    ]


_DP = ctypes.POINTER(ctypes.c_double)


class _GradJob(ctypes.Structure):
    """rdcu_grad_job"""
    _fields_ = [
        ("ctx", ctypes.c_void_p), ("y0", _DP), ("t0", ctypes.c_double),
        ("nsteps", ctypes.c_longlong), ("hs", _DP), ("nout", ctypes.c_int), ("t_out", _DP),
        ("target", _DP), ("loss", _DP), ("grad_coef", _DP), ("grad_y0", _DP),
        ("status", ctypes.c_int),
    ]
# rdcu_dydt_fn: (t, y_dev, f_dev, stream, user) -> status
DYDT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.c_void_p, ctypes.c_void_p)
STEP_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                           ctypes.c_int, ctypes.c_int, _DP, ctypes.c_void_p)

_loaded = {}


def load(path: str) -> ctypes.CDLL:
    """Loads a library exporting the C ABI and declares its signatures."""
    path = os.path.abspath(path)
    if path in _loaded:
        return _loaded[path]
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    lib.rdcu_create.argtypes = [ctypes.POINTER(_Problem), ctypes.c_double, ctypes.c_double,
                                ctypes.c_int, _DP, ctypes.c_int, ctypes.POINTER(vp)]
    lib.rdcu_destroy.argtypes = [vp]
    lib.rdcu_step.argtypes = [vp, _DP, _DP, ctypes.c_double, ctypes.c_double,
                              ctypes.POINTER(ctypes.c_int)]
    lib.rdcu_run.argtypes = [vp, _DP, _DP, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double, STEP_FN, vp]
    lib.rdcu_advance.argtypes = [vp, _DP, _DP, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong)]
    lib.rdcu_time_rhs.argtypes = [vp, ctypes.c_int, _DP, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                  ctypes.c_char_p, ctypes.c_int]
    lib.rdcu_eval.argtypes = [vp, ctypes.c_double, _DP]
    lib.rdcu_get_state.argtypes = [vp, _DP]
    lib.rdcu_set_state.argtypes = [vp, _DP]
    lib.rdcu_rhs.argtypes = [vp, _DP, _DP, _DP]
    lib.rdcu_rhs_noalloc.argtypes = [vp, _DP, _DP, _DP]
    lib.rdcu_rhs_device.argtypes = [vp, vp, vp]
    lib.rdcu_create_fn.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, _DP, DYDT_FN, vp, ctypes.c_int,
                                   ctypes.POINTER(vp)]
    lib.rdcu_frame.argtypes = [vp, ctypes.c_double, _DP, _DP, _DP]
    lib.rdcu_host_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(vp)]
    lib.rdcu_host_free.argtypes = [vp]
    lib.rdcu_get_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
    lib.rdcu_batch_create.argtypes = [ctypes.POINTER(_Problem), ctypes.c_int,
                                      ctypes.POINTER(_DP), ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    lib.rdcu_batch_destroy.argtypes = [vp]
    lib.rdcu_batch_step.argtypes = [vp, _DP, _DP, ctypes.c_double, ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_int)]
    lib.rdcu_batch_run.argtypes = [vp, _DP, _DP, ctypes.c_double, ctypes.c_double, _DP,
                                   ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong)]
    lib.rdcu_batch_member.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
    lib.rdcu_set_coef.argtypes = [vp, _DP]
    lib.rdcu_schedule.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_longlong, _DP,
                                  ctypes.POINTER(ctypes.c_longlong)]
    lib.rdcu_loss_grad.argtypes = [vp, _DP, ctypes.c_double, ctypes.c_longlong, _DP, ctypes.c_int,
                                   _DP, _DP, _DP, _DP, _DP]
    lib.rdcu_loss_grad_device.argtypes = [vp, vp, ctypes.c_double, ctypes.c_longlong, _DP,
                                          ctypes.c_int, _DP, vp, _DP, _DP, vp]
    lib.rdcu_loss_grad_batch.argtypes = [ctypes.POINTER(_GradJob), ctypes.c_int, ctypes.c_int]
    lib.rdcu_last_error.restype = ctypes.c_char_p
    lib.rdcu_backend.restype = ctypes.c_char_p
    for s in SYMBOLS:
        if s not in ("rdcu_last_error", "rdcu_backend"):
            getattr(lib, s).restype = ctypes.c_int
    _loaded[path] = lib
    return lib


def hip_lib() -> ctypes.CDLL:
    """The MI355X library.  Raises if it is not built: there is no fallback."""
    return load(HIP_LIB)


@dataclass
class Term:
    """One dictionary term ``coef * prod_k Y[idx_k]**pow_k`` added to ``field``.

    Same meaning as one ``-c C,field,idx,pow,...`` argument of rdcu
    (rdcu.cu:654-661, help text rdcu.cu:740-745)."""
    coef: float
    field: int
    factors: List[Tuple[int, int]] = field(default_factory=list)

    @staticmethod
    def parse(text: str) -> "Term":
        tok = [t for t in text.split(",") if t != ""]
        ints = [int(float(t)) for t in tok[1:]]
        f = ints[0] if ints else 0
        rest = ints[1:]
        facs = [(rest[2 * j], rest[2 * j + 1] if 2 * j + 1 < len(rest) else 0)
                for j in range((len(rest) + 1) // 2)]
        return Term(float(tok[0]), f, facs)


@dataclass
class Problem:
    """Grid, domain and dictionary (the RHS part of rdcu.cu's ``parameters``)."""
    n: int
    Ns: Sequence[int]
    Ls: Sequence[float]
    terms: List[Term]

    @property
    def ndim(self) -> int:
        return len(self.Ns)

    @property
    def N(self) -> int:
        return int(np.prod(self.Ns))

    @property
    def nblk(self) -> int:
        d = self.ndim
        return 1 + d + d * d

    def index(self, block: int, field: int) -> int:
        """Factor index of (block, field): block 0 = u, 1+a = d/dx_a,
        1+ndim+a*ndim+b = d/dx_b d/dx_a (rdcu.cu:769-785)."""
        return block * self.n + field

    def _c(self):
        p = _Problem()
        p.n = self.n
        p.ndim = self.ndim
        for a in range(3):
            p.Ns[a] = int(self.Ns[a]) if a < self.ndim else 1
            p.Ls[a] = float(self.Ls[a]) if a < self.ndim else 1.0
        p.nterms = len(self.terms)
        coef = np.array([t.coef for t in self.terms] or [0.0], dtype=np.float64)
        fld = np.array([t.field for t in self.terms] or [0], dtype=np.int32)
        nfac = np.array([len(t.factors) for t in self.terms] or [0], dtype=np.int32)
        fac = np.array([v for t in self.terms for pair in t.factors for v in pair] or [0],
                       dtype=np.int32)
        keep = (coef, fld, nfac, fac)
        p.coef = coef.ctypes.data_as(_DP)
        p.field = fld.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        p.nfac = nfac.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        p.fac = fac.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        return p, keep


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_DP)


class Integrator:
    """Handle-based DP45 integrator over one library (dp45.h:7-11)."""

    def __init__(self, lib: ctypes.CDLL, prob: Problem, y0: np.ndarray, atol=1e-8, rtol=1e-8,
                 fixed=False, device=0):
        self.lib = lib
        self.prob = prob
        self.nN = prob.n * prob.N
        y0 = np.ascontiguousarray(y0, dtype=np.float64).reshape(-1)
        if y0.size != self.nN:
            raise ValueError(f"y0 has {y0.size} values, expected {self.nN}")
        p, keep = prob._c()
        h = ctypes.c_void_p()
        self._check(lib.rdcu_create(ctypes.byref(p), atol, rtol, int(bool(fixed)), _ptr(y0),
                                    device, ctypes.byref(h)))
        del keep
        self.h = h

    def _check(self, code):
        if code != OK:
            raise RdcuError(code, self.lib.rdcu_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.rdcu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, t: float, h: float, hmin=0.0, hmax=np.inf):
        """One attempt (dp45_step, dp45.cu:87-147): returns (t, h, accepted)."""
        tt, hh, acc = ctypes.c_double(t), ctypes.c_double(h), ctypes.c_int()
        self._check(self.lib.rdcu_step(self.h, ctypes.byref(tt), ctypes.byref(hh), hmin, hmax,
                                       ctypes.byref(acc)))
        return tt.value, hh.value, bool(acc.value)

    def run(self, t: float, h: float, t1: float, hmin=0.0, hmax=np.inf, stop=np.inf,
            callback: Optional[Callable] = None):
        """dp45_run (dp45.cu:154-169).  callback(t, h, success, current) may
        return a new stop time (or None to keep it).  Returns (t, h)."""
        err = []

        def cb(ctx, t, h, ok, cur, stopp, user):
            try:
                if callback is not None:
                    r = callback(t, h, bool(ok), bool(cur))
                    if r is not None:
                        stopp[0] = float(r)
                return 0
            except Exception as e:  # surface Python errors after the run
                err.append(e)
                return 1

        fn = STEP_FN(cb)
        tt, hh = ctypes.c_double(t), ctypes.c_double(h)
        code = self.lib.rdcu_run(self.h, ctypes.byref(tt), ctypes.byref(hh), hmin, hmax, t1,
                                 stop, fn, None)
        if err:
            raise err[0]
        self._check(code)
        return tt.value, hh.value

    def advance(self, t: float, h: float, t1=np.inf, max_attempts=1, hmin=0.0, hmax=np.inf):
        """Attempts without per-attempt callback; returns (t, h, attempts run)."""
        tt, hh, done = ctypes.c_double(t), ctypes.c_double(h), ctypes.c_longlong()
        self._check(self.lib.rdcu_advance(self.h, ctypes.byref(tt), ctypes.byref(hh), hmin, hmax,
                                          t1, int(max_attempts), ctypes.byref(done)))
        return tt.value, hh.value, done.value

    def time_rhs(self, reps=100):
        """{kernel name: average ms} from HIP events ("rhs" = whole evaluation)."""
        ms = (ctypes.c_double * 16)()
        cnt = ctypes.c_int()
        names = ctypes.create_string_buffer(512)
        self._check(self.lib.rdcu_time_rhs(self.h, int(reps), ms, 16, ctypes.byref(cnt), names,
                                           512))
        return dict(zip(names.value.decode().split(","), list(ms[:cnt.value])))

    def eval(self, t_eval: float) -> np.ndarray:
        out = np.empty(self.nN)
        self._check(self.lib.rdcu_eval(self.h, t_eval, _ptr(out)))
        return out

    def state(self) -> np.ndarray:
        out = np.empty(self.nN)
        self._check(self.lib.rdcu_get_state(self.h, _ptr(out)))
        return out

    def set_state(self, y: np.ndarray):
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1)
        self._check(self.lib.rdcu_set_state(self.h, _ptr(y)))

    def rhs(self, y: np.ndarray, factors=False, noalloc=False):
        """f = dydt(y); with factors=True also every Y block as complex
        [nblk, n, N] in the layout of FILEBASEY.dat.  noalloc=True calls
        rdcu_rhs_noalloc (the same values; no device allocation)."""
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1)
        f = np.empty(self.nN)
        Y = np.empty(2 * self.nN * self.prob.nblk) if factors else None
        fn = self.lib.rdcu_rhs_noalloc if noalloc else self.lib.rdcu_rhs
        self._check(fn(self.h, _ptr(y), _ptr(f), _ptr(Y) if factors else None))
        if factors:
            return f, Y.view(np.complex128).reshape(self.prob.nblk, self.prob.n, self.prob.N)
        return f

    def frame(self, t_eval: float, dense: int = 3):
        """One output frame at t_eval, as the CLI writes it at -D<dense>:
        (y, f or None, Y or None), Y complex [nblk, n, N]."""
        y = np.empty(self.nN)
        f = np.empty(self.nN) if dense >= 2 else None
        Y = np.empty(2 * self.nN * self.prob.nblk) if dense >= 3 else None
        self._check(self.lib.rdcu_frame(self.h, t_eval, _ptr(y),
                                        _ptr(f) if f is not None else None,
                                        _ptr(Y) if Y is not None else None))
        if Y is not None:
            Y = Y.view(np.complex128).reshape(self.prob.nblk, self.prob.n, self.prob.N)
        return y, f, Y

    def set_coef(self, coef):
        """Replaces the dictionary coefficients (term order)."""
        c = np.ascontiguousarray(coef, dtype=np.float64).reshape(-1)
        if c.size != len(self.prob.terms):
            raise ValueError(f"{c.size} coefficients for {len(self.prob.terms)} terms")
        self._check(self.lib.rdcu_set_coef(self.h, _ptr(c)))

    def schedule(self, t0: float, h0: float, t1: float, hmin=0.0, hmax=np.inf,
                 max_steps=1 << 20) -> np.ndarray:
        """Accepted step sizes of an adaptive run from the current state."""
        hs = np.empty(int(max_steps))
        ns = ctypes.c_longlong()
        self._check(self.lib.rdcu_schedule(self.h, t0, h0, hmin, hmax, t1, int(max_steps),
                                           _ptr(hs), ctypes.byref(ns)))
        return hs[:ns.value].copy()

    def loss_grad(self, y0, t0: float, hs, t_out, target, grad=True):
        """Loss over outputs on the frozen schedule hs, and (grad=True) its
        gradients: returns (loss, dloss/dcoef, dloss/dy0)."""
        y0 = np.ascontiguousarray(y0, dtype=np.float64).reshape(-1)
        hs = np.ascontiguousarray(hs, dtype=np.float64).reshape(-1)
        to = np.ascontiguousarray(t_out, dtype=np.float64).reshape(-1)
        tg = np.ascontiguousarray(target, dtype=np.float64).reshape(-1)
        if y0.size != self.nN or tg.size != to.size * self.nN:
            raise ValueError("y0 needs n*N values and target nout*n*N")
        loss = ctypes.c_double()
        gc = np.empty(max(1, len(self.prob.terms))) if grad else None
        gy = np.empty(self.nN) if grad else None
        self._check(self.lib.rdcu_loss_grad(
            self.h, _ptr(y0), t0, hs.size, _ptr(hs), to.size, _ptr(to), _ptr(tg),
            ctypes.byref(loss), _ptr(gc) if grad else None, _ptr(gy) if grad else None))
        if not grad:
            return loss.value
        return loss.value, gc[:len(self.prob.terms)], gy

    def stats(self):
        out = (ctypes.c_longlong * 4)()
        self._check(self.lib.rdcu_get_stats(self.h, out))
        return dict(attempts=out[0], accepted=out[1], rhs=out[2], syncs=out[3])


class _Member(Integrator):
    """Instance i of a Batch: state, eval, rhs and stats of an ordinary
    handle; owned by the batch (closing it does nothing)."""

    def __init__(self, batch: "Batch", i: int):
        self.lib = batch.lib
        self.prob = batch.probs[i]
        self.nN = self.prob.n * self.prob.N
        h = ctypes.c_void_p()
        self._check(self.lib.rdcu_batch_member(batch.h, i, ctypes.byref(h)))
        self.h = h
        self._batch = batch

    def close(self):
        self.h = None


class CallbackIntegrator(Integrator):
    """DP45 on a right-hand side given as a Python callable (rdcu_create_fn):
    the dydt callback of dp45_init (dp45.h:8).  ``fn(t, y_ptr, f_ptr,
    stream)`` receives raw pointers -- device memory on the HIP library (the
    integrator's stream is ``stream``; the callable must leave f complete or
    ordered before later work on it), host memory on the CPU oracle -- and
    returns 0 (nonzero aborts).  The state is n fields of N values.  step,
    run, advance, eval, state, set_state and stats work as for a dictionary
    problem."""

    def __init__(self, lib: ctypes.CDLL, n: int, N: int, y0: np.ndarray, fn: Callable,
                 atol=1e-8, rtol=1e-8, fixed=False, device=0):
        self.lib = lib
        self.prob = None
        self.nN = n * N
        y0 = np.ascontiguousarray(y0, dtype=np.float64).reshape(-1)
        if y0.size != self.nN:
            raise ValueError(f"y0 has {y0.size} values, expected {self.nN}")
        self._err = []

        def cb(t, y, f, stream, user):
            try:
                return int(fn(t, y, f, stream) or 0)
            except Exception as e:  # surfaced after the call
                self._err.append(e)
                return 1

        self._fn = DYDT_FN(cb)  # kept alive with the handle
        h = ctypes.c_void_p()
        self._check(lib.rdcu_create_fn(n, N, atol, rtol, int(bool(fixed)), _ptr(y0), self._fn,
                                       None, device, ctypes.byref(h)))
        self.h = h

    def _check(self, code):
        if getattr(self, "_err", None):
            e = self._err[0]
            self._err.clear()
            raise e
        super()._check(code)


class Batch:
    """Many independent problems integrated side by side (rdcu_batch_*): the
    in-process replacement of a sweep.sh job array (sweep.sh:21-41)."""

    def __init__(self, lib: ctypes.CDLL, probs: Sequence[Problem], y0s: Sequence[np.ndarray],
                 atol=1e-8, rtol=1e-8, fixed=False, device=0):
        self.lib = lib
        self.probs = list(probs)
        if len(y0s) != len(self.probs):
            raise ValueError("one initial state per problem")
        cs, keeps, ys = [], [], []
        for p, y in zip(self.probs, y0s):
            y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1)
            if y.size != p.n * p.N:
                raise ValueError(f"y0 has {y.size} values, expected {p.n * p.N}")
            c, keep = p._c()
            cs.append(c)
            keeps.append(keep)
            ys.append(y)
        arr = (_Problem * len(cs))(*cs)
        yp = (_DP * len(ys))(*[_ptr(y) for y in ys])
        h = ctypes.c_void_p()
        self._check(lib.rdcu_batch_create(arr, len(cs), yp, atol, rtol, int(bool(fixed)), device,
                                          ctypes.byref(h)))
        del keeps, ys
        self.h = h
        self.count = len(cs)

    def _check(self, code):
        if code != OK:
            raise RdcuError(code, self.lib.rdcu_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.rdcu_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __len__(self):
        return self.count

    def member(self, i: int) -> Integrator:
        return _Member(self, i)

    def time_rhs(self, reps=20, max_classes=64):
        """One plain RHS of every instance, timed (rdcu_batch_time): returns
        (ms of all classes back to back, [per class dict: ndim, grid,
        instances, ms and algorithmic bytes of the x-forward, strided and
        x-inverse launches])."""
        ms = np.zeros(1 + 3 * max_classes)
        by = np.zeros(3 * max_classes)
        sh = (ctypes.c_int * (4 * max_classes))()
        nc = ctypes.c_int(0)
        self._check(self.lib.rdcu_batch_time(self.h, reps, _ptr(ms), _ptr(by), sh, max_classes,
                                             ctypes.byref(nc)))
        out = []
        names = ("row_fwd", "col_mult", "row_inv")
        for k in range(min(nc.value, max_classes)):
            out.append({"ndim": sh[4 * k], "grid": [sh[4 * k + 1], sh[4 * k + 2]],
                        "instances": sh[4 * k + 3],
                        "ms": {nm: float(ms[1 + 3 * k + j]) for j, nm in enumerate(names)},
                        "bytes": {nm: float(by[3 * k + j]) for j, nm in enumerate(names)}})
        return float(ms[0]), out

    def step(self, t, h, hmin=0.0, hmax=np.inf):
        """One attempt of every instance: returns (t, h, accepted) arrays."""
        tt = np.array(t, dtype=np.float64, copy=True).reshape(-1)
        hh = np.array(h, dtype=np.float64, copy=True).reshape(-1)
        acc = (ctypes.c_int * self.count)()
        self._check(self.lib.rdcu_batch_step(self.h, _ptr(tt), _ptr(hh), hmin, hmax, acc))
        return tt, hh, np.array(acc[:], dtype=bool)

    def run(self, t, h, t1, max_attempts=1 << 40, hmin=0.0, hmax=np.inf):
        """Every instance while t < t1 and h > hmin: returns (t, h, attempts)."""
        tt = np.array(t, dtype=np.float64, copy=True).reshape(-1)
        hh = np.array(h, dtype=np.float64, copy=True).reshape(-1)
        t1 = np.ascontiguousarray(np.broadcast_to(np.asarray(t1, dtype=np.float64),
                                                  (self.count,)))
        na = (ctypes.c_longlong * self.count)()
        self._check(self.lib.rdcu_batch_run(self.h, _ptr(tt), _ptr(hh), hmin, hmax, _ptr(t1),
                                            int(max_attempts), na))
        return tt, hh, np.array(na[:], dtype=np.int64)


def loss_grad_batch(integrators: Sequence[Integrator], y0s, t0s, hss, t_outs, targets,
                    grad=True, concurrency=4):
    """Integrator.loss_grad for many instances at once (rdcu_loss_grad_batch:
    up to `concurrency` instances on the GPU together, each on its own
    handle).  Returns [(loss, dloss/dcoef, dloss/dy0)] (or [loss] without
    grad), each exactly what loss_grad gives for that instance alone."""
    its = list(integrators)
    if len({id(it) for it in its}) != len(its):
        raise ValueError("every instance needs its own Integrator")
    lib = its[0].lib if its else None
    keep, out = [], []
    jobs = (_GradJob * max(1, len(its)))()
    for i, it in enumerate(its):
        if it.lib is not lib:
            raise ValueError("all instances must use one library")
        y0 = np.ascontiguousarray(y0s[i], dtype=np.float64).reshape(-1)
        hs = np.ascontiguousarray(hss[i], dtype=np.float64).reshape(-1)
        to = np.ascontiguousarray(t_outs[i], dtype=np.float64).reshape(-1)
        tg = np.ascontiguousarray(targets[i], dtype=np.float64).reshape(-1)
        if y0.size != it.nN or tg.size != to.size * it.nN:
            raise ValueError("y0 needs n*N values and target nout*n*N")
        loss = np.zeros(1)
        gc = np.empty(max(1, len(it.prob.terms))) if grad else None
        gy = np.empty(it.nN) if grad else None
        keep += [y0, hs, to, tg, loss, gc, gy]
        jobs[i] = _GradJob(it.h, _ptr(y0), float(t0s[i]), hs.size, _ptr(hs), to.size, _ptr(to),
                           _ptr(tg), _ptr(loss), _ptr(gc) if grad else None,
                           _ptr(gy) if grad else None, 0)
        out.append((loss, gc, gy, len(it.prob.terms)))
    if not its:
        return []
    code = lib.rdcu_loss_grad_batch(jobs, len(its), int(concurrency))
    if code:
        raise RdcuError(code, lib.rdcu_last_error().decode())
    if not grad:
        return [o[0][0] for o in out]
    return [(o[0][0], o[1][:o[3]], o[2]) for o in out]


def glibc_rand_ic(n: int, N: int, seed: int = 1, A: float = 1e-3, A0: float = 0.0) -> np.ndarray:
    """Random initial condition of rdcu.cu:898, 991-1001:
    y = A0 + A/2*(2.0/RAND_MAX*rand()-1) after srand(seed), field-major."""
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand(ctypes.c_uint(seed))
    r = np.fromiter((libc.rand() for _ in range(n * N)), dtype=np.float64, count=n * N)
    return A0 + A / 2 * (2.0 / 2147483647 * r - 1)
