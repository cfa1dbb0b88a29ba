// rd_kernels.h -- gfx950 kernels of the DP45 pseudospectral integrator.
//
// One RHS evaluation du/dt = sum_i C_i prod_k Y_k^eta_ik (rdcu.cu:117-184) is
// three kinds of axis pass over the grid, each one HBM read and one write:
//   k_row_r2c   x-lines: (stage sum ->) real-to-half-complex FFT    -> S
//   k_col       strided lines (y in 2D; y and z in 3D): forward FFT,
//               and on axis 0 the multiplier of every needed factor
//               followed by its inverse FFT                          -> F
//   k_row_c2r   x-lines: half-complex-to-real FFT of every factor, then
//               the dictionary, f, and on the last stage the scaled
//               error with a per-line pairwise sum                   -> k_s
// The DP45 stage sums (dp45.cu:34-69) are fused into k_row_r2c, the error
// kernel (dp45.cu:79-84) into k_row_c2r, and the step control
// (dp45.cu:87-147) runs on the device in k_control, so a whole attempt is a
// fixed kernel sequence that is captured once in a hipGraph.
//
// Layout in HBM: state vectors field-major, row-major grid, x fastest
// (the reference's layout, rdcu.cu:82-86).  Half spectra S and F are
// [block][rows][Nx/2+1] complex.
#pragma once
#include "rd_fft.h"

namespace rd {

constexpr int kNMax = 8;       // fields per problem handled by the kernels
constexpr int kLogCap = 128;   // attempts logged per chunk
constexpr int kRedCap = 4096;  // lines reduced in LDS by one block
// Factor codes at and above kImagBlk select the imaginary part of derivative
// block (code - kImagBlk) in Y (apply_mult); real blocks are below 1+3+9.
constexpr int kImagBlk = 64;

// Device-resident integrator state (replaces dp45.cu's static globals and the
// host-side step control).
struct DevState {
  double t, h, t_last, t_now;
  double hmin, hmax, t1, t_stop;
  double atol, rtol;
  double last_norm;
  int iy, iyt, iyl;  // roles of the three state buffers
  int k1s, k7s;      // physical buffers holding stage slots 1 and 7
  int halted;        // stop executing attempts in this chunk
  int fixed;         // -F
  int nlog;          // attempts logged in this chunk
  int eval_valid;    // last attempt accepted
  int nolog;         // batched runs: no attempt log, no halt when it fills
  long long attempts, accepted;
  // frozen schedule (fixed steps): step i has size hsched[i]; null otherwise
  const double *hsched;
  long long nsched, isched;
  double log_t[kLogCap], log_h[kLogCap], log_err[kLogCap];
  int log_acc[kLogCap];
};

struct Bufs {
  double *y[3];
  double *k[7];
  cx *S;
  cx *F;
  double *part;   // per-line sums of squares
  double *part2;  // block sums when there are more than kRedCap lines
};

struct Grid {
  int n, ndim;
  int Ns[3];
  int Nx, Nxh, L;
  int Nxp;  // row pitch of half spectra (complex): Nxh rounded up to 8 (128 B)
  int nt_store;  // streaming stores in the x-forward and strided passes
  long N, rows, Nh;  // Nh = rows * Nxp
  double invN;
  double invNxy;  // 3-D: 1/(N_y N_x), the scale of the factors without k_z (zfree_blk)
  const double *kt[3];  // wavenumbers per axis (device)
};

// Dictionary: term t adds tcoef[t] * prod_j vals[fslot[j]]^fpow[j] to field
// tfield[t] (j in [tstart[t], tstart[t+1])).  Value slot s < n is field s
// itself; s >= n is transformed factor s - n of the factor list.
struct Dict {
  int nterms;
  const int *tfield;
  const double *tcoef;
  const int *tstart;
  const int *fslot;
  const int *fpow;
};

// Dictionary program of the register path: int32 words, staged into LDS.
//   words [0, n]: start of field f's term list (word n: end of the program)
//   term record : nfac, coef low word, coef high word, then nfac (slot, pow)
// A field's terms keep their input order.  Same meaning as Dict.
struct DictP {
  const int *prog;  // device memory
  int nwords;
};

// Transformed factors, grouped by field: factors of field f are
// [fstart[f], fstart[f+1]); fblk[m] is the derivative block of factor m.
// fslot (null: m itself) is factor m's slot in F, for a sublist of a
// problem's factors.
struct FacList {
  int nf;
  const int *fblk;
  const int *ffield;
  const int *fstart;
  const int *fslot;
};
__host__ __device__ inline int fac_slot(const FacList &fl, int m) {
  return fl.fslot ? fl.fslot[m] : m;
}

// 3-D: a factor whose multiplier has no k_z (first and second derivatives in
// y and x, and their imaginary-part codes).  Its z transforms cancel, so it is
// multiplied in the y pass, right after the forward y transform (scale
// invNxy), and transformed back along y there; the z pass and the inverse y
// pass carry only the other factors.
__host__ __device__ inline bool zfree_blk(int blk, int nd) {
  if (nd != 3) return false;
  if (blk >= kImagBlk) blk -= kImagBlk;  // imaginary-part codes
  if (blk <= nd) return blk - 1 != 0;
  const int q = blk - 1 - nd;
  return q / nd != 0 && q % nd != 0;
}

enum SrcKind { SRC_PTR = 0, SRC_Y = 1, SRC_STAGE = 2, SRC_YT = 3 };
enum DstKind { DST_PTR = 0, DST_K1 = 1, DST_KS = 2, DST_K7 = 3 };

__device__ __forceinline__ bool attempt_active(const DevState *st) {
  return !st->halted && st->t < st->t1 && st->h > st->hmin;
}

// Hermitian part of the reference's multiplier prod (i k_a) (rdcu.cu:78-99),
// times the scale sc (1/N, or invNxy for the 3-D factors without k_z), applied
// to X at wavenumber indices n0..n2 (last axis in
// [0, Nx/2]).  Nyquist rules: d/dx_a vanishes on the N_a/2 plane; mixed
// d2/dx_a dx_b vanishes where exactly one axis is at Nyquist.
__device__ __forceinline__ cx apply_mult(const Grid &g, int blk, const int *nk, cx X,
                                         double sc) {
  const int nd = g.ndim;
  if (blk <= nd) {
    const int a = blk - 1;
    if (nk[a] == g.Ns[a] / 2) return {0.0, 0.0};
    const double val = g.kt[a][nk[a]] * sc;
    return {-(X.im * val), X.re * val};
  }
  const int q = blk - 1 - nd;
  const int a = q / nd, b = q - a * nd;
  double val;
  if (a == b) {
    const double k = g.kt[a][nk[a]];
    val = (k * k) * sc;
  } else {
    const bool na = nk[a] == g.Ns[a] / 2, nb = nk[b] == g.Ns[b] / 2;
    if (na != nb) return {0.0, 0.0};
    val = (g.kt[a][nk[a]] * g.kt[b][nk[b]]) * sc;
  }
  return {-(X.re * val), -(X.im * val)};
}

// apply_mult, and for codes blk >= kImagBlk the imaginary part of Y (the
// factor output of the LDS kernels; the integration path never sees these
// codes and calls apply_mult, which keeps its registers).
__device__ __forceinline__ cx apply_mult_y(const Grid &g, int blk, const int *nk, cx X,
                                           double sc) {
  const int nd = g.ndim;
  if (blk >= kImagBlk) {
    // -i times the anti-Hermitian part of the reference's multiplier m
    // (rdcu.cu:89-95): (m(k) - conj(m(-k)))/2 is m itself where m(-k) is
    // not conj(m(k)), i.e. i*k_a on the Nyquist plane of axis a, and
    // -k_a*k_b where exactly one of a, b is at Nyquist; zero elsewhere.  Its
    // transform is the imaginary part the reference leaves in Y.
    blk -= kImagBlk;
    if (blk <= nd) {
      const int a = blk - 1;
      if (nk[a] != g.Ns[a] / 2) return {0.0, 0.0};
      const double val = g.kt[a][nk[a]] * sc;
      return {X.re * val, X.im * val};
    }
    const int q = blk - 1 - nd;
    const int a = q / nd, b = q - a * nd;
    if (a == b) return {0.0, 0.0};
    const bool na = nk[a] == g.Ns[a] / 2, nb = nk[b] == g.Ns[b] / 2;
    if (na == nb) return {0.0, 0.0};
    const double val = (g.kt[a][nk[a]] * g.kt[b][nk[b]]) * sc;
    return {-(X.im * val), X.re * val};
  }
  return apply_mult(g, blk, nk, X, sc);
}

__device__ __forceinline__ double ipow(double v, int e) {
  if (e == 0) return 1.0;
  const int a = e < 0 ? -e : e;
  double r = v;
  for (int j = 1; j < a; ++j) r = r * v;
  return e < 0 ? 1.0 / r : r;
}

// ------------------------------------------------------------ pow(x, -0.2)
// Correctly rounded x^y, y the double nearest -0.2 (the exponent of
// dp45.cu:120).  x^(-1/5) by Newton on w^-5 = m refined in double-double,
// times 2^(-r/5) and x^(-eps) with eps = 0.2 - (-y).
__device__ __forceinline__ double pow_m02(double x) {
  const double c5hi[5] = {1.0, 0x1.bdb8cdadbe120p-1, 0x1.8406003b2ae5cp-1, 0x1.51cb453b9536cp-1,
                          0x1.2611186bae675p-1};
  const double c5lo[5] = {0.0, 0x1.0a7b0e98593fdp-56, 0x1.ae64d1cc83ab8p-55, 0x1.15a4ea11db2e6p-55,
                          -0x1.a56b786f7e27cp-55};
  if (__builtin_isnan(x)) return x;
  if (x == 0.0) return __builtin_huge_val();
  if (x < 0.0) return __builtin_nan("");
  if (__builtin_isinf(x)) return 0.0;
  int e;
  double m = frexp(x, &e);
  m *= 2.0;
  e -= 1;
  int r = e % 5;
  if (r < 0) r += 5;
  const int q = (e - r) / 5;
  double w = 1.0 - 0.12944943670387588 * (m - 1.0);
  for (int it = 0; it < 5; ++it) {
    const double w2 = w * w;
    const double w5 = (w2 * w2) * w;
    w = __builtin_fma(w * 0.2, 1.0 - m * w5, w);
  }
  const double a2h = w * w, a2l = __builtin_fma(w, w, -a2h);
  const double w4h = a2h * a2h, w4l = __builtin_fma(a2h, a2h, -w4h) + 2.0 * (a2h * a2l);
  const double w5h = w4h * w, w5l = __builtin_fma(w4h, w, -w5h) + w4l * w;
  const double ph = m * w5h, pl = __builtin_fma(m, w5h, -ph) + m * w5l;
  const double rho = (1.0 - ph) - pl;
  const double dlt = (w * rho) * 0.2;
  const double eps = 0x1.999999999999ap-57;
  const double lx = log(x);
  const double el = eps * lx;
  const double c = -el + 0.5 * (el * el);
  double chi = 1.0, clo = 0.0;
  switch (r) {
    case 1: chi = c5hi[1]; clo = c5lo[1]; break;
    case 2: chi = c5hi[2]; clo = c5lo[2]; break;
    case 3: chi = c5hi[3]; clo = c5lo[3]; break;
    case 4: chi = c5hi[4]; clo = c5lo[4]; break;
    default: break;
  }
  const double vh = chi * w;
  double vl = __builtin_fma(chi, w, -vh) + (chi * dlt + clo * w);
  vl = vl + vh * c;
  const double res = vh + vl;
  return ldexp(res, -q);
}

// --------------------------------------------------------------- kernels

__global__ void k_begin(DevState *st, double t_stop, int set_th, double t, double h, double hmin,
                        double hmax, double t1) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (set_th) {
      st->t = t;
      st->h = h;
    }
    st->hmin = hmin;
    st->hmax = hmax;
    st->t1 = t1;
    st->t_stop = t_stop;
    st->halted = 0;
    st->nlog = 0;
    st->nolog = 0;
  }
}

// Stage coefficients of one launch of k_row_r2c: ytemp = y + h*sum_j a[j]*k_{j+1}.
struct StageCoef {
  int ns;  // number of k slots used (s-1), 0 for a plain RHS input
  double a[6];
};

// x-pass forward: rows of Nx reals (optionally formed as a DP45 stage sum)
// -> Nx/2+1 complex per row.  Block = rpb rows; dynamic LDS 2*rpb*L complex.
__global__ void k_row_r2c(Grid g, FftPlan px, const cx *__restrict__ twN, DevState *st,
                          int check_active, Bufs b, int src_kind, const double *src_ptr,
                          StageCoef sc, int rpb) {
  extern __shared__ cx smem[];
  if (check_active && !attempt_active(st)) return;
  const int L = g.L, Nx = g.Nx;
  const long nrows = (long)g.n * g.rows;
  const long row0 = (long)blockIdx.x * rpb;
  const int nr = (int)((nrows - row0) < rpb ? (nrows - row0) : rpb);
  const int tid = threadIdx.x, nt = blockDim.x;
  cx *A = smem, *B = smem + rpb * L;

  if (src_kind == SRC_STAGE) {
    const double h = st->h;
    const double *y = b.y[st->iy];
    double *yt = b.y[st->iyt];
    const double *ks[6];
    ks[0] = b.k[st->k1s];
    for (int j = 1; j < 6; ++j) ks[j] = b.k[j];
    for (int w = tid; w < nr * L; w += nt) {
      const int rr = w / L, j = w - rr * L;
      const long off = (row0 + rr) * Nx + 2 * j;
      const double2 yv = *reinterpret_cast<const double2 *>(y + off);
      const double2 k1 = *reinterpret_cast<const double2 *>(ks[0] + off);
      double ar = sc.a[0] * k1.x, ai = sc.a[0] * k1.y;
      for (int s = 1; s < sc.ns; ++s) {
        if (sc.a[s] == 0.0) continue;
        const double2 kv = *reinterpret_cast<const double2 *>(ks[s] + off);
        ar = __builtin_fma(sc.a[s], kv.x, ar);
        ai = __builtin_fma(sc.a[s], kv.y, ai);
      }
      double2 v;
      v.x = __builtin_fma(h, ar, yv.x);
      v.y = __builtin_fma(h, ai, yv.y);
      *reinterpret_cast<double2 *>(yt + off) = v;
      A[rr * L + j] = cx{v.x, v.y};
    }
  } else {
    const double *src = src_kind == SRC_Y ? b.y[st->iy] : (src_kind == SRC_YT ? b.y[st->iyt] : src_ptr);
    for (int w = tid; w < nr * L; w += nt) {
      const int rr = w / L, j = w - rr * L;
      const long off = (row0 + rr) * Nx + 2 * j;
      const double2 v = *reinterpret_cast<const double2 *>(src + off);
      A[rr * L + j] = cx{v.x, v.y};
    }
  }
  __syncthreads();
  const cx *Z = lds_fft<-1, false>(A, B, px, nr, L, 1, tid, nt);
  for (int w = tid; w < nr * (L + 1); w += nt) {
    const int rr = w / (L + 1), k = w - rr * (L + 1);
    const cx Av = Z[rr * L + (k == L ? 0 : k)];
    const cx Bv = cconj(Z[rr * L + (k == 0 ? 0 : L - k)]);
    const cx E = cadd(Av, Bv), D = csub(Av, Bv);
    const cx t = cmul(D, twN[k]);
    b.S[(row0 + rr) * g.Nxp + k] = cx{0.5 * (E.re + t.im), 0.5 * (E.im - t.re)};
  }
}

// Strided-line pass over tiles of `tc` adjacent half-spectrum columns.
//   mode 0: forward FFT in place on S      (3D axis 1)
//   mode 1: forward FFT of S, then for every factor of the line's field the
//           multiplier and inverse FFT -> F (axis 0)
//   mode 2: inverse FFT in place on F      (3D axis 1)
//   mode 3: mode 0, then for every factor of the line's field (the factors
//           without k_z) the multiplier (scale invNxy) and inverse FFT -> F
//           (3D axis 1)
//   mode 4: the factor part of mode 3 on S already transformed (3D axis 1)
// F positions are the factors' slots (fac_slot).
// Tile id -> (group g, column tile ct); group g -> (g1, g2) = (g / gdiv, g % gdiv);
// line base = g1*gs1 + g2*gs2 + ct*tc, element stride es.  For mode 1, g1 is the
// field and the factor outputs sit at m*Nh + g2*gs2 + ct*tc.
struct ColGeom {
  int L;       // line length
  long es;     // element stride (complex units)
  int gdiv;
  long gs1, gs2;
  int ntc;     // column tiles per group
  int tc;      // columns per tile
  int axis;    // grid axis of the lines (multiplier index)
  int nparts;  // blocks per tile in mode 1 (factor split, register path)
  int nblk;    // logical blocks (register path; the grid is rounded up to 8)
  int bal;     // register path: tiles of even widths (1) or of tc from column 0 (0)
};

// Logical block of a launch whose grid is a multiple of the 8 XCDs: hardware
// hands block b to XCD b % 8, so logical blocks [x*nb/8, (x+1)*nb/8) -- which
// read neighbouring columns and the same spectra -- share XCD x's L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  return (b & 7) * (nb >> 3) + (b >> 3);
}
__device__ __forceinline__ long xcd_item(long j, long nj) {
  return (j & 7) * (nj >> 3) + (j >> 3);
}

__global__ void k_col(Grid g, FftPlan p, ColGeom cg, DevState *st, int check_active, Bufs b,
                      FacList fl, int mode) {
  extern __shared__ cx smem[];
  if (check_active && !attempt_active(st)) return;
  const int tile = blockIdx.x;
  const int grp = tile / cg.ntc, ct = tile - grp * cg.ntc;
  const int g1 = grp / cg.gdiv, g2 = grp - g1 * cg.gdiv;
  const bool fac = mode == 1 || mode == 3 || mode == 4;  // the multiply modes
  if ((mode == 1 || mode == 4) && fl.fstart[g1] == fl.fstart[g1 + 1]) return;  // no factors
  const int c0 = ct * cg.tc;
  const int nc = (g.Nxh - c0) < cg.tc ? (g.Nxh - c0) : cg.tc;
  const int L = cg.L;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int tot = nc * L;
  cx *A = smem, *B = smem + cg.tc * L, *C = smem + 2 * cg.tc * L;
  const long lbase = (long)g2 * cg.gs2 + c0;

  cx *src;
  if (mode == 2) src = b.F + (long)fac_slot(fl, g1) * cg.gs1 + lbase;
  else src = b.S + (long)g1 * cg.gs1 + lbase;
  for (int w = tid; w < tot; w += nt) {
    const int c = w % nc, e = w / nc;
    A[e * nc + c] = src[e * cg.es + c];
  }
  __syncthreads();
  if (mode == 0 || mode == 2) {
    const cx *R = mode == 0 ? lds_fft<-1, true>(A, B, p, nc, 1, nc, tid, nt)
                            : lds_fft<+1, true>(A, B, p, nc, 1, nc, tid, nt);
    for (int w = tid; w < tot; w += nt) {
      const int c = w % nc, e = w / nc;
      src[e * cg.es + c] = R[e * nc + c];
    }
    return;
  }
  (void)fac;
  cx *R = A;
  if (mode != 4) R = lds_fft<-1, true>(A, B, p, nc, 1, nc, tid, nt);
  if (mode == 3) {  // the forward spectra stay in S for the z pass
    for (int w = tid; w < tot; w += nt) {
      const int c = w % nc, e = w / nc;
      src[e * cg.es + c] = R[e * nc + c];
    }
  }
  cx *Cb = (R == A) ? B : A;
  const int f = g1;
  const int m0 = fl.fstart[f], m1 = fl.fstart[f + 1];
  // modes 3 and 4 run along y with the scale of the factors without k_z
  const double msc = mode != 1 ? g.invNxy : g.invN;
  for (int m = m0; m < m1; ++m) {
    const int blk = fl.fblk[m];
    for (int w = tid; w < tot; w += nt) {
      const int c = w % nc, e = w / nc;
      int nk[3];
      if (mode != 1) {
        nk[0] = 0;
        nk[1] = e;
        nk[2] = c0 + c;
      } else {
        nk[0] = e;
        if (g.ndim == 2) {
          nk[1] = c0 + c;
        } else {
          nk[1] = g2;
          nk[2] = c0 + c;
        }
      }
      Cb[e * nc + c] = apply_mult_y(g, blk, nk, R[e * nc + c], msc);
    }
    __syncthreads();
    const cx *Q = lds_fft<+1, true>(Cb, C, p, nc, 1, nc, tid, nt);
    cx *dst = b.F + (long)fac_slot(fl, m) * g.Nh + lbase;
    for (int w = tid; w < tot; w += nt) {
      const int c = w % nc, e = w / nc;
      dst[e * cg.es + c] = Q[e * nc + c];
    }
    __syncthreads();
  }
}

// x-pass inverse for one row position r (all fields): half-complex -> real
// for every factor in groups of `grp` lines, then
//   mode 0 (dictionary): f = sum of terms, stored to the stage output; with
//     err != 0 also the DP45 scaled error and its per-line pairwise sum;
//   mode 1 (store): the factor values are written to Yout[m*N + x].
// In 1D the factor spectra are S itself and the multiplier is applied here.
// Dynamic LDS: vals (n+nf)*Nx doubles, then 2*grp*L complex of FFT scratch.
struct ErrCoef {
  double e[7];
};

__global__ void k_row_c2r(Grid g, FftPlan px, const cx *__restrict__ twN, DevState *st,
                          int check_active, Bufs b, FacList fl, Dict dict, int mode,
                          int src_kind, const double *src_ptr, int dst_kind, int dst_slot,
                          double *dst_ptr, int err, ErrCoef ec, int grp, double *Yout) {
  extern __shared__ double dsm[];
  if (check_active && !attempt_active(st)) return;
  const int n = g.n, Nx = g.Nx, L = g.L;
  const long N = g.N;
  const long r = blockIdx.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nf = fl.nf;
  double *vals = dsm;
  cx *zA = reinterpret_cast<cx *>(dsm + (mode == 0 ? (long)(n + nf) * Nx : 0L));
  cx *zB = zA + grp * L;

  const double *src = nullptr;
  if (mode == 0) {
    src = src_kind == SRC_Y ? b.y[st->iy] : (src_kind == SRC_YT ? b.y[st->iyt] : src_ptr);
    for (int w = tid; w < n * Nx; w += nt) {
      const int f = w / Nx, x = w - f * Nx;
      vals[w] = src[f * N + r * Nx + x];
    }
  }
  for (int g0 = 0; g0 < nf; g0 += grp) {
    const int ng = (nf - g0) < grp ? (nf - g0) : grp;
    for (int w = tid; w < ng * L; w += nt) {
      const int mm = w / L, k = w - mm * L;
      const int m = g0 + mm;
      cx Av, Bv;
      if (g.ndim == 1) {
        const cx *row = b.S + (long)fl.ffield[m] * g.Nh;
        const int blk = fl.fblk[m];
        int nk0 = k, nk1 = L - k;
        Av = apply_mult_y(g, blk, &nk0, row[k], g.invN);
        Bv = apply_mult_y(g, blk, &nk1, row[L - k], g.invN);
      } else {
        const cx *row = b.F + (long)m * g.Nh + r * g.Nxp;
        Av = row[k];
        Bv = row[L - k];
      }
      Bv = cconj(Bv);
      if (k == 0) {
        Av.im = 0.0;
        Bv.im = 0.0;
      }
      const cx E = cadd(Av, Bv), D = csub(Av, Bv);
      const cx t = cmul(D, cconj(twN[k]));
      zA[mm * L + k] = cx{E.re - t.im, E.im + t.re};
    }
    __syncthreads();
    const cx *Z = lds_fft<+1, false>(zA, zB, px, ng, L, 1, tid, nt);
    if (mode == 0) {
      for (int w = tid; w < ng * L; w += nt) {
        const int mm = w / L, j = w - mm * L;
        const cx z = Z[mm * L + j];
        double *v = vals + (long)(n + g0 + mm) * Nx;
        v[2 * j] = z.re;
        v[2 * j + 1] = z.im;
      }
    } else {
      for (int w = tid; w < ng * L; w += nt) {
        const int mm = w / L, j = w - mm * L;
        const cx z = Z[mm * L + j];
        *reinterpret_cast<double2 *>(Yout + (long)(g0 + mm) * N + r * Nx + 2 * j) =
            double2{z.re, z.im};
      }
    }
    __syncthreads();
  }
  if (mode != 0) return;

  double *out;
  if (dst_kind == DST_PTR) out = dst_ptr;
  else if (dst_kind == DST_K1) out = b.k[st->k1s];
  else if (dst_kind == DST_K7) out = b.k[st->k7s];
  else out = b.k[dst_slot];
  double *sq = reinterpret_cast<double *>(zA);  // reuse FFT scratch (n*Nx <= 4*grp*L doubles)
  const double *yv = nullptr, *k1 = nullptr;
  double h = 0, atol = 0, rtol = 0;
  if (err) {
    yv = b.y[st->iy];
    k1 = b.k[st->k1s];
    h = st->h;
    atol = st->atol;
    rtol = st->rtol;
  }
  for (int x = tid; x < Nx; x += nt) {
    double acc[kNMax];
#pragma unroll
    for (int f = 0; f < kNMax; ++f) acc[f] = 0.0;
    for (int t = 0; t < dict.nterms; ++t) {
      double term = 1.0;
      for (int j = dict.tstart[t]; j < dict.tstart[t + 1]; ++j)
        term = term * ipow(vals[(long)dict.fslot[j] * Nx + x], dict.fpow[j]);
      const int fld = dict.tfield[t];
      const double c = dict.tcoef[t];
#pragma unroll
      for (int f = 0; f < kNMax; ++f)
        if (f == fld) acc[f] = __builtin_fma(c, term, acc[f]);
    }
#pragma unroll
    for (int f = 0; f < kNMax; ++f) {
      if (f >= n) break;
      const long idx = f * N + r * Nx + x;
      out[idx] = acc[f];
      if (err) {
        double e = ec.e[0] * k1[idx];
        e = __builtin_fma(ec.e[2], b.k[2][idx], e);
        e = __builtin_fma(ec.e[3], b.k[3][idx], e);
        e = __builtin_fma(ec.e[4], b.k[4][idx], e);
        e = __builtin_fma(ec.e[5], b.k[5][idx], e);
        e = __builtin_fma(ec.e[6], acc[f], e);
        const double wgt = __builtin_fma(rtol, fmax(fabs(yv[idx]), fabs(vals[f * Nx + x])), atol);
        const double qv = (h * e) / wgt;
        sq[f * Nx + x] = qv * qv;
      }
    }
  }
  if (!err) return;
  __syncthreads();
  // pairwise tree per field line: node x covers [x, x+2s) after level s
  for (int s = 1; s < Nx; s <<= 1) {
    const int per = (Nx + 2 * s - 1) / (2 * s);
    for (int w = tid; w < n * per; w += nt) {
      const int f = w / per, i = w - f * per;
      const int x = i * 2 * s;
      if (x + s < Nx) sq[f * Nx + x] = sq[f * Nx + x] + sq[f * Nx + x + s];
    }
    __syncthreads();
  }
  if (tid < n) b.part[(long)tid * g.rows + r] = sq[tid * Nx];
}

// Pairwise tree over v[0..cnt) in LDS (elements past cnt count as 0):
// bottom-up, node i covers [i, i+2s) after level s.  Result in v[0].
__device__ void lds_pairsum(double *v, int cnt, int tid, int nt) {
  for (int s = 1; s < cnt; s <<= 1) {
    const int per = (cnt + 2 * s - 1) / (2 * s);
    for (int i = tid; i < per; i += nt) {
      const int x = i * 2 * s;
      if (x + s < cnt) v[x] = v[x] + v[x + s];
    }
    __syncthreads();
  }
}

// Block sums of kRedCap consecutive lines (used when there are more lines
// than one block can hold); each is the subtree of the global pairwise tree.
__global__ void k_reduce(const double *__restrict__ in, long cnt, double *out, DevState *st,
                         int check_active) {
  __shared__ double v[kRedCap];
  if (check_active && !attempt_active(st)) return;
  if (st->fixed) return;
  const long base = (long)blockIdx.x * kRedCap;
  const int c = (int)((cnt - base) < kRedCap ? (cnt - base) : kRedCap);
  for (int i = threadIdx.x; i < c; i += blockDim.x) v[i] = in[base + i];
  __syncthreads();
  lds_pairsum(v, c, threadIdx.x, blockDim.x);
  if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

// The pairwise tree of lds_pairsum over part[0..cnt), cnt <= 4096, in a
// 256-thread block without a barrier per level: thread i sums its 16 values
// part[16i..16i+15] pairwise in registers, the lanes of a wave combine by
// butterfly shuffles (at distance d, lane x adds lane x+d's sum to its own:
// the tree's left + right; the partner computes right + left, the same bits
// since addition commutes), and the block's four waves through LDS.  Values
// past cnt count as +0, which leaves these sums unchanged: the partials are
// sums of squares (>= +0, or NaN).  Every thread calls it; the result is
// valid in thread 0.  lds: 4 doubles.
__device__ __forceinline__ double tree_sum256(const double *__restrict__ part, int cnt,
                                              double *lds) {
  const int tid = threadIdx.x;
  double v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int i = 16 * tid + q;
    v[q] = i < cnt ? part[i] : 0.0;
  }
#pragma unroll
  for (int s = 1; s < 16; s <<= 1)
#pragma unroll
    for (int q = 0; q < 16; q += 2 * s) v[q] = v[q] + v[q + s];
  double a = v[0];
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) a = a + __shfl_xor(a, d, 64);
  if ((tid & 63) == 0) lds[tid >> 6] = a;
  __syncthreads();
  return tid == 0 ? (lds[0] + lds[1]) + (lds[2] + lds[3]) : 0.0;
}

// Error norm and step control (dp45.cu:117-146) on the device.
// v: LDS for cnt doubles.  Every thread of the block calls it.
__device__ __forceinline__ void control_step(DevState *st, const double *__restrict__ part,
                                             long cnt, long nN, double *v) {
  // with 256 threads the partials are summed by tree_sum256, their loads
  // issued together with the state's
  const bool fast = blockDim.x == 256 && cnt <= 4096;
  double sum = 0.0;
  if (fast) sum = tree_sum256(part, (int)cnt, v);
  if (!attempt_active(st)) return;
  const int fixed = st->fixed;
  if (!fixed && !fast) {
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) v[i] = part[i];
    __syncthreads();
    lds_pairsum(v, (int)cnt, threadIdx.x, blockDim.x);
    sum = v[0];
  }
  if (threadIdx.x != 0) return;
  double t = st->t, h = st->h;
  double norm = 0.0, factor = 1.0;
  bool acc;
  if (fixed) {
    acc = true;
  } else {
    norm = sqrt(sum) / sqrt((double)nN);
    factor = 0.9 * pow_m02(norm);
    if (__builtin_isnan(factor)) factor = 0.2;
    acc = norm < 1;
  }
  if (acc) {
    st->t_last = t;
    t = t + h;
    st->t = t;
    st->t_now = t;
    if (!fixed) {
      if (factor > 10) factor = 10;
      h = h * factor;
      if (h > st->hmax) h = st->hmax;
      st->h = h;
    } else if (st->hsched) {
      const long long i = ++st->isched;
      if (i < st->nsched) {
        h = st->hsched[i];
        st->h = h;
      }
    }
    const int iy = st->iy, iyt = st->iyt, iyl = st->iyl;
    st->iy = iyt;
    st->iyt = iyl;
    st->iyl = iy;
    const int k1 = st->k1s;
    st->k1s = st->k7s;
    st->k7s = k1;
    st->eval_valid = 1;
    st->accepted += 1;
    if (t >= st->t_stop) st->halted = 1;
  } else {
    if (factor < 1) {
      if (factor < 0.2) factor = 0.2;
      h = h * factor;
      st->h = h;
    }
    st->eval_valid = 0;
  }
  st->last_norm = norm;
  st->attempts += 1;
  if (st->nolog) return;
  const int i = st->nlog;
  st->log_t[i] = t;
  st->log_h[i] = h;
  st->log_err[i] = norm;
  st->log_acc[i] = acc ? 1 : 0;
  st->nlog = i + 1;
  if (i + 1 >= kLogCap) st->halted = 1;
}

__device__ __forceinline__ void control_body(DevState *st, const double *__restrict__ part,
                                             long cnt, long nN) {
  __shared__ double v[kRedCap];
  control_step(st, part, cnt, nN, v);
}

__global__ void k_control(DevState *st, const double *__restrict__ part, long cnt, long nN) {
  control_body(st, part, cnt, nN);
}

// ----------------------------------------------------------- callback RHS
// A right-hand side given as a host callback (dp45_init's dydt, dp45.h:8)
// instead of a dictionary: the stage sums, the error norm and the step
// control run here, the RHS wherever the callback puts it.  The state is n
// lines (fields) of N values; the norm sums squares per line by the pairwise
// tree, then over the lines -- the tree of a 1-D problem of n fields.

// Values of one line per block of the error pass: a power of two, so every
// block sums an aligned subtree of the line's pairwise tree.
constexpr int kFnChunk = kRedCap;

// Stage input ytemp = y + h * (sum_j a_j k_j), zero coefficients skipped:
// the operations of k_rowf_reg's stage sum (dp45.cu:34-69).
__global__ void k_fn_stage(const DevState *st, Bufs b, StageCoef sc, long nN) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nN) return;
  const double h = st->h;
  double acc = sc.a[0] * b.k[st->k1s][i];
  for (int j = 1; j < sc.ns; ++j)
    if (sc.a[j] != 0.0) acc = __builtin_fma(sc.a[j], b.k[j][i], acc);
  b.y[st->iyt][i] = __builtin_fma(h, acc, b.y[st->iy][i]);
}

// Scaled error squares of chunk (line, chunk) summed by its subtree
// (dp45.cu:79-84): the operations of the x-inverse pass's error tree.
__global__ void k_fn_err(const DevState *st, Bufs b, ErrCoef ec, long N, int nchunk,
                         double *out) {
  __shared__ double v[kFnChunk];
  const int line = blockIdx.x / nchunk, ch = blockIdx.x - line * nchunk;
  const long off = (long)ch * kFnChunk;
  const int c = (int)((N - off) < kFnChunk ? (N - off) : kFnChunk);
  const long base = (long)line * N + off;
  const double h = st->h, atol = st->atol, rtol = st->rtol;
  const double *y = b.y[st->iy], *yn = b.y[st->iyt], *k1 = b.k[st->k1s], *k7 = b.k[st->k7s];
  for (int x = threadIdx.x; x < c; x += blockDim.x) {
    const long i = base + x;
    double e = ec.e[0] * k1[i];
    e = __builtin_fma(ec.e[2], b.k[2][i], e);
    e = __builtin_fma(ec.e[3], b.k[3][i], e);
    e = __builtin_fma(ec.e[4], b.k[4][i], e);
    e = __builtin_fma(ec.e[5], b.k[5][i], e);
    e = __builtin_fma(ec.e[6], k7[i], e);
    const double w = __builtin_fma(rtol, fmax(fabs(y[i]), fabs(yn[i])), atol);
    const double q = (h * e) / w;
    v[x] = q * q;
  }
  __syncthreads();
  lds_pairsum(v, c, threadIdx.x, blockDim.x);
  if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

// Per line, the pairwise tree over its chunk sums (block = line).
__global__ void k_fn_lines(const double *__restrict__ in, int nchunk, double *out) {
  __shared__ double v[kRedCap];
  const double *p = in + (long)blockIdx.x * nchunk;
  for (int i = threadIdx.x; i < nchunk; i += blockDim.x) v[i] = p[i];
  __syncthreads();
  lds_pairsum(v, nchunk, threadIdx.x, blockDim.x);
  if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

// ---------------------------------------------------------------- batches
// One instance of a batch, as the batched kernels see it (a device table of
// these; each instance owns its buffers, state, factor list and dictionary).
struct Inst {
  Grid g;
  Bufs b;
  DevState *st;
  FacList fl;
  DictP dict;
};

// Start of a batched run or step: per-instance t, h, t1; no output stops.
__global__ void k_begin_b(const Inst *__restrict__ in, int ninst, const double *__restrict__ t,
                          const double *__restrict__ h, const double *__restrict__ t1,
                          double hmin, double hmax, int nolog) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  DevState *st = in[i].st;
  st->t = t[i];
  st->h = h[i];
  st->hmin = hmin;
  st->hmax = hmax;
  st->t1 = t1[i];
  st->t_stop = __builtin_inf();
  st->halted = 0;
  st->nlog = 0;
  st->nolog = nolog;
}

// Step control of every instance (block i = instance i; each instance has at
// most kRedCap error lines), then done[i] = instance i has finished.
__global__ void k_control_b(const Inst *__restrict__ in, int *__restrict__ done) {
  const Inst &I = in[blockIdx.x];
  control_body(I.st, I.b.part, (long)I.g.n * I.g.rows, (long)I.g.n * I.g.N);
  if (threadIdx.x == 0) done[blockIdx.x] = attempt_active(I.st) ? 0 : 1;
}

// Dense output of the last accepted step (dp45.cu:70-76): after the FSAL
// swap the step's k1 is in slot k7s and its k7 in slot k1s.
struct DenseCoef {
  double p[7][4];
};

__global__ void k_dense(const DevState *st, Bufs b, DenseCoef dc, double t_eval, double *out,
                        long nN) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nN) return;
  const double *K[7] = {b.k[st->k7s], b.k[1], b.k[2], b.k[3], b.k[4], b.k[5], b.k[st->k1s]};
  const double hh = st->t_now - st->t_last;
  const double x = (t_eval - st->t_last) / hh;
  const double x2 = x * x, x3 = x2 * x, x4 = x3 * x;
  double Q[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    bool first = true;
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < 7; ++j) {
      const double c = dc.p[j][p];
      if (c == 0.0) continue;
      a = first ? c * K[j][i] : __builtin_fma(c, K[j][i], a);
      first = false;
    }
    Q[p] = a;
  }
  double a = x * Q[0];
  a = __builtin_fma(x2, Q[1], a);
  a = __builtin_fma(x3, Q[2], a);
  a = __builtin_fma(x4, Q[3], a);
  out[i] = __builtin_fma(hh, a, b.y[st->iyl][i]);
}

// The factor blocks in the layout of FILEBASEY.dat (rdcu.cu:339-340): complex
// out[(blk*n + field)*N + x] = (value, 0).  Block 0 is the state y itself;
// block blk >= 1 of `field` is row field*(nblk-1) + blk-1 of Yr (the real-space
// factor list in field-major order).  Grid: (ceil(N/blockDim), nblk*n).
__global__ void k_pack_y(const double *__restrict__ y, const double *__restrict__ Yr,
                         double2 *__restrict__ out, int n, int nblk, long N) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  const int bf = blockIdx.y, blk = bf / n, f = bf - blk * n;
  const double v =
      blk == 0 ? y[(long)f * N + x] : Yr[((long)f * (nblk - 1) + blk - 1) * N + x];
  out[(long)bf * N + x] = make_double2(v, 0.0);
}

// Whether derivative block blk has an imaginary part in Y: first
// derivatives and mixed second derivatives (Nyquist leakage); block 0 and
// the pure second derivatives have none.
__host__ __device__ inline bool y_has_imag(int blk, int nd) {
  if (blk < 1) return false;
  if (blk <= nd) return true;
  const int q = blk - 1 - nd;
  return q / nd != q % nd;
}

// Imaginary parts of Y from the transforms of the kImagBlk factors (Yi, the
// layout of k_pack_y's Yr): written into out's .y of the blocks that have one.
__global__ void k_pack_y_imag(const double *__restrict__ Yi, double2 *__restrict__ out, int n,
                              int nblk, int nd, long N) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  const int bf = blockIdx.y, blk = bf / n, f = bf - blk * n;
  if (!y_has_imag(blk, nd)) return;
  out[(long)bf * N + x].y = Yi[((long)f * (nblk - 1) + blk - 1) * N + x];
}

}  // namespace rd
