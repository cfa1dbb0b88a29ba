// oracle.cpp -- FP64 CPU restatement of rdcu's explicit path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
// implementation in rdcu_amd/; only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it.  Nothing in the product links it.
//
// It exports the same C ABI as the GPU library (include/rdcu_dp45.h), so the
// drop-in CLI can be linked against either, plus a few oracle_* hooks that the
// tests use to pin the pieces one by one.
//
// What it restates (reference file:line):
//   DP45 tableau, error weights, dense output     dp45.cu:6-21
//   stage sums / error / interpolation kernels     dp45.cu:34-84
//   step control (accept, reject, NaN, hmax)       dp45.cu:87-147
//   dense output eval                              dp45.cu:149-152
//   run loop with FSAL                             dp45.cu:154-169
//   wavenumber convention (Nyquist negative)       rdcu.cu:78-99
//   makeY: fwd FFT, d/dx_a, d/dx_b d/dx_a, inv FFT rdcu.cu:117-165
//   dictionary evaluation in term/factor order     rdcu.cu:101-106, 167-184
//
// Arithmetic is fixed operation by operation (explicit fma, no contraction:
// build with -ffp-contract=off) so that the GPU path, which follows the same
// operation sequence, can be compared bit for bit:
//   * complex FFT: Stockham autosort, radix plan [5..][3..][2|4][8..]; twiddles
//     from a host table T_L[m] = exp(-2 pi i m / L);
//   * real transforms: half-length complex FFT plus the standard split step;
//   * spectral multipliers: the Hermitian part of the reference's i*k products,
//     scaled by 1/N (the reference's Re[(1/N) ifft(m * fft(u))]);
//   * error norm: sum of squares by a pairwise tree per grid line, then a
//     pairwise tree over lines; pow(norm, -0.2) correctly rounded.
//
// Parity pins (tests/test_oracle_*.py): brute-force DFT and numpy.fft for the
// transforms, closed-form derivatives of trigonometric polynomials, the CGLE
// formula against the 18-term dictionary, exact solutions of linear ODEs and
// the observed 5th order, Decimal arithmetic for pow(x,-0.2), the glibc rand()
// initial values, and one test per step-control branch.
//
// FILEBASEY.dat imaginary parts: the Nyquist leakage of the first and mixed
// derivative blocks (the transform of the anti-Hermitian part of the
// reference's multiplier); 0 in block 0 (uninitialised memory in the
// reference) and in the pure second derivatives (roundoff there).

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>
#include <omp.h>

#include "rdcu_dp45.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// ---------------------------------------------------------------- complex ops
struct cx {
  double re, im;
};

inline cx add(cx a, cx b) { return {a.re + b.re, a.im + b.im}; }
inline cx sub(cx a, cx b) { return {a.re - b.re, a.im - b.im}; }
inline cx conjg(cx a) { return {a.re, -a.im}; }
// a*w with the two products of each component fused as written.
inline cx cmul(cx a, cx w) {
  return {std::fma(a.re, w.re, -(a.im * w.im)), std::fma(a.re, w.im, a.im * w.re)};
}
// times -i (forward, dir < 0) or +i (inverse, dir > 0)
inline cx rot(cx a, int dir) { return dir < 0 ? cx{a.im, -a.re} : cx{-a.im, a.re}; }

const double kR2 = 0x1.6a09e667f3bcdp-1;  // sqrt(1/2)
const double kS3 = 0x1.bb67ae8584caap-1;  // sqrt(3)/2
const double kC51 = 0x1.3c6ef372fe950p-2;  // cos(2pi/5)
const double kC52 = -0x1.9e3779b97f4a8p-1; // cos(4pi/5)
const double kS51 = 0x1.e6f0e134454ffp-1;  // sin(2pi/5)
const double kS52 = 0x1.2cf2304755a5ep-1;  // sin(4pi/5)

// exp(-+ i pi/4) * a
inline cx rot8_1(cx a, int dir) {
  if (dir < 0) return {(a.re + a.im) * kR2, (a.im - a.re) * kR2};
  return {(a.re - a.im) * kR2, (a.im + a.re) * kR2};
}
// exp(-+ 3 i pi/4) * a
inline cx rot8_3(cx a, int dir) {
  if (dir < 0) return {(a.im - a.re) * kR2, -((a.re + a.im) * kR2)};
  return {-((a.re + a.im) * kR2), (a.re - a.im) * kR2};
}

void dft2(cx *v) {
  cx a = v[0], b = v[1];
  v[0] = add(a, b);
  v[1] = sub(a, b);
}
void dft4(cx *v, int dir) {
  cx a = add(v[0], v[2]), b = sub(v[0], v[2]);
  cx c = add(v[1], v[3]), d = rot(sub(v[1], v[3]), dir);
  v[0] = add(a, c);
  v[1] = add(b, d);
  v[2] = sub(a, c);
  v[3] = sub(b, d);
}
void dft8(cx *v, int dir) {
  cx e[4] = {v[0], v[2], v[4], v[6]};
  cx o[4] = {v[1], v[3], v[5], v[7]};
  dft4(e, dir);
  dft4(o, dir);
  cx t[4] = {o[0], rot8_1(o[1], dir), rot(o[2], dir), rot8_3(o[3], dir)};
  for (int k = 0; k < 4; ++k) {
    v[k] = add(e[k], t[k]);
    v[k + 4] = sub(e[k], t[k]);
  }
}
void dft3(cx *v, int dir) {
  cx t = add(v[1], v[2]), d = sub(v[1], v[2]);
  cx y0 = add(v[0], t);
  cx m = {v[0].re - 0.5 * t.re, v[0].im - 0.5 * t.im};
  cx u = rot(cx{kS3 * d.re, kS3 * d.im}, dir);
  v[0] = y0;
  v[1] = add(m, u);
  v[2] = sub(m, u);
}
void dft5(cx *v, int dir) {
  cx t1 = add(v[1], v[4]), t2 = add(v[2], v[3]);
  cx d1 = sub(v[1], v[4]), d2 = sub(v[2], v[3]);
  cx y0 = add(v[0], add(t1, t2));
  cx a1 = {std::fma(kC52, t2.re, std::fma(kC51, t1.re, v[0].re)),
           std::fma(kC52, t2.im, std::fma(kC51, t1.im, v[0].im))};
  cx a2 = {std::fma(kC51, t2.re, std::fma(kC52, t1.re, v[0].re)),
           std::fma(kC51, t2.im, std::fma(kC52, t1.im, v[0].im))};
  cx b1 = {std::fma(kS51, d1.re, kS52 * d2.re), std::fma(kS51, d1.im, kS52 * d2.im)};
  cx b2 = {std::fma(kS52, d1.re, -(kS51 * d2.re)), std::fma(kS52, d1.im, -(kS51 * d2.im))};
  cx r1 = rot(b1, dir), r2 = rot(b2, dir);
  v[0] = y0;
  v[1] = add(a1, r1);
  v[4] = sub(a1, r1);
  v[2] = add(a2, r2);
  v[3] = sub(a2, r2);
}
void dft(int R, cx *v, int dir) {
  switch (R) {
    case 2: dft2(v); break;
    case 3: dft3(v, dir); break;
    case 4: dft4(v, dir); break;
    case 5: dft5(v, dir); break;
    case 8: dft8(v, dir); break;
  }
}

// exp(-2 pi i m / L), exact at multiples of a quarter turn.
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

// Radix plan: all 5s, all 3s, then one 2 or 4 for the leftover power of two,
// then 8s.  Returns false for lengths with other prime factors.
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

struct FFTPlan {
  int L = 0;
  std::vector<int> rad;
  std::vector<cx> tw;  // T_L
  bool init(int len) {
    L = len;
    if (!make_plan(L, rad)) return false;
    tw.resize(L);
    for (int m = 0; m < L; ++m) tw[m] = unit_root(m, L);
    return true;
  }
  // Out-of-place Stockham; x and y both of length L (x is clobbered).
  // Result is returned in `out` (may alias x).
  void run(cx *x, cx *y, cx *out, int dir) const {
    cx *src = x, *dst = y;
    int Ns = 1;
    cx v[8];
    for (int R : rad) {
      const int stride = L / R;
      const int twstep = L / (Ns * R);
      for (int j = 0; j < stride; ++j) {
        const int k = j % Ns;
        for (int r = 0; r < R; ++r) v[r] = src[j + r * stride];
        if (k != 0) {
          for (int r = 1; r < R; ++r) {
            cx w = tw[(k * r) * twstep];
            if (dir > 0) w = conjg(w);
            v[r] = cmul(v[r], w);
          }
        }
        dft(R, v, dir);
        const int base = (j - k) * R + k;
        for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
      }
      Ns *= R;
      std::swap(src, dst);
    }
    if (src != out) std::memcpy(out, src, sizeof(cx) * L);
  }
};

// Real transforms of even length N through the length-N/2 complex FFT.
struct RealPlan {
  int N = 0, L = 0;
  FFTPlan half;
  std::vector<cx> twN;  // T_N
  bool init(int n) {
    N = n;
    L = n / 2;
    if (n < 2 || n % 2) return false;
    if (!half.init(L)) return false;
    twN.resize(N);
    for (int m = 0; m < N; ++m) twN[m] = unit_root(m, N);
    return true;
  }
  // in: N reals; out: L+1 complex; scratch: 2L complex
  void r2c(const double *in, cx *out, cx *scr) const {
    cx *z = scr, *w = scr + L;
    for (int j = 0; j < L; ++j) z[j] = {in[2 * j], in[2 * j + 1]};
    half.run(z, w, z, -1);
    for (int k = 0; k <= L; ++k) {
      cx A = z[k % L];
      cx B = conjg(z[(L - k) % L]);
      cx E = add(A, B), D = sub(A, B);
      cx t = cmul(D, twN[k]);
      out[k] = {0.5 * (E.re + t.im), 0.5 * (E.im - t.re)};
    }
  }
  // in: L+1 complex (imaginary parts of bins 0 and L ignored); out: N reals
  // equal to N times the inverse transform.
  void c2r(const cx *in, double *out, cx *scr) const {
    cx *z = scr, *w = scr + L;
    for (int k = 0; k < L; ++k) {
      cx A = in[k];
      cx B = conjg(in[L - k]);
      if (k == 0) {
        A.im = 0.0;
        B.im = 0.0;
      }
      cx E = add(A, B), D = sub(A, B);
      cx t = cmul(D, conjg(twN[k]));
      z[k] = {E.re - t.im, E.im + t.re};
    }
    half.run(z, w, z, +1);
    for (int j = 0; j < L; ++j) {
      out[2 * j] = z[j].re;
      out[2 * j + 1] = z[j].im;
    }
  }
};

// Pairwise sum over an implicit zero-padding to the next power of two:
// sum(v[0:P]) = sum(v[0:P/2]) + sum(v[P/2:P]).
double pairsum(const double *v, long n) {
  if (n <= 0) return 0.0;
  if (n == 1) return v[0];
  long P = 1;
  while (P < n) P <<= 1;
  long h = P >> 1;
  if (n <= h) return pairsum(v, n) + 0.0;
  return pairsum(v, h) + pairsum(v + h, n - h);
}

// --------------------------------------------------------- pow(x, -0.2)
// Correctly rounded x^y for y = the double nearest -0.2 (= -1/5 - eps,
// eps = 1/(5*2^54)).  x^y = x^(-1/5) * exp(-eps*ln x); x^(-1/5) from Newton on
// w^-5 = m, refined once in double-double.
const double kC5hi[5] = {1.0, 0x1.bdb8cdadbe120p-1, 0x1.8406003b2ae5cp-1, 0x1.51cb453b9536cp-1,
                         0x1.2611186bae675p-1};
const double kC5lo[5] = {0.0, 0x1.0a7b0e98593fdp-56, 0x1.ae64d1cc83ab8p-55, 0x1.15a4ea11db2e6p-55,
                         -0x1.a56b786f7e27cp-55};

double pow_m02(double x) {
  if (std::isnan(x)) return x;
  if (x == 0.0) return INFINITY;
  if (x < 0.0) return NAN;
  if (std::isinf(x)) return 0.0;
  int e;
  double m = std::frexp(x, &e);  // [0.5, 1)
  m *= 2.0;
  e -= 1;                        // x = m 2^e, m in [1, 2)
  int r = e % 5;
  if (r < 0) r += 5;
  int q = (e - r) / 5;
  double w = 1.0 - 0.12944943670387588 * (m - 1.0);
  for (int it = 0; it < 5; ++it) {
    double w2 = w * w;
    double w5 = (w2 * w2) * w;
    w = std::fma(w * 0.2, 1.0 - m * w5, w);
  }
  // residual rho = 1 - m w^5 in double-double
  double a2h = w * w, a2l = std::fma(w, w, -a2h);
  double w4h = a2h * a2h, w4l = std::fma(a2h, a2h, -w4h) + 2.0 * (a2h * a2l);
  double w5h = w4h * w, w5l = std::fma(w4h, w, -w5h) + w4l * w;
  double ph = m * w5h, pl = std::fma(m, w5h, -ph) + m * w5l;
  double rho = (1.0 - ph) - pl;
  double dlt = (w * rho) * 0.2;  // m^(-1/5) = w + dlt
  // x^(-eps) correction: c = -eps*ln x + (eps ln x)^2/2
  const double eps = 0x1.999999999999ap-57;  // 1/(5*2^54) rounded
  double lx = std::log(x);
  double el = eps * lx;
  double c = -el + 0.5 * (el * el);
  // v = c_r * (w + dlt) * (1 + c)
  double vh = kC5hi[r] * w;
  double vl = std::fma(kC5hi[r], w, -vh) + (kC5hi[r] * dlt + kC5lo[r] * w);
  vl = vl + vh * c;
  double res = vh + vl;
  return std::ldexp(res, -q);
}

// ------------------------------------------------------------- DP45 constants
// Dormand & Prince (1980) tableau, error weights and the dense-output matrix of
// the reference (dp45.cu:6-21).
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

// ----------------------------------------------------------------- problem
struct Term {
  int field;
  double coef;
  std::vector<int> idx, pw;
};

struct Factor {  // a transformed derivative block of one field
  int field, blk;
};

const double kPI = 3.14159265358979323846;  // the reference's PI (rdcu.cu:14)

struct Problem {
  int n = 0, ndim = 0, Ns[3] = {1, 1, 1};
  double Ls[3] = {1, 1, 1};
  long N = 0;          // grid points per field
  int Nx = 0, Nxh = 0; // last axis, half spectrum
  long rows = 0;       // N / Nx
  long Nh = 0;         // rows * Nxh
  int nblk = 0;        // 1 + ndim + ndim^2
  std::vector<Term> terms;
  std::vector<std::vector<double>> kt;  // wavenumbers per axis
  double invN = 0;
  double invNxy = 0;  // 3-D: 1/(N_y N_x), the scale of the factors without k_z
};

int build_problem(const rdcu_problem *p, Problem &P) {
  if (!p) return fail(RDCU_EINVAL, "null problem");
  if (p->n < 1) return fail(RDCU_EINVAL, "n must be >= 1");
  if (p->ndim < 1 || p->ndim > 3) return fail(RDCU_EINVAL, "ndim must be 1, 2 or 3");
  P.n = p->n;
  P.ndim = p->ndim;
  P.N = 1;
  for (int a = 0; a < 3; ++a) {
    P.Ns[a] = a < p->ndim ? p->Ns[a] : 1;
    P.Ls[a] = a < p->ndim ? p->Ls[a] : 1.0;
  }
  for (int a = 0; a < p->ndim; ++a) {
    if (P.Ns[a] < 2 || P.Ns[a] % 2)
      return fail(RDCU_EUNSUPPORTED, "grid sizes must be even and >= 2");
    std::vector<int> rad;
    int len = (a == p->ndim - 1) ? P.Ns[a] / 2 : P.Ns[a];
    if (!make_plan(len, rad))
      return fail(RDCU_EUNSUPPORTED, "grid sizes must have only the prime factors 2, 3, 5");
    P.N *= P.Ns[a];
  }
  P.Nx = P.Ns[P.ndim - 1];
  P.Nxh = P.Nx / 2 + 1;
  P.rows = P.N / P.Nx;
  P.Nh = P.rows * P.Nxh;
  P.nblk = 1 + P.ndim + P.ndim * P.ndim;
  P.invN = 1.0 / (double)P.N;
  P.invNxy = P.ndim == 3 ? 1.0 / (double)((long)P.Ns[1] * P.Ns[2]) : P.invN;
  P.kt.assign(P.ndim, {});
  for (int a = 0; a < P.ndim; ++a) {
    int Na = P.Ns[a];
    P.kt[a].resize(Na);
    for (int na = 0; na < Na; ++na) {
      int nn = na < Na / 2 ? na : na - Na;
      P.kt[a][na] = ((double)(2 * nn) * kPI) / P.Ls[a];
    }
  }
  if (p->nterms < 0 || (p->nterms > 0 && (!p->coef || !p->field || !p->nfac || !p->fac)))
    return fail(RDCU_EINVAL, "bad dictionary arrays");
  long off = 0;
  const int nidx = P.n * P.nblk;
  for (int i = 0; i < p->nterms; ++i) {
    Term t;
    t.field = p->field[i];
    t.coef = p->coef[i];
    if (t.field < 0 || t.field >= P.n) return fail(RDCU_EINVAL, "term field out of range");
    if (p->nfac[i] < 0) return fail(RDCU_EINVAL, "negative factor count");
    for (int j = 0; j < p->nfac[i]; ++j) {
      int ix = p->fac[off + 2 * j], pw = p->fac[off + 2 * j + 1];
      if (ix < 0 || ix >= nidx) return fail(RDCU_EINVAL, "factor index out of range");
      t.idx.push_back(ix);
      t.pw.push_back(pw);
    }
    off += 2 * p->nfac[i];
    P.terms.push_back(t);
  }
  return RDCU_OK;
}

inline double ipow(double v, int e) {
  if (e == 0) return 1.0;
  int a = e < 0 ? -e : e;
  double r = v;
  for (int j = 1; j < a; ++j) r = r * v;
  return e < 0 ? 1.0 / r : r;
}

// Hermitian spectral multiplier of derivative block blk at wavenumber index
// nk[0..ndim-1] (last axis in [0, Nx/2]).
// Codes blk >= kImagBlk: -i times the anti-Hermitian part of the reference's
// multiplier of block blk - kImagBlk, whose transform is the imaginary part
// the reference's complex makeY leaves in Y (rdcu.cu:151-164): i*k_a on the
// Nyquist plane of axis a for first derivatives, -k_a*k_b where exactly one of
// a, b is at Nyquist for mixed ones, zero elsewhere.
constexpr int kImagBlk = 64;
// sc: the scale, 1/N (P.invN), or 1/(N_y N_x) for the 3-D factors without k_z,
// whose z transforms are skipped (zfree)
inline cx apply_mult(const Problem &P, int blk, const int *nk, cx X, double sc) {
  const int nd = P.ndim;
  auto nyq = [&](int a) { return nk[a] == P.Ns[a] / 2; };
  if (blk >= kImagBlk) {
    blk -= kImagBlk;
    if (blk <= nd) {
      int a = blk - 1;
      if (!nyq(a)) return {0.0, 0.0};
      double val = P.kt[a][nk[a]] * sc;
      return {X.re * val, X.im * val};
    }
    int qd = blk - 1 - nd;
    int a = qd / nd, b = qd % nd;
    if (a == b || nyq(a) == nyq(b)) return {0.0, 0.0};
    double val = (P.kt[a][nk[a]] * P.kt[b][nk[b]]) * sc;
    return {-(X.im * val), X.re * val};
  }
  if (blk <= nd) {
    int a = blk - 1;
    if (nyq(a)) return {0.0, 0.0};
    double val = P.kt[a][nk[a]] * sc;
    return {-(X.im * val), X.re * val};
  }
  int qd = blk - 1 - nd;
  int a = qd / nd, b = qd % nd;
  double val;
  if (a == b) {
    double k = P.kt[a][nk[a]];
    val = (k * k) * sc;
  } else {
    if (nyq(a) != nyq(b)) return {0.0, 0.0};
    val = (P.kt[a][nk[a]] * P.kt[b][nk[b]]) * sc;
  }
  return {-(X.re * val), -(X.im * val)};
}


// 3-D: a factor whose multiplier has no k_z (first and second derivatives in
// y and x, and their imaginary-part codes).  The z transforms of such a
// factor cancel, so it is multiplied in the y pass, right after the forward y
// transform (scale 1/(N_y N_x)), and transformed back along y there; the z
// pass and the inverse y pass carry only the other factors.
inline bool zfree(const Problem &P, int blk) {
  if (P.ndim != 3) return false;
  if (blk >= kImagBlk) blk -= kImagBlk;
  const int nd = P.ndim;
  if (blk <= nd) return blk - 1 != 0;
  const int q = blk - 1 - nd;
  return q / nd != 0 && q % nd != 0;
}

// ---------------------------------------------------------------- RHS engine
struct Engine {
  const Problem *P = nullptr;
  RealPlan rx;                  // last axis
  FFTPlan ay, az;               // axis ndim-2 (y) and axis 0 for 3D (z)
  std::vector<Factor> used;     // referenced derivative factors
  std::vector<Factor> all;      // every derivative block
  std::vector<cx> S, F;         // spectra, transformed factors
  std::vector<double> R;        // real-space factors
  long nrhs = 0;

  int init(const Problem &p) {
    P = &p;
    if (!rx.init(p.Nx)) return fail(RDCU_EUNSUPPORTED, "unsupported x length");
    if (p.ndim >= 2 && !ay.init(p.Ns[p.ndim - 2])) return fail(RDCU_EUNSUPPORTED, "bad y length");
    if (p.ndim == 3 && !az.init(p.Ns[0])) return fail(RDCU_EUNSUPPORTED, "bad z length");
    std::vector<char> seen(p.n * p.nblk, 0);
    for (auto &t : p.terms)
      for (int ix : t.idx) seen[ix] = 1;
    // factors grouped by field, block ascending
    for (int f = 0; f < p.n; ++f)
      for (int b = 1; b < p.nblk; ++b) {
        all.push_back({f, b});
        if (seen[b * p.n + f]) used.push_back({f, b});
      }
    S.resize((size_t)p.n * p.Nh);
    return RDCU_OK;
  }

  // Transforms u (n*N) into the real-space factors listed in fl; returns them
  // in R (fl.size() * N).
  void factors(const double *u, const std::vector<Factor> &fl) {
    const Problem &p = *P;
    const long N = p.N, Nh = p.Nh, rows = p.rows;
    const int Nx = p.Nx, Nxh = p.Nxh, nd = p.ndim;
    const int nf = (int)fl.size();
    F.resize((size_t)nf * Nh);
    R.resize((size_t)nf * N);
    if (nf == 0) return;
    // 1. R2C along x, every row of every field
#pragma omp parallel
    {
      std::vector<cx> scr(Nx + 2);
#pragma omp for schedule(static)
      for (long r = 0; r < (long)p.n * rows; ++r)
        rx.r2c(u + r * Nx, &S[(size_t)r * Nxh], scr.data());
    }
    if (nd == 1) {
      // multiply each half spectrum, C2R
#pragma omp parallel
      {
        std::vector<cx> tmp(Nxh), scr(Nx + 2);
#pragma omp for schedule(static)
        for (int m = 0; m < nf; ++m) {
          const cx *s = &S[(size_t)fl[m].field * Nh];
          for (int k = 0; k < Nxh; ++k) {
            int nk[1] = {k};
            tmp[k] = apply_mult(p, fl[m].blk, nk, s[k], p.invN);
          }
          std::memcpy(&F[(size_t)m * Nh], tmp.data(), sizeof(cx) * Nxh);
          rx.c2r(tmp.data(), &R[(size_t)m * N], scr.data());
        }
      }
      return;
    }
    const int Ny = p.Ns[nd - 2];
    const long planeY = (long)Ny * Nxh;  // one (z) slab
    const int Nz = nd == 3 ? p.Ns[0] : 1;
    // 2. forward y for 3D (in place on S); the factors without k_z are
    // multiplied here and transformed back along y at once
    if (nd == 3) {
#pragma omp parallel
      {
        std::vector<cx> a(Ny), b(Ny), c(Ny);
#pragma omp for schedule(static)
        for (long l = 0; l < (long)p.n * Nz * Nxh; ++l) {
          long f = l / ((long)Nz * Nxh), rem = l % ((long)Nz * Nxh);
          long z = rem / Nxh, kx = rem % Nxh;
          cx *base = &S[(size_t)f * Nh + z * planeY + kx];
          for (int y = 0; y < Ny; ++y) a[y] = base[(long)y * Nxh];
          ay.run(a.data(), b.data(), a.data(), -1);
          for (int y = 0; y < Ny; ++y) base[(long)y * Nxh] = a[y];
          for (int m = 0; m < nf; ++m) {
            if (fl[m].field != f || !zfree(p, fl[m].blk)) continue;
            int nk[3] = {0, 0, (int)kx};
            for (int y = 0; y < Ny; ++y) {
              nk[1] = y;
              c[y] = apply_mult(p, fl[m].blk, nk, a[y], p.invNxy);
            }
            ay.run(c.data(), b.data(), c.data(), +1);
            cx *dst = &F[(size_t)m * Nh + z * planeY + kx];
            for (int y = 0; y < Ny; ++y) dst[(long)y * Nxh] = c[y];
          }
        }
      }
    }
    // 3. axis-0 pass: forward, multiply, inverse per factor
    {
      const FFTPlan &a0 = nd == 3 ? az : ay;
      const int L0 = a0.L;
      const long estride = nd == 3 ? planeY : Nxh;   // element stride along axis 0
      const long nlines = nd == 3 ? planeY : Nxh;    // lines per field
#pragma omp parallel
      {
        std::vector<cx> a(L0), b(L0), c(L0);
#pragma omp for schedule(static)
        for (long l = 0; l < (long)p.n * nlines; ++l) {
          int f = (int)(l / nlines);
          long off = l % nlines;  // (y, kx) for 3D, kx for 2D
          bool any = false;
          for (int m = 0; m < nf; ++m) any |= fl[m].field == f && !zfree(p, fl[m].blk);
          if (!any) continue;
          const cx *src = &S[(size_t)f * Nh + off];
          for (int i = 0; i < L0; ++i) a[i] = src[i * estride];
          a0.run(a.data(), b.data(), a.data(), -1);
          int nk[3];
          if (nd == 3) {
            nk[1] = (int)(off / Nxh);
            nk[2] = (int)(off % Nxh);
          } else {
            nk[1] = (int)off;
          }
          for (int m = 0; m < nf; ++m) {
            if (fl[m].field != f || zfree(p, fl[m].blk)) continue;
            for (int i = 0; i < L0; ++i) {
              nk[0] = i;
              c[i] = apply_mult(p, fl[m].blk, nk, a[i], p.invN);
            }
            a0.run(c.data(), b.data(), c.data(), +1);
            cx *dst = &F[(size_t)m * Nh + off];
            for (int i = 0; i < L0; ++i) dst[i * estride] = c[i];
          }
        }
      }
    }
    // 4. inverse y for 3D (in place on F)
    if (nd == 3) {
#pragma omp parallel
      {
        std::vector<cx> a(Ny), b(Ny);
#pragma omp for schedule(static)
        for (long l = 0; l < (long)nf * Nz * Nxh; ++l) {
          long m = l / ((long)Nz * Nxh), rem = l % ((long)Nz * Nxh);
          long z = rem / Nxh, kx = rem % Nxh;
          if (zfree(p, fl[m].blk)) continue;  // already inverse-transformed along y
          cx *base = &F[(size_t)m * Nh + z * planeY + kx];
          for (int y = 0; y < Ny; ++y) a[y] = base[(long)y * Nxh];
          ay.run(a.data(), b.data(), a.data(), +1);
          for (int y = 0; y < Ny; ++y) base[(long)y * Nxh] = a[y];
        }
      }
    }
    // 5. C2R along x
#pragma omp parallel
    {
      std::vector<cx> scr(Nx + 2);
#pragma omp for schedule(static)
      for (long r = 0; r < (long)nf * rows; ++r)
        rx.c2r(&F[(size_t)r * Nxh], &R[(size_t)r * Nx], scr.data());
    }
  }

  // f = sum_terms coef * prod pow(Y, eta), terms in order (rdcu.cu:175-183).
  void dictionary(const double *u, const std::vector<Factor> &fl, double *f) {
    const Problem &p = *P;
    const long N = p.N;
    std::vector<int> slot(p.n * p.nblk, -1);
    for (int m = 0; m < (int)fl.size(); ++m) slot[fl[m].blk * p.n + fl[m].field] = m;
#pragma omp parallel for schedule(static)
    for (long x = 0; x < N; ++x) {
      double acc[64];
      for (int i = 0; i < p.n; ++i) acc[i] = 0.0;
      for (const Term &t : p.terms) {
        double term = 1.0;
        for (size_t j = 0; j < t.idx.size(); ++j) {
          int ix = t.idx[j];
          int blk = ix / p.n, fld = ix % p.n;
          double v = blk == 0 ? u[(size_t)fld * N + x] : R[(size_t)slot[ix] * N + x];
          term = term * ipow(v, t.pw[j]);
        }
        acc[t.field] = std::fma(t.coef, term, acc[t.field]);
      }
      for (int i = 0; i < p.n; ++i) f[(size_t)i * N + x] = acc[i];
    }
  }

  void rhs(const double *u, double *f) {
    factors(u, used);
    dictionary(u, used, f);
    ++nrhs;
  }
};

}  // namespace

// ---------------------------------------------------------------- context
struct rdcu_ctx {
  Problem P;
  Engine E;
  double atol = 1e-8, rtol = 1e-8;
  int fixed = 0;
  std::vector<double> y, yt, yl, k[7], q;
  double t_last = 0, t_now = 0;
  bool k1_valid = false, eval_valid = false;
  double last_norm = 0;
  long long attempts = 0, accepted = 0;
  bool member = false;  // owned by a batch, stepped only through it
  // right-hand side given as a callback (rdcu_create_fn): no dictionary
  rdcu_dydt_fn fn = nullptr;
  void *fn_user = nullptr;
};

// A batch is its instances stepped one after another (the sweep of
// sweep.sh:21-41 without the process per instance).
struct rdcu_batch {
  std::vector<rdcu_ctx *> m;
  ~rdcu_batch() {
    for (rdcu_ctx *c : m) delete c;
  }
};

namespace {

// Stage times of dp45.cu:90-106: t + h*c_s, and t + h for the last stage.
const double STAGE_C[6] = {1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0, 1.0};

// f = dydt(t, y): the dictionary, or the handle's callback.
int rhs_at(rdcu_ctx *c, double t, const double *y, double *f) {
  if (!c->fn) {
    c->E.rhs(y, f);
    return RDCU_OK;
  }
  if (c->fn(t, y, f, nullptr, c->fn_user))
    return fail(RDCU_ECALLBACK, "the right-hand side callback returned an error");
  return RDCU_OK;
}

int no_dict() {
  return fail(RDCU_EUNSUPPORTED, "this call needs a dictionary problem (the handle's RHS is a callback)");
}

// one DP45 attempt (dp45.cu:87-147)
int attempt(rdcu_ctx *c, double *t, double *h, double hmax, int *acc) {
  const Problem &P = c->P;
  const long nN = (long)P.n * P.N;
  if (!c->k1_valid) {
    if (int st = rhs_at(c, *t, c->y.data(), c->k[0].data())) return st;
    c->k1_valid = true;
  }
  const double hh = *h;
  for (int s = 2; s <= 7; ++s) {
    const double *a = STAGE_A[s - 2];
    double *yt = c->yt.data();
    const double *y = c->y.data();
    // ytemp = y + h * (sum_j a_j k_j), zero coefficients skipped (dp45.cu:34-69)
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nN; ++i) {
      double sacc = a[0] * c->k[0][i];
      for (int j = 1; j < s - 1; ++j)
        if (a[j] != 0.0) sacc = std::fma(a[j], c->k[j][i], sacc);
      yt[i] = std::fma(hh, sacc, y[i]);
    }
    const double ts = s < 7 ? *t + hh * STAGE_C[s - 2] : *t + hh;
    if (int st = rhs_at(c, ts, yt, c->k[s - 1].data())) return st;
  }
  ++c->attempts;
  if (c->fixed) {  // dp45.cu:109-115
    c->t_last = *t;
    std::swap(c->yl, c->y);
    std::swap(c->y, c->yt);
    *t = *t + *h;
    c->t_now = *t;
    std::swap(c->k[0], c->k[6]);
    c->eval_valid = true;
    ++c->accepted;
    *acc = 1;
    return RDCU_OK;
  }
  // scaled error, then pairwise sums per x-line and over lines (dp45.cu:79-84, 117-119)
  const long Nx = P.Nx, nlines = nN / Nx;
  c->q.resize(nlines);
  const double atol = c->atol, rtol = c->rtol;
#pragma omp parallel
  {
    std::vector<double> sq(Nx);
#pragma omp for schedule(static)
    for (long l = 0; l < nlines; ++l) {
      for (long x = 0; x < Nx; ++x) {
        long i = l * Nx + x;
        double e = EE[0] * c->k[0][i];
        for (int j = 2; j < 7; ++j) e = std::fma(EE[j], c->k[j][i], e);
        double w = std::fma(rtol, std::fmax(std::fabs(c->y[i]), std::fabs(c->yt[i])), atol);
        double qv = (hh * e) / w;
        sq[x] = qv * qv;
      }
      c->q[l] = pairsum(sq.data(), Nx);
    }
  }
  double ss = pairsum(c->q.data(), nlines);
  double norm = std::sqrt(ss) / std::sqrt((double)nN);
  c->last_norm = norm;
  double factor = 0.9 * pow_m02(norm);
  if (std::isnan(factor)) factor = 0.2;
  if (norm < 1) {
    c->t_last = *t;
    std::swap(c->yl, c->y);
    std::swap(c->y, c->yt);
    *t = *t + *h;
    c->t_now = *t;
    if (factor > 10) factor = 10;
    *h = *h * factor;
    if (*h > hmax) *h = hmax;
    std::swap(c->k[0], c->k[6]);  // FSAL: k1 <- k7; old k1 kept in slot 7 for eval
    c->eval_valid = true;
    ++c->accepted;
    *acc = 1;
    return RDCU_OK;
  }
  if (factor < 1) {
    if (factor < 0.2) factor = 0.2;
    *h = *h * factor;
  }
  c->eval_valid = false;
  *acc = 0;
  return RDCU_OK;
}

}  // namespace


// ------------------------------------------------- differentiation (frozen steps)
// A frozen schedule takes DP45 steps of given sizes, every step accepted with
// no error control.  Stage inputs Ys[s] and stage derivatives K[s] (s = 0..6
// for k1..k7) of one step follow attempt() operation for operation, so the
// trajectory equals a fixed-step run.
namespace adj {

using Vec = std::vector<double>;

void stages(rdcu_ctx *c, const double *y, double h, Vec *Ys, Vec *K) {
  const long nN = (long)c->P.n * c->P.N;
  Ys[0].assign(y, y + nN);
  K[0].resize(nN);
  c->E.rhs(Ys[0].data(), K[0].data());
  for (int s = 2; s <= 7; ++s) {
    const double *a = STAGE_A[s - 2];
    Vec &yt = Ys[s - 1];
    yt.resize(nN);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nN; ++i) {
      double sacc = a[0] * K[0][i];
      for (int j = 1; j < s - 1; ++j)
        if (a[j] != 0.0) sacc = std::fma(a[j], K[j][i], sacc);
      yt[i] = std::fma(h, sacc, y[i]);
    }
    K[s - 1].resize(nN);
    c->E.rhs(yt.data(), K[s - 1].data());
  }
}

// Dense output of a step (the operations of rdcu_eval).
void dense(const double *yl, const Vec *K, double x, double hh, double *out, long nN) {
  const double x2 = x * x, x3 = x2 * x, x4 = x3 * x;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nN; ++i) {
    double Q[4];
    for (int p = 0; p < 4; ++p) {
      bool first = true;
      double acc = 0.0;
      for (int j = 0; j < 7; ++j) {
        if (PP[j][p] == 0.0) continue;
        acc = first ? PP[j][p] * K[j][i] : std::fma(PP[j][p], K[j][i], acc);
        first = false;
      }
      Q[p] = acc;
    }
    double a = x * Q[0];
    a = std::fma(x2, Q[1], a);
    a = std::fma(x3, Q[2], a);
    a = std::fma(x4, Q[3], a);
    out[i] = std::fma(hh, a, yl[i]);
  }
}

// Values of every dictionary slot at u: slot f < n is field f, slot n+m is
// referenced factor m (E.used order).
void slots(rdcu_ctx *c, const double *u, Vec &V) {
  const Problem &P = c->P;
  const long N = P.N;
  const int nf = (int)c->E.used.size();
  c->E.factors(u, c->E.used);
  V.resize((size_t)(P.n + nf) * N);
  std::memcpy(V.data(), u, sizeof(double) * P.n * N);
  if (nf) std::memcpy(V.data() + (size_t)P.n * N, c->E.R.data(), sizeof(double) * nf * N);
}

std::vector<int> slot_of_index(const rdcu_ctx *c) {
  const Problem &P = c->P;
  std::vector<int> so(P.n * P.nblk, -1);
  for (int f = 0; f < P.n; ++f) so[f] = f;
  for (size_t m = 0; m < c->E.used.size(); ++m)
    so[c->E.used[m].blk * P.n + c->E.used[m].field] = P.n + (int)m;
  return so;
}

// D applied to one real field g for derivative block blk (the factor
// pipeline with g in the place of field `fld`).
void apply_block(rdcu_ctx *c, const double *g, int fld, int blk, double *out) {
  const Problem &P = c->P;
  const long N = P.N;
  Vec tmp((size_t)P.n * N, 0.0);
  std::memcpy(tmp.data() + (size_t)fld * N, g, sizeof(double) * N);
  std::vector<Factor> one{{fld, blk}};
  c->E.factors(tmp.data(), one);
  std::memcpy(out, c->E.R.data(), sizeof(double) * N);
}

// out = J(u)^T w and gc[t] += sum_x w[field_t](x) * term_t(u)(x).  Derivative
// blocks are symmetric (second derivatives) or antisymmetric (first) as real
// operators, so D^T g = sign * D g.
void vjp(rdcu_ctx *c, const double *u, const double *w, double *out, Vec &gc) {
  const Problem &P = c->P;
  const long N = P.N;
  const int n = P.n, nf = (int)c->E.used.size(), ns = n + nf;
  Vec V;
  slots(c, u, V);
  const std::vector<int> so = slot_of_index(c);
  Vec G((size_t)ns * N, 0.0);
  const int nt = (int)P.terms.size();
  std::vector<Vec> part(omp_get_max_threads(), Vec(nt, 0.0));
#pragma omp parallel
  {
    Vec &gl = part[omp_get_thread_num()];
#pragma omp for schedule(static)
    for (long x = 0; x < N; ++x) {
      for (int t = 0; t < nt; ++t) {
        const Term &T = P.terms[t];
        const double wv = w[(size_t)T.field * N + x];
        const int nfac = (int)T.idx.size();
        double v[16];
        int sl[16];
        double term = 1.0;
        for (int j = 0; j < nfac && j < 16; ++j) {
          sl[j] = so[T.idx[j]];
          v[j] = V[(size_t)sl[j] * N + x];
          term = term * ipow(v[j], T.pw[j]);
        }
        gl[t] += wv * term;
        for (int j = 0; j < nfac && j < 16; ++j) {
          if (T.pw[j] == 0) continue;
          double d = T.pw[j] * ipow(v[j], T.pw[j] - 1);
          for (int i = 0; i < nfac; ++i)
            if (i != j) d = d * ipow(v[i], T.pw[i]);
          G[(size_t)sl[j] * N + x] += wv * T.coef * d;
        }
      }
    }
  }
  for (const Vec &gl : part)  // fixed order: results do not depend on thread timing
    for (int t = 0; t < nt; ++t) gc[t] += gl[t];
  std::memcpy(out, G.data(), sizeof(double) * n * N);
  Vec D(N);
  for (int m = 0; m < nf; ++m) {
    const Factor &F = c->E.used[m];
    apply_block(c, &G[(size_t)(n + m) * N], F.field, F.blk, D.data());
    const double sgn = F.blk <= P.ndim ? -1.0 : 1.0;
    double *o = out + (size_t)F.field * N;
    for (long x = 0; x < N; ++x) o[x] += sgn * D[x];
  }
}

// Tangent: J(u) du + (df/dcoef) dcoef.
void jvp(rdcu_ctx *c, const double *u, const double *du, const double *dcoef, double *out) {
  const Problem &P = c->P;
  const long N = P.N;
  const int n = P.n;
  Vec V, dV;
  slots(c, u, V);
  slots(c, du, dV);
  const std::vector<int> so = slot_of_index(c);
  const int nt = (int)P.terms.size();
#pragma omp parallel for schedule(static)
  for (long x = 0; x < N; ++x) {
    double acc[64];
    for (int i = 0; i < n; ++i) acc[i] = 0.0;
    for (int t = 0; t < nt; ++t) {
      const Term &T = P.terms[t];
      const int nfac = (int)T.idx.size();
      double term = 1.0, dterm = 0.0;
      for (int j = 0; j < nfac; ++j) {
        const double vj = V[(size_t)so[T.idx[j]] * N + x];
        term = term * ipow(vj, T.pw[j]);
      }
      for (int j = 0; j < nfac; ++j) {
        if (T.pw[j] == 0) continue;
        const double vj = V[(size_t)so[T.idx[j]] * N + x];
        double d = T.pw[j] * ipow(vj, T.pw[j] - 1) * dV[(size_t)so[T.idx[j]] * N + x];
        for (int i = 0; i < nfac; ++i)
          if (i != j) d = d * ipow(V[(size_t)so[T.idx[i]] * N + x], T.pw[i]);
        dterm += d;
      }
      acc[T.field] += T.coef * dterm + (dcoef ? dcoef[t] * term : 0.0);
    }
    for (int i = 0; i < n; ++i) out[(size_t)i * N + x] = acc[i];
  }
}

// Output placement on the schedule: step index and (x, hh) of each output.
int place_outputs(double t0, long long ns, const double *hs, int nout, const double *t_out,
                  std::vector<long long> &ostep, Vec &tn) {
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

}  // namespace adj

extern "C" {

const char *rdcu_last_error(void) { return g_err.c_str(); }
const char *rdcu_backend(void) { return "cpu-oracle"; }

int rdcu_create(const rdcu_problem *prob, double atol, double rtol, int fixedstep,
                const double *y0, int device, rdcu_ctx **out) {
  (void)device;
  if (!out) return fail(RDCU_EINVAL, "null output handle");
  *out = nullptr;
  std::unique_ptr<rdcu_ctx> c(new rdcu_ctx());
  int st = build_problem(prob, c->P);
  if (st) return st;
  st = c->E.init(c->P);
  if (st) return st;
  c->atol = atol;
  c->rtol = rtol;
  c->fixed = fixedstep;
  const size_t nN = (size_t)c->P.n * c->P.N;
  c->y.assign(nN, 0.0);
  if (y0) std::memcpy(c->y.data(), y0, nN * sizeof(double));
  c->yt.assign(nN, 0.0);
  c->yl.assign(nN, 0.0);
  for (auto &kk : c->k) kk.assign(nN, 0.0);
  *out = c.release();
  return RDCU_OK;
}

int rdcu_create_fn(int n, long long N, double atol, double rtol, int fixedstep, const double *y0,
                   rdcu_dydt_fn fn, void *user, int device, rdcu_ctx **out) {
  (void)device;
  if (!out) return fail(RDCU_EINVAL, "null output handle");
  *out = nullptr;
  if (!fn) return fail(RDCU_EINVAL, "null right-hand side callback");
  if (n < 1 || N < 1) return fail(RDCU_EINVAL, "n and N must be >= 1");
  std::unique_ptr<rdcu_ctx> c(new rdcu_ctx());
  c->fn = fn;
  c->fn_user = user;
  Problem &P = c->P;
  P.n = n;
  P.ndim = 1;
  P.N = (long)N;
  P.Ns[0] = (int)std::min<long long>(N, 1 << 30);
  P.Nx = (int)P.Ns[0];
  P.rows = 1;
  P.nblk = 1;
  c->atol = atol;
  c->rtol = rtol;
  c->fixed = fixedstep;
  const size_t nN = (size_t)n * N;
  c->y.assign(nN, 0.0);
  if (y0) std::memcpy(c->y.data(), y0, nN * sizeof(double));
  c->yt.assign(nN, 0.0);
  c->yl.assign(nN, 0.0);
  for (auto &kk : c->k) kk.assign(nN, 0.0);
  *out = c.release();
  return RDCU_OK;
}

int rdcu_destroy(rdcu_ctx *c) {
  delete c;
  return RDCU_OK;
}

int rdcu_step(rdcu_ctx *c, double *t, double *h, double hmin, double hmax, int *accepted) {
  (void)hmin;
  if (!c || !t || !h || !accepted) return fail(RDCU_EINVAL, "null argument");
  if (c->member) return fail(RDCU_ESTATE, "a batch member steps only through its batch");
  return attempt(c, t, h, hmax, accepted);
}

int rdcu_run(rdcu_ctx *c, double *t, double *h, double hmin, double hmax, double t1, double stop,
             rdcu_step_fn fn, void *user) {
  if (!c || !t || !h) return fail(RDCU_EINVAL, "null argument");
  if (c->member) return fail(RDCU_ESTATE, "a batch member steps only through its batch");
  if (int st = rhs_at(c, *t, c->y.data(), c->k[0].data())) return st;  // dp45.cu:157
  c->k1_valid = true;
  while (*t < t1 && *h > hmin) {  // dp45.cu:159
    int acc = 0;
    int st = attempt(c, t, h, hmax, &acc);
    if (st) return st;
    if (fn && fn(c, *t, *h, acc, 1, &stop, user)) return fail(RDCU_ECALLBACK, "callback stop");
  }
  return RDCU_OK;
}

// (a name of its own: this sits in the extern "C" block, where "advance"
// would bind to libc's regexp function of that name)
static int run_attempts(rdcu_ctx *c, double *t, double *h, double hmin, double hmax, double t1,
                        long long max_attempts, long long *done) {
  long long k = 0;
  if (int st = rhs_at(c, *t, c->y.data(), c->k[0].data())) return st;
  c->k1_valid = true;
  while (k < max_attempts && *t < t1 && *h > hmin) {
    int acc = 0;
    int st = attempt(c, t, h, hmax, &acc);
    if (st) return st;
    ++k;
  }
  if (done) *done = k;
  return RDCU_OK;
}

int rdcu_advance(rdcu_ctx *c, double *t, double *h, double hmin, double hmax, double t1,
                 long long max_attempts, long long *done) {
  if (!c || !t || !h) return fail(RDCU_EINVAL, "null argument");
  if (c->member) return fail(RDCU_ESTATE, "a batch member steps only through its batch");
  return run_attempts(c, t, h, hmin, hmax, t1, max_attempts, done);
}

int rdcu_batch_create(const rdcu_problem *probs, int count, const double *const *y0, double atol,
                      double rtol, int fixedstep, int device, rdcu_batch **out) {
  if (!out) return fail(RDCU_EINVAL, "null output handle");
  *out = nullptr;
  if (!probs || count < 1) return fail(RDCU_EINVAL, "a batch needs at least one problem");
  std::unique_ptr<rdcu_batch> b(new rdcu_batch());
  for (int i = 0; i < count; ++i) {
    if (probs[i].ndim > 2) return fail(RDCU_EUNSUPPORTED, "batched instances are 1-D or 2-D");
    rdcu_ctx *c = nullptr;
    int st = rdcu_create(&probs[i], atol, rtol, fixedstep, y0 ? y0[i] : nullptr, device, &c);
    if (st) return fail(st, "instance " + std::to_string(i) + ": " + g_err);
    c->member = true;
    b->m.push_back(c);
  }
  *out = b.release();
  return RDCU_OK;
}

int rdcu_batch_destroy(rdcu_batch *b) {
  delete b;
  return RDCU_OK;
}

int rdcu_batch_member(rdcu_batch *b, int i, rdcu_ctx **out) {
  if (!b || !out) return fail(RDCU_EINVAL, "null argument");
  if (i < 0 || i >= (int)b->m.size()) return fail(RDCU_EINVAL, "instance index out of range");
  *out = b->m[i];
  return RDCU_OK;
}

int rdcu_batch_time(rdcu_batch *b, int reps, double *ms, double *bytes, int *shape,
                    int max_classes, int *nclasses) {
  (void)bytes;
  (void)shape;
  (void)max_classes;
  if (!b || !ms || reps < 1) return fail(RDCU_EINVAL, "bad argument");
  double t0 = omp_get_wtime();
  for (int r = 0; r < reps; ++r)
    for (rdcu_ctx *c : b->m) {
      if (c->fn) return no_dict();
      std::vector<double> f(c->y.size());
      c->E.rhs(c->y.data(), f.data());
    }
  ms[0] = (omp_get_wtime() - t0) * 1e3 / reps;
  if (nclasses) *nclasses = 0;
  return RDCU_OK;
}

int rdcu_batch_step(rdcu_batch *b, double *t, double *h, double hmin, double hmax, int *accepted) {
  (void)hmin;
  if (!b || !t || !h || !accepted) return fail(RDCU_EINVAL, "null argument");
  for (size_t i = 0; i < b->m.size(); ++i) {
    int st = attempt(b->m[i], &t[i], &h[i], hmax, &accepted[i]);
    if (st) return st;
  }
  return RDCU_OK;
}

int rdcu_batch_run(rdcu_batch *b, double *t, double *h, double hmin, double hmax, const double *t1,
                   long long max_attempts, long long *attempts) {
  if (!b || !t || !h || !t1) return fail(RDCU_EINVAL, "null argument");
  if (max_attempts < 0) return fail(RDCU_EINVAL, "max_attempts must be >= 0");
  for (size_t i = 0; i < b->m.size(); ++i) {
    long long k = 0;
    int st = run_attempts(b->m[i], &t[i], &h[i], hmin, hmax, t1[i], max_attempts, &k);
    if (st) return st;
    if (attempts) attempts[i] = k;
  }
  return RDCU_OK;
}


int rdcu_set_coef(rdcu_ctx *c, const double *coef) {
  if (!c || (!coef && !c->P.terms.empty())) return fail(RDCU_EINVAL, "null argument");
  if (c->fn) return no_dict();
  for (size_t t = 0; t < c->P.terms.size(); ++t) c->P.terms[t].coef = coef[t];
  c->k1_valid = false;
  return RDCU_OK;
}

int rdcu_schedule(rdcu_ctx *c, double t0, double h0, double hmin, double hmax, double t1,
                  long long max_steps, double *hs, long long *nsteps) {
  if (!c || !hs || !nsteps) return fail(RDCU_EINVAL, "null argument");
  if (c->fn) return no_dict();
  if (c->member) return fail(RDCU_ESTATE, "a batch member steps only through its batch");
  double t = t0, h = h0;
  long long k = 0;
  c->E.rhs(c->y.data(), c->k[0].data());
  c->k1_valid = true;
  while (t < t1 && h > hmin) {
    const double hu = h;
    int acc = 0;
    int st = attempt(c, &t, &h, hmax, &acc);
    if (st) return st;
    if (acc) {
      if (k >= max_steps) return fail(RDCU_EINVAL, "more accepted steps than max_steps");
      hs[k++] = hu;
    }
  }
  *nsteps = k;
  return RDCU_OK;
}

int rdcu_loss_grad(rdcu_ctx *c, const double *y0, double t0, long long ns, const double *hs,
                   int nout, const double *t_out, const double *target, double *loss,
                   double *grad_coef, double *grad_y0) {
  using adj::Vec;
  if (!c || !y0 || !hs || !t_out || !target || !loss) return fail(RDCU_EINVAL, "null argument");
  if (c->fn) return no_dict();
  if (ns < 1 || nout < 1) return fail(RDCU_EINVAL, "need at least one step and one output");
  const long nN = (long)c->P.n * c->P.N;
  std::vector<long long> ostep;
  Vec tn;
  int st = adj::place_outputs(t0, ns, hs, nout, t_out, ostep, tn);
  if (st) return st;
  const bool grad = grad_coef || grad_y0;
  const double cnt = (double)nout * (double)nN;
  std::vector<Vec> ych;  // start state of every step
  Vec y(y0, y0 + nN), yo(nN);
  Vec Ys[7], K[7];
  Vec gbar((size_t)nout * nN);
  double L = 0.0;
  int o = 0;
  for (long long k = 0; k < ns; ++k) {
    if (grad) ych.push_back(y);
    adj::stages(c, y.data(), hs[k], Ys, K);
    const double hh = tn[k + 1] - tn[k];
    for (; o < nout && ostep[o] == k; ++o) {
      adj::dense(y.data(), K, (t_out[o] - tn[k]) / hh, hh, yo.data(), nN);
      const double *tg = target + (size_t)o * nN;
      for (long i = 0; i < nN; ++i) {
        const double d = yo[i] - tg[i];
        L += d * d;
        gbar[(size_t)o * nN + i] = 2.0 * d / cnt;
      }
    }
    y = Ys[6];
  }
  *loss = L / cnt;
  if (!grad) return RDCU_OK;
  // reverse sweep
  const int nt = (int)c->P.terms.size();
  Vec gc(nt, 0.0), lam(nN, 0.0), ybar(nN), kbar[7], Yb(nN);
  o = nout - 1;
  for (long long k = ns - 1; k >= 0; --k) {
    const double h = hs[k];
    adj::stages(c, ych[k].data(), h, Ys, K);
    const double hh = tn[k + 1] - tn[k];
    for (int j = 0; j < 7; ++j) kbar[j].assign(nN, 0.0);
    ybar.assign(nN, 0.0);
    for (; o >= 0 && ostep[o] == k; --o) {
      const double x = (t_out[o] - tn[k]) / hh;
      const double xp[4] = {x, x * x, x * x * x, x * x * x * x};
      const double *g = &gbar[(size_t)o * nN];
      for (long i = 0; i < nN; ++i) ybar[i] += g[i];
      for (int j = 0; j < 7; ++j) {
        double cj = 0.0;
        for (int p = 0; p < 4; ++p) cj += PP[j][p] * xp[p];
        if (cj == 0.0) continue;
        for (long i = 0; i < nN; ++i) kbar[j][i] += hh * cj * g[i];
      }
    }
    for (int s = 7; s >= 1; --s) {
      adj::vjp(c, Ys[s - 1].data(), kbar[s - 1].data(), Yb.data(), gc);
      if (s == 7)
        for (long i = 0; i < nN; ++i) Yb[i] += lam[i];
      for (long i = 0; i < nN; ++i) ybar[i] += Yb[i];
      if (s >= 2) {
        const double *a = STAGE_A[s - 2];
        for (int j = 0; j < s - 1; ++j) {
          if (a[j] == 0.0) continue;
          for (long i = 0; i < nN; ++i) kbar[j][i] += h * a[j] * Yb[i];
        }
      }
    }
    lam = ybar;
  }
  if (grad_coef)
    for (int t = 0; t < nt; ++t) grad_coef[t] = gc[t];
  if (grad_y0) std::memcpy(grad_y0, lam.data(), sizeof(double) * nN);
  return RDCU_OK;
}

// Batched gradients: the jobs one after the other (each parallel inside).
int rdcu_loss_grad_batch(rdcu_grad_job *jobs, int count, int concurrency) {
  (void)concurrency;
  if (count < 0 || (count > 0 && !jobs)) return fail(RDCU_EINVAL, "bad argument");
  for (int i = 0; i < count; ++i) {
    jobs[i].status = RDCU_OK;
    if (!jobs[i].ctx) return fail(RDCU_EINVAL, "every job needs its own handle");
    for (int k = 0; k < i; ++k)
      if (jobs[k].ctx == jobs[i].ctx) return fail(RDCU_EINVAL, "every job needs its own handle");
  }
  int first = -1;
  std::string msg;
  for (int i = 0; i < count; ++i) {
    rdcu_grad_job &j = jobs[i];
    j.status = rdcu_loss_grad(j.ctx, j.y0, j.t0, j.nsteps, j.hs, j.nout, j.t_out, j.target,
                              j.loss, j.grad_coef, j.grad_y0);
    if (j.status && first < 0) {
      first = i;
      msg = g_err;
    }
  }
  if (first >= 0) return fail(jobs[first].status, msg);
  return RDCU_OK;
}

int rdcu_time_rhs(rdcu_ctx *c, int reps, double *ms, int max, int *count, char *names,
                  int names_len) {
  if (!c || !ms || max < 1 || reps < 1) return fail(RDCU_EINVAL, "bad argument");
  if (c->fn) return no_dict();
  std::vector<double> f(c->y.size());
  double t0 = omp_get_wtime();
  for (int r = 0; r < reps; ++r) c->E.rhs(c->y.data(), f.data());
  ms[0] = (omp_get_wtime() - t0) * 1e3 / reps;
  if (count) *count = 1;
  if (names && names_len > 0) snprintf(names, names_len, "rhs");
  return RDCU_OK;
}

// y_eval = y_old + hh*(x Q0 + x^2 Q1 + x^3 Q2 + x^4 Q3) (dp45.cu:70-76, 149-152)
int rdcu_eval(rdcu_ctx *c, double t_eval, double *y_out) {
  if (!c || !y_out) return fail(RDCU_EINVAL, "null argument");
  if (!c->eval_valid) return fail(RDCU_ESTATE, "no accepted step to interpolate");
  const long nN = (long)c->P.n * c->P.N;
  // after the FSAL swap the step's k1 is in slot 7 and its k7 in slot 1
  const double *K[7] = {c->k[6].data(), c->k[1].data(), c->k[2].data(), c->k[3].data(),
                        c->k[4].data(), c->k[5].data(), c->k[0].data()};
  const double hh = c->t_now - c->t_last;
  const double x = (t_eval - c->t_last) / hh;
  const double x2 = x * x, x3 = x2 * x, x4 = x3 * x;
  const double *yl = c->yl.data();
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nN; ++i) {
    double Q[4];
    for (int p = 0; p < 4; ++p) {
      bool first = true;
      double acc = 0.0;
      for (int j = 0; j < 7; ++j) {
        if (PP[j][p] == 0.0) continue;
        acc = first ? PP[j][p] * K[j][i] : std::fma(PP[j][p], K[j][i], acc);
        first = false;
      }
      Q[p] = acc;
    }
    double a = x * Q[0];
    a = std::fma(x2, Q[1], a);
    a = std::fma(x3, Q[2], a);
    a = std::fma(x4, Q[3], a);
    y_out[i] = std::fma(hh, a, yl[i]);
  }
  return RDCU_OK;
}

int rdcu_get_state(rdcu_ctx *c, double *y_out) {
  if (!c || !y_out) return fail(RDCU_EINVAL, "null argument");
  std::memcpy(y_out, c->y.data(), c->y.size() * sizeof(double));
  return RDCU_OK;
}

int rdcu_set_state(rdcu_ctx *c, const double *y) {
  if (!c || !y) return fail(RDCU_EINVAL, "null argument");
  std::memcpy(c->y.data(), y, c->y.size() * sizeof(double));
  c->k1_valid = false;
  c->eval_valid = false;
  return RDCU_OK;
}

int rdcu_rhs(rdcu_ctx *c, const double *y, double *f_out, double *Y_out) {
  if (!c || !y) return fail(RDCU_EINVAL, "null argument");
  if (c->fn) return no_dict();
  const Problem &P = c->P;
  const long N = P.N;
  const std::vector<Factor> &fl = Y_out ? c->E.all : c->E.used;
  c->E.factors(y, fl);
  std::vector<double> f((size_t)P.n * N);
  c->E.dictionary(y, fl, f.data());
  ++c->E.nrhs;
  if (f_out) std::memcpy(f_out, f.data(), f.size() * sizeof(double));
  if (Y_out) {
    for (int blk = 0; blk < P.nblk; ++blk)
      for (int fld = 0; fld < P.n; ++fld) {
        double *dst = Y_out + 2 * ((size_t)(blk * P.n + fld) * N);
        const double *src;
        if (blk == 0) {
          src = y + (size_t)fld * N;
        } else {
          int m = 0;
          while (!(fl[m].field == fld && fl[m].blk == blk)) ++m;
          src = &c->E.R[(size_t)m * N];
        }
        for (long x = 0; x < N; ++x) {
          dst[2 * x] = src[x];
          dst[2 * x + 1] = 0.0;
        }
      }
    // imaginary parts: the kImagBlk transforms of the blocks that leak
    // (first and mixed derivatives); block 0 and pure second derivatives
    // stay 0 (the reference leaves roundoff, and in block 0 uninitialised
    // memory, there)
    std::vector<Factor> im = c->E.all;
    for (Factor &fc : im) fc.blk += kImagBlk;
    c->E.factors(y, im);
    for (int blk = 1; blk < P.nblk; ++blk) {
      const int q = blk - 1 - P.ndim;
      if (blk > P.ndim && q / P.ndim == q % P.ndim) continue;
      for (int fld = 0; fld < P.n; ++fld) {
        double *dst = Y_out + 2 * ((size_t)(blk * P.n + fld) * N);
        int m = 0;
        while (!(im[m].field == fld && im[m].blk == blk + kImagBlk)) ++m;
        const double *src = &c->E.R[(size_t)m * N];
        for (long x = 0; x < N; ++x) dst[2 * x + 1] = src[x];
      }
    }
  }
  return RDCU_OK;
}

// The CPU build has no device memory: the pointers are host memory.
int rdcu_loss_grad_device(rdcu_ctx *c, const double *y0_dev, double t0, long long ns,
                          const double *hs, int nout, const double *t_out,
                          const double *target_dev, double *loss, double *grad_coef,
                          double *grad_y0_dev) {
  return rdcu_loss_grad(c, y0_dev, t0, ns, hs, nout, t_out, target_dev, loss, grad_coef,
                        grad_y0_dev);
}

int rdcu_rhs_device(rdcu_ctx *c, const double *y_dev, double *f_dev) {
  if (!c || !y_dev || !f_dev) return fail(RDCU_EINVAL, "null argument");
  return rdcu_rhs(c, y_dev, f_dev, nullptr);
}

// The CPU build has no device scratch: the same values as rdcu_rhs.
int rdcu_rhs_noalloc(rdcu_ctx *c, const double *y, double *f_out, double *Y_out) {
  return rdcu_rhs(c, y, f_out, Y_out);
}

// One output frame: the dense output, then the RHS at it (rdcu.cu:270-291).
int rdcu_frame(rdcu_ctx *c, double t_eval, double *y_out, double *f_out, double *Y_out) {
  if (!c || !y_out) return fail(RDCU_EINVAL, "null argument");
  int st = rdcu_eval(c, t_eval, y_out);
  if (st || (!f_out && !Y_out)) return st;
  return rdcu_rhs(c, y_out, f_out, Y_out);
}

int rdcu_host_alloc(size_t bytes, void **out) {
  if (!out) return fail(RDCU_EINVAL, "null argument");
  *out = std::malloc(bytes ? bytes : 16);
  return *out ? RDCU_OK : fail(RDCU_ENOMEM, "malloc failed");
}

int rdcu_host_free(void *p) {
  std::free(p);
  return RDCU_OK;
}

int rdcu_get_stats(rdcu_ctx *c, long long out[4]) {
  if (!c || !out) return fail(RDCU_EINVAL, "null argument");
  out[0] = c->attempts;
  out[1] = c->accepted;
  out[2] = c->E.nrhs;
  out[3] = 0;
  return RDCU_OK;
}

// ------------------------------------------------------------ test hooks
// Complex FFT of length L (dir -1 forward, +1 inverse, unnormalised);
// in/out are interleaved (re, im) arrays of 2L doubles.
int oracle_fft(int L, int dir, const double *in, double *out) {
  FFTPlan pl;
  if (!pl.init(L)) return fail(RDCU_EUNSUPPORTED, "length");
  std::vector<cx> a(L), b(L);
  std::memcpy(a.data(), in, sizeof(cx) * L);
  pl.run(a.data(), b.data(), a.data(), dir);
  std::memcpy(out, a.data(), sizeof(cx) * L);
  return RDCU_OK;
}
// Real-to-half-complex transform of N reals (N/2+1 complex out).
int oracle_r2c(int N, const double *in, double *out) {
  RealPlan pl;
  if (!pl.init(N)) return fail(RDCU_EUNSUPPORTED, "length");
  std::vector<cx> scr(N + 2);
  pl.r2c(in, reinterpret_cast<cx *>(out), scr.data());
  return RDCU_OK;
}
// Half-complex (N/2+1) to N reals, N times the inverse transform.
int oracle_c2r(int N, const double *in, double *out) {
  RealPlan pl;
  if (!pl.init(N)) return fail(RDCU_EUNSUPPORTED, "length");
  std::vector<cx> scr(N + 2);
  pl.c2r(reinterpret_cast<const cx *>(in), out, scr.data());
  return RDCU_OK;
}
int oracle_plan(int L, int *rad, int maxn) {
  std::vector<int> r;
  if (!make_plan(L, r)) return -1;
  for (int i = 0; i < (int)r.size() && i < maxn; ++i) rad[i] = r[i];
  return (int)r.size();
}
void oracle_unit_root(long m, long L, double *out) {
  cx w = unit_root(m, L);
  out[0] = w.re;
  out[1] = w.im;
}
double oracle_pow_m02(double x) { return pow_m02(x); }
double oracle_pairsum(const double *v, long n) { return pairsum(v, n); }
double oracle_last_norm(rdcu_ctx *c) { return c ? c->last_norm : NAN; }
void oracle_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}
int oracle_get_threads(void) { return omp_get_max_threads(); }

// Tangent-linear of rdcu_loss_grad's loss along (dcoef, dy0) on the same
// frozen schedule: *dloss = d loss / d eps at eps = 0 for coef + eps*dcoef,
// y0 + eps*dy0 (forward mode through the same steps).
int oracle_loss_tl(rdcu_ctx *c, const double *y0, double t0, long long ns, const double *hs,
                   int nout, const double *t_out, const double *target, const double *dcoef,
                   const double *dy0, double *loss, double *dloss) {
  using adj::Vec;
  const long nN = (long)c->P.n * c->P.N;
  std::vector<long long> ostep;
  Vec tn;
  int st = adj::place_outputs(t0, ns, hs, nout, t_out, ostep, tn);
  if (st) return st;
  const double cnt = (double)nout * (double)nN;
  Vec y(y0, y0 + nN), dy(dy0, dy0 + nN), yo(nN), dyo(nN);
  Vec Ys[7], K[7], dYs[7], dK[7];
  double L = 0.0, dL = 0.0;
  int o = 0;
  for (long long k = 0; k < ns; ++k) {
    const double h = hs[k];
    adj::stages(c, y.data(), h, Ys, K);
    dYs[0] = dy;
    dK[0].resize(nN);
    adj::jvp(c, Ys[0].data(), dYs[0].data(), dcoef, dK[0].data());
    for (int s = 2; s <= 7; ++s) {
      const double *a = STAGE_A[s - 2];
      dYs[s - 1].resize(nN);
      for (long i = 0; i < nN; ++i) {
        double sacc = 0.0;
        for (int j = 0; j < s - 1; ++j) sacc += a[j] * dK[j][i];
        dYs[s - 1][i] = dy[i] + h * sacc;
      }
      dK[s - 1].resize(nN);
      adj::jvp(c, Ys[s - 1].data(), dYs[s - 1].data(), dcoef, dK[s - 1].data());
    }
    const double hh = tn[k + 1] - tn[k];
    for (; o < nout && ostep[o] == k; ++o) {
      const double x = (t_out[o] - tn[k]) / hh;
      adj::dense(y.data(), K, x, hh, yo.data(), nN);
      const double xp[4] = {x, x * x, x * x * x, x * x * x * x};
      for (long i = 0; i < nN; ++i) {
        double q = 0.0;
        for (int j = 0; j < 7; ++j) {
          double cj = 0.0;
          for (int p = 0; p < 4; ++p) cj += PP[j][p] * xp[p];
          q += cj * dK[j][i];
        }
        dyo[i] = dy[i] + hh * q;
      }
      const double *tg = target + (size_t)o * nN;
      for (long i = 0; i < nN; ++i) {
        const double d = yo[i] - tg[i];
        L += d * d;
        dL += 2.0 * d * dyo[i];
      }
    }
    y = Ys[6];
    dy = dYs[6];
  }
  *loss = L / cnt;
  *dloss = dL / cnt;
  return RDCU_OK;
}
// Real-space factors of u for the given (field, block) list: out[nf*N].
int oracle_factors(rdcu_ctx *c, const double *u, int nf, const int *field, const int *blk,
                   double *out) {
  if (!c) return fail(RDCU_EINVAL, "null ctx");
  std::vector<Factor> fl;
  for (int i = 0; i < nf; ++i) {
    if (field[i] < 0 || field[i] >= c->P.n || blk[i] < 1 || blk[i] >= c->P.nblk)
      return fail(RDCU_EINVAL, "bad factor");
    fl.push_back({field[i], blk[i]});
  }
  c->E.factors(u, fl);
  std::memcpy(out, c->E.R.data(), sizeof(double) * (size_t)nf * c->P.N);
  return RDCU_OK;
}

}  // extern "C"
