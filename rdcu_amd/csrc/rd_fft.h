// rd_fft.h -- FP64 complex arithmetic and the LDS Stockham FFT used by every
// axis pass of the pseudospectral RHS (gfx950, wave64).
//
// The operation sequence of every butterfly, twiddle product and real-split
// step is fixed (explicit fma, the file is compiled with -ffp-contract=off) so
// that results are reproducible bit for bit on any grid tiling and batch
// layout.  Twiddles come from host tables T_L[m] = exp(-2 pi i m / L).
//
// Radix plan for a length L = 5^c 3^b 2^a: c radix-5 passes, b radix-3, one
// radix-2 (a%3==1) or radix-4 (a%3==2), then a/3 radix-8 passes.
#pragma once
#ifndef __HIPCC_RTC__  // hiprtc (run-time specialised kernels) has the HIP builtins already
#include <hip/hip_runtime.h>
#endif

namespace rd {

struct __align__(16) cx {
  double re, im;
};

__device__ __forceinline__ cx cadd(cx a, cx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cx csub(cx a, cx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cx cconj(cx a) { return {a.re, -a.im}; }
__device__ __forceinline__ cx cmul(cx a, cx w) {
  return {__builtin_fma(a.re, w.re, -(a.im * w.im)), __builtin_fma(a.re, w.im, a.im * w.re)};
}
// times -i (DIR < 0, forward) or +i (DIR > 0, inverse)
template <int DIR>
__device__ __forceinline__ cx crot(cx a) {
  if constexpr (DIR < 0) return {a.im, -a.re};
  else return {-a.im, a.re};
}

constexpr double kR2 = 0x1.6a09e667f3bcdp-1;   // sqrt(1/2)
constexpr double kS3 = 0x1.bb67ae8584caap-1;   // sqrt(3)/2
constexpr double kC51 = 0x1.3c6ef372fe950p-2;  // cos(2pi/5)
constexpr double kC52 = -0x1.9e3779b97f4a8p-1; // cos(4pi/5)
constexpr double kS51 = 0x1.e6f0e134454ffp-1;  // sin(2pi/5)
constexpr double kS52 = 0x1.2cf2304755a5ep-1;  // sin(4pi/5)

template <int DIR>
__device__ __forceinline__ cx crot8_1(cx a) {
  if constexpr (DIR < 0) return {(a.re + a.im) * kR2, (a.im - a.re) * kR2};
  else return {(a.re - a.im) * kR2, (a.im + a.re) * kR2};
}
template <int DIR>
__device__ __forceinline__ cx crot8_3(cx a) {
  if constexpr (DIR < 0) return {(a.im - a.re) * kR2, -((a.re + a.im) * kR2)};
  else return {-((a.re + a.im) * kR2), (a.re - a.im) * kR2};
}

template <int DIR>
__device__ __forceinline__ void dft4(cx &x0, cx &x1, cx &x2, cx &x3) {
  cx a = cadd(x0, x2), b = csub(x0, x2);
  cx c = cadd(x1, x3), d = crot<DIR>(csub(x1, x3));
  x0 = cadd(a, c);
  x1 = cadd(b, d);
  x2 = csub(a, c);
  x3 = csub(b, d);
}

template <int R, int DIR>
__device__ __forceinline__ void dft(cx *v) {
  if constexpr (R == 2) {
    cx a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    dft4<DIR>(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 8) {
    cx e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cx o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    cx t1 = crot8_1<DIR>(o1), t2 = crot<DIR>(o2), t3 = crot8_3<DIR>(o3);
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, t1);
    v[5] = csub(e1, t1);
    v[2] = cadd(e2, t2);
    v[6] = csub(e2, t2);
    v[3] = cadd(e3, t3);
    v[7] = csub(e3, t3);
  } else if constexpr (R == 3) {
    cx t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    cx y0 = cadd(v[0], t);
    cx m = {v[0].re - 0.5 * t.re, v[0].im - 0.5 * t.im};
    cx u = crot<DIR>(cx{kS3 * d.re, kS3 * d.im});
    v[0] = y0;
    v[1] = cadd(m, u);
    v[2] = csub(m, u);
  } else if constexpr (R == 5) {
    cx t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    cx d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
    cx y0 = cadd(v[0], cadd(t1, t2));
    cx a1 = {__builtin_fma(kC52, t2.re, __builtin_fma(kC51, t1.re, v[0].re)),
             __builtin_fma(kC52, t2.im, __builtin_fma(kC51, t1.im, v[0].im))};
    cx a2 = {__builtin_fma(kC51, t2.re, __builtin_fma(kC52, t1.re, v[0].re)),
             __builtin_fma(kC51, t2.im, __builtin_fma(kC52, t1.im, v[0].im))};
    cx b1 = {__builtin_fma(kS51, d1.re, kS52 * d2.re), __builtin_fma(kS51, d1.im, kS52 * d2.im)};
    cx b2 = {__builtin_fma(kS52, d1.re, -(kS51 * d2.re)),
             __builtin_fma(kS52, d1.im, -(kS51 * d2.im))};
    cx r1 = crot<DIR>(b1), r2 = crot<DIR>(b2);
    v[0] = y0;
    v[1] = cadd(a1, r1);
    v[4] = csub(a1, r1);
    v[2] = cadd(a2, r2);
    v[3] = csub(a2, r2);
  }
}

constexpr int kMaxPass = 12;

// Radix plan of one transform length, passed by value in kernel arguments.
// The radix of pass i is (code >> 4i) & 15: a packed word rather than an
// array, so that indexing it per pass stays in scalar registers instead of
// a private-memory copy of the argument.
struct FftPlan {
  int L;
  int np;
  unsigned long long code;
  const cx *tw;  // T_L, L entries, device memory
  __host__ __device__ int rad(int i) const { return (int)((code >> (4 * i)) & 15u); }
};

// One Stockham pass over `nl` lines in LDS.  Element e of line l lives at
// buf[l*ls + e*es].  LFAST selects whether consecutive threads walk lines
// (interleaved column tiles) or butterflies (contiguous rows).
template <int R, int DIR, bool LFAST>
__device__ __forceinline__ void stockham_pass(const cx *__restrict__ src, cx *__restrict__ dst,
                                              int L, int Ns, const cx *__restrict__ tw, int nl,
                                              int ls, int es, int tid, int nt) {
  const int stride = L / R;
  const int twstep = L / (Ns * R);
  const int nb = nl * stride;
  for (int w = tid; w < nb; w += nt) {
    int l, j;
    if constexpr (LFAST) {
      l = w % nl;
      j = w / nl;
    } else {
      l = w / stride;
      j = w - l * stride;
    }
    const int k = j % Ns;
    const cx *s = src + l * ls;
    cx v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[(j + r * stride) * es];
    if (k != 0) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cx wv = tw[(k * r) * twstep];
        if constexpr (DIR > 0) wv.im = -wv.im;
        v[r] = cmul(v[r], wv);
      }
    }
    dft<R, DIR>(v);
    cx *d = dst + l * ls;
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[(base + r * Ns) * es] = v[r];
  }
}

// Full transform of nl lines held in LDS buffer a (b is scratch).  Every
// thread of the block must call it; it ends with a barrier and returns the
// buffer holding the result.
template <int DIR, bool LFAST>
__device__ cx *lds_fft(cx *a, cx *b, const FftPlan &p, int nl, int ls, int es, int tid, int nt) {
  cx *src = a, *dst = b;
  int Ns = 1;
  for (int ip = 0; ip < p.np; ++ip) {
    const int R = p.rad(ip);
    switch (R) {
      case 8: stockham_pass<8, DIR, LFAST>(src, dst, p.L, Ns, p.tw, nl, ls, es, tid, nt); break;
      case 4: stockham_pass<4, DIR, LFAST>(src, dst, p.L, Ns, p.tw, nl, ls, es, tid, nt); break;
      case 2: stockham_pass<2, DIR, LFAST>(src, dst, p.L, Ns, p.tw, nl, ls, es, tid, nt); break;
      case 3: stockham_pass<3, DIR, LFAST>(src, dst, p.L, Ns, p.tw, nl, ls, es, tid, nt); break;
      case 5: stockham_pass<5, DIR, LFAST>(src, dst, p.L, Ns, p.tw, nl, ls, es, tid, nt); break;
    }
    __syncthreads();
    Ns *= R;
    cx *t = src;
    src = dst;
    dst = t;
  }
  return src;
}

// ------------------------------------------------------------------------
// Register-resident Stockham.  A line of length L is held by T = L/E threads;
// thread t keeps elements t + T*q (q < E) in registers.  Pass p (radix R,
// Ns = product of the earlier radices) is, for butterfly j = t + T*b
// (b < E/R): inputs j + (L/R)*r = t + T*(b + (E/R)*r) -- always the thread's
// own registers -- then the same twiddle/butterfly arithmetic as
// stockham_pass, outputs to LDS at (j-k)*R + k + r*Ns.  The outputs of the
// last pass land at j + r*Ns = the thread's own slots, so the result stays in
// registers.  A radix that does not divide E (the leading radix-3 pass of a
// length 3*2^k with E = 8) runs through the line in LDS instead (lds_pass).
// LDS slot of element e in the register path: one pad slot per 8 elements,
// so the stride-R scatter of the early passes spreads over all banks.
__device__ __forceinline__ int pad8(int e) { return e + (e >> 3); }
constexpr int padded_len(int L) { return L + (L >> 3); }

template <int R, int E, int DIR>
__device__ __forceinline__ void reg_pass(cx (&v)[E], cx *buf, int es, const cx *tw, int L, int Ns,
                                         int t, int T, bool last) {
  constexpr int B = E / R;
  const int twstep = L / (Ns * R);
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int j = t + T * b;
    const int k = j % Ns;
    cx u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = v[b + B * r];
    if (k != 0) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cx w = tw[(k * r) * twstep];
        if constexpr (DIR > 0) w.im = -w.im;
        u[r] = cmul(u[r], w);
      }
    }
    dft<R, DIR>(u);
    if (last) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + B * r] = u[r];
    } else {
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[pad8(base + r * Ns) * es] = u[r];
    }
  }
}

// A pass whose radix does not divide the values per thread (the radix-3 and
// radix-5 passes that lead the plan of a length 3^b 5^c 2^a): butterfly
// j = t + T*b (b < ceil(E/R), j < L/R) reads its inputs from the line in
// LDS, and after a barrier writes its outputs back, with the twiddle and
// butterfly arithmetic of reg_pass (so the values are the same).  With
// from_regs (the transform's first pass) the thread's values t + T*q are
// stored to the line first.  Never the last pass of a plan (those end in
// radix 2, 4 or 8), so the next pass reads its values back from the line.
template <int R, int E, int DIR>
__device__ __forceinline__ void lds_pass(cx (&v)[E], cx *buf, int es, const cx *tw, int L, int Ns,
                                         int t, int T, bool from_regs) {
  constexpr int NB = (E + R - 1) / R;
  if (from_regs) {
#pragma unroll
    for (int q = 0; q < E; ++q) buf[pad8(t + T * q) * es] = v[q];
  }
  __syncthreads();
  const int nbf = L / R, twstep = L / (Ns * R);
  cx u[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + T * b;
    if (j < nbf) {
#pragma unroll
      for (int r = 0; r < R; ++r) u[b][r] = buf[pad8(j + nbf * r) * es];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + T * b;
    if (j < nbf) {
      const int k = j % Ns;
      if (k != 0) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cx w = tw[(k * r) * twstep];
          if constexpr (DIR > 0) w.im = -w.im;
          u[b][r] = cmul(u[b][r], w);
        }
      }
      dft<R, DIR>(u[b]);
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[pad8(base + r * Ns) * es] = u[b][r];
    }
  }
}

// Full transform of the line whose element e lives at buf[pad8(e)*es] (LDS
// scratch of padded_len(L) elements); input and output in v.  Every thread of the block calls it
// (it contains barriers); on return the scratch may be reused.  With ODD
// (lengths 3*2^k), radix-3 passes run through the line (lds_pass); a
// separate instantiation, so that kernels for powers of two carry none of
// its registers.
template <int E, int DIR, bool ODD = false>
__device__ void reg_fft(cx (&v)[E], cx *buf, int es, const cx *tw, const FftPlan &p, int t, int T) {
  int Ns = 1;
  for (int ip = 0; ip < p.np; ++ip) {
    const int R = p.rad(ip);
    if constexpr (ODD && E % 3 != 0) {
      if (R == 3) { lds_pass<3, E, DIR>(v, buf, es, tw, p.L, Ns, t, T, ip == 0); Ns *= R; continue; }
    }
    if (ip > 0) {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = buf[pad8(t + T * q) * es];
      __syncthreads();
    }
    const bool last = ip == p.np - 1;
    if constexpr (E % 8 == 0) {
      if (R == 8) { reg_pass<8, E, DIR>(v, buf, es, tw, p.L, Ns, t, T, last); Ns *= R; continue; }
    }
    if constexpr (E % 4 == 0) {
      if (R == 4) { reg_pass<4, E, DIR>(v, buf, es, tw, p.L, Ns, t, T, last); Ns *= R; continue; }
    }
    if constexpr (E % 3 == 0) {
      if (R == 3) { reg_pass<3, E, DIR>(v, buf, es, tw, p.L, Ns, t, T, last); Ns *= R; continue; }
    }
    if constexpr (E % 2 == 0) {
      if (R == 2) { reg_pass<2, E, DIR>(v, buf, es, tw, p.L, Ns, t, T, last); Ns *= R; continue; }
    }
    Ns *= R;  // unreachable for plans the host validated
  }
}

// ------------------------------------------------------------------------
// The same transform with the length known at compile time.  The radix plan
// (the host's order: 5s, 3s, one 2 or 4, 8s) is folded in, so every pass's
// span, twiddle stride and index arithmetic are constants and the pass loop
// unrolls; the arithmetic of each butterfly is that of reg_pass, so results
// are identical to reg_fft's.
constexpr unsigned long long static_plan_code(int L) {
  int a = 0, b = 0, c = 0, m = L;
  while (m % 5 == 0) { m /= 5; ++c; }
  while (m % 3 == 0) { m /= 3; ++b; }
  while (m % 2 == 0) { m /= 2; ++a; }
  unsigned long long code = 0;
  int np = 0;
  for (int i = 0; i < c; ++i) code |= 5ull << (4 * np++);
  for (int i = 0; i < b; ++i) code |= 3ull << (4 * np++);
  if (a % 3 == 1) code |= 2ull << (4 * np++);
  if (a % 3 == 2) code |= 4ull << (4 * np++);
  for (int i = 0; i < a / 3; ++i) code |= 8ull << (4 * np++);
  return code;
}
constexpr int static_plan_np(int L) {
  int np = 0;
  for (unsigned long long c = static_plan_code(L); c; c >>= 4) ++np;
  return np;
}
constexpr int static_plan_rad(int L, int i) { return (int)((static_plan_code(L) >> (4 * i)) & 15u); }

template <int R, int E, int DIR, int L, int NS, bool LAST>
__device__ __forceinline__ void reg_pass_s(cx (&v)[E], cx *buf, int es, const cx *tw, int t) {
  constexpr int B = E / R, T = L / E, TWS = L / (NS * R);
  static_assert(E % R == 0, "values per thread must be a multiple of every radix");
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int j = t + T * b;
    const int k = NS == 1 ? 0 : j % NS;
    cx u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = v[b + B * r];
    if (NS > 1 && k != 0) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cx w = tw[(k * r) * TWS];
        if constexpr (DIR > 0) w.im = -w.im;
        u[r] = cmul(u[r], w);
      }
    }
    dft<R, DIR>(u);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + B * r] = u[r];
    } else {
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[pad8(base + r * NS) * es] = u[r];
    }
  }
}

template <int E, int DIR, int L, int IP = 0, int NS = 1>
__device__ __forceinline__ void reg_fft_s(cx (&v)[E], cx *buf, int es, const cx *tw, int t) {
  constexpr int NP = static_plan_np(L);
  if constexpr (IP < NP) {
    constexpr int R = static_plan_rad(L, IP);
    constexpr int T = L / E;
    if constexpr (E % R != 0) {  // through the line (lds_pass); never the last pass
      static_assert(IP < NP - 1, "a plan ends in a radix that divides the values per thread");
      lds_pass<R, E, DIR>(v, buf, es, tw, L, NS, t, T, IP == 0);
    } else {
      if constexpr (IP > 0) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = buf[pad8(t + T * q) * es];
        __syncthreads();
      }
      reg_pass_s<R, E, DIR, L, NS, IP == NP - 1>(v, buf, es, tw, t);
    }
    reg_fft_s<E, DIR, L, IP + 1, NS * R>(v, buf, es, tw, t);
  }
}

// Whether every radix of the plan of L divides E (all passes in registers).
constexpr bool static_plan_in_regs(int L, int E) {
  for (unsigned long long c = static_plan_code(L); c; c >>= 4)
    if (E % (int)(c & 15u) != 0) return false;
  return true;
}

// The same compile-time transform with the twiddles in registers: every
// twiddle a thread multiplies by (passes after the first, in pass order,
// butterfly by butterfly, r = 1..R-1) is loaded from the table once, at the
// start of the kernel, together with its first data loads, instead of being
// read from LDS in every pass.  The values and the arithmetic are those of
// reg_fft_s, so the results are identical; the inverse conjugates at use.
template <int L, int IP = 0>
constexpr int static_tw_count(int E) {
  if constexpr (IP >= static_plan_np(L)) {
    return 0;
  } else {
    constexpr int R = static_plan_rad(L, IP);
    return (IP > 0 ? (E / R) * (R - 1) : 0) + static_tw_count<L, IP + 1>(E);
  }
}

template <int E, int L, int IP = 0, int NS = 1, int OFF = 0>
__device__ __forceinline__ void load_tw_s(cx *w, const cx *__restrict__ tw, int t) {
  constexpr int NP = static_plan_np(L);
  if constexpr (IP < NP) {
    constexpr int R = static_plan_rad(L, IP);
    constexpr int T = L / E, B = E / R, TWS = L / (NS * R);
    if constexpr (IP > 0) {
#pragma unroll
      for (int b = 0; b < B; ++b) {
        const int k = (t + T * b) % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) w[OFF + b * (R - 1) + r - 1] = tw[(k * r) * TWS];
      }
    }
    load_tw_s<E, L, IP + 1, NS * R, OFF + (IP > 0 ? B * (R - 1) : 0)>(w, tw, t);
  }
}

template <int R, int E, int DIR, int L, int NS, bool LAST>
__device__ __forceinline__ void reg_pass_w(cx (&v)[E], cx *buf, int es, const cx *w, int t) {
  constexpr int B = E / R, T = L / E;
  static_assert(E % R == 0, "values per thread must be a multiple of every radix");
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int j = t + T * b;
    const int k = NS == 1 ? 0 : j % NS;
    cx u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = v[b + B * r];
    if (NS > 1 && k != 0) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cx wv = w[b * (R - 1) + r - 1];
        if constexpr (DIR > 0) wv.im = -wv.im;
        u[r] = cmul(u[r], wv);
      }
    }
    dft<R, DIR>(u);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + B * r] = u[r];
    } else {
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[pad8(base + r * NS) * es] = u[r];
    }
  }
}

template <int E, int DIR, int L, int IP = 0, int NS = 1, int OFF = 0>
__device__ __forceinline__ void reg_fft_w(cx (&v)[E], cx *buf, int es, const cx *w, int t) {
  constexpr int NP = static_plan_np(L);
  if constexpr (IP < NP) {
    constexpr int R = static_plan_rad(L, IP);
    constexpr int T = L / E;
    if constexpr (IP > 0) {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = buf[pad8(t + T * q) * es];
      __syncthreads();
    }
    reg_pass_w<R, E, DIR, L, NS, IP == NP - 1>(v, buf, es, w + OFF, t);
    reg_fft_w<E, DIR, L, IP + 1, NS * R, OFF + (IP > 0 ? (E / R) * (R - 1) : 0)>(v, buf, es, w, t);
  }
}

// reg_fft with the length fixed at compile time when LS > 0 (LS must then
// equal p.L), the run-time plan otherwise: LS = 0 for powers of two, LS = -1
// for lengths 3*2^k (radix-3 passes through the line).
template <int E, int DIR, int LS>
__device__ __forceinline__ void fft_any(cx (&v)[E], cx *buf, int es, const cx *tw, const FftPlan &p,
                                        int t, int T) {
  if constexpr (LS > 0) reg_fft_s<E, DIR, LS>(v, buf, es, tw, t);
  else reg_fft<E, DIR, (LS < 0)>(v, buf, es, tw, p, t, T);
}

}  // namespace rd
