// rd_kernels_reg.h -- register-resident variants of the axis passes.
//
// Same arithmetic as rd_kernels.h operation for operation (same radix plan,
// twiddles, real-split steps, multipliers, dictionary, stage sums and error
// tree), so results are bit-identical; only the data movement differs: each
// thread keeps E elements of a line in VGPRs across the Stockham passes, LDS
// carries only the transposes between passes, twiddles are staged in LDS.
//
// k_rowi_reg also fuses the next DP45 stage: after it has produced k_s for a
// row it forms ytemp_{s+1} = y + h*sum_j a_j k_j for that row and transforms
// it, so one RHS is two launches (k_col_reg, k_rowi_reg) instead of three.
#pragma once
#include "rd_kernels.h"

namespace rd {

// Streaming stores for arrays the next kernel reads from HBM: written
// through rather than held dirty in the XCD's L2 until the end-of-kernel
// write-back.  Used by the x-forward and strided passes of large problems
// (Grid::nt_store); measured slower on problems that fit in the caches.
typedef double nt_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_nt(cx *p, cx v) {
  __builtin_nontemporal_store(nt_d2{v.re, v.im}, reinterpret_cast<nt_d2 *>(p));
}
__device__ __forceinline__ void st_nt(double2 *p, double2 v) {
  __builtin_nontemporal_store(nt_d2{v.x, v.y}, reinterpret_cast<nt_d2 *>(p));
}
__device__ __forceinline__ void st_nt(double *p, double v) { __builtin_nontemporal_store(v, p); }
// streaming (NT) or ordinary store, chosen at compile time
template <bool NT, class T>
__device__ __forceinline__ void st_g(T *p, T v) {
  if constexpr (NT) st_nt(p, v);
  else *p = v;
}

// Field f of the dictionary at XM points x0 + stride*i of one row (value rows
// in LDS, vals[slot*Nx + x]): terms of field f in input order, each a product
// of integer powers started from 1, accumulated with fma from 0 -- the
// per-field sequence of operations of rdcu.cu:175-183.
template <int XM>
__device__ __forceinline__ void dict_field(const int *P, int f, const double *vals, int Nx, int x0,
                                           int stride, double (&acc)[XM]) {
#pragma unroll
  for (int i = 0; i < XM; ++i) acc[i] = 0.0;
  int p = P[f];
  const int pe = P[f + 1];
  while (p < pe) {
    const int nfac = P[p];
    const double cc = __hiloint2double(P[p + 2], P[p + 1]);
    p += 3;
    double term[XM];
#pragma unroll
    for (int i = 0; i < XM; ++i) term[i] = 1.0;
    for (int j = 0; j < nfac; ++j, p += 2) {
      const double *vr = vals + P[p] * Nx;
      const int pw = P[p + 1];
#pragma unroll
      for (int i = 0; i < XM; ++i) {
        const int x = x0 + stride * i;
        const double v = x < Nx ? vr[x] : 0.0;
        term[i] = term[i] * ipow(v, pw);
      }
    }
#pragma unroll
    for (int i = 0; i < XM; ++i) acc[i] = __builtin_fma(cc, term[i], acc[i]);
  }
}

// Dictionary evaluator reading the program from LDS (any dictionary).  The
// runtime can replace it by straight-line code generated for one dictionary
// (same operations in the same order).
struct ProgDict {
  static constexpr bool kProg = true;
  template <int XM>
  __device__ __forceinline__ static void eval(const int *P, int f, const double *vals, int Nx,
                                              int x0, int stride, double (&acc)[XM]) {
    dict_field<XM>(P, f, vals, Nx, x0, stride, acc);
  }
};

// Copies n complex values global -> LDS with U loads in flight per thread
// before the first store (a plain strided loop waits out one memory latency
// per element).
template <int U>
__device__ __forceinline__ void stage_cx(cx *dst, const cx *__restrict__ src, int n, int tid,
                                         int nt) {
  for (int i0 = tid; i0 < n; i0 += U * nt) {
    cx tmp[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * nt;
      tmp[u] = i < n ? src[i] : cx{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * nt;
      if (i < n) dst[i] = tmp[u];
    }
  }
}

template <int U, class V>
__device__ __forceinline__ void stage_v(V *dst, const V *__restrict__ src, int n, int tid, int nt) {
  for (int i0 = tid; i0 < n; i0 += U * nt) {
    V tmp[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * nt;
      tmp[u] = i < n ? src[i] : V{};
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * nt;
      if (i < n) dst[i] = tmp[u];
    }
  }
}

// Loads up to U*nt values into registers now and stores them to LDS later,
// so the loads share one memory round trip with the data loads issued in
// between (values past U*nt use a plain loop at store time).
template <int U, class V>
struct Prefetch {
  V r[U];
  __device__ __forceinline__ void load(const V *__restrict__ src, int n, int tid, int nt) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = tid + u * nt;
      r[u] = i < n ? src[i] : V{};
    }
  }
  __device__ __forceinline__ void store(V *dst, const V *__restrict__ src, int n, int tid, int nt) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = tid + u * nt;
      if (i < n) dst[i] = r[u];
    }
    for (int i = tid + U * nt; i < n; i += nt) dst[i] = src[i];
  }
};

// The wavenumber of a strided line's element e: axis 0 (z, or y in 2-D), or
// with ymul the 3-D y axis (constant indices: nk and kv stay in registers).
__device__ __forceinline__ void line_k(bool ymul, int e, const double *ks, int *nk, double *kv) {
  const double k = ks[e];
  if (ymul) {
    nk[1] = e;
    kv[1] = k;
  } else {
    nk[0] = e;
    kv[0] = k;
  }
}

// Multiplier value of derivative block blk at wavenumber indices nk with the
// wavenumbers of each axis already loaded (kv), scaled by sc -- the same
// operations as apply_mult.
__device__ __forceinline__ cx mult_k(const Grid &g, int blk, const int *nk, const double *kv, cx X,
                                     double sc) {
  const int nd = g.ndim;
  if (blk <= nd) {
    const int a = blk - 1;
    if (nk[a] == g.Ns[a] / 2) return {0.0, 0.0};
    const double val = kv[a] * sc;
    return {-(X.im * val), X.re * val};
  }
  const int q = blk - 1 - nd;
  const int a = q / nd, b = q - a * nd;
  double val;
  if (a == b) {
    val = (kv[a] * kv[a]) * sc;
  } else {
    const bool na = nk[a] == g.Ns[a] / 2, nb = nk[b] == g.Ns[b] / 2;
    if (na != nb) return {0.0, 0.0};
    val = (kv[a] * kv[b]) * sc;
  }
  return {-(X.re * val), -(X.im * val)};
}

// Forward x pass of rows (plain input or DP45 stage sum) -> S.
// Block: rpb rows x T threads (T = L/E).  LDS: twiddles L + rpb*padded_len(L)
// complex.
template <int E, bool NT = false, int LS = 0>
__device__ __forceinline__ void rowf_body(long bid, const Grid &g, const FftPlan &px,
                                          const cx *__restrict__ twN, DevState *st, const Bufs &b,
                                          int src_kind, const double *src_ptr, const StageCoef &sc,
                                          int rpb) {
  extern __shared__ cx smem[];
  const int L = LS > 0 ? LS : g.L, Nx = g.Nx, T = L / E;
  const long nrows = (long)g.n * g.rows;
  cx *tw = smem, *buf = smem + L;
  const int tid = threadIdx.x;
  const int rr = tid / T, t = tid - rr * T;
  const long row = bid * rpb + rr;
  const bool act = row < nrows;
  Prefetch<4, cx> pf_tw;
  pf_tw.load(px.tw, L, tid, blockDim.x);

  cx v[E];
  if (act) {
    if (src_kind == SRC_STAGE) {
      const double h = st->h;
      const double *y = b.y[st->iy];
      double *yt = b.y[st->iyt];
      const double *ks[6];
      ks[0] = b.k[st->k1s];
      for (int j = 1; j < 6; ++j) ks[j] = b.k[j];
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const long off = row * Nx + 2 * (t + T * q);
        const double2 yv = *reinterpret_cast<const double2 *>(y + off);
        const double2 k1 = *reinterpret_cast<const double2 *>(ks[0] + off);
        double ar = sc.a[0] * k1.x, ai = sc.a[0] * k1.y;
        for (int s = 1; s < sc.ns; ++s) {
          if (sc.a[s] == 0.0) continue;
          const double2 kv = *reinterpret_cast<const double2 *>(ks[s] + off);
          ar = __builtin_fma(sc.a[s], kv.x, ar);
          ai = __builtin_fma(sc.a[s], kv.y, ai);
        }
        double2 o;
        o.x = __builtin_fma(h, ar, yv.x);
        o.y = __builtin_fma(h, ai, yv.y);
        st_g<NT>(reinterpret_cast<double2 *>(yt + off), o);
        v[q] = cx{o.x, o.y};
      }
    } else {
      const double *src =
          src_kind == SRC_Y ? b.y[st->iy] : (src_kind == SRC_YT ? b.y[st->iyt] : src_ptr);
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const double2 o = *reinterpret_cast<const double2 *>(src + row * Nx + 2 * (t + T * q));
        v[q] = cx{o.x, o.y};
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = cx{0.0, 0.0};
  }
  pf_tw.store(tw, px.tw, L, tid, blockDim.x);
  cx *lb = buf + rr * padded_len(L);
  fft_any<E, -1, LS>(v, lb, 1, tw, px, t, T);
#pragma unroll
  for (int q = 0; q < E; ++q) lb[t + T * q] = v[q];
  __syncthreads();
  if (!act) return;
  cx *S = b.S + row * g.Nxp;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int k = t + T * q;
    const cx Av = v[q];
    const cx Bv = cconj(lb[k == 0 ? 0 : L - k]);
    const cx Ev = cadd(Av, Bv), D = csub(Av, Bv);
    const cx tt = cmul(D, twN[k]);
    st_g<NT>(S + k, cx{0.5 * (Ev.re + tt.im), 0.5 * (Ev.im - tt.re)});
  }
  if (t == 0) {
    const cx Av = lb[0];
    const cx Bv = cconj(lb[0]);
    const cx Ev = cadd(Av, Bv), D = csub(Av, Bv);
    const cx tt = cmul(D, twN[L]);
    st_g<NT>(S + L, cx{0.5 * (Ev.re + tt.im), 0.5 * (Ev.im - tt.re)});
  }
}

template <int E, bool NT, int LS = 0>
__global__ __launch_bounds__(256) void k_rowf_reg(Grid g, FftPlan px, const cx *__restrict__ twN,
                                                  DevState *st, int check_active, Bufs b,
                                                  int src_kind, const double *src_ptr,
                                                  StageCoef sc, int rpb) {
  if (check_active && !attempt_active(st)) return;
  rowf_body<E, NT, LS>(blockIdx.x, g, px, twN, st, b, src_kind, src_ptr, sc, rpb);
}

// Strided-line pass (modes as k_col).  Block: tc columns x T threads; LDS:
// twiddles L + tc*padded_len(L) complex (element e of column c at
// buf[pad8(e)*tc + c]) + tc*L complex holding the forward spectrum (mode 1).
// TG: twiddles and axis-0 wavenumbers read from global memory instead of LDS
// copies (long lines, where the copies would leave room for one column).
// FO: built for modes 0 and 2 only (the transforms of the 3-D y passes): the
// multiply path and its registers are compiled out, so that more blocks fit a
// CU (the runtime launches it for the y forward pass).
// KR: a block with several factors keeps the forward spectrum in registers
// instead of LDS (built for the 3-D y pass with the factors without k_z, 8
// values per thread: less LDS per block, more blocks per CU).
template <int E, bool NT = false, int LS = 0, bool TG = false, bool FO = false, bool KR = false>
__device__ __forceinline__ void col_body(int lblk, const Grid &g, const FftPlan &p,
                                         const ColGeom &cg, const Bufs &b, const FacList &fl,
                                         int mode_arg, int fpw, const DevState *chk = nullptr) {
  extern __shared__ cx smem[];
  const int mode = FO ? (mode_arg == 0 ? 0 : 2) : mode_arg;
  // mode 1: blocks come in groups of nparts, each transforming up to fpw of
  // the field's factors (the forward transform is repeated per part)
  const int nparts = mode == 1 ? cg.nparts : 1;
  const int part = lblk % nparts;
  const int tile = lblk / nparts;
  const int grp = tile / cg.ntc, ct = tile - grp * cg.ntc;
  const int g1 = grp / cg.gdiv, g2 = grp - g1 * cg.gdiv;
  // a part past the field's factors has no work; like an attempt that no
  // longer runs (chk), it returns after the column loads are issued, so
  // that their round trip overlaps the loads the decision waits for
  int m0 = 0, m1 = 0;
  if (mode == 1) {
    m0 = fl.fstart[g1] + part * fpw;
    m1 = fl.fstart[g1 + 1];
    if (m1 > m0 + fpw) m1 = m0 + fpw;
  } else if (mode == 3) {  // every factor of the field (those without k_z)
    m0 = fl.fstart[g1];
    m1 = fl.fstart[g1 + 1];
  }
  // the line axis of the multiply: z (axis 0) in mode 1, y in mode 3, whose
  // factors take the scale without the z round trip
  const bool ymul = !FO && mode == 3;
  const double msc = ymul ? g.invNxy : g.invN;
  const double *ktl = ymul ? g.kt[1] : g.kt[0];
  const bool idle = mode == 1 && m0 >= m1;
  const int L = LS > 0 ? LS : cg.L, T = L / E, tc = cg.tc;
  // tiles of tc columns from column 0 (whole 16-byte-column groups per
  // cache line; the last tile holds the remainder), or with bal the Nxh
  // columns of a group split into ntc tiles whose widths differ by at most
  // one (tc >= the widest), so that a launch can be sized to the CU count:
  // at 512^2 the 257 columns of a field fill 64 tiles of 4 or 5 instead of
  // 64 tiles of 4 plus a 65th of one column
  const int c0 = cg.bal ? (int)((long)ct * g.Nxh / cg.ntc) : ct * tc;
  const int nc = cg.bal ? (int)((long)(ct + 1) * g.Nxh / cg.ntc) - c0
                        : ((g.Nxh - c0) < tc ? (g.Nxh - c0) : tc);
  const int tid = threadIdx.x;
  const int c = tid % tc, t = tid / tc;
  const bool act = c < nc;
  cx *twl = smem, *buf = TG ? smem : smem + L;
  const cx *tw = TG ? p.tw : twl;
  // mode 1: the forward spectrum is kept in LDS when the block has several
  // factors; the wavenumbers of axis 0 follow it
  const bool keep = !KR && m1 - m0 > 1;
  double *k0l = reinterpret_cast<double *>(buf + tc * padded_len(L) + (keep ? tc * L : 0));
  const double *k0s = TG ? ktl : k0l;
  const long lbase = (long)g2 * cg.gs2 + c0 + c;
  cx *src = (mode == 2 ? b.F + (long)fac_slot(fl, g1) * cg.gs1 : b.S + (long)g1 * cg.gs1) + lbase;
  cx v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) v[q] = act ? src[(long)(t + T * q) * cg.es] : cx{0.0, 0.0};
  // everything else the block reads from global memory is issued now, in the
  // same round trip as the column data
  const int nd = g.ndim;
  const int kx = c0 + (act ? c : 0);
  double kv[3];
  int nk[3];
  int blk0 = 0;
  if (mode == 1) {
    if (nd == 2) {
      nk[1] = kx;
      kv[1] = g.kt[1][kx];
    } else {
      nk[1] = g2;
      kv[1] = g.kt[1][g2];
      nk[2] = kx;
      kv[2] = g.kt[2][kx];
    }
    blk0 = idle ? 0 : fl.fblk[m0];
  } else if (ymul) {
    nk[0] = 0;
    kv[0] = 0.0;
    nk[2] = kx;
    kv[2] = g.kt[2][kx];
    blk0 = m0 < m1 ? fl.fblk[m0] : 0;
  }
  Prefetch<4, cx> pf_tw;
  Prefetch<4, double> pf_k;
  if constexpr (!TG) {
    pf_tw.load(p.tw, L, tid, blockDim.x);
    if (mode == 1 || ymul) pf_k.load(ktl, L, tid, blockDim.x);
  }
  if (idle || (chk && !attempt_active(chk))) return;
  if constexpr (!TG) {
    pf_tw.store(twl, p.tw, L, tid, blockDim.x);
    if (mode == 1 || ymul) pf_k.store(k0l, ktl, L, tid, blockDim.x);
  }
  cx *lb = buf + c;
  if (FO || (mode != 1 && mode != 3)) {
    if (mode == 0) fft_any<E, -1, LS>(v, lb, tc, tw, p, t, T);
    else fft_any<E, +1, LS>(v, lb, tc, tw, p, t, T);
    if (act) {
#pragma unroll
      for (int q = 0; q < E; ++q) st_g<NT>(src + (long)(t + T * q) * cg.es, v[q]);
    }
    return;
  }
  fft_any<E, -1, LS>(v, lb, tc, tw, p, t, T);
  if (ymul) {  // the forward spectra stay in S for the z pass
    if (act) {
#pragma unroll
      for (int q = 0; q < E; ++q) st_g<NT>(src + (long)(t + T * q) * cg.es, v[q]);
    }
    if (m0 >= m1) return;
  }
  if constexpr (KR) {  // each factor's multiply from the spectrum in registers
    for (int m = m0; m < m1; ++m) {
      const int blk = m == m0 ? blk0 : fl.fblk[m];
      cx w[E];
#pragma unroll
      for (int q = 0; q < E; ++q) {
        line_k(ymul, t + T * q, k0s, nk, kv);
        w[q] = mult_k(g, blk, nk, kv, v[q], msc);
      }
      fft_any<E, +1, LS>(w, lb, tc, tw, p, t, T);
      if (act) {
        cx *dst = b.F + (long)fac_slot(fl, m) * g.Nh + lbase;
#pragma unroll
        for (int q = 0; q < E; ++q) st_g<NT>(dst + (long)(t + T * q) * cg.es, w[q]);
      }
    }
    return;
  }
  if (!keep) {  // one factor: the spectrum goes from registers to its multiply
    cx w[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      line_k(ymul, t + T * q, k0s, nk, kv);
      w[q] = mult_k(g, blk0, nk, kv, v[q], msc);
    }
    fft_any<E, +1, LS>(w, lb, tc, tw, p, t, T);
    if (act) {
      cx *dst = b.F + (long)fac_slot(fl, m0) * g.Nh + lbase;
#pragma unroll
      for (int q = 0; q < E; ++q) st_g<NT>(dst + (long)(t + T * q) * cg.es, w[q]);
    }
    return;
  }
  // keep the forward spectrum in LDS (each thread reloads only its own
  // elements, so no barrier is needed) instead of registers across factors
  cx *xs = buf + tc * padded_len(L);
#pragma unroll
  for (int q = 0; q < E; ++q) xs[(t + T * q) * tc + c] = v[q];
  for (int m = m0; m < m1; ++m) {
    const int blk = m == m0 ? blk0 : fl.fblk[m];
    cx w[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      line_k(ymul, t + T * q, k0s, nk, kv);
      w[q] = mult_k(g, blk, nk, kv, xs[(t + T * q) * tc + c], msc);
    }
    fft_any<E, +1, LS>(w, lb, tc, tw, p, t, T);
    if (act) {
      cx *dst = b.F + (long)fac_slot(fl, m) * g.Nh + lbase;
#pragma unroll
      for (int q = 0; q < E; ++q) st_g<NT>(dst + (long)(t + T * q) * cg.es, w[q]);
    }
  }
}

// Up to 512 threads per block with 8 values per thread on power-of-two
// lines (204 registers fit two waves per SIMD); 16 values per thread, and 3*2^k
// lines (their radix-3 pass through LDS), need more than 256 registers, so
// their blocks stay at 256 threads and nothing spills.
template <int E, int LS = 0>
constexpr int col_max_threads() { return E == 8 && (LS == 0 || (LS > 0 && LS % 3 != 0)) ? 512 : 256; }

template <int E, bool NT, int LS = 0, bool TG = false, bool FO = false, bool KR = false>
__global__ __launch_bounds__((col_max_threads<E, LS>())) void k_col_reg(Grid g, FftPlan p, ColGeom cg, DevState *st,
                                                 int check_active, Bufs b, FacList fl, int mode,
                                                 int fpw) {
  const int lblk = xcd_block(blockIdx.x, gridDim.x);
  if (lblk >= cg.nblk) return;
  col_body<E, NT, LS, TG, FO, KR>(lblk, g, p, cg, b, fl, mode, fpw, check_active ? st : nullptr);
}

// Inverse x pass for one row position r, all fields: factor lines in groups
// of G (thread (gl, t) owns line g0+gl), dictionary, stage-7 error tree, and
// with nx.ns > 0 the fused next stage (stage sum + forward x pass -> S).
// Block: G*T threads, G >= n.  LDS: twiddles L + G*padded_len(L) complex,
// then (n+nf)*Nx + n*Nx doubles, then the dictionary program.
// Pre-processing of the half-complex -> real transform of factor line m at
// row position r, elements t + T*q: the split step of the real inverse
// transform (and in 1-D the multiplier, the factor spectra being S itself).
// Problem sizes of a kernel compiled for one problem (the generated
// x-inverse kernels): nonzero members (nf: >= 0) are compile-time constants
// equal to the launch's Grid and FacList values; RunGeom reads those instead.
struct RunGeom {
  static constexpr int n = 0, ndim = 0, Nx = 0, Nxp = 0, nf = -1;
  static constexpr long N = 0, rows = 0;
};

template <int E, int LS = 0, class GK = RunGeom>
__device__ __forceinline__ void c2r_pre(const Grid &g, const FacList &fl, const Bufs &b,
                                        const cx *__restrict__ twN, int m, long r, int t, int T,
                                        cx (&v)[E]) {
  const int L = LS > 0 ? LS : g.L;
  const long rows = GK::rows ? GK::rows : g.rows;
  const int Nxp = GK::Nxp ? GK::Nxp : g.Nxp;
  const long Nh = rows * Nxp;
  // the factor tables are read only in 1-D (a per-thread load whose result
  // the row address would otherwise wait for)
  const bool d1 = (GK::ndim ? GK::ndim : g.ndim) == 1;
  const int blk = d1 ? fl.fblk[m] : 0;
  const cx *row = d1 ? b.S + (long)fl.ffield[m] * Nh : b.F + (long)m * Nh + r * Nxp;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int k = t + T * q;
    cx Av = row[k], Bv = row[L - k];
    if (d1) {
      int nk0 = k, nk1 = L - k;
      Av = apply_mult(g, blk, &nk0, Av, g.invN);
      Bv = apply_mult(g, blk, &nk1, Bv, g.invN);
    }
    Bv = cconj(Bv);
    if (k == 0) {
      Av.im = 0.0;
      Bv.im = 0.0;
    }
    const cx Ev = cadd(Av, Bv), D = csub(Av, Bv);
    const cx tt = cmul(D, cconj(twN[k]));
    v[q] = cx{Ev.re - tt.im, Ev.im + tt.re};
  }
}

// LDS of one row of the x-inverse pass (complex units).  The value rows
// (n identity rows, then nf factor rows, Nx doubles each) come first.  The
// factors are transformed in groups of G lines, the first group holding the
// remainder nf - (ng-1)*G, so that the last group is a full one.  The FFT
// lines start where the rows of the last group begin: those rows are written
// only after the group's transform, when the lines are free.  After them come
// the n stage-result (or error-square) rows, then (for the interpreted
// dictionary) the program.  The twiddles are read from global memory
// (cached).  At 256^3 (L = 128, n = 2, nf = 6, G = 4) a generated kernel's
// row takes 20 KB: 8 rows per CU.
struct RowiLayout {
  long early;  // complex offset of the FFT lines: rows written before the last transform
  long late;   // complex size of the rows written after it (inside the lines' span)
  long tail;   // complex size from the lines' start to the program
};
__host__ __device__ inline RowiLayout rowi_layout(int L, int Nx, int n, int nf, int G) {
  const int ng = nf > 0 ? (nf + G - 1) / G : 0;
  const int nearly = n + (ng > 1 ? nf - G : 0);
  RowiLayout o;
  o.early = (long)nearly * Nx / 2;  // Nx is even
  o.late = (long)(n + nf - nearly) * Nx / 2;
  const long lines = (long)G * padded_len(L), rest = o.late + (long)n * Nx / 2;
  o.tail = lines > rest ? lines : rest;
  return o;
}
__host__ __device__ inline long rowi_region(int L, int Nx, int n, int nf, int nwords, int G,
                                            bool prog = true) {
  const RowiLayout o = rowi_layout(L, Nx, n, nf, G);
  return o.early + o.tail + (prog ? (nwords + 3) / 4 : 0);
}

// One grid row r, by the nt threads tid = 0..nt-1 of a row group whose LDS
// region starts at smem.  Every row group of a block runs the same barrier
// sequence.
// LEAN: fewer points and stage-sum items per thread in flight (half the
// registers of those phases), for kernels compiled to keep two wavefronts per
// SIMD (where a CU holds many rows, 3-D).
template <int E, class DE, int LS = 0, class GK = RunGeom, bool TWR = false, bool LEAN = false>
__device__ __forceinline__ void rowi_body(long r, const Grid &g, const FftPlan &px,
                                          const cx *__restrict__ twN, DevState *st, const Bufs &b,
                                          const FacList &fl, const DictP &dict, int src_kind,
                                          const double *src_ptr, int dst_kind, int dst_slot,
                                          double *dst_ptr, int err, const ErrCoef &ec,
                                          const StageCoef &nx, int G, cx *smem, int tid, int nt,
                                          const DevState *chk = nullptr) {
  const int n = GK::n ? GK::n : g.n, Nx = GK::Nx ? GK::Nx : g.Nx;
  const int L = LS > 0 ? LS : g.L, T = L / E;
  const long N = GK::N ? GK::N : g.N, rows = GK::rows ? GK::rows : g.rows;
  const int Nxp = GK::Nxp ? GK::Nxp : g.Nxp;
  const int gl = tid / T, t = tid - gl * T;
  const int nf = GK::nf >= 0 ? GK::nf : fl.nf;
  const int PL = padded_len(L);
  const cx *tw = px.tw;
  const RowiLayout lay = rowi_layout(L, Nx, n, nf, G);
  double *vals = reinterpret_cast<double *>(smem);
  cx *buf = smem + lay.early;
  // the stage results for the fused next stage follow the value rows, in the
  // FFT lines' span (free between the factor transforms and the next stage's
  // transform; the error pass, which uses them for the squares, has no next
  // stage)
  double *fv = reinterpret_cast<double *>(buf + lay.late);
  int *prog = reinterpret_cast<int *>(buf + lay.tail);
  cx *lb = buf + gl * PL;
  const double *src =
      src_kind == SRC_Y ? b.y[st->iy] : (src_kind == SRC_YT ? b.y[st->iyt] : src_ptr);
  // Prologue: every global load is issued before the first LDS store or
  // use, so they share one memory round trip: for E = 8 the first factor
  // group's pre-processed spectrum (its addresses need no other load;
  // computed by every line group, without a divergent branch, and kept by
  // those with a factor), the program and the identity rows.
  constexpr bool kPre = E <= 8;  // register budget: only for 8 values per thread
  // factor groups (rowi_layout): the first holds rem factors, the others G
  const int ng = nf > 0 ? (nf + G - 1) / G : 0;
  const int rem = nf - (ng > 0 ? (ng - 1) * G : 0);
  cx v0[kPre ? E : 1];
  if constexpr (kPre) {
    if (nf > 0) c2r_pre<E, LS, GK>(g, fl, b, twN, gl < rem ? gl : 0, r, t, T, v0);
  }
  // With TWR (a compile-time line length, 8 values per thread) the twiddles
  // of the passes are held in registers, loaded here in the same round trip
  // (otherwise each pass reads them from the table in global memory; the
  // inverse transforms conjugate at use).  The registers pay off where rows
  // are many (3-D), not at 512^2: the tuner picks.
  constexpr bool kTwR = TWR && LS > 0 && E == 8 && static_plan_in_regs(LS > 0 ? LS : 8, E);
  cx twr[kTwR ? static_tw_count<(LS > 0 ? LS : 8)>(E) : 1];
  if constexpr (kTwR) load_tw_s<E, (LS > 0 ? LS : 8)>(twr, tw, t);
  auto fft_c2r = [&](cx(&x)[E]) {
    if constexpr (kTwR) reg_fft_w<E, +1, (LS > 0 ? LS : 8)>(x, lb, 1, twr, t);
    else fft_any<E, +1, LS>(x, lb, 1, tw, px, t, T);
  };
  auto fft_r2c = [&](cx(&x)[E]) {
    if constexpr (kTwR) reg_fft_w<E, -1, (LS > 0 ? LS : 8)>(x, lb, 1, twr, t);
    else fft_any<E, -1, LS>(x, lb, 1, tw, px, t, T);
  };
  Prefetch<4, int> pf_p;
  if constexpr (DE::kProg) pf_p.load(dict.prog, dict.nwords, tid, nt);
  constexpr int IDU = 8;
  double idr[IDU];
#pragma unroll
  for (int u = 0; u < IDU; ++u) {
    const int w = tid + u * nt;
    idr[u] = w < n * Nx ? src[(w / Nx) * N + r * Nx + (w % Nx)] : 0.0;
  }
  // an attempt that no longer runs (chk) returns here, after the loads
  if (chk && !attempt_active(chk)) return;
  if constexpr (DE::kProg) pf_p.store(prog, dict.prog, dict.nwords, tid, nt);
#pragma unroll
  for (int u = 0; u < IDU; ++u) {
    const int w = tid + u * nt;
    if (w < n * Nx) vals[w] = idr[u];
  }
  for (int w = tid + IDU * nt; w < n * Nx; w += nt) vals[w] = src[(w / Nx) * N + r * Nx + (w % Nx)];

  for (int g0 = 0, gsz = rem; g0 < nf; g0 += gsz, gsz = G) {
    const int m = g0 + gl;
    const bool act = gl < gsz;
    cx v[E];
    if (act && kPre && g0 == 0) {
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = v0[kPre ? q : 0];
    } else if (act) {
      c2r_pre<E, LS, GK>(g, fl, b, twN, m, r, t, T, v);
    } else {
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = cx{0.0, 0.0};
    }
    fft_c2r(v);
    if (act) {
      double *vr = vals + (long)(n + m) * Nx;
#pragma unroll
      for (int q = 0; q < E; ++q)
        *reinterpret_cast<double2 *>(vr + 2 * (t + T * q)) = double2{v[q].re, v[q].im};
    }
  }
  __syncthreads();

  double *out;
  if (dst_kind == DST_PTR) out = dst_ptr;
  else if (dst_kind == DST_K1) out = b.k[st->k1s];
  else if (dst_kind == DST_K7) out = b.k[st->k7s];
  else out = b.k[dst_slot];
  double *sq = fv;  // n*Nx doubles
  const double *yv = nullptr, *k1 = nullptr;
  double h = 0, atol = 0, rtol = 0;
  if (err) {
    yv = b.y[st->iy];
    k1 = b.k[st->k1s];
    h = st->h;
    atol = st->atol;
    rtol = st->rtol;
  }
  constexpr int XM = E <= 8 && !LEAN ? 4 : 2;  // points per thread per pass (register budget)
  for (int xb = 0; xb < Nx; xb += XM * nt) {
    const int x0 = xb + tid;
    for (int f = 0; f < n; ++f) {
      double acc[XM];
      double ev[XM][6];
      if (err) {  // error inputs for the XM points, loaded before the dictionary
#pragma unroll
        for (int i = 0; i < XM; ++i) {
          const int x = x0 + nt * i < Nx ? x0 + nt * i : Nx - 1;
          const long idx = f * N + r * Nx + x;
          ev[i][0] = k1[idx];
          ev[i][1] = b.k[2][idx];
          ev[i][2] = b.k[3][idx];
          ev[i][3] = b.k[4][idx];
          ev[i][4] = b.k[5][idx];
          ev[i][5] = yv[idx];
        }
      }
      DE::template eval<XM>(DE::kProg ? prog : dict.prog, f, vals, Nx, x0, nt, acc);
#pragma unroll
      for (int i = 0; i < XM; ++i) {
        const int x = x0 + nt * i;
        if (x >= Nx) continue;
        const long idx = f * N + r * Nx + x;
        out[idx] = acc[i];
        if (nx.ns) fv[f * Nx + x] = acc[i];
        if (err) {
          double e = ec.e[0] * ev[i][0];
          e = __builtin_fma(ec.e[2], ev[i][1], e);
          e = __builtin_fma(ec.e[3], ev[i][2], e);
          e = __builtin_fma(ec.e[4], ev[i][3], e);
          e = __builtin_fma(ec.e[5], ev[i][4], e);
          e = __builtin_fma(ec.e[6], acc[i], e);
          const double wgt =
              __builtin_fma(rtol, fmax(fabs(ev[i][5]), fabs(vals[f * Nx + x])), atol);
          const double qv = (h * e) / wgt;
          sq[f * Nx + x] = qv * qv;
        }
      }
    }
  }
  if (err) {
    __syncthreads();
    for (int s = 1; s < Nx; s <<= 1) {
      const int per = (Nx + 2 * s - 1) / (2 * s);
      for (int w = tid; w < n * per; w += nt) {
        const int f = w / per, i = w - f * per;
        const int x = i * 2 * s;
        if (x + s < Nx) sq[f * Nx + x] = sq[f * Nx + x] + sq[f * Nx + x + s];
      }
      __syncthreads();
    }
    if (tid < n) b.part[(long)tid * rows + r] = sq[tid * Nx];
  }
  if (nx.ns == 0) return;

  // fused next stage for this row: ytemp = y + h*sum a_j k_j (k_s from LDS),
  // spread over all threads of the row (two values each step), staged in LDS
  // (the value rows are free now) for the x forward pass of the n fields.
  __syncthreads();
  double *ys = vals;
  {
    const double hs = st->h;
    const double *y = b.y[st->iy];
    double *yt = b.y[st->iyt];
    const double *ks[6];
    ks[0] = b.k[st->k1s];
    for (int j = 1; j < 6; ++j) ks[j] = b.k[j];
    const int half = Nx / 2;
    // items in flight per thread: 4 (kernels with more than 8 values per
    // thread have no registers left for more than one)
    constexpr int AXU = E <= 8 && !LEAN ? 4 : 1;
#pragma unroll AXU
    for (int w = tid; w < n * half; w += nt) {
      const int f = w / half, j2 = 2 * (w - f * half);
      const long off = f * N + r * Nx + j2;
      const double2 yv2 = *reinterpret_cast<const double2 *>(y + off);
      const double2 a1 = *reinterpret_cast<const double2 *>(ks[0] + off);
      // loads issued unconditionally (absent terms read k1 again and are
      // discarded), so they share one round trip
      double2 kk[5];
#pragma unroll
      for (int s = 1; s < 6; ++s) {
        const bool use = s < nx.ns - 1 && nx.a[s] != 0.0;
        kk[s - 1] = *reinterpret_cast<const double2 *>((use ? ks[s] : ks[0]) + off);
      }
      double ar = nx.a[0] * a1.x, ai = nx.a[0] * a1.y;
#pragma unroll
      for (int s = 1; s < 6; ++s) {  // unrolled: ks[] stays in registers
        const double as = s < nx.ns ? nx.a[s] : 0.0;
        double kx_ = kk[s - 1].x, ky_ = kk[s - 1].y;
        if (s == nx.ns - 1) {  // the stage just computed
          kx_ = fv[f * Nx + j2];
          ky_ = fv[f * Nx + j2 + 1];
        }
        const double nr = __builtin_fma(as, kx_, ar);
        const double ni = __builtin_fma(as, ky_, ai);
        ar = as != 0.0 ? nr : ar;
        ai = as != 0.0 ? ni : ai;
      }
      double2 o;
      o.x = __builtin_fma(hs, ar, yv2.x);
      o.y = __builtin_fma(hs, ai, yv2.y);
      *reinterpret_cast<double2 *>(yt + off) = o;
      *reinterpret_cast<double2 *>(ys + f * Nx + j2) = o;
    }
  }
  __syncthreads();
  const bool act = gl < n;
  const int f = act ? gl : 0;
  cx v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const double2 o = *reinterpret_cast<const double2 *>(ys + f * Nx + 2 * (t + T * q));
    v[q] = act ? cx{o.x, o.y} : cx{0.0, 0.0};
  }
  fft_r2c(v);
#pragma unroll
  for (int q = 0; q < E; ++q) lb[t + T * q] = v[q];
  __syncthreads();
  if (!act) return;
  cx *S = b.S + ((long)f * rows + r) * Nxp;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int k = t + T * q;
    const cx Av = v[q];
    const cx Bv = cconj(lb[k == 0 ? 0 : L - k]);
    const cx Ev = cadd(Av, Bv), D = csub(Av, Bv);
    const cx tt = cmul(D, twN[k]);
    S[k] = cx{0.5 * (Ev.re + tt.im), 0.5 * (Ev.im - tt.re)};
  }
  if (t == 0) {
    const cx Av = lb[0];
    const cx Bv = cconj(lb[0]);
    const cx Ev = cadd(Av, Bv), D = csub(Av, Bv);
    const cx tt = cmul(D, twN[L]);
    S[L] = cx{0.5 * (Ev.re + tt.im), 0.5 * (Ev.im - tt.re)};
  }
}

// Block = R rows x G*T threads (R = blockDim / (G*T)), row r = blockIdx*R +
// sub.  With more than 8 values per thread the per-row thread range and LDS
// offset would cost registers such kernels do not have, so they take R = 1.
template <int E, class DE, int LS = 0, class GK = RunGeom, bool TWR = false, bool LEAN = false>
__device__ __forceinline__ void rowi_rows(const Grid &g, const FftPlan &px,
                                          const cx *__restrict__ twN, DevState *st, const Bufs &b,
                                          const FacList &fl, const DictP &dict, int src_kind,
                                          const double *src_ptr, int dst_kind, int dst_slot,
                                          double *dst_ptr, int err, const ErrCoef &ec,
                                          const StageCoef &nx, int G,
                                          const DevState *chk = nullptr) {
  extern __shared__ cx smem[];
  const int L = LS > 0 ? LS : g.L;
  if constexpr (E <= 8) {
    const int Nx = GK::Nx ? GK::Nx : g.Nx, n = GK::n ? GK::n : g.n;
    const int nf = GK::nf >= 0 ? GK::nf : fl.nf;
    const long region = rowi_region(L, Nx, n, nf, dict.nwords, G, DE::kProg);
    const int gt = G * (L / E), sub = threadIdx.x / gt;
    rowi_body<E, DE, LS, GK, TWR, LEAN>((long)blockIdx.x * (blockDim.x / gt) + sub, g, px, twN, st, b, fl,
                             dict, src_kind, src_ptr, dst_kind, dst_slot, dst_ptr, err, ec, nx, G,
                             smem + sub * region, threadIdx.x - sub * gt, gt, chk);
  } else {
    rowi_body<E, DE, LS, GK, TWR, LEAN>(blockIdx.x, g, px, twN, st, b, fl, dict, src_kind, src_ptr, dst_kind,
                             dst_slot, dst_ptr, err, ec, nx, G, smem, threadIdx.x, blockDim.x,
                             chk);
  }
}

// One whole DP45 attempt of a 1-D problem (a single grid row) in one
// workgroup: the stage-2 x forward pass, the six fused x-inverse passes and
// the step control run as phases separated by barriers instead of as eight
// kernels.  Each phase is the body the separate kernel runs, so the values
// are the same bits.  Block: G*T threads (G = blockDim/T line groups, G >= n);
// LDS: the largest phase (x forward: L + G*padded_len(L) complex; x inverse:
// rowi_region; control: n doubles).
struct AttemptCoef {
  StageCoef sc2;    // stage-2 sum, formed by the x forward phase
  StageCoef nx[6];  // next-stage sums fused after stages 2..7 (nx[5]: none)
  ErrCoef ec;
};
template <int E, class DE, int LS = 0, class GK = RunGeom>
__device__ __forceinline__ void attempt1d_body(const Grid &g, const FftPlan &px,
                                               const cx *__restrict__ twN, DevState *st,
                                               const Bufs &b, const FacList &fl, const DictP &dict,
                                               const AttemptCoef &ac, int fixed) {
  extern __shared__ cx smem[];
  if (!attempt_active(st)) return;
  const int L = LS > 0 ? LS : g.L, T = L / E;
  const int nt = blockDim.x, G = nt / T;
  rowf_body<E, false, LS>(0, g, px, twN, st, b, SRC_STAGE, nullptr, ac.sc2, G);
  __syncthreads();
  // the phases' transforms take their twiddles from registers, loaded with
  // each phase's first loads: in one workgroup nothing hides a read of the
  // table in every pass
  for (int s = 2; s <= 7; ++s) {
    rowi_body<E, DE, LS, GK, true>(0, g, px, twN, st, b, fl, dict, SRC_YT, nullptr,
                             s < 7 ? DST_KS : DST_K7, s - 1, nullptr, s == 7 && !fixed, ac.ec,
                             ac.nx[s - 2], G, smem, threadIdx.x, nt);
    __syncthreads();
  }
  control_step(st, b.part, (long)g.n * g.rows, (long)g.n * g.N, reinterpret_cast<double *>(smem));
}

// LS: 0 for power-of-two lines, -1 for 3*2^k lines (run-time plans)
template <int E, int LS = 0>
__global__ __launch_bounds__(256) void k_rowi_reg(Grid g, FftPlan px, const cx *__restrict__ twN,
                                                  DevState *st, int check_active, Bufs b,
                                                  FacList fl, DictP dict, int src_kind,
                                                  const double *src_ptr, int dst_kind, int dst_slot,
                                                  double *dst_ptr, int err, ErrCoef ec,
                                                  StageCoef nx, int G) {
  rowi_rows<E, ProgDict, LS>(g, px, twN, st, b, fl, dict, src_kind, src_ptr, dst_kind, dst_slot,
                             dst_ptr, err, ec, nx, G, check_active ? st : nullptr);
}

// ---------------------------------------------------------------- batches
// Persistent launches over the instances of one shape class: the grid walks
// the (instance, block) items, skipping instances whose attempts have ended;
// a barrier between items keeps the previous item's LDS reads ahead of the
// next item's LDS writes.

template <int E, int LS = 0>
__global__ __launch_bounds__(256) void k_rowf_b(const Inst *__restrict__ in, int ninst, int per,
                                                FftPlan px, const cx *__restrict__ twN,
                                                int src_kind, StageCoef sc, int rpb, int check) {
  const long total = (long)ninst * per;
  bool first = true;
  for (long it = blockIdx.x; it < total; it += gridDim.x) {
    const int i = (int)(it / per);
    DevState *st = in[i].st;
    if (check && !attempt_active(st)) continue;
    if (!first) __syncthreads();
    first = false;
    rowf_body<E, false, LS>(it - (long)i * per, in[i].g, px, twN, st, in[i].b, src_kind, nullptr,
                            sc, rpb);
  }
}

// One (instance, column tile) item per block (a grid-stride loop around the
// column pass costs registers it does not have).  The grid is a multiple of
// 8 and items are ordered so that each XCD takes a contiguous range of
// logical blocks: neighbouring column tiles of one instance share its L2.
// KR as for col_body (blocks with several factors, 8 values per thread).
template <int E, int LS = 0, bool KR = false>
__global__ __launch_bounds__(256) void k_col_b(const Inst *__restrict__ in, int ninst, int per,
                                               FftPlan p, ColGeom cg, int mode, int fpw,
                                               int check) {
  const long l = xcd_item(blockIdx.x, gridDim.x);
  if (l >= (long)ninst * per) return;
  const int i = (int)(l / per);
  DevState *st = in[i].st;
  if (check && !attempt_active(st)) return;
  col_body<E, false, LS, false, false, KR>((int)(l - (long)i * per), in[i].g, p, cg, in[i].b,
                                           in[i].fl, mode, fpw);
}

// The x-inverse pass takes one (instance, row) item per block: it is the
// kernel with the most live registers, and a grid-stride loop around it
// would cost more.  TWR and LEAN as for rowi_body (generated kernels).
template <int E, class DE, int LS = 0, class GK = RunGeom, bool TWR = false, bool LEAN = false>
__device__ __forceinline__ void rowi_batch(const Inst *__restrict__ in, int ninst, int rows,
                                           const FftPlan &px, const cx *__restrict__ twN,
                                           int src_kind, int dst_kind, int dst_slot, int err,
                                           const ErrCoef &ec, const StageCoef &nx, int G,
                                           int check, long region) {
  extern __shared__ cx smem[];
  // row regions are sized for the largest factor count and program of the class
  if constexpr (E <= 8) {
    const int gt = G * ((LS > 0 ? LS : px.L) / E), R = blockDim.x / gt, sub = threadIdx.x / gt;
    const long it = (long)blockIdx.x * R + sub;  // R divides rows: one instance per block
    if (it >= (long)ninst * rows) return;
    const int i = (int)(it / rows);
    DevState *st = in[i].st;
    if (check && !attempt_active(st)) return;
    const Inst &I = in[i];
    rowi_body<E, DE, LS, GK, TWR, LEAN>(it - (long)i * rows, I.g, px, twN, st, I.b, I.fl, I.dict,
                                        src_kind, nullptr, dst_kind, dst_slot, nullptr, err, ec,
                                        nx, G, smem + sub * region, threadIdx.x - sub * gt, gt);
  } else {  // one row per block (see rowi_rows)
    const long it = blockIdx.x;
    if (it >= (long)ninst * rows) return;
    const int i = (int)(it / rows);
    DevState *st = in[i].st;
    if (check && !attempt_active(st)) return;
    const Inst &I = in[i];
    rowi_body<E, DE, LS, GK, TWR, LEAN>(it - (long)i * rows, I.g, px, twN, st, I.b, I.fl, I.dict,
                                        src_kind, nullptr, dst_kind, dst_slot, nullptr, err, ec,
                                        nx, G, smem, threadIdx.x, blockDim.x);
  }
}

template <int E, int LS = 0>
__global__ __launch_bounds__(256) void k_rowi_b(const Inst *__restrict__ in, int ninst, int rows,
                                                FftPlan px, const cx *__restrict__ twN,
                                                int src_kind, int dst_kind, int dst_slot, int err,
                                                ErrCoef ec, StageCoef nx, int G, int check,
                                                long region) {
  rowi_batch<E, ProgDict, LS>(in, ninst, rows, px, twN, src_kind, dst_kind, dst_slot, err, ec, nx,
                              G, check, region);
}

}  // namespace rd
