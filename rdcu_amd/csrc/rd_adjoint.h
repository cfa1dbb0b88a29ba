// rd_adjoint.h -- device kernels of the discrete adjoint (rdcu_loss_grad).
//
// The reverse sweep of a DP45 step needs, per stage s, the vector-Jacobian
// product J(Y_s)^T w of the dictionary RHS f(u)_i = sum_t C_t prod_j v_j^p_j
// (v_j the factor values, v = D u for a derivative block D), and the
// coefficient gradient sum_x w_field(t) * term_t.  With g_k the weight of
// slot k,
//     g_k(x)  = sum_t sum_{j: slot_j = k} w_field(t) C_t p_j v_j^(p_j-1) prod_{i!=j} v_i^p_i,
//     J^T w   = g_u + sum_k D_k^T g_k,      D_k^T = -D_k (first derivatives)
//                                                  +D_k (second derivatives),
// the derivative blocks being antisymmetric or symmetric real circulant
// operators.  k_rowi_vjp computes the factor values of a row (the inverse x
// pass of the factor spectra, as the RHS does), the slot weights and the
// coefficient partials; the D_k are then applied by the RHS pipeline of a
// companion problem, and k_adj_stage accumulates the stage adjoints.
#pragma once
#include "rd_kernels_reg.h"

namespace rd {

// Sum over the 64 lanes of a wave (butterfly: every lane gets the sum).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Block = G lines of T = L/E threads; one grid row r per block.  LDS:
// G*padded_len(L) complex, then (n+nf)*Nx values, (n+nf)*Nx weights, n*Nx
// adjoint inputs, the dictionary program (padded to 8 bytes), then the
// reduction doubles (max(blockDim, terms * blockDim/64)).  The twiddles are
// read from global memory.
// P[tmap[p]*rows + r] += row sum of w*term for the program's p-th term: with
// whole waves, each term's partial sums are reduced within the wave and the
// waves' sums added at the end (no barrier per term); otherwise by a tree in
// LDS per term.
// The per-term work (VD): ProgVjp interprets the dictionary program; the
// generated JitVjp (hiprtc, one per dictionary structure) is the same
// operations in the same order as straight-line code, with the slot values
// and weights in registers; it needs whole waves.
struct ProgVjp {
  static constexpr bool kGen = false;
  __device__ __forceinline__ static void run(const int *, const double *, const double *,
                                             double *, double *, int, int, int, int) {}
};

template <int E, class VD, int LS = 0>
__device__ __forceinline__ void rowi_vjp_body(const Grid &g, const FftPlan &px,
                                              const cx *__restrict__ twN, const Bufs &b,
                                              const FacList &fl, const DictP &dict,
                                              const int *__restrict__ tmap,
                                              const double *__restrict__ u,
                                              const double *__restrict__ W, double *out_id,
                                              double *Gk, double *P, int G, cx *Spc) {
  extern __shared__ cx smem[];
  const int n = g.n, Nx = g.Nx, L = g.L, T = L / E;
  const long N = g.N, rows = g.rows;
  const long r = blockIdx.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int gl = tid / T, t = tid - gl * T;
  const int nf = fl.nf, ns = n + nf;
  const int PL = padded_len(L);
  const cx *tw = px.tw;
  cx *buf = smem;
  double *vals = reinterpret_cast<double *>(buf + G * PL);
  double *gacc = vals + (long)ns * Nx;
  double *wrow = gacc + (long)ns * Nx;
  int *prog = reinterpret_cast<int *>(wrow + (long)n * Nx);
  double *red = reinterpret_cast<double *>(prog + ((dict.nwords + 1) & ~1));
  const bool waves = (nt & 63) == 0;
  const int nwv = nt >> 6;
  for (int i = tid; i < dict.nwords; i += nt) prog[i] = dict.prog[i];
  for (int w = tid; w < n * Nx; w += nt) {
    const long gi = (long)(w / Nx) * N + r * Nx + (w % Nx);
    vals[w] = u[gi];
    wrow[w] = W[gi];
  }
  for (int w = tid; w < ns * Nx; w += nt) gacc[w] = 0.0;
  __syncthreads();
  cx *lb = buf + gl * PL;
  for (int g0 = 0; g0 < nf; g0 += G) {
    const int m = g0 + gl;
    const bool act = gl < G && m < nf;
    cx v[E];
    if (act) {
      c2r_pre<E>(g, fl, b, twN, m, r, t, T, v);
    } else {
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = cx{0.0, 0.0};
    }
    fft_any<E, +1, LS>(v, lb, 1, tw, px, t, T);
    if (act) {
      double *vr = vals + (long)(n + m) * Nx;
#pragma unroll
      for (int q = 0; q < E; ++q)
        *reinterpret_cast<double2 *>(vr + 2 * (t + T * q)) = double2{v[q].re, v[q].im};
    }
  }
  __syncthreads();
  // terms in program order (fields ascending, input order within a field)
  if constexpr (VD::kGen) VD::run(prog, vals, wrow, gacc, red, Nx, tid, nt, nwv);
  for (int f = 0; f < n && !VD::kGen; ++f) {
    int p = prog[f];
    const int pe = prog[f + 1];
    int pt = 0;  // program-order term index of field f's first term
    for (int f2 = 0; f2 < f; ++f2)
      for (int q = prog[f2]; q < prog[f2 + 1]; q += 3 + 2 * prog[q]) ++pt;
    for (; p < pe; ++pt) {
      const int nfac = prog[p];
      const double cc = __hiloint2double(prog[p + 2], prog[p + 1]);
      const int *fp = prog + p + 3;
      p += 3 + 2 * nfac;
      double acc = 0.0;
      for (int x = tid; x < Nx; x += nt) {
        const double wv = wrow[f * Nx + x];
        double term = 1.0;
        for (int j = 0; j < nfac; ++j) term = term * ipow(vals[fp[2 * j] * Nx + x], fp[2 * j + 1]);
        acc += wv * term;
        for (int j = 0; j < nfac; ++j) {
          const int pj = fp[2 * j + 1];
          if (pj == 0) continue;
          const double vj = vals[fp[2 * j] * Nx + x];
          double d = pj * ipow(vj, pj - 1);
          for (int i = 0; i < nfac; ++i)
            if (i != j) d = d * ipow(vals[fp[2 * i] * Nx + x], fp[2 * i + 1]);
          gacc[fp[2 * j] * Nx + x] += wv * cc * d;
        }
      }
      if (waves) {
        acc = wave_sum(acc);
        if ((tid & 63) == 0) red[pt * nwv + (tid >> 6)] = acc;
        continue;
      }
      red[tid] = acc;
      __syncthreads();
      int p2 = 1;
      while (p2 < nt) p2 <<= 1;
      for (int s = p2 / 2; s > 0; s >>= 1) {  // any block size
        if (tid < s && tid + s < nt) red[tid] += red[tid + s];
        __syncthreads();
      }
      if (tid == 0) P[(long)tmap[pt] * rows + r] += red[0];
      __syncthreads();
    }
  }
  if (waves) {
    int npt = 0;  // terms of all fields
    for (int q = prog[0]; q < prog[n]; q += 3 + 2 * prog[q]) ++npt;
    __syncthreads();
    for (int q = tid; q < npt; q += nt) {
      double sacc = red[q * nwv];
      for (int w = 1; w < nwv; ++w) sacc += red[q * nwv + w];
      P[(long)tmap[q] * rows + r] += sacc;
    }
  }
  for (int w = tid; w < n * Nx; w += nt)
    out_id[(long)(w / Nx) * N + r * Nx + (w % Nx)] = gacc[w];
  for (int w = tid; w < nf * Nx; w += nt)
    Gk[(long)(w / Nx) * N + r * Nx + (w % Nx)] = gacc[n * Nx + w];
  // The x forward pass of the weight rows for the companion problem (its
  // field m is weight row m), with the operations of rowf_body: the real row
  // as L complex values, forward transform, then the split into the
  // half spectrum, written to the companion's S.  Only with 8 values per
  // thread: kernels with more have no registers left for it (Spc is null).
  if constexpr (E != 8) return;
  if (!Spc) return;
  const long Nxp = g.Nxp;
  for (int g0 = 0; g0 < nf; g0 += G) {
    const int m = g0 + gl;
    const bool act = gl < G && m < nf;
    const double *wr = gacc + (long)(n + (act ? m : 0)) * Nx;
    __syncthreads();  // weights complete; lines free
    cx v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const double2 o = *reinterpret_cast<const double2 *>(wr + 2 * (t + T * q));
      v[q] = act ? cx{o.x, o.y} : cx{0.0, 0.0};
    }
    fft_any<E, -1, LS>(v, lb, 1, tw, px, t, T);
#pragma unroll
    for (int q = 0; q < E; ++q) lb[t + T * q] = v[q];
    __syncthreads();
    if (act) {
      cx *S = Spc + ((long)m * rows + r) * Nxp;
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
  }
}

// LS: 0 for power-of-two lines, -1 for 3*2^k lines (run-time plans)
template <int E, int LS = 0>
__global__ __launch_bounds__(256) void k_rowi_vjp(Grid g, FftPlan px, const cx *__restrict__ twN,
                                                  Bufs b, FacList fl, DictP dict,
                                                  const int *__restrict__ tmap,
                                                  const double *__restrict__ u,
                                                  const double *__restrict__ W, double *out_id,
                                                  double *Gk, double *P, int G, cx *Spc) {
  rowi_vjp_body<E, ProgVjp, LS>(g, px, twN, b, fl, dict, tmap, u, W, out_id, Gk, P, G, Spc);
}

// The seven stage values of the step just taken, into dst[j*nN + i]: every
// step of a frozen schedule is accepted, so k1 and k7 have swapped roles
// (K1 now in the k7 slot, K7 in the k1 slot); K2..K6 are slots 1..5.
__global__ void k_save_stages(const DevState *st, Bufs b, double *dst, long nN) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nN) return;
  dst[i] = b.k[st->k7s][i];
  for (int j = 1; j <= 5; ++j) dst[j * nN + i] = b.k[j][i];
  dst[6 * nN + i] = b.k[st->k1s][i];
}

// y_n of the current step (the buffer holding y after the FSAL rotation).
__global__ void k_save_state(const DevState *st, Bufs b, double *dst, long nN) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nN) dst[i] = b.y[st->iy][i];
}

// Stage input y + h*(sum_j a_j k_j), zero coefficients skipped (the
// operations of the forward stage sums).
// The reverse sweep replays one captured step for every step k: k is read
// from the device (*kp, counted down by k_adj_next), and so are the values
// that depend on it (the step size hs[k], the checkpoint, the stored stages,
// the outputs that fall in the step).

// Stage input s of step k: out = y + h*sum_j a_j K_j, K_j at si.k[j] + k*kstride
// (kstride 0: recomputed stages in fixed buffers), h = hs[k] (dp45.cu:34-69).
struct StageIn {
  const double *k[6];
  double a[6];
  int ns;
};
__global__ void k_stage_in(const double *__restrict__ y, StageIn si, const double *__restrict__ hs,
                           const long long *__restrict__ kp, long kstride, double *out, long nN) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nN) return;
  const long long k = *kp;
  const long off = (long)k * kstride + i;
  double sacc = si.a[0] * si.k[0][off];
  for (int j = 1; j < si.ns; ++j)
    if (si.a[j] != 0.0) sacc = __builtin_fma(si.a[j], si.k[j][off], sacc);
  out[i] = __builtin_fma(hs[k], sacc, y[i]);
}

struct AdjK {
  double *k[7];
  double c[7];  // per-array coefficient (0: skip)
};

// The start of a reverse step when the forward pass stored every step's stage
// values (kst: stage j of step k at kst + (7k + j) nN), in one pass over the
// points instead of nine launches: ycur = the checkpoint of step k; the inputs
// of stages 2..7 (out[s-2]) with the operations of k_stage_in; the stage
// adjoints and ybar set to 0 and the dense outputs' adjoints added with the
// operations of k_adj_zero and k_adj_dense.
struct StageIn6 {
  double a[6][6];     // a[s-2][j]: the tableau row of stage s (j < s-1)
  double *out[6];     // the inputs of stages 2..7
};
__global__ void k_adj_begin(const double *__restrict__ ych, const long long *__restrict__ kp,
                            double *__restrict__ ycur, const double *__restrict__ kst,
                            const double *__restrict__ hs, StageIn6 si, double *ybar, AdjK kb,
                            const double *__restrict__ g, const int *__restrict__ obeg,
                            const int *__restrict__ oend, const double *__restrict__ oc,
                            long nN) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nN) return;
  const long long k = *kp;
  const double y = ych[(long)k * nN + i];
  ycur[i] = y;
  const long off = (long)k * 7 * nN + i;
  const double hk = hs[k];
#pragma unroll
  for (int s = 2; s <= 7; ++s) {
    double sacc = si.a[s - 2][0] * kst[off];
    for (int j = 1; j < s - 1; ++j)
      if (si.a[s - 2][j] != 0.0) sacc = __builtin_fma(si.a[s - 2][j], kst[off + (long)j * nN], sacc);
    si.out[s - 2][i] = __builtin_fma(hk, sacc, y);
  }
  double yb = 0.0, kbv[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int o = oend[k] - 1; o >= obeg[k]; --o) {
    const double gv = g[(long)o * nN + i];
    yb += gv;
#pragma unroll
    for (int j = 0; j < 7; ++j) {
      const double cj = oc[7 * o + j];
      if (cj != 0.0) kbv[j] += cj * gv;
    }
  }
  ybar[i] = yb;
#pragma unroll
  for (int j = 0; j < 7; ++j) kb.k[j][i] = kbv[j];
}

// ycur = checkpoint (state at the start) of step k.
__global__ void k_adj_fetch(const double *__restrict__ ych, const long long *__restrict__ kp,
                            double *__restrict__ ycur, long nN) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nN) ycur[i] = ych[(long)*kp * nN + i];
}

__global__ void k_adj_setk(long long *kp, long long k) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *kp = k;
}

__global__ void k_adj_next(long long *kp) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *kp = *kp - 1;
}

// Loss terms of one output: gbar = 2 (y - target) / cnt; per-block sums of
// (y - target)^2 into part[blockIdx].
__global__ void k_loss(const double *__restrict__ y, const double *__restrict__ tg, double cnt,
                       double *gbar, double *part, long nN) {
  __shared__ double s[256];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (i < nN) {
    const double d = y[i] - tg[i];
    d2 = d * d;
    if (gbar) gbar[i] = 2.0 * d / cnt;
  }
  s[threadIdx.x] = d2;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Zeroes the stage adjoints kbar_0..6 (kb.k) and ybar.
__global__ void k_adj_zero(AdjK kb, double *ybar, long nN) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nN) return;
#pragma unroll
  for (int j = 0; j < 7; ++j) kb.k[j][i] = 0.0;
  ybar[i] = 0.0;
}

// Adjoints of the dense outputs o of step k (obeg[k] <= o < oend[k], last
// output first): ybar += g_o, kbar_j += c_oj g_o, with g_o at g + o*nN and
// c_oj = oc[7*o + j] (0: skip).
__global__ void k_adj_dense(double *ybar, AdjK kb, const double *__restrict__ g,
                            const int *__restrict__ obeg, const int *__restrict__ oend,
                            const double *__restrict__ oc, const long long *__restrict__ kp,
                            long nN) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nN) return;
  const long long k = *kp;
  for (int o = oend[k] - 1; o >= obeg[k]; --o) {
    const double gv = g[(long)o * nN + i];
    ybar[i] += gv;
    for (int j = 0; j < 7; ++j) {
      const double cj = oc[7 * o + j];
      if (cj != 0.0) kb.k[j][i] += cj * gv;
    }
  }
}

// Adjoint of stage s: Yb = out_id + sum_{m: fmap[m] = field} Dk[m] (+ lam),
// ybar += Yb, kbar_j += (hs[k] a_j) Yb with a_j = kb.c[j] (0: skip).
__global__ void k_adj_stage(const double *__restrict__ out_id, const double *__restrict__ Dk,
                            const int *__restrict__ fmap, int nf, const double *__restrict__ lam,
                            double *ybar, AdjK kb, const double *__restrict__ hs,
                            const long long *__restrict__ kp, long N, long nN) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nN) return;
  const double h = hs[*kp];
  const int f = (int)(i / N);
  const long x = i - (long)f * N;
  double Yb = out_id[i];
  for (int m = 0; m < nf; ++m)
    if (fmap[m] == f) Yb += Dk[(long)m * N + x];
  if (lam) Yb += lam[i];
  ybar[i] += Yb;
  for (int j = 0; j < 7; ++j)
    if (kb.c[j] != 0.0) kb.k[j][i] += (h * kb.c[j]) * Yb;
}

}  // namespace rd

namespace rd {

// Switches the integrator between adaptive stepping and a frozen schedule
// (fixed steps of sizes hs[], no attempt log).
__global__ void k_set_sched(DevState *st, int fixed, const double *hs, long long ns, int nolog) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->fixed = fixed;
    st->hsched = hs;
    st->nsched = ns;
    st->isched = 0;
    st->nolog = nolog;
  }
}

}  // namespace rd
