// gm_bfab.hip -- f(A) applied to R columns at once, with the adjoint with
// respect to parameters of A (no reference counterpart): the multi-column form
// of gm_fab_* for problems that share A, the contour and the deflation pairs,
// e.g. the stimuli of a monodomain run.  Columns are paired like the batched
// engine's: even columns take the f1 weights and f1(Lambda), odd columns f2.
//
// Forward, per column c (deflated when K > 0):
//   a = Q^T b, b <- b - Q a, Lanczos (m steps, R columns per launch), shifted
//   solves y_s = beta0 w_s (xi_s - T)^{-1} e1 with the column's weights,
//   u = V Im(sum_s y_s) + Q (f(Lambda) a).
// Backward for an upstream G (R columns, same pairing): f(A) G by the same
// steps, and for each parameter theta with E = dA/dtheta (A's pattern,
// symmetric) and each column c the derivative of g^T f(A) b:
//   P-P block   Im sum_s z_s^T E (w_s x_s),  x_s = V_b ytilde_b,s (weighted
//               solves of the b side), z_s = V_g ytilde_g,s (unit weights);
//   Q-Q block   (Q^T g)^T (F o Q^T E Q) (Q^T b), F the divided differences of f;
//   cross       sum_k (Q^T g)_k q_k^T E V_b H_b[:, k] + (b <-> g), with
//               H[:, k] = (lambda_k - T)^{-1} (f(lambda_k) beta0 e1 - C).
// The P-P block is contracted in the Krylov bases, (Vg^T E Vb) against an
// m x m matrix M per column: from 16 columns with m <= 48 as S = Vg^T (E Vb)
// on FP64 MFMA in one pass over the rows, else T = Vg M on FP64 MFMA, then
// (E Vb) . T in one pass over the rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/gpumatfun.h"
#include "gm_bkernels.h"


namespace gmf {
using namespace gm;
typedef double d4 __attribute__((ext_vector_type(4)));

// The P-P block of column c in the Krylov bases:
//   Im sum_s z_s^T E x_s = sum_{k,l} (Vg^T E Vb)[k][l] M_c[k][l],
//   M_c[k][l] = Im sum_s u_s[k] y_s[l] = sum_s (Re u Im y + Im u Re y),
// with z_s = Vg u_s (unit-weight solves of the g side) and x_s = Vb y_s
// (weighted solves of the b side).  M is [R][mk][mp], zero beyond m (mk = m
// rounded up to 4, mp to the l-tile width) so that MFMA operands need no guard.
struct BPPM {
  const double* ur;   // [R][ns][m]
  const double* ui;
  const double* yr;
  const double* yi;
  int R, m, ns, mk, mp;
  double* M;
};

template <int kUnused = 0>
__global__ __launch_bounds__(256) void k_bpp_m(BPPM p) {
  const long long total = (long long)p.R * p.mk * p.mp;
  for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < total;
       t += (long long)gridDim.x * 256) {
    const int l = (int)(t % p.mp);
    const long long ck = t / p.mp;
    const int k = (int)(ck % p.mk), c = (int)(ck / p.mk);
    double v = 0.0;
    if (k < p.m && l < p.m) {
      for (int s = 0; s < p.ns; ++s) {
        const size_t b = ((size_t)c * p.ns + s) * p.m;
        v = fma(p.ur[b + k], p.yi[b + l], v);
        v = fma(p.ui[b + k], p.yr[b + l], v);
      }
    }
    p.M[t] = v;
  }
}

// T[l][i][c] = sum_k M_c[k][l] Vg[k][i][c] on FP64 MFMA (v_mfma_f64_16x16x4f64:
// A = M_c with rows l, B = Vg with columns i; A lane (row lane%16, k lane/16),
// B lane (k lane/16, col lane%16), D register r = (row 4r + lane/16, col
// lane%16)).  A wavefront takes one item (16-row tile, CW columns; the items of
// one tile are consecutive, so a block's wavefronts share Vg's lines) and l in
// blocks of 16 LT; each lane stores its row's CW columns as one vector per l.
// The P-P sum of a parameter is then sum_l sum_i (E Vb[l])[i][c] T[l][i][c]
// (k_bpp_dot).
struct BPPT {
  const double* Vg;   // [m][N][R]
  const double* M;    // [R][mk][mp]
  int N, R, m, mk, mp, items;
  double* T;          // [m][N][R]
};

template <int CW, int LT>
__global__ __launch_bounds__(256) void k_bpp_t(BPPT p) {
  typedef double vcw __attribute__((ext_vector_type(CW)));
  const int lane = threadIdx.x & 63;
  const int it = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (it >= p.items) return;   // whole wavefronts; no barrier follows
  const int CQ = p.R / CW;
  const int c0 = (it % CQ) * CW;
  const int i = (it / CQ) * 16 + (lane & 15);
  const bool rowok = i < p.N;
  const size_t slice = (size_t)p.N * p.R;
  for (int lb = 0; lb < p.m; lb += 16 * LT) {
    d4 acc[LT][CW];
#pragma unroll
    for (int lt = 0; lt < LT; ++lt)
#pragma unroll
      for (int q = 0; q < CW; ++q) acc[lt][q] = d4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < p.mk; k0 += 4) {
      const int k = k0 + (lane >> 4);
      vcw b;
#pragma unroll
      for (int q = 0; q < CW; ++q) b[q] = 0.0;
      if (k < p.m && rowok) b = *reinterpret_cast<const vcw*>(p.Vg + (size_t)k * slice + (size_t)i * p.R + c0);
#pragma unroll
      for (int q = 0; q < CW; ++q) {
        const double* Mq = p.M + ((size_t)(c0 + q) * p.mk + k) * p.mp + lb + (lane & 15);
#pragma unroll
        for (int lt = 0; lt < LT; ++lt)
          acc[lt][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(Mq[lt * 16], b[q], acc[lt][q], 0, 0, 0);
      }
    }
#pragma unroll
    for (int lt = 0; lt < LT; ++lt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int l = lb + lt * 16 + 4 * r + (lane >> 4);
        if (l < p.m && rowok) {
          vcw v;
#pragma unroll
          for (int q = 0; q < CW; ++q) v[q] = acc[lt][q][r];
          *reinterpret_cast<vcw*>(p.T + (size_t)l * slice + (size_t)i * p.R + c0) = v;
        }
      }
  }
}

// sum_l sum_i (E Vb[l])[i][c] T[l][i][c] per column in one pass: a lane takes
// one column of a row (rows interleaved over the whole grid, so the blocks
// work on one front of rows and share the gathered rows of Vb in L2), loads
// the row's nonzeros once, and runs the Krylov vectors l in pairs, the
// gathers of both issued before the first FMA.  Each (row, l) term is summed
// over the nonzeros in order, and a lane adds its terms row by row, l by l.
// partials[G][R] (lanes, then the wavefronts of a block, in a fixed order).
struct BPD {
  const int* rowptr;
  const int* col;
  const double* E;
  const double* Vb;   // [m][N][R]
  const double* T;    // [m][N][R]
  int N, m;
  double* partials;
};

template <int R>
__global__ __launch_bounds__(256) void k_bpp_dot(BPD p) {
  constexpr int CPW = R < 64 ? R : 64;
  constexpr int RPW = 64 / CPW;
  constexpr int CL = R / CPW;
  constexpr int U = 12, LB = 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int slot = lane / CPW, c0 = lane % CPW;
  const size_t slice = (size_t)p.N * R;
  double dot[CL];
#pragma unroll
  for (int q = 0; q < CL; ++q) dot[q] = 0.0;
  for (int i = (blockIdx.x * kWaves + wave) * RPW + slot; i < p.N; i += gridDim.x * kWaves * RPW) {
    const int e0 = p.rowptr[i], e1 = p.rowptr[i + 1];
    if (e1 - e0 <= U) {
      double ev[U];
      int ci[U];
#pragma unroll
      for (int t = 0; t < U; ++t) {
        const bool ok = e0 + t < e1;
        ev[t] = ok ? p.E[e0 + t] : 0.0;
        ci[t] = ok ? p.col[e0 + t] : i;
      }
      for (int l0 = 0; l0 < p.m; l0 += LB) {
        double xv[LB][U][CL], tv[LB][CL];
#pragma unroll
        for (int b = 0; b < LB; ++b) {
          const int l = l0 + b < p.m ? l0 + b : l0;   // (a repeated l is not added)
          const double* Vl = p.Vb + (size_t)l * slice + c0;
#pragma unroll
          for (int t = 0; t < U; ++t)
#pragma unroll
            for (int q = 0; q < CL; ++q) xv[b][t][q] = e0 + t < e1 ? Vl[(size_t)ci[t] * R + q * CPW] : 0.0;
#pragma unroll
          for (int q = 0; q < CL; ++q) tv[b][q] = p.T[(size_t)l * slice + (size_t)i * R + c0 + q * CPW];
        }
#pragma unroll
        for (int b = 0; b < LB; ++b) {
          if (l0 + b >= p.m) break;
#pragma unroll
          for (int q = 0; q < CL; ++q) {
            double sv = 0.0;
#pragma unroll
            for (int t = 0; t < U; ++t)
              if (e0 + t < e1) sv = fma(ev[t], xv[b][t][q], sv);
            dot[q] = fma(sv, tv[b][q], dot[q]);
          }
        }
      }
    } else {
      // a long row: nonzero by nonzero for each l
      for (int l = 0; l < p.m; ++l) {
        const double* Vl = p.Vb + (size_t)l * slice + c0;
#pragma unroll
        for (int q = 0; q < CL; ++q) {
          double sv = 0.0;
          for (int e = e0; e < e1; ++e) sv = fma(p.E[e], Vl[(size_t)p.col[e] * R + q * CPW], sv);
          dot[q] = fma(sv, p.T[(size_t)l * slice + (size_t)i * R + c0 + q * CPW], dot[q]);
        }
      }
    }
  }
  gmb::block_partial_cols<R>(dot, p.partials);
}

// k_bpp_dot with two adjacent columns a lane (16-byte accesses; R / 2 lanes a
// row, 128 / R rows a wavefront, twice k_bpp_dot's rows in flight).  The same
// terms: each (row, l, column) sum over the nonzeros in order; a lane adds its
// terms row by row, l by l, and the block partials [G][R] take the lanes of a
// column in (wavefront, row slot) order.  R = 4 .. 128.
template <int R>
__global__ __launch_bounds__(256) void k_bpp_dot2(BPD p) {
  static_assert(R >= 4 && R <= 128 && (R & (R - 1)) == 0, "two columns a lane");
  typedef double dv2 __attribute__((ext_vector_type(2)));
  constexpr int LPR = R / 2, RPW = 64 / LPR, U = 12;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int slot = lane / LPR, c = 2 * (lane % LPR);
  const size_t slice = (size_t)p.N * R;
  double d0 = 0.0, d1 = 0.0;
  for (int i = (blockIdx.x * kWaves + wave) * RPW + slot; i < p.N; i += gridDim.x * kWaves * RPW) {
    const int e0 = p.rowptr[i], e1 = p.rowptr[i + 1];
    if (e1 - e0 <= U) {
      double ev[U];
      int ci[U];
#pragma unroll
      for (int t = 0; t < U; ++t) {
        const bool ok = e0 + t < e1;
        ev[t] = ok ? p.E[e0 + t] : 0.0;
        ci[t] = ok ? p.col[e0 + t] : i;
      }
      for (int l = 0; l < p.m; ++l) {
        const double* Vl = p.Vb + (size_t)l * slice + c;
        dv2 xv[U];
#pragma unroll
        for (int t = 0; t < U; ++t)
          xv[t] = e0 + t < e1 ? *reinterpret_cast<const dv2*>(Vl + (size_t)ci[t] * R) : dv2{0.0, 0.0};
        const dv2 tv = *reinterpret_cast<const dv2*>(p.T + (size_t)l * slice + (size_t)i * R + c);
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int t = 0; t < U; ++t)
          if (e0 + t < e1) {
            s0 = fma(ev[t], xv[t].x, s0);
            s1 = fma(ev[t], xv[t].y, s1);
          }
        d0 = fma(s0, tv.x, d0);
        d1 = fma(s1, tv.y, d1);
      }
    } else {
      // a long row: nonzero by nonzero for each l
      for (int l = 0; l < p.m; ++l) {
        const double* Vl = p.Vb + (size_t)l * slice + c;
        double s0 = 0.0, s1 = 0.0;
        for (int e = e0; e < e1; ++e) {
          const dv2 x = *reinterpret_cast<const dv2*>(Vl + (size_t)p.col[e] * R);
          s0 = fma(p.E[e], x.x, s0);
          s1 = fma(p.E[e], x.y, s1);
        }
        const dv2 tv = *reinterpret_cast<const dv2*>(p.T + (size_t)l * slice + (size_t)i * R + c);
        d0 = fma(s0, tv.x, d0);
        d1 = fma(s1, tv.y, d1);
      }
    }
  }
  __shared__ double sh[2 * 256];
  sh[2 * threadIdx.x] = d0;
  sh[2 * threadIdx.x + 1] = d1;
  __syncthreads();
  for (int col = threadIdx.x; col < R; col += 256) {
    const int lc = col >> 1, q = col & 1;
    double sum = 0.0;
    for (int w = 0; w < kWaves; ++w)
      for (int sl = 0; sl < RPW; ++sl) sum += sh[2 * (w * 64 + sl * LPR + lc) + q];
    p.partials[(size_t)blockIdx.x * R + col] = sum;
  }
}

// The P-P sums without T: S_c = Vg_c^T (E Vb_c), m x m per column, on FP64
// MFMA, contracted with M_c once a block has run its rows:
//   sum_{k,l} S_c[k][l] M_c[k][l] = sum_l (E Vb[l]) . T[l].
// A block (8 wavefronts) takes 16 columns (one 128-byte line of a row) and a
// group of consecutive 4-row steps.  Per step it stages, for k, l < 16 KT,
//   sV[k][r][j] = Vg[k][i][cb + j],  sE[l][r][j] = (E Vb[l])[i][cb + j]
// (i = 4 step + r; zero beyond m or N) in LDS, a thread a column pair with
// the 8 threads of one (k or l, r) reading a whole line; each E Vb value is
// summed over its row's nonzeros in order.  Wavefront w then runs the MFMA
// chains of columns cb + 2w, cb + 2w + 1 from LDS: A = sV[kt 16 + lane % 16]
// [lane / 16] (rows k of S), B = sE[lt 16 + lane % 16][lane / 16] (columns
// l); D register r of tile (kt, lt) holds S[kt 16 + 4 r + lane / 16][lt 16 +
// lane % 16].  Two LDS buffers alternate, so one barrier a step suffices.
// The wavefront's sum goes to partials[row group][c] (lane order fixed by the
// butterfly).
struct BPS {
  const int* rowptr;
  const int* col;
  const double* E;
  const double* Vg;   // [m][N][R]
  const double* Vb;   // [m][N][R]
  const double* M;    // [R][mk][mp]
  int N, R, m, mk, mp, NG;
  double* partials;   // [NG][R]
};

constexpr int kPsStride = 4 * 16 + 2;   // doubles per k (or l) row of a staged tile

template <int KT>
__global__ __launch_bounds__(512) void k_bpp_s(BPS p) {
  typedef double dv2 __attribute__((ext_vector_type(2)));
  constexpr int U = 9;   // the 9-point stencil; longer rows take the nonzero-by-nonzero path
  constexpr int TILE = KT * 16 * kPsStride;
  constexpr int PASSES = KT * 16 * 4 / 64;   // (k, r) pairs over the block's 64 thread groups
  __shared__ double sV[2][TILE];
  __shared__ double sE[2][TILE];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int halves = p.R >> 4;
  const int cb = (int)(blockIdx.x % halves) * 16, rg = (int)(blockIdx.x / halves);
  const int t16 = lane & 15, r4 = lane >> 4;
  const size_t slice = (size_t)p.N * p.R;
  const int steps = (p.N + 3) >> 2;
  const int sp = (steps + p.NG - 1) / p.NG;
  const int s0 = min(steps, rg * sp), s1 = min(steps, s0 + sp);
  // fill role: column pair jp (columns cb + 2 jp, + 1), row r of the step,
  // and (k or l) = x0 + 16 pass for the passes below
  const int jp = tid & 7, grp = tid >> 3, fr = grp & 3, x0 = grp >> 2;
  d4 acc[KT][KT][2];
#pragma unroll
  for (int kt = 0; kt < KT; ++kt)
#pragma unroll
    for (int lt = 0; lt < KT; ++lt)
#pragma unroll
      for (int q = 0; q < 2; ++q) acc[kt][lt][q] = d4{0.0, 0.0, 0.0, 0.0};
  for (int s = s0; s < s1; ++s) {
    double* bV = sV[s & 1];
    double* bE = sE[s & 1];
    const int i = 4 * s + fr;
    const bool rowok = i < p.N;
    int e0 = 0, e1 = 0;
    if (rowok) {
      e0 = p.rowptr[i];
      e1 = p.rowptr[i + 1];
    }
    const bool shortrow = e1 - e0 <= U;
    // padding slots of a short row point at the row itself, so that the
    // gathers below are issued together without branches; their terms are
    // dropped by a select
    double ev[U];
    int ci[U];
#pragma unroll
    for (int t = 0; t < U; ++t) {
      const bool ok = shortrow && e0 + t < e1;
      ev[t] = ok ? p.E[e0 + t] : 0.0;
      ci[t] = ok ? p.col[e0 + t] : i;
    }
#pragma unroll
    for (int ps = 0; ps < PASSES; ++ps) {
      const int x = x0 + 16 * ps;
      dv2 v = dv2{0.0, 0.0}, e = dv2{0.0, 0.0};
      if (x < p.m && rowok) {
        const size_t off = (size_t)x * slice + cb + 2 * jp;
        v = *reinterpret_cast<const dv2*>(p.Vg + off + (size_t)i * p.R);
        const double* Vx = p.Vb + off;
        if (shortrow) {
          dv2 xv[U];
#pragma unroll
          for (int t = 0; t < U; ++t) xv[t] = *reinterpret_cast<const dv2*>(Vx + (size_t)ci[t] * p.R);
#pragma unroll
          for (int t = 0; t < U; ++t) {
            const bool in = e0 + t < e1;
            const double sx = fma(ev[t], xv[t].x, e.x), sy = fma(ev[t], xv[t].y, e.y);
            e.x = in ? sx : e.x;
            e.y = in ? sy : e.y;
          }
        } else {
          for (int q = e0; q < e1; ++q) {
            const dv2 xq = *reinterpret_cast<const dv2*>(Vx + (size_t)p.col[q] * p.R);
            e.x = fma(p.E[q], xq.x, e.x);
            e.y = fma(p.E[q], xq.y, e.y);
          }
        }
      }
      *reinterpret_cast<dv2*>(bV + x * kPsStride + fr * 16 + 2 * jp) = v;
      *reinterpret_cast<dv2*>(bE + x * kPsStride + fr * 16 + 2 * jp) = e;
    }
    __syncthreads();
    dv2 a[KT], b[KT];
#pragma unroll
    for (int kt = 0; kt < KT; ++kt) {
      a[kt] = *reinterpret_cast<const dv2*>(bV + (kt * 16 + t16) * kPsStride + r4 * 16 + 2 * w);
      b[kt] = *reinterpret_cast<const dv2*>(bE + (kt * 16 + t16) * kPsStride + r4 * 16 + 2 * w);
    }
#pragma unroll
    for (int kt = 0; kt < KT; ++kt)
#pragma unroll
      for (int lt = 0; lt < KT; ++lt) {
        acc[kt][lt][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kt].x, b[lt].x, acc[kt][lt][0], 0, 0, 0);
        acc[kt][lt][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kt].y, b[lt].y, acc[kt][lt][1], 0, 0, 0);
      }
  }
  const int c = cb + 2 * w;
  double d0 = 0.0, d1 = 0.0;
  const double* M0 = p.M + (size_t)c * p.mk * p.mp;
  const double* M1 = M0 + (size_t)p.mk * p.mp;
#pragma unroll
  for (int kt = 0; kt < KT; ++kt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = kt * 16 + 4 * r + r4;
#pragma unroll
      for (int lt = 0; lt < KT; ++lt) {
        const int l = lt * 16 + t16;
        if (k < p.m && l < p.m) {
          d0 = fma(acc[kt][lt][0][r], M0[(size_t)k * p.mp + l], d0);
          d1 = fma(acc[kt][lt][1][r], M1[(size_t)k * p.mp + l], d1);
        }
      }
    }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    d0 += __shfl_xor(d0, off);
    d1 += __shfl_xor(d1, off);
  }
  if (lane == 0) {
    p.partials[(size_t)rg * p.R + c] = d0;
    p.partials[(size_t)rg * p.R + c + 1] = d1;
  }
}

// Block partials of M[j][k][c] = sum_i V[j][i][c] EQ[i][k] for every Krylov
// vector j (block x; the row chunk is block y) on FP64 MFMA: each wavefront holds all KT x CT 16 x 16
// tiles of (k, c), so one EQ load feeds CT chains and one V load KT chains;
// the 4 wavefronts take interleaved 4-row steps of the block's rows and are
// summed in LDS in wavefront order.  partials[G][m][K][R].
struct BVtQ {
  const double* V;    // [m][N][R]
  const double* EQ;   // [N][K]
  int K, R, m, nrows, chunk;
  double* partials;
};

template <int KT, int CT>
__global__ __launch_bounds__(256) void k_bvtq(BVtQ p) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // blockIdx.x = Krylov vector j, blockIdx.y = row chunk: the m blocks of one
  // chunk are consecutive, so they run together and share its EQ rows in L2
  const int j = blockIdx.x, cb = blockIdx.y;
  const int row0 = cb * p.chunk, row1 = min(row0 + p.chunk, p.nrows);
  const int sub = lane >> 4;
  const double* Vj = p.V + (size_t)j * p.nrows * p.R;
  __shared__ double sh[KT * CT * 4 * 64];
  d4 acc[KT][CT];
#pragma unroll
  for (int kt = 0; kt < KT; ++kt)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) acc[kt][ct] = d4{0.0, 0.0, 0.0, 0.0};
  // four 4-row steps a pass, all their loads issued together: out-of-range
  // operands load a valid element (row0, column R - 1, k K - 1) and are
  // replaced by zero afterwards, so no load sits behind a branch; the MFMA
  // chains take the steps in the same order as one step a pass
  constexpr int UR = 4;
  for (int r = row0 + 4 * wave; r < row1; r += 16 * UR) {
    double a[UR][KT], b[UR][CT];
#pragma unroll
    for (int u = 0; u < UR; ++u) {
      const int row = r + 16 * u + sub;
      const size_t rc = row < row1 ? row : row0;
#pragma unroll
      for (int kt = 0; kt < KT; ++kt)
        a[u][kt] = p.EQ[rc * p.K + min(kt * 16 + (lane & 15), p.K - 1)];
#pragma unroll
      for (int ct = 0; ct < CT; ++ct)
        b[u][ct] = Vj[rc * p.R + min(ct * 16 + (lane & 15), p.R - 1)];
    }
#pragma unroll
    for (int u = 0; u < UR; ++u) {
      const bool ok = r + 16 * u + sub < row1;
#pragma unroll
      for (int kt = 0; kt < KT; ++kt) a[u][kt] = (ok && kt * 16 + (lane & 15) < p.K) ? a[u][kt] : 0.0;
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) b[u][ct] = (ok && ct * 16 + (lane & 15) < p.R) ? b[u][ct] : 0.0;
#pragma unroll
      for (int kt = 0; kt < KT; ++kt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
          acc[kt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][kt], b[u][ct], acc[kt][ct], 0, 0, 0);
    }
  }
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int kt = 0; kt < KT; ++kt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            double& d = sh[((kt * CT + ct) * 4 + i) * 64 + lane];
            d = w == 0 ? acc[kt][ct][i] : d + acc[kt][ct][i];
          }
    }
    __syncthreads();
  }
  // D register i of lane l = (k 16 kt + 4 i + l/16, c 16 ct + l%16)
  for (int t = threadIdx.x; t < KT * CT * 4 * 64; t += 256) {
    const int ln = t & 63, i = (t >> 6) & 3, tile = t >> 8;
    const int k2 = (tile / CT) * 16 + 4 * i + (ln >> 4), c = (tile % CT) * 16 + (ln & 15);
    if (k2 < p.K && c < p.R)
      p.partials[(((size_t)cb * p.m + j) * p.K + k2) * p.R + c] = sh[t];
  }
}

// y += x (n values).
template <int kUnused = 0>
__global__ __launch_bounds__(256) void k_add(const double* x, double* y, size_t n) {
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (size_t)gridDim.x * 256)
    y[t] += x[t];
}

}  // namespace gmf

namespace {

thread_local std::string g_ferr;
std::mutex g_fmu;
std::unordered_set<const void*> g_freg;

gm_status ffail(gm_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_ferr = buf;
  return st;
}

#define F_HIP(call)                                                                      \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ffail(e_ == hipErrorOutOfMemory ? GM_ENOMEM : GM_EDEVICE, "%s: %s (%s:%d)", \
                   #call, hipGetErrorString(e_), __FILE__, __LINE__);                    \
  } while (0)
#define F_TRY(x)              \
  do {                        \
    gm_status s_ = (x);       \
    if (s_ != GM_OK) return s_; \
  } while (0)

inline int bdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

struct FSide {
  double *U = nullptr, *V = nullptr, *W0 = nullptr, *W1 = nullptr, *al = nullptr, *be = nullptr,
         *yr = nullptr, *yi = nullptr, *ur = nullptr, *ui = nullptr, *C = nullptr,
         *out = nullptr, *Zd = nullptr, *upd = nullptr;
  // the stream this side's deflation and Krylov run are enqueued on, and the
  // block partials and solve scratch they use (the g side's are the engine's
  // own; the b side has its stream and buffers, so that the b side's
  // recomputed forward run can overlap the g side's run of a backward)
  hipStream_t st = nullptr;
  double* partials = nullptr;
  double2 *wD = nullptr, *wDU = nullptr, *wDU2 = nullptr, *wX = nullptr;
};

}  // namespace

struct gm_bfab {
  int dev = 0;
  hipStream_t stream = nullptr;
  // the b side's stream; ev2 marks the end of its last enqueued run, and
  // pend2 that a forward pass without output has not been joined yet (the
  // next entry point makes the engine stream wait for it); ev1 orders the b
  // side after the engine stream's earlier work
  hipStream_t stream2 = nullptr;
  hipEvent_t ev1 = nullptr, ev2 = nullptr;
  bool pend2 = false;
  int n = 0, R = 0, m = 0, ns = 0, K = 0;
  long long nnz = 0;
  bool have_forward = false;
  int G_sp = 1, chunk_sp = 1, nrowblk = 0, G_df = 1, chunk_df = 1;
  int CW = 4, LT = 1, mk = 4, mp = 16, pp_items = 0;   // T kernel shape
  int G_pd = 1;                                         // one-pass P-P dot grid
  int *rowptr = nullptr, *col = nullptr, *rowblk = nullptr;
  // slab SpMM of the forward Krylov runs at >= 32 columns (gmb::k_bspmm_xs):
  // row-block descriptors and grid
  int* rowblk_xs = nullptr;
  int nrowblk_xs = 0, G_xs = 8;
  double *val = nullptr, *xi = nullptr, *wts1 = nullptr, *wts2 = nullptr, *ones = nullptr,
         *gsr = nullptr, *gsi = nullptr, *Q = nullptr, *el = nullptr, *pl = nullptr,
         *partials = nullptr, *ppart = nullptr, *Mpp = nullptr, *Tpp = nullptr, *pdot = nullptr;
  double2 *wD = nullptr, *wDU = nullptr, *wDU2 = nullptr, *wX = nullptr;
  unsigned* status = nullptr;
  std::vector<double> lam, el_h, pl_h, dd1, dd2;   // host (dd K x K column-major)
  std::vector<double*> pdA, pEQ;                    // stored parameter derivatives
  std::vector<std::vector<double>> pQEQ;
  FSide sb, sg;
};

namespace {

bool falive(const gm_bfab* f) {
  std::lock_guard<std::mutex> lk(g_fmu);
  return f && g_freg.count(f);
}

template <typename T>
gm_status falloc(T** p, size_t count) {
  *p = nullptr;
  F_HIP(hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T)));
  return GM_OK;
}

void ffree(void* p) {
  if (p) (void)hipFree(p);
}

void free_bfab(gm_bfab* f) {
  void* bufs[] = {f->rowptr, f->col, f->rowblk, f->rowblk_xs, f->val, f->xi, f->wts1, f->wts2, f->ones, f->gsr,
                  f->gsi, f->Q, f->el, f->pl, f->partials, f->ppart, f->wD, f->wDU, f->wDU2,
                  f->wX, f->status, f->Mpp, f->Tpp, f->pdot};
  for (void* p : bufs) ffree(p);
  for (double* p : f->pdA) ffree(p);
  for (double* p : f->pEQ) ffree(p);
  {
    // the b side's own partials and solve scratch (the g side's are the engine's)
    void* bbufs[] = {f->sb.partials, f->sb.wD, f->sb.wDU, f->sb.wDU2, f->sb.wX};
    for (void* p : bbufs) ffree(p);
  }
  for (FSide* s : {&f->sb, &f->sg}) {
    void* sbufs[] = {s->U, s->V, s->W0, s->W1, s->al, s->be, s->yr, s->yi, s->ur, s->ui, s->C,
                     s->out, s->Zd, s->upd};
    for (void* p : sbufs) ffree(p);
  }
  if (f->ev1) (void)hipEventDestroy(f->ev1);
  if (f->ev2) (void)hipEventDestroy(f->ev2);
  if (f->stream2) (void)hipStreamDestroy(f->stream2);
  if (f->stream) (void)hipStreamDestroy(f->stream);
}

gm_status freduce_on(gm_bfab* f, double* partials, hipStream_t st, int G, int nv, double* out,
                     int do_sqrt, int flag_zero) {
  gm::ReduceArgs r{};
  r.partials = partials;
  r.G = G;
  r.nv = nv;
  r.out = out;
  r.stride = 1;
  r.do_sqrt = do_sqrt;
  r.flag_zero = flag_zero;
  r.status = f->status;
  hipLaunchKernelGGL(gm::k_reduce<0>, dim3(gm::reduce_grid(r.nv)), dim3(gm::kReduceBlock), 0, st, r);
  F_HIP(hipGetLastError());
  return GM_OK;
}

gm_status freduce(gm_bfab* f, int G, int nv, double* out, int do_sqrt, int flag_zero) {
  gm::ReduceArgs r{};
  r.partials = f->partials;
  r.G = G;
  r.nv = nv;
  r.out = out;
  r.stride = 1;
  r.do_sqrt = do_sqrt;
  r.flag_zero = flag_zero;
  r.status = f->status;
  hipLaunchKernelGGL(gm::k_reduce<0>, dim3(gm::reduce_grid(r.nv)), dim3(gm::kReduceBlock), 0, f->stream, r);
  F_HIP(hipGetLastError());
  return GM_OK;
}

// Q^T U -> s.Zd (K x R); U <- U - Q Zd; s.upd = Q (f(Lambda) Zd) by column parity.
gm_status fdeflate(gm_bfab* f, FSide& s) {
  gmb::BDeflQtU q{};
  q.Q = f->Q;
  q.U = s.U;
  q.K = f->K;
  q.R = f->R;
  q.nrows = f->n;
  q.chunk = f->chunk_df;
  q.KT = bdiv(f->K, 16);
  q.CT = bdiv(f->R, 16);
  q.partials = s.partials;
  gmb::launch_bdefl_qtu(q, f->G_df, s.st);
  F_HIP(hipGetLastError());
  F_TRY(freduce_on(f, s.partials, s.st, f->G_df, f->K * f->R, s.Zd, 0, 0));
  gmb::BDeflApply a{};
  a.Q = f->Q;
  a.Z = s.Zd;
  a.el = f->el;
  a.pl = f->pl;
  a.Usol = s.U;
  a.Ud = s.U;    // in place: each element is read and written by the same thread
  a.upd = s.upd;
  a.K = f->K;
  a.R = f->R;
  a.nrows = f->n;
  hipLaunchKernelGGL(gmb::k_bdefl_apply<0>, dim3(std::min(4096, bdiv(f->n, gm::kWaves)), bdiv(f->R, 64)),
                     dim3(gm::kBlock), 2 * (size_t)f->K * 64 * sizeof(double), s.st, a);
  F_HIP(hipGetLastError());
  return GM_OK;
}

// Lanczos on s.U (R columns), the weighted shifted solves (s.yr, s.yi), and
// with want_out s.out = V C (+ s.upd); with unit also the unit-weight solves
// (s.ur, s.ui).
template <int R>
gm_status fkrylov(gm_bfab* f, FSide& s, bool want_out, bool unit) {
  const size_t N = (size_t)f->n;
  gmb::BNorm nn{};
  nn.u = s.U;
  nn.ldu = R;
  nn.nrows = f->n;
  nn.chunk = f->chunk_sp;
  nn.partials = s.partials;
  nn.status = f->status;
  hipLaunchKernelGGL((gmb::k_bnorm<R>), dim3(f->G_sp), dim3(gm::kBlock), 0, s.st, nn);
  F_HIP(hipGetLastError());
  F_TRY(freduce_on(f, s.partials, s.st, f->G_sp, R, s.be, 1, 0));
  gmb::BSpmm base{};
  base.rowptr = f->rowptr;
  base.col = f->col;
  base.val = f->val;
  base.nrows = f->n;
  base.chunk = f->chunk_sp;
  base.rowblk = f->rowblk;
  base.nrowblk = f->nrowblk;
  base.al = s.al;
  base.be = s.be;
  base.partials = s.partials;
  base.g0 = 1.0;
  base.status = f->status;
  double* Wb[2] = {s.W0, s.W1};
  for (int k = 0; k < f->m; ++k) {
    gmb::BSpmm a = base;
    a.k = k;
    a.vk = s.V + (size_t)k * N * R;
    a.vkm1 = k > 0 ? s.V + (size_t)(k - 1) * N * R : nullptr;
    a.x = k == 0 ? s.U : Wb[(k - 1) & 1];
    a.ldx = R;
    a.y = Wb[k & 1];
    a.ldy = R;
    int gpart = f->G_sp;
    if constexpr (R >= 32) {
      // column slabs of 32 (measured faster than whole rows at 32 columns on
      // the 512^2 9-point operator: 78.6 against 96 us a launch)
      a.rowblk = f->rowblk_xs;
      a.nrowblk = f->nrowblk_xs;
      hipLaunchKernelGGL((gmb::k_bspmm_xs<R, gmb::kOnly, false>), dim3(f->G_xs), dim3(gm::kBlock),
                         0, s.st, a);
      gpart = f->G_xs / (R / 32);
    } else {
      hipLaunchKernelGGL((gmb::k_bspmm<R, gmb::kOnly, true>), dim3(f->G_sp), dim3(gm::kBlock), 0,
                         s.st, a);
    }
    F_HIP(hipGetLastError());
    F_TRY(freduce_on(f, s.partials, s.st, gpart, R, s.al + (size_t)k * R, 0, 0));
    gmb::BFinish fn{};
    fn.w = Wb[k & 1];
    fn.vk = a.vk;
    fn.al = s.al;
    fn.be = s.be;
    fn.k = k;
    fn.nrows = f->n;
    fn.chunk = f->chunk_sp;
    fn.partials = s.partials;
    fn.status = f->status;
    if constexpr (R >= 4 && R <= 64) {   // measured: 34.3 against 35.9 us at 32 columns; 2.43 against 2.40 ms at 256
      hipLaunchKernelGGL((gmb::k_bfinish2<R>), dim3(f->G_sp), dim3(gm::kBlock), 0, s.st, fn);
    } else {
      hipLaunchKernelGGL((gmb::k_bfinish<R>), dim3(f->G_sp), dim3(gm::kBlock), 0, s.st, fn);
    }
    F_HIP(hipGetLastError());
    F_TRY(freduce_on(f, s.partials, s.st, f->G_sp, R, s.be + (size_t)(k + 1) * R, 1, 1));
  }
  for (int pass = 0; pass < (unit ? 2 : 1); ++pass) {
    gmb::BSolve sv{};
    sv.al = s.al;
    sv.be = s.be;
    sv.R = R;
    sv.m = f->m;
    sv.ns = f->ns;
    sv.wts1 = pass ? f->ones : f->wts1;
    sv.wts2 = pass ? f->ones : f->wts2;
    sv.etaxi = f->xi;
    sv.wD = s.wD;
    sv.wDU = s.wDU;
    sv.wDU2 = s.wDU2;
    sv.wX = s.wX;
    sv.yr = pass ? s.ur : s.yr;
    sv.yi = pass ? s.ui : s.yi;
    sv.status = f->status;
    hipLaunchKernelGGL(gmb::k_bsolves<0>, dim3(bdiv((long long)R * f->ns, 64)), dim3(64), 0,
                       s.st, sv);
    F_HIP(hipGetLastError());
  }
  gmb::BCollapse cc{};
  cc.yr = s.yr;
  cc.yi = s.yi;
  cc.gsr = f->gsr;
  cc.gsi = f->gsi;
  cc.R = R;
  cc.m = f->m;
  cc.ns = f->ns;
  cc.nd = 1;
  cc.C = s.C;
  hipLaunchKernelGGL(gmb::k_bcollapse<0>, dim3(std::max(1, std::min(1024, bdiv((long long)f->m * R, gm::kBlock)))),
                     dim3(gm::kBlock), 0, s.st, cc);
  F_HIP(hipGetLastError());
  if (want_out) {
    gmb::BVC v{};
    v.V = s.V;
    v.C = s.C;
    v.R = R;
    v.m = f->m;
    v.nd = 1;
    v.nrows = f->n;
    v.Z = s.out;
    v.ldz = R;
    v.accumulate = 0;
    if (R % 4 == 0 && reinterpret_cast<uintptr_t>(s.out) % 16 == 0) {
      // four adjacent columns a thread, 16-byte loads (7% faster at 256 columns)
      hipLaunchKernelGGL((gmb::k_bvc<1, 4>), dim3(bdiv((long long)f->n * R / 4, gm::kBlock)),
                         dim3(gm::kBlock), 0, s.st, v);
    } else {
      hipLaunchKernelGGL((gmb::k_bvc<1>), dim3(bdiv((long long)f->n * R, gm::kBlock)), dim3(gm::kBlock), 0,
                         s.st, v);
    }
    F_HIP(hipGetLastError());
    if (f->K) {
      hipLaunchKernelGGL(gmf::k_add<0>, dim3(std::min(8192, bdiv((long long)N * R, 256))), dim3(256), 0,
                         s.st, s.upd, s.out, N * R);
      F_HIP(hipGetLastError());
    }
  }
  return GM_OK;
}

gm_status fkrylov_dispatch(gm_bfab* f, FSide& s, bool want_out, bool unit) {
  switch (f->R) {
    case 2: return fkrylov<2>(f, s, want_out, unit);
    case 4: return fkrylov<4>(f, s, want_out, unit);
    case 8: return fkrylov<8>(f, s, want_out, unit);
    case 16: return fkrylov<16>(f, s, want_out, unit);
    case 32: return fkrylov<32>(f, s, want_out, unit);
    default: return fkrylov<64>(f, s, want_out, unit);
  }
}

// M for the P-P block from the unit-weight solves of the g side and the
// weighted solves of the b side, then T = Vg M (both per backward, for all
// parameters).
template <int CW>
void launch_bpp_t(const gmf::BPPT& p, int LT, hipStream_t st) {
  const dim3 grid(bdiv(p.items, 4));
  switch (LT) {
    case 1: hipLaunchKernelGGL((gmf::k_bpp_t<CW, 1>), grid, dim3(256), 0, st, p); break;
    case 2: hipLaunchKernelGGL((gmf::k_bpp_t<CW, 2>), grid, dim3(256), 0, st, p); break;
    case 3: hipLaunchKernelGGL((gmf::k_bpp_t<CW, 3>), grid, dim3(256), 0, st, p); break;
    default: hipLaunchKernelGGL((gmf::k_bpp_t<CW, 4>), grid, dim3(256), 0, st, p); break;
  }
}

// The P-P block without T (k_bpp_s) while all S tiles fit one wavefront's
// accumulators (m <= 48: 3 x 3 tiles of 16 for each of two columns) and a
// block's 16 columns fit R (measured on 512^2 at 32 columns, m = 40: 2.79 ms
// against 1.99 + 1.38 ms for T and the dot; and no N R m buffer for T).
bool fpp_fused(const gm_bfab* f) {
  return f->m <= 48 && f->R >= 16;
}

gm_status fppm(gm_bfab* f) {
  const size_t N = (size_t)f->n;
  if (!f->Mpp) F_TRY(falloc(&f->Mpp, (size_t)f->R * f->mk * f->mp));
  if (!f->pdot) F_TRY(falloc(&f->pdot, ((size_t)f->G_pd + 1) * f->R));
  gmf::BPPM a{f->sg.ur, f->sg.ui, f->sb.yr, f->sb.yi, f->R, f->m, f->ns, f->mk, f->mp, f->Mpp};
  hipLaunchKernelGGL(gmf::k_bpp_m<0>, dim3(std::min(4096, bdiv((long long)f->R * f->mk * f->mp, 256))),
                     dim3(256), 0, f->stream, a);
  F_HIP(hipGetLastError());
  if (fpp_fused(f)) return GM_OK;
  if (!f->Tpp) F_TRY(falloc(&f->Tpp, N * f->R * f->m));
  gmf::BPPT t{f->sg.V, f->Mpp, f->n, f->R, f->m, f->mk, f->mp, f->pp_items, f->Tpp};
  if (f->CW == 2)
    launch_bpp_t<2>(t, f->LT, f->stream);
  else
    launch_bpp_t<4>(t, f->LT, f->stream);
  F_HIP(hipGetLastError());
  return GM_OK;
}

// The P-P sums of one parameter (E on A's pattern) for all R columns:
// sum_{k,l} (Vg^T E Vb)[k][l] M_c[k][l] by k_bpp_s, or sum_l (E Vb[l]) . T[l]
// per column by k_bpp_dot; partials [G_pd][R] summed into pdot[G_pd R .. + R).
// (m launches of the batched SpMM, one per l, measured 3.1 against 2.25 ms
// for the one-pass dot on 512^2 at 32 columns, m = 40.)
template <int R>
gm_status fpp(gm_bfab* f, const double* E) {
  gmf::BPD d{f->rowptr, f->col, E, f->sb.V, f->Tpp, f->n, f->m, f->pdot};
  if (fpp_fused(f)) {
    // G_pd row groups, R / 16 column groups of 16, 8 wavefronts a block
    gmf::BPS b{f->rowptr, f->col, E, f->sg.V, f->sb.V, f->Mpp, f->n, R, f->m, f->mk, f->mp, f->G_pd, f->pdot};
    const dim3 grid(f->G_pd * (R / 16));
    if (f->m <= 16)
      hipLaunchKernelGGL((gmf::k_bpp_s<1>), grid, dim3(512), 0, f->stream, b);
    else if (f->m <= 32)
      hipLaunchKernelGGL((gmf::k_bpp_s<2>), grid, dim3(512), 0, f->stream, b);
    else
      hipLaunchKernelGGL((gmf::k_bpp_s<3>), grid, dim3(512), 0, f->stream, b);
  } else if constexpr (R >= 4)
    hipLaunchKernelGGL((gmf::k_bpp_dot2<R>), dim3(f->G_pd), dim3(256), 0, f->stream, d);
  else
    hipLaunchKernelGGL((gmf::k_bpp_dot<R>), dim3(f->G_pd), dim3(256), 0, f->stream, d);
  F_HIP(hipGetLastError());
  gm::ReduceArgs r{};
  r.partials = f->pdot;
  r.G = f->G_pd;
  r.nv = R;
  r.out = f->pdot + (size_t)f->G_pd * R;
  r.stride = 1;
  hipLaunchKernelGGL(gm::k_reduce<0>, dim3(gm::reduce_grid(r.nv)), dim3(gm::kReduceBlock), 0, f->stream, r);
  F_HIP(hipGetLastError());
  return GM_OK;
}

gm_status fpp_dispatch(gm_bfab* f, const double* E) {
  switch (f->R) {
    case 2: return fpp<2>(f, E);
    case 4: return fpp<4>(f, E);
    case 8: return fpp<8>(f, E);
    case 16: return fpp<16>(f, E);
    case 32: return fpp<32>(f, E);
    default: return fpp<64>(f, E);
  }
}

// EQ = E Q ([N][K]) for E on A's pattern, and QEQ = Q^T E Q (host, K x K row-major).
gm_status feq(gm_bfab* f, const double* E, double* EQ, std::vector<double>& QEQ) {
  const int G1 = std::max(1, std::min(4096, bdiv(f->n, 256)));   // one column: 256 rows a pass
  const int chunk1 = bdiv(f->n, G1);
  for (int k = 0; k < f->K; ++k) {
    gmb::BSpmm a{};
    a.rowptr = f->rowptr;
    a.col = f->col;
    a.val = E;
    a.nrows = f->n;
    a.chunk = chunk1;
    a.x = f->Q + k;
    a.ldx = (size_t)f->K;
    a.y = EQ + k;
    a.ldy = (size_t)f->K;
    a.g0 = 1.0;
    a.ca = 1.0;
    hipLaunchKernelGGL((gmb::k_bspmm<1, gmb::kCheb>), dim3(bdiv(f->n, chunk1)), dim3(gm::kBlock), 0,
                       f->stream, a);
    F_HIP(hipGetLastError());
  }
  // Q^T (E Q): the Gram kernel with U = EQ (K columns)
  gmb::BDeflQtU q{};
  q.Q = f->Q;
  q.U = EQ;
  q.K = f->K;
  q.R = f->K;
  q.nrows = f->n;
  q.chunk = f->chunk_df;
  q.KT = bdiv(f->K, 16);
  q.CT = bdiv(f->K, 16);
  q.partials = f->partials;
  gmb::launch_bdefl_qtu(q, f->G_df, f->stream);
  F_HIP(hipGetLastError());
  F_TRY(freduce(f, f->G_df, f->K * f->K, f->ppart, 0, 0));
  QEQ.resize((size_t)f->K * f->K);
  F_HIP(hipMemcpyAsync(QEQ.data(), f->ppart, QEQ.size() * sizeof(double), hipMemcpyDeviceToHost,
                       f->stream));
  F_HIP(hipStreamSynchronize(f->stream));
  return GM_OK;
}

// Same pivoted elimination as the single-column engine: (shift I - T) h = r for
// the m x m tridiagonal of column c (al[k][R], be[k][R]).
void fshift_solve(int m, int R, int c, const double* al, const double* be, double shift,
                  std::vector<double>& r) {
  std::vector<double> d(m), du(m), dl(m), du2(m, 0.0);
  for (int i = 0; i < m; ++i) {
    d[i] = shift - al[(size_t)i * R + c];
    const double o = i + 1 < m ? -be[(size_t)(i + 1) * R + c] : 0.0;
    dl[i] = o;
    du[i] = o;
  }
  for (int i = 0; i + 1 < m; ++i) {
    if (std::fabs(d[i]) >= std::fabs(dl[i])) {
      const double fct = d[i] != 0.0 ? dl[i] / d[i] : 0.0;
      d[i + 1] -= fct * du[i];
      r[i + 1] -= fct * r[i];
    } else {
      const double fct = d[i] / dl[i];
      d[i] = dl[i];
      const double t = d[i + 1];
      d[i + 1] = du[i] - fct * t;
      if (i + 2 < m) {
        du2[i] = du[i + 1];
        du[i + 1] = -fct * du2[i];
      }
      du[i] = t;
      std::swap(r[i], r[i + 1]);
      r[i + 1] -= fct * r[i];
    }
  }
  for (int i = m - 1; i >= 0; --i) {
    double v = r[i];
    if (i + 1 < m) v -= du[i] * r[i + 1];
    if (i + 2 < m) v -= du2[i] * r[i + 2];
    r[i] = d[i] != 0.0 ? v / d[i] : 0.0;
  }
}

template <int KT>
void launch_bvtq_kt(const gmf::BVtQ& v, int G, hipStream_t st) {
  const dim3 grid(v.m, G);
  switch ((v.R + 15) / 16) {
    case 1: hipLaunchKernelGGL((gmf::k_bvtq<KT, 1>), grid, dim3(256), 0, st, v); break;
    case 2: hipLaunchKernelGGL((gmf::k_bvtq<KT, 2>), grid, dim3(256), 0, st, v); break;
    default: hipLaunchKernelGGL((gmf::k_bvtq<KT, 4>), grid, dim3(256), 0, st, v); break;
  }
}

void launch_bvtq(const gmf::BVtQ& v, int G, hipStream_t st) {
  switch ((v.K + 15) / 16) {
    case 1: launch_bvtq_kt<1>(v, G, st); break;
    case 2: launch_bvtq_kt<2>(v, G, st); break;
    case 3: launch_bvtq_kt<3>(v, G, st); break;
    default: launch_bvtq_kt<4>(v, G, st); break;
  }
}

// The engine stream waits for a b-side run that a forward pass without
// output left in flight.
gm_status fjoin(gm_bfab* f) {
  if (f->pend2) {
    F_HIP(hipStreamWaitEvent(f->stream, f->ev2, 0));
    f->pend2 = false;
  }
  return GM_OK;
}

gm_status fcheck_status(gm_bfab* f, const char* what) {
  unsigned stv = 0;
  F_HIP(hipMemcpyAsync(&stv, f->status, sizeof(unsigned), hipMemcpyDeviceToHost, f->stream));
  F_HIP(hipStreamSynchronize(f->stream));
  if (stv & 6u) return ffail(GM_EBREAKDOWN, "Lanczos breakdown in the %s (%u)", what, stv);
  return GM_OK;
}

}  // namespace

extern "C" {

const char* gm_bfab_last_error(void) { return g_ferr.c_str(); }

gm_status gm_bfab_create(const gm_bfab_desc* d, gm_bfab** out) {
  g_ferr.clear();
  if (!d || !out) return ffail(GM_EINVAL, "null argument");
  *out = nullptr;
  if (d->n <= 0 || !d->A_ptr || !d->A_ind || !d->A_values || d->subspace_size < 1 ||
      d->contour_size < 1 || !d->xi || !d->wts1 || !d->wts2)
    return ffail(GM_EINVAL, "bad sizes or null arrays");
  const int R = d->columns;
  if (R != 2 && R != 4 && R != 8 && R != 16 && R != 32 && R != 64)
    return ffail(GM_EINVAL, "columns must be 2, 4, 8, 16, 32 or 64");
  if (d->sparse_format != GM_SPARSE_CSR) return ffail(GM_EINVAL, "A must be CSR");
  if (d->nnz > INT32_MAX) return ffail(GM_EINVAL, "nnz beyond 32-bit row pointers");
  if (d->A_ptr[0] != 0 || d->A_ptr[d->n] != d->nnz) return ffail(GM_EINVAL, "bad pointer array");
  for (int i = 0; i < d->n; ++i)
    if (d->A_ptr[i + 1] < d->A_ptr[i]) return ffail(GM_EINVAL, "pointer array not monotone");
  for (long long j = 0; j < d->nnz; ++j)
    if (d->A_ind[j] < 0 || d->A_ind[j] >= d->n) return ffail(GM_EINVAL, "index out of range");
  if (d->num_eigs < 0 || d->num_eigs > 64 || d->num_eigs > d->n)
    return ffail(GM_EINVAL, "num_eigs must be in [0, min(64, n)]");
  if (d->num_eigs > 0 && (!d->Q || !d->lambda || !d->f1_lambda || !d->f2_lambda))
    return ffail(GM_EINVAL, "deflation needs Q, lambda, f1_lambda and f2_lambda");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= d->device || d->device < 0)
    return ffail(GM_EDEVICE, "no HIP device %d", d->device);
  gm_bfab* f = new gm_bfab();
  auto bail = [&](gm_status st) {
    std::string keep = g_ferr;
    free_bfab(f);
    delete f;
    g_ferr = keep;
    return st;
  };
#define F_CREATE(x)                  \
  do {                               \
    gm_status s_ = (x);              \
    if (s_ != GM_OK) return bail(s_); \
  } while (0)
  f->dev = d->device;
  if (hipSetDevice(f->dev) != hipSuccess ||
      hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(ffail(GM_EDEVICE, "device/stream"));
  f->n = d->n;
  f->nnz = d->nnz;
  f->R = R;
  f->m = d->subspace_size;
  f->ns = d->contour_size;
  f->K = d->num_eigs;
  const size_t N = (size_t)f->n;
  const int rows_per_pass = gm::kWaves * (R < 64 ? 64 / R : 1);
  // at most 1024 blocks: few block partials per Lanczos sum (as in the batched form)
  f->G_sp = std::max(1, std::min(1024, bdiv(f->n, rows_per_pass)));
  f->chunk_sp = bdiv(f->n, f->G_sp);
  f->G_sp = bdiv(f->n, f->chunk_sp);
  f->G_df = std::max(1, std::min(256, bdiv(f->n, 1024)));
  f->chunk_df = bdiv(f->n, f->G_df);
  f->G_df = bdiv(f->n, f->chunk_df);
  // T kernel: items (16-row tile, CW columns), l in tiles of 16 (at most 4 a pass)
  f->CW = R == 2 ? 2 : 4;
  f->LT = std::min(4, bdiv(f->m, 16));
  f->mk = bdiv(f->m, 4) * 4;
  f->mp = bdiv(f->m, 16 * f->LT) * 16 * f->LT;
  f->pp_items = bdiv(f->n, 16) * (R / f->CW);
  f->G_pd = bdiv(std::max(1, std::min(2048, bdiv(f->n, gm::kWaves * (R < 64 ? 64 / R : 1)))), 8) * 8;
  hipStream_t st = f->stream;
  auto up = [&](auto** p, const auto* src, size_t cnt) -> gm_status {
    F_TRY(falloc(p, cnt));
    F_HIP(hipMemcpyAsync(*p, src, cnt * sizeof(**p), hipMemcpyHostToDevice, st));
    return GM_OK;
  };
  F_CREATE(up(&f->rowptr, d->A_ptr, N + 1));
  F_CREATE(up(&f->col, d->A_ind, (size_t)f->nnz));
  F_CREATE(up(&f->val, d->A_values, (size_t)f->nnz));
  std::vector<int> rb;
  gmb::bspmm_row_blocks(d->A_ptr, f->n, rb);
  f->nrowblk = (int)rb.size() - 1;
  F_CREATE(up(&f->rowblk, rb.data(), rb.size()));
  std::vector<int> rbx;   // its own vector: rb's upload may still be in flight
  gmb::bspmm_xs_blocks(d->A_ptr, f->n, rbx);
  f->nrowblk_xs = (int)rbx.size() / 4;
  f->G_xs = gmb::bspmm_xs_grid(f->dev, f->nrowblk_xs);
  F_CREATE(up(&f->rowblk_xs, rbx.data(), rbx.size()));
  if (hipStreamSynchronize(st) != hipSuccess) return bail(ffail(GM_EDEVICE, "sync"));
  // nodes and weights: xi interleaved; weights ns x 2 [Re Im] for the solves
  std::vector<double> w1(2 * (size_t)f->ns), w2(2 * (size_t)f->ns), ones(2 * (size_t)f->ns, 0.0),
      gr(f->ns, 1.0), gi(f->ns, 0.0);
  for (int s = 0; s < f->ns; ++s) {
    w1[s] = d->wts1[2 * s];
    w1[s + f->ns] = d->wts1[2 * s + 1];
    w2[s] = d->wts2[2 * s];
    w2[s + f->ns] = d->wts2[2 * s + 1];
    ones[s] = 1.0;
  }
  F_CREATE(up(&f->xi, d->xi, 2 * (size_t)f->ns));
  F_CREATE(up(&f->wts1, w1.data(), w1.size()));
  F_CREATE(up(&f->wts2, w2.data(), w2.size()));
  F_CREATE(up(&f->ones, ones.data(), ones.size()));
  F_CREATE(up(&f->gsr, gr.data(), gr.size()));
  F_CREATE(up(&f->gsi, gi.data(), gi.size()));
  if (f->K) {
    std::vector<double> qr(N * f->K);   // column-major n x K -> row-major [N][K]
    for (int k = 0; k < f->K; ++k)
      for (size_t i = 0; i < N; ++i) qr[i * f->K + k] = d->Q[(size_t)k * N + i];
    F_CREATE(up(&f->Q, qr.data(), qr.size()));
    F_CREATE(up(&f->el, d->f1_lambda, (size_t)f->K));
    F_CREATE(up(&f->pl, d->f2_lambda, (size_t)f->K));
    f->lam.assign(d->lambda, d->lambda + f->K);
    f->el_h.assign(d->f1_lambda, d->f1_lambda + f->K);
    f->pl_h.assign(d->f2_lambda, d->f2_lambda + f->K);
    if (d->f1_divdiff) f->dd1.assign(d->f1_divdiff, d->f1_divdiff + (size_t)f->K * f->K);
    if (d->f2_divdiff) f->dd2.assign(d->f2_divdiff, d->f2_divdiff + (size_t)f->K * f->K);
    if (hipStreamSynchronize(st) != hipSuccess) return bail(ffail(GM_EDEVICE, "sync"));
  }
  const size_t ny = (size_t)f->m * R * f->ns;
  const size_t np = std::max<size_t>({(size_t)f->G_sp * R, (size_t)f->G_xs * R, (size_t)f->G_df * std::max(f->K, 1) * R,
                                      (size_t)f->G_df * std::max(f->K, 1) * std::max(f->K, 1),
                                      (size_t)f->G_df * f->m * std::max(f->K, 1) * R});
  F_CREATE(falloc(&f->partials, np));
  F_CREATE(falloc(&f->ppart, (size_t)f->m * std::max(f->K, 1) * R));
  F_CREATE(falloc(&f->wD, ny));
  F_CREATE(falloc(&f->wDU, ny));
  F_CREATE(falloc(&f->wDU2, ny));
  F_CREATE(falloc(&f->wX, ny));
  F_CREATE(falloc(&f->status, 1));
  if (hipMemsetAsync(f->status, 0, sizeof(unsigned), st) != hipSuccess)
    return bail(ffail(GM_EDEVICE, "memset"));
  for (FSide* s : {&f->sb, &f->sg}) {
    F_CREATE(falloc(&s->U, N * R));
    F_CREATE(falloc(&s->V, N * R * f->m));
    F_CREATE(falloc(&s->W0, N * R));
    F_CREATE(falloc(&s->W1, N * R));
    F_CREATE(falloc(&s->al, (size_t)f->m * R));
    F_CREATE(falloc(&s->be, (size_t)(f->m + 1) * R));
    F_CREATE(falloc(&s->yr, ny));
    F_CREATE(falloc(&s->yi, ny));
    F_CREATE(falloc(&s->ur, ny));
    F_CREATE(falloc(&s->ui, ny));
    F_CREATE(falloc(&s->C, (size_t)f->m * R));
    F_CREATE(falloc(&s->out, N * R));
    F_CREATE(falloc(&s->Zd, (size_t)std::max(f->K, 1) * R));
    F_CREATE(falloc(&s->upd, f->K ? N * R : 1));
  }
  // the b side runs on its own stream with its own partials and solve
  // scratch; the g side uses the engine's
  if (hipStreamCreateWithFlags(&f->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&f->ev1, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&f->ev2, hipEventDisableTiming) != hipSuccess)
    return bail(ffail(GM_EDEVICE, "b-side stream/events"));
  F_CREATE(falloc(&f->sb.partials, np));
  F_CREATE(falloc(&f->sb.wD, ny));
  F_CREATE(falloc(&f->sb.wDU, ny));
  F_CREATE(falloc(&f->sb.wDU2, ny));
  F_CREATE(falloc(&f->sb.wX, ny));
  f->sb.st = f->stream2;
  f->sg.st = f->stream;
  f->sg.partials = f->partials;
  f->sg.wD = f->wD;
  f->sg.wDU = f->wDU;
  f->sg.wDU2 = f->wDU2;
  f->sg.wX = f->wX;
  if (hipStreamSynchronize(st) != hipSuccess) return bail(ffail(GM_EDEVICE, "sync"));
  {
    std::lock_guard<std::mutex> lk(g_fmu);
    g_freg.insert(f);
  }
  *out = f;
  return GM_OK;
#undef F_CREATE
}

gm_status gm_bfab_delete(gm_bfab* f) {
  {
    std::lock_guard<std::mutex> lk(g_fmu);
    if (!f || !g_freg.count(f)) return ffail(GM_EBADHANDLE, "invalid bfab handle");
    g_freg.erase(f);
  }
  (void)hipSetDevice(f->dev);
  if (f->stream2) (void)hipStreamSynchronize(f->stream2);
  if (f->stream) (void)hipStreamSynchronize(f->stream);
  free_bfab(f);
  delete f;
  return GM_OK;
}

gm_status gm_bfab_forward(gm_bfab* f, const double* B, double* U) {
  if (!falive(f)) return ffail(GM_EBADHANDLE, "invalid bfab handle");
  if (!B) return ffail(GM_EINVAL, "null input");
  F_HIP(hipSetDevice(f->dev));
  const size_t NR = (size_t)f->n * f->R;
  // the b side (its own stream) after the engine stream's earlier work
  F_HIP(hipEventRecord(f->ev1, f->stream));
  F_HIP(hipStreamWaitEvent(f->stream2, f->ev1, 0));
  F_HIP(hipMemcpyAsync(f->sb.U, B, NR * sizeof(double), hipMemcpyDefault, f->stream2));
  if (!U) {
    // B is the caller's until this returns: wait for its copy only, and leave
    // the run in flight; the next entry point joins it, so a backward's g-side
    // run overlaps it.  Its breakdown flags are checked with that call's.
    F_HIP(hipStreamSynchronize(f->stream2));
  }
  if (f->K) F_TRY(fdeflate(f, f->sb));
  F_TRY(fkrylov_dispatch(f, f->sb, U != nullptr, false));
  if (U) F_HIP(hipMemcpyAsync(U, f->sb.out, NR * sizeof(double), hipMemcpyDefault, f->stream2));
  F_HIP(hipEventRecord(f->ev2, f->stream2));
  f->have_forward = true;
  if (!U) {
    f->pend2 = true;
    return GM_OK;
  }
  F_HIP(hipStreamWaitEvent(f->stream, f->ev2, 0));
  f->pend2 = false;   // the engine stream now waits for all b-side work
  return fcheck_status(f, "forward pass");
}

gm_status gm_bfab_set_param_derivs(gm_bfab* f, int32_t nparams, const double* const* dA_values) {
  if (!falive(f)) return ffail(GM_EBADHANDLE, "invalid bfab handle");
  if (nparams < 0 || (nparams > 0 && !dA_values)) return ffail(GM_EINVAL, "bad parameter arrays");
  for (int q = 0; q < nparams; ++q)
    if (!dA_values[q]) return ffail(GM_EINVAL, "null dA/dtheta values for parameter %d", q);
  if (f->K && nparams > 0 && (f->dd1.empty() || f->dd2.empty()))
    return ffail(GM_EINVAL, "the parameter gradient with deflation needs f1_divdiff and f2_divdiff");
  F_HIP(hipSetDevice(f->dev));
  F_TRY(fjoin(f));
  F_HIP(hipStreamSynchronize(f->stream));
  for (double* p : f->pdA) ffree(p);
  for (double* p : f->pEQ) ffree(p);
  f->pdA.clear();
  f->pEQ.clear();
  f->pQEQ.clear();
  for (int q = 0; q < nparams; ++q) {
    double* E = nullptr;
    F_TRY(falloc(&E, (size_t)f->nnz));
    f->pdA.push_back(E);
    F_HIP(hipMemcpyAsync(E, dA_values[q], (size_t)f->nnz * sizeof(double), hipMemcpyHostToDevice,
                         f->stream));
    if (f->K) {
      double* EQ = nullptr;
      F_TRY(falloc(&EQ, (size_t)f->n * f->K));
      f->pEQ.push_back(EQ);
      f->pQEQ.emplace_back();
      F_TRY(feq(f, E, EQ, f->pQEQ.back()));
    }
  }
  F_HIP(hipStreamSynchronize(f->stream));
  return GM_OK;
}

gm_status gm_bfab_backward_param(gm_bfab* f, const double* G, double* grad_B, double* grad_theta) {
  if (!falive(f)) return ffail(GM_EBADHANDLE, "invalid bfab handle");
  if (!f->have_forward) return ffail(GM_ESTATE, "backward needs a forward pass first");
  if (!G) return ffail(GM_EINVAL, "null upstream gradient");
  const int P = (int)f->pdA.size();
  if (P > 0 && !grad_theta) return ffail(GM_EINVAL, "null grad_theta");
  F_HIP(hipSetDevice(f->dev));
  const int R = f->R, m = f->m, K = f->K;
  const size_t N = (size_t)f->n, NR = N * R;
  hipStream_t st = f->stream;
  F_HIP(hipMemcpyAsync(f->sg.U, G, NR * sizeof(double), hipMemcpyDefault, st));
  if (K) F_TRY(fdeflate(f, f->sg));
  F_TRY(fkrylov_dispatch(f, f->sg, grad_B != nullptr, P > 0));
  // everything below reads the b side: wait for its run if a forward pass
  // without output left it in flight (it overlapped the g side's run above)
  F_TRY(fjoin(f));
  if (grad_B) F_HIP(hipMemcpyAsync(grad_B, f->sg.out, NR * sizeof(double), hipMemcpyDefault, st));
  if (P > 0) {
    F_TRY(fppm(f));
    // deflation data of both sides, and H per column (host)
    std::vector<double> Hb, Hg, ab, ag;
    if (K) {
      for (int side = 0; side < 2; ++side) {
        FSide& sd = side ? f->sg : f->sb;
        std::vector<double> al((size_t)m * R), be((size_t)(m + 1) * R), Cv((size_t)m * R),
            a((size_t)K * R);
        F_HIP(hipMemcpyAsync(al.data(), sd.al, al.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        F_HIP(hipMemcpyAsync(be.data(), sd.be, be.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        F_HIP(hipMemcpyAsync(Cv.data(), sd.C, Cv.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        F_HIP(hipMemcpyAsync(a.data(), sd.Zd, a.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        F_HIP(hipStreamSynchronize(st));
        std::vector<double> H((size_t)R * m * K), r(m);   // [c][j][k]
        for (int c = 0; c < R; ++c) {
          const std::vector<double>& fl = (c & 1) ? f->pl_h : f->el_h;
          for (int k = 0; k < K; ++k) {
            for (int j = 0; j < m; ++j) r[j] = -Cv[(size_t)j * R + c];
            r[0] += fl[k] * be[c];
            fshift_solve(m, R, c, al.data(), be.data(), f->lam[k], r);
            for (int j = 0; j < m; ++j) H[((size_t)c * m + j) * K + k] = r[j];
          }
        }
        (side ? Hg : Hb).swap(H);
        (side ? ag : ab).swap(a);
      }
    }
    for (int q = 0; q < P; ++q) {
      F_TRY(fpp_dispatch(f, f->pdA[q]));
      F_HIP(hipMemcpyAsync(grad_theta + (size_t)q * R, f->pdot + (size_t)f->G_pd * R, R * sizeof(double),
                           hipMemcpyDeviceToHost, st));
      F_HIP(hipStreamSynchronize(st));
      if (K) {
        std::vector<double> Mb, Mg;
        for (int side = 0; side < 2; ++side) {
          FSide& sd = side ? f->sg : f->sb;
          gmf::BVtQ v{};
          v.V = sd.V;
          v.EQ = f->pEQ[q];
          v.K = K;
          v.R = R;
          v.m = m;
          v.nrows = f->n;
          v.chunk = f->chunk_df;
          v.partials = f->partials;
          launch_bvtq(v, f->G_df, st);
          F_HIP(hipGetLastError());
          F_TRY(freduce(f, f->G_df, m * K * R, f->ppart, 0, 0));
          std::vector<double>& M = side ? Mg : Mb;   // [j][k][c]
          M.resize((size_t)m * K * R);
          F_HIP(hipMemcpyAsync(M.data(), f->ppart, M.size() * sizeof(double), hipMemcpyDeviceToHost, st));
          F_HIP(hipStreamSynchronize(st));
        }
        const std::vector<double>& QEQ = f->pQEQ[q];
        for (int c = 0; c < R; ++c) {
          const std::vector<double>& F = (c & 1) ? f->dd2 : f->dd1;
          double g = 0.0;
          for (int k = 0; k < K; ++k)
            for (int l = 0; l < K; ++l)
              g += ag[(size_t)k * R + c] * F[(size_t)l * K + k] * QEQ[(size_t)k * K + l] *
                   ab[(size_t)l * R + c];
          for (int k = 0; k < K; ++k) {
            double tb = 0.0, tg = 0.0;
            for (int j = 0; j < m; ++j) {
              tb += Mb[((size_t)j * K + k) * R + c] * Hb[((size_t)c * m + j) * K + k];
              tg += Mg[((size_t)j * K + k) * R + c] * Hg[((size_t)c * m + j) * K + k];
            }
            g += ag[(size_t)k * R + c] * tb + ab[(size_t)k * R + c] * tg;
          }
          grad_theta[(size_t)q * R + c] += g;
        }
      }
    }
  }
  return fcheck_status(f, "backward pass");
}

gm_status gm_bfab_info(gm_bfab* f, int32_t* columns, int32_t* num_params) {
  if (!falive(f)) return ffail(GM_EBADHANDLE, "invalid bfab handle");
  if (columns) *columns = f->R;
  if (num_params) *num_params = (int32_t)f->pdA.size();
  return GM_OK;
}

}  // extern "C"
