// gm_kernels.h -- gfx950 device kernels of the f(A)b engine (FP64 throughout).
//
// Layout in HBM (chosen for 64-wide wavefronts, not the reference's):
//   * the two columns of every N x 2 quantity (exp path, phi path) are
//     interleaved as double2, so a CSR gather fetches both columns of x[col]
//     with one 16-byte load;
//   * the Lanczos basis is V[k][i] (double2), k = 0..m-1, each slice
//     contiguous in i: the three-term recurrence and the V*C recombination
//     stream it with fully coalesced loads;
//   * A is CSR (int32 row pointers and columns, FP64 values);
//   * Q is stored row-major (N x K) so a thread reads one row's K entries
//     contiguously in both deflation passes;
//   * aabb, Yreal, Yimag keep the reference's column-major layout
//     (lanczos.h:51-55) so the getters return them unchanged.
//
// Reductions (dot products, norms, Q^T U) are deterministic: each block
// reduces its rows in a fixed tree and writes one partial; k_reduce, launched
// after the kernel, sums the partials in a fixed order.  Results do not
// depend on scheduling, so repeated runs are bit-identical.
#pragma once
#include "gm_device.h"

// Every kernel here is a template (kUnused where nothing else varies) so that
// several translation units may include this header without duplicate symbols.


namespace gm {

constexpr int kMaxDegree = 16;        // gamma_size <= 17
constexpr int kDeflTile = 8;          // deflation vectors per register tile

// ---------------------------------------------------------------------------
// CSR SpMM with R = 2 interleaved columns and fused Lanczos / Horner
// epilogues, in two schedules (row groups, row blocks) picked per matrix.
// ---------------------------------------------------------------------------
enum SpmmMode : int {
  kFirst = 0,   // v_k = w/beta_k (stored); y = A(g0 v_k) + g1 v_k
  kMid = 1,     // y = A x + g_i v_k
  kLast = 2,    // y = A x - beta_k v_{k-1};  alpha_k = y . v_k
  kOnly = 3,    // (degree 0) v_k = w/beta_k; y = A(g0 v_k) - beta_k v_{k-1}; alpha_k
  kHorner = 4   // y = A x + z  (optionally y += previous y)
};

struct SpmmArgs {
  const int* __restrict__ rowptr;
  const int* __restrict__ col;
  const double* __restrict__ val;
  const int* __restrict__ rowblk;  // row-block bounds (k_spmm_stream only)
  const int* __restrict__ blkj;    // rowptr at each row-block bound (k_spmm_stream only)
  const short* __restrict__ col16;  // C16: column - first row of the nonzero's row block
  const int* __restrict__ blk_prob;  // SEG: problem of each row block
  int aabb_stride;             // SEG: doubles between problems' aabb
  int nrows;
  int chunk;                   // rows per block (k_spmm only)
  const double2* x;            // gather source
  double2* y;                  // output (row aligned)
  double2* vk;                 // V[k]: written by kFirst/kOnly, read by kMid/kLast
  const double2* vkm1;         // V[k-1] or null (k == 0)
  const double2* z;            // kHorner addend
  const double* aabb;          // (m+1) x 4 column-major
  int ld;                      // m + 1
  int k;                       // iteration
  double g0, gi;               // gamma_0, gamma_i
  int accumulate;              // kHorner: y += result instead of y = result
  double* partials;
  unsigned* status;            // bit 0: beta = 0 seen, bit 2: non-finite beta
  // kFirst/kOnly with bpart: beta_k = sqrt(sum of the previous finish's bG
  // partials), formed in every block (no separate reduction launch); block 0
  // stores it at beta_out[0] / beta_out[ld] and flags it in status
  const double* bpart;
  int bG;
  double* beta_out;
};

// Where kFirst/kOnly take beta_k from when it is formed in the kernel.
struct FusedBeta {
  const double* part;   // null: read beta_k from aabb
  int G;
  double* out;
  bool writer;
};

// Per-launch constants of the row epilogue: 1/beta_k, beta_k (kFirst, kOnly),
// beta_k for the -beta_k v_{k-1} term (kLast, kOnly), and the scale of A x.
// VW: the basis holds the unnormalised w_j (the segmented form): v_k = w_k /
// beta_k and v_{k-1} are formed where read, the bits the stored v_k would
// have, and kFirst / kOnly store nothing (vk = null) or, at k = 0, the raw w_0.
struct SpmmScalars {
  double2 inv, bet, bprev, sx;
  double2 bk, bkm1;   // VW: beta_k (kMid, kLast) and beta_{k-1} (kLast, kOnly)
};

template <int MODE, bool VW = false>
__device__ __forceinline__ SpmmScalars spmm_scalars(const SpmmArgs& p, const double* aabb,
                                                    const FusedBeta& fb) {
  SpmmScalars c;
  c.inv = d2(1.0, 1.0);
  c.bet = d2(0.0, 0.0);
  c.bprev = d2(0.0, 0.0);
  if (MODE == kFirst || MODE == kOnly) {
    double b0, b1;
    if (fb.part) {
      const double2 t = block_reduce_partials2(fb.part, fb.G);
      b0 = sqrt(t.x);
      b1 = sqrt(t.y);
      if (fb.writer && threadIdx.x == 0) {
        fb.out[0] = b0;
        fb.out[p.ld] = b1;
        if (!isfinite(b0) || !isfinite(b1)) atomicOr(p.status, 4u);
        else if (b0 == 0.0 || b1 == 0.0) atomicOr(p.status, 1u);
      }
    } else {
      b0 = aabb[p.k + 2 * p.ld];
      b1 = aabb[p.k + 3 * p.ld];
    }
    c.inv = d2(safe_inv(b0), safe_inv(b1));
    c.bet = d2(b0, b1);
  }
  if ((MODE == kLast || MODE == kOnly) && p.k > 0)
    c.bprev = (MODE == kOnly && fb.part) ? c.bet
                                         : d2(aabb[p.k + 2 * p.ld], aabb[p.k + 3 * p.ld]);
  c.sx = (MODE == kFirst || MODE == kOnly) ? d2(p.g0 * c.inv.x, p.g0 * c.inv.y) : d2(1.0, 1.0);
  c.bk = d2(1.0, 1.0);
  c.bkm1 = d2(1.0, 1.0);
  if (VW && (MODE == kMid || MODE == kLast)) c.bk = d2(aabb[p.k + 2 * p.ld], aabb[p.k + 3 * p.ld]);
  if (VW && (MODE == kLast || MODE == kOnly) && p.k > 0)
    c.bkm1 = d2(aabb[p.k - 1 + 2 * p.ld], aabb[p.k - 1 + 3 * p.ld]);
  return c;
}

// Row r's operands besides A x: w_r (kFirst, kOnly), v_k,r (kMid, kLast),
// v_{k-1},r (kLast, kOnly with k > 0), z_r and the old y_r (kHorner).
template <int MODE>
__device__ __forceinline__ void spmm_row_load(const SpmmArgs& p, int r, double2& o0, double2& o1) {
  o0 = d2(0.0, 0.0);
  o1 = d2(0.0, 0.0);
  if (MODE == kFirst || MODE == kOnly) o0 = p.x[r];
  if (MODE == kMid || MODE == kLast) o0 = p.vk[r];
  if ((MODE == kLast || MODE == kOnly) && p.k > 0) o1 = p.vkm1[r];
  if (MODE == kHorner) {
    o0 = p.z[r];
    if (p.accumulate) o1 = p.y[r];
  }
}

// Row r's output from acc = (A x)_r and the operands of spmm_row_load;
// kLast/kOnly add y_r v_r to dot.
template <int MODE, bool VW = false>
__device__ __forceinline__ void spmm_row_store(const SpmmArgs& p, const SpmmScalars& c, int r,
                                               double2 acc, double2 o0, double2 o1, double2& dot) {
  if (VW) {
    if (MODE == kMid || MODE == kLast) o0 = d2(safe_div(o0.x, c.bk.x), safe_div(o0.y, c.bk.y));
    if ((MODE == kLast || MODE == kOnly) && p.k > 0)
      o1 = d2(safe_div(o1.x, c.bkm1.x), safe_div(o1.y, c.bkm1.y));
  }
  if (MODE == kFirst) {
    const double2 v = d2(safe_div(o0.x, c.bet.x), safe_div(o0.y, c.bet.y));
    if (!VW) p.vk[r] = v;
    else if (p.vk) p.vk[r] = o0;
    p.y[r] = d2(fma(p.gi, v.x, acc.x * c.sx.x), fma(p.gi, v.y, acc.y * c.sx.y));
  } else if (MODE == kMid) {
    p.y[r] = d2(fma(p.gi, o0.x, acc.x), fma(p.gi, o0.y, acc.y));
  } else if (MODE == kLast || MODE == kOnly) {
    double2 v;
    if (MODE == kOnly) {
      v = d2(safe_div(o0.x, c.bet.x), safe_div(o0.y, c.bet.y));
      if (!VW) p.vk[r] = v;
      else if (p.vk) p.vk[r] = o0;
      acc.x *= c.sx.x;
      acc.y *= c.sx.y;
    } else {
      v = o0;
    }
    double2 yv = acc;
    if (p.k > 0) {
      yv.x = fma(-c.bprev.x, o1.x, yv.x);
      yv.y = fma(-c.bprev.y, o1.y, yv.y);
    }
    p.y[r] = yv;
    dot.x = fma(yv.x, v.x, dot.x);
    dot.y = fma(yv.y, v.y, dot.y);
  } else {  // kHorner
    double2 yv = d2(acc.x + o0.x, acc.y + o0.y);
    if (p.accumulate) {
      yv.x += o1.x;
      yv.y += o1.y;
    }
    p.y[r] = yv;
  }
}

template <int MODE, bool VW = false>
__device__ __forceinline__ void spmm_row_out(const SpmmArgs& p, const SpmmScalars& c, int r,
                                             double2 acc, double2& dot) {
  double2 o0, o1;
  spmm_row_load<MODE>(p, r, o0, o1);
  spmm_row_store<MODE, VW>(p, c, r, acc, o0, o1, dot);
}

// Row-group schedule: LANES consecutive lanes share one row (coalesced
// val/col loads, butterfly reduction inside the lane group); each block owns
// a contiguous chunk of rows.  Suits rows of tens of nonzeros and more.
template <int LANES, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmm(SpmmArgs p) {
  constexpr int GROUPS = kBlock / LANES;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int row0 = b * p.chunk;
  const int row1 = min(row0 + p.chunk, p.nrows);
  const int sub = threadIdx.x % LANES;
  const int grp = threadIdx.x / LANES;
  const FusedBeta fb{p.bpart, p.bG, p.beta_out, blockIdx.x == 0};
  const SpmmScalars c = spmm_scalars<MODE>(p, p.aabb, fb);
  double2 dot = d2(0.0, 0.0);

  for (int r = row0 + grp; r < row1; r += GROUPS) {
    const int j0 = p.rowptr[r], j1 = p.rowptr[r + 1];
    double2 acc = d2(0.0, 0.0);
    for (int j = j0 + sub; j < j1; j += LANES) {
      const double a = p.val[j];
      const double2 xv = p.x[p.col[j]];
      acc.x = fma(a, xv.x, acc.x);
      acc.y = fma(a, xv.y, acc.y);
    }
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) {
      acc.x += __shfl_xor(acc.x, off, LANES);
      acc.y += __shfl_xor(acc.y, off, LANES);
    }
    if (sub == 0) spmm_row_out<MODE>(p, c, r, acc, dot);
  }

  // alpha_k = sum of these partials, formed by k_reduce into aabb(k, 0:1)
  if (MODE == kLast || MODE == kOnly) block_partial2(dot, p.partials);
}

// Row-block ("CSR-stream") schedule for short rows.  The host cuts the rows
// into blocks of at most kBlock rows and at most TILE nonzeros (a row longer
// than TILE forms a block of its own).  Phase 1 streams the block's
// nonzeros with all threads -- consecutive lanes take consecutive nonzeros,
// so the val/col loads are fully coalesced and every thread has TILE/kBlock
// independent gathers in flight -- and stages the products a_j x[col_j] in
// LDS.  Phase 2 gives each row to one thread, which sums its products in
// column order and runs the epilogue; the row-aligned loads and stores of
// the epilogue are coalesced too.  One long row is summed by the whole block
// instead (strided FMAs, then a fixed-order block tree).
// SEG: the rows belong to independent problems, no block straddles two, and
// each block takes its Lanczos scalars from its problem's aabb and writes its
// partial at its row-block index (so a problem's partials are contiguous);
// the basis holds the unnormalised w_k (VW of spmm_row_store).
// C16: column indices are stored as 16-bit offsets from the block's first row
// (banded orderings such as symrcm keep them small), 2 bytes per nonzero
// instead of 4.  PRE: the epilogue's own-row operands are loaded before the
// gathers (their latency overlaps phase 1) instead of after the LDS sums;
// which is faster depends on the matrix, so the autotuner times both.
template <int TILE, int MODE, bool SEG = false, bool C16 = false, bool PRE = true>
__global__ __launch_bounds__(kBlock) void k_spmm_stream(SpmmArgs p) {
  static_assert(TILE % kBlock == 0, "TILE must be a multiple of the block size");
  constexpr int PER = TILE / kBlock;
  __shared__ double2 prod[TILE];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  // the block's row and nonzero ranges are four independent loads, so the
  // nonzero stream starts one round trip after launch, not two
  const int r0 = p.rowblk[b], r1 = p.rowblk[b + 1];
  const int j0 = p.blkj[b], j1 = p.blkj[b + 1];
  const int t = threadIdx.x;
  const double* aabb = SEG ? p.aabb + (size_t)p.blk_prob[b] * p.aabb_stride : p.aabb;
  const FusedBeta fb{SEG ? nullptr : p.bpart, p.bG, p.beta_out, blockIdx.x == 0};
  const SpmmScalars c = spmm_scalars<MODE, SEG>(p, aabb, fb);
  double2 dot = d2(0.0, 0.0);

  if (j1 - j0 <= TILE) {
    const int r = r0 + t;
    int a0 = 0, a1 = 0;
    double2 o0 = d2(0.0, 0.0), o1 = d2(0.0, 0.0);
    if (r < r1) {
      a0 = p.rowptr[r] - j0;
      a1 = p.rowptr[r + 1] - j0;
      if (PRE) spmm_row_load<MODE>(p, r, o0, o1);
    }
    // Slots past the block's last nonzero load that nonzero again (the same
    // line for every such lane) and are dropped below, so the PER index loads
    // and then the PER gathers each go out together, without a branch per
    // slot (with one, each index load waited for its predecessor).  A block
    // with no nonzeros loads nothing.
    double a[PER];
    int ci[PER];
    double2 xv[PER];
    if (j1 > j0) {
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int j = min(j0 + t + u * kBlock, j1 - 1);
        a[u] = p.val[j];
        ci[u] = C16 ? r0 + (int)p.col16[j] : p.col[j];
      }
#pragma unroll
      for (int u = 0; u < PER; ++u) xv[u] = p.x[ci[u]];
    } else {
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        a[u] = 0.0;
        xv[u] = d2(0.0, 0.0);
      }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int j = j0 + t + u * kBlock;
      if (j < j1) prod[t + u * kBlock] = d2(a[u] * xv[u].x, a[u] * xv[u].y);
    }
    __syncthreads();
    if (r < r1) {
      double2 acc = d2(0.0, 0.0);
      for (int i = a0; i < a1; ++i) {
        const double2 q = prod[i];
        acc.x += q.x;
        acc.y += q.y;
      }
      if (!PRE) spmm_row_load<MODE>(p, r, o0, o1);
      spmm_row_store<MODE, SEG>(p, c, r, acc, o0, o1, dot);
    }
  } else {
    double2 acc = d2(0.0, 0.0);
    for (int j = j0 + t; j < j1; j += kBlock) {
      const double av = p.val[j];
      const double2 xv = p.x[C16 ? r0 + (int)p.col16[j] : p.col[j]];
      acc.x = fma(av, xv.x, acc.x);
      acc.y = fma(av, xv.y, acc.y);
    }
    acc = block_sum2(acc);
    if (t == 0) spmm_row_out<MODE, SEG>(p, c, r0, acc, dot);
  }

  if (MODE == kLast || MODE == kOnly) block_partial2_at(dot, p.partials, SEG ? b : (int)blockIdx.x);
}

// w <- w - alpha_k v_k (in place), beta_{k+1} = ||w||.   build_subspace.cu:63-76
// With apart, alpha_k = sum of the SpMM's aG partials is formed in every block
// and block 0 stores it in aabb (no separate reduction launch).
struct FinishArgs {
  double2* w;
  const double2* vk;
  double* aabb;
  int ld, k, nrows, chunk;
  double* partials;
  unsigned* status;
  const double* apart;
  int aG;
};

template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_lanczos_finish(FinishArgs p) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int row0 = b * p.chunk;
  const int row1 = min(row0 + p.chunk, p.nrows);
  double2 al;
  if (p.apart) {
    al = block_reduce_partials2(p.apart, p.aG);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      p.aabb[p.k] = al.x;
      p.aabb[p.k + p.ld] = al.y;
    }
  } else {
    al = d2(p.aabb[p.k], p.aabb[p.k + p.ld]);
  }
  double2 nrm = d2(0.0, 0.0);
  for (int r = row0 + threadIdx.x; r < row1; r += kBlock) {
    const double2 v = p.vk[r];
    double2 w = p.w[r];
    w.x = fma(-al.x, v.x, w.x);
    w.y = fma(-al.y, v.y, w.y);
    p.w[r] = w;
    nrm.x = fma(w.x, w.x, nrm.x);
    nrm.y = fma(w.y, w.y, nrm.y);
  }
  // beta_{k+1} = sqrt(sum of these partials), formed by k_reduce into aabb(k+1, 2:3)
  block_partial2(nrm, p.partials);
}

// ---------------------------------------------------------------------------
// Shifted complex tridiagonal solves, one lane per system, LAPACK ZGTSV-style
// partial pivoting (what the CPU path's sparse backslash runs,
// nicksmethodcpu.m:19-20).  System s < n_s uses column 0 (exp), s >= n_s
// column 1 (phi); (T - shift_s I) y = -beta_0 w_s e1 as in
// lin_sys_solves.cu:5-21.  Workspaces are [row][system] so a wavefront's
// accesses are contiguous.  LDS = 1: a block solves p.sb systems (lanes
// 0..sb-1) and keeps the elimination's factors in LDS (4 m sb double2, sized
// by the launch) instead of the global workspaces, so the back substitution
// reads them at LDS latency; the arithmetic is the same either way.
// ---------------------------------------------------------------------------

struct SolveArgs {
  const double* aabb;
  int ld, m, ns;
  const double* wts1;   // ns x 2
  const double* wts2;
  const double* shifts; // 4 m ns, complex interleaved, system s at s*m
  double2* wD;
  double2* wDU;
  double2* wDU2;
  double2* wX;
  double* yr;           // m x 2ns column-major
  double* yi;
  unsigned* status;     // bit 1: singular pivot
  // independent problems (blockIdx.y = problem): element strides between
  // consecutive problems' aabb, wts1/wts2, shifts, yr/yi and workspaces
  size_t aabb_ps, wts_ps, shifts_ps, y_ps, work_ps;
  int sb;               // LDS form: systems per block
};

// Systems per block for the LDS form of k_shifted_solves (factors 64 m bytes a
// system, within kSolveLds bytes), or 0 when one system's factors do not fit.
constexpr int kSolveLds = 48 * 1024;
inline int solve_systems_per_block(int m) {
  const int sb = kSolveLds / (64 * (m > 0 ? m : 1));
  return sb < 1 ? 0 : (sb > 64 ? 64 : sb);
}

template <int LDS>
__global__ __launch_bounds__(64) void k_shifted_solves(SolveArgs p) {
  extern __shared__ double2 sfac[];
  constexpr bool lds = LDS != 0;
  const int t = threadIdx.x;
  const int s = lds ? blockIdx.x * p.sb + t : blockIdx.x * blockDim.x + t;
  const int S = 2 * p.ns;
  if (s >= S || (lds && t >= p.sb)) return;
  const size_t q = blockIdx.y;
  p.aabb += q * p.aabb_ps;
  p.wts1 += q * p.wts_ps;
  p.wts2 += q * p.wts_ps;
  p.shifts += q * p.shifts_ps;
  p.yr += q * p.y_ps;
  p.yi += q * p.y_ps;
  p.wD += q * p.work_ps;
  p.wDU += q * p.work_ps;
  p.wDU2 += q * p.work_ps;
  p.wX += q * p.work_ps;
  // factor k of this system: LDS [f][k][t] (f = D, DU, DU2, X), or global [k][s]
  const size_t fs = lds ? (size_t)p.m * p.sb : 0;
  double2* const fD = lds ? sfac : p.wD;
  double2* const fDU = lds ? sfac + fs : p.wDU;
  double2* const fDU2 = lds ? sfac + 2 * fs : p.wDU2;
  double2* const fX = lds ? sfac + 3 * fs : p.wX;
  auto fo = [&](int k) { return lds ? (size_t)k * p.sb + t : (size_t)k * S + s; };
  const int c = s < p.ns ? 0 : 1;
  const int node = s - c * p.ns;
  const double* w = c == 0 ? p.wts1 : p.wts2;
  const double b0 = p.aabb[2 * p.ld + c * p.ld];
  const double2 wt = d2(w[node], w[node + p.ns]);
  const double* al = p.aabb + c * p.ld;
  const double* be = p.aabb + (2 + c) * p.ld;
  const double* sh = p.shifts + 2 * (size_t)s * p.m;
  const int m = p.m;
  auto diag = [&](int k) { return d2(al[k] - sh[2 * k], -sh[2 * k + 1]); };
  auto off = [&](int k) { return d2(be[k + 1], 0.0); };  // T(k+1,k) = T(k,k+1)

  // Both sweeps load the operands of kSolveSteps steps together (the steps
  // themselves are unchanged), so the dependent chain waits on memory once per
  // kSolveSteps steps rather than on every step.
  constexpr int U = kSolveSteps;
  double2 cur_d = diag(0);
  double2 cur_x = d2(-(b0 * wt.x), -(b0 * wt.y));
  double2 cur_du = m > 1 ? off(0) : d2(0.0, 0.0);
  bool singular = false;
  for (int k0 = 0; k0 < m - 1; k0 += U) {
    double2 dv[U];     // diag(k0 + 1 + u)
    double bv[U + 1];  // be[k0 + 1 + u]
#pragma unroll
    for (int u = 0; u < U; ++u) dv[u] = k0 + 1 + u < m ? diag(k0 + 1 + u) : d2(0.0, 0.0);
#pragma unroll
    for (int u = 0; u <= U; ++u) bv[u] = k0 + 1 + u < m ? be[k0 + 1 + u] : 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u;
      if (k >= m - 1) break;
      const double2 dl = d2(bv[u], 0.0);                                        // off(k)
      double2 next_d = dv[u];                                                   // diag(k + 1)
      double2 next_x = d2(0.0, 0.0);
      double2 next_du = (k + 1 < m - 1) ? d2(bv[u + 1], 0.0) : d2(0.0, 0.0);   // off(k + 1)
      double2 sD, sDU, sDU2 = d2(0.0, 0.0), sX;
      if (dl.x == 0.0 && dl.y == 0.0) {
        if (cur_d.x == 0.0 && cur_d.y == 0.0) singular = true;
        sD = cur_d; sDU = cur_du; sX = cur_x;
      } else if (cabs1(cur_d) >= cabs1(dl)) {
        const double2 mult = cdiv(dl, cur_d);
        next_d = csub(next_d, cmul(mult, cur_du));
        next_x = csub(next_x, cmul(mult, cur_x));
        sD = cur_d; sDU = cur_du; sX = cur_x;
      } else {
        const double2 mult = cdiv(cur_d, dl);
        sD = dl;
        const double2 temp = next_d;
        next_d = csub(cur_du, cmul(mult, temp));
        if (k < m - 2) {
          sDU2 = next_du;
          next_du = d2(-(mult.x * sDU2.x - mult.y * sDU2.y), -(mult.x * sDU2.y + mult.y * sDU2.x));
        }
        sDU = temp;
        sX = next_x;
        next_x = csub(cur_x, cmul(mult, next_x));
      }
      const size_t o = fo(k);
      fD[o] = sD; fDU[o] = sDU; fDU2[o] = sDU2; fX[o] = sX;
      cur_d = next_d; cur_x = next_x; cur_du = next_du;
    }
  }
  if (cur_d.x == 0.0 && cur_d.y == 0.0) singular = true;
  if (singular) {
    atomicOr(p.status, 2u);
    for (int k = 0; k < m; ++k) { p.yr[k + (size_t)m * s] = 0.0; p.yi[k + (size_t)m * s] = 0.0; }
    return;
  }
  // back substitution
  double2 x2 = cdiv(cur_x, cur_d);           // x_{m-1}
  p.yr[(m - 1) + (size_t)m * s] = x2.x;
  p.yi[(m - 1) + (size_t)m * s] = x2.y;
  if (m > 1) {
    const size_t o = fo(m - 2);
    double2 x1 = cdiv(csub(fX[o], cmul(fDU[o], x2)), fD[o]);
    p.yr[(m - 2) + (size_t)m * s] = x1.x;
    p.yi[(m - 2) + (size_t)m * s] = x1.y;
    for (int k1 = m - 3; k1 >= 0; k1 -= U) {
      double2 bX[U], bDU[U], bDU2[U], bD[U];   // step k = k1 - u
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (k1 - u >= 0) {
          const size_t o2 = fo(k1 - u);
          bX[u] = fX[o2]; bDU[u] = fDU[o2]; bDU2[u] = fDU2[o2]; bD[u] = fD[o2];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k1 - u;
        if (k < 0) break;
        double2 tt = csub(bX[u], cmul(bDU[u], x1));
        tt = csub(tt, cmul(bDU2[u], x2));
        const double2 xk = cdiv(tt, bD[u]);
        p.yr[k + (size_t)m * s] = xk.x;
        p.yi[k + (size_t)m * s] = xk.y;
        x2 = x1;
        x1 = xk;
      }
    }
  }
}

// The same systems, one wavefront per system: parallel cyclic reduction
// (Hockney's odd-even reduction applied to every equation at once).  Lane l
// holds equations l + 64 j (j < Q <= 2, m <= 64 Q), the rows past m padded as
// decoupled identity rows.  Step h eliminates the couplings at distance h
// with the neighbours' rows at i - h and i + h (wave shuffles; a row outside
// [0, 64 Q) is an identity row); after the steps up to h >= 64 Q every row is
// decoupled and y_i = d_i / b_i.  No pivoting: for T real symmetric and the
// shift z off the real axis, or real and outside T's spectrum, the reduced
// diagonals stay away from zero (on the contour systems of the test problems
// the solutions agree with the pivoted elimination to 2e-13).  The steps are
// unrolled so every register index is a constant.
template <int Q>
__global__ __launch_bounds__(64) void k_shifted_solves_pcr(SolveArgs p) {
  const int s = blockIdx.x;
  const int S = 2 * p.ns;
  if (s >= S) return;
  const size_t q = blockIdx.y;
  p.aabb += q * p.aabb_ps;
  p.wts1 += q * p.wts_ps;
  p.wts2 += q * p.wts_ps;
  p.shifts += q * p.shifts_ps;
  p.yr += q * p.y_ps;
  p.yi += q * p.y_ps;
  const int lane = threadIdx.x;
  const int c = s < p.ns ? 0 : 1;
  const int node = s - c * p.ns;
  const double* w = c == 0 ? p.wts1 : p.wts2;
  const double b0 = p.aabb[2 * p.ld + c * p.ld];
  const double2 wt = d2(w[node], w[node + p.ns]);
  const double* al = p.aabb + c * p.ld;
  const double* be = p.aabb + (2 + c) * p.ld;
  const double* sh = p.shifts + 2 * (size_t)s * p.m;
  const int m = p.m;
  const double2 zero = d2(0.0, 0.0), one = d2(1.0, 0.0);
  // row i: A y_{i-1} + B y_i + C y_{i+1} = D, (T - shift) y = -beta_0 w e1
  double2 A[Q], B[Q], C[Q], D[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const int i = lane + 64 * j;
    if (i < m) {
      B[j] = d2(al[i] - sh[2 * i], -sh[2 * i + 1]);
      A[j] = i > 0 ? d2(be[i], 0.0) : zero;
      C[j] = i + 1 < m ? d2(be[i + 1], 0.0) : zero;
      D[j] = i == 0 ? d2(-(b0 * wt.x), -(b0 * wt.y)) : zero;
    } else {
      A[j] = zero;
      B[j] = one;
      C[j] = zero;
      D[j] = zero;
    }
  }
  auto shfl2 = [](double2 v, int src) {
    return d2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
  };
  static_assert(Q == 1 || Q == 2, "m <= 128");
  constexpr int kSteps = Q == 1 ? 6 : 7;   // h = 1 .. 32 Q
#pragma unroll
  for (int st = 0; st < kSteps; ++st) {
    const int h = 1 << st;
    const int hq = h >> 6, hr = h & 63;
    const int sm = (lane - hr) & 63, sp = (lane + hr) & 63;
    const bool borrow = lane < hr;          // row i - h lies in register j - hq - 1
    const bool carry = lane + hr >= 64;     // row i + h lies in register j + hq + 1
    double2 Am[Q], Bm[Q], Cm[Q], Dm[Q], Ap[Q], Bp[Q], Cp[Q], Dp[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      // row i - h: register j - hq (lanes >= hr) or j - hq - 1 (lanes < hr)
      const int r1 = j - hq, r2 = j - hq - 1;
      double2 a1 = zero, b1 = one, c1 = zero, d1 = zero, a2 = zero, b2 = one, c2 = zero, d2v = zero;
      if (r1 >= 0) {
        a1 = hr ? shfl2(A[r1], sm) : A[r1];
        b1 = hr ? shfl2(B[r1], sm) : B[r1];
        c1 = hr ? shfl2(C[r1], sm) : C[r1];
        d1 = hr ? shfl2(D[r1], sm) : D[r1];
      }
      if (hr && r2 >= 0) {
        a2 = shfl2(A[r2], sm);
        b2 = shfl2(B[r2], sm);
        c2 = shfl2(C[r2], sm);
        d2v = shfl2(D[r2], sm);
      }
      const bool use2 = hr && borrow;
      Am[j] = use2 ? a2 : a1;
      Bm[j] = use2 ? b2 : b1;
      Cm[j] = use2 ? c2 : c1;
      Dm[j] = use2 ? d2v : d1;
      // row i + h: register j + hq (lanes + hr < 64) or j + hq + 1
      const int r3 = j + hq, r4 = j + hq + 1;
      double2 a3 = zero, b3 = one, c3 = zero, d3 = zero, a4 = zero, b4 = one, c4 = zero, d4 = zero;
      if (r3 < Q) {
        a3 = hr ? shfl2(A[r3], sp) : A[r3];
        b3 = hr ? shfl2(B[r3], sp) : B[r3];
        c3 = hr ? shfl2(C[r3], sp) : C[r3];
        d3 = hr ? shfl2(D[r3], sp) : D[r3];
      }
      if (hr && r4 < Q) {
        a4 = shfl2(A[r4], sp);
        b4 = shfl2(B[r4], sp);
        c4 = shfl2(C[r4], sp);
        d4 = shfl2(D[r4], sp);
      }
      const bool use4 = hr && carry;
      Ap[j] = use4 ? a4 : a3;
      Bp[j] = use4 ? b4 : b3;
      Cp[j] = use4 ? c4 : c3;
      Dp[j] = use4 ? d4 : d3;
    }
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      const double2 fa = cdiv(A[j], Bm[j]);   // row i - h times -fa
      const double2 fc = cdiv(C[j], Bp[j]);   // row i + h times -fc
      A[j] = csub(zero, cmul(fa, Am[j]));
      C[j] = csub(zero, cmul(fc, Cp[j]));
      B[j] = csub(csub(B[j], cmul(fa, Cm[j])), cmul(fc, Ap[j]));
      D[j] = csub(csub(D[j], cmul(fa, Dm[j])), cmul(fc, Dp[j]));
    }
  }
  bool bad = false;
  double2 y[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    y[j] = cdiv(D[j], B[j]);
    if ((B[j].x == 0.0 && B[j].y == 0.0) || !isfinite(y[j].x) || !isfinite(y[j].y)) bad = true;
  }
  if (__any(bad)) {
    if (lane == 0) atomicOr(p.status, 2u);
#pragma unroll
    for (int j = 0; j < Q; ++j) y[j] = zero;
  }
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const int i = lane + 64 * j;
    if (i < m) {
      p.yr[i + (size_t)m * s] = y[j].x;
      p.yi[i + (size_t)m * s] = y[j].y;
    }
  }
}

// ---------------------------------------------------------------------------
// apply_V, collapsed over the nodes:
//   x_c = sum_s Im(q_s(A) V_c y_s) = sum_j A^(d-j) V_c c_j,
//   c_j = sum_s (gsr_{j,s} Im y_s + gsi_{j,s} Re y_s)
// so one pass over V (V*C, tall-skinny) plus d SpMMs replace the reference's
// 2 n_s d SpMMs and 4 n_s gemv (apply_V.cu:5-95).
// ---------------------------------------------------------------------------
struct CollapseArgs {
  const double* yr;    // m x 2ns
  const double* yi;
  const double* gsr;   // (d+1) x ns
  const double* gsi;
  int m, ns, nd;       // nd = d + 1
  double2* C;          // [nd][m] double2 (x: exp, y: phi)
  // independent problems (blockIdx.y = problem): strides of yr/yi, gsr/gsi, C
  size_t y_ps, g_ps, c_ps;
};

template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_collapse(CollapseArgs p) {
  const size_t q = blockIdx.y;
  p.yr += q * p.y_ps;
  p.yi += q * p.y_ps;
  p.gsr += q * p.g_ps;
  p.gsi += q * p.g_ps;
  p.C += q * p.c_ps;
  const int total = p.nd * p.m;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < total; t += gridDim.x * kBlock) {
    const int j = t / p.m, k = t - j * p.m;
    double cx = 0.0, cy = 0.0;
    for (int s = 0; s < p.ns; ++s) {
      const double gr = p.gsr[j + (size_t)p.nd * s], gim = p.gsi[j + (size_t)p.nd * s];
      const size_t e = k + (size_t)p.m * s, f = k + (size_t)p.m * (s + p.ns);
      cx = fma(gr, p.yi[e], cx);
      cx = fma(gim, p.yr[e], cx);
      cy = fma(gr, p.yi[f], cy);
      cy = fma(gim, p.yr[f], cy);
    }
    p.C[t] = d2(cx, cy);
  }
}

// Z_j = V C_j for j < nd, one row per thread; C staged in LDS.
struct VCArgs {
  const double2* V;    // [m][N]
  const double2* C;    // [nd][m]
  int m, nd, nrows;
  size_t ldv;          // N
  double2* Z;          // [nd][N]; nd == 1 -> Z is the output X
  int accumulate;      // nd == 1: X += V c_0
};

template <int ND>
__global__ __launch_bounds__(kBlock) void k_vc(VCArgs p) {
  extern __shared__ double2 sC[];
  for (int t = threadIdx.x; t < ND * p.m; t += kBlock) sC[t] = p.C[t];
  __syncthreads();
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= p.nrows) return;
  double2 acc[ND];
#pragma unroll
  for (int j = 0; j < ND; ++j) acc[j] = d2(0.0, 0.0);
  for (int k = 0; k < p.m; ++k) {
    const double2 v = p.V[(size_t)k * p.ldv + r];
#pragma unroll
    for (int j = 0; j < ND; ++j) {
      const double2 c = sC[j * p.m + k];
      acc[j].x = fma(v.x, c.x, acc[j].x);
      acc[j].y = fma(v.y, c.y, acc[j].y);
    }
  }
#pragma unroll
  for (int j = 0; j < ND; ++j) {
    double2 o = acc[j];
    if (ND == 1 && p.accumulate) {
      const double2 old = p.Z[r];
      o.x += old.x;
      o.y += old.y;
    }
    p.Z[(size_t)j * p.ldv + r] = o;
  }
}

// ---------------------------------------------------------------------------
// Time-step update and deflation.
// ---------------------------------------------------------------------------
// kUpdMS: Mitchell-Schaeffer ionic model (no reference counterpart; the
// reference's monodomain reaction is FitzHugh-Nagumo).  The gate h lives in
// the vector the FHN path uses for its recovery variable:
//   h <- h + dt ((1 - h)/tau_open)  if u < v_gate,  h - dt h/tau_close otherwise
//   g  = h u^2 (1 - u)/tau_in - u/tau_out   (with the updated h)
enum UpdateMode : int { kUpdFisher = 0, kUpdMono = 1, kUpdPrepare = 2, kUpdMS = 3 };

struct UpdateArgs {
  const double2* X;
  const double2* upd;
  const double* mass;
  double* ode;
  double dt, aode_dt, bode_dt, code, vrest, vpeak, cion;
  double ms_tau_in, ms_tau_out, ms_tau_open, ms_tau_close, ms_v_gate;
  int clamp, x_zero;
  double2* Usol;
  double2* U;          // K == 0: U = Usol
  const double* Q;     // row-major N x K
  int K, nrows, chunk;
  double* partials;    // K == 0: 2 per block; K > 0: 2K per block
  double* beta0;       // &aabb[2 ld] and &aabb[3 ld] -> beta0[0], beta0[ld]
  int ld;
  double* Zout;        // 2K (K > 0): Z = Q^T Usol, [k][c]
  unsigned* status;
};

// KT deflation vectors per register tile of pass 2: kDeflTile, or kDeflTileWide
// when K fits in one wide tile (one pass over the rows and Q instead of
// ceil(K / 8)); each Z partial is summed in the same order either way.
constexpr int kDeflTileWide = 24;

template <int MODE, int KT = kDeflTile>
__global__ __launch_bounds__(kBlock) void k_update1(UpdateArgs p) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int row0 = b * p.chunk;
  const int row1 = min(row0 + p.chunk, p.nrows);
  // pass 1: the elementwise update (update_U.cu:8-16, update_U_monodomain.cu:8-25)
  double2 nrm = d2(0.0, 0.0);
  for (int r = row0 + threadIdx.x; r < row1; r += kBlock) {
    double2 us;
    if (MODE == kUpdPrepare) {
      us = p.Usol[r];
    } else {
      const double2 x = p.x_zero ? d2(0.0, 0.0) : p.X[r];
      const double2 up = p.upd[r];
      double u = ((x.x + x.y) + up.x) + up.y;
      double g;
      if (MODE == kUpdFisher) {
        if (p.clamp) u = u * (u >= 0.0 ? 1.0 : 0.0);
        const double ms = p.mass[r];
        const double ub = u / ms;
        g = (ms * ub) * (1.0 - ub);
      } else if (MODE == kUpdMono) {
        const double o = p.ode[r];
        const double on = o + ((u * p.aode_dt + o * p.bode_dt) + p.code);
        p.ode[r] = on;
        const double src = ((u - p.vrest) * u) * (u - p.vpeak) * p.cion;
        g = src - on;
      } else {   // kUpdMS
        const double h = p.ode[r];
        const double dh = u < p.ms_v_gate ? (1.0 - h) / p.ms_tau_open : -h / p.ms_tau_close;
        const double hn = h + p.dt * dh;
        p.ode[r] = hn;
        g = (hn * (u * u) * (1.0 - u)) / p.ms_tau_in - u / p.ms_tau_out;
      }
      us = d2(u, g * p.dt);
      p.Usol[r] = us;
    }
    if (p.K == 0) {
      p.U[r] = us;
      nrm.x = fma(us.x, us.x, nrm.x);
      nrm.y = fma(us.y, us.y, nrm.y);
    }
  }
  if (p.K == 0) {
    block_partial2(nrm, p.partials);   // beta_0 = sqrt(sum), formed by k_reduce
    return;
  }
  // pass 2: block partials of Z = Q^T Usol, KT vectors at a time.
  __shared__ double red[kWaves][2 * KT];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k0 = 0; k0 < p.K; k0 += KT) {
    double acc[2 * KT];
#pragma unroll
    for (int t = 0; t < 2 * KT; ++t) acc[t] = 0.0;
    for (int r = row0 + threadIdx.x; r < row1; r += kBlock) {
      const double2 us = p.Usol[r];
      const double* q = p.Q + (size_t)r * p.K;
#pragma unroll
      for (int t = 0; t < KT; ++t) {
        if (k0 + t < p.K) {
          const double qv = q[k0 + t];
          acc[2 * t] = fma(qv, us.x, acc[2 * t]);
          acc[2 * t + 1] = fma(qv, us.y, acc[2 * t + 1]);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < 2 * KT; ++t) acc[t] = wave_sum(acc[t]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
      for (int t = 0; t < 2 * KT; ++t) red[wv][t] = acc[t];
    }
    __syncthreads();
    if (threadIdx.x < 2 * KT && k0 * 2 + (int)threadIdx.x < 2 * p.K) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < kWaves; ++i) s += red[i][threadIdx.x];
      p.partials[(size_t)blockIdx.x * 2 * p.K + 2 * k0 + threadIdx.x] = s;
    }
  }
  // Z = Q^T Usol = sum of these [G][2K] partials, formed by k_reduce
}

// U = Usol - Q Z; update = Q diag(e^L, phi(L)) Z; beta0 = ||U||.
// lanczos_impl.cu:84-89 / update_U.cu:22-27.
struct Update2Args {
  const double2* Usol;
  double2* U;
  double2* upd;
  const double* Q;
  const double* Z;       // 2K, [k][c]
  const double* el;      // K
  const double* pl;      // K
  int K, nrows, chunk, ld;
  double* partials;
  double* beta0;
  unsigned* status;
};

template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_update2(Update2Args p) {
  extern __shared__ double sz[];   // 4K: Z (2K) then f*Z (2K)
  for (int t = threadIdx.x; t < p.K; t += kBlock) {
    const double zx = p.Z[2 * t], zy = p.Z[2 * t + 1];
    sz[2 * t] = zx;
    sz[2 * t + 1] = zy;
    sz[2 * p.K + 2 * t] = zx * p.el[t];
    sz[2 * p.K + 2 * t + 1] = zy * p.pl[t];
  }
  __syncthreads();
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int row0 = b * p.chunk;
  const int row1 = min(row0 + p.chunk, p.nrows);
  double2 nrm = d2(0.0, 0.0);
  for (int r = row0 + threadIdx.x; r < row1; r += kBlock) {
    const double* q = p.Q + (size_t)r * p.K;
    double cx = 0.0, cy = 0.0, ux = 0.0, uy = 0.0;
    for (int k = 0; k < p.K; ++k) {
      const double qv = q[k];
      cx = fma(qv, sz[2 * k], cx);
      cy = fma(qv, sz[2 * k + 1], cy);
      ux = fma(qv, sz[2 * p.K + 2 * k], ux);
      uy = fma(qv, sz[2 * p.K + 2 * k + 1], uy);
    }
    const double2 us = p.Usol[r];
    const double2 u = d2(us.x - cx, us.y - cy);
    p.U[r] = u;
    p.upd[r] = d2(ux, uy);
    nrm.x = fma(u.x, u.x, nrm.x);
    nrm.y = fma(u.y, u.y, nrm.y);
  }
  block_partial2(nrm, p.partials);   // beta_0 = sqrt(sum), formed by k_reduce
}

// de-interleave an N x 2 double2 array into column-major (for the getters)
template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_deinterleave(const double2* src, double* dst, int n) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n) {
    const double2 v = src[r];
    dst[r] = v.x;
    dst[(size_t)n + r] = v.y;
  }
}

template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_zero2(double2* p, int n) {
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n) p[r] = d2(0.0, 0.0);
}

// Plain streaming copy (16-byte elements), the calibration point for the HBM
// byte counters.
template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_copy16(const double2* __restrict__ src,
                                                   double2* __restrict__ dst, long long n) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    dst[i] = src[i];
}

}  // namespace gm
