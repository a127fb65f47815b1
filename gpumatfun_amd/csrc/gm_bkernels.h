// gm_bkernels.h -- multi-column kernels (R columns per row, lanes over columns).
//
// Shared by the batched form (gm_batch.hip) and the adjoint (gm_adjoint.hip).
// Every N x R quantity is row-major with the R columns contiguous, so a CSR
// gather of x[col] is one coalesced R*8-byte load across a wavefront and each
// matrix entry is read once for all R columns; with R >= 64 a wavefront works
// on one row at a time, with R < 64 on 64/R rows.  All kernels are templates
// so that several translation units may instantiate them.
#pragma once
#include <algorithm>
#include <type_traits>
#include <vector>

#include "gm_device.h"

namespace gmb {
using namespace gm;

constexpr int kMaxND = 17;
constexpr int kUnroll = 8;   // nonzeros per gather batch in k_bspmm
#ifndef GMB_STAGED_UNROLL
#define GMB_STAGED_UNROLL 12
#endif
// nonzeros per gather batch when the matrix is staged in LDS
constexpr int kUnrollStaged = GMB_STAGED_UNROLL;
// rows per pass of a lane in the finish kernel (4 and 8 measured the same)
constexpr int kFinishRows = 2;

// kCheb: y = ca (A x - shift_c x) + cb z (a Chebyshev filter step, or with
// ca = 1, cb = 0 a residual A x - theta_c x); shift_c = shiftc[c], or cc.
enum Mode : int { kFirst = 0, kMid = 1, kLast = 2, kOnly = 3, kHorner = 4, kCheb = 5 };

struct BSpmm {
  const int* __restrict__ rowptr;
  const int* __restrict__ col;
  const double* __restrict__ val;
  int nrows, chunk;
  const int* rowblk;                // staged form: row-block starts (+ n)
  int nrowblk;                      // staged form: number of row blocks
  const double* x;  size_t ldx;     // gather source (+ column offset applied by caller)
  double* y;        size_t ldy;     // output
  double* vk;                       // V[k] (ld = R)
  const double* vkm1;               // V[k-1] or null
  const double* z;  size_t ldz;     // kHorner addend
  const double* al;                 // alpha [k][R]
  double* be;                       // beta  [k][R]
  int k;
  double g0, gi;
  int accumulate;
  double ca, cb, cc;                // kCheb coefficients
  const double* shiftc;             // kCheb per-column shift (or null: cc)
  double* alpha_out;                // kLast/kOnly: alpha[k][R]
  double* partials;                 // [G][R]
  // kFirst/kOnly with bpart: beta_k[c] = sqrt(sum of the previous finish's bG
  // partials [bG][R]) formed in every block; block 0 stores it in be[k][c]
  const double* bpart;
  int bG;
  unsigned* status;
};

// Single-use vector streams of the SpMM (v_k and y written, v_{k-1} read,
// once per launch) as non-temporal accesses when NT.  A compile-time choice:
// a run-time branch between the two forms is merged by the compiler into one
// plain access.
template <bool NT>
__device__ __forceinline__ void st_stream(double* ptr, double v) {
  if constexpr (NT) __builtin_nontemporal_store(v, ptr);
  else *ptr = v;
}
template <bool NT>
__device__ __forceinline__ double ld_stream(const double* ptr) {
  if constexpr (NT) return __builtin_nontemporal_load(ptr);
  else return *ptr;
}

// Column sums of G partials [G][R] (written by the previous kernel) into
// out[0..R) in LDS, the same fixed order in every block: thread t sums
// column t % W over g = t / W, t / W + ns, ... (ns = kBlock / W, W = min(R,
// kBlock)), then a fixed tree over the ns sub-sums.
__device__ __forceinline__ void block_reduce_cols(const double* part, int G, int R, double* out) {
  __shared__ double sub[kBlock];
  for (int v0 = 0; v0 < R; v0 += kBlock) {
    const int W = min(R - v0, kBlock);
    const int ns = kBlock / W;
    const int vv = (int)threadIdx.x % W, sb = (int)threadIdx.x / W;
    double acc = 0.0;
    if (sb < ns)
      for (int g = sb; g < G; g += ns) acc += part[(size_t)g * R + v0 + vv];
    __syncthreads();
    sub[threadIdx.x] = acc;
    __syncthreads();
    int h = 1;
    while (h < ns) h <<= 1;
    for (h >>= 1; h > 0; h >>= 1) {
      if (sb < h && sb + h < ns) sub[threadIdx.x] += sub[threadIdx.x + h * W];
      __syncthreads();
    }
    if ((int)threadIdx.x < W) out[v0 + threadIdx.x] = sub[threadIdx.x];
    __syncthreads();
  }
}

// Per-column block sums (fixed order) written to partials[block][R]; the grid
// total is formed by k_reduce after the kernel.
template <int R>
__device__ void block_partial_cols(const double (&v)[(R < 64 ? 1 : R / 64)], double* partials) {
  constexpr int CPW = R < 64 ? R : 64;
  constexpr int RPW = 64 / CPW;
  constexpr int CL = R / CPW;
  __shared__ double sh[kBlock * CL];
  for (int q = 0; q < CL; ++q) sh[threadIdx.x * CL + q] = v[q];
  __syncthreads();
  for (int c = threadIdx.x; c < R; c += kBlock) {
    const int lanec = c % CPW, q = c / CPW;
    double s = 0.0;
    for (int w = 0; w < kWaves; ++w)
      for (int sl = 0; sl < RPW; ++sl) s += sh[(w * 64 + sl * CPW + lanec) * CL + q];
    partials[(size_t)blockIdx.x * R + c] = s;
  }
}

// Row blocks for the staged form: consecutive rows with at most kBTile
// nonzeros and kBRows rows per block; a longer row is a block of its own
// (read from global memory).  rb gets the block starts and n at the end.
// 512 nonzeros measured best on the 512^2 9-point operator at 32 columns (5%
// over 1024 and 256 for the batched step and the eigensolver; 2048 slower),
// and equal on the 128^3 7-point one at 256 columns.
constexpr int kBTile = 512;
constexpr int kBRows = 256;

inline void bspmm_row_blocks(const int* rowptr, int n, std::vector<int>& rb,
                             int max_rows = kBRows) {
  rb.clear();
  int r = 0;
  while (r < n) {
    rb.push_back(r);
    const int j0 = rowptr[r];
    int e = r + 1;
    while (e < n && e - r < max_rows && rowptr[e + 1] - j0 <= kBTile) ++e;
    r = e;
  }
  rb.push_back(n);
}

// kStaged: the grid's blocks take contiguous runs of the nrowblk row blocks
// (rows rowblk[i] .. rowblk[i+1]-1); each row block's values, column indices
// and row pointers are first staged in LDS by coalesced loads (one HBM
// latency per row block instead of one per row).  Otherwise block b covers
// p.chunk rows read straight from global memory.  Partials are per grid
// block either way.  kNT: the single-use vector streams are non-temporal
// (wide chunks, whose 2 KB rows overflow L2 anyway: 5.22 -> 5.04 ms per launch
// at 256 columns on the 128^3 Laplacian, the kernels that read them next
// unchanged).
template <int R, int MODE, bool kStaged = false, bool kNT = false>
__global__ __launch_bounds__(kBlock) void k_bspmm(BSpmm p) {
  constexpr int CPW = R < 64 ? R : 64;
  constexpr int RPW = 64 / CPW;
  constexpr int CL = R / CPW;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int it0 = 0, it1 = 1;
  if (kStaged) {
    const int per = (p.nrowblk + gridDim.x - 1) / gridDim.x;
    it0 = min(b * per, p.nrowblk);
    it1 = min(it0 + per, p.nrowblk);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int slot = lane / CPW, c0 = lane % CPW;
  double bet[CL], inv[CL], bprev[CL], dot[CL];
  __shared__ double fused[(MODE == kFirst || MODE == kOnly) ? R : 1];
  __shared__ double sval[kStaged ? kBTile : 1];
  __shared__ int scol[kStaged ? kBTile : 1];
  __shared__ int srp[kStaged ? kBRows + 1 : 1];
  if ((MODE == kFirst || MODE == kOnly) && p.bpart) {
    block_reduce_cols(p.bpart, p.bG, R, fused);
    for (int c = threadIdx.x; c < R; c += kBlock) {
      const double bv = sqrt(fused[c]);
      fused[c] = bv;
      if (blockIdx.x == 0) {
        p.be[(size_t)p.k * R + c] = bv;
        if (!isfinite(bv)) atomicOr(p.status, 4u);
        else if (bv == 0.0) atomicOr(p.status, 1u);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < CL; ++q) {
    const int c = c0 + q * CPW;
    bet[q] = (MODE == kFirst || MODE == kOnly)
                 ? (p.bpart ? fused[c] : p.be[(size_t)p.k * R + c]) : 0.0;
    inv[q] = safe_inv(bet[q]) * p.g0;
    bprev[q] = ((MODE == kLast || MODE == kOnly) && p.k > 0)
                   ? ((MODE == kOnly && p.bpart) ? bet[q] : p.be[(size_t)p.k * R + c]) : 0.0;
    dot[q] = 0.0;
  }
  for (int it = it0; it < it1; ++it) {
    int row0, row1;
    if (kStaged) {
      row0 = p.rowblk[it];
      row1 = p.rowblk[it + 1];
    } else {
      row0 = b * p.chunk;
      row1 = min(row0 + p.chunk, p.nrows);
    }
    bool staged = false;
    if (kStaged) {
      const int jbase = p.rowptr[row0];
      const int nz = p.rowptr[row1] - jbase;
      staged = nz <= kBTile;   // uniform over the block
      __syncthreads();         // the previous row block's LDS reads are done
      if (staged) {
        for (int t = threadIdx.x; t < nz; t += kBlock) {
          sval[t] = p.val[jbase + t];
          scol[t] = p.col[jbase + t];
        }
        for (int t = threadIdx.x; t <= row1 - row0; t += kBlock)
          srp[t] = p.rowptr[row0 + t] - jbase;
      }
      __syncthreads();
    }
    // the row loop, with the matrix read from LDS (from_lds) or global memory
    auto rows = [&](auto from_lds) {
      constexpr bool L = decltype(from_lds)::value;
      for (int r = row0 + wave * RPW + slot; r < row1; r += kWaves * RPW) {
        const int j0 = L ? srp[r - row0] : p.rowptr[r];
        const int j1 = L ? srp[r - row0 + 1] : p.rowptr[r + 1];
        double acc[CL];
  #pragma unroll
        for (int q = 0; q < CL; ++q) acc[q] = 0.0;
        // the epilogue's own-row operands, loaded first so that their latency
        // overlaps the gathers
        double o0[CL], o1[CL];
  #pragma unroll
        for (int q = 0; q < CL; ++q) {
          const int c = c0 + q * CPW;
          const size_t iv = (size_t)r * R + c;
          o0[q] = 0.0;
          o1[q] = 0.0;
          if (MODE == kFirst || MODE == kOnly) o0[q] = p.x[(size_t)r * p.ldx + c];
          if (MODE == kMid || MODE == kLast) o0[q] = p.vk[iv];
          if ((MODE == kLast || MODE == kOnly) && p.k > 0) o1[q] = ld_stream<kNT>(p.vkm1 + iv);
          if (MODE == kHorner) {
            o0[q] = p.z[(size_t)r * p.ldz + c];
            if (p.accumulate) o1[q] = p.y[(size_t)r * p.ldy + c];
          }
          if (MODE == kCheb) {
            o0[q] = p.x[(size_t)r * p.ldx + c];
            if (p.z) o1[q] = p.z[(size_t)r * p.ldz + c];
          }
        }
        // kUnroll nonzeros at a time: their val/col loads, then all their
        // gathers, are issued before the first FMA waits, so a wavefront has
        // kUnroll * CL gathers in flight instead of one dependent chain per
        // nonzero.  The FMAs still run in column order.
        constexpr int U = L ? kUnrollStaged : kUnroll;
        for (int jb = j0; jb < j1; jb += U) {
          double a[U];
          int ci[U];
  #pragma unroll
          for (int t = 0; t < U; ++t) {
            const bool ok = jb + t < j1;
            a[t] = ok ? (L ? sval[jb + t] : p.val[jb + t]) : 0.0;
            ci[t] = ok ? (L ? scol[jb + t] : p.col[jb + t]) : -1;
          }
          double xv[U][CL];
  #pragma unroll
          for (int t = 0; t < U; ++t) {
            const double* xr = p.x + (size_t)(ci[t] < 0 ? r : ci[t]) * p.ldx + c0;
  #pragma unroll
            for (int q = 0; q < CL; ++q) xv[t][q] = ci[t] < 0 ? 0.0 : xr[q * CPW];
          }
  #pragma unroll
          for (int t = 0; t < U; ++t) {
            if (ci[t] >= 0) {
  #pragma unroll
              for (int q = 0; q < CL; ++q) acc[q] = fma(a[t], xv[t][q], acc[q]);
            }
          }
        }
  #pragma unroll
        for (int q = 0; q < CL; ++q) {
          const int c = c0 + q * CPW;
          const size_t iy = (size_t)r * p.ldy + c;
          const size_t iv = (size_t)r * R + c;
          if (MODE == kFirst) {
            const double v = safe_div(o0[q], bet[q]);
            st_stream<kNT>(p.vk + iv, v);
            st_stream<kNT>(p.y + iy, fma(p.gi, v, acc[q] * inv[q]));
          } else if (MODE == kMid) {
            st_stream<kNT>(p.y + iy, fma(p.gi, o0[q], acc[q]));
          } else if (MODE == kLast || MODE == kOnly) {
            double v, yv;
            if (MODE == kOnly) {
              v = safe_div(o0[q], bet[q]);
              st_stream<kNT>(p.vk + iv, v);
              yv = acc[q] * inv[q];
            } else {
              v = o0[q];
              yv = acc[q];
            }
            if (p.k > 0) yv = fma(-bprev[q], o1[q], yv);
            st_stream<kNT>(p.y + iy, yv);
            dot[q] = fma(yv, v, dot[q]);
          } else if (MODE == kCheb) {
            const double sh = p.shiftc ? p.shiftc[c] : p.cc;
            p.y[iy] = fma(p.cb, o1[q], p.ca * fma(-sh, o0[q], acc[q]));
          } else {
            double yv = acc[q] + o0[q];
            if (p.accumulate) yv += o1[q];
            p.y[iy] = yv;
          }
        }
      }
    };
    if (kStaged && staged) {
      rows(std::true_type{});
    } else {
      rows(std::false_type{});
    }
  }
  if (MODE == kLast || MODE == kOnly)
    block_partial_cols<R>(dot, p.partials);
}

// Two adjacent doubles as one 16-byte access (plain or non-temporal).
typedef double d2v __attribute__((ext_vector_type(2)));
template <bool NT>
__device__ __forceinline__ void st_stream2(double* ptr, double a, double b) {
  const d2v v = {a, b};
  if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<d2v*>(ptr));
  else *reinterpret_cast<d2v*>(ptr) = v;
}
template <bool NT>
__device__ __forceinline__ d2v ld_stream2(const double* ptr) {
  if constexpr (NT) return __builtin_nontemporal_load(reinterpret_cast<const d2v*>(ptr));
  else return *reinterpret_cast<const d2v*>(ptr);
}

// Row blocks of the slab form: at most kBTile nonzeros and kXsRows rows
// (four passes of a workgroup's 16 row slots), described to the kernel by
// {first row, end row, first nonzero, nonzeros} (bspmm_xs_blocks).
constexpr int kXsRows = 64;
static_assert(kBTile == 2 * kBlock && kXsRows < kBlock, "two staged nonzeros per thread");

inline void bspmm_xs_blocks(const int* rowptr, int n, std::vector<int>& desc) {
  std::vector<int> rb;
  bspmm_row_blocks(rowptr, n, rb, kXsRows);
  desc.clear();
  for (size_t i = 0; i + 1 < rb.size(); ++i) {
    desc.push_back(rb[i]);
    desc.push_back(rb[i + 1]);
    desc.push_back(rowptr[rb[i]]);
    desc.push_back(rowptr[rb[i + 1]] - rowptr[rb[i]]);
  }
}

// Chunks of R = 32 .. 256 columns in column slabs of 32 (the kXs form of
// k_bspmm).
//  - Workgroup b runs on XCD b % 8 (the dispatcher deals workgroups to the
//    eight XCDs in turn) and takes slab (b % 8) % S (S = R / 32 slabs), so at
//    256 columns XCD s owns columns 32s .. 32s + 31 of every row: each XCD's
//    L2 then holds only its own 256-byte slice of an x row, and no line is
//    fetched by two XCDs (a whole 2 KB row leaves a 4 MB L2 before the
//    neighbours that gather it again come by).  With S < 8 each of a slab's
//    8 / S XCDs owns a contiguous part of the row blocks.
//  - The G/8 workgroups of an XCD (all resident: the host sizes G by
//    occupancy) take its row blocks in turn (w, w + G/8, ...), so that they
//    sweep the rows as one front of G/8 row blocks.
//  - Row blocks are double-buffered in LDS: the next block's values, column
//    indices and row pointers are loaded into registers before this block's
//    rows and stored to the other buffer after them; the block descriptors
//    are read two blocks ahead.
//  - A lane holds two adjacent columns of its slab: one 16-byte access per
//    row operand, 16 lanes (256 contiguous bytes) a row, 4 rows a wavefront.
//  - Per row the sums are k_bspmm's (the same nonzeros in the same order, the
//    same epilogue).  Column sums go to partials [G/S][R] (row
//    part * G/8 + b / 8 of the slab's columns), reduced after the kernel.
//  - VW: the basis holds the unnormalised w_j; v_k = w_k / beta_k and v_{k-1}
//    are formed where read (the bits the stored v_k would have), and kFirst /
//    kOnly store nothing (vk = null) or, at k = 0, the raw w_0.
//  - p.rowblk holds p.nrowblk descriptors (bspmm_xs_blocks); a block that does
//    not fit the LDS tile (one row of more than kBTile nonzeros) is read from
//    global memory.  Requires: G a multiple of 8; x, y, vk, vkm1, z 16-byte
//    aligned and every row stride even; no fused beta (bpart).
template <int R, int MODE, bool kNT, bool VW = false>
__global__ __launch_bounds__(kBlock) void k_bspmm_xs(BSpmm p) {
  constexpr int RS = 32, S = R / RS, LPR = RS / 2, RPW = 64 / LPR, RPB = RPW * kWaves;
  static_assert(R % RS == 0 && S >= 1 && 8 % S == 0, "R = 32, 64, 128 or 256 in slabs of 32");
  constexpr int U = kUnroll;
  // workgroup b runs on XCD b % 8; it takes slab (b % 8) % S, and the XCD's
  // share of that slab is the contiguous part (b % 8) / S of its row blocks,
  // swept by the XCD's G/8 workgroups in turn
  constexpr int PARTS = 8 / S;
  const int xcd = (int)(blockIdx.x & 7);
  const int slab = xcd % S, part = xcd / S;
  const int W = (int)(gridDim.x >> 3), lw = (int)(blockIdx.x >> 3);
  const int it_end = (int)((long long)p.nrowblk * (part + 1) / PARTS);
  const int w0 = (int)((long long)p.nrowblk * part / PARTS) + lw;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rslot = wave * RPW + lane / LPR;
  const int c = slab * RS + 2 * (lane % LPR);   // columns c, c + 1
  const int4* desc = reinterpret_cast<const int4*>(p.rowblk);
  double bet[2], inv[2], bprev[2], dot[2], bk[2], bkm1[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    bet[q] = (MODE == kFirst || MODE == kOnly) ? p.be[(size_t)p.k * R + c + q] : 0.0;
    inv[q] = safe_inv(bet[q]) * p.g0;
    bprev[q] = ((MODE == kLast || MODE == kOnly) && p.k > 0) ? p.be[(size_t)p.k * R + c + q] : 0.0;
    dot[q] = 0.0;
    // beta_k and beta_{k-1} that turn the stored w_k, w_{k-1} into v_k, v_{k-1}
    bk[q] = (VW && (MODE == kMid || MODE == kLast)) ? p.be[(size_t)p.k * R + c + q] : 1.0;
    bkm1[q] = (VW && (MODE == kLast || MODE == kOnly) && p.k > 0)
                  ? p.be[(size_t)(p.k - 1) * R + c + q] : 1.0;
  }
  __shared__ double sval[2][kBTile];
  __shared__ int scol[2][kBTile];
  __shared__ int srp[2][kXsRows + 1];
  auto fits = [](const int4 d) { return d.w <= kBTile && d.y - d.x <= kXsRows; };
  // a row block's LDS tile, held in registers between its loads and stores:
  // thread t has nonzeros t and t + kBlock and row pointer t
  double pv[2];
  int pc[2], pr;
  auto fetch = [&](const int4 d) {
    const bool st = fits(d);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int t = (int)threadIdx.x + u * kBlock;
      pv[u] = (st && t < d.w) ? p.val[d.z + t] : 0.0;
      pc[u] = (st && t < d.w) ? p.col[d.z + t] : 0;
    }
    pr = (st && (int)threadIdx.x <= d.y - d.x) ? p.rowptr[d.x + threadIdx.x] - d.z : 0;
  };
  auto commit = [&](const int4 d, int buf) {
    if (!fits(d)) return;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int t = (int)threadIdx.x + u * kBlock;
      if (t < d.w) {
        sval[buf][t] = pv[u];
        scol[buf][t] = pc[u];
      }
    }
    if ((int)threadIdx.x <= d.y - d.x) srp[buf][threadIdx.x] = pr;
  };
  const int4 none = {0, 0, 0, 0};
  int4 dcur = w0 < it_end ? desc[w0] : none;
  int4 dnx = w0 + W < it_end ? desc[w0 + W] : none;
  fetch(dcur);
  commit(dcur, 0);
  __syncthreads();
  int buf = 0;
  for (int it = w0; it < it_end; it += W) {
    const int4 dnx2 = it + 2 * W < it_end ? desc[it + 2 * W] : none;
    fetch(dnx);   // in flight during this block's rows
    const int row0 = dcur.x, row1 = dcur.y;
    auto rows = [&](auto from_lds) {
      constexpr bool L = decltype(from_lds)::value;
      for (int r = row0 + rslot; r < row1; r += RPB) {
        const int j0 = L ? srp[buf][r - row0] : p.rowptr[r];
        const int j1 = L ? srp[buf][r - row0 + 1] : p.rowptr[r + 1];
        const size_t iv = (size_t)r * R + c;
        // the epilogue's own-row operands first, so that their latency
        // overlaps the gathers
        d2v o0 = {0.0, 0.0}, o1 = {0.0, 0.0};
        if (MODE == kFirst || MODE == kOnly) o0 = ld_stream2<false>(p.x + (size_t)r * p.ldx + c);
        if (MODE == kMid || MODE == kLast) o0 = ld_stream2<false>(p.vk + iv);
        if ((MODE == kLast || MODE == kOnly) && p.k > 0) o1 = ld_stream2<kNT>(p.vkm1 + iv);
        if (MODE == kHorner) {
          o0 = ld_stream2<false>(p.z + (size_t)r * p.ldz + c);
          if (p.accumulate) o1 = ld_stream2<false>(p.y + (size_t)r * p.ldy + c);
        }
        if (MODE == kCheb) {
          o0 = ld_stream2<false>(p.x + (size_t)r * p.ldx + c);
          if (p.z) o1 = ld_stream2<false>(p.z + (size_t)r * p.ldz + c);
        }
        double acc[2] = {0.0, 0.0};
        for (int jb = j0; jb < j1; jb += U) {
          double a[U];
          int ci[U];
#pragma unroll
          for (int t = 0; t < U; ++t) {
            const bool ok = jb + t < j1;
            a[t] = ok ? (L ? sval[buf][jb + t] : p.val[jb + t]) : 0.0;
            ci[t] = ok ? (L ? scol[buf][jb + t] : p.col[jb + t]) : -1;
          }
          d2v xv[U];
#pragma unroll
          for (int t = 0; t < U; ++t) {
            const d2v zero = {0.0, 0.0};
            xv[t] = ci[t] < 0 ? zero : ld_stream2<false>(p.x + (size_t)ci[t] * p.ldx + c);
          }
#pragma unroll
          for (int t = 0; t < U; ++t) {
            if (ci[t] >= 0) {
              acc[0] = fma(a[t], xv[t].x, acc[0]);
              acc[1] = fma(a[t], xv[t].y, acc[1]);
            }
          }
        }
        double own0[2] = {o0.x, o0.y}, own1[2] = {o1.x, o1.y};
        if (VW) {
          // the same v_k = w_k / beta_k and v_{k-1} that the normalised basis
          // would hold, bit for bit
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            if (MODE == kMid || MODE == kLast) own0[q] = safe_div(own0[q], bk[q]);
            if ((MODE == kLast || MODE == kOnly) && p.k > 0) own1[q] = safe_div(own1[q], bkm1[q]);
          }
        }
        double vo[2] = {0.0, 0.0}, yo[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          if (MODE == kFirst) {
            vo[q] = safe_div(own0[q], bet[q]);
            yo[q] = fma(p.gi, vo[q], acc[q] * inv[q]);
          } else if (MODE == kMid) {
            yo[q] = fma(p.gi, own0[q], acc[q]);
          } else if (MODE == kLast || MODE == kOnly) {
            double v, yv;
            if (MODE == kOnly) {
              v = safe_div(own0[q], bet[q]);
              vo[q] = v;
              yv = acc[q] * inv[q];
            } else {
              v = own0[q];
              yv = acc[q];
            }
            if (p.k > 0) yv = fma(-bprev[q], own1[q], yv);
            yo[q] = yv;
            dot[q] = fma(yv, v, dot[q]);
          } else if (MODE == kCheb) {
            const double sh = p.shiftc ? p.shiftc[c + q] : p.cc;
            yo[q] = fma(p.cb, own1[q], p.ca * fma(-sh, own0[q], acc[q]));
          } else {
            double yv = acc[q] + own0[q];
            if (p.accumulate) yv += own1[q];
            yo[q] = yv;
          }
        }
        if ((MODE == kFirst || MODE == kOnly) && !VW) st_stream2<kNT>(p.vk + iv, vo[0], vo[1]);
        if ((MODE == kFirst || MODE == kOnly) && VW && p.vk)
          st_stream2<kNT>(p.vk + iv, o0.x, o0.y);   // the raw w_0
        if (MODE == kHorner || MODE == kCheb) {
          *reinterpret_cast<d2v*>(p.y + (size_t)r * p.ldy + c) = d2v{yo[0], yo[1]};
        } else {
          st_stream2<kNT>(p.y + (size_t)r * p.ldy + c, yo[0], yo[1]);
        }
      }
    };
    if (fits(dcur)) rows(std::true_type{});
    else rows(std::false_type{});
    // the other buffer's last readers were this block's predecessor's rows,
    // which ended before the previous barrier
    commit(dnx, buf ^ 1);
    __syncthreads();
    dcur = dnx;
    dnx = dnx2;
    buf ^= 1;
  }
  if (MODE == kLast || MODE == kOnly) {
    // fixed-order block sums per column: row slots of a wavefront in order,
    // then the wavefronts in order
    __shared__ double sh[kBlock * 2];
    sh[threadIdx.x * 2] = dot[0];
    sh[threadIdx.x * 2 + 1] = dot[1];
    __syncthreads();
    if (threadIdx.x < RS) {
      const int j = threadIdx.x;                       // slab column
      const int l = j >> 1, q = j & 1;
      double s = 0.0;
      for (int wv = 0; wv < kWaves; ++wv)
        for (int sl = 0; sl < RPW; ++sl) s += sh[(wv * 64 + sl * LPR + l) * 2 + q];
      p.partials[(size_t)(part * W + lw) * R + slab * RS + j] = s;
    }
  }
}

// Workgroups of the slab SpMM: as many as are resident at once (the fewest
// that any of its instantiations allows per CU), so that each slab's
// workgroups sweep the rows together as one front; a multiple of 8.
template <int R>
void bspmm_xs_occupancy(int* per_cu) {
  auto occ = [&](const void* k) {
    int v = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k, kBlock, 0) == hipSuccess && v > 0)
      *per_cu = std::min(*per_cu, v);
  };
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kFirst, false>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kMid, false>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kLast, false>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kOnly, false>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kHorner, false>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kCheb, false>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kFirst, true>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kMid, true>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kLast, true>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kOnly, true>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kHorner, true>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kFirst, true, true>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kMid, true, true>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kLast, true, true>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kOnly, true, true>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kFirst, false, true>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kMid, false, true>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kLast, false, true>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kOnly, false, true>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kHorner, true, true>));
  occ(reinterpret_cast<const void*>(&k_bspmm_xs<R, kHorner, false, true>));
}

inline int bspmm_xs_grid(int dev, int nrowblk) {
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    ncu = 8;
  int per_cu = 1 << 30;
  bspmm_xs_occupancy<32>(&per_cu);
  bspmm_xs_occupancy<64>(&per_cu);
  bspmm_xs_occupancy<128>(&per_cu);
  bspmm_xs_occupancy<256>(&per_cu);
  if (per_cu == (1 << 30)) per_cu = 1;
  return 8 * std::max(1, std::min(nrowblk, std::max(1, ncu / 8) * per_cu));
}

struct BFinish {
  double* w;          // ld R
  const double* vk;
  const double* al;
  double* be;
  int k, nrows, chunk;
  double* partials;
  unsigned* status;
  // with apart: alpha_k[c] = sum of the SpMM's aG partials [aG][R], formed in
  // every block; block 0 stores it in al_out[k][c]
  const double* apart;
  int aG;
  double* al_out;
  // wout: where w_{k+1} goes (null: over w).  With VW (the kernels' template
  // flag) vk holds the unnormalised w_k and v_k = w_k / beta_k is formed here,
  // bit for bit the normalised basis' value.
  double* wout;
};

template <int R, bool VW = false>
__global__ __launch_bounds__(kBlock) void k_bfinish(BFinish p) {
  constexpr int CPW = R < 64 ? R : 64;
  constexpr int RPW = 64 / CPW;
  constexpr int CL = R / CPW;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int row0 = b * p.chunk;
  const int row1 = min(row0 + p.chunk, p.nrows);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int slot = lane / CPW, c0 = lane % CPW;
  double al[CL], nrm[CL];
  __shared__ double fused[R];
  if (p.apart) {
    block_reduce_cols(p.apart, p.aG, R, fused);
    if (blockIdx.x == 0)
      for (int c = threadIdx.x; c < R; c += kBlock) p.al_out[(size_t)p.k * R + c] = fused[c];
  }
  double bk[CL];
#pragma unroll
  for (int q = 0; q < CL; ++q) {
    al[q] = p.apart ? fused[c0 + q * CPW] : p.al[(size_t)p.k * R + c0 + q * CPW];
    nrm[q] = 0.0;
    bk[q] = VW ? p.be[(size_t)p.k * R + c0 + q * CPW] : 1.0;
  }
  double* wo = p.wout ? p.wout : p.w;
  // NR rows per pass, all loads before the stores (w and vk are distinct
  // buffers, which the compiler cannot assume); a lane's rows are added into
  // nrm in the same order whatever NR is
  constexpr int STEP = kWaves * RPW;
  constexpr int NR = kFinishRows;
  for (int r = row0 + wave * RPW + slot; r < row1; r += NR * STEP) {
    double wv[NR][CL], vv[NR][CL];
#pragma unroll
    for (int h = 0; h < NR; ++h) {
      const int rh = r + h * STEP < row1 ? r + h * STEP : r;
#pragma unroll
      for (int q = 0; q < CL; ++q) {
        const size_t i = (size_t)rh * R + c0 + q * CPW;
        wv[h][q] = p.w[i];
        vv[h][q] = p.vk[i];
      }
    }
#pragma unroll
    for (int h = 0; h < NR; ++h) {
      if (r + h * STEP >= row1) break;
#pragma unroll
      for (int q = 0; q < CL; ++q) {
        const double v = VW ? safe_div(vv[h][q], bk[q]) : vv[h][q];
        const double w = fma(-al[q], v, wv[h][q]);
        wo[(size_t)(r + h * STEP) * R + c0 + q * CPW] = w;
        nrm[q] = fma(w, w, nrm[q]);
      }
    }
  }
  block_partial_cols<R>(nrm, p.partials);
}

// k_bfinish with two adjacent columns a thread (16-byte accesses; R / 2
// threads a row, kBlock / (R / 2) rows a pass) and four rows' loads in flight
// per thread.  The same update w = w - alpha_k v_k per element; the column
// sums of w^2 run over another partition (fixed order).  R >= 4, no apart.
template <int R, bool VW = false>
__global__ __launch_bounds__(kBlock) void k_bfinish2(BFinish p) {
  static_assert(R >= 4 && R <= 2 * kBlock, "two columns a thread");
  constexpr int P2 = R / 2, RPP = kBlock / P2, NR = 4;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int row0 = b * p.chunk;
  const int row1 = min(row0 + p.chunk, p.nrows);
  const int cp = (int)threadIdx.x % P2, rs = (int)threadIdx.x / P2;
  const int c = 2 * cp;
  const double al0 = p.al[(size_t)p.k * R + c], al1 = p.al[(size_t)p.k * R + c + 1];
  const double bk0 = VW ? p.be[(size_t)p.k * R + c] : 1.0;
  const double bk1 = VW ? p.be[(size_t)p.k * R + c + 1] : 1.0;
  double* wo = p.wout ? p.wout : p.w;
  double n0 = 0.0, n1 = 0.0;
  for (int r = row0 + rs; r < row1; r += NR * RPP) {
    d2v wv[NR], vv[NR];
#pragma unroll
    for (int h = 0; h < NR; ++h) {
      const int rh = r + h * RPP < row1 ? r + h * RPP : r;
      wv[h] = ld_stream2<false>(p.w + (size_t)rh * R + c);
      vv[h] = ld_stream2<false>(p.vk + (size_t)rh * R + c);
    }
#pragma unroll
    for (int h = 0; h < NR; ++h) {
      if (r + h * RPP >= row1) break;
      const double v0 = VW ? safe_div(vv[h].x, bk0) : vv[h].x;
      const double v1 = VW ? safe_div(vv[h].y, bk1) : vv[h].y;
      const double w0 = fma(-al0, v0, wv[h].x), w1 = fma(-al1, v1, wv[h].y);
      st_stream2<false>(wo + (size_t)(r + h * RPP) * R + c, w0, w1);
      n0 = fma(w0, w0, n0);
      n1 = fma(w1, w1, n1);
    }
  }
  __shared__ double sh[2 * kBlock];
  sh[2 * threadIdx.x] = n0;
  sh[2 * threadIdx.x + 1] = n1;
  __syncthreads();
  for (int col = threadIdx.x; col < R; col += kBlock) {
    const int pc = col >> 1, q = col & 1;
    double sum = 0.0;
    for (int k = 0; k < RPP; ++k) sum += sh[2 * (k * P2 + pc) + q];
    p.partials[(size_t)blockIdx.x * R + col] = sum;
  }
}

// C[j][k][c] /= beta_k[c] (nd x m x R): the V C product over a basis that holds
// the unnormalised w_k.
template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_bscale_c(double* C, const double* be, int nd, int m,
                                                     int R) {
  const long long e = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (e >= (long long)nd * m * R) return;
  const int c = (int)(e % R), k = (int)((e / R) % m);
  C[e] = safe_div(C[e], be[(size_t)k * R + c]);
}

// Column norms of U (ld ldu, column offset applied) -> be[0][c].
struct BNorm {
  const double* u;
  size_t ldu;
  double* be;
  int nrows, chunk;
  double* partials;
  unsigned* status;
};

template <int R>
__global__ __launch_bounds__(kBlock) void k_bnorm(BNorm p) {
  constexpr int CPW = R < 64 ? R : 64;
  constexpr int RPW = 64 / CPW;
  constexpr int CL = R / CPW;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int row0 = b * p.chunk;
  const int row1 = min(row0 + p.chunk, p.nrows);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int slot = lane / CPW, c0 = lane % CPW;
  double nrm[CL];
#pragma unroll
  for (int q = 0; q < CL; ++q) nrm[q] = 0.0;
  for (int r = row0 + wave * RPW + slot; r < row1; r += kWaves * RPW) {
#pragma unroll
    for (int q = 0; q < CL; ++q) {
      const double u = p.u[(size_t)r * p.ldu + c0 + q * CPW];
      nrm[q] = fma(u, u, nrm[q]);
    }
  }
  block_partial_cols<R>(nrm, p.partials);
}

// (T_c - etaxi_s I) y = -beta0_c w_{f(c),s} e1, one lane per (column, node);
// column c uses weights wts1 (even c, exp) or wts2 (odd c, phi).  Pivoted like
// k_shifted_solves; Y is [c][s][k].
struct BSolve {
  const double* al;   // [k][R]
  const double* be;   // [k][R]
  int R, m, ns;
  const double* wts1, *wts2;   // ns x 2
  const double* etaxi;         // 2 ns (re, im)
  double2 *wD, *wDU, *wDU2, *wX;   // [k][R*ns]
  double* yr;
  double* yi;
  unsigned* status;
};

template <int kUnused = 0>
__global__ __launch_bounds__(64) void k_bsolves(BSolve p) {
  const int sys = blockIdx.x * blockDim.x + threadIdx.x;
  const int S = p.R * p.ns;
  if (sys >= S) return;
  const int c = sys / p.ns, s = sys - c * p.ns;
  const double* w = (c & 1) ? p.wts2 : p.wts1;
  const double2 wt = d2(w[s], w[s + p.ns]);
  const double2 ex = d2(p.etaxi[2 * s], p.etaxi[2 * s + 1]);
  const double b0 = p.be[c];
  const int m = p.m;
  auto diag = [&](int k) { return d2(p.al[(size_t)k * p.R + c] - ex.x, -ex.y); };
  auto off = [&](int k) { return d2(p.be[(size_t)(k + 1) * p.R + c], 0.0); };
  // Both sweeps load the operands of kSolveSteps steps together (the steps
  // themselves are unchanged), so a dependent chain waits on memory once per
  // kSolveSteps steps rather than on every step.
  constexpr int U = kSolveSteps;
  double2 cur_d = diag(0);
  double2 cur_x = d2(-(b0 * wt.x), -(b0 * wt.y));
  double2 cur_du = m > 1 ? off(0) : d2(0.0, 0.0);
  bool singular = false;
  for (int k0 = 0; k0 < m - 1; k0 += U) {
    double av[U], bv[U + 1];   // al[k0 + 1 + u], be[k0 + 1 + u]
#pragma unroll
    for (int u = 0; u < U; ++u) av[u] = k0 + 1 + u < m ? p.al[(size_t)(k0 + 1 + u) * p.R + c] : 0.0;
#pragma unroll
    for (int u = 0; u <= U; ++u) bv[u] = k0 + 1 + u < m ? p.be[(size_t)(k0 + 1 + u) * p.R + c] : 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u;
      if (k >= m - 1) break;
      const double2 dl = d2(bv[u], 0.0);                                   // off(k)
      double2 next_d = d2(av[u] - ex.x, -ex.y), next_x = d2(0.0, 0.0);     // diag(k + 1)
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
      const size_t o = (size_t)k * S + sys;
      p.wD[o] = sD; p.wDU[o] = sDU; p.wDU2[o] = sDU2; p.wX[o] = sX;
      cur_d = next_d; cur_x = next_x; cur_du = next_du;
    }
  }
  if (cur_d.x == 0.0 && cur_d.y == 0.0) singular = true;
  double* yr = p.yr + (size_t)sys * m;
  double* yi = p.yi + (size_t)sys * m;
  if (singular) {
    atomicOr(p.status, 2u);
    for (int k = 0; k < m; ++k) { yr[k] = 0.0; yi[k] = 0.0; }
    return;
  }
  double2 x2 = cdiv(cur_x, cur_d);
  yr[m - 1] = x2.x;
  yi[m - 1] = x2.y;
  if (m > 1) {
    const size_t o = (size_t)(m - 2) * S + sys;
    double2 x1 = cdiv(csub(p.wX[o], cmul(p.wDU[o], x2)), p.wD[o]);
    yr[m - 2] = x1.x;
    yi[m - 2] = x1.y;
    for (int k1 = m - 3; k1 >= 0; k1 -= U) {
      double2 fX[U], fDU[U], fDU2[U], fD[U];   // step k = k1 - u
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (k1 - u >= 0) {
          const size_t q = (size_t)(k1 - u) * S + sys;
          fX[u] = p.wX[q]; fDU[u] = p.wDU[q]; fDU2[u] = p.wDU2[q]; fD[u] = p.wD[q];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k1 - u;
        if (k < 0) break;
        double2 t = csub(fX[u], cmul(fDU[u], x1));
        t = csub(t, cmul(fDU2[u], x2));
        const double2 xk = cdiv(t, fD[u]);
        yr[k] = xk.x;
        yi[k] = xk.y;
        x2 = x1;
        x1 = xk;
      }
    }
  }
}

// C[j][k][c] = sum_s (gsr_{j,s} Im y_{c,s,k} + gsi_{j,s} Re y_{c,s,k})
struct BCollapse {
  const double* yr;
  const double* yi;
  const double* gsr;
  const double* gsi;
  int R, m, ns, nd;
  double* C;
};

template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_bcollapse(BCollapse p) {
  const long long total = (long long)p.nd * p.m * p.R;
  for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < total;
       t += (long long)gridDim.x * kBlock) {
    const int c = (int)(t % p.R);
    const long long jk = t / p.R;
    const int k = (int)(jk % p.m), j = (int)(jk / p.m);
    double acc = 0.0;
    for (int s = 0; s < p.ns; ++s) {
      const size_t e = ((size_t)c * p.ns + s) * p.m + k;
      acc = fma(p.gsr[j + (size_t)p.nd * s], p.yi[e], acc);
      acc = fma(p.gsi[j + (size_t)p.nd * s], p.yr[e], acc);
    }
    p.C[t] = acc;
  }
}

// Z_j[i][c] = sum_k V[k][i][c] C[j][k][c]   (nd == 1: written to X with ld ldx)
struct BVC {
  const double* V;
  const double* C;
  int R, m, nd, nrows;
  double* Z;          // [j][N][R] or X
  size_t ldz;         // R, or the X row stride
  int accumulate;
};

// Z_j = sum_k V[k] C[j][k] per column.  One element per thread (W = 1), or
// W adjacent columns per thread (W = 2 or 4: R, the column offset and the row
// strides are multiples of W) with 16-byte loads, the Krylov slices' loads
// issued eight at a time.
template <int ND, int W = 1>
__global__ __launch_bounds__(kBlock) void k_bvc(BVC p) {
  if constexpr (W == 1) {
    const long long e = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (e >= (long long)p.nrows * p.R) return;
    const int c = (int)(e % p.R);
    const long long i = e / p.R;
    const size_t slice = (size_t)p.nrows * p.R;
    double acc[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) acc[j] = 0.0;
    for (int k = 0; k < p.m; ++k) {
      const double v = p.V[(size_t)k * slice + e];
#pragma unroll
      for (int j = 0; j < ND; ++j) acc[j] = fma(v, p.C[((size_t)j * p.m + k) * p.R + c], acc[j]);
    }
    if (ND == 1) {
      const size_t o = (size_t)i * p.ldz + c;
      p.Z[o] = p.accumulate ? p.Z[o] + acc[0] : acc[0];
    } else {
#pragma unroll
      for (int j = 0; j < ND; ++j) p.Z[(size_t)j * slice + e] = acc[j];
    }
  } else {
    static_assert(W == 2 || W == 4, "W is 1, 2 or 4");
    constexpr int H = W / 2;   // double2 loads per slice
    const long long e = (long long)W * ((long long)blockIdx.x * kBlock + threadIdx.x);
    if (e >= (long long)p.nrows * p.R) return;
    const int c = (int)(e % p.R);
    const long long i = e / p.R;
    const size_t slice = (size_t)p.nrows * p.R;
    double2 acc[ND][H];
#pragma unroll
    for (int j = 0; j < ND; ++j)
#pragma unroll
      for (int h = 0; h < H; ++h) acc[j][h] = d2(0.0, 0.0);
    constexpr int KB = 8;
    for (int k0 = 0; k0 < p.m; k0 += KB) {
      double2 v[KB][H];
#pragma unroll
      for (int t = 0; t < KB; ++t)
#pragma unroll
        for (int h = 0; h < H; ++h)
          v[t][h] = k0 + t < p.m ? *reinterpret_cast<const double2*>(
                                       p.V + (size_t)(k0 + t) * slice + e + 2 * h)
                                 : d2(0.0, 0.0);
#pragma unroll
      for (int t = 0; t < KB; ++t) {
        if (k0 + t >= p.m) break;
#pragma unroll
        for (int j = 0; j < ND; ++j)
#pragma unroll
          for (int h = 0; h < H; ++h) {
            const double2 cc = *reinterpret_cast<const double2*>(
                p.C + ((size_t)j * p.m + k0 + t) * p.R + c + 2 * h);
            acc[j][h].x = fma(v[t][h].x, cc.x, acc[j][h].x);
            acc[j][h].y = fma(v[t][h].y, cc.y, acc[j][h].y);
          }
      }
    }
#pragma unroll
    for (int h = 0; h < H; ++h) {
      if (ND == 1) {
        double2* o = reinterpret_cast<double2*>(p.Z + (size_t)i * p.ldz + c + 2 * h);
        if (p.accumulate) {
          const double2 z = *o;
          *o = d2(z.x + acc[0][h].x, z.y + acc[0][h].y);
        } else {
          *o = acc[0][h];
        }
      } else {
#pragma unroll
        for (int j = 0; j < ND; ++j)
          *reinterpret_cast<double2*>(p.Z + (size_t)j * slice + e + 2 * h) = acc[j][h];
      }
    }
  }
}

// Update per column pair: u = X0 + X1 (+ upd0 + upd1 with deflation), then
// the reaction -- Fisher-KPP (update_U.cu:8-16) or Mitchell-Schaeffer (gate
// h per problem, forward Euler with the new u, g = h u^2 (1-u)/tau_in -
// u/tau_out) -- and Usol = [u, dt g].
struct BUpdate {
  const double* X;    // ld ldu (chunk columns at offset)
  double* Usol;       // ld ldu
  const double* upd;  // ld ldu, or null (no deflation)
  const double* mass;
  double dt;
  int clamp, x_zero, prepare;
  size_t ldu;
  int R, nrows, chunk;
  int reaction;       // GM_REACTION_FISHER or GM_REACTION_MS
  double* ode;        // MS: gate h, [N][ldh] with the chunk's first problem at offset
  size_t ldh;
  double ms_tau_in, ms_tau_out, ms_tau_open, ms_tau_close, ms_v_gate;
  double* be;         // be[0][c]
  double* partials;
  unsigned* status;
};

template <int R>
__global__ __launch_bounds__(kBlock) void k_bupdate(BUpdate p) {
  constexpr int CPW = R < 64 ? R : 64;
  constexpr int RPW = 64 / CPW;
  constexpr int CL = R / CPW;
  static_assert(R % 2 == 0, "column pairs");
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int row0 = b * p.chunk;
  const int row1 = min(row0 + p.chunk, p.nrows);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int slot = lane / CPW, c0 = lane % CPW;
  for (int r = row0 + wave * RPW + slot; r < row1; r += kWaves * RPW) {
    const double ms = p.mass[r];
#pragma unroll
    for (int q = 0; q < CL; ++q) {
      const int c = c0 + q * CPW;
      const size_t i = (size_t)r * p.ldu + c;
      double val;
      if (p.prepare) {
        val = p.Usol[i];
      } else {
        // both columns of the pair are read by both lanes of the pair
        const size_t i0 = i - (c & 1);
        double u = p.x_zero ? 0.0 : (p.X[i0] + p.X[i0 + 1]);
        if (p.upd) u = (u + p.upd[i0]) + p.upd[i0 + 1];
        if (p.reaction == 2) {
          if (c & 1) {
            double* hp = p.ode + (size_t)r * p.ldh + (c >> 1);
            const double h = *hp;
            const double dh = u < p.ms_v_gate ? (1.0 - h) / p.ms_tau_open : -h / p.ms_tau_close;
            const double hn = h + p.dt * dh;
            *hp = hn;
            val = ((hn * (u * u) * (1.0 - u)) / p.ms_tau_in - u / p.ms_tau_out) * p.dt;
          } else {
            val = u;
          }
        } else {
          if (p.clamp) u = u * (u >= 0.0 ? 1.0 : 0.0);
          if (c & 1) {
            const double ub = u / ms;
            val = ((ms * ub) * (1.0 - ub)) * p.dt;
          } else {
            val = u;
          }
        }
      }
      if (!p.prepare) p.Usol[i] = val;   // X and Usol are distinct buffers
    }
  }
  // the column norms of the next step's starting vectors are taken per chunk
  // by k_bnorm
}

// ---------------------------------------------------------------------------
// Deflation of all R columns by the shared (Q, Lambda), lanczos_impl.cu:84-89
// and update_U.cu:18-27 per column:  Z = Q^T Usol,  Ud = Usol - Q Z,
// upd = Q diag(f) Z with f = exp_lambda (even columns) or phi_lambda (odd).
// ---------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));

// Block partials of Z = Q^T Usol on FP64 MFMA (v_mfma_f64_16x16x4f64): block
// x covers rows [x chunk, (x+1) chunk), each wavefront one 16 x 16 tile of Z
// (16 deflation vectors x 16 columns), the row dimension stepped by 4.
// Operand layout: A lane l = (row l%16, k l/16), B lane l = (k l/16, col
// l%16); D register i of lane l = (row 4i + l/16, col l%16).
struct BDeflQtU {
  const double* Q;    // [N][K]
  const double* U;    // [N][R]
  int K, R, nrows, chunk, KT, CT;
  double* partials;   // [G][K][R]
};

// Block partials of Q^T U on FP64 MFMA (v_mfma_f64_16x16x4f64; A = Q rows k,
// B = U columns c, D register i of lane l = (k 4i + l/16, c l%16)).  Each
// wavefront holds all KT k-tiles and CTW c-tiles of the block's column group
// (block y), so one load of Q feeds CTW chains and one of U feeds KT; the 4
// wavefronts take interleaved 4-row steps of the block's rows and are summed
// in LDS in wavefront order.  K <= 64.
template <int KT, int CTW>
__global__ __launch_bounds__(kBlock) void k_bdefl_qtu(BDeflQtU p) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ct0 = blockIdx.y * CTW;
  const int row0 = blockIdx.x * p.chunk, row1 = min(row0 + p.chunk, p.nrows);
  const int sub = lane >> 4;
  __shared__ double sh[KT * CTW * 4 * 64];
  d4 acc[KT][CTW];
#pragma unroll
  for (int kt = 0; kt < KT; ++kt)
#pragma unroll
    for (int ct = 0; ct < CTW; ++ct) acc[kt][ct] = d4{0.0, 0.0, 0.0, 0.0};
  // four 4-row steps a pass, all their loads issued together: out-of-range
  // operands load a valid element (row0, column R - 1, k K - 1) and are
  // replaced by zero afterwards, so no load sits behind a branch; the MFMA
  // chains take the steps in the same order as one step a pass
  constexpr int UR = 4;
  for (int r = row0 + 4 * wave; r < row1; r += 4 * kWaves * UR) {
    double a[UR][KT], b[UR][CTW];
#pragma unroll
    for (int u = 0; u < UR; ++u) {
      const int row = r + 4 * kWaves * u + sub;
      const size_t rc = row < row1 ? row : row0;
#pragma unroll
      for (int kt = 0; kt < KT; ++kt) a[u][kt] = p.Q[rc * p.K + min(kt * 16 + (lane & 15), p.K - 1)];
#pragma unroll
      for (int ct = 0; ct < CTW; ++ct)
        b[u][ct] = p.U[rc * p.R + min((ct0 + ct) * 16 + (lane & 15), p.R - 1)];
    }
#pragma unroll
    for (int u = 0; u < UR; ++u) {
      const bool ok = r + 4 * kWaves * u + sub < row1;
#pragma unroll
      for (int kt = 0; kt < KT; ++kt) a[u][kt] = (ok && kt * 16 + (lane & 15) < p.K) ? a[u][kt] : 0.0;
#pragma unroll
      for (int ct = 0; ct < CTW; ++ct)
        b[u][ct] = (ok && (ct0 + ct) * 16 + (lane & 15) < p.R) ? b[u][ct] : 0.0;
#pragma unroll
      for (int kt = 0; kt < KT; ++kt)
#pragma unroll
        for (int ct = 0; ct < CTW; ++ct)
          acc[kt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][kt], b[u][ct], acc[kt][ct], 0, 0, 0);
    }
  }
  for (int w = 0; w < kWaves; ++w) {
    if (wave == w) {
#pragma unroll
      for (int kt = 0; kt < KT; ++kt)
#pragma unroll
        for (int ct = 0; ct < CTW; ++ct)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            double& d = sh[((kt * CTW + ct) * 4 + i) * 64 + lane];
            d = w == 0 ? acc[kt][ct][i] : d + acc[kt][ct][i];
          }
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < KT * CTW * 4 * 64; t += kBlock) {
    const int ln = t & 63, i = (t >> 6) & 3, tile = t >> 8;
    const int k2 = (tile / CTW) * 16 + 4 * i + (ln >> 4);
    const int c = (ct0 + tile % CTW) * 16 + (ln & 15);
    if (k2 < p.K && c < p.R) p.partials[((size_t)blockIdx.x * p.K + k2) * p.R + c] = sh[t];
  }
}

// Launch over G row chunks (q.chunk rows each) and the column groups of up to
// four 16-column tiles.
template <int KT>
inline void launch_bdefl_qtu_kt(const BDeflQtU& q, int G, hipStream_t st) {
  const int CT = (q.R + 15) / 16;
  if (CT == 1) {
    hipLaunchKernelGGL((k_bdefl_qtu<KT, 1>), dim3(G, 1), dim3(kBlock), 0, st, q);
  } else if (CT == 2) {
    hipLaunchKernelGGL((k_bdefl_qtu<KT, 2>), dim3(G, 1), dim3(kBlock), 0, st, q);
  } else {
    hipLaunchKernelGGL((k_bdefl_qtu<KT, 4>), dim3(G, (CT + 3) / 4), dim3(kBlock), 0, st, q);
  }
}

inline void launch_bdefl_qtu(const BDeflQtU& q, int G, hipStream_t st) {
  switch ((q.K + 15) / 16) {
    case 1: launch_bdefl_qtu_kt<1>(q, G, st); break;
    case 2: launch_bdefl_qtu_kt<2>(q, G, st); break;
    case 3: launch_bdefl_qtu_kt<3>(q, G, st); break;
    default: launch_bdefl_qtu_kt<4>(q, G, st); break;
  }
}

// Ud = Usol - Q Z and upd = Q (f Z); block y covers 64 columns, whose Z and
// f Z are staged in LDS (2 K 64 doubles); a wavefront takes 64 / cpw rows of
// cpw lanes (cpw = the group's width rounded up to a power of two).
struct BDeflApply {
  const double* Q;    // [N][K]
  const double* Z;    // [K][R]
  const double* el;   // K
  const double* pl;   // K
  const double* Usol; // [N][R]
  double* Ud;
  double* upd;
  int K, R, nrows;
};

template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_bdefl_apply(BDeflApply p) {
  extern __shared__ double sz[];     // [K][64] Z, then [K][64] f Z
  const int cg = blockIdx.y * 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = threadIdx.x; t < p.K * 64; t += kBlock) {
    const int k = t / 64, cc = cg + t % 64;
    const double z = cc < p.R ? p.Z[(size_t)k * p.R + cc] : 0.0;
    sz[t] = z;
    sz[p.K * 64 + t] = z * ((cc & 1) ? p.pl[k] : p.el[k]);
  }
  __syncthreads();
  // a group narrower than 64 columns packs 64 / cpw rows into a wavefront
  const int width = min(64, p.R - cg);
  int cpw = 1;
  while (cpw < width) cpw <<= 1;
  const int rpw = 64 / cpw;
  const int slot = lane / cpw, cl = lane % cpw;
  if (cl >= width) return;   // no barrier follows
  const int c = cg + cl;
  for (int r = (blockIdx.x * kWaves + wave) * rpw + slot; r < p.nrows; r += gridDim.x * kWaves * rpw) {
    const double* q = p.Q + (size_t)r * p.K;
    double a = 0.0, u = 0.0;
    for (int k = 0; k < p.K; ++k) {
      const double qv = q[k];
      a = fma(qv, sz[k * 64 + cl], a);
      u = fma(qv, sz[p.K * 64 + k * 64 + cl], u);
    }
    const size_t i = (size_t)r * p.R + c;
    p.Ud[i] = p.Usol[i] - a;
    p.upd[i] = u;
  }
}

}  // namespace gmb
