// gm_device.h -- device helpers shared by the kernels of the f(A)b engine.
//
// Deterministic reductions: each block reduces its rows in a fixed tree and
// writes one partial; the partials are then summed in a fixed order, either by
// k_reduce launched after the producing kernel or, on small grids, by every
// block of the consuming kernel itself (block_reduce_partials2), so results do
// not depend on scheduling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gm {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ double2 d2(double a, double b) { return make_double2(a, b); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block sum of a double2; result valid in thread 0.  Fixed order.
__device__ __forceinline__ double2 block_sum2(double2 v) {
  __shared__ double2 sh[kWaves];
  v.x = wave_sum(v.x);
  v.y = wave_sum(v.y);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double2 r = d2(0.0, 0.0);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kWaves; ++i) { r.x += sh[i].x; r.y += sh[i].y; }
  }
  return r;
}

// Block sum of a double2 written by thread 0 to partials[2b], partials[2b+1].
// The grid-wide total is formed by k_reduce after the kernel ends: a kernel
// boundary is the cheapest cross-XCD synchronisation (an in-kernel ticket
// needs a device-scope release fence per block, i.e. an L2 write-back on
// every XCD).
__device__ __forceinline__ void block_partial2_at(double2 v, double* partials, int idx) {
  const double2 bs = block_sum2(v);
  if (threadIdx.x == 0) {
    partials[2 * idx] = bs.x;
    partials[2 * idx + 1] = bs.y;
  }
}

__device__ __forceinline__ void block_partial2(double2 v, double* partials) {
  block_partial2_at(v, partials, blockIdx.x);
}

// Sum of G [G][2] partials written by the previous kernel, in a fixed order
// (thread t sums g = t, t + kBlock, ..., then the block tree), returned to
// every thread.  Every block that calls it gets the same bits; used instead
// of a separate k_reduce launch when G is small.
__device__ __forceinline__ double2 block_reduce_partials2(const double* part, int G) {
  __shared__ double2 bc;
  double2 s = d2(0.0, 0.0);
  for (int g = threadIdx.x; g < G; g += kBlock) {
    s.x += part[2 * g];
    s.y += part[2 * g + 1];
  }
  s = block_sum2(s);
  if (threadIdx.x == 0) bc = s;
  __syncthreads();
  const double2 r = bc;
  __syncthreads();
  return r;
}

// Sums nv values over G blocks of partials ([G][nv]) in a fixed order and
// writes out[v * stride] (square-rooted when do_sqrt).  One block of
// kReduceBlock threads: ns = kReduceBlock / W threads per value each sum a
// strided subset with four independent accumulators (so four loads are in
// flight), then a fixed-order LDS tree combines the ns sub-sums.  Status
// bits: 4 = non-finite result, 1 = exact zero (when flag_zero).
constexpr int kReduceBlock = 1024;

// Steps of a tridiagonal sweep whose operands are loaded together in the
// shifted solves.
// (16 spills the back-substitution operands to scratch and measured slower)
constexpr int kSolveSteps = 8;

// Segmented form (seg != nullptr): one block per segment q = blockIdx.x sums
// partials seg[q] .. seg[q+1]-1 into out + q * out_pstride.
struct ReduceArgs {
  const double* partials;
  int G, nv;
  double* out;
  int stride;
  int do_sqrt, flag_zero;
  unsigned* status;
  const int* seg;
  size_t out_pstride;
};

// Unsegmented grid for nv values: one block per kReduceBlock of them.
inline int reduce_grid(int nv) {
  const int g = (nv + kReduceBlock - 1) / kReduceBlock;
  return g < 1 ? 1 : (g > 256 ? 256 : g);
}

// (a template so that both translation units may instantiate it)
template <int kUnused = 0>
__global__ __launch_bounds__(kReduceBlock) void k_reduce(ReduceArgs p) {
  __shared__ double sub[kReduceBlock];
  const int t = threadIdx.x;
  const double* part = p.partials;
  double* out = p.out;
  int G = p.G;
  // unsegmented, the blocks split the values in runs of kReduceBlock (each
  // value's sum is the same whatever the grid)
  int vstart = blockIdx.x * kReduceBlock, vstep = gridDim.x * kReduceBlock;
  if (p.seg) {
    const int q = blockIdx.x, g0 = p.seg[q];
    part += (size_t)g0 * p.nv;
    G = p.seg[q + 1] - g0;
    out += (size_t)q * p.out_pstride;
    vstart = 0;
    vstep = kReduceBlock;
  }
  for (int v0 = vstart; v0 < p.nv; v0 += vstep) {
    const int W = min(p.nv - v0, kReduceBlock);
    const int ns = kReduceBlock / W;
    const int vv = t % W, sb = t / W, v = v0 + vv;
    double s = 0.0;
    if (sb < ns) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      int b = sb;
      // 16 loads in flight, then the same additions in the same order as the
      // loop below (the partials may sit in another XCD's L2 or in HBM)
      for (; b + 15 * ns < G; b += 16 * ns) {
        double w[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) w[u] = part[(size_t)(b + u * ns) * p.nv + v];
#pragma unroll
        for (int u = 0; u < 16; u += 4) {
          s0 += w[u];
          s1 += w[u + 1];
          s2 += w[u + 2];
          s3 += w[u + 3];
        }
      }
      for (; b + 3 * ns < G; b += 4 * ns) {
        s0 += part[(size_t)b * p.nv + v];
        s1 += part[(size_t)(b + ns) * p.nv + v];
        s2 += part[(size_t)(b + 2 * ns) * p.nv + v];
        s3 += part[(size_t)(b + 3 * ns) * p.nv + v];
      }
      for (; b < G; b += ns) s0 += part[(size_t)b * p.nv + v];
      s = (s0 + s1) + (s2 + s3);
    }
    __syncthreads();
    sub[t] = s;
    __syncthreads();
    int h = 1;
    while (h < ns) h <<= 1;
    for (h >>= 1; h > 0; h >>= 1) {
      if (sb < h && sb + h < ns) sub[t] += sub[t + h * W];
      __syncthreads();
    }
    if (t < W) {
      double r = sub[t];
      if (p.do_sqrt) {
        r = sqrt(r);
        if (p.status) {
          if (!isfinite(r))
            atomicOr(p.status, 4u);
          else if (p.flag_zero && r == 0.0)
            atomicOr(p.status, 1u);
        }
      }
      out[(size_t)v * p.stride] = r;
    }
  }
}

// Blocks are dispatched round-robin over the 8 XCDs; remap so that the
// blocks resident on one XCD own one contiguous range of rows (their x
// gathers then share that XCD's L2).  Bijective on [0, G).
__device__ __forceinline__ int xcd_remap(int b, int G) {
  const int q = G >> 3, r = G & 7;
  const int x = b & 7, idx = b >> 3;
  return x < r ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

__device__ __forceinline__ double safe_inv(double b) {
  return (b != 0.0 && isfinite(b)) ? 1.0 / b : 0.0;
}
__device__ __forceinline__ double safe_div(double a, double b) {
  return (b != 0.0 && isfinite(b)) ? a / b : 0.0;
}


__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return d2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return d2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double cabs1(double2 a) { return fabs(a.x) + fabs(a.y); }
// Smith's algorithm.
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, den = b.x + b.y * r;
    return d2((a.x + a.y * r) / den, (a.y - a.x * r) / den);
  }
  const double r = b.x / b.y, den = b.x * r + b.y;
  return d2((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}

}  // namespace gm
