// gm_adjoint.hip -- f(A)b with its adjoint (no reference counterpart).
//
// Forward: u = f(A) b by Lanczos on A (m steps, no preconditioner, no
// deflation) and the contour rule u = Im sum_j c_j x_j, x_j = (xi_j - A)^{-1} b
// ~= V ytilde_j, ytilde_j = beta0 (xi_j - T)^{-1} e1, c_j = s f(xi_j) dxi/dt_j.
//
// Backward, for an upstream gradient g = dL/du (A symmetric, so
// (xi - A)^T = xi - A and no conjugate appears):
//   dL/db       = f(A) g                      (the same machinery on g)
//   dL/dkappa_i = Im sum_j c_j z_j^T (dA/dkappa_i) x_j,  z_j = (xi_j - A)^{-1} g
//   dL/dalpha   = Im sum_j (dc_j/dalpha) g^T x_j
//   dL/dtheta   = Im sum_j c_j z_j^T (dA/dtheta) x_j   for a scalar parameter
//                 theta of A given dA/dtheta on A's pattern (gm_fab_backward_param)
// with dA/dkappa of the variable-coefficient model (edge conductance
// (kappa_i + kappa_k)/2, boundary terms linear in kappa_i):
//   z^T dA/dkappa_i x = sum_{k != i} w_ik (z_i - z_k)(x_i - x_k) + s_i z_i x_i,
//   w_ik = -A_ik / (kappa_i + kappa_k),  s_i = (sum_k A_ik) / kappa_i.
// The blocks X = V_b Ytilde_b and Z = V_g Ytilde_g (N x 2 n_s real columns)
// are the one dense contraction here and run on FP64 MFMA
// (v_mfma_f64_16x16x4f64).  With GM_FAB_RECOMPUTE the forward keeps only the
// tridiagonal data and the backward rebuilds V_b (bit-identical: the
// reductions are deterministic) instead of holding it between the passes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/gpumatfun.h"
#include "gm_bkernels.h"

namespace gma {
using namespace gm;

typedef double d4 __attribute__((ext_vector_type(4)));

// X[i][j] = sum_k V[k][i] Y[k][j], V column-major N x m (slices of N), Y row-major
// m x ncol (ncol a multiple of 64), X row-major N x ncol.  One wavefront per
// 16-row tile and 64 columns (4 MFMA 16x16 output tiles); K stepped by 4.
// MFMA f64 16x16x4 operand layout: A lane l = (row l%16, k l/16), B lane l =
// (k l/16, col l%16); D register r of lane l is (row 4r + l/16, col l%16)
// (the f64 form interleaves its output rows, unlike the f32 16x16x4 form,
// whose register r of lane l is row 4(l/16) + r).
__global__ __launch_bounds__(256) void k_vy_mfma(const double* __restrict__ V, int N, int m,
                                                 const double* __restrict__ Y, int ncol,
                                                 double* __restrict__ X) {
  const int lane = threadIdx.x & 63;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int i0 = tile * 16;
  const int ct = blockIdx.y;                 // 64-column group
  if (i0 >= N) return;
  d4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = d4{0.0, 0.0, 0.0, 0.0};
  const int ia = i0 + (lane & 15);
  for (int k0 = 0; k0 < m; k0 += 4) {
    const int k = k0 + (lane >> 4);
    const double a = (k < m && ia < N) ? V[(size_t)k * N + ia] : 0.0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int j = ct * 64 + t * 16 + (lane & 15);
      const double bv = (k < m) ? Y[(size_t)k * ncol + j] : 0.0;
      acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int j = ct * 64 + t * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + 4 * r + (lane >> 4);
      if (i < N) X[(size_t)i * ncol + j] = acc[t][r];
    }
  }
}

// Ytilde in GEMM layout: Yg[k][2s + 0/1] = Re/Im ytilde_s[k]; padding columns 0.
template <int kUnused = 0>
__global__ __launch_bounds__(256) void k_pack_y(const double* yr, const double* yi, int m, int ns,
                                                int ncol, double* Yg) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= m * ncol) return;
  const int k = t / ncol, j = t - k * ncol;
  const int s = j >> 1;
  double v = 0.0;
  if (s < ns) v = (j & 1) ? yi[(size_t)s * m + k] : yr[(size_t)s * m + k];
  Yg[t] = v;
}

// dL/dkappa_i = Im sum_j c_j [sum_k w_ik (z_i - z_k)(x_i - x_k) + s_i z_i x_i]
struct KGrad {
  const int* rowptr;
  const int* col;
  const double* val;
  const double* kappa;
  const double* X;     // [N][ncol] (re, im interleaved per node)
  const double* Z;
  const double* c;     // ns complex
  int N, ns, ncol;
  double* grad;
};

template <int kUnused = 0>
__global__ __launch_bounds__(256) void k_kappa_grad(KGrad p) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.N) return;
  const double ki = p.kappa[i];
  const double* xi = p.X + (size_t)i * p.ncol;
  const double* zi = p.Z + (size_t)i * p.ncol;
  double rowsum = 0.0;
  // accumulate T_j = sum_k w_ik (z_i - z_k)(x_i - x_k) + s_i z_i x_i per node j,
  // then grad = Im sum_j c_j T_j, in fixed order
  double g = 0.0;
  for (int e = p.rowptr[i]; e < p.rowptr[i + 1]; ++e) rowsum += p.val[e];
  const double si = rowsum / ki;
  for (int s = 0; s < p.ns; ++s) {
    const double xr = xi[2 * s], xim = xi[2 * s + 1];
    const double zr = zi[2 * s], zim = zi[2 * s + 1];
    // self term s_i z x
    double tr = si * (zr * xr - zim * xim);
    double tim = si * (zr * xim + zim * xr);
    for (int e = p.rowptr[i]; e < p.rowptr[i + 1]; ++e) {
      const int k = p.col[e];
      if (k == i) continue;
      const double w = -p.val[e] / (ki + p.kappa[k]);
      const double* xk = p.X + (size_t)k * p.ncol;
      const double* zk = p.Z + (size_t)k * p.ncol;
      const double dxr = xr - xk[2 * s], dxi = xim - xk[2 * s + 1];
      const double dzr = zr - zk[2 * s], dzi = zim - zk[2 * s + 1];
      tr = fma(w, dzr * dxr - dzi * dxi, tr);
      tim = fma(w, dzr * dxi + dzi * dxr, tim);
    }
    const double cr = p.c[2 * s], ci = p.c[2 * s + 1];
    g += cr * tim + ci * tr;      // Im(c * T)
  }
  p.grad[i] = g;
}

// Block partials of g^T X_j (2 ns real values per block), [G][2ns].
struct GX {
  const double* g;
  const double* X;
  int N, ncol, nv;   // nv = 2 ns
  double* partials;
};

template <int kUnused = 0>
__global__ __launch_bounds__(256) void k_gx(GX p) {
  __shared__ double sh[256];
  const int per = (p.N + gridDim.x - 1) / gridDim.x;
  const int r0 = blockIdx.x * per, r1 = min(p.N, r0 + per);
  for (int v = 0; v < p.nv; ++v) {
    double s = 0.0;
    for (int i = r0 + threadIdx.x; i < r1; i += 256) s = fma(p.g[i], p.X[(size_t)i * p.ncol + v], s);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) p.partials[(size_t)blockIdx.x * p.nv + v] = sh[0];
    __syncthreads();
  }
}

// dL/dtheta = Im sum_j c_j z_j^T E x_j for E = dA/dtheta (A's pattern): row i
// contributes Im sum_j c_j z_{j,i} (E x_j)_i.  One wavefront per row, lanes over
// the ncol real columns of X and Z (node j = column pair 2j, 2j+1), so each
// gathered row of X is one coalesced load; the nonzeros go 8 at a time (their
// gathers in flight together, the FMAs in column order).  The complex product
// takes the partner lane's value by a lane swap.  Block partials, fixed order.
struct PGrad {
  const int* rowptr;
  const int* col;
  const double* E;
  const double* X;     // [N][ncol]
  const double* Z;
  const double* c;     // ns complex
  int N, ns, ncol;
  double* partials;    // [G][2]
};

template <int kUnused = 0>
__global__ __launch_bounds__(256) void k_param_grad(PGrad p) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double acc = 0.0;
  for (int i = blockIdx.x * 4 + wave; i < p.N; i += gridDim.x * 4) {
    const int e0 = p.rowptr[i], e1 = p.rowptr[i + 1];
    for (int c0 = 0; c0 < p.ncol; c0 += 64) {
      const int cc = c0 + lane;
      double ex = 0.0;
      for (int eb = e0; eb < e1; eb += 8) {
        double ev[8], xv[8];
        int ci[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const bool ok = eb + t < e1;
          ev[t] = ok ? p.E[eb + t] : 0.0;
          ci[t] = ok ? p.col[eb + t] : i;
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) xv[t] = p.X[(size_t)ci[t] * p.ncol + cc];
#pragma unroll
        for (int t = 0; t < 8; ++t)
          if (eb + t < e1) ex = fma(ev[t], xv[t], ex);
      }
      const double z = p.Z[(size_t)i * p.ncol + cc];
      const double ex_o = __shfl_xor(ex, 1);
      const double z_o = __shfl_xor(z, 1);
      const int s = cc >> 1;
      if (!(lane & 1) && s < p.ns) {
        const double tr = z * ex - z_o * ex_o, ti = z * ex_o + z_o * ex;
        acc += p.c[2 * s] * ti + p.c[2 * s + 1] * tr;   // Im(c_j t)
      }
    }
  }
  block_partial2_at(d2(acc, 0.0), p.partials, blockIdx.x);
}

// Block partials of the dot products C[a][b] = sum_i P[a][i] R[b][i] over
// the block's rows, for na P-vectors (this block's 64 of them: blockIdx.y)
// and nb <= 64 R-vectors, each contiguous over N (vector stride ldp / ldr).
// Tiles of 32 rows of both sets are staged in LDS; thread t accumulates
// P-vector t % 64 against R-vectors t / 64, t / 64 + 4, ...
// partials[G][na * nb], index a * nb + b.
struct Dots {
  const double* P;
  size_t ldp;
  int na;
  const double* R;
  size_t ldr;
  int nb;
  int N, chunk;
  double* partials;
};

constexpr int kDotTile = 32;

template <int kUnused = 0>
__global__ __launch_bounds__(256) void k_dots(Dots p) {
  __shared__ double sp[64 * (kDotTile + 1)];
  __shared__ double sr[64 * (kDotTile + 1)];
  const int r0 = blockIdx.x * p.chunk, r1 = min(r0 + p.chunk, p.N);
  const int a0 = blockIdx.y * 64;
  const int na_here = min(64, p.na - a0);
  const int ta = threadIdx.x & 63, tb = threadIdx.x >> 6;
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.0;
  for (int t0 = r0; t0 < r1; t0 += kDotTile) {
    const int nt = min(kDotTile, r1 - t0);
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * kDotTile; e += 256) {
      const int v = e / kDotTile, i = e % kDotTile;
      sp[v * (kDotTile + 1) + i] =
          (v < na_here && i < nt) ? p.P[(size_t)(a0 + v) * p.ldp + t0 + i] : 0.0;
      sr[v * (kDotTile + 1) + i] = (v < p.nb && i < nt) ? p.R[(size_t)v * p.ldr + t0 + i] : 0.0;
    }
    __syncthreads();
    for (int i = 0; i < nt; ++i) {
      const double pv = sp[ta * (kDotTile + 1) + i];
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[q] = fma(pv, sr[(tb + 4 * q) * (kDotTile + 1) + i], acc[q]);
    }
  }
  if (ta < na_here)
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int b = tb + 4 * q;
      if (b < p.nb)
        p.partials[(size_t)blockIdx.x * p.na * p.nb + (size_t)(a0 + ta) * p.nb + b] = acc[q];
    }
}

// y[i] = x[i] + sgn sum_k Q[k][i] a[k] (scale[k] when given): the deflation
// projection (sgn = -1) and the eigen part of f(A) b (scale = f(lambda)).
struct DeflCM {
  const double* Q;
  int K, N;
  const double* a;
  const double* scale;
  double sgn;
  const double* x;
  double* y;
};

template <int kUnused = 0>
__global__ __launch_bounds__(256) void k_defl_cm(DeflCM p) {
  __shared__ double co[64];
  if ((int)threadIdx.x < p.K) co[threadIdx.x] = p.sgn * p.a[threadIdx.x] * (p.scale ? p.scale[threadIdx.x] : 1.0);
  __syncthreads();
  for (int i = blockIdx.x * 256 + threadIdx.x; i < p.N; i += gridDim.x * 256) {
    double acc = p.x[i];
    for (int k = 0; k < p.K; ++k) acc = fma(p.Q[(size_t)k * p.N + i], co[k], acc);
    p.y[i] = acc;
  }
}

}  // namespace gma

namespace {

thread_local std::string g_aerr;
std::mutex g_amu;
std::unordered_set<const void*> g_areg;

gm_status afail(gm_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_aerr = buf;
  return st;
}

#define A_HIP(call)                                                                      \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return afail(e_ == hipErrorOutOfMemory ? GM_ENOMEM : GM_EDEVICE, "%s: %s (%s:%d)", \
                   #call, hipGetErrorString(e_), __FILE__, __LINE__);                    \
  } while (0)
#define A_TRY(x)              \
  do {                        \
    gm_status s_ = (x);       \
    if (s_ != GM_OK) return s_; \
  } while (0)

inline int adiv(long long a, long long b) { return (int)((a + b - 1) / b); }

// One Krylov side (b or g): basis, tridiagonal data, ytilde.
struct Side {
  double *U = nullptr, *V = nullptr, *W0 = nullptr, *W1 = nullptr, *al = nullptr, *be = nullptr,
         *yr = nullptr, *yi = nullptr, *Yg = nullptr, *Xc = nullptr, *out = nullptr,
         *C = nullptr, *a = nullptr;   // V-coefficients of f(A) rhs; Q^T rhs (deflation)
  // the stream a Krylov run of this side is enqueued on and the block partials
  // and solve scratch it uses: the engine's, except while the backward
  // recomputes the b side's basis on the second stream with the b side's own
  hipStream_t st = nullptr;
  double *partials = nullptr, *partials2 = nullptr;
  double2 *wD = nullptr, *wDU = nullptr, *wDU2 = nullptr, *wX = nullptr;
};

}  // namespace

struct gm_fab {
  int dev = 0;
  hipStream_t stream = nullptr;
  // the second stream of the recomputed b-side run (it overlaps the g side's
  // run of a backward); ev1 orders it after the engine stream's earlier work,
  // ev2 marks its end
  hipStream_t stream2 = nullptr;
  hipEvent_t ev1 = nullptr, ev2 = nullptr;
  double *bpart = nullptr, *bpart2 = nullptr;   // the b side's partials while recomputing
  double2 *bwD = nullptr, *bwDU = nullptr, *bwDU2 = nullptr, *bwX = nullptr;
  int n = 0, m = 0, ns = 0, ncol = 0;
  long long nnz = 0;
  bool recompute = false, have_forward = false;
  int G = 1, chunk = 1;
  int *rowptr = nullptr, *col = nullptr;
  double *val = nullptr, *xi = nullptr, *ones = nullptr, *c = nullptr, *gsr = nullptr,
         *gsi = nullptr, *C = nullptr, *partials = nullptr, *kappa = nullptr, *grad = nullptr,
         *gxv = nullptr;
  double2 *wD = nullptr, *wDU = nullptr, *wDU2 = nullptr, *wX = nullptr;
  unsigned* status = nullptr;
  double* dA = nullptr;     // nnz values of dA/dtheta (gm_fab_backward_param)
  double* partials2 = nullptr;  // finish partials when the reductions are fused
  std::vector<double> dc;   // ns complex dc/dalpha (host)
  Side sb, sg;
  // deflation: Q (N x K column-major), f(lambda); E Q, dot partials and sums
  int K = 0, G_dot = 1, chunk_dot = 1;
  double *Q = nullptr, *flam = nullptr, *EQ = nullptr, *dpart = nullptr, *dsum = nullptr;
  std::vector<double> lam, flam_h, fdd;   // host copies (fdd K x K column-major)
  // parameter derivatives kept on the device (gm_fab_set_param_derivs): dA/dtheta_q
  // values, and with deflation E_q Q and Q^T E_q Q
  std::vector<double*> pdA, pEQ;
  std::vector<std::vector<double>> pQEQ;
};

namespace {

bool alive(const gm_fab* f) {
  std::lock_guard<std::mutex> lk(g_amu);
  return f && g_areg.count(f);
}

template <typename T>
gm_status aalloc(T** p, size_t count) {
  *p = nullptr;
  A_HIP(hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T)));
  return GM_OK;
}

void afree(void* p) {
  if (p) (void)hipFree(p);
}

void free_fab(gm_fab* f) {
  void* bufs[] = {f->rowptr, f->col, f->val, f->xi, f->ones, f->c, f->gsr, f->gsi, f->C,
                  f->partials, f->kappa, f->grad, f->gxv, f->wD, f->wDU, f->wDU2, f->wX,
                  f->status, f->dA, f->partials2, f->Q, f->flam, f->EQ, f->dpart, f->dsum};
  for (void* p : bufs) afree(p);
  for (double* p : f->pdA) afree(p);
  for (double* p : f->pEQ) afree(p);
  for (Side* s : {&f->sb, &f->sg}) {
    void* sbufs[] = {s->U, s->V, s->W0, s->W1, s->al, s->be, s->yr, s->yi, s->Yg, s->Xc, s->out,
                     s->C, s->a};
    for (void* p : sbufs) afree(p);
  }
  void* bbufs[] = {f->bpart, f->bpart2, f->bwD, f->bwDU, f->bwDU2, f->bwX};
  for (void* p : bbufs) afree(p);
  if (f->ev1) (void)hipEventDestroy(f->ev1);
  if (f->ev2) (void)hipEventDestroy(f->ev2);
  if (f->stream2) (void)hipStreamDestroy(f->stream2);
  if (f->stream) (void)hipStreamDestroy(f->stream);
}

// k_reduce of a side's partials (or `part`) on the side's stream
gm_status areduce_s(gm_fab* f, Side& s, int G, int nv, double* out, int do_sqrt, int flag_zero,
                    const double* part = nullptr) {
  gm::ReduceArgs r{};
  r.partials = part ? part : s.partials;
  r.G = G;
  r.nv = nv;
  r.out = out;
  r.stride = 1;
  r.do_sqrt = do_sqrt;
  r.flag_zero = flag_zero;
  r.status = f->status;
  hipLaunchKernelGGL(gm::k_reduce<0>, dim3(gm::reduce_grid(r.nv)), dim3(gm::kReduceBlock), 0, s.st, r);
  A_HIP(hipGetLastError());
  return GM_OK;
}

gm_status areduce(gm_fab* f, int G, int nv, double* out, int do_sqrt, int flag_zero,
                  const double* part = nullptr) {
  gm::ReduceArgs r{};
  r.partials = part ? part : f->partials;
  r.G = G;
  r.nv = nv;
  r.out = out;
  r.stride = 1;
  r.do_sqrt = do_sqrt;
  r.flag_zero = flag_zero;
  r.status = f->status;
  hipLaunchKernelGGL(gm::k_reduce<0>, dim3(gm::reduce_grid(r.nv)), dim3(gm::kReduceBlock), 0, f->stream, r);
  A_HIP(hipGetLastError());
  return GM_OK;
}

// Lanczos on side.U, shifted solves with unit weights (ytilde), and
// side.out = V * Im(sum_j c_j ytilde_j) = f(A) rhs.
gm_status krylov(gm_fab* f, Side& s, bool want_out) {
  const size_t N = (size_t)f->n;
  gmb::BNorm nn{};
  nn.u = s.U;
  nn.ldu = 1;
  nn.be = s.be;
  nn.nrows = f->n;
  nn.chunk = f->chunk;
  nn.partials = s.partials;
  nn.status = f->status;
  hipLaunchKernelGGL((gmb::k_bnorm<1>), dim3(f->G), dim3(gm::kBlock), 0, s.st, nn);
  A_HIP(hipGetLastError());
  A_TRY(areduce_s(f, s, f->G, 1, s.be, 1, 0));
  gmb::BSpmm base{};
  base.rowptr = f->rowptr;
  base.col = f->col;
  base.val = f->val;
  base.nrows = f->n;
  base.chunk = f->chunk;
  base.al = s.al;
  base.be = s.be;
  base.partials = s.partials;
  base.g0 = 1.0;
  double* Wb[2] = {s.W0, s.W1};
  // small grids: alpha_k and beta_{k+1} are summed inside the consuming
  // kernels instead of by k_reduce launches
  const bool fuse = f->G <= 1024;
  for (int k = 0; k < f->m; ++k) {
    gmb::BSpmm a = base;
    a.k = k;
    a.x = k == 0 ? s.U : Wb[(k - 1) & 1];
    a.ldx = 1;
    a.y = Wb[k & 1];
    a.ldy = 1;
    a.vk = s.V + (size_t)k * N;
    a.vkm1 = k > 0 ? s.V + (size_t)(k - 1) * N : nullptr;
    a.status = f->status;
    if (fuse && k > 0) {
      a.bpart = s.partials2;
      a.bG = f->G;
    }
    // one column: rows straight from global memory (the LDS-staged form measured
    // slower here: 18.1 vs 14.0 us on the 9-point 512^2 operator)
    hipLaunchKernelGGL((gmb::k_bspmm<1, gmb::kOnly>), dim3(f->G), dim3(gm::kBlock), 0, s.st, a);
    A_HIP(hipGetLastError());
    if (!fuse) A_TRY(areduce_s(f, s, f->G, 1, s.al + k, 0, 0));
    gmb::BFinish fn{};
    fn.w = Wb[k & 1];
    fn.vk = a.vk;
    fn.al = s.al;
    fn.be = s.be;
    fn.k = k;
    fn.nrows = f->n;
    fn.chunk = f->chunk;
    fn.partials = fuse ? s.partials2 : s.partials;
    fn.status = f->status;
    if (fuse) {
      fn.apart = s.partials;
      fn.aG = f->G;
      fn.al_out = s.al;
    }
    hipLaunchKernelGGL((gmb::k_bfinish<1>), dim3(f->G), dim3(gm::kBlock), 0, s.st, fn);
    A_HIP(hipGetLastError());
    if (!fuse || k == f->m - 1)
      A_TRY(areduce_s(f, s, f->G, 1, s.be + k + 1, 1, 1, fuse ? s.partials2 : nullptr));
  }
  gmb::BSolve sv{};
  sv.al = s.al;
  sv.be = s.be;
  sv.R = 1;
  sv.m = f->m;
  sv.ns = f->ns;
  sv.wts1 = f->ones;
  sv.wts2 = f->ones;
  sv.etaxi = f->xi;
  sv.wD = s.wD;
  sv.wDU = s.wDU;
  sv.wDU2 = s.wDU2;
  sv.wX = s.wX;
  sv.yr = s.yr;
  sv.yi = s.yi;
  sv.status = f->status;
  hipLaunchKernelGGL(gmb::k_bsolves<0>, dim3(adiv(f->ns, 64)), dim3(64), 0, s.st, sv);
  A_HIP(hipGetLastError());
  if (want_out) {
    gmb::BCollapse cc{};
    cc.yr = s.yr;
    cc.yi = s.yi;
    cc.gsr = f->gsr;
    cc.gsi = f->gsi;
    cc.R = 1;
    cc.m = f->m;
    cc.ns = f->ns;
    cc.nd = 1;
    cc.C = s.C;
    hipLaunchKernelGGL(gmb::k_bcollapse<0>, dim3(adiv(f->m, 256)), dim3(256), 0, s.st, cc);
    A_HIP(hipGetLastError());
    gmb::BVC v{};
    v.V = s.V;
    v.C = s.C;
    v.R = 1;
    v.m = f->m;
    v.nd = 1;
    v.nrows = f->n;
    v.Z = s.out;
    v.ldz = 1;
    v.accumulate = 0;
    hipLaunchKernelGGL((gmb::k_bvc<1>), dim3(adiv(f->n, 256)), dim3(256), 0, s.st, v);
    A_HIP(hipGetLastError());
  }
  return GM_OK;
}

// side.Xc = V * Ytilde (N x ncol) on FP64 MFMA.
// The backward's recomputed b-side run on the second stream, after the engine
// stream's earlier work, with the b side's own partials and solve scratch;
// join_b makes the engine stream wait for it.
gm_status recompute_b(gm_fab* f) {
  A_HIP(hipEventRecord(f->ev1, f->stream));
  A_HIP(hipStreamWaitEvent(f->stream2, f->ev1, 0));
  Side& s = f->sb;
  s.st = f->stream2;
  s.partials = f->bpart;
  s.partials2 = f->bpart2;
  s.wD = f->bwD;
  s.wDU = f->bwDU;
  s.wDU2 = f->bwDU2;
  s.wX = f->bwX;
  const gm_status rc = krylov(f, s, false);
  s.st = f->stream;
  s.partials = f->partials;
  s.partials2 = f->partials2;
  s.wD = f->wD;
  s.wDU = f->wDU;
  s.wDU2 = f->wDU2;
  s.wX = f->wX;
  if (rc != GM_OK) return rc;
  A_HIP(hipEventRecord(f->ev2, f->stream2));
  return GM_OK;
}

gm_status join_b(gm_fab* f) {
  A_HIP(hipStreamWaitEvent(f->stream, f->ev2, 0));
  return GM_OK;
}

gm_status resolvent_block(gm_fab* f, Side& s) {
  hipLaunchKernelGGL(gma::k_pack_y<0>, dim3(adiv((long long)f->m * f->ncol, 256)), dim3(256), 0,
                     f->stream, s.yr, s.yi, f->m, f->ns, f->ncol, s.Yg);
  A_HIP(hipGetLastError());
  dim3 grid(adiv(adiv(f->n, 16), 4), f->ncol / 64);
  hipLaunchKernelGGL(gma::k_vy_mfma, grid, dim3(256), 0, f->stream, s.V, f->n, f->m, s.Yg,
                     f->ncol, s.Xc);
  A_HIP(hipGetLastError());
  return GM_OK;
}

gm_status alloc_side(gm_fab* f, Side& s) {
  const size_t N = (size_t)f->n;
  A_TRY(aalloc(&s.U, N));
  A_TRY(aalloc(&s.V, N * f->m));
  A_TRY(aalloc(&s.W0, N));
  A_TRY(aalloc(&s.W1, N));
  A_TRY(aalloc(&s.al, (size_t)f->m));
  A_TRY(aalloc(&s.be, (size_t)f->m + 1));
  A_TRY(aalloc(&s.yr, (size_t)f->m * f->ns));
  A_TRY(aalloc(&s.yi, (size_t)f->m * f->ns));
  A_TRY(aalloc(&s.Yg, (size_t)f->m * f->ncol));
  A_TRY(aalloc(&s.Xc, N * f->ncol));
  A_TRY(aalloc(&s.out, N));
  A_TRY(aalloc(&s.C, (size_t)f->m));
  A_TRY(aalloc(&s.a, (size_t)std::max(f->K, 1)));
  return GM_OK;
}

// out (host, na x nb row-major) = P^T R for na P-vectors and nb R-vectors
// (contiguous over N, strides ldp / ldr), in a fixed summation order.
gm_status dots(gm_fab* f, const double* P, size_t ldp, int na, const double* R, size_t ldr, int nb,
               double* out_dev, std::vector<double>* out_host) {
  gma::Dots d{};
  d.P = P;
  d.ldp = ldp;
  d.na = na;
  d.R = R;
  d.ldr = ldr;
  d.nb = nb;
  d.N = f->n;
  d.chunk = f->chunk_dot;
  d.partials = f->dpart;
  hipLaunchKernelGGL(gma::k_dots<0>, dim3(f->G_dot, adiv(na, 64)), dim3(256), 0, f->stream, d);
  A_HIP(hipGetLastError());
  A_TRY(areduce(f, f->G_dot, na * nb, out_dev, 0, 0, f->dpart));
  if (out_host) {
    out_host->resize((size_t)na * nb);
    A_HIP(hipMemcpyAsync(out_host->data(), out_dev, out_host->size() * sizeof(double),
                         hipMemcpyDeviceToHost, f->stream));
    A_HIP(hipStreamSynchronize(f->stream));
  }
  return GM_OK;
}

// Deflation of a side's start vector: a = Q^T U, U <- U - Q a.
gm_status deflate(gm_fab* f, Side& s) {
  A_TRY(dots(f, f->Q, (size_t)f->n, f->K, s.U, (size_t)f->n, 1, s.a, nullptr));
  gma::DeflCM d{f->Q, f->K, f->n, s.a, nullptr, -1.0, s.U, s.U};
  hipLaunchKernelGGL(gma::k_defl_cm<0>, dim3(std::min(4096, adiv(f->n, 256))), dim3(256), 0,
                     f->stream, d);
  A_HIP(hipGetLastError());
  return GM_OK;
}

// s.out += Q (f(lambda) * a): the eigen part of f(A) rhs.
gm_status add_eigen_part(gm_fab* f, Side& s) {
  gma::DeflCM d{f->Q, f->K, f->n, s.a, f->flam, 1.0, s.out, s.out};
  hipLaunchKernelGGL(gma::k_defl_cm<0>, dim3(std::min(4096, adiv(f->n, 256))), dim3(256), 0,
                     f->stream, d);
  A_HIP(hipGetLastError());
  return GM_OK;
}

// Solves (shift I - T) h = r for the m x m Lanczos tridiagonal T (diagonal
// al, off-diagonal be[1..m-1]) by Gaussian elimination with partial pivoting
// (row interchanges between neighbours, as in LAPACK's dgtsv).
void tridiag_shift_solve(int m, const double* al, const double* be, double shift,
                         std::vector<double>& r) {
  std::vector<double> dl(m), d(m), du(m), du2(m, 0.0);
  for (int i = 0; i < m; ++i) {
    d[i] = shift - al[i];
    dl[i] = i + 1 < m ? -be[i + 1] : 0.0;   // below the diagonal, row i+1
    du[i] = i + 1 < m ? -be[i + 1] : 0.0;   // above the diagonal, row i
  }
  for (int i = 0; i + 1 < m; ++i) {
    if (std::fabs(d[i]) >= std::fabs(dl[i])) {
      const double fct = d[i] != 0.0 ? dl[i] / d[i] : 0.0;
      d[i + 1] -= fct * du[i];
      r[i + 1] -= fct * r[i];
      dl[i] = 0.0;
    } else {   // swap rows i and i+1
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

// EQ = E Q (K SpMVs with E's values on A's pattern) and, on the host,
// QEQ = Q^T E Q (K x K row-major).
gm_status eq_products(gm_fab* f, const double* E, double* EQ, std::vector<double>& QEQ) {
  const size_t N = (size_t)f->n;
  for (int k = 0; k < f->K; ++k) {
    gmb::BSpmm a{};
    a.rowptr = f->rowptr;
    a.col = f->col;
    a.val = E;
    a.nrows = f->n;
    a.chunk = f->chunk;
    a.x = f->Q + (size_t)k * N;
    a.ldx = 1;
    a.y = EQ + (size_t)k * N;
    a.ldy = 1;
    a.g0 = 1.0;
    a.ca = 1.0;
    hipLaunchKernelGGL((gmb::k_bspmm<1, gmb::kCheb>), dim3(f->G), dim3(gm::kBlock), 0, f->stream, a);
    A_HIP(hipGetLastError());
  }
  return dots(f, f->Q, N, f->K, EQ, N, f->K, f->dsum, &QEQ);
}

}  // namespace

extern "C" {

const char* gm_fab_last_error(void) { return g_aerr.c_str(); }

gm_status gm_fab_create(const gm_fab_desc* d, gm_fab** out) {
  g_aerr.clear();
  if (!d || !out) return afail(GM_EINVAL, "null argument");
  *out = nullptr;
  if (d->n <= 0 || !d->A_ptr || !d->A_ind || !d->A_values || d->subspace_size < 1 ||
      d->contour_size < 1 || !d->xi || !d->wts)
    return afail(GM_EINVAL, "bad sizes or null arrays");
  if (d->sparse_format != GM_SPARSE_CSR)
    return afail(GM_EINVAL, "the adjoint needs A in CSR with its diagonal stored");
  if (d->A_ptr[0] != 0 || d->A_ptr[d->n] != d->nnz) return afail(GM_EINVAL, "bad pointer array");
  for (int i = 0; i < d->n; ++i)
    if (d->A_ptr[i + 1] < d->A_ptr[i]) return afail(GM_EINVAL, "pointer array not monotone");
  for (long long j = 0; j < d->nnz; ++j)
    if (d->A_ind[j] < 0 || d->A_ind[j] >= d->n) return afail(GM_EINVAL, "index out of range");
  if (d->num_eigs < 0 || d->num_eigs > 64 || d->num_eigs > d->n)
    return afail(GM_EINVAL, "num_eigs must be in [0, min(64, n)]");
  if (d->num_eigs > 0 && (!d->Q || !d->lambda || !d->f_lambda))
    return afail(GM_EINVAL, "deflation needs Q, lambda and f_lambda");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= d->device || d->device < 0)
    return afail(GM_EDEVICE, "no HIP device %d", d->device);
  gm_fab* f = new gm_fab();
  auto bail = [&](gm_status st) {
    std::string keep = g_aerr;
    free_fab(f);
    delete f;
    g_aerr = keep;
    return st;
  };
#define A_CREATE(x)                  \
  do {                               \
    gm_status s_ = (x);              \
    if (s_ != GM_OK) return bail(s_); \
  } while (0)
  f->dev = d->device;
  if (hipSetDevice(f->dev) != hipSuccess || hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(afail(GM_EDEVICE, "device/stream"));
  f->n = d->n;
  f->nnz = d->nnz;
  f->m = d->subspace_size;
  f->ns = d->contour_size;
  f->ncol = adiv(2 * f->ns, 64) * 64;
  f->recompute = (d->flags & GM_FAB_RECOMPUTE) != 0;
  f->G = std::max(1, std::min(4096, adiv(f->n, 256)));
  f->chunk = adiv(f->n, f->G);
  hipStream_t st = f->stream;
  auto up = [&](auto** p, const auto* src, size_t cnt) -> gm_status {
    A_TRY(aalloc(p, cnt));
    A_HIP(hipMemcpyAsync(*p, src, cnt * sizeof(**p), hipMemcpyHostToDevice, st));
    return GM_OK;
  };
  A_CREATE(up(&f->rowptr, d->A_ptr, (size_t)f->n + 1));
  A_CREATE(up(&f->col, d->A_ind, (size_t)f->nnz));
  A_CREATE(up(&f->val, d->A_values, (size_t)f->nnz));
  A_CREATE(up(&f->xi, d->xi, 2 * (size_t)f->ns));
  std::vector<double> ones(2 * (size_t)f->ns, 0.0), gr(f->ns), gi(f->ns);
  for (int s = 0; s < f->ns; ++s) {
    ones[s] = 1.0;            // wts layout ns x 2 [Re Im]: Re = 1, Im = 0
    gr[s] = d->wts[2 * s];
    gi[s] = d->wts[2 * s + 1];
  }
  A_CREATE(up(&f->ones, ones.data(), ones.size()));
  A_CREATE(up(&f->c, d->wts, 2 * (size_t)f->ns));
  A_CREATE(up(&f->gsr, gr.data(), gr.size()));
  A_CREATE(up(&f->gsi, gi.data(), gi.size()));
  f->dc.assign(2 * (size_t)f->ns, 0.0);
  if (d->dwts) std::copy(d->dwts, d->dwts + 2 * f->ns, f->dc.begin());
  A_CREATE(aalloc(&f->C, (size_t)f->m));
  A_CREATE(aalloc(&f->partials, (size_t)4096 * std::max(2, f->ncol)));
  A_CREATE(aalloc(&f->partials2, (size_t)4096));
  A_CREATE(aalloc(&f->kappa, (size_t)f->n));
  A_CREATE(aalloc(&f->grad, (size_t)f->n));
  A_CREATE(aalloc(&f->gxv, (size_t)f->ncol));
  const size_t ny = (size_t)f->m * f->ns;
  A_CREATE(aalloc(&f->wD, ny));
  A_CREATE(aalloc(&f->wDU, ny));
  A_CREATE(aalloc(&f->wDU2, ny));
  A_CREATE(aalloc(&f->wX, ny));
  A_CREATE(aalloc(&f->status, 1));
  if (hipMemsetAsync(f->status, 0, sizeof(unsigned), st) != hipSuccess)
    return bail(afail(GM_EDEVICE, "memset"));
  f->K = d->num_eigs;
  if (f->K > 0) {
    const size_t NK = (size_t)f->n * f->K;
    A_CREATE(up(&f->Q, d->Q, NK));
    A_CREATE(up(&f->flam, d->f_lambda, (size_t)f->K));
    A_CREATE(aalloc(&f->EQ, NK));
    f->lam.assign(d->lambda, d->lambda + f->K);
    f->flam_h.assign(d->f_lambda, d->f_lambda + f->K);
    if (d->f_divdiff) f->fdd.assign(d->f_divdiff, d->f_divdiff + (size_t)f->K * f->K);
  }
  f->G_dot = std::max(1, std::min(2048, adiv(f->n, 128)));
  f->chunk_dot = adiv(f->n, f->G_dot);
  f->G_dot = adiv(f->n, f->chunk_dot);
  if (f->K > 0) {
    const int na = std::max(f->m, f->K);
    A_CREATE(aalloc(&f->dpart, (size_t)f->G_dot * na * f->K));
    A_CREATE(aalloc(&f->dsum, (size_t)na * f->K));
  }
  A_CREATE(alloc_side(f, f->sb));
  A_CREATE(alloc_side(f, f->sg));
  for (Side* sd : {&f->sb, &f->sg}) {
    sd->st = f->stream;
    sd->partials = f->partials;
    sd->partials2 = f->partials2;
    sd->wD = f->wD;
    sd->wDU = f->wDU;
    sd->wDU2 = f->wDU2;
    sd->wX = f->wX;
  }
  A_CREATE(aalloc(&f->bpart, (size_t)4096 * std::max(2, f->ncol)));
  A_CREATE(aalloc(&f->bpart2, (size_t)4096));
  A_CREATE(aalloc(&f->bwD, ny));
  A_CREATE(aalloc(&f->bwDU, ny));
  A_CREATE(aalloc(&f->bwDU2, ny));
  A_CREATE(aalloc(&f->bwX, ny));
  if (hipStreamCreateWithFlags(&f->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&f->ev1, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&f->ev2, hipEventDisableTiming) != hipSuccess)
    return bail(afail(GM_EDEVICE, "second stream/events"));
  if (hipStreamSynchronize(st) != hipSuccess) return bail(afail(GM_EDEVICE, "sync"));
  {
    std::lock_guard<std::mutex> lk(g_amu);
    g_areg.insert(f);
  }
  *out = f;
  return GM_OK;
#undef A_CREATE
}

gm_status gm_fab_delete(gm_fab* f) {
  {
    std::lock_guard<std::mutex> lk(g_amu);
    if (!f || !g_areg.count(f)) return afail(GM_EBADHANDLE, "invalid fab handle");
    g_areg.erase(f);
  }
  (void)hipSetDevice(f->dev);
  if (f->stream2) (void)hipStreamSynchronize(f->stream2);
  if (f->stream) (void)hipStreamSynchronize(f->stream);
  free_fab(f);
  delete f;
  return GM_OK;
}

gm_status gm_fab_set_matrix(gm_fab* f, const double* A_values) {
  if (!alive(f)) return afail(GM_EBADHANDLE, "invalid fab handle");
  if (!A_values) return afail(GM_EINVAL, "null values");
  A_HIP(hipSetDevice(f->dev));
  A_HIP(hipMemcpyAsync(f->val, A_values, (size_t)f->nnz * sizeof(double), hipMemcpyHostToDevice,
                       f->stream));
  A_HIP(hipStreamSynchronize(f->stream));
  f->have_forward = false;
  return GM_OK;
}

gm_status gm_fab_set_weights(gm_fab* f, const double* wts, const double* dwts) {
  if (!alive(f)) return afail(GM_EBADHANDLE, "invalid fab handle");
  if (!wts) return afail(GM_EINVAL, "null weights");
  A_HIP(hipSetDevice(f->dev));
  std::vector<double> gr(f->ns), gi(f->ns);
  for (int s = 0; s < f->ns; ++s) {
    gr[s] = wts[2 * s];
    gi[s] = wts[2 * s + 1];
  }
  A_HIP(hipMemcpyAsync(f->c, wts, 2 * (size_t)f->ns * sizeof(double), hipMemcpyHostToDevice, f->stream));
  A_HIP(hipMemcpyAsync(f->gsr, gr.data(), f->ns * sizeof(double), hipMemcpyHostToDevice, f->stream));
  A_HIP(hipMemcpyAsync(f->gsi, gi.data(), f->ns * sizeof(double), hipMemcpyHostToDevice, f->stream));
  A_HIP(hipStreamSynchronize(f->stream));
  if (dwts) std::copy(dwts, dwts + 2 * f->ns, f->dc.begin());
  f->have_forward = false;
  return GM_OK;
}

gm_status gm_fab_forward(gm_fab* f, const double* b, double* u) {
  if (!alive(f)) return afail(GM_EBADHANDLE, "invalid fab handle");
  if (!b) return afail(GM_EINVAL, "null vector");
  A_HIP(hipSetDevice(f->dev));
  const size_t N = (size_t)f->n;
  A_HIP(hipMemcpyAsync(f->sb.U, b, N * sizeof(double), hipMemcpyDefault, f->stream));
  if (f->K) A_TRY(deflate(f, f->sb));
  A_TRY(krylov(f, f->sb, true));
  if (u) {
    if (f->K) A_TRY(add_eigen_part(f, f->sb));
    A_HIP(hipMemcpyAsync(u, f->sb.out, N * sizeof(double), hipMemcpyDefault, f->stream));
  }
  A_HIP(hipStreamSynchronize(f->stream));
  f->have_forward = true;
  unsigned stv = 0;
  A_HIP(hipMemcpy(&stv, f->status, sizeof(unsigned), hipMemcpyDeviceToHost));
  if (stv & 6u) return afail(GM_EBREAKDOWN, "Lanczos breakdown in the forward pass (%u)", stv);
  return GM_OK;
}

gm_status gm_fab_backward(gm_fab* f, const double* g, const double* kappa, double* grad_b,
                          double* grad_kappa, double* grad_alpha) {
  if (!alive(f)) return afail(GM_EBADHANDLE, "invalid fab handle");
  if (!f->have_forward) return afail(GM_ESTATE, "backward needs a forward pass first");
  if (!g) return afail(GM_EINVAL, "null upstream gradient");
  if (grad_kappa && !kappa) return afail(GM_EINVAL, "kappa is needed for the kappa gradient");
  if (f->K && (grad_kappa || grad_alpha))
    return afail(GM_EINVAL, "the kappa and alpha gradients are built without deflation");
  A_HIP(hipSetDevice(f->dev));
  const size_t N = (size_t)f->n;
  hipStream_t st = f->stream;
  // with recompute, V_b is rebuilt (deterministic) on the second stream while
  // the z side runs on the engine stream
  const bool rb = f->recompute && (grad_kappa || grad_alpha);
  if (rb) A_TRY(recompute_b(f));
  A_HIP(hipMemcpyAsync(f->sg.U, g, N * sizeof(double), hipMemcpyDefault, st));
  // z side: f(A) g is dL/db
  if (f->K) A_TRY(deflate(f, f->sg));
  A_TRY(krylov(f, f->sg, grad_b != nullptr));
  if (f->K && grad_b) A_TRY(add_eigen_part(f, f->sg));
  if (grad_b)
    A_HIP(hipMemcpyAsync(grad_b, f->sg.out, N * sizeof(double), hipMemcpyDefault, st));
  if (grad_kappa || grad_alpha) {
    if (rb) A_TRY(join_b(f));
    A_TRY(resolvent_block(f, f->sb));
  }
  if (grad_kappa) {
    A_TRY(resolvent_block(f, f->sg));
    A_HIP(hipMemcpyAsync(f->kappa, kappa, N * sizeof(double), hipMemcpyDefault, st));
    gma::KGrad kg{};
    kg.rowptr = f->rowptr;
    kg.col = f->col;
    kg.val = f->val;
    kg.kappa = f->kappa;
    kg.X = f->sb.Xc;
    kg.Z = f->sg.Xc;
    kg.c = f->c;
    kg.N = f->n;
    kg.ns = f->ns;
    kg.ncol = f->ncol;
    kg.grad = f->grad;
    hipLaunchKernelGGL(gma::k_kappa_grad<0>, dim3(adiv(f->n, 256)), dim3(256), 0, st, kg);
    A_HIP(hipGetLastError());
    A_HIP(hipMemcpyAsync(grad_kappa, f->grad, N * sizeof(double), hipMemcpyDefault, st));
  }
  if (grad_alpha) {
    const int G = std::max(1, std::min(1024, adiv(f->n, 256)));
    gma::GX gx{};
    gx.g = f->sg.U;
    gx.X = f->sb.Xc;
    gx.N = f->n;
    gx.ncol = f->ncol;
    gx.nv = 2 * f->ns;
    gx.partials = f->partials;
    hipLaunchKernelGGL(gma::k_gx<0>, dim3(G), dim3(256), 0, st, gx);
    A_HIP(hipGetLastError());
    A_TRY(areduce(f, G, 2 * f->ns, f->gxv, 0, 0));
    std::vector<double> gxv(2 * (size_t)f->ns);
    A_HIP(hipMemcpyAsync(gxv.data(), f->gxv, gxv.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    A_HIP(hipStreamSynchronize(st));
    double ga = 0.0;
    for (int s = 0; s < f->ns; ++s) {
      const double dr = f->dc[2 * s], di = f->dc[2 * s + 1];
      const double xr = gxv[2 * s], xi = gxv[2 * s + 1];
      ga += dr * xi + di * xr;       // Im(dc * g^T x)
    }
    *grad_alpha = ga;
  }
  A_HIP(hipStreamSynchronize(st));
  unsigned stv = 0;
  A_HIP(hipMemcpy(&stv, f->status, sizeof(unsigned), hipMemcpyDeviceToHost));
  if (stv & 6u) return afail(GM_EBREAKDOWN, "Lanczos breakdown in the backward pass (%u)", stv);
  return GM_OK;
}

gm_status gm_fab_backward_param(gm_fab* f, const double* g, int32_t nparams,
                                const double* const* dA_values, double* grad_b,
                                double* grad_theta) {
  if (!alive(f)) return afail(GM_EBADHANDLE, "invalid fab handle");
  if (!f->have_forward) return afail(GM_ESTATE, "backward needs a forward pass first");
  if (!g || nparams < 0 || (nparams > 0 && !grad_theta))
    return afail(GM_EINVAL, "null upstream gradient or parameter arrays");
  const bool stored = nparams > 0 && !dA_values;
  if (stored && nparams != (int)f->pdA.size())
    return afail(GM_EINVAL, "%d parameters asked, %d stored by gm_fab_set_param_derivs", nparams,
                 (int)f->pdA.size());
  for (int q = 0; q < nparams && !stored; ++q)
    if (!dA_values[q]) return afail(GM_EINVAL, "null dA/dtheta values for parameter %d", q);
  if (f->K && nparams > 0 && f->fdd.empty())
    return afail(GM_EINVAL, "the parameter gradient with deflation needs f_divdiff");
  A_HIP(hipSetDevice(f->dev));
  const size_t N = (size_t)f->n;
  hipStream_t st = f->stream;
  A_HIP(hipMemcpyAsync(f->sg.U, g, N * sizeof(double), hipMemcpyDefault, st));
  if (f->K) A_TRY(deflate(f, f->sg));
  A_TRY(krylov(f, f->sg, grad_b != nullptr || (f->K && nparams > 0)));
  if (f->K && grad_b) A_TRY(add_eigen_part(f, f->sg));
  if (grad_b)
    A_HIP(hipMemcpyAsync(grad_b, f->sg.out, N * sizeof(double), hipMemcpyDefault, st));
  // deflation: the Krylov data of both sides, Q^T b, Q^T g and, per side,
  // H[:, k] = (lambda_k - T)^{-1} (f(lambda_k) beta0 e1 - C), the V-coefficients
  // of g_k(A) P rhs with g_k(x) = (f(lambda_k) - f(x)) / (lambda_k - x)
  const int m = f->m, K = f->K;
  std::vector<double> Hb, Hg, ab, ag;
  if (K && nparams > 0) {
    for (int side = 0; side < 2; ++side) {
      Side& sd = side ? f->sg : f->sb;
      std::vector<double> al(m), be(m + 1), Cv(m), a(K);
      A_HIP(hipMemcpyAsync(al.data(), sd.al, m * sizeof(double), hipMemcpyDeviceToHost, st));
      A_HIP(hipMemcpyAsync(be.data(), sd.be, (m + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
      A_HIP(hipMemcpyAsync(Cv.data(), sd.C, m * sizeof(double), hipMemcpyDeviceToHost, st));
      A_HIP(hipMemcpyAsync(a.data(), sd.a, K * sizeof(double), hipMemcpyDeviceToHost, st));
      A_HIP(hipStreamSynchronize(st));
      std::vector<double> H((size_t)m * K), r(m);
      for (int k = 0; k < K; ++k) {
        for (int j = 0; j < m; ++j) r[j] = -Cv[j];
        r[0] += f->flam_h[k] * be[0];
        tridiag_shift_solve(m, al.data(), be.data(), f->lam[k], r);
        for (int j = 0; j < m; ++j) H[(size_t)j * K + k] = r[j];
      }
      (side ? Hg : Hb).swap(H);
      (side ? ag : ab).swap(a);
    }
  }
  if (nparams > 0) {
    if (f->recompute) A_TRY(krylov(f, f->sb, false));
    A_TRY(resolvent_block(f, f->sb));
    A_TRY(resolvent_block(f, f->sg));
    if (!stored && !f->dA) A_TRY(aalloc(&f->dA, (size_t)f->nnz));
    const int G = std::max(1, std::min(4096, adiv(f->n, 32)));   // ~8 rows per wavefront
    for (int q = 0; q < nparams; ++q) {
      const double* Edev = stored ? f->pdA[q] : f->dA;
      if (!stored)
        A_HIP(hipMemcpyAsync(f->dA, dA_values[q], (size_t)f->nnz * sizeof(double),
                             hipMemcpyHostToDevice, st));
      gma::PGrad pg{};
      pg.rowptr = f->rowptr;
      pg.col = f->col;
      pg.E = Edev;
      pg.X = f->sb.Xc;
      pg.Z = f->sg.Xc;
      pg.c = f->c;
      pg.N = f->n;
      pg.ns = f->ns;
      pg.ncol = f->ncol;
      pg.partials = f->partials;
      hipLaunchKernelGGL(gma::k_param_grad<0>, dim3(G), dim3(256), 0, st, pg);
      A_HIP(hipGetLastError());
      A_TRY(areduce(f, G, 2, f->gxv, 0, 0));
      double out[2];
      A_HIP(hipMemcpyAsync(out, f->gxv, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
      A_HIP(hipStreamSynchronize(st));   // dA_values[q] and out are host memory
      double gt = out[0];
      if (K) {
        std::vector<double> QEQt, Mb, Mg;
        const double* EQ = stored ? f->pEQ[q] : f->EQ;
        if (!stored) A_TRY(eq_products(f, f->dA, f->EQ, QEQt));
        const std::vector<double>& QEQ = stored ? f->pQEQ[q] : QEQt;
        A_TRY(dots(f, f->sb.V, N, m, EQ, N, K, f->dsum, &Mb));
        A_TRY(dots(f, f->sg.V, N, m, EQ, N, K, f->dsum, &Mg));
        // Q-Q block: sum_kl (Q^T g)_k f[l_k, l_l] (Q^T E Q)_kl (Q^T b)_l
        for (int k = 0; k < K; ++k)
          for (int l = 0; l < K; ++l)
            gt += ag[k] * f->fdd[(size_t)l * K + k] * QEQ[(size_t)k * K + l] * ab[l];
        // cross blocks: (Q^T g)_k q_k^T E V_b H_b[:, k] + (Q^T b)_k q_k^T E V_g H_g[:, k]
        for (int k = 0; k < K; ++k) {
          double tb = 0.0, tg = 0.0;
          for (int j = 0; j < m; ++j) {
            tb += Mb[(size_t)j * K + k] * Hb[(size_t)j * K + k];
            tg += Mg[(size_t)j * K + k] * Hg[(size_t)j * K + k];
          }
          gt += ag[k] * tb + ab[k] * tg;
        }
      }
      grad_theta[q] = gt;
    }
  }
  A_HIP(hipStreamSynchronize(st));
  unsigned stv = 0;
  A_HIP(hipMemcpy(&stv, f->status, sizeof(unsigned), hipMemcpyDeviceToHost));
  if (stv & 6u) return afail(GM_EBREAKDOWN, "Lanczos breakdown in the backward pass (%u)", stv);
  return GM_OK;
}

gm_status gm_fab_set_param_derivs(gm_fab* f, int32_t nparams, const double* const* dA_values) {
  if (!alive(f)) return afail(GM_EBADHANDLE, "invalid fab handle");
  if (nparams < 0 || (nparams > 0 && !dA_values)) return afail(GM_EINVAL, "bad parameter arrays");
  for (int q = 0; q < nparams; ++q)
    if (!dA_values[q]) return afail(GM_EINVAL, "null dA/dtheta values for parameter %d", q);
  A_HIP(hipSetDevice(f->dev));
  A_HIP(hipStreamSynchronize(f->stream));
  for (double* p : f->pdA) afree(p);
  for (double* p : f->pEQ) afree(p);
  f->pdA.clear();
  f->pEQ.clear();
  f->pQEQ.clear();
  for (int q = 0; q < nparams; ++q) {
    double* E = nullptr;
    A_TRY(aalloc(&E, (size_t)f->nnz));
    f->pdA.push_back(E);
    A_HIP(hipMemcpyAsync(E, dA_values[q], (size_t)f->nnz * sizeof(double), hipMemcpyHostToDevice,
                         f->stream));
    if (f->K) {
      double* EQ = nullptr;
      A_TRY(aalloc(&EQ, (size_t)f->n * f->K));
      f->pEQ.push_back(EQ);
      f->pQEQ.emplace_back();
      A_TRY(eq_products(f, E, EQ, f->pQEQ.back()));
    }
  }
  A_HIP(hipStreamSynchronize(f->stream));
  return GM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// gm_fab_forward_noalloc: u = f(A) b exactly as gm_fab_forward computes it
// (K = 0), for callers that cannot allocate device memory.  Every buffer
// beyond the caller's inputs and outputs lives in `scratch`, a file-scope
// device array of kNoallocScratchBytes that each GPU holds from the library's
// first use on it until the process ends; its device address comes from
// hipGetSymbolAddress once per device, with that device current.  All calls
// on one GPU share the array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kNoallocScratchBytes = size_t(64) << 20;
constexpr size_t kNoallocScratchDoubles = kNoallocScratchBytes / sizeof(double);
constexpr int kNoallocMaxDevices = 64;

}  // namespace

static __device__ __align__(256) double scratch[kNoallocScratchDoubles];

namespace {

std::mutex g_scratch_mu;
double* g_scratch_ptr[kNoallocMaxDevices] = {};

// The device address of `scratch` on the current device (`dev`).
gm_status scratch_address(int dev, double** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return afail(GM_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                   hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return GM_OK;
}

// Clears the status word and derives the unit solve weights (ns x 2 [Re Im])
// and the split contour weights gsr, gsi from the caller's wts: a fill kernel
// rather than hipMemsetAsync, so the sequence can be captured into a graph.
__global__ __launch_bounds__(256) void k_noalloc_prepare(const double* __restrict__ wts, int ns,
                                                         double* ones, double* gsr, double* gsi,
                                                         unsigned* status) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) *status = 0u;
  if (s < ns) {
    ones[s] = 1.0;
    ones[ns + s] = 0.0;
    gsr[s] = wts[2 * s];
    gsi[s] = wts[2 * s + 1];
  }
}

// Carves `count` doubles from the scratch cursor, 256-byte aligned.
double* carve(double*& cur, size_t count) {
  double* p = cur;
  cur += (std::max<size_t>(count, 1) + 31) & ~size_t(31);
  return p;
}

size_t carve_size(size_t count) { return (std::max<size_t>(count, 1) + 31) & ~size_t(31); }

}  // namespace

extern "C" gm_status gm_fab_forward_noalloc(int32_t n, int64_t nnz, const int32_t* A_ptr,
                                            const int32_t* A_ind, const double* A_values,
                                            int32_t subspace_size, int32_t contour_size,
                                            const double* xi, const double* wts, const double* b,
                                            double* u, void* stream, int32_t device,
                                            uint32_t* status) {
  g_aerr.clear();
  if (n <= 0 || nnz < 0 || subspace_size < 1 || contour_size < 1)
    return afail(GM_EINVAL, "bad sizes (n %d, nnz %lld, m %d, n_s %d)", n, (long long)nnz,
                 subspace_size, contour_size);
  if (!A_ptr || !A_ind || !A_values || !xi || !wts || !b || !u)
    return afail(GM_EINVAL, "null device array");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count ||
      device >= kNoallocMaxDevices)
    return afail(GM_EDEVICE, "no HIP device %d", device);
  const size_t N = (size_t)n, m = (size_t)subspace_size, ns = (size_t)contour_size;
  const size_t ny = m * ns;
  const size_t need = carve_size(N * m) + 2 * carve_size(N) + carve_size(m) + carve_size(m + 1) +
                      2 * carve_size(ny) + carve_size(m) + 2 * carve_size(4096) +
                      4 * carve_size(2 * ny) + carve_size(2 * ns) + 2 * carve_size(ns) +
                      carve_size(1);
  if (need > kNoallocScratchDoubles)
    return afail(GM_EINVAL,
                 "n = %d, m = %d, n_s = %d needs %zu bytes of scratch; the limit is %zu bytes "
                 "(about n (m + 2) <= %zu doubles)",
                 n, subspace_size, contour_size, need * sizeof(double), kNoallocScratchBytes,
                 kNoallocScratchDoubles);
  int prev = -1;
  A_HIP(hipGetDevice(&prev));
  if (prev != device) A_HIP(hipSetDevice(device));
  struct Restore {
    int prev, dev;
    ~Restore() {
      if (prev >= 0 && prev != dev) (void)hipSetDevice(prev);
    }
  } restore{prev, device};
  double* base = nullptr;
  A_TRY(scratch_address(device, &base));
  hipStream_t st = static_cast<hipStream_t>(stream);

  // the buffers of gm_fab's Side and engine, carved from the scratch array
  double* cur = base;
  double* V = carve(cur, N * m);
  double* W0 = carve(cur, N);
  double* W1 = carve(cur, N);
  double* al = carve(cur, m);
  double* be = carve(cur, m + 1);
  double* yr = carve(cur, ny);
  double* yi = carve(cur, ny);
  double* C = carve(cur, m);
  double* partials = carve(cur, 4096);
  double* partials2 = carve(cur, 4096);
  double2* wD = reinterpret_cast<double2*>(carve(cur, 2 * ny));
  double2* wDU = reinterpret_cast<double2*>(carve(cur, 2 * ny));
  double2* wDU2 = reinterpret_cast<double2*>(carve(cur, 2 * ny));
  double2* wX = reinterpret_cast<double2*>(carve(cur, 2 * ny));
  double* ones = carve(cur, 2 * ns);
  double* gsr = carve(cur, ns);
  double* gsi = carve(cur, ns);
  unsigned* stat = reinterpret_cast<unsigned*>(carve(cur, 1));

  const int G = std::max(1, std::min(4096, adiv(n, 256)));
  const int chunk = adiv(n, G);
  hipLaunchKernelGGL(k_noalloc_prepare, dim3(adiv(contour_size, 256)), dim3(256), 0, st, wts,
                     contour_size, ones, gsr, gsi, stat);
  A_HIP(hipGetLastError());
  auto reduce = [&](int nv, double* out, int do_sqrt, int flag_zero,
                    const double* part) -> gm_status {
    gm::ReduceArgs r{};
    r.partials = part;
    r.G = G;
    r.nv = nv;
    r.out = out;
    r.stride = 1;
    r.do_sqrt = do_sqrt;
    r.flag_zero = flag_zero;
    r.status = stat;
    hipLaunchKernelGGL(gm::k_reduce<0>, dim3(gm::reduce_grid(r.nv)), dim3(gm::kReduceBlock), 0,
                       st, r);
    A_HIP(hipGetLastError());
    return GM_OK;
  };

  // the launch sequence of krylov() in gm_fab_forward, reading b in place
  gmb::BNorm nn{};
  nn.u = b;
  nn.ldu = 1;
  nn.be = be;
  nn.nrows = n;
  nn.chunk = chunk;
  nn.partials = partials;
  nn.status = stat;
  hipLaunchKernelGGL((gmb::k_bnorm<1>), dim3(G), dim3(gm::kBlock), 0, st, nn);
  A_HIP(hipGetLastError());
  A_TRY(reduce(1, be, 1, 0, partials));
  gmb::BSpmm sb{};
  sb.rowptr = A_ptr;
  sb.col = A_ind;
  sb.val = A_values;
  sb.nrows = n;
  sb.chunk = chunk;
  sb.al = al;
  sb.be = be;
  sb.partials = partials;
  sb.g0 = 1.0;
  double* Wb[2] = {W0, W1};
  const bool fuse = G <= 1024;
  for (int k = 0; k < subspace_size; ++k) {
    gmb::BSpmm a = sb;
    a.k = k;
    a.x = k == 0 ? b : Wb[(k - 1) & 1];
    a.ldx = 1;
    a.y = Wb[k & 1];
    a.ldy = 1;
    a.vk = V + (size_t)k * N;
    a.vkm1 = k > 0 ? V + (size_t)(k - 1) * N : nullptr;
    a.status = stat;
    if (fuse && k > 0) {
      a.bpart = partials2;
      a.bG = G;
    }
    hipLaunchKernelGGL((gmb::k_bspmm<1, gmb::kOnly>), dim3(G), dim3(gm::kBlock), 0, st, a);
    A_HIP(hipGetLastError());
    if (!fuse) A_TRY(reduce(1, al + k, 0, 0, partials));
    gmb::BFinish fn{};
    fn.w = Wb[k & 1];
    fn.vk = a.vk;
    fn.al = al;
    fn.be = be;
    fn.k = k;
    fn.nrows = n;
    fn.chunk = chunk;
    fn.partials = fuse ? partials2 : partials;
    fn.status = stat;
    if (fuse) {
      fn.apart = partials;
      fn.aG = G;
      fn.al_out = al;
    }
    hipLaunchKernelGGL((gmb::k_bfinish<1>), dim3(G), dim3(gm::kBlock), 0, st, fn);
    A_HIP(hipGetLastError());
    if (!fuse || k == subspace_size - 1)
      A_TRY(reduce(1, be + k + 1, 1, 1, fuse ? partials2 : partials));
  }
  gmb::BSolve sv{};
  sv.al = al;
  sv.be = be;
  sv.R = 1;
  sv.m = subspace_size;
  sv.ns = contour_size;
  sv.wts1 = ones;
  sv.wts2 = ones;
  sv.etaxi = xi;
  sv.wD = wD;
  sv.wDU = wDU;
  sv.wDU2 = wDU2;
  sv.wX = wX;
  sv.yr = yr;
  sv.yi = yi;
  sv.status = stat;
  hipLaunchKernelGGL(gmb::k_bsolves<0>, dim3(adiv(contour_size, 64)), dim3(64), 0, st, sv);
  A_HIP(hipGetLastError());
  gmb::BCollapse cc{};
  cc.yr = yr;
  cc.yi = yi;
  cc.gsr = gsr;
  cc.gsi = gsi;
  cc.R = 1;
  cc.m = subspace_size;
  cc.ns = contour_size;
  cc.nd = 1;
  cc.C = C;
  hipLaunchKernelGGL(gmb::k_bcollapse<0>, dim3(adiv(subspace_size, 256)), dim3(256), 0, st, cc);
  A_HIP(hipGetLastError());
  gmb::BVC v{};
  v.V = V;
  v.C = C;
  v.R = 1;
  v.m = subspace_size;
  v.nd = 1;
  v.nrows = n;
  v.Z = u;
  v.ldz = 1;
  v.accumulate = 0;
  hipLaunchKernelGGL((gmb::k_bvc<1>), dim3(adiv(n, 256)), dim3(256), 0, st, v);
  A_HIP(hipGetLastError());
  if (status) {
    unsigned sv_host = 0;
    A_HIP(hipMemcpyAsync(&sv_host, stat, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    A_HIP(hipStreamSynchronize(st));
    *status = sv_host;
    if (sv_host & 6u) return afail(GM_EBREAKDOWN, "Lanczos breakdown (%u)", sv_host);
  }
  return GM_OK;
}
