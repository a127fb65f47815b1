// gm_eigs.hip -- the smallest eigenpairs of a symmetric sparse A on the GPU,
// for the deflation vectors Q and Lambda (no reference counterpart: the
// reference takes them from MATLAB's eigs(K, M, 'sm') with a sparse
// factorisation, process_nonuniform.m:106-110).
//
// Chebyshev-filtered subspace iteration (Zhou & Saad, "A Chebyshev-Davidson
// algorithm for large symmetric eigenproblems", SIAM J. Matrix Anal. Appl. 29,
// 2007; the scaled three-term filter of Zhou, Saad, Tiago & Chelikowsky, J.
// Comput. Phys. 219, 2006) on a block of s >= K columns:
//   1. Y = p(A) X with p the degree-d Chebyshev polynomial of the first kind
//      on [a, b] (damping the spectrum above a, the largest Ritz value of the
//      block, up to the upper bound b; amplifying below a), by the scaled
//      three-term recurrence, each term one batched SpMM (gmb::k_bspmm kCheb,
//      row blocks staged in LDS);
//   2. orthonormalise: G = Y^T Y on FP64 MFMA, G = V diag(g) V^T on the
//      host, X = Y V diag(g)^(-1/2);
//   3. Rayleigh-Ritz: H = X^T (A X) on FP64 MFMA, H = W diag(theta) W^T,
//      X <- X W;
//   4. residuals ||A x_i - theta_i x_i|| of the K wanted columns; stop when
//      every one is below tol |theta_i|.
// The s x s eigenproblems are solved on the host by cyclic Jacobi rotations.
// The filter separates lambda_K from the rest of the block at a rate set by
// (lambda_s - lambda_K) / (b - lambda_s): fast for grid operators, slow when
// the largest eigenvalue is many orders above the wanted ones (sliver
// elements), where a shift-invert method with a factorisation is the tool.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/gpumatfun.h"
#include "gm_bkernels.h"

namespace gme {
using namespace gm;

// out[r][c] = sum_k in[r][k] W[k][c] (s x s W staged in LDS): a wavefront
// covers 64 / s rows, lane = (row slot, column).
struct Combine {
  const double* in;
  const double* W;
  double* out;
  int s, nrows;
};

template <int S>
__global__ __launch_bounds__(kBlock) void k_combine(Combine p) {
  __shared__ double sW[S * S];
  for (int t = threadIdx.x; t < S * S; t += kBlock) sW[t] = p.W[t];
  __syncthreads();
  constexpr int RPW = 64 / S;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int slot = lane / S, c = lane % S;
  for (int r = (blockIdx.x * kWaves + wave) * RPW + slot; r < p.nrows;
       r += gridDim.x * kWaves * RPW) {
    const double* row = p.in + (size_t)r * S;
    double acc = 0.0;
#pragma unroll 8
    for (int k = 0; k < S; ++k) acc = fma(row[k], sW[k * S + c], acc);
    p.out[(size_t)r * S + c] = acc;
  }
}

// Start block: x[i] uniform in [-1, 1) from a counter-based generator
// (splitmix64 of seed + i; Steele, Lea & Flood, OOPSLA 2014), the same for
// a given seed on any device or grid.
struct Fill {
  double* x;
  size_t count;
  unsigned long long seed;
};

template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_fill_uniform(Fill p) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < p.count;
       i += (size_t)gridDim.x * kBlock) {
    unsigned long long z = p.seed + 0x9e3779b97f4a7c15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    p.x[i] = (double)(z >> 11) * 0x1.0p-52 - 1.0;
  }
}

// q[k][i] = x[i][k] for k < K (the first K columns of the N x S block as a
// column-major n x K array).
struct Extract {
  const double* x;
  double* q;
  int S, K, nrows;
};

template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_extract(Extract p) {
  const int k = blockIdx.y;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < p.nrows; i += gridDim.x * kBlock)
    p.q[(size_t)k * p.nrows + i] = p.x[(size_t)i * p.S + k];
}

}  // namespace gme

namespace {

thread_local std::string g_eerr;

gm_status efail(gm_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_eerr = buf;
  return st;
}

#define E_HIP(call)                                                                      \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return efail(e_ == hipErrorOutOfMemory ? GM_ENOMEM : GM_EDEVICE, "%s: %s (%s:%d)", \
                   #call, hipGetErrorString(e_), __FILE__, __LINE__);                    \
  } while (0)
#define E_TRY(x)              \
  do {                        \
    gm_status s_ = (x);       \
    if (s_ != GM_OK) return s_; \
  } while (0)

inline int ediv(long long a, long long b) { return (int)((a + b - 1) / b); }

// Symmetric eigendecomposition A = V diag(w) V^T (n x n row-major) by
// cyclic Jacobi rotations (Golub & Van Loan, Matrix Computations, 4th ed.,
// Alg. 8.5.3); w ascending, V's columns the eigenvectors.
void sym_eig(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  double norm2 = 0.0;
  for (double v : A) norm2 += v * v;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += A[(size_t)i * n + j] * A[(size_t)i * n + j];
    if (off <= 1e-34 * norm2) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(tau * tau + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {   // A <- A J (columns p, q)
          const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * akp - s * akq;
          A[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {   // A <- J^T A (rows p, q)
          const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * apk - s * aqk;
          A[(size_t)q * n + k] = s * apk + c * aqk;
        }
        A[(size_t)p * n + q] = A[(size_t)q * n + p] = 0.0;
        for (int k = 0; k < n; ++k) {   // V <- V J
          const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(),
                   [&](int a, int b) { return A[(size_t)a * n + a] < A[(size_t)b * n + b]; });
  std::vector<double> Vs((size_t)n * n);
  w.resize(n);
  for (int j = 0; j < n; ++j) {
    w[j] = A[(size_t)idx[j] * n + idx[j]];
    for (int k = 0; k < n; ++k) Vs[(size_t)k * n + j] = V[(size_t)k * n + idx[j]];
  }
  V.swap(Vs);
}

// Device state of one solve; X, Y, Z and AX are N x s row-major blocks whose
// roles rotate between stages (the pointers are permuted, never the data).
struct Eigs {
  hipStream_t st = nullptr;
  int n = 0, s = 0, G_sp = 1, chunk_sp = 1, G_df = 1, chunk_df = 1, G_rb = 1;
  int *rowptr = nullptr, *col = nullptr, *rowblk = nullptr;
  // slab SpMM at >= 32 columns (gmb::k_bspmm_xs): row-block descriptors, grid
  int* rowblk_xs = nullptr;
  int nrb_xs = 0, G_xs = 8;
  double *val = nullptr, *X = nullptr, *Y = nullptr, *Z = nullptr, *AX = nullptr,
         *partials = nullptr, *small = nullptr, *wmat = nullptr, *shift = nullptr;
  ~Eigs() {
    void* bufs[] = {rowptr, col, rowblk, rowblk_xs, val, X, Y, Z, AX, partials, small, wmat, shift};
    for (void* p : bufs)
      if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  }
};

template <typename T>
gm_status ealloc(T** p, size_t count) {
  E_HIP(hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T)));
  return GM_OK;
}

// out (s x s, host, row-major) = P^T U for P, U N x s row-major, summed in
// a fixed order (block partials on MFMA, then k_reduce) and symmetrised.
gm_status gram(Eigs& e, const double* P, const double* U, std::vector<double>& out) {
  gmb::BDeflQtU q{};
  q.Q = P;
  q.U = U;
  q.K = e.s;
  q.R = e.s;
  q.nrows = e.n;
  q.chunk = e.chunk_df;
  q.KT = ediv(e.s, 16);
  q.CT = ediv(e.s, 16);
  q.partials = e.partials;
  gmb::launch_bdefl_qtu(q, e.G_df, e.st);
  E_HIP(hipGetLastError());
  gm::ReduceArgs r{};
  r.partials = e.partials;
  r.G = e.G_df;
  r.nv = e.s * e.s;
  r.out = e.small;
  r.stride = 1;
  hipLaunchKernelGGL(gm::k_reduce<0>, dim3(gm::reduce_grid(r.nv)), dim3(gm::kReduceBlock), 0, e.st, r);
  E_HIP(hipGetLastError());
  out.resize((size_t)e.s * e.s);
  E_HIP(hipMemcpyAsync(out.data(), e.small, out.size() * sizeof(double), hipMemcpyDeviceToHost, e.st));
  E_HIP(hipStreamSynchronize(e.st));
  for (int i = 0; i < e.s; ++i)
    for (int j = i + 1; j < e.s; ++j) {
      const double v = 0.5 * (out[(size_t)i * e.s + j] + out[(size_t)j * e.s + i]);
      out[(size_t)i * e.s + j] = out[(size_t)j * e.s + i] = v;
    }
  return GM_OK;
}

// out = in W (N x S times S x S; W from the host, row-major).
template <int S>
gm_status combine(Eigs& e, const double* in, const std::vector<double>& W, double* out) {
  E_HIP(hipMemcpyAsync(e.wmat, W.data(), W.size() * sizeof(double), hipMemcpyHostToDevice, e.st));
  gme::Combine c{};
  c.in = in;
  c.W = e.wmat;
  c.out = out;
  c.s = S;
  c.nrows = e.n;
  const int rows_per_block = gm::kWaves * (64 / S);
  hipLaunchKernelGGL(gme::k_combine<S>, dim3(std::max(1, std::min(8192, ediv(e.n, rows_per_block)))),
                     dim3(gm::kBlock), 0, e.st, c);
  E_HIP(hipGetLastError());
  E_HIP(hipStreamSynchronize(e.st));   // W is host memory the caller reuses
  return GM_OK;
}

// y = ca (A x - shift x) + cb z, all N x S row-major; shift = shiftc[c] or cc.
template <int S>
gm_status cheb_step(Eigs& e, const double* x, const double* z, double* y, double ca, double cb,
                    double cc, const double* shiftc) {
  gmb::BSpmm a{};
  a.rowptr = e.rowptr;
  a.col = e.col;
  a.val = e.val;
  a.nrows = e.n;
  a.chunk = e.chunk_sp;
  a.rowblk = e.rowblk;
  a.nrowblk = e.G_rb;
  a.x = x;
  a.ldx = S;
  a.y = y;
  a.ldy = S;
  a.z = z;
  a.ldz = S;
  a.g0 = 1.0;
  a.ca = ca;
  a.cb = cb;
  a.cc = cc;
  a.shiftc = shiftc;
  if constexpr (S >= 32) {
    // column slabs of 32 (16-byte accesses, double-buffered row blocks)
    a.rowblk = e.rowblk_xs;
    a.nrowblk = e.nrb_xs;
    hipLaunchKernelGGL((gmb::k_bspmm_xs<S, gmb::kCheb, false>), dim3(e.G_xs), dim3(gm::kBlock), 0,
                       e.st, a);
  } else {
    hipLaunchKernelGGL((gmb::k_bspmm<S, gmb::kCheb, true>), dim3(std::min(e.G_rb, 8192)),
                       dim3(gm::kBlock), 0, e.st, a);
  }
  E_HIP(hipGetLastError());
  return GM_OK;
}

// Column 2-norms of the N x S array y.
template <int S>
gm_status colnorms(Eigs& e, const double* y, std::vector<double>& nrm) {
  gmb::BNorm b{};
  b.u = y;
  b.ldu = S;
  b.nrows = e.n;
  b.chunk = e.chunk_sp;
  b.partials = e.partials;
  hipLaunchKernelGGL((gmb::k_bnorm<S>), dim3(e.G_sp), dim3(gm::kBlock), 0, e.st, b);
  E_HIP(hipGetLastError());
  gm::ReduceArgs r{};
  r.partials = e.partials;
  r.G = e.G_sp;
  r.nv = S;
  r.out = e.small;
  r.stride = 1;
  r.do_sqrt = 1;
  hipLaunchKernelGGL(gm::k_reduce<0>, dim3(gm::reduce_grid(r.nv)), dim3(gm::kReduceBlock), 0, e.st, r);
  E_HIP(hipGetLastError());
  nrm.resize(S);
  E_HIP(hipMemcpyAsync(nrm.data(), e.small, S * sizeof(double), hipMemcpyDeviceToHost, e.st));
  E_HIP(hipStreamSynchronize(e.st));
  return GM_OK;
}

template <int S>
gm_status run(const gm_eigs_desc* d, Eigs& e, double* lambda, double* Qout, int32_t* passes_out,
              double* res_out) {
  const size_t N = (size_t)e.n;
  const int K = d->num_eigs;
  // upper bound of the spectrum: the caller's, or the Gershgorin bound
  double b = d->upper_bound;
  if (!(b > 0.0)) {
    b = 0.0;
    for (int i = 0; i < e.n; ++i) {
      double r = 0.0;
      for (int j = d->A_ptr[i]; j < d->A_ptr[i + 1]; ++j) r += std::fabs(d->A_values[j]);
      b = std::max(b, r);
    }
  }
  {
    gme::Fill f{e.X, N * S, (unsigned long long)d->seed};
    hipLaunchKernelGGL(gme::k_fill_uniform<0>, dim3(std::min<size_t>(8192, (N * S + 255) / 256)),
                       dim3(gm::kBlock), 0, e.st, f);
    E_HIP(hipGetLastError());
  }
  std::vector<double> G, W, w, V, theta(S, 0.0), nrm;
  // src -> dst with orthonormal columns spanning the same space, by scaled
  // SVQB (Stathopoulos & Wu, SIAM J. Sci. Comput. 23, 2002): G = Y^T Y,
  // D = diag(G)^(-1/2), D G D = V diag(g) V^T with g clamped below at
  // 1e-15 max(g), dst = src D V diag(g)^(-1/2).
  auto svqb = [&](const double* src, double* dst) -> gm_status {
    E_TRY(gram(e, src, src, G));
    std::vector<double> dsc(S);
    for (int i = 0; i < S; ++i) {
      const double gi = G[(size_t)i * S + i];
      if (!(gi > 0.0) || !std::isfinite(gi))
        return efail(GM_EBREAKDOWN, "block column %d has squared norm %g", i, gi);
      dsc[i] = 1.0 / std::sqrt(gi);
    }
    for (int i = 0; i < S; ++i)
      for (int j = 0; j < S; ++j) G[(size_t)i * S + j] *= dsc[i] * dsc[j];
    sym_eig(S, G, w, V);
    const double floor = 1e-15 * std::max(w[S - 1], 1e-300);
    W.assign((size_t)S * S, 0.0);
    for (int i = 0; i < S; ++i)
      for (int j = 0; j < S; ++j)
        W[(size_t)i * S + j] = dsc[i] * V[(size_t)i * S + j] / std::sqrt(std::max(w[j], floor));
    return combine<S>(e, src, W, dst);
  };
  // X <- orth(src) through tmp: two SVQB passes, the second restoring
  // orthogonality to rounding after the first lost digits to the block's
  // condition number.  src, tmp and the old X are the three blocks X, Y, Z.
  auto orthonormalise = [&](double* src, double* tmp) -> gm_status {
    double* dst = (e.X != src && e.X != tmp) ? e.X : (e.Y != src && e.Y != tmp) ? e.Y : e.Z;
    E_TRY(svqb(src, tmp));
    E_TRY(svqb(tmp, dst));
    e.X = dst;
    e.Y = src;
    e.Z = tmp;
    return GM_OK;
  };
  // Rayleigh-Ritz on the orthonormal X: AX = A X, H = X^T A X = V diag(theta)
  // V^T, X <- X V and AX <- AX V.
  auto rayleigh_ritz = [&]() -> gm_status {
    E_TRY(cheb_step<S>(e, e.X, nullptr, e.AX, 1.0, 0.0, 0.0, nullptr));
    E_TRY(gram(e, e.X, e.AX, G));
    sym_eig(S, G, theta, V);
    E_TRY(combine<S>(e, e.X, V, e.Y));
    E_TRY(combine<S>(e, e.AX, V, e.Z));
    std::swap(e.X, e.Y);
    std::swap(e.AX, e.Z);
    return GM_OK;
  };
  E_TRY(orthonormalise(e.X, e.Y));
  E_TRY(rayleigh_ritz());
  const int deg = std::max(1, d->degree);
  double worst = INFINITY;
  int pass = 0;
  for (;; ++pass) {
    // relative residuals ||A x_i - theta_i x_i|| / |theta_i| of the wanted pairs
    E_HIP(hipMemcpyAsync(e.shift, theta.data(), S * sizeof(double), hipMemcpyHostToDevice, e.st));
    E_TRY(cheb_step<S>(e, e.X, nullptr, e.Y, 1.0, 0.0, 0.0, e.shift));
    E_TRY(colnorms<S>(e, e.Y, nrm));
    worst = 0.0;
    for (int i = 0; i < K; ++i)
      worst = std::max(worst, nrm[i] / std::max(std::fabs(theta[i]), 1e-300));
    if (worst <= d->tol || pass >= d->max_passes) break;
    // filter: Y_1 = (sigma_1 / e)(A - c) X, Y_{i+1} = (2 sigma_{i+1} / e)(A -
    // c) Y_i - sigma_i sigma_{i+1} Y_{i-1}, the scaled Chebyshev recurrence on
    // [a, b] with a the block's largest Ritz value, normalised at theta_0
    const double a = theta[S - 1], a0 = theta[0];
    if (!(b > a)) return efail(GM_EINVAL, "upper bound %g is not above the Ritz value %g", b, a);
    const double ec = 0.5 * (b - a), c = 0.5 * (b + a);
    double sigma = ec / (a0 - c);
    const double tau = 2.0 / sigma;
    double *Yo = e.X, *Yc = e.Y, *Yn = e.Z;
    E_TRY(cheb_step<S>(e, Yo, nullptr, Yc, sigma / ec, 0.0, c, nullptr));
    for (int i = 2; i <= deg; ++i) {
      const double sn = 1.0 / (tau - sigma);
      E_TRY(cheb_step<S>(e, Yc, Yo, Yn, 2.0 * sn / ec, -sigma * sn, c, nullptr));
      double* t = Yo;
      Yo = Yc;
      Yc = Yn;
      Yn = t;
      sigma = sn;
    }
    E_TRY(orthonormalise(Yc, Yn));
    E_TRY(rayleigh_ritz());
  }
  // the K smallest Ritz pairs; Q column-major n x K
  if (Qout) {
    gme::Extract x{e.X, e.Y, S, K, e.n};
    hipLaunchKernelGGL(gme::k_extract<0>, dim3(std::min(4096, ediv(e.n, gm::kBlock)), K),
                       dim3(gm::kBlock), 0, e.st, x);
    E_HIP(hipGetLastError());
    E_HIP(hipMemcpyAsync(Qout, e.Y, N * K * sizeof(double), hipMemcpyDeviceToHost, e.st));
    E_HIP(hipStreamSynchronize(e.st));
  }
  for (int k = 0; k < K; ++k) lambda[k] = theta[k];
  if (passes_out) *passes_out = pass;
  if (res_out) *res_out = worst;
  if (!(worst <= d->tol))
    return efail(GM_EBREAKDOWN, "not converged in %d passes (worst relative residual %.3e)", pass,
                 worst);
  return GM_OK;
}

}  // namespace

extern "C" {

const char* gm_eigs_last_error(void) { return g_eerr.c_str(); }

gm_status gm_sym_eig(int32_t n, const double* A, double* w, double* V) {
  g_eerr.clear();
  if (n < 1 || n > 4096 || !A || !w) return efail(GM_EINVAL, "need 1 <= n <= 4096 and A, w");
  std::vector<double> a(A, A + (size_t)n * n), wv, vv;
  sym_eig(n, a, wv, vv);
  std::copy(wv.begin(), wv.end(), w);
  if (V) std::copy(vv.begin(), vv.end(), V);
  return GM_OK;
}

gm_status gm_eigs_smallest(const gm_eigs_desc* d, double* lambda, double* Q, int32_t* passes,
                           double* max_residual) {
  g_eerr.clear();
  if (!d || !lambda) return efail(GM_EINVAL, "null descriptor or output");
  if (d->n < 1 || d->nnz < 1 || !d->A_ptr || !d->A_ind || !d->A_values)
    return efail(GM_EINVAL, "bad matrix");
  if (d->sparse_format != GM_SPARSE_CSR && d->sparse_format != GM_SPARSE_CSC)
    return efail(GM_EINVAL, "unknown sparse format %d", d->sparse_format);
  const int s = d->block_size;
  if (s != 16 && s != 32 && s != 64) return efail(GM_EINVAL, "block_size must be 16, 32 or 64");
  if (d->num_eigs < 1 || d->num_eigs >= s || s > d->n)
    return efail(GM_EINVAL, "need 1 <= num_eigs < block_size <= n");
  if (!(d->tol > 0.0) || d->max_passes < 0 || d->degree < 1)
    return efail(GM_EINVAL, "bad tolerance, pass limit or filter degree");
  if (d->nnz > INT32_MAX) return efail(GM_EINVAL, "nnz beyond 32-bit row pointers");
  if (d->A_ptr[0] != 0 || d->A_ptr[d->n] != d->nnz) return efail(GM_EINVAL, "bad pointer array");
  for (int i = 0; i < d->n; ++i)
    if (d->A_ptr[i + 1] < d->A_ptr[i]) return efail(GM_EINVAL, "pointer array not monotone");
  for (long long j = 0; j < d->nnz; ++j)
    if (d->A_ind[j] < 0 || d->A_ind[j] >= d->n) return efail(GM_EINVAL, "index out of range");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || d->device < 0 || count <= d->device)
    return efail(GM_EDEVICE, "no HIP device %d", d->device);
  E_HIP(hipSetDevice(d->device));
  Eigs e;
  e.n = d->n;
  e.s = s;
  E_HIP(hipStreamCreateWithFlags(&e.st, hipStreamNonBlocking));
  const size_t N = (size_t)e.n;
  const int rows_per_pass = gm::kWaves * (64 / s);
  e.G_sp = std::max(1, std::min(4096, ediv(e.n, rows_per_pass)));
  e.chunk_sp = ediv(e.n, e.G_sp);
  e.G_sp = ediv(e.n, e.chunk_sp);
  e.G_df = std::max(1, std::min(256, ediv(e.n, 1024)));
  e.chunk_df = ediv(e.n, e.G_df);
  e.G_df = ediv(e.n, e.chunk_df);
  E_TRY(ealloc(&e.rowptr, N + 1));
  E_TRY(ealloc(&e.col, (size_t)d->nnz));
  E_TRY(ealloc(&e.val, (size_t)d->nnz));
  // a symmetric matrix has the same CSR and CSC arrays
  E_HIP(hipMemcpyAsync(e.rowptr, d->A_ptr, (N + 1) * sizeof(int), hipMemcpyHostToDevice, e.st));
  E_HIP(hipMemcpyAsync(e.col, d->A_ind, (size_t)d->nnz * sizeof(int), hipMemcpyHostToDevice, e.st));
  E_HIP(hipMemcpyAsync(e.val, d->A_values, (size_t)d->nnz * sizeof(double), hipMemcpyHostToDevice,
                       e.st));
  {
    std::vector<int> rb;
    gmb::bspmm_row_blocks(d->A_ptr, e.n, rb);
    e.G_rb = (int)rb.size() - 1;
    E_TRY(ealloc(&e.rowblk, rb.size()));
    E_HIP(hipMemcpyAsync(e.rowblk, rb.data(), rb.size() * sizeof(int), hipMemcpyHostToDevice, e.st));
    E_HIP(hipStreamSynchronize(e.st));   // rb is freed at the end of the scope
    gmb::bspmm_xs_blocks(d->A_ptr, e.n, rb);
    e.nrb_xs = (int)rb.size() / 4;
    e.G_xs = gmb::bspmm_xs_grid(d->device, e.nrb_xs);
    E_TRY(ealloc(&e.rowblk_xs, rb.size()));
    E_HIP(hipMemcpyAsync(e.rowblk_xs, rb.data(), rb.size() * sizeof(int), hipMemcpyHostToDevice, e.st));
    E_HIP(hipStreamSynchronize(e.st));
  }
  E_TRY(ealloc(&e.X, N * s));
  E_TRY(ealloc(&e.Y, N * s));
  E_TRY(ealloc(&e.Z, N * s));
  E_TRY(ealloc(&e.AX, N * s));
  E_TRY(ealloc(&e.partials, std::max<size_t>((size_t)e.G_sp * s, (size_t)e.G_df * s * s)));
  E_TRY(ealloc(&e.small, (size_t)s * s));
  E_TRY(ealloc(&e.wmat, (size_t)s * s));
  E_TRY(ealloc(&e.shift, (size_t)s));
  E_HIP(hipStreamSynchronize(e.st));
  switch (s) {
    case 16: return run<16>(d, e, lambda, Q, passes, max_residual);
    case 32: return run<32>(d, e, lambda, Q, passes, max_residual);
    default: return run<64>(d, e, lambda, Q, passes, max_residual);
  }
}

}  // extern "C"
