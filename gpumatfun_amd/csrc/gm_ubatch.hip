// gm_ubatch.hip -- independent problems of different sizes, one launch per
// stage ("uneven batch"; no reference counterpart).
//
// P problems, each with its own matrix A_p (n_p rows), its own two columns
// [u_p, dt g_p] and its own contour (weights, shifts, shifted gammas), share
// the subspace size m, the polynomial gamma, the number of nodes n_s and dt.
// The matrices are concatenated into one block-diagonal CSR and the vectors
// into one N = sum n_p row range, laid out exactly as the plain path lays out
// one problem (double2 rows [exp column, phi column], V[k][i]).  Rows are cut
// into blocks of at most kBlock rows and TILE nonzeros that never straddle two
// problems, so every kernel of a step covers all problems in one launch:
//   * SpMM: k_spmm_stream<TILE, MODE, SEG> (each block reads its problem's
//     alpha/beta, writes its partial at its row-block index); TILE (512, 1024
//     or 2048), 16-bit column offsets and early epilogue loads are picked per
//     batch by timing whole steps at create (GM_FLAG_AUTOTUNE), else 1024
//     nonzeros, 16-bit offsets where they fit, early loads;
//   * reductions: k_reduce in segmented form, one block per problem;
//   * finish, V*C, update: row-block kernels below (one thread per row);
//   * shifted solves and collapse: the plain kernels with blockIdx.y = problem.
// Small problems that are launch-bound on their own thus share each launch
// with the large ones.  No deflation in this form.
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
#include "gm_kernels.h"

namespace gmu {
using namespace gm;

constexpr int kTile = 1024;   // nonzeros per row block without the autotuner

// w <- w - alpha_k v_k with the problem's alpha; partial ||w||^2 per block.
// build_subspace.cu:63-76 (per problem).
// vk holds the unnormalised w_k (v_k = w_k / beta_k formed here, the bits of
// the normalised basis) and w_{k+1} goes to wout (null: over w).
struct UFinish {
  double2* w;
  const double2* vk;
  const double* aabb;
  size_t aabb_stride;
  int ld, k;
  const int* rowblk;
  const int* blk_prob;
  double* partials;
  double2* wout;
};

template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_ufinish(UFinish p) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int r = p.rowblk[b] + (int)threadIdx.x;
  const double* aabb = p.aabb + (size_t)p.blk_prob[b] * p.aabb_stride;
  const double2 al = d2(aabb[p.k], aabb[p.k + p.ld]);
  double2 nrm = d2(0.0, 0.0);
  if (r < p.rowblk[b + 1]) {
    double2 v = p.vk[r];
    v = d2(safe_div(v.x, aabb[p.k + 2 * p.ld]), safe_div(v.y, aabb[p.k + 3 * p.ld]));
    double2 w = p.w[r];
    w.x = fma(-al.x, v.x, w.x);
    w.y = fma(-al.y, v.y, w.y);
    (p.wout ? p.wout : p.w)[r] = w;
    nrm.x = w.x * w.x;
    nrm.y = w.y * w.y;
  }
  block_partial2_at(nrm, p.partials, b);
}

// Z_j = V C_j (j < ND) with the block's problem's C staged in LDS.
struct UVC {
  const double2* V;    // [m][N]
  const double2* C;    // [P][ND][m]
  int m;
  size_t ldv;          // N
  const int* rowblk;
  const int* blk_prob;
  double2* Z;          // [ND][N]; ND == 1: the output X
};

template <int ND>
__global__ __launch_bounds__(kBlock) void k_uvc(UVC p) {
  extern __shared__ double2 sC[];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const double2* C = p.C + (size_t)p.blk_prob[b] * ND * p.m;
  for (int t = threadIdx.x; t < ND * p.m; t += kBlock) sC[t] = C[t];
  __syncthreads();
  const int r = p.rowblk[b] + (int)threadIdx.x;
  if (r >= p.rowblk[b + 1]) return;
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
  for (int j = 0; j < ND; ++j) p.Z[(size_t)j * p.ldv + r] = acc[j];
}

// C[p][j][k] /= beta_k of problem p (both columns): V C over a basis that
// holds the unnormalised w_k.
template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_uscale_c(double2* C, const double* aabb,
                                                     size_t aabb_stride, int ld, int nd, int m,
                                                     int P) {
  const long long e = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (e >= (long long)P * nd * m) return;
  const int k = (int)(e % m), q = (int)(e / ((long long)nd * m));
  const double* ab = aabb + (size_t)q * aabb_stride;
  const double2 cv = C[e];
  C[e] = d2(safe_div(cv.x, ab[k + 2 * ld]), safe_div(cv.y, ab[k + 3 * ld]));
}

// Fisher-KPP update (update_U.cu:8-16, no deflation) or, when prepare, the
// create-time copy; U = Usol and the partial norms for beta_0.
struct UUpdate {
  const double2* X;
  const double* mass;
  double2* Usol;
  double2* U;
  double dt;
  int clamp, prepare;
  const int* rowblk;
  double* partials;
};

template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_uupdate(UUpdate p) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int r = p.rowblk[b] + (int)threadIdx.x;
  double2 nrm = d2(0.0, 0.0);
  if (r < p.rowblk[b + 1]) {
    double2 us;
    if (p.prepare) {
      us = p.Usol[r];
    } else {
      const double2 x = p.X[r];
      double u = x.x + x.y;
      if (p.clamp) u = u * (u >= 0.0 ? 1.0 : 0.0);
      const double ms = p.mass[r];
      const double ub = u / ms;
      const double g = (ms * ub) * (1.0 - ub);
      us = d2(u, g * p.dt);
      p.Usol[r] = us;
    }
    p.U[r] = us;
    nrm.x = us.x * us.x;
    nrm.y = us.y * us.y;
  }
  block_partial2_at(nrm, p.partials, b);
}

}  // namespace gmu

// ====================================================================== host
namespace {

thread_local std::string g_uerr;
std::mutex g_umu;
std::unordered_set<const void*> g_ureg;

gm_status ufail(gm_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_uerr = buf;
  return st;
}

#define U_HIP(call)                                                                      \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ufail(e_ == hipErrorOutOfMemory ? GM_ENOMEM : GM_EDEVICE, "%s: %s (%s:%d)", \
                   #call, hipGetErrorString(e_), __FILE__, __LINE__);                    \
  } while (0)
#define U_TRY(x)              \
  do {                        \
    gm_status s_ = (x);       \
    if (s_ != GM_OK) return s_; \
  } while (0)

}  // namespace

struct gm_ubatch {
  int dev = 0;
  hipStream_t stream = nullptr;
  int P = 0, N = 0, m = 0, nd = 0, ns = 0, G = 0;
  long long nnz = 0;
  double dt = 0.0;
  bool clamp = true;
  std::vector<double> gamma;
  std::vector<int> off;          // P + 1 row offsets
  int *rowptr = nullptr, *col = nullptr, *rowblk = nullptr, *blkj = nullptr,
      *blk_prob = nullptr, *prob_blk = nullptr;
  short* col16 = nullptr;        // 16-bit column offsets from each block's first row
  int tile = gmu::kTile;         // nonzeros per row block (SpMM LDS tile)
  bool use16 = false, pre = true;  // SpMM: 16-bit offsets; epilogue loads before gathers
  int solve_sb = 0;              // shifted solves: factors in global memory (0) or LDS
                                 // (> 0 systems a block), or a wavefront a system (-1)
  double *val = nullptr, *mass = nullptr, *aabb = nullptr, *wts1 = nullptr, *wts2 = nullptr,
         *shifts = nullptr, *gsr = nullptr, *gsi = nullptr, *yr = nullptr, *yi = nullptr,
         *partials = nullptr;
  double2 *U = nullptr, *Usol = nullptr, *X = nullptr, *W[2] = {nullptr, nullptr},
          *T[2] = {nullptr, nullptr}, *V = nullptr, *Z = nullptr, *C = nullptr, *wD = nullptr,
          *wDU = nullptr, *wDU2 = nullptr, *wX = nullptr;
  unsigned* status = nullptr;
  hipGraphExec_t graph = nullptr;
};

namespace {

bool ulive(const gm_ubatch* u) {
  std::lock_guard<std::mutex> lk(g_umu);
  return u && g_ureg.count(u);
}

template <typename T>
gm_status ualloc(T** p, size_t count) {
  *p = nullptr;
  U_HIP(hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T)));
  return GM_OK;
}

template <typename T>
gm_status uupload(T** p, const T* src, size_t count, hipStream_t st) {
  U_TRY(ualloc(p, count));
  if (count) U_HIP(hipMemcpyAsync(*p, src, count * sizeof(T), hipMemcpyHostToDevice, st));
  return GM_OK;
}

void ufree_all(gm_ubatch* u) {
  if (u->graph) (void)hipGraphExecDestroy(u->graph);
  void* bufs[] = {u->col16, u->rowptr, u->col, u->rowblk, u->blkj, u->blk_prob, u->prob_blk, u->val,
                  u->mass,
                  u->aabb, u->wts1, u->wts2, u->shifts, u->gsr, u->gsi, u->yr, u->yi,
                  u->partials, u->U, u->Usol, u->X, u->W[0], u->W[1], u->T[0], u->T[1], u->V,
                  u->Z, u->C, u->wD, u->wDU, u->wDU2, u->wX, u->status};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (u->stream) (void)hipStreamDestroy(u->stream);
}

size_t aabb_stride(const gm_ubatch* u) { return 4 * (size_t)(u->m + 1); }

// Per-problem sums of the [G][2] partials into out_q[0], out_q[stride].
gm_status ureduce(gm_ubatch* u, double* out, int stride, int do_sqrt, int flag_zero) {
  gm::ReduceArgs r{};
  r.partials = u->partials;
  r.G = u->G;
  r.nv = 2;
  r.out = out;
  r.stride = stride;
  r.do_sqrt = do_sqrt;
  r.flag_zero = flag_zero;
  r.status = u->status;
  r.seg = u->prob_blk;
  r.out_pstride = aabb_stride(u);
  hipLaunchKernelGGL(gm::k_reduce<0>, dim3(u->P), dim3(gm::kReduceBlock), 0, u->stream, r);
  U_HIP(hipGetLastError());
  return GM_OK;
}

template <int TILE, int MODE, bool C16, bool PRE>
void uspmm_launch(gm_ubatch* u, const gm::SpmmArgs& a) {
  hipLaunchKernelGGL((gm::k_spmm_stream<TILE, MODE, true, C16, PRE>), dim3(u->G),
                     dim3(gm::kBlock), 0, u->stream, a);
}

// The schedules the autotuner chooses from: every tile with 16-bit offsets or
// not and early epilogue loads, and the 1024-nonzero tile with late loads.
template <int MODE>
gm_status uspmm(gm_ubatch* u, const gm::SpmmArgs& a) {
  const bool c16 = a.col16 != nullptr;
  if (u->tile == 512) {
    if (c16) uspmm_launch<512, MODE, true, true>(u, a);
    else uspmm_launch<512, MODE, false, true>(u, a);
  } else if (u->tile == 2048) {
    if (c16) uspmm_launch<2048, MODE, true, true>(u, a);
    else uspmm_launch<2048, MODE, false, true>(u, a);
  } else if (u->pre) {
    if (c16) uspmm_launch<1024, MODE, true, true>(u, a);
    else uspmm_launch<1024, MODE, false, true>(u, a);
  } else {
    if (c16) uspmm_launch<1024, MODE, true, false>(u, a);
    else uspmm_launch<1024, MODE, false, false>(u, a);
  }
  U_HIP(hipGetLastError());
  return GM_OK;
}

gm::SpmmArgs uspmm_base(gm_ubatch* u) {
  gm::SpmmArgs a{};
  a.rowptr = u->rowptr;
  a.col = u->col;
  a.val = u->val;
  a.rowblk = u->rowblk;
  a.blkj = u->blkj;
  a.col16 = u->use16 ? u->col16 : nullptr;
  a.blk_prob = u->blk_prob;
  a.aabb_stride = (int)aabb_stride(u);
  a.nrows = u->N;
  a.aabb = u->aabb;
  a.ld = u->m + 1;
  a.partials = u->partials;
  a.status = u->status;
  a.g0 = u->gamma[0];
  return a;
}

gm_status enqueue_update(gm_ubatch* u, int prepare) {
  gmu::UUpdate p{};
  p.X = u->X;
  p.mass = u->mass;
  p.Usol = u->Usol;
  p.U = u->U;
  p.dt = u->dt;
  p.clamp = u->clamp ? 1 : 0;
  p.prepare = prepare;
  p.rowblk = u->rowblk;
  p.partials = u->partials;
  hipLaunchKernelGGL(gmu::k_uupdate<0>, dim3(u->G), dim3(gm::kBlock), 0, u->stream, p);
  U_HIP(hipGetLastError());
  return ureduce(u, u->aabb + 2 * (u->m + 1), u->m + 1, 1, 0);   // beta_0 per problem
}

template <int ND>
void launch_uvc(gm_ubatch* u, const gmu::UVC& a) {
  hipLaunchKernelGGL((gmu::k_uvc<ND>), dim3(u->G), dim3(gm::kBlock), ND * u->m * sizeof(double2),
                     u->stream, a);
}

// One time step of every problem: build_subspace, lin_sys_solves, apply_V,
// update_U of the plain path, each stage one launch over all problems.
gm_status enqueue_ustep(gm_ubatch* u) {
  const int d = u->nd - 1;
  const size_t N = (size_t)u->N;
  const int ld = u->m + 1;
  // The basis V holds the unnormalised w_k: the finish writes w_{k+1} straight
  // into V[k+1], the SpMM gathers it there and forms v_k = w_k / beta_k where
  // it needs it (the same bits), so v_k is never stored; V C takes C / beta.
  auto Vk = [&](int k) { return u->V + (size_t)k * N; };
  for (int k = 0; k < u->m; ++k) {
    const double2* wsrc = k == 0 ? u->U : Vk(k);
    double2* wnext = u->W[k & 1];
    gm::SpmmArgs a = uspmm_base(u);
    a.k = k;
    a.vk = k == 0 ? Vk(0) : nullptr;   // kFirst/kOnly: only the raw w_0 is stored
    a.vkm1 = k > 0 ? Vk(k - 1) : nullptr;
    a.x = wsrc;
    if (d == 0) {
      a.y = wnext;
      U_TRY(uspmm<gm::kOnly>(u, a));
    } else {
      a.y = u->T[0];
      a.gi = u->gamma[1];
      U_TRY(uspmm<gm::kFirst>(u, a));
      a.vk = Vk(k);
      for (int i = 2; i <= d; ++i) {
        a.x = u->T[(i - 2) & 1];
        a.y = u->T[(i - 1) & 1];
        a.gi = u->gamma[i];
        U_TRY(uspmm<gm::kMid>(u, a));
      }
      a.x = u->T[(d - 1) & 1];
      a.y = wnext;
      U_TRY(uspmm<gm::kLast>(u, a));
    }
    U_TRY(ureduce(u, u->aabb + k, ld, 0, 0));                          // alpha_k
    gmu::UFinish f{};
    f.w = wnext;
    f.vk = Vk(k);
    f.wout = k + 1 < u->m ? Vk(k + 1) : nullptr;
    f.aabb = u->aabb;
    f.aabb_stride = aabb_stride(u);
    f.ld = ld;
    f.k = k;
    f.rowblk = u->rowblk;
    f.blk_prob = u->blk_prob;
    f.partials = u->partials;
    hipLaunchKernelGGL(gmu::k_ufinish<0>, dim3(u->G), dim3(gm::kBlock), 0, u->stream, f);
    U_HIP(hipGetLastError());
    U_TRY(ureduce(u, u->aabb + (k + 1) + 2 * ld, ld, 1, 1));          // beta_{k+1}
  }
  // shifted solves, every problem's 2 n_s systems (blockIdx.y = problem)
  gm::SolveArgs s{};
  s.aabb = u->aabb;
  s.ld = ld;
  s.m = u->m;
  s.ns = u->ns;
  s.wts1 = u->wts1;
  s.wts2 = u->wts2;
  s.shifts = u->shifts;
  s.wD = u->wD;
  s.wDU = u->wDU;
  s.wDU2 = u->wDU2;
  s.wX = u->wX;
  s.yr = u->yr;
  s.yi = u->yi;
  s.status = u->status;
  s.aabb_ps = aabb_stride(u);
  s.wts_ps = 2 * (size_t)u->ns;
  s.shifts_ps = 4 * (size_t)u->m * u->ns;
  s.y_ps = 2 * (size_t)u->m * u->ns;
  s.work_ps = 2 * (size_t)u->m * u->ns;
  s.sb = u->solve_sb;
  if (s.sb < 0 && u->m <= 64)
    hipLaunchKernelGGL(gm::k_shifted_solves_pcr<1>, dim3(2 * u->ns, u->P), dim3(64), 0, u->stream, s);
  else if (s.sb < 0)
    hipLaunchKernelGGL(gm::k_shifted_solves_pcr<2>, dim3(2 * u->ns, u->P), dim3(64), 0, u->stream, s);
  else if (s.sb > 0)
    hipLaunchKernelGGL(gm::k_shifted_solves<1>, dim3((2 * u->ns + s.sb - 1) / s.sb, u->P),
                       dim3(64), 4 * (size_t)u->m * s.sb * sizeof(double2), u->stream, s);
  else
    hipLaunchKernelGGL(gm::k_shifted_solves<0>, dim3((2 * u->ns + 63) / 64, u->P), dim3(64), 0,
                       u->stream, s);
  U_HIP(hipGetLastError());
  // collapse over the nodes, then V*C and the Horner SpMMs (apply_V)
  gm::CollapseArgs c{};
  c.yr = u->yr;
  c.yi = u->yi;
  c.gsr = u->gsr;
  c.gsi = u->gsi;
  c.m = u->m;
  c.ns = u->ns;
  c.nd = u->nd;
  c.C = u->C;
  c.y_ps = 2 * (size_t)u->m * u->ns;
  c.g_ps = (size_t)u->nd * u->ns;
  c.c_ps = (size_t)u->nd * u->m;
  hipLaunchKernelGGL(gm::k_collapse<0>, dim3(std::max(1, std::min(64, (u->nd * u->m + 255) / 256)), u->P),
                     dim3(gm::kBlock), 0, u->stream, c);
  U_HIP(hipGetLastError());
  hipLaunchKernelGGL(gmu::k_uscale_c<0>,
                     dim3((int)(((long long)u->P * u->nd * u->m + gm::kBlock - 1) / gm::kBlock)),
                     dim3(gm::kBlock), 0, u->stream, u->C, u->aabb, aabb_stride(u), ld, u->nd, u->m,
                     u->P);
  U_HIP(hipGetLastError());
  gmu::UVC v{};
  v.V = u->V;
  v.C = u->C;
  v.m = u->m;
  v.ldv = N;
  v.rowblk = u->rowblk;
  v.blk_prob = u->blk_prob;
  v.Z = d == 0 ? u->X : u->Z;
  switch (u->nd) {
#define U_VC(K) \
  case K: launch_uvc<K>(u, v); break;
    U_VC(1) U_VC(2) U_VC(3) U_VC(4) U_VC(5) U_VC(6) U_VC(7) U_VC(8) U_VC(9)
    U_VC(10) U_VC(11) U_VC(12) U_VC(13) U_VC(14) U_VC(15) U_VC(16) U_VC(17)
#undef U_VC
    default: return ufail(GM_EINVAL, "gamma_size %d exceeds %d", u->nd, gm::kMaxDegree + 1);
  }
  U_HIP(hipGetLastError());
  const double2* src = u->Z;
  for (int j = 1; j <= d; ++j) {
    gm::SpmmArgs a = uspmm_base(u);
    double2* dst = (j == d) ? u->X : u->T[(j - 1) & 1];
    a.x = src;
    a.y = dst;
    a.z = u->Z + (size_t)j * N;
    a.accumulate = 0;
    U_TRY(uspmm<gm::kHorner>(u, a));
    src = dst;
  }
  return enqueue_update(u, 0);
}

// Rows of problem q cut into blocks of <= kBlock rows and <= kTile nonzeros
// (a longer row alone), appended to blk with global row numbers.
void problem_blocks(const std::vector<int>& rp, int r0, int r1, int tile, std::vector<int>& blk) {
  int r = r0;
  while (r < r1) {
    const int start = r;
    const long long base = rp[r];
    if (rp[r + 1] - base > tile) {
      ++r;
    } else {
      while (r < r1 && r - start < gm::kBlock && rp[r + 1] - base <= tile) ++r;
    }
    blk.push_back(r);
  }
}

// Cuts the concatenated rows into row blocks of at most `tile` nonzeros that never
// straddle two problems and uploads them (row-block starts, each block's problem,
// each problem's block range, and the 16-bit column offsets where they all fit).
gm_status set_blocks(gm_ubatch* u, const std::vector<int>& rp, const std::vector<int>& ci,
                     int tile) {
  std::vector<int> blk{0}, bprob, pblk{0};
  for (int q = 0; q < u->P; ++q) {
    const size_t before = blk.size();
    problem_blocks(rp, u->off[q], u->off[q + 1], tile, blk);
    for (size_t t = before; t < blk.size(); ++t) bprob.push_back(q);
    pblk.push_back((int)blk.size() - 1);
  }
  for (void* p : {(void*)u->rowblk, (void*)u->blkj, (void*)u->blk_prob, (void*)u->prob_blk,
                  (void*)u->col16})
    if (p) U_HIP(hipFree(p));
  u->rowblk = u->blkj = u->blk_prob = u->prob_blk = nullptr;
  u->col16 = nullptr;
  u->G = (int)bprob.size();
  u->tile = tile;
  hipStream_t st = u->stream;
  U_TRY(uupload(&u->rowblk, blk.data(), blk.size(), st));
  std::vector<int> bj(blk.size());
  for (size_t b = 0; b < blk.size(); ++b) bj[b] = rp[blk[b]];
  U_TRY(uupload(&u->blkj, bj.data(), bj.size(), st));
  U_TRY(uupload(&u->blk_prob, bprob.data(), bprob.size(), st));
  U_TRY(uupload(&u->prob_blk, pblk.data(), pblk.size(), st));
  std::vector<short> c16(ci.size());
  bool fits = true;
  for (int b = 0; b < u->G && fits; ++b)
    for (int j = rp[blk[b]]; j < rp[blk[b + 1]]; ++j) {
      const int off = ci[j] - blk[b];
      if (off < -32768 || off > 32767) {
        fits = false;
        break;
      }
      c16[j] = (short)off;
    }
  if (fits) U_TRY(uupload(&u->col16, c16.data(), c16.size(), st));
  U_HIP(hipStreamSynchronize(st));   // the host vectors go out of scope
  return GM_OK;
}


// Times two whole steps (after one untimed) under every candidate schedule and
// keeps the fastest; the state the steps advance is restored afterwards.
gm_status autotune(gm_ubatch* u, const std::vector<int>& rp, const std::vector<int>& ci) {
  const size_t N = (size_t)u->N;
  double2* keep = nullptr;
  U_TRY(ualloc(&keep, N));
  struct Cand { int tile; bool c16, pre; };
  std::vector<Cand> cands;
  for (int tile : {512, 1024, 2048})
    for (int c16 = 1; c16 >= 0; --c16) cands.push_back({tile, c16 != 0, true});
  for (int c16 = 1; c16 >= 0; --c16) cands.push_back({1024, c16 != 0, false});
  hipEvent_t e0 = nullptr, e1 = nullptr;
  gm_status st = GM_OK;
  float best = 1e30f;
  Cand pick{gmu::kTile, u->use16, true};
  if (hipMemcpyAsync(keep, u->Usol, N * sizeof(double2), hipMemcpyDeviceToDevice, u->stream) !=
          hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    st = ufail(GM_EDEVICE, "autotune set-up");
  for (const Cand& c : cands) {
    if (st != GM_OK) break;
    if (c.tile != u->tile) st = set_blocks(u, rp, ci, c.tile);
    if (st != GM_OK) break;
    if (c.c16 && !u->col16) continue;   // offsets do not fit this cut
    u->use16 = c.c16;
    u->pre = c.pre;
    st = enqueue_ustep(u);
    if (st == GM_OK && hipEventRecord(e0, u->stream) != hipSuccess) st = ufail(GM_EDEVICE, "event");
    for (int i = 0; i < 2 && st == GM_OK; ++i) st = enqueue_ustep(u);
    if (st == GM_OK && (hipEventRecord(e1, u->stream) != hipSuccess ||
                        hipEventSynchronize(e1) != hipSuccess))
      st = ufail(GM_EDEVICE, "event");
    float ms = 0.f;
    if (st == GM_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = ufail(GM_EDEVICE, "event");
    if (st == GM_OK && ms < best) {
      best = ms;
      pick = c;
    }
  }
  if (st == GM_OK && pick.tile != u->tile) st = set_blocks(u, rp, ci, pick.tile);
  if (st == GM_OK) {
    u->use16 = pick.c16 && u->col16 != nullptr;
    u->pre = pick.pre;
  }
  // then the shifted solves' forms under the kept schedule: one lane per
  // system with the factors in global memory or in LDS, or one wavefront per
  // system (cyclic reduction, m <= 128)
  std::vector<int> scand = {0};
  if (gm::solve_systems_per_block(u->m) > 0) scand.push_back(gm::solve_systems_per_block(u->m));
  if (u->m <= 128) scand.push_back(-1);
  if (st == GM_OK && scand.size() > 1) {
    float sbest = 1e30f;
    int spick = 0;
    for (int sb : scand) {
      u->solve_sb = sb;
      st = enqueue_ustep(u);
      if (st == GM_OK && hipEventRecord(e0, u->stream) != hipSuccess) st = ufail(GM_EDEVICE, "event");
      for (int i = 0; i < 2 && st == GM_OK; ++i) st = enqueue_ustep(u);
      float ms = 0.f;
      if (st == GM_OK && (hipEventRecord(e1, u->stream) != hipSuccess ||
                          hipEventSynchronize(e1) != hipSuccess ||
                          hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        st = ufail(GM_EDEVICE, "event");
      if (st != GM_OK) break;
      if (ms < sbest) {
        sbest = ms;
        spick = sb;
      }
    }
    u->solve_sb = spick;
  }
  if (st == GM_OK) {
    if (hipMemcpyAsync(u->Usol, keep, N * sizeof(double2), hipMemcpyDeviceToDevice, u->stream) !=
        hipSuccess)
      st = ufail(GM_EDEVICE, "restore state");
  }
  if (st == GM_OK) st = enqueue_update(u, 1);   // U, beta_0 and X from the restored state
  if (st == GM_OK && hipMemsetAsync(u->status, 0, sizeof(unsigned), u->stream) != hipSuccess)
    st = ufail(GM_EDEVICE, "status");
  if (st == GM_OK && hipStreamSynchronize(u->stream) != hipSuccess) st = ufail(GM_EDEVICE, "sync");
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(keep);
  return st;
}

gm_status create_impl(const gm_ubatch_desc* d, gm_ubatch* u) {
  u->dev = d->device;
  U_HIP(hipSetDevice(u->dev));
  U_HIP(hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking));
  hipStream_t st = u->stream;
  u->P = d->num_problems;
  u->m = d->subspace_size;
  u->nd = d->gamma_size;
  u->ns = d->contour_size;
  u->dt = d->dt;
  u->clamp = !(d->flags & GM_FLAG_NO_CLAMP);
  if ((d->flags & GM_FLAG_WAVE_SOLVES) && d->subspace_size <= 128) u->solve_sb = -1;
  u->gamma.assign(d->gamma, d->gamma + u->nd);
  u->off.assign(u->P + 1, 0);
  long long nnz = 0;
  for (int q = 0; q < u->P; ++q) {
    u->off[q + 1] = u->off[q] + d->n[q];
    nnz += d->nnz[q];
  }
  u->N = u->off[u->P];
  u->nnz = nnz;
  const size_t N = (size_t)u->N;
  // block-diagonal CSR, row blocks that never straddle two problems
  std::vector<int> rp(N + 1), ci((size_t)nnz);
  std::vector<double> vv((size_t)nnz), mass(N, 1.0);
  std::vector<double2> u2(N);
  long long e = 0;
  for (int q = 0; q < u->P; ++q) {
    const int o = u->off[q], n = d->n[q];
    for (int i = 0; i < n; ++i) {
      rp[o + i] = (int)(e + d->A_ptr[q][i]);
      if (d->mass && d->mass[q]) mass[o + i] = d->mass[q][i];
      u2[o + i] = make_double2(d->U[q][i], d->U[q][n + i] * u->dt);   // lanczos_impl.cu:79
    }
    for (long long j = 0; j < d->nnz[q]; ++j) {
      ci[e + j] = d->A_ind[q][j] + o;
      vv[e + j] = d->A_values[q][j];
    }
    e += d->nnz[q];
    rp[o + n] = (int)e;
  }
  U_TRY(uupload(&u->rowptr, rp.data(), N + 1, st));
  U_TRY(uupload(&u->col, ci.data(), (size_t)nnz, st));
  U_TRY(uupload(&u->val, vv.data(), (size_t)nnz, st));
  // the most row blocks any candidate tile cuts (partials are per row block)
  int maxG = 0;
  for (int tile : {512, 1024, 2048}) {
    U_TRY(set_blocks(u, rp, ci, tile));
    maxG = std::max(maxG, u->G);
  }
  U_TRY(set_blocks(u, rp, ci, gmu::kTile));
  u->use16 = u->col16 != nullptr;
  U_TRY(uupload(&u->mass, mass.data(), N, st));
  U_TRY(uupload(&u->Usol, u2.data(), N, st));
  const size_t P = (size_t)u->P, m = (size_t)u->m, ns = (size_t)u->ns, nd = (size_t)u->nd;
  U_TRY(uupload(&u->wts1, d->wts1, P * 2 * ns, st));
  U_TRY(uupload(&u->wts2, d->wts2, P * 2 * ns, st));
  U_TRY(uupload(&u->shifts, d->shifts, P * 4 * m * ns, st));
  U_TRY(uupload(&u->gsr, d->gamma_s_real, P * nd * ns, st));
  U_TRY(uupload(&u->gsi, d->gamma_s_imag, P * nd * ns, st));
  // host vectors must outlive the asynchronous uploads
  U_HIP(hipStreamSynchronize(st));
  U_TRY(ualloc(&u->U, N));
  U_TRY(ualloc(&u->X, N));
  U_TRY(ualloc(&u->W[0], N));
  U_TRY(ualloc(&u->W[1], N));
  if (u->nd > 1) {
    U_TRY(ualloc(&u->T[0], N));
    U_TRY(ualloc(&u->T[1], N));
    U_TRY(ualloc(&u->Z, N * nd));
  }
  U_TRY(ualloc(&u->V, N * m));
  U_TRY(ualloc(&u->C, P * nd * m));
  U_TRY(ualloc(&u->aabb, P * aabb_stride(u)));
  U_HIP(hipMemsetAsync(u->aabb, 0, P * aabb_stride(u) * sizeof(double), st));
  U_TRY(ualloc(&u->yr, P * 2 * m * ns));
  U_TRY(ualloc(&u->yi, P * 2 * m * ns));
  U_TRY(ualloc(&u->wD, P * 2 * m * ns));
  U_TRY(ualloc(&u->wDU, P * 2 * m * ns));
  U_TRY(ualloc(&u->wDU2, P * 2 * m * ns));
  U_TRY(ualloc(&u->wX, P * 2 * m * ns));
  U_TRY(ualloc(&u->partials, 2 * (size_t)maxG));
  U_TRY(ualloc(&u->status, 1));
  U_HIP(hipMemsetAsync(u->status, 0, sizeof(unsigned), st));
  U_TRY(enqueue_update(u, 1));   // U = Usol and beta_0 (lanczos_impl.cu:79-91, K = 0)
  U_HIP(hipStreamSynchronize(st));
  if (d->flags & GM_FLAG_AUTOTUNE) U_TRY(autotune(u, rp, ci));
  return GM_OK;
}

gm_status validate(const gm_ubatch_desc* d) {
  if (!d) return ufail(GM_EINVAL, "null descriptor");
  if (d->num_problems < 1 || !d->n || !d->nnz || !d->A_values || !d->A_ptr || !d->A_ind || !d->U)
    return ufail(GM_EINVAL, "bad problem count or null per-problem arrays");
  if (d->sparse_format != GM_SPARSE_CSR)
    return ufail(GM_EINVAL, "the uneven batch takes each A in CSR");
  if (d->subspace_size < 1 || d->gamma_size < 1 || d->gamma_size > gm::kMaxDegree + 1 || !d->gamma)
    return ufail(GM_EINVAL, "bad subspace size or gamma");
  if (d->contour_size < 1 || !d->gamma_s_real || !d->gamma_s_imag || !d->wts1 || !d->wts2 ||
      !d->shifts)
    return ufail(GM_EINVAL, "bad contour data");
  long long rows = 0, nnz = 0;
  for (int q = 0; q < d->num_problems; ++q) {
    const int n = d->n[q];
    if (n < 1 || !d->A_ptr[q] || !d->A_ind[q] || !d->A_values[q] || !d->U[q])
      return ufail(GM_EINVAL, "problem %d: bad size or null array", q);
    if (d->A_ptr[q][0] != 0 || d->A_ptr[q][n] != d->nnz[q])
      return ufail(GM_EINVAL, "problem %d: bad pointer array", q);
    for (int i = 0; i < n; ++i)
      if (d->A_ptr[q][i + 1] < d->A_ptr[q][i])
        return ufail(GM_EINVAL, "problem %d: pointer array not monotone", q);
    for (long long j = 0; j < d->nnz[q]; ++j)
      if (d->A_ind[q][j] < 0 || d->A_ind[q][j] >= n)
        return ufail(GM_EINVAL, "problem %d: column index out of range", q);
    rows += n;
    nnz += d->nnz[q];
  }
  if (rows > 0x7fffffffLL || nnz > 0x7fffffffLL)
    return ufail(GM_EINVAL, "total rows and nonzeros must fit int32");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= d->device || d->device < 0)
    return ufail(GM_EDEVICE, "no HIP device %d", d->device);
  return GM_OK;
}

}  // namespace

extern "C" {

const char* gm_ubatch_last_error(void) { return g_uerr.c_str(); }

gm_status gm_ubatch_create(const gm_ubatch_desc* desc, gm_ubatch** out) {
  g_uerr.clear();
  if (!out) return ufail(GM_EINVAL, "null output handle");
  *out = nullptr;
  U_TRY(validate(desc));
  gm_ubatch* u = new (std::nothrow) gm_ubatch();
  if (!u) return ufail(GM_ENOMEM, "host allocation failed");
  const gm_status s = create_impl(desc, u);
  if (s != GM_OK) {
    const std::string keep = g_uerr;
    ufree_all(u);
    delete u;
    g_uerr = keep;
    return s;
  }
  {
    std::lock_guard<std::mutex> lk(g_umu);
    g_ureg.insert(u);
  }
  *out = u;
  return GM_OK;
}

gm_status gm_ubatch_delete(gm_ubatch* u) {
  {
    std::lock_guard<std::mutex> lk(g_umu);
    if (!u || !g_ureg.count(u)) return ufail(GM_EBADHANDLE, "invalid uneven-batch handle");
    g_ureg.erase(u);
  }
  (void)hipSetDevice(u->dev);
  if (u->stream) (void)hipStreamSynchronize(u->stream);
  ufree_all(u);
  delete u;
  return GM_OK;
}

gm_status gm_ubatch_step(gm_ubatch* u, int32_t nsteps, int32_t use_graph) {
  if (!ulive(u)) return ufail(GM_EBADHANDLE, "invalid uneven-batch handle");
  U_HIP(hipSetDevice(u->dev));
  if (nsteps < 0) return ufail(GM_EINVAL, "nsteps < 0");
  if (!use_graph) {
    for (int i = 0; i < nsteps; ++i) U_TRY(enqueue_ustep(u));
    return GM_OK;
  }
  if (!u->graph) {
    hipGraph_t g = nullptr;
    U_HIP(hipStreamBeginCapture(u->stream, hipStreamCaptureModeThreadLocal));
    const gm_status s = enqueue_ustep(u);
    hipError_t e = hipStreamEndCapture(u->stream, &g);
    if (s != GM_OK) {
      if (g) (void)hipGraphDestroy(g);
      return s;
    }
    if (e != hipSuccess) return ufail(GM_EDEVICE, "capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&u->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      u->graph = nullptr;
      return ufail(GM_EDEVICE, "instantiate: %s", hipGetErrorString(e));
    }
  }
  for (int i = 0; i < nsteps; ++i) U_HIP(hipGraphLaunch(u->graph, u->stream));
  return GM_OK;
}

gm_status gm_ubatch_get_U(gm_ubatch* u, int32_t problem, double* dst) {
  if (!ulive(u)) return ufail(GM_EBADHANDLE, "invalid uneven-batch handle");
  if (!dst || problem < 0 || problem >= u->P) return ufail(GM_EINVAL, "bad problem index or null");
  U_HIP(hipSetDevice(u->dev));
  const int o = u->off[problem], n = u->off[problem + 1] - o;
  std::vector<double2> tmp(n);
  U_HIP(hipMemcpyAsync(tmp.data(), u->Usol + o, n * sizeof(double2), hipMemcpyDeviceToHost,
                       u->stream));
  U_HIP(hipStreamSynchronize(u->stream));
  for (int i = 0; i < n; ++i) {
    dst[i] = tmp[i].x;
    dst[n + i] = tmp[i].y;
  }
  return GM_OK;
}

gm_status gm_ubatch_synchronize(gm_ubatch* u) {
  if (!ulive(u)) return ufail(GM_EBADHANDLE, "invalid uneven-batch handle");
  U_HIP(hipSetDevice(u->dev));
  unsigned stv = 0;
  U_HIP(hipMemcpyAsync(&stv, u->status, sizeof(unsigned), hipMemcpyDeviceToHost, u->stream));
  U_HIP(hipStreamSynchronize(u->stream));
  if (stv & 6u) return ufail(GM_EBREAKDOWN, "Lanczos breakdown in the uneven batch (status %u)", stv);
  return GM_OK;
}

gm_status gm_ubatch_info(gm_ubatch* u, int32_t* num_problems, int64_t* total_rows,
                         int32_t* row_blocks) {
  if (!ulive(u)) return ufail(GM_EBADHANDLE, "invalid uneven-batch handle");
  if (num_problems) *num_problems = u->P;
  if (total_rows) *total_rows = u->N;
  if (row_blocks) *row_blocks = u->G;
  return GM_OK;
}

gm_status gm_ubatch_get_solve_form(gm_ubatch* u, int32_t* form) {
  if (!ulive(u)) return ufail(GM_EBADHANDLE, "invalid uneven-batch handle");
  if (form) *form = u->solve_sb < 0 ? 2 : (u->solve_sb > 0 ? 1 : 0);
  return GM_OK;
}

gm_status gm_ubatch_get_schedule(gm_ubatch* u, int32_t* tile, int32_t* index_bytes,
                                 int32_t* early_loads) {
  if (!ulive(u)) return ufail(GM_EBADHANDLE, "invalid uneven-batch handle");
  if (tile) *tile = u->tile;
  if (index_bytes) *index_bytes = u->use16 ? 2 : 4;
  if (early_loads) *early_loads = u->pre ? 1 : 0;
  return GM_OK;
}

}  // extern "C"
