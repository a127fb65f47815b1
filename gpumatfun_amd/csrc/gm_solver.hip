// gm_solver.hip -- solver state and the C ABI (include/gpumatfun.h).
//
// One gm_solver owns every device buffer of one problem (the reference's
// lanczos_solver, lanczos.h:16-86) and one HIP stream.  The stepping entry
// points only enqueue kernels; nothing inside a step waits on the host, so a
// whole step can be captured into a HIP graph (gm_step).
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>

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

using gm::kBlock;

namespace {

thread_local std::string g_last_error;
std::mutex g_registry_mu;
std::unordered_set<const void*> g_registry;

gm_status fail(gm_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return st;
}

#define GM_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(e_ == hipErrorOutOfMemory ? GM_ENOMEM : GM_EDEVICE, "%s: %s (%s:%d)", \
                  #call, hipGetErrorString(e_), __FILE__, __LINE__);                   \
  } while (0)

#define GM_TRY(expr)             \
  do {                           \
    gm_status st_ = (expr);      \
    if (st_ != GM_OK) return st_; \
  } while (0)

inline int cdiv_i(long long a, long long b) { return (int)((a + b - 1) / b); }

// LDS tile sizes (nonzeros) of the row-block SpMM schedule.
constexpr int kNumTiles = 3;
constexpr int kTiles[kNumTiles] = {512, 1024, 2048};

}  // namespace

struct gm_solver {
  int dev = 0;
  hipStream_t stream = nullptr;
  int n = 0;
  long long nnz = 0;
  int m = 0, nd = 0, ns = 0, K = 0;
  double dt = 0, aode = 0, bode = 0, code = 0, vrest = 0, vamp = 0, vth = 0, vpeak = 0,
         c1ion = 0, c2ion = 0;
  bool clamp = true;
  // Mitchell-Schaeffer parameters (gm_set_ms_params); tau_in, tau_out, tau_open,
  // tau_close, v_gate
  double ms[5] = {0.3, 6.0, 120.0, 150.0, 0.13};
  std::vector<double> gamma;
  // schedules
  // SpMM schedule: tile > 0 selects the row-block kernel with that many
  // nonzeros per LDS tile (row blocks rowblk[tile_idx]); tile == 0 the
  // row-group kernel with `lanes` lanes per row.
  int lanes = 8, tile = 0, tile_idx = -1;
  bool use16 = false;   // row blocks: 16-bit column offsets (col16[tile_idx])
  bool pre = false;     // row blocks: epilogue operands loaded before the gathers
  int G_sp = 1, chunk_sp = 1, G_el = 1, chunk_el = 1, G_df = 1, chunk_df = 1;
  int* rowblk[kNumTiles] = {};
  int* blkj[kNumTiles] = {};    // rowptr at each row-block bound
  int nrowblk[kNumTiles] = {};
  // 16-bit column offsets from each nonzero's row-block start, per tile size
  // (null when some offset does not fit, or GM_FLAG_INDEX32)
  short* col16[kNumTiles] = {};
  // device buffers
  int* rowptr = nullptr;
  int* col = nullptr;
  double* val = nullptr;
  double2 *U = nullptr, *Usol = nullptr, *X = nullptr, *upd = nullptr;
  double2 *W[2] = {nullptr, nullptr}, *T[2] = {nullptr, nullptr};
  double2* V = nullptr;
  double2* Z = nullptr;
  double2* C = nullptr;
  double *Q = nullptr, *el = nullptr, *pl = nullptr, *mass = nullptr, *ode = nullptr;
  int solve_sb = 0;   // shifted solves: 0 = factors in global memory, > 0 = systems per
                      // block with the factors in LDS, -1 = one wavefront per system
                      // (cyclic reduction, m <= 128); chosen by the autotuner
  double *aabb = nullptr, *Yr = nullptr, *Yi = nullptr, *wts1 = nullptr, *wts2 = nullptr,
         *shifts = nullptr, *gsr = nullptr, *gsi = nullptr, *Zdefl = nullptr;
  double2 *wD = nullptr, *wDU = nullptr, *wDU2 = nullptr, *wX = nullptr;
  double* partials = nullptr;
  double* partials2 = nullptr;   // finish norm partials when reductions are fused
  unsigned* status = nullptr;
  double* stage = nullptr;  // n x 2 column-major staging for the getters
  // host-side state machine
  bool built = false, solved = false, x_zero = true;
  hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
};

namespace {

bool live(const gm_solver* s) {
  std::lock_guard<std::mutex> lk(g_registry_mu);
  return s != nullptr && g_registry.count(s) != 0;
}

#define GM_CHECK_HANDLE(s)                                                              \
  do {                                                                                  \
    if (!live(s)) return fail(GM_EBADHANDLE, "invalid, deleted or null solver handle"); \
    GM_HIP(hipSetDevice((s)->dev));                                                     \
  } while (0)

template <typename T>
gm_status dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  GM_HIP(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
  return GM_OK;
}

template <typename T>
gm_status upload(T** p, const T* src, size_t count, hipStream_t st) {
  GM_TRY(dalloc(p, count));
  if (count && src) GM_HIP(hipMemcpyAsync(*p, src, count * sizeof(T), hipMemcpyHostToDevice, st));
  return GM_OK;
}

void free_all(gm_solver* s) {
  for (auto& g : s->graph)
    if (g) hipGraphExecDestroy(g);
  for (int* rb : s->rowblk)
    if (rb) hipFree(rb);
  for (int* bj : s->blkj)
    if (bj) hipFree(bj);
  for (short* c : s->col16)
    if (c) hipFree(c);
  void* bufs[] = {s->rowptr, s->col, s->val, s->U, s->Usol, s->X, s->upd, s->W[0], s->W[1],
                  s->T[0], s->T[1], s->V, s->Z, s->C, s->Q, s->el, s->pl, s->mass, s->ode,
                  s->aabb, s->Yr, s->Yi, s->wts1, s->wts2, s->shifts, s->gsr, s->gsi,
                  s->Zdefl, s->wD, s->wDU, s->wDU2, s->wX, s->partials, s->partials2,
                  s->status, s->stage};
  for (void* b : bufs)
    if (b) hipFree(b);
  if (s->stream) hipStreamDestroy(s->stream);
}

// ----------------------------------------------------------------- launches
using gm::SpmmArgs;

template <int MODE>
gm_status launch_spmm_mode(gm_solver* s, const SpmmArgs& a) {
  dim3 grid(s->G_sp), block(kBlock);
  if (s->tile > 0) {
    const int v = (a.col16 != nullptr ? 1 : 0) + (s->pre ? 2 : 0);
#define GM_STREAM(T)                                                                                  \
  case T:                                                                                             \
    switch (v) {                                                                                      \
      case 0: hipLaunchKernelGGL((gm::k_spmm_stream<T, MODE, false, false, false>), grid, block, 0, s->stream, a); break; \
      case 1: hipLaunchKernelGGL((gm::k_spmm_stream<T, MODE, false, true, false>), grid, block, 0, s->stream, a); break; \
      case 2: hipLaunchKernelGGL((gm::k_spmm_stream<T, MODE, false, false, true>), grid, block, 0, s->stream, a); break; \
      default: hipLaunchKernelGGL((gm::k_spmm_stream<T, MODE, false, true, true>), grid, block, 0, s->stream, a); break; \
    }                                                                                                 \
    break;
    switch (s->tile) {
      GM_STREAM(512)
      GM_STREAM(1024)
      default:
      GM_STREAM(2048)
    }
#undef GM_STREAM
    GM_HIP(hipGetLastError());
    return GM_OK;
  }
  switch (s->lanes) {
    case 1: hipLaunchKernelGGL((gm::k_spmm<1, MODE>), grid, block, 0, s->stream, a); break;
    case 2: hipLaunchKernelGGL((gm::k_spmm<2, MODE>), grid, block, 0, s->stream, a); break;
    case 4: hipLaunchKernelGGL((gm::k_spmm<4, MODE>), grid, block, 0, s->stream, a); break;
    case 8: hipLaunchKernelGGL((gm::k_spmm<8, MODE>), grid, block, 0, s->stream, a); break;
    case 16: hipLaunchKernelGGL((gm::k_spmm<16, MODE>), grid, block, 0, s->stream, a); break;
    case 32: hipLaunchKernelGGL((gm::k_spmm<32, MODE>), grid, block, 0, s->stream, a); break;
    default: hipLaunchKernelGGL((gm::k_spmm<64, MODE>), grid, block, 0, s->stream, a); break;
  }
  GM_HIP(hipGetLastError());
  return GM_OK;
}

// Sum [G][nv] block partials into out[v * stride] (k_reduce, a block per 1024 values).
gm_status launch_reduce(gm_solver* s, int G, int nv, double* out, int stride, int do_sqrt,
                        int flag_zero, const double* part = nullptr) {
  gm::ReduceArgs r{};
  r.partials = part ? part : s->partials;
  r.G = G;
  r.nv = nv;
  r.out = out;
  r.stride = stride;
  r.do_sqrt = do_sqrt;
  r.flag_zero = flag_zero;
  r.status = s->status;
  hipLaunchKernelGGL(gm::k_reduce<0>, dim3(gm::reduce_grid(r.nv)), dim3(gm::kReduceBlock), 0, s->stream, r);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

// Measurement kernels of gm_profile_barrier: `iters` cooperative grid barriers,
// each after every thread stores one value; and the same store as a kernel of
// its own (the kernel boundary the barrier would replace).
__global__ __launch_bounds__(kBlock) void k_grid_barriers(int iters, double* buf) {
  cooperative_groups::grid_group g = cooperative_groups::this_grid();
  const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  for (int it = 0; it < iters; ++it) {
    buf[i] = (double)it;
    g.sync();
  }
}

__global__ __launch_bounds__(kBlock) void k_touch(double* buf) {
  buf[(size_t)blockIdx.x * kBlock + threadIdx.x] += 1.0;
}

SpmmArgs spmm_base(gm_solver* s) {
  SpmmArgs a{};
  a.rowptr = s->rowptr;
  a.col = s->col;
  a.val = s->val;
  a.rowblk = s->tile > 0 ? s->rowblk[s->tile_idx] : nullptr;
  a.blkj = s->tile > 0 ? s->blkj[s->tile_idx] : nullptr;
  a.col16 = (s->tile > 0 && s->use16) ? s->col16[s->tile_idx] : nullptr;
  a.nrows = s->n;
  a.chunk = s->chunk_sp;
  a.aabb = s->aabb;
  a.ld = s->m + 1;
  a.partials = s->partials;
  a.status = s->status;
  return a;
}

// Row-group schedule with `lanes` lanes per row.
void set_schedule(gm_solver* s, int lanes) {
  s->lanes = lanes;
  s->tile = 0;
  s->tile_idx = -1;
  s->use16 = false;
  s->pre = false;
  const int groups = kBlock / lanes;
  s->G_sp = std::max(1, std::min(4096, cdiv_i(s->n, groups)));
  s->chunk_sp = cdiv_i(s->n, s->G_sp);
}

// Row-block schedule with kTiles[idx] nonzeros per LDS tile, with 16-bit
// column offsets when use16 and they fit, epilogue operands loaded early
// when pre.
void set_tile_schedule(gm_solver* s, int idx, bool use16 = false, bool pre = false) {
  s->tile = kTiles[idx];
  s->tile_idx = idx;
  s->use16 = use16 && s->col16[idx] != nullptr;
  s->pre = pre;
  s->G_sp = s->nrowblk[idx];
  s->chunk_sp = 0;
}

// Cut rows 0..n-1 into consecutive blocks of at most kBlock rows and at most
// `tile` nonzeros; a row with more than `tile` nonzeros is a block alone.
std::vector<int> row_blocks(const std::vector<int>& rp, int n, int tile) {
  std::vector<int> blk{0};
  if (n == 0) blk.push_back(0);  // one empty block
  int r = 0;
  while (r < n) {
    const int start = r;
    const long long base = rp[r];
    if (rp[r + 1] - base > tile) {
      ++r;
    } else {
      while (r < n && r - start < kBlock && rp[r + 1] - base <= tile) ++r;
    }
    blk.push_back(r);
  }
  return blk;
}

// build_subspace.cu:5-91.  Iteration k reads w_k (U for k = 0) and writes
// w_{k+1} into W[k & 1]; beta_k lives in aabb and is applied on the fly.
// When both grids are small (at most kFuseG blocks), alpha_k and beta_{k+1}
// are summed inside the kernels that consume them (every block forms the
// same fixed-order sum of the producer's partials) instead of by k_reduce
// launches: two launches fewer per Lanczos step on launch-bound sizes.
constexpr int kFuseG = 1024;

gm_status enqueue_build(gm_solver* s) {
  const int d = s->nd - 1;
  const size_t N = (size_t)s->n;
  const int ld = s->m + 1;
  const bool fuse = s->G_sp <= kFuseG && s->G_el <= kFuseG;
  for (int k = 0; k < s->m; ++k) {
    const double2* wsrc = k == 0 ? s->U : s->W[(k - 1) & 1];
    double2* wnext = s->W[k & 1];
    double2* vk = s->V + (size_t)k * N;
    const double2* vkm1 = k > 0 ? s->V + (size_t)(k - 1) * N : nullptr;
    SpmmArgs a = spmm_base(s);
    a.k = k;
    a.vk = vk;
    a.vkm1 = vkm1;
    a.g0 = s->gamma[0];
    if (fuse && k > 0) {   // beta_k from the previous finish's partials
      a.bpart = s->partials2;
      a.bG = s->G_el;
      a.beta_out = s->aabb + k + 2 * ld;
    }
    if (d == 0) {
      a.x = wsrc;
      a.y = wnext;
      GM_TRY(launch_spmm_mode<gm::kOnly>(s, a));
    } else {
      a.x = wsrc;
      a.y = s->T[0];
      a.gi = s->gamma[1];
      GM_TRY(launch_spmm_mode<gm::kFirst>(s, a));
      a.bpart = nullptr;
      for (int i = 2; i <= d; ++i) {
        a.x = s->T[(i - 2) & 1];
        a.y = s->T[(i - 1) & 1];
        a.gi = s->gamma[i];
        GM_TRY(launch_spmm_mode<gm::kMid>(s, a));
      }
      a.x = s->T[(d - 1) & 1];
      a.y = wnext;
      GM_TRY(launch_spmm_mode<gm::kLast>(s, a));
    }
    if (!fuse) GM_TRY(launch_reduce(s, s->G_sp, 2, s->aabb + k, ld, 0, 0));   // alpha_k
    gm::FinishArgs f{};
    f.w = wnext;
    f.vk = vk;
    f.aabb = s->aabb;
    f.ld = s->m + 1;
    f.k = k;
    f.nrows = s->n;
    f.chunk = s->chunk_el;
    f.partials = fuse ? s->partials2 : s->partials;
    f.status = s->status;
    if (fuse) {
      f.apart = s->partials;
      f.aG = s->G_sp;
    }
    hipLaunchKernelGGL(gm::k_lanczos_finish<0>, dim3(s->G_el), dim3(kBlock), 0, s->stream, f);
    GM_HIP(hipGetLastError());
    // beta_{k+1} = ||w||, flagged when zero or non-finite (fused: formed by
    // the next iteration's first SpMM, except after the last iteration)
    if (!fuse || k == s->m - 1)
      GM_TRY(launch_reduce(s, s->G_el, 2, s->aabb + (k + 1) + 2 * ld, ld, 1, 1,
                           fuse ? s->partials2 : nullptr));
  }
  return GM_OK;
}

gm_status enqueue_solves(gm_solver* s) {
  gm::SolveArgs a{};
  a.aabb = s->aabb;
  a.ld = s->m + 1;
  a.m = s->m;
  a.ns = s->ns;
  a.wts1 = s->wts1;
  a.wts2 = s->wts2;
  a.shifts = s->shifts;
  a.wD = s->wD;
  a.wDU = s->wDU;
  a.wDU2 = s->wDU2;
  a.wX = s->wX;
  a.yr = s->Yr;
  a.yi = s->Yi;
  a.status = s->status;
  a.sb = s->solve_sb;
  if (a.sb < 0 && s->m <= 64)
    hipLaunchKernelGGL(gm::k_shifted_solves_pcr<1>, dim3(2 * s->ns), dim3(64), 0, s->stream, a);
  else if (a.sb < 0)
    hipLaunchKernelGGL(gm::k_shifted_solves_pcr<2>, dim3(2 * s->ns), dim3(64), 0, s->stream, a);
  else if (a.sb > 0)
    hipLaunchKernelGGL(gm::k_shifted_solves<1>, dim3(cdiv_i(2 * s->ns, a.sb)), dim3(64),
                       4 * (size_t)s->m * a.sb * sizeof(double2), s->stream, a);
  else
    hipLaunchKernelGGL(gm::k_shifted_solves<0>, dim3(cdiv_i(2 * s->ns, 64)), dim3(64), 0,
                       s->stream, a);
  GM_HIP(hipGetLastError());
  return GM_OK;
}

template <int ND>
void launch_vc(gm_solver* s, const gm::VCArgs& a) {
  hipLaunchKernelGGL((gm::k_vc<ND>), dim3(cdiv_i(s->n, kBlock)), dim3(kBlock),
                     ND * s->m * sizeof(double2), s->stream, a);
}

gm_status enqueue_apply(gm_solver* s) {
  const int d = s->nd - 1;
  const size_t N = (size_t)s->n;
  gm::CollapseArgs c{};
  c.yr = s->Yr;
  c.yi = s->Yi;
  c.gsr = s->gsr;
  c.gsi = s->gsi;
  c.m = s->m;
  c.ns = s->ns;
  c.nd = s->nd;
  c.C = s->C;
  hipLaunchKernelGGL(gm::k_collapse<0>, dim3(std::min(64, cdiv_i(s->nd * s->m, kBlock))), dim3(kBlock),
                     0, s->stream, c);
  GM_HIP(hipGetLastError());
  gm::VCArgs v{};
  v.V = s->V;
  v.C = s->C;
  v.m = s->m;
  v.nd = s->nd;
  v.nrows = s->n;
  v.ldv = N;
  v.Z = d == 0 ? s->X : s->Z;
  v.accumulate = (d == 0 && !s->x_zero) ? 1 : 0;
  switch (s->nd) {
#define GM_VC(ND) \
  case ND: launch_vc<ND>(s, v); break;
    GM_VC(1) GM_VC(2) GM_VC(3) GM_VC(4) GM_VC(5) GM_VC(6) GM_VC(7) GM_VC(8) GM_VC(9)
    GM_VC(10) GM_VC(11) GM_VC(12) GM_VC(13) GM_VC(14) GM_VC(15) GM_VC(16) GM_VC(17)
#undef GM_VC
    default: return fail(GM_EINVAL, "gamma_size %d exceeds %d", s->nd, gm::kMaxDegree + 1);
  }
  GM_HIP(hipGetLastError());
  // Horner over the collapsed coefficients: x = Z_0; x = A x + Z_j.
  const double2* src = s->Z;
  for (int j = 1; j <= d; ++j) {
    SpmmArgs a = spmm_base(s);
    double2* dst = (j == d) ? s->X : s->T[(j - 1) & 1];
    a.x = src;
    a.y = dst;
    a.z = s->Z + (size_t)j * N;
    a.accumulate = (j == d && !s->x_zero) ? 1 : 0;
    GM_TRY(launch_spmm_mode<gm::kHorner>(s, a));
    src = dst;
  }
  s->x_zero = false;
  return GM_OK;
}

// update_U.cu:5-32 (mode kUpdFisher), update_U_monodomain.cu:5-43 (kUpdMono),
// or the constructor's deflation of a given Usol (kUpdPrepare,
// lanczos_impl.cu:79-89).
gm_status enqueue_update(gm_solver* s, int mode) {
  gm::UpdateArgs a{};
  a.X = s->X;
  a.upd = s->upd;
  a.mass = s->mass;
  a.ode = s->ode;
  a.dt = s->dt;
  a.aode_dt = s->dt * s->aode;
  a.bode_dt = s->dt * s->bode;
  a.code = s->code;
  a.vrest = s->vrest;
  a.vpeak = s->vpeak;
  a.cion = s->c1ion / s->vamp / s->vamp;
  a.ms_tau_in = s->ms[0];
  a.ms_tau_out = s->ms[1];
  a.ms_tau_open = s->ms[2];
  a.ms_tau_close = s->ms[3];
  a.ms_v_gate = s->ms[4];
  a.clamp = s->clamp ? 1 : 0;
  a.x_zero = s->x_zero ? 1 : 0;
  a.Usol = s->Usol;
  a.U = s->U;
  a.Q = s->Q;
  a.K = s->K;
  a.nrows = s->n;
  a.partials = s->partials;
  a.beta0 = s->aabb + 2 * (s->m + 1);
  a.ld = s->m + 1;
  a.Zout = s->Zdefl;
  a.status = s->status;
  const int G = s->K == 0 ? s->G_el : s->G_df;
  a.chunk = s->K == 0 ? s->chunk_el : s->chunk_df;
  const bool wide = s->K > gm::kDeflTile && s->K <= gm::kDeflTileWide;
#define GM_UPD1(M)                                                                        \
  if (wide)                                                                               \
    hipLaunchKernelGGL((gm::k_update1<M, gm::kDeflTileWide>), dim3(G), dim3(kBlock), 0,     \
                       s->stream, a);                                                     \
  else                                                                                    \
    hipLaunchKernelGGL((gm::k_update1<M>), dim3(G), dim3(kBlock), 0, s->stream, a);
  switch (mode) {
    case gm::kUpdFisher:
      GM_UPD1(gm::kUpdFisher)
      break;
    case gm::kUpdMono:
      GM_UPD1(gm::kUpdMono)
      break;
    case gm::kUpdMS:
      GM_UPD1(gm::kUpdMS)
      break;
    default:
      GM_UPD1(gm::kUpdPrepare)
      break;
  }
#undef GM_UPD1
  GM_HIP(hipGetLastError());
  double* beta0 = s->aabb + 2 * (s->m + 1);
  if (s->K == 0) {
    GM_TRY(launch_reduce(s, G, 2, beta0, s->m + 1, 1, 0));              // beta_0
  } else {
    GM_TRY(launch_reduce(s, G, 2 * s->K, s->Zdefl, 1, 0, 0));           // Z = Q^T Usol
  }
  if (s->K > 0) {
    gm::Update2Args b{};
    b.Usol = s->Usol;
    b.U = s->U;
    b.upd = s->upd;
    b.Q = s->Q;
    b.Z = s->Zdefl;
    b.el = s->el;
    b.pl = s->pl;
    b.K = s->K;
    b.nrows = s->n;
    b.chunk = s->chunk_el;
    b.ld = s->m + 1;
    b.partials = s->partials;
    b.beta0 = s->aabb + 2 * (s->m + 1);
    b.status = s->status;
    hipLaunchKernelGGL(gm::k_update2<0>, dim3(s->G_el), dim3(kBlock), 4 * s->K * sizeof(double),
                       s->stream, b);
    GM_HIP(hipGetLastError());
    GM_TRY(launch_reduce(s, s->G_el, 2, beta0, s->m + 1, 1, 0));        // beta_0
  } else {
    // update = 0 when nothing is deflated (Q is N x 0)
  }
  if (mode != gm::kUpdPrepare) s->x_zero = true;
  return GM_OK;
}

gm_status enqueue_step(gm_solver* s, int mono) {
  GM_TRY(enqueue_build(s));
  GM_TRY(enqueue_solves(s));
  GM_TRY(enqueue_apply(s));
  GM_TRY(enqueue_update(s, mono == 2 ? gm::kUpdMS : (mono ? gm::kUpdMono : gm::kUpdFisher)));
  return GM_OK;
}

gm_status check_status(gm_solver* s) {
  unsigned st = 0;
  GM_HIP(hipMemcpyAsync(&st, s->status, sizeof(unsigned), hipMemcpyDeviceToHost, s->stream));
  GM_HIP(hipStreamSynchronize(s->stream));
  if (st & 6u)
    return fail(GM_EBREAKDOWN, "Lanczos breakdown: %s%s",
                (st & 2u) ? "singular shifted tridiagonal system; " : "",
                (st & 4u) ? "non-finite beta (NaN/Inf in the input?)" : "");
  return GM_OK;
}

gm_status fetch_pair(gm_solver* s, const double2* src, double* dst) {
  if (!dst) return fail(GM_EINVAL, "null destination pointer");
  hipLaunchKernelGGL(gm::k_deinterleave<0>, dim3(cdiv_i(s->n, kBlock)), dim3(kBlock), 0, s->stream,
                     src, s->stage, s->n);
  GM_HIP(hipGetLastError());
  GM_HIP(hipMemcpyAsync(dst, s->stage, 2 * (size_t)s->n * sizeof(double), hipMemcpyDeviceToHost,
                        s->stream));
  GM_HIP(hipStreamSynchronize(s->stream));
  return GM_OK;
}

// ------------------------------------------------------------ validation
gm_status validate(const gm_create_desc* d) {
  const char* P = "MATLAB:lanczos_create:";
  if (!d) return fail(GM_EINVAL, "%sinvalidNumInputs: null create descriptor", P);
  if (d->n <= 0 || !d->U) return fail(GM_EINVAL, "%sbadU: Invalid U input argument.", P);
  if (!d->A_values && d->nnz > 0)
    return fail(GM_EINVAL, "%sinputNotSparse: Invalid sparse matrix input argument.", P);
  if (!d->A_ptr || (d->nnz > 0 && !d->A_ind) || d->nnz < 0 || d->nnz > 0x7fffffffLL)
    return fail(GM_EINVAL, "%sinputNotSparse: Invalid sparse matrix input argument.", P);
  if (d->sparse_format != GM_SPARSE_CSC && d->sparse_format != GM_SPARSE_CSR)
    return fail(GM_EINVAL, "%sinputNotSparse: unknown sparse format %d", P, d->sparse_format);
  if (d->A_ptr[0] != 0 || d->A_ptr[d->n] != d->nnz)
    return fail(GM_EINVAL, "%sinputNotSparse: pointer array must run from 0 to nnz", P);
  for (int i = 0; i < d->n; ++i)
    if (d->A_ptr[i + 1] < d->A_ptr[i])
      return fail(GM_EINVAL, "%sinputNotSparse: pointer array not monotone at %d", P, i);
  for (long long j = 0; j < d->nnz; ++j)
    if (d->A_ind[j] < 0 || d->A_ind[j] >= d->n)
      return fail(GM_EINVAL, "%sinputNotSquare: index %d out of range at %lld (matrix must be %d x %d)",
                  P, d->A_ind[j], j, d->n, d->n);
  if (d->num_eigs < 0 || (d->num_eigs > 0 && !d->Q))
    return fail(GM_EINVAL, "%sbadQ: Invalid Q input argument.", P);
  if (d->num_eigs > 0 && !d->lambda)
    return fail(GM_EINVAL, "%sbad_lambda: Invalid lambda input argument.", P);
  if (d->subspace_size < 1) return fail(GM_EINVAL, "%sbad_N: Invalid N input argument.", P);
  if (d->gamma_size < 1 || !d->gamma)
    return fail(GM_EINVAL, "%sbad_gamma: Invalid gamma input argument.", P);
  if (d->gamma_size > gm::kMaxDegree + 1)
    return fail(GM_EINVAL, "%sbad_gamma: polynomial degree %d exceeds %d", P, d->gamma_size - 1,
                gm::kMaxDegree);
  if (d->contour_size < 1 || !d->gamma_s_real)
    return fail(GM_EINVAL, "%sbad_gamma_s_real: Invalid gamma_s_real input argument.", P);
  if (!d->gamma_s_imag)
    return fail(GM_EINVAL, "%sbad_gamma_s_imag: Invalid gamma_s_imag input argument.", P);
  if (d->num_eigs > 0 && !d->exp_lambda)
    return fail(GM_EINVAL, "%sbad_exp_lambda: Invalid exp_lambda input argument.", P);
  if (d->num_eigs > 0 && !d->phi_lambda)
    return fail(GM_EINVAL, "%sbad_phi_lambda: Invalid phi_lambda input argument.", P);
  if (!std::isfinite(d->dt)) return fail(GM_EINVAL, "%sbad_dt: Invalid dt input argument.", P);
  if (!d->wts1) return fail(GM_EINVAL, "%sbad_wts1: Invalid wts1 input argument.", P);
  if (!d->wts2) return fail(GM_EINVAL, "%sbad_wts2: Invalid wts2 input argument.", P);
  if (!d->shifts) return fail(GM_EINVAL, "%sbad_shifts: Invalid shifts input argument c.", P);
  if (d->valvec && (d->valvec[0] != -1.0 || d->valvec[1] != 0.0 || d->valvec[2] != 1.0))
    return fail(GM_EINVAL, "%sbad_valvec: valvec must be [-1 0 1]", P);
  if ((long long)d->gamma_size * d->subspace_size * (long long)sizeof(double2) > 65536)
    return fail(GM_EINVAL, "%sbad_N: (d+1)*m = %d exceeds the 4096 coefficients kept in LDS", P,
                d->gamma_size * d->subspace_size);
  if (d->mass)
    for (int i = 0; i < d->n; ++i)
      if (!(d->mass[i] != 0.0))
        return fail(GM_EINVAL, "%sbad_mass: mass must be nonzero (entry %d)", P, i);
  return GM_OK;
}

gm_status check_device(int dev) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return fail(GM_EDEVICE, "no HIP device available (%s)", hipGetErrorString(e));
  if (dev < 0 || dev >= count) return fail(GM_EDEVICE, "device %d out of range (%d)", dev, count);
  hipDeviceProp_t prop;
  GM_HIP(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(GM_EDEVICE, "device %d is %s; this library is built for gfx950 only", dev,
                prop.gcnArchName);
  return GM_OK;
}

gm_status create_impl(const gm_create_desc* d, gm_solver* s) {
  GM_TRY(check_device(d->device));
  s->dev = d->device;
  GM_HIP(hipSetDevice(s->dev));
  GM_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  hipStream_t st = s->stream;
  s->n = d->n;
  s->nnz = d->nnz;
  s->m = d->subspace_size;
  s->nd = d->gamma_size;
  s->ns = d->contour_size;
  s->K = d->num_eigs;
  s->dt = d->dt;
  s->aode = d->aode; s->bode = d->bode; s->code = d->code;
  s->vrest = d->vrest; s->vamp = d->vamp; s->vth = d->vth; s->vpeak = d->vpeak;
  s->c1ion = d->c1ion; s->c2ion = d->c2ion;
  s->clamp = !(d->flags & GM_FLAG_NO_CLAMP);
  s->gamma.assign(d->gamma, d->gamma + s->nd);
  const size_t N = (size_t)s->n;

  // A -> CSR (a CSC input is the CSR of A^T: transpose its index structure,
  // which is what the reference gets by flipping the cuSPARSE op,
  // cusparse_wrapper.cpp:83-84).
  std::vector<int> rp(N + 1, 0), ci((size_t)s->nnz);
  std::vector<double> vv((size_t)s->nnz);
  if (d->sparse_format == GM_SPARSE_CSR) {
    std::copy(d->A_ptr, d->A_ptr + N + 1, rp.begin());
    std::copy(d->A_ind, d->A_ind + s->nnz, ci.begin());
    std::copy(d->A_values, d->A_values + s->nnz, vv.begin());
  } else {
    for (long long j = 0; j < s->nnz; ++j) rp[d->A_ind[j] + 1]++;
    for (size_t i = 0; i < N; ++i) rp[i + 1] += rp[i];
    std::vector<int> pos(rp.begin(), rp.end() - 1);
    for (int c = 0; c < s->n; ++c)
      for (int j = d->A_ptr[c]; j < d->A_ptr[c + 1]; ++j) {
        const int r = d->A_ind[j];
        ci[pos[r]] = c;
        vv[pos[r]] = d->A_values[j];
        pos[r]++;
      }
  }
  GM_TRY(upload(&s->rowptr, rp.data(), N + 1, st));
  GM_TRY(upload(&s->col, ci.data(), (size_t)s->nnz, st));
  GM_TRY(upload(&s->val, vv.data(), (size_t)s->nnz, st));
  int max_blocks = 1;
  for (int i = 0; i < kNumTiles; ++i) {
    const std::vector<int> blk = row_blocks(rp, s->n, kTiles[i]);
    s->nrowblk[i] = (int)blk.size() - 1;
    max_blocks = std::max(max_blocks, s->nrowblk[i]);
    GM_TRY(upload(&s->rowblk[i], blk.data(), blk.size(), st));
    std::vector<int> bj(blk.size());
    for (size_t b = 0; b < blk.size(); ++b) bj[b] = (int)rp[blk[b]];
    GM_TRY(upload(&s->blkj[i], bj.data(), bj.size(), st));
    // column offsets from each block's first row, if every one fits 16 bits
    std::vector<short> c16;
    if (!(d->flags & GM_FLAG_INDEX32)) {
      c16.resize((size_t)s->nnz);
      bool fits = true;
      for (int b = 0; b < s->nrowblk[i] && fits; ++b)
        for (long long j = rp[blk[b]]; j < rp[blk[b + 1]]; ++j) {
          const int off = ci[j] - blk[b];
          if (off < -32768 || off > 32767) {
            fits = false;
            break;
          }
          c16[j] = (short)off;
        }
      if (fits) GM_TRY(upload(&s->col16[i], c16.data(), c16.size(), st));
    }
    // upload() returns before the copy ends; blk and c16 must outlive it
    GM_HIP(hipStreamSynchronize(st));
  }

  // U (n x 2 column-major) -> interleaved, column 1 scaled by dt
  // (lanczos_impl.cu:79); Usol = U (lanczos_impl.cu:81-82).
  std::vector<double2> u2(N);
  for (size_t i = 0; i < N; ++i) u2[i] = make_double2(d->U[i], d->U[N + i] * s->dt);
  GM_TRY(upload(&s->Usol, u2.data(), N, st));
  GM_TRY(dalloc(&s->U, N));
  GM_TRY(dalloc(&s->X, N));
  GM_TRY(dalloc(&s->upd, N));
  GM_HIP(hipMemsetAsync(s->upd, 0, N * sizeof(double2), st));
  GM_HIP(hipMemsetAsync(s->X, 0, N * sizeof(double2), st));
  for (int i = 0; i < 2; ++i) {
    GM_TRY(dalloc(&s->W[i], N));
    GM_TRY(dalloc(&s->T[i], N));
  }
  GM_TRY(dalloc(&s->V, N * s->m));
  GM_TRY(dalloc(&s->Z, s->nd > 1 ? N * s->nd : 1));
  GM_TRY(dalloc(&s->C, (size_t)s->nd * s->m));
  // Q column-major n x K -> row-major.  The host copy lives until the stream
  // is synchronised below: upload() returns before the copy has finished.
  std::vector<double> qr;
  if (s->K > 0) {
    qr.resize(N * s->K);
    for (int k = 0; k < s->K; ++k)
      for (size_t i = 0; i < N; ++i) qr[i * s->K + k] = d->Q[(size_t)k * N + i];
    GM_TRY(upload(&s->Q, qr.data(), qr.size(), st));
    GM_TRY(upload(&s->el, d->exp_lambda, (size_t)s->K, st));
    GM_TRY(upload(&s->pl, d->phi_lambda, (size_t)s->K, st));
  }
  std::vector<double> ones;
  if (!d->mass) ones.assign(N, 1.0);
  GM_TRY(upload(&s->mass, d->mass ? d->mass : ones.data(), N, st));
  std::vector<double> zeros;
  if (!d->ode) zeros.assign(N, 0.0);
  GM_TRY(upload(&s->ode, d->ode ? d->ode : zeros.data(), N, st));
  const size_t ld = (size_t)s->m + 1;
  GM_TRY(dalloc(&s->aabb, 4 * ld));
  GM_HIP(hipMemsetAsync(s->aabb, 0, 4 * ld * sizeof(double), st));
  const size_t ny = (size_t)s->m * 2 * s->ns;
  GM_TRY(dalloc(&s->Yr, ny));
  GM_TRY(dalloc(&s->Yi, ny));
  GM_HIP(hipMemsetAsync(s->Yr, 0, ny * sizeof(double), st));
  GM_HIP(hipMemsetAsync(s->Yi, 0, ny * sizeof(double), st));
  GM_TRY(upload(&s->wts1, d->wts1, 2 * (size_t)s->ns, st));
  GM_TRY(upload(&s->wts2, d->wts2, 2 * (size_t)s->ns, st));
  GM_TRY(upload(&s->shifts, d->shifts, 4 * (size_t)s->m * s->ns, st));
  GM_TRY(upload(&s->gsr, d->gamma_s_real, (size_t)s->nd * s->ns, st));
  GM_TRY(upload(&s->gsi, d->gamma_s_imag, (size_t)s->nd * s->ns, st));
  GM_TRY(dalloc(&s->Zdefl, 2 * (size_t)std::max(1, s->K)));
  GM_TRY(dalloc(&s->wD, ny));
  GM_TRY(dalloc(&s->wDU, ny));
  GM_TRY(dalloc(&s->wDU2, ny));
  GM_TRY(dalloc(&s->wX, ny));
  GM_TRY(dalloc(&s->stage, 2 * N));

  // schedules
  const double avg = s->n ? (double)s->nnz / s->n : 1.0;
  int lanes = 1;
  while (lanes < 64 && lanes < avg) lanes <<= 1;
  const bool blocks = (d->flags & GM_FLAG_ROW_BLOCKS) ||
                      (!(d->flags & GM_FLAG_ROW_GROUPS) && avg <= 16.0);
  if (blocks)
    set_tile_schedule(s, 1, (d->flags & GM_FLAG_INDEX16) != 0,   // 1024-nonzero tiles
                      (d->flags & GM_FLAG_EARLY_LOADS) != 0);
  else
    set_schedule(s, lanes);
  // shifted solves: one wavefront per system when asked (and m allows)
  if ((d->flags & GM_FLAG_WAVE_SOLVES) && s->m <= 128) s->solve_sb = -1;
  s->G_el = std::max(1, std::min(4096, cdiv_i(s->n, kBlock)));
  s->chunk_el = cdiv_i(s->n, s->G_el);
  s->G_df = std::max(1, std::min(512, cdiv_i(s->n, kBlock)));
  s->chunk_df = cdiv_i(s->n, s->G_df);
  const size_t npart = std::max<size_t>(2 * (size_t)std::max(4096, max_blocks),
                                        (size_t)s->G_df * 2 * std::max(1, s->K));
  GM_TRY(dalloc(&s->partials, npart));
  GM_TRY(dalloc(&s->partials2, 2 * (size_t)kFuseG));
  GM_TRY(dalloc(&s->status, 1));
  GM_HIP(hipMemsetAsync(s->status, 0, sizeof(unsigned), st));

  // deflation of the initial vectors (lanczos_impl.cu:84-89) and beta_0
  s->x_zero = true;
  GM_TRY(enqueue_update(s, gm::kUpdPrepare));
  GM_HIP(hipStreamSynchronize(st));

  if (d->flags & GM_FLAG_AUTOTUNE) {
    // Time the build-subspace SpMM for each lane width on this matrix and
    // keep the fastest.  Each candidate runs a short Lanczos run; the solver
    // state is rebuilt by the next build_subspace anyway.
    // Candidates: row groups of 1..32 lanes (c = lanes), and row blocks
    // encoded as c = -(1 + tile index + 100 c16 + 1000 pre): every tile size
    // with 32-bit indices or 16-bit column offsets (where they fit), with the
    // epilogue operands loaded early or late.
    std::vector<int> cands = {1, 2, 4, 8, 16, 32};
    for (int pre = 0; pre < 2; ++pre)
      for (int c16 = 0; c16 < 2; ++c16)
        for (int i = 0; i < kNumTiles; ++i) cands.push_back(-(1 + i + 100 * c16 + 1000 * pre));
    auto apply = [&](int c) {
      if (c > 0) {
        set_schedule(s, c);
      } else {
        const int e = -c - 1;
        set_tile_schedule(s, e % 100, (e / 100) % 10 == 1, e / 1000 == 1);
      }
    };
    float best = 1e30f;
    int best_c = s->tile > 0 ? -(1 + s->tile_idx + (s->use16 ? 100 : 0) + (s->pre ? 1000 : 0))
                             : s->lanes;
    hipEvent_t e0, e1;
    GM_HIP(hipEventCreate(&e0));
    GM_HIP(hipEventCreate(&e1));
    for (int l : cands) {
      if (l > 0 && l > 4 * avg + 4) continue;
      if (l < 0 && ((-l - 1) / 100) % 10 == 1 && !s->col16[(-l - 1) % 100]) continue;
      apply(l);
      GM_TRY(enqueue_build(s));  // warm-up
      GM_HIP(hipEventRecord(e0, st));
      GM_TRY(enqueue_build(s));
      GM_HIP(hipEventRecord(e1, st));
      GM_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      GM_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) { best = ms; best_c = l; }
    }
    apply(best_c);
    // The shifted solves in three forms, each timed together with apply_V:
    // one lane per system with the factors in global memory, or in LDS (a
    // few systems a block, which spreads Y over several blocks and XCDs, paid
    // for by the collapse that reads Y), or one wavefront per system (cyclic
    // reduction, m <= 128).  Timed on the last build's tridiagonals; apply_V's
    // output X is cleared afterwards.
    {
      std::vector<int> cand = {0};
      if (gm::solve_systems_per_block(s->m) > 0) cand.push_back(gm::solve_systems_per_block(s->m));
      if (s->m <= 128) cand.push_back(-1);
      if (cand.size() > 1) {
        float tbest = 1e30f;
        int pick = 0;
        for (int sb : cand) {
          s->solve_sb = sb;
          GM_TRY(enqueue_solves(s));  // warm-up
          GM_TRY(enqueue_apply(s));
          GM_HIP(hipEventRecord(e0, st));
          for (int i = 0; i < 3; ++i) {
            GM_TRY(enqueue_solves(s));
            GM_TRY(enqueue_apply(s));
          }
          GM_HIP(hipEventRecord(e1, st));
          GM_HIP(hipEventSynchronize(e1));
          float ms = 0.f;
          GM_HIP(hipEventElapsedTime(&ms, e0, e1));
          if (ms < tbest) { tbest = ms; pick = sb; }
        }
        s->solve_sb = pick;
        s->x_zero = true;   // X holds the timing runs' output: the next apply_V overwrites it
      }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    GM_HIP(hipMemsetAsync(s->status, 0, sizeof(unsigned), st));
    GM_HIP(hipStreamSynchronize(st));
  }
  return GM_OK;
}

}  // namespace

// ================================================================ C ABI
extern "C" {

const char* gm_last_error(void) { return g_last_error.c_str(); }

gm_status gm_lanczos_create(const gm_create_desc* desc, gm_solver** out) {
  g_last_error.clear();
  if (!out) return fail(GM_EINVAL, "MATLAB:lanczos_create:maxlhs: One output argument is required.");
  *out = nullptr;
  GM_TRY(validate(desc));
  gm_solver* s = new (std::nothrow) gm_solver();
  if (!s) return fail(GM_ENOMEM, "host allocation of the solver failed");
  gm_status st = create_impl(desc, s);
  if (st != GM_OK) {
    std::string keep = g_last_error;
    free_all(s);
    delete s;
    g_last_error = keep;
    return st;
  }
  {
    std::lock_guard<std::mutex> lk(g_registry_mu);
    g_registry.insert(s);
  }
  *out = s;
  return GM_OK;
}

gm_status gm_lanczos_delete(gm_solver* s) {
  {
    std::lock_guard<std::mutex> lk(g_registry_mu);
    if (!s || !g_registry.count(s))
      return fail(GM_EBADHANDLE, "MATLAB:lanczos_delete:inputNotHandle: invalid or deleted handle");
    g_registry.erase(s);
  }
  hipSetDevice(s->dev);
  if (s->stream) hipStreamSynchronize(s->stream);
  free_all(s);
  delete s;
  return GM_OK;
}

gm_status gm_build_subspace(gm_solver* s) {
  GM_CHECK_HANDLE(s);
  GM_TRY(enqueue_build(s));
  s->built = true;
  s->solved = false;
  return GM_OK;
}

gm_status gm_lin_sys_solves(gm_solver* s) {
  GM_CHECK_HANDLE(s);
  if (!s->built) return fail(GM_ESTATE, "lin_sys_solves called before build_subspace");
  GM_TRY(enqueue_solves(s));
  s->solved = true;
  return GM_OK;
}

gm_status gm_apply_V(gm_solver* s) {
  GM_CHECK_HANDLE(s);
  if (!s->built) return fail(GM_ESTATE, "apply_V called before build_subspace");
  return enqueue_apply(s);
}

gm_status gm_update_U(gm_solver* s) {
  GM_CHECK_HANDLE(s);
  return enqueue_update(s, gm::kUpdFisher);
}

gm_status gm_update_U_monodomain(gm_solver* s) {
  GM_CHECK_HANDLE(s);
  return enqueue_update(s, gm::kUpdMono);
}

gm_status gm_set_ms_params(gm_solver* s, const gm_ms_params* p) {
  GM_CHECK_HANDLE(s);
  if (!p || !(p->tau_in > 0) || !(p->tau_out > 0) || !(p->tau_open > 0) || !(p->tau_close > 0))
    return fail(GM_EINVAL, "Mitchell-Schaeffer time constants must be positive");
  s->ms[0] = p->tau_in;
  s->ms[1] = p->tau_out;
  s->ms[2] = p->tau_open;
  s->ms[3] = p->tau_close;
  s->ms[4] = p->v_gate;
  for (auto& g : s->graph)   // captured steps hold the old parameters
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  return GM_OK;
}

gm_status gm_update_U_ms(gm_solver* s) {
  GM_CHECK_HANDLE(s);
  return enqueue_update(s, gm::kUpdMS);
}

gm_status gm_get_U(gm_solver* s, double* dst) {
  GM_CHECK_HANDLE(s);
  return fetch_pair(s, s->Usol, dst);
}

gm_status gm_get_xsol(gm_solver* s, double* dst) {
  GM_CHECK_HANDLE(s);
  if (s->x_zero) {
    if (!dst) return fail(GM_EINVAL, "null destination pointer");
    GM_HIP(hipStreamSynchronize(s->stream));
    std::fill(dst, dst + 2 * (size_t)s->n, 0.0);
    return GM_OK;
  }
  return fetch_pair(s, s->X, dst);
}

gm_status gm_get_aabb(gm_solver* s, double* dst) {
  GM_CHECK_HANDLE(s);
  if (!dst) return fail(GM_EINVAL, "null destination pointer");
  GM_HIP(hipMemcpyAsync(dst, s->aabb, 4 * ((size_t)s->m + 1) * sizeof(double),
                        hipMemcpyDeviceToHost, s->stream));
  GM_HIP(hipStreamSynchronize(s->stream));
  return GM_OK;
}

gm_status gm_set_y(gm_solver* s, const double* yreal, const double* yimag) {
  GM_CHECK_HANDLE(s);
  if (!yreal || !yimag) return fail(GM_EINVAL, "null y pointer");
  const size_t ny = (size_t)s->m * 2 * s->ns;
  GM_HIP(hipMemcpyAsync(s->Yr, yreal, ny * sizeof(double), hipMemcpyHostToDevice, s->stream));
  GM_HIP(hipMemcpyAsync(s->Yi, yimag, ny * sizeof(double), hipMemcpyHostToDevice, s->stream));
  GM_HIP(hipStreamSynchronize(s->stream));
  s->solved = true;
  return GM_OK;
}

gm_status gm_get_y(gm_solver* s, double* yreal, double* yimag) {
  GM_CHECK_HANDLE(s);
  if (!yreal || !yimag) return fail(GM_EINVAL, "null y pointer");
  const size_t ny = (size_t)s->m * 2 * s->ns;
  GM_HIP(hipMemcpyAsync(yreal, s->Yr, ny * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  GM_HIP(hipMemcpyAsync(yimag, s->Yi, ny * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  GM_HIP(hipStreamSynchronize(s->stream));
  return GM_OK;
}

gm_status gm_get_state(gm_solver* s, double* u, double* dtg, double* ode) {
  GM_CHECK_HANDLE(s);
  if (u || dtg) {
    std::vector<double> tmp(2 * (size_t)s->n);
    GM_TRY(fetch_pair(s, s->Usol, tmp.data()));
    if (u) std::copy(tmp.begin(), tmp.begin() + s->n, u);
    if (dtg) std::copy(tmp.begin() + s->n, tmp.end(), dtg);
  }
  if (ode) {
    GM_HIP(hipMemcpyAsync(ode, s->ode, (size_t)s->n * sizeof(double), hipMemcpyDeviceToHost,
                          s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
  }
  return GM_OK;
}

gm_status gm_set_state(gm_solver* s, const double* u, const double* dtg, const double* ode) {
  GM_CHECK_HANDLE(s);
  const size_t N = (size_t)s->n;
  if (u || dtg) {
    std::vector<double> cur(2 * N);
    GM_TRY(fetch_pair(s, s->Usol, cur.data()));
    std::vector<double2> u2(N);
    for (size_t i = 0; i < N; ++i)
      u2[i] = make_double2(u ? u[i] : cur[i], dtg ? dtg[i] : cur[N + i]);
    GM_HIP(hipMemcpyAsync(s->Usol, u2.data(), N * sizeof(double2), hipMemcpyHostToDevice, s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
  }
  if (ode) {
    GM_HIP(hipMemcpyAsync(s->ode, ode, N * sizeof(double), hipMemcpyHostToDevice, s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
  }
  s->x_zero = true;
  GM_TRY(enqueue_update(s, gm::kUpdPrepare));
  GM_HIP(hipStreamSynchronize(s->stream));
  return GM_OK;
}

gm_status gm_step(gm_solver* s, int32_t nsteps, int32_t monodomain, int32_t use_graph) {
  GM_CHECK_HANDLE(s);
  if (nsteps < 0) return fail(GM_EINVAL, "nsteps must be >= 0");
  if (monodomain < 0 || monodomain > 2) return fail(GM_EINVAL, "reaction model must be 0, 1 or 2");
  const int mono = monodomain;
  if (!s->x_zero) {
    // a partial step is pending (apply_V without update_U): finish it eagerly
    use_graph = 0;
  }
  if (!use_graph) {
    for (int i = 0; i < nsteps; ++i) GM_TRY(enqueue_step(s, mono));
  } else {
    if (!s->graph[mono]) {
      hipGraph_t g = nullptr;
      GM_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
      gm_status st = enqueue_step(s, mono);
      hipError_t e = hipStreamEndCapture(s->stream, &g);
      if (st != GM_OK) {
        if (g) hipGraphDestroy(g);
        return st;
      }
      if (e != hipSuccess) return fail(GM_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&s->graph[mono], g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) {
        s->graph[mono] = nullptr;
        return fail(GM_EDEVICE, "graph instantiate failed: %s", hipGetErrorString(e));
      }
    }
    for (int i = 0; i < nsteps; ++i) GM_HIP(hipGraphLaunch(s->graph[mono], s->stream));
  }
  if (nsteps > 0) {
    s->built = true;
    s->solved = true;
    s->x_zero = true;
  }
  return GM_OK;
}

gm_status gm_synchronize(gm_solver* s) {
  GM_CHECK_HANDLE(s);
  return check_status(s);
}

gm_status gm_get_info(gm_solver* s, int64_t* n, int64_t* nnz, int32_t* m, int32_t* gamma_size,
                      int32_t* contour_size, int32_t* num_eigs, int32_t* spmm_lanes) {
  GM_CHECK_HANDLE(s);
  if (n) *n = s->n;
  if (nnz) *nnz = s->nnz;
  if (m) *m = s->m;
  if (gamma_size) *gamma_size = s->nd;
  if (contour_size) *contour_size = s->ns;
  if (num_eigs) *num_eigs = s->K;
  if (spmm_lanes) *spmm_lanes = s->tile > 0 ? -s->tile : s->lanes;
  return GM_OK;
}

gm_status gm_get_stream(gm_solver* s, void** stream) {
  GM_CHECK_HANDLE(s);
  if (!stream) return fail(GM_EINVAL, "null stream pointer");
  *stream = reinterpret_cast<void*>(s->stream);
  return GM_OK;
}

gm_status gm_get_schedule(gm_solver* s, int32_t* kind, int32_t* param, int32_t* index_bytes) {
  GM_CHECK_HANDLE(s);
  if (kind) *kind = s->tile > 0 ? (s->pre ? 2 : 1) : 0;
  if (param) *param = s->tile > 0 ? s->tile : s->lanes;
  if (index_bytes) *index_bytes = (s->tile > 0 && s->use16) ? 2 : 4;
  return GM_OK;
}

gm_status gm_get_solve_form(gm_solver* s, int32_t* form, int32_t* systems_per_block) {
  GM_CHECK_HANDLE(s);
  if (form) *form = s->solve_sb < 0 ? 2 : (s->solve_sb > 0 ? 1 : 0);
  if (systems_per_block) *systems_per_block = s->solve_sb < 0 ? 1 : (s->solve_sb > 0 ? s->solve_sb : 64);
  return GM_OK;
}

gm_status gm_profile_spmm(gm_solver* s, int32_t mode, int32_t reps, double* avg_ms) {
  GM_CHECK_HANDLE(s);
  if (!avg_ms || reps < 1) return fail(GM_EINVAL, "avg_ms must be non-null and reps >= 1");
  if (mode != 0 && mode != 1) return fail(GM_EINVAL, "mode must be 0 or 1");
  if (!s->built) return fail(GM_ESTATE, "profile_spmm needs a built subspace (run a step first)");
  const size_t N = (size_t)s->n;
  SpmmArgs a = spmm_base(s);
  a.g0 = s->gamma[0];
  a.gi = s->nd > 1 ? s->gamma[1] : 0.0;
  // Inputs are V[0] (unit norm) and scratch T/W buffers; k = 1 so the fused
  // beta/dot terms are live.  The alpha written to aabb row 1 is overwritten
  // by the next build_subspace.
  const int k = s->m > 1 ? 1 : 0;
  a.k = k;
  a.vk = s->V + (size_t)k * N;
  a.vkm1 = k > 0 ? s->V : nullptr;
  hipEvent_t e0, e1;
  GM_HIP(hipEventCreate(&e0));
  GM_HIP(hipEventCreate(&e1));
  auto one = [&](int i) -> gm_status {
    if (mode == 0) {
      a.x = (i & 1) ? s->T[1] : s->V;
      a.y = (i & 1) ? s->T[0] : s->T[1];
      return launch_spmm_mode<gm::kMid>(s, a);
    }
    a.x = s->V + (size_t)k * N;
    a.y = s->W[i & 1];
    return launch_spmm_mode<gm::kLast>(s, a);
  };
  gm_status st = one(0);  // warm-up
  if (st == GM_OK) {
    GM_HIP(hipEventRecord(e0, s->stream));
    for (int i = 0; i < reps && st == GM_OK; ++i) st = one(i);
    GM_HIP(hipEventRecord(e1, s->stream));
    GM_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GM_HIP(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = (double)ms / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return st;
}

gm_status gm_profile_spmm_flushed(gm_solver* s, int32_t mode, int32_t reps, int64_t flush_bytes,
                                  double* avg_ms) {
  GM_CHECK_HANDLE(s);
  if (!avg_ms || reps < 1 || flush_bytes < 32)
    return fail(GM_EINVAL, "avg_ms must be non-null, reps >= 1 and flush_bytes >= 32");
  if (mode != 0 && mode != 1) return fail(GM_EINVAL, "mode must be 0 or 1");
  if (!s->built) return fail(GM_ESTATE, "profile_spmm_flushed needs a built subspace (run a step first)");
  const size_t N = (size_t)s->n;
  SpmmArgs a = spmm_base(s);
  a.g0 = s->gamma[0];
  a.gi = s->nd > 1 ? s->gamma[1] : 0.0;
  const int k = s->m > 1 ? 1 : 0;   // as in gm_profile_spmm
  a.k = k;
  a.vk = s->V + (size_t)k * N;
  a.vkm1 = k > 0 ? s->V : nullptr;
  const long long n = flush_bytes / 32;   // half read, half written
  double2 *src = nullptr, *dst = nullptr;
  GM_TRY(dalloc(&src, (size_t)n));
  gm_status st = dalloc(&dst, (size_t)n);
  if (st != GM_OK) {
    (void)hipFree(src);
    return st;
  }
  hipEvent_t e0, e1;
  GM_HIP(hipEventCreate(&e0));
  GM_HIP(hipEventCreate(&e1));
  const int grid = (int)std::min<long long>(8192, (n + kBlock - 1) / kBlock);
  double total = 0.0;
  for (int i = 0; i <= reps && st == GM_OK; ++i) {   // i = 0: warm-up, not timed
    hipLaunchKernelGGL(gm::k_copy16<0>, dim3(grid), dim3(kBlock), 0, s->stream, src, dst, n);
    GM_HIP(hipEventRecord(e0, s->stream));
    if (mode == 0) {
      a.x = (i & 1) ? s->T[1] : s->V;
      a.y = (i & 1) ? s->T[0] : s->T[1];
      st = launch_spmm_mode<gm::kMid>(s, a);
    } else {
      a.x = s->V + (size_t)k * N;
      a.y = s->W[i & 1];
      st = launch_spmm_mode<gm::kLast>(s, a);
    }
    GM_HIP(hipEventRecord(e1, s->stream));
    GM_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GM_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (i > 0) total += ms;
  }
  *avg_ms = total / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(src);
  (void)hipFree(dst);
  return st;
}

gm_status gm_profile_barrier(int32_t device, int32_t blocks, int32_t iters, double* grid_sync_us,
                             double* launch_us) {
  if (!grid_sync_us || !launch_us || blocks < 1 || iters < 1)
    return fail(GM_EINVAL, "bad profile_barrier arguments");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return fail(GM_EDEVICE, "no HIP device %d", device);
  GM_HIP(hipSetDevice(device));
  int coop = 0, per_cu = 0, cus = 0;
  GM_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device));
  GM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  GM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_grid_barriers,
                                                       kBlock, 0));
  if (!coop) return fail(GM_EDEVICE, "device %d has no cooperative launch", device);
  if (blocks > per_cu * cus)
    return fail(GM_EINVAL, "%d blocks cannot all be resident (at most %d)", blocks, per_cu * cus);
  hipStream_t st = nullptr;
  GM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  double* buf = nullptr;
  gm_status rc = dalloc(&buf, (size_t)blocks * kBlock);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == GM_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(GM_EDEVICE, "events");
  auto timed = [&](bool coop_kernel, int n, float* ms) -> gm_status {
    void* args[] = {&n, &buf};
    if (hipEventRecord(e0, st) != hipSuccess) return fail(GM_EDEVICE, "event");
    if (coop_kernel) {
      const hipError_t e = hipLaunchCooperativeKernel((const void*)k_grid_barriers, dim3(blocks),
                                                      dim3(kBlock), args, 0, st);
      if (e != hipSuccess) return fail(GM_EDEVICE, "cooperative launch: %s", hipGetErrorString(e));
    } else {
      for (int i = 0; i < n; ++i)
        hipLaunchKernelGGL(k_touch, dim3(blocks), dim3(kBlock), 0, st, buf);
    }
    if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(ms, e0, e1) != hipSuccess)
      return fail(GM_EDEVICE, "timing");
    return GM_OK;
  };
  float ms = 0.f;
  if (rc == GM_OK) rc = timed(true, 1, &ms);            // warm-up
  if (rc == GM_OK) rc = timed(true, iters, &ms);
  if (rc == GM_OK) *grid_sync_us = 1000.0 * ms / iters;
  if (rc == GM_OK) rc = timed(false, 1, &ms);
  if (rc == GM_OK) rc = timed(false, iters, &ms);
  if (rc == GM_OK) *launch_us = 1000.0 * ms / iters;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (buf) (void)hipFree(buf);
  (void)hipStreamDestroy(st);
  return rc;
}

gm_status gm_profile_copy(gm_solver* s, int64_t bytes, int32_t reps, double* avg_ms) {
  GM_CHECK_HANDLE(s);
  if (!avg_ms || reps < 1 || bytes < 16) return fail(GM_EINVAL, "bad profile_copy arguments");
  const long long n = bytes / 16;
  double2 *src = nullptr, *dst = nullptr;
  GM_TRY(dalloc(&src, (size_t)n));
  gm_status st = dalloc(&dst, (size_t)n);
  if (st != GM_OK) {
    (void)hipFree(src);
    return st;
  }
  hipEvent_t e0, e1;
  GM_HIP(hipEventCreate(&e0));
  GM_HIP(hipEventCreate(&e1));
  const int grid = (int)std::min<long long>(8192, (n + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(gm::k_copy16<0>, dim3(grid), dim3(kBlock), 0, s->stream, src, dst, n);
  GM_HIP(hipEventRecord(e0, s->stream));
  for (int i = 0; i < reps; ++i)
    hipLaunchKernelGGL(gm::k_copy16<0>, dim3(grid), dim3(kBlock), 0, s->stream, src, dst, n);
  GM_HIP(hipEventRecord(e1, s->stream));
  GM_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  GM_HIP(hipEventElapsedTime(&ms, e0, e1));
  *avg_ms = (double)ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(src);
  (void)hipFree(dst);
  return GM_OK;
}

}  // extern "C"
