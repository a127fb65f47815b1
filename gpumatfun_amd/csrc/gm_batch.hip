// gm_batch.hip -- batched form: many right-hand sides sharing one matrix.
//
// P independent problems share A (and the contour nodes, weights and
// polynomial).  Problem p owns the column pair [u_p, dt g_p] -- the same two
// columns as the plain path (exp and phi) -- so a batch of P problems is R = 2P
// columns, each with its own Lanczos process (its own alpha, beta).
//
// Layout: every N x R quantity is row-major, columns contiguous
// (X[i*ld + c]); a CSR gather of x[col] is then one coalesced R*8-byte load
// across the lanes of a wavefront, and each matrix entry is read once for all
// R columns.  Lanes map to columns: with R >= 64 a wavefront works on one row
// at a time (64 columns per pass), with R < 64 it works on 64/R rows.
// Columns are processed in chunks of Rc (a power of two); the Lanczos basis of
// one chunk, m x N x Rc doubles, is what bounds Rc.
//
// Per step and chunk: build (m iterations of SpMM kernels + finish), shifted
// solves (Rc x n_s systems, one lane each), collapse, V*C, Horner SpMMs, and
// the update: Fisher-KPP (update_U.cu:8-16, optional clamp) or
// Mitchell-Schaeffer (gate h per problem).  With K shared deflation vectors
// the step ends with Z = Q^T Usol on FP64 MFMA over all R columns, then the
// deflated starting vectors Ud = Usol - Q Z and upd = Q f(Lambda) Z
// (lanczos_impl.cu:84-89, update_U.cu:18-27).
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

// ====================================================================== host
namespace {

constexpr int kBatchGrid = 1024;   // most blocks of the column kernels (set_bschedule)
// leaf of the internal locality order: on the 128^3 Laplacian, leaves of 1024
// to 65536 nodes leave the same fraction of gathers (~16%) with no reuse within
// 2048 rows, against 43% in natural order; timed with 256 columns, leaves of
// 512, 1024, 4096 and 16384 nodes gave 730, 729, 725 and 721 solves/s
constexpr int kLocalityLeaf = 512;


thread_local std::string g_berr;
std::mutex g_bmu;
std::unordered_set<const void*> g_breg;

gm_status bfail(gm_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_berr = buf;
  return st;
}

#define B_HIP(call)                                                                      \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return bfail(e_ == hipErrorOutOfMemory ? GM_ENOMEM : GM_EDEVICE, "%s: %s (%s:%d)", \
                   #call, hipGetErrorString(e_), __FILE__, __LINE__);                    \
  } while (0)
#define B_TRY(x)              \
  do {                        \
    gm_status s_ = (x);       \
    if (s_ != GM_OK) return s_; \
  } while (0)

inline int cdivi(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace

struct gm_batch {
  int dev = 0;
  hipStream_t stream = nullptr;
  int n = 0, P = 0, R = 0, Rc = 0, m = 0, nd = 0, ns = 0;
  long long nnz = 0;
  double dt = 0.0;
  bool clamp = true, x_zero = true;
  std::vector<double> gamma;
  int *rowptr = nullptr, *col = nullptr, *rowblk = nullptr;
  int nrowblk = 0;                  // LDS-staged SpMM row blocks
  // slab SpMM of chunks of >= 32 columns (gmb::k_bspmm_xs): its row-block
  // descriptors (4 ints each), grid (a multiple of 8) and whether the step
  // uses it (with it, the basis is kept unnormalised)
  int* rowblk_xs = nullptr;
  int nrowblk_xs = 0, G_xs = 8;
  bool xs = false;
  // GM_FLAG_LOCALITY_ORDER: device row q holds the caller's node perm[q]
  std::vector<int> perm;
  double *val = nullptr, *Usol = nullptr, *X = nullptr, *W0 = nullptr, *W1 = nullptr, *T0 = nullptr,
         *T1 = nullptr, *V = nullptr, *Z = nullptr, *C = nullptr, *al = nullptr, *be = nullptr,
         *yr = nullptr, *yi = nullptr, *gsr = nullptr, *gsi = nullptr, *wts1 = nullptr,
         *wts2 = nullptr, *etaxi = nullptr, *mass = nullptr, *partials = nullptr, *stage = nullptr;
  double2 *wD = nullptr, *wDU = nullptr, *wDU2 = nullptr, *wX = nullptr;
  unsigned* status = nullptr;
  int G_sp = 1, chunk_sp = 1;
  // deflation by the shared (Q, Lambda): Q row-major [N][K], Zd [K][R],
  // Ud = deflated starting vectors, upd = the deflated part's update
  int K = 0, G_df = 1, chunk_df = 1;
  double *Q = nullptr, *el = nullptr, *pl = nullptr, *Zd = nullptr, *Ud = nullptr,
         *upd = nullptr;
  // reaction: GM_REACTION_FISHER or GM_REACTION_MS (gate h in ode, [N][P])
  int reaction = 0;
  double* ode = nullptr;
  double ms[5] = {0.3, 6.0, 120.0, 150.0, 0.13};
  hipGraphExec_t graph = nullptr;
};

namespace {

bool blive(const gm_batch* b) {
  std::lock_guard<std::mutex> lk(g_bmu);
  return b && g_breg.count(b);
}

template <typename T>
gm_status balloc(T** p, size_t count) {
  *p = nullptr;
  B_HIP(hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T)));
  return GM_OK;
}

void bfree_all(gm_batch* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  void* bufs[] = {b->rowptr, b->col, b->rowblk, b->rowblk_xs, b->val, b->Usol, b->X, b->W0, b->W1, b->T0, b->T1, b->V,
                  b->Z, b->C, b->al, b->be, b->yr, b->yi, b->gsr, b->gsi, b->wts1, b->wts2,
                  b->etaxi, b->mass, b->partials, b->stage, b->wD, b->wDU, b->wDU2, b->wX,
                  b->status, b->Q, b->el, b->pl, b->Zd, b->Ud, b->upd, b->ode};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
}

gm_status breduce(gm_batch* b, int nv, double* out, int do_sqrt, int flag_zero, int G = 0) {
  gm::ReduceArgs r{};
  r.partials = b->partials;
  r.G = G > 0 ? G : b->G_sp;
  r.nv = nv;
  r.out = out;
  r.stride = 1;
  r.do_sqrt = do_sqrt;
  r.flag_zero = flag_zero;
  r.status = b->status;
  hipLaunchKernelGGL(gm::k_reduce<0>, dim3(gm::reduce_grid(r.nv)), dim3(gm::kReduceBlock), 0, b->stream, r);
  B_HIP(hipGetLastError());
  return GM_OK;
}

// Rows of block partials that a kLast/kOnly SpMM of a chunk of R columns
// leaves for k_reduce.
int spmm_partial_rows(const gm_batch* b, int R) {
  return (R >= 32 && b->xs) ? b->G_xs / (R / 32) : b->G_sp;
}

template <int R>
struct Launch {
  // chunks of >= 128 columns stream v_k, v_{k-1} and y non-temporally
  static constexpr bool kNT = R >= 128;
  static gm_status spmm(gm_batch* b, int mode, const gmb::BSpmm& a0) {
    if constexpr (R >= 32) {
      if (b->xs) {
        gmb::BSpmm a = a0;
        a.rowblk = b->rowblk_xs;
        a.nrowblk = b->nrowblk_xs;
        dim3 g(b->G_xs), t(gm::kBlock);
        switch (mode) {
          case gmb::kFirst: hipLaunchKernelGGL((gmb::k_bspmm_xs<R, gmb::kFirst, kNT, true>), g, t, 0, b->stream, a); break;
          case gmb::kMid: hipLaunchKernelGGL((gmb::k_bspmm_xs<R, gmb::kMid, kNT, true>), g, t, 0, b->stream, a); break;
          case gmb::kLast: hipLaunchKernelGGL((gmb::k_bspmm_xs<R, gmb::kLast, kNT, true>), g, t, 0, b->stream, a); break;
          case gmb::kOnly: hipLaunchKernelGGL((gmb::k_bspmm_xs<R, gmb::kOnly, kNT, true>), g, t, 0, b->stream, a); break;
          default: hipLaunchKernelGGL((gmb::k_bspmm_xs<R, gmb::kHorner, kNT, true>), g, t, 0, b->stream, a); break;
        }
        B_HIP(hipGetLastError());
        return GM_OK;
      }
    }
    const gmb::BSpmm& a = a0;
    dim3 g(b->G_sp), t(gm::kBlock);
    switch (mode) {
      case gmb::kFirst: hipLaunchKernelGGL((gmb::k_bspmm<R, gmb::kFirst, true, kNT>), g, t, 0, b->stream, a); break;
      case gmb::kMid: hipLaunchKernelGGL((gmb::k_bspmm<R, gmb::kMid, true, kNT>), g, t, 0, b->stream, a); break;
      case gmb::kLast: hipLaunchKernelGGL((gmb::k_bspmm<R, gmb::kLast, true, kNT>), g, t, 0, b->stream, a); break;
      case gmb::kOnly: hipLaunchKernelGGL((gmb::k_bspmm<R, gmb::kOnly, true, kNT>), g, t, 0, b->stream, a); break;
      default: hipLaunchKernelGGL((gmb::k_bspmm<R, gmb::kHorner, true, kNT>), g, t, 0, b->stream, a); break;
    }
    B_HIP(hipGetLastError());
    return GM_OK;
  }
  // vw: the basis holds the unnormalised w_k (the slab SpMM's dataflow)
  template <bool VW>
  static void finish_launch(gm_batch* b, const gmb::BFinish& a) {
    if constexpr (R >= 4 && R <= 64) {   // measured: 34.3 against 35.9 us at 32 columns; 2.43 against 2.40 ms at 256
      hipLaunchKernelGGL((gmb::k_bfinish2<R, VW>), dim3(b->G_sp), dim3(gm::kBlock), 0, b->stream, a);
    } else {
      hipLaunchKernelGGL((gmb::k_bfinish<R, VW>), dim3(b->G_sp), dim3(gm::kBlock), 0, b->stream, a);
    }
  }
  static gm_status finish(gm_batch* b, const gmb::BFinish& a, bool vw) {
    if (vw) finish_launch<true>(b, a);
    else finish_launch<false>(b, a);
    B_HIP(hipGetLastError());
    return GM_OK;
  }
  static gm_status update(gm_batch* b, const gmb::BUpdate& a) {
    hipLaunchKernelGGL((gmb::k_bupdate<R>), dim3(b->G_sp), dim3(gm::kBlock), 0, b->stream, a);
    B_HIP(hipGetLastError());
    return GM_OK;
  }
};

// Chunk widths are powers of two >= 2: four columns per thread from 4 up.
template <int ND>
void launch_bvc(gm_batch* b, const gmb::BVC& a) {
  if (b->Rc >= 4) {
    const long long e = (long long)b->n * b->Rc / 4;
    hipLaunchKernelGGL((gmb::k_bvc<ND, 4>), dim3(cdivi(e, gm::kBlock)), dim3(gm::kBlock), 0, b->stream, a);
  } else {
    const long long e = (long long)b->n * b->Rc / 2;
    hipLaunchKernelGGL((gmb::k_bvc<ND, 2>), dim3(cdivi(e, gm::kBlock)), dim3(gm::kBlock), 0, b->stream, a);
  }
}

// One step of one chunk of Rc columns starting at column c0.
template <int R>
gm_status enqueue_chunk(gm_batch* b, int c0) {
  const size_t N = (size_t)b->n;
  const int d = b->nd - 1;
  const size_t ldu = (size_t)b->R;
  double* Wb[2] = {b->W0, b->W1};
  double* Tb[2] = {b->T0, b->T1};
  gmb::BSpmm base{};
  base.rowptr = b->rowptr;
  base.col = b->col;
  base.val = b->val;
  base.nrows = b->n;
  base.chunk = b->chunk_sp;
  base.rowblk = b->rowblk;
  base.nrowblk = b->nrowblk;
  base.al = b->al;
  base.be = b->be;
  base.alpha_out = b->al;
  base.partials = b->partials;
  base.g0 = b->gamma[0];
  // the chunk's starting vectors: Usol columns c0..c0+R-1 (row stride R_total)
  {
    gmb::BNorm nn{};
    nn.u = (b->K ? b->Ud : b->Usol) + c0;
    nn.ldu = ldu;
    nn.be = b->be;
    nn.nrows = b->n;
    nn.chunk = b->chunk_sp;
    nn.partials = b->partials;
    nn.status = b->status;
    hipLaunchKernelGGL((gmb::k_bnorm<R>), dim3(b->G_sp), dim3(gm::kBlock), 0, b->stream, nn);
    B_HIP(hipGetLastError());
    B_TRY(breduce(b, R, b->be, 1, 0));                                   // beta_0
  }
  // With the slab SpMM the basis V holds the unnormalised w_k (the finish
  // writes w_{k+1} straight into V[k+1], the SpMM gathers it from there and
  // forms v_k = w_k / beta_k where it needs it), so v_k is never stored: one
  // vector pass less per Lanczos step.  V C then takes C / beta.
  const bool ub = R >= 32 && b->xs;
  auto Vk = [&](int k) { return b->V + (size_t)k * N * R; };
  for (int k = 0; k < b->m; ++k) {
    gmb::BSpmm a = base;
    const double* wsrc = k == 0 ? (b->K ? b->Ud : b->Usol) + c0 : (ub ? Vk(k) : Wb[(k - 1) & 1]);
    const size_t ldw = k == 0 ? ldu : (size_t)R;
    double* wnext = Wb[k & 1];
    a.k = k;
    a.vk = Vk(k);
    a.vkm1 = k > 0 ? Vk(k - 1) : nullptr;
    a.x = wsrc;
    a.ldx = ldw;
    // kFirst/kOnly: with ub only k = 0 stores (the raw w_0)
    double* vk_first = (ub && k > 0) ? nullptr : Vk(k);
    if (d == 0) {
      a.vk = vk_first;
      a.y = wnext;
      a.ldy = R;
      B_TRY(Launch<R>::spmm(b, gmb::kOnly, a));
      B_TRY(breduce(b, R, b->al + (size_t)k * R, 0, 0, spmm_partial_rows(b, R)));   // alpha_k
    } else {
      a.y = Tb[0];
      a.ldy = R;
      a.gi = b->gamma[1];
      a.vk = vk_first;
      B_TRY(Launch<R>::spmm(b, gmb::kFirst, a));
      a.vk = Vk(k);
      for (int i = 2; i <= d; ++i) {
        a.x = Tb[(i - 2) & 1];
        a.ldx = R;
        a.y = Tb[(i - 1) & 1];
        a.gi = b->gamma[i];
        B_TRY(Launch<R>::spmm(b, gmb::kMid, a));
      }
      a.x = Tb[(d - 1) & 1];
      a.ldx = R;
      a.y = wnext;
      B_TRY(Launch<R>::spmm(b, gmb::kLast, a));
      B_TRY(breduce(b, R, b->al + (size_t)k * R, 0, 0, spmm_partial_rows(b, R)));   // alpha_k
    }
    gmb::BFinish f{};
    f.w = wnext;
    f.vk = Vk(k);
    f.wout = (ub && k + 1 < b->m) ? Vk(k + 1) : nullptr;
    f.al = b->al;
    f.be = b->be;
    f.k = k;
    f.nrows = b->n;
    f.chunk = b->chunk_sp;
    f.partials = b->partials;
    f.status = b->status;
    B_TRY(Launch<R>::finish(b, f, ub));
    B_TRY(breduce(b, R, b->be + (size_t)(k + 1) * R, 1, 1));          // beta_{k+1}
  }
  // shifted solves
  gmb::BSolve s{};
  s.al = b->al;
  s.be = b->be;
  s.R = R;
  s.m = b->m;
  s.ns = b->ns;
  s.wts1 = b->wts1;
  s.wts2 = b->wts2;
  s.etaxi = b->etaxi;
  s.wD = b->wD;
  s.wDU = b->wDU;
  s.wDU2 = b->wDU2;
  s.wX = b->wX;
  s.yr = b->yr;
  s.yi = b->yi;
  s.status = b->status;
  hipLaunchKernelGGL(gmb::k_bsolves<0>, dim3(cdivi((long long)R * b->ns, 64)), dim3(64), 0, b->stream, s);
  B_HIP(hipGetLastError());
  gmb::BCollapse cc{};
  cc.yr = b->yr;
  cc.yi = b->yi;
  cc.gsr = b->gsr;
  cc.gsi = b->gsi;
  cc.R = R;
  cc.m = b->m;
  cc.ns = b->ns;
  cc.nd = b->nd;
  cc.C = b->C;
  hipLaunchKernelGGL(gmb::k_bcollapse<0>, dim3(std::min(1024, cdivi((long long)b->nd * b->m * R, gm::kBlock))),
                     dim3(gm::kBlock), 0, b->stream, cc);
  B_HIP(hipGetLastError());
  if (ub) {
    hipLaunchKernelGGL(gmb::k_bscale_c<0>, dim3(cdivi((long long)b->nd * b->m * R, gm::kBlock)),
                       dim3(gm::kBlock), 0, b->stream, b->C, b->be, b->nd, b->m, R);
    B_HIP(hipGetLastError());
  }
  gmb::BVC v{};
  v.V = b->V;
  v.C = b->C;
  v.R = R;
  v.m = b->m;
  v.nd = b->nd;
  v.nrows = b->n;
  v.Z = d == 0 ? b->X + c0 : b->Z;
  v.ldz = d == 0 ? ldu : (size_t)R;
  v.accumulate = 0;
  switch (b->nd) {
#define B_VC(K) \
  case K: launch_bvc<K>(b, v); break;
    B_VC(1) B_VC(2) B_VC(3) B_VC(4) B_VC(5) B_VC(6) B_VC(7) B_VC(8) B_VC(9)
    B_VC(10) B_VC(11) B_VC(12) B_VC(13) B_VC(14) B_VC(15) B_VC(16) B_VC(17)
#undef B_VC
    default: return bfail(GM_EINVAL, "gamma_size too large");
  }
  B_HIP(hipGetLastError());
  const double* src = b->Z;
  size_t lds = R;
  for (int j = 1; j <= d; ++j) {
    gmb::BSpmm a = base;
    double* dst = (j == d) ? b->X + c0 : Tb[(j - 1) & 1];
    a.x = src;
    a.ldx = lds;
    a.y = dst;
    a.ldy = (j == d) ? ldu : (size_t)R;
    a.z = b->Z + (size_t)j * N * R;
    a.ldz = R;
    a.accumulate = 0;
    B_TRY(Launch<R>::spmm(b, gmb::kHorner, a));
    src = dst;
    lds = a.ldy;
  }
  return GM_OK;
}

template <int R>
gm_status enqueue_update_chunk(gm_batch* b, int c0, int prepare) {
  gmb::BUpdate u{};
  u.X = b->X + c0;
  u.Usol = b->Usol + c0;
  u.upd = b->K ? b->upd + c0 : nullptr;
  u.reaction = b->reaction;
  u.ode = b->ode ? b->ode + c0 / 2 : nullptr;
  u.ldh = (size_t)b->P;
  u.ms_tau_in = b->ms[0];
  u.ms_tau_out = b->ms[1];
  u.ms_tau_open = b->ms[2];
  u.ms_tau_close = b->ms[3];
  u.ms_v_gate = b->ms[4];
  u.mass = b->mass;
  u.dt = b->dt;
  u.clamp = b->clamp ? 1 : 0;
  u.x_zero = 0;
  u.prepare = prepare;
  u.ldu = (size_t)b->R;
  u.R = R;
  u.nrows = b->n;
  u.chunk = b->chunk_sp;
  u.be = b->be;   // overwritten by the next chunk's norm kernel anyway
  u.partials = b->partials;
  u.status = b->status;
  return Launch<R>::update(b, u);
}

gm_status dispatch_chunk(gm_batch* b, int c0) {
  switch (b->Rc) {
    case 2: return enqueue_chunk<2>(b, c0);
    case 4: return enqueue_chunk<4>(b, c0);
    case 8: return enqueue_chunk<8>(b, c0);
    case 16: return enqueue_chunk<16>(b, c0);
    case 32: return enqueue_chunk<32>(b, c0);
    case 64: return enqueue_chunk<64>(b, c0);
    case 128: return enqueue_chunk<128>(b, c0);
    case 256: return enqueue_chunk<256>(b, c0);
  }
  return bfail(GM_EINVAL, "unsupported chunk width %d", b->Rc);
}

gm_status dispatch_update(gm_batch* b, int c0, int prepare) {
  switch (b->Rc) {
    case 2: return enqueue_update_chunk<2>(b, c0, prepare);
    case 4: return enqueue_update_chunk<4>(b, c0, prepare);
    case 8: return enqueue_update_chunk<8>(b, c0, prepare);
    case 16: return enqueue_update_chunk<16>(b, c0, prepare);
    case 32: return enqueue_update_chunk<32>(b, c0, prepare);
    case 64: return enqueue_update_chunk<64>(b, c0, prepare);
    case 128: return enqueue_update_chunk<128>(b, c0, prepare);
    case 256: return enqueue_update_chunk<256>(b, c0, prepare);
  }
  return bfail(GM_EINVAL, "unsupported chunk width %d", b->Rc);
}

// Z = Q^T Usol (MFMA partials + fixed-order reduce), then Ud = Usol - Q Z
// and upd = Q f(Lambda) Z for all R columns.
gm_status enqueue_deflate(gm_batch* b) {
  gmb::BDeflQtU q{};
  q.Q = b->Q;
  q.U = b->Usol;
  q.K = b->K;
  q.R = b->R;
  q.nrows = b->n;
  q.chunk = b->chunk_df;
  q.KT = cdivi(b->K, 16);
  q.CT = cdivi(b->R, 16);
  q.partials = b->partials;
  gmb::launch_bdefl_qtu(q, b->G_df, b->stream);
  B_HIP(hipGetLastError());
  B_TRY(breduce(b, b->K * b->R, b->Zd, 0, 0, b->G_df));
  gmb::BDeflApply a{};
  a.Q = b->Q;
  a.Z = b->Zd;
  a.el = b->el;
  a.pl = b->pl;
  a.Usol = b->Usol;
  a.Ud = b->Ud;
  a.upd = b->upd;
  a.K = b->K;
  a.R = b->R;
  a.nrows = b->n;
  hipLaunchKernelGGL(gmb::k_bdefl_apply<0>, dim3(std::min(4096, cdivi(b->n, gm::kWaves)), cdivi(b->R, 64)),
                     dim3(gm::kBlock), 2 * (size_t)b->K * 64 * sizeof(double), b->stream, a);
  B_HIP(hipGetLastError());
  return GM_OK;
}

// A whole step: every chunk's f(A) products, then every chunk's update, then
// the deflation of the new starting vectors.
gm_status enqueue_bstep(gm_batch* b) {
  for (int c0 = 0; c0 < b->R; c0 += b->Rc) B_TRY(dispatch_chunk(b, c0));
  for (int c0 = 0; c0 < b->R; c0 += b->Rc) B_TRY(dispatch_update(b, c0, 0));
  if (b->K) B_TRY(enqueue_deflate(b));
  b->x_zero = true;
  return GM_OK;
}

// Grid of the column kernels: at most 1024 blocks (each SpMM block takes a run
// of the LDS-staged row blocks), so that the per-column block partials of the
// Lanczos sums stay few: on the 512^2 9-point operator at 32 columns a cap of
// 1024 measured 16% faster per step than 4096 blocks (whose k_reduce of 4096
// x 32 partials took 22 us twice per Lanczos step), 2048 and 512 in between.
void set_bschedule(gm_batch* b) {
  const int rows_per_pass = gm::kWaves * (b->Rc < 64 ? 64 / b->Rc : 1);
  b->G_sp = std::max(1, std::min(kBatchGrid, cdivi(b->n, rows_per_pass)));
  b->chunk_sp = cdivi(b->n, b->G_sp);
  b->G_sp = cdivi(b->n, b->chunk_sp);
}

gm_status alloc_chunk_buffers(gm_batch* b) {
  const size_t N = (size_t)b->n, Rc = (size_t)b->Rc;
  B_TRY(balloc(&b->W0, N * Rc));
  B_TRY(balloc(&b->W1, N * Rc));
  B_TRY(balloc(&b->T0, N * Rc));
  B_TRY(balloc(&b->T1, N * Rc));
  B_TRY(balloc(&b->V, N * Rc * b->m));
  B_TRY(balloc(&b->Z, b->nd > 1 ? N * Rc * b->nd : 1));
  B_TRY(balloc(&b->C, (size_t)b->nd * b->m * Rc));
  B_TRY(balloc(&b->al, (size_t)b->m * Rc));
  B_TRY(balloc(&b->be, (size_t)(b->m + 1) * Rc));
  const size_t ny = (size_t)b->m * Rc * b->ns;
  B_TRY(balloc(&b->yr, ny));
  B_TRY(balloc(&b->yi, ny));
  B_TRY(balloc(&b->wD, ny));
  B_TRY(balloc(&b->wDU, ny));
  B_TRY(balloc(&b->wDU2, ny));
  B_TRY(balloc(&b->wX, ny));
  B_TRY(balloc(&b->partials, std::max<size_t>(4096 * std::max<size_t>(Rc, 2),
                                              (size_t)b->G_df * b->K * b->R)));
  return GM_OK;
}

void free_chunk_buffers(gm_batch* b) {
  double** bufs[] = {&b->W0, &b->W1, &b->T0, &b->T1, &b->V, &b->Z, &b->C, &b->al, &b->be,
                     &b->yr, &b->yi, &b->partials};
  for (double** p : bufs) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  double2** cb[] = {&b->wD, &b->wDU, &b->wDU2, &b->wX};
  for (double2** p : cb) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
}

}  // namespace

extern "C" {

const char* gm_batch_last_error(void) { return g_berr.c_str(); }

gm_status gm_batch_create(const gm_batch_desc* d, gm_batch** out) {
  g_berr.clear();
  if (!out || !d) return bfail(GM_EINVAL, "null argument");
  *out = nullptr;
  if (d->n <= 0 || d->num_problems <= 0 || !d->U || !d->A_ptr || !d->A_values || !d->A_ind)
    return bfail(GM_EINVAL, "bad sizes or null arrays");
  if (d->subspace_size < 1 || d->gamma_size < 1 || d->gamma_size > gmb::kMaxND || !d->gamma)
    return bfail(GM_EINVAL, "bad subspace size or gamma");
  if (d->contour_size < 1 || !d->gamma_s_real || !d->gamma_s_imag || !d->wts1 || !d->wts2 ||
      !d->etaxi)
    return bfail(GM_EINVAL, "bad contour data");
  if (d->sparse_format != GM_SPARSE_CSR && d->sparse_format != GM_SPARSE_CSC)
    return bfail(GM_EINVAL, "bad sparse format");
  if (d->num_eigs < 0 || d->num_eigs > 64 ||
      (d->num_eigs > 0 && (!d->Q || !d->exp_lambda || !d->phi_lambda)))
    return bfail(GM_EINVAL, "num_eigs must be 0..64, with Q, exp_lambda and phi_lambda");
  if (d->reaction != GM_REACTION_FISHER && d->reaction != GM_REACTION_MS)
    return bfail(GM_EINVAL, "reaction must be GM_REACTION_FISHER or GM_REACTION_MS");
  if (d->reaction == GM_REACTION_MS &&
      !(d->ms.tau_in > 0 && d->ms.tau_out > 0 && d->ms.tau_open > 0 && d->ms.tau_close > 0))
    return bfail(GM_EINVAL, "Mitchell-Schaeffer time constants must be positive");
  if (d->A_ptr[0] != 0 || d->A_ptr[d->n] != d->nnz) return bfail(GM_EINVAL, "bad pointer array");
  for (int i = 0; i < d->n; ++i)
    if (d->A_ptr[i + 1] < d->A_ptr[i]) return bfail(GM_EINVAL, "pointer array not monotone");
  for (long long j = 0; j < d->nnz; ++j)
    if (d->A_ind[j] < 0 || d->A_ind[j] >= d->n) return bfail(GM_EINVAL, "index out of range");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= d->device || d->device < 0)
    return bfail(GM_EDEVICE, "no HIP device %d", d->device);
  gm_batch* b = new gm_batch();
  auto fail_clean = [&](gm_status st) {
    std::string keep = g_berr;
    bfree_all(b);
    delete b;
    g_berr = keep;
    return st;
  };
#define B_CREATE(x)                       \
  do {                                    \
    gm_status s_ = (x);                   \
    if (s_ != GM_OK) return fail_clean(s_); \
  } while (0)
  b->dev = d->device;
  if (hipSetDevice(b->dev) != hipSuccess) return fail_clean(bfail(GM_EDEVICE, "hipSetDevice"));
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return fail_clean(bfail(GM_EDEVICE, "stream create"));
  b->n = d->n;
  b->nnz = d->nnz;
  b->P = d->num_problems;
  b->R = 2 * b->P;
  b->m = d->subspace_size;
  b->nd = d->gamma_size;
  b->ns = d->contour_size;
  b->dt = d->dt;
  b->clamp = !(d->flags & GM_FLAG_NO_CLAMP);
  b->gamma.assign(d->gamma, d->gamma + b->nd);
  const size_t N = (size_t)b->n;
  // chunk width: a power of two dividing R, at most 256 and at most chunk_columns
  int rc = 2;
  while (rc * 2 <= std::min(256, b->R) && b->R % (rc * 2) == 0) rc *= 2;
  if (d->chunk_columns > 0)
    while (rc > d->chunk_columns && rc > 2) rc /= 2;
  if (b->R % rc != 0) return fail_clean(bfail(GM_EINVAL, "chunk width must divide 2*num_problems"));
  b->Rc = rc;
  // A -> CSR
  std::vector<int> rp(N + 1, 0), ci((size_t)b->nnz);
  std::vector<double> vv((size_t)b->nnz);
  if (d->sparse_format == GM_SPARSE_CSR) {
    std::copy(d->A_ptr, d->A_ptr + N + 1, rp.begin());
    std::copy(d->A_ind, d->A_ind + b->nnz, ci.begin());
    std::copy(d->A_values, d->A_values + b->nnz, vv.begin());
  } else {
    for (long long j = 0; j < b->nnz; ++j) rp[d->A_ind[j] + 1]++;
    for (size_t i = 0; i < N; ++i) rp[i + 1] += rp[i];
    std::vector<int> pos(rp.begin(), rp.end() - 1);
    for (int c = 0; c < b->n; ++c)
      for (int j = d->A_ptr[c]; j < d->A_ptr[c + 1]; ++j) {
        const int r = d->A_ind[j];
        ci[pos[r]] = c;
        vv[pos[r]] = d->A_values[j];
        pos[r]++;
      }
  }
  // GM_FLAG_LOCALITY_ORDER: the nodes renumbered by gm_locality_order; row q
  // of every device array is the caller's node perm[q], and a row keeps its
  // entries in the caller's order (so its sums are the same).
  if (d->flags & GM_FLAG_LOCALITY_ORDER) {
    b->perm.resize(N);
    if (gm_locality_order(b->n, rp.data(), ci.data(), kLocalityLeaf, b->perm.data()) != GM_OK)
      return fail_clean(bfail(GM_EINVAL, "locality order"));
    std::vector<int> iperm(N), rp2(N + 1), ci2(ci.size());
    std::vector<double> vv2(vv.size());
    for (size_t q = 0; q < N; ++q) iperm[b->perm[q]] = (int)q;
    rp2[0] = 0;
    for (size_t q = 0; q < N; ++q) {
      const int r = b->perm[q];
      int o = rp2[q];
      for (int j = rp[r]; j < rp[r + 1]; ++j, ++o) {
        ci2[o] = iperm[ci[j]];
        vv2[o] = vv[j];
      }
      rp2[q + 1] = o;
    }
    rp.swap(rp2);
    ci.swap(ci2);
    vv.swap(vv2);
  }
  auto src = [&](size_t q) { return b->perm.empty() ? q : (size_t)b->perm[q]; };
  hipStream_t st = b->stream;
  auto up = [&](double** p, const double* src, size_t cnt) -> gm_status {
    B_TRY(balloc(p, cnt));
    B_HIP(hipMemcpyAsync(*p, src, cnt * sizeof(double), hipMemcpyHostToDevice, st));
    return GM_OK;
  };
  B_CREATE(balloc(&b->rowptr, N + 1));
  B_CREATE(balloc(&b->col, (size_t)b->nnz));
  if (hipMemcpyAsync(b->rowptr, rp.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(b->col, ci.data(), ci.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess)
    return fail_clean(bfail(GM_EDEVICE, "upload A"));
  B_CREATE(up(&b->val, vv.data(), vv.size()));
  {
    std::vector<int> rb;
    gmb::bspmm_row_blocks(rp.data(), b->n, rb);
    b->nrowblk = (int)rb.size() - 1;
    B_CREATE(balloc(&b->rowblk, rb.size()));
    if (hipMemcpyAsync(b->rowblk, rb.data(), rb.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail_clean(bfail(GM_EDEVICE, "upload row blocks"));
    gmb::bspmm_xs_blocks(rp.data(), b->n, rb);
    b->nrowblk_xs = (int)rb.size() / 4;
    b->G_xs = gmb::bspmm_xs_grid(b->dev, b->nrowblk_xs);
    B_CREATE(balloc(&b->rowblk_xs, rb.size()));
    if (hipMemcpyAsync(b->rowblk_xs, rb.data(), rb.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail_clean(bfail(GM_EDEVICE, "upload row blocks"));
  }
  b->xs = (d->flags & GM_FLAG_COLUMN_SLABS) != 0;
  // U (n x 2P column-major) -> row-major Usol with dt on the g columns
  std::vector<double> u((size_t)N * b->R);
  for (int c = 0; c < b->R; ++c)
    for (size_t i = 0; i < N; ++i)
      u[i * b->R + c] = d->U[(size_t)c * N + src(i)] * ((c & 1) ? b->dt : 1.0);
  B_CREATE(up(&b->Usol, u.data(), u.size()));
  B_CREATE(balloc(&b->X, N * b->R));
  B_CREATE(up(&b->gsr, d->gamma_s_real, (size_t)b->nd * b->ns));
  B_CREATE(up(&b->gsi, d->gamma_s_imag, (size_t)b->nd * b->ns));
  B_CREATE(up(&b->wts1, d->wts1, 2 * (size_t)b->ns));
  B_CREATE(up(&b->wts2, d->wts2, 2 * (size_t)b->ns));
  B_CREATE(up(&b->etaxi, d->etaxi, 2 * (size_t)b->ns));
  std::vector<double> massv(N, 1.0);
  if (d->mass)
    for (size_t i = 0; i < N; ++i) massv[i] = d->mass[src(i)];
  B_CREATE(up(&b->mass, massv.data(), N));
  b->reaction = d->reaction;
  if (b->reaction == GM_REACTION_MS) {
    b->ms[0] = d->ms.tau_in;
    b->ms[1] = d->ms.tau_out;
    b->ms[2] = d->ms.tau_open;
    b->ms[3] = d->ms.tau_close;
    b->ms[4] = d->ms.v_gate;
    // gate h, n x P column-major -> [N][P]
    std::vector<double> h((size_t)N * b->P, 1.0);
    if (d->ode)
      for (int q = 0; q < b->P; ++q)
        for (size_t i = 0; i < N; ++i) h[i * b->P + q] = d->ode[(size_t)q * N + src(i)];
    B_CREATE(up(&b->ode, h.data(), h.size()));
    if (hipStreamSynchronize(st) != hipSuccess) return fail_clean(bfail(GM_EDEVICE, "sync"));
  }
  b->K = d->num_eigs;
  b->G_df = std::max(1, std::min(256, cdivi(b->n, 1024)));
  b->chunk_df = cdivi(b->n, b->G_df);
  if (b->K) {
    // Q n x K column-major -> [N][K]
    std::vector<double> qr(N * b->K);
    for (int k = 0; k < b->K; ++k)
      for (size_t i = 0; i < N; ++i) qr[i * b->K + k] = d->Q[(size_t)k * N + src(i)];
    B_CREATE(up(&b->Q, qr.data(), qr.size()));
    B_CREATE(up(&b->el, d->exp_lambda, (size_t)b->K));
    B_CREATE(up(&b->pl, d->phi_lambda, (size_t)b->K));
    B_CREATE(balloc(&b->Zd, (size_t)b->K * b->R));
    B_CREATE(balloc(&b->Ud, N * b->R));
    B_CREATE(balloc(&b->upd, N * b->R));
    if (hipStreamSynchronize(st) != hipSuccess) return fail_clean(bfail(GM_EDEVICE, "sync"));
  }
  B_CREATE(balloc(&b->stage, N * b->R));
  B_CREATE(balloc(&b->status, 1));
  if (hipMemsetAsync(b->status, 0, sizeof(unsigned), st) != hipSuccess)
    return fail_clean(bfail(GM_EDEVICE, "memset"));
  // chunk buffers: halve the chunk until they fit
  for (;;) {
    set_bschedule(b);
    gm_status s = alloc_chunk_buffers(b);
    if (s == GM_OK) break;
    free_chunk_buffers(b);
    (void)hipGetLastError();
    if (s != GM_ENOMEM || b->Rc <= 2) return fail_clean(s);
    b->Rc /= 2;
  }
  // deflate the initial vectors (lanczos_impl.cu:84-89)
  if (b->K) B_CREATE(enqueue_deflate(b));
  if (hipStreamSynchronize(st) != hipSuccess) return fail_clean(bfail(GM_EDEVICE, "sync"));
  if (d->flags & GM_FLAG_AUTOTUNE) {
    // Time one chunk of the step for each chunk width that the allocated
    // buffers hold (every width <= Rc) and keep the cheapest per column.  The
    // runs overwrite only X and scratch; Usol is untouched.
    // From 32 columns up both SpMM forms (whole rows, column slabs) are timed.
    const int rc_max = b->Rc;
    double best = 1e300;
    int best_rc = rc_max;
    bool best_xs = false;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
      return fail_clean(bfail(GM_EDEVICE, "event create"));
    for (int rc = rc_max; rc >= 8 || rc == rc_max; rc /= 2) {
      b->Rc = rc;
      set_bschedule(b);
      for (int xs = 0; xs <= (rc >= 32 ? 1 : 0); ++xs) {
        b->xs = xs != 0;
        gm_status s1 = dispatch_chunk(b, 0);                       // warm-up
        if (s1 != GM_OK) return fail_clean(s1);
        // the faster of two timed runs (the widths' costs per column lie
        // within a few per cent of each other on large operators)
        double ms_best = 1e300;
        for (int rep = 0; rep < 2; ++rep) {
          (void)hipEventRecord(e0, st);
          gm_status s2 = dispatch_chunk(b, 0);
          (void)hipEventRecord(e1, st);
          if (s2 != GM_OK || hipEventSynchronize(e1) != hipSuccess)
            return fail_clean(s2 != GM_OK ? s2 : bfail(GM_EDEVICE, "autotune"));
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, e0, e1);
          ms_best = std::min(ms_best, (double)ms);
        }
        const double per_col = ms_best / rc;
        if (per_col < best) {
          best = per_col;
          best_rc = rc;
          best_xs = b->xs;
        }
      }
      if (rc <= 2) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->Rc = best_rc;
    b->xs = best_xs;
    set_bschedule(b);
    if (hipMemsetAsync(b->status, 0, sizeof(unsigned), st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail_clean(bfail(GM_EDEVICE, "autotune sync"));
  }
  {
    std::lock_guard<std::mutex> lk(g_bmu);
    g_breg.insert(b);
  }
  *out = b;
  return GM_OK;
#undef B_CREATE
}

gm_status gm_batch_delete(gm_batch* b) {
  {
    std::lock_guard<std::mutex> lk(g_bmu);
    if (!b || !g_breg.count(b)) return bfail(GM_EBADHANDLE, "invalid batch handle");
    g_breg.erase(b);
  }
  (void)hipSetDevice(b->dev);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  bfree_all(b);
  delete b;
  return GM_OK;
}

gm_status gm_batch_step(gm_batch* b, int32_t nsteps, int32_t use_graph) {
  if (!blive(b)) return bfail(GM_EBADHANDLE, "invalid batch handle");
  B_HIP(hipSetDevice(b->dev));
  if (nsteps < 0) return bfail(GM_EINVAL, "nsteps < 0");
  if (!use_graph) {
    for (int i = 0; i < nsteps; ++i) B_TRY(enqueue_bstep(b));
    return GM_OK;
  }
  if (!b->graph) {
    hipGraph_t g = nullptr;
    B_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    gm_status s = enqueue_bstep(b);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (s != GM_OK) {
      if (g) (void)hipGraphDestroy(g);
      return s;
    }
    if (e != hipSuccess) return bfail(GM_EDEVICE, "capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      b->graph = nullptr;
      return bfail(GM_EDEVICE, "instantiate: %s", hipGetErrorString(e));
    }
  }
  for (int i = 0; i < nsteps; ++i) B_HIP(hipGraphLaunch(b->graph, b->stream));
  return GM_OK;
}

gm_status gm_batch_get_U(gm_batch* b, double* dst) {
  if (!blive(b)) return bfail(GM_EBADHANDLE, "invalid batch handle");
  if (!dst) return bfail(GM_EINVAL, "null destination");
  B_HIP(hipSetDevice(b->dev));
  const size_t N = (size_t)b->n;
  std::vector<double> tmp(N * b->R);
  B_HIP(hipMemcpyAsync(tmp.data(), b->Usol, tmp.size() * sizeof(double), hipMemcpyDeviceToHost,
                       b->stream));
  B_HIP(hipStreamSynchronize(b->stream));
  for (int c = 0; c < b->R; ++c)
    for (size_t i = 0; i < N; ++i)
      dst[(size_t)c * N + (b->perm.empty() ? i : (size_t)b->perm[i])] = tmp[i * b->R + c];
  return GM_OK;
}

gm_status gm_batch_synchronize(gm_batch* b) {
  if (!blive(b)) return bfail(GM_EBADHANDLE, "invalid batch handle");
  B_HIP(hipSetDevice(b->dev));
  unsigned stv = 0;
  B_HIP(hipMemcpyAsync(&stv, b->status, sizeof(unsigned), hipMemcpyDeviceToHost, b->stream));
  B_HIP(hipStreamSynchronize(b->stream));
  if (stv & 6u) return bfail(GM_EBREAKDOWN, "batched Lanczos breakdown (status %u)", stv);
  return GM_OK;
}

gm_status gm_batch_info(gm_batch* b, int32_t* num_problems, int32_t* chunk_columns) {
  if (!blive(b)) return bfail(GM_EBADHANDLE, "invalid batch handle");
  if (num_problems) *num_problems = b->P;
  if (chunk_columns) *chunk_columns = b->Rc;
  return GM_OK;
}

gm_status gm_batch_get_column_slabs(gm_batch* b, int32_t* on) {
  if (!blive(b)) return bfail(GM_EBADHANDLE, "invalid batch handle");
  if (!on) return bfail(GM_EINVAL, "null argument");
  *on = (b->Rc >= 32 && b->xs) ? 1 : 0;
  return GM_OK;
}

gm_status gm_batch_get_ode(gm_batch* b, double* dst) {
  if (!blive(b)) return bfail(GM_EBADHANDLE, "invalid batch handle");
  if (!dst) return bfail(GM_EINVAL, "null destination");
  if (!b->ode) return bfail(GM_ESTATE, "the batch has no gate state (reaction is not MS)");
  B_HIP(hipSetDevice(b->dev));
  const size_t N = (size_t)b->n;
  std::vector<double> tmp(N * b->P);
  B_HIP(hipMemcpyAsync(tmp.data(), b->ode, tmp.size() * sizeof(double), hipMemcpyDeviceToHost,
                       b->stream));
  B_HIP(hipStreamSynchronize(b->stream));
  for (int q = 0; q < b->P; ++q)
    for (size_t i = 0; i < N; ++i)
      dst[(size_t)q * N + (b->perm.empty() ? i : (size_t)b->perm[i])] = tmp[i * b->P + q];
  return GM_OK;
}

gm_status gm_batch_get_stream(gm_batch* b, void** stream) {
  if (!blive(b)) return bfail(GM_EBADHANDLE, "invalid batch handle");
  if (!stream) return bfail(GM_EINVAL, "null pointer");
  *stream = b->stream;
  return GM_OK;
}

}  // extern "C"
