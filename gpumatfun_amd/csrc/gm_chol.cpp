// gm_chol.cpp -- sparse Cholesky factorisation of A - sigma I and the
// deflation eigensolver built on it (host code, threads; no GPU needed).
//
// It replaces the factorisation inside eigs(K, M, k, 'sm') of
// process_nonuniform.m:38,106 (the smallest eigenpairs by shift-invert) and
// eigs(K, M, 1, 'lm') of :40 (lambda_max) for meshes whose lambda_max /
// lambda_K is too large for a polynomial filter on the GPU (sliver elements).
// A = M^(-1/2) K M^(-1/2) with the lumped M, so the orthonormal eigenvectors of
// A are the M-orthonormal ones of (K, M) scaled by M^(1/2) (:107).
//
//   ordering   nested dissection on breadth-first level structures (George's
//              automatic nested dissection): the separator is the level that
//              halves the part, trimmed to the nodes that touch the next
//              level; leaves of at most kLeaf nodes in reverse BFS order
//   symbolic   elimination tree (Liu's path compression), postorder, column
//              counts by row subtrees, fundamental supernodes merged with
//              their parents while few explicit zeros are added
//   numeric    multifrontal: each supernode's front is assembled from A and
//              its children's update matrices (extend-add) and factored by a
//              left-looking blocked Cholesky, then its Schur complement
//              (SYRK) becomes the update matrix for the parent; the subtrees
//              below a fixed cut run on separate threads, the fronts above it
//              split their dense products over the threads
//   solve      supernodal forward and back substitution on many right-hand
//              sides held row-major (n x nrhs, a row's values contiguous)
//   eigen      block Lanczos on (A - sigma I)^(-1) with full
//              reorthogonalisation (two classical Gram-Schmidt passes) and
//              explicit restarts from the Ritz block; lambda = sigma + 1/theta
//
// Every reduction runs in an order fixed by the matrix alone (the subtree cut
// and the row chunks of the tall products do not depend on the thread count),
// so the results are bitwise independent of the number of threads.
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gpumatfun.h"

struct gm_chol;

namespace {

using i64 = int64_t;

std::mutex g_err_mu;
std::string g_err;

void set_err(const std::string& s) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err = s;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// threads <= 0: the CPUs this process may run on (its affinity mask), capped
// by OMP_NUM_THREADS when that is set and by 64.
int resolve_threads(int t) {
  if (t > 0) return std::min(t, 256);
  int n = 0;
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
  if (n <= 0) n = (int)std::thread::hardware_concurrency();
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const int k = std::atoi(e);
    if (k > 0) n = std::min(n > 0 ? n : k, k);
  }
  return std::max(1, std::min(n, 64));
}

// body(i, tid) for i in [0, count) on up to nthreads threads, dynamic order.
// Returns false if a body threw (e.g. bad_alloc).
template <class F>
bool parallel_for(int nthreads, i64 count, F&& body) {
  if (count <= 0) return true;
  const int nt = (int)std::min<i64>(std::max(nthreads, 1), count);
  std::atomic<i64> next(0);
  std::atomic<bool> ok(true);
  auto work = [&](int tid) {
    try {
      for (;;) {
        const i64 i = next.fetch_add(1, std::memory_order_relaxed);
        if (i >= count || !ok.load(std::memory_order_relaxed)) break;
        body(i, tid);
      }
    } catch (...) {
      ok = false;
    }
  };
  std::vector<std::thread> th;
  try {
    for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
  } catch (...) {
    // fewer threads than asked: the remaining ones share the work
  }
  work(0);
  for (auto& x : th) x.join();
  return ok.load();
}

// ---------------------------------------------------------------------------
// Dense kernels (column-major).  C -= A * op(B): A is m x k (lda); op(B) is
// k x n, given as B n x k (ldb, bt = true: B^T) or B k x n (ldb, bt = false).
// Blocked as in BLIS: kNC columns, kKC depth, kMC rows; packed panels of kMR
// rows and kNR columns; the micro-kernel sums each kKC slice in p order, so an
// element's value depends only on k (not on how m and n are split).
// ---------------------------------------------------------------------------
constexpr int kMR = 16, kNR = 6, kKC = 256, kMC = 96, kNC = 1536;

// Micro-kernels: out (kMR x kNR, column-major) = sum over p < kc of the packed
// A column a[p*kMR ..] times the packed B row b[p*kNR ..], p in order.  The
// vector forms keep the 16 x 6 tile in registers (AVX-512: 12 zmm; AVX2: two
// 8-row halves of 12 ymm each); every element is one FMA chain over p.
typedef double v8d __attribute__((vector_size(64)));
typedef double v4d __attribute__((vector_size(32)));

__attribute__((target("avx512f,fma"))) void micro_avx512(int kc, const double* __restrict a,
                                                         const double* __restrict b,
                                                         double* __restrict out) {
  v8d c[kNR][2];
  for (int j = 0; j < kNR; ++j) c[j][0] = c[j][1] = (v8d){0, 0, 0, 0, 0, 0, 0, 0};
  for (int p = 0; p < kc; ++p) {
    v8d a0, a1;
    std::memcpy(&a0, a + (size_t)p * kMR, sizeof a0);
    std::memcpy(&a1, a + (size_t)p * kMR + 8, sizeof a1);
    const double* bp = b + (size_t)p * kNR;
#pragma GCC unroll 6
    for (int j = 0; j < kNR; ++j) {
      const double s = bp[j];
      const v8d bj = {s, s, s, s, s, s, s, s};
      c[j][0] += a0 * bj;
      c[j][1] += a1 * bj;
    }
  }
  for (int j = 0; j < kNR; ++j) {
    std::memcpy(out + j * kMR, &c[j][0], sizeof(v8d));
    std::memcpy(out + j * kMR + 8, &c[j][1], sizeof(v8d));
  }
}

__attribute__((target("avx2,fma"))) void micro_avx2(int kc, const double* __restrict a,
                                                   const double* __restrict b,
                                                   double* __restrict out) {
  for (int h = 0; h < 2; ++h) {
    v4d c[kNR][2];
    for (int j = 0; j < kNR; ++j) c[j][0] = c[j][1] = (v4d){0, 0, 0, 0};
    for (int p = 0; p < kc; ++p) {
      v4d a0, a1;
      std::memcpy(&a0, a + (size_t)p * kMR + 8 * h, sizeof a0);
      std::memcpy(&a1, a + (size_t)p * kMR + 8 * h + 4, sizeof a1);
      const double* bp = b + (size_t)p * kNR;
#pragma GCC unroll 6
      for (int j = 0; j < kNR; ++j) {
        const double s = bp[j];
        const v4d bj = {s, s, s, s};
        c[j][0] += a0 * bj;
        c[j][1] += a1 * bj;
      }
    }
    for (int j = 0; j < kNR; ++j) {
      std::memcpy(out + j * kMR + 8 * h, &c[j][0], sizeof(v4d));
      std::memcpy(out + j * kMR + 8 * h + 4, &c[j][1], sizeof(v4d));
    }
  }
}

void micro_generic(int kc, const double* __restrict a, const double* __restrict b,
                   double* __restrict out) {
  double acc[kNR * kMR];
  for (int t = 0; t < kNR * kMR; ++t) acc[t] = 0.0;
  for (int p = 0; p < kc; ++p) {
    const double* ap = a + (size_t)p * kMR;
    const double* bp = b + (size_t)p * kNR;
    for (int j = 0; j < kNR; ++j)
      for (int i = 0; i < kMR; ++i) acc[j * kMR + i] += ap[i] * bp[j];
  }
  for (int t = 0; t < kNR * kMR; ++t) out[t] = acc[t];
}

using micro_fn = void (*)(int, const double*, const double*, double*);

// GM_CHOL_ISA=avx2 or generic forces a narrower kernel (for testing).
micro_fn pick_micro() {
  __builtin_cpu_init();
  if (const char* isa = std::getenv("GM_CHOL_ISA")) {
    if (!std::strcmp(isa, "generic")) return micro_generic;
    if (!std::strcmp(isa, "avx2") && __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma"))
      return micro_avx2;
  }
  if (__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("fma")) return micro_avx512;
  if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) return micro_avx2;
  return micro_generic;
}

micro_fn micro() {
  static const micro_fn f = pick_micro();
  return f;
}

thread_local std::vector<double> tl_pa, tl_pb;

void gemm_sub(int m, int n, int k, const double* A, i64 lda, const double* B, i64 ldb, bool bt,
              double* C, i64 ldc) {
  if (m <= 0 || n <= 0 || k <= 0) return;
  const micro_fn mk = micro();
  double out[kNR * kMR];
  for (int j0 = 0; j0 < n; j0 += kNC) {
    const int nc = std::min(kNC, n - j0);
    const int npan = (nc + kNR - 1) / kNR;
    for (int p0 = 0; p0 < k; p0 += kKC) {
      const int kc = std::min(kKC, k - p0);
      tl_pb.resize((size_t)npan * kc * kNR);
      double* pb = tl_pb.data();
      for (int jp = 0; jp < npan; ++jp) {
        double* dst = pb + (size_t)jp * kc * kNR;
        for (int jj = 0; jj < kNR; ++jj) {
          const int j = j0 + jp * kNR + jj;
          if (j < j0 + nc) {
            if (bt) {
              const double* src = B + j + (i64)p0 * ldb;
              for (int p = 0; p < kc; ++p) dst[(size_t)p * kNR + jj] = src[(i64)p * ldb];
            } else {
              const double* src = B + p0 + (i64)j * ldb;
              for (int p = 0; p < kc; ++p) dst[(size_t)p * kNR + jj] = src[p];
            }
          } else {
            for (int p = 0; p < kc; ++p) dst[(size_t)p * kNR + jj] = 0.0;
          }
        }
      }
      for (int i0 = 0; i0 < m; i0 += kMC) {
        const int mc = std::min(kMC, m - i0);
        const int mpan = (mc + kMR - 1) / kMR;
        tl_pa.resize((size_t)mpan * kc * kMR);
        double* pa = tl_pa.data();
        for (int ip = 0; ip < mpan; ++ip) {
          double* dst = pa + (size_t)ip * kc * kMR;
          const int ib = i0 + ip * kMR;
          const int iv = std::min(kMR, i0 + mc - ib);
          for (int p = 0; p < kc; ++p) {
            const double* src = A + ib + (i64)(p0 + p) * lda;
            double* d = dst + (size_t)p * kMR;
            int ii = 0;
            for (; ii < iv; ++ii) d[ii] = src[ii];
            for (; ii < kMR; ++ii) d[ii] = 0.0;
          }
        }
        for (int jp = 0; jp < npan; ++jp) {
          const int jb = j0 + jp * kNR;
          const int jv = std::min(kNR, j0 + nc - jb);
          for (int ip = 0; ip < mpan; ++ip) {
            const int ib = i0 + ip * kMR;
            const int iv = std::min(kMR, i0 + mc - ib);
            mk(kc, pa + (size_t)ip * kc * kMR, pb + (size_t)jp * kc * kNR, out);
            for (int jj = 0; jj < jv; ++jj) {
              double* c = C + ib + (i64)(jb + jj) * ldc;
              const double* o = out + jj * kMR;
              for (int ii = 0; ii < iv; ++ii) c[ii] -= o[ii];
            }
          }
        }
      }
    }
  }
}

// gemm_sub split over threads by column chunks (or row chunks when n is small).
void gemm_sub_par(int threads, int m, int n, int k, const double* A, i64 lda, const double* B,
                  i64 ldb, bool bt, double* C, i64 ldc) {
  const double flops = 2.0 * m * n * (double)k;
  if (threads <= 1 || flops < 4e6) {
    gemm_sub(m, n, k, A, lda, B, ldb, bt, C, ldc);
    return;
  }
  constexpr int kColChunk = 16 * kNR, kRowChunk = 16 * kMR;
  if (n >= 2 * kColChunk) {
    const int nt = (n + kColChunk - 1) / kColChunk;
    parallel_for(threads, nt, [&](i64 t, int) {
      const int j0 = (int)t * kColChunk, nn = std::min(kColChunk, n - j0);
      const double* Bj = bt ? B + j0 : B + (i64)j0 * ldb;
      gemm_sub(m, nn, k, A, lda, Bj, ldb, bt, C + (i64)j0 * ldc, ldc);
    });
  } else {
    const int nt = (m + kRowChunk - 1) / kRowChunk;
    parallel_for(threads, nt, [&](i64 t, int) {
      const int i0 = (int)t * kRowChunk, mm = std::min(kRowChunk, m - i0);
      gemm_sub(mm, n, k, A + i0, lda, B, ldb, bt, C + i0, ldc);
    });
  }
}

constexpr int kNB = 64;   // column block of the front factorisation

// Partial Cholesky of a front.  Lb (nr x nc, ld nr) holds the front's first nc
// columns (rows 0..nc-1 are the pivots); U (mu x mu, ld mu, lower) the rest of
// the front.  On exit Lb = [L11; L21] and U -= L21 L21^T.  Returns the local
// column of a non-positive pivot, or -1.
int factor_front(int nr, int nc, double* Lb, double* U, int threads) {
  const int mu = nr - nc;
  for (int j0 = 0; j0 < nc; j0 += kNB) {
    const int jb = std::min(kNB, nc - j0);
    // left-looking update of the block column by the finished columns
    if (j0 > 0)
      gemm_sub_par(threads, nr - j0, jb, j0, Lb + j0, nr, Lb + j0, nr, true,
                   Lb + j0 + (i64)j0 * nr, nr);
    // unblocked Cholesky of the diagonal block
    for (int j = 0; j < jb; ++j) {
      double* cj = Lb + (i64)(j0 + j) * nr;
      for (int q = 0; q < j; ++q) {
        const double* cq = Lb + (i64)(j0 + q) * nr;
        const double l = cq[j0 + j];
        for (int i = j0 + j; i < j0 + jb; ++i) cj[i] -= l * cq[i];
      }
      const double d = cj[j0 + j];
      if (!(d > 0.0) || !std::isfinite(d)) return j0 + j;
      const double r = std::sqrt(d);
      cj[j0 + j] = r;
      for (int i = j0 + j + 1; i < j0 + jb; ++i) cj[i] /= r;
    }
    // the rows below: B <- B L_dd^(-T), in row chunks
    const int r0 = j0 + jb;
    const int rows = nr - r0;
    if (rows > 0) {
      constexpr int kChunk = 256;
      const int nchunk = (rows + kChunk - 1) / kChunk;
      auto solve_rows = [&](i64 t, int) {
        const int a0 = r0 + (int)t * kChunk, a1 = std::min(nr, a0 + kChunk);
        for (int j = 0; j < jb; ++j) {
          double* cj = Lb + (i64)(j0 + j) * nr;
          for (int q = 0; q < j; ++q) {
            const double* cq = Lb + (i64)(j0 + q) * nr;
            const double l = cq[j0 + j];
            for (int i = a0; i < a1; ++i) cj[i] -= l * cq[i];
          }
          const double r = cj[j0 + j];
          for (int i = a0; i < a1; ++i) cj[i] /= r;
        }
      };
      if (threads > 1 && (double)rows * jb * jb > 2e6)
        parallel_for(threads, nchunk, solve_rows);
      else
        for (int t = 0; t < nchunk; ++t) solve_rows(t, 0);
    }
  }
  if (mu > 0) {
    // U -= L21 L21^T, lower part, by column blocks of U
    const double* L21 = Lb + nc;
    constexpr int kSB = 64;
    const int nb = (mu + kSB - 1) / kSB;
    auto syrk_block = [&](i64 t, int) {
      const int c0 = (int)t * kSB, cw = std::min(kSB, mu - c0);
      gemm_sub(mu - c0, cw, nc, L21 + c0, nr, L21 + c0, nr, true, U + c0 + (i64)c0 * mu, mu);
    };
    if (threads > 1 && (double)mu * mu * nc > 4e6)
      parallel_for(threads, nb, syrk_block);
    else
      for (int t = 0; t < nb; ++t) syrk_block(t, 0);
  }
  return -1;
}

// ---------------------------------------------------------------------------
// Nested dissection.  perm[q] = old index placed at q.
// ---------------------------------------------------------------------------
constexpr int kLeaf = 64;

struct NDTask {
  std::vector<int> nodes;
  int lo;
};

// Independent parts are dissected concurrently.  A part's nodes carry its id
// in owner[] (atomic: a breadth-first search reads the marks of neighbours in
// other parts while those are being relabelled); the positions a part fills in
// perm are fixed by its offset lo, so the order does not depend on timing.
class Dissector {
 public:
  Dissector(int n, const int* ptr, const int* ind, std::vector<int>& perm)
      : n_(n), ptr_(ptr), ind_(ind), perm_(perm), owner_(n), level_(n, -1), deg_(n, 0) {
    perm_.assign(n, -1);
    for (int v = 0; v < n; ++v) {
      owner_[v].store(1, std::memory_order_relaxed);
      for (int j = ptr[v]; j < ptr[v + 1]; ++j)
        if (ind[j] != v) ++deg_[v];
    }
  }

  void run(int threads) {
    NDTask t;
    t.nodes.resize(n_);
    for (int v = 0; v < n_; ++v) t.nodes[v] = v;
    t.lo = 0;
    next_id_.store(2);
    queue_.push_back(std::move(t));
    const int nt = std::max(1, threads);
    std::vector<std::thread> th;
    try {
      for (int i = 1; i < nt; ++i) th.emplace_back([this] { worker(); });
    } catch (...) {
    }
    worker();
    for (auto& x : th) x.join();
    if (failed_) throw std::bad_alloc();
  }

 private:
  struct Buf {
    std::vector<int> queue, lsize;
  };

  int owner(int v) const { return owner_[v].load(std::memory_order_relaxed); }
  void set_owner(const std::vector<int>& nodes, int id) {
    for (int v : nodes) owner_[v].store(id, std::memory_order_relaxed);
  }

  void worker() {
    Buf b;
    b.queue.resize(n_);
    for (;;) {
      NDTask t;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !queue_.empty() || busy_ == 0; });
        if (queue_.empty()) {
          cv_.notify_all();
          return;
        }
        t = std::move(queue_.back());
        queue_.pop_back();
        ++busy_;
      }
      std::vector<NDTask> out;
      try {
        process(t, b, out);
      } catch (...) {
        failed_ = true;
        out.clear();
      }
      {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& c : out) queue_.push_back(std::move(c));
        --busy_;
      }
      cv_.notify_all();
    }
  }

  // BFS over the part `id` from root; fills b.queue[0..cnt), level_ and
  // b.lsize; returns cnt.
  int bfs(int id, int root, Buf& b) {
    int head = 0, tail = 0;
    b.queue[tail++] = root;
    level_[root] = 0;
    b.lsize.assign(1, 1);
    while (head < tail) {
      const int v = b.queue[head++];
      const int lv = level_[v];
      for (int j = ptr_[v]; j < ptr_[v + 1]; ++j) {
        const int w = ind_[j];
        if (owner(w) != id || level_[w] >= 0) continue;
        level_[w] = lv + 1;
        if ((int)b.lsize.size() <= lv + 1) b.lsize.push_back(0);
        ++b.lsize[lv + 1];
        b.queue[tail++] = w;
      }
    }
    return tail;
  }

  void reset(const NDTask& t) {
    for (int v : t.nodes) level_[v] = -1;
  }

  // reverse breadth-first order over each connected piece
  void place_leaf(const NDTask& t, int id, Buf& b) {
    reset(t);
    int pos = t.lo + (int)t.nodes.size();
    for (int v : t.nodes) {
      if (level_[v] >= 0) continue;
      const int cnt = bfs(id, v, b);
      for (int q = 0; q < cnt; ++q) perm_[--pos] = b.queue[q];
    }
  }

  void push_part(std::vector<NDTask>& out, NDTask&& part) {
    if (part.nodes.empty()) return;
    set_owner(part.nodes, next_id_.fetch_add(1));
    out.push_back(std::move(part));
  }

  void process(NDTask& t, Buf& b, std::vector<NDTask>& out) {
    const int id = owner(t.nodes[0]);
    const int m = (int)t.nodes.size();
    if (m <= kLeaf) {
      place_leaf(t, id, b);
      return;
    }
    reset(t);
    const int cnt = bfs(id, t.nodes[0], b);
    if (cnt < m) {
      // disconnected: the reached component and the rest become two parts
      NDTask a, c;
      a.nodes.assign(b.queue.begin(), b.queue.begin() + cnt);
      c.nodes.reserve(m - cnt);
      for (int v : t.nodes)
        if (level_[v] < 0) c.nodes.push_back(v);
      a.lo = t.lo;
      c.lo = t.lo + cnt;
      push_part(out, std::move(a));
      push_part(out, std::move(c));
      return;
    }
    // pseudo-peripheral root: restart from a least-degree node of the last
    // level while the level structure gets deeper
    int root = t.nodes[0];
    int height = (int)b.lsize.size();
    for (int it = 0; it < 8; ++it) {
      int best = -1;
      for (int q = cnt - b.lsize.back(); q < cnt; ++q) {
        const int v = b.queue[q];
        if (best < 0 || deg_[v] < deg_[best]) best = v;
      }
      reset(t);
      bfs(id, best, b);
      const int h2 = (int)b.lsize.size();
      if (h2 > height) {
        root = best;
        height = h2;
      } else {
        reset(t);
        bfs(id, root, b);
        break;
      }
    }
    if (height < 3) {
      place_leaf(t, id, b);
      return;
    }
    // separator level: the smallest level that leaves 40%..60% of the other
    // nodes on each side (the median level if none does), never the first or
    // the last
    std::vector<int> cum(height + 1, 0);
    for (int l = 0; l < height; ++l) cum[l + 1] = cum[l] + b.lsize[l];
    int kbest = -1;
    for (int l = 1; l < height - 1; ++l) {
      const double below = cum[l], above = m - cum[l + 1];
      if (std::min(below, above) >= 0.4 * (below + above))
        if (kbest < 0 || b.lsize[l] < b.lsize[kbest]) kbest = l;
    }
    if (kbest < 0) {
      kbest = 1;
      while (kbest < height - 2 && cum[kbest + 1] < m / 2) ++kbest;
    }
    // the separator is that level's nodes that touch the next level; the
    // rest of the level joins the near side
    NDTask a, c;
    std::vector<int> sep;
    for (int v : t.nodes) {
      const int lv = level_[v];
      if (lv < kbest) {
        a.nodes.push_back(v);
      } else if (lv > kbest) {
        c.nodes.push_back(v);
      } else {
        bool touches = false;
        for (int j = ptr_[v]; j < ptr_[v + 1] && !touches; ++j) {
          const int w = ind_[j];
          touches = owner(w) == id && level_[w] == kbest + 1;
        }
        if (touches)
          sep.push_back(v);
        else
          a.nodes.push_back(v);
      }
    }
    const int ns = (int)sep.size();
    for (int q = 0; q < ns; ++q) perm_[t.lo + m - ns + q] = sep[q];
    set_owner(sep, 0);
    a.lo = t.lo;
    c.lo = t.lo + (int)a.nodes.size();
    push_part(out, std::move(a));
    push_part(out, std::move(c));
  }

  const int n_;
  const int* ptr_;
  const int* ind_;
  std::vector<int>& perm_;
  std::vector<std::atomic<int>> owner_;
  std::vector<int> level_, deg_;
  std::atomic<int> next_id_{2};
  std::mutex mu_;
  std::condition_variable cv_;
  std::vector<NDTask> queue_;
  int busy_ = 0;
  std::atomic<bool> failed_{false};
};

void nested_dissection(int n, const int* ptr, const int* ind, std::vector<int>& perm,
                       int threads) {
  Dissector d(n, ptr, ind, perm);
  d.run(threads);
}

}  // namespace

// ---------------------------------------------------------------------------
// The factor.
// ---------------------------------------------------------------------------
struct gm_chol {
  int n = 0;
  int threads = 1;
  double sigma = 0.0;
  std::vector<int> perm, iperm;      // perm[new] = old
  int nsuper = 0;
  std::vector<int> sfirst;           // nsuper + 1: supernode s has columns sfirst[s]..sfirst[s+1]-1
  std::vector<int> sparent;          // parent supernode or -1
  std::vector<i64> rptr;             // nsuper + 1: rows of s are rows[rptr[s] .. rptr[s+1])
  std::vector<int> rows;             // sorted; the first ncol are the supernode's columns
  std::vector<i64> lptr;             // nsuper + 1: L block of s at lval[lptr[s]], nr x nc, ld nr
  std::vector<double> lval;
  // fixed cut of the supernodal tree: subtree t covers supernodes
  // sub_lo[t] .. sub_root[t] (postorder is contiguous); the rest are "top"
  std::vector<int> sub_lo, sub_root, top;
  i64 nnz_l = 0;
  double flops = 0.0;
  double t_order = 0, t_symbolic = 0, t_numeric = 0;
};

namespace {

int s_ncol(const gm_chol* f, int s) { return f->sfirst[s + 1] - f->sfirst[s]; }
int s_nrow(const gm_chol* f, int s) { return (int)(f->rptr[s + 1] - f->rptr[s]); }

// Symbolic analysis after perm is set.  Returns false with the error set.
bool symbolic(gm_chol* f, const int* ptr, const int* ind) {
  const int n = f->n;
  std::vector<int>& perm = f->perm;
  std::vector<int> iperm(n);
  for (int q = 0; q < n; ++q) iperm[perm[q]] = q;
  // elimination tree of C = A(perm, perm): row k's entries left of the diagonal
  std::vector<int> parent(n, -1), anc(n, -1);
  for (int k = 0; k < n; ++k) {
    const int o = perm[k];
    for (int j = ptr[o]; j < ptr[o + 1]; ++j) {
      int i = iperm[ind[j]];
      if (i >= k) continue;
      while (anc[i] != -1 && anc[i] != k) {
        const int nx = anc[i];
        anc[i] = k;
        i = nx;
      }
      if (anc[i] == -1) {
        anc[i] = k;
        parent[i] = k;
      }
    }
  }
  // postorder (children in increasing label order)
  std::vector<int> head(n, -1), nxt(n, -1);
  for (int v = n - 1; v >= 0; --v)
    if (parent[v] >= 0) {
      nxt[v] = head[parent[v]];
      head[parent[v]] = v;
    }
  std::vector<int> post(n), stk;
  stk.reserve(n);
  int q = 0;
  for (int r = 0; r < n; ++r) {
    if (parent[r] != -1) continue;
    stk.push_back(r);
    while (!stk.empty()) {
      const int v = stk.back();
      if (head[v] >= 0) {
        const int c = head[v];
        head[v] = nxt[c];
        stk.push_back(c);
      } else {
        stk.pop_back();
        post[q++] = v;
      }
    }
  }
  if (q != n) {
    set_err("gm_chol: elimination tree traversal failed");
    return false;
  }
  // relabel: new label j <- post[j]
  std::vector<int> ipost(n);
  for (int j = 0; j < n; ++j) ipost[post[j]] = j;
  std::vector<int> perm2(n), parent2(n);
  for (int j = 0; j < n; ++j) {
    perm2[j] = perm[post[j]];
    parent2[j] = parent[post[j]] >= 0 ? ipost[parent[post[j]]] : -1;
  }
  perm.swap(perm2);
  parent.swap(parent2);
  for (int j = 0; j < n; ++j) iperm[perm[j]] = j;
  // column counts (rows of L incl. the diagonal) by row subtrees
  std::vector<i64> cc(n, 1);
  std::vector<int> mark(n, -1), nchild(n, 0);
  for (int j = 0; j < n; ++j)
    if (parent[j] >= 0) ++nchild[parent[j]];
  for (int k = 0; k < n; ++k) {
    mark[k] = k;
    const int o = perm[k];
    for (int j = ptr[o]; j < ptr[o + 1]; ++j) {
      int i = iperm[ind[j]];
      if (i >= k) continue;
      while (mark[i] != k) {
        mark[i] = k;
        ++cc[i];
        i = parent[i];
      }
    }
  }
  // fundamental supernodes
  std::vector<int> sf;
  for (int j = 0; j < n; ++j) {
    const bool join = j > 0 && parent[j - 1] == j && nchild[j] == 1 && cc[j - 1] == cc[j] + 1;
    if (!join) sf.push_back(j);
  }
  sf.push_back(n);
  int ns = (int)sf.size() - 1;
  // relaxed amalgamation: merge a supernode into its parent when the parent
  // starts right after it and the merged block adds few explicit zeros
  std::vector<int> sup_of(n);
  for (int s = 0; s < ns; ++s)
    for (int j = sf[s]; j < sf[s + 1]; ++j) sup_of[j] = s;
  std::vector<int> first(ns), last(ns);
  std::vector<i64> nrow(ns), truenz(ns), stored(ns);
  std::vector<int> alive(ns, 1), into(ns);
  for (int s = 0; s < ns; ++s) {
    first[s] = sf[s];
    last[s] = sf[s + 1] - 1;
    nrow[s] = cc[sf[s]];
    i64 tn = 0;
    for (int j = sf[s]; j < sf[s + 1]; ++j) tn += cc[j];
    truenz[s] = tn;
    const i64 c = sf[s + 1] - sf[s];
    stored[s] = c * nrow[s] - c * (c - 1) / 2;
    into[s] = s;
  }
  auto find = [&](int s) {
    while (into[s] != s) {
      into[s] = into[into[s]];
      s = into[s];
    }
    return s;
  };
  for (int s = 0; s < ns; ++s) {
    const int r = find(s);
    if (r != s) continue;
    const int pj = parent[last[s]];
    if (pj < 0) continue;
    const int p = find(sup_of[pj]);
    if (first[p] != last[s] + 1) continue;
    const i64 c = (last[s] - first[s] + 1) + (last[p] - first[p] + 1);
    const i64 nr = (last[s] - first[s] + 1) + nrow[p];
    const i64 st = c * nr - c * (c - 1) / 2;
    const i64 tn = truenz[s] + truenz[p];
    const double z = (double)(st - tn) / (double)st;
    const bool merge = c <= 4 || (c <= 16 && z < 0.8) || (c <= 48 && z < 0.1) || z < 0.05;
    if (!merge) continue;
    into[s] = p;
    first[p] = first[s];
    nrow[p] = nr;
    truenz[p] = tn;
    stored[p] = st;
    alive[s] = 0;
  }
  f->sfirst.clear();
  for (int s = 0; s < ns; ++s)
    if (alive[s]) f->sfirst.push_back(first[s]);
  // alive supernodes in column order: their firsts are increasing
  std::sort(f->sfirst.begin(), f->sfirst.end());
  f->sfirst.push_back(n);
  ns = (int)f->sfirst.size() - 1;
  f->nsuper = ns;
  for (int s = 0; s < ns; ++s)
    for (int j = f->sfirst[s]; j < f->sfirst[s + 1]; ++j) sup_of[j] = s;
  f->sparent.assign(ns, -1);
  for (int s = 0; s < ns; ++s) {
    const int pj = parent[f->sfirst[s + 1] - 1];
    f->sparent[s] = pj >= 0 ? sup_of[pj] : -1;
  }
  // row structures: columns, A's entries below, children's rows below
  std::vector<int> chead(ns, -1), cnext(ns, -1);
  for (int s = ns - 1; s >= 0; --s)
    if (f->sparent[s] >= 0) {
      cnext[s] = chead[f->sparent[s]];
      chead[f->sparent[s]] = s;
    }
  f->rptr.assign(ns + 1, 0);
  f->rows.clear();
  std::fill(mark.begin(), mark.end(), -1);
  std::vector<int> tmp;
  for (int s = 0; s < ns; ++s) {
    const int c0 = f->sfirst[s], c1 = f->sfirst[s + 1];
    tmp.clear();
    for (int j = c0; j < c1; ++j) {
      tmp.push_back(j);
      mark[j] = s;
    }
    for (int j = c0; j < c1; ++j) {
      const int o = perm[j];
      for (int e = ptr[o]; e < ptr[o + 1]; ++e) {
        const int i = iperm[ind[e]];
        if (i >= c1 && mark[i] != s) {
          mark[i] = s;
          tmp.push_back(i);
        }
      }
    }
    for (int c = chead[s]; c >= 0; c = cnext[c]) {
      const int cnc = f->sfirst[c + 1] - f->sfirst[c];
      for (i64 e = f->rptr[c] + cnc; e < f->rptr[c + 1]; ++e) {
        const int i = f->rows[e];
        if (i >= c1 && mark[i] != s) {
          mark[i] = s;
          tmp.push_back(i);
        }
      }
    }
    std::sort(tmp.begin() + (c1 - c0), tmp.end());
    f->rows.insert(f->rows.end(), tmp.begin(), tmp.end());
    f->rptr[s + 1] = (i64)f->rows.size();
  }
  f->lptr.assign(ns + 1, 0);
  f->nnz_l = 0;
  f->flops = 0.0;
  std::vector<double> work(ns, 0.0);
  for (int s = 0; s < ns; ++s) {
    const i64 nc = s_ncol(f, s), nr = s_nrow(f, s), mu = nr - nc;
    f->lptr[s + 1] = f->lptr[s] + nc * nr;
    f->nnz_l += nc * nr - nc * (nc - 1) / 2;
    work[s] = (double)nc * nc * nc / 3.0 + (double)nc * nc * mu + (double)nc * mu * mu;
    f->flops += work[s];
  }
  // fixed subtree cut: split every subtree above total / 96 work into its
  // children (which then become subtrees) and the node itself (top)
  std::vector<double> sw(work);
  for (int s = 0; s < ns; ++s)
    if (f->sparent[s] >= 0) sw[f->sparent[s]] += sw[s];
  double total = 0;
  for (int s = 0; s < ns; ++s)
    if (f->sparent[s] < 0) total += sw[s];
  const double cut = total / 96.0;
  std::vector<int> desc_lo(ns);
  for (int s = 0; s < ns; ++s) desc_lo[s] = s;
  for (int s = 0; s < ns; ++s)
    if (f->sparent[s] >= 0) desc_lo[f->sparent[s]] = std::min(desc_lo[f->sparent[s]], desc_lo[s]);
  std::vector<int> cand;
  for (int s = 0; s < ns; ++s)
    if (f->sparent[s] < 0) cand.push_back(s);
  std::vector<char> is_top(ns, 0);
  f->sub_lo.clear();
  f->sub_root.clear();
  while (!cand.empty()) {
    const int s = cand.back();
    cand.pop_back();
    if (sw[s] > cut && chead[s] >= 0) {
      is_top[s] = 1;
      for (int c = chead[s]; c >= 0; c = cnext[c]) cand.push_back(c);
    } else if (sw[s] > cut) {
      is_top[s] = 1;   // a heavy leaf: factored as a top node
    } else {
      f->sub_root.push_back(s);
    }
  }
  std::sort(f->sub_root.begin(), f->sub_root.end());
  for (int r : f->sub_root) f->sub_lo.push_back(desc_lo[r]);
  f->top.clear();
  for (int s = 0; s < ns; ++s)
    if (is_top[s]) f->top.push_back(s);
  f->iperm.swap(iperm);
  return true;
}

// Extend-add of child c's update matrix into the front of s.  pos[] maps a
// global row of s to its local position.
void extend_add(const gm_chol* f, int s, int c, const double* Uc, const int* pos, double* Lb,
                double* U) {
  const int nc = s_ncol(f, s), nr = s_nrow(f, s), mu = nr - nc;
  const int cnc = s_ncol(f, c), cmu = s_nrow(f, c) - cnc;
  const int* cr = f->rows.data() + f->rptr[c] + cnc;
  for (int b = 0; b < cmu; ++b) {
    const int pb = pos[cr[b]];
    const double* src = Uc + (i64)b * cmu;
    if (pb < nc) {
      double* dst = Lb + (i64)pb * nr;
      for (int a = b; a < cmu; ++a) dst[pos[cr[a]]] += src[a];
    } else {
      double* dst = U + (i64)(pb - nc) * mu - nc;
      for (int a = b; a < cmu; ++a) dst[pos[cr[a]]] += src[a];
    }
  }
}

// Assemble and factor supernode s.  updates[c] holds the update matrices of
// finished children (freed here).  Returns the failing global column or -1.
int factor_super(gm_chol* f, int s, const int* ptr, const int* ind, const double* val,
                 std::vector<std::vector<double>>& updates, std::vector<int>& pos,
                 const std::vector<int>& chead, const std::vector<int>& cnext, int threads) {
  const int c0 = f->sfirst[s], nc = s_ncol(f, s), nr = s_nrow(f, s), mu = nr - nc;
  const int* rs = f->rows.data() + f->rptr[s];
  for (int a = 0; a < nr; ++a) pos[rs[a]] = a;
  double* Lb = f->lval.data() + f->lptr[s];
  std::fill(Lb, Lb + (i64)nr * nc, 0.0);
  std::vector<double> U((size_t)mu * mu, 0.0);
  for (int j = 0; j < nc; ++j) {
    const int o = f->perm[c0 + j];
    double* col = Lb + (i64)j * nr;
    for (int e = ptr[o]; e < ptr[o + 1]; ++e) {
      const int i = f->iperm[ind[e]];
      if (i >= c0 + j) col[pos[i]] += val[e];
    }
    col[j] -= f->sigma;
  }
  for (int c = chead[s]; c >= 0; c = cnext[c]) {
    if (!updates[c].empty()) extend_add(f, s, c, updates[c].data(), pos.data(), Lb, U.data());
    std::vector<double>().swap(updates[c]);
  }
  const int bad = factor_front(nr, nc, Lb, U.data(), threads);
  if (bad >= 0) return c0 + bad;
  if (mu > 0) updates[s].swap(U);
  return -1;
}

bool numeric(gm_chol* f, const int* ptr, const int* ind, const double* val) {
  const int ns = f->nsuper, n = f->n;
  f->lval.assign((size_t)f->lptr[ns], 0.0);
  std::vector<int> chead(ns, -1), cnext(ns, -1);
  for (int s = ns - 1; s >= 0; --s)
    if (f->sparent[s] >= 0) {
      cnext[s] = chead[f->sparent[s]];
      chead[f->sparent[s]] = s;
    }
  std::vector<std::vector<double>> updates(ns);
  const int T = f->threads;
  std::vector<std::vector<int>> pos(T);
  std::atomic<int> bad(-1);
  const int nsub = (int)f->sub_root.size();
  // largest subtrees first
  std::vector<int> order(nsub);
  std::vector<double> sw(nsub, 0.0);
  for (int t = 0; t < nsub; ++t) {
    order[t] = t;
    sw[t] = (double)(f->lptr[f->sub_root[t] + 1] - f->lptr[f->sub_lo[t]]);
  }
  std::sort(order.begin(), order.end(), [&](int a, int b) { return sw[a] > sw[b]; });
  const bool ok = parallel_for(T, nsub, [&](i64 k, int tid) {
    if (bad.load() >= 0) return;
    const int t = order[k];
    if (pos[tid].empty()) pos[tid].assign(n, 0);
    for (int s = f->sub_lo[t]; s <= f->sub_root[t]; ++s) {
      const int b = factor_super(f, s, ptr, ind, val, updates, pos[tid], chead, cnext, 1);
      if (b >= 0) {
        bad = b;
        return;
      }
    }
  });
  if (!ok) {
    set_err("gm_chol: out of host memory in the numeric factorisation");
    return false;
  }
  if (pos[0].empty()) pos[0].assign(n, 0);
  for (int s : f->top) {
    if (bad.load() >= 0) break;
    const int b = factor_super(f, s, ptr, ind, val, updates, pos[0], chead, cnext, T);
    if (b >= 0) bad = b;
  }
  if (bad.load() >= 0) {
    set_err("gm_chol: A - sigma I is not positive definite (pivot " +
            std::to_string(bad.load()) + " of the nested-dissection order)");
    return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Solves on Y (n x nr row-major, permuted order).
// ---------------------------------------------------------------------------

// x (nc x nr row-major) <- L11^(-1) x
void trsv_lower_rows(int nc, int nrh, const double* Lb, int ld, double* x) {
  for (int j = 0; j < nc; ++j) {
    const double* cj = Lb + (i64)j * ld;
    double* xj = x + (i64)j * nrh;
    const double d = cj[j];
    for (int r = 0; r < nrh; ++r) xj[r] /= d;
    for (int i = j + 1; i < nc; ++i) {
      const double l = cj[i];
      double* xi = x + (i64)i * nrh;
      for (int r = 0; r < nrh; ++r) xi[r] -= l * xj[r];
    }
  }
}

// x (nc x nr row-major) <- L11^(-T) x
void trsv_upper_rows(int nc, int nrh, const double* Lb, int ld, double* x) {
  for (int j = nc - 1; j >= 0; --j) {
    const double* cj = Lb + (i64)j * ld;
    double* xj = x + (i64)j * nrh;
    for (int i = j + 1; i < nc; ++i) {
      const double l = cj[i];
      const double* xi = x + (i64)i * nrh;
      for (int r = 0; r < nrh; ++r) xj[r] -= l * xi[r];
    }
    const double d = cj[j];
    for (int r = 0; r < nrh; ++r) xj[r] /= d;
  }
}

// Forward step of supernode s.  Updates for rows inside [lo_row, n) that
// belong to `own` (pos_out < 0) go straight into Y; rows mapped by
// out_pos >= 0 go into out (mu_out x nrh row-major) instead.
void fwd_super(const gm_chol* f, int s, int nrh, double* Y, const int* out_pos, double* out,
               std::vector<double>& t2, int threads = 1) {
  const int c0 = f->sfirst[s], nc = s_ncol(f, s), nr = s_nrow(f, s), mu = nr - nc;
  const double* Lb = f->lval.data() + f->lptr[s];
  double* x = Y + (i64)c0 * nrh;
  trsv_lower_rows(nc, nrh, Lb, nr, x);
  if (mu == 0) return;
  t2.assign((size_t)mu * nrh, 0.0);
  // t2 (mu x nrh col-major) -= L21 x
  gemm_sub_par(threads, mu, nrh, nc, Lb + nc, nr, x, nrh, true, t2.data(), mu);
  const int* rs = f->rows.data() + f->rptr[s] + nc;
  for (int a = 0; a < mu; ++a) {
    const int row = rs[a];
    const int op = out_pos ? out_pos[row] : -1;
    double* dst = op >= 0 ? out + (i64)op * nrh : Y + (i64)row * nrh;
    for (int r = 0; r < nrh; ++r) dst[r] += t2[a + (size_t)r * mu];
  }
}

void bwd_super(const gm_chol* f, int s, int nrh, double* Y, std::vector<double>& g,
               int threads = 1) {
  const int c0 = f->sfirst[s], nc = s_ncol(f, s), nr = s_nrow(f, s), mu = nr - nc;
  const double* Lb = f->lval.data() + f->lptr[s];
  double* x = Y + (i64)c0 * nrh;
  if (mu > 0) {
    const int* rs = f->rows.data() + f->rptr[s] + nc;
    g.resize((size_t)mu * nrh);
    for (int a = 0; a < mu; ++a) std::memcpy(&g[(size_t)a * nrh], Y + (i64)rs[a] * nrh,
                                             sizeof(double) * nrh);
    // x^T (nrh x nc, ld nrh) -= g^T (nrh x mu) L21 (mu x nc)
    gemm_sub_par(threads, nrh, nc, mu, g.data(), nrh, Lb + nc, nr, false, x, nrh);
  }
  trsv_upper_rows(nc, nrh, Lb, nr, x);
}

void solve_permuted(const gm_chol* f, int nrh, double* Y) {
  const int nsub = (int)f->sub_root.size();
  const int T = f->threads;
  // forward: subtrees in parallel; updates leaving a subtree land in rows of
  // its root's update matrix, collected per subtree and added in a fixed order
  std::vector<std::vector<double>> outs(nsub);
  std::vector<std::vector<int>> posv(T);
  parallel_for(T, nsub, [&](i64 t, int tid) {
    const int root = f->sub_root[t];
    const int rnc = s_ncol(f, root), rmu = s_nrow(f, root) - rnc;
    const int* rr = f->rows.data() + f->rptr[root] + rnc;
    if (posv[tid].empty()) posv[tid].assign(f->n, -1);
    std::vector<int>& op = posv[tid];
    for (int a = 0; a < rmu; ++a) op[rr[a]] = a;
    outs[t].assign((size_t)rmu * nrh, 0.0);
    std::vector<double> t2;
    for (int s = f->sub_lo[t]; s <= root; ++s) fwd_super(f, s, nrh, Y, op.data(), outs[t].data(), t2);
    for (int a = 0; a < rmu; ++a) op[rr[a]] = -1;
  });
  for (int t = 0; t < nsub; ++t) {
    const int root = f->sub_root[t];
    const int rnc = s_ncol(f, root), rmu = s_nrow(f, root) - rnc;
    const int* rr = f->rows.data() + f->rptr[root] + rnc;
    for (int a = 0; a < rmu; ++a) {
      double* dst = Y + (i64)rr[a] * nrh;
      const double* src = outs[t].data() + (size_t)a * nrh;
      for (int r = 0; r < nrh; ++r) dst[r] += src[r];
    }
  }
  std::vector<double> t2;
  for (int s : f->top) fwd_super(f, s, nrh, Y, nullptr, nullptr, t2, T);
  // backward: top nodes in reverse, then subtrees in parallel
  for (auto it = f->top.rbegin(); it != f->top.rend(); ++it) bwd_super(f, *it, nrh, Y, t2, T);
  parallel_for(T, nsub, [&](i64 t, int) {
    std::vector<double> g;
    for (int s = f->sub_root[t]; s >= f->sub_lo[t]; --s) bwd_super(f, s, nrh, Y, g);
  });
}

bool check_csr(int32_t n, const int32_t* ptr, const int32_t* ind) {
  if (n <= 0 || !ptr || !ind) return false;
  if (ptr[0] != 0) return false;
  for (int i = 0; i < n; ++i)
    if (ptr[i + 1] < ptr[i]) return false;
  for (int j = 0; j < ptr[n]; ++j)
    if (ind[j] < 0 || ind[j] >= n) return false;
  return true;
}

// ---------------------------------------------------------------------------
// Tall-skinny products with a fixed row chunking (thread-count independent).
// V is n x ldv row-major (the basis), W n x s row-major.
// ---------------------------------------------------------------------------
constexpr int kRowChunk = 4096;

// Row-streaming kernels on row-major blocks (no packing: a row of V or W is
// contiguous).  gram: C (a x s, col-major) += V[r0:r1, 0:a]^T W[r0:r1, 0:s],
// eight columns of V against eight of W at a time, rows in order.  axpy:
// W[r0:r1, 0:s] += alpha V[r0:r1, 0:a] Ct (Ct a x s row-major), eight rows at
// a time, i in order.  One body, compiled for three instruction sets.
#define GM_GRAM_BODY                                                                  \
  const int a8 = a & ~7, s8 = s & ~7;                                                 \
  for (int i0 = 0; i0 < a8; i0 += 8)                                                  \
    for (int j0 = 0; j0 < s8; j0 += 8) {                                              \
      v8d acc[8];                                                                     \
      for (int j = 0; j < 8; ++j) acc[j] = (v8d){0, 0, 0, 0, 0, 0, 0, 0};             \
      for (int r = r0; r < r1; ++r) {                                                 \
        v8d v;                                                                        \
        std::memcpy(&v, V + (i64)r * ldv + i0, sizeof v);                             \
        const double* w = W + (i64)r * s + j0;                                        \
        for (int j = 0; j < 8; ++j) {                                                 \
          const double t = w[j];                                                      \
          acc[j] += v * (v8d){t, t, t, t, t, t, t, t};                                \
        }                                                                             \
      }                                                                               \
      for (int j = 0; j < 8; ++j)                                                     \
        for (int q = 0; q < 8; ++q) C[(i0 + q) + (i64)(j0 + j) * a] += acc[j][q];     \
    }                                                                                 \
  for (int i = 0; i < a; ++i)                                                         \
    for (int j = (i < a8 ? s8 : 0); j < s; ++j) {                                     \
      double acc = 0.0;                                                               \
      for (int r = r0; r < r1; ++r) acc += V[(i64)r * ldv + i] * W[(i64)r * s + j];   \
      C[i + (i64)j * a] += acc;                                                       \
    }

#define GM_AXPY_BODY                                                                  \
  const int s8 = s & ~7;                                                              \
  int r = r0;                                                                         \
  for (; r + 8 <= r1; r += 8)                                                         \
    for (int j0 = 0; j0 < s8; j0 += 8) {                                              \
      v8d acc[8];                                                                     \
      for (int q = 0; q < 8; ++q) std::memcpy(&acc[q], W + (i64)(r + q) * s + j0, sizeof(v8d)); \
      for (int i = 0; i < a; ++i) {                                                   \
        v8d c;                                                                        \
        std::memcpy(&c, Ct + (i64)i * s + j0, sizeof c);                              \
        c *= (v8d){alpha, alpha, alpha, alpha, alpha, alpha, alpha, alpha};           \
        for (int q = 0; q < 8; ++q) {                                                 \
          const double t = V[(i64)(r + q) * ldv + i];                                 \
          acc[q] += c * (v8d){t, t, t, t, t, t, t, t};                                \
        }                                                                             \
      }                                                                               \
      for (int q = 0; q < 8; ++q) std::memcpy(W + (i64)(r + q) * s + j0, &acc[q], sizeof(v8d)); \
    }                                                                                 \
  for (int rr = r0; rr < r1; ++rr)                                                    \
    for (int j = (rr < r ? s8 : 0); j < s; ++j) {                                     \
      double acc = W[(i64)rr * s + j];                                                \
      for (int i = 0; i < a; ++i) acc += V[(i64)rr * ldv + i] * (alpha * Ct[(i64)i * s + j]); \
      W[(i64)rr * s + j] = acc;                                                       \
    }

#define GM_TALL_ARGS int r0, int r1, int a, const double* __restrict V, int ldv, int s
__attribute__((target("avx512f,fma"))) void gram_avx512(GM_TALL_ARGS, const double* __restrict W,
                                                        double* __restrict C) {
  GM_GRAM_BODY
}
__attribute__((target("avx2,fma"))) void gram_avx2(GM_TALL_ARGS, const double* __restrict W,
                                                  double* __restrict C) {
  GM_GRAM_BODY
}
void gram_generic(GM_TALL_ARGS, const double* __restrict W, double* __restrict C) { GM_GRAM_BODY }
__attribute__((target("avx512f,fma"))) void axpy_avx512(GM_TALL_ARGS, double alpha,
                                                        const double* __restrict Ct,
                                                        double* __restrict W) {
  GM_AXPY_BODY
}
__attribute__((target("avx2,fma"))) void axpy_avx2(GM_TALL_ARGS, double alpha,
                                                  const double* __restrict Ct,
                                                  double* __restrict W) {
  GM_AXPY_BODY
}
void axpy_generic(GM_TALL_ARGS, double alpha, const double* __restrict Ct, double* __restrict W) {
  GM_AXPY_BODY
}
#undef GM_GRAM_BODY
#undef GM_AXPY_BODY
#undef GM_TALL_ARGS

void tall_gram_rows(int r0, int r1, int a, const double* V, int ldv, int s, const double* W,
                    double* C) {
  const micro_fn m = micro();
  if (m == micro_avx512) gram_avx512(r0, r1, a, V, ldv, s, W, C);
  else if (m == micro_avx2) gram_avx2(r0, r1, a, V, ldv, s, W, C);
  else gram_generic(r0, r1, a, V, ldv, s, W, C);
}

void tall_axpy_rows(int r0, int r1, int a, const double* V, int ldv, int s, double alpha,
                    const double* Ct, double* W) {
  const micro_fn m = micro();
  if (m == micro_avx512) axpy_avx512(r0, r1, a, V, ldv, s, alpha, Ct, W);
  else if (m == micro_avx2) axpy_avx2(r0, r1, a, V, ldv, s, alpha, Ct, W);
  else axpy_generic(r0, r1, a, V, ldv, s, alpha, Ct, W);
}

// C (a x s, col-major) = V[:, 0:a]^T W, chunk partials summed in chunk order.
void tall_gram(int threads, int n, int a, const double* V, int ldv, int s, const double* W,
               std::vector<double>& C) {
  const int nch = (n + kRowChunk - 1) / kRowChunk;
  std::vector<double> part((size_t)nch * a * s, 0.0);
  parallel_for(threads, nch, [&](i64 t, int) {
    const int r0 = (int)t * kRowChunk, r1 = std::min(n, r0 + kRowChunk);
    tall_gram_rows(r0, r1, a, V, ldv, s, W, part.data() + (size_t)t * a * s);
  });
  C.assign((size_t)a * s, 0.0);
  for (int t = 0; t < nch; ++t) {
    const double* c = part.data() + (size_t)t * a * s;
    for (size_t e = 0; e < (size_t)a * s; ++e) C[e] += c[e];
  }
}

// W += alpha V[:, 0:a] C  (C a x s col-major)
void tall_axpy(int threads, int n, int a, const double* V, int ldv, int s, double alpha,
               const double* C, double* W) {
  std::vector<double> Ct((size_t)a * s);
  for (int i = 0; i < a; ++i)
    for (int j = 0; j < s; ++j) Ct[(size_t)i * s + j] = C[i + (size_t)j * a];
  const int nch = (n + kRowChunk - 1) / kRowChunk;
  parallel_for(threads, nch, [&](i64 t, int) {
    const int r0 = (int)t * kRowChunk, r1 = std::min(n, r0 + kRowChunk);
    tall_axpy_rows(r0, r1, a, V, ldv, s, alpha, Ct.data(), W);
  });
}

// W -= V[:, 0:a] C
void tall_update(int threads, int n, int a, const double* V, int ldv, int s, const double* C,
                 double* W) {
  tall_axpy(threads, n, a, V, ldv, s, -1.0, C, W);
}

// Symmetric eigen-decomposition of a small dense matrix (col-major a, n x n):
// Householder tridiagonalisation, then implicit QL with Wilkinson shifts.
// w ascending; a is overwritten by the eigenvectors (columns).
bool sym_eig_small(int n, std::vector<double>& a, std::vector<double>& w) {
  w.assign(n, 0.0);
  std::vector<double> e(n, 0.0);
  auto A = [&](int i, int j) -> double& { return a[i + (size_t)j * n]; };
  // tridiagonalisation (EISPACK tred2 order: rows from the bottom)
  for (int i = n - 1; i > 0; --i) {
    const int l = i - 1;
    double h = 0.0, scale = 0.0;
    if (l > 0) {
      for (int k = 0; k <= l; ++k) scale += std::fabs(A(i, k));
      if (scale == 0.0) {
        e[i] = A(i, l);
      } else {
        for (int k = 0; k <= l; ++k) {
          A(i, k) /= scale;
          h += A(i, k) * A(i, k);
        }
        double fv = A(i, l);
        const double g = fv >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
        e[i] = scale * g;
        h -= fv * g;
        A(i, l) = fv - g;
        fv = 0.0;
        for (int j = 0; j <= l; ++j) {
          A(j, i) = A(i, j) / h;
          double gg = 0.0;
          for (int k = 0; k <= j; ++k) gg += A(j, k) * A(i, k);
          for (int k = j + 1; k <= l; ++k) gg += A(k, j) * A(i, k);
          e[j] = gg / h;
          fv += e[j] * A(i, j);
        }
        const double hh = fv / (h + h);
        for (int j = 0; j <= l; ++j) {
          const double fj = A(i, j);
          const double gj = e[j] - hh * fj;
          e[j] = gj;
          for (int k = 0; k <= j; ++k) A(j, k) -= fj * e[k] + gj * A(i, k);
        }
      }
    } else {
      e[i] = A(i, l);
    }
    w[i] = h;
  }
  w[0] = 0.0;
  e[0] = 0.0;
  for (int i = 0; i < n; ++i) {
    const int l = i - 1;
    if (w[i] != 0.0) {
      for (int j = 0; j <= l; ++j) {
        double g = 0.0;
        for (int k = 0; k <= l; ++k) g += A(i, k) * A(k, j);
        for (int k = 0; k <= l; ++k) A(k, j) -= g * A(k, i);
      }
    }
    w[i] = A(i, i);
    A(i, i) = 1.0;
    for (int j = 0; j <= l; ++j) A(j, i) = A(i, j) = 0.0;
  }
  // implicit QL on (w, e)
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(w[m]) + std::fabs(w[m + 1]);
        if (std::fabs(e[m]) <= 1e-300 || std::fabs(e[m]) <= 2.2e-16 * dd) break;
      }
      if (m != l) {
        if (iter++ == 60) return false;
        double g = (w[l + 1] - w[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = w[m] - w[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double fv = s * e[i];
          const double b = c * e[i];
          r = std::hypot(fv, g);
          e[i + 1] = r;
          if (r == 0.0) {
            w[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = fv / r;
          c = g / r;
          g = w[i + 1] - p;
          r = (w[i] - g) * s + 2.0 * c * b;
          p = s * r;
          w[i + 1] = g + p;
          g = c * r - b;
          for (int k = 0; k < n; ++k) {
            fv = A(k, i + 1);
            A(k, i + 1) = s * A(k, i) + c * fv;
            A(k, i) = c * A(k, i) - s * fv;
          }
        }
        if (r == 0.0 && i >= l) continue;
        w[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  // ascending order
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int x, int y) { return w[x] < w[y]; });
  std::vector<double> w2(n), a2((size_t)n * n);
  for (int j = 0; j < n; ++j) {
    w2[j] = w[idx[j]];
    std::memcpy(&a2[(size_t)j * n], &a[(size_t)idx[j] * n], sizeof(double) * n);
  }
  w.swap(w2);
  a.swap(a2);
  return true;
}

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// uniform [-1, 1) deviate for (seed, original row, column)
double deviate(uint64_t seed, i64 row, int col) {
  const uint64_t h = splitmix64(seed ^ splitmix64((uint64_t)row * 1024u + (uint64_t)col));
  return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// Orthonormalise the s columns of W (n x s row-major) against V[:, 0:a]
// (gs_passes classical Gram-Schmidt passes) and among themselves (Cholesky
// QR, twice).  R (s x s,
// col-major, upper) gets the coefficients of the first pass's QR when not
// null.  Columns that vanish are replaced by fresh deviates.  Returns false
// if that does not recover.
bool orthonormalise(int threads, int n, const double* V, int ldv, int a, int s, double* W,
                    double* R, uint64_t seed, int gs_passes, int* refills) {
  std::vector<double> C, G;
  for (int attempt = 0; attempt < 4; ++attempt) {
    for (int pass = 0; pass < (attempt == 0 ? gs_passes : 2) && a > 0; ++pass) {
      tall_gram(threads, n, a, V, ldv, s, W, C);
      tall_update(threads, n, a, V, ldv, s, C.data(), W);
    }
    std::vector<double> Rtot((size_t)s * s, 0.0);
    for (int i = 0; i < s; ++i) Rtot[i + (size_t)i * s] = 1.0;
    bool ok = true;
    int badcol = -1;
    for (int pass = 0; pass < 2 && ok; ++pass) {
      tall_gram(threads, n, s, W, s, s, W, G);
      // Cholesky G = R^T R (R upper, col-major)
      std::vector<double> Rp((size_t)s * s, 0.0);
      double gmax = 0.0;
      for (int i = 0; i < s; ++i) gmax = std::max(gmax, G[i + (size_t)i * s]);
      for (int j = 0; j < s && ok; ++j) {
        for (int i = 0; i <= j; ++i) {
          double v = G[i + (size_t)j * s];
          for (int k = 0; k < i; ++k) v -= Rp[k + (size_t)i * s] * Rp[k + (size_t)j * s];
          if (i == j) {
            if (!(v > 1e-24 * std::max(gmax, 1e-300)) || !std::isfinite(v)) {
              ok = false;
              badcol = j;
              break;
            }
            Rp[j + (size_t)j * s] = std::sqrt(v);
          } else {
            Rp[i + (size_t)j * s] = v / Rp[i + (size_t)i * s];
          }
        }
      }
      if (!ok) break;
      // W <- W Rp^(-1), row by row
      const int nch = (n + kRowChunk - 1) / kRowChunk;
      parallel_for(threads, nch, [&](i64 t, int) {
        const int r0 = (int)t * kRowChunk, r1 = std::min(n, r0 + kRowChunk);
        for (int r = r0; r < r1; ++r) {
          double* w = W + (i64)r * s;
          for (int j = 0; j < s; ++j) {
            double v = w[j];
            for (int k = 0; k < j; ++k) v -= w[k] * Rp[k + (size_t)j * s];
            w[j] = v / Rp[j + (size_t)j * s];
          }
        }
      });
      // Rtot <- Rp Rtot
      std::vector<double> Rn((size_t)s * s, 0.0);
      for (int j = 0; j < s; ++j)
        for (int i = 0; i <= j; ++i) {
          double v = 0.0;
          for (int k = i; k <= j; ++k) v += Rp[i + (size_t)k * s] * Rtot[k + (size_t)j * s];
          Rn[i + (size_t)j * s] = v;
        }
      Rtot.swap(Rn);
    }
    if (ok) {
      if (R) std::memcpy(R, Rtot.data(), sizeof(double) * s * s);
      return true;
    }
    // a (near) dependent column: replace it by a fresh deviate and retry
    if (refills) ++*refills;
    for (int r = 0; r < n; ++r)
      W[(i64)r * s + badcol] = deviate(seed + 7919u * (attempt + 1), r, badcol);
    if (R) std::fill(R, R + (size_t)s * s, 0.0);
  }
  return false;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* gm_chol_last_error(void) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  return g_err.c_str();
}

gm_status gm_nd_order(int32_t n, const int32_t* ptr, const int32_t* ind, int32_t* perm) {
  if (!perm || !check_csr(n, ptr, ind)) return GM_EINVAL;
  try {
    std::vector<int> p;
    nested_dissection(n, ptr, ind, p, resolve_threads(0));
    std::memcpy(perm, p.data(), sizeof(int) * n);
  } catch (...) {
    set_err("gm_nd_order: out of host memory");
    return GM_ENOMEM;
  }
  return GM_OK;
}

gm_status gm_chol_factor(int32_t n, const int32_t* ptr, const int32_t* ind, const double* val,
                         double sigma, int32_t threads, gm_chol** out) {
  if (!out || !val || !check_csr(n, ptr, ind)) {
    set_err("gm_chol_factor: invalid arguments");
    return GM_EINVAL;
  }
  *out = nullptr;
  gm_chol* f = nullptr;
  try {
    f = new gm_chol();
    f->n = n;
    f->threads = resolve_threads(threads);
    f->sigma = sigma;
    double t0 = now_s();
    nested_dissection(n, ptr, ind, f->perm, f->threads);
    double t1 = now_s();
    f->t_order = t1 - t0;
    if (!symbolic(f, ptr, ind)) {
      delete f;
      return GM_EINVAL;
    }
    double t2 = now_s();
    f->t_symbolic = t2 - t1;
    if (!numeric(f, ptr, ind, val)) {
      delete f;
      return GM_EBREAKDOWN;
    }
    f->t_numeric = now_s() - t2;
  } catch (const std::bad_alloc&) {
    delete f;
    set_err("gm_chol_factor: out of host memory");
    return GM_ENOMEM;
  } catch (...) {
    delete f;
    set_err("gm_chol_factor: unexpected failure");
    return GM_EINVAL;
  }
  *out = f;
  return GM_OK;
}

gm_status gm_chol_solve(gm_chol* f, int32_t nrhs, double* X) {
  if (!f) return GM_EBADHANDLE;
  if (nrhs <= 0 || !X) return GM_EINVAL;
  try {
    const int n = f->n;
    std::vector<double> Y((size_t)n * nrhs);
    for (int q = 0; q < n; ++q)
      for (int r = 0; r < nrhs; ++r) Y[(size_t)q * nrhs + r] = X[f->perm[q] + (i64)r * n];
    solve_permuted(f, nrhs, Y.data());
    for (int q = 0; q < n; ++q)
      for (int r = 0; r < nrhs; ++r) X[f->perm[q] + (i64)r * n] = Y[(size_t)q * nrhs + r];
  } catch (...) {
    set_err("gm_chol_solve: out of host memory");
    return GM_ENOMEM;
  }
  return GM_OK;
}

gm_status gm_chol_info(gm_chol* f, int64_t* nnz_l, double* flops, int32_t* nsuper,
                       double* seconds) {
  if (!f) return GM_EBADHANDLE;
  if (nnz_l) *nnz_l = f->nnz_l;
  if (flops) *flops = f->flops;
  if (nsuper) *nsuper = f->nsuper;
  if (seconds) {
    seconds[0] = f->t_order;
    seconds[1] = f->t_symbolic;
    seconds[2] = f->t_numeric;
  }
  return GM_OK;
}

gm_status gm_chol_perm(gm_chol* f, int32_t* perm) {
  if (!f) return GM_EBADHANDLE;
  if (!perm) return GM_EINVAL;
  std::memcpy(perm, f->perm.data(), sizeof(int) * f->n);
  return GM_OK;
}

gm_status gm_chol_delete(gm_chol* f) {
  delete f;
  return GM_OK;
}

gm_status gm_eigs_shift_invert(const gm_sieigs_desc* d, double* lambda, double* Q,
                               int32_t* blocks_out, double* max_residual) {
  if (!d || !lambda || !check_csr(d->n, d->A_ptr, d->A_ind) || !d->A_values) {
    set_err("gm_eigs_shift_invert: invalid arguments");
    return GM_EINVAL;
  }
  const int n = d->n, K = d->num_eigs, s = d->block_size;
  const int restarts = d->max_restarts >= 0 ? d->max_restarts : 4;
  if (K < 1 || s < K || s > 512 || 2 * s > n) {
    set_err("gm_eigs_shift_invert: need 1 <= num_eigs <= block_size <= min(512, n / 2)");
    return GM_EINVAL;
  }
  // at least two blocks; the basis never exceeds n columns
  const int maxb = std::max(2, std::min(d->max_blocks > 0 ? d->max_blocks : 12, n / s));
  const double tol = d->tol > 0 ? d->tol : 1e-14;
  const int T = resolve_threads(d->threads);
  // ||A||_inf (largest absolute row sum), the scale of the residual test
  double anorm = 0.0;
  for (int r = 0; r < n; ++r) {
    double acc = 0.0;
    for (int e = d->A_ptr[r]; e < d->A_ptr[r + 1]; ++e) acc += std::fabs(d->A_values[e]);
    anorm = std::max(anorm, acc);
  }
  if (!(anorm > 0.0)) anorm = 1.0;
  gm_chol* f = nullptr;
  gm_status st = gm_chol_factor(n, d->A_ptr, d->A_ind, d->A_values, d->sigma, T, &f);
  if (st != GM_OK) return st;
  int blocks_done = 0;
  double worst = 0.0;
  bool converged = false;
  // phase seconds (GM_CHOL_TRACE=1 prints them): solve, first Gram-Schmidt
  // pass, second pass + QR, Rayleigh-Ritz
  double ph[4] = {0, 0, 0, 0};
  try {
    const int ldv = s * maxb;
    std::vector<double> V((size_t)n * ldv, 0.0);   // basis, permuted rows, row-major
    std::vector<double> W((size_t)n * s), X((size_t)n * s);
    // start block: deviates indexed by the original row
    for (int q = 0; q < n; ++q)
      for (int c = 0; c < s; ++c) X[(size_t)q * s + c] = deviate(d->seed, f->perm[q], c);
    std::vector<double> ritz_vals;
    std::vector<double> Yk;       // Ritz vectors in basis coordinates (a x K)
    int a_final = 0;
    for (int rs = 0; rs <= restarts && !converged; ++rs) {
      // first block from X
      if (!orthonormalise(T, n, nullptr, ldv, 0, s, X.data(), nullptr, d->seed + rs, 0, nullptr)) {
        set_err("gm_eigs_shift_invert: could not orthonormalise the start block");
        st = GM_EBREAKDOWN;
        break;
      }
      for (int q = 0; q < n; ++q)
        std::memcpy(&V[(size_t)q * ldv], &X[(size_t)q * s], sizeof(double) * s);
      std::vector<double> Tm;   // block tridiagonal, (maxb*s)^2 col-major
      const int tdim = maxb * s;
      Tm.assign((size_t)tdim * tdim, 0.0);
      std::vector<double> Bprev((size_t)s * s, 0.0), C, R((size_t)s * s);
      for (int j = 0; j < maxb; ++j) {
        // W = (A - sigma I)^(-1) V_j
        for (int q = 0; q < n; ++q)
          std::memcpy(&W[(size_t)q * s], &V[(size_t)q * ldv + (size_t)j * s], sizeof(double) * s);
        double tp = now_s();
        solve_permuted(f, s, W.data());
        ph[0] += now_s() - tp;
        tp = now_s();
        ++blocks_done;
        const int a = (j + 1) * s;
        // A_j = V_j^T W; full reorthogonalisation against V[:, 0:a]
        tall_gram(T, n, a, V.data(), ldv, s, W.data(), C);
        for (int c = 0; c < s; ++c)
          for (int r = 0; r < s; ++r)
            Tm[(j * s + r) + (size_t)(j * s + c) * tdim] = C[(j * s + r) + (size_t)c * a];
        // symmetrise the diagonal block
        for (int c = 0; c < s; ++c)
          for (int r = c + 1; r < s; ++r) {
            double& x1 = Tm[(j * s + r) + (size_t)(j * s + c) * tdim];
            double& x2 = Tm[(j * s + c) + (size_t)(j * s + r) * tdim];
            const double avg = 0.5 * (x1 + x2);
            x1 = x2 = avg;
          }
        tall_update(T, n, a, V.data(), ldv, s, C.data(), W.data());
        ph[1] += now_s() - tp;
        tp = now_s();
        int refills = 0;
        // (the pass above and one more: "twice is enough")
        if (!orthonormalise(T, n, V.data(), ldv, a, s, W.data(), R.data(),
                            d->seed + 1000003u * (j + 1) + rs, 1, &refills)) {
          st = GM_EBREAKDOWN;
          set_err("gm_eigs_shift_invert: block Lanczos breakdown");
          break;
        }
        // Rayleigh-Ritz on T_j (a x a)
        std::vector<double> Ta((size_t)a * a);
        for (int c = 0; c < a; ++c)
          for (int r = 0; r < a; ++r) Ta[r + (size_t)c * a] = Tm[r + (size_t)c * tdim];
        std::vector<double> w;
        ph[2] += now_s() - tp;
        tp = now_s();
        if (!sym_eig_small(a, Ta, w)) {
          st = GM_EBREAKDOWN;
          set_err("gm_eigs_shift_invert: tridiagonal QL did not converge");
          break;
        }
        // estimated residual of Ritz pair i in the inverted operator,
        // || R y_i[last block] || / theta_i, the largest theta first
        bool all = true;
        for (int i = 0; i < K; ++i) {
          const int col = a - 1 - i;
          const double th = w[col];
          double rr = 0.0;
          for (int r = 0; r < s; ++r) {
            double v = 0.0;
            for (int c = r; c < s; ++c) v += R[r + (size_t)c * s] * Ta[(a - s + c) + (size_t)col * a];
            rr += v * v;
          }
          const double res = std::sqrt(rr) / std::fabs(th);
          if (!(res <= tol)) all = false;
        }
        ph[3] += now_s() - tp;
        // the next block and the off-diagonal coefficients (unless last)
        const bool last = all || j == maxb - 1;
        if (!last) {
          for (int q = 0; q < n; ++q)
            std::memcpy(&V[(size_t)q * ldv + (size_t)(j + 1) * s], &W[(size_t)q * s],
                        sizeof(double) * s);
          // T[(j+1) block, j block] = R; symmetric position R^T
          for (int c = 0; c < s; ++c)
            for (int r = 0; r < s; ++r) {
              const double v = R[r + (size_t)c * s];
              Tm[((j + 1) * s + r) + (size_t)(j * s + c) * tdim] = v;
              Tm[(j * s + c) + (size_t)((j + 1) * s + r) * tdim] = v;
            }
        }
        if (last) {
          ritz_vals.assign(w.begin(), w.end());
          Yk.assign((size_t)a * s, 0.0);
          // the s largest Ritz vectors, largest theta first
          for (int i = 0; i < s; ++i)
            std::memcpy(&Yk[(size_t)i * a], &Ta[(size_t)(a - 1 - i) * a], sizeof(double) * a);
          a_final = a;
          break;
        }
      }
      if (st != GM_OK) break;
      // X = V[:, 0:a] Yk (the Ritz block: the restart block, or the answer)
      std::fill(X.begin(), X.end(), 0.0);
      tall_axpy(T, n, a_final, V.data(), ldv, s, 1.0, Yk.data(), X.data());
      // the estimate is a proxy (it can undershoot once new blocks are
      // rounding-dominated): accept only on the explicit residuals
      // ||A x_i - lambda_i x_i|| / ||A||_inf of the K wanted pairs
      std::vector<double> Xo((size_t)n * K), lam_k(K);
      for (int i = 0; i < K; ++i) lam_k[i] = d->sigma + 1.0 / ritz_vals[a_final - 1 - i];
      for (int q = 0; q < n; ++q)
        std::memcpy(&Xo[(size_t)f->perm[q] * K], &X[(size_t)q * s], sizeof(double) * K);
      const int nch = (n + kRowChunk - 1) / kRowChunk;
      std::vector<double> part((size_t)nch * K, 0.0);
      parallel_for(T, nch, [&](i64 t, int) {
        const int r0 = (int)t * kRowChunk, r1 = std::min(n, r0 + kRowChunk);
        double* acc2 = part.data() + (size_t)t * K;
        std::vector<double> y(K);
        for (int r = r0; r < r1; ++r) {
          for (int i = 0; i < K; ++i) y[i] = -lam_k[i] * Xo[(size_t)r * K + i];
          for (int e = d->A_ptr[r]; e < d->A_ptr[r + 1]; ++e) {
            const double v = d->A_values[e];
            const double* xr = &Xo[(size_t)d->A_ind[e] * K];
            for (int i = 0; i < K; ++i) y[i] += v * xr[i];
          }
          for (int i = 0; i < K; ++i) acc2[i] += y[i] * y[i];
        }
      });
      worst = 0.0;
      for (int i = 0; i < K; ++i) {
        double ss = 0.0;
        for (int t = 0; t < nch; ++t) ss += part[(size_t)t * K + i];
        worst = std::max(worst, std::sqrt(ss) / anorm);
      }
      converged = worst <= tol;
      if (converged) break;
    }
    if (st == GM_OK) {
      // lambda_i = sigma + 1 / theta_i, ascending (largest theta first)
      const int a = a_final;
      for (int i = 0; i < K; ++i) lambda[i] = d->sigma + 1.0 / ritz_vals[a - 1 - i];
      if (Q)
        for (int q = 0; q < n; ++q)
          for (int i = 0; i < K; ++i) Q[f->perm[q] + (i64)i * n] = X[(size_t)q * s + i];
      if (!converged) {
        set_err("gm_eigs_shift_invert: tolerance not met in the allowed restarts");
        st = GM_EBREAKDOWN;
      }
    }
  } catch (const std::bad_alloc&) {
    set_err("gm_eigs_shift_invert: out of host memory");
    st = GM_ENOMEM;
  }
  if (blocks_out) *blocks_out = blocks_done;
  if (max_residual) *max_residual = worst;
  if (const char* tr = std::getenv("GM_CHOL_TRACE"))
    if (tr[0] == '1')
      std::fprintf(stderr,
                   "[gm_eigs_shift_invert] n=%d nnz(L)=%lld order %.2fs symbolic %.2fs numeric "
                   "%.2fs (%.3g flop); %d block solves %.2fs, GS pass 1 %.2fs, pass 2 + QR "
                   "%.2fs, Rayleigh-Ritz %.2fs\n",
                   n, (long long)f->nnz_l, f->t_order, f->t_symbolic, f->t_numeric, f->flops,
                   blocks_done, ph[0], ph[1], ph[2], ph[3]);
  gm_chol_delete(f);
  return st;
}

gm_status gm_lambda_max(int32_t n, const int32_t* ptr, const int32_t* ind, const double* val,
                        double tol, int32_t max_iter, uint64_t seed, int32_t threads,
                        double* lmax, int32_t* iters) {
  if (!lmax || !val || !check_csr(n, ptr, ind)) return GM_EINVAL;
  const int T = resolve_threads(threads);
  const int maxit = max_iter > 0 ? max_iter : 300;
  const double tl = tol > 0 ? tol : 1e-12;
  try {
    std::vector<double> v(n), vp(n, 0.0), w(n), al, be;
    const int nch = (n + kRowChunk - 1) / kRowChunk;
    std::vector<double> part(nch);
    auto dot = [&](const std::vector<double>& x, const std::vector<double>& y) {
      parallel_for(T, nch, [&](i64 t, int) {
        const int r0 = (int)t * kRowChunk, r1 = std::min(n, r0 + kRowChunk);
        double acc = 0.0;
        for (int r = r0; r < r1; ++r) acc += x[r] * y[r];
        part[t] = acc;
      });
      double s = 0.0;
      for (int t = 0; t < nch; ++t) s += part[t];
      return s;
    };
    for (int i = 0; i < n; ++i) v[i] = deviate(seed, i, 0);
    double nv = std::sqrt(dot(v, v));
    for (int i = 0; i < n; ++i) v[i] /= nv;
    double beta = 0.0, prev = 0.0, theta = 0.0;
    int stable = 0, it = 0;
    // largest eigenvalue of the leading k x k tridiagonal block by bisection
    auto tri_max = [&](int k) {
      double lo = 1e300, hi = -1e300;
      for (int i = 0; i < k; ++i) {
        const double r = (i > 0 ? std::fabs(be[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(be[i]) : 0.0);
        lo = std::min(lo, al[i] - r);
        hi = std::max(hi, al[i] + r);
      }
      // count of eigenvalues below x (Sturm sequence)
      auto below = [&](double x) {
        int c = 0;
        double q = 1.0;
        for (int i = 0; i < k; ++i) {
          const double b2 = i > 0 ? be[i - 1] * be[i - 1] : 0.0;
          q = al[i] - x - (i > 0 ? b2 / q : 0.0);
          if (q == 0.0) q = -1e-300;
          if (q < 0.0) ++c;
        }
        return c;
      };
      for (int b = 0; b < 200 && hi - lo > 1e-15 * std::max(std::fabs(hi), std::fabs(lo)); ++b) {
        const double mid = 0.5 * (lo + hi);
        if (below(mid) >= k) hi = mid;
        else lo = mid;
      }
      return hi;
    };
    for (it = 0; it < maxit; ++it) {
      parallel_for(T, nch, [&](i64 t, int) {
        const int r0 = (int)t * kRowChunk, r1 = std::min(n, r0 + kRowChunk);
        for (int r = r0; r < r1; ++r) {
          double acc = 0.0;
          for (int e = ptr[r]; e < ptr[r + 1]; ++e) acc += val[e] * v[ind[e]];
          w[r] = acc - beta * vp[r];
        }
      });
      const double alpha = dot(w, v);
      for (int i = 0; i < n; ++i) w[i] -= alpha * v[i];
      al.push_back(alpha);
      beta = std::sqrt(dot(w, w));
      theta = tri_max((int)al.size());
      if (it > 0 && std::fabs(theta - prev) <= tl * std::fabs(theta)) {
        if (++stable >= 3) break;
      } else {
        stable = 0;
      }
      prev = theta;
      if (beta == 0.0) break;
      be.push_back(beta);
      for (int i = 0; i < n; ++i) {
        vp[i] = v[i];
        v[i] = w[i] / beta;
      }
    }
    *lmax = theta;
    if (iters) *iters = it + 1;
  } catch (...) {
    set_err("gm_lambda_max: out of host memory");
    return GM_ENOMEM;
  }
  return GM_OK;
}

}  // extern "C"
