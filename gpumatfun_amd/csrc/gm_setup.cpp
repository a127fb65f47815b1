// gm_setup.cpp -- host set-up of the lanczos_create inputs (no device code).
//
// Computes in C++ what problem_startup.m:71-144 computes in MATLAB before it
// creates the engine, so a C/C++/MEX host needs neither MATLAB nor Python:
//   * gm_polynomial_coeffs   -- polynomial_coeffs.m:11-28
//   * gm_contour_setup       -- problem_startup.m:80-103 (nodes, weights)
//   * gm_shift_setup         -- problem_startup.m:106-123, 140-144
//   * gm_eig_weights         -- problem_startup.m:72-77
//   * gm_subspace_size       -- problem_startup_nonuniform.m:101-113
//   * gm_rcm_order           -- the symrcm reordering of process_nonuniform.m:15
//                               (reverse Cuthill-McKee, George & Liu)
// Jacobi elliptic functions of real argument come from the arithmetic-
// geometric mean (A&S 16.4); at the complex points x + iK'/2 they follow from
// the addition formulas A&S 16.21.1-3 with sn, cn, dn(iK'/2 | m) in closed form
// (A&S 16.5 with the complementary parameter).  K(m) = pi / (2 AGM(1, sqrt(1-m))).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#include "../../include/gpumatfun.h"

namespace {

using cplx = std::complex<double>;

// K(m) from the complementary parameter m1 = 1 - m (no cancellation near m = 1).
double ellipk_m1(double m1) {
  double a = 1.0, b = std::sqrt(m1);
  for (int it = 0; it < 64 && std::fabs(a - b) > 1e-16 * a; ++it) {
    const double an = 0.5 * (a + b);
    b = std::sqrt(a * b);
    a = an;
  }
  return M_PI / (2.0 * a);
}

// sn, cn, dn of real u for parameter m (m1 = 1 - m), descending Landen / AGM.
void ellipj_real(double u, double m, double m1, double* sn, double* cn, double* dn) {
  double a[65], c[65];
  a[0] = 1.0;
  c[0] = std::sqrt(m);
  double b = std::sqrt(m1);
  int n = 0;
  while (std::fabs(c[n]) > 1e-17 * a[n] && n < 63) {
    a[n + 1] = 0.5 * (a[n] + b);
    c[n + 1] = 0.5 * (a[n] - b);
    b = std::sqrt(a[n] * b);
    ++n;
  }
  double phi = std::ldexp(a[n] * u, n);
  double phi_prev = phi;
  for (int lev = n; lev > 0; --lev) {
    phi_prev = phi;
    phi = 0.5 * (std::asin(c[lev] / a[lev] * std::sin(phi)) + phi);
  }
  *sn = std::sin(phi);
  *cn = std::cos(phi);
  *dn = n > 0 ? std::cos(phi) / std::cos(phi_prev - phi) : 1.0;
}

// sn, cn, dn at x + i K'(m)/2, k = sqrt(m).
void ellipj_half_period(double x, double m, double k, double m1, cplx* sn, cplx* cn, cplx* dn) {
  double s, c, d;
  ellipj_real(x, m, m1, &s, &c, &d);
  const double s1 = 1.0 / std::sqrt(1.0 + k);
  const double c1 = std::sqrt(k / (1.0 + k));
  const double d1 = std::sqrt(k);
  const double den = c1 * c1 + m * s * s * s1 * s1;
  *sn = cplx(s * d1, c * d * s1 * c1) / den;
  *cn = cplx(c * c1, -s * d * s1 * d1) / den;
  *dn = cplx(d * c1 * d1, -m * s * c * s1) / den;
}

// exp(z) - 1 without cancellation for small |z|.
cplx expm1c(cplx z) {
  const double x = z.real(), y = z.imag();
  const double sh = std::sin(0.5 * y);
  return cplx(std::expm1(x) * std::cos(y) - 2.0 * sh * sh, std::exp(x) * std::sin(y));
}

cplx f1c(cplx x, double eta, double alpha) { return std::exp(-eta * std::pow(x, 0.5 * alpha)); }

cplx f2c(cplx x, double eta, double alpha) {
  const cplx z = -eta * std::pow(x, 0.5 * alpha);
  return expm1c(z) / z;
}

double binom(int n, int k) {
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// Horner value of sum_i c[i] x^(deg - i).
double polyval(const std::vector<double>& c, double x) {
  double v = 0.0;
  for (double ci : c) v = v * x + ci;
  return v;
}

}  // namespace

extern "C" {

gm_status gm_polynomial_coeffs(int32_t degree, double prob_dim, double* gamma) {
  if (degree < 0 || degree > 16 || !gamma || !(prob_dim > 0.0)) return GM_EINVAL;
  const int d = degree;
  std::vector<double> R(d + 2, 0.0);
  for (int j = 1; j <= d + 2; ++j) {
    const int nb = d - j + 3;
    double kap = binom(d + 1, j - 1);
    for (int t = 0; t < j - 1; ++t) kap *= (d + 0.5 - t) / (t + 1.5);
    const double sc = std::pow(0.25, j - 1) * ((j - 1) % 2 ? -1.0 : 1.0) * kap;
    for (int i = 0; i < nb; ++i) R[j - 1 + i] += sc * binom(d - j + 2, i) * std::pow(-4.0, -i);
  }
  for (int i = 0; i <= d; ++i) gamma[i] = -R[d + 1 - i] * (3.0 + 2.0 * d) / std::pow(prob_dim, d - i);
  return GM_OK;
}

gm_status gm_contour_setup(double l1, double lN, int32_t n, double eta, double alpha, double* xi,
                           double* dxidt, double* wts1, double* wts2, double* s_out) {
  if (!(l1 > 0.0) || !(lN > l1) || n < 1 || !xi || !wts1 || !wts2) return GM_EINVAL;
  const double r = std::sqrt(lN / l1);
  const double k = (r - 1.0) / (r + 1.0);
  const double one_minus_k = 2.0 / (r + 1.0);
  const double m = k * k;
  const double m1 = one_minus_k * (1.0 + k);
  const double K = ellipk_m1(m1);
  const double scale = std::sqrt(l1 * lN);
  const double s = -4.0 * K * scale / (k * M_PI * n);
  for (int j = 0; j < n; ++j) {
    const double x = -K + (n - 0.5 - j) * 2.0 * K / n;
    cplx sn, cn, dn;
    ellipj_half_period(x, m, k, m1, &sn, &cn, &dn);
    const cplx z = scale * (1.0 / k + sn) / (1.0 / k - sn);
    const cplx dz = cn * dn / ((1.0 / k - sn) * (1.0 / k - sn));
    const cplx w1 = s * f1c(z, eta, alpha) * dz;
    const cplx w2 = s * f2c(z, eta, alpha) * dz;
    xi[2 * j] = z.real();
    xi[2 * j + 1] = z.imag();
    if (dxidt) {
      dxidt[2 * j] = dz.real();
      dxidt[2 * j + 1] = dz.imag();
    }
    wts1[j] = w1.real();
    wts1[n + j] = w1.imag();
    wts2[j] = w2.real();
    wts2[n + j] = w2.imag();
  }
  if (s_out) *s_out = s;
  return GM_OK;
}

gm_status gm_shift_setup(const double* gamma, int32_t gamma_size, const double* xi, int32_t n,
                         int32_t m, double* gamma_s_real, double* gamma_s_imag, double* etaxi,
                         double* shifts) {
  if (!gamma || gamma_size < 1 || !xi || n < 1 || m < 1 || !gamma_s_real || !gamma_s_imag)
    return GM_EINVAL;
  const int nd = gamma_size;
  for (int j = 0; j < n; ++j) {
    const cplx z(xi[2 * j], xi[2 * j + 1]);
    cplx g = gamma[0];
    gamma_s_real[(size_t)j * nd] = g.real();
    gamma_s_imag[(size_t)j * nd] = g.imag();
    for (int i = 1; i < nd; ++i) {
      g = gamma[i] + z * g;
      gamma_s_real[(size_t)j * nd + i] = g.real();
      gamma_s_imag[(size_t)j * nd + i] = g.imag();
    }
    const cplx ez = z * g;
    if (etaxi) {
      etaxi[2 * j] = ez.real();
      etaxi[2 * j + 1] = ez.imag();
    }
    if (shifts) {
      // exp systems then phi systems; system j holds etaxi_j in all m slots
      for (int half = 0; half < 2; ++half)
        for (int t = 0; t < m; ++t) {
          const size_t o = 2 * ((size_t)half * n * m + (size_t)j * m + t);
          shifts[o] = ez.real();
          shifts[o + 1] = ez.imag();
        }
    }
  }
  return GM_OK;
}

gm_status gm_eig_weights(const double* lambda, int32_t K, double eta, double alpha,
                         double* exp_lambda, double* phi_lambda) {
  if (K < 0 || (K > 0 && (!lambda || !exp_lambda || !phi_lambda))) return GM_EINVAL;
  for (int i = 0; i < K; ++i) {
    // A Neumann operator's zero mode comes out of an eigensolver as a tiny
    // eigenvalue of either sign; a negative one is that zero (its real power
    // would be NaN), and phi(0) = 1.
    const double z = -eta * std::pow(std::fmax(lambda[i], 0.0), 0.5 * alpha);
    exp_lambda[i] = std::exp(z);
    phi_lambda[i] = z == 0.0 ? 1.0 : std::expm1(z) / z;
  }
  if (K > 0) phi_lambda[0] = 1.0;  // problem_startup.m:77
  return GM_OK;
}

gm_status gm_subspace_size(const double* gamma, int32_t gamma_size, double l1, double lN,
                           double tol, int32_t* m) {
  if (!gamma || gamma_size < 1 || !(l1 > 0.0) || !(lN > l1) || !(tol > 0.0) || !m) return GM_EINVAL;
  // q(x) = x p(x); its extrema on [l1, lN] are at the end points or at real
  // roots of q'(x) = sum_i (d+1-i) gamma_i x^(d-i) inside (l1, lN).
  const int d = gamma_size - 1;
  std::vector<double> q(gamma, gamma + gamma_size);
  q.push_back(0.0);
  std::vector<double> dq(gamma_size);
  for (int i = 0; i < gamma_size; ++i) dq[i] = (d + 1 - i) * gamma[i];
  double lo = std::min(polyval(q, l1), polyval(q, lN));
  double hi = std::max(polyval(q, l1), polyval(q, lN));
  // bracket sign changes of q' on a fine grid, then bisect each root
  const int grid = 1 << 16;
  double xa = l1, fa = polyval(dq, xa);
  for (int t = 1; t <= grid; ++t) {
    const double xb = l1 + (lN - l1) * t / grid, fb = polyval(dq, xb);
    if (fa == 0.0 || (fa < 0.0) != (fb < 0.0)) {
      double a = xa, b = xb, ya = fa;
      for (int it = 0; it < 200 && b - a > 1e-15 * std::fabs(b); ++it) {
        const double c = 0.5 * (a + b), yc = polyval(dq, c);
        if ((yc < 0.0) == (ya < 0.0)) {
          a = c;
          ya = yc;
        } else {
          b = c;
        }
      }
      const double v = polyval(q, 0.5 * (a + b));
      lo = std::min(lo, v);
      hi = std::max(hi, v);
    }
    xa = xb;
    fa = fb;
  }
  if (!(lo > 0.0)) return GM_EINVAL;
  const double rr = std::sqrt(hi / lo);
  const double kp = (rr - 1.0) / (rr + 1.0);
  if (!(kp > 0.0 && kp < 1.0)) {
    *m = 1;
    return GM_OK;
  }
  *m = (int32_t)std::ceil(std::log(tol / 2.0) / std::log(kp));
  return GM_OK;
}

// Reverse Cuthill-McKee ordering of a symmetric sparsity pattern (George & Liu,
// "Computer Solution of Large Sparse Positive Definite Systems", 1981): per
// connected component, a pseudo-peripheral start node (repeated breadth-first
// searches from a minimum-degree node, moving to a minimum-degree node of the
// last level while the eccentricity grows), then a breadth-first numbering that
// takes each node's unnumbered neighbours in increasing degree (ties by index),
// and the whole order reversed.  Diagonal entries are ignored.
gm_status gm_rcm_order(int32_t n, const int32_t* ptr, const int32_t* ind, int32_t* perm) {
  if (n < 0 || (n > 0 && (!ptr || !ind || !perm))) return GM_EINVAL;
  if (n == 0) return GM_OK;
  if (ptr[0] != 0) return GM_EINVAL;
  for (int i = 0; i < n; ++i)
    if (ptr[i + 1] < ptr[i]) return GM_EINVAL;
  for (int j = 0; j < ptr[n]; ++j)
    if (ind[j] < 0 || ind[j] >= n) return GM_EINVAL;
  std::vector<int> deg(n, 0);
  for (int i = 0; i < n; ++i)
    for (int j = ptr[i]; j < ptr[i + 1]; ++j)
      if (ind[j] != i) ++deg[i];
  std::vector<int> level(n, -1), order;
  order.reserve(n);
  std::vector<char> numbered(n, 0);
  // breadth-first levels from `root` over the unnumbered nodes; returns the last level
  std::vector<int> queue(n), last;
  auto bfs_levels = [&](int root, int* height) {
    std::vector<int> touched;
    int head = 0, tail = 0;
    queue[tail++] = root;
    level[root] = 0;
    touched.push_back(root);
    int h = 0;
    while (head < tail) {
      const int v = queue[head++];
      h = level[v];
      for (int j = ptr[v]; j < ptr[v + 1]; ++j) {
        const int w = ind[j];
        if (w == v || numbered[w] || level[w] >= 0) continue;
        level[w] = level[v] + 1;
        queue[tail++] = w;
        touched.push_back(w);
      }
    }
    last.clear();
    for (int t = 0; t < tail; ++t)
      if (level[queue[t]] == h) last.push_back(queue[t]);
    for (int v : touched) level[v] = -1;
    *height = h;
  };
  std::vector<int> byd(n);
  for (int i = 0; i < n; ++i) byd[i] = i;
  std::stable_sort(byd.begin(), byd.end(), [&](int a, int b) { return deg[a] < deg[b]; });
  std::vector<int> nbr;
  for (int seed : byd) {
    if (numbered[seed]) continue;
    int root = seed, h = 0;
    bfs_levels(root, &h);
    for (int it = 0; it < 16; ++it) {   // pseudo-peripheral node
      int cand = last[0];
      for (int v : last)
        if (deg[v] < deg[cand] || (deg[v] == deg[cand] && v < cand)) cand = v;
      int h2 = 0;
      bfs_levels(cand, &h2);
      if (h2 <= h) break;
      root = cand;
      h = h2;
    }
    // Cuthill-McKee numbering of this component from root
    size_t head = order.size();
    order.push_back(root);
    numbered[root] = 1;
    while (head < order.size()) {
      const int v = order[head++];
      nbr.clear();
      for (int j = ptr[v]; j < ptr[v + 1]; ++j) {
        const int w = ind[j];
        if (w != v && !numbered[w]) {
          numbered[w] = 1;
          nbr.push_back(w);
        }
      }
      std::sort(nbr.begin(), nbr.end(),
                [&](int a, int b) { return deg[a] < deg[b] || (deg[a] == deg[b] && a < b); });
      for (int w : nbr) order.push_back(w);
    }
  }
  for (int i = 0; i < n; ++i) perm[i] = order[n - 1 - i];
  return GM_OK;
}

// Cache-locality order of a symmetric sparsity pattern by recursive
// bisection: a part of more than `leaf` nodes is numbered breadth-first from a
// pseudo-peripheral node (found as in gm_rcm_order, within the part; each
// further component of the part from its first node) and cut into the first
// and the second half of that numbering; the halves are ordered depth-first,
// first before second, and a part of at most `leaf` nodes keeps its nodes in
// increasing index.  On a mesh every part is a compact piece of it and parts
// consecutive in the order are neighbours, so that rows numbered close
// together gather nearby rows (the bisection tree of a space-filling order).
gm_status gm_locality_order(int32_t n, const int32_t* ptr, const int32_t* ind, int32_t leaf,
                            int32_t* perm) {
  if (n < 0 || leaf < 1 || (n > 0 && (!ptr || !ind || !perm))) return GM_EINVAL;
  if (n == 0) return GM_OK;
  if (ptr[0] != 0) return GM_EINVAL;
  for (int i = 0; i < n; ++i)
    if (ptr[i + 1] < ptr[i]) return GM_EINVAL;
  for (int j = 0; j < ptr[n]; ++j)
    if (ind[j] < 0 || ind[j] >= n) return GM_EINVAL;
  std::vector<int> deg(n, 0);
  for (int i = 0; i < n; ++i)
    for (int j = ptr[i]; j < ptr[i + 1]; ++j)
      if (ind[j] != i) ++deg[i];
  std::vector<int> vs(n), part(n, 0), level(n, -1), queue(n), last;
  for (int i = 0; i < n; ++i) vs[i] = i;
  int stamp = 0;
  // breadth-first numbering of the part's nodes (part[v] == id) from root,
  // into queue[0..); returns the count and the height, last level in `last`
  auto bfs = [&](int root, int id, int* height) {
    int head = 0, tail = 0;
    queue[tail++] = root;
    level[root] = 0;
    int h = 0;
    while (head < tail) {
      const int v = queue[head++];
      h = level[v];
      for (int j = ptr[v]; j < ptr[v + 1]; ++j) {
        const int w = ind[j];
        if (part[w] != id || level[w] >= 0) continue;
        level[w] = level[v] + 1;
        queue[tail++] = w;
      }
    }
    last.clear();
    for (int t = tail - 1; t >= 0 && level[queue[t]] == h; --t) last.push_back(queue[t]);
    for (int t = 0; t < tail; ++t) level[queue[t]] = -1;
    *height = h;
    return tail;
  };
  std::vector<std::pair<int, int>> stack{{0, n}};
  std::vector<int> numbering;
  int out = 0;
  while (!stack.empty()) {
    const auto [lo, hi] = stack.back();
    stack.pop_back();
    const int sz = hi - lo;
    if (sz <= leaf) {
      std::sort(vs.begin() + lo, vs.begin() + hi);
      for (int t = lo; t < hi; ++t) perm[out++] = vs[t];
      continue;
    }
    const int id = ++stamp;
    for (int t = lo; t < hi; ++t) part[vs[t]] = id;
    numbering.clear();
    // every component of the part, from its first node in vs order
    for (int t = lo; t < hi && (int)numbering.size() < sz; ++t) {
      const int seed = vs[t];
      if (part[seed] != id) continue;   // numbered already
      int root = seed, h = 0;
      bfs(root, id, &h);
      for (int it = 0; it < 16; ++it) {
        int cand = last[0];
        for (int v : last)
          if (deg[v] < deg[cand] || (deg[v] == deg[cand] && v < cand)) cand = v;
        int h2 = 0;
        bfs(cand, id, &h2);
        if (h2 <= h) break;
        root = cand;
        h = h2;
      }
      const int cnt = bfs(root, id, &h);
      for (int q = 0; q < cnt; ++q) {
        numbering.push_back(queue[q]);
        part[queue[q]] = -id;   // out of the part for the next components
      }
    }
    std::copy(numbering.begin(), numbering.end(), vs.begin() + lo);
    const int mid = lo + sz / 2;
    stack.push_back({mid, hi});
    stack.push_back({lo, mid});
  }
  return GM_OK;
}

}  // extern "C"
