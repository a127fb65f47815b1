/*
 * fab_oracle.c -- C restatement of the reference's MATLAB CPU path, for
 * TESTS AND THE CPU BASELINE ONLY (never linked into the product library).
 * Built twice: in double (the restatement) and in long double (see FAB_REAL).
 *
 *   lanczos            lanczos.m:1-54
 *   jacobi_poly_eval   jacobi_poly_eval.m:1-19
 *   shifted solves     nicksmethodcpu.m:17-21 (sparse backslash on a
 *                      tridiagonal matrix = LAPACK ZGTSV: Gaussian elimination
 *                      with partial pivoting)
 *   recombination      nicksmethodcpu.m:33-43 (node by node, as written)
 *   step loop          problem_startup_CPU_nonuniform.m:110-122 (with mass = 1
 *                      it is problem_startup_CPU.m:119-129)
 *
 * Sparse products are row-parallel (OpenMP) with each row summed in storage
 * order, so they are bit-identical for any thread count; dot products and
 * norms are summed in fixed 4096-element chunks, then over chunks in order,
 * so they are also independent of the thread count.
 *
 * Arithmetic variants (fab_set_variant, all off by default).  Each switches
 * one rounding choice of the restatement to the one a fused GPU evaluation
 * makes, or adds full reorthogonalisation, so that the effect of each on the
 * result can be measured on its own:
 *   FAB_V_SCALE   the first polynomial term (and, at degree 0, the Lanczos
 *                 product) as (A w) * (gamma_0 * (1/beta)) from the
 *                 unnormalised w, in place of A (gamma_0 * (w / beta));
 *   FAB_V_FMA     fused multiply-adds in the Horner terms, the three-term
 *                 recurrence, the dot products and the norms;
 *   FAB_V_TREE    dot products and norms summed per 256-element block by a
 *                 pairwise tree, then the block sums pairwise, in place of
 *                 sequential 4096-element chunks;
 *   FAB_V_DEFL    the deflation split U - Q Z and Q f(L) Z accumulated with
 *                 fused multiply-adds;
 *   FAB_V_REORTH  full reorthogonalisation of every new Lanczos vector
 *                 against all earlier ones (classical Gram-Schmidt, twice):
 *                 in exact arithmetic the same Krylov approximation, so it
 *                 gives the answer that rounding-free Lanczos would.
 * Two more take effect in the long-double build only, where they round one
 * stage's results to double and so show how much of the difference between
 * the two builds that stage's rounding makes:
 *   FAB_V_RDEFL   the deflated vectors and the deflation update;
 *   FAB_V_RLANCZOS every new Lanczos vector and coefficient;
 *   FAB_V_RSPMV   every sparse product: operand, products and row sums in
 *                 double.
 */
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* Working precision: double, or long double (x87, 64-bit significand) for the
 * extended-precision build (FAB_REAL=long double), whose rounding errors are
 * 2^11 times smaller -- the reference point that the double-precision
 * evaluation orders are measured against.  The interface is double either way. */
#ifndef FAB_REAL
#define FAB_REAL double
#endif
typedef FAB_REAL real;
typedef FAB_REAL complex zd;

#define CHUNK 4096

enum {
  FAB_V_SCALE = 1, FAB_V_FMA = 2, FAB_V_TREE = 4, FAB_V_DEFL = 8, FAB_V_REORTH = 16,
  FAB_V_RDEFL = 32, FAB_V_RLANCZOS = 64, FAB_V_RSPMV = 128
};

/* x rounded to double (a no-op in the double build) */
static inline FAB_REAL to_double(FAB_REAL x) { return (FAB_REAL)(double)x; }
static unsigned g_variant = 0;

unsigned fab_set_variant(unsigned flags) {
  unsigned old = g_variant;
  g_variant = flags;
  return old;
}

#define TREE_BLOCK 256

/* pairwise sum of v[0..n) (n >= 1) */
static real pairwise(const real* v, long n) {
  if (n == 1) return v[0];
  long h = n / 2;
  return pairwise(v, h) + pairwise(v + h, n - h);
}

/* sum of a[i] * b[i] per 256-element block by a pairwise tree, then the block
 * sums pairwise */
static real dot_tree(long n, const real* a, const real* b, int fused) {
  if (n <= 0) return 0.0;
  long nb = (n + TREE_BLOCK - 1) / TREE_BLOCK;
  real* part = (real*)malloc(sizeof(real) * nb);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nb; ++c) {
    long lo = c * TREE_BLOCK, hi = lo + TREE_BLOCK < n ? lo + TREE_BLOCK : n;
    real t[TREE_BLOCK];
    for (long i = lo; i < hi; ++i) t[i - lo] = fused ? fma(a[i], b[i], 0.0) : a[i] * b[i];
    part[c] = pairwise(t, hi - lo);
  }
  real s = pairwise(part, nb);
  free(part);
  return s;
}

static real dot_fixed(long n, const real* a, const real* b) {
  if (g_variant & FAB_V_TREE) return dot_tree(n, a, b, (g_variant & FAB_V_FMA) != 0);
  const int fused = (g_variant & FAB_V_FMA) != 0;
  long nc = (n + CHUNK - 1) / CHUNK;
  real* part = (real*)malloc(sizeof(real) * (nc > 0 ? nc : 1));
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    long lo = c * CHUNK, hi = lo + CHUNK < n ? lo + CHUNK : n;
    real s = 0.0;
    if (fused)
      for (long i = lo; i < hi; ++i) s = fma(a[i], b[i], s);
    else
      for (long i = lo; i < hi; ++i) s += a[i] * b[i];
    part[c] = s;
  }
  real s = 0.0;
  for (long c = 0; c < nc; ++c) s += part[c];
  free(part);
  return s;
}

/* y(:,c) = A x(:,c) for ncol columns, column-major with leading dim n. */
static void spmm(long n, const int* rp, const int* ci, const double* va, int ncol,
                 const real* x, real* y) {
  if (g_variant & FAB_V_RSPMV) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
      for (int c = 0; c < ncol; ++c) {
        const real* xc = x + (size_t)c * n;
        double s = 0.0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) s += va[j] * (double)xc[ci[j]];
        y[(size_t)c * n + i] = s;
      }
    }
    return;
  }
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    for (int c = 0; c < ncol; ++c) {
      const real* xc = x + (size_t)c * n;
      real s = 0.0;
      for (int j = rp[i]; j < rp[i + 1]; ++j) s += va[j] * xc[ci[j]];
      y[(size_t)c * n + i] = s;
    }
  }
}

/* jacobi_poly_eval.m with z = 1 column (real and imaginary parts). */
static void poly_eval(long n, const int* rp, const int* ci, const double* va, int nd,
                      const double* gr, const double* gi, const real* br, const real* bi,
                      real* xr, real* xi, real* work /* 4n */) {
  for (long k = 0; k < n; ++k) {
    xr[k] = br[k] * gr[0] - bi[k] * gi[0];
    xi[k] = br[k] * gi[0] + bi[k] * gr[0];
  }
  real* in = work;
  real* out = work + 2 * n;
  for (int i = 1; i < nd; ++i) {
    memcpy(in, xr, sizeof(real) * n);
    memcpy(in + n, xi, sizeof(real) * n);
    spmm(n, rp, ci, va, 2, in, out);
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; ++k) {
      xr[k] = out[k] + br[k] * gr[i] - bi[k] * gi[i];
      xi[k] = out[n + k] + br[k] * gi[i] + bi[k] * gr[i];
    }
  }
}

/* w[:, 0..1] -= V1/V2[:, 0..cols) (V^T w), twice: full reorthogonalisation */
static void reorth2(long n, int cols, const real* V1, const real* V2, real* w) {
  real* h = (real*)malloc(sizeof(real) * (cols > 0 ? cols : 1));
  for (int pass = 0; pass < 2; ++pass) {
    for (int c = 0; c < 2; ++c) {
      const real* V = c == 0 ? V1 : V2;
      real* wc = w + (size_t)c * n;
      for (int i = 0; i < cols; ++i) h[i] = dot_fixed(n, V + (size_t)i * n, wc);
#pragma omp parallel for schedule(static)
      for (long k = 0; k < n; ++k) {
        real s = 0.0;
        for (int i = 0; i < cols; ++i) s += V[(size_t)i * n + k] * h[i];
        wc[k] -= s;
      }
    }
  }
  free(h);
}

/* a * x + y, fused when FAB_V_FMA is set */
static inline real axpy1(real a, real x, real y, int fused) {
  return fused ? fma(a, x, y) : y + x * a;
}

/* lanczos.m for two columns; V1, V2 n x m column-major; al (m), be (m+1). */
static void lanczos2(long n, const int* rp, const int* ci, const double* va, int m, int nd,
                     const double* gamma, const real* a, const real* b, real* V1,
                     real* V2, real* al1, real* be1, real* al2, real* be2,
                     real* work /* 8n */) {
  const int fused = (g_variant & FAB_V_FMA) != 0;
  const int scale = (g_variant & FAB_V_SCALE) != 0;
  real* w = work;               /* n x 2 */
  real* t = work + 2 * n;       /* n x 2 */
  real* u = work + 4 * n;       /* n x 2 */
  real* zero = work + 6 * n;    /* 2n zeros */
  memset(zero, 0, sizeof(real) * 2 * n);
  be1[0] = sqrt(dot_fixed(n, a, a));
  be2[0] = sqrt(dot_fixed(n, b, b));
  for (long k = 0; k < n; ++k) {
    V1[k] = a[k] / be1[0];
    V2[k] = b[k] / be2[0];
  }
  for (int j = 0; j < m; ++j) {
    const real* v1 = V1 + (size_t)j * n;
    const real* v2 = V2 + (size_t)j * n;
    int first = 1;       /* first Horner term still to form */
    int product = 0;     /* w already holds A p(A) v */
    if (scale) {
      /* (A w)(gamma_0 (1/beta_j)) from the unnormalised w (a, b at j = 0) */
      if (j == 0) {
        memcpy(u, a, sizeof(real) * n);
        memcpy(u + n, b, sizeof(real) * n);
      } else {
        memcpy(u, w, sizeof(real) * 2 * n);
      }
      spmm(n, rp, ci, va, 2, u, t);
      const real s1 = gamma[0] * (1.0 / be1[j]), s2 = gamma[0] * (1.0 / be2[j]);
      if (nd >= 2) {
#pragma omp parallel for schedule(static)
        for (long k = 0; k < n; ++k) {
          u[k] = axpy1(gamma[1], v1[k], t[k] * s1, fused);
          u[n + k] = axpy1(gamma[1], v2[k], t[n + k] * s2, fused);
        }
      } else {
#pragma omp parallel for schedule(static)
        for (long k = 0; k < n; ++k) {
          w[k] = t[k] * s1;
          w[n + k] = t[n + k] * s2;
        }
        product = 1;
      }
      first = 0;
    }
    if (first) {
      /* w = p(A) [v1 v2]  (jacobi_poly_eval with zero imaginary parts) */
      for (long k = 0; k < n; ++k) {
        u[k] = v1[k] * gamma[0];
        u[n + k] = v2[k] * gamma[0];
      }
    }
    for (int i = first ? 1 : 2; i < nd; ++i) {
      spmm(n, rp, ci, va, 2, u, t);
#pragma omp parallel for schedule(static)
      for (long k = 0; k < n; ++k) {
        u[k] = axpy1(gamma[i], v1[k], t[k], fused);
        u[n + k] = axpy1(gamma[i], v2[k], t[n + k], fused);
      }
    }
    if (!product) spmm(n, rp, ci, va, 2, u, w);                    /* lanczos.m:27 */
    if (j != 0) {                                                  /* :30-32 */
      const real* p1 = V1 + (size_t)(j - 1) * n;
      const real* p2 = V2 + (size_t)(j - 1) * n;
#pragma omp parallel for schedule(static)
      for (long k = 0; k < n; ++k) {
        w[k] = axpy1(-be1[j], p1[k], w[k], fused);
        w[n + k] = axpy1(-be2[j], p2[k], w[n + k], fused);
      }
    }
    al1[j] = dot_fixed(n, w, v1);                                  /* :35-36 */
    al2[j] = dot_fixed(n, w + n, v2);
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; ++k) {                                 /* :38 */
      w[k] = axpy1(-al1[j], v1[k], w[k], fused);
      w[n + k] = axpy1(-al2[j], v2[k], w[n + k], fused);
    }
    if (g_variant & FAB_V_REORTH) reorth2(n, j + 1, V1, V2, w);
    be1[j + 1] = sqrt(dot_fixed(n, w, w));                         /* :40-41 */
    be2[j + 1] = sqrt(dot_fixed(n, w + n, w + n));
    if (g_variant & FAB_V_RLANCZOS) {
      al1[j] = to_double(al1[j]);
      al2[j] = to_double(al2[j]);
      be1[j + 1] = to_double(be1[j + 1]);
      be2[j + 1] = to_double(be2[j + 1]);
      for (long k = 0; k < 2 * n; ++k) w[k] = to_double(w[k]);
    }
    if (j < m - 1) {                                               /* :44-47 */
      real* n1 = V1 + (size_t)(j + 1) * n;
      real* n2 = V2 + (size_t)(j + 1) * n;
#pragma omp parallel for schedule(static)
      for (long k = 0; k < n; ++k) {
        n1[k] = w[k] / be1[j + 1];
        n2[k] = w[n + k] / be2[j + 1];
      }
    }
  }
}

static real cabs1(zd z) { return fabs(creal(z)) + fabs(cimag(z)); }

/* ZGTSV: returns 0 or the 1-based index of a zero pivot. */
static int zgtsv1(int n, zd* dl, zd* d, zd* du, zd* b) {
  zd* du2 = (zd*)calloc(n > 2 ? n - 2 : 1, sizeof(zd));
  int info = 0;
  for (int k = 0; k < n - 1 && !info; ++k) {
    if (dl[k] == 0) {
      if (d[k] == 0) info = k + 1;
    } else if (cabs1(d[k]) >= cabs1(dl[k])) {
      zd mult = dl[k] / d[k];
      d[k + 1] = d[k + 1] - mult * du[k];
      b[k + 1] = b[k + 1] - mult * b[k];
    } else {
      zd mult = d[k] / dl[k];
      d[k] = dl[k];
      zd temp = d[k + 1];
      d[k + 1] = du[k] - mult * temp;
      if (k < n - 2) {
        du2[k] = du[k + 1];
        du[k + 1] = -mult * du2[k];
      }
      du[k] = temp;
      temp = b[k];
      b[k] = b[k + 1];
      b[k + 1] = temp - mult * b[k + 1];
    }
  }
  if (!info && d[n - 1] == 0) info = n;
  if (!info) {
    b[n - 1] = b[n - 1] / d[n - 1];
    if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / d[n - 2];
    for (int k = n - 3; k >= 0; --k)
      b[k] = (b[k] - du[k] * b[k + 1] - du2[k] * b[k + 2]) / d[k];
  }
  free(du2);
  return info;
}

/* y(:,j) = s*w_j*((etaxi_j I - H) \ beta0 e1), real/imag parts m x ns. */
static int shifted(int m, int ns, const real* al, const real* be, const double* etaxi,
                   const double* w, real s, real* yr, real* yi) {
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int j = 0; j < ns; ++j) {
    zd* dl = (zd*)malloc(sizeof(zd) * m);
    zd* d = (zd*)malloc(sizeof(zd) * m);
    zd* du = (zd*)malloc(sizeof(zd) * m);
    zd* x = (zd*)malloc(sizeof(zd) * m);
    zd ex = etaxi[2 * j] + I * etaxi[2 * j + 1];
    for (int k = 0; k < m; ++k) {
      d[k] = ex - al[k];
      x[k] = 0;
    }
    for (int k = 0; k + 1 < m; ++k) dl[k] = du[k] = -be[k + 1];
    x[0] = be[0];
    if (zgtsv1(m, dl, d, du, x)) bad = 1;
    zd sc = s * (w[2 * j] + I * w[2 * j + 1]);
    for (int k = 0; k < m; ++k) {
      zd v = sc * x[k];
      yr[(size_t)j * m + k] = creal(v);
      yi[(size_t)j * m + k] = cimag(v);
    }
    free(dl); free(d); free(du); free(x);
  }
  return bad;
}

/* out = V y  (n x m times m) */
static void gemv(long n, int m, const real* V, const real* y, real* out) {
#pragma omp parallel for schedule(static)
  for (long k = 0; k < n; ++k) {
    real s = 0.0;
    for (int j = 0; j < m; ++j) s += V[(size_t)j * n + k] * y[j];
    out[k] = s;
  }
}

/*
 * nicksmethodcpu.m: v1 ~ f1(A) a, v2 ~ f2(A) b.
 *   wts1/wts2/etaxi: complex interleaved (ns); gsr/gsi: nd x ns column-major.
 */
static int fab_nicksmethodcpu(long n, const int* rp, const int* ci, const double* va, int m, int nd,
                       const double* gamma, int ns, const double* gsr, const double* gsi,
                       const double* wts1, const double* wts2, double s, const double* etaxi,
                       const real* a, const real* b, real* v1, real* v2) {
  real* V1 = (real*)malloc(sizeof(real) * (size_t)n * m);
  real* V2 = (real*)malloc(sizeof(real) * (size_t)n * m);
  real* work = (real*)malloc(sizeof(real) * 8 * (size_t)n);
  real* al1 = (real*)malloc(sizeof(real) * m);
  real* al2 = (real*)malloc(sizeof(real) * m);
  real* be1 = (real*)malloc(sizeof(real) * (m + 1));
  real* be2 = (real*)malloc(sizeof(real) * (m + 1));
  real* y1r = (real*)malloc(sizeof(real) * (size_t)m * ns);
  real* y1i = (real*)malloc(sizeof(real) * (size_t)m * ns);
  real* y2r = (real*)malloc(sizeof(real) * (size_t)m * ns);
  real* y2i = (real*)malloc(sizeof(real) * (size_t)m * ns);
  real* Yr = (real*)malloc(sizeof(real) * (size_t)n);
  real* Yi = (real*)malloc(sizeof(real) * (size_t)n);
  real* xr = (real*)malloc(sizeof(real) * (size_t)n);
  real* xi = (real*)malloc(sizeof(real) * (size_t)n);
  lanczos2(n, rp, ci, va, m, nd, gamma, a, b, V1, V2, al1, be1, al2, be2, work);
  int bad = shifted(m, ns, al1, be1, etaxi, wts1, s, y1r, y1i);
  bad |= shifted(m, ns, al2, be2, etaxi, wts2, s, y2r, y2i);
  memset(v1, 0, sizeof(real) * n);
  memset(v2, 0, sizeof(real) * n);
  for (int i = 0; i < ns; ++i) {                                  /* nicksmethodcpu.m:33 */
    for (int col = 0; col < 2; ++col) {
      const real* V = col == 0 ? V1 : V2;
      const real* yr = (col == 0 ? y1r : y2r) + (size_t)i * m;
      const real* yi = (col == 0 ? y1i : y2i) + (size_t)i * m;
      real* v = col == 0 ? v1 : v2;
      gemv(n, m, V, yr, Yr);                                      /* :34-37 */
      gemv(n, m, V, yi, Yi);
      poly_eval(n, rp, ci, va, nd, gsr + (size_t)i * nd, gsi + (size_t)i * nd, Yr, Yi, xr,
                xi, work);                                        /* :39-40 */
#pragma omp parallel for schedule(static)
      for (long k = 0; k < n; ++k) v[k] = v[k] + xi[k];            /* :41-42 */
    }
  }
  free(V1); free(V2); free(work); free(al1); free(al2); free(be1); free(be2);
  free(y1r); free(y1i); free(y2r); free(y2i); free(Yr); free(Yi); free(xr); free(xi);
  return bad;
}

/*
 * `steps` time steps of problem_startup_CPU_nonuniform.m:110-122:
 *   deflate (Q n x K column-major), f(A) products, u = x1 + dt x2 + up1 + dt up2,
 *   g = mass ubar (1 - ubar), ubar = u / mass.
 * u and gu are updated in place.
 */
int fab_cpu_steps(long n, const int* rp, const int* ci, const double* va, int K,
                  const double* Q, const double* el, const double* pl, double dt, int m,
                  int nd, const double* gamma, int ns, const double* gsr, const double* gsi,
                  const double* wts1, const double* wts2, double s, const double* etaxi,
                  const double* mass, int steps, double* u_io, double* gu_io) {
  /* Q in the working precision (a cast when that is double) */
  const int same = sizeof(real) == sizeof(double);
  real* Qw = NULL;
  if (!same && K > 0) {
    Qw = (real*)malloc(sizeof(real) * (size_t)n * K);
    for (size_t i = 0; i < (size_t)n * K; ++i) Qw[i] = Q[i];
  }
  const real* Qr = same ? (const real*)(const void*)Q : Qw;
  real* u = (real*)malloc(sizeof(real) * n);
  real* gu = (real*)malloc(sizeof(real) * n);
  for (long i = 0; i < n; ++i) {
    u[i] = u_io[i];
    gu[i] = gu_io[i];
  }
  real* xb = (real*)malloc(sizeof(real) * n);
  real* yb = (real*)malloc(sizeof(real) * n);
  real* up1 = (real*)calloc(n, sizeof(real));
  real* up2 = (real*)calloc(n, sizeof(real));
  real* x1 = (real*)malloc(sizeof(real) * n);
  real* x2 = (real*)malloc(sizeof(real) * n);
  real* Qu = (real*)malloc(sizeof(real) * (K > 0 ? K : 1));
  real* Qg = (real*)malloc(sizeof(real) * (K > 0 ? K : 1));
  int bad = 0;
  for (int st = 0; st < steps; ++st) {
    for (int k = 0; k < K; ++k) {                                   /* Q'*u, Q'*gu */
      Qu[k] = dot_fixed(n, Qr + (size_t)k * n, u);
      Qg[k] = dot_fixed(n, Qr + (size_t)k * n, gu);
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
      real cu = 0.0, cg = 0.0, e1 = 0.0, e2 = 0.0;
      if (g_variant & FAB_V_DEFL) {
        for (int k = 0; k < K; ++k) {
          const real q = Qr[(size_t)k * n + i];
          cu = fma(q, Qu[k], cu);
          cg = fma(q, Qg[k], cg);
          e1 = fma(q, el[k] * Qu[k], e1);
          e2 = fma(q, pl[k] * Qg[k], e2);
        }
      } else {
        for (int k = 0; k < K; ++k) {
          const real q = Qr[(size_t)k * n + i];
          cu += q * Qu[k];
          cg += q * Qg[k];
          e1 += q * (el[k] * Qu[k]);
          e2 += q * (pl[k] * Qg[k]);
        }
      }
      xb[i] = u[i] - cu;
      yb[i] = gu[i] - cg;
      up1[i] = e1;
      up2[i] = e2;
      if (g_variant & FAB_V_RDEFL) {
        xb[i] = to_double(xb[i]);
        yb[i] = to_double(yb[i]);
        up1[i] = to_double(up1[i]);
        up2[i] = to_double(up2[i]);
      }
    }
    bad |= fab_nicksmethodcpu(n, rp, ci, va, m, nd, gamma, ns, gsr, gsi, wts1, wts2, s, etaxi,
                              xb, yb, x1, x2);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
      u[i] = x1[i] + dt * x2[i] + up1[i] + dt * up2[i];
      const real ub = u[i] / mass[i];
      gu[i] = mass[i] * ub * (1 - ub);
    }
  }
  for (long i = 0; i < n; ++i) {
    u_io[i] = (double)u[i];
    gu_io[i] = (double)gu[i];
  }
  free(xb); free(yb); free(up1); free(up2); free(x1); free(x2); free(Qu); free(Qg);
  free(u); free(gu); free(Qw);
  return bad;
}

int fab_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
  return omp_get_max_threads();
#else
  (void)t;
  return 1;
#endif
}
