/*
 * Intermediate-representation (IR) basis: singular-value expansion of the
 * logistic kernel, computed from scratch so that imaginary-time grids exist
 * for any cutoff Lambda (the reference only reads precomputed grid files,
 * SymmetryMBPT/src/common/transformer_t.cpp:48-82, and ships Lambda = 1e3).
 *
 * Kernels on x, y in [-1, 1] (Shinaoka et al., PRB 96, 035147 (2017);
 * Chikano et al., Comput. Phys. Commun. 240, 181 (2019)):
 *   fermionic  K_F(x, y) = exp(-Lambda y (x + 1) / 2) / (1 + exp(-Lambda y))
 *   bosonic    K_B(x, y) = y exp(-Lambda y (x + 1) / 2) / (1 - exp(-Lambda y))
 * K = sum_l s_l u_l(x) v_l(y), with int u_l^2 dx = 1, u_l(1) > 0 and parity
 * u_l(-x) = (-1)^l u_l(x) (both kernels satisfy K(x, y) = K(-x, -y)).
 *
 * Method (own implementation):
 *  1. The kernel splits into even and odd parts on [0, 1]^2,
 *     K_pm(x, y) = K(x, y) pm K(x, -y); each is discretized with
 *     Gauss-Legendre rules on segments refined geometrically towards x = 1
 *     and y = 0, where the functions vary on the scale 1 / Lambda:
 *     A_ij = sqrt(w_i) K_pm(x_i, y_j) sqrt(w_j).
 *  2. A is factored in binary128 arithmetic (__float128): Householder QR with
 *     column pivoting truncated at the numerical rank, then a one-sided
 *     Jacobi SVD of the small triangular factor.  Double precision cannot
 *     resolve singular values near 1e-15 s_0, which the last basis functions
 *     of a grid have.
 *  3. u_l at the nodes, u_l(x_i) = U_il / sqrt(w_i), becomes a Legendre
 *     series per segment (the Gauss rule is exact for the projection).
 *  4. Sampling points: tau nodes are the midpoints between consecutive zeros
 *     of u_{L-1} with -1 and 1 added (L points); Matsubara nodes are where
 *     the smooth part of u^_{L-1}(i w_n) changes sign.
 *  5. u^_l(i w_n) = (1/sqrt 2) int exp(i pi (n + zeta/2) (x + 1)) u_l(x) dx
 *     (zeta = 1 fermions, 0 bosons), segment by segment in closed form:
 *     int_{-1}^{1} P_k(t) exp(i z t) dt = 2 i^k j_k(z) (spherical Bessel).
 */
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef __float128 qd;

static const double kPi = 3.14159265358979323846;

/* Gauss-Legendre nodes and weights on [-1, 1] in binary128 (Newton on P_n). */
static void gauss_legendre_q(int n, qd *x, qd *w) {
  for (int i = 0; i < n; ++i) {
    qd z = cosq(M_PIq * ((qd)i + 0.75Q) / ((qd)n + 0.5Q));
    for (int it = 0; it < 100; ++it) {
      qd p0 = 1, p1 = z;
      for (int k = 2; k <= n; ++k) {
        qd p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      /* p1 = P_n(z), p0 = P_{n-1}(z) */
      qd dp = n * (z * p1 - p0) / (z * z - 1);
      qd dz = p1 / dp;
      z -= dz;
      if (fabsq(dz) < 1e-33Q) {
        p0 = 1;
        p1 = z;
        for (int k = 2; k <= n; ++k) {
          qd p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
          p0 = p1;
          p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1);
        x[n - 1 - i] = z;
        w[n - 1 - i] = 2 / ((1 - z * z) * dp * dp);
        break;
      }
    }
  }
}

/* K(x, y) for y >= 0 and x in [-1, 1]; K(x, -y) = K(-x, y). */
static qd kernel_q(int fermi, qd lam, qd x, qd y) {
  qd a = lam * y;
  if (fermi) return expq(-a * (x + 1) / 2) / (1 + expq(-a));
  /* y / (1 - exp(-a)) = 1 / lam * a / (1 - exp(-a)), -> 1 / lam as a -> 0 */
  qd den = -expm1q(-a);
  qd f = a > 0 ? a / den : 1;
  return expq(-a * (x + 1) / 2) * f / lam;
}

/* Segment edges on [0, 1]: `nuni` uniform pieces below `mid`, then
 * geometric pieces towards `end` (0 or 1) down to a width 1 / (16 Lambda),
 * each split in `split` parts. */
static int half_edges(double lam, int towards_one, int nuni, int split, double *e) {
  int n = 0;
  double g[128];
  int ng = 0;
  /* geometric distances from the refined end: 1/2, 1/4, ..., <= 1/(16 Lambda) */
  double d = 0.5;
  while (d > 1.0 / (16.0 * lam) && ng < 120) {
    g[ng++] = d;
    d *= 0.5;
  }
  g[ng++] = 0.0;
  if (towards_one) {
    for (int i = 0; i < nuni; ++i) e[n++] = 0.5 * i / nuni;
    for (int i = 0; i + 1 < ng; ++i)
      for (int s = 0; s < split; ++s) e[n++] = 1.0 - (g[i] + (g[i + 1] - g[i]) * s / split);
    e[n++] = 1.0;
  } else {
    e[n++] = 0.0;
    for (int i = ng - 2; i >= 0; --i)
      for (int s = 1; s <= split; ++s) e[n++] = g[i + 1] + (g[i] - g[i + 1]) * s / split;
    /* e[n-1] == 0.5 now */
    for (int i = 1; i <= nuni; ++i) e[n++] = 0.5 + 0.5 * i / nuni;
  }
  return n - 1; /* number of segments */
}

/* Nodes and weights of a composite rule with `deg` points per segment. */
static void composite(const double *edges, int nseg, int deg, const qd *gx, const qd *gw, qd *x, qd *w) {
  for (int s = 0; s < nseg; ++s) {
    qd a = edges[s], b = edges[s + 1], c = (a + b) / 2, h = (b - a) / 2;
    for (int i = 0; i < deg; ++i) {
      x[s * deg + i] = c + h * gx[i];
      w[s * deg + i] = h * gw[i];
    }
  }
}

/* Householder QR with column pivoting of A (m x n, column-major, lda = m),
 * stopped at rank r when the largest remaining column norm falls below
 * tol * (largest initial column norm) or r == rmax.  On return the first r
 * columns of A below the diagonal hold the reflectors (v_k with v_k[k] = 1
 * implied... stored explicitly in V), R (r x n, row-major) the triangular
 * factor in pivoted column order, perm the column order. */
static int rrqr_q(int m, int n, qd *A, int rmax, qd tol, qd *V /* m x rmax */, qd *beta, qd *R /* rmax x n */,
                  int *perm) {
  qd *cn = malloc(sizeof(qd) * n);
  qd cmax0 = 0;
  for (int j = 0; j < n; ++j) {
    qd s = 0;
    for (int i = 0; i < m; ++i) s += A[(size_t)j * m + i] * A[(size_t)j * m + i];
    cn[j] = s;
    perm[j] = j;
    if (s > cmax0) cmax0 = s;
  }
  int r = 0;
  for (int k = 0; k < rmax && k < m && k < n; ++k) {
    /* pivot: largest remaining column norm (recomputed exactly) */
    int p = k;
    qd best = -1;
#pragma omp parallel for schedule(static)
    for (int j = k; j < n; ++j) {
      qd s = 0;
      for (int i = k; i < m; ++i) s += A[(size_t)j * m + i] * A[(size_t)j * m + i];
      cn[j] = s;
    }
    for (int j = k; j < n; ++j)
      if (cn[j] > best) {
        best = cn[j];
        p = j;
      }
    if (best <= tol * tol * cmax0) break;
    if (p != k) {
      for (int i = 0; i < m; ++i) {
        qd t = A[(size_t)k * m + i];
        A[(size_t)k * m + i] = A[(size_t)p * m + i];
        A[(size_t)p * m + i] = t;
      }
      int t = perm[k];
      perm[k] = perm[p];
      perm[p] = t;
    }
    /* reflector for column k, rows k..m-1 */
    qd *a = A + (size_t)k * m;
    qd alpha = sqrtq(best);
    if (a[k] > 0) alpha = -alpha;
    qd *v = V + (size_t)k * m;
    for (int i = 0; i < k; ++i) v[i] = 0;
    v[k] = a[k] - alpha;
    for (int i = k + 1; i < m; ++i) v[i] = a[i];
    qd vv = 0;
    for (int i = k; i < m; ++i) vv += v[i] * v[i];
    beta[k] = vv > 0 ? 2 / vv : 0;
    /* apply to the remaining columns */
#pragma omp parallel for schedule(static)
    for (int j = k; j < n; ++j) {
      qd *c = A + (size_t)j * m;
      qd d = 0;
      for (int i = k; i < m; ++i) d += v[i] * c[i];
      d *= beta[k];
      for (int i = k; i < m; ++i) c[i] -= d * v[i];
    }
    r = k + 1;
  }
  for (int i = 0; i < r; ++i)
    for (int j = 0; j < n; ++j) R[(size_t)i * n + j] = j >= i ? A[(size_t)j * m + i] : 0;
  free(cn);
  return r;
}

/* One-sided Jacobi on the columns of X (n x r, column-major): X J = W with
 * orthogonal columns; J (r x r, column-major) accumulates the rotations.
 * Pairs are visited in round-robin (tournament) order: the r/2 pairs of one
 * round are disjoint and rotate in parallel. */
static void rotate_pair(int n, int r, qd *X, qd *J, int p, int q, int *rotated) {
  qd *xp = X + (size_t)p * n, *xq = X + (size_t)q * n;
  qd a = 0, b = 0, c = 0;
  for (int i = 0; i < n; ++i) {
    a += xp[i] * xp[i];
    b += xq[i] * xq[i];
    c += xp[i] * xq[i];
  }
  if (c == 0 || fabsq(c) <= 1e-32Q * sqrtq(a * b)) return;
  *rotated = 1;
  qd zeta = (b - a) / (2 * c);
  qd t = (zeta >= 0 ? 1 : -1) / (fabsq(zeta) + sqrtq(1 + zeta * zeta));
  qd cs = 1 / sqrtq(1 + t * t), sn = cs * t;
  for (int i = 0; i < n; ++i) {
    qd u = xp[i], v = xq[i];
    xp[i] = cs * u - sn * v;
    xq[i] = sn * u + cs * v;
  }
  qd *jp = J + (size_t)p * r, *jq = J + (size_t)q * r;
  for (int i = 0; i < r; ++i) {
    qd u = jp[i], v = jq[i];
    jp[i] = cs * u - sn * v;
    jq[i] = sn * u + cs * v;
  }
}

static void jacobi_q(int n, int r, qd *X, qd *J) {
  for (int i = 0; i < r; ++i)
    for (int j = 0; j < r; ++j) J[(size_t)j * r + i] = i == j;
  if (r < 2) return;
  const int rr = r + (r & 1);  /* players incl. a bye when r is odd */
  int *ring = malloc(sizeof(int) * rr);
  for (int sweep = 0; sweep < 60; ++sweep) {
    int rotated = 0;
    for (int i = 0; i < rr; ++i) ring[i] = i;
    for (int round = 0; round < rr - 1; ++round) {
#pragma omp parallel for schedule(dynamic, 1) reduction(| : rotated)
      for (int h = 0; h < rr / 2; ++h) {
        int p = ring[h], q = ring[rr - 1 - h];
        if (p > q) { int t = p; p = q; q = t; }
        if (q < r) rotate_pair(n, r, X, J, p, q, &rotated);
      }
      /* rotate every player but the first */
      const int last = ring[rr - 1];
      for (int i = rr - 1; i > 1; --i) ring[i] = ring[i - 1];
      ring[1] = last;
    }
    if (!rotated) break;
  }
  free(ring);
}

typedef struct {
  double s;
  int parity;  /* 0 even, 1 odd */
  int idx;     /* column inside its parity block */
} sv_entry;

static int cmp_sv(const void *a, const void *b) {
  double x = ((const sv_entry *)a)->s, y = ((const sv_entry *)b)->s;
  return x < y ? 1 : x > y ? -1 : 0;
}

/* Spherical Bessel j_k(z), k < nk, z >= 0. */
static void sph_bessel(int nk, double z, double *j) {
  if (z == 0.0) {
    j[0] = 1.0;
    for (int k = 1; k < nk; ++k) j[k] = 0.0;
    return;
  }
  if (z > nk + 2) {
    j[0] = sin(z) / z;
    if (nk > 1) j[1] = sin(z) / (z * z) - cos(z) / z;
    for (int k = 1; k + 1 < nk; ++k) j[k + 1] = (2 * k + 1) / z * j[k] - j[k - 1];
    return;
  }
  /* Miller: downward recurrence from well above nk, normalized by j_0 or j_1 */
  int N = nk + 40 + (int)z;
  double fp1 = 0.0, f = 1e-280;
  double buf[256];
  for (int k = N; k >= 1; --k) {
    double fm1 = (2 * k + 1) / z * f - fp1;
    fp1 = f;
    f = fm1;
    if (k - 1 < nk) buf[k - 1] = f;
    if (fabs(f) > 1e250) {
      f *= 1e-250;
      fp1 *= 1e-250;
      for (int i = k - 1; i < nk; ++i) buf[i] *= 1e-250;
    }
  }
  /* buf[0] ~ j_0, buf[1] ~ j_1 up to one common factor */
  double j0 = sin(z) / z, j1 = sin(z) / (z * z) - cos(z) / z;
  double scale = fabs(j0) >= fabs(j1) ? j0 / buf[0] : j1 / buf[1];
  for (int k = 0; k < nk; ++k) j[k] = buf[k] * scale;
}

/* Smooth part of the Matsubara transform of one basis function:
 * int_{-1}^{1} exp(i a x) u(x) dx for real a; u given by Legendre series
 * coef[seg][deg] on the segments of edges.  Returns re + i im. */
static void fourier_segments(int nseg, int deg, const double *edges, const double *coef, double a, double *re,
                             double *im) {
  double jb[64];
  double sr = 0, si = 0;
  for (int s = 0; s < nseg; ++s) {
    const double c = 0.5 * (edges[s] + edges[s + 1]), h = 0.5 * (edges[s + 1] - edges[s]);
    const double z = fabs(a) * h;
    sph_bessel(deg, z, jb);
    /* sum_k c_k 2 i^k j_k(a h); j_k(-z) = (-1)^k j_k(z) */
    double pr = 0, pi = 0;
    const double *ck = coef + (size_t)s * deg;
    for (int k = 0; k < deg; ++k) {
      double v = 2.0 * ck[k] * jb[k] * ((a < 0 && (k & 1)) ? -1.0 : 1.0);
      switch (k & 3) {
        case 0: pr += v; break;
        case 1: pi += v; break;
        case 2: pr -= v; break;
        default: pi -= v; break;
      }
    }
    /* times h exp(i a c) */
    const double er = cos(a * c), ei = sin(a * c);
    sr += h * (pr * er - pi * ei);
    si += h * (pr * ei + pi * er);
  }
  *re = sr;
  *im = si;
}

static double legendre_series(int deg, const double *c, double t) {
  /* Clenshaw for sum_k c_k P_k(t) */
  double b1 = 0, b2 = 0;
  for (int k = deg - 1; k >= 0; --k) {
    double alpha = (2.0 * k + 1) / (k + 1) * t, beta = -(double)(k + 1) / (k + 2);
    double b0 = c[k] + alpha * b1 + beta * b2;
    b2 = b1;
    b1 = b0;
  }
  /* P_0 = 1: result = b1 (with the recurrence P_{k+1} = (2k+1)/(k+1) t P_k - k/(k+1) P_{k-1}) */
  return b1;
}

static double eval_u(int nseg, int deg, const double *edges, const double *coef, double x) {
  int s = 0;
  while (s + 1 < nseg && x > edges[s + 1]) ++s;
  const double c = 0.5 * (edges[s] + edges[s + 1]), h = 0.5 * (edges[s + 1] - edges[s]);
  return legendre_series(deg, coef + (size_t)s * deg, (x - c) / h);
}

/*
 * statistics: 1 fermionic, 0 bosonic.  nl basis functions (even for
 * fermions, odd for bosons: the sampling frequencies are symmetric, the
 * bosonic ones about n = 0).  Outputs (caller
 * allocated): s[nl]; *nseg_out and edges[nseg + 1] of the segments of [-1, 1]
 * (at most max_seg); coef[nl][max_seg][deg] Legendre coefficients of u_l on
 * each segment; xsample[nl]; wsample[nl] Matsubara indices n (w = (2n + zeta)
 * pi / beta); uwl[2 * (n_idx * nl + l)] the complex u^_l at wsample[n_idx].
 * Returns 0, or 1 when nl exceeds the numerical rank, 2 on bad arguments,
 * 3 when the Matsubara sampling points cannot be found, 4 when u_{nl-1} does
 * not have nl - 1 zeros (nl beyond what the discretization resolves).
 */
int symgw_ir_basis(int statistics, double lambda, int nl, int max_seg, int deg, int *nseg_out, double *edges,
                        double *coef, double *s_out, double *xsample, int64_t *wsample, double *uwl) {
  const int fermi = statistics != 0;
  if (nl < 2 || lambda <= 0 || deg < 4 || deg > 48 || max_seg < 4) return 2;
  if ((fermi && (nl & 1)) || (!fermi && !(nl & 1))) return 2;
  const int gdeg = deg;
  qd *gx = malloc(sizeof(qd) * gdeg), *gw = malloc(sizeof(qd) * gdeg);
  gauss_legendre_q(gdeg, gx, gw);
  double ex[512], ey[512];
  const int nsx = half_edges(lambda, 1, 4, 2, ex);
  const int nsy = half_edges(lambda, 0, 4, 2, ey);
  if (2 * nsx > max_seg) {
    free(gx);
    free(gw);
    return 2;
  }
  const int m = nsx * gdeg, n = nsy * gdeg;
  qd *x = malloc(sizeof(qd) * m), *wx = malloc(sizeof(qd) * m);
  qd *y = malloc(sizeof(qd) * n), *wy = malloc(sizeof(qd) * n);
  composite(ex, nsx, gdeg, gx, gw, x, wx);
  composite(ey, nsy, gdeg, gx, gw, y, wy);
  const qd lam = lambda;
  const int rmax = nl / 2 + 24 < m ? nl / 2 + 24 : m;
  /* per parity: u at the x nodes (m x r) and singular values */
  qd *U[2] = {NULL, NULL};
  qd *S[2] = {NULL, NULL};
  int rank[2] = {0, 0};
  for (int par = 0; par < 2; ++par) {
    qd *A = malloc(sizeof(qd) * (size_t)m * n);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i) {
        qd kp = kernel_q(fermi, lam, x[i], y[j]), km = kernel_q(fermi, lam, -x[i], y[j]);
        A[(size_t)j * m + i] = sqrtq(wx[i]) * (par == 0 ? kp + km : kp - km) * sqrtq(wy[j]);
      }
    qd *V = malloc(sizeof(qd) * (size_t)m * rmax), *beta = malloc(sizeof(qd) * rmax);
    qd *R = malloc(sizeof(qd) * (size_t)rmax * n);
    int *perm = malloc(sizeof(int) * n);
    int r = rrqr_q(m, n, A, rmax, 1e-30Q, V, beta, R, perm);
    /* X = (R P^T)^T (n x r, column-major): X[j][i] = R[i][perm^-1(j)] */
    qd *X = malloc(sizeof(qd) * (size_t)n * r), *J = malloc(sizeof(qd) * (size_t)r * r);
    for (int i = 0; i < r; ++i)
      for (int c = 0; c < n; ++c) X[(size_t)i * n + perm[c]] = R[(size_t)i * n + c];
    jacobi_q(n, r, X, J);
    qd *sv = malloc(sizeof(qd) * r);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < r; ++i) {
      qd t = 0;
      for (int c = 0; c < n; ++c) t += X[(size_t)i * n + c] * X[(size_t)i * n + c];
      sv[i] = sqrtq(t);
    }
    /* left vectors: Q * J, Q = H_0 H_1 ... H_{r-1} applied to [J; 0] */
    qd *Ul = calloc((size_t)m * r, sizeof(qd));
    for (int c = 0; c < r; ++c)
      for (int i = 0; i < r; ++i) Ul[(size_t)c * m + i] = J[(size_t)c * r + i];
    for (int k = r - 1; k >= 0; --k) {
      const qd *v = V + (size_t)k * m;
#pragma omp parallel for schedule(static)
      for (int c = 0; c < r; ++c) {
        qd *col = Ul + (size_t)c * m;
        qd d = 0;
        for (int i = k; i < m; ++i) d += v[i] * col[i];
        d *= beta[k];
        for (int i = k; i < m; ++i) col[i] -= d * v[i];
      }
    }
    U[par] = Ul;
    S[par] = sv;
    rank[par] = r;
    free(A); free(V); free(beta); free(R); free(perm); free(X); free(J);
  }
  /* merge by singular value */
  sv_entry *all = malloc(sizeof(sv_entry) * (rank[0] + rank[1]));
  int na = 0;
  for (int par = 0; par < 2; ++par)
    for (int i = 0; i < rank[par]; ++i) all[na++] = (sv_entry){(double)S[par][i], par, i};
  qsort(all, na, sizeof(sv_entry), cmp_sv);
  int status = 0;
  if (na < nl) status = 1;
  /* segments of [-1, 1]: mirror of the [0, 1] x segments */
  const int nseg = 2 * nsx;
  for (int i = 0; i <= nsx; ++i) {
    edges[nsx - i] = -ex[i];
    edges[nsx + i] = ex[i];
  }
  *nseg_out = nseg;
  /* Legendre coefficients per segment: c_k = (2k + 1)/2 sum_i w_i u(x_i) P_k(t_i) */
  double *Pk = malloc(sizeof(double) * gdeg * gdeg);
  for (int i = 0; i < gdeg; ++i) {
    double t = (double)gx[i], p0 = 1, p1 = t;
    Pk[i * gdeg + 0] = 1;
    if (gdeg > 1) Pk[i * gdeg + 1] = t;
    for (int k = 2; k < gdeg; ++k) {
      double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
      Pk[i * gdeg + k] = p2;
    }
  }
  for (int l = 0; l < nl && l < na; ++l) {
    const sv_entry e = all[l];
    s_out[l] = e.s;
    const qd *col = U[e.parity] + (size_t)e.idx * m;
    /* sign: u(1) > 0 -> value at the last node of the last x segment */
    const double sg = (double)col[m - 1] >= 0 ? 1.0 : -1.0;
    const double psign = e.parity ? -1.0 : 1.0;
    for (int sgi = 0; sgi < nsx; ++sgi) {
      double *cp = coef + ((size_t)l * max_seg + nsx + sgi) * deg;   /* right half */
      double *cm = coef + ((size_t)l * max_seg + nsx - 1 - sgi) * deg; /* mirrored left half */
      for (int k = 0; k < deg; ++k) {
        double acc = 0;
        for (int i = 0; i < gdeg; ++i) {
          const int node = sgi * gdeg + i;
          const double u = sg * (double)(col[node] / sqrtq(wx[node])) / sqrt(2.0);
          acc += (double)gw[i] * u * Pk[i * gdeg + k];
        }
        cp[k] = 0.5 * (2 * k + 1) * acc;
        /* u(-x) = psign u(x): on the mirrored segment t -> -t, P_k(-t) = (-1)^k P_k(t) */
        cm[k] = cp[k] * psign * ((k & 1) ? -1.0 : 1.0);
      }
    }
  }
  free(Pk);
  if (status == 0) {
    /* tau sampling: midpoints of the zeros of u_{nl-1} with -1, 1 added.  The
     * zeros are searched on (0, 1) and mirrored (u_{nl-1} has parity
     * (-1)^(nl-1); an odd function also vanishes at x = 0, a segment edge). */
    const double *cl = coef + (size_t)(nl - 1) * max_seg * deg;
    const int odd_last = (nl - 1) & 1;
    double *zr = malloc(sizeof(double) * (nl + 8));
    int nz = 0;
    const int fine = 64;
    double xa = 1e-13, fa = eval_u(nseg, deg, edges, cl, xa);
    for (int sgi = nsx; sgi < nseg && nz <= nl; ++sgi) {
      const double a = edges[sgi], b = edges[sgi + 1];
      for (int i = 1; i <= fine; ++i) {
        const double xb = a + (b - a) * i / fine, fb = eval_u(nseg, deg, edges, cl, xb);
        if ((fa < 0) != (fb < 0) && fa != 0 && fb != 0) {
          double lo = xa, hi = xb, flo = fa;
          for (int it = 0; it < 200 && hi - lo > 1e-16 * (1 + fabs(lo)); ++it) {
            const double mid = 0.5 * (lo + hi), fm = eval_u(nseg, deg, edges, cl, mid);
            if ((fm < 0) == (flo < 0)) { lo = mid; flo = fm; } else hi = mid;
          }
          if (nz < nl + 8) zr[nz++] = 0.5 * (lo + hi);
        }
        xa = xb;
        fa = fb;
      }
    }
    const int ntot = 2 * nz + odd_last;
    if (ntot != nl - 1) status = 4;
    else {
      double *zeros = malloc(sizeof(double) * (ntot + 2));
      int o = 0;
      zeros[o++] = -1.0;
      for (int i = nz - 1; i >= 0; --i) zeros[o++] = -zr[i];
      if (odd_last) zeros[o++] = 0.0;
      for (int i = 0; i < nz; ++i) zeros[o++] = zr[i];
      zeros[o++] = 1.0;
      for (int i = 0; i < nl; ++i) xsample[i] = 0.5 * (zeros[i] + zeros[i + 1]);
      free(zeros);
    }
    free(zr);
  }
  if (status == 0) {
    /* Matsubara sampling: the discrete extrema of u^_{nl-1}(i w_n) over
     * n >= 0 (n where the function is above or below both neighbours: every
     * n while it alternates at low frequency, then one point per interval
     * between its sign changes and one after the last), mirrored to negative
     * frequencies (fermions n -> -n-1, bosons n -> -n with n = 0 once).
     * u^_{nl-1} is purely imaginary or purely real (the phase
     * exp(i pi (n + zeta/2)) cancels the alternation of the Fourier integral
     * of a function of definite parity): fermions even l imaginary, odd l
     * real; bosons the other way.  The fermionic n = 0 is compared with its
     * mirror n = -1 (f(-1) = f(0) for a real part, -f(0) for an imaginary one). */
    const int zeta = fermi ? 1 : 0;
    const int odd = (nl - 1) & 1;
    const int use_im = fermi ? !odd : odd;
    const double *cl = coef + (size_t)(nl - 1) * max_seg * deg;
    const int64_t nmax = (int64_t)(40.0 * lambda) + 2000;
    const int npos = fermi ? nl / 2 : (nl - 1) / 2;
    int64_t *ext = malloc(sizeof(int64_t) * (nl + 8));
    int ne = 0;
    double f[3] = {0, 0, 0};
    for (int64_t nn = 0; nn <= nmax + 1 && ne < npos + 2; ++nn) {
      double re, im;
      const double a = kPi * (nn + 0.5 * zeta);
      fourier_segments(nseg, deg, edges, cl, a, &re, &im);
      f[0] = f[1];
      f[1] = f[2];
      f[2] = use_im ? re * sin(a) + im * cos(a) : re * cos(a) - im * sin(a);
      if (nn == 0) {
        f[1] = fermi ? (use_im ? -f[2] : f[2]) : f[2];  /* the mirror of n = 0 */
        continue;
      }
      const int64_t c = nn - 1;  /* candidate with neighbours f[0], f[2] */
      if (c == 0 && !fermi) continue;  /* the bosonic n = 0 is added once below */
      const int is_ext = c == 0 ? ((f[1] >= f[0] && f[1] >= f[2]) || (f[1] <= f[0] && f[1] <= f[2]))
                                : ((f[1] > f[0] && f[1] >= f[2]) || (f[1] < f[0] && f[1] <= f[2]));
      if (is_ext) ext[ne++] = c;
    }
    if (ne < npos) status = 3;
    else {
      int o = 0;
      if (fermi) {
        for (int i = npos - 1; i >= 0; --i) wsample[o++] = -ext[i] - 1;
        for (int i = 0; i < npos; ++i) wsample[o++] = ext[i];
      } else {
        for (int i = npos - 1; i >= 0; --i) wsample[o++] = -ext[i];
        wsample[o++] = 0;
        for (int i = 0; i < npos; ++i) wsample[o++] = ext[i];
      }
      if (o != nl) status = 3;
    }
    free(ext);
  }
  if (status == 0) {
    /* u^_l(i w_n) = (1/sqrt 2) exp(i pi (n + zeta/2)) int exp(i pi (n + zeta/2) x) u_l(x) dx */
    const int zeta = fermi ? 1 : 0;
    for (int in = 0; in < nl; ++in) {
      const double a = kPi * (wsample[in] + 0.5 * zeta);
      const double er = cos(a), ei = sin(a);
      for (int l = 0; l < nl; ++l) {
        double re, im;
        fourier_segments(nseg, deg, edges, coef + (size_t)l * max_seg * deg, a, &re, &im);
        uwl[2 * ((size_t)in * nl + l)] = (re * er - im * ei) / sqrt(2.0);
        uwl[2 * ((size_t)in * nl + l) + 1] = (re * ei + im * er) / sqrt(2.0);
      }
    }
  }
  free(all);
  for (int p = 0; p < 2; ++p) { free(U[p]); free(S[p]); }
  free(x); free(wx); free(y); free(wy); free(gx); free(gw);
  return status;
}

/* u^_l(i w_n) of a basis returned by symgw_ir_basis at caller-given
 * Matsubara indices n[i] (w = (2 n + zeta) pi / beta): out[2 * (i * nl + l)]
 * (re, im).  statistics: 1 fermionic (zeta = 1), 0 bosonic (zeta = 0). */
int symgw_ir_uhat(int statistics, int nl, int nseg, int deg, int max_seg, const double *edges, const double *coef,
                  int nfreq, const int64_t *n, double *out) {
  if (nl < 1 || nseg < 1 || deg < 1 || deg > 48 || nseg > max_seg || nfreq < 0) return 2;
  const int zeta = statistics ? 1 : 0;
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < nfreq; ++i) {
    const double a = kPi * (n[i] + 0.5 * zeta);
    const double er = cos(a), ei = sin(a);
    for (int l = 0; l < nl; ++l) {
      double re, im;
      fourier_segments(nseg, deg, edges, coef + (size_t)l * max_seg * deg, a, &re, &im);
      out[2 * ((size_t)i * nl + l)] = (re * er - im * ei) / sqrt(2.0);
      out[2 * ((size_t)i * nl + l) + 1] = (re * ei + im * er) / sqrt(2.0);
    }
  }
  return 0;
}
