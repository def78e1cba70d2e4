/*
 * FP64 CPU oracle — TEST INFRASTRUCTURE ONLY (see oracle.h for the list of
 * reference functions restated here and for who may call it).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- derived tables (own derivation, symmetry_utils.cpp:96-191) ---------- */
typedef struct {
  int nk, ink, ns, nts;
  int *nb_ao, *nb_aux;          /* blocks per k                          */
  int **ao_sz, **aux_sz;        /* block sizes                           */
  int64_t **ao_off, **aux_off;  /* flat offsets inside the k slice       */
  int64_t *kao_off, *kao_sz, *qaux_off, *qaux_sz;          /* full BZ     */
  int64_t *kao_off_i, *qaux_off_i;                         /* iBZ         */
  int64_t flat_ao, flat_ao_full, flat_aux, flat_aux_full;
  int64_t *kp_off;              /* [nk*nk] V slice offset, -1 if absent  */
  int *kp_nb;                   /* [nk*nk]                               */
  int **kp_irr;                 /* [nk*nk] -> 3*nb                       */
  int64_t **kp_boff;            /* [nk*nk] -> block offsets in the slice */
  int **kp_q;                   /* [nk*nk] -> aux size of each block      */
  int64_t flat_V;
  int64_t *orep_koff, *auxrep_koff;  /* per k: start of its rep slice     */
  int64_t **orep_boff, **auxrep_boff;
} tables;

static void free_tables(tables *T) {
  int nk = T->nk;
  for (int k = 0; k < nk; ++k) {
    free(T->ao_sz[k]); free(T->aux_sz[k]); free(T->ao_off[k]); free(T->aux_off[k]);
    free(T->orep_boff[k]); free(T->auxrep_boff[k]);
  }
  for (int p = 0; p < nk * nk; ++p) { free(T->kp_irr[p]); free(T->kp_boff[p]); free(T->kp_q[p]); }
  free(T->nb_ao); free(T->nb_aux); free(T->ao_sz); free(T->aux_sz); free(T->ao_off); free(T->aux_off);
  free(T->kao_off); free(T->kao_sz); free(T->qaux_off); free(T->qaux_sz); free(T->kao_off_i); free(T->qaux_off_i);
  free(T->kp_off); free(T->kp_nb); free(T->kp_irr); free(T->kp_boff); free(T->kp_q);
  free(T->orep_koff); free(T->auxrep_koff); free(T->orep_boff); free(T->auxrep_boff);
}

static void make_tables(const orc_desc *d, tables *T) {
  int nk = d->nk;
  T->nk = nk; T->ink = d->ink; T->ns = d->ns; T->nts = d->nts;
  T->nb_ao = calloc(nk, sizeof(int)); T->nb_aux = calloc(nk, sizeof(int));
  T->ao_sz = calloc(nk, sizeof(int *)); T->aux_sz = calloc(nk, sizeof(int *));
  T->ao_off = calloc(nk, sizeof(int64_t *)); T->aux_off = calloc(nk, sizeof(int64_t *));
  T->kao_off = calloc(nk, 8); T->kao_sz = calloc(nk, 8); T->qaux_off = calloc(nk, 8); T->qaux_sz = calloc(nk, 8);
  int64_t pa = 0, px = 0, oa = 0, ox = 0;
  for (int k = 0; k < nk; ++k) {
    int na = d->kao_nblocks[k], nx = d->qaux_nblocks[k];
    T->nb_ao[k] = na; T->nb_aux[k] = nx;
    T->ao_sz[k] = malloc(sizeof(int) * na); T->ao_off[k] = malloc(8 * na);
    T->aux_sz[k] = malloc(sizeof(int) * nx); T->aux_off[k] = malloc(8 * nx);
    int64_t o = 0;
    for (int b = 0; b < na; ++b) { int s = d->ao_sizes[pa + b]; T->ao_sz[k][b] = s; T->ao_off[k][b] = o; o += (int64_t)s * s; }
    T->kao_off[k] = oa; T->kao_sz[k] = o; oa += o; pa += na;
    o = 0;
    for (int b = 0; b < nx; ++b) { int s = d->aux_sizes[px + b]; T->aux_sz[k][b] = s; T->aux_off[k][b] = o; o += (int64_t)s * s; }
    T->qaux_off[k] = ox; T->qaux_sz[k] = o; ox += o; px += nx;
  }
  T->flat_ao_full = oa; T->flat_aux_full = ox;
  T->kao_off_i = calloc(d->ink, 8); T->qaux_off_i = calloc(d->ink, 8);
  oa = 0; ox = 0;
  for (int i = 0; i < d->ink; ++i) {
    int k = d->irre_list[i];
    T->kao_off_i[i] = oa; oa += T->kao_sz[k];
    T->qaux_off_i[i] = ox; ox += T->qaux_sz[k];
  }
  T->flat_ao = oa; T->flat_aux = ox;
  /* V: the aux size of a block is that of q with k2 = k1 + q */
  T->kp_off = malloc(8 * (size_t)nk * nk); T->kp_nb = calloc((size_t)nk * nk, sizeof(int));
  T->kp_irr = calloc((size_t)nk * nk, sizeof(int *)); T->kp_boff = calloc((size_t)nk * nk, sizeof(int64_t *));
  T->kp_q = calloc((size_t)nk * nk, sizeof(int *));
  int64_t pv = 0, vo = 0;
  for (int p = 0; p < nk * nk; ++p) {
    T->kp_off[p] = -1;
    int nb = d->kpair_nblocks[p];
    T->kp_nb[p] = nb;
    if (nb == 0) continue;
    int k1 = p / nk, k2 = p % nk, q = -1;
    for (int qq = 0; qq < nk; ++qq) if (d->kpq[k1 * nk + qq] == k2) { q = qq; break; }
    T->kp_irr[p] = malloc(sizeof(int) * 3 * nb); T->kp_boff[p] = malloc(8 * nb); T->kp_q[p] = malloc(sizeof(int) * nb);
    int64_t o = 0;
    for (int b = 0; b < nb; ++b) {
      int iq = d->V_irreps[3 * (pv + b)], i1 = d->V_irreps[3 * (pv + b) + 1], i2 = d->V_irreps[3 * (pv + b) + 2];
      T->kp_irr[p][3 * b] = iq; T->kp_irr[p][3 * b + 1] = i1; T->kp_irr[p][3 * b + 2] = i2;
      T->kp_boff[p][b] = o;
      T->kp_q[p][b] = T->aux_sz[q][iq];
      o += (int64_t)T->aux_sz[q][iq] * T->ao_sz[k1][i1] * T->ao_sz[k2][i2];
    }
    pv += nb;
    T->kp_off[p] = vo; vo += o;
  }
  T->flat_V = vo;
  /* rotation blocks */
  T->orep_koff = calloc(nk, 8); T->auxrep_koff = calloc(nk, 8);
  T->orep_boff = calloc(nk, sizeof(int64_t *)); T->auxrep_boff = calloc(nk, sizeof(int64_t *));
  int64_t po = 0, pp = 0, ro = 0, rx = 0;
  for (int k = 0; k < nk; ++k) {
    int kr = d->irre_list[d->irre_index[k]];
    int nbo = d->orep_nblocks ? d->orep_nblocks[k] : 0, nbx = d->auxrep_nblocks ? d->auxrep_nblocks[k] : 0;
    T->orep_boff[k] = malloc(8 * (nbo + 1)); T->auxrep_boff[k] = malloc(8 * (nbx + 1));
    T->orep_koff[k] = ro;
    int64_t o = 0;
    for (int b = 0; b < nbo; ++b) {
      int a1 = d->orep_irreps[2 * (po + b)], a2 = d->orep_irreps[2 * (po + b) + 1];
      T->orep_boff[k][b] = o; o += (int64_t)T->ao_sz[k][a1] * T->ao_sz[kr][a2];
    }
    ro += o; po += nbo;
    T->auxrep_koff[k] = rx;
    o = 0;
    for (int b = 0; b < nbx; ++b) {
      int a1 = d->auxrep_irreps[2 * (pp + b)], a2 = d->auxrep_irreps[2 * (pp + b) + 1];
      T->auxrep_boff[k][b] = o; o += (int64_t)T->aux_sz[k][a1] * T->aux_sz[kr][a2];
    }
    rx += o; pp += nbx;
  }
}

int64_t orc_flat_ao(const orc_desc *d) { tables T; make_tables(d, &T); int64_t r = T.flat_ao; free_tables(&T); return r; }
int64_t orc_flat_ao_full(const orc_desc *d) { tables T; make_tables(d, &T); int64_t r = T.flat_ao_full; free_tables(&T); return r; }
int64_t orc_flat_aux(const orc_desc *d) { tables T; make_tables(d, &T); int64_t r = T.flat_aux; free_tables(&T); return r; }
int64_t orc_flat_aux_full(const orc_desc *d) { tables T; make_tables(d, &T); int64_t r = T.flat_aux_full; free_tables(&T); return r; }
int64_t orc_flat_V(const orc_desc *d) { tables T; make_tables(d, &T); int64_t r = T.flat_V; free_tables(&T); return r; }
int64_t orc_kpair_offset(const orc_desc *d, int k1, int k2) {
  tables T; make_tables(d, &T); int64_t r = T.kp_off[k1 * d->nk + k2]; free_tables(&T); return r;
}

/* ---- small complex GEMM: C (MxN) (+)= op(A) op(B), row-major --------------
 * ta/tb: 0 = as is, 1 = transpose, 2 = conjugate (no transpose), 3 = conj-transpose.
 * lda/ldb are the row lengths of A and B as stored. */
static void zgemm(int M, int N, int K, const double *A, int lda, int ta, const double *B, int ldb, int tb,
                  double *C, int ldc, int accumulate) {
  for (int i = 0; i < M; ++i) {
    double *c = C + 2 * (size_t)i * ldc;
    if (!accumulate) memset(c, 0, sizeof(double) * 2 * N);
    for (int k = 0; k < K; ++k) {
      const double *a = (ta & 1) ? A + 2 * ((size_t)k * lda + i) : A + 2 * ((size_t)i * lda + k);
      double ar = a[0], ai = (ta & 2) ? -a[1] : a[1];
      if (tb & 1) {
        for (int j = 0; j < N; ++j) {
          const double *b = B + 2 * ((size_t)j * ldb + k);
          double br = b[0], bi = (tb & 2) ? -b[1] : b[1];
          c[2 * j] += ar * br - ai * bi;
          c[2 * j + 1] += ar * bi + ai * br;
        }
      } else {
        const double *b = B + 2 * (size_t)k * ldb;
        if (tb & 2) {
          for (int j = 0; j < N; ++j) {
            double br = b[2 * j], bi = -b[2 * j + 1];
            c[2 * j] += ar * br - ai * bi;
            c[2 * j + 1] += ar * bi + ai * br;
          }
        } else {
          for (int j = 0; j < N; ++j) {
            double br = b[2 * j], bi = b[2 * j + 1];
            c[2 * j] += ar * br - ai * bi;
            c[2 * j + 1] += ar * bi + ai * br;
          }
        }
      }
    }
  }
}

/* ---- rotations (cugw_routines_block.cu:134-186, cugw_qkpt_block.cu:294-339) */
static void rotate_block(int m, int n, const double *O, const double *X, double *Y, int conj) {
  /* Y (m x m) = O (m x n) X (n x n) O^H */
  double *tmp = malloc(sizeof(double) * 2 * (size_t)m * n);
  zgemm(m, n, n, O, n, 0, X, n, 0, tmp, n, 0);
  zgemm(m, m, n, tmp, n, 0, O, n, 3, Y, m, 0);
  if (conj) for (int i = 0; i < m * m; ++i) Y[2 * i + 1] = -Y[2 * i + 1];
  free(tmp);
}

int orc_rotate_G(const orc_desc *d, const double *G_ibz, const double *orep, double *G_full) {
  tables T; make_tables(d, &T);
  const int nk = d->nk, nts_ns = d->nts * d->ns;
  for (int k = 0; k < nk; ++k) {
    int ck = d->conj_index[k], ick = d->irre_index[ck], kr = d->irre_list[ick];
    int cj = (k != ck);
    for (int ts = 0; ts < nts_ns; ++ts) {
      const double *src = G_ibz + 2 * ((size_t)ts * T.flat_ao + T.kao_off_i[ick]);
      double *dst = G_full + 2 * ((size_t)ts * T.flat_ao_full + T.kao_off[k]);
      if (d->index[ck] == ck) {
        for (int64_t x = 0; x < T.kao_sz[k]; ++x) { dst[2 * x] = src[2 * x]; dst[2 * x + 1] = cj ? -src[2 * x + 1] : src[2 * x + 1]; }
      } else {
        int nb = d->orep_nblocks[ck];
        int64_t pre = 0;
        for (int kk = 0; kk < ck; ++kk) pre += d->orep_nblocks[kk];
        for (int b = 0; b < nb; ++b) {
          int a1 = d->orep_irreps[2 * (pre + b)], a2 = d->orep_irreps[2 * (pre + b) + 1];
          int m = T.ao_sz[ck][a1], n = T.ao_sz[kr][a2];
          const double *O = orep + 2 * (T.orep_koff[ck] + T.orep_boff[ck][b]);
          rotate_block(m, n, O, src + 2 * T.ao_off[kr][a2], dst + 2 * T.ao_off[ck][a1], cj);
        }
      }
    }
  }
  free_tables(&T);
  return 0;
}

int orc_rotate_P(const orc_desc *d, const double *P_ibz, const double *auxrep, double *P_full) {
  tables T; make_tables(d, &T);
  const int nk = d->nk, nts = d->nts;
  for (int q = 0; q < nk; ++q) {
    int cq = d->conj_index[q], icq = d->irre_index[cq], qr = d->irre_list[icq];
    int cj = (q != cq);
    for (int t = 0; t < nts; ++t) {
      const double *src = P_ibz + 2 * ((size_t)nts * T.qaux_off_i[icq] + (size_t)t * T.qaux_sz[qr]);
      double *dst = P_full + 2 * ((size_t)nts * T.qaux_off[q] + (size_t)t * T.qaux_sz[q]);
      if (d->index[cq] == cq) {
        for (int64_t x = 0; x < T.qaux_sz[q]; ++x) { dst[2 * x] = src[2 * x]; dst[2 * x + 1] = cj ? -src[2 * x + 1] : src[2 * x + 1]; }
      } else {
        int nb = d->auxrep_nblocks[cq];
        int64_t pre = 0;
        for (int kk = 0; kk < cq; ++kk) pre += d->auxrep_nblocks[kk];
        for (int b = 0; b < nb; ++b) {
          int a1 = d->auxrep_irreps[2 * (pre + b)], a2 = d->auxrep_irreps[2 * (pre + b) + 1];
          int m = T.aux_sz[cq][a1], n = T.aux_sz[qr][a2];
          const double *O = auxrep + 2 * (T.auxrep_koff[cq] + T.auxrep_boff[cq][b]);
          rotate_block(m, n, O, src + 2 * T.aux_off[qr][a2], dst + 2 * T.aux_off[cq][a1], cj);
        }
      }
    }
  }
  free_tables(&T);
  return 0;
}

/* ---- P0 (gw_block.cpp:6-84) restricted to the iBZ q ------------------------ */
int orc_P0(const orc_desc *d, const double *V, const double *G_full, int t0, int t1, double *P0, int nthreads) {
  return orc_P0_sample(d, V, G_full, t0, t1, d->ink, d->nk, P0, nthreads);
}

int orc_P0_sample(const orc_desc *d, const double *V, const double *G_full, int t0, int t1, int nq, int nkk,
                  double *P0, int nthreads) {
  tables T; make_tables(d, &T);
  const int nk = d->nk, ns = d->ns, nts = d->nts;
  const double scale = -1.0 / nk * (ns == 1 ? 2.0 : 1.0);
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (int iqr = 0; iqr < nq; ++iqr) {
    const int q = d->irre_list[iqr];
    const int64_t qsl = T.qaux_sz[q];
    double *acc = calloc(2 * (size_t)qsl * (t1 - t0), sizeof(double));
    size_t cap = 0;
    double *Y = NULL, *X1 = NULL, *X2 = NULL;
    for (int k1 = 0; k1 < nkk; ++k1) {
      const int k2 = d->kpq[k1 * nk + q];
      const int p = k1 * nk + k2;
      if (T.kp_off[p] < 0) continue;
      const double *Vp = V + 2 * T.kp_off[p];
      for (int s = 0; s < ns; ++s) {
        for (int b = 0; b < T.kp_nb[p]; ++b) {
          const int iq = T.kp_irr[p][3 * b], i1 = T.kp_irr[p][3 * b + 1], i2 = T.kp_irr[p][3 * b + 2];
          const int qs = T.aux_sz[q][iq], a1 = T.ao_sz[k1][i1], a2 = T.ao_sz[k2][i2];
          const size_t K = (size_t)a1 * a2, sz = K * qs;
          if (sz > cap) {
            cap = sz;
            Y = realloc(Y, 16 * cap); X1 = realloc(X1, 16 * cap); X2 = realloc(X2, 16 * cap);
          }
          const double *Vb = Vp + 2 * T.kp_boff[p][b];
          /* Y (k1k2 x q) = V^T */
          for (int Q = 0; Q < qs; ++Q)
            for (size_t x = 0; x < K; ++x) { Y[2 * (x * qs + Q)] = Vb[2 * (Q * K + x)]; Y[2 * (x * qs + Q) + 1] = Vb[2 * (Q * K + x) + 1]; }
          for (int t = t0; t < t1; ++t) {
            const double *Gtp = G_full + 2 * (((size_t)(nts - 1 - t) * ns + s) * T.flat_ao_full + T.kao_off[k1] + T.ao_off[k1][i1]);
            const double *G1 = G_full + 2 * (((size_t)t * ns + s) * T.flat_ao_full + T.kao_off[k2] + T.ao_off[k2][i2]);
            /* X1 (k1 x k2 q) = G_tp (k1 x k1) * Y (k1 x k2 q) */
            zgemm(a1, a2 * qs, a1, Gtp, a1, 0, Y, a2 * qs, 0, X1, a2 * qs, 0);
            /* X2 (q k1 x k2) = conj(V) (q k1 x k2) * G1^T (k2 x k2) */
            zgemm(qs * a1, a2, a2, Vb, a2, 2, G1, a2, 1, X2, a2, 0);
            /* P0^T (q x q) += X2 (q x k1 k2) * X1 (k1 k2 x q):  P0[Q][P] += sum_x X2[P][x] X1[x][Q] */
            double *Pq = acc + 2 * ((size_t)(t - t0) * qsl + T.aux_off[q][iq]);
            for (int P = 0; P < qs; ++P)
              for (size_t x = 0; x < K; ++x) {
                const double xr = X2[2 * (P * K + x)], xi = X2[2 * (P * K + x) + 1];
                const double *x1 = X1 + 2 * (x * qs);
                for (int Q = 0; Q < qs; ++Q) {
                  double *o = Pq + 2 * ((size_t)Q * qs + P);
                  o[0] += xr * x1[2 * Q] - xi * x1[2 * Q + 1];
                  o[1] += xr * x1[2 * Q + 1] + xi * x1[2 * Q];
                }
              }
          }
        }
      }
    }
    for (int t = t0; t < t1; ++t) {
      double *dst = P0 + 2 * ((size_t)nts * T.qaux_off_i[iqr] + (size_t)t * qsl);
      const double *src = acc + 2 * ((size_t)(t - t0) * qsl);
      for (int64_t x = 0; x < 2 * qsl; ++x) dst[x] = scale * src[x];
    }
    free(acc); free(Y); free(X1); free(X2);
  }
  free_tables(&T);
  return 0;
}

/* ---- Sigma (gw_block.cpp:86-158) restricted to the iBZ k ------------------- */
int orc_Sigma(const orc_desc *d, const double *V, const double *G_full, const double *P_full, int t0, int t1,
              double *Sigma, int nthreads) {
  tables T; make_tables(d, &T);
  const int nk = d->nk, ns = d->ns, nts = d->nts;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (int ikr = 0; ikr < d->ink; ++ikr) {
    const int k1 = d->irre_list[ikr];
    const int64_t ksl = T.kao_sz[k1];
    double *acc = calloc(2 * (size_t)ksl * (t1 - t0) * ns, sizeof(double));
    size_t cap = 0;
    double *Y1 = NULL, *Y2 = NULL;
    for (int q = 0; q < nk; ++q) {
      const int k2 = d->kmq[k1 * nk + q];
      const int p = k1 * nk + k2;
      if (T.kp_off[p] < 0) continue;
      const double *Vp = V + 2 * T.kp_off[p];
      for (int s = 0; s < ns; ++s) {
        for (int b = 0; b < T.kp_nb[p]; ++b) {
          const int iq = T.kp_irr[p][3 * b], i1 = T.kp_irr[p][3 * b + 1], i2 = T.kp_irr[p][3 * b + 2];
          const int qs = T.aux_sz[q][iq], a1 = T.ao_sz[k1][i1], a2 = T.ao_sz[k2][i2];
          const size_t K = (size_t)a1 * a2, sz = K * qs;
          if (sz > cap) { cap = sz; Y1 = realloc(Y1, 16 * cap); Y2 = realloc(Y2, 16 * cap); }
          const double *Vb = Vp + 2 * T.kp_boff[p][b];
          for (int t = t0; t < t1; ++t) {
            const double *G1 = G_full + 2 * (((size_t)t * ns + s) * T.flat_ao_full + T.kao_off[k2] + T.ao_off[k2][i2]);
            const double *Pqp = P_full + 2 * ((size_t)nts * T.qaux_off[q] + (size_t)t * T.qaux_sz[q] + T.aux_off[q][iq]);
            /* Y1 (q k1 x k2) = V (q k1 x k2) * G1 (k2 x k2) */
            zgemm(qs * a1, a2, a2, Vb, a2, 0, G1, a2, 0, Y1, a2, 0);
            /* Y2 (k1 k2 x q) = Y1^T (k1k2 x q) * P (q x q) */
            zgemm((int)K, qs, qs, Y1, (int)K, 1, Pqp, qs, 0, Y2, qs, 0);
            /* Sigma (k1 x k1) += Y2 (k1 x k2 q) * V_nPj (k2 q x k1), V_nPj[(n,P)][j] = conj(V[P][j][n]) */
            double *Sg = acc + 2 * (((size_t)(t - t0) * ns + s) * ksl + T.ao_off[k1][i1]);
            for (int i = 0; i < a1; ++i)
              for (int n = 0; n < a2; ++n)
                for (int P = 0; P < qs; ++P) {
                  const double yr = Y2[2 * ((size_t)(i * a2 + n) * qs + P)], yi = Y2[2 * ((size_t)(i * a2 + n) * qs + P) + 1];
                  const double *vrow = Vb + 2 * ((size_t)P * K);
                  for (int j = 0; j < a1; ++j) {
                    const double vr = vrow[2 * (j * a2 + n)], vi = -vrow[2 * (j * a2 + n) + 1];
                    Sg[2 * (i * a1 + j)] += yr * vr - yi * vi;
                    Sg[2 * (i * a1 + j) + 1] += yr * vi + yi * vr;
                  }
                }
          }
        }
      }
    }
    for (int t = t0; t < t1; ++t)
      for (int s = 0; s < ns; ++s) {
        double *dst = Sigma + 2 * (((size_t)t * ns + s) * T.flat_ao + T.kao_off_i[ikr]);
        const double *src = acc + 2 * (((size_t)(t - t0) * ns + s) * ksl);
        for (int64_t x = 0; x < 2 * ksl; ++x) dst[x] = -src[x] / nk;
      }
    free(acc); free(Y1); free(Y2);
  }
  free_tables(&T);
  return 0;
}

/* ---- single (q, t) / (k, t, s) slices summed over the whole mesh ----------
 * The same loop nests as orc_P0 / orc_Sigma (gw_block.cpp:6-84, :86-158) for
 * one output slice, with the k (P0) or q (Sigma) sum split over threads in
 * contiguous ranges and the partial sums added in thread order.  V is given
 * per pair: the slice of pair p = k1*nk + k2 starts at V + 2*voff[p]. */
static int nthreads_of(int n) {
#ifdef _OPENMP
  return n < 1 ? 1 : n;
#else
  (void)n;
  return 1;
#endif
}

int orc_P0_at(const orc_desc *d, const double *V, const int64_t *voff, const double *G_full, int iqr, int t,
              double *out, int nthreads) {
  tables T; make_tables(d, &T);
  const int nk = d->nk, ns = d->ns, nts = d->nts;
  if (iqr < 0 || iqr >= d->ink || t < 0 || t >= nts) { free_tables(&T); return 1; }
  const int q = d->irre_list[iqr];
  const int64_t qsl = T.qaux_sz[q];
  const double scale = -1.0 / nk * (ns == 1 ? 2.0 : 1.0);
  const int nth = nthreads_of(nthreads);
  double *part = calloc(2 * (size_t)qsl * nth, sizeof(double));
  int bad = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nth) reduction(| : bad)
#endif
  {
#ifdef _OPENMP
    const int th = omp_get_thread_num(), nt = omp_get_num_threads();
#else
    const int th = 0, nt = 1;
#endif
    double *acc = part + 2 * (size_t)qsl * th;
    size_t cap = 0;
    double *X1 = NULL, *X2 = NULL;
    const int kb = (int)((int64_t)nk * th / nt), ke = (int)((int64_t)nk * (th + 1) / nt);
    for (int k1 = kb; k1 < ke; ++k1) {
      const int k2 = d->kpq[k1 * nk + q];
      const int p = k1 * nk + k2;
      if (T.kp_off[p] < 0) continue;
      if (voff[p] < 0) { bad = 1; continue; }
      const double *Vp = V + 2 * voff[p];
      for (int s = 0; s < ns; ++s)
        for (int b = 0; b < T.kp_nb[p]; ++b) {
          const int iq = T.kp_irr[p][3 * b], i1 = T.kp_irr[p][3 * b + 1], i2 = T.kp_irr[p][3 * b + 2];
          const int qs = T.aux_sz[q][iq], a1 = T.ao_sz[k1][i1], a2 = T.ao_sz[k2][i2];
          const size_t K = (size_t)a1 * a2, sz = K * qs;
          if (sz > cap) { cap = sz; X1 = realloc(X1, 16 * cap); X2 = realloc(X2, 16 * cap); }
          const double *Vb = Vp + 2 * T.kp_boff[p][b];
          const double *Gtp = G_full + 2 * (((size_t)(nts - 1 - t) * ns + s) * T.flat_ao_full + T.kao_off[k1] + T.ao_off[k1][i1]);
          const double *G1 = G_full + 2 * (((size_t)t * ns + s) * T.flat_ao_full + T.kao_off[k2] + T.ao_off[k2][i2]);
          /* X1[Q][p][m] = sum_a Gtp[p][a] V[Q][a][m]  (G(beta - t) V) */
          for (int Q = 0; Q < qs; ++Q) {
            zgemm(a1, a2, a1, Gtp, a1, 0, Vb + 2 * (size_t)Q * K, a2, 0, X2, a2, 0);
            memcpy(X1 + 2 * (size_t)Q * K, X2, 16 * K);
          }
          /* Y[P][p][n] = sum_m conj(V[P][p][m]) G1[n][m]  (conj(V) G(t)^T, as orc_P0_sample) */
          double *Y = X2;
          zgemm(qs * a1, a2, a2, Vb, a2, 2, G1, a2, 1, Y, a2, 0);
          /* P0[Q][P] += sum_x Y[P][x] X1[Q][x] over x = (p, m) */
          for (int Q = 0; Q < qs; ++Q)
            for (int P = 0; P < qs; ++P) {
              double re = 0, im = 0;
              const double *x1 = X1 + 2 * (size_t)Q * K, *x2 = Y + 2 * (size_t)P * K;
              for (size_t x = 0; x < K; ++x) {
                re += x2[2 * x] * x1[2 * x] - x2[2 * x + 1] * x1[2 * x + 1];
                im += x2[2 * x] * x1[2 * x + 1] + x2[2 * x + 1] * x1[2 * x];
              }
              double *o = acc + 2 * (T.aux_off[q][iq] + (size_t)Q * qs + P);
              o[0] += re;
              o[1] += im;
            }
        }
    }
    free(X1); free(X2);
  }
  for (int64_t x = 0; x < 2 * qsl; ++x) {
    double v = 0;
    for (int th = 0; th < nth; ++th) v += part[2 * (size_t)qsl * th + x];
    out[x] = scale * v;
  }
  free(part);
  free_tables(&T);
  return bad ? 2 : 0;
}

int orc_rotate_P_t(const orc_desc *d, const double *P_ibz, const double *auxrep, int t, double *P_full_t) {
  tables T; make_tables(d, &T);
  const int nk = d->nk, nts = d->nts;
  for (int q = 0; q < nk; ++q) {
    int cq = d->conj_index[q], icq = d->irre_index[cq], qr = d->irre_list[icq];
    int cj = (q != cq);
    const double *src = P_ibz + 2 * ((size_t)nts * T.qaux_off_i[icq] + (size_t)t * T.qaux_sz[qr]);
    double *dst = P_full_t + 2 * T.qaux_off[q];
    if (d->index[cq] == cq) {
      for (int64_t x = 0; x < T.qaux_sz[q]; ++x) { dst[2 * x] = src[2 * x]; dst[2 * x + 1] = cj ? -src[2 * x + 1] : src[2 * x + 1]; }
    } else {
      int nb = d->auxrep_nblocks[cq];
      int64_t pre = 0;
      for (int kk = 0; kk < cq; ++kk) pre += d->auxrep_nblocks[kk];
      for (int b = 0; b < nb; ++b) {
        int a1 = d->auxrep_irreps[2 * (pre + b)], a2 = d->auxrep_irreps[2 * (pre + b) + 1];
        int m = T.aux_sz[cq][a1], n = T.aux_sz[qr][a2];
        const double *O = auxrep + 2 * (T.auxrep_koff[cq] + T.auxrep_boff[cq][b]);
        rotate_block(m, n, O, src + 2 * T.aux_off[qr][a2], dst + 2 * T.aux_off[cq][a1], cj);
      }
    }
  }
  free_tables(&T);
  return 0;
}

int orc_Sigma_at(const orc_desc *d, const double *V, const int64_t *voff, const double *G_full, const double *P_full_t,
                 int ikr, int t, int s, double *out, int nthreads) {
  tables T; make_tables(d, &T);
  const int nk = d->nk, ns = d->ns, nts = d->nts;
  if (ikr < 0 || ikr >= d->ink || t < 0 || t >= nts || s < 0 || s >= ns) { free_tables(&T); return 1; }
  const int k1 = d->irre_list[ikr];
  const int64_t ksl = T.kao_sz[k1];
  const int nth = nthreads_of(nthreads);
  double *part = calloc(2 * (size_t)ksl * nth, sizeof(double));
  int bad = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nth) reduction(| : bad)
#endif
  {
#ifdef _OPENMP
    const int th = omp_get_thread_num(), nt = omp_get_num_threads();
#else
    const int th = 0, nt = 1;
#endif
    double *acc = part + 2 * (size_t)ksl * th;
    size_t cap = 0;
    double *Y1 = NULL, *Y2 = NULL;
    const int qb = (int)((int64_t)nk * th / nt), qe = (int)((int64_t)nk * (th + 1) / nt);
    for (int q = qb; q < qe; ++q) {
      const int k2 = d->kmq[k1 * nk + q];
      const int p = k1 * nk + k2;
      if (T.kp_off[p] < 0) continue;
      if (voff[p] < 0) { bad = 1; continue; }
      const double *Vp = V + 2 * voff[p];
      for (int b = 0; b < T.kp_nb[p]; ++b) {
        const int iq = T.kp_irr[p][3 * b], i1 = T.kp_irr[p][3 * b + 1], i2 = T.kp_irr[p][3 * b + 2];
        const int qs = T.aux_sz[q][iq], a1 = T.ao_sz[k1][i1], a2 = T.ao_sz[k2][i2];
        const size_t K = (size_t)a1 * a2, sz = K * qs;
        if (sz > cap) { cap = sz; Y1 = realloc(Y1, 16 * cap); Y2 = realloc(Y2, 16 * cap); }
        const double *Vb = Vp + 2 * T.kp_boff[p][b];
        const double *G1 = G_full + 2 * (((size_t)t * ns + s) * T.flat_ao_full + T.kao_off[k2] + T.ao_off[k2][i2]);
        const double *Pqp = P_full_t + 2 * (T.qaux_off[q] + T.aux_off[q][iq]);
        /* Y1 (q k1 x k2) = V G1;  Y2 (k1 k2 x q) = Y1^T P;  Sigma += Y2 conj(V) (gw_block.cpp:86-158) */
        zgemm(qs * a1, a2, a2, Vb, a2, 0, G1, a2, 0, Y1, a2, 0);
        zgemm((int)K, qs, qs, Y1, (int)K, 1, Pqp, qs, 0, Y2, qs, 0);
        double *Sg = acc + 2 * T.ao_off[k1][i1];
        for (int i = 0; i < a1; ++i)
          for (int n = 0; n < a2; ++n)
            for (int P = 0; P < qs; ++P) {
              const double yr = Y2[2 * ((size_t)(i * a2 + n) * qs + P)], yi = Y2[2 * ((size_t)(i * a2 + n) * qs + P) + 1];
              const double *vrow = Vb + 2 * ((size_t)P * K);
              for (int j = 0; j < a1; ++j) {
                const double vr = vrow[2 * (j * a2 + n)], vi = -vrow[2 * (j * a2 + n) + 1];
                Sg[2 * (i * a1 + j)] += yr * vr - yi * vi;
                Sg[2 * (i * a1 + j) + 1] += yr * vi + yi * vr;
              }
            }
      }
    }
    free(Y1); free(Y2);
  }
  for (int64_t x = 0; x < 2 * ksl; ++x) {
    double v = 0;
    for (int th = 0; th < nth; ++th) v += part[2 * (size_t)ksl * th + x];
    out[x] = -v / nk;
  }
  free(part);
  free_tables(&T);
  return bad ? 2 : 0;
}

/* ---- dense inverses -------------------------------------------------------- */
int orc_inverse_lu(int n, const double *A, double *Ainv) {
  /* LU with partial pivoting (Doolittle), then solve for each unit vector */
  double *LU = malloc(sizeof(double) * 2 * (size_t)n * n);
  int *piv = malloc(sizeof(int) * n);
  memcpy(LU, A, sizeof(double) * 2 * (size_t)n * n);
  for (int k = 0; k < n; ++k) {
    int p = k; double best = -1;
    for (int i = k; i < n; ++i) {
      double v = LU[2 * (i * n + k)] * LU[2 * (i * n + k)] + LU[2 * (i * n + k) + 1] * LU[2 * (i * n + k) + 1];
      if (v > best) { best = v; p = i; }
    }
    piv[k] = p;
    if (best == 0) { free(LU); free(piv); return 1; }
    if (p != k)
      for (int j = 0; j < n; ++j) {
        double a = LU[2 * (k * n + j)], b = LU[2 * (k * n + j) + 1];
        LU[2 * (k * n + j)] = LU[2 * (p * n + j)]; LU[2 * (k * n + j) + 1] = LU[2 * (p * n + j) + 1];
        LU[2 * (p * n + j)] = a; LU[2 * (p * n + j) + 1] = b;
      }
    const double dr = LU[2 * (k * n + k)], di = LU[2 * (k * n + k) + 1], dd = dr * dr + di * di;
    for (int i = k + 1; i < n; ++i) {
      double ar = LU[2 * (i * n + k)], ai = LU[2 * (i * n + k) + 1];
      double lr = (ar * dr + ai * di) / dd, li = (ai * dr - ar * di) / dd;
      LU[2 * (i * n + k)] = lr; LU[2 * (i * n + k) + 1] = li;
      for (int j = k + 1; j < n; ++j) {
        double ur = LU[2 * (k * n + j)], ui = LU[2 * (k * n + j) + 1];
        LU[2 * (i * n + j)] -= lr * ur - li * ui;
        LU[2 * (i * n + j) + 1] -= lr * ui + li * ur;
      }
    }
  }
  double *col = malloc(sizeof(double) * 2 * n);
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) { col[2 * i] = 0; col[2 * i + 1] = 0; }
    col[2 * c] = 1.0;
    /* apply the row permutation in order */
    for (int k = 0; k < n; ++k) {
      int p = piv[k];
      if (p != k) { double a = col[2 * k], b = col[2 * k + 1]; col[2 * k] = col[2 * p]; col[2 * k + 1] = col[2 * p + 1]; col[2 * p] = a; col[2 * p + 1] = b; }
    }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) {
        double lr = LU[2 * (i * n + j)], li = LU[2 * (i * n + j) + 1];
        col[2 * i] -= lr * col[2 * j] - li * col[2 * j + 1];
        col[2 * i + 1] -= lr * col[2 * j + 1] + li * col[2 * j];
      }
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) {
        double ur = LU[2 * (i * n + j)], ui = LU[2 * (i * n + j) + 1];
        col[2 * i] -= ur * col[2 * j] - ui * col[2 * j + 1];
        col[2 * i + 1] -= ur * col[2 * j + 1] + ui * col[2 * j];
      }
      double dr = LU[2 * (i * n + i)], di = LU[2 * (i * n + i) + 1], dd = dr * dr + di * di;
      double xr = col[2 * i], xi = col[2 * i + 1];
      col[2 * i] = (xr * dr + xi * di) / dd; col[2 * i + 1] = (xi * dr - xr * di) / dd;
    }
    for (int i = 0; i < n; ++i) { Ainv[2 * (i * n + c)] = col[2 * i]; Ainv[2 * (i * n + c) + 1] = col[2 * i + 1]; }
  }
  free(col); free(LU); free(piv);
  return 0;
}

int orc_inverse_chol(int n, const double *A, double *Ainv) {
  /* A = L L^H (Hermitian positive definite), A^-1 = L^-H L^-1 */
  double *L = calloc(2 * (size_t)n * n, sizeof(double));
  for (int j = 0; j < n; ++j) {
    double s = A[2 * (j * n + j)];
    for (int k = 0; k < j; ++k) s -= L[2 * (j * n + k)] * L[2 * (j * n + k)] + L[2 * (j * n + k) + 1] * L[2 * (j * n + k) + 1];
    if (!(s > 0)) { free(L); return 1; }
    double ljj = sqrt(s);
    L[2 * (j * n + j)] = ljj;
    for (int i = j + 1; i < n; ++i) {
      double zr = A[2 * (i * n + j)], zi = A[2 * (i * n + j) + 1];
      for (int k = 0; k < j; ++k) {
        double ar = L[2 * (i * n + k)], ai = L[2 * (i * n + k) + 1];
        double br = L[2 * (j * n + k)], bi = -L[2 * (j * n + k) + 1];
        zr -= ar * br - ai * bi; zi -= ar * bi + ai * br;
      }
      L[2 * (i * n + j)] = zr / ljj; L[2 * (i * n + j) + 1] = zi / ljj;
    }
  }
  /* Linv (lower) by forward substitution */
  double *Li = calloc(2 * (size_t)n * n, sizeof(double));
  for (int c = 0; c < n; ++c) {
    for (int i = c; i < n; ++i) {
      double zr = (i == c) ? 1.0 : 0.0, zi = 0.0;
      for (int k = c; k < i; ++k) {
        double ar = L[2 * (i * n + k)], ai = L[2 * (i * n + k) + 1];
        double br = Li[2 * (k * n + c)], bi = Li[2 * (k * n + c) + 1];
        zr -= ar * br - ai * bi; zi -= ar * bi + ai * br;
      }
      double d = L[2 * (i * n + i)];
      Li[2 * (i * n + c)] = zr / d; Li[2 * (i * n + c) + 1] = zi / d;
    }
  }
  /* Ainv = Li^H Li */
  zgemm(n, n, n, Li, n, 3, Li, n, 0, Ainv, n, 0);
  free(L); free(Li);
  return 0;
}

/* ---- W (cugw_solver_block.cpp:385-439) ------------------------------------- */
int orc_W(const orc_desc *d, const double *P0, const double *Tnt_BF, const double *Ttn_FB, double *P) {
  tables T; make_tables(d, &T);
  const int nts = d->nts, nwb = d->nw_b, ni = nts - 2;
  int rc = 0;
  for (int iqr = 0; iqr < d->ink; ++iqr) {
    const int q = d->irre_list[iqr];
    const int64_t qsl = T.qaux_sz[q];
    const double *src = P0 + 2 * ((size_t)nts * T.qaux_off_i[iqr]);
    double *Pt = malloc(sizeof(double) * 2 * (size_t)nts * qsl);
    for (int t = 0; t < nts / 2; ++t) {
      memcpy(Pt + 2 * (size_t)t * qsl, src + 2 * (size_t)t * qsl, sizeof(double) * 2 * qsl);
      memcpy(Pt + 2 * (size_t)(nts - 1 - t) * qsl, src + 2 * (size_t)t * qsl, sizeof(double) * 2 * qsl);
    }
    double *Pw = malloc(sizeof(double) * 2 * (size_t)nwb * qsl);
    zgemm(nwb, (int)qsl, ni, Tnt_BF, ni, 0, Pt + 2 * qsl, (int)qsl, 0, Pw, (int)qsl, 0);
    for (int n = 0; n < nwb; ++n)
      for (int b = 0; b < T.nb_aux[q]; ++b) {
        const int m = T.aux_sz[q][b];
        double *X = Pw + 2 * ((size_t)n * qsl + T.aux_off[q][b]);
        double *A = malloc(sizeof(double) * 2 * m * m), *Ai = malloc(sizeof(double) * 2 * m * m);
        double *Y = malloc(sizeof(double) * 2 * m * m);
        for (int i = 0; i < m; ++i)
          for (int j = 0; j < m; ++j) {
            A[2 * (i * m + j)] = (i == j ? 1.0 : 0.0) - 0.5 * (X[2 * (i * m + j)] + X[2 * (j * m + i)]);
            A[2 * (i * m + j) + 1] = -0.5 * (X[2 * (i * m + j) + 1] - X[2 * (j * m + i) + 1]);
          }
        if (orc_inverse_chol(m, A, Ai)) rc = 1;
        for (int i = 0; i < m; ++i)
          for (int j = i; j < m; ++j) {
            double hr = 0.5 * (Ai[2 * (i * m + j)] + Ai[2 * (j * m + i)]);
            double hi = 0.5 * (Ai[2 * (i * m + j) + 1] - Ai[2 * (j * m + i) + 1]);
            Ai[2 * (i * m + j)] = hr; Ai[2 * (i * m + j) + 1] = hi;
            Ai[2 * (j * m + i)] = hr; Ai[2 * (j * m + i) + 1] = -hi;
          }
        zgemm(m, m, m, Ai, m, 0, X, m, 0, Y, m, 0);
        for (int i = 0; i < m; ++i) memcpy(X + 2 * (size_t)i * m, Y + 2 * (size_t)i * m, sizeof(double) * 2 * m);
        free(A); free(Ai); free(Y);
      }
    double *dst = P + 2 * ((size_t)nts * T.qaux_off_i[iqr]);
    zgemm(nts, (int)qsl, nwb, Ttn_FB, nwb, 0, Pw, (int)qsl, 0, dst, (int)qsl, 0);
    free(Pt); free(Pw);
  }
  free_tables(&T);
  return rc;
}

/* ---- Dyson (dyson_block.cpp:20-66 / sc_type_block.cpp:136-175) ------------- */
int orc_dyson(const orc_desc *d, const double *Sigma, const double *F, const double *S, double mu,
              const double *Tnt, const double *Ttn, const double *iwn, double *G) {
  tables T; make_tables(d, &T);
  const int nts = d->nts, ns = d->ns, nw = d->nw, ni = nts - 2;
  int rc = 0;
  for (int s = 0; s < ns; ++s)
    for (int ikr = 0; ikr < d->ink; ++ikr) {
      const int k = d->irre_list[ikr];
      const int64_t ksl = T.kao_sz[k], koff = T.kao_off_i[ikr];
      double *Sk = malloc(sizeof(double) * 2 * (size_t)nts * ksl);
      for (int t = 0; t < nts; ++t)
        memcpy(Sk + 2 * (size_t)t * ksl, Sigma + 2 * (((size_t)t * ns + s) * T.flat_ao + koff), sizeof(double) * 2 * ksl);
      double *Sw = malloc(sizeof(double) * 2 * (size_t)nw * ksl);
      zgemm(nw, (int)ksl, ni, Tnt, ni, 0, Sk + 2 * ksl, (int)ksl, 0, Sw, (int)ksl, 0);
      const double *Fk = F + 2 * ((size_t)s * T.flat_ao + koff), *Sk_ = S + 2 * ((size_t)s * T.flat_ao + koff);
      for (int w = 0; w < nw; ++w) {
        const double zr = iwn[2 * w] + mu, zi = iwn[2 * w + 1];
        for (int b = 0; b < T.nb_ao[k]; ++b) {
          const int m = T.ao_sz[k][b];
          const int64_t o = T.ao_off[k][b];
          double *A = malloc(sizeof(double) * 2 * m * m);
          for (int e = 0; e < m * m; ++e) {
            const double sr = Sk_[2 * (o + e)], si = Sk_[2 * (o + e) + 1];
            A[2 * e] = zr * sr - zi * si - Fk[2 * (o + e)] - Sw[2 * ((size_t)w * ksl + o + e)];
            A[2 * e + 1] = zr * si + zi * sr - Fk[2 * (o + e) + 1] - Sw[2 * ((size_t)w * ksl + o + e) + 1];
          }
          if (orc_inverse_lu(m, A, Sw + 2 * ((size_t)w * ksl + o))) rc = 1;
          free(A);
        }
      }
      double *Gt = malloc(sizeof(double) * 2 * (size_t)nts * ksl);
      zgemm(nts, (int)ksl, nw, Ttn, nw, 0, Sw, (int)ksl, 0, Gt, (int)ksl, 0);
      for (int t = 0; t < nts; ++t)
        memcpy(G + 2 * (((size_t)t * ns + s) * T.flat_ao + koff), Gt + 2 * (size_t)t * ksl, sizeof(double) * 2 * ksl);
      free(Sk); free(Sw); free(Gt);
    }
  free_tables(&T);
  return rc;
}

double orc_density(const orc_desc *d, const double *G, const double *S, const double *kw) {
  tables T; make_tables(d, &T);
  const int nts = d->nts, ns = d->ns;
  double tot = 0;
  for (int s = 0; s < ns; ++s)
    for (int ikr = 0; ikr < d->ink; ++ikr) {
      const int k = d->irre_list[ikr];
      for (int b = 0; b < T.nb_ao[k]; ++b) {
        const int m = T.ao_sz[k][b];
        const double *Gb = G + 2 * (((size_t)(nts - 1) * ns + s) * T.flat_ao + T.kao_off_i[ikr] + T.ao_off[k][b]);
        const double *Sb = S + 2 * ((size_t)s * T.flat_ao + T.kao_off_i[ikr] + T.ao_off[k][b]);
        for (int i = 0; i < m; ++i)
          for (int j = 0; j < m; ++j)
            tot += kw[ikr] * (Gb[2 * (i * m + j)] * Sb[2 * (j * m + i)] - Gb[2 * (i * m + j) + 1] * Sb[2 * (j * m + i) + 1]);
      }
    }
  free_tables(&T);
  return -(2.0 / ns) * tot / d->nk;
}
