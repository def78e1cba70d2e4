/*
 * FP64 CPU oracle of the symmetry-adapted GW hot path — TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
 * library, and only as the checker / the timed CPU restatement; the product
 * path (libsymgw) never calls it.
 *
 * It restates the reference's CPU algorithms (paths relative to the
 * CQMP/MBSymmetry tree):
 *   P0, Sigma block contractions   SymmetryMBPT/src/cpu_ref/gw_block.cpp:6-158
 *   iBZ -> full-BZ G / P rotation  SymmetryMBPT/src/cuda/cugw_routines_block.cu:100-201,
 *                                  SymmetryMBPT/src/cuda/cugw_qkpt_block.cu:287-340
 *   W (screened interaction)       SymmetryMBPT/src/solver/cugw_solver_block.cpp:385-439
 *   Dyson                          SymmetryMBPT/src/cpu_ref/dyson_block.cpp:20-66
 *   block/dense scatter            SymmetryMBPT/src/common/block_utils.cpp:6-143
 * with its own offsets (cumulative b*b, as symmetry_utils.cpp:96-191 reads
 * them) so that a mistake in the product's layout code is caught.
 *
 * The factorizations differ from the reference's Eigen calls (LDLT for W,
 * FullPivLU for Dyson): W uses a Cholesky factorization of the Hermitian
 * positive-definite I - (P0 + P0^H)/2 and Dyson an LU with partial pivoting.
 * Both are pinned against numpy.linalg and residual checks in tests/.
 *
 * Complex arrays are interleaved (re, im) doubles.  Layout names as in
 * include/symgw.h; "ts_kij_full" is ts_kij over the full Brillouin zone.
 */
#ifndef SYMGW_ORACLE_H
#define SYMGW_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same memory layout as symgw_layout_desc. */
typedef struct orc_desc {
  int32_t nk, ink, ns, nts, nw, nw_b;
  const int32_t *kao_nblocks, *ao_sizes, *qaux_nblocks, *aux_sizes;
  const int32_t *irre_list, *index, *irre_index, *conj_index, *kpq, *kmq;
  const int32_t *orep_nblocks, *orep_irreps, *auxrep_nblocks, *auxrep_irreps;
  const int32_t *kpair_nblocks, *V_irreps;
} orc_desc;

/* Sizes derived by the oracle itself. */
int64_t orc_flat_ao(const orc_desc *d);        /* iBZ        */
int64_t orc_flat_ao_full(const orc_desc *d);
int64_t orc_flat_aux(const orc_desc *d);       /* iBZ        */
int64_t orc_flat_aux_full(const orc_desc *d);
int64_t orc_flat_V(const orc_desc *d);
int64_t orc_kpair_offset(const orc_desc *d, int k1, int k2);

/* G_ibz (ts_kij) -> G_full (ts_kij_full): copy / O G O^H / conj. */
int orc_rotate_G(const orc_desc *d, const double *G_ibz, const double *orep, double *G_full);
/* P_ibz (q_t_QP) -> P_full (q_t_QP over the full BZ). */
int orc_rotate_P(const orc_desc *d, const double *P_ibz, const double *auxrep, double *P_full);

/* P0 for the iBZ q and t in [t0, t1), written into q_t_QP (other rows kept).
 * Scaled as the CPU reference: -(2/ns)/nk (gw_block.cpp:80-81). nthreads > 1
 * parallelises over q (deterministic). */
int orc_P0(const orc_desc *d, const double *V, const double *G_full, int t0, int t1, double *P0, int nthreads);
/* Bounded sample of the same loop nest for CPU timing: only the iBZ q with
 * index < nq and the k < nkk contribute. */
int orc_P0_sample(const orc_desc *d, const double *V, const double *G_full, int t0, int t1, int nq, int nkk,
                  double *P0, int nthreads);
/* Sigma for the iBZ k and t in [t0, t1) into ts_kij (other rows kept), scaled
 * -1/nk (gw_block.cpp:156); P_full is the q_t_QP over the full BZ. */
int orc_Sigma(const orc_desc *d, const double *V, const double *G_full, const double *P_full, int t0, int t1,
              double *Sigma, int nthreads);

/* One output slice summed over the whole mesh, for parity checks of systems
 * whose full sweeps are too large for the CPU: P0 of iBZ q index iqr at time t
 * (out: qslice complex, scaled as orc_P0), Sigma of iBZ k index ikr at (t, s)
 * (out: kslice complex, scaled as orc_Sigma).  V is given per pair: pair
 * p = k1*nk + k2 at V + 2*voff[p]; a needed pair with voff[p] < 0 returns 2.
 * P_full_t: one time slice of P over the full BZ ([q][qslice(q)]). */
int orc_P0_at(const orc_desc *d, const double *V, const int64_t *voff, const double *G_full, int iqr, int t,
              double *out, int nthreads);
int orc_Sigma_at(const orc_desc *d, const double *V, const int64_t *voff, const double *G_full, const double *P_full_t,
                 int ikr, int t, int s, double *out, int nthreads);
/* Time slice t of orc_rotate_P. */
int orc_rotate_P_t(const orc_desc *d, const double *P_ibz, const double *auxrep, int t, double *P_full_t);

/* W: P0 (q_t_QP, already /nk, rows t < nts/2 used) -> P (q_t_QP all nts). */
int orc_W(const orc_desc *d, const double *P0, const double *Tnt_BF, const double *Ttn_FB, double *P);
/* Dyson: Sigma (ts_kij, final) -> G (ts_kij). iwn: nw complex i*omega_n. */
int orc_dyson(const orc_desc *d, const double *Sigma, const double *F, const double *S, double mu,
              const double *Tnt, const double *Ttn, const double *iwn, double *G);
/* -(2/ns) sum_{s,k} w_k Tr[G_k(beta) S_k] / nk */
double orc_density(const orc_desc *d, const double *G, const double *S, const double *kw);

/* Dense complex helpers used by the tests' closed-form checks. */
int orc_inverse_lu(int n, const double *A, double *Ainv);
int orc_inverse_chol(int n, const double *A, double *Ainv);

#ifdef __cplusplus
}
#endif
#endif
