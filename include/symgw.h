/*
 * symgw.h — C ABI of the MI355X symmetry-adapted GW hot path.
 *
 * Every entry point returns an int status (SYMGW_OK == 0) and never throws; the
 * message of the last failure on a context is returned by symgw_last_error().
 * All complex arrays are interleaved complex128 (re, im) given as double*.
 * Layout names follow SymmetryMBPT:
 *   ts_kij  : [t][s][flat_ao(iBZ k)]           host G / Sigma (sc_loop_block_t)
 *   k_ts_ij : [k][t][s][flat_ao(k)]            device G / Sigma
 *   q_t_QP  : [q_ir][t][flat_aux(q_ir)]        P0 / P (row-major [Q][P] per block)
 * A "flat" slice of one k is the concatenation of its row-major b x b irrep blocks.
 *
 * Reference interfaces replaced (paths relative to the CQMP/MBSymmetry tree):
 *   cugw_routine_block<prec>   SymmetryMBPT/include/cuda/cugw_routines_block.h:22-114
 *   cugw_solver_block_t        SymmetryMBPT/src/solver/cugw_solver_block.cpp:34-469
 *   dyson_sc_block_t           SymmetryMBPT/src/scmbpt/sc_type_block.cpp:6-87
 */
#ifndef SYMGW_H
#define SYMGW_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SYMGW_OK 0
#define SYMGW_ERR_ARG 1        /* bad argument / layout inconsistency      */
#define SYMGW_ERR_HIP 2        /* HIP runtime failure                      */
#define SYMGW_ERR_NOMEM 3      /* device or host allocation failed          */
#define SYMGW_ERR_STATE 4      /* call out of order (e.g. no G uploaded)   */
#define SYMGW_ERR_NODEVICE 5   /* no usable gfx950 device                   */

/* GEMM precision of the P0 and Sigma contractions.  Transforms, W, Dyson and
 * the mu search always run in FP64. */
#define SYMGW_PREC_FP64 0      /* plain: complex FP64                          */
#define SYMGW_PREC_FP32 1      /* complex FP32 operands, FP32 accumulation   */
/* Mixed: the P0 contraction and the two short products of Sigma in FP64; the
 * q x q x k1 k2 product of Sigma (Z = P^H V) on INT8 matrix cores as exact
 * int32 products of 7-bit slices (Ozaki scheme, 5 slices per operand, a
 * power-of-two exponent per row of P and per column of V), reconstructed in
 * FP64.  V and P are also kept as slices: V8 once per set_V / generate_V,
 * P8 after each rotation of P. */
#define SYMGW_PREC_MIXED 2
/* Flag OR-ed into the precision argument of symgw_create /
 * symgw_layout_footprint: V is not kept resident.  symgw_generate_V records
 * the generator (symgw_set_V is refused) and every P0 / Sigma sweep
 * regenerates V on the device one batch at a time, just ahead of that batch's
 * contraction: the pairs (k, k + q) of one iBZ q for P0, (k, k - q) of one
 * iBZ k for Sigma, into a buffer of the largest batch.  Values are those of
 * the resident V bit for bit; the device needs the largest batch instead of
 * the whole V (for systems whose V exceeds HBM).  With SYMGW_PREC_MIXED the
 * INT8 slices of each Sigma batch's V are made right after it is generated,
 * into a buffer of the largest batch's slices. */
#define SYMGW_FLAG_STREAM_V 0x100

/*
 * Block-sparse layout of one system.  Mirrors the tables that
 * symmetry_utils_t reads from symmetry_info.h5
 * (SymmetryMBPT/src/common/symmetry_utils.cpp:6-218).  Offsets are derived
 * inside the library (cumulative b*b sums), as the reference stores them.
 *
 * Orbital (ao) and auxiliary (aux) irrep blocks are given for every k of the
 * full Brillouin zone.  The irreducible set and its maps follow
 * symmetry_utils.h:171-187: irre_list[ik] is the full-BZ index of the
 * ik-th stored point, index[k] the full-BZ index of the star representative
 * of k, irre_index[k] the iBZ position of that representative and
 * conj_index[k] the time-reversal partner (k itself when none).
 * kpq[k1*nk + q] is the full-BZ index of k1 + q and kmq[k1*nk + q] that of
 * k1 - q (momentum_conservation, symmetry_utils.cpp:315-341).
 *
 * Rotation blocks (orep / auxrep) are given per full-BZ k as a list of
 * (ia1, ia2) pairs: block ia1 of k is  O * X(block ia2 of the iBZ point) * O^H
 * with O a sizes_k(ia1) x sizes_ik(ia2) row-major matrix
 * (cugw_routines_block.cu:147-179, cugw_qkpt_block.cu:302-332).
 *
 * Interaction tensor V: for each ordered k-pair p = k1*nk + k2 that is
 * stored, kpair_nblocks[p] blocks (iq, ik1, ik2) with q x k1 x k2 row-major
 * elements V[Q][a][b] (a in k1 block ik1, b in k2 block ik2).  The aux block
 * index iq refers to the aux blocks of the momentum transfer of the pair.
 * Pairs with kpair_nblocks[p] == 0 are not stored.
 */
typedef struct symgw_layout_desc {
  int32_t nk;          /* k points in the full Brillouin zone            */
  int32_t ink;         /* stored (irreducible) k points                  */
  int32_t ns;          /* spins, 1 or 2                                  */
  int32_t nts;         /* imaginary-time points incl. 0 and beta; even  */
  int32_t nw;          /* fermionic Matsubara samples                    */
  int32_t nw_b;        /* bosonic Matsubara samples                      */
  const int32_t *kao_nblocks;    /* [nk]                                 */
  const int32_t *ao_sizes;       /* [sum kao_nblocks]                    */
  const int32_t *qaux_nblocks;   /* [nk]                                 */
  const int32_t *aux_sizes;      /* [sum qaux_nblocks]                   */
  const int32_t *irre_list;      /* [ink]                                */
  const int32_t *index;          /* [nk]                                 */
  const int32_t *irre_index;     /* [nk]                                 */
  const int32_t *conj_index;     /* [nk]                                 */
  const int32_t *kpq;            /* [nk*nk]                              */
  const int32_t *kmq;            /* [nk*nk]                              */
  const int32_t *orep_nblocks;   /* [nk]                                 */
  const int32_t *orep_irreps;    /* [sum orep_nblocks][2]                */
  const int32_t *auxrep_nblocks; /* [nk]                                 */
  const int32_t *auxrep_irreps;  /* [sum auxrep_nblocks][2]              */
  const int32_t *kpair_nblocks;  /* [nk*nk]                              */
  const int32_t *V_irreps;       /* [sum kpair_nblocks][3]               */
} symgw_layout_desc;

typedef struct symgw_ctx symgw_ctx;

/* Derived sizes of a layout (no device needed). */
typedef struct symgw_layout_info {
  int64_t flat_ao_size;        /* iBZ ts_kij slice length per (t,s)     */
  int64_t flat_ao_size_full;   /* full-BZ flat ao length                */
  int64_t flat_aux_size;       /* iBZ flat aux length (per t)           */
  int64_t flat_aux_size_full;  /* full-BZ flat aux length (per t)       */
  int64_t flat_V_size;         /* complex elements of all stored pairs  */
  int64_t flat_orep_size;      /* complex elements of all orep blocks   */
  int64_t flat_auxrep_size;    /* complex elements of all auxrep blocks */
  int64_t n_P0_blocks;         /* V-block contractions of one P0 sweep  */
  int64_t n_Sigma_blocks;      /* V-block contractions of one Sigma sweep */
  double flop_P0;              /* 8*M*N*K FLOPs of one P0 sweep (t < nts/2, all spins) */
  double flop_Sigma;           /* 8*M*N*K FLOPs of one Sigma sweep      */
  int64_t device_bytes;        /* device footprint of a context          */
} symgw_layout_info;

/* GW_complexity_estimation_block (analysis.cpp:212-298) with the 8MNK
 * convention, plus the footprint model used for admission. */
int symgw_layout_query(const symgw_layout_desc *desc, symgw_layout_info *info);

/* Device bytes a context of this layout allocates in the given precision
 * (SYMGW_PREC_*): the footprint model behind admission control.  Counterpart
 * of gw_qkpt_block::size (SymmetryMBPT/include/cuda/cugw_qkpt_block.h:45-53),
 * for the whole resident system instead of one stream's scratch.  Runs the
 * same work planner as symgw_create, so a layout the kernels cannot take
 * (a block above 256, a plan beyond the CU's LDS) fails here with the same
 * message (symgw_create_error). */
int symgw_layout_footprint(const symgw_layout_desc *desc, int precision, int64_t *bytes);

/* Free and total memory of a device (hipMemGetInfo); the check
 * GW_check_devices_free_space (SymmetryMBPT/src/solver/cugw_solver_block.cpp:471-493)
 * makes before sizing its streams. */
int symgw_device_memory(int device, int64_t *free_bytes, int64_t *total_bytes);

/* Measured device peaks, for the roofline next to the datasheet values (no
 * counterpart in the reference): out[0] FP64 and out[1] FP32 matrix-core
 * TFLOP/s of the contraction kernels' complex 16x16x4 step (four real MFMAs,
 * operands in registers, 12 waves per workgroup, 4 workgroups per CU), out[2]
 * HBM GB/s of a 1 GiB device-to-device copy (bytes read + bytes written;
 * a copy reaches about two thirds of the 8 TB/s datasheet figure).
 * Allocates ~2 GiB for the duration of the call. */
int symgw_measure_peaks(int device, double *out);

/* Intermediate-representation basis of the logistic kernel (no counterpart in
 * the reference, which reads precomputed grids: transformer_t.cpp:48-82 reads
 * xsample, wsample, uxl, uwl, ux1l of ir_grid/<Lambda>_<nl>.h5).  Computes the
 * first nl singular functions u_l(x), x in [-1, 1], of
 *   statistics 1 (fermionic): K(x, y) = e^{-Lambda y (x+1)/2} / (1 + e^{-Lambda y})
 *   statistics 0 (bosonic):   K(x, y) = y e^{-Lambda y (x+1)/2} / (1 - e^{-Lambda y})
 * (binary128 SVD of a Gauss-Legendre discretization with deg points per
 * segment; deg = 24 resolves the functions to ~1e-12), normalized with
 * u_l(1) > 0, as per-segment Legendre series: *nseg_out segments with
 * edges[0 .. nseg], coef[l][seg][k] (coef holds nl x max_seg x deg doubles).
 * Also returns the singular values s[nl], the tau sampling points
 * xsample[nl] (midpoints between the zeros of u_{nl-1} with -1 and 1 added),
 * the Matsubara sampling indices wsample[nl] (the discrete extrema of
 * u^_{nl-1}, mirrored; fermionic nl even, bosonic nl odd) and
 * uwl[2 * (i * nl + l)] = u^_l(i w_{wsample[i]}) (re, im), where
 * u^_l(i w_n) = 2^{-1/2} int_{-1}^{1} e^{i pi (n + zeta/2)(x + 1)} u_l(x) dx.
 * Returns 0; 1 when nl exceeds the numerical rank; 2 on bad arguments (or
 * max_seg too small); 3 / 4 when the Matsubara / tau sampling points cannot
 * be placed (nl beyond what the discretization resolves).  Host only, uses
 * the calling process's OpenMP threads; Lambda = 1e4, nl = 104 takes ~15 s
 * on 8 cores. */
int symgw_ir_basis(int statistics, double lambda, int nl, int max_seg, int deg, int *nseg_out, double *edges,
                   double *coef, double *s, double *xsample, int64_t *wsample, double *uwl);
/* u^_l(i w_n) of such a basis at nfreq caller-given indices n[i]:
 * out[2 * (i * nl + l)] (re, im).  Returns 0, or 2 on bad arguments. */
int symgw_ir_uhat(int statistics, int nl, int nseg, int deg, int max_seg, const double *edges, const double *coef,
                  int nfreq, const int64_t *n, double *out);

/* Offset of the flat V slice of ordered pair (k1, k2) in complex elements, or
 * -1 when the pair is not stored (symmetry_utils.h:114-123). */
int64_t symgw_kpair_offset(const symgw_ctx *ctx, int k1, int k2);

/* cugw_routine_block ctor + allocate_memory (cugw_routines_block.cu:6-91):
 * select `device`, allocate every device buffer of the system, upload the
 * work lists.  precision: SYMGW_PREC_*.  The context owns its own stream. */
int symgw_create(const symgw_layout_desc *desc, int device, int precision, symgw_ctx **out);
int symgw_destroy(symgw_ctx *ctx);
const char *symgw_last_error(const symgw_ctx *ctx);
/* Message of the last failure on the calling thread of symgw_create,
 * symgw_layout_query, symgw_layout_footprint, symgw_device_memory or
 * symgw_measure_peaks (no context exists then). */
const char *symgw_create_error(void);
int symgw_layout_info_ctx(const symgw_ctx *ctx, symgw_layout_info *info);

/* V: the whole flat interaction tensor (df_integral_block_t::read_entire,
 * df_integral_block_t.h:38-48), complex128, uploaded once and kept resident. */
int symgw_set_V(symgw_ctx *ctx, const double *V_flat, int64_t n_complex);
/* V generated on the device from a counter-based generator (no host copy):
 * V[Q][a][b] = (g1 + i g2)/sqrt(2) * scale * exp(-|a_glob - b_glob|/4),
 * g1, g2 unit-variance Irwin-Hall(4) deviates of Philox4x32-10(seed, element).
 * decay[d] = exp(-d/4) for d < n_decay (host-computed so host == device). */
int symgw_generate_V(symgw_ctx *ctx, uint64_t seed, double scale,
                     const double *decay, int n_decay);
/* Host twin of symgw_generate_V for one stored pair (k1, k2): writes the
 * pair's flat slice (complex128) into out (kpair slice size elements). */
int symgw_generate_V_pair_host(const symgw_layout_desc *desc, uint64_t seed, double scale,
                               const double *decay, int n_decay, int k1, int k2, double *out);

/* The same for npairs stored pairs at once, spread over the host's threads:
 * pair i = (pairs[2i], pairs[2i+1]) is written at out + 2 * out_off[i]
 * (complex128, its kpair slice size elements). */
int symgw_generate_V_pairs_host(const symgw_layout_desc *desc, uint64_t seed, double scale, const double *decay,
                                int n_decay, int npairs, const int32_t *pairs, const int64_t *out_off, double *out);

/* copy_ORep_host_to_device / copy_AuxRep_host_to_device
 * (cugw_routines_block.cu:214-234): flat rotation blocks, complex128. */
int symgw_set_reps(symgw_ctx *ctx, const double *orep_flat, const double *auxrep_flat);

/* copy_G_host_to_device (cugw_routines_block.cu:100-201): G_ts_kij
 * (nts, ns, flat_ao iBZ) complex128 -> device G over the full BZ (rotated,
 * conjugated for time-reversal partners).  The iBZ G also stays resident. */
int symgw_upload_G(symgw_ctx *ctx, const double *G_ts_kij);
/* Device iBZ G -> host ts_kij (overwrite). */
int symgw_download_G(symgw_ctx *ctx, double *G_ts_kij);

/* cugw_routine_block::solve_P0 (cugw_routines_block.cu:276-294): P0 over every
 * (q_ir, k) pair, t < nts/2, spin prefactor -2/ns applied, NOT divided by nk
 * (the caller divides, cugw_solver_block.cpp:302-304). Deterministic. */
int symgw_solve_P0(symgw_ctx *ctx);
/* copy_P0_device_to_host (cugw_routines_block.cu:236-247): ACCUMULATES (+=)
 * the device P0/P (q_t_QP, nts*flat_aux complex128) into host memory. */
int symgw_copy_P0_to_host(const symgw_ctx *ctx, double *P0_q_t_QP);
/* copy_P_host_to_device (cugw_routines_block.cu:203-212): overwrite device P. */
int symgw_upload_P(symgw_ctx *ctx, const double *P_q_t_QP);

/* Transform matrices (transformer_t.cpp:75-81), complex128 row-major:
 * Tnt (nw x (nts-2)), Ttn (nts x nw), Tnt_BF (nw_b x (nts-2)),
 * Ttn_FB (nts x nw_b); wsample_fermi: the nw values i*omega_n (complex). */
int symgw_set_transforms(symgw_ctx *ctx, const double *Tnt, const double *Ttn,
                         const double *Tnt_BF, const double *Ttn_FB,
                         const double *iwn_fermi);

/* cugw_solver_block_t::compute_Pq_from_P0q (cugw_solver_block.cpp:385-439) on
 * the device: mirror P0(beta - t) = P0(t), scale by 1/nk, P0(iW) = Tnt_BF P0,
 * per (q, n, block) Pt = [I - (P0 + P0^H)/2]^-1 (hermitized) * P0, back with
 * Ttn_FB.  Overwrites the device P0 buffer with P (as the reference aliases). */
int symgw_solve_W(symgw_ctx *ctx);

/* symgw_solve_W without allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same results as
 * symgw_solve_W, bit for bit; the solver scratch comes from a 32 MiB array
 * that the library holds on each GPU it runs on until the process ends, and
 * the bosonic frequencies are solved in groups that fit it.  A system whose
 * solves of one frequency need more than 32 MiB (2 n^2 complex128 per aux
 * block of every iBZ q) gets SYMGW_ERR_ARG.  The calls of one process are
 * serialised (they share each GPU's array); allocates nothing, neither on the
 * device nor through a library routine. */
int symgw_solve_W_noalloc(symgw_ctx *ctx);

/* cugw_routine_block::solve_Sigma (cugw_routines_block.cu:296-318): rotates P
 * to every q of the full BZ and contracts Sigma for the iBZ k, all t;
 * prefactor -1, NOT divided by nk (cugw_solver_block.cpp:50-52). */
int symgw_solve_Sigma(symgw_ctx *ctx);
/* copy_Sigma_device_to_host (cugw_routines_block.cu:249-274): ACCUMULATES
 * (+=) into host Sigma_ts_kij (nts, ns, flat_ao iBZ). */
int symgw_copy_Sigma_to_host(const symgw_ctx *ctx, double *Sigma_ts_kij);
/* Overwrite the device Sigma (host ts_kij, already divided by nk). */
int symgw_upload_Sigma(symgw_ctx *ctx, const double *Sigma_ts_kij);

/* F and S (ns, flat_ao iBZ) complex128 (sc_loop_block_t::read_input_data,
 * sc_loop_block_t.cpp:64-109 after the Bz shift and hermitization). */
int symgw_set_fock(symgw_ctx *ctx, const double *F, const double *S);
/* dyson_sc_block_t::compute_G (sc_type_block.cpp:6-87) on the device:
 * Sigma(iw) = Tnt Sigma(t); per (s, k, w, block)
 * G = [(iw + mu) S - F - Sigma(iw)]^-1; G(t) = Ttn G(iw).  The device Sigma
 * used is the one left by symgw_solve_Sigma divided by nk (sigma_scale =
 * 1/nk), or an uploaded one (sigma_scale = 1).  Result: device iBZ G and the
 * full-BZ rotated G used by the next symgw_solve_P0. */
int symgw_dyson(symgw_ctx *ctx, double mu, double sigma_scale);
/* Electron count N(mu) = -(2/ns) * sum_{k,s} w_k Tr[G_k(beta) S_k] / nk of
 * the device G (sc_loop_gf_block.cpp:18-20 density matrix, traced). */
int symgw_density(symgw_ctx *ctx, const double *kweights, double *nel);
/* Chemical-potential search (absent from the reference, whose mu is fixed):
 * bisection of N(mu) = nel_target through repeated device Dyson solves,
 * starting from [mu_lo, mu_hi] and widening it until it brackets the target
 * (SYMGW_ERR_ARG if it never does); leaves G of the final mu on the device. */
int symgw_mu_search(symgw_ctx *ctx, double nel_target, const double *kweights,
                    double sigma_scale, double mu_lo, double mu_hi, double tol,
                    double *mu_out, double *nel_out);

/* One GW iteration on the device: P0 -> W -> Sigma -> Dyson(mu), ending with
 * the rotated full-BZ G ready for the next iteration.  use_graph != 0
 * captures the launch sequence of each phase once as a hipGraph (re-captured
 * when mu changes) and replays the four graphs; with timing on, HIP events
 * between the replays give the per-phase device times. */
int symgw_gw_iteration(symgw_ctx *ctx, double mu, int use_graph);

/* Restrict the P0 and Sigma sweeps to the iBZ points with index in
 * [first, first + count): P0 of those q, Sigma of those k (all k, resp. q,
 * summed); the other rows of P0 / Sigma are left as they are.  This is one
 * GPU's share when the iBZ points are divided over GPUs (the reference
 * strides (q, k) pairs over devices instead, cugw_routines_block.cu:279,302);
 * W, Dyson and the rotations still cover every point.  Default: every
 * point.  Drops the captured iteration graphs. */
int symgw_set_shard(symgw_ctx *ctx, int first, int count);

/* Rules of the work-list planner that splits every P0 and Sigma block
 * contraction into LDS chunks and slabs: no counterpart in the reference
 * (which hands each block to cuBLAS, cugw_qkpt_block.cu:233-260); the search
 * space of the autotuner (mbsymmetry_amd/autotune.py), which times candidate
 * plans on the device for a system's block-size histogram.  Results do not
 * depend on the plan beyond rounding (the order of the k sums is fixed by
 * the plan).  symgw_plan_defaults gives the built-in rules; symgw_replan
 * rebuilds the work lists of a context under another plan (the data on the
 * device is kept; SYMGW_ERR_ARG if the plan is out of range or would change
 * a buffer size); symgw_get_plan returns the context's plan. */
typedef struct symgw_plan {
  double sig_chunk_cost;   /* fixed cost of one Sigma LDS chunk, MFMA FLOPs (default 2.5e5) */
  int32_t sig_rows;        /* most aux rows of one Sigma chunk, 1..64 (64)                   */
  int32_t sig_max_times;   /* most times per Sigma slab: 1, 2 or 4 (4)                       */
  double p0_chunk_cost;    /* fixed cost of one P0 LDS chunk (2.5e5)                          */
  int32_t p0_band_cap;     /* most output rows of one P0 slab, 1..16 (16)                    */
  int32_t reserved;        /* 0 */
} symgw_plan;
int symgw_plan_defaults(symgw_plan *plan);
int symgw_replan(symgw_ctx *ctx, const symgw_plan *plan);
int symgw_get_plan(const symgw_ctx *ctx, symgw_plan *plan);

/* Wait for all work queued on the context's stream. */
int symgw_sync(symgw_ctx *ctx);
/* Device time (ms) of the phases of the last timed symgw_gw_iteration,
 * from HIP events on the context stream: ms5[0] P0 contraction kernel,
 * [1] W (transforms, solves, rotation of P to the full BZ), [2] Sigma
 * contraction kernel, [3] Dyson (transforms, solves, rotation of G), [4] the
 * whole iteration.  symgw_solve_P0 / symgw_solve_Sigma with timing on set
 * [0] / [2]. */
int symgw_last_times(const symgw_ctx *ctx, double *ms5);
/* Enable/disable per-phase HIP event timing (off by default). */
int symgw_set_timing(symgw_ctx *ctx, int on);

#ifdef __cplusplus
}
#endif
#endif /* SYMGW_H */
