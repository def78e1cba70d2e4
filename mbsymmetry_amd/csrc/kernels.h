// Launch interface of the device kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "common.h"
#include "layout.h"

namespace symgw {

// One batch of the transform GEMM: rows of `in` start at in_off with stride
// in_row_stride, rows of `out` at out_off with out_row_stride; xlen columns
// per split group.
struct TBatch {
  int64_t in_off, out_off, in_row_stride, out_row_stride;
  int64_t in_split, out_split;  // offset between column groups (nsplit > 1)
  int32_t xlen, pad;
};

// One V block for the device generator.
struct GenBlk {
  int64_t v_off;      // absolute complex offset in the V buffer
  int64_t local_off;  // offset inside the pair slice (generator counter base)
  int32_t q, k1, k2;
  int32_t a_base;     // dense orbital index of the block's first row (k1 side)
  int32_t b_base;     // dense orbital index of the block's first column (k2 side)
  int32_t pair;       // k1 * nk + k2
};

// Items with part_out >= 0 write unscaled partial sums to `part`; launch_p0_reduce
// then adds the parts of each split item in order into P0 (times pref).
template <typename T>
hipError_t launch_p0(const P0Item *items, const P0Blk *blks, int n_items, int nts, int ns, const cplx<T> *V,
                     const cplx<T> *G, cplx<double> *P0, cplx<double> *part, double pref, size_t lds, hipStream_t st,
                     int mode);
hipError_t launch_p0_reduce(const P0Split *sp, int nsp, int shard_first, int shard_end, const cplx<double> *part,
                            cplx<double> *P0, double pref, int nts, hipStream_t st);
// INT8 slice copies read by the mixed-precision Z product of Sigma.
struct Z8Ops {
  const int8_t *P8 = nullptr, *V8 = nullptr;
  const int *PE = nullptr, *VE = nullptr;
};
// z8 == nullptr: every Z product on the matrix cores of T; otherwise the
// blocks with SigBlk::kp > 0 run it on INT8 slices (T = double only).
// Items with part_out >= 0 write unscaled partial sums to `part`; launch_sig_reduce
// then adds the parts of each split item in order into Sig (times pref).
template <typename T>
hipError_t launch_sigma(const SigItem *items, const SigBlk *blks, int n_items, int nts, int ns, const cplx<T> *V,
                        const cplx<T> *G, const cplx<T> *Pf, cplx<double> *Sig, cplx<double> *part, double pref,
                        size_t lds, hipStream_t st, int times, const Z8Ops *z8);
hipError_t launch_sig_reduce(const SigSplit *sp, int nsp, int shard_first, int shard_end, const cplx<double> *part,
                             cplx<double> *Sig, double pref, int nts, hipStream_t st);
// Split rows of complex128 matrices into INT8 slices (Split8Blk);
// row_prefix[b] is the first flat row of block b (nsb + 1 entries).
hipError_t launch_split8(const Split8Blk *sb, int nsb, const int64_t *row_prefix, int64_t nrows_total,
                         const cplx<double> *src, int8_t *dst, int *ex, hipStream_t st);
// peaks.hip: device peak microbenchmarks
template <typename T>
hipError_t launch_mfma_peak(void *out, int n_wg, int iters, hipStream_t st);
int mfma_peak_threads();
hipError_t launch_hbm_copy(const void *src, void *dst, int64_t bytes, int n_wg, hipStream_t st);
template <typename TO>
hipError_t launch_rotate(const RotBlk *rb, int n_rb, int nts_ns, const cplx<double> *src, const int64_t *src_ts_stride,
                         const int64_t *src_slice, const cplx<double> *rep, cplx<TO> *dst, const int64_t *dst_slice,
                         const int64_t *dst_slice_size, hipStream_t st);
hipError_t launch_tgemm(const TBatch *tb, int nbatch, int max_cols, int M, int KT, int nsplit,
                        const cplx<double> *Tm, const cplx<double> *in, cplx<double> *out, int mode, int row_shift,
                        int nts, double alpha, int64_t in_split_stride, int64_t out_split_stride, hipStream_t st);
// W solves of the bosonic frequencies [nw_first, nw_first + nw_count);
// scratch holds nw_count groups of scr_group complex elements.
hipError_t launch_wsolve(const SolveBlk *blks, int nblk, int nw_first, int nw_count, cplx<double> *Pw,
                         const int64_t *w_off, const int64_t *qslice, cplx<double> *scratch, int64_t scr_group,
                         hipStream_t st);
hipError_t launch_dyson(const SolveBlk *blks, int nblk, int nw, int ns, const cplx<double> *Sw, cplx<double> *Gw,
                        const int64_t *d_off, const int64_t *kslice, const int64_t *f_off, const cplx<double> *F,
                        const cplx<double> *S, int64_t flat_ao, const cplx<double> *iwn, double mu,
                        double sigma_scale, cplx<double> *scratch, int64_t scr_group, hipStream_t st);
hipError_t launch_density(const SolveBlk *blks, int nblk, int ns, const cplx<double> *Gbeta, const cplx<double> *S,
                          const int64_t *f_off, int64_t flat_ao, const double *kw, double *out, hipStream_t st);
template <typename T>
hipError_t launch_genv(const GenBlk *gb, int ngb, uint64_t seed, double scale, const double *decay, int n_decay,
                       cplx<T> *V, hipStream_t st);
template <typename TO>
hipError_t launch_convert(const cplx<double> *in, cplx<TO> *out, int64_t n, hipStream_t st);
hipError_t launch_reorder(const cplx<double> *in, cplx<double> *out, const int64_t *koff, const int64_t *ksize,
                          int ink, int nts_ns, int64_t max_ksize, int64_t flat_ao, int dir, double scale,
                          hipStream_t st);

}  // namespace symgw
