// Host-side block-sparse layout of one GW system: the derived tables of
// symmetry_utils_t (SymmetryMBPT/src/common/symmetry_utils.cpp:96-191) and the
// per-contraction work lists that the device kernels walk.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/symgw.h"

namespace symgw {

struct Layout {
  int nk = 0, ink = 0, ns = 0, nts = 0, nw = 0, nw_b = 0;

  // full-BZ irrep blocks: sizes, flat offsets (cumulative b*b), offsets of the
  // block inside the dense nao / naux index range (cumulative b)
  std::vector<std::vector<int>> ao_sizes, aux_sizes;
  std::vector<std::vector<int64_t>> ao_offsets, aux_offsets;
  std::vector<std::vector<int>> ao_block_offsets, aux_block_offsets;
  std::vector<int64_t> kao_off, kao_size, qaux_off, qaux_size;          // full BZ
  std::vector<int64_t> kao_off_irre, kao_size_irre, qaux_off_irre, qaux_size_irre;

  std::vector<int> irre_list, index, irre_index, conj_index, kpq, kmq;

  // rotation blocks per full-BZ k (used through the TR partner ck)
  std::vector<std::vector<std::array<int, 2>>> orep_irreps, auxrep_irreps;
  std::vector<std::vector<int64_t>> orep_offsets, auxrep_offsets;
  std::vector<int64_t> korep_off, korep_size, kauxrep_off, kauxrep_size;

  // interaction tensor tables per ordered pair p = k1*nk + k2
  std::vector<int64_t> kpair_off, kpair_size;           // -1 / 0 when absent
  std::vector<std::vector<std::array<int, 3>>> V_irreps;
  std::vector<std::vector<int64_t>> V_offsets;
  std::vector<std::vector<int>> V_qsize;                // aux size of each block

  int64_t flat_ao = 0, flat_ao_full = 0, flat_aux = 0, flat_aux_full = 0;
  int64_t flat_V = 0, flat_orep = 0, flat_auxrep = 0;
  int max_ao_block = 0, max_aux_block = 0;
};

// Build and validate; returns false with a message on inconsistent input.
bool build_layout(const symgw_layout_desc *d, Layout &L, std::string &err);

// ---- device work lists ------------------------------------------------------
// Contraction workgroup shape: one workgroup per CU on the whole 160 KiB LDS,
// so large blocks still get aux-row chunks that fill the 16-row MFMA tiles;
// 16 waves (four per SIMD, at most 128 VGPRs each) with complex128 operands,
// 12 (three per SIMD, 168 VGPRs) with complex64, whose kernels need the
// registers.
constexpr int contract_lds(int elem_bytes) { return elem_bytes > 0 ? 160 * 1024 : 0; }
// The contraction kernels' LDS starts with the job counter of their dynamic
// deal; the data (LDS chunks, split-K buffers) follow it.
constexpr int kDealBytes = 16;
constexpr int contract_waves(int elem_bytes) { return elem_bytes == 16 ? 16 : elem_bytes > 0 ? 12 : 0; }
// Accumulator tiles each contraction wave keeps across a slab: three with
// complex64 operands, two with complex128 (the register file holds no more).
constexpr int contract_tiles(int elem_bytes) { return elem_bytes == 8 ? 3 : 2; }
// Most output rows of one P0 slab (one MFMA tile row); the rows of an aux
// block are split into balanced bands.
constexpr int kP0Band = 16;
// Consecutive times of one P0 slab (they share every V fragment), per kernel
// mode (p0_mode below): with complex64 operands six in the stacked-only
// kernel (the split-K buffer of 12 waves x 6 accumulator tiles then takes 144
// of the 160 KiB of LDS; 3.8% faster than four) and four in the others (six
// spill there); two with complex128 (four spill registers there).
constexpr int p0_times(int elem_bytes, int mode) { return elem_bytes == 8 ? (mode == 1 ? 6 : 4) : 2; }
// P0 slabs of aux blocks whose column tiles fit one per wave run the times of
// a slab in one pass (stacked); larger aux blocks one time after the other.
// A system with one kind runs the kernel instantiation for that kind alone,
// so that the registers of the other path do not burden it; a system with
// both runs one launch of the instantiation that chooses per slab (two
// launches in sequence would leave the GPU to the tail of the first).
constexpr bool p0_stacked(int elem_bytes, int qs) { return (qs + 15) / 16 <= contract_waves(elem_bytes); }
// Widest ao block that one Sigma slab covers whole (3 x 3 tiles of 16 per time).
constexpr int kSigBand = 48;
// Most aux rows Q of one Sigma LDS chunk.
constexpr int kSigRows = 64;
// Shape of the Sigma slabs of an ao block of k1 orbitals: `times` consecutive
// times per slab (they share every V fragment; the last slab takes the
// remaining times) and square `band` x `band` output bands, with
// times * ceil(band / 16)^2 within the workgroup's contract_waves x
// contract_tiles accumulator tiles (32 for complex128, 36 for complex64).
//  - k1 <= 32: the whole block, four times.
//  - k1 <= 48: the whole block, four times (complex64) or two (complex128).
//  - wider blocks are split into bands of whole tiles, and every pair of bands
//    recomputes its Z (once per row band) and E (once per column band), so
//    fewer, wider bands save work while fewer times per slab cost V loads:
//    one time of the widest bands (6 tiles for complex64, 5 for complex128)
//    where that cuts the band count to 2/3 or less of two times of the
//    widest bands that fit (4 tiles), else the latter (the faster of four,
//    two and one times on MI355X at every k1 of 56..256 measured).
struct SigShape {
  int times, band;
};
// Slab depths (times per slab) with a Sigma kernel instantiation, item list
// and launch each (eight times per slab for blocks up to 32 measured 8-11%
// slower: the deeper slabs leave smaller LDS chunks).
constexpr int kNumSigDepths = 3;
constexpr int kSigDepths[kNumSigDepths] = {1, 2, 4};
constexpr int sig_depth_index(int times) { return times == 1 ? 0 : times == 2 ? 1 : 2; }
inline SigShape sig_shape(int elem_bytes, int k1) {
  if (k1 <= 32) return {4, k1};
  if (k1 <= kSigBand) return {elem_bytes == 8 ? 4 : 2, k1};
  const int tiles = (k1 + 15) / 16, cap2 = 4, cap1 = elem_bytes == 8 ? 6 : 5;
  auto bands = [&](int cap) { return (tiles + cap - 1) / cap; };
  auto even = [&](int cap) { const int nb = bands(cap); return 16 * ((tiles + nb - 1) / nb); };
  if (3 * bands(cap1) <= 2 * bands(cap2)) return {1, even(cap1)};
  return {2, even(cap2)};
}

// One V-block contraction that feeds a P0 output slab (cugw_qkpt_block.cu:219-261).
struct P0Blk {
  int64_t v_off;       // V block start (complex elements)
  int64_t g1_base;     // G_full offset of block ik1 of k      (t = 0, s = 0)
  int64_t g2_base;     // G_full offset of block ik2 of k + q  (t = 0, s = 0)
  int32_t g1_tstride;  // ns * kslice(k)
  int32_t g1_sstride;  // kslice(k)
  int32_t g2_tstride;
  int32_t g2_sstride;
  int32_t k1, k2;
  int32_t nsub;        // orbitals n of k + q per LDS chunk
  int32_t lde;         // LDS row stride of E and C (k1*nsub, odd)
};
// Rows [Q0, Q0 + nband) of aux block iq of iBZ point q_ir, summed over the k
// of its blocks [blk_begin, blk_end): every k, or one part of them when the
// item is split (part_out >= 0).
struct P0Item {
  int32_t blk_begin, blk_end;
  int32_t q, Q0, nband, iqr;
  int64_t out_base;          // q_t_QP offset of row Q0 at t = 0
  int32_t out_tstride;       // qslice(q_ir)
  int32_t pad2;
  int64_t part_out;          // -1: the slab writes P0; else this part's partial
                             // sums, [slab][time][nband][q] complex128, start here
};
// One split P0 item: its parts' partial sums are added in part order, scaled
// and written to P0 by the reduction kernel.
struct P0Split {
  int64_t part_out;          // partial buffer of part 0 (complex128 elements)
  int64_t part_stride;       // elements per part: np * TB * nband * q
  int64_t out_base;
  int32_t out_tstride, nparts;
  int32_t TB, np, nband, q;  // np: time slabs of the item
  int32_t iqr, pad;
};

// One V-block contraction that feeds a Sigma output slab
// (cugw_qkpt_block.cu:369-407).
struct SigBlk {
  int64_t v_off;
  int64_t g2_base;     // G_full offset of block ik2 of k - q (t = 0, s = 0)
  int64_t p_base;      // P_full offset of aux block iq of q (t = 0)
  int32_t g2_tstride, g2_sstride;
  int32_t p_tstride;   // qslice(q)
  int32_t q, k2;
  int32_t qsub;        // Q rows per LDS chunk
  int32_t ldz, lde;    // LDS row strides of the Z and E chunks
  int32_t nsub;        // orbitals n of k - q per LDS chunk
  // mixed mode (Z product on INT8 slices): split copies of this V block
  // (columns (a, b) over P) and of P of aux block iq of q (rows (t, Q) over P)
  int32_t kp;          // q rounded up to 16 (row length of the split copies)
  int64_t v8_off;      // bytes: V8 of this block, [comp][slice][col][kp]
  int64_t ve_off;      // int32 index: column exponents, [comp][col]
  int64_t p8_off;      // bytes: P8 of (q, iq) at t = 0, [t][comp][slice][Q][kp]
  int64_t pe_off;      // int32 index: row exponents at t = 0, [t][comp][Q]
};
// Rows [I0, I0 + Ib) x columns [J0, J0 + Jb) of ao block ik1 of iBZ point
// k_ir, summed over the q of its blocks [blk_begin, blk_end): every q, or one
// part of them when the item is split (part_out >= 0).
struct SigItem {
  int32_t blk_begin, blk_end;
  int32_t k1, I0, Ib, J0, Jb, ikr;
  int64_t out_base;          // k_ts_ij (iBZ) offset of row I0 at t = 0, s = 0
  int32_t out_tstride;       // ns * kslice(k)
  int32_t out_sstride;       // kslice(k)
  int64_t part_out;          // -1: the slab writes Sigma; else this part's partial
                             // sums, [slab][time][Ib][Jb] complex128, start here
};
// One split Sigma item: its parts' partial sums are added in part order, scaled
// and written to Sigma by the reduction kernel.
struct SigSplit {
  int64_t part_out;          // partial buffer of part 0 (complex128 elements)
  int64_t part_stride;       // elements per part: per * TB * Ib * Jb
  int64_t out_base;
  int32_t out_tstride, out_sstride;
  int32_t nparts, TB, per, ns;  // per: slabs of the item (time slabs x spins)
  int32_t k1, J0, Ib, Jb;
  int32_t ikr, pad;
};
// Sigma items (P0 items) are split over their q (k), i.e. over their blocks,
// until the sweep has about this many workgroups, and any item above a
// quarter of one CU's share of the sweep's work is split too; every part
// holds at least kSplitMinBlocks blocks.
constexpr int kSplitTarget = 1024;
constexpr int kSplitMinBlocks = 4;

// Rotation of one full-BZ k (G) or q (P) slice from its iBZ representative:
// block ia1 of the target = O * X(block ia2 of the source) * O^H, or a copy
// (cugw_routines_block.cu:134-186, cugw_qkpt_block.cu:288-340).
struct RotBlk {
  int64_t dst_off;     // target block offset inside the target slice
  int64_t src_off;     // source block offset inside the source slice
  int64_t rep_off;     // rotation block offset in the flat rep array, -1 => copy
  int32_t m, n;        // target block m x m, source block n x n (O is m x n)
  int32_t target;      // full-BZ index of the target
  int32_t conj;        // 1 => complex-conjugate the result (k != conj_index[k])
};

// INT8 slice split of one real-matrix family for the mixed-precision Z
// product of Sigma (Ozaki scheme: x = 2^e sum_s x_s 128^-s, x_s int8,
// s < kZ8Slices).
// Rows of `nsub` sub-blocks of `nrows` rows each, K elements per row (row r
// of sub-block u at src_off + u * sub_stride + r * row_stride, element k at
// + k * k_stride, complex128).  Output per sub-block u:
// dst_off + u * kZ8Comps * kZ8Slices * nrows * kp bytes, [comp][slice][row][kp] (k >= K zero);
// exponents e_off + u * 3 * nrows, [comp][row].  comps: re, im and
// re + im (kind 0, the A operand of the 3M product) or re - im (kind 1, the
// conjugated B operand).
// Five slices: the truncation error of a product is ~2^-35 of |A|max |B|max
// per term (four slices, 2^-28, left the Dyson step of the LiH fixture at
// |dG| ~ 5e-8).
constexpr int kZ8Slices = 5;
constexpr int kZ8Comps = 3;
// Aux blocks up to this size keep the FP64 Z product in the mixed mode: the
// exact reconstruction of the four int32 accumulators costs about as much as
// a short FP64 reduction (measured per block size on MI355X).
constexpr int kZ8MinQ = 32;
// streamed V: target size of one batch's V buffer (a batch of one iBZ point
// may exceed it).  Several points per batch keep each launch's grid large: one
// point of a 6x6x6 mesh gives a few hundred workgroups per Sigma launch, under
// one per CU, and runs at a quarter of the rate of a whole-sweep launch.
constexpr int64_t kStreamBatchBytes = int64_t(32) << 30;
// Workgroup of the Sigma kernel with the INT8 Z product: the FP64 shape, 16
// waves of two accumulator tiles.  At 128 VGPRs its slices and integer sums
// spill 13 registers, and it still runs 6% faster than 12 waves of three
// tiles without spills (config 2: 3813 against 4065 ms per Sigma sweep).
constexpr int kZ8Waves = 16;
constexpr int kZ8Tiles = 2;
struct Split8Blk {
  int64_t src_off, sub_stride, dst_off, e_off;
  int32_t row_stride, k_stride, nrows, K, kp, nsub, kind, pad;
};

// Largest source block a rotation reads (O is m x n with n <= kRotMaxBlock):
// the rotation workgroup holds 16 columns of X O^H, n x 16 complex128, in
// its kRotLdsBytes of LDS.
constexpr int kRotLdsBytes = 64 * 1024;
constexpr int kRotMaxBlock = kRotLdsBytes / 16 / 16;

// Contiguous aux-block of one W solve or ao-block of one Dyson solve.
struct SolveBlk {
  int64_t off;         // flat offset inside the slice
  int64_t scr;         // offset of this block's scratch inside one frequency group
  int32_t n;           // block size
  int32_t slice;       // index of the k (Dyson) or q (W) slice
};

struct WorkLists {
  std::vector<P0Blk> p0_blks;
  std::vector<P0Item> p0_items;
  std::vector<P0Split> p0_splits;                 // split P0 items
  int64_t p0_part_elems = 0;                      // complex128 elements of their partial sums
  int p0_mode = 0;  // 1: every slab stacked, 0: none, 2: both kinds (chosen per slab)
  std::vector<SigBlk> sig_blks;
  std::vector<SigItem> sig_items[kNumSigDepths];  // slabs of kSigDepths[d] times
  std::vector<SigSplit> sig_splits;               // split items of every depth
  int64_t sig_part_elems = 0;                     // complex128 elements of the partial sums
  std::vector<RotBlk> g_rot;      // targets over the full BZ (ao)
  std::vector<RotBlk> p_rot;      // targets over the full BZ (aux)
  std::vector<SolveBlk> w_blocks; // aux blocks of every iBZ q
  std::vector<SolveBlk> d_blocks; // ao blocks of every iBZ k
  int64_t w_scr_group = 0;        // scratch elements per bosonic frequency (2 n^2 per block)
  int64_t d_scr_group = 0;        // scratch elements per (w, s) (n^2 per block)
  int64_t n_p0_contractions = 0, n_sig_contractions = 0;
  double flop_p0 = 0, flop_sig = 0;
  size_t p0_lds = 0, sig_lds[kNumSigDepths] = {};  // dynamic LDS bytes of the contraction kernels (job counter included)
  // mixed mode: INT8 split tables (V blocks read by Sigma; P blocks of every q)
  std::vector<Split8Blk> v8_split, p8_split;
  // with a streamed V the V splits of each Sigma batch start at
  // v8_batch_begin[batch] and write a batch-sized slice buffer (v8_bytes and
  // ve_count are then the largest batch's)
  std::vector<int> v8_batch_begin;
  int64_t v8_bytes = 0, ve_count = 0, p8_bytes = 0, pe_count = 0;
  // streamed V (not resident): P0 and Sigma run one batch per group of
  // consecutive iBZ points (q for P0, k for Sigma; as many as keep the group's
  // V within kStreamBatchBytes, at least one); each batch's V pairs are
  // generated into one buffer just ahead of its contraction, and the blocks'
  // v_off are offsets in that buffer.
  bool stream = false;
  std::vector<int> p0_group_first, sig_group_first;  // [ngroups + 1] first iBZ index of each group
  // items of iBZ q (P0) / iBZ k (Sigma) index b: [prefix[b], prefix[b + 1])
  // (every mode: they also delimit the shards of symgw_set_shard)
  std::vector<int> p0_batch_items;                 // [ink + 1] item prefix
  std::vector<int> sig_batch_items[kNumSigDepths]; // [ink + 1] item prefix per slab depth
  // pairs of each batch with their offsets in the batch buffer
  std::vector<std::pair<int, int64_t>> p0_batch_pairs, sig_batch_pairs;
  std::vector<int> p0_batch_pair_begin, sig_batch_pair_begin;  // [ngroups + 1]
  int64_t batch_V_max = 0;                         // complex elements of the largest batch
};

// elem_bytes: bytes of one complex element in the contraction precision;
// z8: also plan the INT8 slice copies of the mixed-precision Z product.
// Tunable rules of the work-list planner (the autotuner's search space;
// the defaults are the hand-measured rules).
struct Plan {
  double sig_chunk_cost = 2.5e5;  // fixed cost of one Sigma LDS chunk (barriers, staging), in MFMA FLOPs
  int sig_rows = kSigRows;        // most aux rows Q of one Sigma chunk (1..kSigRows)
  int sig_max_times = 4;          // most times per Sigma slab (1, 2 or 4)
  double p0_chunk_cost = 2.5e5;   // the same for one P0 chunk
  int p0_band_cap = kP0Band;      // most output rows of one P0 slab (1..kP0Band)
};
bool valid_plan(const Plan &p);
bool build_worklists(const Layout &L, int elem_bytes, WorkLists &W, std::string &err, bool z8 = false,
                     bool stream = false, const Plan &plan = Plan());

// Per-iteration FLOPs of P0 and Sigma, 8*M*N*K per complex GEMM
// (analysis.cpp:229-255 and :271-298 without the +12MN epilogue).
void flop_model(const Layout &L, double &p0, double &sig, int64_t &nb_p0, int64_t &nb_sig);

}  // namespace symgw
