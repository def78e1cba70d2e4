// P0 and Sigma block contractions on the CDNA4 matrix cores.
//
// Every workgroup owns one output slab and walks its V-block list in
// a fixed order ("owner computes"), so results are bitwise reproducible with
// no atomics or locks (the reference serialises per-q / per-k AXPYs behind
// spin locks, cugw_qkpt_block.cu:269-285 and :414-436).  All complex products
// are 16x16 tiles of real MFMAs (v_mfma_f64_16x16x4f64 for the plain path,
// v_mfma_f32_16x16x4f32 for the FP32 path): four per complex k-step, or three
// (3M) in the products whose tiles live for one job; operands stream from
// memory or from LDS, intermediates never leave the workgroup's LDS.  A wave
// works on up to four persistent tiles, or on one or two job tiles, that share
// one operand fragment per k-step.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

namespace symgw {

// LDS of one rotation workgroup (the W = X O^H column chunk); the planner
// rejects source blocks above kRotMaxBlock, so a chunk is never narrower than
// 16 columns.
constexpr int kRotLds = kRotLdsBytes;
static_assert(kRotMaxBlock * 16 * 16 <= kRotLds, "rotation chunk of 16 columns must fit the LDS");

namespace {

// Waves per workgroup (16 for FP64, 12 for FP32) on the full 160 KiB LDS of
// a CU.
template <typename T>
struct Cfg {
  static constexpr int NW = contract_waves(2 * (int)sizeof(T));
  static constexpr int NT = 64 * NW;
  // accumulator tiles a wave keeps across the whole slab
  static constexpr int MT = contract_tiles((int)sizeof(cplx<T>));
};

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 op(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  // result row held by register r of a lane (lane & 15 is the column)
  __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 op(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// x / d for a divisor fixed over a product: one multiply-high by
// m = ceil(2^32 / d) instead of the ~25-instruction integer division.  Exact
// for 0 <= x with x * d < 2^32 (x * m / 2^32 = x / d + x * e / 2^32 with
// 0 <= e < 1, which stays below the next multiple of 1/d); every index here is
// below 2^17 and every divisor below 2^13.
struct FDiv {
  uint32_t m, d;
  __device__ __forceinline__ explicit FDiv(int dd) : m(dd > 1 ? 0xFFFFFFFFu / (uint32_t)dd + 1u : 0u), d((uint32_t)dd) {}
  __device__ __forceinline__ int div(int x) const { return d == 1 ? x : (int)__umulhi((uint32_t)x, m); }
};

template <typename T>
struct Tile {
  using elem = T;
  typename Mfma<T>::v4 re, im;
};
// A complex tile as the three real products of the 3M (Gauss) scheme:
// P1 = sum Ar Br, P2 = sum Ai Bi, P3 = sum (Ar + Ai)(Br + Bi) (Br - Bi when
// B is conjugated); A B = (P1 - P2) + i (P3 - P1 - P2), A conj(B) =
// (P1 + P2) + i (P3 - P1 + P2).  Three MFMAs per complex k-step instead of
// four, for 50% more accumulator registers: used by the products whose tiles
// live only for one job (the Z and E products of Sigma, E and C of P0).
template <typename T>
struct Tile3 {
  using elem = T;
  typename Mfma<T>::v4 p1, p2, p3;
};
// Tiles of one job (and of one pass over the persistent tiles) sharing one
// operand fragment per k-step: 1 for FP64 and 2 for FP32, the most that three
// waves per SIMD hold without spilling in the Sigma kernel.
template <typename T>
struct Group { static constexpr int n = sizeof(T) == 8 ? 1 : 2; };
template <typename T>
using Acc = Tile<T>[Cfg<T>::MT];

template <typename T>
__device__ __forceinline__ void tile_zero(Tile<T> &t) {
  t.re = typename Mfma<T>::v4{T(0), T(0), T(0), T(0)};
  t.im = t.re;
}
template <typename T>
__device__ __forceinline__ void tile_zero(Tile3<T> &t) {
  t.p1 = typename Mfma<T>::v4{T(0), T(0), T(0), T(0)};
  t.p2 = t.p1;
  t.p3 = t.p1;
}

// acc += a * b (CB: acc += a * conj(b)) for one 16x16x4 complex step, four
// real MFMAs.  Lane l holds A[l%16][k0+l/16] and B[k0+l/16][l%16] (operand
// layout of the 16x16x4 MFMAs).  The conjugate of B is folded into the signs
// so that a freshly loaded operand is never touched before its MFMA.
template <bool CB, typename T>
__device__ __forceinline__ void mma4(Tile<T> &acc, const cplx<T> &a, const cplx<T> &b) {
  if (CB) {
    acc.re = Mfma<T>::op(a.re, b.re, acc.re);
    acc.im = Mfma<T>::op(-a.re, b.im, acc.im);
    acc.re = Mfma<T>::op(a.im, b.im, acc.re);
    acc.im = Mfma<T>::op(a.im, b.re, acc.im);
  } else {
    acc.re = Mfma<T>::op(a.re, b.re, acc.re);
    acc.im = Mfma<T>::op(a.re, b.im, acc.im);
    acc.re = Mfma<T>::op(-a.im, b.im, acc.re);
    acc.im = Mfma<T>::op(a.im, b.re, acc.im);
  }
}
template <bool CB, typename T>
__device__ __forceinline__ void mma4(Tile3<T> &acc, const cplx<T> &a, const cplx<T> &b) {
  acc.p1 = Mfma<T>::op(a.re, b.re, acc.p1);
  acc.p2 = Mfma<T>::op(a.im, b.im, acc.p2);
  acc.p3 = Mfma<T>::op(a.re + a.im, CB ? b.re - b.im : b.re + b.im, acc.p3);
}

// The k-loops below keep AH k-steps of operands in flight: the fragments of
// step k+AH are loaded while the MFMAs of step k run, in a ring of AH + 1
// register sets (no copies), and every load is unconditional.  Loaders
// therefore clamp their indices and return some element of their operand for
// any row, column or k; garbage in rows or columns outside the product only
// reaches accumulator rows or columns that are never stored, and the A
// operand is zeroed where k >= ke or the lane's A row is masked (arow false)
// at the moment it is used.

// Keeps the scheduler from sinking the next step's loads below this step's
// MFMAs (it does so under register pressure, which serialises every k-step on
// the memory latency).
__device__ __forceinline__ void prefetch_fence() { __builtin_amdgcn_sched_barrier(0); }

// The k-step loop shared by the products below: ceil((ke - kb) / 4) steps,
// step s reading register set s % (AH + 1); ld(set, k) loads the fragments of
// the step at k into a set, mm(set, k) runs that step's MFMAs.  The loop exit
// is tested once per AH + 1 steps, after the last of them: a branch between a
// prefetch and the step that consumes it lets the compiler sink the load into
// the consumer's block, right before its use, which takes the prefetch away
// (measured: every second step then waited on its own loads).  The last
// (steps mod (AH + 1)) steps run from sets already loaded.
template <int AH, typename LD, typename MM>
__device__ __forceinline__ void kloop(int kb, int ke, const LD &ld, const MM &mm) {
  constexpr int R = AH + 1;
  const int nst = (ke - kb + 3) >> 2;
#pragma unroll
  for (int d = 0; d < AH; ++d) ld(d, kb + 4 * d);
  int k0 = kb;
  for (int g = nst / R; g > 0; --g) {
#pragma unroll
    for (int h = 0; h < R; ++h) {
      ld((h + AH) % R, k0 + 4 * AH);
      prefetch_fence();
      mm(h, k0);
      k0 += 4;
    }
  }
  const int rem = nst - (nst / R) * R;
#pragma unroll
  for (int h = 0; h < AH; ++h)
    if (h < rem) {
      mm(h, k0);
      k0 += 4;
    }
}

template <typename T>
__device__ __forceinline__ cplx<T> amask(const cplx<T> &a, bool ok) {
  return cplx<T>{ok ? a.re : T(0), ok ? a.im : T(0)};
}

// Operand k-steps kept in flight ahead of the MFMAs (register sets used in a
// ring of kAhead + 1).
constexpr int kAhead = 2;
// The same for the job products (3M tiles): one step in both precisions
// (measured per kernel: FP64 Sigma +1-2% and FP32 Sigma +3.8% against two).
template <typename T>
constexpr int kAheadJob = 1;
// P0's E and C jobs under the dynamic deal (FP32: P0 +1.6%) or the static
// round-robin one (FP64: the dynamic deal measured 2.5-3.4% slower).
template <typename T>
constexpr bool kP0Dyn = sizeof(T) == 4;
// Sigma's Z and E products with one job layout each (column- and row-share)
// instead of a per-chunk choice: FP64 (Sigma +1.2%, 5 -> 1 spilled registers).
template <typename T>
constexpr bool kSigOneLayout = sizeof(T) == 8;

// acc[C0+c] += A (16 x K) * B_c (K x 16) for c < NG with C0 + c < n, over k in
// [kb, ke): one A fragment per k-step shared by the tiles.  la(i, k), lb(c, k, j)
// with c the absolute tile index.  n must be uniform across the wave; a group
// with fewer than NG tiles left runs as the next smaller group, so that the
// k-loop body holds no per-tile branch.
template <int C0, int NG, bool CB, int AH = kAhead, typename TT, int N, typename LA, typename LB>
__device__ __forceinline__ void mma_rowshare(TT (&acc)[N], int n, int kb, int ke, bool arow, const LA &la,
                                             const LB &lb) {
  using T = typename TT::elem;
  static_assert(C0 + NG <= N, "tile group outside the accumulator array");
  if constexpr (NG > 1) {
    if (C0 + NG > n) {
      mma_rowshare<C0, NG - 1, CB, AH>(acc, n, kb, ke, arow, la, lb);
      return;
    }
  }
  if (kb >= ke || C0 >= n) return;
  const int lane = threadIdx.x & 63, i = lane & 15, kk = lane >> 4;
  cplx<T> a[AH + 1], b[AH + 1][NG];
  kloop<AH>(
      kb, ke,
      [&](int r, int k) {
        a[r] = la(i, k + kk);
#pragma unroll
        for (int c = 0; c < NG; ++c) b[r][c] = lb(C0 + c, k + kk, i);
      },
      [&](int r, int k) {
        const cplx<T> am = amask(a[r], arow && k + kk < ke);
#pragma unroll
        for (int c = 0; c < NG; ++c) mma4<CB>(acc[C0 + c], am, b[r][c]);
      });
}

// acc[C0+c] += A_c (16 x K) * B (K x 16): one B fragment shared.  la(c, i, k), lb(k, j).
template <int C0, int NG, bool CB, int AH = kAhead, typename TT, int N, typename LA, typename LB>
__device__ __forceinline__ void mma_colshare(TT (&acc)[N], int n, int kb, int ke, const LA &la, const LB &lb) {
  using T = typename TT::elem;
  static_assert(C0 + NG <= N, "tile group outside the accumulator array");
  if constexpr (NG > 1) {
    if (C0 + NG > n) {
      mma_colshare<C0, NG - 1, CB, AH>(acc, n, kb, ke, la, lb);
      return;
    }
  }
  if (kb >= ke || C0 >= n) return;
  const int lane = threadIdx.x & 63, i = lane & 15, kk = lane >> 4;
  cplx<T> a[AH + 1][NG], b[AH + 1];
  kloop<AH>(
      kb, ke,
      [&](int r, int k) {
        b[r] = lb(k + kk, i);
#pragma unroll
        for (int c = 0; c < NG; ++c) a[r][c] = la(C0 + c, i, k + kk);
      },
      [&](int r, int k) {
        const bool kin = k + kk < ke;
#pragma unroll
        for (int c = 0; c < NG; ++c) mma4<CB>(acc[C0 + c], amask(a[r][c], kin), b[r]);
      });
}

// acc[C0+c] += A_c * B_c with independent operands.  la(c, i, k), lb(c, k, j).
template <int C0, int NG, bool CB, typename T, int N, typename LA, typename LB>
__device__ __forceinline__ void mma_multi(Tile<T> (&acc)[N], int n, int kb, int ke, const LA &la, const LB &lb) {
  static_assert(C0 + NG <= N, "tile group outside the accumulator array");
  if constexpr (NG > 1) {
    if (C0 + NG > n) {
      mma_multi<C0, NG - 1, CB>(acc, n, kb, ke, la, lb);
      return;
    }
  }
  if (kb >= ke || C0 >= n) return;
  const int lane = threadIdx.x & 63, i = lane & 15, kk = lane >> 4;
  cplx<T> a[kAhead + 1][NG], b[kAhead + 1][NG];
  kloop<kAhead>(
      kb, ke,
      [&](int r, int k) {
#pragma unroll
        for (int c = 0; c < NG; ++c) {
          a[r][c] = la(C0 + c, i, k + kk);
          b[r][c] = lb(C0 + c, k + kk, i);
        }
      },
      [&](int r, int k) {
        const bool kin = k + kk < ke;
#pragma unroll
        for (int c = 0; c < NG; ++c) mma4<CB>(acc[C0 + c], amask(a[r][c], kin), b[r][c]);
      });
}

// All persistent tiles of a wave, in groups of Group<T>::n.
template <bool CB, typename T, int NA, typename LA, typename LB>
__device__ __forceinline__ void mma_rowshare_all(Tile<T> (&acc)[NA], int n, int kb, int ke, bool arow, const LA &la,
                                                 const LB &lb) {
  constexpr int NG = Group<T>::n, MT = Cfg<T>::MT;
  mma_rowshare<0, NG < MT ? NG : MT, CB>(acc, n, kb, ke, arow, la, lb);
  if (NG < MT) mma_rowshare<NG < MT ? NG : 0, NG < MT ? MT - NG : 1, CB>(acc, n, kb, ke, arow, la, lb);
}
template <bool CB, typename T, int MT, typename LA, typename LB>
__device__ __forceinline__ void mma_multi_all(Tile<T> (&acc)[MT], int n, int kb, int ke, const LA &la, const LB &lb) {
  constexpr int NG = Group<T>::n;
  mma_multi<0, NG < MT ? NG : MT, CB>(acc, n, kb, ke, la, lb);
  if (NG < MT) mma_multi<NG < MT ? NG : 0, NG < MT ? MT - NG : 1, CB>(acc, n, kb, ke, la, lb);
}

// Split of K into S parts for part p (boundaries multiples of 4).
__device__ __forceinline__ void k_part(int K, int S, int p, int &kb, int &ke) {
  const int q4 = ((K + 4 * S - 1) / (4 * S)) * 4;
  kb = min(K, p * q4);
  ke = min(K, kb + q4);
}

// Work split of nt accumulator tiles over the NW waves: S waves share each
// tile over S parts of the reduction (split-K), so nt tiles are spread over
// floor(NW / S) tile groups with at most MT tiles per wave; the NW mod S
// waves past the last group hold no tiles (callers give them none).  S
// minimises the per-wave work ceil(nt / floor(NW / S)) / S; ties keep the
// smaller S.
template <int NW, int MT>
__device__ __forceinline__ int choose_split(int nt) {
  int best = 1, num = 1 << 30, den = 1;
  for (int s = 1; s <= NW; ++s) {
    const int g = NW / s, tpw = (nt + g - 1) / g;
    if (tpw > MT) continue;
    if (tpw * den < num * s) {
      num = tpw;
      den = s;
      best = s;
    }
  }
  return best;
}

// Sum the S split-K partial copies of each wave's tiles (waves g*S .. g*S+S-1
// hold the parts of tile group g) in part order 0, 1, ..., S-1, so the result
// is bitwise reproducible; part 0 of each group ends with the totals.  Parts
// 1.. S-1 pass through LDS slot `wave` at `red` (NW * N * 8 * 64 elements of
// T); callers sync before reusing it.
template <typename T, int N>
__device__ __forceinline__ void split_reduce(Tile<T> (&acc)[N], int nmine, int S, T *red) {
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, kp = wave % S;
  __syncthreads();
  if (kp != 0) {
    T *slot = red + wave * (N * 8 * 64);
#pragma unroll
    for (int c = 0; c < N; ++c)
      if (c < nmine)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          slot[(c * 8 + r) * 64 + lane] = acc[c].re[r];
          slot[(c * 8 + 4 + r) * 64 + lane] = acc[c].im[r];
        }
  }
  __syncthreads();
  if (kp == 0) {
    for (int p = 1; p < S; ++p) {
      const T *slot = red + (wave + p) * (N * 8 * 64);
#pragma unroll
      for (int c = 0; c < N; ++c)
        if (c < nmine)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            acc[c].re[r] += slot[(c * 8 + r) * 64 + lane];
            acc[c].im[r] += slot[(c * 8 + 4 + r) * 64 + lane];
          }
    }
  }
}

// f(row, col, value) for the 4 elements a lane holds.
template <typename T, typename F>
__device__ __forceinline__ void tile_store(const Tile<T> &acc, const F &f) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < 4; ++r) f(Mfma<T>::row(lane, r), lane & 15, cplx<T>{acc.re[r], acc.im[r]});
}
// The same for a 3M tile (CB: the product conjugated B).
template <bool CB, typename T, typename F>
__device__ __forceinline__ void job_store(const Tile3<T> &acc, const F &f) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const T p1 = acc.p1[r], p2 = acc.p2[r], p3 = acc.p3[r];
    f(Mfma<T>::row(lane, r), lane & 15, CB ? cplx<T>{p1 + p2, p3 - p1 + p2} : cplx<T>{p1 - p2, p3 - p1 - p2});
  }
}

// C (M x N) = A (M x K) B (K x N) into LDS as independent jobs.  Row-share
// jobs are (row tile, group of NG column tiles) sharing the A fragment of each
// k-step; column-share jobs are (group of NG row tiles, column tile) sharing
// B.  la(row, k), lb(k, col), st(row, col, value) (st only for row < M,
// col < N); loaders clamp (see above).
template <typename T, int NG, bool CB, int AH, typename LA, typename LB, typename ST>
__device__ __forceinline__ void job_rs(int M, int N, int K, int mt, int gi, int ntn, const LA &la, const LB &lb,
                                       const ST &st) {
  const int g = gi * NG, n = min(NG, ntn - g), m0 = mt * 16;
  Tile3<T> acc[NG];
#pragma unroll
  for (int c = 0; c < NG; ++c) tile_zero(acc[c]);
  mma_rowshare<0, NG, CB, AH>(acc, n, 0, K, true, [&](int i, int k) { return la(m0 + i, k); },
                             [&](int c, int k, int jj) { return lb(k, (g + c) * 16 + jj); });
#pragma unroll
  for (int c = 0; c < NG; ++c)
    if (c < n)
      job_store<CB>(acc[c], [&](int i, int jj, cplx<T> v) {
        const int row = m0 + i, col = (g + c) * 16 + jj;
        if (row < M && col < N) st(row, col, v);
      });
}
template <typename T, int NG, bool CB, int AH, typename LA, typename LB, typename ST>
__device__ __forceinline__ void job_cs(int M, int N, int K, int gt, int nt, int ntm, const LA &la, const LB &lb,
                                       const ST &st) {
  const int g = gt * NG, n = min(NG, ntm - g), n0 = nt * 16;
  Tile3<T> acc[NG];
#pragma unroll
  for (int c = 0; c < NG; ++c) tile_zero(acc[c]);
  mma_colshare<0, NG, CB, AH>(acc, n, 0, K, [&](int c, int i, int k) { return la((g + c) * 16 + i, k); },
                             [&](int k, int jj) { return lb(k, n0 + jj); });
#pragma unroll
  for (int c = 0; c < NG; ++c)
    if (c < n)
      job_store<CB>(acc[c], [&](int i, int jj, cplx<T> v) {
        const int row = (g + c) * 16 + i, col = n0 + jj;
        if (row < M && col < N) st(row, col, v);
      });
}

// Static round-robin deal: job j to wave (first + j) % NW; returns the job
// count so that a following product can continue the deal.
template <typename T, int NG, bool CB, int AH, int NW, typename LA, typename LB, typename ST>
__device__ __forceinline__ int gemm_jobs_rs_g(int M, int N, int K, int first, const LA &la, const LB &lb,
                                              const ST &st) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ntm = (M + 15) / 16, ntn = (N + 15) / 16, ngr = (ntn + NG - 1) / NG, njobs = ntm * ngr;
  // job j = mt * ngr + gi, advanced by NW per round without dividing
  int j = (wave - first % NW + NW) % NW, mt = j / ngr, gi = j - mt * ngr;
  const int dmt = NW / ngr, dgi = NW - dmt * ngr;
  for (; j < njobs; j += NW, gi += dgi, mt += dmt + (gi >= ngr ? 1 : 0), gi -= gi >= ngr ? ngr : 0)
    job_rs<T, NG, CB, AH>(M, N, K, mt, gi, ntn, la, lb, st);
  return njobs;
}

template <typename T, int NG, bool CB, int AH, int NW, typename LA, typename LB, typename ST>
__device__ __forceinline__ int gemm_jobs_cs_g(int M, int N, int K, int first, const LA &la, const LB &lb,
                                              const ST &st) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ntm = (M + 15) / 16, ntn = (N + 15) / 16, ngr = (ntm + NG - 1) / NG, njobs = ngr * ntn;
  // job j = gt * ntn + nt, advanced by NW per round without dividing
  int j = (wave - first % NW + NW) % NW, gt = j / ntn, nt = j - gt * ntn;
  const int dgt = NW / ntn, dnt = NW - dgt * ntn;
  for (; j < njobs; j += NW, nt += dnt, gt += dgt + (nt >= ntn ? 1 : 0), nt -= nt >= ntn ? ntn : 0)
    job_cs<T, NG, CB, AH>(M, N, K, gt, nt, ntm, la, lb, st);
  return njobs;
}

// Grouped jobs when there are enough of them to keep every wave busy, else
// single-tile jobs (less operand sharing, shorter critical path).  NW: waves
// of the calling workgroup.
template <typename T, bool CB = false, int AH = kAheadJob<T>, int NW = Cfg<T>::NW, typename LA, typename LB,
          typename ST>
__device__ __forceinline__ int gemm_jobs_rs(int M, int N, int K, int first, const LA &la, const LB &lb,
                                            const ST &st) {
  constexpr int NG = Group<T>::n;
  const int ntm = (M + 15) / 16, ntn = (N + 15) / 16;
  if (ntm * ((ntn + NG - 1) / NG) < NW) return gemm_jobs_rs_g<T, 1, CB, AH, NW>(M, N, K, first, la, lb, st);
  return gemm_jobs_rs_g<T, NG, CB, AH, NW>(M, N, K, first, la, lb, st);
}
template <typename T, bool CB = false, int AH = kAheadJob<T>, int NW = Cfg<T>::NW, typename LA, typename LB,
          typename ST>
__device__ __forceinline__ int gemm_jobs_cs(int M, int N, int K, int first, const LA &la, const LB &lb,
                                            const ST &st) {
  constexpr int NG = Group<T>::n;
  const int ntm = (M + 15) / 16, ntn = (N + 15) / 16;
  if (((ntm + NG - 1) / NG) * ntn < NW) return gemm_jobs_cs_g<T, 1, CB, AH, NW>(M, N, K, first, la, lb, st);
  return gemm_jobs_cs_g<T, NG, CB, AH, NW>(M, N, K, first, la, lb, st);
}

// Dynamic deal of the jobs of one phase (the products between two barriers),
// numbered across the phase's products in order: a wave takes its next job
// index from a counter in LDS, one atomic per job, so a wave whose jobs ran
// short (fewer k-steps, cache hits) takes more and the waves reach the
// phase's barrier together.  Which wave runs a job changes nothing in its
// arithmetic, so the results are the same bits as with any other deal.  The
// counter is set to zero by one thread before the barrier that opens the
// phase (no wave takes jobs of the phase before it, and every wave has
// taken its last job of the previous phase before that barrier).
struct Deal {
  int *ctr;  // the counter, in LDS
  int j;     // this wave's current job index (wave-uniform)
  __device__ __forceinline__ int take() {
    int v = 0;
    if ((threadIdx.x & 63) == 0) v = atomicAdd(ctr, 1);
    return __builtin_amdgcn_readfirstlane(v);
  }
  __device__ __forceinline__ void open() { j = take(); }
};
// The product's jobs are [base, base + njobs) of the phase; returns njobs.
template <typename T, int NG, bool CB, int AH, typename LA, typename LB, typename ST>
__device__ __forceinline__ int gemm_jobs_rs_dyn(int M, int N, int K, int base, Deal &d, const LA &la, const LB &lb,
                                                const ST &st) {
  const int ntm = (M + 15) / 16, ntn = (N + 15) / 16, ngr = (ntn + NG - 1) / NG, njobs = ntm * ngr;
  for (; d.j - base < njobs; d.j = d.take()) {
    const int jj = d.j - base, mt = jj / ngr;
    job_rs<T, NG, CB, AH>(M, N, K, mt, jj - mt * ngr, ntn, la, lb, st);
  }
  return njobs;
}
template <typename T, int NG, bool CB, int AH, typename LA, typename LB, typename ST>
__device__ __forceinline__ int gemm_jobs_cs_dyn(int M, int N, int K, int base, Deal &d, const LA &la, const LB &lb,
                                                const ST &st) {
  const int ntm = (M + 15) / 16, ntn = (N + 15) / 16, ngr = (ntm + NG - 1) / NG, njobs = ngr * ntn;
  for (; d.j - base < njobs; d.j = d.take()) {
    const int jj = d.j - base, gt = jj / ntn;
    job_cs<T, NG, CB, AH>(M, N, K, gt, jj - gt * ntn, ntm, la, lb, st);
  }
  return njobs;
}
// Grouped or single-tile jobs, as gemm_jobs_rs / gemm_jobs_cs choose them.
template <typename T, bool CB = false, int AH = kAheadJob<T>, int NW = Cfg<T>::NW, typename LA, typename LB,
          typename ST>
__device__ __forceinline__ int gemm_jobs_rs_d(int M, int N, int K, int base, Deal &d, const LA &la, const LB &lb,
                                              const ST &st) {
  constexpr int NG = Group<T>::n;
  const int ntm = (M + 15) / 16, ntn = (N + 15) / 16;
  if (ntm * ((ntn + NG - 1) / NG) < NW) return gemm_jobs_rs_dyn<T, 1, CB, AH>(M, N, K, base, d, la, lb, st);
  return gemm_jobs_rs_dyn<T, NG, CB, AH>(M, N, K, base, d, la, lb, st);
}
template <typename T, bool CB = false, int AH = kAheadJob<T>, int NW = Cfg<T>::NW, typename LA, typename LB,
          typename ST>
__device__ __forceinline__ int gemm_jobs_cs_d(int M, int N, int K, int base, Deal &d, const LA &la, const LB &lb,
                                              const ST &st) {
  constexpr int NG = Group<T>::n;
  const int ntm = (M + 15) / 16, ntn = (N + 15) / 16;
  if (((ntm + NG - 1) / NG) * ntn < NW) return gemm_jobs_cs_dyn<T, 1, CB, AH>(M, N, K, base, d, la, lb, st);
  return gemm_jobs_cs_dyn<T, NG, CB, AH>(M, N, K, base, d, la, lb, st);
}

// ---- mixed precision: the Z product of Sigma on INT8 slices ----------------
// Ozaki scheme (Ozaki, Ogita, Oishi, Rump, Numer. Algorithms 59 (2012) 95):
// each real operand row (A) or column (B) is x = 2^e sum_s x_s 128^-s with
// int8 slices x_s (split8_kernel); a real product sum_k A B is then
// 2^(eA + eB) sum_{a + b < kZ8Slices} 128^-(a + b) sum_k A_a B_b, each inner sum
// exact in int32 on v_mfma_i32_16x16x64_i8.  The slice pairs of one
// diagonal d = a + b are concatenated along K (K' = (d + 1) kp), so short K
// still fills the 64-deep instruction.  The three real products of the 3M
// complex scheme use their own slices: A comps (re, im, re + im), B comps
// (re, im, re - im) for A conj(B).
typedef int v4i __attribute__((ext_vector_type(4)));

// sum_d 128^-d R_d of the five diagonal accumulators, times 2^ex, as
// 128^-1 (head + tail / 128): the head 128 R0 + R1 is exact in int32 for
// K <= 256 and up to three products summed (|R_0| <= 3 K 127^2, 128 |R_0| +
// |R_1| < 1.6e9); the tail R2 + R3/128 + R4/128^2 is formed in FP32, whose
// rounding (~1e-7 of the tail, ~1e-11 of the value) stays below the 128^-5
// truncation of the slices.
static_assert(kZ8Slices == 5, "z8_value combines five diagonals");
__device__ __forceinline__ double z8_value(const v4i (&acc)[kZ8Slices], int e, int ex) {
  const int head = acc[0][e] * 128 + acc[1][e];
  const float tail = fmaf(fmaf((float)acc[4][e], 1.0f / 128.0f, (float)acc[3][e]), 1.0f / 128.0f, (float)acc[2][e]);
  return ldexp(fma((double)tail, 1.0 / 128.0, (double)head), ex - 7);
}

// Z jobs of one chunk: Z[row][col] = conj(sum_P P[row][P] conj(V[P][col])),
// rows (tb, r) -> P row Q0 + r at time t0 + tb, cols (jj, l) -> V column
// (J0 + jj) k2 + n0 + l.  Single-tile jobs, the first product of the phase
// (jobs [0, njobs) of the dynamic deal); returns the job count.  Per real product
// and 64-wide chunk of K the five A and five B slices are loaded once and the
// 15 slice pairs with a + b < 5 issued from registers; the three products'
// sums combine in int32 and each Z element is reconstructed and written once.
template <typename T>
__device__ __forceinline__ int z8_jobs(int M, int N, Deal &d, const int8_t *__restrict__ P8,
                                       const int *__restrict__ PE, int64_t p8_tstride, int pe_tstride, int q,
                                       int kp, int Q0, FDiv dnr, int nr, const int8_t *__restrict__ V8,
                                       const int *__restrict__ VE, int ncol, int J0, int n0, int k2, FDiv dnn,
                                       int nn, int Jb, int ldz, cplx<T> *bufZ) {
  constexpr int NS = kZ8Slices;
  const int lane = threadIdx.x & 63;
  const int ntm = (M + 15) / 16, ntn = (N + 15) / 16, njobs = ntm * ntn;
  const int64_t ss_a = (int64_t)q * kp, ss_b = (int64_t)ncol * kp;  // slice strides
  const int nch = (kp + 63) / 64, kg = 16 * (lane >> 4);
  for (; d.j < njobs; d.j = d.take()) {
    const int j = d.j, mt = j / ntn, nt = j - mt * ntn;
    // this lane's operand row and column
    const int arow = min(mt * 16 + (lane & 15), M - 1), acol = min(nt * 16 + (lane & 15), N - 1);
    const int atb = dnr.div(arow), ar = arow - atb * nr;
    const int bjj = dnn.div(acol), bl = acol - bjj * nn;
    const int8_t *Arow = P8 + atb * p8_tstride + (int64_t)(Q0 + ar) * kp + kg;
    const int8_t *Bcol = V8 + (int64_t)((J0 + bjj) * k2 + n0 + bl) * kp + kg;
    // output element e of this lane: row 4 (lane / 16) + e, column lane % 16
    const int col = nt * 16 + (lane & 15), cc = min(col, N - 1), ojj = dnn.div(cc), ol = cc - ojj * nn;
    const int vcol = (J0 + ojj) * k2 + n0 + ol;
    // Z = conj(A conj(B)) = (P1 + P2) - i (P3 - P1 + P2).  A row (B column)
    // has one exponent for its three components, so the products' integer
    // sums combine exactly before the one reconstruction per part:
    // X = P1 + P2 (the real part) and Y = P3 - P1 + P2 (minus the imaginary
    // part), Y started as -P1 and the P2 MFMAs issued into both.
    v4i X[NS], Y[NS];
#pragma unroll
    for (int dd = 0; dd < NS; ++dd) X[dd] = v4i{0, 0, 0, 0};
    auto comp = [&](int c, v4i (&A1)[NS], v4i (&A2)[NS], bool two) {
      const int8_t *Ac = Arow + c * NS * ss_a, *Bc = Bcol + c * NS * ss_b;
      for (int ch = 0; ch < nch; ++ch) {
        // unconditional loads (k beyond kp reads the row's first bytes and is
        // zeroed at use), all ten issued before the MFMAs wait on them
        const bool in = ch * 64 + kg < kp;
        const int off = in ? ch * 64 : -kg;
        v4i a[NS], b[NS];
#pragma unroll
        for (int sl = 0; sl < NS; ++sl) {
          a[sl] = *reinterpret_cast<const v4i *>(Ac + sl * ss_a + off);
          b[sl] = *reinterpret_cast<const v4i *>(Bc + sl * ss_b + off);
        }
#pragma unroll
        for (int sl = 0; sl < NS; ++sl)
          if (!in) a[sl] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int sa = 0; sa < NS; ++sa)
#pragma unroll
          for (int sb = 0; sa + sb < NS; ++sb) {
            A1[sa + sb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[sa], b[sb], A1[sa + sb], 0, 0, 0);
            if (two) A2[sa + sb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[sa], b[sb], A2[sa + sb], 0, 0, 0);
          }
      }
    };
    comp(0, X, Y, false);                       // X = P1
#pragma unroll
    for (int dd = 0; dd < NS; ++dd) Y[dd] = v4i{0, 0, 0, 0} - X[dd];  // Y = -P1
    comp(1, X, Y, true);                        // X = P1 + P2, Y = P2 - P1
    comp(2, Y, Y, false);                       // Y = P3 - P1 + P2
    const int eb = VE[vcol];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = mt * 16 + 4 * (lane >> 4) + e;
      if (col >= N || row >= M) continue;
      const int otb = dnr.div(row), orr = row - otb * nr;
      const int ex = PE[otb * pe_tstride + Q0 + orr] + eb;
      bufZ[(otb * Jb + ojj) * ldz + orr * nn + ol] = cplx<T>{(T)z8_value(X, e, ex), (T)-z8_value(Y, e, ex)};
    }
  }
  return njobs;
}

}  // namespace

// Split of real-matrix rows into INT8 slices (Split8Blk), one wave per row:
// the row maximum m over its three components gives one e with m 2^-e <= 127
// for all three (so that the Z product can add their integer sums), then
// x 2^-e = sum_{s < kZ8Slices} x_s 128^-s + r with |r| <= 128^-(kZ8Slices-1)/2:
// rint, an exact subtraction and an exact scaling per slice.
__global__ __launch_bounds__(256) void split8_kernel(const Split8Blk *__restrict__ sb, int nsb,
                                                     const int64_t *__restrict__ row_prefix, int64_t nrows_total,
                                                     const cplx<double> *__restrict__ src, int8_t *__restrict__ dst,
                                                     int *__restrict__ ex) {
  const int lane = threadIdx.x & 63;
  for (int64_t row = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6; row < nrows_total;
       row += ((int64_t)gridDim.x * 256) >> 6) {
    int lo = 0, hi = nsb - 1;  // block: row_prefix[b] <= row < row_prefix[b + 1]
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (row_prefix[mid] <= row) lo = mid; else hi = mid - 1;
    }
    const Split8Blk S = sb[lo];
    const int64_t local = row - row_prefix[lo];
    const int u = (int)(local / S.nrows), r = (int)(local - (int64_t)u * S.nrows);
    const cplx<double> *x = src + S.src_off + (int64_t)u * S.sub_stride + (int64_t)r * S.row_stride;
    const double sg = S.kind == 0 ? 1.0 : -1.0;
    double m[kZ8Comps] = {0.0, 0.0, 0.0};
    for (int k = lane; k < S.K; k += 64) {
      const cplx<double> v = x[(int64_t)k * S.k_stride];
      m[0] = fmax(m[0], fabs(v.re));
      m[1] = fmax(m[1], fabs(v.im));
      m[2] = fmax(m[2], fabs(v.re + sg * v.im));
    }
#pragma unroll
    for (int c = 0; c < kZ8Comps; ++c)
      for (int o = 32; o > 0; o >>= 1) m[c] = fmax(m[c], __shfl_xor(m[c], o));
    const double mm = fmax(m[0], fmax(m[1], m[2]));
    int xe = 0;
    if (mm > 0.0) (void)frexp(mm / 127.0, &xe);
    const int e[kZ8Comps] = {xe, xe, xe};
    int8_t *d = dst + S.dst_off + (int64_t)u * kZ8Comps * kZ8Slices * S.nrows * S.kp + (int64_t)r * S.kp;
    if (lane < kZ8Comps) ex[S.e_off + (int64_t)u * 3 * S.nrows + (int64_t)lane * S.nrows + r] = e[lane];
    for (int k = lane; k < S.kp; k += 64) {
      cplx<double> v{0.0, 0.0};
      if (k < S.K) v = x[(int64_t)k * S.k_stride];
#pragma unroll
      for (int c = 0; c < kZ8Comps; ++c) {
        double t = ldexp(c == 0 ? v.re : c == 1 ? v.im : v.re + sg * v.im, -e[c]);
#pragma unroll
        for (int sl = 0; sl < kZ8Slices; ++sl) {
          const double w = rint(t);
          d[(int64_t)(c * kZ8Slices + sl) * S.nrows * S.kp + k] = (int8_t)(int)w;
          t = (t - w) * 128.0;
        }
      }
    }
  }
}

hipError_t launch_split8(const Split8Blk *sb, int nsb, const int64_t *row_prefix, int64_t nrows_total,
                         const cplx<double> *src, int8_t *dst, int *ex, hipStream_t st) {
  if (nsb == 0 || nrows_total == 0) return hipSuccess;
  const int64_t waves = nrows_total, grid = std::min<int64_t>((waves + 3) / 4, 65536);
  hipLaunchKernelGGL(split8_kernel, dim3((unsigned)grid), dim3(256), 0, st, sb, nsb, row_prefix, nrows_total, src,
                     dst, ex);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// P0.  Slab = rows [Q0, Q0 + nband) (nband <= 16) of aux block iq of iBZ point
// q, for p0_times(MODE) consecutive times t < nts/2.  For every V block (k, ik1, ik2)
// and spin:
//   E_Q = V_Q G^{k+q}(t),  C_Q = G^k(beta - t) E_Q,  P0[Q][P] += sum_x C_Q[x] conj(V_P[x])
// i.e. the three GEMMs of cugw_qkpt_block.cu:231-260 (X1 = G V, X2 = conj(V) G^T,
// P0^T += X2 X1) regrouped around one intermediate per Q row.  All rows of the
// slab go through each product together, in chunks of nsub orbitals n of k+q
// sized for the LDS (E and C of a chunk stay in LDS; V is read from memory).
// When every wave can hold one column tile (q <= 16 * waves) the times share
// every V fragment: their G^{k+q} blocks are stacked as columns of one E
// product and their C rows as the row tiles of the rank-K update; otherwise
// the times are processed one after the other.
// ---------------------------------------------------------------------------
template <typename T, int MODE>
__global__ __launch_bounds__(Cfg<T>::NT) void p0_mfma_kernel(const P0Item *__restrict__ items, const P0Blk *__restrict__ blks,
                                                    int64_t n_work, int nt_half, int nts, int ns,
                                                    const cplx<T> *__restrict__ V, const cplx<T> *__restrict__ G,
                                                    cplx<double> *__restrict__ P0, cplx<double> *__restrict__ part,
                                                    double pref) {
  constexpr int TB = p0_times((int)sizeof(cplx<T>), MODE);
  const int64_t w = xcd_logical(blockIdx.x, gridDim.x);
  if (w >= n_work) return;
  const int np = (nt_half + TB - 1) / TB;
  const int t0 = (int)(w % np) * TB, ntimes = min(TB, nt_half - t0);
  const P0Item it = items[w / np];
  constexpr int NW = Cfg<T>::NW, NT = Cfg<T>::NT, MT = Cfg<T>::MT;
  // accumulators: one per stacked time, at least the MT column tiles of the unstacked passes
  constexpr int MTP = TB > MT ? TB : MT;
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // the job counter of the dynamic deal, then the data (from the start of
  // the LDS under the static deal)
  Deal deal{reinterpret_cast<int *>(smem_raw), 0};
  cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw + (kP0Dyn<T> ? kDealBytes : 0));
  const int q = it.q, nband = it.nband, nct = (q + 15) / 16;
  // MODE 1: every slab stacked, 0: none, 2: per slab
  const bool stack = MODE == 2 ? nct <= NW : MODE == 1;
  const int npass = stack ? 1 : ntimes, tbn = stack ? ntimes : 1;
  // column tiles: S waves share a tile over S parts of x.  Stacked: one column
  // tile per wave (tile tg) with one accumulator per time; otherwise up to MT
  // column tiles per wave (tg, tg + TG, ...) for one time.
  const int S = stack ? choose_split<NW, 1>(nct) : choose_split<NW, MT>(nct);
  const int TG = NW / S, kp = wave % S, tg = wave / S;
  const int nmine = stack ? (tg < nct ? tbn : 0) : (tg < TG && nct > tg ? (nct - tg + TG - 1) / TG : 0);

  for (int pass = 0; pass < npass; ++pass) {
    const int ta = t0 + pass;  // first time of this pass
    Tile<T> acc[MTP];
#pragma unroll
    for (int c = 0; c < MTP; ++c) tile_zero(acc[c]);
    for (int b = it.blk_begin; b < it.blk_end; ++b) {
      const P0Blk B = blks[b];
      const int k1 = B.k1, k2 = B.k2, K = k1 * k2, nsub = B.nsub, lde = B.lde;
      const FDiv dk1(k1);
      const cplx<T> *Vb = V + B.v_off;
      const cplx<T> *Vr = Vb + (int64_t)it.Q0 * K;        // V rows of the band
      cplx<T> *bufE = lds;                                // [tbn][nband][lde]  E[r][a][l] at a*nn + l
      cplx<T> *bufC = lds + TB * nband * lde;             // [tbn][nband][lde]  C[r][p][l] at p*nn + l
      const int g1ts = B.g1_tstride, g2ts = B.g2_tstride;
      for (int s = 0; s < ns; ++s) {
        // time ta + tb: G1 at nts - 1 - (ta + tb), G2 at ta + tb
        const cplx<T> *G1 = G + B.g1_base + (int64_t)(nts - 1 - ta) * g1ts + (int64_t)s * B.g1_sstride;
        const cplx<T> *G2 = G + B.g2_base + (int64_t)ta * g2ts + (int64_t)s * B.g2_sstride;
        for (int n0 = 0; n0 < k2; n0 += nsub) {
          const int nn = min(nsub, k2 - n0), X = k1 * nn;
          const FDiv dnn(nn);
          // under the dynamic deal the E and C jobs of the chunk are one run
          // of it (C's continue E's numbering after the barrier between them)
          if constexpr (kP0Dyn<T>) {
            if (tid == 0) *deal.ctr = 0;
          }
          __syncthreads();
          if constexpr (kP0Dyn<T>) deal.open();
          int ne;
          // E_tb[(r,a)][l] = sum_m V[Q0+r][a][m] G2_tb[m][n0+l], columns (tb, l)
          {
            auto la = [&](int row, int m) {
              row = min(row, nband * k1 - 1);
              return Vr[(int64_t)row * k2 + min(m, k2 - 1)];
            };
            auto lb = [&](int m, int col) {
              col = min(col, tbn * nn - 1);
              const int tb = dnn.div(col), l = col - tb * nn;
              return G2[(int64_t)tb * g2ts + min(m, k2 - 1) * k2 + n0 + l];
            };
            auto st = [&](int row, int col, cplx<T> v) {
              const int r = dk1.div(row), a = row - r * k1, tb = dnn.div(col), l = col - tb * nn;
              bufE[(tb * nband + r) * lde + a * nn + l] = v;
            };
            if constexpr (kP0Dyn<T>) {
              if (tbn * nn > 16) ne = gemm_jobs_rs_d<T, false, 1>(nband * k1, tbn * nn, k2, 0, deal, la, lb, st);
              else ne = gemm_jobs_cs_d<T, false, 1>(nband * k1, tbn * nn, k2, 0, deal, la, lb, st);
            } else {
              if (tbn * nn > 16) ne = gemm_jobs_rs<T, false, 1>(nband * k1, tbn * nn, k2, 0, la, lb, st);
              else ne = gemm_jobs_cs<T, false, 1>(nband * k1, tbn * nn, k2, 0, la, lb, st);
            }
          }
          __syncthreads();
          // C_tb[r][p][l] = sum_a G1_tb[p][a] E_tb[r][a][l]  (the times' jobs numbered in
          // turn; the static deal starts over at wave 0)
          {
            int base = kP0Dyn<T> ? ne : 0;
            for (int tb = 0; tb < tbn; ++tb) {
              const cplx<T> *G1t = G1 - (int64_t)tb * g1ts;
              const cplx<T> *Et = bufE + tb * nband * lde;
              cplx<T> *Ct = bufC + tb * nband * lde;
              auto la = [&](int p, int a) { return G1t[min(p, k1 - 1) * k1 + min(a, k1 - 1)]; };
              auto lb = [&](int a, int col) {
                col = min(col, nband * nn - 1);
                const int r = dnn.div(col), l = col - r * nn;
                return Et[r * lde + min(a, k1 - 1) * nn + l];
              };
              auto st = [&](int p, int col, cplx<T> v) {
                const int r = dnn.div(col), l = col - r * nn;
                Ct[r * lde + p * nn + l] = v;
              };
              if constexpr (kP0Dyn<T>) base += gemm_jobs_rs_d<T, false, 1>(k1, nband * nn, k1, base, deal, la, lb, st);
              else base += gemm_jobs_rs<T, false, 1>(k1, nband * nn, k1, base, la, lb, st);
            }
          }
          __syncthreads();
          // acc[r][P] += sum_{(p,l)} C[r][p][l] conj(V[P][p][n0+l]) over this wave's part of (p, l)
          {
            int kb, ke;
            k_part(X, S, kp, kb, ke);
            const int rc = min(threadIdx.x & 15, nband - 1);   // rows >= nband are never stored
            if (stack) {
              // one V fragment per k-step for the tbn time tiles of column tile tg, one step
              // ahead (with two, 4 registers spilled in FP64 and 8-11 in FP32; FP32 P0 +2%)
              if (nmine > 0)
                mma_colshare<0, TB, true, 1>(
                    acc, tbn, kb, ke, [&](int c, int i, int x) { return bufC[(c * nband + rc) * lde + min(x, X - 1)]; },
                    [&](int x, int j) {
                      x = min(x, X - 1);
                      const int p = dnn.div(x), l = x - p * nn;
                      return Vb[(int64_t)min(tg * 16 + j, q - 1) * K + p * k2 + n0 + l];
                    });
            } else {
              mma_rowshare_all<true>(
                  acc, nmine, kb, ke, true, [&](int i, int x) { return bufC[rc * lde + min(x, X - 1)]; },
                  [&](int c, int x, int j) {
                    x = min(x, X - 1);
                    const int p = dnn.div(x), l = x - p * nn;
                    const int P = min((tg + TG * c) * 16 + j, q - 1);
                    return Vb[(int64_t)P * K + p * k2 + n0 + l];
                  });
            }
          }
        }
      }
    }
    if (S > 1) split_reduce(acc, nmine, S, reinterpret_cast<T *>(lds));
    if (kp == 0) {
#pragma unroll
      for (int c = 0; c < MTP; ++c) {
        if (c < nmine) {
          const int t = stack ? ta + c : ta, jt = stack ? tg : tg + TG * c;
          if (it.part_out < 0) {
            cplx<double> *out = P0 + it.out_base + (int64_t)t * it.out_tstride;
            tile_store(acc[c], [&](int i, int j, cplx<T> v) {
              const int P = jt * 16 + j;
              if (i < nband && P < q) out[i * q + P] = cplx<double>{pref * (double)v.re, pref * (double)v.im};
            });
          } else {
            // one part of a split item: the unscaled partial sums of time t of this slab
            cplx<double> *out = part + it.part_out + ((int64_t)(w % np) * TB + (t - t0)) * nband * q;
            tile_store(acc[c], [&](int i, int j, cplx<T> v) {
              const int P = jt * 16 + j;
              if (i < nband && P < q) out[i * q + P] = cplx<double>{(double)v.re, (double)v.im};
            });
          }
        }
      }
    }
    // rows t >= nts/2 of the reference's device P0 stay zero (cugw_routines_block.cu:79)
    for (int tb = 0; tb < tbn; ++tb) {
      cplx<double> *mir = P0 + it.out_base + (int64_t)(nts - 1 - (ta + tb)) * it.out_tstride;
      for (int e = tid; e < nband * q; e += NT) mir[e] = cplx<double>{0.0, 0.0};
    }
    __syncthreads();
  }
}

// The parts of every split P0 item of the shard [shard_first, shard_end) of
// iBZ points: P0 = pref * (part 0 + part 1 + ...), summed in part order, one
// workgroup per split item (times t < nts/2; the kernel zeroes the others).
__global__ __launch_bounds__(256) void p0_reduce_kernel(const P0Split *__restrict__ sp, int shard_first,
                                                        int shard_end, const cplx<double> *__restrict__ part,
                                                        cplx<double> *__restrict__ P0, double pref, int nt_half) {
  const P0Split S = sp[blockIdx.x];
  if (S.iqr < shard_first || S.iqr >= shard_end) return;
  const int rows = S.nband * S.q;
  const int64_t n = (int64_t)S.np * S.TB * rows;
  for (int64_t e = threadIdx.x; e < n; e += 256) {
    const int t = (int)(e / rows), r = (int)(e - (int64_t)t * rows);
    if (t >= nt_half) continue;  // past the last time of the last slab: never written
    double re = 0.0, im = 0.0;
    for (int p = 0; p < S.nparts; ++p) {
      const cplx<double> v = part[S.part_out + (int64_t)p * S.part_stride + e];
      re += v.re;
      im += v.im;
    }
    P0[S.out_base + (int64_t)t * S.out_tstride + r] = cplx<double>{pref * re, pref * im};
  }
}

hipError_t launch_p0_reduce(const P0Split *sp, int nsp, int shard_first, int shard_end, const cplx<double> *part,
                            cplx<double> *P0, double pref, int nts, hipStream_t st) {
  if (nsp == 0) return hipSuccess;
  hipLaunchKernelGGL(p0_reduce_kernel, dim3((unsigned)nsp), dim3(256), 0, st, sp, shard_first, shard_end, part, P0,
                     pref, nts / 2);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Sigma.  Slab = rows [I0, I0+Ib) x columns [J0, J0+Jb) of ao block ik1 of iBZ
// point k, spin s, for TB consecutive times t0 + tb.  For every V block
// (q, iq, ik2), in chunks of qsub aux rows Q by nsub orbitals n (E and Z stored
// with the (Q, n) reduction index contiguous):
//   Z_Q(t) = sum_P conj(P_{QP}(t)) V_P,  E_Q(t) = V_Q G^{k-q}(t),
//   Sigma(t)[i][j] += sum_{Q,n} E_Q(t)[i][n] conj(Z_Q(t)[j][n])
// the contraction of cugw_qkpt_block.cu:381-407 (Y1 = V G, Y2 = Y1 P,
// Sigma += Y2 V^H) with the aux contraction moved onto V so that the largest
// GEMM (q x q x k1 k2) reads V straight from memory.  The times of a slab
// share every V fragment: their P rows are stacked as rows of one Z product
// and their G^{k-q} blocks as columns of one E product.
// ---------------------------------------------------------------------------
// Sigma workgroup shape: the contraction shape of T, or, with the INT8 Z
// product, kZ8Waves waves of kZ8Tiles accumulator tiles.
template <typename T, bool Z8>
struct SigCfg {
  static constexpr int NW = Z8 ? kZ8Waves : Cfg<T>::NW;
  static constexpr int MT = Z8 ? kZ8Tiles : Cfg<T>::MT;
  static constexpr int NT = 64 * NW;
};

template <typename T, int TB, bool Z8>
__global__ __launch_bounds__((SigCfg<T, Z8>::NT)) void sigma_mfma_kernel(const SigItem *__restrict__ items,
                                                       const SigBlk *__restrict__ blks, int64_t n_work, int nts,
                                                       int ns, const cplx<T> *__restrict__ V,
                                                       const cplx<T> *__restrict__ G, const cplx<T> *__restrict__ Pf,
                                                       cplx<double> *__restrict__ Sig, cplx<double> *__restrict__ part,
                                                       double pref, Z8Ops z8) {
  const int64_t w = xcd_logical(blockIdx.x, gridDim.x);
  if (w >= n_work) return;
  const int np = (nts + TB - 1) / TB, per = np * ns;
  const SigItem it = items[w / per];
  const int ts = (int)(w % per), t0 = (ts / ns) * TB, s = ts - (ts / ns) * ns;
  const int tbn = min(TB, nts - t0);  // times of this slab
  constexpr int NW = SigCfg<T, Z8>::NW, MT = SigCfg<T, Z8>::MT, NT = SigCfg<T, Z8>::NT;
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // the job counter of the dynamic deal, then the data
  Deal deal{reinterpret_cast<int *>(smem_raw), 0};
  cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw + kDealBytes);
  const int k1 = it.k1, I0 = it.I0, Ib = it.Ib, J0 = it.J0, Jb = it.Jb;
  const int nti = (Ib + 15) / 16, ntj = (Jb + 15) / 16, ntile = nti * ntj, nall = tbn * ntile;
  // output tiles u = tb * ntile + tt: S waves share a tile over S parts of the
  // (Q, n) reduction
  const int S = choose_split<NW, MT>(nall), TG = NW / S, kp = wave % S, tg = wave / S;
  const int nmine = tg < TG && nall > tg ? (nall - tg + TG - 1) / TG : 0;  // tiles tg, tg + TG, ...
  Tile<T> acc[MT];
#pragma unroll
  for (int c = 0; c < MT; ++c) tile_zero(acc[c]);

  int zend_prev = 0;  // end of the previous block's Z and E buffers (elements)
  for (int b = it.blk_begin; b < it.blk_end; ++b) {
    const SigBlk B = blks[b];
    const int q = B.q, k2 = B.k2, K = k1 * k2, qsub = B.qsub, nsub = B.nsub, ldz = B.ldz, lde = B.lde;
    const cplx<T> *Vb = V + B.v_off;
    const cplx<T> *G2 = G + B.g2_base + (int64_t)t0 * B.g2_tstride + (int64_t)s * B.g2_sstride;
    const cplx<T> *Pb = Pf + B.p_base + (int64_t)t0 * B.p_tstride;
    const int g2ts = B.g2_tstride, pts = B.p_tstride;
    // y = r*nn + l indexes (aux row of the chunk, orbital n0 + l of k - q)
    cplx<T> *bufZ = lds;                      // [TB][Jb][ldz]  Z[r][j][n] at j*ldz + y
    cplx<T> *bufE = bufZ + TB * Jb * ldz;     // [TB][Ib][lde]  E[r][i][n] at i*lde + y
    cplx<T> *bufG = bufE + TB * Ib * lde;     // [k2][tbn * nn]  G2_tb[m][n0 + l] at m * tbn * nn + tb * nn + l
    // The G chunk is staged before the chunk's first barrier, while slower
    // waves may still run the previous block's final product on its Z and E
    // buffers: where this block's G buffer starts below their end, wait.
    const int zend = TB * Jb * ldz + TB * Ib * lde;
    if (zend < zend_prev) __syncthreads();
    zend_prev = zend;
    const cplx<T> *Vj = Vb + J0 * k2;
    const FDiv dIb(Ib);
    for (int Q0 = 0; Q0 < q; Q0 += qsub) {
      const int nr = min(qsub, q - Q0);
      const FDiv dnr(nr);
      for (int n0 = 0; n0 < k2; n0 += nsub) {
        const int nn = min(nsub, k2 - n0), zc = Jb * nn, NC = tbn * nn;
        const FDiv dnn(nn);
        // G chunk of the E product: every wave has passed the barrier after
        // the previous chunk's E, the only reader of bufG; the barrier below
        // also ends the previous chunk's final product, the reader of bufZ and
        // bufE.  The Z and E jobs that follow it are one phase of the deal.
        if (tid == 0) *deal.ctr = 0;
        {
          const FDiv dNC(NC);
          for (int e = tid; e < k2 * NC; e += NT) {
            const int m = dNC.div(e), c = e - m * NC, tb = dnn.div(c), l = c - tb * nn;
            bufG[e] = G2[(int64_t)tb * g2ts + m * k2 + n0 + l];
          }
        }
        __syncthreads();
        deal.open();
        // Z[(tb,r)][(j,l)] = sum_P conj(P_tb[Q0+r][P]) V[P][J0+j][n0+l], formed as the
        // conjugate of sum_P P_tb[Q0+r][P] conj(V[P][J0+j][n0+l]) with P read
        // straight from memory (rows of the chunk, contiguous in P)
        int nj;
        {
          auto la = [&](int row, int P) {
            row = min(row, tbn * nr - 1);
            const int tb = dnr.div(row), r = row - tb * nr;
            return Pb[(int64_t)tb * pts + (int64_t)(Q0 + r) * q + min(P, q - 1)];
          };
          auto lb = [&](int P, int x) {
            x = min(x, zc - 1);
            const int jj = dnn.div(x), l = x - jj * nn;
            return Vj[(int64_t)min(P, q - 1) * K + jj * k2 + n0 + l];
          };
          auto st = [&](int row, int x, cplx<T> v) {
            const int tb = dnr.div(row), r = row - tb * nr, jj = dnn.div(x), l = x - jj * nn;
            bufZ[(tb * Jb + jj) * ldz + r * nn + l] = cplx<T>{v.re, -v.im};
          };
          // several row tiles: share each V fragment among them (always in
          // FP64, where one job layout per product saves 4 spilled registers)
          if (Z8 && B.kp > 0) {
            // mixed precision: INT8 slices of P (this slab's first time) and V
            const int64_t p8t = (int64_t)kZ8Comps * kZ8Slices * q * B.kp;
            nj = z8_jobs(tbn * nr, zc, deal, z8.P8 + B.p8_off + t0 * p8t, z8.PE + B.pe_off + (int64_t)t0 * 3 * q,
                             p8t, 3 * q, q, B.kp, Q0, dnr, nr, z8.V8 + B.v8_off, z8.VE + B.ve_off, K, J0, n0, k2,
                             dnn, nn, Jb, ldz, bufZ);
          } else if (kSigOneLayout<T> || tbn * nr > 16) {
            nj = gemm_jobs_cs_d<T, true, kAheadJob<T>, NW>(tbn * nr, zc, q, 0, deal, la, lb, st);
          } else {
            nj = gemm_jobs_rs_d<T, true, kAheadJob<T>, NW>(tbn * nr, zc, q, 0, deal, la, lb, st);
          }
        }
        // E[(r,i)][(tb,l)] = sum_m V[Q0+r][I0+i][m] G2_tb[m][n0+l]  (jobs dealt on after Z's)
        {
          auto la = [&](int row, int m) {
            row = min(row, nr * Ib - 1);
            const int r = dIb.div(row), ii = row - r * Ib;
            return Vb[(int64_t)(Q0 + r) * K + (I0 + ii) * k2 + min(m, k2 - 1)];
          };
          auto lb = [&](int m, int col) { return bufG[min(m, k2 - 1) * NC + min(col, NC - 1)]; };
          auto st = [&](int row, int col, cplx<T> v) {
            const int r = dIb.div(row), ii = row - r * Ib, tb = dnn.div(col), l = col - tb * nn;
            bufE[(tb * Ib + ii) * lde + r * nn + l] = v;
          };
          if (kSigOneLayout<T> || tbn * nn > 16)
            gemm_jobs_rs_d<T, false, kAheadJob<T>, NW>(nr * Ib, tbn * nn, k2, nj, deal, la, lb, st);
          else
            gemm_jobs_cs_d<T, false, kAheadJob<T>, NW>(nr * Ib, tbn * nn, k2, nj, deal, la, lb, st);
        }
        __syncthreads();
        // acc[i][j] += sum_y E_tb[i][y] conj(Z_tb[j][y]) over this wave's part of y
        const int Y = nr * nn;
        int yb, ye;
        k_part(Y, S, kp, yb, ye);
        mma_multi_all<true>(
            acc, nmine, yb, ye,
            [&](int c, int i, int y) {
              const int u = tg + TG * c, tb = u / ntile, i0 = ((u - tb * ntile) / ntj) * 16;
              return bufE[(tb * Ib + min(i0 + i, Ib - 1)) * lde + min(y, Y - 1)];
            },
            [&](int c, int y, int j) {
              const int u = tg + TG * c, tb = u / ntile, j0 = ((u - tb * ntile) % ntj) * 16;
              return bufZ[(tb * Jb + min(j0 + j, Jb - 1)) * ldz + min(y, Y - 1)];
            });
      }
    }
  }
  if (S > 1) split_reduce(acc, nmine, S, reinterpret_cast<T *>(lds));
  if (kp == 0) {
#pragma unroll
    for (int c = 0; c < MT; ++c) {
      const int u = tg + TG * c, tb = u / ntile, tt = u - tb * ntile;
      if (c < nmine) {
        const int i0 = (tt / ntj) * 16, j0 = (tt % ntj) * 16;
        if (it.part_out < 0) {
          cplx<double> *out = Sig + it.out_base + (int64_t)(t0 + tb) * it.out_tstride + (int64_t)s * it.out_sstride;
          tile_store(acc[c], [&](int i, int j, cplx<T> v) {
            if (i0 + i < Ib && j0 + j < Jb)
              out[(int64_t)(i0 + i) * k1 + J0 + j0 + j] = cplx<double>{pref * (double)v.re, pref * (double)v.im};
          });
        } else {
          // one part of a split item: the unscaled partial sums of this slab
          cplx<double> *out = part + it.part_out + ((int64_t)ts * TB + tb) * Ib * Jb;
          tile_store(acc[c], [&](int i, int j, cplx<T> v) {
            if (i0 + i < Ib && j0 + j < Jb) out[(i0 + i) * Jb + j0 + j] = cplx<double>{(double)v.re, (double)v.im};
          });
        }
      }
    }
  }
}

// The parts of every split Sigma item of the shard [shard_first, shard_end)
// of iBZ points: Sigma = pref * (part 0 + part 1 + ...), summed in part order
// (so the result does not depend on which part finished first), one
// workgroup per split item.
__global__ __launch_bounds__(256) void sig_reduce_kernel(const SigSplit *__restrict__ sp, int shard_first,
                                                         int shard_end, const cplx<double> *__restrict__ part,
                                                         cplx<double> *__restrict__ Sig, double pref, int nts) {
  const SigSplit S = sp[blockIdx.x];
  if (S.ikr < shard_first || S.ikr >= shard_end) return;
  const int ij = S.Ib * S.Jb;
  const int64_t n = (int64_t)S.per * S.TB * ij;
  for (int64_t e = threadIdx.x; e < n; e += 256) {
    const int64_t st = e / ij;
    const int r = (int)(e - st * ij), i = r / S.Jb, j = r - i * S.Jb;
    const int ts = (int)(st / S.TB), tb = (int)(st - (int64_t)ts * S.TB);
    const int t = (ts / S.ns) * S.TB + tb, s = ts - (ts / S.ns) * S.ns;
    if (t >= nts) continue;  // past the last time of the last slab: never written
    double re = 0.0, im = 0.0;
    for (int p = 0; p < S.nparts; ++p) {
      const cplx<double> v = part[S.part_out + (int64_t)p * S.part_stride + e];
      re += v.re;
      im += v.im;
    }
    Sig[S.out_base + (int64_t)t * S.out_tstride + (int64_t)s * S.out_sstride + (int64_t)i * S.k1 + S.J0 + j] =
        cplx<double>{pref * re, pref * im};
  }
}

hipError_t launch_sig_reduce(const SigSplit *sp, int nsp, int shard_first, int shard_end, const cplx<double> *part,
                             cplx<double> *Sig, double pref, int nts, hipStream_t st) {
  if (nsp == 0) return hipSuccess;
  hipLaunchKernelGGL(sig_reduce_kernel, dim3((unsigned)nsp), dim3(256), 0, st, sp, shard_first, shard_end, part, Sig,
                     pref, nts);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Rotation of one block of one (t, s) slice from the iBZ representative to a
// full-BZ point (cugw_routines_block.cu:134-186, cugw_qkpt_block.cu:288-340):
//   Y = O X O^H (O is m x n), or a copy, conjugated for time-reversal partners.
// On the matrix cores: W = X O^H into LDS (in column chunks of O's rows), then
// Y = O W straight to memory.  src: iBZ array; the (t,s) slice of the
// representative of target k starts at src_slice[k] + ts * src_ts_stride[k].
// dst: k_ts_ij over the full BZ.
// ---------------------------------------------------------------------------
template <typename TO>
__global__ __launch_bounds__(Cfg<double>::NT) void rotate_mfma_kernel(
    const RotBlk *__restrict__ rb, int n_rb, int nts_ns, const cplx<double> *__restrict__ src,
    const int64_t *__restrict__ src_ts_stride, const int64_t *__restrict__ src_slice,
    const cplx<double> *__restrict__ rep, cplx<TO> *__restrict__ dst, const int64_t *__restrict__ dst_slice,
    const int64_t *__restrict__ dst_slice_size) {
  constexpr int NT = Cfg<double>::NT;
  const int64_t w = xcd_logical(blockIdx.x, gridDim.x);
  if (w >= (int64_t)n_rb * nts_ns) return;
  const int r = (int)(w / nts_ns), ts = (int)(w % nts_ns);
  const RotBlk R = rb[r];
  const int tid = threadIdx.x;
  const cplx<double> *X = src + (int64_t)ts * src_ts_stride[R.target] + src_slice[R.target] + R.src_off;
  cplx<TO> *Y = dst + dst_slice[R.target] * nts_ns + (int64_t)ts * dst_slice_size[R.target] + R.dst_off;
  const int m = R.m, n = R.n;
  const double sgn = R.conj ? -1.0 : 1.0;
  if (R.rep_off < 0) {
    for (int i = tid; i < m * m; i += NT) {
      const cplx<double> v = X[i];
      Y[i] = cplx<TO>{(TO)v.re, (TO)(sgn * v.im)};
    }
    return;
  }
  const cplx<double> *O = rep + R.rep_off;  // m x n
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<double> *Wb = reinterpret_cast<cplx<double> *>(smem_raw);  // [n][jc]
  const int jc_max = max(16, min(m, (kRotLds / 16) / n));
  for (int j0 = 0; j0 < m; j0 += jc_max) {
    const int jc = min(jc_max, m - j0);
    __syncthreads();
    // W[a][jj] = sum_b X[a][b] conj(O[j0+jj][b])
    gemm_jobs_rs<double, true>(
        n, jc, n, 0, [&](int a, int b) { return X[min(a, n - 1) * n + min(b, n - 1)]; },
        [&](int b, int jj) { return O[(j0 + min(jj, jc - 1)) * n + min(b, n - 1)]; },
        [&](int a, int jj, cplx<double> v) { Wb[a * jc + jj] = v; });
    __syncthreads();
    // Y[i][j0+jj] = sum_a O[i][a] W[a][jj]
    gemm_jobs_rs<double>(
        m, jc, n, 0, [&](int i, int a) { return O[min(i, m - 1) * n + min(a, n - 1)]; },
        [&](int a, int jj) { return Wb[min(a, n - 1) * jc + min(jj, jc - 1)]; },
        [&](int i, int jj, cplx<double> v) { Y[i * m + j0 + jj] = cplx<TO>{(TO)v.re, (TO)(sgn * v.im)}; });
  }
}

template <typename TO>
hipError_t launch_rotate(const RotBlk *rb, int n_rb, int nts_ns, const cplx<double> *src, const int64_t *src_ts_stride,
                         const int64_t *src_slice, const cplx<double> *rep, cplx<TO> *dst, const int64_t *dst_slice,
                         const int64_t *dst_slice_size, hipStream_t st) {
  const int64_t n_work = (int64_t)n_rb * nts_ns;
  if (n_work == 0) return hipSuccess;
  (void)hipFuncSetAttribute((const void *)rotate_mfma_kernel<TO>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            kRotLds);
  hipLaunchKernelGGL((rotate_mfma_kernel<TO>), dim3((unsigned)xcd_grid(n_work)), dim3(Cfg<double>::NT), kRotLds, st,
                     rb, n_rb, nts_ns, src, src_ts_stride, src_slice, rep, dst, dst_slice, dst_slice_size);
  return hipGetLastError();
}

template hipError_t launch_rotate<double>(const RotBlk *, int, int, const cplx<double> *, const int64_t *,
                                          const int64_t *, const cplx<double> *, cplx<double> *, const int64_t *,
                                          const int64_t *, hipStream_t);
template hipError_t launch_rotate<float>(const RotBlk *, int, int, const cplx<double> *, const int64_t *,
                                         const int64_t *, const cplx<double> *, cplx<float> *, const int64_t *,
                                         const int64_t *, hipStream_t);

template <typename T, int MODE>
hipError_t launch_p0_k(const P0Item *items, const P0Blk *blks, int n_items, int nts, int ns, const cplx<T> *V,
                       const cplx<T> *G, cplx<double> *P0, cplx<double> *part, double pref, size_t lds,
                       hipStream_t st) {
  const int nt_half = nts / 2;
  constexpr int TB = p0_times((int)sizeof(cplx<T>), MODE);
  const int64_t n_work = (int64_t)n_items * ((nt_half + TB - 1) / TB);
  if (n_work == 0) return hipSuccess;
  (void)hipFuncSetAttribute((const void *)p0_mfma_kernel<T, MODE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
  hipLaunchKernelGGL((p0_mfma_kernel<T, MODE>), dim3((unsigned)xcd_grid(n_work)), dim3(Cfg<T>::NT), lds, st, items,
                     blks, n_work, nt_half, nts, ns, V, G, P0, part, pref);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_p0(const P0Item *items, const P0Blk *blks, int n_items, int nts, int ns, const cplx<T> *V,
                     const cplx<T> *G, cplx<double> *P0, cplx<double> *part, double pref, size_t lds, hipStream_t st,
                     int mode) {
  if (mode == 1) return launch_p0_k<T, 1>(items, blks, n_items, nts, ns, V, G, P0, part, pref, lds, st);
  if (mode == 0) return launch_p0_k<T, 0>(items, blks, n_items, nts, ns, V, G, P0, part, pref, lds, st);
  return launch_p0_k<T, 2>(items, blks, n_items, nts, ns, V, G, P0, part, pref, lds, st);
}

template <typename T, int TB, bool Z8>
hipError_t launch_sigma_tb(const SigItem *items, const SigBlk *blks, int n_items, int nts, int ns, const cplx<T> *V,
                           const cplx<T> *G, const cplx<T> *Pf, cplx<double> *Sig, cplx<double> *part, double pref,
                           size_t lds, hipStream_t st, const Z8Ops &z8) {
  const int64_t n_work = (int64_t)n_items * ((nts + TB - 1) / TB) * ns;
  if (n_work == 0) return hipSuccess;
  (void)hipFuncSetAttribute((const void *)sigma_mfma_kernel<T, TB, Z8>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
  hipLaunchKernelGGL((sigma_mfma_kernel<T, TB, Z8>), dim3((unsigned)xcd_grid(n_work)), dim3(SigCfg<T, Z8>::NT), lds, st,
                     items, blks, n_work, nts, ns, V, G, Pf, Sig, part, pref, z8);
  return hipGetLastError();
}

template <typename T, bool Z8>
hipError_t launch_sigma_z(const SigItem *items, const SigBlk *blks, int n_items, int nts, int ns, const cplx<T> *V,
                          const cplx<T> *G, const cplx<T> *Pf, cplx<double> *Sig, cplx<double> *part, double pref,
                          size_t lds, hipStream_t st, int times, const Z8Ops &z8) {
  if (times == 4)
    return launch_sigma_tb<T, 4, Z8>(items, blks, n_items, nts, ns, V, G, Pf, Sig, part, pref, lds, st, z8);
  if (times == 2)
    return launch_sigma_tb<T, 2, Z8>(items, blks, n_items, nts, ns, V, G, Pf, Sig, part, pref, lds, st, z8);
  if (times == 1)
    return launch_sigma_tb<T, 1, Z8>(items, blks, n_items, nts, ns, V, G, Pf, Sig, part, pref, lds, st, z8);
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_sigma(const SigItem *items, const SigBlk *blks, int n_items, int nts, int ns, const cplx<T> *V,
                        const cplx<T> *G, const cplx<T> *Pf, cplx<double> *Sig, cplx<double> *part, double pref,
                        size_t lds, hipStream_t st, int times, const Z8Ops *z8) {
  if constexpr (sizeof(T) == 8) {
    if (z8)
      return launch_sigma_z<T, true>(items, blks, n_items, nts, ns, V, G, Pf, Sig, part, pref, lds, st, times, *z8);
  } else {
    if (z8) return hipErrorInvalidValue;  // the INT8 Z product reads FP64 P and V
  }
  return launch_sigma_z<T, false>(items, blks, n_items, nts, ns, V, G, Pf, Sig, part, pref, lds, st, times, Z8Ops{});
}

template hipError_t launch_p0<double>(const P0Item *, const P0Blk *, int, int, int, const cplx<double> *,
                                      const cplx<double> *, cplx<double> *, cplx<double> *, double, size_t,
                                      hipStream_t, int);
template hipError_t launch_p0<float>(const P0Item *, const P0Blk *, int, int, int, const cplx<float> *,
                                     const cplx<float> *, cplx<double> *, cplx<double> *, double, size_t, hipStream_t,
                                     int);
template hipError_t launch_sigma<double>(const SigItem *, const SigBlk *, int, int, int, const cplx<double> *,
                                         const cplx<double> *, const cplx<double> *, cplx<double> *, cplx<double> *,
                                         double, size_t, hipStream_t, int, const Z8Ops *);
template hipError_t launch_sigma<float>(const SigItem *, const SigBlk *, int, int, int, const cplx<float> *,
                                        const cplx<float> *, const cplx<float> *, cplx<double> *, cplx<double> *,
                                        double, size_t, hipStream_t, int, const Z8Ops *);

}  // namespace symgw
