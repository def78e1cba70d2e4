// P0 and Sigma block contractions on the CDNA4 matrix cores.
//
// Every workgroup owns one output slab and walks its V-block list in
// a fixed order ("owner computes"), so results are bitwise reproducible with
// no atomics or locks (the reference serialises per-q / per-k AXPYs behind
// spin locks, cugw_qkpt_block.cu:269-285 and :414-436).  All complex products
// are 16x16 tiles built from four real MFMAs (v_mfma_f64_16x16x4f64 for the
// plain path, v_mfma_f32_16x16x4f32 for the FP32 path); operands are staged in
// LDS, intermediates never leave the workgroup.  A wave works on up to four
// tiles at a time that share one operand fragment per k-step.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

namespace symgw {

namespace {

// Waves per workgroup: FP64 runs 8 waves on the full 160 KiB LDS of a CU,
// FP32 runs 4 waves on 80 KiB (two workgroups per CU).  Either way two waves
// share each SIMD.
template <typename T>
struct Cfg {
  static constexpr int NW = sizeof(T) == 8 ? 8 : 4;
  static constexpr int NT = 64 * NW;
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

template <typename T>
struct Tile {
  typename Mfma<T>::v4 re, im;
};
// Tiles a wave advances together, sharing one operand fragment per k-step:
// 2 for FP64 and 4 for FP32 (two waves per SIMD within the register file).
template <typename T>
struct Group { static constexpr int n = sizeof(T) == 8 ? 2 : 4; };
template <typename T>
using Tiles = Tile<T>[4];

template <typename T>
__device__ __forceinline__ void tile_zero(Tile<T> &t) {
  t.re = typename Mfma<T>::v4{T(0), T(0), T(0), T(0)};
  t.im = t.re;
}

template <typename T>
__device__ __forceinline__ cplx<T> cz() { return cplx<T>{T(0), T(0)}; }

// acc += a * b for one 16x16x4 complex step (four real MFMAs).  Lane l holds
// A[l%16][k0+l/16] and B[k0+l/16][l%16] (operand layout of the 16x16x4 MFMAs).
template <typename T>
__device__ __forceinline__ void mma4(Tile<T> &acc, const cplx<T> &a, const cplx<T> &b) {
  acc.re = Mfma<T>::op(a.re, b.re, acc.re);
  acc.im = Mfma<T>::op(a.re, b.im, acc.im);
  acc.re = Mfma<T>::op(-a.im, b.im, acc.re);
  acc.im = Mfma<T>::op(a.im, b.re, acc.im);
}

// The loaders below return the operand element or zero outside the valid
// range, so the one-step look-ahead never reads out of bounds.

// acc[C0+c] += A (16 x K) * B_c (K x 16) for c < NG with C0 + c < n, over k in
// [kb, ke): one A fragment per k-step shared by the tiles.  la(i, k), lb(c, k, j)
// with c the absolute tile index.
template <int C0, int NG, typename T, typename LA, typename LB>
__device__ __forceinline__ void mma_rowshare(Tiles<T> &acc, int n, int kb, int ke, const LA &la, const LB &lb) {
  if (kb >= ke || C0 >= n) return;
  const int lane = threadIdx.x & 63, i = lane & 15, kk = lane >> 4;
  cplx<T> a0 = la(i, kb + kk), b0[NG];
#pragma unroll
  for (int c = 0; c < NG; ++c) b0[c] = C0 + c < n ? lb(C0 + c, kb + kk, i) : cz<T>();
  for (int k0 = kb; k0 < ke; k0 += 4) {
    const cplx<T> a1 = la(i, k0 + 4 + kk);
    cplx<T> b1[NG];
#pragma unroll
    for (int c = 0; c < NG; ++c) b1[c] = C0 + c < n ? lb(C0 + c, k0 + 4 + kk, i) : cz<T>();
#pragma unroll
    for (int c = 0; c < NG; ++c)
      if (C0 + c < n) mma4(acc[C0 + c], a0, b0[c]);
    a0 = a1;
#pragma unroll
    for (int c = 0; c < NG; ++c) b0[c] = b1[c];
  }
}

// acc[C0+c] += A_c (16 x K) * B (K x 16): one B fragment shared.  la(c, i, k), lb(k, j).
template <int C0, int NG, typename T, typename LA, typename LB>
__device__ __forceinline__ void mma_colshare(Tiles<T> &acc, int n, int kb, int ke, const LA &la, const LB &lb) {
  if (kb >= ke || C0 >= n) return;
  const int lane = threadIdx.x & 63, i = lane & 15, kk = lane >> 4;
  cplx<T> b0 = lb(kb + kk, i), a0[NG];
#pragma unroll
  for (int c = 0; c < NG; ++c) a0[c] = C0 + c < n ? la(C0 + c, i, kb + kk) : cz<T>();
  for (int k0 = kb; k0 < ke; k0 += 4) {
    const cplx<T> b1 = lb(k0 + 4 + kk, i);
    cplx<T> a1[NG];
#pragma unroll
    for (int c = 0; c < NG; ++c) a1[c] = C0 + c < n ? la(C0 + c, i, k0 + 4 + kk) : cz<T>();
#pragma unroll
    for (int c = 0; c < NG; ++c)
      if (C0 + c < n) mma4(acc[C0 + c], a0[c], b0);
    b0 = b1;
#pragma unroll
    for (int c = 0; c < NG; ++c) a0[c] = a1[c];
  }
}

// acc[C0+c] += A_c * B_c with independent operands.  la(c, i, k), lb(c, k, j).
template <int C0, int NG, typename T, typename LA, typename LB>
__device__ __forceinline__ void mma_multi(Tiles<T> &acc, int n, int kb, int ke, const LA &la, const LB &lb) {
  if (kb >= ke || C0 >= n) return;
  const int lane = threadIdx.x & 63, i = lane & 15, kk = lane >> 4;
  cplx<T> a0[NG], b0[NG];
#pragma unroll
  for (int c = 0; c < NG; ++c) {
    a0[c] = C0 + c < n ? la(C0 + c, i, kb + kk) : cz<T>();
    b0[c] = C0 + c < n ? lb(C0 + c, kb + kk, i) : cz<T>();
  }
  for (int k0 = kb; k0 < ke; k0 += 4) {
    cplx<T> a1[NG], b1[NG];
#pragma unroll
    for (int c = 0; c < NG; ++c) {
      a1[c] = C0 + c < n ? la(C0 + c, i, k0 + 4 + kk) : cz<T>();
      b1[c] = C0 + c < n ? lb(C0 + c, k0 + 4 + kk, i) : cz<T>();
    }
#pragma unroll
    for (int c = 0; c < NG; ++c)
      if (C0 + c < n) mma4(acc[C0 + c], a0[c], b0[c]);
#pragma unroll
    for (int c = 0; c < NG; ++c) {
      a0[c] = a1[c];
      b0[c] = b1[c];
    }
  }
}

// All four persistent tiles, in groups of Group<T>::n.
template <typename T, typename LA, typename LB>
__device__ __forceinline__ void mma_rowshare4(Tiles<T> &acc, int n, int kb, int ke, const LA &la, const LB &lb) {
  constexpr int NG = Group<T>::n;
  mma_rowshare<0, NG>(acc, n, kb, ke, la, lb);
  if (NG == 2) mma_rowshare<2, 2>(acc, n, kb, ke, la, lb);
}
template <typename T, typename LA, typename LB>
__device__ __forceinline__ void mma_multi4(Tiles<T> &acc, int n, int kb, int ke, const LA &la, const LB &lb) {
  constexpr int NG = Group<T>::n;
  mma_multi<0, NG>(acc, n, kb, ke, la, lb);
  if (NG == 2) mma_multi<2, 2>(acc, n, kb, ke, la, lb);
}

// Split-K share of wave w out of NW over [0, K) (boundaries multiples of 4).
template <int NW>
__device__ __forceinline__ void wave_krange(int K, int wave, int &kb, int &ke) {
  const int q4 = ((K + 4 * NW - 1) / (4 * NW)) * 4;
  kb = min(K, wave * q4);
  ke = min(K, kb + q4);
}

// Contiguous share [b, e) of n items for wave w out of NW.
template <int NW>
__device__ __forceinline__ void wave_share(int n, int wave, int &b, int &e) {
  const int per = (n + NW - 1) / NW;
  b = min(n, wave * per);
  e = min(n, b + per);
}

// Sum the NW waves' copies of ntile accumulator tiles in wave order (fixed
// order: bitwise reproducible) and leave the totals in wave 0.  Uses the LDS at
// `red` (NW * 4 * 8 * 64 elements of T); callers sync before reusing it.
template <typename T>
__device__ __forceinline__ void wave_reduce(Tiles<T> &acc, int ntile, T *red) {
  constexpr int NW = Cfg<T>::NW;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 4; ++c)
    if (c < ntile)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        red[((wave * 4 + c) * 8 + r) * 64 + lane] = acc[c].re[r];
        red[((wave * 4 + c) * 8 + 4 + r) * 64 + lane] = acc[c].im[r];
      }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (c < ntile)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          T sre = red[((0 * 4 + c) * 8 + r) * 64 + lane], sim = red[((0 * 4 + c) * 8 + 4 + r) * 64 + lane];
          for (int w2 = 1; w2 < NW; ++w2) {
            sre += red[((w2 * 4 + c) * 8 + r) * 64 + lane];
            sim += red[((w2 * 4 + c) * 8 + 4 + r) * 64 + lane];
          }
          acc[c].re[r] = sre;
          acc[c].im[r] = sim;
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

// C (M x N) = A (M x K) B (K x N) into LDS by the waves, for products with
// at most two column tiles (N <= 32): row tiles are split among the waves and
// processed in groups of four sharing the B fragment.  la(i_row, k), lb(k, j_col),
// st(row, col, value) (called only for row < M, col < N).
template <typename T, typename LA, typename LB, typename ST>
__device__ __forceinline__ void wg_gemm_rows(int M, int N, int K, const LA &la, const LB &lb, const ST &st) {
  const int wave = threadIdx.x >> 6;
  const int ntm = (M + 15) / 16, ntn = (N + 15) / 16;
  int rb, re;
  wave_share<Cfg<T>::NW>(ntm, wave, rb, re);
  for (int nt = 0; nt < ntn; ++nt) {
    const int n0 = nt * 16;
    constexpr int NG = Group<T>::n;
    for (int g = rb; g < re; g += NG) {
      const int n = min(NG, re - g);
      Tiles<T> acc;
#pragma unroll
      for (int c = 0; c < NG; ++c) tile_zero(acc[c]);
      mma_colshare<0, NG>(acc, n, 0, K, [&](int c, int i, int k) { return la((g + c) * 16 + i, k); },
                          [&](int k, int j) { return lb(k, n0 + j); });
#pragma unroll
      for (int c = 0; c < NG; ++c)
        if (c < n)
          tile_store(acc[c], [&](int i, int j, cplx<T> v) {
            const int row = (g + c) * 16 + i, col = n0 + j;
            if (row < M && col < N) st(row, col, v);
          });
    }
  }
}

// Same for products with few row tiles and many column tiles: column tiles
// split among the waves, groups of four sharing the A fragment.
template <typename T, typename LA, typename LB, typename ST>
__device__ __forceinline__ void wg_gemm_cols(int M, int N, int K, const LA &la, const LB &lb, const ST &st) {
  const int wave = threadIdx.x >> 6;
  const int ntm = (M + 15) / 16, ntn = (N + 15) / 16;
  int cb, ce;
  wave_share<Cfg<T>::NW>(ntn, wave, cb, ce);
  for (int mt = 0; mt < ntm; ++mt) {
    const int m0 = mt * 16;
    constexpr int NG = Group<T>::n;
    for (int g = cb; g < ce; g += NG) {
      const int n = min(NG, ce - g);
      Tiles<T> acc;
#pragma unroll
      for (int c = 0; c < NG; ++c) tile_zero(acc[c]);
      mma_rowshare<0, NG>(acc, n, 0, K, [&](int i, int k) { return la(m0 + i, k); },
                          [&](int c, int k, int j) { return lb(k, (g + c) * 16 + j); });
#pragma unroll
      for (int c = 0; c < NG; ++c)
        if (c < n)
          tile_store(acc[c], [&](int i, int j, cplx<T> v) {
            const int row = m0 + i, col = (g + c) * 16 + j;
            if (row < M && col < N) st(row, col, v);
          });
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// P0.  Slab = rows [Q0, Q0 + nband) (nband <= 16) of aux block iq of iBZ point
// q, time t < nts/2.  For every V block (k, ik1, ik2) and spin:
//   E_Q = V_Q G^{k+q}(t),  C_Q = G^k(beta - t) E_Q,  P0[Q][P] += sum_x C_Q[x] conj(V_P[x])
// i.e. the three GEMMs of cugw_qkpt_block.cu:231-260 (X1 = G V, X2 = conj(V) G^T,
// P0^T += X2 X1) regrouped around one intermediate per Q row, processed in row
// chunks of qsub rows sized for the LDS.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(Cfg<T>::NT) void p0_mfma_kernel(const P0Item *__restrict__ items, const P0Blk *__restrict__ blks,
                                                    int64_t n_work, int nt_half, int nts, int ns,
                                                    const cplx<T> *__restrict__ V, const cplx<T> *__restrict__ G,
                                                    cplx<double> *__restrict__ P0, double pref) {
  const int64_t w = xcd_logical(blockIdx.x, gridDim.x);
  if (w >= n_work) return;
  const int t = (int)(w % nt_half);
  const P0Item it = items[w / nt_half];
  constexpr int NW = Cfg<T>::NW, NT = Cfg<T>::NT;
  const int tid = threadIdx.x, wave = tid >> 6;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw);
  const int q = it.q, nband = it.nband, nct = (q + 15) / 16;
  // <= 4 column tiles: every wave takes all of them over 1/NW of x
  // (split-K); otherwise wave w owns column tiles w, w+NW, ...
  const bool split = nct <= 4;
  const int nmine = split ? nct : (nct - wave + NW - 1) / NW;
  Tiles<T> acc;
#pragma unroll
  for (int c = 0; c < 4; ++c) tile_zero(acc[c]);

  for (int b = it.blk_begin; b < it.blk_end; ++b) {
    const P0Blk B = blks[b];
    const int k1 = B.k1, k2 = B.k2, K = k1 * k2, ldk = B.ldk, qsub = B.qsub;
    const cplx<T> *Vb = V + B.v_off;
    cplx<T> *bufV = lds;               // V rows, then C
    cplx<T> *bufE = lds + qsub * ldk;  // E = V G2
    for (int s = 0; s < ns; ++s) {
      const cplx<T> *G1 = G + B.g1_base + (int64_t)(nts - 1 - t) * B.g1_tstride + (int64_t)s * B.g1_sstride;
      const cplx<T> *G2 = G + B.g2_base + (int64_t)t * B.g2_tstride + (int64_t)s * B.g2_sstride;
      for (int r0 = 0; r0 < nband; r0 += qsub) {
        const int nr = min(qsub, nband - r0);
        const cplx<T> *Vr = Vb + (int64_t)(it.Q0 + r0) * K;
        __syncthreads();
        for (int e = tid; e < nr * K; e += NT) {
          const int r = e / K, x = e - r * K;
          bufV[r * ldk + x] = Vr[e];
        }
        __syncthreads();
        // E[(r,a)][n] = sum_m V[r][a][m] G2[m][n]
        wg_gemm_rows<T>(
            nr * k1, k2, k2,
            [&](int row, int m) {
              if (row >= nr * k1 || m >= k2) return cz<T>();
              const int r = row / k1, a = row - r * k1;
              return bufV[r * ldk + a * k2 + m];
            },
            [&](int m, int n) { return (m < k2 && n < k2) ? G2[m * k2 + n] : cz<T>(); },
            [&](int row, int n, cplx<T> v) {
              const int r = row / k1, a = row - r * k1;
              bufE[r * ldk + a * k2 + n] = v;
            });
        __syncthreads();
        // C[r][p][n] = sum_a G1[p][a] E[r][a][n]  (into bufV)
        wg_gemm_cols<T>(
            k1, nr * k2, k1, [&](int p, int a) { return (p < k1 && a < k1) ? G1[p * k1 + a] : cz<T>(); },
            [&](int a, int col) {
              if (a >= k1 || col >= nr * k2) return cz<T>();
              const int r = col / k2, n = col - r * k2;
              return bufE[r * ldk + a * k2 + n];
            },
            [&](int p, int col, cplx<T> v) {
              const int r = col / k2, n = col - r * k2;
              bufV[r * ldk + p * k2 + n] = v;
            });
        __syncthreads();
        // acc[r0 + r][P] += sum_x C[r][x] conj(V[P][x])
        int kb = 0, ke = K;
        if (split) wave_krange<NW>(K, wave, kb, ke);
        mma_rowshare4(
            acc, nmine, kb, ke,
            [&](int i, int x) {
              const int rr = i - r0;
              return (rr >= 0 && rr < nr && x < ke) ? bufV[rr * ldk + x] : cz<T>();
            },
            [&](int c, int x, int j) {
              const int P = (split ? c : wave + NW * c) * 16 + j;
              if (x >= ke || P >= q) return cz<T>();
              const cplx<T> v = Vb[(int64_t)P * K + x];
              return cplx<T>{v.re, -v.im};
            });
      }
    }
  }
  cplx<double> *out = P0 + it.out_base + (int64_t)t * it.out_tstride;
  if (split) wave_reduce(acc, nct, reinterpret_cast<T *>(lds));
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int jt = split ? c : wave + NW * c;
    if (c < nmine && (!split || wave == 0))
      tile_store(acc[c], [&](int i, int j, cplx<T> v) {
        const int P = jt * 16 + j;
        if (i < nband && P < q) out[i * q + P] = cplx<double>{pref * (double)v.re, pref * (double)v.im};
      });
  }
  // rows t >= nts/2 of the reference's device P0 stay zero (cugw_routines_block.cu:79)
  cplx<double> *mir = P0 + it.out_base + (int64_t)(nts - 1 - t) * it.out_tstride;
  for (int e = tid; e < nband * q; e += NT) mir[e] = cplx<double>{0.0, 0.0};
}

// ---------------------------------------------------------------------------
// Sigma.  Slab = rows [I0, I0+Ib) x columns [J0, J0+Jb) of ao block ik1 of iBZ
// point k, time t, spin s.  For every V block (q, iq, ik2), in chunks of qsub
// aux rows Q (E and Z stored with the (Q, n) reduction index contiguous):
//   Z_Q = sum_P conj(P_{QP}(t)) V_P,  E_Q = V_Q G^{k-q}(t),
//   Sigma[i][j] += sum_{Q,n} E_Q[i][n] conj(Z_Q[j][n])
// the contraction of cugw_qkpt_block.cu:381-407 (Y1 = V G, Y2 = Y1 P,
// Sigma += Y2 V^H) with the aux contraction moved onto V so that the largest
// GEMM (q x q x k1 k2) reads V straight from memory.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(Cfg<T>::NT) void sigma_mfma_kernel(const SigItem *__restrict__ items,
                                                       const SigBlk *__restrict__ blks, int64_t n_work, int nts,
                                                       int ns, const cplx<T> *__restrict__ V,
                                                       const cplx<T> *__restrict__ G, const cplx<T> *__restrict__ Pf,
                                                       cplx<double> *__restrict__ Sig, double pref) {
  const int64_t w = xcd_logical(blockIdx.x, gridDim.x);
  if (w >= n_work) return;
  const int per = nts * ns;
  const SigItem it = items[w / per];
  const int ts = (int)(w % per), t = ts / ns, s = ts - (ts / ns) * ns;
  constexpr int NW = Cfg<T>::NW, NT = Cfg<T>::NT;
  const int tid = threadIdx.x, wave = tid >> 6;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw);
  const int k1 = it.k1, I0 = it.I0, Ib = it.Ib, J0 = it.J0, Jb = it.Jb;
  const int nti = (Ib + 15) / 16, ntj = (Jb + 15) / 16, ntile = nti * ntj;
  // <= 4 output tiles: every wave takes all of them over 1/NW of the (Q, n)
  // reduction (split-K); otherwise wave w owns tiles w, w+NW, ...
  const bool split = ntile <= 4;
  const int nmine = split ? ntile : (ntile - wave + NW - 1) / NW;
  Tiles<T> acc;
#pragma unroll
  for (int c = 0; c < 4; ++c) tile_zero(acc[c]);

  for (int b = it.blk_begin; b < it.blk_end; ++b) {
    const SigBlk B = blks[b];
    const int q = B.q, k2 = B.k2, K = k1 * k2, qsub = B.qsub, ldz = B.ldz, lde = B.lde;
    const cplx<T> *Vb = V + B.v_off;
    const cplx<T> *G2 = G + B.g2_base + (int64_t)t * B.g2_tstride + (int64_t)s * B.g2_sstride;
    const cplx<T> *Pb = Pf + B.p_base + (int64_t)t * B.p_tstride;
    // y = r*k2 + n indexes (aux row of the chunk, orbital n of k - q)
    cplx<T> *bufZ = lds;                 // [Jb][ldz]  Z[r][j][n] at j*ldz + y
    cplx<T> *bufE = bufZ + Jb * ldz;     // [Ib][lde]  E[r][i][n] at i*lde + y
    cplx<T> *bufP = bufE + Ib * lde;     // [qsub][q]  conj(P) rows
    const int zc = Jb * k2;
    const cplx<T> *Vj = Vb + J0 * k2;
    for (int Q0 = 0; Q0 < q; Q0 += qsub) {
      const int nr = min(qsub, q - Q0);
      __syncthreads();
      for (int e = tid; e < nr * q; e += NT) {
        const cplx<T> p = Pb[(int64_t)Q0 * q + e];
        bufP[e] = cplx<T>{p.re, -p.im};
      }
      __syncthreads();
      // Z[r][x] = sum_P conj(P[Q0+r][P]) V[P][J0*k2 + x],  x < Jb*k2
      wg_gemm_cols<T>(
          nr, zc, q, [&](int r, int P) { return (r < nr && P < q) ? bufP[r * q + P] : cz<T>(); },
          [&](int P, int x) { return (P < q && x < zc) ? Vj[(int64_t)P * K + x] : cz<T>(); },
          [&](int r, int x, cplx<T> v) {
            const int jj = x / k2, n = x - jj * k2;
            bufZ[jj * ldz + r * k2 + n] = v;
          });
      // E[(r,i)][n] = sum_m V[Q0+r][I0+i][m] G2[m][n]
      wg_gemm_rows<T>(
          nr * Ib, k2, k2,
          [&](int row, int m) {
            if (row >= nr * Ib || m >= k2) return cz<T>();
            const int r = row / Ib, ii = row - r * Ib;
            return Vb[(int64_t)(Q0 + r) * K + (I0 + ii) * k2 + m];
          },
          [&](int m, int n) { return (m < k2 && n < k2) ? G2[m * k2 + n] : cz<T>(); },
          [&](int row, int n, cplx<T> v) {
            const int r = row / Ib, ii = row - r * Ib;
            bufE[ii * lde + r * k2 + n] = v;
          });
      __syncthreads();
      // acc[i][j] += sum_y E[i][y] conj(Z[j][y])
      const int Y = nr * k2;
      int yb = 0, ye = Y;
      if (split) wave_krange<NW>(Y, wave, yb, ye);
      mma_multi4(
          acc, nmine, yb, ye,
          [&](int c, int i, int y) {
            const int i0 = ((split ? c : wave + NW * c) / ntj) * 16;
            return (i0 + i < Ib && y < ye) ? bufE[(i0 + i) * lde + y] : cz<T>();
          },
          [&](int c, int y, int j) {
            const int j0 = ((split ? c : wave + NW * c) % ntj) * 16;
            if (j0 + j >= Jb || y >= ye) return cz<T>();
            const cplx<T> z = bufZ[(j0 + j) * ldz + y];
            return cplx<T>{z.re, -z.im};
          });
    }
  }
  cplx<double> *out = Sig + it.out_base + (int64_t)t * it.out_tstride + (int64_t)s * it.out_sstride;
  if (split) wave_reduce(acc, ntile, reinterpret_cast<T *>(lds));
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int tt = split ? c : wave + NW * c;
    if (c < nmine && (!split || wave == 0)) {
      const int i0 = (tt / ntj) * 16, j0 = (tt % ntj) * 16;
      tile_store(acc[c], [&](int i, int j, cplx<T> v) {
        if (i0 + i < Ib && j0 + j < Jb)
          out[(int64_t)(i0 + i) * k1 + J0 + j0 + j] = cplx<double>{pref * (double)v.re, pref * (double)v.im};
      });
    }
  }
  (void)I0;
}

template <typename T>
hipError_t launch_p0(const P0Item *items, const P0Blk *blks, int n_items, int nts, int ns, const cplx<T> *V,
                     const cplx<T> *G, cplx<double> *P0, double pref, size_t lds, hipStream_t st) {
  const int nt_half = nts / 2;
  const int64_t n_work = (int64_t)n_items * nt_half;
  if (n_work == 0) return hipSuccess;
  (void)hipFuncSetAttribute((const void *)p0_mfma_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((p0_mfma_kernel<T>), dim3((unsigned)xcd_grid(n_work)), dim3(Cfg<T>::NT), lds, st, items, blks,
                     n_work,
                     nt_half, nts, ns, V, G, P0, pref);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_sigma(const SigItem *items, const SigBlk *blks, int n_items, int nts, int ns, const cplx<T> *V,
                        const cplx<T> *G, const cplx<T> *Pf, cplx<double> *Sig, double pref, size_t lds,
                        hipStream_t st) {
  const int64_t n_work = (int64_t)n_items * nts * ns;
  if (n_work == 0) return hipSuccess;
  (void)hipFuncSetAttribute((const void *)sigma_mfma_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
  hipLaunchKernelGGL((sigma_mfma_kernel<T>), dim3((unsigned)xcd_grid(n_work)), dim3(Cfg<T>::NT), lds, st, items, blks,
                     n_work, nts, ns, V, G, Pf, Sig, pref);
  return hipGetLastError();
}

template hipError_t launch_p0<double>(const P0Item *, const P0Blk *, int, int, int, const cplx<double> *,
                                      const cplx<double> *, cplx<double> *, double, size_t, hipStream_t);
template hipError_t launch_p0<float>(const P0Item *, const P0Blk *, int, int, int, const cplx<float> *,
                                     const cplx<float> *, cplx<double> *, double, size_t, hipStream_t);
template hipError_t launch_sigma<double>(const SigItem *, const SigBlk *, int, int, int, const cplx<double> *,
                                         const cplx<double> *, const cplx<double> *, cplx<double> *, double, size_t,
                                         hipStream_t);
template hipError_t launch_sigma<float>(const SigItem *, const SigBlk *, int, int, int, const cplx<float> *,
                                        const cplx<float> *, const cplx<float> *, cplx<double> *, double, size_t,
                                        hipStream_t);

}  // namespace symgw
