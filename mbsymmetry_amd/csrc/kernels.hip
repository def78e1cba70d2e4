// Device kernels of the GW hot path (gfx950).
//
// Rotations, transforms, W / Dyson solvers, density, generator.  The P0 and
// Sigma contractions live in contract.hip.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

namespace symgw {

static constexpr int NT = 256;  // threads per workgroup

template <typename T>
__device__ inline cplx<T> czero() { return cplx<T>{T(0), T(0)}; }

// ---------------------------------------------------------------------------
// Batched transform GEMM (FP64): out_b[r][c] = alpha * sum_j T[r][j] in_b[row(j)][c]
// for r < M, j < KT, c < ncol_b.  Columns are split into nsplit groups of
// xlen_b: column c -> (c / xlen) * split_b + c % xlen, with per-batch split
// offsets on each side.
// row(j) = j + row_shift (mode 0) or the P0(beta - t) = P0(t) mirror of
// t = j + row_shift (mode 1).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void tgemm_kernel(const TBatch *__restrict__ tb, int M, int KT, int nsplit,
                                                  const cplx<double> *__restrict__ Tm,
                                                  const cplx<double> *__restrict__ in, cplx<double> *__restrict__ out,
                                                  int mode, int row_shift, int nts, double alpha,
                                                  int64_t in_split_stride, int64_t out_split_stride) {
  const int b = blockIdx.y;
  const TBatch B = tb[b];
  const int ncol = B.xlen * nsplit;
  const int c = blockIdx.x * NT + threadIdx.x;
  const int r0 = blockIdx.z * 8;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<double> *Ts = reinterpret_cast<cplx<double> *>(smem_raw);  // [8][KT]
  const int nr = min(8, M - r0);
  for (int i = threadIdx.x; i < nr * KT; i += NT) Ts[i] = Tm[(int64_t)r0 * KT + i];
  __syncthreads();
  if (c >= ncol) return;
  const int sp = c / B.xlen, cx = c - sp * B.xlen;
  const int64_t in_col = B.in_off + sp * B.in_split + cx;
  cplx<double> acc[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) acc[r] = cplx<double>{0.0, 0.0};
  for (int j = 0; j < KT; ++j) {
    int row = j + row_shift;
    if (mode == 1 && row >= nts / 2) row = nts - 1 - row;
    const cplx<double> v = in[in_col + (int64_t)row * B.in_row_stride];
#pragma unroll
    for (int r = 0; r < 8; ++r)
      if (r < nr) cfma(acc[r], Ts[r * KT + j], v);
  }
  const int64_t out_col = B.out_off + sp * B.out_split + cx;
#pragma unroll
  for (int r = 0; r < 8; ++r)
    if (r < nr) out[out_col + (int64_t)(r0 + r) * B.out_row_stride] = cplx<double>{alpha * acc[r].re, alpha * acc[r].im};
}

// ---------------------------------------------------------------------------
// In-place Gauss-Jordan inversion with partial pivoting of an n x n complex
// matrix A (row-major, any address space), by one workgroup.
// ---------------------------------------------------------------------------
__device__ void gj_invert(cplx<double> *A, int n, int *piv, cplx<double> *fcol, double *red_v, int *red_i) {
  const int tid = threadIdx.x;
  for (int k = 0; k < n; ++k) {
    // pivot search over rows k..n-1
    double best = -1.0;
    int bi = k;
    for (int i = k + tid; i < n; i += NT) {
      cplx<double> a = A[i * n + k];
      double v = a.re * a.re + a.im * a.im;
      if (v > best) { best = v; bi = i; }
    }
    red_v[tid] = best;
    red_i[tid] = bi;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
      if (tid < st) {
        double v2 = red_v[tid + st];
        int i2 = red_i[tid + st];
        if (v2 > red_v[tid] || (v2 == red_v[tid] && i2 < red_i[tid])) { red_v[tid] = v2; red_i[tid] = i2; }
      }
      __syncthreads();
    }
    const int p = red_i[0];
    if (tid == 0) piv[k] = p;
    if (p != k)
      for (int j = tid; j < n; j += NT) {
        cplx<double> x = A[k * n + j];
        A[k * n + j] = A[p * n + j];
        A[p * n + j] = x;
      }
    __syncthreads();
    const cplx<double> d = A[k * n + k];
    const double dn = d.re * d.re + d.im * d.im;
    const cplx<double> inv{d.re / dn, -d.im / dn};
    __syncthreads();
    for (int j = tid; j < n; j += NT) {
      if (j == k) A[k * n + k] = inv;
      else A[k * n + j] = cmul(A[k * n + j], inv);
    }
    for (int i = tid; i < n; i += NT) fcol[i] = A[i * n + k];
    __syncthreads();
    for (int e = tid; e < n * n; e += NT) {
      const int i = e / n, j = e - i * n;
      if (i == k) continue;
      const cplx<double> f = fcol[i];
      const cplx<double> akj = A[k * n + j];
      if (j == k) {
        A[i * n + k] = cplx<double>{-(f.re * akj.re - f.im * akj.im), -(f.re * akj.im + f.im * akj.re)};
      } else {
        cplx<double> v = A[i * n + j];
        v.re -= f.re * akj.re - f.im * akj.im;
        v.im -= f.re * akj.im + f.im * akj.re;
        A[i * n + j] = v;
      }
    }
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {
    const int p = piv[k];
    if (p != k)
      for (int i = tid; i < n; i += NT) {
        cplx<double> x = A[i * n + k];
        A[i * n + k] = A[i * n + p];
        A[i * n + p] = x;
      }
    __syncthreads();
  }
}

// W solve per (q_ir, n, aux block) (cugw_solver_block.cpp:412-424):
//   A = I - (X + X^H)/2,  A <- A^-1,  A <- (A + A^H)/2,  X <- A X
// Pw layout: [q_ir][n][qslice], block at w_off[q_ir] + n*qslice + off.
// Frequencies [nw_first, nw_first + nw_count); frequency n uses scratch group
// n - nw_first.
__global__ __launch_bounds__(NT) void wsolve_kernel(const SolveBlk *__restrict__ blks, int nblk, int nw_first,
                                                   int nw_count, cplx<double> *__restrict__ Pw,
                                                   const int64_t *__restrict__ w_off,
                                                   const int64_t *__restrict__ qslice,
                                                   cplx<double> *__restrict__ scratch, int64_t scr_group) {
  const int b = blockIdx.x / nw_count, nn = nw_first + blockIdx.x % nw_count;
  if (b >= nblk) return;
  const SolveBlk S = blks[b];
  const int n = S.n, tid = threadIdx.x;
  __shared__ int piv[512];
  __shared__ cplx<double> fcol[512];
  __shared__ double red_v[NT];
  __shared__ int red_i[NT];
  cplx<double> *X = Pw + w_off[S.slice] + (int64_t)nn * qslice[S.slice] + S.off;
  cplx<double> *A = scratch + (int64_t)(nn - nw_first) * scr_group + S.scr;
  cplx<double> *Xc = A + (int64_t)n * n;
  for (int e = tid; e < n * n; e += NT) {
    const int i = e / n, j = e - i * n;
    const cplx<double> x = X[e], xt = X[j * n + i];
    Xc[e] = x;
    A[e] = cplx<double>{(i == j ? 1.0 : 0.0) - 0.5 * (x.re + xt.re), -0.5 * (x.im - xt.im)};
  }
  __syncthreads();
  gj_invert(A, n, piv, fcol, red_v, red_i);
  // hermitize in place: pairs (i<j)
  for (int e = tid; e < n * n; e += NT) {
    const int i = e / n, j = e - i * n;
    if (i > j) continue;
    const cplx<double> a = A[i * n + j], bt = A[j * n + i];
    const cplx<double> h{0.5 * (a.re + bt.re), 0.5 * (a.im - bt.im)};
    A[i * n + j] = h;
    A[j * n + i] = cplx<double>{h.re, -h.im};
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += NT) {
    const int i = e / n, j = e - i * n;
    cplx<double> acc{0.0, 0.0};
    for (int kk = 0; kk < n; ++kk) cfma(acc, A[i * n + kk], Xc[kk * n + j]);
    X[e] = acc;
  }
}

// Dyson per (k_ir, w, s, ao block) (sc_type_block.cpp:159-168):
//   G(iw) = [(iw + mu) S - F - sigma_scale * Sigma(iw)]^-1
// Sw/Gw layout: [k_ir][w][s][kslice] at d_off[k_ir] + (w*ns + s)*kslice + off.
// Sw and Gw may be the same buffer (the Dyson phase solves in place): each
// workgroup reads its Sigma block before the barrier and writes G after it.
__global__ __launch_bounds__(NT) void dyson_kernel(const SolveBlk *__restrict__ blks, int nblk, int nw, int ns,
                                                  const cplx<double> *Sw, cplx<double> *Gw,
                                                  const int64_t *__restrict__ d_off,
                                                  const int64_t *__restrict__ kslice,
                                                  const int64_t *__restrict__ f_off,
                                                  const cplx<double> *__restrict__ F, const cplx<double> *__restrict__ S,
                                                  int64_t flat_ao, const cplx<double> *__restrict__ iwn, double mu,
                                                  double sigma_scale, cplx<double> *__restrict__ scratch,
                                                  int64_t scr_group) {
  const int per = nw * ns;
  const int b = blockIdx.x / per, ws = blockIdx.x % per;
  if (b >= nblk) return;
  const int wi = ws / ns, s = ws % ns;
  const SolveBlk B = blks[b];
  const int n = B.n, tid = threadIdx.x;
  __shared__ int piv[512];
  __shared__ cplx<double> fcol[512];
  __shared__ double red_v[NT];
  __shared__ int red_i[NT];
  const int64_t loc = d_off[B.slice] + (int64_t)ws * kslice[B.slice] + B.off;
  const cplx<double> *Sig = Sw + loc;
  const cplx<double> *Fb = F + (int64_t)s * flat_ao + f_off[B.slice] + B.off;
  const cplx<double> *Sb = S + (int64_t)s * flat_ao + f_off[B.slice] + B.off;
  const cplx<double> z{iwn[wi].re + mu, iwn[wi].im};
  cplx<double> *A = scratch + (int64_t)ws * scr_group + B.scr;
  for (int e = tid; e < n * n; e += NT) {
    const cplx<double> zs = cmul(z, Sb[e]);
    // sigma_scale == 0 means no self-energy: Sigma(iw) is not read at all
    const cplx<double> sg = sigma_scale != 0.0 ? Sig[e] : cplx<double>{0.0, 0.0};
    A[e] = cplx<double>{zs.re - Fb[e].re - sigma_scale * sg.re, zs.im - Fb[e].im - sigma_scale * sg.im};
  }
  __syncthreads();
  gj_invert(A, n, piv, fcol, red_v, red_i);
  cplx<double> *Go = Gw + loc;
  for (int e = tid; e < n * n; e += NT) Go[e] = A[e];
}

// Electron count: sum over (s, k_ir, diagonal blocks) of w_k Tr[G(beta) S].
// One workgroup, fixed-order reduction.
__global__ __launch_bounds__(NT) void density_kernel(const SolveBlk *__restrict__ blks, int nblk, int ns,
                                                    const cplx<double> *__restrict__ Gbeta,
                                                    const cplx<double> *__restrict__ S,
                                                    const int64_t *__restrict__ f_off, int64_t flat_ao,
                                                    const double *__restrict__ kw, double *__restrict__ out) {
  __shared__ double red[NT];
  double part = 0.0;
  for (int s = 0; s < ns; ++s)
    for (int b = 0; b < nblk; ++b) {
      const SolveBlk B = blks[b];
      const int n = B.n;
      const cplx<double> *Gb = Gbeta + (int64_t)s * flat_ao + f_off[B.slice] + B.off;
      const cplx<double> *Sb = S + (int64_t)s * flat_ao + f_off[B.slice] + B.off;
      for (int e = threadIdx.x; e < n * n; e += NT) {
        const int i = e / n, j = e - i * n;
        // Tr[G S] = sum_ij G[i][j] S[j][i] (real part)
        const cplx<double> g = Gb[e], sv = Sb[j * n + i];
        part += kw[B.slice] * (g.re * sv.re - g.im * sv.im);
      }
    }
  red[threadIdx.x] = part;
  __syncthreads();
  for (int st = NT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// Synthetic V on the device: one thread per complex element of each block.
template <typename T>
__global__ __launch_bounds__(NT) void genv_kernel(const GenBlk *__restrict__ gb, int ngb, uint64_t seed, double scale,
                                                 const double *__restrict__ decay, int n_decay, cplx<T> *__restrict__ V) {
  const int b = blockIdx.x;
  if (b >= ngb) return;
  const GenBlk B = gb[b];
  const int64_t K = (int64_t)B.k1 * B.k2;
  const int64_t n = (int64_t)B.q * K;
  for (int64_t e = threadIdx.x; e < n; e += NT) {
    const int64_t rem = e % K;
    const int a = (int)(rem / B.k2), m = (int)(rem % B.k2);
    int d = (B.a_base + a) - (B.b_base + m);
    d = d < 0 ? -d : d;
    d = d < n_decay ? d : n_decay - 1;
    const double f = (scale * decay[d]) * 0.70710678118654752440;
    double re, im;
    gen_v_elem(seed, (uint32_t)B.pair, (uint64_t)(B.local_off + e), f, re, im);
    V[B.v_off + e] = cplx<T>{(T)re, (T)im};
  }
}

// Element-wise conversion (used for FP32 copies of FP64 data).
template <typename TO>
__global__ __launch_bounds__(NT) void convert_kernel(const cplx<double> *__restrict__ in, cplx<TO> *__restrict__ out,
                                                    int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < n; e += (int64_t)gridDim.x * NT)
    out[e] = cplx<TO>{(TO)in[e].re, (TO)in[e].im};
}

// Reorder k_ts_ij (iBZ, contraction output) <-> ts_kij: dir 0: k_ts -> ts_k.
__global__ __launch_bounds__(NT) void reorder_kernel(const cplx<double> *__restrict__ in, cplx<double> *__restrict__ out,
                                                    const int64_t *__restrict__ koff, const int64_t *__restrict__ ksize,
                                                    int ink, int nts_ns, int64_t flat_ao, int dir, double scale) {
  const int k = blockIdx.y;
  const int ts = blockIdx.z;
  if (k >= ink || ts >= nts_ns) return;
  const int64_t sz = ksize[k];
  for (int64_t x = (int64_t)blockIdx.x * NT + threadIdx.x; x < sz; x += (int64_t)gridDim.x * NT) {
    const int64_t a = koff[k] * nts_ns + (int64_t)ts * sz + x;   // k_ts_ij
    const int64_t bidx = (int64_t)ts * flat_ao + koff[k] + x;    // ts_kij
    if (dir == 0) out[bidx] = cplx<double>{scale * in[a].re, scale * in[a].im};
    else out[a] = cplx<double>{scale * in[bidx].re, scale * in[bidx].im};
  }
}

// ---- launchers ---------------------------------------------------------------
hipError_t launch_tgemm(const TBatch *tb, int nbatch, int max_cols, int M, int KT, int nsplit,
                        const cplx<double> *Tm, const cplx<double> *in, cplx<double> *out, int mode, int row_shift,
                        int nts, double alpha, int64_t in_split_stride, int64_t out_split_stride, hipStream_t st) {
  if (nbatch == 0 || M == 0) return hipSuccess;
  dim3 grid((max_cols + NT - 1) / NT, nbatch, (M + 7) / 8);
  hipLaunchKernelGGL(tgemm_kernel, grid, dim3(NT), (size_t)8 * KT * sizeof(cplx<double>), st, tb, M, KT, nsplit, Tm,
                     in, out, mode, row_shift, nts, alpha, in_split_stride, out_split_stride);
  return hipGetLastError();
}

hipError_t launch_wsolve(const SolveBlk *blks, int nblk, int nw_first, int nw_count, cplx<double> *Pw,
                         const int64_t *w_off, const int64_t *qslice, cplx<double> *scratch, int64_t scr_group,
                         hipStream_t st) {
  if (nblk == 0 || nw_count <= 0) return hipSuccess;
  hipLaunchKernelGGL(wsolve_kernel, dim3(nblk * nw_count), dim3(NT), 0, st, blks, nblk, nw_first, nw_count, Pw, w_off,
                     qslice, scratch, scr_group);
  return hipGetLastError();
}

hipError_t launch_dyson(const SolveBlk *blks, int nblk, int nw, int ns, const cplx<double> *Sw, cplx<double> *Gw,
                        const int64_t *d_off, const int64_t *kslice, const int64_t *f_off, const cplx<double> *F,
                        const cplx<double> *S, int64_t flat_ao, const cplx<double> *iwn, double mu,
                        double sigma_scale, cplx<double> *scratch, int64_t scr_group, hipStream_t st) {
  if (nblk == 0 || nw == 0) return hipSuccess;
  hipLaunchKernelGGL(dyson_kernel, dim3(nblk * nw * ns), dim3(NT), 0, st, blks, nblk, nw, ns, Sw, Gw, d_off, kslice,
                     f_off, F, S, flat_ao, iwn, mu, sigma_scale, scratch, scr_group);
  return hipGetLastError();
}

hipError_t launch_density(const SolveBlk *blks, int nblk, int ns, const cplx<double> *Gbeta, const cplx<double> *S,
                          const int64_t *f_off, int64_t flat_ao, const double *kw, double *out, hipStream_t st) {
  hipLaunchKernelGGL(density_kernel, dim3(1), dim3(NT), 0, st, blks, nblk, ns, Gbeta, S, f_off, flat_ao, kw, out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_genv(const GenBlk *gb, int ngb, uint64_t seed, double scale, const double *decay, int n_decay,
                       cplx<T> *V, hipStream_t st) {
  if (ngb == 0) return hipSuccess;
  hipLaunchKernelGGL((genv_kernel<T>), dim3(ngb), dim3(NT), 0, st, gb, ngb, seed, scale, decay, n_decay, V);
  return hipGetLastError();
}

template <typename TO>
hipError_t launch_convert(const cplx<double> *in, cplx<TO> *out, int64_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  int64_t blocks = std::min<int64_t>((n + NT - 1) / NT, 65536);
  hipLaunchKernelGGL((convert_kernel<TO>), dim3((unsigned)blocks), dim3(NT), 0, st, in, out, n);
  return hipGetLastError();
}

hipError_t launch_reorder(const cplx<double> *in, cplx<double> *out, const int64_t *koff, const int64_t *ksize,
                          int ink, int nts_ns, int64_t max_ksize, int64_t flat_ao, int dir, double scale,
                          hipStream_t st) {
  dim3 grid((unsigned)std::min<int64_t>((max_ksize + NT - 1) / NT, 1024), ink, nts_ns);
  hipLaunchKernelGGL(reorder_kernel, grid, dim3(NT), 0, st, in, out, koff, ksize, ink, nts_ns, flat_ao, dir, scale);
  return hipGetLastError();
}

template hipError_t launch_genv<double>(const GenBlk *, int, uint64_t, double, const double *, int, cplx<double> *,
                                        hipStream_t);
template hipError_t launch_genv<float>(const GenBlk *, int, uint64_t, double, const double *, int, cplx<float> *,
                                       hipStream_t);
template hipError_t launch_convert<float>(const cplx<double> *, cplx<float> *, int64_t, hipStream_t);

}  // namespace symgw
