// Shared device/host helpers: complex arithmetic, XCD-aware block mapping and
// the counter-based generator used for synthetic interaction tensors.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace symgw {

template <typename T>
struct cplx {
  T re, im;
};

template <typename T>
__host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return cplx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * conj(b)
template <typename T>
__host__ __device__ inline cplx<T> cmulc(cplx<T> a, cplx<T> b) {
  return cplx<T>{a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <typename T>
__host__ __device__ inline void cfma(cplx<T> &acc, cplx<T> a, cplx<T> b) {
  acc.re += a.re * b.re - a.im * b.im;
  acc.im += a.re * b.im + a.im * b.re;
}
// acc += a * conj(b)
template <typename T>
__host__ __device__ inline void cfmac(cplx<T> &acc, cplx<T> a, cplx<T> b) {
  acc.re += a.re * b.re + a.im * b.im;
  acc.im += a.im * b.re - a.re * b.im;
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X.  Map the
// hardware block id so that consecutive logical work items run on the same
// XCD (and share its L2): logical = xcd * per_xcd + slot.  Launch with
// gridDim.x a multiple of 8; logical ids >= n must return.
__device__ inline int64_t xcd_logical(int64_t bid, int64_t grid) {
  const int64_t per = grid / 8;
  return (bid % 8) * per + bid / 8;
}
inline int64_t xcd_grid(int64_t n) { return (n + 7) / 8 * 8; }

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ----------------------
struct u32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline void mulhilo32(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

__host__ __device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(0xD2511F53u, c.x, hi0, lo0);
    mulhilo32(0xCD9E8D57u, c.z, hi1, lo1);
    u32x4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Unit-variance Irwin-Hall(4) deviate from four 16-bit uniforms packed in two
// 32-bit words.  Only exact additions and one rounded multiply: identical on
// host and device.
__host__ __device__ inline double ih4(uint32_t a, uint32_t b) {
  const double s = 1.0 / 65536.0;
  double u0 = (double)(a & 0xFFFFu) * s, u1 = (double)(a >> 16) * s;
  double u2 = (double)(b & 0xFFFFu) * s, u3 = (double)(b >> 16) * s;
  double sum = ((u0 + u1) + (u2 + u3)) + 4.0 * (0.5 / 65536.0);
  return (sum - 2.0) * 1.7320508075688772;  // sqrt(12/4)
}

// One complex V element: counter = (element lo, element hi, pair, tag).
__host__ __device__ inline void gen_v_elem(uint64_t seed, uint32_t pair, uint64_t elem, double factor,
                                           double &re, double &im) {
  u32x4 c{(uint32_t)elem, (uint32_t)(elem >> 32), pair, 0x53594D47u};
  u32x4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  double g1 = ih4(r.x, r.y), g2 = ih4(r.z, r.w);
  re = g1 * factor;
  im = g2 * factor;
}

}  // namespace symgw
