// Device peak microbenchmarks: the matrix-core rate of the contraction
// kernels' complex 16x16x4 step (four real MFMAs, two independent tiles per
// wave, operands in registers) in FP64 and FP32, and the HBM copy bandwidth.
// They give the measured peaks that the contraction rates are compared with
// beside the datasheet values.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace symgw {
namespace {

constexpr int kPeakThreads = 768;  // 12 waves, three per SIMD (the FP32 contraction workgroup)

template <typename T>
struct PeakOp;
template <>
struct PeakOp<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 op(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};
template <>
struct PeakOp<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 op(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

template <typename T>
__global__ __launch_bounds__(kPeakThreads) void mfma_peak_kernel(T *out, int iters, T s) {
  typedef typename PeakOp<T>::v4 v4;
  const int t = threadIdx.x;
  T ar[2], ai[2], br[2], bi[2];
  v4 re[2], im[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    ar[c] = s * T(t + c);
    ai[c] = s * T(t - c);
    br[c] = s * T(2 * t + c);
    bi[c] = s * T(c - t);
    re[c] = v4{T(0), T(0), T(0), T(0)};
    im[c] = re[c];
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      re[c] = PeakOp<T>::op(ar[c], br[c], re[c]);
      im[c] = PeakOp<T>::op(ar[c], bi[c], im[c]);
      re[c] = PeakOp<T>::op(-ai[c], bi[c], re[c]);
      im[c] = PeakOp<T>::op(ai[c], br[c], im[c]);
    }
  }
  T acc = T(0);
#pragma unroll
  for (int c = 0; c < 2; ++c) acc += re[c][0] + re[c][1] + re[c][2] + re[c][3] + im[c][0] + im[c][1] + im[c][2] + im[c][3];
  out[(int64_t)blockIdx.x * kPeakThreads + t] = acc;
}

// dst[i] = src[i] over n 16-byte elements: workgroup b copies the contiguous
// range [b * per, (b + 1) * per), four loads of a thread in flight before its
// stores (the fastest of the copy shapes measured on MI355X).
__global__ __launch_bounds__(256) void hbm_copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                       int64_t n, int64_t per) {
  const int64_t b = (int64_t)blockIdx.x * per, e = b + per < n ? b + per : n;
  for (int64_t i = b + threadIdx.x; i < e; i += 4 * 256) {
    uint4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i + u * 256 < e) v[u] = src[i + u * 256];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i + u * 256 < e) dst[i + u * 256] = v[u];
  }
}

}  // namespace

template <typename T>
hipError_t launch_mfma_peak(void *out, int n_wg, int iters, hipStream_t st) {
  hipLaunchKernelGGL(mfma_peak_kernel<T>, dim3((unsigned)n_wg), dim3(kPeakThreads), 0, st, static_cast<T *>(out), iters,
                     T(1e-3));
  return hipGetLastError();
}
template hipError_t launch_mfma_peak<double>(void *, int, int, hipStream_t);
template hipError_t launch_mfma_peak<float>(void *, int, int, hipStream_t);

int mfma_peak_threads() { return kPeakThreads; }

hipError_t launch_hbm_copy(const void *src, void *dst, int64_t bytes, int n_wg, hipStream_t st) {
  const int64_t n = bytes / 16;
  hipLaunchKernelGGL(hbm_copy_kernel, dim3((unsigned)n_wg), dim3(256), 0, st, static_cast<const uint4 *>(src),
                     static_cast<uint4 *>(dst), n, (n + n_wg - 1) / n_wg);
  return hipGetLastError();
}

}  // namespace symgw
