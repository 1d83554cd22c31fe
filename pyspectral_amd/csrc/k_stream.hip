// Streaming bandwidth probes (no counterpart in the reference): what a plain HBM-bound
// kernel reaches on this part, the practical ceiling next to the published peak.  Each
// lane moves 16-byte values, four loads in flight before its stores, over a grid of a
// few workgroups per CU that strides through the array.
#include <hip/hip_runtime.h>

#include "../../include/psx.h"

namespace {

constexpr int kThreads = 256;
constexpr int kUnroll = 4;

__global__ __launch_bounds__(kThreads) void stream_copy(const double2* __restrict__ src, double2* __restrict__ dst,
                                                        long long n) {
  const long long stride = (long long)gridDim.x * kThreads;
  long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
    double2 v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) v[u] = src[i + u * stride];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) dst[i + u * stride] = v[u];
  }
  for (; i < n; i += stride) dst[i] = src[i];
}

// block-contiguous: workgroup b copies chunks of kThreads * kUnroll values at
// (b + j gridDim) * chunk; NT = 1 stores with the non-temporal hint
template <int NTS>
__global__ __launch_bounds__(kThreads) void stream_copy_chunk(const double2* __restrict__ src,
                                                              double2* __restrict__ dst, long long n) {
  constexpr int chunk = kThreads * kUnroll;
  for (long long base = (long long)blockIdx.x * chunk; base < n; base += (long long)gridDim.x * chunk) {
    double2 v[kUnroll];
    const long long i0 = base + threadIdx.x;
    if (base + chunk <= n) {
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) v[u] = src[i0 + u * kThreads];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        if constexpr (NTS) {
          __builtin_nontemporal_store(v[u].x, &dst[i0 + u * kThreads].x);
          __builtin_nontemporal_store(v[u].y, &dst[i0 + u * kThreads].y);
        } else {
          dst[i0 + u * kThreads] = v[u];
        }
      }
    } else {
      for (int u = 0; u < kUnroll; ++u)
        if (i0 + u * kThreads < n) dst[i0 + u * kThreads] = src[i0 + u * kThreads];
    }
  }
}

// read-only: per-block sums of the real parts into part[blockIdx.x] (so that the loads
// are used)
__global__ __launch_bounds__(kThreads) void stream_read(const double2* __restrict__ src, double* __restrict__ part,
                                                        long long n) {
  __shared__ double sh[kThreads];
  const long long stride = (long long)gridDim.x * kThreads;
  long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  double acc = 0.0;
  for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
    double2 v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) v[u] = src[i + u * stride];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) acc += v[u].x + v[u].y;
  }
  for (; i < n; i += stride) acc += src[i].x + src[i].y;
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

}  // namespace

extern "C" int psx_stream_probe(const double* src, double* dst, long long n, int mode, int blocks, void* stream) {
  if (!src || !dst || n < 1 || blocks < 1 || mode < 0 || mode > 3) return PSX_E_ARG;
  const hipStream_t s = (hipStream_t)stream;
  if (mode == 0) {
    hipLaunchKernelGGL(stream_copy, dim3(blocks), dim3(kThreads), 0, s, (const double2*)src, (double2*)dst, n);
  } else if (mode == 2) {
    hipLaunchKernelGGL(stream_copy_chunk<0>, dim3(blocks), dim3(kThreads), 0, s, (const double2*)src, (double2*)dst, n);
  } else if (mode == 3) {
    hipLaunchKernelGGL(stream_copy_chunk<1>, dim3(blocks), dim3(kThreads), 0, s, (const double2*)src, (double2*)dst, n);
  } else {
    hipLaunchKernelGGL(stream_read, dim3(blocks), dim3(kThreads), 0, s, (const double2*)src, dst, n);
  }
  return hipGetLastError() == hipSuccess ? PSX_OK : PSX_E_HIP;
}
