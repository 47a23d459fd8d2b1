// Streaming-bandwidth probe: the HBM rate plain kernels reach on this
// device, measured beside the stage kernels so that their fraction of the
// bandwidth roofline can be read against the vendor peak and against what a
// memory-only kernel achieves.  Three access patterns, all 16-byte vector
// accesses with the non-temporal hint, many waves in flight per CU:
//   kind 0  copy:          dst[i] = src[i]
//   kind 1  5 reads + 1 write (the fused advection stage's stream count):
//                          dst[i] = s0[i] + s1[i] + s2[i] + s3[i] + s4[i],
//                          s_q = src + q * n
//   kind 2  reads only:    per-thread sums of src, one float per thread to dst
#include <string>

#include "../../include/lsm3d_amd.h"
#include "lsm_runtime.h"

namespace lsm {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kProbeThreads = 256;
constexpr int kProbeUnroll = 4;

__device__ __forceinline__ v4f ld(const v4f* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st(v4f v, v4f* p) { __builtin_nontemporal_store(v, p); }

__global__ void __launch_bounds__(kProbeThreads) k_probe_copy(const v4f* __restrict__ s, v4f* __restrict__ d,
                                                             long long n) {
  const long long stride = (long long)gridDim.x * kProbeThreads;
  long long i = (long long)blockIdx.x * kProbeThreads + threadIdx.x;
  for (; i + (kProbeUnroll - 1) * stride < n; i += kProbeUnroll * stride) {
    v4f v[kProbeUnroll];
#pragma unroll
    for (int u = 0; u < kProbeUnroll; u++) v[u] = ld(s + i + u * stride);
#pragma unroll
    for (int u = 0; u < kProbeUnroll; u++) st(v[u], d + i + u * stride);
  }
  for (; i < n; i += stride) st(ld(s + i), d + i);
}

__global__ void __launch_bounds__(kProbeThreads) k_probe_5r1w(const v4f* __restrict__ s, v4f* __restrict__ d,
                                                             long long n) {
  const long long stride = (long long)gridDim.x * kProbeThreads;
  for (long long i = (long long)blockIdx.x * kProbeThreads + threadIdx.x; i < n; i += stride) {
    v4f a = ld(s + i), b = ld(s + n + i), c = ld(s + 2 * n + i), e = ld(s + 3 * n + i), f = ld(s + 4 * n + i);
    st(((a + b) + (c + e)) + f, d + i);
  }
}

__global__ void __launch_bounds__(kProbeThreads) k_probe_read(const v4f* __restrict__ s, float* __restrict__ d,
                                                             long long n) {
  const long long stride = (long long)gridDim.x * kProbeThreads;
  long long i = (long long)blockIdx.x * kProbeThreads + threadIdx.x;
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  for (; i + (kProbeUnroll - 1) * stride < n; i += kProbeUnroll * stride) {
#pragma unroll
    for (int u = 0; u < kProbeUnroll; u++) acc += ld(s + i + u * stride);
  }
  for (; i < n; i += stride) acc += ld(s + i);
  d[(long long)blockIdx.x * kProbeThreads + threadIdx.x] = (acc.x + acc.y) + (acc.z + acc.w);
}

}  // namespace
}  // namespace lsm

using namespace lsm;

extern "C" {

double lsm_amd_bandwidth_probe(int kind, const void* src, void* dst, long long bytes, int workgroups, int reps) {
  const char* fn = "lsm_amd_bandwidth_probe";
  Call c(fn);
  if (kind < 0 || kind > 2) raise_error(std::string(fn) + ": kind is 0, 1 or 2");
  if (bytes <= 0 || bytes % 16 != 0) raise_error(std::string(fn) + ": bytes must be a positive multiple of 16");
  if (workgroups <= 0 || reps <= 0) raise_error(std::string(fn) + ": workgroups and reps must be positive");
  if (!src || !dst) raise_error(std::string(fn) + ": missing array");
  if (!c.ok()) return -1.0;
  if (!(is_device_pointer(src) && is_device_pointer(dst))) {
    raise_error(std::string(fn) + ": src and dst must be device memory");
    return -1.0;
  }
  const long long n = bytes / 16;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  check_hip(hipEventCreate(&e0), "hipEventCreate");
  check_hip(hipEventCreate(&e1), "hipEventCreate");
  if (!c.ok()) return -1.0;
  const v4f* s = static_cast<const v4f*>(src);
  auto launch = [&]() {
    if (kind == 0)
      hipLaunchKernelGGL(k_probe_copy, dim3((unsigned)workgroups), dim3(kProbeThreads), 0, c.stream(), s,
                         static_cast<v4f*>(dst), n);
    else if (kind == 1)
      hipLaunchKernelGGL(k_probe_5r1w, dim3((unsigned)workgroups), dim3(kProbeThreads), 0, c.stream(), s,
                         static_cast<v4f*>(dst), n);
    else
      hipLaunchKernelGGL(k_probe_read, dim3((unsigned)workgroups), dim3(kProbeThreads), 0, c.stream(), s,
                         static_cast<float*>(dst), n);
    c.launched("k_probe");
  };
  launch();  // warm-up
  check_hip(hipEventRecord(e0, c.stream()), "hipEventRecord");
  for (int r = 0; r < reps && c.ok(); r++) launch();
  check_hip(hipEventRecord(e1, c.stream()), "hipEventRecord");
  check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
  float ms = 0.f;
  check_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c.finish();
  return c.ok() ? (double)ms / reps : -1.0;
}

}  // extern "C"
