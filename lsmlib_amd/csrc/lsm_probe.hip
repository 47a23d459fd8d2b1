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
//   kind 3  5 reads + 1 write as kind 1 with 4-byte accesses (one float per
//           lane, the width the stage kernels load with)
// lsm_amd_march_probe: the stage kernels' tile march with the arithmetic
// taken out -- 32 x 8 column tiles, z-chunks, the stage launch's tile order,
// 5 fields read and 1 written at every cell, no halo -- the streaming rate the
// access pattern itself allows.
// lsm_amd_valu_probe: the FP32 vector-ALU issue rate -- eight independent
// v_fma_f32 (or v_rcp_f32) chains per lane, no memory traffic in the loop --
// the ceiling the stage kernels' VALU lane-instruction counts are compared
// with.
#include <string>

#include "../../include/lsm3d_amd.h"
#include "lsm_march.h"
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

__global__ void __launch_bounds__(kProbeThreads) k_probe_5r1w_b4(const float* __restrict__ s, float* __restrict__ d,
                                                                long long n) {
  const long long stride = (long long)gridDim.x * kProbeThreads;
  for (long long i = (long long)blockIdx.x * kProbeThreads + threadIdx.x; i < n; i += stride) {
    const float a = __builtin_nontemporal_load(s + i), b = __builtin_nontemporal_load(s + n + i);
    const float c = __builtin_nontemporal_load(s + 2 * n + i), e = __builtin_nontemporal_load(s + 3 * n + i);
    const float f = __builtin_nontemporal_load(s + 4 * n + i);
    __builtin_nontemporal_store(((a + b) + (c + e)) + f, d + i);
  }
}

constexpr int kMarchTX = 32, kMarchTY = 8;

// Fields of (n + 6)^3 cells with 3 ghost cells per side, rows of px floats
// and planes of px * py floats (interior cell (0,0,0) at row offset 32 of
// row 3 of plane 3, the padded layout of the stage workloads).
__global__ void __launch_bounds__(kMarchTX* kMarchTY)
    k_probe_march(const float* __restrict__ s, float* __restrict__ d, long long field, TileGrid tg, int px,
                  long long plane) {
  int bx, by, bz;
  launch_tile_coords(tg, (int)blockIdx.x, (int)gridDim.x, bx, by, bz);
  const int t = (int)threadIdx.x;
  const int i = bx * kMarchTX + t % kMarchTX, j = by * kMarchTY + t / kMarchTX;
  long long o = (long long)(j + 3) * px + (i + 32) + (long long)(bz * tg.kz + 3) * plane;
  const int nz = tg.nkz * tg.kz;  // n is a multiple of kz
  const int k1 = min(tg.kz, nz - bz * tg.kz);
  for (int k = 0; k < k1; k++, o += plane) {
    const float v = (__builtin_nontemporal_load(s + o) + __builtin_nontemporal_load(s + field + o)) +
                    (__builtin_nontemporal_load(s + 2 * field + o) + __builtin_nontemporal_load(s + 3 * field + o)) +
                    __builtin_nontemporal_load(s + 4 * field + o);
    __builtin_nontemporal_store(v, d + o);
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

// Eight independent chains per lane, kValuUnroll x 8 instructions per trip:
// OP 0 v_fma_f32 with scalar b and c (a = a * b + c), OP 1 v_rcp_f32 (a = 1 /
// a, a transcendental instruction), OP 2 v_mul_f32 (a = a * b), OP 3
// v_add_f32 (a = a + c), OP 4 v_fma_f32 with per-lane operands, OP 5 the
// accumulating v_fmac_f32 (a = b' * c' + a, per-lane b', c'), OP 6 the
// three-operand v_max3_f32 (per-lane operands), OP 7 v_fmamk_f32 (literal
// constants).  b and c are kernel arguments, so OP 0, 2 and 3 read scalar
// registers; b', c' are per-lane copies.  One float
// per thread is written so nothing is dead; b and c come from the host, so no
// chain folds.
constexpr int kValuChains = 8, kValuUnroll = 4;

template <int OP>
__global__ void __launch_bounds__(kProbeThreads) k_probe_valu(float* __restrict__ d, int trips, float b, float c) {
  float a[kValuChains];
#pragma unroll
  for (int q = 0; q < kValuChains; q++) a[q] = (float)(threadIdx.x + q) * 1e-3f + 1.f;
  const float bl = b + (float)threadIdx.x * 1e-9f, cl = c - (float)threadIdx.x * 1e-9f;  // per lane (OP 4)
  for (int t = 0; t < trips; t++) {
#pragma unroll
    for (int u = 0; u < kValuUnroll; u++) {
#pragma unroll
      for (int q = 0; q < kValuChains; q++) {
        if (OP == 0) a[q] = __builtin_fmaf(a[q], b, c);
        else if (OP == 1) a[q] = __builtin_amdgcn_rcpf(a[q]);
        else if (OP == 2) a[q] = a[q] * b;
        else if (OP == 3) a[q] = a[q] + c;
        else if (OP == 4) a[q] = __builtin_fmaf(a[q], bl, cl);
        else if (OP == 5) a[q] = __builtin_fmaf(bl, cl, a[q]);
        else if (OP == 6) a[q] = __builtin_fmaxf(__builtin_fmaxf(a[q], bl), cl);
        else a[q] = __builtin_fmaf(a[q], 0.999f, 1e-4f);
      }
    }
  }
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < kValuChains; q++) s += a[q];
  d[(long long)blockIdx.x * kProbeThreads + threadIdx.x] = s;
}

}  // namespace
}  // namespace lsm

using namespace lsm;

extern "C" {

double lsm_amd_bandwidth_probe(int kind, const void* src, void* dst, long long bytes, int workgroups, int reps) {
  const char* fn = "lsm_amd_bandwidth_probe";
  Call c(fn);
  if (kind < 0 || kind > 3) raise_error(std::string(fn) + ": kind is 0, 1, 2 or 3");
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
    else if (kind == 3)
      hipLaunchKernelGGL(k_probe_5r1w_b4, dim3((unsigned)workgroups), dim3(kProbeThreads), 0, c.stream(),
                         static_cast<const float*>(src), static_cast<float*>(dst), bytes / 4);
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

// src: 5 consecutive fields, dst: 1 field, each of px * py * (n + 6) floats;
// n a multiple of 32 (x), 64 (y) and kz.  Returns ms per launch (average of
// reps after one warm-up launch), or -1 on error.
double lsm_amd_march_probe(const void* src, void* dst, int n, int px, int py, int kz, int reps) {
  const char* fn = "lsm_amd_march_probe";
  Call c(fn);
  if (n <= 0 || n % kMarchTX || n % (kMarchTY * kNumXcd) || kz <= 0 || n % kz || reps <= 0)
    raise_error(std::string(fn) + ": n must be a positive multiple of 64 and of kz");
  if (px < n + 64 || py < n + 6) raise_error(std::string(fn) + ": rows of at least n + 64 floats, n + 6 rows");
  if (!src || !dst) raise_error(std::string(fn) + ": missing array");
  if (!c.ok()) return -1.0;
  if (!(is_device_pointer(src) && is_device_pointer(dst))) {
    raise_error(std::string(fn) + ": src and dst must be device memory");
    return -1.0;
  }
  const long long plane = (long long)px * py, field = plane * (n + 6);
  const TileGrid tg = make_tile_grid(Box{0, n - 1, 0, n - 1, 0, n - 1}, kMarchTX, kMarchTY, kz);
  const int nblk = tg.count();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  check_hip(hipEventCreate(&e0), "hipEventCreate");
  check_hip(hipEventCreate(&e1), "hipEventCreate");
  if (!c.ok()) return -1.0;
  auto launch = [&]() {
    hipLaunchKernelGGL(k_probe_march, dim3((unsigned)nblk), dim3(kMarchTX * kMarchTY), 0, c.stream(),
                       static_cast<const float*>(src), static_cast<float*>(dst), field, tg, px, plane);
    c.launched("k_probe_march");
  };
  launch();
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

// dst: workgroups * 256 floats.  Each launch runs workgroups x 256 lanes of
// trips x 32 instructions (op as k_probe_valu's OP).  Returns ms per launch
// (average of reps after one warm-up launch), or -1 on error.
double lsm_amd_valu_probe(void* dst, int op, int workgroups, int trips, int reps) {
  const char* fn = "lsm_amd_valu_probe";
  Call c(fn);
  if (op < 0 || op > 7) raise_error(std::string(fn) + ": op is 0 to 7");
  if (workgroups <= 0 || trips <= 0 || reps <= 0) raise_error(std::string(fn) + ": sizes must be positive");
  if (!dst) raise_error(std::string(fn) + ": missing array");
  if (!c.ok()) return -1.0;
  if (!is_device_pointer(dst)) {
    raise_error(std::string(fn) + ": dst must be device memory");
    return -1.0;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  check_hip(hipEventCreate(&e0), "hipEventCreate");
  check_hip(hipEventCreate(&e1), "hipEventCreate");
  if (!c.ok()) return -1.0;
  auto launch = [&]() {
    auto* k = op == 0 ? k_probe_valu<0> : op == 1 ? k_probe_valu<1> : op == 2 ? k_probe_valu<2>
            : op == 3 ? k_probe_valu<3> : op == 4 ? k_probe_valu<4> : op == 5 ? k_probe_valu<5>
            : op == 6 ? k_probe_valu<6> : k_probe_valu<7>;
    hipLaunchKernelGGL(k, dim3((unsigned)workgroups), dim3(kProbeThreads), 0, c.stream(), static_cast<float*>(dst),
                       trips, 0.999f, 1e-4f);
    c.launched("k_probe_valu");
  };
  launch();
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
