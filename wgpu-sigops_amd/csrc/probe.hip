// probe.hip -- VALU issue-rate microbenchmark: the peak of the roofline this engine is bound by.
//
// The path is exact integer multiply-accumulate work, so its ceiling is the rate of
// v_mad_u64_u32 (32x32 -> 64-bit multiply with a 64-bit addend) over all SIMDs. The probe runs
// NCH independent accumulator chains per lane, each step exactly one v_mad_u64_u32 whose addend and
// result are the same aligned VGPR pair (the form the field arithmetic uses), on a grid that fills
// every SIMD, and reports the achieved operations per second. It also reports the rate of
// v_fma_f64 for comparison.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/sigops.h"

namespace {

constexpr int NCH = 16;
constexpr int ITERS = 4096;

__global__ __launch_bounds__(256) void k_probe_mad(uint32_t* out, uint32_t b) {
  uint64_t acc[NCH];
#pragma unroll
  for (int c = 0; c < NCH; c++) acc[c] = threadIdx.x * 131u + c;
  for (int i = 0; i < ITERS; i++) {
#pragma unroll
    for (int c = 0; c < NCH; c++) acc[c] = (uint64_t)(uint32_t)(acc[c] >> 32) * b + acc[c];
  }
  uint64_t r = 0;
#pragma unroll
  for (int c = 0; c < NCH; c++) r ^= acc[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)r ^ (uint32_t)(r >> 32);
}

__global__ __launch_bounds__(256) void k_probe_fma(uint32_t* out, double b) {
  double acc[NCH];
#pragma unroll
  for (int c = 0; c < NCH; c++) acc[c] = threadIdx.x + c;
  for (int i = 0; i < ITERS; i++) {
#pragma unroll
    for (int c = 0; c < NCH; c++) acc[c] = __builtin_fma(acc[c], b, 0.25);
  }
  double r = 0;
#pragma unroll
  for (int c = 0; c < NCH; c++) r += acc[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)r;
}

}  // namespace

extern "C" int sigops_probe_valu(int device, double* mad_per_s, double* fma64_per_s) {
  if (hipSetDevice(device) != hipSuccess) return SIGOPS_E_DEVICE;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return SIGOPS_E_DEVICE;
  const int blocks = cus * 8;  // 8 waves per SIMD
  uint32_t* out = nullptr;
  if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint32_t)) != hipSuccess) return SIGOPS_E_NOMEM;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const double ops = (double)blocks * 256 * ITERS * NCH;
  float best_mad = 1e30f, best_fma = 1e30f;
  for (int rep = 0; rep < 4; rep++) {
    float ms;
    (void)hipEventRecord(e0, nullptr);
    k_probe_mad<<<blocks, 256>>>(out, 0x9E3779B1u + rep);
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (rep) best_mad = ms < best_mad ? ms : best_mad;
    (void)hipEventRecord(e0, nullptr);
    k_probe_fma<<<blocks, 256>>>(out, 1.0000001 + rep * 1e-9);
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (rep) best_fma = ms < best_fma ? ms : best_fma;
  }
  const hipError_t err = hipGetLastError();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(out);
  if (err != hipSuccess) return SIGOPS_E_DEVICE;
  if (mad_per_s) *mad_per_s = ops / (best_mad * 1e-3);
  if (fma64_per_s) *fma64_per_s = ops / (best_fma * 1e-3);
  return SIGOPS_OK;
}
