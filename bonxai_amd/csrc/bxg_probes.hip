// Measurement instruments (include/bonxai_gpu_probes.h), built into their own
// library so that the map library carries only the map's kernels.  Each
// probe enqueues its kernel on the stream of a map context and returns.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/bonxai_gpu_probes.h"

namespace bxg {

constexpr uint32_t kProbeBlock = 256;
constexpr uint32_t kReadsPerLane = 4;  // as the lookup kernel: 4 lookups a lane

// a plain 16-byte-per-lane copy: calibration of the HBM byte counters
__global__ void __launch_bounds__(256) k_copy_probe(const uint4* __restrict__ src,
                                                    uint4* __restrict__ dst, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// n scattered 4-byte reads from a table of mask + 1 words (a power of two),
// kReadsPerLane a lane as in the lookup kernel: the rate the memory system
// gives the cell reads of a lookup batch whose coordinates come in no
// particular order
__device__ __forceinline__ uint64_t probe_mix(uint64_t x) {
  x ^= x >> 31;
  x *= 0x7fb5d329728ea185ull;
  x ^= x >> 27;
  x *= 0x81dadef4bc2dd44dull;
  return x ^ (x >> 33);
}

__global__ void __launch_bounds__(256) k_gather_probe(const uint32_t* __restrict__ table,
                                                      uint64_t mask, uint64_t n,
                                                      uint32_t* __restrict__ out) {
  const uint64_t base = (uint64_t)blockIdx.x * blockDim.x * kReadsPerLane + threadIdx.x;
  uint32_t v[kReadsPerLane];
#pragma unroll
  for (uint32_t j = 0; j < kReadsPerLane; ++j) {
    const uint64_t i = base + (uint64_t)j * blockDim.x;
    v[j] = i < n ? table[probe_mix(i) & mask] : 0u;
  }
#pragma unroll
  for (uint32_t j = 0; j < kReadsPerLane; ++j) {
    const uint64_t i = base + (uint64_t)j * blockDim.x;
    if (i < n) out[i] = v[j];
  }
}

// The FP64 operations of the scan front end, element-wise, for the
// correctly-rounded checks of sqrt, division, multiply-add order and floor.
__global__ void k_fp64_probe(const double* __restrict__ a, const double* __restrict__ b,
                             const double* __restrict__ c, uint64_t n, double* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[5 * i + 0] = __builtin_sqrt(a[i]);
  out[5 * i + 1] = a[i] / b[i];
  out[5 * i + 2] = c[i] + ((a[i] / b[i]) * c[i]);
  out[5 * i + 3] = __builtin_floor(a[i] * b[i]);
  out[5 * i + 4] = (a[i] * a[i] + b[i] * b[i]) + c[i] * c[i];
}

namespace {

int stream_of(bxg_ctx* c, hipStream_t* s) {
  void* p = nullptr;
  const int rc = bxg_ctx_stream(c, &p);
  *s = (hipStream_t)p;
  return rc;
}

// the launch's error as a BXG status
int launched() { return hipGetLastError() == hipSuccess ? BXG_OK : BXG_ERR_HIP; }

uint32_t blocks(uint64_t lanes, uint64_t most) {
  return (uint32_t)std::min<uint64_t>((lanes + kProbeBlock - 1) / kProbeBlock, most);
}

}  // namespace
}  // namespace bxg

using namespace bxg;

extern "C" {

int bxg_copy_probe(bxg_ctx* c, const void* src, void* dst, uint64_t bytes) {
  if (!c || (bytes && (!src || !dst)) || (bytes % 16)) return BXG_ERR_INVALID_ARG;
  if (!bytes) return BXG_OK;
  hipStream_t s;
  if (const int rc = stream_of(c, &s)) return rc;
  const uint64_t n = bytes / 16;
  hipLaunchKernelGGL(k_copy_probe, dim3(blocks(n, 8192)), dim3(kProbeBlock), 0, s, (const uint4*)src,
                     (uint4*)dst, n);
  return launched();
}

int bxg_gather_probe(bxg_ctx* c, const void* table, uint64_t table_bytes, uint64_t n, void* out) {
  const uint64_t words = table_bytes / 4;
  if (!c || table_bytes % 4 || words == 0 || (words & (words - 1)) || (n && (!table || !out)))
    return BXG_ERR_INVALID_ARG;
  if (!n) return BXG_OK;
  const uint64_t per_block = (uint64_t)kProbeBlock * kReadsPerLane;
  if ((n + per_block - 1) / per_block > UINT32_MAX) return BXG_ERR_INVALID_ARG;
  hipStream_t s;
  if (const int rc = stream_of(c, &s)) return rc;
  hipLaunchKernelGGL(k_gather_probe, dim3((uint32_t)((n + per_block - 1) / per_block)), dim3(kProbeBlock),
                     0, s, (const uint32_t*)table, words - 1, n, (uint32_t*)out);
  return launched();
}

int bxg_fp64_probe(bxg_ctx* c, const double* a, const double* b, const double* cc, uint64_t n,
                   double* out) {
  if (!c || (n && (!a || !b || !cc || !out))) return BXG_ERR_INVALID_ARG;
  if (!n) return BXG_OK;
  if ((n + kProbeBlock - 1) / kProbeBlock > UINT32_MAX) return BXG_ERR_INVALID_ARG;
  hipStream_t s;
  if (const int rc = stream_of(c, &s)) return rc;
  hipLaunchKernelGGL(k_fp64_probe, dim3((uint32_t)((n + kProbeBlock - 1) / kProbeBlock)),
                     dim3(kProbeBlock), 0, s, a, b, cc, n, out);
  return launched();
}

}  // extern "C"
