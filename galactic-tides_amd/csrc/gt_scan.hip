// Exclusive prefix sums: per-tile totals, then each tile scanned with its
// offset.  Up to kSummedTiles tiles each workgroup adds its offset up from the
// totals of the tiles before it (a few loads per thread, from L2) and the last
// tile writes the grand total: two passes, one for a single tile.  Beyond
// that a single workgroup scans the tile totals in a pass of its own, so the
// offset sums do not grow with the square of the tile count.  Sums are formed
// in 64 bits so callers can detect results that no longer fit 32 bits.
#include "gt_internal.h"

namespace gth {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;
constexpr int kWave = 64;
constexpr size_t kSummedTiles = 1024;

// Inclusive scan of one value per thread across the workgroup.
__device__ inline unsigned long long block_inclusive_scan(unsigned long long v,
                                                          unsigned long long *wave_tot) {
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const unsigned long long u = __shfl_up(v, off, kWave);
    if (lane >= off) v += u;
  }
  if (lane == kWave - 1) wave_tot[wave] = v;
  __syncthreads();
  unsigned long long before = 0;
  for (int w = 0; w < wave; ++w) before += wave_tot[w];
  __syncthreads();
  return v + before;
}

__global__ __launch_bounds__(kBlock) void tile_totals(const uint32_t *__restrict__ in, size_t n,
                                                      unsigned long long *__restrict__ tot) {
  __shared__ unsigned long long wt[kBlock / kWave];
  const size_t base = static_cast<size_t>(blockIdx.x) * kTile + threadIdx.x * kItems;
  unsigned long long s = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k)
    if (base + k < n) s += in[base + k];
  s = block_inclusive_scan(s, wt);
  if (threadIdx.x == kBlock - 1) tot[blockIdx.x] = s;
}

// One workgroup: exclusive scan of the tile totals in place, grand total to *total.
__global__ __launch_bounds__(kBlock) void scan_totals(unsigned long long *__restrict__ tot,
                                                      size_t ntiles,
                                                      unsigned long long *__restrict__ total) {
  __shared__ unsigned long long wt[kBlock / kWave];
  __shared__ unsigned long long carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (size_t base = 0; base < ntiles; base += kBlock) {
    const size_t k = base + threadIdx.x;
    const unsigned long long v = k < ntiles ? tot[k] : 0ull;
    const unsigned long long inc = block_inclusive_scan(v, wt);
    const unsigned long long c = carry;
    if (k < ntiles) tot[k] = c + inc - v;
    __syncthreads();
    if (threadIdx.x == kBlock - 1) carry = c + inc;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// kSum: tot holds the tiles' totals and the workgroup sums those before it
// (the last tile writes the grand total); otherwise tot holds the offsets.
// compact (may be null): for every position p whose value is 1 (flags),
// compact[out[p]] = p as well.
template <bool kSum>
__global__ __launch_bounds__(kBlock) void tile_scan(const uint32_t *in, uint32_t *out, size_t n,
                                                    const unsigned long long *__restrict__ tot,
                                                    unsigned long long *__restrict__ total,
                                                    uint32_t *__restrict__ compact) {
  __shared__ unsigned long long wt[kBlock / kWave];
  __shared__ unsigned long long offset_s;
  unsigned long long offset;
  if constexpr (kSum) {
    unsigned long long part = 0;
    for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kBlock) part += tot[t];
    part = block_inclusive_scan(part, wt);
    if (threadIdx.x == kBlock - 1) offset_s = part;
    __syncthreads();
    offset = offset_s;
  } else {
    offset = tot[blockIdx.x];
  }
  const size_t base = static_cast<size_t>(blockIdx.x) * kTile + threadIdx.x * kItems;
  uint32_t v[kItems];
  unsigned long long s = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    s += v[k];
  }
  const unsigned long long inc = block_inclusive_scan(s, wt);
  unsigned long long run = offset + inc - s;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    if (base + k < n) {
      out[base + k] = static_cast<uint32_t>(run);
      if (compact && v[k]) compact[run] = static_cast<uint32_t>(base + k);
    }
    run += v[k];
  }
  if (kSum && blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) *total = offset + inc;
}

__global__ void zero_total(unsigned long long *total) { *total = 0ull; }

}  // namespace

size_t scan_tmp_words(size_t n) { return (n + kTile - 1) / kTile + 1; }

void exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, unsigned long long *total,
                        unsigned long long *tmp, hipStream_t s, uint32_t *compact) {
  const size_t ntiles = (n + kTile - 1) / kTile;
  if (ntiles == 0) {
    hipLaunchKernelGGL(zero_total, dim3(1), dim3(1), 0, s, total);
    GTH_LAUNCH_CHECK();
    return;
  }
  if (ntiles > 1) {
    hipLaunchKernelGGL(tile_totals, dim3(static_cast<uint32_t>(ntiles)), dim3(kBlock), 0, s, in,
                       n, tmp);
    GTH_LAUNCH_CHECK();
  }
  if (ntiles <= kSummedTiles) {
    hipLaunchKernelGGL(tile_scan<true>, dim3(static_cast<uint32_t>(ntiles)), dim3(kBlock), 0, s,
                       in, out, n, tmp, total, compact);
    GTH_LAUNCH_CHECK();
    return;
  }
  hipLaunchKernelGGL(scan_totals, dim3(1), dim3(kBlock), 0, s, tmp, ntiles, total);
  GTH_LAUNCH_CHECK();
  hipLaunchKernelGGL(tile_scan<false>, dim3(static_cast<uint32_t>(ntiles)), dim3(kBlock), 0, s, in,
                     out, n, tmp, total, compact);
  GTH_LAUNCH_CHECK();
}

}  // namespace gth
