// Stable LSD radix sort of (uint64 key, uint32 value) pairs, 8 bits per pass.
//
// Per pass: each workgroup histograms a tile of 2048 keys; then one workgroup
// per digit scans that digit's row of the digit-major tile histogram (its
// counts over the tiles before each tile), starting from the keys of smaller
// digits, which it adds up from the totals the other digits' workgroups
// publish; then each workgroup ranks its tile's keys among equal digits with
// wave ballots (match masks) in eight rounds of 256 consecutive keys, which
// keeps equal keys in input order, stages them in LDS in digit order and
// writes them out in that order.  Used for the octree's octant-path keys,
// where stability is what reproduces the reference's stable counting sort
// (src/octree.cpp:106-121).
#include <algorithm>
#include <utility>

#include "gt_internal.h"

namespace gth {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;
constexpr int kWave = 64;
constexpr int kWaves = kBlock / kWave;
constexpr int kDigits = 256;

// Pairs to sort: n, or (n_dev set) the device count bounded by n.
__device__ inline size_t sort_count(size_t n, const unsigned long long *n_dev) {
  if (!n_dev) return n;
  const unsigned long long v = *n_dev;
  return v <= n ? static_cast<size_t>(v) : 0;
}

// Each tile's digit counts, digit-major.  Workgroup 0 also clears `clear`
// words (the first pass: the status words the row scans use next).
__global__ __launch_bounds__(kBlock) void tile_histogram(const uint64_t *__restrict__ keys,
                                                         size_t cap,
                                                         const unsigned long long *n_dev,
                                                         int shift, uint32_t ntiles,
                                                         uint32_t *__restrict__ counts,
                                                         unsigned long long *__restrict__ clear,
                                                         uint32_t clear_words) {
  __shared__ uint32_t h[kDigits];
  if (blockIdx.x == 0)
    for (uint32_t t = threadIdx.x; t < clear_words; t += kBlock) clear[t] = 0ull;
  const size_t n = sort_count(cap, n_dev);
  h[threadIdx.x] = 0;
  __syncthreads();
  const size_t base = static_cast<size_t>(blockIdx.x) * kTile;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const size_t i = base + k * kBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xff], 1u);
  }
  __syncthreads();
  counts[static_cast<size_t>(threadIdx.x) * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Workgroup d (by ticket): offsets[d * ntiles + t] becomes the output
// position of digit d's first key of tile t, i.e. the keys of smaller digits
// plus digit d's keys in tiles before t (counts, the tile histogram, is only
// read).  Each workgroup publishes its digit's total in status[d]
// (epoch << 32 | total; status starts cleared and each pass has its own epoch)
// and reads the totals of the smaller digits from there.  Digits are handed
// out by ticket in the order workgroups start, so a workgroup waits only for
// ones already running, which publish before they wait.  A wait is bounded:
// after kSpinLimit polls (it should never come to that) the waiting thread
// adds up that digit's row of the histogram itself, so the result is right
// either way, and counts the event in ctl[0].  ctl (may be null) is two words
// kept by the caller: [0] a count of such recounts, [1] a test setting, 0
// normally; k > 0 makes every wait give up after k polls (it waits for an
// epoch no workgroup publishes), so the recount runs for every smaller digit.
constexpr int kScanItems = 4;
constexpr uint32_t kSpinLimit = 1u << 22;

__global__ __launch_bounds__(kBlock) void digit_row_scan(uint32_t ntiles, uint32_t epoch,
                                                         uint32_t *__restrict__ ticket,
                                                         unsigned long long *status,
                                                         unsigned long long *ctl,
                                                         const uint32_t *__restrict__ counts,
                                                         uint32_t *__restrict__ offsets) {
  __shared__ uint32_t part[kWaves];
  __shared__ uint32_t digit;
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  if (threadIdx.x == 0) digit = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t d = digit;
  const uint32_t *row = counts + static_cast<size_t>(d) * ntiles;
  uint32_t *out_row = offsets + static_cast<size_t>(d) * ntiles;
  // the digit's total, published
  uint32_t v = 0;
  for (uint32_t t = threadIdx.x; t < ntiles; t += kBlock) v += row[t];
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  if (lane == 0) part[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0;
    for (int w = 0; w < kWaves; ++w) total += part[w];
    __hip_atomic_store(status + d, (static_cast<unsigned long long>(epoch) << 32) | total,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  // keys of smaller digits: thread j reads digit j's total
  uint32_t below = 0;
  if (threadIdx.x < d) {
    const unsigned long long test = ctl ? ctl[1] : 0ull;
    const uint32_t limit = test ? static_cast<uint32_t>(min(test, 1ull << 20)) : kSpinLimit;
    const uint32_t want = test ? (epoch ^ 0x80000000u) : epoch;
    unsigned long long w =
        __hip_atomic_load(status + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (uint32_t spin = 0; (w >> 32) != want && spin < limit; ++spin) {
      __builtin_amdgcn_s_sleep(1);
      w = __hip_atomic_load(status + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if ((w >> 32) == want) {
      below = static_cast<uint32_t>(w);
    } else {  // the total, from the histogram row itself
      const uint32_t *r = counts + static_cast<size_t>(threadIdx.x) * ntiles;
      for (uint32_t t = 0; t < ntiles; ++t) below += r[t];
      if (ctl) atomicAdd(ctl, 1ull);
    }
  }
  for (int off = kWave / 2; off > 0; off >>= 1) below += __shfl_xor(below, off, kWave);
  if (lane == 0) part[wave] = below;
  __syncthreads();
  uint32_t carry = 0;
  for (int w = 0; w < kWaves; ++w) carry += part[w];
  __syncthreads();
  // the row's exclusive scan, from carry
  for (uint32_t t0 = 0; t0 < ntiles; t0 += kBlock * kScanItems) {
    uint32_t c[kScanItems], sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      const uint32_t t = t0 + threadIdx.x * kScanItems + k;
      c[k] = t < ntiles ? row[t] : 0u;
      sum += c[k];
    }
    uint32_t incl = sum;
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t u = __shfl_up(incl, off, kWave);
      if (lane >= off) incl += u;
    }
    if (lane == kWave - 1) part[wave] = incl;
    __syncthreads();
    uint32_t before = carry, total = 0;
    for (int w = 0; w < kWaves; ++w) {
      if (w < wave) before += part[w];
      total += part[w];
    }
    uint32_t run = before + incl - sum;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      const uint32_t t = t0 + threadIdx.x * kScanItems + k;
      if (t < ntiles) out_row[t] = run;
      run += c[k];
    }
    carry += total;
    __syncthreads();
  }
}

// The tile's keys are ranked (stably) and placed in LDS in digit order first,
// then written out in that order: consecutive threads write consecutive
// addresses of a digit's run instead of scattering one key per digit bucket.
__global__ __launch_bounds__(kBlock) void tile_scatter(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ vals, size_t cap,
    const unsigned long long *n_dev, int shift, uint32_t ntiles,
    const uint32_t *__restrict__ offsets, uint64_t *__restrict__ keys_out,
    uint32_t *__restrict__ vals_out) {
  const size_t n = sort_count(cap, n_dev);
  __shared__ uint64_t skey[kTile];
  __shared__ uint32_t sval[kTile];
  __shared__ uint32_t run[kDigits];       // next tile-local slot of each digit
  __shared__ uint32_t out_base[kDigits];  // global position of tile-local slot 0 of each digit
  __shared__ uint32_t wave_cnt[kWaves][kDigits];  // per-round, per-wave digit counts
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const size_t base = static_cast<size_t>(blockIdx.x) * kTile;
  const uint32_t in_tile = base < n ? static_cast<uint32_t>(min(n - base, static_cast<size_t>(kTile))) : 0u;
  // the tile's pairs, read once into registers (item k of a thread is the
  // tile's pair k * 256 + thread)
  uint64_t rk[kItems];
  uint32_t rv[kItems];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const uint32_t i = k * kBlock + threadIdx.x;
    rk[k] = i < in_tile ? keys[base + i] : 0ull;
    rv[k] = i < in_tile ? vals[base + i] : 0u;
  }
  // 1. the tile's digit counts, then their exclusive sum: each digit's first
  // tile-local slot
  run[threadIdx.x] = 0;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kItems; ++k)
    if (k * kBlock + threadIdx.x < in_tile) atomicAdd(&run[(rk[k] >> shift) & 0xff], 1u);
  __syncthreads();
  if (threadIdx.x < kWave) {  // one wave: four digits per lane
    const uint32_t a0 = run[4 * lane], a1 = run[4 * lane + 1], a2 = run[4 * lane + 2],
                   a3 = run[4 * lane + 3];
    const uint32_t sum = a0 + a1 + a2 + a3;
    uint32_t incl = sum;
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t u = __shfl_up(incl, off, kWave);
      if (lane >= off) incl += u;
    }
    const uint32_t ex = incl - sum;
    run[4 * lane] = ex;
    run[4 * lane + 1] = ex + a0;
    run[4 * lane + 2] = ex + a0 + a1;
    run[4 * lane + 3] = ex + a0 + a1 + a2;
  }
  __syncthreads();
  out_base[threadIdx.x] =
      offsets[static_cast<size_t>(threadIdx.x) * ntiles + blockIdx.x] - run[threadIdx.x];
  // 2. stable ranks: rounds of 256 consecutive keys, lanes of equal digit
  // matched with wave ballots, waves in order
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
#pragma unroll
  for (int round = 0; round < kItems; ++round) {
    for (int w = 0; w < kWaves; ++w) wave_cnt[w][threadIdx.x] = 0;
    __syncthreads();
    const bool valid = round * kBlock + threadIdx.x < in_tile;
    const uint64_t key = rk[round];
    const uint32_t val = rv[round];
    const uint32_t d = valid ? static_cast<uint32_t>((key >> shift) & 0xff) : 0u;
    // Lanes of this wave holding the same digit.
    uint64_t same = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t set = __ballot(valid && ((d >> b) & 1u));
      same &= ((d >> b) & 1u) ? set : ~set;
    }
    const uint32_t rank_in_wave = __popcll(same & below);
    const bool leader = valid && rank_in_wave == 0;
    if (leader) wave_cnt[wave][d] = __popcll(same);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + rank_in_wave;
      for (int w = 0; w < wave; ++w) pos += wave_cnt[w][d];
      skey[pos] = key;
      sval[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < kWaves; ++w) add += wave_cnt[w][threadIdx.x];
    run[threadIdx.x] += add;
    __syncthreads();
  }
  // 3. out in digit order
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const uint32_t i = k * kBlock + threadIdx.x;
    if (i < in_tile) {
      const uint64_t key = skey[i];
      const uint32_t pos = out_base[(key >> shift) & 0xff] + i;
      keys_out[pos] = key;
      vals_out[pos] = sval[i];
    }
  }
}

__global__ __launch_bounds__(kBlock) void copy_pairs(const uint64_t *__restrict__ kin,
                                                     const uint32_t *__restrict__ vin, size_t cap,
                                                     const unsigned long long *n_dev,
                                                     uint64_t *__restrict__ kout,
                                                     uint32_t *__restrict__ vout) {
  const size_t n = sort_count(cap, n_dev);
  for (size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kBlock) {
    kout[i] = kin[i];
    vout[i] = vin[i];
  }
}

// Scratch (64-bit words): the digit totals' status words and one ticket
// counter per pass (uint32), cleared by the first pass's histogram launch;
// then the digit-major tile histogram of one pass (uint32) and the output
// positions computed from it (uint32, the same size).
constexpr int kMaxPasses = 8;  // 64-bit keys
constexpr size_t kStatusWords = kDigits + kMaxPasses / 2;

size_t histogram_words(size_t ntiles) { return (static_cast<size_t>(kDigits) * ntiles + 1) / 2; }

void sort_passes(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt,
                 size_t n, const unsigned long long *n_dev, int bits, unsigned long long *tmp,
                 unsigned long long *ctl, hipStream_t s) {
  const size_t ntiles = (n + kTile - 1) / kTile;
  unsigned long long *status = tmp;
  uint32_t *tickets = reinterpret_cast<uint32_t *>(tmp + kDigits);
  uint32_t *counts = reinterpret_cast<uint32_t *>(tmp + kStatusWords);
  uint32_t *offsets = reinterpret_cast<uint32_t *>(tmp + kStatusWords + histogram_words(ntiles));
  const int passes = std::min((bits + 7) / 8, kMaxPasses);
  uint64_t *kin = keys, *kout = keys_alt;
  uint32_t *vin = vals, *vout = vals_alt;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    hipLaunchKernelGGL(tile_histogram, dim3(static_cast<uint32_t>(ntiles)), dim3(kBlock), 0, s,
                       kin, n, n_dev, shift, static_cast<uint32_t>(ntiles), counts, tmp,
                       p == 0 ? static_cast<uint32_t>(kStatusWords) : 0u);
    GTH_LAUNCH_CHECK();
    hipLaunchKernelGGL(digit_row_scan, dim3(kDigits), dim3(kBlock), 0, s,
                       static_cast<uint32_t>(ntiles), static_cast<uint32_t>(p + 1), tickets + p,
                       status, ctl, counts, offsets);
    GTH_LAUNCH_CHECK();
    hipLaunchKernelGGL(tile_scatter, dim3(static_cast<uint32_t>(ntiles)), dim3(kBlock), 0, s, kin,
                       vin, n, n_dev, shift, static_cast<uint32_t>(ntiles), offsets, kout, vout);
    GTH_LAUNCH_CHECK();
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  if (kin != keys) {
    const uint32_t grid = static_cast<uint32_t>(std::min<size_t>((n + kBlock - 1) / kBlock, 2048));
    hipLaunchKernelGGL(copy_pairs, dim3(grid), dim3(kBlock), 0, s, kin, vin, n, n_dev, keys, vals);
    GTH_LAUNCH_CHECK();
  }
}

}  // namespace

size_t radix_tmp_words(size_t n) {
  const size_t ntiles = (n + kTile - 1) / kTile;
  return kStatusWords + 2 * histogram_words(ntiles);
}

void launch_row_scan(uint32_t rows, uint32_t ntiles, uint32_t epoch, uint32_t *ticket,
                     unsigned long long *status, unsigned long long *ctl, const uint32_t *counts,
                     uint32_t *offsets, hipStream_t s) {
  if (rows == 0 || ntiles == 0) return;
  if (rows > static_cast<uint32_t>(kDigits)) fail(GT_HIP_ERR_ARGUMENT, "row scan: too many rows");
  hipLaunchKernelGGL(digit_row_scan, dim3(rows), dim3(kBlock), 0, s, ntiles, epoch, ticket, status,
                     ctl, counts, offsets);
  GTH_LAUNCH_CHECK();
}

void radix_sort_u64(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt,
                    size_t n, int bits, unsigned long long *tmp, hipStream_t s,
                    unsigned long long *ctl) {
  if (n <= 1 || bits <= 0) return;
  sort_passes(keys, vals, keys_alt, vals_alt, n, nullptr, bits, tmp, ctl, s);
}

void radix_sort_u64_dev(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt,
                        const unsigned long long *n_dev, size_t cap, int bits,
                        unsigned long long *tmp, hipStream_t s, unsigned long long *ctl) {
  if (cap <= 1 || bits <= 0) return;
  sort_passes(keys, vals, keys_alt, vals_alt, cap, n_dev, bits, tmp, ctl, s);
}

}  // namespace gth
