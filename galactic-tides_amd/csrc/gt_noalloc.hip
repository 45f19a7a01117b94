// The lists-in evaluation for callers that cannot allocate device memory
// (gt_hip_tree_noalloc): the caller's arrays are already in device memory,
// and the evaluation's scratch - its status and count words and the input
// check's flags - is a fixed array in device memory (`scratch` below), one
// per GPU, held until the process ends.  Nothing here allocates.
//
// Host code must not pass `scratch` itself to a kernel: its name is a
// host-side stand-in, not a device address.  The address is looked up once
// per device with hipGetSymbolAddress (noalloc_scratch, called when a context
// is created with that device current) and kept per device.
#include <algorithm>
#include <mutex>
#include <string>

#include "gt_internal.h"

namespace gth {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxDevices = 64;
// Layout (64-bit words): the evaluation frame (frame::kWords), then the
// input check's flags.
constexpr size_t kCheckWord = frame::kWords;
constexpr size_t kScratchWords = 128;
static_assert(kCheckWord < kScratchWords, "scratch layout");

static __device__ __align__(256) unsigned long long scratch[kScratchWords];

std::mutex g_mutex;
void *g_scratch[kMaxDevices] = {};

// Every group and list entry against the sizes: a group's targets within
// [0, n) and at most kMaxGroup of them, its list ranges within the lists,
// every cell index below n_cells and every body index below n.  A failed
// check sets a flag bit in check and kAbortInput in the status word, so the
// evaluation that follows reads nothing.
__global__ __launch_bounds__(kBlock) void check_inputs(
    const gt_group *__restrict__ groups, uint32_t n_groups, uint32_t n, uint64_t n_cells,
    const uint32_t *__restrict__ cell_list, uint64_t n_cell_list,
    const uint32_t *__restrict__ body_list, uint64_t n_body_list,
    unsigned long long *__restrict__ frame, unsigned long long *__restrict__ check) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
  const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  unsigned long long bad = 0;
  for (uint64_t g = t0; g < n_groups; g += stride) {
    const gt_group G = groups[g];
    if (G.begin > G.end || G.end > n) bad |= 1ull;
    else if (G.end - G.begin > static_cast<uint32_t>(kMaxGroup)) bad |= 2ull;
    if (G.cell_begin > G.cell_end || G.cell_end > n_cell_list || G.body_begin > G.body_end ||
        G.body_end > n_body_list)
      bad |= 4ull;
  }
  for (uint64_t k = t0; k < n_cell_list; k += stride)
    if (cell_list[k] >= n_cells) bad |= 8ull;
  for (uint64_t k = t0; k < n_body_list; k += stride)
    if (body_list[k] >= n) bad |= 16ull;
  if (bad) {
    atomicOr(check, bad);
    atomicOr(frame + frame::kStatus, static_cast<unsigned long long>(kAbortInput));
  }
}

}  // namespace

void *noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices)
    fail(GT_HIP_ERR_UNSUPPORTED, "hip: device index beyond the no-allocation table");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
      fail(GT_HIP_ERR_DEVICE, std::string("hip: device scratch address: ") + hipGetErrorString(e));
    g_scratch[device] = ptr;
  }
  return g_scratch[device];
}

size_t noalloc_scratch_bytes() { return kScratchWords * sizeof(unsigned long long); }

void noalloc_tree(void *scratch_ptr, const float4 *bodies, uint32_t n, const Cell *cells,
                  uint64_t n_cells, const uint32_t *cell_list, uint64_t n_cell_list,
                  const uint32_t *body_list, uint64_t n_body_list, const gt_group *groups,
                  uint32_t n_groups, float G, float eps2, uint32_t quad, bool exact,
                  float4 *out, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
  unsigned long long *words = static_cast<unsigned long long *>(scratch_ptr);
  unsigned long long *frame = words;
  // fill kernels, not memsets: the scratch starts every call cleared
  launch_fill_u64(words, kScratchWords, 0ull, s);
  launch_fill_u64(frame + frame::kGroups, 1, n_groups, s);
  const uint64_t most = std::max<uint64_t>(n_groups, std::max(n_cell_list, n_body_list));
  const uint32_t blocks =
      static_cast<uint32_t>(std::min<uint64_t>((most + kBlock - 1) / kBlock, 2048));
  hipLaunchKernelGGL(check_inputs, dim3(std::max(blocks, 1u)), dim3(kBlock), 0, s, groups,
                     n_groups, n, n_cells, cell_list, n_cell_list, body_list, n_body_list, frame,
                     words + kCheckWord);
  GTH_LAUNCH_CHECK();
  GTH_CHECK(hipEventRecord(t0, s));
  launch_eval(bodies, cells, cell_list, body_list, groups, frame, n_groups, G, eps2, quad, exact,
              2, out, s);
  GTH_CHECK(hipEventRecord(t1, s));
  unsigned long long check = 0;
  GTH_CHECK(hipMemcpyAsync(&check, words + kCheckWord, sizeof(check), hipMemcpyDeviceToHost, s));
  GTH_CHECK(hipStreamSynchronize(s));
  if (check & 1ull) fail(GT_HIP_ERR_ARGUMENT, "hip tree: a group's targets are out of range");
  if (check & 2ull)
    fail(GT_HIP_ERR_ARGUMENT,
         "hip tree: a group has more than " + std::to_string(kMaxGroup) + " targets");
  if (check & 4ull) fail(GT_HIP_ERR_ARGUMENT, "hip tree: a group's list range is out of range");
  if (check & 8ull) fail(GT_HIP_ERR_ARGUMENT, "hip tree: cell index out of range");
  if (check & 16ull) fail(GT_HIP_ERR_ARGUMENT, "hip tree: body index out of range");
}

}  // namespace gth
