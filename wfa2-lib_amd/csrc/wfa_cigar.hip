// Run-length CIGARs on the device.  The backtrace leaves one operation byte
// per column ('M', 'X', 'I', 'D', as cigar_t); for the host only the runs
// matter, in the form cigar_get_CIGAR returns with show_mismatches
// (alignment/cigar.c:180-239): (length << 4) | code, code 7 ('=') for
// matches, 8 (X), 1 (I), 2 (D).  A 100-300 bp read at a few percent error
// has a few dozen runs, so the runs are 4-10 times smaller than the bytes.
//
// Two passes, one wave per pair: count the runs (a run starts at the first
// byte and wherever a byte differs from the one before it: one ballot per 64
// bytes), an inclusive scan of the counts (hipcub), then write each run at
// its pair's offset.
#include <hipcub/hipcub.hpp>

#include "wfa_device.h"

namespace wfa {
namespace {

constexpr int kRunsWaves = 4;  // pairs per workgroup

__device__ __forceinline__ uint32_t sam_code(unsigned char op) {
  return op == 'M' ? 7u : op == 'X' ? 8u : op == 'I' ? 1u : op == 'D' ? 2u : 15u;
}

// The pair's CIGAR bytes [begin, end) in the operations arena.
__device__ __forceinline__ void cigar_span(const CigarRunsArgs& a, int p, long long* begin, long long* end) {
  *begin = a.ops_off[p] + a.ops_begin[p];
  *end = a.ops_off[p] + (long long)a.plen[p] + a.tlen[p];
  if (a.status && a.status[p] != 0 && !a.keep_failed) *end = *begin;  // no CIGAR for a failed pair
}

__global__ void __launch_bounds__(64 * kRunsWaves) cigar_count_kernel(CigarRunsArgs a) {
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * kRunsWaves + (threadIdx.x >> 6);
  if (p == 0 && lane == 0) a.offsets[0] = 0;
  if (p >= a.n) return;
  long long begin, end;
  cigar_span(a, p, &begin, &end);
  const unsigned char* ops = reinterpret_cast<const unsigned char*>(a.ops);
  long long cnt = 0;
  for (long long base = begin; base < end; base += 64) {
    const long long i = base + lane;
    const bool valid = i < end;
    const unsigned char c = valid ? ops[i] : 0;
    const unsigned char prev = (valid && i > begin) ? ops[i - 1] : 0;
    const bool start = valid && (i == begin || c != prev);
    cnt += __popcll(__ballot(start));
  }
  if (lane == 0) a.counts[p] = cnt;
}

__global__ void __launch_bounds__(64 * kRunsWaves) cigar_write_kernel(CigarRunsArgs a) {
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * kRunsWaves + (threadIdx.x >> 6);
  if (p >= a.n) return;
  long long begin, end;
  cigar_span(a, p, &begin, &end);
  if (begin >= end) return;
  const unsigned char* ops = reinterpret_cast<const unsigned char*>(a.ops);
  uint32_t* out = a.runs + a.offsets[p];
  const unsigned long long below = (1ull << lane) - 1ull;  // lanes under this one
  long long idx = 0;          // runs started before this chunk
  long long open_start = begin;  // start of the run still open at the chunk's beginning
  for (long long base = begin; base < end; base += 64) {
    const long long i = base + lane;
    const bool valid = i < end;
    const unsigned char c = valid ? ops[i] : 0;
    const unsigned char prev = (valid && i > begin) ? ops[i - 1] : 0;
    const bool start = valid && (i == begin || c != prev);
    const unsigned long long m = __ballot(start);
    if (start && i > begin) {
      // this start closes the run before it: index (its own index - 1)
      const unsigned long long earlier = m & below;
      const long long prev_start = earlier ? base + (63 - __clzll(earlier)) : open_start;
      const long long my = idx + __popcll(earlier);
      out[my - 1] = ((uint32_t)(i - prev_start) << 4) | sam_code(prev);
    }
    idx += __popcll(m);
    if (m) open_start = base + (63 - __clzll(m));
  }
  if (lane == 0) out[idx - 1] = ((uint32_t)(end - open_start) << 4) | sam_code(ops[end - 1]);
}

}  // namespace

hipError_t launch_cigar_runs_count(const CigarRunsArgs& a, void* temp, size_t* temp_bytes, hipStream_t stream) {
  if (temp == nullptr) {
    // scan storage query only
    return hipcub::DeviceScan::InclusiveSum(nullptr, *temp_bytes, a.counts, a.offsets + 1, a.n, stream);
  }
  if (a.n <= 0) return hipSuccess;
  const int grid = (a.n + kRunsWaves - 1) / kRunsWaves;
  hipLaunchKernelGGL(cigar_count_kernel, dim3(grid), dim3(64 * kRunsWaves), 0, stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipcub::DeviceScan::InclusiveSum(temp, *temp_bytes, a.counts, a.offsets + 1, a.n, stream);
}

hipError_t launch_cigar_runs_write(const CigarRunsArgs& a, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  const int grid = (a.n + kRunsWaves - 1) / kRunsWaves;
  hipLaunchKernelGGL(cigar_write_kernel, dim3(grid), dim3(64 * kRunsWaves), 0, stream, a);
  return hipGetLastError();
}

}  // namespace wfa
