// The no-allocation entry point, wfa_device_batch_run_noalloc: the same run
// as wfa_device_batch_run (forward kernels, backtrace, retries) on a batch
// already resident on the device, for callers that cannot allocate device
// memory (code under stream capture, tight serving loops, harnesses that
// forbid hipMalloc).  Every device buffer beyond the batch's own inputs and
// outputs -- the run counters, the plan groups' order arrays, the HBM
// wavefront rings and the backtrace-record arena -- is carved from one
// static device array of kScratchBytes, which each GPU the library runs on
// holds until the process ends.
//
// Calls on one GPU share that GPU's array and race if they overlap (on two
// streams, or two aligners on two host threads): serialise them.  The array
// keeps the last call's contents; a run clears what must start at zero (the
// counters) itself.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "wfa_internal.h"

namespace {

constexpr size_t kScratchBytes = 256ull << 20;
constexpr int kMaxDevices = 64;

// The device array.  In host code its name stands for a host-side shadow,
// not a device address: only the pointer that hipGetSymbolAddress returns
// for the current device is ever handed to kernels or copies.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the array on device `dev`, looked up once per device
// with that device current (each device holds its own copy at its own
// address); the caller's current device is restored.
void* scratch_on_device(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return nullptr;
    if (cur != dev && hipSetDevice(dev) != hipSuccess) return nullptr;
    void* ptr = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (rc != hipSuccess || ptr == nullptr) {
      fprintf(stderr, "[WFA] hipGetSymbolAddress of the scratch array failed on device %d: %s\n", dev,
              hipGetErrorString(rc));
      return nullptr;
    }
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" int wfa_device_batch_run_noalloc(wavefront_aligner_t* const wf, wfa_device_batch_t* batch) {
  if (!wf || !batch) return WF_STATUS_OOM;
  wfa_host::Engine* e = wfa_host::engine_of(wf);
  if (!e || !e->ok) {
    fprintf(stderr, "[WFA] wfa_device_batch_run_noalloc: the batch must be uploaded with this aligner first\n");
    return WF_STATUS_OOM;
  }
  char* base = static_cast<char*>(scratch_on_device(e->device));
  if (!base) return WF_STATUS_OOM;
  return wfa_host::device_batch_run_in(wf, batch, base, kScratchBytes);
}
