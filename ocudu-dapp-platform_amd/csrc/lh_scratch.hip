// Static device scratch of the allocation-free decode entry point
// (lh_pusch_decode_batch_noalloc): the descriptor block of one decoder launch (work-queue
// headers with their counters, work items, codeblock jobs) lives in this file-scope array
// instead of a buffer the library allocates.  Each GPU holds its own copy of the array,
// at its own address, until the process ends.
#include <hip/hip_runtime.h>

#include "lh_internal.h"

static __device__ __align__(256) uint8_t scratch[lh::kScratchBytes];

namespace lh {

// Device address of `scratch` on the current device.  The array's name is a host-side
// stand-in in host code; kernels and copies take this address.
hipError_t scratch_address(uint8_t** ptr)
{
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  *ptr = e == hipSuccess ? static_cast<uint8_t*>(p) : nullptr;
  return e;
}

} // namespace lh
