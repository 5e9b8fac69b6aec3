// Static device scratch of the no-allocation entry point (internal).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace pp2p {

constexpr size_t kScratchBytes = 64ull << 20;   // per GPU, held until the process ends
constexpr int kScratchDevices = 64;

// Device address of this GPU's scratch array.  The first call for a device
// must be made with that device current (hipGetSymbolAddress); later calls
// return the cached address.
hipError_t scratch_address(int device, void **out);

// flag[0] = 1 when any task names a leaf outside [0, nleaf) (flag cleared by
// the caller beforehand).
hipError_t launch_check_tasks(const int *tt, const int *ts, long long n, int nleaf, int *flag, hipStream_t st);

}  // namespace pp2p
