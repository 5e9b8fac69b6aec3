// Static device scratch of the no-allocation entry point
// (pp2p_compute_noalloc): the CSR, the work items, rocPRIM's temporary
// storage, the queue head and an error flag of one list live in a
// file-scope device array, so that a call allocates nothing.  Each GPU that
// the library runs on holds its own copy, at its own address, until the
// process ends.  Every call on one GPU shares that GPU's copy.
#include <hip/hip_runtime.h>

#include <mutex>

#include "p2p_scratch.h"

namespace pp2p {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[kScratchDevices];
}  // namespace

hipError_t scratch_address(int device, void **out)
{
    if (device < 0 || device >= kScratchDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        // the array's name is a host-side stand-in: kernels and copies take
        // the device address, looked up once per device with it current
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess)
            return e;
        if (cur != device)
            return hipErrorInvalidDevice;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

// flag[0] = 1 when a task names a leaf outside [0, nleaf)
__global__ void k_check_tasks(const int *__restrict__ tt, const int *__restrict__ ts, long long n, int nleaf,
                              int *__restrict__ flag)
{
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        const int a = tt[k], b = ts[k];
        if (a < 0 || a >= nleaf || b < 0 || b >= nleaf)
            flag[0] = 1;
    }
}

hipError_t launch_check_tasks(const int *tt, const int *ts, long long n, int nleaf, int *flag, hipStream_t st)
{
    if (n <= 0)
        return hipSuccess;
    long long g = (n + 255) / 256;
    if (g > 4096)
        g = 4096;
    hipLaunchKernelGGL(k_check_tasks, dim3((unsigned)g), dim3(256), 0, st, tt, ts, n, nleaf, flag);
    return hipGetLastError();
}

}  // namespace pp2p
