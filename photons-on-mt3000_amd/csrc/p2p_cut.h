// Source-range cut of a CSR row, shared by the item-build kernels and the
// host (where it is tested at task indices beyond 2^30 and 2^31 - 1 without a
// device or a list of that size).
#pragma once

#include <hip/hip_runtime.h>

namespace pp2p {

// The cut after `beg` in a row whose tasks are [beg, end): the first k in
// (beg, end] with (pairs of tasks [beg, k]) * wi > split, or end when there
// is none.  P(k) is the exclusive prefix of the source sizes over the CSR
// modulo 2^32 (k = 0 .. ntask); differences within a row are exact.  The
// predicate grows with k, so a binary search finds what a sequential scan
// would.  Every P index lies in [beg, end], and the midpoint is formed
// without the sum lo + hi, which overflows int once task indices pass 2^30.
template <class Prefix>
__host__ __device__ inline int range_cut(const Prefix &P, int beg, int end, long long wi, long long split)
{
    const unsigned pb = P(beg);
    int lo = beg + 1, hi = end;
    while (lo < hi) {
        const int mid = lo + ((hi - lo) >> 1);          // beg < lo <= mid < hi <= end
        if ((long long)(P(mid + 1) - pb) * wi > split)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// The prefix as the device stores it: nsx[k].
struct PrefixArray {
    const unsigned *p;
    __host__ __device__ unsigned operator()(int k) const { return p[k]; }
};

}  // namespace pp2p
