// Host entry points of the device kernels (internal to the library).
#pragma once
#include "p2p_device.h"

namespace pp2p {


struct ForceVariant {
    int law;        // LAW_LOGTAB / LAW_POLY
    int soft;       // SOFT_REF / SOFT_PLUMMER
    bool fold64;    // leaf frame + FP64 fold (false: absolute FP32 frame)
    bool indirect;  // i-particles through the active index list
    bool queue;     // persistent queue form
};

hipError_t launch_force(const ForceArgs &a, const LawConst &lc, const ForceVariant &v, int grid, hipStream_t st);
hipError_t force_occupancy(const ForceVariant &v, int table_len, int *groups_per_cu);
hipError_t launch_reset_counter(unsigned *c, hipStream_t st);
// counting pass of the queue form: adds the listed pairs with r_eff < rc to *a.count
hipError_t launch_count(const ForceArgs &a, const LawConst &lc, const ForceVariant &v, int grid, hipStream_t st);

}  // namespace pp2p
