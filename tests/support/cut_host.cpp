// Host build of the work-item cut search (photons-on-mt3000_amd/csrc/
// p2p_cut.h) for tests/test_cut_search.py: the device's binary search run on
// synthetic source-size prefixes at task indices up to 2^31 - 1, beside a
// sequential scan on exact 64-bit sums.  No device is used.
#include "../../photons-on-mt3000_amd/csrc/p2p_cut.h"

namespace {

// every source holds `size` particles: P(k) = k * size, modulo 2^32 as stored
struct ConstPrefix {
    unsigned size;
    unsigned operator()(int k) const { return (unsigned)((unsigned long long)k * size); }
};

// source j holds 1 + (7919 j mod 64) particles (1..64; a permutation of 1..64
// over each block of 64 consecutive j, so a block sums to 2080)
unsigned long long var_prefix_exact(long long k)
{
    unsigned long long s = (unsigned long long)(k / 64) * 2080ull;
    for (long long j = k - k % 64; j < k; ++j)
        s += 1ull + (unsigned long long)((7919ll * j) % 64);
    return s;
}

struct VarPrefix {
    unsigned operator()(int k) const { return (unsigned)var_prefix_exact(k); }
};

}  // namespace

extern "C" {

int cut_const(int beg, int end, long long wi, long long split, unsigned size)
{
    return pp2p::range_cut(ConstPrefix{size}, beg, end, wi, split);
}

int cut_var(int beg, int end, long long wi, long long split)
{
    return pp2p::range_cut(VarPrefix{}, beg, end, wi, split);
}

// the first k in (beg, end] whose pairs of [beg, k] times wi exceed split, by
// a sequential scan on exact sums; end when none does
int cut_var_scan(int beg, int end, long long wi, long long split)
{
    const unsigned long long pb = var_prefix_exact(beg);
    for (long long k = beg + 1; k < end; ++k)
        if ((long long)(var_prefix_exact(k + 1) - pb) * wi > split)
            return (int)k;
    return end;
}

}  // extern "C"
