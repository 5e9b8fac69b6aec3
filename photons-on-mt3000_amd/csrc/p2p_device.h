// Device-side types shared by the kernels (p2p_kernels.hip, p2p_build.hip)
// and the host context (p2p_context.cpp).  gfx950 only: wave64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pp2p {

constexpr int kWave = 64;            // CDNA wavefront width
constexpr int kWavesPerGroup = 4;    // force kernel workgroup = 256 lanes
constexpr int kTileCap = 128;        // j-tile capacity per wave (particles)
constexpr int kLogOctaves = 32;      // LOGTAB: cyclic table spans 32 binary octaves of s
constexpr int kLogSegsMax = 32;      // LOGTAB: segments per octave (16 or 32)
constexpr int kLogTabMax = kLogOctaves * kLogSegsMax;
constexpr int kPolyDeg = 12;         // POLY: degree of h(r)
constexpr int kIplMax = 4;           // i-particles per lane, at most
constexpr int kItemMaxI = kWave * kIplMax;   // i-particles per work item

// Units inside the kernels: lengths are divided by the cutoff radius rc
// (u = ucut * r / rc), so the cutoff sits at s = r^2 = 1 exactly.

enum { SOFT_REF = 0, SOFT_PLUMMER = 1 };
enum { LAW_LOGTAB = 0, LAW_POLY = 1 };

// One unit of the work queue: up to 64 i-particles of one target leaf against
// a contiguous range of that target's CSR source list.
struct alignas(16) Item {
    int tgt;      // target leaf id
    int i_beg;    // first i-particle (index into pos) or into ilist
    int ni;       // number of i-particles, 1..256
    int row;      // CSR row
    int s_beg;    // source range [s_beg, s_end) into src[]
    int s_end;
    int pad0;     // log2 of the number of phases G
    int pad1;     // i-particles per lane (IPL), 1..4
};

// Force-law constants, passed by value.
struct LawConst {
    float smin;       // eps^2 / rc^2 (REF: lower clamp of s; PLUMMER: added to s)
    int   seg_shift;  // LOGTAB: 23 - log2(segments per octave)
    int   seg_bits;   // LOGTAB: 5 + log2(segments per octave)
    float k2;         // POLY: -ucut^2 log2(e), so exp2(k2 s) = exp(-u^2)
    float poly[kPolyDeg + 1];   // POLY: h(r) coefficients, highest degree first
    float pad[3];
};

struct ForceArgs {
    const Item *items;        // work items (queue order)
    int nitems;
    int item_base;            // first item of this launch (GRID / PER_TARGET forms)
    unsigned *counter;        // queue head (QUEUE form)
    const float4 *pos;        // particle positions (scaled by 1/rc) in their leaf frame
    const double4 *ref;       // leaf frame origins (FP64, scaled by 1/rc)
    const int2 *rng;          // leaf (ip, np)
    const int *src;           // CSR source leaves
    const int *ilist;         // active i-particle indices (INDIRECT)
    double *acc;              // 3 planes of npart, accumulated
    long long npart;
    double box;               // periodic box (scaled by 1/rc), 0 = open
    double out_scale;         // m / rc^2: converts the scaled sums to accelerations
    const float4 *table;      // LOGTAB coefficients (kLogOctaves * segs entries)
    int table_len;
    int cull;                 // drop j-particles beyond rc of the item's bounding box
    unsigned long long *count;   // COUNT pass: pairs with s < 1
};

}  // namespace pp2p
