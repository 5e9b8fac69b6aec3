// Device-side types shared by the kernels (p2p_kernels.hip, p2p_build.hip)
// and the host context (p2p_context.cpp).  gfx950 only: wave64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pp2p {

constexpr int kWave = 64;            // CDNA wavefront width
constexpr int kWavesPerGroup = 8;    // force kernel workgroup = 512 lanes
constexpr int kTileCap = 128;        // j-tile capacity per wave (particles)

// LDS of one force-kernel workgroup: one 64 KiB array at LDS address 0.
//   [0, 16K)       j-tiles, kTileCap float4 per wave
//   [24K, 32K)     LOGTAB table, 16 segments per octave
//   [32K, 48K)     source directories, 2 KiB per wave
//   [48K, 64K)     LOGTAB table, 32 segments per octave
// The table sits where the bits of s address it directly: with s in
// [2^-31, 1] (biased exponent e in [96, 127]) the byte address of the entry
// of s is (bits(s) & ~0 << seg_shift) >> (seg_shift - 4) = e * 16 * segs + 16 *
// (top mantissa bits), i.e. 96 * 16 * segs bytes plus the cyclic entry index
// ((e mod 32) * segs + mantissa bits) times 16.  Two workgroups fit a CU's
// 160 KiB, 16 waves per CU.
constexpr int kLdsBytes = 65536;
constexpr int kLdsTileOff = 0;
constexpr int kLdsDirOff = 32768;
constexpr int kLdsDirBytes = 2048;   // sizeof(SrcDir)
__host__ __device__ constexpr int lds_table_off(int seg_shift) { return 96 << (27 - seg_shift); }
static_assert(kLdsTileOff + kWavesPerGroup * kTileCap * 16 <= 24576, "tiles overlap the 16-segment table");
static_assert(kLdsDirOff + kWavesPerGroup * kLdsDirBytes <= 49152, "directories overlap the 32-segment table");
// LINEAR law: F(s) = a + b s on 512 segments per binary octave of s, one
// float2 (a, b) per segment.  Its table outgrows half a CU's LDS, so the
// kernel runs one workgroup of 16 waves per CU with the whole 160 KiB:
//   [0, 24K)       source directories, 1.5 KiB per wave
//   [lo, 124K+8)   the table, where the bits of s address it: the entry of s
//                  is at byte (bits(s) >> 11) & 0x1fff8, i.e. 8 x ((biased
//                  exponent mod 32) * 512 + top 9 mantissa bits); lo = 4 KiB x
//                  (biased exponent of smin mod 32), and the last entry is s = 1
//   [128K, 160K)   j-tiles, kTileCap float4 per wave
// The used octaves must neither wrap (biased exponent of smin >= 96) nor
// reach the directories (lo >= 24K): smin >= 2^-25, eps / rc >= 2^-12.5.
constexpr int kLinSegShift = 14;     // 23 - log2(512)
constexpr int kLinWaves = 16;        // waves per workgroup (one workgroup per CU)
constexpr int kLinLdsBytes = 163840;
constexpr int kLinDirOff = 0;
constexpr int kLinDirStride = 1536;  // 3 x 64 doubles
constexpr int kLinTableLo = 24576;   // lowest byte the table may start at
constexpr int kLinTableEnd = 31 * 4096 + 8;
constexpr int kLinTileOff = 131072;
static_assert(kLinDirOff + kLinWaves * kLinDirStride <= kLinTableLo, "directories overlap the linear table");
static_assert(kLinTableEnd <= kLinTileOff, "the linear table overlaps the tiles");
static_assert(kLinTileOff + kLinWaves * kTileCap * 16 <= kLinLdsBytes, "tiles outgrow the LDS");
constexpr int kLogOctaves = 32;      // LOGTAB: cyclic table spans 32 binary octaves of s
constexpr int kLogSegsMax = 32;      // LOGTAB: segments per octave (16 or 32)
constexpr int kLogTabMax = kLogOctaves * kLogSegsMax;
constexpr int kPolyDeg = 12;         // POLY: degree of h(r)
constexpr int kIplMax = 3;           // i-particles per lane, at most
constexpr int kItemMaxI = kWave * kIplMax;   // i-particles per work item

// Units inside the kernels: lengths are divided by the cutoff radius rc
// (u = ucut * r / rc), so the cutoff sits at s = r^2 = 1 exactly.

enum { SOFT_REF = 0, SOFT_PLUMMER = 1 };
enum { LAW_LOGTAB = 0, LAW_POLY = 1, LAW_LINEAR = 2 };

// workgroup size of the force kernel for a law
__host__ __device__ constexpr int force_waves_per_group(int law) { return law == LAW_LINEAR ? kLinWaves : kWavesPerGroup; }

// One unit of the work queue: up to 64 i-particles of one target leaf against
// a contiguous range of that target's CSR source list.
struct alignas(16) Item {
    int tgt;      // target leaf id
    int i_beg;    // first i-particle (index into pos) or into ilist
    int ni;       // number of i-particles, 1..256
    int row;      // CSR row
    int s_beg;    // source range [s_beg, s_end) into src[]
    int s_end;
    int pad0;     // lanes per phase L (1..64); G = 64 / L phases
    int pad1;     // i-particles per lane (IPL), 1..3
};

// Force-law constants, passed by value.
struct LawConst {
    float smin;       // eps^2 / rc^2 (REF: lower clamp of s; PLUMMER: added to s)
    int   seg_shift;  // LOGTAB: 23 - log2(segments per octave)
    int   seg_bits;   // LOGTAB: 5 + log2(segments per octave)
    float k2;         // POLY: -ucut^2 log2(e), so exp2(k2 s) = exp(-u^2)
    float poly[kPolyDeg + 1];   // POLY: h(r) coefficients, highest degree first
    int   tab_shift;  // LOGTAB: seg_shift - 4 (entry byte address = segment start bits >> tab_shift)
    int   lin_lo;     // LINEAR: LDS byte address of the table's first entry
    float abs_s2;     // absolute FP32 frame: 1 / rc^2 (positions stay in input units)
};

struct ForceArgs {
    const Item *items;        // work items (queue order)
    int nitems;
    int item_base;            // first item of this launch (GRID / PER_TARGET forms)
    unsigned *counter;        // queue head (QUEUE form)
    int deal;                 // QUEUE: > 0: the grid's wave count, its first items dealt out (not from the counter)
    const float4 *pos;        // particle positions (scaled by 1/rc) in their leaf frame
    const double4 *ref;       // leaf frame origins (FP64, scaled by 1/rc)
    const int2 *rng;          // leaf (ip, np)
    const int *src;           // CSR source leaves
    const int *ilist;         // active i-particle indices (INDIRECT)
    double *acc;              // 3 planes of npart, accumulated
    long long npart;
    double box;               // periodic box (scaled by 1/rc), 0 = open
    double out_scale;         // m / rc^2: converts the scaled sums to accelerations
    const float4 *table;      // LOGTAB coefficients (kLogOctaves * segs entries); LINEAR: (a, b) pairs
    int table_len;            // float4 units
    float cull_lim;           // drop j-particles whose squared distance to the item's bounding box
                              // exceeds this (rc^2 (1 + 2^-10) in the frame's units; 3e38: no culling)
    unsigned long long *count;   // COUNT pass: pairs with s < 1
};

}  // namespace pp2p
