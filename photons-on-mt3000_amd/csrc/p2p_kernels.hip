// gfx950 kernels of the short-range P2P path.
//
// Force kernel: persistent waves pop work items (i-particles of one target
// leaf against a range of its CSR source leaves) from a cost-sorted queue.
// i-particles live in registers, up to three per lane; the source particles of
// the item run through a wave-private LDS tile in windows of 64 that span
// consecutive source leaves, each particle moved from its leaf's frame into
// the target leaf's frame by an offset formed in FP64 per leaf pair (minimum
// image when the box is periodic) and added as an FP32 pair hi + lo, culled
// against the item's bounding box and compacted, with the loads of the next
// two windows in flight.  Targets smaller than 64 lane slots split every
// batch of the tile over G lane groups
// ("phases") that meet by a butterfly at the end.  When the phases are whole
// quads a lane reads one tile entry per four steps and the steps share them
// by DPP quad broadcasts folded into the first subtraction (the tile reads
// were 40% of a step's LDS bytes on the linear law at three i-particles per
// lane).  Pair arithmetic is FP32:
// the LOGTAB law needs 13 FP32 and 3 integer VALU operations per pair, the
// LINEAR law 10 and 3, neither a transcendental; a step's coefficient
// gathers issue before the previous step's polynomials.  FP32 partial sums
// fold into FP64 after at least 64 steps.
//
// Reference semantics (what is computed, not how): P2P_core / P2P_core_1 /
// P2P_kernel in src/gpu_kernel.impl.h:50-157, :260-362, :364-437.
#include "p2p_device.h"

#include <type_traits>

namespace pp2p {

// Orders this wave's LDS stores before its later LDS loads (and the loads of
// one batch before the stores of the next).  Waves never share a tile, so no
// workgroup barrier is needed.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Per-wave LDS directory of up to 64 source leaves (see process_item): the
// frame offset of each (target frame - source frame), formed in FP64 with the
// minimum image and kept as an unevaluated FP32 sum hi + lo (x: hi, y: lo).
struct SrcDir {
    float2 ox[kWave], oy[kWave], oz[kWave];
};
static_assert(sizeof(SrcDir) <= kLdsDirBytes, "source directory size");

// inclusive prefix sum over the wave (lane order)
__device__ __forceinline__ int wave_incl_scan(int v, int lane)
{
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int u = __shfl_up(v, o, kWave);
        if (lane >= o)
            v += u;
    }
    return v;
}

// lane k's double, as a wave-uniform value
__device__ __forceinline__ double rl_double(double v, int k)
{
    const unsigned long long b = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, k);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), k);
    return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fminf(v, __shfl_xor(v, o, kWave));
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fmaxf(v, __shfl_xor(v, o, kWave));
    return v;
}

// s = r_eff^2 / rc^2 with the softening rule, clamped to s <= 1 (the cutoff).
// Plummer folds eps^2 into the first product of r^2 (one FMA, no extra add).
// The clamps compare bit patterns as unsigned integers, which orders
// non-negative floats like their values (integer VALU ops issue at twice the
// FP32 rate); any NaN pattern is above the pattern of 1.0, so the result lies
// in [smin, 1] for every input.  SCALE: the differences are in input length
// units (absolute FP32 frame) and r^2 is scaled by 1 / rc^2 first.
template <int SOFT, bool SCALE = false>
__device__ __forceinline__ float soft_s(float dx, float dy, float dz, const LawConst &lc)
{
    constexpr unsigned kOne = 0x3f800000u;
    if (SCALE) {
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        if (SOFT == SOFT_REF)
            return __uint_as_float(min(max(__float_as_uint(r2 * lc.abs_s2), __float_as_uint(lc.smin)), kOne));
        return __uint_as_float(min(__float_as_uint(fmaf(r2, lc.abs_s2, lc.smin)), kOne));
    }
    if (SOFT == SOFT_REF) {                                 // max(r, eps), clamp
        const unsigned r2 = __float_as_uint(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
        return __uint_as_float(min(max(r2, __float_as_uint(lc.smin)), kOne));
    }
    return __uint_as_float(min(__float_as_uint(fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, lc.smin)))), kOne));
}

// LOGTAB law.  F(s) = g(u(s)) / s^{3/2} is a cubic in (s - s0) on
// each of `segs` segments per binary octave of s.  The segment start s0 is s
// with its low mantissa bits cleared, and the same bits shifted right are the
// byte address of the segment's coefficients in LDS (the table lies at the
// LDS offset where s in [2^-31, 1] lands, p2p_device.h), so a pair costs one
// AND, one SUB and one shift besides the FMAs: no square root, reciprocal
// square root or index arithmetic.  The octave [1, 2) holds zero
// coefficients: every pair at or beyond the cutoff contributes nothing.
// s always lies in [smin, 1] (smin >= 2^-31 is checked on the host; the
// clamps map a NaN to smin or 1), so the address stays inside the table.
__device__ __forceinline__ float4 logtab_segment(float s, const char *__restrict__ lds, unsigned seg_mask,
                                                 int tab_shift, float &t)
{
    const unsigned b0 = __float_as_uint(s) & seg_mask;
    t = s - __uint_as_float(b0);
    return *reinterpret_cast<const float4 *>(lds + (b0 >> tab_shift));
}

__device__ __forceinline__ float logtab_eval(float4 c, float t)
{
    return fmaf(fmaf(fmaf(c.w, t, c.z), t, c.y), t, c.x);
}

// LINEAR law.  F(s) = a + b s on 512 segments per binary octave: the low 5
// exponent bits and the top 9 mantissa bits of s, shifted into place, are the
// LDS byte address of the segment's (a, b) pair (one shift and one AND), and
// one FMA evaluates it.  s lies in [smin, 1] and the host places every
// octave from smin's up to the entry of s = 1 (zero) at those addresses, so
// the address stays inside the table.
__device__ __forceinline__ float2 linear_segment(float s, const char *__restrict__ lds)
{
    const unsigned off = (__float_as_uint(s) >> (kLinSegShift - 3)) & 0x1fff8u;
    return *reinterpret_cast<const float2 *>(lds + off);
}

// One pair, POLY law: F = exp(-u^2) h(r) / r^3 with h(r) = erfcx(u) + (2/sqrt(pi)) u
// as a degree-12 polynomial in r = sqrt(s) in [0, 1]; explicit cutoff select.
template <int SOFT, bool SCALE>
__device__ __forceinline__ void pair_poly(float dx, float dy, float dz, const LawConst &lc,
                                          float &ax, float &ay, float &az)
{
    const float s = soft_s<SOFT, SCALE>(dx, dy, dz, lc);
    const float rinv = __builtin_amdgcn_rsqf(s);
    const float r = s * rinv;
    const float e = __builtin_amdgcn_exp2f(s * lc.k2);
    float h = lc.poly[0];
#pragma unroll
    for (int d = 1; d <= kPolyDeg; ++d)
        h = fmaf(h, r, lc.poly[d]);
    float f = (rinv * rinv) * (rinv * (e * h));
    f = (s < 1.0f) ? f : 0.0f;
    ax = fmaf(dx, f, ax);
    ay = fmaf(dy, f, ay);
    az = fmaf(dz, f, az);
}

// A j-particle far beyond the cutoff of any i: pads a batch to a multiple of
// the number of phases and contributes exactly zero.
__device__ __forceinline__ float4 null_particle() { return make_float4(1e18f, 1e18f, 1e18f, 0.f); }

// Lane U of every quad to the four lanes of the quad (DPP quad_perm
// [U,U,U,U]).  Used as the first operand of a subtraction that is its only
// use, the move folds into it (v_sub_f32_dpp) and costs no VALU operation.
template <int U>
__device__ __forceinline__ float quad_bcast(float v)
{
    constexpr int ctrl = U | (U << 2) | (U << 4) | (U << 6);
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xf, 0xf, true));
}

// Step U of a block of four: the j-particle held by lane U of the quad,
// as its differences to i-slot 0
template <int U, typename F>
__device__ __forceinline__ void quad_step(const float4 &mine, float x0, float y0, float z0, F &pairs)
{
    pairs(quad_bcast<U>(mine.x) - x0, quad_bcast<U>(mine.y) - y0, quad_bcast<U>(mine.z) - z0);
}

#ifndef PP2P_WAVES_PER_EU
#define PP2P_WAVES_PER_EU 4
#endif

// One work item with IPL i-particles per lane.  Lane l belongs to phase
// p = l / L (L lanes per phase, G = 64 / L phases; lanes from G * L on are
// idle) and holds the i-slots q, q + L, ..., q + (IPL-1) L of the item,
// q = l % L; slots past n_i hold a far-away dummy.  Source particles are
// staged into the tile and consumed in batches of G * floor(64 / G); phase p
// takes tile entries p, p + G, p + 2G, ... (lanes of one phase read one
// address, neighbouring phases neighbouring addresses), so every j read
// feeds IPL independent pairs.  The G partial sums of an i-particle meet by
// a butterfly over lanes (G a power of two) or a chain of lane reads.
template <int IPL, int LAW, int SOFT, bool FOLD64, bool INDIRECT, bool COUNT>
__device__ __forceinline__ void process_item(const ForceArgs &a, const LawConst &lc, const Item &it,
                                             const char *__restrict__ lds, float4 *tile, SrcDir &dir,
                                             int lane, unsigned long long &n_in_total)
{
    const int ni = it.ni;
    const int L = it.pad0;                          // lanes per phase
    const int G = kWave / L;                        // phases
    const int phase = lane / L;
    const int q = lane - phase * L;
    const bool live = phase < G;
    const int rphase = live ? phase : 0;            // idle lanes shadow phase 0 with dummy i-slots
    // quad broadcast (below) when every quad lies in one phase
    const bool quads = (L & 3) == 0;
    // batch: the largest multiple of G (of 4 G with quads) <= 64
    const int bq = quads ? 4 * G : G;
    const int nb = bq * (kWave / bq);
    const unsigned long long below = (1ull << lane) - 1ull;

    float xi[IPL], yi[IPL], zi[IPL];
    bool valid[IPL];
    float bxl = 3.0e38f, bxh = -3.0e38f, byl = 3.0e38f, byh = -3.0e38f, bzl = 3.0e38f, bzh = -3.0e38f;
    auto i_index = [&](int k) {                     // particle index of slot k (valid slots)
        const int il = q + k * L;
        return INDIRECT ? a.ilist[it.i_beg + il] : it.i_beg + il;
    };
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        const int il = q + k * L;
        valid[k] = live && il < ni;
        xi[k] = yi[k] = zi[k] = 1e18f;
        if (valid[k]) {
            const float4 p = a.pos[i_index(k)];
            xi[k] = p.x; yi[k] = p.y; zi[k] = p.z;
            bxl = fminf(bxl, p.x); bxh = fmaxf(bxh, p.x);
            byl = fminf(byl, p.y); byh = fmaxf(byh, p.y);
            bzl = fminf(bzl, p.z); bzh = fmaxf(bzh, p.z);
        }
    }
    // bounding box of the item's i-particles (lane 0 always holds a valid
    // slot), kept in scalar registers
    auto uniform = [](float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); };
    bxl = wave_min(bxl); bxh = wave_max(bxh);
    byl = wave_min(byl); byh = wave_max(byh);
    bzl = wave_min(bzl); bzh = wave_max(bzh);
    const float cx = uniform(0.5f * (bxl + bxh)), hx = uniform(0.5f * (bxh - bxl));
    const float cy = uniform(0.5f * (byl + byh)), hy = uniform(0.5f * (byh - byl));
    const float cz = uniform(0.5f * (bzl + bzh)), hz = uniform(0.5f * (bzh - bzl));
    const float cull_lim = a.cull_lim;              // squared distance to the item box beyond which no pair reaches

    const double4 rt = a.ref[it.tgt];
    double sx[IPL], sy[IPL], sz[IPL];
    float fx[IPL], fy[IPL], fz[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        sx[k] = sy[k] = sz[k] = 0.0;
        fx[k] = fy[k] = fz[k] = 0.f;
    }
    unsigned n_in = 0;
    int cnt = 0;                                   // j-particles waiting in the tile

    const unsigned seg_mask = 0xffffffffu << lc.seg_shift;
    const int tab_shift = lc.tab_shift;
    // Slot k's differences come from slot 0's: d_k = d_0 + (x_0 - x_k), so a
    // step computes one difference per axis from the j-particle and the DPP
    // broadcast that supplies it has one use.  The extra rounding is that of
    // FP32 coordinates inside one leaf, the precision the positions carry.
    float oxi[IPL], oyi[IPL], ozi[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        oxi[k] = xi[0] - xi[k]; oyi[k] = yi[0] - yi[k]; ozi[k] = zi[0] - zi[k];
    }
    // A table-law step in two halves: `issue` forms the IPL pairs'
    // differences and s and starts their coefficient gathers; `finish`
    // evaluates the polynomials and accumulates.  The quad loop below runs
    // the next step's issue before this step's finish, so a wave has a
    // step's arithmetic to do while its gathers are in flight.
    using Coef = typename std::conditional<LAW == LAW_LINEAR, float2, float4>::type;
    struct Pend {
        float dx[IPL], dy[IPL], dz[IPL], t[IPL];    // t: s (LINEAR) or s - s0 (LOGTAB)
        Coef c[IPL];
    };
    auto issue = [&](float dx0, float dy0, float dz0, Pend &p) {
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            p.dx[k] = k == 0 ? dx0 : dx0 + oxi[k];
            p.dy[k] = k == 0 ? dy0 : dy0 + oyi[k];
            p.dz[k] = k == 0 ? dz0 : dz0 + ozi[k];
        }
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            if constexpr (LAW == LAW_LINEAR) {
                p.t[k] = soft_s<SOFT, !FOLD64>(p.dx[k], p.dy[k], p.dz[k], lc);
                p.c[k] = linear_segment(p.t[k], lds);
            } else {
                p.c[k] = logtab_segment(soft_s<SOFT, !FOLD64>(p.dx[k], p.dy[k], p.dz[k], lc), lds, seg_mask,
                                        tab_shift, p.t[k]);
            }
        }
    };
    auto finish = [&](const Pend &p) {
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            float f;
            if constexpr (LAW == LAW_LINEAR)
                f = fmaf(p.c[k].y, p.t[k], p.c[k].x);
            else
                f = logtab_eval(p.c[k], p.t[k]);
            fx[k] = fmaf(p.dx[k], f, fx[k]);
            fy[k] = fmaf(p.dy[k], f, fy[k]);
            fz[k] = fmaf(p.dz[k], f, fz[k]);
        }
    };
    constexpr bool kPipe = !COUNT && LAW != LAW_POLY;
    // the IPL pairs of one j given its differences to slot 0: all table
    // addresses first, then the gathers, then the polynomials, so that the
    // gathers of a step are in flight together
    auto pairs_of = [&](float dx0, float dy0, float dz0) {
        float dx[IPL], dy[IPL], dz[IPL];
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            dx[k] = k == 0 ? dx0 : dx0 + oxi[k];
            dy[k] = k == 0 ? dy0 : dy0 + oyi[k];
            dz[k] = k == 0 ? dz0 : dz0 + ozi[k];
        }
        if (!COUNT && LAW == LAW_LINEAR) {
            float sv[IPL];
            float2 c[IPL];
#pragma unroll
            for (int k = 0; k < IPL; ++k) {
                sv[k] = soft_s<SOFT, !FOLD64>(dx[k], dy[k], dz[k], lc);
                c[k] = linear_segment(sv[k], lds);
            }
#pragma unroll
            for (int k = 0; k < IPL; ++k) {
                const float f = fmaf(c[k].y, sv[k], c[k].x);
                fx[k] = fmaf(dx[k], f, fx[k]);
                fy[k] = fmaf(dy[k], f, fy[k]);
                fz[k] = fmaf(dz[k], f, fz[k]);
            }
            return;
        }
        if (!COUNT && LAW == LAW_LOGTAB) {
            float t[IPL];
            float4 c[IPL];
#pragma unroll
            for (int k = 0; k < IPL; ++k)
                c[k] = logtab_segment(soft_s<SOFT, !FOLD64>(dx[k], dy[k], dz[k], lc), lds, seg_mask, tab_shift, t[k]);
#pragma unroll
            for (int k = 0; k < IPL; ++k) {
                const float f = logtab_eval(c[k], t[k]);
                fx[k] = fmaf(dx[k], f, fx[k]);
                fy[k] = fmaf(dy[k], f, fy[k]);
                fz[k] = fmaf(dz[k], f, fz[k]);
            }
            return;
        }
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            if (COUNT)
                n_in += valid[k] && soft_s<SOFT, !FOLD64>(dx[k], dy[k], dz[k], lc) < 1.0f;
            else
                pair_poly<SOFT, !FOLD64>(dx[k], dy[k], dz[k], lc, fx[k], fy[k], fz[k]);
        }
    };
    // FP32 partial sums fold into FP64 once at least 64 steps have gone into
    // them (a step adds one pair per slot), so every FP32 sum holds at most
    // 64 + 63 terms whatever the batch length; a batch of a many-phase item
    // is a few steps only.
    int unfolded = 0;                               // steps in the FP32 sums (wave-uniform)
    auto fold = [&]() {
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            sx[k] += (double)fx[k]; sy[k] += (double)fy[k]; sz[k] += (double)fz[k];
            fx[k] = 0.f; fy[k] = 0.f; fz[k] = 0.f;
        }
        unfolded = 0;
    };
    // One batch of n (a multiple of G) tile entries; phase p takes entries
    // p, p + G, p + 2G, ...  When L is a multiple of 4 every quad lies in one
    // phase: per block of four steps each lane reads one entry, lane c of a
    // quad entry (ub + c) G + phase, and step ub + u takes lane u's by a quad
    // broadcast (one tile read per four steps, the pair arithmetic
    // unchanged).  Other shapes read the tile every step, the pointer
    // stepping by G.  One loop per form serves every phase count: specialised
    // copies per count measured slower, the kernel outgrowing the
    // instruction cache.
    const float x0 = xi[0], y0 = yi[0], z0 = zi[0];
    auto run_batch = [&](int n) {
        const int ch = n / G;
        if (kPipe && quads) {
            // as below, software-pipelined: step u + 1's gathers issue before
            // step u's polynomials, and each block reads the next block's
            // entry first (on the last block an entry past the batch, at most
            // n + 4 G - 1 <= 127 < kTileCap: the wave's own tile, unused)
            float4 mine = tile[(lane & 3) * G + rphase];
            for (int ub = 0; ub < ch; ub += 4) {
                const float4 next = tile[(ub + 4 + (lane & 3)) * G + rphase];
                Pend A, B;
                issue(quad_bcast<0>(mine.x) - x0, quad_bcast<0>(mine.y) - y0, quad_bcast<0>(mine.z) - z0, A);
                issue(quad_bcast<1>(mine.x) - x0, quad_bcast<1>(mine.y) - y0, quad_bcast<1>(mine.z) - z0, B);
                finish(A);
                issue(quad_bcast<2>(mine.x) - x0, quad_bcast<2>(mine.y) - y0, quad_bcast<2>(mine.z) - z0, A);
                finish(B);
                issue(quad_bcast<3>(mine.x) - x0, quad_bcast<3>(mine.y) - y0, quad_bcast<3>(mine.z) - z0, B);
                finish(A);
                finish(B);
                mine = next;
            }
        } else if (quads) {
            // n is a multiple of 4 G (batch sizes and the null padding of the
            // last batch): ch is a multiple of 4 and every entry read, u G +
            // phase <= n - 1 < 64, lies in the batch
            for (int ub = 0; ub < ch; ub += 4) {
                const float4 mine = tile[(ub + (lane & 3)) * G + rphase];
                quad_step<0>(mine, x0, y0, z0, pairs_of);
                quad_step<1>(mine, x0, y0, z0, pairs_of);
                quad_step<2>(mine, x0, y0, z0, pairs_of);
                quad_step<3>(mine, x0, y0, z0, pairs_of);
            }
        } else {
            const float4 *tp = tile + rphase;
#pragma unroll 4
            for (int u = 0; u < ch; ++u) {
                const float4 pj = *tp;
                pairs_of(pj.x - x0, pj.y - y0, pj.z - z0);
                tp += G;
            }
        }
        unfolded += ch;
        if (FOLD64 && unfolded >= kWave)
            fold();
    };

    // Sources 64 at a time: lane s holds source s's first particle (relative
    // to the concatenated positions) and the inclusive prefix of the source
    // sizes in registers, its FP64 frame offset in the wave's LDS directory.
    // Particles are staged in windows of 64 that run over the concatenated
    // sources.  The sources that meet a window are found by a wave-uniform
    // walk over the prefix (lane reads, no memory round trip), and the loads
    // of the next two windows run while the current one is staged and
    // computed.
    for (int sb = it.s_beg; sb < it.s_end; sb += kWave) {
        const int nsrc = min(kWave, it.s_end - sb);
        int m_ip = 0, m_np = 0;
        double m_ox = 0.0, m_oy = 0.0, m_oz = 0.0;
        if (lane < nsrc) {
            const int sl = a.src[sb + lane];
            const int2 jr = a.rng[sl];
            m_ip = jr.x;
            m_np = jr.y;
            if (FOLD64) {
                const double4 rsrc = a.ref[sl];
                m_ox = rsrc.x - rt.x; m_oy = rsrc.y - rt.y; m_oz = rsrc.z - rt.z;
                if (a.box > 0.0) {
                    m_ox -= a.box * rint(m_ox / a.box);
                    m_oy -= a.box * rint(m_oy / a.box);
                    m_oz -= a.box * rint(m_oz / a.box);
                }
            }
        }
        const int pre = wave_incl_scan(m_np, lane);          // sources 0..lane
        const int total = __builtin_amdgcn_readlane(pre, kWave - 1);
        const int base = m_ip - (pre - m_np);               // pos index of concatenated position 0, source lane
        auto hi_lo = [](double v) {
            const float h = (float)v;
            return make_float2(h, (float)(v - (double)h));
        };
        dir.ox[lane] = hi_lo(m_ox);
        dir.oy[lane] = hi_lo(m_oy);
        dir.oz[lane] = hi_lo(m_oz);
        wave_lds_sync();

        // source and pos-index base of position w0 + lane; `cur` (uniform) is
        // a source at or before the one holding w0 and moves forward
        int cur = 0;
        auto locate = [&](int w0, int &ks, int &bs) {
            int sc = cur;
            ks = sc;
            bs = __builtin_amdgcn_readlane(base, sc);
            while (sc + 1 < nsrc) {
                const int st = __builtin_amdgcn_readlane(pre, sc);   // first position of source sc + 1
                if (st >= w0 + kWave)
                    break;
                ++sc;
                const int bn = __builtin_amdgcn_readlane(base, sc);
                if (w0 + lane >= st) {
                    ks = sc;
                    bs = bn;
                }
            }
            cur = sc;
        };
        // Two windows in flight: the even windows load into qA, the odd ones
        // into qB, each refilled with the window two ahead once staged, so a
        // load has two windows of batches to arrive in.  Every window issues
        // its refill, also past the block's end, so that the loads in flight
        // are the same on every path and a staging waits for its own load
        // only.  Lanes past the end load a particle of the block they never
        // use: index min(w + lane, total - 1) from the base of the source
        // located there.  Such a lane may land on a trailing empty source,
        // whose base makes the index its first particle minus one: -1 when
        // that empty source starts at particle 0, hence the clamp at 0 (with
        // total = 0 the index is max(first particle - 1, 0); the item's
        // target has particles, so the array is not empty).
        int ksA, ksB, bs;
        auto window_load = [&](int w, int &ks) {
            locate(w, ks, bs);
            return a.pos[max(bs + min(w + lane, total - 1), 0)];
        };
        float4 qA = window_load(0, ksA);
        float4 qB = window_load(kWave, ksB);
        // stage window w0 from q (target frame, cull, compact), then start
        // the load of window w0 + 128 into q
        auto stage = [&](int w0, float4 &q, int &ks) {
            const bool has = w0 + lane < total;
            float4 w = q;
            if (FOLD64) {                              // every lane's ks names a source of the block
                // (x + lo) + hi: the rounding of x + lo is that of a coordinate
                // inside the source leaf, the second that of the result, so
                // the shift stays within about one FP32 rounding of the
                // target-frame coordinate (an FP64 sum rounded once: half of it)
                // at two FP32 additions per axis instead of two conversions and
                // an FP64 addition
                const float2 ox = dir.ox[ks], oy = dir.oy[ks], oz = dir.oz[ks];
                w.x = (q.x + ox.y) + ox.x;
                w.y = (q.y + oy.y) + oy.x;
                w.z = (q.z + oz.y) + oz.x;
            }
            const float ex = fmaxf(fabsf(w.x - cx) - hx, 0.f);
            const float ey = fmaxf(fabsf(w.y - cy) - hy, 0.f);
            const float ez = fmaxf(fabsf(w.z - cz) - hz, 0.f);
            const bool keep = has && fmaf(ez, ez, fmaf(ey, ey, ex * ex)) <= cull_lim;
            const unsigned long long m = __ballot(keep);
            if (keep)
                tile[cnt + __popcll(m & below)] = make_float4(w.x, w.y, w.z, 0.f);
            cnt += __popcll(m);
            q = window_load(w0 + 2 * kWave, ks);
        };
        // cnt < nb before a window is added, so cnt < nb + 64 <= kTileCap and
        // every remainder below is < 64 (one lane per entry)
        auto drain = [&]() {
            while (cnt >= nb) {
                run_batch(nb);
                wave_lds_sync();
                const int rem = cnt - nb;
                float4 v;
                if (lane < rem)
                    v = tile[nb + lane];
                wave_lds_sync();
                if (lane < rem)
                    tile[lane] = v;
                cnt = rem;
                wave_lds_sync();
            }
        };
        // Unrolled by two with one exit at the bottom, so that at the top of
        // every pass qA's load is the older of the two in flight and the wait
        // before its staging leaves qB's outstanding.  A pass whose second
        // window starts at or past the end stages it with no lane in range:
        // it keeps nothing.
        for (int w0 = 0; w0 < total; w0 += 2 * kWave) {
            stage(w0, qA, ksA);
            wave_lds_sync();
            drain();
            stage(w0 + kWave, qB, ksB);
            wave_lds_sync();
            drain();
        }
        wave_lds_sync();                               // directory reused by the next batch
    }
    if (cnt > 0) {
        // cnt < nb, and nb is a multiple of bq: n <= nb <= 64
        const int n = (cnt + bq - 1) / bq * bq;
        if (lane < n - cnt)
            tile[cnt + lane] = null_particle();
        wave_lds_sync();
        run_batch(n);
        wave_lds_sync();
    }
    if (COUNT) {
        n_in_total += n_in;
        return;
    }
    if (FOLD64 && unfolded > 0)
        fold();
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        if (!FOLD64) {
            sx[k] = (double)fx[k]; sy[k] = (double)fy[k]; sz[k] = (double)fz[k];
        }
        // lanes q, q + L, ..., q + (G-1) L hold slot k of one i: sum into lane q
        // (a butterfly needs L and G both powers of two, i.e. G * L = 64)
        if (G * L == kWave) {
            for (int s = L; s < kWave; s <<= 1) {
                sx[k] += __shfl_xor(sx[k], s, kWave);
                sy[k] += __shfl_xor(sy[k], s, kWave);
                sz[k] += __shfl_xor(sz[k], s, kWave);
            }
        } else {
            const double x0 = sx[k], y0 = sy[k], z0 = sz[k];
            for (int m = 1; m < G; ++m) {
                const int src = q + m * L;                  // < G * L <= 64 for phase-0 lanes
                sx[k] += __shfl(x0, src, kWave);
                sy[k] += __shfl(y0, src, kWave);
                sz[k] += __shfl(z0, src, kWave);
            }
        }
        if (phase == 0 && valid[k]) {
            const int ipk = i_index(k);
            unsafeAtomicAdd(a.acc + ipk, sx[k] * a.out_scale);
            unsafeAtomicAdd(a.acc + a.npart + ipk, sy[k] * a.out_scale);
            unsafeAtomicAdd(a.acc + 2 * a.npart + ipk, sz[k] * a.out_scale);
        }
    }
}

// Persistent (QUEUE) or one-wave-per-item kernel.  The item's IPL selects
// one of three specialisations; the branch is uniform per wave.
template <int LAW, int SOFT, bool FOLD64, bool INDIRECT, bool QUEUE, bool COUNT = false>
__global__ __launch_bounds__(kWave * force_waves_per_group(LAW))
__attribute__((amdgpu_waves_per_eu(PP2P_WAVES_PER_EU)))
void p2p_force(ForceArgs a, LawConst lc)
{
    constexpr bool kLin = LAW == LAW_LINEAR;
    constexpr int kWpg = force_waves_per_group(LAW);
    // the workgroup's whole LDS (layouts: p2p_device.h)
    __shared__ float4 s_lds[(kLin ? kLinLdsBytes : kLdsBytes) / 16];
    char *lds = reinterpret_cast<char *>(s_lds);

    const int wv = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;

    if (LAW == LAW_LOGTAB || kLin) {
        float4 *tab = reinterpret_cast<float4 *>(lds + (kLin ? lc.lin_lo : lds_table_off(lc.seg_shift)));
        for (int k = threadIdx.x; k < a.table_len; k += blockDim.x)
            tab[k] = a.table[k];
        __syncthreads();
    }

    float4 *tile = reinterpret_cast<float4 *>(lds + (kLin ? kLinTileOff : kLdsTileOff)) + wv * kTileCap;
    SrcDir &dir = *reinterpret_cast<SrcDir *>(lds + (kLin ? kLinDirOff + wv * kLinDirStride
                                                         : kLdsDirOff + wv * kLdsDirBytes));
    // QUEUE with deal > 0 (the grid's wave count): the grid's first items are
    // dealt out without the counter, wave v of workgroup b taking item
    // v * gridDim.x + b (a workgroup's waves share a CU, which so starts on
    // items spread over the cost order), and later items come from the counter
    // offset by deal.  A launch of a few items per wave then does not start
    // with every wave on the same contended atomic (21 us on average on a
    // 0.6 ms launch).  Long launches take every item from the counter: the
    // staggered starts that the atomic gives measured 1% faster there.  deal
    // and the item index are wave-uniform and stay in scalar registers.
    const int deal = QUEUE ? a.deal : 0;
    int item_id = QUEUE ? (deal ? __builtin_amdgcn_readfirstlane(wv * (deal / kWpg)) + (int)blockIdx.x : 0)
                        : a.item_base + (int)blockIdx.x * kWpg + wv;
    bool first = true;
    unsigned long long n_in = 0;

    for (;;) {
        if (QUEUE && (!first || !deal)) {
            unsigned k = 0;
            if (lane == 0)
                k = atomicAdd(a.counter, 1u);
            item_id = deal + __builtin_amdgcn_readfirstlane((int)k);
        }
        first = false;
        if (item_id >= a.nitems)
            break;
        const Item it = a.items[item_id];
        switch (it.pad1) {
        case 1: process_item<1, LAW, SOFT, FOLD64, INDIRECT, COUNT>(a, lc, it, lds, tile, dir, lane, n_in); break;
        case 2: process_item<2, LAW, SOFT, FOLD64, INDIRECT, COUNT>(a, lc, it, lds, tile, dir, lane, n_in); break;
        default: process_item<3, LAW, SOFT, FOLD64, INDIRECT, COUNT>(a, lc, it, lds, tile, dir, lane, n_in); break;
        }
        if (!QUEUE)
            break;
    }
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1)
            n_in += __shfl_xor(n_in, o, kWave);
        if (lane == 0)
            atomicAdd(a.count, n_in);
    }
}

}  // namespace pp2p

// ---------------------------------------------------------------------------
// Host-side dispatch of the force kernel instantiations.
#include "p2p_launch.h"

namespace pp2p {

template <int LAW, int SOFT, bool FOLD64, bool INDIRECT, bool QUEUE>
static hipError_t launch_one(const ForceArgs &a, const LawConst &lc, int grid, hipStream_t st)
{
    hipLaunchKernelGGL((p2p_force<LAW, SOFT, FOLD64, INDIRECT, QUEUE>), dim3(grid),
                       dim3(kWave * force_waves_per_group(LAW)), 0, st, a, lc);
    return hipGetLastError();
}

template <int LAW, int SOFT, bool FOLD64, bool INDIRECT>
static hipError_t launch_q(const ForceArgs &a, const LawConst &lc, bool queue, int grid, hipStream_t st)
{
    return queue ? launch_one<LAW, SOFT, FOLD64, INDIRECT, true>(a, lc, grid, st)
                 : launch_one<LAW, SOFT, FOLD64, INDIRECT, false>(a, lc, grid, st);
}

template <int LAW, int SOFT, bool FOLD64>
static hipError_t launch_i(const ForceArgs &a, const LawConst &lc, bool indirect, bool queue, int grid, hipStream_t st)
{
    return indirect ? launch_q<LAW, SOFT, FOLD64, true>(a, lc, queue, grid, st)
                    : launch_q<LAW, SOFT, FOLD64, false>(a, lc, queue, grid, st);
}

template <int LAW, int SOFT>
static hipError_t launch_f(const ForceArgs &a, const LawConst &lc, bool fold64, bool indirect, bool queue, int grid, hipStream_t st)
{
    return fold64 ? launch_i<LAW, SOFT, true>(a, lc, indirect, queue, grid, st)
                  : launch_i<LAW, SOFT, false>(a, lc, indirect, queue, grid, st);
}

hipError_t launch_force(const ForceArgs &a, const LawConst &lc, const ForceVariant &v, int grid, hipStream_t st)
{
    if (v.law == LAW_LINEAR)
        return v.soft == SOFT_REF ? launch_f<LAW_LINEAR, SOFT_REF>(a, lc, v.fold64, v.indirect, v.queue, grid, st)
                                  : launch_f<LAW_LINEAR, SOFT_PLUMMER>(a, lc, v.fold64, v.indirect, v.queue, grid, st);
    if (v.law == LAW_LOGTAB)
        return v.soft == SOFT_REF ? launch_f<LAW_LOGTAB, SOFT_REF>(a, lc, v.fold64, v.indirect, v.queue, grid, st)
                                  : launch_f<LAW_LOGTAB, SOFT_PLUMMER>(a, lc, v.fold64, v.indirect, v.queue, grid, st);
    return v.soft == SOFT_REF ? launch_f<LAW_POLY, SOFT_REF>(a, lc, v.fold64, v.indirect, v.queue, grid, st)
                              : launch_f<LAW_POLY, SOFT_PLUMMER>(a, lc, v.fold64, v.indirect, v.queue, grid, st);
}

hipError_t force_occupancy(const ForceVariant &v, int table_len, int *groups_per_cu)
{
    (void)table_len;                                        // the table lives in the static LDS array
    if (v.law == LAW_LINEAR)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(groups_per_cu,
                (const void *)p2p_force<LAW_LINEAR, SOFT_REF, true, false, true>, kWave * kLinWaves, 0);
    if (v.law == LAW_LOGTAB)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(groups_per_cu,
                (const void *)p2p_force<LAW_LOGTAB, SOFT_REF, true, false, true>, kWave * kWavesPerGroup, 0);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(groups_per_cu,
            (const void *)p2p_force<LAW_POLY, SOFT_REF, true, false, true>, kWave * kWavesPerGroup, 0);
}

template <int SOFT, bool FOLD64>
static void launch_count_one(const ForceArgs &a, const LawConst &lc, bool indirect, int grid, hipStream_t st)
{
    if (indirect)
        hipLaunchKernelGGL((p2p_force<LAW_LOGTAB, SOFT, FOLD64, true, true, true>), dim3(grid),
                           dim3(kWave * kWavesPerGroup), 0, st, a, lc);
    else
        hipLaunchKernelGGL((p2p_force<LAW_LOGTAB, SOFT, FOLD64, false, true, true>), dim3(grid),
                           dim3(kWave * kWavesPerGroup), 0, st, a, lc);
}

hipError_t launch_count(const ForceArgs &a, const LawConst &lc, const ForceVariant &v, int grid, hipStream_t st)
{
    if (v.soft == SOFT_REF) {
        if (v.fold64)
            launch_count_one<SOFT_REF, true>(a, lc, v.indirect, grid, st);
        else
            launch_count_one<SOFT_REF, false>(a, lc, v.indirect, grid, st);
    } else {
        if (v.fold64)
            launch_count_one<SOFT_PLUMMER, true>(a, lc, v.indirect, grid, st);
        else
            launch_count_one<SOFT_PLUMMER, false>(a, lc, v.indirect, grid, st);
    }
    return hipGetLastError();
}

__global__ void k_reset_counter(unsigned *c) { *c = 0u; }

hipError_t launch_reset_counter(unsigned *c, hipStream_t st)
{
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, c);
    return hipGetLastError();
}

}  // namespace pp2p
