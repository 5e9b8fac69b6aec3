// gfx950 kernels of the short-range P2P path.
//
// Force kernel: every wave owns one work item at a time (up to 64 i-particles
// of a target leaf against a range of its source leaves).  i-particles live
// in registers; each source leaf is staged through a wave-private LDS tile in
// chunks of 64 particles, converted on the way in from the source leaf's frame
// to the target leaf's frame with an FP64 offset (and a minimum-image wrap
// when the box is periodic).  Targets smaller than 64 particles split the
// j-loop of each chunk over G = 64 / n_i lane groups ("phases"), which are
// summed at the end.  Pair arithmetic is FP32 with the hardware v_rsq_f32;
// the per-chunk partial sums fold into FP64.
//
// Reference semantics (what is computed, not how): P2P_core / P2P_core_1 /
// P2P_kernel in src/gpu_kernel.impl.h:50-157, :260-362, :364-437.
#include "p2p_device.h"

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
template <int SOFT>
__device__ __forceinline__ float soft_s(float r2, const LawConst &lc)
{
    if (SOFT == SOFT_REF)
        return __builtin_amdgcn_fmed3f(r2, lc.smin, 1.0f);     // max(r, eps), clamp
    return fminf(r2 + lc.smin, 1.0f);
}

// One pair, LOGTAB law.  F(s) = g(u(s)) / s^{3/2} is a cubic in (s - s0) on
// each of `segs` segments per binary octave of s; the segment is named by the
// exponent and top mantissa bits of s itself (cyclic over 32 octaves), so no
// square root or reciprocal square root is evaluated.  The octave [1, 2) holds
// zero coefficients: every pair at or beyond the cutoff contributes nothing.
template <int SOFT>
__device__ __forceinline__ void pair_logtab(float xi, float yi, float zi, float4 pj,
                                            const float4 *__restrict__ tab,
                                            const LawConst &lc,
                                            float &ax, float &ay, float &az)
{
    const float dx = pj.x - xi;
    const float dy = pj.y - yi;
    const float dz = pj.z - zi;
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float s = soft_s<SOFT>(r2, lc);
    const unsigned bits = __float_as_uint(s);
    const unsigned seg = __builtin_amdgcn_ubfe(bits, lc.seg_shift, lc.seg_bits);
    const float s0 = __uint_as_float(bits & (0xffffffffu << lc.seg_shift));
    const float t = s - s0;
    const float4 c = tab[seg];
    const float f = fmaf(fmaf(fmaf(c.w, t, c.z), t, c.y), t, c.x);
    ax = fmaf(dx, f, ax);
    ay = fmaf(dy, f, ay);
    az = fmaf(dz, f, az);
}

// One pair, POLY law: F = exp(-u^2) h(r) / r^3 with h(r) = erfcx(u) + (2/sqrt(pi)) u
// as a degree-12 polynomial in r = sqrt(s) in [0, 1]; explicit cutoff select.
template <int SOFT>
__device__ __forceinline__ void pair_poly(float xi, float yi, float zi, float4 pj,
                                          const LawConst &lc,
                                          float &ax, float &ay, float &az)
{
    const float dx = pj.x - xi;
    const float dy = pj.y - yi;
    const float dz = pj.z - zi;
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float s = soft_s<SOFT>(r2, lc);
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

// FOLD64: leaf frames, FP64 fold of every batch's FP32 sums; otherwise one
// FP32 accumulator over absolute coordinates (reference-style form).
template <int LAW, int SOFT, bool FOLD64, bool INDIRECT, bool QUEUE, bool COUNT = false>
__global__ __launch_bounds__(kWave * kWavesPerGroup)
void p2p_force(ForceArgs a, LawConst lc)
{
    extern __shared__ float4 s_tab[];                       // LOGTAB table
    __shared__ float4 s_tile[kWavesPerGroup][kTileCap];     // j-tiles, one per wave

    const int wv = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;

    if (LAW == LAW_LOGTAB) {
        for (int k = threadIdx.x; k < a.table_len; k += blockDim.x)
            s_tab[k] = a.table[k];
        __syncthreads();
    }

    float4 *tile = s_tile[wv];
    const unsigned long long below = (1ull << lane) - 1ull;
    int item_id = QUEUE ? 0 : a.item_base + blockIdx.x * kWavesPerGroup + wv;

    for (;;) {
        if (QUEUE) {
            unsigned k = 0;
            if (lane == 0)
                k = atomicAdd(a.counter, 1u);
            item_id = __builtin_amdgcn_readfirstlane((int)k);
        }
        if (item_id >= a.nitems)
            break;
        const Item it = a.items[item_id];
        const int ni = it.ni;
        const int groups = kWave / ni;                  // phases, >= 1
        const int batch = groups * (kWave / groups);    // j per batch, 60..64
        // lane / ni for lane < 64 and ni <= 64 without a divide
        const int phase = (lane * ((65536 + ni - 1) / ni)) >> 16;
        const int il = lane - phase * ni;               // < ni for every lane
        const bool live = phase < groups;

        const int ip = INDIRECT ? a.ilist[it.i_beg + il] : it.i_beg + il;
        const float4 pi = a.pos[ip];
        const float xi = pi.x, yi = pi.y, zi = pi.z;

        // bounding box of the item's i-particles (all lanes hold some i)
        const float bxl = wave_min(xi), bxh = wave_max(xi);
        const float byl = wave_min(yi), byh = wave_max(yi);
        const float bzl = wave_min(zi), bzh = wave_max(zi);
        const float cx = 0.5f * (bxl + bxh), hx = 0.5f * (bxh - bxl);
        const float cy = 0.5f * (byl + byh), hy = 0.5f * (byh - byl);
        const float cz = 0.5f * (bzl + bzh), hz = 0.5f * (bzh - bzl);
        const float cull_lim = a.cull ? 1.0009765625f : 3.0e38f;   // s > 1 + 2^-10: no pair can reach

        const double4 rt = a.ref[it.tgt];
        double sx = 0.0, sy = 0.0, sz = 0.0;
        float fx = 0.f, fy = 0.f, fz = 0.f;
        int cnt = 0;                                   // j-particles waiting in the tile

        unsigned n_in = 0;                             // COUNT: pairs inside the cutoff
        auto run_batch = [&](int n) {                  // n: multiple of groups
            const int ch = n / groups;
            if (live) {
                const float4 *tp = tile + phase * ch;
#pragma unroll 4
                for (int u = 0; u < ch; ++u) {
                    const float4 pj = tp[u];
                    if (COUNT) {
                        const float dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
                        n_in += soft_s<SOFT>(fmaf(dz, dz, fmaf(dy, dy, dx * dx)), lc) < 1.0f;
                    } else if (LAW == LAW_LOGTAB)
                        pair_logtab<SOFT>(xi, yi, zi, pj, s_tab, lc, fx, fy, fz);
                    else
                        pair_poly<SOFT>(xi, yi, zi, pj, lc, fx, fy, fz);
                }
            }
            if (FOLD64) {
                sx += (double)fx; sy += (double)fy; sz += (double)fz;
                fx = 0.f; fy = 0.f; fz = 0.f;
            }
        };

        for (int s = it.s_beg; s < it.s_end; ++s) {
            const int sl = a.src[s];
            const int2 jr = a.rng[sl];
            double ox = 0.0, oy = 0.0, oz = 0.0;
            if (FOLD64) {
                const double4 rsrc = a.ref[sl];
                ox = rsrc.x - rt.x; oy = rsrc.y - rt.y; oz = rsrc.z - rt.z;
                if (a.box > 0.0) {
                    ox -= a.box * rint(ox / a.box);
                    oy -= a.box * rint(oy / a.box);
                    oz -= a.box * rint(oz / a.box);
                }
            }
            for (int c0 = 0; c0 < jr.y; c0 += kWave) {
                const bool has = lane < jr.y - c0;
                float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
                if (has) {
                    const float4 q = a.pos[jr.x + c0 + lane];
                    if (FOLD64) {
                        w.x = (float)((double)q.x + ox);
                        w.y = (float)((double)q.y + oy);
                        w.z = (float)((double)q.z + oz);
                    } else {
                        w.x = q.x; w.y = q.y; w.z = q.z;
                    }
                }
                const float ex = fmaxf(fabsf(w.x - cx) - hx, 0.f);
                const float ey = fmaxf(fabsf(w.y - cy) - hy, 0.f);
                const float ez = fmaxf(fabsf(w.z - cz) - hz, 0.f);
                const bool keep = has && fmaf(ez, ez, fmaf(ey, ey, ex * ex)) <= cull_lim;
                const unsigned long long m = __ballot(keep);
                if (keep)
                    tile[cnt + __popcll(m & below)] = w;
                cnt += __popcll(m);
                wave_lds_sync();
                while (cnt >= batch) {
                    run_batch(batch);
                    wave_lds_sync();
                    const int rem = cnt - batch;       // < batch: no overlap
                    float4 v;
                    if (lane < rem)
                        v = tile[batch + lane];
                    wave_lds_sync();
                    if (lane < rem)
                        tile[lane] = v;
                    cnt = rem;
                    wave_lds_sync();
                }
            }
        }
        if (cnt > 0) {
            const int n = groups * ((cnt + groups - 1) / groups);
            if (lane < n - cnt)
                tile[cnt + lane] = null_particle();
            wave_lds_sync();
            run_batch(n);
            wave_lds_sync();
        }
        if (!FOLD64) {
            sx = (double)fx; sy = (double)fy; sz = (double)fz;
        }
        if (groups > 1) {                               // sum the phases (tile is free now)
            double *red = reinterpret_cast<double *>(tile);
            red[lane] = sx;
            red[kWave + lane] = sy;
            red[2 * kWave + lane] = sz;
            wave_lds_sync();
            if (phase == 0) {
                for (int g = 1; g < groups; ++g) {
                    sx += red[lane + g * ni];
                    sy += red[kWave + lane + g * ni];
                    sz += red[2 * kWave + lane + g * ni];
                }
            }
            wave_lds_sync();
        }
        if (COUNT) {
            unsigned long long c = n_in;
            for (int o = 32; o > 0; o >>= 1)
                c += __shfl_xor(c, o, kWave);
            if (lane == 0)
                atomicAdd(a.count, c);
            if (!QUEUE)
                break;
            continue;
        }
        if (phase == 0) {
            unsafeAtomicAdd(a.acc + ip, sx * a.out_scale);
            unsafeAtomicAdd(a.acc + a.npart + ip, sy * a.out_scale);
            unsafeAtomicAdd(a.acc + 2 * a.npart + ip, sz * a.out_scale);
        }
        if (!QUEUE)
            break;
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
    const size_t lds = LAW == LAW_LOGTAB ? sizeof(float4) * (size_t)a.table_len : 0;
    hipLaunchKernelGGL((p2p_force<LAW, SOFT, FOLD64, INDIRECT, QUEUE>), dim3(grid),
                       dim3(kWave * kWavesPerGroup), lds, st, a, lc);
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
    if (v.law == LAW_LOGTAB)
        return v.soft == SOFT_REF ? launch_f<LAW_LOGTAB, SOFT_REF>(a, lc, v.fold64, v.indirect, v.queue, grid, st)
                                  : launch_f<LAW_LOGTAB, SOFT_PLUMMER>(a, lc, v.fold64, v.indirect, v.queue, grid, st);
    return v.soft == SOFT_REF ? launch_f<LAW_POLY, SOFT_REF>(a, lc, v.fold64, v.indirect, v.queue, grid, st)
                              : launch_f<LAW_POLY, SOFT_PLUMMER>(a, lc, v.fold64, v.indirect, v.queue, grid, st);
}

hipError_t force_occupancy(const ForceVariant &v, int table_len, int *groups_per_cu)
{
    const size_t lds = v.law == LAW_LOGTAB ? sizeof(float4) * (size_t)table_len : 0;
    if (v.law == LAW_LOGTAB)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(groups_per_cu,
                (const void *)p2p_force<LAW_LOGTAB, SOFT_REF, true, false, true>, kWave * kWavesPerGroup, lds);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(groups_per_cu,
            (const void *)p2p_force<LAW_POLY, SOFT_REF, true, false, true>, kWave * kWavesPerGroup, lds);
}

hipError_t launch_count(const ForceArgs &a, const LawConst &lc, const ForceVariant &v, int grid, hipStream_t st)
{
    const size_t lds = sizeof(float4) * (size_t)a.table_len;
    if (v.soft == SOFT_REF) {
        if (v.indirect)
            hipLaunchKernelGGL((p2p_force<LAW_LOGTAB, SOFT_REF, true, true, true, true>), dim3(grid),
                               dim3(kWave * kWavesPerGroup), lds, st, a, lc);
        else
            hipLaunchKernelGGL((p2p_force<LAW_LOGTAB, SOFT_REF, true, false, true, true>), dim3(grid),
                               dim3(kWave * kWavesPerGroup), lds, st, a, lc);
    } else {
        if (v.indirect)
            hipLaunchKernelGGL((p2p_force<LAW_LOGTAB, SOFT_PLUMMER, true, true, true, true>), dim3(grid),
                               dim3(kWave * kWavesPerGroup), lds, st, a, lc);
        else
            hipLaunchKernelGGL((p2p_force<LAW_LOGTAB, SOFT_PLUMMER, true, false, true, true>), dim3(grid),
                               dim3(kWave * kWavesPerGroup), lds, st, a, lc);
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
