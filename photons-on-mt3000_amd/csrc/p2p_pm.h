// Particle-mesh long-range force on one GPU (internal to the library).
#pragma once
#include <hip/hip_runtime.h>

namespace pp2p {

// The mesh: N^3 cells, C order [x][y][z], periodic; cell-centred CIC.
struct PmArgs {
    const int *px, *py, *pz;   // int32 grid coordinates of particle p at p * stride
    int stride;                // 3: (x, y, z) records; 1: three planes
    const int *tag;       // null: every particle is local; else local iff tag == 1
    const int *tag_idx;   // null: tag[p]; else tag[tag_idx[p]]
    long long n;
    double int2pos;       // FP64 grid step (INT2POS)
    float box, mass;      // BOXSIZE, MASSPART (FP32, as the reference's cpuType)
    int nside;
};

// rho[N^3] += CIC mass of the local particles (FP32 weight products, FP64 sums).
hipError_t pm_deposit(const PmArgs &a, double *rho, hipStream_t st);
// spec[N][N][N/2+1] *= G(k) (the long-range Green's function with the CIC
// deconvolution and the mass-per-cell-to-density factor (N / box)^3; spec[0] = 0).
hipError_t pm_green(int nside, float split_radius, float box, double2 *spec, hipStream_t st);
// grad[3][N^3] = 4-point gradient of pot[N^3] per cell, interpolated with the
// CIC weights: acc32 != null stores acc32[3][n] (SoA, zero for particles that
// are not local); else acc64[3][n] += the local particles' accelerations.
hipError_t pm_interp(const PmArgs &a, const double *pot, double *grad, float *acc32, double *acc64,
                     hipStream_t st);
// p[0..n) = 0
hipError_t pm_zero(double *p, long long n, hipStream_t st);

}  // namespace pp2p
