// Particle-mesh long-range force on one GPU (internal to the library).
#pragma once
#include <hip/hip_runtime.h>

namespace pp2p {

// The mesh: N^3 cells, C order [x][y][z], periodic; cell-centred CIC.
struct PmArgs {
    const int *posi;      // 3 per particle (x, y, z), int32 grid positions
    const int *tag;       // null: every particle is local; else local iff tag == 1
    long long n;
    double int2pos;       // FP64 grid step (INT2POS)
    float box, mass;      // BOXSIZE, MASSPART (FP32, as the reference's cpuType)
    int nside;
};

// rho[N^3] += CIC mass of the local particles (FP32 weight products, FP64 sums).
hipError_t pm_deposit(const PmArgs &a, double *rho, hipStream_t st);
// spec[N][N][N/2+1] *= G(k) (the long-range Green's function with the CIC
// deconvolution; spec[0] = 0).
hipError_t pm_green(int nside, float split_radius, float box, double2 *spec, hipStream_t st);
// grad[3][N^3] = 4-point gradient of pot[N^3] per cell; acc[3][n] (SoA) =
// grad interpolated with the CIC weights; zero for particles that are not
// local.
hipError_t pm_interp(const PmArgs &a, const double *pot, double *grad, float *acc, hipStream_t st);
// p[0..n) = 0
hipError_t pm_zero(double *p, long long n, hipStream_t st);

}  // namespace pp2p
