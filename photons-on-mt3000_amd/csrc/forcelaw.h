// Host-side construction of the force-law constants and tables.
//
// Kernel units: lengths divided by the cutoff radius rc = 2 rs ucut, so the
// pair variable is s = r_eff^2 / rc^2 in (0, 1] and u = ucut sqrt(s).  The
// kernels evaluate F(s) = g(u) / s^{3/2} with
//     g(u) = erfc(u) + (2/sqrt(pi)) u exp(-u^2)
// (the reference's split kernel, src/gpu_kernel.impl.h:131-154 with the
// tables of :374-383), and the acceleration is m / rc^2 * sum dx_s F(s).
#pragma once

#include <vector>

#include "p2p_device.h"

namespace pp2p {

struct LawSetup {
    LawConst lc;                 // by-value kernel constants
    std::vector<float> table;    // LOGTAB: 4 floats per entry, kLogOctaves * segs entries
    int table_len = 0;           // entries
    double rc = 0.0;             // cutoff radius in input length units
    int segs = 0;                // segments per octave
    bool table_finite = true;    // false: a cubic coefficient of a used segment overflows FP32,
                                 // or eps / rc < 2^-15.5 (the used octaves would wrap); POLY only
    std::vector<float> lin;      // LINEAR: (a, b) per segment, from the lowest used octave to s = 1
    bool lin_ok = false;         // false: the used octaves outgrow the LDS (eps / rc < 2^-12.5)
};

// g(u) in FP64 from libm
double law_g(double u);
// F(s) = g(ucut sqrt(s)) / s^{3/2} in FP64
double law_F(double s, double ucut);

// Fills `out` for the given parameters.  segs_per_octave: 16 or 32.
// Returns false (with a message in *err) when the parameters are outside the
// table's range (eps / rc below 2^-15.5) or invalid.
bool law_setup(int softening, double eps, double rs, double ucut, int segs_per_octave, LawSetup *out,
               const char **err);

// FP32 evaluation of the LOGTAB law exactly as the kernel does it (for host
// tests): returns F for a given s already clamped by the softening rule.
float law_logtab_eval(const LawSetup &ls, float s);

// The same for the LINEAR law: F = fma(b, s, a) with the segment's (a, b).
float law_linear_eval(const LawSetup &ls, float s);

}  // namespace pp2p
