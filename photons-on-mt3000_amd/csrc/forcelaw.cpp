// Host-side construction of the force-law constants and tables (FP64 libm,
// Chebyshev interpolation, rounded once to FP32).
#include "forcelaw.h"

#include <cmath>
#include <cstring>

namespace pp2p {

static const double kTwoOverSqrtPi = 1.1283791670955126;

double law_g(double u) { return std::erfc(u) + kTwoOverSqrtPi * u * std::exp(-u * u); }

double law_F(double s, double ucut)
{
    const double r = std::sqrt(s);
    return law_g(ucut * r) / (s * r);
}

// Coefficients (increasing degree) of the degree-`deg` polynomial in
// t in [0, w] interpolating f at the Chebyshev points of that interval.
template <typename Fn>
static void cheb_interp(Fn f, double w, int deg, double *coef)
{
    const int n = deg + 1;
    double x[16], y[16];
    for (int k = 0; k < n; ++k) {
        const double c = std::cos(M_PI * (k + 0.5) / n);   // in (-1, 1)
        x[k] = 0.5 * w * (1.0 + c);
        y[k] = f(x[k]);
    }
    // Newton divided differences, then expand to monomials in t
    double dd[16];
    for (int k = 0; k < n; ++k)
        dd[k] = y[k];
    for (int j = 1; j < n; ++j)
        for (int k = n - 1; k >= j; --k)
            dd[k] = (dd[k] - dd[k - 1]) / (x[k] - x[k - j]);
    double p[16] = {0};
    p[0] = dd[n - 1];
    int len = 1;
    for (int k = n - 2; k >= 0; --k) {
        // p = p * (t - x[k]) + dd[k]
        double q[16] = {0};
        for (int i = 0; i < len; ++i) {
            q[i + 1] += p[i];
            q[i] -= p[i] * x[k];
        }
        q[0] += dd[k];
        ++len;
        std::memcpy(p, q, sizeof(double) * len);
    }
    for (int k = 0; k < n; ++k)
        coef[k] = p[k];
}

bool law_setup(int softening, double eps, double rs, double ucut, int segs, LawSetup *out, const char **err)
{
    *err = nullptr;
    if (!(eps > 0.0) || !(rs > 0.0) || !(ucut > 0.0) || !(ucut <= 6.0)) {
        *err = "force law: eps, rs must be > 0 and 0 < ucut <= 6";
        return false;
    }
    if (segs != 16 && segs != 32) {
        *err = "force law: segments per octave must be 16 or 32";
        return false;
    }
    if (softening != SOFT_REF && softening != SOFT_PLUMMER) {
        *err = "force law: unknown softening";
        return false;
    }
    LawSetup &ls = *out;
    std::memset(&ls.lc, 0, sizeof(ls.lc));
    ls.rc = 2.0 * rs * ucut;
    ls.segs = segs;
    const double smin = (eps / ls.rc) * (eps / ls.rc);
    // the cyclic table holds 32 octaves: the used ones [e_lo, -1] plus the
    // zero octave [1, 2) must not wrap onto each other
    const int e_lo = (int)std::floor(std::log2(smin));
    if (smin >= 1.0) {
        *err = "force law: softening length at or beyond the cutoff radius";
        return false;
    }
    ls.lc.smin = (float)smin;
    if (!(ls.lc.smin > 0.0f)) {
        *err = "force law: eps / rc underflows FP32";
        return false;
    }
    int lg = 0;
    while ((1 << lg) < segs)
        ++lg;
    ls.lc.seg_shift = 23 - lg;
    ls.lc.seg_bits = 5 + lg;
    ls.lc.tab_shift = ls.lc.seg_shift - 4;
    ls.lc.k2 = (float)(-ucut * ucut / std::log(2.0));
    ls.lc.abs_s2 = (float)(1.0 / (ls.rc * ls.rc));

    // LOGTAB: entry index = (biased exponent mod 32) * segs + top mantissa bits
    ls.table_len = kLogOctaves * segs;
    ls.table.assign((size_t)ls.table_len * 4, 0.0f);
    // the used octaves [e_lo, -1] and the zero octave must not wrap onto each
    // other in the cyclic table: below eps / rc = 2^-15.5 only the polynomial
    // law applies
    if (-e_lo + 1 > kLogOctaves)
        ls.table_finite = false;
    for (int e = ls.table_finite ? e_lo : 0; e <= -1; ++e) {
        const int oct = (e + 127) & (kLogOctaves - 1);
        for (int k = 0; k < segs; ++k) {
            const double s0 = std::ldexp(1.0 + (double)k / segs, e);
            const double w = std::ldexp(1.0 / segs, e);
            double c[4];
            cheb_interp([&](double t) { return law_F(s0 + t, ucut); }, w, 3, c);
            float *dst = &ls.table[((size_t)oct * segs + k) * 4];
            for (int d = 0; d < 4; ++d) {
                dst[d] = (float)c[d];
                // the cubic coefficient grows as s^-4.5: below s ~ 3e-9 (eps / rc
                // ~ 6e-5) it leaves the FP32 range, and the table law is unusable
                if (!std::isfinite(dst[d]))
                    ls.table_finite = false;
            }
        }
    }
    // octave [1, 2) (e = 0) stays zero: the cutoff

    // LINEAR: a + b s on each of 512 segments per octave, the line through
    // F at the segment's two Chebyshev points (uncentred in s: a ~ 2.5 F and
    // b s ~ -1.5 F near the power law, so FP32 loses about 2e-7 of F to the
    // cancellation).  Octaves by the biased exponent of the FP32 smin (the
    // kernel's s never lies below it) up to the single entry of s = 1, where
    // every pair at or beyond the cutoff lands; it holds zero.
    {
        unsigned sb;
        std::memcpy(&sb, &ls.lc.smin, 4);
        const int eb_lo = (int)(sb >> 23);                       // biased exponent of smin
        const int nseg = 1 << (23 - kLinSegShift);
        const long long entries = (long long)(127 - eb_lo) * nseg + 1;
        ls.lin_ok = eb_lo >= 96 && (eb_lo & 31) * nseg * 8 >= kLinTableLo;
        if (ls.lin_ok) {
            ls.lin.assign((size_t)entries * 2, 0.0f);
            for (int eb = eb_lo; eb < 127; ++eb) {
                for (int k = 0; k < nseg; ++k) {
                    const double s0 = std::ldexp(1.0 + (double)k / nseg, eb - 127);
                    const double w = std::ldexp(1.0 / nseg, eb - 127);
                    const double x0 = s0 + 0.5 * w * (1.0 - std::sqrt(0.5));
                    const double x1 = s0 + 0.5 * w * (1.0 + std::sqrt(0.5));
                    const double f0 = law_F(x0, ucut), f1 = law_F(x1, ucut);
                    const double b = (f1 - f0) / (x1 - x0);
                    const double a = f0 - b * x0;
                    float *dst = &ls.lin[((size_t)(eb - eb_lo) * nseg + k) * 2];
                    dst[0] = (float)a;
                    dst[1] = (float)b;
                    if (!std::isfinite(dst[0]) || !std::isfinite(dst[1]))
                        ls.lin_ok = false;
                }
            }
            // entry index of s: (biased exponent mod 32) * 512 + top 9 mantissa bits
            ls.lc.lin_lo = (eb_lo & 31) * nseg * 8;
        }
        if (!ls.lin_ok)
            ls.lin.clear();
    }

    // POLY: h(r) = erfcx(ucut r) + (2/sqrt(pi)) ucut r on r in [0, 1]
    double hc[kPolyDeg + 1];
    cheb_interp(
        [&](double r) {
            const double u = ucut * r;
            return std::exp(u * u) * std::erfc(u) + kTwoOverSqrtPi * u;
        },
        1.0, kPolyDeg, hc);
    for (int d = 0; d <= kPolyDeg; ++d)
        ls.lc.poly[d] = (float)hc[kPolyDeg - d];
    return true;
}

float law_logtab_eval(const LawSetup &ls, float s)
{
    unsigned bits;
    std::memcpy(&bits, &s, 4);
    const unsigned seg = (bits >> ls.lc.seg_shift) & ((1u << ls.lc.seg_bits) - 1u);
    const unsigned b0 = bits & (0xffffffffu << ls.lc.seg_shift);
    float s0;
    std::memcpy(&s0, &b0, 4);
    const float t = s - s0;
    const float *c = &ls.table[(size_t)seg * 4];
    return std::fma(std::fma(std::fma(c[3], t, c[2]), t, c[1]), t, c[0]);
}

}  // namespace pp2p

namespace pp2p {

float law_linear_eval(const LawSetup &ls, float s)
{
    unsigned bits;
    std::memcpy(&bits, &s, 4);
    const unsigned byte = (bits >> (kLinSegShift - 3)) & 0x1fff8u;      // the kernel's LDS address
    const float *c = &ls.lin[(byte - (unsigned)ls.lc.lin_lo) / 4];
    return std::fma(c[1], s, c[0]);
}

}  // namespace pp2p
