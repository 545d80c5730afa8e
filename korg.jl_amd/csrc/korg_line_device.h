// Device-side physics of the line-opacity path (Float64), restating Korg.jl's
// src/line_absorption.jl with the same expressions and operation order, plus the Float32
// far-wing forms used by the mixed-precision kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace korg {

// src/constants.jl:3-19
constexpr double kboltz_cgs = 1.380649e-16;
constexpr double c_cgs = 2.99792458e10;
constexpr double electron_mass_cgs = 9.1093897e-28;
constexpr double electron_charge_cgs = 4.80320425e-10;
constexpr double amu_cgs = 1.6605402e-24;
constexpr double kboltz_eV = 8.617333262145e-5;
constexpr double hplanck_eV = 4.135667696e-15;
constexpr double hplanck_cgs = 6.62607015e-27;
constexpr double PI = 3.141592653589793;

// Wavelength offsets in the Float32 kernel are measured in this unit (cm): 0.01 A.
constexpr double DELTA_UNIT = 1e-10;

// Per-line result of the setup pass.  Indices are 0-based global grid indices; lb > ub marks a
// line whose window holds no grid point (skipped, src/line_absorption.jl:117-119).
struct LineMeta {
    double lam0;   // line centre, cm
    int32_t lb;    // first grid index in the window
    int32_t ub;    // last grid index in the window
    int32_t clo;   // grid-index range that may hold a Voigt-core evaluation in some layer
    int32_t chi;   //   (clo > chi: none)
    int32_t jc;    // first grid index with lambda >= lam0 (unclamped; monotone in lam0)
    int32_t cls;   // window-width class, or N_CLASSES when skipped
    int32_t flo;   // grid-index range where the far-wing series needs its R^3, R^4 terms in
    int32_t fhi;   //   some layer (flo > fhi: nowhere)
};
static_assert(sizeof(LineMeta) == 40, "LineMeta layout");

// A line as the profile kernels read it: the LineMeta fields they use plus the line's index,
// stored in partition order so that a batch of 64 candidates is one coalesced load.
struct Cand {
    double lam0;
    int32_t lb, ub, clo, chi, flo, fhi;
    int32_t line;
    int32_t pad;
};
static_assert(sizeof(Cand) == 40, "Cand layout");

// Window-width classes used to bin lines onto wavelength tiles: class c holds lines whose window
// has at most CLASS_WMAX[c] points.
constexpr int N_CLASSES = 6;
__host__ __device__ inline int64_t class_wmax(int c) {
    return c == 0 ? 64 : c == 1 ? 512 : c == 2 ? 4096 : c == 3 ? 32768 : c == 4 ? 262144
                                                                                : (int64_t)1 << 40;
}

// Float32 far-wing forms, per (line, layer):
//   a <= 0.2: phi = d1 R + d2 R^2 + d3 R^3,  R = 1/delta^2 (shared by all layers)
//   a >  0.2: phi = sum_{k=1..5} c_k rho^k,  rho = 1/(delta^2/g2 + 1)

// --------------------------------------------------------------------------------------------
// Float64 reference expressions (device restatement of src/line_absorption.jl)
// --------------------------------------------------------------------------------------------

// src/line_absorption.jl:140-145
__device__ inline double inverse_gaussian_density(double rho, double sigma) {
    if (rho > 1.0 / (sqrt(2.0 * PI) * sigma)) return 0.0;
    return sigma * sqrt(-2.0 * log(sqrt(2.0 * PI) * sigma * rho));
}

// src/line_absorption.jl:156-161
__device__ inline double inverse_lorentz_density(double rho, double gamma) {
    if (rho > 1.0 / (PI * gamma)) return 0.0;
    return sqrt(gamma / (PI * rho) - gamma * gamma);
}

// src/line_absorption.jl:190
__device__ inline double doppler_width(double lambda0, double T, double m, double xi) {
    return lambda0 * sqrt(kboltz_cgs * T / m + (xi * xi) / 2.0) / c_cgs;
}

// src/line_absorption.jl:208-221, ABO branch (the simple branch uses a per-layer table)
__device__ inline double scaled_vdW_abo(double sigma, double alpha, double m, double T) {
    const double v0 = 1e6;
    const double invmu = 1.0 / (1.008 * amu_cgs) + 1.0 / m;
    const double vbar = sqrt(8.0 * kboltz_cgs * T / PI * invmu);
    return 2.0 * pow(4.0 / PI, alpha / 2.0) * tgamma((4.0 - alpha) / 2.0) * v0 * sigma *
           pow(vbar / v0, 1.0 - alpha);
}

// src/line_absorption.jl:229-237
__device__ inline double sigma_line(double lambda) {
    const double e = electron_charge_cgs, me = electron_mass_cgs, c = c_cgs;
    return (PI * (e * e) / me / c) * ((lambda * lambda) / c);
}

// src/line_absorption.jl:251-265
__device__ inline double harris_series64(double v, double& H1, double& H2) {
    const double v2 = v * v;
    const double H0 = exp(-v2);
    if (v < 1.3)
        H1 = -1.12470432 + (-0.15516677 + (3.288675912 + (-2.34357915 + 0.42139162 * v) * v) * v) * v;
    else if (v < 2.4)
        H1 = -4.48480194 + (9.39456063 + (-6.61487486 + (1.98919585 - 0.22041650 * v) * v) * v) * v;
    else
        H1 = ((0.554153432 +
               (0.278711796 + (-0.1883256872 + (0.042991293 - 0.003278278 * v) * v) * v) * v) /
              (v2 - 3.0 / 2.0));
    H2 = (1.0 - 2.0 * v2) * H0;
    return H0;
}

// src/line_absorption.jl:286-310
__device__ inline double voigt_hjerting64(double a, double v) {
    const double sqrt_pi = 1.7724538509055159;  // sqrt(pi) in Float64
    const double v2 = v * v;
    if (a <= 0.2 && v >= 5.0) {
        const double invv2 = 1.0 / v2;
        return (a / sqrt_pi * invv2) * (1.0 + 1.5 * invv2 + 3.75 * (invv2 * invv2));
    } else if (a <= 0.2) {
        double H1, H2;
        const double H0 = harris_series64(v, H1, H2);
        return H0 + (H1 + H2 * a) * a;
    } else if (a <= 1.4 && a + v < 3.2) {
        double H1, H2;
        const double H0 = harris_series64(v, H1, H2);
        const double M0 = H0;
        const double M1 = H1 + 2.0 / sqrt_pi * M0;
        const double M2 = H2 - M0 + 2.0 / sqrt_pi * M1;
        const double M3 = 2.0 / (3.0 * sqrt_pi) * (1.0 - H2) - (2.0 / 3.0) * v2 * M1 + (2.0 / sqrt_pi) * M2;
        const double M4 = 2.0 / 3.0 * v2 * v2 * M0 - 2.0 / (3.0 * sqrt_pi) * M1 + 2.0 / sqrt_pi * M3;
        const double psi = 0.979895023 + (-0.962846325 + (0.532770573 - 0.122727278 * a) * a) * a;
        return psi * (M0 + (M1 + (M2 + (M3 + M4 * a) * a) * a) * a);
    } else {
        const double r2 = v2 / (a * a);
        const double a_invu = 0.7071067811865475 / ((r2 + 1.0) * a);  // 1/sqrt(2)/((r2+1)a)
        const double a2_invu2 = a_invu * a_invu;
        return 0.7978845608028654 * a_invu *  // sqrt(2/pi)
               (1.0 + (3.0 * r2 - 1.0 + ((r2 - 2.0) * 15.0 * r2 + 2.0) * a2_invu2) * a2_invu2);
    }
}

// --------------------------------------------------------------------------------------------
// Float32 Voigt core (branches B and C of voigt_hjerting), branch choices made by the caller
// in Float64.  piece: 0 (v < 1.3), 1 (v < 2.4), 2 (otherwise), decided on the Float64 v.
// --------------------------------------------------------------------------------------------
__device__ inline float harris_h1_f32(float v, float v2, int piece) {
    if (piece == 2)
        return (0.554153432f + (0.278711796f + (-0.1883256872f + (0.042991293f - 0.003278278f * v) * v) * v) * v) *
               __builtin_amdgcn_rcpf(v2 - 1.5f);
    // pieces 0 and 1 are both quartics: one Horner chain with the piece's coefficients (selected,
    // not branched on, so that lanes in different pieces do not diverge); the values are the same
    const bool p0 = piece == 0;
    const float c0 = p0 ? -1.12470432f : -4.48480194f, c1 = p0 ? -0.15516677f : 9.39456063f;
    const float c2 = p0 ? 3.288675912f : -6.61487486f, c3 = p0 ? -2.34357915f : 1.98919585f;
    const float c4 = p0 ? 0.42139162f : -0.22041650f;
    return c0 + (c1 + (c2 + (c3 + c4 * v) * v) * v) * v;
}

// Branch B (a <= 0.2, v < 5) or C (0.2 < a <= 1.4, a + v < 3.2) in Float32.
__device__ inline float voigt_core_f32(float a, float v, int piece, bool branch_c) {
    const float v2 = v * v;
    const float H0 = __expf(-v2);
    const float H1 = harris_h1_f32(v, v2, piece);
    const float H2 = (1.0f - 2.0f * v2) * H0;
    if (!branch_c) return H0 + (H1 + H2 * a) * a;
    const float isp = 0.5641895835477563f;  // 1/sqrt(pi)
    const float M0 = H0;
    const float M1 = H1 + 2.0f * isp * M0;
    const float M2 = H2 - M0 + 2.0f * isp * M1;
    const float M3 = (2.0f / 3.0f) * isp * (1.0f - H2) - (2.0f / 3.0f) * v2 * M1 + 2.0f * isp * M2;
    const float M4 = (2.0f / 3.0f) * v2 * v2 * M0 - (2.0f / 3.0f) * isp * M1 + 2.0f * isp * M3;
    const float psi = 0.979895023f + (-0.962846325f + (0.532770573f - 0.122727278f * a) * a) * a;
    return psi * (M0 + (M1 + (M2 + (M3 + M4 * a) * a) * a) * a);
}


// ---------------------------------------------------------------------------------------------
// Derivatives for the adjoint.  The partial derivatives in a and v of the Hunger branch chosen
// at (a, v) -- the derivative of voigt_hjerting as written (src/line_absorption.jl:286-310),
// which is what forward-mode differentiation of the reference yields.  Returns H.
// ---------------------------------------------------------------------------------------------
__device__ inline void harris_series64_dv(double v, double H0, double& dH0, double& dH1,
                                          double& dH2) {
    const double v2 = v * v;
    dH0 = -2.0 * v * H0;
    if (v < 1.3) {
        dH1 = -0.15516677 + (2.0 * 3.288675912 + (3.0 * -2.34357915 + 4.0 * 0.42139162 * v) * v) * v;
    } else if (v < 2.4) {
        dH1 = 9.39456063 + (2.0 * -6.61487486 + (3.0 * 1.98919585 - 4.0 * 0.22041650 * v) * v) * v;
    } else {
        const double P = 0.554153432 +
                         (0.278711796 + (-0.1883256872 + (0.042991293 - 0.003278278 * v) * v) * v) * v;
        const double dP = 0.278711796 +
                          (2.0 * -0.1883256872 + (3.0 * 0.042991293 - 4.0 * 0.003278278 * v) * v) * v;
        const double q = v2 - 1.5;
        dH1 = (dP * q - 2.0 * v * P) / (q * q);
    }
    dH2 = H0 * (-6.0 * v + 4.0 * v2 * v);
}

// The same partial derivatives evaluated in Float32, the branch and the piece of the Harris series
// chosen on the Float64 a and v (as the Float32 forward chooses them).  Returns H.
__device__ inline float voigt_hjerting_grad_f32(double a64, double v64, float& dHda, float& dHdv) {
    const float isp = 0.5641895835477563f;  // 1/sqrt(pi)
    const float a = (float)a64, v = (float)v64;
    const float v2 = v * v;
    if (a64 <= 0.2 && v64 >= 5.0) {
        const float w = __builtin_amdgcn_rcpf(v2);
        const float poly = 1.0f + 1.5f * w + 3.75f * (w * w);
        dHda = isp * w * poly;
        dHdv = a * isp * (1.0f + 3.0f * w + 11.25f * w * w) * (-2.0f * w * __builtin_amdgcn_rcpf(v));
        return a * dHda;
    }
    if (a64 <= 0.2 || (a64 <= 1.4 && a64 + v64 < 3.2)) {
        // The three pieces of the Harris series' H1 as one form: a quartic Q(v) (the piece's
        // polynomial; for v >= 2.4 the numerator of the rational piece) times iq (1, or
        // 1/(v^2 - 1.5) for v >= 2.4), so that lanes in different pieces take one path:
        // H1 = Q iq, dH1/dv = (Q' - 2 v H1 [v >= 2.4]) iq.
        const bool p0 = v64 < 1.3, p2 = !(v64 < 2.4);
        const float c0 = p0 ? -1.12470432f : (p2 ? 0.554153432f : -4.48480194f);
        const float c1 = p0 ? -0.15516677f : (p2 ? 0.278711796f : 9.39456063f);
        const float c2 = p0 ? 3.288675912f : (p2 ? -0.1883256872f : -6.61487486f);
        const float c3 = p0 ? -2.34357915f : (p2 ? 0.042991293f : 1.98919585f);
        const float c4 = p0 ? 0.42139162f : (p2 ? -0.003278278f : -0.22041650f);
        const float Q = c0 + (c1 + (c2 + (c3 + c4 * v) * v) * v) * v;
        const float dQ = c1 + (2.0f * c2 + (3.0f * c3 + 4.0f * c4 * v) * v) * v;
        const float iq = p2 ? __builtin_amdgcn_rcpf(v2 - 1.5f) : 1.0f;
        const float H0 = __expf(-v2);
        const float H1 = Q * iq;
        const float dH1 = (p2 ? dQ - 2.0f * v * H1 : dQ) * iq;
        const float H2 = (1.0f - 2.0f * v2) * H0;
        const float dH0 = -2.0f * v * H0;
        const float dH2 = H0 * (-6.0f * v + 4.0f * v2 * v);
        if (a64 <= 0.2) {
            dHda = H1 + 2.0f * H2 * a;
            dHdv = dH0 + (dH1 + dH2 * a) * a;
            return H0 + (H1 + H2 * a) * a;
        }
        const float k = 2.0f * isp, k3 = (2.0f / 3.0f) * isp;
        const float M0 = H0;
        const float M1 = H1 + k * M0;
        const float M2 = H2 - M0 + k * M1;
        const float M3 = k3 * (1.0f - H2) - (2.0f / 3.0f) * v2 * M1 + k * M2;
        const float M4 = (2.0f / 3.0f) * v2 * v2 * M0 - k3 * M1 + k * M3;
        const float dM0 = dH0;
        const float dM1 = dH1 + k * dM0;
        const float dM2 = dH2 - dM0 + k * dM1;
        const float dM3 = -k3 * dH2 - (2.0f / 3.0f) * (2.0f * v * M1 + v2 * dM1) + k * dM2;
        const float dM4 = (2.0f / 3.0f) * (4.0f * v2 * v * M0 + v2 * v2 * dM0) - k3 * dM1 + k * dM3;
        const float psi = 0.979895023f + (-0.962846325f + (0.532770573f - 0.122727278f * a) * a) * a;
        const float dpsi = -0.962846325f + (2.0f * 0.532770573f - 3.0f * 0.122727278f * a) * a;
        const float S = M0 + (M1 + (M2 + (M3 + M4 * a) * a) * a) * a;
        const float dSda = M1 + (2.0f * M2 + (3.0f * M3 + 4.0f * M4 * a) * a) * a;
        const float dSdv = dM0 + (dM1 + (dM2 + (dM3 + dM4 * a) * a) * a) * a;
        dHda = dpsi * S + psi * dSda;
        dHdv = psi * dSdv;
        return psi * S;
    }
    // branch D: H = sqrt(2/pi) u P(r^2, u), u = a / (sqrt(2) (v^2 + a^2)), r^2 = v^2 / a^2
    const float ia = __builtin_amdgcn_rcpf(a);
    const float r2 = v2 * ia * ia;
    const float s = v2 + a * a;
    const float is = __builtin_amdgcn_rcpf(s);
    const float u = 0.70710678118654752f * a * is;
    const float u2 = u * u;
    const float c4 = (r2 - 2.0f) * 15.0f * r2 + 2.0f;
    const float P = 1.0f + (3.0f * r2 - 1.0f + c4 * u2) * u2;
    const float iden = 0.70710678118654752f * is * is;  // 1 / (sqrt(2) (v^2 + a^2)^2)
    const float du_da = (v2 - a * a) * iden, du_dv = -2.0f * a * v * iden;
    const float dr2_da = -2.0f * r2 * ia, dr2_dv = 2.0f * v * ia * ia;
    const float dP_dr2 = 3.0f * u2 + (30.0f * r2 - 30.0f) * u2 * u2;
    const float dP_du = 2.0f * (3.0f * r2 - 1.0f) * u + 4.0f * c4 * u2 * u;
    const float s2p = 0.79788456080286536f;  // sqrt(2/pi)
    dHda = s2p * (P * du_da + u * (dP_dr2 * dr2_da + dP_du * du_da));
    dHdv = s2p * (P * du_dv + u * (dP_dr2 * dr2_dv + dP_du * du_dv));
    return s2p * u * P;
}

__device__ inline double voigt_hjerting64_grad(double a, double v, double& dHda, double& dHdv) {
    const double sqrt_pi = 1.7724538509055159;
    const double v2 = v * v;
    if (a <= 0.2 && v >= 5.0) {
        const double w = 1.0 / v2;
        const double poly = 1.0 + 1.5 * w + 3.75 * (w * w);
        dHda = w / sqrt_pi * poly;
        dHdv = a / sqrt_pi * (1.0 + 3.0 * w + 11.25 * w * w) * (-2.0 * w / v);
        return (a / sqrt_pi * w) * poly;
    } else if (a <= 0.2) {
        double H1, H2, dH0, dH1, dH2;
        const double H0 = harris_series64(v, H1, H2);
        harris_series64_dv(v, H0, dH0, dH1, dH2);
        dHda = H1 + 2.0 * H2 * a;
        dHdv = dH0 + (dH1 + dH2 * a) * a;
        return H0 + (H1 + H2 * a) * a;
    } else if (a <= 1.4 && a + v < 3.2) {
        double H1, H2, dH0, dH1, dH2;
        const double H0 = harris_series64(v, H1, H2);
        harris_series64_dv(v, H0, dH0, dH1, dH2);
        const double k = 2.0 / sqrt_pi, k3 = 2.0 / (3.0 * sqrt_pi);
        const double M0 = H0;
        const double M1 = H1 + k * M0;
        const double M2 = H2 - M0 + k * M1;
        const double M3 = k3 * (1.0 - H2) - (2.0 / 3.0) * v2 * M1 + k * M2;
        const double M4 = 2.0 / 3.0 * v2 * v2 * M0 - k3 * M1 + k * M3;
        const double dM0 = dH0;
        const double dM1 = dH1 + k * dM0;
        const double dM2 = dH2 - dM0 + k * dM1;
        const double dM3 = -k3 * dH2 - (2.0 / 3.0) * (2.0 * v * M1 + v2 * dM1) + k * dM2;
        const double dM4 = 2.0 / 3.0 * (4.0 * v2 * v * M0 + v2 * v2 * dM0) - k3 * dM1 + k * dM3;
        const double psi = 0.979895023 + (-0.962846325 + (0.532770573 - 0.122727278 * a) * a) * a;
        const double dpsi = -0.962846325 + (2.0 * 0.532770573 - 3.0 * 0.122727278 * a) * a;
        const double S = M0 + (M1 + (M2 + (M3 + M4 * a) * a) * a) * a;
        const double dSda = M1 + (2.0 * M2 + (3.0 * M3 + 4.0 * M4 * a) * a) * a;
        const double dSdv = dM0 + (dM1 + (dM2 + (dM3 + dM4 * a) * a) * a) * a;
        dHda = dpsi * S + psi * dSda;
        dHdv = psi * dSdv;
        return psi * S;
    } else {
        // u = a / (sqrt(2) (v^2 + a^2)), r^2 = v^2 / a^2; H = sqrt(2/pi) u P(r^2, u)
        const double r2 = v2 / (a * a);
        const double u = 1.0 / 1.4142135623730951 / ((r2 + 1.0) * a);
        const double u2 = u * u;
        const double c4 = (r2 - 2.0) * 15.0 * r2 + 2.0;
        const double P = 1.0 + (3.0 * r2 - 1.0 + c4 * u2) * u2;
        const double den = 1.4142135623730951 * (v2 + a * a) * (v2 + a * a);
        const double du_da = (v2 - a * a) / den, du_dv = -2.0 * a * v / den;
        const double dr2_da = -2.0 * v2 / (a * a * a), dr2_dv = 2.0 * v / (a * a);
        const double dP_dr2 = 3.0 * u2 + (30.0 * r2 - 30.0) * u2 * u2;
        const double dP_du = 2.0 * (3.0 * r2 - 1.0) * u + 4.0 * c4 * u2 * u;
        const double s2p = 0.79788456080286536;  // sqrt(2/pi)
        dHda = s2p * (P * du_da + u * (dP_dr2 * dr2_da + dP_du * du_da));
        dHdv = s2p * (P * du_dv + u * (dP_dr2 * dr2_dv + dP_du * du_dv));
        return s2p * u * P;
    }
}

}  // namespace korg
