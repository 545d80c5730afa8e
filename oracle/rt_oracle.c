/*
 * Float64 CPU restatement of Korg.jl's default radiative transfer for a plane-parallel atmosphere
 * (test infrastructure only: the checker of the device radiative transfer; nothing in the product
 * links it).
 *
 * The path restated is what `radiative_transfer(atm::PlanarAtmosphere, α, S, μ_points; α_ref,
 * τ_scheme="anchored", I_scheme="linear_flux_only")` runs
 * (src/RadiativeTransfer/RadiativeTransfer.jl:62-66, 77-146): with the anchored optical depth
 * and the flux-only linear source it switches to the exponential-integral form (:81-84), a single
 * outward ray μ = 1 with weight 1 and no inward rays, so F = 2π I_surface (:144), where
 *   τ_1 = 0, τ_i = τ_{i-1} + ½ (α_i f_i + α_{i-1} f_{i-1}) (ln τref_i − ln τref_{i-1}),
 *   f_i = τref_i / αref_i                                          (:165-167, 244-253)
 *   I = Σ_i [core(τ_{i+1}, m_i, b_i) − core(τ_i, m_i, b_i)],
 *   m_i = (S_{i+1} − S_i)/(τ_{i+1} − τ_i), b_i = S_i − m_i τ_i        (:379-389)
 *   core(τ, m, b) = (τ E2(τ)(3b + 2mτ) − e^{−τ}(3b + 2m(τ + 1)))/6   (:398-400)
 * with Korg's piecewise approximation of E2 (:408-480) and the source S = blackbody(T, λ)
 * (src/synthesize.jl:285, 412-417).
 *
 * Also the reverse-mode derivative dF/dα (per layer and wavelength) given dL/dF, written by hand
 * from the same expressions (E2 differentiated as approximated).
 *
 * Pins (tests/test_rt.py): E2 against the exact E2 at rtol 1e-3 on [2, 8] (the reference's own
 * test/transfer.jl:4-7) and within 1% everywhere (its docstring claim, :405-407); the derivative
 * of each E2 piece against central differences; with α = αref the anchored τ is the trapezoid
 * integral of τref over ln τref, which converges to τref − τref_1 on a fine grid; a source linear
 * in τ on a deep atmosphere gives F = 2π(b/2 + m/3) (∫E2 = 1/2, ∫τE2 = 1/3) to the accuracy of
 * the E2 approximation; the reverse mode against central differences of F.
 */
#include <math.h>
#include <stdint.h>

/* src/constants.jl:3-19 */
static const double kboltz_cgs = 1.380649e-16;
static const double hplanck_cgs = 6.62607015e-27;
static const double c_cgs = 2.99792458e10;
static const double PI = 3.141592653589793;

/* src/synthesize.jl:412-417 */
double oracle_blackbody(double T, double lam) {
    const double h = hplanck_cgs, c = c_cgs, k = kboltz_cgs;
    return 2 * h * c * c / pow(lam, 5) * 1 / expm1(h * c / lam / k / T);
}

/* Korg's E2 pieces (src/RadiativeTransfer/RadiativeTransfer.jl:427-480): polynomials of degree 5
 * in y = x - n (n = 2..8), their coefficients as published there. */
static const double E2_POLY[7][6] = {
    {0.037534261820486914, -0.04890051070806112, 0.033833820809153176, -0.016916910404576574,
     0.007048712668573576, -0.0026785108140579598},
    {0.010641925085272673, -0.013048381094197039, 0.008297844727977323, -0.003687930990212144,
     0.0013061422257001345, -0.0003995258572729822},
    {0.0031982292493385146, -0.0037793524098489054, 0.0022894548610917728, -0.0009539395254549051,
     0.00031003034577284415, -8.466213288412284e-5},
    {0.000996469042708825, -0.0011482955912753257, 0.0006737946999085467, -0.00026951787996341863,
     8.310134632205409e-5, -2.1202073223788938e-5},
    {0.0003182574636904001, -0.0003600824521626587, 0.00020656268138886323, -8.032993165122457e-5,
     2.390771775334065e-5, -5.8334831318151185e-6},
    {0.00010350984428214624, -0.00011548173161033826, 6.513442611103688e-5, -2.4813114708966427e-5,
     7.200234178941151e-6, -1.7027366981408086e-6},
    {3.413764515111217e-5, -3.76656228439249e-5, 2.096641424390699e-5, -7.862405341465122e-6,
     2.2386015208338193e-6, -5.173353514609864e-7},
};

/* Piece of x (:408-426): -1 x == 0, 0 small, 1..7 the polynomials about 2..8, 8 large. */
static int e2_piece(double x) {
    if (x == 0) return -1;
    if (x < 1.1) return 0;
    if (x < 2.5) return 1;
    if (x < 3.5) return 2;
    if (x < 4.5) return 3;
    if (x < 5.5) return 4;
    if (x < 6.5) return 5;
    if (x < 7.5) return 6;
    if (x < 9) return 7;
    return 8;
}

/* exponential_integral_2 (:408-480) and its derivative (the derivative of the approximation). */
double oracle_expint2(double x, double* dE2) {
    const int p = e2_piece(x);
    double e, d;
    if (p < 0) {
        e = 1.0;
        d = -INFINITY;
    } else if (p == 0) {
        /* 1 + x Q(x), Q = (ln x + γ − 1) + x P(x), P = −1/2 + x(1/12 + x(−1/72 + x/480)) */
        const double g = 0.57721566490153286060651209008240243104215933593992;
        const double P = -0.5 + (0.08333333333333333 + (-0.013888888888888888 +
                                                        0.0020833333333333333 * x) * x) * x;
        const double dP = 0.08333333333333333 + (2 * -0.013888888888888888 +
                                                 3 * 0.0020833333333333333 * x) * x;
        const double Q = (log(x) + g - 1) + P * x;
        e = 1 + Q * x;
        d = Q + x * (1 / x + P + x * dP);
    } else if (p == 8) {
        /* e^{-x} u R(u), u = 1/x, R = 1 + u(−2 + u(6 + u(−24 + 120u))) */
        const double u = 1 / x;
        const double R = 1 + (-2 + (6 + (-24 + 120 * u) * u) * u) * u;
        const double dR = -2 + (12 + (-72 + 480 * u) * u) * u;
        const double ex = exp(-x);
        e = ex * R * u;
        d = -e + ex * (R + u * dR) * (-u * u);
    } else {
        const double* c = E2_POLY[p - 1];
        const double y = x - (p + 1);
        e = c[0] + (c[1] + (c[2] + (c[3] + (c[4] + c[5] * y) * y) * y) * y) * y;
        d = c[1] + (2 * c[2] + (3 * c[3] + (4 * c[4] + 5 * c[5] * y) * y) * y) * y;
    }
    if (dE2) *dE2 = d;
    return e;
}

/* expint_transfer_integral_core (:398-400) and its partial derivatives in τ, m, b. */
double oracle_transfer_core(double tau, double m, double b, double* dtau, double* dm, double* db) {
    double dE;
    const double E = oracle_expint2(tau, &dE);
    const double ex = exp(-tau);
    const double A = 3 * b + 2 * m * tau, B = 3 * b + 2 * m * (tau + 1);
    if (dtau) *dtau = ((E + tau * dE) * A + tau * E * 2 * m + ex * B - ex * 2 * m) / 6;
    if (dm) *dm = (tau * E * 2 * tau - ex * 2 * (tau + 1)) / 6;
    if (db) *db = (tau * E * 3 - ex * 3) / 6;
    return (tau * E * A - ex * B) / 6;
}

/* Anchored optical depth of one wavelength column (:244-253). */
static void tau_anchored(int64_t L, const double* alpha, const double* f, const double* ltau,
                         double* tau) {
    tau[0] = 0.0;
    for (int64_t i = 1; i < L; ++i)
        tau[i] = tau[i - 1] + 0.5 * (alpha[i] * f[i] + alpha[i - 1] * f[i - 1]) * (ltau[i] - ltau[i - 1]);
}

/* F of one column: 2π compute_F_flux_only_expint (:379-389, 144). */
static double flux_column(int64_t L, const double* tau, const double* S) {
    double I = 0;
    for (int64_t i = 0; i + 1 < L; ++i) {
        const double m = (S[i + 1] - S[i]) / (tau[i + 1] - tau[i]);
        const double b = S[i] - m * tau[i];
        I += oracle_transfer_core(tau[i + 1], m, b, 0, 0, 0) - oracle_transfer_core(tau[i], m, b, 0, 0, 0);
    }
    return 2 * PI * I;
}

/*
 * Astrophysical flux F[j] of alpha (column-major [L x nwl]) with the source the blackbody of each
 * layer's temperature at wls[j] (cm), reference optical depth tau_ref[L] and absorption alpha_ref[L]
 * (both of the reference wavelength).  Layers from the surface down.
 */
int oracle_rt_flux(int64_t L, int64_t nwl, const double* alpha, const double* wls,
                   const double* temps, const double* tau_ref, const double* alpha_ref, double* F) {
    if (L < 2 || L > 4096) return 1;
    double f[4096], ltau[4096], tau[4096], S[4096];
    for (int64_t i = 0; i < L; ++i) {
        f[i] = tau_ref[i] / alpha_ref[i];
        ltau[i] = log(tau_ref[i]);
    }
    for (int64_t j = 0; j < nwl; ++j) {
        const double* a = alpha + j * L;
        for (int64_t i = 0; i < L; ++i) S[i] = oracle_blackbody(temps[i], wls[j]);
        tau_anchored(L, a, f, ltau, tau);
        F[j] = flux_column(L, tau, S);
    }
    return 0;
}

/*
 * Reverse mode of oracle_rt_flux: given gF[j] = dL/dF[j], writes galpha[l + L j] = dL/dalpha[l, j].
 */
int oracle_rt_flux_vjp(int64_t L, int64_t nwl, const double* alpha, const double* wls,
                       const double* temps, const double* tau_ref, const double* alpha_ref,
                       const double* gF, double* galpha) {
    if (L < 2 || L > 4096) return 1;
    double f[4096], ltau[4096], tau[4096], S[4096], gtau[4096];
    for (int64_t i = 0; i < L; ++i) {
        f[i] = tau_ref[i] / alpha_ref[i];
        ltau[i] = log(tau_ref[i]);
    }
    for (int64_t j = 0; j < nwl; ++j) {
        const double* a = alpha + j * L;
        double* ga = galpha + j * L;
        for (int64_t i = 0; i < L; ++i) S[i] = oracle_blackbody(temps[i], wls[j]);
        tau_anchored(L, a, f, ltau, tau);
        for (int64_t i = 0; i < L; ++i) gtau[i] = 0.0;
        const double w = 2 * PI * gF[j];
        for (int64_t i = 0; i + 1 < L; ++i) {
            const double dt = tau[i + 1] - tau[i];
            const double m = (S[i + 1] - S[i]) / dt;
            const double b = S[i] - m * tau[i];
            double ct1, cm1, cb1, ct0, cm0, cb0;
            oracle_transfer_core(tau[i + 1], m, b, &ct1, &cm1, &cb1);
            oracle_transfer_core(tau[i], m, b, &ct0, &cm0, &cb0);
            const double Tm = cm1 - cm0, Tb = cb1 - cb0;  /* d term / d m, d term / d b */
            /* m = (S1 − S0)/(τ1 − τ0), b = S0 − m τ0 */
            const double dm_dt1 = -m / dt, dm_dt0 = m / dt;
            const double db_dt1 = -tau[i] * dm_dt1, db_dt0 = -m - tau[i] * dm_dt0;
            gtau[i + 1] += w * (ct1 + Tm * dm_dt1 + Tb * db_dt1);
            gtau[i] += w * (-ct0 + Tm * dm_dt0 + Tb * db_dt0);
        }
        /* τ_k = τ_{k-1} + ½(u_k + u_{k-1}) Δ_k with u = α f: dτ_k/du_j = ½Δ_j [2 ≤ j ≤ k] +
         * ½Δ_{j+1} [j + 1 ≤ k] (0-based: Δ_k = ltau[k] − ltau[k−1] for k ≥ 1) */
        double suffix = 0.0;  /* Σ_{k' ≥ k} gtau[k'] */
        double suffix_next = 0.0;
        for (int64_t k = L - 1; k >= 0; --k) {
            suffix_next = suffix;
            suffix += gtau[k];
            double gu = 0.0;
            if (k >= 1) gu += 0.5 * (ltau[k] - ltau[k - 1]) * suffix;
            if (k + 1 < L) gu += 0.5 * (ltau[k + 1] - ltau[k]) * suffix_next;
            ga[k] = gu * f[k];
        }
    }
    return 0;
}
