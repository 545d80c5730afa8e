/*
 * ORACLE -- test infrastructure only.
 *
 * Float64 CPU restatement of Korg.jl's line-opacity path, used by the test-suite and by the
 * benchmark's CPU-baseline leg as the checker for the device library.  Nothing in the product
 * links or calls this file.
 *
 * It follows, expression by expression (same IEEE operation order, no FMA contraction):
 *   src/line_absorption.jl:33-129   line_absorption!            -> oracle_line_absorption
 *   src/line_absorption.jl:140-145  inverse_gaussian_density    -> oracle_inverse_gaussian_density
 *   src/line_absorption.jl:156-161  inverse_lorentz_density     -> oracle_inverse_lorentz_density
 *   src/line_absorption.jl:190      doppler_width               -> oracle_doppler_width
 *   src/line_absorption.jl:195      scaled_stark                -> oracle_scaled_stark
 *   src/line_absorption.jl:208-221  scaled_vdW                  -> oracle_scaled_vdW
 *   src/line_absorption.jl:229-237  sigma_line                  -> oracle_sigma_line
 *   src/line_absorption.jl:245-249  line_profile                -> oracle_line_profile
 *   src/line_absorption.jl:251-265  harris_series               -> oracle_harris_series
 *   src/line_absorption.jl:286-310  voigt_hjerting              -> oracle_voigt_hjerting
 *   src/wavelengths.jl:188-210      searchsortedfirst/last      -> oracle_searchsorted_first/last
 * with the constants of src/constants.jl:3-19 and the masses of src/atomic_data.jl:16-26.
 *
 * Threading: the reference spawns one task per line chunk, each with a private alpha that is
 * summed afterwards (src/line_absorption.jl:62-69,127).  oracle_line_absorption offers that
 * schedule (mode ORACLE_KORG_CHUNKS, used for timing) and a deterministic schedule
 * (ORACLE_DETERMINISTIC) whose result equals the reference run on one thread: every alpha
 * element receives its lines in list order into a zeroed accumulator, which is then added to
 * alpha once.
 *
 * Transcendentals come from the C library; Julia's own exp/log/pow may differ in the last ulp.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   line_profile            the reference's own assertions (test/line_absorption.jl:4-17)
 *   searchsorted_first/last the reference's known indices (test/wavelengths.jl:60-86)
 *   voigt_hjerting          a = 0 Gaussian and large-a Lorentzian limits; distance from the exact
 *                           Voigt function Re w(v + ia) bounded per Hunger branch
 *   harris_series           H1 against its Dawson-function form; H2 exactly
 *   inverse_*_density       pdf(inverse(rho)) == rho
 *   doppler_width           most-probable-speed known value; microturbulence in quadrature
 *   scaled_stark            the (T/1e4)^(1/6) law
 *   scaled_vdW              simple: (T/1e4)^0.3; ABO: twice the Maxwell average of
 *                           sigma0 (v/v0)^-alpha by quadrature, for several alpha and T
 *   sigma_line              the classical cross-section pi e^2 / (m_e c)
 *   cntm_interp             exact interpolant of a quadratic continuum (same windows as the
 *                           values at line centres)
 *   line_absorption         line integral == amplitude, Doppler and Lorentz FWHMs, both thread
 *                           schedules agree, accumulation, empty list, molecules
 * Parity unpinned: the last ulp of Julia's libm (no reference test pins values below 1e-7
 * relative), and voigt_hjerting branch C beyond Hunger's own accuracy (its coefficients are
 * checked only through the bounded distance from the exact Voigt function and the reference's
 * normalisation test).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/constants.jl:3-19 */
static const double kboltz_cgs = 1.380649e-16;
static const double c_cgs = 2.99792458e10;
static const double electron_mass_cgs = 9.1093897e-28;
static const double electron_charge_cgs = 4.80320425e-10;
static const double amu_cgs = 1.6605402e-24;
static const double kboltz_eV = 8.617333262145e-5;
static const double hplanck_eV = 4.135667696e-15;

enum { ORACLE_DETERMINISTIC = 0, ORACLE_KORG_CHUNKS = 1, ORACLE_WINDOWS_ONLY = 2 };
enum { ORACLE_CNTM_AT_LINES = 0, ORACLE_CNTM_GRID = 1 };

/* src/line_absorption.jl:140-145 */
double oracle_inverse_gaussian_density(double rho, double sigma) {
    if (rho > 1.0 / (sqrt(2.0 * M_PI) * sigma)) return 0.0;
    return sigma * sqrt(-2.0 * log(sqrt(2.0 * M_PI) * sigma * rho));
}

/* src/line_absorption.jl:156-161 */
double oracle_inverse_lorentz_density(double rho, double gamma) {
    if (rho > 1.0 / (M_PI * gamma)) return 0.0;
    return sqrt(gamma / (M_PI * rho) - gamma * gamma);
}

/* src/line_absorption.jl:190 */
double oracle_doppler_width(double lambda0, double T, double m, double xi) {
    return lambda0 * sqrt(kboltz_cgs * T / m + (xi * xi) / 2.0) / c_cgs;
}

/* src/line_absorption.jl:195 */
double oracle_scaled_stark(double gamma_stark, double T) {
    return gamma_stark * pow(T / 10000.0, 1.0 / 6.0);
}

/* src/line_absorption.jl:208-221 */
double oracle_scaled_vdW(double vdW1, double vdW2, double m, double T) {
    if (vdW2 == -1.0) return vdW1 * pow(T / 10000.0, 0.3);
    const double v0 = 1e6;
    const double sigma = vdW1, alpha = vdW2;
    const double invmu = 1.0 / (1.008 * amu_cgs) + 1.0 / m;
    const double vbar = sqrt(8.0 * kboltz_cgs * T / M_PI * invmu);
    return 2.0 * pow(4.0 / M_PI, alpha / 2.0) * tgamma((4.0 - alpha) / 2.0) * v0 * sigma *
           pow(vbar / v0, 1.0 - alpha);
}

/* src/line_absorption.jl:229-237 */
double oracle_sigma_line(double lambda) {
    const double e = electron_charge_cgs, me = electron_mass_cgs, c = c_cgs;
    return (M_PI * (e * e) / me / c) * ((lambda * lambda) / c);
}

/* src/line_absorption.jl:251-265; returns H0, writes H1, H2 */
double oracle_harris_series(double v, double* H1_out, double* H2_out) {
    const double v2 = v * v;
    const double H0 = exp(-v2);
    double H1;
    if (v < 1.3)
        H1 = -1.12470432 + (-0.15516677 + (3.288675912 + (-2.34357915 + 0.42139162 * v) * v) * v) * v;
    else if (v < 2.4)
        H1 = -4.48480194 + (9.39456063 + (-6.61487486 + (1.98919585 - 0.22041650 * v) * v) * v) * v;
    else
        H1 = ((0.554153432 +
               (0.278711796 + (-0.1883256872 + (0.042991293 - 0.003278278 * v) * v) * v) * v) /
              (v2 - 3.0 / 2.0));
    *H1_out = H1;
    *H2_out = (1.0 - 2.0 * v2) * H0;
    return H0;
}

/* src/line_absorption.jl:286-310 */
double oracle_voigt_hjerting(double a, double v) {
    const double sqrt_pi = sqrt(M_PI);
    const double v2 = v * v;
    if (a <= 0.2 && v >= 5.0) {
        const double invv2 = 1.0 / v2;
        return (a / sqrt_pi * invv2) * (1.0 + 1.5 * invv2 + 3.75 * (invv2 * invv2));
    } else if (a <= 0.2) {
        double H1, H2;
        const double H0 = oracle_harris_series(v, &H1, &H2);
        return H0 + (H1 + H2 * a) * a;
    } else if (a <= 1.4 && a + v < 3.2) {
        double H1, H2;
        const double H0 = oracle_harris_series(v, &H1, &H2);
        const double M0 = H0;
        const double M1 = H1 + 2.0 / sqrt_pi * M0;
        const double M2 = H2 - M0 + 2.0 / sqrt_pi * M1;
        const double M3 = 2.0 / (3.0 * sqrt_pi) * (1.0 - H2) - (2.0 / 3.0) * v2 * M1 + (2.0 / sqrt_pi) * M2;
        const double M4 = 2.0 / 3.0 * v2 * v2 * M0 - 2.0 / (3.0 * sqrt_pi) * M1 + 2.0 / sqrt_pi * M3;
        const double psi = 0.979895023 + (-0.962846325 + (0.532770573 - 0.122727278 * a) * a) * a;
        return psi * (M0 + (M1 + (M2 + (M3 + M4 * a) * a) * a) * a);
    } else {
        const double r2 = v2 / (a * a);
        const double a_invu = 1.0 / sqrt(2.0) / ((r2 + 1.0) * a);
        const double a2_invu2 = a_invu * a_invu;
        return sqrt(2.0 / M_PI) * a_invu *
               (1.0 + (3.0 * r2 - 1.0 + ((r2 - 2.0) * 15.0 * r2 + 2.0) * a2_invu2) * a2_invu2);
    }
}

/* src/line_absorption.jl:245-249 */
double oracle_line_profile(double lambda0, double sigma, double gamma, double amplitude,
                           double lambda) {
    const double inv_sigma_sqrt2 = 1.0 / (sigma * sqrt(2.0));
    const double scaling = inv_sigma_sqrt2 / sqrt(M_PI) * amplitude;
    return oracle_voigt_hjerting(gamma * inv_sigma_sqrt2, fabs(lambda - lambda0) * inv_sigma_sqrt2) *
           scaling;
}

/* src/wavelengths.jl:188-198: 1-based index of the first grid point >= lam (n+1 if none).
 * A value >= 1 is read as Angstrom.  NaN compares false everywhere, giving n+1. */
int64_t oracle_searchsorted_first(const double* wls, int64_t n, double lam) {
    if (lam >= 1.0) lam = lam * 1e-8;
    if (!(lam <= wls[n - 1])) return n + 1;
    int64_t lo = 0, hi = n - 1; /* first index with wls[i] >= lam lies in [lo, hi] */
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (wls[mid] < lam) lo = mid + 1; else hi = mid;
    }
    return lo + 1;
}

/* src/wavelengths.jl:200-210: 1-based index of the last grid point <= lam (0 if none). */
int64_t oracle_searchsorted_last(const double* wls, int64_t n, double lam) {
    if (lam >= 1.0) lam = lam * 1e-8;
    if (!(lam <= wls[n - 1])) return n;
    int64_t lo = 0, hi = n; /* count of points <= lam lies in [lo, hi] */
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (wls[mid] <= lam) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* Julia's maximum(): NaN if any element is NaN. */
static double max_nan(const double* x, int64_t n) {
    double m = x[0];
    for (int64_t i = 0; i < n; ++i) {
        if (isnan(x[i])) return x[i];
        if (x[i] > m) m = x[i];
    }
    return m;
}

/* Linear interpolation of one layer's continuum on the coarse grid at x (knots ascending). */
static double cntm_interp(const double* knots, const double* vals, int64_t nk, double x) {
    int64_t lo = 0, hi = nk; /* count of knots <= x */
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (knots[mid] <= x) lo = mid + 1; else hi = mid;
    }
    int64_t i = lo - 1;
    if (i < 0) i = 0;
    if (i > nk - 2) i = nk - 2;
    const double t = (x - knots[i]) / (knots[i + 1] - knots[i]);
    return (1.0 - t) * vals[i] + t * vals[i + 1];
}

typedef struct {
    int64_t L, nwl, nlines;
    const double *wls;
    const double *wl, *log_gf, *E_lower, *gamma_rad, *gamma_stark, *vdW1, *vdW2;
    const int32_t* species_index;
    const double* species_mass;
    const uint8_t* species_is_mol;
    const double* n_div_U;
    const double *temps, *n_e, *xi;
    int cntm_mode;
    const double *cntm_values, *knots;
    int64_t n_knots;
    double cutoff;
    double *beta, *n_eff;
} problem_t;

/* Per-line setup of src/line_absorption.jl:82-119 for one line: per-layer sigma, gamma,
 * amplitude and the 1-based window [lb, ub].  Scratch arrays have length L. */
static void line_setup(const problem_t* P, int64_t i, double* sigma, double* gamma, double* amp,
                       double* rho, double* inv, int64_t* lb, int64_t* ub) {
    const int64_t L = P->L;
    const int32_t s = P->species_index[i];
    const double m = P->species_mass[s];
    const double wl = P->wl[i];
    for (int64_t l = 0; l < L; ++l) sigma[l] = oracle_doppler_width(wl, P->temps[l], m, P->xi[l]);
    for (int64_t l = 0; l < L; ++l) {
        double G = P->gamma_rad[i];
        if (!P->species_is_mol[s]) {
            G = G + P->n_e[l] * oracle_scaled_stark(P->gamma_stark[i], P->temps[l]);
            G = G + P->n_eff[l] * oracle_scaled_vdW(P->vdW1[i], P->vdW2[i], m, P->temps[l]);
        }
        gamma[l] = G * (wl * wl) / (c_cgs * (4.0 * M_PI));
    }
    const double E_upper = P->E_lower[i] + c_cgs * hplanck_eV / wl;
    const double gf = pow(10.0, P->log_gf[i]);
    const double sl = oracle_sigma_line(wl);
    for (int64_t l = 0; l < L; ++l) {
        const double levels = exp(-P->beta[l] * P->E_lower[i]) - exp(-P->beta[l] * E_upper);
        amp[l] = gf * sl * levels * P->n_div_U[(int64_t)s * L + l];
    }
    for (int64_t l = 0; l < L; ++l) {
        double c;
        if (P->cntm_mode == ORACLE_CNTM_AT_LINES)
            c = P->cntm_values[i * L + l];
        else
            c = cntm_interp(P->knots, P->cntm_values + l * P->n_knots, P->n_knots, wl);
        rho[l] = c * P->cutoff / amp[l];
    }
    for (int64_t l = 0; l < L; ++l) inv[l] = oracle_inverse_gaussian_density(rho[l], sigma[l]);
    const double doppler_window = max_nan(inv, L);
    for (int64_t l = 0; l < L; ++l) inv[l] = oracle_inverse_lorentz_density(rho[l], gamma[l]);
    const double lorentz_window = max_nan(inv, L);
    const double window = sqrt(lorentz_window * lorentz_window + doppler_window * doppler_window);
    *lb = oracle_searchsorted_first(P->wls, P->nwl, wl - window);
    *ub = oracle_searchsorted_last(P->wls, P->nwl, wl + window);
}

/*
 * alpha: column-major [L x nwl] (alpha[l + L*j]), accumulated into (src/line_absorption.jl:127).
 * n_div_U: [n_species x L] row-major, n_species(l)/U_species(ln T_l) (src/line_absorption.jl:42-44).
 * n_HI, n_HeI, n_H2: perturber densities for n_eff_vdW (src/line_absorption.jl:54-60).
 * xi: [L] microturbulence in cm/s.
 * cntm_mode 0: cntm_values[i*L + l] is alpha_cntm of layer l at line i's wavelength;
 *           1: cntm_values[l*n_knots + k] on knots[k] (linear interpolation).
 * lb_out/ub_out (optional): the 1-based windows, lb > ub for skipped lines.
 * n_evals_out (optional): sum over lines of L*(ub-lb+1) for lb <= ub.
 * lb_in/ub_in (optional, both or neither): frozen 1-based windows used instead of the computed
 *   ones (clipped to the grid), the windows a ForwardDiff evaluation holds constant.
 * Returns 0.
 */
int oracle_line_absorption(double* alpha, int64_t L, int64_t nwl, const double* wls,
                           int64_t nlines, const double* wl, const double* log_gf,
                           const double* E_lower, const double* gamma_rad,
                           const double* gamma_stark, const double* vdW1, const double* vdW2,
                           const int32_t* species_index, const double* species_mass,
                           const uint8_t* species_is_mol, const double* n_div_U,
                           const double* temps, const double* n_e, const double* n_HI,
                           const double* n_HeI, const double* n_H2, const double* xi,
                           int cntm_mode, const double* cntm_values, int64_t n_knots,
                           const double* knots, double cutoff_threshold, int nthreads, int mode,
                           int64_t* lb_out, int64_t* ub_out, double* n_evals_out,
                           const int64_t* lb_in, const int64_t* ub_in) {
    if (nlines == 0) return 0; /* src/line_absorption.jl:35-37 */
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    problem_t P = {L, nwl, nlines, wls, wl, log_gf, E_lower, gamma_rad, gamma_stark, vdW1, vdW2,
                   species_index, species_mass, species_is_mol, n_div_U, temps, n_e, xi,
                   cntm_mode, cntm_values, knots, n_knots, cutoff_threshold, NULL, NULL};
    double* beta = malloc(sizeof(double) * L);
    double* n_eff = malloc(sizeof(double) * L);
    /* src/line_absorption.jl:39, 54-60 */
    const double a_H = 4.50711, a_He = 1.38375, a_H2 = 5.503;
    const double m_H = 1.008 * amu_cgs, m_He = 4.003 * amu_cgs;
    const double c_He = pow(a_He / a_H, 0.4) * pow(m_He / m_H, -0.3);
    const double c_H2 = pow(a_H2 / a_H, 0.4) * pow(2.0, -0.3);
    for (int64_t l = 0; l < L; ++l) {
        beta[l] = 1.0 / (kboltz_eV * temps[l]);
        n_eff[l] = n_HI[l] + c_He * n_HeI[l] + c_H2 * n_H2[l];
    }
    P.beta = beta;
    P.n_eff = n_eff;

    /* per-line setup: sigma, gamma, amplitude per layer and the window */
    double* sig = malloc(sizeof(double) * nlines * L);
    double* gam = malloc(sizeof(double) * nlines * L);
    double* amp = malloc(sizeof(double) * nlines * L);
    int64_t* lb = malloc(sizeof(int64_t) * nlines);
    int64_t* ub = malloc(sizeof(int64_t) * nlines);
#pragma omp parallel num_threads(nthreads)
    {
        double* rho = malloc(sizeof(double) * L);
        double* inv = malloc(sizeof(double) * L);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nlines; ++i)
            line_setup(&P, i, sig + i * L, gam + i * L, amp + i * L, rho, inv, lb + i, ub + i);
        free(rho);
        free(inv);
    }
    if (lb_in && ub_in) {
        for (int64_t i = 0; i < nlines; ++i) {
            lb[i] = lb_in[i] < 1 ? 1 : lb_in[i];
            ub[i] = ub_in[i] > nwl ? nwl : ub_in[i];
        }
    }
    double nev = 0.0;
    for (int64_t i = 0; i < nlines; ++i)
        if (lb[i] <= ub[i]) nev += (double)L * (double)(ub[i] - lb[i] + 1);
    if (n_evals_out) *n_evals_out = nev;
    if (lb_out) memcpy(lb_out, lb, sizeof(int64_t) * nlines);
    if (ub_out) memcpy(ub_out, ub, sizeof(int64_t) * nlines);

    if (mode == ORACLE_WINDOWS_ONLY) {
        /* windows and evaluation count only; alpha is left untouched */
    } else if (mode == ORACLE_DETERMINISTIC) {
        /* each thread owns a contiguous block of wavelengths (all layers) and adds the lines in
         * list order into a zeroed accumulator: identical to the one-task reference run */
        const int64_t chunk = 512;
        const int64_t nchunks = (nwl + chunk - 1) / chunk;
#pragma omp parallel num_threads(nthreads)
        {
            double* acc = malloc(sizeof(double) * L * chunk);
#pragma omp for schedule(dynamic, 1)
            for (int64_t c = 0; c < nchunks; ++c) {
                const int64_t j0 = c * chunk;
                const int64_t j1 = (j0 + chunk < nwl) ? j0 + chunk : nwl;
                memset(acc, 0, sizeof(double) * L * chunk);
                for (int64_t i = 0; i < nlines; ++i) {
                    if (lb[i] > ub[i]) continue;
                    const int64_t a = (lb[i] - 1 > j0) ? lb[i] - 1 : j0;
                    const int64_t b = (ub[i] < j1) ? ub[i] : j1; /* exclusive, 0-based */
                    for (int64_t j = a; j < b; ++j)
                        for (int64_t l = 0; l < L; ++l)
                            acc[(j - j0) * L + l] +=
                                oracle_line_profile(wl[i], sig[i * L + l], gam[i * L + l],
                                                    amp[i * L + l], wls[j]);
                }
                for (int64_t j = j0; j < j1; ++j)
                    for (int64_t l = 0; l < L; ++l) alpha[j * L + l] += acc[(j - j0) * L + l];
            }
            free(acc);
        }
    } else {
        /* the reference's schedule: line chunks, private alpha per chunk, summed in chunk order */
        const int64_t nch = nthreads;
        const int64_t csize = nlines / nch + (nlines % nch > 0);
        double** parts = malloc(sizeof(double*) * nch);
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
        for (int64_t c = 0; c < nch; ++c) {
            double* a_task = calloc((size_t)(L * nwl), sizeof(double));
            const int64_t i0 = c * csize;
            const int64_t i1 = (i0 + csize < nlines) ? i0 + csize : nlines;
            for (int64_t i = i0; i < i1; ++i) {
                if (lb[i] > ub[i]) continue;
                for (int64_t j = lb[i] - 1; j < ub[i]; ++j)
                    for (int64_t l = 0; l < L; ++l)
                        a_task[j * L + l] += oracle_line_profile(wl[i], sig[i * L + l],
                                                                 gam[i * L + l], amp[i * L + l],
                                                                 wls[j]);
            }
            parts[c] = a_task;
        }
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t k = 0; k < L * nwl; ++k) {
            double s = parts[0][k];
            for (int64_t c = 1; c < nch; ++c) s += parts[c][k];
            alpha[k] += s;
        }
        for (int64_t c = 0; c < nch; ++c) free(parts[c]);
        free(parts);
    }
    free(sig); free(gam); free(amp); free(lb); free(ub); free(beta); free(n_eff);
    return 0;
}


/* ------------------------------------------------------------------------------------------ */
/* Adjoint of the line opacity (the reverse-mode derivative of line_absorption! with respect to
 * its per-layer inputs).  The windows are integer functions of the inputs and are held fixed,
 * as ForwardDiff does through searchsorted; everything else is differentiated exactly,
 * including the branch structure of voigt_hjerting (the derivative of the branch chosen). */

/* d/dv of the Harris series pieces (src/line_absorption.jl:251-265): writes H0', H1', H2'. */
static void harris_series_dv(double v, double H0, double H1, double* dH0, double* dH1,
                             double* dH2) {
    const double v2 = v * v;
    *dH0 = -2.0 * v * H0;
    if (v < 1.3) {
        *dH1 = -0.15516677 + (2.0 * 3.288675912 + (3.0 * -2.34357915 + 4.0 * 0.42139162 * v) * v) * v;
    } else if (v < 2.4) {
        *dH1 = 9.39456063 + (2.0 * -6.61487486 + (3.0 * 1.98919585 - 4.0 * 0.22041650 * v) * v) * v;
    } else {
        const double P = 0.554153432 +
                         (0.278711796 + (-0.1883256872 + (0.042991293 - 0.003278278 * v) * v) * v) * v;
        const double dP = 0.278711796 +
                          (2.0 * -0.1883256872 + (3.0 * 0.042991293 - 4.0 * 0.003278278 * v) * v) * v;
        const double q = v2 - 1.5;
        (void)H1;
        *dH1 = (dP * q - 2.0 * v * P) / (q * q);
    }
    *dH2 = H0 * (-6.0 * v + 4.0 * v2 * v);
}

/* voigt_hjerting (src/line_absorption.jl:286-310) and its partial derivatives in a and v. */
double oracle_voigt_hjerting_grad(double a, double v, double* dHda, double* dHdv) {
    const double sqrt_pi = sqrt(M_PI);
    const double v2 = v * v;
    if (a <= 0.2 && v >= 5.0) {
        const double w = 1.0 / v2;
        const double poly = 1.0 + 1.5 * w + 3.75 * (w * w);
        *dHda = w / sqrt_pi * poly;
        /* d/dv [w (1 + 1.5 w + 3.75 w^2)] = (1 + 3 w + 11.25 w^2) dw/dv, dw/dv = -2 w / v */
        *dHdv = a / sqrt_pi * (1.0 + 3.0 * w + 11.25 * w * w) * (-2.0 * w / v);
        return (a / sqrt_pi * w) * poly;
    } else if (a <= 0.2) {
        double H1, H2, dH0, dH1, dH2;
        const double H0 = oracle_harris_series(v, &H1, &H2);
        harris_series_dv(v, H0, H1, &dH0, &dH1, &dH2);
        *dHda = H1 + 2.0 * H2 * a;
        *dHdv = dH0 + (dH1 + dH2 * a) * a;
        return H0 + (H1 + H2 * a) * a;
    } else if (a <= 1.4 && a + v < 3.2) {
        double H1, H2, dH0, dH1, dH2;
        const double H0 = oracle_harris_series(v, &H1, &H2);
        harris_series_dv(v, H0, H1, &dH0, &dH1, &dH2);
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
        *dHda = dpsi * S + psi * dSda;
        *dHdv = psi * dSdv;
        return psi * S;
    } else {
        const double r2 = v2 / (a * a);
        const double u = 1.0 / sqrt(2.0) / ((r2 + 1.0) * a);
        const double u2 = u * u;
        const double c4 = (r2 - 2.0) * 15.0 * r2 + 2.0; /* 15 r^4 - 30 r^2 + 2 */
        const double P = 1.0 + (3.0 * r2 - 1.0 + c4 * u2) * u2;
        /* u = a / (sqrt(2) (v^2 + a^2)); r^2 = v^2 / a^2 */
        const double den = sqrt(2.0) * (v2 + a * a) * (v2 + a * a);
        const double du_da = (v2 - a * a) / den, du_dv = -2.0 * a * v / den;
        const double dr2_da = -2.0 * v2 / (a * a * a), dr2_dv = 2.0 * v / (a * a);
        const double dP_dr2 = 3.0 * u2 + (30.0 * r2 - 30.0) * u2 * u2;
        const double dP_du = 2.0 * (3.0 * r2 - 1.0) * u + 4.0 * c4 * u2 * u;
        const double s2p = sqrt(2.0 / M_PI);
        *dHda = s2p * (P * du_da + u * (dP_dr2 * dr2_da + dP_du * du_da));
        *dHdv = s2p * (P * du_dv + u * (dP_dr2 * dr2_dv + dP_du * du_dv));
        return s2p * u * P;
    }
}

/* Partial derivatives of one (line, layer)'s profile parameters with respect to the layer's
 * inputs x = (T, n_e, xi, n_HI, n_HeI, n_H2, n/U of the line's species): the line_profile of
 * src/line_absorption.jl:245-249 is scaling * H(a, v) with inv = 1/(sigma sqrt 2), a = gamma inv,
 * v = |lambda - lambda0| inv, scaling = inv/sqrt(pi) * amplitude.  da[q] = da/dx_q,
 * dli[q] = d ln(inv)/dx_q, ds[q] = dscaling/dx_q (src/line_absorption.jl:86-106, 190). */
typedef struct {
    double inv, a, scaling;
    double da[7], dli[7], ds[7];
} layer_partials_t;

static void layer_partials(const problem_t* P, int64_t i, int64_t l, const double* sig,
                           const double* gam, const double* amp, double c_He, double c_H2,
                           layer_partials_t* o) {
    const int32_t s = P->species_index[i];
    const double m = P->species_mass[s];
    const int mol = P->species_is_mol[s] != 0;
    const double lam0 = P->wl[i];
    const double E_lower = P->E_lower[i];
    const double E_up = E_lower + c_cgs * hplanck_eV / lam0;
    const double gf = pow(10.0, P->log_gf[i]);
    const double sl = oracle_sigma_line(lam0);
    const double gscale = (lam0 * lam0) / (c_cgs * (4.0 * M_PI));
    const double T = P->temps[l], xi = P->xi[l];
    const double inv = 1.0 / (sig[l] * sqrt(2.0));
    const double a = gam[l] * inv;
    const double scaling = inv / sqrt(M_PI) * amp[l];
    /* d ln(1/(sigma sqrt 2)) = -d ln sigma, sigma^2 ~ kT/m + xi^2/2 */
    const double q = kboltz_cgs * T / m + xi * xi / 2.0;
    const double dli_dT = -(kboltz_cgs / m) / (2.0 * q);
    const double dli_dxi = -xi / (2.0 * q);
    /* dGamma / dx */
    double dG_dT = 0.0, dG_dne = 0.0, vdw = 0.0;
    if (!mol) {
        const double st = oracle_scaled_stark(P->gamma_stark[i], T);
        vdw = oracle_scaled_vdW(P->vdW1[i], P->vdW2[i], m, T);
        const double dvdw_dT = (P->vdW2[i] == -1.0 ? 0.3 : (1.0 - P->vdW2[i]) / 2.0) * vdw / T;
        dG_dne = st;
        dG_dT = P->n_e[l] * st / (6.0 * T) + P->n_eff[l] * dvdw_dT;
    }
    /* da/dx = inv dgamma/dx + a dln(inv)/dx */
    const double ig = inv * gscale;
    /* scaling = inv / sqrt(pi) * gf * sigma_line * levels * nU */
    const double nU = P->n_div_U[(int64_t)s * P->L + l];
    const double bl = P->beta[l];
    const double el = exp(-bl * E_lower), eu = exp(-bl * E_up);
    const double dlev_dT = (bl / T) * (E_lower * el - E_up * eu);
    const double base = inv / sqrt(M_PI) * gf * sl;
    o->inv = inv;
    o->a = a;
    o->scaling = scaling;
    const double da[7] = {ig * dG_dT + a * dli_dT, ig * dG_dne, a * dli_dxi, ig * vdw,
                          ig * c_He * vdw, ig * c_H2 * vdw, 0.0};
    const double dli[7] = {dli_dT, 0.0, dli_dxi, 0.0, 0.0, 0.0, 0.0};
    const double ds[7] = {scaling * dli_dT + base * nU * dlev_dT, 0.0, scaling * dli_dxi, 0.0, 0.0,
                          0.0, base * (el - eu)};
    memcpy(o->da, da, sizeof da);
    memcpy(o->dli, dli, sizeof dli);
    memcpy(o->ds, ds, sizeof ds);
}

static void vdW_mix(double* c_He, double* c_H2) {
    /* src/line_absorption.jl:54-60 */
    const double a_H = 4.50711, a_He = 1.38375, a_H2 = 5.503;
    const double m_H = 1.008 * amu_cgs, m_He = 4.003 * amu_cgs;
    *c_He = pow(a_He / a_H, 0.4) * pow(m_He / m_H, -0.3);
    *c_H2 = pow(a_H2 / a_H, 0.4) * pow(2.0, -0.3);
}

/* The window of line i: computed (line_setup), or the frozen one when lb_in/ub_in are given. */
static void line_window(const problem_t* P, int64_t i, const int64_t* lb_in, const int64_t* ub_in,
                        int64_t* lb, int64_t* ub) {
    if (lb_in && ub_in) {
        *lb = lb_in[i] < 1 ? 1 : lb_in[i];
        *ub = ub_in[i] > P->nwl ? P->nwl : ub_in[i];
    }
}

/*
 * Vector-Jacobian product: given G = dL/dalpha (column-major [L x nwl]), accumulates
 *   g_T[l], g_ne[l], g_xi[l], g_nHI[l], g_nHeI[l], g_nH2[l]  and  g_nU[s*L + l] (n_div_U).
 * Inputs as oracle_line_absorption (lb_in/ub_in: optional frozen windows).  Each line's
 * contribution per layer is
 *   scaling * (da/dx * S_a + dln(1/(sigma sqrt 2))/dx * S_v) + dscaling/dx * S_H
 * with S_H = sum_j G H, S_a = sum_j G dH/da, S_v = sum_j G v dH/dv over the line's window; the
 * contributions are summed in line order (deterministic).  Returns 0.
 */
int oracle_line_absorption_vjp(const double* G, int64_t L, int64_t nwl, const double* wls,
                               int64_t nlines, const double* wl, const double* log_gf,
                               const double* E_lower, const double* gamma_rad,
                               const double* gamma_stark, const double* vdW1, const double* vdW2,
                               const int32_t* species_index, const double* species_mass,
                               const uint8_t* species_is_mol, int64_t n_species,
                               const double* n_div_U, const double* temps, const double* n_e,
                               const double* n_HI, const double* n_HeI, const double* n_H2,
                               const double* xi, int cntm_mode, const double* cntm_values,
                               int64_t n_knots, const double* knots, double cutoff_threshold,
                               int nthreads, double* g_T, double* g_ne, double* g_xi,
                               double* g_nHI, double* g_nHeI, double* g_nH2, double* g_nU,
                               const int64_t* lb_in, const int64_t* ub_in) {
    for (int64_t l = 0; l < L; ++l) g_T[l] = g_ne[l] = g_xi[l] = g_nHI[l] = g_nHeI[l] = g_nH2[l] = 0.0;
    for (int64_t k = 0; k < n_species * L; ++k) g_nU[k] = 0.0;
    if (nlines == 0) return 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    problem_t P = {L, nwl, nlines, wls, wl, log_gf, E_lower, gamma_rad, gamma_stark, vdW1, vdW2,
                   species_index, species_mass, species_is_mol, n_div_U, temps, n_e, xi,
                   cntm_mode, cntm_values, knots, n_knots, cutoff_threshold, NULL, NULL};
    double* beta = malloc(sizeof(double) * L);
    double* n_eff = malloc(sizeof(double) * L);
    double c_He, c_H2;
    vdW_mix(&c_He, &c_H2);
    for (int64_t l = 0; l < L; ++l) {
        beta[l] = 1.0 / (kboltz_eV * temps[l]);
        n_eff[l] = n_HI[l] + c_He * n_HeI[l] + c_H2 * n_H2[l];
    }
    P.beta = beta;
    P.n_eff = n_eff;
    enum { NG = 7 }; /* T, ne, xi, nHI, nHeI, nH2, nU */
    double* contrib = calloc((size_t)(nlines * L * NG), sizeof(double));
#pragma omp parallel num_threads(nthreads)
    {
        double* sig = malloc(sizeof(double) * L);
        double* gam = malloc(sizeof(double) * L);
        double* amp = malloc(sizeof(double) * L);
        double* rho = malloc(sizeof(double) * L);
        double* tmp = malloc(sizeof(double) * L);
#pragma omp for schedule(dynamic, 4)
        for (int64_t i = 0; i < nlines; ++i) {
            int64_t lb, ub;
            line_setup(&P, i, sig, gam, amp, rho, tmp, &lb, &ub);
            line_window(&P, i, lb_in, ub_in, &lb, &ub);
            if (lb > ub) continue;
            const double lam0 = wl[i];
            for (int64_t l = 0; l < L; ++l) {
                layer_partials_t d;
                layer_partials(&P, i, l, sig, gam, amp, c_He, c_H2, &d);
                double SH = 0.0, Sa = 0.0, Sv = 0.0;
                for (int64_t j = lb - 1; j < ub; ++j) {
                    const double g = G[j * L + l];
                    if (g == 0.0) continue;
                    const double v = fabs(wls[j] - lam0) * d.inv;
                    double Ha, Hv;
                    const double H = oracle_voigt_hjerting_grad(d.a, v, &Ha, &Hv);
                    SH += g * H;
                    Sa += g * Ha;
                    Sv += g * v * Hv;
                }
                double* c = contrib + (i * L + l) * NG;
                for (int q = 0; q < NG; ++q)
                    c[q] = d.scaling * (d.da[q] * Sa + d.dli[q] * Sv) + d.ds[q] * SH;
            }
        }
        free(sig); free(gam); free(amp); free(rho); free(tmp);
    }
    for (int64_t i = 0; i < nlines; ++i) {
        const int32_t s = species_index[i];
        for (int64_t l = 0; l < L; ++l) {
            const double* c = contrib + (i * L + l) * NG;
            g_T[l] += c[0];
            g_ne[l] += c[1];
            g_xi[l] += c[2];
            g_nHI[l] += c[3];
            g_nHeI[l] += c[4];
            g_nH2[l] += c[5];
            g_nU[(int64_t)s * L + l] += c[6];
        }
    }
    free(contrib);
    free(beta);
    free(n_eff);
    return 0;
}

/*
 * Jacobian-vector products (forward mode, what a ForwardDiff.Dual alpha receives): for each of
 * K tangent directions k of the per-layer inputs -- t_T[k*L + l], t_ne, t_xi, t_nHI, t_nHeI,
 * t_nH2 (each [K x L]) and t_nU[(k*n_species + s)*L + l] -- ACCUMULATES into dalpha + k*L*nwl
 * (column-major [L x nwl] each)
 *   sum over lines of  (ds.x) H + scaling ((da.x) dH/da + (dli.x) v dH/dv)
 * over each line's window (computed, or frozen by lb_in/ub_in), the windows held fixed and the
 * branch of voigt_hjerting in use differentiated exactly, lines in list order per element.
 * Returns 0.
 */
int oracle_line_absorption_jvp(double* dalpha, int64_t K, int64_t L, int64_t nwl,
                               const double* wls, int64_t nlines, const double* wl,
                               const double* log_gf, const double* E_lower,
                               const double* gamma_rad, const double* gamma_stark,
                               const double* vdW1, const double* vdW2,
                               const int32_t* species_index, const double* species_mass,
                               const uint8_t* species_is_mol, int64_t n_species,
                               const double* n_div_U, const double* temps, const double* n_e,
                               const double* n_HI, const double* n_HeI, const double* n_H2,
                               const double* xi, int cntm_mode, const double* cntm_values,
                               int64_t n_knots, const double* knots, double cutoff_threshold,
                               int nthreads, const double* t_T, const double* t_ne,
                               const double* t_xi, const double* t_nHI, const double* t_nHeI,
                               const double* t_nH2, const double* t_nU, const int64_t* lb_in,
                               const int64_t* ub_in) {
    if (nlines == 0 || K == 0) return 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    problem_t P = {L, nwl, nlines, wls, wl, log_gf, E_lower, gamma_rad, gamma_stark, vdW1, vdW2,
                   species_index, species_mass, species_is_mol, n_div_U, temps, n_e, xi,
                   cntm_mode, cntm_values, knots, n_knots, cutoff_threshold, NULL, NULL};
    double* beta = malloc(sizeof(double) * L);
    double* n_eff = malloc(sizeof(double) * L);
    double c_He, c_H2;
    vdW_mix(&c_He, &c_H2);
    for (int64_t l = 0; l < L; ++l) {
        beta[l] = 1.0 / (kboltz_eV * temps[l]);
        n_eff[l] = n_HI[l] + c_He * n_HeI[l] + c_H2 * n_H2[l];
    }
    P.beta = beta;
    P.n_eff = n_eff;
    /* per (line, layer, direction): the coefficients of H, dH/da and v dH/dv */
    double* cf = malloc(sizeof(double) * nlines * L * K * 3);
    int64_t* lb = malloc(sizeof(int64_t) * nlines);
    int64_t* ub = malloc(sizeof(int64_t) * nlines);
    double* av = malloc(sizeof(double) * nlines * L * 2); /* (inv, a) */
#pragma omp parallel num_threads(nthreads)
    {
        double* sig = malloc(sizeof(double) * L);
        double* gam = malloc(sizeof(double) * L);
        double* amp = malloc(sizeof(double) * L);
        double* rho = malloc(sizeof(double) * L);
        double* tmp = malloc(sizeof(double) * L);
#pragma omp for schedule(dynamic, 4)
        for (int64_t i = 0; i < nlines; ++i) {
            line_setup(&P, i, sig, gam, amp, rho, tmp, lb + i, ub + i);
            line_window(&P, i, lb_in, ub_in, lb + i, ub + i);
            const int32_t s = species_index[i];
            for (int64_t l = 0; l < L; ++l) {
                layer_partials_t d;
                layer_partials(&P, i, l, sig, gam, amp, c_He, c_H2, &d);
                av[(i * L + l) * 2] = d.inv;
                av[(i * L + l) * 2 + 1] = d.a;
                for (int64_t k = 0; k < K; ++k) {
                    const double x[7] = {t_T ? t_T[k * L + l] : 0.0, t_ne ? t_ne[k * L + l] : 0.0,
                                         t_xi ? t_xi[k * L + l] : 0.0,
                                         t_nHI ? t_nHI[k * L + l] : 0.0,
                                         t_nHeI ? t_nHeI[k * L + l] : 0.0,
                                         t_nH2 ? t_nH2[k * L + l] : 0.0,
                                         t_nU ? t_nU[(k * n_species + s) * L + l] : 0.0};
                    double dsx = 0.0, dax = 0.0, dlx = 0.0;
                    for (int q = 0; q < 7; ++q) {
                        if (x[q] == 0.0) continue; /* a zero tangent adds nothing, NaN partials included */
                        dsx += d.ds[q] * x[q];
                        dax += d.da[q] * x[q];
                        dlx += d.dli[q] * x[q];
                    }
                    double* c = cf + ((i * L + l) * K + k) * 3;
                    c[0] = dsx;
                    c[1] = d.scaling * dax;
                    c[2] = d.scaling * dlx;
                }
            }
        }
        free(sig); free(gam); free(amp); free(rho); free(tmp);
    }
    /* each thread owns a block of wavelengths; lines in list order per element */
    const int64_t chunk = 512;
    const int64_t nchunks = (nwl + chunk - 1) / chunk;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (int64_t c = 0; c < nchunks; ++c) {
        const int64_t j0 = c * chunk;
        const int64_t j1 = (j0 + chunk < nwl) ? j0 + chunk : nwl;
        for (int64_t i = 0; i < nlines; ++i) {
            if (lb[i] > ub[i]) continue;
            const int64_t a0 = (lb[i] - 1 > j0) ? lb[i] - 1 : j0;
            const int64_t b0 = (ub[i] < j1) ? ub[i] : j1;
            for (int64_t j = a0; j < b0; ++j)
                for (int64_t l = 0; l < L; ++l) {
                    const double inv = av[(i * L + l) * 2], a = av[(i * L + l) * 2 + 1];
                    const double v = fabs(wls[j] - wl[i]) * inv;
                    double Ha, Hv;
                    const double H = oracle_voigt_hjerting_grad(a, v, &Ha, &Hv);
                    const double vHv = v * Hv;
                    for (int64_t k = 0; k < K; ++k) {
                        const double* f = cf + ((i * L + l) * K + k) * 3;
                        dalpha[k * L * nwl + j * L + l] += f[0] * H + f[1] * Ha + f[2] * vHv;
                    }
                }
        }
    }
    free(cf); free(lb); free(ub); free(av); free(beta); free(n_eff);
    return 0;
}
