// MI355X (gfx950) line-opacity library: the body of Korg.jl's line_absorption!
// (src/line_absorption.jl:33-129) as HIP kernels behind the C ABI of include/korg_line.h.
//
// Pipeline of one call (one stream, inputs resident in HBM):
//   k_layers      per-layer tables: beta = 1/(kT), n_eff_vdW, (T/1e4)^(1/6), (T/1e4)^0.3; resets
//                 the call's counters
//   k_setup       one wavefront per line, lanes = layers (Float64): Doppler and Lorentz widths,
//                 amplitudes, the window (max over layers) and the per-line records the profile
//                 kernels read
//   k_count / k_scan / k_scatter   stable partitions of the lines into window-width classes and
//                 (Float32 form) near-zone-reach classes
//   k_ranges      per (class, wavelength tile): index range of candidate lines
//   Float64 form:
//     k_profile_f64  line_profile/voigt_hjerting per evaluation (the reference's expressions)
//   Float32 form with Float64 fix-ups:
//     k_wing_ff      the far field: each line's far-wing series where a 64-point sub-tile lies
//                    well away from its centre, as 12-term expansions about the sub-tiles'
//                    centres summed over lines on the matrix cores
//     k_wing_bf16    far wing of every line as a rank-4 update on the matrix cores (bfloat16
//                    MFMA on split high/low operands; k_wing_mfma, FP32 MFMA, when WING_BF16 = 0),
//                    each tile starting from its sub-tiles' far-field sums
//     k_near         the "near zone" of each line (Voigt cores, and where the damped-wing series
//                    is not yet converged): exact branch-by-branch evaluation
// Each profile work item owns its block of alpha outright; lines are added in a fixed order, and
// the block is added into alpha once per kernel.  No atomics touch alpha: results are
// deterministic.
//
// The far-wing expansion.  With w = 1/v^2 the Hunger far-wing branch A (a <= 0.2, v >= 5) is
//   H = (a/sqrt(pi)) (w + 1.5 w^2 + 3.75 w^3),
// and branch D, H = (a/sqrt(pi)) sum_k C_k r^k with r = 1/(v^2 + a^2) = w/(1 + a^2 w) and
// C = (1, 1.5, 3.75 - 2a^2, -15a^2, 11.75a^4), expands to
//   H = (a/sqrt(pi)) (w + (1.5 - a^2) w^2 + (3.75 - 5a^2 + a^4) w^3
//                     + (-26.25a^2 + 10.5a^4 - a^6) w^4 + e5 w^5 + ...),
//   e5 = 94.25a^4 - 18a^6 + a^8.
// Since v = s |delta| (delta the offset in DELTA_UNIT, s = DELTA_UNIT/(sigma sqrt 2)), w = R/s^2 with
// R = 1/delta^2 shared by all layers, so phi = scaling H = sum_{k=1..4} c_k R^k with per-(line,
// layer) coefficients c_k.  Summed over lines this is a matrix product: alpha[layer, point] +=
// C[layer, (line,k)] * P[(line,k), point].  The 4-term truncation is exact for branch A and is
// used for branch D only where |e5| w^4 <= SERIES_EPS; nearer points belong to the near zone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "korg_line.h"
#include "korg_line_device.h"

using namespace korg;

namespace {

constexpr int WAVE = 64;
// Relative truncation error allowed in the far-wing series at the edge of a line's near zone (the
// first dropped term against the leading one).  1e-5 instead of 2e-8 shrinks the near zones of
// strongly damped layers by 2.2x; the measured Float32 alpha error is unchanged (1.3e-5 at the
// 300,000-line workload, 2.3e-5 over a batch of random atmospheres), since other terms dominate
// it; from 2e-5 on it grows (2.0e-5 and 4.9e-5 at 2e-5 and 5e-5, 300,000 lines).
constexpr double SERIES_EPS = 1e-5;
// Relative size below which the R^3 and R^4 terms of the far-wing series are dropped (beyond the
// line's tail radius, where they are smaller than this fraction of the leading term in every
// layer).  1e-5 instead of 1e-6 issues 4% fewer far-wing MFMAs at the 300,000-line workload with
// the alpha error unchanged there (1.3e-5) and 2.27e-5 instead of 2.25e-5 over a batch of random
// atmospheres.
constexpr double SERIES_TAIL_EPS = 1e-5;
// The same bound for the derivatives of the series in the Float32 adjoint: its near zone (exact
// evaluation) extends to where the first dropped term's a- and v-derivatives, relative to the
// leading term's, are below this.
constexpr double SERIES_ADJ_EPS = 2e-7;
// The far wing's matrix products in bfloat16 with hi/lo splits of both operands (three products,
// about 16 significant bits) instead of Float32.
#ifndef KORG_WING_BF16
#define KORG_WING_BF16 1
#endif
constexpr bool WING_BF16 = KORG_WING_BF16;

// --------------------------------------------------------------------------------------------
// wave helpers
// --------------------------------------------------------------------------------------------
__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

__device__ inline double readlane_f64(double x, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
    return __hiloint2double(hi, lo);
}
__device__ inline int readlane_i32(int x, int lane) { return __builtin_amdgcn_readlane(x, lane); }
__device__ inline float readlane_f32(float x, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), lane));
}

// Wavefront-uniform reads of data that no kernel in flight writes: through the constant address
// space they become scalar (SMEM) loads into SGPRs.
// (The host compilation pass of this single-source file sees plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
#define KCONST __attribute__((address_space(4)))
#else
#define KCONST
#endif
template <class T>
__device__ inline const KCONST T* uniform_ptr(const T* p) {
    const uint64_t a = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return (const KCONST T*)(((uint64_t)hi << 32) | lo);
}

// x moved across lanes by a DPP control (both halves of the double); lanes with no source, or in
// rows the row mask leaves out, keep their own x.
template <int CTRL, int ROWS>
__device__ inline double dpp_f64(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)u, (int)(uint32_t)u, CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(u >> 32), (int)(uint32_t)(u >> 32), CTRL, ROWS,
                                               0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// The largest x of the wavefront (no NaN among them), in every lane: prefix maxima within each
// 16-lane row by row shifts, then rows 1 and 3 take row 0's and row 2's last lane (row_bcast:15)
// and rows 2 and 3 lane 31 (row_bcast:31), so lane 63 holds the maximum.  Exact, as a max is.
__device__ inline double wave_max(double x) {
    x = fmax(x, dpp_f64<0x111, 0xF>(x));  // row_shr:1
    x = fmax(x, dpp_f64<0x112, 0xF>(x));  // row_shr:2
    x = fmax(x, dpp_f64<0x114, 0xF>(x));  // row_shr:4
    x = fmax(x, dpp_f64<0x118, 0xF>(x));  // row_shr:8
    x = fmax(x, dpp_f64<0x142, 0xA>(x));  // row_bcast:15
    x = fmax(x, dpp_f64<0x143, 0xC>(x));  // row_bcast:31
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 63);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Per-lane search with a guess: on a (nearly) evenly spaced ascending array, the index estimated
// from g0 + i * step is checked against its neighbours and moved at most 3 places.  upper: the
// first index with a[i] > x, else the first with a[i] >= x.  Returns false when unresolved.
__device__ inline bool lane_bound_guess(const double* __restrict__ a, int64_t n, double g0,
                                        double ginv, double x, bool upper, int64_t& out) {
    if (!(x == x)) return false;
    double t = (x - g0) * ginv;
    t = t < -1.0 ? -1.0 : (t > (double)n ? (double)n : t);
    int64_t j = (int64_t)ceil(t);
    j = j < 0 ? 0 : (j > n ? n : j);
    for (int it = 0; it < 4; ++it) {
        const bool left_ok = j == 0 || (upper ? a[j - 1] <= x : a[j - 1] < x);
        const bool right_ok = j == n || (upper ? a[j] > x : a[j] >= x);
        if (left_ok && right_ok) {
            out = j;
            return true;
        }
        j += left_ok ? 1 : -1;
    }
    return false;
}

// Lower (upper: upper) bound of x on an ascending array for one lane's own query: the guess,
// else a binary search.
__device__ inline int64_t lane_bound(const double* __restrict__ a, int64_t n, double g0, double ginv,
                                     double x, bool upper) {
    if (!(x == x)) return n;  // NaN: nothing compares; lower and upper bound are both n
    int64_t r = 0;
    if (lane_bound_guess(a, n, g0, ginv, x, upper, r)) return r;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (upper ? a[mid] <= x : a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// --------------------------------------------------------------------------------------------
// k_layers: per-layer tables (src/line_absorption.jl:39, 54-60, 195, 210) and counter reset
// --------------------------------------------------------------------------------------------
struct LayerTab {
    double beta, n_eff, t16, t03;
};

// Device counters of one call (reset by k_layers): [0] profile queue, [1] evaluations,
// [2] lines used, [3] error flags, [4] near-zone reach in points, [5] near queue,
// [6..9] evaluations by Hunger branch, [10] near-zone evaluations, [11] far-wing MFMAs per layer
// block, [12] tail-zone evaluations, [13] entries of the per-tile line lists, [14] segments of the
// adjoint, [15] branch-A evaluations inside near zones (with branch counting), [16..23] the
// adjoint moments' queues (one per part of the group list), [24] the far-field queue, [25]
// entries of the far-field line lists.
// Error flags: 1 lines not sorted, 2 species index out of range, 4 the per-tile line lists need
// more entries than their buffer holds.  With any flag set no kernel writes alpha.
constexpr int N_COUNTERS = 26;
constexpr unsigned long long FLAG_UNSORTED = 1, FLAG_SPECIES = 2, FLAG_CAPACITY = 4;
// Largest far-wing series coefficient kept in Float32 (2^120: headroom below FLT_MAX ~ 2^128).
constexpr double F32_COEF_MAX = 1.329227995784916e36;

struct GridInfo {
    double g0, ginv;  // first point and inverse mean spacing of the wavelength grid
    double k0, kinv;  // the same for the continuum knots
    double step;      // mean spacing of the grid (0 for a single point)
};

// One star -- one line_absorption! call -- of a batch.  Grid indices in every kernel are global:
// the star's points occupy [grid_off, grid_off + n_wl) of the concatenation of all stars' grids,
// each padded to a multiple of STAR_PAD points so that no tile of any kernel straddles two
// stars; its lines are [line_off, line_off + n_lines) of the concatenated line lists.
constexpr int STAR_ALIGN = 128;
struct StarDev {
    const double* wls;
    int64_t n_wl, grid_off;
    int64_t line_off, n_lines;
    void* alpha;  // column-major [L x n_wl]
    const double *wl, *log_gf, *E_lower, *gamma_rad, *gamma_stark, *vdW1, *vdW2;
    const int32_t* species;
    const int64_t *win_lb, *win_ub;  // frozen 1-based windows (null: computed)
    const double *temps, *n_e, *n_div_U, *n_HI, *n_HeI, *n_H2, *xi;
    int32_t cntm_mode;
    int32_t ff_irregular;  // far field: nonzero when the grid is not evenly spaced (k_ff_geo)
    int64_t n_knots;
    const double *knots, *cntm;
    double cutoff;
    GridInfo gi;  // written by k_layers
    // adjoint: dL/dalpha in (dtype of the call), and the outputs dL/d(T, n_e, xi, n_HI, n_HeI,
    // n_H2) [L] and dL/d(n_div_U) [S x L] (a null output is not written)
    void* G;
    double* gq[7];
    // fit step: the anchored scheme's reference optical depth and absorption [L], the data and
    // its uncertainty [n_wl]; outputs: the normalized flux [n_wl] (may be null) and chi^2 [1]
    const double *tau_ref, *alpha_ref, *data, *sigma;
    double *flux, *chi2;
    double* flux_c;  // flux of the continuum alone (flux-only transfer)
};

// The star holding global grid point j (the last star with grid_off <= j).
template <class P>
__device__ inline int star_of_point(P stars, int n, int64_t j) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (stars[mid].grid_off <= j) lo = mid; else hi = mid;
    }
    return lo;
}

// The star holding global line i.
template <class P>
__device__ inline int star_of_line(P stars, int n, int64_t i) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (stars[mid].line_off <= i) lo = mid; else hi = mid;
    }
    return lo;
}

// Per-layer tables of every star (block = star); block 0 also clears the call's counters, and
// each star's grid summary is stored in its StarDev.
__global__ void k_layers(int n_stars, int L, StarDev* __restrict__ stars, LayerTab* __restrict__ tab,
                         unsigned long long* __restrict__ counters) {
    const int st = blockIdx.x;
    StarDev& S = stars[st];
    if (st == 0 && threadIdx.x == 0)
        for (int k = 0; k < N_COUNTERS; ++k) counters[k] = 0;
    if (threadIdx.x == 0) {
        const int64_t n_knots = S.cntm_mode == KORG_CNTM_GRID ? S.n_knots : 0;
        GridInfo g;
        g.g0 = S.wls[0];
        g.step = S.n_wl > 1 ? (S.wls[S.n_wl - 1] - S.wls[0]) / (double)(S.n_wl - 1) : 0.0;
        S.ff_irregular = 0;
        g.ginv = S.n_wl > 1 ? (double)(S.n_wl - 1) / (S.wls[S.n_wl - 1] - S.wls[0]) : 0.0;
        g.k0 = n_knots > 1 ? S.knots[0] : 0.0;
        g.kinv = n_knots > 1 ? (double)(n_knots - 1) / (S.knots[n_knots - 1] - S.knots[0]) : 0.0;
        S.gi = g;
    }
    const double a_H = 4.50711, a_He = 1.38375, a_H2 = 5.503;
    const double m_H = 1.008 * amu_cgs, m_He = 4.003 * amu_cgs;
    const double c_He = pow(a_He / a_H, 0.4) * pow(m_He / m_H, -0.3);
    const double c_H2 = pow(a_H2 / a_H, 0.4) * pow(2.0, -0.3);
    for (int l = threadIdx.x; l < L; l += blockDim.x) {
        const double T = S.temps[l];
        LayerTab t;
        t.beta = 1.0 / (kboltz_eV * T);
        t.n_eff = S.n_HI[l] + c_He * S.n_HeI[l] + c_H2 * S.n_H2[l];
        t.t16 = pow(T / 10000.0, 1.0 / 6.0);
        t.t03 = pow(T / 10000.0, 0.3);
        tab[(int64_t)st * L + l] = t;
    }
}

// --------------------------------------------------------------------------------------------
// The far field of the Float32 far wing.  Cut the grid into sub-tiles of FF_SUB points with
// centre c and half-width h (of their points' wavelengths).  Where a sub-tile lies inside a
// line's window, outside its near zone and at least FF_RATIO half-widths from its centre, the
// line's far-wing series sum_k c_k (d + t)^(-2k) (d = c - lam0, t = lambda - c, |t| <= h) is the
// convergent expansion
//   sum_m tau^m sum_k c_k W_km,   tau = t / h,   W_km = binom(-2k, m) (h/d)^m d^(-2k),
// truncated after FF_TERMS terms: at |h/d| <= 1/FF_RATIO the truncation is below 2e-10 of the
// leading (R) term's value at every point, and below 1.5e-6 of the R^4 term's (itself at most
// 1e-5 of the leading one beyond the tail radius).  The sums over lines, T[layer, m] = sum_(line, k) c_k W_km, are a matrix product
// per sub-tile (k_wing_ff, a super-tile of FF_NS sub-tiles per work item, so each line's
// coefficients are read once per FF_NS sub-tiles); the far-wing kernel starts each tile's
// accumulators from T times the powers tau^m of its points and adds only the lines whose
// far field does not cover the tile.  On the 300,000-line workload 93% of the far-wing
// (line, point) pairs fall in far fields, with 12 terms per 64 points in place of the points.
// k_setup decides once per line which sub-tiles take its far field: two runs of whole sub-tiles
// inside the window and clear of the near zone, the lower one starting at the window's first
// whole sub-tile, the upper one ending at its last (disjoint).  Every later kernel tests
// membership of these runs only, so all agree.  The runs are found in closed form on an evenly
// spaced grid (the grids of synthesize); a star whose grid is not (several ranges with gaps)
// takes no far field and is computed term by term.
// --------------------------------------------------------------------------------------------
constexpr int FF_SUB = 64;        // points per sub-tile
constexpr int FF_NS = 8;          // sub-tiles per work item of k_wing_ff
constexpr int FF_TERMS = 12;      // terms of the expansion in tau
constexpr double FF_RATIO = 8.0;  // distance from the line centre, in sub-tile half-widths
constexpr int FF_CB = FF_NS * FF_TERMS / 32;  // 32-column blocks of a work item's T
static_assert(FF_NS * FF_TERMS % 32 == 0, "whole column blocks");
// Stars' grids are padded to whole far-field super-tiles: then every tile and super-tile lies in
// one star and holds, in the same order, what it holds when the star is alone, so a batched star's
// alpha is bitwise its single-star alpha.  (STAR_ALIGN-point blocks divide it.)
constexpr int STAR_PAD = FF_NS * FF_SUB;
static_assert(STAR_PAD % STAR_ALIGN == 0, "padding in whole blocks");

struct FfGeo {
    double c_cm, h_cm;  // centre and half-width of the sub-tile's points (cm)
    double c_u;         // the centre in DELTA_UNIT
    float h_u;          // the half-width in DELTA_UNIT
    int first, last;    // first and last grid point of the sub-tile (global; first > last: none)
    int pad;
};
static_assert(sizeof(FfGeo) == 40, "FfGeo layout");

// Whether global sub-tile g lies in a line's far-field runs (x..y and z..w).
__device__ inline bool ff_in(const int4& r, int64_t g) {
    return (g >= r.x && g <= r.y) || (g >= r.z && g <= r.w);
}
// The list builders' key of a line: window (global grid indices), the lengths of its two runs
// packed in 16 bits each, and the line's index.  The lower run starts at the window's first whole
// sub-tile and the upper one ends at its last, so the window and the lengths give both runs (a
// run longer than 65,535 sub-tiles is shortened: the rest is computed term by term).
__device__ inline int ff_first_whole(int lb) { return (lb + FF_SUB - 1) / FF_SUB; }
__device__ inline int ff_last_whole(int ub) { return (ub + 1) / FF_SUB - 1; }
__device__ inline int4 ff_key(int lb, int ub, const int4& r, int line) {
    const int lo = min(max(r.y - r.x + 1, 0), 0xFFFF), hi = min(max(r.w - r.z + 1, 0), 0xFFFF);
    return make_int4(lb, ub, (lo << 16) | hi, line);
}
// The key's runs, lower then upper (empty when the end is before the start).
__device__ inline int4 ff_key_runs(const int4& k) {
    const int a = ff_first_whole(k.x), z = ff_last_whole(k.y);
    return make_int4(a, a + (int)((unsigned)k.z >> 16) - 1, z - (k.z & 0xFFFF) + 1, z);
}
// The sub-tiles' geometry, one thread per sub-tile.  Sub-tiles never straddle two stars (stars are
// aligned to a multiple of FF_SUB points); one wholly in a star's padding has no points.
__global__ __launch_bounds__(256) void k_ff_geo(int64_t n_sub, int n_stars,
                                                StarDev* __restrict__ stars,
                                                FfGeo* __restrict__ geo) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_sub) return;
    const int64_t j0 = g * FF_SUB;
    StarDev& S = stars[star_of_point(stars, n_stars, j0)];
    const int64_t first = max(j0, S.grid_off);
    const int64_t last = min(j0 + FF_SUB - 1, S.grid_off + S.n_wl - 1);
    FfGeo o = {0.0, 0.0, 0.0, 0.0f, 1, 0, 0};
    if (first <= last) {
        // the star's grid must be evenly spaced for the setup's far-field runs: every point
        // within 1e-3 of a step of the line through its first point with the mean spacing
        const double w0 = S.wls[0], step = S.gi.step;
        bool bad = !(step > 0.0);
        for (int64_t j = first; j <= last && !bad; ++j)
            bad = !(fabs(S.wls[j - S.grid_off] - (w0 + (double)(j - S.grid_off) * step)) <= 1e-3 * step);
        if (bad) atomicOr(&S.ff_irregular, 1);
        const double a = S.wls[first - S.grid_off], b = S.wls[last - S.grid_off];
        o.c_cm = 0.5 * (a + b);
        o.h_cm = 0.5 * (b - a);
        o.c_u = o.c_cm * (1.0 / DELTA_UNIT);
        o.h_u = (float)(o.h_cm * (1.0 / DELTA_UNIT));
        o.first = (int)first;
        o.last = (int)last;
    }
    geo[g] = o;
}

// The geometry of each super-tile (FF_NS whole sub-tiles of one star; first > last when the
// super-tile is not whole: the grid's end, or padding), for the adjoint's far-field steps over
// whole super-tiles.
__global__ __launch_bounds__(256) void k_ff_sgeo(int64_t n_super, int64_t n_sub, int n_stars,
                                                 const StarDev* __restrict__ stars,
                                                 const FfGeo* __restrict__ geo,
                                                 FfGeo* __restrict__ sgeo) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_super) return;
    FfGeo o = {0.0, 0.0, 0.0, 0.0f, 1, 0, 0};
    bool whole = (s + 1) * FF_NS <= n_sub;
    for (int k = 0; k < FF_NS && whole; ++k) {
        const FfGeo& g = geo[s * FF_NS + k];
        whole = g.first <= g.last && g.last - g.first + 1 == FF_SUB;
    }
    if (whole) {
        const int64_t first = geo[s * FF_NS].first, last = geo[s * FF_NS + FF_NS - 1].last;
        const StarDev& S = stars[star_of_point(stars, n_stars, first)];
        const double a = S.wls[first - S.grid_off], b = S.wls[last - S.grid_off];
        o.c_cm = 0.5 * (a + b);
        o.h_cm = 0.5 * (b - a);
        o.c_u = o.c_cm * (1.0 / DELTA_UNIT);
        o.h_u = (float)(o.h_cm * (1.0 / DELTA_UNIT));
        o.first = (int)first;
        o.last = (int)last;
    }
    sgeo[s] = o;
}

// --------------------------------------------------------------------------------------------
// k_setup: one wavefront per line; lane l handles layers l, l+64, ...
// --------------------------------------------------------------------------------------------
struct SetupArgs {
    int64_t n_lines;  // all stars
    int L;
    int Ls;          // rec64 stride per line (L padded to a whole number of Float64 groups)
    int n_lblocks;   // 64-layer blocks of the Float32 records
    int n_stars;
    const StarDev* stars;
    int n_species;
    const double* mass;
    const uint8_t* is_mol;
    const LayerTab* tab;  // [star * L + l]
    LineMeta* meta;
    double4* rec64;  // [line * Ls + l]: (1/(sigma sqrt 2), a, scaling, 0)
    float4* recM;    // [(line * n_lblocks + b) * 64 + lane]: far-wing coefficients, MFMA A layout
    double4* recM64;  // the Float64 series form: the same coefficients in Float64 (recM unused)
    unsigned long long* counters;
    int2* azone;     // Float32 form, when given: each line's adjoint near zone (global indices)
    // far field (null ffrun: off): sub-tile geometry, and each line's two runs of far-field
    // sub-tiles (global sub-tile indices, x..y and z..w; empty when x > y, z > w)
    const FfGeo* geo;
    int64_t n_sub;
    int4* ffrun;
    int4* ffrun_adj;  // Float32 fit step: the runs clear of the adjoint's near zone (null: none)
    int4* azq;       // Float32 fit step, at most 64 layers: each 16-layer block's adjoint near zone
    struct LinePre* pre;  // k_setup_pre -> k_setup: each line's own inputs
    struct LineWin* win;  // k_setup -> k_setup_post: each line's window and zone reaches
};

// The setup runs as three kernels per call: k_setup_pre (one lane per line: star lookup, the
// line's parameters, the continuum-knot interval), k_setup (one wavefront per line, lanes =
// layers: the per-layer records and their maxima over layers) and k_setup_post (one lane per
// line: the grid searches, zones, far-field runs, class and LineMeta).  The searches and loads
// are chains of dependent memory reads, which one lane per line keeps many of in flight.
struct LinePre {
    double wl, E_lower, gf_sl, gamma_rad, gamma_stark, vdW1, vdW2, m;  // gf_sl = 10^log_gf sigma_line
    double E_upper;  // src/line_absorption.jl:101
    double kt;       // continuum on knots: wl's weight in its knot interval kidx
    int32_t st, species, kidx, flags;  // flags: 1 molecule, 2 invalid species (line skipped)
};
struct LineWin {
    double window;                 // NaN when some layer's window is NaN
    double reach, treach, areach;  // zone radii in wavelength (negative: none)
    double ablk[4];                // the adjoint radius of each 16-layer block (with azq)
    int32_t series_off, pad;
};

constexpr double INV_SQRT_PI = 0.5641895835477563;
// (3.75 / SERIES_TAIL_EPS)^(1/4): the tail radius of branch A (a <= 0.2), as v
constexpr double VT_SMALL_A = 24.746160019198825;

// x^p for x >= 0 to about 1e-6 relative (Float32 log2/exp2 of the mantissa-exponent split, so any
// double exponent is fine), enlarged by 1e-5: an upper bound for a zone radius.
__device__ inline double root_up(double x, float p) {
    int e;
    const double m = frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
    const float lg = __builtin_amdgcn_logf((float)m) + (float)e;
    return (double)__builtin_amdgcn_exp2f(lg * p) * 1.00001;
}


// k_setup_pre: one lane per line.  Also flags lines out of order and species out of range (such
// a line gets an empty LineMeta here and is skipped by the later kernels).
__global__ __launch_bounds__(256) void k_setup_pre(SetupArgs A) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n_lines) return;
    const StarDev* __restrict__ stars = A.stars;
    const int st = star_of_line(stars, A.n_stars, i);
    const StarDev& S = stars[st];
    const int64_t li = i - S.line_off;  // index in the star's own line list
    LinePre P;
    P.wl = S.wl[li];
    P.st = st;
    P.species = S.species[li];
    P.kidx = 0;
    P.flags = 0;
    P.E_lower = P.gf_sl = P.gamma_rad = P.gamma_stark = P.vdW1 = P.vdW2 = P.m = 0.0;
    P.E_upper = P.kt = 0.0;
    const int32_t s = P.species;
    if (s < 0 || s >= A.n_species) {
        // invalid species index: flag the call as failed (nothing is added to alpha)
        atomicOr(&A.counters[3], 2ull);
        LineMeta mt = {P.wl, 1, 0, 1, 0, 0, N_CLASSES, 1, 0};
        A.meta[i] = mt;
        P.flags = 2;
        A.pre[i] = P;
        return;
    }
    if (li > 0 && S.wl[li - 1] > P.wl) atomicOr(&A.counters[3], 1ull);
    P.m = A.mass[s];
    P.flags = A.is_mol[s] != 0 ? 1 : 0;
    P.gamma_rad = S.gamma_rad[li];
    P.gamma_stark = S.gamma_stark[li];
    P.vdW1 = S.vdW1[li];
    P.vdW2 = S.vdW2[li];
    P.E_lower = S.E_lower[li];
    P.gf_sl = exp10(S.log_gf[li]) * sigma_line(P.wl);
    P.E_upper = P.E_lower + c_cgs * hplanck_eV / P.wl;  // src/line_absorption.jl:101
    if (S.cntm_mode == KORG_CNTM_GRID) {
        // knot interval holding wl: clamp(searchsortedlast(knots, wl), 1, n_knots - 1) - 1
        int64_t k = lane_bound(S.knots, S.n_knots, S.gi.k0, S.gi.kinv, P.wl, true) - 1;
        k = k < 0 ? 0 : (k > S.n_knots - 2 ? S.n_knots - 2 : k);
        P.kidx = (int32_t)k;
        P.kt = (P.wl - S.knots[k]) / (S.knots[k + 1] - S.knots[k]);
    }
    A.pre[i] = P;
}

// k_setup: one wavefront per line, lanes = layers (blocks of 64).  Three waves per SIMD (at most
// 168 registers): measured faster than two.  The Float32 coefficient block (F32) is a
// compile-time choice, so the Float64 variant carries none of its registers.
#ifndef KORG_SETUP_ATTR
#define KORG_SETUP_ATTR __attribute__((amdgpu_waves_per_eu(3)))
#endif
template <bool F32>
__global__ __launch_bounds__(256) KORG_SETUP_ATTR void k_setup(SetupArgs A) {
    const int lane = threadIdx.x % WAVE;
    // the wave's line, its star and the star's fields are wavefront-uniform: scalar registers
    const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const int64_t i = (int64_t)blockIdx.x * (blockDim.x / WAVE) + wid;
    if (i >= A.n_lines) return;  // whole wavefront
    const int L = A.L;
    const KCONST LinePre* pre = uniform_ptr(A.pre);
    const LinePre P = pre[i];
    if (P.flags & 2) return;  // invalid species: LineMeta written by k_setup_pre
    const KCONST StarDev* stars = uniform_ptr(A.stars);
    const int st = P.st;
    const KCONST StarDev& S = stars[st];
    const int64_t li = i - S.line_off;
    const double wl = P.wl;
    const int32_t s = P.species;
    const double m = P.m;
    const bool mol = (P.flags & 1) != 0;
    const double gamma_rad = P.gamma_rad, gamma_stark = P.gamma_stark;
    const double vdW1 = P.vdW1, vdW2 = P.vdW2;
    const double E_lower = P.E_lower;
    const double E_upper = P.E_upper;
    const double gf_sl = P.gf_sl;
    const int64_t kidx = P.kidx;
    const double kt = P.kt;
    double dmax = 0.0, lmax = 0.0, rmax = -1.0, tmax = -1.0, amax = -1.0;
    bool any_nan = false;
    bool f32_bad = false;  // a far-wing coefficient of some layer is beyond Float32 range
    for (int l = L + lane; l < A.Ls; l += WAVE) A.rec64[i * A.Ls + l] = make_double4(0, 0, 0, 0);
    const int Lm = A.n_lblocks * WAVE;
    for (int l = lane; l < Lm; l += WAVE) {
        double c[4] = {0.0, 0.0, 0.0, 0.0};
        if (l < L) {
            const LayerTab tb = A.tab[(int64_t)st * L + l];
            const double T = S.temps[l];
            // src/line_absorption.jl:86-106
            const double sigma = doppler_width(wl, T, m, S.xi[l]);
            double G = gamma_rad;
            if (!mol) {
                G = G + S.n_e[l] * (gamma_stark * tb.t16);
                const double vdw = (vdW2 == -1.0) ? vdW1 * tb.t03 : scaled_vdW_abo(vdW1, vdW2, m, T);
                G = G + tb.n_eff * vdw;
            }
            const double gamma = G * (wl * wl) / (c_cgs * (4.0 * PI));
            const double levels = exp(-tb.beta * E_lower) - exp(-tb.beta * E_upper);
            const double amp = gf_sl * levels * S.n_div_U[(int64_t)s * L + l];
            // src/line_absorption.jl:108-112
            const double* cv = S.cntm + (int64_t)l * S.n_knots;
            const double cn = (S.cntm_mode == KORG_CNTM_GRID)
                                  ? (1.0 - kt) * cv[kidx] + kt * cv[kidx + 1]  // linear between the knots, weight kt
                                  : S.cntm[li * L + l];
            const double rho = cn * S.cutoff / amp;
            const double dg = inverse_gaussian_density(rho, sigma);
            const double dl = inverse_lorentz_density(rho, gamma);
            any_nan |= (dg != dg) || (dl != dl);
            dmax = dg > dmax ? dg : dmax;
            lmax = dl > lmax ? dl : lmax;
            // profile parameters of line_profile (src/line_absorption.jl:245-249)
            const double s2 = sigma * 1.4142135623730951;  // sigma sqrt 2
            const double inv = 1.0 / s2;
            const double a = gamma * inv;
            const double scaling = inv * INV_SQRT_PI * amp;
            A.rec64[i * A.Ls + l] = make_double4(inv, a, scaling, 0.0);
            // Zone radii of this layer, as v.  Near zone: the Voigt core (branches B and C) and,
            // for a > 0.2, where the 4-term far-wing series is not converged to SERIES_EPS.  Tail
            // radius: beyond it |e3| w^2 + |e4| w^3 <= SERIES_TAIL_EPS (w = 1/v^2; e_k the series
            // coefficients relative to the leading term).  They only place the switch between
            // exact evaluation and the series, so the roots are taken in Float32 and enlarged to
            // stay on the safe side.
            double vr = 5.0, vt = VT_SMALL_A;
            if (!(a <= 0.2)) {
                const double a2 = a * a;
                const double e5 = fabs(94.25 * a2 * a2 - 18.0 * a2 * a2 * a2 + a2 * a2 * a2 * a2);
                const double e3 = fabs(3.75 - 5.0 * a2 + a2 * a2);
                const double e4 = fabs(-26.25 * a2 + 10.5 * a2 * a2 - a2 * a2 * a2);
                vr = fmax(fmax(root_up(e5 * (1.0 / SERIES_EPS), 0.125f), 2.0 * a), 3.2);
                vt = fmax(root_up(e3 * (2.0 / SERIES_TAIL_EPS), 0.25f),
                          root_up(e4 * (2.0 / SERIES_TAIL_EPS), 1.0f / 6.0f));
            }
            const double r = vr * s2;
            rmax = r > rmax ? r : rmax;
            if (F32 && A.azone) {
                // adjoint near zone: the dropped term e5 w^5 of the series, differentiated,
                // (a e5)' w^4 and 5 e5 w^4 (against the leading term's), below SERIES_ADJ_EPS
                double va = vr;
                if (!(a <= 0.2)) {
                    const double a4 = (a * a) * (a * a);
                    const double e5 = fabs(a4 * (94.25 - 18.0 * a * a + a4));
                    const double f5 = fabs(a4 * (471.25 - 126.0 * a * a + 9.0 * a4));
                    va = fmax(va, root_up(fmax(f5, 10.0 * e5) * (1.0 / SERIES_ADJ_EPS), 0.125f));
                }
                amax = fmax(amax, va * s2);
            }
            const double rt = vt * s2;
            tmax = rt > tmax ? rt : tmax;
            if (F32) {
                const double sq = s2 * (1.0 / DELTA_UNIT);
                const double q = sq * sq;  // 1 / (DELTA_UNIT inv)^2
                const double base = scaling * (a * INV_SQRT_PI);
                double e2 = 1.5, e3 = 3.75, e4 = 0.0;
                if (!(a <= 0.2)) {
                    const double a2 = a * a;
                    e2 = 1.5 - a2;
                    e3 = 3.75 - 5.0 * a2 + a2 * a2;
                    e4 = -26.25 * a2 + 10.5 * a2 * a2 - a2 * a2 * a2;
                }
                c[0] = base * q;
                c[1] = base * e2 * q * q;
                c[2] = base * e3 * q * q * q;
                c[3] = base * e4 * (q * q) * (q * q);
#pragma unroll
                for (int k = 0; k < 4; ++k) f32_bad |= !(fabs(c[k]) <= F32_COEF_MAX);  // or NaN
            }
        }
        if (F32 && A.recM64) {
            A.recM64[(i * A.n_lblocks + l / WAVE) * WAVE + l % WAVE] = make_double4(c[0], c[1], c[2], c[3]);
        } else if (F32 && WING_BF16) {
            // bfloat16 form: per (line, layer) the four coefficients' high parts, then their low
            // parts (c - high, rounded again): 16 B at [(line * n_lblocks + block) * 64 + l % 64]
            uint32_t hw[2], lw[2];
#pragma unroll
            for (int k = 0; k < 4; k += 2) {
                const float x0 = (float)c[k], x1 = (float)c[k + 1];
                const __bf16 h0 = (__bf16)x0, h1 = (__bf16)x1;
                const __bf16 g0 = (__bf16)(x0 - (float)h0), g1 = (__bf16)(x1 - (float)h1);
                hw[k / 2] = (uint32_t)__builtin_bit_cast(uint16_t, h0) |
                            ((uint32_t)__builtin_bit_cast(uint16_t, h1) << 16);
                lw[k / 2] = (uint32_t)__builtin_bit_cast(uint16_t, g0) |
                            ((uint32_t)__builtin_bit_cast(uint16_t, g1) << 16);
            }
            ((uint4*)A.recM)[(i * A.n_lblocks + l / WAVE) * WAVE + l % WAVE] =
                make_uint4(hw[0], hw[1], lw[0], lw[1]);
        } else if (F32) {
            // MFMA A-operand layout: lane i of step s holds coefficient k = 2s + i/32 of layer
            // 32 mb + i%32; slot 2s + mb of float4 i.
            float* dst = (float*)(A.recM + (i * A.n_lblocks + l / WAVE) * WAVE);
            const int ll = l % WAVE, mb = ll >> 5, row = ll & 31;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int li = (k & 1) * 32 + row;
                const int slot = (k >> 1) * 2 + mb;
                dst[li * 4 + slot] = (float)c[k];
            }
        }
    }
    // Extreme damping (a ~ 1e6 and more, where e4 ~ a^6) puts the series coefficients out of
    // Float32 range: c_k R^k stays small where the series is used, but neither factor is
    // representable (and inf * 0 is NaN).  Such a line skips the series: its coefficients are
    // zero and the near-zone kernel evaluates its whole window exactly.
    const bool series_off = F32 && !A.recM64 && ballot(f32_bad) != 0;
    if (series_off) {
        // (either layout: every coefficient slot of the line's blocks is zeroed)
        for (int l = lane; l < Lm; l += WAVE)
            A.recM[(i * A.n_lblocks + l / WAVE) * WAVE + l % WAVE] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    // src/line_absorption.jl:109-115: windows are maxima over layers (NaN if any is NaN)
    const bool wnan = ballot(any_nan) != 0;
    const double dwin = wave_max(dmax);
    const double lwin = wave_max(lmax);
    const double reach = wave_max(rmax);
    const double treach = wave_max(tmax);
    const double areach = (F32 && A.azone) ? wave_max(amax) : -1.0;
    // each 16-layer block's own adjoint radius (lanes 16q .. 16q + 15 hold block q's layers)
    double ablk[4] = {areach, areach, areach, areach};
    if (F32 && A.azq && A.n_lblocks == 1) {
        double m16 = amax;
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) m16 = fmax(m16, __shfl_xor(m16, off, WAVE));
#pragma unroll
        for (int q = 0; q < 4; ++q) ablk[q] = readlane_f64(m16, 16 * q);
    }
    const double window = wnan ? __builtin_nan("") : sqrt(lwin * lwin + dwin * dwin);
    if (lane == 0) {
        LineWin w;
        w.window = window;
        w.reach = reach;
        w.treach = treach;
        w.areach = areach;
#pragma unroll
        for (int q = 0; q < 4; ++q) w.ablk[q] = ablk[q];
        w.series_off = series_off ? 1 : 0;
        w.pad = 0;
        A.win[i] = w;
    }
}

// k_setup_post: one lane per line.  Lower/upper bounds on the grid of the window's edges, the
// centre, and the near, tail and adjoint zones' edges (src/line_absorption.jl:109-115), clipped
// as the zones nest; the far-field runs, the window class and the LineMeta.
template <bool F32>
__global__ __launch_bounds__(256) void k_setup_post(SetupArgs A) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n_lines) return;
    const LinePre P = A.pre[i];
    if (P.flags & 2) return;  // invalid species: LineMeta written by k_setup_pre
    const LineWin w = A.win[i];
    const StarDev& S = A.stars[P.st];
    const int64_t li = i - S.line_off;
    const GridInfo gi = S.gi;
    const double wl = P.wl;
    const double window = w.window, reach = w.reach, treach = w.treach, areach = w.areach;
    const bool series_off = w.series_off != 0;
    const int64_t n = S.n_wl;
    const double last = S.wls[n - 1];
    const double h = reach * (1.0 + 1e-6) + 1e-6 * DELTA_UNIT;
    const double ht = treach * (1.0 + 1e-6) + 1e-6 * DELTA_UNIT;
    const double ha = areach * (1.0 + 1e-6) + 1e-6 * DELTA_UNIT;
    // x beyond the grid (or NaN): lower bound n, upper bound n (last index n - 1)
    auto bound = [&](double x, bool upper) -> int64_t {
        if (!(x <= last)) return n;
        return lane_bound(S.wls, n, gi.g0, gi.ginv, x, upper);
    };
    double xlo = wl - window, xhi = wl + window;
    if (xlo >= 1.0) xlo *= 1e-8;  // src/wavelengths.jl:189,201: >= 1 is Angstrom
    if (xhi >= 1.0) xhi *= 1e-8;
    int64_t lb = bound(xlo, false);
    int64_t ub = bound(xhi, true) - 1;
    const int64_t jc = bound(wl, false);
    if (S.win_lb) {  // frozen windows (1-based), clipped to the grid; zones are clipped to them
        lb = S.win_lb[li] - 1;
        ub = S.win_ub[li] - 1;
        lb = lb < 0 ? 0 : lb;
        ub = ub > n - 1 ? n - 1 : ub;
    }
    int64_t clo = 1, chi = 0;
    if (lb <= ub && reach >= 0.0) {
        clo = bound(wl - h, false);
        chi = bound(wl + h, true) - 1;
        clo = clo < lb ? lb : clo;
        chi = chi > ub ? ub : chi;
    }
    int64_t flo = 1, fhi = 0;
    if (lb <= ub && treach >= 0.0) {
        flo = bound(wl - ht, false);
        fhi = bound(wl + ht, true) - 1;
        flo = flo < lb ? lb : flo;
        fhi = fhi > ub ? ub : fhi;
    }
    int64_t alo = clo, ahi = chi;  // the adjoint's near zone contains the forward's
    if (F32 && A.azone && lb <= ub && areach >= 0.0) {
        alo = min(bound(wl - ha, false), clo);
        ahi = max(bound(wl + ha, true) - 1, chi);
        alo = alo < lb ? lb : alo;
        ahi = ahi > ub ? ub : ahi;
    }
    if (series_off && lb <= ub) {
        clo = lb;
        chi = ub;
        flo = 1;
        fhi = 0;
        alo = lb;
        ahi = ub;
    }
    // The far-field runs, in local sub-tile indices q (points 64q .. 64q + 63, whole sub-tiles
    // only): inside the window, clear of the near zone [zlo, zhi], and with the sub-tile's centre
    // 64q + 31.5 at least 8 * 31.5 + 1 points from the line centre x0 (the grid is evenly spaced
    // to 1e-3 of a step, so the expansion's ratio stays above 8 * 252/253).  The lower run starts
    // at the first whole sub-tile of the window, the upper one ends at the last; an empty run has
    // its end before its start.  The forward's runs avoid its near zone, the adjoint's (Float32
    // fit step) the adjoint's wider one.
    const bool ff_ok = lb <= ub && S.ff_irregular == 0 && gi.step > 0.0;
    auto ff_runs = [&](int64_t zlo, int64_t zhi) {
        const double x0 = (wl - S.wls[0]) / gi.step;  // the line centre in points
        const int64_t q0 = (lb + FF_SUB - 1) / FF_SUB;                 // first whole sub-tile
        const int64_t q1 = (ub + 1) / FF_SUB - 1;                      // last one
        const double reach = FF_RATIO * 0.5 * (FF_SUB - 1) + 1.0;      // points, centre to centre
        int64_t lo_end = (int64_t)floor((x0 - reach - 0.5 * (FF_SUB - 1)) / FF_SUB);
        int64_t hi_start = (int64_t)ceil((x0 + reach - 0.5 * (FF_SUB - 1)) / FF_SUB);
        if (zlo <= zhi) {
            lo_end = min(lo_end, (int64_t)floor((double)(zlo - FF_SUB) / FF_SUB));  // 64q + 63 < zlo
            hi_start = max(hi_start, zhi / FF_SUB + 1);                          // 64q > zhi
        }
        lo_end = min(lo_end, q1);
        hi_start = max(hi_start, max(q0, lo_end + 1));
        // runs of at most 65,535 sub-tiles (the list keys hold their lengths in 16 bits)
        lo_end = min(lo_end, q0 + 0xFFFE);
        hi_start = max(hi_start, q1 - 0xFFFE);
        const int64_t gq = S.grid_off / FF_SUB;
        return make_int4((int)(q0 + gq), (int)(lo_end + gq), (int)(hi_start + gq), (int)(q1 + gq));
    };
    const int4 run = A.ffrun && ff_ok ? ff_runs(clo, chi) : make_int4(1, 0, 1, 0);
    const int4 run_adj = F32 && A.ffrun_adj && ff_ok ? ff_runs(alo, ahi) : make_int4(1, 0, 1, 0);
    // each 16-layer block's adjoint near zone inside [alo, ahi] (the whole of it where the block
    // has no narrower radius): beyond it the block's layers take the moments' series
    int4 azq = make_int4(0, 0, 0, 0);
    if (F32 && A.azq && !series_off && lb <= ub && alo <= ahi) {
        int e[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            e[q] = 0;
            if (w.ablk[q] >= 0.0 && w.ablk[q] < areach) {
                const double hq = w.ablk[q] * (1.0 + 1e-6) + 1e-6 * DELTA_UNIT;
                const int64_t qlo = max(bound(wl - hq, false), alo);
                const int64_t qhi = min(bound(wl + hq, true) - 1, ahi);
                e[q] = (int)((uint32_t)min(qlo - alo, (int64_t)0xFFFF) |
                             ((uint32_t)min(ahi - qhi, (int64_t)0xFFFF) << 16));
            }
        }
        azq = make_int4(e[0], e[1], e[2], e[3]);
    }
    int cls = N_CLASSES;
    if (lb <= ub) {
        const int64_t W = ub - lb + 1;
        cls = 0;
        while (cls < N_CLASSES - 1 && W > class_wmax(cls)) ++cls;
    }
    {
        // grid indices are stored global (offset by the star's place in the concatenated grids)
        const int64_t go = S.grid_off;
        LineMeta mt;
        mt.lam0 = wl;
        mt.lb = (int32_t)(lb + go);
        mt.ub = (int32_t)(ub + go);
        mt.clo = (int32_t)(clo + go);
        mt.chi = (int32_t)(chi + go);
        mt.jc = (int32_t)(jc + go);
        mt.cls = cls;
        mt.flo = (int32_t)(flo + go);
        mt.fhi = (int32_t)(fhi + go);
        A.meta[i] = mt;
        if (F32 && A.azone) A.azone[i] = make_int2((int)(alo + go), (int)(ahi + go));
        if (F32 && A.azq) A.azq[i] = azq;
        if (A.ffrun) A.ffrun[i] = run;
        if (F32 && A.ffrun_adj) A.ffrun_adj[i] = run_adj;
    }
}

// Evaluations by Hunger branch (the count_branches diagnostic): one wavefront per line, lanes =
// layers; a layer's core points (branches B and C) lie within its half-width hw of the centre:
// 5 Doppler units for a <= 0.2, 3.2 - a for a <= 1.4, none beyond.  Counted into counters[6..9]
// (A, B, C, D) and [15] (branch-A evaluations inside the near zone).
template <bool F32>
__global__ __launch_bounds__(256) void k_branch_count(SetupArgs A) {
    const int lane = threadIdx.x % WAVE;
    const int64_t i = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (i >= A.n_lines || lane >= A.L) return;
    const LinePre P = A.pre[i];
    if (P.flags & 2) return;
    const StarDev& S = A.stars[P.st];
    const LineMeta mt = A.meta[i];
    const int64_t go = S.grid_off;
    const int64_t lb = mt.lb - go, ub = mt.ub - go, clo = mt.clo - go, chi = mt.chi - go;
    if (lb > ub) return;
    const double4 r = A.rec64[i * A.Ls + lane];
    const double inv = r.x, a = r.y;
    const double hw = (a <= 0.2) ? 5.0 / inv : (a <= 1.4 ? (3.2 - a) / inv : -1.0);
    const int64_t n = S.n_wl;
    const GridInfo gi = S.gi;
    int64_t core_n = 0;
    if (hw >= 0.0) {
        int64_t c0 = 0, c1 = 0;
        if (!lane_bound_guess(S.wls, n, gi.g0, gi.ginv, P.wl - hw, true, c0))
            c0 = lb;  // counting only: fall back to the window edge
        if (!lane_bound_guess(S.wls, n, gi.g0, gi.ginv, P.wl + hw, false, c1))
            c1 = ub + 1;
        c1 -= 1;
        c0 = c0 < lb ? lb : c0;
        c1 = c1 > ub ? ub : c1;
        core_n = c1 >= c0 ? c1 - c0 + 1 : 0;
    }
    const int64_t wing_n = (ub - lb + 1) - core_n;
    const bool small_a = a <= 0.2;
    atomicAdd(&A.counters[small_a ? 6 : 9], (unsigned long long)wing_n);
    if (core_n) atomicAdd(&A.counters[small_a ? 7 : 8], (unsigned long long)core_n);
    // branch-A evaluations inside the near zone (which holds every core point): the rest of
    // the near zone past the core of a layer with a <= 0.2
    if (F32 && small_a && clo <= chi) {
        const int64_t near_wing = (chi - clo + 1) - core_n;
        if (near_wing > 0) atomicAdd(&A.counters[15], (unsigned long long)near_wing);
    }
}

constexpr int PART_BLOCK = 256;

// Partition keys.  WINDOW: lines by window width (for the profile kernels' tiles); NEAR: lines
// by near-zone reach, the largest distance in points from the centre index jc to a near-zone
// point (for k_near's tiles).  Both fit the same class table: a line of class c can only touch
// tiles within class_wmax(c) points of its jc.
enum { KEY_WINDOW = 0, KEY_NEAR = 1 };

__device__ inline int line_class(const LineMeta& m, int key) {
    if (key == KEY_WINDOW) return m.cls;
    if (m.clo > m.chi) return N_CLASSES;
    const int64_t d = max(m.jc - m.clo, m.chi - m.jc) + 1;
    int c = 0;
    while (c < N_CLASSES - 1 && d > class_wmax(c)) ++c;
    return c;
}

// Per block of lines: class counts for the partition, plus the call's totals (evaluations, lines
// used, and the largest distance in points from a line's centre index to its near zone), reduced
// in LDS and added with one atomic per block.
__global__ __launch_bounds__(PART_BLOCK) void k_count(int64_t n, int L, int wing_tile_w,
                                                      const LineMeta* __restrict__ meta,
                                                      int* __restrict__ block_counts,
                                                      int* __restrict__ near_counts,
                                                      unsigned long long* __restrict__ counters) {
    __shared__ int cnt[N_CLASSES + 1], ncnt[N_CLASSES + 1];
    __shared__ unsigned long long evals, used, reach, near_evals, wing_mfma, tail_evals;
    if (threadIdx.x <= N_CLASSES) cnt[threadIdx.x] = ncnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) evals = used = reach = near_evals = wing_mfma = tail_evals = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
    int c = N_CLASSES, cn = N_CLASSES;
    if (i < n) {
        const LineMeta m = meta[i];
        c = m.cls;
        cn = line_class(m, KEY_NEAR);
        if (m.lb <= m.ub) {
            atomicAdd(&evals, (unsigned long long)(m.ub - m.lb + 1) * (unsigned long long)L);
            atomicAdd(&used, 1ull);
            // MFMAs of the far-wing kernel for this line and one 64-layer block: 2 per 32-point
            // block of every tile meeting the window, 2 more where the tile meets the tail zone
            const int tiles = m.ub / wing_tile_w - m.lb / wing_tile_w + 1;
            const int full = m.flo <= m.fhi ? m.fhi / wing_tile_w - m.flo / wing_tile_w + 1 : 0;
            atomicAdd(&wing_mfma, (unsigned long long)(wing_tile_w / 16) * (tiles + full));
            // far-zone evaluations that use the R^3, R^4 terms: tail zone less near zone
            if (m.flo <= m.fhi) {
                int64_t t = m.fhi - m.flo + 1;
                if (m.clo <= m.chi) t -= max(0, min(m.chi, m.fhi) - max(m.clo, m.flo) + 1);
                atomicAdd(&tail_evals, (unsigned long long)t * (unsigned long long)L);
            }
        }
        if (m.clo <= m.chi) {
            const int d = max(m.jc - m.clo, m.chi - m.jc) + 1;
            atomicMax(&reach, (unsigned long long)d);
            atomicAdd(&near_evals, (unsigned long long)(m.chi - m.clo + 1) * (unsigned long long)L);
        }
    }
    for (int k = 0; k < N_CLASSES; ++k) {
        const uint64_t b = ballot(c == k);
        if ((threadIdx.x % WAVE) == 0 && b) atomicAdd(&cnt[k], __popcll(b));
        const uint64_t bn = ballot(cn == k);
        if ((threadIdx.x % WAVE) == 0 && bn) atomicAdd(&ncnt[k], __popcll(bn));
    }
    __syncthreads();
    if (threadIdx.x < N_CLASSES) {
        block_counts[(int64_t)blockIdx.x * N_CLASSES + threadIdx.x] = cnt[threadIdx.x];
        near_counts[(int64_t)blockIdx.x * N_CLASSES + threadIdx.x] = ncnt[threadIdx.x];
    }
    if (threadIdx.x == 0) {
        if (evals) atomicAdd(&counters[1], evals);
        if (used) atomicAdd(&counters[2], used);
        if (reach) atomicMax(&counters[4], reach);
        if (near_evals) atomicAdd(&counters[10], near_evals);
        if (wing_mfma) atomicAdd(&counters[11], wing_mfma);
        if (tail_evals) atomicAdd(&counters[12], tail_evals);
    }
}

// One block: exclusive scan of block_counts ([block][NC]) per class; class_start[c] = first slot of
// class c, class_start[NC] = total.
template <int NC>
__global__ __launch_bounds__(1024) void k_scan(int nblocks, int* __restrict__ block_counts,
                                               int* __restrict__ class_start) {
    __shared__ int part[1024];
    __shared__ int totals[NC];
    const int t = threadIdx.x;
    for (int c = 0; c < NC; ++c) {
        // each thread owns a contiguous run of blocks
        const int per = (nblocks + 1023) / 1024;
        const int b0 = t * per, b1 = min(nblocks, b0 + per);
        int s = 0;
        for (int b = b0; b < b1; ++b) s += block_counts[(int64_t)b * NC + c];
        part[t] = s;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int v = t >= off ? part[t - off] : 0;
            __syncthreads();
            part[t] += v;
            __syncthreads();
        }
        int run = part[t] - s;  // exclusive prefix of this thread's run
        for (int b = b0; b < b1; ++b) {
            const int v = block_counts[(int64_t)b * NC + c];
            block_counts[(int64_t)b * NC + c] = run;
            run += v;
        }
        if (t == 1023) totals[c] = part[t];
        __syncthreads();
    }
    if (t == 0) {
        int acc = 0;
        for (int c = 0; c < NC; ++c) {
            class_start[c] = acc;
            acc += totals[c];
        }
        class_start[NC] = acc;
    }
}

__global__ __launch_bounds__(PART_BLOCK) void k_scatter(int64_t n, const LineMeta* __restrict__ meta,
                                                        const int* __restrict__ block_offsets,
                                                        const int* __restrict__ class_start,
                                                        Cand* __restrict__ cands,
                                                        int* __restrict__ cls_jc,
                                                        int2* __restrict__ keyrng,
                                                        int* __restrict__ part_line, int key,
                                                        const int4* __restrict__ ffrun,
                                                        int4* __restrict__ ffkey) {
    __shared__ int wave_cnt[PART_BLOCK / WAVE][N_CLASSES];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int64_t i = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
    const int c = i < n ? line_class(meta[i], key) : N_CLASSES;
    int rank = 0;
    for (int k = 0; k < N_CLASSES; ++k) {
        const uint64_t b = ballot(c == k);
        if (c == k) rank = __popcll(b & ((1ull << lane) - 1));
        if (lane == 0) wave_cnt[w][k] = __popcll(b);
    }
    __syncthreads();
    if (c < N_CLASSES) {
        int before = 0;
        for (int ww = 0; ww < w; ++ww) before += wave_cnt[ww][c];
        const int pos = class_start[c] + block_offsets[(int64_t)blockIdx.x * N_CLASSES + c] + before + rank;
        const LineMeta m = meta[i];
        Cand cd;
        cd.lam0 = m.lam0;
        cd.lb = m.lb;
        cd.ub = m.ub;
        cd.clo = m.clo;
        cd.chi = m.chi;
        cd.flo = m.flo;
        cd.fhi = m.fhi;
        cd.line = (int)i;
        cd.pad = 0;
        cands[pos] = cd;
        cls_jc[pos] = m.jc;
        // the key's grid-index range and the line index, compact, for the tile-list passes
        keyrng[pos] = key == KEY_WINDOW ? make_int2(m.lb, m.ub) : make_int2(m.clo, m.chi);
        part_line[pos] = (int)i;
        if (ffrun) ffkey[pos] = ff_key(m.lb, m.ub, ffrun[i], (int)i);
    }
}

// Per (class, tile): candidate lines are those of the class whose centre index jc lies in
// [t0 - Wmax, t1 + Wmax]; an overlapping line always does (jc lies in [lb, ub + 1]).
__global__ void k_ranges(int ntiles, int tile_w, int64_t n_wl, const int* __restrict__ class_start,
                         const int* __restrict__ cls_jc, int2* __restrict__ ranges) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles * N_CLASSES) return;
    const int c = t / ntiles, tile = t % ntiles;
    const int64_t t0 = (int64_t)tile * tile_w;
    const int64_t t1 = min(t0 + tile_w, n_wl) - 1;
    const int64_t wmax = class_wmax(c);
    const int64_t lo_key = t0 - wmax, hi_key = t1 + wmax;
    const int s0 = class_start[c], s1 = class_start[c + 1];
    int lo = s0, hi = s1;
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if ((int64_t)cls_jc[mid] < lo_key) lo = mid + 1; else hi = mid;
    }
    const int b = lo;
    hi = s1;
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if ((int64_t)cls_jc[mid] <= hi_key) lo = mid + 1; else hi = mid;
    }
    ranges[(int64_t)c * ntiles + tile] = make_int2(b, lo);
}

// Exact per-tile line lists.  One wavefront per (tile, class) scans the class's candidates for
// the tile (the range from k_ranges) and keeps those whose key range meets the tile -- with the
// far field on (ffkey given), except where the line's far-field runs cover the tile.
// The count pass writes the number per (tile, class); after an exclusive scan over (tile, class)
// in tile-major order, the fill pass writes the line indices: for every tile, class by class and
// in partition order within a class -- the same fixed order as the candidate scans.
// With the far field the same scan also builds the far-field lists, in slots ntiles * N_CLASSES
// + (tile, class): a line joins the list of the super-tile (FF_NS sub-tiles) holding this tile
// from the tile holding the first sub-tile of the super-tile that its runs cover -- once per
// (line, super-tile).  A super-tile's list is then its tiles' slots, in tile, class and
// partition order.
__global__ __launch_bounds__(WAVE) void k_tile_lists(int ntiles, int tile_w, int64_t n_wl,
                                                     const int2* __restrict__ ranges,
                                                     const int2* __restrict__ keyrng,
                                                     const int* __restrict__ part_line,
                                                     const int4* __restrict__ ffkey,
                                                     const int64_t* __restrict__ tc_start,
                                                     int* __restrict__ tc_count,
                                                     int* __restrict__ lists,
                                                     const unsigned long long* __restrict__ counters) {
    if (tc_start && counters[3] != 0) return;  // fill pass: the lists do not fit (or setup failed)
    const int lane = threadIdx.x;
    const int t = blockIdx.x;  // c * ntiles + tile
    const int c = t / ntiles, tile = t % ntiles;
    const int64_t j0 = (int64_t)tile * tile_w;
    const int64_t j1 = min(j0 + tile_w, n_wl) - 1;
    const int2 r = ranges[t];
    const int64_t slot = (int64_t)tile * N_CLASSES + c;
    const int64_t fslot = (int64_t)ntiles * N_CLASSES + slot;
    const int64_t base = tc_start ? tc_start[slot] : 0;
    const int64_t fbase = tc_start && ffkey ? tc_start[fslot] : 0;
    // 32-bit indices from here (grids hold fewer than 2^31 points): the tile's points and
    // sub-tiles, and the sub-tiles of its super-tile
    const int i0 = (int)j0, i1 = (int)j1;
    const int ga = i0 / FF_SUB, gz = i1 / FF_SUB;
    const int s0 = ga / FF_NS * FF_NS, s1 = s0 + FF_NS - 1;
    int64_t cnt = 0, fcnt = 0;
    const uint64_t below = (1ull << lane) - 1;
    if (ffkey) {
        // software-pipelined: the next batch's key (and, filling, line index) is loaded while
        // this batch is tested
        const int4 none = make_int4(1, 0, 0, 0);
        int4 key_n = r.x + lane < r.y ? ffkey[r.x + lane] : none;
        for (int k0 = r.x; k0 < r.y; k0 += WAVE) {
            const int4 key = key_n;
            const int line = key.w;
            const int kn = k0 + WAVE + lane;
            key_n = kn < r.y ? ffkey[kn] : none;  // past the end: no window
            bool ovl = false, far = false;
            if (key.x <= key.y && key.x <= i1 && key.y >= i0) {  // the window meets the tile
                const int4 run = ff_key_runs(key);
                ovl = !((ga >= run.x && gz <= run.y) || (ga >= run.z && gz <= run.w));
                // the first sub-tile of the super-tile in a run; the line joins the far-field
                // list from the tile holding that sub-tile's first point
                int f = max(run.x, s0);
                if (f > min(run.y, s1)) {
                    f = max(run.z, s0);
                    if (f > min(run.w, s1)) f = INT_MAX;
                }
                far = f != INT_MAX && f * FF_SUB >= i0 && f * FF_SUB <= i1;
            }
            const uint64_t b = ballot(ovl);
            const uint64_t fb = ballot(far);
            if (lists && ovl) lists[base + cnt + __popcll(b & below)] = line;
            if (lists && far) lists[fbase + fcnt + __popcll(fb & below)] = line;
            cnt += __popcll(b);
            fcnt += __popcll(fb);
        }
    } else {
        for (int k0 = r.x; k0 < r.y; k0 += WAVE) {
            const int k = k0 + lane;
            bool ovl = false;
            if (k < r.y) {
                const int2 kr = keyrng[k];
                ovl = kr.x <= kr.y && kr.x <= j1 && kr.y >= j0;
            }
            const uint64_t b = ballot(ovl);
            if (lists && ovl) lists[base + cnt + __popcll(b & below)] = part_line[k];
            cnt += __popcll(b);
        }
    }
    if (!tc_start && lane == 0) {
        tc_count[slot] = (int)cnt;
        if (ffkey) tc_count[fslot] = (int)fcnt;
    }
}

// The scans are one wavefront per workgroup: they run beside the near zone's kernel on the other
// stream, whose single-wave workgroups take every slot that frees up, so a scan block that needed
// several waves' slots on one CU at once would wait for that whole kernel.

// Inclusive scan across the wavefront.
__device__ inline int64_t wave_incl_scan(int64_t x, int lane) {
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const int64_t y = __shfl_up((long long)x, off, WAVE);
        if (lane >= off) x += y;
    }
    return x;
}

// One wavefront: exclusive scan of n counts into start[0..n] (start[n] = total; start may be
// count itself), 64 counts per step in order, each step's total carried into the next.  With
// counters given, the total is recorded in counters[slot] and FLAG_CAPACITY is raised when it
// exceeds cap.
template <class T>
__global__ __launch_bounds__(WAVE) void k_scan_flat(int64_t n, const T* count, int64_t* start,
                                                    int64_t cap, unsigned long long* __restrict__ counters,
                                                    int slot) {
    const int t = threadIdx.x;
    int64_t run = 0;
    for (int64_t c0 = 0; c0 < n; c0 += WAVE) {
        const int64_t i = c0 + t;
        const int64_t v = i < n ? (int64_t)count[i] : 0;
        const int64_t inc = wave_incl_scan(v, t);
        if (i < n) start[i] = run + inc - v;
        run += __shfl((long long)inc, WAVE - 1, WAVE);
    }
    if (t == 0) {
        start[n] = run;
        if (counters) {
            counters[slot] = (unsigned long long)run;
            if (run > cap) atomicOr(&counters[3], FLAG_CAPACITY);
        }
    }
}

// The same scan over many workgroups for long inputs: block totals of SCAN_TILE counts
// (k_scan_part), their scan in place (k_scan_flat<int64_t>, which also raises the capacity flag),
// and each block's scan from its offset (k_scan_fill).  Reads are coalesced.
constexpr int SCAN_TILE = 1024;  // counts per workgroup: 16 steps of 64

__global__ __launch_bounds__(WAVE) void k_scan_part(int64_t n, const int* __restrict__ count,
                                                    int64_t* __restrict__ part) {
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int64_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_TILE / WAVE; ++k) {
        const int64_t i = base + k * WAVE + t;
        if (i < n) sum += count[i];
    }
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) sum += __shfl_xor((long long)sum, off, WAVE);
    if (t == 0) part[blockIdx.x] = sum;
}

__global__ __launch_bounds__(WAVE) void k_scan_fill(int64_t n, const int* __restrict__ count,
                                                    const int64_t* __restrict__ part,
                                                    int64_t* __restrict__ start) {
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    constexpr int STEPS = SCAN_TILE / WAVE;
    int v[STEPS];
#pragma unroll
    for (int k = 0; k < STEPS; ++k) {
        const int64_t i = base + k * WAVE + t;
        v[k] = i < n ? count[i] : 0;
    }
    int64_t run = part[blockIdx.x];
#pragma unroll
    for (int k = 0; k < STEPS; ++k) {
        const int64_t i = base + k * WAVE + t;
        const int64_t inc = wave_incl_scan((int64_t)v[k], t);
        if (i < n) start[i] = run + inc - v[k];
        run += __shfl((long long)inc, WAVE - 1, WAVE);
    }
    if (blockIdx.x == 0 && t == 0) start[n] = part[gridDim.x];
}

// --------------------------------------------------------------------------------------------
// profile kernels
// --------------------------------------------------------------------------------------------
struct ProfileArgs {
    int64_t n_wl;
    int L;
    int Ls;          // rec64 stride per line
    int n_lblocks;   // layer blocks of a work item (F64_GROUP layers, or 64 for the MFMA kernel)
    int ntiles, tile_w;
    int64_t n_items;
    const StarDev* stars;  // grid values and alpha of each star
    int n_stars;
    const double4* rec64;
    const float4* recM;
    const double4* recM64;  // the Float64 series form's coefficients (recM's layout, Float64)
    const Cand* cands;   // lines in window-class partition order
    const int2* ranges;
    const LineMeta* meta;
    const int64_t* tile_start;  // [tile * N_CLASSES]: first entry of the tile's list (far wing)
    const int* tile_lines;      // per tile, the lines whose window meets it (far wing)
    // far field (null fft: off): sub-tile geometry and the far-field sums T of every sub-tile
    const FfGeo* geo;
    const float* fft;
    const int4* ffrun;  // each line's far-field runs
    int64_t n_sub;
    int alpha_f32;
    unsigned long long* counters;
};

// The wavefront-uniform fields of one line, taken from the candidate lane that loaded them.
struct LineView {
    int line;
    double lam0;
    int lb, ub, clo, chi, flo, fhi;
};

// s_waitcnt operand that waits for every vector memory load (vmcnt = 0) and leaves the export
// and LDS/scalar counters unconstrained.
constexpr int VMCNT0 = 0x0F70;

// Lowest set lane of `mask`, removed from it; -1 when empty.
__device__ inline int pop_lane(uint64_t& mask) {
    if (!mask) return -1;
    const int b = __ffsll((unsigned long long)mask) - 1;
    mask &= mask - 1;
    return b;
}

// Visit, in a fixed order (class by class, partition order within a class), every line whose
// zone overlaps grid points [j0, j1] of this tile: the window [lb, ub] (KEY_WINDOW) or the near
// zone [clo, chi] (KEY_NEAR).  body(view, operand) is called with wavefront-uniform views.
// The per-line operand load(line) -> T is software-pipelined two lines deep: the load for line
// i + 2 is issued before body() runs for line i, and three operand registers rotate through a
// three-way unrolled loop, so no operand is copied and each wait is for the oldest load only.
// Every path through the loop issues the same loads (past the end of a batch, lane 63's line is
// loaded as a stand-in), which keeps the wait counts exact.  The next batch's 64 candidate
// records are loaded behind the first two operands of the current batch.  With nparts > 1 only
// share `part` of every class's candidates is visited (the shares of a tile's waves).
template <class T, class Load, class Body>
__device__ inline void for_each_line_pf(const int2* __restrict__ ranges, const Cand* __restrict__ cands,
                                        int ntiles, int tile, int64_t j0, int64_t j1, int lane,
                                        int key, Load&& load, Body&& body, int part = 0,
                                        int nparts = 1) {
    for (int c = 0; c < N_CLASSES; ++c) {
        int2 r = ranges[(int64_t)c * ntiles + tile];
        if (nparts > 1) {
            // part `part` of nparts contiguous, near-equal shares of the class's candidates
            const int len = r.y - r.x;
            r = make_int2(r.x + (int)(((int64_t)len * part) / nparts),
                          r.x + (int)(((int64_t)len * (part + 1)) / nparts));
        }
        if (r.x >= r.y) continue;
        Cand cc = cands[min(r.x + lane, r.y - 1)];
        for (int k0 = r.x; k0 < r.y; k0 += WAVE) {
            const bool valid = k0 + lane < r.y;
            const bool ovl = valid && (key == KEY_WINDOW
                                           ? (cc.lb <= j1 && cc.ub >= j0)
                                           : (cc.clo <= cc.chi && cc.clo <= j1 && cc.chi >= j0));
            uint64_t mask = ballot(ovl);
            const int kn = min(k0 + WAVE + lane, r.y - 1);  // next batch (clamped: always valid)
            Cand cn;
            if (!mask) {
                cn = cands[kn];
                cc = cn;
                continue;
            }
            auto view = [&](int b) {
                LineView v;
                v.line = readlane_i32(cc.line, b);
                v.lam0 = readlane_f64(cc.lam0, b);
                v.lb = readlane_i32(cc.lb, b);
                v.ub = readlane_i32(cc.ub, b);
                v.clo = readlane_i32(cc.clo, b);
                v.chi = readlane_i32(cc.chi, b);
                v.flo = readlane_i32(cc.flo, b);
                v.fhi = readlane_i32(cc.fhi, b);
                return v;
            };
            auto line_of = [&](int b) { return readlane_i32(cc.line, b >= 0 ? b : WAVE - 1); };
            int b0 = pop_lane(mask);
            int b1 = pop_lane(mask);
            T x0 = load(line_of(b0));
            T x1 = load(line_of(b1));
            cn = cands[kn];
            for (;;) {
                const int b2 = pop_lane(mask);
                const T x2 = load(line_of(b2));
                body(view(b0), x0);
                if (b1 < 0) break;
                const int b3 = pop_lane(mask);
                x0 = load(line_of(b3));
                body(view(b1), x1);
                if (b2 < 0) break;
                const int b4 = pop_lane(mask);
                x1 = load(line_of(b4));
                body(view(b2), x2);
                if (b3 < 0) break;
                b0 = b3;
                b1 = b4;
            }
            // retire the stand-in loads left in flight by the exit, so that no load into the
            // operand registers is pending when the next batch enters the loop
            __builtin_amdgcn_s_waitcnt(VMCNT0);
            cc = cn;
        }
    }
}

// for_each_line_pf for the near zone with a one-line-deep pipeline (the operand of the next
// visited line is loaded before body() runs for the current one): fewer operand registers than
// the two-deep form, for a kernel whose occupancy is register-bound.  Same visiting order.
template <class T, class Load, class Body>
__device__ inline void for_each_near_line_pf1(const int2* __restrict__ ranges,
                                              const Cand* __restrict__ cands, int ntiles, int tile,
                                              int64_t j0, int64_t j1, int lane, Load&& load,
                                              Body&& body) {
    for (int c = 0; c < N_CLASSES; ++c) {
        const int2 r = ranges[(int64_t)c * ntiles + tile];
        if (r.x >= r.y) continue;
        for (int k0 = r.x; k0 < r.y; k0 += WAVE) {
            const Cand cc = cands[min(k0 + lane, r.y - 1)];
            const bool ovl = k0 + lane < r.y && cc.clo <= cc.chi && cc.clo <= j1 && cc.chi >= j0;
            uint64_t mask = ballot(ovl);
            int b0 = pop_lane(mask);
            if (b0 < 0) continue;
            T x0 = load(readlane_i32(cc.line, b0));
            for (;;) {
                const int b1 = pop_lane(mask);
                // past the last line, this batch's first line again as a stand-in load
                const T x1 = load(readlane_i32(cc.line, b1 >= 0 ? b1 : b0));
                LineView v;
                v.line = readlane_i32(cc.line, b0);
                v.lam0 = readlane_f64(cc.lam0, b0);
                v.lb = v.ub = v.flo = v.fhi = 0;
                v.clo = readlane_i32(cc.clo, b0);
                v.chi = readlane_i32(cc.chi, b0);
                body(v, x0);
                if (b1 < 0) break;
                x0 = x1;
                b0 = b1;
            }
            // retire the stand-in load, so that no load into the operand registers is pending
            // when the next batch enters the loop
            __builtin_amdgcn_s_waitcnt(VMCNT0);
        }
    }
}

__device__ inline int wave_take(unsigned long long* counter, int lane) {
    unsigned long long v = 0;
    if (lane == 0) v = atomicAdd(counter, 1ull);
    return __builtin_amdgcn_readfirstlane((int)v);
}

template <class T>
__device__ inline void add_alpha(void* alpha, int f32, int64_t k, T v) {
    if (f32) ((float*)alpha)[k] += (float)v;
    else ((double*)alpha)[k] += (double)v;
}

// Float64 form: the reference's line_profile / voigt_hjerting per evaluation.  A work item is
// (tile, group of G layers); the G x P accumulators stay in registers.
template <int G, int P>
__global__ __launch_bounds__(WAVE) void k_profile_f64(ProfileArgs A) {
    const int lane = threadIdx.x;
    const int L = A.L;
    if (A.counters[3] != 0) return;  // failed setup: nothing is written
    for (;;) {
        const int item = wave_take(&A.counters[0], lane);
        if (item >= A.n_items) break;
        const int tile = item / A.n_lblocks;
        const int l0 = (item % A.n_lblocks) * G;
        const int64_t j0 = (int64_t)tile * A.tile_w;
        const KCONST StarDev* stars = uniform_ptr(A.stars);
        const KCONST StarDev& S = stars[star_of_point(stars, A.n_stars, j0)];
        const int64_t go = S.grid_off;
        const int64_t j1 = min(j0 + A.tile_w, go + S.n_wl) - 1;  // the star's last point at most
        if (j1 < j0) continue;  // padding between two stars' grids: no points
        double lam[P];
        int jj[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            jj[p] = (int)(j0 + p * WAVE + lane);
            lam[p] = S.wls[min((int64_t)jj[p], j1) - go];  // points past j1 are never used
        }
        double acc[P][G];
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int l = 0; l < G; ++l) acc[p][l] = 0.0;

        auto no_operand = [](int) { return 0; };
        for_each_line_pf<int>(A.ranges, A.cands, A.ntiles, tile, j0, j1, lane, KEY_WINDOW,
                              no_operand, [&](const LineView& v, int) {
            double dl[P];
            bool inw[P];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                inw[p] = jj[p] >= v.lb && jj[p] <= v.ub;
                dl[p] = fabs(lam[p] - v.lam0);  // abs(λ - λ₀), src/line_absorption.jl:248
            }
            const KCONST double4* rp = uniform_ptr(A.rec64 + (int64_t)v.line * A.Ls + l0);
#pragma unroll
            for (int l = 0; l < G; ++l) {
                const double4 r = rp[l];  // (1/(σ√2), a, scaling); zero for padding layers
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    const double h = voigt_hjerting64(r.y, dl[p] * r.x) * r.z;
                    acc[p][l] += inw[p] ? h : 0.0;
                }
            }
        });

#pragma unroll
        for (int p = 0; p < P; ++p) {
            if (jj[p] > j1) continue;
#pragma unroll
            for (int l = 0; l < G; ++l)
                if (l0 + l < L)
                    add_alpha(S.alpha, A.alpha_f32, (int64_t)(l0 + l) + (int64_t)L * (jj[p] - go), acc[p][l]);
        }
    }
}

// Which Hunger branch (src/line_absorption.jl:286-310) applies, decided in Float64 exactly as
// voigt_hjerting does: 0 = A (far wing, a <= 0.2), 1 = B (core, a <= 0.2), 2 = C (core,
// 0.2 < a <= 1.4), 3 = D (far wing / strongly damped).
__device__ inline int hunger_branch(double a, double v) {
    if (a <= 0.2 && v >= 5.0) return 0;
    if (a <= 0.2) return 1;
    if (a <= 1.4 && a + v < 3.2) return 2;
    return 3;
}

// Bits i in [0, 32) with base + i in [lo, hi].
__device__ inline uint32_t span_bits(int lo, int hi, int base) {
    const int a = max(lo - base, 0), z = min(hi - base, 31);
    return a > z ? 0u : (uint32_t)(((2ull << (z - a)) - 1) << a);
}

// ---------------------------------------------------------------------------------------------
// Float32 far wing on the matrix cores.  A work item is (tile of 32*NB points, block of 64
// layers), owned by one workgroup of WW waves; each wave takes a contiguous part of the
// tile's line list and holds the 64 x 32NB block of alpha as 2 x NB 32x32 MFMA accumulators.  For
// each line: A = the line's 64 x 4 coefficient block, B = (R, R^2, R^3, R^4) of the tile's points
// (zero outside the window and inside the near zone), acc += A B with v_mfma_f32_32x32x2_f32
// (K = 2 per instruction, two steps).  The waves' blocks are summed through LDS in wave order and
// added to alpha by wave 0.
// ---------------------------------------------------------------------------------------------
typedef float floatx16 __attribute__((ext_vector_type(16)));

// Registers: three waves per SIMD (at most 168 registers) for 32- and 64-point tiles, measured
// faster than two waves with no spills; two for 128-point tiles, whose 128 accumulator registers
// leave too few for three.
template <int NB, int WW, class AlphaT>
__global__ __launch_bounds__(WAVE * WW) __attribute__((amdgpu_waves_per_eu(NB >= 4 ? 2 : 3)))
void k_wing_mfma(ProfileArgs A) {
    // partial blocks of waves 1..WW-1: [wave][accumulator][register quad][lane]
    __shared__ float4 s_acc[WW > 1 ? WW - 1 : 1][2 * NB][4][WAVE];
    __shared__ int s_item;
    const int lane = threadIdx.x % WAVE;
    const int w = threadIdx.x / WAVE;
    const int n = lane & 31;      // column (point) within a 32-point block
    const bool hi = lane >= 32;   // K half: coefficient/power index k = 2 s + hi
    const int L = A.L;
    if (A.counters[3] != 0) return;  // failed setup: nothing is written (the same for all waves)
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(&A.counters[0], 1ull);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= A.n_items) break;  // the same for the whole workgroup
        const int tile = item / A.n_lblocks;
        const int lblk = item % A.n_lblocks;
        const int64_t j0 = (int64_t)tile * A.tile_w;
        const KCONST StarDev* stars = uniform_ptr(A.stars);
        const KCONST StarDev& S = stars[star_of_point(stars, A.n_stars, j0)];
        const int64_t go = S.grid_off;
        // the star's last point at most; below j0 for padding between two stars' grids (such a
        // tile has no lines and writes nothing)
        const int64_t j1 = min(j0 + A.tile_w, go + S.n_wl) - 1;
        // Even NB: the lanes of a wave hold the 64 consecutive points of a pair of blocks (lane i:
        // point 64 q + i), each lane evaluates its own point, and one lane-half swap turns the
        // pair's (R, R^2) into the two blocks' B operands.  NB = 1: both halves hold the block's
        // 32 points.
        constexpr bool PAIR = NB % 2 == 0;
        constexpr int NL = PAIR ? NB / 2 : NB;  // points evaluated per lane
        double lam[NL];
        int jj[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) jj[b] = (int)(j0 + 32 * b + n);
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int64_t jq = PAIR ? j0 + 64 * q + lane : (int64_t)jj[q];
            // in units of DELTA_UNIT (points past j1 are never used)
            lam[q] = S.wls[min(jq, j1) - go] * (1.0 / DELTA_UNIT);
        }
        floatx16 acc[2][NB];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[mb][b][r] = 0.0f;

        // keep the accumulators in AGPRs on every path (no copies at branch merges)
        auto pin_acc = [&] {
#pragma unroll
            for (int b = 0; b < NB; ++b) asm volatile("" : "+a"(acc[0][b]), "+a"(acc[1][b]));
        };
        const float4* __restrict__ recM = A.recM + (int64_t)lblk * WAVE + lane;
        const int64_t rstride = (int64_t)A.n_lblocks * WAVE;
        auto load_a = [&](int line) { return recM[line * rstride]; };
        // B operands of line `bl` of the batch (R, R^2 | R^3, R^4 per point)
        struct Prep {
            float b0[NB];  // k = 0 | 1: R | R^2 (zero outside the window and in the near zone)
            float b1[NB];  // k = 2 | 3: R^3 | R^4, set when `full`
            bool full;     // the tile meets the line's tail zone: the R^3, R^4 step is needed
        };
        // Per batch of 64 lines, each lane describes its own line: for every 32-point block, the
        // bits of the points in the window and outside the near zone, and (one ballot) whether the
        // tile meets the line's tail zone.  Per line this leaves lane reads and scalar tests.
        const int j0s = (int)j0, j1s = (int)j1;  // grid indices are below 2^31
        uint32_t mk[NB];
        uint64_t full_mask = 0;
        double lam0s = 0.0;  // the lane's line centre in units of DELTA_UNIT
        auto describe = [&](const LineMeta& m) {
            lam0s = m.lam0 * (1.0 / DELTA_UNIT);
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int base = j0s + 32 * b;
                mk[b] = span_bits(m.lb, m.ub, base) & ~span_bits(m.clo, m.chi, base);
            }
            full_mask = ballot(m.flo <= j1s && m.fhi >= j0s);
        };
        // lanes 32..63 of x exchanged with lanes 0..31 of y: x becomes (x_lo | y_lo), y (x_hi | y_hi)
        auto swap_halves = [](float& x, float& y) {
            const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y),
                                                            false, false);
            x = __uint_as_float(r[0]);
            y = __uint_as_float(r[1]);
        };
        auto prep = [&](const LineMeta& m, int bl) {
            Prep p;
            const double lam0 = readlane_f64(lam0s, bl);
            p.full = (full_mask >> bl) & 1;
#pragma unroll
            for (int q = 0; q < NL; ++q) {
                const uint32_t mw = (uint32_t)readlane_i32((int)mk[q * (PAIR ? 2 : 1)], bl);
                // the lane's point in the window and outside the near zone
                const uint32_t mw1 = PAIR ? (uint32_t)readlane_i32((int)mk[2 * q + 1], bl) : mw;
                const bool use = __builtin_amdgcn_inverse_ballot_w64(mw | ((uint64_t)mw1 << 32));
                // Float64 fix-up: lambda_j - lambda0 from the Float64 grid values (both scaled
                // to DELTA_UNIT, which adds a rounding of ~1e-10 of the offset)
                const float d = (float)(lam[q] - lam0);
                float R = __builtin_amdgcn_rcpf(d * d);
                R = use ? R : 0.0f;
                float R2 = R * R;
                if constexpr (PAIR) {
                    if (p.full) {
                        float R3 = R * R2, R4 = R2 * R2;
                        swap_halves(R3, R4);
                        p.b1[2 * q] = R3;
                        p.b1[2 * q + 1] = R4;
                    }
                    swap_halves(R, R2);  // (R, R^2) of block 2q, then of block 2q + 1
                    p.b0[2 * q] = R;
                    p.b0[2 * q + 1] = R2;
                } else {
                    p.b0[q] = hi ? R2 : R;  // lanes n and n + 32 hold the same point
                    if (p.full) p.b1[q] = p.b0[q] * R2;
                }
            }
            return p;
        };
        // acc += A B for one line
        auto mfma = [&](const float4& a4, const Prep& p) {
            // every path waits for the operand here, also when the second step is skipped
            asm volatile("" ::"v"(a4.x), "v"(a4.y), "v"(a4.z), "v"(a4.w));
            // first K step (R, R^2) for every block; a block outside the window has B = 0
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                acc[0][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(a4.x, p.b0[b], acc[0][b], 0, 0, 0);
                acc[1][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(a4.y, p.b0[b], acc[1][b], 0, 0, 0);
            }
            if (p.full) {
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    acc[0][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(a4.z, p.b1[b], acc[0][b], 0, 0, 0);
                    acc[1][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(a4.w, p.b1[b], acc[1][b], 0, 0, 0);
                }
            }
            pin_acc();
        };
        // the tile's lines, 64 per batch: line indices are loaded two batches ahead, their
        // LineMeta one batch ahead, and the operands one pair of lines ahead through four
        // registers (a batch holds a multiple of four lines, so the rotation is in phase at every
        // batch).  Positions past the end of the list prepare a stand-in line and issue nothing.
        // wave w takes the w-th of WW contiguous, near-equal parts of the tile's list
        const int64_t t0 = A.tile_start[(int64_t)tile * N_CLASSES];
        const int64_t t1 = A.tile_start[(int64_t)(tile + 1) * N_CLASSES];
        const int64_t s0 = t0 + ((t1 - t0) * w) / WW;
        const int64_t s1 = t0 + ((t1 - t0) * (w + 1)) / WW;
        if (s0 < s1) {
            const int* __restrict__ tl = A.tile_lines;
            auto ld_idx = [&](int64_t k) { return tl[min(k + lane, s1 - 1)]; };  // clamped
            int idxA = ld_idx(s0);
            int idxB = ld_idx(s0 + WAVE);
            LineMeta mA = A.meta[idxA];
            int idxC = ld_idx(s0 + 2 * WAVE);
            LineMeta mB = A.meta[idxB];
            // line of position b of the current batch (b < 128; past the end: a stand-in)
            auto line_at = [&](int b) {
                return b < WAVE ? readlane_i32(idxA, b) : readlane_i32(idxB, b - WAVE);
            };
            float4 x0 = load_a(line_at(0));
            float4 x1 = load_a(line_at(1));
            float4 x2, x3;
            for (int64_t kb = s0; kb < s1; kb += WAVE) {
                const int nb = (int)min((int64_t)WAVE, s1 - kb);
                // the batch's indices and LineMeta are waited for here, once per batch, not
                // inside the line loop (where a wait would also drain the operand prefetches)
                asm volatile("" ::"v"(idxA), "v"(idxB), "v"(mA.lam0), "v"(mA.lb), "v"(mA.ub),
                             "v"(mA.clo), "v"(mA.chi), "v"(mA.flo), "v"(mA.fhi));
                describe(mA);
                // lines in pairs: the two lines' operand chains are independent and interleave
                for (int b = 0; b < nb; b += 4) {
                    x2 = load_a(line_at(b + 2));
                    x3 = load_a(line_at(b + 3));
                    {
                        const Prep pa = prep(mA, b), pb = prep(mA, b + 1);
                        mfma(x0, pa);
                        if (b + 1 >= nb) break;
                        mfma(x1, pb);
                    }
                    if (b + 2 >= nb) break;
                    x0 = load_a(line_at(b + 4));
                    x1 = load_a(line_at(b + 5));
                    {
                        const Prep pc = prep(mA, b + 2), pd = prep(mA, b + 3);
                        mfma(x2, pc);
                        if (b + 3 >= nb) break;
                        mfma(x3, pd);
                    }
                }
                mA = mB;
                idxA = idxB;
                idxB = idxC;
                mB = A.meta[idxB];
                idxC = ld_idx(kb + 3 * WAVE);
            }
        }

        // sum the waves' blocks in wave order (a fixed order: results are deterministic)
        if (w > 0) {
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        s_acc[w - 1][mb * NB + b][q][lane] =
                            make_float4(acc[mb][b][4 * q], acc[mb][b][4 * q + 1],
                                        acc[mb][b][4 * q + 2], acc[mb][b][4 * q + 3]);
        }
        __syncthreads();
        if (w == 0) {
            // accumulator layout of v_mfma_f32_32x32x2_f32: lane i, register r holds
            // row 8*(r/4) + 4*(i/32) + r%4, column i%32
            AlphaT* __restrict__ alpha = (AlphaT*)S.alpha;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if (jj[b] > j1) continue;
#pragma unroll
                for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        float4 v = make_float4(acc[mb][b][4 * q], acc[mb][b][4 * q + 1],
                                               acc[mb][b][4 * q + 2], acc[mb][b][4 * q + 3]);
#pragma unroll
                        for (int ww = 0; ww < WW - 1; ++ww) {
                            const float4 p = s_acc[ww][mb * NB + b][q][lane];
                            v.x += p.x;
                            v.y += p.y;
                            v.z += p.z;
                            v.w += p.w;
                        }
                        const float vr[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int r = 4 * q + e;
                            const int layer = lblk * WAVE + 32 * mb + 8 * (r >> 2) + 4 * (int)hi + (r & 3);
                            if (layer < L) alpha[(int64_t)layer + (int64_t)L * (jj[b] - go)] += (AlphaT)vr[e];
                        }
                    }
            }
        }
        __syncthreads();  // s_acc and s_item are reused by the next item
    }
}

// ---------------------------------------------------------------------------------------------
// The far wing with bfloat16 matrix products.  The same work items and line lists as
// k_wing_mfma; the tile's lines go four at a time through v_mfma_f32_32x32x16_bf16, K = 16 =
// (4 lines x 4 powers): A = the 4 lines' coefficients of 32 layers (from k_setup's bfloat16
// layout), B = R^(p+1) of the 4 lines at the block's 32 points.  Each operand is split into a
// bfloat16 high part and the bfloat16 of the remainder; acc += A_hi B_hi + A_hi B_lo + A_lo B_hi
// (Float32 accumulation) keeps about 16 significant bits of every term.  Lane i of the wave holds
// K rows 8 (i / 32) .. + 7 -- line slots 2 (i / 32) and 2 (i / 32) + 1, powers 1..4 each -- for
// column (point) i % 32; the R^3, R^4 terms are always included.
// ---------------------------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

// The two selectors of split_pair, bfloat16 pairs (-1, 0) and (0, -1), held in registers.  As
// compile-time constants the compiler folds (-1, 0) into the inline constant -1.0, which the
// hardware reads as the 32-bit pattern of Float32 -1.0, i.e. the pair (0, -1): the wrong element
// (seen on the device with an exhaustive-magnitude check of split_pair).
struct SplitSel {
    uint32_t m0, m1;
};
__device__ inline SplitSel split_selectors() {
    SplitSel s = {0x0000BF80u, 0xBF800000u};
    asm volatile("" : "+v"(s.m0), "+v"(s.m1));
    return s;
}

// High and low bfloat16 parts of a pair of floats, packed as the MFMA operands want them: one
// v_cvt_pk_bf16_f32 (round to nearest even) gives both high parts; x - high is read back from the
// packed pair by v_dot2_f32_bf16 (x + (-1) high + 0 other high, exact: the difference of a float
// and its bfloat16 rounding is a float); one more conversion packs the two remainders.
__device__ inline void split_pair(float x0, float x1, SplitSel sel, uint32_t& hi, uint32_t& lo) {
    const bf16x2 h = __builtin_convertvector((floatx2){x0, x1}, bf16x2);
    const float r0 = __builtin_amdgcn_fdot2_f32_bf16(h, __builtin_bit_cast(bf16x2, sel.m0), x0, false);
    const float r1 = __builtin_amdgcn_fdot2_f32_bf16(h, __builtin_bit_cast(bf16x2, sel.m1), x1, false);
    const bf16x2 l = __builtin_convertvector((floatx2){r0, r1}, bf16x2);
    hi = __builtin_bit_cast(uint32_t, h);
    lo = __builtin_bit_cast(uint32_t, l);
}

__device__ inline void split_bf16(const float (&x)[8], SplitSel sel, bf16x8& hi, bf16x8& lo) {
    uint32_t h[4], l[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) split_pair(x[2 * e], x[2 * e + 1], sel, h[e], l[e]);
    hi = __builtin_bit_cast(bf16x8, make_uint4(h[0], h[1], h[2], h[3]));
    lo = __builtin_bit_cast(bf16x8, make_uint4(l[0], l[1], l[2], l[3]));
}

// Registers: the accumulators (32 NB) plus two lines' operands per block leave room for three
// waves per SIMD only with 32-point tiles.
template <int NB, int WW, class AlphaT>
__global__ __launch_bounds__(WAVE * WW) __attribute__((amdgpu_waves_per_eu(NB >= 2 ? 2 : 3)))
void k_wing_bf16(ProfileArgs A) {
    // partial blocks of waves 1..WW-1: [wave][accumulator][register quad][lane]
    __shared__ float4 s_acc[WW > 1 ? WW - 1 : 1][2 * NB][4][WAVE];
    __shared__ int s_item;
    const int lane = threadIdx.x % WAVE;
    // wavefront-uniform (scalar registers): the wave's share of the line list, the batch and
    // group positions and the lines' readlanes below all stay scalar
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const int n = lane & 31;       // column (point) within a 32-point block
    const int hf = lane >> 5;      // K rows 8 hf .. 8 hf + 7: line slots 2 hf, 2 hf + 1
    const bool hi_half = hf != 0;
    const int L = A.L;
    const SplitSel sel = split_selectors();
    if (A.counters[3] != 0) return;  // failed setup: nothing is written (the same for all waves)
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(&A.counters[0], 1ull);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= A.n_items) break;  // the same for the whole workgroup
        const int tile = item / A.n_lblocks;
        const int lblk = item % A.n_lblocks;
        const int64_t j0 = (int64_t)tile * A.tile_w;
        const KCONST StarDev* stars = uniform_ptr(A.stars);
        const KCONST StarDev& S = stars[star_of_point(stars, A.n_stars, j0)];
        const int64_t go = S.grid_off;
        // the star's last point at most; below j0 for padding between two stars' grids (such a
        // tile has no lines and writes nothing)
        const int64_t j1 = min(j0 + A.tile_w, go + S.n_wl) - 1;
        // wavelengths in DELTA_UNIT relative to the tile's first point, in Float32: the offsets
        // lambda - lambda0 below are then one Float32 subtraction.  A far-zone offset is at least
        // the near-zone radius (3.2 sigma sqrt 2, ten units and more) while the tile spans 32 NB
        // units, so its relative rounding stays below 1e-6 (R's, 2e-6: below the bfloat16 split's)
        const double lamt0 = S.wls[min(j0, j1) - go] * (1.0 / DELTA_UNIT);  // (padding: j1)
        float lam[NB];
        int jj[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            jj[b] = (int)(j0 + 32 * b + n);
            // (points past j1 are never used)
            lam[b] = (float)(S.wls[min((int64_t)jj[b], j1) - go] * (1.0 / DELTA_UNIT) - lamt0);
        }
        floatx16 acc[2][NB];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[mb][b][r] = 0.0f;
        auto pin_acc = [&] {
#pragma unroll
            for (int b = 0; b < NB; ++b) asm volatile("" : "+a"(acc[0][b]), "+a"(acc[1][b]));
        };
        // far field: the sub-tile of each 32-point block (wavefront-uniform)
        FfGeo gb[NB];
        if (A.fft) {
            const KCONST FfGeo* geo = uniform_ptr(A.geo);
#pragma unroll
            for (int b = 0; b < NB; ++b) gb[b] = geo[min((j0 + 32 * b) / FF_SUB, A.n_sub - 1)];
        }
        if (A.fft && w == 0) {
            // wave 0 starts its blocks from the far-field sums: acc = T P with A = T (64 layers x
            // the terms, padded to 16 with zeros) and B = tau^m of the block's points
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int64_t g = (j0 + 32 * b) / FF_SUB;
                if (g >= A.n_sub || gb[b].first > gb[b].last) continue;  // padding: no points
                const int64_t jp = min(max((int64_t)jj[b], (int64_t)gb[b].first), (int64_t)gb[b].last);
                const float tau = gb[b].h_u > 0.0f
                    ? (float)(S.wls[jp - go] * (1.0 / DELTA_UNIT) - gb[b].c_u) / gb[b].h_u : 0.0f;
                float pw[8];
                float p = 1.0f;
                if (hi_half) {
                    const float t2 = tau * tau, t4 = t2 * t2;
                    p = t4 * t4;
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    pw[e] = 8 * hf + e < FF_TERMS ? p : 0.0f;
                    p *= tau;
                }
                bf16x8 ph, pl;
                split_bf16(pw, sel, ph, pl);
                const float* __restrict__ Tg =
                    A.fft + (((g / FF_NS) * A.n_lblocks + lblk) * FF_NS + g % FF_NS) * (WAVE * FF_TERMS);
#pragma unroll
                for (int mb = 0; mb < 2; ++mb) {
                    const float* row = Tg + (32 * mb + n) * FF_TERMS + 8 * hf;
                    const float4 v0 = *(const float4*)row;
                    float4 v1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    if (8 * hf + 4 < FF_TERMS) v1 = *(const float4*)(row + 4);
                    const float tv[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
                    bf16x8 th, tl;
                    split_bf16(tv, sel, th, tl);
                    acc[mb][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(th, ph, acc[mb][b], 0, 0, 0);
                    acc[mb][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(th, pl, acc[mb][b], 0, 0, 0);
                    acc[mb][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(tl, ph, acc[mb][b], 0, 0, 0);
                }
            }
        }
        // coefficients of (line, layer 32 mb + n): 16 B, high parts then low parts
        const uint4* __restrict__ recB = (const uint4*)A.recM + (int64_t)lblk * WAVE + n;
        const int64_t rstride = (int64_t)A.n_lblocks * WAVE;
        // per batch of 64 lines, each lane describes its own line (as k_wing_mfma)
        const int j0s = (int)j0, j1s = (int)j1;
        (void)j1s;
        uint32_t mk[NB];
        float lam0s = 0.0f;
        auto describe = [&](const LineMeta& m, const int4& run) {
            lam0s = (float)(m.lam0 * (1.0 / DELTA_UNIT) - lamt0);
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int base = j0s + 32 * b;
                mk[b] = span_bits(m.lb, m.ub, base) & ~span_bits(m.clo, m.chi, base);
                // a block whose sub-tile takes this line's far field has it already (tiles of
                // at most one sub-tile list no such line)
                if (NB * 32 > FF_SUB && A.fft && ff_in(run, (j0 + 32 * b) / FF_SUB)) mk[b] = 0u;
            }
        };
        const int64_t t0 = A.tile_start[(int64_t)tile * N_CLASSES];
        const int64_t t1 = A.tile_start[(int64_t)(tile + 1) * N_CLASSES];
        const int64_t s0 = t0 + ((t1 - t0) * w) / WW;
        const int64_t s1 = t0 + ((t1 - t0) * (w + 1)) / WW;
        const int* __restrict__ tl = A.tile_lines;
        for (int64_t kb = s0; kb < s1; kb += WAVE) {
            const int nb = (int)min((int64_t)WAVE, s1 - kb);
            const int idx = tl[min(kb + lane, s1 - 1)];  // clamped: a stand-in past the end
            describe(A.meta[idx], NB * 32 > FF_SUB && A.fft ? A.ffrun[idx] : make_int4(1, 0, 1, 0));
            for (int g = 0; g < nb; g += 4) {
                // the group's batch positions g .. g + 3; past nb: stand-ins with no points
                const int q0 = min(g, nb - 1), q1 = min(g + 1, nb - 1);
                const int q2 = min(g + 2, nb - 1), q3 = min(g + 3, nb - 1);
                // both halves' lines are read (uniform readlanes), then each lane selects its own
                const int i0 = readlane_i32(idx, q0), i1 = readlane_i32(idx, q1);
                const int i2 = readlane_i32(idx, q2), i3 = readlane_i32(idx, q3);
                const int la = hi_half ? i2 : i0;
                const int lb2 = hi_half ? i3 : i1;
                uint4 ua[2], ub[2];
#pragma unroll
                for (int mb = 0; mb < 2; ++mb) {
                    ua[mb] = recB[la * rstride + 32 * mb];
                    ub[mb] = recB[lb2 * rstride + 32 * mb];
                }
                const float c0 = readlane_f32(lam0s, q0), c1 = readlane_f32(lam0s, q1);
                const float c2 = readlane_f32(lam0s, q2), c3 = readlane_f32(lam0s, q3);
                const float l0a = hi_half ? c2 : c0;
                const float l0b = hi_half ? c3 : c1;
                bf16x8 bh[NB], bl[NB];
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const uint32_t m0 = (uint32_t)readlane_i32((int)mk[b], q0);
                    const uint32_t m1 = g + 1 < nb ? (uint32_t)readlane_i32((int)mk[b], q1) : 0u;
                    const uint32_t m2 = g + 2 < nb ? (uint32_t)readlane_i32((int)mk[b], q2) : 0u;
                    const uint32_t m3 = g + 3 < nb ? (uint32_t)readlane_i32((int)mk[b], q3) : 0u;
                    const bool ua_ = __builtin_amdgcn_inverse_ballot_w64(m0 | ((uint64_t)m2 << 32));
                    const bool ub_ = __builtin_amdgcn_inverse_ballot_w64(m1 | ((uint64_t)m3 << 32));
                    const float da = lam[b] - l0a, db = lam[b] - l0b;
                    float Ra = __builtin_amdgcn_rcpf(da * da), Rb = __builtin_amdgcn_rcpf(db * db);
                    Ra = ua_ ? Ra : 0.0f;
                    Rb = ub_ ? Rb : 0.0f;
                    const float Ra2 = Ra * Ra, Rb2 = Rb * Rb;
                    const float x[8] = {Ra, Ra2, Ra2 * Ra, Ra2 * Ra2, Rb, Rb2, Rb2 * Rb, Rb2 * Rb2};
                    split_bf16(x, sel, bh[b], bl[b]);
                }
#pragma unroll
                for (int mb = 0; mb < 2; ++mb) {
                    const bf16x8 ah = __builtin_bit_cast(bf16x8, make_uint4(ua[mb].x, ua[mb].y, ub[mb].x, ub[mb].y));
                    const bf16x8 al = __builtin_bit_cast(bf16x8, make_uint4(ua[mb].z, ua[mb].w, ub[mb].z, ub[mb].w));
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        acc[mb][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh[b], acc[mb][b], 0, 0, 0);
                        acc[mb][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl[b], acc[mb][b], 0, 0, 0);
                        acc[mb][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh[b], acc[mb][b], 0, 0, 0);
                    }
                }
                pin_acc();
            }
        }

        // sum the waves' blocks in wave order (a fixed order: results are deterministic)
        if (w > 0) {
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        s_acc[w - 1][mb * NB + b][q][lane] =
                            make_float4(acc[mb][b][4 * q], acc[mb][b][4 * q + 1],
                                        acc[mb][b][4 * q + 2], acc[mb][b][4 * q + 3]);
        }
        __syncthreads();
        if (w == 0) {
            // accumulator layout of the 32x32 MFMAs: lane i, register r holds row
            // 8*(r/4) + 4*(i/32) + r%4, column i%32
            AlphaT* __restrict__ alpha = (AlphaT*)S.alpha;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if (jj[b] > j1) continue;
#pragma unroll
                for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        float4 v = make_float4(acc[mb][b][4 * q], acc[mb][b][4 * q + 1],
                                               acc[mb][b][4 * q + 2], acc[mb][b][4 * q + 3]);
#pragma unroll
                        for (int ww = 0; ww < WW - 1; ++ww) {
                            const float4 p = s_acc[ww][mb * NB + b][q][lane];
                            v.x += p.x;
                            v.y += p.y;
                            v.z += p.z;
                            v.w += p.w;
                        }
                        const float vr[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int r = 4 * q + e;
                            const int layer = lblk * WAVE + 32 * mb + 8 * (r >> 2) + 4 * hf + (r & 3);
                            if (layer < L) alpha[(int64_t)layer + (int64_t)L * (jj[b] - go)] += (AlphaT)vr[e];
                        }
                    }
            }
        }
        __syncthreads();  // s_acc and s_item are reused by the next item
    }
}


// ---------------------------------------------------------------------------------------------
// k_wing_ff: the far-field sums T of one super-tile (FF_NS sub-tiles) and 64-layer block per
// work item, one wave each:  T[layer, (s, m)] = sum over the super-tile's far-field list of
// sum_k c_k(line, layer) W_km(line, s)  on v_mfma_f32_32x32x16_bf16, K = 4 lines x 4 powers:
// A = the lines' coefficients exactly as k_wing_bf16 reads them, B = their W at 32 (sub-tile,
// term) columns (zero for a sub-tile whose far field the line does not take), both split into
// bfloat16 high and low parts (three products).  Column 32 cb + n of an item is sub-tile
// n % FF_NS, term 4 cb + n / FF_NS.  W = binom(-2k, m) u^m R^k with
// R = 1/d^2 and u = h/d (|u| <= 1/FF_RATIO), d = c - lam0 in DELTA_UNIT (a Float64 difference,
// then Float32).  T goes to fft as [super-tile][layer block][sub-tile][64 layers][FF_TERMS].
// ---------------------------------------------------------------------------------------------
struct FarArgs {
    int n_lblocks;
    int64_t n_items;    // super-tiles x layer blocks
    int64_t n_sub;      // sub-tiles of the call
    int64_t base_slot;  // first far-field slot of tile_start
    int ntiles;         // wing tiles of the call
    int per;            // wing tiles per super-tile
    const FfGeo* geo;
    const LineMeta* meta;
    const int4* ffrun;  // each line's far-field runs
    const float4* recM;
    const int64_t* tile_start;
    const int* tile_lines;
    float* fft;
    unsigned long long* counters;
};

// u^m for 0 <= m < 16, by squaring.
__device__ inline float pow_small(float u, int m) {
    const float u2 = u * u, u4 = u2 * u2, u8 = u4 * u4;
    float p = (m & 1) ? u : 1.0f;
    p *= (m & 2) ? u2 : 1.0f;
    p *= (m & 4) ? u4 : 1.0f;
    p *= (m & 8) ? u8 : 1.0f;
    return p;
}

// WW waves share a work item, each taking a contiguous 1/WW of its list; their partial sums are
// added in wave order through LDS (deterministic).  96 accumulator registers and about 150 vector
// registers: two waves per SIMD (the occupancy hint is left at one, so that the compiler need not
// split the register file evenly and spill).
template <int WW>
__global__ __launch_bounds__(WAVE * WW) __attribute__((amdgpu_waves_per_eu(WW > 1 ? 2 : 1))) void k_wing_ff(FarArgs A) {
    __shared__ FfGeo sg[FF_NS];
    __shared__ float4 s_acc[WW > 1 ? WW - 1 : 1][2 * FF_CB][4][WAVE];
    __shared__ int s_item;
    const int lane = threadIdx.x % WAVE;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const int n = lane & 31, hf = lane >> 5;
    const bool hi_half = hf != 0;
    const SplitSel sel = split_selectors();
    if (A.counters[3] != 0) return;  // failed setup: nothing is written
    // Column 32 cb + n of an item is sub-tile n % FF_NS and term 4 cb + n / FF_NS: a lane keeps
    // one sub-tile in every column block, so its line distances and powers are formed once per
    // line, and the term steps up by 4 from one block to the next.  binom(-2k, m) for k = 1..4:
    const int sub = n % FF_NS, m0 = n / FF_NS;
    static_assert(FF_NS * 4 == 32 && FF_CB * 4 == FF_TERMS, "columns: 8 sub-tiles x 4 terms per block");
    float bk[FF_CB][4];
#pragma unroll
    for (int cb = 0; cb < FF_CB; ++cb) {
#pragma unroll
        for (int k = 1; k <= 4; ++k) {
            double b = 1.0;
            for (int i = 1; i <= 4 * cb + m0; ++i) b *= -(double)(2 * k + i - 1) / (double)i;
            bk[cb][k - 1] = (float)b;
        }
    }
    const int64_t rstride = (int64_t)A.n_lblocks * WAVE;
    const int* __restrict__ tl = A.tile_lines;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(&A.counters[24], 1ull);
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= A.n_items) break;  // the same for the whole workgroup
        const int st = item / A.n_lblocks, lblk = item % A.n_lblocks;
        if (item == 0 && threadIdx.x == 0)  // the far-field list entries, for the statistics
            A.counters[25] = (unsigned long long)(A.tile_start[A.base_slot + (int64_t)A.ntiles * N_CLASSES] -
                                                  A.tile_start[A.base_slot]);
        if (threadIdx.x < FF_NS) {
            const int64_t g = (int64_t)st * FF_NS + threadIdx.x;
            const FfGeo none = {0.0, 0.0, 0.0, 0.0f, 1, 0, 0};
            sg[threadIdx.x] = g < A.n_sub ? A.geo[g] : none;
        }
        __syncthreads();
        const double cu = sg[sub].c_u;  // the lane's sub-tile
        const float hu = sg[sub].h_u;
        floatx16 acc[2][FF_CB];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int cb = 0; cb < FF_CB; ++cb)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[mb][cb][r] = 0.0f;
        const uint4* __restrict__ recB = (const uint4*)A.recM + (int64_t)lblk * WAVE + n;
        // the super-tile's list: the far-field slots of its wing tiles
        const int64_t e0 = A.tile_start[A.base_slot + (int64_t)st * A.per * N_CLASSES];
        const int64_t e1 =
            A.tile_start[A.base_slot + (int64_t)min((st + 1) * A.per, A.ntiles) * N_CLASSES];
        const int64_t t0 = e0 + ((e1 - e0) * w) / WW;
        const int64_t t1 = e0 + ((e1 - e0) * (w + 1)) / WW;
        for (int64_t kb = t0; kb < t1; kb += WAVE) {
            const int nb = (int)min((int64_t)WAVE, t1 - kb);
            const int idx = tl[min(kb + lane, t1 - 1)];  // clamped: a stand-in past the end
            // each lane describes its own line: centre in DELTA_UNIT and far-field sub-tiles
            const double lam0u = A.meta[idx].lam0 * (1.0 / DELTA_UNIT);
            const int4 run = A.ffrun[idx];
            uint32_t emask = 0;
#pragma unroll
            for (int s = 0; s < FF_NS; ++s)
                if (ff_in(run, (int64_t)st * FF_NS + s)) emask |= 1u << s;
            for (int g = 0; g < nb; g += 4) {
                const int q0 = min(g, nb - 1), q1 = min(g + 1, nb - 1);
                const int q2 = min(g + 2, nb - 1), q3 = min(g + 3, nb - 1);
                const int i0 = readlane_i32(idx, q0), i1 = readlane_i32(idx, q1);
                const int i2 = readlane_i32(idx, q2), i3 = readlane_i32(idx, q3);
                const int la = hi_half ? i2 : i0;
                const int lb2 = hi_half ? i3 : i1;
                uint4 ua[2], ub[2];
#pragma unroll
                for (int mb = 0; mb < 2; ++mb) {
                    ua[mb] = recB[la * rstride + 32 * mb];
                    ub[mb] = recB[lb2 * rstride + 32 * mb];
                }
                const double l0 = readlane_f64(lam0u, q0), l1 = readlane_f64(lam0u, q1);
                const double l2 = readlane_f64(lam0u, q2), l3 = readlane_f64(lam0u, q3);
                const double lA = hi_half ? l2 : l0, lB = hi_half ? l3 : l1;
                // stand-ins past nb take no far field
                const uint32_t e0 = (uint32_t)readlane_i32((int)emask, q0);
                const uint32_t e1 = g + 1 < nb ? (uint32_t)readlane_i32((int)emask, q1) : 0u;
                const uint32_t e2 = g + 2 < nb ? (uint32_t)readlane_i32((int)emask, q2) : 0u;
                const uint32_t e3 = g + 3 < nb ? (uint32_t)readlane_i32((int)emask, q3) : 0u;
                const uint32_t eA = hi_half ? e2 : e0, eB = hi_half ? e3 : e1;
                bf16x8 ah[2], al[2];
#pragma unroll
                for (int mb = 0; mb < 2; ++mb) {
                    ah[mb] = __builtin_bit_cast(bf16x8, make_uint4(ua[mb].x, ua[mb].y, ub[mb].x, ub[mb].y));
                    al[mb] = __builtin_bit_cast(bf16x8, make_uint4(ua[mb].z, ua[mb].w, ub[mb].z, ub[mb].w));
                }
                // the two lines' 1/d, R^1..R^4 and u^m0 at the lane's sub-tile (zero where the
                // sub-tile is not in the line's far field), once for all column blocks
                float xA = __builtin_amdgcn_rcpf((float)(cu - lA));
                float xB = __builtin_amdgcn_rcpf((float)(cu - lB));
                xA = ((eA >> sub) & 1u) ? xA : 0.0f;
                xB = ((eB >> sub) & 1u) ? xB : 0.0f;
                const float RA = xA * xA, RB = xB * xB;
                float rA[4] = {RA, RA * RA, 0.0f, 0.0f}, rB[4] = {RB, RB * RB, 0.0f, 0.0f};
                rA[2] = rA[1] * RA;
                rA[3] = rA[1] * rA[1];
                rB[2] = rB[1] * RB;
                rB[3] = rB[1] * rB[1];
                const float uA = hu * xA, uB = hu * xB;
                float pA = pow_small(uA, m0), pB = pow_small(uB, m0);  // u^m, m = 4 cb + m0
                const float uA2 = uA * uA, uB2 = uB * uB;
                const float uA4 = uA2 * uA2, uB4 = uB2 * uB2;
#pragma unroll
                for (int cb = 0; cb < FF_CB; ++cb) {
                    // one column block at a time: interleaving them costs registers, not time
                    __builtin_amdgcn_sched_barrier(0);
                    const float x[8] = {bk[cb][0] * pA * rA[0], bk[cb][1] * pA * rA[1],
                                        bk[cb][2] * pA * rA[2], bk[cb][3] * pA * rA[3],
                                        bk[cb][0] * pB * rB[0], bk[cb][1] * pB * rB[1],
                                        bk[cb][2] * pB * rB[2], bk[cb][3] * pB * rB[3]};
                    pA *= uA4;
                    pB *= uB4;
                    bf16x8 bh, bl;
                    split_bf16(x, sel, bh, bl);
#pragma unroll
                    for (int mb = 0; mb < 2; ++mb) {
                        acc[mb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mb], bh, acc[mb][cb], 0, 0, 0);
                        acc[mb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mb], bl, acc[mb][cb], 0, 0, 0);
                        acc[mb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[mb], bh, acc[mb][cb], 0, 0, 0);
                    }
                }
                // the accumulators stay in AGPRs (the vector registers hold the operands)
#pragma unroll
                for (int cb = 0; cb < FF_CB; ++cb) asm volatile("" : "+a"(acc[0][cb]), "+a"(acc[1][cb]));
            }
        }
        // the waves' partial sums, added in wave order by wave 0
        if (w > 0) {
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int cb = 0; cb < FF_CB; ++cb)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        s_acc[w - 1][mb * FF_CB + cb][q][lane] =
                            make_float4(acc[mb][cb][4 * q], acc[mb][cb][4 * q + 1],
                                        acc[mb][cb][4 * q + 2], acc[mb][cb][4 * q + 3]);
        }
        __syncthreads();
        if (w == 0) {
            // accumulator layout of the 32x32 MFMAs: lane i, register r holds row
            // 8*(r/4) + 4*(i/32) + r%4, column i%32
            float* __restrict__ Tst =
                A.fft + ((int64_t)st * A.n_lblocks + lblk) * (FF_NS * WAVE * FF_TERMS);
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int cb = 0; cb < FF_CB; ++cb)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        __builtin_amdgcn_sched_barrier(0);  // a few loads in flight at a time
                        float v[4] = {acc[mb][cb][4 * q], acc[mb][cb][4 * q + 1],
                                      acc[mb][cb][4 * q + 2], acc[mb][cb][4 * q + 3]};
#pragma unroll
                        for (int ww = 0; ww < WW - 1; ++ww) {
                            const float4 p = s_acc[ww][mb * FF_CB + cb][q][lane];
                            v[0] += p.x;
                            v[1] += p.y;
                            v[2] += p.z;
                            v[3] += p.w;
                        }
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int layer = 32 * mb + 8 * q + 4 * hf + e;
                            Tst[(sub * WAVE + layer) * FF_TERMS + 4 * cb + m0] = v[e];
                        }
                    }
        }
        __syncthreads();  // sg, s_acc and s_item are reused by the next item
    }
}

// ---------------------------------------------------------------------------------------------
// The Float64 series form's far wing: k_wing_bf16's work items (tile of 32 points, block of 64
// layers) and per-tile line lists, one wave per item, with v_mfma_f64_16x16x4_f64 taking one line
// per instruction, K = its 4 powers: A = the line's Float64 coefficients of 16 layers, B = (R,
// R^2, R^3, R^4) at 16 points (R = 1 / (lambda - lambda0)^2 in DELTA_UNIT, divided in Float64;
// zero outside the window and inside the near zone).  Lane i holds A(layer i % 16, power i / 16)
// and B(power i / 16, point i % 16); register r of accumulator (row block, column block) holds
// layer i / 16 + 4 r, point i % 16 of the block (the Float64 instruction interleaves its rows over
// the registers; the Float32 16x16x4 one keeps four consecutive rows per lane).  8 instructions
// per line: 4 row blocks x 2 column blocks.
// ---------------------------------------------------------------------------------------------
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <class AlphaT>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(4))) void k_wing_f64(ProfileArgs A) {
    const int lane = threadIdx.x;
    const int col = lane & 15, kq = lane >> 4;
    const int L = A.L;
    if (A.counters[3] != 0) return;  // failed setup: nothing is written
    for (;;) {
        const int item = wave_take(&A.counters[0], lane);
        if (item >= A.n_items) break;
        const int tile = item / A.n_lblocks;
        const int lblk = item % A.n_lblocks;
        const int64_t j0 = (int64_t)tile * A.tile_w;
        const KCONST StarDev* stars = uniform_ptr(A.stars);
        const KCONST StarDev& S = stars[star_of_point(stars, A.n_stars, j0)];
        const int64_t go = S.grid_off;
        const int64_t j1 = min(j0 + A.tile_w, go + S.n_wl) - 1;  // below j0: padding, no lines
        int jj[2];
        double lam[2];
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
            jj[cb] = (int)(j0 + 16 * cb + col);
            lam[cb] = S.wls[min((int64_t)jj[cb], j1) - go] * (1.0 / DELTA_UNIT);  // in bounds
        }
        doublex4 acc[4][2];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) acc[rb][cb] = doublex4{0.0, 0.0, 0.0, 0.0};
        // coefficient kq of layer 16 rb + col of the item's block, for line `line`
        const double* __restrict__ recA =
            (const double*)A.recM64 + ((int64_t)lblk * WAVE + col) * 4 + kq;
        const int64_t rstride = (int64_t)A.n_lblocks * WAVE * 4;
        const int64_t t0 = A.tile_start[(int64_t)tile * N_CLASSES];
        const int64_t t1 = A.tile_start[(int64_t)(tile + 1) * N_CLASSES];
        for (int64_t kb = t0; kb < t1; kb += WAVE) {
            const int nb = (int)min((int64_t)WAVE, t1 - kb);
            const int idx = A.tile_lines[min(kb + lane, t1 - 1)];  // clamped: in bounds
            const LineMeta mm = A.meta[idx];
            // the next line's coefficients are loaded while this one's products run
            double an[4];
            {
                const int line0 = readlane_i32(idx, 0);
#pragma unroll
                for (int rb = 0; rb < 4; ++rb) an[rb] = recA[(int64_t)line0 * rstride + rb * 64];
            }
            for (int b = 0; b < nb; ++b) {
                const double lam0s = readlane_f64(mm.lam0, b) * (1.0 / DELTA_UNIT);
                const int lb = readlane_i32(mm.lb, b), ub = readlane_i32(mm.ub, b);
                const int clo = readlane_i32(mm.clo, b), chi = readlane_i32(mm.chi, b);
                double a[4];
#pragma unroll
                for (int rb = 0; rb < 4; ++rb) a[rb] = an[rb];
                const int next = readlane_i32(idx, min(b + 1, nb - 1));  // past the last: itself
#pragma unroll
                for (int rb = 0; rb < 4; ++rb) an[rb] = recA[(int64_t)next * rstride + rb * 64];
#pragma unroll
                for (int cb = 0; cb < 2; ++cb) {
                    const int j = jj[cb];
                    const bool far = j <= j1 && j >= lb && j <= ub && (j < clo || j > chi);
                    const double d = lam[cb] - lam0s;
                    // 1 / d^2: the hardware reciprocal refined by two Newton steps (to an ulp or
                    // two; the series' truncation is 1e-5)
                    const double d2 = d * d;
                    double y = __builtin_amdgcn_rcp(d2);
                    y = fma(y, fma(-d2, y, 1.0), y);
                    y = fma(y, fma(-d2, y, 1.0), y);
                    const double R = far ? y : 0.0;
                    const double R2 = R * R;
                    const double bv = kq == 0 ? R : (kq == 1 ? R2 : (kq == 2 ? R2 * R : R2 * R2));
#pragma unroll
                    for (int rb = 0; rb < 4; ++rb)
                        acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb], bv, acc[rb][cb], 0, 0, 0);
                }
            }
        }
        AlphaT* __restrict__ alpha = (AlphaT*)S.alpha;
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
            if (jj[cb] > j1) continue;
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int layer = lblk * WAVE + 16 * rb + kq + 4 * r;
                    if (layer < L) alpha[(int64_t)layer + (int64_t)L * (jj[cb] - go)] += (AlphaT)acc[rb][cb][r];
                }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Float32 near zone, lanes = layers.  A work item is (tile of CT points, block of 64 layers);
// for every line whose near zone [clo, chi] meets the tile (in line order), each lane evaluates
// its layer exactly, branch by branch, at the tile's points of the zone.  Runs after
// the far-wing kernel and adds its own sums into alpha.
// ---------------------------------------------------------------------------------------------
constexpr int CT = 32;

struct NearArgs {
    int64_t n_wl, n_lines;
    int L, Ls;
    int n_lblocks, ntiles;
    int64_t n_items;
    const StarDev* stars;
    int n_stars;
    const Cand* cands;      // lines in near-zone reach-class partition order
    const int2* ranges;     // per (class, CT-point tile): candidate index range in cands
    const double4* rec64;
    int alpha_f32;
    unsigned long long* counters;
    float* nbuf;  // when set: the near-zone sums are stored here ([global point][L]), not added
    // nonzero: one item per wavefront (a grid of n_items, items grouped by XCD) instead of
    // persistent wavefronts taking items from a counter, so that the slots of finished waves go
    // to kernels on other streams as well (the near zone running beside the far wing)
    int one_item;
};

// Item of workgroup b in a grid of n: workgroups go to the XCDs round robin (b % 8), and each
// XCD takes a contiguous range of items, so the waves sharing an L2 work on neighbouring tiles.
__device__ inline int64_t xcd_item(int64_t b, int64_t n) {
    const int64_t q = n / 8, r = n % 8;
    const int64_t x = b % 8, k = b / 8;
    return x * q + (x < r ? x : r) + k;
}

__device__ inline float voigt_far_f32(float a, float v, bool branch_d) {
    const float isp = 0.5641895835477563f;  // 1/sqrt(pi)
    if (!branch_d) {
        // branch A: (a/sqrt(pi)) w (1 + 1.5 w + 3.75 w^2), w = 1/v^2
        const float w = __builtin_amdgcn_rcpf(v * v);
        return (a * isp * w) * (1.0f + 1.5f * w + 3.75f * (w * w));
    }
    // branch D
    const float r2 = v * v * __builtin_amdgcn_rcpf(a * a);
    const float a_invu = 0.70710678118654752f * __builtin_amdgcn_rcpf((r2 + 1.0f) * a);
    const float a2 = a_invu * a_invu;
    return 0.79788456080286536f * a_invu *
           (1.0f + (3.0f * r2 - 1.0f + ((r2 - 2.0f) * 15.0f * r2 + 2.0f) * a2) * a2);
}

__global__ __launch_bounds__(WAVE) void k_near(NearArgs A) {
    const int lane = threadIdx.x;
    const int L = A.L;
    if (A.counters[3] != 0) return;
    for (;;) {
        const int item = wave_take(&A.counters[5], lane);
        if (item >= A.n_items) break;
        const int tile = item / A.n_lblocks;
        const int l = (item % A.n_lblocks) * WAVE + lane;
        const bool lv = l < L;
        const int64_t j0 = (int64_t)tile * CT;
        const KCONST StarDev* stars = uniform_ptr(A.stars);
        const KCONST StarDev& S = stars[star_of_point(stars, A.n_stars, j0)];
        const int64_t go = S.grid_off;
        const int64_t j1 = min(j0 + CT, go + S.n_wl) - 1;
        if (j1 < j0) continue;  // padding between two stars' grids: no points
        double lam[CT];  // the tile's wavelengths (wavefront-uniform)
        const KCONST double* wt = uniform_ptr(S.wls + (j0 - go));
#pragma unroll
        for (int t = 0; t < CT; ++t) lam[t] = wt[min((int64_t)t, j1 - j0)];  // in bounds
        float acc[CT];
#pragma unroll
        for (int t = 0; t < CT; ++t) acc[t] = 0.0f;
        // candidates: class by class, line index within a class; 64 checked at once
        const double4* __restrict__ rq = A.rec64 + (lv ? l : 0);
        auto load_q = [&](int line) {
            return lv ? rq[(int64_t)line * A.Ls] : make_double4(0.0, 0.0, 0.0, 0.0);
        };
        const int j0s = (int)j0;
        for_each_line_pf<double4>(A.ranges, A.cands, A.ntiles, tile, j0, j1, lane, KEY_NEAR,
                                  load_q, [&](const LineView& v, const double4& q) {
            const float a32 = (float)q.y, sc = (float)q.z;
            // the line's near zone in tile coordinates (grid indices are below 2^31, so these
            // wavefront-uniform tests stay on the scalar unit)
            const int tlo = v.clo - j0s, thi = v.chi - j0s;
#pragma unroll
            for (int t = 0; t < CT; ++t) {
                if (t < tlo || t > thi) continue;  // wavefront-uniform
                const double v64 = fabs(lam[t] - v.lam0) * q.x;
                const int br = hunger_branch(q.y, v64);
                float h;
                if (br == 1 || br == 2) {
                    const int piece = v64 < 1.3 ? 0 : (v64 < 2.4 ? 1 : 2);
                    h = voigt_core_f32(a32, (float)v64, piece, br == 2);
                } else {
                    h = voigt_far_f32(a32, (float)v64, br == 3);
                }
                acc[t] += h * sc;
            }
        });
        if (lv) {
#pragma unroll
            for (int t = 0; t < CT; ++t) {
                const int64_t j = j0 + t;
                if (j <= j1) add_alpha(S.alpha, A.alpha_f32, (int64_t)l + (int64_t)L * (j - go), acc[t]);
            }
        }
    }
}

// The same near zone with lanes = 8 layers x 8 points: lane (r, c) evaluates layer 8g + r of the
// block at point 8p + c of the tile for each layer group g and point group p.  Layers of one
// group have nearly the same damping and Doppler width, so the Hunger branch (and the Harris
// piece) is uniform over a group's 8 x 8 evaluations away from the few boundaries, where the
// one-point-at-a-time form above runs every branch that any of its 64 layers takes.  The line's
// per-layer records come in as one coalesced load per lane (its own layer, as above) and are
// shared through LDS.  Each (layer, point) accumulates the same terms in the same (line) order as
// k_near: the two forms give identical alpha.
constexpr int NEAR_PGROUPS = CT / 8;  // point groups per tile

// NG: layer groups per 64-layer block (7 when no block has more than 56 layers).  The line's
// records go through LDS as (1/(sigma sqrt 2), a) and the scaling; one line's record load is in
// flight while the previous line is evaluated (register-bound: four waves per SIMD).
// F64 (the Float64 series form): the evaluation is the reference's line_profile in Float64, the
// expression of k_profile_f64, with Float64 sums.
template <int NG, bool F64>
__global__ __launch_bounds__(WAVE) void k_near2(NearArgs A) {
    using AccT = typename std::conditional<F64, double, float>::type;
    __shared__ double2 s_q[WAVE];
    __shared__ AccT s_sc[WAVE];
    const int lane = threadIdx.x;
    const int r = lane & 7, c = lane >> 3;
    const int L = A.L;
    if (A.counters[3] != 0) return;
    for (int pass = 0;; ++pass) {
        int item;
        if (A.one_item) {
            if (pass > 0) break;
            item = (int)xcd_item(blockIdx.x, gridDim.x);
        } else {
            item = wave_take(&A.counters[5], lane);
        }
        if (item >= A.n_items) break;
        const int tile = item / A.n_lblocks;
        const int lbase = (item % A.n_lblocks) * WAVE;
        const int nl = min(WAVE, L - lbase);  // layers of this block
        const int l_own = lbase + lane;       // the layer whose record this lane loads
        const bool own_v = lane < nl;
        const int64_t j0 = (int64_t)tile * CT;
        const KCONST StarDev* stars = uniform_ptr(A.stars);
        const KCONST StarDev& S = stars[star_of_point(stars, A.n_stars, j0)];
        const int64_t go = S.grid_off;
        const int64_t j1 = min(j0 + CT, go + S.n_wl) - 1;
        if (j1 < j0) continue;  // padding between two stars' grids: no points
        double lam[NEAR_PGROUPS];  // this lane's points 8p + c (clamped into the tile: in bounds)
        const double* wt = S.wls + (j0 - go);
#pragma unroll
        for (int p = 0; p < NEAR_PGROUPS; ++p) lam[p] = wt[min((int64_t)(8 * p + c), j1 - j0)];
        AccT acc[NG][NEAR_PGROUPS];
#pragma unroll
        for (int g = 0; g < NG; ++g)
#pragma unroll
            for (int p = 0; p < NEAR_PGROUPS; ++p) acc[g][p] = (AccT)0;
        const double4* __restrict__ rq = A.rec64 + (own_v ? l_own : 0);
        auto load_q = [&](int line) {
            return own_v ? rq[(int64_t)line * A.Ls] : make_double4(0.0, 0.0, 0.0, 0.0);
        };
        const int j0s = (int)j0;
        for_each_near_line_pf1<double4>(A.ranges, A.cands, A.ntiles, tile, j0, j1, lane, load_q,
                                        [&](const LineView& v, const double4& q) {
            // LDS operations of one wave complete in order: the reads below see these writes,
            // and the previous line's reads were issued before them
            s_q[lane] = make_double2(q.x, q.y);
            s_sc[lane] = (AccT)q.z;
            const int tlo = v.clo - j0s, thi = v.chi - j0s;
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                if (8 * g >= nl) break;  // wavefront-uniform
                const double2 qq = s_q[8 * g + r];
                const AccT sc = s_sc[8 * g + r];
                const bool gv = 8 * g + r < nl;
                const float a32 = (float)qq.y;
#pragma unroll
                for (int p = 0; p < NEAR_PGROUPS; ++p) {
                    if (8 * p + 7 < tlo || 8 * p > thi) continue;  // wavefront-uniform
                    const int t = 8 * p + c;
                    if (!gv || t < tlo || t > thi) continue;
                    const double v64 = fabs(lam[p] - v.lam0) * qq.x;
                    if constexpr (F64) {
                        acc[g][p] += voigt_hjerting64(qq.y, v64) * sc;
                        continue;
                    }
                    const int br = hunger_branch(qq.y, v64);
                    float h;
                    if (br == 1 || br == 2) {
                        const int piece = v64 < 1.3 ? 0 : (v64 < 2.4 ? 1 : 2);
                        h = voigt_core_f32(a32, (float)v64, piece, br == 2);
                    } else {
                        h = voigt_far_f32(a32, (float)v64, br == 3);
                    }
                    acc[g][p] += h * sc;
                }
            }
        });
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const int l = lbase + 8 * g + r;
            if (8 * g + r >= nl) continue;
#pragma unroll
            for (int p = 0; p < NEAR_PGROUPS; ++p) {
                const int64_t j = j0 + 8 * p + c;
                if (j > j1) continue;
                if (!F64 && A.nbuf) A.nbuf[(int64_t)l + (int64_t)L * j] = (float)acc[g][p];
                else add_alpha(S.alpha, A.alpha_f32, (int64_t)l + (int64_t)L * (j - go), acc[g][p]);
            }
        }
    }
}

// alpha += the near-zone sums stored by k_near2 (the same addition it makes itself when it runs
// alone, so the two schedules give identical alpha).  One thread per (point, layer) of the
// stars' grids; padding points are skipped.
__global__ __launch_bounds__(256) void k_add_near(int n_stars, int L, int64_t n_total,
                                                  const StarDev* __restrict__ stars,
                                                  const float* __restrict__ nbuf, int alpha_f32,
                                                  const unsigned long long* __restrict__ counters) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    // a raised flag: k_near2 wrote nothing (alpha is left unchanged)
    if (e >= n_total * L || counters[3] != 0) return;
    const int64_t j = e / L;
    const int l = (int)(e - j * L);
    const StarDev& S = stars[star_of_point(stars, n_stars, j)];
    if (j >= S.grid_off + S.n_wl) return;  // padding between two stars' grids
    add_alpha(S.alpha, alpha_f32, (int64_t)l + (int64_t)L * (j - S.grid_off), nbuf[e]);
}

// ---------------------------------------------------------------------------------------------
// Adjoint (vector-Jacobian product) of the line opacity, in Float64, over a batch of stars.  Given
// G = dL/dalpha per star, an item is (line, chunk of up to VJP_CHUNK window points); lanes =
// layers (layers l, l + 64, ... when L > 64).  Each lane sums over the item's points, in point
// order,
//   S_H = sum G H,  S_a = sum G dH/da,  S_v = sum G v dH/dv,
// and turns them into the item's share of dL/dx for its layer,
//   scaling (da/dx S_a + dln(1/(sigma sqrt 2))/dx S_v) + dscaling/dx S_H,
// for x = T, n_e, xi, n_HI, n_HeI, n_H2 and the n_div_U of the line's species.  The windows are
// integer functions of the inputs and are held fixed.
// A segment is up to VJP_SEG consecutive items of one star, handled by one wavefront in item order;
// it leaves one partial sum per output (k_vjp), and k_vjp_reduce sums a star's segments in a fixed
// order: the gradients are deterministic, and the scratch is per segment, not per item.
// ---------------------------------------------------------------------------------------------
// No point of the window lies outside the near zone [z.x, z.y] (or the window is empty).
__device__ inline bool far_zone_empty(const LineMeta& m, int2 z) {
    return m.lb > m.ub || (z.x <= m.lb && z.y >= m.ub);
}

// The adjoint near zone of 16-layer block q of a line whose zone is z, from its azq entry e
// (offsets from z.x in the low 16 bits, from z.y in the high 16 bits).
__device__ inline int2 azone_blk(int2 z, int e) {
    return make_int2(z.x + (e & 0xFFFF), z.y - (int)((uint32_t)e >> 16));
}
// The innermost of a line's block zones (all of them nest about the centre): outside it some
// block's layers take the moments' series.  Without azq, the line's zone.
__device__ inline int2 azone_in(int2 z, const int4* __restrict__ azq, int64_t i) {
    if (!azq) return z;
    const int4 e = azq[i];
    const int lo = max(max(e.x & 0xFFFF, e.y & 0xFFFF), max(e.z & 0xFFFF, e.w & 0xFFFF));
    const int hi = (int)max(max((uint32_t)e.x >> 16, (uint32_t)e.y >> 16),
                            max((uint32_t)e.z >> 16, (uint32_t)e.w >> 16));
    return make_int2(z.x + lo, z.y - hi);
}

constexpr int VJP_CHUNK = 2048;
constexpr int VJP_SEG = 64;
// Points whose loads are in flight together: 8 for the Float32 evaluation, 2 for the Float64
// one (whose registers allow no more at two waves per SIMD).
#ifndef KORG_ADJ_NEAR_F64
#define KORG_ADJ_NEAR_F64 0
#endif
template <bool SERIES>
constexpr int vjp_batch() { return SERIES ? 8 : 2; }

// Items per line: its window in chunks (exact form), or (series form: the far zone comes from the
// moments) its near zone in chunks, at least one item for a line with a window (the first item of
// a line carries the far-zone sums).
// (azone: the adjoint's near zones, given in the series form.)
__global__ __launch_bounds__(256) void k_vjp_count(int64_t n, const LineMeta* __restrict__ meta,
                                                   const int2* __restrict__ azone,
                                                   int* __restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const LineMeta m = meta[i];
    if (m.lb > m.ub) {
        cnt[i] = 0;
    } else if (!azone) {
        cnt[i] = (m.ub - m.lb + VJP_CHUNK) / VJP_CHUNK;
    } else {
        const int2 z = azone[i];
        const int nz = z.x <= z.y ? z.y - z.x + 1 : 0;
        cnt[i] = nz > 0 ? (nz + VJP_CHUNK - 1) / VJP_CHUNK : 1;
    }
}

// Segments per star from the items' exclusive scan over all lines (item_start[N] = total).
__global__ void k_vjp_segs(int n_stars, const StarDev* __restrict__ stars,
                           const int64_t* __restrict__ item_start, int* __restrict__ seg_cnt) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_stars) return;
    const StarDev& S = stars[s];
    const int64_t n = item_start[S.line_off + S.n_lines] - item_start[S.line_off];
    seg_cnt[s] = (int)((n + VJP_SEG - 1) / VJP_SEG);
}

// Partial derivatives of one (line, layer)'s profile parameters -- line_profile is
// scaling * H(a, v) with inv = 1/(sigma sqrt 2), a = gamma inv, v = |lambda - lambda0| inv and
// scaling = inv/sqrt(pi) * amplitude (src/line_absorption.jl:245-249) -- with respect to the
// layer's inputs x = (T, n_e, xi, n_HI, n_HeI, n_H2, n/U of the line's species): da[q] = da/dx_q,
// dli[q] = d ln(inv)/dx_q, ds[q] = dscaling/dx_q (src/line_absorption.jl:86-106, 190).
struct LayerPartials {
    double da[7], dli[7], ds[7];
};

struct LinePartialsIn {  // the line's own constants
    double mass, lam0, E_lower, E_up, gf, sl, gscale, gamma_stark, vdW1, vdW2;
    bool mol;
};

__device__ inline LayerPartials layer_partials(const LinePartialsIn& ln, const LayerTab& tb,
                                               double T, double xi, double n_e, double nU,
                                               double inv, double a, double scaling) {
    const double a_H = 4.50711, a_He = 1.38375, a_H2 = 5.503;
    const double m_H = 1.008 * amu_cgs, m_He = 4.003 * amu_cgs;  // as k_layers
    const double c_He = pow(a_He / a_H, 0.4) * pow(m_He / m_H, -0.3);
    const double c_H2 = pow(a_H2 / a_H, 0.4) * pow(2.0, -0.3);
    // d ln(1/(sigma sqrt 2)) = -d ln sigma, sigma^2 ~ kT/m + xi^2/2 (src/line_absorption.jl:190)
    const double q = kboltz_cgs * T / ln.mass + xi * xi / 2.0;
    const double dli_dT = -(kboltz_cgs / ln.mass) / (2.0 * q);
    const double dli_dxi = -xi / (2.0 * q);
    // dGamma/dx (src/line_absorption.jl:90-99)
    double dG_dT = 0.0, dG_dne = 0.0, vdw = 0.0;
    if (!ln.mol) {
        const double stk = ln.gamma_stark * tb.t16;
        vdw = (ln.vdW2 == -1.0) ? ln.vdW1 * tb.t03 : scaled_vdW_abo(ln.vdW1, ln.vdW2, ln.mass, T);
        const double dvdw_dT = (ln.vdW2 == -1.0 ? 0.3 : (1.0 - ln.vdW2) / 2.0) * vdw / T;
        dG_dne = stk;
        dG_dT = n_e * stk / (6.0 * T) + tb.n_eff * dvdw_dT;
    }
    const double ig = inv * ln.gscale;
    // scaling = inv / sqrt(pi) * gf * sigma_line * levels * n/U (src/line_absorption.jl:101-106)
    const double el = exp(-tb.beta * ln.E_lower), eu = exp(-tb.beta * ln.E_up);
    const double dlev_dT = (tb.beta / T) * (ln.E_lower * el - ln.E_up * eu);
    const double base = inv * INV_SQRT_PI * ln.gf * ln.sl;
    LayerPartials o;
    o.da[0] = ig * dG_dT + a * dli_dT;
    o.da[1] = ig * dG_dne;
    o.da[2] = a * dli_dxi;
    o.da[3] = ig * vdw;
    o.da[4] = ig * c_He * vdw;
    o.da[5] = ig * c_H2 * vdw;
    o.da[6] = 0.0;
    for (int k = 0; k < 7; ++k) o.dli[k] = 0.0;
    o.dli[0] = dli_dT;
    o.dli[2] = dli_dxi;
    for (int k = 0; k < 7; ++k) o.ds[k] = 0.0;
    o.ds[0] = scaling * dli_dT + base * nU * dlev_dT;
    o.ds[2] = scaling * dli_dxi;
    o.ds[6] = base * (el - eu);
    return o;
}

__device__ inline LinePartialsIn line_partials_in(const KCONST StarDev& S, int64_t li, double lam0,
                                                  const double* mass, const uint8_t* is_mol) {
    LinePartialsIn ln;
    const int32_t s = S.species[li];
    ln.mass = mass[s];
    ln.mol = is_mol[s] != 0;
    ln.lam0 = lam0;
    ln.E_lower = S.E_lower[li];
    ln.E_up = ln.E_lower + c_cgs * hplanck_eV / lam0;  // src/line_absorption.jl:101
    ln.gf = exp10(S.log_gf[li]);
    ln.sl = sigma_line(lam0);
    ln.gscale = (lam0 * lam0) / (c_cgs * (4.0 * PI));
    ln.gamma_stark = S.gamma_stark[li];
    ln.vdW1 = S.vdW1[li];
    ln.vdW2 = S.vdW2[li];
    return ln;
}

struct VjpArgs {
    int n_stars;
    int L, Ls, n_species;
    const StarDev* stars;
    const LineMeta* meta;
    const double4* rec64;
    const LayerTab* tab;          // [star * L + l]
    const int64_t* item_start;    // [N + 1]: first item of each line (global)
    const int64_t* seg_start;     // [n_stars + 1]: first segment of each star
    const double* mass;
    const uint8_t* is_mol;
    int g_f32;
    double* partial;              // [segment][6 + n_species][L]
    int64_t seg_cap;              // segments the partials hold
    const unsigned long long* counters;
    // series form: items cover the adjoint's near zones (evaluated in Float32), and a line's
    // first item adds its far-zone sums from the moments (k_vjp_moments' layout, G scaled by
    // 2^gexp[star * L + l])
    int n_lblocks;
    const int2* azone;
    const int4* azq;  // each 16-layer block's adjoint zone (null: the line's)
    const float4* mom;
    const int* gexp;
};

template <bool SERIES>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(SERIES ? 3 : 1)))
void k_vjp(VjpArgs A) {
    const int lane = threadIdx.x;
    const int L = A.L;
    const int NQ = 6 + A.n_species;
    if (A.counters[3] != 0) return;  // failed setup, or the segments do not fit
    const int64_t n_segs = A.seg_start[A.n_stars];
    for (int64_t seg = blockIdx.x; seg < n_segs; seg += gridDim.x) {
        // the star of this segment and its item range
        int st = 0;
        {
            int lo = 0, hi = A.n_stars;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (A.seg_start[mid] <= seg) lo = mid; else hi = mid;
            }
            st = lo;
        }
        const KCONST StarDev* stars = uniform_ptr(A.stars);
        const KCONST StarDev& S = stars[st];
        const int64_t go = S.grid_off;
        const int64_t first = A.item_start[S.line_off];
        const int64_t it0 = first + (seg - A.seg_start[st]) * VJP_SEG;
        const int64_t it1 = min(it0 + VJP_SEG, A.item_start[S.line_off + S.n_lines]);
        double* part = A.partial + seg * NQ * L;
        for (int q = 0; q < NQ; ++q)
            for (int l = lane; l < L; l += WAVE) part[q * L + l] = 0.0;
        // lines of the segment's items: the largest line with item_start <= item, found once
        int64_t i = S.line_off;
        {
            int64_t lo = S.line_off, hi = S.line_off + S.n_lines;
            while (hi - lo > 1) {
                const int64_t mid = (lo + hi) / 2;
                if (A.item_start[mid] <= it0) lo = mid; else hi = mid;
            }
            i = lo;
        }
        for (int64_t item = it0; item < it1; ++item) {
            while (A.item_start[i + 1] <= item) ++i;  // lines with no items are passed over
            const int64_t li = i - S.line_off;
            const LineMeta m = A.meta[i];
            const bool first = item == A.item_start[i];
            // the item's points: a chunk of the window, or (series form) of the near zone
            const int2 z = SERIES ? A.azone[i] : make_int2(m.lb, m.ub);
            const int64_t j0 = z.x + (item - A.item_start[i]) * VJP_CHUNK;
            const int64_t j1 = z.x > z.y ? j0 : min(j0 + VJP_CHUNK, (int64_t)z.y + 1);  // exclusive
            const bool far = SERIES && first && !far_zone_empty(m, z);
            const double lam0 = m.lam0;
            const int32_t s = S.species[li];
            const LinePartialsIn lp = line_partials_in(S, li, lam0, A.mass, A.is_mol);
            for (int l = lane; l < L; l += WAVE) {
                const double4 r = A.rec64[i * A.Ls + l];
                const double inv = r.x, a = r.y, scaling = r.z;
                double SH = 0.0, Sa = 0.0, Sv = 0.0;
                // points in batches: the batch's loads are issued before its evaluations (points
                // past j1 load a stand-in and are skipped)
                constexpr int NBT = vjp_batch<SERIES>();
                for (int64_t jb = j0; jb < j1; jb += NBT) {
                    double gv[NBT], lv[NBT];
#pragma unroll
                    for (int u = 0; u < NBT; ++u) {
                        const int64_t jl = min(jb + u, j1 - 1) - go;  // the star's own index
                        gv[u] = A.g_f32 ? (double)((const float*)S.G)[l + (int64_t)L * jl]
                                        : ((const double*)S.G)[l + (int64_t)L * jl];
                        lv[u] = S.wls[jl];
                    }
#pragma unroll
                    for (int u = 0; u < NBT; ++u) {
                        if (jb + u >= j1) break;
                        const double g = gv[u];
                        const double v = fabs(lv[u] - lam0) * inv;
                        if (SERIES && !KORG_ADJ_NEAR_F64) {
                            float Ha, Hv;
                            const float H = voigt_hjerting_grad_f32(a, v, Ha, Hv);
                            SH += g * H;
                            Sa += g * Ha;
                            Sv += g * (v * Hv);
                        } else {
                            double Ha, Hv;
                            const double H = voigt_hjerting64_grad(a, v, Ha, Hv);
                            SH += g * H;
                            Sa += g * Ha;
                            Sv += g * v * Hv;
                        }
                    }
                }
                if (far) {
                    // far-zone sums from the moments: Q_k = q^k M_k, e_k and (a e_k)' of the
                    // series (branch A: a <= 0.2, else D)
                    const float4 mk = A.mom[(i * A.n_lblocks + l / WAVE) * WAVE + l % WAVE];
                    const int e = A.gexp[(int64_t)st * L + l];
                    const double w = 1.0 / (inv * DELTA_UNIT);
                    const double q = w * w;
                    const double Q1 = ldexp((double)mk.x, e) * q;
                    const double Q2 = ldexp((double)mk.y, e) * (q * q);
                    const double Q3 = ldexp((double)mk.z, e) * (q * q * q);
                    const double Q4 = ldexp((double)mk.w, e) * ((q * q) * (q * q));
                    double e2 = 1.5, e3 = 3.75, e4 = 0.0, f2 = 1.5, f3 = 3.75, f4 = 0.0;
                    if (!(a <= 0.2)) {
                        const double a2 = a * a;
                        e2 = 1.5 - a2;
                        e3 = 3.75 - 5.0 * a2 + a2 * a2;
                        e4 = -26.25 * a2 + 10.5 * a2 * a2 - a2 * a2 * a2;
                        // (a e_k)' = e_k + a e_k'
                        f2 = 1.5 - 3.0 * a2;
                        f3 = 3.75 - 15.0 * a2 + 5.0 * a2 * a2;
                        f4 = -78.75 * a2 + 52.5 * a2 * a2 - 7.0 * a2 * a2 * a2;
                    }
                    SH += a * INV_SQRT_PI * (Q1 + e2 * Q2 + e3 * Q3 + e4 * Q4);
                    Sa += INV_SQRT_PI * (Q1 + f2 * Q2 + f3 * Q3 + f4 * Q4);
                    Sv += a * INV_SQRT_PI * (-2.0 * Q1 - 4.0 * e2 * Q2 - 6.0 * e3 * Q3 - 8.0 * e4 * Q4);
                }
                const LayerTab tb = A.tab[(int64_t)st * L + l];
                const LayerPartials d = layer_partials(lp, tb, S.temps[l], S.xi[l], S.n_e[l],
                                                       S.n_div_U[(int64_t)s * L + l], inv, a,
                                                       scaling);
                part[0 * L + l] += scaling * (d.da[0] * Sa + d.dli[0] * Sv) + d.ds[0] * SH;
                part[1 * L + l] += scaling * (d.da[1] * Sa);
                part[2 * L + l] += scaling * (d.da[2] * Sa + d.dli[2] * Sv) + d.ds[2] * SH;
                part[3 * L + l] += scaling * (d.da[3] * Sa);
                part[4 * L + l] += scaling * (d.da[4] * Sa);
                part[5 * L + l] += scaling * (d.da[5] * Sa);
                part[(6 + s) * L + l] += d.ds[6] * SH;
            }
        }
    }
}

// The series form's near-zone items with lanes over 8 layers x 8 points (blocks of at most 64
// layers): lane (r, c) = (lane % 8, lane / 8) evaluates layer 8 g + r of each layer group g at
// points j0 + 8 p + c, so a wavefront's evaluations share their Hunger branch and Harris piece away
// from the few boundaries (with lanes over layers every point runs every branch its layers take).
// The line's records go through LDS; each lane sums its points in Float64, the 8 point lanes of a
// layer are summed by shuffles (a fixed order), and the group's sums land in lane 8 g + r, which
// then applies the far-zone moments and the chain rule exactly as k_vjp<true> does.
template <int NG>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(4))) void k_vjp_near(VjpArgs A) {
    __shared__ double2 s_q[WAVE];
    const int lane = threadIdx.x;
    const int r = lane & 7, c = lane >> 3;
    const int L = A.L;
    const int NQ = 6 + A.n_species;
    if (A.counters[3] != 0) return;  // failed setup, or the segments do not fit
    const int64_t n_segs = A.seg_start[A.n_stars];
    for (int64_t seg = blockIdx.x; seg < n_segs; seg += gridDim.x) {
        int st = 0;
        {
            int lo = 0, hi = A.n_stars;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (A.seg_start[mid] <= seg) lo = mid; else hi = mid;
            }
            st = lo;
        }
        const KCONST StarDev* stars = uniform_ptr(A.stars);
        const KCONST StarDev& S = stars[st];
        const int64_t go = S.grid_off;
        const int64_t first = A.item_start[S.line_off];
        const int64_t it0 = first + (seg - A.seg_start[st]) * VJP_SEG;
        const int64_t it1 = min(it0 + VJP_SEG, A.item_start[S.line_off + S.n_lines]);
        double* part = A.partial + seg * NQ * L;
        for (int q = 0; q < NQ; ++q)
            for (int l = lane; l < L; l += WAVE) part[q * L + l] = 0.0;
        int64_t i = S.line_off;
        {
            int64_t lo = S.line_off, hi = S.line_off + S.n_lines;
            while (hi - lo > 1) {
                const int64_t mid = (lo + hi) / 2;
                if (A.item_start[mid] <= it0) lo = mid; else hi = mid;
            }
            i = lo;
        }
        for (int64_t item = it0; item < it1; ++item) {
            while (A.item_start[i + 1] <= item) ++i;  // lines with no items are passed over
            const int64_t li = i - S.line_off;
            const LineMeta m = A.meta[i];
            const bool first_item = item == A.item_start[i];
            const int2 z = A.azone[i];
            const int64_t j0 = z.x + (item - A.item_start[i]) * VJP_CHUNK;
            const int64_t j1 = z.x > z.y ? j0 : min(j0 + VJP_CHUNK, (int64_t)z.y + 1);  // exclusive
            const bool far = first_item && !far_zone_empty(m, azone_in(z, A.azq, i));
            const int4 zq = A.azq ? A.azq[i] : make_int4(0, 0, 0, 0);
            const double lam0 = m.lam0;
            const int32_t s = S.species[li];
            const LinePartialsIn lp = line_partials_in(S, li, lam0, A.mass, A.is_mol);
            const double4 rown = lane < L ? A.rec64[i * A.Ls + lane] : make_double4(0, 0, 0, 0);
            // (a wavefront's LDS operations complete in order: the reads below see these writes,
            // and the previous item's reads were issued before them)
            s_q[lane] = make_double2(rown.x, rown.y);
            double mySH = 0.0, mySa = 0.0, mySv = 0.0;  // layer `lane`'s sums
#pragma unroll 1
            for (int g = 0; g < NG; ++g) {
                if (8 * g >= L) break;  // wavefront-uniform
                const int lg = 8 * g + r;
                const bool gvld = lg < L;
                const double2 qq = s_q[lg];
                double SH = 0.0, Sa = 0.0, Sv = 0.0;
                // the points of the chunk in the zone of the group's 16-layer block (beyond it the
                // block's layers take the moments' series)
                const int eq = (g >> 1) == 0 ? zq.x : (g >> 1) == 1 ? zq.y : (g >> 1) == 2 ? zq.z : zq.w;
                const int2 zb = azone_blk(z, eq);
                const int64_t ja = max(j0, (int64_t)zb.x), jz = min(j1, (int64_t)zb.y + 1);
                constexpr int NPB = 4;  // point steps whose loads are in flight together
                for (int64_t jb = ja; jb < jz; jb += 8 * NPB) {
                    double gv[NPB], lv[NPB];
#pragma unroll
                    for (int u = 0; u < NPB; ++u) {
                        const int64_t jl = min(jb + 8 * u + c, jz - 1) - go;  // clamped: in bounds
                        const int64_t idx = (int64_t)(gvld ? lg : 0) + (int64_t)L * jl;
                        gv[u] = A.g_f32 ? (double)((const float*)S.G)[idx] : ((const double*)S.G)[idx];
                        lv[u] = S.wls[jl];
                    }
#pragma unroll
                    for (int u = 0; u < NPB; ++u) {
                        if (!gvld || jb + 8 * u + c >= jz) continue;
                        const double gg = gv[u];
                        const double v = fabs(lv[u] - lam0) * qq.x;
                        float Ha, Hv;
                        const float H = voigt_hjerting_grad_f32(qq.y, v, Ha, Hv);
                        SH += gg * H;
                        Sa += gg * Ha;
                        Sv += gg * (v * Hv);
                    }
                }
                // the 8 point lanes of each layer (lanes r, r + 8, ..., r + 56), summed in a fixed
                // order; lane 8 g + r keeps layer 8 g + r
#pragma unroll
                for (int off = 8; off < WAVE; off <<= 1) {
                    SH += __shfl_xor(SH, off, WAVE);
                    Sa += __shfl_xor(Sa, off, WAVE);
                    Sv += __shfl_xor(Sv, off, WAVE);
                }
                if (c == g) {
                    mySH = SH;
                    mySa = Sa;
                    mySv = Sv;
                }
            }
            const int l = lane;
            if (l < L) {
                const double inv = rown.x, a = rown.y, scaling = rown.z;
                double SH = mySH, Sa = mySa, Sv = mySv;
                if (far) {
                    // far-zone sums from the moments (as k_vjp<true>)
                    const float4 mk = A.mom[(i * A.n_lblocks) * WAVE + l];
                    const int e = A.gexp[(int64_t)st * L + l];
                    const double w = 1.0 / (inv * DELTA_UNIT);
                    const double q = w * w;
                    const double Q1 = ldexp((double)mk.x, e) * q;
                    const double Q2 = ldexp((double)mk.y, e) * (q * q);
                    const double Q3 = ldexp((double)mk.z, e) * (q * q * q);
                    const double Q4 = ldexp((double)mk.w, e) * ((q * q) * (q * q));
                    double e2 = 1.5, e3 = 3.75, e4 = 0.0, f2 = 1.5, f3 = 3.75, f4 = 0.0;
                    if (!(a <= 0.2)) {
                        const double a2 = a * a;
                        e2 = 1.5 - a2;
                        e3 = 3.75 - 5.0 * a2 + a2 * a2;
                        e4 = -26.25 * a2 + 10.5 * a2 * a2 - a2 * a2 * a2;
                        f2 = 1.5 - 3.0 * a2;
                        f3 = 3.75 - 15.0 * a2 + 5.0 * a2 * a2;
                        f4 = -78.75 * a2 + 52.5 * a2 * a2 - 7.0 * a2 * a2 * a2;
                    }
                    SH += a * INV_SQRT_PI * (Q1 + e2 * Q2 + e3 * Q3 + e4 * Q4);
                    Sa += INV_SQRT_PI * (Q1 + f2 * Q2 + f3 * Q3 + f4 * Q4);
                    Sv += a * INV_SQRT_PI * (-2.0 * Q1 - 4.0 * e2 * Q2 - 6.0 * e3 * Q3 - 8.0 * e4 * Q4);
                }
                const LayerTab tb = A.tab[(int64_t)st * L + l];
                const LayerPartials d = layer_partials(lp, tb, S.temps[l], S.xi[l], S.n_e[l],
                                                       S.n_div_U[(int64_t)s * L + l], inv, a,
                                                       scaling);
                part[0 * L + l] += scaling * (d.da[0] * Sa + d.dli[0] * Sv) + d.ds[0] * SH;
                part[1 * L + l] += scaling * (d.da[1] * Sa);
                part[2 * L + l] += scaling * (d.da[2] * Sa + d.dli[2] * Sv) + d.ds[2] * SH;
                part[3 * L + l] += scaling * (d.da[3] * Sa);
                part[4 * L + l] += scaling * (d.da[4] * Sa);
                part[5 * L + l] += scaling * (d.da[5] * Sa);
                part[(6 + s) * L + l] += d.ds[6] * SH;
            }
        }
    }
}

// One block per (star, output): output q < 6 is (quantity q, layer l), q = 6 + s the n_div_U of
// species s at layer l.  The star's segments are summed in a fixed per-thread order and a fixed
// tree (deterministic) and the sum goes to the star's output array of that quantity.
__global__ __launch_bounds__(256) void k_vjp_reduce(int L, int n_species,
                                                    const StarDev* __restrict__ stars,
                                                    const int64_t* __restrict__ seg_start,
                                                    const double* __restrict__ partial,
                                                    const unsigned long long* __restrict__ counters) {
    __shared__ double red[256];
    if (counters[3] != 0) return;  // nothing was summed (the segments did not fit)
    const int NQ = 6 + n_species;
    const int st = blockIdx.y;
    const int b = blockIdx.x, t = threadIdx.x;  // b = q * L + l
    const int64_t s0 = seg_start[st], s1 = seg_start[st + 1];
    double sum = 0.0;
    for (int64_t sg = s0 + t; sg < s1; sg += 256) sum += partial[sg * NQ * L + b];
    red[t] = sum;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) {
        const StarDev& S = stars[st];
        const int q = b / L, l = b % L;
        double* dst = S.gq[q < 6 ? q : 6];
        if (dst) dst[q < 6 ? l : (int64_t)(q - 6) * L + l] = red[0];
    }
}

// ---------------------------------------------------------------------------------------------
// Forward mode (Jacobian-vector products): for up to JVP_DIRS tangent directions x_k of the
// per-layer inputs, d alpha_k = sum over lines of (ds.x_k) H + scaling ((da.x_k) dH/da +
// (dli.x_k) v dH/dv) over each line's window, in Float64, windows held fixed.
// ---------------------------------------------------------------------------------------------
constexpr int JVP_DIRS = 4;

struct JvpTangent {  // device arrays of one direction (null: zero tangent)
    const double *T, *n_e, *xi, *n_HI, *n_HeI, *n_H2, *n_div_U;
};

struct JvpArgs {
    int n_dirs;
    int L, Ls;
    const StarDev* stars;
    const LineMeta* meta;
    const double4* rec64;
    const LayerTab* tab;
    const double* mass;
    const uint8_t* is_mol;
    JvpTangent t[JVP_DIRS];
    double* coef;  // [(line * Ls + l) * JVP_DIRS + k][3]: coefficients of H, dH/da, v dH/dv
    // profile pass
    int64_t n_wl;
    int ntiles, tile_w, n_lblocks;
    int64_t n_items;
    const Cand* cands;
    const int2* ranges;
    void* out[JVP_DIRS];  // d alpha of each direction, column-major [L x n_wl]
    int out_f32;
    unsigned long long* counters;
};

// One wavefront per line, lane = layer: the coefficients (ds.x, scaling da.x, scaling dli.x) of
// each direction.  A zero tangent component contributes nothing (so a NaN partial, e.g. of a
// missing perturber density, only matters where its input moves).
__global__ __launch_bounds__(256) void k_jvp_coef(int64_t n_lines, JvpArgs A) {
    const int lane = threadIdx.x % WAVE;
    const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const int64_t i = (int64_t)blockIdx.x * (blockDim.x / WAVE) + wid;
    if (i >= n_lines) return;
    const KCONST StarDev& S = uniform_ptr(A.stars)[0];
    const int L = A.L;
    const int32_t s = S.species[i];
    const bool ok = A.counters[3] == 0;  // a failed setup: the species index may be invalid
    LinePartialsIn lp = {};
    if (ok) lp = line_partials_in(S, i, A.meta[i].lam0, A.mass, A.is_mol);
    for (int l = lane; l < A.Ls; l += WAVE) {
        double* c = A.coef + ((i * A.Ls + l) * JVP_DIRS) * 3;
        double cf[JVP_DIRS][3] = {};
        if (ok && l < L) {
            const double4 r = A.rec64[i * A.Ls + l];
            const LayerPartials d = layer_partials(lp, A.tab[l], S.temps[l], S.xi[l], S.n_e[l],
                                                   S.n_div_U[(int64_t)s * L + l], r.x, r.y, r.z);
            for (int k = 0; k < A.n_dirs; ++k) {
                const JvpTangent& t = A.t[k];
                const double x[7] = {t.T ? t.T[l] : 0.0,       t.n_e ? t.n_e[l] : 0.0,
                                     t.xi ? t.xi[l] : 0.0,     t.n_HI ? t.n_HI[l] : 0.0,
                                     t.n_HeI ? t.n_HeI[l] : 0.0, t.n_H2 ? t.n_H2[l] : 0.0,
                                     t.n_div_U ? t.n_div_U[(int64_t)s * L + l] : 0.0};
                double dsx = 0.0, dax = 0.0, dlx = 0.0;
#pragma unroll
                for (int q = 0; q < 7; ++q) {
                    if (x[q] == 0.0) continue;
                    dsx += d.ds[q] * x[q];
                    dax += d.da[q] * x[q];
                    dlx += d.dli[q] * x[q];
                }
                cf[k][0] = dsx;
                cf[k][1] = r.z * dax;
                cf[k][2] = r.z * dlx;
            }
        }
#pragma unroll
        for (int k = 0; k < JVP_DIRS; ++k)
#pragma unroll
            for (int q = 0; q < 3; ++q) c[k * 3 + q] = cf[k][q];
    }
}

// The profile pass: a work item is (64-point tile, group of G layers), the lines visited in the
// Float64 form's fixed order (window classes, partition order), one evaluation of H and its two
// partials per (line, layer, point) shared by the directions.
template <int G>
__global__ __launch_bounds__(WAVE) void k_jvp_profile(JvpArgs A) {
    const int lane = threadIdx.x;
    const int L = A.L;
    if (A.counters[3] != 0) return;  // failed setup: nothing is written
    const KCONST StarDev& S = uniform_ptr(A.stars)[0];
    for (;;) {
        const int item = wave_take(&A.counters[0], lane);
        if (item >= A.n_items) break;
        const int tile = item / A.n_lblocks;
        const int l0 = (item % A.n_lblocks) * G;
        const int64_t j0 = (int64_t)tile * A.tile_w;
        const int64_t j1 = min(j0 + A.tile_w, S.n_wl) - 1;
        const int jj = (int)(j0 + lane);
        const double lam = S.wls[min((int64_t)jj, j1)];  // points past j1 are never used
        double acc[JVP_DIRS][G];
#pragma unroll
        for (int k = 0; k < JVP_DIRS; ++k)
#pragma unroll
            for (int l = 0; l < G; ++l) acc[k][l] = 0.0;
        auto no_operand = [](int) { return 0; };
        for_each_line_pf<int>(A.ranges, A.cands, A.ntiles, tile, j0, j1, lane, KEY_WINDOW,
                              no_operand, [&](const LineView& v, int) {
            const bool inw = jj >= v.lb && jj <= v.ub;
            const double dl = fabs(lam - v.lam0);
            const KCONST double4* rp = uniform_ptr(A.rec64 + (int64_t)v.line * A.Ls + l0);
            const KCONST double* cp =
                uniform_ptr(A.coef + ((int64_t)v.line * A.Ls + l0) * JVP_DIRS * 3);
#pragma unroll
            for (int l = 0; l < G; ++l) {
                const double4 r = rp[l];
                const double vv = dl * r.x;
                double Ha, Hv;
                const double H = voigt_hjerting64_grad(r.y, vv, Ha, Hv);
                const double vHv = vv * Hv;
#pragma unroll
                for (int k = 0; k < JVP_DIRS; ++k) {
                    const double* f = cp + (l * JVP_DIRS + k) * 3;
                    const double t = f[0] * H + f[1] * Ha + f[2] * vHv;
                    acc[k][l] += inw ? t : 0.0;
                }
            }
        });
        if (jj > j1) continue;
#pragma unroll
        for (int k = 0; k < JVP_DIRS; ++k) {
            if (k >= A.n_dirs) break;
#pragma unroll
            for (int l = 0; l < G; ++l)
                if (l0 + l < L)
                    add_alpha(A.out[k], A.out_f32, (int64_t)(l0 + l) + (int64_t)L * jj, acc[k][l]);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Radiative transfer on the device, for the fit step: Korg's plane-parallel default
// (src/RadiativeTransfer/RadiativeTransfer.jl:62-66, 77-146): the anchored optical depth (:244-253)
// with the blackbody source (src/synthesize.jl:285, 412-417) and the flux-only linear-source
// solution in exponential integrals (:379-400), with Korg's approximate E2 (:408-480).  One thread
// per wavelength point; the flux of alpha_cntm + alpha_lines and of alpha_cntm alone give the
// normalized flux, chi^2 = sum ((F / F_c - data) / sigma)^2, and by a reverse sweep over the layers
// G = dchi^2/dalpha_lines, which the line adjoint consumes.  Float64.
// ---------------------------------------------------------------------------------------------
// Korg's E2 pieces (src/RadiativeTransfer/RadiativeTransfer.jl:427-480): polynomials of degree 5
// in y = x - n about n = 2..8, their coefficients as published there.
__constant__ double E2_POLY[7][6] = {
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

// Korg's E2 (:408-480) and, when d is given, the derivative of the approximation.
__device__ inline double expint2_korg(double x, double* d) {
    if (x == 0.0) {
        if (d) *d = -INFINITY;
        return 1.0;
    }
    if (x < 1.1) {
        const double g = 0.57721566490153286060651209008240243104215933593992;
        const double P = -0.5 + (0.08333333333333333 + (-0.013888888888888888 +
                                                        0.0020833333333333333 * x) * x) * x;
        const double Q = (log(x) + g - 1) + P * x;
        if (d) {
            const double dP = 0.08333333333333333 + (2 * -0.013888888888888888 +
                                                     3 * 0.0020833333333333333 * x) * x;
            *d = Q + 1 + x * (P + x * dP);
        }
        return 1 + Q * x;
    }
    if (x >= 9.0) {
        const double u = 1 / x;
        const double R = 1 + (-2 + (6 + (-24 + 120 * u) * u) * u) * u;
        const double ex = exp(-x);
        const double e = ex * R * u;
        if (d) {
            const double dR = -2 + (12 + (-72 + 480 * u) * u) * u;
            *d = -e - ex * (R + u * dR) * (u * u);
        }
        return e;
    }
    // polynomials of degree 5 in y = x - n about n = 2..8 (coefficients of :427-480)
    const int n = x < 2.5 ? 2 : x < 3.5 ? 3 : x < 4.5 ? 4 : x < 5.5 ? 5 : x < 6.5 ? 6 : x < 7.5 ? 7 : 8;
    const double* c = E2_POLY[n - 2];
    const double y = x - n;
    if (d) *d = c[1] + (2 * c[2] + (3 * c[3] + (4 * c[4] + 5 * c[5] * y) * y) * y) * y;
    return c[0] + (c[1] + (c[2] + (c[3] + (c[4] + c[5] * y) * y) * y) * y) * y;
}

// Per star and layer: f = tau_ref / alpha_ref and ln tau_ref (the anchored scheme's factors).
struct RtTab {
    double f, ltau;
};

// Also resets the per-(star, layer) scale exponents of G (when given) for k_g_exp's maxima.
__global__ void k_rt_tab(int n_stars, int L, const StarDev* __restrict__ stars, RtTab* __restrict__ tab,
                         int* __restrict__ gexp, int gexp_none) {
    const int st = blockIdx.x;
    const StarDev& S = stars[st];
    for (int l = threadIdx.x; l < L; l += blockDim.x) {
        tab[(int64_t)st * L + l] = RtTab{S.tau_ref[l] / S.alpha_ref[l], log(S.tau_ref[l])};
        if (gexp) gexp[(int64_t)st * L + l] = gexp_none;
    }
}

// Zeroes every star's alpha: thread t of the global (point, layer) space.
__global__ void k_zero_alpha(int n_stars, int L, int64_t n_total, const StarDev* __restrict__ stars,
                             int alpha_f32) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_total * L;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / L;
        const int l = (int)(t % L);
        const StarDev& S = stars[star_of_point(stars, n_stars, j)];
        const int64_t jl = j - S.grid_off;
        if (jl >= S.n_wl) continue;  // padding between stars
        const int64_t k = (int64_t)l + (int64_t)L * jl;
        if (alpha_f32) ((float*)S.alpha)[k] = 0.0f;
        else ((double*)S.alpha)[k] = 0.0;
    }
}

// Each star's alpha set to its continuum on the grid (KORG_CNTM_GRID): per layer, the linear
// interpolation of the knots' values at every grid point -- the alpha that synthesize starts the
// lines from (src/synthesize.jl:225-226, 248-250), with the interpolation the window search uses
// (linear between the bracketing knots; outside the knots, the end intervals extended).  A
// wavefront per 64 points (no chunk straddles two stars: the grids are padded to 512 points):
// lane p finds point p's knot interval
// and weight, then the chunk's alpha -- one contiguous run of 64 L elements -- is written with
// consecutive lanes on consecutive elements (coalesced stores), each element taking its point's
// interval from the lane that found it.
__global__ __launch_bounds__(256) void k_cntm_fill(int n_stars, int L, int64_t n_total,
                                                   const StarDev* __restrict__ stars, int alpha_f32) {
    const int lane = threadIdx.x % WAVE;
    const int64_t j0 = ((int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) * WAVE;
    if (j0 >= n_total) return;
    const StarDev& S = stars[star_of_point(stars, n_stars, j0)];
    const int64_t jl0 = j0 - S.grid_off;
    const int np = (int)(min(jl0 + WAVE, S.n_wl) - jl0);
    if (np <= 0) return;  // padding between two stars' grids
    // lane p: point jl0 + p's knot interval (last knot <= lambda, within [0, n_knots - 2]) and
    // its interpolation weight (x - k_i) / (k_(i+1) - k_i)
    int kidx;
    double tw;
    {
        const double x = S.wls[jl0 + min(lane, np - 1)];
        int64_t lo = 0, hi = S.n_knots;
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) / 2;
            if (S.knots[mid] <= x) lo = mid; else hi = mid;
        }
        kidx = (int)(lo > S.n_knots - 2 ? S.n_knots - 2 : lo);
        tw = (x - S.knots[kidx]) / (S.knots[kidx + 1] - S.knots[kidx]);
    }
    const int64_t n_el = (int64_t)np * L;
    const int64_t base = (int64_t)L * jl0;  // the chunk's first alpha element
    int p = lane / L, l = lane % L;          // element e = p L + l of the chunk
    // every lane runs every step (the shuffles read from all lanes); past the chunk's elements
    // a lane stands in for its last point and stores nothing
    for (int64_t e0 = 0; e0 < n_el; e0 += WAVE) {
        const int64_t e = e0 + lane;
        const int src = min(p, np - 1);
        const int ki = __shfl(kidx, src);
        const double t = __shfl(tw, src);
        if (e < n_el) {
            const double* v = S.cntm + (int64_t)l * S.n_knots + ki;
            const double val = (1.0 - t) * v[0] + t * v[1];
            if (alpha_f32) ((float*)S.alpha)[base + e] = (float)val;
            else ((double*)S.alpha)[base + e] = val;
        }
        l += WAVE;
        while (l >= L) {
            l -= L;
            ++p;
        }
    }
}

struct RtArgs {
    int n_stars, L;
    int64_t n_total;  // global grid points
    const StarDev* stars;
    const RtTab* tab;  // [star * L + l]
    int alpha_f32;
    double* chi2_part;  // [global point]; G (Float64) goes to each star's StarDev::G
    // flux only: alpha holds the total absorption (continuum included); the flux of alpha and of
    // the continuum alone go to StarDev::flux and ::flux_c (either may be null); no chi^2, no G
    int flux_only;
};

__global__ __launch_bounds__(256) void k_rt_fit(RtArgs A) {
    const int L = A.L;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n_total) return;
    // stars' grids are padded to 512 points: the 64 points of a wavefront share one star
    const KCONST StarDev* stars = uniform_ptr(A.stars);
    const int st = star_of_point(stars, A.n_stars, j);
    const KCONST StarDev& S = stars[st];
    const int64_t jl = j - S.grid_off;
    if (jl >= S.n_wl) {
        if (!A.flux_only) A.chi2_part[j] = 0.0;
        return;
    }
    const RtTab* tb = A.tab + (int64_t)st * L;
    const double lam = S.wls[jl];
    // knot interval of lam: clamp(searchsortedlast(knots, lam), 1, n_knots - 1) - 1
    int64_t kidx = 0;
    {
        int64_t lo = 0, hi = S.n_knots;  // last k with knots[k] <= lam
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) / 2;
            if (S.knots[mid] <= lam) lo = mid; else hi = mid;
        }
        kidx = lo > S.n_knots - 2 ? S.n_knots - 2 : lo;
    }
    // the continuum's knot weight and Planck's factors, once per point (FP64 divisions are long
    // sequences): B_lambda(T) = c1 / expm1(c2 / T) with c1 = 2 h c^2 / lam^5, c2 = h c / (k lam)
    const double ckt = (lam - S.knots[kidx]) / (S.knots[kidx + 1] - S.knots[kidx]);
    auto cntm_pt = [&](int l) {
        const double* v = S.cntm + (int64_t)l * S.n_knots + kidx;
        return (1.0 - ckt) * v[0] + ckt * v[1];
    };
    const double il = 1.0 / lam, il2 = il * il;
    const double bb1 = (2 * hplanck_cgs * c_cgs * c_cgs) * (il2 * il2 * il);
    const double bb2 = (hplanck_cgs * c_cgs / kboltz_cgs) * il;
    auto planck = [&](int l) { return bb1 / expm1(bb2 / S.temps[l]); };
    auto alpha_lines = [&](int l) {
        const int64_t k = (int64_t)l + (int64_t)L * jl;
        return A.alpha_f32 ? (double)((const float*)S.alpha)[k] : ((const double*)S.alpha)[k];
    };
    // the total absorption of layer l, given its continuum ac
    auto alpha_total = [&](int l, double ac) { return A.flux_only ? alpha_lines(l) : ac + alpha_lines(l); };
    // forward: flux of the total and of the continuum alone
    // (F and Fc are 6 times the sums of the intervals' integrals until the end)
    double F = 0.0, Fc = 0.0;
    double tau0 = 0.0, tc0 = 0.0;
    double ac0 = cntm_pt(0);
    double u0 = alpha_total(0, ac0) * tb[0].f, uc0 = ac0 * tb[0].f;
    double S0 = planck(0);
    double E0 = 1.0, x0 = 1.0, Ec0 = 1.0, xc0 = 1.0;  // E2 and e^{-tau} at tau = 0
    for (int i = 1; i < L; ++i) {
        const double ac = cntm_pt(i);
        const double u1 = alpha_total(i, ac) * tb[i].f, uc1 = ac * tb[i].f;
        const double dl = tb[i].ltau - tb[i - 1].ltau;
        const double tau1 = tau0 + 0.5 * (u1 + u0) * dl, tc1 = tc0 + 0.5 * (uc1 + uc0) * dl;
        const double S1 = planck(i);
        const double E1 = expint2_korg(tau1, nullptr), x1 = exp(-tau1);
        const double Ec1 = expint2_korg(tc1, nullptr), xc1 = exp(-tc1);
        {
            const double m = (S1 - S0) / (tau1 - tau0), b = S0 - m * tau0;
            F += (tau1 * E1 * (3 * b + 2 * m * tau1) - x1 * (3 * b + 2 * m * (tau1 + 1))) -
                 (tau0 * E0 * (3 * b + 2 * m * tau0) - x0 * (3 * b + 2 * m * (tau0 + 1)));
        }
        {
            const double m = (S1 - S0) / (tc1 - tc0), b = S0 - m * tc0;
            Fc += (tc1 * Ec1 * (3 * b + 2 * m * tc1) - xc1 * (3 * b + 2 * m * (tc1 + 1))) -
                  (tc0 * Ec0 * (3 * b + 2 * m * tc0) - xc0 * (3 * b + 2 * m * (tc0 + 1)));
        }
        tau0 = tau1; tc0 = tc1; u0 = u1; uc0 = uc1; S0 = S1;
        E0 = E1; x0 = x1; Ec0 = Ec1; xc0 = xc1;
    }
    F *= 2 * PI / 6;
    Fc *= 2 * PI / 6;
    if (A.flux_only) {
        // in erg/s/cm^4/A, as synthesize returns them (src/synthesize.jl:314-318)
        if (S.flux) S.flux[jl] = F * 1e-8;
        if (S.flux_c) S.flux_c[jl] = Fc * 1e-8;
        return;
    }
    const double Fn = F / Fc;
    const double sg = S.sigma[jl];
    const double r = (Fn - S.data[jl]) / sg;
    A.chi2_part[j] = r * r;
    if (S.flux) S.flux[jl] = Fn;
    // reverse: G[l] = dchi^2/dalpha_total[l] = dchi^2/dalpha_lines[l]
    // 2 pi dchi^2/dF, over 6: the partials below are 6 times those of the integrals
    const double w = (2 * PI / 6) * (2 * r / sg / Fc);
    double* G = (double*)S.G + (int64_t)L * jl;
    // walk the intervals from the deepest up, recomputing tau_i from tau_{i+1}
    double tau1 = tau0, u1 = u0, S1 = S0;
    double dE1;
    double E1 = expint2_korg(tau1, &dE1), x1 = exp(-tau1);
    double pending = 0.0;  // contribution to gtau[i + 1] from interval i + 1 (already walked)
    double sfx = 0.0;      // sum of gtau[k] for k >= i + 2
    for (int i = L - 2; i >= 0; --i) {
        const double ac = cntm_pt(i);
        const double u_i = (ac + alpha_lines(i)) * tb[i].f;
        const double dl = tb[i + 1].ltau - tb[i].ltau;
        const double tau_i = i == 0 ? 0.0 : tau1 - 0.5 * (u1 + u_i) * dl;
        const double S_i = planck(i);
        double dE_i;
        const double E_i = expint2_korg(tau_i, &dE_i), x_i = exp(-tau_i);
        if (i == 0) dE_i = 0.0;  // tau_0 = 0 is fixed: its derivative is never used
        const double idt = 1.0 / (tau1 - tau_i);
        const double m = (S1 - S_i) * idt, b = S_i - m * tau_i;
        // partials of 6 core(tau, m, b) = tau E (3b + 2m tau) - e^{-tau} (3b + 2m (tau + 1))
        auto partials = [&](double t, double E, double dE, double x, double& ct, double& cm, double& cb) {
            const double Aq = 3 * b + 2 * m * t, Bq = 3 * b + 2 * m * (t + 1);
            ct = (E + t * dE) * Aq + t * E * 2 * m + x * Bq - x * 2 * m;
            cm = t * E * 2 * t - x * 2 * (t + 1);
            cb = t * E * 3 - x * 3;
        };
        double ct1, cm1, cb1, ct0, cm0, cb0;
        partials(tau1, E1, dE1, x1, ct1, cm1, cb1);
        partials(tau_i, E_i, dE_i, x_i, ct0, cm0, cb0);
        const double Tm = cm1 - cm0, Tb = cb1 - cb0;
        const double dm_dt1 = -m * idt, dm_dt0 = m * idt;
        const double db_dt1 = -tau_i * dm_dt1, db_dt0 = -m - tau_i * dm_dt0;
        const double g_up = w * (ct1 + Tm * dm_dt1 + Tb * db_dt1) + pending;  // gtau[i + 1], final
        pending = w * (-ct0 + Tm * dm_dt0 + Tb * db_dt0);
        // dtau_k/du_j = dl_j/2 [1 <= j <= k] + dl_{j+1}/2 [j + 1 <= k]; layer i + 1 is final now
        const double sfx_next = sfx;       // sum over k >= i + 2
        sfx += g_up;                       // sum over k >= i + 1
        double gu = 0.5 * dl * sfx;        // dl = ltau[i + 1] - ltau[i]
        if (i + 2 < L) gu += 0.5 * (tb[i + 2].ltau - tb[i + 1].ltau) * sfx_next;
        G[i + 1] = gu * tb[i + 1].f;
        tau1 = tau_i; u1 = u_i; S1 = S_i; E1 = E_i; dE1 = dE_i; x1 = x_i;
    }
    // layer 0: tau_0 is fixed, u_0 enters tau_k (k >= 1) through the first interval
    G[0] = 0.5 * (tb[1].ltau - tb[0].ltau) * sfx * tb[0].f;
}

// chi^2 of every star: one block per star, its points summed in a fixed order and tree.
__global__ __launch_bounds__(256) void k_star_chi2(const StarDev* __restrict__ stars,
                                                   const double* __restrict__ chi2_part) {
    __shared__ double red[256];
    const StarDev& S = stars[blockIdx.x];
    const int t = threadIdx.x;
    double sum = 0.0;
    for (int64_t j = t; j < S.n_wl; j += 256) sum += chi2_part[S.grid_off + j];
    red[t] = sum;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0 && S.chi2) *S.chi2 = red[0];
}

// ---------------------------------------------------------------------------------------------
// Adjoint of the Float32 form (the fit step at Float32 precision): the far zone of every line
// through moments of the series, the near zone exactly (k_vjp over near-zone items).
//
// In a line's far zone the forward adds phi = scaling H with the series of the header,
// H = (a/sqrt(pi)) sum_k e_k(a) v^-2k, v^-2 = q R (R = 1/delta^2, q = 1/(DELTA_UNIT inv)^2).  The
// sums the chain rule takes over those points therefore follow from four moments per (line, layer),
//   M_k = sum_{j in far zone} G_j R_j^k,  k = 1..4:
//   sum G H = (a/sqrt(pi)) sum_k e_k q^k M_k,   sum G dH/da = (1/sqrt(pi)) sum_k (a e_k)' q^k M_k,
//   sum G v dH/dv = (a/sqrt(pi)) sum_k (-2k) e_k q^k M_k.
// The moments of a group of lines that share points are a matrix product: for 16 layers x
// (4 lines x 4 powers), G[layer, point] P[point, (line, k)] on the matrix cores
// (v_mfma_f32_16x16x4_f32, K = 4 points).  Lines are grouped by window width in classes of a
// factor 2, so the lines of a group cover nearly the same points.  G enters in Float32, scaled
// per (star, layer) by the power of two that brings its largest magnitude into [0.5, 1).  The R^3
// and R^4 moments cover the whole far zone (the forward drops those terms beyond the tail radius,
// where they are below SERIES_TAIL_EPS of the leading one).
// ---------------------------------------------------------------------------------------------
constexpr int ADJ_NC = 24;        // grouping classes: window width W in (2^(c-1), 2^c] points
constexpr int ADJ_GROUP = 4;      // lines per group: the 4 line slots of the 16 B columns
constexpr int GEXP_NONE = -2000;  // scale exponent of an all-zero layer of G

__device__ inline int adj_class(const LineMeta& m, int2 z) {
    if (far_zone_empty(m, z)) return ADJ_NC;
    const uint32_t W = (uint32_t)(m.ub - m.lb + 1);
    const int c = W <= 1 ? 0 : 32 - __clz(W - 1);
    return c < ADJ_NC ? c : ADJ_NC - 1;
}

// Class counts per block of lines (the partition's first pass).
__global__ __launch_bounds__(PART_BLOCK) void k_adj_count(int64_t n, const LineMeta* __restrict__ meta,
                                                          const int2* __restrict__ azone,
                                                          const int4* __restrict__ azq,
                                                          int* __restrict__ block_counts) {
    __shared__ int cnt[ADJ_NC];
    if (threadIdx.x < ADJ_NC) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
    if (i < n) {
        const int c = adj_class(meta[i], azone_in(azone[i], azq, i));
        if (c < ADJ_NC) atomicAdd(&cnt[c], 1);
    }
    __syncthreads();
    if (threadIdx.x < ADJ_NC)
        block_counts[(int64_t)blockIdx.x * ADJ_NC + threadIdx.x] = cnt[threadIdx.x];
}

// Stable scatter of the lines with a far zone by class: within a class, in line order.
// Also marks each group's leader: adj_lead[line] = the group's position in adj_line, else -1.
__global__ __launch_bounds__(PART_BLOCK) void k_adj_scatter(int64_t n, const LineMeta* __restrict__ meta,
                                                            const int2* __restrict__ azone,
                                                            const int4* __restrict__ azq,
                                                            const int* __restrict__ block_offsets,
                                                            const int* __restrict__ class_start,
                                                            int* __restrict__ adj_line,
                                                            int* __restrict__ adj_lead) {
    __shared__ int wave_cnt[PART_BLOCK / WAVE][ADJ_NC];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int64_t i = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
    const int c = i < n ? adj_class(meta[i], azone_in(azone[i], azq, i)) : ADJ_NC;
    int rank = 0;
    for (int k = 0; k < ADJ_NC; ++k) {
        const uint64_t b = ballot(c == k);
        if (c == k) rank = __popcll(b & ((1ull << lane) - 1));
        if (lane == 0) wave_cnt[w][k] = __popcll(b);
    }
    __syncthreads();
    int lead = -1;
    if (c < ADJ_NC) {
        int before = 0;
        for (int ww = 0; ww < w; ++ww) before += wave_cnt[ww][c];
        const int in_class = block_offsets[(int64_t)blockIdx.x * ADJ_NC + c] + before + rank;
        adj_line[class_start[c] + in_class] = (int)i;
        // the first line of each group of ADJ_GROUP within the class leads it
        if (in_class % ADJ_GROUP == 0) lead = class_start[c] + in_class;
    }
    if (i < n) adj_lead[i] = lead;
}

// The groups in line order (star, then wavelength) -- the order the moments kernel takes them in,
// so that the groups in flight together read nearby points of G: a stable compaction of the
// leaders (block counts, their scan by k_scan<1>, the scatter of the leaders' group positions).
__global__ __launch_bounds__(PART_BLOCK) void k_lead_count(int64_t n, const int* __restrict__ adj_lead,
                                                           int* __restrict__ block_counts) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
    const uint64_t b = ballot(i < n && adj_lead[i] >= 0);
    if ((threadIdx.x % WAVE) == 0 && b) atomicAdd(&cnt, __popcll(b));
    __syncthreads();
    if (threadIdx.x == 0) block_counts[blockIdx.x] = cnt;
}

__global__ __launch_bounds__(PART_BLOCK) void k_lead_scatter(int64_t n, const int* __restrict__ adj_lead,
                                                             const int* __restrict__ block_offsets,
                                                             int* __restrict__ groups) {
    __shared__ int wave_cnt[PART_BLOCK / WAVE];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int64_t i = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
    const int lead = i < n ? adj_lead[i] : -1;
    const uint64_t b = ballot(lead >= 0);
    if (lane == 0) wave_cnt[w] = __popcll(b);
    __syncthreads();
    if (lead >= 0) {
        int before = 0;
        for (int ww = 0; ww < w; ++ww) before += wave_cnt[ww];
        groups[block_offsets[blockIdx.x] + before + __popcll(b & ((1ull << lane) - 1))] = lead;
    }
}

// Per (star, layer): the binary exponent e of max_j |G[l, j]| (max |G| 2^-e in [0.5, 1)), by an
// order-free atomicMax over wavefronts.  A wavefront per 128 points (one star: the stars' grids are
// padded to 512 points); lanes = layers.
__global__ __launch_bounds__(256) void k_g_exp(int n_stars, int L, int64_t n_total,
                                               const StarDev* __restrict__ stars, int* __restrict__ gexp) {
    const int lane = threadIdx.x % WAVE;
    const int64_t j0 = ((int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) * STAR_ALIGN;
    if (j0 >= n_total) return;
    const int st = star_of_point(stars, n_stars, j0);
    const StarDev& S = stars[st];
    const int64_t jl0 = j0 - S.grid_off, jl1 = min(jl0 + STAR_ALIGN, S.n_wl);
    const double* __restrict__ G = (const double*)S.G;
    for (int l = lane; l < L; l += WAVE) {
        double mx = 0.0;
        for (int64_t jl = jl0; jl < jl1; ++jl) mx = fmax(mx, fabs(G[l + (int64_t)L * jl]));
        if (mx > 0.0) {
            int e;
            frexp(mx, &e);
            atomicMax(&gexp[(int64_t)st * L + l], e);
        }
    }
}

// K steps per block of the moments kernel (loads in flight together), per Float32 partial sum
// (a multiple of the block), and the points the kernel may read past the last grid point; the
// parts of its item list (one per XCD).
constexpr int MOM_PARTS = 8;
constexpr int MOM_BLOCK = 4;
constexpr int MOM_CHUNK = 16;
constexpr int MOM_PAD = 4 * MOM_BLOCK;
static_assert(4 * MOM_CHUNK == FF_SUB, "a chunk of the moments is one far-field sub-tile");

// Position of layer l (of a 64-layer block) in the moments' A operands: 4 (l % 16) + l / 16, so
// that the float4 of lane column c holds layers c, c + 16, c + 32, c + 48 and the moments' MFMA q
// (the float4's component q) covers the contiguous layers 16q .. 16q + 15.
__device__ inline int mom_pos(int l) { return 4 * (l & 15) + (l >> 4); }

// G in Float32, scaled by 2^-e per (star, layer), as [layer block][global point][64 positions]
// (mom_pos) with MOM_PAD points after the last (layers past L and points outside every star
// zero): the moments' A operand, 1 KB per 4 points.  Also the wavelengths in units of DELTA_UNIT on the same global
// points (zero outside the stars).
__global__ __launch_bounds__(256) void k_g_pack(int n_stars, int L, int n_lblocks, int64_t n_total,
                                                const StarDev* __restrict__ stars,
                                                const int* __restrict__ gexp, float* __restrict__ g32,
                                                double* __restrict__ lam_s) {
    const int lane = threadIdx.x % WAVE;
    const int64_t j0 = ((int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) * STAR_ALIGN;
    if (j0 >= n_total) return;
    const int st = star_of_point(stars, n_stars, j0);
    const StarDev& S = stars[st];
    const int64_t go = S.grid_off;
    const int64_t jl0 = j0 - go;
    // the last chunk also fills the padding after the grid
    const int64_t np = STAR_ALIGN + (j0 + STAR_ALIGN >= n_total ? MOM_PAD : 0);
    const double* __restrict__ G = (const double*)S.G;
    const int64_t stride = (n_total + MOM_PAD) * WAVE;
    for (int b = 0; b < n_lblocks; ++b) {
        const int l = b * WAVE + lane;
        const int e = l < L ? gexp[(int64_t)st * L + l] : 0;
        float* __restrict__ dst = g32 + (int64_t)b * stride + mom_pos(lane);
        for (int64_t p = 0; p < np; ++p) {
            const int64_t jl = jl0 + p;
            dst[(j0 + p) * WAVE] = l < L && jl < S.n_wl ? (float)ldexp(G[l + (int64_t)L * jl], -e) : 0.0f;
        }
    }
    for (int64_t p = lane; p < np; p += WAVE) {
        const int64_t jl = jl0 + p;
        lam_s[j0 + p] = jl < S.n_wl ? S.wls[jl] * (1.0 / DELTA_UNIT) : 0.0;
    }
}

// The far field's moments of G: per sub-tile g and layer, g_m = sum over the sub-tile's points of
// G[layer, j] 2^-e tau_j^m (tau_j = (lambda_j - c) / h, m < FF_TERMS, e the layer's scale exponent
// of k_g_exp), summed in Float64 and stored in Float32 as [layer block][sub-tile][m][64 layers]:
// the A operand of the moments' far-field steps (over a sub-tile in a line's far field,
// sum_j G_j R_j^k = sum_m W_km g_m).  One wavefront per (sub-tile, layer block), lanes = layers;
// sub-tiles with no points (padding) get zeros.
__global__ __launch_bounds__(WAVE) void k_g_ffmom(int n_stars, int L, int n_lblocks, int64_t n_sub,
                                                  const StarDev* __restrict__ stars,
                                                  const FfGeo* __restrict__ geo,
                                                  const int* __restrict__ gexp,
                                                  float* __restrict__ gff) {
    const int lane = threadIdx.x;
    const int64_t item = blockIdx.x;
    const int64_t g = item / n_lblocks;
    const int lblk = (int)(item % n_lblocks);
    if (g >= n_sub) return;
    const FfGeo G = geo[g];
    const int l = lblk * WAVE + lane;
    double acc[FF_TERMS];
#pragma unroll
    for (int m = 0; m < FF_TERMS; ++m) acc[m] = 0.0;
    int e = 0;
    if (G.first <= G.last) {
        const int st = star_of_point(stars, n_stars, G.first);
        const StarDev& S = stars[st];
        const double* __restrict__ Gs = (const double*)S.G;
        e = l < L ? gexp[(int64_t)st * L + l] : 0;
        const double h = (double)G.h_u;
        for (int64_t j = G.first; j <= G.last; ++j) {
            const int64_t jl = j - S.grid_off;
            const double tau = h > 0.0 ? (S.wls[jl] * (1.0 / DELTA_UNIT) - G.c_u) / h : 0.0;
            double p = l < L ? Gs[l + (int64_t)L * jl] : 0.0;
#pragma unroll
            for (int m = 0; m < FF_TERMS; ++m) {
                acc[m] += p;
                p *= tau;
            }
        }
    }
    float* __restrict__ dst = gff + ((int64_t)lblk * n_sub + g) * FF_TERMS * WAVE + mom_pos(lane);
#pragma unroll
    for (int m = 0; m < FF_TERMS; ++m) dst[m * WAVE] = l < L ? (float)ldexp(acc[m], -e) : 0.0f;
}

// The whole super-tiles' moments of G from their sub-tiles' (k_g_ffmom, same layout and scale):
// tau_S = a tau_s + b with a = h_s / h_S and b = (c_s - c_S) / h_S, so the super-tile's moment m is
// sum_s sum_i binom(m, i) b^(m - i) a^i g_s,i.  One wavefront per (super-tile, layer block), lanes
// = the stored positions; super-tiles that are not whole get zeros.
__global__ __launch_bounds__(WAVE) void k_g_ffmom_super(int n_lblocks, int64_t n_sub, int64_t n_super,
                                                        const FfGeo* __restrict__ geo,
                                                        const FfGeo* __restrict__ sgeo,
                                                        const float* __restrict__ gff,
                                                        float* __restrict__ gffs) {
    const int lane = threadIdx.x;
    const int64_t item = blockIdx.x;
    const int64_t sp = item / n_lblocks;
    const int lblk = (int)(item % n_lblocks);
    if (sp >= n_super) return;
    const FfGeo GS = sgeo[sp];
    double acc[FF_TERMS];
#pragma unroll
    for (int m = 0; m < FF_TERMS; ++m) acc[m] = 0.0;
    if (GS.first <= GS.last) {
        const double ih = 1.0 / (double)GS.h_u;
        for (int k = 0; k < FF_NS; ++k) {
            const int64_t g = sp * FF_NS + k;
            const double a = (double)geo[g].h_u * ih, b = (geo[g].c_u - GS.c_u) * ih;
            const float* __restrict__ src = gff + ((int64_t)lblk * n_sub + g) * FF_TERMS * WAVE + lane;
            double c[FF_TERMS];
            double ai = 1.0;
#pragma unroll
            for (int i = 0; i < FF_TERMS; ++i) {
                c[i] = ai * (double)src[i * WAVE];
                ai *= a;
            }
            // v[i] = binom(m, i) b^(m - i), advanced m by m (v_(m+1)[i] = b v_m[i] + v_m[i - 1])
            double v[FF_TERMS];
#pragma unroll
            for (int i = 0; i < FF_TERMS; ++i) v[i] = i == 0 ? 1.0 : 0.0;
#pragma unroll
            for (int m = 0; m < FF_TERMS; ++m) {
                if (m > 0) {
#pragma unroll
                    for (int i = m; i > 0; --i) v[i] = b * v[i] + v[i - 1];
                    v[0] *= b;
                }
                double t = 0.0;
#pragma unroll
                for (int i = 0; i <= m; ++i) t += v[i] * c[i];
                acc[m] += t;
            }
        }
    }
    float* __restrict__ dst = gffs + ((int64_t)lblk * n_super + sp) * FF_TERMS * WAVE + lane;
#pragma unroll
    for (int m = 0; m < FF_TERMS; ++m) dst[m * WAVE] = (float)acc[m];
}

struct MomArgs {
    int n_stars, L, n_lblocks;
    int64_t n_total;           // global grid points
    const StarDev* stars;
    const LineMeta* meta;
    const int2* azone;         // the adjoint's near zones
    const int4* azq;           // each 16-layer block's (null: the line's)
    const int* adj_line;       // the lines with a far zone, by class (ascending width)
    const int* adj_start;      // [ADJ_NC + 1]; adj_start[ADJ_NC] = the number of such lines
    const int* groups;         // the groups' positions in adj_line, in line order
    const int* n_groups;       // their number (device)
    const float* g32;          // k_g_pack's layouts
    const double* lam_s;
    float4* mom;               // [(line * n_lblocks + b) * 64 + l % 64]: (M1, M2, M3, M4) / 2^e
    unsigned long long* counters;
    // far field (null ffrun: none): each line's runs clear of its adjoint near zone, the sub-tile
    // geometry and the sub-tiles' moments of G (k_g_ffmom)
    const int4* ffrun;
    const FfGeo* geo;
    const float* gff;
    int64_t n_sub;
    // the same over whole super-tiles (FF_NS sub-tiles), for lines at least 8 super-tile
    // half-widths away
    const FfGeo* sgeo;
    const float* gffs;
    int64_t n_super;
};

typedef float floatx4 __attribute__((ext_vector_type(4)));

// One wavefront per (group of ADJ_GROUP lines, 64-layer block), widest classes first.  Lane i of a
// K step: A = scaled G of layers 16 m + i % 16 (m = 0..3, one MFMA each: the operands are packed
// in mom_pos order) at point j0 + i / 16; B = R^(p+1) of line slot (i % 16) / 4 at that point,
// p = i % 4, zero outside the line's far zone: its window less its adjoint near zone -- for MFMA m
// the near zone of the line's 16-layer block m (azq), inside which the block's layers are
// evaluated exactly, so that a block whose own radius is narrower than the line's takes the series
// over more of the window.  The lines of a group are handled star by star (a group can straddle
// two stars), each pass over the union of its lines' windows, then (far field) over the union of
// their far-field runs.  Four waves per SIMD (128 VGPRs, a few spilled: faster than three).
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(4))) void k_vjp_moments(MomArgs A) {
    const int lane = threadIdx.x;
    if (A.counters[3] != 0) return;  // failed setup, or a capacity shortfall
    const int64_t n_items = (int64_t)A.n_groups[0] * A.n_lblocks;
    const int col = lane & 15, kk = lane >> 4;
    const int slot = col >> 2, pw = col & 3;
    // far-field steps: binom(-2k, m) of the lane's power k = pw + 1 and terms m = 4t + kk
    float bc[FF_TERMS / 4];
#pragma unroll
    for (int t = 0; t < FF_TERMS / 4; ++t) {
        double b = 1.0;
        for (int i = 1; i <= 4 * t + kk; ++i) b *= -(double)(2 * (pw + 1) + i - 1) / (double)i;
        bc[t] = (float)b;
    }
    const KCONST FfGeo* geo = uniform_ptr(A.geo);
    // The item list in MOM_PARTS contiguous parts, part x first taken by the waves of
    // workgroups x, x + MOM_PARTS, ... (workgroups go to the XCDs round robin, so the waves of
    // one XCD share one L2 and work on nearby points); a wave whose part is used up moves on to
    // the next parts.
    int part = blockIdx.x % MOM_PARTS, parts_left = MOM_PARTS;
    for (;;) {
        const int64_t p_lo = n_items * part / MOM_PARTS, p_hi = n_items * (part + 1) / MOM_PARTS;
        const int64_t item = p_lo + wave_take(&A.counters[16 + part], lane);
        if (item >= p_hi) {
            if (--parts_left == 0) break;
            part = (part + 1) % MOM_PARTS;
            continue;
        }
        const int64_t g = item / A.n_lblocks;
        const int lblk = (int)(item % A.n_lblocks);
        const int64_t p0 = A.groups[g];
        // the group's lines: up to ADJ_GROUP from p0, within the leader's class
        const int l0 = A.adj_line[p0];
        const int64_t cend = A.adj_start[adj_class(A.meta[l0], azone_in(A.azone[l0], A.azq, l0)) + 1];
        const int nl = (int)min((int64_t)ADJ_GROUP, cend - p0);
        int line = -1, st = -1;
        LineMeta m = {0.0, 1, 0, 1, 0, 0, ADJ_NC, 1, 0};
        int2 z = make_int2(1, 0);
        int4 zq = make_int4(0, 0, 0, 0);  // the line's 16-layer blocks' zones (azq)
        int4 fr = make_int4(1, 0, 1, 0);  // the line's far-field runs (empty: none)
        if (slot < nl) {
            line = A.adj_line[p0 + slot];
            m = A.meta[line];
            z = A.azone[line];
            if (A.azq) zq = A.azq[line];
            st = star_of_line(A.stars, A.n_stars, line);
            if (A.ffrun) fr = A.ffrun[line];
        }
        uint32_t todo = (1u << nl) - 1;  // line slots not yet handled
        while (todo) {
            const int sc = readlane_i32(st, 4 * (__ffs(todo) - 1));  // lane 4k holds slot k
            uint32_t pass = 0;
            int u0 = INT_MAX, u1 = INT_MIN;
            for (int k = 0; k < nl; ++k) {
                if (readlane_i32(st, 4 * k) != sc || !((todo >> k) & 1)) continue;
                pass |= 1u << k;
                u0 = min(u0, readlane_i32(m.lb, 4 * k));
                u1 = max(u1, readlane_i32(m.ub, 4 * k));
            }
            todo &= ~pass;
            const bool act = slot < nl && ((pass >> slot) & 1);
            const double lam0s = m.lam0 * (1.0 / DELTA_UNIT);
            // the lane's far zone [lb, clo) and (chi, ub] as bounds on the step's first point j0
            // (the lane's point is j0 + kk); not in this pass: no points
            const int lbr = (act ? m.lb : 1) - kk, ubr = (act ? m.ub : 0) - kk;
            // the innermost of the line's block zones (outside it some block takes the series), and
            // each block's zone: MFMA q's rows are the block's 16 layers, whose B is zero inside it
            int2 zb[4];
            zb[0] = azone_blk(z, zq.x);
            zb[1] = azone_blk(z, zq.y);
            zb[2] = azone_blk(z, zq.z);
            zb[3] = azone_blk(z, zq.w);
            int ilo = z.x, ihi = z.y;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                ilo = max(ilo, zb[q].x);
                ihi = min(ihi, zb[q].y);
            }
            const int clor = ilo - kk, chir = ihi - kk;
            int qlo[4], qhi[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                qlo[q] = zb[q].x - kk;
                qhi[q] = zb[q].y - kk;
            }
            const bool blocks = A.azq != nullptr;
            // Float32 sums over chunks of MOM_CHUNK K steps (one 64-point sub-tile), added into
            // Float64 totals: a far zone spans thousands of points and its moments enter
            // differences (core against wing) in the gradients of the damping.  Chunks end at fixed
            // points, so a line's sums do not depend on its group either.
            floatx4 acc[4];
            double tot[4][4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                acc[q] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int r = 0; r < 4; ++r) tot[q][r] = 0.0;
            }
            // the pass's lines' far-field runs, wavefront-uniform (lane 4k holds slot k; lines not
            // in the pass have none here)
            int rx[ADJ_GROUP], ry[ADJ_GROUP], rz[ADJ_GROUP], rw[ADJ_GROUP];
#pragma unroll
            for (int k = 0; k < ADJ_GROUP; ++k) {
                const bool in = k < nl && ((pass >> k) & 1);
                rx[k] = in ? readlane_i32(fr.x, 4 * k) : 1;
                ry[k] = in ? readlane_i32(fr.y, 4 * k) : 0;
                rz[k] = in ? readlane_i32(fr.z, 4 * k) : 1;
                rw[k] = in ? readlane_i32(fr.w, 4 * k) : 0;
            }
            const int4 lfr = act ? fr : make_int4(1, 0, 1, 0);
            // the pass's lines' windows and adjoint near zones, wavefront-uniform
            int wlo[ADJ_GROUP], whi[ADJ_GROUP], nlo[ADJ_GROUP], nhi[ADJ_GROUP];
#pragma unroll
            for (int k = 0; k < ADJ_GROUP; ++k) {
                const bool in = k < nl && ((pass >> k) & 1);
                wlo[k] = in ? readlane_i32(m.lb, 4 * k) : 1;
                whi[k] = in ? readlane_i32(m.ub, 4 * k) : 0;
                nlo[k] = readlane_i32(ilo, 4 * k);
                nhi[k] = readlane_i32(ihi, 4 * k);
            }
            // K steps start at multiples of 4 (a star's grid starts at one): every point falls in
            // the same step whatever lines share the group, and steps outside a line's far zone add
            // exact zeros, so a line's moments do not depend on its group.  Steps go in blocks of
            // MOM_BLOCK, the block's loads issued before its MFMAs, and blocks in chunks of one
            // 64-point sub-tile, whose Float32 sums are added into the Float64 totals at its end
            // (a block reads at most 4 MOM_BLOCK - 1 points past the union: B = 0 there).  Points in
            // a sub-tile of a line's far-field runs are left to the far-field steps below.  Steps
            // that would add only zeros are skipped: sub-tiles where no line of the pass has a point
            // of its far zone outside its far-field runs (the union of a pass's windows is mostly
            // such sub-tiles: far-field runs and near zones), and blocks with no such point for any
            // lane; the sums are bitwise those of taking every step.
            const int js = u0 & ~(4 * MOM_BLOCK - 1);
            for (int gs = js / FF_SUB; gs <= u1 / FF_SUB; ++gs) {
                const int s0 = gs * FF_SUB, s1 = s0 + FF_SUB - 1;
                bool need = false;
#pragma unroll
                for (int k = 0; k < ADJ_GROUP; ++k)
                    need = need || (wlo[k] <= s1 && whi[k] >= s0 && !(s0 >= nlo[k] && s1 <= nhi[k]) &&
                                    !(A.ffrun && ((gs >= rx[k] && gs <= ry[k]) ||
                                                  (gs >= rz[k] && gs <= rw[k]))));
                if (!need) continue;  // wavefront-uniform
                const bool lane_ff = ff_in(lfr, gs);
                const int jhi = min(gs * FF_SUB + FF_SUB - 4 * MOM_BLOCK, u1);
                for (int jb = max(gs * FF_SUB, js); jb <= jhi; jb += 4 * MOM_BLOCK) {
                    bool farv[MOM_BLOCK];
                    bool anyf = false;
#pragma unroll
                    for (int u = 0; u < MOM_BLOCK; ++u) {
                        const int j0 = jb + 4 * u;
                        farv[u] = j0 >= lbr && j0 <= ubr && (j0 < clor || j0 > chir) && !lane_ff;
                        anyf = anyf || farv[u];
                    }
                    if (ballot(anyf) == 0) continue;  // wavefront-uniform
                    const float4* __restrict__ gp =
                        (const float4*)(A.g32 + ((int64_t)lblk * (A.n_total + MOM_PAD) + jb) * WAVE);
                    const double* __restrict__ lp = A.lam_s + jb;
                    // A: layers 4 col .. 4 col + 3 at point j0 + kk; the wavelength of that point
                    float4 a[MOM_BLOCK];
                    double lam[MOM_BLOCK];
#pragma unroll
                    for (int u = 0; u < MOM_BLOCK; ++u) {
                        a[u] = gp[u * WAVE + lane];
                        lam[u] = lp[4 * u + kk];
                    }
#pragma unroll
                    for (int u = 0; u < MOM_BLOCK; ++u) {
                        const float d = (float)(lam[u] - lam0s);
                        float R = __builtin_amdgcn_rcpf(d * d);
                        R = farv[u] ? R : 0.0f;
                        const float R2 = R * R;
                        const float b = ((pw & 1) ? R2 : R) * ((pw & 2) ? R2 : 1.0f);
                        float bq[4] = {b, b, b, b};
                        if (blocks) {  // wavefront-uniform
                            const int j0 = jb + 4 * u;
#pragma unroll
                            for (int q = 0; q < 4; ++q) bq[q] = (j0 < qlo[q] || j0 > qhi[q]) ? b : 0.0f;
                        }
                        acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].x, bq[0], acc[0], 0, 0, 0);
                        acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].y, bq[1], acc[1], 0, 0, 0);
                        acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].z, bq[2], acc[2], 0, 0, 0);
                        acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].w, bq[3], acc[3], 0, 0, 0);
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) tot[q][r] += (double)acc[q][r];
                    acc[q] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
                }
            }
            if (A.ffrun) {
                // The far-field steps in ascending order (after every direct chunk: a line's
                // Float64 totals add the same values in the same order in any group): over a span
                // of points (a sub-tile, or a whole super-tile of FF_NS) in the line's runs,
                // sum_j G_j R_j^k = sum_m W_km g_m with W_km = binom(-2k, m) u^m R0^k, R0 = 1/d^2,
                // u = h/d, d = c - lam0; K = 4 terms per MFMA, A = the span's moments of G
                // (k_g_ffmom), three MFMAs per layer quarter.  A line whose runs cover a whole
                // super-tile at least FF_RATIO of its half-widths from the centre takes it in one
                // step (the same truncation bound as a sub-tile's) and adds zeros in the sub-tile
                // steps there; the others take the sub-tile steps.
                int glo = INT_MAX, ghi = INT_MIN;
#pragma unroll
                for (int k = 0; k < ADJ_GROUP; ++k) {
                    if (rx[k] <= ry[k]) { glo = min(glo, rx[k]); ghi = max(ghi, ry[k]); }
                    if (rz[k] <= rw[k]) { glo = min(glo, rz[k]); ghi = max(ghi, rw[k]); }
                }
                // one step over a span with centre cu, half-width hu and moments gq (zeros for lanes
                // that are not on), flushed into the totals; acc is zero on entry (every direct
                // chunk ends by clearing it)
                auto ff_step = [&](double cu, float hu, bool on, const float4* __restrict__ gq) {
                    float x = __builtin_amdgcn_rcpf((float)(cu - lam0s));
                    x = on ? x : 0.0f;
                    const float R0 = x * x, R2 = R0 * R0;
                    const float Rk = ((pw & 1) ? R2 : R0) * ((pw & 2) ? R2 : 1.0f);
                    const float u = hu * x, u2 = u * u, u4 = u2 * u2;
                    float w = Rk * ((kk & 1) ? u : 1.0f) * ((kk & 2) ? u2 : 1.0f);
#pragma unroll
                    for (int t = 0; t < FF_TERMS / 4; ++t) {
                        const float4 a4 = gq[t * WAVE + lane];  // terms 4t + kk, layers 4 col ..
                        const float bw = bc[t] * w;
                        acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.x, bw, acc[0], 0, 0, 0);
                        acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.y, bw, acc[1], 0, 0, 0);
                        acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.z, bw, acc[2], 0, 0, 0);
                        acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.w, bw, acc[3], 0, 0, 0);
                        w *= u4;
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) tot[q][r] += (double)acc[q][r];
                        acc[q] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
                    }
                };
                const KCONST FfGeo* sgeo = uniform_ptr(A.sgeo);
                for (int sp = glo / FF_NS; sp <= ghi / FF_NS; ++sp) {
                    const int s0 = sp * FF_NS, s1 = s0 + FF_NS - 1;
                    bool le = false;
                    const double scu = sgeo[sp].c_u;
                    const float shu = sgeo[sp].h_u;
                    if (sgeo[sp].first <= sgeo[sp].last)  // wavefront-uniform: a whole super-tile
                        le = ((s0 >= lfr.x && s1 <= lfr.y) || (s0 >= lfr.z && s1 <= lfr.w)) &&
                             fabs(scu - lam0s) >= FF_RATIO * (double)shu;
                    const uint64_t sb = ballot(le);  // lane 4k holds slot k
                    if (sb)
                        ff_step(scu, shu, le,
                                (const float4*)(A.gffs + ((int64_t)lblk * A.n_super + sp) * FF_TERMS * WAVE));
                    for (int gs = max(s0, glo); gs <= min(s1, ghi); ++gs) {
                        bool any = false;
#pragma unroll
                        for (int k = 0; k < ADJ_GROUP; ++k)
                            any = any || (!((sb >> (4 * k)) & 1) &&
                                          ((gs >= rx[k] && gs <= ry[k]) || (gs >= rz[k] && gs <= rw[k])));
                        if (!any) continue;  // wavefront-uniform
                        ff_step(geo[gs].c_u, geo[gs].h_u, ff_in(lfr, gs) && !le,
                                (const float4*)(A.gff + ((int64_t)lblk * A.n_sub + gs) * FF_TERMS * WAVE));
                    }
                }
            }
            if (act) {
                // accumulator layout of v_mfma_f32_16x16x4_f32: lane i, register r holds row
                // 4 (i / 16) + r, column i % 16; row i of MFMA m is layer 16 m + i of the block
                float* __restrict__ dst = (float*)(A.mom + ((int64_t)line * A.n_lblocks + lblk) * WAVE);
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int r = 0; r < 4; ++r) dst[(16 * q + 4 * kk + r) * 4 + pw] = (float)tot[q][r];
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Hydrogen lines (src/hydrogen_line_absorption.jl:75-150), Float64.  k_hsetup: one thread per
// (layer, transition) -- the table-range check, the line centre and the (T, n_e)-blended log
// profile slice (bilinear over the transition's temperature and density axes: Korg's trilinear
// interpolation, separated), the MHD-weighted amplitude, the grid window and, for H alpha, beta,
// gamma, the resonance-broadened Voigt parameters.  k_hlines: a 32-point tile of the grid per
// workgroup, lanes = layers; the transitions in the tables' order, each adding its resonance
// profile and then its Stark profile; the tile's sums are added into alpha once (no atomics).
// ---------------------------------------------------------------------------------------------
constexpr double RydbergH_eV = 13.598287264;
constexpr double eV_to_cgs = 1.602e-12;
constexpr double bohr_radius_cgs = 5.29177210903e-9;
constexpr double H_MASS = 1.008 * amu_cgs;
constexpr int HCT = 32;  // grid points per k_hlines tile

struct HTrans {  // device copy of korg_htransition
    int32_t lower, upper, n_T, n_ne, n_dnu, pad;
    double log_gf;
    int64_t off_T, off_ne, off_dnu, off_profile, off_lambda0;
};

struct HPair {
    double nu0, F0, amp;
    double lam_res, inv, a, scaling;  // resonance profile (inv == 0: none)
    int32_t lb, ub;                   // grid window, 0-based; lb > ub: the pair adds nothing
};

struct HArgs {
    int L, n_trans, max_dnu, use_MHD;
    int64_t n_wl;
    double window;
    const double *wls, *T, *ne, *nH, *nHe, *U, *xi;
    const HTrans* trans;
    const double *kT, *kne, *kdnu, *logP, *lam0;
    HPair* pairs;
    double* slices;  // [pair][max_dnu]
    int* hull;       // [transition][2]: the union of its windows over the layers (lb > ub: none)
    void* alpha;
    int alpha_f32;
    // Brackett series (lines 4 -> m, m = 5 .. n_max)
    int n_max, n_br;
    struct BRec* brec;  // [layer * n_br + (m - 5)]
    double* bprof;      // [layer * n_br + (m - 5)][BR_POINTS]
    int* bhull;         // [m - 5][2]
};

// Brackett line m of one layer: its convolved Stark profile sampled at start + k step,
// k < BR_POINTS, its amplitude and its window on the grid (0-based; lb > ub: nothing).
constexpr int BR_POINTS = 201;
struct BRec {
    double start, step, amp;
    int32_t lb, ub;
};

// src/statmech.jl:463-497 (without the Hubeny generalization)
__device__ inline double hummer_mihalas_w(double T, double n, double nH, double nHe, double ne) {
    const double r_level = sqrt(5.0 / 2.0 * (n * n * n * n) + 1.0 / 2.0 * (n * n)) * bohr_radius_cgs;
    const double rH = r_level + sqrt(3.0) * bohr_radius_cgs, rHe = r_level + 1.02 * bohr_radius_cgs;
    const double neutral = nH * (rH * rH * rH) + nHe * (rHe * rHe * rHe);
    const double K = n > 3 ? 16.0 / 3.0 * ((n / (n + 1)) * (n / (n + 1))) *
                                 ((n + 7.0 / 6.0) / (n * n + n + 1.0 / 2.0))
                           : 1.0;
    const double chi = RydbergH_eV / (n * n) * eV_to_cgs;
    const double e = electron_charge_cgs;
    const double q = (e * e) / (chi * sqrt(K));
    const double charged = 16.0 * (q * q * q) * ne;
    return exp(-4.0 * PI / 3.0 * (neutral + charged));
}

// the cell i of ascending knots[0..n) with knots[i] <= x < knots[i + 1], for knots[0] < x < knots[n-1]
__device__ inline int knot_cell(const double* k, int n, double x) {
    int i = 0;
    while (i + 2 < n && k[i + 1] <= x) ++i;
    return i;
}

__device__ inline int64_t lower_bound_f64(const double* a, int64_t n, double x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t m = (lo + hi) >> 1;
        if (a[m] < x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

__device__ inline int64_t upper_bound_f64(const double* a, int64_t n, double x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t m = (lo + hi) >> 1;
        if (a[m] <= x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

__global__ __launch_bounds__(64) void k_hinit(HArgs A) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t < A.n_trans) {
        A.hull[2 * t] = INT_MAX;
        A.hull[2 * t + 1] = -1;
    }
}

__global__ __launch_bounds__(64) void k_hsetup(HArgs A) {
    // no fused multiply-adds: the line centre (and so the window edges, which often fall on grid
    // points exactly: the tables' centres and the grids are both in steps of 0.01 A) is rounded
    // as the plain Float64 expression is, operation by operation
#pragma clang fp contract(off)
    const int p = blockIdx.x * 64 + threadIdx.x;
    if (p >= A.L * A.n_trans) return;
    const int l = p / A.n_trans, t = p % A.n_trans;
    const HTrans tr = A.trans[t];
    const double T = A.T[l], ne = A.ne[l];
    HPair h = {};
    h.lb = 1;
    h.ub = 0;
    const double* kT = A.kT + tr.off_T;
    const double* kN = A.kne + tr.off_ne;
    // src/hydrogen_line_absorption.jl:102-104: only inside the tables' (T, n_e) range
    if (!(kT[0] < T && T < kT[tr.n_T - 1] && kN[0] < ne && ne < kN[tr.n_ne - 1])) {
        A.pairs[p] = h;
        return;
    }
    const int i = knot_cell(kT, tr.n_T, T), j = knot_cell(kN, tr.n_ne, ne);
    const double s = (T - kT[i]) / (kT[i + 1] - kT[i]);
    const double u = (ne - kN[j]) / (kN[j + 1] - kN[j]);
    const int nT = tr.n_T;
    auto bilinear = [&](const double* tab) {
        const double lo = tab[j * nT + i] * (1 - s) + tab[j * nT + i + 1] * s;
        const double hi = tab[(j + 1) * nT + i] * (1 - s) + tab[(j + 1) * nT + i + 1] * s;
        return lo * (1 - u) + hi * u;
    };
    double lam0 = bilinear(A.lam0 + tr.off_lambda0);
    // :106-111
    const double beta = 1.0 / (kboltz_eV * T);
    const double Elo = RydbergH_eV * (1.0 - 1.0 / (double)(tr.lower * tr.lower));
    const double Eup = RydbergH_eV * (1.0 - 1.0 / (double)(tr.upper * tr.upper));
    const double w = A.use_MHD ? hummer_mihalas_w(T, (double)tr.upper, A.nH[l], A.nHe[l], ne) : 1.0;
    const double levels = w * (exp(-beta * Elo) - exp(-beta * Eup)) / A.U[l];
    const double amp = pow(10.0, tr.log_gf) * A.nH[l] * sigma_line(lam0) * levels;
    // :113-117
    const int64_t lb = lower_bound_f64(A.wls, A.n_wl, lam0 - A.window);
    const int64_t ub = upper_bound_f64(A.wls, A.n_wl, lam0 + A.window) - 1;
    if (lb >= ub) {
        A.pairs[p] = h;
        return;
    }
    h.lb = (int32_t)lb;
    h.ub = (int32_t)ub;
    h.amp = amp;
    atomicMin(&A.hull[2 * t], (int)lb);
    atomicMax(&A.hull[2 * t + 1], (int)ub);
    // :120-136: H alpha, beta, gamma also get the Barklem et al. resonance profile, centred on
    // fixed wavelengths (which then also centre the Stark profile)
    if (tr.lower == 2 && tr.upper >= 3 && tr.upper <= 5) {
        double sABO, aABO;
        if (tr.upper == 3) {
            lam0 = 6.56460998e-5, sABO = 1180.0, aABO = 0.677;
        } else if (tr.upper == 4) {
            lam0 = 4.8626810200000004e-5, sABO = 2320.0, aABO = 0.455;
        } else {
            lam0 = 4.34168232e-5, sABO = 4208.0, aABO = 0.380;
        }
        const double Gamma = scaled_vdW_abo(sABO * bohr_radius_cgs * bohr_radius_cgs, aABO, H_MASS, T) * A.nH[l];
        const double gamma = Gamma * lam0 * lam0 / (c_cgs * 4.0 * PI);
        const double sigma = doppler_width(lam0, T, H_MASS, A.xi[l]);
        const double inv = 1.0 / (sigma * sqrt(2.0));
        h.lam_res = lam0;
        h.inv = inv;
        h.a = gamma * inv;
        h.scaling = inv / sqrt(PI) * amp;
    }
    // :139-142
    h.nu0 = c_cgs / lam0;
    h.F0 = 1.25e-9 * pow(ne, 2.0 / 3.0);
    A.pairs[p] = h;
    double* sl = A.slices + (int64_t)p * A.max_dnu;
    const double* P = A.logP + tr.off_profile;
    for (int k = 0; k < tr.n_dnu; ++k) sl[k] = bilinear(P + (int64_t)k * tr.n_ne * nT);
}

__global__ __launch_bounds__(64) void k_hlines(HArgs A) {
#pragma clang fp contract(off)
    const int l = blockIdx.y * 64 + threadIdx.x;
    const bool lv = l < A.L;
    const int64_t j0 = (int64_t)blockIdx.x * HCT;
    const int jn = (int)min((int64_t)HCT, A.n_wl - j0);
    double acc[HCT];
#pragma unroll
    for (int q = 0; q < HCT; ++q) acc[q] = 0.0;
    bool touched = false;  // wavefront-uniform: some transition reaches the tile
    for (int t = 0; t < A.n_trans; ++t) {
        // the transition's windows over all layers miss the tile: skip it (uniform loads)
        if (A.hull[2 * t] > j0 + jn - 1 || A.hull[2 * t + 1] < j0) continue;
        const int64_t p = (int64_t)(lv ? l : 0) * A.n_trans + t;
        const HPair h = A.pairs[p];
        const bool ov = lv && h.lb <= h.ub && h.lb < j0 + jn && h.ub >= j0;
        if (!ballot(ov)) continue;  // wavefront-uniform
        touched = true;
        const HTrans tr = A.trans[t];
        const double* kn = A.kdnu + tr.off_dnu;  // kn[0] stands for -floatmax
        const int nd = tr.n_dnu;
        const double* sl = A.slices + p * A.max_dnu;
        const double k1 = kn[1], klast = kn[nd - 1];
#pragma unroll
        for (int q = 0; q < HCT; ++q) {
            const int64_t jj = j0 + q;
            if (!(q < jn && ov && jj >= h.lb && jj <= h.ub)) continue;
            const double lam = A.wls[jj];
            if (h.inv != 0.0) acc[q] += voigt_hjerting64(h.a, fabs(lam - h.lam_res) * h.inv) * h.scaling;
            const double nu = c_cgs / lam;
            double x = fabs(nu - h.nu0) / h.F0;
            if (x == 0.0) x = 2.2250738585072014e-308;  // _zero2epsilon: the smallest normal double
            const double y = log(x);
            double v;
            if (y < k1) {
                v = sl[1];  // the first cell's weight is 1: its left knot is -floatmax
            } else if (y >= klast) {
                v = sl[nd - 1];  // flat beyond the last knot
            } else {
                int lo = 1, hi = nd - 1;  // kn[lo] <= y < kn[hi]
                while (hi - lo > 1) {
                    const int m = (lo + hi) >> 1;
                    if (kn[m] <= y) lo = m;
                    else hi = m;
                }
                const double wgt = (y - kn[lo]) / (kn[lo + 1] - kn[lo]);
                v = sl[lo] * (1 - wgt) + sl[lo + 1] * wgt;
            }
            acc[q] += exp(v) * (c_cgs / (lam * lam)) * h.amp;
        }
    }
    if (!lv || !touched) return;  // a tile no transition reaches leaves alpha as it is
    for (int q = 0; q < jn; ++q) {
        const int64_t k = (int64_t)l + (int64_t)A.L * (j0 + q);
        if (A.alpha_f32) ((float*)A.alpha)[k] += (float)acc[q];
        else ((double*)A.alpha)[k] += acc[q];
    }
}

// ---------------------------------------------------------------------------------------------
// Brackett series (src/hydrogen_line_absorption.jl:152-169): per layer and line 4 -> m, Griem's
// electron-impact and quasistatic (Holtsmark) Stark profiles on BR_POINTS wavelengths across
// 5 max(Doppler width, Stark width) of the line centre (clipped to the grid), convolved with each
// other and, when the Doppler width is at least a quarter of the Stark width, first the
// quasistatic one with the Doppler Gaussian (bracket_line_interpolator, :211-249); the grid
// points within the window then get the linear interpolation of the result times the amplitude.
// ---------------------------------------------------------------------------------------------
__constant__ double c_holt_prob7[15][5] = {
    {0.005, 0.128, 0.260, 0.389, 0.504},     {0.004, 0.109, 0.220, 0.318, 0.389},
    {-0.007, 0.079, 0.162, 0.222, 0.244},    {-0.018, 0.041, 0.089, 0.106, 0.080},
    {-0.026, -0.003, 0.003, -0.023, -0.086}, {-0.025, -0.048, -0.087, -0.148, -0.234},
    {-0.008, -0.085, -0.165, -0.251, -0.343}, {0.018, -0.111, -0.223, -0.321, -0.407},
    {0.032, -0.130, -0.255, -0.354, -0.431}, {0.014, -0.148, -0.269, -0.359, -0.427},
    {-0.005, -0.140, -0.243, -0.323, -0.386}, {0.005, -0.095, -0.178, -0.248, -0.307},
    {-0.002, -0.068, -0.129, -0.187, -0.241}, {-0.007, -0.049, -0.094, -0.139, -0.186},
    {-0.010, -0.036, -0.067, -0.103, -0.143}};
__constant__ double c_holt_c7[5] = {511.318, 1.532, 4.044, 19.266, 41.812};
__constant__ double c_holt_d7[5] = {-6.070, -4.528, -8.759, -14.984, -23.956};
__constant__ double c_holt_pp[5] = {0.0, 0.2, 0.4, 0.6, 0.8};
__constant__ double c_holt_beta[15] = {1.0,  1.259, 1.585, 1.995, 2.512, 3.162, 3.981, 5.012,
                                       6.310, 7.943, 10.0,  12.59, 15.85, 19.95, 25.12};

// src/line_absorption.jl:169-182
__device__ inline double exponential_integral_1(double x) {
#pragma clang fp contract(off)
    if (x < 0) return 0.0;
    if (x <= 0.01) return -log(x) - 0.577215 + x;
    if (x <= 1.0)
        return -log(x) - 0.57721566 +
               x * (0.99999193 + x * (-0.24991055 + x * (0.05519968 + x * (-0.00976004 + x * 0.00107857))));
    if (x <= 30.0) return (x * (x + 2.334733) + 0.25062) / (x * (x + 3.330657) + 1.681534) / x * exp(-x);
    return 0.0;
}

// src/hydrogen_line_absorption.jl:397-445 (0-based table indices)
__device__ inline double holtsmark_profile(double beta, double P) {
#pragma clang fp contract(off)
    if (beta > 500) return (1.5 / sqrt(beta) + 27 / (beta * beta)) / (beta * beta);
    const int im = min((int)floor(5 * P + 1), 4) - 1, ip = im + 1;
    const double WTPP = 5 * (P - c_holt_pp[im]), WTPM = 1 - WTPP;
    if (beta <= 25.12) {
        int jp = 0;
        while (beta > c_holt_beta[jp]) ++jp;  // the first knot >= beta (beta <= 25.12: it exists)
        jp = max(1, jp);
        const int jm = jp - 1;
        const double WTBP = (beta - c_holt_beta[jm]) / (c_holt_beta[jp] - c_holt_beta[jm]);
        const double WTBM = 1 - WTBP;
        const double CBP = c_holt_prob7[jp][ip] * WTPP + c_holt_prob7[jp][im] * WTPM;
        const double CBM = c_holt_prob7[jm][ip] * WTPP + c_holt_prob7[jm][im] * WTPM;
        const double CORR = 1 + CBP * WTBP + CBM * WTBM;
        const double WT = fmax(fmin(0.5 * (10 - beta), 1.0), 0.0);
        const double PR1 = beta <= 10 ? 8 / (83 + (2 + 0.95 * (beta * beta)) * beta) : 0.0;
        const double PR2 = beta >= 8 ? (1.5 / sqrt(beta) + 27 / (beta * beta)) / (beta * beta) : 0.0;
        return (PR1 * WT + PR2 * (1 - WT)) * CORR;
    }
    const double CC = c_holt_c7[ip] * WTPP + c_holt_c7[im] * WTPM;
    const double DD = c_holt_d7[ip] * WTPP + c_holt_d7[im] * WTPM;
    const double CORR = 1 + DD / (CC + beta * sqrt(beta));
    return (1.5 / sqrt(beta) + 27 / (beta * beta)) / (beta * beta) * CORR;
}

// src/hydrogen_line_absorption.jl:370-389 (n = 4)
__device__ inline double greim_1960_Knm_4(int m) {
    if (m - 4 <= 3) return m == 5 ? 0.5820000 : (m == 6 ? 0.8660000 : 1.0200000);
    const double md = m, m4 = (md * md) * (md * md);
    return 5.5e-5 * 256.0 * m4 / (md * md - 16.0) / (1 + 0.13 / (md - 4));
}

// src/hydrogen_line_absorption.jl:180-191 (n = 4)
__device__ inline double brackett_oscillator_strength(int m) {
#pragma clang fp contract(off)
    const double n = 4.0, md = m;
    const double GINF = 0.2027 / pow(n, 0.71), GCA = 0.124 / n, FKN = 1.9603 * n;
    const double WTC = 0.45 - 2.4 / (n * n * n) * (n - 1);
    const double r = md / ((md - n) * (md + n));
    const double FK = FKN * (r * r * r);
    const double XMN12 = pow(md - n, 1.2);
    const double WT = (XMN12 - 1) / (XMN12 + WTC);
    return FK * (1 - WT * GINF - (0.222 + GCA / md) * (1 - WT));
}

// The central BR_POINTS values of the full convolution of a and b (both BR_POINTS long): point k
// sums a[i] b[(BR_POINTS - 1)/2 + k - i] over the valid i, ascending.
__device__ inline double conv_center(const double* a, const double* b, int k) {
    const int s = (BR_POINTS - 1) / 2 + k;
    const int i0 = max(0, s - (BR_POINTS - 1)), i1 = min(BR_POINTS - 1, s);
    double acc = 0.0;
    for (int i = i0; i <= i1; ++i) acc += a[i] * b[s - i];
    return acc;
}

// One block of 256 threads per (layer, line); thread k < BR_POINTS owns sample k.
__global__ __launch_bounds__(256) void k_bsetup(HArgs A) {
#pragma clang fp contract(off)
    __shared__ double s_imp[BR_POINTS], s_q[BR_POINTS], s_g[BR_POINTS];
    const int p = blockIdx.x;
    const int l = p / A.n_br, m = 5 + p % A.n_br;
    const int k = threadIdx.x;
    const double T = A.T[l], ne = A.ne[l], xi = A.xi[l];
    const double lmin = A.wls[0], lmax = A.wls[A.n_wl - 1];
    // :154-159: the line centre and amplitude
    const double E = RydbergH_eV * (1.0 / 16.0 - 1.0 / (double)(m * m));
    const double lam0 = hplanck_eV * c_cgs / E;
    // :218-234: widths, window and the sampled range
    const double F0 = 1.25e-9 * pow(ne, 2.0 / 3.0);
    const double Knm = greim_1960_Knm_4(m);
    const double stark_width = 1.6678E-18 * Knm * F0 * c_cgs;
    const double sdop = doppler_width(lam0, T, H_MASS, xi);
    const double window = 5.0 * fmax(sdop, stark_width);
    const double start = fmax(lmin, lam0 - window), end = fmin(lam0 + window, lmax);
    BRec rec = {0.0, 0.0, 0.0, 1, 0};
    if (start > lmax || end < lmin || start == end) {  // block-uniform: no contribution
        if (k == 0) A.brec[p] = rec;
        return;
    }
    const double step = (end - start) / (BR_POINTS - 1);
    // :272-351: the two Stark profiles at start + k step
    const double nu0 = c_cgs / lam0;
    double lam = 0.0, nu = 0.0, imp = 0.0, q = 0.0;
    bool red = false;
    if (k < BR_POINTS) {
        lam = start + step * k;
        nu = c_cgs / lam;
        red = nu < nu0;
        const double ne_1_6 = pow(ne, 1.0 / 6.0);
        const double GCON1 = 0.2 + 0.09 * sqrt(T / 10000.0) / (1 + ne / 1.0e13);
        const double GCON2 = 0.2 / (1 + ne / 1.0e15);
        const double WTY1 = 1 / (1 + ne / (m - 4 <= 3 ? 1e14 : 1e13));
        const double Y1B = 2 / (1 + 0.012 / T * sqrt(ne / T));
        const double mm = (double)(m * m - 16);
        const double C1CON = Knm / lam0 * (mm * mm) / (16.0 * (double)(m * m)) * 1e-8;
        const double Y1SCAL = 320 * (pow(T / 10000.0, 0.3) / ne_1_6) * WTY1 + Y1B * (1 - WTY1);
        const double C1 = F0 * 78940 / T * C1CON * Y1SCAL;
        const double C2 = (F0 * F0) / (5.96E-23 * ne) * ((Knm / lam0) * (Knm / lam0)) * 1e-16;
        const double beta = fabs(lam - lam0) / F0 / Knm * 1e8;
        const double y1 = C1 * beta, y2 = C2 * (beta * beta);
        const double G1 = 6.77 * sqrt(C1);
        double width;
        if (y2 <= 1e-4 && y1 <= 1e-5) {
            width = G1 * fmax(0.0, 0.2114 + log(sqrt(C2) / C1)) * (1 - GCON1 - GCON2);
        } else {
            const double y1c = 90 * y1;
            const double GAM = G1 *
                               (0.5 * exp(-fmin(80.0, y1)) + exponential_integral_1(y1) -
                                0.5 * exponential_integral_1(y2)) *
                               (1 - GCON1 / (1 + y1c * y1c * y1c) - GCON2 / (1 + 2000 * y1));
            width = GAM <= 1e-20 ? 0.0 : GAM;
        }
        imp = width > 0 ? width / (PI * (width * width + beta * beta)) : 0.0;
        const double PP = ne_1_6 * 0.08989 / sqrt(T);
        const double ps = (0.9 * y1) * (0.9 * y1);
        const double qe = (ps + 0.03 * sqrt(y1)) / (ps + 1.0);
        q = holtsmark_profile(beta, PP) * (1 + qe);
    }
    // every sample is rescaled by the red-wing Boltzmann factor once any sample is redward
    red = __syncthreads_or(red);
    if (k < BR_POINTS) {
        const double dbeta_dlam = 1e8 / (Knm * F0);
        const double sq = sqrt(lam / lam0);
        const double bz = red ? exp((hplanck_cgs * (nu - nu0)) / kboltz_cgs / T) : 1.0;
        imp = imp * sq;
        q = q * sq;
        if (red) {
            imp = imp * bz;
            q = q * bz;
        }
        s_imp[k] = imp * dbeta_dlam;
        s_q[k] = q * dbeta_dlam;
    }
    if (sdop / stark_width > 0.25) {  // block-uniform
        if (k < BR_POINTS) {
            const double d = lam - lam0;
            s_g[k] = exp(-0.5 * (d * d) / (sdop * sdop)) / sqrt(2 * PI) / sdop;
        }
        __syncthreads();
        const double c = k < BR_POINTS ? conv_center(s_q, s_g, k) * step : 0.0;
        __syncthreads();
        if (k < BR_POINTS) s_q[k] = c;
    }
    __syncthreads();
    if (k < BR_POINTS) A.bprof[(int64_t)p * BR_POINTS + k] = conv_center(s_imp, s_q, k) * step;
    if (k == 0) {
        const double beta_T = 1.0 / (kboltz_eV * T);
        const double E_low = RydbergH_eV * (1.0 - 1.0 / 16.0);
        const double w = A.use_MHD ? hummer_mihalas_w(T, (double)m, A.nH[l], A.nHe[l], ne) : 1.0;
        const double levels = w * exp(-beta_T * E_low) * (1 - exp(-beta_T * E)) / A.U[l];
        const double gf = 2.0 * 16.0 * brackett_oscillator_strength(m);
        rec.start = start;
        rec.step = step;
        rec.amp = gf * A.nH[l] * sigma_line(lam0) * levels;
        const int64_t lb = lower_bound_f64(A.wls, A.n_wl, lam0 - window);
        const int64_t ub = upper_bound_f64(A.wls, A.n_wl, lam0 + window) - 1;
        if (lb <= ub) {
            rec.lb = (int32_t)lb;
            rec.ub = (int32_t)ub;
            atomicMin(&A.bhull[2 * (m - 5)], (int)lb);
            atomicMax(&A.bhull[2 * (m - 5) + 1], (int)ub);
        }
        A.brec[p] = rec;
    }
}

__global__ __launch_bounds__(64) void k_binit(HArgs A) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t < A.n_br) {
        A.bhull[2 * t] = INT_MAX;
        A.bhull[2 * t + 1] = -1;
    }
}

// One workgroup per HCT-point tile, lanes = layers: the Brackett lines whose window meets the
// point, in order of m, summed and added into alpha once.
__global__ __launch_bounds__(64) void k_blines(HArgs A) {
#pragma clang fp contract(off)
    const int l = blockIdx.y * 64 + threadIdx.x;
    const bool lv = l < A.L;
    const int64_t j0 = (int64_t)blockIdx.x * HCT;
    const int jn = (int)min((int64_t)HCT, A.n_wl - j0);
    double acc[HCT];
#pragma unroll
    for (int q = 0; q < HCT; ++q) acc[q] = 0.0;
    bool touched = false;  // wavefront-uniform
    for (int mi = 0; mi < A.n_br; ++mi) {
        if (A.bhull[2 * mi] > j0 + jn - 1 || A.bhull[2 * mi + 1] < j0) continue;  // uniform
        const int64_t p = (int64_t)(lv ? l : 0) * A.n_br + mi;
        const BRec r = A.brec[p];
        const bool ov = lv && r.lb <= r.ub && r.lb < j0 + jn && r.ub >= j0;
        if (!ballot(ov)) continue;
        touched = true;
        const double* prof = A.bprof + p * BR_POINTS;
#pragma unroll
        for (int q = 0; q < HCT; ++q) {
            const int64_t jj = j0 + q;
            if (!(q < jn && ov && jj >= r.lb && jj <= r.ub)) continue;
            const double t = (A.wls[jj] - r.start) / r.step;
            int i = (int)floor(t);
            i = i < 0 ? 0 : (i > BR_POINTS - 2 ? BR_POINTS - 2 : i);
            const double w = t - i;
            acc[q] += (prof[i] * (1 - w) + prof[i + 1] * w) * r.amp;
        }
    }
    if (!lv || !touched) return;
    for (int q = 0; q < jn; ++q) {
        const int64_t k = (int64_t)l + (int64_t)A.L * (j0 + q);
        if (A.alpha_f32) ((float*)A.alpha)[k] += (float)acc[q];
        else ((double*)A.alpha)[k] += acc[q];
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Molecular cross-sections (src/molecular_cross_sections.jl:99-114): alpha[l, j] += sigma(vmic_l,
// log10 T_l, lambda_j) n_l for each precomputed table, sigma interpolated linearly in each of the
// three gridded dimensions and zero outside the table (extrapolate(..., 0.0), :76-78).  A block
// covers MOL_TILE grid points x 64 layers; lane = layer.  The table's lambda interval of each
// point is found once per block (lane = point) and shared through LDS.
// ---------------------------------------------------------------------------------------------
constexpr int MOL_TILE = 64;

struct MolTable {  // device copy of korg_mol_xsection
    int n_vmic, n_logT;
    int64_t n_wl;
    const double *vmic, *logT, *wls, *sigma, *n_species;
};

// Interval i (0 .. n-2) of ascending knots x holding v and its weight (v - x_i)/(x_{i+1} - x_i);
// false when v is outside [x_0, x_{n-1}] (or NaN).
__device__ inline bool gridded_cell(const double* x, int64_t n, double v, int64_t& i, double& w) {
    if (!(v >= x[0] && v <= x[n - 1])) return false;
    int64_t lo = 0, hi = n - 1;  // x[lo] <= v, and v < x[hi] unless v is the last knot
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (x[mid] <= v) lo = mid; else hi = mid;
    }
    i = lo;
    w = (v - x[lo]) / (x[lo + 1] - x[lo]);
    return true;
}

__global__ __launch_bounds__(64) void k_mol_interp(int L, int64_t n_wl, const double* __restrict__ wls,
                                                   const double* __restrict__ temps,
                                                   const double* __restrict__ vmic, int n_tables,
                                                   const MolTable* __restrict__ tabs,
                                                   void* __restrict__ alpha, int alpha_f32) {
    __shared__ int64_t s_i[MOL_TILE];
    __shared__ double s_w[MOL_TILE];
    const int lane = threadIdx.x;
    const int l = blockIdx.y * 64 + lane;
    const int64_t j0 = (int64_t)blockIdx.x * MOL_TILE;
    const int np = (int)min((int64_t)MOL_TILE, n_wl - j0);
    for (int k = 0; k < n_tables; ++k) {
        const MolTable tb = tabs[k];
        __syncthreads();  // the previous table's cells are no longer read
        {
            int64_t i = -1;
            double w = 0.0;
            if (lane < np && !gridded_cell(tb.wls, tb.n_wl, wls[j0 + lane], i, w)) i = -1;
            s_i[lane] = i;
            s_w[lane] = w;
        }
        __syncthreads();
        if (l >= L) continue;
        int64_t iv = 0, iT = 0;
        double wv = 0.0, wT = 0.0;
        const bool in = gridded_cell(tb.vmic, tb.n_vmic, vmic[l], iv, wv) &&
                        gridded_cell(tb.logT, tb.n_logT, log10(temps[l]), iT, wT);
        if (!in) continue;
        const double n = tb.n_species[l];
        const int64_t nT = tb.n_logT;
        // the four (vmic, log T) corners of this layer, as offsets into a lambda row
        const double* r0 = tb.sigma + (iv * tb.n_wl) * nT + iT;        // vmic knot iv
        const double* r1 = tb.sigma + ((iv + 1) * tb.n_wl) * nT + iT;  // vmic knot iv + 1
        for (int t = 0; t < np; ++t) {
            const int64_t i = s_i[t];
            if (i < 0) continue;  // wavefront-uniform: the point is outside the table
            const double w = s_w[t];
            auto corner = [&](const double* r, int64_t jj) {  // linear in log T at lambda knot jj
                const double* c = r + jj * nT;
                return (1.0 - wT) * c[0] + wT * c[1];
            };
            const double a0 = (1.0 - w) * corner(r0, i) + w * corner(r0, i + 1);
            const double a1 = (1.0 - w) * corner(r1, i) + w * corner(r1, i + 1);
            const double v = ((1.0 - wv) * a0 + wv * a1) * n;
            const int64_t idx = (int64_t)l + (int64_t)L * (j0 + t);
            if (alpha_f32) ((float*)alpha)[idx] += (float)v;
            else ((double*)alpha)[idx] += v;
        }
    }
}

// ============================================================================================
// host side
// ============================================================================================

// Scratch of korg_line_absorption_device_noalloc: one fixed array per GPU, held from the library's
// load on that device until the process ends.  Host code never passes this name to a kernel (on
// the host it is only a stand-in for the symbol); it uses the device address that
// hipGetSymbolAddress returns for the current device (korg_create resolves and keeps it).
static __device__ __align__(256) unsigned char scratch[KORG_NOALLOC_SCRATCH_BYTES];

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

// Bump allocation from a fixed device block: while a context's `arena` is set, grow() carves
// buffers from it instead of calling hipMalloc (and frees nothing).
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    size_t off = 0;
};

// Device scratch of the forward pipeline (grow-only).  The context owns one for its direct calls;
// every plan owns its own, so the pointers a captured graph holds stay valid whatever later calls
// on the context allocate.
struct Workspace {
    DevBuf nbuf;  // the near zone's sums when it runs beside the far wing
    DevBuf ffgeo, fft;  // far field: sub-tile geometry, the sums T of every sub-tile
    DevBuf ffrun, ffkey;  // far field: each line's runs (line order), list keys (partition order)
    DevBuf ffrun_adj, gff;  // the adjoint's far field: runs clear of its near zone, moments of G
    DevBuf lpre, lwin;      // the setup's per-line records (k_setup_pre, k_setup)
    DevBuf azq;             // Float32 fit step: each 16-layer block's adjoint near zone
    DevBuf gffs, ffsgeo;    // the adjoint's far field over whole super-tiles: G's moments, geometry
    DevBuf meta, rec64, recM, tab, counters, blk, cstart, cls_cand, cls_jc, ranges, nblk, ncstart,
        near_cand, near_jc, near_ranges, cls_rng, cls_line, near_rng, near_line, tc_count,
        tc_start, tile_lines, stars;
    // adjoint and fit step (adj_*, gexp, g32: the series form's moments)
    DevBuf vjp_cnt, vjp_start, seg_cnt, seg_start, vjp_partial, rt_tab, fit_G, chi2_part, adj_blk,
        adj_start, adj_line, gexp, g32, lam_s, azone, adj_lead, adj_groups, adj_gstart, scan_part,
        jvp_coef;
    int64_t tile_cap = 0;             // entries tile_lines holds
    int64_t seg_cap = 0;              // adjoint segments vjp_partial holds
    std::vector<StarDev> stars_host;  // the star table (host copy)
    // Drops every buffer without freeing it (a workspace carved from an Arena).
    void forget() {
        DevBuf* bufs[] = {&meta,     &rec64,     &recM,      &tab,        &counters,  &blk,
                          &cstart,   &cls_cand,  &cls_jc,    &ranges,     &nblk,      &ncstart,
                          &near_cand, &near_jc,  &near_ranges, &cls_rng,  &cls_line,  &near_rng,
                          &near_line, &tc_count, &tc_start,  &tile_lines, &stars,     &scan_part,
                          &nbuf,      &ffgeo,    &fft,       &ffrun,      &ffkey,
                          &ffrun_adj, &gff,      &lpre,      &lwin,      &azq,
                          &gffs,      &ffsgeo};
        for (DevBuf* b : bufs) *b = DevBuf{};
        tile_cap = 0;
        seg_cap = 0;
    }
    void release() {
        DevBuf* bufs[] = {&meta,     &rec64,     &recM,      &tab,        &counters,  &blk,
                          &cstart,   &cls_cand,  &cls_jc,    &ranges,     &nblk,      &ncstart,
                          &near_cand, &near_jc,  &near_ranges, &cls_rng,  &cls_line,  &near_rng,
                          &near_line, &tc_count, &tc_start,  &tile_lines, &stars,     &vjp_cnt,
                          &vjp_start, &seg_cnt,  &seg_start, &vjp_partial, &rt_tab,   &fit_G,
                          &chi2_part, &adj_blk,  &adj_start, &adj_line,  &gexp,       &g32,
                          &lam_s,     &azone,    &adj_lead,  &adj_groups, &adj_gstart,
                          &scan_part, &jvp_coef, &nbuf,      &ffgeo,     &fft,       &ffrun,
                          &ffkey,     &ffrun_adj, &gff,      &lpre,      &lwin,      &azq,
                          &gffs,      &ffsgeo};
        for (DevBuf* b : bufs) {
            if (b->p) (void)hipFree(b->p);
            b->p = nullptr;
            b->cap = 0;
        }
        tile_cap = 0;
        seg_cap = 0;
    }
};

struct korg_ctx {
    int device = 0;
    int n_cu = 0;
    std::string arch;  // gcnArchName of the device
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    // [0] start, [1] setup done, [2] partition done, [3] end, [4] far wing done,
    // [5] tile-list total known, [6] tile-list fill launched (the host sizes the lists between)
    hipEvent_t ev[10] = {};
    // the Float32 form's near zone runs on `side`, beside the far wing (fork/join events)
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    std::string err;
    korg_stats stats{};
    korg_tuning tuning{};
    int64_t last_n_lines = 0;
    int last_precision = KORG_PRECISION_FP64;
    int last_n_lblocks = 0;
    int last_nb = 0;  // 32-point blocks per far-wing tile of the last call
    bool last_ff = false;  // the last forward call ran the far field (ev[9] marks its end)
    int last_ntiles = 0;   // far-wing tiles of the last call
    bool last_sized = false;  // the last forward call sized the tile lists on the host (ev[5], ev[6])
    Workspace ws;
    // the allocation-free entry point: its workspace (carved from the device's fixed scratch at
    // every call), the scratch's device address on this context's device, and whether the last
    // forward call used it (korg_synchronize, korg_get_windows then read ws_fixed)
    Workspace ws_fixed;
    void* scratch_ptr = nullptr;
    std::string scratch_err;
    Arena* arena = nullptr;
    bool last_fixed = false;
    // a batch of several layer counts runs as one pipeline per count: the windows of all its
    // lines, in the caller's star order (korg_get_windows reads them when last_ragged is set)
    bool last_ragged = false;
    bool last_overlap = false;  // the last call's near zone ran beside the wing (its own events)
    std::vector<int64_t> ragged_lb, ragged_ub;
    std::vector<int64_t> last_grid_off, last_line_off;  // per star of the last call
    // adjoint outputs of the host-memory entry point, and its staged G
    DevBuf vjp_out, h_grad;
    // staging for the host-memory entry points
    DevBuf h_alpha, h_wls, h_lines[8], h_win[2], h_species_mass, h_species_mol, h_layers[7],
        h_knots, h_cntm;
    // staging of the host-memory forward-mode entry point: tangents and d alpha
    DevBuf h_tan, h_dalpha;
    // molecular cross-sections: the device table structs, and the host entry point's staging
    DevBuf mol_tabs, h_mol;
    // filter_linelist: search keys, search results, gather runs
    DevBuf flt_keys, flt_found, flt_runs;
    // hydrogen-line tables (korg_set_hydrogen_tables) and the per-call pairs and profile slices
    DevBuf hl_trans, hl_T, hl_ne, hl_dnu, hl_logP, hl_lam0, hl_pairs, hl_slices, hl_hull;
    DevBuf hl_brec, hl_bprof, hl_bhull;  // Brackett series
    int hl_ntrans = 0, hl_max_dnu = 0, hl_nmax = 0;
};

namespace {

int set_err(korg_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

#define HIP_TRY(ctx, call)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_err(ctx, KORG_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

int grow(korg_ctx* ctx, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return KORG_OK;
    if (ctx->arena) {  // carve from the fixed block; nothing is allocated or freed
        Arena& a = *ctx->arena;
        const size_t off = (a.off + 255) & ~(size_t)255;
        if (off > a.cap || bytes > a.cap - off)
            return set_err(ctx, KORG_ERR_CAPACITY, "scratch exhausted (%zu of %zu bytes in use, "
                           "%zu more needed)", a.off, a.cap, bytes);
        b.p = a.base + off;
        b.cap = bytes;
        a.off = off + bytes;
        return KORG_OK;
    }
    if (b.p) {
        HIP_TRY(ctx, hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_err(ctx, KORG_ERR_OUT_OF_MEMORY, "hipMalloc(%zu bytes) failed: %s", bytes,
                       hipGetErrorString(e));
    }
    b.cap = bytes;
    return KORG_OK;
}

#define TRY(x)                      \
    do {                            \
        int rc_ = (x);              \
        if (rc_ != KORG_OK) return rc_; \
    } while (0)

constexpr int F64_GROUP = 8;  // layers per work item of the Float64 form

int validate(korg_ctx* ctx, const korg_grid* g, const korg_lines* ln, const korg_species* sp,
             const korg_layers* ly, const korg_continuum* cn, int alpha_dtype, int precision) {
    if (!g || !ln || !sp || !ly || !cn) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "null struct");
    if (g->n_wl < 1 || !g->all_wls) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "empty wavelength grid");
    if (g->n_wl > 0x7fffffff - 4096)
        return set_err(ctx, KORG_ERR_UNSUPPORTED, "grids of more than 2^31 points are not supported");
    if (ln->n_lines < 0) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "negative line count");
    if (ln->n_lines > 0x7fffffff) return set_err(ctx, KORG_ERR_UNSUPPORTED, "more than 2^31 lines");
    if (ly->n_layers < 1) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "no layers");
    if (sp->n_species < 1 && ln->n_lines > 0)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "no species");
    if (alpha_dtype != KORG_DTYPE_F64 && alpha_dtype != KORG_DTYPE_F32)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "bad alpha dtype %d", alpha_dtype);
    if (precision != KORG_PRECISION_FP64 && precision != KORG_PRECISION_FP32_FIXUP &&
        precision != KORG_PRECISION_FP64_SERIES)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "bad precision %d", precision);
    if (cn->mode != KORG_CNTM_AT_LINES && cn->mode != KORG_CNTM_GRID)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "bad continuum mode %d", cn->mode);
    if (cn->mode == KORG_CNTM_GRID && (cn->n_knots < 2 || !cn->knots))
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "continuum grid needs >= 2 knots");
    if ((int64_t)ly->n_layers * ln->n_lines > ((int64_t)1 << 40))
        return set_err(ctx, KORG_ERR_UNSUPPORTED, "problem too large");
    return KORG_OK;
}

// One star of a call, as the host sees it: device pointers.
struct StarHost {
    void* alpha;
    korg_grid grid;
    korg_lines lines;
    korg_layers layers;
    korg_continuum cntm;
    double cutoff;
    // adjoint input and outputs, fit target (device pointers; null when unused)
    void* G = nullptr;
    double* gq[7] = {};
    const double *tau_ref = nullptr, *alpha_ref = nullptr, *data = nullptr, *sigma = nullptr;
    double *flux = nullptr, *chi2 = nullptr;
    double* flux_c = nullptr;  // flux-only transfer: the continuum's flux
};

// Validates the stars of a batched call (one species table for all; *L_out: star 0's layer
// count -- layer_groups splits a batch of several counts).
int collect_stars(korg_ctx* ctx, int64_t n_stars, const korg_star* stars,
                  const korg_species* species, int alpha_dtype, int precision,
                  std::vector<StarHost>& hs, int* L_out) {
    if (n_stars < 1 || !stars) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "no stars");
    if (n_stars > 0x7fffffff) return set_err(ctx, KORG_ERR_UNSUPPORTED, "too many stars");
    hs.resize((size_t)n_stars);
    const int L = stars[0].layers.n_layers;
    for (int64_t k = 0; k < n_stars; ++k) {
        const korg_star& x = stars[k];
        const int rc = validate(ctx, &x.grid, &x.lines, species, &x.layers, &x.cntm, alpha_dtype,
                                precision);
        if (rc != KORG_OK) {
            ctx->err = "star " + std::to_string(k) + ": " + ctx->err;
            return rc;
        }
        if (!x.alpha && x.lines.n_lines > 0)
            return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "star %lld: alpha is null", (long long)k);
        hs[k] = StarHost{x.alpha, x.grid, x.lines, x.layers, x.cntm, x.cutoff_threshold};
    }
    *L_out = L;
    return KORG_OK;
}

// The stars of a batch grouped by layer count, in order of first appearance (the stars of a group
// in batch order): every pipeline runs over one count.
std::vector<std::vector<int>> layer_groups(const std::vector<StarHost>& hs) {
    std::vector<std::vector<int>> g;
    std::vector<int> count;
    for (int k = 0; k < (int)hs.size(); ++k) {
        const int L = hs[k].layers.n_layers;
        size_t i = 0;
        while (i < count.size() && count[i] != L) ++i;
        if (i == count.size()) {
            count.push_back(L);
            g.emplace_back();
        }
        g[i].push_back(k);
    }
    return g;
}

std::vector<StarHost> subset(const std::vector<StarHost>& hs, const std::vector<int>& idx) {
    std::vector<StarHost> out;
    out.reserve(idx.size());
    for (int k : idx) out.push_back(hs[k]);
    return out;
}

void add_stats(korg_stats& a, const korg_stats& b) {
    a.n_evals += b.n_evals;
    a.n_lines_used += b.n_lines_used;
    a.setup_ms += b.setup_ms;
    a.bin_ms += b.bin_ms;
    a.profile_ms += b.profile_ms;
    a.core_ms += b.core_ms;
    a.total_ms += b.total_ms;
    for (int k = 0; k < 4; ++k) a.n_evals_branch[k] += b.n_evals_branch[k];
    a.n_evals_near += b.n_evals_near;
    a.n_wing_mfma += b.n_wing_mfma;
    a.n_evals_tail += b.n_evals_tail;
    a.n_evals_near_A += b.n_evals_near_A;
}

// Lays the stars out in the global index spaces (grids padded to STAR_PAD points, lines
// concatenated) and uploads the star table to ws.stars on `s`.  Returns the totals.
int upload_stars(korg_ctx* ctx, Workspace& ws, hipStream_t s, int n_stars, const StarHost* st,
                 int64_t* n_lines_total, int64_t* n_wl_total) {
    auto& h = ws.stars_host;
    h.assign(n_stars, StarDev{});
    ctx->last_grid_off.resize(n_stars);
    ctx->last_line_off.resize(n_stars);
    int64_t go = 0, lo = 0;
    for (int k = 0; k < n_stars; ++k) {
        const StarHost& x = st[k];
        StarDev& d = h[k];
        d.wls = x.grid.all_wls;
        d.n_wl = x.grid.n_wl;
        d.grid_off = go;
        d.line_off = lo;
        d.n_lines = x.lines.n_lines;
        d.alpha = x.alpha;
        d.wl = x.lines.wl;
        d.log_gf = x.lines.log_gf;
        d.E_lower = x.lines.E_lower;
        d.gamma_rad = x.lines.gamma_rad;
        d.gamma_stark = x.lines.gamma_stark;
        d.vdW1 = x.lines.vdW1;
        d.vdW2 = x.lines.vdW2;
        d.species = x.lines.species;
        const bool frozen = x.lines.window_lb && x.lines.window_ub;
        d.win_lb = frozen ? x.lines.window_lb : nullptr;
        d.win_ub = frozen ? x.lines.window_ub : nullptr;
        d.temps = x.layers.temps;
        d.n_e = x.layers.n_e;
        d.n_div_U = x.layers.n_div_U;
        d.n_HI = x.layers.n_HI;
        d.n_HeI = x.layers.n_HeI;
        d.n_H2 = x.layers.n_H2;
        d.xi = x.layers.xi;
        d.cntm_mode = x.cntm.mode;
        d.n_knots = x.cntm.n_knots;
        d.knots = x.cntm.knots;
        d.cntm = x.cntm.values;
        d.cutoff = x.cutoff;
        d.G = x.G;
        for (int q = 0; q < 7; ++q) d.gq[q] = x.gq[q];
        d.tau_ref = x.tau_ref;
        d.alpha_ref = x.alpha_ref;
        d.data = x.data;
        d.sigma = x.sigma;
        d.flux = x.flux;
        d.chi2 = x.chi2;
        d.flux_c = x.flux_c;
        ctx->last_grid_off[k] = go;
        ctx->last_line_off[k] = lo;
        go += (x.grid.n_wl + STAR_PAD - 1) / STAR_PAD * STAR_PAD;
        lo += x.lines.n_lines;
    }
    if (go > (int64_t)INT32_MAX || lo > (int64_t)INT32_MAX)
        return set_err(ctx, KORG_ERR_UNSUPPORTED,
                       "more than 2^31 grid points or lines in one call; split the batch");
    TRY(grow(ctx, ws.stars, sizeof(StarDev) * (size_t)n_stars));
    HIP_TRY(ctx, hipMemcpyAsync(ws.stars.p, h.data(), sizeof(StarDev) * (size_t)n_stars,
                                hipMemcpyHostToDevice, s));
    *n_lines_total = lo;
    *n_wl_total = go;
    return KORG_OK;
}

// Sizes and launch shapes of one forward call over a batch of stars.
struct Batch {
    int n_stars = 0, L = 0;
    int64_t N = 0, nwl = 0;
    bool f32 = false, alpha_f32 = false;
    bool s64 = false;  // the Float64 series form (f32 is set too: the same zones and lists)
    bool overlap = false;  // Float32 form: the near zone beside the far wing (second stream)
    bool adj = false;  // Float32 fit step: the adjoint's series form (near zones, moments)
    bool ff = false;   // Float32 form: the far wing's far field (k_wing_ff)
    int64_t n_sub = 0;  // far-field sub-tiles (FF_SUB points) and super-tiles (FF_NS sub-tiles)
    int nsuper = 0;
    int NB = 2, WW = 4, P64 = 1, wpc = 12, near_wpc = 16, near_layout = 2, count_branches = 0;
    int tile_w = 0, ntiles = 0, nt_near = 0, n_groups = 0, Ls = 0, n_lblocks = 0, nblocks = 0;
    korg_species sp{};
};

// The far-wing kernel variants: (32-point blocks per tile, waves per tile).
bool wing_variant_ok(int NB, int WW) {
    return (NB == 1 || NB == 2 || NB == 4) && (WW == 1 || WW == 2 || WW == 4) && !(NB == 4 && WW == 4);
}

int check_tuning(korg_ctx* ctx, const korg_tuning* t) {
    const int tp = t->tile_points, wt = t->waves_per_tile;
    if (t->points_per_lane < 0 || t->points_per_lane > 2 || t->waves_per_cu < 0 ||
        t->waves_per_cu > 64 || t->near_waves_per_cu < 0 || t->near_waves_per_cu > 64 ||
        t->count_branches < 0 || t->count_branches > 1 || t->near_layout < 0 || t->near_layout > 2 ||
        t->near_concurrent < 0 || t->near_concurrent > 2 || t->far_field < 0 || t->far_field > 2 ||
        !(tp == 0 || tp == 32 || tp == 64 || tp == 128) || !(wt == 0 || wt == 1 || wt == 2 || wt == 4))
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "tuning out of range");
    if (!wing_variant_ok(tp ? tp / 32 : 2, wt ? wt : 4))
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT,
                       "tuning: 128-point tiles take at most 2 waves per tile");
    return KORG_OK;
}

// Exclusive scan of n counts into start[0..n] on s (see k_scan_flat); part: scratch of
// scan_blocks(n) + 1 entries.
int64_t scan_blocks(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

void enqueue_scan(hipStream_t s, int64_t n, const int* count, int64_t* start, int64_t cap,
                  unsigned long long* counters, int slot, int64_t* part) {
    if (n <= SCAN_TILE) {
        hipLaunchKernelGGL(k_scan_flat<int>, dim3(1), dim3(WAVE), 0, s, n, count, start, cap,
                           counters, slot);
        return;
    }
    const int64_t nb = scan_blocks(n);
    hipLaunchKernelGGL(k_scan_part, dim3((unsigned)nb), dim3(WAVE), 0, s, n, count, part);
    hipLaunchKernelGGL(k_scan_flat<int64_t>, dim3(1), dim3(WAVE), 0, s, nb, (const int64_t*)part,
                       part, cap, counters, slot);
    hipLaunchKernelGGL(k_scan_fill, dim3((unsigned)nb), dim3(WAVE), 0, s, n, count,
                       (const int64_t*)part, start);
}

// Computes the batch's shapes from the tuning, uploads the star table and grows every buffer of
// ws except the per-tile line lists (sized from the device count, by the caller).
int plan_batch(korg_ctx* ctx, const korg_tuning& t, int n_stars, const StarHost* stars,
               const korg_species* sp, int L, int alpha_dtype, int precision, Workspace& ws,
               hipStream_t s, Batch& b) {
    b = Batch{};
    b.n_stars = n_stars;
    b.L = L;
    b.sp = *sp;
    b.s64 = precision == KORG_PRECISION_FP64_SERIES;
    b.f32 = precision == KORG_PRECISION_FP32_FIXUP || b.s64;
    b.alpha_f32 = alpha_dtype == KORG_DTYPE_F32;
    b.NB = b.s64 ? 1 : (t.tile_points ? t.tile_points / 32 : 2);  // the Float64 wing: 32 points
    b.WW = t.waves_per_tile ? t.waves_per_tile : 4;
    b.P64 = t.points_per_lane == 2 ? 2 : 1;
    b.wpc = t.waves_per_cu > 0 ? t.waves_per_cu : 12;
    b.near_wpc = t.near_waves_per_cu > 0 ? t.near_waves_per_cu : 16;
    b.near_layout = t.near_layout > 0 ? t.near_layout : 2;
    // not in the fixed-scratch path, whose capacity the near buffer would take
    b.overlap = b.f32 && !b.s64 && b.near_layout != 1 && t.near_concurrent != 1 && !ctx->arena;
    b.count_branches = t.count_branches ? 1 : 0;
    b.ff = b.f32 && !b.s64 && WING_BF16 && t.far_field != 1;
    TRY(upload_stars(ctx, ws, s, n_stars, stars, &b.N, &b.nwl));
    b.tile_w = b.f32 ? 32 * b.NB : WAVE * b.P64;
    b.ntiles = (int)((b.nwl + b.tile_w - 1) / b.tile_w);
    if (b.ff) {
        b.n_sub = (b.nwl + FF_SUB - 1) / FF_SUB;
        b.nsuper = (int)((b.n_sub + FF_NS - 1) / FF_NS);
    }
    // scratch of the multi-block scans (tile lists, adjoint items and segments)
    TRY(grow(ctx, ws.scan_part, sizeof(int64_t) *
                                    (size_t)(scan_blocks(std::max<int64_t>({(int64_t)b.ntiles * N_CLASSES * 2, b.N,
                                                                             (int64_t)n_stars})) + 1)));
    if (b.N == 0) return KORG_OK;  // src/line_absorption.jl:35-37
    const int64_t N = b.N;
    b.nt_near = (int)((b.nwl + CT - 1) / CT);
    b.n_groups = (L + F64_GROUP - 1) / F64_GROUP;
    b.Ls = b.n_groups * F64_GROUP;        // rec64 stride: whole Float64 groups
    b.n_lblocks = (L + WAVE - 1) / WAVE;  // 64-layer blocks (Float32 records)
    b.nblocks = (int)((N + PART_BLOCK - 1) / PART_BLOCK);
    TRY(grow(ctx, ws.meta, sizeof(LineMeta) * N));
    TRY(grow(ctx, ws.lpre, sizeof(LinePre) * N));
    TRY(grow(ctx, ws.lwin, sizeof(LineWin) * N));
    TRY(grow(ctx, ws.rec64, sizeof(double4) * N * b.Ls));
    if (b.f32)
        TRY(grow(ctx, ws.recM, (b.s64 ? sizeof(double4) : sizeof(float4)) * (size_t)N * b.n_lblocks * WAVE));
    if (b.overlap) TRY(grow(ctx, ws.nbuf, sizeof(float) * (size_t)L * (size_t)b.nwl));
    TRY(grow(ctx, ws.tab, sizeof(LayerTab) * (size_t)n_stars * L));
    TRY(grow(ctx, ws.counters, sizeof(unsigned long long) * N_COUNTERS));
    TRY(grow(ctx, ws.blk, sizeof(int) * (size_t)b.nblocks * N_CLASSES));
    TRY(grow(ctx, ws.nblk, sizeof(int) * (size_t)b.nblocks * N_CLASSES));
    TRY(grow(ctx, ws.cstart, sizeof(int) * (N_CLASSES + 1)));
    TRY(grow(ctx, ws.cls_cand, sizeof(Cand) * N));
    TRY(grow(ctx, ws.cls_rng, sizeof(int2) * N));
    TRY(grow(ctx, ws.cls_line, sizeof(int) * N));
    TRY(grow(ctx, ws.cls_jc, sizeof(int) * N));
    TRY(grow(ctx, ws.ranges, sizeof(int2) * (size_t)b.ntiles * N_CLASSES));
    if (b.f32) {
        // the wing tiles' lists, then (far field) the super-tiles' lists by their tiles' slots
        const size_t nslots = (size_t)b.ntiles * N_CLASSES * (b.ff ? 2 : 1);
        TRY(grow(ctx, ws.tc_count, sizeof(int) * nslots));
        TRY(grow(ctx, ws.tc_start, sizeof(int64_t) * (nslots + 1)));
        if (b.ff) {
            TRY(grow(ctx, ws.ffrun, sizeof(int4) * (size_t)N));
            TRY(grow(ctx, ws.ffkey, sizeof(int4) * (size_t)N));
            TRY(grow(ctx, ws.ffgeo, sizeof(FfGeo) * (size_t)b.n_sub));
            TRY(grow(ctx, ws.fft, sizeof(float) * (size_t)b.nsuper * b.n_lblocks * FF_NS * WAVE * FF_TERMS));
        }
        TRY(grow(ctx, ws.ncstart, sizeof(int) * (N_CLASSES + 1)));
        TRY(grow(ctx, ws.near_cand, sizeof(Cand) * N));
        TRY(grow(ctx, ws.near_rng, sizeof(int2) * N));
        TRY(grow(ctx, ws.near_line, sizeof(int) * N));
        TRY(grow(ctx, ws.near_jc, sizeof(int) * N));
        TRY(grow(ctx, ws.near_ranges, sizeof(int2) * (size_t)b.nt_near * N_CLASSES));
    }
    return KORG_OK;
}

// Per-layer tables (k_layers, which also clears the counters) and the per-line setup (k_setup):
// windows, zones, the Float64 profile records and, for the Float32 form, the MFMA coefficient
// blocks.  `timed` records ev[0], ev[1].
int enqueue_setup(korg_ctx* ctx, Workspace& ws, const Batch& b, hipStream_t s, bool timed) {
    const int64_t N = b.N;
    auto* counters = (unsigned long long*)ws.counters.p;
    if (timed) HIP_TRY(ctx, hipEventRecord(ctx->ev[0], s));
    hipLaunchKernelGGL(k_layers, dim3(b.n_stars), dim3(64), 0, s, b.n_stars, b.L,
                       (StarDev*)ws.stars.p, (LayerTab*)ws.tab.p, counters);
    SetupArgs sa;
    sa.n_lines = N;
    sa.L = b.L;
    sa.Ls = b.Ls;
    sa.n_lblocks = b.n_lblocks;
    sa.n_stars = b.n_stars;
    sa.stars = (const StarDev*)ws.stars.p;
    sa.n_species = b.sp.n_species;
    sa.mass = b.sp.mass;
    sa.is_mol = b.sp.is_molecule;
    sa.tab = (const LayerTab*)ws.tab.p;
    sa.meta = (LineMeta*)ws.meta.p;
    sa.rec64 = (double4*)ws.rec64.p;
    sa.recM = (float4*)ws.recM.p;
    sa.recM64 = b.s64 ? (double4*)ws.recM.p : nullptr;
    sa.counters = counters;
    sa.azone = b.adj ? (int2*)ws.azone.p : nullptr;
    sa.azq = b.adj && b.L <= WAVE ? (int4*)ws.azq.p : nullptr;
    sa.geo = b.ff ? (const FfGeo*)ws.ffgeo.p : nullptr;
    sa.n_sub = b.n_sub;
    sa.ffrun = b.ff ? (int4*)ws.ffrun.p : nullptr;
    sa.ffrun_adj = b.adj && b.ff ? (int4*)ws.ffrun_adj.p : nullptr;
    if (b.ff)  // the sub-tiles' geometry, which the setup's far-field runs read
        hipLaunchKernelGGL(k_ff_geo, dim3((unsigned)((b.n_sub + 255) / 256)), dim3(256), 0, s,
                           b.n_sub, b.n_stars, (StarDev*)ws.stars.p, (FfGeo*)ws.ffgeo.p);
    sa.pre = (LinePre*)ws.lpre.p;
    sa.win = (LineWin*)ws.lwin.p;
    const dim3 lgrid((unsigned)((N + 255) / 256));                  // one lane per line
    const dim3 sgrid((unsigned)((N + 3) / 4)), sblock(256);         // one wavefront per line
    hipLaunchKernelGGL(k_setup_pre, lgrid, sblock, 0, s, sa);
    if (b.f32) {
        hipLaunchKernelGGL(k_setup<true>, sgrid, sblock, 0, s, sa);
        hipLaunchKernelGGL(k_setup_post<true>, lgrid, sblock, 0, s, sa);
        if (b.count_branches) hipLaunchKernelGGL(k_branch_count<true>, sgrid, sblock, 0, s, sa);
    } else {
        hipLaunchKernelGGL(k_setup<false>, sgrid, sblock, 0, s, sa);
        hipLaunchKernelGGL(k_setup_post<false>, lgrid, sblock, 0, s, sa);
        if (b.count_branches) hipLaunchKernelGGL(k_branch_count<false>, sgrid, sblock, 0, s, sa);
    }
    HIP_TRY(ctx, hipGetLastError());
    if (timed) HIP_TRY(ctx, hipEventRecord(ctx->ev[1], s));
    return KORG_OK;
}

// First half of the pipeline on `s`: setup, the class partitions, and (Float32 form) the per-tile
// list counts and their scan, which flags FLAG_CAPACITY when the lists exceed tile_cap entries.
// `timed` records ev[0], ev[1], ev[5].
int launch_near(korg_ctx* ctx, Workspace& ws, const Batch& b, hipStream_t s, float* nbuf);

// The Float32 near zone on the context's second stream, into ws.nbuf (b.overlap): forked from s
// once its partition is on s, joined by enqueue_back before its sums are added.  `timed` records
// ev[7], ev[8].
int fork_near(korg_ctx* ctx, Workspace& ws, const Batch& b, hipStream_t s, bool timed) {
    HIP_TRY(ctx, hipEventRecord(ctx->fork_ev, s));
    HIP_TRY(ctx, hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0));
    if (timed) HIP_TRY(ctx, hipEventRecord(ctx->ev[7], ctx->side));
    TRY(launch_near(ctx, ws, b, ctx->side, (float*)ws.nbuf.p));
    if (timed) HIP_TRY(ctx, hipEventRecord(ctx->ev[8], ctx->side));
    HIP_TRY(ctx, hipEventRecord(ctx->join_ev, ctx->side));
    return KORG_OK;
}

// early_near: with b.overlap, the near zone is forked as soon as its partition is done (beside the
// far wing's list building as well as the far wing); the caller then runs enqueue_back with the
// same flag, which joins it.
int enqueue_front(korg_ctx* ctx, Workspace& ws, const Batch& b, hipStream_t s, int64_t tile_cap,
                  bool timed, bool early_near = false) {
    TRY(enqueue_setup(ctx, ws, b, s, timed));
    const int64_t N = b.N;
    auto* counters = (unsigned long long*)ws.counters.p;

    hipLaunchKernelGGL(k_count, dim3(b.nblocks), dim3(PART_BLOCK), 0, s, N, b.L, 32 * b.NB,
                       (const LineMeta*)ws.meta.p, (int*)ws.blk.p, (int*)ws.nblk.p, counters);
    hipLaunchKernelGGL(k_scan<N_CLASSES>, dim3(1), dim3(1024), 0, s, b.nblocks, (int*)ws.blk.p,
                       (int*)ws.cstart.p);
    hipLaunchKernelGGL(k_scatter, dim3(b.nblocks), dim3(PART_BLOCK), 0, s, N,
                       (const LineMeta*)ws.meta.p, (const int*)ws.blk.p, (const int*)ws.cstart.p,
                       (Cand*)ws.cls_cand.p, (int*)ws.cls_jc.p, (int2*)ws.cls_rng.p,
                       (int*)ws.cls_line.p, (int)KEY_WINDOW,
                       b.ff ? (const int4*)ws.ffrun.p : nullptr, (int4*)ws.ffkey.p);
    hipLaunchKernelGGL(k_ranges, dim3((b.ntiles * N_CLASSES + 255) / 256), dim3(256), 0, s,
                       b.ntiles, b.tile_w, b.nwl, (const int*)ws.cstart.p, (const int*)ws.cls_jc.p,
                       (int2*)ws.ranges.p);
    if (b.f32) {
        // near-zone partition (by reach class) and its per-tile candidate ranges
        hipLaunchKernelGGL(k_scan<N_CLASSES>, dim3(1), dim3(1024), 0, s, b.nblocks, (int*)ws.nblk.p,
                           (int*)ws.ncstart.p);
        hipLaunchKernelGGL(k_scatter, dim3(b.nblocks), dim3(PART_BLOCK), 0, s, N,
                           (const LineMeta*)ws.meta.p, (const int*)ws.nblk.p,
                           (const int*)ws.ncstart.p, (Cand*)ws.near_cand.p, (int*)ws.near_jc.p,
                           (int2*)ws.near_rng.p, (int*)ws.near_line.p, (int)KEY_NEAR,
                           (const int4*)nullptr, (int4*)nullptr);
        hipLaunchKernelGGL(k_ranges, dim3((b.nt_near * N_CLASSES + 255) / 256), dim3(256), 0, s,
                           b.nt_near, CT, b.nwl, (const int*)ws.ncstart.p,
                           (const int*)ws.near_jc.p, (int2*)ws.near_ranges.p);
        HIP_TRY(ctx, hipGetLastError());
        if (early_near && b.overlap) TRY(fork_near(ctx, ws, b, s, timed));
        // exact per-tile line lists of the far-wing kernels (the wing tiles', then with the far
        // field the super-tiles'): count and one scan over both
        const int64_t ntc = (int64_t)b.ntiles * N_CLASSES;
        const int4* ffkey = b.ff ? (const int4*)ws.ffkey.p : nullptr;
        hipLaunchKernelGGL(k_tile_lists, dim3((unsigned)ntc), dim3(WAVE), 0, s, b.ntiles, b.tile_w,
                           b.nwl, (const int2*)ws.ranges.p, (const int2*)ws.cls_rng.p,
                           (const int*)ws.cls_line.p, ffkey, (const int64_t*)nullptr,
                           (int*)ws.tc_count.p, (int*)nullptr, counters);
        enqueue_scan(s, ntc * (b.ff ? 2 : 1), (const int*)ws.tc_count.p, (int64_t*)ws.tc_start.p,
                     tile_cap, counters, 13, (int64_t*)ws.scan_part.p);
    }
    HIP_TRY(ctx, hipGetLastError());
    if (timed && b.f32) HIP_TRY(ctx, hipEventRecord(ctx->ev[5], s));
    return KORG_OK;
}

template <int NB, int WW>
void launch_wing(const ProfileArgs& pa, dim3 grid, hipStream_t s) {
    if (WING_BF16) {
        if (pa.alpha_f32) hipLaunchKernelGGL((k_wing_bf16<NB, WW, float>), grid, dim3(WAVE * WW), 0, s, pa);
        else hipLaunchKernelGGL((k_wing_bf16<NB, WW, double>), grid, dim3(WAVE * WW), 0, s, pa);
        return;
    }
    if (pa.alpha_f32) hipLaunchKernelGGL((k_wing_mfma<NB, WW, float>), grid, dim3(WAVE * WW), 0, s, pa);
    else hipLaunchKernelGGL((k_wing_mfma<NB, WW, double>), grid, dim3(WAVE * WW), 0, s, pa);
}

// The Float32 (or Float64 series) near-zone kernel on `s`; nbuf: store its sums there instead of
// adding them into alpha.
int launch_near(korg_ctx* ctx, Workspace& ws, const Batch& b, hipStream_t s, float* nbuf) {
    NearArgs na;
    na.n_wl = b.nwl;
    na.n_lines = b.N;
    na.L = b.L;
    na.Ls = b.Ls;
    na.n_lblocks = b.n_lblocks;
    na.ntiles = b.nt_near;
    na.n_items = (int64_t)na.ntiles * b.n_lblocks;
    na.stars = (const StarDev*)ws.stars.p;
    na.n_stars = b.n_stars;
    na.cands = (const Cand*)ws.near_cand.p;
    na.ranges = (const int2*)ws.near_ranges.p;
    na.rec64 = (const double4*)ws.rec64.p;
    na.alpha_f32 = b.alpha_f32;
    na.counters = (unsigned long long*)ws.counters.p;
    na.nbuf = nbuf;
    // beside the far wing (nbuf set): one item per wavefront, in a grid of n_items
    na.one_item = nbuf != nullptr && b.near_layout != 1 ? 1 : 0;
    int64_t cwaves = na.one_item ? na.n_items : (int64_t)ctx->n_cu * b.near_wpc;
    if (cwaves > na.n_items) cwaves = na.n_items;
    if (b.near_layout == 1 && !b.s64)
        hipLaunchKernelGGL(k_near, dim3((unsigned)cwaves), dim3(WAVE), 0, s, na);
    else if (b.s64 && std::min(b.L, WAVE) <= 56)
        hipLaunchKernelGGL((k_near2<7, true>), dim3((unsigned)cwaves), dim3(WAVE), 0, s, na);
    else if (b.s64)
        hipLaunchKernelGGL((k_near2<8, true>), dim3((unsigned)cwaves), dim3(WAVE), 0, s, na);
    else if (std::min(b.L, WAVE) <= 56)
        hipLaunchKernelGGL((k_near2<7, false>), dim3((unsigned)cwaves), dim3(WAVE), 0, s, na);
    else hipLaunchKernelGGL((k_near2<8, false>), dim3((unsigned)cwaves), dim3(WAVE), 0, s, na);
    HIP_TRY(ctx, hipGetLastError());
    return KORG_OK;
}

// Second half on `s`: (Float32 form) the per-tile list fill, then the profile kernels that add
// into alpha.  `timed` records ev[6], ev[2], ev[4], ev[3].
int enqueue_back(korg_ctx* ctx, Workspace& ws, const Batch& b, hipStream_t s, bool timed,
                 bool early_near = false) {
    auto* counters = (unsigned long long*)ws.counters.p;
    if (b.f32) {
        if (timed) HIP_TRY(ctx, hipEventRecord(ctx->ev[6], s));
        const int64_t ntc = (int64_t)b.ntiles * N_CLASSES;
        const int4* ffkey = b.ff ? (const int4*)ws.ffkey.p : nullptr;
        hipLaunchKernelGGL(k_tile_lists, dim3((unsigned)ntc), dim3(WAVE), 0, s, b.ntiles, b.tile_w,
                           b.nwl, (const int2*)ws.ranges.p, (const int2*)ws.cls_rng.p,
                           (const int*)ws.cls_line.p, ffkey, (const int64_t*)ws.tc_start.p,
                           (int*)nullptr, (int*)ws.tile_lines.p, counters);
        HIP_TRY(ctx, hipGetLastError());
    }
    if (timed) HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
    ProfileArgs pa;
    pa.n_wl = b.nwl;
    pa.L = b.L;
    pa.Ls = b.Ls;
    pa.n_lblocks = b.f32 ? b.n_lblocks : b.n_groups;
    pa.ntiles = b.ntiles;
    pa.tile_w = b.tile_w;
    pa.n_items = (int64_t)b.ntiles * pa.n_lblocks;
    pa.stars = (const StarDev*)ws.stars.p;
    pa.n_stars = b.n_stars;
    pa.rec64 = (const double4*)ws.rec64.p;
    pa.recM = (const float4*)ws.recM.p;
    pa.recM64 = b.s64 ? (const double4*)ws.recM.p : nullptr;
    pa.cands = (const Cand*)ws.cls_cand.p;
    pa.meta = (const LineMeta*)ws.meta.p;
    pa.tile_start = (const int64_t*)ws.tc_start.p;
    pa.tile_lines = (const int*)ws.tile_lines.p;
    pa.ranges = (const int2*)ws.ranges.p;
    pa.geo = b.ff ? (const FfGeo*)ws.ffgeo.p : nullptr;
    pa.ffrun = b.ff ? (const int4*)ws.ffrun.p : nullptr;
    pa.fft = b.ff ? (const float*)ws.fft.p : nullptr;
    pa.n_sub = b.n_sub;
    pa.alpha_f32 = b.alpha_f32;
    pa.counters = counters;
    if (!b.f32) {
        int64_t nwaves = (int64_t)ctx->n_cu * b.wpc;
        if (nwaves > pa.n_items) nwaves = pa.n_items;
        const dim3 grid((unsigned)nwaves);
        if (b.P64 == 2) hipLaunchKernelGGL((k_profile_f64<F64_GROUP, 2>), grid, dim3(WAVE), 0, s, pa);
        else hipLaunchKernelGGL((k_profile_f64<F64_GROUP, 1>), grid, dim3(WAVE), 0, s, pa);
        HIP_TRY(ctx, hipGetLastError());
        if (timed) HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
    } else {
        // the near zone on the side stream, into ws.nbuf, while the far wing runs on s (unless
        // enqueue_front forked it already)
        if (b.overlap && !early_near) TRY(fork_near(ctx, ws, b, s, timed));
        if (b.ff) {
            // the far-field sums of every super-tile, which the wing's tiles start from
            FarArgs fa;
            fa.n_lblocks = b.n_lblocks;
            fa.n_items = (int64_t)b.nsuper * b.n_lblocks;
            fa.n_sub = b.n_sub;
            fa.base_slot = (int64_t)b.ntiles * N_CLASSES;
            fa.ntiles = b.ntiles;
            fa.per = FF_NS * FF_SUB / b.tile_w;
            fa.geo = (const FfGeo*)ws.ffgeo.p;
            fa.meta = (const LineMeta*)ws.meta.p;
            fa.ffrun = (const int4*)ws.ffrun.p;
            fa.recM = (const float4*)ws.recM.p;
            fa.tile_start = (const int64_t*)ws.tc_start.p;
            fa.tile_lines = (const int*)ws.tile_lines.p;
            fa.fft = (float*)ws.fft.p;
            fa.counters = counters;
            // two waves per SIMD; when the items are too few to occupy them, four waves share
            // each item
            const int64_t slots = (int64_t)ctx->n_cu * 8;
            if (fa.n_items < 2 * slots) {
                const int64_t nwg = std::min<int64_t>(slots / 4, fa.n_items);
                hipLaunchKernelGGL(k_wing_ff<4>, dim3((unsigned)nwg), dim3(4 * WAVE), 0, s, fa);
            } else {
                hipLaunchKernelGGL(k_wing_ff<1>, dim3((unsigned)slots), dim3(WAVE), 0, s, fa);
            }
            HIP_TRY(ctx, hipGetLastError());
            if (timed) HIP_TRY(ctx, hipEventRecord(ctx->ev[9], s));
        }
        int64_t nwg = (int64_t)ctx->n_cu * ((b.wpc + b.WW - 1) / b.WW);
        if (nwg > pa.n_items) nwg = pa.n_items;
        const dim3 wgrid((unsigned)nwg);
        if (b.s64) {
            int64_t nw = (int64_t)ctx->n_cu * 16;  // four waves per SIMD
            if (nw > pa.n_items) nw = pa.n_items;
            if (pa.alpha_f32) hipLaunchKernelGGL(k_wing_f64<float>, dim3((unsigned)nw), dim3(WAVE), 0, s, pa);
            else hipLaunchKernelGGL(k_wing_f64<double>, dim3((unsigned)nw), dim3(WAVE), 0, s, pa);
        } else switch (b.NB * 8 + b.WW) {
            case 1 * 8 + 1: launch_wing<1, 1>(pa, wgrid, s); break;
            case 1 * 8 + 2: launch_wing<1, 2>(pa, wgrid, s); break;
            case 1 * 8 + 4: launch_wing<1, 4>(pa, wgrid, s); break;
            case 2 * 8 + 1: launch_wing<2, 1>(pa, wgrid, s); break;
            case 2 * 8 + 2: launch_wing<2, 2>(pa, wgrid, s); break;
            case 2 * 8 + 4: launch_wing<2, 4>(pa, wgrid, s); break;
            case 4 * 8 + 1: launch_wing<4, 1>(pa, wgrid, s); break;
            case 4 * 8 + 2: launch_wing<4, 2>(pa, wgrid, s); break;
            default:
                return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "no far-wing kernel for %d-point "
                               "tiles with %d waves per tile", 32 * b.NB, b.WW);
        }
        HIP_TRY(ctx, hipGetLastError());
        if (timed) HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
        if (b.overlap) {
            // the near zone's sums (launched beside the wing, below) are added now
            HIP_TRY(ctx, hipStreamWaitEvent(s, ctx->join_ev, 0));
            const int64_t ne = (int64_t)b.L * b.nwl;
            hipLaunchKernelGGL(k_add_near, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s,
                               b.n_stars, b.L, b.nwl, (const StarDev*)ws.stars.p,
                               (const float*)ws.nbuf.p, (int)b.alpha_f32,
                               (const unsigned long long*)counters);
        } else {
            TRY(launch_near(ctx, ws, b, s, nullptr));
        }
        HIP_TRY(ctx, hipGetLastError());
    }
    if (timed) HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
    return KORG_OK;
}

// Reads the per-tile list total (after enqueue_front on s), waits for it, and grows ws.tile_lines
// to `headroom` times it.
int size_tile_lists(korg_ctx* ctx, Workspace& ws, const Batch& b, hipStream_t s, double headroom,
                    int64_t* total_out) {
    const int64_t ntc = (int64_t)b.ntiles * N_CLASSES * (b.ff ? 2 : 1);
    int64_t total = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&total, (int64_t*)ws.tc_start.p + ntc, sizeof total,
                                hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (total > (int64_t)INT32_MAX)
        return set_err(ctx, KORG_ERR_UNSUPPORTED,
                       "more than 2^31 (line, tile) pairs in one call; split the line list");
    int64_t cap = (int64_t)((double)total * headroom) + (headroom > 1.0 ? 4096 : 0);
    if (cap < total) cap = total;
    if (cap > (int64_t)INT32_MAX) cap = INT32_MAX;
    TRY(grow(ctx, ws.tile_lines, sizeof(int) * (size_t)(cap > 0 ? cap : 1)));
    ws.tile_cap = cap;
    *total_out = total;
    return KORG_OK;
}

// Device address of `scratch` on each device, resolved once per device with that device current.
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;

// Resolves the fixed scratch of ctx's device (ctx->device must be current).  A failure is kept in
// ctx->scratch_err and reported by the allocation-free entry point only.
void resolve_scratch(korg_ctx* ctx) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= ctx->device) g_scratch_ptr.resize(ctx->device + 1, nullptr);
    if (!g_scratch_ptr[ctx->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            ctx->scratch_err = std::string("hipGetSymbolAddress(scratch) failed: ") +
                               hipGetErrorString(e);
            return;
        }
        g_scratch_ptr[ctx->device] = p;
    }
    ctx->scratch_ptr = g_scratch_ptr[ctx->device];
}

// Clears a context's arena when it goes out of scope.
struct ArenaScope {
    korg_ctx* ctx;
    Arena arena;
    ArenaScope(korg_ctx* c, void* base, size_t cap) : ctx(c) {
        arena.base = (char*)base;
        arena.cap = cap;
        ctx->arena = &arena;
    }
    ~ArenaScope() { ctx->arena = nullptr; }
};

// The eager device pipeline over a batch of stars (every pointer a device pointer; one alpha
// dtype and one layer count for the batch), on the context's stream.
//   fixed = false: on ctx->ws (grown with hipMalloc); the host reads the per-tile list total
//                  between the halves and sizes the lists exactly.
//   fixed = true:  on ctx->ws_fixed, carved from the device's fixed scratch, with no allocation
//                  and no host round trip: the per-tile lists take what the other buffers leave of
//                  the scratch, and when they would need more the device raises FLAG_CAPACITY and
//                  no kernel writes alpha (reported by finish_stats).
int run_batch(korg_ctx* ctx, int n_stars, const StarHost* stars, const korg_species* sp, int L,
              int alpha_dtype, int precision, bool fixed = false) {
    hipStream_t s = ctx->stream;
    Workspace& ws = fixed ? ctx->ws_fixed : ctx->ws;
    Batch b;
    std::memset(&ctx->stats, 0, sizeof ctx->stats);
    ctx->last_n_lines = 0;
    ctx->last_fixed = fixed;
    ctx->last_ragged = false;
    if (!fixed) {
        TRY(plan_batch(ctx, ctx->tuning, n_stars, stars, sp, L, alpha_dtype, precision, ws, s, b));
    } else {
        ws.forget();
        ArenaScope as(ctx, ctx->scratch_ptr, (size_t)KORG_NOALLOC_SCRATCH_BYTES);
        TRY(plan_batch(ctx, ctx->tuning, n_stars, stars, sp, L, alpha_dtype, precision, ws, s, b));
        if (b.N > 0 && b.f32) {
            const size_t off = (as.arena.off + 255) & ~(size_t)255;
            const int64_t cap = std::min<int64_t>(
                off < as.arena.cap ? (int64_t)((as.arena.cap - off) / sizeof(int)) : 0, INT32_MAX);
            if (cap < 1)
                return set_err(ctx, KORG_ERR_CAPACITY, "scratch exhausted before the tile lists");
            TRY(grow(ctx, ws.tile_lines, sizeof(int) * (size_t)cap));
            ws.tile_cap = cap;
        }
    }
    ctx->last_n_lines = b.N;
    ctx->last_precision = precision;
    ctx->last_n_lblocks = b.n_lblocks;
    ctx->last_nb = b.NB;
    ctx->last_ff = b.ff;
    ctx->last_ntiles = b.ntiles;
    ctx->last_sized = b.f32 && !fixed;
    ctx->last_overlap = b.overlap;
    if (b.N == 0) return KORG_OK;
    TRY(enqueue_front(ctx, ws, b, s, fixed ? ws.tile_cap : INT64_MAX, true, true));
    if (b.f32 && !fixed) {
        int64_t total = 0;
        TRY(size_tile_lists(ctx, ws, b, s, 1.0, &total));
    }
    return enqueue_back(ctx, ws, b, s, true, true);
}

// One line_absorption! call on device pointers: a batch of one star.
int run_device(korg_ctx* ctx, void* alpha, int alpha_dtype, const korg_grid* g,
               const korg_lines* ln, const korg_species* sp, const korg_layers* ly,
               const korg_continuum* cn, double cutoff, int precision) {
    const StarHost one = {alpha, *g, *ln, *ly, *cn, cutoff};
    return run_batch(ctx, 1, &one, sp, ly->n_layers, alpha_dtype, precision);
}

// Counters of a finished forward call into ctx->stats; an error for a raised flag.
int stats_from_counters(korg_ctx* ctx, const unsigned long long* c, bool f32, int n_lblocks, int nb,
                        bool s64 = false, bool ff = false, int ntiles = 0) {
    if (c[3] & FLAG_SPECIES)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT,
                       "a line's species index is outside [0, n_species); alpha was left unchanged");
    if (c[3] & FLAG_UNSORTED)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT,
                       "the line list is not sorted by wavelength; alpha was left unchanged");
    if (c[3] & FLAG_CAPACITY)
        return set_err(ctx, KORG_ERR_CAPACITY,
                       "per-tile line lists exceeded their buffer; alpha was left unchanged");
    ctx->stats.n_evals = (double)c[1];
    ctx->stats.n_lines_used = (int64_t)c[2];
    for (int k = 0; k < 4; ++k) ctx->stats.n_evals_branch[k] = (double)c[6 + k];
    ctx->stats.n_evals_near = f32 ? (double)c[10] : 0.0;
    // far-wing MFMAs: bfloat16, three per (group of 4 lines, 32-point block, 32-layer half),
    // counted per line-tile entry of the lists (groups rounded up are not counted); with the far
    // field, three per (group of 4 lines, 32-column block, 32-layer half) of a super-tile's list
    // (counters[25] of the c[13] entries) and three per (32-point block, 32-layer half) of every
    // tile for its start from T; Float32, two per 32-point block and K step; the Float64 series
    // form, 8 per (line, 32-point tile, 64-layer block)
    const double far_entries = ff ? (double)c[25] : 0.0;
    ctx->stats.n_wing_mfma = !f32 ? 0.0 : s64 ? (double)c[13] * n_lblocks * 8.0
                             : WING_BF16 ? ((double)c[13] - far_entries) * n_lblocks * 1.5 * nb +
                                               far_entries * n_lblocks * 1.5 * FF_CB +
                                               (ff ? (double)ntiles * n_lblocks * 6.0 * nb : 0.0)
                                         : (double)c[11] * n_lblocks;
    ctx->stats.n_far_entries = far_entries;
    ctx->stats.n_evals_tail = f32 ? (double)c[12] : 0.0;
    ctx->stats.n_evals_near_A = f32 ? (double)c[15] : 0.0;
    return KORG_OK;
}

float elapsed(korg_ctx* ctx, int a, int b) {
    float t = 0;
    if (hipEventElapsedTime(&t, ctx->ev[a], ctx->ev[b]) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0f;
    }
    return t;
}

int finish_stats(korg_ctx* ctx) {
    if (ctx->last_n_lines == 0) return KORG_OK;
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev[3]));
    unsigned long long c[N_COUNTERS];
    const Workspace& ws = ctx->last_fixed ? ctx->ws_fixed : ctx->ws;
    HIP_TRY(ctx, hipMemcpy(c, ws.counters.p, sizeof c, hipMemcpyDeviceToHost));
    const int rc = stats_from_counters(ctx, c, ctx->last_precision != KORG_PRECISION_FP64,
                                       ctx->last_n_lblocks, ctx->last_nb,
                                       ctx->last_precision == KORG_PRECISION_FP64_SERIES,
                                       ctx->last_ff, ctx->last_ntiles);
    if (rc == KORG_ERR_CAPACITY && ctx->last_fixed)
        return set_err(ctx, KORG_ERR_CAPACITY,
                       "the per-tile line lists need %llu entries, more than the %lld the fixed "
                       "scratch (%lld bytes) has left for them; alpha was left unchanged",
                       c[13], (long long)ctx->ws_fixed.tile_cap,
                       (long long)KORG_NOALLOC_SCRATCH_BYTES);
    TRY(rc);
    // device time: the host's sizing of the tile lists (ev[5] -> ev[6]) is left out
    const float gap = ctx->last_sized ? elapsed(ctx, 5, 6) : 0.0f;
    ctx->stats.setup_ms = elapsed(ctx, 0, 1);
    ctx->stats.bin_ms = elapsed(ctx, 1, 2) - gap;
    ctx->stats.profile_ms = elapsed(ctx, 2, 4);
    ctx->stats.far_ms = ctx->last_ff ? elapsed(ctx, 2, 9) : 0.0;
    ctx->stats.core_ms = ctx->last_overlap ? elapsed(ctx, 7, 8) : elapsed(ctx, 4, 3);
    ctx->stats.total_ms = elapsed(ctx, 0, 3) - gap;
    return KORG_OK;
}

// Buffers of the adjoint over the batch b (grown before any capture; enqueue functions allocate
// nothing).
int prepare_vjp(korg_ctx* ctx, Workspace& ws, const Batch& b) {
    TRY(grow(ctx, ws.vjp_cnt, sizeof(int) * (size_t)(b.N > 0 ? b.N : 1)));
    TRY(grow(ctx, ws.vjp_start, sizeof(int64_t) * (size_t)(b.N + 1)));
    TRY(grow(ctx, ws.seg_cnt, sizeof(int) * (size_t)b.n_stars));
    TRY(grow(ctx, ws.seg_start, sizeof(int64_t) * (size_t)(b.n_stars + 1)));
    return KORG_OK;
}

// Adjoint counts on s (the setup already enqueued): items per line, their scan, segments per star,
// their scan, which raises FLAG_CAPACITY when the segments exceed seg_cap.
int enqueue_vjp_counts(korg_ctx* ctx, Workspace& ws, const Batch& b, hipStream_t s,
                       int64_t seg_cap) {
    auto* counters = (unsigned long long*)ws.counters.p;
    if (b.N > 0)
        hipLaunchKernelGGL(k_vjp_count, dim3((unsigned)((b.N + 255) / 256)), dim3(256), 0, s, b.N,
                           (const LineMeta*)ws.meta.p, b.adj ? (const int2*)ws.azone.p : nullptr,
                           (int*)ws.vjp_cnt.p);
    enqueue_scan(s, b.N, (const int*)ws.vjp_cnt.p, (int64_t*)ws.vjp_start.p, INT64_MAX, nullptr, 0,
                 (int64_t*)ws.scan_part.p);
    hipLaunchKernelGGL(k_vjp_segs, dim3((unsigned)((b.n_stars + 255) / 256)), dim3(256), 0, s,
                       b.n_stars, (const StarDev*)ws.stars.p, (const int64_t*)ws.vjp_start.p,
                       (int*)ws.seg_cnt.p);
    enqueue_scan(s, b.n_stars, (const int*)ws.seg_cnt.p, (int64_t*)ws.seg_start.p, seg_cap,
                 counters, 14, (int64_t*)ws.scan_part.p);
    HIP_TRY(ctx, hipGetLastError());
    return KORG_OK;
}

// Reads the segment total (after enqueue_vjp_counts on s), waits for it, and grows the partials to
// headroom times it.
int size_segments(korg_ctx* ctx, Workspace& ws, const Batch& b, hipStream_t s, double headroom) {
    int64_t total = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&total, (int64_t*)ws.seg_start.p + b.n_stars, sizeof total,
                                hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    int64_t cap = (int64_t)((double)total * headroom) + (headroom > 1.0 ? 64 : 0);
    if (cap < total) cap = total;
    const int NQ = 6 + b.sp.n_species;
    TRY(grow(ctx, ws.vjp_partial, sizeof(double) * (size_t)(cap > 0 ? cap : 1) * NQ * b.L));
    ws.seg_cap = cap;
    return KORG_OK;
}

// The adjoint's items and the per-star reduction into the StarDev outputs, on s.
int enqueue_vjp(korg_ctx* ctx, Workspace& ws, const Batch& b, hipStream_t s, int g_f32) {
    VjpArgs va;
    va.n_stars = b.n_stars;
    va.L = b.L;
    va.Ls = b.Ls;
    va.n_species = b.sp.n_species;
    va.stars = (const StarDev*)ws.stars.p;
    va.meta = (const LineMeta*)ws.meta.p;
    va.rec64 = (const double4*)ws.rec64.p;
    va.tab = (const LayerTab*)ws.tab.p;
    va.item_start = (const int64_t*)ws.vjp_start.p;
    va.seg_start = (const int64_t*)ws.seg_start.p;
    va.mass = b.sp.mass;
    va.is_mol = b.sp.is_molecule;
    va.g_f32 = g_f32;
    va.partial = (double*)ws.vjp_partial.p;
    va.seg_cap = ws.seg_cap;
    va.counters = (const unsigned long long*)ws.counters.p;
    va.n_lblocks = b.n_lblocks;
    va.azone = (const int2*)ws.azone.p;
    va.azq = b.L <= WAVE ? (const int4*)ws.azq.p : nullptr;
    va.mom = (const float4*)ws.recM.p;
    va.gexp = (const int*)ws.gexp.p;
    const int64_t grid = (int64_t)ctx->n_cu * 16;
    if (b.adj && b.L <= 56 && !KORG_ADJ_NEAR_F64)
        hipLaunchKernelGGL(k_vjp_near<7>, dim3((unsigned)grid), dim3(WAVE), 0, s, va);
    else if (b.adj && b.L <= WAVE && !KORG_ADJ_NEAR_F64)
        hipLaunchKernelGGL(k_vjp_near<8>, dim3((unsigned)grid), dim3(WAVE), 0, s, va);
    else if (b.adj) hipLaunchKernelGGL(k_vjp<true>, dim3((unsigned)grid), dim3(WAVE), 0, s, va);
    else hipLaunchKernelGGL(k_vjp<false>, dim3((unsigned)grid), dim3(WAVE), 0, s, va);
    HIP_TRY(ctx, hipGetLastError());
    const int NQ = 6 + b.sp.n_species;
    hipLaunchKernelGGL(k_vjp_reduce, dim3((unsigned)(NQ * b.L), (unsigned)b.n_stars), dim3(256), 0, s,
                       b.L, b.sp.n_species, (const StarDev*)ws.stars.p,
                       (const int64_t*)ws.seg_start.p, (const double*)ws.vjp_partial.p,
                       (const unsigned long long*)ws.counters.p);
    HIP_TRY(ctx, hipGetLastError());
    return KORG_OK;
}

// The adjoint pipeline of one star on device pointers: setup (windows and Float64 records), the
// item and segment counts, the items by segments, and the fixed-order reduction into the
// outputs `gq` (device pointers).
int run_vjp(korg_ctx* ctx, const void* G, int g_dtype, const korg_grid* g, const korg_lines* ln,
            const korg_species* sp, const korg_layers* ly, const korg_continuum* cn,
            double cutoff, double* const gq[7]) {
    const int64_t N = ln->n_lines;
    const int L = ly->n_layers;
    hipStream_t s = ctx->stream;
    ctx->last_n_lines = N;
    ctx->last_fixed = false;
    ctx->last_precision = KORG_PRECISION_FP64;
    ctx->last_n_lblocks = 0;
    ctx->last_sized = false;
    std::memset(&ctx->stats, 0, sizeof ctx->stats);
    if (N == 0) {  // no lines: every gradient is zero
        for (int q = 0; q < 7; ++q)
            if (gq[q])
                HIP_TRY(ctx, hipMemsetAsync(gq[q], 0, sizeof(double) * (size_t)(q < 6 ? L : sp->n_species * L), s));
        return KORG_OK;
    }
    Workspace& ws = ctx->ws;
    StarHost one = {nullptr, *g, *ln, *ly, *cn, cutoff};
    one.G = const_cast<void*>(G);
    for (int q = 0; q < 7; ++q) one.gq[q] = gq[q];
    Batch b;
    TRY(plan_batch(ctx, ctx->tuning, 1, &one, sp, L, KORG_DTYPE_F64, KORG_PRECISION_FP64, ws, s, b));
    TRY(prepare_vjp(ctx, ws, b));
    TRY(enqueue_setup(ctx, ws, b, s, true));
    TRY(enqueue_vjp_counts(ctx, ws, b, s, INT64_MAX));
    TRY(size_segments(ctx, ws, b, s, 1.0));
    unsigned long long flags = 0;
    HIP_TRY(ctx, hipMemcpy(&flags, (unsigned long long*)ws.counters.p + 3, sizeof flags,
                           hipMemcpyDeviceToHost));
    if (flags & FLAG_SPECIES)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "a line's species index is outside [0, n_species)");
    if (flags & FLAG_UNSORTED)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "the line list is not sorted by wavelength");
    HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
    TRY(enqueue_vjp(ctx, ws, b, s, g_dtype == KORG_DTYPE_F32));
    HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
    HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
    return KORG_OK;
}

// ---------------------------------------------------------------------------------------------
// fit step: forward, radiative transfer, chi^2 and its gradient through the adjoint
// ---------------------------------------------------------------------------------------------
struct FitHost {
    const double *tau_ref, *alpha_ref, *data, *sigma;
    double *flux, *chi2;
    korg_layer_grads grads;
};

// Points each star's G into one Float64 buffer of the workspace (sum over stars of L x n_wl).
int place_fit_G(korg_ctx* ctx, Workspace& ws, std::vector<StarHost>& hs, int L) {
    size_t total = 0;
    for (const StarHost& x : hs) total += (size_t)L * (size_t)x.grid.n_wl;
    TRY(grow(ctx, ws.fit_G, sizeof(double) * total));
    size_t off = 0;
    for (StarHost& x : hs) {
        x.G = (double*)ws.fit_G.p + off;
        off += (size_t)L * (size_t)x.grid.n_wl;
    }
    return KORG_OK;
}

int prepare_fit(korg_ctx* ctx, Workspace& ws, Batch& b) {
    b.adj = b.f32;
    TRY(prepare_vjp(ctx, ws, b));
    TRY(grow(ctx, ws.rt_tab, sizeof(RtTab) * (size_t)b.n_stars * b.L));
    TRY(grow(ctx, ws.chi2_part, sizeof(double) * (size_t)(b.nwl > 0 ? b.nwl : 1)));
    TRY(grow(ctx, ws.gexp, sizeof(int) * (size_t)b.n_stars * b.L));
    if (b.f32) {
        if ((b.N + ADJ_GROUP - 1) / ADJ_GROUP * b.n_lblocks > INT_MAX)
            return set_err(ctx, KORG_ERR_UNSUPPORTED, "too many lines x layer blocks for the adjoint");
        TRY(grow(ctx, ws.adj_blk, sizeof(int) * (size_t)(b.nblocks > 0 ? b.nblocks : 1) * ADJ_NC));
        TRY(grow(ctx, ws.adj_start, sizeof(int) * (ADJ_NC + 1)));
        TRY(grow(ctx, ws.adj_line, sizeof(int) * (size_t)(b.N > 0 ? b.N : 1)));
        TRY(grow(ctx, ws.azone, sizeof(int2) * (size_t)(b.N > 0 ? b.N : 1)));
        TRY(grow(ctx, ws.azq, sizeof(int4) * (size_t)(b.N > 0 ? b.N : 1)));
        TRY(grow(ctx, ws.adj_lead, sizeof(int) * (size_t)(b.N > 0 ? b.N : 1)));
        TRY(grow(ctx, ws.adj_groups, sizeof(int) * (size_t)(b.N > 0 ? b.N : 1)));
        TRY(grow(ctx, ws.adj_gstart, sizeof(int) * 2));
        TRY(grow(ctx, ws.g32, sizeof(float) * WAVE * (size_t)b.n_lblocks * (size_t)(b.nwl + MOM_PAD)));
        TRY(grow(ctx, ws.lam_s, sizeof(double) * (size_t)(b.nwl + MOM_PAD)));
        if (b.ff) {
            TRY(grow(ctx, ws.ffrun_adj, sizeof(int4) * (size_t)(b.N > 0 ? b.N : 1)));
            TRY(grow(ctx, ws.gff, sizeof(float) * WAVE * FF_TERMS * (size_t)b.n_lblocks *
                                      (size_t)(b.n_sub > 0 ? b.n_sub : 1)));
            TRY(grow(ctx, ws.gffs, sizeof(float) * WAVE * FF_TERMS * (size_t)b.n_lblocks *
                                       (size_t)(b.nsuper > 0 ? b.nsuper : 1)));
            TRY(grow(ctx, ws.ffsgeo, sizeof(FfGeo) * (size_t)(b.nsuper > 0 ? b.nsuper : 1)));
        }
    }
    return KORG_OK;
}

// Series form of the adjoint, first part (G of every star complete): G's scale and Float32 copy,
// the lines with a far zone grouped by window width, and their moments into ws.recM (the forward's
// coefficient blocks, which the step no longer needs).
int enqueue_moments(korg_ctx* ctx, Workspace& ws, const Batch& b, hipStream_t s) {
    const unsigned gblocks = (unsigned)((b.nwl / STAR_ALIGN + 3) / 4);
    hipLaunchKernelGGL(k_g_exp, dim3(gblocks), dim3(256), 0, s, b.n_stars, b.L, b.nwl,
                       (const StarDev*)ws.stars.p, (int*)ws.gexp.p);
    hipLaunchKernelGGL(k_g_pack, dim3(gblocks), dim3(256), 0, s, b.n_stars, b.L, b.n_lblocks, b.nwl,
                       (const StarDev*)ws.stars.p, (const int*)ws.gexp.p, (float*)ws.g32.p,
                       (double*)ws.lam_s.p);
    if (b.ff && b.n_sub > 0) {
        // the sub-tiles' moments of G for the far-field steps, and from them the whole super-tiles'
        hipLaunchKernelGGL(k_g_ffmom, dim3((unsigned)(b.n_sub * b.n_lblocks)), dim3(WAVE), 0, s,
                           b.n_stars, b.L, b.n_lblocks, b.n_sub, (const StarDev*)ws.stars.p,
                           (const FfGeo*)ws.ffgeo.p, (const int*)ws.gexp.p, (float*)ws.gff.p);
        hipLaunchKernelGGL(k_ff_sgeo, dim3((unsigned)((b.nsuper + 255) / 256)), dim3(256), 0, s,
                           (int64_t)b.nsuper, b.n_sub, b.n_stars, (const StarDev*)ws.stars.p,
                           (const FfGeo*)ws.ffgeo.p, (FfGeo*)ws.ffsgeo.p);
        hipLaunchKernelGGL(k_g_ffmom_super, dim3((unsigned)(b.nsuper * b.n_lblocks)), dim3(WAVE), 0,
                           s, b.n_lblocks, b.n_sub, (int64_t)b.nsuper, (const FfGeo*)ws.ffgeo.p,
                           (const FfGeo*)ws.ffsgeo.p, (const float*)ws.gff.p, (float*)ws.gffs.p);
    }
    const int4* azq = b.L <= WAVE ? (const int4*)ws.azq.p : nullptr;
    hipLaunchKernelGGL(k_adj_count, dim3(b.nblocks), dim3(PART_BLOCK), 0, s, b.N,
                       (const LineMeta*)ws.meta.p, (const int2*)ws.azone.p, azq, (int*)ws.adj_blk.p);
    hipLaunchKernelGGL(k_scan<ADJ_NC>, dim3(1), dim3(1024), 0, s, b.nblocks, (int*)ws.adj_blk.p,
                       (int*)ws.adj_start.p);
    hipLaunchKernelGGL(k_adj_scatter, dim3(b.nblocks), dim3(PART_BLOCK), 0, s, b.N,
                       (const LineMeta*)ws.meta.p, (const int2*)ws.azone.p, azq, (const int*)ws.adj_blk.p,
                       (const int*)ws.adj_start.p, (int*)ws.adj_line.p, (int*)ws.adj_lead.p);
    hipLaunchKernelGGL(k_lead_count, dim3(b.nblocks), dim3(PART_BLOCK), 0, s, b.N,
                       (const int*)ws.adj_lead.p, (int*)ws.adj_blk.p);
    hipLaunchKernelGGL(k_scan<1>, dim3(1), dim3(1024), 0, s, b.nblocks, (int*)ws.adj_blk.p,
                       (int*)ws.adj_gstart.p);
    hipLaunchKernelGGL(k_lead_scatter, dim3(b.nblocks), dim3(PART_BLOCK), 0, s, b.N,
                       (const int*)ws.adj_lead.p, (const int*)ws.adj_blk.p, (int*)ws.adj_groups.p);
    MomArgs ma;
    ma.n_stars = b.n_stars;
    ma.L = b.L;
    ma.n_lblocks = b.n_lblocks;
    ma.n_total = b.nwl;
    ma.stars = (const StarDev*)ws.stars.p;
    ma.meta = (const LineMeta*)ws.meta.p;
    ma.azone = (const int2*)ws.azone.p;
    ma.azq = b.L <= WAVE ? (const int4*)ws.azq.p : nullptr;
    ma.adj_line = (const int*)ws.adj_line.p;
    ma.adj_start = (const int*)ws.adj_start.p;
    ma.groups = (const int*)ws.adj_groups.p;
    ma.n_groups = (const int*)ws.adj_gstart.p + 1;
    ma.g32 = (const float*)ws.g32.p;
    ma.lam_s = (const double*)ws.lam_s.p;
    ma.mom = (float4*)ws.recM.p;
    ma.counters = (unsigned long long*)ws.counters.p;
    ma.ffrun = b.ff ? (const int4*)ws.ffrun_adj.p : nullptr;
    ma.geo = (const FfGeo*)ws.ffgeo.p;
    ma.gff = (const float*)ws.gff.p;
    ma.n_sub = b.n_sub;
    ma.sgeo = (const FfGeo*)ws.ffsgeo.p;
    ma.gffs = (const float*)ws.gffs.p;
    ma.n_super = b.nsuper;
    hipLaunchKernelGGL(k_vjp_moments, dim3((unsigned)(ctx->n_cu * 32)), dim3(WAVE), 0, s, ma);
    HIP_TRY(ctx, hipGetLastError());
    return KORG_OK;
}

// Radiative transfer, chi^2 and G of every star (alpha_lines complete), on s.
int enqueue_rt(korg_ctx* ctx, Workspace& ws, const Batch& b, hipStream_t s) {
    hipLaunchKernelGGL(k_rt_tab, dim3(b.n_stars), dim3(64), 0, s, b.n_stars, b.L,
                       (const StarDev*)ws.stars.p, (RtTab*)ws.rt_tab.p, (int*)ws.gexp.p, GEXP_NONE);
    RtArgs ra;
    ra.n_stars = b.n_stars;
    ra.L = b.L;
    ra.n_total = b.nwl;
    ra.stars = (const StarDev*)ws.stars.p;
    ra.tab = (const RtTab*)ws.rt_tab.p;
    ra.alpha_f32 = b.alpha_f32;
    ra.chi2_part = (double*)ws.chi2_part.p;
    ra.flux_only = 0;
    hipLaunchKernelGGL(k_rt_fit, dim3((unsigned)((b.nwl + 255) / 256)), dim3(256), 0, s, ra);
    hipLaunchKernelGGL(k_star_chi2, dim3(b.n_stars), dim3(256), 0, s, (const StarDev*)ws.stars.p,
                       (const double*)ws.chi2_part.p);
    HIP_TRY(ctx, hipGetLastError());
    return KORG_OK;
}

// The whole fit step on s with the capacities of ws (tile lists, adjoint segments): zero the
// alphas, the forward, radiative transfer, chi^2 and G, the adjoint.  With sizing, the host reads
// the list and segment totals in between and grows them (the direct form); without, the
// capacities are fixed and a shortfall raises FLAG_CAPACITY (the captured form).
int enqueue_fit(korg_ctx* ctx, Workspace& ws, const Batch& b, hipStream_t s, bool sizing,
                double headroom) {
    hipLaunchKernelGGL(k_zero_alpha, dim3((unsigned)std::min<int64_t>((b.nwl * b.L + 255) / 256, 1 << 20)),
                       dim3(256), 0, s, b.n_stars, b.L, b.nwl, (const StarDev*)ws.stars.p,
                       b.alpha_f32 ? 1 : 0);
    HIP_TRY(ctx, hipGetLastError());
    if (b.N > 0) {
        TRY(enqueue_front(ctx, ws, b, s, sizing ? INT64_MAX : ws.tile_cap, false, true));
        if (sizing && b.f32) {
            int64_t total = 0;
            TRY(size_tile_lists(ctx, ws, b, s, headroom, &total));
        }
        TRY(enqueue_back(ctx, ws, b, s, false, true));
    } else {
        hipLaunchKernelGGL(k_layers, dim3(b.n_stars), dim3(64), 0, s, b.n_stars, b.L,
                           (StarDev*)ws.stars.p, (LayerTab*)ws.tab.p,
                           (unsigned long long*)ws.counters.p);
    }
    TRY(enqueue_rt(ctx, ws, b, s));
    if (b.adj && b.N > 0) TRY(enqueue_moments(ctx, ws, b, s));
    TRY(enqueue_vjp_counts(ctx, ws, b, s, sizing ? INT64_MAX : ws.seg_cap));
    if (sizing) TRY(size_segments(ctx, ws, b, s, headroom));
    return enqueue_vjp(ctx, ws, b, s, 0);
}

// The stars of korg_continuum_batched / korg_flux_batched: what those read (the grid, the
// continuum on knots, alpha; for the transfer the temperatures and at least 2 layers).
int collect_rt_stars(korg_ctx* ctx, int64_t n_stars, const korg_star* stars, int alpha_dtype,
                     bool transfer, std::vector<StarHost>& hs, int* L_out) {
    if (n_stars < 1 || !stars) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "no stars");
    if (n_stars > 65535) return set_err(ctx, KORG_ERR_UNSUPPORTED, "at most 65535 stars per call");
    if (alpha_dtype != KORG_DTYPE_F64 && alpha_dtype != KORG_DTYPE_F32)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "unknown alpha dtype");
    const int L = stars[0].layers.n_layers;
    hs.resize((size_t)n_stars);
    for (int64_t k = 0; k < n_stars; ++k) {
        const korg_star& x = stars[k];
        if (x.layers.n_layers < (transfer ? 2 : 1))
            return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "star %lld: %d layers (at least %d)",
                           (long long)k, x.layers.n_layers, transfer ? 2 : 1);
        if (x.grid.n_wl < 1 || !x.grid.all_wls || !x.alpha)
            return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "star %lld: empty grid or null alpha",
                           (long long)k);
        if (x.cntm.mode != KORG_CNTM_GRID || x.cntm.n_knots < 2 || !x.cntm.knots || !x.cntm.values)
            return set_err(ctx, KORG_ERR_INVALID_ARGUMENT,
                           "star %lld: the continuum must be given on >= 2 knots (KORG_CNTM_GRID)",
                           (long long)k);
        if (transfer && !x.layers.temps)
            return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "star %lld: no temperatures", (long long)k);
        hs[k] = StarHost{x.alpha, x.grid, x.lines, x.layers, x.cntm, x.cutoff_threshold};
    }
    *L_out = L;
    return KORG_OK;
}

int run_continuum(korg_ctx* ctx, std::vector<StarHost>& hs, int L, int alpha_dtype) {
    hipStream_t s = ctx->stream;
    Workspace& ws = ctx->ws;
    int64_t N = 0, nwl = 0;
    TRY(upload_stars(ctx, ws, s, (int)hs.size(), hs.data(), &N, &nwl));
    hipLaunchKernelGGL(k_cntm_fill, dim3((unsigned)((nwl / WAVE + 3) / 4)), dim3(256), 0, s,
                       (int)hs.size(), L, nwl, (const StarDev*)ws.stars.p,
                       alpha_dtype == KORG_DTYPE_F32 ? 1 : 0);
    HIP_TRY(ctx, hipGetLastError());
    return KORG_OK;
}

int run_flux(korg_ctx* ctx, std::vector<StarHost>& hs, int L, int alpha_dtype) {
    hipStream_t s = ctx->stream;
    Workspace& ws = ctx->ws;
    int64_t N = 0, nwl = 0;
    TRY(upload_stars(ctx, ws, s, (int)hs.size(), hs.data(), &N, &nwl));
    TRY(grow(ctx, ws.rt_tab, sizeof(RtTab) * hs.size() * (size_t)L));
    hipLaunchKernelGGL(k_rt_tab, dim3((unsigned)hs.size()), dim3(64), 0, s, (int)hs.size(), L,
                       (const StarDev*)ws.stars.p, (RtTab*)ws.rt_tab.p, (int*)nullptr, 0);
    RtArgs ra;
    ra.n_stars = (int)hs.size();
    ra.L = L;
    ra.n_total = nwl;
    ra.stars = (const StarDev*)ws.stars.p;
    ra.tab = (const RtTab*)ws.rt_tab.p;
    ra.alpha_f32 = alpha_dtype == KORG_DTYPE_F32 ? 1 : 0;
    ra.chi2_part = nullptr;
    ra.flux_only = 1;
    hipLaunchKernelGGL(k_rt_fit, dim3((unsigned)((nwl + 255) / 256)), dim3(256), 0, s, ra);
    HIP_TRY(ctx, hipGetLastError());
    return KORG_OK;
}

// Times of the last adjoint call (setup, items, reduction, total).
int finish_vjp_stats(korg_ctx* ctx) {
    if (ctx->last_n_lines == 0) return KORG_OK;
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev[3]));
    ctx->stats.setup_ms = elapsed(ctx, 0, 1);
    ctx->stats.profile_ms = elapsed(ctx, 2, 4);
    ctx->stats.core_ms = ctx->last_overlap ? elapsed(ctx, 7, 8) : elapsed(ctx, 4, 3);
    ctx->stats.total_ms = elapsed(ctx, 0, 3);
    return KORG_OK;
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0; }

// The forward-mode pipeline of one star on device pointers: the Float64 form's setup and window
// partition, then per group of up to JVP_DIRS directions the coefficient kernel and the profile
// pass, which add J x_k into out[k].  `tan` and `out` are host arrays of K entries (device
// pointers inside).
int run_jvp(korg_ctx* ctx, void* const* out, int out_dtype, const korg_grid* g,
            const korg_lines* ln, const korg_species* sp, const korg_layers* ly,
            const korg_continuum* cn, double cutoff, int64_t K, const korg_layer_tangents* tan) {
    hipStream_t s = ctx->stream;
    Workspace& ws = ctx->ws;
    std::memset(&ctx->stats, 0, sizeof ctx->stats);
    ctx->last_n_lines = 0;
    ctx->last_fixed = false;
    ctx->last_sized = false;
    korg_tuning t = ctx->tuning;  // the profile pass uses the Float64 form's 64-point tiles
    t.points_per_lane = 1;
    t.count_branches = 0;
    const StarHost one = {nullptr, *g, *ln, *ly, *cn, cutoff};
    Batch b;
    TRY(plan_batch(ctx, t, 1, &one, sp, ly->n_layers, KORG_DTYPE_F64, KORG_PRECISION_FP64, ws, s,
                   b));
    ctx->last_n_lines = b.N;
    ctx->last_precision = KORG_PRECISION_FP64;
    ctx->last_n_lblocks = b.n_groups;
    ctx->last_nb = 0;
    if (b.N == 0 || K == 0) return KORG_OK;
    constexpr int G = 4;  // layers per work item of the profile pass (divides the rec64 stride)
    TRY(grow(ctx, ws.jvp_coef, sizeof(double) * 3 * JVP_DIRS * (size_t)b.N * b.Ls));
    TRY(enqueue_front(ctx, ws, b, s, INT64_MAX, true));
    HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));
    JvpArgs A{};
    A.L = b.L;
    A.Ls = b.Ls;
    A.stars = (const StarDev*)ws.stars.p;
    A.meta = (const LineMeta*)ws.meta.p;
    A.rec64 = (const double4*)ws.rec64.p;
    A.tab = (const LayerTab*)ws.tab.p;
    A.mass = sp->mass;
    A.is_mol = sp->is_molecule;
    A.coef = (double*)ws.jvp_coef.p;
    A.n_wl = b.nwl;
    A.ntiles = b.ntiles;
    A.tile_w = b.tile_w;
    A.n_lblocks = b.Ls / G;
    A.n_items = (int64_t)b.ntiles * A.n_lblocks;
    A.cands = (const Cand*)ws.cls_cand.p;
    A.ranges = (const int2*)ws.ranges.p;
    A.out_f32 = out_dtype == KORG_DTYPE_F32;
    A.counters = (unsigned long long*)ws.counters.p;
    int64_t nwaves = (int64_t)ctx->n_cu * b.wpc;
    if (nwaves > A.n_items) nwaves = A.n_items;
    for (int64_t k0 = 0; k0 < K; k0 += JVP_DIRS) {
        A.n_dirs = (int)std::min<int64_t>(JVP_DIRS, K - k0);
        for (int k = 0; k < JVP_DIRS; ++k) {
            const bool on = k < A.n_dirs;
            const korg_layer_tangents z{};
            const korg_layer_tangents& x = on ? tan[k0 + k] : z;
            A.t[k] = JvpTangent{x.T, x.n_e, x.xi, x.n_HI, x.n_HeI, x.n_H2, x.n_div_U};
            A.out[k] = on ? out[k0 + k] : nullptr;
        }
        if (k0 > 0) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, A.counters);
        hipLaunchKernelGGL(k_jvp_coef, dim3((unsigned)((b.N + 3) / 4)), dim3(256), 0, s, b.N, A);
        hipLaunchKernelGGL(k_jvp_profile<G>, dim3((unsigned)nwaves), dim3(WAVE), 0, s, A);
        HIP_TRY(ctx, hipGetLastError());
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
    HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
    return KORG_OK;
}

int check_jvp_args(korg_ctx* ctx, void* const* dalpha, int64_t K, const korg_layer_tangents* t,
                   int64_t n_lines) {
    if (K < 0 || K > 4096) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "n_tangents %lld out of "
                                          "[0, 4096]", (long long)K);
    if (K > 0 && n_lines > 0) {
        if (!dalpha || !t) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "null dalpha or tangents");
        for (int64_t k = 0; k < K; ++k)
            if (!dalpha[k]) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "dalpha[%lld] is null",
                                           (long long)k);
    }
    return KORG_OK;
}

// Validates the fit targets of a batch and puts them (and the gradient outputs) into the stars.
int collect_fit(korg_ctx* ctx, int64_t n_stars, const korg_fit_target* t,
                std::vector<StarHost>& hs, int precision) {
    if (!t) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "no fit targets");
    if (precision == KORG_PRECISION_FP64_SERIES)
        return set_err(ctx, KORG_ERR_UNSUPPORTED,
                       "the fit step takes KORG_PRECISION_FP64 or KORG_PRECISION_FP32_FIXUP");
    if (n_stars > 65535)
        return set_err(ctx, KORG_ERR_UNSUPPORTED, "at most 65535 stars per fit step");
    for (int64_t k = 0; k < n_stars; ++k) {
        const korg_fit_target& x = t[k];
        StarHost& h = hs[k];
        if (h.layers.n_layers < 2)
            return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "star %lld: radiative transfer needs >= 2 layers",
                           (long long)k);
        if (h.cntm.mode != KORG_CNTM_GRID)
            return set_err(ctx, KORG_ERR_INVALID_ARGUMENT,
                           "star %lld: the fit step needs the continuum on knots (KORG_CNTM_GRID)",
                           (long long)k);
        if (!x.tau_ref || !x.alpha_ref || !x.data || !x.sigma || !x.chi2)
            return set_err(ctx, KORG_ERR_INVALID_ARGUMENT,
                           "star %lld: tau_ref, alpha_ref, data, sigma and chi2 are required", (long long)k);
        h.tau_ref = x.tau_ref;
        h.alpha_ref = x.alpha_ref;
        h.data = x.data;
        h.sigma = x.sigma;
        h.flux = x.flux;
        h.chi2 = x.chi2;
        double* const gq[7] = {x.grads.T, x.grads.n_e, x.grads.xi, x.grads.n_HI, x.grads.n_HeI,
                               x.grads.n_H2, x.grads.n_div_U};
        for (int q = 0; q < 7; ++q) h.gq[q] = gq[q];
    }
    return KORG_OK;
}

// The fit step of a batch, directly (the host sizes the tile lists and the adjoint's segments).
int run_fit(korg_ctx* ctx, std::vector<StarHost>& hs, const korg_species* sp, int L,
            int alpha_dtype, int precision) {
    hipStream_t s = ctx->stream;
    Workspace& ws = ctx->ws;
    std::memset(&ctx->stats, 0, sizeof ctx->stats);
    ctx->last_ragged = false;
    TRY(place_fit_G(ctx, ws, hs, L));
    Batch b;
    TRY(plan_batch(ctx, ctx->tuning, (int)hs.size(), hs.data(), sp, L, alpha_dtype, precision, ws,
                   s, b));
    ctx->last_n_lines = b.N;
    ctx->last_fixed = false;
    ctx->last_precision = precision;
    ctx->last_n_lblocks = b.n_lblocks;
    ctx->last_nb = b.NB;
    ctx->last_ff = b.ff;
    ctx->last_ntiles = b.ntiles;
    ctx->last_sized = false;
    TRY(prepare_fit(ctx, ws, b));
    HIP_TRY(ctx, hipEventRecord(ctx->ev[0], s));
    TRY(enqueue_fit(ctx, ws, b, s, true, 1.0));
    HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
    return KORG_OK;
}

// Counters and time of the last direct fit step (its time includes the host's sizing).
int finish_fit_stats(korg_ctx* ctx) {
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev[3]));
    unsigned long long c[N_COUNTERS];
    HIP_TRY(ctx, hipMemcpy(c, ctx->ws.counters.p, sizeof c, hipMemcpyDeviceToHost));
    TRY(stats_from_counters(ctx, c, ctx->last_precision == KORG_PRECISION_FP32_FIXUP,
                            ctx->last_n_lblocks, ctx->last_nb, false, ctx->last_ff,
                            ctx->last_ntiles));
    ctx->stats.total_ms = elapsed(ctx, 0, 3);
    return KORG_OK;
}

// Windows of the last pipeline's lines, 1-based within each line's star (korg_get_windows).
int fetch_windows(korg_ctx* ctx, int64_t* lb, int64_t* ub, int64_t n_lines) {
    std::vector<LineMeta> m(n_lines);
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const Workspace& ws = ctx->last_fixed ? ctx->ws_fixed : ctx->ws;
    HIP_TRY(ctx, hipMemcpy(m.data(), ws.meta.p, sizeof(LineMeta) * n_lines, hipMemcpyDeviceToHost));
    // grid indices are global over the call's stars: make them 1-based within each line's star
    const auto& lo = ctx->last_line_off;
    const auto& go = ctx->last_grid_off;
    size_t k = 0;
    for (int64_t i = 0; i < n_lines; ++i) {
        while (k + 1 < lo.size() && lo[k + 1] <= i) ++k;
        const int64_t off = k < go.size() ? go[k] : 0;
        lb[i] = (int64_t)m[i].lb - off + 1;
        ub[i] = (int64_t)m[i].ub - off + 1;
    }
    return KORG_OK;
}

// A batch of several layer counts: one pipeline per count (run(stars, L) enqueues it), each
// synchronized and finished (fin()), the statistics summed and the windows kept in the caller's
// star order.  Each star's results are those of its count's pipeline, i.e. of a batch of those
// stars alone.
template <class Run, class Fin>
int run_ragged(korg_ctx* ctx, const std::vector<StarHost>& hs, Run&& run, Fin&& fin) {
    const std::vector<std::vector<int>> groups = layer_groups(hs);
    std::vector<int64_t> loff(hs.size() + 1, 0);
    for (size_t k = 0; k < hs.size(); ++k) loff[k + 1] = loff[k] + hs[k].lines.n_lines;
    std::vector<int64_t> lb((size_t)loff.back(), 1), ub((size_t)loff.back(), 0);
    korg_stats tot{};
    for (const std::vector<int>& g : groups) {
        std::vector<StarHost> sub = subset(hs, g);
        TRY(run(sub, sub[0].layers.n_layers));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        TRY(fin());
        add_stats(tot, ctx->stats);
        const int64_t n = ctx->last_n_lines;
        if (n > 0) {
            std::vector<int64_t> glb((size_t)n), gub((size_t)n);
            TRY(fetch_windows(ctx, glb.data(), gub.data(), n));
            int64_t q = 0;
            for (int k : g)
                for (int64_t i = 0; i < hs[k].lines.n_lines; ++i, ++q) {
                    lb[loff[k] + i] = glb[q];
                    ub[loff[k] + i] = gub[q];
                }
        }
    }
    ctx->stats = tot;
    ctx->last_n_lines = loff.back();
    ctx->ragged_lb.swap(lb);
    ctx->ragged_ub.swap(ub);
    ctx->last_ragged = true;
    return KORG_OK;
}

template <class T>
int stage(korg_ctx* ctx, DevBuf& b, const T* src, size_t n, const T** out) {
    if (!src || n == 0) {
        *out = nullptr;
        return KORG_OK;
    }
    TRY(grow(ctx, b, sizeof(T) * n));
    HIP_TRY(ctx, hipMemcpyAsync(b.p, src, sizeof(T) * n, hipMemcpyHostToDevice, ctx->stream));
    *out = (const T*)b.p;
    return KORG_OK;
}

// Copies the host arrays of one call into the context's staging buffers and points the device-side
// structs (already copied from the host structs) at them.
int stage_inputs(korg_ctx* ctx, const korg_grid* grid, const korg_lines* lines,
                 const korg_species* species, const korg_layers* layers, const korg_continuum* cntm,
                 korg_grid* g, korg_lines* ln, korg_species* sp, korg_layers* ly,
                 korg_continuum* cn) {
    const int64_t N = lines->n_lines, L = layers->n_layers, S = species->n_species;
    TRY(stage(ctx, ctx->h_wls, grid->all_wls, grid->n_wl, &g->all_wls));
    TRY(stage(ctx, ctx->h_lines[0], lines->wl, N, &ln->wl));
    TRY(stage(ctx, ctx->h_lines[1], lines->log_gf, N, &ln->log_gf));
    TRY(stage(ctx, ctx->h_lines[2], lines->E_lower, N, &ln->E_lower));
    TRY(stage(ctx, ctx->h_lines[3], lines->gamma_rad, N, &ln->gamma_rad));
    TRY(stage(ctx, ctx->h_lines[4], lines->gamma_stark, N, &ln->gamma_stark));
    TRY(stage(ctx, ctx->h_lines[5], lines->vdW1, N, &ln->vdW1));
    TRY(stage(ctx, ctx->h_lines[6], lines->vdW2, N, &ln->vdW2));
    TRY(stage(ctx, ctx->h_lines[7], lines->species, N, &ln->species));
    if (lines->window_lb && lines->window_ub) {
        TRY(stage(ctx, ctx->h_win[0], lines->window_lb, N, &ln->window_lb));
        TRY(stage(ctx, ctx->h_win[1], lines->window_ub, N, &ln->window_ub));
    } else {
        ln->window_lb = ln->window_ub = nullptr;
    }
    TRY(stage(ctx, ctx->h_species_mass, species->mass, S, &sp->mass));
    TRY(stage(ctx, ctx->h_species_mol, species->is_molecule, S, &sp->is_molecule));
    TRY(stage(ctx, ctx->h_layers[0], layers->temps, L, &ly->temps));
    TRY(stage(ctx, ctx->h_layers[1], layers->n_e, L, &ly->n_e));
    TRY(stage(ctx, ctx->h_layers[2], layers->n_div_U, S * L, &ly->n_div_U));
    TRY(stage(ctx, ctx->h_layers[3], layers->n_HI, L, &ly->n_HI));
    TRY(stage(ctx, ctx->h_layers[4], layers->n_HeI, L, &ly->n_HeI));
    TRY(stage(ctx, ctx->h_layers[5], layers->n_H2, L, &ly->n_H2));
    TRY(stage(ctx, ctx->h_layers[6], layers->xi, L, &ly->xi));
    if (cntm->mode == KORG_CNTM_GRID) {
        TRY(stage(ctx, ctx->h_knots, cntm->knots, cntm->n_knots, &cn->knots));
        TRY(stage(ctx, ctx->h_cntm, cntm->values, (size_t)L * cntm->n_knots, &cn->values));
    } else {
        TRY(stage(ctx, ctx->h_cntm, cntm->values, (size_t)L * N, &cn->values));
    }
    return KORG_OK;
}

}  // namespace

// A batched forward call captured in a HIP graph for fixed device pointers: the whole pipeline,
// with no host round trip, replayed by korg_plan_launch.  Owns its workspace.
struct korg_plan {
    korg_ctx* ctx = nullptr;
    Workspace ws;
    std::vector<StarHost> stars;
    korg_species sp{};
    int L = 0, alpha_dtype = 0, precision = 0;
    korg_tuning tuning{};
    Batch b;
    double headroom = 1.25;  // tile-list entries allocated per entry measured
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    unsigned long long* h_counters = nullptr;  // pinned; the graph's last node copies the counters
    bool pending = false;                      // launched, counters not yet read
    int n_captures = 0;
    bool fit = false;  // the fit step (korg_plan_fit) rather than the forward alone
};

namespace {

void plan_drop_graph(korg_plan* p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

void plan_free(korg_plan* p) {
    if (!p) return;
    korg_ctx* ctx = p->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    plan_drop_graph(p);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    if (p->h_counters) (void)hipHostFree(p->h_counters);
    p->ws.release();
    delete p;
}

// Captures the plan's pipeline -- front half, back half, and the copy of the counters to pinned
// host memory -- into a graph and instantiates it.
int plan_capture(korg_plan* p) {
    korg_ctx* ctx = p->ctx;
    plan_drop_graph(p);
    HIP_TRY(ctx, hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
    int rc = KORG_OK;
    if (p->fit) {
        rc = enqueue_fit(ctx, p->ws, p->b, p->cap_stream, false, 0.0);
    } else {
        rc = enqueue_front(ctx, p->ws, p->b, p->cap_stream, p->ws.tile_cap, false, true);
        if (rc == KORG_OK) rc = enqueue_back(ctx, p->ws, p->b, p->cap_stream, false, true);
    }
    if (rc == KORG_OK) {
        const hipError_t e = hipMemcpyAsync(p->h_counters, p->ws.counters.p,
                                            sizeof(unsigned long long) * N_COUNTERS,
                                            hipMemcpyDeviceToHost, p->cap_stream);
        if (e != hipSuccess)
            rc = set_err(ctx, KORG_ERR_HIP, "hipMemcpyAsync of the counters failed: %s",
                         hipGetErrorString(e));
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
    if (rc != KORG_OK || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        if (rc != KORG_OK) return rc;
        return set_err(ctx, KORG_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    }
    p->graph = g;
    HIP_TRY(ctx, hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
    ++p->n_captures;
    return KORG_OK;
}

// Shapes the batch, measures the per-tile line lists with one pass of the front half (Float32
// form), allocates them with headroom, and captures the graph.
int plan_build(korg_plan* p) {
    korg_ctx* ctx = p->ctx;
    hipStream_t s = ctx->stream;
    if (p->fit) TRY(place_fit_G(ctx, p->ws, p->stars, p->L));
    TRY(plan_batch(ctx, p->tuning, (int)p->stars.size(), p->stars.data(), &p->sp, p->L,
                   p->alpha_dtype, p->precision, p->ws, s, p->b));
    if (p->fit) {
        // one measuring pass of the setup: tile lists (Float32) and adjoint segments
        TRY(prepare_fit(ctx, p->ws, p->b));
        if (p->b.N > 0) {
            TRY(enqueue_front(ctx, p->ws, p->b, s, INT64_MAX, false));
            if (p->b.f32) {
                int64_t total = 0;
                TRY(size_tile_lists(ctx, p->ws, p->b, s, p->headroom, &total));
            }
        } else {
            hipLaunchKernelGGL(k_layers, dim3(p->b.n_stars), dim3(64), 0, s, p->b.n_stars, p->b.L,
                               (StarDev*)p->ws.stars.p, (LayerTab*)p->ws.tab.p,
                               (unsigned long long*)p->ws.counters.p);
        }
        TRY(enqueue_vjp_counts(ctx, p->ws, p->b, s, INT64_MAX));
        TRY(size_segments(ctx, p->ws, p->b, s, p->headroom));
        return plan_capture(p);
    }
    if (p->b.N == 0) return KORG_OK;
    if (p->b.f32) {
        TRY(enqueue_front(ctx, p->ws, p->b, s, INT64_MAX, false));
        int64_t total = 0;
        TRY(size_tile_lists(ctx, p->ws, p->b, s, p->headroom, &total));
    }
    return plan_capture(p);
}

int plan_launch(korg_plan* p) {
    korg_ctx* ctx = p->ctx;
    HIP_TRY(ctx, hipEventRecord(p->ev0, ctx->stream));
    HIP_TRY(ctx, hipGraphLaunch(p->exec, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(p->ev1, ctx->stream));
    p->pending = true;
    return KORG_OK;
}

// Waits for the last replay and reads its counters.  A replay whose tile lists did not fit wrote
// nothing: the lists are grown to the total it measured, the graph captured again and replayed.
int plan_wait(korg_plan* p) {
    korg_ctx* ctx = p->ctx;
    if (!p->pending) return KORG_OK;
    HIP_TRY(ctx, hipEventSynchronize(p->ev1));
    p->pending = false;
    std::memset(&ctx->stats, 0, sizeof ctx->stats);
    const unsigned long long flags = p->h_counters[3];
    if ((flags & FLAG_CAPACITY) && !(flags & (FLAG_SPECIES | FLAG_UNSORTED))) {
        const int64_t total = (int64_t)p->h_counters[13];
        if (total > (int64_t)INT32_MAX)
            return set_err(ctx, KORG_ERR_UNSUPPORTED,
                           "more than 2^31 (line, tile) pairs in one call; split the line list");
        if (total > p->ws.tile_cap) {
            int64_t cap = (int64_t)((double)total * p->headroom) + 4096;
            if (cap > (int64_t)INT32_MAX) cap = INT32_MAX;
            TRY(grow(ctx, p->ws.tile_lines, sizeof(int) * (size_t)cap));
            p->ws.tile_cap = cap;
        }
        const int64_t segs = (int64_t)p->h_counters[14];
        if (p->fit && segs > p->ws.seg_cap) {
            const int64_t cap = (int64_t)((double)segs * p->headroom) + 64;
            TRY(grow(ctx, p->ws.vjp_partial,
                     sizeof(double) * (size_t)cap * (6 + p->sp.n_species) * p->L));
            p->ws.seg_cap = cap;
        }
        TRY(plan_capture(p));
        TRY(plan_launch(p));
        HIP_TRY(ctx, hipEventSynchronize(p->ev1));
        p->pending = false;
    }
    TRY(stats_from_counters(ctx, p->h_counters, p->b.f32, p->b.n_lblocks, p->b.NB, false, p->b.ff,
                            p->b.ntiles));
    float t = 0;
    if (hipEventElapsedTime(&t, p->ev0, p->ev1) == hipSuccess) ctx->stats.total_ms = t;
    else (void)hipGetLastError();
    return KORG_OK;
}

// ---------------------------------------------------------------------------------------------
// autotuner
// ---------------------------------------------------------------------------------------------
int ilog2(int64_t x) {
    int k = 0;
    while (x > 1) {
        x >>= 1;
        ++k;
    }
    return k;
}

// Key of a tuning-cache entry: device, precision, alpha dtype, layers, and the power-of-two class
// of the star, line and grid-point counts.
std::string tune_key(const korg_ctx* ctx, int precision, int alpha_dtype, int L, int64_t n_stars,
                     int64_t N, int64_t nwl) {
    char k[256];
    snprintf(k, sizeof k, "%s/cu%d/p%d/a%d/L%d/s%d/n%d/w%d", ctx->arch.c_str(), ctx->n_cu,
             precision, alpha_dtype, L, ilog2(n_stars), ilog2(N), ilog2(nwl));
    return k;
}

// Cache file: one entry per line, "key tile_points waves_per_tile waves_per_cu
// near_waves_per_cu points_per_lane ms"; '#' starts a comment; the last entry of a key wins.
bool cache_lookup(const char* path, const std::string& key, korg_tuning* out) {
    FILE* f = fopen(path, "r");
    if (!f) return false;
    char line[512];
    bool found = false;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        char k[256];
        int tp, wt, wc, nw, pl;
        double ms;
        if (sscanf(line, "%255s %d %d %d %d %d %lf", k, &tp, &wt, &wc, &nw, &pl, &ms) != 7) continue;
        if (key != k) continue;
        korg_tuning t{};
        t.tile_points = tp;
        t.waves_per_tile = wt;
        t.waves_per_cu = wc;
        t.near_waves_per_cu = nw;
        t.points_per_lane = pl;
        if (check_tuning(nullptr, &t) == KORG_OK) {
            *out = t;
            found = true;
        }
    }
    fclose(f);
    return found;
}

int cache_store(korg_ctx* ctx, const char* path, const std::string& key, const korg_tuning& t,
                double ms) {
    FILE* f = fopen(path, "a");
    if (!f) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "cannot append to tuning cache %s", path);
    fprintf(f, "%s %d %d %d %d %d %.4f\n", key.c_str(), t.tile_points, t.waves_per_tile,
            t.waves_per_cu, t.near_waves_per_cu, t.points_per_lane, ms);
    fclose(f);
    return KORG_OK;
}

// Device time of the eager pipeline under tuning t: the best of `reps` runs.
int time_tuning(korg_ctx* ctx, const korg_tuning& t, const std::vector<StarHost>& hs,
                const korg_species* sp, int L, int alpha_dtype, int precision, int reps,
                double* ms) {
    ctx->tuning = t;
    double best = 1e300;
    for (int r = 0; r < reps; ++r) {
        TRY(run_batch(ctx, (int)hs.size(), hs.data(), sp, L, alpha_dtype, precision));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        TRY(finish_stats(ctx));
        best = std::min(best, (double)ctx->stats.total_ms);
    }
    *ms = best;
    return KORG_OK;
}

// Coordinate search: tile shape (Float32 form: points per tile x waves per tile; Float64 form:
// points per lane), then persistent waves per CU, then (Float32) near-zone waves per CU.
int autotune(korg_ctx* ctx, const std::vector<StarHost>& hs, const korg_species* sp, int L,
             int alpha_dtype, int precision, korg_tuning* best, double* best_ms) {
    const bool f32 = precision == KORG_PRECISION_FP32_FIXUP;
    korg_tuning cur{};
    cur.waves_per_cu = 12;
    cur.near_waves_per_cu = 16;
    double bms = 1e300;
    korg_tuning bt = cur;
    auto consider = [&](const korg_tuning& t) -> int {
        double ms = 0;
        TRY(time_tuning(ctx, t, hs, sp, L, alpha_dtype, precision, 2, &ms));
        if (ms < bms) {
            bms = ms;
            bt = t;
        }
        return KORG_OK;
    };
    if (f32) {
        const int shapes[][2] = {{32, 1}, {32, 2}, {32, 4}, {64, 1}, {64, 2}, {64, 4}, {128, 1}, {128, 2}};
        for (const auto& sh : shapes) {
            korg_tuning t = cur;
            t.tile_points = sh[0];
            t.waves_per_tile = sh[1];
            TRY(consider(t));
        }
    } else {
        for (int pl : {1, 2}) {
            korg_tuning t = cur;
            t.points_per_lane = pl;
            TRY(consider(t));
        }
    }
    const korg_tuning shape = bt;
    for (int wc : {8, 16, 24}) {
        korg_tuning t = shape;
        t.waves_per_cu = wc;
        TRY(consider(t));
    }
    if (f32) {
        const korg_tuning occ = bt;
        for (int nw : {8, 32}) {
            korg_tuning t = occ;
            t.near_waves_per_cu = nw;
            TRY(consider(t));
        }
    }
    *best = bt;
    *best_ms = bms;
    return KORG_OK;
}

}  // namespace

extern "C" {

int korg_abi_version(void) { return KORG_LINE_ABI_VERSION; }

int korg_create(korg_ctx** out, int device) {
    if (!out) return KORG_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return KORG_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) return KORG_ERR_INVALID_ARGUMENT;
    korg_ctx* ctx = new korg_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete ctx;
        return KORG_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete ctx;
        return KORG_ERR_HIP;
    }
    ctx->n_cu = prop.multiProcessorCount;
    ctx->arch = prop.gcnArchName;
    const size_t colon = ctx->arch.find(':');
    if (colon != std::string::npos) ctx->arch.resize(colon);
    if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return KORG_ERR_HIP;
    }
    ctx->stream = ctx->own_stream;
    for (auto& e : ctx->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            korg_destroy(ctx);
            return KORG_ERR_HIP;
        }
    }
    if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming) != hipSuccess) {
        korg_destroy(ctx);
        return KORG_ERR_HIP;
    }
    // set-up of the allocation-free entry point: the fixed scratch's address on this device, and
    // host-side room for a one-star table, so that its calls allocate nothing
    resolve_scratch(ctx);
    ctx->ws_fixed.stars_host.reserve(1);
    ctx->last_grid_off.reserve(1);
    ctx->last_line_off.reserve(1);
    *out = ctx;
    return KORG_OK;
}

void korg_destroy(korg_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->ws.release();
    DevBuf* bufs[] = {&ctx->vjp_out,  &ctx->h_grad,   &ctx->h_alpha,   &ctx->h_wls,
                      &ctx->h_species_mass, &ctx->h_species_mol, &ctx->h_knots, &ctx->h_cntm,
                      &ctx->hl_trans, &ctx->hl_T,     &ctx->hl_ne,     &ctx->hl_dnu,
                      &ctx->hl_logP,  &ctx->hl_lam0,  &ctx->hl_pairs,  &ctx->hl_slices,
                      &ctx->hl_hull,  &ctx->hl_brec,  &ctx->hl_bprof,  &ctx->hl_bhull};
    for (DevBuf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (auto& b : ctx->h_lines)
        if (b.p) (void)hipFree(b.p);
    for (DevBuf* b : {&ctx->h_win[0], &ctx->h_win[1], &ctx->h_tan, &ctx->h_dalpha, &ctx->mol_tabs,
                      &ctx->h_mol, &ctx->flt_keys, &ctx->flt_found, &ctx->flt_runs})
        if (b->p) (void)hipFree(b->p);
    for (auto& b : ctx->h_layers)
        if (b.p) (void)hipFree(b.p);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

const char* korg_last_error(const korg_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

int korg_set_stream(korg_ctx* ctx, void* stream) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->stream = stream ? (hipStream_t)stream : ctx->own_stream;
    return KORG_OK;
}

int korg_synchronize(korg_ctx* ctx) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return finish_stats(ctx);
}

int korg_get_stats(const korg_ctx* ctx, korg_stats* out) {
    if (!ctx || !out) return KORG_ERR_INVALID_ARGUMENT;
    *out = ctx->stats;
    return KORG_OK;
}

int korg_set_tuning(korg_ctx* ctx, const korg_tuning* t) {
    if (!ctx || !t) return KORG_ERR_INVALID_ARGUMENT;
    TRY(check_tuning(ctx, t));
    ctx->tuning = *t;
    return KORG_OK;
}

int korg_get_tuning(const korg_ctx* ctx, korg_tuning* out) {
    if (!ctx || !out) return KORG_ERR_INVALID_ARGUMENT;
    *out = ctx->tuning;
    return KORG_OK;
}

int korg_line_absorption_device(korg_ctx* ctx, void* alpha, int alpha_dtype, const korg_grid* grid,
                                const korg_lines* lines, const korg_species* species,
                                const korg_layers* layers, const korg_continuum* cntm,
                                double cutoff_threshold, int precision, int synchronize) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    TRY(validate(ctx, grid, lines, species, layers, cntm, alpha_dtype, precision));
    if (!alpha && lines->n_lines > 0) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "alpha is null");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    TRY(run_device(ctx, alpha, alpha_dtype, grid, lines, species, layers, cntm, cutoff_threshold,
                   precision));
    if (synchronize) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        TRY(finish_stats(ctx));
    }
    return KORG_OK;
}

int korg_line_absorption_device_noalloc(korg_ctx* ctx, void* alpha, int alpha_dtype,
                                        const korg_grid* grid, const korg_lines* lines,
                                        const korg_species* species, const korg_layers* layers,
                                        const korg_continuum* cntm, double cutoff_threshold,
                                        int precision, int synchronize) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    TRY(validate(ctx, grid, lines, species, layers, cntm, alpha_dtype, precision));
    if (!alpha && lines->n_lines > 0) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "alpha is null");
    if (!ctx->scratch_ptr)
        return set_err(ctx, KORG_ERR_HIP, "no fixed scratch on device %d: %s", ctx->device,
                       ctx->scratch_err.c_str());
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const StarHost one = {alpha, *grid, *lines, *layers, *cntm, cutoff_threshold};
    const int rc = run_batch(ctx, 1, &one, species, layers->n_layers, alpha_dtype, precision, true);
    if (rc == KORG_ERR_CAPACITY)
        return set_err(ctx, KORG_ERR_CAPACITY,
                       "%lld lines x %d layers on %lld points need more than the fixed scratch of "
                       "%lld bytes (%s); use korg_line_absorption_device",
                       (long long)lines->n_lines, layers->n_layers, (long long)grid->n_wl,
                       (long long)KORG_NOALLOC_SCRATCH_BYTES, ctx->err.c_str());
    TRY(rc);
    if (synchronize) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        TRY(finish_stats(ctx));
    }
    return KORG_OK;
}

int korg_line_absorption_batched(korg_ctx* ctx, int64_t n_stars, const korg_star* stars,
                                 const korg_species* species, int alpha_dtype, int precision,
                                 int synchronize) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    std::vector<StarHost> hs;
    int L = 0;
    TRY(collect_stars(ctx, n_stars, stars, species, alpha_dtype, precision, hs, &L));
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (layer_groups(hs).size() > 1)
        return run_ragged(
            ctx, hs,
            [&](std::vector<StarHost>& sub, int Lg) {
                return run_batch(ctx, (int)sub.size(), sub.data(), species, Lg, alpha_dtype, precision);
            },
            [&] { return finish_stats(ctx); });
    TRY(run_batch(ctx, (int)n_stars, hs.data(), species, L, alpha_dtype, precision));
    if (synchronize) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        TRY(finish_stats(ctx));
    }
    return KORG_OK;
}

int korg_line_absorption(korg_ctx* ctx, void* alpha, int alpha_dtype, const korg_grid* grid,
                         const korg_lines* lines, const korg_species* species,
                         const korg_layers* layers, const korg_continuum* cntm,
                         double cutoff_threshold, int precision) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    TRY(validate(ctx, grid, lines, species, layers, cntm, alpha_dtype, precision));
    if (lines->n_lines == 0) {
        ctx->last_n_lines = 0;
        std::memset(&ctx->stats, 0, sizeof ctx->stats);
        return KORG_OK;
    }
    if (!alpha) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "alpha is null");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int64_t L = layers->n_layers;
    const size_t esz = alpha_dtype == KORG_DTYPE_F32 ? 4 : 8;
    const size_t abytes = esz * (size_t)L * (size_t)grid->n_wl;
    korg_grid g = *grid;
    korg_lines ln = *lines;
    korg_species sp = *species;
    korg_layers ly = *layers;
    korg_continuum cn = *cntm;
    TRY(stage_inputs(ctx, grid, lines, species, layers, cntm, &g, &ln, &sp, &ly, &cn));
    TRY(grow(ctx, ctx->h_alpha, abytes));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->h_alpha.p, alpha, abytes, hipMemcpyHostToDevice, ctx->stream));
    TRY(run_device(ctx, ctx->h_alpha.p, alpha_dtype, &g, &ln, &sp, &ly, &cn, cutoff_threshold,
                   precision));
    HIP_TRY(ctx, hipMemcpyAsync(alpha, ctx->h_alpha.p, abytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return finish_stats(ctx);
}

int korg_line_absorption_vjp_device(korg_ctx* ctx, const void* grad_alpha, int grad_dtype,
                                    const korg_grid* grid, const korg_lines* lines,
                                    const korg_species* species, const korg_layers* layers,
                                    const korg_continuum* cntm, double cutoff_threshold,
                                    korg_layer_grads* grads, int synchronize) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    TRY(validate(ctx, grid, lines, species, layers, cntm, grad_dtype, KORG_PRECISION_FP64));
    if (!grads) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "grads is null");
    if (!grad_alpha && lines->n_lines > 0)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "grad_alpha is null");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    double* const gq[7] = {grads->T, grads->n_e, grads->xi, grads->n_HI, grads->n_HeI, grads->n_H2,
                           grads->n_div_U};
    TRY(run_vjp(ctx, grad_alpha, grad_dtype, grid, lines, species, layers, cntm, cutoff_threshold,
                gq));
    if (synchronize) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        TRY(finish_vjp_stats(ctx));
    }
    return KORG_OK;
}

int korg_line_absorption_vjp(korg_ctx* ctx, const void* grad_alpha, int grad_dtype,
                             const korg_grid* grid, const korg_lines* lines,
                             const korg_species* species, const korg_layers* layers,
                             const korg_continuum* cntm, double cutoff_threshold,
                             korg_layer_grads* grads) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    TRY(validate(ctx, grid, lines, species, layers, cntm, grad_dtype, KORG_PRECISION_FP64));
    if (!grads) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "grads is null");
    if (!grad_alpha && lines->n_lines > 0)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "grad_alpha is null");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int64_t N = lines->n_lines, L = layers->n_layers;
    korg_grid g = *grid;
    korg_lines ln = *lines;
    korg_species sp = *species;
    korg_layers ly = *layers;
    korg_continuum cn = *cntm;
    TRY(stage_inputs(ctx, grid, lines, species, layers, cntm, &g, &ln, &sp, &ly, &cn));
    const void* Gd = nullptr;
    if (N > 0) {
        const size_t gbytes = (grad_dtype == KORG_DTYPE_F32 ? 4 : 8) * (size_t)L * (size_t)grid->n_wl;
        TRY(grow(ctx, ctx->h_grad, gbytes));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->h_grad.p, grad_alpha, gbytes, hipMemcpyHostToDevice,
                                    ctx->stream));
        Gd = ctx->h_grad.p;
    }
    // the gradients go to device scratch, then to the caller's host arrays
    const int64_t S = species->n_species;
    TRY(grow(ctx, ctx->vjp_out, sizeof(double) * (size_t)((6 + S) * L)));
    double* res = (double*)ctx->vjp_out.p;
    double* const dst[7] = {grads->T, grads->n_e, grads->xi, grads->n_HI, grads->n_HeI,
                            grads->n_H2, grads->n_div_U};
    double* gq[7];
    for (int q = 0; q < 7; ++q) gq[q] = dst[q] ? res + (size_t)q * L : nullptr;
    TRY(run_vjp(ctx, Gd, grad_dtype, &g, &ln, &sp, &ly, &cn, cutoff_threshold, gq));
    for (int q = 0; q < 7; ++q) {
        if (!dst[q]) continue;
        const size_t cnt = q < 6 ? (size_t)L : (size_t)(S * L);
        HIP_TRY(ctx, hipMemcpyAsync(dst[q], gq[q], sizeof(double) * cnt, hipMemcpyDeviceToHost,
                                    ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return finish_vjp_stats(ctx);
}

int korg_line_absorption_jvp_device(korg_ctx* ctx, void* const* dalpha, int dalpha_dtype,
                                    const korg_grid* grid, const korg_lines* lines,
                                    const korg_species* species, const korg_layers* layers,
                                    const korg_continuum* cntm, double cutoff_threshold,
                                    int64_t n_tangents, const korg_layer_tangents* tangents,
                                    int synchronize) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    TRY(validate(ctx, grid, lines, species, layers, cntm, dalpha_dtype, KORG_PRECISION_FP64));
    TRY(check_jvp_args(ctx, dalpha, n_tangents, tangents, lines->n_lines));
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    TRY(run_jvp(ctx, dalpha, dalpha_dtype, grid, lines, species, layers, cntm, cutoff_threshold,
                n_tangents, tangents));
    if (synchronize) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        TRY(finish_stats(ctx));
    }
    return KORG_OK;
}

int korg_line_absorption_jvp(korg_ctx* ctx, void* const* dalpha, int dalpha_dtype,
                             const korg_grid* grid, const korg_lines* lines,
                             const korg_species* species, const korg_layers* layers,
                             const korg_continuum* cntm, double cutoff_threshold,
                             int64_t n_tangents, const korg_layer_tangents* tangents) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    TRY(validate(ctx, grid, lines, species, layers, cntm, dalpha_dtype, KORG_PRECISION_FP64));
    TRY(check_jvp_args(ctx, dalpha, n_tangents, tangents, lines->n_lines));
    if (lines->n_lines == 0 || n_tangents == 0) {
        ctx->last_n_lines = 0;
        std::memset(&ctx->stats, 0, sizeof ctx->stats);
        return KORG_OK;
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int64_t L = layers->n_layers, S = species->n_species, K = n_tangents;
    korg_grid g = *grid;
    korg_lines ln = *lines;
    korg_species sp = *species;
    korg_layers ly = *layers;
    korg_continuum cn = *cntm;
    TRY(stage_inputs(ctx, grid, lines, species, layers, cntm, &g, &ln, &sp, &ly, &cn));
    // tangents: per direction 6 per-layer arrays and the n/U block, packed; null stays null
    const size_t per = (size_t)(6 * L + S * L);
    TRY(grow(ctx, ctx->h_tan, sizeof(double) * per * (size_t)K));
    std::vector<korg_layer_tangents> td((size_t)K);
    for (int64_t k = 0; k < K; ++k) {
        double* base = (double*)ctx->h_tan.p + per * (size_t)k;
        const double* src[7] = {tangents[k].T, tangents[k].n_e, tangents[k].xi, tangents[k].n_HI,
                                tangents[k].n_HeI, tangents[k].n_H2, tangents[k].n_div_U};
        const double* dst[7] = {};
        for (int q = 0; q < 7; ++q) {
            if (!src[q]) continue;
            double* d = base + (size_t)q * L;
            const size_t cnt = q < 6 ? (size_t)L : (size_t)(S * L);
            HIP_TRY(ctx, hipMemcpyAsync(d, src[q], sizeof(double) * cnt, hipMemcpyHostToDevice,
                                        ctx->stream));
            dst[q] = d;
        }
        td[k] = korg_layer_tangents{dst[0], dst[1], dst[2], dst[3], dst[4], dst[5], dst[6]};
    }
    const size_t esz = dalpha_dtype == KORG_DTYPE_F32 ? 4 : 8;
    const size_t abytes = esz * (size_t)L * (size_t)grid->n_wl;
    TRY(grow(ctx, ctx->h_dalpha, abytes * (size_t)K));
    std::vector<void*> dd((size_t)K);
    for (int64_t k = 0; k < K; ++k) {
        dd[k] = (char*)ctx->h_dalpha.p + abytes * (size_t)k;
        HIP_TRY(ctx, hipMemcpyAsync(dd[k], dalpha[k], abytes, hipMemcpyHostToDevice, ctx->stream));
    }
    TRY(run_jvp(ctx, dd.data(), dalpha_dtype, &g, &ln, &sp, &ly, &cn, cutoff_threshold, K,
                td.data()));
    for (int64_t k = 0; k < K; ++k)
        HIP_TRY(ctx, hipMemcpyAsync(dalpha[k], dd[k], abytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return finish_stats(ctx);
}

int korg_get_windows(korg_ctx* ctx, int64_t* lb, int64_t* ub, int64_t n_lines) {
    if (!ctx || !lb || !ub) return KORG_ERR_INVALID_ARGUMENT;
    if (n_lines != ctx->last_n_lines)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "n_lines %lld != last call's %lld",
                       (long long)n_lines, (long long)ctx->last_n_lines);
    if (n_lines == 0) return KORG_OK;
    if (ctx->last_ragged) {
        std::copy(ctx->ragged_lb.begin(), ctx->ragged_lb.end(), lb);
        std::copy(ctx->ragged_ub.begin(), ctx->ragged_ub.end(), ub);
        return KORG_OK;
    }
    return fetch_windows(ctx, lb, ub, n_lines);
}

}  // extern "C"

namespace {

// A plan of the forward (targets null) or of the fit step.
int plan_create(korg_ctx* ctx, int64_t n_stars, const korg_star* stars,
                const korg_fit_target* targets, const korg_species* species, int alpha_dtype,
                int precision, double headroom, korg_plan** out) {
    if (!ctx || !out) return KORG_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    ctx->err.clear();
    if (!(headroom == 0.0 || (headroom >= 1.0 && headroom <= 16.0)))
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "headroom must be 0 or in [1, 16]");
    korg_plan* p = new korg_plan();
    p->ctx = ctx;
    int rc = collect_stars(ctx, n_stars, stars, species, alpha_dtype, precision, p->stars, &p->L);
    if (rc == KORG_OK && layer_groups(p->stars).size() > 1)
        rc = set_err(ctx, KORG_ERR_UNSUPPORTED,
                     "a plan needs one layer count for all its stars (capture one plan per count)");
    if (rc == KORG_OK && targets) {
        p->fit = true;
        rc = collect_fit(ctx, n_stars, targets, p->stars, precision);
    }
    if (rc != KORG_OK) {
        delete p;
        return rc;
    }
    p->sp = *species;
    p->alpha_dtype = alpha_dtype;
    p->precision = precision;
    p->tuning = ctx->tuning;
    p->headroom = headroom > 0.0 ? headroom : 1.25;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&p->ev0);
    if (e == hipSuccess) e = hipEventCreate(&p->ev1);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&p->h_counters, sizeof(unsigned long long) * N_COUNTERS,
                          hipHostMallocDefault);
    if (e != hipSuccess) {
        rc = set_err(ctx, KORG_ERR_HIP, "plan resources: %s", hipGetErrorString(e));
    } else {
        std::memset(p->h_counters, 0, sizeof(unsigned long long) * N_COUNTERS);
        rc = plan_build(p);
    }
    if (rc != KORG_OK) {
        plan_free(p);
        return rc;
    }
    *out = p;
    return KORG_OK;
}

}  // namespace

extern "C" {

int korg_plan_batched(korg_ctx* ctx, int64_t n_stars, const korg_star* stars,
                      const korg_species* species, int alpha_dtype, int precision,
                      double headroom, korg_plan** out) {
    return plan_create(ctx, n_stars, stars, nullptr, species, alpha_dtype, precision, headroom, out);
}

int korg_plan_fit(korg_ctx* ctx, int64_t n_stars, const korg_star* stars,
                  const korg_fit_target* targets, const korg_species* species, int alpha_dtype,
                  int precision, double headroom, korg_plan** out) {
    if (!targets) return ctx ? set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "no fit targets") : KORG_ERR_INVALID_ARGUMENT;
    return plan_create(ctx, n_stars, stars, targets, species, alpha_dtype, precision, headroom, out);
}

int korg_fit_gradient_batched(korg_ctx* ctx, int64_t n_stars, const korg_star* stars,
                              const korg_fit_target* targets, const korg_species* species,
                              int alpha_dtype, int precision, int synchronize) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    std::vector<StarHost> hs;
    int L = 0;
    TRY(collect_stars(ctx, n_stars, stars, species, alpha_dtype, precision, hs, &L));
    TRY(collect_fit(ctx, n_stars, targets, hs, precision));
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (layer_groups(hs).size() > 1)
        return run_ragged(
            ctx, hs,
            [&](std::vector<StarHost>& sub, int Lg) {
                return run_fit(ctx, sub, species, Lg, alpha_dtype, precision);
            },
            [&] { return finish_fit_stats(ctx); });
    TRY(run_fit(ctx, hs, species, L, alpha_dtype, precision));
    if (synchronize) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        TRY(finish_fit_stats(ctx));
    }
    return KORG_OK;
}

int korg_continuum_batched(korg_ctx* ctx, int64_t n_stars, const korg_star* stars, int alpha_dtype,
                           int synchronize) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    std::vector<StarHost> hs;
    int L = 0;
    TRY(collect_rt_stars(ctx, n_stars, stars, alpha_dtype, false, hs, &L));
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const std::vector<std::vector<int>> groups = layer_groups(hs);
    for (size_t i = 0; i < groups.size(); ++i) {  // one pass per layer count
        if (i > 0) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // the star table is reused
        std::vector<StarHost> sub = subset(hs, groups[i]);
        TRY(run_continuum(ctx, sub, sub[0].layers.n_layers, alpha_dtype));
    }
    if (synchronize) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return KORG_OK;
}

int korg_flux_batched(korg_ctx* ctx, int64_t n_stars, const korg_star* stars,
                      const korg_flux_target* targets, int alpha_dtype, int synchronize) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    if (!targets) return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "no flux targets");
    std::vector<StarHost> hs;
    int L = 0;
    TRY(collect_rt_stars(ctx, n_stars, stars, alpha_dtype, true, hs, &L));
    for (int64_t k = 0; k < n_stars; ++k) {
        const korg_flux_target& t = targets[k];
        if (!t.tau_ref || !t.alpha_ref)
            return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "star %lld: tau_ref and alpha_ref are required",
                           (long long)k);
        hs[k].tau_ref = t.tau_ref;
        hs[k].alpha_ref = t.alpha_ref;
        hs[k].flux = t.flux;
        hs[k].flux_c = t.cntm;
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const std::vector<std::vector<int>> groups = layer_groups(hs);
    for (size_t i = 0; i < groups.size(); ++i) {
        if (i > 0) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        std::vector<StarHost> sub = subset(hs, groups[i]);
        TRY(run_flux(ctx, sub, sub[0].layers.n_layers, alpha_dtype));
    }
    if (synchronize) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return KORG_OK;
}

int korg_set_hydrogen_tables(korg_ctx* ctx, const korg_htables* t) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    if (!t || t->n_transitions < 1 || t->n_transitions > 4096 || !t->transitions || !t->temps ||
        !t->n_e || !t->log_dnu || !t->log_profile || !t->lambda0)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "hydrogen tables: missing arrays or no transitions");
    std::vector<HTrans> tr(t->n_transitions);
    int max_dnu = 0, n_max = 0;
    auto fits = [](int64_t off, int64_t len, int64_t n) { return off >= 0 && len >= 0 && off <= n - len; };
    auto ascending = [](const double* a, int64_t from, int64_t n) {
        for (int64_t i = from + 1; i < n; ++i)
            if (!(a[i] > a[i - 1])) return false;
        return true;
    };
    for (int k = 0; k < t->n_transitions; ++k) {
        const korg_htransition& x = t->transitions[k];
        const int64_t nP = (int64_t)x.n_dnu * x.n_ne * x.n_T, nL = (int64_t)x.n_ne * x.n_T;
        if (x.lower < 1 || x.upper <= x.lower || x.upper > 64 || x.n_T < 2 || x.n_ne < 2 ||
            x.n_dnu < 2 || x.n_dnu > 1024 || !fits(x.off_T, x.n_T, t->n_temps) ||
            !fits(x.off_ne, x.n_ne, t->n_n_e) || !fits(x.off_dnu, x.n_dnu, t->n_log_dnu) ||
            !fits(x.off_profile, nP, t->n_log_profile) || !fits(x.off_lambda0, nL, t->n_lambda0) ||
            !ascending(t->temps + x.off_T, 0, x.n_T) || !ascending(t->n_e + x.off_ne, 0, x.n_ne) ||
            !ascending(t->log_dnu + x.off_dnu, 1, x.n_dnu))
            return set_err(ctx, KORG_ERR_INVALID_ARGUMENT,
                           "hydrogen tables: transition %d (%d -> %d) has invalid levels, axes or offsets",
                           k, x.lower, x.upper);
        tr[k] = {x.lower, x.upper, x.n_T, x.n_ne, x.n_dnu, 0, x.log_gf,
                 x.off_T, x.off_ne, x.off_dnu, x.off_profile, x.off_lambda0};
        max_dnu = std::max(max_dnu, x.n_dnu);
        n_max = std::max(n_max, x.upper);
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    ctx->hl_ntrans = 0;  // until the upload completes: a failed upload leaves no tables set
    struct {
        DevBuf* b;
        const void* src;
        size_t bytes;
    } up[] = {{&ctx->hl_trans, tr.data(), tr.size() * sizeof(HTrans)},
              {&ctx->hl_T, t->temps, (size_t)t->n_temps * 8},
              {&ctx->hl_ne, t->n_e, (size_t)t->n_n_e * 8},
              {&ctx->hl_dnu, t->log_dnu, (size_t)t->n_log_dnu * 8},
              {&ctx->hl_logP, t->log_profile, (size_t)t->n_log_profile * 8},
              {&ctx->hl_lam0, t->lambda0, (size_t)t->n_lambda0 * 8}};
    for (auto& u : up) {
        TRY(grow(ctx, *u.b, u.bytes));
        HIP_TRY(ctx, hipMemcpyAsync(u.b->p, u.src, u.bytes, hipMemcpyHostToDevice, ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->hl_ntrans = t->n_transitions;
    ctx->hl_max_dnu = max_dnu;
    ctx->hl_nmax = n_max;  // the Brackett series runs to the tables' highest upper level
    return KORG_OK;
}

int korg_hydrogen_line_absorption(korg_ctx* ctx, void* alpha, int alpha_dtype, const korg_grid* grid,
                                  const korg_hlayers* layers, double window_size, int use_MHD,
                                  int synchronize) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    if (ctx->hl_ntrans == 0)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "no hydrogen tables: call korg_set_hydrogen_tables first");
    if (!alpha || !grid || !grid->all_wls || grid->n_wl < 1 || grid->n_wl > INT32_MAX - HCT)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "alpha and a grid of 1 .. 2^31 points are required");
    if (alpha_dtype != KORG_DTYPE_F64 && alpha_dtype != KORG_DTYPE_F32)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "alpha_dtype must be KORG_DTYPE_F64 or _F32");
    if (!layers || layers->n_layers < 1 || layers->n_layers > 65535 || !layers->temps ||
        !layers->n_e || !layers->n_HI || !layers->n_HeI || !layers->U_HI || !layers->xi)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "hydrogen lines: layer arrays missing");
    if (!(window_size >= 0.0) || !std::isfinite(window_size))
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "window_size must be finite and >= 0");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int L = layers->n_layers, NT = ctx->hl_ntrans;
    const int n_br = std::max(0, ctx->hl_nmax - 4);  // Brackett lines 4 -> 5 .. n_max
    TRY(grow(ctx, ctx->hl_pairs, (size_t)L * NT * sizeof(HPair)));
    TRY(grow(ctx, ctx->hl_slices, (size_t)L * NT * ctx->hl_max_dnu * sizeof(double)));
    TRY(grow(ctx, ctx->hl_hull, (size_t)NT * 2 * sizeof(int)));
    TRY(grow(ctx, ctx->hl_brec, (size_t)L * n_br * sizeof(BRec)));
    TRY(grow(ctx, ctx->hl_bprof, (size_t)L * n_br * BR_POINTS * sizeof(double)));
    TRY(grow(ctx, ctx->hl_bhull, (size_t)std::max(n_br, 1) * 2 * sizeof(int)));
    HArgs a;
    a.L = L;
    a.n_trans = NT;
    a.max_dnu = ctx->hl_max_dnu;
    a.use_MHD = use_MHD ? 1 : 0;
    a.n_wl = grid->n_wl;
    a.window = window_size;
    a.wls = grid->all_wls;
    a.T = layers->temps;
    a.ne = layers->n_e;
    a.nH = layers->n_HI;
    a.nHe = layers->n_HeI;
    a.U = layers->U_HI;
    a.xi = layers->xi;
    a.trans = (const HTrans*)ctx->hl_trans.p;
    a.kT = (const double*)ctx->hl_T.p;
    a.kne = (const double*)ctx->hl_ne.p;
    a.kdnu = (const double*)ctx->hl_dnu.p;
    a.logP = (const double*)ctx->hl_logP.p;
    a.lam0 = (const double*)ctx->hl_lam0.p;
    a.pairs = (HPair*)ctx->hl_pairs.p;
    a.slices = (double*)ctx->hl_slices.p;
    a.hull = (int*)ctx->hl_hull.p;
    a.alpha = alpha;
    a.alpha_f32 = alpha_dtype == KORG_DTYPE_F32;
    a.n_max = ctx->hl_nmax;
    a.n_br = n_br;
    a.brec = (BRec*)ctx->hl_brec.p;
    a.bprof = (double*)ctx->hl_bprof.p;
    a.bhull = (int*)ctx->hl_bhull.p;
    const dim3 tiles((unsigned)((grid->n_wl + HCT - 1) / HCT), (L + 63) / 64);
    hipLaunchKernelGGL(k_hinit, dim3((NT + 63) / 64), dim3(64), 0, ctx->stream, a);
    hipLaunchKernelGGL(k_hsetup, dim3((L * NT + 63) / 64), dim3(64), 0, ctx->stream, a);
    hipLaunchKernelGGL(k_hlines, tiles, dim3(64), 0, ctx->stream, a);
    if (n_br > 0) {  // src/hydrogen_line_absorption.jl:152-169, after the tabulated lines
        hipLaunchKernelGGL(k_binit, dim3((n_br + 63) / 64), dim3(64), 0, ctx->stream, a);
        hipLaunchKernelGGL(k_bsetup, dim3((unsigned)(L * n_br)), dim3(256), 0, ctx->stream, a);
        hipLaunchKernelGGL(k_blines, tiles, dim3(64), 0, ctx->stream, a);
    }
    HIP_TRY(ctx, hipGetLastError());
    if (synchronize) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return KORG_OK;
}

}  // extern "C"

namespace {

int check_mol_args(korg_ctx* ctx, const void* alpha, int alpha_dtype, const korg_grid* grid,
                   int32_t n_layers, const double* temps, const double* vmic, int64_t n_tables,
                   const korg_mol_xsection* tables) {
    if (!alpha || !grid || !grid->all_wls || grid->n_wl < 1)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "alpha and a non-empty grid are required");
    if (alpha_dtype != KORG_DTYPE_F64 && alpha_dtype != KORG_DTYPE_F32)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "alpha_dtype must be KORG_DTYPE_F64 or _F32");
    if (n_layers < 1 || n_layers > 65535 * 64 || !temps || !vmic)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "molecular cross-sections: layers missing");
    if (n_tables < 0 || n_tables > 4096 || (n_tables > 0 && !tables))
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "n_tables out of [0, 4096] or tables null");
    for (int64_t k = 0; k < n_tables; ++k) {
        const korg_mol_xsection& t = tables[k];
        if (t.n_vmic < 2 || t.n_logT < 2 || t.n_wl < 2 || !t.vmic || !t.logT || !t.wls ||
            !t.sigma || !t.n_species)
            return set_err(ctx, KORG_ERR_INVALID_ARGUMENT,
                           "table %lld: at least 2 knots per dimension and every array", (long long)k);
    }
    return KORG_OK;
}

// The interpolation on device pointers (tables: a host array of structs of device pointers).
int run_mol(korg_ctx* ctx, void* alpha, int alpha_dtype, const korg_grid* grid, int32_t L,
            const double* temps, const double* vmic, int64_t n_tables,
            const korg_mol_xsection* tables) {
    if (n_tables == 0) return KORG_OK;  // src/molecular_cross_sections.jl:102-104
    std::vector<MolTable> h((size_t)n_tables);
    for (int64_t k = 0; k < n_tables; ++k) {
        const korg_mol_xsection& t = tables[k];
        h[k] = MolTable{t.n_vmic, t.n_logT, t.n_wl, t.vmic, t.logT, t.wls, t.sigma, t.n_species};
    }
    TRY(grow(ctx, ctx->mol_tabs, sizeof(MolTable) * h.size()));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->mol_tabs.p, h.data(), sizeof(MolTable) * h.size(),
                                hipMemcpyHostToDevice, ctx->stream));
    const dim3 g((unsigned)((grid->n_wl + MOL_TILE - 1) / MOL_TILE), (unsigned)((L + 63) / 64));
    hipLaunchKernelGGL(k_mol_interp, g, dim3(64), 0, ctx->stream, (int)L, grid->n_wl,
                       grid->all_wls, temps, vmic, (int)n_tables,
                       (const MolTable*)ctx->mol_tabs.p, alpha, alpha_dtype == KORG_DTYPE_F32 ? 1 : 0);
    HIP_TRY(ctx, hipGetLastError());
    return KORG_OK;
}

}  // namespace

extern "C" {

int korg_molecular_cross_sections_device(korg_ctx* ctx, void* alpha, int alpha_dtype,
                                         const korg_grid* grid, int32_t n_layers,
                                         const double* temps, const double* vmic,
                                         int64_t n_tables, const korg_mol_xsection* tables,
                                         int synchronize) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    TRY(check_mol_args(ctx, alpha, alpha_dtype, grid, n_layers, temps, vmic, n_tables, tables));
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    TRY(run_mol(ctx, alpha, alpha_dtype, grid, n_layers, temps, vmic, n_tables, tables));
    if (synchronize) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return KORG_OK;
}

int korg_molecular_cross_sections(korg_ctx* ctx, void* alpha, int alpha_dtype,
                                  const korg_grid* grid, int32_t n_layers, const double* temps,
                                  const double* vmic, int64_t n_tables,
                                  const korg_mol_xsection* tables) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    TRY(check_mol_args(ctx, alpha, alpha_dtype, grid, n_layers, temps, vmic, n_tables, tables));
    if (n_tables == 0) return KORG_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    // stage: grid, layers, alpha, then each table's arrays, packed into one buffer
    const int64_t L = n_layers, n = grid->n_wl;
    const size_t esz = alpha_dtype == KORG_DTYPE_F32 ? 4 : 8;
    size_t total = (size_t)(n + 2 * L) * 8 + esz * (size_t)(L * n);
    for (int64_t k = 0; k < n_tables; ++k) {
        const korg_mol_xsection& t = tables[k];
        total += 8 * (size_t)(t.n_vmic + t.n_logT + t.n_wl + L +
                              (int64_t)t.n_vmic * t.n_logT * t.n_wl) + 64 * 5;
    }
    TRY(grow(ctx, ctx->h_mol, total));
    char* base = (char*)ctx->h_mol.p;
    size_t off = 0;
    auto put = [&](const void* src, size_t bytes) -> void* {
        void* d = base + off;
        off += (bytes + 63) & ~(size_t)63;
        return hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, ctx->stream) == hipSuccess
                   ? d : nullptr;
    };
    korg_grid g{n, (const double*)put(grid->all_wls, 8 * (size_t)n)};
    const double* dT = (const double*)put(temps, 8 * (size_t)L);
    const double* dv = (const double*)put(vmic, 8 * (size_t)L);
    void* da = put(alpha, esz * (size_t)(L * n));
    std::vector<korg_mol_xsection> dt((size_t)n_tables);
    bool ok = g.all_wls && dT && dv && da;
    for (int64_t k = 0; k < n_tables && ok; ++k) {
        const korg_mol_xsection& t = tables[k];
        korg_mol_xsection& d = dt[k];
        d = t;
        d.vmic = (const double*)put(t.vmic, 8 * (size_t)t.n_vmic);
        d.logT = (const double*)put(t.logT, 8 * (size_t)t.n_logT);
        d.wls = (const double*)put(t.wls, 8 * (size_t)t.n_wl);
        d.sigma = (const double*)put(t.sigma, 8 * (size_t)t.n_vmic * t.n_logT * t.n_wl);
        d.n_species = (const double*)put(t.n_species, 8 * (size_t)L);
        ok = d.vmic && d.logT && d.wls && d.sigma && d.n_species;
    }
    if (!ok) {
        (void)hipGetLastError();
        return set_err(ctx, KORG_ERR_HIP, "staging the molecular cross-sections failed");
    }
    TRY(run_mol(ctx, da, alpha_dtype, &g, n_layers, dT, dv, n_tables, dt.data()));
    HIP_TRY(ctx, hipMemcpyAsync(alpha, da, esz * (size_t)(L * n), hipMemcpyDeviceToHost,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return KORG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// filter_linelist (src/synthesize.jl:331-350) for stars drawn from one master list in HBM.
// k_filter_search: one thread per (star, range) key pair, the two binary searches of the master's
// wavelengths; the host then applies the reference's no-double-count rule in range order, merges
// adjacent runs, and k_filter_gather copies the lines of stars with more than one run into the
// caller's pool (one thread per pool line, its run found by binary search over the run table).
// ---------------------------------------------------------------------------------------------
namespace {

__global__ __launch_bounds__(256) void k_filter_search(const double* wl, int64_t n,
                                                       const double* keys, int64_t n_pairs,
                                                       int64_t* found) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n_pairs) return;
    // searchsortedfirst (1-based) of first - buffer, searchsortedlast of last + buffer
    found[2 * t] = lower_bound_f64(wl, n, keys[2 * t]) + 1;
    found[2 * t + 1] = upper_bound_f64(wl, n, keys[2 * t + 1]);
}

struct FilterRun {
    int64_t src, dst;  // 0-based first master line and first pool line of a gathered run
};

__global__ __launch_bounds__(256) void k_filter_gather(korg_lines m, const FilterRun* runs,
                                                       int64_t n_runs, int64_t total,
                                                       int64_t cap, char* pool) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    int64_t lo = 0, hi = n_runs - 1;  // the last run with dst <= t (runs[0].dst == 0)
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (runs[mid].dst <= t) lo = mid;
        else hi = mid - 1;
    }
    const int64_t s = runs[lo].src + (t - runs[lo].dst);
    double* col = (double*)pool;
    const double* src[7] = {m.wl, m.log_gf, m.E_lower, m.gamma_rad, m.gamma_stark, m.vdW1, m.vdW2};
#pragma unroll
    for (int f = 0; f < 7; ++f) col[f * cap + t] = src[f][s];
    ((int32_t*)(col + 7 * cap))[t] = m.species[s];
}

}  // namespace

extern "C" {

int korg_filter_linelist_batched(korg_ctx* ctx, const korg_lines* master, int64_t n_stars,
                                 const korg_filter_star* stars, void* pool, int64_t pool_lines,
                                 korg_lines* out, int64_t* pool_lines_needed) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    if (!master || master->n_lines < 0 || n_stars < 0 || (n_stars > 0 && (!stars || !out)) ||
        pool_lines < 0)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "filter_linelist: master, stars and out are required");
    const korg_lines& m = *master;
    if (m.n_lines > 0 && (!m.wl || !m.log_gf || !m.E_lower || !m.gamma_rad || !m.gamma_stark ||
                          !m.vdW1 || !m.vdW2 || !m.species))
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "filter_linelist: a master column is missing");
    if (m.window_lb || m.window_ub)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT,
                       "filter_linelist: the master list cannot carry frozen windows");
    std::vector<double> keys;
    std::vector<int64_t> first_pair((size_t)n_stars + 1, 0);
    for (int64_t k = 0; k < n_stars; ++k) {
        const korg_filter_star& s = stars[k];
        if (s.n_ranges < 0 || (s.n_ranges > 0 && !s.ranges) || !std::isfinite(s.line_buffer))
            return set_err(ctx, KORG_ERR_INVALID_ARGUMENT,
                           "filter_linelist: star %lld needs its ranges and a finite buffer", (long long)k);
        for (int32_t r = 0; r < s.n_ranges; ++r) {
            const double a = s.ranges[2 * r], b = s.ranges[2 * r + 1];
            if (!(a <= b) || !std::isfinite(a) || !std::isfinite(b) ||
                (r > 0 && !(s.ranges[2 * r - 1] <= a)))
                return set_err(ctx, KORG_ERR_INVALID_ARGUMENT,
                               "filter_linelist: star %lld: ranges must be finite and ascending",
                               (long long)k);
            keys.push_back(a - s.line_buffer);  // src/synthesize.jl:336
            keys.push_back(b + s.line_buffer);  // :340
        }
        first_pair[k + 1] = first_pair[k] + s.n_ranges;
    }
    const int64_t n_pairs = first_pair[n_stars];
    std::vector<int64_t> found((size_t)(2 * n_pairs));
    if (n_pairs > 0 && m.n_lines > 0) {
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        TRY(grow(ctx, ctx->flt_keys, sizeof(double) * keys.size()));
        TRY(grow(ctx, ctx->flt_found, sizeof(int64_t) * found.size()));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->flt_keys.p, keys.data(), sizeof(double) * keys.size(),
                                    hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_filter_search, dim3((unsigned)((n_pairs + 255) / 256)), dim3(256), 0,
                           ctx->stream, m.wl, m.n_lines, (const double*)ctx->flt_keys.p, n_pairs,
                           (int64_t*)ctx->flt_found.p);
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, hipMemcpyAsync(found.data(), ctx->flt_found.p, sizeof(int64_t) * found.size(),
                                    hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        for (int64_t i = 0; i < n_pairs; ++i) {  // an empty master: searchsortedfirst 1, last 0
            found[2 * i] = 1;
            found[2 * i + 1] = 0;
        }
    }
    // the reference's selection, as 0-based runs [first, end) of the master per star
    struct Run { int64_t first, end; };
    std::vector<std::vector<Run>> runs((size_t)n_stars);
    std::vector<FilterRun> gather;
    int64_t need = 0;
    for (int64_t k = 0; k < n_stars; ++k) {
        int64_t last = 0;  // last_line_index, 1-based (src/synthesize.jl:334)
        for (int64_t p = first_pair[k]; p < first_pair[k + 1]; ++p) {
            const int64_t f = std::max(found[2 * p], last + 1);  // :338
            last = found[2 * p + 1];                             // :340
            if (f > last) continue;
            std::vector<Run>& rk = runs[k];
            if (!rk.empty() && rk.back().end == f - 1) rk.back().end = last;  // adjacent: one run
            else rk.push_back({f - 1, last});
        }
        if (runs[k].size() > 1)
            for (const Run& r : runs[k]) {
                gather.push_back({r.first, need});
                need += r.end - r.first;
            }
    }
    if (pool_lines_needed) *pool_lines_needed = need;
    if (need > pool_lines)
        return set_err(ctx, KORG_ERR_CAPACITY,
                       "filter_linelist: the pool holds %lld lines, %lld are needed",
                       (long long)pool_lines, (long long)need);
    if (need > 0 && !pool)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "filter_linelist: pool is NULL");
    if (need > 0) {
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        TRY(grow(ctx, ctx->flt_runs, sizeof(FilterRun) * gather.size()));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->flt_runs.p, gather.data(), sizeof(FilterRun) * gather.size(),
                                    hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_filter_gather, dim3((unsigned)((need + 255) / 256)), dim3(256), 0,
                           ctx->stream, m, (const FilterRun*)ctx->flt_runs.p,
                           (int64_t)gather.size(), need, pool_lines, (char*)pool);
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    const double* col = (const double*)pool;
    int64_t g = 0;
    for (int64_t k = 0; k < n_stars; ++k) {
        korg_lines o = m;
        o.window_lb = o.window_ub = nullptr;
        const std::vector<Run>& rk = runs[k];
        if (rk.size() <= 1) {  // a view into the master (nothing kept: zero lines at its start)
            const int64_t f = rk.empty() ? 0 : rk[0].first;
            o.n_lines = rk.empty() ? 0 : rk[0].end - f;
            if (m.n_lines > 0) {
                o.wl += f; o.log_gf += f; o.E_lower += f; o.gamma_rad += f; o.gamma_stark += f;
                o.vdW1 += f; o.vdW2 += f; o.species += f;
            }
        } else {
            const int64_t d = gather[g].dst;
            o.n_lines = 0;
            for (const Run& r : rk) o.n_lines += r.end - r.first;
            g += (int64_t)rk.size();
            const double* c[7];
            for (int f = 0; f < 7; ++f) c[f] = col + f * pool_lines + d;
            o.wl = c[0]; o.log_gf = c[1]; o.E_lower = c[2]; o.gamma_rad = c[3];
            o.gamma_stark = c[4]; o.vdW1 = c[5]; o.vdW2 = c[6];
            o.species = (const int32_t*)(col + 7 * pool_lines) + d;
        }
        out[k] = o;
    }
    return KORG_OK;
}

int korg_plan_launch(korg_plan* p, int synchronize) {
    if (!p) return KORG_ERR_INVALID_ARGUMENT;
    korg_ctx* ctx = p->ctx;
    ctx->err.clear();
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (!p->exec) {  // a forward plan without lines: nothing to do
        std::memset(&ctx->stats, 0, sizeof ctx->stats);
        return KORG_OK;
    }
    // the previous replay's counters are read before this one overwrites them
    if (p->pending) TRY(plan_wait(p));
    TRY(plan_launch(p));
    if (synchronize) TRY(plan_wait(p));
    return KORG_OK;
}

int korg_plan_wait(korg_plan* p) {
    if (!p) return KORG_ERR_INVALID_ARGUMENT;
    p->ctx->err.clear();
    HIP_TRY(p->ctx, hipSetDevice(p->ctx->device));
    return plan_wait(p);
}

int korg_plan_get_info(const korg_plan* p, korg_plan_info* out) {
    if (!p || !out) return KORG_ERR_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof *out);
    out->n_stars = (int64_t)p->stars.size();
    out->n_lines = p->b.N;
    out->n_points = p->b.nwl;
    out->tile_capacity = p->ws.tile_cap;
    out->tile_entries = p->b.f32 ? (int64_t)p->h_counters[13] : 0;
    out->n_captures = p->n_captures;
    out->tuning = p->tuning;
    return KORG_OK;
}

void korg_plan_destroy(korg_plan* p) { plan_free(p); }

int korg_autotune_batched(korg_ctx* ctx, int64_t n_stars, const korg_star* stars,
                          const korg_species* species, int alpha_dtype, int precision,
                          const char* cache_path, korg_tuning* out) {
    if (!ctx) return KORG_ERR_INVALID_ARGUMENT;
    ctx->err.clear();
    std::vector<StarHost> hs;
    int L = 0;
    TRY(collect_stars(ctx, n_stars, stars, species, alpha_dtype, precision, hs, &L));
    if (layer_groups(hs).size() > 1)
        return set_err(ctx, KORG_ERR_UNSUPPORTED,
                       "autotuning needs one layer count for all the stars (tune per count)");
    int64_t N = 0, nwl = 0;
    for (const StarHost& x : hs) {
        N += x.lines.n_lines;
        nwl += x.grid.n_wl;
    }
    const korg_tuning keep = ctx->tuning;
    const std::string key = tune_key(ctx, precision, alpha_dtype, L, n_stars, N, nwl);
    korg_tuning best{};
    if (cache_path && cache_lookup(cache_path, key, &best)) {
        best.count_branches = keep.count_branches;
        ctx->tuning = best;
        if (out) *out = best;
        return KORG_OK;
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (N == 0) {  // nothing to time
        if (out) *out = keep;
        return KORG_OK;
    }
    // the candidates add into a scratch copy of the alphas, never into the caller's
    const size_t esz = alpha_dtype == KORG_DTYPE_F32 ? 4 : 8;
    size_t bytes = 0;
    for (const StarHost& x : hs) bytes += esz * (size_t)L * (size_t)x.grid.n_wl;
    DevBuf scratch;
    TRY(grow(ctx, scratch, bytes));
    size_t off = 0;
    for (StarHost& x : hs) {
        x.alpha = (char*)scratch.p + off;
        off += esz * (size_t)L * (size_t)x.grid.n_wl;
    }
    int rc = KORG_OK;
    hipError_t e = hipMemsetAsync(scratch.p, 0, bytes, ctx->stream);
    if (e != hipSuccess) rc = set_err(ctx, KORG_ERR_HIP, "hipMemsetAsync: %s", hipGetErrorString(e));
    double ms = 0;
    if (rc == KORG_OK) rc = autotune(ctx, hs, species, L, alpha_dtype, precision, &best, &ms);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(scratch.p);
    ctx->ws.release();  // sized for the largest candidate; the next call grows what it needs
    ctx->last_n_lines = 0;
    ctx->tuning = keep;
    if (rc != KORG_OK) return rc;
    best.count_branches = keep.count_branches;
    if (cache_path) TRY(cache_store(ctx, cache_path, key, best, ms));
    ctx->tuning = best;
    std::memset(&ctx->stats, 0, sizeof ctx->stats);
    ctx->stats.total_ms = ms;
    if (out) *out = best;
    return KORG_OK;
}

}  // extern "C"
