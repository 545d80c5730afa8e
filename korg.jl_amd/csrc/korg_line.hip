// MI355X (gfx950) line-opacity library: the body of Korg.jl's line_absorption!
// (src/line_absorption.jl:33-129) as HIP kernels behind the C ABI of include/korg_line.h.
//
// Pipeline of one call (all on one stream, inputs resident in HBM):
//   k_layers     per-layer tables: beta = 1/(kT), n_eff_vdW, (T/1e4)^(1/6), (T/1e4)^0.3
//   k_setup      one wavefront per line, lanes = layers (Float64): widths, amplitudes, the window
//                (max over layers), per-(line, layer) profile records, width class
//   k_count / k_scan / k_scatter   stable partition of the lines into window-width classes
//   k_ranges     per (class, wavelength tile): index range of candidate lines
//   k_profile_*  persistent wavefronts take (tile, layer block) items from a queue; each owns its
//                alpha block outright, adds the overlapping lines in a fixed order in registers,
//                then adds the block into alpha once.  No atomics on alpha: deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "korg_line.h"
#include "korg_line_device.h"

using namespace korg;

namespace {

constexpr int WAVE = 64;

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

__device__ inline double wave_max_nan(double x) {
    // Julia's maximum: NaN as soon as one element is NaN.
    if (ballot(x != x)) return __builtin_nan("");
    for (int off = 32; off >= 1; off >>= 1) {
        double y = __shfl_xor(x, off, WAVE);
        x = y > x ? y : x;
    }
    return x;
}

// First index i in [0, n) with a[i] >= x (n if none): a 64-way search, every lane probes.
// Requires a ascending and the whole wavefront active.
template <class T>
__device__ inline int64_t wave_lower_bound(const T* __restrict__ a, int64_t n, T x, int lane) {
    int64_t lo = 0, hi = n;  // answer in [lo, hi]
    while (hi - lo > WAVE) {
        const int64_t step = (hi - lo + WAVE - 1) / WAVE;
        const int64_t p = lo + (int64_t)(lane + 1) * step - 1;
        const bool below = p < hi && a[p] < x;
        const int k = __popcll(ballot(below));
        const int64_t phi = lo + (int64_t)(k + 1) * step - 1;
        hi = (k < WAVE && phi < hi) ? phi : hi;
        lo = lo + (int64_t)k * step;
    }
    const int64_t p = lo + lane;
    const bool below = p < hi && a[p] < x;
    return lo + __popcll(ballot(below));
}
// First index i with a[i] > x (n if none).
template <class T>
__device__ inline int64_t wave_upper_bound(const T* __restrict__ a, int64_t n, T x, int lane) {
    int64_t lo = 0, hi = n;
    while (hi - lo > WAVE) {
        const int64_t step = (hi - lo + WAVE - 1) / WAVE;
        const int64_t p = lo + (int64_t)(lane + 1) * step - 1;
        const bool le = p < hi && a[p] <= x;
        const int k = __popcll(ballot(le));
        const int64_t phi = lo + (int64_t)(k + 1) * step - 1;
        hi = (k < WAVE && phi < hi) ? phi : hi;
        lo = lo + (int64_t)k * step;
    }
    const int64_t p = lo + lane;
    const bool le = p < hi && a[p] <= x;
    return lo + __popcll(ballot(le));
}
__device__ inline int64_t wave_lower_bound_i32(const int* a, int64_t n, int x, int lane) {
    return wave_lower_bound<int>(a, n, x, lane);
}
__device__ inline int64_t wave_upper_bound_i32(const int* a, int64_t n, int x, int lane) {
    return wave_upper_bound<int>(a, n, x, lane);
}

// searchsortedfirst / searchsortedlast on a Wavelengths (src/wavelengths.jl:188-210), 0-based:
// values >= 1 are read as Angstrom; NaN gives an empty window (first = n, last = n - 1).
__device__ inline int64_t ss_first0(const double* wls, int64_t n, double x, int lane) {
    if (x >= 1.0) x = x * 1e-8;
    if (!(x <= wls[n - 1])) return n;
    return wave_lower_bound<double>(wls, n, x, lane);
}
__device__ inline int64_t ss_last0(const double* wls, int64_t n, double x, int lane) {
    if (x >= 1.0) x = x * 1e-8;
    if (!(x <= wls[n - 1])) return n - 1;
    return wave_upper_bound<double>(wls, n, x, lane) - 1;
}

// --------------------------------------------------------------------------------------------
// k_layers: per-layer tables (src/line_absorption.jl:39, 54-60, 195, 210) and queue reset
// --------------------------------------------------------------------------------------------
struct LayerTab {
    double beta, n_eff, t16, t03;
};

// Device counters of one call (reset by k_layers): [0] profile queue, [1] evaluations,
// [2] lines used, [3] error flags, [4] core reach, [5] core queue, [6..9] evaluations by branch.
constexpr int N_COUNTERS = 16;

__global__ void k_layers(int L, const double* __restrict__ temps, const double* __restrict__ n_HI,
                         const double* __restrict__ n_HeI, const double* __restrict__ n_H2,
                         LayerTab* __restrict__ tab, unsigned long long* __restrict__ counters) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l == 0) {
        counters[0] = 0;  // profile work queue
        counters[1] = 0;  // evaluation count
        counters[2] = 0;  // lines used
        counters[3] = 0;  // set when the line list is not sorted by wavelength
        counters[4] = 0;  // max distance (grid points) from jc to a core point
        counters[5] = 0;  // core work queue
        counters[6] = 0;  // evaluations in branch A (far wing, a <= 0.2)
        counters[7] = 0;  //                      B (core, a <= 0.2)
        counters[8] = 0;  //                      C (core, 0.2 < a <= 1.4)
        counters[9] = 0;  //                      D (far wing, a > 0.2)
    }
    if (l >= L) return;
    const double a_H = 4.50711, a_He = 1.38375, a_H2 = 5.503;
    const double m_H = 1.008 * amu_cgs, m_He = 4.003 * amu_cgs;
    const double c_He = pow(a_He / a_H, 0.4) * pow(m_He / m_H, -0.3);
    const double c_H2 = pow(a_H2 / a_H, 0.4) * pow(2.0, -0.3);
    const double T = temps[l];
    LayerTab t;
    t.beta = 1.0 / (kboltz_eV * T);
    t.n_eff = n_HI[l] + c_He * n_HeI[l] + c_H2 * n_H2[l];
    t.t16 = pow(T / 10000.0, 1.0 / 6.0);
    t.t03 = pow(T / 10000.0, 0.3);
    tab[l] = t;
}

// --------------------------------------------------------------------------------------------
// k_setup: one wavefront per line; lane l handles layers l, l+64, ...
// --------------------------------------------------------------------------------------------
struct SetupArgs {
    int64_t n_lines, n_wl;
    int L;
    int Ls;  // layers per line in the record arrays (L padded with zero records)
    const double* wls;
    const double *wl, *log_gf, *E_lower, *gamma_rad, *gamma_stark, *vdW1, *vdW2;
    const int32_t* species;
    int n_species;
    const double* mass;
    const uint8_t* is_mol;
    const double* n_div_U;
    const double *temps, *n_e, *xi;
    const LayerTab* tab;
    int cntm_mode;
    int64_t n_knots;
    const double *knots, *cntm;
    double cutoff;
    int want_rec32;
    int n_lblocks;  // layer blocks (each <= 64 layers, aligned with 64-layer chunks)
    int count_branches;
    LineMeta* meta;
    int* jc_all;
    double4* rec64;  // [line * L + l]: (inv_sigma_sqrt2, a, scaling, 0)
    float4* recA;    // [line * Ls + l]: (d1, d2, d3, 0) for a <= 0.2, (c1, c2, c3, c4) otherwise
    float2* recD;    // [line * Ls + l]: (c5, 1/g2) for a > 0.2
    uint64_t* dmask; // [line * n_lblocks + b]: bit k set when layer 64b+k has a > 0.2
    unsigned long long* counters;
};

// Continuum at x for layer l on the coarse grid: linear interpolation between the bracketing
// knots (index found once per line by the caller).
__device__ inline double cntm_lerp(const double* knots, const double* vals, int64_t i, double x) {
    const double t = (x - knots[i]) / (knots[i + 1] - knots[i]);
    return (1.0 - t) * vals[i] + t * vals[i + 1];
}

__global__ __launch_bounds__(256) void k_setup(SetupArgs A) {
    const int lane = threadIdx.x % WAVE;
    const int64_t i = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (i >= A.n_lines) return;  // whole wavefront
    const int L = A.L;
    const double wl = A.wl[i];
    const int32_t s = A.species[i];
    if (s < 0 || s >= A.n_species) {
        // invalid species index: flag the call as failed (nothing is added to alpha)
        if (lane == 0) {
            atomicOr(&A.counters[3], 2ull);
            LineMeta mt = {wl, 1, 0, 1, 0, 0, N_CLASSES};
            A.meta[i] = mt;
            A.jc_all[i] = 0;
        }
        return;
    }
    const double m = A.mass[s];
    const bool mol = A.is_mol[s] != 0;
    const double gamma_rad = A.gamma_rad[i], gamma_stark = A.gamma_stark[i];
    const double vdW1 = A.vdW1[i], vdW2 = A.vdW2[i];
    const double E_lower = A.E_lower[i];
    if (lane == 0 && i > 0 && A.wl[i - 1] > wl) atomicOr(&A.counters[3], 1ull);
    const double E_upper = E_lower + c_cgs * hplanck_eV / wl;  // src/line_absorption.jl:101
    const double gf = pow(10.0, A.log_gf[i]);
    const double sl = sigma_line(wl);
    int64_t kidx = 0;
    if (A.cntm_mode == KORG_CNTM_GRID) {
        // knot interval holding wl: clamp(searchsortedlast(knots, wl), 1, n_knots - 1) - 1
        kidx = wave_upper_bound<double>(A.knots, A.n_knots, wl, lane) - 1;
        kidx = kidx < 0 ? 0 : (kidx > A.n_knots - 2 ? A.n_knots - 2 : kidx);
    }
    double dmax = 0.0, lmax = 0.0, hwmax = -1.0;
    bool any_nan = false;
    double my_hw = -1.0, my_a = 0.0;  // layer `lane` (first 64 layers), for branch counting
    for (int l = L + lane; l < A.Ls; l += WAVE) {
        A.rec64[i * A.Ls + l] = make_double4(0.0, 0.0, 0.0, 0.0);
        if (A.want_rec32) {
            A.recA[i * A.Ls + l] = make_float4(0.f, 0.f, 0.f, 0.f);
            A.recD[i * A.Ls + l] = make_float2(0.f, 0.f);
        }
    }
    for (int l = lane; l < L; l += WAVE) {
        const LayerTab tb = A.tab[l];
        const double T = A.temps[l];
        // src/line_absorption.jl:86-106
        const double sigma = doppler_width(wl, T, m, A.xi[l]);
        double G = gamma_rad;
        if (!mol) {
            G = G + A.n_e[l] * (gamma_stark * tb.t16);
            const double vdw = (vdW2 == -1.0) ? vdW1 * tb.t03 : scaled_vdW_abo(vdW1, vdW2, m, T);
            G = G + tb.n_eff * vdw;
        }
        const double gamma = G * (wl * wl) / (c_cgs * (4.0 * PI));
        const double levels = exp(-tb.beta * E_lower) - exp(-tb.beta * E_upper);
        const double amp = gf * sl * levels * A.n_div_U[(int64_t)s * L + l];
        // src/line_absorption.jl:108-112
        const double c = (A.cntm_mode == KORG_CNTM_GRID)
                              ? cntm_lerp(A.knots, A.cntm + (int64_t)l * A.n_knots, kidx, wl)
                              : A.cntm[i * L + l];
        const double rho = c * A.cutoff / amp;
        const double dg = inverse_gaussian_density(rho, sigma);
        const double dl = inverse_lorentz_density(rho, gamma);
        any_nan |= (dg != dg) || (dl != dl);
        dmax = dg > dmax ? dg : dmax;
        lmax = dl > lmax ? dl : lmax;
        // profile parameters of line_profile (src/line_absorption.jl:245-249)
        const double inv = 1.0 / (sigma * 1.4142135623730951);
        const double a = gamma * inv;
        const double scaling = inv / 1.7724538509055159 * amp;
        A.rec64[i * A.Ls + l] = make_double4(inv, a, scaling, 0.0);
        // half-width (cm) of the Voigt-core region of this layer (branches B and C)
        const double hw = (a <= 0.2) ? 5.0 / inv : (a <= 1.4 ? (3.2 - a) / inv : -1.0);
        hwmax = hw > hwmax ? hw : hwmax;
        if (l == lane) {
            my_hw = hw;
            my_a = a;
        }
        if (A.want_rec32) {
            // Float32 far-wing coefficients, formed in Float64 (offsets in DELTA_UNIT):
            //   v = |delta| s, s = DELTA_UNIT * inv;  base = scaling * a / sqrt(pi)
            //   A (a <= 0.2): H = (a/sqrt(pi)) w (1 + 1.5 w + 3.75 w^2), w = 1/v^2 = R/s^2
            //   D (a > 0.2):  H = (a/sqrt(pi)) sum_k C_k r^k, r = 1/(v^2 + a^2) = rho'/a^2,
            //                 C = (1, 1.5, 3.75 - 2a^2, -15 a^2, 11.75 a^4),
            //                 rho' = 1/(delta^2/g2 + 1), g2 = (a/s)^2
            const double sdel = DELTA_UNIT * inv;
            const double base = scaling * (a / 1.7724538509055159);
            const bool kind_d = !(a <= 0.2);
            if (!kind_d) {
                const double q = 1.0 / (sdel * sdel);
                A.recA[i * A.Ls + l] = make_float4((float)(base * q), (float)(1.5 * base * q * q),
                                                   (float)(3.75 * base * q * q * q), 0.f);
                A.recD[i * A.Ls + l] = make_float2(0.f, 0.f);
            } else {
                const double ia2 = 1.0 / (a * a);
                const double g2 = a / sdel * (a / sdel);
                const double ia6 = ia2 * ia2 * ia2;
                A.recA[i * A.Ls + l] =
                    make_float4((float)(base * ia2), (float)(1.5 * base * ia2 * ia2),
                                (float)((3.75 - 2.0 * a * a) * base * ia6), (float)(-15.0 * base * ia6));
                A.recD[i * A.Ls + l] = make_float2((float)(11.75 * base * ia6), (float)(1.0 / g2));
            }
            const uint64_t dm = ballot(kind_d);
            if (lane == 0) A.dmask[i * A.n_lblocks + l / WAVE] = dm;
        }
    }
    // src/line_absorption.jl:109-115: windows are maxima over layers
    const bool wnan = ballot(any_nan) != 0;
    double dwin = wave_max_nan(dmax);
    double lwin = wave_max_nan(lmax);
    double hw = hwmax;
    for (int off = 32; off >= 1; off >>= 1) {
        double y = __shfl_xor(hw, off, WAVE);
        hw = y > hw ? y : hw;
    }
    const double window = wnan ? __builtin_nan("") : sqrt(lwin * lwin + dwin * dwin);
    const int64_t n = A.n_wl;
    const int64_t lb = ss_first0(A.wls, n, wl - window, lane);
    const int64_t ub = ss_last0(A.wls, n, wl + window, lane);
    const int64_t jc = wave_lower_bound<double>(A.wls, n, wl, lane);
    int64_t clo = 1, chi = 0;
    if (lb <= ub && hw >= 0.0) {
        // conservative superset of the core region; grid points outside it are far wing in
        // every layer
        const double h = hw * (1.0 + 1e-6) + 1e-6 * DELTA_UNIT;
        clo = wave_lower_bound<double>(A.wls, n, wl - h, lane);
        chi = wave_upper_bound<double>(A.wls, n, wl + h, lane) - 1;
        clo = clo < lb ? lb : clo;
        chi = chi > ub ? ub : chi;
    }
    if (A.count_branches && lb <= ub && lane < L && lane < WAVE) {
        // evaluations of this layer by branch: core points lie within my_hw of lam0
        int64_t core_n = 0;
        if (my_hw >= 0.0) {
            int64_t lo = 0, hi = n;
            const double xl = wl - my_hw, xr = wl + my_hw;
            while (lo < hi) {  // first index with wls >= xl... strictly inside: > xl
                const int64_t mid = lo + (hi - lo) / 2;
                if (A.wls[mid] <= xl) lo = mid + 1; else hi = mid;
            }
            int64_t c0 = lo;
            lo = 0;
            hi = n;
            while (lo < hi) {  // first index with wls >= xr
                const int64_t mid = lo + (hi - lo) / 2;
                if (A.wls[mid] < xr) lo = mid + 1; else hi = mid;
            }
            int64_t c1 = lo - 1;
            c0 = c0 < lb ? lb : c0;
            c1 = c1 > ub ? ub : c1;
            core_n = c1 >= c0 ? c1 - c0 + 1 : 0;
        }
        const int64_t wing_n = (ub - lb + 1) - core_n;
        const bool small_a = my_a <= 0.2;
        atomicAdd(&A.counters[small_a ? 6 : 9], (unsigned long long)wing_n);
        if (core_n) atomicAdd(&A.counters[small_a ? 7 : 8], (unsigned long long)core_n);
    }
    int cls = N_CLASSES;
    if (lb <= ub) {
        const int64_t W = ub - lb + 1;
        cls = 0;
        while (cls < N_CLASSES - 1 && W > class_wmax(cls)) ++cls;
    }
    if (lane == 0) {
        LineMeta mt;
        mt.lam0 = wl;
        mt.lb = (int32_t)lb;
        mt.ub = (int32_t)ub;
        mt.clo = (int32_t)clo;
        mt.chi = (int32_t)chi;
        mt.jc = (int32_t)jc;
        mt.cls = cls;
        A.meta[i] = mt;
        A.jc_all[i] = (int32_t)jc;
        if (clo <= chi) {
            const int64_t d = max(jc - clo, chi - jc) + 1;
            atomicMax(&A.counters[4], (unsigned long long)d);
        }
        if (lb <= ub) {
            atomicAdd(&A.counters[1], (unsigned long long)((ub - lb + 1) * (int64_t)L));
            atomicAdd(&A.counters[2], 1ull);
        }
    }
}

// --------------------------------------------------------------------------------------------
// Stable partition of lines by class: counts per block, scan, scatter (deterministic).
// --------------------------------------------------------------------------------------------
constexpr int PART_BLOCK = 256;

__global__ __launch_bounds__(PART_BLOCK) void k_count(int64_t n, const LineMeta* __restrict__ meta,
                                                      int* __restrict__ block_counts) {
    __shared__ int cnt[N_CLASSES + 1];
    if (threadIdx.x <= N_CLASSES) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
    const int c = i < n ? meta[i].cls : N_CLASSES;
    for (int k = 0; k < N_CLASSES; ++k) {
        const uint64_t b = ballot(c == k);
        if ((threadIdx.x % WAVE) == 0 && b) atomicAdd(&cnt[k], __popcll(b));
    }
    __syncthreads();
    if (threadIdx.x < N_CLASSES) block_counts[(int64_t)blockIdx.x * N_CLASSES + threadIdx.x] = cnt[threadIdx.x];
}

// One block: exclusive scan of block_counts per class; class_start[c] = first slot of class c,
// class_start[N_CLASSES] = total.
__global__ __launch_bounds__(1024) void k_scan(int nblocks, int* __restrict__ block_counts,
                                               int* __restrict__ class_start) {
    __shared__ int part[1024];
    __shared__ int totals[N_CLASSES];
    const int t = threadIdx.x;
    for (int c = 0; c < N_CLASSES; ++c) {
        // each thread owns a contiguous run of blocks
        const int per = (nblocks + 1023) / 1024;
        const int b0 = t * per, b1 = min(nblocks, b0 + per);
        int s = 0;
        for (int b = b0; b < b1; ++b) s += block_counts[(int64_t)b * N_CLASSES + c];
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
            const int v = block_counts[(int64_t)b * N_CLASSES + c];
            block_counts[(int64_t)b * N_CLASSES + c] = run;
            run += v;
        }
        if (t == 1023) totals[c] = part[t];
        __syncthreads();
    }
    if (t == 0) {
        int acc = 0;
        for (int c = 0; c < N_CLASSES; ++c) {
            class_start[c] = acc;
            acc += totals[c];
        }
        class_start[N_CLASSES] = acc;
    }
}

__global__ __launch_bounds__(PART_BLOCK) void k_scatter(int64_t n, const LineMeta* __restrict__ meta,
                                                        const int* __restrict__ block_offsets,
                                                        const int* __restrict__ class_start,
                                                        int* __restrict__ cls_lines,
                                                        int* __restrict__ cls_jc) {
    __shared__ int wave_cnt[PART_BLOCK / WAVE][N_CLASSES];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int64_t i = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
    const int c = i < n ? meta[i].cls : N_CLASSES;
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
        cls_lines[pos] = (int)i;
        cls_jc[pos] = meta[i].jc;
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

// --------------------------------------------------------------------------------------------
// profile kernels
// --------------------------------------------------------------------------------------------
struct ProfileArgs {
    int64_t n_wl;
    int L;
    int Ls;  // record stride per line (layers padded to whole blocks)
    int n_lblocks;
    int ntiles, tile_w;
    int64_t n_items;
    const double* wls;
    const LineMeta* meta;
    const double4* rec64;
    const float4* recA;
    const float2* recD;
    const uint64_t* dmask;
    const int* cls_lines;
    const int2* ranges;
    void* alpha;
    int alpha_f32;
    unsigned long long* counters;
};

// Visit, in a fixed order (class by class, line index within a class), every line whose window
// overlaps grid points [j0, j1] of this tile; calls f(line) with a wave-uniform line index.
template <class F>
__device__ inline void for_each_line(const ProfileArgs& A, int tile, int64_t j0, int64_t j1,
                                     int lane, F&& f) {
    for (int c = 0; c < N_CLASSES; ++c) {
        const int2 r = A.ranges[(int64_t)c * A.ntiles + tile];
        for (int k0 = r.x; k0 < r.y; k0 += WAVE) {
            const int k = k0 + lane;
            int line = 0;
            bool ovl = false;
            if (k < r.y) {
                line = A.cls_lines[k];
                const LineMeta m = A.meta[line];
                ovl = m.lb <= j1 && m.ub >= j0;
            }
            uint64_t mask = ballot(ovl);
            while (mask) {
                const int b = __ffsll((unsigned long long)mask) - 1;
                mask &= mask - 1;
                f(readlane_i32(line, b));
            }
        }
    }
}

__device__ inline int wave_take(unsigned long long* counter, int lane) {
    unsigned long long v = 0;
    if (lane == 0) v = atomicAdd(counter, 1ull);
    return __builtin_amdgcn_readfirstlane((int)v);
}

// Float64 form: the reference's line_profile / voigt_hjerting per evaluation.  A work item is
// (tile, group of G layers); the G x P accumulators stay in registers.
template <int G, int P>
__global__ __launch_bounds__(WAVE) void k_profile_f64(ProfileArgs A) {
    const int lane = threadIdx.x;
    const int L = A.L;
    if (A.counters[3] != 0) return;  // unsorted line list: nothing is written
    for (;;) {
        const int item = wave_take(&A.counters[0], lane);
        if (item >= A.n_items) break;
        const int tile = item / A.n_lblocks;
        const int l0 = (item % A.n_lblocks) * G;
        const int64_t j0 = (int64_t)tile * A.tile_w;
        const int64_t j1 = min(j0 + A.tile_w, A.n_wl) - 1;
        double lam[P];
        int jj[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            jj[p] = (int)(j0 + p * WAVE + lane);
            lam[p] = jj[p] <= j1 ? A.wls[jj[p]] : 0.0;
        }
        double acc[P][G];
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int l = 0; l < G; ++l) acc[p][l] = 0.0;

        for_each_line(A, tile, j0, j1, lane, [&](int line) {
            const KCONST LineMeta* mp = uniform_ptr(A.meta + line);
            const double lam0 = mp->lam0;
            const int lb = mp->lb, ub = mp->ub;
            double dl[P];
            bool inw[P];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                inw[p] = jj[p] >= lb && jj[p] <= ub;
                dl[p] = fabs(lam[p] - lam0);  // abs(λ - λ₀), src/line_absorption.jl:248
            }
            const KCONST double4* rp = uniform_ptr(A.rec64 + (int64_t)line * A.Ls + l0);
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
            for (int l = 0; l < G; ++l) {
                if (l0 + l < L) {
                    const int64_t k = (int64_t)(l0 + l) + (int64_t)L * jj[p];
                    if (A.alpha_f32) ((float*)A.alpha)[k] += (float)acc[p][l];
                    else ((double*)A.alpha)[k] += acc[p][l];
                }
            }
        }
    }
}

// Which Hunger branch (src/line_absorption.jl:286-310) applies, decided in Float64 exactly as
// voigt_hjerting does: true for the Voigt-core branches (B: a <= 0.2 and v < 5;
// C: 0.2 < a <= 1.4 and a + v < 3.2), false for the far-wing branches (A and D).
__device__ inline bool is_core_branch(double a, double v) {
    if (a <= 0.2) return !(v >= 5.0);
    if (a <= 1.4) return a + v < 3.2;
    return false;
}

// Float32 form with Float64 fix-ups, far-wing branches only (A and D of voigt_hjerting).
// Lanes = wavelength points, LB layers per item in registers.  Where a Voigt core may lie in the
// wavefront's points, core evaluations are masked out with the exact Float64 branch test; they
// are added by k_core_f32.
template <int LB, int P>
__global__ __launch_bounds__(WAVE) void k_profile_f32(ProfileArgs A) {
    const int lane = threadIdx.x;
    const int L = A.L;
    if (A.counters[3] != 0) return;  // unsorted line list: nothing is written
    for (;;) {
        const int item = wave_take(&A.counters[0], lane);
        if (item >= A.n_items) break;
        const int tile = item / A.n_lblocks;
        const int l0 = (item % A.n_lblocks) * LB;
        const int64_t j0 = (int64_t)tile * A.tile_w;
        const int64_t j1 = min(j0 + A.tile_w, A.n_wl) - 1;
        double lam[P];
        int jj[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            jj[p] = (int)(j0 + p * WAVE + lane);
            lam[p] = jj[p] <= j1 ? A.wls[jj[p]] : 0.0;
        }
        float acc[P][LB];
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int l = 0; l < LB; ++l) acc[p][l] = 0.0f;

        for_each_line(A, tile, j0, j1, lane, [&](int line) {
            const KCONST LineMeta* mp = uniform_ptr(A.meta + line);
            const double lam0 = mp->lam0;
            const int lb = mp->lb, ub = mp->ub, clo = mp->clo, chi = mp->chi;
            // Float64 fix-up: lambda_j - lambda0 from the exact Float64 grid values, scaled and
            // rounded once to Float32.  R = 1/delta^2 is shared by all layers.
            double dl[P];
            float d2[P], R[P], R2[P], R3[P];
            bool inw[P];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                inw[p] = jj[p] >= lb && jj[p] <= ub;
                dl[p] = lam[p] - lam0;
                const float d = (float)(dl[p] * (1.0 / DELTA_UNIT));
                const float dd = d * d;
                const float r = __builtin_amdgcn_rcpf(dd);
                R[p] = inw[p] ? r : 0.0f;
                R2[p] = R[p] * R[p];
                R3[p] = R2[p] * R[p];
                d2[p] = inw[p] ? dd : __builtin_huge_valf();
            }
            const bool core = clo <= chi && !(chi < j0 || clo > j1);
            const KCONST float4* ra = uniform_ptr(A.recA + (int64_t)line * A.Ls + l0);
            const KCONST float2* rd = uniform_ptr(A.recD + (int64_t)line * A.Ls + l0);
            const uint64_t dm = *uniform_ptr(A.dmask + (int64_t)line * A.n_lblocks + l0 / WAVE);
            if (!core) {
                // records of 8 layers are fetched together: one scalar-load wait per 8 layers
#pragma unroll
                for (int g = 0; g < LB; g += 8) {
                    float4 r[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) r[k] = ra[g + k];
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const int l = g + k;
                        if (!((dm >> l) & 1)) {
#pragma unroll
                            for (int p = 0; p < P; ++p) {
                                float a = acc[p][l];
                                a = fmaf(r[k].x, R[p], a);
                                a = fmaf(r[k].y, R2[p], a);
                                a = fmaf(r[k].z, R3[p], a);
                                acc[p][l] = a;
                            }
                        } else {
                            const float2 e = rd[l];
#pragma unroll
                            for (int p = 0; p < P; ++p) {
                                const float rho = __builtin_amdgcn_rcpf(fmaf(d2[p], e.y, 1.0f));
                                float h = fmaf(rho, e.x, r[k].w);
                                h = fmaf(h, rho, r[k].z);
                                h = fmaf(h, rho, r[k].y);
                                h = fmaf(h, rho, r[k].x);
                                acc[p][l] = fmaf(h, rho, acc[p][l]);
                            }
                        }
                    }
                }
            } else {
                double adl[P];
#pragma unroll
                for (int p = 0; p < P; ++p) adl[p] = fabs(dl[p]);
                const KCONST double4* rq = uniform_ptr(A.rec64 + (int64_t)line * A.Ls + l0);
#pragma unroll
                for (int l = 0; l < LB; ++l) {
                    const float4 r = ra[l];
                    const double4 q = rq[l];
                    const bool kind_d = (dm >> l) & 1;
                    float2 e = make_float2(0.f, 0.f);
                    if (kind_d) e = rd[l];
#pragma unroll
                    for (int p = 0; p < P; ++p) {
                        float w;
                        if (!kind_d) {
                            w = fmaf(r.z, R3[p], fmaf(r.y, R2[p], r.x * R[p]));
                        } else {
                            const float rho = __builtin_amdgcn_rcpf(fmaf(d2[p], e.y, 1.0f));
                            float h = fmaf(rho, e.x, r.w);
                            h = fmaf(h, rho, r.z);
                            h = fmaf(h, rho, r.y);
                            h = fmaf(h, rho, r.x);
                            w = h * rho;
                        }
                        const bool wing = inw[p] && !is_core_branch(q.y, adl[p] * q.x);
                        acc[p][l] += wing ? w : 0.0f;
                    }
                }
            }
        });

#pragma unroll
        for (int p = 0; p < P; ++p) {
            if (jj[p] > j1) continue;
#pragma unroll
            for (int l = 0; l < LB; ++l) {
                if (l0 + l < L) {
                    const int64_t k = (int64_t)(l0 + l) + (int64_t)L * jj[p];
                    if (A.alpha_f32) ((float*)A.alpha)[k] += acc[p][l];
                    else ((double*)A.alpha)[k] += (double)acc[p][l];
                }
            }
        }
    }
}

// Voigt-core evaluations of the Float32 form (branches B and C), lanes = layers.  A work item is
// (tile of CT points, block of 64 layers); for every line whose core range [clo, chi] meets the
// tile, in line order, each lane evaluates its layer at the tile's points that pass the exact
// Float64 branch test.  Runs after k_profile_f32 and adds its own sums into alpha.
constexpr int CT = 32;

struct CoreArgs {
    int64_t n_wl, n_lines;
    int L;
    int Ls;
    int n_lblocks, ntiles;
    int64_t n_items;
    const double* wls;
    const LineMeta* meta;
    const int* jc_all;
    const double4* rec64;
    void* alpha;
    int alpha_f32;
    unsigned long long* counters;
};

__global__ __launch_bounds__(WAVE) void k_core_f32(CoreArgs A) {
    const int lane = threadIdx.x;
    const int L = A.L;
    if (A.counters[3] != 0) return;
    const int cw = (int)A.counters[4];  // max distance from jc to a core point, in points
    for (;;) {
        const int item = wave_take(&A.counters[5], lane);
        if (item >= A.n_items) break;
        const int tile = item / A.n_lblocks;
        const int l = (item % A.n_lblocks) * WAVE + lane;
        const bool lv = l < L;
        const int64_t j0 = (int64_t)tile * CT;
        const int64_t j1 = min(j0 + CT, A.n_wl) - 1;
        float acc[CT];
#pragma unroll
        for (int t = 0; t < CT; ++t) acc[t] = 0.0f;
        // candidate lines: jc in [j0 - cw - 1, j1 + cw + 1] (jc is monotone in the line index)
        const int64_t i0 = wave_lower_bound_i32(A.jc_all, A.n_lines, (int)(j0 - cw - 1), lane);
        const int64_t i1 = wave_upper_bound_i32(A.jc_all, A.n_lines, (int)(j1 + cw + 1), lane);
        for (int64_t i = i0; i < i1; ++i) {
            const int line = __builtin_amdgcn_readfirstlane((int)i);
            const LineMeta m = *uniform_ptr(A.meta + line);
            if (m.clo > m.chi || m.chi < j0 || m.clo > j1) continue;
            const double4 q = lv ? A.rec64[(int64_t)line * A.Ls + l] : make_double4(0, 0, 0, 0);
            const float a32 = (float)q.y, sc = (float)q.z;
#pragma unroll
            for (int t = 0; t < CT; ++t) {
                const int64_t j = j0 + t;
                if (j < m.clo || j > m.chi) continue;  // wavefront-uniform
                const double v64 = fabs(*uniform_ptr(A.wls + j) - m.lam0) * q.x;
                if (lv && is_core_branch(q.y, v64)) {
                    const int piece = v64 < 1.3 ? 0 : (v64 < 2.4 ? 1 : 2);
                    acc[t] += voigt_core_f32(a32, (float)v64, piece, q.y > 0.2) * sc;
                }
            }
        }
        if (lv) {
#pragma unroll
            for (int t = 0; t < CT; ++t) {
                const int64_t j = j0 + t;
                if (j <= j1) {
                    const int64_t k = (int64_t)l + (int64_t)L * j;
                    if (A.alpha_f32) ((float*)A.alpha)[k] += acc[t];
                    else ((double*)A.alpha)[k] += (double)acc[t];
                }
            }
        }
    }
}

}  // namespace

// ============================================================================================
// host side
// ============================================================================================
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct korg_ctx {
    int device = 0;
    int n_cu = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev[5] = {};
    std::string err;
    korg_stats stats{};
    korg_tuning tuning{};
    int64_t last_n_lines = 0;
    // device scratch (grow-only)
    DevBuf meta, jc_all, rec64, recA, recD, dmask, tab, counters, blk, cstart, cls_lines, cls_jc, ranges;
    // staging for the host-memory entry point
    DevBuf h_alpha, h_wls, h_lines[8], h_species_mass, h_species_mol, h_layers[7], h_knots,
        h_cntm;
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

struct Plan {
    int P;     // wavelength points per lane
    int LB;    // layers per work item
    int wpc;   // persistent wavefronts per compute unit
};

constexpr int F64_GROUP = 8;  // layers per work item of the Float64 form

Plan choose_plan(const korg_ctx* ctx, int L, int precision) {
    Plan p;
    if (precision == KORG_PRECISION_FP64) {
        p.P = ctx->tuning.points_per_lane == 2 ? 2 : 1;
        p.LB = F64_GROUP;
    } else {
        p.P = ctx->tuning.points_per_lane > 0 ? (ctx->tuning.points_per_lane >= 2 ? 2 : 1) : 2;
        p.LB = L <= 8 ? 8 : L <= 16 ? 16 : L <= 32 ? 32 : L <= 56 ? 56 : 64;
    }
    p.wpc = ctx->tuning.waves_per_cu > 0 ? ctx->tuning.waves_per_cu : 12;
    return p;
}

template <int LB>
void launch_profile_f32(int P, dim3 grid, hipStream_t s, const ProfileArgs& A) {
    if (P == 2) hipLaunchKernelGGL((k_profile_f32<LB, 2>), grid, dim3(WAVE), 0, s, A);
    else hipLaunchKernelGGL((k_profile_f32<LB, 1>), grid, dim3(WAVE), 0, s, A);
}

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
    if (precision != KORG_PRECISION_FP64 && precision != KORG_PRECISION_FP32_FIXUP)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "bad precision %d", precision);
    if (cn->mode != KORG_CNTM_AT_LINES && cn->mode != KORG_CNTM_GRID)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "bad continuum mode %d", cn->mode);
    if (cn->mode == KORG_CNTM_GRID && (cn->n_knots < 2 || !cn->knots))
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "continuum grid needs >= 2 knots");
    if ((int64_t)ly->n_layers * ln->n_lines > ((int64_t)1 << 40))
        return set_err(ctx, KORG_ERR_UNSUPPORTED, "problem too large");
    return KORG_OK;
}

// The device pipeline; every pointer is a device pointer.
int run_device(korg_ctx* ctx, void* alpha, int alpha_dtype, const korg_grid* g,
               const korg_lines* ln, const korg_species* sp, const korg_layers* ly,
               const korg_continuum* cn, double cutoff, int precision) {
    const int64_t N = ln->n_lines;
    const int L = ly->n_layers;
    const int64_t nwl = g->n_wl;
    ctx->last_n_lines = N;
    std::memset(&ctx->stats, 0, sizeof ctx->stats);
    if (N == 0) return KORG_OK;  // src/line_absorption.jl:35-37
    hipStream_t s = ctx->stream;
    const Plan plan = choose_plan(ctx, L, precision);
    const int tile_w = WAVE * plan.P;
    const int ntiles = (int)((nwl + tile_w - 1) / tile_w);
    const int n_lblocks = (L + plan.LB - 1) / plan.LB;
    const int Ls = n_lblocks * plan.LB;  // records padded to whole layer blocks
    const int nblocks = (int)((N + PART_BLOCK - 1) / PART_BLOCK);
    const bool f32 = precision == KORG_PRECISION_FP32_FIXUP;

    TRY(grow(ctx, ctx->meta, sizeof(LineMeta) * N));
    TRY(grow(ctx, ctx->jc_all, sizeof(int) * N));
    TRY(grow(ctx, ctx->rec64, sizeof(double4) * N * Ls));
    if (f32) {
        TRY(grow(ctx, ctx->recA, sizeof(float4) * N * Ls));
        TRY(grow(ctx, ctx->recD, sizeof(float2) * N * Ls));
        TRY(grow(ctx, ctx->dmask, sizeof(uint64_t) * N * n_lblocks));
    }
    TRY(grow(ctx, ctx->tab, sizeof(LayerTab) * L));
    TRY(grow(ctx, ctx->counters, sizeof(unsigned long long) * N_COUNTERS));
    TRY(grow(ctx, ctx->blk, sizeof(int) * (size_t)nblocks * N_CLASSES));
    TRY(grow(ctx, ctx->cstart, sizeof(int) * (N_CLASSES + 1)));
    TRY(grow(ctx, ctx->cls_lines, sizeof(int) * N));
    TRY(grow(ctx, ctx->cls_jc, sizeof(int) * N));
    TRY(grow(ctx, ctx->ranges, sizeof(int2) * (size_t)ntiles * N_CLASSES));

    auto* counters = (unsigned long long*)ctx->counters.p;
    HIP_TRY(ctx, hipEventRecord(ctx->ev[0], s));
    hipLaunchKernelGGL(k_layers, dim3((L + 63) / 64), dim3(64), 0, s, L, ly->temps, ly->n_HI,
                       ly->n_HeI, ly->n_H2, (LayerTab*)ctx->tab.p, counters);
    SetupArgs sa;
    sa.n_lines = N;
    sa.n_wl = nwl;
    sa.L = L;
    sa.Ls = Ls;
    sa.wls = g->all_wls;
    sa.wl = ln->wl;
    sa.log_gf = ln->log_gf;
    sa.E_lower = ln->E_lower;
    sa.gamma_rad = ln->gamma_rad;
    sa.gamma_stark = ln->gamma_stark;
    sa.vdW1 = ln->vdW1;
    sa.vdW2 = ln->vdW2;
    sa.species = ln->species;
    sa.n_species = sp->n_species;
    sa.mass = sp->mass;
    sa.is_mol = sp->is_molecule;
    sa.n_div_U = ly->n_div_U;
    sa.temps = ly->temps;
    sa.n_e = ly->n_e;
    sa.xi = ly->xi;
    sa.tab = (const LayerTab*)ctx->tab.p;
    sa.cntm_mode = cn->mode;
    sa.n_knots = cn->n_knots;
    sa.knots = cn->knots;
    sa.cntm = cn->values;
    sa.cutoff = cutoff;
    sa.want_rec32 = f32 ? 1 : 0;
    sa.n_lblocks = n_lblocks;
    sa.count_branches = ctx->tuning.count_branches ? 1 : 0;
    sa.meta = (LineMeta*)ctx->meta.p;
    sa.jc_all = (int*)ctx->jc_all.p;
    sa.rec64 = (double4*)ctx->rec64.p;
    sa.recA = (float4*)ctx->recA.p;
    sa.recD = (float2*)ctx->recD.p;
    sa.dmask = (uint64_t*)ctx->dmask.p;
    sa.counters = counters;
    hipLaunchKernelGGL(k_setup, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, s, sa);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipEventRecord(ctx->ev[1], s));

    hipLaunchKernelGGL(k_count, dim3(nblocks), dim3(PART_BLOCK), 0, s, N,
                       (const LineMeta*)ctx->meta.p, (int*)ctx->blk.p);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, nblocks, (int*)ctx->blk.p,
                       (int*)ctx->cstart.p);
    hipLaunchKernelGGL(k_scatter, dim3(nblocks), dim3(PART_BLOCK), 0, s, N,
                       (const LineMeta*)ctx->meta.p, (const int*)ctx->blk.p,
                       (const int*)ctx->cstart.p, (int*)ctx->cls_lines.p, (int*)ctx->cls_jc.p);
    hipLaunchKernelGGL(k_ranges, dim3((ntiles * N_CLASSES + 255) / 256), dim3(256), 0, s,
                       ntiles, tile_w, nwl, (const int*)ctx->cstart.p, (const int*)ctx->cls_jc.p,
                       (int2*)ctx->ranges.p);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipEventRecord(ctx->ev[2], s));

    ProfileArgs pa;
    pa.n_wl = nwl;
    pa.L = L;
    pa.Ls = Ls;
    pa.n_lblocks = n_lblocks;
    pa.ntiles = ntiles;
    pa.tile_w = tile_w;
    pa.n_items = (int64_t)ntiles * n_lblocks;
    pa.wls = g->all_wls;
    pa.meta = (const LineMeta*)ctx->meta.p;
    pa.rec64 = (const double4*)ctx->rec64.p;
    pa.recA = (const float4*)ctx->recA.p;
    pa.recD = (const float2*)ctx->recD.p;
    pa.dmask = (const uint64_t*)ctx->dmask.p;
    pa.cls_lines = (const int*)ctx->cls_lines.p;
    pa.ranges = (const int2*)ctx->ranges.p;
    pa.alpha = alpha;
    pa.alpha_f32 = alpha_dtype == KORG_DTYPE_F32;
    pa.counters = counters;
    int64_t nwaves = (int64_t)ctx->n_cu * plan.wpc;
    if (nwaves > pa.n_items) nwaves = pa.n_items;
    const dim3 grid((unsigned)nwaves);
    if (!f32) {
        if (plan.P == 2) hipLaunchKernelGGL((k_profile_f64<F64_GROUP, 2>), grid, dim3(WAVE), 0, s, pa);
        else hipLaunchKernelGGL((k_profile_f64<F64_GROUP, 1>), grid, dim3(WAVE), 0, s, pa);
        HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
    } else {
        switch (plan.LB) {
            case 8: launch_profile_f32<8>(plan.P, grid, s, pa); break;
            case 16: launch_profile_f32<16>(plan.P, grid, s, pa); break;
            case 32: launch_profile_f32<32>(plan.P, grid, s, pa); break;
            case 56: launch_profile_f32<56>(plan.P, grid, s, pa); break;
            default: launch_profile_f32<64>(plan.P, grid, s, pa); break;
        }
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, hipEventRecord(ctx->ev[4], s));
        CoreArgs ca;
        ca.n_wl = nwl;
        ca.n_lines = N;
        ca.L = L;
        ca.Ls = Ls;
        ca.n_lblocks = (L + WAVE - 1) / WAVE;
        ca.ntiles = (int)((nwl + CT - 1) / CT);
        ca.n_items = (int64_t)ca.ntiles * ca.n_lblocks;
        ca.wls = g->all_wls;
        ca.meta = (const LineMeta*)ctx->meta.p;
        ca.jc_all = (const int*)ctx->jc_all.p;
        ca.rec64 = (const double4*)ctx->rec64.p;
        ca.alpha = alpha;
        ca.alpha_f32 = alpha_dtype == KORG_DTYPE_F32;
        ca.counters = counters;
        int64_t cwaves = (int64_t)ctx->n_cu * 16;
        if (cwaves > ca.n_items) cwaves = ca.n_items;
        hipLaunchKernelGGL(k_core_f32, dim3((unsigned)cwaves), dim3(WAVE), 0, s, ca);
    }
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipEventRecord(ctx->ev[3], s));
    return KORG_OK;
}

int finish_stats(korg_ctx* ctx) {
    if (ctx->last_n_lines == 0) return KORG_OK;
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev[3]));
    float t01 = 0, t12 = 0, t24 = 0, t43 = 0, t03 = 0;
    HIP_TRY(ctx, hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]));
    HIP_TRY(ctx, hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]));
    HIP_TRY(ctx, hipEventElapsedTime(&t24, ctx->ev[2], ctx->ev[4]));
    HIP_TRY(ctx, hipEventElapsedTime(&t43, ctx->ev[4], ctx->ev[3]));
    HIP_TRY(ctx, hipEventElapsedTime(&t03, ctx->ev[0], ctx->ev[3]));
    unsigned long long c[10];
    static_assert(sizeof c <= sizeof(unsigned long long) * N_COUNTERS, "counters");
    HIP_TRY(ctx, hipMemcpy(c, ctx->counters.p, sizeof c, hipMemcpyDeviceToHost));
    if (c[3] & 2)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT,
                       "a line's species index is outside [0, n_species); alpha was left unchanged");
    if (c[3] & 1)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT,
                       "the line list is not sorted by wavelength; alpha was left unchanged");
    ctx->stats.n_evals = (double)c[1];
    ctx->stats.n_lines_used = (int64_t)c[2];
    ctx->stats.setup_ms = t01;
    ctx->stats.bin_ms = t12;
    ctx->stats.profile_ms = t24;
    ctx->stats.core_ms = t43;
    ctx->stats.total_ms = t03;
    for (int k = 0; k < 4; ++k) ctx->stats.n_evals_branch[k] = (double)c[6 + k];
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
    *out = ctx;
    return KORG_OK;
}

void korg_destroy(korg_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->own_stream) (void)hipStreamSynchronize(ctx->own_stream);
    DevBuf* bufs[] = {&ctx->meta, &ctx->jc_all, &ctx->rec64, &ctx->recA, &ctx->recD, &ctx->dmask, &ctx->tab, &ctx->counters, &ctx->blk,
                      &ctx->cstart, &ctx->cls_lines, &ctx->cls_jc, &ctx->ranges, &ctx->h_alpha,
                      &ctx->h_wls, &ctx->h_species_mass, &ctx->h_species_mol, &ctx->h_knots,
                      &ctx->h_cntm};
    for (DevBuf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (auto& b : ctx->h_lines)
        if (b.p) (void)hipFree(b.p);
    for (auto& b : ctx->h_layers)
        if (b.p) (void)hipFree(b.p);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
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
    if (t->points_per_lane < 0 || t->points_per_lane > 2 || t->waves_per_cu < 0 ||
        t->waves_per_cu > 64 || t->waves_per_tile != 0 || t->count_branches < 0 ||
        t->count_branches > 1)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "tuning out of range");
    ctx->tuning = *t;
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
    const int64_t N = lines->n_lines, L = layers->n_layers, S = species->n_species;
    const size_t esz = alpha_dtype == KORG_DTYPE_F32 ? 4 : 8;
    const size_t abytes = esz * (size_t)L * (size_t)grid->n_wl;
    korg_grid g = *grid;
    korg_lines ln = *lines;
    korg_species sp = *species;
    korg_layers ly = *layers;
    korg_continuum cn = *cntm;
    TRY(stage(ctx, ctx->h_wls, grid->all_wls, grid->n_wl, &g.all_wls));
    TRY(stage(ctx, ctx->h_lines[0], lines->wl, N, &ln.wl));
    TRY(stage(ctx, ctx->h_lines[1], lines->log_gf, N, &ln.log_gf));
    TRY(stage(ctx, ctx->h_lines[2], lines->E_lower, N, &ln.E_lower));
    TRY(stage(ctx, ctx->h_lines[3], lines->gamma_rad, N, &ln.gamma_rad));
    TRY(stage(ctx, ctx->h_lines[4], lines->gamma_stark, N, &ln.gamma_stark));
    TRY(stage(ctx, ctx->h_lines[5], lines->vdW1, N, &ln.vdW1));
    TRY(stage(ctx, ctx->h_lines[6], lines->vdW2, N, &ln.vdW2));
    TRY(stage(ctx, ctx->h_lines[7], lines->species, N, &ln.species));
    TRY(stage(ctx, ctx->h_species_mass, species->mass, S, &sp.mass));
    TRY(stage(ctx, ctx->h_species_mol, species->is_molecule, S, &sp.is_molecule));
    TRY(stage(ctx, ctx->h_layers[0], layers->temps, L, &ly.temps));
    TRY(stage(ctx, ctx->h_layers[1], layers->n_e, L, &ly.n_e));
    TRY(stage(ctx, ctx->h_layers[2], layers->n_div_U, S * L, &ly.n_div_U));
    TRY(stage(ctx, ctx->h_layers[3], layers->n_HI, L, &ly.n_HI));
    TRY(stage(ctx, ctx->h_layers[4], layers->n_HeI, L, &ly.n_HeI));
    TRY(stage(ctx, ctx->h_layers[5], layers->n_H2, L, &ly.n_H2));
    TRY(stage(ctx, ctx->h_layers[6], layers->xi, L, &ly.xi));
    if (cntm->mode == KORG_CNTM_GRID) {
        TRY(stage(ctx, ctx->h_knots, cntm->knots, cntm->n_knots, &cn.knots));
        TRY(stage(ctx, ctx->h_cntm, cntm->values, (size_t)L * cntm->n_knots, &cn.values));
    } else {
        TRY(stage(ctx, ctx->h_cntm, cntm->values, (size_t)L * N, &cn.values));
    }
    TRY(grow(ctx, ctx->h_alpha, abytes));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->h_alpha.p, alpha, abytes, hipMemcpyHostToDevice, ctx->stream));
    TRY(run_device(ctx, ctx->h_alpha.p, alpha_dtype, &g, &ln, &sp, &ly, &cn, cutoff_threshold,
                   precision));
    HIP_TRY(ctx, hipMemcpyAsync(alpha, ctx->h_alpha.p, abytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return finish_stats(ctx);
}

int korg_get_windows(korg_ctx* ctx, int64_t* lb, int64_t* ub, int64_t n_lines) {
    if (!ctx || !lb || !ub) return KORG_ERR_INVALID_ARGUMENT;
    if (n_lines != ctx->last_n_lines)
        return set_err(ctx, KORG_ERR_INVALID_ARGUMENT, "n_lines %lld != last call's %lld",
                       (long long)n_lines, (long long)ctx->last_n_lines);
    if (n_lines == 0) return KORG_OK;
    std::vector<LineMeta> m(n_lines);
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipMemcpy(m.data(), ctx->meta.p, sizeof(LineMeta) * n_lines, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n_lines; ++i) {
        lb[i] = (int64_t)m[i].lb + 1;
        ub[i] = (int64_t)m[i].ub + 1;
    }
    return KORG_OK;
}

}  // extern "C"
