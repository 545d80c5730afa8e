/*
 * korg_line.h -- C ABI of the MI355X (gfx950) line-opacity library.
 *
 * The library computes the line opacity of Korg.jl's `line_absorption!`
 * (src/line_absorption.jl:33-129): for every line, per-layer Doppler and Lorentz widths and
 * amplitudes, a window cut where the profile falls below cutoff_threshold * alpha_cntm
 * (src/line_absorption.jl:108-119), and the sum of Hunger-approximation Voigt profiles
 * (src/line_absorption.jl:245-310) added into alpha[layer, wavelength].
 *
 * Every entry point replaces the body of that one Julia function; the Julia side flattens its
 * arguments (`linelist`, `λs::Wavelengths`, `temps`, `nₑ`, `n_densities`, `partition_fns`, `ξ`,
 * `α_cntm`, `cutoff_threshold`) into the structs below.  Plain pointers and sizes only.
 *
 * Conventions
 *   - Units: cm, s^-1, eV, K, cm^-3 (as in the reference).
 *   - alpha is column-major [n_layers x n_wl] (Julia's Matrix): alpha[l + n_layers * j].
 *     Lines are ADDED into alpha (src/line_absorption.jl:127); an empty line list leaves
 *     alpha untouched (src/line_absorption.jl:35-37).
 *   - Return value: KORG_OK (0) or a KORG_ERR_* code; korg_last_error() gives the message.
 *     No C++ exception crosses the ABI.
 *   - A context owns one HIP stream and its device scratch; it is not thread-safe (use one
 *     context per host thread).  Results are deterministic: no atomics touch alpha and the
 *     order in which lines are summed into an element is fixed.
 */
#ifndef KORG_LINE_H
#define KORG_LINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KORG_LINE_ABI_VERSION 17

enum korg_status {
    KORG_OK = 0,
    KORG_ERR_INVALID_ARGUMENT = 1,
    KORG_ERR_HIP = 2,
    KORG_ERR_OUT_OF_MEMORY = 3,
    KORG_ERR_NO_DEVICE = 4,
    KORG_ERR_UNSUPPORTED = 5,
    /* a plan's per-tile line lists did not fit their buffer (alpha was left unchanged) */
    KORG_ERR_CAPACITY = 6,
};

/* Arithmetic of the profile evaluation. */
enum korg_precision {
    /* Float64 throughout, evaluating line_profile/voigt_hjerting with the reference's own
     * expressions (src/line_absorption.jl:245-310): the device form of the Float64 oracle. */
    KORG_PRECISION_FP64 = 0,
    /* Float32 profile evaluation and accumulation with Float64 fix-ups: per-(line, layer)
     * setup, window, branch selection and the wavelength offset lambda - lambda0 stay Float64. */
    KORG_PRECISION_FP32_FIXUP = 1,
    /* No counterpart in the reference: the Float32 form's algorithm in Float64, for a
     * like-for-like comparison of the two precisions.  The same windows, near zones and
     * per-tile line lists; the far zone as the four-term series on the Float64 matrix cores
     * (v_mfma_f64_16x16x4_f64), the near zone with the reference's own Float64 expressions.  Its
     * difference from KORG_PRECISION_FP64 is the series' truncation.  Forward entry points only
     * (the adjoint and the fit step take FP64 or FP32_FIXUP). */
    KORG_PRECISION_FP64_SERIES = 2,
};

enum korg_dtype { KORG_DTYPE_F64 = 0, KORG_DTYPE_F32 = 1 };

/* How alpha_cntm(lambda0) of each layer is supplied (src/line_absorption.jl:108). */
enum korg_cntm_mode {
    /* values[i * n_layers + l] = alpha_cntm of layer l at line i's wavelength */
    KORG_CNTM_AT_LINES = 0,
    /* values[l * n_knots + k] at knots[k] (cm, ascending); linear interpolation, the form of
     * the interpolators synthesize builds (src/synthesize.jl:226-229) */
    KORG_CNTM_GRID = 1,
};

/* The line list, struct of arrays (fields of Korg.Line, src/linelist.jl:5-14). */
typedef struct {
    int64_t n_lines;
    const double* wl;          /* cm */
    const double* log_gf;
    const double* E_lower;     /* eV */
    const double* gamma_rad;   /* s^-1 */
    const double* gamma_stark; /* s^-1 per electron at 10^4 K */
    const double* vdW1;        /* gamma_vdW at 10^4 K if vdW2 == -1, else ABO sigma */
    const double* vdW2;        /* -1, or ABO alpha */
    const int32_t* species;    /* index into korg_species */
    /* Optional frozen windows, 1-based as korg_get_windows returns them (lb > ub skips the line):
     * when both are given they replace the windows the call would compute
     * (src/line_absorption.jl:108-119), so that alpha is a smooth function of the per-layer
     * inputs -- as under ForwardDiff, whose duals carry no derivative through the integer
     * window -- for finite-difference checks and fits that hold the windows of an iterate.
     * Both NULL: computed, as the reference does. */
    const int64_t* window_lb;
    const int64_t* window_ub;
} korg_lines;

/* Species of the lines: mass (get_mass, src/species.jl:175) and ismolecule (:168). */
typedef struct {
    int32_t n_species;
    const double* mass;          /* g */
    const uint8_t* is_molecule;  /* 1: Stark and vdW broadening are skipped */
} korg_species;

/* Atmospheric layers (the per-layer vectors of line_absorption!). */
typedef struct {
    int32_t n_layers;
    const double* temps;    /* K */
    const double* n_e;      /* cm^-3 */
    /* [n_species * n_layers]: n_densities[s][l] / partition_fns[s](log(temps[l]))
     * (src/line_absorption.jl:42-44) */
    const double* n_div_U;
    const double* n_HI;     /* perturbers for n_eff_vdW (src/line_absorption.jl:54-60) */
    const double* n_HeI;
    const double* n_H2;
    const double* xi;       /* [n_layers] microturbulence, cm/s */
} korg_layers;

/* The wavelength grid: every point of a Korg.Wavelengths (its all_wls), cm, ascending. */
typedef struct {
    int64_t n_wl;
    const double* all_wls;
} korg_grid;

typedef struct {
    int32_t mode;           /* korg_cntm_mode */
    int64_t n_knots;        /* KORG_CNTM_GRID only */
    const double* knots;    /* KORG_CNTM_GRID only */
    const double* values;
} korg_continuum;

/* Statistics of the last call on a context. */
typedef struct {
    double n_evals;          /* profile evaluations: sum over lines of n_layers*(ub-lb+1) */
    int64_t n_lines_used;    /* lines with a non-empty window */
    double setup_ms;         /* device time of the per-line setup kernels */
    double bin_ms;           /* device time of the binning kernels */
    double profile_ms;       /* device time of the profile-accumulation kernel */
    double core_ms;          /* device time of the Voigt-core kernel (Float32 form only) */
    double total_ms;         /* device time of the whole call (excludes host<->device copies) */
    /* evaluations by Hunger branch (src/line_absorption.jl:288-309), filled when
     * korg_tuning.count_branches is set: far wing a <= 0.2 (A), core a <= 0.2 (B),
     * core 0.2 < a <= 1.4 (C), far wing / strongly damped (D) */
    double n_evals_branch[4];
    /* Float32 form only (0 for Float64): evaluations done exactly by the near-zone kernel
     * (the rest of n_evals are far-wing series terms on the matrix cores), and the
     * v_mfma_f32_32x32x16_bf16 instructions the far-wing kernel issued (32768 flops each; three
     * per group of four lines, 32-point block and 32-layer half, counted per line) */
    double n_evals_near;
    double n_wing_mfma;
    /* Float32 form only: far-zone evaluations inside the lines' tail radii, where the R^3, R^4
     * series terms exceed 1e-5 of the leading term (the bfloat16 far wing includes them
     * everywhere) */
    double n_evals_tail;
    /* Float32 form with korg_tuning.count_branches: near-zone evaluations in Hunger branch A
     * (the near zone holds every branch-B and -C evaluation; its remaining evaluations are D) */
    double n_evals_near_A;
    /* Float32 form with the far field: device time of the far-field kernel (included in
     * profile_ms), and the (line, 512-point super-tile) entries of its line lists */
    double far_ms;
    double n_far_entries;
} korg_stats;

typedef struct korg_ctx korg_ctx;

/* Create a context on HIP device `device` (its own stream, scratch grown on demand). */
int korg_create(korg_ctx** out, int device);
void korg_destroy(korg_ctx* ctx);
/* Message of the last error on ctx (never NULL; "" when none). ctx may be NULL. */
const char* korg_last_error(const korg_ctx* ctx);
int korg_abi_version(void);

/*
 * Drop-in body of line_absorption! (src/line_absorption.jl:33-129) on HOST memory.
 * alpha: host pointer, column-major [n_layers x n_wl], dtype korg_dtype; updated in place.
 * All other pointers are host pointers.  Synchronous: returns after alpha is updated.
 */
int korg_line_absorption(korg_ctx* ctx, void* alpha, int alpha_dtype, const korg_grid* grid,
                         const korg_lines* lines, const korg_species* species,
                         const korg_layers* layers, const korg_continuum* cntm,
                         double cutoff_threshold, int precision);

/*
 * The same on DEVICE memory: every pointer in the structs and alpha are device pointers
 * (HBM-resident inputs).  Runs on the context's stream; returns without synchronising when
 * `synchronize` is 0 (use korg_synchronize).
 */
int korg_line_absorption_device(korg_ctx* ctx, void* alpha, int alpha_dtype,
                                const korg_grid* grid, const korg_lines* lines,
                                const korg_species* species, const korg_layers* layers,
                                const korg_continuum* cntm, double cutoff_threshold,
                                int precision, int synchronize);

/*
 * korg_line_absorption_device with no device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory during the call (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels and results
 * (bitwise) as korg_line_absorption_device.
 *
 * Its scratch is a fixed device array of KORG_NOALLOC_SCRATCH_BYTES per GPU, held from the
 * library's load on that GPU until the process ends; korg_create resolves its address.  The call
 * carves the pipeline's buffers from it, makes no host round trip before returning (synchronize
 * = 0 leaves everything enqueued) and never calls hipMalloc/hipFree.  Per line it needs about
 * 32 B x (n_layers rounded up to 8) + 1 KB per 64 layers (Float32) + 0.2 KB, plus about 4 B per grid
 * point and 4 B per (line, 64-point tile) pair: for 56 layers about 20,000 lines on 100,000 points
 * in Float32.  A problem that does not fit returns KORG_ERR_CAPACITY with alpha unchanged (checked
 * on the host before any launch, or for the per-tile lists on the device, in which case no kernel
 * writes alpha).
 *
 * Every call on one GPU, from any context, shares that GPU's array, which keeps the last call's
 * contents: calls that overlap in time on one GPU (other streams, other contexts, other host
 * threads) race and give wrong results.  Calls on one context's stream are ordered and safe.
 */
#define KORG_NOALLOC_SCRATCH_BYTES (64ll << 20)
int korg_line_absorption_device_noalloc(korg_ctx* ctx, void* alpha, int alpha_dtype,
                                        const korg_grid* grid, const korg_lines* lines,
                                        const korg_species* species, const korg_layers* layers,
                                        const korg_continuum* cntm, double cutoff_threshold,
                                        int precision, int synchronize);

/*
 * One star -- one line_absorption! call -- of a batch (device pointers throughout).  Its line
 * list is sorted by wavelength and indexes the batch's shared species table.  Stars may have
 * different layer counts (Korg's interpolate_marcs drops layers, src/atmosphere.jl:461-479): a
 * batch of several counts runs as one pipeline per count, in order of first appearance, each
 * synchronized before the next (so such a call returns synchronized whatever `synchronize` asks);
 * plans and the autotuner take one count.
 */
typedef struct {
    void* alpha;              /* column-major [n_layers x grid.n_wl], the batch's alpha dtype */
    korg_grid grid;
    korg_lines lines;
    korg_layers layers;
    korg_continuum cntm;
    double cutoff_threshold;
} korg_star;

/*
 * Many line_absorption! calls in one pipeline (DEVICE memory): all stars' lines are set up in one
 * launch, binned onto the stars' wavelength tiles, and every profile kernel is a single persistent
 * launch over a (star, wavelength tile) work queue.  Each star's alpha receives exactly what
 * korg_line_absorption_device would add for it alone under the same tuning (bitwise).  `stars`
 * is a host array of n_stars structs.
 */
int korg_line_absorption_batched(korg_ctx* ctx, int64_t n_stars, const korg_star* stars,
                                 const korg_species* species, int alpha_dtype, int precision,
                                 int synchronize);

/* Use an existing HIP stream (hipStream_t passed as void*) instead of the context's own. */
int korg_set_stream(korg_ctx* ctx, void* stream);
int korg_synchronize(korg_ctx* ctx);
int korg_get_stats(const korg_ctx* ctx, korg_stats* out);

/*
 * Per-line windows of the last call, 1-based as the reference's lb/ub
 * (src/line_absorption.jl:114-115); lb > ub marks a skipped line.  Host arrays of n_lines.
 */
int korg_get_windows(korg_ctx* ctx, int64_t* lb, int64_t* ub, int64_t n_lines);

/* Kernel tuning (0 = library default).  No counterpart in the reference, whose only knob is
 * tasks_per_thread (src/line_absorption.jl:34); korg_autotune_batched chooses these. */
typedef struct {
    int32_t points_per_lane;   /* Float64 form: wavelength points per lane, 1 or 2 (default 1) */
    int32_t waves_per_tile;    /* Float32 far wing: waves sharing a tile's line list, 1, 2 or 4
                                  (default 4; at most 2 with 128-point tiles) */
    int32_t waves_per_cu;      /* persistent profile waves launched per compute unit (default 12) */
    int32_t count_branches;    /* 1: fill korg_stats.n_evals_branch (extra setup cost) */
    int32_t tile_points;       /* Float32 far wing: wavelength points per tile, 32, 64 or 128
                                  (default 64) */
    int32_t near_waves_per_cu; /* Float32 near zone: persistent waves per compute unit (default 16) */
    int32_t near_layout;       /* Float32 near zone: 1 = lanes over 64 layers, one point at a time;
                                  2 = lanes over 8 layers x 8 points (default); same alpha */
    int32_t near_concurrent;   /* Float32 near zone: 1 = after the far wing on the call's stream;
                                  2 (default) = beside it on a second stream of the context, into
                                  a buffer of n_layers x points floats added afterwards; same
                                  alpha (the allocation-free entry point always runs 1) */
    int32_t far_field;         /* Float32 far wing: 1 = every (line, tile) pair term by term;
                                  2 (default) = the far field, where a 64-point sub-tile lies
                                  inside a line's window, outside its near zone and more than 8
                                  half-widths from its centre, as a 12-term expansion about the
                                  sub-tile's centre (truncation below 2e-10 of the leading term);
                                  grids that are not evenly spaced take no far field */
} korg_tuning;
int korg_set_tuning(korg_ctx* ctx, const korg_tuning* tuning);
int korg_get_tuning(const korg_ctx* ctx, korg_tuning* out);

/*
 * Gradients of the adjoint: dL/dx for each per-layer input x of korg_layers, given
 * G = dL/dalpha.  A NULL member is not written.
 */
typedef struct {
    double* T;        /* [n_layers] dL/dT */
    double* n_e;      /* [n_layers] dL/dn_e */
    double* xi;       /* [n_layers] dL/dxi (xi in cm/s) */
    double* n_HI;     /* [n_layers] dL/dn(H I) (through the vdW perturber density) */
    double* n_HeI;    /* [n_layers] dL/dn(He I) */
    double* n_H2;     /* [n_layers] dL/dn(H2) */
    double* n_div_U;  /* [n_species * n_layers] dL/d(n_s / U_s), laid out as korg_layers.n_div_U */
} korg_layer_grads;

/*
 * Vector-Jacobian product of line_absorption! (src/line_absorption.jl:33-129) with respect to
 * its per-layer inputs: the reverse-mode derivative a ChainRulesCore rrule of the line opacity
 * needs (the reference differentiates this call with ForwardDiff duals, src/synthesize.jl:305-306
 * under Fit).  grad_alpha: G = dL/dalpha, column-major [n_layers x n_wl], dtype korg_dtype.
 * Windows are held fixed (they are integer functions of the inputs); every other dependence,
 * including the branch of voigt_hjerting in use, is differentiated exactly.  Computed in
 * Float64; deterministic.  Host pointers; synchronous.
 */
int korg_line_absorption_vjp(korg_ctx* ctx, const void* grad_alpha, int grad_dtype,
                             const korg_grid* grid, const korg_lines* lines,
                             const korg_species* species, const korg_layers* layers,
                             const korg_continuum* cntm, double cutoff_threshold,
                             korg_layer_grads* grads);

/* The same on DEVICE memory (grad_alpha, the structs' arrays and the grads members). */
int korg_line_absorption_vjp_device(korg_ctx* ctx, const void* grad_alpha, int grad_dtype,
                                    const korg_grid* grid, const korg_lines* lines,
                                    const korg_species* species, const korg_layers* layers,
                                    const korg_continuum* cntm, double cutoff_threshold,
                                    korg_layer_grads* grads, int synchronize);

/*
 * Tangents of the per-layer inputs: one direction of a forward-mode derivative.  Members as
 * korg_layer_grads (n_div_U [n_species * n_layers]); a NULL member is a zero tangent.
 */
typedef struct {
    const double* T;
    const double* n_e;
    const double* xi;
    const double* n_HI;
    const double* n_HeI;
    const double* n_H2;
    const double* n_div_U;
} korg_layer_tangents;

/*
 * Jacobian-vector products of line_absorption! (src/line_absorption.jl:33-129): for each of
 * n_tangents directions x_k of the per-layer inputs, ADDS J x_k -- the derivative of alpha along
 * x_k -- into dalpha[k] (column-major [n_layers x n_wl], dtype korg_dtype).  This is what a
 * ForwardDiff.Dual alpha receives in its partials when Korg's Fit differentiates synthesize
 * (src/Fit/fit_via_synthesis.jl:345-355 with autodiff=:forwarddiff; src/synthesize.jl:242-247,
 * 305-306): one direction per partial.  As ForwardDiff does, the windows are held at the values'
 * windows (their dependence on the inputs is integer) and the branch of voigt_hjerting in use is
 * differentiated exactly.  Float64 throughout; deterministic.  A zero tangent component adds
 * nothing.  The host form takes host pointers (dalpha: a host array of n_tangents host
 * pointers) and is synchronous; the _device form takes device pointers (dalpha: a host array of
 * device pointers; tangents: a host array of structs of device pointers).
 */
int korg_line_absorption_jvp(korg_ctx* ctx, void* const* dalpha, int dalpha_dtype,
                             const korg_grid* grid, const korg_lines* lines,
                             const korg_species* species, const korg_layers* layers,
                             const korg_continuum* cntm, double cutoff_threshold,
                             int64_t n_tangents, const korg_layer_tangents* tangents);
int korg_line_absorption_jvp_device(korg_ctx* ctx, void* const* dalpha, int dalpha_dtype,
                                    const korg_grid* grid, const korg_lines* lines,
                                    const korg_species* species, const korg_layers* layers,
                                    const korg_continuum* cntm, double cutoff_threshold,
                                    int64_t n_tangents, const korg_layer_tangents* tangents,
                                    int synchronize);

/*
 * Plans: a batched forward call (korg_line_absorption_batched) captured once in a HIP graph for
 * fixed device pointers and replayed with no host round trip -- the form for a fit that
 * re-evaluates the same stars with new inputs in the same buffers (Korg's Fit re-runs synthesize
 * per iteration, src/Fit/fit_via_synthesis.jl:304).  No counterpart in the reference.
 *
 * korg_plan_batched validates the stars, shapes the pipeline under the context's current
 * tuning, measures the per-tile line lists once on the current inputs, allocates them `headroom`
 * times larger (0: 1.25; else in [1, 16]) and captures the graph.  The plan owns its device
 * scratch; it keeps `ctx` (destroy plans before their context).  Every replay ADDS into the
 * alphas, as the direct call does, and reads the arrays' contents at replay time.
 *
 * korg_plan_launch replays on the context's stream (synchronize = 0: returns at once; call
 * korg_plan_wait before reading the results).  A replay whose inputs need more tile-list entries
 * than allocated writes nothing; korg_plan_wait then grows the lists, captures the graph again
 * and replays it, so a finished launch has always added the full result.  korg_get_stats gives
 * the counts of the last finished replay (total_ms: the replay's stream time).
 */
typedef struct korg_plan korg_plan;
typedef struct {
    int64_t n_stars;
    int64_t n_lines;        /* lines of all stars */
    int64_t n_points;       /* grid points of all stars, each star padded to a multiple of 512 */
    int64_t tile_capacity;  /* entries allocated for the per-tile line lists (Float32 form) */
    int64_t tile_entries;   /* entries the last finished replay needed */
    int32_t n_captures;     /* graph captures so far (1 + regrowths) */
    int32_t pad;
    korg_tuning tuning;     /* the tuning the plan was captured with */
} korg_plan_info;
int korg_plan_batched(korg_ctx* ctx, int64_t n_stars, const korg_star* stars,
                      const korg_species* species, int alpha_dtype, int precision,
                      double headroom, korg_plan** out);
int korg_plan_launch(korg_plan* plan, int synchronize);
int korg_plan_wait(korg_plan* plan);
int korg_plan_get_info(const korg_plan* plan, korg_plan_info* out);
void korg_plan_destroy(korg_plan* plan);

/*
 * Autotuner: times the batched pipeline on these stars (into scratch alphas; the caller's are not
 * touched) over the tile shape (Float32: 32/64/128-point tiles x 1/2/4 waves per tile; Float64:
 * 1/2 points per lane), then persistent waves per CU, then near-zone waves per CU, and makes the
 * fastest the context's tuning.  cache_path (may be NULL): a text file of results keyed by device,
 * precision, alpha dtype, layer count and the power-of-two classes of the star, line and point
 * counts; a cached key is used without timing, a new result is appended.  No counterpart in the
 * reference.
 */
int korg_autotune_batched(korg_ctx* ctx, int64_t n_stars, const korg_star* stars,
                          const korg_species* species, int alpha_dtype, int precision,
                          const char* cache_path, korg_tuning* out);

/*
 * The fit step of a batch of stars: the line opacity (forward) into each star's alpha (zeroed
 * first), radiative transfer, chi^2 between the normalized flux and the data, and the gradient of
 * chi^2 with respect to each star's per-layer inputs through the line-opacity adjoint.  This is
 * what one Newton/LBFGS iteration of Korg's Fit needs of the line opacity
 * (src/Fit/fit_via_synthesis.jl:304; the reference differentiates `synthesize` with ForwardDiff,
 * src/synthesize.jl:305-306, 310).
 *
 * Radiative transfer is Korg's plane-parallel default (src/RadiativeTransfer/RadiativeTransfer.jl
 * :62-66, 77-146): anchored optical depth (:244-253) on tau_ref and alpha_ref, the blackbody of each
 * layer's temperature as source (src/synthesize.jl:285, 412-417), the flux-only linear-source
 * solution in exponential integrals (:379-400) with Korg's approximate E2 (:408-480), in Float64.
 * The normalized flux is F(alpha_cntm + alpha_lines) / F(alpha_cntm); alpha_cntm is each star's
 * continuum on knots (KORG_CNTM_GRID required).  chi^2 = sum_j ((F_norm_j - data_j) / sigma_j)^2.
 * The gradient through alpha_lines is the line-opacity adjoint of G = dchi^2/dalpha_lines from the
 * transfer's reverse mode, windows held fixed: at KORG_PRECISION_FP64 that of
 * korg_line_absorption_vjp (Float64 throughout); at KORG_PRECISION_FP32_FIXUP the far zone of each
 * line through Float32 moments of the far-wing series on the matrix cores, its near zone by exact
 * Float32 evaluation (gradient within a few 1e-5 of the Float64 step).  Device pointers.
 */
typedef struct {
    const double* tau_ref;    /* [n_layers] reference optical depth of the anchored scheme */
    const double* alpha_ref;  /* [n_layers] absorption at the reference wavelength */
    const double* data;       /* [n_wl] observed normalized flux */
    const double* sigma;      /* [n_wl] its uncertainties */
    double* flux;             /* out [n_wl]: the model's normalized flux (may be NULL) */
    double* chi2;             /* out [1] */
    korg_layer_grads grads;   /* out: dchi^2/d(per-layer inputs), device arrays (NULL: skipped) */
} korg_fit_target;
int korg_fit_gradient_batched(korg_ctx* ctx, int64_t n_stars, const korg_star* stars,
                              const korg_fit_target* targets, const korg_species* species,
                              int alpha_dtype, int precision, int synchronize);
/* The fit step captured once as a plan (replayed by korg_plan_launch, as korg_plan_batched; a
 * replay recomputes everything from the inputs in the captured buffers). */
int korg_plan_fit(korg_ctx* ctx, int64_t n_stars, const korg_star* stars,
                  const korg_fit_target* targets, const korg_species* species, int alpha_dtype,
                  int precision, double headroom, korg_plan** out);

/*
 * The continuum of every star on its grid: alpha[l, j] = the linear interpolation of layer l's
 * continuum on its knots (KORG_CNTM_GRID required) at lambda_j, overwriting alpha -- the alpha
 * that synthesize fills before the lines are added (src/synthesize.jl:225-226, 248-250; outside
 * the knots the end intervals are extended).  The line opacity is then added by
 * korg_line_absorption_batched or _device, which accumulate.  Device pointers; the stars' line
 * lists are not read.  No counterpart as a separate call in the reference.
 */
int korg_continuum_batched(korg_ctx* ctx, int64_t n_stars, const korg_star* stars, int alpha_dtype,
                           int synchronize);

/*
 * The emergent flux of every star by Korg's plane-parallel default transfer (the scheme of
 * korg_fit_gradient_batched, in Float64): flux from each star's alpha taken as the total
 * absorption (the continuum included, as korg_continuum_batched and the lines leave it), cntm
 * from the continuum alone -- synthesize's flux and cntm in erg/s/cm^4/A
 * (src/synthesize.jl:286-290, 310-318).  Device pointers.
 */
typedef struct {
    const double* tau_ref;    /* [n_layers] reference optical depth of the anchored scheme */
    const double* alpha_ref;  /* [n_layers] absorption at the reference wavelength */
    double* flux;             /* out [n_wl] (may be NULL) */
    double* cntm;             /* out [n_wl] (may be NULL) */
} korg_flux_target;
int korg_flux_batched(korg_ctx* ctx, int64_t n_stars, const korg_star* stars,
                      const korg_flux_target* targets, int alpha_dtype, int synchronize);

/*
 * Hydrogen lines (src/hydrogen_line_absorption.jl:75-169): the Stehle & Hutcheon (1999) Stark
 * profiles of the Lyman, Balmer and Paschen series to n = 30, plus the Barklem et al. (2000)
 * resonance profiles of H alpha, beta and gamma, with the MHD occupation probabilities
 * (src/statmech.jl:463-497); and the Brackett series 4 -> m, m = 5 .. the tables' highest upper
 * level, whose Griem electron-impact and quasistatic (Holtsmark) Stark profiles are convolved
 * with each other and with the Doppler profile on 201 wavelengths per line and layer
 * (:180-470), over windows of 5 max(Doppler width, Stark width) that can reach far from the
 * lines (into the optical, for high m and n_e).
 *
 * korg_set_hydrogen_tables copies the tables to the device once per context, replacing Korg's
 * load-time _load_stark_profiles (src/hydrogen_line_absorption.jl:6-39).  Per transition, in
 * Korg's (the file's) order: its levels and log gf, and offsets into packed host arrays of its
 * table axes -- temperatures [n_T], electron densities [n_ne], log(delta_nu / F0) [n_dnu] (entry 0
 * is not read: Korg's first knot is -floatmax) -- its log profile [n_dnu][n_ne][n_T] (log 0 given
 * as -700, as Korg clips it) and its line centre [n_ne][n_T] in cm.
 */
typedef struct {
    int32_t lower, upper;      /* principal quantum numbers */
    int32_t n_T, n_ne, n_dnu;  /* axis lengths */
    int32_t reserved;
    double log_gf;
    int64_t off_T, off_ne, off_dnu, off_profile, off_lambda0;
} korg_htransition;
typedef struct {
    int32_t n_transitions;
    const korg_htransition* transitions;
    const double* temps;
    const double* n_e;
    const double* log_dnu;
    const double* log_profile;
    const double* lambda0;
    int64_t n_temps, n_n_e, n_log_dnu, n_log_profile, n_lambda0;  /* packed array lengths */
} korg_htables;
int korg_set_hydrogen_tables(korg_ctx* ctx, const korg_htables* tables);

/*
 * synthesize's hydrogen-line loop (src/synthesize.jl:293-301): for every layer l,
 * hydrogen_line_absorption!(alpha[l, :], wls, T_l, n_e_l, n_HI_l, n_HeI_l, U_HI_l, xi_l,
 * window_size; use_MHD), accumulating into alpha (column-major [n_layers x n_wl], f64 or f32).
 * Device pointers throughout; U_HI is the H I partition function at each layer's temperature.
 * window_size in cm (synthesize's default: 150e-8).  Float64 arithmetic.
 */
typedef struct {
    int32_t n_layers;
    const double* temps;
    const double* n_e;
    const double* n_HI;
    const double* n_HeI;
    const double* U_HI;
    const double* xi;  /* cm/s */
} korg_hlayers;
int korg_hydrogen_line_absorption(korg_ctx* ctx, void* alpha, int alpha_dtype,
                                  const korg_grid* grid, const korg_hlayers* layers,
                                  double window_size, int use_MHD, int synchronize);

/*
 * Precomputed molecular cross-sections (Korg's MolecularCrossSection,
 * src/molecular_cross_sections.jl:3-8): the absorption cross-section of one molecular species on a
 * (vmic, log10 T, wavelength) grid, built by the precompute (:48-83) from line_absorption! --
 * every call of which this library can run (korg_line_absorption_batched over the vmic knots, the
 * log T knots as layers).  sigma is laid out lambda-row by lambda-row within each vmic knot:
 * sigma[(iv * n_wl + j) * n_logT + iT] (Korg's itp.coefs[iv, iT, j] permuted).
 */
typedef struct {
    int32_t n_vmic, n_logT;   /* at least 2 knots each */
    int64_t n_wl;             /* at least 2 */
    const double* vmic;       /* [n_vmic] km/s, ascending */
    const double* logT;       /* [n_logT] log10(T / K), ascending */
    const double* wls;        /* [n_wl] cm, ascending: the table's wavelengths */
    const double* sigma;      /* cm^2 */
    const double* n_species;  /* [n_layers]: the species' number density in each layer */
} korg_mol_xsection;

/*
 * interpolate_molecular_cross_sections! (src/molecular_cross_sections.jl:99-114): for every table
 * and layer l, alpha[l, j] += sigma(vmic[l], log10 temps[l], lambda_j) * n_species[l], sigma
 * linear in each gridded dimension and zero outside the table (extrapolate(..., 0.0), :76-78).
 * vmic in km/s per layer.  alpha column-major [n_layers x n_wl], f64 or f32.  No tables: alpha
 * untouched.  The host form takes host pointers and is synchronous; the _device form takes device
 * pointers (tables: a host array of structs of device pointers).  Float64 arithmetic.
 */
int korg_molecular_cross_sections(korg_ctx* ctx, void* alpha, int alpha_dtype,
                                  const korg_grid* grid, int32_t n_layers, const double* temps,
                                  const double* vmic, int64_t n_tables,
                                  const korg_mol_xsection* tables);
int korg_molecular_cross_sections_device(korg_ctx* ctx, void* alpha, int alpha_dtype,
                                         const korg_grid* grid, int32_t n_layers,
                                         const double* temps, const double* vmic,
                                         int64_t n_tables, const korg_mol_xsection* tables,
                                         int synchronize);

/*
 * filter_linelist (src/synthesize.jl:331-350) on the device, for a batch of stars drawn from one
 * master list in HBM.  Star k keeps, for each of its wavelength ranges r in order, the master lines
 * searchsortedfirst(wl, first_r - line_buffer) .. searchsortedlast(wl, last_r + line_buffer), a
 * range starting no earlier than one past the previous range's searchsortedlast (so no line is
 * kept twice), in master order -- exactly the reference's selection.
 *
 * master: device arrays, sorted by wavelength (window_lb/ub must be NULL: frozen windows belong
 * to one star's list).  stars: host array of n_stars; their ranges are host arrays.  out: host
 * array of n_stars korg_lines, filled with device pointers and species indices of the master's
 * species table, windows NULL.  A star whose kept lines form one contiguous run of the master
 * (every single-range star) gets pointers into the master's arrays (no copy); the others get their
 * lines gathered into `pool`, a caller-owned device buffer of KORG_FILTER_POOL_BYTES(pool_lines)
 * bytes (7 Float64 columns wl, log_gf, E_lower, gamma_rad, gamma_stark, vdW1, vdW2, then the
 * int32 species column, each pool_lines long).  *pool_lines_needed receives the pool size the call
 * needs; when pool_lines is smaller the call returns KORG_ERR_CAPACITY and writes neither pool nor
 * out (pool may be NULL then: a size query).  The outputs stay valid while master and pool do.
 * Synchronous (the selection sizes come back to the host).
 */
typedef struct {
    int32_t n_ranges;
    const double* ranges;  /* host [2 * n_ranges]: first and last wavelength of each range, cm,
                            * ascending (eachwindow(wls), src/wavelengths.jl:159) */
    double line_buffer;    /* cm */
} korg_filter_star;
#define KORG_FILTER_POOL_BYTES(n) ((int64_t)(n) * 60)
int korg_filter_linelist_batched(korg_ctx* ctx, const korg_lines* master, int64_t n_stars,
                                 const korg_filter_star* stars, void* pool, int64_t pool_lines,
                                 korg_lines* out, int64_t* pool_lines_needed);

#ifdef __cplusplus
}
#endif
#endif /* KORG_LINE_H */
