# KorgLineGPU.jl -- Julia binding of libkorgline.so (include/korg_line.h).
#
# Loading this module adds more specific methods of Korg.line_absorption! for Float64 lines, so
# that every one of Korg's calls dispatches to the GPU:
#   - Float64 alpha, as a Matrix or any view of one: synthesize's main call
#     (src/synthesize.jl:305-306), its reference-wavelength call on view(α_ref, :, 1)
#     (src/synthesize.jl:266-276) and the molecular cross-section precompute on view(α, i, :, :)
#     (src/molecular_cross_sections.jl:75-76);
#   - interpolate_molecular_cross_sections! for Float64 alpha (src/synthesize.jl:277-282, 307-308);
#   - ForwardDiff.Dual alpha, as when Korg's Fit differentiates synthesize with
#     autodiff=:forwarddiff (src/Fit/fit_via_synthesis.jl:345-355, src/synthesize.jl:242-247):
#     the values through korg_line_absorption, the partials through korg_line_absorption_jvp
#     (one tangent direction per partial, windows held at the values' windows as ForwardDiff's
#     integer searchsorted holds them).
#
#     using Korg, KorgLineGPU
#     KorgLineGPU.set_precision!(:fp32)   # or :fp64 (default)
#     sol = synthesize(atm, linelist, A_X, 5000, 5100)
module KorgLineGPU

using Korg
using Interpolations
using ChainRulesCore
using ForwardDiff

const LIB = get(ENV, "KORG_LINE_LIB", joinpath(@__DIR__, "..", "libkorgline.so"))

struct KorgLines
    n_lines::Int64
    wl::Ptr{Float64}
    log_gf::Ptr{Float64}
    E_lower::Ptr{Float64}
    gamma_rad::Ptr{Float64}
    gamma_stark::Ptr{Float64}
    vdW1::Ptr{Float64}
    vdW2::Ptr{Float64}
    species::Ptr{Int32}
    window_lb::Ptr{Int64}   # frozen windows (unused here: C_NULL, the windows are computed)
    window_ub::Ptr{Int64}
end

struct KorgSpecies
    n_species::Int32
    mass::Ptr{Float64}
    is_molecule::Ptr{UInt8}
end

struct KorgLayers
    n_layers::Int32
    temps::Ptr{Float64}
    n_e::Ptr{Float64}
    n_div_U::Ptr{Float64}
    n_HI::Ptr{Float64}
    n_HeI::Ptr{Float64}
    n_H2::Ptr{Float64}
    xi::Ptr{Float64}
end

struct KorgGrid
    n_wl::Int64
    all_wls::Ptr{Float64}
end

struct KorgContinuum
    mode::Int32
    n_knots::Int64
    knots::Ptr{Float64}
    values::Ptr{Float64}
end

const PRECISION = Ref{Int32}(0)       # 0 = KORG_PRECISION_FP64, 1 = KORG_PRECISION_FP32_FIXUP
const CTX = Ref{Ptr{Cvoid}}(C_NULL)   # one context per process; guarded by LOCK
const LOCK = ReentrantLock()

set_precision!(p::Symbol) = (PRECISION[] = p === :fp32 ? Int32(1) : Int32(0); nothing)

function context()
    if CTX[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:korg_create, LIB), Cint, (Ref{Ptr{Cvoid}}, Cint), h, 0)
        rc == 0 || error("korg_create failed with code $rc (no HIP device?)")
        CTX[] = h[]
    end
    CTX[]
end

function check(rc)
    if rc != 0
        msg = unsafe_string(ccall((:korg_last_error, LIB), Cstring, (Ptr{Cvoid},), CTX[]))
        error("korg_line error $rc: $msg")
    end
end

# alpha_cntm: the per-layer interpolators synthesize builds (src/synthesize.jl:226-229) are
# passed as their knots and values; any other callable is evaluated at each line centre.
function continuum_arrays(α_cntm, linelist, L)
    itps = α_cntm
    if all(f -> f isa Interpolations.Extrapolation, itps)
        knots = collect(Float64, first(itps).itp.knots[1])
        if all(f -> f.itp.knots[1] == knots, itps)
            vals = Matrix{Float64}(undef, length(knots), L)  # column l = layer l
            for (l, f) in enumerate(itps)
                vals[:, l] .= f.itp.coefs
            end
            return Int32(1), knots, vals    # KORG_CNTM_GRID, values[l*n_knots + k]
        end
    end
    vals = Matrix{Float64}(undef, L, length(linelist))   # values[i*L + l]
    for (i, line) in enumerate(linelist), l in 1:L
        vals[l, i] = α_cntm[l](line.wl)
    end
    Int32(0), Float64[], vals
end

to_layers(x::Number, L) = fill(Float64(x), L)
to_layers(x, L) = Float64.(collect(x))

# The C-ABI view of one call: arrays (kept alive by the returned NamedTuple) and the structs that
# point into them.  `ll` is sorted by wavelength; `specs` lists its species (index = table row).
function flatten(ll, specs, λs, temps, nₑ, n_densities, partition_fns, ξ, α_cntm)
    L = length(temps)
    sidx = Dict(s => Int32(i - 1) for (i, s) in enumerate(specs))
    wl = Float64[l.wl for l in ll]
    log_gf = Float64[l.log_gf for l in ll]
    El = Float64[l.E_lower for l in ll]
    grad = Float64[l.gamma_rad for l in ll]
    gstark = Float64[l.gamma_stark for l in ll]
    vdW1 = Float64[l.vdW[1] for l in ll]
    vdW2 = Float64[l.vdW[2] for l in ll]
    sp = Int32[sidx[l.species] for l in ll]
    mass = Float64[Korg.get_mass(s) for s in specs]
    ismol = UInt8[Korg.ismolecule(s) for s in specs]
    # n/U per species and layer (src/line_absorption.jl:42-44), column s = species
    n_div_U = Matrix{Float64}(undef, L, length(specs))
    for (i, s) in enumerate(specs)
        n_div_U[:, i] .= n_densities[s] ./ partition_fns[s].(log.(temps))
    end
    # densities may be per-layer vectors or one number for every layer (the molecular
    # precompute passes scalars, src/molecular_cross_sections.jl:66-67)
    nHI = to_layers(get(n_densities, Korg.species"H I", NaN), L)
    nHeI = to_layers(get(n_densities, Korg.species"He I", NaN), L)
    nH2 = to_layers(get(n_densities, Korg.species"H2", NaN), L)
    xi = to_layers(ξ, L)
    ne = to_layers(nₑ, L)
    T = Float64.(temps)
    mode, knots, cvals = continuum_arrays(α_cntm, ll, L)
    all_wls = λs.all_wls
    arrays = (wl, log_gf, El, grad, gstark, vdW1, vdW2, sp, mass, ismol, n_div_U, nHI, nHeI, nH2,
              xi, ne, T, knots, cvals, all_wls)
    grid = KorgGrid(length(all_wls), pointer(all_wls))
    lines = KorgLines(length(wl), pointer(wl), pointer(log_gf), pointer(El), pointer(grad),
                      pointer(gstark), pointer(vdW1), pointer(vdW2), pointer(sp), C_NULL, C_NULL)
    species = KorgSpecies(length(specs), pointer(mass), pointer(ismol))
    layers = KorgLayers(L, pointer(T), pointer(ne), pointer(n_div_U), pointer(nHI),
                        pointer(nHeI), pointer(nH2), pointer(xi))
    cntm = KorgContinuum(mode, length(knots), isempty(knots) ? C_NULL : pointer(knots),
                         pointer(cvals))
    (arrays=arrays, grid=grid, lines=lines, species=species, layers=layers, cntm=cntm)
end

const Float64Line = Korg.Line{Float64,Float64,Float64,Float64,Float64,Float64}

# The sorted line list and its species table; logs Korg's error for H I (src/line_absorption.jl:45-47).
function prepare_lines(linelist)
    ll = issorted(linelist; by=l -> l.wl) ? linelist : sort(linelist; by=l -> l.wl)
    specs = unique(l.species for l in ll)
    if Korg.species"H I" in specs
        @error "Atomic hydrogen should not be in the linelist. Korg has built-in hydrogen lines."
    end
    ll, specs
end

# alpha as the library's column-major n_layers x n_wl block: α itself when it is one (a Matrix,
# or a view with unit stride down a layer column and columns n_layers apart, e.g.
# view(α_ref, :, 1) of a vector), else a zeroed scratch matrix whose sum the caller adds back
# (e.g. the strided view(α, i, :, :) of the molecular precompute).
is_library_layout(α) = α isa StridedArray && stride(α, 1) == 1 &&
                       (ndims(α) == 1 || stride(α, 2) == size(α, 1))

function add_lines!(α::AbstractVecOrMat{Float64}, args, cutoff_threshold)
    direct = is_library_layout(α)
    buf = direct ? α : zeros(Float64, size(α, 1), size(α, 2))
    GC.@preserve buf args begin   # args references every array the structs point into
        lock(LOCK) do
            rc = ccall((:korg_line_absorption, LIB), Cint,
                       (Ptr{Cvoid}, Ptr{Float64}, Cint, Ref{KorgGrid}, Ref{KorgLines},
                        Ref{KorgSpecies}, Ref{KorgLayers}, Ref{KorgContinuum}, Cdouble, Cint),
                       context(), pointer(buf), Cint(0), args.grid, args.lines, args.species,
                       args.layers, args.cntm, Float64(cutoff_threshold), PRECISION[])
            check(rc)
        end
    end
    direct || (α .+= buf)
    nothing
end

function Korg.line_absorption!(α::AbstractVecOrMat{Float64},
                               linelist::AbstractVector{<:Float64Line},
                               λs::Korg.Wavelengths, temps::AbstractVector{Float64}, nₑ,
                               n_densities, partition_fns, ξ, α_cntm;
                               cutoff_threshold=3e-4, tasks_per_thread=1)
    isempty(linelist) && return nothing                       # src/line_absorption.jl:35-37
    size(α, 1) == length(temps) && size(α, 2) == length(λs.all_wls) ||
        throw(DimensionMismatch("α must be n_layers × n_wavelengths"))
    ll, specs = prepare_lines(linelist)
    args = flatten(ll, specs, λs, temps, nₑ, n_densities, partition_fns, ξ, α_cntm)
    add_lines!(α, args, cutoff_threshold)
end

# ---------------------------------------------------------------------------------------------
# Forward mode: ForwardDiff.Dual alpha.  The values go through korg_line_absorption (as the
# Float64 method), the partials through korg_line_absorption_jvp: partial k of alpha receives
# J x_k, where x_k is partial k of each per-layer input (T, n_e, xi, the vdW perturbers, and n/U,
# whose Dual arithmetic carries both the densities' and the partition functions' derivatives).

struct KorgLayerTangents
    T::Ptr{Float64}
    n_e::Ptr{Float64}
    xi::Ptr{Float64}
    n_HI::Ptr{Float64}
    n_HeI::Ptr{Float64}
    n_H2::Ptr{Float64}
    n_div_U::Ptr{Float64}
end

value_of(x) = ForwardDiff.value(x)          # Duals, possibly nested in arrays
partial_of(x::ForwardDiff.Dual, k) = ForwardDiff.partials(x, k)
partial_of(x::Real, k) = 0.0

struct ValuedContinuum{F}   # a continuum callable's value (windows use values only)
    f::F
end
(c::ValuedContinuum)(λ) = value_of(c.f(λ))

function Korg.line_absorption!(α::AbstractVecOrMat{<:ForwardDiff.Dual{Tg,Float64,N}},
                               linelist::AbstractVector{<:Float64Line},
                               λs::Korg.Wavelengths, temps, nₑ, n_densities, partition_fns, ξ,
                               α_cntm; cutoff_threshold=3e-4, tasks_per_thread=1) where {Tg,N}
    isempty(linelist) && return nothing                       # src/line_absorption.jl:35-37
    L, nwl = length(temps), length(λs.all_wls)
    size(α, 1) == L && size(α, 2) == nwl ||
        throw(DimensionMismatch("α must be n_layers × n_wavelengths"))
    ll, specs = prepare_lines(linelist)
    # per-layer inputs as Duals (plain numbers have zero partials)
    nd = Dict(s => (v isa Number ? fill(v, L) : collect(v)) for (s, v) in n_densities)
    nU = [nd[s] ./ partition_fns[s].(log.(temps)) for s in specs]   # Dual arithmetic
    ξv = ξ isa Number ? fill(ξ, L) : collect(ξ)
    get_nd(code) = haskey(nd, code) ? nd[code] : fill(NaN, L)
    # values: the Float64 call's arguments
    vals = Dict(s => Float64.(value_of.(v)) for (s, v) in nd)
    pf_vals = Dict(s => (lnT -> value_of(partition_fns[s](lnT))) for s in specs)
    cntm_vals = all(f -> f isa Interpolations.Extrapolation && eltype(f.itp.coefs) == Float64, α_cntm) ?
                α_cntm : [ValuedContinuum(f) for f in α_cntm]
    args = flatten(ll, specs, λs, Float64.(value_of.(temps)), Float64.(value_of.(nₑ)), vals,
                   pf_vals, Float64.(value_of.(ξv)), cntm_vals)
    αv = zeros(Float64, L, nwl)
    add_lines!(αv, args, cutoff_threshold)
    # tangents, one direction per partial; layout of korg_layer_tangents
    S = length(specs)
    tT = [Float64[partial_of(x, k) for x in temps] for k in 1:N]
    tne = [Float64[partial_of(x, k) for x in nₑ] for k in 1:N]
    txi = [Float64[partial_of(x, k) for x in ξv] for k in 1:N]
    tHI = [Float64[partial_of(x, k) for x in get_nd(Korg.species"H I")] for k in 1:N]
    tHeI = [Float64[partial_of(x, k) for x in get_nd(Korg.species"He I")] for k in 1:N]
    tH2 = [Float64[partial_of(x, k) for x in get_nd(Korg.species"H2")] for k in 1:N]
    tnU = [reduce(hcat, [Float64[partial_of(x, k) for x in nU[i]] for i in 1:S]) for k in 1:N]
    dα = [zeros(Float64, L, nwl) for _ in 1:N]
    GC.@preserve args tT tne txi tHI tHeI tH2 tnU dα begin
        tangents = [KorgLayerTangents(pointer(tT[k]), pointer(tne[k]), pointer(txi[k]),
                                      pointer(tHI[k]), pointer(tHeI[k]), pointer(tH2[k]),
                                      pointer(tnU[k])) for k in 1:N]
        ptrs = [pointer(d) for d in dα]
        lock(LOCK) do
            check(ccall((:korg_line_absorption_jvp, LIB), Cint,
                        (Ptr{Cvoid}, Ptr{Ptr{Float64}}, Cint, Ref{KorgGrid}, Ref{KorgLines},
                         Ref{KorgSpecies}, Ref{KorgLayers}, Ref{KorgContinuum}, Cdouble, Int64,
                         Ptr{KorgLayerTangents}),
                        context(), ptrs, Cint(0), args.grid, args.lines, args.species,
                        args.layers, args.cntm, Float64(cutoff_threshold), Int64(N), tangents))
        end
    end
    for j in 1:nwl, l in 1:L
        a = α[l, j]
        α[l, j] = ForwardDiff.Dual{Tg}(ForwardDiff.value(a) + αv[l, j],
                                       ForwardDiff.Partials(ntuple(k -> ForwardDiff.partials(a, k) +
                                                                        dα[k][l, j], N)))
    end
    nothing
end

# ---------------------------------------------------------------------------------------------
# Molecular cross-sections: Korg's interpolate_molecular_cross_sections!
# (src/molecular_cross_sections.jl:99-114) for Float64 alpha through korg_molecular_cross_sections.
# (The precompute, MolecularCrossSection(linelist, wls), calls line_absorption! on views of its
# table and so already reaches the GPU through the methods above.)

struct KorgMolXsection
    n_vmic::Int32
    n_logT::Int32
    n_wl::Int64
    vmic::Ptr{Float64}
    logT::Ptr{Float64}
    wls::Ptr{Float64}
    sigma::Ptr{Float64}
    n_species::Ptr{Float64}
end

function Korg.interpolate_molecular_cross_sections!(α::AbstractVecOrMat{Float64},
                                                    molecular_cross_sections, λs::Korg.Wavelengths,
                                                    Ts::AbstractVector{Float64}, vmic,
                                                    number_densities)
    isempty(molecular_cross_sections) && return nothing     # src/molecular_cross_sections.jl:102-104
    L = length(Ts)
    vm = to_layers(vmic, L)
    T = Float64.(Ts)
    all_wls = collect(Float64, λs)
    keep = Any[vm, T, all_wls]
    tabs = KorgMolXsection[]
    for σ in molecular_cross_sections
        knots = σ.itp.itp.knots
        v, lt, w = collect(Float64, knots[1]), collect(Float64, knots[2]), collect(Float64, knots[3])
        # Korg's coefs[iv, iT, j] -> the library's sigma[(iv * n_wl + j) * n_logT + iT]
        coefs = permutedims(σ.itp.itp.coefs, (2, 3, 1))
        n = to_layers(number_densities[σ.species], L)
        push!(keep, v, lt, w, coefs, n)
        push!(tabs, KorgMolXsection(length(v), length(lt), length(w), pointer(v), pointer(lt),
                                    pointer(w), pointer(coefs), pointer(n)))
    end
    direct = is_library_layout(α)
    buf = direct ? α : zeros(Float64, size(α, 1), size(α, 2))
    grid = KorgGrid(length(all_wls), pointer(all_wls))
    GC.@preserve buf keep tabs begin
        lock(LOCK) do
            check(ccall((:korg_molecular_cross_sections, LIB), Cint,
                        (Ptr{Cvoid}, Ptr{Float64}, Cint, Ref{KorgGrid}, Int32, Ptr{Float64},
                         Ptr{Float64}, Int64, Ptr{KorgMolXsection}),
                        context(), pointer(buf), Cint(0), grid, Int32(L), pointer(T), pointer(vm),
                        Int64(length(tabs)), tabs))
        end
    end
    direct || (α .+= buf)
    nothing
end

# ---------------------------------------------------------------------------------------------
# Reverse mode.  line_absorption! mutates its first argument, so the differentiable form is the
# non-mutating line opacity; its rrule calls korg_line_absorption_vjp (include/korg_line.h).
# The windows are held fixed (they are integer functions of the inputs, as under ForwardDiff).

struct KorgLayerGrads
    T::Ptr{Float64}
    n_e::Ptr{Float64}
    xi::Ptr{Float64}
    n_HI::Ptr{Float64}
    n_HeI::Ptr{Float64}
    n_H2::Ptr{Float64}
    n_div_U::Ptr{Float64}
end

"""
    line_opacity(linelist, λs, temps, nₑ, n_densities, partition_fns, ξ, α_cntm; cutoff_threshold)

The line opacity that `line_absorption!` adds, as a new `n_layers × n_wavelengths` matrix.
"""
function line_opacity(linelist, λs, temps, nₑ, n_densities, partition_fns, ξ, α_cntm;
                      cutoff_threshold=3e-4)
    α = zeros(Float64, length(temps), length(λs.all_wls))
    Korg.line_absorption!(α, linelist, λs, temps, nₑ, n_densities, partition_fns, ξ, α_cntm;
                          cutoff_threshold=cutoff_threshold)
    α
end

function ChainRulesCore.rrule(::typeof(line_opacity), linelist, λs, temps, nₑ, n_densities,
                              partition_fns, ξ, α_cntm; cutoff_threshold=3e-4)
    α = line_opacity(linelist, λs, temps, nₑ, n_densities, partition_fns, ξ, α_cntm;
                     cutoff_threshold=cutoff_threshold)
    function pullback(ᾱ)
        G = Matrix{Float64}(unthunk(ᾱ))
        ll = issorted(linelist; by=l -> l.wl) ? linelist : sort(linelist; by=l -> l.wl)
        L = length(temps)
        specs = unique(l.species for l in ll)
        gT, gne, gxi = zeros(L), zeros(L), zeros(L)
        gHI, gHeI, gH2 = zeros(L), zeros(L), zeros(L)
        gnU = zeros(L, length(specs))          # column i = species i (n_div_U[s*L + l])
        # the same flattening as the forward method (sorted lines, species table, n/U, ...)
        args = flatten(ll, specs, λs, temps, nₑ, n_densities, partition_fns, ξ, α_cntm)
        GC.@preserve G gT gne gxi gHI gHeI gH2 gnU args begin
            grads = KorgLayerGrads(pointer(gT), pointer(gne), pointer(gxi), pointer(gHI),
                                   pointer(gHeI), pointer(gH2), pointer(gnU))
            lock(LOCK) do
                check(ccall((:korg_line_absorption_vjp, LIB), Cint,
                            (Ptr{Cvoid}, Ptr{Float64}, Cint, Ref{KorgGrid}, Ref{KorgLines},
                             Ref{KorgSpecies}, Ref{KorgLayers}, Ref{KorgContinuum}, Cdouble,
                             Ref{KorgLayerGrads}),
                            context(), G, Cint(0), args.grid, args.lines, args.species,
                            args.layers, args.cntm, Float64(cutoff_threshold), grads))
            end
        end
        # n/U -> n_densities, and the partition functions' dependence on T
        dn = Dict{Any,Vector{Float64}}()
        for (i, s) in enumerate(specs)
            U = partition_fns[s].(log.(temps))
            dU = [ForwardDiff.derivative(partition_fns[s], log(T)) for T in temps]
            dn[s] = gnU[:, i] ./ U
            gT .+= gnU[:, i] .* (-n_densities[s] .* dU ./ (U .^ 2 .* temps))
        end
        for (code, g) in ((Korg.species"H I", gHI), (Korg.species"He I", gHeI),
                          (Korg.species"H2", gH2))
            haskey(n_densities, code) && (dn[code] = get(dn, code, zeros(L)) .+ g)
        end
        ξ̄ = ξ isa Number ? sum(gxi) : gxi
        return (NoTangent(), NoTangent(), NoTangent(), gT, gne, Tangent{typeof(n_densities)}(dn),
                NoTangent(), ξ̄, NoTangent())
    end
    return α, pullback
end

end # module
