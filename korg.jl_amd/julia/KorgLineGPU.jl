# KorgLineGPU.jl -- Julia binding of libkorgline.so (include/korg_line.h).
#
# Loading this module adds a more specific method of Korg.line_absorption! for Float64 alpha
# matrices and Float64 lines, so that Korg's own calls inside `synthesize`
# (src/synthesize.jl:266-276 and :305-306) dispatch to the GPU.  Other element types (e.g.
# ForwardDiff.Dual during fitting) keep using Korg's CPU method.
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
    nan = fill(NaN, L)
    nHI = Float64.(get(n_densities, Korg.species"H I", nan))
    nHeI = Float64.(get(n_densities, Korg.species"He I", nan))
    nH2 = Float64.(get(n_densities, Korg.species"H2", nan))
    xi = ξ isa Number ? fill(Float64(ξ), L) : Float64.(ξ)
    ne = Float64.(nₑ)
    T = Float64.(temps)
    mode, knots, cvals = continuum_arrays(α_cntm, ll, L)
    all_wls = λs.all_wls
    arrays = (wl, log_gf, El, grad, gstark, vdW1, vdW2, sp, mass, ismol, n_div_U, nHI, nHeI, nH2,
              xi, ne, T, knots, cvals, all_wls)
    grid = KorgGrid(length(all_wls), pointer(all_wls))
    lines = KorgLines(length(wl), pointer(wl), pointer(log_gf), pointer(El), pointer(grad),
                      pointer(gstark), pointer(vdW1), pointer(vdW2), pointer(sp))
    species = KorgSpecies(length(specs), pointer(mass), pointer(ismol))
    layers = KorgLayers(L, pointer(T), pointer(ne), pointer(n_div_U), pointer(nHI),
                        pointer(nHeI), pointer(nH2), pointer(xi))
    cntm = KorgContinuum(mode, length(knots), isempty(knots) ? C_NULL : pointer(knots),
                         pointer(cvals))
    (arrays=arrays, grid=grid, lines=lines, species=species, layers=layers, cntm=cntm)
end

function Korg.line_absorption!(α::Matrix{Float64},
                               linelist::AbstractVector{<:Korg.Line{Float64,Float64,Float64,
                                                                     Float64,Float64,Float64}},
                               λs::Korg.Wavelengths, temps::AbstractVector{Float64}, nₑ,
                               n_densities, partition_fns, ξ, α_cntm;
                               cutoff_threshold=3e-4, tasks_per_thread=1)
    isempty(linelist) && return nothing                       # src/line_absorption.jl:35-37
    if !issorted(linelist; by=l -> l.wl)
        linelist = sort(linelist; by=l -> l.wl)
    end
    specs = unique(l.species for l in linelist)
    if Korg.species"H I" in specs
        @error "Atomic hydrogen should not be in the linelist. Korg has built-in hydrogen lines."
    end
    args = flatten(linelist, specs, λs, temps, nₑ, n_densities, partition_fns, ξ, α_cntm)
    GC.@preserve α args begin   # args references every array the structs point into
        lock(LOCK) do
            rc = ccall((:korg_line_absorption, LIB), Cint,
                       (Ptr{Cvoid}, Ptr{Float64}, Cint, Ref{KorgGrid}, Ref{KorgLines},
                        Ref{KorgSpecies}, Ref{KorgLayers}, Ref{KorgContinuum}, Cdouble, Cint),
                       context(), α, Cint(0), args.grid, args.lines, args.species, args.layers,
                       args.cntm, Float64(cutoff_threshold), PRECISION[])
            check(rc)
        end
    end
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
