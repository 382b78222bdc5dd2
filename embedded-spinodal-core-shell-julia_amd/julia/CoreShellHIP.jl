# CoreShellHIP: Julia binding of libcsg.so (C ABI in include/csg.h).
#
# Drop-in for the time-step loop of core_shell.jl: the functions below keep
# the names, arguments, return values and error behaviour of
# core_shell.jl:4-103 (Parameters, make_grid, initial_fields, geometry,
# bulk_energy, bulk_derivative, chemical_potential, transport_rate, step,
# energy, run_simulation) and run the arithmetic on one MI355X through ccall.
# Extra keyword arguments (defaults reproduce the reference): precision
# (:fp64 or :mixed), history_every, device.
module CoreShellHIP

using DelimitedFiles
using Libdl

export Parameters, make_grid, initial_fields, geometry, bulk_energy, bulk_derivative,
       chemical_potential, transport_rate, step, energy, run_simulation, run_ensemble

const libcsg = joinpath(@__DIR__, "..", "libcsg.so")

const CSG_OK = Cint(0)
const CSG_ERR_NONFINITE = Cint(3)

Base.@kwdef struct Parameters          # core_shell.jl:4-15, same field order as csg_params
    A::Float64 = 2.0
    B::Float64 = 4.0
    chi::Float64 = 0.5
    P::Float64 = 5.0
    ca::Float64 = 0.5
    cb::Float64 = 0.0
    cc::Float64 = 1.0
    kappa::Float64 = 2.0
    mobility::Float64 = 1.0
    Q::Float64 = 0.5
end

lasterror() = unsafe_string(ccall((:csg_last_error, libcsg), Cstring, ()))
check(rc) = rc == CSG_OK || error("csg: " * lasterror())

# One device context (csg_create).  `generation` counts field uploads and
# parameter changes, so a holder of checkpoints (the adjoint rule) can tell
# whether the context has been used for something else since.
mutable struct Context
    ptr::Ptr{Cvoid}
    n::Int
    p::Parameters
    phi_key::Union{Nothing,UInt}   # fingerprint of the resident phi (nothing: unknown)
    generation::Int
    function Context(n::Integer, spacing::Real, p::Parameters, precision::Symbol, device::Integer)
        out = Ref{Ptr{Cvoid}}(C_NULL)
        prec = precision === :fp64 ? Cint(0) : precision === :mixed ? Cint(1) :
               throw(ArgumentError("precision must be :fp64 or :mixed"))
        pref = Ref(p)
        check(ccall((:csg_create, libcsg), Cint,
                    (Ref{Ptr{Cvoid}}, Int32, Float64, Ref{Parameters}, Int32, Int32),
                    out, n, spacing, pref, prec, device))
        ctx = new(out[], Int(n), p, nothing, 0)
        finalizer(destroy!, ctx)
        return ctx
    end
end

# Free the device context now (also the finalizer); idempotent.
function destroy!(ctx::Context)
    if ctx.ptr != C_NULL
        ccall((:csg_destroy, libcsg), Cint, (Ptr{Cvoid},), ctx.ptr)
        ctx.ptr = C_NULL
    end
    return nothing
end

# Swap the Parameters in place (csg_set_params: nothing is reallocated).
function set_params!(ctx::Context, p::Parameters)
    ctx.p == p && return ctx
    pref = Ref(p)
    check(ccall((:csg_set_params, libcsg), Cint, (Ptr{Cvoid}, Ref{Parameters}), ctx.ptr, pref))
    ctx.p = p
    ctx.generation += 1
    return ctx
end

function set_fields!(ctx::Context, c0::Matrix{Float64}, phi::Matrix{Float64})
    GC.@preserve c0 phi check(ccall((:csg_set_fields, libcsg), Cint,
                                    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), ctx.ptr, c0, phi))
    ctx.phi_key = nothing   # not fingerprinted here; the next `context` call uploads its phi
    ctx.generation += 1
    return ctx
end

function field(ctx::Context)
    c = Matrix{Float64}(undef, ctx.n, ctx.n)
    GC.@preserve c check(ccall((:csg_get_field, libcsg), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx.ptr, c))
    return c
end

# make_grid (core_shell.jl:17-25): the device computes wavenumbers on the fly;
# the arrays are returned for callers that use them.
function make_grid(n, spacing)
    m = [i < (n + 1) ÷ 2 ? i : i - n for i in 0:n-1]
    k = collect(2pi .* (m .* ((1 / spacing) / n)))
    derivative_k = copy(k)
    iseven(n) && (derivative_k[n ÷ 2 + 1] = 0.0)
    kx = reshape(derivative_k, n, 1)
    ky = reshape(derivative_k, 1, n)
    k2 = reshape(k .^ 2, n, 1) .+ reshape(k .^ 2, 1, n)
    return (; kx, ky, k2, spacing)
end

# initial_fields (core_shell.jl:27-37), Julia-identical noise (computed in C++).
function initial_fields(n; spacing=1.0, radius=70.0, width=4.0, seed=5749)
    c = Matrix{Float64}(undef, n, n)
    phi = Matrix{Float64}(undef, n, n)
    GC.@preserve c phi check(ccall((:csg_initial_fields, libcsg), Cint,
                                   (Int32, Float64, Float64, Float64, Int64, Ptr{Float64}, Ptr{Float64}),
                                   n, spacing, radius, width, seed, c, phi))
    return c, phi
end

# Contexts reused by the function API: one per (n, spacing, precision,
# device), Parameters swapped in place and phi re-uploaded only when it
# changes; at most MAX_CONTEXTS are kept, the least recently used is freed
# first.  `cache` and `order` let the adjoint rule keep its own, smaller set.
const MAX_CONTEXTS = 8
const ContextKey = Tuple{Int,Float64,Symbol,Int}
const contexts = Dict{ContextKey,Context}()
const context_order = ContextKey[]   # least recently used first

function cached_context(n, spacing, p, precision, device; cache=contexts, order=context_order,
                        limit=MAX_CONTEXTS)
    key = (Int(n), Float64(spacing), precision, Int(device))
    ctx = get(cache, key, nothing)
    if ctx === nothing || ctx.ptr == C_NULL
        while length(order) >= limit
            destroy!(pop!(cache, popfirst!(order)))
        end
        filter!(!=(key), order)
        ctx = Context(n, spacing, p, precision, device)
        cache[key] = ctx
    else
        filter!(!=(key), order)
        set_params!(ctx, p)
    end
    push!(order, key)
    return ctx
end

# Fingerprint of every element of phi.  Base.hash of an array of 8192 or more
# elements hashes only about log2(length) of them, mostly near its end, so two
# particles of one size that differ only near the centre could hash alike and
# the device would keep the wrong phi.
phi_fingerprint(phi) = foldl((h, x) -> hash(x, h), phi; init=hash(size(phi)))

# The fingerprint a shape carries (taken once by `geometry`), else computed.
shape_key(shape) = hasproperty(shape, :key) ? shape.key : phi_fingerprint(shape.phi)

function context(phi, spacing, p, precision, device; key=phi_fingerprint(phi))
    ctx = cached_context(size(phi, 1), spacing, p, precision, device)
    if ctx.phi_key != key
        m = Matrix{Float64}(phi)
        set_fields!(ctx, m, m)
        ctx.phi_key = key
    end
    return ctx
end

# geometry (core_shell.jl:39-43); the shape keeps a snapshot of phi, and its
# fingerprint, for the device calls.
function geometry(phi, p; precision=:fp64, device=0)
    snap = Matrix{Float64}(phi)
    key = phi_fingerprint(snap)
    ctx = context(snap, 1.0, p, precision, device; key)
    n = size(phi, 1)
    h, g, M = (Matrix{Float64}(undef, n, n) for _ in 1:3)
    GC.@preserve h g M check(ccall((:csg_geometry, libcsg), Cint,
                                   (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), ctx.ptr, h, g, M))
    return (; h, g, M, phi=snap, key)
end

# entry points of the form (ctx, const double* c, double* out)
function pointwise(sym, c, shape, p, spacing, precision, device)
    ctx = context(shape.phi, spacing, p, precision, device; key=shape_key(shape))
    fptr = Libdl.dlsym(Libdl.dlopen(libcsg), sym)
    cc = Matrix{Float64}(c)
    out = similar(cc)
    GC.@preserve cc out check(ccall(fptr, Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), ctx.ptr, cc, out))
    return out
end

bulk_energy(c, shape, p; precision=:fp64, device=0) =
    pointwise(:csg_bulk_energy, c, shape, p, 1.0, precision, device)          # core_shell.jl:45-49
bulk_derivative(c, shape, p; precision=:fp64, device=0) =
    pointwise(:csg_bulk_derivative, c, shape, p, 1.0, precision, device)      # core_shell.jl:51-55
chemical_potential(c, shape, grid, p; precision=:fp64, device=0) =
    pointwise(:csg_chemical_potential, c, shape, p, grid.spacing, precision, device)  # :57-60

# transport_rate (core_shell.jl:62-70): full n x n FFTW-convention spectrum.
function transport_rate(c, shape, grid, p; precision=:fp64, device=0)
    ctx = context(shape.phi, grid.spacing, p, precision, device; key=shape_key(shape))
    cc = Matrix{Float64}(c)
    re, im = similar(cc), similar(cc)
    d = Ref(0.0)
    GC.@preserve cc re im check(ccall((:csg_transport_rate, libcsg), Cint,
                                      (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{Float64}),
                                      ctx.ptr, cc, re, im, d))
    return complex.(re, im), d[]
end

# energy (core_shell.jl:78-82)
function energy(c, shape, grid, p; precision=:fp64, device=0)
    ctx = context(shape.phi, grid.spacing, p, precision, device; key=shape_key(shape))
    cc = Matrix{Float64}(c)
    F = Ref(0.0)
    GC.@preserve cc check(ccall((:csg_energy, libcsg), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ref{Float64}),
                                ctx.ptr, cc, F))
    return F[]
end

# step (core_shell.jl:72-76)
function step(c, shape, grid, p, dt; precision=:fp64, device=0)
    ctx = cached_context(size(c, 1), grid.spacing, p, precision, device)
    set_fields!(ctx, Matrix{Float64}(c), Matrix{Float64}(shape.phi))
    check(ccall((:csg_step, libcsg), Cint, (Ptr{Cvoid}, Float64, Int64), ctx.ptr, dt, 1))
    return field(ctx)
end

# run_simulation (core_shell.jl:84-103)
function run_simulation(c0, phi; dt=0.01, steps=1000, spacing=1.0, p=Parameters(),
                        directory=nothing, precision=:fp64, history_every=1, device=0)
    ctx = cached_context(size(c0, 1), spacing, p, precision, device)
    set_fields!(ctx, Matrix{Float64}(c0), Matrix{Float64}(phi))
    rows = steps ÷ history_every + 1
    history = zeros(rows, 6)
    rc = GC.@preserve history ccall((:csg_run, libcsg), Cint, (Ptr{Cvoid}, Float64, Int64, Int64, Ptr{Float64}),
                                    ctx.ptr, dt, steps, history_every, history)
    if rc == CSG_ERR_NONFINITE
        s = ccall((:csg_nonfinite_step, libcsg), Int64, (Ptr{Cvoid},), ctx.ptr)
        error("Non-finite composition at step $s")
    end
    check(rc)
    c = field(ctx)
    if directory !== nothing
        mkpath(directory)
        writedlm(joinpath(directory, "history.csv"), history, ',')
        writedlm(joinpath(directory, "composition.csv"), c, ',')
        writedlm(joinpath(directory, "particle.csv"), phi, ',')
    end
    return c, history
end

# ---- ensembles (csg_batch_*): run_simulation of many particles at once ----
# Returns [(c, history), ...]; a non-finite member raises after the run with
# every member's first non-finite step (-1: finite).
function run_ensemble(c0s::Vector, phis::Vector; dt=0.01, steps=1000, spacings=nothing,
                      params=nothing, precision=:mixed, history_every=1, device=0)
    k = length(c0s)
    ns = Int32[size(c, 1) for c in c0s]
    sp = spacings === nothing ? C_NULL : Float64.(spacings)
    ps = params === nothing ? C_NULL : Parameters[p for p in params]
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve ns sp ps check(ccall((:csg_batch_create, libcsg), Cint,
                                     (Ref{Ptr{Cvoid}}, Int32, Ptr{Int32}, Ptr{Float64}, Ptr{Parameters}, Int32, Int32),
                                     out, k, ns, sp, ps, precision === :mixed ? 1 : 0, device))
    b = out[]
    try
        for i in 1:k
            c0m, phim = Matrix{Float64}(c0s[i]), Matrix{Float64}(phis[i])
            GC.@preserve c0m phim check(ccall((:csg_batch_set_fields, libcsg), Cint,
                                              (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}), b, i - 1, c0m, phim))
        end
        rows = steps ÷ history_every + 1
        hs = [zeros(rows, 6) for _ in 1:k]
        ptrs = [pointer(h) for h in hs]
        rc = GC.@preserve hs ptrs ccall((:csg_batch_run, libcsg), Cint,
                                        (Ptr{Cvoid}, Float64, Int64, Int64, Ptr{Ptr{Float64}}),
                                        b, dt, steps, history_every, ptrs)
        if rc == CSG_ERR_NONFINITE
            bad = [ccall((:csg_batch_nonfinite_step, libcsg), Int64, (Ptr{Cvoid}, Int32), b, i - 1) for i in 1:k]
            error("Non-finite composition in ensemble members $(findall(>=(0), bad)) at steps $(filter(>=(0), bad))")
        end
        check(rc)
        res = Tuple{Matrix{Float64}, Matrix{Float64}}[]
        for i in 1:k
            c = Matrix{Float64}(undef, ns[i], ns[i])
            GC.@preserve c check(ccall((:csg_batch_get_field, libcsg), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                                       b, i - 1, c))
            push!(res, (c, hs[i]))
        end
        return res
    finally
        ccall((:csg_batch_destroy, libcsg), Cint, (Ptr{Cvoid},), b)
    end
end

# ---- kernel schedules: tune a context's step passes (csg_autotune) ----
# Returns (selection per pass, device microseconds per loop iteration).
function autotune!(ctx::Context; dt=0.01, rounds=2)
    sel = zeros(Int32, 4)
    us = Ref{Float64}(0.0)
    GC.@preserve sel check(ccall((:csg_autotune, libcsg), Cint, (Ptr{Cvoid}, Float64, Int32, Ptr{Int32}, Ref{Float64}),
                                 ctx.ptr, dt, rounds, sel, us))
    return Tuple(sel), us[]
end

function set_schedule!(ctx::Context, sel)
    s = Int32.(collect(sel))
    GC.@preserve s check(ccall((:csg_set_schedule, libcsg), Cint, (Ptr{Cvoid}, Ptr{Int32}), ctx.ptr, s))
end

end # module
