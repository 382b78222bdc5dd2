# ChainRulesCore rule for the final-state loss of the HIP time loop.
#
#   final_loss(c0, kappa, P, chi; phi, target, lam, dt, steps, spacing, base)
#     = 1/2 ||c(T) - target||^2 + lam * energy(c(T))
#
# where c(T) is `steps` applications of step (core_shell.jl:72-76) from c0
# with Parameters `base` except kappa, P, chi.  The pullback runs the
# hand-written adjoint on the GPU (csg_adjoint_forward / csg_adjoint_backward;
# every state kept when the run fits KEEP_ALL_BYTES of device memory, else
# sqrt(steps) checkpoints) and returns cotangents for c0, kappa, P and chi.
#
# Device contexts are reused, not created per call: the forward runs on one
# FP64 context per (n, spacing, device) from a small cache of its own (its
# checkpoints can hold gigabytes), with the parameters swapped in place; the
# energies and chemical potentials go through CoreShellHIP's function-API
# cache.  If the adjoint context has been reused or evicted between the rule
# and its pullback, the pullback runs the forward again before the backward.
module CoreShellHIPChainRules

using ChainRulesCore
using ..CoreShellHIP: Parameters, Context, ContextKey, cached_context, set_fields!, field, check, libcsg, energy,
                      geometry, make_grid, chemical_potential

# Device memory the pullback may spend on keeping every state of the forward
# run ((steps + 1) FP64 fields of N^2 values), as adjoint.py does.
const KEEP_ALL_BYTES = 32 * 2^30
const CSG_ERR_HIP = Cint(2)

default_every(n, steps) = (steps + 1) * n * n * 8 <= KEEP_ALL_BYTES ? 1 : max(1, round(Int, sqrt(steps)))

const MAX_ADJOINT_CONTEXTS = 2
const adjoint_contexts = Dict{ContextKey,Context}()
const adjoint_order = ContextKey[]

adjoint_context(n, spacing, p, device) =
    cached_context(n, spacing, p, :fp64, device; cache=adjoint_contexts, order=adjoint_order,
                   limit=MAX_ADJOINT_CONTEXTS)

with_kpc(b::Parameters, kappa, P, chi) =
    Parameters(b.A, b.B, chi, P, b.ca, b.cb, b.cc, kappa, b.mobility, b.Q)

# Parameters `p` with field `f` set to `v`.
with_field(p::Parameters, f::Symbol, v) =
    Parameters(; (k => (k == f ? Float64(v) : getfield(p, k)) for k in fieldnames(Parameters))...)

# Checkpointed forward run on the adjoint context; when keeping every state
# does not fit the device, the library frees what it allocated and the run
# is repeated with ~sqrt(steps) checkpoints.  Returns (ctx, c(T), generation).
function forward_ckpt(c0, phi, p, dt, steps, spacing; device=0)
    n = size(c0, 1)
    ctx = adjoint_context(n, spacing, p, device)
    every = default_every(n, steps)
    c0m, phim = Matrix{Float64}(c0), Matrix{Float64}(phi)
    set_fields!(ctx, c0m, phim)
    rc = ccall((:csg_adjoint_forward, libcsg), Cint, (Ptr{Cvoid}, Float64, Int64, Int64), ctx.ptr, dt, steps, every)
    if rc == CSG_ERR_HIP && every == 1
        set_fields!(ctx, c0m, phim)
        rc = ccall((:csg_adjoint_forward, libcsg), Cint, (Ptr{Cvoid}, Float64, Int64, Int64), ctx.ptr, dt, steps,
                   max(1, round(Int, sqrt(steps))))
    end
    check(rc)
    return ctx, field(ctx), ctx.generation
end

# The loss alone: plain steps on a function-API context, no checkpoints.
function final_loss(c0, kappa, P, chi; phi, target, lam=0.0, dt=0.01, steps=1000, spacing=1.0,
                    base=Parameters(), device=0)
    p = with_kpc(base, kappa, P, chi)
    n = size(c0, 1)
    ctx = cached_context(n, spacing, p, :fp64, device)
    set_fields!(ctx, Matrix{Float64}(c0), Matrix{Float64}(phi))
    check(ccall((:csg_step, libcsg), Cint, (Ptr{Cvoid}, Float64, Int64), ctx.ptr, dt, steps))
    cT = field(ctx)
    shape, grid = geometry(phi, p), make_grid(n, spacing)
    return 0.5 * sum(abs2, cT .- target) + lam * energy(cT, shape, grid, p)
end

function ChainRulesCore.rrule(::typeof(final_loss), c0, kappa, P, chi; phi, target, lam=0.0, dt=0.01,
                              steps=1000, spacing=1.0, base=Parameters(), device=0)
    p = with_kpc(base, kappa, P, chi)
    ctx, cT, gen = forward_ckpt(c0, phi, p, dt, steps, spacing; device=device)
    shape, grid = geometry(phi, p), make_grid(size(c0, 1), spacing)
    L = 0.5 * sum(abs2, cT .- target) + lam * energy(cT, shape, grid, p)
    function pullback(dL)
        actx = ctx
        if actx.ptr == C_NULL || actx.generation != gen
            actx, _, _ = forward_ckpt(c0, phi, p, dt, steps, spacing; device=device)
        end
        gbar = Matrix{Float64}(cT .- target)
        if lam != 0
            gbar .+= lam * spacing^2 .* chemical_potential(cT, shape, grid, p)
        end
        dc0 = similar(gbar)
        dp = zeros(3)
        GC.@preserve gbar dc0 dp check(ccall((:csg_adjoint_backward, libcsg), Cint,
                                             (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                                             actx.ptr, gbar, dc0, dp))
        # the loss's direct parameter terms: energy is affine in kappa, P and
        # chi (geometry does not depend on them, so `shape` is reused)
        direct = zeros(3)
        if lam != 0
            for (i, f) in enumerate((:kappa, :P, :chi))
                hi, lo = with_field(p, f, 1.0), with_field(p, f, 0.0)
                direct[i] = lam * (energy(cT, shape, grid, hi) - energy(cT, shape, grid, lo))
            end
        end
        g = dp .+ direct
        return NoTangent(), dL .* dc0, dL * g[1], dL * g[2], dL * g[3]
    end
    return L, pullback
end

end # module
