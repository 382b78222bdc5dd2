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
module CoreShellHIPChainRules

using ChainRulesCore
using ..CoreShellHIP: Parameters, Context, set_fields!, field, check, libcsg, energy, geometry,
                      make_grid, chemical_potential

# Device memory the pullback may spend on keeping every state of the forward
# run ((steps + 1) FP64 fields of N^2 values), as adjoint.py does.
const KEEP_ALL_BYTES = 32 * 2^30

default_every(n, steps) = (steps + 1) * n * n * 8 <= KEEP_ALL_BYTES ? 1 : max(1, round(Int, sqrt(steps)))

function forward_ckpt(c0, phi, p, dt, steps, every, spacing)
    ctx = Context(size(c0, 1), spacing, p, :fp64, 0)
    set_fields!(ctx, Matrix{Float64}(c0), Matrix{Float64}(phi))
    check(ccall((:csg_adjoint_forward, libcsg), Cint, (Ptr{Cvoid}, Float64, Int64, Int64),
                ctx.ptr, dt, steps, every))
    return ctx, field(ctx)
end

function final_loss(c0, kappa, P, chi; phi, target, lam=0.0, dt=0.01, steps=1000, spacing=1.0,
                    base=Parameters())
    p = Parameters(base.A, base.B, chi, P, base.ca, base.cb, base.cc, kappa, base.mobility, base.Q)
    ctx, cT = forward_ckpt(c0, phi, p, dt, steps, max(1, round(Int, sqrt(steps))), spacing)
    shape, grid = geometry(phi, p), make_grid(size(c0, 1), spacing)
    return 0.5 * sum(abs2, cT .- target) + lam * energy(cT, shape, grid, p)
end

function ChainRulesCore.rrule(::typeof(final_loss), c0, kappa, P, chi; phi, target, lam=0.0, dt=0.01,
                              steps=1000, spacing=1.0, base=Parameters())
    p = Parameters(base.A, base.B, chi, P, base.ca, base.cb, base.cc, kappa, base.mobility, base.Q)
    ctx, cT = forward_ckpt(c0, phi, p, dt, steps, default_every(size(c0, 1), steps), spacing)
    shape, grid = geometry(phi, p), make_grid(size(c0, 1), spacing)
    L = 0.5 * sum(abs2, cT .- target) + lam * energy(cT, shape, grid, p)
    function pullback(dL)
        gbar = Matrix{Float64}(cT .- target)
        direct = zeros(3)
        if lam != 0
            gbar .+= lam * spacing^2 .* chemical_potential(cT, shape, grid, p)
            # energy is affine in kappa, P and chi
            for (i, f) in enumerate((:kappa, :P, :chi))
                hi = Parameters(; (k => (k == f ? 1.0 : getfield(p, k)) for k in fieldnames(Parameters))...)
                lo = Parameters(; (k => (k == f ? 0.0 : getfield(p, k)) for k in fieldnames(Parameters))...)
                direct[i] = lam * (energy(cT, geometry(phi, hi), grid, hi) - energy(cT, geometry(phi, lo), grid, lo))
            end
        end
        dc0 = similar(gbar)
        dp = zeros(3)
        GC.@preserve gbar dc0 dp check(ccall((:csg_adjoint_backward, libcsg), Cint,
                                             (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                                             ctx.ptr, gbar, dc0, dp))
        g = dp .+ direct
        return NoTangent(), dL .* dc0, dL * g[1], dL * g[2], dL * g[3]
    end
    return L, pullback
end

end # module
