# FunicularHIPChainRulesExt: reverse-mode rules for cholqr2! and project on the
# HIP backend, as ChainRulesCore rrules. A Funicular.jl checkout lists it as a
# package extension triggered by ChainRulesCore together with the package that
# triggers FunicularHIPExt. The reference has no derivatives (ChainRulesCore is
# not among its dependencies, Project.toml:6-9); the rules follow the QR
# adjoint (Walter and Lehmann; the complex form with a real-diagonal R) and
# the product rule for S = Q^H G Q.
#
# Every pullback is a sweep over the forward pass's row blocks or panels, on
# the library's kernels (fhip_gram, fhip_gemm_nn, fhip_axpby, fhip_trsm_runc,
# fhip_rowdot), and stores no intermediate beyond its outputs: Q is the matrix
# the forward pass left behind, read panel by panel.
#
#   QR (Y = Q R):  M = R Rbar^H - Qbar^H Q,
#                  N = tril(M, -1) + tril(M, -1)^H + Re diag(M),
#                  Ybar = (Qbar + Q N) R^-H
#   S = Q^H G Q:   Qbar = G (Q Sbar^H) + G^H (Q Sbar),
#                  dbar_i = Re sum_a conj(Q_ia) (Q Sbar^H)_ia for an operator
#                  whose real diagonal d is a parameter (TridiagonalStencil's).
#
# This image has no Julia, so this file is not exercised here; the Python rules
# funicular_hip.adjoint are the same sweeps and are what the tests run.
module FunicularHIPChainRulesExt

using ChainRulesCore
using LinearAlgebra
import Funicular
using Funicular: PanelMatrix, cholqr2!, project, gram, panelmul!, rightmul!, sweep!, readpanels, writepanels,
                 updatepanels, rowtraversal, checkout_device_buffers!, checkin_device_buffers!, panelwidth, free!

# The HIP extension's module, looked up at call time: the two extensions can
# come up in either order when their trigger packages load together.
hip() = Base.get_extension(Funicular, :FunicularHIPExt)

# cholqr2 without the mutation, for callers that differentiate through Q as
# well as R: a fresh PanelMatrix Q and R with Q R = Y; Y is left as it was.
function cholqr2_qr(Y::PanelMatrix{T}; nbuffers=nothing) where {T}
    Q = PanelMatrix{T}(undef, Y.N, Y.k; plan=Y.plan, w=panelwidth(Y))
    copyto!(Q, Y)
    R = cholqr2!(Q; nbuffers=nbuffers)
    Q, R
end

function qr_pullback(Q::PanelMatrix{T}, R::Matrix{T}, Qbar, Rbar; nbuffers=nothing) where {T}
    plan = Q.plan
    b = plan.backend
    k = Q.k
    M = zeros(T, k, k)
    Rbar isa AbstractZero || (M .+= R * Rbar')
    Qbar isa AbstractZero || (M .-= gram(Qbar, Q; nbuffers=nbuffers))
    L = tril(M, -1)
    N = Matrix{T}(L + L' + Diagonal(real(diag(M))))
    Ybar = PanelMatrix{T}(undef, Q.N, k; plan=plan, w=panelwidth(Q))
    Nd, Rd = checkout_device_buffers!(plan, T, (k, k), 2)
    try
        Funicular.h2d!(Nd, N, b)
        Funicular.h2d!(Rd, R, b)
        operands = Qbar isa AbstractZero ? (readpanels(Q), writepanels(Ybar)) :
                   (readpanels(Q), readpanels(Qbar), writepanels(Ybar))
        sweep!(operands...; nbuffers=nbuffers, traversal=rowtraversal(Q)) do block, buffers...
            Qb, Yb = first(buffers), last(buffers)
            Funicular.rightmul_gemm!(b, Yb, Qb, Nd)             # Q N
            Qbar isa AbstractZero || axpy!(one(T), buffers[2], Yb)  # + Qbar
            hip().rdiv_upper_adjoint!(b, Yb, Rd, block)           # R^-H, folded on the first block
        end
    finally
        checkin_device_buffers!(plan, (Nd, Rd))
    end
    Ybar
end

function ChainRulesCore.rrule(::typeof(cholqr2!), Y::PanelMatrix{T}; nbuffers=nothing) where {T}
    R = cholqr2!(Y; nbuffers=nbuffers)  # Y holds Q from here on
    cholqr2_inplace_pullback(Rbar) = (NoTangent(), qr_pullback(Y, R, ZeroTangent(), unthunk(Rbar); nbuffers=nbuffers))
    R, cholqr2_inplace_pullback
end

function ChainRulesCore.rrule(::typeof(cholqr2_qr), Y::PanelMatrix{T}; nbuffers=nothing) where {T}
    Q, R = cholqr2_qr(Y; nbuffers=nbuffers)
    function cholqr2_qr_pullback(QRbar)
        Qbar, Rbar = unthunk(QRbar)
        (NoTangent(), qr_pullback(Q, R, unthunk(Qbar), unthunk(Rbar); nbuffers=nbuffers))
    end
    (Q, R), cholqr2_qr_pullback
end

function project_pullback(Q::PanelMatrix{T}, G, Sbar::AbstractMatrix; nbuffers=nothing) where {T}
    plan = Q.plan
    b = plan.backend
    w = panelwidth(Q)
    S1 = Matrix{T}(Sbar')
    S2 = Matrix{T}(Sbar)
    Z1 = PanelMatrix{T}(undef, Q.N, Q.k; plan=plan, w=w)
    Z2 = PanelMatrix{T}(undef, Q.N, Q.k; plan=plan, w=w)
    Qbar = PanelMatrix{T}(undef, Q.N, Q.k; plan=plan, w=w)
    dbar = Funicular.alloc_device(b, Float64, (Q.N, 1))
    try
        rightmul!(Z1, Q, S1; nbuffers=nbuffers)      # Q Sbar^H
        rightmul!(Z2, Q, S2; nbuffers=nbuffers)      # Q Sbar
        sweep!(readpanels(Q), readpanels(Z1); nbuffers=nbuffers) do j, Qb, Zb
            hip().rowdot!(b, dbar, Qb, Zb, j > 1)
        end
        panelmul!(Qbar, G, Z1; nbuffers=nbuffers)    # G (Q Sbar^H)
        panelmul!(Z1, G', Z2; nbuffers=nbuffers)     # G^H (Q Sbar), over Z1
        sweep!(readpanels(Z1), updatepanels(Qbar); nbuffers=nbuffers) do _, Wb, Qb
            axpy!(one(T), Wb, Qb)
        end
    finally
        free!(Z1)
        free!(Z2)
    end
    d = Vector{Float64}(undef, Q.N)
    Funicular.d2h!(reshape(d, :, 1), dbar, b)
    Qbar, d
end

# The operator's tangent: the sensitivity to the real diagonal parameters of
# an operator that carries them in a `diag` field (the HIP extension's
# TridiagonalStencil), none for an operator without any.
operator_tangent(G, d) =
    (hasproperty(G, :diag) && getproperty(G, :diag) !== nothing) ? Tangent{typeof(G)}(diag=d) : NoTangent()

function ChainRulesCore.rrule(::typeof(project), Q::PanelMatrix{T}, G; nbuffers=nothing) where {T}
    S = project(Q, G; nbuffers=nbuffers)
    function project_pullback_rule(Sbar)
        Qbar, d = project_pullback(Q, G, unthunk(Sbar); nbuffers=nbuffers)
        (NoTangent(), Qbar, operator_tangent(G, d))
    end
    S, project_pullback_rule
end

end # module FunicularHIPChainRulesExt
