# HydrogradHIP.jl -- drop-in MI355X path for Hydrograd.jl's SWE right-hand
# side and Euler step, through libhgswe.so (include/hgswe.h).
#
# Usage (inside a Hydrograd session):
#     using Hydrograd, HydrogradHIP
#     ode_f = HydrogradHIP.make_ode_f(swe2d_extra_params)   # replaces the closure of solve_swe_2D.jl:281-294
#     Q = HydrogradHIP.custom_ODE_update_cells(ode_f, Q, params_vector, t, dt, swe2d_extra_params)
#
# The functions keep the reference's names and signatures:
#   swe_2d_rhs(dQdt, Q, params_vector, t, p_extra)           semi_discretize_swe_2D.jl:18-19
#   custom_ODE_update_cells(ode_f, Q, params_vector, t, dt, p)  custom_ODE_solvers.jl:5
#   custom_ODE_solve(ode_f, Q0, params_vector, p)             custom_ODE_solvers.jl:36
# and read the reference's own structs (mesh_2D, BoundaryConditions2D,
# SWE2D_Extra_Parameters) unchanged.  ForwardDiff.Dual values cannot cross
# the C ABI: they are rejected with an error.  Reverse mode (Zygote and the
# SciMLSensitivity VJP choices that call it) is served by ChainRulesCore
# rules backed by the library's hand-written adjoint (hg_rhs_vjp, hg_adjoint),
# with the cotangent of params_vector formed for the active parameter exactly
# as Zygote would through semi_discretize_swe_2D.jl:114-199.
module HydrogradHIP

using ChainRulesCore

export swe_2d_rhs, custom_ODE_update_cells, fields_context, custom_ODE_solve, make_ode_f, tsit5_solve, set_ude!, ude_grad,
       sensitivity_jvp

const libhgswe = joinpath(@__DIR__, "..", "lib", "libhgswe.so")

const HG_BC = Dict("inlet-q" => Int32(0), "exit-h" => Int32(1), "wall" => Int32(2), "symm" => Int32(3))
const HG_PARAM_NONE, HG_PARAM_ZB, HG_PARAM_N_ZONE, HG_PARAM_N_CELL, HG_PARAM_Q, HG_PARAM_UDE = Int32.(0:5)

struct HgMesh
    n_cells::Int64
    face_off::Ptr{Int32}
    face_nbr::Ptr{Int32}
    face_nx::Ptr{Float64}
    face_ny::Ptr{Float64}
    face_len::Ptr{Float64}
    area::Ptr{Float64}
    centroid_xy::Ptr{Float64}
end

struct HgBcs
    n_ghosts::Int64
    ghost_slot::Ptr{Int32}
    ghost_kind::Ptr{Int32}
    ghost_bc::Ptr{Int32}
    n_inlet::Int32
    n_exit::Int32
    inlet_off::Ptr{Int32}
    inlet_ghost::Ptr{Int32}
end

struct HgConsts
    g::Float64
    k_n::Float64
    h_small::Float64
end

mutable struct Context
    ptr::Ptr{Cvoid}
    N::Int
    active::Int32
    lux_hidden::Int  # hidden layers of the UDE network when its weights are params_vector
    ude_sets::Int    # weight sets of the MLP closure (0: off)
end

# hg_lux_net: the reference's UDE network (create_NN_model, process_UDE.jl:2-65)
struct HgLuxNet
    choice::Int32
    n_hidden::Int32
    width::NTuple{4,Int32}
    act::NTuple{4,Int32}
    output_bounds::NTuple{2,Float64}
    h_bounds::NTuple{2,Float64}
    u_bounds::NTuple{2,Float64}
    ks_bounds::NTuple{2,Float64}
    ln_epsilon::Float64
end

check(ctx, rc) = rc == 0 || error("libhgswe: " * unsafe_string(ccall((:hg_last_error, libhgswe), Cstring, (Ptr{Cvoid},), ctx)))

# Flatten mesh_2D (mesh_2D.jl:2-71) + BoundaryConditions2D (bc_2D.jl:3-47) to
# the CSR arrays of hg_mesh / hg_bcs (0-based).
function flatten(m, bcs)
    N = m.numOfCells
    off = zeros(Int32, N + 1)
    for i in 1:N
        off[i+1] = off[i] + m.cellNodesCount[i]
    end
    M = off[end]
    nbr = zeros(Int32, M); nx = zeros(M); ny = zeros(M); len = zeros(M)
    slot_of_face = Dict{Int,Int}()
    for i in 1:N, k in 1:m.cellNodesCount[i]
        s = off[i] + k                              # 1-based slot
        f = m.cellFacesList[i, k]
        j = m.cellNeighbors_Dict[i][k]
        nbr[s] = m.bFace_is_boundary[f] ? ~Int32(j - 1) : Int32(j - 1)
        nx[s], ny[s] = m.cell_normals[i][k]
        len[s] = m.face_lengths[f]
        m.bFace_is_boundary[f] && (slot_of_face[f] = s)
    end
    G = m.numOfAllBounaryFaces
    gslot = zeros(Int32, G); gkind = zeros(Int32, G); gbc = fill(Int32(-1), G)
    function mark!(faceIDs, ghostIDs, kind, idx)
        for (f, g) in zip(faceIDs, ghostIDs)
            gslot[g] = slot_of_face[f] - 1; gkind[g] = kind; gbc[g] = idx
        end
    end
    for b in 1:bcs.nInletQ_BCs; mark!(bcs.inletQ_faceIDs[b], bcs.inletQ_ghostCellIDs[b], 0, b - 1); end
    for b in 1:bcs.nExitH_BCs;  mark!(bcs.exitH_faceIDs[b], bcs.exitH_ghostCellIDs[b], 1, b - 1); end
    for b in 1:bcs.nWall_BCs;   mark!(bcs.wall_faceIDs[b], bcs.wall_ghostCellIDs[b], 2, -1); end
    for b in 1:bcs.nSymm_BCs;   mark!(bcs.symm_faceIDs[b], bcs.symm_ghostCellIDs[b], 3, -1); end
    inlet_off = zeros(Int32, bcs.nInletQ_BCs + 1)
    inlet_ghost = Int32[]
    for b in 1:bcs.nInletQ_BCs                      # NodeString order
        append!(inlet_ghost, Int32.(bcs.inletQ_ghostCellIDs[b] .- 1))
        inlet_off[b+1] = length(inlet_ghost)
    end
    cen = vec(permutedims(m.cell_centroids[:, 1:2]))
    return (; N, off, nbr, nx, ny, len, area = Vector{Float64}(m.cell_areas), cen, G, gslot, gkind, gbc,
              nI = bcs.nInletQ_BCs, nE = bcs.nExitH_BCs, inlet_off, inlet_ghost = isempty(inlet_ghost) ? Int32[0] : inlet_ghost)
end

const CONTEXTS = IdDict{Any,Context}()

# One context per mesh object (objectid of p_extra.my_mesh_2D).
function context(p)
    get!(CONTEXTS, p.my_mesh_2D) do
        d = flatten(p.my_mesh_2D, p.boundary_conditions)
        c = p.swe_2D_constants
        out = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve d begin
            mesh = Ref(HgMesh(d.N, pointer(d.off), pointer(d.nbr), pointer(d.nx), pointer(d.ny), pointer(d.len),
                              pointer(d.area), pointer(d.cen)))
            bc = Ref(HgBcs(d.G, pointer(d.gslot), pointer(d.gkind), pointer(d.gbc), d.nI, d.nE,
                           pointer(d.inlet_off), pointer(d.inlet_ghost)))
            k = Ref(HgConsts(c.g, c.k_n, c.h_small))
            rc = ccall((:hg_create, libhgswe), Cint, (Ref{Ptr{Cvoid}}, Ref{HgMesh}, Ref{HgBcs}, Ref{HgConsts}, Int32),
                       out, mesh, bc, k, Int32(1))
            check(C_NULL, rc)
        end
        ctx = Context(out[], d.N, HG_PARAM_NONE, 0, 0)
        finalizer(x -> ccall((:hg_destroy, libhgswe), Cint, (Ptr{Cvoid},), x.ptr), ctx)
        S0 = p.S0_cells
        rc = ccall((:hg_set_fields, libhgswe), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                    Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   ctx.ptr, Vector{Float64}(p.hstill), Vector{Float64}(p.hstill_ghostCells), Vector{Float64}(p.zb_cells),
                   Vector{Float64}(p.zb_ghostCells), Vector{Float64}(S0[:, 1]), Vector{Float64}(S0[:, 2]),
                   Vector{Float64}(p.ManningN_cells), Vector{Float64}(p.inletQ_TotalQ), Vector{Float64}(p.exitH_WSE))
        check(ctx.ptr, rc)
        set_active!(ctx, p)
        set_n_law!(ctx, p)
        set_lux_ude!(ctx, p)
        ctx
    end
end

# K Manning-n fields on the mesh of p (hg_fields_create): the mesh is stored
# once and every call advances all K fields, with the arrays of the K fields
# laid end to end (field f after field f-1); ManningN is an N x K matrix.
# Returns a Context for the same calls as a single mesh's (hg_step,
# hg_grad_probes, ...), with K-fold arrays.
function fields_context(p, ManningN::AbstractMatrix)
    K = size(ManningN, 2)
    d = flatten(p.my_mesh_2D, p.boundary_conditions)
    c = p.swe_2D_constants
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve d begin
        mesh = Ref(HgMesh(d.N, pointer(d.off), pointer(d.nbr), pointer(d.nx), pointer(d.ny), pointer(d.len),
                          pointer(d.area), pointer(d.cen)))
        bc = Ref(HgBcs(d.G, pointer(d.gslot), pointer(d.gkind), pointer(d.gbc), d.nI, d.nE,
                       pointer(d.inlet_off), pointer(d.inlet_ghost)))
        k = Ref(HgConsts(c.g, c.k_n, c.h_small))
        rc = ccall((:hg_fields_create, libhgswe), Cint,
                   (Ref{Ptr{Cvoid}}, Int32, Ref{HgMesh}, Ref{HgBcs}, Ref{HgConsts}, Int32), out, Int32(K), mesh, bc, k,
                   Int32(1))
        check(C_NULL, rc)
    end
    ctx = Context(out[], K * d.N, HG_PARAM_NONE, 0, 0)
    finalizer(x -> ccall((:hg_destroy, libhgswe), Cint, (Ptr{Cvoid},), x.ptr), ctx)
    rep(v) = repeat(Vector{Float64}(v), K)
    S0 = p.S0_cells
    check(ctx.ptr, ccall((:hg_set_fields, libhgswe), Cint,
                         (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                          Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                         ctx.ptr, rep(p.hstill), rep(p.hstill_ghostCells), rep(p.zb_cells), rep(p.zb_ghostCells),
                         rep(S0[:, 1]), rep(S0[:, 2]), Vector{Float64}(vec(ManningN)), rep(p.inletQ_TotalQ),
                         rep(p.exitH_WSE)))
    return ctx
end

# The forward simulation's variable Manning n (ManningN_option "variable",
# semi_discretize_swe_2D.jl:129-135): the law of create_manning_function
# (process_ManningN_2D.jl:122-134) runs on the device before every RHS.
const HG_NLAW = Dict("power_law" => Int32(1), "sigmoid" => Int32(2), "inverse" => Int32(3), "h_Umag_ks" => Int32(4))
function set_n_law!(ctx, p)
    s = p.settings
    (s.bPerform_Forward_Simulation && s.forward_settings.ManningN_option == "variable") || return nothing
    fs = s.forward_settings
    kind = HG_NLAW[fs.ManningN_function_type]
    prm = fs.ManningN_function_parameters
    get1(k) = Float64(get(prm, k, 1.0))
    params = [Float64(get(prm, "n_lower", 0.0)), Float64(get(prm, "n_upper", 0.0)), get1("k"), get1("h_mid")]
    ks = kind == HG_NLAW["h_Umag_ks"] ? Vector{Float64}(p.ks_cells) : C_NULL
    check(ctx.ptr, ccall((:hg_set_n_law, libhgswe), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}),
                         ctx.ptr, kind, params, ks))
    return nothing
end

# active_param_name -> HG_PARAM_* (semi_discretize_swe_2D.jl:114-199)
function set_active!(ctx, p)
    s = p.settings
    inv = s.bPerform_Inversion || s.bPerform_Sensitivity_Analysis
    kind, zone, nz = HG_PARAM_NONE, Int32[0], Int32(0)
    if inv && p.active_param_name == "zb"
        kind = HG_PARAM_ZB
    elseif inv && p.active_param_name == "ManningN"
        kind = HG_PARAM_N_ZONE
        zone = Int32.(p.srh_all_Dict["matID_cells"])
        nz = Int32(maximum(zone) + 1)
    elseif inv && p.active_param_name == "Q"
        kind = HG_PARAM_Q
    end
    check(ctx.ptr, ccall((:hg_set_active_param, libhgswe), Cint, (Ptr{Cvoid}, Int32, Ptr{Int32}, Int32),
                         ctx.ptr, kind, zone, nz))
    ctx.active = kind
end

# The reference's UDE network on the device (UDE_choice ManningN_h,
# ManningN_h_Umag_ks or FlowResistance; semi_discretize_swe_2D.jl:145-158,
# 521-531): its weights are params_vector (HG_PARAM_UDE).
const HG_LUX = Dict("ManningN_h" => Int32(1), "ManningN_h_Umag_ks" => Int32(2), "FlowResistance" => Int32(3))
const HG_ACT = Dict("relu" => Int32(0), "leakyrelu" => Int32(1), "sigmoid" => Int32(2), "tanh" => Int32(3),
                    "softplus" => Int32(4))
function set_lux_ude!(ctx, p)
    s = p.settings
    (s.bPerform_UDE && haskey(HG_LUX, s.UDE_settings.UDE_choice)) || return nothing
    cfg = s.UDE_settings.UDE_NN_config
    hidden = Int32.(cfg["hidden_layers"])
    acts = [HG_ACT[a] for a in cfg["activations"]]
    pad(v, z) = ntuple(i -> i <= length(v) ? v[i] : z, 4)
    b2(k) = haskey(cfg, k) ? (Float64(cfg[k][1]), Float64(cfg[k][2])) : (0.0, 1.0)
    ub = haskey(cfg, "Umag_bounds") ? b2("Umag_bounds") : b2("velocity_magnitude_bounds")
    net = HgLuxNet(HG_LUX[s.UDE_settings.UDE_choice], Int32(length(hidden)), pad(hidden, Int32(0)),
                   pad(acts, Int32(0)), b2("output_bounds"), b2("h_bounds"), ub, b2("ks_bounds"), 0.0)
    w = lux_flat(p.ude_model_params, length(hidden))
    ks = s.UDE_settings.UDE_choice == "ManningN_h_Umag_ks" ? Vector{Float64}(p.ks_cells) : C_NULL
    check(ctx.ptr, ccall((:hg_set_lux_ude, libhgswe), Cint, (Ptr{Cvoid}, Ref{HgLuxNet}, Ptr{Float64}, Ptr{Float64}),
                         ctx.ptr, Ref(net), w, ks))
    check(ctx.ptr, ccall((:hg_set_active_param, libhgswe), Cint, (Ptr{Cvoid}, Int32, Ptr{Int32}, Int32),
                         ctx.ptr, HG_PARAM_UDE, C_NULL, Int32(0)))
    ctx.active = HG_PARAM_UDE
    ctx.lux_hidden = length(hidden)
    return nothing
end

# Lux parameters (layer_1 Dense, layer_2 LayerNorm, ..., last Dense) in the
# library's flat order, by field name.
function lux_flat(ps, nh)
    v = Float64[]
    for l in 1:nh
        d = getproperty(ps, Symbol(:layer_, 2l - 1))
        n = getproperty(ps, Symbol(:layer_, 2l))
        append!(v, vec(d.weight), vec(d.bias), vec(n.scale), vec(n.bias))
    end
    o = getproperty(ps, Symbol(:layer_, 2nh + 1))
    append!(v, vec(o.weight), vec(o.bias))
    return v
end

# The flat weight cotangent back into the structure of ps (a NamedTuple of
# arrays or a ComponentArray; every array is overwritten).
function lux_unflat(ps, g, nh)
    pb = deepcopy(ps)
    k = 0
    put!(a) = (n = length(a); a .= reshape(g[k+1:k+n], size(a)); k += n)
    for l in 1:nh
        d = getproperty(pb, Symbol(:layer_, 2l - 1))
        n = getproperty(pb, Symbol(:layer_, 2l))
        put!(d.weight); put!(d.bias); put!(n.scale); put!(n.bias)
    end
    o = getproperty(pb, Symbol(:layer_, 2nh + 1))
    put!(o.weight); put!(o.bias)
    return pb
end

flatparams(ctx, p) = ctx.active == HG_PARAM_UDE ? lux_flat(p, ctx.lux_hidden) : Vector{Float64}(p)

nodual(x) = eltype(x) <: Float64 || error("HydrogradHIP: only Float64 arrays cross the C ABI (got $(eltype(x)))")

function swe_2d_rhs(dQdt::AbstractVector, Q::AbstractVector, params_vector::AbstractVector, t::Float64, p_extra)
    nodual(Q); nodual(params_vector)
    ctx = context(p_extra)
    q = Vector{Float64}(Q)
    prm = flatparams(ctx, params_vector)
    out = p_extra.bInPlaceODE ? dQdt : similar(q)
    check(ctx.ptr, ccall((:hg_rhs, libhgswe), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}),
                         ctx.ptr, q, prm, length(prm), out))
    return out
end

# The ODEFunction closure of solve_swe_2D.jl:281-294, backed by the device.
struct DeviceRHS{P}
    p::P
end
(f::DeviceRHS)(u, p, t) = swe_2d_rhs(similar(u), u, p, t, f.p)
make_ode_f(p_extra) = DeviceRHS(p_extra)

function custom_ODE_update_cells(ode_f::DeviceRHS, Q, params_vector, t, dt, swe2d_extra_params)
    nodual(Q); nodual(params_vector)
    ctx = context(swe2d_extra_params)
    q = Vector{Float64}(Q)
    prm = flatparams(ctx, params_vector)
    check(ctx.ptr, ccall((:hg_step, libhgswe), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Int64),
                         ctx.ptr, q, prm, length(prm), dt, 1))
    return q
end

# Same saves as the reference (one column per step, custom_ODE_solvers.jl:61-86),
# the steps themselves run on the device.
function custom_ODE_solve(ode_f::DeviceRHS, Q0, params_vector, swe2d_extra_params)
    c = swe2d_extra_params.swe_2D_constants
    steps = c.tspan[1]:c.dt:c.tspan[2]
    ctx = context(swe2d_extra_params)
    prm = flatparams(ctx, params_vector)
    check(ctx.ptr, ccall((:hg_set_params, libhgswe), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64), ctx.ptr, prm, length(prm)))
    check(ctx.ptr, ccall((:hg_state_upload, libhgswe), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx.ptr, Vector{Float64}(Q0)))
    sol = Matrix{Float64}(undef, length(Q0), length(steps))
    for k in eachindex(steps)
        check(ctx.ptr, ccall((:hg_advance, libhgswe), Cint, (Ptr{Cvoid}, Float64, Int64), ctx.ptr, c.dt, 1))
        check(ctx.ptr, ccall((:hg_state_download, libhgswe), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx.ptr, view(sol, :, k)))
    end
    return sol
end

# Adaptive Tsit5 on the device, for solve(ODEProblem(ode_f, Q0, tspan, p),
# Tsit5(); saveat = range(tspan..., length = nsave + 1)) as
# swe_2D_forward_simulation.jl:24-53 runs it.  Returns a 3N x (nsave + 1)
# matrix (one column per saved time, like Array(sol)).
function tsit5_solve(ode_f::DeviceRHS, Q0, params_vector, tspan, nsave; abstol = 1e-6, reltol = 1e-3, dt0 = 1e-3)
    nodual(Q0); nodual(params_vector)
    ctx = context(ode_f.p)
    prm = flatparams(ctx, params_vector)
    sol = Matrix{Float64}(undef, length(Q0), nsave + 1)
    stats = zeros(Int64, 3)
    check(ctx.ptr, ccall((:hg_tsit5_solve, libhgswe), Cint,
                         (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Float64, Int64, Float64, Float64,
                          Float64, Ptr{Float64}, Ptr{Int64}),
                         ctx.ptr, Vector{Float64}(Q0), prm, length(prm), Float64(tspan[1]), Float64(tspan[2]), nsave,
                         abstol, reltol, dt0, sol, stats))
    return sol
end

# Forward-mode sensitivity on the device (the ForwardDiff.jacobian columns of
# swe_2D_sensitivity.jl:33-80, one tangent direction per call): returns the
# state after nsteps Euler steps and its derivative along (dparams, dQ0).
function sensitivity_jvp(ode_f::DeviceRHS, Q0, params_vector, dparams, dt, nsteps; dQ0 = nothing)
    nodual(Q0); nodual(params_vector)
    ctx = context(ode_f.p)
    prm = flatparams(ctx, params_vector)
    dp = dparams === nothing ? C_NULL : Vector{Float64}(dparams)
    d0 = dQ0 === nothing ? C_NULL : Vector{Float64}(dQ0)
    QN = zeros(length(Q0)); dQN = zeros(length(Q0))
    check(ctx.ptr, ccall((:hg_jvp, libhgswe), Cint,
                         (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}, Float64, Int64,
                          Ptr{Float64}, Ptr{Float64}),
                         ctx.ptr, Vector{Float64}(Q0), d0, prm, length(prm), dp, dt, nsteps, QN, dQN))
    return QN, dQN
end

# UDE closure for Manning's n on the device (3 -> 16 -> 16 -> 1 tanh network,
# 353 weights; the layout is documented in include/hgswe.h).  weights = nothing
# turns it off.
function set_ude!(p_extra, weights; precision = 32, set_of_cell = nothing)
    ctx = context(p_extra)
    if weights === nothing
        check(ctx.ptr, ccall((:hg_set_ude, libhgswe), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Int32}, Int32),
                             ctx.ptr, 0, C_NULL, C_NULL, 32))
        ctx.ude_sets = 0
    else
        # one set per column of a 353 x n_sets matrix; set_of_cell (0-based, per cell) picks each cell's set
        w = Vector{Float64}(vec(weights))
        s = set_of_cell === nothing ? C_NULL : Vector{Int32}(set_of_cell)
        check(ctx.ptr, ccall((:hg_set_ude, libhgswe), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Int32}, Int32),
                             ctx.ptr, Int32(length(w) ÷ 353), w, s, Int32(precision)))
        ctx.ude_sets = length(w) ÷ 353
    end
    return nothing
end

# Cotangents of the closure weights from the last reverse pass (hg_ude_grad):
# a 353-vector for one weight set, a 353 x n_sets matrix for several.
function ude_grad(p_extra)
    ctx = context(p_extra)
    ctx.ude_sets > 0 || error("libhgswe: the UDE closure is off")
    g = zeros(353, ctx.ude_sets)
    check(ctx.ptr, ccall((:hg_ude_grad, libhgswe), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx.ptr, g))
    return ctx.ude_sets == 1 ? vec(g) : g
end

# ---------------------------------------------------------------- reverse mode

# Cotangent of params_vector from the per-cell n, zb and per-inlet Q
# cotangents (semi_discretize_swe_2D.jl:114-199: zb -> bed data, zone n ->
# cell n gather, Q -> inlet totals).
function params_bar(ctx, p_extra, prm, nb, zbb, qb, p = nothing)
    k = ctx.active
    if k == HG_PARAM_UDE  # weight cotangent of the network (hg_lux_grad)
        g = zeros(length(prm))
        check(ctx.ptr, ccall((:hg_lux_grad, libhgswe), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx.ptr, g))
        return p === nothing ? g : lux_unflat(p, g, ctx.lux_hidden)
    end
    k == HG_PARAM_ZB && return zbb
    k == HG_PARAM_N_CELL && return nb
    k == HG_PARAM_Q && return qb
    if k == HG_PARAM_N_ZONE
        zone = p_extra.srh_all_Dict["matID_cells"]
        out = zeros(length(prm))
        for i in eachindex(nb)
            out[zone[i]+1] += nb[i]
        end
        return out
    end
    return zero(prm)
end

dense(x, n) = x isa AbstractZero ? zeros(n) : Vector{Float64}(unthunk(x))

# The per-cell n and zb cotangent buffers a pullback asks the library for:
# only the active parameter's (a NULL output skips that cotangent's
# accumulation in the reverse step; the inlet-Q cotangent is always formed).
function cotangent_buffers(ctx, N)
    k = ctx.active
    nb = (k == HG_PARAM_N_CELL || k == HG_PARAM_N_ZONE) ? zeros(N) : nothing
    zbb = k == HG_PARAM_ZB ? zeros(N) : nothing
    return nb, zbb
end
orNULL(x) = x === nothing ? Ptr{Float64}(C_NULL) : x

# Pullback of one RHS evaluation: (Q, params) -> dQdt.
function rhs_pullback(ctx, p_extra, q, prm, ybar, p = nothing)
    N = ctx.N
    nI = length(p_extra.inletQ_TotalQ)
    Qb = zeros(3N); qb = zeros(max(nI, 1))
    nb, zbb = cotangent_buffers(ctx, N)
    check(ctx.ptr, ccall((:hg_rhs_vjp, libhgswe), Cint,
                         (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                          Ptr{Float64}, Ptr{Float64}),
                         ctx.ptr, q, prm, length(prm), dense(ybar, 3N), Qb, orNULL(nb), orNULL(zbb), qb))
    return Qb, params_bar(ctx, p_extra, prm, nb, zbb, qb[1:nI], p)
end

function ChainRulesCore.rrule(f::DeviceRHS, u, p, t)
    y = f(u, p, t)
    ctx = context(f.p)
    q = Vector{Float64}(u); prm = flatparams(ctx, p)
    function DeviceRHS_pullback(ybar)
        Qb, pb = rhs_pullback(ctx, f.p, q, prm, ybar, p)
        return NoTangent(), Qb, pb, NoTangent()
    end
    return y, DeviceRHS_pullback
end

function ChainRulesCore.rrule(::typeof(swe_2d_rhs), dQdt, Q, params_vector, t, p_extra)
    y = swe_2d_rhs(dQdt, Q, params_vector, t, p_extra)
    ctx = context(p_extra)
    q = Vector{Float64}(Q); prm = flatparams(ctx, params_vector)
    function swe_2d_rhs_pullback(ybar)
        Qb, pb = rhs_pullback(ctx, p_extra, q, prm, ybar, params_vector)
        return NoTangent(), NoTangent(), Qb, pb, NoTangent(), NoTangent()
    end
    return y, swe_2d_rhs_pullback
end

# Pullback of one Euler step with its dry clamp (custom_ODE_solvers.jl:5-33):
# the clamp's mask is recorded in the forward pass of hg_adjoint.
function ChainRulesCore.rrule(::typeof(custom_ODE_update_cells), ode_f::DeviceRHS, Q, params_vector, t, dt,
                              swe2d_extra_params)
    y = custom_ODE_update_cells(ode_f, Q, params_vector, t, dt, swe2d_extra_params)
    ctx = context(swe2d_extra_params)
    q = Vector{Float64}(Q); prm = flatparams(ctx, params_vector)
    N = ctx.N
    nI = length(swe2d_extra_params.inletQ_TotalQ)
    function update_cells_pullback(ybar)
        Qb, nb, zbb, qb = step_adjoint(ctx, q, prm, dt, 1, dense(ybar, 3N), nI)
        pb = params_bar(ctx, swe2d_extra_params, prm, nb, zbb, qb, params_vector)
        return NoTangent(), NoTangent(), Qb, pb, NoTangent(), NoTangent(), NoTangent()
    end
    return y, update_cells_pullback
end

# hg_adjoint over nsteps Euler steps from q with final-state cotangent
# lamN (checkpoint interval: the library's choice): the cotangents of q, of
# the active parameter's per-cell field (or nothing) and of the inlet Qs.
function step_adjoint(ctx, q, prm, dt, nsteps, lamN, nI)
    N = ctx.N
    Qb = zeros(3N); qb = zeros(max(nI, 1))
    nb, zbb = cotangent_buffers(ctx, N)
    check(ctx.ptr, ccall((:hg_adjoint, libhgswe), Cint,
                         (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Int64, Int64, Ptr{Float64},
                          Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                         ctx.ptr, q, prm, length(prm), dt, nsteps, 0, lamN, Qb, orNULL(nb), orNULL(zbb), qb))
    return Qb, nb, zbb, qb[1:nI]
end

addto(a, b) = a === nothing ? b : (b === nothing ? a : a .+ b)

# Pullback of the whole fixed-dt solve (custom_ODE_solvers.jl:36-95; column k
# of the result is the state after k steps).  When the cotangent reaches the
# final state only (the inversion and UDE losses read the last column), one
# device-resident reverse sweep over all the steps; otherwise the steps are
# reversed one at a time, each column's cotangent joining the state's
# cotangent after its step, and the parameter cotangents summed.
function ChainRulesCore.rrule(::typeof(custom_ODE_solve), ode_f::DeviceRHS, Q0, params_vector, swe2d_extra_params)
    sol = custom_ODE_solve(ode_f, Q0, params_vector, swe2d_extra_params)
    ctx = context(swe2d_extra_params)
    dt = swe2d_extra_params.swe_2D_constants.dt
    q0 = Vector{Float64}(Q0); prm = flatparams(ctx, params_vector)
    nI = length(swe2d_extra_params.inletQ_TotalQ)
    K = size(sol, 2)
    function solve_pullback(sbar)
        S = sbar isa AbstractZero ? zeros(size(sol)) : Matrix{Float64}(unthunk(sbar))
        if K == 1 || all(iszero, view(S, :, 1:K-1))
            Qb, nb, zbb, qb = step_adjoint(ctx, q0, prm, dt, K, S[:, K], nI)
            pb = params_bar(ctx, swe2d_extra_params, prm, nb, zbb, qb, params_vector)
        else
            # the parameter cotangent is linear in each step's outputs (and a
            # network's weight cotangent is read right after its step), so the
            # flat per-step cotangents add up
            lam = S[:, K]
            pflat = nothing
            for k in K:-1:1
                qk = k == 1 ? q0 : sol[:, k-1]
                lam, nbk, zbbk, qbk = step_adjoint(ctx, qk, prm, dt, 1, lam, nI)
                pflat = addto(pflat, params_bar(ctx, swe2d_extra_params, prm, nbk, zbbk, qbk))
                k > 1 && (lam .+= view(S, :, k-1))
            end
            Qb = lam
            pb = ctx.active == HG_PARAM_UDE ? lux_unflat(params_vector, pflat, ctx.lux_hidden) : pflat
        end
        return NoTangent(), NoTangent(), Qb, pb, NoTangent()
    end
    return sol, solve_pullback
end

end # module
