"""
    MaviHIP

Julia side of the MI355X step hot path of Mavi.jl.  Adds the device mode
`GPU <: Mavi.Configs.DeviceMode` and methods on Mavi's own generic functions,
so that `System(...; int_cfg=IntCfg(dt=..., chunks_cfg=..., device=GPU()))`
runs `newton_step!` / `Rings.Integration.step!` through libmavi_hip.so
(`ccall`, C ABI in include/mavi_hip.h).  Built-in potentials (LenJonesCfg,
HarmTruncCfg, RingsCfg with HarmTruncCfg/InteractionMatrix) run on the GPU;
any other DynamicCfg (user forces, Szabo, RunTumble) and non-rectangular
geometries keep Mavi's Sequencial CPU methods.

Reference dispatch points used (Mavi.jl repository paths):
  calc_forces!(system, chunks, device)      src/integration.jl:112, 159, 197, 226
  update_chunks!(chunks)                     src/chunks.jl:150-161
  get_step_function(sys_type, system)        src/integration.jl:537-548, src/rings/integration.jl:545
  run_system(system; tf, num_steps)          src/run_system.jl:7-23
  kinetic_energy / potential_energy          src/quantities.jl:12-18, 46-66
"""
module MaviHIP

using Mavi
using Mavi.Configs
using Mavi.Systems
using Mavi.ChunksMod
using Mavi.Integration
using StaticArrays

export GPU

const LIB = joinpath(@__DIR__, "..", "mavi_hip", "libmavi_hip.so")

struct GPU <: Mavi.Configs.DeviceMode
    device::Int
end
GPU() = GPU(0)

# --- C ABI ---------------------------------------------------------------

const MAVI_OK = Cint(0)
const MAVI_WALL_RIGID = Cint(0)
const MAVI_WALL_PERIODIC = Cint(1)
const MAVI_POT_LJ = Cint(0)
const MAVI_POT_HARMTRUNC = Cint(1)
const MAVI_POT_RINGS = Cint(2)
const MAVI_POT_SZABO = Cint(3)
const MAVI_POT_RTP = Cint(4)

struct SysDesc                      # mavi_sys_desc
    n::Int64
    bottom_left::NTuple{2,Float64}
    size::NTuple{2,Float64}
    wall::Int32
    potential::Int32
    pot::NTuple{4,Float64}
    dt::Float64
    cols::Int32
    rows::Int32
    nc::Int32
    reserved0::Int32
    radius::Float64
    nmax::Int32
    nrings::Int32
    ntypes::Int32
    reserved1::Int32
    ring_types::Ptr{Int32}
    type_np::Ptr{Int32}
    type_par::Ptr{Float64}
    interaction::Ptr{Float64}
    seed::UInt64
    active::NTuple{4,Float64}    # Szabo: vo, mobility, relax_time, rot_diff; RTP: vo, tumble_rate
    nwalls::Int32
    reserved2::Int32
    walls::Ptr{Cvoid}            # Ptr{MaviWall}
end

struct MaviWall                     # mavi_wall
    kind::Int32
    mode::Int32
    pot_kind::Int32
    nlines::Int32
    geom::NTuple{4,Float64}
    pot::NTuple{4,Float64}
    lines::Ptr{Float64}
end

const WMODE = Dict(Configs.Outside => Int32(0), Configs.Inside => Int32(1), Configs.Repulsion => Int32(2))
wall_pot(p::LenJonesCfg) = (MAVI_POT_LJ, (p.sigma, p.epsilon, 0.0, 0.0))
wall_pot(p::HarmTruncCfg) = (MAVI_POT_HARMTRUNC, (p.k_rep, p.k_atr, p.dist_eq, p.dist_max))

# One mavi_wall per (wall, geometry) pair; `keep` holds the line tables alive.
function wall_record(w, g::RectangleCfg, keep)
    kind = w isa PeriodicWalls ? Int32(1) : Int32(0)
    MaviWall(kind, 0, 0, 0, (g.bottom_left[1], g.bottom_left[2], g.length, g.height), (0.0, 0.0, 0.0, 0.0), C_NULL)
end
function wall_record(w, g::Configs.CircleCfg, keep)
    kind = w isa RigidWalls ? Int32(2) : w isa Configs.SlipperyWalls ? Int32(4) : Int32(5)
    mode, pk, pot = w isa Configs.PotentialWalls ? (WMODE[typeof(w.mode)], wall_pot(w.potential)...) :
                    (Int32(0), Cint(0), (0.0, 0.0, 0.0, 0.0))
    MaviWall(kind, mode, pk, 0, (g.center[1], g.center[2], g.radius, 0.0), pot, C_NULL)
end
function wall_record(w, g::Configs.LinesCfg, keep)
    kind = w isa Configs.SlipperyWalls ? Int32(3) : Int32(6)
    tab = Float64[]
    for l in g.lines
        append!(tab, (l.p1[1], l.p1[2], l.p2[1], l.p2[2]))
    end
    push!(keep, tab)
    mode, pk, pot = w isa Configs.PotentialWalls ? (WMODE[typeof(w.mode)], wall_pot(w.potential)...) :
                    (Int32(0), Cint(0), (0.0, 0.0, 0.0, 0.0))
    MaviWall(kind, mode, pk, length(g.lines), (0.0, 0.0, 0.0, 0.0), pot, pointer(tab))
end
wall_records(s::SpaceCfg{<:Configs.ManyWalls}, keep) =
    [wall_record(w, g, keep) for (w, g) in zip(s.wall_type.list, s.geometry_cfg.list)]
wall_records(s::SpaceCfg{<:Union{RigidWalls,PeriodicWalls},<:RectangleCfg}, keep) = MaviWall[]
wall_records(s::SpaceCfg, keep) = [wall_record(s.wall_type, s.geometry_cfg, keep)]

mutable struct DeviceSys
    ctx::Ptr{Cvoid}
    sys::Ptr{Cvoid}
    host_dirty::Bool
end

function check(ctx, st)
    st == MAVI_OK && return nothing
    msg = unsafe_string(ccall((:mavi_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx))
    error("libmavi_hip status $st: $msg")
end

# One context per device, one device system per Mavi System (side table keyed
# by objectid: System is immutable, its arrays are not).
const CONTEXTS = Dict{Int,Ptr{Cvoid}}()
const SYSTEMS = Dict{UInt,DeviceSys}()

function context(dev::Int)
    get!(CONTEXTS, dev) do
        out = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:mavi_ctx_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), dev, out)
        st == MAVI_OK || error("mavi_ctx_create failed ($st)")
        out[]
    end
end

# --- which configurations take the GPU path --------------------------------

gpu_kind(::Any) = nothing
gpu_kind(::LenJonesCfg) = MAVI_POT_LJ
gpu_kind(::HarmTruncCfg) = MAVI_POT_HARMTRUNC
gpu_kind(::SzaboCfg) = MAVI_POT_SZABO
gpu_kind(::RunTumbleCfg) = MAVI_POT_RTP
gpu_kind(::Mavi.Rings.Configs.RingsCfg{U,T,F}) where {U,T<:Mavi.Rings.Configs.HarmTruncCfg,F} = MAVI_POT_RINGS

wall_kind(::SpaceCfg{RigidWalls,<:RectangleCfg}) = MAVI_WALL_RIGID
wall_kind(::SpaceCfg{PeriodicWalls,<:RectangleCfg}) = MAVI_WALL_PERIODIC
wall_kind(s::SpaceCfg{<:Configs.ManyWalls}) = Configs.get_main_wall(s) isa PeriodicWalls ? MAVI_WALL_PERIODIC :
                                               MAVI_WALL_RIGID
wall_kind(::SpaceCfg{<:Union{RigidWalls,Configs.SlipperyWalls,Configs.PotentialWalls}}) = MAVI_WALL_RIGID
wall_kind(::Any) = nothing

on_gpu(system) = system.int_cfg.device isa GPU &&
                 gpu_kind(system.dynamic_cfg) !== nothing && wall_kind(system.space_cfg) !== nothing

function device_system(system)
    key = objectid(system)
    haskey(SYSTEMS, key) && return SYSTEMS[key]
    ctx = context(system.int_cfg.device.device)
    g = Configs.get_bounding_box(system.space_cfg.geometry_cfg)  # the chunk space
    dyn = system.dynamic_cfg
    chunks = system.chunks
    cols = isnothing(chunks) ? 0 : chunks.num_cols
    rows = isnothing(chunks) ? 0 : chunks.num_rows
    nc = isnothing(chunks) ? 0 : size(chunks.chunk_particles, 1)
    pot = dyn isa LenJonesCfg ? (dyn.sigma, dyn.epsilon, 0.0, 0.0) :
          dyn isa HarmTruncCfg ? (dyn.k_rep, dyn.k_atr, dyn.dist_eq, dyn.dist_max) :
          dyn isa SzaboCfg ? (dyn.k_rep, dyn.k_adh, dyn.r_eq, dyn.r_max) :
          dyn isa RunTumbleCfg ? (dyn.sigma, dyn.epsilon, 0.0, 0.0) : (0.0, 0.0, 0.0, 0.0)
    active = dyn isa SzaboCfg ? (dyn.vo, dyn.mobility, dyn.relax_time, dyn.rot_diff) :
             dyn isa RunTumbleCfg ? (dyn.vo, dyn.tumble_rate, 0.0, 0.0) : (0.0, 0.0, 0.0, 0.0)
    desc = SysDesc(length(system.state.pos), Tuple(g.bottom_left), (g.length, g.height),
                   wall_kind(system.space_cfg), gpu_kind(dyn), pot, system.int_cfg.dt,
                   cols, rows, nc, 0, minimum(particle_radius(dyn)),
                   0, 0, 0, 0, C_NULL, C_NULL, C_NULL, C_NULL, UInt64(0), active, 0, 0, C_NULL)
    keep = Any[]
    walls = wall_records(system.space_cfg, keep)
    if !isempty(walls)
        desc = SysDesc(ntuple(i -> getfield(desc, i), fieldcount(SysDesc) - 3)..., Int32(length(walls)), 0,
                       Ptr{Cvoid}(pointer(walls)))
    end
    # (RingsSystem fills nmax/nrings/ntypes and the type tables the same way,
    #  from RingsCfg fields p0, relax_time, vo, mobility, rot_diff, k_area,
    #  k_spring, l_spring, num_particles and interaction_finder.)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve walls keep check(ctx, ccall((:mavi_sys_create, LIB), Cint,
                                             (Ptr{Cvoid}, Ref{SysDesc}, Ref{Ptr{Cvoid}}), ctx, desc, out))
    d = DeviceSys(ctx, out[], true)
    SYSTEMS[key] = d
    finalizer(system.state) do _
        ccall((:mavi_sys_destroy, LIB), Cint, (Ptr{Cvoid},), d.sys)
        delete!(SYSTEMS, key)
    end
    return d
end

# SelfPropelledState.pol_angle travels through the `pol` argument
function upload!(system, d::DeviceSys)
    pos = system.state.pos
    vel = hasproperty(system.state, :vel) ? system.state.vel : nothing
    pol = hasproperty(system.state, :pol_angle) ? system.state.pol_angle : nothing
    GC.@preserve pos vel pol begin
        check(d.ctx, ccall((:mavi_sys_upload, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                           d.sys, pointer(pos), isnothing(vel) ? C_NULL : pointer(vel),
                           isnothing(pol) ? C_NULL : pointer(pol)))
    end
    d.host_dirty = false
end

function download!(system, d::DeviceSys)
    pos = system.state.pos
    vel = hasproperty(system.state, :vel) ? system.state.vel : nothing
    pol = hasproperty(system.state, :pol_angle) ? system.state.pol_angle : nothing
    forces = get_forces(system)
    GC.@preserve pos vel pol forces begin
        check(d.ctx, ccall((:mavi_sys_download, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                           d.sys, pointer(pos), isnothing(vel) ? C_NULL : pointer(vel),
                           isnothing(pol) ? C_NULL : pointer(pol)))
        check(d.ctx, ccall((:mavi_get_forces, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), d.sys, pointer(forces)))
    end
end

# --- methods on Mavi's generic functions ------------------------------------

function Integration.calc_forces!(system::System, chunks, device::GPU)
    on_gpu(system) || return Integration.calc_forces!(system, chunks, Sequencial())
    d = device_system(system)
    upload!(system, d)
    forces = get_forces(system)
    GC.@preserve forces check(d.ctx, ccall((:mavi_forces, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), d.sys,
                                           pointer(forces)))
    return nothing
end

"""newton_step! (or szabo_step! / rtp_step!, by the system's DynamicCfg) on
the GPU: upload, one step, download (per-step sync: user code may read the
state).  Self-propelled steps draw their noise from the device stream."""
function gpu_newton_step!(system)
    d = device_system(system)
    upload!(system, d)
    check(d.ctx, ccall((:mavi_step, LIB), Cint, (Ptr{Cvoid}, Int64), d.sys, 1))
    download!(system, d)
    system.time_info.time += system.int_cfg.dt
    system.time_info.num_steps += 1
end

function Integration.get_step_function(t::StandardSys, system::System)
    on_gpu(system) && return gpu_newton_step!
    return Integration.get_step_function(t, system, Val(:cpu))
end
# the reference's selection, src/integration.jl:538-548
function Integration.get_step_function(::StandardSys, system, ::Val{:cpu})
    cfg = system.dynamic_cfg
    cfg isa SzaboCfg && return Integration.szabo_step!
    cfg isa RunTumbleCfg && return Integration.rtp_step!
    return Integration.newton_step!
end

# --- rings: invasions (InvasionsInfo, src/rings/rings.jl:86-100) -------------

"""Hand a RingsSystem's r_chunks and InvasionsCfg to the device
(src/rings/rings.jl:188-202, 276-284): later device steps check for
invasions on the reference's schedule."""
function invasions_cfg!(system, d::DeviceSys)
    extra = system.int_cfg.extra
    rc = system.info.r_chunks
    sched = !isnothing(extra.invasions_cfg)
    (sched || !isnothing(rc)) || return nothing
    cols = isnothing(rc) ? 0 : rc.num_cols
    rows = isnothing(rc) ? 0 : rc.num_rows
    nc = isnothing(rc) ? 0 : size(rc.chunk_particles, 1)
    every = sched ? extra.invasions_cfg.steps_to_update : 0
    check(d.ctx, ccall((:mavi_rings_invasions_cfg, LIB), Cint, (Ptr{Cvoid}, Cint, Int64, Int32, Int32, Int32),
                       d.sys, sched ? 1 : 0, every, cols, rows, nc))
end

"""system.info.invasions.list <- the device's last check (the same set as
check_invasions!, ordered by invasor, invaded, p_id)."""
function invasions_download!(system, d::DeviceSys)
    count = Ref{Int64}(0)
    check(d.ctx, ccall((:mavi_rings_invasions_get, LIB), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Int64}, Int64),
                       d.sys, count, C_NULL, 0))
    rec = Vector{Int64}(undef, 3 * count[])
    GC.@preserve rec check(d.ctx, ccall((:mavi_rings_invasions_get, LIB), Cint,
                                        (Ptr{Cvoid}, Ref{Int64}, Ptr{Int64}, Int64), d.sys, count, pointer(rec),
                                        count[]))
    list = system.info.invasions.list
    empty!(list)
    for k in 0:count[]-1
        push!(list, Mavi.Rings.Invasion(invasor=rec[3k+1], invaded=rec[3k+2], p_id=rec[3k+3]))
    end
    return list
end

# --- rings: variable population, sources and sinks (src/rings/sources.jl) ----

struct MaviRingSource
    kind::Int32
    geom::Int32
    g::NTuple{4,Float64}
    nbox::Int32
    random_pol::Int32
    pad::Float64
    pol::Float64
    bbox::Ptr{Float64}
    spawn::Ptr{Float64}
    chunk_off::Ptr{Int32}
    chunk_ids::Ptr{Int32}
end

const MAVI_SRC_SOURCE = Int32(0)
const MAVI_SRC_SINK = Int32(1)
const MAVI_GEOM_RECT = Int32(0)
const MAVI_GEOM_CIRCLE = Int32(1)

sink_geom(g::RectangleCfg) = (MAVI_GEOM_RECT, (g.bottom_left[1], g.bottom_left[2], g.length, g.height))
sink_geom(g::Configs.CircleCfg) = (MAVI_GEOM_CIRCLE, (g.center[1], g.center[2], g.radius, 0.0))

"""Records of system.info.sources (Source / Sink objects the reference built,
src/rings/rings.jl:154-186); the arrays they point into go to keep."""
function source_records(system, keep)
    recs = MaviRingSource[]
    isnothing(system.info.sources) && return recs
    cols = isnothing(system.chunks) ? 0 : system.chunks.num_cols
    for s in system.info.sources
        if s isa Mavi.Rings.Sources.Sink
            geom, g = sink_geom(s.cfg.geometry_cfg)
            push!(recs, MaviRingSource(MAVI_SRC_SINK, geom, g, 0, 0, 0.0, 0.0, C_NULL, C_NULL, C_NULL, C_NULL))
            continue
        bbox = Float64[]
        for b in s.bbox
            append!(bbox, (b.bottom_left[1], b.bottom_left[2], b.length, b.height))
        end
        spawn = Float64[]
        for pts in s.bbox_spawn_pos, p in pts
            append!(spawn, (p[1], p[2]))
        end
        off, ids = C_NULL, C_NULL
        if s.checker isa Mavi.Rings.Sources.ChunksChecker
            o = Int32[0]
            l = Int32[]
            for lst in s.checker.ids
                for c in lst
                    push!(l, Int32((c[1] - 1) * cols + (c[2] - 1)))
                end
                push!(o, length(l))
            end
            push!(keep, o, l)
            off, ids = pointer(o), pointer(l)
        end
        push!(keep, bbox, spawn)
        rnd = !(s.cfg.spawn_pol isa Number)
        push!(recs, MaviRingSource(MAVI_SRC_SOURCE, 0, (0.0, 0.0, 0.0, 0.0), length(s.bbox), rnd ? 1 : 0,
                                   s.cfg.pad, rnd ? 0.0 : s.cfg.spawn_pol, pointer(bbox), pointer(spawn), off, ids))
    end
    return recs
end

"""Hand a RingsSystem's VarRingsIds mask and sources to the device; random
spawn polarisations come from `pol_draws!` (drawn from system.rng)."""
function rings_var_cfg!(system, d::DeviceSys)
    ids = system.state.rings_ids
    ids isa Mavi.Rings.States.VarRingsIds || return nothing
    keep = Any[]
    recs = source_records(system, keep)
    mask = Int32.(ids.mask)
    GC.@preserve keep mask recs check(d.ctx, ccall((:mavi_rings_var_cfg, LIB), Cint,
                                                   (Ptr{Cvoid}, Ptr{Int32}, Int32, Ptr{MaviRingSource}), d.sys,
                                                   mask, length(recs), recs))
end

"Append n uniform draws of system.rng to the device's spawn polarisation pool (the library draws from its own stream once the pool is empty)."
function pol_draws!(system, d::DeviceSys, n)
    u = rand(system.rng, Float64, n)
    check(d.ctx, ccall((:mavi_rings_pol_draws, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64), d.sys, u, n))
end

"run_system with a resident fast path: upload once, all steps on the device, download once."
function run_gpu!(system; tf=nothing, num_steps=nothing)
    n = if isnothing(tf)
        num_steps
    else
        t, k = system.time_info.time, 0
        while t < tf
            t += system.int_cfg.dt
            k += 1
        end
        k
    end
    d = device_system(system)
    upload!(system, d)
    check(d.ctx, ccall((:mavi_step, LIB), Cint, (Ptr{Cvoid}, Int64), d.sys, n))
    download!(system, d)
    for _ in 1:n
        system.time_info.time += system.int_cfg.dt
    end
    system.time_info.num_steps += n
end

end # module
