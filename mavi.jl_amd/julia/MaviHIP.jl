"""
    MaviHIP

Julia side of the MI355X step hot path of Mavi.jl.  Adds the device mode
`GPU <: Mavi.Configs.DeviceMode` and methods on Mavi's own generic functions,
so that `System(...; int_cfg=IntCfg(dt=..., chunks_cfg=..., device=GPU()))`
or `RingsSystem(...; int_cfg=RingsIntCfg(dt=..., p_chunks_cfg=..., device=GPU()))`
runs `newton_step!`, `szabo_step!`, `rtp_step!` and `Rings.Integration.step!`
through libmavi_hip.so (`ccall`, C ABI in include/mavi_hip.h).  Built-in
dynamics (LenJonesCfg, HarmTruncCfg, SzaboCfg, RunTumbleCfg, RingsCfg with
HarmTruncCfg/InteractionMatrix) run on the GPU; any other DynamicCfg (user
forces) and spaces the device does not take keep Mavi's CPU methods.

Reference dispatch points used (Mavi.jl repository paths):
  calc_forces!(system, chunks, device)      src/integration.jl:112, 159, 197, 226
  get_step_function(sys_type, system)        src/integration.jl:537-548, src/rings/integration.jl:545
  run_system(system; tf, num_steps)          src/run_system.jl:7-23
  run_experiment_batch(batch, get_system)    src/experiments.jl:246-300
"""
module MaviHIP

using Mavi
using Mavi.Configs
using Mavi.Systems
using Mavi.ChunksMod
using Mavi.Integration
using Serialization
using StaticArrays
import Mavi.Rings
import Mavi.Experiments

export GPU

const LIB = joinpath(@__DIR__, "..", "mavi_hip", "libmavi_hip.so")

"""
    GPU(; device=0, seed=0, lj_rcut=0.0)

Device mode of the MI355X path.  `seed` keys the device noise stream
(Philox4x32-10 over (seed, step count)) of resident runs and batches;
`lj_rcut > 0` cuts LenJonesCfg at that distance (an extension: the
reference's LenJonesCfg has no cutoff, src/configs.jl:380-397).
"""
struct GPU <: Mavi.Configs.DeviceMode
    device::Int
    seed::UInt64
    lj_rcut::Float64
end
GPU(; device=0, seed=0, lj_rcut=0.0) = GPU(device, UInt64(seed), Float64(lj_rcut))

# ParticleNeighbors of a GPU system keep one buffer slice (neighbors.jl:17-18).
Rings.NeighborsMod.get_slide_idx(::GPU) = 1

# --- C ABI ---------------------------------------------------------------

const MAVI_OK = Cint(0)
const MAVI_E_UNSUPPORTED = Cint(6)
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

struct MaviRingSource               # mavi_ring_source
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

function check(ctx, st)
    st == MAVI_OK && return nothing
    msg = unsafe_string(ccall((:mavi_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx))
    error("libmavi_hip status $st: $msg")
end

# One context per device, one device system per Mavi System (side table keyed
# by objectid: System is immutable, its arrays are not).
const CONTEXTS = Dict{Int,Ptr{Cvoid}}()
const SYSTEMS = Dict{UInt,Any}()

function context(dev::Int)
    get!(CONTEXTS, dev) do
        out = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:mavi_ctx_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), dev, out)
        st == MAVI_OK || error("mavi_ctx_create failed ($st)")
        out[]
    end
end

# --- walls --------------------------------------------------------------------

const WMODE = Dict(Configs.Outside => Int32(0), Configs.Inside => Int32(1), Configs.Repulsion => Int32(2))
wall_pot(p::LenJonesCfg) = (MAVI_POT_LJ, (Float64(p.sigma), Float64(p.epsilon), 0.0, 0.0))
wall_pot(p::HarmTruncCfg) = (MAVI_POT_HARMTRUNC, (Float64(p.k_rep), Float64(p.k_atr), Float64(p.dist_eq),
                                                  Float64(p.dist_max)))

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

# --- which configurations take the GPU path --------------------------------

gpu_kind(::Any) = nothing
gpu_kind(::LenJonesCfg) = MAVI_POT_LJ
gpu_kind(::HarmTruncCfg) = MAVI_POT_HARMTRUNC
gpu_kind(::SzaboCfg) = MAVI_POT_SZABO
gpu_kind(::RunTumbleCfg) = MAVI_POT_RTP
gpu_kind(::Rings.Configs.RingsCfg{U,T,F}) where {U,T<:Rings.Configs.HarmTruncCfg,F} = MAVI_POT_RINGS

wall_kind(::SpaceCfg{RigidWalls,<:RectangleCfg}) = MAVI_WALL_RIGID
wall_kind(::SpaceCfg{PeriodicWalls,<:RectangleCfg}) = MAVI_WALL_PERIODIC
wall_kind(s::SpaceCfg{<:Configs.ManyWalls}) = Configs.get_main_wall(s) isa PeriodicWalls ? MAVI_WALL_PERIODIC :
                                               MAVI_WALL_RIGID
wall_kind(::SpaceCfg{<:Union{RigidWalls,Configs.SlipperyWalls,Configs.PotentialWalls}}) = MAVI_WALL_RIGID
wall_kind(::Any) = nothing

is_rings(system) = system.type isa Rings.RingsSys

# Rings run on the device with a periodic main wall (their walls! wraps
# positions; rigid walls would need state.vel) and without ring neighbours
# (r_neigh), which the device does not compute.
function on_gpu(system)
    system.int_cfg.device isa GPU || return false
    gpu_kind(system.dynamic_cfg) === nothing && return false
    wall_kind(system.space_cfg) === nothing && return false
    if is_rings(system)
        Configs.get_main_wall(system.space_cfg) isa PeriodicWalls || return false
        isnothing(system.info.r_neigh) || return false
    end
    return true
end

# --- descriptor -------------------------------------------------------------

# RingsCfg fields in the order of mavi_sys_desc.type_par (one row per type)
const RING_FIELDS = (:p0, :relax_time, :vo, :mobility, :rot_diff, :k_area, :k_spring, :l_spring)

per_type(v::Number, t) = Float64(v)
per_type(v::AbstractVector, t) = Float64(v[t])

ht_params(c) = (Float64(c.k_rep), Float64(c.k_atr), Float64(c.dist_eq), Float64(c.dist_max))
inter_cfg(f::Rings.Configs.InteractionMatrix, t1, t2) = f.matrix[t1, t2]
inter_cfg(f, t1, t2) = f

"""Per-type tables of a RingsCfg / RingsState: particles per ring, the
RING_FIELDS rows, the ntypes x ntypes interaction matrix (column-major, four
HarmTrunc parameters per entry) and the ring types; all pushed to `keep`."""
function rings_tables(system, keep)
    dyn, st = system.dynamic_cfg, system.state
    nt = dyn.num_types
    npv = st.num_particles isa Integer ? Int32[st.num_particles] : Int32.(st.num_particles)
    par = Float64[]
    for t in 1:nt, f in RING_FIELDS
        push!(par, per_type(getfield(dyn, f), t))
    end
    inter = Float64[]
    for t2 in 1:nt, t1 in 1:nt
        append!(inter, ht_params(inter_cfg(dyn.interaction_finder, t1, t2)))
    end
    types = isnothing(st.types) ? Int32[] : Int32.(st.types)
    push!(keep, npv, par, inter, types)
    return npv, par, inter, types
end

# minimum(particle_radius(dynamic_cfg)) of get_chunks (src/systems.jl:14-28)
min_radius(dyn::Rings.Configs.RingsCfg) =
    minimum(c.dist_eq / 2 for c in Rings.Configs.list_self_interactions(dyn.interaction_finder))
min_radius(dyn) = minimum(particle_radius(dyn))

"""mavi_sys_desc of a System or RingsSystem; the host tables it points into
are pushed to `keep`, which must outlive the call that reads it."""
function describe(system, keep; seed=system.int_cfg.device.seed)
    g = Configs.get_bounding_box(system.space_cfg.geometry_cfg)  # the chunk space
    dyn = system.dynamic_cfg
    chunks = system.chunks
    cols = isnothing(chunks) ? 0 : chunks.num_cols
    rows = isnothing(chunks) ? 0 : chunks.num_rows
    nc = isnothing(chunks) ? 0 : size(chunks.chunk_particles, 1)
    pot = dyn isa LenJonesCfg ? (Float64(dyn.sigma), Float64(dyn.epsilon), system.int_cfg.device.lj_rcut, 0.0) :
          dyn isa HarmTruncCfg ? ht_params(dyn) :
          dyn isa SzaboCfg ? (Float64(dyn.k_rep), Float64(dyn.k_adh), Float64(dyn.r_eq), Float64(dyn.r_max)) :
          dyn isa RunTumbleCfg ? (Float64(dyn.sigma), Float64(dyn.epsilon), 0.0, 0.0) : (0.0, 0.0, 0.0, 0.0)
    active = dyn isa SzaboCfg ? (Float64(dyn.vo), Float64(dyn.mobility), Float64(dyn.relax_time),
                                 Float64(dyn.rot_diff)) :
             dyn isa RunTumbleCfg ? (Float64(dyn.vo), Float64(dyn.tumble_rate), 0.0, 0.0) : (0.0, 0.0, 0.0, 0.0)
    nmax = nrings = ntypes = Int32(0)
    p_types = p_np = Ptr{Int32}(C_NULL)
    p_par = p_inter = Ptr{Float64}(C_NULL)
    if is_rings(system)
        npv, par, inter, types = rings_tables(system, keep)
        nmax = Int32(size(system.state.rings_pos, 1))
        nrings = Int32(size(system.state.rings_pos, 2))
        ntypes = Int32(dyn.num_types)
        p_np, p_par, p_inter = pointer(npv), pointer(par), pointer(inter)
        p_types = isnothing(system.state.types) ? Ptr{Int32}(C_NULL) : pointer(types)
    end
    walls = wall_records(system.space_cfg, keep)
    push!(keep, walls)
    return SysDesc(length(system.state.pos), (Float64(g.bottom_left[1]), Float64(g.bottom_left[2])),
                   (Float64(g.length), Float64(g.height)), wall_kind(system.space_cfg), gpu_kind(dyn), pot,
                   Float64(system.int_cfg.dt), cols, rows, nc, 0, Float64(min_radius(dyn)),
                   nmax, nrings, ntypes, 0, p_types, p_np, p_par, p_inter, UInt64(seed), active,
                   Int32(length(walls)), 0, isempty(walls) ? C_NULL : Ptr{Cvoid}(pointer(walls)))
end

mutable struct DeviceSys
    ctx::Ptr{Cvoid}
    sys::Ptr{Cvoid}
    variable::Bool     # VarRingsIds population
    neighbors::Bool    # ParticleNeighbors on the device
    invasions::Bool    # r_chunks / InvasionsCfg on the device
end

function device_system(system)
    key = objectid(system)
    haskey(SYSTEMS, key) && return SYSTEMS[key]::DeviceSys
    ctx = context(system.int_cfg.device.device)
    keep = Any[]
    desc = describe(system, keep)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve keep check(ctx, ccall((:mavi_sys_create, LIB), Cint, (Ptr{Cvoid}, Ref{SysDesc}, Ref{Ptr{Cvoid}}),
                                       ctx, desc, out))
    d = DeviceSys(ctx, out[], false, false, false)
    SYSTEMS[key] = d
    finalizer(system.state) do _
        ccall((:mavi_sys_destroy, LIB), Cint, (Ptr{Cvoid},), d.sys)
        delete!(SYSTEMS, key)
    end
    upload!(system, d)
    if is_rings(system)
        rings_var_cfg!(system, d)
        invasions_cfg!(system, d)
        neighbors_cfg!(system, d)
    end
    return d
end

# pos is state.pos (for rings the reshape of rings_pos); the third array is
# SelfPropelledState.pol_angle or RingsState.pol
host_vel(state) = hasproperty(state, :vel) ? state.vel : nothing
host_pol(state) = hasproperty(state, :pol_angle) ? state.pol_angle : hasproperty(state, :pol) ? state.pol : nothing
as_ptr(a) = isnothing(a) ? Ptr{Float64}(C_NULL) : Ptr{Float64}(pointer(a))

function upload!(system, d::DeviceSys)
    pos, vel, pol = system.state.pos, host_vel(system.state), host_pol(system.state)
    GC.@preserve pos vel pol begin
        check(d.ctx, ccall((:mavi_sys_upload, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                           d.sys, as_ptr(pos), as_ptr(vel), as_ptr(pol)))
    end
end

function download!(system, d::DeviceSys)
    pos, vel, pol = system.state.pos, host_vel(system.state), host_pol(system.state)
    forces = get_forces(system)
    GC.@preserve pos vel pol forces begin
        check(d.ctx, ccall((:mavi_sys_download, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                           d.sys, as_ptr(pos), as_ptr(vel), as_ptr(pol)))
        check(d.ctx, ccall((:mavi_get_forces, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), d.sys, as_ptr(forces)))
    end
    if is_rings(system)
        rings_info_download!(system, d)
        d.variable && var_download!(system, d)
        d.neighbors && neighbors_download!(system, d)
        d.invasions && invasions_download!(system, d)
    end
end

# --- methods on Mavi's generic functions ------------------------------------

"""calc_forces! on the GPU for plain systems.  A rings system keeps the CPU
pair loop here: its forces! adds springs and area terms after this call
(src/rings/integration.jl:197-226), and its GPU step function replaces the
whole step! instead."""
function Integration.calc_forces!(system::System, chunks, device::GPU)
    if !on_gpu(system) || is_rings(system)
        return Integration.calc_forces!(system, chunks, Sequencial())
    end
    d = device_system(system)
    upload!(system, d)
    forces = get_forces(system)
    GC.@preserve forces check(d.ctx, ccall((:mavi_forces, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), d.sys,
                                           as_ptr(forces)))
    return nothing
end

"""newton_step! (or szabo_step! / rtp_step!, by the system's DynamicCfg) on
the GPU: upload, one step, download (per-step sync: user code may read and
edit the state between steps).  Self-propelled steps draw their noise from
the device stream at (seed, num_steps)."""
function gpu_newton_step!(system)
    d = device_system(system)
    upload!(system, d)
    set_noise_position!(system, d)
    check(d.ctx, ccall((:mavi_step, LIB), Cint, (Ptr{Cvoid}, Int64), d.sys, 1))
    download!(system, d)
    system.time_info.time += system.int_cfg.dt
    system.time_info.num_steps += 1
end

"""Rings.Integration.step! (src/rings/integration.jl:522-543) on the GPU with
per-step sync.  A fixed population takes its rotational noise from
system.rng, one randn per ring in ring order, the draws the reference makes
(src/rings/integration.jl:348); a variable population draws it from the
device stream at (seed, num_steps)."""
function gpu_rings_step!(system)
    d = device_system(system)
    upload!(system, d)
    if d.variable
        set_noise_position!(system, d)
        check(d.ctx, ccall((:mavi_step, LIB), Cint, (Ptr{Cvoid}, Int64), d.sys, 1))
    else
        noise = randn(system.rng, size(system.state.rings_pos, 2))
        check(d.ctx, ccall((:mavi_rings_step, LIB), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}), d.sys, 1, noise))
    end
    download!(system, d)
    system.time_info.num_steps += 1
    system.time_info.time += system.int_cfg.dt
end

function Integration.get_step_function(t::StandardSys, system::System)
    on_gpu(system) && return gpu_newton_step!
    # the reference's selection, src/integration.jl:538-548
    return invoke(Integration.get_step_function, Tuple{StandardSys,Any}, t, system)
end

function Integration.get_step_function(t::Rings.RingsSys, system::System)
    on_gpu(system) && return gpu_rings_step!
    # the reference's step!, src/rings/integration.jl:545
    return invoke(Integration.get_step_function, Tuple{Rings.RingsSys,Any}, t, system)
end

# The device noise stream at step k is a function of (seed, k): keying it by
# the step count makes a run resumed from a saved state continue the stream.
function set_noise_position!(system, d::DeviceSys)
    seed = system.int_cfg.device.seed
    check(d.ctx, ccall((:mavi_sys_noise_set, LIB), Cint, (Ptr{Cvoid}, UInt64, UInt64), d.sys, seed,
                       UInt64(system.time_info.num_steps)))
end

"""run_system with a resident fast path: upload once, every step on the
device (hipGraph replays; rings and self-propelled systems draw their noise
from the device stream), download once."""
function run_gpu!(system; tf=nothing, num_steps=nothing)
    on_gpu(system) || return Mavi.run_system(system; tf=tf, num_steps=num_steps)
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
    set_noise_position!(system, d)
    check(d.ctx, ccall((:mavi_step, LIB), Cint, (Ptr{Cvoid}, Int64), d.sys, n))
    download!(system, d)
    for _ in 1:n
        system.time_info.time += system.int_cfg.dt
    end
    system.time_info.num_steps += n
    return system
end

# --- rings: RingsInfo, neighbours, invasions, variable populations ----------

"""system.info.areas, cms and continuos_pos (RingsInfo, src/rings/rings.jl:118-128)
as the device's last forces! left them."""
function rings_info_download!(system, d::DeviceSys)
    info = system.info
    areas, cms, cpos = info.areas, info.cms, info.continuos_pos
    GC.@preserve areas cms cpos check(d.ctx, ccall((:mavi_rings_info, LIB), Cint,
                                                   (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                                                   d.sys, as_ptr(areas), as_ptr(cms), as_ptr(cpos)))
end

"""ParticleNeighbors of RingsInfo.p_neigh (src/rings/neighbors.jl:47-137) on
the device: calc_interaction's neigh_update! inside every forces!."""
function neighbors_cfg!(system, d::DeviceSys)
    pn = system.info.p_neigh
    isnothing(pn) && return nothing
    inner = pn.neighbors
    inner.cfg.only_count && return nothing  # counts only: the CPU keeps them (no list on the device side)
    check(d.ctx, ccall((:mavi_rings_neighbors_cfg, LIB), Cint, (Ptr{Cvoid}, Cint, Float64, Cint), d.sys, 1,
                       inner.cfg.tol, pn.type == :rings ? 1 : 0))
    d.neighbors = true
end

"""The device's neighbour lists into p_neigh: count (nentities x 1) and list
(15 x nentities x 1) share the layout of the C arrays."""
function neighbors_download!(system, d::DeviceSys)
    inner = system.info.p_neigh.neighbors
    count, list = inner.count, inner.list
    GC.@preserve count list check(d.ctx, ccall((:mavi_rings_neighbors_get, LIB), Cint,
                                               (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}), d.sys, pointer(count),
                                               pointer(list)))
end

"""Hand a RingsSystem's r_chunks and InvasionsCfg to the device
(src/rings/rings.jl:188-202, 276-284): later device steps check for
invasions on the reference's schedule."""
function invasions_cfg!(system, d::DeviceSys)
    extra = system.int_cfg.extra
    rc = system.info.r_chunks
    sched = !isnothing(extra) && !isnothing(extra.invasions_cfg)
    (sched || !isnothing(rc)) || return nothing
    cols = isnothing(rc) ? 0 : rc.num_cols
    rows = isnothing(rc) ? 0 : rc.num_rows
    nc = isnothing(rc) ? 0 : size(rc.chunk_particles, 1)
    every = sched ? extra.invasions_cfg.steps_to_update : 0
    check(d.ctx, ccall((:mavi_rings_invasions_cfg, LIB), Cint, (Ptr{Cvoid}, Cint, Int64, Int32, Int32, Int32),
                       d.sys, sched ? 1 : 0, every, cols, rows, nc))
    d.invasions = true
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
        push!(list, Rings.Invasion(invasor=rec[3k+1], invaded=rec[3k+2], p_id=rec[3k+3]))
    end
    return list
end

sink_geom(g::RectangleCfg) = (MAVI_GEOM_RECT, (g.bottom_left[1], g.bottom_left[2], g.length, g.height))
sink_geom(g::Configs.CircleCfg) = (MAVI_GEOM_CIRCLE, (g.center[1], g.center[2], g.radius, 0.0))

"""Records of system.info.sources (Source / Sink objects the reference built,
src/rings/rings.jl:154-186); the arrays they point into go to keep."""
function source_records(system, keep)
    recs = MaviRingSource[]
    isnothing(system.info.sources) && return recs
    cols = isnothing(system.chunks) ? 0 : system.chunks.num_cols
    for s in system.info.sources
        if s isa Rings.Sources.Sink
            geom, g = sink_geom(s.cfg.geometry_cfg)
            push!(recs, MaviRingSource(MAVI_SRC_SINK, geom, g, 0, 0, 0.0, 0.0, C_NULL, C_NULL, C_NULL, C_NULL))
            continue
        end
        bbox = Float64[]
        for b in s.bbox
            append!(bbox, (b.bottom_left[1], b.bottom_left[2], b.length, b.height))
        end
        spawn = Float64[]
        for pts in s.bbox_spawn_pos, p in pts
            append!(spawn, (p[1], p[2]))
        end
        off, ids = Ptr{Int32}(C_NULL), Ptr{Int32}(C_NULL)
        if s.checker isa Rings.Sources.ChunksChecker
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
    ids isa Rings.States.VarRingsIds || return nothing
    keep = Any[]
    recs = source_records(system, keep)
    mask = Int32.(ids.mask)
    GC.@preserve keep mask recs check(d.ctx, ccall((:mavi_rings_var_cfg, LIB), Cint,
                                                   (Ptr{Cvoid}, Ptr{Int32}, Int32, Ptr{MaviRingSource}), d.sys,
                                                   mask, length(recs), recs))
    d.variable = true
end

"Append n uniform draws of system.rng to the device's spawn polarisation pool (the library draws from its own stream once the pool is empty)."
function pol_draws!(system, d::DeviceSys, n)
    u = rand(system.rng, Float64, n)
    check(d.ctx, ccall((:mavi_rings_pol_draws, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64), d.sys, u, n))
end

"""VarRingsIds (mask, ids, uids, num_active) from the device, then
update_ids! (calc_active_ids!, src/rings/states.jl:200-223) for p_ids."""
function var_download!(system, d::DeviceSys)
    ids = system.state.rings_ids
    nr = length(ids.mask)
    mask = Vector{Int32}(undef, nr)
    rid = Vector{Int64}(undef, nr)
    uid = Vector{Int64}(undef, nr)
    na = Ref{Int64}(0)
    check(d.ctx, ccall((:mavi_rings_var_get, LIB), Cint,
                       (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int64}, Ptr{Int64}, Ref{Int64}, Ptr{Int64}, Ptr{Int64}),
                       d.sys, mask, rid, uid, na, C_NULL, C_NULL))
    ids.mask .= mask .!= 0
    ids.ids .= rid
    ids.uids .= uid
    Mavi.States.update_ids!(system.state)
    return ids
end

# --- run_experiment_batch on a device batch ---------------------------------

# Whether a collector is due at time t, for Mavi's own collectors whose
# schedule is known (src/experiments.jl:364-443); any other collector may
# look at every step.
col_due(col::Experiments.DelayedCol, t) = t - col.state.last_save_time > col.cfg.delay_time
col_due(col::Experiments.ManyCols, t) = any(c -> col_due(c, t), values(col.cols))
col_due(col, t) = true
cp_due(::Nothing, ex, t) = false
cp_due(cfg::Experiments.CheckpointCfg, ex, t) = !(t - ex.checkpoint.last_time < cfg.delta_time)

"""Steps until the next one whose state the host must see: the final time,
a collection or a checkpoint (the conditions of run_experiment and
check_checkpoint, src/experiments.jl:117-131, 448-455, times accumulated
step by step as the reference does)."""
function steps_to_event(ex, tf)
    t = ex.system.time_info.time
    dt = ex.system.int_cfg.dt
    k = 0
    while true
        t += dt
        k += 1
        (t >= tf || col_due(ex.col, t) || cp_due(ex.cfg.checkpoint_cfg, ex, t)) && return k
    end
end

gpu_step_function(f) = f === gpu_newton_step! || f === gpu_rings_step!

"""
    MaviHIP.run_experiment_batch(batch, get_system, stop_func=nothing; verbose=true, use_threads=true)

run_experiment_batch (src/experiments.jl:246-300) with the experiments of a
GPU batch stepped together on the device: experiments whose systems share
dt, time and step count (and take a replica: chunks, a single rectangle
wall, LJ / HarmTrunc gases or rings with at most 32 particles per ring)
run as one mavi_batch, with the reference's collections, checkpoints and
files.  Experiments with a custom step or a stop function, and systems the
batch does not take, run through run_experiment as the reference does.
Replica k's device noise is keyed by hash((seed, k)) and the step count.
"""
function run_experiment_batch(experiment_batch, get_system, stop_func=nothing; verbose=true, use_threads=true)
    init = experiment_batch.init_system
    if !(init.int_cfg.device isa GPU)
        return Experiments.run_experiment_batch(experiment_batch, get_system, stop_func; verbose=verbose,
                                                use_threads=use_threads)
    end
    exp_cfg = experiment_batch.exp_cfg
    col_cfg = experiment_batch.col_cfg
    values = experiment_batch.values
    mkpath(exp_cfg.root)
    Experiments.save_experiment_configs(col_cfg, exp_cfg)
    serialize(joinpath(exp_cfg.root, "values.bin"), values)
    Mavi.save_system(init, joinpath(exp_cfg.root, "init_system"))

    results = Vector{Union{Nothing,Exception}}(nothing, length(values))
    pending = Tuple{Int,Any}[]
    for idx in eachindex(values)
        exp_root = joinpath(exp_cfg.root, "data", string(idx))
        isfile(joinpath(exp_root, ".done")) && continue
        try
            ex = try
                Experiments.load_experiment(exp_root, experiment_batch.custom_step)
            catch
                cfg_i = Experiments.ExperimentCfg(tf=exp_cfg.tf, root=exp_root,
                                                  save_final_state=exp_cfg.save_final_state,
                                                  checkpoint_cfg=exp_cfg.checkpoint_cfg)
                Experiments.Experiment(cfg_i, col_cfg, get_system(init, values[idx], idx),
                                       experiment_batch.custom_step)
            end
            push!(pending, (idx, ex))
        catch e
            results[idx] = e
            verbose && println("$idx: Error setting up experiment: $e")
        end
    end

    groups = Dict{Any,Vector{Tuple{Int,Any}}}()
    solo = Tuple{Int,Any}[]
    for (idx, ex) in pending
        s = ex.system
        if isnothing(stop_func) && gpu_step_function(ex.custom_step) && on_gpu(s)
            key = (s.int_cfg.dt, s.time_info.time, s.time_info.num_steps, typeof(s.type))
            push!(get!(groups, key, Tuple{Int,Any}[]), (idx, ex))
        else
            push!(solo, (idx, ex))
        end
    end
    for members in values_of(groups)
        rest = try
            run_group!(members, exp_cfg.tf)
        catch e
            for (idx, _) in members
                results[idx] = e
                verbose && println("$idx: Error during experiment: $e")
            end
            Tuple{Int,Any}[]
        end
        append!(solo, rest)
    end
    for (idx, ex) in solo
        try
            Experiments.run_experiment(ex, stop_func)
        catch e
            results[idx] = e
            verbose && println("$idx: Error during experiment: $e")
            Mavi.save_system(ex.system, mkpath(joinpath(ex.cfg.root, "error_system")))
        end
    end
    if verbose
        failed = findall(x -> x isa Exception, results)
        isempty(failed) || println("\nFailed experiments: ", failed)
    end
    return results
end

values_of(d::Dict) = collect(Base.values(d))

mutable struct DeviceBatch
    ctx::Ptr{Cvoid}
    b::Ptr{Cvoid}
end

function batch_destroy!(db::DeviceBatch)
    db.b == C_NULL && return nothing
    ccall((:mavi_batch_destroy, LIB), Cint, (Ptr{Cvoid},), db.b)
    db.b = C_NULL
end

"""One device batch for experiments that share dt, time and step count;
returns the members the batch did not take (to run on their own)."""
function run_group!(members, tf)
    length(members) < 2 && return members
    lead = members[1][2].system
    ctx = context(lead.int_cfg.device.device)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ctx, ccall((:mavi_batch_create, LIB), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), ctx, out))
    db = DeviceBatch(ctx, out[])
    try
        taken = Tuple{Int,Any,Int32}[]
        rest = Tuple{Int,Any}[]
        for (idx, ex) in members
            keep = Any[]
            s = ex.system
            desc = describe(s, keep; seed=hash((s.int_cfg.device.seed, idx)) % UInt64)
            ri = Ref{Int32}(0)
            st = GC.@preserve keep ccall((:mavi_batch_add_replica, LIB), Cint, (Ptr{Cvoid}, Ref{SysDesc}, Ref{Int32}),
                                         db.b, desc, ri)
            if st == MAVI_E_UNSUPPORTED
                push!(rest, (idx, ex))
            else
                check(ctx, st)
                push!(taken, (idx, ex, ri[]))
            end
        end
        if length(taken) < 2
            # a batch of one is the system stepped alone
            append!(rest, [(idx, ex) for (idx, ex, _) in taken])
            return rest
        end
        check(ctx, ccall((:mavi_batch_finalize, LIB), Cint, (Ptr{Cvoid},), db.b))
        for (_, ex, ri) in taken
            st = ex.system.state
            pos, vel, pol = st.pos, host_vel(st), host_pol(st)
            GC.@preserve pos vel pol check(ctx, ccall((:mavi_batch_upload, LIB), Cint,
                                                      (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                                                      db.b, ri, as_ptr(pos), as_ptr(vel), as_ptr(pol)))
        end
        check(ctx, ccall((:mavi_batch_noise_set_counter, LIB), Cint, (Ptr{Cvoid}, UInt64), db.b,
                         UInt64(lead.time_info.num_steps)))
        for (_, ex, _) in taken
            # the preamble of run_experiment, src/experiments.jl:105-115
            Mavi.MaviSerder.save_system_configs(ex.system, ex.cfg.root)
            Experiments.save_experiment_configs(ex.col.cfg, ex.cfg)
            mkpath(joinpath(ex.cfg.root, Experiments.COL_DIRNAME))
            if !isnothing(ex.checkpoint)
                mkpath(joinpath(ex.cfg.root, Experiments.CHECKPOINT_DIRNAME, "1"))
                mkpath(joinpath(ex.cfg.root, Experiments.CHECKPOINT_DIRNAME, "2"))
            end
        end
        lead_ex = taken[1][2]
        while lead_ex.system.time_info.time < tf
            k = minimum(steps_to_event(ex, tf) for (_, ex, _) in taken)
            check(ctx, ccall((:mavi_batch_run, LIB), Cint, (Ptr{Cvoid}, Int64), db.b, k))
            for (_, ex, ri) in taken
                s = ex.system
                st = s.state
                pos, vel, pol = st.pos, host_vel(st), host_pol(st)
                forces = get_forces(s)
                GC.@preserve pos vel pol forces begin
                    check(ctx, ccall((:mavi_batch_download, LIB), Cint,
                                     (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                                     db.b, ri, as_ptr(pos), as_ptr(vel), as_ptr(pol)))
                    check(ctx, ccall((:mavi_batch_get_forces, LIB), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                                     db.b, ri, as_ptr(forces)))
                end
                for _ in 1:k
                    s.time_info.time += s.int_cfg.dt
                end
                s.time_info.num_steps += k
                Experiments.collect(ex.col, s)
                Experiments.check_checkpoint(ex.cfg.checkpoint_cfg, ex)
            end
        end
        for (_, ex, _) in taken
            Experiments.save_data(ex.col, mkpath(joinpath(ex.cfg.root, Experiments.COL_DIRNAME)))
            ex.cfg.save_final_state && Mavi.save_system(ex.system, joinpath(ex.cfg.root, Experiments.FINAL_STATE_NAME))
            open(joinpath(ex.cfg.root, ".done"), "w") do f
                write(f, "Experiment concluded on device batch\n")
            end
        end
        return rest
    finally
        batch_destroy!(db)
    end
end

end # module
