# ==============================================================================
# backends/mi355x.jl
#
# MI355X (gfx950) backend for Elastic2D: the time-stepping path runs in
# libfomo_mi355x.so, bound here with ccall (C ABI: include/fomo_mi355x.h of
# the fomo_julia_amd repository).
#
# Included from src/Elastic2D.jl after backends/backend.jl and the core,
# kernels, simulation and visualization files, with one added branch in
# backend(::Symbol) (src/backends/backend.jl:39-51):
#     elseif name in (:mi355x, :MI355X, :amd)
#         return MI355X_BACKEND
#
# What dispatches here, so that init_medium / init_habc / setup_receivers /
# run_shot! / run_shots! / create_geometry run unchanged:
#   to_device(::Array, ::MI355XBackend)      arrays stay host Arrays: Medium{Matrix{Float32}}
#                                           and HABCConfig{Matrix{Float32}} are uploaded once
#                                           per device model (init.jl:93-97,199-201)
#   Wavefield(nx, nz, ::MI355XBackend)       a Wavefield{HipMatrix}: ten HipMatrix fields
#                                           (HipMatrix <: AbstractMatrix{Float32}) that view
#                                           one device wavefield (Structures.jl:18-55)
#   reset!, synchronize, _create_source, _create_receivers, _get_gather
#                                           (shot_manager.jl:51-72,196-222)
#   run_time_loop!(::MI355XBackend, ...)     more specific than time_stepper.jl:97: replayed
#                                           HIP graphs; per-step segments with an on_step
#   run_shots!(::MI355XBackend, ...)         batched: one launch per kernel for a group of
#                                           shots; the reference's loop when on_step is given
#   time_step! and the per-stage kernels     one library call each (Elastic2D.jl:55-62)
#
# Device objects belong to the wavefield's FomoDevice (model, wavefield, a
# cached stage shot) and are freed by its finalizer or when a run binds the
# wavefield to another medium; the shot of a run_time_loop! is freed when the
# loop returns.
# ==============================================================================

const LIBFOMO = get(ENV, "FOMO_MI355X_LIB", "libfomo_mi355x.so")

struct MI355XBackend <: AbstractBackend end
const MI355X_BACKEND = MI355XBackend()

const FOMO_SRC_PRESSURE = Int32(0)
const FOMO_REC_CODE = Dict(:vz => Int32(0), :vx => Int32(1), :p => Int32(2))
const FOMO_FRAME_CODE = Dict(:p => Int32(0), :vel => Int32(1), :vx => Int32(2), :vz => Int32(3))

# fomo_model_desc, field for field
struct FomoModelDesc
    nx::Int32
    nz::Int32
    pad::Int32
    nbc::Int32
    fd_order::Int32
    free_surface::Int32
    dx::Float32
    dz::Float32
    dt::Float32
    dtx::Float32
    dtz::Float32
    qx::Float32
    qz::Float32
    qt_x::Float32
    qt_z::Float32
    qxt::Float32
    storage::Int32
end

# ------------------------------------------------------------------ context
const FOMO_CTX = Ref{Ptr{Cvoid}}(C_NULL)

# Step forms (fomo_ctx_set_step_mode): single shots take the double step (two
# steps per pass over the interior, split steps where a run cannot), batches
# the whole-grid fused step; both give the same results.
const FOMO_STEP_FUSED = Int32(1)
const FOMO_STEP_DOUBLE = Int32(3)

function fomo_ctx()
    if FOMO_CTX[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:fomo_ctx_create, LIBFOMO), Int32, (Int32, Ref{Ptr{Cvoid}}), Int32(0), h)
        fomo_check(rc, "fomo_ctx_create")
        FOMO_CTX[] = h[]
        # two columns per lane, automatic tile height, 16 steps per graph;
        # split-step edge chains of 16-row tiles, one interior block per CU
        rc = ccall((:fomo_ctx_set_tiling, LIBFOMO), Int32, (Ptr{Cvoid}, Int32, Int32, Int32),
                   h[], Int32(2), Int32(0), Int32(16))
        fomo_check(rc, "fomo_ctx_set_tiling")
        rc = ccall((:fomo_ctx_set_split, LIBFOMO), Int32, (Ptr{Cvoid}, Int32, Int32), h[], Int32(16), Int32(1))
        fomo_check(rc, "fomo_ctx_set_split")
        fomo_step_mode!(FOMO_STEP_DOUBLE)
    end
    return FOMO_CTX[]
end

function fomo_step_mode!(mode::Int32)
    rc = ccall((:fomo_ctx_set_step_mode, LIBFOMO), Int32, (Ptr{Cvoid}, Int32), FOMO_CTX[], mode)
    fomo_check(rc, "fomo_ctx_set_step_mode")
end

function fomo_check(rc::Integer, what::AbstractString)
    rc == 0 && return nothing
    msg = ccall((:fomo_ctx_last_error, LIBFOMO), Cstring, (Ptr{Cvoid},), FOMO_CTX[])
    error("$what failed ($rc): $(unsafe_string(msg))")
end

to_device(data::Array, ::MI355XBackend) = data

function synchronize(::MI355XBackend)
    rc = ccall((:fomo_ctx_synchronize, LIBFOMO), Int32, (Ptr{Cvoid},), fomo_ctx())
    fomo_check(rc, "fomo_ctx_synchronize")
end

# ------------------------------------------------------------------ device state
# The device objects behind one Wavefield{HipMatrix}.
mutable struct FomoDevice
    nx::Int
    nz::Int
    model::Ptr{Cvoid}       # fomo_model the wavefield is bound to (C_NULL: none yet)
    model_key::Any          # (M, H, a, params) that model was built from
    wavefield::Ptr{Cvoid}   # fomo_wavefield
    shot::Ptr{Cvoid}        # fomo_shot of the per-stage calls
    shot_key::Any
    version::Int            # bumped whenever the device state may have changed
    function FomoDevice(nx::Int, nz::Int)
        d = new(nx, nz, C_NULL, nothing, C_NULL, C_NULL, nothing, 1)
        finalizer(fomo_release!, d)
        return d
    end
end

function fomo_destroy_shot!(d::FomoDevice)
    if d.shot != C_NULL
        ccall((:fomo_shot_destroy, LIBFOMO), Int32, (Ptr{Cvoid},), d.shot)
        d.shot = C_NULL
        d.shot_key = nothing
    end
end

function fomo_release!(d::FomoDevice)
    fomo_destroy_shot!(d)
    if d.wavefield != C_NULL
        ccall((:fomo_wavefield_destroy, LIBFOMO), Int32, (Ptr{Cvoid},), d.wavefield)
        d.wavefield = C_NULL
    end
    if d.model != C_NULL
        ccall((:fomo_model_destroy, LIBFOMO), Int32, (Ptr{Cvoid},), d.model)
        d.model = C_NULL
        d.model_key = nothing
    end
    return nothing
end

"""
    HipMatrix <: AbstractMatrix{Float32}

One field of a wavefield that lives in HBM.  Reads download the whole field
once per change of the device state (host mirror); writes upload it.
"""
mutable struct HipMatrix <: AbstractMatrix{Float32}
    dev::FomoDevice
    field::Int32            # FOMO_FIELD_* (vx vz txx tzz txz, then the *_old of each)
    cache::Matrix{Float32}  # host mirror
    cached::Int             # dev.version the mirror holds
    dirty::Bool             # mirror written before the device state existed
end

Base.size(A::HipMatrix) = (A.dev.nx, A.dev.nz)
Base.IndexStyle(::Type{HipMatrix}) = IndexCartesian()

function fomo_host!(A::HipMatrix)
    d = A.dev
    if A.cached != d.version && d.wavefield != C_NULL
        rc = ccall((:fomo_wavefield_download, LIBFOMO), Int32, (Ptr{Cvoid}, Int32, Ptr{Float32}),
                   d.wavefield, A.field, A.cache)
        fomo_check(rc, "fomo_wavefield_download")
    end
    A.cached = d.version
    return A.cache
end

function fomo_upload!(A::HipMatrix)
    d = A.dev
    if d.wavefield == C_NULL
        A.dirty = true       # uploaded when a run first binds the wavefield
        return nothing
    end
    rc = ccall((:fomo_wavefield_upload, LIBFOMO), Int32, (Ptr{Cvoid}, Int32, Ptr{Float32}),
               d.wavefield, A.field, A.cache)
    fomo_check(rc, "fomo_wavefield_upload")
    A.dirty = false
    d.version += 1
    A.cached = d.version
    return nothing
end

Base.getindex(A::HipMatrix, i::Int, j::Int) = fomo_host!(A)[i, j]

function Base.setindex!(A::HipMatrix, v, i::Int, j::Int)
    fomo_host!(A)[i, j] = Float32(v)
    fomo_upload!(A)
    return v
end

function Base.fill!(A::HipMatrix, x)
    fomo_host!(A)
    fill!(A.cache, Float32(x))
    fomo_upload!(A)
    return A
end

function Base.copyto!(dest::Matrix{Float32}, src::HipMatrix)
    size(dest) == size(src) || throw(DimensionMismatch("copyto!: $(size(dest)) vs $(size(src))"))
    return copyto!(dest, fomo_host!(src))
end

Base.Array(A::HipMatrix) = copy(fomo_host!(A))

fomo_fields(W::Wavefield{HipMatrix}) =
    (W.vx, W.vz, W.txx, W.tzz, W.txz, W.vx_old, W.vz_old, W.txx_old, W.tzz_old, W.txz_old)

function Wavefield(nx::Int, nz::Int, ::MI355XBackend)
    d = FomoDevice(nx, nz)
    f(id) = HipMatrix(d, Int32(id), zeros(Float32, nx, nz), d.version, false)
    return Wavefield(f(0), f(1), f(2), f(3), f(4), f(5), f(6), f(7), f(8), f(9))
end

# The device model of (M, H, a, params) with W's device wavefield on it.  A
# new medium replaces the previous model; the state carries over.
function fomo_bind!(W::Wavefield{HipMatrix}, M::Medium, H::HABCConfig, a, params::SimParams)
    d = W.vx.dev
    key = (M, H, a, params)
    d.model != C_NULL && d.model_key === key && return d
    (M.nx, M.nz) == (d.nx, d.nz) || throw(DimensionMismatch("wavefield and medium sizes differ"))
    fs = fomo_fields(W)
    if d.wavefield != C_NULL
        for A in fs
            fomo_host!(A)
            A.dirty = true
        end
    end
    fomo_release!(d)
    desc = FomoModelDesc(M.nx, M.nz, M.pad, H.nbc, params.fd_order, M.is_free_surface,
                         M.dx, M.dz, params.dt, params.dtx, params.dtz,
                         H.qx, H.qz, H.qt_x, H.qt_z, H.qxt, Int32(0))
    coef = Vector{Float32}(a)
    med = map(x -> Matrix{Float32}(x), (M.rho_vx, M.rho_vz, M.lam, M.mu_txx, M.mu_txz))
    wts = map(x -> Matrix{Float32}(x), (H.w_vx, H.w_vz, H.w_tau))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:fomo_model_create, LIBFOMO), Int32,
               (Ptr{Cvoid}, Ref{FomoModelDesc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
                Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
                Ptr{Float32}, Ref{Ptr{Cvoid}}),
               fomo_ctx(), desc, coef, med[1], med[2], med[3], med[4], med[5],
               wts[1], wts[2], wts[3], h)
    fomo_check(rc, "fomo_model_create")
    d.model = h[]
    d.model_key = key
    w = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:fomo_wavefield_create, LIBFOMO), Int32, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), d.model, w)
    fomo_check(rc, "fomo_wavefield_create")
    d.wavefield = w[]
    d.version += 1
    for A in fs
        if A.dirty
            fomo_upload!(A)
        else
            fill!(A.cache, 0.0f0)   # the new device state is zero
            A.cached = d.version
        end
    end
    return d
end

function fomo_bound(W::Wavefield{HipMatrix})
    d = W.vx.dev
    d.model == C_NULL && error("this wavefield has no device model yet: run time_step! or " *
                               "run_time_loop!, or call fomo_bind!(W, M, H, a, params)")
    return d
end

# ------------------------------------------------------------------ shots
function fomo_shot_new(model::Ptr{Cvoid}, src, rec, nt::Int)
    wav = src === nothing ? Float32[] : Vector{Float32}(src.wavelet)
    si = src === nothing ? Int32(1) : Int32(src.i)
    sj = src === nothing ? Int32(1) : Int32(src.j)
    ri = rec === nothing ? Int32[] : Vector{Int32}(collect(rec.i))
    rj = rec === nothing ? Int32[] : Vector{Int32}(collect(rec.j))
    code = rec === nothing ? Int32(0) :
           get(() -> error("unknown receiver type $(rec.type)"), FOMO_REC_CODE, rec.type)
    rt = fill(code, length(ri))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:fomo_shot_create, LIBFOMO), Int32,
               (Ptr{Cvoid}, Int32, Int32, Int32, Ptr{Float32}, Int32, Int32, Ptr{Int32},
                Ptr{Int32}, Ptr{Int32}, Int32, Ref{Ptr{Cvoid}}),
               model, si, sj, FOMO_SRC_PRESSURE, wav, Int32(length(wav)), Int32(nt), ri, rj, rt,
               Int32(length(ri)), h)
    fomo_check(rc, "fomo_shot_create")
    return h[]
end

# [nt x n_rec] gather of a shot into a host matrix
function fomo_gather!(data::AbstractMatrix{Float32}, shot::Ptr{Cvoid})
    buf = data isa Matrix{Float32} ? data : Matrix{Float32}(undef, size(data))
    rc = ccall((:fomo_shot_gather, LIBFOMO), Int32, (Ptr{Cvoid}, Ptr{Float32}), shot, buf)
    fomo_check(rc, "fomo_shot_gather")
    buf === data || copyto!(data, buf)
    return data
end

# The per-stage calls' shot, kept while (src, rec) stay the same objects.
function fomo_stage_shot!(d::FomoDevice, src, rec, nt::Int)
    key = (src, rec, nt)
    d.shot != C_NULL && d.shot_key === key && return d.shot
    fomo_destroy_shot!(d)
    d.shot = fomo_shot_new(d.model, src, rec, nt)
    d.shot_key = key
    return d.shot
end

# ------------------------------------------------------------------ reset, helpers of run_shot!
function reset!(::MI355XBackend, W::Wavefield{HipMatrix})
    d = W.vx.dev
    if d.wavefield != C_NULL
        rc = ccall((:fomo_wavefield_reset, LIBFOMO), Int32, (Ptr{Cvoid},), d.wavefield)
        fomo_check(rc, "fomo_wavefield_reset")
    end
    d.version += 1
    for A in fomo_fields(W)
        fill!(A.cache, 0.0f0)
        A.cached = d.version
        A.dirty = false
    end
    return nothing
end
reset!(W::Wavefield{HipMatrix}) = reset!(MI355X_BACKEND, W)

_create_source(::MI355XBackend, M::Medium, shot::ShotConfig, wavelet::AbstractVector{Float32}) =
    Source(_coord_to_index(shot.source_x, M.dx, M.pad), _coord_to_index(shot.source_z, M.dz, M.pad),
           Vector{Float32}(wavelet))

_create_receivers(::MI355XBackend, template::Receivers, nt::Int) =
    Receivers(copy(template.i), copy(template.j), zeros(Float32, nt, length(template.i)), template.type)

_get_gather(::MI355XBackend, rec::Receivers) = copy(rec.data)

# ------------------------------------------------------------------ time stepping
function time_step!(::MI355XBackend, W::Wavefield{HipMatrix}, M::Medium, H::HABCConfig,
                    a, src::Source, rec::Receivers, k::Int, params::SimParams)
    d = fomo_bind!(W, M, H, a, params)
    nt = size(rec.data, 1)
    s = fomo_stage_shot!(d, src, rec, nt)
    rc = ccall((:fomo_time_step, LIBFOMO), Int32, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32),
               d.model, d.wavefield, s, Int32(k))
    fomo_check(rc, "fomo_time_step")
    d.version += 1
    if 1 <= k <= nt && length(rec.i) > 0   # row k of the device gather, as record_receivers! writes it
        g = fomo_gather!(Matrix{Float32}(undef, nt, length(rec.i)), s)
        rec.data[k, :] .= view(g, k, :)
    end
    return nothing
end

function run_time_loop!(backend::MI355XBackend, W::Wavefield{HipMatrix}, M::Medium,
                        H::HABCConfig, a, src::Source, rec::Receivers, params::SimParams;
                        progress::Bool=true, on_step=nothing)
    nt = params.nt
    size(rec.data) == (nt, length(rec.i)) ||
        throw(DimensionMismatch("rec.data must be [nt x n_rec] = ($nt, $(length(rec.i)))"))
    d = fomo_bind!(W, M, H, a, params)
    s = fomo_shot_new(d.model, src, rec, nt)
    try
        if on_step isa MultiFieldRecorder
            fomo_record_frames!(on_step, d, s, params)
        elseif on_step === nothing
            rc = ccall((:fomo_run_time_loop, LIBFOMO), Int32,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32),
                       d.model, d.wavefield, s, Int32(1), Int32(nt))
            fomo_check(rc, "fomo_run_time_loop")
        else
            for k in 1:nt
                rc = ccall((:fomo_run_time_loop, LIBFOMO), Int32,
                           (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32),
                           d.model, d.wavefield, s, Int32(k), Int32(1))
                fomo_check(rc, "fomo_run_time_loop")
                d.version += 1
                if on_step(W, TimeStepInfo(k, k * params.dt, nt)) === false
                    @info "Simulation stopped early at step $k"
                    break
                end
            end
        end
        d.version += 1
        synchronize(backend)
        fomo_gather!(rec.data, s)
    finally
        ccall((:fomo_shot_destroy, LIBFOMO), Int32, (Ptr{Cvoid},), s)
    end
    return nothing
end

# Batched shots: `batch` shots at a time, each on its own device wavefield, one
# launch per kernel for the group.  Results and on_shot_complete calls come in
# shot order, as in shot_manager.jl:145-180.  With an on_step callback (or
# batch = 1) the reference's one-shot-at-a-time loop runs instead.  W ends
# holding the binding to the medium; its fields are not the last shot's state.
function run_shots!(backend::MI355XBackend, W::Wavefield{HipMatrix}, M::Medium, H::HABCConfig,
                    a, rec_template::Receivers, shot_config::MultiShotConfig, params::SimParams;
                    on_shot_complete=nothing, on_step=nothing, progress::Bool=true, batch::Int=16)
    if on_step !== nothing || batch <= 1
        return invoke(run_shots!, Tuple{AbstractBackend, Wavefield, Medium, HABCConfig, Any,
                                        Receivers, MultiShotConfig, SimParams},
                      backend, W, M, H, a, rec_template, shot_config, params;
                      on_shot_complete=on_shot_complete, on_step=on_step, progress=progress)
    end
    d = fomo_bind!(W, M, H, a, params)
    nt = params.nt
    shots = shot_config.shots
    results = Vector{ShotResult}(undef, length(shots))
    n_rec = length(rec_template.i)
    for g0 in 1:batch:length(shots)
        group = g0:min(g0 + batch - 1, length(shots))
        n = length(group)
        wfs = fill(C_NULL, n)
        shs = fill(C_NULL, n)
        srcs = Vector{Any}(undef, n)
        bh = Ref{Ptr{Cvoid}}(C_NULL)
        try
            for (t, idx) in enumerate(group)
                srcs[t] = _create_source(backend, M, shots[idx], shot_config.wavelet)
                w = Ref{Ptr{Cvoid}}(C_NULL)
                rc = ccall((:fomo_wavefield_create, LIBFOMO), Int32, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                           d.model, w)
                fomo_check(rc, "fomo_wavefield_create")
                wfs[t] = w[]
                shs[t] = fomo_shot_new(d.model, srcs[t], rec_template, nt)
            end
            models = fill(d.model, n)
            fomo_step_mode!(FOMO_STEP_FUSED)  # a batch's strips are too large a share to split
            rc = try
                ccall((:fomo_batch_create, LIBFOMO), Int32,
                      (Ptr{Cvoid}, Int32, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}},
                       Ref{Ptr{Cvoid}}),
                      fomo_ctx(), Int32(n), models, wfs, shs, bh)
            finally
                fomo_step_mode!(FOMO_STEP_DOUBLE)
            end
            fomo_check(rc, "fomo_batch_create")
            rc = ccall((:fomo_batch_run, LIBFOMO), Int32, (Ptr{Cvoid}, Int32, Int32),
                       bh[], Int32(1), Int32(nt))
            fomo_check(rc, "fomo_batch_run")
            synchronize(backend)
            for (t, idx) in enumerate(group)
                gather = fomo_gather!(Matrix{Float32}(undef, nt, n_rec), shs[t])
                results[idx] = ShotResult(gather, shots[idx].shot_id, srcs[t].i, srcs[t].j,
                                          _to_cpu_vec(rec_template.i), _to_cpu_vec(rec_template.j))
            end
        finally
            bh[] != C_NULL && ccall((:fomo_batch_destroy, LIBFOMO), Int32, (Ptr{Cvoid},), bh[])
            for s in shs
                s != C_NULL && ccall((:fomo_shot_destroy, LIBFOMO), Int32, (Ptr{Cvoid},), s)
            end
            for w in wfs
                w != C_NULL && ccall((:fomo_wavefield_destroy, LIBFOMO), Int32, (Ptr{Cvoid},), w)
            end
        end
        if on_shot_complete !== nothing
            for idx in group
                on_shot_complete(results[idx])
            end
        end
    end
    return results
end

# ------------------------------------------------------------------ per-stage kernels
function fomo_stage_done(d::FomoDevice, rc::Integer, what::AbstractString)
    fomo_check(rc, what)
    d.version += 1
    return nothing
end

function backup_boundary!(::MI355XBackend, W::Wavefield{HipMatrix}, H::HABCConfig, M::Medium)
    d = fomo_bound(W)
    rc = ccall((:fomo_backup_boundary, LIBFOMO), Int32, (Ptr{Cvoid}, Ptr{Cvoid}), d.model, d.wavefield)
    fomo_stage_done(d, rc, "fomo_backup_boundary")
end

function update_velocity!(::MI355XBackend, W::Wavefield{HipMatrix}, M::Medium, a, params::SimParams)
    d = fomo_bound(W)
    rc = ccall((:fomo_update_velocity, LIBFOMO), Int32, (Ptr{Cvoid}, Ptr{Cvoid}), d.model, d.wavefield)
    fomo_stage_done(d, rc, "fomo_update_velocity")
end

function apply_habc_velocity!(::MI355XBackend, W::Wavefield{HipMatrix}, H::HABCConfig, M::Medium)
    d = fomo_bound(W)
    rc = ccall((:fomo_apply_habc_velocity, LIBFOMO), Int32, (Ptr{Cvoid}, Ptr{Cvoid}), d.model, d.wavefield)
    fomo_stage_done(d, rc, "fomo_apply_habc_velocity")
end

function update_stress!(::MI355XBackend, W::Wavefield{HipMatrix}, M::Medium, a, params::SimParams)
    d = fomo_bound(W)
    rc = ccall((:fomo_update_stress, LIBFOMO), Int32, (Ptr{Cvoid}, Ptr{Cvoid}), d.model, d.wavefield)
    fomo_stage_done(d, rc, "fomo_update_stress")
end

function apply_habc_stress!(::MI355XBackend, W::Wavefield{HipMatrix}, H::HABCConfig, M::Medium)
    d = fomo_bound(W)
    rc = ccall((:fomo_apply_habc_stress, LIBFOMO), Int32, (Ptr{Cvoid}, Ptr{Cvoid}), d.model, d.wavefield)
    fomo_stage_done(d, rc, "fomo_apply_habc_stress")
end

function apply_free_surface!(::MI355XBackend, W::Wavefield{HipMatrix}, M::Medium)
    d = fomo_bound(W)
    rc = ccall((:fomo_apply_free_surface, LIBFOMO), Int32, (Ptr{Cvoid}, Ptr{Cvoid}), d.model, d.wavefield)
    fomo_stage_done(d, rc, "fomo_apply_free_surface")
end

function inject_source!(::MI355XBackend, W::Wavefield{HipMatrix}, src::Source, k::Int, dt)
    d = fomo_bound(W)
    s = fomo_stage_shot!(d, src, nothing, max(k, length(src.wavelet), 1))
    rc = ccall((:fomo_inject_source, LIBFOMO), Int32, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32),
               d.model, d.wavefield, s, Int32(k))
    fomo_stage_done(d, rc, "fomo_inject_source")
end

function record_receivers!(::MI355XBackend, W::Wavefield{HipMatrix}, rec::Receivers, k::Int)
    d = fomo_bound(W)
    nt = size(rec.data, 1)
    s = fomo_stage_shot!(d, nothing, rec, nt)
    rc = ccall((:fomo_record_receivers, LIBFOMO), Int32, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32),
               d.model, d.wavefield, s, Int32(k))
    fomo_check(rc, "fomo_record_receivers")
    if 1 <= k <= nt && length(rec.i) > 0
        g = fomo_gather!(Matrix{Float32}(undef, nt, length(rec.i)), s)
        rec.data[k, :] .= view(g, k, :)
    end
    return nothing
end

# ------------------------------------------------------------------ video frames
# A per-step on_step callback reads W through the host mirrors:
function _get_field_data(W::Wavefield{HipMatrix}, field::Symbol)
    field == :p && return @. (fomo_host!(W.txx) + fomo_host!(W.tzz)) * 0.5f0
    field == :vx && return copy(fomo_host!(W.vx))
    field == :vz && return copy(fomo_host!(W.vz))
    field == :vel && return @. sqrt(fomo_host!(W.vx)^2 + fomo_host!(W.vz)^2)
    error("Unknown field: $field. Use :p, :vel, :vx, or :vz")
end

# A MultiFieldRecorder as on_step (time_stepper.jl:112-119) keeps the graph
# replay: frames are taken on the device after every skip-th step into a
# frame ring (fomo_run_time_loop_frames), drained once per ring, and handed
# to each _SingleRecorder as its own extraction would (video_recorder.jl:207-228).
function fomo_record_frames!(rec::MultiFieldRecorder, d::FomoDevice, s::Ptr{Cvoid}, p::SimParams)
    cfg = rec.config
    skip = cfg.skip
    nf = length(cfg.fields)
    ds = cfg.downsample
    nx_d, nz_d = cld(d.nx, ds), cld(d.nz, ds)
    cap = max(1, min(cld(p.nt, skip), (1 << 30) ÷ (4 * nf * nx_d * nz_d)))
    codes = Int32[FOMO_FRAME_CODE[f] for f in cfg.fields]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:fomo_frames_create, LIBFOMO), Int32,
               (Ptr{Cvoid}, Int32, Int32, Ptr{Int32}, Int32, Int32, Int32, Int32, Ref{Ptr{Cvoid}}),
               fomo_ctx(), Int32(d.nx), Int32(d.nz), codes, Int32(nf), Int32(skip), Int32(ds),
               Int32(cap), h)
    fomo_check(rc, "fomo_frames_create")
    buf = Array{Float32}(undef, nx_d, nz_d, nf, cap)   # [slot][field][j][i] in C order
    try
        k = 1
        while k <= p.nt
            n = min(cap * skip, p.nt - k + 1)
            rc = ccall((:fomo_run_time_loop_frames, LIBFOMO), Int32,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32, Ptr{Cvoid}),
                       d.model, d.wavefield, s, Int32(k), Int32(n), h[])
            fomo_check(rc, "fomo_run_time_loop_frames")
            rc = ccall((:fomo_frames_download, LIBFOMO), Int32, (Ptr{Cvoid}, Int32, Int32, Ptr{Float32}),
                       h[], Int32(0), Int32(cap), buf)
            fomo_check(rc, "fomo_frames_download")
            for fr in cld(k, skip):((k + n - 1) ÷ skip)
                kk = fr * skip
                slot = (fr - 1) % cap + 1
                for (q, f) in enumerate(cfg.fields)
                    sr = rec.recorders[f]
                    sr.frame_buffer .= view(buf, :, :, q, slot)
                    if sr.clim === nothing && !sr.auto_clim_done && sr.frame_count >= 5
                        rms = sqrt(sum(sr.frame_buffer .^ 2) / length(sr.frame_buffer))
                        if rms > 1e-10
                            sr.hm.colorrange = (-3.0f0 * rms, 3.0f0 * rms)
                            sr.auto_clim_done = true
                        end
                    end
                    sr.hm[1] = sr.frame_buffer
                    sr.title_obs[] = _field_title(f, Float32(kk * p.dt), kk, p.nt)
                    recordframe!(sr.video_stream)
                    sr.frame_count += 1
                end
            end
            k += n
        end
    finally
        ccall((:fomo_frames_destroy, LIBFOMO), Int32, (Ptr{Cvoid},), h[])
    end
    d.version += 1
    return nothing
end
