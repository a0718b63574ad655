# FomoHIP.jl -- Fomo.jl's forward-modeling API over the MI355X engine's C-ABI (include/fomo_hip.h,
# fomo.jl_amd/libfomo_hip.so).
#
# A drop-in for the CUDA.jl solver bodies: a Fomo.jl maintainer includes this file inside the Fomo
# module (so the reference's own _check_geometry / _check_numerics / _check_numerics_centered
# (src/utils/checks.jl), ricker_wavelet (src/utils/ricker_wavelet.jl) and get_fd_coefficients /
# get_centered_d1 / get_centered_d2 (src/utils/FD_utils.jl, FD_centered.jl) are in scope) and exports these
# functions in place of the CUDA.jl ones (src/Fomo.jl:107-117).  Names, positional arguments,
# keywords, defaults, validation messages and returns are the reference's:
#   acoustic2d        src/equations/acoustic2d/acoustic2d.jl:47-64, 187-191  -> NamedTuple
#   elastic2d         src/equations/elastic2d/elastic2d.jl:45-163            -> NamedTuple
#   scalar2d          src/equations/scalar2d/scalar2d.jl:104-189             -> NamedTuple
#   coupled2d         src/equations/coupled2d/coupled2d.jl:53-183            -> NamedTuple
#   acoustic2d_batch  src/equations/acoustic2d/batch.jl:180-261              -> NamedTuple
#   elastic2d_batch   src/equations/elastic2d/batch.jl:197-261               -> NamedTuple
#   acoustic3d        src/equations/acoustic3d/acoustic3d.jl:31-129          -> plain 4-tuple
#   elastic3d         src/equations/elastic3d/elastic3d.jl:31-117            -> plain 4-tuple
# Every host array stays Julia-owned (column-major Float32 / Int32, passed as pointers; ccall keeps
# its arguments alive for the call); the library keeps no pointer after a call returns.  Status
# codes map to exceptions: FOMO_EINVAL -> ArgumentError, anything else -> ErrorException, with
# fomo_last_error() as the message.  `use_cuda_graph` means "use a hipGraph"; `verbose` is
# accepted and prints the reference's two progress lines.
#
# Extensions beyond the reference are keywords with reference-preserving defaults: `storage`
# (:f32 bit-identical to the reference; :f16 / :bf16 16-bit storage with Float32 arithmetic;
# :mat16 Float32 fields with Float16 materials, 2D only).

const LIB = get(ENV, "FOMO_HIP_LIB", joinpath(@__DIR__, "..", "libfomo_hip.so"))

# status codes and enums (include/fomo_hip.h)
const FOMO_OK = Int32(0)
const FOMO_EINVAL = Int32(1)
const FOMO_EOVERFLOW = Int32(5)
const FOMO_ACOUSTIC2D = Int32(1)
const FOMO_ELASTIC2D = Int32(2)
const FOMO_ACOUSTIC3D = Int32(3)
const FOMO_ELASTIC3D = Int32(4)
const FOMO_SCALAR2D = Int32(5)
const FOMO_COUPLED2D = Int32(6)
const FOMO_BOUNDARY_HABC = Int32(0)
const FOMO_BOUNDARY_SPONGE = Int32(1)
const STORAGE_CODES = Dict(:f32 => Int32(0), :f16 => Int32(1), :bf16 => Int32(2), :mat16 => Int32(3))
const SNAP_PLANES = Dict(:xz => Int32(0), :xy => Int32(1), :yz => Int32(2))

# struct fomo_desc, field by field (4-byte fields, then one Float64 at an 8-byte-aligned offset, so
# the Julia and C layouts agree)
struct FomoDesc
    eq::Int32
    nx::Int32
    ny::Int32
    nz::Int32
    nbc::Int32
    fd_order::Int32
    nt::Int32
    dh::Float32
    dt::Float32
    v_ref::Float32
    n_src::Int32
    n_rec::Int32
    source_type::Int32
    storage::Int32
    snap_interval::Int32
    snap_plane::Int32
    snap_index::Int32
    use_graph::Int32
    record_p::Int32
    tile_cfg::Int32
    boundary::Int32
    sponge_factor::Float32
    smooth_sigma::Float64
    v_ref_s::Float32
    reserved::NTuple{1,Int32}
end

# struct fomo_shot (one window of a windowed batch plan, fomo_batch_create)
struct FomoShot
    x0::Int32
    nx::Int32
    n_src::Int32
    sx::Ptr{Int32}
    sz::Ptr{Int32}
    wavelet::Ptr{Float32}
    n_rec::Int32
    rx::Ptr{Int32}
    rz::Ptr{Int32}
    v_ref::Float32
end

function _desc(eq, nx, ny, nz, nbc, fd_order, nt, dh, dt, v_ref, n_src, n_rec; source_type=0, storage=:f32,
               snap_interval=0, snap_plane=:xz, snap_index=0, use_graph=true, boundary=:habc,
               sponge_factor=0.015f0, smooth_sigma=0.0, v_ref_s=0.0f0)
    haskey(STORAGE_CODES, storage) || throw(ArgumentError("storage must be :f32, :f16, :bf16 or :mat16, got $storage"))
    return FomoDesc(Int32(eq), Int32(nx), Int32(ny), Int32(nz), Int32(nbc), Int32(fd_order), Int32(nt),
                    Float32(dh), Float32(dt), Float32(v_ref), Int32(n_src), Int32(n_rec), Int32(source_type),
                    STORAGE_CODES[storage], Int32(snap_interval), SNAP_PLANES[snap_plane], Int32(snap_index),
                    Int32(use_graph), Int32(0), Int32(0),
                    boundary === :sponge ? FOMO_BOUNDARY_SPONGE : FOMO_BOUNDARY_HABC, Float32(sponge_factor),
                    Float64(smooth_sigma), Float32(v_ref_s), (Int32(0),))
end

last_error() = unsafe_string(ccall((:fomo_last_error, LIB), Cstring, ()))

function _check(rc::Cint)
    rc == FOMO_OK && return nothing
    msg = last_error()
    rc == FOMO_EINVAL && throw(ArgumentError(msg))
    error("fomo error $rc: $msg")
end

_i32(v) = Int32.(collect(v))
_snapbuf(cells, nsnap) = nsnap > 0 ? zeros(Float32, cells, nsnap) : zeros(Float32, 0, 0)
_snapptr(buf) = isempty(buf) ? Ptr{Float32}(C_NULL) : pointer(buf)

# acoustic2d.jl:91-97: Ricker(f0) or the user's wavelet, one row per source, optionally * dt
function _wavelet_rows(wavelet, f0, dt, nt, n_src, scale_source_by_dt)
    w = isnothing(wavelet) ? ricker_wavelet(f0, dt, nt) : Float32.(wavelet)
    length(w) == nt || throw(ArgumentError("wavelet 长度 $(length(w)) ≠ nt=$nt"))
    m = repeat(reshape(w, 1, nt), n_src, 1)
    scale_source_by_dt && (m .*= dt)
    return m
end

function acoustic2d(vp::Matrix{Float32}, rho::Matrix{Float32}, dh::Float32, dt::Float32, nt::Int, f0::Float32;
                    sx::Vector{Int}, sz::Vector{Int}, rx::AbstractVector{Int}, rz::AbstractVector{Int},
                    nbc::Int=50, fd_order::Int=8, snap_interval::Int=0, boundary::Symbol=:habc,
                    v_ref::Float32=Float32(minimum(x for x in vp if x > 0.0f0)), sponge_factor::Float32=0.015f0,
                    wavelet::Union{Nothing,AbstractVector{<:Real}}=nothing, scale_source_by_dt::Bool=false,
                    use_cuda_graph::Bool=true, verbose::Bool=true, storage::Symbol=:f32)
    boundary in (:habc, :sponge) || throw(ArgumentError("boundary must be :habc or :sponge, got $boundary"))
    nx, nz = size(vp)
    _check_geometry(nx, nz, sx, sz, rx, rz)
    _check_numerics(maximum(vp), minimum(x for x in vp if x > 0.0f0), dh, dt, f0, fd_order)
    get_fd_coefficients(fd_order)
    wm = _wavelet_rows(wavelet, f0, dt, nt, length(sx), scale_source_by_dt)
    d = _desc(FOMO_ACOUSTIC2D, nx, 1, nz, nbc, fd_order, nt, dh, dt, v_ref, length(sx), length(rx);
              storage=storage, snap_interval=snap_interval, use_graph=use_cuda_graph, boundary=boundary,
              sponge_factor=sponge_factor)
    nrec = length(rx)
    seis_p, seis_vx, seis_vz = (zeros(Float32, nrec, nt) for _ in 1:3)
    nsnap = snap_interval > 0 ? nt ÷ snap_interval : 0
    sb = _snapbuf(nx * nz, nsnap)
    t = Ref{Float64}(0.0)
    verbose && @info "Starting acoustic2d simulation... (nx=$nx, nz=$nz, nt=$nt, boundary=$boundary)"
    GC.@preserve sb begin
        _check(ccall((:fomo_acoustic2d, LIB), Cint,
                     (Ref{FomoDesc}, Ptr{Float32}, Ptr{Float32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float32}, Ptr{Int32},
                      Ptr{Int32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ref{Float64}),
                     d, vp, rho, _i32(sx), _i32(sz), wm, _i32(rx), _i32(rz), seis_p, seis_vx, seis_vz,
                     _snapptr(sb), t))
    end
    verbose && @info "Simulation complete! GPU time: $(round(t[], digits=3))s"
    snaps = Matrix{Float32}[reshape(sb[:, k], nx, nz) for k in 1:nsnap]
    return (seis_p=seis_p, seis_vx=seis_vx, seis_vz=seis_vz, snaps=snaps, stats=(kernel_time_s=t[],))
end

function elastic2d(vp::Matrix{Float32}, vs::Matrix{Float32}, rho::Matrix{Float32}, dh::Float32, dt::Float32,
                   nt::Int, f0::Float32;
                   sx::Vector{Int}, sz::Vector{Int}, rx::AbstractVector{Int}, rz::AbstractVector{Int},
                   nbc::Int=50, fd_order::Int=8, snap_interval::Int=0, boundary::Symbol=:habc,
                   v_ref::Float32=Float32(minimum(x for x in vp if x > 0.0f0)), sponge_factor::Float32=0.015f0,
                   wavelet::Union{Nothing,AbstractVector{<:Real}}=nothing, scale_source_by_dt::Bool=false,
                   use_cuda_graph::Bool=true, verbose::Bool=true, storage::Symbol=:f32)
    boundary in (:habc, :sponge) || throw(ArgumentError("boundary must be :habc or :sponge, got $boundary"))
    nx, nz = size(vp)
    _check_geometry(nx, nz, sx, sz, rx, rz)
    vmin_disp = any(x -> x > 0.0f0, vs) ? minimum(x for x in vs if x > 0.0f0) : minimum(x for x in vp if x > 0.0f0)
    _check_numerics(maximum(vp), vmin_disp, dh, dt, f0, fd_order)
    get_fd_coefficients(fd_order)
    wm = _wavelet_rows(wavelet, f0, dt, nt, length(sx), scale_source_by_dt)
    d = _desc(FOMO_ELASTIC2D, nx, 1, nz, nbc, fd_order, nt, dh, dt, v_ref, length(sx), length(rx);
              storage=storage, snap_interval=snap_interval, use_graph=use_cuda_graph, boundary=boundary,
              sponge_factor=sponge_factor)
    nrec = length(rx)
    seis_vx, seis_vz = (zeros(Float32, nrec, nt) for _ in 1:2)
    nsnap = snap_interval > 0 ? nt ÷ snap_interval : 0
    sb = _snapbuf(nx * nz, nsnap)
    t = Ref{Float64}(0.0)
    verbose && @info "Starting elastic2d simulation... (nx=$nx, nz=$nz, nt=$nt, boundary=$boundary)"
    GC.@preserve sb begin
        _check(ccall((:fomo_elastic2d, LIB), Cint,
                     (Ref{FomoDesc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float32},
                      Ptr{Int32}, Ptr{Int32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ref{Float64}),
                     d, vp, vs, rho, _i32(sx), _i32(sz), wm, _i32(rx), _i32(rz), seis_vx, seis_vz, _snapptr(sb), t))
    end
    verbose && @info "Simulation complete! GPU time: $(round(t[], digits=3))s"
    snaps = Matrix{Float32}[reshape(sb[:, k], nx, nz) for k in 1:nsnap]
    return (seis_vx=seis_vx, seis_vz=seis_vz, snaps=snaps, stats=(kernel_time_s=t[],))
end

function scalar2d(vp::AbstractMatrix{Float32}, dh::Float32, dt::Float32, nt::Int, f0::Float32;
                  sx::AbstractVector{Int}, sz::AbstractVector{Int}, rx::AbstractVector{Int},
                  rz::AbstractVector{Int}, nbc::Int=50, fd_order::Int=8,
                  v_ref::Float32=Float32(minimum(x for x in vp if x > 0.0f0)),
                  wavelet::Union{Nothing,AbstractVector{<:Real}}=nothing, source_scale::Symbol=:v2dt2,
                  verbose::Bool=true)
    source_scale in (:v2dt2, :dt, :none) ||
        throw(ArgumentError("source_scale 须为 :v2dt2 / :dt / :none,got $source_scale"))
    nx, nz = size(vp)
    _check_geometry(nx, nz, sx, sz, rx, rz)
    _check_numerics_centered(maximum(vp), minimum(x for x in vp if x > 0.0f0), dh, dt, f0, fd_order)
    get_centered_d2(fd_order)
    w = isnothing(wavelet) ? ricker_wavelet(f0, dt, nt) : Float32.(wavelet)
    length(w) == nt || throw(ArgumentError("wavelet 长度 $(length(w)) ≠ nt=$nt"))
    wm = repeat(reshape(w, 1, nt), length(sx), 1)
    if source_scale === :v2dt2          # scalar2d.jl:139-146
        for q in 1:length(sx)
            wm[q, :] .*= Float32(vp[sx[q], sz[q]]^2 * dt^2)
        end
    elseif source_scale === :dt
        wm .*= dt
    end
    d = _desc(FOMO_SCALAR2D, nx, 1, nz, nbc, fd_order, nt, dh, dt, v_ref, length(sx), length(rx))
    seis_u = zeros(Float32, length(rx), nt)
    t = Ref{Float64}(0.0)
    verbose && @info "Starting scalar2d... (nx=$nx, nz=$nz, nt=$nt)"
    _check(ccall((:fomo_scalar2d, LIB), Cint,
                 (Ref{FomoDesc}, Ptr{Float32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float32}, Ptr{Int32}, Ptr{Int32},
                  Ptr{Float32}, Ref{Float64}),
                 d, Matrix{Float32}(vp), _i32(sx), _i32(sz), wm, _i32(rx), _i32(rz), seis_u, t))
    verbose && @info "Scalar complete! GPU time: $(round(t[], digits=3))s"
    return (seis_u=seis_u, stats=(kernel_time_s=t[],))
end

function coupled2d(vp::Matrix{Float32}, vs::Matrix{Float32}, dh::Float32, dt::Float32, nt::Int, f0::Float32;
                   sx::Vector{Int}, sz::Vector{Int}, rx::AbstractVector{Int}, rz::AbstractVector{Int},
                   nbc::Int=50, fd_order::Int=8, snap_interval::Int=0,
                   v_ref_p::Float32=Float32(minimum(x for x in vp if x > 0.0f0)),
                   v_ref_s::Float32=Float32(minimum(x for x in vs if x > 0.0f0)),
                   smooth_sigma::Float64=3.0,
                   wavelet::Union{Nothing,AbstractVector{<:Real}}=nothing,
                   verbose::Bool=true)
    nx, nz = size(vp)
    _check_geometry(nx, nz, sx, sz, rx, rz)
    vmin_disp = any(x -> x > 0.0f0, vs) ? minimum(x for x in vs if x > 0.0f0) : minimum(x for x in vp if x > 0.0f0)
    _check_numerics_centered(maximum(vp), vmin_disp, dh, dt, f0, fd_order)
    get_centered_d1(fd_order)
    get_centered_d2(fd_order)
    w = isnothing(wavelet) ? ricker_wavelet(f0, dt, nt) : Float32.(wavelet)
    length(w) == nt || throw(ArgumentError("wavelet 长度 $(length(w)) ≠ nt=$nt"))
    wm = repeat(reshape(w, 1, nt), length(sx), 1)
    d = _desc(FOMO_COUPLED2D, nx, 1, nz, nbc, fd_order, nt, dh, dt, v_ref_p, length(sx), length(rx);
              snap_interval=snap_interval, smooth_sigma=smooth_sigma, v_ref_s=v_ref_s)
    nrec = length(rx)
    seis_P, seis_S = (zeros(Float32, nrec, nt) for _ in 1:2)
    nsnap = snap_interval > 0 ? nt ÷ snap_interval : 0
    sbP, sbS = _snapbuf(nx * nz, nsnap), _snapbuf(nx * nz, nsnap)
    t = Ref{Float64}(0.0)
    verbose && @info "Starting coupled2d simulation... (nx=$nx, nz=$nz, nt=$nt)"
    GC.@preserve sbP sbS begin
        _check(ccall((:fomo_coupled2d, LIB), Cint,
                     (Ref{FomoDesc}, Ptr{Float32}, Ptr{Float32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float32}, Ptr{Int32},
                      Ptr{Int32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ref{Float64}),
                     d, vp, vs, _i32(sx), _i32(sz), wm, _i32(rx), _i32(rz), seis_P, seis_S, _snapptr(sbP),
                     _snapptr(sbS), t))
    end
    verbose && @info "Simulation complete! GPU time: $(round(t[], digits=3))s"
    snaps_P = Matrix{Float32}[reshape(sbP[:, k], nx, nz) for k in 1:nsnap]
    snaps_S = Matrix{Float32}[reshape(sbS[:, k], nx, nz) for k in 1:nsnap]
    return (seis_P=seis_P, seis_S=seis_S, snaps_P=snaps_P, snaps_S=snaps_S, stats=(kernel_time_s=t[],))
end

# batch.jl:193-199: a vector means one source per shot; a matrix is (n_shots, n_src_per_shot).  The
# C-ABI takes (n_src, n_shots) column-major, as init_batched_source stores them (batch_acq.jl:27-37).
function _shot_matrix(sx, sz)
    sxm = sx isa AbstractVector ? reshape(collect(Int, sx), :, 1) : collect(Int, sx)
    szm = sz isa AbstractVector ? reshape(collect(Int, sz), :, 1) : collect(Int, sz)
    size(sxm) == size(szm) || throw(ArgumentError("sx 与 sz 尺寸不一致: $(size(sxm)) vs $(size(szm))"))
    size(sxm, 1) >= 1 || throw(ArgumentError("至少需要 1 炮"))
    return sxm, szm
end

function acoustic2d_batch(vp::AbstractMatrix{Float32}, rho::AbstractMatrix{Float32}, dh::Float32, dt::Float32,
                          nt::Int, f0::Float32; sx, sz, rx::AbstractVector{Int}, rz::AbstractVector{Int},
                          nbc::Int=50, fd_order::Int=8,
                          v_ref::Float32=Float32(minimum(x for x in vp if x > 0.0f0)),
                          wavelet::Union{Nothing,AbstractVector{<:Real}}=nothing, scale_source_by_dt::Bool=false,
                          verbose::Bool=true, storage::Symbol=:f32)
    nx, nz = size(vp)
    sxm, szm = _shot_matrix(sx, sz)
    ns, n_src = size(sxm)
    for s in 1:ns
        _check_geometry(nx, nz, vec(sxm[s, :]), vec(szm[s, :]), rx, rz)
    end
    _check_numerics(maximum(vp), minimum(x for x in vp if x > 0.0f0), dh, dt, f0, fd_order)
    get_fd_coefficients(fd_order)
    w = _wavelet_rows(wavelet, f0, dt, nt, n_src, scale_source_by_dt)
    d = _desc(FOMO_ACOUSTIC2D, nx, 1, nz, nbc, fd_order, nt, dh, dt, v_ref, n_src, length(rx); storage=storage)
    nrec = length(rx)
    seis_p, seis_vx, seis_vz = (zeros(Float32, nrec, nt, ns) for _ in 1:3)
    t = Ref{Float64}(0.0)
    verbose && @info "Starting acoustic2d_batch... (nx=$nx, nz=$nz, nt=$nt, shots=$ns)"
    _check(ccall((:fomo_acoustic2d_batch, LIB), Cint,
                 (Ref{FomoDesc}, Ptr{Float32}, Ptr{Float32}, Int32, Ptr{Int32}, Ptr{Int32}, Ptr{Float32},
                  Ptr{Int32}, Ptr{Int32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ref{Float64}),
                 d, Matrix{Float32}(vp), Matrix{Float32}(rho), Int32(ns), Int32.(permutedims(sxm)),
                 Int32.(permutedims(szm)), w, _i32(rx), _i32(rz), seis_p, seis_vx, seis_vz, t))
    return (seis_p=seis_p, seis_vx=seis_vx, seis_vz=seis_vz, stats=(kernel_time_s=t[], n_shots=ns))
end

function elastic2d_batch(vp::AbstractMatrix{Float32}, vs::AbstractMatrix{Float32}, rho::AbstractMatrix{Float32},
                         dh::Float32, dt::Float32, nt::Int, f0::Float32; sx, sz, rx::AbstractVector{Int},
                         rz::AbstractVector{Int}, nbc::Int=50, fd_order::Int=8,
                         v_ref::Float32=Float32(minimum(x for x in vp if x > 0.0f0)),
                         wavelet::Union{Nothing,AbstractVector{<:Real}}=nothing, scale_source_by_dt::Bool=false,
                         verbose::Bool=true, storage::Symbol=:f32)
    nx, nz = size(vp)
    sxm, szm = _shot_matrix(sx, sz)
    ns, n_src = size(sxm)
    for s in 1:ns
        _check_geometry(nx, nz, vec(sxm[s, :]), vec(szm[s, :]), rx, rz)
    end
    vmin_disp = any(x -> x > 0.0f0, vs) ? minimum(x for x in vs if x > 0.0f0) : minimum(x for x in vp if x > 0.0f0)
    _check_numerics(maximum(vp), vmin_disp, dh, dt, f0, fd_order)
    get_fd_coefficients(fd_order)
    w = _wavelet_rows(wavelet, f0, dt, nt, n_src, scale_source_by_dt)
    d = _desc(FOMO_ELASTIC2D, nx, 1, nz, nbc, fd_order, nt, dh, dt, v_ref, n_src, length(rx); storage=storage)
    nrec = length(rx)
    seis_vx, seis_vz = (zeros(Float32, nrec, nt, ns) for _ in 1:2)
    t = Ref{Float64}(0.0)
    verbose && @info "Starting elastic2d_batch... (nx=$nx, nz=$nz, nt=$nt, shots=$ns)"
    _check(ccall((:fomo_elastic2d_batch, LIB), Cint,
                 (Ref{FomoDesc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Int32, Ptr{Int32}, Ptr{Int32},
                  Ptr{Float32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float32}, Ptr{Float32}, Ref{Float64}),
                 d, Matrix{Float32}(vp), Matrix{Float32}(vs), Matrix{Float32}(rho), Int32(ns),
                 Int32.(permutedims(sxm)), Int32.(permutedims(szm)), w, _i32(rx), _i32(rz), seis_vx, seis_vz, t))
    return (seis_vx=seis_vx, seis_vz=seis_vz, stats=(kernel_time_s=t[], n_shots=ns))
end

# acoustic3d.jl:47-52 / elastic3d.jl:47-51: snap_index 0 selects ny ÷ 2 (for every plane, as the
# reference does); snapshots are 2D slices of p (acoustic) or vz (elastic) in snap_plane.
function _slice_dims(snap_plane, nx, ny, nz)
    haskey(SNAP_PLANES, snap_plane) || throw(ArgumentError("snap_plane must be :xy, :xz or :yz, got $snap_plane"))
    return snap_plane === :xz ? (nx, nz) : (snap_plane === :xy ? (nx, ny) : (ny, nz))
end

function acoustic3d(vp::Array{Float32,3}, rho::Array{Float32,3}, dh::Float32, dt::Float32, nt::Int, f0::Float32;
                    sx::Vector{Int}, sy::Vector{Int}, sz::Vector{Int}, rx::AbstractVector{Int},
                    ry::AbstractVector{Int}, rz::AbstractVector{Int}, nbc::Int=50, fd_order::Int=8,
                    snap_interval::Int=0, snap_plane::Symbol=:xz, snap_index::Int=0,
                    v_ref::Float32=minimum(vp[vp.>0.0f0]), storage::Symbol=:f32)
    nx, ny, nz = size(vp)
    snap_index == 0 && (snap_index = ny ÷ 2)
    get_fd_coefficients(fd_order)
    w = ricker_wavelet(f0, dt, nt)
    wm = repeat(reshape(w, 1, nt), length(sx), 1)
    d = _desc(FOMO_ACOUSTIC3D, nx, ny, nz, nbc, fd_order, nt, dh, dt, v_ref, length(sx), length(rx);
              storage=storage, snap_interval=snap_interval, snap_plane=snap_plane, snap_index=snap_index)
    nrec = length(rx)
    seis_vx, seis_vy, seis_vz = (zeros(Float32, nrec, nt) for _ in 1:3)
    a, b = _slice_dims(snap_plane, nx, ny, nz)
    nsnap = snap_interval > 0 ? nt ÷ snap_interval : 0
    sb = _snapbuf(a * b, nsnap)
    t = Ref{Float64}(0.0)
    @info "Starting acoustic3d simulation... (nx=$nx, ny=$ny, nz=$nz, nt=$nt)"
    GC.@preserve sb begin
        _check(ccall((:fomo_acoustic3d, LIB), Cint,
                     (Ref{FomoDesc}, Ptr{Float32}, Ptr{Float32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float32},
                      Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
                      Ptr{Float32}, Ref{Float64}),
                     d, vp, rho, _i32(sx), _i32(sy), _i32(sz), wm, _i32(rx), _i32(ry), _i32(rz), seis_vx, seis_vy,
                     seis_vz, Ptr{Float32}(C_NULL), _snapptr(sb), t))
    end
    @info "Simulation complete! GPU time: $(round(t[], digits=3))s"
    snaps = Matrix{Float32}[reshape(sb[:, k], a, b) for k in 1:nsnap]
    return seis_vx, seis_vy, seis_vz, snaps
end

function elastic3d(vp::Array{Float32,3}, vs::Array{Float32,3}, rho::Array{Float32,3}, dh::Float32, dt::Float32,
                   nt::Int, f0::Float32; sx::Vector{Int}, sy::Vector{Int}, sz::Vector{Int},
                   rx::AbstractVector{Int}, ry::AbstractVector{Int}, rz::AbstractVector{Int}, nbc::Int=50,
                   fd_order::Int=8, snap_interval::Int=0, snap_plane::Symbol=:xz, snap_index::Int=0,
                   v_ref::Float32=minimum(vp[vp.>0.0f0]), storage::Symbol=:f32)
    nx, ny, nz = size(vp)
    snap_index == 0 && (snap_index = ny ÷ 2)
    get_fd_coefficients(fd_order)
    w = ricker_wavelet(f0, dt, nt)
    wm = repeat(reshape(w, 1, nt), length(sx), 1)
    d = _desc(FOMO_ELASTIC3D, nx, ny, nz, nbc, fd_order, nt, dh, dt, v_ref, length(sx), length(rx);
              storage=storage, snap_interval=snap_interval, snap_plane=snap_plane, snap_index=snap_index)
    nrec = length(rx)
    seis_vx, seis_vy, seis_vz = (zeros(Float32, nrec, nt) for _ in 1:3)
    a, b = _slice_dims(snap_plane, nx, ny, nz)
    nsnap = snap_interval > 0 ? nt ÷ snap_interval : 0
    sb = _snapbuf(a * b, nsnap)
    t = Ref{Float64}(0.0)
    @info "Starting elastic3d simulation... (nx=$nx, ny=$ny, nz=$nz, nt=$nt)"
    GC.@preserve sb begin
        _check(ccall((:fomo_elastic3d, LIB), Cint,
                     (Ref{FomoDesc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32},
                      Ptr{Float32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
                      Ptr{Float32}, Ref{Float64}),
                     d, vp, vs, rho, _i32(sx), _i32(sy), _i32(sz), wm, _i32(rx), _i32(ry), _i32(rz), seis_vx, seis_vy,
                     seis_vz, _snapptr(sb), t))
    end
    @info "Simulation complete! GPU time: $(round(t[], digits=3))s"
    snaps = Matrix{Float32}[reshape(sb[:, k], a, b) for k in 1:nsnap]
    return seis_vx, seis_vy, seis_vz, snaps
end

# FP64 per-trace misfit of two device-resident plans (fomo_trace_misfit): num[r] = sum (a - b)^2 and
# den[r] = sum b^2 over samples t0 <= t < t1 of trace slot `slot` (the reflection measure of
# test/verify_patch.jl:86-87 without downloading the traces).
function trace_misfit(plan_a::Ptr{Cvoid}, plan_b::Ptr{Cvoid}, nrec::Int, slot::Int, t0::Int, t1::Int)
    num = zeros(Float64, nrec)
    den = zeros(Float64, nrec)
    _check(ccall((:fomo_trace_misfit, LIB), Cint,
                 (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Int32, Int32, Int32, Int32, Ptr{Float64}, Ptr{Float64}),
                 plan_a, Int32(0), plan_b, Int32(0), Int32(slot), Int32(t0), Int32(t1), num, den))
    return num, den
end
