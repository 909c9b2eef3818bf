# HOMBackend.jl -- Julia ccall shim over libhom.so (include/hom_abi.h).
#
# Included from HighOrderMRISim's module root (src/HighOrderMRISim.jl:21-41),
# it replaces the CUDA kernel internals of src/simulation/kernels/*.jl behind
# the unchanged simulate(obj, hoseq, sys; sim_params) API:
#
#   HighOrderGPUWorkspace(...)            -> HOMWorkspace(...)        (HighOrderGPU.jl:16-58)
#   run_high_order_precession!(ws, ...)   -> hom_precession_block     (HighOrderGPU.jl:124-140)
#   run_high_order_excitation!(ws, ...)   -> hom_excitation_block     (HighOrderGPU.jl:223-233)
#   run_sim_time_iter! without callbacks  -> hom_run                  (SimulatorCore.jl:96-152)
#
# sim_params keep their meaning ("precision" => "f32" | "f64",
# "coil_combination", "return_type", "max_block_length", ...).  On this
# backend "f32" runs the mixed-precision kernels (FP64 phase anchors,
# split-FP16 MFMA coil GEMM, FP64 signal; NRMSE ~1e-6 against the f64 CPU
# path, inside the reference's f32 tolerance of 3e-5) and returns
# Complex{Float32}; "hip_precision" => "f32" selects the literal f32 kernels
# instead, and "mixed" (accepted once setup_simulation lets it through)
# returns Complex{Float64}.  Host arrays are passed as plain pointers in
# Julia's own layouts (column-major, interleaved Complex{Float64}); BitVector
# flags are widened to UInt8 first.
module HOMBackend

export HOMWorkspace, hom_simulate_shard, hom_available, hom_get_kspace, HOMServer,
       hom_put_phantom!, hom_put_sequence!, hom_submit!, hom_result!

const libhom = get(ENV, "HOM_LIBRARY",
    joinpath(@__DIR__, "..", "libhom.so"))

const HOM_OK = Cint(0)
const PRECISIONS = Dict("f32" => Cint(0), "f64" => Cint(1), "mixed" => Cint(2))
const COIL_MODES = Dict("gemm" => Cint(0), "reduction" => Cint(1),
                        "fused_warp" => Cint(2), "fused_tiled" => Cint(3))

struct HOMSequenceHost
    n_nodes::Int64
    t::Ptr{Float64}
    dt::Ptr{Float64}
    gx::Ptr{Float64}
    gy::Ptr{Float64}
    gz::Ptr{Float64}
    b1::Ptr{Float64}
    df::Ptr{Float64}
    psi::Ptr{Float64}
    adc::Ptr{UInt8}
    excitation::Ptr{UInt8}
    h::NTuple{16,Ptr{Float64}}
end

hom_available() = isfile(libhom)

last_error(plan::Ptr{Cvoid}) = unsafe_string(ccall((:hom_last_error, libhom), Cstring, (Ptr{Cvoid},), plan))

# Status -> the reference's exception types (src/simulation/Simulate.jl:70-135 conventions).
function check(status::Cint, plan::Ptr{Cvoid})
    status == HOM_OK && return nothing
    msg = last_error(plan)
    status == Cint(1) && throw(ArgumentError(msg))
    status == Cint(2) && throw(DimensionMismatch(msg))
    error(msg)
end

mutable struct HOMWorkspace
    handle::Ptr{Cvoid}
    precision::String
    nspins::Int
    ncoils::Int
    nnodes::Int
    keep::Vector{Any}          # host arrays pinned for the duration of a call
    function HOMWorkspace(handle, precision, nspins, ncoils, nnodes)
        ws = new(handle, precision, nspins, ncoils, nnodes, Any[])
        finalizer(close!, ws)
        return ws
    end
end

function close!(ws::HOMWorkspace)
    if ws.handle != C_NULL
        ccall((:hom_plan_destroy, libhom), Cint, (Ptr{Cvoid},), ws.handle)
        ws.handle = C_NULL
    end
    return nothing
end

f64vec(x) = Vector{Float64}(x)
u8vec(x) = Vector{UInt8}(collect(x))

"""
    HOMWorkspace(obj, hoseqd, max_adc; gamma, precision="f64", coil_combination="gemm",
                 max_rf_adc=max_adc, groupsize=256, adc_tile=8, coil_tile=1, device=0)

Device-resident plan for one phantom and one sampled sequence (`gamma`, the
gyromagnetic ratio in Hz/T, is required: the caller passes the package's γ):
the argument
checks of HighOrderGPUWorkspace (HighOrderGPU.jl:19-35, including the fused
groupsize / tile / shared-memory checks through hom_check_workspace_options)
before any device work, then the upload of the phantom, CSM and sequence
(Simulate.jl:151-157).
"""
function HOMWorkspace(obj, hoseqd, max_adc; precision="f64", coil_combination="gemm",
                      max_rf_adc=max_adc, groupsize=256, adc_tile=8, coil_tile=1, device=0, gamma)
    max_adc isa Integer && max_adc >= 0 ||
        throw(ArgumentError("The maximum ADC count must be a nonnegative integer."))
    max_rf_adc isa Integer && 0 <= max_rf_adc <= max_adc ||
        throw(ArgumentError("The RF ADC count must be an integer between zero and max_adc."))
    haskey(PRECISIONS, precision) || throw(ArgumentError("The precision must be f32, f64 or mixed."))
    haskey(COIL_MODES, coil_combination) ||
        throw(ArgumentError("Unknown coil_combination: $coil_combination."))
    groupsize isa Integer && adc_tile isa Integer && coil_tile isa Integer ||
        throw(ArgumentError("Fused ADC/coil tiles must be selected from 1/2/4/8 and 1/2/4."))
    check(ccall((:hom_check_workspace_options, libhom), Cint,
                (Cint, Cint, Cint, Int64, Int64, Int64), device, PRECISIONS[precision],
                COIL_MODES[coil_combination], groupsize, adc_tile, coil_tile), C_NULL)
    seqd = hoseqd.seqd
    ns, nc, nn = length(obj), size(obj.csm, 2), length(seqd.t)
    handle = Ref{Ptr{Cvoid}}(C_NULL)
    st = ccall((:hom_plan_create, libhom), Cint,
        (Ptr{Ptr{Cvoid}}, Cint, Cint, Int64, Int64, Int64, Int64, Int64, Float64, Cint),
        handle, device, PRECISIONS[precision], ns, nc, nn, max_adc, max_rf_adc, Float64(gamma),
        COIL_MODES[coil_combination])
    check(st, C_NULL)
    ws = HOMWorkspace(handle[], precision, ns, nc, nn)
    upload_phantom!(ws, obj)
    upload_sequence!(ws, hoseqd)
    return ws
end

function upload_phantom!(ws::HOMWorkspace, obj)
    x, y, z, ρ, T1, T2, Δw = map(f64vec, (obj.x, obj.y, obj.z, obj.ρ, obj.T1, obj.T2, obj.Δw))
    check(ccall((:hom_upload_phantom, libhom), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}), ws.handle, x, y, z, ρ, T1, T2, Δw), ws.handle)
    csm = Matrix{ComplexF64}(obj.csm)     # dense column-major copy of a possibly strided view
    check(ccall((:hom_upload_csm, libhom), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Int64),
        ws.handle, csm, max(size(csm, 1), 1)), ws.handle)
    return ws
end

function upload_sequence!(ws::HOMWorkspace, hoseqd)
    s = hoseqd.seqd
    arrays = (f64vec(s.t), f64vec(s.Δt), f64vec(s.Gx), f64vec(s.Gy), f64vec(s.Gz),
              Vector{ComplexF64}(s.B1), f64vec(s.Δf), f64vec(s.ψ))
    adc, exc = u8vec(s.ADC), u8vec(s.excitation_bool)
    fields = ntuple(j -> f64vec(getfield(hoseqd, j + 1)), 16)
    GC.@preserve arrays adc exc fields begin
        host = HOMSequenceHost(length(s.t), pointer(arrays[1]), pointer(arrays[2]),
            pointer(arrays[3]), pointer(arrays[4]), pointer(arrays[5]),
            Ptr{Float64}(pointer(arrays[6])), pointer(arrays[7]), pointer(arrays[8]),
            pointer(adc), pointer(exc), ntuple(j -> pointer(fields[j]), 16))
        check(ccall((:hom_upload_sequence, libhom), Cint, (Ptr{Cvoid}, Ref{HOMSequenceHost}),
            ws.handle, host), ws.handle)
    end
    return ws
end

method_flags(m) = Cint((m.ho0 ? 1 : 0) | (m.ho1 ? 2 : 0) | (m.ho2 ? 4 : 0) | (m.ho3 ? 8 : 0) |
                       (m.Δw_excitation ? 16 : 0) | (m.Δw_precession ? 32 : 0))

function set_state!(ws::HOMWorkspace, M)
    xy, z = Vector{ComplexF64}(M.xy), f64vec(M.z)
    check(ccall((:hom_set_state, libhom), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Float64}),
        ws.handle, xy, z), ws.handle)
end

function get_state!(M, ws::HOMWorkspace)
    xy, z = zeros(ComplexF64, ws.nspins), zeros(Float64, ws.nspins)
    check(ccall((:hom_get_state, libhom), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Float64}),
        ws.handle, xy, z), ws.handle)
    M.xy .= xy
    M.z .= z
    return M
end

# Kernel precision behind a requested sim_params["precision"] (see the
# module header): "f32" -> mixed unless "hip_precision" => "f32".
function kernel_precision(params)
    requested = get(params, "precision", "f32")
    requested == "f32" && return get(params, "hip_precision", "mixed") == "f32" ? "f32" : "mixed"
    return requested
end

"""
    hom_simulate_shard(obj, hoseqd, params, plan; callbacks=(), gamma, M)

Drop-in for simulate_shard (Simulate.jl:138-193) on one MI355X: one device
plan, the whole block loop through hom_run when there are no callbacks
(mixed precision fuses precession blocks), per-block entry points with state
round trips when callbacks must observe each block (SimulatorCore.jl:109-111).
M is the initial state (initialize_spins_state).  Returns simulate_shard's
(; signal, state, params, device, seconds): the signal as (sum(ADC), Nc, 1)
of the requested precision's complex type, seconds the wall time of the
block loop (Simulate.jl:181-185 times the same span).
"""
function hom_simulate_shard(obj, hoseqd, params, plan; callbacks=(), gamma, M)
    (; parts, excitation_bool, adc_counts, max_adc, max_rf_adc) = plan
    precision = kernel_precision(params)
    ws = HOMWorkspace(obj, hoseqd, max_adc; precision, max_rf_adc, gamma,
                      coil_combination=get(params, "coil_combination", "gemm"),
                      device=something(get(params, "gpu_device", 0), 0))
    dev = Ref{Ptr{Cvoid}}(C_NULL)
    try
        set_state!(ws, M)
        flags = method_flags(params["sim_method"])
        pf = Int64[first(p) - 1 for p in parts]           # 0-based node ranges
        pl = Int64[last(p) - 1 for p in parts]
        rf = UInt8.(excitation_bool)
        T = precision == "f32" ? ComplexF32 : ComplexF64   # the plan's signal type
        nc = size(obj.csm, 2)
        total = sum(adc_counts)
        t0 = time_ns()
        # the reference's fence policy (SimulatorCore.jl:107-111): per-block
        # steps when callbacks must see each block or synchronization="block"
        if isempty(callbacks) && get(params, "synchronization", "auto") != "block"
            check(ccall((:hom_run, libhom), Cint,
                (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{UInt8}, Int64, Cint),
                ws.handle, pf, pl, rf, length(pf), flags), ws.handle)
            sig = zeros(T, total, nc)
            check(ccall((:hom_download_signal, libhom), Cint, (Ptr{Cvoid}, Ptr{T}, Int64),
                ws.handle, sig, max(total, 1)), ws.handle)
        else
            sig = zeros(T, total, nc)
            # one block's rows (at most max_adc x Nc, column stride = its ADC
            # count) go to a device buffer and are copied into sig's rows
            check(ccall((:hom_buffer_alloc, libhom), Cint, (Ptr{Cvoid}, Int64, Ptr{Ptr{Cvoid}}),
                ws.handle, sizeof(T) * max(max_adc, 1) * nc, dev), ws.handle)
            row = 0
            rfs = 0
            for b in eachindex(parts)
                n = adc_counts[b]
                out = n == 0 ? C_NULL : dev[]
                st = excitation_bool[b] ?
                    ccall((:hom_excitation_block, libhom), Cint,
                        (Ptr{Cvoid}, Int64, Int64, Cint, Ptr{Cvoid}, Int64),
                        ws.handle, pf[b], pl[b], flags, out, max(n, 1)) :
                    ccall((:hom_precession_block, libhom), Cint,
                        (Ptr{Cvoid}, Int64, Int64, Cint, Ptr{Cvoid}, Int64),
                        ws.handle, pf[b], pl[b], flags, out, max(n, 1))
                check(st, ws.handle)
                if n > 0
                    rows = Matrix{T}(undef, n, nc)
                    check(ccall((:hom_buffer_download, libhom), Cint,
                        (Ptr{Cvoid}, Ptr{T}, Ptr{Cvoid}, Int64), ws.handle, rows, dev[],
                        sizeof(T) * n * nc), ws.handle)
                    sig[row+1:row+n, :] .= rows
                end
                row += n
                rfs += excitation_bool[b] ? 1 : 0
                get_state!(M, ws)
                progress = (; block=b, Nblocks=length(parts), rfs, samples=row + 1)
                for cb in callbacks
                    cb(progress, (; excitation_bool, parts), (; obj, seqd=hoseqd, sig, Xt=M), params)
                end
                set_state!(ws, M)           # callbacks may edit the state
            end
        end
        get_state!(M, ws)
        seconds = (time_ns() - t0) / 1e9
        out = get(params, "precision", "f32") == "f32" ? ComplexF32.(sig) : sig
        params["implementation"] = "kernel"
        params["adc_state_buffer_samples"] =
            (precision == "mixed" || get(params, "coil_combination", "gemm") in ("fused_warp", "fused_tiled")) ?
            max_rf_adc : max_adc
        return (; signal=reshape(out, total, nc, 1), state=M, params, device="MI355X (HIP)", seconds)
    finally
        if dev[] != C_NULL
            ccall((:hom_buffer_free, libhom), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ws.handle, dev[])
        end
        close!(ws)
    end
end

# Host view of a discretized sequence for the duration of `f(host)`.
function with_sequence_host(f, hoseqd)
    s = hoseqd.seqd
    arrays = (f64vec(s.t), f64vec(s.Δt), f64vec(s.Gx), f64vec(s.Gy), f64vec(s.Gz),
              Vector{ComplexF64}(s.B1), f64vec(s.Δf), f64vec(s.ψ))
    adc, exc = u8vec(s.ADC), u8vec(s.excitation_bool)
    fields = ntuple(j -> f64vec(getfield(hoseqd, j + 1)), 16)
    GC.@preserve arrays adc exc fields begin
        host = HOMSequenceHost(length(s.t), pointer(arrays[1]), pointer(arrays[2]),
            pointer(arrays[3]), pointer(arrays[4]), pointer(arrays[5]),
            Ptr{Float64}(pointer(arrays[6])), pointer(arrays[7]), pointer(arrays[8]),
            pointer(adc), pointer(exc), ntuple(j -> pointer(fields[j]), 16))
        return f(host)
    end
end

"""
    hom_get_kspace(hoseqd; gamma, skip_rf=nothing)

`get_kspace` on an already discretized sequence through `hom_kspace`
(reference src/datatypes/HO_Sequence.jl:221-287): returns
`(K_nominal, K_nominal_adc, K_dfc, K_dfc_adc)`.
"""
function hom_get_kspace(hoseqd; gamma, skip_rf=nothing)
    n = length(hoseqd.seqd.t)
    nadc = count(hoseqd.seqd.ADC)
    kn, kna = zeros(n - 1, 3), zeros(nadc, 3)
    kd, kda = zeros(n - 1, 16), zeros(nadc, 16)
    skip = skip_rf === nothing ? UInt8[] : UInt8.(collect(skip_rf))
    st = with_sequence_host(hoseqd) do host
        ccall((:hom_kspace, libhom), Cint,
              (Ref{HOMSequenceHost}, Float64, Ptr{UInt8}, Int64, Ptr{Float64}, Ptr{Float64},
               Ptr{Float64}, Ptr{Float64}),
              host, gamma, skip_rf === nothing ? C_NULL : pointer(skip), length(skip),
              kn, kna, kd, kda)
    end
    st == Cint(2) && throw(DimensionMismatch("skip_rf must contain one flag per RF event."))
    st == HOM_OK || throw(ArgumentError("Cannot integrate an empty sequence."))
    return kn, kna, kd, kda
end

# ---- continuous-batching server (include/hom_serve.h) ----

struct HOMServerConfig
    device::Cint
    device_bytes::Int64
    host_bytes::Int64
    max_active::Cint
    step_budget::Int64
    use_graphs::Cint
    gpt::Cint
    main_grid::Cint
    threaded::Cint
end

struct HOMRequest
    phantom::Int64
    spin_first::Int64
    spin_count::Int64
    csm::Int64
    coil_count::Int64
    sequence::Int64
    node_first::Int64
    node_last::Int64
    flags::Cint
end

mutable struct HOMServer
    handle::Ptr{Cvoid}
end

server_error(h::Ptr{Cvoid}) =
    unsafe_string(ccall((:hom_server_last_error, libhom), Cstring, (Ptr{Cvoid},), h))

function server_check(status::Cint, h::Ptr{Cvoid})
    status == HOM_OK && return nothing
    msg = server_error(h)
    status == Cint(1) && throw(ArgumentError(msg))
    status == Cint(2) && throw(DimensionMismatch(msg))
    error(msg)
end

function HOMServer(; device=0, device_bytes=8 << 30, host_bytes=1 << 30, max_active=64,
                   step_budget=0, use_graphs=true, threaded=true)
    cfg = HOMServerConfig(device, device_bytes, host_bytes, max_active, step_budget,
                          use_graphs, 0, 0, threaded)
    ref = Ref{Ptr{Cvoid}}(C_NULL)
    server_check(ccall((:hom_server_create, libhom), Cint,
                       (Ref{Ptr{Cvoid}}, Ref{HOMServerConfig}), ref, cfg), C_NULL)
    srv = HOMServer(ref[])
    finalizer(srv) do s
        s.handle == C_NULL || ccall((:hom_server_destroy, libhom), Cint, (Ptr{Cvoid},), s.handle)
        s.handle = C_NULL
    end
    return srv
end

"Uploads an HO_Phantom and its CSM; returns (phantom id, csm id)."
function hom_put_phantom!(srv::HOMServer, obj)
    id, cid = Ref{Int64}(0), Ref{Int64}(0)
    a = map(f64vec, (obj.x, obj.y, obj.z, obj.ρ, obj.T1, obj.T2, obj.Δw))
    server_check(ccall((:hom_server_put_phantom, libhom), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{Int64}),
        srv.handle, length(a[1]), a..., id), srv.handle)
    csm = Matrix{ComplexF64}(obj.csm)
    server_check(ccall((:hom_server_put_csm, libhom), Cint,
        (Ptr{Cvoid}, Int64, Int64, Ptr{ComplexF64}, Int64, Ref{Int64}),
        srv.handle, id[], size(csm, 2), csm, size(csm, 1), cid), srv.handle)
    return id[], cid[]
end

function hom_put_sequence!(srv::HOMServer, hoseqd; gamma)
    id = Ref{Int64}(0)
    st = with_sequence_host(hoseqd) do host
        ccall((:hom_server_put_sequence, libhom), Cint,
              (Ptr{Cvoid}, Ref{HOMSequenceHost}, Float64, Ref{Int64}), srv.handle, host, gamma, id)
    end
    server_check(st, srv.handle)
    return id[]
end

"Queues a request; returns its id (`flags` as for hom_run)."
function hom_submit!(srv::HOMServer, phantom, csm, sequence; spins, coils, nodes, flags)
    req = HOMRequest(phantom, first(spins) - 1, length(spins), csm, coils, sequence,
                     first(nodes) - 1, last(nodes) - 1, flags)
    rid = Ref{Int64}(0)
    server_check(ccall((:hom_server_submit, libhom), Cint,
                       (Ptr{Cvoid}, Ref{HOMRequest}, Ref{Int64}), srv.handle, req, rid), srv.handle)
    return rid[]
end

"Waits for a request and returns its (ΣADC, Nc) signal."
function hom_result!(srv::HOMServer, rid)
    state = Ref{Cint}(0)
    server_check(ccall((:hom_server_wait, libhom), Cint, (Ptr{Cvoid}, Int64, Float64, Ref{Cint}),
                       srv.handle, rid, -1.0, state), srv.handle)
    rows, cols = Ref{Int64}(0), Ref{Int64}(0)
    server_check(ccall((:hom_server_result_shape, libhom), Cint,
                       (Ptr{Cvoid}, Int64, Ref{Int64}, Ref{Int64}), srv.handle, rid, rows, cols),
                 srv.handle)
    sig = zeros(ComplexF64, rows[], cols[])
    server_check(ccall((:hom_server_result, libhom), Cint,
                       (Ptr{Cvoid}, Int64, Ptr{ComplexF64}, Int64), srv.handle, rid, sig,
                       max(rows[], 1)), srv.handle)
    return sig
end

end # module
