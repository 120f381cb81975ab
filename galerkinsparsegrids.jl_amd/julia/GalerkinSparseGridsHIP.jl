# GalerkinSparseGridsHIP.jl -- Julia binding of libgsg (include/gsg.h).
#
# A drop-in for the operator path of GalerkinSparseGrids.jl: the matrices returned by
# D_matrix / grad_matrix / laplacian_matrix (multidim_derivative.jl:61-79) and the
# wave RHS of wave_data (pdes.jl:22-49) become `HIPOperator`s, and wave_evolve /
# transport_evolve integrate on the GPU with fixed-step RK4.  Coefficient vectors keep
# the package's vhier layout (dg_vmethods.jl:123-142), so coeffs_DG / vcoeffs_DG /
# tensor_construct output is passed in unchanged.
#
# Load with:  ENV["GSG_LIBRARY"] = "/path/to/libgsg.so"; include("GalerkinSparseGridsHIP.jl")
module GalerkinSparseGridsHIP

using LinearAlgebra
using SparseArrays

const libgsg = get(ENV, "GSG_LIBRARY", joinpath(@__DIR__, "..", "libgsg.so"))

const GSG_SCHEME = Dict("sparse" => Int32(0), "full" => Int32(1))
const GSG_OP_DERIV, GSG_OP_GRAD_SUM, GSG_OP_LAPLACIAN, GSG_OP_WAVE, GSG_OP_TRANSPORT = Int32.(0:4)
const GSG_PREC = Dict("fp64" => Int32(0), "fp32" => Int32(1))

# gsg_op_desc
struct OpDesc
    D::Int32; k::Int32; n::Int32; scheme::Int32; kind::Int32; dir::Int32
    precision::Int32; reserved::Int32; coef::NTuple{8,Float64}
end

function check(rc)
    rc < 0 && throw(ErrorException("libgsg error $rc: " * unsafe_string(ccall((:gsg_last_error, libgsg), Cstring, ()))))
    return rc
end

mutable struct HIPOperator <: AbstractMatrix{Float64}
    h::Ptr{Cvoid}
    len::Int
    function HIPOperator(h::Ptr{Cvoid})
        len = Int(ccall((:gsg_op_len, libgsg), Int64, (Ptr{Cvoid},), h))
        A = new(h, len)
        finalizer(A) do a
            a.h != C_NULL && ccall((:gsg_op_destroy, libgsg), Cint, (Ptr{Cvoid},), a.h)
            a.h = C_NULL
        end
        return A
    end
end

function HIPOperator(desc::OpDesc)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:gsg_op_create, libgsg), Cint, (Ref{OpDesc}, Ref{Ptr{Cvoid}}), desc, out))
    return HIPOperator(out[])
end

Base.size(A::HIPOperator) = (A.len, A.len)
Base.getindex(::HIPOperator, ::Int, ::Int) = throw(ArgumentError("HIPOperator is matrix-free"))

# `A*x`, the product of the reference's time loop (pdes.jl:63,65,179-191)
function Base.:*(A::HIPOperator, x::AbstractVector{Float64})
    xv = Vector{Float64}(x)
    y = similar(xv)
    GC.@preserve xv y check(ccall((:gsg_op_apply, libgsg), Cint,
                                  (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64), A.h, xv, y, A.len))
    return y
end

desc(D, k, n, kind; dir=0, coef=zeros(D), scheme="sparse", precision="fp64") =
    OpDesc(D, k, n, GSG_SCHEME[scheme], kind, dir, GSG_PREC[precision], 0,
           ntuple(i -> i <= length(coef) ? Float64(coef[i]) : 0.0, 8))

D_matrix(D::Int, d::Int, k::Int, n::Int; scheme="sparse", precision="fp64") =
    HIPOperator(desc(D, k, n, GSG_OP_DERIV; dir=d, scheme=scheme, precision=precision))
grad_matrix(D::Int, k::Int, n::Int; kw...) = [D_matrix(D, d, k, n; kw...) for d in 1:D]
laplacian_matrix(D::Int, k::Int, n::Int; scheme="sparse", precision="fp64") =
    HIPOperator(desc(D, k, n, GSG_OP_LAPLACIAN; scheme=scheme, precision=precision))

# fixed-step RK4 on the device; returns (tout, yout) like ODE.jl
function rk4(A::HIPOperator, y0::Vector{Float64}, t0, t1, nsteps; save_every=0)
    y = copy(y0)
    dt = (t1 - t0) / nsteps
    if save_every > 0
        nsnap = nsteps ÷ save_every
        snaps = Vector{Float64}(undef, nsnap * A.len)
        GC.@preserve y snaps check(ccall((:gsg_op_rk4, libgsg), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Int64, Float64, Float64, Int64, Int64, Ptr{Float64}),
            A.h, y, A.len, t0, dt, nsteps, save_every, snaps))
        tout = [t0; [t0 + i * save_every * dt for i in 1:nsnap]]
        yout = [[y0]; [snaps[(i-1)*A.len+1:i*A.len] for i in 1:nsnap]]
        return tout, yout
    end
    GC.@preserve y check(ccall((:gsg_op_rk4, libgsg), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Float64, Float64, Int64, Int64, Ptr{Float64}),
        A.h, y, A.len, t0, dt, nsteps, 0, C_NULL))
    return [t0, t1], [y0, y]
end

# pdes.jl:54-70 with order="rk4"
function wave_evolve(D::Int, k::Int, n::Int, f0coeffs::Vector{Float64}, v0coeffs::Vector{Float64},
                     time0::Real, time1::Real; nsteps::Int, order="rk4", scheme="sparse",
                     precision="fp64", save_every=0)
    order == "rk4" || throw(ArgumentError(:order))
    A = HIPOperator(desc(D, k, n, GSG_OP_WAVE; scheme=scheme, precision=precision))
    return rk4(A, [f0coeffs; v0coeffs], time0, time1, nsteps; save_every=save_every)
end

# pdes.jl:74-83: the initial position and velocity profiles as functions, projected with
# the package's own vcoeffs_DG (dg_vmethods.jl:149-179), then the vector form above
function wave_evolve(D::Int, k::Int, n::Int, f0::Function, v0::Function, time0::Real, time1::Real;
                     nsteps::Int, scheme="sparse", kwargs...)
    f0coeffs = Main.GalerkinSparseGrids.vcoeffs_DG(D, k, n, f0; scheme=scheme)
    v0coeffs = Main.GalerkinSparseGrids.vcoeffs_DG(D, k, n, v0; scheme=scheme)
    return wave_evolve(D, k, n, f0coeffs, v0coeffs, time0, time1; nsteps=nsteps, scheme=scheme, kwargs...)
end

# u_t + a . grad u = 0 (the composition of pdes.jl:179,186,191 with constant a)
function transport_evolve(D::Int, k::Int, n::Int, u0::Vector{Float64}, a::Vector{Float64},
                          time0::Real, time1::Real; nsteps::Int, scheme="sparse",
                          precision="fp64", save_every=0)
    A = HIPOperator(desc(D, k, n, GSG_OP_TRANSPORT; coef=a, scheme=scheme, precision=precision))
    return rk4(A, u0, time0, time1, nsteps; save_every=save_every)
end

# Any SparseMatrixCSC the Julia code builds (e.g. the Vlasov transforms written by
# pdes.jl:149-155) can be applied on the device as well.
function csc_operator(A; precision="fp64")
    out = Ref{Ptr{Cvoid}}(C_NULL)
    colptr = Vector{Int64}(A.colptr); rowval = Vector{Int64}(A.rowval); nzval = Vector{Float64}(A.nzval)
    GC.@preserve colptr rowval nzval check(ccall((:gsg_csc_create, libgsg), Cint,
        (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Int64, Int32, Ref{Ptr{Cvoid}}),
        A.m, A.n, colptr, rowval, nzval, length(nzval), GSG_PREC[precision], out))
    return HIPOperator(out[])
end

# ---- Vlasov path (1d_nodal_basis.jl, multidim_nodal_basis.jl, pdes.jl:131-227) ----
const GSG_BASIS = Dict("modal" => Int32(0), "nodal" => Int32(1), "points" => Int32(2))

# transform_1D(k, n, from, to) (1d_nodal_basis.jl:383-414), dense; k in {2, 3, 5}
function transform_1D(k::Int, n::Int, from::String, to::String)
    N1 = k << n
    out = zeros(Float64, N1, N1)   # row-major from the library: transpose below
    GC.@preserve out check(ccall((:gsg_transform_1d, libgsg), Cint,
        (Int32, Int32, Int32, Int32, Ptr{Float64}, Int64), k, n, GSG_BASIS[from], GSG_BASIS[to], out, N1 * N1))
    return sparse(permutedims(out))
end

# transform(D, k, n, from, to; scheme) (multidim_nodal_basis.jl:111-122) as a device
# operator: `T * x` replaces the restricted-Kronecker SparseMatrixCSC product
function transform(D::Int, k::Int, n::Int, from::String, to::String; scheme="sparse")
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:gsg_transform_create, libgsg), Cint,
        (Int32, Int32, Int32, Int32, Int32, Int32, Ref{Ptr{Cvoid}}),
        D, k, n, scheme == "full" ? 1 : 0, GSG_BASIS[from], GSG_BASIS[to], out))
    return HIPOperator(out[])
end
# The same transform as a SparseMatrixCSC (gsg_transform_csc: the matrix the device
# operator applies), for callers that write the operators out as pdes.jl:145-163 does
function transform_csc(D::Int, k::Int, n::Int, from::String, to::String; scheme="sparse")
    sch = scheme == "full" ? Int32(1) : Int32(0)
    nnz = Ref{Int64}(0)
    check(ccall((:gsg_transform_csc, libgsg), Cint,
        (Int32, Int32, Int32, Int32, Int32, Int32, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ref{Int64}),
        D, k, n, sch, GSG_BASIS[from], GSG_BASIS[to], C_NULL, C_NULL, C_NULL, nnz))
    N = Int(check(ccall((:gsg_get_size, libgsg), Int64, (Int32, Int32, Int32, Int32), D, k, n, sch)))
    colptr = Vector{Int64}(undef, N + 1); rowval = Vector{Int64}(undef, nnz[]); nzval = Vector{Float64}(undef, nnz[])
    check(ccall((:gsg_transform_csc, libgsg), Cint,
        (Int32, Int32, Int32, Int32, Int32, Int32, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ref{Int64}),
        D, k, n, sch, GSG_BASIS[from], GSG_BASIS[to], colptr, rowval, nzval, nnz))
    return SparseMatrixCSC(N, N, colptr, rowval, nzval)   # 1-based already
end

# D_matrix(D, d, k, n; scheme) (multidim_derivative.jl:61-65) as the reference returns it
function D_matrix_csc(D::Int, d::Int, k::Int, n::Int; scheme="sparse")
    sch = scheme == "full" ? Int32(1) : Int32(0)
    nnz = Ref{Int64}(0)
    check(ccall((:gsg_d_matrix_csc, libgsg), Cint,
        (Int32, Int32, Int32, Int32, Int32, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ref{Int64}),
        D, d, k, n, sch, C_NULL, C_NULL, C_NULL, nnz))
    N = Int(check(ccall((:gsg_get_size, libgsg), Int64, (Int32, Int32, Int32, Int32), D, k, n, sch)))
    colptr = Vector{Int64}(undef, N + 1); rowval = Vector{Int64}(undef, nnz[]); nzval = Vector{Float64}(undef, nnz[])
    check(ccall((:gsg_d_matrix_csc, libgsg), Cint,
        (Int32, Int32, Int32, Int32, Int32, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ref{Int64}),
        D, d, k, n, sch, colptr, rowval, nzval, nnz))
    return SparseMatrixCSC(N, N, colptr, rowval, nzval)
end

make_modal2point_matrices(D, k, n) = (transform(D, k, n, "modal", "nodal"), transform(D, k, n, "nodal", "points"))
make_point2modal_matrices(D, k, n) = (transform(D, k, n, "points", "nodal"), transform(D, k, n, "nodal", "modal"))

struct VlasovDesc        # gsg_vlasov_desc
    Dx::Int32; k::Int32; n::Int32; scheme::Int32; precision::Int32; reserved::Int32
end

# vlasov_evolve(D, k, n, m2n, n2p, p2n, n2m, f0_modal, F_point, time0, time1, nout)
# (pdes.jl:131-227): the steprule runs on the device; fixed-step RK4 with nsteps steps
# between outputs.  The transform arguments are accepted for signature compatibility
# (the operator builds its own).  To write vlasov.h5 as the reference does, take the
# matrices from transform_csc and D_matrix_csc.
function vlasov_evolve(D::Int, k::Int, n::Int, m2n, n2p, p2n, n2m, f0_modal::Vector{Float64},
                       F_point::Vector{Vector{Float64}}, time0::Real, time1::Real, nout::Int=2;
                       nsteps::Int, scheme="sparse", precision="fp64")
    out = Ref{Ptr{Cvoid}}(C_NULL)
    d = VlasovDesc(D, k, n, scheme == "full" ? 1 : 0, GSG_PREC[precision], 0)
    F = isempty(F_point) ? Float64[] : reduce(vcat, F_point)
    GC.@preserve F check(ccall((:gsg_vlasov_create, libgsg), Cint,
        (Ref{VlasovDesc}, Ptr{Float64}, Int64, Ref{Ptr{Cvoid}}),
        d, isempty(F) ? C_NULL : pointer(F), length(F), out))
    A = HIPOperator(out[])
    return rk4(A, f0_modal, time0, time1, nsteps * (nout - 1); save_every=nsteps)
end

export HIPOperator, csc_operator, D_matrix, grad_matrix, laplacian_matrix, wave_evolve, transport_evolve, rk4,
       transform_1D, transform, transform_csc, D_matrix_csc, make_modal2point_matrices, make_point2modal_matrices,
       vlasov_evolve

end # module
