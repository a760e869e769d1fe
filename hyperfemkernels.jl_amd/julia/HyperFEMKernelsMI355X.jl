# Julia side of the `backend = :mi355x` drop-in: thin `ccall` bindings of libhfk (include/hfk.h) and the methods
# HyperFEMKernels.jl would route to them.  Not exercised in this repository (no Julia toolchain here); the same
# C ABI is exercised by the Python mirror (hyperfemkernels.jl_amd/hfk) and its tests.
#
# Conventions: Gridap's cell dof ids are passed as they are (Int32, > 0 free, < 0 Dirichlet); table indices
# (ldof, nodeperm) are passed 0-based (the 1-based Julia arrays minus one); CSC arrays come back 1-based, ready
# for SparseMatrixCSC.
module HyperFEMKernelsMI355X

using SparseArrays

const libhfk = get(ENV, "HFK_LIBHFK", joinpath(@__DIR__, "..", "hfk", "libhfk.so"))

const HFK_STORE = Dict(:full64 => Cint(0), :sym64 => Cint(1), :full32 => Cint(2), :sym32 => Cint(3), :fac64 => Cint(4),
                       :fac32 => Cint(5))
const HFK_LAW = Dict(:svk => Cint(0), :neohookean => Cint(1), :linear => Cint(2))

struct HFKError <: Exception
    status::Cint
    msg::String
end

function check(status::Cint, ctx::Ptr{Cvoid}=C_NULL)
    status == 0 && return nothing
    msg = unsafe_string(ccall((:hfk_last_error, libhfk), Cstring, (Ptr{Cvoid},), ctx))
    status == -1 && throw(ArgumentError(msg))
    status == -2 && throw(DimensionMismatch(msg))
    status == -3 && error(msg)
    throw(HFKError(status, msg))
end

mutable struct Context
    ptr::Ptr{Cvoid}
end
function Context(device::Integer=0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hfk_open, libhfk), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, r))
    c = Context(r[])
    finalizer(c -> ccall((:hfk_close, libhfk), Cint, (Ptr{Cvoid},), c.ptr), c)
    c
end

"Device operator of one volume group (what `KernelOperator(res, jac, U, V; backend = :mi355x)` builds)."
mutable struct DeviceOperator
    ptr::Ptr{Cvoid}
    ctx::Context
    nfree::Int
    law::Any      # a CompiledLaw in use (kept alive with the operator), or nothing
end

"""
    DeviceOperator(ctx, cell_dofs, ldof, nodeperm, coords, nfree, ndir; law, λ, μ, P)

`cell_dofs`: `Matrix{Int32}` nd × ncells from `cell_dof_matrix` (src/tables.jl:329-342); `ldof`: FieldTable.ldof
(nn × 3, 1-based); `nodeperm`: TPTable.nodeperm (1-based); `coords`: 3 × 8 × ncells vertices in tensor order.
"""
function DeviceOperator(ctx::Context, cell_dofs::Matrix{Int32}, ldof::Matrix{Int}, nodeperm::Vector{Int},
                        coords::Array{Float64,3}, nfree::Integer, ndir::Integer; law::Symbol=:neohookean,
                        λ::Real=2.0, μ::Real=1.0, P::Integer)
    nn = size(ldof, 1)
    Q = round(Int, nn^(1 / 3))
    ld = Int32[ldof[n, c] - 1 for n in 1:nn, c in 1:3]          # [c*nn + n] layout: column-major nn × 3
    np_ = Int32.(nodeperm .- 1)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hfk_op_create, libhfk), Cint,
                (Ptr{Cvoid}, Cint, Cint, Int64, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float64}, Int64, Int64, Cint,
                 Cdouble, Cdouble, Ptr{Ptr{Cvoid}}),
                ctx.ptr, Q, P, size(cell_dofs, 2), cell_dofs, ld, np_, coords, nfree, ndir, HFK_LAW[law], λ, μ, r), ctx.ptr)
    op = DeviceOperator(r[], ctx, nfree, nothing)
    finalizer(o -> ccall((:hfk_op_destroy, libhfk), Cint, (Ptr{Cvoid},), o.ptr), op)
    op
end

set_dirichlet!(op::DeviceOperator, d::Vector{Float64}) =
    check(ccall((:hfk_op_set_dirichlet, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float64}), op.ptr, d), op.ctx.ptr)

"The fast path of `update(op, res, jac)` for a form of the same structure (src/assembly.jl:693-744): λ and μ rebound, nothing rebuilt; prepare! again."
update!(op::DeviceOperator; λ::Real, μ::Real) =
    (check(ccall((:hfk_op_set_parameters, libhfk), Cint, (Ptr{Cvoid}, Cdouble, Cdouble), op.ptr, λ, μ), op.ctx.ptr); op)

"""
    set_traction!(op, face_cell, face_local, traction)

A dead traction on faces of the cells (the Neumann load `-Λ ∫ t⋅v dΓ` of a boundary group, src/assembly.jl:539-578):
`face_cell` 0-based cells of the faces of `BoundaryTriangulation(model, tags=...)`, `face_local` the local face 0..5
(ξ_d = s ↦ 2d + s), `traction` 3 × nfaces.  Enters residual! and the energy; the Jacobian does not change.
"""
set_traction!(op::DeviceOperator, face_cell::Vector{Int32}, face_local::Vector{Int32}, traction::Matrix{Float64}) =
    check(ccall((:hfk_op_set_traction, libhfk), Cint, (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Int32}, Ptr{Float64}), op.ptr,
                length(face_cell), face_cell, face_local, traction), op.ctx.ptr)
"""
    set_robin!(op, face_cell, face_local, k)

A Robin (spring) term `k ∫ u⋅v dΓ` on faces of the cells (faces as in `set_traction!`, `k` one stiffness per face):
in residual!, Jv, the assembled Jacobian, the node blocks of BlockJacobi and the energy; prepare! again afterwards.
"""
set_robin!(op::DeviceOperator, face_cell::Vector{Int32}, face_local::Vector{Int32}, k::Vector{Float64}) =
    check(ccall((:hfk_op_set_robin, libhfk), Cint, (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Int32}, Ptr{Float64}), op.ptr,
                length(face_cell), face_cell, face_local, k), op.ctx.ptr)
"""
    set_skeleton!(op, plus_cell, plus_face, minus_cell, minus_face, orient; source=nothing, name="", params=[1.0])

Skeleton integrals on the interior facets of a `SkeletonTriangulation` (src/skeleton.jl): per facet the 0-based plus
and minus cells (the `SkeletonPair` glue), their local faces 0..5 and the orientation code 0..7 of the minus face's
coordinates relative to the plus face's (the permutation the glue's minus map carries).  `source === nothing`: the
built-in gradient-jump penalty `(γ h/2)|∇u⁺ − ∇u⁻|²`, γ = params[1]; otherwise the HIP source of a facet point form
(struct `name` with static `residual`, `jvp`, `energy` of the 24 facet slots; the captured integrand emitted as HIP).
The sparsity pattern gains the cross blocks; rebuild the matrix after this call.
"""
set_skeleton!(op::DeviceOperator, plus_cell::Vector{Int32}, plus_face::Vector{Int32}, minus_cell::Vector{Int32},
              minus_face::Vector{Int32}, orient::Vector{Int32}; source::Union{Nothing,AbstractString}=nothing,
              name::AbstractString="", params::Vector{Float64}=[1.0]) =
    check(ccall((:hfk_op_set_skeleton, libhfk), Cint,
                (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Cstring, Cstring,
                 Ptr{Float64}, Cint), op.ptr, length(plus_cell), plus_cell, plus_face, minus_cell, minus_face, orient,
                source === nothing ? C_NULL : source, source === nothing ? C_NULL : name, params, length(params)),
          op.ctx.ptr)

"""
    set_boundary_form!(op, face_cell, face_local; source=nothing, name="", params=[p])

A boundary group with a captured integrand (src/assembly.jl:539-578) on faces of the cells (as in `set_traction!`):
`source === nothing` is the built-in follower pressure `p = params[1]` (nominal traction `-p cof(F) N`); otherwise the
HIP source of a point form of the cell's 12 slots at the face point (struct `name`).
"""
set_boundary_form!(op::DeviceOperator, face_cell::Vector{Int32}, face_local::Vector{Int32};
                   source::Union{Nothing,AbstractString}=nothing, name::AbstractString="", params::Vector{Float64}=[1.0]) =
    check(ccall((:hfk_op_set_boundary_form, libhfk), Cint,
                (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Int32}, Cstring, Cstring, Ptr{Float64}, Cint), op.ptr,
                length(face_cell), face_cell, face_local, source === nothing ? C_NULL : source,
                source === nothing ? C_NULL : name, params, length(params)), op.ctx.ptr)

"The load factor Λ of the traction (a number of the form, as in `en(Λ)` of test/solvers/test_functionals.jl:40-45)."
set_load_factor!(op::DeviceOperator, Λ::Real) =
    check(ccall((:hfk_op_set_load_factor, libhfk), Cint, (Ptr{Cvoid}, Cdouble), op.ptr, Λ), op.ctx.ptr)

"Pointwise shear modulus at the quadrature points (np × ncells, tensor point order); `nothing`: the constant μ."
set_mu_field!(op::DeviceOperator, μ::Union{Nothing,Matrix{Float64}}) =
    check(ccall((:hfk_op_set_mu_field, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float64}), op.ptr, μ === nothing ? C_NULL : μ),
          op.ctx.ptr)

"""
    set_mu_function!(op, source, name, params)

A function-of-x leaf (`CoordArray`, src/coordfunctions.jl:150-213) as the shear modulus, evaluated on the device at the
physical quadrature points: `source` is HIP C++ defining `__device__ double name(const double* x, const double* p)`,
compiled with hiprtc (cached by source text); `params` is the array `p`.
"""
set_mu_function!(op::DeviceOperator, source::AbstractString, name::AbstractString, params::Vector{Float64}=Float64[]) =
    check(ccall((:hfk_op_set_mu_function, libhfk), Cint, (Ptr{Cvoid}, Cstring, Cstring, Ptr{Float64}, Int64), op.ptr,
                source, name, isempty(params) ? C_NULL : params, length(params)), op.ctx.ptr)

"""
    CompiledLaw(ctx, source, name)

A pointwise constitutive law compiled at run time with hiprtc (the role of the generated `PointForm`,
src/codegen.jl:166-198): `source` is HIP C++ defining a struct `name` with `setup(F, λ, μ)`, `stress(c, J)` and
`tangent(c, J, k, L)` (`__device__`, F column-major, indices 0-based).  Code objects are cached per process by
source hash.  A law that does not compile throws an `ArgumentError` with the compiler log.
"""
mutable struct CompiledLaw
    ptr::Ptr{Cvoid}
    ctx::Context
end
function CompiledLaw(ctx::Context, source::AbstractString, name::AbstractString)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hfk_law_create, libhfk), Cint, (Ptr{Cvoid}, Cstring, Cstring, Ptr{Ptr{Cvoid}}), ctx.ptr, source, name, r),
          ctx.ptr)
    l = CompiledLaw(r[], ctx)
    finalizer(l -> ccall((:hfk_law_destroy, libhfk), Cint, (Ptr{Cvoid},), l.ptr), l)
    l
end

"Compile a law without a device: `nothing`, or an `ArgumentError` with the compiler log."
check_law(source::AbstractString, name::AbstractString) =
    check(ccall((:hfk_law_check, libhfk), Cint, (Cstring, Cstring, Ptr{UInt8}, Int64), source, name, C_NULL, 0))

"Evaluate `op` with a compiled law (`nothing`: its built-in law); prepared tangents are invalidated."
function set_law!(op::DeviceOperator, law::Union{Nothing,CompiledLaw})
    check(ccall((:hfk_op_set_law, libhfk), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), op.ptr, law === nothing ? C_NULL : law.ptr),
          op.ctx.ptr)
    op.law = law
    op
end

"""
    DeviceFunctional(op; reduce = :sum)
    DeviceFunctional(op, source, name; reduce = :min, params = Float64[])

`KernelFunctional` (src/functionals.jl:9-154) on the device: the strain energy of `op`'s law (`:sum`), or the
extremum over the quadrature points of a pointwise function `g(Gu, Gdu, p)` given as HIP source (`:min`/`:max`), with
the increment set by `set_increment!` (free values and Dirichlet values, changeable between calls like the
reference's `duh`).  Call it on the free values: `fun(x) -> Float64`.
"""
mutable struct DeviceFunctional
    op::Any
    reduce::Symbol
    fn::Ptr{Cvoid}
    params::Vector{Float64}
    dx::Vector{Float64}
    dx_dir::Union{Nothing,Vector{Float64}}
end
function DeviceFunctional(op; reduce::Symbol=:sum)
    reduce === :sum || throw(ArgumentError("HyperFEMKernels: a :min/:max functional needs a pointwise function"))
    DeviceFunctional(op, :sum, C_NULL, Float64[], zeros(op.nfree), nothing)
end
function DeviceFunctional(op, source::AbstractString, name::AbstractString; reduce::Symbol=:min,
                          params::Vector{Float64}=Float64[])
    reduce in (:min, :max) || throw(ArgumentError("HyperFEMKernels: reduce must be :min or :max"))
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hfk_pointfn_create, libhfk), Cint, (Ptr{Cvoid}, Cstring, Cstring, Ptr{Ptr{Cvoid}}), op.ctx.ptr, source,
                name, r), op.ctx.ptr)
    f = DeviceFunctional(op, reduce, r[], params, zeros(op.nfree), nothing)
    finalizer(f -> ccall((:hfk_pointfn_destroy, libhfk), Cint, (Ptr{Cvoid},), f.fn), f)
    f
end
set_increment!(f::DeviceFunctional, dx::Vector{Float64}, dx_dir::Union{Nothing,Vector{Float64}}=nothing) =
    (f.dx = dx; f.dx_dir = dx_dir; f)
function (f::DeviceFunctional)(x::Vector{Float64})
    v = Ref{Float64}(0.0)
    if f.reduce === :sum
        check(ccall((:hfk_energy, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), f.op.ptr, x, v), f.op.ctx.ptr)
    else
        check(ccall((:hfk_extremum, libhfk), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Cint, Ptr{Float64}),
                    f.op.ptr, f.fn, x, f.dx, f.dx_dir === nothing ? C_NULL : f.dx_dir,
                    isempty(f.params) ? C_NULL : f.params, length(f.params), f.reduce === :min ? 0 : 1, v), f.op.ctx.ptr)
    end
    v[]
end

"""
    energy_noalloc!(energy_dev, f::DeviceFunctional, x_dev; stream = C_NULL)

The energy of `f` (reduce = :sum) from device pointers without any device allocation (no counterpart in the
reference: for a line search captured in a HIP graph).  The partial sums use the library's static 1 MiB scratch per
device: at most 131,072 of them (batches + 512 + Robin faces); calls that overlap on one device race.
"""
function energy_noalloc!(energy_dev::Ptr{Float64}, f::DeviceFunctional, x_dev::Ptr{Float64}; stream::Ptr{Cvoid}=C_NULL)
    f.reduce === :sum || throw(ArgumentError("HyperFEMKernels: energy_noalloc! needs a reduce = :sum functional"))
    check(ccall((:hfk_energy_noalloc_dev, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                f.op.ptr, x_dev, energy_dev, stream), f.op.ctx.ptr)
    energy_dev
end

"Algebra.residual!(b, op, x) (src/assembly.jl:909-916)"
function residual!(b::Vector{Float64}, op::DeviceOperator, x::Vector{Float64})
    length(x) == op.nfree == length(b) || throw(DimensionMismatch("HyperFEMKernels: vector lengths"))
    check(ccall((:hfk_residual, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), op.ptr, x, b), op.ctx.ptr)
    b
end

"sparsity_pattern(op) and nonzero_positions(op) (src/assembly.jl:768-771)"
function sparsity_pattern(op::DeviceOperator, nd::Integer, ncells::Integer)
    nnz = Ref{Int64}(0)
    check(ccall((:hfk_pattern_nnz, libhfk), Cint, (Ptr{Cvoid}, Ptr{Int64}), op.ptr, nnz), op.ctx.ptr)
    colptr = Vector{Int64}(undef, op.nfree + 1)
    rowval = Vector{Int64}(undef, nnz[])
    nzpos = Matrix{Int32}(undef, nd * nd, ncells)
    check(ccall((:hfk_pattern, libhfk), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Int32}), op.ptr, colptr, rowval,
                nzpos), op.ctx.ptr)
    SparseMatrixCSC(op.nfree, op.nfree, colptr, rowval, zeros(nnz[])), nzpos
end

"jacobian!(A, op, x) (src/assembly.jl:918-927): the nonzeros of A overwritten"
function jacobian!(A::SparseMatrixCSC{Float64,Int}, op::DeviceOperator, x::Vector{Float64})
    check(ccall((:hfk_jacobian, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), op.ptr, x, nonzeros(A)), op.ctx.ptr)
    A
end

"MatrixFreeJacobian on the device: prepare!(mf, x) and mul!(y, mf, dx) (src/matrixfree.jl:140-260)"
mutable struct DeviceMatrixFree
    op::DeviceOperator
    storage::Symbol
end
function prepare!(mf::DeviceMatrixFree, x::Vector{Float64})
    check(ccall((:hfk_mf_prepare, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint), mf.op.ptr, x, HFK_STORE[mf.storage]),
          mf.op.ctx.ptr)
    mf
end
function mul!(y::Vector{Float64}, mf::DeviceMatrixFree, dx::Vector{Float64})
    check(ccall((:hfk_mf_apply, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), mf.op.ptr, dx, y), mf.op.ctx.ptr)
    y
end

"PMultigrid(spaces; smoother = Chebyshev(degree, range, inflate, eig_iters)) on device operators, finest first"
mutable struct DevicePMultigrid
    ptr::Ptr{Cvoid}
    levels::Vector{DeviceOperator}
end
function DevicePMultigrid(levels::Vector{DeviceOperator}; degree=2, range=20.0, inflate=1.1, eig_iters=10,
                          smoother_storage::Union{Nothing,Symbol}=nothing, coarse_rtol=1e-2, coarse_maxiter=500,
                          coarse_every::Integer=1)
    ptrs = [l.ptr for l in levels]
    r = Ref{Ptr{Cvoid}}(C_NULL)
    sst = smoother_storage === nothing ? Cint(-1) : HFK_STORE[smoother_storage]
    check(ccall((:hfk_pmg_create, libhfk), Cint,
                (Ptr{Ptr{Cvoid}}, Cint, Cint, Cdouble, Cdouble, Cint, Cint, Cint, Cdouble, Cint, Ptr{Ptr{Cvoid}}),
                ptrs, length(levels), degree, range, inflate, eig_iters, sst, 0, coarse_rtol, coarse_maxiter, r),
          levels[1].ctx.ptr)
    g = DevicePMultigrid(r[], levels)
    finalizer(g -> ccall((:hfk_pmg_destroy, libhfk), Cint, (Ptr{Cvoid},), g.ptr), g)
    check(ccall((:hfk_pmg_set_coarse_every, libhfk), Cint, (Ptr{Cvoid}, Cint), g.ptr, coarse_every), levels[1].ctx.ptr)
    g
end
"preconditioner_update!: returns (ok, indefinite)"
function update!(g::DevicePMultigrid)
    st = zeros(Cint, 2)
    check(ccall((:hfk_pmg_update, libhfk), Cint, (Ptr{Cvoid}, Ptr{Cint}), g.ptr, st), g.levels[1].ctx.ptr)
    (st[1] == 1, st[2] == 1)
end

"MatrixFreeSolver solve! (src/solvers.jl:111-158): returns (converged, iterations, indefinite, residual norm)"
const KRYLOV = Dict(:cg => 0, :fcg => 1, :gmres => 2, :fgmres => 3)
function solve!(x::Vector{Float64}, mf::DeviceMatrixFree, b::Vector{Float64}; pmg::Union{Nothing,DevicePMultigrid}=nothing,
                krylov::Symbol=:gmres, rtol=1e-8, atol=0.0, maxiter=500, restart=50)
    st = zeros(Cint, 3)
    rn = zeros(2)
    prec = pmg === nothing ? Cint(1) : Cint(2)
    pmg === nothing && check(ccall((:hfk_block_jacobi_update, libhfk), Cint, (Ptr{Cvoid}, Ptr{Cint}), mf.op.ptr, st),
                             mf.op.ctx.ptr)
    check(ccall((:hfk_krylov_solve, libhfk), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Cint, Cint, Ptr{Cint},
                 Ptr{Float64}),
                mf.op.ptr, pmg === nothing ? C_NULL : pmg.ptr, KRYLOV[krylov], prec, b, x, rtol, atol, maxiter, restart,
                st, rn),
          mf.op.ctx.ptr)
    (st[1] == 1, Int(st[2]), st[3] == 1, rn[1])
end

# --- learned coarse correction (no counterpart in the reference) ---------------------------------------------------
# Device buffers (training samples) are passed as raw device pointers, e.g. `pointer(::ROCArray{Float64})` from
# AMDGPU.jl; `stream = C_NULL` runs on the context's stream.

"U-Net on the vertex lattice of the Q1 coarse space (n_lattice^3 vertices, lattice_dofs: 3 x n_lattice^3 signed ids)"
mutable struct CoarseNetwork
    ptr::Ptr{Cvoid}
    ctx::Context
end
function CoarseNetwork(ctx::Context, n_lattice::Integer, lattice_dofs::Matrix{Int32}, ncoarse::Integer, h::Real;
                       seed::Integer=0, coarsen::Bool=false)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:hfk_net_create, libhfk), Cint,
                (Ptr{Cvoid}, Cint, Cint, Ptr{Int32}, Int64, Cdouble, UInt64, Ptr{Ptr{Cvoid}}),
                ctx.ptr, n_lattice, coarsen ? 1 : 0, lattice_dofs, ncoarse, h, seed, r), ctx.ptr)
    net = CoarseNetwork(r[], ctx)
    finalizer(n -> ccall((:hfk_net_destroy, libhfk), Cint, (Ptr{Cvoid},), n.ptr), net)
    net
end
set_aux!(net::CoarseNetwork, aux::Union{Nothing,Vector{Float32}}) =
    check(ccall((:hfk_net_set_aux, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float32}), net.ptr,
                aux === nothing ? Ptr{Float32}(C_NULL) : pointer(aux)), net.ctx.ptr)

"restricted (and coarse-smoothed) p-MG residuals of the device right-hand sides b_dev ([k][nfine]) into bc_dev"
coarse_pairs!(g::DevicePMultigrid, b_dev::Ptr{Float64}, nsamples::Integer, bc_dev::Ptr{Float64};
              coarse_smoothed::Bool=true, stream::Ptr{Cvoid}=C_NULL) =
    check(ccall((:hfk_pmg_coarse_pairs_dev, libhfk), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Ptr{Float64}, Ptr{Float64}, Cint, Cdouble, Cint, Ptr{Cvoid}),
                g.ptr, b_dev, nsamples, bc_dev, C_NULL, coarse_smoothed ? 1 : 0, 1e-10, 5000, stream), g.levels[1].ctx.ptr)

"energy training of net on the coarse level of g; returns the loss per step"
function train!(g::DevicePMultigrid, net::CoarseNetwork, bc_dev::Ptr{Float64}, nsamples::Integer, steps::Integer;
                batch::Integer=4, lr=1e-3, seed::Integer=0, stream::Ptr{Cvoid}=C_NULL)
    hist = zeros(max(steps, 1))
    check(ccall((:hfk_pmg_train_coarse_network, libhfk), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Cint, Cint, Cint, Cdouble, UInt64, Ptr{Float64}, Ptr{Cvoid}),
                g.ptr, net.ptr, bc_dev, nsamples, steps, batch, lr, seed, hist, stream), g.levels[1].ctx.ptr)
    hist[1:steps]
end

const COARSE_MODES = Dict(:off => 0, :warmstart => 1, :replace => 2, :precondition => 3, :cycle => 4)
"attach net to the coarse level of g (mode in COARSE_MODES); ldiv!/solve! are unchanged"
set_coarse_network!(g::DevicePMultigrid, net::Union{Nothing,CoarseNetwork}, mode::Symbol=:precondition;
                    smoother_degree::Integer=0, scaled::Bool=false) =
    check(ccall((:hfk_pmg_set_coarse_network, libhfk), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint),
                g.ptr, net === nothing ? C_NULL : net.ptr, COARSE_MODES[mode], smoother_degree, scaled ? 1 : 0),
          g.levels[1].ctx.ptr)

# --- Oseen-BDF2 step on tetrahedra (P2-P1, SUPG/PSPG/LSIC; test/common.jl:165-228) ---------------------------------
"""KernelOperator of the Oseen step on the device: cell_dofs 34 x ncells (Gridap's MultiFieldFESpace ids in Gridap's
local order), coords 3 x 4 x ncells, quadrature points 3 x np and weights np of the unit tetrahedron.  The local
order is Gridap's, passed as tables (0-based): `nodeperm` (10) the P2 table node of each of the library's nodes (the
cell's vertices in the order of `coords`, then the midpoints of edges (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)),
`pnodeperm` (4) the P1 table node of each vertex, `ldof` (34) the cell-local dof of (velocity component c, P2 node n)
at c*10 + n and of the pressure at P1 node n at 30 + n (FieldTable ldof, src/tables.jl:171-237); `nothing` is the
library's own order."""
mutable struct DeviceOseen
    ptr::Ptr{Cvoid}
    ctx::Context
    nfree::Int
    pattern::Union{Nothing,Tuple{Vector{Int},Vector{Int},Matrix{Int32}}}
end
function DeviceOseen(ctx::Context, cell_dofs::Matrix{Int32}, coords::Array{Float64,3}, qpoints::Matrix{Float64},
                     qweights::Vector{Float64}, nfree::Integer, ndir::Integer; rho=1.0, mu=0.01, c0=150.0,
                     nodeperm::Union{Nothing,Vector{Int32}}=nothing, pnodeperm::Union{Nothing,Vector{Int32}}=nothing,
                     ldof::Union{Nothing,Vector{Int32}}=nothing)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    orptr(a) = a === nothing ? Ptr{Int32}(C_NULL) : pointer(a)
    GC.@preserve nodeperm pnodeperm ldof check(ccall((:hfk_oseen_create, libhfk), Cint,
                (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float64}, Cint, Ptr{Float64},
                 Ptr{Float64}, Int64, Int64, Cdouble, Cdouble, Cdouble, Ptr{Ptr{Cvoid}}),
                ctx.ptr, size(cell_dofs, 2), cell_dofs, orptr(nodeperm), orptr(pnodeperm), orptr(ldof), coords,
                length(qweights), qpoints, qweights, nfree, ndir, rho, mu, c0, r), ctx.ptr)
    op = DeviceOseen(r[], ctx, nfree, nothing)
    finalizer(o -> ccall((:hfk_oseen_destroy, libhfk), Cint, (Ptr{Cvoid},), o.ptr), op)
    op
end
"""
    set_form!(op::DeviceOseen, source, name, params = Float64[])

The volume integrand as a point form compiled with hiprtc (the PointForm of the captured integrand,
src/codegen.jl:122-198): HIP source of a struct `name` with `residual` and `tangent` (see include/hfk.h); the Julia
host emits it from the captured tree as hfk/formgen.py does from a Python one.  `nothing` restores the built-in
SUPG/PSPG/LSIC integrand.
"""
set_form!(op::DeviceOseen, source::Union{Nothing,AbstractString}, name::AbstractString="", params::Vector{Float64}=Float64[]) =
    check(ccall((:hfk_oseen_set_form, libhfk), Cint, (Ptr{Cvoid}, Cstring, Cstring, Ptr{Float64}, Cint), op.ptr,
                source === nothing ? C_NULL : source, source === nothing ? C_NULL : name,
                isempty(params) ? C_NULL : params, length(params)), op.ctx.ptr)

"""
    set_blocks!(op::DeviceOseen, mask)            # mask::BitMatrix 2 x 2, or a 2 x 2 x ncells array per cell

The (test field, trial field) blocks that carry a term (`form_blocks`/`cell_block_mask`, src/assembly.jl:288-303,
825-832), field 1 the velocity and 2 the pressure; `nothing`: every block.  The pattern is rebuilt at its next use.
"""
function set_blocks!(op::DeviceOseen, mask)
    op.pattern = nothing
    mask === nothing && return check(ccall((:hfk_oseen_set_blocks, libhfk), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int32}),
                                           op.ptr, C_NULL, C_NULL), op.ctx.ptr)
    m = Int32.(mask)  # column-major: [fi + 2 fj] (0-based fields) per cell
    ndims(m) == 2 ? check(ccall((:hfk_oseen_set_blocks, libhfk), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int32}), op.ptr, m,
                                C_NULL), op.ctx.ptr) :
                    check(ccall((:hfk_oseen_set_blocks, libhfk), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int32}), op.ptr,
                                C_NULL, m), op.ctx.ptr)
end

"point data 9 x np x ncells: advecting velocity w, stabilisation velocity un, history f"
set_fields!(op::DeviceOseen, fields::Array{Float64,3}) =
    check(ccall((:hfk_oseen_set_fields, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float64}), op.ptr, fields), op.ctx.ptr)
set_dirichlet!(op::DeviceOseen, dvals::Vector{Float64}) =
    check(ccall((:hfk_oseen_set_dirichlet, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float64}), op.ptr, dvals), op.ctx.ptr)
"Elementwise Laplacian of uⁿ at the points (3 × np × ncells), the `lapn` of the stabilisation load (test/common.jl:191-195); `nothing` removes it."
set_laplacian!(op::DeviceOseen, lap::Union{Nothing,Array{Float64,3}}) =
    check(ccall((:hfk_oseen_set_laplacian, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float64}), op.ptr,
                lap === nothing ? C_NULL : lap), op.ctx.ptr)

"""
    step_fields!(op, un, un_dir, unm1, unm1_dir; aw, bw, c1, c2, laplacian=true)

The point data of a step from the FE functions uⁿ (`xh⁻`) and uⁿ⁻¹ (`xh⁻⁻`) of test/common.jl:183-196, computed on
the device: their free values (`get_free_dof_values`, pressure entries not read) and the Dirichlet values of their
velocity spaces; w = aw uⁿ + bw uⁿ⁻¹, f = c1 uⁿ - c2 uⁿ⁻¹, the elementwise Laplacian of uⁿ when `laplacian`, and w at the
backflow points of the outlet faces.
"""
step_fields!(op::DeviceOseen, un::Vector{Float64}, un_dir::Vector{Float64}, unm1::Vector{Float64},
             unm1_dir::Vector{Float64}; aw::Real, bw::Real, c1::Real, c2::Real, laplacian::Bool=true) =
    check(ccall((:hfk_oseen_step_fields, libhfk), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Cdouble, Cdouble,
                 Cint),
                op.ptr, un, un_dir, unm1, unm1_dir, aw, bw, c1, c2, laplacian), op.ctx.ptr)

"""
    set_outlet!(op, face_cell, normals, xo, xb, Pc, ZA)

Outlet terms (test/common.jl:212-213, 219) on the faces of `BoundaryTriangulation(model, tags="outlet")`:
`face_cell` 0-based cells, `normals` 3 × nfaces, `xo` 4 × nqo × nfaces (reference coordinates in the cell and dS of
`Measure(Γo, 4)`), `xb` 7 × nqb × nfaces (the same for `Measure(Γo, 5)` with w appended).
"""
set_outlet!(op::DeviceOseen, face_cell::Vector{Int32}, normals::Matrix{Float64}, xo::Array{Float64,3},
            xb::Array{Float64,3}, Pc::Real, ZA::Real) =
    check(ccall((:hfk_oseen_set_outlet, libhfk), Cint,
                (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Cdouble, Cdouble),
                op.ptr, length(face_cell), face_cell, normals, size(xo, 2), xo, size(xb, 2), xb, Pc, ZA), op.ctx.ptr)

"`update(op, res, jac)` of the Oseen step with other numbers (ρ, μ, c0: load and time stepping, the BDF2 switch)."
update!(op::DeviceOseen; rho::Real, mu::Real, c0::Real) =
    (check(ccall((:hfk_oseen_set_parameters, libhfk), Cint, (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble), op.ptr, rho, mu, c0),
           op.ctx.ptr); op)

"The Windkessel pressure `Pc[]` of the next step."
set_outlet_pressure!(op::DeviceOseen, Pc::Real) =
    check(ccall((:hfk_oseen_set_outlet_pressure, libhfk), Cint, (Ptr{Cvoid}, Cdouble), op.ptr, Pc), op.ctx.ptr)

function residual!(b::Vector{Float64}, op::DeviceOseen, x::Vector{Float64})
    check(ccall((:hfk_oseen_residual, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), op.ptr, x, b), op.ctx.ptr)
    b
end
function sparsity_pattern(op::DeviceOseen)
    if op.pattern === nothing
        nnz = Ref{Int64}(0)
        check(ccall((:hfk_oseen_pattern_nnz, libhfk), Cint, (Ptr{Cvoid}, Ptr{Int64}), op.ptr, nnz), op.ctx.ptr)
        info = zeros(Int64, 6)
        check(ccall((:hfk_oseen_info, libhfk), Cint, (Ptr{Cvoid}, Ptr{Int64}), op.ptr, info), op.ctx.ptr)
        colptr = zeros(Int64, op.nfree + 1); rowval = zeros(Int64, nnz[]); nzpos = zeros(Int32, 34 * 34, info[3])
        check(ccall((:hfk_oseen_pattern, libhfk), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Int32}),
                    op.ptr, colptr, rowval, nzpos), op.ctx.ptr)
        op.pattern = (colptr, rowval, nzpos)
    end
    colptr, rowval, _ = op.pattern
    SparseMatrixCSC(op.nfree, op.nfree, colptr, rowval, zeros(length(rowval)))
end
function jacobian!(A::SparseMatrixCSC{Float64,Int}, op::DeviceOseen, x::Vector{Float64})
    check(ccall((:hfk_oseen_jacobian, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), op.ptr, x, nonzeros(A)),
          op.ctx.ptr)
    A
end

"""
    DeviceSystem(ctx, A::SparseMatrixCSC)        # the pattern of an assembled matrix (e.g. `sparsity_pattern(op)`)

An assembled square matrix on the device for Krylov solves of a multi-field system: `set_blocks!` with nodal blocks
(the reference's `field_blocks` per field, src/preconditioners.jl:151-171, or blocks coupling the fields at a node;
6 × nblocks free ids, 0 = empty slot), `set_patches!` with Vanka patches (restricted additive Schwarz), `update!`
with new values (extraction and inversion on the device), `solve!` as `MatrixFreeSolver` (src/solvers.jl:111-158).
"""
mutable struct DeviceSystem
    ptr::Ptr{Cvoid}
    ctx::Context
    n::Int
end
function DeviceSystem(ctx::Context, A::SparseMatrixCSC)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    colptr = Int64.(A.colptr); rowval = Int64.(A.rowval)
    check(ccall((:hfk_sys_create, libhfk), Cint, (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Ptr{Cvoid}}),
                ctx.ptr, size(A, 2), colptr, rowval, r), ctx.ptr)
    S = DeviceSystem(r[], ctx, size(A, 2))
    finalizer(o -> ccall((:hfk_sys_destroy, libhfk), Cint, (Ptr{Cvoid},), o.ptr), S)
    S
end
set_blocks!(S::DeviceSystem, blocks::Matrix{Int32}) =
    check(ccall((:hfk_sys_set_blocks, libhfk), Cint, (Ptr{Cvoid}, Int64, Ptr{Int32}), S.ptr, size(blocks, 2), blocks),
          S.ctx.ptr)
"Vanka patches: the dofs of patch p are dofs[ptr[p]+1 : ptr[p+1]] (ptr 0-based offsets), owned[k] ≠ 0 where patch p writes dofs[k]."
set_patches!(S::DeviceSystem, ptr::Vector{Int64}, dofs::Vector{Int32}, owned::Vector{Int8}) =
    check(ccall((:hfk_sys_set_patches, libhfk), Cint, (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int32}, Ptr{Int8}),
                S.ptr, length(ptr) - 1, ptr, dofs, owned), S.ctx.ptr)
"New values on the device (`nonzeros` of the CSC matrix, device pointer); returns the number of singular blocks/patches."
function update!(S::DeviceSystem, nzval_dev::Ptr{Float64}; stream::Ptr{Cvoid}=C_NULL)
    bad = Ref{Int64}(0)
    check(ccall((:hfk_sys_update_dev, libhfk), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Int64}, Ptr{Cvoid}),
                S.ptr, nzval_dev, bad, stream), S.ctx.ptr)
    bad[]
end
"""solve!(x_dev, S, b_dev; krylov = :gmres, preconditioner = :blocks)   # or :sweeps, :patches, :none"""
function solve!(x_dev::Ptr{Float64}, S::DeviceSystem, b_dev::Ptr{Float64}; krylov::Symbol=:gmres,
                preconditioner::Symbol=:blocks, sweeps::Integer=2, omega::Real=0.8, rtol::Real=1e-8, atol::Real=0.0,
                maxiter::Integer=1000, restart::Integer=50, stream::Ptr{Cvoid}=C_NULL)
    kind = krylov === :gmres ? 2 : krylov === :fgmres ? 3 : throw(ArgumentError("krylov must be :gmres or :fgmres"))
    pc = Dict(:none => 0, :blocks => 1, :sweeps => 2, :patches => 3)[preconditioner]
    st = zeros(Int32, 3); rn = zeros(2)
    check(ccall((:hfk_sys_solve_dev, libhfk), Cint,
                (Ptr{Cvoid}, Cint, Cint, Cint, Cdouble, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Cint, Cint,
                 Ptr{Int32}, Ptr{Float64}, Ptr{Cvoid}),
                S.ptr, kind, pc, sweeps, omega, b_dev, x_dev, rtol, atol, maxiter, restart, st, rn, stream), S.ctx.ptr)
    (converged = st[1] == 1, iterations = Int(st[2]), residual = rn[1], rhs_norm = rn[2])
end

end # module
