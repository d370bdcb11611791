# backend_amd.jl -- AMD (MI355X, gfx950) backend for LibNEGF.jl.
#
# Selected like the reference's backends, by LIBNEGF_HW = "amd"
# (src/LibNEGF.jl:54-56 includes "backend_" * ENV["LIBNEGF_HW"] * ".jl").
# Device memory comes from AMDGPU.jl (ROCArray); all arithmetic is a ccall
# into libnegf_hip.so (include/negf_hip.h).  Mirrors src/backend_cpu.jl
# function for function; the sweep-level entry points at the end replace
# whole bndiag_of_inv_rgf_local! / keldyshndiag! / bndiag_of_inv_ddrgf!
# calls when the caller opts in.
#
# This file is the binding a LibNEGF.jl maintainer adds.  The C ABI it
# calls is the one the Python host mirror (libnegf_amd) and its tests
# exercise; this file itself has not been run under Julia.

using AMDGPU
import LinearAlgebra

const libnegf = joinpath(@__DIR__, "libnegf_amd", "libnegf_hip.so")

# the tuner's measured choices (GEMM tiles, inverse schedules, RGF vs DD-RGF levels)
function __init_tune_table__()
    path = get(ENV, "LIBNEGF_AMD_TUNE_FILE", joinpath(@__DIR__, "libnegf_amd", "tune_table.txt"))
    if !isempty(path) && isfile(path)
        ccall((:negf_tune_load, libnegf), Cint, (Cstring,), path)
    end
end

# ----------------------------------------------------------------------------
# types (src/backend_cpu.jl:7-20).  LibNEGF.jl:48 must become
#   const ArrayWithType = ROCArray{FieldType,2}
# for this backend (the comment at LibNEGF.jl:49-50 anticipates it).
"""
    AmdLU

Holds what the AMD backend needs to apply op(S)^{-1}: the LAPACK-layout
factors and 1-based pivots on the device (the counterpart of CpuLU).
"""
mutable struct AmdLU
    A::ROCArray{FieldType,2}
    piv::ROCArray{Int32,1}
end

const ArrayOrLUUnder_ = Union{ArrayWithType,AmdLU,Nothing}
const ArrayOrLU_ = Matrix{ArrayOrLUUnder_}
const ArrayOrLUView_ = SubArray{ArrayOrLUUnder_,2,ArrayOrLU_,Tuple{UnitRange{Int},UnitRange{Int}},false}

_stream() = AMDGPU.stream().stream          # hipStream_t of the current task
_ld(M) = max(stride(M, 2), 1)

function _check(rc::Cint, where::String)
    if rc < 0
        msg = unsafe_string(ccall((:negf_status_string, libnegf), Cstring, (Cint,), rc))
        error("$where: $msg ($rc)")
    end
end

# The reference prints and exits on a zero pivot (src/backend_cpu.jl:151-155).
function _check_info(info::ROCArray{Int32}, where::String)
    if any(!=(0), Array(info))
        println(Core.stdout, "ERROR: LAPACK lu returned an error info")
        exit()
    end
end

# ----------------------------------------------------------------------------
# memory (src/backend_cpu.jl:26-56, 82-88)
be_copy_to_hw!(Mout::ROCArray, Min::Array) = copyto!(Mout, Min)
be_copy_to_hw(M::Array)::ROCArray = ROCArray(M)
be_copy_from_hw!(Mout::Array, Min::ROCArray) = copyto!(Mout, Min)
be_copy_from_hw(M::ROCArray)::Array = Array(M)

function be_copy_in_hw!(Mout::ROCArray, Min::ROCArray)
    m, n = size(Min)
    _check(ccall((:negf_zcopy_strided_batched, libnegf), Cint,
                 (Int32, Int32, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Int64, Int64, Int32, Ptr{Cvoid}),
                 m, n, pointer(Min), _ld(Min), 0, pointer(Mout), _ld(Mout), 0, 1, _stream()),
           "be_copy_in_hw!")
end
be_copy_in_hw(M::ROCArray)::ROCArray = copy(M)

be_zero_array(dims::Tuple{Int,Int})::ROCArray = AMDGPU.zeros(FieldType, dims)

function be_identity(n::Int)::ROCArray
    M = be_zero_array((n, n))
    one_ = FieldType[1.0]; zero_ = FieldType[0.0]
    _check(ccall((:negf_zfill_strided_batched, libnegf), Cint,
                 (Int32, Int32, Ptr{FieldType}, Ptr{FieldType}, Ptr{Cvoid}, Int64, Int64, Int32, Ptr{Cvoid}),
                 n, n, zero_, one_, pointer(M), n, 0, 1, _stream()), "be_identity")
    return M
end

function be_fill!(M::ROCArray, x::Number)
    v = FieldType[x]
    _check(ccall((:negf_zfill_strided_batched, libnegf), Cint,
                 (Int32, Int32, Ptr{FieldType}, Ptr{FieldType}, Ptr{Cvoid}, Int64, Int64, Int32, Ptr{Cvoid}),
                 size(M, 1), size(M, 2), v, v, pointer(M), _ld(M), 0, 1, _stream()), "be_fill!")
end

be_random_array(dims::Tuple{Int,Int})::ROCArray = ROCArray(rand(FieldType, dims))

function be_ctranspose!(Mout::ROCArray, Min::ROCArray, td::TimingData, cd::CountingData)
    m, n = size(Min)
    _check(ccall((:negf_zadjoint_strided_batched, libnegf), Cint,
                 (Int32, Int32, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Int64, Int64, Int32, Ptr{Cvoid}),
                 m, n, pointer(Min), _ld(Min), 0, pointer(Mout), _ld(Mout), 0, 1, _stream()),
           "be_ctranspose!")
end

# non-mutating adjoint (src/backend_cpu.jl:70-80): a new device matrix; bm_convert
# reads Hermitian storage's lower blocks through it (src/blockmatrix.jl:108)
function be_ctranspose(Min::ROCArray, td::TimingData, cd::CountingData)::ROCArray
    Mout = be_zero_array((size(Min, 2), size(Min, 1)))
    be_ctranspose!(Mout, Min, td, cd)
    return Mout
end

function be_zero_lu(n::Int)::AmdLU
    return AmdLU(be_zero_array((n, n)), AMDGPU.ones(Int32, n))
end

# ----------------------------------------------------------------------------
# factorizations and solves (src/backend_cpu.jl:143-265)
function be_lu!(Mout::AmdLU, Min::ROCArray, td::TimingData, cd::CountingData)
    be_copy_in_hw!(Mout.A, Min)
    info = AMDGPU.zeros(Int32, 1)
    n = size(Min, 1)
    _check(ccall((:negf_zgetrf, libnegf), Cint,
                 (Int32, Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Int32}, Ptr{Cvoid}),
                 n, pointer(Mout.A), _ld(Mout.A), pointer(Mout.piv), pointer(info), _stream()),
           "be_lu!")
    _check_info(info, "be_lu!")
end

function be_mldivide!(trans::Char, Mout::ROCArray, Min::ROCArray, Mlu::AmdLU,
    td::TimingData, cd::CountingData)
    be_copy_in_hw!(Mout, Min)
    n, nrhs = size(Mout)
    _check(ccall((:negf_zgetrs, libnegf), Cint,
                 (UInt8, Int32, Int32, Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
                 UInt8(trans), n, nrhs, pointer(Mlu.A), _ld(Mlu.A), pointer(Mlu.piv),
                 pointer(Mout), _ld(Mout), _stream()), "be_mldivide!")
end

function be_gemm!(tA::Char, tB::Char, alpha::Number, A::ROCArray, B::ROCArray, beta::Number,
    C::ROCArray, td::TimingData, cd::CountingData)
    m, n = size(C)
    k = tA == 'N' ? size(A, 2) : size(A, 1)
    a = FieldType[alpha]; b = FieldType[beta]
    _check(ccall((:negf_zgemm, libnegf), Cint,
                 (UInt8, UInt8, Int32, Int32, Int32, Ptr{FieldType}, Ptr{Cvoid}, Int64, Ptr{Cvoid},
                  Int64, Ptr{FieldType}, Ptr{Cvoid}, Int64, Ptr{Cvoid}),
                 UInt8(tA), UInt8(tB), m, n, k, a, pointer(A), _ld(A), pointer(B), _ld(B), b,
                 pointer(C), _ld(C), _stream()), "be_gemm!")
end

function be_inv(M::ROCArray)::ROCArray
    n = size(M, 1)
    out = be_zero_array((n, n))
    info = AMDGPU.zeros(Int32, 1)
    _check(ccall((:negf_zinv_strided_batched, libnegf), Cint,
                 (Int32, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Int64, Int64, Int32, Ptr{Int32}, Ptr{Cvoid}),
                 n, pointer(M), _ld(M), 0, pointer(out), n, 0, 1, pointer(info), _stream()), "be_inv")
    _check_info(info, "be_inv")
    return out
end

# selected inverse ops: pivots discarded / identity pivots, as the reference
function be_getrf!(A::ROCArray, td::TimingData, cd::CountingData)
    n = size(A, 1)
    piv = AMDGPU.zeros(Int32, n); info = AMDGPU.zeros(Int32, 1)
    _check(ccall((:negf_zgetrf, libnegf), Cint,
                 (Int32, Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Int32}, Ptr{Cvoid}),
                 n, pointer(A), _ld(A), pointer(piv), pointer(info), _stream()), "be_getrf!")
end

function be_getri!(A::ROCArray, td::TimingData, cd::CountingData)
    n = size(A, 1)
    work = be_zero_array((n, n))
    _check(ccall((:negf_zgetri, libnegf), Cint,
                 (Int32, Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Cvoid}, Ptr{Cvoid}),
                 n, pointer(A), _ld(A), C_NULL, pointer(work), _stream()), "be_getri!")
end

function be_trsm!(side::Char, ul::Char, tA::Char, dA::Char, alpha::Number, A::ROCArray,
    B::ROCArray, td::TimingData, cd::CountingData)
    m, n = size(B)
    a = FieldType[alpha]
    _check(ccall((:negf_ztrsm, libnegf), Cint,
                 (UInt8, UInt8, UInt8, UInt8, Int32, Int32, Ptr{FieldType}, Ptr{Cvoid}, Int64,
                  Ptr{Cvoid}, Int64, Ptr{Cvoid}),
                 UInt8(side), UInt8(ul), UInt8(tA), UInt8(dA), m, n, a, pointer(A), _ld(A),
                 pointer(B), _ld(B), _stream()), "be_trsm!")
end

# ----------------------------------------------------------------------------
# sweep-level replacements (one call per sweep instead of one per block op).
# They take BlockMatrix objects whose blocks are packed contiguously; see
# INTEGRATION.md for the packing helper bm_pack_amd.
# bndiag_of_inv_ddrgf! on any blockSizes (levels: the tuner's blockSizeD1 per level)
function bndiag_of_inv_ddrgf_amd!(gd, gu, gl, ad, au, al, blockSizes::Vector{Int32}, batch::Int,
    levels::Vector{Int32})
    info = AMDGPU.zeros(Int32, batch)
    nd = sum(Int64.(blockSizes) .^ 2)
    no = sum(Int64(blockSizes[i]) * blockSizes[i+1] for i in 1:length(blockSizes)-1; init=0)
    _check(ccall((:negf_ddrgf_blocksizes, libnegf), Cint,
                 (Int32, Ptr{Int32}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64,
                  Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Int32}, Int32, Ptr{Int32}, Ptr{Cvoid}),
                 length(blockSizes), blockSizes, pointer(ad), nd, pointer(au), max(no, 1),
                 pointer(al), max(no, 1), pointer(gd), pointer(gu), pointer(gl), batch, levels,
                 length(levels), pointer(info), _stream()), "bndiag_of_inv_ddrgf_amd!")
    _check_info(info, "bndiag_of_inv_ddrgf_amd!")
end

function bndiag_of_inv_rgf_amd!(gd, gu, gl, ad, au, al, blockSizes::Vector{Int32}, batch::Int)
    info = AMDGPU.zeros(Int32, batch)
    nd = sum(Int64.(blockSizes) .^ 2)
    no = sum(Int64(blockSizes[i]) * blockSizes[i+1] for i in 1:length(blockSizes)-1; init=0)
    _check(ccall((:negf_rgf, libnegf), Cint,
                 (Int32, Ptr{Int32}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64,
                  Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Int32}, Ptr{Cvoid}),
                 length(blockSizes), blockSizes, pointer(ad), nd, pointer(au), max(no, 1),
                 pointer(al), max(no, 1), pointer(gd), pointer(gu), pointer(gl), batch,
                 pointer(info), _stream()), "bndiag_of_inv_rgf_amd!")
    _check_info(info, "bndiag_of_inv_rgf_amd!")
end
