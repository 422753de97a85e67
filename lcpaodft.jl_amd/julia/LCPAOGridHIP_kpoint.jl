# LCPAOGridHIP_kpoint.jl -- drop-in methods for the steps on either side of
# LCPAODFT.jl's k-point diagonalisation, calling liblcpao_grid.so.
#
# `include` after LCPAOGridHIP.jl, for runs on ONE rank only: the library
# sums over the pairs its context holds, and a rank's context holds only that
# rank's pairs, while these reference routines loop over every pair.
# Untested here: Julia is not available in the build environment; the C ABI
# underneath is tested from Python (tests/test_gpu_bloch.py).
#
# Replaced methods (reference file:line):
#   HS_matrix!(S, H, OLP::Vector{Float64}, Hks::Vector{Float64}, ..., kpts)   HS_matrix.jl:143
#   Calc_DM_Crystal_Collinear_nopol!                                        Crystal_DFT.jl:219

function _kpoint_context()
    MPI.Comm_size(MPI.COMM_WORLD) == 1 || error("LCPAOGridHIP_kpoint.jl: one rank only")
    last = LAST_GRID[]
    last === nothing && error("no grid context yet (build one through a grid routine first)")
    return last
end

function HS_matrix!(S, H, OLP::Vector{Float64}, Hks::Vector{Float64}, Natom, Total_NumOrbs, MP, FNAN, natn, ncn,
                    atv_ijk, kpts::AbstractVector{Float64})
    ucell, h = _kpoint_context()
    k = Float64[kpts[1], kpts[2], kpts[3]]
    fp = [pointer(OLP), pointer(Hks)]
    dp = [Ptr{Float64}(pointer(S)), Ptr{Float64}(pointer(H))]   # ComplexF64 column-major = interleaved doubles
    GC.@preserve S H OLP Hks fp dp k begin
        rc = ccall((:lcpao_grid_bloch_sum, LIBGRID), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Int32, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}), h.ptr, k, 2, fp, dp)
        _grid_check(rc, h.ptr)
    end
    return nothing
end

function Calc_DM_Crystal_Collinear_nopol!(electron::CrystalBloch, kpoints::KPoints, system_grid::System_Grid,
                                          DM::Vector{Vector{Float64}})
    ucell, h = _kpoint_context()
    Calc_fnkCnk!(electron, kpoints)                    # scales the states (Crystal_DFT.jl:243)
    Cnk = electron.Cnk[1]
    nk = kpoints.MPI_Nkpt
    kp = Float64[x for ik in 1:nk for x in (kpoints.MPI_kpts[ik][1], kpoints.MPI_kpts[ik][2], kpoints.MPI_kpts[ik][3])]
    nst = size(Cnk[1], 2)
    cp = [Ptr{Float64}(pointer(Cnk[ik])) for ik in 1:nk]
    GC.@preserve Cnk cp kp DM begin
        rc = ccall((:lcpao_grid_dm_states, LIBGRID), Cint,
                   (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64, Ptr{Ptr{Float64}}, Float64, Ptr{Float64}),
                   h.ptr, nk, kp, nst, cp, 1.0 / kpoints.AllNkpt, DM[1])
        _grid_check(rc, h.ptr)
    end
    return nothing
end
