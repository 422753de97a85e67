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
#   HS_matrix!(S, OLP::Vector{Float64}, ..., kpts)                          HS_matrix.jl:88
#   HS_matrix!(S, H, OLP::Vector{Float64}, Hks::Vector{Float64}, ..., natn) HS_matrix.jl:114 (no k; real S, H)
#   HS_matrix!(S, OLP::Vector{Float64}, ..., natn)                          HS_matrix.jl:67 (no k; real S)
#   Calc_DM_Crystal_Collinear_nopol!                                        Crystal_DFT.jl:219
#   Calc_DM_Crystal_Collinear_pol!                                          Crystal_DFT.jl:276
#   Calc_DM_Crystal_NonCollinear!                                           Crystal_DFT.jl:342
#   Calc_iDM_Crystal_NonCollinear!                                          Crystal_DFT.jl:413
#   HS_matrix_NC!(H, Hks::Vector{Vector{Float64}}, iHks, ..., kpts)         HS_matrix.jl:211
#   HS_matrix_NC!(H, Hks nested, iHks, ..., kpts)                           HS_matrix.jl:260
#   HS_matrix_NC!(H, Hks::Vector{Vector{Float64}}, iHks, ..., natn)         HS_matrix.jl:167 (no k)

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

# one to four flat arrays -> dense complex matrices (column-major ComplexF64)
function _bloch_sum!(h, k::Vector{Float64}, flats, denses)
    n = length(flats)
    fp = [pointer(f) for f in flats]
    dp = [Ptr{Float64}(pointer(d)) for d in denses]
    GC.@preserve flats denses fp dp k begin
        rc = ccall((:lcpao_grid_bloch_sum, LIBGRID), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Int32, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}), h.ptr, k, n, fp, dp)
        _grid_check(rc, h.ptr)
    end
end

function HS_matrix!(S, OLP::Vector{Float64}, Natom, Total_NumOrbs, MP, FNAN, natn, ncn, atv_ijk,
                    kpts::AbstractVector{Float64})
    ucell, h = _kpoint_context()
    _bloch_sum!(h, Float64[kpts[1], kpts[2], kpts[3]], [OLP], [S])
    return nothing
end

# Without k (the cluster forms): the same sum at k = 0, where every phase is
# exactly 1 + 0im, so the real parts are the reference's sums bit for bit.
function HS_matrix!(S, H, OLP::Vector{Float64}, Hks::Vector{Float64}, Natom, Total_NumOrbs, MP, FNAN, natn)
    ucell, h = _kpoint_context()
    Sc = zeros(ComplexF64, size(S))
    Hc = zeros(ComplexF64, size(H))
    _bloch_sum!(h, zeros(3), [OLP, Hks], [Sc, Hc])
    S .= real.(Sc)
    H .= real.(Hc)
    return nothing
end

function HS_matrix!(S, OLP::Vector{Float64}, Natom, Total_NumOrbs, MP, FNAN, natn)
    ucell, h = _kpoint_context()
    Sc = zeros(ComplexF64, size(S))
    _bloch_sum!(h, zeros(3), [OLP], [Sc])
    S .= real.(Sc)
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

# states of one spin channel -> the k-point list, pointers and state count
function _dm_states_call(h, kpoints::KPoints, Cnk_s, out::Vector{Float64})
    nk = kpoints.MPI_Nkpt
    kp = Float64[x for ik in 1:nk for x in (kpoints.MPI_kpts[ik][1], kpoints.MPI_kpts[ik][2], kpoints.MPI_kpts[ik][3])]
    nst = size(Cnk_s[1], 2)
    cp = [Ptr{Float64}(pointer(Cnk_s[ik])) for ik in 1:nk]
    GC.@preserve Cnk_s cp kp out begin
        rc = ccall((:lcpao_grid_dm_states, LIBGRID), Cint,
                   (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64, Ptr{Ptr{Float64}}, Float64, Ptr{Float64}),
                   h.ptr, nk, kp, nst, cp, 1.0 / kpoints.AllNkpt, out)
        _grid_check(rc, h.ptr)
    end
end

function Calc_DM_Crystal_Collinear_pol!(electron::CrystalBloch, kpoints::KPoints, system_grid::System_Grid,
                                        DM::Vector{Vector{Float64}})
    ucell, h = _kpoint_context()
    Calc_fnkCnk!(electron, kpoints)                    # scales the states (Crystal_DFT.jl:302)
    _dm_states_call(h, kpoints, electron.Cnk[1], DM[1])
    _dm_states_call(h, kpoints, electron.Cnk[2], DM[2])
    return nothing
end

# spinor states (Nfsize x nstates per k-point) -> DM[1..4] and, when iout is
# not nothing, the two imaginary spin-diagonal sums
function _dm_states_nc_call(h, kpoints::KPoints, Cnk_s, out::Vector{Vector{Float64}}, iout)
    nk = kpoints.MPI_Nkpt
    kp = Float64[x for ik in 1:nk for x in (kpoints.MPI_kpts[ik][1], kpoints.MPI_kpts[ik][2], kpoints.MPI_kpts[ik][3])]
    nst = size(Cnk_s[1], 2)
    cp = [Ptr{Float64}(pointer(Cnk_s[ik])) for ik in 1:nk]
    op = [pointer(out[s]) for s in 1:4]
    ip = iout === nothing ? Ptr{Ptr{Float64}}(C_NULL) : [pointer(iout[s]) for s in 1:2]
    GC.@preserve Cnk_s cp kp out op iout ip begin
        rc = ccall((:lcpao_grid_dm_states_nc, LIBGRID), Cint,
                   (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64, Ptr{Ptr{Float64}}, Float64, Ptr{Ptr{Float64}},
                    Ptr{Ptr{Float64}}),
                   h.ptr, nk, kp, nst, cp, 1.0 / kpoints.AllNkpt, op, ip)
        _grid_check(rc, h.ptr)
    end
end

function Calc_DM_Crystal_NonCollinear!(electron::CrystalBloch, kpoints::KPoints, system_grid::System_Grid,
                                       DM::Vector{Vector{Float64}})
    ucell, h = _kpoint_context()
    Calc_fnkCnk!(electron, kpoints)                    # Crystal_DFT.jl:369
    _dm_states_nc_call(h, kpoints, electron.Cnk[1], DM, nothing)
    return nothing
end

# nested iDM[s][atom][Rn][ist][jst] <-> flat pair order (atom, Rn, ist, jst)
function _flatten_nested(X, Natom, FNAN, Total_NumOrbs, natn)
    out = Float64[]
    for atom = 1:Natom, Rn = 1:FNAN[atom]+1, ist = 1:Total_NumOrbs[atom]
        append!(out, X[atom][Rn][ist])
    end
    return out
end

function _unflatten_nested!(X, flat, Natom, FNAN, Total_NumOrbs)
    st = 0
    for atom = 1:Natom, Rn = 1:FNAN[atom]+1, ist = 1:Total_NumOrbs[atom]
        row = X[atom][Rn][ist]
        n = length(row)
        copyto!(row, 1, flat, st + 1, n)
        st += n
    end
end

function Calc_iDM_Crystal_NonCollinear!(electron::CrystalBloch, kpoints::KPoints, system_grid::System_Grid, iDM)
    electron.cal_mode == 2 && error("LCPAOGridHIP_kpoint.jl: cal_mode 2 (sum only) stays the reference's")
    ucell, h = _kpoint_context()
    Natom, FNAN, NO = system_grid.Natom, system_grid.FNAN, system_grid.Total_NumOrbs
    hs = sum(NO[a] * NO[system_grid.natn[a][r]] for a in 1:Natom for r in 1:FNAN[a]+1)
    dm = [zeros(hs) for _ in 1:4]
    idm = [zeros(hs) for _ in 1:2]
    _dm_states_nc_call(h, kpoints, electron.Cnk[1], dm, idm)   # states already scaled by Calc_DM_Crystal_NonCollinear!
    _unflatten_nested!(iDM[1], idm[1], Natom, FNAN, NO)
    _unflatten_nested!(iDM[2], idm[2], Natom, FNAN, NO)
    return nothing
end

function HS_matrix_NC!(H, Hks::Vector{Vector{Float64}}, iHks::Vector{Vector{Vector{Vector{Vector{Float64}}}}},
                       Natom, Total_NumOrbs, MP, FNAN, natn, ncn, atv_ijk, kpts::AbstractVector{Float64})
    ucell, h = _kpoint_context()
    k = Float64[kpts[1], kpts[2], kpts[3]]
    ih = [_flatten_nested(iHks[s], Natom, FNAN, Total_NumOrbs, natn) for s in 1:3]
    hp = [pointer(Hks[s]) for s in 1:4]
    ip = [pointer(ih[s]) for s in 1:3]
    GC.@preserve H Hks ih hp ip k begin
        rc = ccall((:lcpao_grid_bloch_sum_nc, LIBGRID), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}, Ptr{Float64}),
                   h.ptr, k, hp, ip, Ptr{Float64}(pointer(H)))
        _grid_check(rc, h.ptr)
    end
    return nothing
end

# nested Hks[s][atom][Rn][ist][jst] (HS_matrix.jl:260): flattened in pair order
function HS_matrix_NC!(H, Hks::Vector{Vector{Vector{Vector{Vector{Float64}}}}},
                       iHks::Vector{Vector{Vector{Vector{Vector{Float64}}}}},
                       Natom, Total_NumOrbs, MP, FNAN, natn, ncn, atv_ijk, kpts::AbstractVector{Float64})
    flat = [_flatten_nested(Hks[s], Natom, FNAN, Total_NumOrbs, natn) for s in 1:4]
    HS_matrix_NC!(H, flat, iHks, Natom, Total_NumOrbs, MP, FNAN, natn, ncn, atv_ijk, kpts)
end

# the form without k (HS_matrix.jl:167): every phase is 1, i.e. the k form at
# k = 0.  The reference reads the fourth channel as Hks[4][ist][jst] (:199),
# which indexes a number for jst > 1; the flat Hks[4][hst] of its k form
# (:246) is taken here.
function HS_matrix_NC!(H, Hks::Vector{Vector{Float64}}, iHks::Vector{Vector{Vector{Vector{Vector{Float64}}}}},
                       Natom, Total_NumOrbs, MP, FNAN, natn)
    ucell, _ = _kpoint_context()
    sg = ucell.system_grid
    HS_matrix_NC!(H, Hks, iHks, Natom, Total_NumOrbs, MP, FNAN, natn, sg.ncn, sg.atv_ijk, Float64[0.0, 0.0, 0.0])
end
