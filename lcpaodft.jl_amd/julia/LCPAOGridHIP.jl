# LCPAOGridHIP.jl -- drop-in methods for LCPAODFT.jl's real-space grid path,
# calling liblcpao_grid.so (MI355X / gfx950) through `ccall`.
#
# Meant to be `include`d inside `module LCPAODFT` after the reference's own
# definitions: each method below has the reference's exact signature and
# replaces its body.  Untested here: Julia is not available in the build
# environment; the C ABI underneath is tested from Python (tests/).
#
# Replaced methods (reference file:line):
#   Calc_MatrixElements_dVH_Vxc_{off,on,nc}!   Hamiltonian.jl:235,280,327
#   Set_Density_Grid_{nonpol,pol,nc}!          Set_Density_Grid.jl:60,112,172
#   Force3_{nospin,spin}, Force3_nc            Force.jl:407,454,506
#   Set_AdenPCC_Grid                           Set_AdenPCC_Grid.jl:1
#   Solve_Poisson!(SpinPol, dft_mixing, ...)   Poisson.jl:1
# Force3 receives the nested DM_Vec in the reference; the wrapper flattens it
# back to the global flat layout (Convert_Matrix.jl:1-24 order).
# The k-point routines (HS_matrix!, Calc_DM_Crystal_Collinear_nopol!) are in
# LCPAOGridHIP_kpoint.jl: they need a context that holds every pair.

const LIBGRID = get(ENV, "LCPAO_GRID_LIB", joinpath(@__DIR__, "..", "lcpao_grid", "liblcpao_grid.so"))
const GRID_PRECISION = Ref{Int32}(0)   # 0 FP64, 1 FP32, 2 split-BF16x3

struct LcpaoGeom
    natom::Int32
    atom_base::Int32
    ngrid::NTuple{3,Int32}
    cpycell::Int32
    gridvol::Float64
    no::Ptr{Int32}
    gridn::Ptr{Int32}
    grid_list::Ptr{Ptr{Int32}}
    cell_list::Ptr{Ptr{Int32}}
    npairs::Int64
    pair_atom::Ptr{Int32}
    pair_natn::Ptr{Int32}
    pair_ncn::Ptr{Int32}
    numolg::Ptr{Int32}
    olg1::Ptr{Ptr{Int32}}
    olg2::Ptr{Ptr{Int32}}
    latvecs::Ptr{Float64}
    origin::Ptr{Float64}
    gxyz::Ptr{Float64}
    atom2spe::Ptr{Int32}
end

mutable struct GridHandle
    ptr::Ptr{Cvoid}
    orbs::Any       # the Orbs_Grid whose tables were uploaded
    dorbs::Any
end

# Context lifetime: one live device context per process, that of the latest
# UCell.  The reference builds a new UCell per geometry (DFT.jl:219, every
# step of a geometry optimisation); when a grid routine first sees a new
# UCell the previous context is destroyed before the new one is created, so
# its tables, block plan and tiles (several GB at 216 atoms) are released
# at once.  UCell is an immutable struct, so it cannot key a WeakKeyDict;
# the identity test is `===`.  grid_release!() frees the context explicitly.
const LAST_GRID = Ref{Any}(nothing)   # (ucell, handle) of the live context: Solve_Poisson! has no ucell

function _grid_destroy!(h::GridHandle)
    if h.ptr != C_NULL
        ccall((:lcpao_grid_destroy, LIBGRID), Cvoid, (Ptr{Cvoid},), h.ptr)
        h.ptr = C_NULL
    end
    h.orbs = nothing
    h.dorbs = nothing
    return nothing
end

function grid_release!()
    last = LAST_GRID[]
    last === nothing || _grid_destroy!(last[2])
    LAST_GRID[] = nothing
    return nothing
end

function _grid_check(rc, ctx)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:lcpao_grid_last_error, LIBGRID), Cstring, (Ptr{Cvoid},), ctx))
    error("lcpao_grid error $rc: $msg")
end

function _grid_context(ucell::UCell)
    last = LAST_GRID[]
    last !== nothing && last[1] === ucell && return last[2]
    grid_release!()   # a new geometry: free the previous context first
    sg = ucell.system_grid
    glp = [pointer(v) for v in ucell.GridListAtom]
    clp = [pointer(v) for v in ucell.CellListAtom]
    o1p = [pointer(v) for v in ucell.MPI_GListTAtoms1]
    o2p = [pointer(v) for v in ucell.MPI_GListTAtoms2]
    lat = Matrix{Float64}(transpose(sg.Latvecs))       # row-major 3x3
    gxyz = reduce(vcat, sg.Gxyz)
    a2s = Int32.(sg.atom2spe .- 1)
    no = Int32.(sg.Total_NumOrbs)
    ctxref = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve glp clp o1p o2p lat gxyz a2s no ucell begin
        g = LcpaoGeom(sg.Natom, 1, Int32.(sg.Ngrid), sg.CpyCell, sg.GridVol, pointer(no),
                      pointer(ucell.GridN_Atom), pointer(glp), pointer(clp), sg.MPI_size,
                      pointer(sg.MPI_atom), pointer(sg.MPI_natn), pointer(sg.MPI_ncn),
                      pointer(ucell.MPI_NumOLG), pointer(o1p), pointer(o2p),
                      pointer(lat), pointer(sg.Grid_Origin), pointer(gxyz), pointer(a2s))
        rc = ccall((:lcpao_grid_create, LIBGRID), Cint, (Ref{Ptr{Cvoid}}, Ref{LcpaoGeom}, Cint), ctxref, g, -1)
        _grid_check(rc, C_NULL)
    end
    h = GridHandle(ctxref[], nothing, nothing)
    finalizer(_grid_destroy!, h)   # at process exit; a no-op after grid_release!()
    LAST_GRID[] = (ucell, h)
    return h
end

function _grid_tables!(h::GridHandle, Orbs_Grid)
    h.orbs === Orbs_Grid && return nothing
    ptrs = [pointer(m) for m in Orbs_Grid.data]
    GC.@preserve Orbs_Grid ptrs begin
        _grid_check(ccall((:lcpao_grid_set_tables, LIBGRID), Cint, (Ptr{Cvoid}, Ptr{Ptr{Float64}}), h.ptr, ptrs), h.ptr)
    end
    h.orbs = Orbs_Grid
    return nothing
end

function _grid_hks!(ucell, Orbs_Grid, Vpot_Grid, MPI_Hks, nch)
    h = _grid_context(ucell)
    _grid_tables!(h, Orbs_Grid)
    vp = [pointer(Vpot_Grid[c]) for c in 1:nch]
    hp = [pointer(MPI_Hks[c]) for c in 1:nch]
    GC.@preserve Vpot_Grid MPI_Hks vp hp begin
        rc = ccall((:lcpao_grid_hamiltonian, LIBGRID), Cint,
                   (Ptr{Cvoid}, Int32, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}, Int32), h.ptr, nch, vp, hp, GRID_PRECISION[])
        _grid_check(rc, h.ptr)
    end
    return nothing
end

Calc_MatrixElements_dVH_Vxc_off!(ucell::UCell, Orbs_Grid, Vpot_Grid::Vector{Vector{Float64}}, MPI_Hks) =
    _grid_hks!(ucell, Orbs_Grid, Vpot_Grid, MPI_Hks, 1)
Calc_MatrixElements_dVH_Vxc_on!(ucell::UCell, Orbs_Grid, Vpot_Grid::Vector{Vector{Float64}}, MPI_Hks) =
    _grid_hks!(ucell, Orbs_Grid, Vpot_Grid, MPI_Hks, 2)
Calc_MatrixElements_dVH_Vxc_nc!(ucell::UCell, Orbs_Grid, Vpot_Grid::Vector{Vector{Float64}}, MPI_Hks) =
    _grid_hks!(ucell, Orbs_Grid, Vpot_Grid, MPI_Hks, 4)

function _grid_density!(ucell, Orbs_Grid, DM, Density_Grid, nspin)
    comm = MPI.COMM_WORLD
    myrank = MPI.Comm_rank(comm)
    h = _grid_context(ucell)
    _grid_tables!(h, Orbs_Grid)
    dp = [pointer(DM[s]) for s in 1:nspin]
    rp = [pointer(Density_Grid[s]) for s in 1:nspin]
    offset = Int64(ucell.system_grid.MPHks[myrank+1])
    GC.@preserve DM Density_Grid dp rp begin
        rc = ccall((:lcpao_grid_density, LIBGRID), Cint,
                   (Ptr{Cvoid}, Int32, Ptr{Ptr{Float64}}, Int64, Ptr{Ptr{Float64}}, Int32),
                   h.ptr, nspin, dp, offset, rp, GRID_PRECISION[])
        _grid_check(rc, h.ptr)
    end
    for s in 1:nspin
        MPI.Allreduce!(Density_Grid[s], MPI.SUM, comm)   # Set_Density_Grid.jl:108
    end
    return nothing
end

Set_Density_Grid_nonpol!(ucell::UCell, Orbs_Grid, DM, Density_Grid) = _grid_density!(ucell, Orbs_Grid, DM, Density_Grid, 1)
Set_Density_Grid_pol!(ucell::UCell, Orbs_Grid, DM, Density_Grid) = _grid_density!(ucell, Orbs_Grid, DM, Density_Grid, 2)
Set_Density_Grid_nc!(ucell::UCell, Orbs_Grid, DM, Density_Grid) = _grid_density!(ucell, Orbs_Grid, DM, Density_Grid, 4)

# DM_Vec[s][atom][Rn][ist][jst] -> global flat DM[s] (Convert_Matrix.jl:1-24)
_flat_dm(DM_Vec, ucell) = Set_DM_Vec2DM(DM_Vec, ucell.system_grid)

function _grid_force3(pao, Orbs_Grid, Vs, DM_Vec, ucell, weights)
    comm = MPI.COMM_WORLD
    myrank = MPI.Comm_rank(comm)
    h = _grid_context(ucell)
    _grid_tables!(h, Orbs_Grid)
    if h.dorbs === nothing
        dOrbs = Set_dOrbitals_Grid(pao, ucell)         # Force.jl:427
        dx = [pointer(m) for m in dOrbs[1].data]
        dy = [pointer(m) for m in dOrbs[2].data]
        dz = [pointer(m) for m in dOrbs[3].data]
        GC.@preserve dOrbs dx dy dz begin
            rc = ccall((:lcpao_grid_set_dtables, LIBGRID), Cint,
                       (Ptr{Cvoid}, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}), h.ptr, dx, dy, dz)
            _grid_check(rc, h.ptr)
        end
        h.dorbs = dOrbs
    end
    DM = _flat_dm(DM_Vec, ucell)
    nch = length(Vs)
    VpotForce = zeros(Float64, ucell.system_grid.Natom, 3)
    vp = [pointer(Vs[c]) for c in 1:nch]
    dp = [pointer(DM[c]) for c in 1:nch]
    w = Float64.(weights)
    offset = Int64(ucell.system_grid.MPHks[myrank+1])
    GC.@preserve Vs DM vp dp w VpotForce begin
        rc = ccall((:lcpao_grid_force3, LIBGRID), Cint,
                   (Ptr{Cvoid}, Int32, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}, Int64, Ptr{Float64}, Ptr{Float64}, Int32),
                   h.ptr, nch, vp, dp, offset, w, VpotForce, GRID_PRECISION[])
        _grid_check(rc, h.ptr)
    end
    MPI.Allreduce!(VpotForce, MPI.SUM, comm)          # Force.jl:447
    return VpotForce
end

Force3_nospin(pao::Vector{PAO}, Orbs_Grid, Vpot_Grid, DM, ucell::UCell) =
    _grid_force3(pao, Orbs_Grid, [Vpot_Grid[1]], DM, ucell, (4.0,))
Force3_spin(pao::Vector{PAO}, Orbs_Grid, Vpot_Grid, DM, ucell::UCell) =
    _grid_force3(pao, Orbs_Grid, [Vpot_Grid[1], Vpot_Grid[2]], DM, ucell, (2.0, 2.0))
Force3_nc(pao::Vector{PAO}, Orbs_Grid, dVHart_Grid, Vxc_Grid, DM, ucell::UCell) =
    _grid_force3(pao, Orbs_Grid, [dVHart_Grid + Vxc_Grid[1], dVHart_Grid + Vxc_Grid[2], Vxc_Grid[3], Vxc_Grid[4]],
                 DM, ucell, (2.0, 2.0, 4.0, -4.0))

# ---- Set_AdenPCC_Grid (Set_AdenPCC_Grid.jl:1-142) ---------------------------
# Every atom contributes on every rank, so the result equals the reference's
# after its MPI.Allreduce! calls; no reduction follows here.
struct LcpaoAtomicRadial
    nmesh_pao::Int32
    pao_xv::Ptr{Float64}
    pao_rv::Ptr{Float64}
    atomic_den::Ptr{Float64}
    nmesh_vps::Int32
    vps_xv::Ptr{Float64}
    vps_rv::Ptr{Float64}
    atomic_pcc::Ptr{Float64}
end

function Set_AdenPCC_Grid(SpinPol::AbstractString, Init_Atoms_Nspin, Init_Atoms_Angle, pao::Vector{PAO},
                          pspot::Vector{Pspot}, ucell::UCell)
    h = _grid_context(ucell)
    Nspin = SpinPol == "off" ? 1 : SpinPol == "on" ? 2 : SpinPol == "nc" ? 4 : error("please check SpinPol")
    natom = ucell.system_grid.Natom
    NN = prod(ucell.system_grid.Ngrid)
    rad = [LcpaoAtomicRadial(p.Spe_Num_Mesh_PAO, pointer(p.Spe_PAO_XV), pointer(p.Spe_PAO_RV), pointer(p.Spe_Atomic_Den),
                             v.Spe_Num_Mesh_VPS, pointer(v.Spe_VPS_XV), pointer(v.Spe_VPS_RV), pointer(v.Spe_Atomic_PCC))
           for (p, v) in zip(pao, pspot)]
    ins = Float64[x for a in 1:natom for x in (Init_Atoms_Nspin[a][1], Init_Atoms_Nspin[a][2])]
    ang = Float64[x for a in 1:natom for x in (Init_Atoms_Angle[a][1], Init_Atoms_Angle[a][2])]
    ADensity_Grid = zeros(Float64, NN)
    PCCDensity_Grid = zeros(Float64, NN)
    Density_Grid = [zeros(Float64, NN) for _ in 1:Nspin]
    rp = [pointer(d) for d in Density_Grid]
    GC.@preserve pao pspot rad ins ang ADensity_Grid PCCDensity_Grid Density_Grid rp begin
        rc = ccall((:lcpao_grid_aden_pcc, LIBGRID), Cint,
                   (Ptr{Cvoid}, Int32, Int32, Ptr{LcpaoAtomicRadial}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                    Ptr{Float64}, Ptr{Ptr{Float64}}),
                   h.ptr, Nspin, length(rad), rad, ins, ang, ADensity_Grid, PCCDensity_Grid, rp)
        _grid_check(rc, h.ptr)
    end
    return ADensity_Grid, PCCDensity_Grid, Density_Grid
end

# ---- Solve_Poisson! (Poisson.jl:1-108) -------------------------------------
# Runs on the context of the latest UCell (the method receives no UCell); its
# grid must be dft_mixing.Ngrid.
function Solve_Poisson!(SpinPol::String, dft_mixing::Rhor_Mixing, Density_Grid, ADensity_Grid, dVHart_Grid)
    last = LAST_GRID[]
    last === nothing && error("Solve_Poisson!: no grid context yet (build one through a grid routine first)")
    ucell, h = last
    Tuple(ucell.system_grid.Ngrid) == Tuple(dft_mixing.Ngrid) || error("Solve_Poisson!: grid differs from the context's")
    nspin = SpinPol == "off" ? 1 : 2
    rp = [pointer(Density_Grid[s]) for s in 1:nspin]
    rec = vec(Matrix{Float64}(dft_mixing.Recvecs))      # column-major, as the library expects
    GC.@preserve Density_Grid ADensity_Grid dVHart_Grid rp rec begin
        rc = ccall((:lcpao_grid_poisson, LIBGRID), Cint,
                   (Ptr{Cvoid}, Int32, Ptr{Ptr{Float64}}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   h.ptr, nspin, rp, ADensity_Grid, rec, dVHart_Grid)
        _grid_check(rc, h.ptr)
    end
    return nothing
end
