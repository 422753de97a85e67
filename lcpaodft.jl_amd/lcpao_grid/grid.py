"""Host-side mirror of the reference's grid-path routines over the C ABI.

The functions keep the reference's names and argument order, with the
Julia mutation mark "!" dropped:

  Set_Orbitals_Grid(pao, ucell)                       Set_Orbitals_Grid.jl:26-42
  Set_dOrbitals_Grid(pao, ucell, Orbs_Grid)           Set_dOrbitals_Grid.jl:1-20
  Calc_MatrixElements_dVH_Vxc_{off,on,nc}(ucell, Orbs_Grid, Vpot_Grid, MPI_Hks)
                                                      Hamiltonian.jl:235-380
  Set_Density_Grid(SpinPol, ucell, Orbs_Grid, DM, Density_Grid) and
  Set_Density_Grid_{nonpol,pol,nc}(...)               Set_Density_Grid.jl:1-240
  Force3_{nospin,spin}(pao, Orbs_Grid, Vpot_Grid, DM, ucell),
  Force3_nc(pao, Orbs_Grid, dVHart_Grid, Vxc_Grid, DM, ucell)
                                                      Force.jl:407-566

Orbs_Grid is a GridContext: the tables live in device memory.  DM is the
global flat density matrix per spin (the layout of DM before
Set_DM2DM_Vec, Convert_Matrix.jl:27-62); Force3 takes it flat as well.
MPI_Hks / Density_Grid are filled in place (numpy arrays).  Errors raise
GridError.  There is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import GridError, check

WEIGHTS_FORCE3 = {1: (4.0,), 2: (2.0, 2.0), 4: (2.0, 2.0, 4.0, -4.0)}


def _ptrs(arrs, t=_lib.dp):
    return (t * len(arrs))(*[a.ctypes.data_as(t) if a is not None else None for a in arrs])


def _prec(p):
    if isinstance(p, str):
        if p not in _lib.PRECISIONS:
            raise ValueError("precision must be one of %s" % list(_lib.PRECISIONS))
        return _lib.PRECISIONS[p]
    return int(p)


def _geom(u, with_lattice=True, with_cell_lists=True):
    """lcpao_geom of a UCell (and the arrays it points into, to keep alive)."""
    keep = []
    no = np.ascontiguousarray(u.Total_NumOrbs, np.int32)
    gridn = np.ascontiguousarray(u.GridN_Atom, np.int32)
    glist = np.ascontiguousarray(u.glist, np.int32)
    clist = np.ascontiguousarray(u.clist, np.int32)
    olg1 = np.ascontiguousarray(u.olg1, np.int32)
    olg2 = np.ascontiguousarray(u.olg2, np.int32)
    keep += [no, gridn, glist, clist, olg1, olg2]
    isz = glist.itemsize
    base_g, base_c = glist.ctypes.data, clist.ctypes.data
    gptr = (_lib.ip * u.Natom)(*[C.cast(base_g + int(u.gl_off[a]) * isz, _lib.ip) for a in range(u.Natom)])
    cptr = (_lib.ip * u.Natom)(*[C.cast(base_c + int(u.gl_off[a]) * isz, _lib.ip) for a in range(u.Natom)])
    b1, b2 = olg1.ctypes.data, olg2.ctypes.data
    P = u.MPI_size
    o1p = (_lib.ip * max(P, 1))(*[C.cast(b1 + int(u.olg_off[p]) * isz, _lib.ip) for p in range(P)])
    o2p = (_lib.ip * max(P, 1))(*[C.cast(b2 + int(u.olg_off[p]) * isz, _lib.ip) for p in range(P)])
    pa = np.ascontiguousarray(u.MPI_atom, np.int32)
    pn = np.ascontiguousarray(u.MPI_natn, np.int32)
    pc = np.ascontiguousarray(u.MPI_ncn, np.int32)
    nolg = np.ascontiguousarray(u.MPI_NumOLG, np.int32)
    lat = np.ascontiguousarray(u.Latvecs, np.float64)
    org = np.ascontiguousarray(u.Grid_Origin, np.float64)
    gx = np.ascontiguousarray(u.Gxyz, np.float64)
    a2s = np.ascontiguousarray(u.atom2spe, np.int32)
    keep += [gptr, cptr, o1p, o2p, pa, pn, pc, nolg, lat, org, gx, a2s]
    g = _lib.Geom()
    g.natom, g.atom_base = u.Natom, 0
    g.ngrid[:] = list(u.Ngrid)
    g.cpycell, g.gridvol = u.CpyCell, u.GridVol
    g.no, g.gridn = no.ctypes.data_as(_lib.ip), gridn.ctypes.data_as(_lib.ip)
    g.grid_list = gptr
    if with_cell_lists:  # optional in the C ABI (needed by the block forms and device tables)
        g.cell_list = cptr
    g.npairs = P
    g.pair_atom, g.pair_natn, g.pair_ncn = pa.ctypes.data_as(_lib.ip), pn.ctypes.data_as(_lib.ip), pc.ctypes.data_as(_lib.ip)
    g.numolg, g.olg1, g.olg2 = nolg.ctypes.data_as(_lib.ip), o1p, o2p
    if with_lattice:
        g.latvecs, g.origin, g.gxyz = lat.ctypes.data_as(_lib.dp), org.ctypes.data_as(_lib.dp), gx.ctypes.data_as(_lib.dp)
        g.atom2spe = a2s.ctypes.data_as(_lib.ip)
    return g, keep


def block_layout(ucell):
    """Host-only check of the block-form limits for a geometry (no device):
    (fits, largest padded row count, largest sphere-image count, blocks)."""
    L = _lib.load()
    g, keep = _geom(ucell)
    mr, ms, nb = C.c_int32(), C.c_int32(), C.c_int64()
    rc = L.lcpao_grid_block_layout(C.byref(g), C.byref(mr), C.byref(ms), C.byref(nb))
    if rc not in (0, _lib.LCPAO_ERANGE):
        raise GridError(rc, "lcpao_grid_block_layout failed")
    return rc == 0, mr.value, ms.value, nb.value


class GridContext:
    """Device context of one geometry (lists, tables, work decomposition)."""

    def __init__(self, ucell, device=0, with_lattice=True, with_cell_lists=True):
        L = _lib.load()
        u = ucell
        self.ucell = u
        g, self._keep = _geom(u, with_lattice, with_cell_lists)
        no = np.ascontiguousarray(u.Total_NumOrbs, np.int32)
        gridn = np.ascontiguousarray(u.GridN_Atom, np.int32)
        self.h = C.c_void_p()
        rc = L.lcpao_grid_create(C.byref(self.h), C.byref(g), int(device))
        if rc != 0:
            raise GridError(rc, L.lcpao_grid_last_error(None).decode())
        self._keep = []  # the library copied everything
        self.L = L
        self.device = int(device)
        self._stream = None  # caller stream bound by set_stream, else None
        hs, nn, to, tg = (C.c_int64() for _ in range(4))
        L.lcpao_grid_sizes(self.h, C.byref(hs), C.byref(nn), C.byref(to), C.byref(tg))
        self.my_hsize, self.nn, self.total_olg, self.total_gridn = hs.value, nn.value, to.value, tg.value
        self.natom = u.Natom
        self.no = no.copy()
        self.gridn = gridn.copy()

    def block_forms(self):
        """-1 not decided yet, 1 the split forms run the block kernels, 0 the
        geometry is beyond the block limits and pair-list kernels serve them."""
        v = C.c_int32()
        self._check(self.L.lcpao_grid_block_forms(self.h, C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            self.L.lcpao_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        check(rc, self.h)

    # tables -----------------------------------------------------------------
    def set_tables(self, tabs):
        """tabs[a]: (GridN_a, NO_a) C-order == Julia NO_a x GridN_a column-major."""
        tabs = [np.ascontiguousarray(t, np.float64) for t in tabs]
        self._check(self.L.lcpao_grid_set_tables(self.h, _ptrs(tabs)))

    def set_dtables(self, dx, dy, dz):
        dx, dy, dz = ([np.ascontiguousarray(t, np.float64) for t in d] for d in (dx, dy, dz))
        self._check(self.L.lcpao_grid_set_dtables(self.h, _ptrs(dx), _ptrs(dy), _ptrs(dz)))

    def build_tables(self, paos, with_derivatives=False):
        """Tables generated on the device from the radial data of each species."""
        arr = (_lib.Pao * len(paos))()
        keep = []
        for s, p in enumerate(paos):
            arr[s].maxl = p.Spe_MaxL_Basis
            nb = list(p.Spe_Num_Basis) + [0] * (4 - len(p.Spe_Num_Basis))
            arr[s].num_basis[:] = nb
            arr[s].nmesh = p.Spe_Num_Mesh_PAO
            rv = np.ascontiguousarray(p.Spe_PAO_RV)
            keep.append(rv)
            arr[s].rv = rv.ctypes.data_as(_lib.dp)
            rw = (_lib.dp * 64)()
            for l in range(p.Spe_MaxL_Basis + 1):
                for q in range(p.Spe_Num_Basis[l]):
                    w = np.ascontiguousarray(p.Spe_PAO_RWF[l][q])
                    keep.append(w)
                    rw[l * 16 + q] = w.ctypes.data_as(_lib.dp)
            keep.append(rw)
            arr[s].rwf = C.cast(rw, _lib.dpp)
        self._check(self.L.lcpao_grid_build_tables(self.h, len(paos), arr, int(bool(with_derivatives))))

    # grid fields next to the contractions ------------------------------------
    @staticmethod
    def _radial(paos, vpss):
        """lcpao_atomic_radial per species (PAO with Spe_Atomic_Den, VPS)."""
        arr = (_lib.AtomicRadial * len(paos))()
        keep = []
        for s, (p, v) in enumerate(zip(paos, vpss)):
            fields = [np.ascontiguousarray(x, np.float64) for x in
                      (p.Spe_PAO_XV, p.Spe_PAO_RV, p.Spe_Atomic_Den, v.Spe_VPS_XV, v.Spe_VPS_RV, v.Spe_Atomic_PCC)]
            if fields[2].size != fields[0].size + 4 or fields[5].size != fields[3].size + 4:
                raise ValueError("atomic densities need nmesh + 4 values")
            keep += fields
            arr[s].nmesh_pao, arr[s].nmesh_vps = fields[0].size, fields[3].size
            arr[s].pao_xv, arr[s].pao_rv, arr[s].atomic_den = (f.ctypes.data_as(_lib.dp) for f in fields[:3])
            arr[s].vps_xv, arr[s].vps_rv, arr[s].atomic_pcc = (f.ctypes.data_as(_lib.dp) for f in fields[3:])
        return arr, keep

    @staticmethod
    def _spins(natom, init_nspin, init_angle):
        ns = np.ascontiguousarray(np.asarray(init_nspin, np.float64)[:, :2]).reshape(natom, 2)
        an = None if init_angle is None else np.ascontiguousarray(np.asarray(init_angle, np.float64)[:, :2]).reshape(natom, 2)
        return ns, an

    def aden_pcc(self, nspin, paos, vpss, init_nspin, init_angle=None):
        """(ADensity, PCCDensity, [Density per spin]) on the grid, as
        Set_AdenPCC_Grid returns them; paos carry Spe_Atomic_Den (read with
        the core charge), vpss the partial-core density."""
        arr, keep = self._radial(paos, vpss)
        ns, an = self._spins(self.natom, init_nspin, init_angle)
        A, P = np.zeros(self.nn), np.zeros(self.nn)
        R = [np.zeros(self.nn) for _ in range(nspin)]
        self._check(self.L.lcpao_grid_aden_pcc(self.h, nspin, len(paos), arr, ns.ctypes.data_as(_lib.dp),
                                               None if an is None else an.ctypes.data_as(_lib.dp),
                                               A.ctypes.data_as(_lib.dp), P.ctypes.data_as(_lib.dp), _ptrs(R)))
        return A, P, R

    def aden_pcc_dev(self, nspin, paos, vpss, init_nspin, dA, dP, dR, init_angle=None):
        arr, keep = self._radial(paos, vpss)
        ns, an = self._spins(self.natom, init_nspin, init_angle)
        self._order()
        self._check(self.L.lcpao_grid_aden_pcc_dev(self.h, nspin, len(paos), arr, ns.ctypes.data_as(_lib.dp),
                                                   None if an is None else an.ctypes.data_as(_lib.dp),
                                                   C.cast(dA.data_ptr(), _lib.dp), C.cast(dP.data_ptr(), _lib.dp),
                                                   self._dptrs(dR)))

    @staticmethod
    def _recvecs(recvecs):
        """Recvecs (rows = reciprocal vectors) in Julia's column-major order."""
        r = np.asarray(recvecs, np.float64).reshape(3, 3)
        return np.ascontiguousarray(r.T).reshape(-1)

    def poisson(self, rho, aden, recvecs, nspin=None):
        """dVHart of Solve_Poisson! from the spin densities rho (list) and
        the atomic density; nspin 0 (rho = [Density_xyz], aden None) runs the
        inner solver on rho[0] as given."""
        nspin = len(rho) if nspin is None else nspin
        R = [np.ascontiguousarray(r, np.float64) for r in rho]
        a = None if aden is None else np.ascontiguousarray(aden, np.float64)
        rec = self._recvecs(recvecs)
        V = np.zeros(self.nn)
        self._check(self.L.lcpao_grid_poisson(self.h, nspin, _ptrs(R), None if a is None else a.ctypes.data_as(_lib.dp),
                                              rec.ctypes.data_as(_lib.dp), V.ctypes.data_as(_lib.dp)))
        return V

    def poisson_dev(self, drho, daden, recvecs, dV, nspin=None):
        nspin = len(drho) if nspin is None else nspin
        rec = self._recvecs(recvecs)
        self._order()
        self._check(self.L.lcpao_grid_poisson_dev(self.h, nspin, self._dptrs(drho),
                                                  None if daden is None else C.cast(daden.data_ptr(), _lib.dp),
                                                  rec.ctypes.data_as(_lib.dp), C.cast(dV.data_ptr(), _lib.dp)))

    # either side of the k-point diagonalisation --------------------------------
    def bloch_sizes(self):
        fs, ng = C.c_int64(), C.c_int64()
        self._check(self.L.lcpao_grid_bloch_sizes(self.h, C.byref(fs), C.byref(ng)))
        return fs.value, ng.value

    def bloch_sum(self, kpt, flats):
        """HS_matrix! with a k-point: flat pair arrays (my_hsize) -> dense
        complex (fsize, fsize) matrices."""
        fsize, _ = self.bloch_sizes()
        F = [np.ascontiguousarray(f, np.float64) for f in flats]
        D = [np.zeros((fsize, fsize), np.complex128, order="F") for _ in F]
        k = np.ascontiguousarray(kpt, np.float64)
        self._check(self.L.lcpao_grid_bloch_sum(self.h, k.ctypes.data_as(_lib.dp), len(F), _ptrs(F),
                                                _ptrs([d.reshape(-1, order="F").view(np.float64) for d in D])))
        return D

    def bloch_sum_dev(self, kpt, dflats, ddense):
        """Device form: ddense are complex128 torch tensors (fsize, fsize)
        stored column-major (a transposed view of a contiguous tensor, or any
        tensor whose memory is in Julia's order)."""
        k = np.ascontiguousarray(kpt, np.float64)
        self._order()
        self._check(self.L.lcpao_grid_bloch_sum_dev(self.h, k.ctypes.data_as(_lib.dp), len(dflats),
                                                    self._dptrs(dflats), self._dptrs(ddense)))

    def dm_states(self, kpts, states, scale):
        """Calc_DM_Crystal_Collinear_nopol!: states per k-point (complex
        (fsize, nstates), columns scaled) -> flat DM (my_hsize)."""
        kp = np.ascontiguousarray(kpts, np.float64).reshape(-1, 3)
        cs = [np.asfortranarray(c, np.complex128) for c in states]
        nst = cs[0].shape[1]
        flat = [np.ascontiguousarray(c.reshape(-1, order="F")).view(np.float64) for c in cs]
        DM = np.zeros(self.my_hsize)
        self._check(self.L.lcpao_grid_dm_states(self.h, len(cs), kp.ctypes.data_as(_lib.dp), nst, _ptrs(flat),
                                                float(scale), DM.ctypes.data_as(_lib.dp)))
        return DM

    def dm_states_dev(self, kpts, dstates, nstates, scale, dDM):
        kp = np.ascontiguousarray(kpts, np.float64).reshape(-1, 3)
        self._order()
        self._check(self.L.lcpao_grid_dm_states_dev(self.h, kp.shape[0], kp.ctypes.data_as(_lib.dp), int(nstates),
                                                    self._dptrs(dstates), float(scale),
                                                    C.cast(dDM.data_ptr(), _lib.dp)))

    def bloch_sum_nc(self, kpt, hks, ihks):
        """HS_matrix_NC! with a k-point: four flat Hks channels and the three
        iHks arrays flattened in pair order -> dense complex (2 fsize, 2 fsize)."""
        fsize, _ = self.bloch_sizes()
        Hs = [np.ascontiguousarray(f, np.float64) for f in hks]
        Is = [np.ascontiguousarray(f, np.float64) for f in ihks]
        if len(Hs) != 4 or len(Is) != 3:
            raise ValueError("HS_matrix_NC needs four Hks and three iHks arrays")
        H = np.zeros((2 * fsize, 2 * fsize), np.complex128, order="F")
        k = np.ascontiguousarray(kpt, np.float64)
        self._check(self.L.lcpao_grid_bloch_sum_nc(self.h, k.ctypes.data_as(_lib.dp), _ptrs(Hs), _ptrs(Is),
                                                   H.reshape(-1, order="F").view(np.float64).ctypes.data_as(_lib.dp)))
        return H

    def bloch_sum_nc_dev(self, kpt, dhks, dihks, dH):
        """Device form; dH is a complex128 tensor (2 fsize, 2 fsize) in
        column-major memory."""
        k = np.ascontiguousarray(kpt, np.float64)
        self._order()
        self._check(self.L.lcpao_grid_bloch_sum_nc_dev(self.h, k.ctypes.data_as(_lib.dp), self._dptrs(dhks),
                                                       self._dptrs(dihks), C.cast(dH.data_ptr(), _lib.dp)))

    def dm_states_nc(self, kpts, states, scale, with_idm=False):
        """Calc_DM_Crystal_NonCollinear!: spinor states per k-point (complex
        (2 fsize, nstates), columns scaled) -> (DM[4], iDM[2] or None)."""
        kp = np.ascontiguousarray(kpts, np.float64).reshape(-1, 3)
        cs = [np.asfortranarray(c, np.complex128) for c in states]
        nst = cs[0].shape[1]
        flat = [np.ascontiguousarray(c.reshape(-1, order="F")).view(np.float64) for c in cs]
        DM = [np.zeros(self.my_hsize) for _ in range(4)]
        iDM = [np.zeros(self.my_hsize) for _ in range(2)] if with_idm else None
        self._check(self.L.lcpao_grid_dm_states_nc(self.h, len(cs), kp.ctypes.data_as(_lib.dp), nst, _ptrs(flat),
                                                   float(scale), _ptrs(DM), _ptrs(iDM) if with_idm else None))
        return DM, iDM

    def dm_states_nc_dev(self, kpts, dstates, nstates, scale, dDM, diDM=None):
        kp = np.ascontiguousarray(kpts, np.float64).reshape(-1, 3)
        self._order()
        self._check(self.L.lcpao_grid_dm_states_nc_dev(self.h, kp.shape[0], kp.ctypes.data_as(_lib.dp), int(nstates),
                                                       self._dptrs(dstates), float(scale), self._dptrs(dDM),
                                                       None if diDM is None else self._dptrs(diDM)))

    def set_positions(self, gxyz):
        """Atom positions (natom x 3, bohr) for the next build_tables; lists kept."""
        g = np.ascontiguousarray(gxyz, np.float64).reshape(-1, 3)
        self._check(self.L.lcpao_grid_set_positions(self.h, g.ctypes.data_as(_lib.dp)))

    def get_tables(self, which=0, atoms=None):
        """Device tables copied back (which: 0 phi, 1..3 d/dx, d/dy, d/dz);
        with `atoms`, only those atoms' tables (the others are None)."""
        keep = set(range(self.natom)) if atoms is None else {int(a) for a in atoms}
        out = [np.zeros((int(self.gridn[a]), int(self.no[a]))) if a in keep else None for a in range(self.natom)]
        self._check(self.L.lcpao_grid_get_tables(self.h, which, _ptrs(out)))
        return out

    # contractions on host arrays ----------------------------------------------
    def hamiltonian(self, V, precision="fp64"):
        V = [np.ascontiguousarray(v, np.float64) for v in V]
        H = [np.zeros(self.my_hsize) for _ in V]
        self._check(self.L.lcpao_grid_hamiltonian(self.h, len(V), _ptrs(V), _ptrs(H), _prec(precision)))
        return H

    def density(self, DM, dm_offset=0, precision="fp64"):
        DM = [np.ascontiguousarray(d, np.float64) for d in DM]
        R = [np.zeros(self.nn) for _ in DM]
        self._check(self.L.lcpao_grid_density(self.h, len(DM), _ptrs(DM), int(dm_offset), _ptrs(R), _prec(precision)))
        return R

    def force3(self, V, DM, weights=None, dm_offset=0, precision="fp64", force=None):
        V = [np.ascontiguousarray(v, np.float64) for v in V]
        DM = [np.ascontiguousarray(d, np.float64) for d in DM]
        w = np.ascontiguousarray(weights if weights is not None else WEIGHTS_FORCE3[len(V)], np.float64)
        F = np.zeros((3, self.natom)) if force is None else np.ascontiguousarray(force.T)  # column-major natom x 3
        self._check(self.L.lcpao_grid_force3(self.h, len(V), _ptrs(V), _ptrs(DM), int(dm_offset),
                                             w.ctypes.data_as(_lib.dp), F.ctypes.data_as(_lib.dp), _prec(precision)))
        return F.T.copy()

    # contractions on device arrays (torch tensors, inputs resident in HBM) ------
    @staticmethod
    def _dptrs(ts):
        return (_lib.dp * len(ts))(*[C.cast(t.data_ptr(), _lib.dp) for t in ts])

    TUNE_OPS = {"rho64": 0, "f364": 1, "rhobf": 2, "rho32": 3}
    TUNE_NCAND = {"rho64": 9, "f364": 2, "rhobf": 4, "rho32": 4}

    def tune(self, op, variant=-1):
        """Fix the kernel variant of a tunable operation ("rho64", "f364", "rhobf", "rho32"), or
        -1 to let the next call time the candidates and keep the fastest.
        FP64 density: 0 matrix cores, 1..5 vector FMA with unit shapes
        (waves x grid-list entries) 4x2048, 4x1024, 8x512, 2x1024, 4x512,
        6 block form with the orbital columns staged in LDS, 7 block form
        reading the columns from the tables;
        bf16x3 density: 0 single-buffered chunk planes, 1 double-buffered,
        2 double-buffered with the chunks loaded straight into LDS, 3 items
        of up to 8 super-tiles, two workgroups per CU;
        FP32 density: 0 pair lists 4x2048, 1 block form, 2 pair lists
        4x1024, 3 pair lists 8x512."""
        self._check(self.L.lcpao_grid_tune(self.h, self.TUNE_OPS[op], variant))

    def tuned(self, op):
        """(chosen variant or -1, measured ms of each candidate)."""
        v = C.c_int32()
        ms = (C.c_float * 16)()
        self._check(self.L.lcpao_grid_tuned(self.h, self.TUNE_OPS[op], C.byref(v), ms, 16))
        return v.value, [round(float(x), 4) for x in ms[:self.TUNE_NCAND[op]]]

    def tune_load(self, path):
        """Fix the variants stored for this geometry's key in the tuner cache
        file `path` (JSON); returns how many operations were set."""
        n = C.c_int32()
        self._check(self.L.lcpao_grid_tune_load(self.h, str(path).encode(), C.byref(n)))
        return n.value

    def tune_save(self, path):
        """Write this context's tuned variants into the cache file `path`."""
        self._check(self.L.lcpao_grid_tune_save(self.h, str(path).encode()))

    def set_stream(self, stream):
        """Launch on `stream` (a torch.cuda.Stream, or None for the context's
        own stream).  The default stream of torch is the HIP null stream, which
        the library takes as "own stream": work on it is then ordered by a
        wait before each device call."""
        handle = 0 if stream is None else int(stream.cuda_stream)
        self._check(self.L.lcpao_grid_set_stream(self.h, C.c_void_p(handle or None)))
        self._stream = stream if handle else None

    def _order(self):
        """The device calls read their inputs on the context's stream, which
        does not follow the torch stream that produced them: wait for that
        stream unless the context launches on it."""
        if self._stream is None:
            import torch
            torch.cuda.current_stream(self.device).synchronize()

    def hamiltonian_dev(self, dV, dH, precision="fp64"):
        self._order()
        self._check(self.L.lcpao_grid_hamiltonian_dev(self.h, len(dV), self._dptrs(dV), self._dptrs(dH), _prec(precision)))

    def density_dev(self, dDM, drho, dm_offset=0, precision="fp64"):
        self._order()
        self._check(self.L.lcpao_grid_density_dev(self.h, len(dDM), self._dptrs(dDM), int(dm_offset), self._dptrs(drho),
                                                  _prec(precision)))

    def force3_dev(self, dV, dDM, dforce, weights=None, dm_offset=0, precision="fp64"):
        w = np.ascontiguousarray(weights if weights is not None else WEIGHTS_FORCE3[len(dV)], np.float64)
        self._order()
        self._check(self.L.lcpao_grid_force3_dev(self.h, len(dV), self._dptrs(dV), self._dptrs(dDM), int(dm_offset),
                                                 w.ctypes.data_as(_lib.dp), C.cast(dforce.data_ptr(), _lib.dp),
                                                 _prec(precision)))

    def force3_dev_noalloc(self, dV, dDM, dforce, weights=None, dm_offset=0):
        """FP64 Force3 of force3_dev through the allocation-free entry point
        (its partial sums in the library's fixed per-GPU scratch array)."""
        w = np.ascontiguousarray(weights if weights is not None else WEIGHTS_FORCE3[len(dV)], np.float64)
        self._order()
        self._check(self.L.lcpao_grid_force3_dev_noalloc(self.h, len(dV), self._dptrs(dV), self._dptrs(dDM),
                                                         int(dm_offset), w.ctypes.data_as(_lib.dp),
                                                         C.cast(dforce.data_ptr(), _lib.dp)))


# ----------------------------------------------------------------------------
# reference-named entry points
# ----------------------------------------------------------------------------
def Set_Orbitals_Grid(pao, ucell, device=0):
    """Set_Orbitals_Grid (Set_Orbitals_Grid.jl:26-42): tables computed on the device."""
    ctx = GridContext(ucell, device)
    ctx.build_tables(pao, with_derivatives=False)
    return ctx


def Set_dOrbitals_Grid(pao, ucell, Orbs_Grid):
    """Set_dOrbitals_Grid (Set_dOrbitals_Grid.jl:1-20): gradient tables into the same context."""
    Orbs_Grid.build_tables(pao, with_derivatives=True)
    return Orbs_Grid


def _hks(Orbs_Grid, Vpot_Grid, MPI_Hks, nch, precision):
    if len(Vpot_Grid) < nch or len(MPI_Hks) < nch:
        raise ValueError("need %d channels" % nch)
    H = Orbs_Grid.hamiltonian(Vpot_Grid[:nch], precision)
    for c in range(nch):
        MPI_Hks[c][:] = H[c]


def Calc_MatrixElements_dVH_Vxc_off(ucell, Orbs_Grid, Vpot_Grid, MPI_Hks, precision="fp64"):
    _hks(Orbs_Grid, Vpot_Grid, MPI_Hks, 1, precision)


def Calc_MatrixElements_dVH_Vxc_on(ucell, Orbs_Grid, Vpot_Grid, MPI_Hks, precision="fp64"):
    _hks(Orbs_Grid, Vpot_Grid, MPI_Hks, 2, precision)


def Calc_MatrixElements_dVH_Vxc_nc(ucell, Orbs_Grid, Vpot_Grid, MPI_Hks, precision="fp64"):
    _hks(Orbs_Grid, Vpot_Grid, MPI_Hks, 4, precision)


def _dens(ucell, Orbs_Grid, DM, Density_Grid, nspin, precision):
    R = Orbs_Grid.density(DM[:nspin], ucell.MPHks_rank, precision)
    for s in range(nspin):
        Density_Grid[s][:] = R[s]


def Set_Density_Grid_nonpol(ucell, Orbs_Grid, DM, Density_Grid, precision="fp64"):
    _dens(ucell, Orbs_Grid, DM, Density_Grid, 1, precision)


def Set_Density_Grid_pol(ucell, Orbs_Grid, DM, Density_Grid, precision="fp64"):
    _dens(ucell, Orbs_Grid, DM, Density_Grid, 2, precision)


def Set_Density_Grid_nc(ucell, Orbs_Grid, DM, Density_Grid, precision="fp64"):
    _dens(ucell, Orbs_Grid, DM, Density_Grid, 4, precision)


def Set_Density_Grid(SpinPol, ucell, Orbs_Grid, DM, Density_Grid, precision="fp64"):
    """Dispatcher of Set_Density_Grid.jl:1-9."""
    {"off": Set_Density_Grid_nonpol, "on": Set_Density_Grid_pol, "nc": Set_Density_Grid_nc}[SpinPol](
        ucell, Orbs_Grid, DM, Density_Grid, precision)


def Force3_nospin(pao, Orbs_Grid, Vpot_Grid, DM, ucell, precision="fp64"):
    """VpotForce (natom x 3) = 4*GridVol*sum (Force.jl:407-451); the gradient
    tables must be in Orbs_Grid (Set_dOrbitals_Grid)."""
    return Orbs_Grid.force3([Vpot_Grid[0]], [DM[0]], WEIGHTS_FORCE3[1], ucell.MPHks_rank, precision)


def Force3_spin(pao, Orbs_Grid, Vpot_Grid, DM, ucell, precision="fp64"):
    return Orbs_Grid.force3(list(Vpot_Grid[:2]), list(DM[:2]), WEIGHTS_FORCE3[2], ucell.MPHks_rank, precision)


def Force3_nc(pao, Orbs_Grid, dVHart_Grid, Vxc_Grid, DM, ucell, precision="fp64"):
    V = [dVHart_Grid + Vxc_Grid[0], dVHart_Grid + Vxc_Grid[1], Vxc_Grid[2], Vxc_Grid[3]]
    return Orbs_Grid.force3(V, list(DM[:4]), WEIGHTS_FORCE3[4], ucell.MPHks_rank, precision)


def Set_AdenPCC_Grid(SpinPol, Init_Atoms_Nspin, Init_Atoms_Angle, pao, pspot, ucell, Orbs_Grid):
    """Set_AdenPCC_Grid (Set_AdenPCC_Grid.jl:1-142) on the device of the
    context Orbs_Grid: returns (ADensity_Grid, PCCDensity_Grid, Density_Grid)."""
    nspin = {"off": 1, "on": 2, "nc": 4}[SpinPol]
    angle = None if Init_Atoms_Angle is None else [list(x)[:2] for x in Init_Atoms_Angle]
    return Orbs_Grid.aden_pcc(nspin, pao, pspot, [list(x)[:2] for x in Init_Atoms_Nspin], angle)


def Solve_Poisson(SpinPol, Recvecs, Density_Grid, ADensity_Grid, dVHart_Grid, Orbs_Grid):
    """Solve_Poisson! (Poisson.jl:1-7): dVHart_Grid is overwritten."""
    nspin = {"off": 1, "on": 2, "nc": 4}[SpinPol]
    dVHart_Grid[:] = Orbs_Grid.poisson(list(Density_Grid[:min(nspin, 2)]), ADensity_Grid, Recvecs, nspin)


def HS_matrix(S, H, OLP, Hks, kpts, Orbs_Grid):
    """HS_matrix!(S, H, OLP::Vector{Float64}, Hks::Vector{Float64}, ..., kpts)
    (HS_matrix.jl:143-166) on the device of the context Orbs_Grid; S and H
    (complex fsize x fsize) are overwritten."""
    Sk, Hk = Orbs_Grid.bloch_sum(kpts, [OLP, Hks])
    S[:, :] = Sk
    H[:, :] = Hk


def HS_matrix_cluster(S, H, OLP, Hks, Orbs_Grid):
    """HS_matrix!(S, H, OLP::Vector{Float64}, Hks::Vector{Float64}, Natom, ...,
    natn) (HS_matrix.jl:114-136), the form without k: real S and H are
    overwritten with the sums at k = 0 (every phase exactly 1), H may be None
    for the one-matrix form (HS_matrix.jl:67-85)."""
    flats = [OLP] if H is None else [OLP, Hks]
    out = Orbs_Grid.bloch_sum([0.0, 0.0, 0.0], flats)
    S[:, :] = out[0].real
    if H is not None:
        H[:, :] = out[1].real


def Calc_DM_Crystal_Collinear_nopol(Cnk, MPI_kpts, AllNkpt, DM, Orbs_Grid):
    """The accumulation of Calc_DM_Crystal_Collinear_nopol! (Crystal_DFT.jl:219-270)
    for this rank's k-points (states scaled as Calc_fnkCnk! leaves them);
    DM[0] is overwritten with this rank's part (the reference then sums the
    ranks with MPI.Allreduce!)."""
    DM[0][:] = Orbs_Grid.dm_states(MPI_kpts, Cnk, 1.0 / AllNkpt)


def Calc_DM_Crystal_Collinear_pol(Cnk, MPI_kpts, AllNkpt, DM, Orbs_Grid):
    """The accumulation of Calc_DM_Crystal_Collinear_pol! (Crystal_DFT.jl:276-339):
    Cnk[s] are the scaled states of spin s per k-point; DM[0] and DM[1] are
    overwritten with this rank's part, one device pass per spin."""
    for s in range(2):
        DM[s][:] = Orbs_Grid.dm_states(MPI_kpts, Cnk[s], 1.0 / AllNkpt)


def Calc_DM_Crystal_NonCollinear(Cnk, MPI_kpts, AllNkpt, DM, Orbs_Grid):
    """The accumulation of Calc_DM_Crystal_NonCollinear! (Crystal_DFT.jl:342-410):
    Cnk are the scaled spinor states (2 fsize, nstates) per k-point; DM[0..3]
    (up-up, down-down, Re up-down, Im up-down) are overwritten."""
    out, _ = Orbs_Grid.dm_states_nc(MPI_kpts, Cnk, 1.0 / AllNkpt)
    for s in range(4):
        DM[s][:] = out[s]


def Calc_iDM_Crystal_NonCollinear(Cnk, MPI_kpts, AllNkpt, iDM, Orbs_Grid):
    """The k-point sums of Calc_iDM_Crystal_NonCollinear! (Crystal_DFT.jl:413-486,
    cal_mode != 2): iDM[0], iDM[1] (imaginary parts of the up-up and down-down
    blocks, the nested arrays flattened in pair order) are overwritten."""
    _, out = Orbs_Grid.dm_states_nc(MPI_kpts, Cnk, 1.0 / AllNkpt, with_idm=True)
    iDM[0][:] = out[0]
    iDM[1][:] = out[1]


def HS_matrix_NC(H, Hks, iHks, kpts, Orbs_Grid):
    """HS_matrix_NC!(H, Hks::Vector{Vector{Float64}}, iHks, ..., kpts)
    (HS_matrix.jl:211-257): Hks are the four flat channels, iHks the three
    nested arrays flattened in pair order; H (complex 2 fsize x 2 fsize) is
    overwritten (its down-up block zero, as in the reference)."""
    H[:, :] = Orbs_Grid.bloch_sum_nc(kpts, Hks, iHks)


def flatten_nested(X, ucell):
    """A nested per-pair array X[atom][Rn][ist][jst] (the layout of the
    reference's nested Hks / iHks / DM_Vec, Convert_Matrix.jl:27-62) as the
    flat pair-block array of the pairs of `ucell`, in pair order (atom, Rn,
    ist, jst)."""
    out = []
    for a in range(ucell.Natom):
        for r in range(int(ucell.nb_off[a + 1] - ucell.nb_off[a])):
            out.append(np.asarray(X[a][r], np.float64).reshape(-1))
    return np.concatenate(out) if out else np.zeros(0)


def HS_matrix_NC_nested(H, Hks, iHks, kpts, Orbs_Grid):
    """HS_matrix_NC!(H, Hks::Vector{Vector{Vector{Vector{Vector{Float64}}}}},
    iHks, ..., kpts) (HS_matrix.jl:260-305): the four Hks channels and the
    three iHks arrays are nested [atom][Rn][ist][jst]; flattened in pair
    order they are the flat form's inputs.  H (complex 2 fsize x 2 fsize) is
    overwritten."""
    u = Orbs_Grid.ucell
    H[:, :] = Orbs_Grid.bloch_sum_nc(kpts, [flatten_nested(x, u) for x in Hks], [flatten_nested(x, u) for x in iHks])


def HS_matrix_NC_cluster(H, Hks, iHks, Orbs_Grid):
    """HS_matrix_NC!(H, Hks::Vector{Vector{Float64}}, iHks, Natom, ..., natn)
    (HS_matrix.jl:167-208), the form without k: every phase is 1, so H is the
    flat form's sum at k = 0.  The reference reads the fourth channel as
    Hks[4][ist][jst] (:199), which indexes a number for jst > 1; the flat
    element Hks[4][hst] that its k form reads (:246) is taken here."""
    H[:, :] = Orbs_Grid.bloch_sum_nc([0.0, 0.0, 0.0], Hks, iHks)
