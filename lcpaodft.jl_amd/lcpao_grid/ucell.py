"""UCell lists (UCell.jl:1-46, 210-294) built by the library's host code.

``UCell(Nspin, Latvecs, Natom, atom2spe, Gxyz, Atoms_Cut1, Ngrid,
Grid_Origin; Total_NumOrbs)`` keeps the reference's argument order.  Lists
are stored flat (CSR) and exposed under the reference's field names as
lists of arrays; atom indices inside them are 0-based.
"""
import ctypes as C

import numpy as np

from . import _lib


class UCell:
    def __init__(self, Nspin, Latvecs, Natom, atom2spe, Gxyz, Atoms_Cut1, Ngrid, Grid_Origin=(0.0, 0.0, 0.0),
                 Total_NumOrbs=None, nprocs=1, rank=0, nthreads=0):
        L = _lib.load()
        self.Nspin = Nspin
        self.Latvecs = np.ascontiguousarray(Latvecs, dtype=np.float64).reshape(3, 3)
        self.Natom = int(Natom)
        self.atom2spe = np.ascontiguousarray(atom2spe, dtype=np.int32)
        self.Gxyz = np.ascontiguousarray(Gxyz, dtype=np.float64).reshape(self.Natom, 3)
        self.Atoms_Cut1 = np.ascontiguousarray(Atoms_Cut1, dtype=np.float64)
        self.Ngrid = tuple(int(x) for x in Ngrid)
        self.Grid_Origin = np.ascontiguousarray(Grid_Origin, dtype=np.float64)
        no = np.zeros(self.Natom, np.int32) if Total_NumOrbs is None else np.ascontiguousarray(Total_NumOrbs, dtype=np.int32)
        self.Total_NumOrbs = no
        ng = np.ascontiguousarray(self.Ngrid, dtype=np.int32)
        h = C.c_void_p()
        rc = L.lcpao_ucell_build(C.byref(h), self.Latvecs.ctypes.data_as(_lib.dp), self.Natom,
                                 self.Gxyz.ctypes.data_as(_lib.dp), self.Atoms_Cut1.ctypes.data_as(_lib.dp),
                                 no.ctypes.data_as(_lib.ip), ng.ctypes.data_as(_lib.ip),
                                 self.Grid_Origin.ctypes.data_as(_lib.dp), nprocs, rank, nthreads)
        if rc != 0:
            raise _lib.GridError(rc, L.lcpao_ucell_last_error().decode())
        try:
            cpy = C.c_int32()
            nloop, npairs, ths, mhs, mho, tgn, tolg = (C.c_int64() for _ in range(7))
            gv = C.c_double()
            L.lcpao_ucell_scalars(h, C.byref(cpy), C.byref(nloop), C.byref(npairs), C.byref(ths), C.byref(mhs),
                                  C.byref(mho), C.byref(gv), C.byref(tgn), C.byref(tolg))
            self.CpyCell, self.Nloop, self.MPI_size = cpy.value, nloop.value, npairs.value
            self.Total_Hsize, self.myHsize, self.MPHks_rank = ths.value, mhs.value, mho.value
            self.GridVol = gv.value
            n = self.Natom
            self.FNAN = np.zeros(n, np.int32)
            self.nb_off = np.zeros(n + 1, np.int64)
            self.GridN_Atom = np.zeros(n, np.int32)
            self.gl_off = np.zeros(n + 1, np.int64)
            # neighbour count is known only after FNAN; copy in two steps
            L.lcpao_ucell_copy(h, self.FNAN.ctypes.data_as(_lib.ip), self.nb_off.ctypes.data_as(_lib.lp), None, None,
                               self.GridN_Atom.ctypes.data_as(_lib.ip), self.gl_off.ctypes.data_as(_lib.lp),
                               *([None] * 10))
            nnb = int(self.nb_off[-1])
            self.natn_flat = np.zeros(nnb, np.int32)
            self.ncn_flat = np.zeros(nnb, np.int32)
            self.glist = np.zeros(tgn.value, np.int32)
            self.clist = np.zeros(tgn.value, np.int32)
            P = npairs.value
            self.MPI_atom = np.zeros(P, np.int32)
            self.MPI_natn = np.zeros(P, np.int32)
            self.MPI_ncn = np.zeros(P, np.int32)
            self.pair_hoff = np.zeros(P + 1, np.int64)
            self.MPI_NumOLG = np.zeros(P, np.int32)
            self.olg_off = np.zeros(P + 1, np.int64)
            self.olg1 = np.zeros(tolg.value, np.int32)
            self.olg2 = np.zeros(tolg.value, np.int32)
            a = lambda x, t: x.ctypes.data_as(t)
            L.lcpao_ucell_copy(h, None, None, a(self.natn_flat, _lib.ip), a(self.ncn_flat, _lib.ip), None, None,
                               a(self.glist, _lib.ip), a(self.clist, _lib.ip), a(self.MPI_atom, _lib.ip),
                               a(self.MPI_natn, _lib.ip), a(self.MPI_ncn, _lib.ip), a(self.pair_hoff, _lib.lp),
                               a(self.MPI_NumOLG, _lib.ip), a(self.olg_off, _lib.lp), a(self.olg1, _lib.ip),
                               a(self.olg2, _lib.ip))
        finally:
            L.lcpao_ucell_destroy(h)
        self.nprocs, self.rank = nprocs, rank
        self.NN = int(np.prod(self.Ngrid))

    # reference-shaped views ------------------------------------------------
    @property
    def natn(self):
        return [self.natn_flat[self.nb_off[a]:self.nb_off[a + 1]] for a in range(self.Natom)]

    @property
    def ncn(self):
        return [self.ncn_flat[self.nb_off[a]:self.nb_off[a + 1]] for a in range(self.Natom)]

    @property
    def GridListAtom(self):
        return [self.glist[self.gl_off[a]:self.gl_off[a + 1]] for a in range(self.Natom)]

    @property
    def CellListAtom(self):
        return [self.clist[self.gl_off[a]:self.gl_off[a + 1]] for a in range(self.Natom)]

    @property
    def MPI_GListTAtoms1(self):
        return [self.olg1[self.olg_off[p]:self.olg_off[p + 1]] for p in range(self.MPI_size)]

    @property
    def MPI_GListTAtoms2(self):
        return [self.olg2[self.olg_off[p]:self.olg_off[p + 1]] for p in range(self.MPI_size)]

    @property
    def GridVolume(self):
        return self.GridVol
