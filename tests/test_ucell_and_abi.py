"""CPU tests of the product's host code: the UCell list builder equals the
oracle's literal restatement of the reference (exact integers), the PAO
reader equals the oracle's, and the C-ABI library loads and exports every
symbol that include/lcpao_grid.h declares (no compute calls without a GPU).
"""
import ctypes
import json
import os
import re

import numpy as np
import pytest

import oracle as O
from lcpao_grid import _lib, systems as S
from lcpao_grid.pao import PAO
from lcpao_grid.ucell import UCell

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _same(u, g):
    pairs = [(u.FNAN, g.fnan), (u.natn_flat, g.natn), (u.ncn_flat, g.ncn), (u.GridN_Atom, g.gridn),
             (u.glist, g.glist), (u.clist, g.clist), (u.MPI_atom, g.pair_atom), (u.MPI_natn, g.pair_natn),
             (u.MPI_ncn, g.pair_ncn), (u.pair_hoff, g.pair_hoff), (u.MPI_NumOLG, g.numolg), (u.olg1, g.olg1),
             (u.olg2, g.olg2)]
    return all(np.array_equal(a, b) for a, b in pairs) and u.CpyCell == g.cpycell and u.GridVol == g.gridvol \
        and u.Total_Hsize == g.total_hsize and u.MPHks_rank == g.my_hoff


STRUCTS = {
    "si2_g16": lambda: S.si2(16),
    "si8_g20": lambda: S.si(1, 20),
    "nacl8_g18": lambda: S.nacl(1, 18),
    "si2_tri": lambda: S.Structure("tri", [[7.1, 0.3, 0.2], [0.5, 6.8, -0.4], [0.1, 0.6, 7.3]],
                                   [[0.05, 0.1, 0.02], [0.3, 0.26, 0.21]], ["Si", "Na"], (15, 14, 13), 7.0),
}


@pytest.mark.parametrize("name", list(STRUCTS))
@pytest.mark.parametrize("nprocs,rank", [(1, 0), (3, 2)])
def test_ucell_builder_equals_oracle(name, nprocs, rank):
    st = STRUCTS[name]()
    u = st.ucell(nprocs, rank, nthreads=4)
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid, (0, 0, 0), nprocs, rank)
    assert _same(u, g)


def test_ucell_builder_golden():
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "ucell_si2_g3.json")))
    inp = gold["inputs"]
    u = UCell(1, inp["Latvecs"], 2, [0, 0], inp["Gxyz"], inp["Atom_Cut1"], inp["Ngrid"], inp["Grid_Origin"])
    assert u.CpyCell == gold["CpyCell"] and list(u.FNAN) == gold["FNAN"]
    assert [list(x + 1) for x in u.natn] == gold["natn"] and [list(x) for x in u.ncn] == gold["ncn"]
    assert list(u.GridN_Atom) == gold["GridN_Atom"]
    assert [list(x) for x in u.GridListAtom] == gold["GridListAtom"]
    assert [list(x) for x in u.CellListAtom] == gold["CellListAtom"]


def test_ucell_builder_errors():
    with pytest.raises(_lib.GridError):
        UCell(1, np.eye(3) * 10, 2, [0, 0], [[0, 0, 0], [1, 1, 1]], [7.0, 7.0], (8, 8, 8), nprocs=10 ** 6)
    with pytest.raises(_lib.GridError):
        UCell(1, np.eye(3) * 10, 1, [0], [[0, 0, 0]], [7.0], (0, 8, 8))


def test_ucell_one_pair_per_rank_boundary():
    """Si2 at 16^3 has 174 pairs: 174 ranks each get exactly one, in order,
    and 175 ranks are refused as the reference refuses them (UCell.jl:52)."""
    st = S.si2(16)
    full = st.ucell()
    assert full.MPI_size == 174
    got = []
    for r in (0, 1, 96, 97, 172, 173):
        u = st.ucell(174, r)
        assert u.MPI_size == 1
        got.append((int(u.MPI_atom[0]), int(u.MPI_natn[0]), int(u.MPI_ncn[0])))
    ref = [(int(full.MPI_atom[r]), int(full.MPI_natn[r]), int(full.MPI_ncn[r])) for r in (0, 1, 96, 97, 172, 173)]
    assert got == ref
    with pytest.raises(_lib.GridError):
        st.ucell(175, 0)


@pytest.mark.parametrize("sym,cut,basis", [("Si", 7.0, "s2p2d1"), ("Ga", 7.0, "s3p2d2"), ("Na", 9.0, "s3p2d1")])
def test_pao_reader_equals_oracle(sym, cut, basis):
    p, q = PAO(sym, cut, basis), O.Pao(sym, cut, basis)
    assert p.Spe_Total_NumOrbs == q.total_orbs and p.Spe_Num_Mesh_PAO == q.nmesh
    assert np.array_equal(p.Spe_PAO_RV, q.rv)
    for l in range(p.Spe_MaxL_Basis + 1):
        for m in range(p.Spe_Num_Basis[l]):
            assert np.array_equal(p.Spe_PAO_RWF[l][m], q.rwf(l, m))
    with pytest.raises(ValueError):
        PAO(sym, cut, "s2p2d")


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    hdr = open(os.path.join(ROOT, "include", "lcpao_grid.h")).read()
    declared = set(re.findall(r"^(?:int|void|const char \*)\s*(lcpao_\w+)\(", hdr, re.M))
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert getattr(L, name) is not None


def test_create_without_device_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = _lib.load()
    g = _lib.Geom()
    h = ctypes.c_void_p()
    rc = L.lcpao_grid_create(ctypes.byref(h), ctypes.byref(g), 0)
    assert rc != 0 and L.lcpao_grid_last_error(None)
    rc = L.lcpao_grid_create(ctypes.byref(h), None, 0)
    assert rc == _lib.LCPAO_EINVAL


def test_tune_rejects_null_context():
    L = _lib.load()
    assert L.lcpao_grid_tune(None, 0, 0) == _lib.LCPAO_EINVAL
    v = ctypes.c_int32()
    assert L.lcpao_grid_tuned(None, 0, ctypes.byref(v), None, 0) == _lib.LCPAO_EINVAL


def test_loader_keeps_one_hip_runtime_in_the_process():
    """Loading the library maps exactly one HIP and one HSA runtime (torch's,
    when torch is installed), whatever is imported afterwards."""
    import subprocess
    import sys
    code = ("import re, sys; sys.path.insert(0, %r); from lcpao_grid import _lib; _lib.load(); import torch; "
            "m = {l.split()[-1] for l in open('/proc/self/maps') if re.search('libamdhip64|libhsa-runtime', l)}; "
            "print(len(m))") % os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lcpaodft.jl_amd")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().splitlines()[-1] == "2"
