"""Bases and geometries beyond the block forms' limits, on the GPU.

  * Si2 with the 41-orbital Si7.0-s3p3d3f2 basis (16^3): device tables
    against the oracle; FP64 H, rho and Force3 against the oracle within
    1e-12 (the FP64 H kernel in 32 x 32 chunks, the vector-FMA density and
    Force3 kernels); the FP32 and bf16x3 forms, beyond the block forms'
    32 orbitals per atom, run the pair-list kernels (block_forms() == 0)
    and meet their bounds; Bloch sum and density matrix from states in
    32 x 32 tiles against the oracle.
  * One Si atom in a 3.5 bohr cubic cell (14^3): 59 sphere images and 768
    padded rows on a block, beyond the block forms; every form against the
    oracle.
"""
import numpy as np
import pytest
import torch

import oracle as O
from lcpao_grid import GridError
from lcpao_grid import systems as S
from lcpao_grid.grid import GridContext, WEIGHTS_FORCE3, block_layout
from tolerance import within, worst_ratio

pytestmark = pytest.mark.gpu

RTOL = {"fp64": 1e-12, "fp32": 1e-5, "bf16x3": 1e-4}


class Case:
    def __init__(self, st, nwaves, nocc, seed):
        self.st = st
        self.u = st.ucell()
        self.g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
        paos = [O.Pao(*S.BASES[s]) for s in st.species_names]
        self.tabs = O.orbitals(self.g, paos, st.atom2spe)
        self.dt = O.dorbitals(self.g, paos, st.atom2spe)
        self.V, self.DM = S.inputs(st, self.u, nwaves=nwaves, nocc=nocc, seed=seed)

    def force_bound(self, Vs, Ds):
        absd = [[np.abs(x) for x in d] for d in self.dt]
        atabs = [np.abs(t) for t in self.tabs]
        w = WEIGHTS_FORCE3[len(Vs)]
        return sum(abs(w[k]) * O.force3_sums(self.g, atabs, absd, np.abs(Vs[k]), np.abs(Ds[k]))
                   for k in range(len(Vs))) * self.g.gridvol


_cache = {}


def case(name):
    if name not in _cache:
        _cache[name] = {"si2f": lambda: Case(S.si2_large_basis(16), 10, 6, 2),
                        "dense": lambda: Case(S.dense_cell(), 8, 3, 1)}[name]()
    return _cache[name]


def _check(got, ref, mag, prec, what):
    r = worst_ratio(got, ref, mag)
    print(what, prec, "worst ratio %.3g" % r, flush=True)
    if prec == "fp64":
        ok, msg = within(got, ref, mag)
        assert ok, what + ": " + msg
    assert r <= RTOL[prec], what


def test_large_basis_tables_on_device():
    c = case("si2f")
    ctx = GridContext(c.u)
    ctx.build_tables(c.st.paos(), with_derivatives=True)
    for which, ref in ((0, c.tabs), (1, c.dt[0]), (2, c.dt[1]), (3, c.dt[2])):
        got = ctx.get_tables(which)
        for a in range(c.u.Natom):
            assert got[a].shape == (c.u.GridN_Atom[a], 41)
            assert np.abs(got[a] - ref[a]).max() <= (1e-14 if which == 0 else 1e-13) * np.abs(ref[a]).max()


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("name", ["si2f", "dense"])
def test_beyond_block_limits_all_forms(name, prec):
    c = case(name)
    assert not block_layout(c.u)[0]
    ctx = GridContext(c.u)
    ctx.set_tables(c.tabs)
    ctx.set_dtables(*c.dt)
    V2 = 0.5 * c.V + 0.01
    H = ctx.hamiltonian([c.V, V2], prec)
    Href, M = O.hamiltonian(c.g, c.tabs, [c.V, V2], with_abs=True)
    for k in range(2):
        _check(H[k], Href[k], M[k], prec, "%s H[%d]" % (name, k))
    rho = ctx.density([c.DM], precision=prec)[0]
    Rref, MR = O.density(c.g, c.tabs, [c.DM], with_abs=True)
    _check(rho, Rref[0], MR[0], prec, "%s rho" % name)
    F = ctx.force3([c.V], [c.DM], precision=prec)
    Fref = O.force3(c.g, c.tabs, c.dt, [c.V], [c.DM])
    _check(F, Fref, c.force_bound([c.V], [c.DM]), prec, "%s Force3" % name)
    if prec != "fp64":
        assert ctx.block_forms() == 0


def test_large_basis_fp64_matrix_core_candidates_refuse():
    """The FP64 matrix-core density and Force3 kernels hold 32 orbitals per
    atom: fixed to them, a 41-orbital basis is refused; left open, the
    vector-FMA kernels run (tested above)."""
    c = case("si2f")
    ctx = GridContext(c.u)
    ctx.set_tables(c.tabs)
    ctx.set_dtables(*c.dt)
    ctx.tune("rho64", 0)
    with pytest.raises(GridError):
        ctx.density([c.DM])
    ctx.tune("f364", 0)
    with pytest.raises(GridError):
        ctx.force3([c.V], [c.DM])
    ctx.tune("rho64", -1)
    ctx.density([c.DM])
    # the matrix-core kernel (0) and the block forms (6, 7, 8: no block plan
    # beyond 32 orbitals) are not run; one of the vector-FMA pair-list shapes wins
    v, ms = ctx.tuned("rho64")
    assert v in range(1, 6) and ms[0] == -1 and ms[6] == ms[7] == ms[8] == -1 and all(m > 0 for m in ms[1:6])


def test_large_basis_bloch_sum_and_dm_in_tiles():
    c = case("si2f")
    ctx = GridContext(c.u)
    fsize, ngroups = ctx.bloch_sizes()
    assert fsize == 82
    rng = np.random.default_rng(3)
    flats = [rng.standard_normal(c.u.myHsize) for _ in range(2)]
    k = [0.13, -0.22, 0.31]
    got = ctx.bloch_sum(k, flats)
    want = O.hs_matrix(c.g, k, flats)
    for a, b in zip(got, want):
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()
    kpts = np.array([[0.0, 0.0, 0.0], [0.25, 0.1, -0.2]])
    Cs = [rng.standard_normal((fsize, 9)) + 1j * rng.standard_normal((fsize, 9)) for _ in kpts]
    DM = ctx.dm_states(kpts, Cs, 0.5)
    DMo = O.dm_collinear(c.g, kpts, Cs, 0.5)
    assert np.abs(DM - DMo).max() <= 1e-13 * np.abs(DMo).max()


def test_large_basis_noalloc_force3():
    c = case("si2f")
    ctx = GridContext(c.u)
    ctx.set_tables(c.tabs)
    ctx.set_dtables(*c.dt)
    dV = torch.tensor(c.V, device="cuda")
    dD = torch.tensor(c.DM, device="cuda")
    f = torch.zeros(3 * c.u.Natom, dtype=torch.float64, device="cuda")
    ctx.force3_dev_noalloc([dV], [dD], f)
    F = f.cpu().numpy().reshape(3, c.u.Natom).T
    Fref = O.force3(c.g, c.tabs, c.dt, [c.V], [c.DM])
    ok, msg = within(F, Fref, c.force_bound([c.V], [c.DM]))
    assert ok, msg
