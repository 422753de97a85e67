"""Pins the CPU oracle (oracle/lcpao_oracle.c) before it is trusted as the
checker of the HIP library.

  UCell lists / GridVol        test_UCell.jl:19-50 golden integers (exact)
  orbital tables (a2)          test_Set_Orbs_Grid.jl:22-50 golden values
  Ylm_real                     textbook Cartesian real harmonics
  dYlm/dtheta, dYlm/dphi       central differences of Ylm_real (dphi form
                               is (1/sin theta) dY/dphi, as the reference)
  radial spline                mesh-node known answer (value = RWF[m])
  gradient tables (a3)         central differences of the orbital tables
  overlap merge (a5)           set definition (brute force)
  blocked H / rho (a7-a10)     point-major brute force from the set
                               definition; duality; Hermiticity; linearity
  Force3 (a11)                 central differences of E = 2 GridVol sum V rho
A planted mistake in any of these functions fails at least one check
(test_planted_mistakes_are_caught).
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from lcpao_grid import systems as S
from tolerance import within

HERE = os.path.dirname(os.path.abspath(__file__))


def golden(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def si2_golden_geom():
    u = golden("ucell_si2_g3.json")
    inp = u["inputs"]
    return u, O.Geom(inp["Latvecs"], inp["Gxyz"], inp["Atom_Cut1"], [13, 13], inp["Ngrid"], inp["Grid_Origin"])


def test_ucell_golden(si2_golden_geom):
    u, g = si2_golden_geom
    assert g.cpycell == u["CpyCell"]
    assert g.gridvol == pytest.approx(u["GridVol"], rel=1e-15)
    assert list(g.fnan) == u["FNAN"]
    for a in range(2):
        s, e = g.nb_off[a], g.nb_off[a + 1]
        assert list(g.natn[s:e] + 1) == u["natn"][a]
        assert list(g.ncn[s:e]) == u["ncn"][a]
        s, e = g.gl_off[a], g.gl_off[a + 1]
        assert g.gridn[a] == u["GridN_Atom"][a]
        assert list(g.glist[s:e]) == u["GridListAtom"][a]
        assert list(g.clist[s:e]) == u["CellListAtom"][a]
    assert np.allclose(g.atv[-1], u["atv_end"], rtol=1e-15)
    assert list(g.atv_ijk[-1]) == u["atv_ijk_end"]


def test_pao_metadata():
    # test_Read_PAO.jl:8-24: Si7.0 s2p2d1 -> 500-point mesh, 13 orbitals
    p = O.Pao("Si", 7.0, "s2p2d1")
    assert p.nmesh == 500 and p.total_orbs == 13 and p.maxl == 2 and p.num_basis == [2, 2, 1]
    g = O.Pao("Ga", 7.0, "s3p2d2")
    assert g.total_orbs == 19


def test_orbital_tables_golden(si2_golden_geom):
    _, g = si2_golden_geom
    o = golden("orbs_si2_g3.json")
    tabs = O.orbitals(g, [O.Pao("Si", 7.0, "s2p2d1")], [0, 0])
    for a in range(2):
        for i in range(13):
            gold = np.array(o["orbs"][str(a + 1)][str(i + 1)])
            mine = tabs[a][:, i]
            # the golden values are printed to 17 significant digits
            assert np.max(np.abs(mine - gold)) <= 1e-15 * max(1.0, np.abs(gold).max()) + 1e-16


def _ylm_textbook(l, m, x, y, z):
    r = np.sqrt(x * x + y * y + z * z)
    x, y, z = x / r, y / r, z / r
    pi = np.pi
    table = {
        (0, 0): 0.5 * np.sqrt(1 / pi),
        (1, 1): np.sqrt(3 / (4 * pi)) * x, (1, -1): np.sqrt(3 / (4 * pi)) * y, (1, 0): np.sqrt(3 / (4 * pi)) * z,
        (2, -2): 0.5 * np.sqrt(15 / pi) * x * y, (2, -1): 0.5 * np.sqrt(15 / pi) * y * z,
        (2, 0): 0.25 * np.sqrt(5 / pi) * (3 * z * z - 1), (2, 1): 0.5 * np.sqrt(15 / pi) * x * z,
        (2, 2): 0.25 * np.sqrt(15 / pi) * (x * x - y * y),
        (3, -3): 0.25 * np.sqrt(35 / (2 * pi)) * y * (3 * x * x - y * y),
        (3, -2): 0.5 * np.sqrt(105 / pi) * x * y * z,
        (3, -1): 0.25 * np.sqrt(21 / (2 * pi)) * y * (5 * z * z - 1),
        (3, 0): 0.25 * np.sqrt(7 / pi) * z * (5 * z * z - 3),
        (3, 1): 0.25 * np.sqrt(21 / (2 * pi)) * x * (5 * z * z - 1),
        (3, 2): 0.25 * np.sqrt(105 / pi) * z * (x * x - y * y),
        (3, 3): 0.25 * np.sqrt(35 / (2 * pi)) * x * (x * x - 3 * y * y),
    }
    return table[(l, m)]


def test_ylm_real_closed_forms():
    L = O.lib()
    rng = np.random.default_rng(7)
    for _ in range(50):
        th, ph = rng.uniform(0, np.pi), rng.uniform(0, 2 * np.pi)
        x, y, z = np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)
        for l in range(4):
            for m in range(-l, l + 1):
                assert L.orc_ylm_real(l, m, th, ph) == pytest.approx(_ylm_textbook(l, m, x, y, z), abs=1e-13)
                h = 1e-6
                dth = (L.orc_ylm_real(l, m, th + h, ph) - L.orc_ylm_real(l, m, th - h, ph)) / (2 * h)
                dph = (L.orc_ylm_real(l, m, th, ph + h) - L.orc_ylm_real(l, m, th, ph - h)) / (2 * h)
                if l > 0:
                    assert L.orc_dylm_dtheta(l, m, th, ph) == pytest.approx(dth, abs=1e-8)
                    # dYlmdphi_real carries the 1/sin(theta) of the spherical gradient
                    # (Set_dOrbitals_Grid.jl:271 uses dPx = -sin(phi)/R without it)
                    assert L.orc_dylm_dphi(l, m, th, ph) == pytest.approx(dph / np.sin(th), abs=1e-7)


def test_spline_known_answer_at_mesh_nodes():
    """An atom placed so that a grid point sits exactly at distance RV[m]
    along z: the cubic interpolation returns RWF[m] (y2 = 0, y1 = 1)."""
    p = O.Pao("Si", 7.0, "s2p2d1")
    for m in (40, 200, 400):
        R = p.rv[m]
        # one atom, a huge cell so that only the home image matters; grid point 0 at the origin
        L = 40.0
        g = O.Geom(np.eye(3) * L, [[0.0, 0.0, -R]], [7.0], [13], [8, 8, 8])
        tabs = O.orbitals(g, [p], [0])
        t = list(g.glist[: g.gridn[0]]).index(0)
        assert g.clist[t] == 0
        # s orbitals: RF * Y00;  pz: RF * sqrt(3/4pi) * cos(theta) with theta = 0
        y00 = 0.5 * np.sqrt(1 / np.pi)
        assert tabs[0][t, 0] == pytest.approx(p.rwf(0, 0)[m] * y00, rel=1e-14)
        assert tabs[0][t, 1] == pytest.approx(p.rwf(0, 1)[m] * y00, rel=1e-14)
        assert tabs[0][t, 4] == pytest.approx(p.rwf(1, 0)[m] * np.sqrt(3 / (4 * np.pi)), rel=1e-14)


def test_gradient_tables_match_finite_differences():
    st = S.si2(12)
    geo = lambda gx: O.Geom(st.Latvecs, gx, st.cutoffs(), st.norbs(), st.Ngrid)
    paos = [O.Pao("Si", 7.0, "s2p2d1")]
    g0 = geo(st.Gxyz)
    d = O.dorbitals(g0, paos, [0, 0])
    # theta = acos(z/r) loses digits near the poles (as in the reference), so
    # a step much below 1e-4 is dominated by that noise
    h = 1e-4
    for xyz in range(3):
        gp, gm = st.Gxyz.copy(), st.Gxyz.copy()
        gp[1, xyz] += h
        gm[1, xyz] -= h
        Gp, Gm = geo(gp), geo(gm)
        tp, tm = O.orbitals(Gp, paos, [0, 0]), O.orbitals(Gm, paos, [0, 0])
        # compare on entries present in all three lists (same (g, cell))
        k0 = {(int(a), int(b)): t for t, (a, b) in enumerate(zip(g0.glist[g0.gl_off[1]:], g0.clist[g0.gl_off[1]:]))}
        kp = {(int(a), int(b)): t for t, (a, b) in enumerate(zip(Gp.glist[Gp.gl_off[1]:], Gp.clist[Gp.gl_off[1]:]))}
        km = {(int(a), int(b)): t for t, (a, b) in enumerate(zip(Gm.glist[Gm.gl_off[1]:], Gm.clist[Gm.gl_off[1]:]))}
        common = [k for k in k0 if k in kp and k in km]
        fd = np.array([(tp[1][kp[k]] - tm[1][km[k]]) / (2 * h) for k in common])
        an = np.array([d[xyz][1][k0[k]] for k in common])
        assert np.max(np.abs(fd - an)) <= 1e-6 * np.abs(an).max()


@pytest.mark.parametrize("mk", [lambda: S.si2(16), lambda: S.nacl(1, 16)])
def test_overlap_merge_equals_set_definition(mk):
    st = mk()
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
    for q in range(0, g.npair, max(1, g.npair // 40)):
        t1, t2 = g.overlap_bruteforce(q)
        s, e = g.olg_off[q], g.olg_off[q + 1]
        assert np.array_equal(g.olg1[s:e], t1) and np.array_equal(g.olg2[s:e], t2)


class Small:
    def __init__(self, st, nprocs=1, rank=0):
        self.st = st
        self.g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid, (0, 0, 0), nprocs, rank)
        self.paos = [O.Pao(*S.BASES[s]) for s in st.species_names]
        self.tabs = O.orbitals(self.g, self.paos, st.atom2spe)
        full = st.ucell()
        self.V, self.DM = S.inputs(st, full, nwaves=10, nocc=6, seed=5)


@pytest.fixture(scope="module")
def si2s():
    return Small(S.si2(16))


@pytest.fixture(scope="module")
def nacl8s():
    return Small(S.nacl(1, 16))


@pytest.mark.parametrize("which", ["si2s", "nacl8s"])
def test_blocked_contractions_equal_bruteforce(which, request):
    c = request.getfixturevalue(which)
    H, M = O.hamiltonian(c.g, c.tabs, [c.V], with_abs=True)
    R, MR = O.density(c.g, c.tabs, [c.DM], with_abs=True)
    Hb, Rb, npp = O.bruteforce(c.g, c.tabs, c.V, c.DM)
    assert npp == c.g.olg1.size
    assert within(Hb, H[0], M[0])[0]
    assert within(Rb, R[0], MR[0])[0]


@pytest.mark.parametrize("which", ["si2s", "nacl8s"])
def test_duality_hermiticity_linearity(which, request):
    c = request.getfixturevalue(which)
    g = c.g
    H = O.hamiltonian(g, c.tabs, [c.V])[0]
    R = O.density(g, c.tabs, [c.DM])[0]
    # sum_pairs sum_ij D_ij H_ij = GridVol sum_g rho(g) V(g)
    lhs, rhs = float(np.dot(c.DM, H)), float(g.gridvol * np.dot(R, c.V))
    assert abs(lhs - rhs) <= 1e-13 * (np.abs(c.DM) @ np.abs(H))
    # H(A,B,Rn) = H(B,A,-Rn)^T
    key = {(int(a), int(b), tuple(g.atv_ijk[c_])): q for q, (a, b, c_) in enumerate(zip(g.pair_atom, g.pair_natn, g.pair_ncn))}
    for q in range(0, g.npair, 7):
        a, b = int(g.pair_atom[q]), int(g.pair_natn[q])
        qt = key[(b, a, tuple(-g.atv_ijk[g.pair_ncn[q]]))]
        na, nb = g.no[a], g.no[b]
        Hq = H[g.pair_hoff[q]:g.pair_hoff[q + 1]].reshape(na, nb)
        Ht = H[g.pair_hoff[qt]:g.pair_hoff[qt + 1]].reshape(nb, na)
        assert np.allclose(Hq, Ht.T, rtol=1e-12, atol=1e-15 * np.abs(H).max())
    # linearity in V and D
    H2 = O.hamiltonian(g, c.tabs, [2.0 * c.V - 0.5])[0]
    H1 = O.hamiltonian(g, c.tabs, [np.ones_like(c.V)])[0]
    assert np.allclose(H2, 2.0 * H - 0.5 * H1, rtol=1e-11, atol=1e-14 * np.abs(H1).max())
    R2 = O.density(g, c.tabs, [3.0 * c.DM])[0]
    assert np.allclose(R2, 3.0 * R, rtol=1e-12, atol=1e-16)
    # V == 1: grid overlap, charge = sum D o S_grid
    assert float(np.dot(c.DM, H1)) == pytest.approx(float(g.gridvol * R.sum()), rel=1e-12)


def _nonsymmetric_dm(st, seed):
    """Flat DM from a random non-symmetric Gamma matrix: as the up-down
    channels of the non-collinear form, D(B,A,-Rn) != D(A,B,Rn)^T."""
    n = int(st.norbs().sum())
    rng = S.SplitMix64(seed)
    Dg = rng.normal(n * n).reshape(n, n) / n
    return S.density_matrix_flat(st.ucell(), Dg)


@pytest.mark.parametrize("which", ["si2s", "nacl8s"])
def test_blocked_density_equals_bruteforce_for_nonsymmetric_dm(which, request):
    """The blocked density restatement takes each pair's own block (no
    symmetry assumed), which the non-collinear channels DM3/DM4 need."""
    c = request.getfixturevalue(which)
    D = _nonsymmetric_dm(c.st, 11)
    R, MR = O.density(c.g, c.tabs, [D], with_abs=True)
    _, Rb, _ = O.bruteforce(c.g, c.tabs, c.V, D)
    ok, msg = within(Rb, R[0], MR[0])
    assert ok, msg
    # four spins are four independent densities
    R4 = O.density(c.g, c.tabs, [c.DM, D, -D, 0.5 * c.DM])
    assert np.array_equal(R4[1], R[0]) and np.array_equal(R4[2], -R[0])


@pytest.fixture(scope="module")
def nacl8s_r1of2():
    return Small(S.nacl(1, 16), nprocs=2, rank=1)


@pytest.mark.parametrize("nworkers,use_blas,which", [(1, True, "nacl8s"), (3, True, "nacl8s"), (2, False, "nacl8s"),
                                                     (3, True, "nacl8s_r1of2")])
def test_cpu_baseline_equals_blocked_restatement(nworkers, use_blas, which, request):
    """The CPU baseline (the same blocked algorithm, pairs split over
    threads, DGEMM from OpenBLAS) gives the restatement's H and rho within
    the FP64 bar, also on a rank's share of the pairs, so it can check
    full-size FP64 results."""
    c = request.getfixturevalue(which)
    H, M = O.hamiltonian(c.g, c.tabs, [c.V], with_abs=True)
    R, MR = O.density(c.g, c.tabs, [c.DM], with_abs=True)
    _, done, Hc, Rc = O.cpu_baseline(c.g, c.tabs, c.V, c.DM, nworkers=nworkers, use_blas=use_blas)
    assert done == c.g.npair
    ok, msg = within(Hc, H[0], M[0])
    assert ok, msg
    ok, msg = within(Rc, R[0], MR[0])
    assert ok, msg
    # with absolute inputs it returns the sums of absolute terms
    _, _, Ha, Ra = O.cpu_baseline(c.g, [np.abs(t) for t in c.tabs], np.abs(c.V), np.abs(c.DM), nworkers=nworkers,
                                  use_blas=use_blas)
    assert np.allclose(Ha, M[0], rtol=1e-13, atol=0) and np.allclose(Ra, MR[0], rtol=1e-13, atol=0)


def _energy(st, gxyz, V, Dg):
    s2 = S.Structure("fd", st.Latvecs, np.linalg.solve(st.Latvecs.T, np.asarray(gxyz).T).T, st.species, st.Ngrid, st.a_cell)
    g = O.Geom(s2.Latvecs, s2.Gxyz, s2.cutoffs(), s2.norbs(), s2.Ngrid)
    tabs = O.orbitals(g, [O.Pao(*S.BASES[s]) for s in s2.species_names], s2.atom2spe)
    DM = S.density_matrix_flat(s2.ucell(), Dg)
    R = O.density(g, tabs, [DM])[0]
    return 2.0 * g.gridvol * float(np.dot(V, R))


def test_force3_matches_finite_difference_of_grid_energy():
    st = S.si(1, 16, disp=0.1, seed=4)
    u = st.ucell()
    V = S.potential(st, 10, 4)
    Dg = S.density_gamma(int(u.Total_NumOrbs.sum()), 8, 4)
    DM = S.density_matrix_flat(u, Dg)
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
    paos = [O.Pao("Si", 7.0, "s2p2d1")]
    tabs = O.orbitals(g, paos, st.atom2spe)
    dt = O.dorbitals(g, paos, st.atom2spe)
    F = O.force3(g, tabs, dt, [V], [DM])
    h = 1e-4
    for atom, xyz in ((0, 0), (3, 2), (5, 1)):
        gp, gm = st.Gxyz.copy(), st.Gxyz.copy()
        gp[atom, xyz] += h
        gm[atom, xyz] -= h
        fd = (_energy(st, gp, V, Dg) - _energy(st, gm, V, Dg)) / (2 * h)
        # VpotForce is dE/dR (the reference's force arrays are gradients)
        assert F[atom, xyz] == pytest.approx(fd, rel=2e-4, abs=2e-6)


def test_split_ranks_partition_the_pairs():
    st = S.nacl(1, 12)
    full = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
    parts = [O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid, (0, 0, 0), 3, r) for r in range(3)]
    assert sum(p.npair for p in parts) == full.npair
    assert sum(p.my_hsize for p in parts) == full.total_hsize == full.my_hsize
    assert [p.my_hoff for p in parts] == [0, parts[0].my_hsize, parts[0].my_hsize + parts[1].my_hsize]
    assert np.array_equal(np.concatenate([p.numolg for p in parts]), full.numolg)


def test_planted_mistakes_are_caught(si2s):
    """Mistakes planted in the oracle's outputs (on the CPU) are detected by
    the same checks that pin it: a dropped term, a wrong sign, a transposed
    block and a wrong V index each fail one of them."""
    c = si2s
    g = c.g
    H, M = O.hamiltonian(g, c.tabs, [c.V], with_abs=True)
    Hb, Rb, _ = O.bruteforce(g, c.tabs, c.V, c.DM)
    R = O.density(g, c.tabs, [c.DM])[0]
    # dropped term: one pair block zeroed
    bad = H[0].copy()
    q = int(np.argmax(g.numolg))
    bad[g.pair_hoff[q]:g.pair_hoff[q + 1]] = 0.0
    assert not within(Hb, bad, M[0])[0]
    # wrong sign of V
    assert not within(Hb, -H[0], M[0])[0]
    # transposed non-symmetric block
    na, nb = g.no[g.pair_atom[q]], g.no[g.pair_natn[q]]
    blk = H[0][g.pair_hoff[q]:g.pair_hoff[q + 1]].reshape(na, nb)
    bad = H[0].copy()
    bad[g.pair_hoff[q]:g.pair_hoff[q + 1]] = blk.T.reshape(-1)
    assert not within(Hb, bad, M[0])[0] or np.allclose(blk, blk.T)
    # V read at the wrong grid point (shifted by one) breaks duality
    Hs = O.hamiltonian(g, c.tabs, [np.roll(c.V, 1)])[0]
    assert abs(np.dot(c.DM, Hs) - g.gridvol * np.dot(R, c.V)) > 1e-6 * abs(np.dot(c.DM, Hs))
