"""Grid fields next to the contractions, CPU side: the PAO/VPS readers, the
oracle's KumoF, Set_AdenPCC_Grid and Solve_Poisson! against the reference's
golden vectors (test_Set_AdenPCC_Grid.jl, test_Poisson.jl) and against
closed forms that a wrong index, sign or factor would fail."""
import json
import os

import numpy as np
import pytest

import oracle as O
from lcpao_grid.pao import PAO, Read_PAO
from lcpao_grid.vps import Read_VPS

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rec(lat):
    return 2 * np.pi * np.linalg.inv(np.asarray(lat, float).T)


# ---- readers ----------------------------------------------------------------
def test_vps_reader_fields():
    v = Read_VPS("Si", "", "LDA")
    assert v.Spe_Core_Charge == 4.0 and v.is_pcc and v.Spe_Num_Mesh_VPS == 500
    assert v.Spe_Atomic_PCC.size == 504
    p = v.Spe_Atomic_PCC
    assert p[0] == 2 * p[1] - p[2] and p[501] == 2 * p[500] - p[499] and p[502] == 0 and p[503] == 0
    dx = np.diff(v.Spe_VPS_XV)
    assert np.allclose(dx, dx[0], rtol=1e-9)
    assert np.allclose(np.exp(v.Spe_VPS_XV), v.Spe_VPS_RV, rtol=1e-9)
    assert Read_VPS("Fe", "H", "GGA-PBE").Spe_Core_Charge == 16.0
    with pytest.raises(ValueError):
        Read_VPS("Si", "", "B3LYP")
    with pytest.raises(ValueError):
        Read_VPS("Xx", "", "LDA")


@pytest.mark.parametrize("sym,cut,basis,extra,cc", [("Si", 7.0, "s2p2d1", "", 4.0), ("Ga", 7.0, "s3p2d2", "", 13.0),
                                                    ("As", 7.0, "s3p2d2", "", 15.0), ("Fe", 6.0, "s2p2d1", "H", 16.0)])
def test_atomic_density_holds_the_core_charge(sym, cut, basis, extra, cc):
    """Read_PAO.jl:288-303: after renormalisation 4 pi dx sum rho r^3 = core charge."""
    p = Read_PAO(cc, sym, cut, basis, extra)
    n, xv, den = p.Spe_Num_Mesh_PAO, p.Spe_PAO_XV, p.Spe_Atomic_Den
    total = 4 * np.pi * (xv[1] - xv[0]) * np.sum(den[1:n + 1] * np.exp(3 * xv))
    assert abs(total - cc) < 1e-12 * cc
    assert den[0] == 2 * den[1] - den[2] and den[n + 1] == 2 * den[n] - den[n - 1]
    raw = PAO(sym, cut, basis, extra=extra).Spe_Atomic_Den[1:n + 1]
    nz = raw != 0
    assert np.array_equal(den[1:n + 1] == 0, ~nz)
    assert np.allclose(den[1:n + 1][nz] / raw[nz], den[1] / raw[0], rtol=1e-14)


# ---- KumoF ------------------------------------------------------------------
def _mesh(n=200):
    xv = np.linspace(-7.0, 2.5, n)
    return xv, np.exp(xv)


def test_kumof_reproduces_data_linear_in_the_index():
    """The four-point cubic (KumoF.jl:41-60) is exact for data linear in the
    mesh index; a shifted index or a dropped term breaks this."""
    xv, rv = _mesh()
    n = xv.size
    yv = np.zeros(n + 4)
    yv[1:n + 1] = 0.3 + 0.01 * np.arange(n)
    yv[0], yv[n + 1] = 2 * yv[1] - yv[2], 2 * yv[n] - yv[n - 1]
    rng = np.random.default_rng(1)
    for x in np.concatenate([rng.uniform(xv[0], xv[-1], 50), xv[::17]]):
        t = (n - 1) * (x - xv[0]) / (xv[-1] - xv[0])
        assert abs(O.kumof(n, x, xv, rv, yv) - (0.3 + 0.01 * t)) < 1e-13
    assert O.kumof(n, xv[-1] + 3.0, xv, rv, yv) == yv[n]  # clamped at the mesh end


def test_kumof_below_the_mesh_is_a_parabola_in_r():
    """KumoF.jl:3-40: a r^2 + b below the mesh, equal to a constant for
    constant data."""
    xv, rv = _mesh()
    n = xv.size
    yv = np.zeros(n + 4)
    yv[1:n + 1] = np.cos(rv)
    yv[0], yv[n + 1] = 2 * yv[1] - yv[2], 2 * yv[n] - yv[n - 1]
    b = O.kumof(n, -np.inf, xv, rv, yv)
    r1, r2 = 0.3 * rv[0], 0.8 * rv[0]
    f1, f2 = (O.kumof(n, np.log(r), xv, rv, yv) for r in (r1, r2))
    assert abs((f1 - b) / (f2 - b) - (r1 / r2) ** 2) < 1e-9
    yc = np.zeros(n + 4)
    yc[: n + 2] = 0.7
    assert abs(O.kumof(n, np.log(0.5 * rv[0]), xv, rv, yc) - 0.7) < 1e-14


# ---- Set_AdenPCC_Grid ----------------------------------------------------------
def _case(c):
    P = Read_PAO(c["pao"]["core_charge"], c["pao"]["symbol"], c["pao"]["cutoff"], c["pao"]["basis"], c["pao"]["extra"])
    V = Read_VPS(c["vps"]["symbol"], c["vps"]["extra"], c["vps"]["xc"])
    nat = len(c["Gxyz"])
    g = O.Geom(c["Latvecs"], c["Gxyz"], c["Atom_Cut1"], [P.Spe_Total_NumOrbs] * nat, c["Ngrid"], c["Grid_Origin"])
    rad = [(P.Spe_PAO_XV, P.Spe_PAO_RV, P.Spe_Atomic_Den, V.Spe_VPS_XV, V.Spe_VPS_RV, V.Spe_Atomic_PCC)]
    return P, V, g, rad


def _golden_cases():
    return json.load(open(os.path.join(GOLD, "aden_pcc_g3.json")))["cases"]


@pytest.mark.parametrize("k", range(3))
def test_aden_pcc_oracle_matches_the_reference_vectors(k):
    """PCCDensity agrees to round-off.  ADensity (and the spin densities made
    from it) agree to a constant factor within 6e-11 of 1 over every point:
    the vectors were written with a renormalisation that differs in the 11th
    digit; the reference's own test compares at rtol sqrt(eps) = 1.5e-8."""
    c = _golden_cases()[k]
    P, V, g, rad = _case(c)
    nspin = {"off": 1, "on": 2, "nc": 4}[c["SpinPol"]]
    A, Pc, R = O.aden_pcc(g, nspin, [s - 1 for s in c["atom2spe"]], rad, c["Init_Atoms_Nspin"],
                          [a[:2] for a in c["Init_Atoms_Angle"]])
    gp, ga = np.array(c["PCCDensity_Grid"]), np.array(c["ADensity_Grid"])
    assert np.abs(Pc - gp).max() <= 1e-14 * np.abs(gp).max()
    ratio = A / ga
    assert np.abs(ratio - 1).max() < 6e-11 and ratio.max() - ratio.min() < 2e-12
    for s in range(nspin):
        gd = np.array(c["Density_Grid"][s])
        assert np.abs(R[s] - gd).max() <= 6e-11 * np.abs(gd).max()


def test_aden_pcc_spin_modes_are_consistent():
    """nc rotation (Set_AdenPCC_Grid.jl:125-138) against the collinear modes:
    theta = 0 gives rho_up/down of "on"; theta = pi/2 puts the moment in x."""
    c = _golden_cases()[2]
    P, V, g, rad = _case(c)
    ns = [[9.0, 5.0]]
    A1, _, R1 = O.aden_pcc(g, 1, [0], rad, ns)
    A2, _, R2 = O.aden_pcc(g, 2, [0], rad, ns)
    A4, _, R4 = O.aden_pcc(g, 4, [0], rad, ns, [[0.0, 0.3]])
    assert np.array_equal(A1, A2) and np.array_equal(A1, A4) and np.array_equal(R1[0], A1)
    assert np.allclose(R4[0], R2[0], rtol=1e-14) and np.allclose(R4[1], R2[1], rtol=1e-14)
    assert np.abs(R4[2]).max() == 0 and np.abs(R4[3]).max() == 0
    _, _, Rx = O.aden_pcc(g, 4, [0], rad, ns, [[np.pi / 2, 0.0]])
    tot = 2 * A1
    assert np.allclose(Rx[0], 0.5 * tot, rtol=1e-12) and np.allclose(Rx[1], 0.5 * tot, rtol=1e-12)
    assert np.allclose(Rx[2], 0.5 * (4.0 / 14.0) * tot, rtol=1e-12)


# ---- Solve_Poisson! ------------------------------------------------------------
def test_poisson_oracle_matches_the_reference_vector():
    d = json.load(open(os.path.join(GOLD, "poisson_si2_g3.json")))
    V = O.poisson(d["inputs"]["Ngrid"], _rec(d["inputs"]["Latvecs"]), d["ADensity_Grid"])
    assert np.abs(V - np.array(d["dVHart_Grid"])).max() < 1e-15


def _grid_coords(N):
    n1, n2, n3 = np.meshgrid(*[np.arange(n) for n in N], indexing="ij")
    return n1.reshape(-1), n2.reshape(-1), n3.reshape(-1)


def test_poisson_plane_wave_closed_form():
    """cos(G.r) -> 4 pi cos(G.r) / |G|^2 for G on the grid (a skew cell)."""
    lat = np.array([[6.0, 0.4, 0.0], [0.5, 5.5, 0.3], [0.2, -0.1, 7.0]])
    N = (12, 10, 8)
    rec = _rec(lat)
    n1, n2, n3 = _grid_coords(N)
    for m in [(1, 0, 0), (2, -3, 1), (-5, 4, 3)]:
        ph = 2 * np.pi * (m[0] * n1 / N[0] + m[1] * n2 / N[1] + m[2] * n3 / N[2])
        G = m[0] * rec[0] + m[1] * rec[1] + m[2] * rec[2]
        V = O.poisson(N, rec, np.cos(ph) + 0.25)  # the G = 0 part drops out
        assert np.abs(V - 4 * np.pi / (G @ G) * np.cos(ph)).max() < 1e-13


def test_poisson_middle_index_of_an_odd_axis_counts_as_negative():
    """Poisson.jl:65-72: on an axis of 5 points index 2 is taken as -3 and
    index 3 as -2, so cos(2 pi 2 n / 5) gets the mean of the two factors."""
    lat = np.diag([5.0, 6.0, 7.0])
    N = (4, 4, 5)
    rec = _rec(lat)
    _, _, n3 = _grid_coords(N)
    ph = 2 * np.pi * 2 * n3 / 5
    V = O.poisson(N, rec, np.cos(ph))
    b2 = rec[2] @ rec[2]
    want = 0.5 * (4 * np.pi / (9 * b2) + 4 * np.pi / (4 * b2)) * np.cos(ph)
    assert np.abs(V - want).max() < 1e-13
