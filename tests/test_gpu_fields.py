"""Grid fields on the device against the oracle and the reference's vectors:
Set_AdenPCC_Grid (lcpao_grid_aden_pcc) and Solve_Poisson!
(lcpao_grid_poisson, the library's own LDS-resident mixed-radix FFT passes).

Atomic densities: every grid point's sum runs in the reference's atom order
on both sides, so the device agrees with the oracle to the last few ulps of
log/exp (bound 1e-13 of the field's maximum).  Poisson: the oracle's FFT and
the device FFT differ in rounding only (bound 1e-13 of the maximum); the reference's
vector at 1e-15; at the 144^3 GaAs grid a plane wave against its closed
form."""
import json
import os

import numpy as np
import pytest

import oracle as O
from lcpao_grid import systems as S
from lcpao_grid.grid import GridContext
from lcpao_grid.pao import Read_PAO
from lcpao_grid.ucell import UCell
from lcpao_grid.vps import Read_VPS

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def _rad(paos, vpss):
    return [(p.Spe_PAO_XV, p.Spe_PAO_RV, p.Spe_Atomic_Den, v.Spe_VPS_XV, v.Spe_VPS_RV, v.Spe_Atomic_PCC)
            for p, v in zip(paos, vpss)]


@pytest.mark.parametrize("k", range(3))
def test_aden_pcc_reference_vectors(k):
    c = json.load(open(os.path.join(GOLD, "aden_pcc_g3.json")))["cases"][k]
    print("case", c["name"], flush=True)
    P = Read_PAO(c["pao"]["core_charge"], c["pao"]["symbol"], c["pao"]["cutoff"], c["pao"]["basis"], c["pao"]["extra"])
    V = Read_VPS(c["vps"]["symbol"], c["vps"]["extra"], c["vps"]["xc"])
    nat = len(c["Gxyz"])
    a2s = [s - 1 for s in c["atom2spe"]]
    u = UCell(1, c["Latvecs"], nat, a2s, c["Gxyz"], c["Atom_Cut1"], c["Ngrid"], c["Grid_Origin"],
              Total_NumOrbs=[P.Spe_Total_NumOrbs] * nat)
    ctx = GridContext(u)
    nspin = {"off": 1, "on": 2, "nc": 4}[c["SpinPol"]]
    ins, ang = c["Init_Atoms_Nspin"], [a[:2] for a in c["Init_Atoms_Angle"]]
    A, Pc, R = ctx.aden_pcc(nspin, [P], [V], ins, ang)
    g = O.Geom(c["Latvecs"], c["Gxyz"], c["Atom_Cut1"], [P.Spe_Total_NumOrbs] * nat, c["Ngrid"], c["Grid_Origin"])
    A0, P0, R0 = O.aden_pcc(g, nspin, a2s, _rad([P], [V]), ins, ang)
    assert _rel(A, A0) < 1e-13 and _rel(Pc, P0) < 1e-13
    for s in range(nspin):
        assert _rel(R[s], R0[s]) < 1e-13
    assert _rel(Pc, c["PCCDensity_Grid"]) < 1e-13
    assert _rel(A, c["ADensity_Grid"]) < 6e-11  # constant-factor offset of the vectors, see test_fields_oracle
    ctx.close()


@pytest.fixture(scope="module")
def si64_fields():
    st = S.si64()
    u = st.ucell()
    ctx = GridContext(u)
    paos, vpss = st.atomic()
    g = O.Geom(st.Latvecs, st.Gxyz, st.cutoffs(), st.norbs(), st.Ngrid)
    yield st, u, ctx, paos, vpss, g
    ctx.close()


@pytest.mark.parametrize("nspin", [1, 2, 4])
def test_aden_pcc_si64_matches_oracle(si64_fields, nspin):
    st, u, ctx, paos, vpss, g = si64_fields
    rng = np.random.default_rng(nspin)
    ins = np.column_stack([rng.uniform(1, 3, st.Natom), rng.uniform(1, 3, st.Natom)])
    ang = np.column_stack([rng.uniform(0, np.pi, st.Natom), rng.uniform(0, 2 * np.pi, st.Natom)])
    print("si64 nspin", nspin, flush=True)
    A, Pc, R = ctx.aden_pcc(nspin, paos, vpss, ins, ang)
    A0, P0, R0 = O.aden_pcc(g, nspin, st.atom2spe, _rad(paos, vpss), ins, ang)
    assert _rel(A, A0) < 1e-13 and _rel(Pc, P0) < 1e-13
    for s in range(nspin):
        assert _rel(R[s], R0[s]) < 1e-13
    # the atomic charge: GridVol sum of 2 ADensity = sum of the core charges (grid quadrature)
    q = 2 * A.sum() * u.GridVol
    assert abs(q - 4.0 * st.Natom) < 1e-5 * 4.0 * st.Natom


def test_poisson_reference_vector():
    d = json.load(open(os.path.join(GOLD, "poisson_si2_g3.json")))
    inp = d["inputs"]
    u = UCell(1, inp["Latvecs"], 2, [0, 0], inp["Gxyz"], inp["Atom_Cut1"], inp["Ngrid"], inp["Grid_Origin"],
              Total_NumOrbs=[13, 13])
    ctx = GridContext(u)
    rec = 2 * np.pi * np.linalg.inv(np.asarray(inp["Latvecs"]).T)
    V = ctx.poisson([np.array(d["ADensity_Grid"])], None, rec, nspin=0)
    assert np.abs(V - np.array(d["dVHart_Grid"])).max() < 1e-15
    ctx.close()


@pytest.mark.parametrize("ngrid", [(32, 32, 32), (15, 9, 5), (14, 11, 13)])
def test_poisson_matches_oracle(ngrid):
    """Spin modes of Set_Density_xyz!, an odd grid (middle index rule) and
    radices 7, 11, 13."""
    st = S.si2()
    u = UCell(1, st.Latvecs, 2, [0, 0], st.Gxyz, [7.0, 7.0], ngrid, Total_NumOrbs=[13, 13])
    ctx = GridContext(u)
    rec = 2 * np.pi * np.linalg.inv(st.Latvecs.T)
    nn = int(np.prod(ngrid))
    rng = np.random.default_rng(7)
    r0, r1, ra = rng.uniform(0, 0.1, nn), rng.uniform(0, 0.1, nn), rng.uniform(0, 0.1, nn)
    for nspin, dens in [(1, 2 * r0 - 2 * ra), (2, r0 + r1 - 2 * ra), (4, r0 + r1 - 2 * ra)]:
        print("poisson", ngrid, "nspin", nspin, flush=True)
        V = ctx.poisson([r0, r1][:min(nspin, 2)], ra, rec, nspin=nspin)
        Vr = O.poisson(ngrid, rec, dens)
        assert _rel(V, Vr) < 1e-13
    ctx.close()


def test_poisson_plane_wave_gaas216_device_resident():
    import torch
    st = S.gaas216()
    u = st.ucell()
    ctx = GridContext(u)
    rec = 2 * np.pi * np.linalg.inv(st.Latvecs.T)
    N = st.Ngrid
    n1, n2, n3 = np.meshgrid(*[np.arange(n) for n in N], indexing="ij")
    m = (3, -2, 5)
    ph = 2 * np.pi * (m[0] * n1 / N[0] + m[1] * n2 / N[1] + m[2] * n3 / N[2])
    G = m[0] * rec[0] + m[1] * rec[1] + m[2] * rec[2]
    rho = (0.5 * np.cos(ph) + 0.03).reshape(-1)
    dev = torch.device("cuda:0")
    dr = torch.from_numpy(rho).to(dev)
    da = torch.zeros_like(dr)
    dV = torch.empty_like(dr)
    ctx.poisson_dev([dr], da, rec, dV, nspin=1)  # Density_xyz = 2 rho
    want = 2 * 0.5 * 4 * np.pi / (G @ G) * np.cos(ph).reshape(-1)
    assert np.abs(dV.cpu().numpy() - want).max() < 1e-12 * np.abs(want).max()
    ctx.close()


def test_atomic_density_to_hartree_stays_on_device():
    """aden_pcc_dev then poisson_dev: Density_xyz = 2 rho - 2 rho_atom is
    zero for rho = rho_atom ("off"), so dVHart vanishes."""
    import torch
    st = S.si2()
    u = st.ucell()
    ctx = GridContext(u)
    paos, vpss = st.atomic()
    dev = torch.device("cuda:0")
    dA, dP, dR = (torch.empty(u.NN, dtype=torch.float64, device=dev) for _ in range(3))
    ctx.aden_pcc_dev(1, paos, vpss, [[2.0, 2.0]] * 2, dA, dP, [dR])
    assert torch.equal(dA, dR)
    dV = torch.empty_like(dA)
    ctx.poisson_dev([dR], dA, 2 * np.pi * np.linalg.inv(st.Latvecs.T), dV)
    assert float(dV.abs().max()) == 0.0
    ctx.close()
